"""Writes tests/golden/reference_luts.npz: the reference's six BxDF LUT data
files (bin/albedoConductorLUT.bin, albedoDielectric{Outer,Inner,
ReflectiveOuter,ReflectiveInner}LUT.bin, sheenLUT.bin; raw float32, loaded by
rendering_interface.cpp:518-622) as arrays in this project's table layout,
for the test that renders with the reference's tables as inputs. Run in a
checkout that has the reference next to it; the fixture is data only."""
import os
import sys

import numpy as np

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/bin"
NAMES = {"conductor": ("albedoConductorLUT", (32, 32)),
         "diel_outer": ("albedoDielectricOuterLUT", (32, 32, 32)),
         "diel_inner": ("albedoDielectricInnerLUT", (32, 32, 32)),
         "refl_outer": ("albedoDielectricReflectiveOuterLUT", (32, 32, 32)),
         "refl_inner": ("albedoDielectricReflectiveInnerLUT", (32, 32, 32)),
         "sheen": ("sheenLUT", (32, 32, 4))}
out = {}
for key, (name, shape) in NAMES.items():
    a = np.fromfile(os.path.join(SRC, name + ".bin"), np.float32)
    out[key] = a.reshape(shape)
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_luts.npz"), **out)
print({k: v.shape for k, v in out.items()})
