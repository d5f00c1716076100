"""Builds the gfx950 HIP library ratbite_amd/libratbite_hip.so in-tree.

Every source under csrc/ is compiled by hipcc (device code for gfx950 only)
into csrc/../_build/ and linked into one shared library with a C ABI
(include/ratbite_hip.h). Run as a module or call build().
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libratbite_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-I" + os.path.join(HERE, "..", "include")]
# Shading, light-tree and intersection arithmetic is evaluated as written (no
# implicit FMA contraction), the same operation sequence the CPU restatement
# uses: near-specular lobe PDFs, sharp spherical-Gaussian amplitudes and shared
# triangle edges all amplify a one-ulp rounding difference. Kernels that want
# fused multiply-adds spell them out with fmaf().
DEVICE = ["-x", "hip", "--offload-arch=gfx950", "-munsafe-fp-atomics",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=off"]


def sources():
    out = []
    for root, _, files in os.walk(CSRC):
        for f in sorted(files):
            if f.endswith(".hip") or f.endswith(".cpp"):
                out.append(os.path.join(root, f))
    return sorted(out)


def headers():
    out = []
    for root, _, files in os.walk(CSRC):
        out += [os.path.join(root, f) for f in files if f.endswith((".hpp", ".h"))]
    out.append(os.path.join(HERE, "..", "include", "ratbite_hip.h"))
    return out


def _obj(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    return os.path.join(OUT_DIR, rel + ".o")


def _compile(src, force):
    obj = _obj(src)
    newest_dep = max([os.path.getmtime(src)] + [os.path.getmtime(h) for h in headers()])
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= newest_dep:
        return obj, None
    if os.sep + "host" + os.sep in src:
        # host-only C++: no device pass, no floating-point contraction
        cmd = [HIPCC] + COMMON + ["-ffp-contract=off", "-c", src, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + DEVICE + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        return obj, "compile failed: %s\n%s\n%s" % (" ".join(cmd), p.stdout, p.stderr)
    return obj, None


def build(force=True, jobs=8, verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, force), srcs))
    errs = [e for _, e in results if e]
    if errs:
        raise RuntimeError("\n".join(errs))
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), p.stdout, p.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
