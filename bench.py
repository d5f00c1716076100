#!/usr/bin/env python3
"""Throughput of the spectral path-tracing hot path on one MI355X.

Workload: the seeded 1920x1080 scene of 200 displaced level-5 icospheres
(4.1M triangles), 10,000 textured emissive triangles behind the Gaussian light
tree, a 2048x1024 log-normal env map with a sun, the four-way material mix,
64 spp per frame and path depth 16. One step renders one full frame
(rbh_render with sampleCount 64) into the FP64 accumulation buffer; the scene,
LBVH, light tree and env-map tables are built once before the timed region.

Printed: one JSON line with samples/s (whole frames), Mrays/s, the roofline
of the dominant kernel measured with HIP events on the render stream, the
same frames replayed as HIP graphs, the guided study on the same scene
(relMSE of guided vs plain at equal wall time, one-sample MIS at selection
0.5, against a plain FP64-accumulated reference), and the CPU baseline: the
scalar C++ restatement (oracle/, built -march=native on this host) on the
centre 256x256 crop of the same frame, timed on this process's host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per unit of work, from the reference's own data: an
# extension ray moves a 32 B ray in and a 16 B hit out, a shadow ray 32 B in
# and a 4 B flag out, a shaded vertex reads and writes the 136 B path state.
BYTES_CLOSEST = 48
BYTES_SHADOW = 36
BYTES_VERTEX = 272
BYTES_PIXEL_LAUNCH = 64      # FP64 RGBW read-modify-write per pixel per pass
KERNEL_NAMES = {"trace": "k_trace", "shade": "k_shade", "shadow": "k_shadow"}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed frames")
    ap.add_argument("--warmup", type=int, default=1, help="untimed frames")
    ap.add_argument("--spp", type=int, default=64, help="samples per pixel per frame")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-spp", type=int, default=256, help="spp of the CPU baseline crop")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0: the host cores this process may use (OMP_NUM_THREADS, else the affinity mask)")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the launch shape first (short 8-spp renders) instead of the library default, "
                         "which was picked by a sweep of full frames of this workload")
    ap.add_argument("--copy-mib", type=int, default=1024,
                    help="size of the device-to-device copy timed first (the measured bandwidth ceiling; "
                         "its k_copy dispatch also scales PMC byte counters)")
    ap.add_argument("--workload", choices=["spheres", "batch", "guided", "large", "cdf"], default="spheres",
                    help="spheres: the 1080p frame (default); batch: 64 seeded scenes in one batched launch; "
                         "guided: relMSE at equal time, guided vs plain; large: the 8.4M-triangle precision-split "
                         "scene at 256 spp, depth 24, plain and FP8-guided; cdf: the env-map CDF build at "
                         "2048x1024 and 8192x4096")
    ap.add_argument("--scenes", type=int, default=64, help="batch workload: number of scenes")
    ap.add_argument("--budget-ms", type=float, default=1000.0, help="guided workload: time budget per method")
    ap.add_argument("--ref-spp", type=int, default=None,
                    help="guided / large workloads: spp of the plain reference (default 8192 / 4096)")
    ap.add_argument("--scene", choices=["spheres", "bounce"], default="spheres", help="guided workload: scene")
    ap.add_argument("--chunk-spp", type=int, default=4, help="guided workload: spp per render call")
    ap.add_argument("--train-iters", type=int, default=8, help="guided workload: collect+train rounds")
    ap.add_argument("--train-steps", type=int, default=8, help="guided workload: Adam steps per round")
    ap.add_argument("--train-batch", type=int, default=1 << 18, help="guided workload: records per Adam step")
    ap.add_argument("--train-lr", type=float, default=1e-3, help="guided study: Adam learning rate")
    ap.add_argument("--fixed-selection", type=int, default=1,
                    help="guided study: 1 = one-sample MIS at selection 0.5, 0 = learned selection")
    ap.add_argument("--guided-budget-ms", type=float, default=500.0,
                    help="default line: wall time per method of the guided study (0: no study)")
    ap.add_argument("--guided-ref-spp", type=int, default=4096, help="default line: spp of the study's reference")
    ap.add_argument("--graph-steps", type=int, default=2,
                    help="default line: frames timed again with the passes replayed as HIP graphs (0: none)")
    return ap.parse_args()


def traffic_per_launch(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE of one frame
    of this workload, scaled by the counters' reading of a plain copy), from
    the committed PMC summary profiles/spheres1080p_pmc.txt (a line
    '<kernel> traffic_bytes_per_launch: X'), or None when it holds none."""
    path = os.path.join(ROOT, "profiles", "spheres1080p_pmc.txt")
    if not os.path.exists(path):
        return None
    key = "%s traffic_bytes_per_launch:" % kernel
    with open(path) as f:
        for line in f:
            if line.startswith(key):
                return float(line.split(":", 1)[1])
    return None


def host_threads():
    """The host cores this process may use: the CPU share the machine gives
    it (OMP_NUM_THREADS where the environment sets one, else the affinity
    mask)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(sc, args, tree):
    """The oracle (scalar C++ restatement, OpenMP over 16x16 tiles) on the centre
    256x256 crop of the frame; returns samples/s of path tracing alone."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import tempfile
    import oracle_lib as O
    threads = args.cpu_threads or host_threads()
    build_s = O.use_native_build(os.path.join(tempfile.gettempdir(), "rbh_oracle_native_%d" % os.getpid()),
                                 jobs=threads)
    cw, ch = min(256, sc.width), min(256, sc.height)
    crop = ((sc.width - cw) // 2, (sc.height - ch) // 2, cw, ch)
    lp = sc.launch_params(offset=0, count=args.cpu_spp, depth=args.depth)
    _, st = O.render(sc, lp, threads=threads, crop=crop, oscene=O.OracleScene(sc, tree=tree))
    v = cw * ch * args.cpu_spp / st["renderSeconds"]
    model = None
    try:
        with open("/proc/cpuinfo") as fi:
            for line in fi:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    c1 = cpu_c1(threads)
    return {"value": v, "unit": "samples/s", "cores": threads, "kind": "port", "nproc": os.cpu_count(),
            "cpu_model": model, "c1": c1, "build": "g++ -O3 -march=native -ffp-contract=off -fopenmp (%.0f s)" % build_s,
            "sample": "centre %dx%d crop of the same frame at %d spp, depth %d (%.1f s of path tracing; "
                      "CPU BVH build of %.1f s excluded)" % (cw, ch, args.cpu_spp, args.depth, st["renderSeconds"],
                                                             st["bvhSeconds"])}


def cpu_c1(threads):
    """The CPU-runnable configuration in full (256x256 Cornell box with 8 UV
    spheres and a 4000 K blackbody emitter, 16 spp, depth 8): the GPU render
    (best of 3) beside the oracle on the same host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_lib as O
    from ratbite_amd import scenes
    from ratbite_amd.renderer import Renderer
    sc = scenes.cornell()
    r = Renderer(sc)

    def best_of():
        best = None
        for _ in range(4):
            r.ctx.synchronize()
            t = time.perf_counter()
            r.render(spp=sc.spp, offset=0)
            r.ctx.synchronize()
            el = time.perf_counter() - t
            best = el if best is None else min(best, el)
        return best

    best = best_of()
    t = r.tuning()
    r.set_tuning(path_slots=t["pathSlots"], trace_blocks=t["traceBlocksPerCU"], shade_blocks=t["shadeBlocksPerCU"],
                 use_graph=True, shade_waves=t["shadeWaves"], trace_fetch=t["traceFetch"],
                 trace_waves=t["traceWaves"], quant_nodes=t["quantNodes"], sort_hits=t["sortHits"])
    best_graph = best_of()
    lp = sc.launch_params(offset=0, count=sc.spp)
    _, st = O.render(sc, lp, threads=threads, oscene=O.OracleScene(sc, tree=r.light_tree()))
    r.close()
    n = sc.width * sc.height * sc.spp
    return {"workload": "cornell 256x256, 8 UV spheres, blackbody 4000 K emitter, 16 spp, depth 8", "samples": n,
            "gpu_ms": best * 1e3, "gpu_samples_per_s": n / best, "gpu_graph_ms": best_graph * 1e3,
            "gpu_graph_samples_per_s": n / best_graph, "cpu_seconds": st["renderSeconds"],
            "cpu_samples_per_s": n / st["renderSeconds"]}


def batch_workload(args):
    """64 seeded scenes of uneven size (triangles log-uniform in [1k, 500k],
    films from 64^2 to 1024^2, 1..5000 emissive triangles, the material mix of
    the spheres scene, 32 spp, depth 12) rendered in one batched launch
    sequence after autotuning the launch shape; plus 64 env-map CDF builds
    from 256x128 to 4096x2048. Reports aggregate samples/s and the p50/p99
    per-scene completion times."""
    import numpy as np
    from ratbite_amd import scenes
    from ratbite_amd.envmaps import lognormal_envmap
    from ratbite_amd.renderer import Renderer, render_batch
    t0 = time.time()
    rs = [Renderer(sc) for sc in scenes.batch(args.scenes)]
    print("batch: %d scenes built in %.1f s" % (len(rs), time.time() - t0), file=sys.stderr, flush=True)
    tuning = rs[0].autotune(trial_spp=2)
    print("autotuned launch shape %s" % tuning, file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        render_batch(rs)
    rs[0].ctx.synchronize()
    t1 = time.perf_counter()
    done = None
    for _ in range(args.steps):
        done = render_batch(rs)
    rs[0].ctx.synchronize()
    elapsed = time.perf_counter() - t1
    samples = sum(r.scene.width * r.scene.height * r.scene.spp for r in rs) * args.steps
    # env-map CDF builds alongside: 64 maps from 256x128 to 4096x2048
    ctx = rs[0].ctx
    cdf_ms = []
    for i in range(64):
        w = 256 << (i % 5)
        tex = lognormal_envmap(w, w // 2, sigma=2.0, seed=1000 + i, sun=False)
        ctx.envmap_build(tex)
        cdf_ms.append(ctx.stats()["cdfMs"])
    st = rs[0].ctx.stats()
    out = {
        "metric": "batched samples/s", "value": samples / elapsed, "unit": "samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": "%d seeded scenes, 32 spp, depth 12, batched" % len(rs),
                   "triangles": int(sum(len(r.scene.indices) for r in rs)),
                   "pixels": int(sum(r.scene.width * r.scene.height for r in rs))},
        "completion_ms_p50": float(np.percentile(done, 50)), "completion_ms_p99": float(np.percentile(done, 99)),
        "mrays_per_s": (st["closestRays"] + st["shadowRays"]) / (st["renderMs"] * 1e-3) / 1e6,
        "tuning": tuning,
        "cdf_builds": {"count": 64, "total_ms": float(sum(cdf_ms)), "max_ms": float(max(cdf_ms))},
    }
    print(json.dumps(out), flush=True)
    for r in rs:
        r.close()


FLOP_FORWARD = 2 * (80 * 64 + 3 * 64 * 64 + 64 * 40)   # 39,936 per query
FLOP_TRAIN = 3 * FLOP_FORWARD                           # forward + backward data + weight gradients
MFMA_PEAK_TFLOPS = {"bf16": 2500.0, "fp8": 5000.0}


def mlp_rates(r, chunk):
    """MFMA rates of the guiding network: 32 Adam steps on 2^18 synthetic
    records (staged once), and inference inside timed guided renders (BF16
    and FP8; one query per shaded vertex)."""
    import ctypes
    import numpy as np
    from ratbite_amd import _lib as L
    rng = np.random.default_rng(7)
    n, steps = 1 << 18, 32
    rec = np.zeros((n, L.GUIDE_REC), np.float32)
    rec[:, 0:3] = rng.uniform(0, 1, (n, 3))
    for a, b in ((3, 6), (6, 9), (11, 14)):
        v = rng.normal(size=(n, 3))
        rec[:, a:b] = v / np.linalg.norm(v, axis=1, keepdims=True)
    rec[:, 9] = rng.uniform(0.05, 1, n)
    rec[:, 10] = rng.uniform(0, 1, n)
    rec[:, 14] = rng.exponential(1.0, n)
    loss = ctypes.c_float()
    r.ctx.call("rbh_guide_train", L.ptr(rec), n, n, 1, ctypes.byref(loss))     # warm-up
    r.ctx.synchronize()
    t = time.perf_counter()
    r.ctx.call("rbh_guide_train", L.ptr(rec), n, n, steps, ctypes.byref(loss))
    train_s = time.perf_counter() - t
    out = {"train_step_ms": train_s * 1e3 / steps, "train_records_per_step": n,
           "train_tflops": steps * n * FLOP_TRAIN / train_s / 1e12}
    out["train_frac"] = out["train_tflops"] / MFMA_PEAK_TFLOPS["bf16"]
    for name, flag in (("bf16", 0), ("fp8", L.RBH_RENDER_GUIDE_FP8)):
        st = r.render(spp=chunk, offset=1 << 22, guided=True, no_drop=True, flags=flag | L.RBH_RENDER_TIMING)
        q = st["pathVertices"]
        tf = q * FLOP_FORWARD / (st["guideMs"] * 1e-3) / 1e12
        out["infer_%s" % name] = {"queries": q, "ms": st["guideMs"], "tflops": tf,
                                  "frac": tf / MFMA_PEAK_TFLOPS[name]}
    return out


def guided_study(r, budget_ms, ref_spp, chunk, train_iters, train_steps, train_batch, lr, fixed_selection):
    """Equal-time study of the learned guide on the renderer's scene: the
    plain tracer renders for the budget; the guided tracer spends the same
    wall time on collecting training records, training and rendering. The
    collecting renders are guided renders of the image's own sample range,
    so their samples count; only the Adam steps are overhead. relMSE (raw,
    and without the 0.1% largest per-pixel errors) against a plain render of
    ref_spp samples from sample 2^22 (FP64-accumulated) is taken at eight
    checkpoints of the budget with the clock paused, giving the final ratio
    and the time at which the guided tracer reaches the plain tracer's final
    error. Both sides use the no-drop sample semantics, so they converge to
    the same image. fixed_selection: one-sample MIS at selection 0.5 with the
    40-output network; otherwise the learned selection in [0.1, 0.9]."""
    import numpy as np

    def errs(a, ref, eps=1e-2):
        e = ((a - ref) ** 2 / (ref ** 2 + eps)).mean(axis=2).ravel()
        k = max(1, int(np.ceil(1e-3 * e.size)))
        return float(e.mean()), float(np.sort(e)[:-k].mean())

    def image():
        a = r.accum()
        return a[..., :3] / a[..., 3:4]

    t0 = time.perf_counter()
    done = 0
    while done < ref_spp:
        step = min(256, ref_spp - done)
        # an independent part of the sequence, below the 2^24 region where
        # the open fault was seen on other scenes
        r.render(spp=step, offset=(1 << 22) + done, no_drop=True, fresh=(done == 0))
        done += step
    ref = image()
    ref_s = time.perf_counter() - t0
    print("guided study: reference %d spp in %.1f s" % (ref_spp, ref_s), file=sys.stderr, flush=True)

    def run(method):
        checkpoints = [budget_ms * k / 8.0 for k in range(1, 9)]
        spent, spp, trained, curve = 0.0, 0, 0, []
        train_ms = 0.0
        r.ctx.synchronize()
        while checkpoints:
            t = time.perf_counter()
            if method == "guided" and trained < train_iters:
                # collecting renders are guided renders too (unbiased for any
                # state of the network), so their samples join the image
                r.render(spp=chunk, offset=spp, guided=True, collect=True)
                spp += chunk
                t1 = time.perf_counter()
                r.train_guide(batch=train_batch, steps=train_steps)
                trained += 1
                train_ms += (time.perf_counter() - t1) * 1e3
            else:
                r.render(spp=chunk, offset=spp, guided=(method == "guided"), no_drop=True)
                spp += chunk
            spent += (time.perf_counter() - t) * 1e3
            while checkpoints and spent >= checkpoints[0]:
                checkpoints.pop(0)
                curve.append((spent, spp) + (errs(image(), ref) if spp else (float("inf"), float("inf"))))
        return curve, train_ms

    plain, _ = run("plain")
    r.configure_guide(seed=1, lr=lr, capacity=1 << 22, fixed_selection=fixed_selection)
    guided, train_ms = run("guided")

    def reach(curve, target, idx):
        for c in curve:
            if c[idx] <= target:
                return c[0]
        return None

    fp, fg = plain[-1], guided[-1]
    t_raw, t_trim = reach(guided, fp[2], 2), reach(guided, fp[3], 3)
    return {
        "relmse_ratio": fg[2] / fp[2], "relmse_ratio_trimmed": fg[3] / fp[3],
        "time_to_plain_relmse": None if t_raw is None else t_raw / budget_ms,
        "time_to_plain_relmse_trimmed": None if t_trim is None else t_trim / budget_ms,
        "relmse_plain": fp[2], "relmse_guided": fg[2], "relmse_trimmed_plain": fp[3], "relmse_trimmed_guided": fg[3],
        "spp_plain": fp[1], "spp_guided": fg[1], "guided_training_ms": train_ms, "budget_ms": budget_ms,
        "reference_spp": ref_spp, "reference_s": ref_s,
        "selection": "fixed 0.5 (40 trained outputs)" if fixed_selection else "learned in [0.1, 0.9] (41 outputs)",
        "training": "%d x (%d spp guided collect + %d Adam steps of %d records), lr %g" % (
            train_iters, chunk, train_steps, train_batch, lr),
        "curve_plain": [[round(c[0], 1), c[1], c[2], c[3]] for c in plain],
        "curve_guided": [[round(c[0], 1), c[1], c[2], c[3]] for c in guided],
    }


def guided_workload(args):
    """The guided study as its own bench line (metric: relMSE ratio at equal
    time). --scene spheres is the 1080p workload's scene (at --width x
    --height)."""
    from ratbite_amd import scenes
    from ratbite_amd.renderer import Renderer

    if args.scene == "bounce":
        raise SystemExit("bench.py --workload guided --scene bounce is switched off: its plain no-drop reference "
                         "renders from sample 2^24 ended in an illegal memory access on the GPU whose cause has not "
                         "been found (a sanitized CPU build of the same kernel bodies replays such ranges cleanly)")
    sc = scenes.spheres(width=args.width, height=args.height, spp=16, depth=args.depth)
    r = Renderer(sc)
    g = guided_study(r, args.budget_ms, args.ref_spp or 8192, args.chunk_spp, args.train_iters, args.train_steps,
                     args.train_batch, args.train_lr, args.fixed_selection)
    mlp = mlp_rates(r, args.chunk_spp)
    out = {
        "metric": "relMSE at equal time, guided vs plain", "value": g["relmse_ratio"], "unit": "ratio",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": args.budget_ms, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16", "data": "synthetic (seeded scene generator)",
        "config": {"workload": "%s %dx%d, %d ms per method, reference %d spp plain, guide %s, %s"
                               % (args.scene, sc.width, sc.height, args.budget_ms, g["reference_spp"],
                                  g["selection"], g["training"])},
        "guided": g, "mlp": mlp,
    }
    print(json.dumps(out), flush=True)
    r.close()


def cdf_timings(steps):
    """Env-map CDF builds (k_envmap_rows + k_envmap_marginal) of seeded
    log-normal maps with a sun at 2048x1024 and 8192x4096, FP32 and FP64
    normalisation, each rebuilt `steps` times on a context of its own and
    timed with HIP events on its stream around both launches; GB/s from the
    algorithmic 14 B per RGB texel (12 B read, 2 B index written)."""
    from ratbite_amd import Context
    from ratbite_amd.envmaps import lognormal_envmap
    res = {}
    with Context(0) as ctx:
        for (w, h) in ((2048, 1024), (8192, 4096)):
            tex = lognormal_envmap(w, h, sigma=2.0, seed=3, sun=True)
            for fp64 in (False, True):
                ctx.envmap_build(tex, fp64=fp64)
                ctx.envmap_rebuild(fp64=fp64, iterations=steps)
                ms = ctx.stats()["cdfMs"]
                gbs = 14.0 * w * h / (ms * 1e-3) / 1e9
                key = "%dx%d_%s" % (w, h, "fp64" if fp64 else "fp32")
                res[key] = {"ms": ms, "GB_s": gbs, "frac": gbs / HBM_PEAK_GBS}
                print("cdf %s: %.4f ms, %.0f GB/s, %.1f%% of peak" % (key, ms, gbs, 100 * gbs / HBM_PEAK_GBS),
                      file=sys.stderr, flush=True)
    return res


def cdf_workload(args):
    """The env-map CDF build (k_envmap_rows: row CDFs and their inversion,
    then k_envmap_marginal) on seeded log-normal maps with a sun at the two
    sizes the configurations use, 2048x1024 and 8192x4096, FP32 and FP64
    normalisation; each build repeated --steps times after a warm-up, timed
    with HIP events on the build stream around both launches. Algorithmic
    bytes: 12 B of RGB texel read + 2 B of index written per texel."""
    res = cdf_timings(max(1, args.steps))
    top = res["2048x1024_fp32"]
    out = {"metric": "env-map CDF build GB/s", "value": top["GB_s"], "unit": "GB/s", "n_gpus": 1,
           "steps": args.steps, "warmup": 1, "ms_per_step": top["ms"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded log-normal env maps with a sun)",
           "config": {"workload": "env-map CDF build, 2048x1024 and 8192x4096 RGB FP32"},
           "roofline": {"bound": "hbm", "kernel": "k_envmap_rows + k_envmap_marginal", "achieved": top["GB_s"], "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": top["frac"], "traffic": None},
           "sizes": res}
    print(json.dumps(out), flush=True)


def large_workload(args):
    """The large precision-split scene (8.4M triangles: 800 displaced level-4
    icospheres over a 1449^2-vertex heightfield, 60% translucent or metal,
    100,000 textured emissive triangles, 8192x4096 log-normal env map with
    sigma 3) at 1920x1080, 256 spp, depth 24. The env-map CDF and the light
    tree's flux sums are built in their FP64 modes; pixels accumulate in FP64.
    Timed: one plain frame and one guided frame with FP8 E4M3 inference (after
    a short collect+train phase, not timed). Reported: samples/s and Mrays/s
    of both; MFMA rates of the guiding kernels; relMSE of both frames against
    a plain render of --ref-spp samples (an independent part of the sample
    sequence); the FP32 CDF build's deviation from the FP64 build."""
    raise SystemExit("bench.py --workload large is switched off: its plain reference renders ended in an "
                     "illegal memory access in the shade kernel whose cause has not been found yet")
    import numpy as np
    from ratbite_amd import _lib as L
    from ratbite_amd import scenes
    from ratbite_amd.renderer import Renderer

    ref_spp = args.ref_spp or 4096
    spp = 256 if args.spp == 64 else args.spp
    depth = 24 if args.depth == 16 else args.depth
    t0 = time.time()
    sc = scenes.large(width=args.width, height=args.height, spp=spp, depth=depth)
    print("scene: %d triangles, %d emissive, env %s (%.1f s)" % (len(sc.indices), len(sc.lights), sc.env.shape[:2],
                                                                 time.time() - t0), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    r = Renderer(sc, fp64_cdf=True, fp64_light_tree=True)
    ctx = r.ctx
    st0 = ctx.stats()
    print("uploads + builds %.1f s; LBVH %.1f ms depth %d; env CDF (FP64) %.3f ms"
          % (time.perf_counter() - t0, st0["bvhMs"], st0["bvhDepth"], st0["cdfMs"]), file=sys.stderr, flush=True)

    # FP32 CDF build of the same map against the FP64 build (both on the GPU)
    cond64, _, cdfy64, int64 = ctx.envmap_read()
    ctx.envmap_rebuild(fp64=False)
    cdf32_ms = ctx.stats()["cdfMs"]
    cond32, _, cdfy32, int32 = ctx.envmap_read()
    cdf_err = {"marginal_max_abs": float(np.max(np.abs(cdfy32.astype(np.float64) - cdfy64.astype(np.float64)))),
               "integral_rel": abs(int32 - int64) / int64,
               "conditional_index_equal_frac": float(np.mean(cond32 == cond64)),
               "conditional_index_max_diff": int(np.max(np.abs(cond32.astype(np.int64) - cond64.astype(np.int64)))),
               "fp64_ms": st0["cdfMs"], "fp32_ms": cdf32_ms}
    ctx.envmap_rebuild(fp64=True)
    print("CDF FP32 vs FP64: %s" % cdf_err, file=sys.stderr, flush=True)

    def image():
        a = r.accum()
        return a[..., :3] / a[..., 3:4]

    def errs(a, ref, eps=1e-2):
        e = ((a - ref) ** 2 / (ref ** 2 + eps)).mean(axis=2).ravel()
        k = max(1, int(np.ceil(1e-3 * e.size)))
        return float(e.mean()), float(np.sort(e)[:-k].mean())

    t0 = time.perf_counter()
    done = 0
    while done < ref_spp:
        step = min(256, ref_spp - done)
        r.render(spp=step, offset=1 << 24 | done, no_drop=True, fresh=(done == 0))
        done += step
        print("reference: %d / %d spp" % (done, ref_spp), file=sys.stderr, flush=True)
    ref = image()
    ref_s = time.perf_counter() - t0

    def frame(guided, flags):
        ctx.synchronize()
        t = time.perf_counter()
        st = r.render(spp=spp, offset=0, guided=guided, no_drop=True, flags=flags | L.RBH_RENDER_TIMING, fresh=True)
        ctx.synchronize()
        el = time.perf_counter() - t
        return el, st, errs(image(), ref)

    r.render(spp=4, offset=0, no_drop=True)           # warm-up
    el_p, st_p, e_p = frame(False, 0)
    print("plain: %.1f ms, relMSE %.4g (trimmed %.4g)" % (el_p * 1e3, e_p[0], e_p[1]), file=sys.stderr, flush=True)

    r.configure_guide(seed=1, lr=args.train_lr, capacity=1 << 22)
    t0 = time.perf_counter()
    for it in range(args.train_iters):
        r.render(spp=args.chunk_spp, offset=(1 << 23) + it * args.chunk_spp, guided=True, collect=True)
        r.train_guide(batch=args.train_batch, steps=args.train_steps)
    ctx.synchronize()
    train_s = time.perf_counter() - t0
    el_g, st_g, e_g = frame(True, L.RBH_RENDER_GUIDE_FP8)
    print("guided FP8: %.1f ms, relMSE %.4g (trimmed %.4g); training %.1f s" % (el_g * 1e3, e_g[0], e_g[1], train_s),
          file=sys.stderr, flush=True)
    mlp = mlp_rates(r, args.chunk_spp)
    q = st_g["pathVertices"]
    fp8_tf = q * FLOP_FORWARD / (st_g["guideMs"] * 1e-3) / 1e12 if st_g["guideMs"] > 0 else None
    samples = sc.width * sc.height * spp
    out = {
        "metric": "1080p samples/s (large precision-split scene)", "value": samples / el_p, "unit": "samples/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": el_p * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 CDF, light flux sums, accumulation; fp8 guide)",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": "large %dx%d: 800 level-4 icospheres + 1449^2 heightfield (%d triangles), %d textured "
                               "emissive triangles, env 8192x4096 log-normal sigma 3, %d spp, depth %d"
                               % (sc.width, sc.height, len(sc.indices), len(sc.lights), spp, depth),
                   "width": sc.width, "height": sc.height, "spp": spp, "depth": depth},
        "mrays_per_s": (st_p["closestRays"] + st_p["shadowRays"]) / el_p / 1e6,
        "rays": {"closest": st_p["closestRays"], "shadow": st_p["shadowRays"], "vertices": st_p["pathVertices"]},
        "kernel_ms": {"k_trace": st_p["traceMs"], "k_shade": st_p["shadeMs"], "k_shadow": st_p["shadowMs"]},
        "guided_fp8": {"samples_per_s": samples / el_g, "ms": el_g * 1e3,
                       "mrays_per_s": (st_g["closestRays"] + st_g["shadowRays"]) / el_g / 1e6,
                       "guide_infer_ms": st_g["guideMs"], "guide_infer_tflops": fp8_tf,
                       "guide_infer_frac": None if fp8_tf is None else fp8_tf / MFMA_PEAK_TFLOPS["fp8"],
                       "training_s": train_s},
        "relmse_vs_reference": {"plain": e_p[0], "plain_trimmed": e_p[1], "guided_fp8": e_g[0],
                                "guided_fp8_trimmed": e_g[1], "reference_spp": ref_spp,
                                "reference_s": ref_s},
        "cdf_fp32_vs_fp64": cdf_err,
        "build": {"bvh_ms": st0["bvhMs"], "bvh_depth": st0["bvhDepth"], "light_tree": "fp64 power mode"},
        "mlp": mlp,
        "nonfinite_samples": st_p["nonFiniteSamples"] + st_g["nonFiniteSamples"],
    }
    print(json.dumps(out), flush=True)
    r.close()


def graph_rate(r, ctx, sc, args, eager_rate):
    """The same frames rendered with each pass captured as a HIP graph and
    replayed (no per-kernel timing), timed on the host around whole frames."""
    t = r.tuning()
    r.set_tuning(path_slots=t["pathSlots"], trace_blocks=t["traceBlocksPerCU"], shade_blocks=t["shadeBlocksPerCU"],
                 use_graph=True, shade_waves=t["shadeWaves"], trace_fetch=t["traceFetch"],
                 trace_waves=t["traceWaves"], quant_nodes=t["quantNodes"], sort_hits=t["sortHits"])
    params = sc.launch_params(offset=0, count=args.spp)
    ctx.render(params)
    ctx.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.graph_steps):
        ctx.render(params)
    ctx.synchronize()
    el = time.perf_counter() - t1
    r.set_tuning(path_slots=t["pathSlots"], trace_blocks=t["traceBlocksPerCU"], shade_blocks=t["shadeBlocksPerCU"],
                 use_graph=bool(t["useGraph"]), shade_waves=t["shadeWaves"], trace_fetch=t["traceFetch"],
                 trace_waves=t["traceWaves"], quant_nodes=t["quantNodes"], sort_hits=t["sortHits"])
    rate = sc.width * sc.height * args.spp * args.graph_steps / el
    print("graph replay: %.1f ms per frame" % (el * 1e3 / args.graph_steps), file=sys.stderr, flush=True)
    return {"samples_per_s": rate, "ms_per_step": el * 1e3 / args.graph_steps, "steps": args.graph_steps,
            "over_eager": rate / eager_rate}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        if int(os.environ.get("RANK", "0")) == 0:
            print("bench.py measures one GPU: the path tracer runs on a single MI355X here, "
                  "so --gpus %d (world size %d) is not supported" % (args.gpus, world), file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        return batch_workload(args)
    if args.workload == "guided":
        return guided_workload(args)
    if args.workload == "large":
        return large_workload(args)
    if args.workload == "cdf":
        return cdf_workload(args)

    from ratbite_amd import _lib as L
    from ratbite_amd import scenes
    from ratbite_amd.renderer import Renderer

    t0 = time.time()
    sc = scenes.spheres(width=args.width, height=args.height, spp=args.spp, depth=args.depth)
    print("scene: %d triangles, %d emissive, env %s (%.1f s)" % (len(sc.indices), len(sc.lights),
          None if sc.env is None else sc.env.shape[:2], time.time() - t0), file=sys.stderr, flush=True)
    r = Renderer(sc)
    ctx = r.ctx
    st0 = ctx.stats()
    print("LBVH %.1f ms depth %d; env CDF %.3f ms" % (st0["bvhMs"], st0["bvhDepth"], st0["cdfMs"]),
          file=sys.stderr, flush=True)
    tuning = r.autotune(trial_spp=8, depth=args.depth) if args.tune else r.tuning()
    print("launch shape (%s) %s" % ("autotuned" if args.tune else "library default", tuning), file=sys.stderr,
          flush=True)
    copy_gbs = None
    if args.copy_mib > 0:
        import ctypes
        ms = ctypes.c_float()
        nbytes = args.copy_mib << 20
        ctx.check(L.lib().rbh_test_copy(ctx.h, nbytes, 5, ctypes.byref(ms)))
        copy_gbs = 2.0 * nbytes / (ms.value * 1e-3) / 1e9
        print("device copy %d MiB: %.3f ms, %.0f GB/s read+write" % (args.copy_mib, ms.value, copy_gbs),
              file=sys.stderr, flush=True)

    params = sc.launch_params(offset=0, count=args.spp, flags=L.RBH_RENDER_TIMING)
    for _ in range(args.warmup):
        ctx.render(params)
    ctx.synchronize()

    tot = {k: 0 for k in ("closestRays", "shadowRays", "pathVertices", "traceMs", "shadeMs", "shadowMs",
                          "renderMs", "traceLaunches")}
    ctx.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        ctx.render(params)
        s = ctx.stats()
        for k in tot:
            tot[k] += s[k]
        print("step: %.1f ms, %d closest + %d shadow rays" % (s["renderMs"], s["closestRays"], s["shadowRays"]),
              file=sys.stderr, flush=True)
    ctx.synchronize()
    elapsed = time.perf_counter() - t1

    pixels = sc.width * sc.height
    samples = pixels * args.spp * args.steps
    rays = tot["closestRays"] + tot["shadowRays"]
    kms = {"trace": tot["traceMs"], "shade": tot["shadeMs"], "shadow": tot["shadowMs"]}
    kbytes = {"trace": BYTES_CLOSEST * tot["closestRays"], "shade": BYTES_VERTEX * tot["pathVertices"],
              "shadow": BYTES_SHADOW * tot["shadowRays"]}
    dom = max(kms, key=kms.get)
    launches = tot["traceLaunches"]
    achieved = kbytes[dom] / (kms[dom] * 1e-3) / 1e9
    traffic = traffic_per_launch(KERNEL_NAMES[dom])
    # every traced ray is a path vertex (misses are shaded by the miss kernel);
    # the FP64 accumulation is read and written once per pixel per pass, and a
    # pass traces `depth` bounces
    passes = tot["traceLaunches"] // max(1, args.depth)
    path_bytes = (BYTES_CLOSEST + BYTES_SHADOW + BYTES_VERTEX) * tot["closestRays"] + \
        BYTES_PIXEL_LAUNCH * pixels * passes
    graph = graph_rate(r, ctx, sc, args, samples / elapsed) if args.graph_steps > 0 else None

    out = {
        "metric": "1080p samples/s", "value": samples / elapsed, "unit": "samples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generator)",
        "config": {"workload": "spheres %dx%d: 200 level-5 icospheres (%d triangles), %d emissive triangles "
                               "sharing 16 seeded 256x256 RGBA8 value-noise emission textures, env 2048x1024 "
                               "log-normal + sun, %d spp, depth %d" % (sc.width, sc.height, len(sc.indices),
                                                                       len(sc.lights), args.spp, args.depth),
                   "width": sc.width, "height": sc.height, "spp": args.spp, "depth": args.depth},
        "mrays_per_s": rays / elapsed / 1e6,
        "rays": {"closest": tot["closestRays"], "shadow": tot["shadowRays"], "vertices": tot["pathVertices"]},
        "kernel_ms": {KERNEL_NAMES[k]: v for k, v in kms.items()},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "avg_launch_ms": kms[dom] / max(1, launches),
                     "algorithmic_bytes_per_launch": kbytes[dom] / max(1, launches),
                     "traffic_gbs": None if traffic is None else traffic / (kms[dom] / max(1, launches) * 1e-3) / 1e9,
                     "copy_gbs_measured": copy_gbs},
        "path_roofline": {"achieved": path_bytes / elapsed / 1e9, "unit": "GB/s",
                          "frac": path_bytes / elapsed / 1e9 / HBM_PEAK_GBS, "passes": passes},
        "graph": graph,
        "cpu_baseline": None,
        "tuning": tuning,
    }
    if args.guided_budget_ms > 0:
        out["guided"] = guided_study(r, args.guided_budget_ms, args.guided_ref_spp, args.chunk_spp,
                                     args.train_iters, args.train_steps, args.train_batch, args.train_lr,
                                     bool(args.fixed_selection))
        out["guided"]["workload"] = "the same scene and frame size, plain vs guided, no-drop sample semantics"
    # the env-map CDF build at the configurations' map sizes (its own
    # context; the frame above is unaffected)
    out["envmap_cdf"] = {"kernels": "k_envmap_rows + k_envmap_marginal", "bytes_per_texel": 14,
                         "sizes": cdf_timings(20)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, args, r.light_tree())
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    r.close()


if __name__ == "__main__":
    main()
