#!/usr/bin/env python3
"""Headline benchmark: direct N^2 FP32 gravity with Velocity Verlet on one MI355X.

Workload: uniform-density sphere of radius 1, N = 262,144 equal masses 1/N,
Gaussian velocities sigma = 0.05 per axis, G = 1, eps = 0.005, dt = 5e-4
(synthetic, seed 42). A step is one Velocity Verlet step
(store+drift kernel, then the force kernel with the kick in its epilogue),
all on one HIP stream with the inputs resident in HBM.

Prints one JSON line with the whole-job rate in Ginteractions/s, the roofline
of the dominant kernel (the direct force kernel, timed per launch with HIP
events on its stream; 20 flop per interaction against the 157.3 TFLOP/s FP32
vector peak), the FP64-oracle force error on sample targets, and the CPU
baseline (the oracle's restatement of the reference CPU force path, timed on
this host on a bounded sample).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nbody-gpu-sim_amd"))

METRIC = "direct N² Ginteractions/s; BH ms/step @1M; batched-ensemble speedup @fixed err"
FLOP_PER_INTERACTION = 20.0
PEAK_FP32_TFLOPS = 157.3
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "direct_forces_uniform_sphere_n262144_traffic.md")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profile():
    """HBM bytes per launch of the force kernel, from the committed rocprofv3 --pmc summary."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def cpu_baseline(h, G, eps, n_targets):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc

    n = len(h["mass"])
    t = np.sort(np.random.default_rng(7).choice(n, n_targets, replace=False))
    threads = orc.max_threads()
    t0 = time.perf_counter()
    orc.direct_f32(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], G, eps, targets=t,
                   nthreads=threads)
    sec = time.perf_counter() - t0
    rate = n_targets * (n - 1) / sec / 1e9
    return {"value": rate, "unit": "Ginteractions/s", "cores": threads, "kind": "port",
            "sample": f"{n_targets} of {n} targets x all {n} sources, one FP32 force "
                      f"evaluation (reference CPU pair function summed in j order), "
                      f"{sec:.2f} s wall"}


def force_error(nbx, d, h_now, G, eps, n_targets=256):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc

    n = len(h_now["mass"])
    t = np.sort(np.random.default_rng(42).choice(n, min(n, n_targets), replace=False))
    ref = orc.direct_f64(h_now["pos_x"], h_now["pos_y"], h_now["pos_z"], h_now["mass"], G,
                         orc.eps2_of(eps), targets=t)
    got = d.acc()[t].astype(np.float64)
    e = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return {"targets": int(len(t)), "median": float(np.median(e)),
            "p99": float(np.quantile(e, 0.99)), "max": float(e.max())}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--variant", type=int, default=-1, help="force a direct tile variant")
    ap.add_argument("--autotune", action="store_true", help="time all tile variants first")
    ap.add_argument("--cpu-targets", type=int, default=16384)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this engine runs on a single MI355X "
              "(no multi-GPU path is built); run with --gpus 1", file=sys.stderr)
        return 2
    import nbx

    n, G, eps, dt = args.n, 1.0, 0.005, 5e-4
    eps2 = float(np.float32(np.float32(eps) * np.float32(eps)))
    info = nbx.device_info()
    log(f"device {info['name']} ({info['cu']} CUs, {info['clock_khz'] / 1e3:.0f} MHz)")

    if args.autotune:
        best, ms = nbx.C.c_int(), nbx.C.c_double()
        nbx.check(nbx.lib.nbx_direct_autotune(n, 3, nbx.C.byref(best), nbx.C.byref(ms)))
        log(f"autotune: best variant {nbx.variant_names()[best.value]} {ms.value:.3f} ms")
    if args.variant >= 0:
        nbx.check(nbx.lib.nbx_direct_set_variant(args.variant))

    h = nbx.gen_uniform_sphere(n, seed=42, R=1.0, M=1.0, sigma=0.05)
    d = nbx.DeviceParticles.from_host(h)
    s = nbx.Stream()
    lib = nbx.lib
    nbx.check(lib.nbx_direct_forces(d.ptr, G, eps2, 256, s.h))  # a(0)

    def step(ev_a=None, ev_b=None):
        nbx.check(lib.nbx_verlet_drift(d.ptr, dt, s.h))
        if ev_a is not None:
            ev_a.record(s)
        nbx.check(lib.nbx_direct_forces_kick(d.ptr, G, eps2, dt, s.h))
        if ev_b is not None:
            ev_b.record(s)

    for _ in range(args.warmup):
        step()
    s.synchronize()
    log(f"warmup done ({args.warmup} steps)")

    K = args.steps
    fa = [nbx.Event() for _ in range(K)]
    fb = [nbx.Event() for _ in range(K)]
    e0, e1 = nbx.Event(), nbx.Event()
    s.synchronize()
    wall0 = time.perf_counter()
    e0.record(s)
    for k in range(K):
        step(fa[k], fb[k])
    e1.record(s)
    s.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = e0.elapsed_ms(e1)
    force_ms = np.array([fa[k].elapsed_ms(fb[k]) for k in range(K)])
    log(f"timed {K} steps: {total_ms:.3f} ms (wall {wall * 1e3:.3f} ms), force kernel "
        f"mean {force_ms.mean():.3f} ms min {force_ms.min():.3f} ms")

    inter = float(n) * float(n - 1)
    value = K * inter / (total_ms / 1e3) / 1e9
    achieved = FLOP_PER_INTERACTION * inter / (force_ms.mean() / 1e3) / 1e12
    traffic = traffic_from_profile()

    h_now = d.download()
    err = force_error(nbx, d, h_now, G, eps)
    log(f"force error vs FP64 on {err['targets']} targets: median {err['median']:.2e} "
        f"p99 {err['p99']:.2e} max {err['max']:.2e}")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(h, G, eps, args.cpu_targets)
        log(f"cpu baseline: {cpu['value']:.4f} Ginteractions/s on {cpu['cores']} threads")

    names = nbx.variant_names()
    variant = args.variant
    if variant < 0:
        # the variant the dispatcher picked for this size
        variant = None
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "Ginteractions/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": total_ms / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "direct N^2 FP32 + Velocity Verlet, uniform sphere R=1, "
                               "N=262144, m=1/N, sigma_v=0.05, eps=0.005, dt=5e-4",
                   "n": n, "eps": eps, "dt": dt, "G": G, "parallelism": "single GPU",
                   "tile_variant": names[variant] if variant is not None else "auto"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic,
                     "kernel": "direct_forces_kernel (forces + kick epilogue)",
                     "flop_per_interaction": FLOP_PER_INTERACTION,
                     "interactions_per_launch": inter,
                     "launch_ms_mean": float(force_ms.mean())},
        "force_error_fp64": err,
        "cpu_baseline": cpu,
        "device": info["name"],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
