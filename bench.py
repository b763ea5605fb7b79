#!/usr/bin/env python3
"""Benchmarks of the N-body engine on one MI355X.

Headline (the JSON line's `value`): direct N^2 FP32 gravity with Velocity
Verlet on a uniform-density sphere of radius 1, N = 262,144 equal masses 1/N,
Gaussian velocities sigma = 0.05 per axis, G = 1, eps = 0.005, dt = 5e-4
(synthetic, seed 42). A step is one Velocity Verlet step (store+drift kernel,
then the force kernel with the kick in its epilogue) on one HIP stream, with
the inputs resident in HBM. Reported as whole-job interactions/s, with the
roofline of the force kernel (timed per launch with HIP events on its stream;
20 flop per interaction against the 157.3 TFLOP/s FP32 peak), its HBM traffic
from the committed rocprofv3 counter summary, the FP64-oracle force error on
sample targets, and the CPU baseline (the oracle's restatement of the
reference CPU force path, timed on this host on a bounded sample).

The same line carries the other two parts of the metric:
  barnes_hut - ms/step at N = 1,048,576 (two Plummer spheres, theta 0.5) with
               per-phase times and achieved GB/s of the HBM-bound phases;
  ensemble   - 512 Plummer systems, N_s log-uniform in [256, 16384]: batched
               (persistent launch, hipGraph chunks of 100 steps) vs the same
               steps advancing each system alone; per-system forces are
               bit-identical to the standalone kernel.

It also carries spatial_hash: the short-range operator on the reference's own
force.spatial_hash workload (1,048,576 bodies uniform in a sphere of radius 10,
unit masses, cell = cutoff = 2, eps = 0.1), ms per computeForces call (grid
build + forces) and the roofline of the force kernel.

cpu_path_n4096 is the reference's CPU-runnable case (4,096-body Plummer, a(0) +
100 steps) timed on the GPU and on the oracle's restatement of the reference CPU
path, with the final positions compared. Every CPU baseline is reported on all
host threads and on one.

--only {direct,barnes-hut,ensemble,precision,spatial-hash,cpu-path} runs one part
(for profiling); precision compares plain FP32 with the double-single/Kahan form
at N=65,536.
"""
import argparse
import itertools
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
PEAK_HBM_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "direct_forces_uniform_sphere_n262144_traffic.md")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    return orc


def eps2_of(eps):
    return float(np.float32(np.float32(eps) * np.float32(eps)))


def host_info(orc) -> dict:
    """CPU of the host the baselines ran on."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "threads_used": orc.max_threads()}


def traffic_from_profile():
    """HBM bytes per launch of the force kernel, from the committed rocprofv3 summary."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


# ------------------------------------------------------------------ direct

def run_direct(nbx, args):
    n, G, eps, dt = args.n, 1.0, 0.005, 5e-4
    e2 = eps2_of(eps)
    if args.calibrate_copy:
        C = nbx.C
        nb = 256 << 20
        a, b = C.c_void_p(), C.c_void_p()
        nbx.check(nbx.lib.nbx_malloc(C.byref(a), nb))
        nbx.check(nbx.lib.nbx_malloc(C.byref(b), nb))
        nbx.check(nbx.lib.nbx_memcpy_dtod(b, a, nb))
        nbx.synchronize()
        nbx.check(nbx.lib.nbx_free(a))
        nbx.check(nbx.lib.nbx_free(b))
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
    ws = nbx.DirectWorkspace(n)  # DirectForceCalculator's workspace, allocated before timing
    plan = ws.plan(n)
    nbx.check(lib.nbx_direct_compute(ws.h, d.ptr, G, e2, s.h))  # a(0)

    def step(ev_a=None, ev_b=None):
        nbx.check(lib.nbx_verlet_drift(d.ptr, dt, s.h))
        if ev_a is not None:
            ev_a.record(s)
        nbx.check(lib.nbx_direct_compute_kick(ws.h, d.ptr, G, e2, dt, s.h))
        if ev_b is not None:
            ev_b.record(s)

    for _ in range(args.warmup):
        step()
    s.synchronize()
    K = args.steps
    fa = [nbx.Event() for _ in range(K)]
    fb = [nbx.Event() for _ in range(K)]
    e0, e1 = nbx.Event(), nbx.Event()
    wall0 = time.perf_counter()
    e0.record(s)
    for k in range(K):
        step(fa[k], fb[k])
    e1.record(s)
    s.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = e0.elapsed_ms(e1)
    force_ms = np.array([fa[k].elapsed_ms(fb[k]) for k in range(K)])
    log(f"direct: {K} steps {total_ms:.3f} ms (wall {wall * 1e3:.3f} ms), force kernel mean "
        f"{force_ms.mean():.3f} ms")
    inter = float(n) * float(n - 1)
    value = K * inter / (total_ms / 1e3) / 1e9
    achieved = FLOP_PER_INTERACTION * inter / (force_ms.mean() / 1e3) / 1e12

    err = None
    if not args.no_check:
        orc = oracle()
        hn = d.download()
        t = np.sort(np.random.default_rng(42).choice(n, min(n, args.check_targets), replace=False))
        ref = orc.direct_f64(hn["pos_x"], hn["pos_y"], hn["pos_z"], hn["mass"], G, e2, targets=t)
        got = d.acc()[t].astype(np.float64)
        e = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
        err = {"targets": int(len(t)), "median": float(np.median(e)),
               "p99": float(np.quantile(e, 0.99)), "max": float(e.max())}
        log(f"direct force error vs FP64: {err}")

    cpu = None
    if not args.no_cpu_baseline:
        orc = oracle()
        t = np.sort(np.random.default_rng(7).choice(n, args.cpu_targets, replace=False))
        threads = orc.max_threads()
        t0 = time.perf_counter()
        orc.direct_f32(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], G, eps, targets=t,
                       nthreads=threads)
        sec = time.perf_counter() - t0
        t1 = t[:args.cpu1_targets]
        t0 = time.perf_counter()
        orc.direct_f32(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], G, eps, targets=t1,
                       nthreads=1)
        sec1 = time.perf_counter() - t0
        cpu = {"value": args.cpu_targets * (n - 1) / sec / 1e9, "unit": "Ginteractions/s",
               "cores": threads, "kind": "port",
               "sample": f"{args.cpu_targets} of {n} targets x all {n} sources: one FP32 force "
                         f"evaluation with the reference CPU pair function summed in j order "
                         f"(oracle), {sec:.2f} s wall on {threads} threads",
               "single_thread": {"value": len(t1) * (n - 1) / sec1 / 1e9,
                                 "unit": "Ginteractions/s",
                                 "sample": f"{len(t1)} targets, {sec1:.2f} s on 1 thread"},
               "host": host_info(orc)}
        log(f"direct cpu baseline: {cpu['value']:.3f} Ginteractions/s on {threads} threads, "
            f"{cpu['single_thread']['value']:.3f} on 1")
    names = nbx.variant_names()
    v = plan["variant"]
    sym = names[v].startswith("sym_")
    return {
        "value": value, "ms_per_step": total_ms / K,
        "config": {"workload": "direct N^2 FP32 + Velocity Verlet, uniform sphere R=1, "
                               f"N={n}, m=1/N, sigma_v=0.05, eps=0.005, dt=5e-4",
                   "n": n, "eps": eps, "dt": dt, "G": G, "parallelism": "single GPU",
                   "tile_variant": names[v], "rounds": plan["rounds"],
                   "partial_scratch_bytes": plan["partial_bytes"]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic_from_profile(),
                     "kernel": ("direct_sym_kernel + sym_reduce_kernel (forces + kick)" if sym
                                else "direct_forces_kernel (forces + kick epilogue)"),
                     "evaluation": ("pair-symmetric: each of the N(N-1)/2 pairs evaluated once "
                                    "and applied to both bodies; achieved counts the N(N-1) "
                                    "interactions of the reference's algorithm" if sym else
                                    "one-sided: N(N-1) pair evaluations"),
                     "flop_per_interaction": FLOP_PER_INTERACTION,
                     "interactions_per_launch": inter,
                     "pair_evaluations_per_launch": inter / 2 if sym else inter,
                     "launch_ms_mean": float(force_ms.mean())},
        "force_error_fp64": err,
        "cpu_baseline": cpu,
    }


# ------------------------------------------------------------------ CPU-path config

def run_cpu_path(nbx, args):
    """The reference's CPU-runnable case: direct N^2 FP32 on a 4,096-body Plummer sphere
    (G = M = a = 1, m = 1/N, seed 42), eps 0.01, dt 1e-3, a(0) + 100 Verlet steps, on the
    GPU and on the oracle's restatement of the reference CPU path (all threads and 1)."""
    n, G, eps, dt, steps = 4096, 1.0, 0.01, 1e-3, 100
    e2 = eps2_of(eps)
    h = nbx.gen_plummer(n, seed=42)
    d = nbx.DeviceParticles.from_host(h)
    s = nbx.Stream()
    e0, e1 = nbx.Event(), nbx.Event()
    e0.record(s)
    nbx.check(nbx.lib.nbx_direct_forces(d.ptr, G, e2, 256, s.h))
    for _ in range(steps):
        nbx.check(nbx.lib.nbx_direct_step(d.ptr, G, e2, dt, s.h))
    e1.record(s)
    s.synchronize()
    gpu_ms = e0.elapsed_ms(e1)
    inter = (steps + 1) * float(n) * (n - 1)
    got = d.download()
    out = {"n": n, "steps": steps, "eps": eps, "dt": dt,
           "workload": "Plummer a=1, M=1, m=1/N, seed 42; a(0) + 100 Velocity Verlet steps",
           "gpu_ms": gpu_ms, "gpu_Ginteractions_per_s": inter / (gpu_ms * 1e-3) / 1e9}
    orc = oracle()
    ref = {k: v.copy() for k, v in h.items()}
    threads = orc.max_threads()
    t0 = time.perf_counter()
    orc.run_direct(ref, G, eps, dt, steps, nthreads=threads)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    one = {k: v.copy() for k, v in h.items()}
    t0 = time.perf_counter()
    orc.run_direct(one, G, eps, dt, steps, nthreads=1)
    cpu1_ms = (time.perf_counter() - t0) * 1e3
    pos_diff = max(float(np.abs(got[k] - ref[k]).max()) for k in ("pos_x", "pos_y", "pos_z"))
    eg, er = orc.energy_f64(got, G, e2), orc.energy_f64(ref, G, e2)
    e_gpu, e_ref = eg["ke"] + eg["pe"], er["ke"] + er["pe"]
    out.update({"cpu_baseline": {"value": cpu_ms, "unit": "ms/run", "cores": threads,
                                 "kind": "port",
                                 "sample": "the full run with the oracle's restatement of the "
                                           "reference CPU path (pair function summed in j order, "
                                           "Verlet as integrator.cu)",
                                 "single_thread": {"value": cpu1_ms, "unit": "ms/run"},
                                 "host": host_info(orc)},
                "max_position_diff_vs_cpu": pos_diff,
                "energy_rel_diff_vs_cpu": abs(e_gpu - e_ref) / abs(e_ref)})
    log(f"cpu path N={n}: GPU {gpu_ms:.2f} ms, CPU {cpu_ms:.0f} ms on {threads} threads, "
        f"{cpu1_ms:.0f} ms on 1; max |dx| {pos_diff:.2e}")
    return out


# ------------------------------------------------------------------ Barnes-Hut

# algorithmic bytes per particle of the HBM-bound build phases:
#   bbox    reads x, y, z
#   morton  reads x, y, z, m; writes the Morton code, the index, the Hilbert
#           index and the packed body (x, y, z, m)
#   sort    stable radix sort of (code, index): 4 passes of 8 bits, each reading
#           the keys to count, then reading and writing keys and values
#   paths   reads the sorted index and the packed body at it; writes the body in
#           sorted order, the octant path (16 B), both path-sort keys and the rank
#   export  the traversal's task order: gathers an order index and a Hilbert
#           index, writes both, and sorts them as above
BH_BYTES = {"bbox": 12.0, "morton": 16.0 + 28.0, "sort": 4 * 20.0,
            "paths": 4.0 + 16.0 + 16.0 + 16.0 + 12.0, "export": 16.0 + 4 * 20.0}
# bytes of one traversal record (centre of mass and G m, centre and half size,
# skip and thresholds, mass); the walk visits one per step
BH_RECORD_BYTES = 64


def run_bh(nbx, args):
    n, G, eps, theta, dt = args.bh_n, 1.0, 0.01, 0.5, 1e-3
    h = nbx.gen_two_plummer(n, seed=42)
    d = nbx.DeviceParticles.from_host(h)
    t = nbx.BHTree(n)
    s = nbx.Stream()
    t.build(d, s)
    t.forces(d, theta, G, eps, s)

    def step():
        nbx.check(nbx.lib.nbx_verlet_drift(d.ptr, dt, s.h))
        t.build(d, s)
        t.forces(d, theta, G, eps, s)
        nbx.check(nbx.lib.nbx_update_velocities(d.ptr, dt, s.h))

    step()
    s.synchronize()
    K = args.bh_steps
    e0, e1 = nbx.Event(), nbx.Event()
    e0.record(s)
    for _ in range(K):
        step()
    e1.record(s)
    s.synchronize()
    ms = e0.elapsed_ms(e1) / K
    # one more step with phase events for the breakdown
    t.set_timing(True)
    step()
    s.synchronize()
    phases = t.phase_ms()
    t.set_timing(False)
    nodes, depth = t.node_count()
    gbs = {k: BH_BYTES[k] * n / (phases[k] * 1e-3) / 1e9 for k in BH_BYTES if phases[k] > 0}
    # the traversal alone on the current tree: HIP events around R launches on s,
    # and the walk's own counts (nbx_bh_walk_census: the same tasks and decisions)
    R = 5
    f0, f1 = nbx.Event(), nbx.Event()
    f0.record(s)
    for _ in range(R):
        t.forces(d, theta, G, eps, s)
    f1.record(s)
    s.synchronize()
    trav_ms = f0.elapsed_ms(f1) / R
    census = t.walk_census(d, theta, G, eps)
    rec_bytes = census["wave_steps"] * BH_RECORD_BYTES
    flops = census["interactions"] * FLOP_PER_INTERACTION
    trav_gbs = rec_bytes / (trav_ms * 1e-3) / 1e9
    trav_tf = flops / (trav_ms * 1e-3) / 1e12
    waves = -(-n // 64)
    roofline = {
        "kernel": "bh_traverse", "bound": "hbm", "achieved": trav_gbs, "peak": PEAK_HBM_GBS,
        "unit": "GB/s", "frac": trav_gbs / PEAK_HBM_GBS, "traffic": None,
        "launch_ms_mean": trav_ms,
        "bytes_per_launch": rec_bytes,
        "bytes_rule": f"{BH_RECORD_BYTES} B per record a wave's walk visits (one per step)",
        "fp32": {"achieved_TFLOPs": trav_tf, "peak_TFLOPs": PEAK_FP32_TFLOPS,
                 "frac": trav_tf / PEAK_FP32_TFLOPS,
                 "rule": f"{FLOP_PER_INTERACTION:.0f} flop per particle-node force evaluation"},
        "census": census, "waves": waves,
        "steps_per_wave": census["wave_steps"] / waves,
        "lane_utilisation": census["active_lane_steps"] / (64.0 * census["wave_steps"]),
        "note": "neither roof binds: each wave step issues about 22 VALU and 14 SALU "
                "instructions for 64 lanes, of which about 65 % are active, and needs the record "
                "the previous step chose; wave wait-state counters put 44 % of the waves' cycles "
                "on those loads and 32 % on dependencies between their own instructions"}
    log(f"bh_traverse: {trav_ms:.3f} ms, {census['wave_steps']} wave-steps "
        f"({census['wave_steps'] / waves:.0f} per wave), {census['interactions']} interactions; "
        f"{trav_gbs:.0f} GB/s of records, {trav_tf:.2f} TFLOP/s")
    log(f"barnes-hut N={n}: {ms:.3f} ms/step, phases {phases}, nodes {nodes}, depth {depth}")
    out = {"n": n, "theta": theta, "eps": eps, "dt": dt, "steps": K, "ms_per_step": ms,
           "workload": "two Plummer spheres (a=1, M=1 each, truncated at 99.9% mass) at x=+-5, "
                       "relative velocity 0.5 along x",
           "phases_ms": phases, "nodes": nodes, "max_depth": depth,
           "builder": t.build_info()["builder"],
           "achieved_GBps": gbs, "build_bytes_per_particle": BH_BYTES,
           "hbm_peak_GBps": PEAK_HBM_GBS, "roofline": roofline}
    if not args.no_cpu_baseline:
        orc = oracle()
        hn = d.download()
        x, y, z, m = hn["pos_x"], hn["pos_y"], hn["pos_z"], hn["mass"]
        t0 = time.perf_counter()
        tree = orc.bh_pipeline(hn)
        build_s = time.perf_counter() - t0
        sample = min(n, args.bh_cpu_targets)
        tg = np.sort(np.random.default_rng(3).choice(n, sample, replace=False))
        threads = orc.max_threads()
        t0 = time.perf_counter()
        orc.bh_forces(tree["nodes"], x, y, z, m, theta, G, eps, targets=tg, nthreads=threads)
        trav_s = time.perf_counter() - t0
        est = (build_s + trav_s * n / sample) * 1e3
        tg1 = tg[:args.bh_cpu1_targets]
        t0 = time.perf_counter()
        orc.bh_forces(tree["nodes"], x, y, z, m, theta, G, eps, targets=tg1, nthreads=1)
        trav1_s = time.perf_counter() - t0
        est1 = (build_s + trav1_s * n / len(tg1)) * 1e3
        out["cpu_baseline"] = {
            "value": est, "unit": "ms/step", "cores": threads, "kind": "port",
            "sample": f"oracle build (bbox, Morton, stable sort, single-thread insertion + COM, "
                      f"as the reference host build) {build_s * 1e3:.0f} ms + traversal of "
                      f"{sample} of {n} targets on {threads} threads {trav_s * 1e3:.0f} ms, "
                      f"scaled to all targets",
            "single_thread": {"value": est1, "unit": "ms/step",
                              "sample": f"same build + traversal of {len(tg1)} targets on 1 "
                                        f"thread {trav1_s * 1e3:.0f} ms, scaled"},
            "host": host_info(orc)}
        log(f"barnes-hut cpu baseline: {est:.1f} ms/step on {threads} threads, {est1:.1f} on 1")
    return out


# ------------------------------------------------------------------ spatial hash

# Algorithmic flop of the spatial-hash force kernel: every candidate pair (a body
# of the 27 neighbour cells, itself excluded) costs the distance test, 3 sub +
# 3 mul + 2 add = 8 flop; a kept pair adds + eps^2, rsqrt, the force factor
# (G m) (inv^3) and the three accumulations, 12 flop more (20 in all, as the
# direct kernel counts an interaction).
SH_FLOP_CANDIDATE = 8.0
SH_FLOP_KEPT = 12.0


def grid_pair_counts(orc, h, cell, cutoff, eps, sample):
    """Candidate pairs of one force call (exact, from the cell counts) and kept pairs
    (pairs inside the cutoff, from the oracle on `sample` random targets)."""
    x, y, z, m = h["pos_x"], h["pos_y"], h["pos_z"], h["mass"]
    g = orc.grid_build(x, y, z, cell)
    dx, dy, dz = (int(v) for v in g["dims"])
    counts = (g["cell_end"] - g["cell_start"]).astype(np.int64).reshape(dz, dy, dx)
    p = np.pad(counts, 1)
    near = sum(p[1 + a:1 + a + dz, 1 + b:1 + b + dy, 1 + c:1 + c + dx]
               for a, b, c in itertools.product((-1, 0, 1), repeat=3))
    candidates = int((counts * (near - 1)).sum())
    n = len(x)
    tg = np.sort(np.random.default_rng(5).choice(n, min(sample, n), replace=False))
    _, _, kept = orc.grid_forces_f64(g, x, y, z, m, cutoff, 1.0, eps2_of(eps), targets=tg)
    return g, candidates, float(kept.mean()) * n, tg


def run_spatial_hash(nbx, args):
    n, G, eps, cell, cutoff, R = args.sh_n, 1.0, 0.1, 2.0, 2.0, 10.0
    h = nbx.gen_uniform_sphere(n, seed=42, R=R, M=float(n), sigma=0.0)
    d = nbx.DeviceParticles.from_host(h)
    grid = nbx.Grid(n, cell)
    s = nbx.Stream()
    grid.build(d, s)
    grid.forces(d, cutoff, G, eps, s)
    s.synchronize()
    K = args.sh_steps
    e0, e1, f0, f1 = nbx.Event(), nbx.Event(), nbx.Event(), nbx.Event()
    e0.record(s)
    for _ in range(K):  # SpatialHashCalculator::computeForces: build, then forces
        grid.build(d, s)
        grid.forces(d, cutoff, G, eps, s)
    e1.record(s)
    s.synchronize()
    ms_call = e0.elapsed_ms(e1) / K
    f0.record(s)
    for _ in range(K):
        grid.forces(d, cutoff, G, eps, s)
    f1.record(s)
    s.synchronize()
    ms_forces = f0.elapsed_ms(f1) / K
    b0, b1 = nbx.Event(), nbx.Event()
    b0.record(s)
    for _ in range(K):
        grid.build(d, s)
    b1.record(s)
    s.synchronize()
    ms_build = b0.elapsed_ms(b1) / K  # timed on its own: the difference of two loops is noise
    info = grid.info()
    orc = oracle()
    g, cand, kept, tg = grid_pair_counts(orc, h, cell, cutoff, eps, args.sh_check_targets)
    flop = SH_FLOP_CANDIDATE * cand + SH_FLOP_KEPT * kept
    tflops = flop / (ms_forces * 1e-3) / 1e12
    acc = d.acc()
    a64, mag, npairs = orc.grid_forces_f64(g, h["pos_x"], h["pos_y"], h["pos_z"], h["mass"],
                                           cutoff, G, eps2_of(eps), targets=tg)
    err = np.abs(acc[tg].astype(np.float64) - a64).max(axis=1) / np.maximum(mag, 1e-300)
    log(f"spatial hash N={n}: {ms_call:.3f} ms/call (forces {ms_forces:.3f} ms, "
        f"build {ms_build:.3f} ms), "
        f"{cand / n:.0f} candidates and {kept / n:.0f} kept pairs per target, "
        f"{tflops:.1f} TFLOP/s = {tflops / PEAK_FP32_TFLOPS:.1%} of FP32 peak, "
        f"err/sum|terms| max {err.max():.2e}")
    out = {"n": n, "cell": cell, "cutoff": cutoff, "eps": eps, "G": G, "calls": K,
           "workload": f"uniform sphere R={R:g}, unit masses (the reference's force.spatial_hash "
                       f"benchmark: spherical init, default cell, cutoff and softening)",
           "grid_dims": info["dims"].tolist(), "total_cells": info["total"],
           "ms_per_call": ms_call, "ms_forces": ms_forces, "ms_build": ms_build,
           "candidate_pairs": cand, "kept_pairs_est": kept,
           "roofline": {"bound": "valu-fp32", "kernel": "grid_forces (with the body gather)",
                        "achieved": tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": tflops / PEAK_FP32_TFLOPS,
                        "flop": f"{SH_FLOP_CANDIDATE:g} per candidate pair + "
                                f"{SH_FLOP_KEPT:g} per kept pair"},
           "force_error": {"targets": len(tg), "max_err_over_sum_abs_terms": float(err.max()),
                           "median": float(np.median(err))}}
    if not args.no_cpu_baseline:
        threads = orc.max_threads()
        x, y, z, m = h["pos_x"], h["pos_y"], h["pos_z"], h["mass"]
        t0 = time.perf_counter()
        gb = orc.grid_build(x, y, z, cell)
        build_s = time.perf_counter() - t0
        sample = min(n, args.sh_cpu_targets)
        st = np.sort(np.random.default_rng(6).choice(n, sample, replace=False))
        t0 = time.perf_counter()
        orc.grid_forces_f32(gb, x, y, z, m, cutoff, G, eps, targets=st, nthreads=threads)
        f_s = time.perf_counter() - t0
        est = (build_s + f_s * n / sample) * 1e3
        st1 = st[:args.sh_cpu1_targets]
        t0 = time.perf_counter()
        orc.grid_forces_f32(gb, x, y, z, m, cutoff, G, eps, targets=st1, nthreads=1)
        f1_s = time.perf_counter() - t0
        est1 = (build_s + f1_s * n / len(st1)) * 1e3
        out["cpu_baseline"] = {
            "value": est, "unit": "ms/call", "cores": threads, "kind": "port",
            "sample": f"oracle grid build (single thread) {build_s * 1e3:.0f} ms + FP32 forces of "
                      f"{sample} of {n} targets on {threads} threads {f_s * 1e3:.0f} ms, scaled "
                      f"to all targets",
            "single_thread": {"value": est1, "unit": "ms/call",
                              "sample": f"same build + forces of {len(st1)} targets on 1 thread "
                                        f"{f1_s * 1e3:.0f} ms, scaled"},
            "host": host_info(orc)}
        log(f"spatial hash cpu baseline: {est:.1f} ms/call on {threads} threads, {est1:.1f} on 1")
    return out


# ------------------------------------------------------------------ ensemble

def run_ensemble(nbx, args):
    counts = nbx.gen_ensemble_sizes(512, seed=42, nmin=256, nmax=16384)
    h = nbx.gen_ensemble(counts, seed=1000)
    G, eps, dt = 1.0, 0.01, 1e-3
    e2 = eps2_of(eps)
    pairs = int((counts * (counts - 1)).sum())
    ens = nbx.Ensemble(counts)
    ens.upload(h)
    s = nbx.Stream()
    v, tune_ms = ens.autotune(2)
    ens.forces(G, e2, s)
    chunk = args.ens_chunk
    ens.steps(G, e2, dt, chunk, chunk, s)  # capture + first replay
    s.synchronize()
    e0, e1 = nbx.Event(), nbx.Event()
    K = args.ens_steps
    e0.record(s)
    ens.steps(G, e2, dt, K, chunk, s)
    e1.record(s)
    s.synchronize()
    batched = e0.elapsed_ms(e1) / K
    # baseline: the same steps, one system at a time with the single-system entry points
    seq = nbx.Ensemble(counts)
    seq.upload(h)
    seq.forces(G, e2, s)
    seq.steps_sequential(G, e2, dt, 1, s)
    s.synchronize()
    Ks = max(1, args.ens_seq_steps)
    e0.record(s)
    seq.steps_sequential(G, e2, dt, Ks, s)
    e1.record(s)
    s.synchronize()
    sequential = e0.elapsed_ms(e1) / Ks
    err = None
    if not args.no_check:
        orc = oracle()
        ens.forces(G, e2, s)
        s.synchronize()
        acc = np.stack([ens.read("acc_x"), ens.read("acc_y"), ens.read("acc_z")], 1)
        pos = {k: ens.read(k) for k in ("pos_x", "pos_y", "pos_z", "mass")}
        worst = 0.0
        for q in np.linspace(0, len(counts) - 1, 8).astype(int):
            sl = ens.system(int(q))
            ref = orc.direct_f64(pos["pos_x"][sl], pos["pos_y"][sl], pos["pos_z"][sl],
                                 pos["mass"][sl], G, e2)
            e = np.linalg.norm(acc[sl] - ref, axis=1) / np.linalg.norm(ref, axis=1)
            worst = max(worst, float(np.quantile(e, 0.99)))
        err = {"systems_checked": 8, "worst_p99_rel_err_fp64": worst}
    names = nbx.variant_names()
    out = {"systems": int(len(counts)), "total_particles": int(counts.sum()),
           "interactions_per_step": pairs, "batched_ms_per_step": batched,
           "sequential_ms_per_step": sequential, "speedup": sequential / batched,
           "batched_Ginteractions_per_s": pairs / batched / 1e6,
           "steps_timed": K, "graph_chunk": chunk, "graph_launches": -(-K // chunk),
           "tile_variant": names[v], "force_error": err,
           "workload": "512 Plummer systems (a=1, M=1 each), N_s log-uniform in [256, 16384], "
                       "eps=0.01, dt=1e-3"}
    log(f"ensemble: batched {batched:.3f} ms/step, sequential {sequential:.3f} ms/step, "
        f"speedup {out['speedup']:.2f}, tile {names[v]}")
    if not args.no_cpu_baseline:
        # one ensemble step on the oracle's restatement of the reference CPU path: the force
        # evaluation of 8 systems spread over the size range (all targets), scaled by the
        # step's total interaction count
        orc = oracle()
        threads = orc.max_threads()
        pos = {k: h[k] for k in ("pos_x", "pos_y", "pos_z", "mass")}
        offs = np.concatenate([[0], np.cumsum(counts)])

        def timed(systems, nthreads):
            t0 = time.perf_counter()
            done = 0
            for q in systems:
                sl = slice(int(offs[q]), int(offs[q + 1]))
                orc.direct_f32(pos["pos_x"][sl], pos["pos_y"][sl], pos["pos_z"][sl],
                               pos["mass"][sl], G, eps, nthreads=nthreads)
                done += int(counts[q]) * (int(counts[q]) - 1)
            return done / (time.perf_counter() - t0)

        order = np.argsort(counts)
        sample = order[np.linspace(0, len(counts) - 1, 8).astype(int)]
        rate = timed(sample, threads)
        rate1 = timed(order[np.linspace(0, len(counts) // 2, 4).astype(int)], 1)
        out["cpu_baseline"] = {
            "value": pairs / rate * 1e3, "unit": "ms/step", "cores": threads, "kind": "port",
            "sample": f"oracle FP32 forces of 8 systems spread over the size range on {threads} "
                      f"threads ({rate / 1e9:.2f} Ginteractions/s), scaled to the step's "
                      f"{pairs:.3e} interactions",
            "single_thread": {"value": pairs / rate1 * 1e3, "unit": "ms/step",
                              "sample": f"4 systems on 1 thread ({rate1 / 1e9:.3f} "
                                        f"Ginteractions/s), scaled"},
            "host": host_info(orc)}
        log(f"ensemble cpu baseline: {out['cpu_baseline']['value']:.0f} ms/step on {threads} "
            f"threads, {out['cpu_baseline']['single_thread']['value']:.0f} on 1")
    return out


# ------------------------------------------------------------------ precision

def run_precision(nbx, args):
    n, G, eps, dt = args.prec_n, 1.0, 0.01, 2.5e-4
    e2 = eps2_of(eps)
    K = args.prec_steps
    every = max(1, K // 20)
    h = nbx.gen_plummer(n, seed=42)
    plain = nbx.DeviceParticles.from_host(h)
    ds = nbx.DSDeviceParticles.from_host(h)
    ws = nbx.DirectWorkspace(n)
    s = nbx.Stream()
    C = nbx.C

    def ds_energy():
        ke, pe = C.c_double(), C.c_double()
        mom = (C.c_double * 3)()
        nbx.check(nbx.lib.nbx_ds_energy(ds.ptr, G, e2, C.byref(ke), C.byref(pe), mom, s.h))
        return ke.value + pe.value, list(mom)

    def plain_energy():
        e = nbx.energy(plain, G, eps)
        return e["ke"] + e["pe"], e["mom"]

    orc = oracle()
    tg = np.sort(np.random.default_rng(5).choice(n, 512, replace=False))

    def force_err(acc, pos):
        ref = orc.direct_f64_pos64(pos[0], pos[1], pos[2], h["mass"], G, e2, targets=tg)
        e = np.linalg.norm(acc[tg] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        return {"median": float(np.median(e)), "p99": float(np.quantile(e, 0.99))}

    res = {}
    for name in ("plain", "double_single"):
        if name == "plain":
            nbx.check(nbx.lib.nbx_direct_compute(ws.h, plain.ptr, G, e2, s.h))
            energy = plain_energy
            def step(): nbx.check(nbx.lib.nbx_direct_compute_step(ws.h, plain.ptr, G, e2, dt, s.h))
            def state():
                a = plain.acc().astype(np.float64)
                return a, [plain.read(k).astype(np.float64) for k in ("pos_x", "pos_y", "pos_z")]
        else:
            nbx.check(nbx.lib.nbx_ds_forces(ws.h, ds.ptr, G, e2, s.h))
            energy = ds_energy
            def step(): nbx.check(nbx.lib.nbx_ds_step(ws.h, ds.ptr, G, e2, dt, s.h))
            def state(): return ds.acc().astype(np.float64), ds.positions64()
        s.synchronize()
        E0, P0 = energy()
        a0, p0 = state()
        err0 = force_err(a0, p0)
        drift = []
        e_start, e_stop = nbx.Event(), nbx.Event()
        run_ms = 0.0
        for k in range(0, K, every):
            e_start.record(s)
            for _ in range(min(every, K - k)):
                step()
            e_stop.record(s)
            run_ms += e_start.elapsed_ms(e_stop)
            E, P = energy()
            drift.append([k + min(every, K - k), (E - E0) / abs(E0)])
        a1, p1 = state()
        res[name] = {"ms_per_step": run_ms / K, "E0": E0,
                     "rel_energy_drift_final": drift[-1][1],
                     "rel_energy_drift_max": max(abs(v) for _, v in drift),
                     "momentum_final": energy()[1], "force_err_start": err0,
                     "force_err_end": force_err(a1, p1), "drift_series": drift}
        log(f"precision {name}: {res[name]['ms_per_step']:.3f} ms/step, final drift "
            f"{res[name]['rel_energy_drift_final']:.3e}, max {res[name]['rel_energy_drift_max']:.3e}")
    out = {"n": n, "steps": K, "dt": dt, "eps": eps,
           "workload": "Plummer sphere a=1, M=1, FP32 pair forces; double_single = hi/lo "
                       "positions, Kahan sums, FP64 renormalised drift", **res}
    if not args.no_cpu_baseline:
        # a step of the reference CPU path (the plain FP32 form; the reference has no
        # precision form) is one force evaluation: timed on a target sample and scaled
        threads = orc.max_threads()
        ts = np.sort(np.random.default_rng(8).choice(n, 4096, replace=False))
        t0 = time.perf_counter()
        orc.direct_f32(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], G, eps, targets=ts,
                       nthreads=threads)
        sec = time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.direct_f32(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], G, eps, targets=ts[:256],
                       nthreads=1)
        sec1 = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": sec * n / len(ts) * 1e3, "unit": "ms/step", "cores": threads,
            "kind": "port",
            "sample": f"oracle FP32 force evaluation of 4096 of {n} targets on {threads} threads "
                      f"{sec * 1e3:.0f} ms, scaled (the Verlet update is O(N) and left out)",
            "single_thread": {"value": sec1 * n / 256 * 1e3, "unit": "ms/step",
                              "sample": f"256 targets on 1 thread {sec1 * 1e3:.0f} ms, scaled"},
            "host": host_info(orc)}
        log(f"precision cpu baseline (plain FP32 path): {out['cpu_baseline']['value']:.0f} "
            f"ms/step on {threads} threads, {out['cpu_baseline']['single_thread']['value']:.0f} "
            f"on 1")
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--only", choices=["direct", "barnes-hut", "ensemble", "precision",
                                       "spatial-hash", "cpu-path"])
    ap.add_argument("--variant", type=int, default=-1, help="force a direct tile variant")
    ap.add_argument("--check-targets", type=int, default=4096,
                    help="sample targets of the FP64 force check")
    ap.add_argument("--autotune", action="store_true", help="time all direct tile variants first")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="one 256 MiB device copy first (reference launch for counter scaling)")
    ap.add_argument("--cpu-targets", type=int, default=16384)
    ap.add_argument("--cpu1-targets", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--bh-n", type=int, default=1 << 20)
    ap.add_argument("--bh-steps", type=int, default=10)
    ap.add_argument("--bh-cpu-targets", type=int, default=65536)
    ap.add_argument("--bh-cpu1-targets", type=int, default=8192)
    ap.add_argument("--ens-steps", type=int, default=1000)
    ap.add_argument("--ens-chunk", type=int, default=100)
    ap.add_argument("--ens-seq-steps", type=int, default=5)
    ap.add_argument("--sh-n", type=int, default=1 << 20)
    ap.add_argument("--sh-steps", type=int, default=10)
    ap.add_argument("--sh-check-targets", type=int, default=2048)
    ap.add_argument("--sh-cpu-targets", type=int, default=8192)
    ap.add_argument("--sh-cpu1-targets", type=int, default=1024)
    ap.add_argument("--prec-n", type=int, default=65536)
    ap.add_argument("--prec-steps", type=int, default=10000)
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this engine runs on a single MI355X "
              "(no multi-GPU path is built); run with --gpus 1", file=sys.stderr)
        return 2
    import nbx

    info = nbx.device_info()
    log(f"device {info['name']} ({info['cu']} CUs, {info['clock_khz'] / 1e3:.0f} MHz)")
    if args.only == "precision":
        print(json.dumps({"precision": run_precision(nbx, args), "device": info["name"]}),
              flush=True)
        return 0
    if args.only == "barnes-hut":
        print(json.dumps({"barnes_hut": run_bh(nbx, args), "device": info["name"]}), flush=True)
        return 0
    if args.only == "cpu-path":
        print(json.dumps({"cpu_path": run_cpu_path(nbx, args), "device": info["name"]}),
              flush=True)
        return 0
    if args.only == "spatial-hash":
        print(json.dumps({"spatial_hash": run_spatial_hash(nbx, args), "device": info["name"]}),
              flush=True)
        return 0
    if args.only == "ensemble":
        print(json.dumps({"ensemble": run_ensemble(nbx, args), "device": info["name"]}),
              flush=True)
        return 0
    direct = run_direct(nbx, args)
    line = {
        "metric": METRIC,
        "value": direct["value"],
        "unit": "Ginteractions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": direct["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": direct["config"],
        "roofline": direct["roofline"],
        "cpu_baseline": direct["cpu_baseline"],
        "force_error_fp64": direct["force_error_fp64"],
        "device": info["name"],
    }
    if args.only != "direct":
        line["barnes_hut"] = run_bh(nbx, args)
        line["ensemble"] = run_ensemble(nbx, args)
        line["spatial_hash"] = run_spatial_hash(nbx, args)
        line["precision"] = run_precision(nbx, args)
        if not args.no_cpu_baseline:
            line["cpu_path_n4096"] = run_cpu_path(nbx, args)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
