"""Barnes-Hut at the 1,048,576-body collision workload, stepped.

Two Plummer spheres (a = 1, M = 1 each, 524,288 bodies each) at x = -+5,
approaching at 0.5; theta = 0.5, eps = 0.01, dt = 1e-3. The reference rebuilds
the tree on every force call (src/cuda/force_barnes_hut.cu:577-583), so every
step's build is part of the contract: after each drift the GPU's own positions
are fed to the oracle pipeline (bbox, Morton, stable sort, insertion build,
src/cuda/force_barnes_hut.cu:266-543), and the bounding box, the sorted Morton
codes, the permutation and the whole 76-byte node array must be bit-exact.
The drift and kick are checked bit for bit against the oracle's IEEE
restatement of src/cuda/integrator.cu:12-37 on the same state, and the forces
on 1,024 sample targets against the oracle's traversal of its own tree.
"""
import numpy as np
import pytest

import bh_model as M

pytestmark = pytest.mark.gpu

N = 1 << 20
THETA, G, EPS, DT = 0.5, 1.0, 0.01, 1e-3
STEPS = 3


def state(d):
    return {k: d.read(k) for k in ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z",
                                   "acc_x", "acc_y", "acc_z", "acc_old_x", "acc_old_y",
                                   "acc_old_z", "mass")}


def check_build(nbx, orc, t, h, step):
    ref = orc.bh_pipeline(h)
    lo, hi, codes, idx = t.sort_result(N)
    assert np.array_equal(lo, ref["bbox_min"]) and np.array_equal(hi, ref["bbox_max"]), step
    assert np.array_equal(codes, ref["codes"]), f"step {step}: Morton codes differ"
    assert np.array_equal(idx, ref["sorted_idx"]), f"step {step}: permutation differs"
    cnt, depth = t.node_count()
    assert (cnt, depth) == (ref["node_count"], ref["max_depth"]), step
    ok, why = M.same_nodes(t.nodes(), ref["nodes"])
    assert ok, f"step {step}: {why}"
    return ref


def check_forces(orc, d, ref, h, step):
    tg = np.sort(np.random.default_rng(step).choice(N, 1024, replace=False))
    got = d.acc().astype(np.float64)[tg]
    a = orc.bh_forces(ref["nodes"], h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], THETA, G, EPS,
                      targets=tg).astype(np.float64)
    err = np.linalg.norm(got - a, axis=1) / np.linalg.norm(a, axis=1)
    print(f"step {step}: nodes {ref['node_count']}, depth {ref['max_depth']}, "
          f"force rel err median {np.median(err):.2e} max {err.max():.2e}", flush=True)
    assert np.median(err) < 1e-6 and err.max() < 1e-4, step


def test_collision_steps_bit_exact_every_build(nbx, orc):
    h0 = nbx.gen_two_plummer(N, seed=42)
    d = nbx.DeviceParticles.from_host(h0)
    t = nbx.BHTree(N)
    # a(0), as ParticleSystem::finishInitialization (src/core/particle_system.cpp:48-52)
    t.build(d)
    t.forces(d, THETA, G, EPS)
    nbx.synchronize()
    h = state(d)
    ref = check_build(nbx, orc, t, h, 0)
    check_forces(orc, d, ref, h, 0)
    for step in range(1, STEPS + 1):
        before = state(d)
        nbx.check(nbx.lib.nbx_verlet_drift(d.ptr, DT, None))
        t.build(d)
        t.forces(d, THETA, G, EPS)
        nbx.check(nbx.lib.nbx_update_velocities(d.ptr, DT, None))
        nbx.synchronize()
        h = state(d)
        # drift and kick against the oracle on the same state and the GPU's forces
        exp = {k: v.copy() for k, v in before.items()}
        orc.store_old(exp)
        orc.drift(exp, DT)
        for k in ("pos_x", "pos_y", "pos_z", "acc_old_x", "acc_old_y", "acc_old_z"):
            assert np.array_equal(h[k], exp[k]), (step, k)
        for k in ("acc_x", "acc_y", "acc_z"):
            exp[k] = h[k].copy()
        orc.kick(exp, DT)
        for k in ("vel_x", "vel_y", "vel_z"):
            assert np.array_equal(h[k], exp[k]), (step, k)
        ref = check_build(nbx, orc, t, h, step)
        check_forces(orc, d, ref, h, step)
