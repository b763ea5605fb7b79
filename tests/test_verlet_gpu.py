"""GPU parity of the Velocity Verlet kernels and the FP64 energy reductions.

Drift, kick and acc_old store are bit-exact against the oracle (same IEEE
float operations in the same order). Whole steps differ from the oracle only
through the force sum, so trajectories are compared with a tolerance."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def random_state(nbx, n, seed):
    r = np.random.default_rng(seed)
    h = nbx.empty_host(n)
    for k in h:
        h[k][:] = r.standard_normal(n).astype(np.float32)
    h["mass"][:] = np.abs(h["mass"]) + np.float32(0.1)
    return h


def test_drift_store_kick_bit_exact(nbx, orc):
    h = random_state(nbx, 1537, 3)
    d = nbx.DeviceParticles.from_host(h)
    d.write("acc_old_x", h["acc_old_x"]); d.write("acc_old_y", h["acc_old_y"])
    d.write("acc_old_z", h["acc_old_z"])
    dt = 0.0137
    nbx.check(nbx.lib.nbx_verlet_drift(d.ptr, dt, None))
    nbx.synchronize()
    ref = {k: v.copy() for k, v in h.items()}
    orc.store_old(ref)
    orc.drift(ref, dt)
    for k in ("pos_x", "pos_y", "pos_z", "acc_old_x", "acc_old_y", "acc_old_z"):
        assert np.array_equal(d.read(k), ref[k]), k
    nbx.check(nbx.lib.nbx_update_velocities(d.ptr, dt, None))
    nbx.synchronize()
    orc.kick(ref, dt)
    for k in ("vel_x", "vel_y", "vel_z"):
        assert np.array_equal(d.read(k), ref[k]), k
    nbx.check(nbx.lib.nbx_update_positions(d.ptr, dt, None))
    nbx.synchronize()
    orc.drift(ref, dt)
    assert np.array_equal(d.read("pos_x"), ref["pos_x"])


def test_integrator_closed_forms(nbx):
    # reference tests/test_integrator.cpp:14-74
    h = nbx.empty_host(1)
    for k, v in dict(pos_x=1.0, pos_y=-2.0, pos_z=0.5, vel_x=0.3, vel_z=-0.2, acc_x=2.0,
                     acc_y=-1.0, acc_z=4.0, mass=1.0).items():
        h[k][0] = v
    d = nbx.DeviceParticles.from_host(h)
    dt = 0.01
    nbx.check(nbx.lib.nbx_update_positions(d.ptr, dt, None))
    nbx.synchronize()
    exp = [1.0 + 0.3 * dt + dt * dt, -2.0 - 0.5 * dt * dt, 0.5 - 0.2 * dt + 2.0 * dt * dt]
    got = [d.read("pos_x")[0], d.read("pos_y")[0], d.read("pos_z")[0]]
    assert np.allclose(got, exp, atol=1e-6)


def test_direct_step_tracks_oracle_run(nbx, orc):
    h = nbx.gen_plummer(1024, seed=21)
    G, eps, dt, steps = 1.0, 0.01, 1e-3, 10
    d = nbx.DeviceParticles.from_host(h)
    nbx.direct_forces(d, G, eps)
    for _ in range(steps):
        nbx.direct_step(d, G, eps, dt)
    nbx.synchronize()
    ref = {k: v.copy() for k, v in h.items()}
    orc.run_direct(ref, G, eps, dt, steps)
    got = d.download()
    for k in ("pos_x", "pos_y", "pos_z"):
        assert np.max(np.abs(got[k] - ref[k])) < 1e-6, k
    for k in ("vel_x", "vel_y", "vel_z"):
        assert np.max(np.abs(got[k] - ref[k])) < 1e-5, k


def test_plummer_4096_100_steps_match_cpu_path(nbx, orc):
    # the reference's CPU-runnable case: equal-mass (1/N) Plummer sphere, eps 0.01,
    # dt 1e-3, 100 Velocity Verlet steps, against the oracle's restatement of the
    # reference CPU path (FP32 sequential j sums)
    h = nbx.gen_plummer(4096, seed=42)
    G, eps, dt, steps = 1.0, 0.01, 1e-3, 100
    d = nbx.DeviceParticles.from_host(h)
    nbx.direct_forces(d, G, eps)
    for _ in range(steps):
        nbx.direct_step(d, G, eps, dt)
    nbx.synchronize()
    ref = {k: v.copy() for k, v in h.items()}
    orc.run_direct(ref, G, eps, dt, steps)
    got = d.download()
    for k in ("pos_x", "pos_y", "pos_z"):
        assert np.max(np.abs(got[k] - ref[k])) < 1e-5, k
    for k in ("vel_x", "vel_y", "vel_z"):
        assert np.max(np.abs(got[k] - ref[k])) < 1e-4, k
    e2 = orc.eps2_of(eps)
    eg, er = orc.energy_f64(got, G, e2), orc.energy_f64(ref, G, e2)
    assert abs((eg["ke"] + eg["pe"]) - (er["ke"] + er["pe"])) <= 1e-6 * abs(er["ke"] + er["pe"])


def test_energy_matches_oracle(nbx, orc):
    h = nbx.gen_plummer(3000, seed=8)
    d = nbx.DeviceParticles.from_host(h)
    e = nbx.energy(d, 1.0, 0.01)
    r = orc.energy_f64(h, 1.0, orc.eps2_of(0.01))
    assert abs(e["ke"] - r["ke"]) <= 1e-13 * abs(r["ke"])
    assert abs(e["pe"] - r["pe"]) <= 1e-12 * abs(r["pe"])
    assert np.allclose(e["mom"], r["mom"], atol=1e-15)


def test_kinetic_energy_known_answer(nbx):
    h = nbx.empty_host(2)
    h["vel_x"][0], h["vel_y"][1] = 1.0, 2.0
    h["mass"][:] = [1.0, 2.0]
    d = nbx.DeviceParticles.from_host(h)
    assert nbx.energy(d, 1.0, 0.1, potential=False)["ke"] == 4.5


def test_two_body_orbit_energy_drift(nbx):
    # reference tests/test_integrator.cpp:107-168
    h = nbx.empty_host(2)
    r, m = 5.0, 1.0
    v = math.sqrt(m / (4 * r))
    h["pos_x"][:] = [-r, r]
    h["vel_y"][:] = [-v, v]
    h["mass"][:] = m
    d = nbx.DeviceParticles.from_host(h)
    nbx.direct_forces(d, 1.0, 0.01)
    e0 = nbx.energy(d, 1.0, 0.01)
    for _ in range(100):
        nbx.direct_step(d, 1.0, 0.01, 1e-3)
    e1 = nbx.energy(d, 1.0, 0.01)
    E0, E1 = e0["ke"] + e0["pe"], e1["ke"] + e1["pe"]
    assert abs(E1 - E0) / abs(E0) < 1e-3
