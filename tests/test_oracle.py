"""Pins of the CPU oracle (oracle/oracle.c) — runs on the CPU.

The oracle is only trusted after it reproduces, independently of its own
formulas: the reference's closed-form known-answer tests
(tests/test_force_calculation.cpp, tests/test_integrator.cpp of the
reference), float64 numpy implementations, and textbook routines.
"""
import math

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st


# ------------------------------------------------------------ pair function
# reference tests/test_force_calculation.cpp:17-34
def test_two_body_acceleration(orc):
    a = orc.pair_accel([0, 0, 0], [1, 0, 0], 1.0, 1.0, 0.0)
    assert a[0] > 0
    assert abs(a[1]) < 1e-6 and abs(a[2]) < 1e-6
    assert abs(np.linalg.norm(a.astype(np.float64)) - 1.0) < 1e-5


# :36-48
def test_softening_prevents_infinity(orc):
    a = orc.pair_accel([0, 0, 0], [0.001, 0, 0], 1.0, 1.0, 0.1)
    assert np.all(np.isfinite(a))


# :50-64
def test_acceleration_direction(orc):
    a = orc.pair_accel([0, 0, 0], [1, 1, 1], 1.0, 1.0, 0.01).astype(np.float64)
    d = np.ones(3) / math.sqrt(3.0)
    assert np.allclose(a / np.linalg.norm(a), d, atol=1e-5)


# :181-204 magnitude property, :206-231 direction property, :233-253 finiteness
@settings(max_examples=200, deadline=None)
@given(m2=st.floats(0.01, 100.0), r=st.floats(0.05, 100.0))
def test_force_magnitude_property(orc, m2, r):
    m2, r = float(np.float32(m2)), float(np.float32(r))
    a = orc.pair_accel([0, 0, 0], [r, 0, 0], m2, 1.0, 0.01).astype(np.float64)
    expect = m2 * r / (r * r + 1e-4) ** 1.5
    assert abs(np.linalg.norm(a) - expect) / expect < 0.01


@settings(max_examples=200, deadline=None)
@given(r=st.floats(0.05, 100.0), dx=st.floats(-1, 1), dy=st.floats(-1, 1), dz=st.floats(-1, 1))
def test_force_direction_property(orc, r, dx, dy, dz):
    d = np.array([dx, dy, dz], np.float64)
    if d @ d <= 0.01:
        return
    p2 = (d / np.linalg.norm(d) * r).astype(np.float32)
    a = orc.pair_accel([0, 0, 0], p2, 1.0, 1.0, 0.01).astype(np.float64)
    e = p2.astype(np.float64) / np.linalg.norm(p2.astype(np.float64))
    assert e @ (a / np.linalg.norm(a)) > 0.999


@settings(max_examples=200, deadline=None)
@given(eps=st.floats(0.001, 10.0), r=st.floats(0.0, 100.0))
def test_softening_finiteness_property(orc, eps, r):
    a = orc.pair_accel([0, 0, 0], [r, 0, 0], 1.0, 1.0, eps)
    assert np.all(np.isfinite(a)) and np.linalg.norm(a.astype(np.float64)) < 1e10


# ------------------------------------------------------------ direct sums
def grid32():
    """reference tests/test_force_calculation.cpp:112-121 input."""
    i = np.arange(32)
    x = (((i * 7) % 11).astype(np.float32) * np.float32(0.3)).astype(np.float32)
    y = (((i * 3) % 13).astype(np.float32) * np.float32(0.3)).astype(np.float32)
    z = (((i * 5) % 9).astype(np.float32) * np.float32(0.3)).astype(np.float32)
    m = (np.float32(0.5) + np.float32(0.1) * (i % 5).astype(np.float32)).astype(np.float32)
    return x, y, z, m


def numpy_f64_direct(x, y, z, m, G, eps2):
    """Independent float64 all-pairs sum (vectorised, not the oracle's loop)."""
    p = np.stack([x, y, z], 1).astype(np.float64)
    d = p[None, :, :] - p[:, None, :]
    r2 = (d * d).sum(-1) + eps2
    w = G * m.astype(np.float64)[None, :] * r2 ** -1.5
    np.fill_diagonal(w, 0.0)
    return (w[:, :, None] * d).sum(1)


def test_direct_f32_single_particle_zero(orc):
    a = orc.direct_f32([1.0], [2.0], [3.0], [5.0], 1.0, 0.1)
    assert np.all(a == 0.0)


def test_direct_f32_two_body_antisymmetric(orc):
    a = orc.direct_f32([-0.5, 0.5], [0, 0], [0, 0], [1, 1], 1.0, 0.01)
    assert np.allclose(a[0], -a[1], atol=1e-5) and a[0, 0] > 0


def test_direct_grid32_matches_float64(orc):
    x, y, z, m = grid32()
    a32 = orc.direct_f32(x, y, z, m, 1.0, 0.1).astype(np.float64)
    a64 = numpy_f64_direct(x, y, z, m, 1.0, orc.eps2_of(0.1))
    assert np.all(np.abs(a32 - a64) <= 1e-5 * (np.abs(a64) + 1e-3))
    f64 = orc.direct_f64(x, y, z, m, 1.0, orc.eps2_of(0.1))
    assert np.allclose(f64, a64, rtol=1e-12, atol=1e-13)


def test_direct_f64_matches_numpy_plummer(orc, nbx):
    h = nbx.gen_plummer(700, seed=7)
    x, y, z, m = h["pos_x"], h["pos_y"], h["pos_z"], h["mass"]
    eps2 = orc.eps2_of(0.01)
    ref = numpy_f64_direct(x, y, z, m, 1.0, eps2)
    got = orc.direct_f64(x, y, z, m, 1.0, eps2)
    scale = np.abs(ref).sum(1, keepdims=True)
    assert np.max(np.abs(got - ref) / scale) < 1e-13
    t = np.array([0, 5, 699, 350])
    assert np.array_equal(orc.direct_f64(x, y, z, m, 1.0, eps2, targets=t), got[t])


def test_direct_f32_targets_subset(orc, nbx):
    h = nbx.gen_plummer(300, seed=3)
    args = (h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 1.0, 0.05)
    full = orc.direct_f32(*args)
    t = np.array([7, 0, 299])
    assert np.array_equal(orc.direct_f32(*args, targets=t), full[t])


# ------------------------------------------------------------ integrator
# reference tests/test_integrator.cpp:14-41
def test_single_step_position_update(orc, nbx):
    h = nbx.empty_host(1)
    h["vel_x"][0] = 1.0
    h["mass"][0] = 1.0
    orc.drift(h, 0.1)
    assert abs(h["pos_x"][0] - 0.1) < 1e-5 and h["pos_y"][0] == 0 and h["pos_z"][0] == 0


# :43-74
def test_constant_acceleration_closed_form(orc, nbx):
    h = nbx.empty_host(1)
    for k, v in dict(pos_x=1.0, pos_y=-2.0, pos_z=0.5, vel_x=0.3, vel_z=-0.2, acc_x=2.0,
                     acc_y=-1.0, acc_z=4.0, mass=1.0).items():
        h[k][0] = v
    dt = 0.01
    orc.drift(h, dt)
    exp = [1.0 + 0.3 * dt + 0.5 * 2.0 * dt * dt, -2.0 - 0.5 * dt * dt, 0.5 - 0.2 * dt + 2.0 * dt * dt]
    got = [h["pos_x"][0], h["pos_y"][0], h["pos_z"][0]]
    assert np.allclose(got, exp, atol=1e-6)


def test_kick_closed_form(orc, nbx):
    h = nbx.empty_host(2)
    h["vel_x"][:] = [1.0, -2.0]
    h["acc_old_x"][:] = [2.0, 0.5]
    h["acc_x"][:] = [4.0, 0.25]
    orc.kick(h, 0.1)
    assert np.allclose(h["vel_x"], [1.0 + 0.05 * 6.0, -2.0 + 0.05 * 0.75], atol=1e-6)


# :76-102
def test_kinetic_energy_known_answer(orc, nbx):
    h = nbx.empty_host(2)
    h["vel_x"][0], h["vel_y"][1] = 1.0, 2.0
    h["mass"][:] = [1.0, 2.0]
    e = orc.energy_f64(h, 1.0, orc.eps2_of(0.1))
    assert abs(e["ke"] - 4.5) < 1e-12
    assert np.allclose(e["mom"], [1.0, 4.0, 0.0])


def test_potential_matches_numpy(orc, nbx):
    h = nbx.gen_plummer(400, seed=11)
    eps2 = orc.eps2_of(0.01)
    p = np.stack([h["pos_x"], h["pos_y"], h["pos_z"]], 1).astype(np.float64)
    m = h["mass"].astype(np.float64)
    d = p[None] - p[:, None]
    r = np.sqrt((d * d).sum(-1) + eps2)
    iu = np.triu_indices(len(m), 1)
    pe = -(m[:, None] * m[None, :] / r)[iu].sum()
    e = orc.energy_f64(h, 1.0, eps2)
    assert abs(e["pe"] - pe) < 1e-12 * abs(pe)


# :107-168
def test_energy_conservation_two_body_orbit(orc, nbx):
    h = nbx.empty_host(2)
    r, m = 5.0, 1.0
    v = math.sqrt(m / (4 * r))
    h["pos_x"][:] = [-r, r]
    h["vel_y"][:] = [-v, v]
    h["mass"][:] = [m, m]
    eps2 = orc.eps2_of(0.01)
    h0 = {k: a.copy() for k, a in h.items()}
    e0 = orc.energy_f64(h0, 1.0, eps2)
    orc.run_direct(h, 1.0, 0.01, 0.001, 100)
    e1 = orc.energy_f64(h, 1.0, eps2)
    E0, E1 = e0["ke"] + e0["pe"], e1["ke"] + e1["pe"]
    assert abs(E1 - E0) / abs(E0) < 1e-3
    # the orbit moved the bodies and kept them symmetric about the origin
    assert h["pos_y"][1] > 0 and abs(h["pos_y"][0] + h["pos_y"][1]) < 1e-6
