"""The precision form at its full workload: 65,536-body Plummer sphere.

Plummer a = 1, M = 1, m = 1/N, eps = 0.01, dt = 2.5e-4. Double-single (hi + lo)
positions, FP32 pair forces with compensated sums, FP64 energies. Checked:

* forces of both double-single kernels against the oracle's FP64 direct sum
  over the hi + lo positions, on 4,096 sample targets, at t = 0 and after
  200 Verlet steps (the GPU's own positions then): median <= 3e-7,
  p99 <= 1e-6, the plain path's bounds;
* over the workload's 10,000 steps the relative energy drift of the
  double-single run stays within that of the plain FP32 run on the same input
  (energies in FP64 from the softened pair potential of
  src/cuda/integrator.cu:82-104, evaluated every 500 steps).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 65536
G, EPS, DT = 1.0, 0.01, 2.5e-4


@pytest.fixture(scope="module")
def plummer(nbx):
    return nbx.gen_plummer(N, seed=42)


@pytest.fixture(scope="module")
def ws(nbx):
    return nbx.DirectWorkspace(N)


@pytest.fixture(scope="module")
def targets():
    return np.sort(np.random.default_rng(11).choice(N, 4096, replace=False))


def fp64_err(orc, h, acc, pos, tg):
    ref = orc.direct_f64_pos64(pos[0], pos[1], pos[2], h["mass"], G, orc.eps2_of(EPS), targets=tg)
    e = np.linalg.norm(acc[tg].astype(np.float64) - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return float(np.median(e)), float(np.quantile(e, 0.99)), float(e.max())


@pytest.mark.parametrize("kind", [0, 1], ids=["one_sided", "pair_symmetric"])
def test_ds_forces_vs_fp64_start_and_after_200_steps(nbx, orc, ws, plummer, targets, kind):
    e2 = orc.eps2_of(EPS)
    d = nbx.DSDeviceParticles.from_host(plummer)
    nbx.check(nbx.lib.nbx_ds_set_kernel(kind))
    try:
        nbx.check(nbx.lib.nbx_ds_forces(ws.h, d.ptr, G, e2, None))
        nbx.synchronize()
        med, p99, mx = fp64_err(orc, plummer, d.acc(), d.positions64(), targets)
        print(f"kernel {kind} t=0: median {med:.2e} p99 {p99:.2e} max {mx:.2e}", flush=True)
        assert med <= 3e-7 and p99 <= 1e-6
        for _ in range(200):
            nbx.check(nbx.lib.nbx_ds_step(ws.h, d.ptr, G, e2, DT, None))
        nbx.synchronize()
    finally:
        nbx.check(nbx.lib.nbx_ds_set_kernel(-1))
    pos = d.positions64()
    assert np.any(d.read("pos_lo_x") != 0)
    med, p99, mx = fp64_err(orc, plummer, d.acc(), pos, targets)
    print(f"kernel {kind} after 200 steps: median {med:.2e} p99 {p99:.2e} max {mx:.2e}",
          flush=True)
    assert med <= 3e-7 and p99 <= 1e-6


def test_ds_energy_drift_within_plain(nbx, orc, ws, plummer):
    e2 = orc.eps2_of(EPS)
    C = nbx.C
    plain = nbx.DeviceParticles.from_host(plummer)
    ds = nbx.DSDeviceParticles.from_host(plummer)
    nbx.check(nbx.lib.nbx_direct_compute(ws.h, plain.ptr, G, e2, None))
    nbx.check(nbx.lib.nbx_ds_forces(ws.h, ds.ptr, G, e2, None))

    def e_ds():
        ke, pe = C.c_double(), C.c_double()
        mom = (C.c_double * 3)()
        nbx.check(nbx.lib.nbx_ds_energy(ds.ptr, G, e2, C.byref(ke), C.byref(pe), mom, None))
        return ke.value + pe.value

    def e_plain():
        e = nbx.energy(plain, G, EPS)
        return e["ke"] + e["pe"]

    E0p, E0d = e_plain(), e_ds()
    assert abs(E0p - E0d) <= 1e-12 * abs(E0p)  # same state, lo parts zero
    worst_p = worst_d = 0.0
    for _ in range(20):
        for _ in range(500):
            nbx.check(nbx.lib.nbx_direct_compute_step(ws.h, plain.ptr, G, e2, DT, None))
            nbx.check(nbx.lib.nbx_ds_step(ws.h, ds.ptr, G, e2, DT, None))
        worst_p = max(worst_p, abs(e_plain() - E0p) / abs(E0p))
        worst_d = max(worst_d, abs(e_ds() - E0d) / abs(E0d))
    print(f"max |dE/E| over 10000 steps: plain {worst_p:.3e}, double-single {worst_d:.3e}",
          flush=True)
    assert worst_d <= worst_p
