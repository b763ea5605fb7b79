"""GPU checks of the precision form (double-single positions, Kahan pair sums).

* forces vs the FP64 oracle with positions hi + lo (same bounds as the plain
  direct kernel: median <= 3e-7, p99 <= 1e-6);
* the drift is bit-exact against the same FP64 arithmetic in numpy;
* energies use hi + lo positions (FP64 oracle in numpy);
* over 300 steps a double-single run stays closer to an all-FP64 integration
  than the plain FP32 run does.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def eps2(orc, eps):
    return orc.eps2_of(eps)


@pytest.fixture(scope="module")
def ws(nbx):
    """Workspace of the pair-symmetric kernels for up to 8,192 bodies."""
    return nbx.DirectWorkspace(8192)


@pytest.fixture(params=[0, 1], ids=["one_sided", "pair_symmetric"])
def ds_kernel(nbx, request):
    """Runs a test with each double-single force kernel forced."""
    nbx.check(nbx.lib.nbx_ds_set_kernel(request.param))
    yield request.param
    nbx.check(nbx.lib.nbx_ds_set_kernel(-1))


def test_ds_forces_match_fp64(nbx, orc, ws, ds_kernel):
    h = nbx.gen_plummer(4096, seed=42)
    rng = np.random.default_rng(0)
    x64 = [h[k].astype(np.float64) + rng.uniform(-1, 1, 4096) * 2.0 ** -26 * np.abs(h[k])
           for k in ("pos_x", "pos_y", "pos_z")]
    d = nbx.DSDeviceParticles.from_host(h, pos64=x64)
    lo = d.read("pos_lo_x")
    assert np.any(lo != 0)
    nbx.check(nbx.lib.nbx_ds_forces(ws.h, d.ptr, 1.0, eps2(orc, 0.01), None))
    nbx.synchronize()
    got = d.acc().astype(np.float64)
    pos = d.positions64()
    ref = orc.direct_f64_pos64(pos[0], pos[1], pos[2], h["mass"], 1.0, eps2(orc, 0.01))
    e = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"ds forces: median {np.median(e):.2e} p99 {np.quantile(e, 0.99):.2e} max {e.max():.2e}")
    assert np.median(e) <= 3e-7 and np.quantile(e, 0.99) <= 1e-6


def test_ds_forces_ragged_and_zero_softening(nbx, orc, ws, ds_kernel):
    # 1000 and 2500 bodies give 2 and 5 super-blocks of 512 in the symmetric kernel
    for n, eps in ((1, 0.01), (65, 0.01), (1000, 0.0), (2500, 0.01), (2500, 0.0)):
        h = nbx.gen_uniform_sphere(n, seed=n)
        d = nbx.DSDeviceParticles.from_host(h)
        nbx.check(nbx.lib.nbx_ds_forces(ws.h, d.ptr, 1.0, eps2(orc, eps), None))
        nbx.synchronize()
        got = d.acc().astype(np.float64)
        ref = orc.direct_f64(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 1.0, eps2(orc, eps))
        if n == 1:
            assert np.all(got == 0)
            continue
        e = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
        assert np.quantile(e, 0.99) <= 1e-5, (n, eps)


def test_ds_symmetric_mixed_masses_and_kick(nbx, orc, ws):
    # per-pair masses on both sides; then one step, whose kick is
    # v += (a_prev + a_new) (0.5 dt) in IEEE single rounding
    n = 5000
    h = nbx.gen_plummer(n, seed=8)
    h["mass"][:] = np.random.default_rng(4).uniform(0.2, 1.8, n).astype(np.float32) / np.float32(n)
    rng = np.random.default_rng(5)
    x64 = [h[k].astype(np.float64) + rng.uniform(-1, 1, n) * 2.0 ** -26 * np.abs(h[k])
           for k in ("pos_x", "pos_y", "pos_z")]
    d = nbx.DSDeviceParticles.from_host(h, pos64=x64)
    e2 = eps2(orc, 0.01)
    dt = np.float32(1e-3)
    nbx.check(nbx.lib.nbx_ds_set_kernel(1))
    try:
        nbx.check(nbx.lib.nbx_ds_forces(ws.h, d.ptr, 1.0, e2, None))
        nbx.synchronize()
        got = d.acc().astype(np.float64)
        pos = d.positions64()
        ref = orc.direct_f64_pos64(pos[0], pos[1], pos[2], h["mass"], 1.0, e2)
        e = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
        assert np.median(e) <= 3e-7 and np.quantile(e, 0.99) <= 1e-6
        a_prev = [d.read(f) for f in ("acc_x", "acc_y", "acc_z")]
        v_prev = [d.read(f) for f in ("vel_x", "vel_y", "vel_z")]
        nbx.check(nbx.lib.nbx_ds_step(ws.h, d.ptr, 1.0, e2, float(dt), None))
        nbx.synchronize()
    finally:
        nbx.check(nbx.lib.nbx_ds_set_kernel(-1))
    half = np.float32(np.float32(0.5) * dt)
    for k, c in enumerate("xyz"):
        a_new = d.read(f"acc_{c}")
        want = (v_prev[k] + ((a_prev[k] + a_new).astype(np.float32) * half).astype(np.float32))
        assert np.array_equal(d.read(f"vel_{c}"), want.astype(np.float32)), c


def test_ds_drift_bit_exact(nbx):
    n = 3001
    h = nbx.gen_plummer(n, seed=3)
    r = np.random.default_rng(1)
    for k in ("acc_x", "acc_y", "acc_z"):
        h[k][:] = r.standard_normal(n).astype(np.float32)
    x64 = [h[k].astype(np.float64) + r.uniform(-1, 1, n) * 1e-9 for k in ("pos_x", "pos_y", "pos_z")]
    d = nbx.DSDeviceParticles.from_host(h, pos64=x64)
    hi0 = [d.read(k) for k in ("pos_x", "pos_y", "pos_z")]
    lo0 = [d.read(k) for k in ("pos_lo_x", "pos_lo_y", "pos_lo_z")]
    dt = np.float32(2.5e-4)
    nbx.check(nbx.lib.nbx_ds_drift(d.ptr, float(dt), None))
    nbx.synchronize()
    ddt = np.float64(dt)
    for a, (vk, ak) in enumerate((("vel_x", "acc_x"), ("vel_y", "acc_y"), ("vel_z", "acc_z"))):
        inc = h[vk].astype(np.float64) * ddt + h[ak].astype(np.float64) * (0.5 * ddt * ddt)
        x = (hi0[a].astype(np.float64) + lo0[a].astype(np.float64)) + inc
        hi = x.astype(np.float32)
        lo = (x - hi.astype(np.float64)).astype(np.float32)
        assert np.array_equal(d.read(("pos_x", "pos_y", "pos_z")[a]), hi)
        assert np.array_equal(d.read(("pos_lo_x", "pos_lo_y", "pos_lo_z")[a]), lo)
        assert np.array_equal(d.read(("acc_old_x", "acc_old_y", "acc_old_z")[a]), h[ak])


def test_ds_energy(nbx, orc):
    n = 1500
    h = nbx.gen_plummer(n, seed=9)
    r = np.random.default_rng(2)
    x64 = [h[k].astype(np.float64) + r.uniform(-1, 1, n) * 1e-8 for k in ("pos_x", "pos_y", "pos_z")]
    d = nbx.DSDeviceParticles.from_host(h, pos64=x64)
    ke, pe = nbx.C.c_double(), nbx.C.c_double()
    mom = (nbx.C.c_double * 3)()
    e2 = eps2(orc, 0.01)
    nbx.check(nbx.lib.nbx_ds_energy(d.ptr, 1.0, e2, nbx.C.byref(ke), nbx.C.byref(pe), mom, None))
    p = np.stack(d.positions64(), 1)
    m = h["mass"].astype(np.float64)
    dd = p[None] - p[:, None]
    rr = np.sqrt((dd * dd).sum(-1) + e2)
    iu = np.triu_indices(n, 1)
    pe_ref = -(m[:, None] * m[None, :] / rr)[iu].sum()
    v = np.stack([h["vel_x"], h["vel_y"], h["vel_z"]], 1).astype(np.float64)
    assert abs(pe.value - pe_ref) <= 1e-12 * abs(pe_ref)
    assert abs(ke.value - 0.5 * (m * (v * v).sum(1)).sum()) <= 1e-12 * ke.value


def fp64_run(h, G, eps2_, dt, steps):
    p = np.stack([h["pos_x"], h["pos_y"], h["pos_z"]], 1).astype(np.float64)
    v = np.stack([h["vel_x"], h["vel_y"], h["vel_z"]], 1).astype(np.float64)
    m = h["mass"].astype(np.float64)

    def acc(p):
        d = p[None] - p[:, None]
        w = G * m[None] * ((d * d).sum(-1) + eps2_) ** -1.5
        np.fill_diagonal(w, 0)
        return (w[..., None] * d).sum(1)

    a = acc(p)
    for _ in range(steps):
        p = p + v * dt + a * (0.5 * dt * dt)
        an = acc(p)
        v = v + (a + an) * (0.5 * dt)
        a = an
    return p


def test_ds_tracks_fp64_integration_better(nbx, orc, ws):
    n, dt, steps, eps = 256, 1e-3, 300, 0.01
    h = nbx.gen_plummer(n, seed=5)
    e2 = eps2(orc, eps)
    ref = fp64_run(h, 1.0, e2, float(np.float32(dt)), steps)
    plain = nbx.DeviceParticles.from_host(h)
    nbx.direct_forces(plain, 1.0, eps, ws=ws)
    ds = nbx.DSDeviceParticles.from_host(h)
    nbx.check(nbx.lib.nbx_ds_forces(ws.h, ds.ptr, 1.0, e2, None))
    for _ in range(steps):
        nbx.direct_step(plain, 1.0, eps, dt, ws=ws)
        nbx.check(nbx.lib.nbx_ds_step(ws.h, ds.ptr, 1.0, e2, dt, None))
    nbx.synchronize()
    pp = np.stack([plain.read(k) for k in ("pos_x", "pos_y", "pos_z")], 1).astype(np.float64)
    pd = np.stack(ds.positions64(), 1)
    dev_plain = np.abs(pp - ref).max()
    dev_ds = np.abs(pd - ref).max()
    print(f"max |x - x_fp64| after {steps} steps: plain {dev_plain:.2e}, double-single {dev_ds:.2e}")
    assert dev_ds < dev_plain


def test_ds_symmetric_needs_workspace(nbx, orc):
    d = nbx.DSDeviceParticles.from_host(nbx.gen_plummer(2000, seed=1))
    nbx.check(nbx.lib.nbx_ds_set_kernel(1))
    try:
        with pytest.raises(nbx.NbxError):
            nbx.check(nbx.lib.nbx_ds_forces(None, d.ptr, 1.0, eps2(orc, 0.01), None))
    finally:
        nbx.check(nbx.lib.nbx_ds_set_kernel(-1))
    small = nbx.DirectWorkspace(1000)
    with pytest.raises(nbx.NbxError):
        nbx.check(nbx.lib.nbx_ds_forces(small.h, d.ptr, 1.0, eps2(orc, 0.01), None))


def test_ds_symmetric_rounds_bit_identical(nbx, orc):
    # a workspace with room for three partial slots runs the tiles in rounds of one
    # distance; the accelerations equal the one-round run bit for bit
    n = 7 * 1024 + 300
    h = nbx.gen_plummer(n, seed=4)
    x64 = [h[k].astype(np.float64) + np.random.default_rng(3).uniform(-1, 1, n) * 1e-8
           for k in ("pos_x", "pos_y", "pos_z")]
    e2 = eps2(orc, 0.01)
    res = []
    # the smallest workspace accepted: three slots for the body count padded to the
    # largest super-block (8,192); with the kernel's 512-body super-blocks (15 of
    # them) that still holds only rounds of one distance
    for scratch in (0, 9 * 4 * 8 * 1024):
        w = nbx.DirectWorkspace(n, scratch)
        d = nbx.DSDeviceParticles.from_host(h, pos64=x64)
        nbx.check(nbx.lib.nbx_ds_set_kernel(1))
        try:
            nbx.check(nbx.lib.nbx_ds_forces(w.h, d.ptr, 1.0, e2, None))
            nbx.synchronize()
        finally:
            nbx.check(nbx.lib.nbx_ds_set_kernel(-1))
        res.append(d.acc())
    assert np.array_equal(res[0], res[1])
