"""GPU parity of the direct N^2 force kernel against the CPU oracle.

Tolerances:
  * reference per-component bound (tests/test_force_calculation.cpp:141-146 of
    the reference): |gpu - ref| <= 1e-3 (|ref| + 1e-3) vs the FP32 sequential sum;
  * FP64 oracle: per particle e_i = |a_gpu - a_64| / |a_64|. Asserted: median
    <= 3e-7 (float output rounding alone is ~6e-8) and p99 <= 1e-6, and the
    cancellation-free form
    |a_gpu - a_64| / sum_j |a_ij| <= 1e-6 for every particle (|a_i| -> 0 at
    the centre of a symmetric system makes e_i unbounded under any summation).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_forces(nbx, h, G, eps, variant=-1, scratch=0, workspace=True):
    """Forces through DirectForceCalculator's workspace (or the workspace-free launcher)."""
    d = nbx.DeviceParticles.from_host(h)
    ws = nbx.DirectWorkspace(len(h["mass"]), scratch) if workspace else None
    nbx.check(nbx.lib.nbx_direct_set_variant(variant))
    try:
        nbx.direct_forces(d, G, eps, ws=ws)
        nbx.synchronize()
        return d.acc().astype(np.float64)
    finally:
        nbx.check(nbx.lib.nbx_direct_set_variant(-1))


def abs_pair_sum(h, G, eps2, targets):
    p = np.stack([h["pos_x"], h["pos_y"], h["pos_z"]], 1).astype(np.float64)
    m = h["mass"].astype(np.float64)
    out = np.zeros(len(targets))
    for k, i in enumerate(targets):
        d = p - p[i]
        r2 = (d * d).sum(1) + eps2
        w = G * m * np.sqrt((d * d).sum(1)) * r2 ** -1.5
        w[i] = 0.0
        out[k] = w.sum()
    return out


def error_stats(got, ref):
    e = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return {"median": float(np.median(e)), "p99": float(np.quantile(e, 0.99)),
            "max": float(e.max())}


def check_fp64(nbx, orc, h, G, eps, targets=None, variant=-1, got=None):
    n = len(h["mass"])
    if got is None:
        got = run_forces(nbx, h, G, eps, variant)
    if targets is None:
        targets = np.arange(n) if n <= 4096 else np.random.default_rng(0).choice(n, 4096, False)
    eps2 = orc.eps2_of(eps)
    ref = orc.direct_f64(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], G, eps2, targets)
    st = error_stats(got[targets], ref)
    norm = np.linalg.norm(got[targets] - ref, axis=1) / abs_pair_sum(h, G, eps2, targets)
    print(f"n={n} eps={eps} variant={variant} rel-err {st} normalised max {norm.max():.3e}")
    assert st["median"] <= 3e-7, st
    assert st["p99"] <= 1e-6, st
    assert norm.max() <= 1e-6, norm.max()
    return st


def test_single_particle_zero(nbx):
    h = nbx.empty_host(1)
    h["pos_x"][0], h["pos_y"][0], h["pos_z"][0], h["mass"][0] = 1.0, 2.0, 3.0, 5.0
    a = run_forces(nbx, h, 1.0, 0.1)
    assert np.all(a == 0.0)


def test_two_body_equal_and_opposite(nbx):
    h = nbx.empty_host(2)
    h["pos_x"][:] = [-0.5, 0.5]
    h["mass"][:] = 1.0
    a = run_forces(nbx, h, 1.0, 0.01)
    assert np.allclose(a[0], -a[1], atol=1e-5) and a[0, 0] > 0


def test_grid32_matches_cpu_reference(nbx, orc):
    i = np.arange(32)
    h = nbx.empty_host(32)
    h["pos_x"][:] = ((i * 7) % 11).astype(np.float32) * np.float32(0.3)
    h["pos_y"][:] = ((i * 3) % 13).astype(np.float32) * np.float32(0.3)
    h["pos_z"][:] = ((i * 5) % 9).astype(np.float32) * np.float32(0.3)
    h["mass"][:] = np.float32(0.5) + np.float32(0.1) * (i % 5).astype(np.float32)
    got = run_forces(nbx, h, 1.0, 0.1)
    ref = orc.direct_f32(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 1.0, 0.1).astype(np.float64)
    assert np.all(np.abs(got - ref) <= 1e-3 * (np.abs(ref) + 1e-3))
    check_fp64(nbx, orc, h, 1.0, 0.1, got=got)


def test_plummer_4096_fp64(nbx, orc):
    h = nbx.gen_plummer(4096, seed=42)
    check_fp64(nbx, orc, h, 1.0, 0.01)


@pytest.mark.parametrize("n", [3, 63, 64, 65, 127, 129, 1000, 4097])
def test_ragged_sizes(nbx, orc, n):
    h = nbx.gen_plummer(n, seed=n)
    check_fp64(nbx, orc, h, 1.0, 0.01)


def mixed_masses(h, seed):
    """Masses drawn from [0.1, 2), as the reference initialisers do for min_mass < max_mass."""
    n = len(h["mass"])
    h["mass"][:] = np.random.default_rng(seed).uniform(0.1, 2.0, n).astype(np.float32) / np.float32(n)
    return h


@pytest.mark.parametrize("n", [4096, 20000])
def test_mixed_masses_fp64(nbx, orc, n):
    # per-pair mass factor (folds whose masses differ)
    h = mixed_masses(nbx.gen_plummer(n, seed=5), seed=n)
    check_fp64(nbx, orc, h, 1.0, 0.01)


@pytest.mark.parametrize("where", [0, 1500, 4095])
def test_one_heavy_body(nbx, orc, where):
    # equal masses except one: the j ranges holding it take the per-pair path,
    # the others apply the shared mass once per fold
    h = nbx.gen_plummer(4096, seed=11)
    h["mass"][where] = np.float32(50.0) * h["mass"][0]
    check_fp64(nbx, orc, h, 1.0, 0.01)


@pytest.mark.parametrize("masses", ["equal", "mixed"])
def test_every_variant_matches(nbx, orc, masses):
    h = nbx.gen_plummer(3000, seed=9)
    if masses == "mixed":
        mixed_masses(h, seed=3)
    names = nbx.variant_names()
    results = []
    for v in range(len(names)):
        got = run_forces(nbx, h, 0.7, 0.02, variant=v)
        check_fp64(nbx, orc, h, 0.7, 0.02, variant=v, got=got)
        results.append(got)
    # the same tile shape is deterministic
    again = run_forces(nbx, h, 0.7, 0.02, variant=len(names) - 1)
    assert np.array_equal(again, results[-1])


def test_zero_softening_skips_self(nbx, orc):
    # eps = 0 is rejected by the reference validators but the kernel must
    # still skip j == i by index (force_direct.cu:60) for distinct positions.
    h = nbx.gen_uniform_sphere(777, seed=4)
    got = run_forces(nbx, h, 1.0, 0.0)
    assert np.all(np.isfinite(got))
    t = np.arange(777)
    ref = orc.direct_f64(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 1.0, 0.0, t)
    st = error_stats(got, ref)
    assert st["p99"] <= 1e-5, st


def test_range_entry(nbx):
    h = nbx.gen_plummer(2000, seed=1)
    full = run_forces(nbx, h, 1.0, 0.01)
    d = nbx.DeviceParticles.from_host(h)
    nbx.check(nbx.lib.nbx_direct_forces_range(d.ptr, 300, 555, 1.0,
                                              float(np.float32(0.01) * np.float32(0.01)), None))
    nbx.synchronize()
    a = d.acc()
    assert np.array_equal(a[300:855], full[300:855].astype(np.float32))
    assert np.all(a[:300] == 0) and np.all(a[855:] == 0)


def test_invalid_arguments(nbx):
    d = nbx.DeviceParticles(16)
    with pytest.raises(nbx.NbxError):
        nbx.check(nbx.lib.nbx_direct_forces(d.ptr, 1.0, 0.01, 2048, None))
    with pytest.raises(nbx.NbxError):
        nbx.check(nbx.lib.nbx_direct_forces(d.ptr, 1.0, float("nan"), 256, None))
    empty = nbx.Particles()
    with pytest.raises(nbx.NbxError):
        nbx.check(nbx.lib.nbx_direct_forces(nbx.C.byref(empty), 1.0, 0.01, 256, None))


def test_uniform_sphere_262144_sample(nbx, orc):
    # the 262,144-body workload; 4,096 seeded sample targets vs the FP64 sum
    h = nbx.gen_uniform_sphere(262144, seed=42, R=1.0, sigma=0.05)
    t = np.sort(np.random.default_rng(42).choice(262144, 4096, replace=False))
    got = run_forces(nbx, h, 1.0, 0.005)
    check_fp64(nbx, orc, h, 1.0, 0.005, targets=t, got=got)


@pytest.mark.parametrize("n", [20000, 70000, 150000])
def test_default_shape_by_size_fp64(nbx, orc, n):
    # the default path at one size of each symmetric tier below the 262,144-body one
    # (sym_ipl4_ws1, sym_ipl6_ws1, sym_ipl8_ws1): 1,024 seeded sample targets vs FP64
    names = nbx.variant_names()
    expect = {20000: "sym_ipl4_ws1", 70000: "sym_ipl6_ws1", 150000: "sym_ipl8_ws1"}[n]
    assert names[nbx.lib.nbx_direct_variant_for(n)] == expect
    h = mixed_masses(nbx.gen_plummer(n, seed=n), n)
    t = np.sort(np.random.default_rng(n).choice(n, 1024, replace=False))
    got = run_forces(nbx, h, 1.0, 0.01)
    check_fp64(nbx, orc, h, 1.0, 0.01, targets=t, got=got)


# ---------------------------------------------------------------- pair-symmetric kernel
# Whole-system evaluations with each unordered pair computed once and applied to both
# bodies (direct_sym.hip). Same FP64 bounds as the one-sided kernel.

def sym_variants(nbx):
    return [(v, nm) for v, nm in enumerate(nbx.variant_names()) if nm.startswith("sym_")]


def sym_block(name):
    ipl, ws = (int(t[3:] if t.startswith("ipl") else t[2:]) for t in name.split("_")[1:])
    return 64 * ipl * ws


@pytest.mark.parametrize("blocks", [0.6, 1, 2, 2.4, 3, 4, 4.5, 7])
def test_sym_super_block_counts(nbx, orc, blocks):
    # K = ceil(n / S) super-blocks: one (diagonal only), even and odd K, ragged last block
    for v, name in sym_variants(nbx):
        n = max(3, int(blocks * sym_block(name)))
        if n > 20000:
            continue
        h = nbx.gen_plummer(n, seed=n + v)
        check_fp64(nbx, orc, h, 1.0, 0.01, variant=v)


@pytest.mark.parametrize("masses", ["mixed", "one_heavy"])
def test_sym_masses(nbx, orc, masses):
    for v, name in sym_variants(nbx):
        n = 3 * sym_block(name) + 77
        h = nbx.gen_plummer(n, seed=21)
        if masses == "mixed":
            mixed_masses(h, seed=v)
        else:
            h["mass"][n // 2] = np.float32(40.0) * h["mass"][0]
        check_fp64(nbx, orc, h, 1.0, 0.01, variant=v)


def test_sym_zero_softening(nbx, orc):
    # eps = 0: the diagonal tiles zero the self pair by index; off-diagonal tiles have none
    for v, name in sym_variants(nbx):
        n = 2 * sym_block(name) + 5
        h = nbx.gen_uniform_sphere(n, seed=8)
        got = run_forces(nbx, h, 1.0, 0.0, variant=v)
        assert np.all(np.isfinite(got)), name
        t = np.random.default_rng(1).choice(n, 512, False)
        ref = orc.direct_f64(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 1.0, 0.0, t)
        st = error_stats(got[t], ref)
        assert st["p99"] <= 1e-5, (name, st)


def test_sym_deterministic(nbx):
    for v, name in sym_variants(nbx):
        h = nbx.gen_uniform_sphere(3 * sym_block(name) + 1, seed=3)
        a = run_forces(nbx, h, 1.0, 0.01, variant=v)
        b = run_forces(nbx, h, 1.0, 0.01, variant=v)
        assert np.array_equal(a, b), name


def test_sym_kick_epilogue(nbx):
    # v_new = v + (a_old + a_new) * (0.5 dt), IEEE single rounding, with a_new as returned
    dt = np.float32(1e-3)
    for v, name in sym_variants(nbx):
        n = 2 * sym_block(name) + 300
        h = nbx.gen_plummer(n, seed=17)
        rng = np.random.default_rng(2)
        for f in ("acc_old_x", "acc_old_y", "acc_old_z"):
            h[f][:] = rng.normal(size=n).astype(np.float32)
        d = nbx.DeviceParticles.from_host(h)
        for f in ("acc_old_x", "acc_old_y", "acc_old_z"):
            d.write(f, h[f])  # upload() leaves acc_old alone, as the reference copy does
        ws = nbx.DirectWorkspace(n)
        nbx.check(nbx.lib.nbx_direct_set_variant(v))
        try:
            ws.forces_kick(d, 1.0, 0.01, float(dt))
            nbx.synchronize()
        finally:
            nbx.check(nbx.lib.nbx_direct_set_variant(-1))
        half = np.float32(np.float32(0.5) * dt)
        for c in "xyz":
            a_new = d.read(f"acc_{c}")
            s = (h[f"acc_old_{c}"] + a_new).astype(np.float32)
            want = (h[f"vel_{c}"] + (s * half).astype(np.float32)).astype(np.float32)
            assert np.array_equal(d.read(f"vel_{c}"), want), (name, c)


def test_sym_range_call_uses_one_sided(nbx):
    # target sub-ranges are evaluated by the one-sided kernel whatever variant is forced
    v, name = sym_variants(nbx)[0]
    h = nbx.gen_plummer(2 * sym_block(name), seed=6)
    eps2 = float(np.float32(0.01) * np.float32(0.01))
    res = []
    for forced in (v, -1):
        d = nbx.DeviceParticles.from_host(h)
        nbx.check(nbx.lib.nbx_direct_set_variant(forced))
        try:
            nbx.check(nbx.lib.nbx_direct_forces_range(d.ptr, 100, 900, 1.0, eps2, None))
            nbx.synchronize()
        finally:
            nbx.check(nbx.lib.nbx_direct_set_variant(-1))
        res.append(d.acc())
    assert np.array_equal(res[0], res[1])


def test_workspace_free_launcher_runs_one_sided(nbx):
    # nbx_direct_forces takes no workspace: a forced symmetric shape falls back to the
    # one-sided default, bit for bit
    n = 70000
    h = nbx.gen_plummer(n, seed=2)
    sym = [v for v, _ in sym_variants(nbx)][1]
    a = run_forces(nbx, h, 1.0, 0.01, variant=sym, workspace=False)
    b = run_forces(nbx, h, 1.0, 0.01, variant=3, workspace=True)  # ipl1_nb8_ws8 at 2^16..2^17
    assert np.array_equal(a, b)


@pytest.mark.parametrize("blocks", [3, 4.5, 8, 11])
def test_sym_rounds_bit_identical(nbx, blocks):
    # partial scratch of 3, 5 and 8 slots gives rounds of 1, 2 and 3 tile distances;
    # every round split sums each body's partials in the same order
    for v, name in sym_variants(nbx):
        S = sym_block(name)
        n = int(blocks * S) + 17
        if n > 24000:
            continue
        h = nbx.gen_uniform_sphere(n, seed=v)
        whole = run_forces(nbx, h, 1.0, 0.01, variant=v)
        # the workspace's padded length: n rounded up to each symmetric shape's
        # super-block, the largest of them
        npad = max(-(-n // sym_block(nm)) * sym_block(nm) for _, nm in sym_variants(nbx))
        for slots in (3, 5, 8):
            ws = nbx.DirectWorkspace(n, slots * 3 * npad * 4)
            nbx.check(nbx.lib.nbx_direct_set_variant(v))
            try:
                plan = ws.plan(n)
                d = nbx.DeviceParticles.from_host(h)
                ws.forces(d, 1.0, 0.01)
                nbx.synchronize()
            finally:
                nbx.check(nbx.lib.nbx_direct_set_variant(-1))
            K = -(-n // S)
            if slots * npad // (K * S) < 2 * (K // 2) + 1:
                assert plan["rounds"] > 1, (name, slots, plan)
            assert np.array_equal(d.acc().astype(np.float64), whole), (name, n, slots)


def test_workspace_errors(nbx):
    with pytest.raises(nbx.NbxError):
        nbx.DirectWorkspace(0)
    with pytest.raises(nbx.NbxError):
        nbx.DirectWorkspace(100000, 1000)  # below three partial slots
    ws = nbx.DirectWorkspace(1000)
    d = nbx.DeviceParticles(1001)
    with pytest.raises(nbx.NbxError):
        ws.forces(d, 1.0, 0.01)
    with pytest.raises(nbx.NbxError):
        ws.plan(1001)
    assert ws.plan(1000)["rounds"] == 1


def test_sym_one_million_in_rounds(nbx, orc):
    # 1,048,576 bodies: the default workspace (1 GiB of partials) runs the
    # pair-symmetric tiles in rounds; 4,096 sample targets vs the FP64 sum
    n = 1 << 20
    h = nbx.gen_uniform_sphere(n, seed=42, R=1.0, sigma=0.05)
    ws = nbx.DirectWorkspace(n)
    plan = ws.plan(n)
    print(f"1M plan: {plan}", flush=True)
    assert nbx.variant_names()[plan["variant"]].startswith("sym_")
    assert plan["rounds"] > 1 and plan["partial_bytes"] <= 1 << 30
    d = nbx.DeviceParticles.from_host(h)
    ws.forces(d, 1.0, 0.005)
    nbx.synchronize()
    t = np.sort(np.random.default_rng(7).choice(n, 4096, replace=False))
    ref = orc.direct_f64(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 1.0, orc.eps2_of(0.005), t)
    st = error_stats(d.acc().astype(np.float64)[t], ref)
    print(f"1M in rounds: {st}", flush=True)
    assert st["median"] <= 3e-7 and st["p99"] <= 1e-6
