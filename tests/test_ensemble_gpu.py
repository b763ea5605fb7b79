"""GPU checks of the batched ensemble.

Each system's accelerations from the persistent batched launch are
bit-identical to the single-system kernel with the same tile shape (both run
direct_tile() on the same inputs in the same order), so the batched form meets
any force-error bound the standalone path meets (<= 1e-6 vs FP64). Steps
replayed from a captured hipGraph equal eagerly launched steps bit for bit,
over several replays with host work and synchronisations in between.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G, EPS, DT = 1.0, 0.01, 1e-3


def e2():
    return float(np.float32(np.float32(EPS) * np.float32(EPS)))


@pytest.fixture(scope="module")
def small(nbx):
    counts = nbx.gen_ensemble_sizes(24, seed=7, nmin=64, nmax=3000)
    counts[0] = 1  # a single-particle system
    counts[1] = 65  # ragged tile
    h = nbx.gen_ensemble(counts, seed=100)
    return counts, h


def standalone(nbx, h, sl, variant):
    sub = {k: np.ascontiguousarray(v[sl]) for k, v in h.items()}
    d = nbx.DeviceParticles.from_host(sub)
    ws = nbx.DirectWorkspace(len(sub["mass"]))
    nbx.check(nbx.lib.nbx_direct_set_variant(variant))
    try:
        nbx.direct_forces(d, G, EPS, ws=ws)
        nbx.synchronize()
    finally:
        nbx.check(nbx.lib.nbx_direct_set_variant(-1))
    return d.acc()


# one-sided shapes 0, 4, 6, 7 and the pair-symmetric shapes 8 (sym_ipl8_ws4),
# 9 (sym_ipl4_ws4), 12 (sym_ipl4_ws2, 512-body super-blocks) and 15
# (sym_ipl4_ws1, 256): the systems of 1, 65 and up to 3000 bodies give 1 to 12
# super-blocks, so diagonal-only, even and odd tile sets all occur
@pytest.mark.parametrize("variant", [0, 4, 6, 7, 8, 9, 12, 15])
def test_batched_forces_bit_identical_to_standalone(nbx, small, variant):
    counts, h = small
    ens = nbx.Ensemble(counts)
    ens.upload(h)
    ens.set_variant(variant)
    ens.forces(G, e2())
    nbx.synchronize()
    acc = np.stack([ens.read("acc_x"), ens.read("acc_y"), ens.read("acc_z")], 1)
    for q in range(len(counts)):
        ref = standalone(nbx, h, ens.system(q), variant)
        assert np.array_equal(acc[ens.system(q)], ref), f"system {q} (n={counts[q]})"


@pytest.mark.parametrize("variant", [4, 9, 15])
def test_batched_forces_vs_fp64(nbx, orc, small, variant):
    counts, h = small
    ens = nbx.Ensemble(counts)
    ens.upload(h)
    ens.set_variant(variant)
    ens.forces(G, e2())
    nbx.synchronize()
    acc = np.stack([ens.read("acc_x"), ens.read("acc_y"), ens.read("acc_z")], 1).astype(np.float64)
    worst = 0.0
    for q in range(2, len(counts)):
        sl = ens.system(q)
        ref = orc.direct_f64(h["pos_x"][sl], h["pos_y"][sl], h["pos_z"][sl], h["mass"][sl], G, e2())
        e = np.linalg.norm(acc[sl] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        worst = max(worst, float(np.quantile(e, 0.99)))
    print(f"worst per-system p99 rel err vs FP64: {worst:.2e}")
    assert worst <= 1e-6
    assert np.all(acc[ens.system(0)] == 0)  # lone particle


def test_variant_names_match_single_system(nbx):
    # a batched shape index names the same tile shape as the single-system one
    assert nbx.lib.nbx_ens_variant_count() == nbx.lib.nbx_direct_variant_count()
    assert nbx.variant_names()[9] == "sym_ipl4_ws4" and nbx.variant_names()[8] == "sym_ipl8_ws4"
    assert nbx.variant_names()[12] == "sym_ipl4_ws2" and nbx.variant_names()[15] == "sym_ipl4_ws1"
    assert nbx.variant_names()[18] == "sym_ipl6_ws1" and nbx.variant_names()[19] == "sym_ipl8_ws1"


@pytest.mark.parametrize("variant", [2, 9])
def test_graph_steps_equal_eager_and_sequential(nbx, small, variant):
    counts, h = small
    a, b, c = nbx.Ensemble(counts), nbx.Ensemble(counts), nbx.Ensemble(counts)
    for e in (a, b, c):
        e.upload(h)
        e.set_variant(variant)
        e.forces(G, e2())
    nbx.synchronize()
    s = nbx.Stream()
    # graph replays interleaved with host reads and synchronisations
    for rep in range(3):
        a.steps(G, e2(), DT, 4, graph_chunk=4, stream=s)
        s.synchronize()
        b.steps(G, e2(), DT, 4, graph_chunk=0, stream=s)
        s.synchronize()
        for f in ("pos_x", "vel_z", "acc_y"):
            assert np.array_equal(a.read(f), b.read(f)), (rep, f)
    nbx.check(nbx.lib.nbx_direct_set_variant(variant))
    try:
        c.steps_sequential(G, e2(), DT, 12, stream=s)
        s.synchronize()
    finally:
        nbx.check(nbx.lib.nbx_direct_set_variant(-1))
    for f in ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "acc_x"):
        va, vb, vc = a.read(f), b.read(f), c.read(f)
        assert np.array_equal(va, vb), f
        assert np.array_equal(va, vc), f


def test_queue_resets_between_launches(nbx, small):
    counts, h = small
    ens = nbx.Ensemble(counts)
    ens.upload(h)
    ens.forces(G, e2())
    nbx.synchronize()
    first = ens.read("acc_y")
    for _ in range(5):
        ens.forces(G, e2())
    nbx.synchronize()
    assert np.array_equal(ens.read("acc_y"), first)


def test_energy_per_system(nbx, orc, small):
    counts, h = small
    ens = nbx.Ensemble(counts)
    ens.upload(h)
    ke, pe = ens.energy(G, e2())
    for q in (1, 5, len(counts) - 1):
        sl = ens.system(q)
        sub = {k: np.ascontiguousarray(v[sl]) for k, v in h.items()}
        r = orc.energy_f64(sub, G, e2())
        assert abs(ke[q] - r["ke"]) <= 1e-12 * abs(r["ke"])
        assert abs(pe[q] - r["pe"]) <= 1e-11 * abs(r["pe"])
    assert ke[0] >= 0 and pe[0] == 0


def test_autotune_and_bounds(nbx, small):
    counts, h = small
    ens = nbx.Ensemble(counts)
    ens.upload(h)
    v, ms = ens.autotune(2)
    assert 0 <= v < nbx.lib.nbx_ens_variant_count() and ms > 0
    with pytest.raises(nbx.NbxError):
        ens.set_variant(99)
    with pytest.raises(nbx.NbxError):
        ens.steps(G, e2(), DT, 5, graph_chunk=2)
