"""GPU parity of the Barnes-Hut pipeline against the CPU oracle.

Bit-exact: bounding box, Morton codes, the stable sort permutation, and the
whole 76-byte node array (numbering, centres, half sizes, centres of mass,
masses, children, particle indices, leaf flags, counts), node count and max
depth. Forces: the traversal makes the same accept/open decisions as the
oracle (same IEEE arithmetic), so accelerations differ only by v_rsq_f32 vs
1/sqrtf and FMA accumulation: |a_gpu - a_orc| / sum|terms| <= 2e-6 per particle.
"""
import numpy as np
import pytest

import bh_model as M
from test_bh_oracle import CASES, tree_inputs

pytestmark = pytest.mark.gpu


def gpu_build(nbx, h, capacity=None, builder=0):
    n = len(h["mass"])
    d = nbx.DeviceParticles.from_host(h)
    t = nbx.BHTree(capacity or n)
    t.set_builder(builder)
    t.build(d)
    return d, t


def check_exact(nbx, orc, h, capacity=None, builder=0):
    n = len(h["mass"])
    d, t = gpu_build(nbx, h, capacity, builder)
    ref = orc.bh_pipeline(h, capacity)
    lo, hi, codes, idx = t.sort_result(n)
    assert np.array_equal(lo, ref["bbox_min"]) and np.array_equal(hi, ref["bbox_max"])
    assert np.array_equal(codes, ref["codes"]), "Morton codes / sort order differ"
    assert np.array_equal(idx, ref["sorted_idx"]), "sort permutation differs"
    cnt, depth = t.node_count()
    assert cnt == ref["node_count"] and depth == ref["max_depth"], (cnt, depth, ref["node_count"], ref["max_depth"])
    ok, why = M.same_nodes(t.nodes(), ref["nodes"])
    assert ok, why
    return d, t, ref


@pytest.mark.parametrize("kind,n", CASES)
def test_tree_bit_exact(nbx, orc, kind, n):
    _, t, ref = check_exact(nbx, orc, tree_inputs(nbx, kind, n))
    info = t.build_info()
    expect_seq = ref["node_count"] + 33 >= 2 * n
    assert info["builder"] == ("sequential" if expect_seq else "parallel"), info


def test_parallel_builder_handles_depth32_cells(nbx, orc):
    h = tree_inputs(nbx, "pairs", 4000)
    _, t, ref = check_exact(nbx, orc, h)
    assert ref["depth_merges"] == 6 and ref["max_depth"] == 32
    info = t.build_info()
    assert info["builder"] == "parallel" and info["depth32_leaves"] == 2


def test_sequential_builder_agrees_with_parallel(nbx, orc):
    h = nbx.gen_plummer(5000, seed=12)
    _, tp = gpu_build(nbx, h)
    _, ts = gpu_build(nbx, h, builder=1)
    assert tp.build_info()["builder"] == "parallel"
    assert ts.build_info()["builder"] == "sequential"
    ok, why = M.same_nodes(tp.nodes(), ts.nodes())
    assert ok, why


@pytest.mark.parametrize("n,capacity", [(700, 1000), (3000, 3000)])
def test_capacity_larger_than_n(nbx, orc, n, capacity):
    # the node budget is 2 * the capacity the tree was created with (:252-258)
    check_exact(nbx, orc, nbx.gen_plummer(n, seed=n), capacity=capacity)


def test_two_plummer_65536_exact(nbx, orc):
    check_exact(nbx, orc, nbx.gen_two_plummer(65536, seed=42))


def clustered(nbx, n_bg, n_cl, side, seed):
    """A Plummer sphere whose last n_cl bodies sit in a cube of the given side."""
    h = nbx.gen_plummer(n_bg + n_cl, seed=seed)
    rng = np.random.default_rng(seed)
    for a, k in enumerate(("pos_x", "pos_y", "pos_z")):
        off = rng.uniform(0, side, n_cl).astype(np.float32)
        h[k][n_bg:] = (np.float32((0.3, -0.2, 0.1)[a]) + off).astype(np.float32)
    return h


def prefix_runs(h, ref, digits=20):
    pos = np.stack([h["pos_x"], h["pos_y"], h["pos_z"]], 1)
    return M.shared_prefix_runs(pos, ref["sorted_idx"], ref["bbox_min"], ref["bbox_max"], digits)


def test_runs_sharing_sorted_digits(nbx, orc):
    # the path sort orders by the first 20 octant digits; bodies closer than that
    # resolves (here 192 and 108 of them, leaves down to depth 27) are ordered
    # by their remaining digits in one workgroup per run
    h = clustered(nbx, 5000, 300, 4e-5, seed=7)
    _, t, ref = check_exact(nbx, orc, h)
    runs = prefix_runs(h, ref)
    assert 2 <= runs[0] <= 1024 and ref["max_depth"] > 20, (runs[:4], ref["max_depth"])
    assert t.build_info()["builder"] == "parallel"


@pytest.mark.parametrize("n_cl,clusters,side,seed", [
    (1025, 1, 1e-5, 10),  # one more than LDS holds
    (1025, 1, 1e-7, 9),   # mostly exact duplicates: depth-32 leaves, ties kept in rank order
    (1500, 1, 1e-5, 8), (5000, 1, 1e-5, 8),
    (1800, 3, 1e-6, 8),   # three long runs in one build
])
def test_runs_longer_than_a_workgroup_stay_parallel(nbx, orc, n_cl, clusters, side, seed):
    # more bodies share their first 20 digits than one workgroup orders in LDS:
    # path_refine_long sorts each such run (one workgroup per run, through
    # global scratch) and the parallel builder keeps the build, bit-exact
    h = clustered(nbx, 3000, n_cl * clusters, side, seed=seed)
    for c in range(1, clusters):  # further clusters elsewhere in the sphere
        sl = slice(3000 + c * n_cl, 3000 + (c + 1) * n_cl)
        h["pos_x"][sl] = (h["pos_x"][sl] + np.float32(0.4 * c)).astype(np.float32)
    _, t, ref = check_exact(nbx, orc, h)
    runs = prefix_runs(h, ref)
    assert sum(r > 1024 for r in runs) >= clusters, runs[:8]
    assert t.build_info()["builder"] == "parallel"


def test_one_million_with_dense_cluster_stays_parallel(nbx, orc):
    # the 1,048,576-body collision workload with 50,000 of its bodies moved into
    # a cube of side 1e-5 (runs of tens of thousands sharing 20 digits, leaves to
    # depth ~30): parallel build, node array bit-exact, forces vs the oracle
    n, n_cl = 1 << 20, 50000
    h = nbx.gen_two_plummer(n, seed=42)
    rng = np.random.default_rng(3)
    for a, k in enumerate(("pos_x", "pos_y", "pos_z")):
        h[k][-n_cl:] = (np.float32((0.3, -0.2, 0.1)[a]) +
                        rng.uniform(0, 1e-5, n_cl).astype(np.float32)).astype(np.float32)
    d, t, ref = check_exact(nbx, orc, h)
    assert t.build_info()["builder"] == "parallel"
    cl = np.stack([h["pos_x"][-n_cl:], h["pos_y"][-n_cl:], h["pos_z"][-n_cl:]], 1)
    runs = M.shared_prefix_runs(cl, np.arange(n_cl), ref["bbox_min"], ref["bbox_max"], 20)
    assert runs[0] > 1024 and ref["max_depth"] > 25, (runs[:4], ref["max_depth"])
    t.forces(d, 0.5, 1.0, 0.01)
    nbx.synchronize()
    tg = np.sort(np.concatenate([np.random.default_rng(4).choice(n - n_cl, 512, replace=False),
                                 n - n_cl + np.arange(0, n_cl, n_cl // 512)]))
    got = d.acc().astype(np.float64)[tg]
    a = orc.bh_forces(ref["nodes"], h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 0.5, 1.0, 0.01,
                      targets=tg).astype(np.float64)
    err = np.linalg.norm(got - a, axis=1) / np.linalg.norm(a, axis=1)
    print(f"1M with a 50,000-body cluster: depth {ref['max_depth']}, nodes {ref['node_count']}, "
          f"force rel err median {np.median(err):.2e} max {err.max():.2e}", flush=True)
    assert np.median(err) < 1e-6 and err.max() < 1e-4


def test_block_with_more_cells_than_lds_holds(nbx, orc):
    # 300 pairs 2e-7 apart inside a 3e-4 cube: their chains put over 2,048 cells
    # in some blocks of 1,024 positions, which the bottom-up pass then reduces in
    # global memory instead of LDS
    h = nbx.gen_plummer(10600, seed=5)
    rng = np.random.default_rng(5)
    base = (rng.uniform(0, 3e-4, (300, 3)).astype(np.float32) + np.float32(0.2)).astype(np.float32)
    for a, k in enumerate(("pos_x", "pos_y", "pos_z")):
        h[k][10000::2] = base[:, a]
        h[k][10001::2] = (base[:, a] + np.float32(2e-7)).astype(np.float32)
    _, t, ref = check_exact(nbx, orc, h)
    assert t.build_info()["builder"] == "parallel" and ref["max_depth"] == 27


def test_plummer_4m_three_levels_of_straddling_cells(nbx, orc):
    # 4,300,000 bodies: 4,200 blocks, so straddling cells reach the third group
    # level (runs longer than 64 x 64 blocks); the node array is bit-exact
    check_exact(nbx, orc, nbx.gen_plummer(4_300_000, seed=9))


def force_check(nbx, orc, h, theta, G, eps, targets=None):
    d, t = gpu_build(nbx, h)
    t.forces(d, theta, G, eps)
    nbx.synchronize()
    got = d.acc().astype(np.float64)
    ref = orc.bh_pipeline(h)
    n = len(h["mass"])
    tg = np.arange(n) if targets is None else targets
    a = orc.bh_forces(ref["nodes"], h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], theta, G, eps,
                      targets=tg).astype(np.float64)
    # normalise by the sum of |pair terms| bound: G * sum_j m_j / (r^2+eps^2) of the direct sum
    p = np.stack([h["pos_x"], h["pos_y"], h["pos_z"]], 1).astype(np.float64)
    m = h["mass"].astype(np.float64)
    scale = np.array([G * np.sum(m * np.sqrt(((p - p[i]) ** 2).sum(1)) /
                                 (((p - p[i]) ** 2).sum(1) + eps * eps) ** 1.5) for i in tg])
    err = np.linalg.norm(got[tg] - a, axis=1) / scale
    print(f"n={n} theta={theta}: max normalised |gpu - oracle| = {err.max():.2e}")
    assert err.max() <= 2e-6
    return got, d, t


@pytest.mark.parametrize("theta", [0.0, 0.3, 0.5, 1.0])
def test_forces_match_oracle(nbx, orc, theta):
    force_check(nbx, orc, nbx.gen_plummer(3000, seed=3), theta, 1.0, 0.01)


def test_forces_degenerate_inputs(nbx, orc):
    for kind, n in (("coincident", 16), ("collinear", 16), ("pairs", 400), ("plummer", 1)):
        h = tree_inputs(nbx, kind, n)
        d, t = gpu_build(nbx, h)
        t.forces(d, 0.5, 1.0, 0.1)
        nbx.synchronize()
        got = d.acc()
        ref = orc.bh_pipeline(h)
        a = orc.bh_forces(ref["nodes"], h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 0.5, 1.0, 0.1)
        assert np.all(np.isfinite(got))
        assert np.allclose(got, a, rtol=1e-5, atol=1e-6), kind
    # N=1: the particle's own leaf is skipped (m_eff = 0): zero acceleration
    h = tree_inputs(nbx, "plummer", 1)
    d, t = gpu_build(nbx, h)
    t.forces(d, 0.5, 1.0, 0.1)
    nbx.synchronize()
    assert np.all(d.acc() == 0)


def theta_on_boundary(h, nodes, i, eps, rng):
    """A theta with fl(theta * theta) == fl(size2 / d2) for an internal cell that
    particle i's walk must visit (a sibling of one of i's ancestors): the
    decision size2 / d2 < theta2 (force_barnes_hut.cu:218-224) is then false by
    equality, and true for the next float theta. Values are computed in float32
    in the traversal's operation order."""
    f = np.float32
    x = np.array([h["pos_x"][i], h["pos_y"][i], h["pos_z"][i]], f)
    eps2 = f(eps) * f(eps)
    cand, v = [], 0
    while not nodes["is_leaf"][v]:
        nxt = None
        for c in nodes["children"][v]:
            if c < 0:
                continue
            if np.all(np.abs(x - nodes["center"][c]) <= nodes["half_size"][c]):
                nxt = c if nxt is None else nxt
            elif not nodes["is_leaf"][c] and nodes["total_mass"][c] > 0:
                cand.append(int(c))
        if nxt is None:
            break
        v = nxt
    rng.shuffle(cand)
    for c in cand:
        dd = (nodes["com"][c] - x).astype(f)
        d2 = f(f(f(dd[0] * dd[0]) + f(dd[1] * dd[1])) + f(dd[2] * dd[2])) + eps2
        hh = nodes["half_size"][c]
        r = f(f(f(4.0) * hh) * hh / d2)
        if not 0.04 < r < 1.0:
            continue
        t = f(np.sqrt(r))
        for _ in range(4):
            t = np.nextafter(t, f(0))
        for _ in range(9):
            if f(t * t) == r:
                return t, c
            t = np.nextafter(t, f(1))
    return None, None


def test_theta_decisions_on_the_boundary(nbx, orc):
    # A decision exactly on the boundary goes through the division in the walk
    # (the thresholds S_open/S_acc leave it undecided); one float up, the
    # opposite decision. Cases are kept where the flip moves the oracle's force
    # on particle i by > 1e-3 relative, 50x the tolerance of the comparison.
    h = nbx.gen_plummer(3000, seed=21)
    ref = orc.bh_pipeline(h)["nodes"]
    rng = np.random.default_rng(5)
    found = 0
    for i in rng.choice(3000, 400, replace=False):
        th, c = theta_on_boundary(h, ref, int(i), 0.01, rng)
        if th is None:
            continue
        thetas = (float(th), float(np.nextafter(th, np.float32(1))))
        want = [orc.bh_forces(ref, h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], theta, 1.0,
                              0.01, targets=np.array([i])).astype(np.float64)[0]
                for theta in thetas]
        if np.linalg.norm(want[0] - want[1]) < 1e-3 * np.linalg.norm(want[0]):
            continue
        for theta, a in zip(thetas, want):
            d, t = gpu_build(nbx, h)
            t.forces(d, theta, 1.0, 0.01)
            nbx.synchronize()
            g = d.acc()[i].astype(np.float64)
            assert np.linalg.norm(g - a) <= 2e-5 * np.linalg.norm(a), (i, c, theta)
        found += 1
        if found == 3:
            break
    assert found == 3


def test_force_calls_reuse_one_build(nbx, orc):
    # Per-call thresholds and G * m: rewritten when (theta, G) changes, written by
    # the build for the last (theta, G). Any order of calls and rebuilds gives
    # the accelerations of a fresh build.
    h = nbx.gen_plummer(5000, seed=11)
    d, t = gpu_build(nbx, h)

    def acc(tree, theta, G):
        tree.forces(d, theta, G, 0.01)
        nbx.synchronize()
        return d.acc().copy()

    a1 = acc(t, 0.5, 1.0)
    a2 = acc(t, 1.0, 1.0)
    a3 = acc(t, 0.5, 1.0)
    a4 = acc(t, 0.5, 2.0)
    assert np.array_equal(a1.view(np.uint32), a3.view(np.uint32))
    _, t2 = gpu_build(nbx, h)
    assert np.array_equal(acc(t2, 1.0, 1.0).view(np.uint32), a2.view(np.uint32))
    # G = 2 scales every rounded product by a power of two: exactly twice
    assert np.array_equal(a4.view(np.uint32), (2 * a1).view(np.uint32))
    # a rebuild writes the fields for the last (theta, G) itself
    t.build(d)
    assert np.array_equal(acc(t, 0.5, 2.0).view(np.uint32), a4.view(np.uint32))
    t.build(d)
    assert np.array_equal(acc(t, 0.5, 1.0).view(np.uint32), a1.view(np.uint32))


def test_theta_zero_equals_direct(nbx, orc):
    h = nbx.gen_plummer(2000, seed=8)
    d, t = gpu_build(nbx, h)
    t.forces(d, 0.0, 1.0, 0.01)
    nbx.synchronize()
    got = d.acc().astype(np.float64)
    ref = orc.direct_f64(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 1.0, orc.eps2_of(0.01))
    e = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(e) < 1e-6 and e.max() < 1e-4


def test_rebuild_and_errors(nbx, orc):
    h = nbx.gen_plummer(2000, seed=1)
    d, t = gpu_build(nbx, h, capacity=2000)
    h2 = nbx.gen_plummer(2000, seed=2)
    d.upload(h2)
    t.build(d)
    ok, why = M.same_nodes(t.nodes(), orc.bh_pipeline(h2)["nodes"])
    assert ok, why
    big = nbx.DeviceParticles(2001)
    with pytest.raises(nbx.NbxError):
        t.build(big)
    t2 = nbx.BHTree(100)
    small = nbx.DeviceParticles.from_host(nbx.gen_plummer(100))
    with pytest.raises(nbx.NbxError):
        t2.forces(small, 0.5, 1.0, 0.01)


def test_two_plummer_1m_exact_and_forces(nbx, orc):
    # the 1,048,576-body collision workload: Morton codes, permutation and the
    # full node array bit-exact; forces on 2,048 sample targets vs the oracle
    h = nbx.gen_two_plummer(1 << 20, seed=42)
    d, t, ref = check_exact(nbx, orc, h)
    t.forces(d, 0.5, 1.0, 0.01)
    nbx.synchronize()
    got = d.acc().astype(np.float64)
    tg = np.sort(np.random.default_rng(1).choice(1 << 20, 2048, replace=False))
    a = orc.bh_forces(ref["nodes"], h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 0.5, 1.0, 0.01,
                      targets=tg).astype(np.float64)
    err = np.linalg.norm(got[tg] - a, axis=1) / np.linalg.norm(a, axis=1)
    print(f"1M two-Plummer: nodes {ref['node_count']}, depth {ref['max_depth']}, "
          f"force rel err median {np.median(err):.2e} max {err.max():.2e}")
    assert np.median(err) < 1e-6 and err.max() < 1e-4


def test_node_budget_exhausted_is_reported(nbx, orc):
    # the input on which the reference's insertion writes past its 2N nodes
    # (tests/bh_model.py budget_overrun_input): the sequential builder stops at the
    # budget and every synchronous query reports NBX_ERR_CAPACITY
    from test_bh_oracle import overrun_host
    h = overrun_host(nbx)
    n = len(h["mass"])
    d, t = gpu_build(nbx, h)
    for query in (t.node_count, t.build_info, t.nodes):
        with pytest.raises(nbx.NbxError) as e:
            query()
        assert e.value.status == 4 and "node budget" in str(e.value)
    # with room for four times the bodies the same input builds bit-exact
    check_exact(nbx, orc, h, capacity=4 * n)


def test_walk_census_closed_form_at_theta_zero(nbx, orc):
    # theta = 0 opens every cell with mass, so each wave visits every record:
    # wave steps = tasks * nodes, openings = n * internal cells, force evaluations
    # = n * leaves - n (a body's own single-body leaf gives no force), active
    # lane-steps = n * nodes. 3,001 bodies: a last wave with 57 lanes.
    h = nbx.gen_plummer(3001, seed=5)
    d, t = gpu_build(nbx, h)
    nodes = t.nodes()
    n, nn = 3001, len(nodes)
    leaves = int(nodes["is_leaf"].sum())
    assert np.all(nodes["particle_count"][nodes["is_leaf"] == 1] == 1)
    c = t.walk_census(d, 0.0, 1.0, 0.01)
    assert c["wave_steps"] == -(-n // 64) * nn
    assert c["openings"] == n * (nn - leaves)
    assert c["interactions"] == n * leaves - n
    assert c["active_lane_steps"] == n * nn
    # at theta = 0.5 the walk does less, and the census agrees with itself
    c5 = t.walk_census(d, 0.5, 1.0, 0.01)
    assert 0 < c5["wave_steps"] < c["wave_steps"] and c5["interactions"] < c["interactions"]
    # each active lane at a step opens the node, takes its force or skips its own leaf
    assert c5["active_lane_steps"] >= c5["interactions"] + c5["openings"]
    assert c5["active_lane_steps"] - c5["interactions"] - c5["openings"] <= n * 40
    # forces after a census are the walk's own (the census shares its prepared records)
    t.forces(d, 0.5, 1.0, 0.01)
    nbx.synchronize()
    ref = orc.bh_forces(nodes, h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 0.5, 1.0, 0.01)
    e = np.linalg.norm(d.acc().astype(np.float64) - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(e) < 1e-6


def test_forces_do_not_depend_on_task_order(nbx, orc):
    # Tasks of 64 particles in Hilbert order (the default) or in the tree's own
    # order: every lane sums its interactions in the records' order and every force
    # term uses the same d2, so the accelerations agree bit for bit; the Hilbert
    # order's tasks are more compact, so its waves visit fewer records.
    h = nbx.gen_two_plummer(1 << 17, seed=21)
    d = nbx.DeviceParticles.from_host(h)
    t = nbx.BHTree(1 << 17)
    res = {}
    for mode in (1, 0):
        t.set_task_order(mode)
        t.build(d)
        t.forces(d, 0.5, 1.0, 0.01)
        nbx.synchronize()
        res[mode] = (d.acc().copy(), t.walk_census(d, 0.5, 1.0, 0.01))
    assert np.array_equal(res[0][0], res[1][0])
    c0, c1 = res[0][1], res[1][1]
    assert c0["interactions"] == c1["interactions"] and c0["openings"] == c1["openings"]
    assert c0["active_lane_steps"] == c1["active_lane_steps"]
    assert c0["wave_steps"] < c1["wave_steps"]
    ref = orc.bh_pipeline(h)
    sel = np.random.default_rng(3).choice(1 << 17, 512, replace=False)
    sel = np.sort(sel)
    a = orc.bh_forces(ref["nodes"], h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 0.5, 1.0, 0.01,
                      targets=sel)
    e = np.linalg.norm(res[0][0][sel].astype(np.float64) - a, axis=1) / np.linalg.norm(a, axis=1)
    assert np.median(e) < 1e-6
    with pytest.raises(nbx.NbxError):
        t.set_task_order(2)
