"""Pins of the CPU oracle's Barnes-Hut pieces (CPU only).

Morton codes: exhaustive expandBits check and a naive bit interleave.
Sort: numpy's stable argsort. Tree: an independent pure-Python restatement
of the insertion build, PR-octree invariants, the reference's mass
conservation KATs (tests/test_barnes_hut.cpp), and the numbering rule the GPU
builder relies on. Forces: theta -> 0 equals the direct sum; the reference's
2% DirectVsBarnesHut KAT and its convergence KAT.
"""
import numpy as np
import pytest

import bh_model as M


def naive_interleave(ix, iy, iz):
    code = 0
    for b in range(10):
        code |= ((ix >> b) & 1) << (3 * b + 2)
        code |= ((iy >> b) & 1) << (3 * b + 1)
        code |= ((iz >> b) & 1) << (3 * b)
    return code


def naive_spread(v):
    out = 0
    for b in range(10):
        out |= ((v >> b) & 1) << (3 * b)
    return out


def test_expand_bits_exhaustive(orc):
    for v in range(1024):
        assert orc.lib.orc_expand_bits(v) == naive_spread(v)


def test_morton_code_matches_naive(orc):
    rng = np.random.default_rng(0)
    pts = rng.uniform(-0.2, 1.2, size=(2000, 3)).astype(np.float32)
    pts[:5] = [[0, 0, 0], [1, 1, 1], [0.5, 0.25, 0.125], [-1, 2, 0.999], [1e-8, 0.9995, 1.0]]
    for p in pts:
        q = [int(np.float32(min(max(v, np.float32(0)), np.float32(1))) * np.float32(1023)) for v in p]
        assert orc.lib.orc_morton_code(*map(float, p)) == naive_interleave(*q)


def test_bbox_and_codes(orc, nbx):
    h = nbx.gen_plummer(3000, seed=1)
    lo, hi = orc.bbox(h["pos_x"], h["pos_y"], h["pos_z"])
    for a, k in enumerate(("pos_x", "pos_y", "pos_z")):
        assert lo[a] == h[k].min() and hi[a] == h[k].max()
    codes, idx = orc.morton_codes(h["pos_x"], h["pos_y"], h["pos_z"], lo, hi)
    assert np.array_equal(idx, np.arange(3000))
    sizes = (hi - lo + np.float32(0.001)).astype(np.float32)
    inv = (np.float32(1) / sizes).astype(np.float32)
    for i in range(0, 3000, 97):
        n = [np.float32(np.float32(h[k][i] - lo[a]) * inv[a])
             for a, k in enumerate(("pos_x", "pos_y", "pos_z"))]
        q = [int(np.float32(min(max(v, np.float32(0)), np.float32(1))) * np.float32(1023)) for v in n]
        assert codes[i] == naive_interleave(*q)


def test_stable_sort_matches_numpy(orc):
    rng = np.random.default_rng(3)
    for n in (1, 2, 17, 1000, 50000):
        keys = rng.integers(0, 50, size=n).astype(np.uint32) << rng.integers(0, 25)
        c, i = orc.stable_sort(keys, np.arange(n, dtype=np.int32))
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(i, order) and np.array_equal(c, keys[order])


def tree_inputs(nbx, kind, n, seed=42):
    if kind == "plummer":
        h = nbx.gen_plummer(n, seed=seed)
    elif kind == "sphere":
        h = nbx.gen_uniform_sphere(n, seed=seed, R=5.0, sigma=0.0)
    elif kind == "coincident":
        h = nbx.empty_host(n)
        h["pos_x"][:], h["pos_y"][:], h["pos_z"][:], h["mass"][:] = 1.0, 2.0, 3.0, 1.0
    elif kind == "collinear":
        h = nbx.empty_host(n)
        h["pos_x"][:] = (np.float32(0.0001) * np.arange(n, dtype=np.float32)).astype(np.float32)
        h["mass"][:] = 1.0
    elif kind == "pairs":  # exact duplicates among spread points -> depth-32 cells
        h = nbx.gen_plummer(n, seed=seed)
        for k in ("pos_x", "pos_y", "pos_z"):
            h[k][n // 2: n // 2 + 5] = h[k][n // 2]
            h[k][10:12] = h[k][3]
    else:
        raise ValueError(kind)
    return h


CASES = [("plummer", 1), ("plummer", 2), ("plummer", 3), ("coincident", 16), ("collinear", 16),
         ("sphere", 50), ("sphere", 100), ("plummer", 300), ("pairs", 400), ("plummer", 1000)]


@pytest.mark.parametrize("kind,n", CASES)
def test_tree_oracle_matches_python_restatement(orc, nbx, kind, n):
    h = tree_inputs(nbx, kind, n)
    t = orc.bh_pipeline(h)
    pos = np.stack([h["pos_x"], h["pos_y"], h["pos_z"]], 1)
    nodes, depth = M.insert_build(pos, h["mass"], t["sorted_idx"], t["bbox_min"], t["bbox_max"],
                                  2 * n)
    ok, why = M.same_nodes(t["nodes"], M.nodes_to_array(nodes, orc.NODE_DTYPE))
    assert ok, why
    assert t["max_depth"] == depth


@pytest.mark.parametrize("kind,n", [c for c in CASES if c[0] not in ("coincident", "collinear")])
def test_numbering_rule_reproduces_insertion_order(orc, nbx, kind, n):
    h = tree_inputs(nbx, kind, n)
    t = orc.bh_pipeline(h)
    if t["budget_merges"]:
        pytest.skip("node budget fired: the sequential builder handles this input")
    pos = np.stack([h["pos_x"], h["pos_y"], h["pos_z"]], 1)
    index, members = M.numbering_build(pos, h["mass"], t["sorted_idx"], t["bbox_min"],
                                       t["bbox_max"])
    nodes = t["nodes"]
    assert len(index) == len(nodes)
    # every cell's index matches the node at the same place in the oracle's tree
    for pre, i in index.items():
        j = 0
        for o in pre:
            j = nodes[j]["children"][o]
        assert j == i, (pre, i, j)
        assert nodes[i]["particle_count"] == len(members[pre])


@pytest.mark.parametrize("kind,n", [c for c in CASES if c[0] not in ("coincident", "collinear")])
def test_path_order_construction_reproduces_the_tree(orc, nbx, kind, n):
    # the construction the GPU builder uses (cells from adjacent common prefixes
    # of the path-sorted particles, preorder numbering from one scan) gives the
    # oracle's node array field for field, and its cell order is the preorder
    # with children 7..0 that the reference's stack walk visits
    h = tree_inputs(nbx, kind, n)
    t = orc.bh_pipeline(h)
    pos = np.stack([h["pos_x"], h["pos_y"], h["pos_z"]], 1)
    r = M.path_order_build(pos, h["mass"], t["sorted_idx"], t["bbox_min"], t["bbox_max"], 2 * n)
    if r is None:
        assert t["node_count"] + 33 >= 2 * n
        pytest.skip("node budget could fire: the sequential builder handles this input")
    nodes, depth, pre_of = r
    ok, why = M.same_nodes(t["nodes"], M.nodes_to_array(nodes, orc.NODE_DTYPE))
    assert ok, why
    assert depth == t["max_depth"]
    walk, stack = [], [0]  # children pushed 0..7, popped 7..0 (:160-241)
    while stack:
        v = stack.pop()
        walk.append(v)
        if not t["nodes"][v]["is_leaf"]:
            stack += [c for c in t["nodes"][v]["children"] if c >= 0]
    assert [pre_of[v] for v in walk] == list(range(len(walk)))


@pytest.mark.parametrize("kind,n", CASES)
def test_tree_invariants(orc, nbx, kind, n):
    h = tree_inputs(nbx, kind, n)
    t = orc.bh_pipeline(h)
    nodes = t["nodes"]
    # mass conservation within 0.1% (reference verifyMassConservation, :562-570)
    total = np.float32(0)
    for v in h["mass"]:
        total = np.float32(total + v)
    assert abs(total - nodes[0]["total_mass"]) < 0.001 * total
    # every particle is the first particle of exactly one leaf or merged into one
    leaves = nodes[nodes["is_leaf"] == 1]
    assert leaves["particle_count"].sum() == n
    assert len(set(leaves["particle_index"].tolist())) == len(leaves)
    # children are nested cells of half the size
    for i, nd in enumerate(nodes):
        for o, c in enumerate(nd["children"]):
            if c < 0:
                continue
            assert c > i  # created after the parent
            assert nodes[c]["half_size"] == np.float32(nd["half_size"] * np.float32(0.5))
            assert nd["particle_count"] >= nodes[c]["particle_count"]


def test_degenerate_inputs_use_budget_merges(orc, nbx):
    # reference tests/test_barnes_hut.cpp:53-89: coincident particles exhaust the budget
    t = orc.bh_pipeline(tree_inputs(nbx, "coincident", 16))
    assert t["budget_merges"] > 0 and t["node_count"] <= 32


def test_depth32_cells_merge(orc, nbx):
    t = orc.bh_pipeline(tree_inputs(nbx, "pairs", 400))
    # 5 copies of one point (4 merges) and 3 copies of another (2 merges)
    assert t["depth_merges"] == 6 and t["budget_merges"] == 0 and t["max_depth"] == 32


def test_bh_theta_zero_is_direct_sum(orc, nbx):
    h = nbx.gen_plummer(400, seed=5)
    t = orc.bh_pipeline(h)
    a = orc.bh_forces(t["nodes"], h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 0.0, 1.0, 0.01)
    ref = orc.direct_f64(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 1.0, orc.eps2_of(0.01))
    err = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert err.max() < 1e-5


def test_direct_vs_barnes_hut_kat(orc, nbx):
    # reference tests/test_spatial_hash.cpp:171-227: theta 0.1, per component < 2%
    # (floor 1e-4), unit masses as SphericalDistParams defaults (types.hpp:157-158).
    # The reference runs it at N=30, where its own node budget (2N with a
    # 33-node reserve, force_barnes_hut.cu:449-453) fires and merges distant
    # particles into one leaf; see test_budget_merges_break_small_kat. The
    # bound is therefore checked at N=300, where the budget cannot fire.
    h = nbx.gen_uniform_sphere(300, seed=42, R=5.0, M=300.0, sigma=0.0)
    t = orc.bh_pipeline(h)
    assert t["budget_merges"] == 0
    bh = orc.bh_forces(t["nodes"], h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 0.1, 1.0, 0.1)
    d = orc.direct_f32(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 1.0, 0.1)
    mag = np.linalg.norm(d, axis=1, keepdims=True)
    assert np.max(np.linalg.norm(bh - d, axis=1) / mag[:, 0]) < 0.02
    # per component, with the floor scaled to the particle's |a| (not an absolute 1e-4)
    assert np.max(np.abs(bh - d) / np.maximum(0.05 * mag, np.abs(d))) < 0.02


def test_budget_merges_break_small_kat(orc, nbx):
    # At N=30 the budget fires and even theta=0 differs from the direct sum.
    h = nbx.gen_uniform_sphere(30, seed=42, R=5.0, M=30.0, sigma=0.0)
    t = orc.bh_pipeline(h)
    assert t["budget_merges"] > 0
    bh = orc.bh_forces(t["nodes"], h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 0.0, 1.0, 0.1)
    d = orc.direct_f32(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 1.0, 0.1)
    assert np.max(np.abs(bh - d) / np.maximum(1e-4, np.abs(d))) > 0.02


def test_approximation_convergence_kat(orc, nbx):
    # reference tests/test_barnes_hut.cpp:168-220 (unit masses; N=500 keeps the
    # node budget from firing, as above)
    h = nbx.gen_uniform_sphere(500, seed=42, R=10.0, M=500.0, sigma=0.0)
    t = orc.bh_pipeline(h)
    assert t["budget_merges"] == 0
    d = orc.direct_f32(h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 1.0, 0.1)

    def err(theta):
        a = orc.bh_forces(t["nodes"], h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], theta, 1.0, 0.1)
        return np.linalg.norm(a - d, axis=1).sum()

    assert err(0.3) <= 1.1 * err(0.8)
    assert err(0.3) < err(0.8)


def overrun_host(nbx):
    pts = np.array(M.budget_overrun_input(), np.float32)
    h = nbx.empty_host(len(pts))
    h["pos_x"][:], h["pos_y"][:], h["pos_z"][:] = pts[:, 0], pts[:, 1], pts[:, 2]
    h["mass"][:] = np.float32(1.0 / len(pts))
    return h


def test_node_budget_overrun_detected(orc, nbx):
    # The reference's insertion makes no budget check when it creates a leaf in an
    # empty child; on this input it needs more than its 2N nodes. The independent
    # Python restatement (no bound on its node list) shows the count past 2N, and
    # the oracle refuses the build instead of writing past its array.
    h = overrun_host(nbx)
    n = len(h["mass"])
    x, y, z = h["pos_x"], h["pos_y"], h["pos_z"]
    lo, hi = orc.bbox(x, y, z)
    codes, idx = orc.morton_codes(x, y, z, lo, hi)
    _, sidx = orc.stable_sort(codes, idx)
    pos = np.stack([x, y, z], 1)
    nodes, _ = M.insert_build(pos, h["mass"], sidx, lo, hi, 2 * n)
    assert len(nodes) > 2 * n
    with pytest.raises(RuntimeError, match="overrun"):
        orc.bh_pipeline(h)
    # with room for every node (capacity for more bodies) the same input builds
    t = orc.bh_pipeline(h, max_particles=4 * n)
    assert t["node_count"] == len(M.insert_build(pos, h["mass"], sidx, lo, hi, 8 * n)[0])
