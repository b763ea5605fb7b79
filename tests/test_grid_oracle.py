"""CPU checks of the spatial-hash oracle (no GPU).

What pins each oracle function, besides its restatement of
src/cuda/force_spatial_hash.cu:
  * getCellIndex            known answers of tests/test_spatial_hash.cpp:34-47
  * hashCell                Python's modulo over every offset of a small box
  * grid sizing             a numpy float32 restatement of :232-260
  * build                   numpy bincount/cumsum and a stable argsort of the (cell,
                            sub-cell) key; the one-cell-per-particle property of
                            :133-167; every particle inside its cell
                            (verifyCellAssignment :342-368) and inside its sub-cell
  * forces                  a brute-force cutoff sum without any grid (same pair set,
                            FP64), the cutoff-boundary known answers of :81-127, and the
                            FP64 direct sum once the cutoff exceeds every distance
"""
import numpy as np
import pytest


def cube(n, half, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-half, half, n).astype(np.float32) for _ in range(3)]


def numpy_dims(x, y, z, cell):
    """:232-252 in float32: pad the exact bbox by 0.001f, dims = (int)ceilf(ext/cell)+1."""
    f = np.float32
    lo = np.array([x.min(), y.min(), z.min()], f) - f(0.001)
    hi = np.array([x.max(), y.max(), z.max()], f) + f(0.001)
    r = (hi - lo) / f(cell)
    assert r.dtype == np.float32
    return lo, hi, np.ceil(r).astype(np.int64) + 1


def numpy_cells(x, y, z, lo, dims, cell):
    """:36-49 in float32 with the clamp."""
    f = np.float32
    c = [np.clip(np.floor((p - lo[a]) / f(cell)).astype(np.int64), 0, dims[a] - 1)
         for a, p in enumerate((x, y, z))]
    return c[0] + c[1] * dims[0] + c[2] * dims[0] * dims[1], c


def sub_cells_per_axis(dims):
    """16 sub-cells per axis up to 2^20 cells, 4 up to 2^26, none beyond."""
    total = int(np.prod(dims))
    return 16 if total <= 1 << 20 else (4 if total <= 1 << 26 else 0)


def numpy_sort_key(x, y, z, lo, dims, cell):
    """Order within a cell: Morton index of the position's sub-cell (floor(ns u) - ns c
    per axis, u = (p - min) / cell in float32, ns = sub_cells_per_axis(dims))."""
    f = np.float32
    pc, c = numpy_cells(x, y, z, lo, dims, cell)
    ns = sub_cells_per_axis(dims)
    if ns == 0:
        return pc, None
    s = []
    for a, p in enumerate((x, y, z)):
        u = (p - lo[a]) / f(cell)
        assert u.dtype == np.float32
        s.append(np.clip(np.floor(u * f(ns)).astype(np.int64) - ns * c[a], 0, ns - 1))
    def spread(v):  # bit b of v to bit 3b
        return sum(((v >> b) & 1) << (3 * b) for b in range(4))
    sub = (spread(s[0]) << 2) | (spread(s[1]) << 1) | spread(s[2])
    return pc * (ns ** 3) + sub, s


def test_cell_index_known_answers(orc):
    # tests/test_spatial_hash.cpp:34-47
    assert orc.grid_cell_index(0.5, 0.5, 0.5, 2.0) == (0, 0, 0)
    assert orc.grid_cell_index(2.5, 4.5, 6.5, 2.0) == (1, 2, 3)
    # floor, not truncation, below zero
    assert orc.grid_cell_index(-0.5, -2.0, -2.5, 2.0) == (-1, -1, -2)


def test_hash_cell_wraps_like_python_modulo(orc):
    dims = (3, 4, 5)
    for cx in range(-7, 8):
        for cy in range(-9, 10, 3):
            for cz in range(-11, 12, 4):
                want = (cx % 3) + (cy % 4) * 3 + (cz % 5) * 12
                assert orc.grid_hash_cell((cx, cy, cz), dims) == want


@pytest.mark.parametrize("cell", [0.7, 1.0, 2.0, 3.3, 25.0])
def test_dims_follow_float32_formula(orc, cell):
    x, y, z = cube(3000, 10.0, 1)
    bbox, dims, total = orc.grid_dims(x, y, z, cell)
    lo, hi, want = numpy_dims(x, y, z, cell)
    np.testing.assert_array_equal(bbox[:3], lo)
    np.testing.assert_array_equal(bbox[3:], hi)
    np.testing.assert_array_equal(dims, want)
    assert total == int(np.prod(want))


def test_grid_too_large_is_refused(orc):
    x, y, z = cube(100, 1e4, 2)
    assert orc.grid_dims(x, y, z, 0.01) is None      # ~2e6 cells per axis
    with pytest.raises(RuntimeError, match="too large"):
        orc.grid_build(x, y, z, 0.01)
    # the same box with a coarser cell fits (41^3 cells)
    assert orc.grid_dims(x, y, z, 500.0) is not None


def test_build_matches_numpy(orc):
    x, y, z = cube(5000, 10.0, 3)
    for cell in (0.09, 0.9, 2.0, 4.5):  # 0.09: above 2^20 cells, 4 sub-cells per axis
        g = orc.grid_build(x, y, z, cell)
        assert (g["total"] > 1 << 20) == (cell < 0.1)
        lo = g["bbox"][:3]
        pc, _ = numpy_cells(x, y, z, lo, g["dims"].astype(np.int64), cell)
        np.testing.assert_array_equal(g["particle_cell"], pc)
        counts = np.bincount(pc, minlength=g["total"])
        start = np.concatenate([[0], np.cumsum(counts)[:-1]])
        np.testing.assert_array_equal(g["cell_start"], start)
        np.testing.assert_array_equal(g["cell_end"], start + counts)
        key, _ = numpy_sort_key(x, y, z, lo, g["dims"].astype(np.int64), cell)
        np.testing.assert_array_equal(g["sorted"], np.argsort(key, kind="stable"))


def test_each_particle_in_exactly_one_cell(orc):
    # Property 5 of tests/test_spatial_hash.cpp:133-167 over cell sizes in (0.5, 10).
    rng = np.random.default_rng(5)
    x, y, z = cube(50, 10.0, 6)
    for cell in rng.uniform(0.5, 10.0, 25):
        g = orc.grid_build(x, y, z, float(cell))
        seen = []
        for c in range(g["total"]):
            seen.extend(g["sorted"][g["cell_start"][c]:g["cell_end"][c]].tolist())
        assert sorted(seen) == list(range(50))


def test_every_particle_inside_its_cell(orc):
    # verifyCellAssignment (:342-368) in float32, and the same for the sub-cell.
    f = np.float32
    x, y, z = cube(2000, 7.0, 7)
    for cell in (0.6, 1.5, 2.0):
        g = orc.grid_build(x, y, z, cell)
        dims = g["dims"].astype(np.int64)
        _, coords = numpy_cells(x, y, z, g["bbox"][:3], dims, cell)
        _, subs = numpy_sort_key(x, y, z, g["bbox"][:3], dims, cell)
        for a, p in enumerate((x, y, z)):
            cmin = g["bbox"][a] + coords[a].astype(f) * f(cell)
            assert np.all(p >= cmin) and np.all(p <= cmin + f(cell))
            q = f(cell) / f(sub_cells_per_axis(dims))
            smin = cmin + subs[a].astype(f) * q
            assert np.all(p >= smin - 1e-5 * f(cell)) and np.all(p <= smin + q + 1e-5 * f(cell))
        # bodies of one cell are listed in ascending sub-cell order
        key = np.asarray(numpy_sort_key(x, y, z, g["bbox"][:3], dims, cell)[0])
        assert np.all(np.diff(key[g["sorted"]]) >= 0)


def _cases(nbx):
    out = {}
    x, y, z = cube(2000, 10.0, 11)
    out["cube_cell2"] = (x, y, z, np.ones(2000, np.float32), 2.0, 2.0)
    out["cube_cell3_cut2"] = (x, y, z, np.ones(2000, np.float32), 3.0, 2.0)
    h = nbx.gen_plummer(3000, seed=12)
    out["plummer_cell0.3"] = (h["pos_x"], h["pos_y"], h["pos_z"], h["mass"], 0.3, 0.25)
    rng = np.random.default_rng(13)
    cl = np.zeros((3, 300), np.float32)
    cl[:, 200:] = rng.normal(0, 0.5, (3, 100)).astype(np.float32)
    out["coincident"] = (cl[0], cl[1], cl[2], rng.uniform(0.5, 2, 300).astype(np.float32), 0.5,
                         0.5)
    return out


@pytest.mark.parametrize("case", ["cube_cell2", "cube_cell3_cut2", "plummer_cell0.3",
                                  "coincident"])
def test_grid_sum_equals_brute_force(nbx, orc, case):
    x, y, z, m, cell, cutoff = _cases(nbx)[case]
    g = orc.grid_build(x, y, z, cell)
    eps2 = orc.eps2_of(0.05)
    a, mag, npg = orc.grid_forces_f64(g, x, y, z, m, cutoff, 1.0, eps2)
    b, npb = orc.cutoff_forces_f64(x, y, z, m, cutoff, 1.0, eps2)
    np.testing.assert_array_equal(npg, npb)
    assert npg.sum() > 0
    scale = np.maximum(mag, 1e-300)[:, None]
    assert np.all(np.abs(a - b) <= 1e-12 * scale)
    # and the FP32 restatement of the kernel stays within FP32 rounding of it: about
    # 5 ulp per contribution (r^2, sqrt, divide, cube, products) plus one per addition
    a32 = orc.grid_forces_f32(g, x, y, z, m, cutoff, 1.0, 0.05)
    bound = (16 + 2 * npg)[:, None] * 2.0 ** -24 * scale
    assert np.all(np.abs(a32 - a) <= bound)


def test_cutoff_boundary_known_answers(orc):
    # tests/test_spatial_hash.cpp:81-127: G=1, eps=0.01, cutoff = cell = 2, delta = 0.25.
    G, eps, cutoff, delta = 1.0, 0.01, 2.0, 0.25
    m = np.ones(2, np.float32)
    z0 = np.zeros(2, np.float32)
    x = np.array([0.0, cutoff + delta], np.float32)
    g = orc.grid_build(x, z0, z0, cutoff)
    a = orc.grid_forces_f32(g, x, z0, z0, m, cutoff, G, eps)
    assert np.all(a == 0.0)
    r = np.float32(cutoff - delta)
    x = np.array([0.0, r], np.float32)
    g = orc.grid_build(x, z0, z0, cutoff)
    a = orc.grid_forces_f32(g, x, z0, z0, m, cutoff, G, eps)
    d2 = r * r + np.float32(eps) * np.float32(eps)
    want = G * r / (d2 * np.sqrt(d2))
    assert abs(a[0, 0] - want) < 1e-4 and abs(a[1, 0] + want) < 1e-4
    assert np.all(a[:, 1:] == 0.0)


def test_cutoff_beyond_every_distance_is_the_direct_sum(nbx, orc):
    h = nbx.gen_plummer(600, seed=21)
    x, y, z, m = h["pos_x"], h["pos_y"], h["pos_z"], h["mass"]
    span = float(max(np.ptp(x), np.ptp(y), np.ptp(z)))
    cutoff = 2.0 * span  # every pair kept
    g = orc.grid_build(x, y, z, cutoff)
    eps2 = orc.eps2_of(0.01)
    a, mag, npg = orc.grid_forces_f64(g, x, y, z, m, cutoff, 1.0, eps2)
    assert np.all(npg == 599)
    d = orc.direct_f64(x, y, z, m, 1.0, eps2)
    assert np.all(np.abs(a - d) <= 1e-12 * mag[:, None])
