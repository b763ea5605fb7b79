"""GPU parity of the spatial-hash operator (nbx_grid_*) against the CPU oracle.

Bit-exact: padded bbox, dims, total cells, the cell of every particle,
cell_start / cell_end of every cell and the sorted indices (ascending within a
cell; the reference's atomic scatter leaves that order open and ascending is one
of its outcomes). Forces: the cutoff test is evaluated with the same IEEE
arithmetic as the oracle, so the GPU keeps exactly the oracle's pairs; the kept
pairs use v_rsq_f32 and FMA accumulation. Bound per target and component:
|a_gpu - a_f64| <= (16 + 2K) 2^-24 sum_j |a_ij|, K the kept pairs of the target
(about 5 ulp per contribution plus one per addition), the same bound the oracle's
own FP32 restatement meets.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def host_from(x, y, z, m):
    n = len(x)
    h = {f: np.zeros(n, np.float32) for f in ("vel_x", "vel_y", "vel_z", "acc_x", "acc_y",
                                              "acc_z", "acc_old_x", "acc_old_y", "acc_old_z")}
    h.update(pos_x=np.ascontiguousarray(x, np.float32), pos_y=np.ascontiguousarray(y, np.float32),
             pos_z=np.ascontiguousarray(z, np.float32), mass=np.ascontiguousarray(m, np.float32))
    return h


def cube(n, half, seed, mass=1.0):
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(-half, half, n).astype(np.float32) for _ in range(3))
    return host_from(x, y, z, np.full(n, mass, np.float32))


def sphere(nbx, n, R, seed=42):
    h = nbx.gen_uniform_sphere(n, seed=seed, R=R, M=float(n))  # unit masses
    return h


def case(nbx, name):
    """(host particles, cell, cutoff)."""
    if name == "uniform100":            # tests/test_spatial_hash.cpp:18-32
        return cube(100, 10.0, 1), 2.0, 2.0
    if name == "cube2000":
        return cube(2000, 10.0, 2), 2.0, 2.0
    if name == "cube2000_cell3":
        return cube(2000, 10.0, 3), 3.0, 2.0
    if name == "cube2000_cut3":         # cutoff beyond the cell: only the 27 cells are walked
        return cube(2000, 10.0, 9), 2.0, 3.0
    if name == "sphere4096":            # the reference benchmark's distribution at its default N
        return sphere(nbx, 4096, 10.0), 2.0, 2.0
    if name == "dense50000":            # hundreds of bodies per cell: shared walks
        return sphere(nbx, 50000, 4.0), 2.0, 2.0
    if name == "plummer65536":          # sparse: 3.7 M cells, most of them empty
        return nbx.gen_plummer(65536, seed=7), 0.5, 0.5
    if name == "coincident":
        rng = np.random.default_rng(4)
        p = np.zeros((3, 300), np.float32)
        p[:, 200:] = rng.normal(0, 0.5, (3, 100)).astype(np.float32)
        return host_from(p[0], p[1], p[2], rng.uniform(0.5, 2, 300)), 0.5, 0.5
    if name == "collinear":
        t = np.linspace(-5, 5, 700, dtype=np.float32)
        return host_from(t, 0.5 * t, np.zeros_like(t), np.ones_like(t)), 0.4, 0.3
    if name == "single":
        return host_from([1.5], [-2.0], [0.25], [3.0]), 1.0, 1.0
    if name == "huge_cell":             # cell larger than the box: 2 x 2 x 2 cells
        return cube(500, 3.0, 5), 50.0, 50.0
    if name == "sparse_69m":            # 411^3 cells, above 2^26: cell order alone, no sub-cells
        rng = np.random.default_rng(8)
        p = rng.uniform(-205.0, 205.0, (3, 1000)).astype(np.float32)
        p[:, 900:] = rng.normal(0.0, 1.0, (3, 100)).astype(np.float32)  # a cluster with pairs
        return host_from(p[0], p[1], p[2], rng.uniform(0.5, 2.0, 1000)), 1.0, 1.0
    raise KeyError(name)


BUILD_CASES = ["uniform100", "cube2000", "cube2000_cell3", "cube2000_cut3", "sphere4096",
               "dense50000",
               "plummer65536", "coincident", "collinear", "single", "huge_cell", "sparse_69m"]


def gpu_grid(nbx, h, cell, capacity=None):
    d = nbx.DeviceParticles.from_host(h)
    g = nbx.Grid(capacity or len(h["mass"]), cell)
    g.build(d)
    return d, g


def check_build(nbx, orc, h, cell, g):
    n = len(h["mass"])
    ref = orc.grid_build(h["pos_x"], h["pos_y"], h["pos_z"], cell)
    info = g.info()
    assert np.array_equal(info["bbox"], ref["bbox"])
    assert np.array_equal(info["dims"], ref["dims"]) and info["total"] == ref["total"]
    cells = g.cells()
    assert np.array_equal(cells["particle_cell"][:n], ref["particle_cell"])
    assert np.array_equal(cells["cell_start"], ref["cell_start"])
    assert np.array_equal(cells["cell_end"], ref["cell_end"])
    assert np.array_equal(cells["sorted"][:n], ref["sorted"])
    return ref


def check_forces(orc, h, ref, acc, cutoff, G, eps, targets=None):
    eps2 = orc.eps2_of(eps)
    a64, mag, npairs = orc.grid_forces_f64(ref, h["pos_x"], h["pos_y"], h["pos_z"], h["mass"],
                                           cutoff, G, eps2, targets=targets)
    got = acc if targets is None else acc[targets]
    bound = (16 + 2 * npairs)[:, None] * 2.0 ** -24 * mag[:, None]
    err = np.abs(got.astype(np.float64) - a64)
    bad = np.argwhere(err > bound)
    assert len(bad) == 0, (bad[:5], err[tuple(bad[0])], bound[bad[0][0]], npairs[bad[0][0]])
    return npairs


@pytest.mark.parametrize("name", BUILD_CASES)
def test_build_bit_exact(nbx, orc, name):
    h, cell, _ = case(nbx, name)
    print(f"grid build {name} n={len(h['mass'])} cell={cell}", flush=True)
    _, g = gpu_grid(nbx, h, cell)
    check_build(nbx, orc, h, cell, g)


@pytest.mark.parametrize("name", BUILD_CASES)
def test_forces_match_oracle(nbx, orc, name):
    h, cell, cutoff = case(nbx, name)
    print(f"grid forces {name} n={len(h['mass'])} cell={cell} cutoff={cutoff}", flush=True)
    d, g = gpu_grid(nbx, h, cell)
    ref = check_build(nbx, orc, h, cell, g)
    g.forces(d, cutoff, 1.0, 0.05)
    nbx.synchronize()
    acc = d.acc()
    assert np.all(np.isfinite(acc))
    check_forces(orc, h, ref, acc, cutoff, 1.0, 0.05)


@pytest.mark.parametrize("name", ["cube2000", "cube2000_cut3", "sphere4096", "dense50000",
                                  "plummer65536", "coincident"])
def test_walks_are_bit_identical(nbx, name):
    h, cell, cutoff = case(nbx, name)
    print(f"grid walks {name}", flush=True)
    d, g = gpu_grid(nbx, h, cell)
    g.forces(d, cutoff, 1.0, 0.05)
    nbx.synchronize()
    a0 = d.acc().copy()
    g.set_walk(1)
    g.forces(d, cutoff, 1.0, 0.05)
    nbx.synchronize()
    assert np.array_equal(a0, d.acc())
    g.set_walk(0)
    g.forces(d, cutoff, 1.0, 0.05)
    nbx.synchronize()
    assert np.array_equal(a0, d.acc()), "repeat call differs"


def test_cutoff_boundary_known_answers(nbx):
    # tests/test_spatial_hash.cpp:81-127
    G, eps, cutoff, cell, delta = 1.0, 0.01, 2.0, 2.0, 0.25
    z0 = [0.0, 0.0]
    for sep, inside in ((cutoff + delta, False), (cutoff - delta, True)):
        h = host_from([0.0, sep], z0, z0, [1.0, 1.0])
        d, g = gpu_grid(nbx, h, cell)
        g.forces(d, cutoff, G, eps)
        nbx.synchronize()
        a = d.acc()
        if not inside:
            assert np.all(a == 0.0)
        else:
            r = np.float32(sep)
            d2 = r * r + np.float32(eps) ** 2
            want = G * r / (d2 * np.sqrt(d2))
            assert abs(a[0, 0] - want) < 1e-4 and abs(a[1, 0] + want) < 1e-4
            assert np.all(a[:, 1:] == 0.0)


def test_compute_forces_finite(nbx):
    # tests/test_spatial_hash.cpp:49-73: 100 bodies in [-5, 5]^3, cell = cutoff = 2, eps 0.1.
    h = cube(100, 5.0, 9)
    d, g = gpu_grid(nbx, h, 2.0)
    assert g.info()["total"] > 0
    g.forces(d, 2.0, 1.0, 0.1)
    nbx.synchronize()
    assert np.all(np.isfinite(d.acc()))


def test_zero_softening_skips_self(nbx, orc):
    # eps = 0: the self pair would be 0 * inf = NaN if it were not skipped by index.
    h = cube(3000, 6.0, 10)
    d, g = gpu_grid(nbx, h, 1.5)
    ref = check_build(nbx, orc, h, 1.5, g)
    g.forces(d, 1.5, 1.0, 0.0)
    nbx.synchronize()
    acc = d.acc()
    assert np.all(np.isfinite(acc))
    check_forces(orc, h, ref, acc, 1.5, 1.0, 0.0)


def test_mixed_masses_and_G(nbx, orc):
    h = cube(5000, 8.0, 11)
    h["mass"] = np.random.default_rng(12).uniform(0.1, 3.0, 5000).astype(np.float32)
    d, g = gpu_grid(nbx, h, 2.0)
    ref = check_build(nbx, orc, h, 2.0, g)
    g.forces(d, 1.7, 6.674e-3, 0.02)
    nbx.synchronize()
    check_forces(orc, h, ref, d.acc(), 1.7, 6.674e-3, 0.02)


def test_positions_moved_without_rebuild(nbx, orc):
    # The force call reads the current positions and recomputes each target's cell
    # with the last build's grid (:94-108), as the reference kernel does.
    h, cell, cutoff = case(nbx, "sphere4096")
    d, g = gpu_grid(nbx, h, cell)
    ref = check_build(nbx, orc, h, cell, g)
    rng = np.random.default_rng(13)
    moved = {k: v.copy() for k, v in h.items()}
    for k in ("pos_x", "pos_y", "pos_z"):
        moved[k] = (moved[k] + rng.normal(0, 0.3, len(moved[k]))).astype(np.float32)
    d.upload(moved)
    g.forces(d, cutoff, 1.0, 0.05)
    nbx.synchronize()
    check_forces(orc, moved, ref, d.acc(), cutoff, 1.0, 0.05)


def test_rebuild_smaller_box_and_count(nbx, orc):
    h = cube(4000, 10.0, 14)
    d, g = gpu_grid(nbx, h, 1.0, capacity=4000)
    check_build(nbx, orc, h, 1.0, g)
    h2 = cube(1500, 3.0, 15)
    d2 = nbx.DeviceParticles.from_host(h2)
    g.build(d2)
    ref = check_build(nbx, orc, h2, 1.0, g)
    cells = g.cells()
    assert np.all(cells["sorted"][1500:] == -1) and np.all(cells["particle_cell"][1500:] == -1)
    g.forces(d2, 1.0, 1.0, 0.05)
    nbx.synchronize()
    check_forces(orc, h2, ref, d2.acc(), 1.0, 1.0, 0.05)


def test_errors(nbx):
    h = cube(100, 1e4, 16)
    d = nbx.DeviceParticles.from_host(h)
    g = nbx.Grid(100, 0.01)
    with pytest.raises(nbx.NbxError, match="Spatial hash grid too large") as e:
        g.build(d)
    assert e.value.status == 4
    with pytest.raises(nbx.NbxError) as e:   # no successful build yet
        g.forces(d, 0.01, 1.0, 0.1)
    assert e.value.status == 5
    g2 = nbx.Grid(50, 1.0)
    with pytest.raises(nbx.NbxError) as e:   # more particles than the grid was made for
        g2.build(d)
    assert e.value.status == 4
    g3 = nbx.Grid(200, 5000.0)
    g3.build(d)
    d_small = nbx.DeviceParticles.from_host(cube(60, 1.0, 17))
    with pytest.raises(nbx.NbxError) as e:   # forces on another particle count
        g3.forces(d_small, 1.0, 1.0, 0.1)
    assert e.value.status == 5
    for bad in (0.0, -1.0, float("inf"), float("nan")):
        with pytest.raises(nbx.NbxError) as e:
            nbx.Grid(10, bad)
        assert e.value.status == 1


def test_benchmark_sphere_million_sampled(nbx, orc):
    # 1,048,576 bodies in the reference benchmark's sphere (R = 10, cell = cutoff = 2):
    # the whole build bit-exact, forces on 2,048 sampled targets.
    h = sphere(nbx, 1 << 20, 10.0, seed=42)
    print("grid 1M build", flush=True)
    d, g = gpu_grid(nbx, h, 2.0)
    ref = check_build(nbx, orc, h, 2.0, g)
    print("grid 1M forces", flush=True)
    g.forces(d, 2.0, 1.0, 0.05)
    nbx.synchronize()
    acc = d.acc()
    assert np.all(np.isfinite(acc))
    targets = np.random.default_rng(18).choice(1 << 20, 2048, replace=False)
    npairs = check_forces(orc, h, ref, acc, 2.0, 1.0, 0.05, targets=targets)
    assert npairs.mean() > 1000
