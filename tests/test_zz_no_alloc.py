"""The allocation-free entry point nbx_direct_compute_noalloc, through the C ABI.

Its pair-symmetric scratch is a fixed 64 MiB device array inside the library.
Two calls in a row (300,000 bodies, which runs 21 rounds of distances in it,
then 70,000, which fits one round) must give accelerations bit-identical to
nbx_direct_compute with an allocated workspace on the same inputs; a call
beyond what the array holds fails with NBX_ERR_CAPACITY and leaves the
accelerations untouched.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G, EPS = 1.0, 0.005
NBX_ERR_CAPACITY = 4


def workspace_forces(nbx, h):
    d = nbx.DeviceParticles.from_host(h)
    ws = nbx.DirectWorkspace(len(h["mass"]))
    ws.forces(d, G, EPS)
    nbx.synchronize()
    return d.acc()


def noalloc_forces(nbx, h):
    d = nbx.DeviceParticles.from_host(h)
    nbx.check(nbx.lib.nbx_direct_compute_noalloc(d.ptr, G, nbx.eps2_of(EPS), None))
    nbx.synchronize()
    return d.acc()


def test_noalloc_equals_workspace_path_twice(nbx):
    nbx.lib.nbx_direct_compute_noalloc.restype = nbx.C.c_int
    nbx.lib.nbx_direct_compute_noalloc.argtypes = [nbx.PP, nbx.C.c_float, nbx.C.c_float, nbx.VP]
    for n, seed in ((300000, 1), (70000, 2)):
        print(f"noalloc n={n}", flush=True)
        h = nbx.gen_uniform_sphere(n, seed=seed, R=1.0, sigma=0.05)
        assert nbx.variant_names()[nbx.lib.nbx_direct_variant_for(n)].startswith("sym_")
        a = noalloc_forces(nbx, h)
        b = workspace_forces(nbx, h)
        assert np.all(np.isfinite(a)) and np.any(a != 0)
        assert np.array_equal(a, b), n


def test_noalloc_rejects_beyond_its_scratch(nbx):
    nbx.lib.nbx_direct_compute_noalloc.restype = nbx.C.c_int
    nbx.lib.nbx_direct_compute_noalloc.argtypes = [nbx.PP, nbx.C.c_float, nbx.C.c_float, nbx.VP]
    d = nbx.DeviceParticles(1200000)
    st = nbx.lib.nbx_direct_compute_noalloc(d.ptr, G, nbx.eps2_of(EPS), None)
    assert st == NBX_ERR_CAPACITY, st
    assert b"scratch" in nbx.lib.nbx_last_error()
    nbx.synchronize()
    assert np.all(d.read("acc_x") == 0)  # nothing was launched
