"""CPU checks of the C ABI library: it loads, exports every entry point that
include/nbx.h declares, and its host-only helpers (status strings,
generators) behave. No GPU compute is called here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "nbx.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(nbx_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol(nbx):
    syms = declared_symbols()
    assert len(syms) > 30
    missing = [s for s in syms if not hasattr(nbx.lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", nbx.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (nbx_[a-z0-9_]+)", out))
    assert set(syms) <= exported, sorted(set(syms) - exported)


def test_python_binding_covers_header(nbx):
    # every symbol the binding lists is declared in the header
    assert set(nbx.EXPORTED) <= set(declared_symbols())


def test_status_strings(nbx):
    assert nbx.lib.nbx_status_string(0) == b"ok"
    assert nbx.lib.nbx_status_string(1) == b"invalid argument"
    assert nbx.lib.nbx_status_string(99) == b"unknown status"


def test_struct_layout_matches_particle_data(nbx):
    # nbody::ParticleData: 13 pointers then size_t count (types.hpp:96-131)
    assert ctypes.sizeof(nbx.Particles) == 14 * 8
    assert nbx.Particles.count.offset == 13 * 8


def test_generators_deterministic_and_centred(nbx):
    a = nbx.gen_plummer(2048, seed=42)
    b = nbx.gen_plummer(2048, seed=42)
    c = nbx.gen_plummer(2048, seed=43)
    for k in a:
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["pos_x"], c["pos_x"])
    for k in ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z"):
        assert abs(a[k].astype(np.float64).mean()) < 1e-6
    assert np.all(a["mass"] == np.float32(1.0 / 2048))
    r = np.sqrt(a["pos_x"].astype(np.float64) ** 2 + a["pos_y"] ** 2 + a["pos_z"] ** 2)
    # Plummer enclosed-mass fraction at r: r^3 / (1 + r^2)^{3/2}; median radius ~1.305
    assert 1.15 < np.median(r) < 1.45
    assert r.max() < 40.0  # truncated at 99.9% of the mass


def test_plummer_virial_ratio(nbx, orc):
    h = nbx.gen_plummer(3000, seed=5)
    e = orc.energy_f64(h, 1.0, 0.0)
    q = 2 * e["ke"] / abs(e["pe"])
    assert 0.85 < q < 1.15  # equilibrium model: 2K/|W| ~ 1


def test_uniform_sphere_and_two_plummer(nbx):
    u = nbx.gen_uniform_sphere(4096, seed=1, R=1.0, sigma=0.05)
    r = np.sqrt(u["pos_x"].astype(np.float64) ** 2 + u["pos_y"] ** 2 + u["pos_z"] ** 2)
    assert r.max() <= 1.0 + 1e-6
    assert abs(np.median(r) - 0.5 ** (1 / 3)) < 0.03
    assert abs(u["vel_x"].std() - 0.05) < 0.005
    t = nbx.gen_two_plummer(4096, seed=2)
    assert abs(t["pos_x"][:2048].mean() + 5.0) < 1e-4 and abs(t["pos_x"][2048:].mean() - 5.0) < 1e-4
    assert abs(t["vel_x"][:2048].mean() - 0.25) < 1e-5 and abs(t["vel_x"][2048:].mean() + 0.25) < 1e-5
    assert np.all(t["mass"] == np.float32(1.0 / 2048))
    with pytest.raises(nbx.NbxError):
        nbx.gen_two_plummer(3)


def test_ensemble_sizes(nbx):
    s = nbx.gen_ensemble_sizes(512, seed=42)
    assert s.min() >= 256 and s.max() <= 16384
    assert np.array_equal(s, nbx.gen_ensemble_sizes(512, seed=42))
    lg = np.log(s.astype(np.float64))
    assert abs(lg.mean() - 0.5 * (np.log(256) + np.log(16384))) < 0.2


def test_default_direct_shape_by_size(nbx):
    # the shape a call without tuning launches, at each size boundary (host logic only)
    names = nbx.variant_names()
    assert nbx.lib.nbx_direct_set_variant(-1) == 0
    expect = [(4096, "ipl1_nb8_ws16"), (16384, "ipl1_nb8_ws16"), (16385, "sym_ipl4_ws1"),
              (65535, "sym_ipl4_ws1"), (65536, "sym_ipl6_ws1"), (131071, "sym_ipl6_ws1"),
              (131072, "sym_ipl8_ws1"), (196607, "sym_ipl8_ws1"), (196608, "sym_ipl6_ws4"),
              (1 << 20, "sym_ipl6_ws4")]
    got = [(n, names[nbx.lib.nbx_direct_variant_for(n)]) for n, _ in expect]
    assert got == expect
