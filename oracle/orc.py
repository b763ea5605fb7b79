"""ctypes wrapper of the CPU oracle (liborc.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to check or to time the CPU baseline. See oracle.h for
the reference file:line each function restates and for what pins it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborc.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"oracle not built at {LIB_PATH}; run `make -C oracle`")
lib = C.CDLL(LIB_PATH)

F32P = C.POINTER(C.c_float)
F64P = C.POINTER(C.c_double)
I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
U32P = C.POINTER(C.c_uint32)

# 76-byte node, layout of nbody::OctreeNode (barnes_hut_tree.hpp:9-30)
NODE_DTYPE = np.dtype([
    ("center", "<f4", 3), ("half_size", "<f4"), ("com", "<f4", 3), ("total_mass", "<f4"),
    ("children", "<i4", 8), ("particle_index", "<i4"), ("is_leaf", "u1"), ("pad", "u1", 3),
    ("particle_count", "<i4"),
])
assert NODE_DTYPE.itemsize == 76


def _s(name, restype, *args):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(args)
    return f


_s("orc_pair_accel", None, F32P, F32P, C.c_float, C.c_float, C.c_float, F32P)
_s("orc_direct_f32", None, C.c_size_t, F32P, F32P, F32P, F32P, C.c_float, C.c_float, I64P,
   C.c_size_t, F32P, F32P, F32P, C.c_int)
_s("orc_direct_f64", None, C.c_size_t, F32P, F32P, F32P, F32P, C.c_double, C.c_double, I64P,
   C.c_size_t, F64P, F64P, F64P, C.c_int)
_s("orc_direct_f64_pos64", None, C.c_size_t, F64P, F64P, F64P, F32P, C.c_double, C.c_double,
   I64P, C.c_size_t, F64P, F64P, F64P, C.c_int)
_s("orc_store_old", None, C.c_size_t, F32P, F32P, F32P, F32P, F32P, F32P)
_s("orc_drift", None, C.c_size_t, F32P, F32P, F32P, F32P, F32P, F32P, F32P, F32P, F32P, C.c_float)
_s("orc_kick", None, C.c_size_t, F32P, F32P, F32P, F32P, F32P, F32P, F32P, F32P, F32P, C.c_float)
_s("orc_run_direct", None, C.c_size_t, *([F32P] * 12), F32P, C.c_float, C.c_float, C.c_float,
   C.c_int, C.c_int)
_s("orc_energy_f64", None, C.c_size_t, *([F32P] * 7), C.c_double, C.c_double, F64P, F64P, F64P,
   C.c_int)
_s("orc_expand_bits", C.c_uint32, C.c_uint32)
_s("orc_morton_code", C.c_uint32, C.c_float, C.c_float, C.c_float)
_s("orc_bbox", None, C.c_size_t, F32P, F32P, F32P, F32P, F32P)
_s("orc_morton_codes", None, C.c_size_t, F32P, F32P, F32P, F32P, F32P, U32P, I32P)
_s("orc_stable_sort", None, C.c_size_t, U32P, I32P)
_s("orc_build_tree", C.c_int, C.c_size_t, F32P, F32P, F32P, F32P, I32P, F32P, F32P, C.c_size_t,
   C.c_void_p, I32P, I32P, I32P, I32P)
_s("orc_bh_forces", None, C.c_size_t, C.c_void_p, C.c_int32, F32P, F32P, F32P, F32P, C.c_float,
   C.c_float, C.c_float, I64P, C.c_size_t, F32P, F32P, F32P, C.c_int)
_s("orc_grid_cell_index", None, C.c_float, C.c_float, C.c_float, C.c_float, I32P)
_s("orc_grid_hash_cell", C.c_int32, I32P, I32P)
_s("orc_grid_dims", C.c_int, C.c_size_t, F32P, F32P, F32P, C.c_float, F32P, I32P, I64P)
_s("orc_grid_build", None, C.c_size_t, F32P, F32P, F32P, C.c_float, F32P, I32P, I32P, I32P, I32P,
   I32P)
_s("orc_grid_forces_f32", None, C.c_size_t, F32P, F32P, F32P, F32P, F32P, I32P, C.c_float, I32P,
   I32P, I32P, C.c_float, C.c_float, C.c_float, I64P, C.c_size_t, F32P, F32P, F32P, C.c_int)
_s("orc_grid_forces_f64", None, C.c_size_t, F32P, F32P, F32P, F32P, F32P, I32P, C.c_float, I32P,
   I32P, I32P, C.c_float, C.c_double, C.c_double, I64P, C.c_size_t, F64P, F64P, F64P, F64P, I64P,
   C.c_int)
_s("orc_cutoff_forces_f64", None, C.c_size_t, F32P, F32P, F32P, F32P, C.c_float, C.c_double,
   C.c_double, I64P, C.c_size_t, F64P, F64P, F64P, I64P, C.c_int)
_s("orc_max_threads", C.c_int)


def _f(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(F32P)


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(F64P)


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def eps2_of(eps: float) -> float:
    """The float eps*eps the reference stores (force_calculator.hpp:15-18)."""
    return float(np.float32(np.float32(eps) * np.float32(eps)))


def max_threads() -> int:
    return lib.orc_max_threads()


def pair_accel(p1, p2, m2, G, eps) -> np.ndarray:
    a, b, out = f32(p1), f32(p2), np.zeros(3, np.float32)
    lib.orc_pair_accel(_f(a), _f(b), m2, G, eps, _f(out))
    return out


def _targets(targets, n):
    if targets is None:
        return None, 0, n
    t = np.ascontiguousarray(targets, dtype=np.int64)
    return t, len(t), len(t)


def direct_f32(x, y, z, m, G, eps, targets=None, nthreads=0) -> np.ndarray:
    x, y, z, m = map(f32, (x, y, z, m))
    t, nt, nout = _targets(targets, len(x))
    out = [np.zeros(nout, np.float32) for _ in range(3)]
    lib.orc_direct_f32(len(x), _f(x), _f(y), _f(z), _f(m), G, eps,
                       t.ctypes.data_as(I64P) if t is not None else None, nt,
                       *map(_f, out), nthreads)
    return np.stack(out, axis=1)


def direct_f64(x, y, z, m, G, eps2, targets=None, nthreads=0) -> np.ndarray:
    x, y, z, m = map(f32, (x, y, z, m))
    t, nt, nout = _targets(targets, len(x))
    out = [np.zeros(nout, np.float64) for _ in range(3)]
    lib.orc_direct_f64(len(x), _f(x), _f(y), _f(z), _f(m), G, eps2,
                       t.ctypes.data_as(I64P) if t is not None else None, nt,
                       *map(_d, out), nthreads)
    return np.stack(out, axis=1)


def direct_f64_pos64(x, y, z, m, G, eps2, targets=None, nthreads=0) -> np.ndarray:
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    m = f32(m)
    t, nt, nout = _targets(targets, len(x))
    out = [np.zeros(nout, np.float64) for _ in range(3)]
    lib.orc_direct_f64_pos64(len(x), _d(x), _d(y), _d(z), _f(m), G, eps2,
                             t.ctypes.data_as(I64P) if t is not None else None, nt,
                             *map(_d, out), nthreads)
    return np.stack(out, axis=1)


def drift(h: dict, dt: float) -> None:
    lib.orc_drift(len(h["mass"]), *(_f(h[k]) for k in ("pos_x", "pos_y", "pos_z", "vel_x",
                  "vel_y", "vel_z", "acc_x", "acc_y", "acc_z")), dt)


def store_old(h: dict) -> None:
    lib.orc_store_old(len(h["mass"]), *(_f(h[k]) for k in ("acc_x", "acc_y", "acc_z",
                      "acc_old_x", "acc_old_y", "acc_old_z")))


def kick(h: dict, dt: float) -> None:
    lib.orc_kick(len(h["mass"]), *(_f(h[k]) for k in ("vel_x", "vel_y", "vel_z", "acc_old_x",
                 "acc_old_y", "acc_old_z", "acc_x", "acc_y", "acc_z")), dt)


def run_direct(h: dict, G: float, eps: float, dt: float, steps: int, nthreads=0) -> None:
    keys = ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "acc_x", "acc_y", "acc_z",
            "acc_old_x", "acc_old_y", "acc_old_z")
    lib.orc_run_direct(len(h["mass"]), *(_f(h[k]) for k in keys), _f(h["mass"]), G, eps, dt,
                       steps, nthreads)


def energy_f64(h: dict, G: float, eps2: float, nthreads=0) -> dict:
    ke, pe = C.c_double(), C.c_double()
    mom = np.zeros(3, np.float64)
    lib.orc_energy_f64(len(h["mass"]), *(_f(h[k]) for k in ("pos_x", "pos_y", "pos_z", "vel_x",
                       "vel_y", "vel_z", "mass")), G, eps2, C.byref(ke), C.byref(pe), _d(mom),
                       nthreads)
    return {"ke": ke.value, "pe": pe.value, "mom": mom}


def bbox(x, y, z):
    x, y, z = map(f32, (x, y, z))
    lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib.orc_bbox(len(x), _f(x), _f(y), _f(z), _f(lo), _f(hi))
    return lo, hi


def morton_codes(x, y, z, lo, hi):
    x, y, z, lo, hi = map(f32, (x, y, z, lo, hi))
    n = len(x)
    codes, idx = np.zeros(n, np.uint32), np.zeros(n, np.int32)
    lib.orc_morton_codes(n, _f(x), _f(y), _f(z), _f(lo), _f(hi), codes.ctypes.data_as(U32P),
                         idx.ctypes.data_as(I32P))
    return codes, idx


def stable_sort(codes, idx):
    c = np.ascontiguousarray(codes, np.uint32).copy()
    i = np.ascontiguousarray(idx, np.int32).copy()
    lib.orc_stable_sort(len(c), c.ctypes.data_as(U32P), i.ctypes.data_as(I32P))
    return c, i


def build_tree(x, y, z, m, sorted_idx, lo, hi, max_nodes):
    x, y, z, m, lo, hi = map(f32, (x, y, z, m, lo, hi))
    si = np.ascontiguousarray(sorted_idx, np.int32)
    nodes = np.zeros(max(int(max_nodes), 1), NODE_DTYPE)
    cnt, depth, nb, nd = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    rc = lib.orc_build_tree(len(x), _f(x), _f(y), _f(z), _f(m), si.ctypes.data_as(I32P), _f(lo),
                            _f(hi), int(max_nodes), nodes.ctypes.data_as(C.c_void_p),
                            C.byref(cnt), C.byref(depth), C.byref(nb), C.byref(nd))
    if rc != 0:
        raise RuntimeError("oracle tree build would overrun the node budget")
    return {"nodes": nodes[:cnt.value].copy(), "node_count": cnt.value, "max_depth": depth.value,
            "budget_merges": nb.value, "depth_merges": nd.value}


def bh_forces(nodes, x, y, z, m, theta, G, eps, targets=None, nthreads=0) -> np.ndarray:
    x, y, z, m = map(f32, (x, y, z, m))
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    t, nt, nout = _targets(targets, len(x))
    out = [np.zeros(nout, np.float32) for _ in range(3)]
    lib.orc_bh_forces(len(x), nodes.ctypes.data_as(C.c_void_p), len(nodes), _f(x), _f(y), _f(z),
                      _f(m), theta, G, eps, t.ctypes.data_as(I64P) if t is not None else None,
                      nt, *map(_f, out), nthreads)
    return np.stack(out, axis=1)


def bh_pipeline(h: dict, max_particles: int | None = None):
    """bbox -> Morton -> stable sort -> tree, as BarnesHutTree::build (:330-336)."""
    x, y, z, m = h["pos_x"], h["pos_y"], h["pos_z"], h["mass"]
    lo, hi = bbox(x, y, z)
    codes, idx = morton_codes(x, y, z, lo, hi)
    scodes, sidx = stable_sort(codes, idx)
    cap = 2 * (max_particles if max_particles is not None else len(x))
    tree = build_tree(x, y, z, m, sidx, lo, hi, cap)
    tree.update({"bbox_min": lo, "bbox_max": hi, "codes": scodes, "sorted_idx": sidx,
                 "unsorted_codes": codes})
    return tree


# ------------------------------------------------------------ spatial hash

def _i32(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(I32P)


def grid_cell_index(x, y, z, cell) -> tuple:
    out = np.zeros(3, np.int32)
    lib.orc_grid_cell_index(x, y, z, cell, _i32(out))
    return tuple(int(v) for v in out)


def grid_hash_cell(cell, dims) -> int:
    c, d = np.asarray(cell, np.int32).copy(), np.asarray(dims, np.int32).copy()
    return int(lib.orc_grid_hash_cell(_i32(c), _i32(d)))


def grid_dims(x, y, z, cell):
    """Padded bbox (6 floats), dims (3 ints) and total cells, or None if the grid
    would exceed 1e8 cells (SpatialHashGrid::build, force_spatial_hash.cu:241-260)."""
    x, y, z = map(f32, (x, y, z))
    bbox, dims, total = np.zeros(6, np.float32), np.zeros(3, np.int32), C.c_int64()
    rc = lib.orc_grid_dims(len(x), _f(x), _f(y), _f(z), cell, _f(bbox), _i32(dims),
                           C.byref(total))
    if rc != 0:
        return None
    return bbox, dims, total.value


def grid_build(x, y, z, cell):
    """Everything SpatialHashGrid::build leaves behind, as a dict."""
    x, y, z = map(f32, (x, y, z))
    sized = grid_dims(x, y, z, cell)
    if sized is None:
        raise RuntimeError("Spatial hash grid too large: reduce cell_size or bounding box")
    bbox, dims, total = sized
    n = len(x)
    pc, srt = np.zeros(n, np.int32), np.zeros(n, np.int32)
    cs, ce = np.zeros(total, np.int32), np.zeros(total, np.int32)
    lib.orc_grid_build(n, _f(x), _f(y), _f(z), cell, _f(bbox), _i32(dims), _i32(pc), _i32(cs),
                       _i32(ce), _i32(srt))
    return {"bbox": bbox, "dims": dims, "total": total, "cell": np.float32(cell),
            "particle_cell": pc, "cell_start": cs, "cell_end": ce, "sorted": srt}


def grid_forces_f32(g: dict, x, y, z, m, cutoff, G, eps, targets=None, nthreads=0):
    x, y, z, m = map(f32, (x, y, z, m))
    t, nt, nout = _targets(targets, len(x))
    out = [np.zeros(nout, np.float32) for _ in range(3)]
    lib.orc_grid_forces_f32(len(x), _f(x), _f(y), _f(z), _f(m), _f(g["bbox"]), _i32(g["dims"]),
                            float(g["cell"]), _i32(g["cell_start"]), _i32(g["cell_end"]),
                            _i32(g["sorted"]), cutoff, G, eps,
                            t.ctypes.data_as(I64P) if t is not None else None, nt,
                            *map(_f, out), nthreads)
    return np.stack(out, axis=1)


def grid_forces_f64(g: dict, x, y, z, m, cutoff, G, eps2, targets=None, nthreads=0):
    """(acc (n,3) float64, sum of |contribution| per target, kept pairs per target)."""
    x, y, z, m = map(f32, (x, y, z, m))
    t, nt, nout = _targets(targets, len(x))
    out = [np.zeros(nout, np.float64) for _ in range(4)]
    np_ = np.zeros(nout, np.int64)
    lib.orc_grid_forces_f64(len(x), _f(x), _f(y), _f(z), _f(m), _f(g["bbox"]), _i32(g["dims"]),
                            float(g["cell"]), _i32(g["cell_start"]), _i32(g["cell_end"]),
                            _i32(g["sorted"]), cutoff, G, eps2,
                            t.ctypes.data_as(I64P) if t is not None else None, nt,
                            *map(_d, out), np_.ctypes.data_as(I64P), nthreads)
    return np.stack(out[:3], axis=1), out[3], np_


def cutoff_forces_f64(x, y, z, m, cutoff, G, eps2, targets=None, nthreads=0):
    """Brute-force cutoff sum (no grid): (acc (n,3) float64, kept pairs per target)."""
    x, y, z, m = map(f32, (x, y, z, m))
    t, nt, nout = _targets(targets, len(x))
    out = [np.zeros(nout, np.float64) for _ in range(3)]
    np_ = np.zeros(nout, np.int64)
    lib.orc_cutoff_forces_f64(len(x), _f(x), _f(y), _f(z), _f(m), cutoff, G, eps2,
                              t.ctypes.data_as(I64P) if t is not None else None, nt,
                              *map(_d, out), np_.ctypes.data_as(I64P), nthreads)
    return np.stack(out, axis=1), np_
