"""ctypes binding of libnbx.so (the C ABI declared in include/nbx.h).

Host-side mirror used by the Python tests, bench.py and smoke(): it calls the
HIP kernels through the same extern "C" entry points a C++ or other FFI caller
would bind. There is no fallback: if libnbx.so is missing or fails to load,
importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnbx.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libnbx.so not built at {LIB_PATH}; run __graft_entry__.build()")
lib = C.CDLL(LIB_PATH)

F32P = C.POINTER(C.c_float)
I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
U32P = C.POINTER(C.c_uint32)
F64P = C.POINTER(C.c_double)
VP = C.c_void_p

FIELDS = ["pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "acc_x", "acc_y", "acc_z",
          "acc_old_x", "acc_old_y", "acc_old_z", "mass"]


class Particles(C.Structure):
    """nbx_particles == nbody::ParticleData (13 float pointers + size_t count)."""
    _fields_ = [(f, F32P) for f in FIELDS] + [("count", C.c_size_t)]


class DSParticles(C.Structure):
    _fields_ = [("p", Particles), ("pos_lo_x", F32P), ("pos_lo_y", F32P), ("pos_lo_z", F32P)]


# 76-byte node, layout of nbody::OctreeNode (include/nbody/barnes_hut_tree.hpp:9-30)
NODE_DTYPE = np.dtype([
    ("center", "<f4", 3), ("half_size", "<f4"), ("com", "<f4", 3), ("total_mass", "<f4"),
    ("children", "<i4", 8), ("particle_index", "<i4"), ("is_leaf", "u1"), ("pad", "u1", 3),
    ("particle_count", "<i4"),
])
assert NODE_DTYPE.itemsize == 76


class NbxError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"nbx status {status}: {msg}")
        self.status = status


def _sig(name, restype, *argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)
    return fn


PP = C.POINTER(Particles)
DSP = C.POINTER(DSParticles)
_sig("nbx_status_string", C.c_char_p, C.c_int)
_sig("nbx_last_error", C.c_char_p)
_sig("nbx_device_info", C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_int),
     C.POINTER(C.c_size_t))
_sig("nbx_malloc", C.c_int, C.POINTER(VP), C.c_size_t)
_sig("nbx_free", C.c_int, VP)
_sig("nbx_memcpy_htod", C.c_int, VP, VP, C.c_size_t)
_sig("nbx_memcpy_dtoh", C.c_int, VP, VP, C.c_size_t)
_sig("nbx_memcpy_dtod", C.c_int, VP, VP, C.c_size_t)
_sig("nbx_synchronize", C.c_int)
_sig("nbx_stream_create", C.c_int, C.POINTER(VP))
_sig("nbx_stream_destroy", C.c_int, VP)
_sig("nbx_stream_synchronize", C.c_int, VP)
_sig("nbx_event_create", C.c_int, C.POINTER(VP))
_sig("nbx_event_destroy", C.c_int, VP)
_sig("nbx_event_record", C.c_int, VP, VP)
_sig("nbx_event_elapsed_ms", C.c_int, VP, VP, C.POINTER(C.c_float))
_sig("nbx_particles_alloc", C.c_int, PP, C.c_size_t)
_sig("nbx_particles_free", C.c_int, PP)
_sig("nbx_particles_upload", C.c_int, PP, PP)
_sig("nbx_particles_download", C.c_int, PP, PP)
_sig("nbx_direct_forces", C.c_int, PP, C.c_float, C.c_float, C.c_int, VP)
_sig("nbx_direct_forces_range", C.c_int, PP, C.c_size_t, C.c_size_t, C.c_float, C.c_float, VP)
_sig("nbx_direct_step", C.c_int, PP, C.c_float, C.c_float, C.c_float, VP)
_sig("nbx_direct_forces_kick", C.c_int, PP, C.c_float, C.c_float, C.c_float, VP)
_sig("nbx_direct_create", C.c_int, C.c_size_t, C.c_size_t, C.POINTER(VP))
_sig("nbx_direct_destroy", C.c_int, VP)
_sig("nbx_direct_compute", C.c_int, VP, PP, C.c_float, C.c_float, VP)
_sig("nbx_direct_compute_kick", C.c_int, VP, PP, C.c_float, C.c_float, C.c_float, VP)
_sig("nbx_direct_compute_step", C.c_int, VP, PP, C.c_float, C.c_float, C.c_float, VP)
_sig("nbx_direct_compute_noalloc", C.c_int, PP, C.c_float, C.c_float, VP)
_sig("nbx_direct_plan", C.c_int, VP, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_int),
     C.POINTER(C.c_size_t))
_sig("nbx_direct_variant_count", C.c_int)
_sig("nbx_direct_variant_name", C.c_int, C.c_int, C.c_char_p, C.c_size_t)
_sig("nbx_direct_set_variant", C.c_int, C.c_int)
_sig("nbx_direct_variant_for", C.c_int, C.c_size_t)
_sig("nbx_direct_autotune", C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_int), F64P)
_sig("nbx_store_old_accelerations", C.c_int, PP, VP)
_sig("nbx_update_positions", C.c_int, PP, C.c_float, VP)
_sig("nbx_update_velocities", C.c_int, PP, C.c_float, VP)
_sig("nbx_verlet_drift", C.c_int, PP, C.c_float, VP)
_sig("nbx_energy", C.c_int, PP, C.c_float, C.c_float, F64P, F64P, F64P, VP)
_sig("nbx_bh_create", C.c_int, C.c_size_t, C.POINTER(VP))
_sig("nbx_bh_destroy", C.c_int, VP)
_sig("nbx_bh_build", C.c_int, VP, PP, VP)
_sig("nbx_bh_forces", C.c_int, VP, PP, C.c_float, C.c_float, C.c_float, VP)
_sig("nbx_bh_node_count", C.c_int, VP, I32P, I32P)
_sig("nbx_bh_copy_nodes", C.c_int, VP, VP, C.c_size_t)
_sig("nbx_bh_copy_sort", C.c_int, VP, F32P, U32P, I32P, C.c_size_t)
_sig("nbx_bh_build_info", C.c_int, VP, I32P, I32P)
_sig("nbx_bh_set_timing", C.c_int, VP, C.c_int)
_sig("nbx_bh_set_builder", C.c_int, VP, C.c_int)
_sig("nbx_bh_set_task_order", C.c_int, VP, C.c_int)
_sig("nbx_bh_phase_ms", C.c_int, VP, F64P)
_sig("nbx_bh_walk_census", C.c_int, VP, PP, C.c_float, C.c_float, C.c_float, I64P)
_sig("nbx_grid_create", C.c_int, C.c_size_t, C.c_float, C.POINTER(VP))
_sig("nbx_grid_destroy", C.c_int, VP)
_sig("nbx_grid_build", C.c_int, VP, PP, VP)
_sig("nbx_grid_forces", C.c_int, VP, PP, C.c_float, C.c_float, C.c_float, VP)
_sig("nbx_grid_info", C.c_int, VP, I32P, F32P, I32P, F32P)
_sig("nbx_grid_copy_cells", C.c_int, VP, I32P, I32P, I32P, I32P)
_sig("nbx_grid_set_walk", C.c_int, VP, C.c_int)
_sig("nbx_ds_alloc", C.c_int, DSP, C.c_size_t)
_sig("nbx_ds_free", C.c_int, DSP)
_sig("nbx_ds_forces", C.c_int, VP, DSP, C.c_float, C.c_float, VP)
_sig("nbx_ds_set_kernel", C.c_int, C.c_int)
_sig("nbx_ds_drift", C.c_int, DSP, C.c_float, VP)
_sig("nbx_ds_step", C.c_int, VP, DSP, C.c_float, C.c_float, C.c_float, VP)
_sig("nbx_ds_energy", C.c_int, DSP, C.c_float, C.c_float, F64P, F64P, F64P, VP)
_sig("nbx_ens_create", C.c_int, C.c_int, I64P, C.POINTER(VP))
_sig("nbx_ens_destroy", C.c_int, VP)
_sig("nbx_ens_particles", C.c_int, VP, PP, I64P)
_sig("nbx_ens_forces", C.c_int, VP, C.c_float, C.c_float, VP)
_sig("nbx_ens_steps", C.c_int, VP, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, VP)
_sig("nbx_ens_steps_sequential", C.c_int, VP, C.c_float, C.c_float, C.c_float, C.c_int, VP)
_sig("nbx_ens_variant_count", C.c_int)
_sig("nbx_ens_autotune", C.c_int, VP, C.c_int, C.POINTER(C.c_int), F64P)
_sig("nbx_ens_set_variant", C.c_int, VP, C.c_int)
_sig("nbx_ens_energy", C.c_int, VP, C.c_float, C.c_float, F64P, F64P, VP)
_sig("nbx_gen_plummer", C.c_int, C.c_size_t, C.c_uint64, C.c_double, C.c_double, C.c_double,
     *([F32P] * 7))
_sig("nbx_gen_uniform_sphere", C.c_int, C.c_size_t, C.c_uint64, C.c_double, C.c_double,
     C.c_double, *([F32P] * 7))
_sig("nbx_gen_two_plummer", C.c_int, C.c_size_t, C.c_uint64, C.c_double, C.c_double, C.c_double,
     C.c_double, *([F32P] * 7))
_sig("nbx_gen_ensemble_sizes", C.c_int, C.c_int, C.c_uint64, C.c_int64, C.c_int64, I64P)

EXPORTED = [
    "nbx_status_string", "nbx_last_error", "nbx_device_info", "nbx_malloc", "nbx_free",
    "nbx_memcpy_htod", "nbx_memcpy_dtoh", "nbx_memcpy_dtod", "nbx_synchronize",
    "nbx_stream_create", "nbx_stream_destroy", "nbx_stream_synchronize", "nbx_event_create",
    "nbx_event_destroy", "nbx_event_record", "nbx_event_elapsed_ms", "nbx_particles_alloc",
    "nbx_particles_free", "nbx_particles_upload", "nbx_particles_download", "nbx_direct_forces",
    "nbx_direct_forces_range", "nbx_direct_step", "nbx_direct_forces_kick",
    "nbx_direct_create", "nbx_direct_destroy", "nbx_direct_compute", "nbx_direct_compute_kick",
    "nbx_direct_compute_step", "nbx_direct_plan", "nbx_direct_compute_noalloc",
    "nbx_direct_variant_count",
    "nbx_direct_variant_name", "nbx_direct_set_variant", "nbx_direct_variant_for",
    "nbx_direct_autotune",
    "nbx_store_old_accelerations", "nbx_update_positions", "nbx_update_velocities",
    "nbx_verlet_drift", "nbx_energy", "nbx_bh_create", "nbx_bh_destroy", "nbx_bh_build",
    "nbx_bh_forces", "nbx_bh_node_count", "nbx_bh_copy_nodes", "nbx_bh_copy_sort",
    "nbx_bh_build_info", "nbx_bh_set_timing", "nbx_bh_set_builder", "nbx_bh_set_task_order", "nbx_bh_phase_ms",
    "nbx_bh_walk_census",
    "nbx_grid_create", "nbx_grid_destroy", "nbx_grid_build", "nbx_grid_forces", "nbx_grid_info",
    "nbx_grid_copy_cells", "nbx_grid_set_walk",
    "nbx_ds_alloc", "nbx_ds_free", "nbx_ds_forces", "nbx_ds_set_kernel", "nbx_ds_drift",
    "nbx_ds_step",
    "nbx_ds_energy", "nbx_ens_create", "nbx_ens_destroy", "nbx_ens_particles", "nbx_ens_forces",
    "nbx_ens_steps", "nbx_ens_steps_sequential", "nbx_ens_variant_count", "nbx_ens_autotune",
    "nbx_ens_set_variant", "nbx_ens_energy", "nbx_gen_plummer", "nbx_gen_uniform_sphere",
    "nbx_gen_two_plummer", "nbx_gen_ensemble_sizes",
]


def check(status: int) -> None:
    if status != 0:
        raise NbxError(status, lib.nbx_last_error().decode(errors="replace"))


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(F32P)


# --------------------------------------------------------------- host data

def empty_host(n: int) -> dict:
    return {f: np.zeros(n, dtype=np.float32) for f in FIELDS}


def host_struct(h: dict) -> Particles:
    n = len(h["mass"])
    p = Particles()
    for f in FIELDS:
        arr = h.setdefault(f, np.zeros(n, dtype=np.float32))
        setattr(p, f, fptr(arr))
    p.count = n
    return p


def _gen(fn, n, *args) -> dict:
    h = empty_host(n)
    outs = [fptr(h[f]) for f in ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "mass")]
    st = fn(n, *args, *outs)
    if st != 0:
        raise NbxError(st, "generator rejected its arguments")
    return h


def gen_plummer(n: int, seed: int = 42, a: float = 1.0, M: float = 1.0,
                mass_cut: float = 0.999) -> dict:
    return _gen(lib.nbx_gen_plummer, n, seed, a, M, mass_cut)


def gen_uniform_sphere(n: int, seed: int = 42, R: float = 1.0, M: float = 1.0,
                       sigma: float = 0.05) -> dict:
    return _gen(lib.nbx_gen_uniform_sphere, n, seed, R, M, sigma)


def gen_two_plummer(n: int, seed: int = 42, a: float = 1.0, M_each: float = 1.0,
                    sep: float = 10.0, vrel: float = 0.5) -> dict:
    return _gen(lib.nbx_gen_two_plummer, n, seed, a, M_each, sep, vrel)


def gen_ensemble_sizes(nsys: int, seed: int = 42, nmin: int = 256, nmax: int = 16384) -> np.ndarray:
    out = np.zeros(nsys, dtype=np.int64)
    st = lib.nbx_gen_ensemble_sizes(nsys, seed, nmin, nmax, out.ctypes.data_as(I64P))
    if st != 0:
        raise NbxError(st, "bad ensemble size arguments")
    return out


# --------------------------------------------------------------- device data

class DeviceParticles:
    """Owns an nbx_particles set in device memory (ParticleDataManager semantics)."""

    def __init__(self, n: int):
        self.n = n
        self.s = Particles()
        check(lib.nbx_particles_alloc(C.byref(self.s), n))

    @classmethod
    def from_host(cls, h: dict) -> "DeviceParticles":
        d = cls(len(h["mass"]))
        d.upload(h)
        return d

    def upload(self, h: dict) -> None:
        hs = host_struct(h)
        check(lib.nbx_particles_upload(C.byref(self.s), C.byref(hs)))

    def download(self) -> dict:
        h = empty_host(self.n)
        hs = host_struct(h)
        check(lib.nbx_particles_download(C.byref(hs), C.byref(self.s)))
        return h

    def read(self, field: str) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float32)
        check(lib.nbx_memcpy_dtoh(out.ctypes.data, C.cast(getattr(self.s, field), VP), self.n * 4))
        return out

    def write(self, field: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float32)
        assert a.shape == (self.n,)
        check(lib.nbx_memcpy_htod(C.cast(getattr(self.s, field), VP), a.ctypes.data, self.n * 4))

    def acc(self) -> np.ndarray:
        return np.stack([self.read("acc_x"), self.read("acc_y"), self.read("acc_z")], axis=1)

    @property
    def ptr(self):
        return C.byref(self.s)

    def free(self) -> None:
        if self.s.count or self.s.pos_x:
            check(lib.nbx_particles_free(C.byref(self.s)))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.h = VP()
        check(lib.nbx_stream_create(C.byref(self.h)))

    def synchronize(self) -> None:
        check(lib.nbx_stream_synchronize(self.h))

    def __del__(self):
        try:
            lib.nbx_stream_destroy(self.h)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.h = VP()
        check(lib.nbx_event_create(C.byref(self.h)))

    def record(self, stream=None) -> None:
        check(lib.nbx_event_record(self.h, stream.h if stream is not None else None))

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        check(lib.nbx_event_elapsed_ms(self.h, later.h, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib.nbx_event_destroy(self.h)
        except Exception:
            pass


def synchronize() -> None:
    check(lib.nbx_synchronize())


def device_info() -> dict:
    name = C.create_string_buffer(256)
    cu, clk, mem = C.c_int(), C.c_int(), C.c_size_t()
    check(lib.nbx_device_info(name, 256, C.byref(cu), C.byref(clk), C.byref(mem)))
    return {"name": name.value.decode(), "cu": cu.value, "clock_khz": clk.value,
            "mem": mem.value}


def eps2_of(eps: float) -> float:
    """softening_eps2_ = eps * eps in FP32 (include/nbody/force_calculator.hpp:15-18)."""
    return float(np.float32(np.float32(eps) * np.float32(eps)))


class DirectWorkspace:
    """nbx_direct: the pair-symmetric kernel's workspace, as DirectForceCalculator owns it.

    Allocated once for up to max_particles bodies (scratch_bytes bounds its partial
    sums; 0 takes the one-round need, at most 1 GiB) and freed with the object.
    """

    def __init__(self, max_particles: int, scratch_bytes: int = 0):
        self.h = VP()
        self.max_particles = max_particles
        check(lib.nbx_direct_create(max_particles, scratch_bytes, C.byref(self.h)))

    def forces(self, d: "DeviceParticles", G: float, eps: float, stream=None) -> None:
        check(lib.nbx_direct_compute(self.h, d.ptr, G, eps2_of(eps), _stream(stream)))

    def forces_kick(self, d: "DeviceParticles", G: float, eps: float, dt: float,
                    stream=None) -> None:
        check(lib.nbx_direct_compute_kick(self.h, d.ptr, G, eps2_of(eps), dt, _stream(stream)))

    def step(self, d: "DeviceParticles", G: float, eps: float, dt: float, stream=None) -> None:
        check(lib.nbx_direct_compute_step(self.h, d.ptr, G, eps2_of(eps), dt, _stream(stream)))

    def plan(self, n: int) -> dict:
        v, r, b = C.c_int(), C.c_int(), C.c_size_t()
        check(lib.nbx_direct_plan(self.h, n, C.byref(v), C.byref(r), C.byref(b)))
        return {"variant": v.value, "rounds": r.value, "partial_bytes": b.value}

    def __del__(self):
        try:
            lib.nbx_direct_destroy(self.h)
        except Exception:
            pass


def _stream(stream):
    if stream is None:
        return None
    return stream.h if isinstance(stream, Stream) else stream


def direct_forces(d: DeviceParticles, G: float, eps: float, block_size: int = 256,
                  stream=None, ws: DirectWorkspace | None = None) -> None:
    """Forces of every body; with a workspace the pair-symmetric kernel may run."""
    if ws is not None:
        ws.forces(d, G, eps, stream)
        return
    check(lib.nbx_direct_forces(d.ptr, G, eps2_of(eps), block_size, _stream(stream)))


def direct_step(d: DeviceParticles, G: float, eps: float, dt: float, stream=None,
                ws: DirectWorkspace | None = None) -> None:
    if ws is not None:
        ws.step(d, G, eps, dt, stream)
        return
    check(lib.nbx_direct_step(d.ptr, G, eps2_of(eps), dt, _stream(stream)))


def energy(d: DeviceParticles, G: float, eps: float, potential: bool = True) -> dict:
    eps2 = np.float32(np.float32(eps) * np.float32(eps))
    ke, pe = C.c_double(), C.c_double()
    mom = (C.c_double * 3)()
    check(lib.nbx_energy(d.ptr, G, float(eps2), C.byref(ke), C.byref(pe) if potential else None,
                         mom, None))
    return {"ke": ke.value, "pe": pe.value if potential else None, "mom": list(mom)}


def variant_names() -> list:
    out = []
    buf = C.create_string_buffer(64)
    for v in range(lib.nbx_direct_variant_count()):
        check(lib.nbx_direct_variant_name(v, buf, 64))
        out.append(buf.value.decode())
    return out


BH_PHASES = ("bbox", "morton", "sort", "paths", "tree", "com", "export", "forces")


class BHTree:
    """nbx_bh handle: BarnesHutTree(max_particles) with a 2*max_particles node budget."""

    def __init__(self, max_particles: int):
        self.h = VP()
        check(lib.nbx_bh_create(max_particles, C.byref(self.h)))

    def build(self, d: DeviceParticles, stream=None) -> None:
        check(lib.nbx_bh_build(self.h, d.ptr, stream.h if stream is not None else None))

    def forces(self, d: DeviceParticles, theta: float, G: float, eps: float, stream=None) -> None:
        check(lib.nbx_bh_forces(self.h, d.ptr, theta, G, eps,
                                stream.h if stream is not None else None))

    def node_count(self):
        c, dep = C.c_int32(), C.c_int32()
        check(lib.nbx_bh_node_count(self.h, C.byref(c), C.byref(dep)))
        return c.value, dep.value

    def nodes(self) -> np.ndarray:
        cnt, _ = self.node_count()
        out = np.zeros(max(cnt, 1), NODE_DTYPE)
        check(lib.nbx_bh_copy_nodes(self.h, out.ctypes.data, len(out)))
        return out[:cnt]

    def sort_result(self, n: int):
        bbox = np.zeros(6, np.float32)
        codes = np.zeros(n, np.uint32)
        idx = np.zeros(n, np.int32)
        check(lib.nbx_bh_copy_sort(self.h, fptr(bbox), codes.ctypes.data_as(U32P),
                                   idx.ctypes.data_as(I32P), n))
        return bbox[:3], bbox[3:], codes, idx

    def build_info(self):
        b, m = C.c_int32(), C.c_int32()
        check(lib.nbx_bh_build_info(self.h, C.byref(b), C.byref(m)))
        return {"builder": "sequential" if b.value else "parallel", "depth32_leaves": m.value}

    def set_timing(self, on: bool) -> None:
        check(lib.nbx_bh_set_timing(self.h, 1 if on else 0))

    def set_builder(self, mode: int) -> None:
        check(lib.nbx_bh_set_builder(self.h, mode))

    def set_task_order(self, mode: int) -> None:
        check(lib.nbx_bh_set_task_order(self.h, mode))

    def walk_census(self, d: DeviceParticles, theta: float, G: float, eps: float) -> dict:
        c = np.zeros(4, np.int64)
        check(lib.nbx_bh_walk_census(self.h, d.ptr, theta, G, eps, c.ctypes.data_as(I64P)))
        return {"wave_steps": int(c[0]), "interactions": int(c[1]), "openings": int(c[2]),
                "active_lane_steps": int(c[3])}

    def phase_ms(self) -> dict:
        ms = (C.c_double * 8)()
        check(lib.nbx_bh_phase_ms(self.h, ms))
        return dict(zip(BH_PHASES, list(ms)))

    def __del__(self):
        try:
            lib.nbx_bh_destroy(self.h)
        except Exception:
            pass


class Grid:
    """nbx_grid handle: SpatialHashGrid(max_particles, cell_size)."""

    def __init__(self, max_particles: int, cell_size: float):
        self.h = VP()
        self.max_particles = max_particles
        check(lib.nbx_grid_create(max_particles, cell_size, C.byref(self.h)))

    def build(self, d: DeviceParticles, stream=None) -> None:
        check(lib.nbx_grid_build(self.h, d.ptr, stream.h if stream is not None else None))

    def forces(self, d: DeviceParticles, cutoff: float, G: float, eps: float, stream=None) -> None:
        check(lib.nbx_grid_forces(self.h, d.ptr, cutoff, G, eps,
                                  stream.h if stream is not None else None))

    def info(self) -> dict:
        dims, cell, total = (C.c_int32 * 3)(), C.c_float(), C.c_int32()
        bbox = np.zeros(6, np.float32)
        check(lib.nbx_grid_info(self.h, C.cast(dims, I32P), C.byref(cell), C.byref(total),
                                fptr(bbox)))
        return {"dims": np.array(list(dims), np.int32), "cell": np.float32(cell.value),
                "total": total.value, "bbox": bbox}

    def cells(self) -> dict:
        """copyCellDataToHost: cell_start/cell_end (total) and particle_cells/sorted (max_particles)."""
        total = self.info()["total"]
        cs, ce = np.zeros(total, np.int32), np.zeros(total, np.int32)
        pc, srt = np.zeros(self.max_particles, np.int32), np.zeros(self.max_particles, np.int32)
        p = lambda a: a.ctypes.data_as(I32P)
        check(lib.nbx_grid_copy_cells(self.h, p(cs), p(ce), p(pc), p(srt)))
        return {"cell_start": cs, "cell_end": ce, "particle_cell": pc, "sorted": srt}

    def set_walk(self, mode: int) -> None:
        check(lib.nbx_grid_set_walk(self.h, mode))

    def __del__(self):
        try:
            lib.nbx_grid_destroy(self.h)
        except Exception:
            pass


class DSDeviceParticles:
    """Precision-form particles: nbx_ds_particles (pos hi arrays + lo arrays)."""

    def __init__(self, n: int):
        self.n = n
        self.s = DSParticles()
        check(lib.nbx_ds_alloc(C.byref(self.s), n))

    @classmethod
    def from_host(cls, h: dict, pos64=None) -> "DSDeviceParticles":
        d = cls(len(h["mass"]))
        hs = host_struct(h)
        check(lib.nbx_particles_upload(C.byref(d.s.p), C.byref(hs)))
        if pos64 is not None:
            d.set_positions64(*pos64)
        return d

    def set_positions64(self, x, y, z) -> None:
        for hi_name, lo_name, v in (("pos_x", "pos_lo_x", x), ("pos_y", "pos_lo_y", y),
                                    ("pos_z", "pos_lo_z", z)):
            v = np.asarray(v, np.float64)
            hi = v.astype(np.float32)
            lo = (v - hi.astype(np.float64)).astype(np.float32)
            self._write(getattr(self.s.p, hi_name), hi)
            self._write(getattr(self.s, lo_name), lo)

    def _write(self, ptr, arr):
        a = np.ascontiguousarray(arr, np.float32)
        check(lib.nbx_memcpy_htod(C.cast(ptr, VP), a.ctypes.data, self.n * 4))

    def _read(self, ptr):
        out = np.empty(self.n, np.float32)
        check(lib.nbx_memcpy_dtoh(out.ctypes.data, C.cast(ptr, VP), self.n * 4))
        return out

    def read(self, field: str) -> np.ndarray:
        if field.startswith("pos_lo"):
            return self._read(getattr(self.s, field))
        return self._read(getattr(self.s.p, field))

    def positions64(self):
        return [self.read(h).astype(np.float64) + self.read(l).astype(np.float64)
                for h, l in (("pos_x", "pos_lo_x"), ("pos_y", "pos_lo_y"), ("pos_z", "pos_lo_z"))]

    def acc(self) -> np.ndarray:
        return np.stack([self.read("acc_x"), self.read("acc_y"), self.read("acc_z")], axis=1)

    @property
    def ptr(self):
        return C.byref(self.s)

    def __del__(self):
        try:
            lib.nbx_ds_free(C.byref(self.s))
        except Exception:
            pass


class Ensemble:
    """nbx_ensemble: independent systems concatenated in one device SoA set."""

    def __init__(self, counts):
        self.counts = np.ascontiguousarray(counts, np.int64)
        self.nsys = len(self.counts)
        self.h = VP()
        check(lib.nbx_ens_create(self.nsys, self.counts.ctypes.data_as(I64P), C.byref(self.h)))
        self.view = Particles()
        self.offsets = np.zeros(self.nsys + 1, np.int64)
        check(lib.nbx_ens_particles(self.h, C.byref(self.view), self.offsets.ctypes.data_as(I64P)))
        self.total = int(self.offsets[-1])

    def upload(self, h: dict) -> None:
        hs = host_struct(h)
        check(lib.nbx_particles_upload(C.byref(self.view), C.byref(hs)))

    def read(self, field: str) -> np.ndarray:
        out = np.empty(self.total, np.float32)
        check(lib.nbx_memcpy_dtoh(out.ctypes.data, C.cast(getattr(self.view, field), VP),
                                  self.total * 4))
        return out

    def system(self, q: int) -> slice:
        return slice(int(self.offsets[q]), int(self.offsets[q + 1]))

    def forces(self, G: float, eps2: float, stream=None) -> None:
        check(lib.nbx_ens_forces(self.h, G, eps2, stream.h if stream is not None else None))

    def steps(self, G, eps2, dt, steps, graph_chunk=0, stream=None) -> None:
        check(lib.nbx_ens_steps(self.h, G, eps2, dt, steps, graph_chunk,
                                stream.h if stream is not None else None))

    def steps_sequential(self, G, eps2, dt, steps, stream=None) -> None:
        check(lib.nbx_ens_steps_sequential(self.h, G, eps2, dt, steps,
                                           stream.h if stream is not None else None))

    def set_variant(self, v: int) -> None:
        check(lib.nbx_ens_set_variant(self.h, v))

    def autotune(self, reps=3):
        b, ms = C.c_int(), C.c_double()
        check(lib.nbx_ens_autotune(self.h, reps, C.byref(b), C.byref(ms)))
        return b.value, ms.value

    def energy(self, G, eps2):
        ke = np.zeros(self.nsys)
        pe = np.zeros(self.nsys)
        check(lib.nbx_ens_energy(self.h, G, eps2, ke.ctypes.data_as(F64P), pe.ctypes.data_as(F64P),
                                 None))
        return ke, pe

    def __del__(self):
        try:
            lib.nbx_ens_destroy(self.h)
        except Exception:
            pass


def gen_ensemble(counts, seed: int = 42) -> dict:
    """Concatenated Plummer systems (a=1, M=1 each), system q seeded with seed + q."""
    parts = [gen_plummer(int(c), seed=seed + q) for q, c in enumerate(counts)]
    return {f: np.concatenate([p[f] for p in parts]) for f in FIELDS}
