"""ctypes binding of the hipgran C ABI (include/hipgran.h).

The library is libgran_amd/libhipgran.so, built in-tree for gfx950 by
`make -C libgran_amd/csrc`. Loading fails loudly when it is missing: there is
no CPU fallback for any device entry point.
"""
import ctypes as C
import os
import re

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libhipgran.so")
HEADER = os.path.join(ROOT, "include", "hipgran.h")

HG_OK, HG_ERR_INVALID, HG_ERR_OOM, HG_ERR_HIP, HG_ERR_NUMERIC, HG_ERR_CAPACITY, HG_ERR_NO_DEVICE, HG_ERR_IO = range(8)
BINARY_NONE, BINARY_CONTACT, BINARY_SINTER, BINARY_SINTER_LEGACY = 0, 1, 2, 3
HAMAKER_NONE, HAMAKER_FULL, HAMAKER_LIMITED, HAMAKER_SIMPLIFIED = 0, 1, 2, 3
PAIRS_ALL, PAIRS_NEIGHBOR_LIST = 0, 1
PRECISION_FP64, PRECISION_MIXED = 0, 1

_CONTACT_FIELDS = ("k", "gamma_n", "k_t", "gamma_t", "mu_s", "phi_d", "k_r", "gamma_r", "mu_r", "phi_r",
                   "k_o", "gamma_o", "mu_o", "phi_o", "r_part", "mass", "inertia", "dt")
_SINTER_FIELDS = ("k", "gamma_n", "k_t", "gamma_t", "k_r", "gamma_r", "k_o", "gamma_o",
                  "r_part", "mass", "inertia", "dt", "d_crit")
_HAMAKER_FIELDS = ("A", "h0", "r_cutoff", "r_part", "mass")


class ContactParams(C.Structure):
    _fields_ = [(f, C.c_double) for f in _CONTACT_FIELDS]


class SinterParams(C.Structure):
    _fields_ = [(f, C.c_double) for f in _SINTER_FIELDS]


class HamakerParams(C.Structure):
    _fields_ = [(f, C.c_double) for f in _HAMAKER_FIELDS]


class Params(C.Structure):
    _fields_ = [("binary_model", C.c_int32), ("hamaker_model", C.c_int32), ("pair_mode", C.c_int32),
                ("precision", C.c_int32), ("facet_contact", C.c_int32), ("facet_hamaker", C.c_int32),
                ("neighbor_capacity", C.c_uint32), ("pair_capacity", C.c_uint32), ("facet_advance", C.c_int32),
                ("sinter_breaking", C.c_int32), ("r_verlet", C.c_double),
                ("contact", ContactParams), ("sinter", SinterParams), ("hamaker", HamakerParams),
                ("surface_contact", ContactParams), ("surface_hamaker", HamakerParams)]


class ServerConfig(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("max_requests", C.c_uint32), ("max_facets", C.c_uint32),
                ("quantum", C.c_uint32), ("rebuild_every", C.c_uint32), ("dt", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("submitted", "queued", "running", "done", "failed", "rounds",
                                           "particle_steps")] + [("device_ms", C.c_double), ("elapsed_s", C.c_double)]


HG_REQ_QUEUED, HG_REQ_RUNNING, HG_REQ_DONE, HG_REQ_FAILED = range(4)
HG_FAULT_NEIGHBOR_CAPACITY, HG_FAULT_PAIR_CAPACITY, HG_FAULT_NUMERIC = 1, 2, 4
HG_NOALLOC_MAX_PARTICLES = 131072


def fill(struct, values):
    """Sets a params sub-struct from a sequence in ctor-argument order."""
    for (name, _), val in zip(struct._fields_, values):
        setattr(struct, name, float(val))
    return struct


_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_szp = C.POINTER(C.c_size_t)
_sysp = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "hg_last_error": (C.c_char_p, []),
    "hg_create": (C.c_int, [C.POINTER(Params), C.c_size_t, _dp, _dp, _dp, _dp, C.POINTER(_sysp)]),
    "hg_destroy": (C.c_int, [_sysp]),
    "hg_set_facets": (C.c_int, [_sysp, C.c_size_t, _dp, _dp]),
    "hg_get_facets": (C.c_int, [_sysp, _dp, _dp]),
    "hg_advance_facet": (C.c_int, [_sysp, C.c_size_t, _dp, C.c_double]),
    "hg_rebuild_neighbors": (C.c_int, [_sysp]),
    "hg_compute_accelerations": (C.c_int, [_sysp]),
    "hg_step": (C.c_int, [_sysp, C.c_double, C.c_uint64, C.c_uint32]),
    "hg_step_async": (C.c_int, [_sysp, C.c_double, C.c_uint64, C.c_uint32]),
    "hg_get_state": (C.c_int, [_sysp, _dp, _dp, _dp, _dp]),
    "hg_set_state": (C.c_int, [_sysp, _dp, _dp, _dp, _dp]),
    "hg_get_accelerations": (C.c_int, [_sysp, _dp, _dp]),
    "hg_get_time": (C.c_int, [_sysp, _dp, _u64p]),
    "hg_reset_accelerations": (C.c_int, [_sysp]),
    "hg_get_neighbors": (C.c_int, [_sysp, _u64p, _u32p, C.c_size_t, _szp]),
    "hg_get_bonds": (C.c_int, [_sysp, _u32p, C.c_size_t, _szp, _szp]),
    "hg_get_springs": (C.c_int, [_sysp, C.c_int, _u32p, _dp, C.c_size_t, _szp]),
    "hg_diagnostics": (C.c_int, [_sysp, C.c_double, C.c_double, _dp]),
    "hg_energy": (C.c_int, [_sysp, C.c_double, C.c_double, _dp]),
    "hg_sinter_topology": (C.c_int, [C.c_size_t, _dp, C.c_double, C.c_double, _u32p, C.c_size_t, _szp, _szp]),
    "hg_set_neck_strengths": (C.c_int, [_sysp, _dp]),
    "hg_breaking_sinter_topology": (C.c_int, [C.c_size_t, _dp, C.c_double, C.c_double, _u32p, C.c_size_t, _szp]),
    "hg_pool_set_neck_strengths": (C.c_int, [_sysp, C.c_int32, _dp]),
    "hg_profile_enable": (C.c_int, [_sysp, C.c_int]),
    "hg_set_per_step_launches": (C.c_int, [_sysp, C.c_int]),
    "hg_set_dissipation_tracking": (C.c_int, [_sysp, C.c_int]),
    "hg_get_dissipated_work": (C.c_int, [_sysp, _dp]),
    "hg_pool_set_chunk_shape": (C.c_int, [_sysp, C.c_int, C.c_int]),
    "hg_pool_get_chunk_shape": (C.c_int, [_sysp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "hg_profile_read": (C.c_int, [_sysp, _u64p, _dp]),
    "hg_generate_aggregate": (C.c_int, [C.c_size_t, C.c_double, C.c_double, C.c_double, C.c_uint64, _dp]),
    "hg_generate_aggregates": (C.c_int, [C.c_size_t, _u32p, _dp, C.c_double, _u64p, C.c_double, _dp, C.c_int]),
    "hg_legacy_sinter_frames": (C.c_int, [C.c_size_t, _dp, C.c_double, C.c_double, C.c_void_p, C.c_void_p, _u32p,
                                          _dp, C.c_size_t, _szp]),
    "hg_set_legacy_connectors": (C.c_int, [_sysp, C.c_size_t, _u32p, _dp]),
    "hg_get_legacy_connectors": (C.c_int, [_sysp, _dp, C.c_size_t, _szp]),
    "hg_write_vtk_particles": (C.c_int, [C.c_char_p, C.c_size_t, _dp, _dp, C.c_double]),
    "hg_write_vtk_necks": (C.c_int, [C.c_char_p, C.c_size_t, _dp, C.c_size_t, _u32p, C.c_double]),
    "hg_pool_create":(C.c_int, [C.POINTER(Params), C.c_size_t, C.c_size_t, C.c_size_t, C.POINTER(_sysp)]),
    "hg_pool_destroy": (C.c_int, [_sysp]),
    "hg_pool_add": (C.c_int, [_sysp, C.c_size_t, _dp, _dp, _dp, _dp, C.POINTER(C.c_int32)]),
    "hg_pool_remove": (C.c_int, [_sysp, C.c_int32]),
    "hg_pool_set_facets": (C.c_int, [_sysp, C.c_int32, C.c_size_t, _dp, _dp]),
    "hg_pool_step": (C.c_int, [_sysp, C.c_double, C.c_uint64, C.c_uint32]),
    "hg_pool_get_state": (C.c_int, [_sysp, C.c_int32, _dp, _dp, _dp, _dp]),
    "hg_pool_energy": (C.c_int, [_sysp, C.c_int32, C.c_double, C.c_double, _dp]),
    "hg_pool_energy_noalloc": (C.c_int, [_sysp, C.c_int32, C.c_double, C.c_double, _dp]),
    "hg_pool_get_accelerations": (C.c_int, [_sysp, C.c_int32, _dp, _dp]),
    "hg_pool_get_neighbors": (C.c_int, [_sysp, C.c_int32, _u64p, _u32p, C.c_size_t, _szp]),
    "hg_pool_get_springs": (C.c_int, [_sysp, C.c_int32, C.c_int, _u32p, _dp, C.c_size_t, _szp]),
    "hg_pool_get_bonds": (C.c_int, [_sysp, C.c_int32, _u32p, C.c_size_t, _szp, _szp]),
    "hg_pool_info": (C.c_int, [_sysp, C.c_int32, _szp, _dp, _u64p]),
    "hg_pool_bytes_per_particle": (C.c_size_t, [C.POINTER(Params), C.c_size_t]),
    "hg_pool_last_step_profile": (C.c_int, [_sysp, _dp, _dp, _u64p, _u64p]),
    "hg_pool_fault": (C.c_int, [_sysp, C.c_int32, C.POINTER(C.c_int)]),
    "hg_save": (C.c_int, [_sysp, C.c_void_p, C.c_size_t, _szp]),
    "hg_load": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(_sysp)]),
    "hg_pool_save": (C.c_int, [_sysp, C.c_int32, C.c_void_p, C.c_size_t, _szp]),
    "hg_pool_load": (C.c_int, [_sysp, C.c_void_p, C.c_size_t, C.POINTER(C.c_int32)]),
    "hg_pool_set_wide_threshold": (C.c_int, [_sysp, C.c_size_t]),
    "hg_pool_last_wide_profile": (C.c_int, [_sysp, _dp, _u64p]),
    "hg_pool_last_busy": (C.c_int, [_sysp, _dp]),
    "hg_pool_set_wide_cus": (C.c_int, [_sysp, C.c_int]),
    "hg_pool_set_wide_graph": (C.c_int, [_sysp, C.c_int]),
    "hg_pool_set_packing": (C.c_int, [_sysp, C.c_int]),
    "hg_pool_autotune": (C.c_int, [_sysp, C.c_int]),
    "hg_pool_get_tuning": (C.c_int, [_sysp, _szp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "hg_server_create": (C.c_int, [C.POINTER(Params), C.POINTER(ServerConfig), C.POINTER(_sysp)]),
    "hg_server_destroy": (C.c_int, [_sysp]),
    "hg_server_submit": (C.c_int, [_sysp, C.c_size_t, _dp, _dp, _dp, _dp, C.c_size_t, _dp, _dp, C.c_uint64, _u64p]),
    "hg_server_status": (C.c_int, [_sysp, C.c_uint64, C.POINTER(C.c_int)]),
    "hg_server_wait": (C.c_int, [_sysp, C.c_uint64, C.c_double, _dp, _dp, _dp, _dp, _dp]),
    "hg_server_get_stats": (C.c_int, [_sysp, C.POINTER(ServerStats)]),
}

_lib = None


def header_symbols():
    """Function names declared in include/hipgran.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hg_[a-z0-9_]+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"hipgran library not built: {LIB_PATH} is missing (run `make -C libgran_amd/csrc` "
                              "or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class HipgranError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"hipgran status {status}: {msg}")
        self.status = status


def check(status):
    if status != HG_OK:
        raise HipgranError(status, lib().hg_last_error().decode())
    return status


def dptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def u32ptr(a):
    return a.ctypes.data_as(_u32p)


def c64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)
