"""ctypes access to the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker, never the product.
"""
import ctypes as C
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def u32ptr(a):
    return a.ctypes.data_as(_u32p)


# ------------------------------------------------------------ constants -----
# Material and model constants of the reference drivers (test/contact.cpp:31-51,
# test/alt_sintered.cpp:38-66, test/hamaker.cpp:55-57).
R_PART = 1.4e-8
RHO = 1700.0
MASS = 4.0 / 3.0 * math.pi * R_PART ** 3.0 * RHO
INERTIA = 2.0 / 5.0 * MASS * R_PART ** 2.0


def contact_params(k=10000.0, gamma_n=5.0e-9, mu=1.0, phi=1.0, mu_o=0.1, dt=1e-13,
                   r_part=R_PART, mass=MASS, inertia=INERTIA):
    """contact_force_functor ctor order (contact_force.h:10-30) without n_part."""
    gamma_t, gamma_r, gamma_o = 0.2 * gamma_n, 0.05 * gamma_n, 0.05 * gamma_n
    return np.array([k, gamma_n, k, gamma_t, mu, phi, k, gamma_r, mu_o, phi, k, gamma_o, mu_o, phi,
                     r_part, mass, inertia, dt], dtype=np.float64)


def sinter_params(k_bond=1000.0, d_crit=1.0e-9, dt=1e-13, r_part=R_PART, mass=MASS, inertia=INERTIA):
    """alt_sinter_functor ctor order (alt_sinter_bridge.h:19-36) + critical_separation."""
    g = 2.0 * math.sqrt(2.0 * mass * k_bond)
    return np.array([k_bond, g, k_bond, 0.2 * g, k_bond, 0.05 * g, k_bond, 0.05 * g,
                     r_part, mass, inertia, dt, d_crit], dtype=np.float64)


def hamaker_params(A=1.0e-20, h0=1.0e-9, r_cutoff=3.0 * R_PART, r_part=R_PART, mass=MASS):
    return np.array([A, h0, r_cutoff, r_part, mass], dtype=np.float64)


# ------------------------------------------------------- system params ------
class ContactP(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "k", "gamma_n", "k_t", "gamma_t", "mu_s", "phi_d", "k_r", "gamma_r", "mu_r", "phi_r",
        "k_o", "gamma_o", "mu_o", "phi_o", "r_part", "mass", "inertia", "dt")]


class SinterP(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "k", "gamma_n", "k_t", "gamma_t", "k_r", "gamma_r", "k_o", "gamma_o",
        "r_part", "mass", "inertia", "dt", "d_crit")]


class HamakerP(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("A", "h0", "r_cutoff", "r_part", "mass")]


class SystemP(C.Structure):
    _fields_ = [("binary_model", C.c_int), ("hamaker_model", C.c_int), ("neighbor_mode", C.c_int),
                ("facet_contact", C.c_int), ("facet_hamaker", C.c_int), ("sinter_breaking", C.c_int),
                ("r_verlet", C.c_double),
                ("contact", ContactP), ("sinter", SinterP), ("hamaker", HamakerP),
                ("surface_contact", ContactP), ("surface_hamaker", HamakerP)]


def _fill(struct, arr):
    for (name, _), val in zip(struct._fields_, arr):
        setattr(struct, name, float(val))
    return struct


BIN_NONE, BIN_CONTACT, BIN_SINTER, BIN_SINTER_LEGACY = 0, 1, 2, 3
HAM_NONE, HAM_FULL, HAM_LIMITED, HAM_SIMPLIFIED = 0, 1, 2, 3


def system_params(binary_model, hamaker_model, neighbor_mode=0, r_verlet=0.0, contact=None, sinter=None,
                  hamaker=None, facet_contact=0, facet_hamaker=0, surface_contact=None, surface_hamaker=None,
                  sinter_breaking=0):
    p = SystemP()
    p.sinter_breaking = sinter_breaking
    p.binary_model, p.hamaker_model, p.neighbor_mode = binary_model, hamaker_model, neighbor_mode
    p.facet_contact, p.facet_hamaker, p.r_verlet = facet_contact, facet_hamaker, r_verlet
    _fill(p.contact, contact if contact is not None else contact_params())
    _fill(p.sinter, sinter if sinter is not None else sinter_params())
    _fill(p.hamaker, hamaker if hamaker is not None else hamaker_params())
    _fill(p.surface_contact, surface_contact if surface_contact is not None else contact_params())
    _fill(p.surface_hamaker, surface_hamaker if surface_hamaker is not None else hamaker_params())
    return p


# ------------------------------------------------------------- loading ------
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        L.or_sys_create.restype = C.c_void_p
        L.or_sys_create.argtypes = [C.POINTER(SystemP), C.c_size_t, _dp, _dp, _dp, _dp]
        for name in ("or_sys_destroy", "or_sys_compute_accelerations"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.or_sys_rebuild_neighbors.argtypes = [C.c_void_p, C.c_int]
        L.or_sys_set_facets.argtypes = [C.c_void_p, C.c_size_t, _dp, _dp]
        L.or_sys_step.argtypes = [C.c_void_p, C.c_double, C.c_uint64, C.c_uint32, _u64p]
        L.or_sys_get_state.argtypes = [C.c_void_p] + [_dp] * 6
        L.or_sys_set_state.argtypes = [C.c_void_p] + [_dp] * 4
        L.or_sys_time.restype = C.c_double
        L.or_sys_time.argtypes = [C.c_void_p]
        L.or_sys_get_neighbors.restype = C.c_size_t
        L.or_sys_get_neighbors.argtypes = [C.c_void_p, _u64p, _u32p, C.c_size_t]
        L.or_sys_get_bonds.restype = C.c_size_t
        L.or_sys_get_bonds.argtypes = [C.c_void_p, _u32p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.or_sys_get_springs.restype = C.c_size_t
        L.or_sys_get_springs.argtypes = [C.c_void_p, C.c_int, _u32p, _dp, C.c_size_t]
        L.or_sys_diagnostics.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
        L.or_sys_pair_magnitudes.argtypes = [C.c_void_p, _dp]
        L.or_sys_functor_magnitudes.argtypes = [C.c_void_p, _dp]
        L.or_sys_energy_terms.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
        L.or_contact_seq.argtypes = [_dp, C.c_size_t, _dp, _dp, _dp]
        L.or_sinter_seq.argtypes = [_dp, _dp, C.c_size_t, _dp, _dp, _dp, C.POINTER(C.c_int)]
        L.or_hamaker_seq.argtypes = [C.c_int, _dp, C.c_size_t, _dp, _dp]
        L.or_fold_seq.argtypes = [_dp, _dp, _dp, C.c_size_t, _dp, _dp]
        L.or_facet_seq.argtypes = [_dp, _dp, _dp, _dp, C.c_double, C.c_size_t, _dp, _dp]
        L.or_closest_point.argtypes = [_dp, C.c_size_t, _dp, _dp]
        L.or_breaking_seq.argtypes = [_dp, _dp, C.c_double, C.c_size_t, _dp, _dp, C.POINTER(C.c_int)]
        L.or_sys_set_neck_strengths.argtypes = [C.c_void_p, _dp]
        L.or_sys_set_dissipation_tracking.argtypes = [C.c_void_p, C.c_int]
        L.or_sys_dissipated_work.argtypes = [C.c_void_p, _dp]
        _szp = C.POINTER(C.c_size_t)
        L.or_legacy_frames.restype = C.c_size_t
        L.or_legacy_frames.argtypes = [C.c_size_t, _dp, C.c_double, C.c_double, _dp, C.c_size_t, _u32p, _dp,
                                       C.c_size_t, _szp]
        L.or_sys_legacy_set.argtypes = [C.c_void_p, C.c_size_t, _u32p, _dp]
        L.or_sys_legacy_get.argtypes = [C.c_void_p, _dp]
        L.or_legacy_seq.argtypes = [C.c_size_t, _dp, _dp, _dp, _dp, _dp, C.c_size_t, C.c_size_t, _dp, _dp, _szp,
                                    _szp]
        L.or_sinter_topology.restype = C.c_size_t
        L.or_sinter_topology.argtypes = [C.c_size_t, _dp, C.c_double, C.c_double, _u32p, C.c_size_t,
                                         C.POINTER(C.c_size_t)]
        _lib = L
    return _lib


# ------------------------------------------------------ sequence drivers ----
def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def oracle_contact_seq(params, states):
    states = c64(states)
    n = states.shape[0]
    out = np.zeros((n, 12))
    springs = np.zeros((n, 18))
    lib().or_contact_seq(dptr(c64(params)), n, dptr(states), dptr(out), dptr(springs))
    return out, springs


def oracle_sinter_seq(sinter, contact, states):
    states = c64(states)
    n = states.shape[0]
    out = np.zeros((n, 12))
    springs = np.zeros((n, 18))
    b = C.c_int(0)
    lib().or_sinter_seq(dptr(c64(sinter)), dptr(c64(contact)), n, dptr(states), dptr(out), dptr(springs),
                        C.byref(b))
    return out, springs, b.value


def _breaking(fn, sinter, contact, strength, states):
    states = c64(states)
    n = states.shape[0]
    out = np.zeros((n, 6))
    broken = np.zeros(n, dtype=np.int32)
    fn(dptr(c64(sinter)), dptr(c64(contact)), float(strength), n, dptr(states), dptr(out),
       broken.ctypes.data_as(C.POINTER(C.c_int)))
    return out, broken


def oracle_breaking_seq(sinter, contact, strength, states):
    """alt_breaking_sinter on the ordered pair (0, 1): (a, alpha) per step and
    whether the bond has broken after that step."""
    return _breaking(lib().or_breaking_seq, sinter, contact, strength, states)


def oracle_hamaker_seq(mode, ham, states):
    states = c64(states)
    n = states.shape[0]
    out = np.zeros((n, 12))
    lib().or_hamaker_seq(mode, dptr(c64(ham)), n, dptr(states), dptr(out))
    return out


def oracle_fold_seq(sinter, contact, ham, states):
    states = c64(states)
    n = states.shape[0]
    out = np.zeros((n, 12))
    lib().or_fold_seq(dptr(c64(sinter)), dptr(c64(contact)), dptr(c64(ham)), n, dptr(states), dptr(out))
    return out


def oracle_facet_seq(verts, vf, contact, ham, dt_adv, states):
    states = c64(states)
    n = states.shape[0]
    out = np.zeros((n, 6))
    lib().or_facet_seq(dptr(c64(verts)), dptr(c64(vf)), dptr(c64(contact)), dptr(c64(ham)), dt_adv, n,
                       dptr(states), dptr(out))
    return out


def oracle_closest_point(verts, pts):
    pts = c64(pts)
    out = np.zeros_like(pts)
    lib().or_closest_point(dptr(c64(verts)), pts.shape[0], dptr(pts), dptr(out))
    return out


def oracle_topology(x0, r_part, d_crit):
    x0 = c64(x0)
    n = x0.shape[0]
    cap = 16 * n + 16
    pairs = np.zeros((cap, 2), dtype=np.uint32)
    sk = C.c_size_t(0)
    nb = lib().or_sinter_topology(n, dptr(x0), r_part, d_crit, u32ptr(pairs), cap, C.byref(sk))
    return pairs[:nb].copy(), sk.value


# -------------------------------------------------------- system driver -----
def legacy_frames(x0, r_part, d_crit, vecs):
    """sinter_functor's constructor (or_legacy_frames): bonds (i < j, row
    order), connectors (nb, 2, 5, 3) = i's five points then j's, and the
    number of unit vectors drawn from `vecs` (ValueError when they ran out)."""
    x0, vecs = c64(x0), c64(vecs).reshape(-1, 3)
    used = C.c_size_t(0)
    nb = lib().or_legacy_frames(len(x0), dptr(x0), r_part, d_crit, dptr(vecs), len(vecs), None, None, 0,
                                C.byref(used))
    pairs = np.zeros((max(nb, 1), 2), dtype=np.uint32)
    conn = np.zeros((max(nb, 1), 2, 5, 3))
    lib().or_legacy_frames(len(x0), dptr(x0), r_part, d_crit, dptr(vecs), len(vecs), u32ptr(pairs), dptr(conn), nb,
                           C.byref(used))
    if used.value > len(vecs):
        raise ValueError("legacy_frames: the unit vectors ran out")
    return pairs[:nb], conn[:nb], used.value


def _legacy_seq(fn, n, x0, th0, sinter13, contact18, vecs, incs, with_bonds):
    x0, th0, vecs, incs = c64(x0), c64(th0), c64(vecs).reshape(-1, 3), c64(incs)
    nsteps = incs.shape[0]
    out = np.zeros((nsteps, n, n, 6))
    used = C.c_size_t(0)
    args = [n, dptr(x0), dptr(th0), dptr(c64(sinter13)), dptr(c64(contact18)), dptr(vecs), len(vecs), nsteps,
            dptr(incs), dptr(out)]
    if with_bonds:
        nb = C.c_size_t(0)
        fn(*args, C.byref(nb), C.byref(used))
    else:
        fn(*args, C.byref(used))
    return out, used.value


def oracle_legacy_seq(n, x0, th0, sinter13, contact18, vecs, incs):
    """or_legacy_seq: incs is (nsteps, n, 12) = dx, dtheta, v, omega."""
    return _legacy_seq(lib().or_legacy_seq, n, x0, th0, sinter13, contact18, vecs, incs, True)


class OracleSystem:
    """The restated libgran system (functor fold + libtimestep contract)."""

    def __init__(self, params, x0, v0=None, th0=None, w0=None):
        n = x0.shape[0]
        z = np.zeros((n, 3))
        self.n = n
        self.params = params
        self.step_index = C.c_uint64(0)
        self._h = lib().or_sys_create(C.byref(params), n, dptr(c64(x0)), dptr(c64(v0 if v0 is not None else z)),
                                      dptr(c64(th0 if th0 is not None else z)),
                                      dptr(c64(w0 if w0 is not None else z)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_sys_destroy(self._h)
            self._h = None

    def track_dissipation(self, enable=True):
        """Start (zeroing it) or stop the cumulative dissipated-work ledger."""
        lib().or_sys_set_dissipation_tracking(self._h, 1 if enable else 0)

    def dissipated_work(self):
        out = np.zeros(5)
        lib().or_sys_dissipated_work(self._h, dptr(out))
        return dict(zip(("normal", "tangential", "rolling", "torsion", "total"), out))

    def set_neck_strengths(self, strengths):
        """alt_breaking_sinter neck strengths, one per bond in bonds() order."""
        lib().or_sys_set_neck_strengths(self._h, dptr(c64(strengths)))

    def set_legacy(self, pairs, conn):
        """sinter_functor bonds and connectors (the layout of legacy_frames)."""
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32)
        lib().or_sys_legacy_set(self._h, len(pairs), u32ptr(pairs), dptr(c64(conn)))
        self._nlegacy = len(pairs)

    def legacy_connectors(self):
        out = np.zeros((max(self._nlegacy, 1), 2, 5, 3))
        lib().or_sys_legacy_get(self._h, dptr(out))
        return out[:self._nlegacy]

    def set_facets(self, verts, vels):
        verts, vels = c64(verts), c64(vels)
        lib().or_sys_set_facets(self._h, verts.shape[0], dptr(verts), dptr(vels))

    def rebuild(self, brute=True):
        lib().or_sys_rebuild_neighbors(self._h, 1 if brute else 0)

    def compute_accelerations(self):
        lib().or_sys_compute_accelerations(self._h)

    def step(self, dt, nsteps=1, rebuild_every=0):
        lib().or_sys_step(self._h, dt, nsteps, rebuild_every, C.byref(self.step_index))

    def state(self):
        arrs = [np.zeros((self.n, 3)) for _ in range(6)]
        lib().or_sys_get_state(self._h, *[dptr(a) for a in arrs])
        return dict(zip(("x", "v", "theta", "omega", "a", "alpha"), arrs))

    def set_state(self, x=None, v=None, theta=None, omega=None):
        lib().or_sys_set_state(self._h, *[dptr(c64(a)) if a is not None else None for a in (x, v, theta, omega)])

    @property
    def t(self):
        return lib().or_sys_time(self._h)

    def neighbors(self):
        off = np.zeros(self.n + 1, dtype=np.uint64)
        tot = lib().or_sys_get_neighbors(self._h, off.ctypes.data_as(_u64p), u32ptr(np.zeros(1, np.uint32)), 0)
        idx = np.zeros(max(tot, 1), dtype=np.uint32)
        lib().or_sys_get_neighbors(self._h, off.ctypes.data_as(_u64p), u32ptr(idx), tot)
        return off, idx[:tot]

    def bonds(self):
        cap = 16 * self.n + 16
        pairs = np.zeros((cap, 2), dtype=np.uint32)
        sk = C.c_size_t(0)
        nb = lib().or_sys_get_bonds(self._h, u32ptr(pairs), cap, C.byref(sk))
        return pairs[:nb].copy(), sk.value

    def springs(self, which):
        tot = lib().or_sys_get_springs(self._h, which, u32ptr(np.zeros(2, np.uint32)), dptr(np.zeros(9)), 0)
        ij = np.zeros((max(tot, 1), 2), dtype=np.uint32)
        s = np.zeros((max(tot, 1), 9))
        lib().or_sys_get_springs(self._h, which, u32ptr(ij), dptr(s), tot)
        return ij[:tot], s[:tot]

    def pair_magnitudes(self):
        """(n, 2): sum_j |a_ij| and sum_j |alpha_ij| at the current state."""
        out = np.zeros((self.n, 2))
        lib().or_sys_pair_magnitudes(self._h, dptr(out))
        return out

    def functor_magnitudes(self):
        """(n, 2): sum over pairs and over each pair's functors of the functor's
        |a| and |alpha| (no cancellation inside a pair) at the current state."""
        out = np.zeros((self.n, 2))
        lib().or_sys_functor_magnitudes(self._h, dptr(out))
        return out

    def energy_terms(self, mass=MASS, inertia=INERTIA):
        """(n, 8) per-particle energy ledger terms (dem_system::energy_terms)."""
        out = np.zeros((self.n, 8))
        lib().or_sys_energy_terms(self._h, mass, inertia, dptr(out))
        return out

    def energy(self, mass=MASS, inertia=INERTIA):
        """The ledger's seven components summed in particle order, and their total."""
        t = self.energy_terms(mass, inertia)
        out = np.zeros(8)
        for i in range(self.n):
            out[:7] += t[i, :7]
        out[7] = sum(out[:7])
        return out

    def diagnostics(self, mass=MASS, inertia=INERTIA):
        out = np.zeros(9)
        lib().or_sys_diagnostics(self._h, mass, inertia, dptr(out))
        return dict(ke=out[0], ke_trans=out[1], ke_rot=out[2], P=out[3], L=out[4], rg=out[5], cm=out[6:9].copy())
