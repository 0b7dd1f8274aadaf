"""libgran_amd: the libgran DEM per-timestep path (bonded/sintered contact,
Luding friction with persistent springs, Hamaker attraction, neighbour-list
rebuild, rotational velocity Verlet) on AMD MI355X (gfx950).

The product is the C ABI in include/hipgran.h, implemented by hand-written HIP
kernels in libgran_amd/csrc and built into libgran_amd/libhipgran.so. The C++
drop-in headers (libgran_amd/include/libgran/...) and this Python module are
thin hosts over that ABI. Importing this module loads the library and fails
loudly when it is missing.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import (BINARY_CONTACT, BINARY_NONE, BINARY_SINTER, BINARY_SINTER_LEGACY, HAMAKER_FULL, HAMAKER_LIMITED, HAMAKER_NONE,
                   HAMAKER_SIMPLIFIED,
                   PAIRS_ALL, PAIRS_NEIGHBOR_LIST, PRECISION_FP64, PRECISION_MIXED, HipgranError, Params)

_abi.lib()  # fail at import time when the HIP library is absent

# hg_energy's terms, in its output order
ENERGY_TERMS = ("ke_trans", "ke_rot", "u_normal", "u_tangential", "u_rolling", "u_torsion", "u_hamaker", "total")

__all__ = ["System", "Pool", "Params", "make_params", "generate_aggregate", "sinter_topology", "HipgranError",
           "BINARY_NONE", "BINARY_CONTACT", "BINARY_SINTER", "BINARY_SINTER_LEGACY", "legacy_sinter_frames",
           "HAMAKER_NONE", "HAMAKER_FULL", "HAMAKER_LIMITED",
           "HAMAKER_SIMPLIFIED", "write_vtk_particles", "write_vtk_necks",
           "PAIRS_ALL", "PAIRS_NEIGHBOR_LIST", "PRECISION_FP64", "PRECISION_MIXED"]


def make_params(binary_model, hamaker_model, pair_mode=PAIRS_ALL, r_verlet=0.0, contact=None, sinter=None,
                hamaker=None, facet_contact=0, facet_hamaker=0, surface_contact=None, surface_hamaker=None,
                precision=PRECISION_FP64, neighbor_capacity=0, pair_capacity=0, sinter_breaking=0):
    """Builds hg_params; functor parameter sequences are in the reference ctor
    order (contact_force.h:10-30, alt_sinter_bridge.h:19-36 + d_crit,
    limited_range_hamaker_force.h:20-27 as A, h0, r_cutoff, r_part, mass).
    sinter_breaking=1 with BINARY_SINTER selects alt_breaking_sinter."""
    p = Params()
    p.sinter_breaking = sinter_breaking
    p.binary_model, p.hamaker_model, p.pair_mode, p.precision = binary_model, hamaker_model, pair_mode, precision
    p.facet_contact, p.facet_hamaker = facet_contact, facet_hamaker
    p.neighbor_capacity, p.pair_capacity, p.r_verlet = neighbor_capacity, pair_capacity, r_verlet
    for name, vals in (("contact", contact), ("sinter", sinter), ("hamaker", hamaker),
                       ("surface_contact", surface_contact), ("surface_hamaker", surface_hamaker)):
        if vals is not None:
            _abi.fill(getattr(p, name), vals)
    return p


def generate_aggregate(n, r_part, df, kf, seed):
    """Seeded TD-CCA aggregate, (n, 3) positions in metres (host only)."""
    x = np.zeros((n, 3))
    _abi.check(_abi.lib().hg_generate_aggregate(n, r_part, df, kf, seed, _abi.dptr(x)))
    return x


def generate_aggregates(sizes, dfs, kf, seeds, r_part, threads=0):
    """Many seeded TD-CCA aggregates on host threads; returns a list of (n_k, 3)
    arrays, each identical to generate_aggregate(n_k, r_part, df_k, kf, seed_k)."""
    sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
    dfs = np.ascontiguousarray(dfs, dtype=np.float64)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    x = np.zeros((int(sizes.sum()), 3))
    _abi.check(_abi.lib().hg_generate_aggregates(len(sizes), _abi.u32ptr(sizes), _abi.dptr(dfs), kf,
                                                 seeds.ctypes.data_as(_abi._u64p), r_part, _abi.dptr(x), threads))
    return np.split(x, np.cumsum(sizes)[:-1])


_UNIT_VECTOR_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double))


def legacy_sinter_frames(x0, r_part, critical_separation, next_unit_vector):
    """The legacy sinter_functor's bonds and connectors (hg_legacy_sinter_frames):
    next_unit_vector() returns the reference constructor's random unit
    vectors, drawn in its order. Returns (pairs (nb, 2), connectors
    (nb, 2, 5, 3)): per bond i's five points, then j's."""
    x0 = _abi.c64(x0)
    nb = C.c_size_t(0)
    _abi.check(_abi.lib().hg_legacy_sinter_frames(x0.shape[0], _abi.dptr(x0), r_part, critical_separation, None, None,
                                                  None, None, 0, C.byref(nb)))
    pairs = np.zeros((max(nb.value, 1), 2), dtype=np.uint32)
    conn = np.zeros((max(nb.value, 1), 2, 5, 3))
    errors = []

    def trampoline(_user, out):
        if errors:  # the source failed: end every further draw at once (x or y
            # axis in turn, never both collinear with one normal); the frames
            # are discarded and the error raised after the call
            k = len(errors) % 2
            errors.append(None)
            out[0], out[1], out[2] = float(k == 0), float(k == 1), 0.0
            return
        try:
            v = next_unit_vector()
            out[0], out[1], out[2] = float(v[0]), float(v[1]), float(v[2])
        except Exception as e:  # noqa: BLE001  (reported after the call)
            errors.append(e)
            out[0], out[1], out[2] = 1.0, 0.0, 0.0

    cb = _UNIT_VECTOR_FN(trampoline)
    st = _abi.lib().hg_legacy_sinter_frames(x0.shape[0], _abi.dptr(x0), r_part, critical_separation,
                                            C.cast(cb, C.c_void_p), None, _abi.u32ptr(pairs), _abi.dptr(conn),
                                            nb.value, C.byref(nb))
    if errors:
        raise errors[0]
    _abi.check(st)
    return pairs[:nb.value], conn[:nb.value]


def write_vtk_particles(path, x, theta, r_part):
    """Legacy-VTK particle file (hg_write_vtk_particles): positions / r_part
    and the body axes of orientation theta, as writer.cpp:51-118 writes it."""
    x, theta = _abi.c64(x), _abi.c64(theta)
    if x.shape != theta.shape:
        raise ValueError("x and theta must have the same shape")
    _abi.check(_abi.lib().hg_write_vtk_particles(os.fsencode(path), x.shape[0], _abi.dptr(x), _abi.dptr(theta),
                                                 r_part))


def write_vtk_necks(path, x, pairs, r_part):
    """Legacy-VTK neck file (hg_write_vtk_necks): midpoints / r_part and
    normals of the bonded pairs, in the given order (writer.cpp:120-172)."""
    x = _abi.c64(x)
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    _abi.check(_abi.lib().hg_write_vtk_necks(os.fsencode(path), x.shape[0], _abi.dptr(x), pairs.shape[0],
                                             _abi.u32ptr(pairs), r_part))


def sinter_topology(x0, r_part, d_crit):
    """Bonds (i<j, reference insertion order) and skipped-cycle count (host only)."""
    x0 = _abi.c64(x0)
    n = x0.shape[0]
    nb, sk = C.c_size_t(0), C.c_size_t(0)
    _abi.check(_abi.lib().hg_sinter_topology(n, _abi.dptr(x0), r_part, d_crit, None, 0, C.byref(nb), C.byref(sk)))
    pairs = np.zeros((max(nb.value, 1), 2), dtype=np.uint32)
    _abi.check(_abi.lib().hg_sinter_topology(n, _abi.dptr(x0), r_part, d_crit, _abi.u32ptr(pairs), nb.value,
                                             C.byref(nb), C.byref(sk)))
    return pairs[:nb.value], sk.value


def breaking_sinter_topology(x0, r_part, d_crit):
    """alt_breaking_sinter bonds: every pair i<j within d_crit (host only)."""
    x0 = _abi.c64(x0)
    n = x0.shape[0]
    nb = C.c_size_t(0)
    _abi.check(_abi.lib().hg_breaking_sinter_topology(n, _abi.dptr(x0), r_part, d_crit, None, 0, C.byref(nb)))
    pairs = np.zeros((max(nb.value, 1), 2), dtype=np.uint32)
    _abi.check(_abi.lib().hg_breaking_sinter_topology(n, _abi.dptr(x0), r_part, d_crit, _abi.u32ptr(pairs),
                                                      nb.value, C.byref(nb)))
    return pairs[:nb.value]


class System:
    """One DEM system on the current GPU (hg_system)."""

    def __init__(self, params, x0, v0=None, theta0=None, omega0=None):
        L = _abi.lib()
        x0 = _abi.c64(x0)
        self.n = x0.shape[0]
        self.params = params
        self._h = C.c_void_p()
        _abi.check(L.hg_create(C.byref(params), self.n, _abi.dptr(x0), _abi.dptr(_abi.c64(v0)),
                               _abi.dptr(_abi.c64(theta0)), _abi.dptr(_abi.c64(omega0)), C.byref(self._h)))

    @classmethod
    def load(cls, snapshot):
        """A system rebuilt from save() (or Pool.save()) bytes (hg_load)."""
        self = cls.__new__(cls)
        self._h = C.c_void_p()
        buf = bytes(snapshot)
        _abi.check(_abi.lib().hg_load(buf, len(buf), C.byref(self._h)))
        self.params = None
        self.n = int(np.frombuffer(buf, dtype=np.uint64, count=1, offset=16 + C.sizeof(Params))[0])
        return self

    def save(self):
        """The system's snapshot as bytes (hg_save)."""
        size = C.c_size_t(0)
        _abi.check(_abi.lib().hg_save(self._h, None, 0, C.byref(size)))
        buf = C.create_string_buffer(size.value)
        _abi.check(_abi.lib().hg_save(self._h, buf, size.value, C.byref(size)))
        return buf.raw[:size.value]

    def close(self):
        # at interpreter exit the module globals may already be gone
        if getattr(self, "_h", None) and self._h.value and _abi is not None and C is not None:
            _abi.lib().hg_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close

    def set_facets(self, verts, v_facet):
        verts, v_facet = _abi.c64(verts), _abi.c64(v_facet)
        _abi.check(_abi.lib().hg_set_facets(self._h, verts.shape[0], _abi.dptr(verts), _abi.dptr(v_facet)))

    def set_neck_strengths(self, strengths):
        """alt_breaking_sinter neck strengths, one per bond in bonds() order."""
        _abi.check(_abi.lib().hg_set_neck_strengths(self._h, _abi.dptr(_abi.c64(strengths))))

    def facets(self, nf):
        verts, vel = np.zeros((nf, 3, 3)), np.zeros((nf, 3))
        _abi.check(_abi.lib().hg_get_facets(self._h, _abi.dptr(verts), _abi.dptr(vel)))
        return verts, vel

    def rebuild_neighbors(self):
        _abi.check(_abi.lib().hg_rebuild_neighbors(self._h))

    def compute_accelerations(self):
        _abi.check(_abi.lib().hg_compute_accelerations(self._h))

    def step(self, dt, nsteps=1, rebuild_every=0):
        _abi.check(_abi.lib().hg_step(self._h, dt, nsteps, rebuild_every))

    def step_async(self, dt, nsteps=1, rebuild_every=0):
        """hg_step_async: queue the steps and return; errors surface at the next read."""
        _abi.check(_abi.lib().hg_step_async(self._h, dt, nsteps, rebuild_every))

    def state(self):
        arrs = [np.zeros((self.n, 3)) for _ in range(4)]
        _abi.check(_abi.lib().hg_get_state(self._h, *[_abi.dptr(a) for a in arrs]))
        out = dict(zip(("x", "v", "theta", "omega"), arrs))
        a, al = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        _abi.check(_abi.lib().hg_get_accelerations(self._h, _abi.dptr(a), _abi.dptr(al)))
        out["a"], out["alpha"] = a, al
        return out

    def set_state(self, x=None, v=None, theta=None, omega=None):
        arrs = [_abi.c64(q) for q in (x, v, theta, omega)]
        _abi.check(_abi.lib().hg_set_state(self._h, *[_abi.dptr(q) for q in arrs]))

    def time(self):
        t, k = C.c_double(0), C.c_uint64(0)
        _abi.check(_abi.lib().hg_get_time(self._h, C.byref(t), C.byref(k)))
        return t.value, k.value

    def neighbors(self):
        off = np.zeros(self.n + 1, dtype=np.uint64)
        tot = C.c_size_t(0)
        _abi.check(_abi.lib().hg_get_neighbors(self._h, off.ctypes.data_as(_abi._u64p), None, 0, C.byref(tot)))
        idx = np.zeros(max(tot.value, 1), dtype=np.uint32)
        _abi.check(_abi.lib().hg_get_neighbors(self._h, off.ctypes.data_as(_abi._u64p), _abi.u32ptr(idx), tot.value,
                                               C.byref(tot)))
        return off, idx[:tot.value]

    def bonds(self):
        nb, sk = C.c_size_t(0), C.c_size_t(0)
        _abi.check(_abi.lib().hg_get_bonds(self._h, None, 0, C.byref(nb), C.byref(sk)))
        pairs = np.zeros((max(nb.value, 1), 2), dtype=np.uint32)
        _abi.check(_abi.lib().hg_get_bonds(self._h, _abi.u32ptr(pairs), nb.value, C.byref(nb), C.byref(sk)))
        return pairs[:nb.value], sk.value

    def springs(self, which):
        tot = C.c_size_t(0)
        _abi.check(_abi.lib().hg_get_springs(self._h, which, None, None, 0, C.byref(tot)))
        ij = np.zeros((max(tot.value, 1), 2), dtype=np.uint32)
        s = np.zeros((max(tot.value, 1), 9))
        _abi.check(_abi.lib().hg_get_springs(self._h, which, _abi.u32ptr(ij), _abi.dptr(s), tot.value, C.byref(tot)))
        return ij[:tot.value], s[:tot.value]

    def set_legacy_connectors(self, pairs, connectors):
        """Connectors of an HG_BINARY_SINTER_LEGACY system (hg_set_legacy_connectors),
        in the layout legacy_sinter_frames returns."""
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        _abi.check(_abi.lib().hg_set_legacy_connectors(self._h, pairs.shape[0], _abi.u32ptr(pairs),
                                                       _abi.dptr(_abi.c64(connectors))))

    def legacy_connectors(self):
        """The connectors now, (nb, 2, 5, 3) in bonds() order."""
        nb = C.c_size_t(0)
        _abi.check(_abi.lib().hg_get_legacy_connectors(self._h, None, 0, C.byref(nb)))
        out = np.zeros((max(nb.value, 1), 2, 5, 3))
        _abi.check(_abi.lib().hg_get_legacy_connectors(self._h, _abi.dptr(out), nb.value, C.byref(nb)))
        return out[:nb.value]

    def diagnostics(self, mass, inertia):
        out = np.zeros(9)
        _abi.check(_abi.lib().hg_diagnostics(self._h, mass, inertia, _abi.dptr(out)))
        return dict(ke=out[0], ke_trans=out[1], ke_rot=out[2], P=out[3], L=out[4], rg=out[5], cm=out[6:9].copy())

    def energy(self, mass, inertia):
        """FP64 energy ledger (hg_energy): kinetic, pair elastic, spring and
        Hamaker terms and their total, in joules."""
        out = np.zeros(8)
        _abi.check(_abi.lib().hg_energy(self._h, mass, inertia, _abi.dptr(out)))
        return dict(zip(ENERGY_TERMS, out))

    def run(self, dt, nsteps, mass, inertia, thermo_period, thermo_path="energies.dat", rebuild_every=0,
            dump_period=0, dump_dir=None, r_part=None):
        """The driver loop of main.cpp:142-161: before the steps n with
        n % thermo_period == 0, one line of t, E_trs, E_rot, E_tot, |P|, |L|
        (tab-separated, six significant digits as a default C++ stream writes
        them) into `thermo_path`, under the header "t E_trs E_rot E_tot P L";
        with dump_period, a VTK particle file dump_dir/particles_<k>.vtk before
        the steps n with n % dump_period == 0 (writer.cpp:51-54). The steps in
        between run as one hg_step call each, so a thermo period costs one
        device reduction, not a host loop. t is n * dt, as the reference
        prints it. Returns the number of thermo lines."""
        marks = set()
        if thermo_period:
            marks.update(range(0, nsteps, thermo_period))
        if dump_period:
            if dump_dir is None or r_part is None:
                raise ValueError("dump_period needs dump_dir and r_part")
            marks.update(range(0, nsteps, dump_period))
        marks = sorted(marks) + [nsteps]
        lines = dumps = 0
        with open(thermo_path, "w") as ofs:
            ofs.write("t\tE_trs\tE_rot\tE_tot\tP\tL\n")
            for a, b in zip(marks[:-1], marks[1:]):
                if dump_period and a % dump_period == 0:
                    st = self.state()
                    write_vtk_particles(os.path.join(dump_dir, f"particles_{dumps}.vtk"), st["x"], st["theta"],
                                        r_part)
                    dumps += 1
                if thermo_period and a % thermo_period == 0:
                    d = self.diagnostics(mass, inertia)
                    ofs.write("\t".join(f"{q:g}" for q in (float(a) * dt, d["ke_trans"], d["ke_rot"], d["ke"],
                                                           d["P"], d["L"])) + "\n")
                    lines += 1
                self.step(dt, b - a, rebuild_every)
        return lines

    def profile(self, enable=True):
        _abi.check(_abi.lib().hg_profile_enable(self._h, 1 if enable else 0))

    def track_dissipation(self, enable=True):
        """Start (zeroing it) or stop the cumulative dissipated-work ledger."""
        _abi.check(_abi.lib().hg_set_dissipation_tracking(self._h, 1 if enable else 0))

    def dissipated_work(self):
        """Dissipated work since tracking started (J): normal, tangential,
        rolling, torsion and total."""
        out = np.zeros(5)
        _abi.check(_abi.lib().hg_get_dissipated_work(self._h, _abi.dptr(out)))
        return dict(zip(("normal", "tangential", "rolling", "torsion", "total"), out))

    def per_step_launches(self, enable=True):
        """Per-step kernels for a small FP64 system instead of its one
        launch per rebuild interval (same results; a timing baseline)."""
        _abi.check(_abi.lib().hg_set_per_step_launches(self._h, 1 if enable else 0))

    def profile_read(self):
        launches = np.zeros(4, dtype=np.uint64)
        ms = np.zeros(4)
        _abi.check(_abi.lib().hg_profile_read(self._h, launches.ctypes.data_as(_abi._u64p), _abi.dptr(ms)))
        return launches, ms


class Pool:
    """Many independent systems stepped together (hg_pool)."""

    def __init__(self, params, max_particles, max_systems, max_facets=0):
        self.params = params
        self._h = C.c_void_p()
        _abi.check(_abi.lib().hg_pool_create(C.byref(params), max_particles, max_systems, max_facets,
                                             C.byref(self._h)))

    def close(self):
        # at interpreter exit the module globals may already be gone
        if getattr(self, "_h", None) and self._h.value and _abi is not None and C is not None:
            _abi.lib().hg_pool_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close

    @staticmethod
    def bytes_per_particle(params, max_facets=0):
        return int(_abi.lib().hg_pool_bytes_per_particle(C.byref(params), max_facets))

    def add(self, x0, v0=None, theta0=None, omega0=None):
        x0 = _abi.c64(x0)
        slot = C.c_int32(-1)
        _abi.check(_abi.lib().hg_pool_add(self._h, x0.shape[0], _abi.dptr(x0), _abi.dptr(_abi.c64(v0)),
                                          _abi.dptr(_abi.c64(theta0)), _abi.dptr(_abi.c64(omega0)), C.byref(slot)))
        return slot.value

    def try_add(self, x0, v0=None, theta0=None, omega0=None):
        """add(), or None when the arena or the slot table is full."""
        try:
            return self.add(x0, v0, theta0, omega0)
        except HipgranError as e:
            if e.status == _abi.HG_ERR_CAPACITY:
                return None
            raise

    def remove(self, slot):
        _abi.check(_abi.lib().hg_pool_remove(self._h, slot))

    def set_neck_strengths(self, slot, strengths):
        """alt_breaking_sinter neck strengths of one system, in bonds(slot) order."""
        _abi.check(_abi.lib().hg_pool_set_neck_strengths(self._h, slot, _abi.dptr(_abi.c64(strengths))))

    def set_facets(self, slot, verts, v_facet):
        verts, v_facet = _abi.c64(verts), _abi.c64(v_facet)
        _abi.check(_abi.lib().hg_pool_set_facets(self._h, slot, verts.shape[0], _abi.dptr(verts), _abi.dptr(v_facet)))

    def step(self, dt, nsteps, rebuild_every=0):
        _abi.check(_abi.lib().hg_pool_step(self._h, dt, nsteps, rebuild_every))

    def info(self, slot):
        n, t, k = C.c_size_t(0), C.c_double(0), C.c_uint64(0)
        _abi.check(_abi.lib().hg_pool_info(self._h, slot, C.byref(n), C.byref(t), C.byref(k)))
        return n.value, t.value, k.value

    def energy(self, slot, mass, inertia):
        """hg_energy for one system of the pool."""
        out = np.zeros(8)
        _abi.check(_abi.lib().hg_pool_energy(self._h, slot, mass, inertia, _abi.dptr(out)))
        return dict(zip(ENERGY_TERMS, out))

    def state(self, slot):
        n = self.info(slot)[0]
        arrs = [np.zeros((n, 3)) for _ in range(6)]
        _abi.check(_abi.lib().hg_pool_get_state(self._h, slot, *[_abi.dptr(a) for a in arrs[:4]]))
        _abi.check(_abi.lib().hg_pool_get_accelerations(self._h, slot, _abi.dptr(arrs[4]), _abi.dptr(arrs[5])))
        return dict(zip(("x", "v", "theta", "omega", "a", "alpha"), arrs))

    def neighbors(self, slot):
        n = self.info(slot)[0]
        off = np.zeros(n + 1, dtype=np.uint64)
        tot = C.c_size_t(0)
        _abi.check(_abi.lib().hg_pool_get_neighbors(self._h, slot, off.ctypes.data_as(_abi._u64p), None, 0,
                                                    C.byref(tot)))
        idx = np.zeros(max(tot.value, 1), dtype=np.uint32)
        _abi.check(_abi.lib().hg_pool_get_neighbors(self._h, slot, off.ctypes.data_as(_abi._u64p), _abi.u32ptr(idx),
                                                    tot.value, C.byref(tot)))
        return off, idx[:tot.value]

    def springs(self, slot, which):
        tot = C.c_size_t(0)
        _abi.check(_abi.lib().hg_pool_get_springs(self._h, slot, which, None, None, 0, C.byref(tot)))
        ij = np.zeros((max(tot.value, 1), 2), dtype=np.uint32)
        s = np.zeros((max(tot.value, 1), 9))
        _abi.check(_abi.lib().hg_pool_get_springs(self._h, slot, which, _abi.u32ptr(ij), _abi.dptr(s), tot.value,
                                                  C.byref(tot)))
        return ij[:tot.value], s[:tot.value]

    def bonds(self, slot):
        nb, sk = C.c_size_t(0), C.c_size_t(0)
        _abi.check(_abi.lib().hg_pool_get_bonds(self._h, slot, None, 0, C.byref(nb), C.byref(sk)))
        pairs = np.zeros((max(nb.value, 1), 2), dtype=np.uint32)
        _abi.check(_abi.lib().hg_pool_get_bonds(self._h, slot, _abi.u32ptr(pairs), nb.value, C.byref(nb),
                                                C.byref(sk)))
        return pairs[:nb.value], sk.value

    def last_step_profile(self):
        c, r, k, ks = C.c_double(0), C.c_double(0), C.c_uint64(0), C.c_uint64(0)
        _abi.check(_abi.lib().hg_pool_last_step_profile(self._h, C.byref(c), C.byref(r), C.byref(k), C.byref(ks)))
        w, wl = C.c_double(0), C.c_uint64(0)
        _abi.check(_abi.lib().hg_pool_last_wide_profile(self._h, C.byref(w), C.byref(wl)))
        b = C.c_double(0)
        _abi.check(_abi.lib().hg_pool_last_busy(self._h, C.byref(b)))
        return dict(chunk_ms=c.value, rebuild_ms=r.value, chunks=k.value, chunk_steps=ks.value, wide_ms=w.value,
                    wide_launches=wl.value, busy_ms=b.value)

    def save(self, slot):
        """One system's snapshot as bytes (hg_pool_save)."""
        size = C.c_size_t(0)
        _abi.check(_abi.lib().hg_pool_save(self._h, slot, None, 0, C.byref(size)))
        buf = C.create_string_buffer(size.value)
        _abi.check(_abi.lib().hg_pool_save(self._h, slot, buf, size.value, C.byref(size)))
        return buf.raw[:size.value]

    def load(self, snapshot):
        """Adds a system from save() bytes (hg_pool_load); returns its slot."""
        buf = bytes(snapshot)
        slot = C.c_int32(-1)
        _abi.check(_abi.lib().hg_pool_load(self._h, buf, len(buf), C.byref(slot)))
        return slot.value

    def fault(self, slot):
        """HG_FAULT_* bits of a system that stopped (0 while it runs)."""
        f = C.c_int(0)
        _abi.check(_abi.lib().hg_pool_fault(self._h, slot, C.byref(f)))
        return f.value

    def set_wide_threshold(self, wide_min):
        """Systems of at least wide_min particles (0: none) use the wide launches."""
        _abi.check(_abi.lib().hg_pool_set_wide_threshold(self._h, int(wide_min)))

    def set_wide_cus(self, wide_cus):
        """CUs given to the wide launches while both kinds of systems run."""
        _abi.check(_abi.lib().hg_pool_set_wide_cus(self._h, int(wide_cus)))

    def set_chunk_shape(self, threads, wg_per_cu):
        """Chunk-kernel threads per workgroup (256 or 512) and workgroups per CU
        (each with that share of the LDS)."""
        _abi.check(_abi.lib().hg_pool_set_chunk_shape(self._h, int(threads), int(wg_per_cu)))

    def set_wide_graph(self, enable):
        """Replay each interval's wide launches as a HIP graph (default) or not."""
        _abi.check(_abi.lib().hg_pool_set_wide_graph(self._h, 1 if enable else 0))

    def set_packing(self, enable):
        """Step small systems side by side in one chunk workgroup (default) or
        give each its own workgroup; results are the same."""
        _abi.check(_abi.lib().hg_pool_set_packing(self._h, 1 if enable else 0))

    def autotune(self, enable=True):
        """Tune the wide threshold and room on the next live intervals."""
        _abi.check(_abi.lib().hg_pool_autotune(self._h, 1 if enable else 0))

    def tuning(self):
        wm, wc, done = C.c_size_t(0), C.c_int(0), C.c_int(0)
        _abi.check(_abi.lib().hg_pool_get_tuning(self._h, C.byref(wm), C.byref(wc), C.byref(done)))
        tpb, wgs = C.c_int(0), C.c_int(0)
        _abi.check(_abi.lib().hg_pool_get_chunk_shape(self._h, C.byref(tpb), C.byref(wgs)))
        return dict(wide_min=wm.value, wide_cus=wc.value, tuned=bool(done.value), chunk_threads=tpb.value,
                    chunk_wg_per_cu=wgs.value)


class Server:
    """Continuous-batching request server (hg_server): requests queue in
    arrival order, are admitted while they fit the device budget, run in one
    shared pool and are retired with their final state on the host."""

    def __init__(self, params, budget_bytes, max_requests, max_facets=0, quantum=1000, rebuild_every=50,
                 dt=1e-13):
        self.params = params
        cfg = _abi.ServerConfig(int(budget_bytes), int(max_requests), int(max_facets), int(quantum),
                                int(rebuild_every), float(dt))
        self._h = C.c_void_p()
        _abi.check(_abi.lib().hg_server_create(C.byref(params), C.byref(cfg), C.byref(self._h)))
        self._sizes = {}

    def close(self):
        # at interpreter exit the module globals may already be gone
        if getattr(self, "_h", None) and self._h.value and _abi is not None and C is not None:
            _abi.lib().hg_server_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close

    def submit(self, x0, nsteps, v0=None, theta0=None, omega0=None, facets=None):
        """Queues one request; facets = (verts (nf, 3, 3), v_facet (nf, 3)) or None. Returns its id."""
        x0 = _abi.c64(x0)
        nf, verts, vel = 0, None, None
        if facets is not None:
            verts, vel = _abi.c64(facets[0]), _abi.c64(facets[1])
            nf = verts.shape[0]
        rid = C.c_uint64(0)
        _abi.check(_abi.lib().hg_server_submit(self._h, x0.shape[0], _abi.dptr(x0), _abi.dptr(_abi.c64(v0)),
                                               _abi.dptr(_abi.c64(theta0)), _abi.dptr(_abi.c64(omega0)), nf,
                                               _abi.dptr(verts), _abi.dptr(vel), int(nsteps), C.byref(rid)))
        self._sizes[rid.value] = x0.shape[0]
        return rid.value

    def status(self, rid):
        st = C.c_int(0)
        _abi.check(_abi.lib().hg_server_status(self._h, rid, C.byref(st)))
        return st.value

    def wait(self, rid, timeout=-1.0):
        """Final state and times (submitted, admitted, done; seconds since the
        server started) of a request; releases it."""
        n = self._sizes[rid]
        arrs = [np.zeros((n, 3)) for _ in range(4)]
        times = np.zeros(3)
        st = _abi.lib().hg_server_wait(self._h, rid, float(timeout), *[_abi.dptr(a) for a in arrs],
                                       _abi.dptr(times))
        if st != _abi.HG_ERR_CAPACITY or b"timed out" not in _abi.lib().hg_last_error():
            del self._sizes[rid]  # finished (done or failed): the server has released it
        _abi.check(st)
        out = dict(zip(("x", "v", "theta", "omega"), arrs))
        out["times"] = times
        return out

    def stats(self):
        s = _abi.ServerStats()
        _abi.check(_abi.lib().hg_server_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}
