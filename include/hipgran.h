/*
 * hipgran — C ABI of the MI355X (gfx950) DEM per-timestep path.
 *
 * The reference (egor-demidov/libgran) is a header-only C++ template library
 * with no binary interface; its drop-in surface is the granular_system family
 * of templates. Each entry point below replaces one operation that the
 * reference performs inside those templates or inside its time-stepping
 * dependency; the citation says which. The C++ drop-in headers under
 * libgran_amd/include/libgran/ call these functions, and so can any FFI
 * (ctypes, cgo, JNI): plain pointers and sizes only, no exceptions.
 *
 * Conventions
 *   - Vectors cross the ABI as AoS triples: double[3*n] (x0,y0,z0,x1,...).
 *   - Every function returns hg_status; on failure hg_last_error() returns a
 *     thread-local message for the last failing call on this thread.
 *   - One hg_system is not internally synchronised: call it from one thread at
 *     a time. Distinct systems may be used concurrently; each owns its own
 *     HIP stream on the device that was current at hg_create.
 */
#ifndef HIPGRAN_H
#define HIPGRAN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hg_status {
    HG_OK = 0,
    HG_ERR_INVALID = 1,     /* bad argument or unsupported combination */
    HG_ERR_OOM = 2,         /* device or host allocation failed */
    HG_ERR_HIP = 3,         /* a HIP runtime call failed */
    HG_ERR_NUMERIC = 4,     /* NaN/Inf reached the particle state */
    HG_ERR_CAPACITY = 5,    /* a neighbour or pair-state table overflowed */
    HG_ERR_NO_DEVICE = 6,   /* no gfx950 device visible */
    HG_ERR_IO = 7           /* a file could not be created or written */
} hg_status;

/* contact_force_functor ctor arguments, same order and meaning
 * (include/libgran/contact_force/contact_force.h:10-30, n_part omitted).
 * Also used for surface_contact_force_functor (surface_contact_force.h:10-30). */
typedef struct hg_contact_params {
    double k, gamma_n, k_t, gamma_t, mu_s, phi_d, k_r, gamma_r, mu_r, phi_r,
           k_o, gamma_o, mu_o, phi_o, r_part, mass, inertia, dt;
} hg_contact_params;

/* alt_sinter_functor ctor arguments (alt_sinter_bridge.h:19-36): bonded
 * stiffness/damping, then critical_separation as d_crit. */
typedef struct hg_sinter_params {
    double k, gamma_n, k_t, gamma_t, k_r, gamma_r, k_o, gamma_o,
           r_part, mass, inertia, dt, d_crit;
} hg_sinter_params;

/* hamaker_functor (hamaker_force.h:20-26) / limited_range_hamaker_functor
 * (limited_range_hamaker_force.h:20-27) / simplified_hamaker_functor
 * (simplified_hamaker_force.h:10-16) / surface_hamaker_functor
 * (surface_hamaker_force.h:12-18). r_cutoff is read only in LIMITED mode. */
typedef struct hg_hamaker_params {
    double A, h0, r_cutoff, r_part, mass;
} hg_hamaker_params;

/* HG_BINARY_SINTER_LEGACY: the legacy sinter_functor with its
 * sinter_step_handler (sinter_bridge.h:13-272), FP64 single systems only.
 * Every pair i < j with |(|x_j - x_i| - 2r)| <= critical_separation is
 * bonded (no cycle check) and held by five connector springs whose ends move
 * with the particles (hg_legacy_sinter_frames, hg_set_legacy_connectors);
 * other pairs take the embedded contact functor. Parameters: sinter.k = k_d,
 * sinter.gamma_n = gamma_d, sinter.d_crit = critical_separation; the contact
 * functor, r_part, mass and inertia from `contact`; other sinter fields are
 * not read. */
enum { HG_BINARY_NONE = 0, HG_BINARY_CONTACT = 1, HG_BINARY_SINTER = 2, HG_BINARY_SINTER_LEGACY = 3 };
/* HG_HAMAKER_SIMPLIFIED: simplified_hamaker_functor
 * (simplified_hamaker_force.h:18-30), 2 A r / h^2 along x_j - x_i. */
enum { HG_HAMAKER_NONE = 0, HG_HAMAKER_FULL = 1, HG_HAMAKER_LIMITED = 2, HG_HAMAKER_SIMPLIFIED = 3 };
/* HG_PAIRS_ALL: every ordered pair j != i, as granular_system
 *   (granular_system.h:25-86).
 * HG_PAIRS_NEIGHBOR_LIST: listed pairs only, as granular_system_neighbor_list
 *   (granular_system_neighbor_list.h:26-87) with radius r_verlet. */
enum { HG_PAIRS_ALL = 0, HG_PAIRS_NEIGHBOR_LIST = 1 };
/* FP64: every operation in double, in the reference's arithmetic order.
 * MIXED: distances, overlaps and gaps in double; force/torque arithmetic in
 * float; positions, orientations, velocities and spring vectors kept and
 * integrated in double. */
enum { HG_PRECISION_FP64 = 0, HG_PRECISION_MIXED = 1 };

typedef struct hg_params {
    int32_t binary_model;        /* HG_BINARY_*: first functor of the binary container */
    int32_t hamaker_model;       /* HG_HAMAKER_*: second functor of the binary container */
    int32_t pair_mode;           /* HG_PAIRS_* */
    int32_t precision;           /* HG_PRECISION_* */
    int32_t facet_contact;       /* facets carry a surface_contact_force_functor */
    int32_t facet_hamaker;       /* facets carry a surface_hamaker_functor */
    uint32_t neighbor_capacity;  /* max listed neighbours per particle; 0 = derived from r_verlet */
    uint32_t pair_capacity;      /* max bonded + touching pairs per particle; 0 = default */
    int32_t facet_advance;       /* 0: facets advance by v_facet*dt after every step;
                                    1: only through hg_advance_facet (the driver moves them,
                                    as triangular_facet::update_positions does) */
    int32_t sinter_breaking;     /* with HG_BINARY_SINTER: alt_breaking_sinter_functor instead of
                                    alt_sinter_functor (alt_breaking_sinter_bridge.h:15-166): every
                                    pair within d_crit is bonded (no cycle check), and an ordered
                                    bonded pair (i, j) becomes a contact pair once |f_n n + f_t|
                                    exceeds its bond's neck strength (hg_set_neck_strengths;
                                    infinite until set) */
    double r_verlet;             /* neighbour-list radius (granular_system_neighbor_list.h:54) */
    hg_contact_params contact;           /* contact functor (also the one alt_sinter embeds) */
    hg_sinter_params sinter;             /* alt_sinter bonded model + d_crit */
    hg_hamaker_params hamaker;           /* binary hamaker functor */
    hg_contact_params surface_contact;   /* facet surface contact */
    hg_hamaker_params surface_hamaker;   /* facet surface hamaker */
} hg_params;

typedef struct hg_system hg_system;

/* Thread-local message of the last failing call on this thread. */
const char *hg_last_error(void);

/* Creates a system on the current HIP device and uploads the initial state.
 * Replaces: the granular_system / granular_system_neighbor_list ctors
 * (granular_system.h:53-63, granular_system_neighbor_list.h:54-64) together
 * with the alt_sinter_functor ctor's bond search (alt_sinter_bridge.h:19-90),
 * which runs here on x0 when binary_model == HG_BINARY_SINTER (and the legacy
 * sinter_functor's, sinter_bridge.h:146-161, for HG_BINARY_SINTER_LEGACY). In
 * neighbour-list mode the list is built once here. theta0/omega0/v0 may be
 * NULL (zeros). */
hg_status hg_create(const hg_params *params, size_t n, const double *x0, const double *v0,
                    const double *theta0, const double *omega0, hg_system **out);

/* Frees all device and host memory of the system. */
hg_status hg_destroy(hg_system *sys);

/* Installs nf triangular facets (unary container of triangular_facet,
 * triangular_facet.h:18-30): verts = nf*9 doubles (3 vertices each),
 * v_facet = nf*3 doubles. Replaces earlier facets and clears their springs. */
hg_status hg_set_facets(hg_system *sys, size_t nf, const double *verts, const double *v_facet);
/* Reads the facets' current vertices (nf*9) and velocities (nf*3). */
hg_status hg_get_facets(hg_system *sys, double *verts, double *v_facet);
/* Sets facet f's velocity to v_facet[3] and moves its vertices by
 * v_facet*dt. Replaces: triangular_facet::update_positions(dt) with the
 * driver's current v_facet (triangular_facet.h:62-69). */
hg_status hg_advance_facet(hg_system *sys, size_t f, const double *v_facet, double dt);

/* Rebuilds the neighbour list from the current positions.
 * Replaces: update_neighbor_list() of rotational_binary_system_neighbors_omp,
 * called by the driver (test/hamaker_neighbor.cpp:121-122). */
hg_status hg_rebuild_neighbors(hg_system *sys);

/* One acceleration pass at the current state: every pair through the functor
 * fold, then the facets. Mutates spring state exactly as the reference's pass
 * does. Replaces: the acceleration pass of rotational_binary_system_omp
 * calling granular_system::compute_accelerations (granular_system.h:65-81). */
hg_status hg_compute_accelerations(hg_system *sys);

/* Advances nsteps velocity-Verlet steps of size dt. When rebuild_every > 0
 * and the system is in neighbour-list mode, the list is rebuilt before every
 * step whose global index is a multiple of rebuild_every. Facets advance by
 * v_facet*dt after each step (triangular_facet.h:62-66).
 * Returns HG_ERR_NUMERIC when the closing kick leaves a velocity or angular
 * velocity NaN/Inf, the overflow and invalid-operation conditions the
 * reference's tests trap with feenableexcept (test/contact.cpp:9-18); the
 * state is left as the steps made it.
 * Replaces: do_step(dt) of the libtimestep system with
 * rotational_velocity_verlet_half + rotational_step_handler
 * (test/contact.cpp:98-106). */
hg_status hg_step(hg_system *sys, double dt, uint64_t nsteps, uint32_t rebuild_every);
/* hg_step without waiting for the device: the steps are queued and the call
 * returns. Table-capacity errors of the queued steps are reported by the next
 * call that reads results or state (hg_get_state, hg_get_accelerations,
 * hg_get_neighbors, hg_get_springs, hg_get_legacy_connectors, hg_diagnostics,
 * hg_energy, hg_set_state, hg_step). Replaces do_step as a driver calls it,
 * once per step with reads in between (Writerside/topics/Granular-system.md:65,
 * test/contact.cpp:105). */
hg_status hg_step_async(hg_system *sys, double dt, uint64_t nsteps, uint32_t rebuild_every);

/* Copies the state out (any pointer may be NULL).
 * Replaces: get_x/get_v/get_theta/get_omega (Granular-system.md:65-226). */
hg_status hg_get_state(hg_system *sys, double *x, double *v, double *theta, double *omega);
/* Overwrites the state (any pointer may be NULL); marks accelerations stale. */
hg_status hg_set_state(hg_system *sys, const double *x, const double *v, const double *theta,
                       const double *omega);
/* Last computed accelerations a (linear) and alpha (angular). */
hg_status hg_get_accelerations(hg_system *sys, double *a, double *alpha);
/* Sets both acceleration buffers to zero; the next step's first half kick
 * uses these zeros. Replaces: reset_acceleration_buffers()
 * (Granular-system.md:228-244). */
hg_status hg_reset_accelerations(hg_system *sys);
/* Simulation time and number of steps taken. */
hg_status hg_get_time(hg_system *sys, double *t, uint64_t *step_index);

/* Snapshot of the whole system (state, accelerations, the current pair-state
 * table, the neighbour list, bonds, facets with their springs, neck strengths,
 * legacy connectors, time and step count) as a self-checking byte string:
 * hg_load builds a system that continues exactly as the saved one would
 * (bit for bit). *size receives the byte count; the bytes are written when
 * buf is given and cap suffices (HG_ERR_CAPACITY otherwise). Not in the
 * reference, whose only dump is the VTK file of positions and orientation
 * axes (writer.cpp:51-118), which cannot restart a run. The format is
 * described in libgran_amd/csrc/snapshot.h; snapshots of pool systems
 * (hg_pool_save) load here too. */
hg_status hg_save(hg_system *sys, void *buf, size_t cap, size_t *size);
/* Creates a system on the current device from hg_save / hg_pool_save bytes,
 * with the parameters they carry. HG_ERR_INVALID when the bytes are damaged
 * (checksum) or inconsistent. */
hg_status hg_load(const void *buf, size_t size, hg_system **out);

/* Neighbour list as CSR: offsets[n+1]; idx written when cap suffices;
 * *total = number of entries. Each row is ascending. */
hg_status hg_get_neighbors(hg_system *sys, uint64_t *offsets, uint32_t *idx, size_t cap, size_t *total);

/* Bonds (i<j, in the lexicographic order the reference inserts them) and the
 * number of cycle-closing candidates skipped. Mirrors alt_sinter_functor's
 * public bonded_contacts table (alt_sinter_bridge.h:163). */
hg_status hg_get_bonds(hg_system *sys, uint32_t *pairs, size_t cap, size_t *nbonds, size_t *nskipped);

/* Active spring entries, ordered by (i, j): which = 0 contact springs
 * (contact_force.h:53-54,156), 1 bonded springs (alt_sinter_bridge.h:158-161).
 * ij = 2 per entry, springs = 9 per entry (tangential, rolling, torsion). */
hg_status hg_get_springs(hg_system *sys, int which, uint32_t *ij, double *springs, size_t cap,
                         size_t *total);

/* Device reductions of the reference diagnostics (test/compute_energy.cpp:7-58,
 * test/mass_distribution.cpp:7-27). out[9] = KE_total, KE_trans, KE_rot, |P|,
 * |L|, Rg, centre of mass (3). */
hg_status hg_diagnostics(hg_system *sys, double mass, double inertia, double *out);

/* Energy ledger in FP64 (the reference defines no potential energy; it
 * reports kinetic energy only, test/compute_energy.cpp:7-58). out[8] =
 * KE_trans, KE_rot, U_normal, U_tangential, U_rolling, U_torsion, U_hamaker,
 * E_total. Pairs are the evaluated ones (the neighbour list in that mode) and
 * each counts once: U_normal = k/2 overlap^2 for overlapping contacts and for
 * every bond (either sign), with the contact or sinter stiffness; the springs
 * k/2 |xi|^2 with the pair's two ordered copies averaged; U_hamaker the
 * README.md:154-157 potential (or -2Ar/h for HG_HAMAKER_SIMPLIFIED),
 * continued linearly below h0 and, for HG_HAMAKER_LIMITED, shifted to zero at
 * r_cutoff, so that its derivative is the force the functors apply. With
 * HG_BINARY_SINTER_LEGACY a bond's U_normal is its five connector springs,
 * k_d/2 (|s| - L0)^2 each (the potential of sinter_bridge.h:251), and the
 * spring terms hold the contact springs of unbonded pairs. Facet
 * terms are not included. Per-particle terms are summed in particle order. */
hg_status hg_energy(hg_system *sys, double mass, double inertia, double *out);

/* The alt_sinter bond search on its own (host only, no device needed): the
 * same bonds and skip count hg_create derives for HG_BINARY_SINTER.
 * Replaces: the topology part of the alt_sinter_functor ctor
 * (alt_sinter_bridge.h:56-90). pairs = 2*cap uint32. */
hg_status hg_sinter_topology(size_t n, const double *x0, double r_part, double d_crit, uint32_t *pairs, size_t cap,
                             size_t *nbonds, size_t *nskipped);

/* alt_breaking_sinter neck strengths, one per bond in hg_get_bonds order
 * (pairs i < j in lexicographic order, the order in which the reference's
 * constructor draws them, alt_breaking_sinter_bridge.h:63-75). The reference
 * stores each draw at neck_strengths[i] and [j] but reads [i*n + j]
 * (:71-72, :118); each bond keeps its own draw here, which that read intends.
 * Needs params.sinter_breaking. Replaces: the neck_strength_distribution draws
 * of the alt_breaking_sinter_functor ctor. */
hg_status hg_set_neck_strengths(hg_system *sys, const double *strengths);
/* The bond search of alt_breaking_sinter_functor on its own (host only): every
 * pair i < j with |norm(x_i - x_j) - 2 r| < d_crit, lexicographic order.
 * Replaces: alt_breaking_sinter_bridge.h:63-75. pairs = 2*cap uint32. */
hg_status hg_breaking_sinter_topology(size_t n, const double *x0, double r_part, double d_crit, uint32_t *pairs,
                                      size_t cap, size_t *nbonds);

/* ------------------------------------------------------------------------
 * Pools: many independent systems stepped together.
 *
 * Not in the reference, which simulates one granular_system per process. A
 * pool holds up to max_systems systems that share one hg_params (functor set,
 * pair mode, precision, dt) in one device arena of max_particles particles.
 * hg_pool_step advances every active system by the same number of steps with
 * one persistent kernel launch per rebuild interval: each workgroup takes one
 * whole system at a time (largest first), or in neighbour-list mode a pack of
 * small systems that together have at most one particle per thread, and runs
 * all of their steps with workgroup barriers between the kick/drift and force
 * phases. A system's
 * results do not depend on which other systems share the pool. Systems can be
 * added and removed between steps (continuous batching); a system's
 * per-step behaviour is that of an hg_system with the same params.
 * ------------------------------------------------------------------------ */
typedef struct hg_pool hg_pool;

/* max_particles < 2^31, max_systems < 2^24, max_facets: facets per system
 * (0..16). */
hg_status hg_pool_create(const hg_params *params, size_t max_particles, size_t max_systems, size_t max_facets,
                         hg_pool **out);
hg_status hg_pool_destroy(hg_pool *pool);
/* Adds a system; its particles must fit in one free range of the arena,
 * otherwise HG_ERR_CAPACITY (nothing is changed). Bonds are found on x0 as in
 * hg_create. *slot receives the system's id. */
hg_status hg_pool_add(hg_pool *pool, size_t n, const double *x0, const double *v0, const double *theta0,
                      const double *omega0, int32_t *slot);
/* Removes a system and frees its particle range. */
hg_status hg_pool_remove(hg_pool *pool, int32_t slot);
/* Facets of one system (nf <= max_facets), as hg_set_facets. */
hg_status hg_pool_set_facets(hg_pool *pool, int32_t slot, size_t nf, const double *verts, const double *v_facet);
/* Advances every active, unfaulted system by nsteps of size dt (any count).
 * In neighbour-list mode with rebuild_every > 0 each system's list is rebuilt
 * before every step whose index (the system's own step count) is a multiple
 * of rebuild_every, as hg_step does; systems at different phases of that
 * cadence share the launches up to the nearest rebuild of any of them.
 * A system whose neighbour list or pair-state table overflows, or whose
 * velocities become NaN/Inf, is marked faulted (hg_pool_fault) at the end of
 * the rebuild interval in which it happened and is not stepped again; the
 * other systems complete all nsteps unaffected. The call then returns
 * HG_ERR_NUMERIC if any system went non-finite, else HG_ERR_CAPACITY, and
 * hg_last_error() names the slots. Remove a faulted system with
 * hg_pool_remove; its state stays readable until then. */
hg_status hg_pool_step(hg_pool *pool, double dt, uint64_t nsteps, uint32_t rebuild_every);
/* Why a system stopped (0: it has not): HG_FAULT_* bits. Not in the
 * reference, whose tests stop the whole process on a floating-point trap
 * (test/contact.cpp:9-18) or an allocation failure (contact_force.h:53). */
enum { HG_FAULT_NEIGHBOR_CAPACITY = 1, HG_FAULT_PAIR_CAPACITY = 2, HG_FAULT_NUMERIC = 4 };
hg_status hg_pool_fault(hg_pool *pool, int32_t slot, int *fault);
/* Neck strengths of one system, as hg_set_neck_strengths. */
hg_status hg_pool_set_neck_strengths(hg_pool *pool, int32_t slot, const double *strengths);
/* hg_save for one system of the pool: saving, removing and later reloading a
 * request (hg_pool_load, into this pool, another pool with the same params,
 * or as a single system with hg_load) pre-empts it without changing its
 * result. */
hg_status hg_pool_save(hg_pool *pool, int32_t slot, void *buf, size_t cap, size_t *size);
/* Adds a system from hg_save / hg_pool_save bytes taken with this pool's
 * params (HG_ERR_INVALID otherwise); HG_ERR_CAPACITY when it does not fit. */
hg_status hg_pool_load(hg_pool *pool, const void *buf, size_t size, int32_t *slot);
/* Per-system copies, as the hg_system getters. */
hg_status hg_pool_get_state(hg_pool *pool, int32_t slot, double *x, double *v, double *theta, double *omega);
/* hg_energy for one system of the pool. */
hg_status hg_pool_energy(hg_pool *pool, int32_t slot, double mass, double inertia, double *out);
/* hg_pool_energy without allocating anything, for callers that cannot
 * allocate: serving loops that must not hipMalloc/hipFree between steps (the
 * ledger of each request as it retires), code that runs where allocation is
 * forbidden. No counterpart in the reference. hg_pool_energy allocates and
 * grows a per-pool buffer for the per-particle terms; this one writes them to
 * a fixed device array of the library (one per GPU, 8 MiB, held until the
 * process ends) and sums them on the device in particle order, so its result
 * equals hg_pool_energy's bit for bit. Systems of at most
 * HG_NOALLOC_MAX_PARTICLES particles (HG_ERR_CAPACITY beyond). All calls on
 * one GPU share that array: calls that overlap on one GPU (from two threads,
 * or on pools whose streams run concurrently) race, and the caller must
 * serialise them. */
enum { HG_NOALLOC_MAX_PARTICLES = 131072 };
hg_status hg_pool_energy_noalloc(hg_pool *pool, int32_t slot, double mass, double inertia, double *out);
hg_status hg_pool_get_accelerations(hg_pool *pool, int32_t slot, double *a, double *alpha);
hg_status hg_pool_get_neighbors(hg_pool *pool, int32_t slot, uint64_t *offsets, uint32_t *idx, size_t cap,
                                size_t *total);
hg_status hg_pool_get_springs(hg_pool *pool, int32_t slot, int which, uint32_t *ij, double *springs, size_t cap,
                              size_t *total);
hg_status hg_pool_get_bonds(hg_pool *pool, int32_t slot, uint32_t *pairs, size_t cap, size_t *nbonds,
                            size_t *nskipped);
/* Particle count, simulation time and step count of one system. */
hg_status hg_pool_info(hg_pool *pool, int32_t slot, size_t *n, double *t, uint64_t *step_index);
/* Bytes of device memory one system of n particles occupies in a pool with
 * these params (state, neighbour list, pair-state table, facet springs):
 * the admission unit of an HBM budget. */
size_t hg_pool_bytes_per_particle(const hg_params *params, size_t max_facets);
/* Device time of the last hg_pool_step's chunk launches (milliseconds, HIP
 * events on the pool's stream), the number of launches and the steps they ran
 * in total (a launch runs one rebuild interval, or the part of one that the
 * call covers). */
hg_status hg_pool_last_step_profile(hg_pool *pool, double *chunk_ms, double *rebuild_ms, uint64_t *chunks,
                                    uint64_t *chunk_steps);
/* Systems of at least wide_min particles (default 2048; 0: none) are stepped
 * by per-step launches over all their particles, one thread each, on a second
 * stream beside the chunk kernel, instead of by one workgroup: a large system
 * then no longer sets the critical path of the interval. The per-particle
 * operations are the same, so results do not change. Takes effect at the next
 * hg_pool_step. */
hg_status hg_pool_set_wide_threshold(hg_pool *pool, size_t wide_min);
/* While both kinds of systems are present, the chunk kernel's persistent grid
 * shrinks to wg_per_cu * (CUs - wide_cus) workgroups, leaving room worth
 * wide_cus CUs (default a quarter) for the wide launches that run beside it
 * on a second stream. wide_cus = 0 runs them one after the other instead:
 * the interval's wide steps, then the chunk kernel, each on the whole GPU.
 * 0 <= wide_cus < CU count. */
hg_status hg_pool_set_wide_cus(hg_pool *pool, int wide_cus);
/* The wide launches of one interval (2 per step + 1) are recorded once into a
 * HIP graph per (interval length, dt, wide set) and replayed; enable = 0
 * launches them one by one instead. Default on. */
hg_status hg_pool_set_wide_graph(hg_pool *pool, int enable);
/* Packing (neighbour-list pools, default on): the chunk kernel steps small
 * systems side by side in one workgroup, one particle per thread, instead of
 * one system per workgroup. Results are the same either way; enable = 0
 * gives every system a workgroup of its own. */
hg_status hg_pool_set_packing(hg_pool *pool, int enable);
/* Online autotuner: from the next hg_pool_step on, first the chunk kernel's
 * shape (256 threads x 2 workgroups per CU, or 512 x 1) at the current launch
 * split, then the launch split at the better shape (wide threshold 0, or
 * 1024..8192 up to the largest system, run one after the other or beside the
 * chunk kernel with room for CUs/32, /16, /8 or /4). Each
 * candidate runs two rebuild intervals of live steps, scored by the second
 * one's host-measured time per step; the fastest is kept. Neither choice
 * changes results, so no step is wasted. enable = 0 stops it and keeps the
 * current setting. */
hg_status hg_pool_autotune(hg_pool *pool, int enable);
/* Current split and whether the autotuner has finished (tuned = 1). */
hg_status hg_pool_get_tuning(hg_pool *pool, size_t *wide_min, int *wide_cus, int *tuned);
/* Device-busy time of the last hg_pool_step (milliseconds): over its
 * launches, the time from an interval's first device operation to its last,
 * both streams joined (HIP events); the rest of the call's wall time is host
 * work between intervals. */
hg_status hg_pool_last_busy(hg_pool *pool, double *busy_ms);
/* Shape of the chunk kernel: threads per workgroup (256 or 512, default 256)
 * and workgroups per CU (default 2; at most 2048 / threads). The grid is
 * wg_per_cu per CU, and each workgroup gets that fraction of the CU's LDS,
 * which decides which systems keep their state in LDS. Results do not depend
 * on it; hg_pool_autotune tries the shapes. */
hg_status hg_pool_set_chunk_shape(hg_pool *pool, int threads, int wg_per_cu);
hg_status hg_pool_get_chunk_shape(hg_pool *pool, int *threads, int *wg_per_cu);
/* Device time of the last hg_pool_step's wide launches (second stream, HIP
 * events; milliseconds) and their number. */
hg_status hg_pool_last_wide_profile(hg_pool *pool, double *wide_ms, uint64_t *launches);

/* ------------------------------------------------------------------------
 * Server: continuous batching of independent simulation requests.
 *
 * Not in the reference, whose drivers run one granular_system to completion
 * per process. A server owns one pool (arena sized by budget_bytes through
 * hg_pool_bytes_per_particle: state, neighbour list, pair-state table and
 * facet springs) and one worker thread. Requests (initial state, facets, step
 * count) are submitted by one producer thread and queue in arrival order. The
 * worker admits the head of the queue as soon as its particles fit a free
 * range of the arena and a slot is free. Each worker round queues one
 * interval of the resident requests (at most `quantum` steps, and no further
 * than the nearest rebuild or the end of a request) and, while it runs on the
 * device, reads back the requests that finished in the previous round and
 * uploads the admitted ones on a second stream; the host work of an admission
 * (bond search, pair-state rows) is done by hg_server_submit on the
 * producer's thread. A request whose state becomes NaN or infinite, or whose
 * neighbour list or pair-state table overflows, fails alone (hg_server_wait
 * returns HG_ERR_NUMERIC or HG_ERR_CAPACITY) and the others go on. A
 * request's result does not depend on the other requests on the device.
 * Callers are the serving stream's producers.
 * ------------------------------------------------------------------------ */
typedef struct hg_server hg_server;

typedef struct {
    uint64_t budget_bytes;  /* device memory for the pool arena */
    uint32_t max_requests;  /* requests resident on the device at once */
    uint32_t max_facets;    /* facets per request (0..16) */
    uint32_t quantum;       /* most steps per worker round (any nsteps is accepted) */
    uint32_t rebuild_every; /* neighbour-list cadence */
    double dt;              /* time step of every request */
} hg_server_config;

typedef struct {
    uint64_t submitted, queued, running, done, failed;
    uint64_t rounds;         /* worker rounds that stepped requests */
    uint64_t particle_steps; /* sum over rounds of stepped particles x the round's steps */
    double device_ms;        /* device-busy time of the rounds (first to last operation, HIP events) */
    double elapsed_s;        /* seconds since hg_server_create */
} hg_server_stats;

enum { HG_REQ_QUEUED = 0, HG_REQ_RUNNING = 1, HG_REQ_DONE = 2, HG_REQ_FAILED = 3 };

hg_status hg_server_create(const hg_params *params, const hg_server_config *cfg, hg_server **out);
/* Stops the worker after its current round; unfinished requests are dropped. */
hg_status hg_server_destroy(hg_server *srv);
/* Queues a request (inputs are copied; v0, theta0, omega0 may be NULL for
 * zeros; nf facets as hg_set_facets). *id identifies it in later calls. A
 * request larger than the whole budget is refused with HG_ERR_CAPACITY. */
hg_status hg_server_submit(hg_server *srv, size_t n, const double *x0, const double *v0, const double *theta0,
                           const double *omega0, size_t nf, const double *verts, const double *v_facet,
                           uint64_t nsteps, uint64_t *id);
/* HG_REQ_* state of a request not yet released by hg_server_wait. */
hg_status hg_server_status(hg_server *srv, uint64_t id, int *state);
/* Blocks until the request has finished (timeout_s < 0: no limit; on time-out
 * HG_ERR_CAPACITY and the request stays), copies its final x, v, theta, omega
 * (any may be NULL) and times[3] = {submitted, admitted, done} in seconds
 * since hg_server_create, and releases it. A failed request is released too
 * and returns its own status (HG_ERR_NUMERIC, HG_ERR_CAPACITY, HG_ERR_HIP). */
hg_status hg_server_wait(hg_server *srv, uint64_t id, double timeout_s, double *x, double *v, double *theta,
                         double *omega, double *times);
hg_status hg_server_get_stats(hg_server *srv, hg_server_stats *out);

/* Device-time profile of a system's launches, measured with HIP events
 * recorded on the system's stream around each launch (not in the reference,
 * which only times whole runs with std::chrono, main.cpp:146-164).
 * hg_profile_enable(sys, 1) resets and starts; 0 stops. hg_profile_read
 * synchronises the stream and returns, per phase, the launch count and the
 * summed device milliseconds: [0] acceleration pass, [1] kick/drift and
 * final kick, [2] neighbour rebuild (all of its kernels), [3] reserved. */
hg_status hg_profile_enable(hg_system *sys, int enable);
hg_status hg_profile_read(hg_system *sys, uint64_t *launches, double *device_ms);
/* Small FP64 systems (up to 256 particles) step through one launch per
 * rebuild interval; enable = 1 makes hg_step use the per-step launches (the
 * path every larger system takes) without the event timing of
 * hg_profile_enable, enable = 0 restores the choice by size. Results are the
 * same either way. Not in the reference (one code path per functor there):
 * the per-step baseline against which the one-launch path is timed. */
hg_status hg_set_per_step_launches(hg_system *sys, int enable);
/* Cumulative dissipated work, the ledger's counterpart of the energy the
 * damping and friction take out (not in the reference, which reports kinetic
 * energy only). enable = 1 zeroes the accumulator and from the next step adds
 * each ordered pair's half of its pair's loss: gamma_n v_n^2 dt, and for the
 * tangential, rolling and torsion springs the work their forces take from the
 * relative motion (the rolling term weighted r/rpp, as the reference's rolling
 * torque and velocity differ by that factor) minus the change of the spring's
 * energy k/2 |xi|^2; a separating contact loses its springs whole. With it,
 * hg_energy's total plus this work stays constant to the integrator's error.
 * FP64 systems with the contact or alt_sinter functor (bonds that do not
 * break) only; facet terms are not counted, as in hg_energy; steps run as
 * per-step launches while it is on. enable = 0 stops adding. */
hg_status hg_set_dissipation_tracking(hg_system *sys, int enable);
/* out[5]: dissipated work since tracking was enabled (joules): normal,
 * tangential, rolling, torsion and their total, summed in particle order. */
hg_status hg_get_dissipated_work(hg_system *sys, double *out);

/* Seeded tunable-dimension cluster-cluster aggregate of n touching spheres of
 * radius r_part (fractal dimension df, prefactor kf). x_out = 3*n doubles.
 * Not in the reference: the input generator of the aggregate workloads. */
hg_status hg_generate_aggregate(size_t n, double r_part, double df, double kf, uint64_t seed, double *x_out);
/* count aggregates at once on host threads (threads <= 0: all cores);
 * aggregate k has n[k] particles, dimension df[k], seed seeds[k]; positions
 * are concatenated in x_out (3 * sum(n) doubles). Each equals
 * hg_generate_aggregate with the same arguments. */
hg_status hg_generate_aggregates(size_t count, const uint32_t *n, const double *df, double kf, const uint64_t *seeds,
                                 double r_part, double *x_out, int threads);

/* A caller's source of random unit vectors, called as the reference's
 * std::function<field_value_t()> random_unit_vector_generator
 * (sinter_bridge.h:125): writes one vector to out[0..2]. */
typedef void (*hg_unit_vector_fn)(void *user, double out[3]);
/* The legacy sinter_functor constructor's set-up (sinter_bridge.h:138-211),
 * on the host: bonds are the pairs i < j, in row order, with
 * |(|x_j - x_i| - 2 r_part)| <= critical_separation; for each, unit vectors
 * are drawn from `next` until one is not collinear with the bond normal, and
 * the bond's frame gives ten connector points: connectors[30 b + 3 q + c] is
 * point q (0..4) attached to particle pairs[2b], connectors[30 b + 15 + 3 q
 * + c] to particle pairs[2b + 1]. *nbonds gets the bond count; pairs and
 * connectors are written, and `next` called, only when both are given and
 * cap >= *nbonds. The draws happen in the reference's order, so a generator
 * that the reference's constructor would consume gives its connectors. */
hg_status hg_legacy_sinter_frames(size_t n, const double *x0, double r_part, double critical_separation,
                                  hg_unit_vector_fn next, void *user, uint32_t *pairs, double *connectors, size_t cap,
                                  size_t *nbonds);
/* Gives an HG_BINARY_SINTER_LEGACY system its connectors, in the layout of
 * hg_legacy_sinter_frames; pairs must be the system's bonds (hg_get_bonds)
 * in order. Required before the first acceleration pass; later calls replace
 * the connectors. Replaces the connector state sinter_functor keeps for its
 * step handler (sinter_bridge.h:261-264). */
hg_status hg_set_legacy_connectors(hg_system *sys, size_t nbonds, const uint32_t *pairs, const double *connectors);
/* The connectors as they are now (moved by every step), in the same layout. */
hg_status hg_get_legacy_connectors(hg_system *sys, double *connectors, size_t cap, size_t *nbonds);

/* Legacy-VTK ASCII POLYDATA of n particles at path: positions divided by
 * r_part, and fields thetax, thetay, thetaz holding each particle's body x,
 * y and z axes for the orientation theta (rotations about the fixed x, y, z
 * axes, composed left to right). Numbers as an std::ostream prints doubles
 * (printf "%g"), the reference's byte layout. Replaces write_particles
 * (writer.cpp:51-118), which names the file itself (dir/particles_<k>.vtk);
 * here the caller passes the whole path. HG_ERR_IO if it cannot be written. */
hg_status hg_write_vtk_particles(const char *path, size_t n, const double *x, const double *theta, double r_part);
/* Legacy-VTK ASCII POLYDATA of sinter necks at path: for each of the nbonds
 * pairs (i, j) = (pairs[2b], pairs[2b+1]), in the given order, the midpoint
 * of x_i and x_j divided by r_part and the field `normals` holding
 * (x_i - x_j) / |x_i - x_j|. Replaces write_necks (writer.cpp:120-172), which
 * takes the n*n enabled_contacts matrix and lists its pairs i < j row by
 * row; hg_get_bonds returns the alt_sinter bonds in that order. */
hg_status hg_write_vtk_necks(const char *path, size_t n, const double *x, size_t nbonds, const uint32_t *pairs,
                             double r_part);

#ifdef __cplusplus
}
#endif

#endif /* HIPGRAN_H */
