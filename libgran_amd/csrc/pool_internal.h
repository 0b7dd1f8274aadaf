// Internal interface between the pool (pool.hip) and its callers inside the
// library (server.cpp): adding a system in two parts, so that the host work
// can run on another thread ahead of the device work.
#pragma once

#include <utility>
#include <vector>

#include "hg_internal.h"

namespace hg {

// The host work of adding a system: the bond search and the sorted pair-state
// rows (local indices | bond bit), and the state in the device's 4-double
// records. Needs no device and no pool, only the pool's parameters.
struct pool_prep {
    size_t n = 0;
    std::vector<double> q4[4];  // x, v, theta, omega
    std::vector<std::pair<uint32_t, uint32_t>> bonds;
    size_t skipped = 0;
    std::vector<std::vector<uint32_t>> rows;
};

// given_bonds (with given_skipped) replaces the bond search on x0.
hg_status pool_prepare(const hg_params &p, size_t n, const double *x0, const double *v0, const double *theta0,
                       const double *omega0, const std::vector<std::pair<uint32_t, uint32_t>> *given_bonds,
                       size_t given_skipped, pool_prep &out);
// Places a prepared system in a free slot and particle range of the pool and
// uploads it on stream cs (HG_ERR_CAPACITY, with nothing changed, when none
// is free). The new system joins the launch order at the next pool_begin;
// while an interval is queued, cs must not be the pool's own stream for the
// copies to overlap it (pool_copy_stream).
hg_status pool_commit(hg_pool *Q, pool_prep &prep, int32_t *slot_out, hipStream_t cs);
// hg_pool_set_facets with its copies on stream cs.
hg_status pool_set_facets(hg_pool *Q, int32_t slot, size_t nf, const double *verts, const double *v_facet,
                          hipStream_t cs);

// One rebuild interval (at most max_len steps, fewer when a system is due
// for a rebuild sooner), queued and not waited for; *len_out receives its
// length (0: nothing to step). pool_end waits for it, updates the host mirror
// and appends the systems that faulted to `faulted`. In between, the host may
// read back or release parked systems and commit new ones on another stream.
// That is safe because the queued kernels reach particles through the launch
// order, the wide block map and the neighbour lists of the systems in the
// interval, except the rebuild's binning pass, which scans owner[] below the
// high-water mark but skips every particle whose slot has no rebuild flag in
// the device meta: the flag is clear for every slot outside the interval.
hg_status pool_begin(hg_pool *Q, double dt, uint64_t max_len, uint32_t rebuild_every, uint32_t *len_out);
hg_status pool_end(hg_pool *Q, std::vector<int32_t> &faulted);
// HG_OK, or the hg_pool_step status naming the faulted systems.
hg_status pool_fault_status(hg_pool *Q, const std::vector<int32_t> &faulted, const char *where);
void pool_reset_profile(hg_pool *Q);
// Takes a system out of the launch order from the next pool_begin on,
// without freeing it (a finished request whose state is still to be read).
void pool_park(hg_pool *Q, int32_t slot);
// Frees a system's slot and range (hg_pool_remove without waiting for the
// pool's stream: the system must not be in a queued interval).
hg_status pool_release(hg_pool *Q, int32_t slot, hipStream_t cs);
// hg_pool_set_neck_strengths with the copies on stream cs.
hg_status pool_set_necks(hg_pool *Q, int32_t slot, const double *strengths, hipStream_t cs);
// hg_pool_get_state with the copies on stream cs.
hg_status pool_read_state(hg_pool *Q, int32_t slot, double *x, double *v, double *theta, double *omega,
                          hipStream_t cs);
hipStream_t pool_copy_stream(hg_pool *Q);

}  // namespace hg
