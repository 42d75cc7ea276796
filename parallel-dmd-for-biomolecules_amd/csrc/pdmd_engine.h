// pdmd_engine.h -- launch wrappers of the gfx950 DMD kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "pdmd_device.h"

namespace pdmd {

// LDS bytes of one system's workgroup of nw wavefronts for a calendar of
// nblk 16-entry blocks, work lists of maxw entries and leaf_cap leaves in LDS.
size_t engine_lds_bytes(int nblk, int maxw, int leaf_cap, int nw);
// most wavefronts per system whose work areas fit in lds_limit
int engine_max_waves(int max_n, int maxw, size_t lds_limit);
// work-list capacity of the event kernels for a list capacity
int engine_maxw(int cap);
// Opt the event-loop kernels in to lds_limit bytes of LDS per workgroup.
hipError_t engine_prepare(size_t lds_limit);
// Lists, pseudo-event seeds and events() for n systems (one workgroup of up
// to 4 wavefronts each, as many as fit in lds_limit).
hipError_t engine_setup(const SysDev* d_sys, const int* d_ids, int n, int max_n, int maxd, size_t lds_limit,
                        hipStream_t s);
// A list update forced at the current time for n systems (origin shift,
// nbor + events), one workgroup of up to 4 wavefronts each.
hipError_t engine_rebuild(const SysDev* d_sys, const int* d_ids, int n, int max_n, int maxd, size_t lds_limit,
                          hipStream_t s);
// Event loop: every listed system runs until its SysState::target.
// budget_ticks: wall_clock64 ticks after the launch's start after which
// each system stops between rounds (0 = run to the event targets), with
// launch_t0 a device word that takes the start (needed with a budget);
// waves: wavefronts per system (1 to 8, lowered to what fits in lds_limit),
// each round runs up to that many conflict-free pair events side by side
constexpr int kStopAtOutput = 1;  // engine_run flag: stop after the next output pseudo-event
hipError_t engine_run(const SysDev* d_sys, const int* d_ids, const long long* d_inc, int n, int max_n, int maxw,
                      size_t lds_limit, hipStream_t s, long long budget_ticks = 0, int flags = 0, int waves = 1,
                      unsigned long long* launch_t0 = nullptr);
// Debug checks (checkover.f, check_nc_int.f) of system id's n beads: out[8]
// counters (zeroed by the caller, out[3] preset to all ones); alt replaces
// the resident state records when not null.
hipError_t engine_check(const SysDev* d_sys, int id, int n, const Rec* alt, unsigned long long* out, hipStream_t s);
// Observables of system id (energy.f, radgyr.f, end2end.f) on one
// workgroup: outd[7] ered, tred, sumvel, ehh_ii, ehh_ij, rg_avg, e2e_avg;
// outi[3] hb_alpha, hb_ii, hb_ij.  cstart/clen: nch chains (device arrays);
// nfirst: chains of the leading species block; boxl in Angstrom.
hipError_t engine_observables(const SysDev* d_sys, int id, const int* cstart, const int* clen, int nch, int nfirst,
                              double boxl, double* outd, int* outi, hipStream_t s);
// The same without device allocation (pdmd_observables_noalloc): the chain
// tables (host arrays, nch <= 4096) and the results pass through a fixed
// 33 KB device array per GPU; returns after the results reached h_outd[7]
// and h_outi[3].  Calls that overlap on one GPU race on that array.
hipError_t engine_observables_noalloc(const SysDev* d_sys, int id, int device, const int* h_cstart, const int* h_clen,
                                      int nch, int nfirst, double boxl, double* h_outd, int* h_outi, hipStream_t s);
// Looks up this GPU's copy of that array (device must be current); done
// once per device at context creation.
hipError_t engine_scratch_init(int device);
// Test hook: collision times of (i, j) pairs (1-based) of system id.
hipError_t engine_predict(const SysDev* d_sys, int id, int n, const int* pi, const int* pj, double* tij, int* type,
                          hipStream_t s);

}  // namespace pdmd
