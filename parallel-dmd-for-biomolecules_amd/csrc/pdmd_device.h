// pdmd_device.h -- device-resident layout of one PRIME20 system (shared by
// the host set-up code and the gfx950 kernels).
//
// Layout in HBM, one allocation set per system, structure-of-arrays:
//   bead[N]        64 B  {x,y,z,vx,vy,vz,mass,-}: one 64-byte segment per
//                        gather, the dominant access of the predictor
//   ident/bptnr    int32 per bead, extra/ovr int4 per bead
//   nb_up/nb_dn    int32 [N*cap]: partner | static pair code << 24
//   tim/nptnr/coltype/cbin over N+3 calendar entries (beads, then the
//                        ghost, sync and output pseudo-events)
// Indices are 0-based on the device; the C ABI converts to the reference's
// 1-based numbering.
#pragma once
#include <cstdint>

namespace pdmd {

struct __attribute__((aligned(16))) Bead {
    double x, y, z, vx, vy, vz, bm, pad;
};

// Scaled parameter tables of one system (box units), from
// scale_down.f / make_code.f / nbor_setup.f.
struct DevTables {
    double sigma_sq[29][29];
    double welldia_sq[29][29];
    double ep_sqrt[29][29];
    double shlddia_sq[29][29];
    double ev_param[4][51];
    double sqz610[6][29];
    double rlsq[51];
    double shder[5];
    double epsilon1;
};

// Mutable per-system scalars (the reference's module GLOBAL scalars).
struct SysState {
    double tfalse, t, interval, interval_max, t_fact, n_forced, sortsize, tbin_off;
    double old_tfalse, pad0;
    long long coll, numghosts, nupdates, nforced;
    unsigned long long rng_ctr[2];
    long long nevents[32];
    int nbin, status, log_count, rebuild_pending;
    long long target;     // stop once coll reaches this
    long long max_nb_seen;
};

struct LogRec {
    long long coll;
    int i, j, type, pad;
    double t;
};

struct SysDev {
    int N, nres, cap, canon;
    int M, Mq, ncoarse, log_cap;      // fine cells per dim (interior), coarse cells per dim
    double width, half, setemp, avegtime, hdelr;
    unsigned long long seed;
    const DevTables* tab;
    Bead* bead;
    int* ident;
    int* bptnr;
    int4* extra;
    int4* ovr;
    const int* chain;
    const int* loc;
    const int* clen;
    const int* res;
    const int* restype;
    const int* hp;
    const double* bondw;  // per residue: blmin/blmax for codes 10, 11, 12
    double4* old_r;
    int* nb_up;
    int* nup;
    int* nb_dn;
    int* ndn;
    double* tim;
    int* nptnr;
    int* coltype;
    int* cbin;
    int* cfine;    // packed fine cell coordinates per bead
    int* ckey;     // coarse cell per bead
    int* cstart;   // [ncoarse+1]
    int* crun;     // [ncoarse]
    int* csorted;  // [N]
    SysState* st;
    LogRec* log;
};

// status codes (also returned through the C ABI, negated)
enum : int {
    kOk = 0,
    kErrListCapacity = 1,
    kErrNegativeTime = 2,
    kErrBadCode = 3,
    kErrOverrideSlots = 4,
    kErrRng = 5,
    kErrBin = 6,
    kErrDyn = 7,
};

}  // namespace pdmd
