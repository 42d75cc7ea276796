// pdmd_device.h -- device-resident layout of one PRIME20 system (shared by
// the host set-up code and the gfx950 kernels).
//
// Layout in HBM, one allocation set per system, structure-of-arrays:
//   rec[N]      64 B  {x,y,z,vx,vy,vz,mass, topo, dyn}: one 64-byte segment
//                     per gather carries everything a collision-time
//                     prediction needs from a bead
//   bw[N]       16 B  bond window of the bead's own side-chain bond
//                     (Calpha: code 10, N: code 11, C: code 12)
//   ovr/extra   int4  pair-code overrides and H-bond shoulder partners
//   nb_up/nb_dn int32 [N*cap]: partner | static pair code << 24
//   leaf[N+3]   16 B  {tim, nptnr, coltype|bin} per calendar entry: beads,
//                     then the ghost, sync and output pseudo-events
// Indices are 0-based on the device; the C ABI converts to the reference's
// 1-based numbering.
#pragma once
#include <cstdint>

namespace pdmd {

// topo word: chain (11 bits: at most 2047 chains) | bead of a chain built
// as -Dchaptype=2 builds the second species (1: no bonds or intra-chain
// wells) | system built with -Dno_h (1) | kind (2: 0 Calpha, 1 N, 2 C, 3 R)
// | residue-1 (7) | chain length-1 (7) | hp (1) | proline residue (1) |
// system built with -Dno_hbs (1, the sign bit)
// dyn word:  identity (5 bits) | (H-bond partner + 1) << 5
struct __attribute__((aligned(16))) Rec {
    double x, y, z, vx, vy, vz, bm;
    int topo, dyn;
};

struct __attribute__((aligned(16))) Leaf {
    double tim;
    int nptnr;
    int meta;  // (bin + 1) in the low 16 bits, coltype in the high 16 (signed)
};

__host__ __device__ inline int topo_chain(int t) { return t & 0x7FF; }
// -Dchaptype=2: the bead's chain is of the unbonded second species
__host__ __device__ inline int topo_free(int t) { return (t >> 11) & 1; }
// -Dno_h: side chains of different chains never form code-16 wells
__host__ __device__ inline int topo_noh(int t) { return (t >> 12) & 1; }
__host__ __device__ inline int topo_kind(int t) { return (t >> 13) & 3; }
__host__ __device__ inline int topo_res(int t) { return ((t >> 15) & 127) + 1; }
__host__ __device__ inline int topo_len(int t) { return ((t >> 22) & 127) + 1; }
__host__ __device__ inline int topo_hp(int t) { return (t >> 29) & 1; }
__host__ __device__ inline int topo_pro(int t) { return (t >> 30) & 1; }
__host__ __device__ inline int topo_nohbs(int t) { return t < 0 ? 1 : 0; }
__host__ __device__ inline int make_topo(int chain, int kind, int r, int L, int hp, int pro, int nohbs, int noh,
                                         int free_ = 0) {
    return (int)((unsigned)((chain & 0x7FF) | (free_ << 11) | (noh << 12) | (kind << 13) | ((r - 1) << 15) |
                            ((L - 1) << 22) | (hp << 29) | (pro << 30)) |
                 ((unsigned)nohbs << 31));
}
__host__ __device__ inline int dyn_ident(int d) { return d & 31; }
__host__ __device__ inline int dyn_bptnr(int d) { return (d >> 5) - 1; }
__host__ __device__ inline int make_dyn(int ident, int bptnr) { return ident | ((bptnr + 1) << 5); }
__host__ __device__ inline int leaf_bin(int meta) { return (meta & 0xFFFF) - 1; }
__host__ __device__ inline int leaf_ct(int meta) { return meta >> 16; }
__host__ __device__ inline int make_meta(int bin, int ct) { return ((bin + 1) & 0xFFFF) | (ct << 16); }

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
    // nbor.f's scan order of neighbour cells (cell_link.f:13-94): for the
    // offset (dx, dy, dz) in [-2, 2]^3 at (dx+2) + 5 (dy+2) + 25 (dz+2), its
    // place 1..62 in the half stencil, 0 where the opposite offset is in it
    signed char halfst[128];
};

// Slots of SysState::phase (pdmd_stats_t::phase_cycles).
constexpr int kPhaseSlots = 40;

// Mutable per-system scalars (the reference's module GLOBAL scalars).
struct SysState {
    double tfalse, t, interval, interval_max, t_fact, n_forced, sortsize, tbin_off;
    double old_tfalse, pad0;
    long long coll, numghosts, nupdates, nforced;
    unsigned long long rng_ctr[2];
    long long nevents[32];
    int nbin, status, log_count, pad1;
    long long target;
    long long max_nb_seen;
    long long phase[kPhaseSlots];  // device cycles and counts per engine phase (PDMD_PHASE_TIMERS builds only)
    long long run_ticks;      // wall_clock64 ticks this system spent in event-loop launches
    long long rebuild_ticks;  // ... of which in list rebuilds (with events())
    long long rebuilds;       // list rebuilds those ticks cover
};

struct LogRec {
    long long coll;
    int i, j, type, pad;
    double t;
};

struct SysDev {
    int N, nres, cap, canon;
    int M, Mq, ncoarse, log_cap;  // fine cells per axis (interior), coarse cells per axis
    int chaptype2;                // built as -Dchaptype=2 (checkover.f:67 then skips bond windows)
    int waves_cap;                // most pair events per round for this system (0: the launch's wavefronts)
    int cq, pad;                  // fine cells per coarse cell of the list rebuild's search (2..4)
    double width, half, setemp, avegtime, hdelr;
    unsigned long long seed;
    const DevTables* tab;
    Rec* rec;
    double2* bw;
    int4* extra;
    int4* ovr;
    double4* old_r;
    int* nb_up;
    int* nup;
    int* nb_dn;
    int* ndn;
    Leaf* leaf;
    int* cfine;    // packed fine cell coordinates per bead
    int* ckey;     // coarse cell per bead
    int2* cfp;     // [N] fine cell and the place in its list as cell_add.f builds it (1 = highest index)
    int* cstart;   // [ncoarse+1]
    int* crun;     // [ncoarse]
    int* csorted;  // [N]
    double4* spos; // [N] positions in coarse-cell order (list rebuild)
    int4* sinfo;   // [N] bead, topo word, fine cell, 0 in coarse-cell order
    unsigned* mark;  // [N] per-round footprint marks of a multi-wavefront system (zero between rounds)
    // [N+3] when each calendar entry was last added (add_tbin.f): the
    // committed-entry count of the event that added it << 20 | the add's
    // place within that event; of two entries with equal bin and time the
    // later-added one comes first, as in the reference's LIFO bucket lists
    unsigned long long* stamp;
    SysState* st;
    LogRec* log;
};

// status codes (returned through the C ABI)
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
