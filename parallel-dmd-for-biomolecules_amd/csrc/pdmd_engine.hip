// pdmd_engine.hip -- gfx950 kernels of the PRIME20 discontinuous-MD event
// loop.  One workgroup of W wavefronts (64 lanes each) owns one system for a
// whole launch and commits its events in the reference's serial order.
//   * Each round pops up to W calendar entries; the pair events among them
//     run side by side, one per wavefront, and the round commits the prefix
//     the serial order would have run the same way (per-bead footprint marks
//     in HBM, no new calendar key ahead of a later popped one).
//   * Inside an event the wavefront issues every collision-time prediction
//     the event triggers (full recomputes of the colliding pair and of the
//     down beads that pointed at it, single-pair updates of the down set) as
//     one batch of 64-byte gathers, one lane per pair, FP64 in the
//     reference's evaluation order (no contraction).
//   * The calendar is a three-level min tree (16-entry leaf blocks, 64-block
//     superblocks) whose block/superblock minima carry partner and type, so
//     the next entry comes out of one LDS min-reduction; leaves live in LDS
//     when the system fits, else in HBM.
//   * Neighbour-list rebuilds (coarse-cell counting sort + per-bead search),
//     free flight and displacement checks at the sync event run on all W
//     wavefronts.
// Many systems in one launch run concurrently, one workgroup each.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <mutex>

#include "pdmd_device.h"
#include "pdmd_engine.h"

namespace pdmd {
namespace {

constexpr int kW = 64;
constexpr int kBlk = 16;
constexpr int kSup = 64;
constexpr int kNumbin = 2000;
constexpr int kMaxBlist = 256;
constexpr double kLt = 1e-10;
constexpr double kSm = 5e-12;
constexpr int kMaxWaves = 8;  // most wavefronts in one system's workgroup

__device__ inline int lane_id() { return (int)(threadIdx.x & (kW - 1)); }
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// Phase timers (device clock cycles summed into SysState::phase[k]) when
// built with -DPDMD_PHASE_TIMERS; no code otherwise.  Phases: 0 selection,
// 1 pair event before rescheduling, 2 reschedule loads and work list,
// 3 reschedule predictions, 4 down-set updates, 5 leaf writes, 6 block
// rescans, 7 sync event without rebuild, 8 list rebuild, 9 full events(),
// 10 ghost event before rescheduling, 11 reschedule rows and work list,
// 12 reschedule down-set scan, 13 rebuild count, 14 sync count; list
// rebuild steps: 16 cells and histogram, 17 scan, 18 scatter, 19 sorted
// copies, 20 neighbour search; multi-wavefront rounds: 21 rounds of two or
// more pair events, 22 entries popped in them, 23 entries committed, 24
// footprint marks up to the first round barrier, 25 footprint checks up to
// the second, 26 new keys up to the third, 27 the round's rescans and merge,
// 28 rounds cut by a footprint conflict or an event that cannot speculate,
// 29 rounds cut by the key order; round selection: 30 the pops, 31 the
// scalar copies after them; rescheduling batches: 32 pair predictions made
// in them, 33 their 64-lane passes (32 / (64 * 33) is their lane occupancy);
// down-set work outside the batches: 34 single-pair updates predicted one
// per lane, 35 full recomputes run one bead at a time, 36 events rescheduled
// by the generic form.
enum { kPhSelect = 0, kPhPair = 1, kPhLoads = 2, kPhPred = 3, kPhDown = 4, kPhLeaf = 5, kPhRescan = 6,
       kPhSync = 7, kPhLists = 8, kPhEvents = 9, kPhGhost = 10, kPhRows = 11, kPhScan = 12, kPhNRebuild = 13,
       kPhNSync = 14, kPhLCells = 16, kPhLScan = 17, kPhLScatter = 18, kPhLCopy = 19, kPhLSearch = 20,
       kPhNRound = 21, kPhNPopped = 22, kPhNCommit = 23, kPhBMark = 24, kPhBCheck = 25, kPhBKeys = 26,
       kPhBTail = 27, kPhNCutMark = 28, kPhNCutKey = 29, kPhSelPop = 30, kPhSelCopy = 31, kPhNPred = 32,
       kPhNPredPass = 33, kPhNSingle = 34, kPhNFull = 35, kPhNGeneric = 36 };
#ifdef PDMD_PHASE_TIMERS
#define PH_BEGIN(v) long long v = clock64()
#define PH_END(S, k, v)                                                                                  \
    do {                                                                                                 \
        if (threadIdx.x == 0)                                                                            \
            atomicAdd((unsigned long long*)&(S).st->phase[k], (unsigned long long)(clock64() - (v)));   \
        v = clock64();                                                                                   \
    } while (0)
#define PH_COUNT(S, k)                                                                      \
    do {                                                                                    \
        if (threadIdx.x == 0) atomicAdd((unsigned long long*)&(S).st->phase[k], 1ull);     \
    } while (0)
#define PH_RESET(v) v = clock64()
#define PH_ADDN(S, k, n)                                                                                  \
    do {                                                                                                  \
        if (threadIdx.x == 0) atomicAdd((unsigned long long*)&(S).st->phase[k], (unsigned long long)(n)); \
    } while (0)
// a per-lane count, summed over the wavefront where it is added
#define PH_LANE_DECL(v) int v = 0
#define PH_LANE_INC(v) ++v
#define PH_LANE_ADD(S, k, v)                                            \
    do {                                                                \
        int ph_s_ = v;                                                  \
        for (int ph_o_ = 32; ph_o_ > 0; ph_o_ >>= 1) ph_s_ += __shfl_xor(ph_s_, ph_o_, 64); \
        PH_ADDN(S, k, ph_s_);                                           \
    } while (0)
#else
#define PH_BEGIN(v) \
    do {            \
    } while (0)
#define PH_END(S, k, v) \
    do {                \
    } while (0)
#define PH_COUNT(S, k) \
    do {               \
    } while (0)
#define PH_RESET(v) \
    do {            \
    } while (0)
#define PH_ADDN(S, k, n) \
    do {                 \
    } while (0)
#define PH_LANE_DECL(v) \
    do {                \
    } while (0)
#define PH_LANE_INC(v) \
    do {               \
    } while (0)
#define PH_LANE_ADD(S, k, v) \
    do {                     \
    } while (0)
#endif

// Global-memory access through the generic pointers of SysDev, as global_*
// operations (a flat operation would also hold back every LDS wait).
#define GLB __attribute__((address_space(1)))
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <class T>
__device__ inline T gld(const T* p) {
    static_assert(sizeof(T) % 16 == 0, "16-byte multiple");
    u32x4 t[sizeof(T) / 16];
    const GLB u32x4* q = (const GLB u32x4*)p;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 16); ++k) t[k] = q[k];
    T r;
    __builtin_memcpy(&r, t, sizeof(T));
    return r;
}
template <class T>
__device__ inline void gst(T* p, const T& v) {
    static_assert(sizeof(T) % 16 == 0, "16-byte multiple");
    u32x4 t[sizeof(T) / 16];
    __builtin_memcpy(t, &v, sizeof(T));
    GLB u32x4* q = (GLB u32x4*)p;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 16); ++k) q[k] = t[k];
}
template <class T>
__device__ inline T gl1(const T* p) {
    return *(const GLB T*)p;
}
template <class T>
__device__ inline void gs1(T* p, T v) {
    *(GLB T*)p = v;
}
__device__ inline int2 gl2(const int2* p) {
    const unsigned long long v = gl1(reinterpret_cast<const unsigned long long*>(p));
    return make_int2((int)(unsigned)v, (int)(unsigned)(v >> 32));
}
__device__ inline void gs2(int2* p, int2 v) {
    gs1(reinterpret_cast<unsigned long long*>(p), (unsigned long long)(unsigned)v.x | ((unsigned long long)(unsigned)v.y << 32));
}

// The workgroup's dynamic LDS.  Views into it are kept as offsets and
// turned into pointers from this array where they are used, so the compiler
// keeps LDS accesses as ds_* operations (a pointer stored in a struct and
// passed on would become a flat pointer).
extern __shared__ __attribute__((aligned(16))) unsigned char smem[];

// the scaled tables sit at the start of the LDS, the system's SysDev right
// after them (field reads become LDS reads in every function)
constexpr int kSdOff = (int)((sizeof(DevTables) + 15) / 16 * 16);
__device__ inline const DevTables& tabs() { return *reinterpret_cast<const DevTables*>(smem); }
__device__ inline const SysDev& sdev() { return *reinterpret_cast<const SysDev*>(smem + kSdOff); }

// ------------------------------------------------------------------------
// project RNG (counter-based splitmix64) and a logarithm from basic ops
// ------------------------------------------------------------------------
__device__ inline uint64_t sm64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ inline double rng_u01(uint64_t seed, int stream, uint64_t ctr) {
    const uint64_t key = sm64(seed ^ (0x5851F42D4C957F2Dull * (uint64_t)(stream + 1)));
    return (double)(sm64(key + ctr) >> 11) * (1.0 / 9007199254740992.0);
}
// log x = e ln2 + 2 atanh((m-1)/(m+1)), m in [1/sqrt2, sqrt2]; identical
// bits to the CPU oracle's restatement
__device__ double dlog(double x) {
    uint64_t b = (uint64_t)__double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7ff);
    int adj = 0;
    if (e == 0) {
        x = x * 18014398509481984.0;
        b = (uint64_t)__double_as_longlong(x);
        e = (int)((b >> 52) & 0x7ff);
        adj = -54;
    }
    e = e - 1023 + adj;
    b = (b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m = __longlong_as_double((long long)b);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double p = 2.0 / 25.0;
    p = p * z + 2.0 / 23.0;
    p = p * z + 2.0 / 21.0;
    p = p * z + 2.0 / 19.0;
    p = p * z + 2.0 / 17.0;
    p = p * z + 2.0 / 15.0;
    p = p * z + 2.0 / 13.0;
    p = p * z + 2.0 / 11.0;
    p = p * z + 2.0 / 9.0;
    p = p * z + 2.0 / 7.0;
    p = p * z + 2.0 / 5.0;
    p = p * z + 2.0 / 3.0;
    const double lm = 2.0 * s + (s * z) * p;
    const double de = (double)e;
    return de * 6.93147180369123816490e-01 + (de * 1.90821492927058770002e-10 + lm);
}

// ------------------------------------------------------------------------
// pair codes: closed form of make_code.f from two topo words, overrides
// ------------------------------------------------------------------------
__device__ inline int static_t0(int ta, int tb) {
    int ka = topo_kind(ta), ra = topo_res(ta), kb = topo_kind(tb), rb = topo_res(tb);
    if (topo_chain(ta) == topo_chain(tb)) {
        // -Dchaptype=2 builds no code inside a second-species chain
        // (make_code.f:349,366; its side-chain loop finds no firstside2 entry)
        if (topo_free(ta)) return 1;
        if (ka > kb || (ka == kb && ra > rb)) {
            int t = ka;
            ka = kb;
            kb = t;
            t = ra;
            ra = rb;
            rb = t;
        }
        const int d = rb - ra;
        switch (ka * 4 + kb) {
            case 0: return d == 1 ? 9 : 1;
            case 1: return d == 0 ? 4 : (d == 1 ? 7 : (d == -1 ? 18 : 1));
            case 2: return d == 0 ? 5 : (d == -1 ? 8 : (d == 1 ? 17 : 1));
            case 3: return d == 0 ? 10 : (d == 1 ? 24 : (d == -1 ? 25 : 1));
            case 5: return d == 1 ? 20 : 1;
            case 6:
                if (d == -1) return 6;
                if (d == 0) return 8;
                if (d == -2) return 19;
                return (d >= -3 && d <= 3) ? 1 : 15;
            case 7: return d == 0 ? 11 : (d == -1 ? 23 : 1);
            case 10: return d == 1 ? 21 : 1;
            case 11: return d == 0 ? 12 : (d == 1 ? 22 : (d == 2 ? 26 : 1));
            case 15: return (d >= 4 && topo_hp(ta) && topo_hp(tb)) ? 16 : 1;
            default: return 1;
        }
    }
    // -Dno_h zeroes hp only where make_code.f:80-104 reads it (different
    // chains); the same-chain loops read hp1/hp2 (make_code.f:174-186)
    // and none between two second-species chains (make_code.f:519-547)
    if (ka == 3 && kb == 3) return (topo_hp(ta) && topo_hp(tb) && !topo_noh(ta) && !(topo_free(ta) && topo_free(tb))) ? 16 : 1;
    if (ka == 1 && kb == 2) return (topo_pro(ta) && topo_chain(ta) < topo_chain(tb)) ? 1 : 15;
    if (ka == 2 && kb == 1) return (topo_pro(tb) && topo_chain(tb) < topo_chain(ta)) ? 1 : 15;
    return 1;
}
// under -Dno_hbs make_code.f skips every code-15 assignment (:114-170,
// :192-208, :348-365): those pairs keep the hard-core code 1
__device__ inline int static_t(int ta, int tb) {
    const int r = static_t0(ta, tb);
    return (r == 15 && topo_nohbs(ta)) ? 1 : r;
}
// Where nbor.f finds the pair {a, b}, as one 64-bit key: the cell it scans
// (cell_add.f's linear index, cells in ascending order), the place of that
// cell's own bead in its list, the other cell's place in the scan
// (cell_link.f's half stencil; 0 for the own cell), the other bead's place.
// A bead's neighbour rows in nbor.f's order are its rows sorted by this key.
// a and b: the beads' fine cells (10 bits per axis) and list places.
__device__ inline unsigned long long nbor_key(int M, const DevTables& T, int2 a, int2 b) {
    const int fa = a.x, fb = b.x, pa = a.y, pb = b.y, nc = M + 4;
    int d[3];
    for (int q = 0; q < 3; ++q) {
        // periodic offset in [-(M / 2), M - M / 2), in [-2, 2] for neighbours
        int t = ((fb >> (10 * q)) & 1023) - ((fa >> (10 * q)) & 1023);
        t += t < 0 ? M : 0;
        d[q] = t < M - M / 2 ? t : t - M;
    }
    auto lin = [&](int f) {
        return 1ull + (unsigned long long)((f & 1023) + 2) + (unsigned long long)(((f >> 10) & 1023) + 2) * nc +
               (unsigned long long)((f >> 20) + 2) * nc * nc;
    };
    if (d[0] == 0 && d[1] == 0 && d[2] == 0)
        return (lin(fa) << 26) | ((unsigned long long)min(pa, pb) << 16) | (unsigned long long)max(pa, pb);
    const int h = T.halfst[(d[0] + 2) + 5 * (d[1] + 2) + 25 * (d[2] + 2)];
    if (h > 0) return (lin(fa) << 26) | ((unsigned long long)pa << 16) | ((unsigned long long)h << 10) | (unsigned)pb;
    const int g = T.halfst[(2 - d[0]) + 5 * (2 - d[1]) + 25 * (2 - d[2])];
    return (lin(fb) << 26) | ((unsigned long long)pb << 16) | ((unsigned long long)g << 10) | (unsigned)pa;
}

// ev_code(row, col) with row's override slots o
__device__ inline int dyn_code(int4 o, int col, int stat) {
    if (o.x >= 0 && (o.x >> 8) == col) return o.x & 0xff;
    if (o.y >= 0 && (o.y >> 8) == col) return o.y & 0xff;
    if (o.z >= 0 && (o.z >> 8) == col) return o.z & 0xff;
    if (o.w >= 0 && (o.w >> 8) == col) return o.w & 0xff;
    return stat;
}
// single-lane writer of ev_code(row, col) = code (keeps an override only
// where it differs from the static code)
__device__ void set_code(const SysDev&, int row, int col, int code, int& status) {
    const SysDev& S = sdev();

    int4 o = gld(S.ovr + (row));
    const int stat = static_t(gl1(&S.rec[row].topo), gl1(&S.rec[col].topo));
    if (o.x >= 0 && (o.x >> 8) == col) o.x = -1;
    if (o.y >= 0 && (o.y >> 8) == col) o.y = -1;
    if (o.z >= 0 && (o.z >> 8) == col) o.z = -1;
    if (o.w >= 0 && (o.w >> 8) == col) o.w = -1;
    if (code != stat) {
        const int v = (col << 8) | code;
        if (o.x < 0) o.x = v;
        else if (o.y < 0) o.y = v;
        else if (o.z < 0) o.z = v;
        else if (o.w < 0) o.w = v;
        else status = kErrOverrideSlots;
    }
    gst(S.ovr + (row), o);
}

// ------------------------------------------------------------------------
// predictors: core.f, bond.f, sqwel.f, sqshlder.f, nc_sqwel.f
// ------------------------------------------------------------------------
struct Geo {
    double rx, ry, rz, b, r2, v2;
};
__device__ inline void geo(const Rec& a, const Rec& c, double tf, Geo& g) {
    const double vx = a.vx - c.vx, vy = a.vy - c.vy, vz = a.vz - c.vz;
    double rx = a.x - c.x + vx * tf;
    double ry = a.y - c.y + vy * tf;
    double rz = a.z - c.z + vz * tf;
    rx = rx - round(rx);
    ry = ry - round(ry);
    rz = rz - round(rz);
    g.rx = rx;
    g.ry = ry;
    g.rz = rz;
    g.b = rx * vx + ry * vy + rz * vz;
    g.r2 = rx * rx + ry * ry + rz * rz;
    g.v2 = vx * vx + vy * vy + vz * vz;
}
__device__ inline void bond_window(const DevTables& T, int code, double2 bw, double& blmin, double& blmax) {
    if (code >= 10) {
        blmin = bw.x;
        blmax = bw.y;
    } else {
        blmin = T.ev_param[2][code];
        blmax = T.ev_param[3][code];
    }
}
// collision time of pair (a, c) with ev_code `code` at false time tf; tij
// keeps its preset (1e9) and type its value when no event is predicted.
// Every predictor of core.f, bond.f, sqwel.f, sqshlder.f and nc_sqwel.f has
// the same shape: an inner radius^2 A and an outer radius^2 B with
//   dA = b^2 - v^2 (r^2 - A),  dB = b^2 - v^2 (r^2 - B)
// and one of four roots: approach to A, (-b - sqrt dA)/v^2; exit through B,
// (-b + sqrt dB)/v^2; entry through B, (-b - sqrt dB)/v^2; and the stable
// exit form of bond.f:63, -(r^2 - B)/(sqrt dB + b).  The per-class rules
// below only choose A, B, the root and the type, so lanes of different
// classes share one square root and one division, evaluated exactly as
// each reference routine writes them.
__device__ void predict(const DevTables& T, int code, const Rec& a, const Rec& c, bool bonded, double2 bwa,
                        double tf, double& tij, int& type) {
    const int ia = dyn_ident(a.dyn), ic = dyn_ident(c.dyn);
    Geo g;
    geo(a, c, tf, g);
    const double b = g.b, r2 = g.r2, v2 = g.v2;
    const bool is_core = code <= 3 || (code >= 17 && code <= 26);
    const bool is_bond = code >= 4 && code <= 12;
    const bool is_nc = code == 15, is_well = code == 16, is_shl = code >= 40;
    if (!(is_core || is_bond || is_nc || is_well || is_shl)) {
        type = -99;  // codes 13, 14, 27..39: error in ev_code matrix
        return;
    }
    // inner and outer radii squared
    const double ssq = T.sigma_sq[ia][ic];
    double A, B = 0.0;
    if (is_core) {  // core.f:27-31
        const double pe = T.ev_param[1][code];
        A = ssq * (pe * pe);
        if (code >= 22) {
            const double q = T.sqz610[code - 21][ia > ic ? ia : ic];
            A = A * (q * q);
        }
    } else if (is_bond) {  // bond.f:19-31,70-81
        double blmin, blmax;
        bond_window(T, code, bwa, blmin, blmax);
        A = blmin * blmin;
        B = blmax * blmax;
    } else {
        // nc_sqwel.f / sqwel.f / sqshlder.f: hard core unscaled, except the
        // bonded N-C pair (core scaled by ev_param(1,15) twice)
        A = (is_nc && ia + ic != 5 && bonded) ? ssq * T.ev_param[1][15] * T.ev_param[1][15] : ssq;
        B = is_shl ? T.shlddia_sq[ia][ic] : T.welldia_sq[ia][ic];
    }
    const double dA = b * b - v2 * (r2 - A);
    const double diff = r2 - B;
    const double dB = b * b - v2 * diff;
    // root: 0 none, 1 approach A, 2 exit B, 3 enter B, 4 stable exit B
    int root = 0, ty = 0;
    const bool in = b < 0.0;
    if (is_core) {
        if (in && dA > 0.0) { root = 1; ty = 1; }
    } else if (is_bond) {
        if (in) {
            if (dA > 0.0) { root = 1; ty = 2; }
            else if (dB > 0.0) { root = 2; ty = 3; }
        } else if (dB > 0.0) { root = 4; ty = 3; }
    } else if (is_nc && ia + ic != 5) {
        if (bonded) {  // nc_sqwel.f:76-101: inside the well by construction
            if (in && dA > 0.0) { root = 1; ty = 1; }
            else { root = 2; ty = 8; }
        } else if (in && dB > 0.0) { root = 3; ty = 9; }
    } else {
        // H-bond N-C well (types 1/16/7), side-chain well (1/8/4), shoulder (1/10/12)
        const int ex = is_nc ? 16 : (is_well ? 8 : 10);
        const int en = is_nc ? 7 : (is_well ? 4 : 12);
        if (in) {
            if (diff < 0.0) {
                if (dA > 0.0) { root = 1; ty = 1; }
                else { root = 2; ty = ex; }
            } else if (dB > 0.0) { root = 3; ty = en; }
        } else if (diff < 0.0) { root = 2; ty = ex; }
    }
    if (root == 0) return;
    const double sq = sqrt(root == 1 ? dA : dB);
    double num, den = v2;
    if (root == 1 || root == 3)
        num = -b - sq;
    else if (root == 2)
        num = -b + sq;
    else {
        num = -diff;
        den = sq + b;
    }
    tij = num / den;
    type = ty;
}

// ------------------------------------------------------------------------
// wave helpers
// ------------------------------------------------------------------------
__device__ inline Rec shfl_rec(const Rec& r, int src) {
    Rec o;
    o.x = __shfl(r.x, src, kW);
    o.y = __shfl(r.y, src, kW);
    o.z = __shfl(r.z, src, kW);
    o.vx = __shfl(r.vx, src, kW);
    o.vy = __shfl(r.vy, src, kW);
    o.vz = __shfl(r.vz, src, kW);
    o.bm = __shfl(r.bm, src, kW);
    o.topo = __shfl(r.topo, src, kW);
    o.dyn = __shfl(r.dyn, src, kW);
    return o;
}
// Hand-off through LDS between the lanes of one wavefront: LDS operations
// of a wavefront complete in program order, so only the compiler's ordering
// is needed (no wait for outstanding global loads).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}
// Hand-off of global-memory and LDS writes between the lanes of one
// wavefront (the stores complete before any lane reads past this point).
// Event-level code synchronises only its own wavefront: the other
// wavefronts of a system's workgroup meet it at the round barriers.
__device__ inline void wave_sync_mem() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}
// Barrier of every wavefront of the system's workgroup, for the steps that
// all of them run together (list rebuilds, full rescans).
__device__ inline void wg_sync(int nw) {
    if (nw > 1)
        __syncthreads();
    else
        wave_sync_mem();
}

// DPP lane moves: CTRL row_shr:n = 0x110+n, row_bcast:15 = 0x142,
// row_bcast:31 = 0x143; lanes without a valid source keep their own value
template <int CTRL, int RM>
__device__ inline int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, RM, 0xf, false);
}
template <int CTRL, int RM>
__device__ inline double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<CTRL, RM>((int)b), hi = dpp_i<CTRL, RM>((int)(b >> 32));
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}
__device__ inline double readlane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}
// first non-zero status of the wave (0 if none); wave-uniform
__device__ inline int wave_status(int v) {
    const unsigned long long m = __ballot(v != 0);
    return m ? __builtin_amdgcn_readlane(v, __ffsll((long long)m) - 1) : 0;
}
__device__ inline double wave_max(double v) {
    double o;
    o = dpp_d<0x111, 0xf>(v); v = o > v ? o : v;
    o = dpp_d<0x112, 0xf>(v); v = o > v ? o : v;
    o = dpp_d<0x114, 0xf>(v); v = o > v ? o : v;
    o = dpp_d<0x118, 0xf>(v); v = o > v ? o : v;
    o = dpp_d<0x142, 0xa>(v); v = o > v ? o : v;
    o = dpp_d<0x143, 0xc>(v); v = o > v ? o : v;
    return readlane_d(v, 63);
}
__device__ inline int wave_imax(int v) {
    v = max(v, dpp_i<0x111, 0xf>(v));
    v = max(v, dpp_i<0x112, 0xf>(v));
    v = max(v, dpp_i<0x114, 0xf>(v));
    v = max(v, dpp_i<0x118, 0xf>(v));
    v = max(v, dpp_i<0x142, 0xa>(v));
    v = max(v, dpp_i<0x143, 0xc>(v));
    return __builtin_amdgcn_readlane(v, 63);
}
// min of predictions (t, position) carrying partner and type
__device__ inline void wave_min_pred(double& t, int& p, int& m, int& ty) {
    for (int s = 1; s < kW; s <<= 1) {
        const double ot = __shfl_xor(t, s, kW);
        const int op = __shfl_xor(p, s, kW);
        const int om = __shfl_xor(m, s, kW);
        const int oty = __shfl_xor(ty, s, kW);
        if (ot < t || (ot == t && op < p)) {
            t = ot;
            p = op;
            m = om;
            ty = oty;
        }
    }
}

// Calendar key of one entry: order (bin, time, later add first), payload
// partner and type.  bc = bin << 8 | (coltype + 2); ip = index << 16 |
// (partner + 2).  Equal bins and times, which only bit-identical event
// times give, are settled by the entries' add stamps: the reference's
// bucket lists put each add at the head (add_tbin.f:28-32) and its stable
// sort keeps list order among equal times (main.F90:521-545), so the entry
// added last is taken first.
struct Key {
    double t;
    int bc;
    unsigned ip;
};
constexpr int kEmptyBc = INT_MAX;
constexpr int kStampBits = 20;  // adds per event: up to N + 4 in a sync's list rebuild
__device__ inline unsigned long long mk_stamp(long long coll, int sub) {
    return ((unsigned long long)coll << kStampBits) | (unsigned long long)(unsigned)sub;
}
// stamps live in HBM and are read and written at the device's L2 (other
// wavefronts of the workgroup write them)
__device__ inline void stamp_set(int k, unsigned long long v) {
    __hip_atomic_store((GLB unsigned long long*)(sdev().stamp + k), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long stamp_get(int k) {
    return __hip_atomic_load((GLB unsigned long long*)(sdev().stamp + k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// a before b in the order above, with a's and b's stamps given (entries not
// yet written: a round's new keys)
__device__ inline bool key_less_st(const Key& a, unsigned long long sa, const Key& b, unsigned long long sb) {
    const int ab = a.bc >> 8, bb = b.bc >> 8;
    if (ab != bb) return ab < bb;
    if (a.t != b.t) return a.t < b.t;
    if (a.bc == kEmptyBc || sa == sb) return a.ip < b.ip;
    return sa > sb;
}
// a before b, both entries of the calendar (their stamps written)
__device__ inline bool key_less(const Key& a, const Key& b) {
    const int ab = a.bc >> 8, bb = b.bc >> 8;
    if (ab != bb) return ab < bb;
    if (a.t != b.t) return a.t < b.t;
    if (a.bc == kEmptyBc || a.ip == b.ip) return a.ip < b.ip;
    const unsigned long long sa = stamp_get((int)(a.ip >> 16)), sb = stamp_get((int)(b.ip >> 16));
    return sa != sb ? sa > sb : a.ip < b.ip;
}
__device__ inline Key empty_key() {
    Key e;
    e.t = 0.0;
    e.bc = kEmptyBc;
    e.ip = 0xFFFFFFFFu;
    return e;
}
__device__ inline int key_bin(const Key& k) { return k.bc >> 8; }
__device__ inline int key_ct(const Key& k) { return (k.bc & 0xff) - 2; }
__device__ inline int key_idx(const Key& k) { return (int)(k.ip >> 16); }
__device__ inline int key_p(const Key& k) { return (int)(k.ip & 0xFFFFu) - 2; }
__device__ inline Key leaf_key(const Leaf& lf, int idx) {
    const int bin = leaf_bin(lf.meta);
    if (bin < 0) return empty_key();
    Key k;
    k.t = lf.tim;
    k.bc = (bin << 8) | (leaf_ct(lf.meta) + 2);
    k.ip = ((unsigned)idx << 16) | (unsigned)(lf.nptnr + 2);
    return k;
}
template <int CTRL, int RM>
__device__ inline void key_step(Key& k) {
    Key o;
    o.t = dpp_d<CTRL, RM>(k.t);
    o.bc = dpp_i<CTRL, RM>(k.bc);
    o.ip = (unsigned)dpp_i<CTRL, RM>((int)k.ip);
    if (key_less(o, k)) k = o;
}
// exact minimum over each row of 16 lanes, in the row's last lane
__device__ inline void key_min_rows_exact(Key& k) {
    key_step<0x111, 0xf>(k);
    key_step<0x112, 0xf>(k);
    key_step<0x114, 0xf>(k);
    key_step<0x118, 0xf>(k);
}
__device__ inline Key readlane_key(const Key& k, int l) {
    Key r;
    r.t = readlane_d(k.t, l);
    r.bc = __builtin_amdgcn_readlane(k.bc, l);
    r.ip = (unsigned)__builtin_amdgcn_readlane((int)k.ip, l);
    return r;
}

// per-system register state
struct Sc {
    double tfalse, t, interval, interval_max, t_fact, n_forced, sortsize, tbin_off, old_tfalse;
    long long coll, numghosts, nupdates, nforced, target;
    unsigned long long rc0, rc1;
    int nbin, status;
    long long nev[32];  // tally by resolved coltype (main.F90:926)
};

__device__ inline int calc_bin(const Sc& c, double tim, bool uncond, int& status) {
    if (!uncond && !(tim < c.interval_max)) return -1;
    int j = (int)((tim + c.tbin_off) / c.sortsize) + 1;
    if (j > kNumbin) j = kNumbin;  // -Ddebugging clamp, add_tbin.f:20-23
    if (j < 0) {
        // a negative time already reported (events.f:59-73 stops there) keeps its status
        if (!status) status = kErrBin;
        return -1;
    }
    // inserted behind the scan position: the reference's bucket scan
    // (main.F90:496-498) never returns to it before the next rebuild
    return j < c.nbin ? -1 : j;
}

// flags: kPairReadsSp = the outcome resolution read the shoulder set or an
// H-bond partner; kPairWroteSp = the event rewrote them
constexpr int kPairReadsSp = 1, kPairWroteSp = 2;
struct Pair {  // the event's colliding pair (j < 0 for a ghost event)
    int i, j, flags, pad;
    Rec ri, rj;
    double2 bwi, bwj;
    int4 xi, xj;
};

constexpr int kMaxSpec = 8;  // full recomputes batched with the pair's own
// per-owner minimum of a prediction batch: LDS atomics on (time, position)
struct MinSlot {
    unsigned long long tkey;
    unsigned packed;  // position << 21 | (type + 1) << 16 | partner
    int pad;
};
// a down-set bead whose full recompute is batched with the pair's
struct SpecInfo {
    Rec r;
    double2 bw;
    int4 x;
    int l, nu, nx, off;
};
// a down-set bead's own data, read by the lanes that predict its single-pair
// updates inside the batch
struct DownInfo {
    Rec r;
    int4 o;
    double2 bw;
};
// time as an unsigned key with the order of the doubles (-0 as +0)
__device__ inline unsigned long long t_order(double t) {
    if (t == 0.0) t = 0.0;
    const unsigned long long b = (unsigned long long)__double_as_longlong(t);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}


// One round of a multi-wavefront system: the popped calendar entries (in
// key order, one per wavefront) and what each wavefront reports back.
struct Batch {
    Key key[kMaxWaves];   // popped entries
    Key knew[kMaxWaves];  // least key among the wavefront's new calendar entries
    unsigned long long knew_st[kMaxWaves];  // its add stamp
    int ty[kMaxWaves];    // resolved type of each event
    int nbl[kMaxWaves];   // blocks to rescan, in each wavefront's list
    int st[kMaxWaves];    // status after each event
    int wdone[kMaxWaves]; // the wavefront's committed entries are written
    int K;                // entries popped this round (0: stop)
    int cut;              // first wavefront whose event is not committed
    int rebuild;          // the round's sync event rebuilds the lists
    int gbead;            // the bead of a ghost event leading the round
    long long t0;         // start of a budgeted launch (wall_clock64)
};
// Order proxy of a calendar key: the bin in the top 12 bits, then the top
// 52 bits of the ordered time.  Proxies order keys as key_less does except
// that distinct keys may share one; the reductions below take the minimum
// proxy with 64-bit lane moves and settle shared proxies with key_less.
__device__ inline unsigned long long key_proxy(const Key& k) {
    return k.bc == kEmptyBc ? ~0ull
                            : (((unsigned long long)(unsigned)(k.bc >> 8) << 52) | (t_order(k.t) >> 12));
}
template <int CTRL, int RM>
__device__ inline unsigned long long dpp_u64(unsigned long long v) {
    const int lo = dpp_i<CTRL, RM>((int)v), hi = dpp_i<CTRL, RM>((int)(v >> 32));
    return (unsigned long long)(unsigned)lo | ((unsigned long long)(unsigned)hi << 32);
}
template <int CTRL, int RM>
__device__ inline void umin_step(unsigned long long& v) {
    const unsigned long long o = dpp_u64<CTRL, RM>(v);
    v = o < v ? o : v;
}
// minimum over the wave, returned to every lane
__device__ inline Key key_min_wave(Key k) {
    const unsigned long long p = key_proxy(k);
    unsigned long long m = p;
    umin_step<0x111, 0xf>(m);
    umin_step<0x112, 0xf>(m);
    umin_step<0x114, 0xf>(m);
    umin_step<0x118, 0xf>(m);
    umin_step<0x142, 0xa>(m);
    umin_step<0x143, 0xc>(m);
    const unsigned long long mn = (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)m, 63) |
                                  ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(m >> 32), 63) << 32);
    if (mn == ~0ull) return empty_key();
    const unsigned long long tie = __ballot(p == mn);
    Key r = readlane_key(k, __ffsll((long long)tie) - 1);
    unsigned long long rest = tie & (tie - 1);
    while (rest) {
        const Key o = readlane_key(k, __ffsll((long long)rest) - 1);
        rest &= rest - 1;
        if (key_less(o, r)) r = o;
    }
    return r;
}
// minimum over the wave of keys with their add stamps given per lane (new
// entries, not yet written), returned to every lane with its stamp
__device__ inline Key key_min_wave_st(Key k, unsigned long long st, unsigned long long& out_st) {
    const unsigned long long p = key_proxy(k);
    unsigned long long m = p;
    umin_step<0x111, 0xf>(m);
    umin_step<0x112, 0xf>(m);
    umin_step<0x114, 0xf>(m);
    umin_step<0x118, 0xf>(m);
    umin_step<0x142, 0xa>(m);
    umin_step<0x143, 0xc>(m);
    const unsigned long long mn = (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)m, 63) |
                                  ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(m >> 32), 63) << 32);
    out_st = 0;
    if (mn == ~0ull) return empty_key();
    const unsigned long long tie = __ballot(p == mn);
    int src = __ffsll((long long)tie) - 1;
    Key r = readlane_key(k, src);
    unsigned long long rs = (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)st, src) |
                            ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(st >> 32), src) << 32);
    unsigned long long rest = tie & (tie - 1);
    while (rest) {
        src = __ffsll((long long)rest) - 1;
        rest &= rest - 1;
        const Key o = readlane_key(k, src);
        const unsigned long long os = (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)st, src) |
                                      ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(st >> 32), src) << 32);
        if (key_less_st(o, os, r, rs)) {
            r = o;
            rs = os;
        }
    }
    out_st = rs;
    return r;
}
// Minimum over each row of 16 lanes.  Returns true on the lane that holds
// its row's minimum in k afterwards (the row's last lane when the row is
// empty or when shared proxies need the exact scan).
__device__ inline bool key_min_rows(Key& k) {
    const int lane = lane_id();
    const unsigned long long p = key_proxy(k);
    unsigned long long m = p;
    umin_step<0x111, 0xf>(m);
    umin_step<0x112, 0xf>(m);
    umin_step<0x114, 0xf>(m);
    umin_step<0x118, 0xf>(m);
    const unsigned long long rm = dpp_u64<0x15F, 0xf>(m);  // the row's last lane, to the whole row
    const unsigned long long tie = __ballot(p == rm && rm != ~0ull);
    const unsigned rb = (unsigned)(tie >> (lane & 48)) & 0xFFFFu;
    if (__ballot(__popc(rb) > 1)) {
        key_min_rows_exact(k);
        return (lane & 15) == 15;
    }
    if (rm == ~0ull) {
        k = empty_key();
        return (lane & 15) == 15;
    }
    return p == rm;
}

// LDS views, one set per wavefront: the calendar tree, the master scalars
// and the round record are shared, the work lists are the wavefront's own
struct Lds {
    int o_blk, o_sup, o_wm, o_wx, o_blist, o_pair, o_sc, o_leaf, o_wd, o_wt, o_wy, o_slot, o_spec, o_down;
    int o_msc, o_batch;
    Leaf* gleaf;  // calendar leaves in HBM when they do not fit in LDS
    // the round's footprint marks: a flat pointer into LDS (behind the
    // leaves, when both fit there) or to SysDev::mark
    unsigned* mark;
    int nblk, nsup, maxw, nw;
    bool leaf_lds, mark_lds;
    __device__ Key* blk() const { return reinterpret_cast<Key*>(smem + o_blk); }
    __device__ Key* sup() const { return reinterpret_cast<Key*>(smem + o_sup); }
    __device__ int* wm() const { return reinterpret_cast<int*>(smem + o_wm); }  // work list: partner / bead
    __device__ int* wx() const { return reinterpret_cast<int*>(smem + o_wx); }  // work list: owner / op tag
    __device__ int* wd() const { return reinterpret_cast<int*>(smem + o_wd); }
    __device__ double* wt() const { return reinterpret_cast<double*>(smem + o_wt); }  // predicted times
    __device__ int* wy() const { return reinterpret_cast<int*>(smem + o_wy); }  // (type + 1) << 16 | partner
    __device__ int* blist() const { return reinterpret_cast<int*>(smem + o_blist); }  // blocks to rescan
    __device__ MinSlot* slot() const { return reinterpret_cast<MinSlot*>(smem + o_slot); }
    __device__ SpecInfo* spec() const { return reinterpret_cast<SpecInfo*>(smem + o_spec); }
    __device__ DownInfo* down() const { return reinterpret_cast<DownInfo*>(smem + o_down); }
    __device__ Pair* pair() const { return reinterpret_cast<Pair*>(smem + o_pair); }
    __device__ Sc* sc() const { return reinterpret_cast<Sc*>(smem + o_sc); }
    __device__ Sc* msc() const { return reinterpret_cast<Sc*>(smem + o_msc); }
    __device__ Batch* batch() const { return reinterpret_cast<Batch*>(smem + o_batch); }
    __device__ Leaf* lleaf() const { return reinterpret_cast<Leaf*>(smem + o_leaf); }
};
// the Lds views themselves (one per wavefront), after the SysDev copy
constexpr int kLdOff = kSdOff + (int)((sizeof(SysDev) + 15) / 16 * 16);
constexpr int kLdStride = (int)((sizeof(Lds) + 15) / 16 * 16);
__device__ inline const Lds& ldsv() { return *reinterpret_cast<const Lds*>(smem + kLdOff + wave_id() * kLdStride); }
template <class T>
__device__ inline T* sgpr_ptr(T* p) {
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v), hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (T*)(((unsigned long long)hi << 32) | lo);
}
// The wavefront's views as a value in scalar registers, for functions with
// many wave fences (after each one the compiler would read a view's offset
// from LDS again before using it).
__device__ inline Lds lds_regs() {
    const Lds& m = ldsv();
    Lds r;
    r.o_blk = __builtin_amdgcn_readfirstlane(m.o_blk);
    r.o_sup = __builtin_amdgcn_readfirstlane(m.o_sup);
    r.o_wm = __builtin_amdgcn_readfirstlane(m.o_wm);
    r.o_wx = __builtin_amdgcn_readfirstlane(m.o_wx);
    r.o_blist = __builtin_amdgcn_readfirstlane(m.o_blist);
    r.o_pair = __builtin_amdgcn_readfirstlane(m.o_pair);
    r.o_sc = __builtin_amdgcn_readfirstlane(m.o_sc);
    r.o_leaf = __builtin_amdgcn_readfirstlane(m.o_leaf);
    r.o_wd = __builtin_amdgcn_readfirstlane(m.o_wd);
    r.o_wt = __builtin_amdgcn_readfirstlane(m.o_wt);
    r.o_wy = __builtin_amdgcn_readfirstlane(m.o_wy);
    r.o_slot = __builtin_amdgcn_readfirstlane(m.o_slot);
    r.o_spec = __builtin_amdgcn_readfirstlane(m.o_spec);
    r.o_down = __builtin_amdgcn_readfirstlane(m.o_down);
    r.o_msc = __builtin_amdgcn_readfirstlane(m.o_msc);
    r.o_batch = __builtin_amdgcn_readfirstlane(m.o_batch);
    r.gleaf = sgpr_ptr(m.gleaf);
    r.nblk = __builtin_amdgcn_readfirstlane(m.nblk);
    r.nsup = __builtin_amdgcn_readfirstlane(m.nsup);
    r.maxw = __builtin_amdgcn_readfirstlane(m.maxw);
    r.nw = __builtin_amdgcn_readfirstlane(m.nw);
    r.leaf_lds = __builtin_amdgcn_readfirstlane(m.leaf_lds ? 1 : 0) != 0;
    r.mark = sgpr_ptr(m.mark);
    r.mark_lds = __builtin_amdgcn_readfirstlane(m.mark_lds ? 1 : 0) != 0;
    return r;
}
__device__ inline Leaf leaf_get(const Lds& L, int k) { return L.leaf_lds ? L.lleaf()[k] : gld(L.gleaf + k); }
// a leaf another wavefront of the workgroup has just written (its HBM copy
// read past the vector L1)
__device__ inline Leaf leaf_get_fresh(const Lds& L, int k) {
    if (L.leaf_lds) return L.lleaf()[k];
    unsigned long long* q = reinterpret_cast<unsigned long long*>(L.gleaf + k);
    const unsigned long long a = __hip_atomic_load((GLB unsigned long long*)q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long b = __hip_atomic_load((GLB unsigned long long*)(q + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    Leaf r;
    r.tim = __longlong_as_double((long long)a);
    r.nptnr = (int)(unsigned)b;
    r.meta = (int)(unsigned)(b >> 32);
    return r;
}
__device__ inline void leaf_set(const Lds& L, int k, const Leaf& v) {
    if (L.leaf_lds)
        L.lleaf()[k] = v;
    else
        gst(L.gleaf + k, v);
}

// block minima of the listed blocks (4 blocks per pass, one row of 16 lanes
// each); a block listed twice gets the same value twice
__device__ void rescan_blocks(const Lds&, int N3, const int* list, int nb) {
    const Lds L = lds_regs();

    const int lane = lane_id(), k = lane & 15;
    // 16 blocks per step: the leaf loads of four passes are in flight together
    for (int base = 0; base < nb; base += 16) {
        Key e[4];
        int bq[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = base + 4 * u + (lane >> 4);
            bq[u] = q < nb ? list[q] : -1;
            const int idx = bq[u] * kBlk + k;
            e[u] = empty_key();
            if (bq[u] >= 0 && idx < N3) e[u] = leaf_key(leaf_get(L, idx), idx);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (base + 4 * u < nb && key_min_rows(e[u]) && bq[u] >= 0) L.blk()[bq[u]] = e[u];
    }
}
// superblock minima over the superblocks of the listed blocks (at most 64),
// each reduced once
__device__ void rescan_sups(const Lds&, const int* list, int nb) {
    const Lds L = lds_regs();

    const int lane = lane_id();
    unsigned long long touched = 0;
    for (int base = 0; base < nb; base += kW) {
        const int q = base + lane;
        const int sq = q < nb ? list[q] / kSup : 0;
        const int nr = min(nb - base, kW);
        for (int r = 0; r < nr; ++r) touched |= 1ull << __builtin_amdgcn_readlane(sq, r);
    }
    while (touched) {
        const int s = __ffsll((long long)touched) - 1;
        touched &= touched - 1;
        const int b = s * kSup + lane;
        const Key e = key_min_wave(b < L.nblk ? L.blk()[b] : empty_key());
        if (lane == 0) L.sup()[s] = e;
    }
}
// both levels for the wavefront's block list
__device__ void rescan_block_list(const Lds&, int N3, int nb) {
    const Lds& L = ldsv();

    rescan_blocks(L, N3, L.blist(), nb);
    wave_sync();
    rescan_sups(L, L.blist(), nb);
    wave_sync();
}
// every block and superblock minimum; all wavefronts of the system take part
__device__ void rescan_all(const Lds&, int N3) {
    const Lds& L = ldsv();

    const int lane = lane_id(), wave = wave_id(), nw = L.nw;
    for (int base = 4 * wave; base < L.nblk; base += 4 * nw) {
        const int b = base + (lane >> 4), k = lane & 15;
        Key e = empty_key();
        const int idx = b * kBlk + k;
        if (b < L.nblk && idx < N3) e = leaf_key(leaf_get(L, idx), idx);
        if (key_min_rows(e) && b < L.nblk) L.blk()[b] = e;
    }
    wg_sync(nw);
    for (int s = wave; s < L.nsup; s += nw) {
        const int b = s * kSup + lane;
        const Key e = key_min_wave(b < L.nblk ? L.blk()[b] : empty_key());
        if (lane == 0) L.sup()[s] = e;
    }
    wg_sync(nw);
}

// ------------------------------------------------------------------------
// multi-wavefront rounds: footprint marks and the accepted prefix
// ------------------------------------------------------------------------
// shoulder beads of an H-bond pair (repuls_add.f:16-25): N bead i, C bead j
__device__ inline void shoulder_idx(int i, int Li, int j, int Lj, int& ncim1, int& ncai, int& ncaj, int& nnjp1) {
    ncim1 = i + Li - 1;
    ncai = i - Li;
    ncaj = j - 2 * Lj;
    nnjp1 = j - Lj + 1;
}

// Lane's bead of the shoulder set an event on (i, j) may read or rewrite
// (main.F90:1487-1634 via repuls_check/_3, repuls_add/_del_a/_del_b and the
// H-bond partner's identity): the H-bond partners of i and j (lanes 0, 1)
// and the shoulder beads of (i,j), (j,i), (i,hbp(i)), (hbp(i),i),
// (j,hbp(j)), (hbp(j),j) (lanes 2..25); -1 where there is none
__device__ inline int shoulder_bead(const SysDev&, const Pair& P, int lane) {
    const SysDev& S = sdev();
    const int hi = P.xi.w, hj = P.j >= 0 ? P.xj.w : -1;
    if (lane == 0) return hi;
    if (lane == 1) return hj;
    if (lane >= 26 || P.j < 0) return -1;
    const int g = (lane - 2) >> 2, q = (lane - 2) & 3;
    const int x = g < 2 ? P.i : (g < 4 ? P.i : P.j);
    const int y = g < 2 ? P.j : (g < 4 ? hi : hj);
    if (y < 0) return -1;
    const int tx = x == P.i ? P.ri.topo : P.rj.topo;
    const int ty = y == P.i ? P.ri.topo : (y == P.j ? P.rj.topo : gl1(&S.rec[y].topo));
    const bool swap = (g & 1) != 0;
    const int n = swap ? y : x, cb = swap ? x : y;
    const int Ln = topo_len(swap ? ty : tx), Lc = topo_len(swap ? tx : ty);
    int a[4];
    shoulder_idx(n, Ln, cb, Lc, a[0], a[1], a[2], a[3]);
    const int v = q == 0 ? a[0] : (q == 1 ? a[1] : (q == 2 ? a[2] : a[3]));
    return (v >= 0 && v < S.N) ? v : -1;
}
// the marks live in HBM and are read and written at the device's L2
// Footprint marks of bead b through Lds::mark (flat: LDS when the system's
// leaves and marks fit there, else SysDev::mark in HBM).
__device__ inline void mark_or(const Lds& L, const SysDev&, int b, unsigned bits) {
    __hip_atomic_fetch_or(L.mark + b, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned mark_get(const Lds& L, const SysDev&, int b) {
    return __hip_atomic_load(L.mark + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void mark_clear(const Lds& L, const SysDev&, int b) {
    __hip_atomic_store(L.mark + b, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// A wavefront's place in a multi-wavefront round: -1 alone; 0 the round's
// first event (always committed, writes as it goes); k >= 1 a speculative
// event whose writes wait for the round's checks.  In a round every
// wavefront passes the kRoundBars round barriers once (footprints marked,
// new calendar keys known), whichever way it leaves.
constexpr int kRoundBars = 2;
__device__ inline void bar_to(int& nb, int n) {
    while (nb < n) {
        __syncthreads();
        ++nb;
    }
}
__device__ inline int spec_abort(const Lds& L, int bk, int& nb) {
    if (lane_id() == 0) atomicMin(&L.batch()->cut, bk);
    bar_to(nb, kRoundBars);
    return -1;
}

__device__ inline void round_clear_marks(const Lds& L, const SysDev& S, int dent, int i, int j, int sp) {
    const int lane = lane_id();
    if (dent >= 0) mark_clear(L, S, dent);
    if (lane == 0) mark_clear(L, S, i);
    if (lane == 1 && j >= 0) mark_clear(L, S, j);
    if (sp >= 0) mark_clear(L, S, sp);
}
// Events of a round committed: the prefix before the first cut-off event,
// before the first event after an error, and before the first event whose
// key is not below every new calendar key of the events before it (the
// serial order would have run one of those first).
__device__ inline int batch_accepted(const Batch& B) {
    int A = min(B.K, B.cut);
    Key lo = B.knew[0];
    unsigned long long lo_st = B.knew_st[0];
    if (B.st[0] != 0) A = min(A, 1);
    for (int k = 1; k < A; ++k) {
        const Key kk = B.key[k];
        // kk was added before the round began, lo during it: on equal bin
        // and time lo comes first (stamp 0 is older than every add)
        if (!key_less_st(kk, 0ull, lo, lo_st)) return k;
        if (B.st[k] != 0) return k + 1;
        if (key_less_st(B.knew[k], B.knew_st[k], lo, lo_st)) {
            lo = B.knew[k];
            lo_st = B.knew_st[k];
        }
    }
    return A;
}

// ------------------------------------------------------------------------
// rescheduling after an event: partial_events.f:16-201
// ------------------------------------------------------------------------
// The list rows of i and j (lane k: slot k) with their lengths; they do not
// change between list rebuilds, so a pair event loads them before it
// resolves its outcome
struct Rows {
    int eui, edi, euj, edj;
    int nui, ndi, nuj, ndj;
};
__device__ inline Rows load_rows(const SysDev& S, int i, int j) {
    const int lane = lane_id(), cap = S.cap;
    const int jr = j >= 0 ? j : i;
    Rows r;
    r.nui = gl1(S.nup + (i));
    r.ndi = gl1(S.ndn + (i));
    r.nuj = j >= 0 ? gl1(S.nup + (j)) : 0;
    r.ndj = j >= 0 ? gl1(S.ndn + (j)) : 0;
    r.eui = lane < cap ? gl1(S.nb_up + ((size_t)i * cap + lane)) : 0;
    r.edi = lane < cap ? gl1(S.nb_dn + ((size_t)i * cap + lane)) : 0;
    r.euj = lane < cap ? gl1(S.nb_up + ((size_t)jr * cap + lane)) : 0;
    r.edj = lane < cap ? gl1(S.nb_dn + ((size_t)jr * cap + lane)) : 0;
    return r;
}
__device__ inline Rec load_rec(const SysDev&, const Pair& P, int m) {
    const SysDev& S = sdev();

    if (m == P.i) return P.ri;
    if (m == P.j) return P.rj;
    return gld(S.rec + (m));
}

// Cooperative full recompute of bead l (rl, bwl, xl wave-uniform): min over
// l's up list and upper shoulder partners (partial_events.f:121-139).
__device__ void full_one(const SysDev&, const Lds&, int l, const Rec& rl, double2 bwl, int4 xl, double tf,
                         double T0, double& tim, int& np, int& ct, int& status) {
    const Lds L = lds_regs();

    const SysDev& S = sdev();

    const DevTables& T = tabs();
    const Pair& P = *L.pair();
    const int lane = lane_id(), cap = S.cap;
    const int nu = gl1(S.nup + (l));
    const int nx = (xl.x > l ? 1 : 0) + (xl.y > l ? 1 : 0) + (xl.z > l ? 1 : 0);
    const int tot = nu + nx;
    double bt = 1e300;
    int bp = INT_MAX, bm = -1, bty = -1;
    for (int base = 0; base < tot; base += kW) {
        const int p = base + lane;
        if (p < tot) {
            int m, stat;
            if (p < nu) {
                const int ent = gl1(S.nb_up + ((size_t)l * cap + p));
                m = ent & 0xFFFFFF;
                stat = ent >> 24;
            } else {
                const int k = p - nu;
                int cnt = 0;
                m = -1;
                if (xl.x > l) {
                    if (cnt == k) m = xl.x;
                    ++cnt;
                }
                if (xl.y > l) {
                    if (cnt == k) m = xl.y;
                    ++cnt;
                }
                if (xl.z > l) {
                    if (cnt == k) m = xl.z;
                    ++cnt;
                }
                stat = -1;
            }
            const Rec rm = load_rec(S, P, m);
            const int4 om = gld(S.ovr + (m));
            if (stat < 0) stat = static_t(rm.topo, rl.topo);
            const int code = dyn_code(om, l, stat);  // ev_code(m, l), eventredo_up.f:28
            double tij = 1000000000.0;
            int ty = 0;
            predict(T, code, rl, rm, dyn_bptnr(rl.dyn) == m, bwl, tf, tij, ty);
            if (ty == -99) status = kErrBadCode;
            if (tij < 0.0 && fabs(tij) > 1.0e-10) status = kErrNegativeTime;
            if (tij < bt) {
                bt = tij;
                bp = p;
                bm = m;
                bty = ty;
            }
        }
    }
    wave_min_pred(bt, bp, bm, bty);
    if (bt < T0) {
        tim = bt + tf;
        np = bm;
        ct = bty;
    } else {
        tim = T0 + tf;
        np = -1;
        ct = -1;
    }
}

// eventredo_down.f: single-pair update of (st_t, st_p, st_c) of bead l
__device__ inline void single_update(const DevTables& T, const Rec& rl, int4 ol, double2 bwl, int p, const Rec& rp,
                                     double tf, double& st_t, int& st_p, int& st_c, bool& modified, int& status) {
    const int code = dyn_code(ol, p, static_t(rl.topo, rp.topo));  // ev_code(l, p)
    double tij = 1000000000.0;
    int ty = 0;
    predict(T, code, rl, rp, dyn_bptnr(rl.dyn) == p, bwl, tf, tij, ty);
    if (ty == -99) status = kErrBadCode;
    if (tij < 0.0 && fabs(tij) > 1.0e-10) status = kErrNegativeTime;
    tij = tij + tf;
    if (tij < st_t) {
        st_t = tij;
        st_p = p;
        st_c = ty;
        modified = true;
    }
}

// Generic form for any down-set size: full recomputes one at a time.
// extra_blk: a further calendar block to rescan (the ghost entry), or -1
__device__ void reschedule_generic(const SysDev&, const DevTables&, const Lds&, Sc&, const Pair&,
                                   int extra_blk, bool xpulse_del) {
    const Lds& L = ldsv();

    const SysDev& S = sdev();

    const DevTables& T = tabs();
    Sc& c = *L.sc();
    const Pair& P = *L.pair();

    const int lane = lane_id(), cap = S.cap, N3 = S.N + 3;
    const int i = P.i, j = P.j;
    const double tf = c.tfalse;
    const double T0 = c.interval_max + kLt - tf;
    int status = 0;
    PH_BEGIN(tph);
    // ---- list rows of i and j, one entry per lane (cap <= 64)
    const int nui = gl1(S.nup + (i)), ndi = gl1(S.ndn + (i));
    const int nuj = j >= 0 ? gl1(S.nup + (j)) : 0, ndj = j >= 0 ? gl1(S.ndn + (j)) : 0;
    const int eui = lane < nui ? gl1(S.nb_up + ((size_t)i * cap + lane)) : 0;
    const int edi = lane < ndi ? gl1(S.nb_dn + ((size_t)i * cap + lane)) : 0;
    const int euj = lane < nuj ? gl1(S.nb_up + ((size_t)j * cap + lane)) : 0;
    const int edj = lane < ndj ? gl1(S.nb_dn + ((size_t)j * cap + lane)) : 0;
    // ---- work list: candidates of i (tag 0) and j (tag 1), then the down
    //      set with its op partner (tag 2: i, tag 3: j), in reference order
    const int xi0 = P.xi.x, xi1 = P.xi.y, xi2 = P.xi.z;
    const int xj0 = P.xj.x, xj1 = P.xj.y, xj2 = P.xj.z;
    int w = 0;
    if (lane < nui) {
        L.wm()[lane] = eui;
        L.wx()[lane] = 0;
    }
    w = nui;
    const int kNoStat = 0xFF << 24;
    if (xi0 > i) { if (lane == 0) { L.wm()[w] = xi0 | kNoStat; L.wx()[w] = 0; } ++w; }
    if (xi1 > i) { if (lane == 0) { L.wm()[w] = xi1 | kNoStat; L.wx()[w] = 0; } ++w; }
    if (xi2 > i) { if (lane == 0) { L.wm()[w] = xi2 | kNoStat; L.wx()[w] = 0; } ++w; }
    if (j >= 0) {
        if (lane < nuj) {
            L.wm()[w + lane] = euj;
            L.wx()[w + lane] = 1;
        }
        w += nuj;
        if (xj0 > j) { if (lane == 0) { L.wm()[w] = xj0 | kNoStat; L.wx()[w] = 1; } ++w; }
        if (xj1 > j) { if (lane == 0) { L.wm()[w] = xj1 | kNoStat; L.wx()[w] = 1; } ++w; }
        if (xj2 > j) { if (lane == 0) { L.wm()[w] = xj2 | kNoStat; L.wx()[w] = 1; } ++w; }
    }
    const int ncand = w;
    if (lane < ndi) {
        L.wm()[w + lane] = edi & 0xFFFFFF;
        L.wx()[w + lane] = 2;
    }
    w += ndi;
    if (xi0 >= 0 && xi0 < i) { if (lane == 0) { L.wm()[w] = xi0; L.wx()[w] = 2; } ++w; }
    if (xi1 >= 0 && xi1 < i) { if (lane == 0) { L.wm()[w] = xi1; L.wx()[w] = 2; } ++w; }
    if (xi2 >= 0 && xi2 < i) { if (lane == 0) { L.wm()[w] = xi2; L.wx()[w] = 2; } ++w; }
    if (j >= 0) {
        const int l = edj & 0xFFFFFF;
        const bool keep = lane < ndj && l != i;
        const unsigned long long bal = __ballot(keep);
        const int pre = __popcll(bal & ((1ull << lane) - 1ull));
        if (keep) {
            L.wm()[w + pre] = l;
            L.wx()[w + pre] = 3;
        }
        w += __popcll(bal);
        if (xj0 >= 0 && xj0 < j) { if (lane == 0) { L.wm()[w] = xj0; L.wx()[w] = 3; } ++w; }
        if (xj1 >= 0 && xj1 < j) { if (lane == 0) { L.wm()[w] = xj1; L.wx()[w] = 3; } ++w; }
        if (xj2 >= 0 && xj2 < j) { if (lane == 0) { L.wm()[w] = xj2; L.wx()[w] = 3; } ++w; }
    }
    const int nD = w - ncand;
    if (w > L.maxw) status = kErrListCapacity;
    wave_sync_mem();
    PH_END(S, kPhLoads, tph);
    // ---- candidate predictions of the full recomputes of i and j
    PH_ADDN(S, kPhNPred, min(ncand, L.maxw));
    PH_ADDN(S, kPhNPredPass, (ncand + kW - 1) / kW);
    double bt0 = 1e300, bt1 = 1e300;
    int bp0 = INT_MAX, bp1 = INT_MAX, bm0 = -1, bm1 = -1, bty0 = -1, bty1 = -1;
    for (int base = 0; base < ncand; base += kW) {
        const int q = base + lane;
        if (q < ncand && q < L.maxw) {
            const int ent = L.wm()[q];
            const int owner = L.wx()[q];
            const int m = ent & 0xFFFFFF;
            int stat = (ent >> 24) & 0xFF;
            const Rec rx = owner == 0 ? P.ri : P.rj;
            const int x = owner == 0 ? i : j;
            const Rec rm = load_rec(S, P, m);
            const int4 om = gld(S.ovr + (m));
            if (stat == 0xFF) stat = static_t(rm.topo, rx.topo);
            const int code = dyn_code(om, x, stat);  // ev_code(m, x)
            double tij = 1000000000.0;
            int ty = 0;
            predict(T, code, rx, rm, dyn_bptnr(rx.dyn) == m, owner == 0 ? P.bwi : P.bwj, tf, tij, ty);
            if (ty == -99) status = kErrBadCode;
            if (tij < 0.0 && fabs(tij) > 1.0e-10) status = kErrNegativeTime;
            if (owner == 0) {
                if (tij < bt0) {
                    bt0 = tij;
                    bp0 = q;
                    bm0 = m;
                    bty0 = ty;
                }
            } else if (tij < bt1) {
                bt1 = tij;
                bp1 = q;
                bm1 = m;
                bty1 = ty;
            }
        }
    }
    wave_min_pred(bt0, bp0, bm0, bty0);
    wave_min_pred(bt1, bp1, bm1, bty1);
    double ti = T0 + tf, tj = T0 + tf;
    int pi = -1, ci = -1, pj = -1, cj = -1;
    if (bt0 < T0) {
        ti = bt0 + tf;
        pi = bm0;
        ci = bty0;
    }
    if (bt1 < T0) {
        tj = bt1 + tf;
        pj = bm1;
        cj = bty1;
    }
    PH_END(S, kPhPred, tph);
    // ---- the down set: per unique bead, its I-ops then its J-ops
    if (lane == 0) {
        L.blist()[0] = i / kBlk;
        L.blist()[1] = j >= 0 ? j / kBlk : i / kBlk;
        L.blist()[2] = extra_blk >= 0 ? extra_blk : i / kBlk;
    }
    int nbl = 3;
    for (int base = 0; base < nD; base += kW) {
        const int q = base + lane;
        bool first = false, is_i = false;
        int l = -1, cI = 0, cJ = 0;
        if (q < nD && ncand + q < L.maxw) {
            l = L.wm()[ncand + q];
            first = true;
            for (int r = 0; r < q; ++r)
                if (L.wm()[ncand + r] == l) {
                    first = false;
                    break;
                }
            if (first)
                for (int r = 0; r < nD && ncand + r < L.maxw; ++r)
                    if (L.wm()[ncand + r] == l) {
                        if (L.wx()[ncand + r] == 2) ++cI;
                        else ++cJ;
                    }
            is_i = (l == i);
        }
        Rec rl = P.ri;
        int4 ol = make_int4(-1, -1, -1, -1), xl = make_int4(-1, -1, -1, -1);
        double2 bwl = make_double2(0.0, 0.0);
        double st_t = 0.0;
        int st_p = -1, st_c = -1;
        if (first) {
            rl = load_rec(S, P, l);
            ol = gld(S.ovr + (l));
            bwl = gld(S.bw + (l));
            xl = gld(S.extra + (l));
            if (is_i) {  // bead i, reached through extra(j,:): its fresh result
                st_t = ti;
                st_p = pi;
                st_c = ci;
            } else {
                const Leaf lf = leaf_get(L, l);
                st_t = lf.tim;
                st_p = lf.nptnr;
                st_c = leaf_ct(lf.meta);
            }
        }
        bool modified = false, pending = false;
        int rem = first ? cI + cJ : 0;
        for (int guard = 0; guard < 64 * 9; ++guard) {
            while (rem > 0 && !pending) {
                const bool opi = rem > cJ;
                const int p = opi ? i : j;
                --rem;
                if (st_p != p)
                    single_update(T, rl, ol, bwl, p, opi ? P.ri : P.rj, tf, st_t, st_p, st_c, modified, status);
                else
                    pending = true;
            }
            const unsigned long long pm = __ballot(pending);
            if (pm == 0) break;
            const int leader = __ffsll((long long)pm) - 1;
            const int lf_l = __shfl(l, leader, kW);
            const Rec rlf = shfl_rec(rl, leader);
            const double2 bwlf = make_double2(__shfl(bwl.x, leader, kW), __shfl(bwl.y, leader, kW));
            const int4 xlf = make_int4(__shfl(xl.x, leader, kW), __shfl(xl.y, leader, kW), __shfl(xl.z, leader, kW),
                                       __shfl(xl.w, leader, kW));
            double ft;
            int fp, fc;
            full_one(S, L, lf_l, rlf, bwlf, xlf, tf, T0, ft, fp, fc, status);
            if (lane == leader) {
                st_t = ft;
                st_p = fp;
                st_c = fc;
                modified = true;
                pending = false;
            }
        }
        // bead i's state lives in (ti, pi, ci) and is written with i's leaf
        const unsigned long long im = __ballot(first && is_i);
        if (im) {
            const int src = __ffsll((long long)im) - 1;
            ti = __shfl(st_t, src, kW);
            pi = __shfl(st_p, src, kW);
            ci = __shfl(st_c, src, kW);
        }
        if (is_i) modified = false;
        if (modified) {
            Leaf lf;
            lf.tim = st_t;
            lf.nptnr = st_p;
            lf.meta = make_meta(calc_bin(c, st_t, false, status), st_c);
            leaf_set(L, l, lf);
            stamp_set(l, mk_stamp(c.coll, (j >= 0 ? 3 : 4) + q));
        }
        const unsigned long long mm = __ballot(modified);
        const int pre = __popcll(mm & ((1ull << lane) - 1ull));
        if (modified && nbl + pre < kMaxBlist) L.blist()[nbl + pre] = l / kBlk;
        nbl += __popcll(mm);
        if (nbl > kMaxBlist) nbl = kMaxBlist;
    }
    PH_END(S, kPhDown, tph);
    if (lane == 0) {
        // adds in partial_events.f order: i, j, then the down set (a ghost
        // event's own entry came first)
        Leaf a;
        a.tim = ti;
        a.nptnr = pi;
        a.meta = make_meta(calc_bin(c, ti, false, status), ci);
        leaf_set(L, i, a);
        stamp_set(i, mk_stamp(c.coll, j >= 0 ? 1 : 2));
        if (j >= 0) {
            Leaf b;
            b.tim = tj;
            b.nptnr = pj;
            b.meta = make_meta(calc_bin(c, tj, false, status), cj);
            leaf_set(L, j, b);
            stamp_set(j, mk_stamp(c.coll, 2));
        }
    }
    // repuls_del_b.f once the rescheduling is done
    if (xpulse_del && lane == 0) {
        const bool iN = dyn_ident(P.ri.dyn) < dyn_ident(P.rj.dyn);
        const int a = iN ? i : j, b = iN ? j : i;
        const int La = topo_len(iN ? P.ri.topo : P.rj.topo), Lb = topo_len(iN ? P.rj.topo : P.ri.topo);
        const int ncim1 = a + La - 1, ncai = a - La, ncaj = b - 2 * Lb, nnjp1 = b - Lb + 1;
        int4 ea = gld(S.extra + (a)), eb = gld(S.extra + (b));
        ea.x = -1;
        ea.y = -1;
        ea.w = -1;
        eb.x = -1;
        eb.y = -1;
        eb.w = -1;
        gst(S.extra + (a), ea);
        gst(S.extra + (b), eb);
        gs1(&S.extra[ncaj].z, -1);
        gs1(&S.extra[nnjp1].z, -1);
        gs1(&S.extra[ncai].z, -1);
        gs1(&S.extra[ncim1].z, -1);
    }
    status = wave_status(status);
    if (status && !c.status) c.status = status;
    if (!L.leaf_lds)
        wave_sync_mem();
    else
        wave_sync();
    PH_END(S, kPhLeaf, tph);
    rescan_block_list(L, N3, nbl);
    PH_END(S, kPhRescan, tph);
}

// partial_events.f:16-201 after an event of pair (i, j) (j < 0: ghost of i).
// One batch of predictions covers the full recomputes of i and j and of the
// down-set beads whose calendar partner is i or j (speculatively: the
// recompute happens only if their op sequence reaches it, and is then the
// same value); each owner's minimum over (time, list position) comes out of
// LDS atomics.  The remaining single-pair updates run one down bead per
// lane.  Three dependent rounds of global loads: the rows of i and j; the
// down beads, the rows of the batched beads and the candidates of i and j;
// the partners of the batched beads.  Down sets above one wavefront take
// the generic path.  extra_blk: a further calendar block to rescan (the
// ghost entry), or -1
__device__ void reschedule(const SysDev&, const DevTables&, const Lds&, Sc&, const Pair&, int extra_blk,
                           bool xpulse_del, int bk, const Rows& rw) {
    const Lds L = lds_regs();

    const SysDev& S = sdev();

    const DevTables& T = tabs();
    Sc& c = *L.sc();
    const Pair& P = *L.pair();
    // the wavefront's work-area views, held in scalar registers: the wave
    // fences below would make the compiler read them from LDS again
    int* const v_wm = reinterpret_cast<int*>(smem + __builtin_amdgcn_readfirstlane(L.o_wm));
    int* const v_wx = reinterpret_cast<int*>(smem + __builtin_amdgcn_readfirstlane(L.o_wx));
    int* const v_wd = reinterpret_cast<int*>(smem + __builtin_amdgcn_readfirstlane(L.o_wd));
    int* const v_wy = reinterpret_cast<int*>(smem + __builtin_amdgcn_readfirstlane(L.o_wy));
    double* const v_wt = reinterpret_cast<double*>(smem + __builtin_amdgcn_readfirstlane(L.o_wt));
    MinSlot* const v_slot = reinterpret_cast<MinSlot*>(smem + __builtin_amdgcn_readfirstlane(L.o_slot));
    SpecInfo* const v_spec = reinterpret_cast<SpecInfo*>(smem + __builtin_amdgcn_readfirstlane(L.o_spec));
    DownInfo* const v_down = reinterpret_cast<DownInfo*>(smem + __builtin_amdgcn_readfirstlane(L.o_down));
    const int maxw = __builtin_amdgcn_readfirstlane(L.maxw);

    const int lane = lane_id(), cap = S.cap, N3 = S.N + 3;
    const int i = P.i, j = P.j;
    const double tf = c.tfalse;
    const double T0 = c.interval_max + kLt - tf;
    const int dbase = maxw - (2 * cap + 8);  // down-set region of the work list
    const bool batch = bk >= 0;
    // place of i's add among this event's calendar adds (a ghost event adds
    // its own entry first); j follows, then the down set in list order
    const int sub0 = j >= 0 ? 1 : 2;
    int status = 0, nb = 0;
    PH_BEGIN(tph);
    // ---- first loads: list rows of i and j, whole rows with their lengths
    const int nui = rw.nui, ndi = rw.ndi, nuj = rw.nuj, ndj = rw.ndj;
    const int eui = rw.eui, edi = rw.edi, euj = rw.euj, edj = rw.edj;
    const int xi0 = P.xi.x, xi1 = P.xi.y, xi2 = P.xi.z;
    const int xj0 = P.xj.x, xj1 = P.xj.y, xj2 = P.xj.z;
    const int kNoStat = 0xFF << 24;
    // ---- work list: candidates of i (owner 0) and j (owner 1) at 0..
    int w = 0;
    if (lane < nui) {
        v_wm[lane] = eui;
        v_wx[lane] = 0;
    }
    w = nui;
    if (xi0 > i) { if (lane == 0) { v_wm[w] = xi0 | kNoStat; v_wx[w] = 0; } ++w; }
    if (xi1 > i) { if (lane == 0) { v_wm[w] = xi1 | kNoStat; v_wx[w] = 0; } ++w; }
    if (xi2 > i) { if (lane == 0) { v_wm[w] = xi2 | kNoStat; v_wx[w] = 0; } ++w; }
    if (j >= 0) {
        if (lane < nuj) {
            v_wm[w + lane] = euj;
            v_wx[w + lane] = 1;
        }
        w += nuj;
        if (xj0 > j) { if (lane == 0) { v_wm[w] = xj0 | kNoStat; v_wx[w] = 1; } ++w; }
        if (xj1 > j) { if (lane == 0) { v_wm[w] = xj1 | kNoStat; v_wx[w] = 1; } ++w; }
        if (xj2 > j) { if (lane == 0) { v_wm[w] = xj2 | kNoStat; v_wx[w] = 1; } ++w; }
    }
    const int ncand = w;
    // ---- the down set at dbase.., tag 2 (op with i) / 3 (op with j), in
    //      reference order: i's down list and lower shoulder partners, then j's
    int d = 0;
    if (lane < ndi) {
        v_wm[dbase + lane] = edi & 0xFFFFFF;
        v_wx[dbase + lane] = 2;
    }
    d = ndi;
    if (xi0 >= 0 && xi0 < i) { if (lane == 0) { v_wm[dbase + d] = xi0; v_wx[dbase + d] = 2; } ++d; }
    if (xi1 >= 0 && xi1 < i) { if (lane == 0) { v_wm[dbase + d] = xi1; v_wx[dbase + d] = 2; } ++d; }
    if (xi2 >= 0 && xi2 < i) { if (lane == 0) { v_wm[dbase + d] = xi2; v_wx[dbase + d] = 2; } ++d; }
    if (j >= 0) {
        const int l = edj & 0xFFFFFF;
        const bool keep = lane < ndj && l != i;
        const unsigned long long bal = __ballot(keep);
        const int pre = __popcll(bal & ((1ull << lane) - 1ull));
        if (keep) {
            v_wm[dbase + d + pre] = l;
            v_wx[dbase + d + pre] = 3;
        }
        d += __popcll(bal);
        if (xj0 >= 0 && xj0 < j) { if (lane == 0) { v_wm[dbase + d] = xj0; v_wx[dbase + d] = 3; } ++d; }
        if (xj1 >= 0 && xj1 < j) { if (lane == 0) { v_wm[dbase + d] = xj1; v_wx[dbase + d] = 3; } ++d; }
        if (xj2 >= 0 && xj2 < j) { if (lane == 0) { v_wm[dbase + d] = xj2; v_wx[dbase + d] = 3; } ++d; }
    }
    const int nD = d;
    if (nD > kW) {
        // the generic form writes as it goes: it runs only as a round's
        // first event, and no other event of the round is committed
        if (bk >= 1) {
            spec_abort(L, bk, nb);
            return;
        }
        if (batch && lane == 0) atomicMin(&L.batch()->cut, 1);
        wave_sync_mem();
        PH_COUNT(S, kPhNGeneric);
        reschedule_generic(S, T, L, c, P, extra_blk, xpulse_del);
        if (batch) {
            if (lane == 0) L.batch()->nbl[bk] = 0;
            bar_to(nb, kRoundBars);
        }
        return;
    }
    if (lane < 2 + kMaxSpec) {
        v_slot[lane].tkey = ~0ull;
        v_slot[lane].packed = ~0u;
    }
    wave_sync();
    PH_END(S, kPhRows, tph);
    // ---- one lane per down-set entry: first occurrence of its bead, the
    //      bead's op counts, and whether its full recompute is batched
    int l = -1, cI = 0, cJ = 0;
    bool first = false;
    const int dent = lane < nD ? v_wm[dbase + lane] : -1;
    {
        // every lane compares its bead with each entry, read from lanes
        const int lv = dent;
        const int tv = lane < nD ? v_wx[dbase + lane] : 0;
        l = lv;
        first = lane < nD;
        for (int r = 0; r < nD; ++r) {
            const int lr = __builtin_amdgcn_readlane(lv, r);
            const int tr = __builtin_amdgcn_readlane(tv, r);
            if (lr == lv) {
                if (r < lane) first = false;
                if (tr == 2) ++cI;
                else ++cJ;
            }
        }
        if (!first) {
            cI = 0;
            cJ = 0;
        }
    }
    const bool is_i = first && l == i;
    Leaf lf0;
    lf0.tim = 0.0;
    lf0.nptnr = -1;
    lf0.meta = 0;
    if (first && !is_i) lf0 = leaf_get(L, l);
    bool spec = first && !is_i && ((lf0.nptnr == i && cI > 0) || (lf0.nptnr == j && cJ > 0));
    const unsigned long long smask = __ballot(spec);
    const int myslot = spec ? __popcll(smask & ((1ull << lane) - 1ull)) : -1;
    if (myslot >= kMaxSpec) spec = false;
    const int nspec = min(__popcll(smask), kMaxSpec);
    if (spec) v_spec[myslot].l = l;
    wave_sync();
    PH_END(S, kPhScan, tph);
    // ---- second loads: the down beads' own data, the rows of the batched beads
    //      and the candidates of i and j, all in flight together
    Rec rl = P.ri;
    int4 ol = make_int4(-1, -1, -1, -1), xl = make_int4(-1, -1, -1, -1);
    double2 bwl = make_double2(0.0, 0.0);
    int nul = 0;
    if (first) {
        rl = load_rec(S, P, l);
        ol = gld(S.ovr + (l));
        bwl = gld(S.bw + (l));
        xl = gld(S.extra + (l));
    }
    if (spec) nul = gl1(S.nup + (l));
    int sent[kMaxSpec];
#pragma unroll
    for (int s2 = 0; s2 < kMaxSpec; ++s2) {
        sent[s2] = 0;
        if (s2 < nspec && lane < cap) sent[s2] = gl1(S.nb_up + ((size_t)v_spec[s2].l * cap + lane));
    }
    const int cent = lane < ncand ? v_wm[lane] : 0;
    const int cm = cent & 0xFFFFFF;
    Rec crm = P.ri;
    int4 com = make_int4(-1, -1, -1, -1);
    if (lane < ncand) {
        crm = load_rec(S, P, cm);
        com = gld(S.ovr + (cm));
    }
    // the down beads' own data for their single-pair predictions
    if (first && cI + cJ > 0) {
        DownInfo& di = v_down[lane];
        di.r = rl;
        di.o = ol;
        di.bw = bwl;
    }
    // the batched beads' pair entries, after the candidates
    if (spec) {
        SpecInfo& si = v_spec[myslot];
        si.r = rl;
        si.bw = bwl;
        si.x = xl;
        si.nu = nul;
        si.nx = (xl.x > l ? 1 : 0) + (xl.y > l ? 1 : 0) + (xl.z > l ? 1 : 0);
    }
    wave_sync();
    if (lane == 0) {
        int off = ncand;
        for (int s2 = 0; s2 < nspec; ++s2) {
            SpecInfo& si = v_spec[s2];
            const int n = si.nu + si.nx;
            if (off + n <= dbase) {
                si.off = off;
                off += n;
            } else
                si.off = -1;  // no room: this bead's recompute runs on its own
        }
        v_wd[0] = off;
    }
    wave_sync();
    const int npairs = v_wd[0];
#pragma unroll
    for (int s2 = 0; s2 < kMaxSpec; ++s2) {
        if (s2 < nspec) {
            const int off = v_spec[s2].off, nu = v_spec[s2].nu;
            if (off >= 0) {
                if (lane < nu) {
                    v_wm[off + lane] = sent[s2];
                    v_wx[off + lane] = 2 + s2;
                }
                if (lane == 0) {
                    const int4 x = v_spec[s2].x;
                    const int ls = v_spec[s2].l;
                    int k = off + nu;
                    if (x.x > ls) { v_wm[k] = x.x | kNoStat; v_wx[k] = 2 + s2; ++k; }
                    if (x.y > ls) { v_wm[k] = x.y | kNoStat; v_wx[k] = 2 + s2; ++k; }
                    if (x.z > ls) { v_wm[k] = x.z | kNoStat; v_wx[k] = 2 + s2; ++k; }
                }
            }
        }
    }
    int sp = -1;  // the wavefront's bead of the shoulder set, marked in a round
    bool sp_w = false;
    int rep = -1;  // earlier event whose calendar-entry write this bead's updates replay on
    // the down bead's updates may take a full recompute (a pending op)
    const bool mfull = first && !is_i && ((lf0.nptnr == i && cI > 0) || (lf0.nptnr == j && cJ > 0) || cI >= 2 || cJ >= 2);
    // single-pair predictions of the down beads (placed after the pairs)
    const bool wantI = first && cI > 0, wantJ = first && cJ > 0;
    const unsigned long long mI = __ballot(wantI), mJ = __ballot(wantJ);
    const int nI = __popcll(mI), nsing = nI + __popcll(mJ);
    const bool sing_fit = npairs + nsing <= maxw;
    // ---- a round's footprint check.  Marks per bead: bit k = wavefront k
    //      rewrites the bead's state, bit 8+k = it rewrites the bead's
    //      calendar entry, bit 16+k = that rewrite may be a full recompute.
    //      Event k commits only if no earlier event of the round rewrites a
    //      state it reads or an entry it rewrites; a later event reads the
    //      state before this one's writes (they wait for the round's last
    //      barrier), as in the serial order.  One exception: a down bead
    //      whose entry exactly one earlier event rewrote by single-pair
    //      updates only, and whose own updates here are single-pair ones,
    //      replays those updates on that event's entry once it is written.
    if (batch) {
        PH_RESET(tph);
        Batch& B = *L.batch();
        const unsigned sbit = 1u << bk, lbit = 1u << (8 + bk), fbit = 1u << (16 + bk);
        sp = shoulder_bead(S, P, lane);
        if (lane == 0) B.wdone[bk] = 0;
        if (dent >= 0) mark_or(L, S, dent, lbit | (mfull ? fbit : 0u));
        if (lane == 0) mark_or(L, S, i, sbit | lbit | fbit);
        if (lane == 1 && j >= 0) mark_or(L, S, j, sbit | lbit | fbit);
        sp_w = (P.flags & kPairWroteSp) != 0;
        const bool sp_r = (P.flags & kPairReadsSp) != 0;
        if (bk == 0 && sp_w && sp >= 0) mark_or(L, S, sp, sbit);
        bar_to(nb, 1);
        PH_END(S, kPhBMark, tph);
        if (bk >= 1) {
            const unsigned lowS = sbit - 1u, lowL = lowS << 8;
            bool bad = false;
            for (int base = 0; base < npairs; base += kW) {
                const int q = base + lane;
                if (q < npairs) {
                    const int m = (base == 0 && q < ncand ? cent : v_wm[q]) & 0xFFFFFF;
                    if (mark_get(L, S, m) & lowS) bad = true;
                }
            }
            if (dent >= 0) {
                const unsigned mk = mark_get(L, S, dent);
                if (mk & lowS) {
                    bad = true;
                } else if ((mk & lowL) && first) {
                    // (other lanes of the same bead leave it to its first)
                    const unsigned wl = (mk & lowL) >> 8;
                    const int m = __ffs(wl) - 1;
                    if (!is_i && __popc(wl) == 1 && !(mk & (1u << (16 + m))) && !mfull && sing_fit)
                        rep = m;
                    else
                        bad = true;
                }
            }
            if (lane == 0 && (mark_get(L, S, i) & (lowS | lowL))) bad = true;
            if (lane == 1 && j >= 0 && (mark_get(L, S, j) & (lowS | lowL))) bad = true;
            if (sp_r && sp >= 0 && (mark_get(L, S, sp) & lowS)) bad = true;
            // a down bead whose full recompute is not in the batch would
            // read rows that were not checked
            if (mfull && !(spec && v_spec[myslot].off >= 0)) bad = true;
            if (__ballot(bad) && lane == 0) atomicMin(&B.cut, bk);
        }
        PH_END(S, kPhBCheck, tph);
        // the cut only moves down: a wavefront already past it stops here
        // (the others learn theirs at the next barrier)
        if (bk >= B.cut) {
            bar_to(nb, kRoundBars);
            round_clear_marks(L, S, dent, i, j, bk == 0 && sp_w ? sp : -1);
            return;
        }
    }
    // ---- the single-pair predictions (l, i) and (l, j) of each down bead,
    //      after the pair entries: each is the same for every op of l with
    //      that bead, so one per (bead, partner) serves the whole op sequence
    int sqI = -1, sqJ = -1;
    if (sing_fit) {
        const unsigned long long below = (1ull << lane) - 1ull;
        if (wantI) {
            sqI = npairs + __popcll(mI & below);
            v_wm[sqI] = lane;
        }
        if (wantJ) {
            sqJ = npairs + nI + __popcll(mJ & below);
            v_wm[sqJ] = lane | 0x100;
        }
    }
    const int ntot = sing_fit ? npairs + nsing : npairs;
    wave_sync();
    PH_END(S, kPhLoads, tph);
    // ---- the prediction batch: the candidates of i and j, the batched
    //      beads' partners, then the single-pair predictions
    PH_ADDN(S, kPhNPred, ntot);
    PH_ADDN(S, kPhNPredPass, (ntot + kW - 1) / kW);
    for (int base = 0; base < ntot; base += kW) {
        const int q = base + lane;
        if (q < ntot) {
            const bool single = q >= npairs;
            const int ent = base == 0 && q < ncand ? cent : v_wm[q];
            const int owner = single ? 0 : v_wx[q];
            int m, stat, xc;
            Rec ra, rb;
            int4 oc;
            double2 bwa;
            if (single) {
                // eventredo_down.f: ev_code(l, p) with l's record first
                const DownInfo& di = v_down[ent & 0xFF];
                const bool oj = (ent & 0x100) != 0;
                m = oj ? j : i;
                ra = di.r;
                rb = oj ? P.rj : P.ri;
                oc = di.o;
                bwa = di.bw;
                xc = m;
                stat = static_t(ra.topo, rb.topo);
            } else {
                m = ent & 0xFFFFFF;
                stat = (ent >> 24) & 0xFF;
                if (owner == 0) {
                    xc = i;
                    ra = P.ri;
                    bwa = P.bwi;
                } else if (owner == 1) {
                    xc = j;
                    ra = P.rj;
                    bwa = P.bwj;
                } else {
                    const SpecInfo& si = v_spec[owner - 2];
                    xc = si.l;
                    ra = si.r;
                    bwa = si.bw;
                }
                if (base == 0 && q < ncand) {
                    rb = crm;
                    oc = com;
                } else {
                    rb = load_rec(S, P, m);
                    oc = gld(S.ovr + (m));
                }
                if (stat == 0xFF) stat = static_t(rb.topo, ra.topo);
            }
            const int code = dyn_code(oc, xc, stat);  // ev_code(m, x), or ev_code(l, p) for a single
            double tij = 1000000000.0;
            int ty = 0;
            predict(T, code, ra, rb, dyn_bptnr(ra.dyn) == m, bwa, tf, tij, ty);
            v_wt[q] = tij;
            if (single) {
                v_wy[q] = ty;  // checked where the op sequence uses it
            } else {
                if (ty == -99) status = kErrBadCode;
                if (tij < 0.0 && fabs(tij) > 1.0e-10) status = kErrNegativeTime;
                v_wy[q] = (((ty + 1) & 31) << 16) | (m & 0xFFFF);  // fields masked: position stays in bits 21..31
                atomicMin(&v_slot[owner].tkey, t_order(tij));
            }
        }
    }
    wave_sync();
    for (int base = 0; base < npairs; base += kW) {
        const int q = base + lane;
        if (q < npairs) {
            const int owner = v_wx[q];
            if (t_order(v_wt[q]) == v_slot[owner].tkey) atomicMin(&v_slot[owner].packed, ((unsigned)q << 21) | (unsigned)v_wy[q]);
        }
    }
    wave_sync();
    PH_END(S, kPhPred, tph);
    // owner results (full recompute rule: an event only below T0)
    double ti = T0 + tf, tj = T0 + tf;
    int pi = -1, ci = -1, pj = -1, cj = -1;
    {
        const unsigned a = v_slot[0].packed;
        if (a != ~0u) {
            const double bt = v_wt[a >> 21];
            if (bt < T0) {
                ti = bt + tf;
                pi = (int)(a & 0xFFFF);
                ci = (int)((a >> 16) & 31) - 1;
            }
        }
        const unsigned b = v_slot[1].packed;
        if (j >= 0 && b != ~0u) {
            const double bt = v_wt[b >> 21];
            if (bt < T0) {
                tj = bt + tf;
                pj = (int)(b & 0xFFFF);
                cj = (int)((b >> 16) & 31) - 1;
            }
        }
    }
    const double ti0 = ti;
    const int pi0 = pi, ci0 = ci;
    double st_t = lf0.tim;
    int st_p = lf0.nptnr, st_c = leaf_ct(lf0.meta);
    bool modified = false, pending = false;
    int rem = first ? cI + cJ : 0;
    if (is_i) {
        st_t = ti;
        st_p = pi;
        st_c = ci;
    }
    // ---- the rest of each down bead's ops; full recomputes from the batch,
    //      others (not foreseen) cooperatively one at a time
    const bool have_full = spec && v_spec[myslot].off >= 0;
    PH_LANE_DECL(ph_single);
    for (int guard = 0; guard < 64 * 9; ++guard) {
        while (first && (rem > 0 || pending)) {
            if (pending) {
                if (is_i) {
                    st_t = ti0;
                    st_p = pi0;
                    st_c = ci0;
                } else if (have_full) {
                    const unsigned a = v_slot[2 + myslot].packed;
                    st_t = T0 + tf;
                    st_p = -1;
                    st_c = -1;
                    if (a != ~0u) {
                        const double bt = v_wt[a >> 21];
                        if (bt < T0) {
                            st_t = bt + tf;
                            st_p = (int)(a & 0xFFFF);
                            st_c = (int)((a >> 16) & 31) - 1;
                        }
                    }
                } else
                    break;
                modified = true;
                pending = false;
            } else {
                const bool opi = rem > cJ;
                const int p = opi ? i : j;
                const int sq = opi ? sqI : sqJ;
                --rem;
                if (st_p == p)
                    pending = true;
                else if (sq >= 0) {
                    // eventredo_down.f from the batch's prediction
                    double tij = v_wt[sq];
                    const int ty = v_wy[sq];
                    if (ty == -99) status = kErrBadCode;
                    if (tij < 0.0 && fabs(tij) > 1.0e-10) status = kErrNegativeTime;
                    tij = tij + tf;
                    if (tij < st_t) {
                        st_t = tij;
                        st_p = p;
                        st_c = ty;
                        modified = true;
                    }
                } else {
                    single_update(T, rl, ol, bwl, p, opi ? P.ri : P.rj, tf, st_t, st_p, st_c, modified, status);
                    PH_LANE_INC(ph_single);
                }
            }
        }
        const unsigned long long pm = __ballot(pending);
        if (pm == 0) break;
        PH_COUNT(S, kPhNFull);
        const int leader = __ffsll((long long)pm) - 1;
        const int lf_l = __shfl(l, leader, kW);
        const Rec rlf = shfl_rec(rl, leader);
        const double2 bwlf = make_double2(__shfl(bwl.x, leader, kW), __shfl(bwl.y, leader, kW));
        const int4 xlf = make_int4(__shfl(xl.x, leader, kW), __shfl(xl.y, leader, kW), __shfl(xl.z, leader, kW),
                                   __shfl(xl.w, leader, kW));
        double ft;
        int fp, fc;
        full_one(S, L, lf_l, rlf, bwlf, xlf, tf, T0, ft, fp, fc, status);
        if (lane == leader) {
            st_t = ft;
            st_p = fp;
            st_c = fc;
            modified = true;
            pending = false;
        }
    }
    PH_LANE_ADD(S, kPhNSingle, ph_single);
    // bead i's state lives in (ti, pi, ci) and is written with i's leaf
    const unsigned long long im = __ballot(is_i);
    if (im) {
        const int src = __ffsll((long long)im) - 1;
        ti = __shfl(st_t, src, kW);
        pi = __shfl(st_p, src, kW);
        ci = __shfl(st_c, src, kW);
    }
    const bool wr = modified && !is_i;
    Leaf lf, la, lb;
    if (wr) {
        lf.tim = st_t;
        lf.nptnr = st_p;
        lf.meta = make_meta(calc_bin(c, st_t, false, status), st_c);
    }
    if (lane == 0) {
        la.tim = ti;
        la.nptnr = pi;
        la.meta = make_meta(calc_bin(c, ti, false, status), ci);
        if (j >= 0) {
            lb.tim = tj;
            lb.nptnr = pj;
            lb.meta = make_meta(calc_bin(c, tj, false, status), cj);
        }
    }
    status = wave_status(status);
    if (status && !c.status) c.status = status;
    if (batch) {
        // report the least new calendar key and the status, then commit
        // only within the round's accepted prefix
        Batch& B = *L.batch();
        PH_END(S, kPhDown, tph);
        Key kn = wr ? leaf_key(lf, l) : empty_key();
        unsigned long long ks = wr ? mk_stamp(c.coll, sub0 + 2 + lane) : 0ull;
        if (lane == 0) {
            const Key ka = leaf_key(la, i);
            const unsigned long long sa = mk_stamp(c.coll, sub0);
            if (key_less_st(ka, sa, kn, ks)) {
                kn = ka;
                ks = sa;
            }
            if (j >= 0) {
                const Key kb = leaf_key(lb, j);
                const unsigned long long sb = mk_stamp(c.coll, sub0 + 1);
                if (key_less_st(kb, sb, kn, ks)) {
                    kn = kb;
                    ks = sb;
                }
            } else {
                // a ghost event leading the round: its next ghost time
                const Key kg = leaf_key(leaf_get(L, S.N), S.N);
                const unsigned long long sg = mk_stamp(c.coll, 1);  // re-added by this ghost event
                if (key_less_st(kg, sg, kn, ks)) {
                    kn = kg;
                    ks = sg;
                }
            }
        }
        unsigned long long kns = 0;
        kn = key_min_wave_st(kn, ks, kns);
        if (lane == 0) {
            B.knew[bk] = kn;
            B.knew_st[bk] = kns;
            B.st[bk] = c.status;
        }
        bar_to(nb, kRoundBars);
        PH_END(S, kPhBKeys, tph);
        // every wavefront has checked: clear this one's marks
        round_clear_marks(L, S, dent, i, j, bk == 0 && sp_w ? sp : -1);
        if (bk >= batch_accepted(B)) return;
        if (bk >= 1 && lane == 0) {
            gst(S.rec + (i), P.ri);
            gst(S.rec + (j), P.rj);
        }
    }
    bool wb = wr;  // the lane's bead takes a new entry (its block is rescanned)
    if (rep >= 0) {
        // the earlier event's entry first, then this event's single-pair
        // updates on it (i before j, eventredo_down.f), as the serial order
        // would have applied them
        Batch& B = *L.batch();
        while (__hip_atomic_load(&B.wdone[rep], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
            __builtin_amdgcn_s_sleep(1);
        const Leaf cur = leaf_get_fresh(L, l);
        double rt = cur.tim;
        int rp = cur.nptnr, rc = leaf_ct(cur.meta), st2 = 0;
        bool mod = false;
        if (cI > 0) {
            const double t = v_wt[sqI] + tf;
            if (t < rt) {
                rt = t;
                rp = i;
                rc = v_wy[sqI];
                mod = true;
            }
        }
        if (cJ > 0) {
            const double t = v_wt[sqJ] + tf;
            if (t < rt) {
                rt = t;
                rp = j;
                rc = v_wy[sqJ];
                mod = true;
            }
        }
        if (mod) {
            Leaf nl;
            nl.tim = rt;
            nl.nptnr = rp;
            nl.meta = make_meta(calc_bin(c, rt, false, st2), rc);
            leaf_set(L, l, nl);
            stamp_set(l, mk_stamp(c.coll, sub0 + 2 + lane));
        }
        st2 = wave_status(st2);
        if (st2 && lane == 0 && B.st[bk] == 0) B.st[bk] = st2;
        wb = mod;
    } else if (wr) {
        leaf_set(L, l, lf);
        stamp_set(l, mk_stamp(c.coll, sub0 + 2 + lane));
    }
    if (lane == 0) {
        L.blist()[0] = i / kBlk;
        L.blist()[1] = j >= 0 ? j / kBlk : i / kBlk;
        L.blist()[2] = extra_blk >= 0 ? extra_blk : i / kBlk;
    }
    const unsigned long long mm = __ballot(wb);
    const int pre = __popcll(mm & ((1ull << lane) - 1ull));
    if (wb) L.blist()[3 + pre] = l / kBlk;
    const int nbl = 3 + __popcll(mm);
    PH_END(S, kPhDown, tph);
    if (lane == 0) {
        leaf_set(L, i, la);
        stamp_set(i, mk_stamp(c.coll, sub0));
        if (j >= 0) {
            leaf_set(L, j, lb);
            stamp_set(j, mk_stamp(c.coll, sub0 + 1));
        }
    }
    // repuls_del_b.f once the rescheduling is done
    if (xpulse_del && lane == 0) {
        const bool iN = dyn_ident(P.ri.dyn) < dyn_ident(P.rj.dyn);
        const int a = iN ? i : j, b = iN ? j : i;
        const int La = topo_len(iN ? P.ri.topo : P.rj.topo), Lb = topo_len(iN ? P.rj.topo : P.ri.topo);
        const int ncim1 = a + La - 1, ncai = a - La, ncaj = b - 2 * Lb, nnjp1 = b - Lb + 1;
        int4 ea = gld(S.extra + (a)), eb = gld(S.extra + (b));
        ea.x = -1;
        ea.y = -1;
        ea.w = -1;
        eb.x = -1;
        eb.y = -1;
        eb.w = -1;
        gst(S.extra + (a), ea);
        gst(S.extra + (b), eb);
        gs1(&S.extra[ncaj].z, -1);
        gs1(&S.extra[nnjp1].z, -1);
        gs1(&S.extra[ncai].z, -1);
        gs1(&S.extra[ncim1].z, -1);
    }
    if (batch) {
        // the round's block rescans run after every committed event's
        // writes; a later event that replays on these entries waits for
        // the flag
        wave_sync_mem();
        if (lane == 0) {
            L.batch()->nbl[bk] = nbl;
            __hip_atomic_store(&L.batch()->wdone[bk], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        return;
    }
    if (!L.leaf_lds || xpulse_del) {
        wave_sync_mem();
    } else
        wave_sync();
    PH_END(S, kPhLeaf, tph);
    rescan_block_list(L, N3, nbl);
    PH_END(S, kPhRescan, tph);
}

// ------------------------------------------------------------------------
// event dynamics: eventdyn.f, bumped.f, repuls_*.f (uniform over lanes)
// ------------------------------------------------------------------------
__device__ inline void nudge(Rec& bi, Rec& bj, double bd, const Geo& g, int sign) {
    if (sign > 0) {
        bi.x = bi.x + bd * g.rx;
        bi.y = bi.y + bd * g.ry;
        bi.z = bi.z + bd * g.rz;
        bj.x = bj.x - bd * g.rx;
        bj.y = bj.y - bd * g.ry;
        bj.z = bj.z - bd * g.rz;
    } else {
        bi.x = bi.x - bd * g.rx;
        bi.y = bi.y - bd * g.ry;
        bi.z = bi.z - bd * g.rz;
        bj.x = bj.x + bd * g.rx;
        bj.y = bj.y + bd * g.ry;
        bj.z = bj.z + bd * g.rz;
    }
}

__device__ void eventdyn(const DevTables& T, Rec& bi, Rec& bj, double2 bwi, bool bonded_ij, int evcode, double tf,
                         int& ct, int& status) {
    const int a = dyn_ident(bi.dyn), c = dyn_ident(bj.dyn);
    Geo g;
    geo(bi, bj, tf, g);
    const double bmi = bi.bm, bmj = bj.bm;
    const double rmass = 2 * bmi * bmj / (bmi + bmj);
    double ratio = 0.0, wellsq, epsave, del_pe, bd;
    switch (ct) {
        case 2:
        case 3: {
            double blmin, blmax;
            bond_window(T, evcode, bwi, blmin, blmax);
            ratio = ct == 2 ? rmass * g.b / (blmin * blmin) : rmass * g.b / (blmax * blmax);
            break;
        }
        case 1: {
            double sigsq;
            if (evcode == 15) {
                if (bonded_ij) {
                    const double pe = T.ev_param[1][evcode];
                    sigsq = T.sigma_sq[a][c] * (pe * pe);
                } else
                    sigsq = T.sigma_sq[a][c];
            } else {
                const double pe = T.ev_param[1][evcode];
                sigsq = T.sigma_sq[a][c] * (pe * pe);
                if (evcode >= 22 && evcode <= 26) {
                    const double q = T.sqz610[evcode - 21][a > c ? a : c];
                    sigsq = sigsq * (q * q);
                }
            }
            ratio = rmass * g.b / sigsq;
            break;
        }
        case 4:
            wellsq = T.welldia_sq[a][c];
            epsave = T.ep_sqrt[a][c];
            del_pe = 4.0 * wellsq * epsave / rmass;
            bd = kSm * sqrt(wellsq);
            if (g.b * g.b + del_pe > 0.0) {
                ratio = rmass * (sqrt((4.0 * wellsq * epsave / rmass) + g.b * g.b) + g.b) / (2.0 * wellsq);
                ct = 20;
                nudge(bi, bj, bd, g, -1);
            } else {
                ratio = rmass * g.b / wellsq;
                ct = 22;
                nudge(bi, bj, bd, g, +1);
            }
            break;
        case 8:
            wellsq = T.welldia_sq[a][c];
            epsave = T.ep_sqrt[a][c];
            del_pe = 4.0 * wellsq * epsave / rmass;
            bd = kSm * sqrt(wellsq);
            if (g.b * g.b > del_pe) {
                ratio = rmass * (-sqrt(-del_pe + g.b * g.b) + g.b) / (2.0 * wellsq);
                ct = 21;
                nudge(bi, bj, bd, g, +1);
            } else {
                ratio = rmass * g.b / wellsq;
                ct = 22;
                nudge(bi, bj, bd, g, -1);
            }
            break;
        case 9:
            wellsq = T.welldia_sq[a][c];
            ratio = rmass * g.b / wellsq;
            ct = 23;
            bd = kSm * sqrt(wellsq);
            nudge(bi, bj, bd, g, +1);
            break;
        case 5:
            wellsq = T.shlddia_sq[a][c];
            epsave = -T.epsilon1;
            del_pe = 4.0 * wellsq * epsave / rmass;
            bd = kSm * sqrt(wellsq);
            ratio = rmass * (-sqrt(-del_pe + g.b * g.b) + g.b) / (2.0 * wellsq);
            ct = 24;
            nudge(bi, bj, bd, g, +1);
            break;
        case 6:
            wellsq = T.shlddia_sq[a][c];
            epsave = -T.epsilon1;
            del_pe = 4.0 * wellsq * epsave / rmass;
            bd = kSm * sqrt(wellsq);
            if (g.b * g.b > -del_pe) {
                ratio = rmass * (sqrt(del_pe + g.b * g.b) + g.b) / (2.0 * wellsq);
                ct = 25;
                nudge(bi, bj, bd, g, -1);
            } else {
                ratio = rmass * g.b / wellsq;
                ct = 26;
                nudge(bi, bj, bd, g, +1);
            }
            break;
        case 13:
            wellsq = T.shlddia_sq[a][c];
            ratio = rmass * g.b / wellsq;
            ct = 27;
            bd = kSm * sqrt(wellsq);
            nudge(bi, bj, bd, g, +1);
            break;
        default:
            status = kErrDyn;
            return;
    }
    const double dvx = ratio * g.rx, dvy = ratio * g.ry, dvz = ratio * g.rz;
    bi.vx = bi.vx - dvx / bmi;
    bj.vx = bj.vx + dvx / bmj;
    bi.vy = bi.vy - dvy / bmi;
    bj.vy = bj.vy + dvy / bmj;
    bi.vz = bi.vz - dvz / bmi;
    bj.vz = bj.vz + dvz / bmj;
    bi.x = bi.x + dvx * tf / bmi;
    bi.y = bi.y + dvy * tf / bmi;
    bi.z = bi.z + dvz * tf / bmi;
    bj.x = bj.x - dvx * tf / bmj;
    bj.y = bj.y - dvy * tf / bmj;
    bj.z = bj.z - dvz * tf / bmj;
}

__device__ inline double dist_rec(const Rec& a, const Rec& b, double tf) {
    Geo g;
    geo(a, b, tf, g);
    return sqrt(g.rx * g.rx + g.ry * g.ry + g.rz * g.rz);
}
// repuls_check.f:16-81 (n = the N bead, cb = the C bead)
__device__ double repuls_check(const SysDev&, const DevTables& T, const Pair& P, int n, int cb, double tf) {
    const SysDev& S = sdev();

    const Rec rn = load_rec(S, P, n), rc = load_rec(S, P, cb);
    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(n, topo_len(rn.topo), cb, topo_len(rc.topo), ncim1, ncai, ncaj, nnjp1);
    const double d1 = dist_rec(rn, load_rec(S, P, ncaj), tf), d2 = dist_rec(rn, load_rec(S, P, nnjp1), tf);
    const double d3 = dist_rec(rc, load_rec(S, P, ncai), tf), d4 = dist_rec(rc, load_rec(S, P, ncim1), tf);
    if (d1 > T.shder[1] && d2 > T.shder[2] && d3 > T.shder[3] && d4 > T.shder[4]) return 10.0;
    return 11.0;
}
// repuls_check_3.f:16-102
__device__ double repuls_check_3(const SysDev&, const DevTables& T, const Pair& P, int n, int cb, int k,
                                 double tf) {
    const SysDev& S = sdev();

    const Rec rn = load_rec(S, P, n), rc = load_rec(S, P, cb);
    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(n, topo_len(rn.topo), cb, topo_len(rc.topo), ncim1, ncai, ncaj, nnjp1);
    int m = 0;
    if (ncaj != k && dist_rec(rn, load_rec(S, P, ncaj), tf) > T.shder[1]) ++m;
    if (nnjp1 != k && dist_rec(rn, load_rec(S, P, nnjp1), tf) > T.shder[2]) ++m;
    if (ncai != k && dist_rec(rc, load_rec(S, P, ncai), tf) > T.shder[3]) ++m;
    if (ncim1 != k && dist_rec(rc, load_rec(S, P, ncim1), tf) > T.shder[4]) ++m;
    return m == 3 ? 10.0 : 11.0;
}
// repuls_add.f:14-47 / repuls_del_a.f:14-37 (single lane; n = N bead, cb = C bead)
__device__ void repuls_add(const SysDev&, int n, int Ln, int cb, int Lc, int& status) {
    const SysDev& S = sdev();

    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(n, Ln, cb, Lc, ncim1, ncai, ncaj, nnjp1);
    set_code(S, ncaj, n, 50, status);
    set_code(S, n, ncaj, 40, status);
    set_code(S, n, nnjp1, 40, status);
    set_code(S, nnjp1, n, 50, status);
    set_code(S, ncai, cb, 50, status);
    set_code(S, cb, ncai, 40, status);
    set_code(S, cb, ncim1, 40, status);
    set_code(S, ncim1, cb, 50, status);
    int4 en = gld(S.extra + (n)), ec = gld(S.extra + (cb));
    en.x = ncaj;
    en.y = nnjp1;
    en.w = cb;
    ec.x = ncai;
    ec.y = ncim1;
    ec.w = n;
    gst(S.extra + (n), en);
    gst(S.extra + (cb), ec);
    gs1(&S.extra[ncaj].z, n);
    gs1(&S.extra[nnjp1].z, n);
    gs1(&S.extra[ncai].z, cb);
    gs1(&S.extra[ncim1].z, cb);
}
__device__ void repuls_del_a(const SysDev&, int n, int Ln, int cb, int Lc, int& status) {
    const SysDev& S = sdev();

    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(n, Ln, cb, Lc, ncim1, ncai, ncaj, nnjp1);
    set_code(S, ncaj, n, 1, status);
    set_code(S, n, ncaj, 1, status);
    set_code(S, n, nnjp1, 1, status);
    set_code(S, nnjp1, n, 1, status);
    set_code(S, ncai, cb, 1, status);
    set_code(S, cb, ncai, 1, status);
    set_code(S, ncim1, cb, 1, status);
    set_code(S, cb, ncim1, 1, status);
}
__device__ inline bool terminal_t(int t) {
    const int k = topo_kind(t), r = topo_res(t), L = topo_len(t);
    return (k == 1 && r == 1) || (k == 2 && r == L);
}

// One pair event: outcome resolution (main.F90:1487-1634), dynamics, H-bond
// and shoulder bookkeeping (main.F90:641-921), then partial_events.
__device__ int pair_event(const SysDev&, const DevTables&, const Lds&, Sc&, int i, int j, int ct, int bk) {
    const Lds L = lds_regs();

    const SysDev& S = sdev();

    const DevTables& T = tabs();
    Sc& c = *L.sc();

    const int lane = lane_id();
    const double tf = c.tfalse;
    PH_BEGIN(tpe);
    Pair P;
    P.i = i;
    P.j = j;
    P.flags = 0;
    P.pad = 0;
    P.ri = gld(S.rec + (i));
    P.rj = gld(S.rec + (j));
    P.bwi = gld(S.bw + (i));
    P.bwj = gld(S.bw + (j));
    P.xi = gld(S.extra + (i));
    P.xj = gld(S.extra + (j));
    const int4 oi = gld(S.ovr + (i));
    const Rows rows = load_rows(S, i, j);  // in flight while the outcome is resolved
    const int evcode = dyn_code(oi, j, static_t(P.ri.topo, P.rj.topo));  // ev_code(i, j)
    const bool ok = !terminal_t(P.ri.topo) && !terminal_t(P.rj.topo);
    const int idi = dyn_ident(P.ri.dyn), idj = dyn_ident(P.rj.dyn);
    int status = 0;
    bool xpulse = false, fence = false;
    if (ct == 7 || ct == 10 || ct == 12) P.flags |= kPairReadsSp;
    if (ct == 7) {
        if (P.xi.w < 0 && P.xj.w < 0) {
            if (ok) {
                const double rating = idi < idj ? repuls_check(S, T, P, i, j, tf) : repuls_check(S, T, P, j, i, tf);
                ct = rating <= 10.0 ? 4 : 14;
            } else {
                // a random draw: its counter value is known only in order
                int nb = 0;
                if (bk >= 1) return spec_abort(L, bk, nb);
                const double ran_non = rng_u01(S.seed, 1, c.rc1++);
                ct = ran_non <= 0.2 ? 4 : 9;
            }
        } else
            ct = 9;
    } else if (ct == 10) {
        double rating;
        if (evcode < 45) {
            const int hbp = P.xi.w;
            const int idh = dyn_ident(load_rec(S, P, hbp).dyn);
            rating = idi < idh ? repuls_check_3(S, T, P, i, hbp, j, tf) : repuls_check_3(S, T, P, hbp, i, j, tf);
        } else {
            const int hbp = P.xj.w;
            const int idh = dyn_ident(load_rec(S, P, hbp).dyn);
            rating = idj < idh ? repuls_check_3(S, T, P, j, hbp, i, tf) : repuls_check_3(S, T, P, hbp, j, i, tf);
        }
        ct = rating <= 10.0 ? 5 : 15;
    } else if (ct == 12) {
        const bool xi_ = evcode < 45;
        const Rec rx = xi_ ? P.ri : P.rj;
        const int hbp = xi_ ? P.xi.w : P.xj.w;
        if (dyn_bptnr(rx.dyn) == hbp) {
            Geo g;
            const Rec rh = load_rec(S, P, hbp);
            geo(rx, rh, tf, g);
            const double d2 = g.rx * g.rx + g.ry * g.ry + g.rz * g.rz;
            const double rating = (d2 - T.sigma_sq[dyn_ident(rx.dyn)][dyn_ident(rh.dyn)]) < 0.0 ? 10.0 : 11.0;
            ct = rating > 10.0 ? 6 : 13;
        } else
            ct = 15;
    }
    if (ct < 14) eventdyn(T, P.ri, P.rj, P.bwi, dyn_bptnr(P.ri.dyn) == j, evcode, tf, ct, status);
    const int Li = topo_len(P.ri.topo), Lj = topo_len(P.rj.topo);
    if (ct == 20) {
        if (idi + idj == 5) {
            P.ri.dyn = make_dyn(idi + 4, j);
            P.rj.dyn = make_dyn(idj + 4, i);
            if (ok) {
                int nb = 0;
                if (bk >= 1) return spec_abort(L, bk, nb);
                if (lane == 0) {
                    if (idi < idj) repuls_add(S, i, Li, j, Lj, status);
                    else repuls_add(S, j, Lj, i, Li, status);
                }
                fence = true;
            }
        }
    } else if (ct == 21) {
        if (idi <= 8) {
            if (ok) {
                int nb = 0;
                if (bk >= 1) return spec_abort(L, bk, nb);
                if (lane == 0) {
                    if (idi < idj) repuls_del_a(S, i, Li, j, Lj, status);
                    else repuls_del_a(S, j, Lj, i, Li, status);
                }
                xpulse = true;
                fence = true;
            }
            if (dyn_bptnr(P.ri.dyn) == j) {
                P.ri.dyn = make_dyn(idi - 4, -1);
                P.rj.dyn = make_dyn(idj - 4, -1);
            }
        }
    } else if (ct == 24 || ct == 25) {
        P.flags |= kPairReadsSp;
        const bool xi_ = evcode < 45;
        const int hbp = xi_ ? P.xi.w : P.xj.w;
        const int x = xi_ ? i : j;
        int xdyn = xi_ ? P.ri.dyn : P.rj.dyn;
        int hdyn = load_rec(S, P, hbp).dyn;
        const int idx = dyn_ident(xdyn), idh = dyn_ident(hdyn);
        bool changed = false;
        if (ct == 24 && idx + idh == 5) {
            xdyn = make_dyn(idx + 4, hbp);
            hdyn = make_dyn(idh + 4, x);
            changed = true;
        } else if (ct == 25 && idx >= 5) {
            xdyn = make_dyn(idx - 4, -1);
            hdyn = make_dyn(idh - 4, -1);
            changed = true;
        }
        if (changed) {
            if (xi_) P.ri.dyn = xdyn;
            else P.rj.dyn = xdyn;
            if (hbp == i) P.ri.dyn = hdyn;
            else if (hbp == j) P.rj.dyn = hdyn;
            else {
                int nb = 0;
                if (bk >= 1) return spec_abort(L, bk, nb);
                if (lane == 0) gs1(&S.rec[hbp].dyn, hdyn);
                fence = true;
            }
        }
    } else if (ct == 14 || ct == 15 || ct == 16) {
        // bumped.f: push the pair off the discontinuity
        Geo g;
        geo(P.ri, P.rj, tf, g);
        const double bd = evcode >= 40 ? kSm * sqrt(T.shlddia_sq[idi][idj]) : kSm * sqrt(T.welldia_sq[idi][idj]);
        nudge(P.ri, P.rj, bd, g, g.b < 0.0 ? -1 : +1);
        if ((ct == 14 || ct == 16) && ok && bk >= 1) {
            int nb = 0;
            return spec_abort(L, bk, nb);
        }
        if (ct == 14 && ok) {
            if (lane == 0) {
                if (idi < idj) repuls_add(S, i, Li, j, Lj, status);
                else repuls_add(S, j, Lj, i, Li, status);
            }
            fence = true;
        } else if (ct == 16 && ok) {
            if (lane == 0) {
                if (idi < idj) repuls_del_a(S, i, Li, j, Lj, status);
                else repuls_del_a(S, j, Lj, i, Li, status);
            }
            xpulse = true;
            fence = true;
        }
    }
    if (lane == 0 && bk < 1) {  // a speculative event writes its pair once committed
        gst(S.rec + (i), P.ri);
        gst(S.rec + (j), P.rj);
    }
    if (lane == 0 && bk >= 0) L.batch()->ty[bk] = ct;
    if (fence) P.flags |= kPairWroteSp;
    if (fence) {
        // shoulder tables changed: make them visible before the predictions
        wave_sync_mem();
        P.xi = gld(S.extra + (i));
        P.xj = gld(S.extra + (j));
    }
    status = wave_status(status);
    if (status && !c.status) c.status = status;
    if (lane == 0) *L.pair() = P;
    wave_sync();
    PH_END(S, kPhPair, tpe);
    reschedule(S, T, L, c, *L.pair(), -1, xpulse, bk, rows);
    return ct;
}

// Andersen ghost collision, main.F90:997-1049
__device__ int ghost_event(const SysDev&, const DevTables&, const Lds&, Sc&, int bk) {
    const Lds& L = ldsv();

    const SysDev& S = sdev();

    const DevTables& T = tabs();
    Sc& c = *L.sc();

    const int lane = lane_id();
    const int N = S.N;
    PH_BEGIN(tg0);
    int i = 0, guard = 0;
    do {
        i = (int)(rng_u01(S.seed, 0, c.rc0++) * N);
    } while (i == N && ++guard < 1000);
    c.numghosts++;
    Pair P;
    P.i = i;
    P.j = -1;
    P.flags = 0;
    P.pad = 0;
    P.ri = gld(S.rec + (i));
    P.bwi = gld(S.bw + (i));
    P.xi = gld(S.extra + (i));
    P.rj = P.ri;
    P.bwj = P.bwi;
    P.xj = make_int4(-1, -1, -1, -1);
    Rec b = P.ri;
    const double tf = c.tfalse, bm = b.bm;
    b.x = b.x + b.vx * tf;
    b.y = b.y + b.vy * tf;
    b.z = b.z + b.vz * tf;
    double v1, v2, r, fact;
    do {
        v1 = 2.0 * rng_u01(S.seed, 0, c.rc0++) - 1.0;
        v2 = 2.0 * rng_u01(S.seed, 0, c.rc0++) - 1.0;
        r = v1 * v1 + v2 * v2;
    } while ((r == 0.0 || r >= 1.0) && ++guard < 1000);
    fact = sqrt(-2.0 * S.setemp * bm * dlog(r) / r);
    b.vx = v1 * fact / bm;
    b.vy = v2 * fact / bm;
    do {
        v1 = 2.0 * rng_u01(S.seed, 0, c.rc0++) - 1.0;
        v2 = 2.0 * rng_u01(S.seed, 0, c.rc0++) - 1.0;
        r = v1 * v1 + v2 * v2;
    } while ((r == 0.0 || r >= 1.0) && ++guard < 2000);
    fact = sqrt(-2.0 * S.setemp * bm * dlog(r) / r);
    b.vz = v1 * fact / bm;
    b.x = b.x - b.vx * tf;
    b.y = b.y - b.vy * tf;
    b.z = b.z - b.vz * tf;
    double tgho;
    do {
        tgho = rng_u01(S.seed, 0, c.rc0++);
    } while ((tgho < 1.e-18 || tgho == 1.0) && ++guard < 3000);
    if (guard >= 1000) c.status = kErrRng;
    const double tg = -1.0 * dlog(tgho) * S.avegtime + tf;
    int status = 0;
    P.ri = b;
    P.rj = b;
    if (lane == 0) {
        gst(S.rec + (i), b);
        Leaf g;
        g.tim = tg;
        g.nptnr = -2;
        g.meta = make_meta(calc_bin(c, tg, false, status), -2);
        leaf_set(L, N, g);
        stamp_set(N, mk_stamp(c.coll, 1));  // added before partial_events' adds (main.F90:1041-1049)
    }
    if (status && !c.status) c.status = status;
    if (c.tfalse < c.old_tfalse) c.tfalse = c.old_tfalse;
    if (lane == 0) *L.pair() = P;
    if (lane == 0 && bk >= 0) {
        L.batch()->ty[bk] = 0;
        L.batch()->gbead = i;
    }
    wave_sync();
    PH_END(S, kPhGhost, tg0);
    reschedule(S, T, L, c, *L.pair(), N / kBlk, false, bk, load_rows(S, i, -1));
    return i;
}

// ------------------------------------------------------------------------
// neighbour lists: nbor.f membership restated for a coarse-cell search
// ------------------------------------------------------------------------
// Pair (i, j) is listed when its fine cells (cell_add.f binning) are within
// the n_wrap=2 stencil (|d| <= 2 per axis, periodic) and either its code is
// a bond/local one (4-12, 17-26) or r^2 <= rlsq(code) (nbor.f:63-128).
// Steps: fine and coarse cell of every bead; stable counting sort of the
// beads by coarse cell; copies of position, topology and fine cell in that
// order; then one lane per bead scans the (up to 27) neighbouring coarse
// cells, all cell ranges requested before the candidates, each candidate
// read with two independent loads.  Candidates come in a fixed cell order
// and ascending bead index within a cell, so the lists are deterministic.
// v mod m for v in [-1, m]
__device__ inline int wrap1(int v, int m) { return v < 0 ? v + m : (v >= m ? v - m : v); }
__device__ void rebuild_lists(const SysDev&, const Lds&, Sc& c) {
    const Lds& L = ldsv();

    const SysDev& S = sdev();

    const DevTables& T = tabs();
    const int lane = lane_id(), wave = wave_id(), nw = L.nw;
    const int tid = (int)threadIdx.x, nthr = kW * nw;
    // coarse cells of cq >= 2 fine cells each (the last one takes the
    // remainder): every fine cell within the n_wrap = 2 stencil of a bead's
    // lies in the 27 coarse cells around the bead's own
    const int N = S.N, M = S.M, cq = S.cq, Mq = M / cq, NC = Mq * Mq * Mq, cap = S.cap;
    PH_BEGIN(trl);
#pragma unroll 4
    for (int k = tid; k < N; k += nthr) {
        const Rec b = gld(S.rec + (k));
        int cx = (int)((b.x + S.half) / S.width);
        int cy = (int)((b.y + S.half) / S.width);
        int cz = (int)((b.z + S.half) / S.width);
        cx = ((cx % M) + M) % M;
        cy = ((cy % M) + M) % M;
        cz = ((cz % M) + M) % M;
        gs1(S.cfine + (k), cx | (cy << 10) | (cz << 20));
        const int qx = min(cx / cq, Mq - 1), qy = min(cy / cq, Mq - 1), qz = min(cz / cq, Mq - 1);
        gs1(S.ckey + (k), qx + Mq * (qy + Mq * qz));
    }
    for (int q = tid; q < NC; q += nthr) {
        gs1(S.cstart + (q), 0);
        gs1(S.crun + (q), 0);
    }
    wg_sync(nw);
    for (int k = tid; k < N; k += nthr) atomicAdd(&S.cstart[gl1(S.ckey + k)], 1);
    wg_sync(nw);
    PH_END(S, kPhLCells, trl);
    // exclusive scan of the cell counts (wavefront 0)
    int carry = 0;
    for (int base = 0; wave == 0 && base < NC; base += kW) {
        const int q = base + lane;
        const int v = q < NC ? gl1(S.cstart + (q)) : 0;
        int incl = v;
        incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xf, 0xf, false);
        incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xf, 0xf, false);
        incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xf, 0xf, false);
        incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xf, 0xf, false);
        incl += __builtin_amdgcn_update_dpp(0, incl, 0x142, 0xa, 0xf, false);
        incl += __builtin_amdgcn_update_dpp(0, incl, 0x143, 0xc, 0xf, false);
        if (q < NC) gs1(S.cstart + (q), carry + incl - v);
        carry += __builtin_amdgcn_readlane(incl, kW - 1);
    }
    if (wave == 0 && lane == 0) gs1(S.cstart + (NC), carry);
    wg_sync(nw);
    PH_END(S, kPhLScan, trl);
    // stable scatter in bead order (wavefront 0): rank among the lower lanes
    // of the same cell
    for (int base = 0; wave == 0 && base < N; base += kW) {
        const int k = base + lane;
        const int key = k < N ? gl1(S.ckey + (k)) : -1;
        int rank = 0, cnt = 0;
        const int nl = min(N - base, kW);
        for (int r = 0; r < nl; ++r) {
            const int o = __builtin_amdgcn_readlane(key, r);
            if (o == key) {
                if (r < lane) ++rank;
                ++cnt;
            }
        }
        const int basepos = k < N ? gl1(S.cstart + (key)) + gl1(S.crun + (key)) : 0;
        wave_sync_mem();
        if (k < N) {
            gs1(S.csorted + (basepos + rank), k);
            if (rank == cnt - 1) gs1(S.crun + (key), gl1(S.crun + (key)) + cnt);
        }
        wave_sync_mem();
    }
    wg_sync(nw);
    PH_END(S, kPhLScatter, trl);
    // positions, topology and fine cells in sorted order
#pragma unroll 4
    for (int p2 = tid; p2 < N; p2 += nthr) {
        const int jb = gl1(S.csorted + (p2));
        const Rec b = gld(S.rec + (jb));
        gst(S.spos + (p2), make_double4(b.x, b.y, b.z, 0.0));
        gst(S.sinfo + (p2), make_int4(jb, b.topo, gl1(S.cfine + jb), 0));
    }
    wg_sync(nw);
    // each bead's place in its fine cell's list as cell_add.f builds it
    // (head first, the highest index at 1): the beads of a fine cell sit
    // in its coarse cell's sorted run in index order
    for (int p2 = tid; p2 < N; p2 += nthr) {
        const int4 si = gld(S.sinfo + (p2));
        const int end = gl1(S.cstart + (gl1(S.ckey + si.x) + 1));
        int pos = 1;
#pragma unroll 4
        for (int q2 = p2 + 1; q2 < end; ++q2) pos += gl1(&S.sinfo[q2].z) == si.z ? 1 : 0;
        gs2(S.cfp + (si.x), make_int2(si.z, pos));
    }
    PH_END(S, kPhLCopy, trl);
    int mx = 0, status = 0;
    const int noff = Mq > 2 ? 3 : Mq;  // distinct coarse offsets {0, +1, -1}
    // per lane, its bead's non-empty neighbour cell ranges, [t][lane] in the
    // wavefront's own work area; wavefronts take 64-bead chunks in turn
    int2* rng = reinterpret_cast<int2*>(smem + L.o_wm);
    // beads are taken in cell order (64 consecutive sorted places per
    // chunk): the lanes of a chunk sit in the same or adjacent cells, so
    // their candidate reads fall on shared cache lines
    for (int i0 = wave * kW; i0 < N; i0 += nw * kW) {
        const int ps = i0 + lane;
        const bool act = ps < N;
        const int4 si = act ? gld(S.sinfo + (ps)) : make_int4(-1, 0, 0, 0);
        const double4 bi = act ? gld(S.spos + (ps)) : make_double4(0.0, 0.0, 0.0, 0.0);
        const int i = si.x;  // -1 on an idle lane
        const int fi = si.z;
        const int ix = fi & 1023, iy = (fi >> 10) & 1023, iz = fi >> 20;
        // the coarse cell, as the binning step derived it from the fine one
        const int qx = min(ix / cq, Mq - 1), qy = min(iy / cq, Mq - 1), qz = min(iz / cq, Mq - 1);
        const int4 oi = gld(S.ovr + (act ? i : 0));
        int nr = 0;
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            const int a = t / 9, bb = (t / 3) % 3, cc = t % 3;
            if (act && a < noff && bb < noff && cc < noff) {
                const int oa = a == 2 ? -1 : a, ob = bb == 2 ? -1 : bb, oc = cc == 2 ? -1 : cc;
                // periodic neighbour of a coarse cell in [0, Mq) by one step
                const int nx = wrap1(qx + oa, Mq), ny = wrap1(qy + ob, Mq), nz = wrap1(qz + oc, Mq);
                const int cell = nx + Mq * (ny + Mq * nz);
                const int c0 = gl1(S.cstart + (cell)), c1 = gl1(S.cstart + (cell + 1));
                if (c1 > c0) {
                    rng[nr * kW + lane] = make_int2(c0, c1);
                    ++nr;
                }
            }
        }
        wave_sync();
        // candidates in (cell, index) order, four in flight per lane
        int ri = 0, p = 0, pe = 0;
        if (nr > 0) {
            const int2 r0 = rng[lane];
            p = r0.x;
            pe = r0.y;
        }
        int nu = 0, nd = 0;
        while (__ballot(ri < nr)) {
            int pu[4];
            bool vu[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                vu[u] = ri < nr;
                pu[u] = p;
                if (vu[u]) {
                    ++p;
                    if (p >= pe) {
                        ++ri;
                        if (ri < nr) {
                            const int2 r1 = rng[ri * kW + lane];
                            p = r1.x;
                            pe = r1.y;
                        }
                    }
                }
            }
            int4 inf[4];
            double4 pj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (vu[u]) {
                    inf[u] = gld(S.sinfo + (pu[u]));
                    pj[u] = gld(S.spos + (pu[u]));
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (!vu[u]) continue;
                const int jb = inf[u].x;
                if (jb == i) continue;
                const int fj = inf[u].z;
                // fine cells lie in [0, M): differences in (-M, M), taken mod M
                int dx = (fj & 1023) - ix, dy = ((fj >> 10) & 1023) - iy, dz = (fj >> 20) - iz;
                dx += dx < 0 ? M : 0;
                dy += dy < 0 ? M : 0;
                dz += dz < 0 ? M : 0;
                if ((dx > 2 && dx < M - 2) || (dy > 2 && dy < M - 2) || (dz > 2 && dz < M - 2)) continue;
                const int stat = static_t(inf[u].y, si.y);
                // ev_code(j,i), nbor.f:63: j's overrides exist only towards i's
                int code = stat;
                if ((oi.x >= 0 && (oi.x >> 8) == jb) || (oi.y >= 0 && (oi.y >> 8) == jb) ||
                    (oi.z >= 0 && (oi.z >> 8) == jb) || (oi.w >= 0 && (oi.w >> 8) == jb))
                    code = dyn_code(gld(S.ovr + (jb)), i, stat);
                bool take;
                if ((code >= 4 && code <= 12) || (code >= 17 && code < 27)) {
                    take = true;
                } else {
                    double rx = bi.x - pj[u].x, ry = bi.y - pj[u].y, rz = bi.z - pj[u].z;
                    rx = rx - round(rx);
                    ry = ry - round(ry);
                    rz = rz - round(rz);
                    const double r2 = rx * rx + ry * ry + rz * rz;
                    take = code >= 1 && code <= 50 && r2 <= T.rlsq[code];
                }
                if (!take) continue;
                if (jb > i) {
                    if (nu < cap) gs1(S.nb_up + ((size_t)i * cap + nu), jb | (stat << 24));
                    ++nu;
                } else {
                    if (nd < cap) gs1(S.nb_dn + ((size_t)i * cap + nd), jb | (stat << 24));
                    ++nd;
                }
            }
        }
        if (act) {
            if (nu > cap || nd > cap) status = kErrListCapacity;
            mx = max(mx, max(nu, nd));
            gs1(S.nup + (i), min(nu, cap));
            gs1(S.ndn + (i), min(nd, cap));
        }
        wave_sync();
    }
    PH_END(S, kPhLSearch, trl);
    status = wave_status(status);
    if (status && !c.status) c.status = status;
    mx = wave_imax(mx);
    if (lane == 0) atomicMax(&S.st->max_nb_seen, (long long)mx);
    wg_sync(nw);
}

// Neighbour lists of a small system (N <= kDenseN) without the cell sort.
// Positions, topology words, override slots and fine cells are staged in
// LDS (wavefront 0's work area is free during a rebuild); each wavefront
// takes beads in turn, tests the fine-cell stencil of all candidates one per
// lane, and queues the (bead, candidate) pairs that pass; full 64-lane
// batches of the queue then take the pair-code and cutoff test.  Membership
// is rebuild_lists' own (the n_wrap = 2 stencil, then the bonded/local codes
// or the cutoff, nbor.f:63-128); rows come in bead-index order.
// Wavefront 0's area holds at least 24 * maxw bytes (its work lists, maxw >=
// kMinMaxw); the staging takes 60 bytes per bead and its queue 512 bytes
// after it.
constexpr int kMinMaxw = 768;
constexpr int kDenseN = 256;
constexpr int kDenseQ = 128;  // queue ring: under 64 entries + one pass of 64
static_assert(60 * kDenseN + 4 * kDenseQ <= 24 * kMinMaxw, "dense list staging exceeds wavefront 0's work area");
static_assert(4 * kDenseQ <= 4 * kMinMaxw, "dense list queue exceeds a wavefront's work list");
struct __attribute__((aligned(16))) Stage {
    double x, y, z;
    int topo, pad;
};
__device__ void rebuild_lists_dense(const SysDev&, const Lds&, Sc& c) {
    const Lds& L = ldsv();

    const SysDev& S = sdev();

    const DevTables& T = tabs();
    const int lane = lane_id(), wave = wave_id(), nw = L.nw;
    const int tid = (int)threadIdx.x, nthr = kW * nw;
    const int N = S.N, M = S.M, cap = S.cap;
    PH_BEGIN(trl);
    unsigned char* const w0 = smem + reinterpret_cast<const Lds*>(smem + kLdOff)->o_wm;
    Stage* const stg = reinterpret_cast<Stage*>(w0);
    int4* const sovr = reinterpret_cast<int4*>(w0 + 32 * N);
    int2* const cnt = reinterpret_cast<int2*>(w0 + 48 * N);
    int* const sfine = reinterpret_cast<int*>(w0 + 56 * N);
    int* const q = wave == 0 ? reinterpret_cast<int*>(w0 + 60 * N) : reinterpret_cast<int*>(smem + L.o_wm);
    for (int k = tid; k < N; k += nthr) {
        const Rec b = gld(S.rec + (k));
        int cx = (int)((b.x + S.half) / S.width);
        int cy = (int)((b.y + S.half) / S.width);
        int cz = (int)((b.z + S.half) / S.width);
        cx = ((cx % M) + M) % M;
        cy = ((cy % M) + M) % M;
        cz = ((cz % M) + M) % M;
        Stage g;
        g.x = b.x;
        g.y = b.y;
        g.z = b.z;
        g.topo = b.topo;
        g.pad = 0;
        stg[k] = g;
        sovr[k] = gld(S.ovr + (k));
        sfine[k] = cx | (cy << 10) | (cz << 20);
    }
    for (int i = wave + nw * lane; i < N; i += nw * kW) cnt[i] = make_int2(0, 0);
    wg_sync(nw);
    // fine cells and each bead's place in its fine cell's list (cell_add.f:
    // head first, the highest index at 1), for the rows' nbor.f order: lane
    // k counts the later beads of its cell, 64 candidates per LDS read
    for (int k0 = wave * kW; k0 < N; k0 += nw * kW) {
        const int k = k0 + lane;
        const int f = k < N ? sfine[k] : -1;
        int pos = 1;
        for (int q0 = k0; q0 < N; q0 += kW) {
            const int v = q0 + lane < N ? sfine[q0 + lane] : -2;
#pragma unroll
            for (int m = 0; m < kW; ++m) pos += (__builtin_amdgcn_readlane(v, m) == f && q0 + m > k) ? 1 : 0;
        }
        if (k < N) {
            gs1(S.cfine + (k), f);
            gs2(S.cfp + (k), make_int2(f, pos));
        }
    }
    PH_END(S, kPhLCells, trl);
    const unsigned long long below = (1ull << lane) - 1ull;
    int qh = 0, qn = 0;  // queue head and length (wavefront-uniform)
    // the pair test on the next min(qn, 64) queued pairs, in queue order
    auto drain = [&]() {
        const int nb = min(qn, kW);
        const bool v = lane < nb;
        const int e = v ? q[(qh + lane) & (kDenseQ - 1)] : 0;
        const int i = e >> 16, jb = e & 0xFFFF;
        bool take = false;
        int stat = 0;
        if (v) {
            const Stage si = stg[i], sj = stg[jb];
            const int4 oi = sovr[i];
            stat = static_t(sj.topo, si.topo);
            // ev_code(j,i), nbor.f:63: j's overrides exist only towards i's
            int code = stat;
            if ((oi.x >= 0 && (oi.x >> 8) == jb) || (oi.y >= 0 && (oi.y >> 8) == jb) ||
                (oi.z >= 0 && (oi.z >> 8) == jb) || (oi.w >= 0 && (oi.w >> 8) == jb))
                code = dyn_code(sovr[jb], i, stat);
            if ((code >= 4 && code <= 12) || (code >= 17 && code < 27)) {
                take = true;
            } else {
                double rx = si.x - sj.x, ry = si.y - sj.y, rz = si.z - sj.z;
                rx = rx - round(rx);
                ry = ry - round(ry);
                rz = rz - round(rz);
                const double r2 = rx * rx + ry * ry + rz * rz;
                take = code >= 1 && code <= 50 && r2 <= T.rlsq[code];
            }
        }
        const bool tu = take && jb > i, td = take && jb < i;
        const unsigned long long up = __ballot(tu), dn = __ballot(td);
        // a bead's pairs are contiguous in the queue: lanes [s, lane) of
        // this lane's segment precede it in its row
        const int ip = __shfl_up(i, 1);
        const unsigned long long hm = __ballot(v && (lane == 0 || ip != i));
        const unsigned long long hl = hm & ((2ull << lane) - 1ull);
        const int sh = hl ? 63 - __clzll(hl) : 0;
        const unsigned long long seg = below & ~((1ull << sh) - 1ull);
        const bool tail = v && (lane + 1 == nb || ((hm >> (lane + 1)) & 1ull));
        int2 b0 = make_int2(0, 0);
        if (v) b0 = cnt[i];
        if (tu) {
            const int p = b0.x + __popcll(up & seg);
            if (p < cap) gs1(S.nb_up + ((size_t)i * cap + p), jb | (stat << 24));
        }
        if (td) {
            const int p = b0.y + __popcll(dn & seg);
            if (p < cap) gs1(S.nb_dn + ((size_t)i * cap + p), jb | (stat << 24));
        }
        const unsigned long long segi = seg | (1ull << lane);
        if (tail) cnt[i] = make_int2(b0.x + __popcll(up & segi), b0.y + __popcll(dn & segi));
        qh = (qh + nb) & (kDenseQ - 1);
        qn -= nb;
    };
    for (int i = wave; i < N; i += nw) {
        const int fi = sfine[i];
        const int ix = fi & 1023, iy = (fi >> 10) & 1023, iz = fi >> 20;
        for (int base = 0; base < N; base += kW) {
            const int jb = base + lane;
            bool pass = false;
            if (jb < N && jb != i) {
                const int fj = sfine[jb];
                // fine cells lie in [0, M): differences in (-M, M), taken mod M
                int dx = (fj & 1023) - ix, dy = ((fj >> 10) & 1023) - iy, dz = (fj >> 20) - iz;
                dx += dx < 0 ? M : 0;
                dy += dy < 0 ? M : 0;
                dz += dz < 0 ? M : 0;
                pass = !((dx > 2 && dx < M - 2) || (dy > 2 && dy < M - 2) || (dz > 2 && dz < M - 2));
            }
            const unsigned long long m = __ballot(pass);
            if (pass) q[(qh + qn + __popcll(m & below)) & (kDenseQ - 1)] = (i << 16) | jb;
            qn += __popcll(m);
            if (qn >= kW) drain();
        }
    }
    while (qn > 0) drain();
    int mx = 0, status = 0;
    for (int i = wave + nw * lane; i < N; i += nw * kW) {
        const int2 n2 = cnt[i];
        gs1(S.nup + (i), min(n2.x, cap));
        gs1(S.ndn + (i), min(n2.y, cap));
        if (n2.x > cap || n2.y > cap) status = kErrListCapacity;
        mx = max(mx, max(n2.x, n2.y));
    }
    PH_END(S, kPhLSearch, trl);
    status = wave_status(status);
    if (status && !c.status) c.status = status;
    mx = wave_imax(mx);
    if (lane == 0) atomicMax(&S.st->max_nb_seen, (long long)mx);
    wg_sync(nw);
}

// events.f:11-127 with tfalse = 0: every bead's next event, then the
// calendar; all wavefronts of the system take part (the master scalars'
// scan position is at bin 1 already)
__device__ void events_all(const SysDev&, const DevTables&, const Lds&, Sc&) {
    const Lds& L = ldsv();

    const SysDev& S = sdev();

    const DevTables& T = tabs();
    Sc& c = *L.msc();

    const int tid = (int)threadIdx.x, nthr = kW * L.nw;
    const int N = S.N, cap = S.cap;
    int status = 0;
    for (int i = tid; i < N; i += nthr) {
        const Rec bi = gld(S.rec + (i));
        const double2 bwi = gld(S.bw + (i));
        double ti = c.interval_max + kLt;
        int pi = -1, ci = -1;
        const int n = gl1(S.nup + (i));
        for (int k = 0; k < n; ++k) {
            const int ent = gl1(S.nb_up + ((size_t)i * cap + k));
            const int m = ent & 0xFFFFFF;
            const Rec bm = gld(S.rec + (m));
            const int code = dyn_code(gld(S.ovr + (m)), i, ent >> 24);
            double tij = 1000000000.0;
            int ty = 0;
            predict(T, code, bi, bm, dyn_bptnr(bi.dyn) == m, bwi, 0.0, tij, ty);
            if (ty == -99) status = kErrBadCode;
            if (tij < ti) {
                ti = tij;
                pi = m;
                ci = ty;
            }
            if (tij < 0.0 && fabs(tij) > 1.0e-10) status = kErrNegativeTime;
        }
        const int4 e = gld(S.extra + (i));
        for (int k = 0; k < 3; ++k) {
            const int m = k == 0 ? e.x : (k == 1 ? e.y : e.z);
            if (m > i) {
                const Rec bm = gld(S.rec + (m));
                const int code = dyn_code(gld(S.ovr + (m)), i, static_t(bm.topo, bi.topo));
                double tij = 1000000000.0;
                int ty = 0;
                // events.f:79-83: code 1 -> core, anything else -> shoulder
                predict(T, code == 1 ? 1 : 40, bi, bm, dyn_bptnr(bi.dyn) == m, bwi, 0.0, tij, ty);
                if (tij < 0.0 && fabs(tij) > 1.0e-10) status = kErrNegativeTime;
                if (tij < ti) {
                    pi = m;
                    ci = ty;
                    ti = tij;
                }
            }
        }
        Leaf lf;
        lf.tim = ti;
        lf.nptnr = pi;
        lf.meta = make_meta(calc_bin(c, ti, false, status), ci);
        leaf_set(L, i, lf);
        stamp_set(i, mk_stamp(c.coll, i + 1));  // events.f:111-122 adds in index order
    }
    for (int e = N + tid; e < N + 3; e += nthr) {
        Leaf lf = leaf_get(L, e);
        lf.meta = make_meta(calc_bin(c, lf.tim, false, status), -2);
        leaf_set(L, e, lf);
        stamp_set(e, mk_stamp(c.coll, e + 1));
    }
    status = wave_status(status);
    if (status && !c.status) c.status = status;
    wg_sync(L.nw);
    rescan_all(L, N + 3);
}

// Rows in nbor.f's order: every bead's up and down rows sorted by
// nbor_key, so that the lowest list position among partners tied in time
// is the one eventredo_up.f's strict < takes in nbor.f's row, and down
// rows are walked in partial_events.f's order.  One lane per row (row r:
// bead r / 2, its up row for even r, its down row for odd r): the row's
// entries and keys in the wavefront's work area (kSortMax each per lane,
// [slot][lane]), the row in up to five 16-byte loads and all its partner
// loads in flight, then every entry written at its rank; a longer row is
// sorted by the whole wavefront, one lane per entry.
constexpr int kSortMax = 20;
static_assert(12 * kSortMax * kW <= 20 * kMinMaxw, "row sort staging exceeds a wavefront's work area");
__device__ void sort_rows(const SysDev&, const DevTables&, const Lds&) {
    const Lds& L = ldsv();
    const SysDev& S = sdev();
    const DevTables& T = tabs();
    const int lane = lane_id(), wave = wave_id(), nw = L.nw, N = S.N, cap = S.cap, M = S.M;
    unsigned long long* const keys = reinterpret_cast<unsigned long long*>(smem + L.o_wm);  // [kSortMax][kW]
    int* const ents = reinterpret_cast<int*>(keys + kSortMax * kW);                        // [kSortMax][kW]
    const bool vec = (cap & 3) == 0;  // rows of whole int4s (row bases are 16-byte aligned)
    for (int r0 = wave * kW; r0 < 2 * N; r0 += nw * kW) {
        const int r = r0 + lane;
        const bool act = r < 2 * N;
        const int x = act ? r >> 1 : 0;
        int* const row = ((r & 1) == 0 ? S.nb_up : S.nb_dn) + (size_t)x * cap;
        const int n = act ? min(gl1(((r & 1) == 0 ? S.nup : S.ndn) + (x)), cap) : 0;
        const int2 cx = gl2(S.cfp + (x));
        if (n >= 2 && n <= kSortMax) {
            if (vec) {
                // reads up to the next multiple of 4 (<= kSortMax, <= cap)
                int4 v[kSortMax / 4];
#pragma unroll
                for (int u = 0; u < kSortMax / 4; ++u)
                    if (4 * u < n) v[u] = gld(reinterpret_cast<const int4*>(row) + u);
#pragma unroll
                for (int u = 0; u < kSortMax / 4; ++u)
                    if (4 * u < n) {
                        ents[(4 * u) * kW + lane] = v[u].x;
                        ents[(4 * u + 1) * kW + lane] = v[u].y;
                        ents[(4 * u + 2) * kW + lane] = v[u].z;
                        ents[(4 * u + 3) * kW + lane] = v[u].w;
                    }
            } else {
                for (int k = 0; k < n; ++k) ents[k * kW + lane] = gl1(row + (k));
            }
            {
                int e[kSortMax];
                int2 cb[kSortMax];
#pragma unroll
                for (int u = 0; u < kSortMax; ++u) e[u] = u < n ? ents[u * kW + lane] & 0xFFFFFF : x;
#pragma unroll
                for (int u = 0; u < kSortMax; ++u)
                    if (u < n) cb[u] = gl2(S.cfp + (e[u]));
#pragma unroll
                for (int u = 0; u < kSortMax; ++u)
                    if (u < n) keys[u * kW + lane] = nbor_key(M, T, cx, cb[u]);
            }
            for (int k = 0; k < n; ++k) {
                const unsigned long long kk = keys[k * kW + lane];
                int rk = 0;
#pragma unroll
                for (int m = 0; m < kSortMax; ++m) rk += m < n && keys[m * kW + lane] < kk ? 1 : 0;
                gs1(row + (rk), ents[k * kW + lane]);
            }
        }
        // rows longer than the staging: the wavefront sorts each in turn
        unsigned long long big = __ballot(n > kSortMax);
        while (big) {
            const int src = __ffsll((long long)big) - 1;
            big &= big - 1;
            const int rb_ = r0 + src, xb = rb_ >> 1, nbig = __shfl(n, src, kW);
            int* const rb = ((rb_ & 1) == 0 ? S.nb_up : S.nb_dn) + (size_t)xb * cap;
            const int2 cxb = gl2(S.cfp + (xb));
            const int e = lane < nbig ? gl1(rb + (lane)) : 0;
            const unsigned long long kk =
                lane < nbig ? nbor_key(M, T, cxb, gl2(S.cfp + (e & 0xFFFFFF))) : ~0ull;
            int rk = 0;
            for (int m = 0; m < nbig; ++m) {
                const unsigned long long km = (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)kk, m) |
                                              ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(kk >> 32), m) << 32);
                rk += km < kk ? 1 : 0;
            }
            wave_sync_mem();  // every lane has read its entry before any is written
            if (lane < nbig) gs1(rb + (rk), e);
            wave_sync_mem();
        }
    }
    wg_sync(nw);
}

// Wrap and lists; all wavefronts of the system take part.
__device__ void wrap_and_lists(const SysDev&, const Lds&, Sc&) {
    const Lds& L = ldsv();

    const SysDev& S = sdev();

    Sc& c = *L.msc();

    const int tid = (int)threadIdx.x, nthr = kW * L.nw;
    if (tid == 0) c.nbin = 1;
    for (int k = tid; k < S.N; k += nthr) {
        Rec b = gld(S.rec + (k));
        b.x = b.x - round(b.x);
        b.y = b.y - round(b.y);
        b.z = b.z - round(b.z);
        gst(S.rec + (k), b);
        gst(S.old_r + (k), make_double4(b.x, b.y, b.z, 0.0));
    }
    wg_sync(L.nw);
    if (S.N <= kDenseN)
        rebuild_lists_dense(S, L, c);
    else
        rebuild_lists(S, L, c);
    sort_rows(S, tabs(), L);
}

// Wrap, lists and full reschedule; all wavefronts of the system take part.
__device__ void rebuild(const SysDev&, const DevTables&, const Lds&, Sc&) {
    const Lds& L = ldsv();
    const SysDev& S = sdev();
    const DevTables& T = tabs();
    Sc& c = *L.msc();
    PH_BEGIN(trb);
    wrap_and_lists(S, L, c);
    PH_END(S, kPhLists, trb);
    events_all(S, T, L, c);
    PH_END(S, kPhEvents, trb);
    PH_COUNT(S, kPhNRebuild);
}

// sync pseudo-event, main.F90:1126-1187, in two parts around the list
// rebuild that all wavefronts run: sync_begin (wavefront 0) advances the
// free flight, shifts the calendar and returns whether the lists are
// rebuilt; sync_finish (wavefront 0) schedules the next sync.
__device__ bool sync_begin(const SysDev&, const DevTables&, const Lds&, Sc&) {
    const Lds& L = ldsv();

    const SysDev& S = sdev();

    const DevTables& T = tabs();
    Sc& c = *L.sc();

    const int lane = lane_id();
    const int N = S.N;
    const double tf = c.tfalse;
    PH_BEGIN(tsy);
    PH_COUNT(S, kPhNSync);
    c.t = c.t + tf;
    for (int e = lane; e < N + 3; e += kW) {
        Leaf lf = leaf_get(L, e);
        lf.tim = lf.tim - tf;
        leaf_set(L, e, lf);
    }
    c.interval_max = c.interval_max - tf;
    c.tbin_off = c.tbin_off + tf;
    double moved_far = 0.0;
    // free flight to real positions and the displacement check (displ.f),
    // four beads per lane in flight at a time
#pragma unroll 4
    for (int k = lane; k < N; k += kW) {
        Rec b = gld(S.rec + (k));
        const double4 o = gld(S.old_r + (k));
        b.x = b.x + b.vx * tf;
        b.y = b.y + b.vy * tf;
        b.z = b.z + b.vz * tf;
        gst(S.rec + (k), b);
        const double a = o.x - b.x, bb = o.y - b.y, cc = o.z - b.z;
        const double dis = a * a + bb * bb + cc * cc;
        const double moved = dis / S.hdelr;
        if (moved > moved_far) moved_far = moved;
    }
    moved_far = wave_max(moved_far);
    c.tfalse = 0.0;
    bool update = false;
    if (moved_far >= 0.1) {  // displ.f:37-46
        update = true;
        if (moved_far >= 1.25 * 1.25) {
            c.t_fact = c.t_fact / 1.01;
            c.interval = c.t_fact / sqrt(S.setemp);
        }
    }
    wave_sync_mem();
    if (update || c.interval > c.interval_max) {
        if (!update) {
            c.nforced++;
            c.n_forced = c.n_forced * 1.01;
        }
        c.interval_max = c.interval * c.n_forced;
        c.sortsize = c.interval_max / (double)kNumbin;
        c.tbin_off = 0.0;
        c.nupdates++;
        PH_END(S, kPhSync, tsy);
        return true;
    } else {
        // every entry moved by the same -tf: the cached block and superblock
        // minima are shifted exactly as their leaves were
        for (int b = lane; b < L.nblk; b += kW)
            if (L.blk()[b].bc != kEmptyBc) L.blk()[b].t = L.blk()[b].t - tf;
        if (lane < L.nsup && L.sup()[lane].bc != kEmptyBc) L.sup()[lane].t = L.sup()[lane].t - tf;
        wave_sync_mem();
    }
    PH_END(S, kPhSync, tsy);
    return false;
}
__device__ void sync_finish(const SysDev&, const Lds&, Sc&) {
    const Lds& L = ldsv();

    const SysDev& S = sdev();

    Sc& c = *L.sc();

    const int lane = lane_id();
    const int N = S.N;
    PH_BEGIN(tsy);
    int status = 0;
    const double t2 = c.interval * 0.999;
    if (lane == 0) {
        Leaf lf;
        lf.tim = t2;
        lf.nptnr = -2;
        lf.meta = make_meta(calc_bin(c, t2, true, status), -2);
        leaf_set(L, N + 1, lf);
        stamp_set(N + 1, mk_stamp(c.coll, N + 4));  // after a rebuild's events() adds
        L.blist()[0] = (N + 1) / kBlk;
    }
    if (status && !c.status) c.status = status;
    if (!L.leaf_lds)
        wave_sync_mem();
    else
        wave_sync();
    rescan_block_list(L, N + 3, 1);
    PH_END(S, kPhSync, tsy);
}

__device__ void output_event(const SysDev&, const Lds&, Sc&) {
    const Lds& L = ldsv();

    const SysDev& S = sdev();

    Sc& c = *L.sc();

    const int lane = lane_id();
    const int N = S.N;
    int status = 0;
    const double t3 = 3.3 / (sqrt(S.setemp)) + 5 + c.tfalse;
    if (lane == 0) {
        Leaf lf;
        lf.tim = t3;
        lf.nptnr = -2;
        lf.meta = make_meta(calc_bin(c, t3, false, status), -2);
        leaf_set(L, N + 2, lf);
        stamp_set(N + 2, mk_stamp(c.coll, 1));
        L.blist()[0] = (N + 2) / kBlk;
    }
    if (status && !c.status) c.status = status;
    if (!L.leaf_lds)
        wave_sync_mem();
    else
        wave_sync();
    rescan_block_list(L, N + 3, 1);
}

// LDS plan: tables, SysDev, the views, the superblock minima, master
// scalars and round record, then each wavefront's work area (wavefront 0's
// is the larger: the list rebuild and the generic rescheduling use it as
// scratch; its scalars are the master copy), then this system's block
// minima, then as many calendar leaves as the workgroup's lds_bytes hold:
// a small system in a launch sized for a large one keeps all its leaves in
// LDS
__device__ Lds make_lds(const SysDev&, int maxw, int lds_bytes, int nw, int wave) {
    const SysDev& S = sdev();

    Lds L;
    const int N3 = S.N + 3;
    L.nblk = (N3 + kBlk - 1) / kBlk;
    L.nsup = (L.nblk + kSup - 1) / kSup;
    L.maxw = maxw;
    L.nw = nw;
    int p = 0;
    auto take = [&](int bytes) {
        const int o = p;
        p = (p + bytes + 15) / 16 * 16;
        return o;
    };
    take((int)sizeof(DevTables));        // the tables, at offset 0
    take((int)sizeof(SysDev));           // the system's SysDev, at kSdOff
    take(kLdStride * kMaxWaves);         // these views, at kLdOff
    L.o_sup = take((int)sizeof(Key) * kSup);
    L.o_msc = take((int)sizeof(Sc));
    L.o_batch = take((int)sizeof(Batch));
    for (int w = 0; w <= wave; ++w) {
        L.o_wm = take((int)sizeof(int) * maxw);
        L.o_wx = take((int)sizeof(int) * maxw);
        L.o_blist = take((int)sizeof(int) * kMaxBlist);
        L.o_wd = take((int)sizeof(int) * (w == 0 ? maxw : 16));
        L.o_wy = take((int)sizeof(int) * maxw);
        L.o_wt = take((int)sizeof(double) * maxw);
        L.o_slot = take((int)sizeof(MinSlot) * (2 + kMaxSpec));
        L.o_spec = take((int)sizeof(SpecInfo) * kMaxSpec);
        L.o_down = take((int)sizeof(DownInfo) * kW);
        L.o_pair = take((int)sizeof(Pair));
        L.o_sc = w == 0 ? L.o_msc : take((int)sizeof(Sc));
    }
    for (int w = wave + 1; w < nw; ++w) {
        take((int)sizeof(int) * maxw);
        take((int)sizeof(int) * maxw);
        take((int)sizeof(int) * kMaxBlist);
        take((int)sizeof(int) * 16);
        take((int)sizeof(int) * maxw);
        take((int)sizeof(double) * maxw);
        take((int)sizeof(MinSlot) * (2 + kMaxSpec));
        take((int)sizeof(SpecInfo) * kMaxSpec);
        take((int)sizeof(DownInfo) * kW);
        take((int)sizeof(Pair));
        take((int)sizeof(Sc));
    }
    L.o_blk = take((int)sizeof(Key) * L.nblk);
    L.o_leaf = p;
    L.leaf_lds = N3 <= (lds_bytes - p) / (int)sizeof(Leaf);
    L.gleaf = S.leaf;
    // footprint marks behind the leaves, when both fit
    const int o_mark = (p + N3 * (int)sizeof(Leaf) + 15) / 16 * 16;
    L.mark_lds = L.leaf_lds && o_mark + 4 * S.N <= lds_bytes;
    L.mark = L.mark_lds ? reinterpret_cast<unsigned*>(smem + o_mark) : S.mark;
    return L;
}

__device__ inline void load_sc(const SysState* st, Sc& c) {
    c.tfalse = st->tfalse;
    c.t = st->t;
    c.interval = st->interval;
    c.interval_max = st->interval_max;
    c.t_fact = st->t_fact;
    c.n_forced = st->n_forced;
    c.sortsize = st->sortsize;
    c.tbin_off = st->tbin_off;
    c.old_tfalse = st->old_tfalse;
    c.coll = st->coll;
    c.numghosts = st->numghosts;
    c.nupdates = st->nupdates;
    c.nforced = st->nforced;
    c.target = st->target;
    c.rc0 = st->rng_ctr[0];
    c.rc1 = st->rng_ctr[1];
    c.nbin = st->nbin;
    c.status = st->status;
    for (int k = 0; k < 32; ++k) c.nev[k] = st->nevents[k];
}
__device__ inline void store_sc(SysState* st, const Sc& c) {
    st->tfalse = c.tfalse;
    st->t = c.t;
    st->interval = c.interval;
    st->interval_max = c.interval_max;
    st->t_fact = c.t_fact;
    st->n_forced = c.n_forced;
    st->sortsize = c.sortsize;
    st->tbin_off = c.tbin_off;
    st->old_tfalse = c.old_tfalse;
    st->coll = c.coll;
    st->numghosts = c.numghosts;
    st->nupdates = c.nupdates;
    st->nforced = c.nforced;
    st->target = c.target;
    st->rng_ctr[0] = c.rc0;
    st->rng_ctr[1] = c.rc1;
    st->nbin = c.nbin;
    st->status = c.status;
    for (int k = 0; k < 32; ++k) st->nevents[k] = c.nev[k];
}
__device__ void copy_tables(const DevTables* src, DevTables* dst) {
    const double* s = (const double*)src;
    double* d = (double*)dst;
    const int n = (int)(sizeof(DevTables) / sizeof(double));
    for (int k = (int)threadIdx.x; k < n; k += (int)blockDim.x) d[k] = s[k];
}

}  // namespace

// ------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------
// Start-up of one system per wavefront: wrap, lists, pseudo-event seeds and
// events() (main.F90:205-424 after the host-side H-bond restore).
// One workgroup of up to kSetupWaves wavefronts per system: the list
// rebuild and events() run on all of them, the rest on the first lane.
constexpr int kSetupWaves = 4;
__global__ __launch_bounds__(64 * kSetupWaves) void k_setup(const SysDev* __restrict__ sys, const int* ids, int maxw) {
    const int nw = (int)(blockDim.x >> 6), wave = wave_id();
    if (threadIdx.x == 0) *reinterpret_cast<SysDev*>(smem + kSdOff) = sys[ids[blockIdx.x]];
    __syncthreads();
    const SysDev& S = sdev();
    if (lane_id() == 0)
        *reinterpret_cast<Lds*>(smem + kLdOff + wave * kLdStride) = make_lds(S, maxw, 0, nw, wave);
    __syncthreads();
    const Lds& L = ldsv();
    copy_tables(S.tab, const_cast<DevTables*>(&tabs()));
    __syncthreads();
    const DevTables& T = tabs();
    Sc& c = *L.sc();
    if (threadIdx.x == 0) load_sc(S.st, c);
    __syncthreads();
    const int N = S.N;
    if (threadIdx.x == 0) {
        Leaf g, s2, s3;
        if (S.canon) {
            double tgho = 0.0;
            int guard = 0;
            do {
                tgho = rng_u01(S.seed, 0, c.rc0++);
            } while ((tgho < 1.e-18 || tgho == 1.0) && ++guard < 1000);
            g.tim = -1.0 * dlog(tgho) * S.avegtime * .0000001;
        } else
            g.tim = 1000000000.0;
        s2.tim = c.interval;
        s3.tim = 3.3 / (sqrt(S.setemp)) + 5;
        g.nptnr = s2.nptnr = s3.nptnr = -2;
        g.meta = s2.meta = s3.meta = make_meta(-1, -2);
        gst(S.leaf + (N), g);
        gst(S.leaf + (N + 1), s2);
        gst(S.leaf + (N + 2), s3);
    }
    __threadfence_block();
    __syncthreads();
    rebuild(S, T, L, c);
    __syncthreads();
    if (threadIdx.x == 0) store_sc(S.st, c);
}

// Workgroup entry of the rebuild kernels: the system's descriptor, the work
// areas and the tables into LDS; the first lane loads the master scalars
// (each wavefront's own copy is not loaded here).
__device__ void enter_system(const SysDev* __restrict__ sys, const int* ids, int maxw) {
    const int nw = (int)(blockDim.x >> 6), wave = wave_id();
    if (threadIdx.x == 0) *reinterpret_cast<SysDev*>(smem + kSdOff) = sys[ids[blockIdx.x]];
    __syncthreads();
    const SysDev& S = sdev();
    if (lane_id() == 0)
        *reinterpret_cast<Lds*>(smem + kLdOff + wave * kLdStride) = make_lds(S, maxw, 0, nw, wave);
    __syncthreads();
    copy_tables(S.tab, const_cast<DevTables*>(&tabs()));
    if (threadIdx.x == 0) load_sc(S.st, *ldsv().msc());
    __syncthreads();
}

// A list update forced at the current time (pdmd_rebuild), in two
// dispatches so that the list build and the full reschedule are timed and
// counted apart.  This first one runs the sync pseudo-event's shift of the
// time origin and its update branch (main.F90:1126-1187) up to the new
// lists, on every wavefront of the system's workgroup.  No displacement
// test, no rate adaptation and no new sync entry: the pending one keeps its
// shifted time.
__global__ __launch_bounds__(64 * kSetupWaves) void k_rebuild_lists(const SysDev* __restrict__ sys, const int* ids,
                                                                    int maxw) {
    enter_system(sys, ids, maxw);
    const SysDev& S = sdev();
    const Lds& L = ldsv();
    Sc& c = *L.msc();
    const int N = S.N;
    const double tf = c.tfalse;
    __syncthreads();  // every wavefront holds tf before the first lane moves the origin
    for (int e = (int)threadIdx.x; e < N + 3; e += (int)blockDim.x) {
        Leaf lf = gld(S.leaf + (e));
        lf.tim = lf.tim - tf;
        gst(S.leaf + (e), lf);
    }
    for (int k = (int)threadIdx.x; k < N; k += (int)blockDim.x) {
        Rec b = gld(S.rec + (k));
        b.x = b.x + b.vx * tf;
        b.y = b.y + b.vy * tf;
        b.z = b.z + b.vz * tf;
        gst(S.rec + (k), b);
    }
    if (threadIdx.x == 0) {
        c.t = c.t + tf;
        c.tfalse = 0.0;
        c.interval_max = c.interval * c.n_forced;
        c.sortsize = c.interval_max / (double)kNumbin;
        c.tbin_off = 0.0;
        c.nupdates++;
    }
    __threadfence_block();
    __syncthreads();
    wrap_and_lists(S, L, c);
    __syncthreads();
    if (threadIdx.x == 0) store_sc(S.st, c);
}

// The second dispatch of a forced list update: events() over the new lists
// (events.f), every bead's next event and the pseudo-events' bins.
__global__ __launch_bounds__(64 * kSetupWaves) void k_rebuild_events(const SysDev* __restrict__ sys, const int* ids,
                                                                     int maxw) {
    enter_system(sys, ids, maxw);
    const SysDev& S = sdev();
    const Lds& L = ldsv();
    Sc& c = *L.msc();
    events_all(S, tabs(), L, c);
    __syncthreads();
    if (threadIdx.x == 0) store_sc(S.st, c);
}

// Pop up to kmax calendar entries in key order (wavefront 0): pair events
// only after the first, since a pseudo-event runs alone.  Each popped entry
// but the last is taken out of the min tree until the round's rescans.
// Every speculative wavefront k gets a copy of the master scalars (which it
// sets to its event's time and bin after the round's first barrier); the
// master takes the first event.  Returns the number popped (0: stop, or an
// empty calendar).
__device__ int select_batch(const Lds&, Sc& c, int N, int N3, int kmax) {
    const Lds& L = ldsv();
    const SysDev& S = sdev();
    (void)S;
    PH_BEGIN(tsp);

    Batch& B = *L.batch();
    const int lane = lane_id();
    // the views this loop uses, read once (wave fences below would make the
    // compiler read them from LDS again)
    Key* const blk = L.blk();
    Key* const sup = L.sup();
    const int nsup = L.nsup, nblk = L.nblk;
    const bool leaf_lds = L.leaf_lds;
    const Leaf* const lleaf = L.lleaf();
    const Leaf* const gleaf = L.gleaf;
    int K = 0;
    unsigned pip = 0xFFFFFFFFu;  // lane m: index << 16 | partner of popped entry m
    // one superblock (up to 1024 entries): the block minima stay in lane
    // registers (lane b: block b) and an exclusion replaces one of them; the
    // tree in LDS keeps its values, which the round's rescans of the popped
    // blocks confirm or replace
    const bool one_sup = nsup == 1;
    Key bmin = empty_key();
    if (one_sup && lane < nblk) bmin = blk[lane];
    while (true) {
        const Key k = one_sup ? key_min_wave(bmin) : key_min_wave(lane < nsup ? sup[lane] : empty_key());
        if (k.bc == kEmptyBc) {
            if (K == 0) c.status = kErrBin;
            break;
        }
        const int e = key_idx(k);
        if (K > 0 && e >= N) break;
        if (lane == 0) B.key[K] = k;
        if (lane == K) pip = k.ip;
        ++K;
        // a ghost event may lead a round (its draws come first in the
        // serial order); sync and output events run alone
        if (e > N || K >= kmax) break;
        // take e out of the tree: its block without the popped entries, then
        // its superblock
        const int b = e / kBlk, sp = b / kSup;
        Key x = empty_key();
        const int idx = b * kBlk + (lane & 15);
        if (lane < kBlk && idx < N3) {
            bool popped = false;
            for (int m = 0; m < K; ++m) popped |= (int)(__builtin_amdgcn_readlane((int)pip, m) >> 16) == idx;
            if (!popped) x = leaf_key(leaf_lds ? lleaf[idx] : gld(gleaf + idx), idx);
        }
        const bool win = key_min_rows(x);
        if (one_sup) {
            const unsigned long long wm = __ballot(win && lane < kBlk);
            const Key nb = readlane_key(x, __ffsll((long long)wm) - 1);
            if (lane == b) bmin = nb;
            continue;
        }
        if (win && lane < kBlk) blk[b] = x;
        wave_sync();
        const int bb = sp * kSup + lane;
        const Key y = key_min_wave(bb < nblk ? blk[bb] : empty_key());
        if (lane == 0) sup[sp] = y;
        wave_sync();
    }
    PH_END(S, kPhSelPop, tsp);
    wave_sync();
    if (K == 0) return 0;
    const long long cw = lane < (int)(sizeof(Sc) / 8) ? reinterpret_cast<const long long*>(&c)[lane] : 0;
    for (int k = 1; k < K; ++k) {
        const Lds& Lk = *reinterpret_cast<const Lds*>(smem + kLdOff + k * kLdStride);
        if (lane < (int)(sizeof(Sc) / 8)) reinterpret_cast<long long*>(Lk.sc())[lane] = cw;
    }
    wave_sync();
    const Key k0 = B.key[0];
    c.old_tfalse = c.tfalse;
    c.nbin = key_bin(k0);
    c.tfalse = k0.t;
    c.coll++;
    PH_END(S, kPhSelCopy, tsp);
    return K;
}

__device__ inline void log_event(const SysDev& S, int& log_count, long long coll, int e, int jj, int ty, double t) {
    if (S.log && log_count < S.log_cap) {
        if (lane_id() == 0) {
            LogRec r;
            r.coll = coll;
            r.i = e + 1;
            r.j = jj + 1;
            r.type = ty;
            r.pad = 0;
            r.t = t;
            gst(S.log + (log_count), r);
        }
        ++log_count;
    }
}

// The event loop: each workgroup commits events of its system until the
// system's coll counter has advanced by inc[k].  With one wavefront per
// workgroup each round is one event.  With W wavefronts a round pops up to
// W entries; pair events run side by side, one per wavefront, and the
// round commits the prefix of them that the serial order would have run
// the same way (disjoint footprints, no new entry ahead of a later one).
// A wall-clock budget (ticks of wall_clock64, 0 = none) stops a system early
// between rounds, so that one launch over systems of uneven speed ends
// evenly; flags & kStopAtOutput stops it right after its next output
// pseudo-event (where the reference writes observables and configurations).
__global__ __launch_bounds__(64 * kMaxWaves) void k_run(const SysDev* __restrict__ sys, const int* ids, const long long* inc,
                                                       int maxw, int lds_bytes, long long budget, int flags,
                                                       unsigned long long* launch_t0) {
    const int nw = (int)(blockDim.x >> 6), wave = wave_id(), lane = lane_id();
    if (threadIdx.x == 0) *reinterpret_cast<SysDev*>(smem + kSdOff) = sys[ids[blockIdx.x]];
    __syncthreads();
    const SysDev& S = sdev();
    if (lane == 0)
        *reinterpret_cast<Lds*>(smem + kLdOff + wave * kLdStride) = make_lds(S, maxw, lds_bytes, nw, wave);
    __syncthreads();
    const Lds& L = ldsv();
    const int N = S.N, N3 = N + 3;
    copy_tables(S.tab, const_cast<DevTables*>(&tabs()));
    if (L.leaf_lds)
        for (int e = (int)threadIdx.x; e < N3; e += (int)blockDim.x) L.lleaf()[e] = gld(S.leaf + (e));
    if (L.mark_lds)
        for (int b = (int)threadIdx.x; b < N; b += (int)blockDim.x) L.mark[b] = 0u;
    __syncthreads();
    const DevTables& T = tabs();
    Sc& c = *L.sc();  // wavefront 0: the master scalars; others: their own copy
    Batch& B = *L.batch();
    if (threadIdx.x == 0) {
        load_sc(S.st, c);
        c.target = c.coll + inc[blockIdx.x];
    }
    __syncthreads();
    rescan_all(L, N3);
    int log_count = S.st->log_count;
    // the budget runs from the launch's start: the earliest start among the
    // workgroups (launch_t0 is reset to all ones before a budgeted launch),
    // so that a workgroup dispatched late does not get a later deadline
    long long w0 = wall_clock64();
    const long long t_begin = w0;
    long long t_rebuild = 0, n_rebuild = 0;
    // the system's own limit on pair events per round (its tuned wavefronts)
    const int kcap = S.waves_cap > 0 && S.waves_cap < nw ? S.waves_cap : nw;
    if (budget > 0) {
        if (threadIdx.x == 0) {
            atomicMin(launch_t0, (unsigned long long)w0);
            B.t0 = (long long)__hip_atomic_load(launch_t0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        w0 = B.t0;
    }
    bool stop_out = false;
    while (true) {
        if (wave == 0) {
            PH_BEGIN(tsel);
            int K = 0;
            if (c.coll < c.target && c.status == 0 && !stop_out && !(budget > 0 && wall_clock64() - w0 > budget)) {
                long long left = c.target - c.coll;
                K = select_batch(L, c, N, N3, (int)min((long long)kcap, left));
            }
            if (lane == 0) {
                B.K = K;
                B.cut = K;
            }
            PH_END(S, kPhSelect, tsel);
        }
        if (nw > 1)
            __syncthreads();
        else
            wave_sync();
        const int K = B.K;
        if (K == 0) break;
        if (K == 1) {
            const Key k = B.key[0];
            const int e = key_idx(k);
            if (e == N + 1) {
                // the sync event; every wavefront takes part in a list rebuild
                if (wave == 0) {
                    const bool rb = sync_begin(S, T, L, c);
                    if (lane == 0) B.rebuild = rb ? 1 : 0;
                }
                if (nw > 1)
                    __syncthreads();
                else
                    wave_sync();
                if (B.rebuild) {
                    const long long tr = wall_clock64();
                    rebuild(S, T, L, c);
                    t_rebuild += wall_clock64() - tr;
                    ++n_rebuild;
                }
                if (wave == 0) {
                    sync_finish(S, L, c);
                    log_event(S, log_count, c.coll, e, -1, 0, k.t);
                }
            } else if (wave == 0) {
                int jj = -1, ty = 0;
                if (e < N) {
                    jj = key_p(k);
                    ty = pair_event(S, T, L, c, e, jj, key_ct(k), -1);
                    if (ty >= 1 && ty <= 30 && lane == 0) c.nev[ty]++;
                } else if (e == N) {
                    jj = ghost_event(S, T, L, c, -1);
                } else {
                    output_event(S, L, c);
                    stop_out = (flags & kStopAtOutput) != 0;
                }
                log_event(S, log_count, c.coll, e, jj, ty, k.t);
            }
            if (nw > 1) __syncthreads();
            continue;
        }
        // a round of K >= 2 pair events, one per wavefront; a speculative
        // wavefront's scalars are the master's as its event comes up
        if (wave >= 1 && wave < K && lane == 0) {
            // the copy was taken before the master counted the round's first
            // event: the event number comes from the master, so that every
            // event of the round stamps its adds with its own number (equal
            // stamps would order tied entries by index, not last-added first)
            Sc& ck = *L.sc();
            ck.old_tfalse = B.key[wave - 1].t;
            ck.tfalse = B.key[wave].t;
            ck.nbin = key_bin(B.key[wave]);
            ck.coll = L.msc()->coll + wave;
            ck.status = 0;
        }
        wave_sync();
        if (wave < K) {
            const Key k = B.key[wave];
            if (key_idx(k) == N)
                ghost_event(S, T, L, c, 0);
            else
                pair_event(S, T, L, *L.sc(), key_idx(k), key_p(k), key_ct(k), wave);
        } else {
            int nb = 0;
            bar_to(nb, kRoundBars);
        }
        PH_BEGIN(ttail);
        const int A = batch_accepted(B);
        if (wave == 0) {
            PH_COUNT(S, kPhNRound);
            PH_ADDN(S, kPhNPopped, K);
            PH_ADDN(S, kPhNCommit, A);
            PH_ADDN(S, kPhNCutMark, min(K, B.cut) < K ? 1 : 0);
            PH_ADDN(S, kPhNCutKey, A < min(K, B.cut) ? 1 : 0);
        }
        __syncthreads();
        // block minima of every committed event's blocks and of the popped
        // entries that were not committed, then the superblock minima
        if (wave == 0) {
            int nbl = B.nbl[0];
            for (int k = A; k < K; ++k) {
                if (lane == 0) L.blist()[nbl] = key_idx(B.key[k]) / kBlk;
                ++nbl;
            }
            if (lane == 0) B.nbl[0] = nbl;
            wave_sync();
        }
        if (wave < A) rescan_blocks(L, N3, L.blist(), B.nbl[wave]);
        // a single superblock is never read: selection pops from the block
        // minima (its value is recomputed by the single-event paths)
        if (L.nsup > 1) {
            __syncthreads();
            if (wave < A) rescan_sups(L, L.blist(), B.nbl[wave]);
        }
        if (wave == 0) {
            // the master takes the committed events in order
            for (int k = 0; k < A; ++k) {
                const Key kk = B.key[k];
                const int ty = B.ty[k];
                if (ty >= 1 && ty <= 30 && lane == 0) c.nev[ty]++;
                log_event(S, log_count, c.coll + k, key_idx(kk), key_idx(kk) == N ? B.gbead : key_p(kk), ty, kk.t);
            }
            if (A >= 2) {
                if (lane == 0) {
                    for (int k = 1; k < A; ++k)
                        if (B.st[k] != 0 && c.status == 0) c.status = B.st[k];
                    c.coll += A - 1;
                    c.old_tfalse = B.key[A - 2].t;
                    c.tfalse = B.key[A - 1].t;
                    c.nbin = key_bin(B.key[A - 1]);
                }
            }
        }
        __syncthreads();
        PH_END(S, kPhBTail, ttail);
    }
    __syncthreads();
    if (L.leaf_lds)
        for (int e2 = (int)threadIdx.x; e2 < N3; e2 += (int)blockDim.x) gst(S.leaf + (e2), L.lleaf()[e2]);
    if (threadIdx.x == 0) {
        S.st->log_count = log_count;
        store_sc(S.st, c);
        S.st->run_ticks += wall_clock64() - t_begin;
        S.st->rebuild_ticks += t_rebuild;
        S.st->rebuilds += n_rebuild;
    }
}

// ------------------------------------------------------------------------
// debug checks of one system's state (one lane per bead i)
// ------------------------------------------------------------------------
// repuls_check.f:16-81 on a state array (n = the N bead, cb = the C bead);
// 11 when a shoulder bead falls outside the system
__device__ double check_rating(const DevTables& T, const Rec* rec, int N, int n, int cb, double tf) {
    const Rec rn = rec[n], rc = rec[cb];
    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(n, topo_len(rn.topo), cb, topo_len(rc.topo), ncim1, ncai, ncaj, nnjp1);
    if (ncim1 < 0 || ncim1 >= N || ncai < 0 || ncai >= N || ncaj < 0 || ncaj >= N || nnjp1 < 0 || nnjp1 >= N)
        return 11.0;
    const double d1 = dist_rec(rn, rec[ncaj], tf), d2 = dist_rec(rn, rec[nnjp1], tf);
    const double d3 = dist_rec(rc, rec[ncai], tf), d4 = dist_rec(rc, rec[ncim1], tf);
    if (d1 > T.shder[1] && d2 > T.shder[2] && d3 > T.shder[3] && d4 > T.shder[4]) return 10.0;
    return 11.0;
}
// out[0..2]: checkover.f:21-131 over all pairs (i < j) at tfalse: cores
// closer than sigma (codes 1 and >= 15, squeezed for 22-26), bonds longer /
// shorter than their window (4-12); out[3]: least i*N + j among them;
// out[4..7]: check_nc_int.f:25-354 over i's listed code-15 partners:
// boundbad, unboundbad, n_ss, m_ss.  alt: positions and velocities to check
// in place of the resident ones (topology and codes stay the system's).
__global__ void k_check(const SysDev* sys, int id, const Rec* alt, unsigned long long* out) {
    const SysDev& S = sys[id];
    const DevTables& T = *S.tab;
    const int N = S.N, cap = S.cap;
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= N) return;
    const Rec* rec = alt ? alt : S.rec;
    const double tf = S.st->tfalse;
    const Rec ri = rec[i];
    const int idi = dyn_ident(ri.dyn);
    unsigned long long over = 0, lng = 0, sht = 0, first = ~0ull;
    for (int j = i + 1; j < N; ++j) {
        const Rec rj = rec[j];
        const int idj = dyn_ident(rj.dyn);
        const int code = dyn_code(S.ovr[j], i, static_t(rj.topo, ri.topo));  // ev_code(j, i)
        if (code == 1 || code >= 15) {
            const double pe = T.ev_param[1][code];
            double sigsq = T.sigma_sq[idi][idj] * (pe * pe);
            if (code >= 22 && code <= 26) {
                const double q = T.sqz610[code - 21][idi > idj ? idi : idj];
                sigsq = sigsq * (q * q);
            }
            Geo g;
            geo(ri, rj, tf, g);
            const double r2 = g.r2 * 1.0000000001;
            if (r2 <= sigsq) {
                ++over;
                first = min(first, (unsigned long long)i * (unsigned long long)N + (unsigned long long)j);
            }
        } else if (code >= 4 && code <= 12 && !S.chaptype2) {  // checkover.f:67: chaptype 1 only
            double blmin, blmax;
            bond_window(T, code, S.bw[i], blmin, blmax);
            blmin = blmin * blmin;
            blmax = blmax * blmax;
            Geo g;
            geo(ri, rj, tf, g);
            const bool lo = g.r2 * 1.0000000001 < blmin, hi = g.r2 * 0.9999999999 > blmax;
            if (hi) ++lng;
            if (lo) ++sht;
            if (lo || hi) first = min(first, (unsigned long long)i * (unsigned long long)N + (unsigned long long)j);
        }
    }
    unsigned long long bb = 0, ub = 0, nss = 0, mss = 0;
    const int nu = S.nup[i];
    const int4 xi = S.extra[i];
    for (int k = 0; k < nu && k < cap; ++k) {
        const int ent = S.nb_up[(size_t)i * cap + k];
        const int j = ent & 0xFFFFFF;
        if (j < 0 || j >= N) continue;
        const Rec rj = rec[j];
        const int idj = dyn_ident(rj.dyn);
        if (dyn_code(S.ovr[j], i, ent >> 24) != 15) continue;
        const bool term = terminal_t(ri.topo) || terminal_t(rj.topo);
        if (dyn_bptnr(ri.dyn) == j) {
            const double rating = term ? 10.0 : (idi < idj ? check_rating(T, rec, N, i, j, tf) : check_rating(T, rec, N, j, i, tf));
            if (rating > 10.0) ++bb;
            if (!term && xi.w == j) ++nss;
        } else {
            Geo g;
            geo(ri, rj, tf, g);
            if (g.r2 - T.welldia_sq[idi][idj] < 0.0) {
                const double rating = term ? 11.0 : (idi < idj ? check_rating(T, rec, N, i, j, tf) : check_rating(T, rec, N, j, i, tf));
                if (rating <= 10.0) ++ub;
                if (!term && xi.w == j) ++nss;
            }
        }
        if (xi.w == j) ++mss;
    }
    if (over) atomicAdd(out + 0, over);
    if (lng) atomicAdd(out + 1, lng);
    if (sht) atomicAdd(out + 2, sht);
    if (first != ~0ull) atomicMin(out + 3, first);
    if (bb) atomicAdd(out + 4, bb);
    if (ub) atomicAdd(out + 5, ub);
    if (nss) atomicAdd(out + 6, nss);
    if (mss) atomicAdd(out + 7, mss);
}

// ------------------------------------------------------------------------
// observables of one system: energy.f:25-101, radgyr.f:18-135,
// end2end.f:14-44, on one workgroup of kObsT lanes
// ------------------------------------------------------------------------
constexpr int kObsT = 256;
// sum over the workgroup in a fixed tree (the same result on every call)
__device__ inline double obs_sum(double v, double* red) {
    const int t = (int)threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int w = kObsT / 2; w > 0; w >>= 1) {
        if (t < w) red[t] = red[t] + red[t + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}
// Bead sums in a fixed order per lane (kinetic energy, H-bond counts,
// side-chain well energies over the up lists: every side-chain pair inside
// a well is listed, wells lie inside the list cutoff); chain walks one lane
// per chain (backbone unwrapping into q, then radius of gyration and
// end-to-end distance).  nfirst: chains of the leading species block.
// outd: ered, tred, sumvel, ehh_ii, ehh_ij, rg_avg, e2e_avg; outi: hb_alpha,
// hb_ii, hb_ij.
__global__ __launch_bounds__(kObsT) void k_obs(const SysDev* sys, int id, const int* cstart, const int* clen, int nch,
                                              int nfirst, double boxl, double* outd, int* outi) {
    __shared__ double red[kObsT];
    const SysDev& S = sys[id];
    const DevTables& T = *S.tab;
    const int N = S.N, cap = S.cap, t = (int)threadIdx.x;
    const double tf = S.st->tfalse;
    double sumvel = 0.0, eii = 0.0, eij = 0.0;
    double hii = 0.0, hij = 0.0, halpha = 0.0;
    for (int i = t; i < N; i += kObsT) {
        const Rec r = S.rec[i];
        sumvel = sumvel + r.bm * (r.vx * r.vx + r.vy * r.vy + r.vz * r.vz);
        const int j = dyn_bptnr(r.dyn);
        if (j > i) {
            if (topo_chain(r.topo) == topo_chain(S.rec[j].topo))
                hii = hii + 1.0;
            else
                hij = hij + 1.0;
        }
        const int kind = topo_kind(r.topo), res = topo_res(r.topo), L = topo_len(r.topo);
        if (kind == 1 && res >= 5 && j == i + L - 4) halpha = halpha + 1.0;
        if (kind == 3) {
            const int nu = min(S.nup[i], cap), idi = dyn_ident(r.dyn);
            for (int k = 0; k < nu; ++k) {
                const int ent = S.nb_up[(size_t)i * cap + k];
                const int m = ent & 0xFFFFFF;
                if ((ent >> 24) != 16 || m == j) continue;
                const Rec rm = S.rec[m];
                Geo g;
                geo(r, rm, tf, g);
                const int idm = dyn_ident(rm.dyn);
                if (g.r2 <= T.welldia_sq[idi][idm]) {
                    if (topo_chain(r.topo) == topo_chain(rm.topo))
                        eii = eii + T.ep_sqrt[idi][idm];
                    else
                        eij = eij + T.ep_sqrt[idi][idm];
                }
            }
        }
    }
    // backbone unwrapping along each chain (radgyr.f:48-104) into q
    double4* q = S.spos;
    for (int c = t; c < nch; c += kObsT) {
        const int s = cstart[c], L = clen[c];
        // every bead of the chain: the second species' sums below can reach
        // into an earlier chain's side-chain beads
        for (int k = s; k < s + 4 * L; ++k) {
            const Rec b = S.rec[k];
            q[k] = make_double4(b.x + b.vx * tf, b.y + b.vy * tf, b.z + b.vz * tf, 0.0);
        }
        auto unwrap = [&](int a, int b) {
            double4 qa = q[a];
            const double4 qb = q[b];
            double dd = qa.x - qb.x;
            if (dd > 0.5) qa.x = qa.x - 1.0;
            else if (dd < -1.0 * 0.5) qa.x = qa.x + 1.0;
            dd = qa.y - qb.y;
            if (dd > 0.5) qa.y = qa.y - 1.0;
            else if (dd < -1.0 * 0.5) qa.y = qa.y + 1.0;
            dd = qa.z - qb.z;
            if (dd > 0.5) qa.z = qa.z - 1.0;
            else if (dd < -1.0 * 0.5) qa.z = qa.z + 1.0;
            q[a] = qa;
        };
        for (int k = s + 1; k < s + L; ++k) unwrap(k, k - 1);
        for (int k = s; k < s + L; ++k) unwrap(L + k, k);
        for (int k = s; k < s + L; ++k) unwrap(2 * L + k, k);
    }
    __threadfence_block();
    __syncthreads();
    double rg = 0.0, e2e = 0.0, nbb = 0.0;
    for (int c = t; c < nch; c += kObsT) {
        const int s = cstart[c], L = clen[c];
        // radgyr.f:117-125 sums a chain of the second species block over
        // beads (c-1)*numbeads2+1 .. +3*chnln2 without the first block's
        // offset its unwrapping uses (:78), while those beads exist
        int rb = s;
        if (c >= nfirst && c * 4 * L + 3 * L <= s + 4 * L) rb = c * 4 * L;
        double cx = 0, cy = 0, cz = 0, rgs = 0;
        for (int k = rb; k < rb + 3 * L; ++k) {
            const double4 v = q[k];
            cx = cx + v.x;
            cy = cy + v.y;
            cz = cz + v.z;
        }
        cx = cx / (double)(3 * L);
        cy = cy / (double)(3 * L);
        cz = cz / (double)(3 * L);
        for (int k = rb; k < rb + 3 * L; ++k) {
            const double4 v = q[k];
            const double rx = (v.x - cx) * (v.x - cx), ry = (v.y - cy) * (v.y - cy), rz = (v.z - cz) * (v.z - cz);
            rgs = rgs + rx + ry + rz;
        }
        rg = rg + rgs;
        nbb = nbb + (double)(3 * L);
        e2e = e2e + dist_rec(S.rec[s + L], S.rec[s + 3 * L - 1], tf);  // end2end.f: N of residue 1 to C of residue L
    }
    sumvel = obs_sum(sumvel, red);
    eii = obs_sum(eii, red);
    eij = obs_sum(eij, red);
    hii = obs_sum(hii, red);
    hij = obs_sum(hij, red);
    halpha = obs_sum(halpha, red);
    rg = obs_sum(rg, red);
    e2e = obs_sum(e2e, red);
    nbb = obs_sum(nbb, red);
    if (t == 0) {
        const double eps_hb = T.ep_sqrt[5][8];
        const int sum_hb = (int)hii + (int)hij;
        const double sumeps = -(sum_hb * eps_hb + (eii + eij));
        outd[0] = 0.5 * sumvel + sumeps;
        outd[1] = sumvel / 3.0 / (double)N;
        outd[2] = sumvel;
        outd[3] = eii;
        outd[4] = eij;
        outd[5] = sqrt(rg / nbb) * boxl;
        outd[6] = e2e / (double)nch * boxl;
        outi[0] = (int)halpha;
        outi[1] = (int)hii;
        outi[2] = (int)hij;
    }
}

// test hook: predicted collision time of listed pairs (1-based in, code
// from ev_code(j,i) as events.f/eventredo_up.f read it)
__global__ void k_predict(const SysDev* sys, int id, int n, const int* pi, const int* pj, double* tij, int* type) {
    const SysDev& S = sys[id];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = pi[k] - 1, j = pj[k] - 1;
    if (i < 0 || j < 0 || i >= S.N || j >= S.N || i == j) {
        tij[k] = 0.0;
        type[k] = -100;
        return;
    }
    const Rec ri = gld(S.rec + (i)), rj = gld(S.rec + (j));
    const int code = dyn_code(gld(S.ovr + (j)), i, static_t(rj.topo, ri.topo));
    double t = 1000000000.0;
    int ty = 0;
    predict(*S.tab, code, ri, rj, dyn_bptnr(ri.dyn) == j, gld(S.bw + (i)), S.st->tfalse, t, ty);
    tij[k] = t;
    type[k] = ty;
}

// also sized for the list rebuild's per-lane cell ranges (27 x 64 int2 over
// the work-list arrays)
int engine_maxw(int cap) { return max((4 + kMaxSpec) * (cap + 3), kMinMaxw); }

static size_t r16(size_t b) { return (b + 15) / 16 * 16; }
// the bytes make_lds lays out before the leaves, for nw wavefronts
static size_t lds_fixed(int nblk_max, int maxw, int nw) {
    size_t b = r16(sizeof(DevTables)) + r16(sizeof(SysDev)) + (size_t)kLdStride * kMaxWaves +
               r16(sizeof(Key) * nblk_max) + r16(sizeof(Key) * kSup) + r16(sizeof(Sc)) + r16(sizeof(Batch));
    const size_t common = 3 * r16(sizeof(int) * maxw) + r16(sizeof(int) * kMaxBlist) + r16(sizeof(double) * maxw) +
                          r16(sizeof(MinSlot) * (2 + kMaxSpec)) + r16(sizeof(SpecInfo) * kMaxSpec) +
                          r16(sizeof(DownInfo) * kW) + r16(sizeof(Pair));
    b += common + r16(sizeof(int) * maxw);                                          // wavefront 0
    b += (size_t)(nw - 1) * (common + r16(sizeof(int) * 16) + r16(sizeof(Sc)));   // the others
    return b;
}
size_t engine_lds_bytes(int nblk_max, int maxw, int leaf_cap, int nw) {
    return lds_fixed(nblk_max, maxw, nw) + sizeof(Leaf) * (size_t)leaf_cap;
}
static int nblk_for(int n) { return (n + 3 + kBlk - 1) / kBlk; }

hipError_t engine_setup(const SysDev* d_sys, const int* d_ids, int n, int max_n, int maxw, size_t lds_limit,
                        hipStream_t s) {
    int nw = kSetupWaves;
    while (nw > 1 && engine_lds_bytes(nblk_for(max_n), maxw, 0, nw) > lds_limit) --nw;
    const size_t lds = engine_lds_bytes(nblk_for(max_n), maxw, 0, nw);
    hipLaunchKernelGGL(k_setup, dim3(n), dim3(kW * nw), lds, s, d_sys, d_ids, maxw);
    return hipGetLastError();
}
hipError_t engine_rebuild(const SysDev* d_sys, const int* d_ids, int n, int max_n, int maxw, size_t lds_limit,
                          hipStream_t s) {
    int nw = kSetupWaves;
    while (nw > 1 && engine_lds_bytes(nblk_for(max_n), maxw, 0, nw) > lds_limit) --nw;
    const size_t lds = engine_lds_bytes(nblk_for(max_n), maxw, 0, nw);
    hipLaunchKernelGGL(k_rebuild_lists, dim3(n), dim3(kW * nw), lds, s, d_sys, d_ids, maxw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rebuild_events, dim3(n), dim3(kW * nw), lds, s, d_sys, d_ids, maxw);
    return hipGetLastError();
}
// clears the start word of a budgeted launch (a kernel rather than a memset,
// so that a captured graph replays it faithfully)
__global__ void k_clock_reset(unsigned long long* w) { *w = ~0ull; }

hipError_t engine_run(const SysDev* d_sys, const int* d_ids, const long long* d_inc, int n, int max_n, int maxw,
                      size_t lds_limit, hipStream_t s, long long budget_ticks, int flags, int waves,
                      unsigned long long* launch_t0) {
    const int nblk = nblk_for(max_n);
    int nw = waves < 1 ? 1 : (waves > kMaxWaves ? kMaxWaves : waves);
    while (nw > 1 && engine_lds_bytes(nblk, maxw, 0, nw) > lds_limit) --nw;
    const size_t base = engine_lds_bytes(nblk, maxw, 0, nw);
    // up to one workgroup per compute unit: each takes the whole LDS
    // budget, so every system keeps as many leaves in LDS as fit beside
    // its own block minima; beyond that, the size the largest system needs,
    // so that several workgroups share a compute unit
    static int ncu = 0;
    if (ncu == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
            ncu = v > 0 ? v : 1;
        else
            ncu = 1;
    }
    size_t lds = lds_limit;
    if (n > ncu) {
        size_t want = base + sizeof(Leaf) * (size_t)(max_n + 3);
        lds = want < lds_limit ? want : lds_limit;
    }
    if (lds < base) lds = base;
    if (budget_ticks > 0) {
        if (!launch_t0) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_clock_reset, dim3(1), dim3(1), 0, s, launch_t0);
    }
    hipLaunchKernelGGL(k_run, dim3(n), dim3(kW * nw), lds, s, d_sys, d_ids, d_inc, maxw, (int)lds, budget_ticks, flags,
                       launch_t0);
    return hipGetLastError();
}
int engine_max_waves(int max_n, int maxw, size_t lds_limit) {
    int nw = kMaxWaves;
    while (nw > 1 && engine_lds_bytes(nblk_for(max_n), maxw, 0, nw) > lds_limit) --nw;
    return nw;
}
hipError_t engine_predict(const SysDev* d_sys, int id, int n, const int* pi, const int* pj, double* tij, int* type,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_predict, dim3((n + 255) / 256), dim3(256), 0, s, d_sys, id, n, pi, pj, tij, type);
    return hipGetLastError();
}
hipError_t engine_check(const SysDev* d_sys, int id, int n, const Rec* alt, unsigned long long* out, hipStream_t s) {
    hipLaunchKernelGGL(k_check, dim3((n + 255) / 256), dim3(256), 0, s, d_sys, id, alt, out);
    return hipGetLastError();
}
hipError_t engine_observables(const SysDev* d_sys, int id, const int* cstart, const int* clen, int nch, int nfirst,
                              double boxl, double* outd, int* outi, hipStream_t s) {
    hipLaunchKernelGGL(k_obs, dim3(1), dim3(kObsT), 0, s, d_sys, id, cstart, clen, nch, nfirst, boxl, outd, outi);
    return hipGetLastError();
}
// ------------------------------------------------------------------------
// observables without device allocation: the chain tables and the results
// are staged in one fixed device array per GPU
// ------------------------------------------------------------------------
constexpr int kScratchChains = 4096;  // a system holds at most 2047 chains
constexpr size_t kScratchClen = sizeof(int) * kScratchChains;
constexpr size_t kScratchOutd = 2 * kScratchClen;
constexpr size_t kScratchOuti = kScratchOutd + 64;
constexpr size_t kScratchBytes = kScratchOuti + 256;  // 33 024 bytes
// Host code must not use this name as a device address: each GPU holds its
// own copy, whose address engine_scratch_init looks up.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices];
}  // namespace

hipError_t engine_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<unsigned char*>(p);
    return hipSuccess;
}

hipError_t engine_observables_noalloc(const SysDev* d_sys, int id, int device, const int* h_cstart, const int* h_clen,
                                      int nch, int nfirst, double boxl, double* h_outd, int* h_outi, hipStream_t s) {
    if (nch < 1 || nch > kScratchChains) return hipErrorInvalidValue;
    hipError_t e = engine_scratch_init(device);
    if (e != hipSuccess) return e;
    unsigned char* base;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = g_scratch[device];
    }
    int* cs = reinterpret_cast<int*>(base);
    int* cl = reinterpret_cast<int*>(base + kScratchClen);
    double* od = reinterpret_cast<double*>(base + kScratchOutd);
    int* oi = reinterpret_cast<int*>(base + kScratchOuti);
    e = hipMemcpyAsync(cs, h_cstart, sizeof(int) * (size_t)nch, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(cl, h_clen, sizeof(int) * (size_t)nch, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_obs, dim3(1), dim3(kObsT), 0, s, d_sys, id, cs, cl, nch, nfirst, boxl, od, oi);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(h_outd, od, sizeof(double) * 7, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(h_outi, oi, sizeof(int) * 3, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

hipError_t engine_prepare(size_t lds_limit) {
    hipError_t e = hipFuncSetAttribute((const void*)k_run, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_setup, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_rebuild_lists, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)k_rebuild_events, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds_limit);
}

}  // namespace pdmd
