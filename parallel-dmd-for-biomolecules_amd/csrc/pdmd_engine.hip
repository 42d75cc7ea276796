// pdmd_engine.hip -- gfx950 kernels of the PRIME20 discontinuous-MD event
// loop.  One wavefront (64 lanes) owns one system for the whole launch and
// commits its events in the reference's serial order; the wave
// parallelises everything inside an event:
//   * collision-time prediction over the candidate lists of the beads an
//     event touches (FP64, reference evaluation order, no contraction),
//   * the calendar: entries are grouped in 16-entry blocks whose minima
//     live in LDS; the next event is a wave min-reduction over the block
//     table, touched blocks are rescanned four at a time,
//   * neighbour-list rebuilds (coarse-cell counting sort + per-bead scan),
//     displacement checks and free-flight advances at the sync event.
// Many systems in one launch run concurrently, one wavefront each.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "pdmd_device.h"
#include "pdmd_engine.h"

namespace pdmd {
namespace {

constexpr int kW = 64;
constexpr int kBlk = 16;
constexpr int kNumbin = 2000;
constexpr double kLt = 1e-10;
constexpr double kSm = 5e-12;

// ------------------------------------------------------------------------
// project RNG (counter-based splitmix64) and logarithm from basic ops
// ------------------------------------------------------------------------
__device__ __host__ inline uint64_t sm64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ inline double rng_u01(uint64_t seed, int stream, uint64_t ctr) {
    const uint64_t key = sm64(seed ^ (0x5851F42D4C957F2Dull * (uint64_t)(stream + 1)));
    return (double)(sm64(key + ctr) >> 11) * (1.0 / 9007199254740992.0);
}
__device__ double dlog(double x) {
    uint64_t b = __double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7ff);
    int adj = 0;
    if (e == 0) {
        x = x * 18014398509481984.0;
        b = __double_as_longlong(x);
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
// topology: static pair codes (closed form of make_code.f) and overrides
// ------------------------------------------------------------------------
__device__ int static_code(const SysDev& S, int a, int b) {
    const int La = S.clen[a], Lb = S.clen[b];
    int ka = (S.loc[a] - 1) / La, ra = (S.loc[a] - 1) % La + 1;
    int kb = (S.loc[b] - 1) / Lb, rb = (S.loc[b] - 1) % Lb + 1;
    if (S.chain[a] == S.chain[b]) {
        if (ka > kb || (ka == kb && ra > rb)) {
            int t = ka; ka = kb; kb = t;
            t = ra; ra = rb; rb = t;
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
            case 15: return (d >= 4 && S.hp[S.res[a]] == 1 && S.hp[S.res[b]] == 1) ? 16 : 1;
            default: return 1;
        }
    }
    if (ka == 3 && kb == 3) return (S.hp[S.res[a]] == 1 && S.hp[S.res[b]] == 1) ? 16 : 1;
    if ((ka == 1 && kb == 2) || (ka == 2 && kb == 1)) {
        const int nbd = ka == 1 ? a : b, cbd = ka == 1 ? b : a;
        if (S.restype[S.res[nbd]] == 17 && S.chain[nbd] < S.chain[cbd]) return 1;
        return 15;
    }
    return 1;
}

// ev_code(row, col): a dynamic override written by the H-bond shoulder
// bookkeeping, else the static code.
__device__ inline int dyn_code(const SysDev& S, int row, int col, int stat) {
    const int4 o = S.ovr[row];
    if (o.x >= 0 && (o.x >> 8) == col) return o.x & 0xff;
    if (o.y >= 0 && (o.y >> 8) == col) return o.y & 0xff;
    if (o.z >= 0 && (o.z >> 8) == col) return o.z & 0xff;
    if (o.w >= 0 && (o.w >> 8) == col) return o.w & 0xff;
    return stat;
}
// single-lane writer of ev_code(row, col) = code
__device__ void set_code(const SysDev& S, int row, int col, int code, int& status) {
    int4 o = S.ovr[row];
    int* s = &o.x;
    const int stat = static_code(S, row, col);
    for (int k = 0; k < 4; ++k)
        if (s[k] >= 0 && (s[k] >> 8) == col) s[k] = -1;
    if (code != stat) {
        int k = 0;
        while (k < 4 && s[k] >= 0) ++k;
        if (k == 4) {
            status = kErrOverrideSlots;
            return;
        }
        s[k] = (col << 8) | code;
    }
    S.ovr[row] = o;
}

// ------------------------------------------------------------------------
// predictors: core.f, bond.f, sqwel.f, sqshlder.f, nc_sqwel.f
// ------------------------------------------------------------------------
struct Geo {
    double rx, ry, rz, b, r2, v2;
};
__device__ inline void geo(const Bead& a, const Bead& c, double tf, Geo& g) {
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

__device__ inline void bond_window(const SysDev& S, const DevTables& T, int code, int res_i, double& blmin,
                                   double& blmax) {
    blmin = T.ev_param[2][code];
    blmax = T.ev_param[3][code];
    if (code >= 10 && code <= 12) {
        const double* w = S.bondw + (size_t)res_i * 6 + (code - 10) * 2;
        blmin = w[0];
        blmax = w[1];
    }
}

// Collision time of pair (i, j) at false time tf with ev_code `code`; tij
// keeps its preset (1e9) and type its value when no event is predicted.
__device__ void predict(const SysDev& S, const DevTables& T, int code, const Bead& bi, const Bead& bj, int idi,
                        int idj, bool bonded_ij, int res_i, double tf, double& tij, int& type) {
    Geo g;
    geo(bi, bj, tf, g);
    const double b = g.b, r2 = g.r2, v2 = g.v2;
    if (code <= 3 || (code >= 17 && code <= 26)) {  // core.f:14-40
        if (b < 0.0) {
            const double pe = T.ev_param[1][code];
            double sigsq = T.sigma_sq[idi][idj] * (pe * pe);
            if (code >= 22 && code <= 26) {
                const double q = T.sqz610[code - 21][idi > idj ? idi : idj];
                sigsq = sigsq * (q * q);
            }
            const double discr = b * b - v2 * (r2 - sigsq);
            if (discr > 0.0) {
                tij = (-b - sqrt(discr)) / v2;
                type = 1;
            }
        }
    } else if (code <= 12) {  // bond.f:17-117
        double blmin, blmax;
        bond_window(S, T, code, res_i, blmin, blmax);
        if (b < 0.0) {
            const double d1 = b * b - v2 * (r2 - blmin * blmin);
            if (d1 > 0.0) {
                tij = (-b - sqrt(d1)) / v2;
                type = 2;
            } else {
                const double d2 = b * b - v2 * (r2 - blmax * blmax);
                if (d2 > 0.0) {
                    tij = (-b + sqrt(d2)) / v2;
                    type = 3;
                }
            }
        } else {
            const double d2 = b * b - v2 * (r2 - blmax * blmax);
            if (d2 > 0.0) {
                tij = -(r2 - blmax * blmax) / (sqrt(d2) + b);
                type = 3;
            }
        }
    } else if (code == 15) {  // nc_sqwel.f:19-122
        const double diff = r2 - T.welldia_sq[idi][idj];
        if (idi + idj == 5) {
            if (b < 0.0) {
                if (diff < 0.0) {
                    const double cd = b * b - v2 * (r2 - T.sigma_sq[idi][idj]);
                    if (cd > 0.0) {
                        tij = (-b - sqrt(cd)) / v2;
                        type = 1;
                    } else {
                        const double wd = b * b - v2 * diff;
                        tij = (-b + sqrt(wd)) / v2;
                        type = 16;
                    }
                } else {
                    const double wd = b * b - v2 * diff;
                    if (wd > 0.0) {
                        tij = (-b - sqrt(wd)) / v2;
                        type = 7;
                    }
                }
            } else if (diff < 0.0) {
                const double wd = b * b - v2 * diff;
                tij = (-b + sqrt(wd)) / v2;
                type = 16;
            }
        } else if (bonded_ij) {
            if (b < 0.0) {
                const double fs = T.sigma_sq[idi][idj] * T.ev_param[1][15] * T.ev_param[1][15];
                const double fc = b * b - v2 * (r2 - fs);
                if (fc > 0.0) {
                    tij = (-b - sqrt(fc)) / v2;
                    type = 1;
                } else {
                    const double wd = b * b - v2 * diff;
                    tij = (-b + sqrt(wd)) / v2;
                    type = 8;
                }
            } else {
                const double wd = b * b - v2 * diff;
                tij = (-b + sqrt(wd)) / v2;
                type = 8;
            }
        } else if (b < 0.0) {
            const double wd = b * b - v2 * diff;
            if (wd > 0.0) {
                tij = (-b - sqrt(wd)) / v2;
                type = 9;
            }
        }
    } else if (code == 16) {  // sqwel.f:17-63
        const double diff = r2 - T.welldia_sq[idi][idj];
        if (b < 0.0) {
            if (diff < 0.0) {
                const double cd = b * b - v2 * (r2 - T.sigma_sq[idi][idj]);
                if (cd > 0.0) {
                    tij = (-b - sqrt(cd)) / v2;
                    type = 1;
                } else {
                    const double wd = b * b - v2 * diff;
                    tij = (-b + sqrt(wd)) / v2;
                    type = 8;
                }
            } else {
                const double wd = b * b - v2 * diff;
                if (wd > 0.0) {
                    tij = (-b - sqrt(wd)) / v2;
                    type = 4;
                }
            }
        } else if (diff < 0.0) {
            const double wd = b * b - v2 * diff;
            tij = (-b + sqrt(wd)) / v2;
            type = 8;
        }
    } else if (code >= 40) {  // sqshlder.f:17-62
        const double diff = r2 - T.shlddia_sq[idi][idj];
        if (b < 0.0) {
            if (diff < 0.0) {
                const double cd = b * b - v2 * (r2 - T.sigma_sq[idi][idj]);
                if (cd > 0.0) {
                    tij = (-b - sqrt(cd)) / v2;
                    type = 1;
                } else {
                    const double sd = b * b - v2 * diff;
                    tij = (-b + sqrt(sd)) / v2;
                    type = 10;
                }
            } else {
                const double sd = b * b - v2 * diff;
                if (sd > 0.0) {
                    tij = (-b - sqrt(sd)) / v2;
                    type = 12;
                }
            }
        } else if (diff < 0.0) {
            const double sd = b * b - v2 * diff;
            tij = (-b + sqrt(sd)) / v2;
            type = 10;
        }
    } else {
        type = -99;  // codes 13, 14, 27..39: error in ev_code matrix
    }
}

// prediction of pair (x, m) with the reference's code lookup: up-pairs read
// ev_code(m, x) (eventredo_up.f:28), down-pairs ev_code(x, m) (eventredo_down.f:28)
__device__ inline void predict_pair(const SysDev& S, const DevTables& T, int x, int m, int code, double tf,
                                    double& tij, int& type) {
    const Bead bx = S.bead[x];
    const Bead bm = S.bead[m];
    tij = 1000000000.0;
    type = 0;
    predict(S, T, code, bx, bm, S.ident[x], S.ident[m], S.bptnr[x] == m, S.res[x], tf, tij, type);
}

// ------------------------------------------------------------------------
// wave helpers
// ------------------------------------------------------------------------
__device__ inline bool key_less(int b1, double t1, int i1, int b2, double t2, int i2) {
    if (b1 != b2) return b1 < b2;
    if (t1 != t2) return t1 < t2;
    return i1 < i2;
}
__device__ inline void wave_min_key(int& b, double& t, int& i, int width) {
    for (int m = 1; m < width; m <<= 1) {
        const int ob = __shfl_xor(b, m, kW);
        const double ot = __shfl_xor(t, m, kW);
        const int oi = __shfl_xor(i, m, kW);
        if (key_less(ob, ot, oi, b, t, i)) {
            b = ob;
            t = ot;
            i = oi;
        }
    }
}
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
__device__ inline double wave_max(double v) {
    for (int s = 1; s < kW; s <<= 1) {
        const double o = __shfl_xor(v, s, kW);
        v = o > v ? o : v;
    }
    return v;
}
__device__ inline int wave_or(int v) {
    for (int s = 1; s < kW; s <<= 1) v |= __shfl_xor(v, s, kW);
    return v;
}
__device__ inline int wave_sum(int v) {
    for (int s = 1; s < kW; s <<= 1) v += __shfl_xor(v, s, kW);
    return v;
}
__device__ inline unsigned long long ballot(bool p) { return __ballot(p); }

// ------------------------------------------------------------------------
// per-system register state
// ------------------------------------------------------------------------
struct Sc {
    double tfalse, t, interval, interval_max, t_fact, n_forced, sortsize, tbin_off, old_tfalse;
    long long coll, numghosts, nupdates, nforced, target;
    unsigned long long rc0, rc1;
    int nbin, status;
};

__device__ inline int calc_bin(const Sc& c, double tim, bool uncond, int& status) {
    if (!uncond && !(tim < c.interval_max)) return -1;
    int j = (int)((tim + c.tbin_off) / c.sortsize) + 1;
    if (j > kNumbin) j = kNumbin;  // -Ddebugging clamp, add_tbin.f:20-23
    if (j < 0) {
        status = kErrBin;
        return -1;
    }
    return j < c.nbin ? -1 : j;  // behind the scan position: never reached before the next rebuild
}

struct Lds {
    double* bt_t;
    int* bt_b;
    int* bt_i;
    int* dl_l;
    int* dl_tag;
    int* blist;
    int* misc;
    int nblk, maxd, maxblist;
};

// rescan the listed calendar blocks, four 16-entry blocks per wave pass
__device__ void rescan_blocks(const SysDev& S, const Lds& L, int N3, int nb) {
    const int lane = threadIdx.x;
    for (int base = 0; base < nb; base += 4) {
        const int q = base + (lane >> 4), k = lane & 15;
        const int blk = q < nb ? L.blist[q] : -1;
        int kb = INT_MAX, ki = INT_MAX;
        double kt = 0.0;
        const int e = blk * kBlk + k;
        if (blk >= 0 && e < N3) {
            const int cb = S.cbin[e];
            if (cb >= 0) {
                kb = cb;
                kt = S.tim[e];
                ki = e;
            }
        }
        wave_min_key(kb, kt, ki, 16);
        if (k == 0 && blk >= 0) {
            L.bt_t[blk] = kt;
            L.bt_b[blk] = kb;
            L.bt_i[blk] = ki;
        }
    }
    __syncthreads();
}
__device__ void rescan_all(const SysDev& S, const Lds& L, int N3) {
    const int lane = threadIdx.x;
    for (int base = 0; base < L.nblk; base += 4) {
        const int blk = base + (lane >> 4), k = lane & 15;
        int kb = INT_MAX, ki = INT_MAX;
        double kt = 0.0;
        const int e = blk * kBlk + k;
        if (blk < L.nblk && e < N3) {
            const int cb = S.cbin[e];
            if (cb >= 0) {
                kb = cb;
                kt = S.tim[e];
                ki = e;
            }
        }
        wave_min_key(kb, kt, ki, 16);
        if (k == 0 && blk < L.nblk) {
            L.bt_t[blk] = kt;
            L.bt_b[blk] = kb;
            L.bt_i[blk] = ki;
        }
    }
    __syncthreads();
}

// Full recompute of the next event of up to two beads x1, x2 (x2 < 0: none)
// over their up lists and upper shoulder partners: the loops of
// partial_events.f:26-46 / events.f:24-104.  Results are wave-uniform.
__device__ void wave_full2(const SysDev& S, const DevTables& T, int x1, int x2, double tf, double T0,
                           double& tim1, int& np1, int& ct1, double& tim2, int& np2, int& ct2, int& status) {
    const int lane = threadIdx.x;
    const int cap = S.cap;
    int nx[2], ex[2][3], nex[2];
    const int xs[2] = {x1, x2};
    for (int q = 0; q < 2; ++q) {
        nx[q] = 0;
        nex[q] = 0;
        if (xs[q] < 0) continue;
        nx[q] = S.nup[xs[q]];
        const int4 e = S.extra[xs[q]];
        const int ev[3] = {e.x, e.y, e.z};
        for (int k = 0; k < 3; ++k)
            if (ev[k] > xs[q]) ex[q][nex[q]++] = ev[k];
    }
    const int tot1 = nx[0] + nex[0], tot = tot1 + nx[1] + nex[1];
    double bt[2] = {1e300, 1e300};
    int bp[2] = {INT_MAX, INT_MAX}, bm[2] = {-1, -1}, bty[2] = {-1, -1};
    for (int base = 0; base < tot; base += kW) {
        const int p = base + lane;
        if (p < tot) {
            const int q = p < tot1 ? 0 : 1;
            const int pp = q == 0 ? p : p - tot1;
            const int x = xs[q];
            int m, stat;
            if (pp < nx[q]) {
                const int ent = S.nb_up[(size_t)x * cap + pp];
                m = ent & 0xFFFFFF;
                stat = ent >> 24;
            } else {
                m = ex[q][pp - nx[q]];
                stat = static_code(S, m, x);
            }
            const int code = dyn_code(S, m, x, stat);
            double tij;
            int ty;
            predict_pair(S, T, x, m, code, tf, tij, ty);
            if (ty == -99) status = kErrBadCode;
            if (tij < 0.0 && fabs(tij) > 1.0e-10) status = kErrNegativeTime;
            if (tij < bt[q]) {  // lanes see their candidates in list order
                bt[q] = tij;
                bp[q] = p;
                bm[q] = m;
                bty[q] = ty;
            }
        }
    }
    for (int q = 0; q < 2; ++q) wave_min_pred(bt[q], bp[q], bm[q], bty[q]);
    // tim = min(T0, candidates) + tf, strict '<' keeps the first minimum
    if (bt[0] < T0) {
        tim1 = bt[0] + tf;
        np1 = bm[0];
        ct1 = bty[0];
    } else {
        tim1 = T0 + tf;
        np1 = -1;
        ct1 = -1;
    }
    if (bt[1] < T0) {
        tim2 = bt[1] + tf;
        np2 = bm[1];
        ct2 = bty[1];
    } else {
        tim2 = T0 + tf;
        np2 = -1;
        ct2 = -1;
    }
}

// eventredo_down.f: single pair update of l's next event with partner p
__device__ inline void single_update(const SysDev& S, const DevTables& T, int l, int p, double tf, double& st_t,
                                     int& st_p, int& st_c, bool& modified, int& status) {
    const int code = dyn_code(S, l, p, static_code(S, l, p));
    double tij;
    int ty;
    predict_pair(S, T, l, p, code, tf, tij, ty);
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

// partial_events.f:16-201 for event (i, j); j < 0 for a ghost event.
// Returns (through L.blist / *nblist) the calendar blocks to rescan.
__device__ void partial_events(const SysDev& S, const DevTables& T, const Lds& L, Sc& c, int i, int j,
                               bool xpulse_del) {
    const int lane = threadIdx.x;
    const int N3 = S.N + 3;
    const double tf = c.tfalse;
    const double T0 = c.interval_max + kLt - tf;
    int status = 0;
    // (a) full recompute of i and j
    double ti, tj;
    int pi, ci, pj, cj;
    wave_full2(S, T, i, j, tf, T0, ti, pi, ci, tj, pj, cj, status);
    int nbl = 0;
    if (lane == 0) {
        S.tim[i] = ti;
        S.nptnr[i] = pi;
        S.coltype[i] = ci;
        S.cbin[i] = calc_bin(c, ti, false, status);
        L.blist[nbl] = i / kBlk;
        if (j >= 0) {
            S.tim[j] = tj;
            S.nptnr[j] = pj;
            S.coltype[j] = cj;
            S.cbin[j] = calc_bin(c, tj, false, status);
            L.blist[nbl + 1] = j / kBlk;
        }
    }
    nbl = j >= 0 ? 2 : 1;
    // (b) the down set: dnnab(i), extra(i,:)<i, dnnab(j)\{i}, extra(j,:)<j
    const int cap = S.cap;
    int nD = 0;
    const int ndi = S.ndn[i];
    for (int base = 0; base < ndi; base += kW) {
        const int p = base + lane;
        if (p < ndi && p < L.maxd) {
            L.dl_l[p] = S.nb_dn[(size_t)i * cap + p] & 0xFFFFFF;
            L.dl_tag[p] = 0;
        }
    }
    nD = ndi;
    {
        const int4 e = S.extra[i];
        const int ev[3] = {e.x, e.y, e.z};
        for (int k = 0; k < 3; ++k)
            if (ev[k] >= 0 && ev[k] < i) {
                if (lane == 0 && nD < L.maxd) {
                    L.dl_l[nD] = ev[k];
                    L.dl_tag[nD] = 0;
                }
                ++nD;
            }
    }
    if (j >= 0) {
        const int ndj = S.ndn[j];
        for (int base = 0; base < ndj; base += kW) {
            const int p = base + lane;
            const int l = p < ndj ? (S.nb_dn[(size_t)j * cap + p] & 0xFFFFFF) : -1;
            const bool keep = p < ndj && l != i;
            const unsigned long long bal = ballot(keep);
            const int pre = __popcll(bal & ((1ull << lane) - 1ull));
            if (keep && nD + pre < L.maxd) {
                L.dl_l[nD + pre] = l;
                L.dl_tag[nD + pre] = 1;
            }
            nD += __popcll(bal);
        }
        const int4 e = S.extra[j];
        const int ev[3] = {e.x, e.y, e.z};
        for (int k = 0; k < 3; ++k)
            if (ev[k] >= 0 && ev[k] < j) {
                if (lane == 0 && nD < L.maxd) {
                    L.dl_l[nD] = ev[k];
                    L.dl_tag[nD] = 1;
                }
                ++nD;
            }
    }
    if (nD > L.maxd) {
        status = kErrListCapacity;
        nD = L.maxd;
    }
    __syncthreads();
    // (c) per unique l: apply its I-ops then J-ops in reference order
    for (int base = 0; base < nD; base += kW) {
        const int q = base + lane;
        bool first = false;
        int l = -1, cI = 0, cJ = 0;
        if (q < nD) {
            l = L.dl_l[q];
            first = true;
            for (int r = 0; r < q; ++r)
                if (L.dl_l[r] == l) {
                    first = false;
                    break;
                }
            if (first)
                for (int r = 0; r < nD; ++r)
                    if (L.dl_l[r] == l) {
                        if (L.dl_tag[r] == 0) ++cI;
                        else ++cJ;
                    }
        }
        double st_t = 0.0;
        int st_p = -1, st_c = -1;
        bool modified = false, pending = false;
        if (first) {
            st_t = S.tim[l];
            st_p = S.nptnr[l];
            st_c = S.coltype[l];
        }
        int rem = first ? cI + cJ : 0;
        for (int guard = 0; guard < 64 * 9; ++guard) {
            while (rem > 0 && !pending) {
                const int p = (rem > cJ) ? i : j;
                --rem;
                if (st_p != p)
                    single_update(S, T, l, p, tf, st_t, st_p, st_c, modified, status);
                else
                    pending = true;
            }
            const unsigned long long pm = ballot(pending);
            if (pm == 0) break;
            const int leader = __ffsll((long long)pm) - 1;
            const int lf = __shfl(l, leader, kW);
            double ft, dt;
            int fp, fc, dp, dc;
            wave_full2(S, T, lf, -1, tf, T0, ft, fp, fc, dt, dp, dc, status);
            if (lane == leader) {
                st_t = ft;
                st_p = fp;
                st_c = fc;
                modified = true;
                pending = false;
            }
        }
        if (modified) {
            S.tim[l] = st_t;
            S.nptnr[l] = st_p;
            S.coltype[l] = st_c;
            S.cbin[l] = calc_bin(c, st_t, false, status);
        }
        const unsigned long long mm = ballot(modified);
        const int pre = __popcll(mm & ((1ull << lane) - 1ull));
        if (modified && nbl + pre < L.maxblist) L.blist[nbl + pre] = l / kBlk;
        nbl += __popcll(mm);
        if (nbl > L.maxblist) nbl = L.maxblist;
    }
    // (d) repuls_del_b.f once the rescheduling is done
    if (xpulse_del && lane == 0) {
        const int a = S.ident[i] < S.ident[j] ? i : j, b = a == i ? j : i;
        const int ncim1 = a + S.clen[a] - 1, ncai = a - S.clen[a];
        const int ncaj = b - 2 * S.clen[b], nnjp1 = b - S.clen[b] + 1;
        int4 ea = S.extra[a], eb = S.extra[b];
        ea.x = -1; ea.y = -1; ea.w = -1;
        eb.x = -1; eb.y = -1; eb.w = -1;
        S.extra[a] = ea;
        S.extra[b] = eb;
        S.extra[ncaj].z = -1;
        S.extra[nnjp1].z = -1;
        S.extra[ncai].z = -1;
        S.extra[ncim1].z = -1;
    }
    status = wave_or(status);
    if (status && !c.status) c.status = status;
    __syncthreads();
    rescan_blocks(S, L, N3, nbl);
}

// ------------------------------------------------------------------------
// event dynamics: eventdyn.f, bumped.f, repuls_*.f (single lane)
// ------------------------------------------------------------------------
__device__ inline void nudge(Bead& bi, Bead& bj, double bd, const Geo& g, int sign) {
    if (sign > 0) {
        bi.x = bi.x + bd * g.rx; bi.y = bi.y + bd * g.ry; bi.z = bi.z + bd * g.rz;
        bj.x = bj.x - bd * g.rx; bj.y = bj.y - bd * g.ry; bj.z = bj.z - bd * g.rz;
    } else {
        bi.x = bi.x - bd * g.rx; bi.y = bi.y - bd * g.ry; bi.z = bi.z - bd * g.rz;
        bj.x = bj.x + bd * g.rx; bj.y = bj.y + bd * g.ry; bj.z = bj.z + bd * g.rz;
    }
}

__device__ void eventdyn(const SysDev& S, const DevTables& T, Bead& bi, Bead& bj, int i, int j, int a, int c,
                         int evcode, double tf, int& ct, int& status) {
    Geo g;
    geo(bi, bj, tf, g);
    const double bmi = bi.bm, bmj = bj.bm;
    const double rmass = 2 * bmi * bmj / (bmi + bmj);
    double ratio = 0.0, wellsq, epsave, del_pe, bd;
    switch (ct) {
        case 2:
        case 3: {
            double blmin, blmax;
            bond_window(S, T, evcode, S.res[i], blmin, blmax);
            ratio = ct == 2 ? rmass * g.b / (blmin * blmin) : rmass * g.b / (blmax * blmax);
            break;
        }
        case 1: {
            double sigsq;
            if (evcode == 15) {
                if (S.bptnr[i] == j) {
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

__device__ inline double dist_tf(const SysDev& S, int a, int b, double tf) {
    Geo g;
    const Bead ba = S.bead[a], bb = S.bead[b];
    geo(ba, bb, tf, g);
    return sqrt(g.rx * g.rx + g.ry * g.ry + g.rz * g.rz);
}
__device__ inline void shoulder_idx(const SysDev& S, int i, int j, int& ncim1, int& ncai, int& ncaj, int& nnjp1) {
    ncim1 = i + S.clen[i] - 1;
    ncai = i - S.clen[i];
    ncaj = j - 2 * S.clen[j];
    nnjp1 = j - S.clen[j] + 1;
}
__device__ double repuls_check(const SysDev& S, const DevTables& T, int i, int j, double tf) {
    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(S, i, j, ncim1, ncai, ncaj, nnjp1);
    const double d1 = dist_tf(S, i, ncaj, tf), d2 = dist_tf(S, i, nnjp1, tf);
    const double d3 = dist_tf(S, j, ncai, tf), d4 = dist_tf(S, j, ncim1, tf);
    if (d1 > T.shder[1] && d2 > T.shder[2] && d3 > T.shder[3] && d4 > T.shder[4]) return 10.0;
    return 11.0;
}
__device__ double repuls_check_3(const SysDev& S, const DevTables& T, int i, int j, int k, double tf) {
    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(S, i, j, ncim1, ncai, ncaj, nnjp1);
    int m = 0;
    if (ncaj != k && dist_tf(S, i, ncaj, tf) > T.shder[1]) ++m;
    if (nnjp1 != k && dist_tf(S, i, nnjp1, tf) > T.shder[2]) ++m;
    if (ncai != k && dist_tf(S, j, ncai, tf) > T.shder[3]) ++m;
    if (ncim1 != k && dist_tf(S, j, ncim1, tf) > T.shder[4]) ++m;
    return m == 3 ? 10.0 : 11.0;
}
__device__ void repuls_add(const SysDev& S, int i, int j, int& status) {
    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(S, i, j, ncim1, ncai, ncaj, nnjp1);
    set_code(S, ncaj, i, 50, status);
    set_code(S, i, ncaj, 40, status);
    set_code(S, i, nnjp1, 40, status);
    set_code(S, nnjp1, i, 50, status);
    set_code(S, ncai, j, 50, status);
    set_code(S, j, ncai, 40, status);
    set_code(S, j, ncim1, 40, status);
    set_code(S, ncim1, j, 50, status);
    int4 ei = S.extra[i], ej = S.extra[j];
    ei.x = ncaj; ei.y = nnjp1; ei.w = j;
    ej.x = ncai; ej.y = ncim1; ej.w = i;
    S.extra[i] = ei;
    S.extra[j] = ej;
    S.extra[ncaj].z = i;
    S.extra[nnjp1].z = i;
    S.extra[ncai].z = j;
    S.extra[ncim1].z = j;
}
__device__ void repuls_del_a(const SysDev& S, int i, int j, int& status) {
    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(S, i, j, ncim1, ncai, ncaj, nnjp1);
    set_code(S, ncaj, i, 1, status);
    set_code(S, i, ncaj, 1, status);
    set_code(S, i, nnjp1, 1, status);
    set_code(S, nnjp1, i, 1, status);
    set_code(S, ncai, j, 1, status);
    set_code(S, j, ncai, 1, status);
    set_code(S, ncim1, j, 1, status);
    set_code(S, j, ncim1, 1, status);
}
__device__ inline bool terminal(const SysDev& S, int k) {
    return S.loc[k] == S.clen[k] + 1 || S.loc[k] == 3 * S.clen[k];
}

// One pair event: outcome resolution (main.F90:1487-1634), dynamics, H-bond
// and shoulder bookkeeping (main.F90:641-921), then partial_events.
__device__ int pair_event(const SysDev& S, const DevTables& T, const Lds& L, Sc& c, int i) {
    const int lane = threadIdx.x;
    const int j = S.nptnr[i];
    int ct = S.coltype[i];
    const double tf = c.tfalse;
    const int evcode = dyn_code(S, i, j, static_code(S, i, j));
    const bool ok = !terminal(S, i) && !terminal(S, j);
    int status = 0;
    bool xpulse = false;
    const int idi = S.ident[i], idj = S.ident[j];
    if (ct == 7) {
        const int4 ei = S.extra[i], ej = S.extra[j];
        if (ei.w < 0 && ej.w < 0) {
            if (ok) {
                const double rating = idi < idj ? repuls_check(S, T, i, j, tf) : repuls_check(S, T, j, i, tf);
                ct = rating <= 10.0 ? 4 : 14;
            } else {
                const double ran_non = rng_u01(S.seed, 1, c.rc1++);
                ct = ran_non <= 0.2 ? 4 : 9;
            }
        } else
            ct = 9;
    } else if (ct == 10) {
        double rating;
        if (evcode < 45) {
            const int hbp = S.extra[i].w;
            rating = idi < S.ident[hbp] ? repuls_check_3(S, T, i, hbp, j, tf) : repuls_check_3(S, T, hbp, i, j, tf);
        } else {
            const int hbp = S.extra[j].w;
            rating = idj < S.ident[hbp] ? repuls_check_3(S, T, j, hbp, i, tf) : repuls_check_3(S, T, hbp, j, i, tf);
        }
        ct = rating <= 10.0 ? 5 : 15;
    } else if (ct == 12) {
        const int x = evcode < 45 ? i : j;
        const int hbp = S.extra[x].w;
        if (S.bptnr[x] == hbp) {
            const double d2 = [&] {
                Geo g;
                const Bead bx = S.bead[x], bh = S.bead[hbp];
                geo(bx, bh, tf, g);
                return g.rx * g.rx + g.ry * g.ry + g.rz * g.rz;
            }();
            const double rating = (d2 - T.sigma_sq[S.ident[x]][S.ident[hbp]]) < 0.0 ? 10.0 : 11.0;
            ct = rating > 10.0 ? 6 : 13;
        } else
            ct = 15;
    }
    Bead bi = S.bead[i], bj = S.bead[j];
    if (ct < 14) eventdyn(S, T, bi, bj, i, j, idi, idj, evcode, tf, ct, status);
    __syncthreads();
    if (lane == 0) {
        if (ct == 20) {
            if (idi + idj == 5) {
                S.bptnr[i] = j;
                S.bptnr[j] = i;
                S.ident[i] = idi + 4;
                S.ident[j] = idj + 4;
                if (ok) {
                    if (idi + 4 < idj + 4) repuls_add(S, i, j, status);
                    else repuls_add(S, j, i, status);
                }
            }
        } else if (ct == 21) {
            if (idi <= 8) {
                if (ok) {
                    if (idi < idj) repuls_del_a(S, i, j, status);
                    else repuls_del_a(S, j, i, status);
                    xpulse = true;
                }
                if (S.bptnr[i] == j) {
                    S.bptnr[i] = -1;
                    S.bptnr[j] = -1;
                    S.ident[i] = idi - 4;
                    S.ident[j] = idj - 4;
                }
            }
        } else if (ct == 24) {
            const int x = evcode < 45 ? i : j;
            const int hbp = S.extra[x].w;
            if (S.ident[x] + S.ident[hbp] == 5) {
                S.bptnr[x] = hbp;
                S.bptnr[hbp] = x;
                S.ident[x] += 4;
                S.ident[hbp] += 4;
            }
        } else if (ct == 25) {
            const int x = evcode < 45 ? i : j;
            const int hbp = S.extra[x].w;
            if (S.ident[x] >= 5) {
                S.bptnr[x] = -1;
                S.bptnr[hbp] = -1;
                S.ident[x] -= 4;
                S.ident[hbp] -= 4;
            }
        } else if (ct == 14 || ct == 15 || ct == 16) {
            // bumped.f: push the pair off the discontinuity
            Geo g;
            geo(bi, bj, tf, g);
            const double bd = evcode >= 40 ? kSm * sqrt(T.shlddia_sq[idi][idj]) : kSm * sqrt(T.welldia_sq[idi][idj]);
            nudge(bi, bj, bd, g, g.b < 0.0 ? -1 : +1);
            if (ct == 14 && ok) {
                if (idi < idj) repuls_add(S, i, j, status);
                else repuls_add(S, j, i, status);
            } else if (ct == 16 && ok) {
                if (idi < idj) repuls_del_a(S, i, j, status);
                else repuls_del_a(S, j, i, status);
                xpulse = true;
            }
        }
        S.bead[i] = bi;
        S.bead[j] = bj;
    }
    xpulse = __shfl(xpulse ? 1 : 0, 0, kW) != 0;
    status = wave_or(status);
    if (status && !c.status) c.status = status;
    __syncthreads();
    partial_events(S, T, L, c, i, j, xpulse);
    return ct;
}

// Andersen ghost collision, main.F90:997-1049
__device__ int ghost_event(const SysDev& S, const DevTables& T, const Lds& L, Sc& c) {
    const int lane = threadIdx.x;
    const int N = S.N;
    int i = 0, guard = 0;
    do {
        i = (int)(rng_u01(S.seed, 0, c.rc0++) * N);
    } while (i == N && ++guard < 1000);
    c.numghosts++;
    Bead b = S.bead[i];
    const double tf = c.tfalse, bm = b.bm;
    b.x = b.x + b.vx * tf;
    b.y = b.y + b.vy * tf;
    b.z = b.z + b.vz * tf;
    double v1, v2, r, fact;
    guard = 0;
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
    if (lane == 0) {
        S.bead[i] = b;
        S.tim[N] = tg;
        S.cbin[N] = calc_bin(c, tg, false, status);
    }
    if (status && !c.status) c.status = status;
    if (c.tfalse < c.old_tfalse) c.tfalse = c.old_tfalse;
    __syncthreads();
    if (lane == 0) L.blist[0] = N / kBlk;
    __syncthreads();
    rescan_blocks(S, L, S.N + 3, 1);
    partial_events(S, T, L, c, i, -1, false);
    return i;
}

// ------------------------------------------------------------------------
// neighbour lists: nbor.f membership restated for a coarse-cell search
// ------------------------------------------------------------------------
// Pair (i, j) is listed when its fine cells (the reference's cell_add.f
// binning) are within the n_wrap=2 stencil (|d| <= 2 per axis, periodic),
// and either its code is a bond/local one (4-12, 17-26) or r^2 <= rlsq(code).
__device__ void rebuild_lists(const SysDev& S, const DevTables& T, const Lds& L, Sc& c) {
    const int lane = threadIdx.x;
    const int N = S.N, M = S.M, Mq = S.Mq, NC = S.ncoarse, cap = S.cap;
    // fine and coarse cells
    for (int k = lane; k < N; k += kW) {
        const Bead b = S.bead[k];
        int cx = (int)((b.x + S.half) / S.width) + 2 - 2;
        int cy = (int)((b.y + S.half) / S.width) + 2 - 2;
        int cz = (int)((b.z + S.half) / S.width) + 2 - 2;
        cx = ((cx % M) + M) % M;
        cy = ((cy % M) + M) % M;
        cz = ((cz % M) + M) % M;
        S.cfine[k] = cx | (cy << 10) | (cz << 20);
        int qx = cx >> 1, qy = cy >> 1, qz = cz >> 1;
        if (qx >= Mq) qx = Mq - 1;
        if (qy >= Mq) qy = Mq - 1;
        if (qz >= Mq) qz = Mq - 1;
        S.ckey[k] = qx + Mq * (qy + Mq * qz);
    }
    for (int q = lane; q < NC; q += kW) {
        S.cstart[q] = 0;
        S.crun[q] = 0;
    }
    __syncthreads();
    for (int k = lane; k < N; k += kW) atomicAdd(&S.cstart[S.ckey[k]], 1);
    __syncthreads();
    // exclusive scan of cstart (wave-wide, chunked)
    int carry = 0;
    for (int base = 0; base < NC; base += kW) {
        const int q = base + lane;
        const int v = q < NC ? S.cstart[q] : 0;
        int incl = v;
        for (int s = 1; s < kW; s <<= 1) {
            const int o = __shfl_up(incl, s, kW);
            if (lane >= s) incl += o;
        }
        if (q < NC) S.cstart[q] = carry + incl - v;
        carry += __shfl(incl, kW - 1, kW);
    }
    if (lane == 0) S.cstart[NC] = carry;
    __syncthreads();
    // stable scatter in bead order
    for (int base = 0; base < N; base += kW) {
        const int k = base + lane;
        const int key = k < N ? S.ckey[k] : -1;
        int rank = 0, cnt = 0;
        for (int s = 0; s < kW; ++s) {
            const int o = __shfl(key, s, kW);
            if (o == key) {
                if (s < lane) ++rank;
                ++cnt;
            }
        }
        int basepos = 0;
        if (k < N) basepos = S.cstart[key] + S.crun[key];
        __syncthreads();
        if (k < N) {
            S.csorted[basepos + rank] = k;
            if (rank == cnt - 1) S.crun[key] += cnt;
        }
        __syncthreads();
    }
    // per-bead scan of the coarse neighbourhood
    int mx = 0;
    int status = 0;
    for (int i = lane; i < N; i += kW) {
        const int fi = S.cfine[i];
        const int ix = fi & 1023, iy = (fi >> 10) & 1023, iz = fi >> 20;
        const int k = S.ckey[i];
        const int qx = k % Mq, qy = (k / Mq) % Mq, qz = k / (Mq * Mq);
        const Bead bi = S.bead[i];
        int nu = 0, nd = 0;
        int offs[3];
        int noff = 0;
        offs[noff++] = 0;
        if (Mq > 1) offs[noff++] = 1;
        if (Mq > 2) offs[noff++] = -1;
        for (int a = 0; a < noff; ++a)
            for (int bb = 0; bb < noff; ++bb)
                for (int cc = 0; cc < noff; ++cc) {
                    const int nx = (qx + offs[a] + Mq) % Mq, ny = (qy + offs[bb] + Mq) % Mq,
                              nz = (qz + offs[cc] + Mq) % Mq;
                    const int cell = nx + Mq * (ny + Mq * nz);
                    const int s0 = S.cstart[cell], s1 = S.cstart[cell + 1];
                    for (int p = s0; p < s1; ++p) {
                        const int jb = S.csorted[p];
                        if (jb == i) continue;
                        const int fj = S.cfine[jb];
                        int dx = (fj & 1023) - ix, dy = ((fj >> 10) & 1023) - iy, dz = (fj >> 20) - iz;
                        dx = ((dx % M) + M) % M;
                        dy = ((dy % M) + M) % M;
                        dz = ((dz % M) + M) % M;
                        if (dx > 2 && dx < M - 2) continue;
                        if (dy > 2 && dy < M - 2) continue;
                        if (dz > 2 && dz < M - 2) continue;
                        const int stat = static_code(S, jb, i);
                        const int code = dyn_code(S, jb, i, stat);  // ev_code(j,i), nbor.f:63
                        bool take;
                        if ((code >= 4 && code <= 12) || (code >= 17 && code < 27)) {
                            take = true;
                        } else {
                            const Bead bj = S.bead[jb];
                            double rx = bi.x - bj.x, ry = bi.y - bj.y, rz = bi.z - bj.z;
                            rx = rx - round(rx);
                            ry = ry - round(ry);
                            rz = rz - round(rz);
                            const double r2 = rx * rx + ry * ry + rz * rz;
                            take = code >= 1 && code <= 50 && r2 <= T.rlsq[code];
                        }
                        if (!take) continue;
                        if (jb > i) {
                            if (nu < cap) S.nb_up[(size_t)i * cap + nu] = jb | (stat << 24);
                            ++nu;
                        } else {
                            if (nd < cap) S.nb_dn[(size_t)i * cap + nd] = jb | (stat << 24);
                            ++nd;
                        }
                    }
                }
        if (nu > cap || nd > cap) status = kErrListCapacity;
        mx = max(mx, max(nu, nd));
        S.nup[i] = min(nu, cap);
        S.ndn[i] = min(nd, cap);
    }
    status = wave_or(status);
    if (status && !c.status) c.status = status;
    for (int s = 1; s < kW; s <<= 1) mx = max(mx, __shfl_xor(mx, s, kW));
    if (lane == 0 && mx > S.st->max_nb_seen) S.st->max_nb_seen = mx;
    __syncthreads();
}

// events.f:11-127 with tfalse = 0: every bead's next event, then the calendar
__device__ void events_all(const SysDev& S, const DevTables& T, const Lds& L, Sc& c) {
    const int lane = threadIdx.x;
    const int N = S.N, cap = S.cap;
    int status = 0;
    for (int i = lane; i < N; i += kW) {
        double ti = c.interval_max + kLt;
        int pi = -1, ci = -1;
        const int n = S.nup[i];
        for (int k = 0; k < n; ++k) {
            const int ent = S.nb_up[(size_t)i * cap + k];
            const int m = ent & 0xFFFFFF;
            const int code = dyn_code(S, m, i, ent >> 24);
            double tij;
            int ty;
            predict_pair(S, T, i, m, code, 0.0, tij, ty);
            if (ty == -99) status = kErrBadCode;
            if (tij < ti) {
                ti = tij;
                pi = m;
                ci = ty;
            }
            if (tij < 0.0 && fabs(tij) > 1.0e-10) status = kErrNegativeTime;
        }
        const int4 e = S.extra[i];
        const int ev[3] = {e.x, e.y, e.z};
        for (int k = 0; k < 3; ++k) {
            const int m = ev[k];
            if (m > i) {
                const int code = dyn_code(S, m, i, static_code(S, m, i));
                double tij = 1000000000.0;
                int ty = 0;
                const Bead bi = S.bead[i], bm = S.bead[m];
                // events.f:79-83: code 1 -> core, anything else -> shoulder
                predict(S, T, code == 1 ? 1 : 40, bi, bm, S.ident[i], S.ident[m], S.bptnr[i] == m, S.res[i], 0.0,
                        tij, ty);
                if (tij < 0.0 && fabs(tij) > 1.0e-10) status = kErrNegativeTime;
                if (tij < ti) {
                    pi = m;
                    ci = ty;
                    ti = tij;
                }
            }
        }
        S.tim[i] = ti;
        S.nptnr[i] = pi;
        S.coltype[i] = ci;
    }
    __syncthreads();
    c.nbin = 1;
    for (int e = lane; e < N + 3; e += kW) S.cbin[e] = calc_bin(c, S.tim[e], false, status);
    status = wave_or(status);
    if (status && !c.status) c.status = status;
    __syncthreads();
    rescan_all(S, L, N + 3);
}

__device__ void rebuild(const SysDev& S, const DevTables& T, const Lds& L, Sc& c) {
    const int lane = threadIdx.x;
    for (int k = lane; k < S.N; k += kW) {
        Bead b = S.bead[k];
        b.x = b.x - round(b.x);
        b.y = b.y - round(b.y);
        b.z = b.z - round(b.z);
        S.bead[k] = b;
        S.old_r[k] = make_double4(b.x, b.y, b.z, 0.0);
    }
    __syncthreads();
    rebuild_lists(S, T, L, c);
    events_all(S, T, L, c);
}

// sync pseudo-event, main.F90:1126-1187
__device__ void sync_event(const SysDev& S, const DevTables& T, const Lds& L, Sc& c) {
    const int lane = threadIdx.x;
    const int N = S.N;
    const double tf = c.tfalse;
    c.t = c.t + tf;
    for (int e = lane; e < N + 3; e += kW) S.tim[e] = S.tim[e] - tf;
    c.interval_max = c.interval_max - tf;
    c.tbin_off = c.tbin_off + tf;
    double moved_far = 0.0;
    for (int k = lane; k < N; k += kW) {
        Bead b = S.bead[k];
        b.x = b.x + b.vx * tf;
        b.y = b.y + b.vy * tf;
        b.z = b.z + b.vz * tf;
        S.bead[k] = b;
        const double4 o = S.old_r[k];
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
    __syncthreads();
    if (update || c.interval > c.interval_max) {
        if (!update) {
            c.nforced++;
            c.n_forced = c.n_forced * 1.01;
        }
        c.interval_max = c.interval * c.n_forced;
        c.sortsize = c.interval_max / (double)kNumbin;
        c.tbin_off = 0.0;
        c.nupdates++;
        rebuild(S, T, L, c);
    } else {
        rescan_all(S, L, N + 3);
    }
    int status = 0;
    const double t2 = c.interval * 0.999;
    if (lane == 0) {
        S.tim[N + 1] = t2;
        S.cbin[N + 1] = calc_bin(c, t2, true, status);
        L.blist[0] = (N + 1) / kBlk;
    }
    if (status && !c.status) c.status = status;
    __syncthreads();
    rescan_blocks(S, L, N + 3, 1);
}

__device__ void output_event(const SysDev& S, const Lds& L, Sc& c) {
    const int lane = threadIdx.x;
    const int N = S.N;
    int status = 0;
    const double t3 = 3.3 / (sqrt(S.setemp)) + 5 + c.tfalse;
    if (lane == 0) {
        S.tim[N + 2] = t3;
        S.cbin[N + 2] = calc_bin(c, t3, false, status);
        L.blist[0] = (N + 2) / kBlk;
    }
    if (status && !c.status) c.status = status;
    __syncthreads();
    rescan_blocks(S, L, N + 3, 1);
}

__device__ Lds make_lds(int nblk, int maxd) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    Lds L;
    L.nblk = nblk;
    L.maxd = maxd;
    L.maxblist = 256;
    unsigned char* p = smem;
    L.bt_t = (double*)p;
    p += sizeof(double) * nblk;
    L.bt_b = (int*)p;
    p += sizeof(int) * nblk;
    L.bt_i = (int*)p;
    p += sizeof(int) * nblk;
    L.dl_l = (int*)p;
    p += sizeof(int) * maxd;
    L.dl_tag = (int*)p;
    p += sizeof(int) * maxd;
    L.blist = (int*)p;
    p += sizeof(int) * L.maxblist;
    L.misc = (int*)p;
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
    st->rng_ctr[0] = c.rc0;
    st->rng_ctr[1] = c.rc1;
    st->nbin = c.nbin;
    st->status = c.status;
}

}  // namespace

// ------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------
// Set-up of one system per wavefront: wrap, lists, pseudo-event seeds and
// events() (main.F90:205-424 after the host-side H-bond restore).
__global__ __launch_bounds__(64) void k_setup(const SysDev* sys, const int* ids, int nblk, int maxd) {
    const SysDev& S = sys[ids[blockIdx.x]];
    const DevTables& T = *S.tab;
    Lds L = make_lds(nblk, maxd);
    Sc c;
    load_sc(S.st, c);
    const int N = S.N;
    for (int k = threadIdx.x; k < N; k += kW) {
        S.tim[k] = c.interval_max + kLt;
        S.coltype[k] = -1;
        S.nptnr[k] = -1;
    }
    if (threadIdx.x == 0) {
        if (S.canon) {
            double tgho = 0.0;
            int guard = 0;
            do {
                tgho = rng_u01(S.seed, 0, c.rc0++);
            } while ((tgho < 1.e-18 || tgho == 1.0) && ++guard < 1000);
            S.tim[N] = -1.0 * dlog(tgho) * S.avegtime * .0000001;
        } else
            S.tim[N] = 1000000000.0;
        S.tim[N + 1] = c.interval;
        S.tim[N + 2] = 3.3 / (sqrt(S.setemp)) + 5;
        for (int e = N; e < N + 3; ++e) {
            S.nptnr[e] = -2;
            S.coltype[e] = -2;
        }
    }
    c.rc0 = __shfl(c.rc0, 0, kW);
    __syncthreads();
    rebuild(S, T, L, c);
    if (threadIdx.x == 0) store_sc(S.st, c);
}

// The event loop: each wavefront commits events of its system until the
// system's coll counter reaches its target.
__global__ __launch_bounds__(64) void k_run(const SysDev* sys, const int* ids, const long long* inc, int nblk,
                                           int maxd) {
    const SysDev& S = sys[ids[blockIdx.x]];
    const DevTables& T = *S.tab;
    Lds L = make_lds(nblk, maxd);
    Sc c;
    load_sc(S.st, c);
    c.target = c.coll + inc[blockIdx.x];
    const int N = S.N, N3 = N + 3;
    rescan_all(S, L, N3);
    const int lane = threadIdx.x;
    int log_count = S.st->log_count;
    while (c.coll < c.target && c.status == 0) {
        // next event: min over the block table of (bin, time, index)
        int kb = INT_MAX, ki = INT_MAX;
        double kt = 0.0;
        for (int b = lane; b < L.nblk; b += kW) {
            const int ob = L.bt_b[b];
            const double ot = L.bt_t[b];
            const int oi = L.bt_i[b];
            if (key_less(ob, ot, oi, kb, kt, ki)) {
                kb = ob;
                kt = ot;
                ki = oi;
            }
        }
        wave_min_key(kb, kt, ki, kW);
        if (kb == INT_MAX) {
            c.status = kErrBin;
            break;
        }
        const int e = ki;
        c.old_tfalse = c.tfalse;
        c.nbin = kb;
        c.tfalse = kt;
        c.coll++;
        int jj = -1, ty = 0;
        if (e < N) {
            jj = S.nptnr[e];
            ty = pair_event(S, T, L, c, e);
            if (ty >= 1 && ty <= 30 && lane == 0) S.st->nevents[ty]++;
        } else if (e == N) {
            jj = ghost_event(S, T, L, c);
        } else if (e == N + 1) {
            sync_event(S, T, L, c);
        } else {
            output_event(S, L, c);
        }
        if (S.log && log_count < S.log_cap) {
            if (lane == 0) {
                LogRec r;
                r.coll = c.coll;
                r.i = e + 1;
                r.j = jj + 1;
                r.type = ty;
                r.pad = 0;
                r.t = kt;
                S.log[log_count] = r;
            }
            ++log_count;
        }
    }
    if (lane == 0) {
        S.st->log_count = log_count;
        store_sc(S.st, c);
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
    const int code = dyn_code(S, j, i, static_code(S, j, i));
    double t;
    int ty;
    predict_pair(S, *S.tab, i, j, code, S.st->tfalse, t, ty);
    tij[k] = t;
    type[k] = ty;
}

size_t engine_lds_bytes(int nblk, int maxd) {
    return sizeof(double) * nblk + 2 * sizeof(int) * nblk + 2 * sizeof(int) * maxd + sizeof(int) * 256 + 64;
}

hipError_t engine_setup(const SysDev* d_sys, const int* d_ids, int n, int max_n, int maxd, hipStream_t s) {
    const int nblk = (max_n + 3 + kBlk - 1) / kBlk;
    const size_t lds = engine_lds_bytes(nblk, maxd);
    hipLaunchKernelGGL(k_setup, dim3(n), dim3(kW), lds, s, d_sys, d_ids, nblk, maxd);
    return hipGetLastError();
}
hipError_t engine_run(const SysDev* d_sys, const int* d_ids, const long long* d_inc, int n, int max_n, int maxd,
                      hipStream_t s) {
    const int nblk = (max_n + 3 + kBlk - 1) / kBlk;
    const size_t lds = engine_lds_bytes(nblk, maxd);
    hipLaunchKernelGGL(k_run, dim3(n), dim3(kW), lds, s, d_sys, d_ids, d_inc, nblk, maxd);
    return hipGetLastError();
}
hipError_t engine_predict(const SysDev* d_sys, int id, int n, const int* pi, const int* pj, double* tij, int* type,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_predict, dim3((n + 255) / 256), dim3(256), 0, s, d_sys, id, n, pi, pj, tij, type);
    return hipGetLastError();
}

}  // namespace pdmd
