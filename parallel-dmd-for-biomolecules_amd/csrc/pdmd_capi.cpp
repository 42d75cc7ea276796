// pdmd_capi.cpp -- C ABI of the MI355X PRIME20 DMD engine (include/pdmd.h):
// contexts, system upload with the reference's start-up sequence, device
// launches, downloads, event logs and observables.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pdmd.h"
#include "pdmd_device.h"
#include "pdmd_engine.h"
#include "pdmd_host.h"
#include "pdmd_internal.h"

namespace pdmd {
namespace {
thread_local std::string g_err;
}  // namespace

int report_error(const char* msg) {
    g_err = msg;
    return -1;
}

namespace {

int fail(const std::string& msg, int code = -1) {
    g_err = msg;
    return code;
}
#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return fail(std::string(#expr ": ") + hipGetErrorString(e_), -2); \
    } while (0)

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t count) {
        n = count;
        return hipMalloc((void**)&p, sizeof(T) * (count ? count : 1));
    }
};

struct HostSys {
    Topology topo;
    Params p;  // scaled to this box
    Cutoffs cut{};
    int N = 0, cap = 64, canon = 1;
    double boxl = 0, tstar = 0;
    uint64_t seed = 0;
    // device memory
    DevBuf<DevTables> tab;
    DevBuf<Rec> rec;
    DevBuf<double2> bw;
    DevBuf<int> nb_up, nup, nb_dn, ndn, cfine, ckey, cstart, crun, csorted;
    DevBuf<int4> extra, ovr;
    DevBuf<int2> cfp;
    DevBuf<Leaf> leaf;
    DevBuf<double4> old_r, spos;
    DevBuf<int4> sinfo;
    DevBuf<unsigned> mark;
    DevBuf<unsigned long long> stamp;
    DevBuf<SysState> st;
    DevBuf<LogRec> log;
    SysDev dev{};
    bool alive = false;
};

}  // namespace
}  // namespace pdmd

struct pdmd_ctx {
    pdmd::Params raw;
    std::string param_dir;
    std::vector<std::unique_ptr<pdmd::HostSys>> sys;
    pdmd::DevBuf<pdmd::SysDev> d_sys;
    pdmd::DevBuf<int> d_ids;
    pdmd::DevBuf<long long> d_inc;
    pdmd::DevBuf<unsigned long long> d_clock;  // start of a budgeted launch
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    size_t lds_limit = 65536;
    int device = 0;
    long long wall_khz = 100000;  // wall_clock64 rate
    int waves = 1;                // wavefronts per system
    // graph launches: the event-loop launch captured once per launch shape
    // and replayed while the shape stays the same
    bool graphs = false;
    hipGraphExec_t gexec = nullptr;
    struct Shape {
        const void *sys, *ids, *inc;
        int nsys, max_n, maxw, flags, waves;
        long long ticks;
        bool operator==(const Shape& o) const {
            return sys == o.sys && ids == o.ids && inc == o.inc && nsys == o.nsys && max_n == o.max_n &&
                   maxw == o.maxw && flags == o.flags && waves == o.waves && ticks == o.ticks;
        }
    } gshape{};
};

namespace pdmd {
namespace {

// ev_code writes of the host-side start-up (same slot semantics as the
// device): an override is kept only where it differs from the static code.
struct HostCodes {
    const Topology& t;
    std::vector<int4>& ovr;
    int dyn(int row, int col) const {
        const int4 o = ovr[row];
        const int s[4] = {o.x, o.y, o.z, o.w};
        for (int k = 0; k < 4; ++k)
            if (s[k] >= 0 && (s[k] >> 8) == col) return s[k] & 0xff;
        return static_code(t, row, col);
    }
    bool set(int row, int col, int code) {
        int4& o = ovr[row];
        int* s = &o.x;
        for (int k = 0; k < 4; ++k)
            if (s[k] >= 0 && (s[k] >> 8) == col) s[k] = -1;
        if (code != static_code(t, row, col)) {
            int k = 0;
            while (k < 4 && s[k] >= 0) ++k;
            if (k == 4) return false;
            s[k] = (col << 8) | code;
        }
        return true;
    }
};

bool host_repuls_add(const Topology& t, HostCodes& hc, std::vector<int4>& extra, int i, int j) {
    const int Li = t.chain_len[t.chain[i]], Lj = t.chain_len[t.chain[j]];
    const int ncim1 = i + Li - 1, ncai = i - Li, ncaj = j - 2 * Lj, nnjp1 = j - Lj + 1;
    bool ok = hc.set(ncaj, i, 50) && hc.set(i, ncaj, 40) && hc.set(i, nnjp1, 40) && hc.set(nnjp1, i, 50) &&
              hc.set(ncai, j, 50) && hc.set(j, ncai, 40) && hc.set(j, ncim1, 40) && hc.set(ncim1, j, 50);
    extra[i].x = ncaj;
    extra[i].y = nnjp1;
    extra[ncaj].z = i;
    extra[nnjp1].z = i;
    extra[j].x = ncai;
    extra[j].y = ncim1;
    extra[ncai].z = j;
    extra[ncim1].z = j;
    extra[i].w = j;
    extra[j].w = i;
    return ok;
}

bool terminal_bead(const Topology& t, int k) {
    const int L = t.chain_len[t.chain[k]];
    return t.loc[k] == L + 1 || t.loc[k] == 3 * L;
}

template <class T>
hipError_t upload(DevBuf<T>& b, const std::vector<T>& v) {
    hipError_t e = b.alloc(v.size());
    if (e != hipSuccess) return e;
    if (!v.empty()) e = hipMemcpy(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
    return e;
}

int validate_desc(const pdmd_system_desc* d) {
    if (!d || d->nchains <= 0 || !d->chain_len || !d->restype || !d->hp || !d->sv)
        return fail("pdmd_system_desc: missing arrays");
    for (int c = 0; c < d->nchains; ++c)
        if (d->chain_len[c] < 4) return fail("pdmd_system_desc: chains need at least 4 residues");
    if (!(d->boxl > 0) || !(d->tstar > 0)) return fail("pdmd_system_desc: boxl and tstar must be positive");
    if (d->flags & ~(PDMD_NO_HBS | PDMD_NO_H | PDMD_CHAPTYPE2)) return fail("pdmd_system_desc: unknown flags");
    return 0;
}

// Host half of the start-up of pdmd_system_add: validation, topology, the
// -Drunr H-bond restore, tables, cutoffs and cell geometry.  No device call.
struct Prep {
    std::unique_ptr<HostSys> hs;
    std::vector<Rec> bead;
    std::vector<int4> extra, ovr;
    std::vector<double2> bw;
    DevTables tab{};
    SysState st{};
    int M = 0, Mq = 0, nres = 0;
    double setemp = 0;
};

int prepare_system(const Params& raw, const pdmd_system_desc* d, Prep& out) {
    if (int e = validate_desc(d)) return e;
    out.hs = std::make_unique<HostSys>();
    HostSys& H = *out.hs;
    const int nc = d->nchains;
    std::vector<int> clen(d->chain_len, d->chain_len + nc), spec(nc);
    for (int c = 0; c < nc; ++c) spec[c] = d->species ? d->species[c] : c;
    int nres = 0;
    for (int c = 0; c < nc; ++c) nres += clen[c];
    std::vector<int> rt(d->restype, d->restype + nres), hp(d->hp, d->hp + nres);
    for (int q = 0; q < nres; ++q)
        if (rt[q] < 9 || rt[q] > 28) return fail("pdmd_system_add: restype must be 9..28");
    H.topo.build(clen, spec, rt, hp);
    H.topo.no_hbs = (d->flags & PDMD_NO_HBS) != 0;
    H.topo.no_h = (d->flags & PDMD_NO_H) != 0;
    H.topo.chaptype2 = (d->flags & PDMD_CHAPTYPE2) != 0;
    const Topology& t = H.topo;
    const int N = t.N;
    H.N = N;
    H.cap = d->list_capacity > 0 ? d->list_capacity : 64;
    if (H.cap > 64) return fail("pdmd_system_add: list_capacity above 64 (one wavefront lane per list slot)");
    if (N > 65533) return fail("pdmd_system_add: at most 65533 beads per system");
    if (nc > 2047) return fail("pdmd_system_add: at most 2047 chains per system");
    for (int c = 0; c < nc; ++c)
        if (clen[c] > 128) return fail("pdmd_system_add: at most 128 residues per chain");
    H.canon = d->canon ? 1 : 0;
    H.boxl = d->boxl;
    H.tstar = d->tstar;
    H.seed = d->seed;
    H.p = raw;
    scale_params(H.p, d->boxl);
    const Params& P = H.p;
    // state: positions wrapped as inputinfo.f:90-95 and main.F90:210-216 do
    std::vector<Rec>& bead = out.bead;
    bead.assign(N, Rec{});
    for (int k = 0; k < N; ++k) {
        Rec b{};
        double x = d->sv[6 * k], y = d->sv[6 * k + 1], z = d->sv[6 * k + 2];
        x = x - std::round(x);
        y = y - std::round(y);
        z = z - std::round(z);
        x = x - std::round(x);
        y = y - std::round(y);
        z = z - std::round(z);
        b.x = x;
        b.y = y;
        b.z = z;
        b.vx = d->sv[6 * k + 3];
        b.vy = d->sv[6 * k + 4];
        b.vz = d->sv[6 * k + 5];
        b.bm = raw.bmass[t.ident0[k]];  // inputinfo.f:400-403, initial identities
        bead[k] = b;
    }
    std::vector<int> ident = t.ident0, bptnr(N, -1);
    std::vector<int4>& extra = out.extra;
    std::vector<int4>& ovr = out.ovr;
    extra.assign(N, make_int4(-1, -1, -1, -1));
    ovr.assign(N, make_int4(-1, -1, -1, -1));
    if (d->bptnr)
        for (int k = 0; k < N; ++k) {
            const int b = d->bptnr[k];
            if (b < 0 || b > N) return fail("pdmd_system_add: bptnr out of range");
            bptnr[k] = b - 1;
        }
    // -Drunr restore of H-bond identities and shoulders, main.F90:240-321.
    // The reference walks every pair k < kj in order, but only two kinds of
    // pair act: an N-C pair (identities 1 + 4) inside its well, and
    // (k, bptnr(k)).  The N-C pairs closer than the wider of the two N-C
    // wells come from a cell grid; sorted into (k, kj) order with the
    // bptnr pairs, they are visited in the reference's order, so identity
    // changes made along the way reach the later pairs as they do there.
    HostCodes hc{t, ovr};
    std::vector<std::pair<int, int>> cand;
    {
        const double w2 = std::max(P.welldia_sq[1][4], P.welldia_sq[4][1]);
        auto r2_of = [&](int a, int b) {
            double rx = bead[a].x - bead[b].x, ry = bead[a].y - bead[b].y, rz = bead[a].z - bead[b].z;
            rx = rx - std::round(rx);
            ry = ry - std::round(ry);
            rz = rz - std::round(rz);
            return rx * rx + ry * ry + rz * rz;
        };
        std::vector<int> nbead, cbead;
        for (int k = 0; k < N; ++k) {
            const int L = t.chain_len[t.chain[k]], kind = (t.loc[k] - 1) / L;
            if (kind == 1) nbead.push_back(k);
            if (kind == 2) cbead.push_back(k);
        }
        // cells at least 1 + 1e-6 well radii wide (a pair closer than the
        // well sits in neighbouring cells; positions are wrapped to
        // [-1/2, 1/2]), and about four beads per cell at most
        const int nc = std::min((int)std::floor(1.0 / (std::sqrt(w2) * (1.0 + 1e-6))), (int)std::cbrt(4.0 * N) + 1);
        if (nc >= 3) {
            auto cell_of = [&](double x) { return std::min(nc - 1, std::max(0, (int)std::floor((x + 0.5) * nc))); };
            std::vector<int> head((size_t)nc * nc * nc, -1), next(N, -1);
            for (int q : cbead) {
                const int c = cell_of(bead[q].x) + nc * (cell_of(bead[q].y) + nc * cell_of(bead[q].z));
                next[q] = head[c];
                head[c] = q;
            }
            for (int k : nbead) {
                const int cx = cell_of(bead[k].x), cy = cell_of(bead[k].y), cz = cell_of(bead[k].z);
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int c = (cx + dx + nc) % nc + nc * ((cy + dy + nc) % nc + nc * ((cz + dz + nc) % nc));
                            for (int q = head[c]; q >= 0; q = next[q])
                                if (r2_of(k, q) < w2) cand.emplace_back(std::min(k, q), std::max(k, q));
                        }
            }
        } else {
            for (int k : nbead)
                for (int q : cbead)
                    if (r2_of(k, q) < w2) cand.emplace_back(std::min(k, q), std::max(k, q));
        }
        for (int k = 0; k < N; ++k)
            if (bptnr[k] > k) cand.emplace_back(k, bptnr[k]);
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    }
    for (const auto& pr : cand) {
        const int k = pr.first, kj = pr.second;
        if (ident[k] + ident[kj] == 5 && hc.dyn(k, kj) == 15) {
            double rx = bead[k].x - bead[kj].x, ry = bead[k].y - bead[kj].y, rz = bead[k].z - bead[kj].z;
            rx = rx - std::round(rx);
            ry = ry - std::round(ry);
            rz = rz - std::round(rz);
            const double r2 = rx * rx + ry * ry + rz * rz;
            const double diff = r2 - P.welldia_sq[ident[k]][ident[kj]];
            if (diff < 0.0 && !terminal_bead(t, k) && !terminal_bead(t, kj)) {
                const bool ok = ident[k] == 1 ? host_repuls_add(t, hc, extra, k, kj)
                                              : host_repuls_add(t, hc, extra, kj, k);
                if (!ok) return fail("pdmd_system_add: override slots exhausted in H-bond restore");
            }
        }
        if (kj == bptnr[k]) {
            if (ident[k] == 1) {
                ident[k] = 5;
                ident[kj] = 8;
            } else {
                ident[k] = 8;
                ident[kj] = 5;
            }
        }
    }
    nbor_setup(P, t, ident, H.cut);
    const int M = (int)(1.0 / (H.cut.sig_max_all * kRlConst) * kNWrap);  // main.F90:390, interior cells
    if (M < 2 * kNWrap + 1) return fail("pdmd_system_add: box too small for the n_wrap=2 cell stencil");
    if (M >= 1024) return fail("pdmd_system_add: box too large for packed cell coordinates");
    const int Mq = M / 2;
    // device tables
    DevTables& tab = out.tab;
    tab = DevTables{};
    for (int a = 0; a < 29; ++a)
        for (int b = 0; b < 29; ++b) {
            tab.sigma_sq[a][b] = P.sigma_sq[a][b];
            tab.welldia_sq[a][b] = P.welldia_sq[a][b];
            tab.ep_sqrt[a][b] = P.ep_sqrt[a][b];
            tab.shlddia_sq[a][b] = P.shlddia_sq[a][b];
        }
    std::memcpy(tab.ev_param, P.ev_param, sizeof tab.ev_param);
    std::memcpy(tab.sqz610, P.sqz610, sizeof tab.sqz610);
    std::memcpy(tab.rlsq, H.cut.rlsq, sizeof tab.rlsq);
    for (int k = 0; k < 5; ++k) tab.shder[k] = P.shder_dist[k];
    tab.epsilon1 = P.epsilon[1];
    half_stencil(tab.halfst);
    // per-bead topology/identity words and own side-chain bond windows
    // (inputinfo.f:300-349, bond.f:21-31: Calpha owns code 10, N code 11, C code 12)
    std::vector<double2>& bw = out.bw;
    bw.assign(N, make_double2(0.0, 0.0));
    for (int k = 0; k < N; ++k) {
        const int c = t.chain[k], L = t.chain_len[c], kind = (t.loc[k] - 1) / L, r = (t.loc[k] - 1) % L + 1;
        const int q = t.res[k];
        const int pro = rt[q] == 17 ? 1 : 0;
        bead[k].topo = make_topo(c, kind, r, L, hp[q] ? 1 : 0, pro, t.no_hbs ? 1 : 0, t.no_h ? 1 : 0,
                                 t.free_bead(k) ? 1 : 0);
        bead[k].dyn = make_dyn(ident[k], bptnr[k]);
        if (kind <= 2) {
            const int iii = rt[q] - 8;
            const double bl = (kind == 0 ? raw.drca[iii] : kind == 1 ? raw.drnh[iii] : raw.drco[iii]) /
                              P.boxl_orig;
            const double dl = kind == 0 ? raw.del_rca[iii] : kind == 1 ? raw.del_rnh[iii] : raw.del_rco[iii];
            bw[k] = make_double2((1.0 - dl) * bl, (1.0 + dl) * bl);
        }
    }
    SysState& st = out.st;
    st = SysState{};
    const double setemp = d->tstar * 12.0;  // main.F90:126-127
    st.t_fact = 0.00005;
    st.n_forced = 150.0;
    st.interval = st.t_fact / std::sqrt(setemp);
    st.interval_max = st.n_forced * st.interval;
    st.sortsize = st.interval_max / (double)kNumbin;
    st.nbin = 1;
    out.M = M;
    out.Mq = Mq;
    out.nres = nres;
    out.setemp = setemp;
    return 0;
}

// Device bytes of one prepared system: every buffer pdmd_system_add
// allocates, each rounded up to the allocator's 4 KiB granularity.
int64_t device_bytes(const Prep& pp) {
    const HostSys& H = *pp.hs;
    const int64_t N = H.N, cap = H.cap, nc = (int64_t)pp.Mq * pp.Mq * pp.Mq;
    const int64_t sizes[] = {(int64_t)sizeof(DevTables), N * (int64_t)sizeof(Rec), N * 16 /*bw*/, N * 16 /*extra*/,
                             N * 16 /*ovr*/, N * 32 /*old_r*/, N * cap * 4 /*nb_up*/, N * cap * 4 /*nb_dn*/,
                             N * 4 /*nup*/, N * 4 /*ndn*/, (N + 3) * (int64_t)sizeof(Leaf), N * 4 /*cfine*/,
                             N * 4 /*ckey*/, N * 8 /*cfp*/, N * 4 /*csorted*/, N * 32 /*spos*/, N * 16 /*sinfo*/, N * 4 /*mark*/,
                             (nc + 1) * 4 /*cstart*/, nc * 4 /*crun*/, (N + 3) * 8 /*stamp*/,
                             (int64_t)sizeof(SysState)};
    int64_t b = 0;
    for (int64_t x : sizes) b += (x + 4095) / 4096 * 4096;
    return b;
}

// the parameter tables that ship beside the library (parameters/)
const Params* shipped_params() {
    static Params p;
    static int state = 0;  // 0 unread, 1 loaded, -1 failed
    if (state == 0) {
        state = -1;
        Dl_info info{};
        if (dladdr((const void*)&shipped_params, &info) && info.dli_fname) {
            std::string dir(info.dli_fname);
            const size_t cut = dir.find_last_of('/');
            dir = (cut == std::string::npos ? std::string(".") : dir.substr(0, cut)) + "/parameters";
            std::string err;
            if (load_params(p, dir, err)) state = 1;
        }
    }
    return state == 1 ? &p : nullptr;
}

}  // namespace

struct PreparedSystem {
    Prep pp;
};

}  // namespace pdmd

using namespace pdmd;

extern "C" {

const char* pdmd_version(void) { return "pdmd-mi355x 0.1 (gfx950)"; }
const char* pdmd_last_error(void) { return g_err.c_str(); }

int pdmd_ctx_create(const char* param_dir, pdmd_ctx** out) {
    if (!out || !param_dir) return fail("pdmd_ctx_create: null argument");
    auto ctx = std::make_unique<pdmd_ctx>();
    std::string err;
    if (!load_params(ctx->raw, param_dir, err)) return fail(err);
    ctx->param_dir = param_dir;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail("no HIP device", -3);
    HIPCHK(hipGetDevice(&ctx->device));
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&ctx->ev0));
    HIPCHK(hipEventCreate(&ctx->ev1));
    // LDS per workgroup: opt in to the device maximum (160 KB on gfx950)
    {
        int lds = 0, optin = 0;
        (void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
        (void)hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, ctx->device);
        size_t want = (size_t)std::max(lds, optin);
        if (want < 65536) want = 65536;
        if (engine_prepare(want) == hipSuccess)
            ctx->lds_limit = want;
        else {
            (void)hipGetLastError();
            ctx->lds_limit = (size_t)std::max(lds, 32768);
        }
    }
    // the fixed staging array of pdmd_observables_noalloc on this device
    HIPCHK(engine_scratch_init(ctx->device));
    HIPCHK(ctx->d_clock.alloc(1));
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) == hipSuccess && khz > 0)
            ctx->wall_khz = khz;
        else
            (void)hipGetLastError();
    }
    *out = ctx.release();
    return 0;
}

int pdmd_ctx_destroy(pdmd_ctx* ctx) {
    if (!ctx) return 0;
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    delete ctx;
    return 0;
}

int64_t pdmd_system_bytes(const pdmd_system_desc* d) {
    const Params* raw = shipped_params();
    if (!raw) return fail("pdmd_system_bytes: parameter tables beside the library not found");
    Prep pp;
    if (int e = prepare_system(*raw, d, pp)) return e;
    return device_bytes(pp);
}

int64_t pdmd_system_bytes_ctx(pdmd_ctx* ctx, const pdmd_system_desc* d) {
    if (!ctx) return fail("pdmd_system_bytes_ctx: null context");
    Prep pp;
    if (int e = prepare_system(ctx->raw, d, pp)) return e;
    return device_bytes(pp);
}

int pdmd_system_add(pdmd_ctx* ctx, const pdmd_system_desc* d, int32_t* sys_id) {
    if (!ctx || !sys_id) return fail("pdmd_system_add: null argument");
    PreparedSystem ps;
    if (int e = prepare_system(ctx->raw, d, ps.pp)) return e;
    return prepared_add(ctx, &ps, sys_id);
}

}  // extern "C"

namespace pdmd {

PreparedSystem* prepared_new(pdmd_ctx* ctx, const pdmd_system_desc* d) {
    if (!ctx) {
        fail("pdmd_system_add: null context");
        return nullptr;
    }
    auto ps = std::make_unique<PreparedSystem>();
    if (prepare_system(ctx->raw, d, ps->pp)) return nullptr;
    return ps.release();
}

void prepared_free(PreparedSystem* p) { delete p; }

int64_t prepared_bytes(const PreparedSystem* p) { return device_bytes(p->pp); }

int prepared_add(pdmd_ctx* ctx, PreparedSystem* ps, int32_t* sys_id) {
    if (!ctx || !ps || !sys_id || !ps->pp.hs) return fail("pdmd_system_add: null argument");
    Prep pp = std::move(ps->pp);
    ps->pp = Prep{};
    // an advance may still be running on the context's stream: it reads the
    // descriptor table and the id buffer that the start-up below rewrites
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HostSys& H = *pp.hs;
    const int N = H.N, nres = pp.nres, M = pp.M, Mq = pp.Mq;
    const double setemp = pp.setemp;
    const std::vector<Rec>& bead = pp.bead;
    const std::vector<int4>& extra = pp.extra;
    const std::vector<int4>& ovr = pp.ovr;
    const std::vector<double2>& bw = pp.bw;
    const DevTables& tab = pp.tab;
    const SysState& st = pp.st;
    {
        std::vector<DevTables> tv(1, tab);
        HIPCHK(upload(H.tab, tv));
    }
    HIPCHK(upload(H.rec, bead));
    HIPCHK(upload(H.bw, bw));
    HIPCHK(upload(H.extra, extra));
    HIPCHK(upload(H.ovr, ovr));
    HIPCHK(H.old_r.alloc(N));
    HIPCHK(H.nb_up.alloc((size_t)N * H.cap));
    HIPCHK(H.nb_dn.alloc((size_t)N * H.cap));
    HIPCHK(H.nup.alloc(N));
    HIPCHK(H.ndn.alloc(N));
    HIPCHK(H.leaf.alloc(N + 3));
    HIPCHK(H.cfine.alloc(N));
    HIPCHK(H.ckey.alloc(N));
    HIPCHK(H.cfp.alloc(N));
    HIPCHK(H.csorted.alloc(N));
    HIPCHK(H.spos.alloc(N));
    HIPCHK(H.sinfo.alloc(N));
    HIPCHK(H.mark.alloc(N));
    HIPCHK(hipMemsetAsync(H.mark.p, 0, sizeof(unsigned) * (size_t)N, ctx->stream));
    HIPCHK(H.stamp.alloc(N + 3));
    HIPCHK(hipMemsetAsync(H.stamp.p, 0, sizeof(unsigned long long) * (size_t)(N + 3), ctx->stream));
    HIPCHK(H.cstart.alloc((size_t)Mq * Mq * Mq + 1));
    HIPCHK(H.crun.alloc((size_t)Mq * Mq * Mq));
    {
        std::vector<SysState> sv(1, st);
        HIPCHK(upload(H.st, sv));
    }
    SysDev& D = H.dev;
    D.N = N;
    D.nres = nres;
    D.cap = H.cap;
    D.canon = H.canon;
    D.chaptype2 = H.topo.chaptype2 ? 1 : 0;
    D.waves_cap = 0;
    D.cq = 2;
    D.M = M;
    D.Mq = Mq;
    D.ncoarse = Mq * Mq * Mq;
    D.log_cap = 0;
    D.width = 1.0 / (double)M;
    D.half = 1.0 / 2.0;
    D.setemp = setemp;
    D.avegtime = 0.00005 / std::sqrt(setemp);
    D.hdelr = H.cut.hdelr;
    D.seed = H.seed;
    D.tab = H.tab.p;
    D.rec = H.rec.p;
    D.bw = H.bw.p;
    D.extra = H.extra.p;
    D.ovr = H.ovr.p;
    D.old_r = H.old_r.p;
    D.nb_up = H.nb_up.p;
    D.nup = H.nup.p;
    D.nb_dn = H.nb_dn.p;
    D.ndn = H.ndn.p;
    D.leaf = H.leaf.p;
    D.cfine = H.cfine.p;
    D.ckey = H.ckey.p;
    D.cfp = H.cfp.p;
    D.cstart = H.cstart.p;
    D.crun = H.crun.p;
    D.csorted = H.csorted.p;
    D.spos = H.spos.p;
    D.sinfo = H.sinfo.p;
    D.mark = H.mark.p;
    D.stamp = H.stamp.p;
    D.st = H.st.p;
    D.log = nullptr;
    H.alive = true;
    // slot: the first free one; the system is registered there only once
    // its start-up has succeeded
    int id = -1;
    for (size_t k = 0; k < ctx->sys.size(); ++k)
        if (!ctx->sys[k]) {
            id = (int)k;
            break;
        }
    const size_t nslots = id < 0 ? ctx->sys.size() + 1 : ctx->sys.size();
    if (id < 0) id = (int)ctx->sys.size();
    // device descriptor table
    if (ctx->d_sys.n < nslots) {
        DevBuf<SysDev> nb;
        HIPCHK(nb.alloc(nslots + 16));
        if (ctx->d_sys.p) HIPCHK(hipMemcpy(nb.p, ctx->d_sys.p, sizeof(SysDev) * ctx->d_sys.n, hipMemcpyDeviceToDevice));
        std::swap(ctx->d_sys.p, nb.p);
        std::swap(ctx->d_sys.n, nb.n);
    }
    HIPCHK(hipMemcpy(ctx->d_sys.p + id, &H.dev, sizeof(SysDev), hipMemcpyHostToDevice));
    // device start-up: lists, pseudo-event seeds, events()
    if (ctx->d_ids.n < 1) HIPCHK(ctx->d_ids.alloc(4096));
    if (ctx->d_inc.n < 1) HIPCHK(ctx->d_inc.alloc(4096));
    HIPCHK(hipMemcpy(ctx->d_ids.p, &id, sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(engine_setup(ctx->d_sys.p, ctx->d_ids.p, 1, N, engine_maxw(H.cap), ctx->lds_limit, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    SysState chk;
    HIPCHK(hipMemcpy(&chk, H.st.p, sizeof chk, hipMemcpyDeviceToHost));
    if (chk.status)
        return fail("pdmd_system_add: device start-up failed with status " + std::to_string(chk.status), -4);
    if ((size_t)id == ctx->sys.size()) ctx->sys.emplace_back();
    ctx->sys[id] = std::move(pp.hs);
    *sys_id = id + 1;
    return 0;
}

}  // namespace pdmd

extern "C" {


static HostSys* get_sys(pdmd_ctx* ctx, int32_t id) {
    if (!ctx || id < 1 || id > (int)ctx->sys.size() || !ctx->sys[id - 1]) {
        fail("invalid system id");
        return nullptr;
    }
    return ctx->sys[id - 1].get();
}

int pdmd_system_remove(pdmd_ctx* ctx, int32_t id) {
    if (!get_sys(ctx, id)) return -1;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->sys[id - 1].reset();
    return 0;
}

int pdmd_nbeads(pdmd_ctx* ctx, int32_t id) {
    HostSys* H = get_sys(ctx, id);
    return H ? H->N : -1;
}

namespace pdmd {
namespace {
int advance_launch(pdmd_ctx* ctx, int32_t nsys, const int32_t* ids, const int64_t* nev, double budget_ms,
                   int flags = 0);
}
}  // namespace pdmd

int pdmd_advance_async(pdmd_ctx* ctx, int32_t nsys, const int32_t* ids, const int64_t* nev) {
    return advance_launch(ctx, nsys, ids, nev, 0.0);
}

int pdmd_advance_budget_async(pdmd_ctx* ctx, int32_t nsys, const int32_t* ids, const int64_t* max_nev,
                              double budget_ms) {
    if (!(budget_ms > 0.0)) return fail("pdmd_advance_budget_async: budget must be positive");
    return advance_launch(ctx, nsys, ids, max_nev, budget_ms);
}

int pdmd_advance_to_output(pdmd_ctx* ctx, int32_t nsys, const int32_t* ids, const int64_t* max_nev) {
    if (int rc = advance_launch(ctx, nsys, ids, max_nev, 0.0, kStopAtOutput)) return rc;
    return pdmd_wait(ctx);
}

}  // extern "C"

namespace pdmd {
namespace {
int advance_launch(pdmd_ctx* ctx, int32_t nsys, const int32_t* ids, const int64_t* nev, double budget_ms, int flags) {
    if (!ctx || nsys <= 0 || !ids || !nev) return fail("pdmd_advance: bad arguments");
    if (nsys > 4096) return fail("pdmd_advance: at most 4096 systems per launch");
    std::vector<int> dids(nsys);
    std::vector<long long> inc(nsys);
    int max_n = 0, max_cap = 0;
    for (int k = 0; k < nsys; ++k) {
        HostSys* H = get_sys(ctx, ids[k]);
        if (!H) return -1;
        for (int q = 0; q < k; ++q)
            if (ids[q] == ids[k]) return fail("pdmd_advance: duplicate system id");
        if (nev[k] < 0) return fail("pdmd_advance: negative event count");
        dids[k] = ids[k] - 1;
        inc[k] = nev[k];
        max_n = std::max(max_n, H->N);
        max_cap = std::max(max_cap, H->cap);
    }
    HIPCHK(hipMemcpyAsync(ctx->d_ids.p, dids.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_inc.p, inc.data(), sizeof(long long) * nsys, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    long long ticks = 0;
    if (budget_ms > 0.0) ticks = (long long)(budget_ms * (double)ctx->wall_khz);
    const int maxw = engine_maxw(max_cap);
    if (!ctx->graphs) {
        HIPCHK(engine_run(ctx->d_sys.p, ctx->d_ids.p, ctx->d_inc.p, nsys, max_n, maxw, ctx->lds_limit, ctx->stream,
                          ticks, flags, ctx->waves, ctx->d_clock.p));
    } else {
        // the system list and event targets live in device buffers, so one
        // captured launch serves every step of the same shape
        const pdmd_ctx::Shape shape{ctx->d_sys.p, ctx->d_ids.p, ctx->d_inc.p, nsys, max_n, maxw, flags, ctx->waves,
                                    ticks};
        if (!ctx->gexec || !(shape == ctx->gshape)) {
            if (ctx->gexec) {
                HIPCHK(hipGraphExecDestroy(ctx->gexec));
                ctx->gexec = nullptr;
            }
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
            const hipError_t er = engine_run(ctx->d_sys.p, ctx->d_ids.p, ctx->d_inc.p, nsys, max_n, maxw,
                                             ctx->lds_limit, ctx->stream, ticks, flags, ctx->waves, ctx->d_clock.p);
            const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
            const hipError_t ei = er == hipSuccess && ec == hipSuccess ? hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0)
                                                                       : hipSuccess;
            if (g) (void)hipGraphDestroy(g);
            HIPCHK(er);
            HIPCHK(ec);
            HIPCHK(ei);
            ctx->gshape = shape;
        }
        HIPCHK(hipGraphLaunch(ctx->gexec, ctx->stream));
    }
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    ctx->timed = true;
    // the copies above read pageable memory synchronously before returning
    return 0;
}
}  // namespace
}  // namespace pdmd

extern "C" {

int pdmd_rebuild(pdmd_ctx* ctx, int32_t nsys, const int32_t* ids) {
    if (!ctx || nsys <= 0 || !ids) return fail("pdmd_rebuild: bad arguments");
    if (nsys > 4096) return fail("pdmd_rebuild: at most 4096 systems per call");
    std::vector<int> dids(nsys);
    int max_n = 0, max_cap = 0;
    for (int k = 0; k < nsys; ++k) {
        HostSys* H = get_sys(ctx, ids[k]);
        if (!H) return -1;
        for (int q = 0; q < k; ++q)
            if (ids[q] == ids[k]) return fail("pdmd_rebuild: duplicate system id");
        dids[k] = ids[k] - 1;
        max_n = std::max(max_n, H->N);
        max_cap = std::max(max_cap, H->cap);
    }
    HIPCHK(hipMemcpyAsync(ctx->d_ids.p, dids.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(engine_rebuild(ctx->d_sys.p, ctx->d_ids.p, nsys, max_n, engine_maxw(max_cap), ctx->lds_limit, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < nsys; ++k) {
        SysState chk;
        HIPCHK(hipMemcpy(&chk, get_sys(ctx, ids[k])->st.p, sizeof chk, hipMemcpyDeviceToHost));
        if (chk.status) return fail("pdmd_rebuild: system " + std::to_string(ids[k]) + " status " +
                                    std::to_string(chk.status), -3);
    }
    return 0;
}

int pdmd_set_graph_launch(pdmd_ctx* ctx, int32_t on) {
    if (!ctx) return fail("pdmd_set_graph_launch: null context");
    ctx->graphs = on != 0;
    return 0;
}

int pdmd_set_waves(pdmd_ctx* ctx, int32_t waves_per_system) {
    if (!ctx) return fail("pdmd_set_waves: null context");
    if (waves_per_system < 1 || waves_per_system > 8) return fail("pdmd_set_waves: waves per system must be 1 to 8");
    ctx->waves = waves_per_system;
    return 0;
}

int pdmd_wait(pdmd_ctx* ctx) {
    if (!ctx) return fail("pdmd_wait: null context");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int pdmd_advance(pdmd_ctx* ctx, int32_t nsys, const int32_t* ids, const int64_t* nev) {
    int e = pdmd_advance_async(ctx, nsys, ids, nev);
    if (e) return e;
    e = pdmd_wait(ctx);
    if (e) return e;
    for (int k = 0; k < nsys; ++k) {
        HostSys* H = get_sys(ctx, ids[k]);
        int status = 0;
        HIPCHK(hipMemcpy(&status, &H->st.p->status, sizeof(int), hipMemcpyDeviceToHost));
        if (status) return fail("system " + std::to_string(ids[k]) + " stopped with device status " + std::to_string(status), -10 - status);
    }
    return 0;
}

int pdmd_autotune_waves(pdmd_ctx* ctx, int32_t nsys, const int32_t* ids, int64_t probe_events, int32_t* waves_out,
                        double* rates_out) {
    if (!ctx || nsys <= 0 || !ids || probe_events <= 0) return fail("pdmd_autotune_waves: bad arguments");
    int max_n = 0, max_cap = 0;
    for (int k = 0; k < nsys; ++k) {
        HostSys* H = get_sys(ctx, ids[k]);
        if (!H) return -1;
        max_n = std::max(max_n, H->N);
        max_cap = std::max(max_cap, H->cap);
    }
    const int top = engine_max_waves(max_n, engine_maxw(max_cap), ctx->lds_limit);
    const int cand[4] = {1, 2, 4, 8};
    std::vector<int64_t> nev(nsys, probe_events);
    int best = 1;
    double best_rate = -1.0;
    for (int q = 0; q < 4; ++q) {
        if (rates_out) rates_out[q] = 0.0;
        if (cand[q] > top) continue;
        ctx->waves = cand[q];
        int e = pdmd_advance(ctx, nsys, ids, nev.data());
        if (e) return e;
        float ms = 0.0f;
        e = pdmd_last_launch_ms(ctx, &ms);
        if (e) return e;
        const double rate = (double)probe_events * nsys / std::max(1e-3, (double)ms) * 1e3;
        if (rates_out) rates_out[q] = rate;
        if (rate > best_rate) {
            best_rate = rate;
            best = cand[q];
        }
    }
    ctx->waves = best;
    if (waves_out) *waves_out = best;
    return 0;
}

int pdmd_set_system_tuning(pdmd_ctx* ctx, int32_t id, int32_t waves, int32_t cell_factor) {
    HostSys* H = get_sys(ctx, id);
    if (!H) return -1;
    if (waves < 0 || waves > 8) return fail("pdmd_set_system_tuning: waves must be 0 to 8");
    if (cell_factor != 0 && (cell_factor < 2 || cell_factor > 4))
        return fail("pdmd_set_system_tuning: cell factor must be 2 to 4 (0 keeps it)");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    H->dev.waves_cap = waves;
    if (cell_factor) H->dev.cq = cell_factor;
    HIPCHK(hipMemcpy(ctx->d_sys.p + (id - 1), &H->dev, sizeof(SysDev), hipMemcpyHostToDevice));
    return 0;
}

int pdmd_autotune_systems(pdmd_ctx* ctx, int32_t nsys, const int32_t* ids, int64_t probe_events, int32_t* waves_out,
                          int32_t* cell_factor_out, double* rates_out) {
    if (!ctx || nsys <= 0 || !ids || probe_events <= 0 || !waves_out || !cell_factor_out)
        return fail("pdmd_autotune_systems: bad arguments");
    int max_n = 0, max_cap = 0;
    for (int k = 0; k < nsys; ++k) {
        HostSys* H = get_sys(ctx, ids[k]);
        if (!H) return -1;
        max_n = std::max(max_n, H->N);
        max_cap = std::max(max_cap, H->cap);
    }
    const int top = engine_max_waves(max_n, engine_maxw(max_cap), ctx->lds_limit);
    int cand[4] = {1, 2, 4, top > 4 ? top : 0};
    std::vector<int64_t> nev(nsys, probe_events);
    std::vector<SysState> st0(nsys), st1(nsys);
    auto snap = [&](std::vector<SysState>& v) -> int {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        for (int k = 0; k < nsys; ++k)
            HIPCHK(hipMemcpy(&v[k], get_sys(ctx, ids[k])->st.p, sizeof(SysState), hipMemcpyDeviceToHost));
        return 0;
    };
    // wavefronts: every system at each candidate, timed by its own clock
    std::vector<double> best(nsys, -1.0);
    for (int k = 0; k < nsys; ++k) waves_out[k] = 1;
    for (int q = 0; q < 4; ++q) {
        for (int k = 0; k < nsys; ++k)
            if (rates_out) rates_out[4 * k + q] = 0.0;
        if (cand[q] == 0 || cand[q] > top) continue;
        for (int k = 0; k < nsys; ++k)
            if (int e = pdmd_set_system_tuning(ctx, ids[k], cand[q], 0)) return e;
        ctx->waves = cand[q];
        if (int e = snap(st0)) return e;
        if (int e = pdmd_advance(ctx, nsys, ids, nev.data())) return e;
        if (int e = snap(st1)) return e;
        for (int k = 0; k < nsys; ++k) {
            const double sec = (double)(st1[k].run_ticks - st0[k].run_ticks) / ((double)ctx->wall_khz * 1e3);
            const double rate = sec > 0.0 ? (double)(st1[k].coll - st0[k].coll) / sec : 0.0;
            if (rates_out) rates_out[4 * k + q] = rate;
            if (rate > best[k]) {
                best[k] = rate;
                waves_out[k] = cand[q];
            }
        }
    }
    int wmax = 1;
    for (int k = 0; k < nsys; ++k) wmax = std::max(wmax, waves_out[k]);
    for (int k = 0; k < nsys; ++k)
        if (int e = pdmd_set_system_tuning(ctx, ids[k], waves_out[k], 0)) return e;
    ctx->waves = wmax;
    // rebuild cell factor: mean rebuild time at each factor, per system
    std::vector<double> rb(nsys, 1e300);
    for (int k = 0; k < nsys; ++k) cell_factor_out[k] = get_sys(ctx, ids[k])->dev.cq;
    bool any = false;
    for (int k = 0; k < nsys; ++k) any |= get_sys(ctx, ids[k])->N > 256;
    if (any)
        for (int cf = 2; cf <= 4; ++cf) {
            for (int k = 0; k < nsys; ++k)
                if (int e = pdmd_set_system_tuning(ctx, ids[k], waves_out[k], cf)) return e;
            if (int e = snap(st0)) return e;
            if (int e = pdmd_advance(ctx, nsys, ids, nev.data())) return e;
            if (int e = snap(st1)) return e;
            for (int k = 0; k < nsys; ++k) {
                const long long n = st1[k].rebuilds - st0[k].rebuilds;
                if (get_sys(ctx, ids[k])->N <= 256 || n <= 0) continue;
                const double per = (double)(st1[k].rebuild_ticks - st0[k].rebuild_ticks) / (double)n;
                if (per < rb[k]) {
                    rb[k] = per;
                    cell_factor_out[k] = cf;
                }
            }
        }
    for (int k = 0; k < nsys; ++k)
        if (int e = pdmd_set_system_tuning(ctx, ids[k], waves_out[k], cell_factor_out[k])) return e;
    return 0;
}

int pdmd_last_launch_ms(pdmd_ctx* ctx, float* ms) {
    if (!ctx || !ms) return fail("pdmd_last_launch_ms: null argument");
    if (!ctx->timed) return fail("pdmd_last_launch_ms: no launch yet");
    HIPCHK(hipEventSynchronize(ctx->ev1));
    HIPCHK(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return 0;
}

int pdmd_download(pdmd_ctx* ctx, int32_t id, double* sv, int32_t* identity, int32_t* bptnr, int32_t* extra_repuls,
                  double* tim, int32_t* nptnr, int32_t* coltype) {
    HostSys* H = get_sys(ctx, id);
    if (!H) return -1;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int N = H->N;
    std::vector<Rec> b(N);
    HIPCHK(hipMemcpy(b.data(), H->rec.p, sizeof(Rec) * N, hipMemcpyDeviceToHost));
    for (int k = 0; k < N; ++k) {
        if (sv) {
            sv[6 * k] = b[k].x;
            sv[6 * k + 1] = b[k].y;
            sv[6 * k + 2] = b[k].z;
            sv[6 * k + 3] = b[k].vx;
            sv[6 * k + 4] = b[k].vy;
            sv[6 * k + 5] = b[k].vz;
        }
        if (identity) identity[k] = dyn_ident(b[k].dyn);
        if (bptnr) bptnr[k] = dyn_bptnr(b[k].dyn) + 1;
    }
    if (extra_repuls) {
        std::vector<int4> e(N);
        HIPCHK(hipMemcpy(e.data(), H->extra.p, sizeof(int4) * N, hipMemcpyDeviceToHost));
        for (int k = 0; k < N; ++k) {
            extra_repuls[4 * k] = e[k].x + 1;
            extra_repuls[4 * k + 1] = e[k].y + 1;
            extra_repuls[4 * k + 2] = e[k].z + 1;
            extra_repuls[4 * k + 3] = e[k].w + 1;
        }
    }
    if (tim || nptnr || coltype) {
        std::vector<Leaf> lf(N + 3);
        HIPCHK(hipMemcpy(lf.data(), H->leaf.p, sizeof(Leaf) * (N + 3), hipMemcpyDeviceToHost));
        for (int k = 0; k < N + 3; ++k) {
            if (tim) tim[k] = lf[k].tim;
            if (nptnr) nptnr[k] = lf[k].nptnr >= 0 ? lf[k].nptnr + 1 : lf[k].nptnr;
            if (coltype) coltype[k] = leaf_ct(lf[k].meta);
        }
    }
    return 0;
}

int pdmd_stats(pdmd_ctx* ctx, int32_t id, pdmd_stats_t* out) {
    HostSys* H = get_sys(ctx, id);
    if (!H || !out) return -1;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    SysState st;
    HIPCHK(hipMemcpy(&st, H->st.p, sizeof st, hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof *out);
    out->coll = st.coll;
    long long pe = 0;
    for (int k = 1; k <= 30; ++k) {
        out->nevents[k] = st.nevents[k];
        pe += st.nevents[k];
    }
    out->pair_events = pe;
    out->numghosts = st.numghosts;
    out->nupdates = st.nupdates;
    out->nforced = st.nforced;
    out->t = st.t;
    out->tfalse = st.tfalse;
    out->interval = st.interval;
    out->interval_max = st.interval_max;
    out->status = st.status;
    out->max_list = (int32_t)st.max_nb_seen;
    out->nbeads = H->N;
    std::vector<int> nu(H->N), nd(H->N);
    HIPCHK(hipMemcpy(nu.data(), H->nup.p, sizeof(int) * H->N, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(nd.data(), H->ndn.p, sizeof(int) * H->N, hipMemcpyDeviceToHost));
    long long su = 0, sd = 0;
    for (int k = 0; k < H->N; ++k) {
        su += nu[k];
        sd += nd[k];
    }
    out->avg_up = (double)su / H->N;
    out->avg_dn = (double)sd / H->N;
    for (int k = 0; k < kPhaseSlots; ++k) out->phase_cycles[k] = st.phase[k];
    out->device_ms = (double)st.run_ticks / (double)ctx->wall_khz;
    out->rebuild_ms = (double)st.rebuild_ticks / (double)ctx->wall_khz;
    out->rebuilds = st.rebuilds;
    out->waves = H->dev.waves_cap;
    out->cell_factor = H->dev.cq;
    out->cells_axis = H->dev.M;
    return 0;
}

int pdmd_log_enable(pdmd_ctx* ctx, int32_t id, int32_t capacity) {
    HostSys* H = get_sys(ctx, id);
    if (!H) return -1;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (capacity < 0) return fail("pdmd_log_enable: negative capacity");
    if (H->log.p) {
        (void)hipFree(H->log.p);
        H->log.p = nullptr;
        H->log.n = 0;
    }
    if (capacity > 0) HIPCHK(H->log.alloc(capacity));
    H->dev.log = capacity > 0 ? H->log.p : nullptr;
    H->dev.log_cap = capacity;
    int zero = 0;
    HIPCHK(hipMemcpy(&H->st.p->log_count, &zero, sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_sys.p + (id - 1), &H->dev, sizeof(SysDev), hipMemcpyHostToDevice));
    return 0;
}

int pdmd_log_read(pdmd_ctx* ctx, int32_t id, int32_t max, pdmd_event_t* out, int32_t* n) {
    HostSys* H = get_sys(ctx, id);
    if (!H || !out || !n) return -1;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int cnt = 0;
    HIPCHK(hipMemcpy(&cnt, &H->st.p->log_count, sizeof(int), hipMemcpyDeviceToHost));
    cnt = std::min(cnt, H->dev.log_cap);
    cnt = std::min(cnt, max);
    if (cnt > 0) {
        std::vector<LogRec> r(cnt);
        HIPCHK(hipMemcpy(r.data(), H->log.p, sizeof(LogRec) * cnt, hipMemcpyDeviceToHost));
        for (int k = 0; k < cnt; ++k) {
            out[k].coll = r[k].coll;
            out[k].i = r[k].i;
            out[k].j = r[k].j;
            out[k].type = r[k].type;
            out[k].pad = 0;
            out[k].t = r[k].t;
        }
    }
    *n = cnt;
    return 0;
}

int pdmd_list_capacity(pdmd_ctx* ctx, int32_t id) {
    HostSys* H = get_sys(ctx, id);
    return H ? H->cap : -1;
}

int pdmd_lists(pdmd_ctx* ctx, int32_t id, int32_t* n_up, int32_t* up, int32_t* n_dn, int32_t* dn) {
    HostSys* H = get_sys(ctx, id);
    if (!H) return -1;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const size_t N = (size_t)H->N, cap = (size_t)H->cap;
    if (n_up) HIPCHK(hipMemcpy(n_up, H->nup.p, sizeof(int) * N, hipMemcpyDeviceToHost));
    if (n_dn) HIPCHK(hipMemcpy(n_dn, H->ndn.p, sizeof(int) * N, hipMemcpyDeviceToHost));
    // rows hold partner | static code << 24; the ABI gives 1-based partners
    auto rows = [&](const int* src, int32_t* dst) -> int {
        std::vector<int> r(N * cap);
        HIPCHK(hipMemcpy(r.data(), src, sizeof(int) * N * cap, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < N * cap; ++k) dst[k] = (r[k] & 0xFFFFFF) + 1;
        return 0;
    };
    if (up)
        if (int e = rows(H->nb_up.p, up)) return e;
    if (dn)
        if (int e = rows(H->nb_dn.p, dn)) return e;
    return 0;
}

int pdmd_predict_pairs(pdmd_ctx* ctx, int32_t id, int32_t n, const int32_t* pi, const int32_t* pj, double* tij,
                       int32_t* type) {
    HostSys* H = get_sys(ctx, id);
    if (!H || n < 0 || (n > 0 && (!pi || !pj || !tij || !type))) return fail("pdmd_predict_pairs: bad arguments");
    if (n == 0) return 0;
    for (int k = 0; k < n; ++k)
        if (pi[k] < 1 || pj[k] < 1 || pi[k] > H->N || pj[k] > H->N || pi[k] == pj[k])
            return fail("pdmd_predict_pairs: pair index out of range");
    DevBuf<int> di, dj, dt;
    DevBuf<double> dtij;
    HIPCHK(di.alloc(n));
    HIPCHK(dj.alloc(n));
    HIPCHK(dt.alloc(n));
    HIPCHK(dtij.alloc(n));
    HIPCHK(hipMemcpy(di.p, pi, sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dj.p, pj, sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(engine_predict(ctx->d_sys.p, id - 1, n, di.p, dj.p, dtij.p, dt.p, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(tij, dtij.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(type, dt.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    return 0;
}

int pdmd_check(pdmd_ctx* ctx, int32_t id, const double* sv, pdmd_check_t* out) {
    HostSys* H = get_sys(ctx, id);
    if (!H || !out) return fail("pdmd_check: bad arguments");
    const int N = H->N;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    DevBuf<Rec> alt;
    if (sv) {
        // the caller's sv(6,N) in place of the resident positions/velocities
        std::vector<Rec> b(N);
        HIPCHK(hipMemcpy(b.data(), H->rec.p, sizeof(Rec) * N, hipMemcpyDeviceToHost));
        for (int k = 0; k < N; ++k) {
            const double* q = sv + 6 * (size_t)k;
            b[k].x = q[0];
            b[k].y = q[1];
            b[k].z = q[2];
            b[k].vx = q[3];
            b[k].vy = q[4];
            b[k].vz = q[5];
        }
        HIPCHK(alt.alloc(N));
        HIPCHK(hipMemcpy(alt.p, b.data(), sizeof(Rec) * N, hipMemcpyHostToDevice));
    }
    DevBuf<unsigned long long> cnt;
    HIPCHK(cnt.alloc(8));
    unsigned long long init[8] = {0, 0, 0, ~0ull, 0, 0, 0, 0};
    HIPCHK(hipMemcpy(cnt.p, init, sizeof(init), hipMemcpyHostToDevice));
    HIPCHK(engine_check(ctx->d_sys.p, id - 1, N, sv ? alt.p : nullptr, cnt.p, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    unsigned long long r[8];
    HIPCHK(hipMemcpy(r, cnt.p, sizeof(r), hipMemcpyDeviceToHost));
    out->overlaps = (int64_t)r[0];
    out->bonds_long = (int64_t)r[1];
    out->bonds_short = (int64_t)r[2];
    out->first_i = r[3] == ~0ull ? 0 : (int32_t)(r[3] / (unsigned long long)N) + 1;
    out->first_j = r[3] == ~0ull ? 0 : (int32_t)(r[3] % (unsigned long long)N) + 1;
    out->boundbad = (int64_t)r[4];
    out->unboundbad = (int64_t)r[5];
    out->ss_counted = (int64_t)r[6];
    out->ss_total = (int64_t)r[7];
    return 0;
}

int pdmd_observables(pdmd_ctx* ctx, int32_t id, pdmd_obs_t* o) {
    HostSys* H = get_sys(ctx, id);
    if (!H || !o) return fail("pdmd_observables: bad arguments");
    const Topology& t = H->topo;
    const int nch = (int)t.chain_len.size();
    int nfirst = 0;  // chains of the leading species block: the reference's species 1
    while (nfirst < nch && t.species[nfirst] == t.species[0]) nfirst++;
    DevBuf<int> cs, cl, oi;
    DevBuf<double> od;
    HIPCHK(cs.alloc(nch));
    HIPCHK(cl.alloc(nch));
    HIPCHK(od.alloc(7));
    HIPCHK(oi.alloc(3));
    HIPCHK(hipMemcpyAsync(cs.p, t.chain_start.data(), sizeof(int) * nch, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(cl.p, t.chain_len.data(), sizeof(int) * nch, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(engine_observables(ctx->d_sys.p, id - 1, cs.p, cl.p, nch, nfirst, H->p.boxl_orig, od.p, oi.p, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double d[7];
    int n[3];
    HIPCHK(hipMemcpy(d, od.p, sizeof d, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(n, oi.p, sizeof n, hipMemcpyDeviceToHost));
    std::memset(o, 0, sizeof *o);
    o->ered = d[0];
    o->tred = d[1];
    o->sumvel = d[2];
    o->ehh_ii = d[3];
    o->ehh_ij = d[4];
    o->rg_avg = d[5];
    o->e2e_avg = d[6];
    o->hb_alpha = n[0];
    o->hb_ii = n[1];
    o->hb_ij = n[2];
    return 0;
}

int pdmd_observables_noalloc(pdmd_ctx* ctx, int32_t id, pdmd_obs_t* o) {
    HostSys* H = get_sys(ctx, id);
    if (!H || !o) return fail("pdmd_observables_noalloc: bad arguments");
    const Topology& t = H->topo;
    const int nch = (int)t.chain_len.size();
    if (nch > 4096) return fail("pdmd_observables_noalloc: at most 4096 chains (the staging array's size)");
    int nfirst = 0;  // chains of the leading species block: the reference's species 1
    while (nfirst < nch && t.species[nfirst] == t.species[0]) nfirst++;
    double d[7];
    int n[3];
    HIPCHK(engine_observables_noalloc(ctx->d_sys.p, id - 1, ctx->device, t.chain_start.data(), t.chain_len.data(), nch,
                                      nfirst, H->p.boxl_orig, d, n, ctx->stream));
    std::memset(o, 0, sizeof *o);
    o->ered = d[0];
    o->tred = d[1];
    o->sumvel = d[2];
    o->ehh_ii = d[3];
    o->ehh_ij = d[4];
    o->rg_avg = d[5];
    o->e2e_avg = d[6];
    o->hb_alpha = n[0];
    o->hb_ii = n[1];
    o->hb_ij = n[2];
    return 0;
}

int pdmd_pair_codes(const pdmd_system_desc* d, int32_t n, const int32_t* a, const int32_t* b, int32_t* code) {
    if (int e = validate_desc(d)) return e;
    if (n < 0 || (n > 0 && (!a || !b || !code))) return fail("pdmd_pair_codes: bad arguments");
    const int nc = d->nchains;
    std::vector<int> clen(d->chain_len, d->chain_len + nc), spec(nc);
    for (int c = 0; c < nc; ++c) spec[c] = d->species ? d->species[c] : c;
    int nres = 0;
    for (int c = 0; c < nc; ++c) nres += clen[c];
    std::vector<int> rt(d->restype, d->restype + nres), hp(d->hp, d->hp + nres);
    Topology t;
    t.build(clen, spec, rt, hp);
    t.no_hbs = (d->flags & PDMD_NO_HBS) != 0;
    t.no_h = (d->flags & PDMD_NO_H) != 0;
    t.chaptype2 = (d->flags & PDMD_CHAPTYPE2) != 0;
    for (int k = 0; k < n; ++k) {
        if (a[k] < 1 || a[k] > t.N || b[k] < 1 || b[k] > t.N) return fail("pdmd_pair_codes: bead index out of range");
        code[k] = a[k] == b[k] ? 1 : static_code(t, a[k] - 1, b[k] - 1);
    }
    return 0;
}

int pdmd_generate_seq(const char* param_dir, int32_t nchains, const int32_t* chain_len, const int32_t* restype_in,
                      double density, double tstar, uint64_t seed, int32_t* restype, int32_t* hp, double* sv,
                      double* boxl) {
    if (!param_dir || nchains <= 0 || !chain_len || !restype || !hp || !sv || !boxl)
        return fail("pdmd_generate: bad arguments");
    Params raw;
    std::string err;
    if (!load_params(raw, param_dir, err)) return fail(err);
    GenSpec g;
    g.chain_len.assign(chain_len, chain_len + nchains);
    if (restype_in) {
        int nres = 0;
        for (int c = 0; c < nchains; ++c) nres += chain_len[c];
        g.restype.assign(restype_in, restype_in + nres);
    }
    g.density = density;
    g.tstar = tstar;
    g.seed = seed;
    GenOut o;
    if (!generate(raw, g, o, err)) return fail(err);
    std::memcpy(restype, o.restype.data(), sizeof(int) * o.restype.size());
    std::memcpy(hp, o.hp.data(), sizeof(int) * o.hp.size());
    std::memcpy(sv, o.sv.data(), sizeof(double) * o.sv.size());
    *boxl = o.boxl;
    return 0;
}

int pdmd_generate(const char* param_dir, int32_t nchains, const int32_t* chain_len, double density, double tstar,
                  uint64_t seed, int32_t* restype, int32_t* hp, double* sv, double* boxl) {
    return pdmd_generate_seq(param_dir, nchains, chain_len, nullptr, density, tstar, seed, restype, hp, sv, boxl);
}

}  // extern "C"
