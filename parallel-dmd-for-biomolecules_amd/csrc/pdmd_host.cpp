// pdmd_host.cpp -- host-side set-up of PRIME20 systems (see pdmd_host.h).
// Floating-point expressions keep the reference's evaluation order; the
// library is compiled with -ffp-contract=off.
#include "pdmd_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_map>

namespace pdmd {

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// ---------------------------------------------------------------------------
// Parameter files.  Fortran fixed-width fields: blanks ignored, explicit
// decimal point present in every shipped value.
// ---------------------------------------------------------------------------
static double field(const std::string& line, int col1, int width) {
    std::string s;
    for (int c = col1 - 1; c < col1 - 1 + width && c < (int)line.size(); ++c)
        if (line[c] != ' ' && line[c] != '\r' && line[c] != '\n' && line[c] != '\t') s.push_back(line[c]);
    if (s.empty()) return 0.0;
    return std::strtod(s.c_str(), nullptr);
}

static bool read_lines(const std::string& path, std::vector<std::string>& out) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    char buf[1024];
    while (std::fgets(buf, sizeof buf, f)) out.emplace_back(buf);
    std::fclose(f);
    return true;
}

bool load_params(Params& p, const std::string& dir, std::string& err) {
    std::memset(&p, 0, sizeof p);
    std::vector<std::string> L;
    // protein.data: 12 list-directed reals (inputinfo.f:141-176)
    {
        FILE* f = std::fopen((dir + "/protein.data").c_str(), "r");
        if (!f) { err = "missing protein.data in " + dir; return false; }
        double v[12];
        for (int k = 0; k < 12; ++k)
            if (std::fscanf(f, "%lf", &v[k]) != 1) { std::fclose(f); err = "bad protein.data"; return false; }
        std::fclose(f);
        for (int k = 1; k <= 4; ++k) {
            p.sigma[k] = v[k - 1];
            p.welldia[k] = v[3 + k];
            p.epsilon[k] = v[7 + k];
        }
        for (int k = 1; k <= 4; ++k) {  // inputinfo.f:178-189
            p.sigma[k + 4] = p.sigma[k];
            p.epsilon[k + 4] = p.epsilon[k];
            p.welldia[k + 4] = p.welldia[k];
        }
    }
    // ep19p_ha55a_weakhp.data (2(2x,i2,2x),f8.3), stored negated (inputinfo.f:282-288)
    L.clear();
    if (!read_lines(dir + "/ep19p_ha55a_weakhp.data", L) || L.size() < 400) { err = "bad ep19p_ha55a_weakhp.data"; return false; }
    for (int k = 0; k < 400; ++k) {
        int a = (int)field(L[k], 3, 2), b = (int)field(L[k], 9, 2);
        if (a < 1 || a > 28 || b < 1 || b > 28) { err = "bad ep table index"; return false; }
        p.ep[a][b] = -field(L[k], 13, 8);
    }
    // rcarnrco.data (6(f6.3,2x)); tolerances floored at del (inputinfo.f:292-298)
    L.clear();
    if (!read_lines(dir + "/rcarnrco.data", L) || L.size() < 20) { err = "bad rcarnrco.data"; return false; }
    for (int k = 1; k <= 20; ++k) {
        const std::string& s = L[k - 1];
        p.drca[k] = field(s, 1, 6);
        p.drnh[k] = field(s, 9, 6);
        p.drco[k] = field(s, 17, 6);
        p.del_rca[k] = field(s, 25, 6);
        p.del_rnh[k] = field(s, 33, 6);
        p.del_rco[k] = field(s, 41, 6);
        if (p.del_rca[k] < kDel) p.del_rca[k] = kDel;
        if (p.del_rnh[k] < kDel) p.del_rnh[k] = kDel;
        if (p.del_rco[k] < kDel) p.del_rco[k] = kDel;
    }
    // beadwell_ha55a.data (2(2x,i2,2x),2(f6.3,2x)) (inputinfo.f:352-363)
    L.clear();
    if (!read_lines(dir + "/beadwell_ha55a.data", L) || L.size() < 400) { err = "bad beadwell_ha55a.data"; return false; }
    for (int k = 0; k < 400; ++k) {
        int a = (int)field(L[k], 3, 2), b = (int)field(L[k], 9, 2);
        if (a < 1 || a > 28 || b < 1 || b > 28) { err = "bad beadwell index"; return false; }
        p.bds[a][b] = field(L[k], 13, 6);
        p.wel[a][b] = field(L[k], 21, 6);
    }
    for (int i = 9; i <= 28; ++i) {  // inputinfo.f:365-370
        p.sigma[i] = 1.00 * p.bds[i][i];
        p.welldia[i] = 1.5 * p.sigma[i];
    }
    p.sigma[9] = 1.00 * p.bds[9][9] - 1.2;
    // sqz6to10.data read as sz8,sz6,sz7,sz9,sz10 (inputinfo.f:371-382)
    L.clear();
    if (!read_lines(dir + "/sqz6to10.data", L) || L.size() < 20) { err = "bad sqz6to10.data"; return false; }
    for (int i = 1; i <= 20; ++i) {
        const std::string& s = L[i - 1];
        double sz8 = field(s, 1, 6), sz6 = field(s, 9, 6), sz7 = field(s, 17, 6), sz9 = field(s, 25, 6),
               sz10 = field(s, 33, 6);
        p.sqz610[1][i + 8] = sz6 * 2 / (p.sigma[i + 8] + p.sigma[4]);
        p.sqz610[2][i + 8] = sz7 * 2 / (p.sigma[i + 8] + p.sigma[1]);
        p.sqz610[3][i + 8] = sz8 * 2 / (p.sigma[i + 8] + p.sigma[2]);
        p.sqz610[4][i + 8] = sz9 * 2 / (p.sigma[i + 8] + p.sigma[2]);
        p.sqz610[5][i + 8] = sz10 * 2 / (p.sigma[i + 8] + p.sigma[4]);
    }
    // mass.data (4x,i2,2x,f8.3) (inputinfo.f:385-398)
    L.clear();
    if (!read_lines(dir + "/mass.data", L) || L.size() < 23) { err = "bad mass.data"; return false; }
    for (int i = 0; i < 23; ++i) {
        int a = (int)field(L[i], 5, 2);
        if (a < 1 || a > 28) { err = "bad mass index"; return false; }
        p.bmass[a] = field(L[i], 9, 8);
    }
    p.bmass[3] = p.bmass[20];
    for (int i = 1; i <= 4; ++i) p.bmass[i + 4] = p.bmass[i];
    return true;
}

void scale_params(Params& p, double boxl) {
    // scale_down.f:16-83
    const double B = boxl;
    p.boxl_orig = B;
    for (int k = 1; k <= 28; ++k) {
        p.sigma[k] = p.sigma[k] / B;
        p.welldia[k] = p.welldia[k] / B;
    }
    p.shder_dist[1] = 5.00 / B;
    p.shder_dist[2] = 4.74 / B;
    p.shder_dist[3] = 4.86 / B;
    p.shder_dist[4] = 4.83 / B;
    for (int k = 1; k <= 28; ++k)
        for (int kk = 1; kk <= 28; ++kk) {
            const double s = p.sigma[k] + p.sigma[kk], w = p.welldia[k] + p.welldia[kk];
            p.sigma_sq[k][kk] = 0.25 * (s * s);
            p.sigma_2b[k][kk] = 0.50 * s;
            p.welldia_sq[k][kk] = 0.25 * (w * w);
            p.ep_sqrt[k][kk] = std::sqrt(p.epsilon[k] * p.epsilon[kk]);
            p.shlddia_sq[k][kk] = p.shder_dist[4] * p.shder_dist[4];
        }
    for (int k = 9; k <= 28; ++k)
        for (int kk = 9; kk <= 28; ++kk) {
            p.ep_sqrt[k][kk] = p.epsilon[1] * p.ep[k][kk];
            p.sigma_sq[k][kk] = p.bds[k][kk] * p.bds[k][kk] / (B * B);
            p.sigma_2b[k][kk] = p.bds[k][kk] / B;
            p.welldia_sq[k][kk] = p.wel[k][kk] * p.wel[k][kk] / (B * B);
        }
    const double d1 = p.shder_dist[1] * p.shder_dist[1], d2 = p.shder_dist[2] * p.shder_dist[2],
                 d3 = p.shder_dist[3] * p.shder_dist[3];
    p.shlddia_sq[1][2] = p.shlddia_sq[2][1] = p.shlddia_sq[5][2] = p.shlddia_sq[2][5] = d1;
    p.shlddia_sq[1][1] = p.shlddia_sq[5][5] = p.shlddia_sq[1][5] = p.shlddia_sq[5][1] = d2;
    p.shlddia_sq[2][4] = p.shlddia_sq[4][2] = p.shlddia_sq[2][8] = p.shlddia_sq[8][2] = d3;
    // make_code.f:18-69
    for (int k = 1; k <= 3; ++k)
        for (int l = 1; l <= 50; ++l) p.ev_param[k][l] = 1;
    p.ev_param[1][15] = 1.05 * ((2.24 / B) / ((p.sigma[1] + p.sigma[4]) / 2.0));
    p.ev_param[1][17] = 1.1436;
    p.ev_param[1][18] = 0.88;
    p.ev_param[1][19] = 0.87829;
    p.ev_param[1][20] = 0.8;
    p.ev_param[1][21] = 0.7713;
    p.ev_param[1][27] = 1.0;
    const double lo = 1.0 - kDel, hi = 1.0 + kDel;
    p.ev_param[2][4] = 1.46 * lo / B;
    p.ev_param[2][5] = 1.51 * lo / B;
    p.ev_param[2][6] = 1.33 * lo / B;
    p.ev_param[2][7] = 2.41 * lo / B;
    p.ev_param[2][8] = 2.45 * lo / B;
    p.ev_param[2][9] = 3.8 * lo / B;
    p.ev_param[2][10] = p.ev_param[2][11] = p.ev_param[2][12] = lo;
    p.ev_param[3][4] = 1.46 * hi / B;
    p.ev_param[3][5] = 1.51 * hi / B;
    p.ev_param[3][6] = 1.33 * hi / B;
    p.ev_param[3][7] = hi * 2.41 / B;
    p.ev_param[3][8] = hi * 2.45 / B;
    p.ev_param[3][9] = hi * 3.8 / B;
    p.ev_param[3][10] = p.ev_param[3][11] = p.ev_param[3][12] = hi;
}

// ---------------------------------------------------------------------------
// Topology
// ---------------------------------------------------------------------------
void Topology::build(const std::vector<int>& clen, const std::vector<int>& spec, const std::vector<int>& rt,
                     const std::vector<int>& hpin) {
    chain_len = clen;
    species = spec;
    restype = rt;
    hp = hpin;
    const int nc = (int)clen.size();
    chain_start.assign(nc, 0);
    N = 0;
    nres = 0;
    for (int c = 0; c < nc; ++c) {
        chain_start[c] = N;
        N += 4 * clen[c];
        nres += clen[c];
    }
    nfirst = 0;
    while (nfirst < nc && spec[nfirst] == spec[0]) ++nfirst;
    chain.assign(N, 0);
    loc.assign(N, 0);
    res.assign(N, 0);
    ident0.assign(N, 0);
    int r0 = 0;
    for (int c = 0; c < nc; ++c) {
        const int L = clen[c], s = chain_start[c];
        for (int k = 1; k <= 4 * L; ++k) {
            const int b = s + k - 1, rr = (k - 1) % L;
            chain[b] = c;
            loc[b] = k;
            res[b] = r0 + rr;
            ident0[b] = k <= L ? 2 : (k <= 2 * L ? 1 : (k <= 3 * L ? 4 : rt[r0 + rr]));
        }
        r0 += L;
    }
}

// kind: 0 Calpha, 1 N, 2 C, 3 R; residue index 1..L within the chain
static inline void kind_of(const Topology& t, int b, int& kind, int& r) {
    const int L = t.chain_len[t.chain[b]];
    const int k = t.loc[b] - 1;
    kind = k / L;
    r = k % L + 1;
}

int static_code(const Topology& t, int a, int b) {
    // Closed form of the dense matrix built by make_code.f:72-563 (upper
    // triangle then mirrored), initial (unbonded) identities.
    int ka, ra, kb, rb;
    kind_of(t, a, ka, ra);
    kind_of(t, b, kb, rb);
    if (t.chain[a] == t.chain[b]) {
        // -Dchaptype=2: nothing inside a second-species chain (make_code.f:
        // 349,366; no firstside2.data is read, inputinfo.f:120-132)
        if (t.free_bead(a)) return 1;
        if (ka > kb || (ka == kb && ra > rb)) {
            std::swap(ka, kb);
            std::swap(ra, rb);
        }
        const int d = rb - ra;  // residue offset of the second kind relative to the first
        switch (ka * 4 + kb) {
            case 0 * 4 + 0: return (d == 1) ? 9 : 1;                                     // Ca-Ca
            case 0 * 4 + 1: return d == 0 ? 4 : (d == 1 ? 7 : (d == -1 ? 18 : 1));       // Ca-N
            case 0 * 4 + 2: return d == 0 ? 5 : (d == -1 ? 8 : (d == 1 ? 17 : 1));       // Ca-C
            case 0 * 4 + 3: return d == 0 ? 10 : (d == 1 ? 24 : (d == -1 ? 25 : 1));     // Ca-R
            case 1 * 4 + 1: return (d == 1) ? 20 : 1;                                    // N-N
            case 1 * 4 + 2:                                                              // N-C
                if (d == -1) return 6;
                if (d == 0) return 8;
                if (d == -2) return 19;
                return (d >= -3 && d <= 3) || t.no_hbs ? 1 : 15;
            case 1 * 4 + 3: return d == 0 ? 11 : (d == -1 ? 23 : 1);                    // N-R
            case 2 * 4 + 2: return (d == 1) ? 21 : 1;                                    // C-C
            case 2 * 4 + 3: return d == 0 ? 12 : (d == 1 ? 22 : (d == 2 ? 26 : 1));     // C-R
            case 3 * 4 + 3:                                                              // R-R
                return (d >= 4 && t.hp[t.res[a]] == 1 && t.hp[t.res[b]] == 1) ? 16 : 1;
            default: return 1;
        }
    }
    // different chains
    // -Dno_h zeroes hp for make_code.f:80-104 only (inputinfo.f:270-280);
    // the same-chain 16s above come from hp1/hp2, which it leaves alone
    // -Dchaptype=2: none between two second-species chains (make_code.f:519-547)
    if (ka == 3 && kb == 3)
        return (!t.no_h && t.hp[t.res[a]] == 1 && t.hp[t.res[b]] == 1 && !(t.free_bead(a) && t.free_bead(b))) ? 16 : 1;
    if ((ka == 1 && kb == 2) || (ka == 2 && kb == 1)) {
        if (t.no_hbs) return 1;  // make_code.f:114-170 skipped under -Dno_hbs
        // proline NH rule (make_code.f:139-161): kept only where the Pro side
        // chain precedes the CO bead in index order, i.e. the N's chain first.
        const int nb = ka == 1 ? a : b, cb = ka == 1 ? b : a;
        if (t.restype[t.res[nb]] == 17 && t.chain[nb] < t.chain[cb]) return 1;
        return 15;
    }
    return 1;
}

void half_stencil(signed char out[125]) {
    for (int k = 0; k < 125; ++k) out[k] = 0;
    // the offsets in cell_link.f's two loops, in its order: first the
    // nearest shell (dz in 0, 1, -1; dy in 0, 1; dx in 0, 1, -1), then the
    // second shell (dz in 0, 1, -1, 2, -2; dy in 0, 1, 2; dx in 2, -2 only
    // where |dz| and dy stay within the first shell); with dy = 0 it keeps
    // dx >= 1 for dz <= 0 and dx >= 0 for dz >= 1, so that each pair of
    // cells is scanned from one side
    static const int o[5] = {0, 1, -1, 2, -2}, oy[3] = {0, 1, 2};
    int n = 0;
    auto put = [&](int dx, int dy, int dz) {
        if (dy == 0 && (dz < 1 ? dx < 1 : dx < 0)) return;
        out[(dx + 2) + 5 * (dy + 2) + 25 * (dz + 2)] = (signed char)++n;
    };
    for (int iz = 0; iz < 3; ++iz)
        for (int iy = 0; iy < 2; ++iy)
            for (int ix = 0; ix < 3; ++ix) put(o[ix], o[iy], o[iz]);
    for (int iz = 0; iz < 5; ++iz)
        for (int iy = 0; iy < 3; ++iy)
            for (int ix = (iz < 3 && iy != 2) ? 3 : 0; ix < 5; ++ix) put(o[ix], oy[iy], o[iz]);
}

void nbor_setup(const Params& p, const Topology& t, const std::vector<int>& ident, Cutoffs& c) {
    // nbor_setup.f:13-118.  Scan domain: the first chain of every species
    // (intra-chain pairs) and the first chains of every pair of species.
    double sig_max[51];
    for (int i = 0; i <= 50; ++i) sig_max[i] = 0.0;
    c.sig_max_all = 0.0;
    auto consider = [&](int i, int j, int code) {
        if (code > 26) return;
        const int a = ident[i], b = ident[j];
        double s;
        if (code <= 3) s = p.sigma_2b[a][b];
        else if (code == 15) s = 0.5 * (p.welldia[a] + p.welldia[b]);
        else if (code == 16) s = p.wel[a][b] / p.boxl_orig;
        else s = 0.0;
        if (s > sig_max[code]) {
            sig_max[code] = s;
            if (sig_max[code] > c.sig_max_all) c.sig_max_all = sig_max[code];
        }
    };
    std::vector<int> firsts;
    {
        std::unordered_map<int, int> seen;
        for (int ch = 0; ch < (int)t.chain_len.size(); ++ch)
            if (seen.emplace(t.species[ch], ch).second) firsts.push_back(ch);
    }
    for (int ch : firsts) {
        const int s = t.chain_start[ch], n = 4 * t.chain_len[ch];
        for (int i = s; i < s + n - 1; ++i)
            for (int j = i + 1; j < s + n; ++j) consider(i, j, static_code(t, j, i));
    }
    // Cross pairs between species: the code of an inter-chain pair depends
    // on the two bead kinds, side-chain hp flags and the proline rule only,
    // so the max over all pairs equals the max over the distinct
    // (identity, hp/proline class) combinations present in each first chain.
    struct Cls { int ident, kind, hp, pro; };
    std::vector<std::vector<Cls>> cls(firsts.size());
    for (size_t q = 0; q < firsts.size(); ++q) {
        const int ch = firsts[q], s = t.chain_start[ch], n = 4 * t.chain_len[ch];
        std::vector<Cls>& v = cls[q];
        for (int i = s; i < s + n; ++i) {
            int kind, r;
            kind_of(t, i, kind, r);
            Cls x{ident[i], kind, kind == 3 ? t.hp[t.res[i]] : 0,
                  kind == 1 ? (t.restype[t.res[i]] == 17 ? 1 : 0) : 0};
            bool dup = false;
            for (const Cls& y : v)
                if (y.ident == x.ident && y.kind == x.kind && y.hp == x.hp && y.pro == x.pro) dup = true;
            if (!dup) v.push_back(x);
        }
    }
    for (size_t qa = 0; qa < firsts.size(); ++qa)
        for (size_t qb = qa + 1; qb < firsts.size(); ++qb)
            for (const Cls& x : cls[qa])
                for (const Cls& y : cls[qb]) {
                    int code = 1;
                    if (x.kind == 3 && y.kind == 3) code = (!t.no_h && x.hp == 1 && y.hp == 1) ? 16 : 1;
                    else if (x.kind == 1 && y.kind == 2) code = x.pro || t.no_hbs ? 1 : 15;  // N in the earlier chain
                    else if (x.kind == 2 && y.kind == 1) code = t.no_hbs ? 1 : 15;
                    const int a = x.ident, b = y.ident;
                    if (code > 26) continue;
                    double s;
                    if (code <= 3) s = p.sigma_2b[b][a];
                    else if (code == 15) s = 0.5 * (p.welldia[a] + p.welldia[b]);
                    else s = p.wel[a][b] / p.boxl_orig;
                    if (s > sig_max[code]) {
                        sig_max[code] = s;
                        if (sig_max[code] > c.sig_max_all) c.sig_max_all = sig_max[code];
                    }
                }
    double sigij = 0.0, sig = 0.0;
    for (int i = 1; i <= 4; ++i)
        for (int j = 1; j <= 4; ++j) {
            if (i != 3 || j != 3) sig = p.sigma[i] + p.sigma[j];
            if (sig > sigij) sigij = sig;
        }
    for (int i = 40; i <= 50; ++i) {
        sig_max[i] = 0.5 * sigij * 1.0;
        if (sig_max[i] > c.sig_max_all) c.sig_max_all = sig_max[i];
    }
    for (int i = 0; i <= 50; ++i) {
        c.rlsq[i] = 0.0;
        if (i >= 1 && sig_max[i] != 0.0) {
            const double v = (kRlConst - 1) * c.sig_max_all + sig_max[i];
            c.rlsq[i] = v * v;
        }
    }
    const double rl = kRlConst * c.sig_max_all;
    const double h = 0.4 * (rl - c.sig_max_all);
    c.hdelr = h * h;
}

// ---------------------------------------------------------------------------
// Seeded generator
// ---------------------------------------------------------------------------
namespace {
struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 unit(V3 a) { return (1.0 / norm(a)) * a; }

struct Rng {
    uint64_t key, ctr = 0;
    explicit Rng(uint64_t seed) : key(splitmix64(seed ^ 0xA0761D6478BD642Full)) {}
    double u() { return (double)(splitmix64(key + ctr++) >> 11) * (1.0 / 9007199254740992.0); }
    double gauss() {
        double u1 = u(), u2 = u();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// place D with |CD| = l, angle BCD = theta, dihedral ABCD = phi
V3 nerf(V3 A, V3 B, V3 C, double l, double theta, double phi) {
    V3 bc = unit(C - B);
    V3 n = unit(cross(B - A, bc));
    V3 m = cross(n, bc);
    const double dx = -l * std::cos(theta), dy = l * std::sin(theta) * std::cos(phi),
                 dz = l * std::sin(theta) * std::sin(phi);
    return C + dx * bc + dy * m + dz * n;
}
double tri_angle(double a, double l, double d) {  // angle between sides a and l opposite d
    double c = (a * a + l * l - d * d) / (2 * a * l);
    c = std::max(-1.0, std::min(1.0, c));
    return std::acos(c);
}
bool trilat(V3 p1, V3 p2, V3 p3, double r1, double r2, double r3, int sign, V3& out) {
    V3 ex = unit(p2 - p1);
    const double i = dot(ex, p3 - p1);
    V3 ey = unit(p3 - p1 - i * ex);
    V3 ez = cross(ex, ey);
    const double d = norm(p2 - p1), j = dot(ey, p3 - p1);
    const double x = (r1 * r1 - r2 * r2 + d * d) / (2 * d);
    const double y = (r1 * r1 - r3 * r3 + i * i + j * j) / (2 * j) - (i / j) * x;
    double z2 = r1 * r1 - x * x - y * y;
    // slightly inconsistent distance triples (glycine) fall back to the
    // in-plane point; the bond-window test then decides
    const double z = z2 > 0 ? std::sqrt(z2) : 0.0;
    out = p1 + x * ex + y * ey + (sign * z) * ez;
    return true;
}

struct Grid {
    double cell, boxl;
    int m;
    std::unordered_map<long long, std::vector<int>> cells;
    long long key(int x, int y, int z) const { return ((long long)x * m + y) * m + z; }
    void coords(V3 p, int& x, int& y, int& z) const {
        auto w = [&](double v) {
            double f = v - boxl * std::floor(v / boxl);
            int c = (int)(f / cell);
            return c >= m ? m - 1 : c;
        };
        x = w(p.x);
        y = w(p.y);
        z = w(p.z);
    }
    void add(int id, V3 p) {
        int x, y, z;
        coords(p, x, y, z);
        cells[key(x, y, z)].push_back(id);
    }
    void remove(int id, V3 p) {
        int x, y, z;
        coords(p, x, y, z);
        auto& v = cells[key(x, y, z)];
        v.erase(std::remove(v.begin(), v.end(), id), v.end());
    }
};
}  // namespace

bool generate(const Params& raw, const GenSpec& spec, GenOut& out, std::string& err) {
    Rng rng(spec.seed);
    const int nc = (int)spec.chain_len.size();
    int N = 0, nres = 0;
    for (int L : spec.chain_len) {
        if (L < 4) { err = "chains need at least 4 residues"; return false; }
        N += 4 * L;
        nres += L;
    }
    out.restype.assign(nres, 9);
    out.hp.assign(nres, 1);  // hp1.inp: every side chain hydrophobic-capable, as shipped
    if (!spec.restype.empty()) {
        if ((int)spec.restype.size() != nres) { err = "given sequence length differs from the chain lengths"; return false; }
        for (int q = 0; q < nres; ++q) {
            if (spec.restype[q] < 9 || spec.restype[q] > 28) { err = "side-chain identities must lie in 9..28"; return false; }
            out.restype[q] = spec.restype[q];
        }
    } else {
        for (int q = 0; q < nres; ++q) out.restype[q] = 9 + (int)(rng.u() * 20.0) % 20;
    }
    const double sig1 = raw.sigma[1];
    out.boxl = std::cbrt((double)N * sig1 * sig1 * sig1 / spec.density);
    const double B = out.boxl;

    Params p = raw;
    scale_params(p, 1.0);  // Angstrom units (boxl_orig = 1)
    Topology t;
    std::vector<int> species(nc);
    for (int c = 0; c < nc; ++c) species[c] = c;
    t.build(spec.chain_len, species, out.restype, out.hp);

    std::vector<V3> pos(N);
    std::vector<char> placed(N, 0);
    const double gcell = 9.0;
    Grid grid{gcell, B, std::max(1, (int)(B / gcell)), {}};
    grid.cell = B / grid.m;

    auto min_image = [&](V3 d) {
        d.x -= B * std::round(d.x / B);
        d.y -= B * std::round(d.y / B);
        d.z -= B * std::round(d.z / B);
        return d;
    };
    // pair acceptance: bond windows strictly inside, cores with 1% margin,
    // N-C H-bond pairs outside their well so no shoulder state is needed.
    auto pair_ok = [&](int a, int b) {
        const int code = static_code(t, a, b);
        const double d = norm(min_image(pos[a] - pos[b]));
        const int ia = t.ident0[a], ib = t.ident0[b];
        if (code >= 4 && code <= 12) {
            const int lo = a < b ? a : b;
            double blmin = p.ev_param[2][code], blmax = p.ev_param[3][code];
            const int rt = out.restype[t.res[lo]] - 8;
            if (code == 10) { blmin = (1 - std::max(raw.del_rca[rt], kDel)) * raw.drca[rt]; blmax = (1 + std::max(raw.del_rca[rt], kDel)) * raw.drca[rt]; }
            if (code == 11) { blmin = (1 - std::max(raw.del_rnh[rt], kDel)) * raw.drnh[rt]; blmax = (1 + std::max(raw.del_rnh[rt], kDel)) * raw.drnh[rt]; }
            if (code == 12) { blmin = (1 - std::max(raw.del_rco[rt], kDel)) * raw.drco[rt]; blmax = (1 + std::max(raw.del_rco[rt], kDel)) * raw.drco[rt]; }
            return d > blmin * (1 + 1e-7) && d < blmax * (1 - 1e-7);
        }
        if (code == 15) return d > std::sqrt(p.welldia_sq[ia][ib]) * 1.01;
        double s2 = p.sigma_sq[ia][ib];
        if (code != 16) {
            const double pe = p.ev_param[1][code];
            s2 = s2 * (pe * pe);
            if (code >= 22 && code <= 26) {
                const double q = p.sqz610[code - 21][std::max(ia, ib)];
                s2 = s2 * (q * q);
            }
        }
        return d > std::sqrt(s2) * 1.0005;
    };
    auto bead_ok = [&](int a) {
        const int c = t.chain[a], s = t.chain_start[c], n = 4 * t.chain_len[c];
        for (int b = s; b < s + n; ++b)
            if (b != a && placed[b] && !pair_ok(a, b)) return false;
        int x, y, z;
        grid.coords(pos[a], x, y, z);
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    const int m = grid.m;
                    auto it = grid.cells.find(grid.key(((x + dx) % m + m) % m, ((y + dy) % m + m) % m, ((z + dz) % m + m) % m));
                    if (it == grid.cells.end()) continue;
                    for (int b : it->second)
                        if (t.chain[b] != c && !pair_ok(a, b)) return false;
                }
        return true;
    };
    const double th_ca = tri_angle(1.46, 1.51, 2.45);   // N-Ca-C
    const double th_c = tri_angle(1.51, 1.33, 2.41);    // Ca-C-N(+1)
    const double th_n = tri_angle(1.33, 1.46, 2.45);    // C-N-Ca
    auto torsion_pair = [&](double& phi, double& psi, bool pro) {
        const double u = rng.u();
        const double deg = 3.141592653589793 / 180.0;
        // beta, polyproline-II and alpha basins; the squeeze cores of
        // make_code.f (codes 17-26) reject whatever falls outside
        if (u < 0.55) { phi = (-120 + 40 * (rng.u() - 0.5)) * deg; psi = (130 + 40 * (rng.u() - 0.5)) * deg; }
        else if (u < 0.9) { phi = (-75 + 40 * (rng.u() - 0.5)) * deg; psi = (145 + 40 * (rng.u() - 0.5)) * deg; }
        else if (u < 0.97) { phi = (-63 + 30 * (rng.u() - 0.5)) * deg; psi = (-42 + 30 * (rng.u() - 0.5)) * deg; }
        else { phi = (360 * rng.u() - 180) * deg; psi = (360 * rng.u() - 180) * deg; }
        if (pro) phi = (-75 + 12 * (rng.u() - 0.5)) * deg;  // squeeze cores leave proline phi near -75
    };
    auto place_side = [&](int s, int L, int r, V3 N_, V3 CA, V3 C_) {
        const int rt = out.restype[t.res[s + r - 1]] - 8;
        V3 R;
        if (!trilat(CA, N_, C_, raw.drca[rt], raw.drnh[rt], raw.drco[rt], +1, R)) return false;
        pos[s + 3 * L + r - 1] = R;
        return true;
    };
    for (int c = 0; c < nc; ++c) {
        const int L = spec.chain_len[c], s = t.chain_start[c];
        auto ca = [&](int r) { return s + r - 1; };
        auto nn = [&](int r) { return s + L + r - 1; };
        auto cc = [&](int r) { return s + 2 * L + r - 1; };
        auto rr = [&](int r) { return s + 3 * L + r - 1; };
        auto unplace = [&](int r) {
            for (int b : {ca(r), nn(r), cc(r), rr(r)})
                if (placed[b]) { grid.remove(b, pos[b]); placed[b] = 0; }
        };
        auto commit = [&](int r) {
            for (int b : {ca(r), nn(r), cc(r), rr(r)}) { placed[b] = 1; grid.add(b, pos[b]); }
        };
        auto check_res = [&](int r) {
            for (int b : {nn(r), ca(r), cc(r), rr(r)}) {
                placed[b] = 1;
                if (!bead_ok(b)) {
                    for (int q : {nn(r), ca(r), cc(r), rr(r)}) placed[q] = 0;
                    return false;
                }
            }
            for (int q : {nn(r), ca(r), cc(r), rr(r)}) placed[q] = 0;
            return true;
        };
        bool chain_done = false;
        for (int attempt = 0; attempt < 400 && !chain_done; ++attempt) {
            for (int r = 1; r <= L; ++r) unplace(r);
            // residue 1
            bool ok1 = false;
            for (int tr = 0; tr < 200 && !ok1; ++tr) {
                V3 n1{rng.u() * B, rng.u() * B, rng.u() * B};
                V3 u1 = unit(V3{rng.gauss(), rng.gauss(), rng.gauss()});
                V3 a1 = n1 + 1.46 * u1;
                V3 ref = a1 + V3{rng.gauss(), rng.gauss(), rng.gauss()};
                V3 c1 = nerf(ref, n1, a1, 1.51, th_ca, 6.283185307179586 * rng.u());
                pos[nn(1)] = n1;
                pos[ca(1)] = a1;
                pos[cc(1)] = c1;
                if (!place_side(s, L, 1, n1, a1, c1)) continue;
                if (check_res(1)) { commit(1); ok1 = true; }
            }
            if (!ok1) continue;
            int r = 2, backtracks = 0;
            bool fail = false;
            while (r <= L) {
                bool ok = false;
                for (int tr = 0; tr < 60 && !ok; ++tr) {
                    double phi, psi;
                    torsion_pair(phi, psi, out.restype[t.res[s + r - 1]] == 17);
                    V3 Np = pos[nn(r - 1)], Ap = pos[ca(r - 1)], Cp = pos[cc(r - 1)];
                    V3 Nr = nerf(Np, Ap, Cp, 1.33, th_c, psi);
                    // omega chosen so |Ca(r-1) - Ca(r)| = 3.8 when reachable
                    V3 a0 = nerf(Ap, Cp, Nr, 1.46, th_n, 0.0), a180 = nerf(Ap, Cp, Nr, 1.46, th_n, 3.141592653589793);
                    const double d0 = dot(a0 - Ap, a0 - Ap), d180 = dot(a180 - Ap, a180 - Ap);
                    const double P = 0.5 * (d0 + d180), Q = 0.5 * (d0 - d180);
                    double cw = (3.8 * 3.8 - P) / Q;
                    double omega = cw < -1 ? 3.141592653589793 : (cw > 1 ? 0.0 : std::acos(cw));
                    V3 Ar = nerf(Ap, Cp, Nr, 1.46, th_n, omega);
                    V3 Cr = nerf(Cp, Nr, Ar, 1.51, th_ca, phi);
                    pos[nn(r)] = Nr;
                    pos[ca(r)] = Ar;
                    pos[cc(r)] = Cr;
                    if (!place_side(s, L, r, Nr, Ar, Cr)) continue;
                    if (check_res(r)) ok = true;
                }
                if (ok) {
                    commit(r);
                    ++r;
                } else {
                    if (++backtracks > 40) { fail = true; break; }
                    const int back = std::min(r - 2, 2);
                    for (int q = 0; q < back; ++q) unplace(--r);
                }
            }
            if (!fail) chain_done = true;
        }
        if (!chain_done) { err = "generator could not place chain"; return false; }
    }
    // Maxwell-Boltzmann velocities at setemp = 12 T*, zero momentum, exact T
    const double setemp = spec.tstar * 12.0;
    std::vector<double> bm(N);
    for (int b = 0; b < N; ++b) bm[b] = raw.bmass[t.ident0[b]];
    std::vector<V3> v(N);
    V3 P{0, 0, 0};
    double M = 0;
    for (int b = 0; b < N; ++b) {
        const double sd = std::sqrt(setemp / bm[b]);
        v[b] = {sd * rng.gauss(), sd * rng.gauss(), sd * rng.gauss()};
        P = P + bm[b] * v[b];
        M += bm[b];
    }
    double ke = 0;
    for (int b = 0; b < N; ++b) {
        v[b] = v[b] - (1.0 / M) * P;
        ke += bm[b] * dot(v[b], v[b]);
    }
    const double scale = std::sqrt(setemp * 3.0 * N / ke);
    out.sv.assign(6 * (size_t)N, 0.0);
    for (int b = 0; b < N; ++b) {
        V3 x = (1.0 / B) * pos[b];
        double xs[3] = {x.x - 0.5, x.y - 0.5, x.z - 0.5};
        for (int k = 0; k < 3; ++k) xs[k] = xs[k] - std::round(xs[k]);
        out.sv[6 * b + 0] = xs[0];
        out.sv[6 * b + 1] = xs[1];
        out.sv[6 * b + 2] = xs[2];
        out.sv[6 * b + 3] = scale * v[b].x;
        out.sv[6 * b + 4] = scale * v[b].y;
        out.sv[6 * b + 5] = scale * v[b].z;
    }
    return true;
}

}  // namespace pdmd
