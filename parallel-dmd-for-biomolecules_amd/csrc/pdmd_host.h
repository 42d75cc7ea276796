// pdmd_host.h -- host-side system set-up for the MI355X PRIME20 DMD engine:
// parameter tables (reference file formats), unit scaling, the compact
// interaction topology that replaces the dense ev_code matrix, the
// neighbour-cutoff set-up and the seeded input generator.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pdmd {

constexpr double kLtstep = 1e-10;   // code/def.h:1
constexpr double kSmdist = 5e-12;   // code/def.h:2
constexpr double kDel = 0.02375;    // code/def.h:38
constexpr double kRlConst = 1.2;    // code/def.h:39
constexpr int kNWrap = 2;           // qfile/script.sh:7 -Dn_wrap=2
constexpr int kNumbin = 2000;       // qfile/script.sh:7 -Dnumbin=2000
constexpr int kXrepuls1 = 40;       // code/def.h:54
constexpr int kXrepuls2 = 50;       // code/def.h:55

// Tables of code/header.f:42-50 after inputinfo.f + scale_down.f + make_code.f:18-69.
struct Params {
    double sigma[29], welldia[29], epsilon[29];
    double sigma_sq[29][29], sigma_2b[29][29], welldia_sq[29][29], ep_sqrt[29][29], shlddia_sq[29][29];
    double ep[29][29], bds[29][29], wel[29][29];
    double ev_param[4][51];
    double sqz610[6][29];
    double bmass[29];
    double drca[21], drnh[21], drco[21], del_rca[21], del_rnh[21], del_rco[21];
    double shder_dist[5];
    double boxl_orig;
};

// Reads the reference's parameter files (inputinfo.f:141-398 formats).
bool load_params(Params& p, const std::string& dir, std::string& err);
// scale_down.f:16-83 plus make_code.f:18-69 (ev_param) for a box of boxl Angstrom.
void scale_params(Params& p, double boxl);

// One system's chain topology.  Beads of chain c are laid out as the
// reference does per chain: Calpha_1..L, N_1..L, C_1..L, R_1..L.
struct Topology {
    int N = 0, nres = 0;
    std::vector<int> chain_start, chain_len, species;    // per chain
    std::vector<int> restype, hp;                        // per residue (side-chain identity 9..28, hp flag)
    std::vector<int> chain, loc, res, ident0;            // per bead (0-based chain/res; loc 1-based)
    bool no_hbs = false;                                 // -Dno_hbs: no code-15 N-C wells
    bool no_h = false;                                   // -Dno_h: no inter-chain code-16 wells
    bool chaptype2 = false;                              // -Dchaptype=2: chains after the first species block unbonded
    int nfirst = 0;                                      // chains of the first species block
    bool free_bead(int b) const { return chaptype2 && chain[b] >= nfirst; }
    void build(const std::vector<int>& chain_len, const std::vector<int>& species,
               const std::vector<int>& restype, const std::vector<int>& hp);
};

// Static pair code ev_code(a,b) as left by make_code.f (symmetric, a != b,
// 0-based bead indices).  Closed form of make_code.f:72-563.
int static_code(const Topology& t, int a, int b);

// cell_link.f:13-94 (n_wrap = 2): the place 1..62 of each neighbour-cell
// offset in the order nbor.f scans them, at (dx+2) + 5 (dy+2) + 25 (dz+2);
// 0 for the offsets whose opposite is scanned instead, and for (0, 0, 0).
void half_stencil(signed char out[125]);

// nbor_setup.f:13-118: neighbour cutoffs per code, sig_max_all and hdelr.
// Dynamic overrides present at set-up are ignored there as in the reference
// (ev_code <= 26 filter), so only static codes matter.
struct Cutoffs { double rlsq[51]; double sig_max_all; double hdelr; };
void nbor_setup(const Params& p, const Topology& t, const std::vector<int>& ident, Cutoffs& c);

// Seeded synthetic input generator: random sequences, self-avoiding coils
// satisfying every bond window, squeeze core and hard core, Maxwell-Boltzmann
// velocities at T* with zero momentum.  Density rho* = N sigma_1^3 / V.
struct GenSpec {
    std::vector<int> chain_len;
    double density = 0.01;
    double tstar = 0.2;
    uint64_t seed = 1;
    std::vector<int> restype;  // optional given sequence (9..28 per residue); empty = drawn from the seed
};
struct GenOut {
    double boxl = 0;
    std::vector<int> restype, hp;   // per residue
    std::vector<double> sv;         // 6*N, reference sv(6,N) layout, box units
};
bool generate(const Params& raw, const GenSpec& spec, GenOut& out, std::string& err);

uint64_t splitmix64(uint64_t x);

// Sets the text pdmd_last_error() returns on this thread; returns -1.
int report_error(const char* msg);

}  // namespace pdmd
