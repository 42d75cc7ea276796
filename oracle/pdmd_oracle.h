/*
 * pdmd_oracle.h -- CPU oracle for the parallel PRIME20 DMD event loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a serial C restatement of the
 * reference's DMD semantics (parallel-dmd-PRIME20/code/*.f, main.F90) used
 * as the checker for the HIP engine and as the CPU baseline in bench.py.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product never links, loads or calls anything here.
 *
 * Indices across this API are 1-based, like the reference (bead 1..N,
 * pseudo-events N+1 ghost, N+2 sync, N+3 output).
 */
#ifndef PDMD_ORACLE_H
#define PDMD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;

typedef struct {
    int64_t coll;   /* reference coll counter after this event (1-based) */
    int32_t i;      /* event owner: bead 1..N or pseudo N+1..N+3 */
    int32_t j;      /* partner (pair event) / ghost bead (N+1) / 0 */
    int32_t type;   /* resolved coltype (pair) ; 0 for pseudo events */
    int32_t pad;
    double  t;      /* tfalse at the event (time since last sync) */
} orc_event;

typedef struct {
    double ered, tred, sumvel;
    int32_t hb_alpha, hb_ii, hb_ij, pad;
    double ehh_ii, ehh_ij, rg_avg, e2e_avg;
} orc_obs;

typedef struct {
    int64_t coll, numghosts, nupdates, nforced;
    int64_t nevents[31];
    double t, tfalse, interval, interval_max, t_fact, n_forced, sortsize, tbin_off;
    int32_t nbin, num_cell;
    double sig_max_all, hdelr, avg_nbrs;
    int64_t rng_ctr[2];
} orc_stats;

/* Build a system.  chain_len[c] residues per chain, species[c] groups chains
 * sharing one sequence (the reference's nop1/nop2 species), restype[] is the
 * side-chain identity (9..28) of every residue (chains concatenated), hp[]
 * the side-chain hydrophobic flag per residue.  sv is the reference's
 * sv(6,N) layout (x,y,z,vx,vy,vz per bead, positions in box units) and
 * bptnr the 1-based H-bond partner table (0 = none), as restored by -Drunr.
 * canon != 0 enables Andersen ghost collisions.  Returns NULL on error. */
orc_sys* orc_create(const char* param_dir, int nchains, const int32_t* chain_len,
                    const int32_t* species, const int32_t* restype, const int32_t* hp,
                    double boxl_angstrom, double tstar, uint64_t seed, int canon,
                    const double* sv, const int32_t* bptnr, char* err, int errlen);
/* The same with build flags: ORC_NO_HBS restates -Dno_hbs (no N-C
 * hydrogen-bond wells, make_code.f:114-170,192-208,348-365). */
#define ORC_NO_HBS 1
/* ORC_NO_H restates -Dno_h: hp zeroed for the side chains (inputinfo.f:270-280),
 * which removes the code-16 wells between chains (make_code.f:80-104); the
 * same-chain ones read hp1/hp2 and stay. */
#define ORC_NO_H 2
/* ORC_CHAPTYPE2 restates -Dchaptype=2: the chains after the first species
 * block (the reference's nop2 block) get no intra-chain codes (no bonds,
 * local cores, H-bond or side-chain wells, make_code.f:349-484) and no
 * side-chain wells with each other (make_code.f:519-547); checkover.f skips
 * its bond-window checks (checkover.f:67). */
#define ORC_CHAPTYPE2 4
orc_sys* orc_create_ex(const char* param_dir, int nchains, const int32_t* chain_len,
                       const int32_t* species, const int32_t* restype, const int32_t* hp,
                       double boxl_angstrom, double tstar, uint64_t seed, int canon, int flags,
                       const double* sv, const int32_t* bptnr, char* err, int errlen);
void orc_destroy(orc_sys* s);

int32_t orc_nbeads(const orc_sys* s);
/* Process exactly nev events.  Logs up to log_cap of them.  Returns the
 * number processed, or a negative status on a self-check failure. */
int64_t orc_run(orc_sys* s, int64_t nev, orc_event* log, int64_t log_cap);
/* A list update forced at the current time (the sync pseudo-event's origin
 * shift and update branch, main.F90:1126-1187, without its displacement
 * test, rate adaptation or new sync entry). */
void orc_rebuild(orc_sys* s);

void orc_get_state(const orc_sys* s, double* sv, int32_t* identity, int32_t* bptnr,
                   int32_t* extra_repuls /* N*4 */, double* tim /* N+3 */,
                   int32_t* nptnr /* N+3 */, int32_t* coltype /* N+3 */);
void orc_stats_get(const orc_sys* s, orc_stats* st);
void orc_observables(const orc_sys* s, orc_obs* o);
int32_t orc_ev_code(const orc_sys* s, int32_t a, int32_t b);
/* Neighbour lists (up: partner > i, down: partner < i) in the reference's
 * nbor order.  Returns count; copies up to cap entries. */
int32_t orc_nb_up(const orc_sys* s, int32_t i, int32_t* out, int32_t cap);
int32_t orc_nb_dn(const orc_sys* s, int32_t i, int32_t* out, int32_t cap);
int32_t orc_cell_of(const orc_sys* s, int32_t i);
/* Predict pair (i,j) exactly as events()/eventredo_up would (code taken
 * from ev_code(j,i)); returns tij (1e9 when no event) and *type. */
double orc_predict(const orc_sys* s, int32_t i, int32_t j, int32_t* type);
/* Overlap check restating checkover.f; returns number of violations. */
int32_t orc_checkover(const orc_sys* s);
/* checkover.f split by kind plus check_nc_int.f (see pdmd_oracle.c) */
void orc_check(orc_sys* s, const double* sv, int64_t out[9]);
/* Table access for tests: which = 0 sigma_sq, 1 welldia_sq, 2 ep_sqrt,
 * 3 shlddia_sq, 4 sigma_2b, 5 bmass(k), 6 ev_param(k=row,kk=col),
 * 7 sqz610(k,kk), 8 rlsq(k). */
double orc_table(const orc_sys* s, int which, int k, int kk);
/* Per-residue bond windows (box units): which 0..5 = blmin10, blmax10,
 * blmin11, blmax11, blmin12, blmax12. */
double orc_bond(const orc_sys* s, int which, int residue);

/* Project RNG and logarithm (defined identically in the product). */
double orc_rng_u01(uint64_t seed, int stream, uint64_t ctr);
double orc_log(double x);
/* Disable the debugging clamp of bin indices (script default keeps it). */
void orc_set_max_log_obs(orc_sys* s, int enable);

#ifdef __cplusplus
}
#endif
#endif
