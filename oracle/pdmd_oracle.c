/*
 * pdmd_oracle.c -- serial CPU restatement of the reference's PRIME20 DMD
 * event loop.  TEST INFRASTRUCTURE: the checker for the HIP engine and the
 * CPU baseline ("port") timed by bench.py.  Nothing in the product uses it.
 *
 * Every routine names the reference file:line it follows (paths relative to
 * parallel-dmd-PRIME20/code/).  The serial semantics are those of main.F90
 * with the worker branch (main.F90:1432-1960) executed on current state, as
 * a one-worker run would, and a project counter-based RNG in place of the
 * absent IFPORT drandm.  Generalisation: every chain may carry its own
 * sequence; chains with equal `species` id form one of the reference's
 * species (nop1/nop2 blocks).  Every residue carries a side-chain bead
 * (firstside*.data all 1, as shipped).
 *
 * Compile with -ffp-contract=off: floating-point expressions below keep the
 * reference's evaluation order (Fortran left-to-right for equal precedence).
 *
 * Parity status per function:
 *   pinned bit for bit against the reference's own subroutines (compiled by
 *   oracle/ref from the unmodified sources; fixtures in tests/golden/):
 *     load_params + scale_down (every table), make_code (ev_code, also with
 *     -Dno_hbs), nbor_setup/cell_link/nbor (lists, in the reference's
 *     order), events (calendar), core/bond/sqwel/sqshlder/nc_sqwel (every
 *     pair), eventdyn for coltypes 1-3 and side-chain wells 4/8,
 *     partial_events/eventredo_up/eventredo_down/add_tbin/del_tbin, displ
 *     and the sync and output pseudo-events over 1.8e5-2e5 calendar entries,
 *     energy/radgyr/end_to_end;
 *   pinned by invariants only (NVE energy conservation to 1e-10, calendar
 *   equal to a brute-force minimum, list membership by brute force):
 *     the worker's resolution of N-C well events (coltypes 7/10/12 ->
 *     4/5/6/9/13/14/15/16; repuls_check, repuls_check_3, check_sigma,
 *     repuls_add, repuls_del_a/b, bumpoff, H-bond bookkeeping) -- parity
 *     against the reference unpinned for these;
 *   ghost and terminal-capture random draws: project-defined RNG in place
 *     of IFPORT drandm (absent), pinned by the splitmix64 KAT; parity
 *     against the reference unpinned by construction.
 */
#include "pdmd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* code/def.h:1-57 constants */
#define LTSTEP 1e-10
#define SMDIST 5e-12
#define DNC 1.46
#define DCC 1.51
#define DCN 1.33
#define DCACA 3.8
#define DTIE 2.41
#define DTIE2 2.45
#define SQZ1 1.1436
#define SQZ2 0.88
#define SQZ3 0.87829
#define SQZ4 0.8
#define SQZ5 0.7713
#define SQZ11 1.0
#define N_B_HYDRO 3
#define N_B_HBOND 3
#define DEL 0.02375
#define RL_CONST 1.2
#define N_WRAP 2
#define N_NAB_CELL 62
#define XREPULS1 40
#define XREPULS2 50
#define NUMBIN 2000 /* qfile/script.sh:7 -Dnumbin=2000 */
#define MAXNBS 512  /* reference uses 2000 slots (def.h:57); overflow is an error here */

typedef struct {
    double sigma[29], welldia[29], epsilon[29];
    double sigma_sq[29][29], sigma_2b[29][29], welldia_sq[29][29], ep_sqrt[29][29],
        shlddia_sq[29][29];
    double ep[29][29], bds[29][29], wel[29][29];
    double ev_param[4][51];
    double sqz610[6][29];
    double bmass[29];
    double drca[21], drnh[21], drco[21], del_rca[21], del_rnh[21], del_rco[21];
    double shder_dist1, shder_dist2, shder_dist3, shder_dist4;
    double boxl_orig;
} params_t;

struct orc_sys {
    params_t P;
    int N, nchains;
    int *chain_start, *chain_len, *species; /* per chain (0-based chain id) */
    /* per bead, 1-based */
    int *chnnum, *loc, *cl, *res; /* chain (1-based), local index, chain length, residue (1-based global) */
    int *hp_bead;
    double *sv; /* sv[(k-1)*6 + c] */
    double *bm, *old_r;
    int *identity, *bptnr, *extra; /* extra[(k-1)*4 + (m-1)] */
    signed char* ev; /* dense ev_code, ev[(a-1)*N + (b-1)] */
    /* per residue bond data (box units) */
    double *bdln, *bl_rn, *bl_rc, *del_bdln, *del_blrn, *del_blrc;
    /* neighbour lists */
    int *nb, *na_npt, *dnnab, *nnabdn;
    /* cells */
    int num_cell, *cell, *wrap_map, *clinks, map[N_NAB_CELL + 1];
    double width, half, boxl;
    /* calendar and times, 1-based N+3 */
    double* tim;
    int *nptnr, *coltype, *tlinks, *tlinks2, bin[NUMBIN + 2], nbin;
    double tfalse, t, interval, interval_max, t_fact, n_forced, sortsize, tbin_off;
    double setemp, avegtime, sig_max_all, hdelr, rlsq[51];
    int canon;
    int no_hbs; /* -Dno_hbs: no N-C hydrogen-bond wells (make_code.f:114-170,192-208,348-365) */
    int no_h;   /* -Dno_h: hp zeroed where make_code.f:80-104 reads it (inputinfo.f:270-280) */
    int chaptype2; /* -Dchaptype=2 for the chains after the first species block */
    int nfirst;    /* chains of the first species block (the reference's nop1 block) */
    int64_t coll, numghosts, nupdates, nforced, nevents[31];
    uint64_t seed, rng_ctr[2];
    double old_tfalse;
    int err;
    double avg_nbrs;
    int log_obs;
};

#define EV(S, a, b) ((S)->ev[(size_t)((a)-1) * (size_t)(S)->N + (size_t)((b)-1)])
#define SV(S, c, k) ((S)->sv[(size_t)((k)-1) * 6 + ((c)-1)])
#define XR(S, k, m) ((S)->extra[(size_t)((k)-1) * 4 + ((m)-1)])

/* ------------------------------------------------------------------ */
/* Project RNG: counter-based splitmix64.  Stream 0 = ghost events,    */
/* stream 1 = terminal-bead capture draws.                             */
/* ------------------------------------------------------------------ */
static uint64_t sm64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
double orc_rng_u01(uint64_t seed, int stream, uint64_t ctr) {
    uint64_t key = sm64(seed ^ (0x5851F42D4C957F2Dull * (uint64_t)(stream + 1)));
    return (double)(sm64(key + ctr) >> 11) * (1.0 / 9007199254740992.0);
}
static double drand(orc_sys* S, int stream) {
    return orc_rng_u01(S->seed, stream, S->rng_ctr[stream]++);
}

/* Natural logarithm from basic IEEE operations only (atanh series), so the
 * CPU and GPU produce identical bits: x = m 2^e, m in [1/sqrt2, sqrt2],
 * log m = 2 atanh(s), s = (m-1)/(m+1). */
double orc_log(double x) {
    uint64_t b;
    memcpy(&b, &x, 8);
    int e = (int)((b >> 52) & 0x7ff);
    int adj = 0;
    if (e == 0) { /* subnormal */
        x = x * 18014398509481984.0; /* 2^54 */
        memcpy(&b, &x, 8);
        e = (int)((b >> 52) & 0x7ff);
        adj = -54;
    }
    e = e - 1023 + adj;
    b = (b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m;
    memcpy(&m, &b, 8);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
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
    double lm = 2.0 * s + (s * z) * p;
    double de = (double)e;
    return de * 6.93147180369123816490e-01 + (de * 1.90821492927058770002e-10 + lm);
}

static double dnint(double x) { return round(x); }

/* ------------------------------------------------------------------ */
/* Parameter files: inputinfo.f:105-411 read formats, scale_down.f     */
/* ------------------------------------------------------------------ */
static double fld(const char* line, int c0, int w) {
    /* Fortran fixed-width numeric field starting at column c0 (1-based). */
    char buf[64];
    int n = (int)strlen(line), k = 0;
    for (int c = c0 - 1; c < c0 - 1 + w && c < n; ++c)
        if (line[c] != ' ' && line[c] != '\n' && line[c] != '\r') buf[k++] = line[c];
    buf[k] = 0;
    if (k == 0) return 0.0;
    return strtod(buf, NULL);
}
static FILE* popen_param(const char* dir, const char* name) {
    char path[4096];
    snprintf(path, sizeof path, "%s/%s", dir, name);
    return fopen(path, "r");
}

static int load_params(params_t* P, const char* dir, char* err, int errlen) {
    char line[512];
    FILE* f;
    memset(P, 0, sizeof *P);
    /* protein.data: 12 list-directed reals (inputinfo.f:141-176) */
    if (!(f = popen_param(dir, "protein.data"))) goto fail;
    double pv[12];
    for (int k = 0; k < 12; ++k)
        if (fscanf(f, "%lf", &pv[k]) != 1) { fclose(f); goto fail; }
    fclose(f);
    for (int k = 0; k < 4; ++k) {
        P->sigma[k + 1] = pv[k];
        P->welldia[k + 1] = pv[4 + k];
        P->epsilon[k + 1] = pv[8 + k];
    }
    for (int k = 1; k <= 4; ++k) { /* inputinfo.f:178-189 */
        P->sigma[k + 4] = P->sigma[k];
        P->epsilon[k + 4] = P->epsilon[k];
        P->welldia[k + 4] = P->welldia[k];
    }
    /* ep19p_ha55a_weakhp.data, format (2(2x,i2,2x),f8.3) (inputinfo.f:282-288) */
    if (!(f = popen_param(dir, "ep19p_ha55a_weakhp.data"))) goto fail;
    for (int k = 0; k < 400; ++k) {
        if (!fgets(line, sizeof line, f)) { fclose(f); goto fail; }
        int a = (int)fld(line, 3, 2), b = (int)fld(line, 9, 2);
        if (a < 1 || a > 28 || b < 1 || b > 28) { fclose(f); goto fail; }
        P->ep[a][b] = -fld(line, 13, 8);
    }
    fclose(f);
    /* rcarnrco.data, format (6(f6.3,2x)) (inputinfo.f:292-298) */
    if (!(f = popen_param(dir, "rcarnrco.data"))) goto fail;
    for (int k = 1; k <= 20; ++k) {
        if (!fgets(line, sizeof line, f)) { fclose(f); goto fail; }
        P->drca[k] = fld(line, 1, 6);
        P->drnh[k] = fld(line, 9, 6);
        P->drco[k] = fld(line, 17, 6);
        P->del_rca[k] = fld(line, 25, 6);
        P->del_rnh[k] = fld(line, 33, 6);
        P->del_rco[k] = fld(line, 41, 6);
        if (P->del_rca[k] < DEL) P->del_rca[k] = DEL;
        if (P->del_rnh[k] < DEL) P->del_rnh[k] = DEL;
        if (P->del_rco[k] < DEL) P->del_rco[k] = DEL;
    }
    fclose(f);
    /* beadwell_ha55a.data, format (2(2x,i2,2x),2(f6.3,2x)) (inputinfo.f:352-363) */
    if (!(f = popen_param(dir, "beadwell_ha55a.data"))) goto fail;
    for (int k = 0; k < 400; ++k) {
        if (!fgets(line, sizeof line, f)) { fclose(f); goto fail; }
        int a = (int)fld(line, 3, 2), b = (int)fld(line, 9, 2);
        if (a < 1 || a > 28 || b < 1 || b > 28) { fclose(f); goto fail; }
        P->bds[a][b] = fld(line, 13, 6);
        P->wel[a][b] = fld(line, 21, 6);
    }
    fclose(f);
    for (int i = 9; i <= 28; ++i) { /* inputinfo.f:365-370 */
        P->sigma[i] = 1.00 * P->bds[i][i];
        P->welldia[i] = 1.5 * P->sigma[i];
    }
    P->sigma[9] = 1.00 * P->bds[9][9] - 1.2;
    /* sqz6to10.data, format (5(f6.3,2x)) read as sz8,sz6,sz7,sz9,sz10 (inputinfo.f:371-382) */
    if (!(f = popen_param(dir, "sqz6to10.data"))) goto fail;
    for (int i = 1; i <= 20; ++i) {
        if (!fgets(line, sizeof line, f)) { fclose(f); goto fail; }
        double sz8 = fld(line, 1, 6), sz6 = fld(line, 9, 6), sz7 = fld(line, 17, 6),
               sz9 = fld(line, 25, 6), sz10 = fld(line, 33, 6);
        P->sqz610[1][i + 8] = sz6 * 2 / (P->sigma[i + 8] + P->sigma[4]);
        P->sqz610[2][i + 8] = sz7 * 2 / (P->sigma[i + 8] + P->sigma[1]);
        P->sqz610[3][i + 8] = sz8 * 2 / (P->sigma[i + 8] + P->sigma[2]);
        P->sqz610[4][i + 8] = sz9 * 2 / (P->sigma[i + 8] + P->sigma[2]);
        P->sqz610[5][i + 8] = sz10 * 2 / (P->sigma[i + 8] + P->sigma[4]);
    }
    fclose(f);
    /* mass.data, format (4x,i2,2x,f8.3) (inputinfo.f:385-398) */
    if (!(f = popen_param(dir, "mass.data"))) goto fail;
    for (int i = 0; i < 23; ++i) {
        if (!fgets(line, sizeof line, f)) { fclose(f); goto fail; }
        int a = (int)fld(line, 5, 2);
        if (a < 1 || a > 28) { fclose(f); goto fail; }
        P->bmass[a] = fld(line, 9, 8);
    }
    fclose(f);
    P->bmass[3] = P->bmass[20];
    for (int i = 1; i <= 4; ++i) P->bmass[i + 4] = P->bmass[i];
    return 0;
fail:
    snprintf(err, errlen, "cannot read parameter tables in %s", dir);
    return -1;
}

/* scale_down.f:16-83 (runr build: positions already in box units) */
static void scale_down(params_t* P, double boxl) {
    double B = boxl;
    P->boxl_orig = B;
    for (int k = 1; k <= 28; ++k) {
        P->sigma[k] = P->sigma[k] / B;
        P->welldia[k] = P->welldia[k] / B;
    }
    P->shder_dist1 = 5.00 / B;
    P->shder_dist2 = 4.74 / B;
    P->shder_dist3 = 4.86 / B;
    P->shder_dist4 = 4.83 / B;
    for (int k = 1; k <= 28; ++k)
        for (int kk = 1; kk <= 28; ++kk) {
            double s = P->sigma[k] + P->sigma[kk], w = P->welldia[k] + P->welldia[kk];
            P->sigma_sq[k][kk] = 0.25 * (s * s);
            P->sigma_2b[k][kk] = 0.50 * s;
            P->welldia_sq[k][kk] = 0.25 * (w * w);
            P->ep_sqrt[k][kk] = sqrt(P->epsilon[k] * P->epsilon[kk]);
            P->shlddia_sq[k][kk] = P->shder_dist4 * P->shder_dist4;
        }
    for (int k = 9; k <= 28; ++k)
        for (int kk = 9; kk <= 28; ++kk) {
            P->ep_sqrt[k][kk] = P->epsilon[1] * P->ep[k][kk];
            P->sigma_sq[k][kk] = P->bds[k][kk] * P->bds[k][kk] / (B * B);
            P->sigma_2b[k][kk] = P->bds[k][kk] / B;
            P->welldia_sq[k][kk] = P->wel[k][kk] * P->wel[k][kk] / (B * B);
        }
    double d1 = P->shder_dist1 * P->shder_dist1, d2 = P->shder_dist2 * P->shder_dist2,
           d3 = P->shder_dist3 * P->shder_dist3;
    P->shlddia_sq[1][2] = P->shlddia_sq[2][1] = P->shlddia_sq[5][2] = P->shlddia_sq[2][5] = d1;
    P->shlddia_sq[1][1] = P->shlddia_sq[5][5] = P->shlddia_sq[1][5] = P->shlddia_sq[5][1] = d2;
    P->shlddia_sq[2][4] = P->shlddia_sq[4][2] = P->shlddia_sq[2][8] = P->shlddia_sq[8][2] = d3;
}

/* make_code.f:18-69: ev_param */
static void make_ev_param(params_t* P) {
    for (int k = 1; k <= 3; ++k)
        for (int l = 1; l <= 50; ++l) P->ev_param[k][l] = 1;
    double B = P->boxl_orig;
    P->ev_param[1][15] = 1.05 * ((2.24 / B) / ((P->sigma[1] + P->sigma[4]) / 2.0));
    P->ev_param[1][17] = SQZ1;
    P->ev_param[1][18] = SQZ2;
    P->ev_param[1][19] = SQZ3;
    P->ev_param[1][20] = SQZ4;
    P->ev_param[1][21] = SQZ5;
    P->ev_param[1][27] = SQZ11;
    P->ev_param[2][4] = DNC * (1.0 - DEL) / B;
    P->ev_param[2][5] = DCC * (1.0 - DEL) / B;
    P->ev_param[2][6] = DCN * (1.0 - DEL) / B;
    P->ev_param[2][7] = DTIE * (1.0 - DEL) / B;
    P->ev_param[2][8] = DTIE2 * (1.0 - DEL) / B;
    P->ev_param[2][9] = DCACA * (1.0 - DEL) / B;
    P->ev_param[2][10] = (1.0 - DEL);
    P->ev_param[2][11] = (1.0 - DEL);
    P->ev_param[2][12] = (1.0 - DEL);
    P->ev_param[3][4] = DNC * (1.0 + DEL) / B;
    P->ev_param[3][5] = DCC * (1.0 + DEL) / B;
    P->ev_param[3][6] = DCN * (1.0 + DEL) / B;
    P->ev_param[3][7] = (1.0 + DEL) * DTIE / B;
    P->ev_param[3][8] = (1.0 + DEL) * DTIE2 / B;
    P->ev_param[3][9] = (1.0 + DEL) * DCACA / B;
    P->ev_param[3][10] = (1.0 + DEL);
    P->ev_param[3][11] = (1.0 + DEL);
    P->ev_param[3][12] = (1.0 + DEL);
}

/* make_code.f:72-566, generalised from two species blocks to any chains */
static void make_code(orc_sys* S) {
    int N = S->N;
    memset(S->ev, 1, (size_t)N * (size_t)N);
    int* id = S->identity;
    /* all r's: 16 for hp side chains on different chains (:74-104) */
    for (int m = 1; m <= N - 1; ++m)
        for (int l = m + 1; l <= N; ++l)
            if (!S->no_h && id[m] > 8 && id[l] > 8 && S->hp_bead[m] == 1 && S->hp_bead[l] == 1 &&
                S->chnnum[l] != S->chnnum[m])
                EV(S, m, l) = 16;
    if (!S->no_hbs) {
        /* all nc's: 15 (:108-117) */
        for (int m = 1; m <= N - 1; ++m)
            for (int l = m + 1; l <= N; ++l)
                if ((id[m] == 1 && id[l] == 4) || (id[m] == 4 && id[l] == 1)) EV(S, m, l) = 15;
        /* same-chain reset (:120-136) */
        for (int c = 0; c < S->nchains; ++c) {
            int kk = S->chain_start[c], nbc = 4 * S->chain_len[c];
            for (int m = kk + 1; m <= kk + nbc - 1; ++m)
                for (int l = m + 1; l <= kk + nbc; ++l) EV(S, m, l) = 1;
        }
        /* proline NH rule (:139-161) */
        for (int m = 1; m <= N - 1; ++m)
            for (int l = m + 1; l <= N; ++l) {
                if (id[m] == 17 && id[l] == 4)
                    EV(S, m - 2 * S->cl[m], l) = 1;
                else if (id[m] == 4 && id[l] == 17)
                    EV(S, l - 2 * S->cl[l], m) = 1;
            }
    }
    /* per chain block (:167-329 species 1, :331-484 species 2; with
     * -Dchaptype=2 the second block's side chains have no firstside2.data
     * read (inputinfo.f:120-132) and its H-bond and bond codes are skipped
     * (:349,366), so it sets nothing) */
    for (int c = 0; c < S->nchains; ++c) {
        if (S->chaptype2 && c >= S->nfirst) continue;
        int lll = S->chain_start[c], L = S->chain_len[c];
#define FS(r) (3 * L + (r)) /* fside(r): firstside data all 1 */
        int ncount = 2 + N_B_HYDRO;
        for (int k = 1; k <= L; ++k) {
            for (int l = ncount; l <= L; ++l)
                if (S->hp_bead[lll + FS(k)] == 1 && S->hp_bead[lll + FS(l)] == 1)
                    EV(S, lll + FS(k), lll + FS(l)) = 16;
            ncount = ncount + 1;
        }
        if (!S->no_hbs) {
            for (int k = lll + L + 1; k <= lll + 2 * L; ++k)
                for (int l = lll + 2 * L + 1; l <= lll + 3 * L; ++l) EV(S, k, l) = 15;
            for (int k = lll + L + 1; k <= lll + 2 * L; ++k)
                for (int l = k + L - N_B_HBOND; l <= k + L + N_B_HBOND; ++l)
                    if (l >= 1 && l <= N) EV(S, k, l) = 1;
        }
        for (int k = lll + 1; k <= lll + L; ++k) EV(S, k, L + k) = 4;
        for (int k = lll + 1; k <= lll + L; ++k) EV(S, k, 2 * L + k) = 5;
        for (int k = lll + L + 2; k <= lll + 2 * L; ++k) EV(S, k, k + L - 1) = 6;
        for (int k = lll + 1; k <= lll + L - 1; ++k) EV(S, k, L + k + 1) = 7;
        for (int k = lll + 2; k <= lll + L; ++k) EV(S, k, 2 * L + k - 1) = 8;
        for (int k = lll + L + 1; k <= lll + 2 * L; ++k) EV(S, k, k + L) = 8;
        for (int k = lll + 1; k <= lll + L - 1; ++k) EV(S, k, k + 1) = 9;
        for (int k = lll + 1; k <= lll + L; ++k) EV(S, k, lll + FS(k - lll)) = 10;
        for (int k = lll + L + 1; k <= lll + 2 * L; ++k) EV(S, k, lll + FS(k - lll - L)) = 11;
        for (int k = lll + 2 * L + 1; k <= lll + 3 * L; ++k)
            EV(S, k, lll + FS(k - lll - 2 * L)) = 12;
        for (int k = lll + 1; k <= lll + L - 1; ++k) EV(S, k, k + 2 * L + 1) = 17;
        for (int k = lll + 2; k <= lll + L; ++k) EV(S, k, k + L - 1) = 18;
        for (int k = lll + L + 3; k <= lll + 2 * L; ++k) EV(S, k, k + L - 2) = 19;
        for (int k = lll + L + 1; k <= lll + 2 * L - 1; ++k) EV(S, k, k + 1) = 20;
        for (int k = lll + 2 * L + 1; k <= lll + 3 * L - 1; ++k) EV(S, k, k + 1) = 21;
        for (int k = lll + 2 * L + 1; k <= lll + 3 * L - 1; ++k)
            EV(S, k, lll + FS(k - lll - 2 * L + 1)) = 22;
        for (int k = lll + L + 2; k <= lll + 2 * L; ++k) EV(S, k, lll + FS(k - lll - L - 1)) = 23;
        for (int k = lll + 1; k <= lll + L - 1; ++k) EV(S, k, lll + FS(k - lll + 1)) = 24;
        for (int k = lll + 2; k <= lll + L; ++k) EV(S, k, lll + FS(k - lll - 1)) = 25;
        for (int k = lll + 2 * L + 1; k <= lll + 3 * L - 2; ++k)
            EV(S, k, lll + FS(k - lll - 2 * L + 2)) = 26;
#undef FS
    }
    if (S->chaptype2 && S->nfirst < S->nchains) {
        /* :489-517: side-chain wells between chains again, over all beads */
        for (int m = 1; m <= N - 1; ++m)
            for (int l = m + 1; l <= N; ++l)
                if (!S->no_h && id[m] > 8 && id[l] > 8 && S->hp_bead[m] == 1 && S->hp_bead[l] == 1 &&
                    S->chnnum[l] != S->chnnum[m])
                    EV(S, m, l) = 16;
        /* :519-547: none between two chains of the second block */
        for (int m = S->chain_start[S->nfirst] + 1; m <= N - 1; ++m)
            for (int l = m + 1; l <= N; ++l)
                if (!S->no_h && id[m] > 8 && id[l] > 8 && S->hp_bead[m] == 1 && S->hp_bead[l] == 1 &&
                    S->chnnum[l] != S->chnnum[m])
                    EV(S, m, l) = 1;
    }
    /* symmetrise from the upper triangle (:558-563) */
    for (int i = 1; i <= N - 1; ++i)
        for (int j = i + 1; j <= N; ++j) EV(S, j, i) = EV(S, i, j);
}

/* ------------------------------------------------------------------ */
/* H-bond shoulder bookkeeping: repuls_add.f, repuls_del_a/b.f          */
/* ------------------------------------------------------------------ */
static void shoulder_idx(orc_sys* S, int i, int j, int* ncim1, int* ncai, int* ncaj, int* nnjp1) {
    *ncim1 = i + S->cl[i] - 1; /* repuls_add.f:14-28 */
    *ncai = i - S->cl[i];
    *ncaj = j - 2 * S->cl[j];
    *nnjp1 = j - S->cl[j] + 1;
}
static void repuls_add(orc_sys* S, int i, int j) { /* repuls_add.f:14-47 */
    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(S, i, j, &ncim1, &ncai, &ncaj, &nnjp1);
    EV(S, ncaj, i) = XREPULS2;
    EV(S, i, ncaj) = XREPULS1;
    EV(S, i, nnjp1) = XREPULS1;
    EV(S, nnjp1, i) = XREPULS2;
    EV(S, ncai, j) = XREPULS2;
    EV(S, j, ncai) = XREPULS1;
    EV(S, j, ncim1) = XREPULS1;
    EV(S, ncim1, j) = XREPULS2;
    XR(S, i, 1) = ncaj;
    XR(S, i, 2) = nnjp1;
    XR(S, ncaj, 3) = i;
    XR(S, nnjp1, 3) = i;
    XR(S, j, 1) = ncai;
    XR(S, j, 2) = ncim1;
    XR(S, ncai, 3) = j;
    XR(S, ncim1, 3) = j;
    XR(S, i, 4) = j;
    XR(S, j, 4) = i;
}
static void repuls_del_a(orc_sys* S, int i, int j) { /* repuls_del_a.f:14-37 */
    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(S, i, j, &ncim1, &ncai, &ncaj, &nnjp1);
    EV(S, ncaj, i) = 1;
    EV(S, i, ncaj) = 1;
    EV(S, i, nnjp1) = 1;
    EV(S, nnjp1, i) = 1;
    EV(S, ncai, j) = 1;
    EV(S, j, ncai) = 1;
    EV(S, ncim1, j) = 1;
    EV(S, j, ncim1) = 1;
}
static void repuls_del_b(orc_sys* S, int i, int j) { /* repuls_del_b.f:14-39 */
    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(S, i, j, &ncim1, &ncai, &ncaj, &nnjp1);
    XR(S, i, 1) = 0;
    XR(S, i, 2) = 0;
    XR(S, ncaj, 3) = 0;
    XR(S, nnjp1, 3) = 0;
    XR(S, j, 1) = 0;
    XR(S, j, 2) = 0;
    XR(S, ncai, 3) = 0;
    XR(S, ncim1, 3) = 0;
    XR(S, i, 4) = 0;
    XR(S, j, 4) = 0;
}

/* ------------------------------------------------------------------ */
/* Pair geometry shared by the predictors (core.f:14-24 etc.)           */
/* ------------------------------------------------------------------ */
typedef struct { double rx, ry, rz, vx, vy, vz, b, r2, v2; } pairgeo;
static void geo(const orc_sys* S, int i, int j, double tf, pairgeo* g) {
    g->vx = SV(S, 4, i) - SV(S, 4, j);
    g->vy = SV(S, 5, i) - SV(S, 5, j);
    g->vz = SV(S, 6, i) - SV(S, 6, j);
    double rx = SV(S, 1, i) - SV(S, 1, j) + g->vx * tf;
    double ry = SV(S, 2, i) - SV(S, 2, j) + g->vy * tf;
    double rz = SV(S, 3, i) - SV(S, 3, j) + g->vz * tf;
    g->rx = rx - dnint(rx);
    g->ry = ry - dnint(ry);
    g->rz = rz - dnint(rz);
    g->b = g->rx * g->vx + g->ry * g->vy + g->rz * g->vz;
    g->r2 = g->rx * g->rx + g->ry * g->ry + g->rz * g->rz;
    g->v2 = g->vx * g->vx + g->vy * g->vy + g->vz * g->vz;
}

static void core(const orc_sys* S, int i, int j, int evcode, double* tij, int* type) {
    /* core.f:14-40 */
    const params_t* P = &S->P;
    pairgeo g;
    geo(S, i, j, S->tfalse, &g);
    if (g.b < 0.0) {
        double pe = P->ev_param[1][evcode];
        double sigsq = P->sigma_sq[S->identity[i]][S->identity[j]] * (pe * pe);
        if (evcode >= 22 && evcode <= 26) {
            int k = S->identity[i] > S->identity[j] ? S->identity[i] : S->identity[j];
            double q = P->sqz610[evcode - 21][k];
            sigsq = sigsq * (q * q);
        }
        double discr = g.b * g.b - g.v2 * (g.r2 - sigsq);
        if (discr > 0.0) {
            *tij = (-g.b - sqrt(discr)) / g.v2;
            *type = 1;
        }
    }
}

static void bond_window(const orc_sys* S, int i, int evcode, double* blmin, double* blmax) {
    /* bond.f:19-31 / eventdyn.f:38-58: windows per owner's residue */
    const params_t* P = &S->P;
    int r = S->res[i];
    *blmin = P->ev_param[2][evcode];
    *blmax = P->ev_param[3][evcode];
    if (evcode == 10) {
        *blmin = (1.0 - S->del_bdln[r]) * S->bdln[r];
        *blmax = (1.0 + S->del_bdln[r]) * S->bdln[r];
    } else if (evcode == 11) {
        *blmin = (1.0 - S->del_blrn[r]) * S->bl_rn[r];
        *blmax = (1.0 + S->del_blrn[r]) * S->bl_rn[r];
    } else if (evcode == 12) {
        *blmin = (1.0 - S->del_blrc[r]) * S->bl_rc[r];
        *blmax = (1.0 + S->del_blrc[r]) * S->bl_rc[r];
    }
}

static void bond(const orc_sys* S, int i, int j, int evcode, double* tij, int* type) {
    /* bond.f:17-117 */
    double blmin, blmax;
    bond_window(S, i, evcode, &blmin, &blmax);
    pairgeo g;
    geo(S, i, j, S->tfalse, &g);
    if (g.b < 0.0) {
        double d1 = g.b * g.b - g.v2 * (g.r2 - blmin * blmin);
        if (d1 > 0.0) {
            *tij = (-g.b - sqrt(d1)) / g.v2;
            *type = 2;
        } else {
            double d2 = g.b * g.b - g.v2 * (g.r2 - blmax * blmax);
            if (d2 > 0.0) {
                *tij = (-g.b + sqrt(d2)) / g.v2;
                *type = 3;
            }
        }
    } else {
        double d2 = g.b * g.b - g.v2 * (g.r2 - blmax * blmax);
        if (d2 > 0.0) {
            *tij = -(g.r2 - blmax * blmax) / (sqrt(d2) + g.b);
            *type = 3;
        }
    }
}

static void sqwel(const orc_sys* S, int i, int j, double* tij, int* type) {
    /* sqwel.f:17-63 */
    const params_t* P = &S->P;
    pairgeo g;
    geo(S, i, j, S->tfalse, &g);
    int a = S->identity[i], c = S->identity[j];
    if (g.b < 0.0) {
        double diff = g.r2 - P->welldia_sq[a][c];
        if (diff < 0.0) {
            double cd = g.b * g.b - g.v2 * (g.r2 - P->sigma_sq[a][c]);
            if (cd > 0.0) {
                *tij = (-g.b - sqrt(cd)) / g.v2;
                *type = 1;
            } else {
                double wd = g.b * g.b - g.v2 * diff;
                *tij = (-g.b + sqrt(wd)) / g.v2;
                *type = 8;
            }
        } else {
            double wd = g.b * g.b - g.v2 * diff;
            if (wd > 0.0) {
                *tij = (-g.b - sqrt(wd)) / g.v2;
                *type = 4;
            }
        }
    } else {
        double diff = g.r2 - P->welldia_sq[a][c];
        if (diff < 0.0) {
            double wd = g.b * g.b - g.v2 * diff;
            *tij = (-g.b + sqrt(wd)) / g.v2;
            *type = 8;
        }
    }
}

static void sqshlder(const orc_sys* S, int i, int j, double* tij, int* type) {
    /* sqshlder.f:17-62 */
    const params_t* P = &S->P;
    pairgeo g;
    geo(S, i, j, S->tfalse, &g);
    int a = S->identity[i], c = S->identity[j];
    if (g.b < 0.0) {
        double diff = g.r2 - P->shlddia_sq[a][c];
        if (diff < 0.0) {
            double cd = g.b * g.b - g.v2 * (g.r2 - P->sigma_sq[a][c]);
            if (cd > 0.0) {
                *tij = (-g.b - sqrt(cd)) / g.v2;
                *type = 1;
            } else {
                double sd = g.b * g.b - g.v2 * diff;
                *tij = (-g.b + sqrt(sd)) / g.v2;
                *type = 10;
            }
        } else {
            double sd = g.b * g.b - g.v2 * diff;
            if (sd > 0.0) {
                *tij = (-g.b - sqrt(sd)) / g.v2;
                *type = 12;
            }
        }
    } else {
        double diff = g.r2 - P->shlddia_sq[a][c];
        if (diff < 0.0) {
            double sd = g.b * g.b - g.v2 * diff;
            *tij = (-g.b + sqrt(sd)) / g.v2;
            *type = 10;
        }
    }
}

static void nc_sqwel(const orc_sys* S, int i, int j, double* tij, int* type) {
    /* nc_sqwel.f:19-122 */
    const params_t* P = &S->P;
    pairgeo g;
    geo(S, i, j, S->tfalse, &g);
    int a = S->identity[i], c = S->identity[j];
    double diff = g.r2 - P->welldia_sq[a][c];
    if (a + c == 5) {
        if (g.b < 0.0) {
            if (diff < 0.0) {
                double cd = g.b * g.b - g.v2 * (g.r2 - P->sigma_sq[a][c]);
                if (cd > 0.0) {
                    *tij = (-g.b - sqrt(cd)) / g.v2;
                    *type = 1;
                } else {
                    double wd = g.b * g.b - g.v2 * diff;
                    *tij = (-g.b + sqrt(wd)) / g.v2;
                    *type = 16;
                }
            } else {
                double wd = g.b * g.b - g.v2 * diff;
                if (wd > 0.0) {
                    *tij = (-g.b - sqrt(wd)) / g.v2;
                    *type = 7;
                }
            }
        } else {
            if (diff < 0.0) {
                double wd = g.b * g.b - g.v2 * diff;
                *tij = (-g.b + sqrt(wd)) / g.v2;
                *type = 16;
            }
        }
    } else if (S->bptnr[i] == j) {
        if (g.b < 0.0) {
            double fs = P->sigma_sq[a][c] * P->ev_param[1][15] * P->ev_param[1][15];
            double fc = g.b * g.b - g.v2 * (g.r2 - fs);
            if (fc > 0.0) {
                *tij = (-g.b - sqrt(fc)) / g.v2;
                *type = 1;
            } else {
                double wd = g.b * g.b - g.v2 * diff;
                *tij = (-g.b + sqrt(wd)) / g.v2;
                *type = 8;
            }
        } else {
            double wd = g.b * g.b - g.v2 * diff;
            *tij = (-g.b + sqrt(wd)) / g.v2;
            *type = 8;
        }
    } else {
        if (g.b < 0.0) {
            double wd = g.b * g.b - g.v2 * diff;
            if (wd > 0.0) {
                *tij = (-g.b - sqrt(wd)) / g.v2;
                *type = 9;
            }
        }
    }
}

/* dispatch of events.f:28-47 / eventredo_up.f:27-49 */
static int predict_code(const orc_sys* S, int i, int j, int evcode, double* tij, int* type) {
    if (evcode <= 3)
        core(S, i, j, evcode, tij, type);
    else if (evcode >= 4 && evcode <= 12)
        bond(S, i, j, evcode, tij, type);
    else if (evcode == 15)
        nc_sqwel(S, i, j, tij, type);
    else if (evcode == 16)
        sqwel(S, i, j, tij, type);
    else if (evcode >= 17 && evcode <= 26)
        core(S, i, j, evcode, tij, type);
    else if (evcode >= 40)
        sqshlder(S, i, j, tij, type);
    else
        return -1; /* error in ev_code matrix */
    return 0;
}

/* ------------------------------------------------------------------ */
/* Calendar: add_tbin.f:12-32, del_tbin.f:12-20                         */
/* ------------------------------------------------------------------ */
static void add_tbin(orc_sys* S, int i) {
    int N3 = S->N + 3;
    int j = (int)((S->tim[i] + S->tbin_off) / S->sortsize) + 1;
    if (j > NUMBIN) j = NUMBIN; /* -Ddebugging clamp */
    if (j < 0) {
        if (!S->err) S->err = -10; /* a negative time (-12) reported first keeps its code */
        j = 0;
    }
    S->tlinks[i] = S->bin[j];
    S->tlinks2[i] = j + N3;
    if (S->bin[j] != 0) S->tlinks2[S->bin[j]] = i;
    S->bin[j] = i;
}
static void del_tbin(orc_sys* S, int i) {
    int N3 = S->N + 3;
    if (S->tlinks2[i] > N3)
        S->bin[S->tlinks2[i] - N3] = S->tlinks[i];
    else
        S->tlinks[S->tlinks2[i]] = S->tlinks[i];
    if (S->tlinks[i] != 0) S->tlinks2[S->tlinks[i]] = S->tlinks2[i];
    S->tlinks2[i] = 0;
}

/* eventredo_up.f:25-56 */
static void eventredo_up(orc_sys* S, int i, int j) {
    double tij = 1000000000.0;
    int type = 0;
    int evcode = EV(S, j, i);
    if (predict_code(S, i, j, evcode, &tij, &type)) S->err = -11;
    if (tij < S->tim[i]) {
        S->tim[i] = tij;
        S->nptnr[i] = j;
        S->coltype[i] = type;
    }
    if (tij < 0.0 && fabs(tij) > 1.0e-10) S->err = -12;
}
/* eventredo_down.f:25-78 */
static void eventredo_down(orc_sys* S, int i, int j) {
    double tij = 1000000000.0;
    int type = 0;
    int evcode = EV(S, i, j);
    if (predict_code(S, i, j, evcode, &tij, &type)) S->err = -11;
    if (tij < 0.0 && fabs(tij) > 1.0e-10) S->err = -12;
    tij = tij + S->tfalse;
    if (tij < S->tim[i]) {
        if (S->tlinks2[i] != 0) del_tbin(S, i);
        S->tim[i] = tij;
        S->nptnr[i] = j;
        S->coltype[i] = type;
        if (S->tim[i] < S->interval_max) add_tbin(S, i);
    }
}
/* extra-partner prediction in events.f:77-103 / partial_events.f:37-41 */
static void eventredo_extra(orc_sys* S, int i, int j, int absolute) {
    (void)absolute;
    eventredo_up(S, i, j);
}

/* full recompute of bead l (partial_events.f:121-139 pattern) */
static void redo_full(orc_sys* S, int l) {
    if (S->tlinks2[l] != 0) del_tbin(S, l);
    S->tim[l] = S->interval_max + LTSTEP - S->tfalse;
    S->coltype[l] = -1;
    S->nptnr[l] = -1;
    int ks = (l - 1) * MAXNBS;
    for (int k = 0; k < S->na_npt[l]; ++k) eventredo_up(S, l, S->nb[ks + k]);
    for (int k = 1; k <= 3; ++k)
        if (XR(S, l, k) > l) eventredo_extra(S, l, XR(S, l, k), 0);
    S->tim[l] = S->tim[l] + S->tfalse;
    if (S->tim[l] < S->interval_max) add_tbin(S, l);
}

/* partial_events.f:16-201 */
static void partial_events(orc_sys* S, int i, int j, int xpulse_del) {
    redo_full(S, i);
    if (!(S->canon && j == 0)) redo_full(S, j);
    /* i's down neighbours */
    {
        int ks = (i - 1) * MAXNBS;
        for (int kk = 0; kk < S->nnabdn[i]; ++kk) {
            int l = S->dnnab[ks + kk];
            if (S->nptnr[l] != i)
                eventredo_down(S, l, i);
            else
                redo_full(S, l);
        }
        for (int kk = 1; kk <= 3; ++kk) {
            int l = XR(S, i, kk);
            if (l < i && l != 0) {
                if (S->nptnr[l] != i)
                    eventredo_down(S, l, i);
                else
                    redo_full(S, l);
            }
        }
    }
    if (!(S->canon && j == 0)) {
        int ks = (j - 1) * MAXNBS;
        for (int kk = 0; kk < S->nnabdn[j]; ++kk) {
            int l = S->dnnab[ks + kk];
            if (l != i) {
                if (S->nptnr[l] != j)
                    eventredo_down(S, l, j);
                else
                    redo_full(S, l);
            }
        }
        for (int kk = 1; kk <= 3; ++kk) {
            int l = XR(S, j, kk);
            if (l < j && l != 0) {
                if (S->nptnr[l] != j)
                    eventredo_down(S, l, j);
                else
                    redo_full(S, l);
            }
        }
    }
    if (xpulse_del) {
        if (S->identity[i] < S->identity[j])
            repuls_del_b(S, i, j);
        else
            repuls_del_b(S, j, i);
    }
}

/* events.f:11-127 */
static void events(orc_sys* S) {
    int N = S->N;
    for (int i = 1; i <= N; ++i) {
        int ks = (i - 1) * MAXNBS;
        for (int k = 0; k < S->na_npt[i]; ++k) {
            int j = S->nb[ks + k];
            double tij = 1000000000.0;
            int type = 0;
            int evcode = EV(S, j, i);
            if (predict_code(S, i, j, evcode, &tij, &type)) S->err = -11;
            if (tij < S->tim[i]) {
                S->tim[i] = tij;
                S->nptnr[i] = j;
                S->coltype[i] = type;
            }
            if (tij < 0.0 && fabs(tij) > 1.0e-10) S->err = -12;
        }
        for (int k = 1; k <= 3; ++k) {
            int j = XR(S, i, k);
            if (j > i) {
                double tij = 1000000000.0;
                int type = 0;
                int evcode = EV(S, j, i);
                if (evcode == 1)
                    core(S, i, j, evcode, &tij, &type);
                else
                    sqshlder(S, i, j, &tij, &type);
                if (tij < 0.0 && fabs(tij) > 1.0e-10) S->err = -12;
                if (tij < S->tim[i]) {
                    S->nptnr[i] = j;
                    S->coltype[i] = type;
                    S->tim[i] = tij;
                }
            }
        }
    }
    for (int i = 1; i <= NUMBIN + 1; ++i) S->bin[i] = 0;
    for (int i = 1; i <= N + 3; ++i) {
        S->tlinks[i] = 0;
        S->tlinks2[i] = 0;
        if (S->tim[i] < S->interval_max) add_tbin(S, i);
    }
}

/* ------------------------------------------------------------------ */
/* nbor_setup.f:13-118, cell_link.f:13-94, cell_add.f:12-28, nbor.f     */
/* ------------------------------------------------------------------ */
static void nbor_setup(orc_sys* S) {
    const params_t* P = &S->P;
    double sig_max[51];
    S->sig_max_all = 0.0;
    for (int i = 1; i <= 50; ++i) sig_max[i] = 0.0;
    /* Scan domain: first chain of every species (intra), and the first
     * chains of every pair of species (cross), as nbor_setup.f:20-95 does
     * for its two species. */
    int firsts[4096], nsp = 0;
    for (int c = 0; c < S->nchains; ++c) {
        int seen = 0;
        for (int k = 0; k < nsp; ++k)
            if (S->species[firsts[k]] == S->species[c]) seen = 1;
        if (!seen && nsp < 4096) firsts[nsp++] = c;
    }
#define SCAN_PAIR(i, j)                                                              \
    do {                                                                             \
        double welli = P->welldia[S->identity[i]], wellj = P->welldia[S->identity[j]]; \
        int evcode = EV(S, j, i);                                                    \
        if (evcode <= 26) {                                                          \
            double sig_ij;                                                           \
            if (evcode <= 3)                                                         \
                sig_ij = P->sigma_2b[S->identity[i]][S->identity[j]];                \
            else if (evcode == 15)                                                   \
                sig_ij = 0.5 * (welli + wellj);                                      \
            else if (evcode == 16)                                                   \
                sig_ij = P->wel[S->identity[i]][S->identity[j]] / P->boxl_orig;      \
            else                                                                     \
                sig_ij = 0.0;                                                        \
            if (sig_ij > sig_max[evcode]) {                                          \
                sig_max[evcode] = sig_ij;                                            \
                if (sig_max[evcode] > S->sig_max_all) S->sig_max_all = sig_max[evcode]; \
            }                                                                        \
        }                                                                            \
    } while (0)
    for (int a = 0; a < nsp; ++a) {
        int c = firsts[a], s0 = S->chain_start[c], nbc = 4 * S->chain_len[c];
        for (int i = s0 + 1; i <= s0 + nbc - 1; ++i)
            for (int j = i + 1; j <= s0 + nbc; ++j) SCAN_PAIR(i, j);
    }
    for (int a = 0; a < nsp; ++a)
        for (int b = a + 1; b < nsp; ++b) {
            int c1 = firsts[a], c2 = firsts[b];
            int s1 = S->chain_start[c1], n1 = 4 * S->chain_len[c1];
            int s2 = S->chain_start[c2], n2 = 4 * S->chain_len[c2];
            for (int i = s1 + 1; i <= s1 + n1; ++i)
                for (int j = s2 + 1; j <= s2 + n2; ++j) SCAN_PAIR(i, j);
        }
#undef SCAN_PAIR
    double sigij = 0.0, sig = 0.0;
    for (int i = 1; i <= 4; ++i)
        for (int j = 1; j <= 4; ++j) {
            if (i != 3 || j != 3) sig = P->sigma[i] + P->sigma[j];
            if (sig > sigij) sigij = sig;
        }
    for (int i = 40; i <= 50; ++i) {
        double sig_ij = 0.5 * sigij * 1.0;
        sig_max[i] = sig_ij;
        if (sig_max[i] > S->sig_max_all) S->sig_max_all = sig_max[i];
    }
    for (int i = 1; i <= 50; ++i) {
        S->rlsq[i] = 0.0;
        if (sig_max[i] != 0.0) {
            double v = (RL_CONST - 1) * S->sig_max_all + sig_max[i];
            S->rlsq[i] = v * v;
        }
    }
    double rl = RL_CONST * S->sig_max_all;
    double h = 0.4 * (rl - S->sig_max_all);
    S->hdelr = h * h;
}

static int i_cell(int nc, int ix, int iy, int iz) {
    return 1 + (ix - 1 + nc) % nc + ((iy - 1 + nc) % nc) * nc + ((iz - 1 + nc) % nc) * nc * nc;
}
static void cell_link(orc_sys* S) { /* cell_link.f:13-94 (n_wrap == 2) */
    int nc = S->num_cell;
    static const int o[6] = {0, 0, 1, -1, 2, -2};
    static const int oy[4] = {0, 0, 1, 2};
    int d_cell = 1;
    for (int iz = 1; iz <= 3; ++iz) {
        int dz = o[iz];
        for (int iy = 1; iy <= 2; ++iy) {
            int dy = o[iy];
            for (int ix = 1; ix <= 3; ++ix) {
                int dx = o[ix];
                if (dy == 0) {
                    if (dz < 1) {
                        if (dx < 1) continue;
                    } else if (dx < 0)
                        continue;
                }
                S->map[d_cell++] = dx + (dy + dz * nc) * nc;
            }
        }
    }
    for (int iz = 1; iz <= 5; ++iz) {
        int dz = o[iz];
        for (int iy = 1; iy <= 3; ++iy) {
            int dy = oy[iy];
            int min_ix = (iz < 4 && iy != 3) ? 4 : 1;
            for (int ix = min_ix; ix <= 5; ++ix) {
                int dx = o[ix];
                if (dy == 0) {
                    if (dz < 1) {
                        if (dx < 1) continue;
                    } else if (dx < 0)
                        continue;
                }
                S->map[d_cell++] = dx + (dy + dz * nc) * nc;
            }
        }
    }
    for (int ix = 1; ix <= nc; ++ix) {
        int x = ix <= N_WRAP ? ix + nc - 2 * N_WRAP : (nc - ix < N_WRAP ? ix - nc + 2 * N_WRAP : ix);
        for (int iy = 1; iy <= nc; ++iy) {
            int y = iy <= N_WRAP ? iy + nc - 2 * N_WRAP
                                 : (nc - iy < N_WRAP ? iy - nc + 2 * N_WRAP : iy);
            for (int iz = 1; iz <= nc; ++iz) {
                int z = iz <= N_WRAP ? iz + nc - 2 * N_WRAP
                                     : (nc - iz < N_WRAP ? iz - nc + 2 * N_WRAP : iz);
                S->wrap_map[i_cell(nc, ix, iy, iz)] = i_cell(nc, x, y, z);
            }
        }
    }
}
static int cell_index(const orc_sys* S, int k) { /* cell_add.f:24-27 */
    int nc = S->num_cell;
    int x = (int)((SV(S, 1, k) + S->half) / S->width) + N_WRAP;
    int y = (int)((SV(S, 2, k) + S->half) / S->width) + N_WRAP;
    int z = (int)((SV(S, 3, k) + S->half) / S->width) + N_WRAP;
    return 1 + x + y * nc + z * nc * nc;
}
int32_t orc_cell_of(const orc_sys* S, int32_t k) { return cell_index(S, k); }

static void cell_add(orc_sys* S) {
    int N = S->N, nc = S->num_cell;
    for (int k = 1; k <= N; ++k) S->clinks[k] = 0;
    for (int k = 1; k <= nc * nc * nc + 1; ++k) S->cell[k] = 0;
    for (int k = 1; k <= N; ++k) {
        int c = cell_index(S, k);
        S->clinks[k] = S->cell[c];
        S->cell[c] = k;
    }
}

static void nb_add(orc_sys* S, int lo, int hi) {
    int l = S->na_npt[lo];
    if (l < MAXNBS) {
        S->nb[(lo - 1) * MAXNBS + l] = hi;
        S->na_npt[lo] = l + 1;
    } else
        S->err = -13;
    l = S->nnabdn[hi];
    if (l < MAXNBS) {
        S->dnnab[(hi - 1) * MAXNBS + l] = lo;
        S->nnabdn[hi] = l + 1;
    } else
        S->err = -13;
}

static void nbor(orc_sys* S) { /* nbor.f:23-137 */
    int N = S->N, nc = S->num_cell;
    for (int l = 1; l <= N; ++l) {
        S->na_npt[l] = 0;
        S->nnabdn[l] = 0;
    }
    cell_add(S);
    int* n_cell = (int*)malloc(sizeof(int) * (size_t)(N + 1));
    for (int c = N_WRAP * nc * nc; c <= nc * nc * (nc - N_WRAP); ++c) {
        int n_bead = 0, bead = S->cell[c];
        if (bead == 0) continue;
        while (bead != 0) {
            n_cell[++n_bead] = bead;
            bead = S->clinks[bead];
        }
        int f_bead = n_bead;
        for (int n = 1; n <= N_NAB_CELL; ++n) {
            int ncid = S->wrap_map[c + S->map[n]];
            bead = S->cell[ncid];
            while (bead != 0) {
                n_cell[++n_bead] = bead;
                bead = S->clinks[bead];
            }
        }
        for (int i_n = 1; i_n <= f_bead; ++i_n) {
            int i = n_cell[i_n];
            for (int j_n = i_n + 1; j_n <= n_bead; ++j_n) {
                int j = n_cell[j_n];
                int evcode = EV(S, j, i);
                if ((evcode >= 4 && evcode <= 12) || (evcode >= 17 && evcode < 27)) {
                    if (j > i)
                        nb_add(S, i, j);
                    else
                        nb_add(S, j, i);
                } else {
                    double rx = SV(S, 1, i) - SV(S, 1, j), ry = SV(S, 2, i) - SV(S, 2, j),
                           rz = SV(S, 3, i) - SV(S, 3, j);
                    rx = rx - dnint(rx);
                    ry = ry - dnint(ry);
                    rz = rz - dnint(rz);
                    double r2 = rx * rx + ry * ry + rz * rz;
                    if (evcode >= 1 && evcode <= 50 && r2 <= S->rlsq[evcode]) {
                        if (j > i)
                            nb_add(S, i, j);
                        else
                            nb_add(S, j, i);
                    }
                }
            }
        }
    }
    free(n_cell);
}

/* displ.f:20-46 */
static int displ(orc_sys* S) {
    double moved_far = 0.0;
    for (int i = 1; i <= S->N; ++i) {
        double a = S->old_r[3 * i - 3] - SV(S, 1, i);
        double b = S->old_r[3 * i - 2] - SV(S, 2, i);
        double c = S->old_r[3 * i - 1] - SV(S, 3, i);
        double dis = a * a + b * b + c * c;
        double moved = dis / S->hdelr;
        if (moved > moved_far) moved_far = moved;
    }
    if (moved_far >= 0.1) {
        if (moved_far >= 1.25 * 1.25) {
            S->t_fact = S->t_fact / 1.01;
            S->interval = S->t_fact / sqrt(S->setemp);
        }
        return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Collision dynamics: eventdyn.f:4-385, bumped.f:1-47                  */
/* ------------------------------------------------------------------ */
static void nudge(orc_sys* S, int i, int j, double bumpdist, const pairgeo* g, int sign) {
    /* sign=+1: i moves along +r, j along -r; sign=-1 the reverse */
    if (sign > 0) {
        SV(S, 1, i) = SV(S, 1, i) + bumpdist * g->rx;
        SV(S, 2, i) = SV(S, 2, i) + bumpdist * g->ry;
        SV(S, 3, i) = SV(S, 3, i) + bumpdist * g->rz;
        SV(S, 1, j) = SV(S, 1, j) - bumpdist * g->rx;
        SV(S, 2, j) = SV(S, 2, j) - bumpdist * g->ry;
        SV(S, 3, j) = SV(S, 3, j) - bumpdist * g->rz;
    } else {
        SV(S, 1, i) = SV(S, 1, i) - bumpdist * g->rx;
        SV(S, 2, i) = SV(S, 2, i) - bumpdist * g->ry;
        SV(S, 3, i) = SV(S, 3, i) - bumpdist * g->rz;
        SV(S, 1, j) = SV(S, 1, j) + bumpdist * g->rx;
        SV(S, 2, j) = SV(S, 2, j) + bumpdist * g->ry;
        SV(S, 3, j) = SV(S, 3, j) + bumpdist * g->rz;
    }
}

static void eventdyn(orc_sys* S, int i, int j, int evcode, int* ct) {
    const params_t* P = &S->P;
    pairgeo g;
    geo(S, i, j, S->tfalse, &g);
    double bmi = S->bm[i], bmj = S->bm[j];
    double rmass = 2 * bmi * bmj / (bmi + bmj);
    int a = S->identity[i], c = S->identity[j];
    double ratio = 0.0, wellsq, epsave, del_pe, bumpdist;
    switch (*ct) {
    case 2: {
        double blmin, blmax;
        bond_window(S, i, evcode, &blmin, &blmax);
        ratio = rmass * g.b / (blmin * blmin);
        break;
    }
    case 3: {
        double blmin, blmax;
        bond_window(S, i, evcode, &blmin, &blmax);
        ratio = rmass * g.b / (blmax * blmax);
        break;
    }
    case 1: {
        double sigsq;
        if (evcode == 15) {
            if (S->bptnr[i] == j) {
                double pe = P->ev_param[1][evcode];
                sigsq = P->sigma_sq[a][c] * (pe * pe);
            } else
                sigsq = P->sigma_sq[a][c];
        } else {
            double pe = P->ev_param[1][evcode];
            sigsq = P->sigma_sq[a][c] * (pe * pe);
            if (evcode >= 22 && evcode <= 26) {
                int k = a > c ? a : c;
                double q = P->sqz610[evcode - 21][k];
                sigsq = sigsq * (q * q);
            }
        }
        ratio = rmass * g.b / sigsq;
        break;
    }
    case 4:
        wellsq = P->welldia_sq[a][c];
        epsave = P->ep_sqrt[a][c];
        del_pe = 4.0 * wellsq * epsave / rmass;
        if (g.b * g.b + del_pe > 0.0) {
            ratio = rmass * (sqrt((4.0 * wellsq * epsave / rmass) + g.b * g.b) + g.b) / (2.0 * wellsq);
            *ct = 20;
            bumpdist = SMDIST * sqrt(wellsq);
            nudge(S, i, j, bumpdist, &g, -1);
        } else {
            ratio = rmass * g.b / wellsq;
            *ct = 22;
            bumpdist = SMDIST * sqrt(wellsq);
            nudge(S, i, j, bumpdist, &g, +1);
        }
        break;
    case 8:
        wellsq = P->welldia_sq[a][c];
        epsave = P->ep_sqrt[a][c];
        del_pe = 4.0 * wellsq * epsave / rmass;
        bumpdist = SMDIST * sqrt(wellsq);
        if (g.b * g.b > del_pe) {
            ratio = rmass * (-sqrt(-del_pe + g.b * g.b) + g.b) / (2.0 * wellsq);
            *ct = 21;
            nudge(S, i, j, bumpdist, &g, +1);
        } else {
            ratio = rmass * g.b / wellsq;
            *ct = 22;
            nudge(S, i, j, bumpdist, &g, -1);
        }
        break;
    case 9:
        wellsq = P->welldia_sq[a][c];
        ratio = rmass * g.b / wellsq;
        *ct = 23;
        bumpdist = SMDIST * sqrt(wellsq);
        nudge(S, i, j, bumpdist, &g, +1);
        break;
    case 5:
        wellsq = P->shlddia_sq[a][c];
        epsave = -P->epsilon[1];
        del_pe = 4.0 * wellsq * epsave / rmass;
        bumpdist = SMDIST * sqrt(wellsq);
        ratio = rmass * (-sqrt(-del_pe + g.b * g.b) + g.b) / (2.0 * wellsq);
        *ct = 24;
        nudge(S, i, j, bumpdist, &g, +1);
        break;
    case 6:
        wellsq = P->shlddia_sq[a][c];
        epsave = -P->epsilon[1];
        del_pe = 4.0 * wellsq * epsave / rmass;
        bumpdist = SMDIST * sqrt(wellsq);
        if (g.b * g.b > -del_pe) {
            ratio = rmass * (sqrt(del_pe + g.b * g.b) + g.b) / (2.0 * wellsq);
            *ct = 25;
            nudge(S, i, j, bumpdist, &g, -1);
        } else {
            ratio = rmass * g.b / wellsq;
            *ct = 26;
            nudge(S, i, j, bumpdist, &g, +1);
        }
        break;
    case 13:
        wellsq = P->shlddia_sq[a][c];
        ratio = rmass * g.b / wellsq;
        *ct = 27;
        bumpdist = SMDIST * sqrt(wellsq);
        nudge(S, i, j, bumpdist, &g, +1);
        break;
    default:
        S->err = -14;
        return;
    }
    double tf = S->tfalse;
    double dvx = ratio * g.rx, dvy = ratio * g.ry, dvz = ratio * g.rz;
    SV(S, 4, i) = SV(S, 4, i) - dvx / bmi;
    SV(S, 4, j) = SV(S, 4, j) + dvx / bmj;
    SV(S, 5, i) = SV(S, 5, i) - dvy / bmi;
    SV(S, 5, j) = SV(S, 5, j) + dvy / bmj;
    SV(S, 6, i) = SV(S, 6, i) - dvz / bmi;
    SV(S, 6, j) = SV(S, 6, j) + dvz / bmj;
    SV(S, 1, i) = SV(S, 1, i) + dvx * tf / bmi;
    SV(S, 2, i) = SV(S, 2, i) + dvy * tf / bmi;
    SV(S, 3, i) = SV(S, 3, i) + dvz * tf / bmi;
    SV(S, 1, j) = SV(S, 1, j) - dvx * tf / bmj;
    SV(S, 2, j) = SV(S, 2, j) - dvy * tf / bmj;
    SV(S, 3, j) = SV(S, 3, j) - dvz * tf / bmj;
}

static void bumpoff(orc_sys* S, int i, int j, int evcode) { /* bumped.f:1-47 */
    const params_t* P = &S->P;
    pairgeo g;
    geo(S, i, j, S->tfalse, &g);
    double bumpdist;
    if (evcode >= 40)
        bumpdist = SMDIST * sqrt(P->shlddia_sq[S->identity[i]][S->identity[j]]);
    else
        bumpdist = SMDIST * sqrt(P->welldia_sq[S->identity[i]][S->identity[j]]);
    nudge(S, i, j, bumpdist, &g, g.b < 0.0 ? -1 : +1);
}

/* distance at tfalse used by repuls_check*.f / check_sigma.f */
static double dist_tf(const orc_sys* S, int a, int b) {
    pairgeo g;
    geo(S, a, b, S->tfalse, &g);
    return sqrt(g.rx * g.rx + g.ry * g.ry + g.rz * g.rz);
}
static double repuls_check(orc_sys* S, int i, int j) { /* repuls_check.f:16-81 */
    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(S, i, j, &ncim1, &ncai, &ncaj, &nnjp1);
    double d1 = dist_tf(S, i, ncaj), d2 = dist_tf(S, i, nnjp1);
    double d3 = dist_tf(S, j, ncai), d4 = dist_tf(S, j, ncim1);
    const params_t* P = &S->P;
    if (d1 > P->shder_dist1 && d2 > P->shder_dist2 && d3 > P->shder_dist3 && d4 > P->shder_dist4)
        return 10.0;
    return 11.0;
}
static double repuls_check_3(orc_sys* S, int i, int j, int k) { /* repuls_check_3.f:16-102 */
    int ncim1, ncai, ncaj, nnjp1;
    shoulder_idx(S, i, j, &ncim1, &ncai, &ncaj, &nnjp1);
    const params_t* P = &S->P;
    int m = 0;
    double d1 = dist_tf(S, i, ncaj);
    if (ncaj != k && d1 > P->shder_dist1) m++;
    double d2 = dist_tf(S, i, nnjp1);
    if (nnjp1 != k && d2 > P->shder_dist2) m++;
    double d3 = dist_tf(S, j, ncai);
    if (ncai != k && d3 > P->shder_dist3) m++;
    double d4 = dist_tf(S, j, ncim1);
    if (ncim1 != k && d4 > P->shder_dist4) m++;
    return m == 3 ? 10.0 : 11.0;
}
static double check_sigma(orc_sys* S, int i, int j) { /* check_sigma.f:12-28 */
    pairgeo g;
    geo(S, i, j, S->tfalse, &g);
    double rijsq = g.rx * g.rx + g.ry * g.ry + g.rz * g.rz;
    double diff = rijsq - S->P.sigma_sq[S->identity[i]][S->identity[j]];
    return diff < 0.0 ? 10.0 : 11.0;
}

static int terminal(const orc_sys* S, int k) { /* main.F90:1493-1494 end-bead test */
    return S->loc[k] == S->cl[k] + 1 || S->loc[k] == 3 * S->cl[k];
}

/* ------------------------------------------------------------------ */
/* One pair event: worker resolution (main.F90:1432-1960) then commit   */
/* (main.F90:641-943) on current state.                                 */
/* ------------------------------------------------------------------ */
static int pair_event(orc_sys* S, int i) {
    int j = S->nptnr[i];
    int ct = S->coltype[i];
    int evcode = EV(S, i, j);
    int ok = !terminal(S, i) && !terminal(S, j);
    int hb_partner = 0, xpulse = 0;
    if (ct == 7) {
        if (XR(S, i, 4) == 0 && XR(S, j, 4) == 0) {
            if (ok) {
                double rating = S->identity[i] < S->identity[j] ? repuls_check(S, i, j)
                                                                : repuls_check(S, j, i);
                ct = rating <= 10.0 ? 4 : 14;
            } else {
                double ran_non = drand(S, 1);
                ct = ran_non <= 0.2 ? 4 : 9;
            }
        } else
            ct = 9;
    } else if (ct == 10) {
        double rating;
        if (evcode < 45) {
            hb_partner = XR(S, i, 4);
            rating = S->identity[i] < S->identity[hb_partner] ? repuls_check_3(S, i, hb_partner, j)
                                                              : repuls_check_3(S, hb_partner, i, j);
        } else {
            hb_partner = XR(S, j, 4);
            rating = S->identity[j] < S->identity[hb_partner] ? repuls_check_3(S, j, hb_partner, i)
                                                              : repuls_check_3(S, hb_partner, j, i);
        }
        ct = rating <= 10.0 ? 5 : 15;
    } else if (ct == 12) {
        int x = evcode < 45 ? i : j;
        hb_partner = XR(S, x, 4);
        if (S->bptnr[x] == hb_partner) {
            double rating = check_sigma(S, x, hb_partner);
            ct = rating > 10.0 ? 6 : 13;
        } else
            ct = 15;
    }
    if (ct < 14) eventdyn(S, i, j, evcode, &ct);
    if (ct == 20) {
        if (S->identity[i] + S->identity[j] == 5) {
            S->bptnr[i] = j;
            S->bptnr[j] = i;
            S->identity[i] += 4;
            S->identity[j] += 4;
            if (ok) {
                if (S->identity[i] < S->identity[j])
                    repuls_add(S, i, j);
                else
                    repuls_add(S, j, i);
            }
        }
    } else if (ct == 21) {
        if (S->identity[i] <= 8) {
            if (ok) {
                if (S->identity[i] < S->identity[j])
                    repuls_del_a(S, i, j);
                else
                    repuls_del_a(S, j, i);
                xpulse = 1;
            }
            if (S->bptnr[i] == j) {
                S->bptnr[i] = 0;
                S->bptnr[j] = 0;
                S->identity[i] -= 4;
                S->identity[j] -= 4;
            }
        }
    } else if (ct == 24) {
        int x = evcode < 45 ? i : j;
        hb_partner = XR(S, x, 4);
        if (S->identity[x] + S->identity[hb_partner] == 5) {
            S->bptnr[x] = hb_partner;
            S->bptnr[hb_partner] = x;
            S->identity[x] += 4;
            S->identity[hb_partner] += 4;
        }
    } else if (ct == 25) {
        int x = evcode < 45 ? i : j;
        hb_partner = XR(S, x, 4);
        if (S->identity[x] >= 5) {
            S->bptnr[x] = 0;
            S->bptnr[hb_partner] = 0;
            S->identity[x] -= 4;
            S->identity[hb_partner] -= 4;
        }
    } else if (ct == 14) {
        bumpoff(S, i, j, evcode);
        if (ok) {
            if (S->identity[i] < S->identity[j])
                repuls_add(S, i, j);
            else
                repuls_add(S, j, i);
        }
    } else if (ct == 15) {
        bumpoff(S, i, j, evcode);
    } else if (ct == 16) {
        bumpoff(S, i, j, evcode);
        if (ok) {
            if (S->identity[i] < S->identity[j])
                repuls_del_a(S, i, j);
            else
                repuls_del_a(S, j, i);
            xpulse = 1;
        }
    }
    S->coltype[i] = ct;
    if (ct >= 1 && ct <= 30) S->nevents[ct]++;
    partial_events(S, i, j, xpulse);
    return ct;
}

/* Andersen ghost, main.F90:997-1049 */
static int ghost_event(orc_sys* S) {
    int N = S->N, i;
    S->numghosts++;
    do {
        i = (int)(drand(S, 0) * N) + 1;
    } while (i == N + 1);
    double tf = S->tfalse, bm = S->bm[i], v1, v2, r, fact;
    SV(S, 1, i) = SV(S, 1, i) + SV(S, 4, i) * tf;
    SV(S, 2, i) = SV(S, 2, i) + SV(S, 5, i) * tf;
    SV(S, 3, i) = SV(S, 3, i) + SV(S, 6, i) * tf;
    do {
        v1 = 2.0 * drand(S, 0) - 1.0;
        v2 = 2.0 * drand(S, 0) - 1.0;
        r = v1 * v1 + v2 * v2;
    } while (r == 0.0 || r >= 1.0);
    fact = sqrt(-2.0 * S->setemp * bm * orc_log(r) / r);
    SV(S, 4, i) = v1 * fact / bm;
    SV(S, 5, i) = v2 * fact / bm;
    do {
        v1 = 2.0 * drand(S, 0) - 1.0;
        v2 = 2.0 * drand(S, 0) - 1.0;
        r = v1 * v1 + v2 * v2;
    } while (r == 0.0 || r >= 1.0);
    fact = sqrt(-2.0 * S->setemp * bm * orc_log(r) / r);
    SV(S, 6, i) = v1 * fact / bm;
    SV(S, 1, i) = SV(S, 1, i) - SV(S, 4, i) * tf;
    SV(S, 2, i) = SV(S, 2, i) - SV(S, 5, i) * tf;
    SV(S, 3, i) = SV(S, 3, i) - SV(S, 6, i) * tf;
    if (S->tlinks2[N + 1] != 0) del_tbin(S, N + 1);
    double tgho;
    do {
        tgho = drand(S, 0);
    } while (tgho < 1.e-18 || tgho == 1.0);
    S->tim[N + 1] = -1.0 * orc_log(tgho) * S->avegtime + S->tfalse;
    if (S->tim[N + 1] < S->interval_max) add_tbin(S, N + 1);
    if (S->tfalse < S->old_tfalse) S->tfalse = S->old_tfalse;
    partial_events(S, i, 0, 0);
    return i;
}

static void rebuild_all(orc_sys* S) {
    int N = S->N;
    for (int k = 1; k <= N; ++k) {
        SV(S, 1, k) = SV(S, 1, k) - round(SV(S, 1, k));
        SV(S, 2, k) = SV(S, 2, k) - round(SV(S, 2, k));
        SV(S, 3, k) = SV(S, 3, k) - round(SV(S, 3, k));
        S->old_r[3 * k - 3] = SV(S, 1, k);
        S->old_r[3 * k - 2] = SV(S, 2, k);
        S->old_r[3 * k - 1] = SV(S, 3, k);
    }
    nbor(S);
    for (int k = 1; k <= N; ++k) {
        S->tim[k] = S->interval_max + LTSTEP;
        S->coltype[k] = -1;
        S->nptnr[k] = -1;
    }
    events(S);
    S->nbin = 1;
}

/* sync pseudo-event, main.F90:1126-1187 */
static void sync_event(orc_sys* S) {
    int N = S->N, i = N + 2;
    double tf = S->tfalse;
    S->t = S->t + tf;
    for (int k = 1; k <= N + 3; ++k) S->tim[k] = S->tim[k] - tf;
    S->interval_max = S->interval_max - tf;
    S->tbin_off = S->tbin_off + tf;
    for (int k = 1; k <= N; ++k) {
        SV(S, 1, k) = SV(S, 1, k) + SV(S, 4, k) * tf;
        SV(S, 2, k) = SV(S, 2, k) + SV(S, 5, k) * tf;
        SV(S, 3, k) = SV(S, 3, k) + SV(S, 6, k) * tf;
    }
    S->tfalse = 0.0;
    int update = displ(S);
    if (update || S->interval > S->interval_max) {
        if (!update) {
            S->nforced++;
            S->n_forced = S->n_forced * 1.01;
        }
        S->interval_max = S->interval * S->n_forced;
        S->sortsize = S->interval_max / (double)NUMBIN;
        S->tbin_off = 0.0;
        S->nupdates++;
        rebuild_all(S);
    }
    if (S->tlinks2[i] != 0) del_tbin(S, i);
    S->tim[i] = S->interval * 0.999;
    add_tbin(S, i);
}

/* A list update forced at the current time: the sync pseudo-event's shift of
 * the time origin and its update branch (main.F90:1126-1187: positions to
 * the current time, times shifted, calendar rescaled, nbor.f + events.f),
 * without the displacement test, the rate adaptation or a new sync entry. */
void orc_rebuild(orc_sys* S) {
    int N = S->N;
    double tf = S->tfalse;
    S->t = S->t + tf;
    for (int k = 1; k <= N + 3; ++k) S->tim[k] = S->tim[k] - tf;
    for (int k = 1; k <= N; ++k) {
        SV(S, 1, k) = SV(S, 1, k) + SV(S, 4, k) * tf;
        SV(S, 2, k) = SV(S, 2, k) + SV(S, 5, k) * tf;
        SV(S, 3, k) = SV(S, 3, k) + SV(S, 6, k) * tf;
    }
    S->tfalse = 0.0;
    S->interval_max = S->interval * S->n_forced;
    S->sortsize = S->interval_max / (double)NUMBIN;
    S->tbin_off = 0.0;
    S->nupdates++;
    rebuild_all(S);
}

static void energy(const orc_sys* S, orc_obs* o);
/* output pseudo-event, main.F90:1191-1246 */
static void output_event(orc_sys* S) {
    int i = S->N + 3;
    if (S->tlinks2[i] != 0) del_tbin(S, i);
    S->tim[i] = 3.3 / (sqrt(S->setemp)) + 5 + S->tfalse;
    if (S->tim[i] < S->interval_max) add_tbin(S, i);
}

int64_t orc_run(orc_sys* S, int64_t nev, orc_event* log, int64_t log_cap) {
    int N = S->N;
    int64_t done = 0;
    while (done < nev) {
        S->old_tfalse = S->tfalse;
        while (S->bin[S->nbin] == 0) {
            S->nbin++;
            if (S->nbin > NUMBIN + 1) return -20;
        }
        /* earliest entry of the first non-empty bin, first in list order on ties
         * (stable insertion sort of main.F90:521-545, first element) */
        int best = S->bin[S->nbin];
        for (int k = S->tlinks[best]; k != 0; k = S->tlinks[k])
            if (S->tim[k] < S->tim[best]) best = k;
        int i = best;
        const double tsel = S->tim[i];
        S->tfalse = tsel;
        S->coll++;
        int j = 0, type = 0;
        if (i <= N) {
            j = S->nptnr[i];
            type = pair_event(S, i);
        } else if (i == N + 1) {
            j = ghost_event(S);
        } else if (i == N + 2) {
            sync_event(S);
        } else {
            output_event(S);
        }
        if (log && done < log_cap) {
            log[done].coll = S->coll;
            log[done].i = i;
            log[done].j = j;
            log[done].type = type;
            log[done].pad = 0;
            log[done].t = tsel;
        }
        done++;
        if (S->err) return S->err;
    }
    return done;
}

/* ------------------------------------------------------------------ */
/* Observables: energy.f:25-101, radgyr.f:18-135, end2end.f:14-44       */
/* ------------------------------------------------------------------ */
static void energy(const orc_sys* S, orc_obs* o) {
    const params_t* P = &S->P;
    int N = S->N;
    int hb_ii = 0, hb_ij = 0, hb_alpha = 0;
    double ehh_ii = 0.0, ehh_ij = 0.0;
    for (int i = 1; i <= N - 1; ++i)
        for (int j = i + 1; j <= N; ++j) {
            if (S->bptnr[i] == j) {
                if (S->chnnum[i] == S->chnnum[j])
                    hb_ii++;
                else
                    hb_ij++;
            } else if (EV(S, j, i) == 16) {
                pairgeo g;
                geo(S, i, j, S->tfalse, &g);
                double r2 = g.rx * g.rx + g.ry * g.ry + g.rz * g.rz;
                double wellsq = P->welldia_sq[S->identity[i]][S->identity[j]];
                double dep = P->ep_sqrt[S->identity[i]][S->identity[j]];
                if (r2 <= wellsq) {
                    if (S->chnnum[i] == S->chnnum[j])
                        ehh_ii += dep;
                    else
                        ehh_ij += dep;
                }
            }
        }
    for (int c = 0; c < S->nchains; ++c) {
        int lll = S->chain_start[c], L = S->chain_len[c];
        for (int i = lll + L + 5; i <= lll + 2 * L; ++i)
            if (S->bptnr[i] == i + L - 4) hb_alpha++;
    }
    double eps_hb = P->ep_sqrt[5][8];
    int sum_hb = hb_ii + hb_ij;
    double sum_ehh = ehh_ii + ehh_ij;
    double sumeps = -(sum_hb * eps_hb + sum_ehh);
    double sumvel = 0.0;
    for (int i = 1; i <= N; ++i)
        sumvel = sumvel + S->bm[i] * (SV(S, 4, i) * SV(S, 4, i) + SV(S, 5, i) * SV(S, 5, i) +
                                      SV(S, 6, i) * SV(S, 6, i));
    o->ered = 0.5 * sumvel + sumeps;
    o->tred = sumvel / 3.0 / (double)N;
    o->sumvel = sumvel;
    o->hb_alpha = hb_alpha;
    o->hb_ii = hb_ii;
    o->hb_ij = hb_ij;
    o->ehh_ii = ehh_ii;
    o->ehh_ij = ehh_ij;
}

void orc_observables(const orc_sys* S, orc_obs* o) {
    memset(o, 0, sizeof *o);
    energy(S, o);
    int N = S->N;
    double tf = S->tfalse, half = S->half, boxl = S->boxl;
    double* st = (double*)malloc(sizeof(double) * 3 * (size_t)(N + 1));
    double rg = 0.0, e2e = 0.0;
    int nbb = 0;
    /* chains of the leading species block are the reference's species 1 */
    int nfirst = 0;
    while (nfirst < S->nchains && S->species[nfirst] == S->species[0]) nfirst++;
    for (int c = 0; c < S->nchains; ++c) {
        int ii = S->chain_start[c], L = S->chain_len[c];
        /* radgyr.f:117-125 sums a species-2 chain i over beads
         * (i-1)*numbeads2+1 .. (i-1)*numbeads2+3*chnln2, without the nop1
         * offset its unwrapping uses (:78); kept as the reference computes
         * it while those beads are already filled, the chain's own beads
         * otherwise (the reference would read unset memory there) */
        int rb = ii;
        if (c >= nfirst && c * 4 * L + 3 * L <= ii + 4 * L) rb = c * 4 * L;
        for (int jb = ii + 1; jb <= ii + 4 * L; ++jb)
            for (int k = 1; k <= 3; ++k) st[3 * jb + k - 1] = SV(S, k, jb) + SV(S, k + 3, jb) * tf;
#define UNWRAP(a, b)                                               \
    for (int k = 1; k <= 3; ++k) {                                 \
        double d = st[3 * (a) + k - 1] - st[3 * (b) + k - 1];      \
        if (d > half)                                              \
            st[3 * (a) + k - 1] = st[3 * (a) + k - 1] - boxl;      \
        else if (d < -1.0 * half)                                  \
            st[3 * (a) + k - 1] = st[3 * (a) + k - 1] + boxl;      \
    }
        for (int jb = ii + 2; jb <= ii + L; ++jb) UNWRAP(jb, jb - 1);
        for (int jb = ii + 1; jb <= ii + L; ++jb) UNWRAP(L + jb, jb);
        for (int jb = ii + 1; jb <= ii + L; ++jb) UNWRAP(2 * L + jb, jb);
#undef UNWRAP
        double cx = 0.0, cy = 0.0, cz = 0.0, rgsum = 0.0;
        for (int jb = rb + 1; jb <= rb + 3 * L; ++jb) {
            cx = cx + st[3 * jb];
            cy = cy + st[3 * jb + 1];
            cz = cz + st[3 * jb + 2];
        }
        cx = cx / (double)(3 * L);
        cy = cy / (double)(3 * L);
        cz = cz / (double)(3 * L);
        for (int jb = rb + 1; jb <= rb + 3 * L; ++jb) {
            double rx = (st[3 * jb] - cx) * (st[3 * jb] - cx);
            double ry = (st[3 * jb + 1] - cy) * (st[3 * jb + 1] - cy);
            double rz = (st[3 * jb + 2] - cz) * (st[3 * jb + 2] - cz);
            rgsum = rgsum + rx + ry + rz;
        }
        rg = rg + rgsum;
        nbb += 3 * L;
        /* end2end.f: N of residue 1 to C of residue L */
        pairgeo g;
        geo(S, ii + L + 1, ii + 3 * L, tf, &g);
        double ex = g.rx, ey = g.ry, ez = g.rz;
        e2e = e2e + sqrt(ex * ex + ey * ey + ez * ez);
    }
    free(st);
    o->rg_avg = sqrt(rg / (double)nbb) * S->P.boxl_orig;
    o->e2e_avg = e2e / (double)S->nchains * S->P.boxl_orig;
}

/* checkover.f:21-131 */
int32_t orc_checkover(const orc_sys* S) {
    const params_t* P = &S->P;
    int N = S->N, bad = 0;
    for (int i = 1; i <= N - 1; ++i)
        for (int j = i + 1; j <= N; ++j) {
            int evcode = EV(S, j, i);
            pairgeo g;
            if (evcode == 1 || evcode >= 15) {
                double pe = P->ev_param[1][evcode];
                double sigsq = P->sigma_sq[S->identity[i]][S->identity[j]] * (pe * pe);
                if (evcode >= 22 && evcode <= 26) {
                    int k = S->identity[i] > S->identity[j] ? S->identity[i] : S->identity[j];
                    double q = P->sqz610[evcode - 21][k];
                    sigsq = sigsq * (q * q);
                }
                geo(S, i, j, S->tfalse, &g);
                double r2 = g.rx * g.rx + g.ry * g.ry + g.rz * g.rz;
                r2 = r2 * 1.0000000001;
                if (r2 <= sigsq) bad++;
            } else if (evcode >= 4 && evcode <= 12 && !S->chaptype2) { /* checkover.f:67: chaptype 1 only */
                double blmin, blmax;
                bond_window(S, i, evcode, &blmin, &blmax);
                blmin = blmin * blmin;
                blmax = blmax * blmax;
                geo(S, i, j, S->tfalse, &g);
                double r2 = g.rx * g.rx + g.ry * g.ry + g.rz * g.rz;
                if (r2 * 0.9999999999 > blmax) bad++;
                if (r2 * 1.0000000001 < blmin) bad++;
            }
        }
    return bad;
}

/* checkover.f:21-131 with the violations split by kind, then
 * check_nc_int.f:25-354 over each bead's up list (code-15 partners):
 * out = {overlaps, bonds too long, bonds too short, first i, first j (1-based,
 * 0 if none), boundbad, unboundbad, n_ss, m_ss}.  sv (the reference layout)
 * replaces the state's positions and velocities when not NULL. */
void orc_check(orc_sys* S, const double* sv, int64_t out[9]) {
    double* keep = S->sv;
    if (sv) S->sv = (double*)sv;
    const params_t* P = &S->P;
    int N = S->N;
    for (int k = 0; k < 9; ++k) out[k] = 0;
    for (int i = 1; i <= N - 1; ++i)
        for (int j = i + 1; j <= N; ++j) {
            int evcode = EV(S, j, i), bad = 0;
            pairgeo g;
            if (evcode == 1 || evcode >= 15) {
                double pe = P->ev_param[1][evcode];
                double sigsq = P->sigma_sq[S->identity[i]][S->identity[j]] * (pe * pe);
                if (evcode >= 22 && evcode <= 26) {
                    int k = S->identity[i] > S->identity[j] ? S->identity[i] : S->identity[j];
                    double q = P->sqz610[evcode - 21][k];
                    sigsq = sigsq * (q * q);
                }
                geo(S, i, j, S->tfalse, &g);
                double r2 = g.rx * g.rx + g.ry * g.ry + g.rz * g.rz;
                r2 = r2 * 1.0000000001;
                if (r2 <= sigsq) {
                    out[0]++;
                    bad = 1;
                }
            } else if (evcode >= 4 && evcode <= 12 && !S->chaptype2) { /* checkover.f:67: chaptype 1 only */
                double blmin, blmax;
                bond_window(S, i, evcode, &blmin, &blmax);
                blmin = blmin * blmin;
                blmax = blmax * blmax;
                geo(S, i, j, S->tfalse, &g);
                double r2 = g.rx * g.rx + g.ry * g.ry + g.rz * g.rz;
                if (r2 * 0.9999999999 > blmax) {
                    out[1]++;
                    bad = 1;
                }
                if (r2 * 1.0000000001 < blmin) {
                    out[2]++;
                    bad = 1;
                }
            }
            if (bad && out[3] == 0) {
                out[3] = i;
                out[4] = j;
            }
        }
    for (int i = 1; i <= N; ++i) {
        int ks = (i - 1) * MAXNBS;
        for (int k = 0; k < S->na_npt[i]; ++k) {
            int j = S->nb[ks + k];
            if (EV(S, j, i) != 15) continue;
            int term = terminal(S, i) || terminal(S, j);
            double rating;
            if (S->bptnr[i] == j) {
                rating = term ? 10.0 : (S->identity[i] < S->identity[j] ? repuls_check(S, i, j) : repuls_check(S, j, i));
                if (rating > 10.0) out[5]++;
                if (!term && XR(S, i, 4) == j) out[7]++;
            } else {
                pairgeo g;
                geo(S, i, j, S->tfalse, &g);
                double rijsq = g.rx * g.rx + g.ry * g.ry + g.rz * g.rz;
                double diff = rijsq - P->welldia_sq[S->identity[i]][S->identity[j]];
                if (diff < 0.0) {
                    rating = term ? 11.0 : (S->identity[i] < S->identity[j] ? repuls_check(S, i, j) : repuls_check(S, j, i));
                    if (rating <= 10.0) out[6]++;
                    if (!term && XR(S, i, 4) == j) out[7]++;
                }
            }
            if (XR(S, i, 4) == j) out[8]++;
        }
    }
    S->sv = keep;
}

/* ------------------------------------------------------------------ */
/* Construction: main.F90:126-424 with -Drunr -Dcanon                   */
/* ------------------------------------------------------------------ */
orc_sys* orc_create(const char* param_dir, int nchains, const int32_t* chain_len,
                    const int32_t* species, const int32_t* restype, const int32_t* hp,
                    double boxl_angstrom, double tstar, uint64_t seed, int canon,
                    const double* sv, const int32_t* bptnr, char* err, int errlen) {
    return orc_create_ex(param_dir, nchains, chain_len, species, restype, hp, boxl_angstrom, tstar, seed, canon, 0,
                         sv, bptnr, err, errlen);
}

orc_sys* orc_create_ex(const char* param_dir, int nchains, const int32_t* chain_len,
                       const int32_t* species, const int32_t* restype, const int32_t* hp,
                       double boxl_angstrom, double tstar, uint64_t seed, int canon, int flags,
                       const double* sv, const int32_t* bptnr, char* err, int errlen) {
    orc_sys* S = (orc_sys*)calloc(1, sizeof(orc_sys));
    char dummy[8];
    if (!err) {
        err = dummy;
        errlen = sizeof dummy;
    }
    if (!S) return NULL;
    if (load_params(&S->P, param_dir, err, errlen)) {
        free(S);
        return NULL;
    }
    S->nchains = nchains;
    int N = 0, nres = 0;
    S->chain_start = (int*)calloc((size_t)nchains, sizeof(int));
    S->chain_len = (int*)calloc((size_t)nchains, sizeof(int));
    S->species = (int*)calloc((size_t)nchains, sizeof(int));
    for (int c = 0; c < nchains; ++c) {
        S->chain_start[c] = N;
        S->chain_len[c] = chain_len[c];
        S->species[c] = species ? species[c] : c;
        N += 4 * chain_len[c];
        nres += chain_len[c];
    }
    S->nfirst = 0;
    while (S->nfirst < nchains && S->species[S->nfirst] == S->species[0]) S->nfirst++;
    S->N = N;
    S->canon = canon;
    S->no_hbs = (flags & ORC_NO_HBS) != 0;
    S->no_h = (flags & ORC_NO_H) != 0;
    S->chaptype2 = (flags & ORC_CHAPTYPE2) != 0;
    S->seed = seed;
    size_t n1 = (size_t)N + 4;
    S->chnnum = (int*)calloc(n1, sizeof(int));
    S->loc = (int*)calloc(n1, sizeof(int));
    S->cl = (int*)calloc(n1, sizeof(int));
    S->res = (int*)calloc(n1, sizeof(int));
    S->hp_bead = (int*)calloc(n1, sizeof(int));
    S->identity = (int*)calloc(n1, sizeof(int));
    S->bptnr = (int*)calloc(n1, sizeof(int));
    S->extra = (int*)calloc(4 * n1, sizeof(int));
    S->sv = (double*)calloc(6 * n1, sizeof(double));
    S->bm = (double*)calloc(n1, sizeof(double));
    S->old_r = (double*)calloc(3 * n1, sizeof(double));
    S->ev = (signed char*)malloc((size_t)N * (size_t)N);
    S->bdln = (double*)calloc((size_t)nres + 1, sizeof(double));
    S->bl_rn = (double*)calloc((size_t)nres + 1, sizeof(double));
    S->bl_rc = (double*)calloc((size_t)nres + 1, sizeof(double));
    S->del_bdln = (double*)calloc((size_t)nres + 1, sizeof(double));
    S->del_blrn = (double*)calloc((size_t)nres + 1, sizeof(double));
    S->del_blrc = (double*)calloc((size_t)nres + 1, sizeof(double));
    S->nb = (int*)calloc((size_t)N * MAXNBS, sizeof(int));
    S->dnnab = (int*)calloc((size_t)N * MAXNBS, sizeof(int));
    S->na_npt = (int*)calloc(n1, sizeof(int));
    S->nnabdn = (int*)calloc(n1, sizeof(int));
    S->clinks = (int*)calloc(n1, sizeof(int));
    S->tim = (double*)calloc(n1, sizeof(double));
    S->nptnr = (int*)calloc(n1, sizeof(int));
    S->coltype = (int*)calloc(n1, sizeof(int));
    S->tlinks = (int*)calloc(n1, sizeof(int));
    S->tlinks2 = (int*)calloc(n1, sizeof(int));
    if (!S->ev || !S->nb || !S->dnnab) {
        snprintf(err, errlen, "out of memory");
        orc_destroy(S);
        return NULL;
    }
    params_t* P = &S->P;
    /* topology (inputinfo.f:193-262) */
    int r = 0;
    for (int c = 0; c < nchains; ++c) {
        int s0 = S->chain_start[c], L = chain_len[c];
        for (int k = 1; k <= 4 * L; ++k) {
            int b = s0 + k;
            S->chnnum[b] = c + 1;
            S->loc[b] = k;
            S->cl[b] = L;
            int rr = (k - 1) % L + 1;
            S->res[b] = r + rr;
            int idn = k <= L ? 2 : (k <= 2 * L ? 1 : (k <= 3 * L ? 4 : restype[r + rr - 1]));
            S->identity[b] = idn;
            S->hp_bead[b] = k > 3 * L ? hp[r + rr - 1] : 0;
        }
        r += L;
    }
    for (int b = 1; b <= N; ++b) {
        if (S->identity[b] < 1 || S->identity[b] > 28) {
            snprintf(err, errlen, "bad residue identity");
            orc_destroy(S);
            return NULL;
        }
        S->bm[b] = P->bmass[S->identity[b]]; /* inputinfo.f:400-403 */
    }
    /* per-residue bond lengths (inputinfo.f:300-349) in Angstrom */
    for (int q = 1; q <= nres; ++q) {
        int iii = restype[q - 1] - 8;
        S->bdln[q] = P->drca[iii];
        S->bl_rn[q] = P->drnh[iii];
        S->bl_rc[q] = P->drco[iii];
        S->del_bdln[q] = P->del_rca[iii];
        S->del_blrn[q] = P->del_rnh[iii];
        S->del_blrc[q] = P->del_rco[iii];
    }
    scale_down(P, boxl_angstrom);
    for (int q = 1; q <= nres; ++q) {
        S->bdln[q] = S->bdln[q] / P->boxl_orig;
        S->bl_rn[q] = S->bl_rn[q] / P->boxl_orig;
        S->bl_rc[q] = S->bl_rc[q] / P->boxl_orig;
    }
    S->boxl = 1.0;
    S->half = S->boxl / 2.0;
    make_ev_param(P);
    make_code(S);
    /* main.F90:126-160 */
    S->setemp = tstar * 12.0;
    S->t_fact = 0.00005;
    S->n_forced = 150.0;
    S->interval = S->t_fact / sqrt(S->setemp);
    S->interval_max = S->n_forced * S->interval;
    S->sortsize = S->interval_max / (double)NUMBIN;
    S->avegtime = 0.00005 / sqrt(S->setemp);
    S->t = 0.0;
    S->tfalse = 0.0;
    S->nbin = 1;
    S->tbin_off = 0.0;
    /* main.F90:205-234 (sv from .config/.lastvel, wrapped with dnint) */
    for (int k = 1; k <= N; ++k) {
        for (int c = 1; c <= 6; ++c) SV(S, c, k) = sv[(size_t)(k - 1) * 6 + (c - 1)];
        SV(S, 1, k) = SV(S, 1, k) - dnint(SV(S, 1, k));
        SV(S, 2, k) = SV(S, 2, k) - dnint(SV(S, 2, k));
        SV(S, 3, k) = SV(S, 3, k) - dnint(SV(S, 3, k));
        SV(S, 1, k) = SV(S, 1, k) - dnint(SV(S, 1, k));
        SV(S, 2, k) = SV(S, 2, k) - dnint(SV(S, 2, k));
        SV(S, 3, k) = SV(S, 3, k) - dnint(SV(S, 3, k));
        S->old_r[3 * k - 3] = SV(S, 1, k);
        S->old_r[3 * k - 2] = SV(S, 2, k);
        S->old_r[3 * k - 1] = SV(S, 3, k);
        S->tim[k] = S->interval_max + LTSTEP;
        S->coltype[k] = -1;
        S->nptnr[k] = -1;
        S->bptnr[k] = 0;
    }
    for (int k = N + 1; k <= N + 3; ++k) {
        S->nptnr[k] = -2;
        S->coltype[k] = -2;
    }
    /* runr restore of H-bonds and shoulders (main.F90:240-321) */
    if (bptnr)
        for (int k = 1; k <= N; ++k) S->bptnr[k] = bptnr[k - 1];
    for (int k = 1; k <= N - 1; ++k)
        for (int kj = k + 1; kj <= N; ++kj) {
            if (S->identity[k] + S->identity[kj] == 5 && EV(S, k, kj) == 15) {
                pairgeo g;
                double rx = SV(S, 1, k) - SV(S, 1, kj), ry = SV(S, 2, k) - SV(S, 2, kj),
                       rz = SV(S, 3, k) - SV(S, 3, kj);
                rx = rx - dnint(rx);
                ry = ry - dnint(ry);
                rz = rz - dnint(rz);
                (void)g;
                double r2 = rx * rx + ry * ry + rz * rz;
                double diff = r2 - P->welldia_sq[S->identity[k]][S->identity[kj]];
                if (diff < 0.0 && !terminal(S, k) && !terminal(S, kj)) {
                    if (S->identity[k] == 1)
                        repuls_add(S, k, kj);
                    else
                        repuls_add(S, kj, k);
                }
            }
            if (kj == S->bptnr[k]) {
                if (S->identity[k] == 1) {
                    S->identity[k] = 5;
                    S->identity[kj] = 8;
                } else {
                    S->identity[k] = 8;
                    S->identity[kj] = 5;
                }
            }
        }
    /* lists (main.F90:389-405) */
    nbor_setup(S);
    S->num_cell = (int)(S->boxl / (S->sig_max_all * RL_CONST) * N_WRAP) + 2 * N_WRAP;
    int nc = S->num_cell;
    if (nc - 2 * N_WRAP < 2 * N_WRAP + 1) {
        snprintf(err, errlen, "box too small for the n_wrap=2 stencil");
        orc_destroy(S);
        return NULL;
    }
    S->cell = (int*)calloc((size_t)nc * nc * nc + 2, sizeof(int));
    S->wrap_map = (int*)calloc((size_t)nc * nc * nc + 2, sizeof(int));
    S->width = S->boxl / (double)(nc - 2 * N_WRAP);
    cell_link(S);
    nbor(S);
    long nbrsum = 0;
    for (int k = 1; k <= N - 1; ++k) nbrsum += S->na_npt[k];
    S->avg_nbrs = (double)nbrsum / (double)N;
    /* pseudo-event seeds (main.F90:407-423) */
    if (canon) {
        double tgho = 0.0;
        do {
            tgho = drand(S, 0);
        } while (tgho < 1.e-18 || tgho == 1.0);
        S->tim[N + 1] = -1.0 * orc_log(tgho) * S->avegtime * .0000001;
    } else
        S->tim[N + 1] = 1000000000.0;
    S->tim[N + 2] = S->interval;
    S->tim[N + 3] = 3.3 / (sqrt(S->setemp)) + 5;
    events(S);
    if (S->err) {
        snprintf(err, errlen, "setup self-check failed (%d)", S->err);
        orc_destroy(S);
        return NULL;
    }
    return S;
}

void orc_destroy(orc_sys* S) {
    if (!S) return;
    free(S->chain_start);
    free(S->chain_len);
    free(S->species);
    free(S->chnnum);
    free(S->loc);
    free(S->cl);
    free(S->res);
    free(S->hp_bead);
    free(S->identity);
    free(S->bptnr);
    free(S->extra);
    free(S->sv);
    free(S->bm);
    free(S->old_r);
    free(S->ev);
    free(S->bdln);
    free(S->bl_rn);
    free(S->bl_rc);
    free(S->del_bdln);
    free(S->del_blrn);
    free(S->del_blrc);
    free(S->nb);
    free(S->dnnab);
    free(S->na_npt);
    free(S->nnabdn);
    free(S->clinks);
    free(S->cell);
    free(S->wrap_map);
    free(S->tim);
    free(S->nptnr);
    free(S->coltype);
    free(S->tlinks);
    free(S->tlinks2);
    free(S);
}

int32_t orc_nbeads(const orc_sys* S) { return S->N; }

void orc_get_state(const orc_sys* S, double* sv, int32_t* identity, int32_t* bptnr,
                   int32_t* extra_repuls, double* tim, int32_t* nptnr, int32_t* coltype) {
    int N = S->N;
    if (sv) memcpy(sv, S->sv, sizeof(double) * 6 * (size_t)N);
    for (int k = 1; k <= N; ++k) {
        if (identity) identity[k - 1] = S->identity[k];
        if (bptnr) bptnr[k - 1] = S->bptnr[k];
        if (extra_repuls)
            for (int m = 1; m <= 4; ++m) extra_repuls[(k - 1) * 4 + m - 1] = XR(S, k, m);
    }
    for (int k = 1; k <= N + 3; ++k) {
        if (tim) tim[k - 1] = S->tim[k];
        if (nptnr) nptnr[k - 1] = S->nptnr[k];
        if (coltype) coltype[k - 1] = S->coltype[k];
    }
}

void orc_stats_get(const orc_sys* S, orc_stats* st) {
    memset(st, 0, sizeof *st);
    st->coll = S->coll;
    st->numghosts = S->numghosts;
    st->nupdates = S->nupdates;
    st->nforced = S->nforced;
    memcpy(st->nevents, S->nevents, sizeof S->nevents);
    st->t = S->t;
    st->tfalse = S->tfalse;
    st->interval = S->interval;
    st->interval_max = S->interval_max;
    st->t_fact = S->t_fact;
    st->n_forced = S->n_forced;
    st->sortsize = S->sortsize;
    st->tbin_off = S->tbin_off;
    st->nbin = S->nbin;
    st->num_cell = S->num_cell;
    st->sig_max_all = S->sig_max_all;
    st->hdelr = S->hdelr;
    st->avg_nbrs = S->avg_nbrs;
    st->rng_ctr[0] = (int64_t)S->rng_ctr[0];
    st->rng_ctr[1] = (int64_t)S->rng_ctr[1];
}

int32_t orc_ev_code(const orc_sys* S, int32_t a, int32_t b) { return EV(S, a, b); }
int32_t orc_nb_up(const orc_sys* S, int32_t i, int32_t* out, int32_t cap) {
    int n = S->na_npt[i];
    for (int k = 0; k < n && k < cap; ++k) out[k] = S->nb[(i - 1) * MAXNBS + k];
    return n;
}
int32_t orc_nb_dn(const orc_sys* S, int32_t i, int32_t* out, int32_t cap) {
    int n = S->nnabdn[i];
    for (int k = 0; k < n && k < cap; ++k) out[k] = S->dnnab[(i - 1) * MAXNBS + k];
    return n;
}
double orc_predict(const orc_sys* S, int32_t i, int32_t j, int32_t* type) {
    double tij = 1000000000.0;
    int ty = 0;
    predict_code(S, i, j, EV(S, j, i), &tij, &ty);
    if (type) *type = ty;
    return tij;
}
double orc_table(const orc_sys* S, int which, int k, int kk) {
    const params_t* P = &S->P;
    switch (which) {
    case 0: return P->sigma_sq[k][kk];
    case 1: return P->welldia_sq[k][kk];
    case 2: return P->ep_sqrt[k][kk];
    case 3: return P->shlddia_sq[k][kk];
    case 4: return P->sigma_2b[k][kk];
    case 5: return P->bmass[k];
    case 6: return P->ev_param[k][kk];
    case 7: return P->sqz610[k][kk];
    case 8: return S->rlsq[k];
    default: return 0.0;
    }
}
double orc_bond(const orc_sys* S, int which, int q) {
    switch (which) {
    case 0: return (1.0 - S->del_bdln[q]) * S->bdln[q];
    case 1: return (1.0 + S->del_bdln[q]) * S->bdln[q];
    case 2: return (1.0 - S->del_blrn[q]) * S->bl_rn[q];
    case 3: return (1.0 + S->del_blrn[q]) * S->bl_rn[q];
    case 4: return (1.0 - S->del_blrc[q]) * S->bl_rc[q];
    case 5: return (1.0 + S->del_blrc[q]) * S->bl_rc[q];
    default: return 0.0;
    }
}
void orc_set_max_log_obs(orc_sys* S, int enable) { S->log_obs = enable; }
