/*
 * pdmd.h -- C ABI of the MI355X PRIME20 discontinuous-MD engine.
 *
 * Drop-in boundary for the reference's event-loop path
 * (parallel-dmd-PRIME20/code/main.F90:484-1258 and the subroutines it
 * calls).  Plain pointers and sizes only; every entry point returns an
 * int status (0 = ok, < 0 = error; pdmd_last_error() gives the text).
 * Bead, partner and event indices are 1-based as in the reference; arrays
 * use the reference layouts (sv(6,N) column-major = x,y,z,vx,vy,vz per
 * bead, positions in box units).  A context is not re-entrant: one host
 * thread per context.
 */
#ifndef PDMD_H
#define PDMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdmd_ctx pdmd_ctx;

/* One PRIME20 system, as the reference reads it with -Drunr
 * (inputinfo.f:28-411, main.F90:240-321): chains of chain_len[c] residues,
 * per-residue side-chain identity restype (9..28, the G..V order of
 * parameters/mass.data) and hp flag, box length in Angstrom, reduced
 * temperature T* (stdin line 1 of temp_XXX), the .config/.lastvel state
 * sv and the .bptnr table (1-based partner, 0 = none; NULL = no H-bonds).
 * species[c] groups chains that share one sequence, the reference's
 * nop1/nop2 blocks (NULL = every chain its own species).  seed keys the
 * project RNG that replaces IFPORT drandm; canon != 0 enables Andersen
 * ghost collisions (-Dcanon). */
#define PDMD_NO_HBS 1 /* no N-C hydrogen-bond wells: the reference's code 15 pairs stay hard cores */
#define PDMD_NO_H 2   /* build flag -Dno_h (inputinfo.f:270-280): no code-16 wells between chains; same-chain ones stay */
/* build flag -Dchaptype=2: the chains after the first species block (the
 * reference's nop2 block) carry no bonds, local cores or intra-chain wells
 * (make_code.f:349-484, inputinfo.f:120-160) and no side-chain wells with
 * each other (make_code.f:519-547); checkover.f skips bond windows (:67).
 * Their beads start apart, as unbonded hard cores must. */
#define PDMD_CHAPTYPE2 4

typedef struct {
    int32_t nchains;
    const int32_t* chain_len;
    const int32_t* species;
    const int32_t* restype;
    const int32_t* hp;
    double boxl;
    double tstar;
    uint64_t seed;
    int32_t canon;
    int32_t list_capacity; /* per-bead up/down list slots, at most 64; 0 = default (64) */
    int32_t flags;         /* PDMD_NO_HBS (-Dno_hbs, make_code.f:114-170,192-208,348-365) | PDMD_NO_H (-Dno_h)
                              | PDMD_CHAPTYPE2 (-Dchaptype=2) */
    int32_t pad;
    const double* sv;
    const int32_t* bptnr;
} pdmd_system_desc;

typedef struct {
    int64_t coll;        /* committed calendar entries (main.F90:639) */
    int64_t pair_events; /* bead-pair events only */
    int64_t numghosts, nupdates, nforced;
    int64_t nevents[31]; /* tally by resolved coltype (main.F90:926) */
    double t, tfalse, interval, interval_max;
    int32_t status, max_list;
    int32_t nbeads, pad;
    double avg_up, avg_dn; /* mean up/down neighbour-list lengths (K_up, K_dn) */
    int64_t phase_cycles[40]; /* device cycles and counts per engine phase; zero unless the
                                 library is built with -DPDMD_PHASE_TIMERS */
    double device_ms;  /* wall-clock device time this system has spent in event-loop launches */
    double rebuild_ms; /* ... of which in neighbour-list rebuilds (nbor.f + events.f at a sync) */
    int64_t rebuilds;  /* the rebuilds rebuild_ms covers */
    int32_t waves;       /* pair events per round this system is limited to (0: the context's setting) */
    int32_t cell_factor; /* fine cells per coarse cell of its rebuild search */
    int32_t cells_axis;  /* fine cells per box edge (main.F90:390, interior cells); the rebuild
                            search grid has (cells_axis / cell_factor)^3 cells */
    int32_t pad2;
} pdmd_stats_t;

typedef struct {
    int64_t coll;
    int32_t i, j, type, pad;
    double t; /* tfalse of the event */
} pdmd_event_t;

typedef struct {
    double ered, tred, sumvel; /* energy.f:25-101 */
    int32_t hb_alpha, hb_ii, hb_ij, pad;
    double ehh_ii, ehh_ij;
    double rg_avg;  /* radgyr.f:18-135 */
    double e2e_avg; /* end2end.f:14-44 */
} pdmd_obs_t;

/* Library version string. */
const char* pdmd_version(void);
/* Text of the last error on this thread. */
const char* pdmd_last_error(void);

/* Context on the current HIP device; param_dir holds the reference's
 * parameter tables (protein.data, beadwell_ha55a.data, rcarnrco.data,
 * sqz6to10.data, mass.data, ep19p_ha55a_weakhp.data).
 * Replaces inputinfo.f:141-398 table reading. */
int pdmd_ctx_create(const char* param_dir, pdmd_ctx** out);
int pdmd_ctx_destroy(pdmd_ctx* ctx);

/* Upload a system and run the reference's start-up on the device:
 * scale_down.f, make_code.f (as a compact topology), the -Drunr H-bond
 * restore (main.F90:240-321), nbor_setup.f, cell_link.f/nbor.f and
 * events.f (main.F90:389-424).  *sys_id receives the system handle. */
int pdmd_system_add(pdmd_ctx* ctx, const pdmd_system_desc* desc, int32_t* sys_id);
int pdmd_system_remove(pdmd_ctx* ctx, int32_t sys_id);
int pdmd_nbeads(pdmd_ctx* ctx, int32_t sys_id);
/* Device bytes a system of this description would occupy: every buffer
 * pdmd_system_add allocates for it (lists at the description's capacity,
 * the coarse cell grid of its box and cutoffs), each rounded up to 4 KiB;
 * the event log of pdmd_log_enable comes on top.  Host-only.  This form
 * reads the parameter tables that ship beside the library (parameters/);
 * pdmd_system_bytes_ctx uses the context's tables.  The serving loop's
 * memory-budget admission uses the latter.  No counterpart in the reference,
 * whose arrays are sized at compile time (code/header.f:11-40). */
int64_t pdmd_system_bytes(const pdmd_system_desc* desc);
int64_t pdmd_system_bytes_ctx(pdmd_ctx* ctx, const pdmd_system_desc* desc);

/* Commit nevents[k] more calendar entries of each listed system, all
 * systems concurrently in one launch (one workgroup per system, of the
 * wavefronts set by pdmd_set_waves / pdmd_autotune_waves), in the
 * reference's serial event order.  Replaces the event loop
 * main.F90:484-1258 (dispatch, commit, partial_events, calendar,
 * ghost/sync/output pseudo-events).  Blocks until done. */
int pdmd_advance(pdmd_ctx* ctx, int32_t nsys, const int32_t* sys_ids, const int64_t* nevents);
/* Asynchronous form for request loops: launch and return; pdmd_wait()
 * blocks until the launch has finished. */
int pdmd_advance_async(pdmd_ctx* ctx, int32_t nsys, const int32_t* sys_ids, const int64_t* nevents);
int pdmd_wait(pdmd_ctx* ctx);
/* The same with a device wall-clock budget: each system stops between
 * events once budget_ms have passed since the launch began (the first of
 * its workgroups started), or at
 * max_nevents[k] more entries, whichever comes first (pdmd_stats gives the
 * entries committed).  For request loops over systems of uneven speed. */
int pdmd_advance_budget_async(pdmd_ctx* ctx, int32_t nsys, const int32_t* sys_ids, const int64_t* max_nevents,
                              double budget_ms);
/* Blocking advance that stops each system right after its next output
 * pseudo-event (entry N+3, main.F90:1191-1246), or at max_nevents[k] more
 * entries: the points where the reference's driver writes observables
 * (energy.f, radgyr.f, end2end.f) and configurations (config.f). */
int pdmd_advance_to_output(pdmd_ctx* ctx, int32_t nsys, const int32_t* sys_ids, const int64_t* max_nevents);
/* Wavefronts per system for the context's later advances (1 to 8; larger
 * values are lowered to what the device's LDS holds).  With W > 1 every
 * round of a system's event loop pops up to W calendar entries and runs
 * the pair events among them side by side, one per wavefront; the round
 * commits the prefix that the serial order would have run identically
 * (disjoint footprints, no new calendar entry ahead of a later popped one),
 * so results are the same for every W.  No counterpart in the reference,
 * whose master/worker speculation (main.F90:484-1258) this replaces.
 * Default 1. */
int pdmd_set_waves(pdmd_ctx* ctx, int32_t waves_per_system);
/* Neighbour lists and every calendar entry rebuilt now (nbor.f:23-137 +
 * events.f:11-127, the reference's nbor() and events() calls at
 * main.F90:1170-1177): the listed systems move their time origin to the
 * current time as the sync pseudo-event does (main.F90:1126-1187: positions
 * advanced, times shifted, calendar rescaled) and take its update branch
 * without the displacement test; the pending sync entry keeps its time.
 * pdmd_system_add does the same at start-up (main.F90:389-424).  A forced
 * update changes the later event sequence as an extra update would in the
 * reference. */
int pdmd_rebuild(pdmd_ctx* ctx, int32_t nsys, const int32_t* sys_ids);
/* Launch mode of pdmd_advance*: with on != 0 the event-loop launch is
 * captured in a HIP graph once per launch shape (systems count, largest
 * system, budget, flags, wavefronts) and replayed while the shape stays the
 * same; the system list and event targets travel in device buffers, so
 * every step of a serving loop with an unchanged resident set is one graph
 * replay.  Results do not depend on the mode.  No counterpart in the
 * reference (its master drives MPI workers instead); the serving loop turns
 * it on.  Default off. */
int pdmd_set_graph_launch(pdmd_ctx* ctx, int32_t on);
/* Autotuner of the wavefronts per system: advances the listed systems by
 * probe_events entries each at 1, 2, 4 and 8 wavefronts (those that fit in
 * LDS), one launch each, and keeps the fastest for later advances.  The
 * probe entries are committed like any others (the event sequence does not
 * depend on the setting).  waves_out gets the choice, rates_out[4] (may be
 * NULL) the committed entries per second of each trial (0 if skipped). */
int pdmd_autotune_waves(pdmd_ctx* ctx, int32_t nsys, const int32_t* sys_ids, int64_t probe_events,
                        int32_t* waves_out, double* rates_out);
/* Per-system tuning, no counterpart in the reference (results do not depend
 * on it): waves caps the pair events one round of this system runs side by
 * side (1..8, 0 = the context's pdmd_set_waves); cell_factor (2..4, 0 =
 * keep) is the coarse cell of its list rebuild's search in fine cells (the
 * fine cells of main.F90:390 fix the list membership; the coarse grid only
 * sets how many candidates each bead scans). */
int pdmd_set_system_tuning(pdmd_ctx* ctx, int32_t sys_id, int32_t waves, int32_t cell_factor);
/* Autotuner of both per system: advances the listed systems by
 * probe_events entries each at 1, 2, 4 and the most wavefronts that fit
 * (one launch each, the per-system device clock timing every system
 * separately), then at cell factors 2, 3 and 4 (systems of more than 256
 * beads; smaller ones search all pairs), and keeps for each system its
 * fastest wavefront count and its cheapest rebuild; the context's launch
 * width becomes the largest chosen.  Probe entries are committed like any
 * others.  waves_out / cell_factor_out (nsys each) get the choices,
 * rates_out (nsys x 4, may be NULL) each system's entries per second at
 * the four wavefront trials (0 if skipped). */
int pdmd_autotune_systems(pdmd_ctx* ctx, int32_t nsys, const int32_t* sys_ids, int64_t probe_events,
                          int32_t* waves_out, int32_t* cell_factor_out, double* rates_out);
/* Device time of the most recent pdmd_advance* launch (HIP events on the
 * context's stream), in milliseconds. */
int pdmd_last_launch_ms(pdmd_ctx* ctx, float* ms);

/* State in the reference layouts (any pointer may be NULL): sv(6,N),
 * identity(N), bptnr(N), extra_repuls(N,4) as extra[4*(k-1)+m-1],
 * tim/nptnr/coltype over N+3 calendar slots (header.f:24-46). */
int pdmd_download(pdmd_ctx* ctx, int32_t sys_id, double* sv, int32_t* identity, int32_t* bptnr,
                  int32_t* extra_repuls, double* tim, int32_t* nptnr, int32_t* coltype);
int pdmd_stats(pdmd_ctx* ctx, int32_t sys_id, pdmd_stats_t* out);

/* Event log of the next `capacity` committed events (coll, i, j, resolved
 * type, event time); for parity checks against the CPU oracle. */
int pdmd_log_enable(pdmd_ctx* ctx, int32_t sys_id, int32_t capacity);
int pdmd_log_read(pdmd_ctx* ctx, int32_t sys_id, int32_t max, pdmd_event_t* out, int32_t* n);

/* Test hook: the neighbour lists as the last rebuild left them (nbor.f's
 * nb / na_npt and dnnab / nnabdn, header.f:21): n_up[k] and n_dn[k] the
 * row lengths of bead k+1, up[k*cap + m] / dn[k*cap + m] its 1-based
 * partners (rows in nbor.f's order, as the reference builds them); cap as
 * pdmd_list_capacity returns.  Any pointer may be NULL. */
int pdmd_list_capacity(pdmd_ctx* ctx, int32_t sys_id);
int pdmd_lists(pdmd_ctx* ctx, int32_t sys_id, int32_t* n_up, int32_t* up, int32_t* n_dn, int32_t* dn);

/* Test hook replacing core.f/bond.f/sqwel.f/sqshlder.f/nc_sqwel.f called
 * as events.f:28-47 does: tij (1e9 when none) and type for pairs (i, j),
 * i < j, at the system's current tfalse. */
int pdmd_predict_pairs(pdmd_ctx* ctx, int32_t sys_id, int32_t n, const int32_t* i, const int32_t* j, double* tij,
                       int32_t* type);

/* Pair codes ev_code(a,b) as make_code.f:18-566 builds them for the system
 * described (before any H-bond restore), for n 1-based bead pairs; host
 * only, needs no device.  The device derives the same codes from compact
 * per-bead topology words instead of the N x N int8 matrix. */
int pdmd_pair_codes(const pdmd_system_desc* desc, int32_t n, const int32_t* a, const int32_t* b, int32_t* code);

/* Observables of the output pseudo-event (energy.f:25-101, radgyr.f:18-135,
 * end2end.f:14-44; main.F90:1191-1246), reduced on the device over the
 * resident state at the current tfalse: side-chain well energies over the
 * neighbour lists instead of energy.f's all-pairs loop. */
int pdmd_observables(pdmd_ctx* ctx, int32_t sys_id, pdmd_obs_t* out);
/* The same observables, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): no counterpart in the reference.  Identical results; nothing
 * is allocated per call -- the chain tables and results pass through one
 * fixed 33 KB device array per GPU, located at context creation, so systems
 * of at most 4096 chains (every system pdmd_system_add accepts).  Calls
 * that overlap on one GPU, from any context, race on that array. */
int pdmd_observables_noalloc(pdmd_ctx* ctx, int32_t sys_id, pdmd_obs_t* out);

typedef struct {
    int64_t overlaps;    /* checkover.f:21-131: pairs with ev_code 1 or >= 15 closer than sigma */
    int64_t bonds_long;  /* checkover.f: bonded pairs (codes 4-12) longer than their window */
    int64_t bonds_short; /* ... and shorter */
    int32_t first_i, first_j; /* 1-based first offending pair in (i, j) order; 0 if none */
    int64_t boundbad;    /* check_nc_int.f:25-354: H-bonded N-C pairs whose angle test fails */
    int64_t unboundbad;  /* ... N-C pairs inside the well, angle test passing, not bonded */
    int64_t ss_counted;  /* ... n_ss: non-terminal such pairs with their shoulder partner set */
    int64_t ss_total;    /* ... m_ss: listed code-15 pairs with extra_repuls(i,4) = j; the
                            reference stops when it differs from n_ss */
} pdmd_check_t;
/* Debug checks of one system on the device: checkover.f (all pairs) and
 * check_nc_int.f (listed code-15 pairs), at the system's current tfalse, on
 * the resident state or, when sv is not NULL, on the caller's sv(6,N) with
 * the system's topology, identities and codes (main.F90:347,425,1194-1204,
 * 1322 call them at the start, at output events under -Ddebugging, at the
 * end). */
int pdmd_check(pdmd_ctx* ctx, int32_t sys_id, const double* sv, pdmd_check_t* out);

/* Seeded synthetic input (no reference counterpart: the shipped generator
 * genconfig/gen_config_random-SQZ.f90 is template-based and IFPORT-seeded):
 * random sequences, self-avoiding coils satisfying every bond window and
 * core, Maxwell-Boltzmann velocities at T*, box from rho* = N sigma_1^3/V.
 * restype/hp hold sum(chain_len) entries, sv 6*N. */
int pdmd_generate(const char* param_dir, int32_t nchains, const int32_t* chain_len, double density, double tstar,
                  uint64_t seed, int32_t* restype, int32_t* hp, double* sv, double* boxl);
/* The same for a given sequence: restype_in holds sum(chain_len) side-chain
 * identities (9..28), copied to restype; NULL draws it as pdmd_generate does.
 * Chains that share a sequence can form one species (the reference's
 * nop1/nop2 blocks, inputinfo.f:230-262). */
int pdmd_generate_seq(const char* param_dir, int32_t nchains, const int32_t* chain_len, const int32_t* restype_in,
                      double density, double tstar, uint64_t seed, int32_t* restype, int32_t* hp, double* sv,
                      double* boxl);

/* ---- serving: continuous batching of trajectory requests (no reference
 * counterpart; the reference runs one system per MPI job).  Requests queue
 * on the host and are admitted as device memory and launch slots free up;
 * each step waits for the launch in flight, retires finished requests,
 * admits queued ones (first fit in arrival order) and launches the next
 * quantum of at most `quantum` calendar entries for every resident request,
 * cut short after slice_ms of device time when slice_ms > 0 (so systems of
 * uneven speed finish a launch together).  Times are seconds on the
 * server's clock (steady clock since creation). */
typedef struct pdmd_server pdmd_server;
typedef struct {
    int32_t state;  /* 0 queued, 1 resident, 2 done, 3 failed */
    int32_t status; /* engine status of a failed request */
    int64_t ncoll, done;
    int32_t nbeads, pad;
    double t_submit, t_admit, t_done;
    pdmd_stats_t stats; /* final statistics (done or failed) */
} pdmd_request_info_t;

int pdmd_server_create(pdmd_ctx* ctx, int64_t mem_budget_bytes, int64_t quantum, double slice_ms, int32_t max_resident,
                       pdmd_server** out);
int pdmd_server_destroy(pdmd_server* srv);
/* Queue a request (the description's arrays are copied). */
int pdmd_server_submit(pdmd_server* srv, const pdmd_system_desc* desc, int64_t ncoll, int64_t* req_id);
/* One scheduling step; *ncompleted = requests finished by the launch it retired. */
int pdmd_server_step(pdmd_server* srv, int32_t* ncompleted);
int pdmd_server_pending(pdmd_server* srv, int32_t* queued, int32_t* resident, int64_t* committed);
/* Record of a request; sv (6*nbeads) receives its final state once done. */
int pdmd_server_result(pdmd_server* srv, int64_t req_id, pdmd_request_info_t* info, double* sv);
int pdmd_server_release(pdmd_server* srv, int64_t req_id);

#ifdef __cplusplus
}
#endif
#endif
