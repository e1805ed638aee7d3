/*
 * neci_gpu.h -- C ABI of the MI355X FCIQMC iteration engine.
 *
 * The engine replaces the body of NECI's PerformFCIMCycPar
 * (src/FciMCPar.F90:1226-1873) for one MPI rank's share of the walkers: the
 * walker list, the walker hash and the spawn buffers stay resident in GPU
 * memory, and one ngpu_iterate() call performs spawn -> merge -> death ->
 * annihilation -> hash statistics.  The five procedure pointers of NECI's
 * plugin surface (src/proc_ptrs.F90:251-258, bound by init_fcimc_fn_pointers,
 * src/fcimc_initialisation.fpp:2029-2283) are selected once by
 * ngpu_set_excitgen()/ngpu_set_integrals()/ngpu_set_lattice(); host-callable
 * batched forms of each pointer are exported for parity testing.
 *
 * Conventions (those of the reference boundary):
 *   - caller-owned host arrays, plain pointers and sizes;
 *   - orbitals are 1-based spin orbitals, odd = beta (macros.h:21);
 *   - determinants are NIfD+1 = nbasis/64+1 int64 words (BitReps.F90:174);
 *   - walker rows use the CurrentDets layout: det words | lenof_sign FP64
 *     signs stored as bit patterns | flag word (BitReps.F90:164-214), with
 *     NECI's flag bit numbers (bit_rep_data.F90:87-118);
 *   - every function returns an ngpu_status; errors mirror the reference's
 *     err/stop_all sites (fcimc_helper.F90:483-510, load_balancer.fpp:540-544,
 *     fcimc_pointed_fns.F90:649-673).  No exceptions cross the ABI;
 *   - one handle per host rank; a handle is not thread-safe.
 */
#ifndef NECI_GPU_H
#define NECI_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NGPU_MAX_REPLICAS 16

typedef struct ngpu_engine ngpu_t;
typedef struct ngpu_batch ngpu_batch_t;

typedef enum {
    NGPU_OK = 0,
    NGPU_ERR_SPAWN_FULL = 1,   /* checkValidSpawnedList, fcimc_helper.F90:483-510 (MEMORYFACSPAWN) */
    NGPU_ERR_WALKERS_FULL = 2, /* AddNewHashDet, load_balancer.fpp:540-544 (MEMORYFACPART) */
    NGPU_ERR_DEATH_PROB = 3,   /* attempt_die_normal, fcimc_pointed_fns.F90:649-673 (tau too big) */
    NGPU_ERR_ARG = 4,
    NGPU_ERR_HIP = 5,
    NGPU_ERR_HASH_FULL = 6,    /* every bucket of the walker hash is taken (MEMORYFACHASH) */
    NGPU_ERR_STATE = 7,        /* call out of order (e.g. iterate before set_reference) */
    NGPU_ERR_EXCITGEN = 8,     /* uniform generator found no allowed orbital within the reference's
                                  attempt bounds (Stop_All sites, GenRandSymExcitNUMod.F90:724-870) */
    NGPU_ERR_CHECK = 9         /* an invariant of ngpu_set_checks failed (the reference's ASSERTs and
                                  CalcHashTableStats stop_all, load_balancer.fpp:796-804) */
} ngpu_status;

enum { NGPU_MODEL_HUBBARD = 0, NGPU_MODEL_MOLECULAR = 1 };
/* singles: "on-the-fly-heat-bath" (excit_gens_int_weighted.F90:442) */
enum { NGPU_SINGLES_HEATBATH = 1 };
/* PCHB doubles particle selection (gasci_pchb_doubles_select_particles.fpp:38-52) */
enum { NGPU_PART_UNIF_UNIF = 0, NGPU_PART_UNIF_FAST = 1, NGPU_PART_UNIF_FULL = 2, NGPU_PART_FULL_FULL = 3 };
/* PCHB doubles hole selection, spin-orbital resolved (pchb_excitgen.F90:216-262):
   FAST-FAST (gasci_pchb_doubles_spinorb_fastweighted.fpp), FULL-FULL
   (gasci_pchb_doubles_spinorb_fullyweighted.fpp) */
enum { NGPU_HOLES_FAST_FAST = 0, NGPU_HOLES_FULL_FULL = 1 };
/* flag bit layout of walker rows at the boundary */
enum { NGPU_FLAGS_MNECI = 0, NGPU_FLAGS_NECI = 1 };

/* System and run parameters (SystemData / CalcData globals the engine needs). */
typedef struct {
    int32_t nel;                 /* electrons */
    int32_t nbasis;              /* spin orbitals */
    int32_t n_replicas;          /* lenof_sign (1 = neci, 2 = dneci, R = mneci) */
    int32_t model;               /* NGPU_MODEL_* */
    int32_t t_trunc_initiator;   /* TRUNCINITIATOR */
    int32_t t_all_real_coeff;    /* ALLREALCOEFF */
    int32_t t_real_spawn_cutoff; /* REALSPAWNCUTOFF on/off */
    int32_t flag_layout;         /* NGPU_FLAGS_* */
    double initiator_walk_no;    /* ADDTOINITIATOR (default 3) */
    double real_spawn_cutoff;    /* REALSPAWNCUTOFF value (default 0.95) */
    double occupied_thresh;      /* SETOCCUPIEDTHRESH (default 1) */
    double av_mc_excits;         /* AvMCExcits (default 1) */
    int64_t max_walkers;         /* MaxWalkersPart = MEMORYFACPART * InitWalkers */
    int64_t max_spawned;         /* MaxSpawned = MEMORYFACSPAWN * InitWalkers * inum_runs */
    int64_t n_hash_buckets;      /* nWalkerHashes = nint(MEMORYFACHASH * InitWalkers), at most 2^29 */
    uint32_t seed;               /* SEED (keys the counter-based RNG) */
    int32_t pad;
} ngpu_sys_t;

/* Per-iteration inputs (Iter, tau, DiagSft relative to Hii). */
typedef struct {
    double tau;
    double diag_sft[NGPU_MAX_REPLICAS];
    uint32_t iter;
    int32_t pad;
} ngpu_iter_in_t;

/* Per-iteration outputs: fcimc_iter_data (FciMCData.F90:445) plus the
   accumulators of SumEContrib, CalcParentFlag and CalcHashTableStats. */
typedef struct {
    double nborn[NGPU_MAX_REPLICAS], ndied[NGPU_MAX_REPLICAS], nannihil[NGPU_MAX_REPLICAS];
    double naborted[NGPU_MAX_REPLICAS], nremoved[NGPU_MAX_REPLICAS];
    double tot_parts[NGPU_MAX_REPLICAS], norm_sq[NGPU_MAX_REPLICAS];
    double enum_cyc[NGPU_MAX_REPLICAS], hf_cyc[NGPU_MAX_REPLICAS], noat_doubs[NGPU_MAX_REPLICAS];
    double spawn_from_sing[NGPU_MAX_REPLICAS], inst_no_at_hf[NGPU_MAX_REPLICAS];
    double no_init_dets[NGPU_MAX_REPLICAS], no_init_walk[NGPU_MAX_REPLICAS];
    double no_noninit_dets[NGPU_MAX_REPLICAS], no_noninit_walk[NGPU_MAX_REPLICAS];
    double n_valid_excits, n_invalid_excits, max_cyc_spawn;
    double bloom_count[3], bloom_size[3];
    int64_t tot_walkers;         /* occupied determinants after the iteration */
    int64_t n_attempts, n_spawned, n_unique, n_new;
    int32_t err, pad;
} ngpu_iter_out_t;

/* ---- lifetime and set-up (replaces the allocations of InitFCIMCCalcPar,
        fcimc_initialisation.fpp:1543-1760) ---- */
int ngpu_create(ngpu_t **h, const ngpu_sys_t *sys);
int ngpu_destroy(ngpu_t *h);

/* FCIDUMP reader (host code, no device work): the &FCI namelist that
   READFCIINT reads (readint.F90:614-636).  orbsym receives up to orbsym_cap
   ORBSYM labels (may be NULL).  Unrestricted (UHF, IUHF) and relativistic
   (TREL) dumps are refused with NGPU_ERR_ARG, as is a missing file. */
int ngpu_fcidump_header(const char *path, int32_t *norb, int32_t *nelec, int32_t *ms2, int32_t *isym,
                        int32_t *orbsym, int32_t orbsym_cap);
/* FCIDUMP integrals (READFCIINT body, readint.F90:700-898) into the arrays
   ngpu_set_integrals takes: tmat2d (nbasis x nbasis, nbasis = 2*NORB), umat
   (umat_len = P(P+1)/2 with P = NORB(NORB+1)/2) and ecore.  Lines with
   |Z| < umat_eps and I > 0 are skipped (UMatEps, 1e-8 by default,
   System_neci.F90:110); *n_two receives the number of two-electron integrals
   stored (the reference's NonZeroInt).  A one-electron pair given twice with
   values more than 1e-7 apart is NGPU_ERR_ARG, as the reference stops there. */
int ngpu_read_fcidump(const char *path, double umat_eps, double *tmat2d, int64_t nbasis, double *umat,
                      int64_t umat_len, double *ecore, int64_t *n_two);

/* Integrals as READFCIINT leaves them (readint.F90:585-950): TMAT2D(nbasis,nbasis)
   column major; UMAT packed by UMatInd_base (UMatCache.F90:257) over spatial
   orbitals, UMAT(1) at umat[0]; ECORE.  Selects get_spawn_helement =
   get_helement_det_only (fcimc_initialisation.fpp:2210). */
int ngpu_set_integrals(ngpu_t *h, const double *tmat2d, const double *umat, int64_t umat_len,
                       double ecore);

/* Real-space Hubbard lattice (init_real_space_hubbard, real_space_hubbard.F90:140):
   nbr[(site-1)*maxnbr + k] = 1-based neighbour site, 0 terminates the row.
   Selects gen_excit_rs_hubbard and get_helement_lattice. */
int ngpu_set_lattice(ngpu_t *h, const int32_t *nbr, int32_t nsites, int32_t maxnbr, double bhub,
                     double uhub);

/* Molecular excitation generator: FCI-PCHB doubles (FAST-FAST holes, spin-orbital
   resolved, gasci_pchb_doubles_spinorb_fastweighted.fpp) with the given particle
   selection, singles with probability psingles (gen_exc_sd,
   excitation_generators.F90:112).  Builds the alias tables (aliasSampling.F90:209)
   on the host and uploads them.  Requires ngpu_set_integrals first. */
int ngpu_set_excitgen(ngpu_t *h, int32_t singles, int32_t particles, double psingles,
                      double pparallel);
/* ngpu_set_excitgen with the hole selection given: NGPU_HOLES_FAST_FAST, or
   NGPU_HOLES_FULL_FULL, the fully weighted holes of
   gasci_pchb_doubles_spinorb_fullyweighted.fpp:138-296 (first hole from
   sum_B |<IJ||AB>|, second from |<IJ||AB>|, each restricted to the unoccupied
   orbitals), which the reference binds when nBasis <= 4 nEl
   (decide_on_PCHB_options, pchb_excitgen.F90:216-223).  Weighted particles
   (UNIF-FULL, FULL-FULL) and weighted holes draw from the attempt's keyed
   stream, redrawing where the reference's constrained samplers redraw. */
int ngpu_set_excitgen_pchb(ngpu_t *h, int32_t singles, int32_t particles, int32_t holes, double psingles,
                           double pparallel);

/* Molecular excitation generator: NECI's uniform random excitation generator
   gen_rand_excit (GenRandSymExcitNUMod.F90:72-153, the default binding of
   generate_excitation, fcimc_initialisation.fpp:2138) for a C1 system: singles
   and doubles chosen uniformly among the spin-allowed ones, doubles with
   probability 1 - psingles.  Replaces ngpu_set_excitgen's binding. */
int ngpu_set_excitgen_uniform(ngpu_t *h, double psingles);

/* Storage precision of the spawned amplitudes.  create_particle
   (fcimc_helper.F90:197) writes each child's sign into SpawnedParts as FP64
   (encode_sign, BitReps.F90:454); fp32 = 1 stores it as FP32 instead (the
   value attempt_create returns is rounded to nearest, so ngpu_attempt_create_batch
   returns it rounded too), and the merge and annihilation still sum in FP64.
   fp32 = 0 (the default) is the reference's FP64.  With FP32, a spawned record
   takes W + 1 words instead of 4 (16 bytes for one-word determinants), and
   ngpu_annihilate rounds its input amplitudes to FP32.  Not while batched. */
int ngpu_set_spawn_precision(ngpu_t *h, int32_t fp32);

/* RandomHash2(1:nbasis) (fcimc_initialisation.fpp:890-915) for FindWalkerHash. */
int ngpu_set_hash(ngpu_t *h, const int32_t *random_hash2);

/* Reference determinant ilutRef (W words) and Hii. */
int ngpu_set_reference(ngpu_t *h, const int64_t *ilut_ref, double hii);

/* Walkers in CurrentDets layout (rows of nbasis/64+1+n_replicas+1 int64).
   The engine computes det_diagH / det_offdiagH itself (AddNewHashDet,
   load_balancer.fpp:514-629). */
int ngpu_upload_walkers(ngpu_t *h, const int64_t *rows, int64_t n);
/* Occupied walkers (and core dets) back in CurrentDets layout, with
   det_diagH (= H_ii - Hii) and det_offdiagH; *n = number of occupied rows. */
int ngpu_download_walkers(ngpu_t *h, int64_t *rows, double *diag_h, double *offdiag_h,
                          int64_t cap, int64_t *n);
/* Allocation-free form of ngpu_download_walkers: the same rows, in the same
   order, compacted on the device into caller-owned DEVICE memory (rows:
   cap rows of nbasis/64+1+n_replicas+1 int64; diag_h, offdiag_h: cap doubles
   or NULL), with the number of occupied rows written to the device int64 *n.
   No counterpart in the reference; it is for callers that cannot allocate
   (code under HIP graph capture, checkpoint writers inside a serving loop):
   it calls no hipMalloc/hipFree and no host synchronisation, and runs
   asynchronously on the engine's stream, ordered after queued iterations.
   Rows past cap are counted in *n but not written.  Its scratch is a fixed
   128 KiB array per GPU, held until the process ends and shared by every
   call on that GPU: calls that overlap on one GPU race.  An engine whose
   walker table holds more than 67,108,864 slots (MaxWalkersPart) is refused
   with NGPU_ERR_ARG. */
int ngpu_download_walkers_dev(ngpu_t *h, int64_t *rows, double *diag_h, double *offdiag_h,
                              int64_t cap, int64_t *n);

/* Semistochastic core space (init_semi_stochastic, semi_stoch_gen.F90:425-445;
   semi_stoch_procs.F90:1095-1120).  The ncore determinants (W words each)
   become core states: resident in the walker list even when empty, flagged
   deterministic and, with the initiator approximation, initiators
   (t_core_inits).  The engine builds the core Hamiltonian on the device
   (calc_determ_hamil, fast_determ_hamil.F90:892-1548: H_ij of connected pairs,
   diagonal H_ii - Hii, off-diagonal zeros dropped) as CSR, values in FP64 or,
   with fp32_values, rounded to FP32 (accumulation stays FP64); *nnz receives
   the number of stored elements.  Every later iteration projects the core
   amplitudes deterministically (determ_projection, semi_stoch_procs.F90:105-241;
   deterministic_annihilation, Annihilation.F90:930-963) and cancels spawns from
   core to core dets (FciMCPar.F90:1659-1668).  One core space per engine: a
   second call with the same determinants (same order) only stores the core
   Hamiltonian again in the precision fp32_values asks for; any other second
   call returns NGPU_ERR_STATE.  Repeated determinants: NGPU_ERR_ARG. */
int ngpu_set_core(ngpu_t *h, const int64_t *dets, int64_t ncore, int32_t fp32_values, int64_t *nnz);

/* How ngpu_set_core builds the core Hamiltonian (call before it).  Every
   method gives the same CSR (rows and columns ascending, the same values):
   NGPU_CORE_BUILD_JOIN finds the connected pairs by joining the core
   determinants on their (nel - 2)-electron subsets (a radix sort of
   ncore * nel(nel-1)/2 entries), the role of calc_determ_hamil_opt
   (fast_determ_hamil.F90:892-1548, semi_stoch_gen.F90:262-268);
   NGPU_CORE_BUILD_SCAN tests every pair (ncore^2 excitation levels), the role
   of calc_determ_hamil_sparse (sparse_arrays.F90:426-572, the reference's
   fallback at semi_stoch_gen.F90:250-260, which is also taken here for fewer
   than 4 electrons); NGPU_CORE_BUILD_AUTO (the default) takes the scan below
   450 * nel(nel-1)/2 core determinants and the join from there on.
   NGPU_ERR_STATE once the core is set. */
enum { NGPU_CORE_BUILD_AUTO = 0, NGPU_CORE_BUILD_JOIN = 1, NGPU_CORE_BUILD_SCAN = 2 };
int ngpu_set_core_build(ngpu_t *h, int32_t method);

/* The core Hamiltonian as stored (print_hamiltonian, semi_stoch_procs.F90:
   2051-2063, PRINT-CORE-HAMIL): rowptr (ncore + 1 entries), cols (0-based core
   indices, ascending in each row) and vals (H_ij, diagonal H_ii - Hii; FP32
   values widened).  With rowptr, cols and vals all NULL only *nnz is set;
   NGPU_ERR_ARG when cap < nnz; NGPU_ERR_STATE without a core.  Waits. */
int ngpu_get_core_hamil(ngpu_t *h, int64_t *rowptr, int32_t *cols, double *vals, int64_t cap, int64_t *nnz);

/* generate_space_most_populated (semi_stoch_gen.F90:1031-1226) over the
   resident walkers: the target determinants of largest weight sum_r |c_r|
   (core_space_weight, semi_stoch_procs.F90:1623-1642, global core space),
   weights below 1e-8 left out, all of them when fewer are occupied; chosen on
   the device (stable radix sorts of the occupied slots).  Equal weights are
   ordered by determinant (ilut_lt), where the reference keeps the first it
   meets in its storage order.  dets (target * W words) receives them, weight
   descending; *n their number.  Runs between iterations (waits). */
int ngpu_most_populated(ngpu_t *h, int64_t target, int64_t *dets, int64_t *n);
/* The semistochastic core space of the target most populated determinants
   (ngpu_most_populated, then ngpu_set_core with them): the POPS-CORE core
   space of init_semi_stochastic (semi_stoch_gen.F90:425-445).  dets (may be
   NULL) receives the chosen determinants, *ncore their number. */
int ngpu_select_core(ngpu_t *h, int64_t target, int32_t fp32_values, int64_t *dets, int64_t *ncore,
                     int64_t *nnz);

/* One FCIQMC iteration (PerformFCIMCycPar, FciMCPar.F90:1177-1920).  Runs
   asynchronously on the engine's stream when out == NULL; otherwise waits and
   fills out.  A failure inside the iteration is reported by the next call that
   waits, and halts the engine (ngpu_get_error). */
int ngpu_iterate(ngpu_t *h, const ngpu_iter_in_t *in, ngpu_iter_out_t *out);

/* Replay one captured graph of the iteration's launches instead of issuing
   them one by one (on != 0).  The kernels read tau, DiagSft and Iter from
   device memory, written by one small launch before each replay, so the graph
   is captured once and recaptured only after a set-up call changes the chain
   (ngpu_set_core, ngpu_set_excitgen, ...).  No counterpart in the reference
   (its iteration is a host loop); results are identical either way. */
int ngpu_set_graph(ngpu_t *h, int32_t on);

/* Launch shape, the knobs of the autotuner (neci_stable_amd/autotune.py):
   blocks of the chunked kernels (prepass, attempt scan, radix passes; 64 to
   16 per compute unit), blocks of the grid-stride kernels (spawn, merge and
   annihilation, insert, statistics; 64 to 64 per compute unit), elements
   per thread of a radix tile (2, 4 or 8), lanes per heat-bath single attempt
   (4, 8 or 16) and blocks of the merge-annihilation kernel (64 to 64 per
   compute unit).  0 keeps a value.  Results do not
   depend on the shape (the kernels' outputs are independent of how work is
   split; only the storage order of new determinants may differ).  No
   counterpart in the reference. */
int ngpu_set_launch_shape(ngpu_t *h, int32_t chunk_blocks, int32_t grid_blocks, int32_t radix_items,
                          int32_t single_lanes, int32_t merge_blocks);
int ngpu_get_launch_shape(ngpu_t *h, int32_t *chunk_blocks, int32_t *grid_blocks, int32_t *radix_items,
                          int32_t *single_lanes, int32_t *merge_blocks, int32_t *n_cu);

/* Batches: several independent engines (different systems, sizes, replica
   counts) iterated together: one launch per kernel of the iteration chain
   for each group of engines that share kernel instances
   (ngpu_batch_get_groups), captured as ONE graph per iteration.  No counterpart in the
   reference, which runs one system per process; each engine's results are
   those of iterating it alone.  While batched, an engine is iterated only
   through ngpu_batch_iterate (ngpu_iterate returns NGPU_ERR_STATE); its
   statistics, downloads and set-up calls work as before and are ordered
   after the batch's iterations.  ngpu_batch_destroy returns the engines to
   independent use; destroy the batch before its engines. */
int ngpu_batch_create(ngpu_batch_t **b, ngpu_t *const *hs, int32_t n);
/* One iteration of every engine of the batch; ins holds one input per engine,
   in the order given to ngpu_batch_create.  Asynchronous. */
int ngpu_batch_iterate(ngpu_batch_t *b, const ngpu_iter_in_t *ins);
int ngpu_batch_destroy(ngpu_batch_t *b);
/* The batch's launch groups: engines whose iteration chains use the same
   kernel instances (determinant words, replica layout, excitation generator,
   sort path, spawn-record and core precision, launch knobs).  Every kernel of
   the chain is ONE launch for all the engines of a group, each engine taking
   a share of its blocks in proportion to its walker capacity; the groups run
   concurrently, each on its own stream.  group_of_engine[i] receives engine
   i's group (either pointer may be NULL).  No counterpart in the reference. */
int ngpu_batch_get_groups(ngpu_batch_t *b, int32_t *group_of_engine, int32_t *n_groups);
/* Properties of launch group `group`: NGPU_GROUP_CORE when its engines have a
   core space; NGPU_GROUP_SHARED_CORE when every engine of the group has the
   same core determinants (replicas of one system, as under the reference's
   global core space, semi_stoch_procs.F90:1623-1642), so one projection reads
   the core Hamiltonian once for all of them (its sums run in another order
   than an engine's own projection: the same amplitudes to rounding). */
enum { NGPU_GROUP_CORE = 1, NGPU_GROUP_SHARED_CORE = 2 };
int ngpu_batch_get_group_flags(ngpu_batch_t *b, int32_t group, int32_t *flags);
/* on != 0 (the default): each batch iteration replays one captured graph;
   0: the group launches are issued directly (needed for stage timing). */
int ngpu_batch_set_graph(ngpu_batch_t *b, int32_t on);
/* Stage timing of the batch's launches (mask as ngpu_set_profiling), while
   the batch launches directly: each group's stage times are read with
   ngpu_get_stage_times on the group's first engine. */
int ngpu_batch_set_profiling(ngpu_batch_t *b, int32_t on);
/* Wait for the last iteration and return its statistics. */
int ngpu_get_stats(ngpu_t *h, ngpu_iter_out_t *out);
/* Update cycle (calculate_new_shift_wrapper, fcimc_iter_utilities.F90:1403):
   wait, return the statistics summed over the iterations since the previous
   call (maxima for max_cyc_spawn/bloom_size, last value for tot_walkers and
   inst_no_at_hf) plus the last iteration's, and start a new cycle. */
int ngpu_get_cycle_stats(ngpu_t *h, ngpu_iter_out_t *sum, ngpu_iter_out_t *last, int32_t *n_iters);

/* Sticky error of the iteration chain (waits for queued work).  The first
   iteration that reports a status (capacity, stability) halts the engine:
   every later iteration returns at once, so the walker list stays as that
   iteration left it, the point at which the reference stops (stop_all).
   ngpu_iterate with an output block, ngpu_get_stats and ngpu_get_cycle_stats
   report the status from then on; this call also gives the number (Iter) of
   the iteration that failed.  *status = NGPU_OK and *iter = 0 while the
   engine runs.  A halted engine refuses uploads, core spaces and explicit
   lists with the same status; destroy it.  Either pointer may be NULL. */
int ngpu_get_error(ngpu_t *h, int32_t *status, uint32_t *iter);

/* Walker-hash diagnostics (waits for queued work): live entries, tombstones
   (entries of removed determinants not yet reclaimed), total entries (4 per
   bucket) and the number of in-place rebuilds since creation.  No counterpart
   in the reference, whose chained buckets (hash.F90) have no tombstones; any
   argument may be NULL. */
int ngpu_get_hash_stats(ngpu_t *h, int64_t *live, int64_t *tombstones, int64_t *entries, int64_t *rebuilds);

/* Invariant checks re-evaluated on the device after every iteration (the
   reference's DEBUG_ ASSERTs and hard checks; off by default, results are the
   same either way):
   NGPU_CHECK_TABLE: every walker slot below the high-water mark is a hole of
     the free queue or a live determinant with nel electrons that a hash
     lookup finds at its own slot; holes == free-queue length (AnnihilatedDet
     == HolesInList, load_balancer.fpp:796-804); live slots == the live count;
   NGPU_CHECK_PGEN: one excitation drawn from every occupied determinant and
     replica has pgen equal to calc_pgen to 1e-12 relative (the ASSERT after
     every draw, gasci_pchb_doubles_spinorb_fastweighted.fpp:191,231);
   NGPU_CHECK_SELFTEST: with NGPU_CHECK_PGEN, the recomputed pgen is scaled by
     1 + 1e-6, so the check fails (for testing the report).
   A failed check halts the engine with NGPU_ERR_CHECK, as any other error of
   the iteration.  Single-engine iterations only (not batches). */
enum { NGPU_CHECK_TABLE = 1, NGPU_CHECK_PGEN = 2, NGPU_CHECK_SELFTEST = 4 };
int ngpu_set_checks(ngpu_t *h, int32_t mask);
/* Cumulative check counts (n <= 7): iterations checked, slots with a wrong
   electron count, live slots the hash does not find at their slot, iterations
   whose holes differ from the free-queue length, iterations whose live slots
   differ from the live count, pgen tests, pgen mismatches. */
int ngpu_get_check_counts(ngpu_t *h, int64_t *counts, int32_t n);

/* Stage timing with events on the engine stream (set_timer/halt_timer analogue,
   lib/timing.F90:122-188).  Each stage is timed from the start of its first
   kernel to the end of its last.  Stages: 0 prepass (initiator, proj-E, death),
   1 attempt scan (2 kernels), 2 spawn (all attempts; molecular: doubles, singles
   queued), 3 queued singles, 4 sort (all radix passes), 5 merge (empty: fused
   into 6), 6 merge and annihilation, 7 insert, 8 hash statistics, 9 the core
   Hamiltonian's H.v (semistochastic projection).  Stages 0, 2, 3, 6, 7, 8 and
   9 are one kernel each.  on is a bit mask of the stages to time
   (bit s for stage s; 0 switches timing off); each timed stage adds two events
   to its launches, so time only the stages needed. */
int ngpu_set_profiling(ngpu_t *h, int32_t on);
int ngpu_get_stage_times(ngpu_t *h, double *ms_total, int32_t nstages, int64_t *n_iters);

/* ---- parity entry points: batched host-callable forms of the plugin surface ---- */

/* FindWalkerHash (hash.F90:23-37) of the determinants in dets (W words each). */
int ngpu_find_walker_hash_batch(ngpu_t *h, const int64_t *dets, int64_t n, int32_t *buckets);

/* get_spawn_helement (proc_ptrs.F90:82-97) for pairs (ilutI, ilutJ): H_IJ. */
int ngpu_get_spawn_helement_batch(ngpu_t *h, const int64_t *ilut_i, const int64_t *ilut_j,
                                  int64_t n, double *hel);

/* get_diagonal_matel / get_off_diagonal_matel (matel_getter.F90:27-107):
   diag = H_ii (total), offdiag = H(det, ref). */
int ngpu_diag_helement_batch(ngpu_t *h, const int64_t *dets, int64_t n, double *diag,
                             double *offdiag);

/* generate_excitation (proc_ptrs.F90:13-32) from each det with six explicit
   uniforms per call (draws[6*i..6*i+5]); ic = 0 marks a null excitation.
   ex[4*i] = (src1, src2 | tgt1, tgt2) for doubles, (src, tgt, 0, 0) for singles. */
int ngpu_generate_excitation_batch(ngpu_t *h, const int64_t *ilut, const double *draws,
                                   int64_t n, int64_t *ilut_j, int32_t *ic, int32_t *ex,
                                   int32_t *parity, double *pgen);

/* calc_pgen of the bound generator for given excitations. */
int ngpu_calc_pgen_batch(ngpu_t *h, const int64_t *ilut, const int32_t *ic, const int32_t *ex,
                         int64_t n, double *pgen);

/* encode_child = FindExcitBitDet (DetBitOps.F90:584-623); pure host function. */
int ngpu_encode_child(const int64_t *ilut_i, int32_t nwords, int32_t ic, const int32_t *ex,
                      int64_t *ilut_j);

/* attempt_die_normal (fcimc_pointed_fns.F90:573-705) for n (Kii, sign) pairs of one
   replica with explicit uniforms; ndie = deaths (negative = births). */
int ngpu_attempt_die_batch(ngpu_t *h, const double *kii, const double *sign, const double *udraw,
                           int64_t n, double tau, double diag_sft, double *ndie);

/* attempt_create_normal (fcimc_pointed_fns.F90:178-491): child amplitude for
   n (hel, pgen, parent sign) triples with explicit rounding uniforms. */
int ngpu_attempt_create_batch(ngpu_t *h, const double *hel, const double *pgen,
                              const double *parent_sign, const double *udraw, int64_t n,
                              double tau, double *child);

/* CompressSpawnedList + AnnihilateSpawnedParts + CalcHashTableStats
   (Annihilation.F90:249-1352, load_balancer.fpp:646-805) on an explicit spawned
   list against the resident walkers.  sinfo: bits 0-7 replica, bit 8 parent
   initiator.  merged_* receive the compressed list (any order). */
int ngpu_annihilate(ngpu_t *h, uint32_t iter, const int64_t *sdet, const double *sval,
                    const uint32_t *sinfo, int64_t n, ngpu_iter_out_t *out, int64_t *merged_det,
                    double *merged_sgn, uint32_t *merged_flags, int64_t *n_merged);

#ifdef __cplusplus
}
#endif
#endif
