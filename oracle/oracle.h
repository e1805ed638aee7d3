/*
 * oracle.h -- CPU restatement of NECI's FCIQMC per-iteration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP engine in neci_stable_amd/ and the CPU baseline timed by bench.py.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product path never does.
 *
 * Every function follows a routine of the reference (file:line cited at the
 * definition in oracle.c).  Random numbers come from the keyed Philox-4x32-10
 * stream described below instead of NECI's sequential dSFMT stream, so that a
 * GPU iteration and an oracle iteration over the same walker set draw the same
 * numbers at the same decision points and can be compared bit for bit.
 *
 * RNG contract (shared with the HIP engine, written independently in each):
 *   key = (seed, replica)
 *   ctr = (lo32(dh), hi32(dh), iter, (k << 3) | purpose)
 *     dh = det word 0                       (1-word dets)
 *        = word0 ^ mix64(word1)             (2-word dets, mix64 = splitmix64 finaliser)
 *   u01(lo, hi) = ((hi << 32 | lo) >> 11) * 2^-53; two uniforms per block:
 *     u_a = u01(out0, out1), u_b = u01(out2, out3)
 *   purpose 0: per det+replica: u_a decide_num_to_spawn, u_b attempt_die
 *   purpose 1,2,3: spawn attempt k: d0,d1 / d2,d3 / d4,d5
 *   purpose 4: per new det+replica: stochRoundSpawn
 *   purpose 5: per det+replica: CalcHashTableStats rounding
 *   purpose 6: spawn attempt k, draws of generators with rejection loops
 *     beyond d1..d4: block b uses key = (seed, replica | (b + 1) << 8)
 */
#ifndef NECI_ORACLE_H
#define NECI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define O_MAX_REPLICAS 16
#define O_MAX_NEL 64
#define O_MAX_W 2

enum { O_MODEL_HUBBARD = 0, O_MODEL_MOLECULAR = 1 };
enum { O_SINGLES_UNIFORM = 0, O_SINGLES_HEATBATH = 1 };
enum { O_PART_UNIF_UNIF = 0, O_PART_UNIF_FAST = 1, O_PART_UNIF_FULL = 2, O_PART_FULL_FULL = 3 };
enum { O_HOLES_FAST_FAST = 0, O_HOLES_FULL_FULL = 1 };

typedef struct {
    int32_t nel, nbasis, nrep;          /* electrons, spin orbitals, lenof_sign */
    int32_t model;
    /* Hubbard lattice: nbr[site*maxnbr + k] = 1-based neighbour site (0 = none) */
    int32_t nsites, maxnbr;
    const int32_t *nbr;
    double bhub, uhub;
    /* molecular: TMAT2D(i,j) stored tmat[(i-1) + (j-1)*nbasis] (column major),
       UMAT packed by UMatInd over spatial orbitals, 1-based index -> umat[idx-1] */
    const double *tmat;
    const double *umat;
    int64_t umat_len;
    double ecore;
    /* run parameters */
    double tau, avmc, init_walk_no, real_spawn_cutoff, occ_thresh;
    int32_t t_trunc_init, t_real_coeff, t_real_spawn_cutoff;
    int32_t spawn_fp32;                 /* 1: spawned amplitudes stored in FP32 (rounded to nearest) */
    double hii;                         /* reference diagonal element (total energy) */
    uint64_t ref[O_MAX_W];              /* reference determinant */
    /* excitation generator */
    int32_t singles, particles;
    double psingles, pparallel;
    /* PCHB tables, built by o_build_pchb */
    int32_t pchb_ready;
    int32_t uniform;                    /* 1: gen_rand_excit instead of PCHB + heat-bath singles */
    /* walker hash */
    const int32_t *random_hash2;        /* RandomHash2(1:nbasis) */
    int64_t nbuckets;
    uint32_t seed, pad2;
    int64_t max_walkers, max_spawned;
    int32_t holes, pad3;                /* PCHB hole selection (O_HOLES_*) */
} osys_t;

typedef struct {
    double nborn[O_MAX_REPLICAS], ndied[O_MAX_REPLICAS], nannihil[O_MAX_REPLICAS];
    double naborted[O_MAX_REPLICAS], nremoved[O_MAX_REPLICAS];
    double tot_parts[O_MAX_REPLICAS], norm_sq[O_MAX_REPLICAS];
    double enum_cyc[O_MAX_REPLICAS], hf_cyc[O_MAX_REPLICAS], noat_doubs[O_MAX_REPLICAS];
    double spawn_from_sing[O_MAX_REPLICAS], inst_no_at_hf[O_MAX_REPLICAS];
    double no_init_dets[O_MAX_REPLICAS], no_init_walk[O_MAX_REPLICAS];
    double no_noninit_dets[O_MAX_REPLICAS], no_noninit_walk[O_MAX_REPLICAS];
    double n_valid_excits, n_invalid_excits, max_cyc_spawn;
    double bloom_count[3], bloom_size[3];
    int64_t tot_walkers;                /* occupied determinants after the iteration */
    int64_t n_attempts, n_spawned, n_unique, n_new;
    int32_t err, pad;
} ostats_t;

typedef struct ostate ostate_t;

/* --- elementary functions (exported for unit tests) --- */
void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
int32_t o_find_walker_hash(const int32_t *nI, int32_t nel, const int32_t *random_hash2,
                           int64_t L);
uint64_t o_umat_ind(int32_t i, int32_t j, int32_t k, int32_t l);
int32_t o_decode(const uint64_t *ilut, int32_t W, int32_t *nI);
int32_t o_make_double(const int32_t *nI, int32_t nel, int32_t e1, int32_t e2, int32_t t1,
                      int32_t t2, int32_t *nJ, int32_t *ex);
int32_t o_make_single(const int32_t *nI, int32_t nel, int32_t elec, int32_t tgt,
                      int32_t *nJ, int32_t *ex);
double o_get_helement(const osys_t *s, const int32_t *nI, const int32_t *nJ, int32_t nel);
double o_get_helement_ex(const osys_t *s, const int32_t *nI, int32_t ic, const int32_t *ex,
                         int32_t tparity);
double o_diag_helement(const osys_t *s, const int32_t *nI);
double o_offdiag_helement(const osys_t *s, const uint64_t *ilut);

/* PCHB alias tables (aliasSampling.F90 + gasci_pchb_doubles_spinorb_fastweighted.fpp) */
int32_t o_build_pchb(osys_t *s);
void o_free_pchb(void);
void o_alias_build(const double *w, int32_t n, double *bias, int32_t *alias, double *prob);
int32_t o_get_pchb_row(int32_t which, int32_t row, double *bias, int32_t *alias, double *prob);

/* one excitation draw from det ilut with explicit uniforms d[6] */
int32_t o_gen_excit(const osys_t *s, const uint64_t *ilut, const double *d, int32_t *nJ,
                    uint64_t *ilutJ, int32_t *ic, int32_t *ex, int32_t *tparity, double *pgen);
double o_calc_pgen(const osys_t *s, const uint64_t *ilut, int32_t ic, const int32_t *ex);
/* as o_gen_excit with the draw context of spawn attempt k of (det, iter, replica):
   generators with rejection loops draw beyond d[1..4] from that context */
/* the heat-bath single weight H(src -> orb) relative to the reference (signed) */
double o_hb_weight(const osys_t *s, const uint64_t *ilut, int32_t src, int32_t orb);
int32_t o_gen_excit_ctx(const osys_t *s, const uint64_t *ilut, const double *d, uint32_t iter, int32_t rep,
                        uint32_t k, int32_t *nJ, uint64_t *ilutJ, int32_t *ic, int32_t *ex, int32_t *tparity,
                        double *pgen, int32_t *fail);

/* --- iteration engine (the restatement of PerformFCIMCycPar) --- */
ostate_t *o_create(const osys_t *sys);
void o_destroy(ostate_t *st);
int32_t o_add_walker(ostate_t *st, const uint64_t *det, const double *sgn, uint32_t flags);
int32_t o_add_walkers(ostate_t *st, const uint64_t *dets, const double *sgns, const uint32_t *flags,
                      int64_t n);
int64_t o_get_walkers(ostate_t *st, uint64_t *dets, double *sgns, uint32_t *flags,
                      double *diagh, double *offdiagh, int64_t cap);
int32_t o_iterate(ostate_t *st, uint32_t iter, const double *diag_sft, ostats_t *out);
/* semistochastic core space: n dets (W words each); CSR of H_ij with diagonal
   H_ii - Hii; fp32 = values rounded to FP32 */
int32_t o_set_core(ostate_t *st, const uint64_t *dets, int64_t n, int32_t fp32, int64_t *nnz);
int64_t o_most_populated(ostate_t *st, int64_t target, uint64_t *out);
void o_core_project(const ostate_t *st, const double *x, const double *shift, double tau, double *y);
int64_t o_core_csr(const ostate_t *st, int64_t *rowptr, int32_t *cols, double *vals);
/* annihilation of an explicit spawned list only (CompressSpawnedList + death-free
   AnnihilateSpawnedParts); spawned rows are (det words | sign | info) */
int32_t o_annihilate_list(ostate_t *st, uint32_t iter, const uint64_t *sdet, const double *sval,
                          const uint32_t *sinfo, int64_t n, ostats_t *out,
                          uint64_t *merged_det, double *merged_sgn, uint32_t *merged_flags,
                          int64_t *n_merged);

/* --- hash-sharded iteration over P threads (the reference's one rank per core) --- */
typedef struct omt omt_t;
omt_t *o_mt_create(const osys_t *sys, int32_t nshards);
void o_mt_destroy(omt_t *m);
int32_t o_mt_add_walkers(omt_t *m, const uint64_t *dets, const double *sgns, const uint32_t *flags, int64_t n);
int64_t o_mt_get_walkers(omt_t *m, uint64_t *dets, double *sgns, uint32_t *flags, int64_t cap);
int32_t o_mt_set_core(omt_t *m, const uint64_t *dets, int64_t n, int32_t fp32, int64_t *nnz);
int32_t o_mt_iterate(omt_t *m, uint32_t iter, const double *diag_sft, ostats_t *out);

#ifdef __cplusplus
}
#endif
#endif
