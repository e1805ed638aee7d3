/*
 * oracle.c -- CPU restatement of NECI's FCIQMC iteration (spawn, merge,
 * death, annihilation, hash statistics).  TEST INFRASTRUCTURE ONLY: the
 * parity checker for the HIP engine and the CPU baseline of bench.py.
 *
 * Data structures deliberately mirror the reference's CPU path:
 *   CurrentDets rows (det words | signs | flags)      FciMCData.F90, BitReps.F90:164-214
 *   chained walker hash HashIndex + FreeSlot list      hash.F90, load_balancer.fpp:514-644
 *   SpawnedParts rows, sorted by ilut_lt and merged    Annihilation.F90:249-634
 * Floating-point expressions are written in the same order as the HIP engine
 * (both built with -ffp-contract=off) so iterations agree bit for bit.
 *
 * Flag word (engine-internal layout; converted to NECI bit positions only at
 * the upload/download boundary):  bit r < 16 initiator of replica r,
 * bit 16 core (deterministic) det, bit 17 removed slot.
 * Spawn info word: bits 0-7 replica, bit 8 parent initiator, bit 9 core parent.
 */
#define _POSIX_C_SOURCE 200809L  /* pthread barriers under -std=c11 */
#include "oracle.h"
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EPS 1e-13
#define FLAG_CORE (1u << 16)
#define FLAG_REMOVED (1u << 17)
#define SINFO_INIT (1u << 8)
#define SINFO_CORE_PARENT (1u << 9)

/* ------------------------------------------------------------------ RNG */

static inline uint32_t mulhilo32(uint32_t a, uint32_t b, uint32_t *hi)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    return (uint32_t)p;
}

/* Philox-4x32-10 (Salmon et al., SC'11): 10 rounds, multipliers
   0xD2511F53 / 0xCD9E8D57, Weyl key increments 0x9E3779B9 / 0xBB67AE85. */
void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint32_t hi0, hi1;
        uint32_t lo0 = mulhilo32(0xD2511F53u, c0, &hi0);
        uint32_t lo1 = mulhilo32(0xCD9E8D57u, c2, &hi1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t det_hash64(const uint64_t *d, int W)
{
    return W == 1 ? d[0] : (d[0] ^ mix64(d[1]));
}

static inline double u01(uint32_t lo, uint32_t hi)
{
    uint64_t x = ((uint64_t)hi << 32) | (uint64_t)lo;
    return (double)(x >> 11) * 0x1.0p-53;
}

static void draw2(const osys_t *s, const uint64_t *det, int W, uint32_t iter, int rep,
                  uint32_t k, uint32_t purpose, double *ua, double *ub)
{
    uint64_t dh = det_hash64(det, W);
    uint32_t ctr[4] = {(uint32_t)dh, (uint32_t)(dh >> 32), iter, (k << 3) | purpose};
    uint32_t key[2] = {s->seed, (uint32_t)rep};
    uint32_t o[4];
    o_philox4x32_10(ctr, key, o);
    *ua = u01(o[0], o[1]);
    *ub = u01(o[2], o[3]);
}

/* ------------------------------------------------------------ bit helpers */

static inline int nwords(int nbasis) { return nbasis / 64 + 1; } /* BitReps.F90:174 NIfD+1 */
static inline int is_beta(int orb) { return orb & 1; }            /* macros.h:21 */
static inline int spat(int orb) { return (orb + 1) / 2; }         /* gtID for RHF */
static inline int test_orb(const uint64_t *d, int orb)
{
    return (int)((d[(orb - 1) >> 6] >> ((orb - 1) & 63)) & 1ull);
}
static inline void set_orb(uint64_t *d, int orb) { d[(orb - 1) >> 6] |= 1ull << ((orb - 1) & 63); }
static inline void clr_orb(uint64_t *d, int orb) { d[(orb - 1) >> 6] &= ~(1ull << ((orb - 1) & 63)); }

/* decode_bit_det (BitReps.F90:931): ascending list of occupied orbitals */
int32_t o_decode(const uint64_t *ilut, int32_t W, int32_t *nI)
{
    int n = 0;
    for (int w = 0; w < W; ++w) {
        uint64_t x = ilut[w];
        while (x) {
            int b = __builtin_ctzll(x);
            nI[n++] = 64 * w + b + 1;
            x &= x - 1;
        }
    }
    return n;
}

static void encode(const int32_t *nI, int nel, int W, uint64_t *d)
{
    for (int w = 0; w < W; ++w) d[w] = 0;
    for (int i = 0; i < nel; ++i) set_orb(d, nI[i]);
}

/* ilut_lt (DetBitOps.F90:431-451): signed 64-bit words, word 0 first */
static inline int ilut_cmp(const uint64_t *a, const uint64_t *b, int W)
{
    for (int w = 0; w < W; ++w) {
        int64_t x = (int64_t)a[w], y = (int64_t)b[w];
        if (x < y) return -1;
        if (x > y) return 1;
    }
    return 0;
}

static inline int det_eq(const uint64_t *a, const uint64_t *b, int W)
{
    for (int w = 0; w < W; ++w)
        if (a[w] != b[w]) return 0;
    return 1;
}

/* FindBitExcitLevel (DetBitOps.F90:140-180): popcount(ilutI & ~ilutJ) */
static inline int excit_level(const uint64_t *a, const uint64_t *b, int W)
{
    int c = 0;
    for (int w = 0; w < W; ++w) c += __builtin_popcountll(a[w] & ~b[w]);
    return c;
}

/* ------------------------------------------------------------ hashing */

/* FindWalkerHash (hash.F90:23-37): h = 1099511628211*h + RandomHash2(nI(i))*i in
   wrapping int64; bucket = |mod(h, L)| + 1 with Fortran (truncating) mod. */
int32_t o_find_walker_hash(const int32_t *nI, int32_t nel, const int32_t *random_hash2, int64_t L)
{
    uint64_t h = 0;
    for (int i = 1; i <= nel; ++i) {
        int32_t term = random_hash2[nI[i - 1] - 1] * i; /* default-integer product */
        h = 1099511628211ull * h + (uint64_t)(int64_t)term;
    }
    int64_t m = (int64_t)h % L; /* C99 % truncates like Fortran mod */
    if (m < 0) m = -m;
    return (int32_t)(m + 1);
}

/* ------------------------------------------------------------ integrals */

/* UMatInd_base (UMatCache.F90:257-303), hermitian real case */
uint64_t o_umat_ind(int32_t i, int32_t j, int32_t k, int32_t l)
{
    int64_t A, B;
    if (i > k) A = ((int64_t)i * (i - 1)) / 2 + k; else A = ((int64_t)k * (k - 1)) / 2 + i;
    if (j > l) B = ((int64_t)j * (j - 1)) / 2 + l; else B = ((int64_t)l * (l - 1)) / 2 + j;
    if (A > B) return (uint64_t)((A * (A - 1)) / 2 + B);
    return (uint64_t)((B * (B - 1)) / 2 + A);
}

static inline double umat(const osys_t *s, int i, int j, int k, int l)
{
    return s->umat[o_umat_ind(i, j, k, l) - 1];
}
static inline double tmat(const osys_t *s, int i, int j)
{
    return s->tmat[(size_t)(i - 1) + (size_t)(j - 1) * (size_t)s->nbasis];
}

/* sltcnd_0_base (sltcnd.fpp:585-622) + ECore (Determinants.F90:544-548) */
static double sltcnd_0(const osys_t *s, const int32_t *nI)
{
    int nel = s->nel;
    double hel_sing = 0.0;
    for (int i = 0; i < nel; ++i) hel_sing += tmat(s, nI[i], nI[i]);
    double hel_doub = 0.0;
    for (int i = 0; i < nel - 1; ++i) {
        double t = 0.0;
        int idi = spat(nI[i]);
        for (int j = i + 1; j < nel; ++j) {
            int idj = spat(nI[j]);
            t += umat(s, idi, idj, idi, idj);
        }
        hel_doub += t;
    }
    double hel_tmp = 0.0;
    for (int i = 0; i < nel - 1; ++i)
        for (int j = i + 1; j < nel; ++j)
            if (is_beta(nI[i]) == is_beta(nI[j]))
                hel_tmp -= umat(s, spat(nI[i]), spat(nI[j]), spat(nI[j]), spat(nI[i]));
    return hel_doub + hel_tmp + hel_sing;
}

/* sltcnd_1_kernel (sltcnd.fpp:643-680) */
static double sltcnd_1(const osys_t *s, const int32_t *nI, int src, int tgt)
{
    int nel = s->nel;
    int a = spat(src), b = spat(tgt);
    double hel = 0.0;
    if (is_beta(src) == is_beta(tgt)) {
        for (int i = 0; i < nel; ++i)
            if (nI[i] != src) {
                int id = spat(nI[i]);
                hel += umat(s, a, id, b, id);
            }
        for (int i = 0; i < nel; ++i)
            if (nI[i] != src && is_beta(src) == is_beta(nI[i])) {
                int id = spat(nI[i]);
                hel -= umat(s, a, id, id, b);
            }
    }
    hel += tmat(s, src, tgt);
    return hel;
}

/* sltcnd_2_kernel (sltcnd.fpp:690-711): ex = [s1 s2 ; t1 t2] */
static double sltcnd_2(const osys_t *s, int s1, int s2, int t1, int t2)
{
    double hel;
    if (is_beta(s1) == is_beta(t1) && is_beta(s2) == is_beta(t2))
        hel = umat(s, spat(s1), spat(s2), spat(t1), spat(t2));
    else
        hel = 0.0;
    if (is_beta(s1) == is_beta(t2) && is_beta(s2) == is_beta(t1))
        hel -= umat(s, spat(s1), spat(s2), spat(t2), spat(t1));
    return hel;
}

/* ------------------------------------------------------ excitations */

/* make_single (excit_parity.F90:15-76), restated; returns tParity */
int32_t o_make_single(const int32_t *nI, int32_t nel, int32_t elec, int32_t tgt, int32_t *nJ,
                      int32_t *ex)
{
    /* 1-based elec */
    int i;
    for (int q = 0; q < nel; ++q) nJ[q] = nI[q];
    int src = nI[elec - 1];
    ex[0] = src;
    ex[1] = tgt;
    if (src < tgt) {
        for (i = elec + 1; i <= nel; ++i) {
            if (tgt < nJ[i - 1]) { nJ[i - 2] = tgt; break; }
            nJ[i - 2] = nJ[i - 1];
        }
        if (i == nel + 1) nJ[nel - 1] = tgt;
    } else {
        for (i = elec - 1; i >= 1; --i) {
            if (tgt > nJ[i - 1]) { nJ[i] = tgt; break; }
            nJ[i] = nJ[i - 1];
        }
        if (i == 0) nJ[0] = tgt;
    }
    int d = elec - i;
    return (d % 2) == 0;
}

/* make_double (excit_parity.F90:78-172), restated; ex = [s1 s2 t1 t2]; returns tParity */
int32_t o_make_double(const int32_t *nI, int32_t nel, int32_t e1, int32_t e2, int32_t t1,
                      int32_t t2, int32_t *nJ, int32_t *ex)
{
    int elecs[2] = {e1 < e2 ? e1 : e2, e1 < e2 ? e2 : e1};
    int tgts[2] = {t1 < t2 ? t1 : t2, t1 < t2 ? t2 : t1};
    int srcs[2] = {nI[elecs[0] - 1], nI[elecs[1] - 1]};
    ex[0] = srcs[0]; ex[1] = srcs[1]; ex[2] = tgts[0]; ex[3] = tgts[1];
    for (int q = 0; q < nel; ++q) nJ[q] = nI[q];
    if (srcs[0] < tgts[0] && srcs[1] < tgts[0]) elecs[1] -= 1;
    int pos_moved = 0, i;
    for (int k = 0; k < 2; ++k) {
        if (srcs[k] < tgts[k]) {
            if (elecs[k] == nel) {
                i = nel + 1;
                nJ[nel - 1] = tgts[k];
            } else {
                for (i = elecs[k] + 1; i <= nel; ++i) {
                    if (tgts[k] < nJ[i - 1]) { nJ[i - 2] = tgts[k]; break; }
                    nJ[i - 2] = nJ[i - 1];
                }
                if (i == nel + 1) nJ[nel - 1] = tgts[k];
            }
        } else {
            if (elecs[k] == 1) {
                i = 0;
                nJ[0] = tgts[k];
            } else {
                for (i = elecs[k] - 1; i >= 1; --i) {
                    if (tgts[k] > nJ[i - 1]) { nJ[i] = tgts[k]; break; }
                    nJ[i] = nJ[i - 1];
                }
                if (i == 0) nJ[0] = tgts[k];
            }
        }
        pos_moved += elecs[k] - i + 1;
    }
    return pos_moved & 1;
}

/* get_helement_det_only (Determinants.F90:508-548) / get_helement_lattice
   (real_space_hubbard.F90:2196-2221) from an excitation matrix */
double o_get_helement_ex(const osys_t *s, const int32_t *nI, int32_t ic, const int32_t *ex,
                         int32_t tparity)
{
    double hel;
    if (s->model == O_MODEL_HUBBARD) {
        if (ic == 0) return o_diag_helement(s, nI);
        if (ic != 1) return 0.0;
        int a = ex[0], b = ex[1];
        /* TMAT2D = bhub between same-spin orbitals of neighbouring sites */
        if (is_beta(a) != is_beta(b)) return 0.0;
        int sa = spat(a), sb = spat(b);
        hel = 0.0;
        for (int k = 0; k < s->maxnbr; ++k)
            if (s->nbr[(sa - 1) * s->maxnbr + k] == sb) { hel = s->bhub; break; }
        if (tparity) hel = -hel;
        return hel;
    }
    if (ic == 0) return sltcnd_0(s, nI) + s->ecore;
    if (ic == 1) hel = sltcnd_1(s, nI, ex[0], ex[1]);
    else if (ic == 2) hel = sltcnd_2(s, ex[0], ex[1], ex[2], ex[3]);
    else return 0.0;
    if (tparity) hel = -hel;
    return hel;
}

/* excitation nI -> nJ: sources ascending, targets ascending, parity of the
   paired moves (equivalent to GetExcitation + make_double ordering) */
static int get_excitation(const int32_t *nI, const int32_t *nJ, int nel, int W, int32_t *ex,
                          int32_t *tparity)
{
    uint64_t dI[O_MAX_W], dJ[O_MAX_W];
    encode(nI, nel, W, dI);
    encode(nJ, nel, W, dJ);
    int ic = excit_level(dI, dJ, W);
    if (ic == 0 || ic > 2) { *tparity = 0; return ic; }
    int ns = 0, nt = 0, src[2], tgt[2];
    for (int i = 0; i < nel; ++i) if (!test_orb(dJ, nI[i])) src[ns++] = nI[i];
    for (int i = 0; i < nel; ++i) if (!test_orb(dI, nJ[i])) tgt[nt++] = nJ[i];
    int32_t nT[O_MAX_NEL], exd[4];
    if (ic == 1) {
        int elec = 0;
        for (int i = 0; i < nel; ++i) if (nI[i] == src[0]) elec = i + 1;
        *tparity = o_make_single(nI, nel, elec, tgt[0], nT, ex);
    } else {
        int e1 = 0, e2 = 0;
        for (int i = 0; i < nel; ++i) {
            if (nI[i] == src[0]) e1 = i + 1;
            if (nI[i] == src[1]) e2 = i + 1;
        }
        *tparity = o_make_double(nI, nel, e1, e2, tgt[0], tgt[1], nT, exd);
        ex[0] = exd[0]; ex[1] = exd[1]; ex[2] = exd[2]; ex[3] = exd[3];
    }
    return ic;
}

double o_get_helement(const osys_t *s, const int32_t *nI, const int32_t *nJ, int32_t nel)
{
    int32_t ex[4], tp;
    int ic = get_excitation(nI, nJ, nel, nwords(s->nbasis), ex, &tp);
    if (ic > 2) return 0.0;
    return o_get_helement_ex(s, nI, ic, ex, tp);
}

/* get_diagonal_matel (matel_getter.F90:27): total H_ii */
double o_diag_helement(const osys_t *s, const int32_t *nI)
{
    if (s->model == O_MODEL_HUBBARD) {
        /* get_diag_helemen_rs_hub (real_space_hubbard.F90:2303-2321): U * doubly occupied */
        int nd = 0;
        for (int i = 0; i + 1 < s->nel; ++i)
            if (is_beta(nI[i]) && nI[i + 1] == nI[i] + 1) nd++;
        return s->uhub * (double)nd;
    }
    return sltcnd_0(s, nI) + s->ecore;
}

/* get_off_diagonal_matel (matel_getter.F90:56-107) with tNoBrillouin: H(det, ref) for
   excitation level 1 or 2 */
double o_offdiag_helement(const osys_t *s, const uint64_t *ilut)
{
    int W = nwords(s->nbasis);
    int ex = excit_level(s->ref, ilut, W);
    if (ex != 1 && ex != 2) return 0.0;
    int32_t nI[O_MAX_NEL], n0[O_MAX_NEL];
    o_decode(ilut, W, nI);
    o_decode(s->ref, W, n0);
    return o_get_helement(s, nI, n0, s->nel);
}

/* ----------------------------------------------------- alias sampling */

/* init_AliasTable_t (aliasSampling.F90:209-297) + init_probs (:412-418) */
void o_alias_build(const double *w, int32_t n, double *bias, int32_t *alias, double *prob)
{
    double sum = 0.0;
    for (int i = 0; i < n; ++i) sum += w[i];
    int *over = (int *)malloc(sizeof(int) * (size_t)n), *under = (int *)malloc(sizeof(int) * (size_t)n);
    int cV = 0, cU = 0;
    for (int i = 0; i < n; ++i) bias[i] = w[i] / sum * (double)n;
    for (int i = 0; i < n; ++i) {
        if (bias[i] > 1.0) over[cV++] = i; else under[cU++] = i;
    }
    for (int a = 0, b = cV - 1; a < b; ++a, --b) { int t = over[a]; over[a] = over[b]; over[b] = t; }
    while (cV != 0 && cU != 0) {
        int i = over[cV - 1], j = under[cU - 1];
        alias[j] = i;
        bias[i] = bias[i] + bias[j] - 1.0;
        cU--; cV--;
        if (bias[i] > 1.0) over[cV++] = i; else under[cU++] = i;
    }
    for (int q = 0; q < cV; ++q) { bias[over[q]] = 1.0; alias[over[q]] = over[q]; }
    for (int q = 0; q < cU; ++q) { bias[under[q]] = 1.0; alias[under[q]] = under[q]; }
    for (int i = 0; i < n; ++i) prob[i] = w[i] / sum;
    free(over);
    free(under);
}

/* sample_AliasTable_t (aliasSampling.F90:310-332), 0-based result */
static inline int alias_sample(const double *bias, const int32_t *alias, int n, double r)
{
    double x = (double)n * r;
    int pos = (int)x + 1;
    double b = x + 1.0 - (double)pos;
    if (b < 0.0) b = 0.0;
    return (b < bias[pos - 1]) ? pos - 1 : alias[pos - 1];
}

/* PCHB tables: AB rows per IJ (fuseIndex over spin orbitals) and J rows per I */
static struct {
    int nbi, ijmax, holes;
    double *ab_bias, *ab_prob;
    int32_t *ab_alias;
    uint8_t *ab_ok;
    double *j_bias, *j_prob;
    int32_t *j_alias;
    uint8_t *j_ok;
    /* FULL-FULL particles: first particle I over all spin orbitals */
    double *i_bias, *i_prob;
    int32_t *i_alias;
    uint8_t i_ok;
    /* FULL-FULL holes: first hole A per IJ, row (IJ-1); second hole B per
       (IJ, A), row (IJ-1)*nbi + (A-1); each row over the nbi spin orbitals */
    double *a_bias, *a_prob, *b_bias, *b_prob;
    int32_t *a_alias, *b_alias;
    uint8_t *a_ok, *b_ok;
} P;

static inline int fuse(int x, int y) { return x < y ? x + y * (y - 1) / 2 : y + x * (x - 1) / 2; }

void o_free_pchb(void)
{
    free(P.ab_bias); free(P.ab_prob); free(P.ab_alias); free(P.ab_ok);
    free(P.j_bias); free(P.j_prob); free(P.j_alias); free(P.j_ok);
    free(P.i_bias); free(P.i_prob); free(P.i_alias);
    free(P.a_bias); free(P.a_prob); free(P.a_alias); free(P.a_ok);
    free(P.b_bias); free(P.b_prob); free(P.b_alias); free(P.b_ok);
    memset(&P, 0, sizeof(P));
}

/* alias table and probabilities of one sampler row; returns 0 (row left
   empty, every probability 0) when the weights sum to zero: init_AliasSampler_t
   (aliasSampling.F90:376-398) then leaves the sampler unassociated */
static int alias_row(const double *w, int n, double *bias, int32_t *alias, double *prob)
{
    double sw = 0.0;
    for (int q = 0; q < n; ++q) sw += w[q];
    if (fabs(sw) <= EPS) return 0;
    o_alias_build(w, n, bias, alias, prob);
    return 1;
}

/* PCHB weight |<IJ||AB>| of a spin-allowed double, 0 otherwise
   (get_PCHB_weight, gasci_pchb_doubles_select_particles.fpp:596-609) */
static double pchb_weight(const osys_t *s, int I, int J, int A, int B)
{
    int t1 = A < B ? A : B, t2 = A < B ? B : A;
    int ms_s = (is_beta(I) ? -1 : 1) + (is_beta(J) ? -1 : 1);
    int ms_t = (is_beta(t1) ? -1 : 1) + (is_beta(t2) ? -1 : 1);
    return ms_s == ms_t ? fabs(sltcnd_2(s, I, J, t1, t2)) : 0.0;
}

/* GAS_doubles_PCHB_compute_samplers of the fully weighted holes
   (gasci_pchb_doubles_spinorb_fullyweighted.fpp:317-391): per (I < J), per A,
   the B row w_B(B) = |<IJ||AB>|; the A row w_A(A) = sum(w_B); the particle
   weight IJ_weights(I, J) = sum(w_A). */
static void build_full_holes(const osys_t *s, double *ijw)
{
    int n = P.nbi, ijmax = P.ijmax;
    P.a_bias = (double *)calloc((size_t)ijmax * n, sizeof(double));
    P.a_prob = (double *)calloc((size_t)ijmax * n, sizeof(double));
    P.a_alias = (int32_t *)calloc((size_t)ijmax * n, sizeof(int32_t));
    P.a_ok = (uint8_t *)calloc((size_t)ijmax, 1);
    P.b_bias = (double *)calloc((size_t)ijmax * n * n, sizeof(double));
    P.b_prob = (double *)calloc((size_t)ijmax * n * n, sizeof(double));
    P.b_alias = (int32_t *)calloc((size_t)ijmax * n * n, sizeof(int32_t));
    P.b_ok = (uint8_t *)calloc((size_t)ijmax * n, 1);
    double *wa = (double *)malloc(sizeof(double) * (size_t)n);
    double *wb = (double *)malloc(sizeof(double) * (size_t)n);
    for (int I = 1; I <= n; ++I)
        for (int J = I + 1; J <= n; ++J) {
            int IJ = fuse(I, J);
            for (int q = 0; q < n; ++q) wa[q] = 0.0;
            for (int A = 1; A <= n; ++A) {
                if (A == I || A == J) continue;
                for (int q = 0; q < n; ++q) wb[q] = 0.0;
                for (int B = 1; B <= n; ++B) {
                    if (B == I || B == J || B == A) continue;
                    wb[B - 1] = pchb_weight(s, I, J, A, B);
                }
                size_t row = (size_t)(IJ - 1) * n + (A - 1);
                P.b_ok[row] = (uint8_t)alias_row(wb, n, P.b_bias + row * n, P.b_alias + row * n,
                                                 P.b_prob + row * n);
                double sw = 0.0;
                for (int q = 0; q < n; ++q) sw += wb[q];
                wa[A - 1] = sw;
            }
            size_t row = (size_t)(IJ - 1);
            P.a_ok[row] = (uint8_t)alias_row(wa, n, P.a_bias + row * n, P.a_alias + row * n, P.a_prob + row * n);
            double sw = 0.0;
            for (int q = 0; q < n; ++q) sw += wa[q];
            ijw[(I - 1) + (size_t)(J - 1) * n] = sw;
            ijw[(J - 1) + (size_t)(I - 1) * n] = sw;
        }
    free(wa);
    free(wb);
}

/* GAS_doubles_PCHB_spinorb_compute_samplers (gasci_pchb_doubles_spinorb_fastweighted.fpp:280-347)
   for the FCI (single supergroup) case, and the J sampler of
   init_PC_WeightedParticles_t (gasci_pchb_doubles_select_particles.fpp:254-301) */
int32_t o_build_pchb(osys_t *s)
{
    o_free_pchb();
    int n = s->nbasis;
    int ijmax = fuse(n, n);
    P.nbi = n;
    P.ijmax = ijmax;
    P.holes = s->holes;
    size_t sz = (size_t)ijmax * (size_t)ijmax;
    P.ab_bias = (double *)calloc(sz, sizeof(double));
    P.ab_prob = (double *)calloc(sz, sizeof(double));
    P.ab_alias = (int32_t *)calloc(sz, sizeof(int32_t));
    P.ab_ok = (uint8_t *)calloc((size_t)ijmax, 1);
    P.j_bias = (double *)calloc((size_t)n * n, sizeof(double));
    P.j_prob = (double *)calloc((size_t)n * n, sizeof(double));
    P.j_alias = (int32_t *)calloc((size_t)n * n, sizeof(int32_t));
    P.j_ok = (uint8_t *)calloc((size_t)n, 1);
    double *w = (double *)malloc(sizeof(double) * (size_t)ijmax);
    double *ijw = (double *)calloc((size_t)n * n, sizeof(double));
    if (s->holes == O_HOLES_FULL_FULL) build_full_holes(s, ijw);
    else for (int I = 1; I <= n; ++I)
        for (int J = I + 1; J <= n; ++J) {
            for (int q = 0; q < ijmax; ++q) w[q] = 0.0;
            for (int A = 1; A <= n; ++A) {
                if (A == I || A == J) continue;
                for (int B = 1; B <= n; ++B) {
                    if (B == I || B == J || B == A) continue;
                    int AB = fuse(A, B);
                    int t1 = A < B ? A : B, t2 = A < B ? B : A;
                    /* spin_allowed: sum of Ms conserved */
                    int ms_s = (is_beta(I) ? -1 : 1) + (is_beta(J) ? -1 : 1);
                    int ms_t = (is_beta(t1) ? -1 : 1) + (is_beta(t2) ? -1 : 1);
                    if (ms_s == ms_t) w[AB - 1] = fabs(sltcnd_2(s, I, J, t1, t2));
                }
            }
            int IJ = fuse(I, J);
            double sw = 0.0;
            for (int q = 0; q < ijmax; ++q) sw += w[q];
            if (!(fabs(sw) <= EPS)) {
                size_t off = (size_t)(IJ - 1) * (size_t)ijmax;
                o_alias_build(w, ijmax, P.ab_bias + off, P.ab_alias + off, P.ab_prob + off);
                P.ab_ok[IJ - 1] = 1;
            }
            ijw[(I - 1) + (size_t)(J - 1) * n] = sw;
            ijw[(J - 1) + (size_t)(I - 1) * n] = sw;
        }
    for (int I = 1; I <= n; ++I) {
        const double *row = ijw + (size_t)(I - 1) * n; /* weights(:, I) */
        double sw = 0.0;
        for (int q = 0; q < n; ++q) sw += row[q];
        if (!(fabs(sw) <= EPS)) {
            size_t off = (size_t)(I - 1) * n;
            o_alias_build(row, n, P.j_bias + off, P.j_alias + off, P.j_prob + off);
            P.j_ok[I - 1] = 1;
        }
    }
    /* the I sampler of the fully weighted particles: sum(weights(:, :), dim=1)
       (init_PC_WeightedParticles_t, gasci_pchb_doubles_select_particles.fpp:287-297) */
    P.i_bias = (double *)calloc((size_t)n, sizeof(double));
    P.i_prob = (double *)calloc((size_t)n, sizeof(double));
    P.i_alias = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    {
        double *wi = (double *)calloc((size_t)n, sizeof(double));
        for (int I = 1; I <= n; ++I) {
            double sw = 0.0;
            for (int q = 0; q < n; ++q) sw += ijw[(size_t)(I - 1) * n + q];
            wi[I - 1] = sw;
        }
        P.i_ok = (uint8_t)alias_row(wi, n, P.i_bias, P.i_alias, P.i_prob);
        free(wi);
    }
    free(w);
    free(ijw);
    s->pchb_ready = 1;
    return 0;
}

int32_t o_get_pchb_row(int32_t which, int32_t row, double *bias, int32_t *alias, double *prob)
{
    int n = which == 0 ? P.ijmax : P.nbi;
    size_t off = (size_t)(row - 1) * n;
    const double *b = which == 0 ? P.ab_bias : P.j_bias;
    const double *p = which == 0 ? P.ab_prob : P.j_prob;
    const int32_t *a = which == 0 ? P.ab_alias : P.j_alias;
    const uint8_t *ok = which == 0 ? P.ab_ok : P.j_ok;
    memcpy(bias, b + off, sizeof(double) * n);
    memcpy(prob, p + off, sizeof(double) * n);
    memcpy(alias, a + off, sizeof(int32_t) * n);
    return ok[row - 1];
}

/* -------------------------------------------------- excitation generation */

/* gen_excit_rs_hubbard (real_space_hubbard.F90:1938-2033) with
   create_cum_list_rs_hubbard (:2080-2142) and pick_from_cum_list
   (lattice_models_utils.F90:150). */
static int gen_hubbard(const osys_t *s, const uint64_t *ilut, const int32_t *nI, const double *d,
                       int32_t *ex, int32_t *nJ, double *pgen)
{
    int nel = s->nel;
    int elec = 1 + (int)(d[0] * (double)nel);
    double p_elec = 1.0 / (double)nel;
    int src = nI[elec - 1];
    int site = spat(src);
    double cum[16], cum_sum = 0.0;
    int orbs[16], nn = 0;
    for (int k = 0; k < s->maxnbr; ++k) {
        int ns = s->nbr[(site - 1) * s->maxnbr + k];
        if (ns == 0) break;
        int orb = is_beta(src) ? 2 * ns - 1 : 2 * ns;
        double elem = 0.0;
        if (!test_orb(ilut, orb)) elem = fabs(s->bhub);
        cum_sum += elem;
        cum[nn] = cum_sum;
        orbs[nn++] = orb;
    }
    if (cum_sum < EPS) return 0;
    double r = d[1] * cum_sum;
    int ind = 0;
    while (ind < nn - 1 && cum[ind] < r) ind++;
    double prev = ind == 0 ? 0.0 : cum[ind - 1];
    if (cum[ind] - prev == 0.0) return 0; /* zero-weight pick: only at r == cum exactly */
    double p_orb = (cum[ind] - prev) / cum_sum;
    *pgen = p_elec * p_orb;
    return 1 + (o_make_single(nI, nel, elec, orbs[ind], nJ, ex) ? 2 : 0);
}

/* The weight |H(src -> orb)| of gen_single_4ind_ex (excit_gens_int_weighted.F90:
   442-620, there sltcnd_1 of the determinant itself) evaluated relative to the
   reference: sltcnd_1(ref, src, orb), plus U(a,j,b,j) - [same spin]
   U(a,j,j,b) for every orbital j the determinant occupies and the reference
   does not (ascending, src skipped), minus the same for every orbital the
   reference occupies and the determinant does not.  Equal to sltcnd_1(nI, src,
   orb) up to FP64 rounding; the device's hb_weight sums in this order.  The
   signed value is also the element a heat-bath single spawns with. */
static double hb_weight(const osys_t *s, const uint64_t *ilut, int src, int orb)
{
    int W = nwords(s->nbasis);
    int32_t nref[O_MAX_NEL];
    o_decode(s->ref, W, nref);
    double v = sltcnd_1(s, nref, src, orb);
    int a = spat(src), b = spat(orb);
    for (int pass = 0; pass < 2; ++pass)
        for (int o = 1; o <= s->nbasis; ++o) {
            if (o == src) continue;
            int in_d = test_orb(ilut, o), in_ref = test_orb(s->ref, o);
            if (pass == 0 && in_d && !in_ref) {
                v += umat(s, a, spat(o), b, spat(o));
                if (is_beta(o) == is_beta(src)) v -= umat(s, a, spat(o), spat(o), b);
            } else if (pass == 1 && !in_d && in_ref) {
                v -= umat(s, a, spat(o), b, spat(o));
                if (is_beta(o) == is_beta(src)) v += umat(s, a, spat(o), spat(o), b);
            }
        }
    return v;
}

double o_hb_weight(const osys_t *s, const uint64_t *ilut, int32_t src, int32_t orb)
{
    return hb_weight(s, ilut, src, orb);
}

/* gen_single_4ind_ex + select_orb_sing (excit_gens_int_weighted.F90:442-620),
   weights by hb_weight */
static int gen_single_heatbath(const osys_t *s, const uint64_t *ilut, const int32_t *nI,
                               double dsel, double dorb, int32_t *ex, int32_t *nJ, double *pgen)
{
    int nel = s->nel;
    int elec = 1 + (int)floor(dsel * (double)nel);
    int src = nI[elec - 1];
    int nsp = s->nbasis / 2;
    double cum[O_MAX_NEL * 4], cpt_arr[O_MAX_NEL * 4], cum_sum = 0.0;
    for (int i = 0; i < nsp; ++i) {
        int orb = 2 * (i + 1) - (is_beta(src) ? 1 : 0);
        double cpt = 0.0;
        if (!test_orb(ilut, orb)) cpt = fabs(hb_weight(s, ilut, src, orb));
        cpt_arr[i] = cpt;
        cum_sum += cpt;
        cum[i] = cum_sum;
    }
    if (fabs(cum_sum) <= EPS) return 0;
    double r = dorb * cum_sum;
    int idx = 0;
    while (idx < nsp - 1 && cum[idx] < r) idx++;
    if (cpt_arr[idx] == 0.0) return 0;
    int orb = 2 * (idx + 1) - (is_beta(src) ? 1 : 0);
    double pg = cpt_arr[idx] / cum_sum;
    pg = pg / (double)nel;
    *pgen = pg;
    return 1 + (o_make_single(nI, nel, elec, orb, nJ, ex) ? 2 : 0);
}

/* pick_biased_elecs (excit_gens_int_weighted.F90:722-848) */
static int pick_biased_elecs(const osys_t *s, const int32_t *nI, double r, int *elecs,
                             double *pgen)
{
    int nel = s->nel;
    int nalpha = 0, nbeta = 0;
    for (int i = 0; i < nel; ++i) { if (is_beta(nI[i])) nbeta++; else nalpha++; }
    int aa = nalpha * (nalpha - 1) / 2, bb = nbeta * (nbeta - 1) / 2;
    int par = aa + bb, ab = nalpha * nbeta;
    double pb = s->pparallel;
    int al_req, be_req, al_num[2] = {0, 0}, be_num[2] = {0, 0};
    if (r < pb ? par == 0 : ab == 0) return 0; /* no pair of the drawn spin type */
    if (r < pb) {
        r = (r / pb) * (double)par;
        int idx = (int)floor(r);
        if (idx < aa) {
            al_req = 2; be_req = 0;
            al_num[0] = (int)ceil((1.0 + sqrt(9.0 + 8.0 * (double)idx)) / 2.0);
            al_num[1] = idx + 1 - ((al_num[0] - 1) * (al_num[0] - 2)) / 2;
        } else {
            al_req = 0; be_req = 2;
            idx -= aa;
            be_num[0] = (int)ceil((1.0 + sqrt(9.0 + 8.0 * (double)idx)) / 2.0);
            be_num[1] = idx + 1 - ((be_num[0] - 1) * (be_num[0] - 2)) / 2;
        }
        *pgen = pb / (double)par;
    } else {
        al_req = 1; be_req = 1;
        *pgen = (1.0 - pb) / (double)ab;
        r = ((r - pb) / (1.0 - pb)) * (double)ab;
        int idx = (int)floor(r);
        al_num[0] = 1 + idx % nalpha;
        be_num[0] = 1 + (int)floor((double)idx / (double)nalpha);
    }
    int al_count = 0, be_count = 0, found = 0;
    for (int i = 1; i <= nel; ++i) {
        if (!is_beta(nI[i - 1])) {
            al_count++;
            if (al_req > 0 && al_count == al_num[al_req - 1]) { elecs[found++] = i; al_req--; }
        } else {
            be_count++;
            if (be_req > 0 && be_count == be_num[be_req - 1]) { elecs[found++] = i; be_req--; }
        }
        if (al_req == 0 && be_req == 0) break;
    }
    return 1;
}

/* GAS_doubles_PCHB_spinorb_gen_exc (gasci_pchb_doubles_spinorb_fastweighted.fpp:131-244)
   with UNIF-FAST (draw_PC_FastWeightedParticles_t, gasci_pchb_doubles_select_particles.fpp:515)
   or UNIF-UNIF particles */
static int gen_double_pchb(const osys_t *s, const uint64_t *ilut, const int32_t *nI, const double *d,
                           int32_t *ex, int32_t *nJ, double *pgen)
{
    int nel = s->nel, n = P.nbi;
    int elecs[2], srcs[2];
    double p;
    if (s->particles == O_PART_UNIF_FAST) {
        elecs[0] = (int)(d[1] * (double)nel) + 1;
        srcs[0] = nI[elecs[0] - 1];
        if (!P.j_ok[srcs[0] - 1]) return 0;
        size_t off = (size_t)(srcs[0] - 1) * n;
        int j = alias_sample(P.j_bias + off, P.j_alias + off, n, d[2]);
        double p_J_I = P.j_prob[off + j];
        srcs[1] = j + 1;
        if (!test_orb(ilut, srcs[1])) return 0;
        elecs[1] = 0;
        for (int q = 0; q < nel; ++q) if (nI[q] == srcs[1]) elecs[1] = q + 1;
        double p_I_J = P.j_ok[srcs[1] - 1] ? P.j_prob[(size_t)(srcs[1] - 1) * n + (srcs[0] - 1)] : 0.0;
        p = (p_J_I + p_I_J) / (double)nel;
        if (srcs[0] > srcs[1]) {
            int t = srcs[0]; srcs[0] = srcs[1]; srcs[1] = t;
            t = elecs[0]; elecs[0] = elecs[1]; elecs[1] = t;
        }
    } else {
        if (!pick_biased_elecs(s, nI, d[1], elecs, &p)) return 0;
        srcs[0] = nI[elecs[0] - 1];
        srcs[1] = nI[elecs[1] - 1];
    }
    int ij = fuse(srcs[0], srcs[1]);
    if (!P.ab_ok[ij - 1]) return 0;
    size_t off = (size_t)(ij - 1) * P.ijmax;
    int ab = alias_sample(P.ab_bias + off, P.ab_alias + off, P.ijmax, d[3]) + 1;
    double pgh = P.ab_prob[off + ab - 1];
    /* tgtOrbs(:, AB) = (B, A) with B <= A */
    int A = (int)floor((1.0 + sqrt(8.0 * (double)ab - 7.0)) / 2.0);
    while (A * (A - 1) / 2 >= ab) A--;
    while ((A + 1) * A / 2 < ab) A++;
    int B = ab - A * (A - 1) / 2;
    int t1 = B, t2 = A;
    if (t1 == t2 || test_orb(ilut, t1) || test_orb(ilut, t2)) return 0;
    if (fabs(pgh) <= EPS) return 0;
    int tp = o_make_double(nI, nel, elecs[0], elecs[1], t1, t2, nJ, ex);
    *pgen = p * pgh;
    return 1 + (tp ? 2 : 0);
}

/* One excitation: returns 0 for a null excitation, else fills outputs. */
/* ---------------------------------------------- uniform molecular generator */

/* The stream of uniforms a generator with rejection loops consumes: d[1..4] of
   the attempt, then purpose-6 Philox blocks (two uniforms each) keyed by
   (seed, replica | (block + 1) << 8). */
typedef struct {
    const osys_t *s;
    const double *d;
    const uint64_t *det;
    uint32_t iter, k;
    int rep, n, blk;
    double c[2];
} draw_seq_t;

static double seq_next(draw_seq_t *q)
{
    q->n++;
    if (q->n <= 4) return q->d[q->n];
    int j = q->n - 5, b = j / 2;
    if (b != q->blk) {
        int W = nwords(q->s->nbasis);
        uint64_t dh = det_hash64(q->det, W);
        uint32_t ctr[4] = {(uint32_t)dh, (uint32_t)(dh >> 32), q->iter, (q->k << 3) | 6u};
        uint32_t key[2] = {q->s->seed, (uint32_t)q->rep | ((uint32_t)(b + 1) << 8)};
        uint32_t o[4];
        o_philox4x32_10(ctr, key, o);
        q->c[0] = u01(o[0], o[1]);
        q->c[1] = u01(o[2], o[3]);
        q->blk = b;
    }
    return q->c[j & 1];
}

/* ------------------------------------ weighted PCHB particles and holes */

static inline int near_zero_d(double x) { return fabs(x) <= EPS; }

/* redraws of a constrained alias draw before it is given up as a null
   excitation (the reference redraws without bound; each redraw succeeds
   with probability >= 0.1, so the bound is never reached in practice) */
#define O_MAX_REDRAW 1024

/* constrained_sample_fast (aliasSampling.F90:505-535) of one sampler row
   (bias, alias, prob over n entries; prob == NULL: an empty sampler) restricted
   to the ascending set contain[0..nc) whose bitmask is mask, with the given
   renormalisation: below redrawing_cutoff = 0.1 (:152) the CDF sampler over
   prob(contain)/renorm (CDF_sampling.fpp:56-110: first cumulated value >= r),
   else alias draws until one lies in the set.  Returns the drawn orbital (0:
   none), its position in contain (1-based) and its probability.  A draw whose
   probability is zero (a cumulated value at r == 0, or r beyond a rounded
   last cumulated value, where the reference's search returns -1) is given up
   as a null excitation. */
static int constrained_sample(const double *bias, const int32_t *alias, const double *prob, int n, const int *contain,
                              int nc, const uint64_t *mask, double renorm, draw_seq_t *q, int *pos, double *p)
{
    if (!prob || near_zero_d(renorm)) return 0;
    if (renorm < 0.1) {
        double r = seq_next(q), cum = 0.0;
        for (int k = 0; k < nc; ++k) {
            double pk = prob[contain[k] - 1] / renorm;
            cum += pk;
            if (!(cum < r)) {
                if (pk == 0.0) return 0;
                *pos = k + 1;
                *p = pk;
                return contain[k];
            }
        }
        return 0;
    }
    for (int t = 0; t < O_MAX_REDRAW; ++t) {
        int tgt = alias_sample(bias, alias, n, seq_next(q)) + 1;
        if (!test_orb(mask, tgt)) continue;
        int k = 0;
        while (contain[k] != tgt) ++k;
        *pos = k + 1;
        *p = prob[tgt - 1] / renorm;
        return *p == 0.0 ? 0 : tgt;
    }
    return 0;
}

/* sum of a row's probabilities over a set, in the set's order (sum(get_prob(..., contain))) */
static double row_sum(const double *prob, const int *set, int ns)
{
    double r = 0.0;
    if (!prob) return 0.0;
    for (int k = 0; k < ns; ++k) r += prob[set[k] - 1];
    return r;
}

static const double *j_row(int I) { return P.j_ok[I - 1] ? P.j_prob + (size_t)(I - 1) * P.nbi : NULL; }

/* draw_PC_WeightedParticles_t (UNIF-FULL, gasci_pchb_doubles_select_particles.fpp:427-470)
   and draw_PC_FullyWeightedParticles_t (FULL-FULL, :333-378): the ordered pair
   (srcs[0], srcs[1]) and p({I, J}); 0 for a null draw */
static int draw_weighted_particles(const osys_t *s, const uint64_t *ilut, const int32_t *nI, draw_seq_t *q,
                                   int *elecs, int *srcs, double *p)
{
    const int nel = s->nel, n = P.nbi, full = s->particles == O_PART_FULL_FULL;
    int set[O_MAX_NEL];
    for (int k = 0; k < nel; ++k) set[k] = nI[k];
    double pf[2] = {0.0, 0.0}, ps[2] = {0.0, 0.0}, rf = 0.0;
    int pos = 0;
    if (full) {
        const double *ip = P.i_ok ? P.i_prob : NULL;
        rf = row_sum(ip, set, nel);
        srcs[0] = constrained_sample(P.i_bias, P.i_alias, ip, n, set, nel, ilut, rf, q, &pos, &pf[0]);
        if (!srcs[0]) return 0;
        elecs[0] = pos;
    } else {
        elecs[0] = (int)(seq_next(q) * (double)nel) + 1;
        srcs[0] = nI[elecs[0] - 1];
    }
    const double *j1 = j_row(srcs[0]);
    double rs1 = row_sum(j1, set, nel);
    size_t o1 = (size_t)(srcs[0] - 1) * n;
    srcs[1] = constrained_sample(P.j_bias + o1, P.j_alias + o1, j1, n, set, nel, ilut, rs1, q, &pos, &ps[0]);
    if (!srcs[1]) return 0;
    elecs[1] = pos;
    const double *j2 = j_row(srcs[1]);
    double rs2 = row_sum(j2, set, nel);
    ps[1] = (!j2 || near_zero_d(rs2)) ? 0.0 : j2[srcs[0] - 1] / rs2;
    if (full) {
        pf[1] = near_zero_d(rf) ? 0.0 : P.i_prob[srcs[1] - 1] / rf;
        *p = pf[0] * ps[0] + pf[1] * ps[1];
    } else {
        *p = (ps[0] + ps[1]) / (double)nel;
    }
    if (srcs[0] > srcs[1]) {
        int t = srcs[0]; srcs[0] = srcs[1]; srcs[1] = t;
        t = elecs[0]; elecs[0] = elecs[1]; elecs[1] = t;
    }
    return 1;
}

/* get_pgen_PC_WeightedParticles_t (:472-500) / get_pgen_PC_FullyWeightedParticles_t (:381-425) */
static double pgen_weighted_particles(const osys_t *s, const int32_t *nI, int I, int J)
{
    const int nel = s->nel, full = s->particles == O_PART_FULL_FULL;
    int set[O_MAX_NEL];
    for (int k = 0; k < nel; ++k) set[k] = nI[k];
    const double *jI = j_row(I), *jJ = j_row(J);
    double rI = row_sum(jI, set, nel), rJ = row_sum(jJ, set, nel);
    double psI = (!jI || near_zero_d(rI)) ? 0.0 : jI[J - 1] / rI;
    double psJ = (!jJ || near_zero_d(rJ)) ? 0.0 : jJ[I - 1] / rJ;
    if (!full) return (psI + psJ) / (double)nel;
    const double *ip = P.i_ok ? P.i_prob : NULL;
    double rf = row_sum(ip, set, nel);
    double pfI = (!ip || near_zero_d(rf)) ? 0.0 : ip[I - 1] / rf;
    double pfJ = (!ip || near_zero_d(rf)) ? 0.0 : ip[J - 1] / rf;
    return pfI * psI + pfJ * psJ;
}

/* unoccupied spin orbitals (ascending) and their bitmask (get_unoccupied,
   gasci_pchb_doubles_spinorb_fullyweighted.fpp:403-412) */
static int unoccupied(const osys_t *s, const uint64_t *ilut, int *un, uint64_t *umask)
{
    const int n = s->nbasis, W = nwords(n);
    int nu = 0;
    for (int w = 0; w < W; ++w) umask[w] = 0;
    for (int o = 1; o <= n; ++o)
        if (!test_orb(ilut, o)) {
            un[nu++] = o;
            set_orb(umask, o);
        }
    return nu;
}

/* renormalisation of a hole row over the unoccupied orbitals: one minus the
   occupied ones' share, or the direct sum when that is 1 or below 1e-5
   (do_direct_calculation, aliasSampling.F90:967-1000) */
static double hole_renorm(const double *row, const int32_t *nI, int nel, const int *un, int nu)
{
    int set[O_MAX_NEL];
    for (int k = 0; k < nel; ++k) set[k] = nI[k];
    double r = 1.0 - row_sum(row, set, nel);
    if (r >= 1.0 || r < 1e-5) r = row_sum(row, un, nu);
    return r;
}

static const double *a_row(int IJ) { return P.a_ok[IJ - 1] ? P.a_prob + (size_t)(IJ - 1) * P.nbi : NULL; }
static const double *b_row(int IJ, int A)
{
    size_t r = (size_t)(IJ - 1) * P.nbi + (A - 1);
    return P.b_ok[r] ? P.b_prob + r * P.nbi : NULL;
}

/* the fully weighted holes of GAS_doubles_PCHB_gen_exc
   (gasci_pchb_doubles_spinorb_fullyweighted.fpp:172-223): (tgt[0], tgt[1]) in
   drawing order and sum(p_first * p_second); 0 for a null draw */
static int draw_full_holes(const osys_t *s, const uint64_t *ilut, const int32_t *nI, int I, int J, draw_seq_t *q,
                           int *tgt, double *ph)
{
    const int n = P.nbi, nel = s->nel;
    int un[128], pos = 0;
    uint64_t umask[O_MAX_W];
    int nu = unoccupied(s, ilut, un, umask);
    int IJ = fuse(I, J);
    const double *ar = a_row(IJ);
    double rf = hole_renorm(ar, nI, nel, un, nu);
    double pf[2], ps[2];
    size_t oa = (size_t)(IJ - 1) * n;
    tgt[0] = constrained_sample(P.a_bias + oa, P.a_alias + oa, ar, n, un, nu, umask, rf, q, &pos, &pf[0]);
    if (!tgt[0]) return 0;
    const double *b1 = b_row(IJ, tgt[0]);
    double rs1 = hole_renorm(b1, nI, nel, un, nu);
    if (near_zero_d(rs1)) return 0;
    size_t ob = ((size_t)(IJ - 1) * n + (tgt[0] - 1)) * n;
    tgt[1] = constrained_sample(P.b_bias + ob, P.b_alias + ob, b1, n, un, nu, umask, rs1, q, &pos, &ps[0]);
    if (!tgt[1]) return 0;
    pf[1] = (!ar || near_zero_d(rf)) ? 0.0 : ar[tgt[1] - 1] / rf;
    const double *b2 = b_row(IJ, tgt[1]);
    double rs2 = hole_renorm(b2, nI, nel, un, nu);
    ps[1] = (!b2 || near_zero_d(rs2)) ? 0.0 : b2[tgt[0] - 1] / rs2;
    *ph = pf[0] * ps[0] + pf[1] * ps[1];
    return 1;
}

/* GAS_doubles_PCHB_get_pgen's hole factor (:246-296) */
static double pgen_full_holes(const osys_t *s, const uint64_t *ilut, const int32_t *nI, int I, int J, int A, int B)
{
    const int nel = s->nel;
    int un[128];
    uint64_t umask[O_MAX_W];
    int nu = unoccupied(s, ilut, un, umask);
    int IJ = fuse(I, J);
    const double *ar = a_row(IJ);
    double rf = hole_renorm(ar, nI, nel, un, nu);
    double pfA = (!ar || near_zero_d(rf)) ? 0.0 : ar[A - 1] / rf;
    double pfB = (!ar || near_zero_d(rf)) ? 0.0 : ar[B - 1] / rf;
    const double *bA = b_row(IJ, A), *bB = b_row(IJ, B);
    double rA = hole_renorm(bA, nI, nel, un, nu), rB = hole_renorm(bB, nI, nel, un, nu);
    double psA = (!bA || near_zero_d(rA)) ? 0.0 : bA[B - 1] / rA;
    double psB = (!bB || near_zero_d(rB)) ? 0.0 : bB[A - 1] / rB;
    return pfA * psA + pfB * psB;
}

/* PCHB doubles with weighted particles (UNIF-FULL, FULL-FULL) or fully
   weighted holes: every uniform from the attempt's draw sequence, particles
   first, then holes */
static int gen_double_pchb_weighted(const osys_t *s, const uint64_t *ilut, const int32_t *nI, draw_seq_t *q,
                                    int32_t *ex, int32_t *nJ, double *pgen)
{
    const int nel = s->nel;
    int elecs[2], srcs[2];
    double p;
    if (s->particles == O_PART_UNIF_FULL || s->particles == O_PART_FULL_FULL) {
        if (!draw_weighted_particles(s, ilut, nI, q, elecs, srcs, &p)) return 0;
    } else if (s->particles == O_PART_UNIF_FAST) {
        elecs[0] = (int)(seq_next(q) * (double)nel) + 1;
        srcs[0] = nI[elecs[0] - 1];
        if (!P.j_ok[srcs[0] - 1]) return 0;
        size_t off = (size_t)(srcs[0] - 1) * P.nbi;
        int j = alias_sample(P.j_bias + off, P.j_alias + off, P.nbi, seq_next(q));
        double p_J_I = P.j_prob[off + j];
        srcs[1] = j + 1;
        if (!test_orb(ilut, srcs[1])) return 0;
        for (int k = 0; k < nel; ++k) if (nI[k] == srcs[1]) elecs[1] = k + 1;
        double p_I_J = P.j_ok[srcs[1] - 1] ? P.j_prob[(size_t)(srcs[1] - 1) * P.nbi + (srcs[0] - 1)] : 0.0;
        p = (p_J_I + p_I_J) / (double)nel;
        if (srcs[0] > srcs[1]) {
            int t = srcs[0]; srcs[0] = srcs[1]; srcs[1] = t;
            t = elecs[0]; elecs[0] = elecs[1]; elecs[1] = t;
        }
    } else {
        if (!pick_biased_elecs(s, nI, seq_next(q), elecs, &p)) return 0;
        srcs[0] = nI[elecs[0] - 1];
        srcs[1] = nI[elecs[1] - 1];
    }
    int t[2];
    double ph;
    if (s->holes == O_HOLES_FULL_FULL) {
        if (!draw_full_holes(s, ilut, nI, srcs[0], srcs[1], q, t, &ph)) return 0;
    } else {
        int ij = fuse(srcs[0], srcs[1]);
        if (!P.ab_ok[ij - 1]) return 0;
        size_t off = (size_t)(ij - 1) * P.ijmax;
        int ab = alias_sample(P.ab_bias + off, P.ab_alias + off, P.ijmax, seq_next(q)) + 1;
        ph = P.ab_prob[off + ab - 1];
        int A = (int)floor((1.0 + sqrt(8.0 * (double)ab - 7.0)) / 2.0);
        while (A * (A - 1) / 2 >= ab) A--;
        while ((A + 1) * A / 2 < ab) A++;
        t[0] = ab - A * (A - 1) / 2;
        t[1] = A;
        if (t[0] == t[1] || test_orb(ilut, t[0]) || test_orb(ilut, t[1])) return 0;
        if (fabs(ph) <= EPS) return 0;
    }
    int tp = o_make_double(nI, nel, elecs[0], elecs[1], t[0], t[1], nJ, ex);
    *pgen = p * ph;
    return 1 + (tp ? 2 : 0);
}

/* spin class of a spin orbital: 1 alpha (even), 2 beta (odd), as ClassCountInd(spin, 0, 0) */
static inline int cls(int orb) { return is_beta(orb) ? 2 : 1; }

/* gen_rand_excit (GenRandSymExcitNUMod.F90:72-153) for a C1 system (one irrep, no Lz).
   ClassCountOcc/Unocc (construct_class_counts :1288) reduce to per-spin counts; the
   orbitals of one spin in symmetry-label order are taken in index order.
   Returns 0 for a null excitation; *fail marks a rejection loop past the bound at
   which the reference stops the run. */
static int gen_uniform(const osys_t *s, const uint64_t *ilut, const int32_t *nI, const double *d,
                       draw_seq_t *q, int32_t *ic, int32_t *ex, int32_t *nJ, double *pgen, int32_t *fail)
{
    const int nel = s->nel, nb = s->nbasis, nsp = nb / 2;
    int occ[3] = {0, 0, 0}, unocc[3];
    for (int i = 0; i < nel; ++i) occ[cls(nI[i])]++;
    unocc[1] = nsp - occ[1];
    unocc[2] = nsp - occ[2];
    double pdoub = 1.0 - s->psingles;
    if (d[0] < pdoub) {
        /* CreateDoubExcit (:155-185) */
        *ic = 2;
        int elecpairs = nel * (nel - 1) / 2;
        /* PickElecPair (:1021-1058) */
        int ind = (int)(elecpairs * seq_next(q)) + 1;
        int X = elecpairs - ind;
        int K = (int)((sqrt(8.0 * (double)X + 1.0) - 1.0) / 2.0);
        int e1 = nel - 1 - K, e2 = nel - X + (K * (K + 1)) / 2;
        int orb1 = nI[e1 - 1], orb2 = nI[e2 - 1];
        int ms1 = is_beta(orb1) ? -1 : 1, ms2 = is_beta(orb2) ? -1 : 1;
        int ispn = (ms1 * ms2 == -1) ? 2 : (ms1 == 1 ? 3 : 1);
        /* FindNumForbiddenOrbs (:385-412, no Lz, one irrep) */
        int forbidden = 0;
        if (ispn == 2) {
            if (unocc[1] == 0) forbidden += unocc[2];
            if (unocc[2] == 0) forbidden += unocc[1];
        } else if (ispn == 1) {
            if (unocc[2] == 1) forbidden = 1;
        } else {
            if (unocc[1] == 1) forbidden = 1;
        }
        /* PickAOrb (:672-870) with IsAOrbSymAllowed (:873-949) */
        int nexcit_a = ispn == 2 ? nb - nel : (ispn == 1 ? nsp - occ[2] : nsp - occ[1]);
        if (nexcit_a == forbidden) return 0;
        int orbA = 0, spinA = 0;
        for (int overall = 0;; ++overall) {
            if (nexcit_a - forbidden <= 3) {
                int chosen = (int)((nexcit_a - forbidden) * seq_next(q)) + 1;
                int z = 1;
                if (ispn == 2) {
                    for (int i = 0; i < nb && !orbA; ++i) {
                        if (test_orb(ilut, i + 1)) continue;
                        int sa = is_beta(i + 1) ? -1 : 1;
                        int allowed = sa == 1 ? unocc[2] != 0 : unocc[1] != 0;
                        if (allowed) {
                            if (z == chosen) { orbA = i + 1; spinA = sa; }
                            z++;
                        }
                    }
                } else {
                    for (int i = 1; i <= nsp && !orbA; ++i) {
                        int o = ispn == 1 ? 2 * i - 1 : 2 * i;
                        if (test_orb(ilut, o)) continue;
                        int allowed = ispn == 1 ? unocc[2] != 1 : unocc[1] != 1;
                        if (allowed) {
                            if (z == chosen) orbA = o;
                            z++;
                        }
                    }
                }
                if (!orbA) { *fail = 1; return 0; }
                break;
            }
            int attempts = 0;
            for (;;) {
                int o;
                if (ispn == 2) {
                    o = (int)(nb * seq_next(q)) + 1;
                } else {
                    o = (int)(nsp * seq_next(q)) + 1;
                    o = ispn == 1 ? 2 * o - 1 : 2 * o;
                }
                if (!test_orb(ilut, o)) { orbA = o; break; }
                if (attempts > 250) { *fail = 1; return 0; }
                attempts++;
            }
            spinA = is_beta(orbA) ? -1 : 1;
            int allowed;
            if (ispn == 2) allowed = spinA == 1 ? unocc[2] != 0 : unocc[1] != 0;
            else if (ispn == 1) allowed = unocc[2] != 1;
            else allowed = unocc[1] != 1;
            if (allowed) break;
            if (overall > nb * 20) { *fail = 1; return 0; }
        }
        /* PickBOrb (:193-306), tCycleOrbs off */
        int nexcit_b, other, spin_b;
        if (ispn == 2) {
            if (is_beta(orbA)) { nexcit_b = unocc[1]; other = unocc[2]; spin_b = 1; }
            else { nexcit_b = unocc[2]; other = unocc[1]; spin_b = 2; }
        } else if (ispn == 1) {
            nexcit_b = unocc[2]; other = unocc[2]; spin_b = 2;
        } else {
            nexcit_b = unocc[1]; other = unocc[1]; spin_b = 1;
        }
        if (ispn != 2) { nexcit_b -= 1; other -= 1; }
        int orbB = 0;
        for (int attempts = 0;; ++attempts) {
            int c = (int)(nsp * seq_next(q));
            int o = 2 * (c + 1) - (spin_b - 1);
            if (!test_orb(ilut, o) && o != orbA) { orbB = o; break; }
            if (attempts > 1000) { *fail = 1; return 0; }
        }
        int par = o_make_double(nI, nel, e1, e2, orbA, orbB, nJ, ex);
        /* FindDoubleProb (:187-191) */
        *pgen = pdoub * ((1.0 / (double)nexcit_b) + (1.0 / (double)other)) /
                (double)(elecpairs * (nexcit_a - forbidden));
        return 1 | (par ? 2 : 0);
    }
    /* CreateSingleExcit (:1118-1286) with CheckIfSingleExcits (:1060-1085) */
    *ic = 1;
    int nowex = 0;
    for (int c = 1; c <= 2; ++c)
        if (occ[c] != 0 && unocc[c] == 0) nowex += occ[c];
    if (nowex == nel) return 0;
    int eleci = 0, nexcit = 0, spin = 0;
    for (int attempts = 0;; ++attempts) {
        eleci = (int)(nel * seq_next(q)) + 1;
        spin = cls(nI[eleci - 1]);
        nexcit = unocc[spin];
        if (nexcit != 0) break;
        if (attempts > 250) { *fail = 1; return 0; }
    }
    int orb = 0;
    for (int attempts = 0;; ++attempts) {
        int c = (int)(nsp * seq_next(q));
        int o = 2 * (c + 1) - (spin - 1);
        if (!test_orb(ilut, o)) { orb = o; break; }
        if (attempts > 250) { *fail = 1; return 0; }
    }
    int par = o_make_single(nI, nel, eleci, orb, nJ, ex);
    *pgen = (1.0 - pdoub) / (double)(nexcit * (nel - nowex));
    return 1 | (par ? 2 : 0);
}

/* calc_pgen_symrandexcit2 (GenRandSymExcitNUMod.F90:1322-1470) for a C1 system */
static double pgen_uniform(const osys_t *s, const int32_t *nI, int32_t ic, const int32_t *ex)
{
    const int nel = s->nel, nb = s->nbasis, nsp = nb / 2;
    int occ[3] = {0, 0, 0}, unocc[3];
    for (int i = 0; i < nel; ++i) occ[cls(nI[i])]++;
    unocc[1] = nsp - occ[1];
    unocc[2] = nsp - occ[2];
    double pdoub = 1.0 - s->psingles;
    if (ic == 1) {
        int nowex = 0;
        for (int c = 1; c <= 2; ++c)
            if (occ[c] != 0 && unocc[c] == 0) nowex += occ[c];
        int nexcit_a = unocc[cls(ex[0])];
        return (1.0 - pdoub) / (double)(nexcit_a * (nel - nowex));
    }
    int msi = is_beta(ex[0]) ? -1 : 1, msj = is_beta(ex[1]) ? -1 : 1;
    int ispn, nexcit_a;
    if (msi * msj == -1) { ispn = 2; nexcit_a = nb - nel; }
    else if (msi == 1) { ispn = 3; nexcit_a = nsp - occ[1]; }
    else { ispn = 1; nexcit_a = nsp - occ[2]; }
    int forbidden = 0;
    if (ispn == 2) {
        if (unocc[1] == 0) forbidden += unocc[2];
        if (unocc[2] == 0) forbidden += unocc[1];
    } else if (ispn == 1) {
        if (unocc[2] == 1) forbidden = 1;
    } else {
        if (unocc[1] == 1) forbidden = 1;
    }
    int nexcit_b, other;
    if (ispn == 2) {
        if (is_beta(ex[2])) { nexcit_b = unocc[1]; other = unocc[2]; }
        else { nexcit_b = unocc[2]; other = unocc[1]; }
    } else if (ispn == 1) {
        nexcit_b = unocc[2]; other = unocc[2];
    } else {
        nexcit_b = unocc[1]; other = unocc[1];
    }
    if (ispn != 2) { nexcit_b -= 1; other -= 1; }
    int elecpairs = nel * (nel - 1) / 2;
    return pdoub * ((1.0 / (double)nexcit_b) + (1.0 / (double)other)) / (double)(elecpairs * (nexcit_a - forbidden));
}

int32_t o_gen_excit(const osys_t *s, const uint64_t *ilut, const double *d, int32_t *nJ,
                    uint64_t *ilutJ, int32_t *ic, int32_t *ex, int32_t *tparity, double *pgen)
{
    int32_t fail = 0;
    return o_gen_excit_ctx(s, ilut, d, 0xFFFFFFFFu, 0, 0, nJ, ilutJ, ic, ex, tparity, pgen, &fail);
}

int32_t o_gen_excit_ctx(const osys_t *s, const uint64_t *ilut, const double *d, uint32_t iter, int32_t rep,
                        uint32_t k, int32_t *nJ, uint64_t *ilutJ, int32_t *ic, int32_t *ex, int32_t *tparity,
                        double *pgen, int32_t *fail)
{
    int W = nwords(s->nbasis);
    int32_t nI[O_MAX_NEL];
    o_decode(ilut, W, nI);
    int res;
    double pg = 0.0;
    if (s->model == O_MODEL_MOLECULAR && s->uniform) {
        draw_seq_t q = {s, d, ilut, iter, k, rep, 0, -1, {0.0, 0.0}};
        res = gen_uniform(s, ilut, nI, d, &q, ic, ex, nJ, &pg, fail);
    } else if (s->model == O_MODEL_HUBBARD) {
        *ic = 1;
        res = gen_hubbard(s, ilut, nI, d, ex, nJ, &pg);
    } else if (d[0] < s->psingles) {
        /* gen_exc_sd (excitation_generators.F90:112-138) */
        *ic = 1;
        res = gen_single_heatbath(s, ilut, nI, d[1], d[2], ex, nJ, &pg);
        pg = pg * s->psingles;
    } else {
        *ic = 2;
        if (s->particles == O_PART_UNIF_FULL || s->particles == O_PART_FULL_FULL || s->holes == O_HOLES_FULL_FULL) {
            draw_seq_t q = {s, d, ilut, iter, k, rep, 0, -1, {0.0, 0.0}};
            res = gen_double_pchb_weighted(s, ilut, nI, &q, ex, nJ, &pg);
        } else {
            res = gen_double_pchb(s, ilut, nI, d, ex, nJ, &pg);
        }
        pg = pg * (1.0 - s->psingles);
    }
    if (!res) return 0;
    *tparity = (res & 2) ? 1 : 0;
    *pgen = pg;
    /* encode_child = FindExcitBitDet (DetBitOps.F90:584-623) */
    for (int w = 0; w < W; ++w) ilutJ[w] = ilut[w];
    if (*ic == 1) { clr_orb(ilutJ, ex[0]); set_orb(ilutJ, ex[1]); }
    else { clr_orb(ilutJ, ex[0]); clr_orb(ilutJ, ex[1]); set_orb(ilutJ, ex[2]); set_orb(ilutJ, ex[3]); }
    return 1;
}

/* calc_pgen for the bound generator: calc_pgen_rs_hubbard (real_space_hubbard.F90:2035),
   pgen_single_4ind (excit_gens_int_weighted.F90:491-545), GAS_doubles_PCHB_spinorb_get_pgen
   (gasci_pchb_doubles_spinorb_fastweighted.fpp:246-268), get_pgen_sd (excitation_generators.F90:141) */
double o_calc_pgen(const osys_t *s, const uint64_t *ilut, int32_t ic, const int32_t *ex)
{
    int W = nwords(s->nbasis);
    if (s->model == O_MODEL_MOLECULAR && s->uniform) {
        int32_t nI[O_MAX_NEL];
        o_decode(ilut, W, nI);
        return pgen_uniform(s, nI, ic, ex);
    }
    int32_t nI[O_MAX_NEL];
    o_decode(ilut, W, nI);
    int nel = s->nel;
    if (s->model == O_MODEL_HUBBARD) {
        if (ic != 1) return 0.0;
        int src = ex[0], tgt = ex[1], site = spat(src);
        double cum_sum = 0.0, cpt = 0.0;
        for (int k = 0; k < s->maxnbr; ++k) {
            int ns = s->nbr[(site - 1) * s->maxnbr + k];
            if (ns == 0) break;
            int orb = is_beta(src) ? 2 * ns - 1 : 2 * ns;
            double elem = test_orb(ilut, orb) ? 0.0 : fabs(s->bhub);
            cum_sum += elem;
            if (orb == tgt) cpt = elem;
        }
        if (cum_sum < EPS) return 0.0;
        return (1.0 / (double)nel) * (cpt / cum_sum);
    }
    if (ic == 1) {
        int src = ex[0], tgt = ex[1];
        double pgen = 1.0 / (double)nel;
        int nsp = s->nbasis / 2;
        double cum_sum = 0.0, cpt_tgt = 0.0;
        for (int i = 0; i < nsp; ++i) {
            int orb = 2 * (i + 1) - (is_beta(tgt) ? 1 : 0);
            if (!test_orb(ilut, orb)) {
                double cpt = fabs(hb_weight(s, ilut, src, orb));
                cum_sum += cpt;
                if (orb == tgt) cpt_tgt = cpt;
            }
        }
        if (fabs(cum_sum) <= EPS) pgen = 0.0; else pgen = pgen * cpt_tgt / cum_sum;
        return s->psingles * pgen;
    }
    if (ic == 2) {
        int I = ex[0], J = ex[1], n = P.nbi;
        double p;
        if (s->particles == O_PART_UNIF_FULL || s->particles == O_PART_FULL_FULL) {
            p = pgen_weighted_particles(s, nI, I, J);
        } else if (s->particles == O_PART_UNIF_FAST) {
            double a = P.j_ok[I - 1] ? P.j_prob[(size_t)(I - 1) * n + (J - 1)] : 0.0;
            double b = P.j_ok[J - 1] ? P.j_prob[(size_t)(J - 1) * n + (I - 1)] : 0.0;
            p = (a + b) / (double)nel;
        } else {
            int nalpha = 0, nbeta = 0;
            for (int i = 0; i < nel; ++i) { if (is_beta(nI[i])) nbeta++; else nalpha++; }
            int par = nalpha * (nalpha - 1) / 2 + nbeta * (nbeta - 1) / 2;
            if (is_beta(I) == is_beta(J)) p = s->pparallel / (double)par;
            else p = (1.0 - s->pparallel) / (double)(nalpha * nbeta);
        }
        double ph;
        if (s->holes == O_HOLES_FULL_FULL) {
            ph = pgen_full_holes(s, ilut, nI, I, J, ex[2], ex[3]);
        } else {
            int ij = fuse(I, J), ab = fuse(ex[2], ex[3]);
            ph = P.ab_ok[ij - 1] ? P.ab_prob[(size_t)(ij - 1) * P.ijmax + ab - 1] : 0.0;
        }
        p = p * ph;
        return (1.0 - s->psingles) * p;
    }
    return 0.0;
}

/* ------------------------------------------------------- iteration state */

struct ostate {
    osys_t s;
    int W, S, rowlen;
    int64_t maxw, maxs;
    /* CurrentDets: rowlen int64 per det = W det words, S signs, 1 flag word */
    uint64_t *cur;
    double *diagH, *offdiagH;
    int64_t totwalkers;
    /* HashIndex chains: head[b] = slot+1, next[slot] = slot+1 */
    int64_t *head, *next;
    int64_t *freeslot, ifree_start, ifree_end;
    /* SpawnedParts */
    uint64_t *spawned, *spawned2;
    int64_t nspawned;
    int64_t *perm, *perm_tmp;
    /* semistochastic core space (cs_replicas: core dets, sparse_core_ham) */
    int64_t ncore;
    uint64_t *core_det;   /* ncore * W */
    int64_t *core_row;    /* CurrentDets row of each core det */
    int64_t *rowptr;      /* CSR of H_ij (diagonal H_ii - Hii), columns ascending */
    int32_t *cols;
    double *vals;         /* FP64 values, or FP64 copies of FP32-rounded values */
    double *cx, *cy;      /* core amplitudes and their projection, ncore * S */
};

static inline double *row_sign(ostate_t *st, uint64_t *row) { (void)st; return (double *)(row + st->W); }
static inline uint64_t *row_flag(ostate_t *st, uint64_t *row) { return row + st->W + st->S; }
static inline uint64_t *cur_row(ostate_t *st, int64_t j) { return st->cur + (size_t)j * st->rowlen; }

ostate_t *o_create(const osys_t *sys)
{
    ostate_t *st = (ostate_t *)calloc(1, sizeof(ostate_t));
    st->s = *sys;
    st->W = nwords(sys->nbasis);
    st->S = sys->nrep;
    st->rowlen = st->W + st->S + 1;
    st->maxw = sys->max_walkers;
    st->maxs = sys->max_spawned;
    st->cur = (uint64_t *)calloc((size_t)st->maxw * st->rowlen, sizeof(uint64_t));
    st->diagH = (double *)calloc((size_t)st->maxw, sizeof(double));
    st->offdiagH = (double *)calloc((size_t)st->maxw, sizeof(double));
    st->head = (int64_t *)calloc((size_t)sys->nbuckets, sizeof(int64_t));
    st->next = (int64_t *)calloc((size_t)st->maxw, sizeof(int64_t));
    st->freeslot = (int64_t *)calloc((size_t)st->maxw, sizeof(int64_t));
    st->spawned = (uint64_t *)calloc((size_t)st->maxs * st->rowlen, sizeof(uint64_t));
    st->spawned2 = (uint64_t *)calloc((size_t)st->maxs * st->rowlen, sizeof(uint64_t));
    st->perm = (int64_t *)calloc((size_t)st->maxs, sizeof(int64_t));
    st->perm_tmp = (int64_t *)calloc((size_t)st->maxs, sizeof(int64_t));
    return st;
}

static void free_core(ostate_t *st)
{
    free(st->core_det); free(st->core_row); free(st->rowptr); free(st->cols); free(st->vals);
    free(st->cx); free(st->cy);
    st->core_det = NULL; st->core_row = NULL; st->rowptr = NULL; st->cols = NULL; st->vals = NULL;
    st->cx = NULL; st->cy = NULL;
    st->ncore = 0;
}

void o_destroy(ostate_t *st)
{
    if (!st) return;
    free(st->cur); free(st->diagH); free(st->offdiagH); free(st->head); free(st->next);
    free(st->freeslot); free(st->spawned); free(st->spawned2); free(st->perm); free(st->perm_tmp);
    free_core(st);
    free(st);
}

static int64_t bucket_of(ostate_t *st, const uint64_t *det)
{
    int32_t nI[O_MAX_NEL];
    o_decode(det, st->W, nI);
    return o_find_walker_hash(nI, st->s.nel, st->s.random_hash2, st->s.nbuckets) - 1;
}

/* hash_table_lookup (hash.F90:294-338) */
static int64_t hash_lookup(ostate_t *st, const uint64_t *det, int64_t b)
{
    int64_t e = st->head[b];
    while (e) {
        if (det_eq(cur_row(st, e - 1), det, st->W)) return e - 1;
        e = st->next[e - 1];
    }
    return -1;
}

/* add_hash_table_entry (hash.F90:216-246): append to the chain */
static void hash_add(ostate_t *st, int64_t slot, int64_t b)
{
    st->next[slot] = 0;
    if (!st->head[b]) { st->head[b] = slot + 1; return; }
    int64_t e = st->head[b];
    while (st->next[e - 1]) e = st->next[e - 1];
    st->next[e - 1] = slot + 1;
}

/* RemoveHashDet (load_balancer.fpp:631-644) */
static void remove_hash_det(ostate_t *st, int64_t slot)
{
    int64_t b = bucket_of(st, cur_row(st, slot));
    int64_t e = st->head[b], prev = 0;
    while (e) {
        if (e - 1 == slot) {
            if (prev) st->next[prev - 1] = st->next[e - 1]; else st->head[b] = st->next[e - 1];
            break;
        }
        prev = e;
        e = st->next[e - 1];
    }
    st->freeslot[st->ifree_end++] = slot;
    *row_flag(st, cur_row(st, slot)) |= FLAG_REMOVED;
}

/* AddNewHashDet (load_balancer.fpp:514-629) */
static int add_new_hash_det(ostate_t *st, const uint64_t *det, const double *sgn, uint64_t flags,
                            double diagH_total, double offdiagH, int64_t b)
{
    int64_t pos;
    if (st->ifree_start < st->ifree_end) {
        pos = st->freeslot[st->ifree_start++];
    } else {
        if (st->totwalkers >= st->maxw) return 2;
        pos = st->totwalkers++;
    }
    uint64_t *row = cur_row(st, pos);
    for (int w = 0; w < st->W; ++w) row[w] = det[w];
    memcpy(row_sign(st, row), sgn, sizeof(double) * st->S);
    *row_flag(st, row) = flags & ~(uint64_t)FLAG_REMOVED;
    st->diagH[pos] = diagH_total - st->s.hii;
    st->offdiagH[pos] = offdiagH;
    hash_add(st, pos, b);
    return 0;
}

int32_t o_add_walker(ostate_t *st, const uint64_t *det, const double *sgn, uint32_t flags)
{
    int32_t nI[O_MAX_NEL];
    o_decode(det, st->W, nI);
    int64_t b = bucket_of(st, det);
    if (hash_lookup(st, det, b) >= 0) return 4;
    return add_new_hash_det(st, det, sgn, flags, o_diag_helement(&st->s, nI),
                            o_offdiag_helement(&st->s, det), b);
}

int32_t o_add_walkers(ostate_t *st, const uint64_t *dets, const double *sgns, const uint32_t *flags,
                      int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        int rc = o_add_walker(st, dets + i * st->W, sgns + i * st->S, flags[i]);
        if (rc) return rc;
    }
    return 0;
}

int64_t o_get_walkers(ostate_t *st, uint64_t *dets, double *sgns, uint32_t *flags, double *diagh,
                      double *offdiagh, int64_t cap)
{
    int64_t n = 0;
    for (int64_t j = 0; j < st->totwalkers; ++j) {
        uint64_t *row = cur_row(st, j);
        uint64_t f = *row_flag(st, row);
        double *sg = row_sign(st, row);
        int occ = 0;
        for (int r = 0; r < st->S; ++r) if (fabs(sg[r]) >= 1e-12) occ = 1;
        if (!occ && !(f & FLAG_CORE)) continue;
        if (n < cap) {
            for (int w = 0; w < st->W; ++w) dets[n * st->W + w] = row[w];
            memcpy(sgns + n * st->S, sg, sizeof(double) * st->S);
            flags[n] = (uint32_t)(f & ~(uint64_t)FLAG_REMOVED);
            if (diagh) diagh[n] = st->diagH[j];
            if (offdiagh) offdiagh[n] = st->offdiagH[j];
        }
        n++;
    }
    return n;
}

/* generate_space_most_populated (semi_stoch_gen.F90:1031-1226) for one rank:
   the target determinants of largest core_space_weight = sum_r |c_r|
   (semi_stoch_procs.F90:1623-1642, the global core space), those below 1e-8
   left out (:1094-1100), all of them when fewer are occupied.  Equal weights
   are ordered by the determinant (ilut_lt ascending), where the reference
   keeps the one it meets first in its storage order
   (proc_most_populated_states, semi_stoch_procs.F90:1310-1421), so that the
   choice is a function of the walker set.  out receives the chosen
   determinants in that order (weight descending); returns their number. */
typedef struct {
    double w;
    uint64_t d[O_MAX_W];
} wdet_t;
static int g_mp_w;
static int cmp_wdet(const void *a, const void *b)
{
    const wdet_t *x = (const wdet_t *)a, *y = (const wdet_t *)b;
    if (x->w > y->w) return -1;
    if (x->w < y->w) return 1;
    return ilut_cmp(x->d, y->d, g_mp_w);
}

int64_t o_most_populated(ostate_t *st, int64_t target, uint64_t *out)
{
    const int W = st->W;
    wdet_t *c = (wdet_t *)malloc(sizeof(wdet_t) * (size_t)(st->totwalkers > 0 ? st->totwalkers : 1));
    int64_t n = 0;
    for (int64_t j = 0; j < st->totwalkers; ++j) {
        uint64_t *row = cur_row(st, j);
        double *sg = row_sign(st, row);
        double w = 0.0;
        for (int r = 0; r < st->S; ++r) w += fabs(sg[r]);
        if (w < 1e-8) continue;
        c[n].w = w;
        for (int q = 0; q < O_MAX_W; ++q) c[n].d[q] = q < W ? row[q] : 0;
        n++;
    }
    g_mp_w = W;
    qsort(c, (size_t)n, sizeof(wdet_t), cmp_wdet);
    int64_t k = n < target ? n : target;
    for (int64_t i = 0; i < k; ++i)
        for (int w = 0; w < W; ++w) out[i * W + w] = c[i].d[w];
    free(c);
    return k;
}

/* ------------------------------------------------------- core space */

/* init_semi_stochastic (semi_stoch_gen.F90:425-445, semi_stoch_procs.F90:1095-1120)
   for a given core space: every core det is in CurrentDets (absent ones with zero
   sign), flagged deterministic and, with the initiator approximation
   (t_core_inits), initiator in every replica.  The core Hamiltonian
   (calc_determ_hamil, fast_determ_hamil.F90:892-1548: H_ij, diagonal H_ii - Hii)
   is stored as CSR, columns ascending, off-diagonal zeros dropped; fp32 keeps
   FP32-rounded values. */
int32_t o_set_core(ostate_t *st, const uint64_t *dets, int64_t n, int32_t fp32, int64_t *nnz_out)
{
    const int W = st->W, S = st->S, nel = st->s.nel;
    free_core(st);
    if (n <= 0) { if (nnz_out) *nnz_out = 0; return 0; }
    st->ncore = n;
    st->core_det = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n * W);
    memcpy(st->core_det, dets, sizeof(uint64_t) * (size_t)n * W);
    st->core_row = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    st->rowptr = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    st->cx = (double *)calloc((size_t)n * S, sizeof(double));
    st->cy = (double *)calloc((size_t)n * S, sizeof(double));
    int32_t *occ = (int32_t *)malloc(sizeof(int32_t) * (size_t)n * nel);
    for (int64_t i = 0; i < n; ++i) o_decode(dets + i * W, W, occ + i * nel);
    for (int pass = 0; pass < 2; ++pass) {
        int64_t k = 0;
        for (int64_t i = 0; i < n; ++i) {
            const uint64_t *di = dets + i * W;
            for (int64_t j = 0; j < n; ++j) {
                const uint64_t *dj = dets + j * W;
                if (excit_level(di, dj, W) > 2) continue;
                double h = (i == j) ? o_diag_helement(&st->s, occ + i * nel) - st->s.hii
                                    : o_get_helement(&st->s, occ + i * nel, occ + j * nel, nel);
                if (i != j && h == 0.0) continue;
                if (pass == 1) {
                    st->cols[k] = (int32_t)j;
                    st->vals[k] = fp32 ? (double)(float)h : h;
                }
                k++;
            }
            if (pass == 0) st->rowptr[i + 1] = k;
        }
        if (pass == 0) {
            st->cols = (int32_t *)malloc(sizeof(int32_t) * (size_t)(k ? k : 1));
            st->vals = (double *)malloc(sizeof(double) * (size_t)(k ? k : 1));
        }
    }
    free(occ);
    if (nnz_out) *nnz_out = st->rowptr[n];
    /* place the core dets in CurrentDets */
    st->ifree_start = st->ifree_end = 0;
    uint64_t cflags = FLAG_CORE;
    if (st->s.t_trunc_init)
        for (int r = 0; r < S; ++r) cflags |= 1ull << r;
    double zero[O_MAX_REPLICAS] = {0};
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t *d = dets + i * W;
        int64_t b = bucket_of(st, d);
        int64_t pos = hash_lookup(st, d, b);
        if (pos < 0) {
            int32_t nI[O_MAX_NEL];
            o_decode(d, W, nI);
            int e = add_new_hash_det(st, d, zero, cflags, o_diag_helement(&st->s, nI),
                                     o_offdiag_helement(&st->s, d), b);
            if (e) return e;
            pos = hash_lookup(st, d, b);
        } else {
            *row_flag(st, cur_row(st, pos)) |= cflags;
        }
        st->core_row[i] = pos;
    }
    return 0;
}

/* is_core_state: the det is in CurrentDets with the deterministic flag */
static int is_core_det(ostate_t *st, const uint64_t *det)
{
    int64_t pos = hash_lookup(st, det, bucket_of(st, det));
    return pos >= 0 && (*row_flag(st, cur_row(st, pos)) & FLAG_CORE);
}

/* determ_projection (semi_stoch_procs.F90:105-241): y_i = tau*(S*x_i - sum_j H_ij x_j).
   Summation order (shared with the GPU engine): the nonzeros of row i are dealt
   to 64 partial sums by their position q in the row (partial q mod 64, each
   summed in ascending q as partial = partial - H*x), the partials are combined
   by the butterfly p[l] = p[l] + p[l xor o] for o = 32,16,8,4,2,1, then
   y = (p[0] + S*x_i) * tau. */
void o_core_project(const ostate_t *st, const double *x, const double *shift, double tau, double *y)
{
    const int S = st->S;
    double p[64], t[64];
    for (int64_t i = 0; i < st->ncore; ++i) {
        for (int r = 0; r < S; ++r) {
            for (int l = 0; l < 64; ++l) p[l] = 0.0;
            for (int64_t q = st->rowptr[i]; q < st->rowptr[i + 1]; ++q) {
                int l = (int)((q - st->rowptr[i]) & 63);
                p[l] = p[l] - st->vals[q] * x[(size_t)st->cols[q] * S + r];
            }
            for (int o = 32; o > 0; o >>= 1) {
                for (int l = 0; l < 64; ++l) t[l] = p[l] + p[l ^ o];
                memcpy(p, t, sizeof(p));
            }
            y[(size_t)i * S + r] = (p[0] + shift[r] * x[(size_t)i * S + r]) * tau;
        }
    }
}

int64_t o_core_csr(const ostate_t *st, int64_t *rowptr, int32_t *cols, double *vals)
{
    if (rowptr) memcpy(rowptr, st->rowptr, sizeof(int64_t) * (size_t)(st->ncore + 1));
    int64_t nnz = st->ncore ? st->rowptr[st->ncore] : 0;
    if (cols) memcpy(cols, st->cols, sizeof(int32_t) * (size_t)nnz);
    if (vals) memcpy(vals, st->vals, sizeof(double) * (size_t)nnz);
    return nnz;
}

/* deterministic_annihilation (Annihilation.F90:930-963) */
static void core_annihilation(ostate_t *st, ostats_t *out)
{
    const int S = st->S;
    for (int64_t i = 0; i < st->ncore; ++i) {
        double *cs = row_sign(st, cur_row(st, st->core_row[i]));
        for (int r = 0; r < S; ++r) {
            double y = st->cy[(size_t)i * S + r], c = cs[r];
            cs[r] = y + c;
            out->nborn[r] += fabs(y);
            if (c * y < 0.0) {
                double a = fabs(c), b = fabs(y);
                out->nannihil[r] += 2.0 * (a < b ? a : b);
            }
        }
    }
}

/* stochastic_round_r (lib/util_mod.fpp:207-224) */
static inline double stochastic_round(double x, double r)
{
    double i = (double)(int64_t)x;
    double res = x - i;
    if (fabs(res) >= 1.0e-12)
        if (fabs(res) > r) i = i + (x >= 0.0 ? 1.0 : -1.0);
    return i;
}

/* ---------------------------------------------------------- the iteration */

static int cmp_rows(ostate_t *st, const uint64_t *base, int64_t a, int64_t b)
{
    return ilut_cmp(base + (size_t)a * st->rowlen, base + (size_t)b * st->rowlen, st->W);
}

/* stable bottom-up merge sort of st->perm[0:n] by ilut_lt */
static void sort_spawned(ostate_t *st, const uint64_t *base, int64_t n)
{
    int64_t *a = st->perm, *t = st->perm_tmp;
    for (int64_t i = 0; i < n; ++i) a[i] = i;
    for (int64_t width = 1; width < n; width *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * width) {
            int64_t mid = lo + width < n ? lo + width : n;
            int64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            int64_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) t[k++] = (cmp_rows(st, base, a[j], a[i]) < 0) ? a[j++] : a[i++];
            while (i < mid) t[k++] = a[i++];
            while (j < hi) t[k++] = a[j++];
        }
        int64_t *x = a; a = t; t = x;
    }
    if (a != st->perm) memcpy(st->perm, a, sizeof(int64_t) * (size_t)n);
}

/* CompressSpawnedList (Annihilation.F90:249-515) with FindResidualParticle (:551-634).
   Input rows in st->spawned[0:n]; merged rows written to st->spawned2, returns count. */
static int64_t compress_spawned(ostate_t *st, int64_t n, ostats_t *out)
{
    const int W = st->W, S = st->S, RL = st->rowlen;
    sort_spawned(st, st->spawned, n);
    int64_t vec = 0, i = 0;
    while (i < n) {
        const uint64_t *first = st->spawned + (size_t)st->perm[i] * RL;
        int64_t j = i + 1;
        while (j < n && det_eq(st->spawned + (size_t)st->perm[j] * RL, first, W)) j++;
        uint64_t *dst = st->spawned2 + (size_t)vec * RL;
        if (j == i + 1) {
            const double *sg = (const double *)(first + W);
            double tot = 0.0;
            for (int r = 0; r < S; ++r) tot += fabs(sg[r]);
            if (tot < 1.e-12) { i = j; continue; }
            memcpy(dst, first, sizeof(uint64_t) * RL);
            vec++;
            i = j;
            continue;
        }
        /* cum_det: det words, zero signs and flags */
        memset(dst, 0, sizeof(uint64_t) * RL);
        for (int w = 0; w < W; ++w) dst[w] = first[w];
        double *cum = (double *)(dst + W);
        uint64_t *cflag = dst + W + S;
        for (int64_t q = i; q < j; ++q) {
            const uint64_t *nw = st->spawned + (size_t)st->perm[q] * RL;
            const double *ns = (const double *)(nw + W);
            uint64_t nf = nw[W + S];
            for (int r = 0; r < S; ++r) {
                double new_sgn = ns[r], cum_sgn = cum[r];
                if (st->s.t_trunc_init) {
                    if ((fabs(cum_sgn) > 1.e-12 && fabs(new_sgn) > 1.e-12) || (nf & (1ull << r)))
                        *cflag |= 1ull << r;
                }
                if (cum_sgn * new_sgn < 0.0) {
                    double a = fabs(cum_sgn), b = fabs(new_sgn);
                    out->nannihil[r] += 2.0 * (a < b ? a : b);
                }
                cum[r] = cum_sgn + new_sgn;
            }
        }
        double tot = 0.0;
        for (int r = 0; r < S; ++r) tot += fabs(cum[r]);
        if (tot > 1.e-12) vec++;
        i = j;
    }
    return vec;
}

/* stochRoundSpawn (Annihilation.F90:1354-1418), non-scaled walkers */
static void stoch_round_spawn(ostate_t *st, const uint64_t *det, uint32_t iter, double *sg,
                              ostats_t *out)
{
    double thr = st->s.occ_thresh;
    for (int r = 0; r < st->S; ++r) {
        if (fabs(sg[r]) > 1.e-12 && fabs(sg[r]) < thr) {
            double pRemove = 1.0 - fabs(sg[r]) / thr;
            double u, dummy;
            draw2(&st->s, det, st->W, iter, r, 0, 4, &u, &dummy);
            if (pRemove > u) {
                out->nremoved[r] += fabs(sg[r]);
                sg[r] = 0.0;
            } else {
                out->nborn[r] += thr - fabs(sg[r]);
                sg[r] = sg[r] >= 0.0 ? thr : -thr;
            }
        }
    }
}

static int is_unocc(const double *sg, int S)
{
    for (int r = 0; r < S; ++r) if (!(fabs(sg[r]) < 1.e-12)) return 0;
    return 1;
}

/* AnnihilateSpawnedParts (Annihilation.F90:965-1352) over merged rows st->spawned2[0:nu] */
static int annihilate_spawned(ostate_t *st, int64_t nu, uint32_t iter, ostats_t *out)
{
    const int W = st->W, S = st->S, RL = st->rowlen;
    int32_t nJ[O_MAX_NEL];
    for (int64_t i = 0; i < nu; ++i) {
        uint64_t *sp = st->spawned2 + (size_t)i * RL;
        double *ss = (double *)(sp + W);
        uint64_t sflag = sp[W + S];
        int64_t b = bucket_of(st, sp);
        int64_t pos = hash_lookup(st, sp, b);
        if (pos >= 0) {
            uint64_t *row = cur_row(st, pos);
            double *cs = row_sign(st, row);
            uint64_t cflag = *row_flag(st, row);
            int determ = (cflag & FLAG_CORE) != 0;
            double tot = 0.0;
            for (int r = 0; r < S; ++r) tot += fabs(cs[r]);
            if (tot >= 1.e-12 || determ) {
                for (int r = 0; r < S; ++r) {
                    double prod = cs[r] * ss[r];
                    if (fabs(cs[r]) < 1.0e-12 && st->s.t_trunc_init) {
                        if (!(sflag & (1ull << r)) && !determ) {
                            out->naborted[r] += fabs(ss[r]);
                            ss[r] = 0.0;
                        }
                    }
                    if (prod < 0.0) {
                        double a = fabs(cs[r]), c = fabs(ss[r]);
                        out->nannihil[r] += 2.0 * (a < c ? a : c);
                    }
                }
                for (int r = 0; r < S; ++r) cs[r] = ss[r] + cs[r];
                if (!determ && is_unocc(cs, S)) remove_hash_det(st, pos);
            }
        } else {
            if (st->s.t_trunc_init)
                for (int r = 0; r < S; ++r)
                    if (!(sflag & (1ull << r))) {
                        out->naborted[r] += fabs(ss[r]);
                        ss[r] = 0.0;
                    }
            if (!is_unocc(ss, S)) {
                stoch_round_spawn(st, sp, iter, ss, out);
                if (!is_unocc(ss, S)) {
                    o_decode(sp, W, nJ);
                    double dh = o_diag_helement(&st->s, nJ);
                    double oh = o_offdiag_helement(&st->s, sp);
                    int err = add_new_hash_det(st, sp, ss, sflag & 0xFFFFull, dh, oh, b);
                    if (err) return err;
                    out->n_new++;
                }
            }
        }
    }
    return 0;
}

/* CalcHashTableStats (load_balancer.fpp:646-805), non-scaled walkers */
static void hash_table_stats(ostate_t *st, uint32_t iter, ostats_t *out)
{
    const int S = st->S;
    double thr = st->s.occ_thresh;
    int64_t nocc = 0;
    for (int64_t j = 0; j < st->totwalkers; ++j) {
        uint64_t *row = cur_row(st, j);
        double *sg = row_sign(st, row);
        uint64_t f = *row_flag(st, row);
        int determ = (f & FLAG_CORE) != 0;
        if (is_unocc(sg, S) && !determ) continue;
        for (int r = 0; r < S; ++r) {
            if (!determ && fabs(sg[r]) > 1.e-12 && fabs(sg[r]) < thr) {
                double pRemove = (thr - fabs(sg[r])) / thr;
                double u, dummy;
                draw2(&st->s, row, st->W, iter, r, 0, 5, &u, &dummy);
                if (pRemove > u) {
                    out->nremoved[r] += fabs(sg[r]);
                    sg[r] = 0.0;
                    if (is_unocc(sg, S)) remove_hash_det(st, j);
                } else {
                    out->nborn[r] += thr - fabs(sg[r]);
                    sg[r] = sg[r] >= 0.0 ? thr : -thr;
                }
            }
        }
        if (is_unocc(sg, S) && !determ) continue;
        nocc++;
        for (int r = 0; r < S; ++r) {
            out->tot_parts[r] += fabs(sg[r]);
            out->norm_sq[r] += sg[r] * sg[r];
        }
        if (det_eq(row, st->s.ref, st->W))
            for (int r = 0; r < S; ++r) out->inst_no_at_hf[r] = sg[r];
    }
    out->tot_walkers = nocc;
}

/* walker_death + attempt_die_normal (fcimc_helper.F90:2279-2407,
   fcimc_pointed_fns.F90:573-705) for slot j; ud is the death uniform per replica */
static int walker_death(ostate_t *st, int64_t j, const double *diag_sft, const double *ud,
                        ostats_t *out)
{
    const int S = st->S;
    uint64_t *row = cur_row(st, j);
    double *sg = row_sign(st, row);
    uint64_t f = *row_flag(st, row);
    int core = (f & FLAG_CORE) != 0;
    double Kii = st->diagH[j];
    double idie[O_MAX_REPLICAS], copy[O_MAX_REPLICAS];
    int err = 0;
    for (int r = 0; r < S; ++r) {
        double fac = st->s.tau * (Kii - diag_sft[r]);
        if (fac > 2.0) err = 3;
        if (st->s.t_real_coeff) {
            idie[r] = fac * fabs(sg[r]);
        } else {
            double rat = fac * fabs(sg[r]);
            idie[r] = (double)(int64_t)rat;
            rat = rat - idie[r];
            if (fabs(rat) > ud[r]) idie[r] = idie[r] + (rat >= 0.0 ? 1.0 : -1.0);
        }
        if (core) idie[r] = 0.0;
    }
    for (int r = 0; r < S; ++r) {
        double a = fabs(sg[r]);
        out->ndied[r] += idie[r] < a ? idie[r] : a;
        double b = idie[r] - a;
        out->nborn[r] += b > 0.0 ? b : 0.0;
        copy[r] = sg[r] - idie[r] * (sg[r] >= 0.0 ? 1.0 : -1.0);
    }
    if (st->s.t_trunc_init) {
        int any = 0;
        for (int r = 0; r < S; ++r) if (fabs(copy[r]) > 1.0e-12) any = 1;
        if (any)
            for (int r = 0; r < S; ++r)
                if ((copy[r] > 0.0) != (sg[r] > 0.0)) {
                    out->naborted[r] += fabs(copy[r]);
                    copy[r] = 0.0;
                }
    }
    int any = 0;
    for (int r = 0; r < S; ++r) if (fabs(copy[r]) > 1.0e-12) any = 1;
    if (any || core) {
        memcpy(sg, copy, sizeof(double) * S);
    } else {
        remove_hash_det(st, j);
        for (int r = 0; r < S; ++r) sg[r] = 0.0;
    }
    return err;
}

/* is_core_state through a sorted list of the core determinants (the lookup of
   the sharded iteration, where a child's owner may be another shard) */
typedef struct {
    const uint64_t *dets;  /* ncore * W, ascending in ilut_cmp order */
    int64_t n;
    int W;
} core_index_t;

static int core_index_has(const core_index_t *ci, const uint64_t *det)
{
    int64_t lo = 0, hi = ci->n - 1;
    while (lo <= hi) {
        int64_t mid = (lo + hi) / 2;
        int c = ilut_cmp(ci->dets + (size_t)mid * ci->W, det, ci->W);
        if (c == 0) return 1;
        if (c < 0) lo = mid + 1; else hi = mid - 1;
    }
    return 0;
}

/* The loop over determinants of PerformFCIMCycPar (FciMCPar.F90:1294-1758):
   initiator flags, SumEContrib, spawning into st->spawned, death.  Children
   that are core dets are recognised through ci when given (sharded runs),
   else through the state's own walker hash. */
static int det_loop(ostate_t *st, uint32_t iter, const double *diag_sft, ostats_t *out, const core_index_t *ci)
{
    const osys_t *s = &st->s;
    const int W = st->W, S = st->S, RL = st->rowlen;
    int err = 0;
    int32_t nJ[O_MAX_NEL], ex[4];
    uint64_t ilutJ[O_MAX_W];
    /* loop_over_determinants (FciMCPar.F90:1294-1758) */
    for (int64_t j = 0; j < st->totwalkers; ++j) {
        uint64_t *row = cur_row(st, j);
        double *sg = row_sign(st, row);
        uint64_t *fl = row_flag(st, row);
        int core = (*fl & FLAG_CORE) != 0;
        if (is_unocc(sg, S)) {
            if (core) continue;
            if (!(*fl & FLAG_REMOVED)) *fl |= FLAG_REMOVED;
            st->freeslot[st->ifree_end++] = j;
            continue;
        }
        int exlevel = excit_level(s->ref, row, W);
        double hoff = st->offdiagH[j];
        /* CalcParentFlag_det / TestInitiator_explicit (fcimc_helper.F90:1036-1243) */
        if (s->t_trunc_init) {
            int isref = det_eq(row, s->ref, W);
            for (int r = 0; r < S; ++r) {
                int was = (*fl >> r) & 1;
                int pop = fabs(sg[r]) > s->init_walk_no;
                int now = was ? (pop || isref || core) : pop;
                if (now) {
                    *fl |= 1ull << r;
                    out->no_init_dets[r] += 1.0;
                    out->no_init_walk[r] += fabs(sg[r]);
                } else {
                    *fl &= ~(1ull << r);
                    out->no_noninit_dets[r] += 1.0;
                    out->no_noninit_walk[r] += fabs(sg[r]);
                }
            }
        }
        /* SumEContrib (fcimc_helper.F90:518-802) */
        for (int r = 0; r < S; ++r) {
            if (exlevel == 0) out->hf_cyc[r] += sg[r];
            if (exlevel == 2) out->noat_doubs[r] += fabs(sg[r]);
            out->enum_cyc[r] += hoff * sg[r];
        }
        double udeath[O_MAX_REPLICAS];
        for (int r = 0; r < S; ++r) {
            double ua, ub;
            draw2(s, row, W, iter, r, 0, 0, &ua, &ub);
            udeath[r] = ub;
            /* decide_num_to_spawn (fcimc_helper.F90:2160-2174) */
            double pa = sg[r] * s->avmc;
            int64_t nsp = (int64_t)pa;
            if (nsp < 0) nsp = -nsp;
            if (fabs(fabs(pa) - (double)nsp) > 1.e-12) {
                double extra = fabs(pa) - (double)nsp;
                if (extra > ua) nsp++;
            }
            out->n_attempts += nsp;
            double wsign = sg[r] >= 0.0 ? 1.0 : -1.0;
            int parent_init = (int)((*fl >> r) & 1);
            for (int64_t p = 0; p < nsp; ++p) {
                /* d0, d1 from purpose 1; the PCHB hole pair d3 and the spawn rounding d5
                   from purpose 2; d2, d4 (FAST particles, heat-bath singles, the uniform
                   generator) from purpose 3 */
                double d[6];
                draw2(s, row, W, iter, r, (uint32_t)p, 1, &d[0], &d[1]);
                draw2(s, row, W, iter, r, (uint32_t)p, 2, &d[3], &d[5]);
                draw2(s, row, W, iter, r, (uint32_t)p, 3, &d[2], &d[4]);
                int32_t ic, tpar, gfail = 0;
                double pgen;
                if (!o_gen_excit_ctx(s, row, d, iter, r, (uint32_t)p, nJ, ilutJ, &ic, ex, &tpar, &pgen, &gfail)) {
                    if (gfail) err = 8;
                    out->n_invalid_excits += 1.0;
                    continue;
                }
                out->n_valid_excits += 1.0;
                /* spawning from and to the core space is cancelled (FciMCPar.F90:1659-1668) */
                if (core && (ci ? core_index_has(ci, ilutJ) : (st->ncore && is_core_det(st, ilutJ)))) continue;
                int32_t nI[O_MAX_NEL];
                o_decode(row, W, nI);
                /* attempt_create_normal (fcimc_pointed_fns.F90:178-491); a heat-bath
                   single spawns with the element its weight was drawn from (hb_weight,
                   the generator's HElGen), a double with sltcnd_2 */
                double hel;
                if (ic == 1 && s->model == O_MODEL_MOLECULAR && !s->uniform) {
                    hel = hb_weight(s, row, ex[0], ex[1]);
                    if (tpar) hel = -hel;
                } else {
                    hel = o_get_helement_ex(s, nI, ic, ex, tpar);
                }
                double prob = pgen * s->avmc;
                double nspawn = -s->tau * hel * wsign / prob;
                if (fabs(nspawn) > out->max_cyc_spawn) out->max_cyc_spawn = fabs(nspawn);
                if (s->t_real_coeff) {
                    if (s->t_real_spawn_cutoff && fabs(nspawn) < s->real_spawn_cutoff)
                        nspawn = s->real_spawn_cutoff * stochastic_round(nspawn / s->real_spawn_cutoff, d[5]);
                } else {
                    nspawn = stochastic_round(nspawn, d[5]);
                }
                /* FP32 storage of the spawned amplitude: C's double -> float conversion
                   rounds to nearest even, as the store does */
                if (s->spawn_fp32) nspawn = (double)(float)nspawn;
                if (fabs(nspawn) <= EPS) continue;
                /* new_child_stats_normal (fcimc_pointed_fns.F90:506-571) */
                out->nborn[r] += fabs(nspawn);
                if (ic == 1) out->spawn_from_sing[r] += fabs(nspawn);
                if (fabs(nspawn) > s->init_walk_no) {
                    out->bloom_count[ic] += 1.0;
                    if (fabs(nspawn) > out->bloom_size[ic]) out->bloom_size[ic] = fabs(nspawn);
                }
                /* create_particle (fcimc_helper.F90:152-308) */
                if (st->nspawned >= st->maxs) { err = 1; continue; }
                uint64_t *sp = st->spawned + (size_t)st->nspawned * RL;
                memset(sp, 0, sizeof(uint64_t) * RL);
                for (int w = 0; w < W; ++w) sp[w] = ilutJ[w];
                ((double *)(sp + W))[r] = nspawn;
                if (s->t_trunc_init && parent_init) sp[W + S] |= 1ull << r;
                st->nspawned++;
            }
        }
        int e = walker_death(st, j, diag_sft, udeath, out);
        if (e) err = e;
    }
    out->n_spawned = st->nspawned;
    return err;
}

int32_t o_iterate(ostate_t *st, uint32_t iter, const double *diag_sft, ostats_t *out)
{
    const osys_t *s = &st->s;
    const int S = st->S;
    memset(out, 0, sizeof(*out));
    st->nspawned = 0;
    st->ifree_start = 0;
    st->ifree_end = 0;
    int err = det_loop(st, iter, diag_sft, out, NULL);
    if (err) { out->err = err; return err; }
    /* determ_projection on the core amplitudes (unchanged by the loop: no death) */
    if (st->ncore) {
        for (int64_t i = 0; i < st->ncore; ++i)
            memcpy(st->cx + (size_t)i * S, row_sign(st, cur_row(st, st->core_row[i])), sizeof(double) * S);
        o_core_project(st, st->cx, diag_sft, s->tau, st->cy);
    }
    /* communicate_and_merge_spawns -> CompressSpawnedList */
    int64_t nu = compress_spawned(st, st->nspawned, out);
    out->n_unique = nu;
    /* DirectAnnihilation: deterministic_annihilation, AnnihilateSpawnedParts, CalcHashTableStats */
    if (st->ncore) core_annihilation(st, out);
    err = annihilate_spawned(st, nu, iter, out);
    if (err) { out->err = err; return err; }
    hash_table_stats(st, iter, out);
    return 0;
}

int32_t o_annihilate_list(ostate_t *st, uint32_t iter, const uint64_t *sdet, const double *sval,
                          const uint32_t *sinfo, int64_t n, ostats_t *out, uint64_t *merged_det,
                          double *merged_sgn, uint32_t *merged_flags, int64_t *n_merged)
{
    const int W = st->W, S = st->S, RL = st->rowlen;
    memset(out, 0, sizeof(*out));
    if (n > st->maxs) return 1;
    st->ifree_start = 0;
    st->ifree_end = 0;
    for (int64_t j = 0; j < st->totwalkers; ++j) {
        uint64_t *row = cur_row(st, j);
        if (is_unocc(row_sign(st, row), S) && !(*row_flag(st, row) & FLAG_CORE))
            st->freeslot[st->ifree_end++] = j;
    }
    for (int64_t i = 0; i < n; ++i) {
        uint64_t *sp = st->spawned + (size_t)i * RL;
        memset(sp, 0, sizeof(uint64_t) * RL);
        for (int w = 0; w < W; ++w) sp[w] = sdet[i * W + w];
        int r = (int)(sinfo[i] & 0xFF);
        ((double *)(sp + W))[r] = sval[i];
        if (sinfo[i] & SINFO_INIT) sp[W + S] |= 1ull << r;
    }
    out->n_spawned = n;
    int64_t nu = compress_spawned(st, n, out);
    out->n_unique = nu;
    for (int64_t i = 0; i < nu; ++i) {
        const uint64_t *sp = st->spawned2 + (size_t)i * RL;
        for (int w = 0; w < W; ++w) merged_det[i * W + w] = sp[w];
        memcpy(merged_sgn + i * S, sp + W, sizeof(double) * S);
        merged_flags[i] = (uint32_t)sp[W + S];
    }
    *n_merged = nu;
    int err = annihilate_spawned(st, nu, iter, out);
    if (err) { out->err = err; return err; }
    hash_table_stats(st, iter, out);
    return 0;
}

/* ===================================================== hash-sharded runs
 *
 * The reference runs one MPI rank per core; each rank owns the determinants
 * that DetermineDetNode (load_balance_calcnodes.F90:25-117) assigns to it by
 * a hash of the occupied orbitals, spawns into per-destination sections of
 * SpawnedParts, exchanges them (SendProcNewParts, Annihilation.F90:150-247)
 * and merges and annihilates its own part.  Here the ranks are P threads of
 * one process sharing memory: shard q is an ostate_t holding the
 * determinants it owns, the exchange is a copy between shards, and the
 * semistochastic core vector is shared (the allgather of
 * semi_stoch_procs.F90:127).  With integer walkers the result is that of the
 * one-state iteration, determinant for determinant: every random number is
 * keyed by determinant, and merged sums of integers do not depend on order.
 */
#include <pthread.h>

struct omt {
    osys_t s;
    int P, W, S, RL;
    ostate_t **sh;
    int64_t *sendoff;  /* P x (P+1): start of each destination's section in a sender's buffer */
    /* shared core space */
    int64_t ncore;
    uint64_t *core_det;
    int64_t *rowptr;
    int32_t *cols;
    double *vals, *cx, *cy;
    int32_t *core_owner;
    int64_t *core_row;
    uint64_t *core_sorted;
    core_index_t ci;
    /* iteration arguments and per-shard results */
    uint32_t iter;
    double sft[O_MAX_REPLICAS];
    ostats_t *st_out;
    int *st_err;
    int phase;  /* 0 iterate, 1 build core rows, -1 exit */
    int core_pass, core_fp32;
    pthread_t *th;
    pthread_barrier_t bar;
};

/* owner shard of a determinant: the walker-hash accumulation of hash.F90:30-35
   over RandomHash2, mixed by the splitmix64 finaliser so that the owner is
   independent of the shard's own bucket index */
static int owner_of(const omt_t *m, const uint64_t *det)
{
    int32_t nI[O_MAX_NEL];
    o_decode(det, m->W, nI);
    uint64_t h = 0;
    for (int i = 0; i < m->s.nel; ++i)
        h = 1099511628211ull * h + (uint64_t)(int64_t)(m->s.random_hash2[nI[i] - 1] * (i + 1));
    return (int)(mix64(h) % (uint64_t)m->P);
}

static int cmp_det_w1(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return ilut_cmp(&x, &y, 1);
}
static int cmp_det_w2(const void *a, const void *b) { return ilut_cmp((const uint64_t *)a, (const uint64_t *)b, 2); }

/* rows [r0, r1) of the core Hamiltonian (the loop of o_set_core, one pass
   with a per-row count) */
static void core_rows_range(omt_t *m, int64_t r0, int64_t r1, int pass, int fp32)
{
    const int W = m->W, nel = m->s.nel;
    int32_t oi[O_MAX_NEL], oj[O_MAX_NEL];
    for (int64_t i = r0; i < r1; ++i) {
        const uint64_t *di = m->core_det + (size_t)i * W;
        o_decode(di, W, oi);
        int64_t k = pass ? m->rowptr[i] : 0;
        for (int64_t j = 0; j < m->ncore; ++j) {
            const uint64_t *dj = m->core_det + (size_t)j * W;
            if (excit_level(di, dj, W) > 2) continue;
            double h;
            if (i == j) {
                h = o_diag_helement(&m->s, oi) - m->s.hii;
            } else {
                o_decode(dj, W, oj);
                h = o_get_helement(&m->s, oi, oj, nel);
            }
            if (i != j && h == 0.0) continue;
            if (pass) {
                m->cols[k] = (int32_t)j;
                m->vals[k] = fp32 ? (double)(float)h : h;
            }
            k++;
        }
        if (!pass) m->rowptr[i + 1] = k;
    }
}

static void range_of(int64_t n, int P, int q, int64_t *a, int64_t *b)
{
    int64_t per = (n + P - 1) / P;
    *a = per * q < n ? per * q : n;
    *b = per * (q + 1) < n ? per * (q + 1) : n;
}

/* one iteration on shard q (the three phases, separated by barriers) */
static void shard_iterate(omt_t *m, int q)
{
    ostate_t *st = m->sh[q];
    ostats_t *out = &m->st_out[q];
    const int P = m->P, RL = m->RL, S = m->S;
    memset(out, 0, sizeof(*out));
    st->nspawned = 0;
    st->ifree_start = st->ifree_end = 0;
    int err = det_loop(st, m->iter, m->sft, out, m->ncore ? &m->ci : NULL);
    /* spawns into sections by destination (a stable counting sort), into spawned2 */
    int64_t *off = m->sendoff + (size_t)q * (P + 1);
    int32_t *own = (int32_t *)malloc(sizeof(int32_t) * (size_t)(st->nspawned ? st->nspawned : 1));
    for (int d = 0; d <= P; ++d) off[d] = 0;
    for (int64_t i = 0; i < st->nspawned; ++i) {
        own[i] = owner_of(m, st->spawned + (size_t)i * RL);
        off[own[i] + 1]++;
    }
    for (int d = 0; d < P; ++d) off[d + 1] += off[d];
    {
        int64_t pos[256];  /* P <= 256 (o_mt_create) */
        for (int d = 0; d < P; ++d) pos[d] = off[d];
        for (int64_t i = 0; i < st->nspawned; ++i)
            memcpy(st->spawned2 + (size_t)(pos[own[i]]++) * RL, st->spawned + (size_t)i * RL, sizeof(uint64_t) * RL);
    }
    free(own);
    /* the core amplitudes this shard holds */
    for (int64_t i = 0; i < m->ncore; ++i)
        if (m->core_owner[i] == q)
            memcpy(m->cx + (size_t)i * S, row_sign(st, cur_row(st, m->core_row[i])), sizeof(double) * S);
    pthread_barrier_wait(&m->bar);
    /* determ_projection, rows split over the shards */
    if (m->ncore) {
        /* rows a..b of y from the whole x, in o_core_project's summation order */
        int64_t a, b;
        range_of(m->ncore, P, q, &a, &b);
        for (int64_t i = a; i < b; ++i) {
            double p[64], t[64];
            for (int r = 0; r < S; ++r) {
                for (int l = 0; l < 64; ++l) p[l] = 0.0;
                for (int64_t k = m->rowptr[i]; k < m->rowptr[i + 1]; ++k) {
                    int l = (int)((k - m->rowptr[i]) & 63);
                    p[l] = p[l] - m->vals[k] * m->cx[(size_t)m->cols[k] * S + r];
                }
                for (int o = 32; o > 0; o >>= 1) {
                    for (int l = 0; l < 64; ++l) t[l] = p[l] + p[l ^ o];
                    memcpy(p, t, sizeof(p));
                }
                m->cy[(size_t)i * S + r] = (p[0] + m->sft[r] * m->cx[(size_t)i * S + r]) * m->s.tau;
            }
        }
    }
    /* receive: the sections addressed to q, in sender order */
    int64_t n = 0;
    for (int src = 0; src < P; ++src) n += m->sendoff[(size_t)src * (P + 1) + q + 1] - m->sendoff[(size_t)src * (P + 1) + q];
    if (n > st->maxs) err = err ? err : 1;
    pthread_barrier_wait(&m->bar);  /* every shard has finished its det loop before spawned is reused */
    if (!err) {
        int64_t k = 0;
        for (int src = 0; src < P; ++src) {
            const int64_t *o = m->sendoff + (size_t)src * (P + 1);
            int64_t c = o[q + 1] - o[q];
            memcpy(st->spawned + (size_t)k * RL, m->sh[src]->spawned2 + (size_t)o[q] * RL, sizeof(uint64_t) * RL * (size_t)c);
            k += c;
        }
        st->nspawned = n;
    }
    pthread_barrier_wait(&m->bar);  /* every section is copied before spawned2 is overwritten */
    if (!err) {
        int64_t nu = compress_spawned(st, st->nspawned, out);
        out->n_unique = nu;
        /* deterministic_annihilation for the core dets of this shard */
        for (int64_t i = 0; i < m->ncore; ++i) {
            if (m->core_owner[i] != q) continue;
            double *cs = row_sign(st, cur_row(st, m->core_row[i]));
            for (int r = 0; r < S; ++r) {
                double y = m->cy[(size_t)i * S + r], c = cs[r];
                cs[r] = y + c;
                out->nborn[r] += fabs(y);
                if (c * y < 0.0) {
                    double a = fabs(c), b = fabs(y);
                    out->nannihil[r] += 2.0 * (a < b ? a : b);
                }
            }
        }
        err = annihilate_spawned(st, nu, m->iter, out);
        if (!err) hash_table_stats(st, m->iter, out);
    }
    m->st_err[q] = err;
}

static void *shard_main(void *arg)
{
    void **a = (void **)arg;
    omt_t *m = (omt_t *)a[0];
    int q = (int)(intptr_t)a[1];
    free(arg);
    for (;;) {
        pthread_barrier_wait(&m->bar);  /* start */
        if (m->phase < 0) break;
        if (m->phase == 0) shard_iterate(m, q);
        if (m->phase == 1) {
            int64_t a, b;
            range_of(m->ncore, m->P, q, &a, &b);
            core_rows_range(m, a, b, m->core_pass, m->core_fp32);
        }
        pthread_barrier_wait(&m->bar);  /* done */
    }
    return NULL;
}

omt_t *o_mt_create(const osys_t *sys, int32_t nshards)
{
    if (nshards < 1 || nshards > 256) return NULL;
    omt_t *m = (omt_t *)calloc(1, sizeof(omt_t));
    m->s = *sys;
    m->P = nshards;
    m->W = nwords(sys->nbasis);
    m->S = sys->nrep;
    m->RL = m->W + m->S + 1;
    m->sh = (ostate_t **)calloc((size_t)nshards, sizeof(ostate_t *));
    osys_t ss = *sys;
    ss.nbuckets = (sys->nbuckets + nshards - 1) / nshards;
    if (ss.nbuckets < 1) ss.nbuckets = 1;
    ss.max_walkers = sys->max_walkers / nshards * 2 + 4096;
    ss.max_spawned = sys->max_spawned / nshards * 2 + 4096;
    for (int q = 0; q < nshards; ++q) m->sh[q] = o_create(&ss);
    m->sendoff = (int64_t *)calloc((size_t)nshards * (nshards + 1), sizeof(int64_t));
    m->st_out = (ostats_t *)calloc((size_t)nshards, sizeof(ostats_t));
    m->st_err = (int *)calloc((size_t)nshards, sizeof(int));
    m->th = (pthread_t *)calloc((size_t)nshards, sizeof(pthread_t));
    pthread_barrier_init(&m->bar, NULL, (unsigned)nshards);
    for (int q = 1; q < nshards; ++q) {
        void **arg = (void **)malloc(2 * sizeof(void *));
        arg[0] = m;
        arg[1] = (void *)(intptr_t)q;
        pthread_create(&m->th[q], NULL, shard_main, arg);
    }
    return m;
}

void o_mt_destroy(omt_t *m)
{
    if (!m) return;
    m->phase = -1;
    pthread_barrier_wait(&m->bar);
    for (int q = 1; q < m->P; ++q) pthread_join(m->th[q], NULL);
    pthread_barrier_destroy(&m->bar);
    for (int q = 0; q < m->P; ++q) o_destroy(m->sh[q]);
    free(m->sh); free(m->sendoff); free(m->st_out); free(m->st_err); free(m->th);
    free(m->core_det); free(m->rowptr); free(m->cols); free(m->vals); free(m->cx); free(m->cy);
    free(m->core_owner); free(m->core_row); free(m->core_sorted);
    free(m);
}

/* the rows of the core Hamiltonian split over the shards' threads */
static void core_rows_all(omt_t *m, int pass, int fp32)
{
    m->core_pass = pass;
    m->core_fp32 = fp32;
    m->phase = 1;
    pthread_barrier_wait(&m->bar);  /* start */
    int64_t a, b;
    range_of(m->ncore, m->P, 0, &a, &b);
    core_rows_range(m, a, b, pass, fp32);
    pthread_barrier_wait(&m->bar);  /* done */
}

int32_t o_mt_add_walkers(omt_t *m, const uint64_t *dets, const double *sgns, const uint32_t *flags, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t *d = dets + i * m->W;
        int rc = o_add_walker(m->sh[owner_of(m, d)], d, sgns + i * m->S, flags[i]);
        if (rc) return rc;
    }
    return 0;
}

int64_t o_mt_get_walkers(omt_t *m, uint64_t *dets, double *sgns, uint32_t *flags, int64_t cap)
{
    int64_t n = 0;
    for (int q = 0; q < m->P; ++q) {
        int64_t c = o_get_walkers(m->sh[q], n < cap ? dets + (size_t)n * m->W : dets, n < cap ? sgns + (size_t)n * m->S : sgns,
                                  n < cap ? flags + n : flags, NULL, NULL, cap - n > 0 ? cap - n : 0);
        n += c;
    }
    return n;
}

/* the core space, shared by the shards: the Hamiltonian rows are built by
   the calling thread (o_set_core's rule), each core det is placed in its
   owner shard */
int32_t o_mt_set_core(omt_t *m, const uint64_t *dets, int64_t n, int32_t fp32, int64_t *nnz_out)
{
    const int W = m->W, S = m->S;
    m->ncore = n;
    m->core_det = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n * W);
    memcpy(m->core_det, dets, sizeof(uint64_t) * (size_t)n * W);
    m->rowptr = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    core_rows_all(m, 0, fp32);
    for (int64_t i = 0; i < n; ++i) m->rowptr[i + 1] += m->rowptr[i];
    int64_t nnz = m->rowptr[n];
    m->cols = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    m->vals = (double *)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    core_rows_all(m, 1, fp32);
    m->cx = (double *)calloc((size_t)n * S, sizeof(double));
    m->cy = (double *)calloc((size_t)n * S, sizeof(double));
    m->core_owner = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    m->core_row = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    uint64_t cflags = FLAG_CORE;
    if (m->s.t_trunc_init)
        for (int r = 0; r < S; ++r) cflags |= 1ull << r;
    double zero[O_MAX_REPLICAS] = {0};
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t *d = dets + (size_t)i * W;
        int q = owner_of(m, d);
        ostate_t *st = m->sh[q];
        st->ifree_start = st->ifree_end = 0;
        int64_t b = bucket_of(st, d);
        int64_t pos = hash_lookup(st, d, b);
        if (pos < 0) {
            int32_t nI[O_MAX_NEL];
            o_decode(d, W, nI);
            int e = add_new_hash_det(st, d, zero, cflags, o_diag_helement(&st->s, nI), o_offdiag_helement(&st->s, d), b);
            if (e) return e;
            pos = hash_lookup(st, d, b);
        } else {
            *row_flag(st, cur_row(st, pos)) |= cflags;
        }
        m->core_owner[i] = q;
        m->core_row[i] = pos;
    }
    m->core_sorted = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n * W);
    memcpy(m->core_sorted, dets, sizeof(uint64_t) * (size_t)n * W);
    qsort(m->core_sorted, (size_t)n, sizeof(uint64_t) * W, W == 1 ? cmp_det_w1 : cmp_det_w2);
    m->ci.dets = m->core_sorted;
    m->ci.n = n;
    m->ci.W = W;
    if (nnz_out) *nnz_out = nnz;
    return 0;
}

int32_t o_mt_iterate(omt_t *m, uint32_t iter, const double *diag_sft, ostats_t *out)
{
    m->iter = iter;
    for (int r = 0; r < m->S; ++r) m->sft[r] = diag_sft[r];
    m->phase = 0;
    pthread_barrier_wait(&m->bar);  /* start */
    shard_iterate(m, 0);
    pthread_barrier_wait(&m->bar);  /* done */
    memset(out, 0, sizeof(*out));
    int err = 0;
    for (int q = 0; q < m->P; ++q) {
        const ostats_t *o = &m->st_out[q];
        const double *a = (const double *)o;
        double *b = (double *)out;
        const int nd = (int)(offsetof(ostats_t, tot_walkers) / sizeof(double));
        const int i_max = (int)(offsetof(ostats_t, max_cyc_spawn) / sizeof(double));
        const int i_bs = (int)(offsetof(ostats_t, bloom_size) / sizeof(double));
        for (int i = 0; i < nd; ++i) {
            if (i == i_max || (i >= i_bs && i < i_bs + 3)) b[i] = a[i] > b[i] ? a[i] : b[i];
            else b[i] += a[i];
        }
        out->tot_walkers += o->tot_walkers;
        out->n_attempts += o->n_attempts;
        out->n_spawned += o->n_spawned;
        out->n_unique += o->n_unique;
        out->n_new += o->n_new;
        if (m->st_err[q] && !err) err = m->st_err[q];
    }
    out->err = err;
    return err;
}
