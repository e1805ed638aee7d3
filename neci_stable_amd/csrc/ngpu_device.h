// ngpu_device.h -- device functions of the FCIQMC iteration engine (gfx950).
//
// Determinants are kept as W 64-bit words (orbital k -> word (k-1)/64, bit
// (k-1)%64, macros.h:16,74-77; odd k = beta, macros.h:21).  Nothing here
// materialises the occupied-orbital list nI: electron selection, Slater-Condon
// sums, fermionic parity and FindWalkerHash all iterate set bits in ascending
// order (the order of nI), so each lane works from its 8W-byte key in
// registers and no per-lane array spills to scratch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ngpu {

constexpr int kMaxRep = 16;
constexpr double kEps = 1e-13;               // constants.F90:28 EPS
constexpr uint32_t kFlagCore = 1u << 16;     // deterministic (core) determinant
constexpr uint32_t kFlagRemoved = 1u << 17;  // empty slot of the walker table
constexpr uint32_t kInfoInit = 1u << 8;      // spawn record: parent was initiator
constexpr uint32_t kInfoCoreParent = 1u << 9;
constexpr int kBucketEntries = 4;            // walker hash: entries per bucket
constexpr int32_t kSlotEmpty = -1;
constexpr int32_t kSlotTomb = -2;

// ------------------------------------------------------------------ RNG
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based.
struct U4 { uint32_t x, y, z, w; };

__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    U4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

__host__ __device__ inline uint64_t splitmix_fin(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <int W>
__host__ __device__ inline uint64_t det_hash64(const uint64_t* d) {
  if (W == 1) return d[0];
  return d[0] ^ splitmix_fin(d[1]);
}

__host__ __device__ inline double u01(uint32_t lo, uint32_t hi) {
  uint64_t x = ((uint64_t)hi << 32) | (uint64_t)lo;
  return (double)(x >> 11) * 0x1.0p-53;
}

// Two uniforms for (det, iteration, replica, k, purpose).  Purposes:
// 0 per-det (num-to-spawn, death); 1-3 spawn attempt k: 1 gives d0, d1, 2
// gives d3 (PCHB hole pair) and d5 (spawn rounding), 3 gives d2 and d4 (read
// only by FAST particle selection, heat-bath singles and the uniform
// generator); 4 new-det rounding; 5 statistics rounding.
template <int W>
__host__ __device__ inline void draw2(uint32_t seed, const uint64_t* det, uint32_t iter, int rep,
                                      uint32_t k, uint32_t purpose, double& ua, double& ub) {
  uint64_t dh = det_hash64<W>(det);
  U4 c{(uint32_t)dh, (uint32_t)(dh >> 32), iter, (k << 3) | purpose};
  U4 o = philox4x32_10(c, seed, (uint32_t)rep);
  ua = u01(o.x, o.y);
  ub = u01(o.z, o.w);
}

// ------------------------------------------------------------ bit helpers
__host__ __device__ inline int popc64(uint64_t x) { return __builtin_popcountll(x); }
__host__ __device__ inline int ctz64(uint64_t x) { return __builtin_ctzll(x); }

template <int W>
__host__ __device__ inline bool test_orb(const uint64_t* d, int orb) {
  return (d[(orb - 1) >> 6] >> ((orb - 1) & 63)) & 1ull;
}
template <int W>
__host__ __device__ inline void flip_orb(uint64_t* d, int orb) {
  d[(orb - 1) >> 6] ^= 1ull << ((orb - 1) & 63);
}
template <int W>
__host__ __device__ inline bool det_eq(const uint64_t* a, const uint64_t* b) {
  bool e = true;
#pragma unroll
  for (int w = 0; w < W; ++w) e = e && (a[w] == b[w]);
  return e;
}
template <int W>
__host__ __device__ inline int popc_det(const uint64_t* d) {
  int c = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) c += popc64(d[w]);
  return c;
}
// FindBitExcitLevel(a, b) = popcount(a & ~b)   (DetBitOps.F90:140-180)
template <int W>
__host__ __device__ inline int excit_level(const uint64_t* a, const uint64_t* b) {
  int c = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) c += popc64(a[w] & ~b[w]);
  return c;
}

// rank-k (0-based) set bit of x, x has more than k bits set
__host__ __device__ inline int select64(uint64_t x, int k) {
  int pos = 0;
  int c = popc64(x & 0xFFFFFFFFull);
  if (k >= c) { k -= c; x >>= 32; pos += 32; }
  c = popc64(x & 0xFFFFull);
  if (k >= c) { k -= c; x >>= 16; pos += 16; }
  c = popc64(x & 0xFFull);
  if (k >= c) { k -= c; x >>= 8; pos += 8; }
  c = popc64(x & 0xFull);
  if (k >= c) { k -= c; x >>= 4; pos += 4; }
  c = popc64(x & 0x3ull);
  if (k >= c) { k -= c; x >>= 2; pos += 2; }
  c = (int)(x & 1ull);
  if (k >= c) { pos += 1; }
  return pos;
}

// orbital (1-based) of the e-th (1-based) set bit of the masked determinant
template <int W>
__host__ __device__ inline int select_orb(const uint64_t* d, uint64_t mask, int e) {
  int k = e - 1;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    uint64_t x = d[w] & mask;
    int c = popc64(x);
    if (k < c) return 64 * w + select64(x, k) + 1;
    k -= c;
  }
  return 0;
}

// number of occupied orbitals strictly between orbitals a and b
template <int W>
__host__ __device__ inline int occ_between(const uint64_t* d, int a, int b) {
  int lo = a < b ? a : b, hi = a < b ? b : a;  // count bits lo+1 .. hi-1 (1-based)
  int c = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    int base = 64 * w;  // orbitals base+1 .. base+64
    int l = lo - base;  // bits with orbital > lo  -> bit index >= l
    int h = hi - base;  // orbital < hi -> bit index <= h-2
    uint64_t m = ~0ull;
    if (l >= 64) m = 0; else if (l > 0) m &= ~0ull << l;
    if (h <= 1) m = 0; else if (h - 1 < 64) m &= (~0ull >> (64 - (h - 1)));
    c += popc64(d[w] & m);
  }
  return c;
}

// make_single parity (excit_parity.F90:15-76): odd number of occupied
// orbitals strictly between src and tgt
template <int W>
__host__ __device__ inline bool single_parity(const uint64_t* d, int src, int tgt) {
  return occ_between<W>(d, src, tgt) & 1;
}
// make_double parity (excit_parity.F90:78-172) for sorted s1<s2, t1<t2:
// move s1->t1 in D, then s2->t2 in the intermediate determinant
template <int W>
__host__ __device__ inline bool double_parity(const uint64_t* d, int s1, int s2, int t1, int t2) {
  uint64_t d1[W];
#pragma unroll
  for (int w = 0; w < W; ++w) d1[w] = d[w];
  int p1 = occ_between<W>(d, s1, t1);
  flip_orb<W>(d1, s1);
  flip_orb<W>(d1, t1);
  int p2 = occ_between<W>(d1, s2, t2);
  return (p1 + p2) & 1;
}

// ----------------------------------------------------------- hashing
// FindWalkerHash (hash.F90:23-37): wrapping int64 accumulation over the
// occupied orbitals in ascending order, bucket = |mod(h, L)| + 1 (1-based).
// The RandomHash2 values of four orbitals are loaded before the four steps
// that use them (the steps themselves stay in orbital order).
template <int W>
__host__ __device__ inline int64_t find_walker_hash(const uint64_t* d, const int32_t* rh, int64_t L) {
  uint64_t h = 0;
  int i = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    uint64_t x = d[w];
    while (x) {
      int orb[4], k = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        orb[q] = 0;
        if (x) {
          orb[q] = 64 * w + ctz64(x) + 1;
          x &= x - 1;
          k = q + 1;
        }
      }
      int32_t r[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) r[q] = q < k ? rh[orb[q] - 1] : 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q < k) {
          ++i;
          int32_t term = r[q] * i;
          h = 1099511628211ull * h + (uint64_t)(int64_t)term;
        }
      }
    }
  }
  int64_t m = (int64_t)h % L;
  if (m < 0) m = -m;
  return m + 1;
}

// --------------------------------------------------------- integrals
// UMatInd_base (UMatCache.F90:257-303), real hermitian integrals, 1-based
__host__ __device__ inline int64_t umat_ind(int i, int j, int k, int l) {
  int64_t A = i > k ? ((int64_t)i * (i - 1)) / 2 + k : ((int64_t)k * (k - 1)) / 2 + i;
  int64_t B = j > l ? ((int64_t)j * (j - 1)) / 2 + l : ((int64_t)l * (l - 1)) / 2 + j;
  return A > B ? (A * (A - 1)) / 2 + B : (B * (B - 1)) / 2 + A;
}

__host__ __device__ inline bool is_beta(int orb) { return orb & 1; }
__host__ __device__ inline int spat(int orb) { return (orb + 1) >> 1; }

struct AliasEnt {  // one alias-table entry (aliasSampling.F90 bias/alias pair)
  double bias;
  double prob_self;   // probs(i) of this entry
  double prob_alias;  // probs(alias(i)): a draw reads one 32-byte entry
  int32_t alias;      // 0-based
  int32_t pad;
};

// Read-only system description passed by value to kernels.
struct DevSys {
  int nel, nbasis, S, model;
  int t_init, t_real, t_rsc, singles;
  int particles, nsites, maxnbr, ijmax;
  int ncore;  // semistochastic core dets (0: none)
  int uniform;  // molecular generator: 0 PCHB + heat-bath singles, 1 gen_rand_excit
  double init_walk_no, rsc, occ_thresh, avmc;
  double hii, ecore, bhub, uhub;
  double psingles, pdoubles, pparallel;
  uint32_t seed;
  // spin counts of the reference determinant (every determinant of the run
  // has them) and pick_biased_elecs's two pair probabilities for those
  // counts, pparallel / #parallel pairs and (1 - pparallel) / #opposite
  // pairs, divided once on the host (the same IEEE quotients)
  int nalpha_ref, nbeta_ref;
  double pg_par_ref, pg_anti_ref;
  uint64_t ref[2];
  int64_t L;                   // walker hash buckets
  const int32_t* rh;           // RandomHash2
  const int32_t* nbr;          // lattice
  const double* tmat;          // TMAT2D column major
  const double* umat;          // packed UMAT, [UMatInd-1]
  const AliasEnt* ab_tab;      // [ijmax][ijmax]
  const double* ab_prob;
  const uint8_t* ab_ok;
  const AliasEnt* j_tab;       // [nbasis][nbasis]
  const double* j_prob;
  const uint8_t* j_ok;
  const int32_t* ab_tgt;       // tgtOrbs: ab_tgt[2*(ab-1)] = B <= A = ab_tgt[2*(ab-1)+1]
  // weighted PCHB samplers: holes 0 FAST-FAST (ab_*), 1 FULL-FULL (a_*, b_*);
  // particles 2 UNIF-FULL, 3 FULL-FULL (i_* and j_*)
  int holes;
  int i_ok;
  const AliasEnt* i_tab;       // first particle over the spin orbitals
  const double* i_prob;
  const AliasEnt* a_tab;       // first hole per IJ: row IJ-1 of nbasis entries
  const double* a_prob;
  const uint8_t* a_ok;
  const AliasEnt* b_tab;       // second hole per (IJ, A): row (IJ-1)*nbasis + A-1
  const double* b_prob;
  const uint8_t* b_ok;
  // heat-bath singles: (U(a,j,b,j), U(a,j,j,b)) at jk[2*(((a-1)*nsp + j-1)*nsp + b-1)]
  // for spatial a, j, b; copies of UMAT entries, rows contiguous in b
  const double* jk;
  // diagonal elements: (U(i,j,i,j), U(i,j,j,i)) at jk0[2*((i-1)*nsp + j-1)]
  const double* jk0;
  // heat-bath single weights relative to the reference determinant:
  // fref[(p-1)*nsp + b-1] = sltcnd_1(ref, p, b) for every spin orbital p and
  // the spatial orbital b taken with p's spin
  const double* fref;
  int nsp;
  // C(p, i) at binom[p*nel1 + i] when spawns are sorted by colex rank, else null
  const uint32_t* binom;
  int nel1;
  // C(p, i) in 64 bits when spawns are sorted by 64-bit colex rank, else null
  const uint64_t* binom64;
  // spawned-list records: FP64 amplitudes in 4 words (key, key or 0, value,
  // info), or FP32 amplitudes in W + 1 words (key words, then info << 32 |
  // FP32 bits); sums over spawns stay FP64 either way
  int spawn_fp32, rec_words;
};

__device__ __host__ inline double U(const DevSys& s, int i, int j, int k, int l) {
  return s.umat[umat_ind(i, j, k, l) - 1];
}
__device__ __host__ inline double T(const DevSys& s, int i, int j) {
  return s.tmat[(size_t)(i - 1) + (size_t)(j - 1) * (size_t)s.nbasis];
}

// Accumulates the values at(o) over the set bits o of x (1-based orbital
// base + bit + 1, ascending), adding (sub = false) or subtracting them in bit
// order; at(o) returns a pointer, or null to skip o.  Four values are loaded
// before any of them is used, so a chain of n loads takes n/4 round trips
// while the sum stays bit-identical to one bit at a time.
template <typename F>
__device__ __host__ inline double accum_bits4(uint64_t x, int base, double acc, bool sub, F at) {
  while (x) {
    const double* p[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      p[q] = nullptr;
      if (x) {
        p[q] = at(base + ctz64(x) + 1);
        x &= x - 1;
      }
    }
    double v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = p[q] ? *p[q] : 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (p[q]) acc = sub ? acc - v[q] : acc + v[q];
  }
  return acc;
}

// sltcnd_0_base (sltcnd.fpp:585-622): ordering of the FP sums follows the
// reference loops exactly (inner partial sum per i, then exchange, then one-body)
template <int W>
__device__ __host__ inline double sltcnd_0(const DevSys& s, const uint64_t* d) {
  double hel_sing = 0.0;
#pragma unroll
  for (int w = 0; w < W; ++w)
    hel_sing = accum_bits4(d[w], 64 * w, hel_sing, false, [&](int o) { return &s.tmat[(size_t)(o - 1) * (s.nbasis + 1)]; });
  double hel_doub = 0.0, hel_tmp = 0.0;
  // i runs over electrons; the inner loops run over electrons above i
#pragma unroll
  for (int wi = 0; wi < W; ++wi) {
    uint64_t xi = d[wi];
    while (xi) {
      int oi = 64 * wi + ctz64(xi) + 1;
      xi &= xi - 1;
      const double* row = s.jk0 + 2 * (size_t)(spat(oi) - 1) * s.nsp;
      double t = 0.0;
      bool any = false;
#pragma unroll
      for (int wj = wi; wj < W; ++wj) {
        uint64_t xj = (wj == wi) ? xi : d[wj];
        any = any || xj != 0;
        t = accum_bits4(xj, 64 * wj, t, false, [&](int oj) { return row + 2 * (spat(oj) - 1); });  // U(i,j,i,j)
      }
      if (any) hel_doub += t;
    }
  }
#pragma unroll
  for (int wi = 0; wi < W; ++wi) {
    uint64_t xi = d[wi];
    while (xi) {
      int oi = 64 * wi + ctz64(xi) + 1;
      xi &= xi - 1;
      const double* row = s.jk0 + 2 * (size_t)(spat(oi) - 1) * s.nsp + 1;
      // exchange only between equal spins: bits of the other spin are masked
      uint64_t same = is_beta(oi) ? 0x5555555555555555ull : 0xAAAAAAAAAAAAAAAAull;
#pragma unroll
      for (int wj = wi; wj < W; ++wj) {
        uint64_t xj = ((wj == wi) ? xi : d[wj]) & same;
        hel_tmp = accum_bits4(xj, 64 * wj, hel_tmp, true, [&](int oj) { return row + 2 * (spat(oj) - 1); });  // U(i,j,j,i)
      }
    }
  }
  return hel_doub + hel_tmp + hel_sing;
}

// sltcnd_1_kernel (sltcnd.fpp:643-680), no parity
template <int W>
__device__ __host__ inline double sltcnd_1(const DevSys& s, const uint64_t* d, int src, int tgt) {
  int a = spat(src), b = spat(tgt);
  double hel = 0.0;
  if (is_beta(src) == is_beta(tgt)) {
    const double* base = s.jk + 2 * ((size_t)(a - 1) * s.nsp * s.nsp + (size_t)(b - 1));
    const size_t step = 2 * (size_t)s.nsp;
    auto coul = [&](int o) -> const double* { return o == src ? nullptr : base + step * (size_t)(spat(o) - 1); };
#pragma unroll
    for (int w = 0; w < W; ++w) hel = accum_bits4(d[w], 64 * w, hel, false, coul);  // U(a,id,b,id)
    uint64_t same = is_beta(src) ? 0x5555555555555555ull : 0xAAAAAAAAAAAAAAAAull;
    auto exch = [&](int o) -> const double* { return o == src ? nullptr : base + step * (size_t)(spat(o) - 1) + 1; };
#pragma unroll
    for (int w = 0; w < W; ++w) hel = accum_bits4(d[w] & same, 64 * w, hel, true, exch);  // U(a,id,id,b)
  }
  hel += T(s, src, tgt);
  return hel;
}

// sltcnd_2_kernel (sltcnd.fpp:690-711), ex = [s1 s2; t1 t2]
__device__ __host__ inline double sltcnd_2(const DevSys& s, int s1, int s2, int t1, int t2) {
  double hel;
  if (is_beta(s1) == is_beta(t1) && is_beta(s2) == is_beta(t2))
    hel = U(s, spat(s1), spat(s2), spat(t1), spat(t2));
  else
    hel = 0.0;
  if (is_beta(s1) == is_beta(t2) && is_beta(s2) == is_beta(t1))
    hel -= U(s, spat(s1), spat(s2), spat(t2), spat(t1));
  return hel;
}

__device__ __host__ inline bool hub_neighbours(const DevSys& s, int sa, int sb) {
  for (int k = 0; k < s.maxnbr; ++k) {
    int n = s.nbr[(sa - 1) * s.maxnbr + k];
    if (n == 0) break;
    if (n == sb) return true;
  }
  return false;
}

// get_helement_det_only / get_helement_lattice for an excitation
// ic = 1: (src, tgt); ic = 2: (s1, s2, t1, t2) sorted
template <int W>
__device__ __host__ inline double helement_ex(const DevSys& s, const uint64_t* d, int ic, int e0,
                                              int e1, int e2, int e3, bool par) {
  double hel;
  if (s.model == 0) {
    if (ic != 1) return 0.0;
    hel = (is_beta(e0) == is_beta(e1) && hub_neighbours(s, spat(e0), spat(e1))) ? s.bhub : 0.0;
  } else if (ic == 1) {
    hel = sltcnd_1<W>(s, d, e0, e1);
  } else {
    hel = sltcnd_2(s, e0, e1, e2, e3);
  }
  return par ? -hel : hel;
}

// get_diagonal_matel (total H_ii)
template <int W>
__device__ __host__ inline double diag_helement(const DevSys& s, const uint64_t* d) {
  if (s.model == 0) {
    // real_space_hubbard.F90:2303-2321: U times doubly occupied sites
    int nd = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) nd += popc64(d[w] & (d[w] >> 1) & 0x5555555555555555ull);
    return s.uhub * (double)nd;
  }
  return sltcnd_0<W>(s, d) + s.ecore;
}

// general H(I, J) through the excitation between two bitstrings
template <int W>
__device__ __host__ inline double helement_pair(const DevSys& s, const uint64_t* dI, const uint64_t* dJ) {
  int ic = excit_level<W>(dI, dJ);
  if (ic == 0) return diag_helement<W>(s, dI);
  if (ic > 2) return 0.0;
  uint64_t src[W], tgt[W];
#pragma unroll
  for (int w = 0; w < W; ++w) { src[w] = dI[w] & ~dJ[w]; tgt[w] = dJ[w] & ~dI[w]; }
  int s1 = select_orb<W>(src, ~0ull, 1), t1 = select_orb<W>(tgt, ~0ull, 1);
  if (ic == 1) return helement_ex<W>(s, dI, 1, s1, t1, 0, 0, single_parity<W>(dI, s1, t1));
  int s2 = select_orb<W>(src, ~0ull, 2), t2 = select_orb<W>(tgt, ~0ull, 2);
  return helement_ex<W>(s, dI, 2, s1, s2, t1, t2, double_parity<W>(dI, s1, s2, t1, t2));
}

// get_off_diagonal_matel (matel_getter.F90:56-107), tNoBrillouin: H(det, ref)
template <int W>
__device__ __host__ inline double offdiag_helement(const DevSys& s, const uint64_t* d) {
  int ex = excit_level<W>(s.ref, d);
  if (ex != 1 && ex != 2) return 0.0;
  return helement_pair<W>(s, d, s.ref);
}

// ------------------------------------------------ excitation generation
struct Excit {
  int ic;          // 0 = null
  int e0, e1, e2, e3;
  bool parity;
  bool hel_ok;     // hel holds the unsigned-parity matrix element
  double pgen;
  double hel;
};

// sample_AliasTable_t (aliasSampling.F90:310-332), 0-based
__device__ __host__ inline int alias_sample(const AliasEnt* tab, int n, double r, double& prob) {
  double x = (double)n * r;
  int pos = (int)x + 1;
  double b = x + 1.0 - (double)pos;
  if (b < 0.0) b = 0.0;
  AliasEnt e = tab[pos - 1];
  bool self = b < e.bias;
  prob = self ? e.prob_self : e.prob_alias;
  return self ? pos - 1 : e.alias;
}

__device__ __host__ inline int fuse_index(int x, int y) {
  return x < y ? x + y * (y - 1) / 2 : y + x * (x - 1) / 2;
}

// gen_excit_rs_hubbard (real_space_hubbard.F90:1938-2033)
template <int W>
__device__ __host__ inline Excit gen_hubbard(const DevSys& s, const uint64_t* d, const double* dr) {
  Excit ex{};
  int nel = s.nel;
  int elec = 1 + (int)(dr[0] * (double)nel);
  double p_elec = 1.0 / (double)nel;
  int src = select_orb<W>(d, ~0ull, elec);
  int site = spat(src);
  double cum_sum = 0.0;
  double a = fabs(s.bhub);
  for (int k = 0; k < s.maxnbr; ++k) {
    int ns = s.nbr[(site - 1) * s.maxnbr + k];
    if (ns == 0) break;
    int orb = is_beta(src) ? 2 * ns - 1 : 2 * ns;
    if (!test_orb<W>(d, orb)) cum_sum += a;
  }
  if (cum_sum < kEps) return ex;
  double r = dr[1] * cum_sum;
  double cum = 0.0, prev = 0.0;
  int orb = 0, k = 0;
  for (; k < s.maxnbr; ++k) {
    int ns = s.nbr[(site - 1) * s.maxnbr + k];
    if (ns == 0) break;
    int o = is_beta(src) ? 2 * ns - 1 : 2 * ns;
    prev = cum;
    if (!test_orb<W>(d, o)) cum += a;
    orb = o;
    bool last = (k + 1 >= s.maxnbr) || s.nbr[(site - 1) * s.maxnbr + k + 1] == 0;
    if (!(cum < r) || last) break;
  }
  if (cum - prev == 0.0) return ex;
  double p_orb = (cum - prev) / cum_sum;
  ex.ic = 1;
  ex.e0 = src;
  ex.e1 = orb;
  ex.parity = single_parity<W>(d, src, orb);
  ex.pgen = p_elec * p_orb;
  return ex;
}

// The heat-bath weight |H(src -> orb)| of gen_single_4ind_ex
// (excit_gens_int_weighted.F90:442-620) evaluated relative to the reference
// determinant: H = sltcnd_1(ref, src, orb) (the fref table) plus, for every
// orbital j that d occupies and the reference does not (ascending, src
// skipped), U(a,j,b,j) - U(a,j,j,b) [same spin as src only for the second],
// then minus the same for every orbital the reference occupies and d does
// not.  The same value as sltcnd_1(d, src, orb) up to FP64 rounding (exact
// zeros stay exact: every term of a zero element is zero); its cost is
// 2 * (excitation level of d) rows instead of nel.  The signed value of the
// selected orbital is also the element the single spawns with (the
// generator's HElGen), so it is not summed a second time.
template <int W>
__device__ __host__ inline double hb_weight(const DevSys& s, const uint64_t* d, int src, int orb) {
  const int a = spat(src), b = spat(orb), nsp = s.nsp;
  double v = s.fref[(size_t)(src - 1) * nsp + (b - 1)];
  const double* base = s.jk + 2 * ((size_t)(a - 1) * nsp * nsp + (size_t)(b - 1));
  const size_t step = 2 * (size_t)nsp;
  const bool beta = is_beta(src);
#pragma unroll
  for (int w = 0; w < W; ++w) {
    uint64_t x = d[w] & ~s.ref[w];
    if ((src - 1) >> 6 == w) x &= ~(1ull << ((src - 1) & 63));
    for (; x; x &= x - 1) {
      const int o = 64 * w + ctz64(x) + 1;
      const double* e = base + step * (size_t)(spat(o) - 1);
      v += e[0];
      if (is_beta(o) == beta) v -= e[1];
    }
  }
#pragma unroll
  for (int w = 0; w < W; ++w) {
    uint64_t x = s.ref[w] & ~d[w];
    if ((src - 1) >> 6 == w) x &= ~(1ull << ((src - 1) & 63));
    for (; x; x &= x - 1) {
      const int o = 64 * w + ctz64(x) + 1;
      const double* e = base + step * (size_t)(spat(o) - 1);
      v -= e[0];
      if (is_beta(o) == beta) v += e[1];
    }
  }
  return v;
}

// gen_single_4ind_ex + select_orb_sing (excit_gens_int_weighted.F90:442-620),
// weights by hb_weight.  The cumulative list is recomputed in a second pass in
// the same order instead of being stored, which reproduces the same FP values.
template <int W>
__device__ __host__ inline Excit gen_single_heatbath(const DevSys& s, const uint64_t* d, double dsel,
                                                     double dorb) {
  Excit ex{};
  int nel = s.nel;
  int elec = 1 + (int)floor(dsel * (double)nel);
  int src = select_orb<W>(d, ~0ull, elec);
  int nsp = s.nbasis / 2;
  int off = is_beta(src) ? 1 : 0;
  double cum_sum = 0.0;
  for (int i = 1; i <= nsp; ++i) {
    int orb = 2 * i - off;
    double cpt = 0.0;
    if (!test_orb<W>(d, orb)) cpt = fabs(hb_weight<W>(s, d, src, orb));
    cum_sum += cpt;
  }
  if (fabs(cum_sum) <= kEps) return ex;
  double r = dorb * cum_sum;
  double cum = 0.0, cpt = 0.0, v = 0.0;
  int orb = 0;
  for (int i = 1; i <= nsp; ++i) {
    orb = 2 * i - off;
    v = test_orb<W>(d, orb) ? 0.0 : hb_weight<W>(s, d, src, orb);
    cpt = fabs(v);
    cum += cpt;
    if (!(cum < r) || i == nsp) break;
  }
  if (cpt == 0.0) return ex;
  double pg = cpt / cum_sum;
  pg = pg / (double)nel;
  ex.ic = 1;
  ex.e0 = src;
  ex.e1 = orb;
  ex.parity = single_parity<W>(d, src, orb);
  ex.pgen = pg;
  ex.hel = v;  // the element the spawn uses (hb_weight's sum, no parity)
  ex.hel_ok = true;
  return ex;
}

// pick_biased_elecs (excit_gens_int_weighted.F90:722-848) on bitstrings: the
// pair index is decoded with integer arithmetic equivalent to the reference's
// ceiling((1+sqrt(9+8*idx))/2)
template <int W>
__device__ __host__ inline void pick_biased_elecs(const DevSys& s, const uint64_t* d, double r, int& o1,
                                                  int& o2, double& pgen) {
  const uint64_t kA = 0xAAAAAAAAAAAAAAAAull, kB = 0x5555555555555555ull;
  int nalpha = 0, nbeta = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) { nalpha += popc64(d[w] & kA); nbeta += popc64(d[w] & kB); }
  int aa = nalpha * (nalpha - 1) / 2, bb = nbeta * (nbeta - 1) / 2;
  int par = aa + bb, ab = nalpha * nbeta;
  double pb = s.pparallel;
  o1 = o2 = 0;
  pgen = 0.0;
  if (r < pb ? par == 0 : ab == 0) return;  // no pair of the drawn spin type
  const bool ref_counts = nalpha == s.nalpha_ref && nbeta == s.nbeta_ref;
  if (r < pb) {
    r = (r / pb) * (double)par;
    int idx = (int)floor(r);
    uint64_t mask = kA;
    if (idx >= aa) { idx -= aa; mask = kB; }
    // the smallest n1 >= 2 with n1 (n1 - 1) / 2 > idx: the root of the
    // quadratic, then one integer correction each way
    int n1 = (int)floorf(0.5f * (1.0f + sqrtf(8.0f * (float)idx + 1.0f))) + 1;
    if (n1 < 2) n1 = 2;
    if (n1 * (n1 - 1) / 2 <= idx) ++n1;
    if (n1 > 2 && (n1 - 1) * (n1 - 2) / 2 > idx) --n1;
    int n2 = idx + 1 - ((n1 - 1) * (n1 - 2)) / 2;
    o1 = select_orb<W>(d, mask, n2);
    o2 = select_orb<W>(d, mask, n1);
    pgen = ref_counts ? s.pg_par_ref : pb / (double)par;
  } else {
    pgen = ref_counts ? s.pg_anti_ref : (1.0 - pb) / (double)ab;
    r = ((r - pb) / (1.0 - pb)) * (double)ab;
    int idx = (int)floor(r);
    int na = 1 + idx % nalpha;
    int nb = 1 + (int)floor((double)idx / (double)nalpha);
    int oa = select_orb<W>(d, kA, na), ob = select_orb<W>(d, kB, nb);
    o1 = oa < ob ? oa : ob;
    o2 = oa < ob ? ob : oa;
  }
}

// GAS_doubles_PCHB_spinorb_gen_exc (gasci_pchb_doubles_spinorb_fastweighted.fpp:131-244)
template <int W>
__device__ __host__ inline Excit gen_double_pchb(const DevSys& s, const uint64_t* d, const double* dr) {
  Excit ex{};
  int nel = s.nel, n = s.nbasis;
  int s1, s2;
  double p;
  if (s.particles == 1) {
    // draw_PC_FastWeightedParticles_t (gasci_pchb_doubles_select_particles.fpp:515-560)
    int e1 = (int)(dr[1] * (double)nel) + 1;
    s1 = select_orb<W>(d, ~0ull, e1);
    if (!s.j_ok[s1 - 1]) return ex;
    const AliasEnt* row = s.j_tab + (size_t)(s1 - 1) * n;
    double p_J_I;
    int j = alias_sample(row, n, dr[2], p_J_I);
    s2 = j + 1;
    if (!test_orb<W>(d, s2)) return ex;
    double p_I_J = s.j_ok[s2 - 1] ? s.j_prob[(size_t)(s2 - 1) * n + (s1 - 1)] : 0.0;
    p = (p_J_I + p_I_J) / (double)nel;
    if (s1 > s2) { int t = s1; s1 = s2; s2 = t; }
  } else {
    pick_biased_elecs<W>(s, d, dr[1], s1, s2, p);
    if (s1 == 0 || s2 == 0) return ex;
  }
  int ij = fuse_index(s1, s2);
  if (!s.ab_ok[ij - 1]) return ex;
  const AliasEnt* row = s.ab_tab + (size_t)(ij - 1) * s.ijmax;
  double pgh;
  int ab = alias_sample(row, s.ijmax, dr[3], pgh) + 1;
  int t1 = s.ab_tgt[2 * (ab - 1)], t2 = s.ab_tgt[2 * (ab - 1) + 1];
  if (t1 == t2 || test_orb<W>(d, t1) || test_orb<W>(d, t2)) return ex;
  if (fabs(pgh) <= kEps) return ex;
  ex.ic = 2;
  ex.e0 = s1; ex.e1 = s2; ex.e2 = t1; ex.e3 = t2;
  ex.parity = double_parity<W>(d, s1, s2, t1, t2);
  ex.pgen = p * pgh;
  return ex;
}

// Uniform draws for generators with rejection loops (a varying number of
// draws): u_1..u_4 are the attempt's d1..d4 (d5 stays the spawn's rounding
// draw); u_5, u_6, ... come from Philox blocks with purpose 6 and key word 1 =
// replica | (block + 1) << 8, two uniforms per block.
template <int W>
struct DrawSeq {
  double d1, d2, d3, d4;  // the attempt's d1..d4, held by value
  uint32_t seed;
  const uint64_t* key;
  uint32_t iter;
  int rep;
  uint32_t k;
  int n;
  int blk;
  double c0, c1;
  __device__ __host__ DrawSeq(const double* d, uint32_t sd, const uint64_t* ky, uint32_t it, int r, uint32_t kk)
      : d1(d[1]), d2(d[2]), d3(d[3]), d4(d[4]), seed(sd), key(ky), iter(it), rep(r), k(kk), n(0), blk(-1),
        c0(0.0), c1(0.0) {}
  __device__ __host__ double next() {
    ++n;
    if (n <= 4) return n == 1 ? d1 : n == 2 ? d2 : n == 3 ? d3 : d4;
    int j = n - 5, b = j >> 1;
    if (b != blk) {
      uint64_t dh = det_hash64<W>(key);
      U4 c{(uint32_t)dh, (uint32_t)(dh >> 32), iter, (k << 3) | 6u};
      U4 o = philox4x32_10(c, seed, (uint32_t)rep | ((uint32_t)(b + 1) << 8));
      c0 = u01(o.x, o.y);
      c1 = u01(o.z, o.w);
      blk = b;
    }
    return (j & 1) ? c1 : c0;
  }
};

// ------------------------------------------ weighted PCHB particles and holes
// Restatement of gasci_pchb_doubles_select_particles.fpp (UNIF-FULL :427-500,
// FULL-FULL :333-425) and gasci_pchb_doubles_spinorb_fullyweighted.fpp
// (GAS_doubles_PCHB_gen_exc / get_pgen :138-296), every uniform from the
// attempt's draw sequence: particles first, then holes.

constexpr int kMaxRedraw = 1024;  // alias redraws before a constrained draw gives up

// spin orbitals not in d (within nbasis)
template <int W>
__device__ __host__ inline void unocc_mask(const DevSys& s, const uint64_t* d, uint64_t* u) {
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const int nb = s.nbasis - 64 * w;
    const uint64_t valid = nb >= 64 ? ~0ull : (nb <= 0 ? 0ull : ((1ull << nb) - 1ull));
    u[w] = ~d[w] & valid;
  }
}

// sum(get_prob(contain)) of one sampler row over the set bits of x, in
// ascending order (row == nullptr: an empty sampler, every probability 0)
template <int W>
__device__ __host__ inline double set_sum(const double* row, const uint64_t* x) {
  if (!row) return 0.0;
  double r = 0.0;
#pragma unroll
  for (int w = 0; w < W; ++w) r = accum_bits4(x[w], 64 * w, r, false, [&](int o) { return &row[o - 1]; });
  return r;
}

// constrained_sample_fast (aliasSampling.F90:505-535): below the redrawing
// cutoff 0.1 the CDF sampler over prob(contain)/renorm (CDF_sampling.fpp:
// first cumulated value >= r), else alias draws until one lies in the set.
// Returns the orbital (0: none) and its probability; a pick of probability
// zero (a cumulated value at r == 0, or r beyond a rounded last cumulated
// value) is a null draw.
template <int W>
__device__ __host__ inline int constrained_draw(const AliasEnt* tab, const double* prob, int n, const uint64_t* set,
                                                double renorm, DrawSeq<W>& u, double& p) {
  if (!prob || fabs(renorm) <= kEps) return 0;
  if (renorm < 0.1) {
    const double r = u.next();
    double cum = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w)
      for (uint64_t x = set[w]; x; x &= x - 1) {
        const int o = 64 * w + ctz64(x) + 1;
        const double pk = prob[o - 1] / renorm;
        cum += pk;
        if (!(cum < r)) {
          if (pk == 0.0) return 0;
          p = pk;
          return o;
        }
      }
    return 0;
  }
  for (int t = 0; t < kMaxRedraw; ++t) {
    double pa;
    const int tgt = alias_sample(tab, n, u.next(), pa) + 1;
    if (!test_orb<W>(set, tgt)) continue;
    p = prob[tgt - 1] / renorm;
    return p == 0.0 ? 0 : tgt;
  }
  return 0;
}

__device__ __host__ inline const double* j_row(const DevSys& s, int I) {
  return s.j_ok[I - 1] ? s.j_prob + (size_t)(I - 1) * s.nbasis : nullptr;
}

// the ordered particle pair and p({I, J}) of UNIF-FULL / FULL-FULL; false: null
template <int W>
__device__ __host__ inline bool draw_weighted_particles(const DevSys& s, const uint64_t* d, DrawSeq<W>& u, int& s1,
                                                        int& s2, double& p) {
  const int n = s.nbasis, nel = s.nel;
  const bool full = s.particles == 3;
  double pf0 = 0.0, ps0 = 0.0, rf = 0.0;
  if (full) {
    const double* ip = s.i_ok ? s.i_prob : nullptr;
    rf = set_sum<W>(ip, d);
    s1 = constrained_draw<W>(s.i_tab, ip, n, d, rf, u, pf0);
    if (!s1) return false;
  } else {
    const int e1 = (int)(u.next() * (double)nel) + 1;
    s1 = select_orb<W>(d, ~0ull, e1);
  }
  const double* j1 = j_row(s, s1);
  const double rs1 = set_sum<W>(j1, d);
  s2 = constrained_draw<W>(s.j_tab + (size_t)(s1 - 1) * n, j1, n, d, rs1, u, ps0);
  if (!s2) return false;
  const double* j2 = j_row(s, s2);
  const double rs2 = set_sum<W>(j2, d);
  const double ps1 = (!j2 || fabs(rs2) <= kEps) ? 0.0 : j2[s1 - 1] / rs2;
  if (full) {
    const double pf1 = fabs(rf) <= kEps ? 0.0 : s.i_prob[s2 - 1] / rf;
    p = pf0 * ps0 + pf1 * ps1;
  } else {
    p = (ps0 + ps1) / (double)nel;
  }
  if (s1 > s2) { int t = s1; s1 = s2; s2 = t; }
  return true;
}

template <int W>
__device__ __host__ inline double pgen_weighted_particles(const DevSys& s, const uint64_t* d, int I, int J) {
  const double *jI = j_row(s, I), *jJ = j_row(s, J);
  const double rI = set_sum<W>(jI, d), rJ = set_sum<W>(jJ, d);
  const double psI = (!jI || fabs(rI) <= kEps) ? 0.0 : jI[J - 1] / rI;
  const double psJ = (!jJ || fabs(rJ) <= kEps) ? 0.0 : jJ[I - 1] / rJ;
  if (s.particles != 3) return (psI + psJ) / (double)s.nel;
  const double* ip = s.i_ok ? s.i_prob : nullptr;
  const double rf = set_sum<W>(ip, d);
  const double pfI = (!ip || fabs(rf) <= kEps) ? 0.0 : ip[I - 1] / rf;
  const double pfJ = (!ip || fabs(rf) <= kEps) ? 0.0 : ip[J - 1] / rf;
  return pfI * psI + pfJ * psJ;
}

// renormalisation of a hole row over the unoccupied set: one minus the
// occupied share, or the direct sum when that is 1 or below 1e-5
// (do_direct_calculation, aliasSampling.F90:967-1000)
template <int W>
__device__ __host__ inline double hole_renorm(const double* row, const uint64_t* d, const uint64_t* un) {
  double r = 1.0 - set_sum<W>(row, d);
  if (r >= 1.0 || r < 1e-5) r = set_sum<W>(row, un);
  return r;
}

__device__ __host__ inline const double* a_row(const DevSys& s, int IJ) {
  return s.a_ok[IJ - 1] ? s.a_prob + (size_t)(IJ - 1) * s.nbasis : nullptr;
}
__device__ __host__ inline const double* b_row(const DevSys& s, int IJ, int A) {
  const size_t r = (size_t)(IJ - 1) * s.nbasis + (size_t)(A - 1);
  return s.b_ok[r] ? s.b_prob + r * s.nbasis : nullptr;
}

// the fully weighted holes: (t0, t1) in drawing order and sum(p_first * p_second)
template <int W>
__device__ __host__ inline bool draw_full_holes(const DevSys& s, const uint64_t* d, int I, int J, DrawSeq<W>& u,
                                                int& t0, int& t1, double& ph) {
  const int n = s.nbasis;
  uint64_t un[W];
  unocc_mask<W>(s, d, un);
  const int IJ = fuse_index(I, J);
  const double* ar = a_row(s, IJ);
  const double rf = hole_renorm<W>(ar, d, un);
  double pf0, ps0;
  t0 = constrained_draw<W>(s.a_tab + (size_t)(IJ - 1) * n, ar, n, un, rf, u, pf0);
  if (!t0) return false;
  const double* b1 = b_row(s, IJ, t0);
  const double rs1 = hole_renorm<W>(b1, d, un);
  if (fabs(rs1) <= kEps) return false;
  t1 = constrained_draw<W>(s.b_tab + ((size_t)(IJ - 1) * n + (size_t)(t0 - 1)) * n, b1, n, un, rs1, u, ps0);
  if (!t1) return false;
  const double pf1 = (!ar || fabs(rf) <= kEps) ? 0.0 : ar[t1 - 1] / rf;
  const double* b2 = b_row(s, IJ, t1);
  const double rs2 = hole_renorm<W>(b2, d, un);
  const double ps1 = (!b2 || fabs(rs2) <= kEps) ? 0.0 : b2[t0 - 1] / rs2;
  ph = pf0 * ps0 + pf1 * ps1;
  return true;
}

template <int W>
__device__ __host__ inline double pgen_full_holes(const DevSys& s, const uint64_t* d, int I, int J, int A, int B) {
  uint64_t un[W];
  unocc_mask<W>(s, d, un);
  const int IJ = fuse_index(I, J);
  const double* ar = a_row(s, IJ);
  const double rf = hole_renorm<W>(ar, d, un);
  const double pfA = (!ar || fabs(rf) <= kEps) ? 0.0 : ar[A - 1] / rf;
  const double pfB = (!ar || fabs(rf) <= kEps) ? 0.0 : ar[B - 1] / rf;
  const double *bA = b_row(s, IJ, A), *bB = b_row(s, IJ, B);
  const double rA = hole_renorm<W>(bA, d, un), rB = hole_renorm<W>(bB, d, un);
  const double psA = (!bA || fabs(rA) <= kEps) ? 0.0 : bA[B - 1] / rA;
  const double psB = (!bB || fabs(rB) <= kEps) ? 0.0 : bB[A - 1] / rB;
  return pfA * psA + pfB * psB;
}

// PCHB doubles with weighted particles or fully weighted holes
template <int W>
__device__ __host__ inline Excit gen_double_pchb_weighted(const DevSys& s, const uint64_t* d, DrawSeq<W>& u) {
  Excit ex{};
  const int nel = s.nel, n = s.nbasis;
  int s1 = 0, s2 = 0;
  double p;
  if (s.particles >= 2) {
    if (!draw_weighted_particles<W>(s, d, u, s1, s2, p)) return ex;
  } else if (s.particles == 1) {
    const int e1 = (int)(u.next() * (double)nel) + 1;
    s1 = select_orb<W>(d, ~0ull, e1);
    if (!s.j_ok[s1 - 1]) return ex;
    double p_J_I;
    const int j = alias_sample(s.j_tab + (size_t)(s1 - 1) * n, n, u.next(), p_J_I);
    s2 = j + 1;
    if (!test_orb<W>(d, s2)) return ex;
    const double p_I_J = s.j_ok[s2 - 1] ? s.j_prob[(size_t)(s2 - 1) * n + (s1 - 1)] : 0.0;
    p = (p_J_I + p_I_J) / (double)nel;
    if (s1 > s2) { int t = s1; s1 = s2; s2 = t; }
  } else {
    pick_biased_elecs<W>(s, d, u.next(), s1, s2, p);
    if (s1 == 0 || s2 == 0) return ex;
  }
  int t1, t2;
  double ph;
  if (s.holes == 1) {
    if (!draw_full_holes<W>(s, d, s1, s2, u, t1, t2, ph)) return ex;
    if (t1 > t2) { int t = t1; t1 = t2; t2 = t; }
  } else {
    const int ij = fuse_index(s1, s2);
    if (!s.ab_ok[ij - 1]) return ex;
    const int ab = alias_sample(s.ab_tab + (size_t)(ij - 1) * s.ijmax, s.ijmax, u.next(), ph) + 1;
    t1 = s.ab_tgt[2 * (ab - 1)];
    t2 = s.ab_tgt[2 * (ab - 1) + 1];
    if (t1 == t2 || test_orb<W>(d, t1) || test_orb<W>(d, t2)) return ex;
    if (fabs(ph) <= kEps) return ex;
  }
  ex.ic = 2;
  ex.e0 = s1; ex.e1 = s2; ex.e2 = t1; ex.e3 = t2;
  ex.parity = double_parity<W>(d, s1, s2, t1, t2);
  ex.pgen = p * ph;
  return ex;
}

// gen_rand_excit (GenRandSymExcitNUMod.F90:72-153) for C1 systems (one
// irrep, no Lz): CreateDoubExcit (:155-185: PickElecPair :1021, the forbidden
// count :385-412, PickAOrb :672-870, IsAOrbSymAllowed :873-949, PickBOrb
// :193-306, FindDoubleProb :187) and CreateSingleExcit (:1118-1286,
// CheckIfSingleExcits :1060).  Class counts are the occupied/unoccupied
// counts per spin (construct_class_counts :1288); the orbitals of one spin in
// symmetry-label order are the orbitals of that spin in index order.  A double
// is chosen when d0 < pDoubles.  The reference stops the run when a rejection
// loop exceeds its bound; here the excitation is null and fail is set.
template <int W>
__device__ __host__ inline Excit gen_uniform(const DevSys& s, const uint64_t* d, const double* dr, DrawSeq<W>& u,
                                             bool& fail) {
  Excit ex{};
  const uint64_t kA = 0xAAAAAAAAAAAAAAAAull, kB = 0x5555555555555555ull;
  const int nel = s.nel, nb = s.nbasis, nsp = nb / 2;
  int na = 0, nbe = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) { na += popc64(d[w] & kA); nbe += popc64(d[w] & kB); }
  const int ua = nsp - na, ub = nsp - nbe;  // unoccupied alpha / beta
  if (dr[0] < s.pdoubles) {
    const int pairs = nel * (nel - 1) / 2;
    int ind = (int)((double)pairs * u.next()) + 1;
    int X = pairs - ind;
    int K = 0;  // INT((SQRT(8X+1)-1)/2): the largest K with K(K+1)/2 <= X
    while ((K + 1) * (K + 2) / 2 <= X) ++K;
    int e1 = nel - 1 - K, e2 = nel - X + K * (K + 1) / 2;
    int o1 = select_orb<W>(d, ~0ull, e1), o2 = select_orb<W>(d, ~0ull, e2);
    bool b1 = is_beta(o1), b2 = is_beta(o2);
    int ispn = (b1 != b2) ? 2 : (!b1 ? 3 : 1);
    int forb = 0;
    if (ispn == 2) {
      if (ua == 0) forb += ub;
      if (ub == 0) forb += ua;
    } else if (ispn == 1) {
      if (ub == 1) forb = 1;
    } else {
      if (ua == 1) forb = 1;
    }
    const int nexa = ispn == 2 ? nb - nel : (ispn == 1 ? ub : ua);
    if (nexa == forb) return ex;
    int orbA = 0;
    for (int overall = 0;; ++overall) {
      if (nexa - forb <= 3) {
        int chosen = (int)((double)(nexa - forb) * u.next()) + 1;
        int z = 1;
        if (ispn == 2) {
          for (int o = 1; o <= nb && !orbA; ++o) {
            if (test_orb<W>(d, o)) continue;
            if (is_beta(o) ? ua != 0 : ub != 0) {
              if (z == chosen) orbA = o;
              z++;
            }
          }
        } else {
          for (int i = 1; i <= nsp && !orbA; ++i) {
            int o = ispn == 1 ? 2 * i - 1 : 2 * i;
            if (test_orb<W>(d, o)) continue;
            if (ispn == 1 ? ub != 1 : ua != 1) {
              if (z == chosen) orbA = o;
              z++;
            }
          }
        }
        if (!orbA) { fail = true; return ex; }
        break;  // the enumerating branch returns without the final check
      }
      for (int att = 0;; ++att) {
        int o = ispn == 2 ? (int)((double)nb * u.next()) + 1
                          : 2 * ((int)((double)nsp * u.next()) + 1) - (ispn == 1 ? 1 : 0);
        if (!test_orb<W>(d, o)) { orbA = o; break; }
        if (att > 250) { fail = true; return ex; }
      }
      bool allowed = ispn == 2 ? (is_beta(orbA) ? ua != 0 : ub != 0) : (ispn == 1 ? ub != 1 : ua != 1);
      if (allowed) break;
      if (overall > nb * 20) { fail = true; return ex; }
    }
    int nexb, other;
    bool bbeta;
    if (ispn == 2) {
      if (is_beta(orbA)) { nexb = ua; other = ub; bbeta = false; }
      else { nexb = ub; other = ua; bbeta = true; }
    } else if (ispn == 1) {
      nexb = ub; other = ub; bbeta = true;
    } else {
      nexb = ua; other = ua; bbeta = false;
    }
    if (ispn != 2) { nexb -= 1; other -= 1; }
    int orbB = 0;
    for (int att = 0;; ++att) {
      int o = 2 * ((int)((double)nsp * u.next()) + 1) - (bbeta ? 1 : 0);
      if (!test_orb<W>(d, o) && o != orbA) { orbB = o; break; }
      if (att > 1000) { fail = true; return ex; }
    }
    ex.ic = 2;
    ex.e0 = o1; ex.e1 = o2;
    ex.e2 = orbA < orbB ? orbA : orbB;
    ex.e3 = orbA < orbB ? orbB : orbA;
    ex.parity = double_parity<W>(d, ex.e0, ex.e1, ex.e2, ex.e3);
    ex.pgen = s.pdoubles * ((1.0 / (double)nexb) + (1.0 / (double)other)) / ((double)(pairs * (nexa - forb)));
    return ex;
  }
  int nowex = 0;
  if (na != 0 && ua == 0) nowex += na;
  if (nbe != 0 && ub == 0) nowex += nbe;
  if (nowex == nel) return ex;
  int src = 0, nexc = 0;
  for (int att = 0;; ++att) {
    int e = (int)((double)nel * u.next()) + 1;
    src = select_orb<W>(d, ~0ull, e);
    nexc = is_beta(src) ? ub : ua;
    if (nexc != 0) break;
    if (att > 250) { fail = true; return ex; }
  }
  int tgt = 0;
  for (int att = 0;; ++att) {
    int o = 2 * ((int)((double)nsp * u.next()) + 1) - (is_beta(src) ? 1 : 0);
    if (!test_orb<W>(d, o)) { tgt = o; break; }
    if (att > 250) { fail = true; return ex; }
  }
  ex.ic = 1;
  ex.e0 = src;
  ex.e1 = tgt;
  ex.parity = single_parity<W>(d, src, tgt);
  ex.pgen = (1.0 - s.pdoubles) / (double)(nexc * (nel - nowex));
  return ex;
}

// calc_pgen_symrandexcit2 (GenRandSymExcitNUMod.F90:1322-1470) for C1 systems
template <int W>
__device__ __host__ inline double pgen_uniform(const DevSys& s, const uint64_t* d, int ic, const int32_t* ex) {
  const uint64_t kA = 0xAAAAAAAAAAAAAAAAull, kB = 0x5555555555555555ull;
  const int nel = s.nel, nb = s.nbasis, nsp = nb / 2;
  int na = 0, nbe = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) { na += popc64(d[w] & kA); nbe += popc64(d[w] & kB); }
  const int ua = nsp - na, ub = nsp - nbe;
  if (ic == 1) {
    int nowex = 0;
    if (na != 0 && ua == 0) nowex += na;
    if (nbe != 0 && ub == 0) nowex += nbe;
    int nexa = is_beta(ex[0]) ? ub : ua;
    return (1.0 - s.pdoubles) / (double)(nexa * (nel - nowex));
  }
  if (ic != 2) return 0.0;
  bool b1 = is_beta(ex[0]), b2 = is_beta(ex[1]);
  int ispn = (b1 != b2) ? 2 : (!b1 ? 3 : 1);
  int nexa = ispn == 2 ? nb - nel : (ispn == 1 ? ub : ua);
  int forb = 0;
  if (ispn == 2) {
    if (ua == 0) forb += ub;
    if (ub == 0) forb += ua;
  } else if (ispn == 1) {
    if (ub == 1) forb = 1;
  } else {
    if (ua == 1) forb = 1;
  }
  int nexb, other;
  if (ispn == 2) {
    if (is_beta(ex[2])) { nexb = ua; other = ub; }
    else { nexb = ub; other = ua; }
  } else if (ispn == 1) {
    nexb = ub; other = ub;
  } else {
    nexb = ua; other = ua;
  }
  if (ispn != 2) { nexb -= 1; other -= 1; }
  const int pairs = nel * (nel - 1) / 2;
  return s.pdoubles * ((1.0 / (double)nexb) + (1.0 / (double)other)) / ((double)(pairs * (nexa - forb)));
}

// generate_excitation as bound by init_fcimc_fn_pointers
template <int W>
__device__ __host__ inline Excit gen_excit(const DevSys& s, const uint64_t* d, const double* dr,
                                           uint32_t iter = 0xFFFFFFFFu, int rep = 0, uint32_t k = 0,
                                           bool* fail = nullptr) {
  if (s.model == 0) return gen_hubbard<W>(s, d, dr);
  Excit ex;
  if (s.uniform) {
    DrawSeq<W> u(dr, s.seed, d, iter, rep, k);
    bool f = false;
    ex = gen_uniform<W>(s, d, dr, u, f);
    if (fail) *fail = f;
    return ex;
  }
  if (dr[0] < s.psingles) {  // gen_exc_sd (excitation_generators.F90:112-138)
    ex = gen_single_heatbath<W>(s, d, dr[1], dr[2]);
    ex.pgen = ex.pgen * s.psingles;
  } else if (s.particles >= 2 || s.holes) {
    DrawSeq<W> u(dr, s.seed, d, iter, rep, k);
    ex = gen_double_pchb_weighted<W>(s, d, u);
    ex.pgen = ex.pgen * s.pdoubles;
  } else {
    ex = gen_double_pchb<W>(s, d, dr);
    ex.pgen = ex.pgen * s.pdoubles;
  }
  return ex;
}

// child determinant: encode_child = FindExcitBitDet (DetBitOps.F90:584-623)
template <int W>
__device__ __host__ inline void apply_excit(const uint64_t* d, const Excit& ex, uint64_t* out) {
#pragma unroll
  for (int w = 0; w < W; ++w) out[w] = d[w];
  if (ex.ic == 1) {
    flip_orb<W>(out, ex.e0);
    flip_orb<W>(out, ex.e1);
  } else {
    flip_orb<W>(out, ex.e0);
    flip_orb<W>(out, ex.e1);
    flip_orb<W>(out, ex.e2);
    flip_orb<W>(out, ex.e3);
  }
}

// stochastic_round_r (lib/util_mod.fpp:207-224)
__device__ __host__ inline double stochastic_round(double x, double r) {
  double i = (double)(int64_t)x;
  double res = x - i;
  if (fabs(res) >= 1.0e-12)
    if (fabs(res) > r) i = i + (x >= 0.0 ? 1.0 : -1.0);
  return i;
}

// attempt_create_normal core (fcimc_pointed_fns.F90:178-491)
__device__ __host__ inline double attempt_create(const DevSys& s, double tau, double hel, double pgen,
                                                 double wsign, double r, double& raw) {
  double prob = pgen * s.avmc;
  double nspawn = -tau * hel * wsign / prob;
  raw = nspawn;
  if (s.t_real) {
    if (s.t_rsc && fabs(nspawn) < s.rsc) nspawn = s.rsc * stochastic_round(nspawn / s.rsc, r);
  } else {
    nspawn = stochastic_round(nspawn, r);
  }
  // FP32 spawned amplitudes: the stored value, rounded to nearest
  if (s.spawn_fp32) nspawn = (double)(float)nspawn;
  return nspawn;
}

// attempt_die_normal core for one replica (fcimc_pointed_fns.F90:573-705)
__device__ __host__ inline double attempt_die(bool real_coeff, double fac, double absc, double r) {
  if (real_coeff) return fac * absc;
  double rat = fac * absc;
  double nd = (double)(int64_t)rat;
  rat = rat - nd;
  if (fabs(rat) > r) nd = nd + (rat >= 0.0 ? 1.0 : -1.0);
  return nd;
}

}  // namespace ngpu
