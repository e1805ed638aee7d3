// ngpu_kernels.h -- kernels of one FCIQMC iteration on the GPU.
//
// Iteration chain (all on the engine stream, no host round trip):
//   k_prepass   per slot: initiator flag, proj-E sums, number of spawn
//               attempts, death (fused: spawning only needs the pre-death sign
//               bit and initiator bit, saved in pinfo)
//   k_scan_*    exclusive scan of attempt counts -> attempt offsets
//   k_spawn     per attempt: parent by binary search, excitation, H_ij,
//               stochastic rounding, wave-aggregated append of spawn records
//   k_sort_*    stable LSD radix sort of (key, spawn index), ilut_lt order
//   k_merge     per run of equal keys: CompressSpawnedList/FindResidualParticle
//   k_annihilate per unique spawned det: hash lookup, add/abort/remove
//   k_insert    new dets: slot from the free queue, bucketed open-addressing insert
//   k_stats     per slot: occupied-threshold rounding, TotParts, norm, HF pop
#pragma once
#include "ngpu_device.h"
#include "../../include/neci_gpu.h"
#include <cstddef>

namespace ngpu {

constexpr int kBlock = 256;
constexpr int kWave = 64;

// statistics fields of ngpu_iter_out_t viewed as a double array
enum {
  F_NBORN = 0, F_NDIED, F_NANNIHIL, F_NABORTED, F_NREMOVED, F_TOTPARTS, F_NORMSQ, F_ENUM,
  F_HFCYC, F_NOATDOUBS, F_SPAWNSING, F_INSTHF, F_NOINITDETS, F_NOINITWALK, F_NONINITDETS,
  F_NONINITWALK, F_NFIELDS
};
constexpr int S_NVALID = F_NFIELDS * kMaxRep;
constexpr int S_NINVALID = S_NVALID + 1;
constexpr int S_MAXSPAWN = S_NVALID + 2;
constexpr int S_BLOOMC = S_NVALID + 3;
constexpr int S_BLOOMS = S_NVALID + 6;
constexpr int S_NDOUBLE = S_NVALID + 9;
__host__ __device__ constexpr int soff(int f, int r) { return f * kMaxRep + r; }
static_assert(sizeof(double) * S_NDOUBLE == offsetof(ngpu_iter_out_t, tot_walkers),
              "statistics block of ngpu_iter_out_t must be S_NDOUBLE doubles");

struct Counters {
  unsigned long long nslots;
  unsigned long long free_head, free_tail;
  unsigned long long n_attempts, n_spawned, n_unique, n_new;
  unsigned long long n_live, n_tomb, n_occ;
  unsigned int err;
  unsigned int rebuild;  // set when tombstones must be reclaimed
};

// probe limit of the walker hash: the load guard (k_hash_check) keeps probe
// sequences short, the cap turns a saturated table into an error, not a stall
constexpr int64_t kMaxProbe = 4096;

struct Walk {
  uint64_t* keys;     // cap * W
  double* sgn;        // cap * S
  uint32_t* flags;    // cap
  double* diagH;      // H_ii - Hii
  double* offdiagH;   // H(det, ref)
  int64_t* hpos;      // hash entry (bucket*4 + e)
  int64_t cap;
  uint64_t* htab;     // L buckets of 8W words
  int64_t L;
  int64_t* freeq;     // free-slot queue (ring, cap entries)
};

struct Shift { double v[kMaxRep]; };

// ----------------------------------------------------------- wave utils
__device__ inline int lane_id() { return __lane_id(); }

__device__ inline uint64_t lanemask_lt() {
  int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// wave-aggregated append: returns the index for lanes with pred, -1 otherwise.
// Must be reached by every active lane of the wave.
__device__ inline long long wave_append(bool pred, unsigned long long* counter, int n = 1) {
  uint64_t m = __ballot(pred);
  if (m == 0) return -1;
  int leader = __ffsll((unsigned long long)m) - 1;
  unsigned long long base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (unsigned long long)(popc64(m) * n));
  base = __shfl(base, leader);
  if (!pred) return -1;
  return (long long)(base + (unsigned long long)(popc64(m & lanemask_lt()) * n));
}

// block-local statistics accumulators in LDS, flushed once per block
struct BlockStats {
  double* acc;  // LDS, S_NDOUBLE entries
  __device__ void zero() {
    for (int i = threadIdx.x; i < S_NDOUBLE; i += blockDim.x) acc[i] = 0.0;
    __syncthreads();
  }
  __device__ void add(int i, double v) {
    if (v != 0.0) atomicAdd(&acc[i], v);
  }
  __device__ void max(int i, double v) {  // v >= 0
    if (v > 0.0) atomicMax((unsigned long long*)&acc[i], (unsigned long long)__double_as_longlong(v));
  }
  __device__ void flush(ngpu_iter_out_t* out) {
    __syncthreads();
    double* g = (double*)out;
    for (int i = threadIdx.x; i < S_NDOUBLE; i += blockDim.x) {
      double v = acc[i];
      if (v == 0.0) continue;
      bool is_max = (i == S_MAXSPAWN) || (i >= S_BLOOMS && i < S_BLOOMS + 3);
      if (is_max)
        atomicMax((unsigned long long*)&g[i], (unsigned long long)__double_as_longlong(v));
      else if (i >= soff(F_INSTHF, 0) && i < soff(F_INSTHF, kMaxRep))
        g[i] = v;  // InstNoAtHF: a single det writes it
      else
        atomicAdd(&g[i], v);
    }
  }
};

// remove a det from the walker hash and the list (RemoveHashDet, load_balancer.fpp:631-644)
template <int W>
__device__ inline void remove_det(const Walk& wk, Counters* cnt, int64_t slot) {
  int64_t hp = wk.hpos[slot];
  int64_t b = hp >> 2;
  int e = (int)(hp & 3);
  int32_t* sl = (int32_t*)(wk.htab + (size_t)b * 8 * W + 4 * W);
  sl[e] = kSlotTomb;
  wk.flags[slot] |= kFlagRemoved;
  unsigned long long q = atomicAdd(&cnt->free_tail, 1ull);
  wk.freeq[q % (unsigned long long)wk.cap] = slot;
  atomicAdd(&cnt->n_tomb, 1ull);
  atomicSub(&cnt->n_live, 1ull);
}

// hash lookup: returns slot or -1
template <int W>
__device__ inline int64_t hash_lookup(const Walk& wk, const uint64_t* key, int64_t home, Counters* cnt) {
  int64_t b = home;
  int64_t maxp = wk.L < kMaxProbe ? wk.L : kMaxProbe;
  for (int64_t probe = 0; probe < maxp; ++probe) {
    const uint64_t* bk = wk.htab + (size_t)b * 8 * W;
    const int32_t* sl = (const int32_t*)(bk + 4 * W);
#pragma unroll
    for (int e = 0; e < kBucketEntries; ++e) {
      int32_t s = sl[e];
      if (s == kSlotEmpty) return -1;
      if (s >= 0) {
        bool eq = true;
#pragma unroll
        for (int w = 0; w < W; ++w) eq = eq && (bk[e * W + w] == key[w]);
        if (eq) return s;
      }
    }
    b = (b + 1 == wk.L) ? 0 : b + 1;
  }
  if (maxp < wk.L) atomicOr(&cnt->err, 1u << NGPU_ERR_HASH_FULL);
  return -1;
}

// bucketed open-addressing insert of an absent key; returns hash position or -1
template <int W>
__device__ inline int64_t hash_insert(const Walk& wk, const uint64_t* key, int64_t home, int32_t slot) {
  int64_t b = home;
  int64_t maxp = wk.L < kMaxProbe ? wk.L : kMaxProbe;
  for (int64_t probe = 0; probe < maxp; ++probe) {
    uint64_t* bk = wk.htab + (size_t)b * 8 * W;
    int32_t* sl = (int32_t*)(bk + 4 * W);
#pragma unroll
    for (int e = 0; e < kBucketEntries; ++e) {
      int32_t cur = sl[e];
      while (cur == kSlotEmpty || cur == kSlotTomb) {
        int32_t prev = atomicCAS(&sl[e], cur, slot);
        if (prev == cur) {
#pragma unroll
          for (int w = 0; w < W; ++w) bk[e * W + w] = key[w];
          return b * 4 + e;
        }
        cur = prev;
      }
    }
    b = (b + 1 == wk.L) ? 0 : b + 1;
  }
  return -1;
}

__device__ inline void chunk_of(int64_t n, int64_t& begin, int64_t& end) {
  int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  begin = (int64_t)blockIdx.x * chunk;
  end = begin + chunk;
  if (end > n) end = n;
  if (begin > n) begin = n;
}

// block exclusive scan of one int64 per thread (blockDim = kBlock)
__device__ inline int64_t block_exclusive_scan(int64_t v, int64_t* lds, int64_t& total) {
  int lane = lane_id(), wv = threadIdx.x / kWave;
  int64_t x = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    int64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) lds[wv] = x;
  __syncthreads();
  int64_t pre = 0, tot = 0;
  for (int q = 0; q < kBlock / kWave; ++q) {
    if (q < wv) pre += lds[q];
    tot += lds[q];
  }
  __syncthreads();
  total = tot;
  return pre + x - v;
}

// ------------------------------------------------------------- prepass
// CalcParentFlag_det + SumEContrib + decide_num_to_spawn + walker_death for
// every occupied slot (FciMCPar.F90:1294-1758, fcimc_helper.F90:518-802,
// 1036-1321, 2160-2407).  Blocks own contiguous slot ranges so that the block
// sums of attempt counts feed the scan directly.
template <int W>
__global__ void __launch_bounds__(kBlock) k_prepass(DevSys s, Walk wk, Counters* cnt,
                                                    ngpu_iter_out_t* out, int32_t* att_cnt,
                                                    uint8_t* pinfo, int64_t* blk_part, Shift sft,
                                                    double tau, uint32_t iter) {
  __shared__ double acc_lds[S_NDOUBLE];
  __shared__ int64_t red[kBlock / kWave];
  BlockStats bs{acc_lds};
  bs.zero();
  const int S = s.S;
  int64_t nslots = (int64_t)cnt->nslots;
  int64_t begin, end;
  chunk_of(nslots, begin, end);
  int64_t my_att = 0;
  bool death_err = false;
  for (int64_t slot = begin + threadIdx.x; slot < end; slot += blockDim.x) {
    uint32_t f = wk.flags[slot];
    if (f & kFlagRemoved) {
      for (int r = 0; r < S; ++r) att_cnt[slot * S + r] = 0;
      continue;
    }
    uint64_t key[W];
#pragma unroll
    for (int w = 0; w < W; ++w) key[w] = wk.keys[slot * W + w];
    double sg[kMaxRep];
    bool occ = false;
    for (int r = 0; r < S; ++r) {
      sg[r] = wk.sgn[slot * S + r];
      occ = occ || !(fabs(sg[r]) < 1.e-12);
    }
    bool core = (f & kFlagCore) != 0;
    if (!occ) {
      for (int r = 0; r < S; ++r) att_cnt[slot * S + r] = 0;
      continue;
    }
    int exlevel = excit_level<W>(s.ref, key);
    double hoff = wk.offdiagH[slot];
    if (s.t_init) {
      bool isref = det_eq<W>(key, s.ref);
      for (int r = 0; r < S; ++r) {
        bool was = (f >> r) & 1u;
        bool pop = fabs(sg[r]) > s.init_walk_no;
        bool now = was ? (pop || isref || core) : pop;
        if (now) {
          f |= 1u << r;
          bs.add(soff(F_NOINITDETS, r), 1.0);
          bs.add(soff(F_NOINITWALK, r), fabs(sg[r]));
        } else {
          f &= ~(1u << r);
          bs.add(soff(F_NONINITDETS, r), 1.0);
          bs.add(soff(F_NONINITWALK, r), fabs(sg[r]));
        }
      }
    }
    for (int r = 0; r < S; ++r) {
      if (exlevel == 0) bs.add(soff(F_HFCYC, r), sg[r]);
      if (exlevel == 2) bs.add(soff(F_NOATDOUBS, r), fabs(sg[r]));
      bs.add(soff(F_ENUM, r), hoff * sg[r]);
    }
    double kii = wk.diagH[slot];
    bool any_left = false;
    double cp[kMaxRep];
    for (int r = 0; r < S; ++r) {
      double ua, ub;
      draw2<W>(s.seed, key, iter, r, 0, 0, ua, ub);
      double pa = sg[r] * s.avmc;
      int64_t nsp = (int64_t)pa;
      if (nsp < 0) nsp = -nsp;
      if (fabs(fabs(pa) - (double)nsp) > 1.e-12) {
        double extra = fabs(pa) - (double)nsp;
        if (extra > ua) nsp++;
      }
      att_cnt[slot * S + r] = (int32_t)nsp;
      my_att += nsp;
      pinfo[slot * S + r] = (uint8_t)((sg[r] < 0.0 ? 1 : 0) | (((f >> r) & 1u) ? 2 : 0));
      // death
      double fac = tau * (kii - sft.v[r]);
      if (fac > 2.0) death_err = true;
      double idie = core ? 0.0 : attempt_die(s.t_real != 0, fac, fabs(sg[r]), ub);
      double a = fabs(sg[r]);
      bs.add(soff(F_NDIED, r), idie < a ? idie : a);
      double bb = idie - a;
      if (bb > 0.0) bs.add(soff(F_NBORN, r), bb);
      cp[r] = sg[r] - idie * (sg[r] >= 0.0 ? 1.0 : -1.0);
      any_left = any_left || (fabs(cp[r]) > 1.0e-12);
    }
    if (s.t_init && any_left) {
      for (int r = 0; r < S; ++r)
        if ((cp[r] > 0.0) != (sg[r] > 0.0)) {
          bs.add(soff(F_NABORTED, r), fabs(cp[r]));
          cp[r] = 0.0;
        }
      any_left = false;
      for (int r = 0; r < S; ++r) any_left = any_left || (fabs(cp[r]) > 1.0e-12);
    }
    if (any_left || core) {
      for (int r = 0; r < S; ++r) wk.sgn[slot * S + r] = cp[r];
      wk.flags[slot] = f;
    } else {
      for (int r = 0; r < S; ++r) wk.sgn[slot * S + r] = 0.0;
      wk.flags[slot] = f;
      remove_det<W>(wk, cnt, slot);
    }
  }
  if (death_err) atomicOr(&cnt->err, 1u << NGPU_ERR_DEATH_PROB);
  // block sum of attempts
  int64_t tot;
  (void)block_exclusive_scan(my_att, red, tot);
  if (threadIdx.x == 0) blk_part[blockIdx.x] = tot;
  bs.flush(out);
}

// exclusive scan of the per-block partials (single block)
__global__ void __launch_bounds__(kBlock) k_scan_partials(int64_t* part, int nparts, Counters* cnt) {
  __shared__ int64_t red[kBlock / kWave];
  int per = (nparts + kBlock - 1) / kBlock;
  int b0 = threadIdx.x * per;
  int64_t loc = 0;
  for (int i = 0; i < per; ++i)
    if (b0 + i < nparts) loc += part[b0 + i];
  int64_t tot;
  int64_t pre = block_exclusive_scan(loc, red, tot);
  for (int i = 0; i < per; ++i)
    if (b0 + i < nparts) {
      int64_t v = part[b0 + i];
      part[b0 + i] = pre;
      pre += v;
    }
  if (threadIdx.x == 0) cnt->n_attempts = (unsigned long long)tot;
}

// attempt offsets per (slot, replica) pair, same chunking as k_prepass
__global__ void __launch_bounds__(kBlock) k_scan_apply(const int32_t* att_cnt, int64_t* att_off,
                                                       const int64_t* part, Counters* cnt, int S) {
  __shared__ int64_t red[kBlock / kWave];
  int64_t nslots = (int64_t)cnt->nslots;
  int64_t begin, end;
  chunk_of(nslots, begin, end);
  int64_t pb = begin * S, pe = end * S;
  int64_t run = part[blockIdx.x];
  for (int64_t base = pb; base < pe; base += blockDim.x) {
    int64_t i = base + threadIdx.x;
    int64_t v = i < pe ? att_cnt[i] : 0;
    int64_t tot;
    int64_t pre = block_exclusive_scan(v, red, tot);
    if (i < pe) att_off[i] = run + pre;
    run += tot;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) att_off[nslots * S] = (int64_t)cnt->n_attempts;
}

// ------------------------------------------------------------- spawning
template <int W>
__global__ void __launch_bounds__(kBlock) k_spawn(DevSys s, Walk wk, Counters* cnt, ngpu_iter_out_t* out,
                                                  const int64_t* att_off, const uint8_t* pinfo,
                                                  uint64_t* sp_key, double* sp_val, uint32_t* sp_info,
                                                  int64_t scap, double tau, uint32_t iter) {
  __shared__ double acc_lds[S_NDOUBLE];
  BlockStats bs{acc_lds};
  bs.zero();
  const int S = s.S;
  const int64_t total = (int64_t)cnt->n_attempts;
  const int64_t npairs = (int64_t)cnt->nslots * S;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  bool overflow = false;
  // every lane runs the same number of loop trips so wave_append stays convergent
  int64_t trips = (total + stride - 1) / stride;
  for (int64_t t = 0; t < trips; ++t) {
    int64_t a = t * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool emit = false;
    uint64_t child[W];
    double nspawn = 0.0;
    uint32_t info = 0;
    if (a < total) {
      // parent pair: last p with att_off[p] <= a
      int64_t lo = 0, hi = npairs - 1;
      while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (att_off[mid] <= a) lo = mid; else hi = mid - 1;
      }
      int64_t p = lo;
      int64_t slot = p / S;
      int r = (int)(p - slot * S);
      uint32_t k = (uint32_t)(a - att_off[p]);
      uint64_t key[W];
#pragma unroll
      for (int w = 0; w < W; ++w) key[w] = wk.keys[slot * W + w];
      uint8_t pi = pinfo[p];
      double dr[6];
      draw2<W>(s.seed, key, iter, r, k, 1, dr[0], dr[1]);
      draw2<W>(s.seed, key, iter, r, k, 2, dr[2], dr[3]);
      draw2<W>(s.seed, key, iter, r, k, 3, dr[4], dr[5]);
      Excit ex = gen_excit<W>(s, key, dr);
      if (ex.ic == 0) {
        bs.add(S_NINVALID, 1.0);
      } else {
        bs.add(S_NVALID, 1.0);
        double hel = helement_ex<W>(s, key, ex.ic, ex.e0, ex.e1, ex.e2, ex.e3, ex.parity);
        double wsign = (pi & 1) ? -1.0 : 1.0;
        double raw;
        nspawn = attempt_create(s, tau, hel, ex.pgen, wsign, dr[5], raw);
        bs.max(S_MAXSPAWN, fabs(raw));
        if (fabs(nspawn) > kEps) {
          emit = true;
          double an = fabs(nspawn);
          bs.add(soff(F_NBORN, r), an);
          if (ex.ic == 1) bs.add(soff(F_SPAWNSING, r), an);
          if (an > s.init_walk_no) {
            bs.add(S_BLOOMC + ex.ic, 1.0);
            bs.max(S_BLOOMS + ex.ic, an);
          }
          apply_excit<W>(key, ex, child);
          info = (uint32_t)r | ((s.t_init && (pi & 2)) ? kInfoInit : 0u);
        }
      }
    }
    long long idx = wave_append(emit, &cnt->n_spawned);
    if (emit) {
      if (idx < scap) {
#pragma unroll
        for (int w = 0; w < W; ++w) sp_key[idx * W + w] = child[w];
        sp_val[idx] = nspawn;
        sp_info[idx] = info;
      } else {
        overflow = true;
      }
    }
  }
  if (overflow) atomicOr(&cnt->err, 1u << NGPU_ERR_SPAWN_FULL);
  bs.flush(out);
}

// ---------------------------------------------------------- radix sort
// Stable LSD radix sort, 8-bit digits, reduce-then-scan: per-block digit
// histograms, one digit-major exclusive scan, stable block-local scatter.
constexpr int kRadixBins = 256;

template <int W>
__device__ inline uint32_t radix_digit(const uint64_t* k, int word, int shift) {
  uint64_t x = k[word];
  if (W == 2 && word == 0) x ^= (1ull << 63);  // ilut_lt compares signed words
  return (uint32_t)(x >> shift) & 0xFFu;
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_radix_hist(const uint64_t* keys, const unsigned long long* n_ptr,
                                                       int64_t ncap, int word, int shift, uint32_t* hist) {
  __shared__ uint32_t h[kRadixBins];
  for (int i = threadIdx.x; i < kRadixBins; i += blockDim.x) h[i] = 0;
  __syncthreads();
  int64_t n = (int64_t)*n_ptr;
  if (n > ncap) n = ncap;
  int64_t begin, end;
  chunk_of(n, begin, end);
  for (int64_t i = begin + threadIdx.x; i < end; i += blockDim.x)
    atomicAdd(&h[radix_digit<W>(keys + i * W, word, shift)], 1u);
  __syncthreads();
  for (int d = threadIdx.x; d < kRadixBins; d += blockDim.x) hist[(size_t)d * gridDim.x + blockIdx.x] = h[d];
}

// digit-major exclusive scan of nb*256 counters (single block)
__global__ void __launch_bounds__(kBlock) k_radix_scan(uint32_t* hist, int nb) {
  __shared__ int64_t red[kBlock / kWave];
  int64_t total_n = (int64_t)nb * kRadixBins;
  int64_t per = (total_n + kBlock - 1) / kBlock;
  int64_t b0 = threadIdx.x * per;
  int64_t loc = 0;
  for (int64_t i = 0; i < per; ++i)
    if (b0 + i < total_n) loc += hist[b0 + i];
  int64_t tot;
  int64_t pre = block_exclusive_scan(loc, red, tot);
  for (int64_t i = 0; i < per; ++i)
    if (b0 + i < total_n) {
      uint32_t v = hist[b0 + i];
      hist[b0 + i] = (uint32_t)pre;
      pre += v;
    }
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_radix_scatter(const uint64_t* kin, const uint32_t* iin,
                                                          uint64_t* kout, uint32_t* iout,
                                                          const unsigned long long* n_ptr, int64_t ncap,
                                                          int word, int shift, const uint32_t* hist) {
  __shared__ uint32_t run[kRadixBins];
  __shared__ uint32_t wcnt[kBlock / kWave][kRadixBins];
  int wv = threadIdx.x / kWave;
  for (int d = threadIdx.x; d < kRadixBins; d += blockDim.x) {
    run[d] = hist[(size_t)d * gridDim.x + blockIdx.x];
    for (int q = 0; q < kBlock / kWave; ++q) wcnt[q][d] = 0;
  }
  __syncthreads();
  int64_t n = (int64_t)*n_ptr;
  if (n > ncap) n = ncap;
  int64_t begin, end;
  chunk_of(n, begin, end);
  uint64_t ltmask = lanemask_lt();
  for (int64_t base = begin; base < end; base += blockDim.x) {
    int64_t i = base + threadIdx.x;
    bool valid = i < end;
    uint64_t k[W];
    uint32_t idx = 0, d = 0;
    if (valid) {
#pragma unroll
      for (int w = 0; w < W; ++w) k[w] = kin[i * W + w];
      idx = iin ? iin[i] : (uint32_t)i;
      d = radix_digit<W>(k, word, shift);
    }
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      uint64_t m = __ballot(valid && ((d >> bit) & 1u));
      peers &= ((d >> bit) & 1u) ? m : ~m;
    }
    int rank = popc64(peers & ltmask);
    if (valid && rank == 0) wcnt[wv][d] = (uint32_t)popc64(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + (uint32_t)rank;
      for (int q = 0; q < wv; ++q) pos += wcnt[q][d];
#pragma unroll
      for (int w = 0; w < W; ++w) kout[(size_t)pos * W + w] = k[w];
      iout[pos] = idx;
    }
    __syncthreads();
    for (int dd = threadIdx.x; dd < kRadixBins; dd += blockDim.x) {
      uint32_t s = 0;
      for (int q = 0; q < kBlock / kWave; ++q) { s += wcnt[q][dd]; wcnt[q][dd] = 0; }
      run[dd] += s;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------- merging
// CompressSpawnedList (Annihilation.F90:249-515) + FindResidualParticle
// (:551-634): the head of each run of equal keys sums the run in sorted
// (= emission, the sort is stable) order.
template <int W>
__global__ void __launch_bounds__(kBlock) k_merge(DevSys s, Counters* cnt, ngpu_iter_out_t* out,
                                                  const uint64_t* skey, const uint32_t* sidx,
                                                  const double* sp_val, const uint32_t* sp_info,
                                                  int64_t scap, uint64_t* u_key, double* u_sgn,
                                                  uint32_t* u_flag) {
  __shared__ double acc_lds[S_NDOUBLE];
  BlockStats bs{acc_lds};
  bs.zero();
  const int S = s.S;
  int64_t n = (int64_t)cnt->n_spawned;
  if (n > scap) n = scap;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t trips = (n + stride - 1) / stride;
  for (int64_t t = 0; t < trips; ++t) {
    int64_t i = t * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool keep = false;
    double sum[kMaxRep];
    uint32_t flag = 0;
    uint64_t key[W];
    if (i < n) {
#pragma unroll
      for (int w = 0; w < W; ++w) key[w] = skey[i * W + w];
      bool head = true;
      if (i > 0) head = !det_eq<W>(key, skey + (i - 1) * W);
      if (head) {
        int64_t j = i + 1;
        while (j < n && det_eq<W>(key, skey + j * W)) ++j;
        for (int r = 0; r < S; ++r) sum[r] = 0.0;
        if (j == i + 1) {
          uint32_t inf = sp_info[sidx[i]];
          int r = (int)(inf & 0xFFu);
          sum[r] = sp_val[sidx[i]];
          if (inf & kInfoInit) flag |= 1u << r;
          keep = !(fabs(sum[r]) < 1.e-12);
        } else {
          for (int64_t q = i; q < j; ++q) {
            uint32_t inf = sp_info[sidx[q]];
            int r = (int)(inf & 0xFFu);
            double nv = sp_val[sidx[q]];
            double cv = sum[r];
            if (s.t_init) {
              if ((fabs(cv) > 1.e-12 && fabs(nv) > 1.e-12) || (inf & kInfoInit)) flag |= 1u << r;
            }
            if (cv * nv < 0.0) {
              double x = fabs(cv), y = fabs(nv);
              bs.add(soff(F_NANNIHIL, r), 2.0 * (x < y ? x : y));
            }
            sum[r] = cv + nv;
          }
          double tot = 0.0;
          for (int r = 0; r < S; ++r) tot += fabs(sum[r]);
          keep = tot > 1.e-12;
        }
      }
    }
    long long u = wave_append(keep, &cnt->n_unique);
    if (keep) {
#pragma unroll
      for (int w = 0; w < W; ++w) u_key[u * W + w] = key[w];
      for (int r = 0; r < S; ++r) u_sgn[u * S + r] = sum[r];
      u_flag[u] = flag;
    }
  }
  bs.flush(out);
}

// ---------------------------------------------------------- annihilation
// AnnihilateSpawnedParts (Annihilation.F90:965-1352): found dets are updated
// in place (each unique key is owned by one lane); absent dets go to the new list.
template <int W>
__global__ void __launch_bounds__(kBlock) k_annihilate(DevSys s, Walk wk, Counters* cnt, ngpu_iter_out_t* out,
                                                       const uint64_t* u_key, const double* u_sgn,
                                                       const uint32_t* u_flag, uint64_t* n_key,
                                                       double* n_sgn, uint32_t* n_flag, uint32_t iter) {
  __shared__ double acc_lds[S_NDOUBLE];
  BlockStats bs{acc_lds};
  bs.zero();
  const int S = s.S;
  int64_t nu = (int64_t)cnt->n_unique;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t trips = (nu + stride - 1) / stride;
  for (int64_t t = 0; t < trips; ++t) {
    int64_t i = t * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool isnew = false;
    uint64_t key[W];
    double ss[kMaxRep];
    uint32_t sflag = 0;
    if (i < nu) {
#pragma unroll
      for (int w = 0; w < W; ++w) key[w] = u_key[i * W + w];
      for (int r = 0; r < S; ++r) ss[r] = u_sgn[i * S + r];
      sflag = u_flag[i];
      int64_t home = find_walker_hash<W>(key, s.rh, s.L) - 1;
      int64_t pos = hash_lookup<W>(wk, key, home, cnt);
      if (pos >= 0) {
        double cs[kMaxRep];
        double tot = 0.0;
        for (int r = 0; r < S; ++r) { cs[r] = wk.sgn[pos * S + r]; tot += fabs(cs[r]); }
        bool determ = (wk.flags[pos] & kFlagCore) != 0;
        if (tot >= 1.e-12 || determ) {
          for (int r = 0; r < S; ++r) {
            double prod = cs[r] * ss[r];
            if (fabs(cs[r]) < 1.0e-12 && s.t_init && !((sflag >> r) & 1u) && !determ) {
              bs.add(soff(F_NABORTED, r), fabs(ss[r]));
              ss[r] = 0.0;
            }
            if (prod < 0.0) {
              double x = fabs(cs[r]), y = fabs(ss[r]);
              bs.add(soff(F_NANNIHIL, r), 2.0 * (x < y ? x : y));
            }
          }
          bool occ = false;
          for (int r = 0; r < S; ++r) {
            cs[r] = ss[r] + cs[r];
            wk.sgn[pos * S + r] = cs[r];
            occ = occ || !(fabs(cs[r]) < 1.e-12);
          }
          if (!determ && !occ) remove_det<W>(wk, cnt, pos);
        }
      } else {
        if (s.t_init)
          for (int r = 0; r < S; ++r)
            if (!((sflag >> r) & 1u)) {
              bs.add(soff(F_NABORTED, r), fabs(ss[r]));
              ss[r] = 0.0;
            }
        bool occ = false;
        for (int r = 0; r < S; ++r) occ = occ || !(fabs(ss[r]) < 1.e-12);
        if (occ) {
          // stochRoundSpawn (Annihilation.F90:1354-1418)
          double thr = s.occ_thresh;
          for (int r = 0; r < S; ++r) {
            if (fabs(ss[r]) > 1.e-12 && fabs(ss[r]) < thr) {
              double pRemove = 1.0 - fabs(ss[r]) / thr;
              double u, dummy;
              draw2<W>(s.seed, key, iter, r, 0, 4, u, dummy);
              if (pRemove > u) {
                bs.add(soff(F_NREMOVED, r), fabs(ss[r]));
                ss[r] = 0.0;
              } else {
                bs.add(soff(F_NBORN, r), thr - fabs(ss[r]));
                ss[r] = ss[r] >= 0.0 ? thr : -thr;
              }
            }
          }
          occ = false;
          for (int r = 0; r < S; ++r) occ = occ || !(fabs(ss[r]) < 1.e-12);
          isnew = occ;
        }
      }
    }
    long long q = wave_append(isnew, &cnt->n_new);
    if (isnew) {
#pragma unroll
      for (int w = 0; w < W; ++w) n_key[q * W + w] = key[w];
      for (int r = 0; r < S; ++r) n_sgn[q * S + r] = ss[r];
      n_flag[q] = sflag & 0xFFFFu;
    }
  }
  bs.flush(out);
}

// AddNewHashDet (load_balancer.fpp:514-629) for the new list
template <int W>
__global__ void __launch_bounds__(kBlock) k_insert(DevSys s, Walk wk, Counters* cnt, const uint64_t* n_key,
                                                   const double* n_sgn, const uint32_t* n_flag,
                                                   const unsigned long long* n_ptr) {
  const int S = s.S;
  int64_t nn = (int64_t)*n_ptr;
  unsigned long long tail = cnt->free_tail;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t key[W];
#pragma unroll
    for (int w = 0; w < W; ++w) key[w] = n_key[i * W + w];
    unsigned long long q = atomicAdd(&cnt->free_head, 1ull);
    int64_t slot;
    if (q < tail) {
      slot = wk.freeq[q % (unsigned long long)wk.cap];
    } else {
      unsigned long long ns = atomicAdd(&cnt->nslots, 1ull);
      if ((int64_t)ns >= wk.cap) {
        atomicOr(&cnt->err, 1u << NGPU_ERR_WALKERS_FULL);
        atomicSub(&cnt->nslots, 1ull);
        continue;
      }
      slot = (int64_t)ns;
    }
    int64_t home = find_walker_hash<W>(key, s.rh, s.L) - 1;
    int64_t hp = hash_insert<W>(wk, key, home, (int32_t)slot);
    if (hp < 0) {
      atomicOr(&cnt->err, 1u << NGPU_ERR_HASH_FULL);
      wk.flags[slot] = kFlagRemoved;
      continue;
    }
#pragma unroll
    for (int w = 0; w < W; ++w) wk.keys[slot * W + w] = key[w];
    for (int r = 0; r < S; ++r) wk.sgn[slot * S + r] = n_sgn[i * S + r];
    wk.flags[slot] = n_flag[i] & ~kFlagRemoved;
    wk.diagH[slot] = diag_helement<W>(s, key) - s.hii;
    wk.offdiagH[slot] = offdiag_helement<W>(s, key);
    wk.hpos[slot] = hp;
    if (s.t_real == 0 && s.t_init == 0) {}
    atomicAdd(&cnt->n_live, 1ull);
  }
}

// free-queue bookkeeping after inserts: pops beyond the tail appended instead
__global__ void k_post_insert(Counters* cnt) {
  if (cnt->free_head > cnt->free_tail) cnt->free_head = cnt->free_tail;
}

// CalcHashTableStats (load_balancer.fpp:646-805)
template <int W>
__global__ void __launch_bounds__(kBlock) k_stats(DevSys s, Walk wk, Counters* cnt, ngpu_iter_out_t* out,
                                                  uint32_t iter) {
  __shared__ double acc_lds[S_NDOUBLE];
  BlockStats bs{acc_lds};
  bs.zero();
  const int S = s.S;
  int64_t nslots = (int64_t)cnt->nslots;
  unsigned long long my_occ = 0;
  for (int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; slot < nslots;
       slot += (int64_t)gridDim.x * blockDim.x) {
    uint32_t f = wk.flags[slot];
    if (f & kFlagRemoved) continue;
    bool determ = (f & kFlagCore) != 0;
    double sg[kMaxRep];
    bool occ = false;
    for (int r = 0; r < S; ++r) { sg[r] = wk.sgn[slot * S + r]; occ = occ || !(fabs(sg[r]) < 1.e-12); }
    if (!occ && !determ) continue;
    uint64_t key[W];
#pragma unroll
    for (int w = 0; w < W; ++w) key[w] = wk.keys[slot * W + w];
    if (s.t_real) {
      double thr = s.occ_thresh;
      bool changed = false;
      for (int r = 0; r < S; ++r) {
        if (!determ && fabs(sg[r]) > 1.e-12 && fabs(sg[r]) < thr) {
          double pRemove = (thr - fabs(sg[r])) / thr;
          double u, dummy;
          draw2<W>(s.seed, key, iter, r, 0, 5, u, dummy);
          if (pRemove > u) {
            bs.add(soff(F_NREMOVED, r), fabs(sg[r]));
            sg[r] = 0.0;
          } else {
            bs.add(soff(F_NBORN, r), thr - fabs(sg[r]));
            sg[r] = sg[r] >= 0.0 ? thr : -thr;
          }
          changed = true;
        }
      }
      if (changed) {
        occ = false;
        for (int r = 0; r < S; ++r) { wk.sgn[slot * S + r] = sg[r]; occ = occ || !(fabs(sg[r]) < 1.e-12); }
        if (!occ && !determ) {
          remove_det<W>(wk, cnt, slot);
          continue;
        }
      }
    }
    my_occ++;
    for (int r = 0; r < S; ++r) {
      bs.add(soff(F_TOTPARTS, r), fabs(sg[r]));
      bs.add(soff(F_NORMSQ, r), sg[r] * sg[r]);
    }
    if (det_eq<W>(key, s.ref))
      for (int r = 0; r < S; ++r) bs.acc[soff(F_INSTHF, r)] = sg[r];
  }
  if (my_occ) atomicAdd(&cnt->n_occ, my_occ);
  bs.flush(out);
}

// ---------------------------------------------------------- hash upkeep
// Tombstones left by removals are reclaimed by rebuilding the index in place
// once live + tombstone entries pass 60% of the entries and tombstones alone
// pass 10%; more than 80% live
// entries is a capacity error (MEMORYFACHASH too small), reported instead of
// letting probe sequences grow without bound.
__global__ void k_hash_check(Counters* cnt, int64_t entries) {
  unsigned long long used = cnt->n_live + cnt->n_tomb;
  cnt->rebuild = (used * 10ull > (unsigned long long)entries * 6ull &&
                  cnt->n_tomb * 10ull > (unsigned long long)entries) ? 1u : 0u;
  if (cnt->n_live * 10ull > (unsigned long long)entries * 8ull) cnt->err |= 1u << NGPU_ERR_HASH_FULL;
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_hash_clear(Walk wk, const Counters* cnt) {
  if (!cnt->rebuild) return;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < wk.L; b += (int64_t)gridDim.x * blockDim.x) {
    int32_t* sl = (int32_t*)(wk.htab + (size_t)b * 8 * W + 4 * W);
#pragma unroll
    for (int e = 0; e < kBucketEntries; ++e) sl[e] = kSlotEmpty;
  }
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_hash_refill(DevSys s, Walk wk, Counters* cnt) {
  if (!cnt->rebuild) return;
  int64_t nslots = (int64_t)cnt->nslots;
  for (int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; slot < nslots;
       slot += (int64_t)gridDim.x * blockDim.x) {
    if (wk.flags[slot] & kFlagRemoved) continue;
    uint64_t key[W];
#pragma unroll
    for (int w = 0; w < W; ++w) key[w] = wk.keys[slot * W + w];
    int64_t home = find_walker_hash<W>(key, s.rh, s.L) - 1;
    int64_t hp = hash_insert<W>(wk, key, home, (int32_t)slot);
    if (hp < 0) atomicOr(&cnt->err, 1u << NGPU_ERR_HASH_FULL);
    else wk.hpos[slot] = hp;
  }
}

__global__ void k_hash_done(Counters* cnt) {
  if (cnt->rebuild) {
    cnt->n_tomb = 0;
    cnt->rebuild = 0;
  }
}

}  // namespace ngpu
