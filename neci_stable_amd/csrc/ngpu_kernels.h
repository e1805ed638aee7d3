// ngpu_kernels.h -- kernels of one FCIQMC iteration on the GPU.
//
// Iteration chain (all on the engine stream, no host round trip):
//   k_prepass    per slot: initiator flag, proj-E sums, number of spawn
//                attempts, death (fused: spawning only needs the pre-death
//                sign bit and initiator bit, saved in pinfo); marks the slots
//                it frees; clears the walker hash when a rebuild is due
//   k_scan_apply exclusive scan of attempt counts -> attempt offsets; the
//                freed slots into the free queue; the hash refill of a rebuild
//   k_spawn      per attempt: parent by binary search, excitation, H_ij,
//                stochastic rounding, LDS-staged append of spawn records;
//                heat-bath singles are queued for k_spawn_singles
//   k_radix_*    stable LSD radix sort of (colex rank or ilut_lt key, spawn index)
//   k_merge_annihilate per run of equal keys: CompressSpawnedList /
//                FindResidualParticle, then hash lookup, add/abort/remove
//   k_insert     new dets: slot from the free queue, bucketed open-addressing insert
//   k_stats      per slot: occupied-threshold rounding, TotParts, norm, HF pop
//   (k_end_iter in the engine: slot and hash bookkeeping, statistics sums)
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
// block statistics beyond the output block: occupied determinants
constexpr int X_NOCC = S_NDOUBLE;
// Blocks flush their statistics into one of kStatCopies partial rows (chosen
// by block index), so that no address takes more than 1/kStatCopies of the
// device-scope atomics; the end-of-iteration kernel sums the rows.
constexpr int kStatStride = 288;  // doubles per row (2304 B, 9 lines of 256 B)
constexpr int kStatCopies = 32;
static_assert(X_NOCC < kStatStride, "statistics row too short");
__host__ __device__ constexpr bool stat_is_max(int i) {
  return i == S_MAXSPAWN || (i >= S_BLOOMS && i < S_BLOOMS + 3);
}

// Device counters.  Appends reserve ranges of the global lists once per LDS
// flush, never per lane or per wave: device-scope atomics on one address are
// serialised, so their count, not their payload, bounds the kernels that issue
// them.  The fields bumped in one kernel sit in separate 256-byte lines.
struct alignas(256) Counters {
  unsigned long long nslots;
  unsigned long long free_head;
  unsigned long long free_tail0;  // free_tail at the start of the iteration
  unsigned long long n_live, n_tomb, n_occ;
  unsigned long long n_attempts;
  unsigned long long n_fail;      // inserts refused (capacity errors)
  unsigned int ins_refused;       // k_insert placed no new determinant (MaxWalkers would be passed)
  unsigned long long free_tail_ins;  // free_tail as k_insert saw it
  unsigned long long rm_pre;         // slots the iteration's prepass freed
  unsigned int err;
  // the walker hash is rebuilt at the start of the next iteration (set by the
  // end-of-iteration accounting; the prepass clears the table, the attempt
  // scan refills it)
  unsigned int rebuild;
  unsigned int rebuilt;  // set by the attempt scan that refilled the table
  // sticky error: the error mask of the first iteration that reported one,
  // and that iteration's number.  While halt is set every kernel of the
  // iteration chain returns at once, so the walker list stays as the failing
  // iteration left it (the reference stops there, stop_all)
  unsigned int halt;
  unsigned int err_iter;
  unsigned long long n_rebuilds;  // walker-hash rebuilds since creation
  alignas(256) unsigned long long free_tail;
  alignas(256) unsigned long long n_spawned;
  alignas(256) unsigned long long n_singles;  // queued single-excitation attempts
  alignas(256) unsigned long long n_unique;
  alignas(256) unsigned long long n_new;
  // per-block counts of tombstone entries taken by inserts, spread over
  // kSpread lines (one add per block)
  struct alignas(256) Spread { unsigned long long v; };
  static constexpr int kSpread = 32;
  Spread reuse_part[kSpread];
  Spread uniq_part[kSpread];  // unique spawned determinants, per block
};

// probe limit of the walker hash: the load guard (k_hash_check) keeps probe
// sequences short, the cap turns a saturated table into an error, not a stall
constexpr int64_t kMaxProbe = 4096;

// Walker table.  The per-slot state that the random-access kernels update
// together -- the S signs, the flag word and the hash entry -- is one row of
// S+1 doubles (signs, then flags in the low and the hash entry in the high
// 32 bits of the last word), so that a lookup-and-update touches one line.
struct Walk {
  uint64_t* keys;     // cap * W
  double* st;         // cap * srow: S signs | flags, hash entry (bucket*4 + e)
  int64_t srow;       // S + 1
  double* diagH;      // H_ii - Hii
  double* offdiagH;   // H(det, ref)
  int64_t cap;
  uint64_t* htab;     // L buckets of 8W words
  int64_t L;
  int64_t* freeq;     // free-slot queue (ring, cap entries)
  __device__ double& sgn(int64_t slot, int r) const { return st[slot * srow + r]; }
  __device__ uint32_t& flags(int64_t slot) const { return *(uint32_t*)(st + slot * srow + srow - 1); }
  __device__ int32_t& hpos(int64_t slot) const { return *((int32_t*)(st + slot * srow + srow - 1) + 1); }
};

struct Shift { double v[kMaxRep]; };
// per-iteration inputs (time step, shifts, iteration number), read by the
// kernels from device memory so that a captured iteration replays unchanged
struct IterParams {
  double tau;
  Shift sft;
  uint32_t iter;
  uint32_t pad;
};

// The blocks a kernel body runs on: this block's index and the number of
// blocks of its launch share.  A plain launch gives the hardware grid; a
// batch launch (one launch covering several engines) gives each engine's
// share of the blocks, so a body sees the same index space either way.
struct Grid {
  int64_t b, nb;
};
__device__ inline Grid hw_grid() { return Grid{(int64_t)blockIdx.x, (int64_t)gridDim.x}; }

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

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  return v;
}
__device__ inline unsigned long long wave_count(bool p) { return (unsigned long long)popc64(__ballot(p)); }

// Block-local statistics in LDS.  Every add/max is a wave-wide call (all 64
// lanes, inactive lanes passing 0): the wave reduces with shuffles and one lane
// updates LDS, so no LDS address is contended by more than one lane per wave.
struct BlockStats {
  double* acc;  // LDS, kStatStride entries
  __device__ void zero() {
    for (int i = threadIdx.x; i < kStatStride; i += blockDim.x) acc[i] = 0.0;
    __syncthreads();
  }
  __device__ void add(int i, double v) {
    double t = wave_sum(v);
    if (lane_id() == 0 && t != 0.0) atomicAdd(&acc[i], t);
  }
  __device__ void max(int i, double v) {  // v >= 0
    double t = wave_max(v);
    if (lane_id() == 0 && t > 0.0)
      atomicMax((unsigned long long*)&acc[i], (unsigned long long)__double_as_longlong(t));
  }
  // into partial row blockIdx % kStatCopies (InstNoAtHF has a single writer,
  // so summing the rows reproduces it)
  __device__ void flush(double* part, int64_t blk) {
    __syncthreads();
    double* g = part + (size_t)(blk % kStatCopies) * kStatStride;
    for (int i = threadIdx.x; i < kStatStride; i += blockDim.x) {
      double v = acc[i];
      if (v == 0.0) continue;
      if (stat_is_max(i))
        atomicMax((unsigned long long*)&g[i], (unsigned long long)__double_as_longlong(v));
      else
        atomicAdd(&g[i], v);
    }
  }
};

// ------------------------------------------------------ LDS-staged appends
// A block stages the records it appends to a global list in LDS and reserves
// space in the list once per flush (one device-scope atomic for up to CAP
// records).  push() is wave-wide; flush_begin()/flush_end() are block-wide and
// must be reached by every thread of the block in every trip.
struct Stage {
  uint32_t n;                 // staged records
  unsigned long long base;    // reserved position of the flush in progress
};

// index in the LDS stage for lanes with pred (-1 otherwise)
__device__ inline int stage_push(bool pred, Stage* st) {
  uint64_t m = __ballot(pred);
  if (m == 0) return -1;
  int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t b = 0;
  if (lane_id() == leader) b = atomicAdd(&st->n, (uint32_t)popc64(m));
  b = __shfl(b, leader);
  return pred ? (int)(b + (uint32_t)popc64(m & lanemask_lt())) : -1;
}

// true when the block must flush now (force, or fewer than blockDim free
// entries left); then cnt = staged records and base = their reserved position
__device__ inline bool stage_flush_begin(Stage* st, unsigned long long* gctr, int cap, bool force,
                                         uint32_t& cnt, unsigned long long& base) {
  __syncthreads();
  cnt = st->n;
  __syncthreads();  // every thread has read n before the next trip pushes again
  if (cnt == 0 || (!force && cnt + blockDim.x <= (uint32_t)cap)) return false;
  if (threadIdx.x == 0) st->base = atomicAdd(gctr, (unsigned long long)cnt);
  __syncthreads();
  base = st->base;
  return true;
}
__device__ inline void stage_flush_end(Stage* st) {
  __syncthreads();
  if (threadIdx.x == 0) st->n = 0;
  __syncthreads();
}
__device__ inline void stage_init(Stage* st) {
  if (threadIdx.x == 0) st->n = 0;
  __syncthreads();
}

// pinfo bit of a slot the prepass freed (first replica's byte)
constexpr uint8_t kPinfoFreed = 8;

// Freed slots staged for the free queue (a ring of cap entries).
constexpr int kFreeStage = 512;
struct FreeStage {
  Stage st;
  int64_t slot[kFreeStage];
};

// RemoveHashDet (load_balancer.fpp:631-644) for the lanes with rm set: the
// hash entry becomes a tombstone, the slot is staged for the free queue.
// Wave-wide call; the live and tombstone counts follow from the free-queue
// tail.  (Releasing the entry as empty where the
// bucket already holds an empty entry would also be valid -- such a bucket
// ends every probe -- but it makes each removal read its bucket and, measured,
// cost more in the removing kernels than the rebuilds it saves.)  The prepass
// removes without this helper: it marks the slot, the attempt scan queues it.
template <int W>
__device__ inline void remove_dets(const Walk& wk, FreeStage* fs, bool rm, int64_t slot) {
  int q = stage_push(rm, &fs->st);
  if (rm) {
    int64_t hp = wk.hpos(slot);
    int32_t* sl = (int32_t*)(wk.htab + (size_t)(hp >> 2) * 8 * W + 4 * W);
    sl[hp & 3] = kSlotTomb;
    wk.flags(slot) |= kFlagRemoved;
    fs->slot[q] = slot;
  }
}
// block-wide: drain the staged slots into the free queue when due
__device__ inline void free_flush(const Walk& wk, Counters* cnt, FreeStage* fs, bool force) {
  uint32_t n;
  unsigned long long base;
  if (!stage_flush_begin(&fs->st, &cnt->free_tail, kFreeStage, force, n, base)) return;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
    wk.freeq[(base + i) % (unsigned long long)wk.cap] = fs->slot[i];
  stage_flush_end(&fs->st);
}


// hash lookup: returns slot or -1.  Entries of a bucket fill in order and
// become empty only when the table is rebuilt, so the first empty entry ends
// the probe sequence.
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

// bucketed open-addressing insert of an absent key; returns hash position or
// -1; reused is set when the entry claimed was a tombstone
template <int W>
__device__ inline int64_t hash_insert(const Walk& wk, const uint64_t* key, int64_t home, int32_t slot,
                                      bool& reused) {
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
          reused = cur == kSlotTomb;
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

// rebuild of the walker hash in two grid-wide phases (a kernel boundary
// between them): every entry emptied, then every live slot reinserted
template <int W>
__device__ inline void hash_clear_part(const Grid& g, const Walk& wk) {
  for (int64_t b = g.b * blockDim.x + threadIdx.x; b < wk.L; b += g.nb * blockDim.x) {
    int32_t* sl = (int32_t*)(wk.htab + (size_t)b * 8 * W + 4 * W);
#pragma unroll
    for (int e = 0; e < kBucketEntries; ++e) sl[e] = kSlotEmpty;
  }
}
template <int W>
__device__ inline void hash_refill_part(const Grid& g, const DevSys& s, const Walk& wk, Counters* cnt,
                                        int64_t nslots) {
  for (int64_t slot = g.b * blockDim.x + threadIdx.x; slot < nslots; slot += g.nb * blockDim.x) {
    if (wk.flags(slot) & kFlagRemoved) continue;
    uint64_t key[W];
#pragma unroll
    for (int w = 0; w < W; ++w) key[w] = wk.keys[slot * W + w];
    int64_t home = find_walker_hash<W>(key, s.rh, s.L) - 1;
    bool reused = false;
    int64_t hp = hash_insert<W>(wk, key, home, (int32_t)slot, reused);
    if (hp < 0) atomicOr(&cnt->err, 1u << NGPU_ERR_HASH_FULL);
    else wk.hpos(slot) = hp;
  }
}

__device__ inline void chunk_of(const Grid& g, int64_t n, int64_t& begin, int64_t& end) {
  int64_t chunk = (n + g.nb - 1) / g.nb;
  begin = g.b * chunk;
  end = begin + chunk;
  if (end > n) end = n;
  if (begin > n) begin = n;
}

// Per-lane statistic accumulator of one field: a register when the kernel is
// built for one replica (the field index is then the same in every trip), the
// block's LDS row otherwise.  add() and done() are wave-wide calls.
template <int R>
struct LaneSum {
  double v = 0.0;
  __device__ void add(BlockStats& bs, int i, double x) {
    if constexpr (R == 1) v += x; else bs.add(i, x);
  }
  __device__ void done(BlockStats& bs, int i) {
    if constexpr (R == 1) bs.add(i, v);
  }
};
template <int R>
struct LaneMax {
  double v = 0.0;
  __device__ void add(BlockStats& bs, int i, double x) {
    if constexpr (R == 1) v = x > v ? x : v; else bs.max(i, x);
  }
  __device__ void done(BlockStats& bs, int i) {
    if constexpr (R == 1) bs.max(i, v);
  }
};

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

// trips of a grid-stride loop that every lane of the grid runs (keeps wave-wide
// calls convergent); the element of this lane in trip t is t*stride + gid
__device__ inline int64_t grid_trips(const Grid& g, int64_t n) {
  int64_t stride = g.nb * blockDim.x;
  return (n + stride - 1) / stride;
}
__device__ inline int64_t grid_elem(const Grid& g, int64_t t) {
  return t * g.nb * blockDim.x + g.b * blockDim.x + threadIdx.x;
}

// ------------------------------------------------------------- prepass
// CalcParentFlag_det + SumEContrib + decide_num_to_spawn + walker_death for
// every occupied slot (FciMCPar.F90:1294-1758, fcimc_helper.F90:518-802,
// 1036-1321, 2160-2407).  Blocks own contiguous slot ranges so that the block
// sums of attempt counts feed the scan directly.
template <int W, int R>
__device__ __forceinline__ void k_prepass_body(const Grid& g_, const DevSys& s, const Walk& wk, Counters* cnt, double* part, int32_t* att_cnt, uint8_t* pinfo, int64_t* blk_part, int64_t* blk_rm, IterParams* prm, const IterParams& v, int use_v) {
  if (cnt->halt) return;
  // the iteration's inputs: from the launch when launches are issued one by
  // one (block 0 publishes them for the later kernels), from device memory
  // in a replayed graph
  const double tau = use_v ? v.tau : prm->tau;
  const uint32_t iter = use_v ? v.iter : prm->iter;
  if (use_v && g_.b == 0 && threadIdx.x == 0) *prm = v;
  __shared__ double acc_lds[kStatStride];
  __shared__ int64_t red[kBlock / kWave];
  BlockStats bs{acc_lds};
  bs.zero();
  const int S = R == 1 ? 1 : s.S;
  int64_t nslots = (int64_t)cnt->nslots;
  const bool rebuild = cnt->rebuild != 0;
  if (rebuild) hash_clear_part<W>(g_, wk);  // the attempt scan refills the table
  int64_t begin, end;
  chunk_of(g_, nslots, begin, end);
  int64_t my_att = 0, my_rm = 0;
  bool death_err = false;
  LaneSum<R> a_idets, a_iwalk, a_ndets, a_nwalk, a_hf, a_doubs, a_enum, a_died, a_born, a_abort;
  for (int64_t base = begin; base < end; base += blockDim.x) {
    int64_t slot = base + threadIdx.x;
    bool valid = slot < end;
    // every field of the slot is loaded at once (lanes past the chunk load
    // the step's first slot) and selected afterwards, so that the loads are
    // one round trip to memory rather than three dependent ones
    const int64_t ls = valid ? slot : base;
    const uint32_t f_ld = wk.flags(ls);
    uint64_t k_ld[W];
    double s_ld[R];
#pragma unroll
    for (int w = 0; w < W; ++w) k_ld[w] = wk.keys[ls * W + w];
    for (int r = 0; r < S; ++r) s_ld[r] = wk.sgn(ls, r);
    const double hoff_ld = wk.offdiagH[ls], kii_ld = wk.diagH[ls];
    uint32_t f = valid ? f_ld : kFlagRemoved;
    bool live = !(f & kFlagRemoved);
    uint64_t key[W];
    double sg[R];
    bool occ = false;
#pragma unroll
    for (int w = 0; w < W; ++w) key[w] = live ? k_ld[w] : 0ull;
    for (int r = 0; r < S; ++r) {
      sg[r] = live ? s_ld[r] : 0.0;
      occ = occ || !(fabs(sg[r]) < 1.e-12);
    }
    bool core = (f & kFlagCore) != 0;
    uint8_t pi0 = 0;
    int exlevel = occ ? excit_level<W>(s.ref, key) : -1;
    double hoff = occ ? hoff_ld : 0.0;
    double kii = occ ? kii_ld : 0.0;
    bool isref = occ && det_eq<W>(key, s.ref);
    bool any_left = false;
    double cp[R];
    for (int r = 0; r < S; ++r) {
      // initiator flag (CalcParentFlag_det / TestInitiator_explicit)
      bool now = false;
      if (occ && s.t_init) {
        bool was = (f >> r) & 1u;
        bool pop = fabs(sg[r]) > s.init_walk_no;
        now = was ? (pop || isref || core) : pop;
        f = now ? (f | (1u << r)) : (f & ~(1u << r));
      }
      if (s.t_init) {
        bool ini = occ && now, non = occ && !now;
        a_idets.add(bs, soff(F_NOINITDETS, r), ini ? 1.0 : 0.0);
        a_iwalk.add(bs, soff(F_NOINITWALK, r), ini ? fabs(sg[r]) : 0.0);
        a_ndets.add(bs, soff(F_NONINITDETS, r), non ? 1.0 : 0.0);
        a_nwalk.add(bs, soff(F_NONINITWALK, r), non ? fabs(sg[r]) : 0.0);
      }
      // SumEContrib
      a_hf.add(bs, soff(F_HFCYC, r), exlevel == 0 ? sg[r] : 0.0);
      a_doubs.add(bs, soff(F_NOATDOUBS, r), exlevel == 2 ? fabs(sg[r]) : 0.0);
      a_enum.add(bs, soff(F_ENUM, r), occ ? hoff * sg[r] : 0.0);
      // decide_num_to_spawn and attempt_die share one Philox block
      double died = 0.0, born = 0.0;
      cp[r] = sg[r];
      if (occ) {
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
        uint8_t pi = (uint8_t)((sg[r] < 0.0 ? 1 : 0) | (((f >> r) & 1u) ? 2 : 0) | (core ? 4 : 0));
        if (r == 0) pi0 = pi;  // written with the removal bit below
        else pinfo[slot * S + r] = pi;
        double fac = tau * (kii - (use_v ? v.sft.v[r] : prm->sft.v[r]));
        if (fac > 2.0) death_err = true;
        double idie = core ? 0.0 : attempt_die(s.t_real != 0, fac, fabs(sg[r]), ub);
        double a = fabs(sg[r]);
        died = idie < a ? idie : a;
        double bb = idie - a;
        born = bb > 0.0 ? bb : 0.0;
        cp[r] = sg[r] - idie * (sg[r] >= 0.0 ? 1.0 : -1.0);
        any_left = any_left || (fabs(cp[r]) > 1.0e-12);
      } else if (valid) {
        att_cnt[slot * S + r] = 0;
      }
      a_died.add(bs, soff(F_NDIED, r), died);
      a_born.add(bs, soff(F_NBORN, r), born);
    }
    bool abort_anti = occ && s.t_init && any_left;
    if (abort_anti) any_left = false;
    for (int r = 0; r < S; ++r) {
      double ab = 0.0;
      if (abort_anti) {
        if ((cp[r] > 0.0) != (sg[r] > 0.0)) {
          ab = fabs(cp[r]);
          cp[r] = 0.0;
        }
        any_left = any_left || (fabs(cp[r]) > 1.0e-12);
      }
      a_abort.add(bs, soff(F_NABORTED, r), ab);
    }
    bool rm = occ && !(any_left || core);
    if (occ) {
      for (int r = 0; r < S; ++r) wk.sgn(slot, r) = rm ? 0.0 : cp[r];
      wk.flags(slot) = rm ? (f | kFlagRemoved) : f;
      if (rm && !rebuild) {  // RemoveHashDet: the entry becomes a tombstone
        int64_t hp = wk.hpos(slot);
        int32_t* sl = (int32_t*)(wk.htab + (size_t)(hp >> 2) * 8 * W + 4 * W);
        sl[hp & 3] = kSlotTomb;
      }
    }
    // bit 3 of the first replica's byte: freed here; the attempt scan queues
    // these slots in the free list (in slot order, no atomics here)
    if (valid) pinfo[slot * S] = (uint8_t)(pi0 | (rm ? kPinfoFreed : 0));
    my_rm += rm ? 1 : 0;
  }
  if (death_err) atomicOr(&cnt->err, 1u << NGPU_ERR_DEATH_PROB);
  a_idets.done(bs, soff(F_NOINITDETS, 0));
  a_iwalk.done(bs, soff(F_NOINITWALK, 0));
  a_ndets.done(bs, soff(F_NONINITDETS, 0));
  a_nwalk.done(bs, soff(F_NONINITWALK, 0));
  a_hf.done(bs, soff(F_HFCYC, 0));
  a_doubs.done(bs, soff(F_NOATDOUBS, 0));
  a_enum.done(bs, soff(F_ENUM, 0));
  a_died.done(bs, soff(F_NDIED, 0));
  a_born.done(bs, soff(F_NBORN, 0));
  a_abort.done(bs, soff(F_NABORTED, 0));
  int64_t tot;
  (void)block_exclusive_scan(my_att, red, tot);
  if (threadIdx.x == 0) blk_part[g_.b] = tot;
  (void)block_exclusive_scan(my_rm, red, tot);
  if (threadIdx.x == 0) blk_rm[g_.b] = tot;
  bs.flush(part, g_.b);
}
template <int W, int R>
__global__ void __launch_bounds__(kBlock) k_prepass(DevSys s, Walk wk, Counters* cnt, double* part, int32_t* att_cnt, uint8_t* pinfo, int64_t* blk_part, int64_t* blk_rm, IterParams* prm, IterParams v, int use_v) {
  k_prepass_body<W, R>(hw_grid(), s, wk, cnt, part, att_cnt, pinfo, blk_part, blk_rm, prm, v, use_v);
}


// attempt offsets per (slot, replica) pair, same chunking as k_prepass; also
// the parent pair of the first attempt of every kBlock-attempt tile (the
// spawn kernel's tiles), written by the pair whose attempts contain it.  When
// a rebuild is due, also the refill of the walker hash that the prepass
// cleared, before any lookup of the iteration.
template <int W>
__device__ __forceinline__ void k_scan_apply_body(const Grid& g_, const DevSys& s, const Walk& wk, const int32_t* att_cnt, int64_t* att_off, const int64_t* part, const int64_t* part_rm, const uint8_t* pinfo, Counters* cnt, int S, int64_t* tile_p, int64_t tcap) {
  if (cnt->halt) return;
  __shared__ int64_t red[kBlock / kWave];
  int64_t nslots = (int64_t)cnt->nslots;
  if (cnt->rebuild) {
    hash_refill_part<W>(g_, s, wk, cnt, nslots);
    // the prepass's removals left no tombstones: the table now holds the live
    // slots only (the end-of-iteration accounting counts the later removals)
    if (g_.b == 0 && threadIdx.x == 0) {
      cnt->n_tomb = 0;
      cnt->rebuilt = 1;
      cnt->n_rebuilds += 1;
    }
  }
  int64_t begin, end;
  chunk_of(g_, nslots, begin, end);
  // this block's first attempt offset and first free-queue position: the
  // counts of the blocks before it.  The slots the prepass freed go, in slot
  // order, after the free queue's tail (the prepass appends nothing itself,
  // so the tail is still free_tail0, the value at the last accounting).
  // With one replica both scans run as one over packed counts (attempts in
  // the low 40 bits, freed slots above).
  const unsigned long long tail0 = cnt->free_tail0;
  constexpr int kPackShift = 40;
  int64_t loc = 0, loc_rm = 0;
  for (int q = threadIdx.x; q < (int)g_.b; q += blockDim.x) {
    loc += part[q];
    loc_rm += part_rm[q];
  }
  int64_t run, run_rm;
  (void)block_exclusive_scan(loc, red, run);
  (void)block_exclusive_scan(loc_rm, red, run_rm);
  if (S == 1) {
    for (int64_t b0 = begin; b0 < end; b0 += blockDim.x) {
      int64_t i = b0 + threadIdx.x;
      int64_t v = 0, f = 0;
      if (i < end) {
        v = att_cnt[i];
        f = (pinfo[i] & kPinfoFreed) ? 1 : 0;
      }
      int64_t tot;
      int64_t pre = block_exclusive_scan(v | (f << kPackShift), red, tot);
      if (i < end) {
        int64_t o = run + (pre & ((1ll << kPackShift) - 1));
        att_off[i] = o;
        for (int64_t t = (o + kBlock - 1) / kBlock; t * kBlock < o + v && t < tcap; ++t) tile_p[t] = i;
        if (f) wk.freeq[(tail0 + (unsigned long long)(run_rm + (pre >> kPackShift))) % (unsigned long long)wk.cap] = i;
      }
      run += tot & ((1ll << kPackShift) - 1);
      run_rm += tot >> kPackShift;
    }
  } else {
    for (int64_t b0 = begin; b0 < end; b0 += blockDim.x) {
      int64_t slot = b0 + threadIdx.x;
      int64_t f = (slot < end && (pinfo[slot * S] & kPinfoFreed)) ? 1 : 0;
      int64_t tot;
      int64_t pre = block_exclusive_scan(f, red, tot);
      if (f) wk.freeq[(tail0 + (unsigned long long)(run_rm + pre)) % (unsigned long long)wk.cap] = slot;
      run_rm += tot;
    }
    int64_t pb = begin * S, pe = end * S;
    for (int64_t base = pb; base < pe; base += blockDim.x) {
      int64_t i = base + threadIdx.x;
      int64_t v = i < pe ? att_cnt[i] : 0;
      int64_t tot;
      int64_t pre = block_exclusive_scan(v, red, tot);
      if (i < pe) {
        int64_t o = run + pre;
        att_off[i] = o;
        for (int64_t t = (o + kBlock - 1) / kBlock; t * kBlock < o + v && t < tcap; ++t) tile_p[t] = i;
      }
      run += tot;
    }
  }
  if (g_.b == g_.nb - 1 && threadIdx.x == 0) {
    att_off[nslots * S] = run;
    cnt->n_attempts = (unsigned long long)run;
    cnt->free_tail = tail0 + (unsigned long long)run_rm;
    cnt->rm_pre = (unsigned long long)run_rm;
  }
}
template <int W>
__global__ void __launch_bounds__(kBlock) k_scan_apply(DevSys s, Walk wk, const int32_t* att_cnt, int64_t* att_off, const int64_t* part, const int64_t* part_rm, const uint8_t* pinfo, Counters* cnt, int S, int64_t* tile_p, int64_t tcap) {
  k_scan_apply_body<W>(hw_grid(), s, wk, att_cnt, att_off, part, part_rm, pinfo, cnt, S, tile_p, tcap);
}

// ------------------------------------------------------------- spawning
// parent (slot, replica) of attempt a: last pair p in [lo, hi] with
// att_off[p] <= a (the caller's bounds bracket the answer)
__device__ inline int64_t parent_pair(const int64_t* att_off, int64_t lo, int64_t hi, int64_t a) {
  while (lo < hi) {
    int64_t mid = (lo + hi + 1) >> 1;
    if (att_off[mid] <= a) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// LDS stages of the spawned list and of the queued single-excitation attempts
constexpr int kSpawnStage = 576;
constexpr int kParentWindow = 512;  // attempt offsets of one tile's parents staged in LDS
constexpr int kQueueStage = 512;
template <int W>
struct SpawnStage {
  Stage st;
  uint64_t key[kSpawnStage * W];
  double val[kSpawnStage];
  uint32_t info[kSpawnStage];
};
struct QueueStage {  // queued attempts as (parent pair, attempt number)
  Stage st;
  int64_t p[kQueueStage];
  uint32_t k[kQueueStage];
};

// A spawned-list record with FP64 amplitudes is kRecWords 64-bit words (32
// bytes, so a record never straddles a cache line): the determinant's words,
// the child amplitude (FP64 bits) in word 2 and the info word (bits 0-7
// replica, bit 8 parent initiator) in word 3.  With FP32 amplitudes it is
// W + 1 words: the determinant's words, then info << 32 | the amplitude's FP32
// bits (16 bytes for one-word determinants).  The merge reads one record per
// spawn and sums the amplitudes in FP64.
constexpr int kRecWords = 4;  // the larger of the two record sizes
template <int W>
__device__ inline double rec_val(const DevSys& s, const uint64_t* r) {
  return s.spawn_fp32 ? (double)__uint_as_float((uint32_t)r[W]) : __longlong_as_double((long long)r[2]);
}
template <int W>
__device__ inline uint32_t rec_info(const DevSys& s, const uint64_t* r) {
  return s.spawn_fp32 ? (uint32_t)(r[W] >> 32) : (uint32_t)r[3];
}

// the spawn sort's 32-bit key is the colex rank times this odd constant: a
// bijection of 32-bit values that spreads the keys' top byte evenly
constexpr uint32_t kRankMul = 0x9E3779B1u;

// colex rank of a determinant: sum over its i-th occupied orbital (bit p,
// ascending) of C(p, i); injective on determinants of nel electrons
template <int W>
__device__ inline uint32_t colex_rank(const uint32_t* binom, int nel1, const uint64_t* d) {
  uint32_t r = 0;
  int c = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    uint64_t x = d[w];
    while (x) {
      int p = 64 * w + ctz64(x);
      x &= x - 1;
      ++c;
      r += binom[p * nel1 + c];
    }
  }
  return r;
}

// binomial table of the colex rank, staged in LDS by the spawn kernels (the
// engine ranks only when nbasis*(nel+1) fits)
constexpr int kBinomLds = 512;
struct BinomLds {
  uint32_t c[kBinomLds];
  __device__ void load(const DevSys& s, const uint32_t* rk) {
    if (rk)
      for (int i = threadIdx.x; i < s.nbasis * s.nel1; i += blockDim.x) c[i] = s.binom[i];
    __syncthreads();
  }
};

// block-wide: drain staged spawns into the spawned list when due; with rk,
// also each spawn's colex rank (the first sort key) at the same index
template <int W>
__device__ inline void spawn_flush(SpawnStage<W>* ss, Counters* cnt, uint64_t* sp_rec, int64_t scap, bool force,
                                   bool& overflow, const DevSys& s, const BinomLds& bl, uint32_t* rk) {
  uint32_t n;
  unsigned long long base;
  if (!stage_flush_begin(&ss->st, &cnt->n_spawned, kSpawnStage, force, n, base)) return;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    unsigned long long d = base + i;
    if (d < (unsigned long long)scap) {
      uint64_t* r = sp_rec + d * (unsigned long long)s.rec_words;
#pragma unroll
      for (int w = 0; w < W; ++w) r[w] = ss->key[i * W + w];
      if (s.spawn_fp32) {  // the value is already FP32-representable (attempt_create)
        r[W] = ((uint64_t)ss->info[i] << 32) | __float_as_uint((float)ss->val[i]);
      } else {
        if (W == 1) r[1] = 0;
        r[2] = (uint64_t)__double_as_longlong(ss->val[i]);
        r[3] = ss->info[i];
      }
      if (rk) rk[d] = colex_rank<W>(bl.c, s.nel1, ss->key + i * W) * kRankMul;
    } else {
      overflow = true;
    }
  }
  stage_flush_end(&ss->st);
}

__device__ inline void queue_flush(QueueStage* qs, Counters* cnt, int64_t* sq, int64_t scap, bool force,
                                   bool& overflow) {
  uint32_t n;
  unsigned long long base;
  if (!stage_flush_begin(&qs->st, &cnt->n_singles, kQueueStage, force, n, base)) return;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    unsigned long long d = base + i;
    if (d < (unsigned long long)scap) {
      sq[2 * d] = qs->p[i];
      sq[2 * d + 1] = (int64_t)qs->k[i];
    } else {
      overflow = true;
    }
  }
  stage_flush_end(&qs->st);
}

// is_core_state (semi_stoch_procs.F90:590-605) of the child of an excitation:
// core dets are always resident, so the walker hash and its core flag answer it
template <int W>
__device__ inline bool core_child(const DevSys& s, const Walk& wk, Counters* cnt, const uint64_t* key,
                                  const Excit& ex) {
  uint64_t child[W];
  apply_excit<W>(key, ex, child);
  int64_t home = find_walker_hash<W>(child, s.rh, s.L) - 1;
  int64_t pos = hash_lookup<W>(wk, child, home, cnt);
  return pos >= 0 && (wk.flags(pos) & kFlagCore);
}

// statistics of the spawning kernels, per lane until the end of the kernel
template <int R>
struct SpawnAcc {
  LaneSum<1> valid, invalid, bloomc1, bloomc2;
  LaneMax<1> maxspawn, blooms1, blooms2;
  LaneSum<R> born, sing;
  __device__ void done(BlockStats& bs, int mode) {
    valid.done(bs, S_NVALID);
    invalid.done(bs, S_NINVALID);
    maxspawn.done(bs, S_MAXSPAWN);
    born.done(bs, soff(F_NBORN, 0));
    if (mode != 1) {
      sing.done(bs, soff(F_SPAWNSING, 0));
      bloomc1.done(bs, S_BLOOMC + 1);
      blooms1.done(bs, S_BLOOMS + 1);
    }
    if (mode != 2) {
      bloomc2.done(bs, S_BLOOMC + 2);
      blooms2.done(bs, S_BLOOMS + 2);
    }
  }
};

// generator bound at compile time, one k_spawn instantiation each, so that the
// code of one generator (the uniform one's private arrays in particular) costs
// the others nothing
// generator instances of the spawn kernel: the lattice generator, PCHB with
// FAST-FAST holes and UNIF-UNIF / UNIF-FAST particles (a fixed number of
// uniforms), the uniform molecular generator, and PCHB with weighted
// particles or holes (constrained samplers that may redraw)
constexpr int kGenHubbard = 0, kGenPchb = 1, kGenUniform = 2, kGenPchbW = 3;

// one spawning attempt (attempt_create_normal + create_particle); mode 0: any
// excitation, 1: doubles only (singles are queued), 2: singles only.  Wave-wide.
template <int W, int R, int G>
__device__ inline void spawn_attempt(const DevSys& s, const Walk& wk, Counters* cnt, BlockStats& bs,
                                     SpawnAcc<R>& acc, const uint8_t* pinfo, bool active, int64_t p, uint32_t k,
                                     int mode, SpawnStage<W>* ss, QueueStage* qs, double tau, uint32_t iter) {
  const int S = R == 1 ? 1 : s.S;
  bool emit = false, queue = false;
  uint64_t child[W];
  double nspawn = 0.0, an = 0.0, raw = 0.0;
  uint32_t info = 0;
  int r = 0, ic = 0;
  double valid = 0.0, invalid = 0.0;
  if (active) {
    int64_t slot = p / S;
    r = (int)(p - slot * S);
    uint64_t key[W];
#pragma unroll
    for (int w = 0; w < W; ++w) key[w] = wk.keys[slot * W + w];
    double dr[6];
    draw2<W>(s.seed, key, iter, r, k, 1, dr[0], dr[1]);
    constexpr bool kPchb = G == kGenPchb || G == kGenPchbW;
    bool single = kPchb && dr[0] < s.psingles;
    if (mode == 1 && single) {
      queue = true;
    } else {
      // purpose 2: the PCHB hole pair (d3) and the spawn rounding (d5); purpose
      // 3 (d2, d4) only for the generators that read them
      draw2<W>(s.seed, key, iter, r, k, 2, dr[3], dr[5]);
      dr[2] = dr[4] = 0.0;
      if (G == kGenUniform || G == kGenPchbW || (G == kGenPchb && (s.particles == 1 || single)))
        draw2<W>(s.seed, key, iter, r, k, 3, dr[2], dr[4]);
      Excit ex;
      if constexpr (G == kGenHubbard) {
        ex = gen_hubbard<W>(s, key, dr);
      } else if constexpr (G == kGenUniform) {
        bool fail = false;
        ex = gen_excit<W>(s, key, dr, iter, r, k, &fail);
        if (fail) atomicOr(&cnt->err, 1u << NGPU_ERR_EXCITGEN);
      } else if (single) {
        ex = gen_single_heatbath<W>(s, key, dr[1], dr[2]);
        ex.pgen = ex.pgen * s.psingles;
      } else if constexpr (G == kGenPchbW) {
        DrawSeq<W> u(dr, s.seed, key, iter, r, k);
        ex = gen_double_pchb_weighted<W>(s, key, u);
        ex.pgen = ex.pgen * s.pdoubles;
      } else {
        ex = gen_double_pchb<W>(s, key, dr);
        ex.pgen = ex.pgen * s.pdoubles;
      }
      if (ex.ic == 0) {
        invalid = 1.0;
      } else if (s.ncore && (pinfo[p] & 4) && core_child<W>(s, wk, cnt, key, ex)) {
        valid = 1.0;  // spawning from and to the core space is cancelled
      } else {
        valid = 1.0;
        ic = ex.ic;
        double hel = ex.hel_ok ? (ex.parity ? -ex.hel : ex.hel)
                               : helement_ex<W>(s, key, ex.ic, ex.e0, ex.e1, ex.e2, ex.e3, ex.parity);
        uint8_t pi = pinfo[p];
        double wsign = (pi & 1) ? -1.0 : 1.0;
        nspawn = attempt_create(s, tau, hel, ex.pgen, wsign, dr[5], raw);
        raw = fabs(raw);
        if (fabs(nspawn) > kEps) {
          emit = true;
          an = fabs(nspawn);
          apply_excit<W>(key, ex, child);
          info = (uint32_t)r | ((s.t_init && (pi & 2)) ? kInfoInit : 0u);
        }
      }
    }
  }
  int qi = qs ? stage_push(queue, &qs->st) : -1;
  if (queue) {
    qs->p[qi] = p;
    qs->k[qi] = k;
  }
  acc.invalid.add(bs, S_NINVALID, invalid);
  acc.valid.add(bs, S_NVALID, valid);
  acc.maxspawn.add(bs, S_MAXSPAWN, raw);
  for (int rr = 0; rr < S; ++rr) {
    acc.born.add(bs, soff(F_NBORN, rr), rr == r ? an : 0.0);
    if (mode != 1) acc.sing.add(bs, soff(F_SPAWNSING, rr), (rr == r && ic == 1) ? an : 0.0);
  }
  bool bloom = emit && an > s.init_walk_no;
  if (mode != 2) {
    acc.bloomc2.add(bs, S_BLOOMC + 2, (bloom && ic == 2) ? 1.0 : 0.0);
    acc.blooms2.add(bs, S_BLOOMS + 2, (bloom && ic == 2) ? an : 0.0);
  }
  if (mode != 1) {
    acc.bloomc1.add(bs, S_BLOOMC + 1, (bloom && ic == 1) ? 1.0 : 0.0);
    acc.blooms1.add(bs, S_BLOOMS + 1, (bloom && ic == 1) ? an : 0.0);
  }
  int idx = stage_push(emit, &ss->st);
  if (emit) {
#pragma unroll
    for (int w = 0; w < W; ++w) ss->key[idx * W + w] = child[w];
    ss->val[idx] = nspawn;
    ss->info[idx] = info;
  }
}

// all attempts: Hubbard spawns directly; molecular doubles spawn here while
// singles (the heat-bath list is ~20x longer per attempt) go to a queue so that
// waves do not idle on a few long lanes
template <int W, int R, int G>
__device__ __forceinline__ void k_spawn_body(const Grid& g_, const DevSys& s, const Walk& wk, Counters* cnt, double* part, const int64_t* att_off, const int64_t* tile_p, int64_t tcap, const uint8_t* pinfo, uint64_t* sp_rec, uint32_t* rk, int64_t scap, int64_t* sq, const IterParams* prm) {
  if (cnt->halt) return;
  const double tau = prm->tau;
  const uint32_t iter = prm->iter;
  __shared__ double acc_lds[kStatStride];
  __shared__ SpawnStage<W> ss;
  __shared__ BinomLds bl;
  bl.load(s, rk);
  __shared__ QueueStage qs;
  BlockStats bs{acc_lds};
  bs.zero();
  stage_init(&ss.st);
  stage_init(&qs.st);
  const int64_t total = (int64_t)cnt->n_attempts;
  const int64_t npairs = (int64_t)cnt->nslots * s.S;
  bool overflow = false;
  constexpr int mode = (G == kGenPchb || G == kGenPchbW) ? 1 : 0;
  SpawnAcc<R> acc;
  __shared__ int64_t pb[2];
  __shared__ int64_t woff[kParentWindow];
  int64_t trips = grid_trips(g_, total);
  for (int64_t t = 0; t < trips; ++t) {
    int64_t a = grid_elem(g_, t);
    // the block's attempts are the consecutive tile a0 .. a0+kBlock-1: its
    // parents lie between the tile's first parent and the next tile's
    int64_t a0 = grid_elem(g_, t) - threadIdx.x;
    int64_t tile = a0 / kBlock;
    if (threadIdx.x == 0) {
      int64_t lo = 0, hi = 0;
      if (a0 < total) {
        bool more = a0 + kBlock < total;
        if (tile + 1 < tcap) {
          lo = tile_p[tile];
          hi = more ? tile_p[tile + 1] : npairs - 1;
        } else {  // beyond the tile table: full searches
          int64_t a1 = (more ? a0 + kBlock : total) - 1;
          lo = parent_pair(att_off, 0, npairs - 1, a0);
          hi = parent_pair(att_off, lo, npairs - 1, a1);
        }
      }
      pb[0] = lo;
      pb[1] = hi;
    }
    __syncthreads();
    const int64_t lo = pb[0], hi = pb[1];
    const int64_t nw = hi - lo + 1;
    bool act = a < total;
    int64_t p = 0;
    uint32_t k = 0;
    if (a0 < total && nw <= kParentWindow) {  // block-uniform: search the window in LDS
      for (int64_t i = threadIdx.x; i < nw; i += blockDim.x) woff[i] = att_off[lo + i];
      __syncthreads();
      if (act) {
        int64_t l = 0, h = nw - 1;
        while (l < h) {
          int64_t mid = (l + h + 1) >> 1;
          if (woff[mid] <= a) l = mid; else h = mid - 1;
        }
        p = lo + l;
        k = (uint32_t)(a - woff[l]);
      }
    } else if (act) {
      p = parent_pair(att_off, lo, hi, a);
      k = (uint32_t)(a - att_off[p]);
    }
    spawn_attempt<W, R, G>(s, wk, cnt, bs, acc, pinfo, act, p, k, mode, &ss, &qs, tau, iter);
    spawn_flush<W>(&ss, cnt, sp_rec, scap, false, overflow, s, bl, rk);
    if (mode == 1) queue_flush(&qs, cnt, sq, scap, false, overflow);
  }
  spawn_flush<W>(&ss, cnt, sp_rec, scap, true, overflow, s, bl, rk);
  if (mode == 1) queue_flush(&qs, cnt, sq, scap, true, overflow);
  if (overflow) atomicOr(&cnt->err, 1u << NGPU_ERR_SPAWN_FULL);
  acc.done(bs, mode);
  bs.flush(part, g_.b);
}
template <int W, int R, int G>
__global__ void __launch_bounds__(kBlock) k_spawn(DevSys s, Walk wk, Counters* cnt, double* part, const int64_t* att_off, const int64_t* tile_p, int64_t tcap, const uint8_t* pinfo, uint64_t* sp_rec, uint32_t* rk, int64_t scap, int64_t* sq, const IterParams* prm) {
  k_spawn_body<W, R, G>(hw_grid(), s, wk, cnt, part, att_off, tile_p, tcap, pinfo, sp_rec, rk, scap, sq, prm);
}

// Queued single excitations (molecular, heat-bath singles), G lanes per
// attempt (G = 4, 8 or 16, a launch-shape choice).  gen_single_heatbath's two
// sequential sweeps need the weight |hb_weight(src -> b)| of every same-spin
// orbital b; lane g of a group evaluates the orbitals b = g+1, g+1+G, ...:
// the reference-relative element fref[src][b] (one contiguous row), then the
// (Coulomb, exchange) pairs jk[a][j][b] of the orbitals j in which the parent
// and the reference differ (contiguous in b, so a group's loads for one j are
// one contiguous run), added in hb_weight's order, and stores it in LDS; both
// sweeps then run in ascending b over LDS on every lane of the group, so the
// selected orbital, pgen and the spawn's element (the selected weight, signed)
// are bit-identical to the one-lane generator.
// The attempt's three Philox blocks are spread over the group's lanes and
// shared.  More lanes per attempt shorten each attempt's chain of loads
// (measured at 20 spatial orbitals: 1, 2 and 4 lanes took 116, 64 and 46 us
// for 1e5 attempts); fewer lanes would repeat the sequential sweeps on fewer
// lanes.

// Adds sign * (U(a,j,b,j) - [same spin] U(a,j,j,b)) into v for the orbitals j
// of x (ascending, src removed), two orbitals per step so that both pairs'
// loads are in flight before the ordered additions.
template <int PL, int G>
__device__ inline void hb_corr(uint64_t x, int base, const double* row, int nsp, int g, int mcount, bool plus,
                               int src, double* v) {
  while (x) {
    int o1 = base + ctz64(x) + 1;
    x &= x - 1;
    bool two = x != 0;
    int o2 = two ? base + ctz64(x) + 1 : o1;
    if (two) x &= x - 1;
    const bool s1 = ((o1 ^ src) & 1) == 0, s2 = ((o2 ^ src) & 1) == 0;
    const double2* r1 = (const double2*)(row + (size_t)2 * (spat(o1) - 1) * nsp) + g;
    const double2* r2 = (const double2*)(row + (size_t)2 * (spat(o2) - 1) * nsp) + g;
    double2 a1[PL], a2[PL];
#pragma unroll
    for (int m = 0; m < PL; ++m) {
      bool in = m < mcount && g + m * G < nsp;
      a1[m] = in ? r1[m * G] : double2{0.0, 0.0};
      a2[m] = (in && two) ? r2[m * G] : double2{0.0, 0.0};
    }
#pragma unroll
    for (int m = 0; m < PL; ++m) {
      if (m < mcount && g + m * G < nsp) {
        if (plus) {
          v[m] += a1[m].x;
          if (s1) v[m] -= a1[m].y;
          if (two) {
            v[m] += a2[m].x;
            if (s2) v[m] -= a2[m].y;
          }
        } else {
          v[m] -= a1[m].x;
          if (s1) v[m] += a1[m].y;
          if (two) {
            v[m] -= a2[m].x;
            if (s2) v[m] += a2[m].y;
          }
        }
      }
    }
  }
}

// PL: spatial orbitals per lane (nsp <= G*PL); dynamic LDS holds the
// (kBlock/G)*nsp values of the block's attempts
template <int W, int R, int PL, int G>
__device__ __forceinline__ void k_spawn_singles_body(const Grid& g_, const DevSys& s, const Walk& wk, Counters* cnt, double* part, const uint8_t* pinfo, uint64_t* sp_rec, uint32_t* rk, int64_t scap, const int64_t* sq, const IterParams* prm) {
  if (cnt->halt) return;
  constexpr int kGroups = kBlock / G;
  const double tau = prm->tau;
  const uint32_t iter = prm->iter;
  __shared__ double acc_lds[kStatStride];
  __shared__ SpawnStage<W> ss;
  __shared__ BinomLds bl;
  bl.load(s, rk);
  // per group: the signed H(src -> b) of its attempt, then their running sums
  extern __shared__ double hv_dyn[];
  BlockStats bs{acc_lds};
  bs.zero();
  stage_init(&ss.st);
  const int S = R == 1 ? 1 : s.S;
  int64_t nq = (int64_t)cnt->n_singles;
  if (nq > scap) nq = scap;
  bool overflow = false;
  SpawnAcc<R> acc;
  const int nsp = s.nbasis / 2;
  const int mcount = (nsp + G - 1) / G;  // wave-uniform
  const int g = threadIdx.x % G;
  const int gl = threadIdx.x / G;        // group within the block
  const int gbase = lane_id() - lane_id() % G;
  const bool leader = g == 0;
  double* myv = hv_dyn + (size_t)gl * 2 * nsp;
  double* mycum = myv + nsp;
  const int64_t groups = (int64_t)g_.nb * kGroups;
  const int64_t grp = (int64_t)g_.b * kGroups + gl;
  const int64_t trips = (nq + groups - 1) / groups;
  for (int64_t t = 0; t < trips; ++t) {
    // blocks whose groups are all past the queue skip the attempt (block-
    // uniform, so the body's barrier is reached by all of the block or none)
    if (t * groups + (int64_t)g_.b * kGroups < nq) {
      int64_t q = t * groups + grp;
      bool act = q < nq;
      int64_t p = act ? sq[2 * q] : 0;
      uint32_t k = act ? (uint32_t)sq[2 * q + 1] : 0u;
      int64_t slot = p / S;
      int r = (int)(p - slot * S);
      uint64_t key[W];
#pragma unroll
      for (int w = 0; w < W; ++w) key[w] = act ? wk.keys[slot * W + w] : s.ref[w];  // idle groups: any valid det
      // lane g < 3 draws Philox block g+1, all three in one pass of the wave;
      // d1 = dsel (purpose 1), d5 = spawn rounding (purpose 2, second value),
      // d2 = dorb (purpose 3, first)
      static_assert(G >= 3, "three lanes of a group draw the attempt's Philox blocks");
      double ua = 0.0, ub = 0.0;
      if (g < 3) draw2<W>(s.seed, key, iter, r, k, (uint32_t)(g + 1), ua, ub);
      const double d1 = __shfl(ub, gbase + 0);
      const double d5 = __shfl(ub, gbase + 1);
      const double d2 = __shfl(ua, gbase + 2);
      int elec = 1 + (int)floor(d1 * (double)s.nel);
      int src = select_orb<W>(key, ~0ull, elec);
      int a = spat(src), off = is_beta(src) ? 1 : 0;
      const double* row = s.jk + (size_t)2 * (a - 1) * nsp * nsp;
      const double* frow = s.fref + (size_t)(src - 1) * nsp;
      double v[PL];
#pragma unroll
      for (int m = 0; m < PL; ++m) v[m] = (m < mcount && g + m * G < nsp) ? frow[g + m * G] : 0.0;
      // the orbitals the parent occupies beyond the reference, then the
      // reference's orbitals the parent lacks (hb_weight's order)
#pragma unroll
      for (int w = 0; w < W; ++w) {
        uint64_t x = key[w] & ~s.ref[w];
        if ((src - 1) >> 6 == w) x &= ~(1ull << ((src - 1) & 63));
        hb_corr<PL, G>(x, 64 * w, row, nsp, g, mcount, true, src, v);
      }
#pragma unroll
      for (int w = 0; w < W; ++w) {
        uint64_t x = s.ref[w] & ~key[w];
        if ((src - 1) >> 6 == w) x &= ~(1ull << ((src - 1) & 63));
        hb_corr<PL, G>(x, 64 * w, row, nsp, g, mcount, false, src, v);
      }
#pragma unroll
      for (int m = 0; m < PL; ++m) {
        int b = g + m * G + 1;
        if (m < mcount && b <= nsp) {
          int orb = 2 * b - off;
          myv[b - 1] = test_orb<W>(key, orb) ? 0.0 : v[m];
        }
      }
      __syncthreads();
      // the normalisation, keeping the running sums (every lane of the group
      // writes the same values, and reads its own writes below); then the
      // first b whose running sum reaches r, by bisection over those sums:
      // the running sums of a second sweep in the same order are the same
      // values, so this is its selection (the last b when none reaches r)
      double cum_sum = 0.0;
      for (int b = 0; b < nsp; ++b) {
        cum_sum += fabs(myv[b]);
        mycum[b] = cum_sum;
      }
      double rr = d2 * cum_sum;
      int lo = 0, hi = nsp - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (!(mycum[mid] < rr)) hi = mid;
        else lo = mid + 1;
      }
      const int bsel = lo + 1;
      const double vsel = myv[lo];
      const double cpt = fabs(vsel);
      bool mine_att = act && leader;
      bool emit = false;
      uint64_t child[W];
      double nspawn = 0.0, an = 0.0, raw = 0.0, valid = 0.0, invalid = 0.0;
      uint32_t info = 0;
      bool cancel = false;
      if (mine_att && s.ncore && (pinfo[p] & 4) && !(fabs(cum_sum) <= kEps || cpt == 0.0)) {
        Excit ex{};
        ex.ic = 1;
        ex.e0 = src;
        ex.e1 = 2 * bsel - off;
        cancel = core_child<W>(s, wk, cnt, key, ex);  // from and to the core space
      }
      if (mine_att) {
        if (fabs(cum_sum) <= kEps || cpt == 0.0) {
          invalid = 1.0;
        } else if (cancel) {
          valid = 1.0;
        } else {
          valid = 1.0;
          int tgt = 2 * bsel - off;
          double pg = cpt / cum_sum;
          pg = pg / (double)s.nel;
          pg = pg * s.psingles;
          double hel = single_parity<W>(key, src, tgt) ? -vsel : vsel;  // the weight's own sum
          uint8_t pi = pinfo[p];
          double wsign = (pi & 1) ? -1.0 : 1.0;
          nspawn = attempt_create(s, tau, hel, pg, wsign, d5, raw);
          raw = fabs(raw);
          if (fabs(nspawn) > kEps) {
            emit = true;
            an = fabs(nspawn);
#pragma unroll
            for (int w = 0; w < W; ++w) child[w] = key[w];
            flip_orb<W>(child, src);
            flip_orb<W>(child, tgt);
            info = (uint32_t)r | ((s.t_init && (pi & 2)) ? kInfoInit : 0u);
          }
        }
      }
      acc.invalid.add(bs, S_NINVALID, invalid);
      acc.valid.add(bs, S_NVALID, valid);
      acc.maxspawn.add(bs, S_MAXSPAWN, raw);
      for (int rr2 = 0; rr2 < S; ++rr2) {
        acc.born.add(bs, soff(F_NBORN, rr2), rr2 == r ? an : 0.0);
        acc.sing.add(bs, soff(F_SPAWNSING, rr2), rr2 == r ? an : 0.0);
      }
      bool bloom = emit && an > s.init_walk_no;
      acc.bloomc1.add(bs, S_BLOOMC + 1, bloom ? 1.0 : 0.0);
      acc.blooms1.add(bs, S_BLOOMS + 1, bloom ? an : 0.0);
      int idx = stage_push(emit, &ss.st);
      if (emit) {
#pragma unroll
        for (int w = 0; w < W; ++w) ss.key[idx * W + w] = child[w];
        ss.val[idx] = nspawn;
        ss.info[idx] = info;
      }
    }
    spawn_flush<W>(&ss, cnt, sp_rec, scap, false, overflow, s, bl, rk);  // its barriers also order the reuse of hv
  }
  spawn_flush<W>(&ss, cnt, sp_rec, scap, true, overflow, s, bl, rk);
  if (overflow) atomicOr(&cnt->err, 1u << NGPU_ERR_SPAWN_FULL);
  acc.done(bs, 2);
  bs.flush(part, g_.b);
}
template <int W, int R, int PL, int G>
__global__ void __launch_bounds__(kBlock) k_spawn_singles(DevSys s, Walk wk, Counters* cnt, double* part, const uint8_t* pinfo, uint64_t* sp_rec, uint32_t* rk, int64_t scap, const int64_t* sq, const IterParams* prm) {
  k_spawn_singles_body<W, R, PL, G>(hw_grid(), s, wk, cnt, part, pinfo, sp_rec, rk, scap, sq, prm);
}

// ---------------------------------------------------------- radix sort
// Stable LSD radix sort, 8-bit digits, reduce-then-scan: per-block digit
// histograms, a per-digit scan over blocks, stable block-local scatter.  The
// merge only needs equal determinants to be adjacent (and, within a run, the
// spawn order, which a stable sort keeps), so the sort key may be any
// injective function of the determinant.  When C(nbasis, nel) < 2^32 the key
// is the determinant's 32-bit colex rank (sum over its i-th occupied bit at
// position p of C(p, i)) times an odd constant; for one-word determinants
// with C(nbasis, nel) < 2^64 whose rank is shorter than the word, the 64-bit
// rank times an odd constant modulo 2^bits.  Both are sorted by one pass on
// their top digit and a per-segment sort (k_seg_sort, k_seg_sort64).
// Otherwise the key is the determinant itself, ordered as ilut_lt orders it,
// in LSD passes.
constexpr int kRadixBins = 1024;  // bins of the widest digit (10 bits)
// radix tiles: IT elements per thread, IT*kBlock per tile (IT = 2, 4 or 8,
// chosen per engine at run time)

// element range of this block in the histogram and scatter passes: whole
// tiles, and no more blocks than there are tiles (blocks past the last tile
// get an empty range), so that a small spawned list is not spread over more
// blocks than it fills
template <int IT>
__device__ inline void radix_chunk(const Grid& g, int64_t n, int64_t& begin, int64_t& end) {
  constexpr int kRadixTile = IT * kBlock;
  int64_t tiles = (n + kRadixTile - 1) / kRadixTile;
  int64_t nb = tiles < g.nb ? tiles : g.nb;
  if (nb < 1) nb = 1;
  int64_t per = (tiles + nb - 1) / nb * kRadixTile;
  begin = g.b * per;
  end = begin + per;
  if (end > n) end = n;
  if (begin > n) begin = n;
}

// sort key: NW words of type K per element; FLIP compares word 0 as signed
// DB-bit digit (DB = 8 or 10; kRadixBins is the largest digit range)
template <typename K, int NW, bool FLIP, int DB = 8>
__device__ inline uint32_t radix_digit(const K* k, int word, int shift) {
  K x = k[word];
  if (FLIP && word == 0) x ^= (K)1 << (sizeof(K) * 8 - 1);
  return (uint32_t)(x >> shift) & ((1u << DB) - 1u);
}

template <typename K, int NW, bool FLIP, int IT, int DB>
__device__ __forceinline__ void k_radix_hist_body(const Grid& g_, const K* keys, int kstride, const unsigned long long* n_ptr, int64_t ncap, int word, int shift, uint32_t* hist) {
  constexpr int RB = 1 << DB;
  __shared__ uint32_t h[RB];
  for (int i = threadIdx.x; i < RB; i += blockDim.x) h[i] = 0;
  __syncthreads();
  int64_t n = (int64_t)*n_ptr;
  if (n > ncap) n = ncap;
  int64_t begin, end;
  radix_chunk<IT>(g_, n, begin, end);
  for (int64_t i = begin + threadIdx.x; i < end; i += blockDim.x)
    atomicAdd(&h[radix_digit<K, NW, FLIP, DB>(keys + i * kstride, word, shift)], 1u);
  __syncthreads();
  for (int d = threadIdx.x; d < RB; d += blockDim.x) hist[(size_t)d * g_.nb + g_.b] = h[d];
}
template <typename K, int NW, bool FLIP, int IT, int DB>
__global__ void __launch_bounds__(kBlock) k_radix_hist(const K* keys, int kstride, const unsigned long long* n_ptr, int64_t ncap, int word, int shift, uint32_t* hist) {
  k_radix_hist_body<K, NW, FLIP, IT, DB>(hw_grid(), keys, kstride, n_ptr, ncap, word, shift, hist);
}

// sort key of each spawned record's determinant: its colex rank
// (binom[p*(nel+1)+i] = C(p, i)) times kRankMul, fused with the histogram of
// the digit at shift
template <int W, int IT, int DB>
__device__ __forceinline__ void k_rank_hist_body(const Grid& g_, const uint64_t* keys, const unsigned long long* n_ptr, int64_t ncap, int rstride, const uint32_t* binom, int nel1, uint32_t* rank, uint32_t* hist, int shift) {
  constexpr int RB = 1 << DB;
  __shared__ uint32_t h[RB];
  for (int i = threadIdx.x; i < RB; i += blockDim.x) h[i] = 0;
  __syncthreads();
  int64_t n = (int64_t)*n_ptr;
  if (n > ncap) n = ncap;
  int64_t begin, end;
  radix_chunk<IT>(g_, n, begin, end);
  for (int64_t i = begin + threadIdx.x; i < end; i += blockDim.x) {
    uint32_t r = colex_rank<W>(binom, nel1, keys + i * rstride) * kRankMul;
    rank[i] = r;
    atomicAdd(&h[(r >> shift) & (RB - 1u)], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < RB; d += blockDim.x) hist[(size_t)d * g_.nb + g_.b] = h[d];
}
template <int W, int IT, int DB>
__global__ void __launch_bounds__(kBlock) k_rank_hist(const uint64_t* keys, const unsigned long long* n_ptr, int64_t ncap, int rstride, const uint32_t* binom, int nel1, uint32_t* rank, uint32_t* hist, int shift) {
  k_rank_hist_body<W, IT, DB>(hw_grid(), keys, n_ptr, ncap, rstride, binom, nel1, rank, hist, shift);
}

// the 64-bit colex-rank sort of one-word determinants (C(nbasis, nel) <
// 2^64, the rank shorter than the word): each record's key, its colex rank
// (binom[p*nel1 + i] = C(p, i) in 64 bits, staged in LDS) times an odd
// constant modulo 2^bits (a bijection of bits-bit values that spreads the
// ranks, which cluster, evenly over the top byte), written to rank[i] and
// fused with the histogram of the top byte (the digit at shift = bits - 8).
// The partition pass on that byte then leaves 256 segments for k_seg_sort64.
constexpr int kBinom64Lds = 64 * 33;
constexpr uint64_t kRank64Mul = 0x9E3779B97F4A7C15ull;
template <int IT, int DB>
__device__ __forceinline__ void k_rank64_hist_body(const Grid& g_, const uint64_t* rec, int rstride, const unsigned long long* n_ptr, int64_t ncap, const uint64_t* binom, int nbasis, int nel1, int bits, uint64_t* rank, uint32_t* hist) {
  constexpr int RB = 1 << DB;
  __shared__ uint32_t h[RB];
  __shared__ uint64_t bt[kBinom64Lds];
  for (int i = threadIdx.x; i < RB; i += blockDim.x) h[i] = 0;
  for (int i = threadIdx.x; i < nbasis * nel1; i += blockDim.x) bt[i] = binom[i];
  __syncthreads();
  const uint64_t mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
  const int shift = bits - DB;
  int64_t n = (int64_t)*n_ptr;
  if (n > ncap) n = ncap;
  int64_t begin, end;
  radix_chunk<IT>(g_, n, begin, end);
  for (int64_t i = begin + threadIdx.x; i < end; i += blockDim.x) {
    uint64_t x = rec[i * rstride], r = 0;
    for (int c = 1; x; ++c) {
      r += bt[ctz64(x) * nel1 + c];
      x &= x - 1;
    }
    r = (r * kRank64Mul) & mask;
    rank[i] = r;
    atomicAdd(&h[(uint32_t)(r >> shift) & (RB - 1u)], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < RB; d += blockDim.x) hist[(size_t)d * g_.nb + g_.b] = h[d];
}
template <int IT, int DB>
__global__ void __launch_bounds__(kBlock) k_rank64_hist(const uint64_t* rec, int rstride, const unsigned long long* n_ptr, int64_t ncap, const uint64_t* binom, int nbasis, int nel1, int bits, uint64_t* rank, uint32_t* hist) {
  k_rank64_hist_body<IT, DB>(hw_grid(), rec, rstride, n_ptr, ncap, binom, nbasis, nel1, bits, rank, hist);
}

// one block per digit: exclusive scan of that digit's per-block counts
__device__ __forceinline__ void k_radix_scan_body(const Grid& g_, uint32_t* hist, int nb, uint32_t* digit_tot) {
  __shared__ int64_t red[kBlock / kWave];
  uint32_t* row = hist + (size_t)g_.b * nb;
  int per = (nb + kBlock - 1) / kBlock;
  int b0 = threadIdx.x * per;
  int64_t loc = 0;
  for (int i = 0; i < per; ++i)
    if (b0 + i < nb) loc += row[b0 + i];
  int64_t tot;
  int64_t pre = block_exclusive_scan(loc, red, tot);
  for (int i = 0; i < per; ++i)
    if (b0 + i < nb) {
      uint32_t v = row[b0 + i];
      row[b0 + i] = (uint32_t)pre;
      pre += v;
    }
  if (threadIdx.x == 0) digit_tot[g_.b] = (uint32_t)tot;
}
__global__ void __launch_bounds__(kBlock) k_radix_scan(uint32_t* hist, int nb, uint32_t* digit_tot) {
  k_radix_scan_body(hw_grid(), hist, nb, digit_tot);
}

// Stable scatter of one digit, in tiles of IT*kBlock elements: the
// tile is ranked by digit (ballot peers within a wave, LDS counts across waves
// and rounds, in element order), reordered by digit in LDS, then written out
// so that consecutive threads write consecutive positions of each digit's run.
template <typename K, int NW, bool FLIP, int IT, int DB>
__device__ __forceinline__ void k_radix_scatter_body(const Grid& g_, const K* kin, int kstride, const uint32_t* iin, K* kout, uint32_t* iout, const unsigned long long* n_ptr, int64_t ncap, int word, int shift, const uint32_t* hist, const uint32_t* digit_tot) {
  constexpr int RB = 1 << DB;          // digit values
  constexpr int DPT = RB / kBlock;     // digits owned by each thread (consecutive)
  constexpr int TILE = IT * kBlock;
  __shared__ uint32_t run[RB];     // next global position of each digit (this block)
  __shared__ uint32_t cnt[RB];     // digit counts of the tile so far
  __shared__ uint32_t tstart[RB];  // start of each digit inside the reordered tile
  __shared__ uint32_t wcnt[kBlock / kWave][RB];
  __shared__ K lk[TILE * NW];
  __shared__ uint32_t lv[TILE];
  __shared__ int64_t red[kBlock / kWave];
  const int wv = threadIdx.x / kWave;
  const int d0 = threadIdx.x * DPT;
  {
    // digit bases: exclusive scan of the digit totals in digit order
    int64_t loc = 0;
#pragma unroll
    for (int q = 0; q < DPT; ++q) loc += digit_tot[d0 + q];
    int64_t tot;
    int64_t base = block_exclusive_scan(loc, red, tot);
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const int d = d0 + q;
      run[d] = (uint32_t)base + hist[(size_t)d * g_.nb + g_.b];
      base += digit_tot[d];
      cnt[d] = 0;
      for (int w = 0; w < kBlock / kWave; ++w) wcnt[w][d] = 0;
    }
  }
  __syncthreads();
  int64_t n = (int64_t)*n_ptr;
  if (n > ncap) n = ncap;
  int64_t begin, end;
  radix_chunk<IT>(g_, n, begin, end);
  const uint64_t ltmask = lanemask_lt();
  for (int64_t tb = begin; tb < end; tb += TILE) {
    K k[IT][NW];
    uint32_t v[IT], d[IT], off[IT];
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      int64_t e = tb + (int64_t)j * kBlock + threadIdx.x;
      bool valid = e < end;
      d[j] = 0;
      v[j] = 0;
      if (valid) {
#pragma unroll
        for (int w = 0; w < NW; ++w) k[j][w] = kin[e * kstride + w];
        v[j] = iin ? iin[e] : (uint32_t)e;
        d[j] = radix_digit<K, NW, FLIP, DB>(k[j], word, shift);
      }
      // lanes of this wave with the same digit, by one ballot per digit bit
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int bit = 0; bit < DB; ++bit) {
        uint64_t m = __ballot(valid && ((d[j] >> bit) & 1u));
        peers &= ((d[j] >> bit) & 1u) ? m : ~m;
      }
      int rank = popc64(peers & ltmask);
      if (valid && rank == 0) wcnt[wv][d[j]] = (uint32_t)popc64(peers);
      __syncthreads();
      if (valid) {
        uint32_t o = cnt[d[j]] + (uint32_t)rank;
        for (int q = 0; q < wv; ++q) o += wcnt[q][d[j]];
        off[j] = o;
      } else {
        off[j] = 0xFFFFFFFFu;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < DPT; ++q) {
        const int dd = d0 + q;
        uint32_t sum = 0;
        for (int w = 0; w < kBlock / kWave; ++w) { sum += wcnt[w][dd]; wcnt[w][dd] = 0; }
        cnt[dd] += sum;
      }
      __syncthreads();
    }
    // digit starts inside the tile
    {
      int64_t loc = 0;
#pragma unroll
      for (int q = 0; q < DPT; ++q) loc += cnt[d0 + q];
      int64_t tot;
      int64_t st = block_exclusive_scan(loc, red, tot);
#pragma unroll
      for (int q = 0; q < DPT; ++q) {
        tstart[d0 + q] = (uint32_t)st;
        st += cnt[d0 + q];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      if (off[j] != 0xFFFFFFFFu) {
        uint32_t lp = tstart[d[j]] + off[j];
#pragma unroll
        for (int w = 0; w < NW; ++w) lk[lp * NW + w] = k[j][w];
        lv[lp] = v[j];
      }
    }
    __syncthreads();
    const int tn = (int)((end - tb) < TILE ? (end - tb) : TILE);
    for (int i = threadIdx.x; i < tn; i += blockDim.x) {
      K kk[NW];
#pragma unroll
      for (int w = 0; w < NW; ++w) kk[w] = lk[i * NW + w];
      uint32_t dd = radix_digit<K, NW, FLIP, DB>(kk, word, shift);
      uint32_t gp = run[dd] + (uint32_t)i - tstart[dd];
#pragma unroll
      for (int w = 0; w < NW; ++w) kout[(size_t)gp * NW + w] = kk[w];
      iout[gp] = lv[i];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      run[d0 + q] += cnt[d0 + q];
      cnt[d0 + q] = 0;
    }
    __syncthreads();
  }
}
template <typename K, int NW, bool FLIP, int IT, int DB>
__global__ void __launch_bounds__(kBlock) k_radix_scatter(const K* kin, int kstride, const uint32_t* iin, K* kout, uint32_t* iout, const unsigned long long* n_ptr, int64_t ncap, int word, int shift, const uint32_t* hist, const uint32_t* digit_tot) {
  k_radix_scatter_body<K, NW, FLIP, IT, DB>(hw_grid(), kin, kstride, iin, kout, iout, n_ptr, ncap, word, shift, hist, digit_tot);
}

// ------------------------------------------------------- partition sort
// Sort of 32-bit keys in two launches after one stable radix pass on the
// keys' top byte: that pass leaves segment b (the keys whose top byte is b)
// contiguous and in input order, and block b of k_seg_sort sorts segment b
// by the low 24 bits with three stable 8-bit passes, so the whole list ends
// sorted and stable.  The passes run in LDS for segments of up to kSegCap
// keys and through global memory (the input buffers serve as scratch) for
// larger ones.  The keys are colex ranks times an odd constant (a bijection
// of 32-bit values), so the top byte spreads the keys evenly over segments.
// Both rank paths partition on a 9-bit top digit (512 segments) and sort up
// to 2048 keys of a segment in LDS, so that two (64-bit keys) or three
// (32-bit keys) 1024-thread blocks fit a compute unit.
constexpr int kSegCap = 2048;
constexpr int kSegBits = 9;  // top digit of the rank partitions
constexpr int kSegs = 1 << kSegBits;

constexpr int kSegThreads = 1024;  // threads of a k_seg_sort block
constexpr int kSegWaves = kSegThreads / kWave;

struct SegScratch {
  uint32_t cnt[256];  // digit counts, then each digit's next position
  uint32_t wcnt[kSegWaves][256];
};

// exclusive scan of the 256 digit counts by the first wave (4 per lane)
__device__ inline void seg_scan_counts(SegScratch* sc) {
  if (threadIdx.x < kWave) {
    const int l = threadIdx.x;
    uint32_t c0 = sc->cnt[4 * l], c1 = sc->cnt[4 * l + 1], c2 = sc->cnt[4 * l + 2], c3 = sc->cnt[4 * l + 3];
    uint32_t s = c0 + c1 + c2 + c3, x = s;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      uint32_t y = __shfl_up(x, o);
      if (l >= o) x += y;
    }
    uint32_t pre = x - s;
    sc->cnt[4 * l] = pre;
    sc->cnt[4 * l + 1] = pre + c0;
    sc->cnt[4 * l + 2] = pre + c0 + c1;
    sc->cnt[4 * l + 3] = pre + c0 + c1 + c2;
  }
}

// one stable 8-bit pass, src -> dst (LDS or global), in tiles of one key per
// thread: ballots rank the keys of a digit within each wave, the waves'
// counts in LDS order the waves.  Every thread of the block calls it.
template <typename K>
__device__ inline void seg_pass(const K* sk, const uint32_t* si, K* dk, uint32_t* di, int64_t n, int shift,
                                SegScratch* sc) {
  const int wv = threadIdx.x / kWave;
  const uint64_t ltmask = lanemask_lt();
  for (int t = threadIdx.x; t < 256; t += blockDim.x) sc->cnt[t] = 0;
  for (int t = threadIdx.x; t < kSegWaves * 256; t += blockDim.x) (&sc->wcnt[0][0])[t] = 0;
  __syncthreads();
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(&sc->cnt[(uint32_t)(sk[i] >> shift) & 255u], 1u);
  __syncthreads();
  seg_scan_counts(sc);
  __syncthreads();
  for (int64_t t0 = 0; t0 < n; t0 += blockDim.x) {
    const int64_t i = t0 + threadIdx.x;
    const bool valid = i < n;
    const K k = valid ? sk[i] : (K)0;
    const uint32_t v = valid ? si[i] : 0u;
    const uint32_t d = (uint32_t)(k >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      uint64_t m = __ballot(valid && ((d >> bit) & 1u));
      peers &= ((d >> bit) & 1u) ? m : ~m;
    }
    const int rank = popc64(peers & ltmask);
    if (valid && rank == 0) sc->wcnt[wv][d] = (uint32_t)popc64(peers);
    __syncthreads();
    if (valid) {
      uint32_t o = sc->cnt[d] + (uint32_t)rank;
      for (int q = 0; q < wv; ++q) o += sc->wcnt[q][d];
      dk[o] = k;
      di[o] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 256; t += blockDim.x) {
      uint32_t add = 0;
      for (int q = 0; q < kSegWaves; ++q) {
        add += sc->wcnt[q][t];
        sc->wcnt[q][t] = 0;
      }
      sc->cnt[t] += add;
    }
    __syncthreads();
  }
}

// block b sorts segment b (one block of kSegThreads per segment) by the npass
// bytes below the top digit; digit_tot holds the segment sizes, the counts of
// the top-digit pass.  Keys of 32 bits (colex rank times an odd constant,
// three passes) or 64 bits (the 64-bit rank path).
template <typename K>
__device__ __forceinline__ void k_seg_sort_body(const Grid& g_, K* kin, uint32_t* iin, K* kout, uint32_t* iout, const uint32_t* digit_tot, int npass) {
  constexpr int CAP = kSegCap;
  __shared__ K ka[CAP], kb[CAP];
  __shared__ uint32_t ia[CAP], ib[CAP];
  __shared__ SegScratch sc;
  __shared__ int64_t seg[2];
  __shared__ int64_t wsum[kSegWaves];
  {  // this segment's start: the sizes of the segments before it
    int64_t v = 0;
    for (int b = threadIdx.x; b < (int)g_.b; b += blockDim.x) v += digit_tot[b];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane_id() == 0) wsum[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t st = 0;
      for (int q = 0; q < kSegWaves; ++q) st += wsum[q];
      seg[0] = st;
      seg[1] = digit_tot[g_.b];
    }
    __syncthreads();
  }
  const int64_t start = seg[0], n = seg[1];
  if (n == 0) return;  // block-uniform
  K *gk = kin + start, *ok = kout + start;
  uint32_t *gi = iin + start, *oi = iout + start;
  if (n <= CAP) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
      ka[i] = gk[i];
      ia[i] = gi[i];
    }
    __syncthreads();
    K *sk = ka, *dk = kb;
    uint32_t *si = ia, *di = ib;
    for (int p = 0; p < npass; ++p) {
      seg_pass<K>(sk, si, dk, di, n, 8 * p, &sc);
      K* tk = sk; sk = dk; dk = tk;
      uint32_t* ti = si; si = di; di = ti;
    }
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
      ok[i] = sk[i];
      oi[i] = si[i];
    }
  } else {  // through global memory, the input serving as scratch
    K *sk = gk, *dk = ok;
    uint32_t *si = gi, *di = oi;
    for (int p = 0; p < npass; ++p) {
      seg_pass<K>(sk, si, dk, di, n, 8 * p, &sc);
      K* tk = sk; sk = dk; dk = tk;
      uint32_t* ti = si; si = di; di = ti;
    }
    if (sk != ok) {  // an even number of passes ended in the input
      for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        ok[i] = sk[i];
        oi[i] = si[i];
      }
    }
  }
}
__global__ void __launch_bounds__(kSegThreads) k_seg_sort(uint32_t* kin, uint32_t* iin, uint32_t* kout, uint32_t* iout, const uint32_t* digit_tot) {
  k_seg_sort_body<uint32_t>(hw_grid(), kin, iin, kout, iout, digit_tot, (32 - kSegBits + 7) / 8);
}
__global__ void __launch_bounds__(kSegThreads) k_seg_sort64(uint64_t* kin, uint32_t* iin, uint64_t* kout, uint32_t* iout, const uint32_t* digit_tot, int npass) {
  k_seg_sort_body<uint64_t>(hw_grid(), kin, iin, kout, iout, digit_tot, npass);
}

// ------------------------------------------------------------- merging
// LDS stage of a determinant list (key, S signs, flag): the merged spawns and
// the new determinants
template <int W, int R>
struct ListStage {
  static constexpr int kCap = R == 1 ? 768 : 256;
  Stage st;
  uint64_t key[kCap * W];
  double sgn[kCap * R];
  uint32_t flag[kCap];
};

template <int W, int R>
__device__ inline int list_push(ListStage<W, R>* ls, bool pred, const uint64_t* key, const double* sg,
                                uint32_t flag, int S) {
  int i = stage_push(pred, &ls->st);
  if (pred) {
#pragma unroll
    for (int w = 0; w < W; ++w) ls->key[i * W + w] = key[w];
    for (int r = 0; r < S; ++r) ls->sgn[i * R + r] = sg[r];
    ls->flag[i] = flag;
  }
  return i;
}

template <int W, int R>
__device__ inline void list_flush(ListStage<W, R>* ls, unsigned long long* gctr, uint64_t* okey, double* osgn,
                                  uint32_t* oflag, int S, bool force) {
  uint32_t n;
  unsigned long long base;
  if (!stage_flush_begin(&ls->st, gctr, ListStage<W, R>::kCap, force, n, base)) return;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    unsigned long long d = base + i;
#pragma unroll
    for (int w = 0; w < W; ++w) okey[d * W + w] = ls->key[i * W + w];
    for (int r = 0; r < S; ++r) osgn[d * S + r] = ls->sgn[i * R + r];
    oflag[d] = ls->flag[i];
  }
  stage_flush_end(&ls->st);
}

// CompressSpawnedList (Annihilation.F90:249-515) + FindResidualParticle
// (:551-634): the head of each run of equal keys sums the run in sorted
// (= emission, the sort is stable) order.
template <typename K, int NW>
__device__ inline bool key_eq(const K* a, const K* b) {
  bool e = true;
#pragma unroll
  for (int w = 0; w < NW; ++w) e = e && (a[w] == b[w]);
  return e;
}

// The merged (compressed) spawned list alone, for the explicit-list entry
// point: run heads sum their runs as k_merge_annihilate does, and the list
// goes to memory; no statistics (k_merge_annihilate counts them).
template <int W, int R, typename K, int NW>
__device__ __forceinline__ void k_merge_list_body(const Grid& g_, const DevSys& s, Counters* cnt, const K* skey, const uint32_t* sidx, const uint64_t* sp_rec, int64_t scap, uint64_t* u_key, double* u_sgn, uint32_t* u_flag, unsigned long long* u_count) {
  if (cnt->halt) return;
  __shared__ ListStage<W, R> ls;
  stage_init(&ls.st);
  const int S = R == 1 ? 1 : s.S;
  int64_t n = (int64_t)cnt->n_spawned;
  if (n > scap) n = scap;
  int64_t trips = grid_trips(g_, n);
  for (int64_t t = 0; t < trips; ++t) {
    int64_t i = grid_elem(g_, t);
    bool keep = false;
    double sum[R];
    uint32_t flag = 0;
    uint64_t key[W];
    if (i < n) {
      const K* ki = skey + i * NW;
      bool head = true;
      if (i > 0) head = !key_eq<K, NW>(ki, skey + (i - 1) * NW);
      if (head) {
        int64_t j = i + 1;
        while (j < n && key_eq<K, NW>(ki, skey + j * NW)) ++j;
        const uint64_t* kd = sp_rec + (size_t)sidx[i] * s.rec_words;
#pragma unroll
        for (int w = 0; w < W; ++w) key[w] = kd[w];
        for (int r = 0; r < S; ++r) sum[r] = 0.0;
        for (int64_t q = i; q < j; ++q) {
          const uint64_t* rq = sp_rec + (size_t)sidx[q] * s.rec_words;
          uint32_t inf = rec_info<W>(s, rq);
          int r = (int)(inf & 0xFFu);
          double nv = rec_val<W>(s, rq);
          double cv = sum[r];
          if (j == i + 1) {
            if (inf & kInfoInit) flag |= 1u << r;
          } else if (s.t_init) {
            if ((fabs(cv) > 1.e-12 && fabs(nv) > 1.e-12) || (inf & kInfoInit)) flag |= 1u << r;
          }
          sum[r] = cv + nv;
        }
        if (j == i + 1) {
          keep = !(fabs(sum[rec_info<W>(s, kd) & 0xFFu]) < 1.e-12);
        } else {
          double tot = 0.0;
          for (int r = 0; r < S; ++r) tot += fabs(sum[r]);
          keep = tot > 1.e-12;
        }
      }
    }
    list_push<W, R>(&ls, keep, key, sum, flag, S);
    list_flush<W, R>(&ls, u_count, u_key, u_sgn, u_flag, S, false);
  }
  list_flush<W, R>(&ls, u_count, u_key, u_sgn, u_flag, S, true);
}
template <int W, int R, typename K, int NW>
__global__ void __launch_bounds__(kBlock) k_merge_list(DevSys s, Counters* cnt, const K* skey, const uint32_t* sidx, const uint64_t* sp_rec, int64_t scap, uint64_t* u_key, double* u_sgn, uint32_t* u_flag, unsigned long long* u_count) {
  k_merge_list_body<W, R, K, NW>(hw_grid(), s, cnt, skey, sidx, sp_rec, scap, u_key, u_sgn, u_flag, u_count);
}

// CompressSpawnedList (Annihilation.F90:249-515) + FindResidualParticle
// (:551-634), then AnnihilateSpawnedParts (:965-1352), in one pass over the
// sorted spawns: the head of each run of equal keys sums the run in sorted
// (= emission, the sort is stable) order, then looks its determinant up in the
// walker hash; found determinants are updated in place (each key is owned by
// one lane), absent ones go to the new list.  The lookups of one wave overlap
// the record gathers of the others, and the merged list never goes to memory.
template <int W, int R, typename K, int NW>
__device__ __forceinline__ void k_merge_annihilate_body(const Grid& g_, const DevSys& s, const Walk& wk, Counters* cnt, double* part, const K* skey, const uint32_t* sidx, const uint64_t* sp_rec, int64_t scap, uint64_t* n_key, double* n_sgn, uint32_t* n_flag, const IterParams* prm) {
  if (cnt->halt) return;
  const uint32_t iter = prm->iter;
  __shared__ double acc_lds[kStatStride];
  __shared__ ListStage<W, R> ls;
  __shared__ FreeStage fs;
  __shared__ unsigned int blk_unique;
  BlockStats bs{acc_lds};
  bs.zero();
  stage_init(&ls.st);
  stage_init(&fs.st);
  if (threadIdx.x == 0) blk_unique = 0;
  const int S = R == 1 ? 1 : s.S;
  int64_t n = (int64_t)cnt->n_spawned;
  if (n > scap) n = scap;
  int64_t trips = grid_trips(g_, n);
  double ann[R], st_ab[R], st_an[R], st_rm[R], st_bn[R];
  for (int r = 0; r < S; ++r) ann[r] = st_ab[r] = st_an[r] = st_rm[r] = st_bn[r] = 0.0;
  for (int64_t t = 0; t < trips; ++t) {
    int64_t i = grid_elem(g_, t);
    bool keep = false;
    double ss[R];
    uint32_t sflag = 0;
    uint64_t key[W];
    if (i < n) {
      const K* ki = skey + i * NW;
      bool head = true;
      if (i > 0) head = !key_eq<K, NW>(ki, skey + (i - 1) * NW);
      if (head) {
        int64_t j = i + 1;
        while (j < n && key_eq<K, NW>(ki, skey + j * NW)) ++j;
        const uint64_t* kd = sp_rec + (size_t)sidx[i] * s.rec_words;
#pragma unroll
        for (int w = 0; w < W; ++w) key[w] = kd[w];
        for (int r = 0; r < S; ++r) ss[r] = 0.0;
        if (j == i + 1) {
          uint32_t inf = rec_info<W>(s, kd);
          int r = (int)(inf & 0xFFu);
          ss[r] = rec_val<W>(s, kd);
          if (inf & kInfoInit) sflag |= 1u << r;
          keep = !(fabs(ss[r]) < 1.e-12);
        } else {
          for (int64_t q = i; q < j; ++q) {
            const uint64_t* rq = sp_rec + (size_t)sidx[q] * s.rec_words;
            uint32_t inf = rec_info<W>(s, rq);
            int r = (int)(inf & 0xFFu);
            double nv = rec_val<W>(s, rq);
            double cv = ss[r];
            if (s.t_init) {
              if ((fabs(cv) > 1.e-12 && fabs(nv) > 1.e-12) || (inf & kInfoInit)) sflag |= 1u << r;
            }
            if (cv * nv < 0.0) {
              double x = fabs(cv), y = fabs(nv);
              ann[r] += 2.0 * (x < y ? x : y);
            }
            ss[r] = cv + nv;
          }
          double tot = 0.0;
          for (int r = 0; r < S; ++r) tot += fabs(ss[r]);
          keep = tot > 1.e-12;
        }
      }
    }
    unsigned int nk = (unsigned int)wave_count(keep);
    if (nk && lane_id() == 0) atomicAdd(&blk_unique, nk);
    bool isnew = false, rm = false;
    int64_t pos = -1;
    if (keep) {
      int64_t home = find_walker_hash<W>(key, s.rh, s.L) - 1;
      pos = hash_lookup<W>(wk, key, home, cnt);
      if (pos >= 0) {
        double cs[R];
        double tot = 0.0;
        for (int r = 0; r < S; ++r) { cs[r] = wk.sgn(pos, r); tot += fabs(cs[r]); }
        bool determ = (wk.flags(pos) & kFlagCore) != 0;
        if (tot >= 1.e-12 || determ) {
          for (int r = 0; r < S; ++r) {
            double prod = cs[r] * ss[r];
            if (fabs(cs[r]) < 1.0e-12 && s.t_init && !((sflag >> r) & 1u) && !determ) {
              st_ab[r] += fabs(ss[r]);
              ss[r] = 0.0;
            }
            if (prod < 0.0) {
              double x = fabs(cs[r]), y = fabs(ss[r]);
              st_an[r] += 2.0 * (x < y ? x : y);
            }
          }
          bool occ = false;
          for (int r = 0; r < S; ++r) {
            cs[r] = ss[r] + cs[r];
            wk.sgn(pos, r) = cs[r];
            occ = occ || !(fabs(cs[r]) < 1.e-12);
          }
          rm = !determ && !occ;
        }
      } else {
        if (s.t_init)
          for (int r = 0; r < S; ++r)
            if (!((sflag >> r) & 1u)) {
              st_ab[r] += fabs(ss[r]);
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
                st_rm[r] += fabs(ss[r]);
                ss[r] = 0.0;
              } else {
                st_bn[r] += thr - fabs(ss[r]);
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
    remove_dets<W>(wk, &fs, rm, pos);
    list_push<W, R>(&ls, isnew, key, ss, sflag & 0xFFFFu, S);
    free_flush(wk, cnt, &fs, false);
    list_flush<W, R>(&ls, &cnt->n_new, n_key, n_sgn, n_flag, S, false);
  }
  free_flush(wk, cnt, &fs, true);
  list_flush<W, R>(&ls, &cnt->n_new, n_key, n_sgn, n_flag, S, true);
  for (int r = 0; r < S; ++r) {
    bs.add(soff(F_NABORTED, r), st_ab[r]);
    bs.add(soff(F_NANNIHIL, r), st_an[r] + ann[r]);
    bs.add(soff(F_NREMOVED, r), st_rm[r]);
    bs.add(soff(F_NBORN, r), st_bn[r]);
  }
  __syncthreads();
  if (threadIdx.x == 0 && blk_unique)
    atomicAdd(&cnt->uniq_part[g_.b % Counters::kSpread].v, (unsigned long long)blk_unique);
  bs.flush(part, g_.b);
}
template <int W, int R, typename K, int NW>
__global__ void __launch_bounds__(kBlock) k_merge_annihilate(DevSys s, Walk wk, Counters* cnt, double* part, const K* skey, const uint32_t* sidx, const uint64_t* sp_rec, int64_t scap, uint64_t* n_key, double* n_sgn, uint32_t* n_flag, const IterParams* prm) {
  k_merge_annihilate_body<W, R, K, NW>(hw_grid(), s, wk, cnt, part, skey, sidx, sp_rec, scap, n_key, n_sgn, n_flag, prm);
}

// ------------------------------------------------------ semistochastic core
// Core amplitudes before the projection (no death acts on core dets, so the
// signs after the prepass are the iteration's starting amplitudes):
// x[i * xs + r] = the amplitude of core det i, replica r (xs = S for an
// engine's own vector; a wider stride for a vector shared by a batch group)
template <int R>
__device__ __forceinline__ void k_core_gather_body(const Grid& g_, const int64_t* core_slot, const Walk& wk, int S, int64_t ncore, double* x, int xs) {
  for (int64_t i = (int64_t)g_.b * blockDim.x + threadIdx.x; i < ncore; i += (int64_t)g_.nb * blockDim.x)
    for (int r = 0; r < S; ++r) x[i * xs + r] = wk.sgn(core_slot[i], r);
}
template <int R>
__global__ void __launch_bounds__(kBlock) k_core_gather(const int64_t* core_slot, Walk wk, int S, int64_t ncore, double* x) {
  k_core_gather_body<R>(hw_grid(), core_slot, wk, S, ncore, x, S);
}

// determ_projection (semi_stoch_procs.F90:105-241), one wave per core row:
// lane l sums the row's nonzeros q = l, l+64, ... as acc = acc - H*x, the
// butterfly acc = acc + shfl_xor(acc, o) for o = 32..1 combines the lanes, and
// y = (acc + S*x_i) * tau.  The CPU restatement sums in exactly this order.
template <int R, typename VT>
__device__ __forceinline__ void k_core_spmv_body(const Grid& g_, const int64_t* rowptr, const int32_t* cols, const VT* vals, const double* x, double* y, int64_t ncore, int S, const IterParams* prm) {
  const double tau = prm->tau;
  const double* sftv = prm->sft.v;
  const int lane = lane_id();
  const int64_t nwaves = (int64_t)g_.nb * blockDim.x / kWave;
  for (int64_t i = ((int64_t)g_.b * blockDim.x + threadIdx.x) / kWave; i < ncore; i += nwaves) {
    double acc[R];
    for (int r = 0; r < S; ++r) acc[r] = 0.0;
    const int64_t e = rowptr[i + 1];
    for (int64_t q = rowptr[i] + lane; q < e; q += kWave) {
      double v = (double)vals[q];
      const double* xc = x + (size_t)cols[q] * S;
      for (int r = 0; r < S; ++r) acc[r] = acc[r] - v * xc[r];
    }
    for (int r = 0; r < S; ++r) {
#pragma unroll
      for (int o = kWave / 2; o > 0; o >>= 1) acc[r] = acc[r] + __shfl_xor(acc[r], o);
    }
    if (lane == 0)
      for (int r = 0; r < S; ++r) y[i * S + r] = (acc[r] + sftv[r] * x[i * S + r]) * tau;
  }
}
template <int R, typename VT>
__global__ void __launch_bounds__(kBlock) k_core_spmv(const int64_t* rowptr, const int32_t* cols, const VT* vals, const double* x, double* y, int64_t ncore, int S, const IterParams* prm) {
  k_core_spmv_body<R, VT>(hw_grid(), rowptr, cols, vals, x, y, ncore, S, prm);
}

// deterministic_annihilation (Annihilation.F90:930-963): the projected
// amplitudes join the core dets' signs (births and annihilations counted)
template <int R>
__device__ __forceinline__ void k_core_apply_body(const Grid& g_, const DevSys& s, const Walk& wk, const Counters* cnt, const int64_t* core_slot, const double* y, int64_t ncore, double* part, int ys) {
  if (cnt->halt) return;
  __shared__ double acc_lds[kStatStride];
  BlockStats bs{acc_lds};
  bs.zero();
  const int S = R == 1 ? 1 : s.S;
  LaneSum<R> born, ann;
  int64_t trips = grid_trips(g_, ncore);
  for (int64_t t = 0; t < trips; ++t) {
    int64_t i = grid_elem(g_, t);
    bool act = i < ncore;
    int64_t slot = act ? core_slot[i] : 0;
    for (int r = 0; r < S; ++r) {
      double b = 0.0, a = 0.0;
      if (act) {
        double yy = y[i * ys + r], c = wk.sgn(slot, r);
        wk.sgn(slot, r) = yy + c;
        b = fabs(yy);
        if (c * yy < 0.0) {
          double x1 = fabs(c), x2 = fabs(yy);
          a = 2.0 * (x1 < x2 ? x1 : x2);
        }
      }
      born.add(bs, soff(F_NBORN, r), b);
      ann.add(bs, soff(F_NANNIHIL, r), a);
    }
  }
  born.done(bs, soff(F_NBORN, 0));
  ann.done(bs, soff(F_NANNIHIL, 0));
  bs.flush(part, g_.b);
}
template <int R>
__global__ void __launch_bounds__(kBlock) k_core_apply(DevSys s, Walk wk, const Counters* cnt, const int64_t* core_slot, const double* y, int64_t ncore, double* part) {
  k_core_apply_body<R>(hw_grid(), s, wk, cnt, core_slot, y, ncore, part, R == 1 ? 1 : s.S);
}

// core Hamiltonian rows (calc_determ_hamil, fast_determ_hamil.F90:892-1548):
// one wave per row i scans every core det j in ascending order; H_ij for
// connected pairs (diagonal H_ii - Hii), off-diagonal zeros dropped.  Pass 0
// counts, pass 1 writes columns and values at rowptr[i].
template <int W, typename VT>
__device__ __forceinline__ void k_core_rows_body(const Grid& g_, const DevSys& s, const uint64_t* keys, int64_t ncore, int pass, int64_t* count, const int64_t* rowptr, int32_t* cols, VT* vals) {
  const int lane = lane_id();
  const int64_t nwaves = (int64_t)g_.nb * blockDim.x / kWave;
  for (int64_t i = ((int64_t)g_.b * blockDim.x + threadIdx.x) / kWave; i < ncore; i += nwaves) {
    const uint64_t* di = keys + i * W;
    int64_t base = pass ? rowptr[i] : 0;
    for (int64_t j0 = 0; j0 < ncore; j0 += kWave) {
      int64_t j = j0 + lane;
      bool keep = false;
      double h = 0.0;
      if (j < ncore) {
        const uint64_t* dj = keys + j * W;
        if (excit_level<W>(di, dj) <= 2) {
          h = (i == j) ? diag_helement<W>(s, di) - s.hii : helement_pair<W>(s, di, dj);
          keep = (i == j) || h != 0.0;
        }
      }
      uint64_t m = __ballot(keep);
      if (pass && keep) {
        int64_t q = base + popc64(m & lanemask_lt());
        cols[q] = (int32_t)j;
        vals[q] = (VT)h;
      }
      base += popc64(m);
    }
    if (!pass && lane == 0) count[i] = base;
  }
}
template <int W, typename VT>
__global__ void __launch_bounds__(kBlock) k_core_rows(DevSys s, const uint64_t* keys, int64_t ncore, int pass, int64_t* count, const int64_t* rowptr, int32_t* cols, VT* vals) {
  k_core_rows_body<W, VT>(hw_grid(), s, keys, ncore, pass, count, rowptr, cols, vals);
}

// place the core dets: resident ones get the core (and initiator) flags and
// their slot; absent ones are appended to the new list with zero signs
template <int W>
__device__ __forceinline__ void k_core_place_body(const Grid& g_, const DevSys& s, const Walk& wk, Counters* cnt, const uint64_t* keys, int64_t ncore, uint32_t cflags, int64_t* core_slot, uint64_t* n_key, double* n_sgn, uint32_t* n_flag) {
  const int S = s.S;
  for (int64_t i = (int64_t)g_.b * blockDim.x + threadIdx.x; i < ncore; i += (int64_t)g_.nb * blockDim.x) {
    const uint64_t* d = keys + i * W;
    int64_t home = find_walker_hash<W>(d, s.rh, s.L) - 1;
    int64_t pos = hash_lookup<W>(wk, d, home, cnt);
    if (pos >= 0) {
      wk.flags(pos) |= cflags;
      core_slot[i] = pos;
    } else {
      unsigned long long q = atomicAdd(&cnt->n_new, 1ull);
#pragma unroll
      for (int w = 0; w < W; ++w) n_key[q * W + w] = d[w];
      for (int r = 0; r < S; ++r) n_sgn[q * S + r] = 0.0;
      n_flag[q] = cflags;
      core_slot[i] = -1;
    }
  }
}
template <int W>
__global__ void __launch_bounds__(kBlock) k_core_place(DevSys s, Walk wk, Counters* cnt, const uint64_t* keys, int64_t ncore, uint32_t cflags, int64_t* core_slot, uint64_t* n_key, double* n_sgn, uint32_t* n_flag) {
  k_core_place_body<W>(hw_grid(), s, wk, cnt, keys, ncore, cflags, core_slot, n_key, n_sgn, n_flag);
}

// AddNewHashDet (load_balancer.fpp:514-629) for the new list
template <int W, int R>
__device__ __forceinline__ void k_insert_body(const Grid& g_, const DevSys& s, const Walk& wk, Counters* cnt, const uint64_t* n_key, const double* n_sgn, const uint32_t* n_flag, const unsigned long long* n_ptr) {
  if (cnt->halt) return;
  const int S = R == 1 ? 1 : s.S;
  int64_t nn = (int64_t)*n_ptr;
  // new det i takes the i-th queued free slot, past the queue a fresh slot
  // at the end of the table: positions follow from i, no atomics
  const unsigned long long head = cnt->free_head;
  const unsigned long long tail = cnt->free_tail;
  const int64_t avail = (int64_t)(tail - head);
  const int64_t ns0 = (int64_t)cnt->nslots;
  if (g_.b == 0 && threadIdx.x == 0) cnt->free_tail_ins = tail;  // for the slot bookkeeping
  // MaxWalkers would be passed (AddNewHashDet's stop, load_balancer.fpp:540-544):
  // no new determinant is placed, so the walker list the halted engine keeps
  // does not depend on the order of the new list
  if (ns0 + (nn > avail ? nn - avail : 0) > wk.cap) {
    if (g_.b == 0 && threadIdx.x == 0) {
      atomicOr(&cnt->err, 1u << NGPU_ERR_WALKERS_FULL);
      cnt->n_fail = (unsigned long long)nn;
      cnt->ins_refused = 1u;
    }
    return;
  }
  unsigned long long my_reused = 0;
  for (int64_t i = (int64_t)g_.b * blockDim.x + threadIdx.x; i < nn;
       i += (int64_t)g_.nb * blockDim.x) {
    int64_t slot = i < avail ? wk.freeq[(head + (unsigned long long)i) % (unsigned long long)wk.cap]
                             : ns0 + (i - avail);
    if (slot >= wk.cap) {
      atomicOr(&cnt->err, 1u << NGPU_ERR_WALKERS_FULL);
      atomicAdd(&cnt->n_fail, 1ull);
      continue;
    }
    uint64_t key[W];
#pragma unroll
    for (int w = 0; w < W; ++w) key[w] = n_key[i * W + w];
    int64_t home = find_walker_hash<W>(key, s.rh, s.L) - 1;
    bool reused = false;
    int64_t hp = hash_insert<W>(wk, key, home, (int32_t)slot, reused);
    my_reused += reused ? 1 : 0;
    if (hp < 0) {
      atomicOr(&cnt->err, 1u << NGPU_ERR_HASH_FULL);
      atomicAdd(&cnt->n_fail, 1ull);
      wk.flags(slot) = kFlagRemoved;
    } else {
#pragma unroll
      for (int w = 0; w < W; ++w) wk.keys[slot * W + w] = key[w];
      for (int r = 0; r < S; ++r) wk.sgn(slot, r) = n_sgn[i * S + r];
      wk.flags(slot) = n_flag[i] & ~kFlagRemoved;
      wk.diagH[slot] = diag_helement<W>(s, key) - s.hii;
      wk.offdiagH[slot] = offdiag_helement<W>(s, key);
      wk.hpos(slot) = hp;
    }
  }
  // one no-return atomic per wave that reused entries
  // one add per block, spread over kSpread lines
  __shared__ unsigned int blk_reused;
  if (threadIdx.x == 0) blk_reused = 0;
  __syncthreads();
  unsigned int c = (unsigned int)wave_sum((double)my_reused);
  if (c && lane_id() == 0) atomicAdd(&blk_reused, c);
  __syncthreads();
  if (threadIdx.x == 0 && blk_reused)
    atomicAdd(&cnt->reuse_part[g_.b % Counters::kSpread].v, (unsigned long long)blk_reused);
}
template <int W, int R>
__global__ void __launch_bounds__(kBlock) k_insert(DevSys s, Walk wk, Counters* cnt, const uint64_t* n_key, const double* n_sgn, const uint32_t* n_flag, const unsigned long long* n_ptr) {
  k_insert_body<W, R>(hw_grid(), s, wk, cnt, n_key, n_sgn, n_flag, n_ptr);
}

// bookkeeping of k_insert: queued slots used, fresh slots appended (beyond
// the capacity refused), live entries added
// slots after k_insert placed nn new determinants: the first ones took the
// queued free slots, the rest fresh slots at the end of the table.  A pure
// function of counters that nothing changes between k_insert and the
// bookkeeping, so k_stats evaluates it for its loop bound while the
// bookkeeping (k_end_iter, or k_post_insert on the set-up paths) applies it.
__device__ inline void slots_after_insert(const Counters* cnt, unsigned long long nn, int64_t cap,
                                          unsigned long long& used, unsigned long long& ns) {
  if (cnt->ins_refused) {  // nothing was placed
    used = 0;
    ns = cnt->nslots;
    return;
  }
  unsigned long long avail = cnt->free_tail_ins - cnt->free_head;
  used = nn < avail ? nn : avail;
  ns = cnt->nslots + (nn - used);
  if (ns > (unsigned long long)cap) ns = (unsigned long long)cap;
}
// reused: tombstone entries the inserts took (the sum of reuse_part, which
// the caller has emptied)
__device__ inline void post_insert(Counters* cnt, unsigned long long nn, int64_t cap, unsigned long long reused) {
  unsigned long long used, ns;
  slots_after_insert(cnt, nn, cap, used, ns);
  cnt->free_head += used;
  cnt->nslots = ns;
  cnt->n_live += nn - cnt->n_fail;
  cnt->n_tomb -= reused;
}
__global__ void k_post_insert(Counters* cnt, const unsigned long long* n_ptr, int64_t cap) {
  unsigned long long reused = 0;
  for (int q = 0; q < Counters::kSpread; ++q) {
    reused += cnt->reuse_part[q].v;
    cnt->reuse_part[q].v = 0;
  }
  post_insert(cnt, *n_ptr, cap, reused);
}

// CalcHashTableStats (load_balancer.fpp:646-805)
template <int W, int R>
__device__ __forceinline__ void k_stats_body(const Grid& g_, const DevSys& s, const Walk& wk, Counters* cnt, double* part, const IterParams* prm) {
  if (cnt->halt) return;
  const uint32_t iter = prm->iter;
  __shared__ double acc_lds[kStatStride];
  __shared__ FreeStage fs;
  BlockStats bs{acc_lds};
  bs.zero();
  stage_init(&fs.st);
  const int S = R == 1 ? 1 : s.S;
  unsigned long long used, ns;
  slots_after_insert(cnt, cnt->n_new, wk.cap, used, ns);  // the bookkeeping applies it later
  const int64_t nslots = (int64_t)ns;
  unsigned long long my_occ = 0;
  double tp[R], nq[R], rmv[R], bn[R];
  for (int r = 0; r < S; ++r) tp[r] = nq[r] = rmv[r] = bn[r] = 0.0;
  int64_t trips = grid_trips(g_, nslots);
  for (int64_t t = 0; t < trips; ++t) {
    int64_t slot = grid_elem(g_, t);
    bool rm = false;
    // the slot's flags, signs and determinant in one round trip to memory
    // (lanes past the table load slot 0, which exists when any lane runs)
    const int64_t ls = slot < nslots ? slot : 0;
    const uint32_t f_ld = wk.flags(ls);
    double s_ld[R];
    uint64_t k_ld[W];
    for (int r = 0; r < S; ++r) s_ld[r] = wk.sgn(ls, r);
#pragma unroll
    for (int w = 0; w < W; ++w) k_ld[w] = wk.keys[ls * W + w];
    if (slot < nslots) {
      uint32_t f = f_ld;
      if (!(f & kFlagRemoved)) {
        bool determ = (f & kFlagCore) != 0;
        double sg[R];
        bool occ = false;
        for (int r = 0; r < S; ++r) { sg[r] = s_ld[r]; occ = occ || !(fabs(sg[r]) < 1.e-12); }
        if (occ || determ) {
          uint64_t key[W];
#pragma unroll
          for (int w = 0; w < W; ++w) key[w] = k_ld[w];
          if (s.t_real) {
            double thr = s.occ_thresh;
            bool changed = false;
            for (int r = 0; r < S; ++r) {
              if (!determ && fabs(sg[r]) > 1.e-12 && fabs(sg[r]) < thr) {
                double pRemove = (thr - fabs(sg[r])) / thr;
                double u, dummy;
                draw2<W>(s.seed, key, iter, r, 0, 5, u, dummy);
                if (pRemove > u) {
                  rmv[r] += fabs(sg[r]);
                  sg[r] = 0.0;
                } else {
                  bn[r] += thr - fabs(sg[r]);
                  sg[r] = sg[r] >= 0.0 ? thr : -thr;
                }
                changed = true;
              }
            }
            if (changed) {
              occ = false;
              for (int r = 0; r < S; ++r) { wk.sgn(slot, r) = sg[r]; occ = occ || !(fabs(sg[r]) < 1.e-12); }
              rm = !occ && !determ;
            }
          }
          if (!rm) {
            my_occ++;
            for (int r = 0; r < S; ++r) {
              tp[r] += fabs(sg[r]);
              nq[r] += sg[r] * sg[r];
            }
            if (det_eq<W>(key, s.ref))
              for (int r = 0; r < S; ++r) bs.acc[soff(F_INSTHF, r)] = sg[r];
          }
        }
      }
    }
    remove_dets<W>(wk, &fs, rm, slot);
    free_flush(wk, cnt, &fs, false);
  }
  free_flush(wk, cnt, &fs, true);
  for (int r = 0; r < S; ++r) {
    bs.add(soff(F_TOTPARTS, r), tp[r]);
    bs.add(soff(F_NORMSQ, r), nq[r]);
    bs.add(soff(F_NREMOVED, r), rmv[r]);
    bs.add(soff(F_NBORN, r), bn[r]);
  }
  bs.add(X_NOCC, (double)my_occ);
  bs.flush(part, g_.b);
}
template <int W, int R>
__global__ void __launch_bounds__(kBlock) k_stats(DevSys s, Walk wk, Counters* cnt, double* part, const IterParams* prm) {
  k_stats_body<W, R>(hw_grid(), s, wk, cnt, part, prm);
}

// ---------------------------------------------------------- hash upkeep
// Tombstones left by removals are reclaimed by rebuilding the index in place
// once live + tombstone entries pass 60% of the entries and tombstones alone
// pass 10%; more than 80% live
// entries is a capacity error (MEMORYFACHASH too small), reported instead of
// letting probe sequences grow without bound.
// the table is rebuilt once live + tombstone entries pass this share (and
// tombstones pass 10 %): later rebuilds against longer probe sequences
constexpr unsigned long long kRebuildUsedPct = 60;

// the rebuild decision, a pure function of the counters after the removals
// since the last accounting (nothing changes them between the removing
// kernels and the accounting), so the in-place clear, refill and bookkeeping
// kernels each evaluate it; an iteration's accounting publishes it for the
// next iteration instead
__device__ inline bool hash_rebuild_due(const Counters* cnt, int64_t entries) {
  unsigned long long removed = cnt->free_tail - cnt->free_tail0;  // each left a tombstone
  unsigned long long live = cnt->n_live - removed, tomb = cnt->n_tomb + removed;
  return (live + tomb) * 100ull > (unsigned long long)entries * kRebuildUsedPct &&
         tomb * 10ull > (unsigned long long)entries;
}

// bookkeeping of the removals since the last call and of an in-place rebuild
// by k_hash_clear/k_hash_refill (one thread; set-up paths)
__device__ inline void hash_done(Counters* cnt, int64_t entries) {
  bool rebuilt = hash_rebuild_due(cnt, entries);
  if (rebuilt) cnt->n_rebuilds += 1;
  unsigned long long removed = cnt->free_tail - cnt->free_tail0;
  cnt->n_live -= removed;
  cnt->n_tomb = rebuilt ? 0ull : cnt->n_tomb + removed;
  cnt->free_tail0 = cnt->free_tail;
  cnt->rebuild = 0;
  if (cnt->n_live * 10ull > (unsigned long long)entries * 8ull) cnt->err |= 1u << NGPU_ERR_HASH_FULL;
}

// end of an iteration (one thread): the removals' tombstones are counted, and
// a rebuild that has become due is left to the next iteration's prepass and
// attempt scan, which run before any other access to the table
__device__ inline void hash_account(Counters* cnt, int64_t entries) {
  unsigned long long removed = cnt->free_tail - cnt->free_tail0;
  cnt->n_live -= removed;
  // after a rebuild in this iteration only the removals after the attempt
  // scan's refill left tombstones; the refill zeroed n_tomb
  cnt->n_tomb += cnt->rebuilt ? removed - cnt->rm_pre : removed;
  cnt->rebuilt = 0;
  cnt->free_tail0 = cnt->free_tail;
  cnt->rebuild = hash_rebuild_due(cnt, entries) ? 1u : 0u;
  if (cnt->n_live * 10ull > (unsigned long long)entries * 8ull) cnt->err |= 1u << NGPU_ERR_HASH_FULL;
}

// in-place rebuild for the set-up paths (upload, core placement, explicit
// lists); iterations rebuild inside k_prepass and k_scan_apply instead
template <int W>
__device__ __forceinline__ void k_hash_clear_body(const Grid& g_, const Walk& wk, const Counters* cnt) {
  if (hash_rebuild_due(cnt, wk.L * (int64_t)kBucketEntries)) hash_clear_part<W>(g_, wk);
}
template <int W>
__global__ void __launch_bounds__(kBlock) k_hash_clear(Walk wk, const Counters* cnt) {
  k_hash_clear_body<W>(hw_grid(), wk, cnt);
}

template <int W>
__device__ __forceinline__ void k_hash_refill_body(const Grid& g_, const DevSys& s, const Walk& wk, Counters* cnt) {
  if (hash_rebuild_due(cnt, wk.L * (int64_t)kBucketEntries)) hash_refill_part<W>(g_, s, wk, cnt, (int64_t)cnt->nslots);
}
template <int W>
__global__ void __launch_bounds__(kBlock) k_hash_refill(DevSys s, Walk wk, Counters* cnt) {
  k_hash_refill_body<W>(hw_grid(), s, wk, cnt);
}

__global__ void k_hash_done(Counters* cnt, int64_t entries) { hash_done(cnt, entries); }

}  // namespace ngpu
