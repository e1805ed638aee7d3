// ngpu_batch.h -- batch launches: one launch of a kernel of the iteration
// chain covers every engine of a batch group.
//
// A group is a set of engines whose chains use the same kernel instances
// (words per determinant, replica layout, generator, sort path and its
// passes, spawn-record precision, core precision, launch knobs).  Each engine
// keeps its own walker table, hash, spawn buffers, counters and statistics;
// an EngDev in device memory describes them, and each launch kind gives every
// engine a contiguous share of the launch's blocks.  A block finds its engine
// from a block map (or from its index, for the kinds with a fixed count per
// engine) and runs the same kernel body as a plain launch, on its share of
// the blocks, so every engine's results are those of iterating it alone.
#pragma once
#include "ngpu_kernels.h"

namespace ngpu {

// launch kinds of a batch chain and how each engine's share is sized
enum {
  kLkChunk = 0,  // chunked kernels: prepass, attempt scan, radix histogram and scatter
  kLkGrid,       // grid-stride kernels: spawns, insert, statistics
  kLkMerge,      // merge and annihilation
  kLkCore,       // core gather and apply (one element per thread)
  kLkCoreRow,    // core H.v (one row per wave)
  kLkKinds
};
// kinds with a fixed number of blocks per engine: the radix digit scan and
// the segment sort (one block per 8-bit digit), the end of the iteration (1)
constexpr int kDigitBlocks = 256;

// One engine of a batch group as its kernels see it.
struct EngDev {
  DevSys s;
  Walk wk;
  Counters* cnt;
  double* part;
  int32_t* att_cnt;
  int64_t* att_off;
  uint8_t* pinfo;
  int64_t* blk_part;
  int64_t* blk_rm;
  int64_t* tile_p;
  int64_t tcap;
  uint64_t* sp_rec;
  uint32_t* rk;
  int64_t scap;
  int64_t* sq;
  uint64_t* sk[2];
  uint32_t* si[2];
  uint32_t* hist;
  uint32_t* digit_tot;
  uint64_t* n_key;
  double* n_sgn;
  uint32_t* n_flag;
  IterParams* prm;
  ngpu_iter_out_t* dout;
  ngpu_iter_out_t* dcyc;
  int64_t ncore;
  int64_t* core_slot;
  int64_t* core_rowptr;
  int32_t* core_cols;
  const void* core_vals;
  double* core_x;
  double* core_y;
  int32_t core_col0;     // shared core space: this engine's first column of the group's vectors
  int32_t b0[kLkKinds];  // first block of this engine's share, per launch kind
  int32_t nb[kLkKinds];  // blocks of the share
};

// this block's engine and its share of a kind with a block map
__device__ inline const EngDev& eng_of(const EngDev* __restrict__ ed, const uint16_t* __restrict__ map, int kind,
                                       Grid& g) {
  const EngDev& a = ed[map[blockIdx.x]];
  g.b = (int64_t)blockIdx.x - a.b0[kind];
  g.nb = a.nb[kind];
  return a;
}

template <int W, int R>
__global__ void __launch_bounds__(kBlock) km_prepass(const EngDev* __restrict__ ed, const uint16_t* __restrict__ map) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkChunk, g);
  const IterParams none{};
  k_prepass_body<W, R>(g, a.s, a.wk, a.cnt, a.part, a.att_cnt, a.pinfo, a.blk_part, a.blk_rm, a.prm, none, 0);
}

template <int W>
__global__ void __launch_bounds__(kBlock) km_scan_apply(const EngDev* __restrict__ ed, const uint16_t* __restrict__ map) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkChunk, g);
  k_scan_apply_body<W>(g, a.s, a.wk, a.att_cnt, a.att_off, a.blk_part, a.blk_rm, a.pinfo, a.cnt, a.s.S, a.tile_p,
                       a.tcap);
}

template <int W, int R, int G>
__global__ void __launch_bounds__(kBlock) km_spawn(const EngDev* __restrict__ ed, const uint16_t* __restrict__ map,
                                                   int ranked) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkGrid, g);
  k_spawn_body<W, R, G>(g, a.s, a.wk, a.cnt, a.part, a.att_off, a.tile_p, a.tcap, a.pinfo, a.sp_rec,
                        ranked ? a.rk : nullptr, a.scap, a.sq, a.prm);
}

template <int W, int R, int PL, int G>
__global__ void __launch_bounds__(kBlock) km_spawn_singles(const EngDev* __restrict__ ed,
                                                           const uint16_t* __restrict__ map, int ranked) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkGrid, g);
  k_spawn_singles_body<W, R, PL, G>(g, a.s, a.wk, a.cnt, a.part, a.pinfo, a.sp_rec, ranked ? a.rk : nullptr, a.scap,
                                    a.sq, a.prm);
}

// radix pass p of the word sort: the first pass reads the keys inside the
// spawn records (first = 1) or the 64-bit rank keys in sk[cur] (first = 2,
// the rank path's partition pass), later passes the previous pass's output
// (ping-pong cur)
template <typename K, int NW, bool FLIP, int IT, int DB>
__global__ void __launch_bounds__(kBlock) km_radix_hist_words(const EngDev* __restrict__ ed,
                                                              const uint16_t* __restrict__ map, int first, int cur,
                                                              int word, int shift) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkChunk, g);
  const K* kin = first == 1 ? (const K*)a.sp_rec : (const K*)a.sk[cur];
  const int kstride = first == 1 ? a.s.rec_words : NW;
  k_radix_hist_body<K, NW, FLIP, IT, DB>(g, kin, kstride, &a.cnt->n_spawned, a.scap, word, shift, a.hist);
}

template <typename K, int NW, bool FLIP, int IT, int DB>
__global__ void __launch_bounds__(kBlock) km_radix_scatter_words(const EngDev* __restrict__ ed,
                                                                 const uint16_t* __restrict__ map, int first,
                                                                 int cur, int word, int shift) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkChunk, g);
  const K* kin = first == 1 ? (const K*)a.sp_rec : (const K*)a.sk[cur];
  const int kstride = first == 1 ? a.s.rec_words : NW;
  const uint32_t* iin = first ? nullptr : a.si[cur];
  k_radix_scatter_body<K, NW, FLIP, IT, DB>(g, kin, kstride, iin, (K*)a.sk[cur ^ 1], a.si[cur ^ 1], &a.cnt->n_spawned,
                                            a.scap, word, shift, a.hist, a.digit_tot);
}

// the 64-bit rank path's keys of the spawned list into sk[0] (the partition
// pass's input) with the histogram of their top byte
template <int IT>
__global__ void __launch_bounds__(kBlock) km_rank64_hist(const EngDev* __restrict__ ed,
                                                         const uint16_t* __restrict__ map, int bits) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkChunk, g);
  k_rank64_hist_body<IT, kSegBits>(g, a.sp_rec, a.s.rec_words, &a.cnt->n_spawned, a.scap, a.s.binom64,
                                     a.s.nbasis, a.s.nel1, bits, a.sk[0], a.hist);
}

// the colex-rank sort: top-byte pass over the ranks the spawn kernels wrote
template <int IT>
__global__ void __launch_bounds__(kBlock) km_radix_hist_ranks(const EngDev* __restrict__ ed,
                                                              const uint16_t* __restrict__ map) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkChunk, g);
  k_radix_hist_body<uint32_t, 1, false, IT, kSegBits>(g, a.rk, 1, &a.cnt->n_spawned, a.scap, 0, 32 - kSegBits,
                                                   a.hist);
}
template <int IT>
__global__ void __launch_bounds__(kBlock) km_radix_scatter_ranks(const EngDev* __restrict__ ed,
                                                                 const uint16_t* __restrict__ map) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkChunk, g);
  k_radix_scatter_body<uint32_t, 1, false, IT, kSegBits>(g, a.rk, 1, nullptr, (uint32_t*)a.sk[1], a.si[1],
                                                         &a.cnt->n_spawned, a.scap, 0, 32 - kSegBits, a.hist,
                                                         a.digit_tot);
}

// one block per digit of every engine: engine blockIdx / digits
__global__ void __launch_bounds__(kBlock) km_radix_scan(const EngDev* __restrict__ ed, int digits) {
  const EngDev& a = ed[blockIdx.x / digits];
  const Grid g{(int64_t)(blockIdx.x % digits), digits};
  k_radix_scan_body(g, a.hist, a.nb[kLkChunk], a.digit_tot);
}
__global__ void __launch_bounds__(kSegThreads) km_seg_sort(const EngDev* __restrict__ ed) {
  const EngDev& a = ed[blockIdx.x / kSegs];
  const Grid g{(int64_t)(blockIdx.x % kSegs), kSegs};
  k_seg_sort_body<uint32_t>(g, (uint32_t*)a.sk[1], a.si[1], (uint32_t*)a.sk[0], a.si[0], a.digit_tot,
                            (32 - kSegBits + 7) / 8);
}
// the 64-bit rank path's kSegs segments per engine: sk[1] -> sk[0],
// npass bytes below the top digit
__global__ void __launch_bounds__(kSegThreads) km_seg_sort64(const EngDev* __restrict__ ed, int npass) {
  const EngDev& a = ed[blockIdx.x / kSegs];
  const Grid g{(int64_t)(blockIdx.x % kSegs), kSegs};
  k_seg_sort_body<uint64_t>(g, a.sk[1], a.si[1], a.sk[0], a.si[0], a.digit_tot, npass);
}

template <int W, int R, typename K, int NW>
__global__ void __launch_bounds__(kBlock) km_merge_annihilate(const EngDev* __restrict__ ed,
                                                              const uint16_t* __restrict__ map, int cur) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkMerge, g);
  k_merge_annihilate_body<W, R, K, NW>(g, a.s, a.wk, a.cnt, a.part, (const K*)a.sk[cur], a.si[cur], a.sp_rec, a.scap,
                                       a.n_key, a.n_sgn, a.n_flag, a.prm);
}

template <int R>
__global__ void __launch_bounds__(kBlock) km_core_gather(const EngDev* __restrict__ ed,
                                                         const uint16_t* __restrict__ map) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkCore, g);
  k_core_gather_body<R>(g, a.core_slot, a.wk, a.s.S, a.ncore, a.core_x, a.s.S);
}

// Core space shared by every engine of the group (the same determinants, so
// the same Hamiltonian, as replicas of one system have under a global core
// space): the engines' amplitudes are the C columns of one interleaved vector
// X[i * C + column], the Hamiltonian is read once for all of them, and
// Y[i * C + column] = (H.X + shift * X) * tau per column.
template <int R>
__global__ void __launch_bounds__(kBlock) km_core_gather_shared(const EngDev* __restrict__ ed,
                                                                const uint16_t* __restrict__ map, double* X, int C) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkCore, g);
  k_core_gather_body<R>(g, a.core_slot, a.wk, a.s.S, a.ncore, X + a.core_col0, C);
}

constexpr int kCoreCols = 16;  // columns per pass of the shared projection

// One wave per row; lane l works on column c0 + (l & 15) and the row's
// nonzeros q = q0 + (l >> 4), +4, ..., so each load instruction of the wave
// reads four whole 128-byte rows of X (16 columns) instead of 64 scattered
// elements; the four partial sums of a column meet in two butterfly steps.
// The order of the additions differs from an engine's own projection
// (k_core_spmv_body), so the results agree with it to rounding.
// colmap[c] = engine << 8 | replica of column c.
template <typename VT>
__global__ void __launch_bounds__(kBlock) km_core_spmv_shared(const EngDev* __restrict__ ed, const int32_t* colmap,
                                                              const double* X, double* Y, int C) {
  const EngDev& a = ed[0];
  const int64_t* rowptr = a.core_rowptr;
  const int32_t* cols = a.core_cols;
  const VT* vals = (const VT*)a.core_vals;
  const int64_t ncore = a.ncore;
  const int lane = lane_id(), k = lane & (kCoreCols - 1), sub = lane / kCoreCols;
  constexpr int kSubs = kWave / kCoreCols;
  const int64_t nwaves = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; i < ncore; i += nwaves) {
    const int64_t q0 = rowptr[i], q1 = rowptr[i + 1];
    for (int c0 = 0; c0 < C; c0 += kCoreCols) {
      const bool on = c0 + k < C;
      double acc = 0.0;
      for (int64_t q = q0 + sub; q < q1; q += kSubs) {
        const double v = (double)vals[q];
        if (on) acc = acc - v * X[(size_t)cols[q] * C + c0 + k];
      }
#pragma unroll
      for (int o = kCoreCols; o < kWave; o <<= 1) acc = acc + __shfl_xor(acc, o);
      if (sub == 0 && on) {
        const int c = c0 + k, m = colmap[c];
        const IterParams* prm = ed[m >> 8].prm;
        Y[i * C + c] = (acc + prm->sft.v[m & 255] * X[i * C + c]) * prm->tau;
      }
    }
  }
}

template <int R>
__global__ void __launch_bounds__(kBlock) km_core_apply_shared(const EngDev* __restrict__ ed,
                                                               const uint16_t* __restrict__ map, const double* Y,
                                                               int C) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkCore, g);
  k_core_apply_body<R>(g, a.s, a.wk, a.cnt, a.core_slot, Y + a.core_col0, a.ncore, a.part, C);
}
template <int R, typename VT>
__global__ void __launch_bounds__(kBlock) km_core_spmv(const EngDev* __restrict__ ed, const uint16_t* __restrict__ map) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkCoreRow, g);
  k_core_spmv_body<R, VT>(g, a.core_rowptr, a.core_cols, (const VT*)a.core_vals, a.core_x, a.core_y, a.ncore, a.s.S,
                          a.prm);
}
template <int R>
__global__ void __launch_bounds__(kBlock) km_core_apply(const EngDev* __restrict__ ed,
                                                        const uint16_t* __restrict__ map) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkCore, g);
  k_core_apply_body<R>(g, a.s, a.wk, a.cnt, a.core_slot, a.core_y, a.ncore, a.part, a.s.S);
}

template <int W, int R>
__global__ void __launch_bounds__(kBlock) km_insert(const EngDev* __restrict__ ed, const uint16_t* __restrict__ map) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkGrid, g);
  k_insert_body<W, R>(g, a.s, a.wk, a.cnt, a.n_key, a.n_sgn, a.n_flag, &a.cnt->n_new);
}

template <int W, int R>
__global__ void __launch_bounds__(kBlock) km_stats(const EngDev* __restrict__ ed, const uint16_t* __restrict__ map) {
  Grid g;
  const EngDev& a = eng_of(ed, map, kLkGrid, g);
  k_stats_body<W, R>(g, a.s, a.wk, a.cnt, a.part, a.prm);
}

}  // namespace ngpu
