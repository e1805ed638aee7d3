// ngpu_engine.hip -- engine state, set-up and the C ABI declared in
// include/neci_gpu.h.  One engine = one rank's walkers on one GPU.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "ngpu_kernels.h"
#include "ngpu_batch.h"

using namespace ngpu;

namespace {

constexpr int kStages = 10;  // stage 9: the core Hamiltonian's H.v
constexpr int kEvRing = 32;  // iterations whose stage events may be in flight

#define HIP_OK(x)                                                        \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      fprintf(stderr, "[neci_gpu] %s:%d %s\n", __FILE__, __LINE__,       \
              hipGetErrorString(e_));                                    \
      return NGPU_ERR_HIP;                                               \
    }                                                                    \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  return hipMalloc((void**)p, n * sizeof(T));
}

// Read-only tables (integrals, heat-bath rows, alias tables, lattice) shared
// by every engine of the process on one device whose table has the same
// bytes: replicas of one system then read one copy, which the L2 and the
// MALL can hold, instead of one copy each.  Reference counted; the last
// engine to let go frees it.
struct SharedTable {
  int dev;
  size_t bytes;
  uint64_t hash;
  std::vector<unsigned char> host;
  void* d;
  int refs;
};
std::mutex g_tab_mu;
std::vector<SharedTable*> g_tabs;

uint64_t table_hash(const unsigned char* p, size_t n) {
  uint64_t h = 0x9E3779B97F4A7C15ull ^ (uint64_t)n;
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    memcpy(&w, p + i, 8);
    h = (h ^ (w * 0xBF58476D1CE4E5B9ull)) * 0x94D049BB133111EBull;
    h ^= h >> 29;
  }
  for (; i < n; ++i) h = (h ^ p[i]) * 0x100000001B3ull;
  return h;
}

// *dst: a device copy of the bytes (an existing identical table, or a new one)
int table_acquire(const void* host, size_t bytes, void** dst) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return NGPU_ERR_HIP;
  if (bytes == 0) bytes = 1;
  const unsigned char* hp = (const unsigned char*)host;
  const uint64_t hh = table_hash(hp, bytes);
  std::lock_guard<std::mutex> lock(g_tab_mu);
  for (SharedTable* t : g_tabs)
    if (t->dev == dev && t->bytes == bytes && t->hash == hh && memcmp(t->host.data(), hp, bytes) == 0) {
      t->refs++;
      *dst = t->d;
      return NGPU_OK;
    }
  void* d = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) return NGPU_ERR_HIP;
  if (hipMemcpy(d, hp, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return NGPU_ERR_HIP;
  }
  SharedTable* t = new SharedTable{dev, bytes, hh, std::vector<unsigned char>(hp, hp + bytes), d, 1};
  g_tabs.push_back(t);
  *dst = d;
  return NGPU_OK;
}

// releases device memory: a shared table when its last user lets go, any
// other allocation at once
void table_release(void* d) {
  if (!d) return;
  {
    std::lock_guard<std::mutex> lock(g_tab_mu);
    for (size_t i = 0; i < g_tabs.size(); ++i)
      if (g_tabs[i]->d == d) {
        if (--g_tabs[i]->refs == 0) {
          (void)hipFree(d);
          delete g_tabs[i];
          g_tabs.erase(g_tabs.begin() + (long)i);
        }
        return;
      }
  }
  (void)hipFree(d);
}

template <typename T>
int table_upload(T** dst, const T* host, size_t n) {
  table_release(*dst);
  *dst = nullptr;
  void* d = nullptr;
  int rc = table_acquire(host, n * sizeof(T), &d);
  if (rc) return rc;
  *dst = (T*)d;
  return NGPU_OK;
}

// NECI flag bit numbers (bit_rep_data.F90:87-118)
const int kMneciInit[16] = {13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 57, 58};
const int kMneciDeterm[16] = {41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 61, 62};
const int kNeciInit[2] = {13, 14};
const int kNeciDeterm[2] = {19, 20};

}  // namespace

// Scratch of the allocation-free download (ngpu_download_walkers_dev): the
// rows to keep per tile of kDlTile walker slots, scanned in place into each
// tile's output offset, then the total at [kDlMaxTiles].  One copy per GPU,
// held until the process ends and shared by every call on that GPU, so calls
// that overlap on one GPU race on it.  Host code never passes this name to a
// kernel: the address each device holds it at comes from hipGetSymbolAddress
// (scratch_setup, called by ngpu_create with the device current).
constexpr int64_t kDlTile = 4096;                       // walker slots per tile (one block)
constexpr int64_t kDlMaxTiles = 16384;                  // tiles the scratch holds (128 KiB)
constexpr int64_t kDlMaxSlots = kDlTile * kDlMaxTiles;  // 67,108,864 walker slots
static __device__ __align__(256) unsigned long long scratch[kDlMaxTiles + 32];

namespace {
constexpr int kMaxDevices = 64;
unsigned long long* g_scratch[kMaxDevices] = {nullptr};  // device address of scratch, per device
std::mutex g_scratch_mu;

// the current device's address of scratch, looked up once per device
int scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return NGPU_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev]) return NGPU_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    fprintf(stderr, "[neci_gpu] hipGetSymbolAddress(scratch) on device %d: %s\n", dev, hipGetErrorString(e));
    return NGPU_ERR_HIP;
  }
  g_scratch[dev] = (unsigned long long*)p;
  return NGPU_OK;
}
}  // namespace

struct ngpu_engine {
  ngpu_sys_t sys{};
  int dev = 0;  // the device the engine's memory and stream live on
  int W = 1, S = 1;
  int nb_chunk = 1024;  // blocks of the chunked kernels (prepass, scans, radix)
  int nb_grid = 1024;   // blocks of the grid-stride kernels
  int radix_items = 4;  // elements per thread of a radix tile (2, 4 or 8)
  int single_lanes = 4;  // lanes per heat-bath single attempt (4, 8 or 16)
  int nb_merge = 0;      // blocks of the merge-annihilation kernel (0: nb_grid)
  int n_cu = 256, max_chunk = 4096;
  hipStream_t stream = nullptr;
  DevSys ds{};
  Walk wk{};
  Counters* cnt = nullptr;
  ngpu_iter_out_t* dout = nullptr;
  double* dpart = nullptr;          // kStatCopies rows of block statistics
  ngpu_iter_out_t* hout = nullptr;  // pinned
  Counters* hcnt = nullptr;         // pinned
  // attempts
  int32_t* att_cnt = nullptr;
  int64_t* att_off = nullptr;
  uint8_t* pinfo = nullptr;
  int64_t* blk_part = nullptr;
  int64_t* blk_rm = nullptr;  // per prepass block: slots freed
  int64_t* tile_p = nullptr;  // parent pair of each spawn tile's first attempt
  int64_t tcap = 0;
  // semistochastic core space
  int64_t ncore = 0, core_nnz = 0;
  bool core_fp32 = false;
  int core_build = NGPU_CORE_BUILD_AUTO;  // how the core Hamiltonian is built
  uint64_t* core_keys = nullptr;
  int64_t* core_slot = nullptr;
  int64_t* core_rowptr = nullptr;
  int32_t* core_cols = nullptr;
  void* core_vals = nullptr;  // double, or float when core_fp32
  double* core_x = nullptr;
  double* core_y = nullptr;
  // spawns, sort, merge, new
  int64_t scap = 0;
  uint64_t* sp_rec = nullptr;  // spawned list, kRecWords words per record
  uint64_t* sk[2] = {nullptr, nullptr};
  uint32_t* si[2] = {nullptr, nullptr};
  uint32_t* hist = nullptr;
  uint32_t* digit_tot = nullptr;
  // sort by 32-bit colex rank when every determinant has one
  bool rank_sort = false;
  int rank_bits = 0;
  uint32_t* d_binom = nullptr;  // C(p, i) at [p*(nel+1)+i]
  uint32_t* rk = nullptr;       // ranks of the spawned list
  // one-word determinants with C(nbasis, nel) < 2^64 whose colex rank needs
  // fewer bytes than the word: the spawn sort partitions on the top byte of
  // the (spread) 64-bit rank, then sorts each segment (k_seg_sort64)
  bool rank64 = false;
  int rank64_bits = 0;
  uint64_t* d_binom64 = nullptr;  // C(p, i) at [p*(nel+1)+i], 64 bits
  int64_t* sq = nullptr;  // queued single-excitation attempts
  // merged spawned list, kept only for the explicit-list entry point
  uint64_t* u_key = nullptr;
  double* u_sgn = nullptr;
  uint32_t* u_flag = nullptr;
  unsigned long long* u_count = nullptr;
  uint64_t* n_key = nullptr;
  double* n_sgn = nullptr;
  uint32_t* n_flag = nullptr;
  // system tables
  int32_t* d_rh = nullptr;
  int32_t* d_nbr = nullptr;
  double* d_tmat = nullptr;
  double* d_umat = nullptr;
  double* d_jk = nullptr;
  double* d_jk0 = nullptr;
  double* d_fref = nullptr;  // reference-relative heat-bath single weights
  AliasEnt* d_ab_tab = nullptr;
  double* d_ab_prob = nullptr;
  uint8_t* d_ab_ok = nullptr;
  AliasEnt* d_j_tab = nullptr;
  double* d_j_prob = nullptr;
  uint8_t* d_j_ok = nullptr;
  int32_t* d_ab_tgt = nullptr;
  AliasEnt* d_i_tab = nullptr;
  double* d_i_prob = nullptr;
  AliasEnt* d_a_tab = nullptr;
  double* d_a_prob = nullptr;
  uint8_t* d_a_ok = nullptr;
  AliasEnt* d_b_tab = nullptr;
  double* d_b_prob = nullptr;
  uint8_t* d_b_ok = nullptr;
  bool have_hash = false, have_ref = false, have_model = false, have_excitgen = false;
  bool pending = false;
  // profiling
  bool prof = false;
  uint32_t prof_mask = 0;  // stages timed while prof is set
  hipEvent_t ev[kEvRing][kStages][2];  // start of a stage's first kernel, end of its last
  uint32_t ev_mask[kEvRing] = {0};      // stages whose both events were recorded
  bool ev_pending[kEvRing] = {false};
  int ev_cur = 0;
  int64_t ev_next = 0;
  bool ev_ok = false;
  double stage_ms[kStages] = {0};
  int64_t prof_iters = 0;
  ngpu_iter_out_t* dcyc = nullptr;  // statistics summed over the update cycle
  // per-iteration inputs on the device; while batched this points into the
  // batch's array and own_params keeps the engine's allocation
  IterParams* dparams = nullptr;
  IterParams* own_params = nullptr;
  // inputs of the next iteration when its first kernel writes them (launches
  // issued one by one); a captured graph reads what k_set_params wrote
  IterParams next_params{};
  bool chain_sets_params = false;
  // replay of the captured iteration chain
  bool use_graph = false;
  bool graph_dirty = true;
  uint32_t checks = 0;                  // NGPU_CHECK_* bits (ngpu_set_checks)
  unsigned long long* d_chk = nullptr;  // per-iteration and cumulative check counts
  hipGraphExec_t gexec = nullptr;
  ngpu_batch* batch = nullptr;
  hipStream_t own_stream = nullptr;
};

// Engines of a batch whose chains use the same kernel instances: each kernel
// of the chain is one launch for all of them (ngpu_batch.h).
struct BatchGroup {
  std::vector<ngpu_t*> hs;
  int W = 1, S = 1;
  int gen = 0;              // kGenHubbard, kGenPchb or kGenUniform
  bool singles = false;     // queued heat-bath singles (PCHB generator)
  int lanes = 4, pl = 4;    // heat-bath singles: lanes per attempt, orbitals per lane
  int items = 4;            // radix tile: elements per thread
  bool rank_sort = false;
  bool rank64 = false;                      // the 64-bit rank partition sort
  int rank64_bits = 0;
  std::vector<std::pair<int, int>> passes;  // (word, shift) of the word-key sort
  bool core = false, core_fp32 = false;
  // every engine of the group has the same core determinants: one projection
  // over the interleaved vectors X, Y of ccols columns (kernels in ngpu_batch.h)
  bool shared_core = false;
  int ccols = 0;
  double* d_X = nullptr;
  double* d_Y = nullptr;
  int32_t* d_colmap = nullptr;
  int rec_words = 0;
  EngDev* d_ed = nullptr;
  uint16_t* d_map[kLkKinds] = {nullptr};
  int tot[kLkKinds] = {0};
};

struct ngpu_batch {
  std::vector<ngpu_t*> hs;
  std::vector<BatchGroup> groups;
  bool groups_ok = false;   // groups and their device descriptors match the engines
  bool use_graph = true;    // replay one captured graph per iteration (else launch directly)
  hipStream_t stream = nullptr;
  std::vector<hipStream_t> side;  // one per group
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  hipGraphExec_t gexec = nullptr;
  IterParams* dparams = nullptr;                 // one entry per engine
  static constexpr int kRing = 4;
  IterParams* hparams[kRing] = {nullptr};        // pinned staging slots
  hipEvent_t slot_done[kRing] = {nullptr};
  bool slot_used[kRing] = {false};
  int next_slot = 0;
};

namespace {

int ready(ngpu_t* h) {
  if (!h) return NGPU_ERR_ARG;
  if (!h->have_hash || !h->have_ref || !h->have_model) return NGPU_ERR_STATE;
  if (h->sys.model == NGPU_MODEL_MOLECULAR && !h->have_excitgen) return NGPU_ERR_STATE;
  return NGPU_OK;
}

int err_from_mask(unsigned int m) {
  for (int e = 1; e < 16; ++e)
    if (m & (1u << e)) return e;
  return NGPU_OK;
}

// the state an iteration starts from: statistics rows and append counters
// empty.  Iterations leave it behind themselves (k_end_iter); this kernel
// restores it after the set-up paths and starts an explicit-list annihilation.
__global__ void k_begin_iter(ngpu_iter_out_t* out, double* part, Counters* cnt, IterParams* prm, IterParams v,
                             int write_params) {
  if (write_params && threadIdx.x == 0) *prm = v;  // the iteration's inputs (not in a captured graph)
  double* g = (double*)out;
  for (int i = threadIdx.x; i < S_NDOUBLE; i += blockDim.x) g[i] = 0.0;
  for (int i = threadIdx.x; i < kStatCopies * kStatStride; i += blockDim.x) part[i] = 0.0;
  if (threadIdx.x == 0) {
    cnt->n_attempts = 0;
    cnt->n_spawned = 0;
    cnt->n_unique = 0;
    cnt->n_new = 0;
    cnt->n_occ = 0;
    cnt->n_singles = 0;
    cnt->n_fail = 0;
    cnt->ins_refused = 0;
    cnt->err = 0;
    out->tot_walkers = 0;
    out->n_attempts = out->n_spawned = out->n_unique = out->n_new = 0;
    out->err = 0;
  }
}

// sum of the block-statistics rows, final counters of the iteration, then
// accumulation into the cycle sums (ENumCyc, HFCyc, NoBorn, ... are cycle
// accumulators in the reference too)
__device__ __forceinline__ void end_iter_body(ngpu_iter_out_t* out, double* part, ngpu_iter_out_t* cyc, Counters* cnt,
                                              int64_t scap, int64_t cap, int64_t hash_entries, const IterParams* prm) {
  // halted: the output block keeps the failing iteration's statistics and
  // error, and nothing is added to the cycle
  if (cnt->halt) return;
  __shared__ unsigned long long n_unique;
  if (threadIdx.x < kWave) {
    // the spread per-block counts, one line per lane, summed across the wave
    const int l = threadIdx.x;
    long long r = 0, u = 0;
    if (l < Counters::kSpread) {
      r = (long long)cnt->reuse_part[l].v;
      u = (long long)cnt->uniq_part[l].v;
      cnt->reuse_part[l].v = 0;
      cnt->uniq_part[l].v = 0;
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
      r += __shfl_xor(r, o);
      u += __shfl_xor(u, o);
    }
    if (l == 0) {
      post_insert(cnt, cnt->n_new, cap, (unsigned long long)r);  // the slots k_insert filled
      hash_account(cnt, hash_entries);
      n_unique = (unsigned long long)u;
    }
  }
  double* g = (double*)out;
  double* c = (double*)cyc;
  for (int i = threadIdx.x; i < kStatStride; i += blockDim.x) {
    bool is_max = stat_is_max(i);
    double v = 0.0;
    for (int q = 0; q < kStatCopies; ++q) {
      double x = part[(size_t)q * kStatStride + i];
      part[(size_t)q * kStatStride + i] = 0.0;  // empty for the next iteration
      v = is_max ? (x > v ? x : v) : v + x;
    }
    if (i == X_NOCC) cnt->n_occ = (unsigned long long)v;
    if (i >= S_NDOUBLE) continue;
    g[i] = v;
    if (is_max) c[i] = c[i] > v ? c[i] : v;
    else if (i >= soff(F_INSTHF, 0) && i < soff(F_INSTHF, kMaxRep)) c[i] = v;
    else c[i] += v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out->tot_walkers = (int64_t)cnt->n_occ;
    out->n_attempts = (int64_t)cnt->n_attempts;
    int64_t ns = (int64_t)cnt->n_spawned;
    out->n_spawned = ns < scap ? ns : scap;
    const unsigned long long nu = n_unique;
    cnt->n_unique = nu;
    out->n_unique = (int64_t)nu;
    out->n_new = (int64_t)cnt->n_new;
    out->err = (int32_t)cnt->err;
    cyc->tot_walkers = out->tot_walkers;
    cyc->n_attempts += out->n_attempts;
    cyc->n_spawned += out->n_spawned;
    cyc->n_unique += out->n_unique;
    cyc->n_new += out->n_new;
    cyc->err |= out->err;
    cyc->pad += 1;  // iterations in the cycle
    if (cnt->err) {  // the chain halts from the next iteration on
      cnt->halt = cnt->err;
      cnt->err_iter = prm->iter;
    }
    // the counters the next iteration appends to start from zero (the chain
    // has no separate reset launch)
    cnt->n_spawned = 0;
    cnt->n_singles = 0;
    cnt->n_new = 0;
    cnt->n_fail = 0;
    cnt->ins_refused = 0;
    cnt->err = 0;
  }
}
__global__ void k_end_iter(ngpu_iter_out_t* out, double* part, ngpu_iter_out_t* cyc, Counters* cnt,
                           int64_t scap, int64_t cap, int64_t hash_entries, const IterParams* prm) {
  end_iter_body(out, part, cyc, cnt, scap, cap, hash_entries, prm);
}
// the end of the iteration of every engine of a batch group: one block each
__global__ void km_end_iter(const EngDev* __restrict__ ed) {
  const EngDev& a = ed[blockIdx.x];
  end_iter_body(a.dout, a.part, a.dcyc, a.cnt, a.scap, a.wk.cap, a.wk.L * (int64_t)kBucketEntries, a.prm);
}

__global__ void k_reset_cycle(ngpu_iter_out_t* cyc) {
  double* c = (double*)cyc;
  for (int i = threadIdx.x; i < S_NDOUBLE; i += blockDim.x) c[i] = 0.0;
  if (threadIdx.x == 0) {
    cyc->tot_walkers = cyc->n_attempts = cyc->n_spawned = cyc->n_unique = cyc->n_new = 0;
    cyc->err = 0;
    cyc->pad = 0;
  }
}

__global__ void k_set_count(unsigned long long* p, unsigned long long v) { *p = v; }
__global__ void k_set_u32(unsigned int* p, unsigned int v) { *p = v; }

// Kernel launch on the engine stream.  With profiling on, the launch that
// opens stage `open` carries that stage's start event and the launch that
// closes stage `close` its stop event (hipExtLaunchKernel records them at the
// kernel's own start and end, not at a neighbouring command).
template <typename F, typename... Args>
void launch_lds(ngpu_t* h, int open, int close, F kern, dim3 grid, dim3 block, size_t lds, Args... args) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->prof && h->ev_ok) {
    if (open >= 0 && ((h->prof_mask >> open) & 1u)) e0 = h->ev[h->ev_cur][open][0];
    if (close >= 0 && ((h->prof_mask >> close) & 1u)) {
      e1 = h->ev[h->ev_cur][close][1];
      h->ev_mask[h->ev_cur] |= 1u << close;
    }
  }
  if (e0 || e1) hipExtLaunchKernelGGL(kern, grid, block, (uint32_t)lds, h->stream, e0, e1, 0, args...);
  else hipLaunchKernelGGL(kern, grid, block, (uint32_t)lds, h->stream, args...);
}

template <typename F, typename... Args>
void launch(ngpu_t* h, int open, int close, F kern, dim3 grid, dim3 block, Args... args) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->prof && h->ev_ok) {
    if (open >= 0 && ((h->prof_mask >> open) & 1u)) e0 = h->ev[h->ev_cur][open][0];
    if (close >= 0 && ((h->prof_mask >> close) & 1u)) {
      e1 = h->ev[h->ev_cur][close][1];
      h->ev_mask[h->ev_cur] |= 1u << close;
    }
  }
  if (e0 || e1) hipExtLaunchKernelGGL(kern, grid, block, 0, h->stream, e0, e1, 0, args...);
  else hipLaunchKernelGGL(kern, grid, block, 0, h->stream, args...);
}

// tombstone reclamation and the load guard after a set-up path (upload, core
// placement): the kernels exit at once unless the counters call for a
// rebuild.  Iterations launch none of them: the end-of-iteration accounting
// flags a due rebuild and the next iteration's prepass and attempt scan do it.
template <int W>
void hash_upkeep(ngpu_t* h) {
  launch(h, -1, -1, k_hash_clear<W>, dim3(h->nb_grid), dim3(kBlock), h->wk, (const Counters*)h->cnt);
  launch(h, -1, -1, k_hash_refill<W>, dim3(h->nb_grid), dim3(kBlock), h->ds, h->wk, h->cnt);
  launch(h, -1, -1, k_hash_done, dim3(1), dim3(1), h->cnt, h->wk.L * (int64_t)kBucketEntries);
}

// (word, shift) passes of the word-key sort of an engine (ilut_lt order:
// least significant word first)
std::vector<std::pair<int, int>> word_passes(const ngpu_t* h) {
  std::vector<std::pair<int, int>> p;
  const int hi = h->W == 1 ? h->sys.nbasis : 64, lo = h->W == 2 ? h->sys.nbasis - 64 : 0;
  if (h->W == 2)
    for (int sh = 0; sh < lo; sh += 8) p.push_back({1, sh});
  for (int sh = 0; sh < hi; sh += 8) p.push_back({0, sh});
  return p;
}

// sort of the spawned list (stage 4) then the merge of equal determinants
// fused with their annihilation into the walker list (stage 6; stage 5, the
// merge alone, is empty since the fusion); ranked: the spawn kernels already
// wrote each record's colex rank
// the rank path: one stable pass on the keys' top byte, then one block per
// top-byte segment sorts it by the low 24 bits (k_seg_sort); the sorted
// (key, spawn index) pairs end in buffer 0
template <int W, int R, int IT>
int sort_ranks(ngpu_t* h, bool ranked) {
  const unsigned long long* n_ptr = &h->cnt->n_spawned;
  constexpr int kTop = 32 - kSegBits;  // the partition digit's shift
  if (ranked)
    launch(h, 4, -1, k_radix_hist<uint32_t, 1, false, IT, kSegBits>, dim3(h->nb_chunk), dim3(kBlock),
           (const uint32_t*)h->rk, 1, n_ptr, h->scap, 0, kTop, h->hist);
  else
    launch(h, 4, -1, k_rank_hist<W, IT, kSegBits>, dim3(h->nb_chunk), dim3(kBlock), (const uint64_t*)h->sp_rec,
           n_ptr, h->scap, h->ds.rec_words, (const uint32_t*)h->d_binom, h->sys.nel + 1, h->rk, h->hist, kTop);
  launch(h, -1, -1, k_radix_scan, dim3(kSegs), dim3(kBlock), h->hist, h->nb_chunk, h->digit_tot);
  launch(h, -1, -1, k_radix_scatter<uint32_t, 1, false, IT, kSegBits>, dim3(h->nb_chunk), dim3(kBlock),
         (const uint32_t*)h->rk, 1, (const uint32_t*)nullptr, (uint32_t*)h->sk[1], h->si[1], n_ptr, h->scap, 0, kTop,
         (const uint32_t*)h->hist, (const uint32_t*)h->digit_tot);
  launch(h, -1, 4, k_seg_sort, dim3(kSegs), dim3(kSegThreads), (uint32_t*)h->sk[1], h->si[1], (uint32_t*)h->sk[0],
         h->si[0], (const uint32_t*)h->digit_tot);
  return 0;
}

template <int W, int R, int IT>
void sort_and_merge_it(ngpu_t* h, bool ranked, bool list_out) {
  const unsigned long long* n_ptr = &h->cnt->n_spawned;
  if (h->rank_sort) {
    const int cur = sort_ranks<W, R, IT>(h, ranked);
    if (list_out)
      launch(h, -1, -1, k_merge_list<W, R, uint32_t, 1>, dim3(h->nb_grid), dim3(kBlock), h->ds, h->cnt,
             (const uint32_t*)h->sk[cur], (const uint32_t*)h->si[cur], (const uint64_t*)h->sp_rec, h->scap, h->u_key,
             h->u_sgn, h->u_flag, h->u_count);
    launch(h, 6, 6, k_merge_annihilate<W, R, uint32_t, 1>, dim3(h->nb_merge ? h->nb_merge : h->nb_grid), dim3(kBlock), h->ds, h->wk, h->cnt,
           h->dpart, (const uint32_t*)h->sk[cur], (const uint32_t*)h->si[cur], (const uint64_t*)h->sp_rec, h->scap,
           h->n_key, h->n_sgn, h->n_flag, (const IterParams*)h->dparams);
    return;
  }
  if (h->rank64) {
    // the 64-bit rank keys (spread over the top byte) into sk[0], one stable
    // pass on the top byte into sk[1], then each top-byte segment sorted by
    // the bytes below it (k_seg_sort64) back into sk[0]
    const int bits = h->rank64_bits;
    launch(h, 4, -1, k_rank64_hist<IT, kSegBits>, dim3(h->nb_chunk), dim3(kBlock), (const uint64_t*)h->sp_rec,
           h->ds.rec_words, n_ptr, h->scap, (const uint64_t*)h->d_binom64, h->sys.nbasis, h->sys.nel + 1, bits,
           h->sk[0], h->hist);
    launch(h, -1, -1, k_radix_scan, dim3(kSegs), dim3(kBlock), h->hist, h->nb_chunk, h->digit_tot);
    launch(h, -1, -1, k_radix_scatter<uint64_t, 1, false, IT, kSegBits>, dim3(h->nb_chunk), dim3(kBlock),
           (const uint64_t*)h->sk[0], 1, (const uint32_t*)nullptr, h->sk[1], h->si[1], n_ptr, h->scap, 0,
           bits - kSegBits, (const uint32_t*)h->hist, (const uint32_t*)h->digit_tot);
    launch(h, -1, 4, k_seg_sort64, dim3(kSegs), dim3(kSegThreads), h->sk[1], h->si[1], h->sk[0], h->si[0],
           (const uint32_t*)h->digit_tot, (bits - kSegBits + 7) / 8);
    if (list_out)
      launch(h, -1, -1, k_merge_list<W, R, uint64_t, W>, dim3(h->nb_grid), dim3(kBlock), h->ds, h->cnt,
             (const uint64_t*)h->sk[0], (const uint32_t*)h->si[0], (const uint64_t*)h->sp_rec, h->scap, h->u_key,
             h->u_sgn, h->u_flag, h->u_count);
    launch(h, 6, 6, k_merge_annihilate<W, R, uint64_t, W>, dim3(h->nb_merge ? h->nb_merge : h->nb_grid),
           dim3(kBlock), h->ds, h->wk, h->cnt, h->dpart, (const uint64_t*)h->sk[0], (const uint32_t*)h->si[0],
           (const uint64_t*)h->sp_rec, h->scap, h->n_key, h->n_sgn, h->n_flag, (const IterParams*)h->dparams);
    return;
  }
  // determinant keys in ilut_lt order: least significant word first
  const std::vector<std::pair<int, int>> passes = word_passes(h);
  int cur = 0;
  const uint64_t* kin = h->sp_rec;
  int kstride = h->ds.rec_words;  // the first pass reads the keys inside the records
  const uint32_t* iin = nullptr;
  for (size_t p = 0; p < passes.size(); ++p) {
    int out = (int)(p & 1);
    bool first = p == 0, last = p + 1 == passes.size();
    launch(h, first ? 4 : -1, -1, k_radix_hist<uint64_t, W, W == 2, IT, 8>, dim3(h->nb_chunk), dim3(kBlock), kin,
           kstride, n_ptr, h->scap, passes[p].first, passes[p].second, h->hist);
    launch(h, -1, -1, k_radix_scan, dim3(256), dim3(kBlock), h->hist, h->nb_chunk, h->digit_tot);
    launch(h, -1, last ? 4 : -1, k_radix_scatter<uint64_t, W, W == 2, IT, 8>, dim3(h->nb_chunk), dim3(kBlock), kin, kstride, iin,
           h->sk[out], h->si[out], n_ptr, h->scap, passes[p].first, passes[p].second, (const uint32_t*)h->hist,
           (const uint32_t*)h->digit_tot);
    kin = h->sk[out];
    kstride = W;
    iin = h->si[out];
    cur = out;
  }
  if (list_out)
    launch(h, -1, -1, k_merge_list<W, R, uint64_t, W>, dim3(h->nb_grid), dim3(kBlock), h->ds, h->cnt,
           (const uint64_t*)h->sk[cur], (const uint32_t*)h->si[cur], (const uint64_t*)h->sp_rec, h->scap, h->u_key,
           h->u_sgn, h->u_flag, h->u_count);
  launch(h, 6, 6, k_merge_annihilate<W, R, uint64_t, W>, dim3(h->nb_merge ? h->nb_merge : h->nb_grid), dim3(kBlock), h->ds, h->wk, h->cnt,
         h->dpart, (const uint64_t*)h->sk[cur], (const uint32_t*)h->si[cur], (const uint64_t*)h->sp_rec, h->scap,
         h->n_key, h->n_sgn, h->n_flag, (const IterParams*)h->dparams);
}

// the sort with the engine's tile size (elements per thread of a radix tile)
template <int W, int R>
void sort_and_merge(ngpu_t* h, bool ranked, bool list_out) {
  if (h->radix_items == 2) sort_and_merge_it<W, R, 2>(h, ranked, list_out);
  else if (h->radix_items == 8) sort_and_merge_it<W, R, 8>(h, ranked, list_out);
  else sort_and_merge_it<W, R, 4>(h, ranked, list_out);
}

// semistochastic projection of the core amplitudes (before spawning)
template <int R>
void core_projection(ngpu_t* h) {
  int nb = (int)std::min<int64_t>((h->ncore + kBlock - 1) / kBlock, h->nb_grid);
  launch(h, -1, -1, k_core_gather<R>, dim3(nb), dim3(kBlock), (const int64_t*)h->core_slot, h->wk,
         h->S, h->ncore, h->core_x);
  int nbw = (int)std::min<int64_t>((h->ncore + kBlock / kWave - 1) / (kBlock / kWave), 8 * h->nb_grid);
  if (h->core_fp32)
    launch(h, 9, 9, k_core_spmv<R, float>, dim3(nbw), dim3(kBlock), (const int64_t*)h->core_rowptr,
           (const int32_t*)h->core_cols, (const float*)h->core_vals, (const double*)h->core_x, h->core_y, h->ncore,
           h->S, (const IterParams*)h->dparams);
  else
    launch(h, 9, 9, k_core_spmv<R, double>, dim3(nbw), dim3(kBlock), (const int64_t*)h->core_rowptr,
           (const int32_t*)h->core_cols, (const double*)h->core_vals, (const double*)h->core_x, h->core_y, h->ncore,
           h->S, (const IterParams*)h->dparams);
}

// merge -> annihilate -> insert -> stats, shared by iterate and the explicit-list entry
template <int W>
void check_chain(ngpu_t* h, bool gen);

template <int W, int R>
void annihilation_chain(ngpu_t* h, bool with_core, bool list_out = false) {
  // the projected core amplitudes join the walkers before the spawns do, as
  // deterministic_annihilation precedes AnnihilateSpawnedParts
  if (h->ncore && with_core)
    launch(h, -1, -1, k_core_apply<R>, dim3((int)std::min<int64_t>((h->ncore + kBlock - 1) / kBlock, h->nb_grid)),
           dim3(kBlock), h->ds, h->wk, (const Counters*)h->cnt, (const int64_t*)h->core_slot, (const double*)h->core_y,
           h->ncore, h->dpart);
  sort_and_merge<W, R>(h, with_core, list_out);  // spawns of the iteration chain carry their ranks
  launch(h, 7, 7, k_insert<W, R>, dim3(h->nb_grid), dim3(kBlock), h->ds, h->wk, h->cnt,
         (const uint64_t*)h->n_key, (const double*)h->n_sgn, (const uint32_t*)h->n_flag,
         (const unsigned long long*)&h->cnt->n_new);
  launch(h, 8, 8, k_stats<W, R>, dim3(h->nb_grid), dim3(kBlock), h->ds, h->wk, h->cnt, h->dpart,
         (const IterParams*)h->dparams);
  launch(h, -1, -1, k_end_iter, dim3(1), dim3(kBlock), h->dout, h->dpart, h->dcyc, h->cnt,
         h->scap, h->wk.cap, h->wk.L * (int64_t)kBucketEntries, (const IterParams*)h->dparams);
  if (h->checks) check_chain<W>(h, with_core);
}

// the generator instance of an engine's spawn kernel
int gen_of(const ngpu_t* h) {
  if (h->sys.model == NGPU_MODEL_HUBBARD) return kGenHubbard;
  if (h->ds.uniform) return kGenUniform;
  return (h->ds.particles >= NGPU_PART_UNIF_FULL || h->ds.holes) ? kGenPchbW : kGenPchb;
}

template <int W, int R>
void iterate_chain(ngpu_t* h) {
  launch(h, 0, 0, k_prepass<W, R>, dim3(h->nb_chunk), dim3(kBlock), h->ds, h->wk, h->cnt, h->dpart, h->att_cnt,
         h->pinfo, h->blk_part, h->blk_rm, h->dparams, h->next_params, h->chain_sets_params ? 1 : 0);
  if (h->ncore) core_projection<R>(h);
  launch(h, 1, 1, k_scan_apply<W>, dim3(h->nb_chunk), dim3(kBlock), h->ds, h->wk, (const int32_t*)h->att_cnt,
         h->att_off, (const int64_t*)h->blk_part, (const int64_t*)h->blk_rm, (const uint8_t*)h->pinfo, h->cnt,
         h->S, h->tile_p, h->tcap);
  const int gen = gen_of(h);
  auto spawn = gen == kGenHubbard ? k_spawn<W, R, kGenHubbard>
             : gen == kGenUniform ? k_spawn<W, R, kGenUniform>
             : gen == kGenPchbW ? k_spawn<W, R, kGenPchbW> : k_spawn<W, R, kGenPchb>;
  launch(h, 2, 2, spawn, dim3(h->nb_grid), dim3(kBlock), h->ds, h->wk, h->cnt, h->dpart,
         (const int64_t*)h->att_off, (const int64_t*)h->tile_p, h->tcap, (const uint8_t*)h->pinfo, h->sp_rec,
         h->rank_sort ? h->rk : nullptr, h->scap, h->sq, (const IterParams*)h->dparams);
  if (h->sys.model == NGPU_MODEL_MOLECULAR && !h->ds.uniform) {
    const int nsp = h->sys.nbasis / 2;
    // G lanes per attempt, PL >= nsp / G orbitals per lane
    const int G = h->single_lanes;
    const size_t lds = 2 * sizeof(double) * (size_t)(kBlock / G) * nsp;  // weights and running sums
    auto kern = G == 16 ? (nsp <= 16 ? k_spawn_singles<W, R, 1, 16> : nsp <= 32 ? k_spawn_singles<W, R, 2, 16>
                           : k_spawn_singles<W, R, 4, 16>)
              : G == 8 ? (nsp <= 16 ? k_spawn_singles<W, R, 2, 8> : nsp <= 24 ? k_spawn_singles<W, R, 3, 8>
                          : nsp <= 32 ? k_spawn_singles<W, R, 4, 8> : nsp <= 40 ? k_spawn_singles<W, R, 5, 8>
                          : k_spawn_singles<W, R, 8, 8>)
              : nsp <= 16 ? k_spawn_singles<W, R, 4, 4> : nsp <= 20 ? k_spawn_singles<W, R, 5, 4>
              : nsp <= 32 ? k_spawn_singles<W, R, 8, 4> : nsp <= 40 ? k_spawn_singles<W, R, 10, 4>
              : k_spawn_singles<W, R, 16, 4>;
    launch_lds(h, 3, 3, kern, dim3(h->nb_grid), dim3(kBlock), lds, h->ds, h->wk, h->cnt, h->dpart,
               (const uint8_t*)h->pinfo, h->sp_rec, h->rank_sort ? h->rk : nullptr, h->scap, (const int64_t*)h->sq,
               (const IterParams*)h->dparams);
  }
  annihilation_chain<W, R>(h, true);
}

void harvest(ngpu_t* h, int slot) {
  if (!h->ev_pending[slot]) return;
  for (int s = 0; s < kStages; ++s) {
    if (!(h->ev_mask[slot] & (1u << s))) continue;
    (void)hipEventSynchronize(h->ev[slot][s][1]);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->ev[slot][s][0], h->ev[slot][s][1]) == hipSuccess) h->stage_ms[s] += ms;
  }
  h->prof_iters++;
  h->ev_pending[slot] = false;
  h->ev_mask[slot] = 0;
}

int finish(ngpu_t* h, ngpu_iter_out_t* out) {
  HIP_OK(hipStreamSynchronize(h->stream));
  h->pending = false;
  if (h->ev_ok)
    for (int q = 0; q < kEvRing; ++q) harvest(h, q);
  if (out) *out = *h->hout;
  return err_from_mask((unsigned)h->hout->err);
}

// ---- host restatement of the PCHB alias tables ----
// init_AliasTable_t (aliasSampling.F90:209-297), init_probs (:412-418)
void build_alias(const std::vector<double>& w, AliasEnt* tab, double* prob) {
  int n = (int)w.size();
  double sum = 0.0;
  for (int i = 0; i < n; ++i) sum += w[i];
  std::vector<double> bias(n);
  std::vector<int> over, under;
  over.reserve(n);
  under.reserve(n);
  std::vector<int> alias(n, 0);
  for (int i = 0; i < n; ++i) bias[i] = w[i] / sum * (double)n;
  for (int i = 0; i < n; ++i) (bias[i] > 1.0 ? over : under).push_back(i);
  std::reverse(over.begin(), over.end());
  while (!over.empty() && !under.empty()) {
    int i = over.back(), j = under.back();
    alias[j] = i;
    bias[i] = bias[i] + bias[j] - 1.0;
    over.pop_back();
    under.pop_back();
    (bias[i] > 1.0 ? over : under).push_back(i);
  }
  for (int i : over) { bias[i] = 1.0; alias[i] = i; }
  for (int i : under) { bias[i] = 1.0; alias[i] = i; }
  for (int i = 0; i < n; ++i) prob[i] = w[i] / sum;
  for (int i = 0; i < n; ++i) {
    tab[i].bias = bias[i];
    tab[i].prob_self = prob[i];
    tab[i].prob_alias = prob[alias[i]];
    tab[i].alias = alias[i];
    tab[i].pad = 0;
  }
}

template <int W>
__global__ void k_batch_hash(DevSys s, const uint64_t* d, int64_t n, int32_t* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (int32_t)find_walker_hash<W>(d + i * W, s.rh, s.L);
}

template <int W>
__global__ void k_batch_hel(DevSys s, const uint64_t* a, const uint64_t* b, int64_t n, double* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = helement_pair<W>(s, a + i * W, b + i * W);
}

template <int W>
__global__ void k_batch_diag(DevSys s, const uint64_t* d, int64_t n, double* diag, double* off) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    diag[i] = diag_helement<W>(s, d + i * W);
    off[i] = offdiag_helement<W>(s, d + i * W);
  }
}

template <int W>
__global__ void k_batch_gen(DevSys s, const uint64_t* d, const double* draws, int64_t n, uint64_t* dj, int32_t* ic,
                            int32_t* ex, int32_t* par, double* pgen) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double dr[6];
    for (int q = 0; q < 6; ++q) dr[q] = draws[i * 6 + q];
    Excit e = gen_excit<W>(s, d + i * W, dr);
    ic[i] = e.ic;
    ex[4 * i] = e.e0; ex[4 * i + 1] = e.e1; ex[4 * i + 2] = e.e2; ex[4 * i + 3] = e.e3;
    par[i] = e.parity ? 1 : 0;
    pgen[i] = e.ic ? e.pgen : 0.0;
    if (e.ic) apply_excit<W>(d + i * W, e, dj + i * W);
    else
      for (int w = 0; w < W; ++w) dj[i * W + w] = 0;
  }
}

// calc_pgen of the bound generator (calc_pgen_rs_hubbard, real_space_hubbard.F90:2035;
// pgen_single_4ind, excit_gens_int_weighted.F90:491; GAS_doubles_PCHB_spinorb_get_pgen,
// gasci_pchb_doubles_spinorb_fastweighted.fpp:246; get_pgen_sd, excitation_generators.F90:141)
template <int W>
__device__ double calc_pgen(const DevSys& s, const uint64_t* d, int ic, const int32_t* ex) {
  if (s.model == 0) {
    if (ic != 1) return 0.0;
    int src = ex[0], tgt = ex[1], site = spat(src);
    double cum_sum = 0.0, cpt = 0.0;
    for (int k = 0; k < s.maxnbr; ++k) {
      int ns = s.nbr[(site - 1) * s.maxnbr + k];
      if (ns == 0) break;
      int orb = is_beta(src) ? 2 * ns - 1 : 2 * ns;
      double elem = test_orb<W>(d, orb) ? 0.0 : fabs(s.bhub);
      cum_sum += elem;
      if (orb == tgt) cpt = elem;
    }
    if (cum_sum < kEps) return 0.0;
    return (1.0 / (double)s.nel) * (cpt / cum_sum);
  }
  if (s.uniform) return pgen_uniform<W>(s, d, ic, ex);
  if (ic == 1) {
    int src = ex[0], tgt = ex[1];
    double pgen = 1.0 / (double)s.nel;
    int nsp = s.nbasis / 2, off = is_beta(tgt) ? 1 : 0;
    double cum_sum = 0.0, cpt_tgt = 0.0;
    for (int i = 1; i <= nsp; ++i) {
      int orb = 2 * i - off;
      if (!test_orb<W>(d, orb)) {
        double cpt = fabs(hb_weight<W>(s, d, src, orb));
        cum_sum += cpt;
        if (orb == tgt) cpt_tgt = cpt;
      }
    }
    if (fabs(cum_sum) <= kEps) pgen = 0.0; else pgen = pgen * cpt_tgt / cum_sum;
    return s.psingles * pgen;
  }
  if (ic == 2) {
    int I = ex[0], J = ex[1], n = s.nbasis;
    double p;
    if (s.particles >= 2) {
      p = pgen_weighted_particles<W>(s, d, I, J);
    } else if (s.particles == 1) {
      double a = s.j_ok[I - 1] ? s.j_prob[(size_t)(I - 1) * n + (J - 1)] : 0.0;
      double b = s.j_ok[J - 1] ? s.j_prob[(size_t)(J - 1) * n + (I - 1)] : 0.0;
      p = (a + b) / (double)s.nel;
    } else {
      const uint64_t kA = 0xAAAAAAAAAAAAAAAAull, kB = 0x5555555555555555ull;
      int na = 0, nb = 0;
      for (int w = 0; w < W; ++w) { na += popc64(d[w] & kA); nb += popc64(d[w] & kB); }
      int par = na * (na - 1) / 2 + nb * (nb - 1) / 2;
      if (is_beta(I) == is_beta(J)) p = s.pparallel / (double)par;
      else p = (1.0 - s.pparallel) / (double)(na * nb);
    }
    double ph;
    if (s.holes == 1) {
      ph = pgen_full_holes<W>(s, d, I, J, ex[2], ex[3]);
    } else {
      int ij = fuse_index(I, J), ab = fuse_index(ex[2], ex[3]);
      ph = s.ab_ok[ij - 1] ? s.ab_prob[(size_t)(ij - 1) * s.ijmax + ab - 1] : 0.0;
    }
    p = p * ph;
    return s.pdoubles * p;
  }
  return 0.0;
}

// ---- invariant checks (ngpu_set_checks) ----
// Counts of one iteration in d_chk[0..7], cumulative counts in d_chk[8..15]:
// holes, live slots, wrong electron counts, hash misses, pgen tests, pgen
// mismatches; then the cumulative ones (ngpu_get_check_counts order).
enum { kChkHoles = 0, kChkLive, kChkNel, kChkHash, kChkPgenN, kChkPgenBad, kChkCum = 8 };

__device__ inline void chk_add(unsigned long long* chk, int i, unsigned long long v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (lane_id() == 0 && v) atomicAdd(&chk[i], v);
}

// the walker table after the iteration's bookkeeping: every slot below the
// high-water mark is either removed (a hole in the free queue) or live, and a
// live one carries nel electrons and is found by a hash lookup at its own slot
template <int W>
__global__ void __launch_bounds__(kBlock) k_check_table(DevSys s, Walk wk, Counters* cnt, unsigned long long* chk) {
  const int64_t nslots = (int64_t)cnt->nslots;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t trips = (nslots + stride - 1) / stride;
  for (int64_t t = 0; t < trips; ++t) {  // every lane runs every trip (wave-wide sums)
    const int64_t slot = t * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long holes = 0, live = 0, bad_nel = 0, bad_hash = 0;
    if (slot < nslots) {
      if (wk.flags(slot) & kFlagRemoved) {
        holes = 1;
      } else {
        live = 1;
        uint64_t key[W];
        int nel = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          key[w] = wk.keys[slot * W + w];
          nel += popc64(key[w]);
        }
        bad_nel = nel != s.nel;
        const int64_t home = find_walker_hash<W>(key, s.rh, s.L) - 1;
        const int64_t pos = hash_lookup<W>(wk, key, home, cnt);
        const int64_t hp = wk.hpos(slot);
        const int32_t* sl = (const int32_t*)(wk.htab + (size_t)(hp >> 2) * 8 * W + 4 * W);
        bad_hash = pos != slot || sl[hp & 3] != (int32_t)slot;
      }
    }
    chk_add(chk, kChkHoles, holes);
    chk_add(chk, kChkLive, live);
    chk_add(chk, kChkNel, bad_nel);
    chk_add(chk, kChkHash, bad_hash);
  }
}

// the generators' pgen against calc_pgen (the reference asserts
// pgen .isclose. get_pgen after every draw, gasci_pchb_doubles_spinorb_
// fastweighted.fpp:191,231): one attempt of every occupied determinant and
// replica, drawn as the iteration's first attempt; off = 1 shifts the
// recomputed value by 1e-6 (a deliberate mismatch, for testing the report)
template <int W>
__global__ void __launch_bounds__(kBlock) k_check_pgen(DevSys s, Walk wk, const Counters* cnt,
                                                       unsigned long long* chk, const IterParams* prm, int off) {
  const int64_t nslots = (int64_t)cnt->nslots;
  const int S = s.S;
  const int64_t n = nslots * S;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t trips = (n + stride - 1) / stride;
  for (int64_t t = 0; t < trips; ++t) {
    const int64_t q = t * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long tests = 0, bad = 0;
    if (q < n) {
      const int64_t slot = q / S;
      const int r = (int)(q - slot * S);
      if (!(wk.flags(slot) & kFlagRemoved) && wk.sgn(slot, r) != 0.0) {
        uint64_t key[W];
#pragma unroll
        for (int w = 0; w < W; ++w) key[w] = wk.keys[slot * W + w];
        double dr[6];
        draw2<W>(s.seed, key, prm->iter, r, 0, 1, dr[0], dr[1]);
        draw2<W>(s.seed, key, prm->iter, r, 0, 2, dr[3], dr[5]);
        draw2<W>(s.seed, key, prm->iter, r, 0, 3, dr[2], dr[4]);
        const Excit ex = gen_excit<W>(s, key, dr, prm->iter, r, 0, nullptr);
        if (ex.ic) {
          const int32_t e[4] = {ex.e0, ex.e1, ex.e2, ex.e3};
          double p2 = calc_pgen<W>(s, key, ex.ic, e);
          if (off) p2 *= 1.0 + 1e-6;
          tests = 1;
          const double a = fabs(ex.pgen), b = fabs(p2);
          bad = !(fabs(ex.pgen - p2) <= 1e-12 * (a > b ? a : b));
        }
      }
    }
    chk_add(chk, kChkPgenN, tests);
    chk_add(chk, kChkPgenBad, bad);
  }
}

// the verdict (one thread): CalcHashTableStats's AnnihilatedDet == HolesInList
// (load_balancer.fpp:796-804) as holes == free-queue length, live slots ==
// the live count, and no bad slot or pgen.  A failure halts the chain like any
// other error of the iteration (NGPU_ERR_CHECK).
__global__ void k_check_final(Counters* cnt, unsigned long long* chk, ngpu_iter_out_t* out, ngpu_iter_out_t* cyc,
                              const IterParams* prm) {
  const unsigned long long queue = cnt->free_tail - cnt->free_head;
  const bool holes_bad = chk[kChkHoles] != queue, live_bad = chk[kChkLive] != cnt->n_live;
  chk[kChkCum + 0] += 1;  // iterations checked
  chk[kChkCum + 1] += chk[kChkNel];
  chk[kChkCum + 2] += chk[kChkHash];
  chk[kChkCum + 3] += holes_bad ? 1 : 0;
  chk[kChkCum + 4] += live_bad ? 1 : 0;
  chk[kChkCum + 5] += chk[kChkPgenN];
  chk[kChkCum + 6] += chk[kChkPgenBad];
  const bool bad = chk[kChkNel] || chk[kChkHash] || chk[kChkPgenBad] || holes_bad || live_bad;
  for (int i = 0; i < kChkCum; ++i) chk[i] = 0;
  if (bad && !cnt->halt) {
    const unsigned int m = 1u << NGPU_ERR_CHECK;
    cnt->halt = m;
    cnt->err_iter = prm->iter;
    out->err |= (int32_t)m;
    cyc->err |= (int32_t)m;
  }
}

template <int W>
void check_chain(ngpu_t* h, bool gen) {
  if (h->checks & NGPU_CHECK_TABLE)
    launch(h, -1, -1, k_check_table<W>, dim3(h->nb_grid), dim3(kBlock), h->ds, h->wk, h->cnt, h->d_chk);
  if (gen && (h->checks & NGPU_CHECK_PGEN))
    launch(h, -1, -1, k_check_pgen<W>, dim3(h->nb_grid), dim3(kBlock), h->ds, h->wk, (const Counters*)h->cnt,
           h->d_chk, (const IterParams*)h->dparams, (h->checks & NGPU_CHECK_SELFTEST) ? 1 : 0);
  launch(h, -1, -1, k_check_final, dim3(1), dim3(1), h->cnt, h->d_chk, h->dout, h->dcyc,
         (const IterParams*)h->dparams);
}

template <int W>
__global__ void k_batch_pgen(DevSys s, const uint64_t* d, const int32_t* ic, const int32_t* ex, int64_t n, double* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = calc_pgen<W>(s, d + i * W, ic[i], ex + 4 * i);
}

__global__ void k_batch_die(bool real_coeff, const double* kii, const double* sg, const double* u, int64_t n,
                            double tau, double sft, double* nd) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    nd[i] = attempt_die(real_coeff, tau * (kii[i] - sft), fabs(sg[i]), u[i]);
}

__global__ void k_batch_create(DevSys s, const double* hel, const double* pgen, const double* psg, const double* u,
                               int64_t n, double tau, double* child) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double raw;
    child[i] = attempt_create(s, tau, hel[i], pgen[i], psg[i] >= 0.0 ? 1.0 : -1.0, u[i], raw);
  }
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t n) { return dalloc(&p, n); }
};

int grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (int)g;
}

int upload_rows_to_new_list(ngpu_t* h, const int64_t* rows, int64_t n) {
  const int W = h->W, S = h->S, RL = W + S + 1;
  std::vector<uint64_t> key((size_t)n * W);
  std::vector<double> sg((size_t)n * S);
  std::vector<uint32_t> fl((size_t)n);
  const int* init_bits = h->sys.flag_layout == NGPU_FLAGS_NECI ? kNeciInit : kMneciInit;
  const int* det_bits = h->sys.flag_layout == NGPU_FLAGS_NECI ? kNeciDeterm : kMneciDeterm;
  int maxr = h->sys.flag_layout == NGPU_FLAGS_NECI ? 2 : 16;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t* row = rows + i * RL;
    for (int w = 0; w < W; ++w) key[i * W + w] = (uint64_t)row[w];
    memcpy(&sg[i * S], row + W, sizeof(double) * S);
    uint64_t nf = (uint64_t)row[W + S];
    uint32_t f = 0;
    for (int r = 0; r < S && r < maxr; ++r) {
      if ((nf >> init_bits[r]) & 1ull) f |= 1u << r;
      if ((nf >> det_bits[r]) & 1ull) f |= kFlagCore;
    }
    fl[i] = f;
  }
  HIP_OK(hipMemcpyAsync(h->n_key, key.data(), key.size() * 8, hipMemcpyHostToDevice, h->stream));
  HIP_OK(hipMemcpyAsync(h->n_sgn, sg.data(), sg.size() * 8, hipMemcpyHostToDevice, h->stream));
  HIP_OK(hipMemcpyAsync(h->n_flag, fl.data(), fl.size() * 4, hipMemcpyHostToDevice, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  return NGPU_OK;
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

int ngpu_create(ngpu_t** out, const ngpu_sys_t* sys) {
  if (!out || !sys) return NGPU_ERR_ARG;
  *out = nullptr;
  if (sys->nel < 1 || sys->nel > 64 || sys->nbasis < 2 || sys->nbasis > 127 || sys->nel > sys->nbasis)
    return NGPU_ERR_ARG;
  if (sys->n_replicas < 1 || sys->n_replicas > NGPU_MAX_REPLICAS) return NGPU_ERR_ARG;
  if (sys->max_walkers < 1 || sys->max_walkers > (int64_t)INT32_MAX || sys->max_spawned < 1 ||
      sys->max_spawned > (int64_t)UINT32_MAX || sys->n_hash_buckets < 1 ||
      sys->n_hash_buckets > (int64_t)INT32_MAX / kBucketEntries)  // entries are stored as int32
    return NGPU_ERR_ARG;
  if (sys->model != NGPU_MODEL_HUBBARD && sys->model != NGPU_MODEL_MOLECULAR) return NGPU_ERR_ARG;
  ngpu_t* h = new ngpu_engine();
  h->sys = *sys;
  h->W = sys->nbasis / 64 + 1;
  h->S = sys->n_replicas;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    delete h;
    return NGPU_ERR_HIP;
  }
  h->dev = dev;
  if (int src = scratch_setup(dev)) {  // the allocation-free download's scratch on this device
    delete h;
    return src;
  }
  h->nb_chunk = std::max(256, 4 * prop.multiProcessorCount);
  h->nb_grid = std::max(512, 4 * prop.multiProcessorCount);
  h->nb_merge = std::max(384, 3 * prop.multiProcessorCount);
  h->n_cu = std::max(1, prop.multiProcessorCount);
  h->max_chunk = 16 * h->n_cu;  // buffers sized by the chunk count allow up to this many blocks
  int rc = NGPU_OK;
  auto fail = [&](hipError_t e) {
    if (e != hipSuccess && rc == NGPU_OK) {
      fprintf(stderr, "[neci_gpu] create: %s\n", hipGetErrorString(e));
      rc = NGPU_ERR_HIP;
    }
  };
  fail(hipStreamCreate(&h->stream));
  const int64_t cap = sys->max_walkers, scap = sys->max_spawned;
  const int W = h->W, S = h->S;
  h->scap = scap;
  Walk& wk = h->wk;
  wk.cap = cap;
  wk.L = sys->n_hash_buckets;
  wk.srow = S + 1;
  fail(dalloc(&wk.keys, (size_t)cap * W));
  fail(dalloc(&wk.st, (size_t)cap * wk.srow));
  fail(dalloc(&wk.diagH, (size_t)cap));
  fail(dalloc(&wk.offdiagH, (size_t)cap));
  fail(dalloc(&wk.freeq, (size_t)cap));
  fail(dalloc(&wk.htab, (size_t)wk.L * 8 * W));
  fail(dalloc(&h->att_cnt, (size_t)cap * S));
  fail(dalloc(&h->att_off, (size_t)cap * S + 1));
  fail(dalloc(&h->pinfo, (size_t)cap * S));
  fail(dalloc(&h->blk_part, (size_t)h->max_chunk));
  fail(dalloc(&h->blk_rm, (size_t)h->max_chunk));
  // tiles of attempts with a recorded first parent: attempts up to 4x the
  // spawn capacity (beyond it the spawn kernel searches the offsets itself)
  h->tcap = (4 * scap) / kBlock + 2;
  fail(dalloc(&h->tile_p, (size_t)h->tcap));
  fail(dalloc(&h->sp_rec, (size_t)scap * kRecWords));
  for (int q = 0; q < 2; ++q) {
    fail(dalloc(&h->sk[q], (size_t)scap * W));
    fail(dalloc(&h->si[q], (size_t)scap));
  }
  fail(dalloc(&h->hist, (size_t)h->max_chunk * kRadixBins));
  {
    // C(nbasis, nel) decides the sort key; the table holds C(p, i) for the
    // colex rank of a determinant (p < nbasis, i <= nel)
    const int nb = sys->nbasis, ne = sys->nel;
    std::vector<unsigned __int128> c((size_t)(nb + 1) * (ne + 1), 0);
    for (int p = 0; p <= nb; ++p) {
      c[(size_t)p * (ne + 1)] = 1;
      for (int i = 1; i <= ne && i <= p; ++i) {
        unsigned __int128 v = c[(size_t)(p - 1) * (ne + 1) + i - 1] + c[(size_t)(p - 1) * (ne + 1) + i];
        c[(size_t)p * (ne + 1) + i] = v > ((unsigned __int128)1 << 64) ? ((unsigned __int128)1 << 64) : v;
      }
    }
    unsigned __int128 total = c[(size_t)nb * (ne + 1) + ne];
    if (total <= ((unsigned __int128)1 << 32) && nb * (ne + 1) <= kBinomLds) {
      h->rank_sort = true;
      while (h->rank_bits < 32 && ((unsigned __int128)1 << h->rank_bits) < total) ++h->rank_bits;
      if (h->rank_bits == 0) h->rank_bits = 1;
      std::vector<uint32_t> b((size_t)nb * (ne + 1));
      for (int p = 0; p < nb; ++p)
        for (int i = 0; i <= ne; ++i) b[(size_t)p * (ne + 1) + i] = (uint32_t)c[(size_t)p * (ne + 1) + i];
      fail(dalloc(&h->d_binom, b.size()));
      if (rc == NGPU_OK) fail(hipMemcpy(h->d_binom, b.data(), b.size() * 4, hipMemcpyHostToDevice));
      h->ds.binom = h->d_binom;
      h->ds.nel1 = sys->nel + 1;
      fail(dalloc(&h->rk, (size_t)scap));
    } else if (h->W == 1 && total < ((unsigned __int128)1 << 64) && nb * (ne + 1) <= kBinom64Lds) {
      int bits = 1;
      while (bits < 64 && ((unsigned __int128)1 << bits) < total) ++bits;
      if ((bits + 7) / 8 < (nb + 7) / 8) {
        h->rank64 = true;
        h->rank64_bits = bits;
        // entries above the total belong to no determinant's rank
        std::vector<uint64_t> b((size_t)nb * (ne + 1));
        for (int p = 0; p < nb; ++p)
          for (int i = 0; i <= ne; ++i) {
            unsigned __int128 v = c[(size_t)p * (ne + 1) + i];
            b[(size_t)p * (ne + 1) + i] = v < total ? (uint64_t)v : 0;
          }
        fail(dalloc(&h->d_binom64, b.size()));
        if (rc == NGPU_OK) fail(hipMemcpy(h->d_binom64, b.data(), b.size() * 8, hipMemcpyHostToDevice));
        h->ds.binom64 = h->d_binom64;
        h->ds.nel1 = sys->nel + 1;
      }
    }
  }
  fail(dalloc(&h->digit_tot, (size_t)kRadixBins));
  fail(dalloc(&h->sq, (size_t)scap * 2));  // (parent pair, attempt number)
  int64_t ncap = std::max<int64_t>(scap, cap);
  fail(dalloc(&h->n_key, (size_t)ncap * W));
  fail(dalloc(&h->n_sgn, (size_t)ncap * S));
  fail(dalloc(&h->n_flag, (size_t)ncap));
  fail(dalloc(&h->cnt, 1));
  fail(dalloc(&h->dout, 1));
  fail(dalloc(&h->dpart, (size_t)kStatCopies * kStatStride));
  fail(dalloc(&h->dcyc, 1));
  fail(dalloc(&h->own_params, 1));
  h->dparams = h->own_params;
  fail(hipHostMalloc((void**)&h->hout, sizeof(ngpu_iter_out_t), hipHostMallocDefault));
  fail(hipHostMalloc((void**)&h->hcnt, sizeof(Counters), hipHostMallocDefault));
  if (rc == NGPU_OK) {
    fail(hipMemsetAsync(wk.htab, 0xFF, (size_t)wk.L * 8 * W * sizeof(uint64_t), h->stream));
    fail(hipMemsetAsync(h->cnt, 0, sizeof(Counters), h->stream));
    fail(hipMemsetAsync(h->dout, 0, sizeof(ngpu_iter_out_t), h->stream));
    fail(hipMemsetAsync(h->dpart, 0, sizeof(double) * kStatCopies * kStatStride, h->stream));
    fail(hipMemsetAsync(h->dcyc, 0, sizeof(ngpu_iter_out_t), h->stream));
    fail(hipMemsetAsync(wk.st, 0, (size_t)cap * wk.srow * sizeof(double), h->stream));
    fail(hipStreamSynchronize(h->stream));
    memset(h->hout, 0, sizeof(ngpu_iter_out_t));
  }
  h->ds.nel = sys->nel;
  h->ds.nbasis = sys->nbasis;
  h->ds.S = S;
  h->ds.model = sys->model;
  h->ds.t_init = sys->t_trunc_initiator;
  h->ds.t_real = sys->t_all_real_coeff;
  h->ds.t_rsc = sys->t_real_spawn_cutoff;
  h->ds.init_walk_no = sys->initiator_walk_no;
  h->ds.rsc = sys->real_spawn_cutoff;
  h->ds.occ_thresh = sys->occupied_thresh;
  h->ds.avmc = sys->av_mc_excits;
  h->ds.seed = sys->seed;
  h->ds.L = sys->n_hash_buckets;
  h->ds.pdoubles = 1.0;
  h->ds.spawn_fp32 = 0;
  h->ds.rec_words = kRecWords;
  if (rc == NGPU_OK) {
    h->ev_ok = true;
    for (int q = 0; q < kEvRing; ++q)
      for (int s = 0; s < kStages; ++s)
        for (int e = 0; e < 2; ++e)
          if (hipEventCreate(&h->ev[q][s][e]) != hipSuccess) h->ev_ok = false;
  }
  if (rc != NGPU_OK) {
    ngpu_destroy(h);
    return rc;
  }
  *out = h;
  return NGPU_OK;
}

int ngpu_destroy(ngpu_t* h) {
  if (!h) return NGPU_ERR_ARG;
  if (h->batch) return NGPU_ERR_STATE;  // ngpu_batch_destroy first
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  void* ptrs[] = {h->wk.keys, h->wk.st, h->wk.diagH, h->wk.offdiagH, h->wk.freeq,
                  h->wk.htab, h->att_cnt, h->att_off, h->pinfo, h->blk_part, h->blk_rm, h->sp_rec,
                  h->sk[0], h->sk[1], h->si[0], h->si[1], h->hist, h->digit_tot, h->sq, h->n_key,
                  h->n_sgn, h->n_flag, h->u_key, h->u_sgn, h->u_flag, h->u_count, h->cnt, h->dout, h->dpart, h->dcyc, h->own_params, h->d_rh, h->d_nbr, h->d_tmat, h->d_umat, h->d_jk, h->d_jk0, h->d_fref, h->d_chk, h->d_binom, h->d_binom64, h->rk, h->tile_p, h->core_keys, h->core_slot, h->core_rowptr, h->core_cols,
                  h->core_vals, h->core_x, h->core_y, h->d_ab_tab,
                  h->d_ab_prob, h->d_ab_ok, h->d_j_tab, h->d_j_prob, h->d_j_ok, h->d_ab_tgt, h->d_i_tab, h->d_i_prob,
                  h->d_a_tab, h->d_a_prob, h->d_a_ok, h->d_b_tab, h->d_b_prob, h->d_b_ok};
  for (void* p : ptrs) table_release(p);
  if (h->hout) (void)hipHostFree(h->hout);
  if (h->hcnt) (void)hipHostFree(h->hcnt);
  if (h->ev_ok)
    for (int q = 0; q < kEvRing; ++q)
      for (int s = 0; s < kStages; ++s)
        for (int e = 0; e < 2; ++e) (void)hipEventDestroy(h->ev[q][s][e]);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return NGPU_OK;
}

extern "C++" {
namespace {
// fref[(p-1)*nsp + b-1] = sltcnd_1(ref, p, b with p's spin) for every spin
// orbital p (the reference-relative heat-bath weights start from it)
template <int W>
__global__ void k_fock_ref(DevSys s, double* out) {
  const int n = s.nbasis * s.nsp;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const int p = 1 + t / s.nsp, b = 1 + t % s.nsp;
    out[t] = sltcnd_1<W>(s, s.ref, p, 2 * b - (is_beta(p) ? 1 : 0));
  }
}

// (re)builds fref once both the integrals and the reference are set
int refresh_fref(ngpu_t* h) {
  if (h->sys.model != NGPU_MODEL_MOLECULAR || !h->have_model || !h->have_ref) return NGPU_OK;
  const size_t n = (size_t)h->sys.nbasis * h->ds.nsp;
  if (h->stream) HIP_OK(hipStreamSynchronize(h->stream));
  table_release(h->d_fref);
  h->d_fref = nullptr;
  h->ds.fref = nullptr;
  DevBuf<double> tmp;
  HIP_OK(tmp.alloc(n));
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
  if (h->W == 1)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fock_ref<1>), grid, dim3(kBlock), 0, h->stream, h->ds, tmp.p);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fock_ref<2>), grid, dim3(kBlock), 0, h->stream, h->ds, tmp.p);
  HIP_OK(hipGetLastError());
  std::vector<double> host(n);
  HIP_OK(hipMemcpyAsync(host.data(), tmp.p, n * 8, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  if (int rc = table_upload(&h->d_fref, host.data(), n)) return rc;
  h->ds.fref = h->d_fref;
  return NGPU_OK;
}
}  // namespace
}  // extern "C++"

int ngpu_set_integrals(ngpu_t* h, const double* tmat2d, const double* umat, int64_t umat_len, double ecore) {
  if (h) h->graph_dirty = true;  // kernel arguments of the captured chain change
  if (!h || !tmat2d || !umat || umat_len < 1 || h->sys.model != NGPU_MODEL_MOLECULAR) return NGPU_ERR_ARG;
  int64_t nsp = h->sys.nbasis / 2;
  int64_t npair = nsp * (nsp + 1) / 2;
  if (umat_len < npair * (npair + 1) / 2) return NGPU_ERR_ARG;
  size_t nt = (size_t)h->sys.nbasis * h->sys.nbasis;
  if (int rc = table_upload(&h->d_tmat, tmat2d, nt)) return rc;
  if (int rc = table_upload(&h->d_umat, umat, (size_t)umat_len)) return rc;
  {
    // (Coulomb, exchange) rows of the heat-bath single-excitation weights
    std::vector<double> jk((size_t)nsp * nsp * nsp * 2);
    for (int a = 1; a <= nsp; ++a)
      for (int j = 1; j <= nsp; ++j)
        for (int b = 1; b <= nsp; ++b) {
          size_t e = 2 * (((size_t)(a - 1) * nsp + (j - 1)) * nsp + (b - 1));
          jk[e] = umat[umat_ind(a, j, b, j) - 1];
          jk[e + 1] = umat[umat_ind(a, j, j, b) - 1];
        }
    if (int rc = table_upload(&h->d_jk, jk.data(), jk.size())) return rc;
    // (J, K) pairs of the diagonal elements
    std::vector<double> jk0((size_t)nsp * nsp * 2);
    for (int i = 1; i <= nsp; ++i)
      for (int j = 1; j <= nsp; ++j) {
        size_t e = 2 * ((size_t)(i - 1) * nsp + (j - 1));
        jk0[e] = umat[umat_ind(i, j, i, j) - 1];
        jk0[e + 1] = umat[umat_ind(i, j, j, i) - 1];
      }
    if (int rc = table_upload(&h->d_jk0, jk0.data(), jk0.size())) return rc;
  }
  h->ds.tmat = h->d_tmat;
  h->ds.umat = h->d_umat;
  h->ds.jk = h->d_jk;
  h->ds.jk0 = h->d_jk0;
  h->ds.nsp = (int)nsp;
  h->ds.ecore = ecore;
  h->have_model = true;
  return refresh_fref(h);
}

int ngpu_set_lattice(ngpu_t* h, const int32_t* nbr, int32_t nsites, int32_t maxnbr, double bhub, double uhub) {
  if (h) h->graph_dirty = true;  // kernel arguments of the captured chain change
  if (!h || !nbr || h->sys.model != NGPU_MODEL_HUBBARD) return NGPU_ERR_ARG;
  if (nsites * 2 != h->sys.nbasis || maxnbr < 1 || maxnbr > 16) return NGPU_ERR_ARG;
  for (int64_t i = 0; i < (int64_t)nsites * maxnbr; ++i)
    if (nbr[i] < 0 || nbr[i] > nsites) return NGPU_ERR_ARG;
  if (int rc = table_upload(&h->d_nbr, nbr, (size_t)nsites * maxnbr)) return rc;
  h->ds.nbr = h->d_nbr;
  h->ds.nsites = nsites;
  h->ds.maxnbr = maxnbr;
  h->ds.bhub = bhub;
  h->ds.uhub = uhub;
  h->have_model = true;
  return NGPU_OK;
}

// GAS_doubles_PCHB_spinorb_compute_samplers (gasci_pchb_doubles_spinorb_fastweighted.fpp:280-347)
// for the FCI case and the J sampler of init_PC_WeightedParticles_t
// (gasci_pchb_doubles_select_particles.fpp:254-301)
extern "C++" {
namespace {
// one sampler row: alias table and probabilities; false (row left empty) when
// the weights sum to zero (init_AliasSampler_t, aliasSampling.F90:376-398)
bool alias_row(const std::vector<double>& w, AliasEnt* tab, double* prob) {
  double sw = 0.0;
  for (double x : w) sw += x;
  if (fabs(sw) <= kEps) return false;
  build_alias(w, tab, prob);
  return true;
}

template <typename T>
int upload_table(T** dst, const std::vector<T>& src) {
  return table_upload(dst, src.data(), src.size());
}
}  // namespace
}  // extern "C++"

// the reference's spin counts and pick_biased_elecs's pair probabilities for
// them (every determinant of the run has the reference's spin counts)
static void update_pair_probs(ngpu_t* h) {
  const uint64_t kA = 0xAAAAAAAAAAAAAAAAull, kB = 0x5555555555555555ull;
  int na = 0, nb = 0;
  for (int w = 0; w < h->W; ++w) {
    na += __builtin_popcountll(h->ds.ref[w] & kA);
    nb += __builtin_popcountll(h->ds.ref[w] & kB);
  }
  h->ds.nalpha_ref = na;
  h->ds.nbeta_ref = nb;
  const int par = na * (na - 1) / 2 + nb * (nb - 1) / 2, ab = na * nb;
  const double pb = h->ds.pparallel;
  h->ds.pg_par_ref = par > 0 ? pb / (double)par : 0.0;
  h->ds.pg_anti_ref = ab > 0 ? (1.0 - pb) / (double)ab : 0.0;
}

int ngpu_set_excitgen_pchb(ngpu_t* h, int32_t singles, int32_t particles, int32_t holes, double psingles,
                           double pparallel) {
  if (h) h->graph_dirty = true;  // kernel arguments of the captured chain change
  if (!h || h->sys.model != NGPU_MODEL_MOLECULAR || !h->have_model) return NGPU_ERR_ARG;
  if (singles != NGPU_SINGLES_HEATBATH) return NGPU_ERR_ARG;
  if (particles < NGPU_PART_UNIF_UNIF || particles > NGPU_PART_FULL_FULL) return NGPU_ERR_ARG;
  if (holes != NGPU_HOLES_FAST_FAST && holes != NGPU_HOLES_FULL_FULL) return NGPU_ERR_ARG;
  if (!(psingles >= 0.0 && psingles <= 1.0) || !(pparallel >= 0.0 && pparallel <= 1.0)) return NGPU_ERR_ARG;
  h->ds.uniform = 0;
  const int n = h->sys.nbasis;
  const int ijmax = fuse_index(n, n);
  std::vector<double> tmat((size_t)n * n);
  int64_t nsp = n / 2, npair = nsp * (nsp + 1) / 2, ulen = npair * (npair + 1) / 2;
  std::vector<double> umat((size_t)ulen);
  HIP_OK(hipMemcpy(tmat.data(), h->d_tmat, tmat.size() * 8, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(umat.data(), h->d_umat, umat.size() * 8, hipMemcpyDeviceToHost));
  DevSys hs = h->ds;
  hs.tmat = tmat.data();
  hs.umat = umat.data();
  // |<IJ||AB>| of a spin-allowed double, 0 otherwise (get_PCHB_weight,
  // gasci_pchb_doubles_select_particles.fpp:596-609)
  auto weight = [&](int I, int J, int A, int B) {
    int t1 = std::min(A, B), t2 = std::max(A, B);
    int ms_s = (is_beta(I) ? -1 : 1) + (is_beta(J) ? -1 : 1);
    int ms_t = (is_beta(t1) ? -1 : 1) + (is_beta(t2) ? -1 : 1);
    return ms_s == ms_t ? fabs(sltcnd_2(hs, I, J, t1, t2)) : 0.0;
  };
  std::vector<double> ijw((size_t)n * n, 0.0);
  if (holes == NGPU_HOLES_FAST_FAST) {
    // GAS_doubles_PCHB_spinorb_compute_samplers (gasci_pchb_doubles_spinorb_fastweighted.fpp:280-347)
    std::vector<AliasEnt> ab_tab((size_t)ijmax * ijmax);
    std::vector<double> ab_prob((size_t)ijmax * ijmax, 0.0);
    std::vector<uint8_t> ab_ok(ijmax, 0);
    std::vector<double> w(ijmax);
    for (int I = 1; I <= n; ++I)
      for (int J = I + 1; J <= n; ++J) {
        std::fill(w.begin(), w.end(), 0.0);
        for (int A = 1; A <= n; ++A) {
          if (A == I || A == J) continue;
          for (int B = 1; B <= n; ++B) {
            if (B == I || B == J || B == A) continue;
            w[fuse_index(A, B) - 1] = weight(I, J, A, B);
          }
        }
        double sw = 0.0;
        for (double x : w) sw += x;
        int IJ = fuse_index(I, J);
        ab_ok[IJ - 1] = alias_row(w, &ab_tab[(size_t)(IJ - 1) * ijmax], &ab_prob[(size_t)(IJ - 1) * ijmax]) ? 1 : 0;
        ijw[(size_t)(I - 1) + (size_t)(J - 1) * n] = sw;
        ijw[(size_t)(J - 1) + (size_t)(I - 1) * n] = sw;
      }
    std::vector<int32_t> tgt(2 * (size_t)ijmax);
    for (int A = 1; A <= n; ++A)
      for (int B = 1; B <= A; ++B) {
        int AB = fuse_index(A, B);
        tgt[2 * (AB - 1)] = B;
        tgt[2 * (AB - 1) + 1] = A;
      }
    if (int rc = upload_table(&h->d_ab_tab, ab_tab)) return rc;
    if (int rc = upload_table(&h->d_ab_prob, ab_prob)) return rc;
    if (int rc = upload_table(&h->d_ab_ok, ab_ok)) return rc;
    if (int rc = upload_table(&h->d_ab_tgt, tgt)) return rc;
  } else {
    // GAS_doubles_PCHB_compute_samplers of the fully weighted holes
    // (gasci_pchb_doubles_spinorb_fullyweighted.fpp:317-391): per (I < J), per
    // A, the B row |<IJ||AB>|; the A row its sums; IJ_weights their sum
    std::vector<AliasEnt> a_tab((size_t)ijmax * n), b_tab((size_t)ijmax * n * n);
    std::vector<double> a_prob((size_t)ijmax * n, 0.0), b_prob((size_t)ijmax * n * n, 0.0);
    std::vector<uint8_t> a_ok(ijmax, 0), b_ok((size_t)ijmax * n, 0);
    std::vector<double> wa(n), wb(n);
    for (int I = 1; I <= n; ++I)
      for (int J = I + 1; J <= n; ++J) {
        const int IJ = fuse_index(I, J);
        std::fill(wa.begin(), wa.end(), 0.0);
        for (int A = 1; A <= n; ++A) {
          if (A == I || A == J) continue;
          std::fill(wb.begin(), wb.end(), 0.0);
          for (int B = 1; B <= n; ++B) {
            if (B == I || B == J || B == A) continue;
            wb[B - 1] = weight(I, J, A, B);
          }
          const size_t row = (size_t)(IJ - 1) * n + (size_t)(A - 1);
          b_ok[row] = alias_row(wb, &b_tab[row * n], &b_prob[row * n]) ? 1 : 0;
          double sw = 0.0;
          for (double x : wb) sw += x;
          wa[A - 1] = sw;
        }
        a_ok[IJ - 1] = alias_row(wa, &a_tab[(size_t)(IJ - 1) * n], &a_prob[(size_t)(IJ - 1) * n]) ? 1 : 0;
        double sw = 0.0;
        for (double x : wa) sw += x;
        ijw[(size_t)(I - 1) + (size_t)(J - 1) * n] = sw;
        ijw[(size_t)(J - 1) + (size_t)(I - 1) * n] = sw;
      }
    if (int rc = upload_table(&h->d_a_tab, a_tab)) return rc;
    if (int rc = upload_table(&h->d_a_prob, a_prob)) return rc;
    if (int rc = upload_table(&h->d_a_ok, a_ok)) return rc;
    if (int rc = upload_table(&h->d_b_tab, b_tab)) return rc;
    if (int rc = upload_table(&h->d_b_prob, b_prob)) return rc;
    if (int rc = upload_table(&h->d_b_ok, b_ok)) return rc;
  }
  // particle samplers (init_PC_WeightedParticles_t, gasci_pchb_doubles_select_particles.fpp:254-301):
  // J rows weights(:, I), the I row sum(weights, dim=1)
  std::vector<AliasEnt> j_tab((size_t)n * n), i_tab(n);
  std::vector<double> j_prob((size_t)n * n, 0.0), i_prob(n, 0.0), wi(n);
  std::vector<uint8_t> j_ok(n, 0);
  for (int I = 1; I <= n; ++I) {
    std::vector<double> row(ijw.begin() + (size_t)(I - 1) * n, ijw.begin() + (size_t)I * n);
    j_ok[I - 1] = alias_row(row, &j_tab[(size_t)(I - 1) * n], &j_prob[(size_t)(I - 1) * n]) ? 1 : 0;
    double sw = 0.0;
    for (double x : row) sw += x;
    wi[I - 1] = sw;
  }
  const bool i_ok = alias_row(wi, i_tab.data(), i_prob.data());
  if (int rc = upload_table(&h->d_j_tab, j_tab)) return rc;
  if (int rc = upload_table(&h->d_j_prob, j_prob)) return rc;
  if (int rc = upload_table(&h->d_j_ok, j_ok)) return rc;
  if (int rc = upload_table(&h->d_i_tab, i_tab)) return rc;
  if (int rc = upload_table(&h->d_i_prob, i_prob)) return rc;
  h->ds.ab_tab = holes == NGPU_HOLES_FAST_FAST ? h->d_ab_tab : nullptr;
  h->ds.ab_prob = holes == NGPU_HOLES_FAST_FAST ? h->d_ab_prob : nullptr;
  h->ds.ab_ok = holes == NGPU_HOLES_FAST_FAST ? h->d_ab_ok : nullptr;
  h->ds.ab_tgt = holes == NGPU_HOLES_FAST_FAST ? h->d_ab_tgt : nullptr;
  h->ds.a_tab = holes == NGPU_HOLES_FULL_FULL ? h->d_a_tab : nullptr;
  h->ds.a_prob = holes == NGPU_HOLES_FULL_FULL ? h->d_a_prob : nullptr;
  h->ds.a_ok = holes == NGPU_HOLES_FULL_FULL ? h->d_a_ok : nullptr;
  h->ds.b_tab = holes == NGPU_HOLES_FULL_FULL ? h->d_b_tab : nullptr;
  h->ds.b_prob = holes == NGPU_HOLES_FULL_FULL ? h->d_b_prob : nullptr;
  h->ds.b_ok = holes == NGPU_HOLES_FULL_FULL ? h->d_b_ok : nullptr;
  h->ds.j_tab = h->d_j_tab;
  h->ds.j_prob = h->d_j_prob;
  h->ds.j_ok = h->d_j_ok;
  h->ds.i_tab = h->d_i_tab;
  h->ds.i_prob = h->d_i_prob;
  h->ds.i_ok = i_ok ? 1 : 0;
  h->ds.holes = holes;
  h->ds.ijmax = ijmax;
  h->ds.singles = singles;
  h->ds.particles = particles;
  h->ds.psingles = psingles;
  h->ds.pdoubles = 1.0 - psingles;
  h->ds.pparallel = pparallel;
  update_pair_probs(h);
  h->have_excitgen = true;
  return NGPU_OK;
}

int ngpu_set_excitgen(ngpu_t* h, int32_t singles, int32_t particles, double psingles, double pparallel) {
  return ngpu_set_excitgen_pchb(h, singles, particles, NGPU_HOLES_FAST_FAST, psingles, pparallel);
}

int ngpu_set_excitgen_uniform(ngpu_t* h, double psingles) {
  if (h) h->graph_dirty = true;  // kernel arguments of the captured chain change
  if (!h || h->sys.model != NGPU_MODEL_MOLECULAR || !h->d_umat) return NGPU_ERR_ARG;
  if (!(psingles >= 0.0 && psingles <= 1.0)) return NGPU_ERR_ARG;
  if ((h->sys.nbasis & 1) || h->sys.nel < 2) return NGPU_ERR_ARG;
  h->ds.uniform = 1;
  h->ds.psingles = psingles;
  h->ds.pdoubles = 1.0 - psingles;
  h->have_excitgen = true;
  return NGPU_OK;
}

int ngpu_set_spawn_precision(ngpu_t* h, int32_t fp32) {
  if (!h || (fp32 != 0 && fp32 != 1)) return NGPU_ERR_ARG;
  if (h->batch) return NGPU_ERR_STATE;  // the batch's graph holds the record format
  h->graph_dirty = true;  // kernel arguments of the captured chain change
  h->ds.spawn_fp32 = fp32;
  h->ds.rec_words = fp32 ? h->W + 1 : kRecWords;
  return NGPU_OK;
}

int ngpu_set_hash(ngpu_t* h, const int32_t* random_hash2) {
  if (h) h->graph_dirty = true;  // kernel arguments of the captured chain change
  if (!h || !random_hash2) return NGPU_ERR_ARG;
  for (int i = 0; i < h->sys.nbasis; ++i)
    if (random_hash2[i] < 1 || random_hash2[i] > 1000 * h->sys.nbasis) return NGPU_ERR_ARG;
  if (!h->d_rh) HIP_OK(dalloc(&h->d_rh, (size_t)h->sys.nbasis));
  HIP_OK(hipMemcpy(h->d_rh, random_hash2, (size_t)h->sys.nbasis * 4, hipMemcpyHostToDevice));
  h->ds.rh = h->d_rh;
  h->have_hash = true;
  return NGPU_OK;
}

int ngpu_set_reference(ngpu_t* h, const int64_t* ilut_ref, double hii) {
  if (h) h->graph_dirty = true;  // kernel arguments of the captured chain change
  if (!h || !ilut_ref) return NGPU_ERR_ARG;
  int nel = 0;
  for (int w = 0; w < h->W; ++w) nel += __builtin_popcountll((uint64_t)ilut_ref[w]);
  if (nel != h->sys.nel) return NGPU_ERR_ARG;
  h->ds.ref[0] = (uint64_t)ilut_ref[0];
  h->ds.ref[1] = h->W > 1 ? (uint64_t)ilut_ref[1] : 0;
  h->ds.hii = hii;
  h->have_ref = true;
  update_pair_probs(h);
  return refresh_fref(h);
}

int ngpu_upload_walkers(ngpu_t* h, const int64_t* rows, int64_t n) {
  int rc = ready(h);
  if (rc) return rc;
  if (n < 0 || (n > 0 && !rows)) return NGPU_ERR_ARG;
  if (n == 0) return NGPU_OK;
  int64_t ncap = std::max<int64_t>(h->scap, h->wk.cap);
  if (n > ncap) return NGPU_ERR_WALKERS_FULL;
  for (int64_t i = 0; i < n; ++i) {
    int nel = 0;
    for (int w = 0; w < h->W; ++w) nel += __builtin_popcountll((uint64_t)rows[i * (h->W + h->S + 1) + w]);
    if (nel != h->sys.nel) return NGPU_ERR_ARG;
  }
  rc = upload_rows_to_new_list(h, rows, n);
  if (rc) return rc;
  hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, h->stream, &h->cnt->n_new, (unsigned long long)n);
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, h->stream, &h->cnt->err, 0u);
  hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, h->stream, &h->cnt->n_fail, 0ull);
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, h->stream, &h->cnt->ins_refused, 0u);
  {
    auto kern = h->W == 1 ? (h->S == 1 ? k_insert<1, 1> : k_insert<1, kMaxRep>)
                          : (h->S == 1 ? k_insert<2, 1> : k_insert<2, kMaxRep>);
    hipLaunchKernelGGL(kern, dim3(h->nb_grid), dim3(kBlock), 0, h->stream, h->ds, h->wk, h->cnt, h->n_key,
                       h->n_sgn, h->n_flag, &h->cnt->n_new);
  }
  hipLaunchKernelGGL(k_post_insert, dim3(1), dim3(1), 0, h->stream, h->cnt, &h->cnt->n_new, h->wk.cap);
  if (h->W == 1) hash_upkeep<1>(h);
  else hash_upkeep<2>(h);
  HIP_OK(hipMemcpyAsync(h->hcnt, h->cnt, sizeof(Counters), hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  // iterations start from empty append counters and statistics rows
  hipLaunchKernelGGL(k_begin_iter, dim3(1), dim3(kBlock), 0, h->stream, h->dout, h->dpart, h->cnt, h->dparams,
                     IterParams{}, 0);
  HIP_OK(hipGetLastError());
  if (h->hcnt->halt) return err_from_mask(h->hcnt->halt);  // halted: nothing was inserted
  return err_from_mask(h->hcnt->err);
}

extern "C++" {
namespace {
// ascending (key, index) radix sort of n elements with the spawn sort's
// kernels: (word, shift) passes, the first reading kin (stride kstride) and
// iin (nullptr: the identity), ping-ponging through kb[2], ib[2]; returns the
// buffer pair holding the result
template <typename K, int NW, bool FLIP, int IT>
int radix_sort_pairs_in(ngpu_t* h, const K* kin, int kstride, const uint32_t* iin,
                        const std::vector<std::pair<int, int>>& passes, const unsigned long long* n_ptr, int64_t n,
                        K* const* kb, uint32_t* const* ib) {
  int cur = -1;
  for (size_t p = 0; p < passes.size(); ++p) {
    const int out = (int)(p & 1);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_radix_hist<K, NW, FLIP, IT, 8>), dim3(h->nb_chunk), dim3(kBlock), 0,
                       h->stream, kin, kstride, n_ptr, n, passes[p].first, passes[p].second, h->hist);
    hipLaunchKernelGGL(k_radix_scan, dim3(256), dim3(kBlock), 0, h->stream, h->hist, h->nb_chunk, h->digit_tot);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_radix_scatter<K, NW, FLIP, IT, 8>), dim3(h->nb_chunk), dim3(kBlock), 0,
                       h->stream, kin, kstride, iin, kb[out], ib[out], n_ptr, n, passes[p].first, passes[p].second,
                       (const uint32_t*)h->hist, (const uint32_t*)h->digit_tot);
    kin = kb[out];
    kstride = NW;
    iin = ib[out];
    cur = out;
  }
  return cur;
}
// the same in the engine's spawn-sort buffers (n <= the spawn capacity)
template <typename K, int NW, bool FLIP, int IT>
int radix_sort_pairs(ngpu_t* h, const K* kin, int kstride, const uint32_t* iin,
                     const std::vector<std::pair<int, int>>& passes, const unsigned long long* n_ptr, int64_t n) {
  K* kb[2] = {(K*)h->sk[0], (K*)h->sk[1]};
  return radix_sort_pairs_in<K, NW, FLIP, IT>(h, kin, kstride, iin, passes, n_ptr, n, kb, h->si);
}

// ---- the core Hamiltonian from its connections ----
// Two determinants with excitation level <= 2 share nel - 2 electrons.  Each
// core determinant is listed once under every (nel - 2)-electron subset of
// its electrons; sorting the list by subset groups the connected pairs, and
// a pair (i, j) of a group is kept in the one group whose subset is the
// canonical one for the pair: their common electrons less the lowest ones, so
// that each connection is found exactly once (level 2: the common electrons
// themselves; level 1: less the lowest).  The pairs and the diagonal are
// sorted by (row, column) and H_ij is evaluated per pair, as the all-pairs
// scan does (off-diagonal zeros dropped), so both builds give the same CSR.
// The role of calc_determ_hamil_opt (fast_determ_hamil.F90:892-1548), which
// joins the core determinants through their alpha and beta strings.

template <int W>
__global__ void k_cs_subsets(const uint64_t* keys, int64_t ncore, int npair, uint64_t* sub, uint32_t* idx) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ncore; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t d[W];
#pragma unroll
    for (int w = 0; w < W; ++w) d[w] = keys[i * W + w];
    int64_t o = i * npair;
    // electron pairs (a < b) in ascending order of a, then b
#pragma unroll
    for (int wa = 0; wa < W; ++wa)
      for (uint64_t xa = d[wa]; xa; xa &= xa - 1) {
        const int ba = ctz64(xa);
        uint64_t rest = xa & (xa - 1);
#pragma unroll
        for (int wb = 0; wb < W; ++wb) {
          if (wb < wa) continue;
          for (uint64_t xb = (wb == wa ? rest : d[wb]); xb; xb &= xb - 1) {
            const int bb = ctz64(xb);
#pragma unroll
            for (int w = 0; w < W; ++w) {
              uint64_t v = d[w];
              if (w == wa) v &= ~(1ull << ba);
              if (w == wb) v &= ~(1ull << bb);
              sub[o * W + w] = v;
            }
            idx[o] = (uint32_t)i;
            ++o;
          }
        }
      }
  }
}

// group starts of the sorted subset list: flag 1 where the subset changes
template <int W>
__global__ void k_cs_starts(const uint64_t* sub, int64_t m, uint32_t* start) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m; t += (int64_t)gridDim.x * blockDim.x) {
    bool s = t == 0;
#pragma unroll
    for (int w = 0; w < W; ++w) s = s || sub[t * W + w] != sub[(t - 1 > 0 ? t - 1 : 0) * W + w];
    start[t] = s ? 1u : 0u;
  }
}

// one wave per group: the canonical connected pairs (i, j), i != j, as
// (i << 32 | j); pass 0 counts them, pass 1 writes them (wave-aggregated
// appends)
template <int W>
__global__ void __launch_bounds__(kBlock) k_cs_pairs(const uint64_t* sub, const uint32_t* idx, int64_t m,
                                                     const int64_t* gstart, int64_t ngroups, const uint64_t* keys,
                                                     int nel, int pass, unsigned long long* total, uint64_t* out) {
  const int lane = lane_id();
  const int64_t nwaves = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; g < ngroups; g += nwaves) {
    const int64_t g0 = gstart[g], g1 = g + 1 < ngroups ? gstart[g + 1] : m;
    const int64_t n = g1 - g0;
    const uint64_t* S = sub + g0 * W;
    for (int64_t base = 0; base < n * n; base += kWave) {
      const int64_t q = base + lane;
      bool emit = false;
      uint64_t rec = 0;
      if (q < n * n) {
        const int64_t a = q / n, b = q - a * n;
        if (a != b) {
          const uint32_t i = idx[g0 + a], j = idx[g0 + b];
          uint64_t I[W];
          int c = 0;
#pragma unroll
          for (int w = 0; w < W; ++w) {
            I[w] = keys[(int64_t)i * W + w] & keys[(int64_t)j * W + w];
            c += popc64(I[w]);
          }
          // the canonical subset: the common electrons less their lowest c - (nel - 2)
          for (int drop = c - (nel - 2); drop > 0; --drop) {
#pragma unroll
            for (int w = 0; w < W; ++w)
              if (I[w]) {
                I[w] &= I[w] - 1;
                break;
              }
          }
          bool same = c <= nel - 1;  // c == nel would be a repeated determinant
#pragma unroll
          for (int w = 0; w < W; ++w) same = same && I[w] == S[w];
          emit = same;
          rec = ((uint64_t)i << 32) | (uint64_t)j;
        }
      }
      const uint64_t mask = __ballot(emit);
      if (!mask) continue;
      const int leader = __ffsll((unsigned long long)mask) - 1;
      unsigned long long pos = 0;
      if (lane == leader) pos = atomicAdd(total, (unsigned long long)popc64(mask));
      pos = __shfl(pos, leader);
      if (pass && emit) out[pos + popc64(mask & lanemask_lt())] = rec;
    }
  }
}

// the diagonal (i, i) of every row
__global__ void k_cs_diag(int64_t ncore, uint64_t* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ncore; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = ((uint64_t)i << 32) | (uint64_t)i;
}

// H_ij of each sorted (row, column) pair, kept when on the diagonal or
// nonzero (the all-pairs scan's rule)
template <int W>
__global__ void k_cs_eval(DevSys s, const uint64_t* keys, const uint64_t* pairs, int64_t np, double* hv,
                          uint32_t* kp) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < np; t += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t pr = pairs[t];
    const int64_t i = (int64_t)(pr >> 32), j = (int64_t)(pr & 0xFFFFFFFFull);
    const uint64_t* di = keys + i * W;
    const double h = (i == j) ? diag_helement<W>(s, di) - s.hii : helement_pair<W>(s, di, keys + j * W);
    hv[t] = h;
    kp[t] = ((i == j) || h != 0.0) ? 1u : 0u;
  }
}

constexpr int64_t kCsTile = 4096;  // pairs per block of the CSR compaction

__global__ void __launch_bounds__(kBlock) k_cs_tilecount(const uint32_t* kp, int64_t np, int64_t* counts) {
  __shared__ int64_t lds[kBlock / kWave];
  const int64_t b0 = (int64_t)blockIdx.x * kCsTile, b1 = std::min<int64_t>(b0 + kCsTile, np);
  int64_t c = 0;
  for (int64_t t = b0 + threadIdx.x; t < b1; t += kBlock) c += kp[t];
  int64_t tot = 0;
  (void)block_exclusive_scan(c, lds, tot);
  if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

// kept pairs to their CSR position (the tile's offset plus the kept pairs
// before them in the tile); the first pair of each row sets its rowptr
template <typename VT>
__global__ void __launch_bounds__(kBlock) k_cs_write(const uint64_t* pairs, const double* hv, const uint32_t* kp,
                                                     int64_t np, const int64_t* offsets, int64_t* rowptr,
                                                     int32_t* cols, VT* vals) {
  __shared__ int64_t lds[kBlock / kWave];
  const int64_t b0 = (int64_t)blockIdx.x * kCsTile;
  int64_t pos0 = offsets[blockIdx.x];
  for (int64_t t0 = 0; t0 < kCsTile && b0 + t0 < np; t0 += kBlock) {
    const int64_t t = b0 + t0 + threadIdx.x;
    const bool keep = t < np && kp[t];
    int64_t tot = 0;
    const int64_t pos = pos0 + block_exclusive_scan(keep ? 1 : 0, lds, tot);
    if (t < np) {
      const uint64_t pr = pairs[t];
      const int64_t i = (int64_t)(pr >> 32);
      if (t == 0 || (int64_t)(pairs[t - 1] >> 32) != i) rowptr[i] = pos;  // a row starts with its first pair
      if (keep) {
        cols[pos] = (int32_t)(pr & 0xFFFFFFFFull);
        vals[pos] = (VT)hv[t];
      }
    }
    pos0 += tot;
  }
}

// the core Hamiltonian by the all-pairs scan (every (i, j) tested): count,
// scan on the host, fill
template <int W>
int core_rows_scan(ngpu_t* h, int64_t ncore, bool fp32, int64_t& nnz) {
  auto grid_rows = [&](int64_t n) { return dim3((unsigned)std::min<int64_t>((n + 3) / 4, 8 * (int64_t)h->nb_grid)); };
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_core_rows<W, double>), grid_rows(ncore), dim3(kBlock), 0, h->stream, h->ds,
                     (const uint64_t*)h->core_keys, ncore, 0, h->core_rowptr + 1, (const int64_t*)nullptr,
                     (int32_t*)nullptr, (double*)nullptr);
  std::vector<int64_t> rp((size_t)ncore + 1, 0);
  HIP_OK(hipMemcpyAsync(rp.data() + 1, h->core_rowptr + 1, (size_t)ncore * 8, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  for (int64_t i = 0; i < ncore; ++i) rp[i + 1] += rp[i];
  nnz = rp[ncore];
  HIP_OK(hipMemcpy(h->core_rowptr, rp.data(), rp.size() * 8, hipMemcpyHostToDevice));
  HIP_OK(dalloc(&h->core_cols, (size_t)nnz));
  if (fp32) {
    float* v = nullptr;
    HIP_OK(dalloc(&v, (size_t)nnz));
    h->core_vals = v;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_core_rows<W, float>), grid_rows(ncore), dim3(kBlock), 0, h->stream, h->ds,
                       (const uint64_t*)h->core_keys, ncore, 1, (int64_t*)nullptr, (const int64_t*)h->core_rowptr,
                       h->core_cols, v);
  } else {
    double* v = nullptr;
    HIP_OK(dalloc(&v, (size_t)nnz));
    h->core_vals = v;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_core_rows<W, double>), grid_rows(ncore), dim3(kBlock), 0, h->stream, h->ds,
                       (const uint64_t*)h->core_keys, ncore, 1, (int64_t*)nullptr, (const int64_t*)h->core_rowptr,
                       h->core_cols, v);
  }
  return NGPU_OK;
}

// the core Hamiltonian from its connections (the subset join above)
template <int W>
int core_rows_join(ngpu_t* h, int64_t ncore, bool fp32, int64_t& nnz) {
  const int nel = h->sys.nel;
  const int64_t npair = (int64_t)nel * (nel - 1) / 2;
  const int64_t m = ncore * npair;
  if (nel < 2 || m > (int64_t)UINT32_MAX || ncore > (int64_t)INT32_MAX) return NGPU_ERR_ARG;
  DevBuf<uint64_t> sk0, sk1;
  DevBuf<uint32_t> si0, si1, start;
  DevBuf<unsigned long long> cnt;
  HIP_OK(sk0.alloc((size_t)m * W));
  HIP_OK(sk1.alloc((size_t)m * W));
  HIP_OK(si0.alloc((size_t)m));
  HIP_OK(si1.alloc((size_t)m));
  HIP_OK(start.alloc((size_t)m));
  HIP_OK(cnt.alloc(2));
  // every core determinant under each of its (nel - 2)-electron subsets, sorted by subset
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cs_subsets<W>), dim3(grid_for(ncore)), dim3(kBlock), 0, h->stream,
                     (const uint64_t*)h->core_keys, ncore, (int)npair, sk0.p, si0.p);
  const unsigned long long mm = (unsigned long long)m;
  HIP_OK(hipMemcpyAsync(cnt.p, &mm, 8, hipMemcpyHostToDevice, h->stream));
  uint64_t* kb[2] = {sk1.p, sk0.p};
  uint32_t* ib[2] = {si1.p, si0.p};
  // the scatter of pass p writes kb[p & 1]; the first pass reads sk0/si0, so
  // the ping-pong starts in the other pair
  const int cur = radix_sort_pairs_in<uint64_t, W, W == 2, 4>(h, sk0.p, W, si0.p, word_passes(h), cnt.p, m, kb, ib);
  const uint64_t* ssub = kb[cur];
  const uint32_t* sidx = ib[cur];
  // group starts
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cs_starts<W>), dim3(grid_for(m)), dim3(kBlock), 0, h->stream, ssub, m,
                     start.p);
  std::vector<uint32_t> hs((size_t)m);
  HIP_OK(hipMemcpyAsync(hs.data(), start.p, (size_t)m * 4, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  std::vector<int64_t> gs;
  for (int64_t t = 0; t < m; ++t)
    if (hs[t]) gs.push_back(t);
  const int64_t ng = (int64_t)gs.size();
  DevBuf<int64_t> gstart;
  HIP_OK(gstart.alloc((size_t)std::max<int64_t>(ng, 1)));
  HIP_OK(hipMemcpy(gstart.p, gs.data(), (size_t)ng * 8, hipMemcpyHostToDevice));
  // the canonical connected pairs: count, then write after the ncore diagonal entries
  const unsigned long long zero = 0;
  HIP_OK(hipMemcpyAsync(cnt.p, &zero, 8, hipMemcpyHostToDevice, h->stream));
  const dim3 gw((unsigned)std::max<int64_t>(1, std::min<int64_t>((ng + 3) / 4, 16 * (int64_t)h->nb_grid)));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cs_pairs<W>), gw, dim3(kBlock), 0, h->stream, ssub, sidx, m,
                     (const int64_t*)gstart.p, ng, (const uint64_t*)h->core_keys, nel, 0, cnt.p, (uint64_t*)nullptr);
  unsigned long long npairs = 0;
  HIP_OK(hipMemcpyAsync(&npairs, cnt.p, 8, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  const int64_t np = (int64_t)npairs + ncore;
  DevBuf<uint64_t> pr0, pr1;
  DevBuf<uint32_t> pi0, pi1;
  HIP_OK(pr0.alloc((size_t)np));
  HIP_OK(pr1.alloc((size_t)np));
  HIP_OK(pi0.alloc((size_t)np));
  HIP_OK(pi1.alloc((size_t)np));
  hipLaunchKernelGGL(k_cs_diag, dim3(grid_for(ncore)), dim3(kBlock), 0, h->stream, ncore, pr0.p);
  HIP_OK(hipMemcpyAsync(cnt.p, &zero, 8, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cs_pairs<W>), gw, dim3(kBlock), 0, h->stream, ssub, sidx, m,
                     (const int64_t*)gstart.p, ng, (const uint64_t*)h->core_keys, nel, 1, cnt.p, pr0.p + ncore);
  // sort by (row, column): the column bits, then the row bits
  int cb = 1;
  while (cb < 32 && ((int64_t)1 << cb) < ncore) ++cb;
  std::vector<std::pair<int, int>> passes;
  for (int sh = 0; sh < cb; sh += 8) passes.push_back({0, sh});
  for (int sh = 32; sh < 32 + cb; sh += 8) passes.push_back({0, sh});
  const unsigned long long npp = (unsigned long long)np;
  HIP_OK(hipMemcpyAsync(cnt.p + 1, &npp, 8, hipMemcpyHostToDevice, h->stream));
  uint64_t* pk[2] = {pr1.p, pr0.p};
  uint32_t* pik[2] = {pi1.p, pi0.p};
  const int pc = radix_sort_pairs_in<uint64_t, 1, false, 4>(h, pr0.p, 1, nullptr, passes, cnt.p + 1, np, pk, pik);
  const uint64_t* pairs = pk[pc];
  // values, the kept entries' positions, CSR
  DevBuf<double> hv;
  DevBuf<uint32_t> kp;
  HIP_OK(hv.alloc((size_t)np));
  HIP_OK(kp.alloc((size_t)np));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cs_eval<W>), dim3(grid_for(np)), dim3(kBlock), 0, h->stream, h->ds,
                     (const uint64_t*)h->core_keys, pairs, np, hv.p, kp.p);
  const int64_t ntiles = (np + kCsTile - 1) / kCsTile;
  DevBuf<int64_t> toff;
  HIP_OK(toff.alloc((size_t)ntiles + 1));
  hipLaunchKernelGGL(k_cs_tilecount, dim3((unsigned)ntiles), dim3(kBlock), 0, h->stream, (const uint32_t*)kp.p, np,
                     toff.p);
  std::vector<int64_t> ht((size_t)ntiles);
  HIP_OK(hipMemcpyAsync(ht.data(), toff.p, (size_t)ntiles * 8, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  int64_t run = 0;
  for (int64_t t = 0; t < ntiles; ++t) {
    const int64_t c = ht[t];
    ht[t] = run;
    run += c;
  }
  nnz = run;
  HIP_OK(hipMemcpy(toff.p, ht.data(), (size_t)ntiles * 8, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(h->core_rowptr + ncore, &nnz, 8, hipMemcpyHostToDevice));
  HIP_OK(dalloc(&h->core_cols, (size_t)nnz));
  if (fp32) {
    float* v = nullptr;
    HIP_OK(dalloc(&v, (size_t)nnz));
    h->core_vals = v;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cs_write<float>), dim3((unsigned)ntiles), dim3(kBlock), 0, h->stream, pairs,
                       (const double*)hv.p, (const uint32_t*)kp.p, np, (const int64_t*)toff.p, h->core_rowptr,
                       h->core_cols, v);
  } else {
    double* v = nullptr;
    HIP_OK(dalloc(&v, (size_t)nnz));
    h->core_vals = v;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cs_write<double>), dim3((unsigned)ntiles), dim3(kBlock), 0, h->stream,
                       pairs, (const double*)hv.p, (const uint32_t*)kp.p, np, (const int64_t*)toff.p, h->core_rowptr,
                       h->core_cols, v);
  }
  HIP_OK(hipStreamSynchronize(h->stream));  // the temporary buffers are freed on return
  return NGPU_OK;
}

template <int W, int R>
int set_core_impl(ngpu_t* h, int64_t ncore, bool fp32, int64_t* nnz_out) {
  const int S = h->S;
  int64_t nnz = 0;
  // fewer than 4 electrons: subsets of nel - 2 <= 1 electrons put most of the
  // core in one group, so the scan is used.  Automatic choice: the scan costs
  // about ncore^2 pair tests, the join about ncore * nel(nel-1)/2 sorted
  // entries; measured at 26 electrons, 40 000 determinants, the scan took
  // 33 ms and the join 123 ms, so the join is taken from ncore >= 450 pairs
  // of electrons per determinant on.
  const int64_t npair = (int64_t)h->sys.nel * (h->sys.nel - 1) / 2;
  const bool scan = h->core_build == NGPU_CORE_BUILD_SCAN || h->sys.nel < 4 ||
                    (h->core_build == NGPU_CORE_BUILD_AUTO && ncore < 450 * npair);
  int rc = scan ? core_rows_scan<W>(h, ncore, fp32, nnz) : core_rows_join<W>(h, ncore, fp32, nnz);
  if (rc) return rc;
  // core dets resident, flagged deterministic (and initiator: t_core_inits)
  uint32_t cflags = kFlagCore | (h->sys.t_trunc_initiator ? ((1u << S) - 1u) : 0u);
  hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, h->stream, &h->cnt->n_new, 0ull);
  hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, h->stream, &h->cnt->n_fail, 0ull);
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, h->stream, &h->cnt->ins_refused, 0u);
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, h->stream, &h->cnt->err, 0u);
  dim3 gp((unsigned)std::min<int64_t>((ncore + kBlock - 1) / kBlock, h->nb_grid));
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_core_place<W>), gp, dim3(kBlock), 0, h->stream, h->ds, h->wk, h->cnt,
                     (const uint64_t*)h->core_keys, ncore, cflags, h->core_slot, h->n_key, h->n_sgn, h->n_flag);
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_insert<W, R>), dim3(h->nb_grid), dim3(kBlock), 0, h->stream, h->ds, h->wk,
                     h->cnt, (const uint64_t*)h->n_key, (const double*)h->n_sgn, (const uint32_t*)h->n_flag,
                     (const unsigned long long*)&h->cnt->n_new);
  hipLaunchKernelGGL(k_post_insert, dim3(1), dim3(1), 0, h->stream, h->cnt, (const unsigned long long*)&h->cnt->n_new,
                     h->wk.cap);
  hash_upkeep<W>(h);
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_core_place<W>), gp, dim3(kBlock), 0, h->stream, h->ds, h->wk, h->cnt,
                     (const uint64_t*)h->core_keys, ncore, cflags, h->core_slot, h->n_key, h->n_sgn, h->n_flag);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(h->hcnt, h->cnt, sizeof(Counters), hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  // iterations start from empty append counters and statistics rows
  hipLaunchKernelGGL(k_begin_iter, dim3(1), dim3(kBlock), 0, h->stream, h->dout, h->dpart, h->cnt, h->dparams,
                     IterParams{}, 0);
  HIP_OK(hipGetLastError());
  if (h->hcnt->halt) return err_from_mask(h->hcnt->halt);
  if (h->hcnt->err) return err_from_mask(h->hcnt->err);
  h->core_nnz = nnz;
  if (nnz_out) *nnz_out = nnz;
  return NGPU_OK;
}

// values of the stored CSR entries, recomputed from the integrals: one wave
// per row
template <int W, typename VT>
__global__ void __launch_bounds__(kBlock) k_core_revalue(DevSys s, const uint64_t* keys, int64_t ncore,
                                                         const int64_t* rowptr, const int32_t* cols, VT* vals) {
  const int lane = lane_id();
  const int64_t nwaves = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; i < ncore; i += nwaves) {
    const uint64_t* di = keys + i * W;
    for (int64_t e = rowptr[i] + lane; e < rowptr[i + 1]; e += kWave) {
      const int64_t j = cols[e];
      vals[e] = (VT)((i == j) ? diag_helement<W>(s, di) - s.hii : helement_pair<W>(s, di, keys + j * W));
    }
  }
}

// the resident core's Hamiltonian values stored again in the other precision
// (same rows and columns, each value recomputed from the integrals)
template <int W>
int restore_core_values(ngpu_t* h, bool fp32) {
  const int64_t ncore = h->ncore;
  dim3 grid((unsigned)std::min<int64_t>((ncore + 3) / 4, 8 * (int64_t)h->nb_grid));
  HIP_OK(hipStreamSynchronize(h->stream));
  HIP_OK(hipFree(h->core_vals));
  h->core_vals = nullptr;
  if (fp32) {
    float* v = nullptr;
    HIP_OK(dalloc(&v, (size_t)h->core_nnz));
    h->core_vals = v;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_core_revalue<W, float>), grid, dim3(kBlock), 0, h->stream, h->ds,
                       (const uint64_t*)h->core_keys, ncore, (const int64_t*)h->core_rowptr,
                       (const int32_t*)h->core_cols, v);
  } else {
    double* v = nullptr;
    HIP_OK(dalloc(&v, (size_t)h->core_nnz));
    h->core_vals = v;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_core_revalue<W, double>), grid, dim3(kBlock), 0, h->stream, h->ds,
                       (const uint64_t*)h->core_keys, ncore, (const int64_t*)h->core_rowptr,
                       (const int32_t*)h->core_cols, v);
  }
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(h->stream));
  h->core_fp32 = fp32;
  return NGPU_OK;
}
}  // namespace
}  // extern "C++"

extern "C++" {
namespace {

// generate_space_most_populated's weight of a slot (core_space_weight,
// semi_stoch_procs.F90:1623-1642, the global core space: sum_r |c_r|), or -1
// for a slot out of use or below 1e-8 (semi_stoch_gen.F90:1094-1100)
__device__ inline double mp_weight(const Walk& wk, int64_t slot, int S) {
  if (wk.flags(slot) & kFlagRemoved) return -1.0;
  double w = 0.0;
  for (int r = 0; r < S; ++r) w += fabs(wk.sgn(slot, r));
  return w < 1e-8 ? -1.0 : w;
}

constexpr int64_t kMpTile = 4096;  // slots per block of the candidate compaction

__global__ void __launch_bounds__(kBlock) k_mp_count(Walk wk, const Counters* cnt, int S, int64_t* counts) {
  __shared__ int64_t lds[kBlock / kWave];
  const int64_t ns = (int64_t)cnt->nslots, b0 = (int64_t)blockIdx.x * kMpTile;
  const int64_t b1 = std::min<int64_t>(b0 + kMpTile, ns);
  int64_t c = 0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kBlock) c += mp_weight(wk, i, S) >= 0.0 ? 1 : 0;
  int64_t tot = 0;
  (void)block_exclusive_scan(c, lds, tot);
  if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

// candidates in slot order: keys (W words), index 0..n-1 and the weight
// complemented (~bits), whose ascending order is the weights' descending order
__global__ void __launch_bounds__(kBlock) k_mp_write(Walk wk, const Counters* cnt, int W, int S,
                                                     const int64_t* offsets, uint64_t* keys, uint64_t* wkey) {
  __shared__ int64_t lds[kBlock / kWave];
  const int64_t ns = (int64_t)cnt->nslots, b0 = (int64_t)blockIdx.x * kMpTile;
  int64_t pos0 = offsets[blockIdx.x];
  for (int64_t t = 0; t < kMpTile && b0 + t < ns; t += kBlock) {
    const int64_t i = b0 + t + threadIdx.x;
    const double w = i < ns ? mp_weight(wk, i, S) : -1.0;
    int64_t tot = 0;
    const int64_t pos = pos0 + block_exclusive_scan(w >= 0.0 ? 1 : 0, lds, tot);
    if (w >= 0.0) {
      for (int q = 0; q < W; ++q) keys[pos * W + q] = wk.keys[i * W + q];
      wkey[pos] = ~(uint64_t)__double_as_longlong(w);
    }
    pos0 += tot;
  }
}

// the weights' keys in the order of a permutation
__global__ void k_mp_gather(const uint64_t* wkey, const uint32_t* perm, int64_t n, uint64_t* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = wkey[perm[i]];
}


// generate_space_most_populated on the resident walkers: the candidates in
// slot order, stably sorted by determinant (ilut_lt), then stably by weight
// descending; the first target are the core space (weight descending, equal
// weights in determinant order)
template <int W>
int most_populated_impl(ngpu_t* h, int64_t target, std::vector<uint64_t>& out) {
  if (h->pending) {
    int rc = finish(h, nullptr);
    if (rc) return rc;
  }
  HIP_OK(hipMemcpy(h->hcnt, h->cnt, sizeof(Counters), hipMemcpyDeviceToHost));
  const int64_t ns = (int64_t)h->hcnt->nslots;
  const int64_t ntiles = std::max<int64_t>(1, (ns + kMpTile - 1) / kMpTile);
  DevBuf<int64_t> counts;
  DevBuf<uint64_t> ckeys, wkey, wperm;
  DevBuf<unsigned long long> n_dev;
  HIP_OK(counts.alloc((size_t)ntiles + 1));
  HIP_OK(ckeys.alloc((size_t)std::max<int64_t>(ns, 1) * W));
  HIP_OK(wkey.alloc((size_t)std::max<int64_t>(ns, 1)));
  HIP_OK(wperm.alloc((size_t)std::max<int64_t>(ns, 1)));
  HIP_OK(n_dev.alloc(1));
  hipLaunchKernelGGL(k_mp_count, dim3((unsigned)ntiles), dim3(kBlock), 0, h->stream, h->wk, (const Counters*)h->cnt,
                     h->S, counts.p);
  std::vector<int64_t> hc((size_t)ntiles + 1, 0);
  HIP_OK(hipMemcpyAsync(hc.data(), counts.p, (size_t)ntiles * 8, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  int64_t n = 0;
  for (int64_t t = 0; t < ntiles; ++t) {
    const int64_t c = hc[t];
    hc[t] = n;
    n += c;
  }
  out.clear();
  if (n == 0) return NGPU_OK;
  if (n > h->scap) return NGPU_ERR_SPAWN_FULL;  // the sort runs in the spawn-sort buffers
  HIP_OK(hipMemcpy(counts.p, hc.data(), (size_t)ntiles * 8, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mp_write, dim3((unsigned)ntiles), dim3(kBlock), 0, h->stream, h->wk, (const Counters*)h->cnt,
                     W, h->S, (const int64_t*)counts.p, ckeys.p, wkey.p);
  const unsigned long long nn = (unsigned long long)n;
  HIP_OK(hipMemcpyAsync(n_dev.p, &nn, 8, hipMemcpyHostToDevice, h->stream));
  // determinant order first (the tie order), then weight (stable)
  int cur = radix_sort_pairs<uint64_t, W, W == 2, 4>(h, ckeys.p, W, nullptr, word_passes(h), n_dev.p, n);
  hipLaunchKernelGGL(k_mp_gather, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, (const uint64_t*)wkey.p,
                     (const uint32_t*)h->si[cur], n, wperm.p);
  std::vector<std::pair<int, int>> wp;
  for (int sh = 0; sh < 64; sh += 8) wp.push_back({0, sh});
  const uint32_t* perm = h->si[cur];
  // the weight passes must not overwrite the permutation they read first
  if (cur == 0) {
    HIP_OK(hipMemcpyAsync(h->si[1], h->si[0], (size_t)n * 4, hipMemcpyDeviceToDevice, h->stream));
    perm = h->si[1];
  }
  cur = radix_sort_pairs<uint64_t, 1, false, 4>(h, wperm.p, 1, perm, wp, n_dev.p, n);
  const int64_t k = std::min(target, n);
  std::vector<uint32_t> idx((size_t)k);
  HIP_OK(hipMemcpyAsync(idx.data(), h->si[cur], (size_t)k * 4, hipMemcpyDeviceToHost, h->stream));
  std::vector<uint64_t> all((size_t)n * W);
  HIP_OK(hipMemcpyAsync(all.data(), ckeys.p, all.size() * 8, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  out.resize((size_t)k * W);
  for (int64_t i = 0; i < k; ++i)
    for (int q = 0; q < W; ++q) out[(size_t)i * W + q] = all[(size_t)idx[i] * W + q];
  return NGPU_OK;
}

}  // namespace
}  // extern "C++"

int ngpu_most_populated(ngpu_t* h, int64_t target, int64_t* dets, int64_t* n) {
  int rc = ready(h);
  if (rc) return rc;
  if (target < 0 || !n || (target > 0 && !dets)) return NGPU_ERR_ARG;
  if (h->batch) return NGPU_ERR_STATE;
  std::vector<uint64_t> out;
  rc = h->W == 1 ? most_populated_impl<1>(h, target, out) : most_populated_impl<2>(h, target, out);
  if (rc) return rc;
  *n = (int64_t)(out.size() / h->W);
  if (!out.empty()) memcpy(dets, out.data(), out.size() * 8);
  return NGPU_OK;
}

int ngpu_select_core(ngpu_t* h, int64_t target, int32_t fp32_values, int64_t* dets, int64_t* ncore, int64_t* nnz) {
  int rc = ready(h);
  if (rc) return rc;
  if (target < 1 || target > (int64_t)INT32_MAX) return NGPU_ERR_ARG;
  if (h->batch) return NGPU_ERR_STATE;
  std::vector<uint64_t> out;
  rc = h->W == 1 ? most_populated_impl<1>(h, target, out) : most_populated_impl<2>(h, target, out);
  if (rc) return rc;
  const int64_t k = (int64_t)(out.size() / h->W);
  if (k == 0) return NGPU_ERR_STATE;  // no occupied determinant to build a core space from
  if (ncore) *ncore = k;
  if (dets) memcpy(dets, out.data(), out.size() * 8);
  return ngpu_set_core(h, (const int64_t*)out.data(), k, fp32_values, nnz);
}

int ngpu_set_core(ngpu_t* h, const int64_t* dets, int64_t ncore, int32_t fp32_values, int64_t* nnz) {
  if (h) h->graph_dirty = true;  // kernel arguments of the captured chain change
  int rc = ready(h);
  if (rc) return rc;
  if (!dets || ncore < 1 || ncore > (int64_t)INT32_MAX || ncore > h->wk.cap) return NGPU_ERR_ARG;
  if (h->ncore) {
    // one core space per engine; called again with the same determinants it
    // changes the storage precision of the core Hamiltonian
    if (ncore != h->ncore || h->batch) return NGPU_ERR_STATE;
    if (h->pending) {
      rc = finish(h, nullptr);
      if (rc) return rc;
    }
    std::vector<int64_t> cur((size_t)ncore * h->W);
    HIP_OK(hipMemcpy(cur.data(), h->core_keys, cur.size() * 8, hipMemcpyDeviceToHost));
    if (memcmp(cur.data(), dets, cur.size() * 8) != 0) return NGPU_ERR_STATE;
    if ((fp32_values != 0) != h->core_fp32) {
      rc = h->W == 1 ? restore_core_values<1>(h, fp32_values != 0) : restore_core_values<2>(h, fp32_values != 0);
      if (rc) return rc;
    }
    if (nnz) *nnz = h->core_nnz;
    return NGPU_OK;
  }
  for (int64_t i = 0; i < ncore; ++i) {
    int nel = 0;
    for (int w = 0; w < h->W; ++w) nel += __builtin_popcountll((uint64_t)dets[i * h->W + w]);
    if (nel != h->sys.nel) return NGPU_ERR_ARG;
  }
  {  // the core space holds each determinant once
    std::vector<int64_t> ord((size_t)ncore);
    for (int64_t i = 0; i < ncore; ++i) ord[i] = i;
    const int W = h->W;
    auto lt = [&](int64_t a, int64_t b) {
      for (int w = W - 1; w >= 0; --w)
        if (dets[a * W + w] != dets[b * W + w]) return (uint64_t)dets[a * W + w] < (uint64_t)dets[b * W + w];
      return false;
    };
    std::sort(ord.begin(), ord.end(), lt);
    for (int64_t t = 1; t < ncore; ++t)
      if (!lt(ord[t - 1], ord[t])) return NGPU_ERR_ARG;
  }
  if (h->pending) {
    rc = finish(h, nullptr);
    if (rc) return rc;
  }
  const int W = h->W, S = h->S;
  HIP_OK(dalloc(&h->core_keys, (size_t)ncore * W));
  HIP_OK(dalloc(&h->core_slot, (size_t)ncore));
  HIP_OK(dalloc(&h->core_rowptr, (size_t)ncore + 1));
  HIP_OK(dalloc(&h->core_x, (size_t)ncore * S));
  HIP_OK(dalloc(&h->core_y, (size_t)ncore * S));
  HIP_OK(hipMemcpy(h->core_keys, dets, (size_t)ncore * W * 8, hipMemcpyHostToDevice));
  h->core_fp32 = fp32_values != 0;
  if (W == 1) rc = S == 1 ? set_core_impl<1, 1>(h, ncore, h->core_fp32, nnz) : set_core_impl<1, kMaxRep>(h, ncore, h->core_fp32, nnz);
  else rc = S == 1 ? set_core_impl<2, 1>(h, ncore, h->core_fp32, nnz) : set_core_impl<2, kMaxRep>(h, ncore, h->core_fp32, nnz);
  if (rc) return rc;
  h->ncore = ncore;
  h->ds.ncore = 1;
  return NGPU_OK;
}

int ngpu_get_core_hamil(ngpu_t* h, int64_t* rowptr, int32_t* cols, double* vals, int64_t cap, int64_t* nnz) {
  int rc = ready(h);
  if (rc) return rc;
  if (!nnz) return NGPU_ERR_ARG;
  if (!h->ncore) return NGPU_ERR_STATE;
  *nnz = h->core_nnz;
  if (!rowptr && !cols && !vals) return NGPU_OK;
  if (!rowptr || !cols || !vals) return NGPU_ERR_ARG;
  if (cap < h->core_nnz) return NGPU_ERR_ARG;
  if (h->pending) {
    rc = finish(h, nullptr);
    if (rc) return rc;
  }
  HIP_OK(hipStreamSynchronize(h->stream));
  const size_t nz = (size_t)h->core_nnz;
  HIP_OK(hipMemcpy(rowptr, h->core_rowptr, ((size_t)h->ncore + 1) * 8, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(cols, h->core_cols, nz * 4, hipMemcpyDeviceToHost));
  if (h->core_fp32) {
    std::vector<float> f(nz);
    HIP_OK(hipMemcpy(f.data(), h->core_vals, nz * 4, hipMemcpyDeviceToHost));
    for (size_t e = 0; e < nz; ++e) vals[e] = (double)f[e];
  } else {
    HIP_OK(hipMemcpy(vals, h->core_vals, nz * 8, hipMemcpyDeviceToHost));
  }
  return NGPU_OK;
}

int ngpu_set_checks(ngpu_t* h, int32_t mask) {
  if (!h || mask < 0 || (mask & ~(NGPU_CHECK_TABLE | NGPU_CHECK_PGEN | NGPU_CHECK_SELFTEST))) return NGPU_ERR_ARG;
  if (h->batch) return NGPU_ERR_STATE;
  if (h->pending) {
    int rc = finish(h, nullptr);
    if (rc) return rc;
  }
  if (!h->d_chk) {
    HIP_OK(dalloc(&h->d_chk, 16));
    HIP_OK(hipMemset(h->d_chk, 0, 16 * sizeof(unsigned long long)));
  }
  h->checks = (uint32_t)mask;
  h->graph_dirty = true;  // the captured chain changes
  return NGPU_OK;
}

int ngpu_get_check_counts(ngpu_t* h, int64_t* counts, int32_t n) {
  if (!h || !counts || n < 1 || n > 7) return NGPU_ERR_ARG;
  for (int i = 0; i < n; ++i) counts[i] = 0;
  if (!h->d_chk) return NGPU_OK;
  if (h->pending) {
    int rc = finish(h, nullptr);
    if (rc && rc != NGPU_ERR_CHECK) return rc;
  }
  unsigned long long c[16];
  HIP_OK(hipStreamSynchronize(h->stream));
  HIP_OK(hipMemcpy(c, h->d_chk, sizeof(c), hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i) counts[i] = (int64_t)c[8 + i];
  return NGPU_OK;
}

int ngpu_set_core_build(ngpu_t* h, int32_t method) {
  if (!h || method < NGPU_CORE_BUILD_AUTO || method > NGPU_CORE_BUILD_SCAN) return NGPU_ERR_ARG;
  if (h->ncore) return NGPU_ERR_STATE;
  h->core_build = method;
  return NGPU_OK;
}

int ngpu_download_walkers(ngpu_t* h, int64_t* rows, double* diag_h, double* offdiag_h, int64_t cap, int64_t* n) {
  if (!h || !n) return NGPU_ERR_ARG;
  HIP_OK(hipStreamSynchronize(h->stream));
  HIP_OK(hipMemcpy(h->hcnt, h->cnt, sizeof(Counters), hipMemcpyDeviceToHost));
  int64_t ns = (int64_t)h->hcnt->nslots;
  const int W = h->W, S = h->S, RL = W + S + 1;
  std::vector<uint64_t> key((size_t)ns * W);
  std::vector<double> sg((size_t)ns * S), dh((size_t)ns), oh((size_t)ns);
  std::vector<uint32_t> fl((size_t)ns);
  if (ns > 0) {
    const int64_t sr = h->wk.srow;
    std::vector<double> st((size_t)ns * sr);
    HIP_OK(hipMemcpy(key.data(), h->wk.keys, key.size() * 8, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(st.data(), h->wk.st, st.size() * 8, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(dh.data(), h->wk.diagH, dh.size() * 8, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(oh.data(), h->wk.offdiagH, oh.size() * 8, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < ns; ++i) {
      for (int r = 0; r < S; ++r) sg[(size_t)i * S + r] = st[(size_t)i * sr + r];
      uint64_t fw;
      memcpy(&fw, &st[(size_t)i * sr + S], 8);
      fl[(size_t)i] = (uint32_t)fw;
    }
  }
  const int* init_bits = h->sys.flag_layout == NGPU_FLAGS_NECI ? kNeciInit : kMneciInit;
  const int* det_bits = h->sys.flag_layout == NGPU_FLAGS_NECI ? kNeciDeterm : kMneciDeterm;
  int maxr = h->sys.flag_layout == NGPU_FLAGS_NECI ? 2 : 16;
  int64_t m = 0;
  for (int64_t i = 0; i < ns; ++i) {
    if (fl[i] & kFlagRemoved) continue;
    bool occ = false;
    for (int r = 0; r < S; ++r) occ = occ || !(fabs(sg[i * S + r]) < 1e-12);
    if (!occ && !(fl[i] & kFlagCore)) continue;
    if (m < cap && rows) {
      int64_t* row = rows + m * RL;
      for (int w = 0; w < W; ++w) row[w] = (int64_t)key[i * W + w];
      memcpy(row + W, &sg[i * S], sizeof(double) * S);
      uint64_t nf = 0;
      for (int r = 0; r < S && r < maxr; ++r) {
        if ((fl[i] >> r) & 1u) nf |= 1ull << init_bits[r];
        if (fl[i] & kFlagCore) nf |= 1ull << det_bits[r];
      }
      row[W + S] = (int64_t)nf;
      if (diag_h) diag_h[m] = dh[i];
      if (offdiag_h) offdiag_h[m] = oh[i];
    }
    ++m;
  }
  *n = m;
  if (!rows) return NGPU_OK;  // size query
  return m > cap ? NGPU_ERR_ARG : NGPU_OK;
}

namespace {

// the row test of ngpu_download_walkers: a slot in use that holds a nonzero
// sign in some replica or is a core determinant
__device__ inline bool dl_keep(const Walk& wk, int64_t slot, int S) {
  const uint32_t f = wk.flags(slot);
  if (f & kFlagRemoved) return false;
  if (f & kFlagCore) return true;
  for (int r = 0; r < S; ++r)
    if (!(fabs(wk.sgn(slot, r)) < 1e-12)) return true;
  return false;
}

// engine flag bits -> NECI flag bit numbers of the boundary layout
struct FlagMap {
  int init[16], determ[16], maxr;
};

// rows kept per tile of kDlTile slots (one block per tile; tiles past the
// used slots count zero)
__global__ void __launch_bounds__(kBlock) k_dl_count(Walk wk, const Counters* cnt, int S, unsigned long long* counts) {
  __shared__ int64_t lds[kBlock / kWave];
  const int64_t ns = (int64_t)cnt->nslots;
  const int64_t b0 = (int64_t)blockIdx.x * kDlTile;
  const int64_t b1 = std::min<int64_t>(b0 + kDlTile, ns);
  int64_t c = 0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kBlock) c += dl_keep(wk, i, S) ? 1 : 0;
  int64_t tot = 0;
  (void)block_exclusive_scan(c, lds, tot);
  if (threadIdx.x == 0) counts[blockIdx.x] = (unsigned long long)tot;
}

// exclusive scan of the ntiles counts in place (one block of 1024 threads,
// ntiles <= kDlMaxTiles: 16 tiles per thread); the total goes to
// counts[kDlMaxTiles] and to *n_out
__global__ void __launch_bounds__(1024) k_dl_scan(unsigned long long* counts, int64_t ntiles, int64_t* n_out) {
  constexpr int kT = 1024, kPer = (int)(kDlMaxTiles / kT);
  __shared__ unsigned long long part[kT];
  const int64_t lo = (int64_t)threadIdx.x * kPer;
  unsigned long long v[kPer], s = 0;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    v[q] = lo + q < ntiles ? counts[lo + q] : 0ull;
    s += v[q];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < kT; o <<= 1) {  // inclusive Hillis-Steele scan of the per-thread sums
    unsigned long long y = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0ull;
    __syncthreads();
    part[threadIdx.x] += y;
    __syncthreads();
  }
  unsigned long long run = part[threadIdx.x] - s;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    if (lo + q < ntiles) counts[lo + q] = run;
    run += v[q];
  }
  if (threadIdx.x == kT - 1) {
    counts[kDlMaxTiles] = part[kT - 1];
    *n_out = (int64_t)part[kT - 1];
  }
}

// the kept rows of each tile, in slot order, from the tile's offset on; rows
// at positions >= cap are counted but not written
__global__ void __launch_bounds__(kBlock) k_dl_write(Walk wk, const Counters* cnt, int W, int S, FlagMap fm,
                                                     const unsigned long long* offsets, int64_t* rows, double* diag_h,
                                                     double* offdiag_h, int64_t cap) {
  __shared__ int64_t lds[kBlock / kWave];
  const int64_t ns = (int64_t)cnt->nslots;
  const int64_t b0 = (int64_t)blockIdx.x * kDlTile;
  const int RL = W + S + 1;
  int64_t pos0 = (int64_t)offsets[blockIdx.x];
  for (int64_t t = 0; t < kDlTile && b0 + t < ns; t += kBlock) {  // block-uniform trip count
    const int64_t i = b0 + t + threadIdx.x;
    const bool keep = i < ns && dl_keep(wk, i, S);
    int64_t tot = 0;
    const int64_t pos = pos0 + block_exclusive_scan(keep ? 1 : 0, lds, tot);
    if (keep && pos < cap) {
      int64_t* row = rows + pos * RL;
      for (int w = 0; w < W; ++w) row[w] = (int64_t)wk.keys[i * W + w];
      for (int r = 0; r < S; ++r) row[W + r] = __double_as_longlong(wk.sgn(i, r));
      const uint32_t f = wk.flags(i);
      uint64_t nf = 0;
      for (int r = 0; r < S && r < fm.maxr; ++r) {
        if ((f >> r) & 1u) nf |= 1ull << fm.init[r];
        if (f & kFlagCore) nf |= 1ull << fm.determ[r];
      }
      row[W + S] = (int64_t)nf;
      if (diag_h) diag_h[pos] = wk.diagH[i];
      if (offdiag_h) offdiag_h[pos] = wk.offdiagH[i];
    }
    pos0 += tot;
  }
}

}  // namespace

int ngpu_download_walkers_dev(ngpu_t* h, int64_t* rows, double* diag_h, double* offdiag_h, int64_t cap, int64_t* n) {
  if (!h || !n || cap < 0 || (cap > 0 && !rows)) return NGPU_ERR_ARG;
  if (h->wk.cap > kDlMaxSlots) {
    fprintf(stderr,
            "[neci_gpu] ngpu_download_walkers_dev: the walker table has %lld slots, more than the %lld its "
            "fixed scratch covers; use ngpu_download_walkers\n",
            (long long)h->wk.cap, (long long)kDlMaxSlots);
    return NGPU_ERR_ARG;
  }
  unsigned long long* sp = (h->dev >= 0 && h->dev < kMaxDevices) ? g_scratch[h->dev] : nullptr;
  if (!sp) return NGPU_ERR_STATE;
  FlagMap fm{};
  const bool neci = h->sys.flag_layout == NGPU_FLAGS_NECI;
  fm.maxr = neci ? 2 : 16;
  for (int r = 0; r < fm.maxr; ++r) {
    fm.init[r] = neci ? kNeciInit[r] : kMneciInit[r];
    fm.determ[r] = neci ? kNeciDeterm[r] : kMneciDeterm[r];
  }
  // every slot in use lies below the table's capacity, so its tiles cover them
  const int64_t ntiles = std::max<int64_t>(1, (h->wk.cap + kDlTile - 1) / kDlTile);
  hipLaunchKernelGGL(k_dl_count, dim3((unsigned)ntiles), dim3(kBlock), 0, h->stream, h->wk, (const Counters*)h->cnt,
                     h->S, sp);
  hipLaunchKernelGGL(k_dl_scan, dim3(1), dim3(1024), 0, h->stream, sp, ntiles, n);
  hipLaunchKernelGGL(k_dl_write, dim3((unsigned)ntiles), dim3(kBlock), 0, h->stream, h->wk, (const Counters*)h->cnt,
                     h->W, h->S, fm, (const unsigned long long*)sp, rows, diag_h, offdiag_h, cap);
  HIP_OK(hipGetLastError());
  return NGPU_OK;
}

namespace {

__global__ void k_set_params(IterParams* p, IterParams v) { *p = v; }

IterParams make_params(const ngpu_t* h, double tau, const double* sft, uint32_t iter) {
  IterParams v{};
  v.tau = tau;
  for (int r = 0; r < kMaxRep; ++r) v.sft.v[r] = (sft && r < h->S) ? sft[r] : 0.0;
  v.iter = iter;
  return v;
}

// the iteration's inputs, stream-ordered before the chain that reads them
void set_params(ngpu_t* h, double tau, const double* sft, uint32_t iter) {
  hipLaunchKernelGGL(k_set_params, dim3(1), dim3(1), 0, h->stream, h->dparams, make_params(h, tau, sft, iter));
}

void run_chain(ngpu_t* h) {
  if (h->W == 1) h->S == 1 ? iterate_chain<1, 1>(h) : iterate_chain<1, kMaxRep>(h);
  else h->S == 1 ? iterate_chain<2, 1>(h) : iterate_chain<2, kMaxRep>(h);
}

// one iteration's launches captured from the engine stream into an
// executable graph; the kernels read tau, shifts and the iteration number
// from h->dparams, so the graph replays every iteration unchanged
int capture_engine(ngpu_t* h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  h->gexec = nullptr;
  hipGraph_t g = nullptr;
  HIP_OK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  bool prof = h->prof;
  h->prof = false;
  run_chain(h);
  h->prof = prof;
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (e != hipSuccess) return NGPU_ERR_HIP;
  e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    h->gexec = nullptr;
    return NGPU_ERR_HIP;
  }
  h->graph_dirty = false;
  return NGPU_OK;
}

}  // namespace

int ngpu_iterate(ngpu_t* h, const ngpu_iter_in_t* in, ngpu_iter_out_t* out) {
  int rc = ready(h);
  if (rc) return rc;
  if (!in || !(in->tau > 0.0)) return NGPU_ERR_ARG;
  if (h->batch) return NGPU_ERR_STATE;  // batched engines iterate through ngpu_batch_iterate
  // iterations queue on the stream back to back: nothing in the chain waits
  // for the host; errors surface at the next synchronising call
  if (h->use_graph && !h->prof) {
    if (h->graph_dirty || !h->gexec) {
      rc = capture_engine(h);
      if (rc) return rc;
    }
    set_params(h, in->tau, in->diag_sft, in->iter);
    HIP_OK(hipGraphLaunch(h->gexec, h->stream));
  } else {
    h->next_params = make_params(h, in->tau, in->diag_sft, in->iter);
    h->chain_sets_params = true;
    if (h->prof && h->ev_ok) {
      h->ev_cur = (int)(h->ev_next % kEvRing);
      harvest(h, h->ev_cur);
      h->ev_mask[h->ev_cur] = 0;
    }
    run_chain(h);
    h->chain_sets_params = false;
    if (h->prof && h->ev_ok) {
      h->ev_pending[h->ev_cur] = true;
      h->ev_next++;
    }
  }
  HIP_OK(hipGetLastError());
  if (out) HIP_OK(hipMemcpyAsync(h->hout, h->dout, sizeof(ngpu_iter_out_t), hipMemcpyDeviceToHost, h->stream));
  h->pending = true;
  if (out) return finish(h, out);
  return NGPU_OK;
}

int ngpu_set_launch_shape(ngpu_t* h, int32_t chunk_blocks, int32_t grid_blocks, int32_t radix_items,
                          int32_t single_lanes, int32_t merge_blocks) {
  if (!h) return NGPU_ERR_ARG;
  if (chunk_blocks < 0 || chunk_blocks > h->max_chunk || (chunk_blocks && chunk_blocks < 64) || grid_blocks < 0 ||
      (grid_blocks && grid_blocks < 64) || grid_blocks > 64 * h->n_cu || merge_blocks < 0 ||
      (merge_blocks && merge_blocks < 64) || merge_blocks > 64 * h->n_cu ||
      !(radix_items == 0 || radix_items == 2 || radix_items == 4 || radix_items == 8) ||
      !(single_lanes == 0 || single_lanes == 4 || single_lanes == 8 || single_lanes == 16))
    return NGPU_ERR_ARG;
  if (h->pending) {
    int rc = finish(h, nullptr);
    if (rc) return rc;
  }
  if (chunk_blocks) h->nb_chunk = chunk_blocks;
  if (grid_blocks) h->nb_grid = grid_blocks;
  if (radix_items) h->radix_items = radix_items;
  if (single_lanes) h->single_lanes = single_lanes;
  if (merge_blocks) h->nb_merge = merge_blocks;
  h->graph_dirty = true;
  return NGPU_OK;
}

int ngpu_get_launch_shape(ngpu_t* h, int32_t* chunk_blocks, int32_t* grid_blocks, int32_t* radix_items,
                          int32_t* single_lanes, int32_t* merge_blocks, int32_t* n_cu) {
  if (!h) return NGPU_ERR_ARG;
  if (merge_blocks) *merge_blocks = h->nb_merge ? h->nb_merge : h->nb_grid;
  if (chunk_blocks) *chunk_blocks = h->nb_chunk;
  if (grid_blocks) *grid_blocks = h->nb_grid;
  if (radix_items) *radix_items = h->radix_items;
  if (single_lanes) *single_lanes = h->single_lanes;
  if (n_cu) *n_cu = h->n_cu;
  return NGPU_OK;
}

int ngpu_set_graph(ngpu_t* h, int32_t on) {
  if (!h) return NGPU_ERR_ARG;
  h->use_graph = on != 0;
  return NGPU_OK;
}

// ---- batches: several engines, one launch per kernel for each group ----
extern "C++" {
namespace {

// orbitals per lane of the heat-bath singles for lanes G and nsp spatial orbitals
int singles_pl(int G, int nsp) {
  if (G == 16) return nsp <= 16 ? 1 : nsp <= 32 ? 2 : 4;
  if (G == 8) return nsp <= 16 ? 2 : nsp <= 24 ? 3 : nsp <= 32 ? 4 : nsp <= 40 ? 5 : 8;
  return nsp <= 16 ? 4 : nsp <= 20 ? 5 : nsp <= 32 ? 8 : nsp <= 40 ? 10 : 16;
}


// an engine's group: same kernel instances and launch knobs
BatchGroup group_of(const ngpu_t* h) {
  BatchGroup g;
  g.W = h->W;
  g.S = h->S == 1 ? 1 : kMaxRep;
  g.gen = gen_of(h);
  g.singles = g.gen == kGenPchb || g.gen == kGenPchbW;
  g.lanes = h->single_lanes;
  g.pl = singles_pl(h->single_lanes, h->sys.nbasis / 2);
  g.items = h->radix_items;
  g.rank_sort = h->rank_sort;
  g.rank64 = h->rank64;
  g.rank64_bits = h->rank64_bits;
  if (!g.rank_sort) g.passes = word_passes(h);
  g.core = h->ncore > 0;
  g.core_fp32 = h->core_fp32;
  g.rec_words = h->ds.rec_words;
  return g;
}

bool same_group(const BatchGroup& a, const BatchGroup& b) {
  return a.W == b.W && a.S == b.S && a.gen == b.gen && a.lanes == b.lanes && a.pl == b.pl && a.items == b.items &&
         a.rank_sort == b.rank_sort && a.rank64 == b.rank64 && a.rank64_bits == b.rank64_bits &&
         a.passes == b.passes && a.core == b.core && a.core_fp32 == b.core_fp32 &&
         a.rec_words == b.rec_words;
}

void free_group(BatchGroup& g) {
  if (g.d_ed) (void)hipFree(g.d_ed);
  g.d_ed = nullptr;
  for (void* p : {(void*)g.d_X, (void*)g.d_Y, (void*)g.d_colmap})
    if (p) (void)hipFree(p);
  g.d_X = g.d_Y = nullptr;
  g.d_colmap = nullptr;
  g.shared_core = false;
  for (int k = 0; k < kLkKinds; ++k) {
    if (g.d_map[k]) (void)hipFree(g.d_map[k]);
    g.d_map[k] = nullptr;
  }
}

EngDev eng_dev(const ngpu_t* h) {
  EngDev e{};
  e.s = h->ds;
  e.wk = h->wk;
  e.cnt = h->cnt;
  e.part = h->dpart;
  e.att_cnt = h->att_cnt;
  e.att_off = h->att_off;
  e.pinfo = h->pinfo;
  e.blk_part = h->blk_part;
  e.blk_rm = h->blk_rm;
  e.tile_p = h->tile_p;
  e.tcap = h->tcap;
  e.sp_rec = h->sp_rec;
  e.rk = h->rk;
  e.scap = h->scap;
  e.sq = h->sq;
  for (int q = 0; q < 2; ++q) {
    e.sk[q] = h->sk[q];
    e.si[q] = h->si[q];
  }
  e.hist = h->hist;
  e.digit_tot = h->digit_tot;
  e.n_key = h->n_key;
  e.n_sgn = h->n_sgn;
  e.n_flag = h->n_flag;
  e.prm = h->dparams;
  e.dout = h->dout;
  e.dcyc = h->dcyc;
  e.ncore = h->ncore;
  e.core_slot = h->core_slot;
  e.core_rowptr = h->core_rowptr;
  e.core_cols = h->core_cols;
  e.core_vals = h->core_vals;
  e.core_x = h->core_x;
  e.core_y = h->core_y;
  return e;
}

// Groups of the batch's engines (in the order given, each group in engine
// order) and their device descriptors.  Each engine's share of a launch is
// its own launch shape scaled by its walker capacity against the group's
// largest, so small problems do not take as many blocks as large ones.
int build_groups(ngpu_batch* b) {
  for (BatchGroup& g : b->groups) free_group(g);
  b->groups.clear();
  for (ngpu_t* h : b->hs) {
    BatchGroup k = group_of(h);
    BatchGroup* g = nullptr;
    for (BatchGroup& x : b->groups)
      if (same_group(x, k)) g = &x;
    if (!g) {
      b->groups.push_back(k);
      g = &b->groups.back();
    }
    g->hs.push_back(h);
  }
  for (BatchGroup& g : b->groups) {
    const size_t E = g.hs.size();
    if (E > 65535) return NGPU_ERR_ARG;
    int64_t wmax = 1;
    for (ngpu_t* h : g.hs) wmax = std::max<int64_t>(wmax, h->wk.cap);
    std::vector<EngDev> ed(E);
    std::vector<uint16_t> map[kLkKinds];
    for (int k = 0; k < kLkKinds; ++k) g.tot[k] = 0;
    for (size_t i = 0; i < E; ++i) {
      const ngpu_t* h = g.hs[i];
      ed[i] = eng_dev(h);
      auto scaled = [&](int nb) {
        int64_t v = ((int64_t)nb * h->wk.cap + wmax - 1) / wmax;
        return (int)std::max<int64_t>(std::min<int64_t>(v, nb), std::min(nb, 16));
      };
      int nb[kLkKinds];
      nb[kLkChunk] = std::min(scaled(h->nb_chunk), h->max_chunk);
      nb[kLkGrid] = scaled(h->nb_grid);
      nb[kLkMerge] = scaled(h->nb_merge ? h->nb_merge : h->nb_grid);
      nb[kLkCore] = (int)std::max<int64_t>(1, std::min<int64_t>((h->ncore + kBlock - 1) / kBlock, h->nb_grid));
      nb[kLkCoreRow] = (int)std::max<int64_t>(
          1, std::min<int64_t>((h->ncore + kBlock / kWave - 1) / (kBlock / kWave), 8 * (int64_t)h->nb_grid));
      for (int k = 0; k < kLkKinds; ++k) {
        if ((k == kLkCore || k == kLkCoreRow) && !g.core) nb[k] = 0;
        ed[i].b0[k] = g.tot[k];
        ed[i].nb[k] = nb[k];
        g.tot[k] += nb[k];
        map[k].insert(map[k].end(), (size_t)nb[k], (uint16_t)i);
      }
    }
    if (g.core && E > 1) {  // the same core space in every engine (replicas of one system)
      const ngpu_t* h0 = g.hs[0];
      bool same = true;
      std::vector<uint64_t> k0((size_t)h0->ncore * h0->W), kx(k0.size());
      HIP_OK(hipMemcpy(k0.data(), h0->core_keys, k0.size() * 8, hipMemcpyDeviceToHost));
      for (size_t i = 1; i < E && same; ++i) {
        const ngpu_t* h = g.hs[i];
        same = h->ncore == h0->ncore && h->core_nnz == h0->core_nnz && h->core_fp32 == h0->core_fp32;
        if (same) {
          HIP_OK(hipMemcpy(kx.data(), h->core_keys, kx.size() * 8, hipMemcpyDeviceToHost));
          same = memcmp(k0.data(), kx.data(), k0.size() * 8) == 0;
        }
      }
      if (same) {
        std::vector<int32_t> colmap;
        int C = 0;
        for (size_t i = 0; i < E; ++i) {
          ed[i].core_col0 = C;
          for (int r = 0; r < g.hs[i]->S; ++r) colmap.push_back((int32_t)(i << 8) | r);
          C += g.hs[i]->S;
        }
        HIP_OK(dalloc(&g.d_X, (size_t)h0->ncore * C));
        HIP_OK(dalloc(&g.d_Y, (size_t)h0->ncore * C));
        HIP_OK(dalloc(&g.d_colmap, colmap.size()));
        HIP_OK(hipMemcpy(g.d_colmap, colmap.data(), colmap.size() * 4, hipMemcpyHostToDevice));
        g.shared_core = true;
        g.ccols = C;
      }
    }
    HIP_OK(dalloc(&g.d_ed, E));
    HIP_OK(hipMemcpy(g.d_ed, ed.data(), sizeof(EngDev) * E, hipMemcpyHostToDevice));
    for (int k = 0; k < kLkKinds; ++k) {
      if (map[k].empty()) continue;
      HIP_OK(dalloc(&g.d_map[k], map[k].size()));
      HIP_OK(hipMemcpy(g.d_map[k], map[k].data(), map[k].size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    }
  }
  while (b->side.size() < b->groups.size()) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    b->side.push_back(st);
    HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    b->join.push_back(ev);
  }
  b->groups_ok = true;
  for (ngpu_t* h : b->hs) h->graph_dirty = false;
  return NGPU_OK;
}

// Launches of one batch group on stream st.  The group's first engine holds
// the stage events (launch() times the stages it is asked to), so its stage
// times are the group's.
template <typename F, typename... Args>
void glaunch(ngpu_t* h0, int open, int close, F kern, int nblocks, int threads, size_t lds, Args... args) {
  if (nblocks <= 0) return;
  launch_lds(h0, open, close, kern, dim3((unsigned)nblocks), dim3((unsigned)threads), lds, args...);
}

template <int W, int R, int IT>
void group_sort_merge(BatchGroup& g) {
  ngpu_t* h0 = g.hs[0];
  const int E = (int)g.hs.size();
  const EngDev* ed = g.d_ed;
  int cur = 0;
  if (g.rank_sort) {
    glaunch(h0, 4, -1, km_radix_hist_ranks<IT>, g.tot[kLkChunk], kBlock, 0, ed, (const uint16_t*)g.d_map[kLkChunk]);
    glaunch(h0, -1, -1, km_radix_scan, E * kSegs, kBlock, 0, ed, kSegs);
    glaunch(h0, -1, -1, km_radix_scatter_ranks<IT>, g.tot[kLkChunk], kBlock, 0, ed,
            (const uint16_t*)g.d_map[kLkChunk]);
    glaunch(h0, -1, 4, km_seg_sort, E * kSegs, kSegThreads, 0, ed);
    glaunch(h0, 6, 6, km_merge_annihilate<W, R, uint32_t, 1>, g.tot[kLkMerge], kBlock, 0, ed,
            (const uint16_t*)g.d_map[kLkMerge], 0);
    return;
  }
  if (g.rank64) {
    // 64-bit rank keys into sk[0], the top-byte partition into sk[1], the
    // segments sorted back into sk[0]
    const int bits = g.rank64_bits;
    glaunch(h0, 4, -1, km_rank64_hist<IT>, g.tot[kLkChunk], kBlock, 0, ed, (const uint16_t*)g.d_map[kLkChunk], bits);
    glaunch(h0, -1, -1, km_radix_scan, E * kSegs, kBlock, 0, ed, kSegs);
    glaunch(h0, -1, -1, km_radix_scatter_words<uint64_t, 1, false, IT, kSegBits>, g.tot[kLkChunk], kBlock, 0, ed,
            (const uint16_t*)g.d_map[kLkChunk], 2, 0, 0, bits - kSegBits);
    glaunch(h0, -1, 4, km_seg_sort64, E * kSegs, kSegThreads, 0, ed, (bits - kSegBits + 7) / 8);
    glaunch(h0, 6, 6, km_merge_annihilate<W, R, uint64_t, W>, g.tot[kLkMerge], kBlock, 0, ed,
            (const uint16_t*)g.d_map[kLkMerge], 0);
    return;
  }
  const int np = (int)g.passes.size();
  for (int p = 0; p < np; ++p) {
    const int in = (p & 1) ^ 1;  // pass p writes sk[p & 1]
    const int first = p == 0 ? 1 : 0;  // the first pass reads the records' words
    glaunch(h0, p == 0 ? 4 : -1, -1, km_radix_hist_words<uint64_t, W, W == 2, IT, 8>, g.tot[kLkChunk], kBlock, 0,
            ed, (const uint16_t*)g.d_map[kLkChunk], first, in, g.passes[p].first, g.passes[p].second);
    glaunch(h0, -1, -1, km_radix_scan, E * kDigitBlocks, kBlock, 0, ed, kDigitBlocks);
    glaunch(h0, -1, p + 1 == np ? 4 : -1, km_radix_scatter_words<uint64_t, W, W == 2, IT, 8>, g.tot[kLkChunk], kBlock,
            0, ed, (const uint16_t*)g.d_map[kLkChunk], first, in, g.passes[p].first, g.passes[p].second);
    cur = p & 1;
  }
  glaunch(h0, 6, 6, km_merge_annihilate<W, R, uint64_t, W>, g.tot[kLkMerge], kBlock, 0, ed,
          (const uint16_t*)g.d_map[kLkMerge], cur);
}

template <int W, int R, int G>
void group_spawn(BatchGroup& g) {
  glaunch(g.hs[0], 2, 2, km_spawn<W, R, G>, g.tot[kLkGrid], kBlock, 0, (const EngDev*)g.d_ed,
          (const uint16_t*)g.d_map[kLkGrid], g.rank_sort ? 1 : 0);
}

template <int W, int R, int PL, int GL>
void group_singles(BatchGroup& g, int nsp) {
  const size_t lds = 2 * sizeof(double) * (size_t)(kBlock / GL) * nsp;  // weights and running sums
  glaunch(g.hs[0], 3, 3, km_spawn_singles<W, R, PL, GL>, g.tot[kLkGrid], kBlock, lds, (const EngDev*)g.d_ed,
          (const uint16_t*)g.d_map[kLkGrid], g.rank_sort ? 1 : 0);
}

template <int W, int R>
void group_chain_wr(BatchGroup& g) {
  ngpu_t* h0 = g.hs[0];
  const EngDev* ed = g.d_ed;
  const uint16_t* const* m = g.d_map;
  glaunch(h0, 0, 0, km_prepass<W, R>, g.tot[kLkChunk], kBlock, 0, ed, (const uint16_t*)m[kLkChunk]);
  if (g.core && g.shared_core) {
    glaunch(h0, -1, -1, km_core_gather_shared<R>, g.tot[kLkCore], kBlock, 0, ed, (const uint16_t*)m[kLkCore], g.d_X,
            g.ccols);
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>((h0->ncore + kBlock / kWave - 1) / (kBlock / kWave),
                                                                8 * (int64_t)h0->nb_grid));
    if (g.core_fp32)
      glaunch(h0, 9, 9, km_core_spmv_shared<float>, nb, kBlock, 0, ed, (const int32_t*)g.d_colmap,
              (const double*)g.d_X, g.d_Y, g.ccols);
    else
      glaunch(h0, 9, 9, km_core_spmv_shared<double>, nb, kBlock, 0, ed, (const int32_t*)g.d_colmap,
              (const double*)g.d_X, g.d_Y, g.ccols);
  } else if (g.core) {
    glaunch(h0, -1, -1, km_core_gather<R>, g.tot[kLkCore], kBlock, 0, ed, (const uint16_t*)m[kLkCore]);
    if (g.core_fp32)
      glaunch(h0, 9, 9, km_core_spmv<R, float>, g.tot[kLkCoreRow], kBlock, 0, ed, (const uint16_t*)m[kLkCoreRow]);
    else
      glaunch(h0, 9, 9, km_core_spmv<R, double>, g.tot[kLkCoreRow], kBlock, 0, ed, (const uint16_t*)m[kLkCoreRow]);
  }
  glaunch(h0, 1, 1, km_scan_apply<W>, g.tot[kLkChunk], kBlock, 0, ed, (const uint16_t*)m[kLkChunk]);
  if (g.gen == kGenHubbard) group_spawn<W, R, kGenHubbard>(g);
  else if (g.gen == kGenUniform) group_spawn<W, R, kGenUniform>(g);
  else if (g.gen == kGenPchbW) group_spawn<W, R, kGenPchbW>(g);
  else group_spawn<W, R, kGenPchb>(g);
  if (g.singles) {
    const int nsp = h0->sys.nbasis / 2;
    switch (g.lanes * 100 + g.pl) {
      case 1601: group_singles<W, R, 1, 16>(g, nsp); break;
      case 1602: group_singles<W, R, 2, 16>(g, nsp); break;
      case 1604: group_singles<W, R, 4, 16>(g, nsp); break;
      case 802: group_singles<W, R, 2, 8>(g, nsp); break;
      case 803: group_singles<W, R, 3, 8>(g, nsp); break;
      case 804: group_singles<W, R, 4, 8>(g, nsp); break;
      case 805: group_singles<W, R, 5, 8>(g, nsp); break;
      case 808: group_singles<W, R, 8, 8>(g, nsp); break;
      case 404: group_singles<W, R, 4, 4>(g, nsp); break;
      case 405: group_singles<W, R, 5, 4>(g, nsp); break;
      case 408: group_singles<W, R, 8, 4>(g, nsp); break;
      case 410: group_singles<W, R, 10, 4>(g, nsp); break;
      default: group_singles<W, R, 16, 4>(g, nsp); break;
    }
  }
  if (g.core && g.shared_core)
    glaunch(h0, -1, -1, km_core_apply_shared<R>, g.tot[kLkCore], kBlock, 0, ed, (const uint16_t*)m[kLkCore],
            (const double*)g.d_Y, g.ccols);
  else if (g.core)
    glaunch(h0, -1, -1, km_core_apply<R>, g.tot[kLkCore], kBlock, 0, ed, (const uint16_t*)m[kLkCore]);
  if (g.items == 2) group_sort_merge<W, R, 2>(g);
  else if (g.items == 8) group_sort_merge<W, R, 8>(g);
  else group_sort_merge<W, R, 4>(g);
  glaunch(h0, 7, 7, km_insert<W, R>, g.tot[kLkGrid], kBlock, 0, ed, (const uint16_t*)m[kLkGrid]);
  glaunch(h0, 8, 8, km_stats<W, R>, g.tot[kLkGrid], kBlock, 0, ed, (const uint16_t*)m[kLkGrid]);
  glaunch(h0, -1, -1, km_end_iter, (int)g.hs.size(), kBlock, 0, ed);
}

void group_chain(BatchGroup& g) {
  if (g.W == 1) g.S == 1 ? group_chain_wr<1, 1>(g) : group_chain_wr<1, kMaxRep>(g);
  else g.S == 1 ? group_chain_wr<2, 1>(g) : group_chain_wr<2, kMaxRep>(g);
}

// one iteration of every group, each on its own stream forked from and
// joined to the batch stream (captured into the batch graph, or launched
// directly)
int batch_chains(ngpu_batch* b) {
  HIP_OK(hipEventRecord(b->fork, b->stream));
  for (size_t i = 0; i < b->groups.size(); ++i) {
    BatchGroup& g = b->groups[i];
    ngpu_t* h0 = g.hs[0];
    HIP_OK(hipStreamWaitEvent(b->side[i], b->fork, 0));
    h0->stream = b->side[i];
    const bool timed = h0->prof && h0->ev_ok && !b->use_graph;
    const bool prof = h0->prof;
    if (timed) {
      h0->ev_cur = (int)(h0->ev_next % kEvRing);
      harvest(h0, h0->ev_cur);
      h0->ev_mask[h0->ev_cur] = 0;
    } else {
      h0->prof = false;  // no events inside a captured graph
    }
    group_chain(g);
    h0->prof = prof;
    if (timed) {
      h0->ev_pending[h0->ev_cur] = true;
      h0->ev_next++;
    }
    h0->stream = b->stream;
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(b->join[i], b->side[i]));
    HIP_OK(hipStreamWaitEvent(b->stream, b->join[i], 0));
  }
  return NGPU_OK;
}

int capture_batch(ngpu_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  hipGraph_t g = nullptr;
  HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = batch_chains(b);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return NGPU_ERR_HIP;
  e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->gexec = nullptr;
    return NGPU_ERR_HIP;
  }
  return NGPU_OK;
}

// a set-up call on an engine (core space, precision, launch shape, ...)
// changes its descriptor and possibly its group: rebuild them, and the graph
int refresh_groups(ngpu_batch* b) {
  bool stale = !b->groups_ok;
  for (ngpu_t* h : b->hs) stale = stale || h->graph_dirty;
  if (!stale) return NGPU_OK;
  HIP_OK(hipStreamSynchronize(b->stream));
  int rc = build_groups(b);
  if (rc) return rc;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  return NGPU_OK;
}

void batch_release(ngpu_batch* b) {
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (ngpu_t* h : b->hs) {
    h->stream = h->own_stream;
    h->dparams = h->own_params;
    h->batch = nullptr;
    h->graph_dirty = true;
  }
  for (BatchGroup& g : b->groups) free_group(g);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  for (hipStream_t st : b->side) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : b->join) (void)hipEventDestroy(ev);
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (int q = 0; q < ngpu_batch::kRing; ++q) {
    if (b->hparams[q]) (void)hipHostFree(b->hparams[q]);
    if (b->slot_done[q]) (void)hipEventDestroy(b->slot_done[q]);
  }
  if (b->dparams) (void)hipFree(b->dparams);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // namespace
}  // extern "C++"

int ngpu_batch_create(ngpu_batch** out, ngpu_t* const* hs, int32_t n) {
  if (!out || !hs || n < 1) return NGPU_ERR_ARG;
  *out = nullptr;
  for (int32_t i = 0; i < n; ++i) {
    int rc = ready(hs[i]);
    if (rc) return rc;
    if (hs[i]->batch) return NGPU_ERR_STATE;
    for (int32_t j = 0; j < i; ++j)
      if (hs[j] == hs[i]) return NGPU_ERR_ARG;
  }
  auto* b = new ngpu_batch;
  bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
            hipMalloc((void**)&b->dparams, sizeof(IterParams) * (size_t)n) == hipSuccess;
  for (int q = 0; ok && q < ngpu_batch::kRing; ++q)
    ok = hipHostMalloc((void**)&b->hparams[q], sizeof(IterParams) * (size_t)n, hipHostMallocDefault) == hipSuccess &&
         hipEventCreateWithFlags(&b->slot_done[q], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    batch_release(b);
    return NGPU_ERR_HIP;
  }
  for (int32_t i = 0; i < n; ++i) {
    ngpu_t* h = hs[i];
    if (h->pending) (void)finish(h, nullptr);
    (void)hipStreamSynchronize(h->stream);
    h->own_stream = h->stream;
    h->stream = b->stream;
    h->dparams = b->dparams + i;
    h->batch = b;
    h->graph_dirty = true;
    b->hs.push_back(h);
  }
  *out = b;
  return NGPU_OK;
}

int ngpu_batch_iterate(ngpu_batch* b, const ngpu_iter_in_t* ins) {
  if (!b || !ins) return NGPU_ERR_ARG;
  const size_t n = b->hs.size();
  for (size_t i = 0; i < n; ++i)
    if (!(ins[i].tau > 0.0)) return NGPU_ERR_ARG;
  // the staging slot is rewritten only after its previous copy has run
  const int q = b->next_slot;
  if (b->slot_used[q]) HIP_OK(hipEventSynchronize(b->slot_done[q]));
  for (size_t i = 0; i < n; ++i)
    b->hparams[q][i] = make_params(b->hs[i], ins[i].tau, ins[i].diag_sft, ins[i].iter);
  HIP_OK(hipMemcpyAsync(b->dparams, b->hparams[q], sizeof(IterParams) * n, hipMemcpyHostToDevice, b->stream));
  HIP_OK(hipEventRecord(b->slot_done[q], b->stream));
  b->slot_used[q] = true;
  b->next_slot = (q + 1) % ngpu_batch::kRing;
  int rc = refresh_groups(b);
  if (rc) return rc;
  if (b->use_graph) {
    if (!b->gexec) {
      rc = capture_batch(b);
      if (rc) return rc;
    }
    HIP_OK(hipGraphLaunch(b->gexec, b->stream));
  } else {
    rc = batch_chains(b);
    if (rc) return rc;
  }
  for (ngpu_t* h : b->hs) h->pending = true;
  return NGPU_OK;
}

int ngpu_batch_set_graph(ngpu_batch* b, int32_t on) {
  if (!b) return NGPU_ERR_ARG;
  b->use_graph = on != 0;
  return NGPU_OK;
}

int ngpu_batch_get_groups(ngpu_batch* b, int32_t* group_of_engine, int32_t* n_groups) {
  if (!b) return NGPU_ERR_ARG;
  int rc = refresh_groups(b);
  if (rc) return rc;
  if (n_groups) *n_groups = (int32_t)b->groups.size();
  if (group_of_engine)
    for (size_t i = 0; i < b->hs.size(); ++i)
      for (size_t g = 0; g < b->groups.size(); ++g)
        for (ngpu_t* h : b->groups[g].hs)
          if (h == b->hs[i]) group_of_engine[i] = (int32_t)g;
  return NGPU_OK;
}

int ngpu_batch_get_group_flags(ngpu_batch* b, int32_t group, int32_t* flags) {
  if (!b || !flags) return NGPU_ERR_ARG;
  int rc = refresh_groups(b);
  if (rc) return rc;
  if (group < 0 || group >= (int32_t)b->groups.size()) return NGPU_ERR_ARG;
  const BatchGroup& g = b->groups[(size_t)group];
  *flags = (g.core ? NGPU_GROUP_CORE : 0) | (g.shared_core ? NGPU_GROUP_SHARED_CORE : 0);
  return NGPU_OK;
}

int ngpu_batch_set_profiling(ngpu_batch* b, int32_t on) {
  if (!b) return NGPU_ERR_ARG;
  for (ngpu_t* h : b->hs) {
    h->prof = on != 0;
    h->prof_mask = (uint32_t)on & ((1u << kStages) - 1u);
    for (int s = 0; s < kStages; ++s) h->stage_ms[s] = 0.0;
    h->prof_iters = 0;
  }
  return NGPU_OK;
}

int ngpu_batch_destroy(ngpu_batch* b) {
  if (!b) return NGPU_ERR_ARG;
  batch_release(b);
  return NGPU_OK;
}

int ngpu_get_stats(ngpu_t* h, ngpu_iter_out_t* out) {
  if (!h) return NGPU_ERR_ARG;
  HIP_OK(hipMemcpyAsync(h->hout, h->dout, sizeof(ngpu_iter_out_t), hipMemcpyDeviceToHost, h->stream));
  return finish(h, out);
}

int ngpu_get_hash_stats(ngpu_t* h, int64_t* live, int64_t* tombstones, int64_t* entries, int64_t* rebuilds) {
  if (!h) return NGPU_ERR_ARG;
  HIP_OK(hipStreamSynchronize(h->stream));
  HIP_OK(hipMemcpy(h->hcnt, h->cnt, sizeof(Counters), hipMemcpyDeviceToHost));
  if (live) *live = (int64_t)h->hcnt->n_live;
  if (tombstones) *tombstones = (int64_t)h->hcnt->n_tomb;
  if (entries) *entries = h->wk.L * (int64_t)kBucketEntries;
  if (rebuilds) *rebuilds = (int64_t)h->hcnt->n_rebuilds;
  return NGPU_OK;
}

int ngpu_get_cycle_stats(ngpu_t* h, ngpu_iter_out_t* sum, ngpu_iter_out_t* last, int32_t* n_iters) {
  if (!h) return NGPU_ERR_ARG;
  HIP_OK(hipMemcpyAsync(h->hout, h->dout, sizeof(ngpu_iter_out_t), hipMemcpyDeviceToHost, h->stream));
  ngpu_iter_out_t tmp;
  HIP_OK(hipMemcpyAsync(&tmp, h->dcyc, sizeof(ngpu_iter_out_t), hipMemcpyDeviceToHost, h->stream));
  hipLaunchKernelGGL(k_reset_cycle, dim3(1), dim3(kBlock), 0, h->stream, h->dcyc);
  HIP_OK(hipMemcpyAsync(h->hcnt, h->cnt, sizeof(Counters), hipMemcpyDeviceToHost, h->stream));
  int rc = finish(h, last);
  if (sum) *sum = tmp;
  if (n_iters) *n_iters = tmp.pad;
  if (rc == NGPU_OK) rc = err_from_mask((unsigned)tmp.err);
  if (rc == NGPU_OK) rc = err_from_mask(h->hcnt->halt);  // halted in an earlier cycle
  return rc;
}

int ngpu_get_error(ngpu_t* h, int32_t* status, uint32_t* iter) {
  if (!h) return NGPU_ERR_ARG;
  HIP_OK(hipStreamSynchronize(h->stream));
  HIP_OK(hipMemcpy(h->hcnt, h->cnt, sizeof(Counters), hipMemcpyDeviceToHost));
  if (status) *status = err_from_mask(h->hcnt->halt);
  if (iter) *iter = h->hcnt->halt ? h->hcnt->err_iter : 0u;
  return NGPU_OK;
}

int ngpu_set_profiling(ngpu_t* h, int32_t on) {
  if (!h) return NGPU_ERR_ARG;
  if (h->batch) return NGPU_ERR_STATE;
  if (h->pending) (void)finish(h, nullptr);
  h->prof = on != 0;
  h->prof_mask = (uint32_t)on & ((1u << kStages) - 1u);
  for (int s = 0; s < kStages; ++s) h->stage_ms[s] = 0.0;
  h->prof_iters = 0;
  return NGPU_OK;
}

int ngpu_get_stage_times(ngpu_t* h, double* ms_total, int32_t nstages, int64_t* n_iters) {
  if (!h || !ms_total) return NGPU_ERR_ARG;
  for (int s = 0; s < nstages && s < kStages; ++s) ms_total[s] = h->stage_ms[s];
  if (n_iters) *n_iters = h->prof_iters;
  return NGPU_OK;
}

// ------------------------------------------------------------ parity entries
int ngpu_find_walker_hash_batch(ngpu_t* h, const int64_t* dets, int64_t n, int32_t* buckets) {
  if (!h || !h->have_hash || n < 0 || (n && (!dets || !buckets))) return NGPU_ERR_ARG;
  if (!n) return NGPU_OK;
  DevBuf<uint64_t> d;
  DevBuf<int32_t> o;
  HIP_OK(d.alloc((size_t)n * h->W));
  HIP_OK(o.alloc((size_t)n));
  HIP_OK(hipMemcpy(d.p, dets, (size_t)n * h->W * 8, hipMemcpyHostToDevice));
  if (h->W == 1) hipLaunchKernelGGL(k_batch_hash<1>, dim3(grid_for(n)), dim3(kBlock), 0, 0, h->ds, d.p, n, o.p);
  else hipLaunchKernelGGL(k_batch_hash<2>, dim3(grid_for(n)), dim3(kBlock), 0, 0, h->ds, d.p, n, o.p);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(buckets, o.p, (size_t)n * 4, hipMemcpyDeviceToHost));
  return NGPU_OK;
}

int ngpu_get_spawn_helement_batch(ngpu_t* h, const int64_t* a, const int64_t* b, int64_t n, double* hel) {
  if (!h || !h->have_model || n < 0 || (n && (!a || !b || !hel))) return NGPU_ERR_ARG;
  if (!n) return NGPU_OK;
  DevBuf<uint64_t> da, db;
  DevBuf<double> o;
  HIP_OK(da.alloc((size_t)n * h->W));
  HIP_OK(db.alloc((size_t)n * h->W));
  HIP_OK(o.alloc((size_t)n));
  HIP_OK(hipMemcpy(da.p, a, (size_t)n * h->W * 8, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(db.p, b, (size_t)n * h->W * 8, hipMemcpyHostToDevice));
  if (h->W == 1) hipLaunchKernelGGL(k_batch_hel<1>, dim3(grid_for(n)), dim3(kBlock), 0, 0, h->ds, da.p, db.p, n, o.p);
  else hipLaunchKernelGGL(k_batch_hel<2>, dim3(grid_for(n)), dim3(kBlock), 0, 0, h->ds, da.p, db.p, n, o.p);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(hel, o.p, (size_t)n * 8, hipMemcpyDeviceToHost));
  return NGPU_OK;
}

int ngpu_diag_helement_batch(ngpu_t* h, const int64_t* dets, int64_t n, double* diag, double* offdiag) {
  if (!h || !h->have_model || !h->have_ref || n < 0 || (n && (!dets || !diag || !offdiag))) return NGPU_ERR_ARG;
  if (!n) return NGPU_OK;
  DevBuf<uint64_t> d;
  DevBuf<double> o1, o2;
  HIP_OK(d.alloc((size_t)n * h->W));
  HIP_OK(o1.alloc((size_t)n));
  HIP_OK(o2.alloc((size_t)n));
  HIP_OK(hipMemcpy(d.p, dets, (size_t)n * h->W * 8, hipMemcpyHostToDevice));
  if (h->W == 1) hipLaunchKernelGGL(k_batch_diag<1>, dim3(grid_for(n)), dim3(kBlock), 0, 0, h->ds, d.p, n, o1.p, o2.p);
  else hipLaunchKernelGGL(k_batch_diag<2>, dim3(grid_for(n)), dim3(kBlock), 0, 0, h->ds, d.p, n, o1.p, o2.p);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(diag, o1.p, (size_t)n * 8, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(offdiag, o2.p, (size_t)n * 8, hipMemcpyDeviceToHost));
  return NGPU_OK;
}

int ngpu_generate_excitation_batch(ngpu_t* h, const int64_t* ilut, const double* draws, int64_t n, int64_t* ilut_j,
                                   int32_t* ic, int32_t* ex, int32_t* parity, double* pgen) {
  if (!h || !h->have_model || n < 0) return NGPU_ERR_ARG;
  if (h->sys.model == NGPU_MODEL_MOLECULAR && !h->have_excitgen) return NGPU_ERR_STATE;
  if (!n) return NGPU_OK;
  if (!ilut || !draws || !ilut_j || !ic || !ex || !parity || !pgen) return NGPU_ERR_ARG;
  for (int64_t i = 0; i < n * 6; ++i)
    if (!(draws[i] >= 0.0 && draws[i] < 1.0)) return NGPU_ERR_ARG;
  for (int64_t i = 0; i < n; ++i) {
    int nel = 0;
    for (int w = 0; w < h->W; ++w) nel += __builtin_popcountll((uint64_t)ilut[i * h->W + w]);
    if (nel != h->sys.nel) return NGPU_ERR_ARG;
  }
  const int W = h->W;
  DevBuf<uint64_t> d, dj;
  DevBuf<double> dr, pg;
  DevBuf<int32_t> dic, dex, dpar;
  HIP_OK(d.alloc((size_t)n * W));
  HIP_OK(dj.alloc((size_t)n * W));
  HIP_OK(dr.alloc((size_t)n * 6));
  HIP_OK(pg.alloc((size_t)n));
  HIP_OK(dic.alloc((size_t)n));
  HIP_OK(dex.alloc((size_t)n * 4));
  HIP_OK(dpar.alloc((size_t)n));
  HIP_OK(hipMemcpy(d.p, ilut, (size_t)n * W * 8, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dr.p, draws, (size_t)n * 6 * 8, hipMemcpyHostToDevice));
  if (W == 1)
    hipLaunchKernelGGL(k_batch_gen<1>, dim3(grid_for(n)), dim3(kBlock), 0, 0, h->ds, d.p, dr.p, n, dj.p, dic.p, dex.p,
                       dpar.p, pg.p);
  else
    hipLaunchKernelGGL(k_batch_gen<2>, dim3(grid_for(n)), dim3(kBlock), 0, 0, h->ds, d.p, dr.p, n, dj.p, dic.p, dex.p,
                       dpar.p, pg.p);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(ilut_j, dj.p, (size_t)n * W * 8, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(ic, dic.p, (size_t)n * 4, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(ex, dex.p, (size_t)n * 16, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(parity, dpar.p, (size_t)n * 4, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(pgen, pg.p, (size_t)n * 8, hipMemcpyDeviceToHost));
  return NGPU_OK;
}

int ngpu_calc_pgen_batch(ngpu_t* h, const int64_t* ilut, const int32_t* ic, const int32_t* ex, int64_t n,
                         double* pgen) {
  if (!h || !h->have_model || n < 0) return NGPU_ERR_ARG;
  if (h->sys.model == NGPU_MODEL_MOLECULAR && !h->have_excitgen) return NGPU_ERR_STATE;
  if (!n) return NGPU_OK;
  if (!ilut || !ic || !ex || !pgen) return NGPU_ERR_ARG;
  const int W = h->W;
  for (int64_t i = 0; i < n; ++i) {
    if (ic[i] != 1 && ic[i] != 2) return NGPU_ERR_ARG;
    for (int q = 0; q < 2 * ic[i]; ++q)
      if (ex[4 * i + q] < 1 || ex[4 * i + q] > h->sys.nbasis) return NGPU_ERR_ARG;
  }
  DevBuf<uint64_t> d;
  DevBuf<int32_t> dic, dex;
  DevBuf<double> o;
  HIP_OK(d.alloc((size_t)n * W));
  HIP_OK(dic.alloc((size_t)n));
  HIP_OK(dex.alloc((size_t)n * 4));
  HIP_OK(o.alloc((size_t)n));
  HIP_OK(hipMemcpy(d.p, ilut, (size_t)n * W * 8, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dic.p, ic, (size_t)n * 4, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dex.p, ex, (size_t)n * 16, hipMemcpyHostToDevice));
  if (W == 1)
    hipLaunchKernelGGL(k_batch_pgen<1>, dim3(grid_for(n)), dim3(kBlock), 0, 0, h->ds, d.p, dic.p, dex.p, n, o.p);
  else
    hipLaunchKernelGGL(k_batch_pgen<2>, dim3(grid_for(n)), dim3(kBlock), 0, 0, h->ds, d.p, dic.p, dex.p, n, o.p);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(pgen, o.p, (size_t)n * 8, hipMemcpyDeviceToHost));
  return NGPU_OK;
}

int ngpu_encode_child(const int64_t* ilut_i, int32_t nwords, int32_t ic, const int32_t* ex, int64_t* ilut_j) {
  if (!ilut_i || !ex || !ilut_j || nwords < 1 || nwords > 2 || (ic != 1 && ic != 2)) return NGPU_ERR_ARG;
  for (int w = 0; w < nwords; ++w) ilut_j[w] = ilut_i[w];
  for (int q = 0; q < 2 * ic; ++q) {
    int orb = ex[q];
    if (orb < 1 || orb > 64 * nwords) return NGPU_ERR_ARG;
    uint64_t bit = 1ull << ((orb - 1) & 63);
    if (q < ic) ilut_j[(orb - 1) >> 6] = (int64_t)((uint64_t)ilut_j[(orb - 1) >> 6] & ~bit);
    else ilut_j[(orb - 1) >> 6] = (int64_t)((uint64_t)ilut_j[(orb - 1) >> 6] | bit);
  }
  return NGPU_OK;
}

int ngpu_attempt_die_batch(ngpu_t* h, const double* kii, const double* sign, const double* udraw, int64_t n,
                           double tau, double diag_sft, double* ndie) {
  if (!h || n < 0 || (n && (!kii || !sign || !udraw || !ndie))) return NGPU_ERR_ARG;
  if (!n) return NGPU_OK;
  DevBuf<double> a, b, c, o;
  HIP_OK(a.alloc((size_t)n));
  HIP_OK(b.alloc((size_t)n));
  HIP_OK(c.alloc((size_t)n));
  HIP_OK(o.alloc((size_t)n));
  HIP_OK(hipMemcpy(a.p, kii, (size_t)n * 8, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(b.p, sign, (size_t)n * 8, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c.p, udraw, (size_t)n * 8, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_batch_die, dim3(grid_for(n)), dim3(kBlock), 0, 0, h->sys.t_all_real_coeff != 0, a.p, b.p, c.p, n,
                     tau, diag_sft, o.p);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(ndie, o.p, (size_t)n * 8, hipMemcpyDeviceToHost));
  return NGPU_OK;
}

int ngpu_attempt_create_batch(ngpu_t* h, const double* hel, const double* pgen, const double* parent_sign,
                              const double* udraw, int64_t n, double tau, double* child) {
  if (!h || n < 0 || (n && (!hel || !pgen || !parent_sign || !udraw || !child))) return NGPU_ERR_ARG;
  if (!n) return NGPU_OK;
  DevBuf<double> a, b, c, d, o;
  HIP_OK(a.alloc((size_t)n));
  HIP_OK(b.alloc((size_t)n));
  HIP_OK(c.alloc((size_t)n));
  HIP_OK(d.alloc((size_t)n));
  HIP_OK(o.alloc((size_t)n));
  HIP_OK(hipMemcpy(a.p, hel, (size_t)n * 8, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(b.p, pgen, (size_t)n * 8, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c.p, parent_sign, (size_t)n * 8, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d.p, udraw, (size_t)n * 8, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_batch_create, dim3(grid_for(n)), dim3(kBlock), 0, 0, h->ds, a.p, b.p, c.p, d.p, n, tau, o.p);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(child, o.p, (size_t)n * 8, hipMemcpyDeviceToHost));
  return NGPU_OK;
}

int ngpu_annihilate(ngpu_t* h, uint32_t iter, const int64_t* sdet, const double* sval, const uint32_t* sinfo,
                    int64_t n, ngpu_iter_out_t* out, int64_t* merged_det, double* merged_sgn, uint32_t* merged_flags,
                    int64_t* n_merged) {
  int rc = ready(h);
  if (rc) return rc;
  if (n < 0 || (n && (!sdet || !sval || !sinfo))) return NGPU_ERR_ARG;
  if (n > h->scap) return NGPU_ERR_SPAWN_FULL;
  for (int64_t i = 0; i < n; ++i)
    if ((int)(sinfo[i] & 0xFFu) >= h->S) return NGPU_ERR_ARG;
  if (h->pending) {
    rc = finish(h, nullptr);
    if (rc) return rc;
  }
  HIP_OK(hipMemcpy(h->hcnt, h->cnt, sizeof(Counters), hipMemcpyDeviceToHost));
  if (h->hcnt->halt) return err_from_mask(h->hcnt->halt);
  const int W = h->W, S = h->S;
  if (!h->u_key) {  // the merged list is kept only for this entry point
    HIP_OK(hipMalloc((void**)&h->u_key, (size_t)h->scap * W * 8));
    HIP_OK(hipMalloc((void**)&h->u_sgn, (size_t)h->scap * S * 8));
    HIP_OK(hipMalloc((void**)&h->u_flag, (size_t)h->scap * 4));
    HIP_OK(hipMalloc((void**)&h->u_count, 8));
  }
  if (n) {
    const int rw = h->ds.rec_words;
    std::vector<uint64_t> rec((size_t)n * rw, 0);
    for (int64_t i = 0; i < n; ++i) {
      uint64_t* r = rec.data() + (size_t)i * rw;
      for (int w = 0; w < W; ++w) r[w] = (uint64_t)sdet[i * W + w];
      if (h->ds.spawn_fp32) {  // FP32 amplitudes: each value rounded to nearest on entry
        float f = (float)sval[i];
        uint32_t fb;
        memcpy(&fb, &f, 4);
        r[W] = ((uint64_t)sinfo[i] << 32) | fb;
      } else {
        uint64_t vb;
        memcpy(&vb, &sval[i], 8);
        r[2] = vb;
        r[3] = sinfo[i];
      }
    }
    HIP_OK(hipMemcpy(h->sp_rec, rec.data(), rec.size() * 8, hipMemcpyHostToDevice));
  }
  hipLaunchKernelGGL(k_begin_iter, dim3(1), dim3(kBlock), 0, h->stream, h->dout, h->dpart, h->cnt, h->dparams,
                     IterParams{}, 0);
  hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, h->stream, &h->cnt->n_spawned, (unsigned long long)n);
  hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, h->stream, h->u_count, 0ull);
  bool p = h->prof;
  h->prof = false;
  // an explicit list carries no deterministic projection
  set_params(h, 1.0, nullptr, iter);
  if (W == 1) h->S == 1 ? annihilation_chain<1, 1>(h, false, true) : annihilation_chain<1, kMaxRep>(h, false, true);
  else h->S == 1 ? annihilation_chain<2, 1>(h, false, true) : annihilation_chain<2, kMaxRep>(h, false, true);
  h->prof = p;
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(h->hout, h->dout, sizeof(ngpu_iter_out_t), hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipMemcpyAsync(h->hcnt, h->cnt, sizeof(Counters), hipMemcpyDeviceToHost, h->stream));
  h->pending = true;
  rc = finish(h, out);
  unsigned long long nu_dev = 0;
  HIP_OK(hipMemcpy(&nu_dev, h->u_count, 8, hipMemcpyDeviceToHost));
  int64_t nu = (int64_t)nu_dev;
  if (n_merged) *n_merged = nu;
  if (nu > 0 && merged_det && merged_sgn && merged_flags) {
    HIP_OK(hipMemcpy(merged_det, h->u_key, (size_t)nu * W * 8, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(merged_sgn, h->u_sgn, (size_t)nu * S * 8, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(merged_flags, h->u_flag, (size_t)nu * 4, hipMemcpyDeviceToHost));
  }
  return rc;
}

}  // extern "C"
