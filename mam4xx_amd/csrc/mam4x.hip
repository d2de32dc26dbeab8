// MI355X (gfx950) implementation of the MAM4 microphysics chain and its C ABI
// (include/mam4x.h).
//
// Cells are enumerated level-major / column-minor (c = k * ncol + i), so
// consecutive lanes read consecutive addresses of each [level][column] plane
// (fully coalesced 64-lane loads).  The chain is FP64 VALU work with no dense
// contraction, so there is no MFMA or LDS tiling; the sigma-only constants
// come from a small device buffer through the scalar cache.
//
// The cost of a cell is dominated by a data-dependent loop: the adaptive SOA
// sub-stepping (gasaerexch_soaexch.hpp:170-335) takes 1 to 1000 sub-steps, a
// heavy-tailed count.  One lane per cell would let every wave run as long as
// its slowest lane, so a step is four launches:
//   1. chain_pre_kernel   per cell: subareas, clear-subarea state, H2SO4
//                         uptake rates and the SOA loop inputs -> workspace;
//                         a predicted sub-step count -> cost bucket
//   2. queue_scatter_kernel  cells with SOA work -> a queue ordered by
//                         bucket, longest first
//   3. soa_queue_kernel   persistent waves drain the queue; a lane takes the
//                         next cell as soon as its current one has converged
//   4. chain_post_kernel  per cell: the rest of the chain, recombination,
//                         calcsize, tendencies
// chain_fused_kernel runs everything in one lane per cell (small domains).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <array>
#include <new>
#include <mutex>
#include <string>
#include <vector>
#include <algorithm>

#include "../../include/mam4x.h"
#include "chain_adjoint.hpp"
#include "chain_cell.hpp"
#include "gas_chem.hpp"
#include "linoz.hpp"
#include "photo.hpp"
#include "sethet.hpp"
#include "setsox.hpp"
#include "water_uptake.hpp"

using namespace mam4x;

namespace {

constexpr int kBlock = 256;
constexpr int kNBucket = 24;  // half-octave cost buckets: 2 log2(1 + predicted sub-steps)
#ifndef MAM4X_QUEUE_CHUNK
#define MAM4X_QUEUE_CHUNK 64
#endif
constexpr int kChunk = MAM4X_QUEUE_CHUNK;  // default queue entries a wave takes per refill
// Waves whose first chunk comes from buckets >= kLongBucket (predicted >= ~180
// sub-steps) raise their issue priority: a cell's sub-steps are a serial
// chain, so the longest cells set the kernel's critical path and should not
// yield issue slots to the bulk.
#ifndef MAM4X_SOA_PRIO
#define MAM4X_SOA_PRIO 1
#endif
constexpr int kLongBucket = 15;
// Refills cost a device-scope atomic on a queue counter, and those atomics
// serialise across the device (tens of ns each), so the kernel keeps their
// number down: every wave's first chunk is assigned statically (chunk = wave
// index: the longest cells go to the first-dispatched, oldest waves), and the
// cells of buckets < kShortBucket (predicted <= ~30 sub-steps) form a second
// part of the queue taken kShortChunk at a time -- holding a few short cells
// back in a busy wave costs little.
#ifndef MAM4X_SHORT_BUCKET
#define MAM4X_SHORT_BUCKET 10
#endif
#ifndef MAM4X_SHORT_CHUNK
#define MAM4X_SHORT_CHUNK 256
#endif
#ifndef MAM4X_STATIC_FIRST
#define MAM4X_STATIC_FIRST 1
#endif
// Minimum waves per SIMD the reverse sweep's heaviest kernels are compiled
// for (1: as many registers as the compiler wants).
#ifndef MAM4X_REPLAY_WAVES
#define MAM4X_REPLAY_WAVES 1
#endif
#ifndef MAM4X_ADJ_PRE_WAVES
#define MAM4X_ADJ_PRE_WAVES 2
#endif
constexpr int kShortBucket = MAM4X_SHORT_BUCKET;
constexpr int kShortChunk = MAM4X_SHORT_CHUNK;
constexpr unsigned char kNoSoa = 0xff;
constexpr unsigned char kCapped = 0x80;  // bucket bit: the SOA loop hit its cap
// bucket bit: the cell's loop left the range where the queue kernel's fast
// quotients are IEEE's (dev::soa_fast_exact); soa_ieee_kernel reruns it
constexpr unsigned char kIeee = 0x40;
static_assert(kNBucket <= kIeee, "bucket indices stay below the flag bits");

// Workspace: 14 planes of one double per cell.  The set-up kernel writes
// them in cell order (coalesced); the queue kernel's chunks hold cells of a
// few set-up blocks, so its gathers stay on few lines.
//   plane 0 SOAG, 1..4 SOA per mode, 5 equilibrium SOAG g0,
//   6..9 H2SO4 uptake rates, 10..13 oxygenated POM per mode.
// The queue kernel overwrites planes 0..5 with its results: SOAG, SOA per
// mode and the time-averaged SOAG.
enum { W_G = 0, W_A = 1, W_G0 = 5, W_UPTK = 6, W_AOPOA = 10, W_NPLANE = 14 };
constexpr int W_GAVG = W_G0;

__device__ __forceinline__ void ws_store_in(double *ws, long n, long c,
                                            const dev::SoaIn<double> &in) {
  ws[W_G * n + c] = in.g;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    ws[(W_A + m) * n + c] = in.a[m];
    ws[(W_UPTK + m) * n + c] = in.uptk[m];
    ws[(W_AOPOA + m) * n + c] = in.aopoa[m];
  }
  ws[W_G0 * n + c] = in.g0;
}

__device__ __forceinline__ void ws_load_in(const double *ws, long n, long c,
                                           dev::SoaIn<double> &in) {
  in.g = ws[W_G * n + c];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    in.a[m] = ws[(W_A + m) * n + c];
    in.uptk[m] = ws[(W_UPTK + m) * n + c];
    in.aopoa[m] = ws[(W_AOPOA + m) * n + c];
  }
  in.g0 = ws[W_G0 * n + c];
}

struct QueueCtl {
  unsigned hist[kNBucket];  // cells per bucket
  unsigned fill[kNBucket];  // scatter cursor per bucket
  unsigned head;            // next free slot of the long part of the queue
  unsigned pad[31];         // the two counters on separate cache lines
  unsigned shead;           // next free slot of the short part
};

// The queue's granularity (mam4x_config.soa_chunk, soa_short_chunk,
// soa_short_bucket; performance only): queue entries a wave takes per refill
// from the long part and from the short part, the first bucket of the long
// part, and the first bucket whose waves raise their issue priority.
struct QueueShape {
  int chunk, short_chunk, short_bucket, long_bucket;
};

static QueueShape queue_shape(const mam4x_config &c) {
  QueueShape q;
  q.chunk = c.soa_chunk > 0 ? c.soa_chunk : kChunk;
  q.short_chunk = c.soa_short_chunk > 0 ? c.soa_short_chunk : kShortChunk;
  q.short_bucket = c.soa_short_bucket > 0 ? c.soa_short_bucket : kShortBucket;
  q.long_bucket = kLongBucket;
  return q;
}

struct Work {
  QueueShape qs;
  double *ws;              // W_NPLANE planes of ncell doubles
  unsigned char *bucket;   // per cell, kNoSoa when the cell has no SOA loop
  int *queue;              // cell indices, longest first
  QueueCtl *ctl;
  long ncell;              // cells of the launch (batches: padded to whole blocks)
  double dtfull;           // SOA loop length: the integer sub-step (:1319)
  unsigned long long *stats;  // SOA sub-step counter (NULL: off)
};

// Where a block's cells come from.  A launch covers either one domain, or a
// batch of scenes each padded to whole blocks, so a block never straddles
// two scenes.  scene() maps the launch-wide cell index (which also indexes
// the workspace) to the scene's argument block and its scene-local index.
struct OneDomain {
  ChainArgs a;
  __device__ __forceinline__ const ChainArgs &scene(long &c) const { return a; }
};

struct SceneBatch {
  const ChainArgs *__restrict__ args;  // one per scene
  const unsigned *__restrict__ blk0;   // first block of each scene, [nscene + 1]
  int nscene;
  __device__ __forceinline__ const ChainArgs &scene(long &c) const {
    const unsigned b = blockIdx.x;
    int lo = 0, hi = nscene;  // blk0[lo] <= b < blk0[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (blk0[mid] <= b)
        lo = mid;
      else
        hi = mid;
    }
    c = (long)(b - blk0[lo]) * kBlock + threadIdx.x;
    return args[lo];
  }
};

// cell c of the memory-order enumeration -> column and plane offset
__device__ __forceinline__ void cell_index(const ChainArgs &a, long c, int &icol, long &off) {
  if (a.layout == MAM4X_LAYOUT_COL_LEV) {
    icol = (int)(c / a.nlev);
    const int k = (int)(c - (long)icol * a.nlev);
    off = (long)icol * a.ldc + k;
  } else if (a.lpw > 1) {
    // each run of 64 cells is a tile of lpw levels x 64 / lpw columns
    const int cpw = 64 / a.lpw;
    const long t = c >> 6;
    const int lane = (int)(c & 63);
    const long tpr = a.ncol / cpw;  // tiles per level row (fill_args: exact)
    const long tl = t / tpr;
    icol = (int)((t - tl * tpr) * cpw + lane % cpw);
    const int k = (int)(tl * a.lpw + lane / cpw);
    off = (long)k * a.ldc + icol;
  } else {
    const int k = (int)(c / a.ncol);
    icol = (int)(c - (long)k * a.ncol);
    off = (long)k * a.ldc + icol;
  }
}

__device__ __forceinline__ void add_stats(unsigned long long *stats, unsigned long long v) {
  if (!stats) return;  // uniform branch: wave-reduce, one atomic per wave
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(stats, v);
}

template <bool PRE, class P, class Src, bool MULTI>
__global__ void __launch_bounds__(kBlock) chain_fused_kernel(Src src, Work w) {
  long c = (long)blockIdx.x * kBlock + threadIdx.x;
  const ChainArgs &a = src.scene(c);
  const long ncell = (long)a.ncol * a.nlev;
  unsigned long long niter = 0;
  if (c < ncell) {
    int icol;
    long off;
    cell_index(a, c, icol, off);
    niter = (unsigned long long)dev::chain_cell_fused<PRE, P, MULTI>(a, off, icol);
  }
  add_stats(w.stats, niter);
}

// One lane per cell on the work queue's order (ntsubstep > 1, and the
// tangent): pass 0 runs the queued cells longest first (by the first
// sub-step's predicted SOA cost), so a wave's lanes run sub-step loops of
// similar length; pass 1 the cells without a SOA loop.  Every cell's result
// is independent of the order.
template <bool PRE, class P, bool MULTI = true>
__global__ void __launch_bounds__(kBlock) chain_fused_order_kernel(OneDomain src, Work w,
                                                                   int pass) {
  const ChainArgs &a = src.a;
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  const long ncell = (long)a.ncol * a.nlev;
  long c = -1;
  if (pass == 0) {
    unsigned nq = 0;
    for (int j = 0; j < kNBucket; ++j) nq += w.ctl->hist[j];
    if (i < (long)nq) c = w.queue[i];
  } else if (i < ncell && w.bucket[i] == kNoSoa) {
    c = i;
  }
  unsigned long long niter = 0;
  if (c >= 0) {
    int icol;
    long off;
    cell_index(a, c, icol, off);
    niter = (unsigned long long)dev::chain_cell_fused<PRE, P, MULTI>(a, off, icol);
  }
  add_stats(w.stats, niter);
}

__global__ void queue_reset_kernel(QueueCtl *q) {
  const int t = threadIdx.x;
  if (t < kNBucket) {
    q->hist[t] = 0;
    q->fill[t] = 0;
  }
  if (t == 0) {
    q->head = 0;
    q->shead = 0;
  }
}

// Phase 1.
template <bool PRE, class P, class Src>
__global__ void __launch_bounds__(kBlock) chain_pre_kernel(Src src, Work w) {
  __shared__ unsigned lh[kNBucket];
  if (threadIdx.x < kNBucket) lh[threadIdx.x] = 0;
  __syncthreads();
  const long g = (long)blockIdx.x * kBlock + threadIdx.x;  // workspace index
  long c = g;
  const ChainArgs &a = src.scene(c);
  if (g < w.ncell && c >= (long)a.ncol * a.nlev) w.bucket[g] = kNoSoa;  // batch padding
  if (g < w.ncell && c < (long)a.ncol * a.nlev) {
    int icol;
    long off;
    cell_index(a, c, icol, off);
    dev::CellCtx<P> X;
    dev::cell_setup<PRE, P>(a, off, icol, X);
    unsigned char b = kNoSoa;
    if (X.cat != 1) {
      dev::Aer<typename P::R> S;
      dev::ClearGas<double> G;
      dev::clear_build<P>(a, off, X, S, G);
      dev::SoaIn<double> in;
      dev::clear_pre<P>(*a.K, X.T, X.p, X.dgw, S, G.g[0], in);
      ws_store_in(w.ws, w.ncell, g, in);
      dev::SoaState<double> st;
      dev::soa_init(st, in);
      const float pred = (float)dev::soa_predict(st, w.dtfull);
      int ib = kNBucket - 1;
      if (pred < 1.0e6f) ib = min(kNBucket - 1, (int)(2.0f * log2f(1.0f + pred)));
      b = (unsigned char)ib;
      atomicAdd(&lh[ib], 1u);
    }
    w.bucket[g] = b;
  }
  __syncthreads();
  if (threadIdx.x < kNBucket && lh[threadIdx.x]) atomicAdd(&w.ctl->hist[threadIdx.x], lh[threadIdx.x]);
}

// Phase 2: queue = cells grouped by bucket, highest bucket first.  Order
// within a bucket is arbitrary; every cell's result is independent of it.
__global__ void __launch_bounds__(kBlock) queue_scatter_kernel(Work w) {
  __shared__ unsigned cnt[kNBucket], base[kNBucket];
  if (threadIdx.x < kNBucket) cnt[threadIdx.x] = 0;
  __syncthreads();
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  unsigned char b = kNoSoa;
  unsigned r = 0;
  if (c < w.ncell) {
    b = w.bucket[c];
    if (b != kNoSoa) r = atomicAdd(&cnt[b], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kNBucket) {
    const int t = threadIdx.x;
    unsigned start = 0;  // cells in buckets above t
    for (int j = kNBucket - 1; j > t; --j) start += w.ctl->hist[j];
    base[t] = cnt[t] ? start + atomicAdd(&w.ctl->fill[t], cnt[t]) : 0;
  }
  __syncthreads();
  if (b != kNoSoa) w.queue[base[b] + r] = (int)c;
}

// Phase 3: persistent SOA loop.  Each wave takes a chunk of queue entries at
// a time (kChunk from the long part, kShortChunk from the short part); a lane
// whose cell has converged takes the next entry of the wave's chunk before
// the following sub-step.  A wave exits once the queue is exhausted and all
// its lanes are idle; the queue is finite and every cell ends after at most
// 1000 sub-steps, so every wave exits.
__global__ void __launch_bounds__(kBlock) soa_queue_kernel(Work w, unsigned long long *stats) {
  unsigned nq = 0, nlong = 0, nsplit = 0;
  for (int j = 0; j < kNBucket; ++j) {
    nq += w.ctl->hist[j];
    if (j >= w.qs.long_bucket) nlong += w.ctl->hist[j];
    if (j >= w.qs.short_bucket) nsplit += w.ctl->hist[j];
  }
  const unsigned qchunk = (unsigned)w.qs.chunk, qschunk = (unsigned)w.qs.short_chunk;
  const unsigned nwave = gridDim.x * (kBlock / 64);
  const unsigned wid = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  bool first = MAM4X_STATIC_FIRST != 0;
  const int lane = threadIdx.x & 63;
  const long n = w.ncell;
  unsigned cbeg = 0, cend = 0;  // the wave's current chunk (wave-uniform)
  bool lexhausted = false, exhausted = false;
  bool active = false;
  int cell = 0;
  int emin = 0;  // the cell's dev::soa_fast_exact record
  dev::SoaState<double> s;
  unsigned long long niter = 0;
  while (true) {
    const unsigned long long need = __ballot(!active);
    if (need && cbeg == cend && !exhausted) {
      // long part: the first chunk of every wave is assigned statically
      // (chunk wid), later ones through the counter; then the short part
      unsigned b0 = nsplit, e0 = nsplit;
      if (first) {
        first = false;
        b0 = wid * qchunk;
      } else if (!lexhausted) {
        if (lane == 0) b0 = atomicAdd(&w.ctl->head, qchunk);
        b0 = (MAM4X_STATIC_FIRST ? nwave * qchunk : 0u) +
             __builtin_amdgcn_readfirstlane(b0);
      }
      if (b0 < nsplit) {
        e0 = min(b0 + qchunk, nsplit);
      } else {
        lexhausted = true;
        if (lane == 0) b0 = atomicAdd(&w.ctl->shead, qschunk);
        b0 = nsplit + __builtin_amdgcn_readfirstlane(b0);
        e0 = min(b0 + qschunk, nq);
      }
      if (MAM4X_SOA_PRIO && b0 < nlong) __builtin_amdgcn_s_setprio(2);
      if (b0 >= nq) {
        exhausted = true;
      } else {
        cbeg = b0;
        cend = e0;
      }
    }
    if (need && cbeg < cend) {
      const unsigned rank = (unsigned)__popcll(need & ((1ull << lane) - 1ull));
      const unsigned avail = cend - cbeg;
      if (!active && rank < avail) {
        cell = w.queue[cbeg + rank];
        dev::SoaIn<double> in;
        ws_load_in(w.ws, n, cell, in);
        dev::soa_init(s, in);
        emin = dev::soa_fast_ok(s, w.dtfull) ? 0 : dev::kSoaExpFail;
        active = true;
      }
      cbeg += min((unsigned)__popcll(need), avail);
    }
    if (!__any(active)) break;
    if (active) {
      emin = min(emin, dev::soa_state_exp(s));
      if (!dev::soa_tick<true>(s, w.dtfull)) {
        // flags in the cell's bucket byte (buckets are < kIeee; the set-up
        // kernel rewrites it every step)
        if (emin < dev::kSoaExpMin) {
          // inputs left untouched for soa_ieee_kernel
          w.bucket[cell] = (unsigned char)(w.bucket[cell] | kIeee);
        } else {
          w.ws[W_G * n + cell] = s.g;
#pragma unroll
          for (int m = 0; m < 4; ++m) w.ws[(W_A + m) * n + cell] = s.a[m];
          w.ws[W_GAVG * n + cell] = dev::soa_gavg(s);
          niter += (unsigned long long)s.niter;
          // the loop stopped at its cap
          if (s.niter > 1000) w.bucket[cell] = (unsigned char)(w.bucket[cell] | kCapped);
        }
        active = false;
      }
    }
  }
  add_stats(stats, niter);
}

// Phase 3b: the cells soa_queue_kernel flagged kIeee (a mass below 2^-400,
// an input out of range; none in ordinary scenes) rerun with IEEE quotients
// from their inputs, results written as the queue kernel writes them.
__global__ void __launch_bounds__(kBlock) soa_ieee_kernel(Work w) {
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  unsigned long long niter = 0;
  if (c < w.ncell) {
    const unsigned char b = w.bucket[c];
    if (b != kNoSoa && (b & kIeee)) {
      const long n = w.ncell;
      dev::SoaIn<double> in;
      ws_load_in(w.ws, n, c, in);
      dev::SoaState<double> s;
      dev::soa_init(s, in);
      while (dev::soa_tick<false>(s, w.dtfull)) {
      }
      w.ws[W_G * n + c] = s.g;
#pragma unroll
      for (int m = 0; m < 4; ++m) w.ws[(W_A + m) * n + c] = s.a[m];
      w.ws[W_GAVG * n + c] = dev::soa_gavg(s);
      niter = (unsigned long long)s.niter;
      if (s.niter > 1000) w.bucket[c] = (unsigned char)(b | kCapped);
    }
  }
  add_stats(w.stats, niter);
}

// Phase 4.
template <bool PRE, class P, class Src>
__global__ void __launch_bounds__(kBlock) chain_post_kernel(Src src, Work w) {
  const long g = (long)blockIdx.x * kBlock + threadIdx.x;  // workspace index
  long c = g;
  const ChainArgs &a = src.scene(c);
  if (g >= w.ncell || c >= (long)a.ncol * a.nlev) return;
  int icol;
  long off;
  cell_index(a, c, icol, off);
  dev::CellCtx<P> X;
  dev::cell_setup<PRE, P>(a, off, icol, X);
  dev::Aer<typename P::R> S;
  dev::ClearGas<double> G;
  [[maybe_unused]] dev::ProcSink D;
  if constexpr (P::diag) {
    D = dev::proc_sink(a, off, X.fclea);
    D.zero();
  }
  if (X.cat != 1) {
    dev::clear_build<P>(a, off, X, S, G);
    double uptk[4] = {0, 0, 0, 0}, sa[4] = {0, 0, 0, 0}, sg = 0.0, sgavg = 0.0;
    if (a.cfg.do_cond) {
      const long n = w.ncell;
      sg = w.ws[W_G * n + g];
      sgavg = w.ws[W_GAVG * n + g];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        sa[m] = w.ws[(W_A + m) * n + g];
        uptk[m] = w.ws[(W_UPTK + m) * n + g];
      }
    }
    dev::clear_post<P>(*a.K, a.cfg, a.dt, X.T, X.p, X.z, X.pblh, X.rh_clear, X.dga, X.dgw, X.wd,
                    G.g1, G.g, S, uptk, sg, sgavg, sa, &D);
  }
  dev::cell_finish<PRE, P>(a, off, X, S, G);
  if (a.soacap)
    a.soacap[off] = (a.cfg.do_cond && X.cat != 1 && (w.bucket[g] & kCapped)) ? 1 : 0;
}

// AeroProcess::validate checks (gasaerexch.hpp:96-120): negative T, p, qv or
// any negative prognostic.  One count per offending cell.
// Reverse sweep (mam4x_compute_adjoint).
//
// A cell without a SOA loop (overcast, or condensation off) runs the whole
// sweep in one lane: chain_adjoint_kernel.  A cell with one runs it in the
// step's own phases, so that the SOA loops, whose length varies from one
// sub-step to 1000 between cells, keep the work queue's lane refill:
//   1. the step's set-up kernel and queue (chain_pre_kernel,
//      queue_scatter_kernel);
//   2. soa_log_kernel: the loop on the queue, each wave logging one row per
//      iteration of its refill loop -- per lane the state a sub-step starts
//      from (6 doubles) and a tag (sub-step of cell c / cell c ended /
//      idle) -- into chunks of kLogRows rows taken from a pool, chained
//      backwards; the loops' final states go to the R_* planes;
//   3. adj_post_kernel: the reverse of cell_finish and clear_post, one lane
//      per cell; its adjoints go to the PostBar planes;
//   4. soa_replay_kernel: wave w replays wave w's log backwards, row by row,
//      so every lane meets its cells' sub-steps in reverse order and the
//      rows are read as they were written, whole;  the adjoints of each
//      loop's inputs go to the I_* planes;
//   5. adj_pre_kernel: the reverse of clear_pre, clear_build and cell_setup,
//      writing the gradients.
struct AdjArgs {
  double *sbar[NSTATE];
  double *Tbar, *pbar, *qvbar;
  const unsigned char *bucket;  // chain_adjoint_kernel: kNoSoa cells only; NULL: all
};

// the SOA loop's final state
enum { R_G = 0, R_A = 1, R_GAVG = 5, R_AVG = 6, R_DTSUM = 7, R_NPLANE = 8 };
// the adjoints of the SOA loop's inputs, from its reverse
enum { I_G = 0, I_G0 = 1, I_A = 2, I_UPTK = 6, I_AOPOA = 10, I_NPLANE = 14 };
// PostBar: the partial gradient (52 + 3), uptk_b (4), soa_g_b, soa_gavg_b,
// soa_a_b (4)
constexpr int kPbPlanes = NSTATE + 3 + 4 + 6;
constexpr int kPbUptk = NSTATE + 3;    // plane of uptk_b[0]
constexpr int kPbSoa = kPbPlanes - 6;  // plane of soa_g_b; soa_gavg_b, soa_a_b follow
constexpr int kAdjPlanes = R_NPLANE + kPbPlanes + I_NPLANE;

// The log: chunks of kLogRows rows; a row holds, for the 64 lanes of a wave,
// kLogVals planes of doubles (lane-contiguous) and a tag each.
constexpr int kLogRows = 32, kLogVals = 6;
constexpr int kTagIdle = (int)0x80000000;  // tag: lane idle; c >= 0: a sub-step
                                           // of cell c; ~c: cell c ended
struct AdjWork {
  QueueShape qs;
  const double *ws;             // the SOA loop's inputs (ws_load_in)
  double *res;                  // R_NPLANE planes
  double *pb;                   // kPbPlanes planes
  double *inb;                  // I_NPLANE planes
  double *rec;                  // log values [cap][kLogRows][kLogVals][64]
  int *tag;                     // log tags [cap][kLogRows][64]
  int *prev;                    // [cap]: the wave's chunk before this one (-1: none)
  int2 *wend;                   // [waves]: (last chunk or -1, rows used in it)
  unsigned *next;               // chunks taken from the pool
  unsigned cap;                 // chunks in the pool
  unsigned nwave;               // waves of the log and replay kernels
  const unsigned char *bucket;  // kNoSoa: no SOA loop
  const int *queue;             // cells with a SOA loop, longest first
  QueueCtl *ctl;
  long ncell;
  double dtfull;
};
static_assert(sizeof(ChainArgs) + sizeof(AdjWork) + sizeof(AdjArgs) <= 4000,
              "the adjoint kernels' arguments fit the kernel-argument segment");

// PostBar -> its planes
__device__ __forceinline__ void pb_store(double *pl, long n, long c,
                                         const dev::rev::PostBar &pb) {
  double *q = pl + c;
#pragma unroll
  for (int f = 0; f < NSTATE; ++f) q[(long)f * n] = pb.g[f];
  q[(long)NSTATE * n] = pb.T_b;
  q[(long)(NSTATE + 1) * n] = pb.p_b;
  q[(long)(NSTATE + 2) * n] = pb.qv_b;
#pragma unroll
  for (int m = 0; m < 4; ++m) q[(long)(kPbUptk + m) * n] = pb.uptk_b[m];
  q[(long)kPbSoa * n] = pb.soa_g_b;
  q[(long)(kPbSoa + 1) * n] = pb.soa_gavg_b;
#pragma unroll
  for (int m = 0; m < 4; ++m) q[(long)(kPbSoa + 2 + m) * n] = pb.soa_a_b[m];
}

// Phase 2: soa_queue_kernel's loop (same refill, same sub-steps) with the
// log written at the end of every iteration.  A pool that runs out stops the
// wave's logging, but the loop still drains the queue and still counts the
// chunks it would have taken: the caller sees *next > cap, the pool this
// run needed, and runs the phase again with a pool that large.
__global__ void __launch_bounds__(kBlock) soa_log_kernel(AdjWork g) {
  unsigned nq = 0, nlong = 0, nsplit = 0;
  for (int j = 0; j < kNBucket; ++j) {
    nq += g.ctl->hist[j];
    if (j >= g.qs.long_bucket) nlong += g.ctl->hist[j];
    if (j >= g.qs.short_bucket) nsplit += g.ctl->hist[j];
  }
  const unsigned qchunk = (unsigned)g.qs.chunk, qschunk = (unsigned)g.qs.short_chunk;
  const unsigned nwave = gridDim.x * (kBlock / 64);
  const unsigned wid = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  bool first = MAM4X_STATIC_FIRST != 0;
  const int lane = threadIdx.x & 63;
  const long n = g.ncell;
  unsigned cbeg = 0, cend = 0;  // the wave's current chunk of the queue (wave-uniform)
  bool lexhausted = false, exhausted = false;
  bool active = false;
  int cell = 0;
  dev::SoaState<double> s;
  int chunk = -1, row = kLogRows;  // the wave's log position (wave-uniform)
  bool logging = true;
  while (true) {
    const unsigned long long need = __ballot(!active);
    if (need && cbeg == cend && !exhausted) {
      unsigned b0 = nsplit, e0 = nsplit;
      if (first) {
        first = false;
        b0 = wid * qchunk;
      } else if (!lexhausted) {
        if (lane == 0) b0 = atomicAdd(&g.ctl->head, qchunk);
        b0 = (MAM4X_STATIC_FIRST ? nwave * qchunk : 0u) +
             __builtin_amdgcn_readfirstlane(b0);
      }
      if (b0 < nsplit) {
        e0 = min(b0 + qchunk, nsplit);
      } else {
        lexhausted = true;
        if (lane == 0) b0 = atomicAdd(&g.ctl->shead, qschunk);
        b0 = nsplit + __builtin_amdgcn_readfirstlane(b0);
        e0 = min(b0 + qschunk, nq);
      }
      if (MAM4X_SOA_PRIO && b0 < nlong) __builtin_amdgcn_s_setprio(2);
      if (b0 >= nq) {
        exhausted = true;
      } else {
        cbeg = b0;
        cend = e0;
      }
    }
    if (need && cbeg < cend) {
      const unsigned rank = (unsigned)__popcll(need & ((1ull << lane) - 1ull));
      const unsigned avail = cend - cbeg;
      if (!active && rank < avail) {
        cell = g.queue[cbeg + rank];
        dev::SoaIn<double> in;
        ws_load_in(g.ws, n, cell, in);
        dev::soa_init(s, in);
        active = true;
      }
      cbeg += min((unsigned)__popcll(need), avail);
    }
    if (!__any(active)) break;
    double r[kLogVals];
    r[0] = s.g;
#pragma unroll
    for (int m = 0; m < 4; ++m) r[1 + m] = s.a[m];
    r[5] = s.tcur;
    int tag = kTagIdle;
    if (active) {
      if (dev::soa_tick<true>(s, g.dtfull)) {
        tag = cell;
      } else {
        tag = ~cell;
        g.res[R_G * n + cell] = s.g;
#pragma unroll
        for (int m = 0; m < 4; ++m) g.res[(R_A + m) * n + cell] = s.a[m];
        g.res[R_GAVG * n + cell] = dev::soa_gavg(s);
        g.res[R_AVG * n + cell] = s.avg;
        g.res[R_DTSUM * n + cell] = s.dtsum;
        active = false;
      }
    }
    if (row == kLogRows) {  // a new chunk, linked back to the last one
      unsigned id = 0;
      if (lane == 0) id = atomicAdd(g.next, 1u);
      id = __builtin_amdgcn_readfirstlane(id);
      row = 0;
      if (id < g.cap && logging) {
        if (lane == 0) g.prev[id] = chunk;
        chunk = (int)id;
      } else {
        logging = false;  // the pool ran out: from here on only count chunks
      }
    }
    if (logging) {
      const long base = (long)chunk * kLogRows + row;
      g.tag[base * 64 + lane] = tag;
      if (tag >= 0) {
#pragma unroll
        for (int j = 0; j < kLogVals; ++j) g.rec[(base * kLogVals + j) * 64 + lane] = r[j];
      }
    }
    ++row;
  }
  if (lane == 0 && wid < g.nwave) g.wend[wid] = make_int2(chunk, chunk >= 0 ? row : 0);
}

// Phase 4.  One lane's reverse of the loops it ran in phase 2: seeds from the
// PostBar planes and the final state, then sub-steps in reverse, then the
// adjoints of the loop's inputs stored to the I_* planes.
struct SoaRev {
  const AdjWork &g;
  int cell;
  dev::SoaState<double> k;
  dev::rev::SoaBar b;
  __device__ explicit SoaRev(const AdjWork &g_) : g(g_), cell(-1) {}
  __device__ __forceinline__ void begin(int c) {
    cell = c;
    const long n = g.ncell;
    dev::SoaIn<double> in;
    ws_load_in(g.ws, n, c, in);
    dev::soa_init(k, in);
    const double *pl = g.pb + (long)kPbSoa * n + c;
    double a_b[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) a_b[m] = pl[(2 + m) * n];
    dev::rev::soa_rev_seed(g.res[R_AVG * n + c], g.res[R_DTSUM * n + c], pl[0], a_b, pl[n], b);
  }
  __device__ __forceinline__ void finish() {
    const long n = g.ncell;
    dev::SoaIn<double> in_b = {};
    dev::rev::soa_rev_init(k, b, in_b);
    double *pl = g.inb + cell;
    pl[I_G * n] = in_b.g;
    pl[I_G0 * n] = in_b.g0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      pl[(I_A + m) * n] = in_b.a[m];
      pl[(I_UPTK + m) * n] = in_b.uptk[m];
      pl[(I_AOPOA + m) * n] = in_b.aopoa[m];
    }
    cell = -1;
  }
};

// one lane's entry of a log row
__device__ __forceinline__ void log_fetch(const AdjWork &g, int chunk, int r, int lane, int &tag,
                                          double v[kLogVals]) {
  const long base = (long)chunk * kLogRows + r;
  tag = g.tag[base * 64 + lane];
#pragma unroll
  for (int j = 0; j < kLogVals; ++j) v[j] = g.rec[(base * kLogVals + j) * 64 + lane];
}

// The wave's rows newest first; the next row's loads are issued before the
// current row's sub-step is reversed, so they are in flight meanwhile.
__global__ void __launch_bounds__(kBlock, MAM4X_REPLAY_WAVES) soa_replay_kernel(AdjWork g) {
  const unsigned wid = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (wid >= g.nwave) return;
  const int2 e = g.wend[wid];
  int chunk = e.x, r = e.y - 1;
  if (chunk < 0 || r < 0) return;
  SoaRev t(g);
  int tag;
  double v[kLogVals];
  log_fetch(g, chunk, r, lane, tag, v);
  while (true) {
    int nchunk = chunk, nr = r - 1;
    if (nr < 0) {
      nchunk = g.prev[chunk];
      nr = kLogRows - 1;
    }
    int ntag = kTagIdle;
    double nv[kLogVals];
    if (nchunk >= 0) log_fetch(g, nchunk, nr, lane, ntag, nv);
    if (tag >= 0) {
      dev::rev::SoaCkpt c;
      c.g = v[0];
#pragma unroll
      for (int m = 0; m < 4; ++m) c.a[m] = v[1 + m];
      c.tcur = v[5];
      dev::rev::soa_tick_rev(t.k, c, g.dtfull, t.b);
    } else {
      if (t.cell >= 0) t.finish();
      if (tag != kTagIdle) t.begin(~tag);
    }
    if (nchunk < 0) break;
    chunk = nchunk;
    r = nr;
    tag = ntag;
#pragma unroll
    for (int j = 0; j < kLogVals; ++j) v[j] = nv[j];
  }
  if (t.cell >= 0) t.finish();
}

// the queue's refill counters back to the start, for another pass over it
__global__ void queue_rewind_kernel(QueueCtl *q, unsigned *next) {
  if (threadIdx.x == 0) {
    q->head = 0;
    q->shead = 0;
    *next = 0;
  }
}

__device__ __forceinline__ void load_tbar(const ChainArgs &a, long off, double tb[NSTATE]) {
#pragma unroll
  for (int f = 0; f < NSTATE; ++f) tb[f] = a.tbar[f] ? a.tbar[f][off] : 0.0;
}

__device__ __forceinline__ void store_grad(const AdjArgs &o, long off, const double sb[NSTATE],
                                           double Tb, double pb, double qb) {
#pragma unroll
  for (int f = 0; f < NSTATE; ++f)
    if (o.sbar[f]) o.sbar[f][off] = sb[f];
  if (o.Tbar) o.Tbar[off] = Tb;
  if (o.pbar) o.pbar[off] = pb;
  if (o.qvbar) o.qvbar[off] = qb;
}

// Phase 3 (cells with a SOA loop).
template <bool PRE>
__global__ void __launch_bounds__(kBlock) adj_post_kernel(ChainArgs a, AdjWork g) {
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  const long n = g.ncell;
  if (c >= n || g.bucket[c] == kNoSoa) return;
  int icol;
  long off;
  cell_index(a, c, icol, off);
  double tb[NSTATE];
  load_tbar(a, off, tb);
  dev::CellCtx<Fp64> X;
  dev::cell_setup<PRE, Fp64>(a, off, icol, X);
  dev::Aer<double> S0;
  dev::ClearGas<double> G0;
  dev::clear_build<Fp64>(a, off, X, S0, G0);
  dev::SoaIn<double> in;
  ws_load_in(g.ws, n, c, in);
  double sa[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) sa[m] = g.res[(R_A + m) * n + c];
  dev::rev::PostBar pb;
  dev::rev::adjoint_post_phase<PRE>(a, off, X, S0, G0, in, g.res[R_G * n + c],
                                    g.res[R_GAVG * n + c], sa, tb, pb);
  pb_store(g.pb, n, c, pb);
}

// Phase 5 (cells with a SOA loop).
template <bool PRE>
__global__ void __launch_bounds__(kBlock, MAM4X_ADJ_PRE_WAVES) adj_pre_kernel(ChainArgs a, AdjWork g, AdjArgs o) {
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  const long n = g.ncell;
  if (c >= n || g.bucket[c] == kNoSoa) return;
  int icol;
  long off;
  cell_index(a, c, icol, off);
  dev::CellCtx<Fp64> X;
  dev::cell_setup<PRE, Fp64>(a, off, icol, X);
  dev::Aer<double> S0;
  dev::ClearGas<double> G0;
  dev::clear_build<Fp64>(a, off, X, S0, G0);
  const double *pl = g.inb + c;
  const double *pp = g.pb + c;
  dev::SoaIn<double> in_b;
  in_b.g = pl[I_G * n];
  in_b.g0 = pl[I_G0 * n];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    in_b.a[m] = pl[(I_A + m) * n];
    in_b.uptk[m] = pp[(long)(kPbUptk + m) * n] + pl[(I_UPTK + m) * n];
    in_b.aopoa[m] = pl[(I_AOPOA + m) * n];
  }
  double sb[NSTATE], Tb, pB, qb;
  dev::rev::adjoint_pre_phase<PRE>(a, off, X, S0, in_b, sb, Tb, pB, qb);
  // plus the post phase's partial gradient
#pragma unroll
  for (int f = 0; f < NSTATE; ++f) sb[f] = pp[(long)f * n] + sb[f];
  Tb = pp[(long)NSTATE * n] + Tb;
  pB = pp[(long)(NSTATE + 1) * n] + pB;
  qb = pp[(long)(NSTATE + 2) * n] + qb;
  store_grad(o, off, sb, Tb, pB, qb);
}

// A SOA loop tape for cells that run none (never called).
struct NoSoaTape {
  __device__ int record(dev::SoaState<double> &, double) { return 0; }
  __device__ dev::rev::SoaCkpt get(int) { return dev::rev::SoaCkpt{}; }
};

// Cells without a SOA loop (one sub-step): the whole sweep in one lane.
template <bool PRE>
__global__ void __launch_bounds__(kBlock) chain_adjoint_kernel(ChainArgs a, AdjArgs o) {
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  const long ncell = (long)a.ncol * a.nlev;
  if (c >= ncell) return;
  if (o.bucket && o.bucket[c] != kNoSoa) return;
  int icol;
  long off;
  cell_index(a, c, icol, off);
  double tb[NSTATE], sb[NSTATE];
  load_tbar(a, off, tb);
#pragma unroll
  for (int f = 0; f < NSTATE; ++f) sb[f] = 0.0;
  double Tb = 0.0, pb = 0.0, qb = 0.0;
  NoSoaTape soa;
  dev::rev::OneSubTape sub;
  dev::rev::adjoint_cell<PRE>(a, off, icol, tb, soa, sub, sb, Tb, pb, qb);
  store_grad(o, off, sb, Tb, pb, qb);
}

// The subarea's aerosol and gases at the start of each sub-step, in HBM:
// kSubVals planes per sub-step, one double per cell.
constexpr int kSubVals = 4 + 28 + 2;
struct PlaneSubTape {
  double *base;
  long n, c;
  __device__ void save(int j, const dev::Aer<double> &S, const double g[2]) {
    double *q = base + (long)j * kSubVals * n + c;
#pragma unroll
    for (int m = 0; m < 4; ++m) q[m * n] = S.n[m];
#pragma unroll
    for (int sp = 0; sp < 7; ++sp)
#pragma unroll
      for (int m = 0; m < 4; ++m) q[(4 + 4 * sp + m) * n] = S.q[sp][m];
    q[32 * n] = g[0];
    q[33 * n] = g[1];
  }
  __device__ void load(int j, dev::Aer<double> &S, double g[2]) const {
    const double *q = base + (long)j * kSubVals * n + c;
#pragma unroll
    for (int m = 0; m < 4; ++m) S.n[m] = q[m * n];
#pragma unroll
    for (int sp = 0; sp < 7; ++sp)
#pragma unroll
      for (int m = 0; m < 4; ++m) S.q[sp][m] = q[(4 + 4 * sp + m) * n];
    g[0] = q[32 * n];
    g[1] = q[33 * n];
  }
};

// ntsubstep > 1: every cell's whole sweep in one lane, a checkpoint per
// sub-step in HBM (subws: ntsubstep x kSubVals planes) and each sub-step's
// SOA loop recomputed inside it from fixed per-lane storage (SoaTape2).
// With a queue (condensation on), pass 0 takes the queued cells longest
// first, so a wave's lanes run SOA loops of similar length, and pass 1 the
// rest; without one, pass 1 takes every cell.
template <bool PRE>
__global__ void __launch_bounds__(kBlock) chain_adjoint_sub_kernel(ChainArgs a, AdjArgs o,
                                                                   double *subws,
                                                                   const int *queue,
                                                                   const QueueCtl *ctl,
                                                                   int pass) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  const long ncell = (long)a.ncol * a.nlev;
  long c;
  if (pass == 0) {
    unsigned nq = 0;
    for (int j = 0; j < kNBucket; ++j) nq += ctl->hist[j];
    if (i >= (long)nq) return;
    c = queue[i];
  } else {
    if (i >= ncell) return;
    if (o.bucket && o.bucket[i] != kNoSoa) return;
    c = i;
  }
  int icol;
  long off;
  cell_index(a, c, icol, off);
  double tb[NSTATE], sb[NSTATE];
  load_tbar(a, off, tb);
#pragma unroll
  for (int f = 0; f < NSTATE; ++f) sb[f] = 0.0;
  double Tb = 0.0, pb = 0.0, qb = 0.0;
  dev::rev::SoaTape2<32> soa;
  PlaneSubTape sub{subws, ncell, c};
  dev::rev::adjoint_cell<PRE>(a, off, icol, tb, soa, sub, sb, Tb, pb, qb);
  store_grad(o, off, sb, Tb, pb, qb);
}

__global__ void __launch_bounds__(kBlock) validate_kernel(ChainArgs a,
                                                          unsigned long long *count) {
  const long ncell = (long)a.ncol * a.nlev;
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  unsigned long long bad = 0;
  if (c < ncell) {
    int icol;
    long off;
    cell_index(a, c, icol, off);
    bool v = (a.T[off] < 0) || (a.p[off] < 0) || (a.qv[off] < 0);
    for (int f = 0; f < NSTATE; ++f) v = v || (a.state[f][off] < 0);
    bad = v ? 1ull : 0ull;
  }
  for (int d = 32; d > 0; d >>= 1) bad += __shfl_xor(bad, d, 64);
  if ((threadIdx.x & 63) == 0 && bad) atomicAdd(count, bad);
}

// Column budgets (mam4x_outputs.column_budgets).  Per column and budget,
// the levels k = l, l + 8, l + 16, ... (l = 0..7) form partial sum l by
// fused multiply-adds in ascending k, each term weighted by dp(k) * (1/g);
// the partials are combined as ((p0+p4)+(p2+p6))+((p1+p5)+(p3+p7)).  Both
// layouts form exactly these operations, so they agree bit for bit:
//  - LEV_COL (columns contiguous): one lane per column holds the eight
//    partials; every load is coalesced over 64 columns;
//  - COL_LEV (levels contiguous): eight lanes per column, lane l walking its
//    levels (each load covers eight contiguous levels of eight columns), the
//    partials combined by a butterfly over the eight lanes (xor 4, 2, 1),
//    which adds the same pairs in the same tree.
// One wave per (column group, budget), four budgets of a group per block;
// the linear block index is mapped so that every budget of one column group
// lands on the same XCD (blocks go round-robin over the eight XCDs), where
// the group's dp lines are shared through that XCD's L2.
constexpr int kNBudget = MAM4X_NBUDGET;
constexpr int kBudgetBlock = 64;  // lanes per budget (one wave)
constexpr int kBudgetWaves = 4;   // budgets of one column group per block
constexpr int kBudgetQuads = (kNBudget + kBudgetWaves - 1) / kBudgetWaves;
constexpr double kInvGravity = 1.0 / 9.80616;  // mam4_constants.hpp:48
constexpr int kClosureN = 7;  // six SO4 (SOA) fields and the gas

struct BudgetArgs {
  const double *tend[NSTATE];
  const double *dp;
  double *out;
  int ncol, nlev, ldc, ngroup;
  int collev;  // MAM4X_LAYOUT_COL_LEV
  // closures (sulfur, organic): their fields in state order and the mol/kg
  // factor of each
  int clo_f[2][kClosureN];
  double clo_c[2][kClosureN];
};

// the per-level value x(k) of budget b: a tendency, or a closure's sum over
// its fields in state order
struct FieldTerm {
  const double *t;
  __device__ double operator()(long off) const { return t[off]; }
};
struct ClosureTerm {
  const double *t[kClosureN];
  double c[kClosureN];
  __device__ double operator()(long off) const {
    double v[kClosureN];
#pragma unroll
    for (int i = 0; i < kClosureN; ++i) v[i] = t[i][off];
    double x = 0.0;
#pragma unroll
    for (int i = 0; i < kClosureN; ++i) x = fma(v[i], c[i], x);
    return x;
  }
};

// LEV_COL: one lane per column, eight partials; whole chunks of sixteen
// (then eight) levels issue their loads before the multiply-adds, the last
// partial chunk is guarded
template <class Term>
__device__ double budget_lane(const BudgetArgs &a, const Term &term, int col) {
  double p[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int k0 = 0;
  for (; k0 + 16 <= a.nlev; k0 += 16) {
    double x[16], w[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) {
      const long off = (long)(k0 + l) * a.ldc + col;
      x[l] = term(off);
      w[l] = a.dp[off];
    }
#pragma unroll
    for (int l = 0; l < 16; ++l) p[l & 7] = fma(x[l], w[l] * kInvGravity, p[l & 7]);
  }
  if (k0 + 8 <= a.nlev) {
    double x[8], w[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      const long off = (long)(k0 + l) * a.ldc + col;
      x[l] = term(off);
      w[l] = a.dp[off];
    }
#pragma unroll
    for (int l = 0; l < 8; ++l) p[l] = fma(x[l], w[l] * kInvGravity, p[l]);
    k0 += 8;
  }
#pragma unroll
  for (int l = 0; l < 8; ++l)
    if (k0 + l < a.nlev) {
      const long off = (long)(k0 + l) * a.ldc + col;
      p[l] = fma(term(off), a.dp[off] * kInvGravity, p[l]);
    }
  return ((p[0] + p[4]) + (p[2] + p[6])) + ((p[1] + p[5]) + (p[3] + p[7]));
}

// COL_LEV: lane l of a column's eight walks k = l, l + 8, ...; eight levels
// per pass load together
template <class Term>
__device__ double budget_octet(const BudgetArgs &a, const Term &term, int col, int l) {
  double p = 0.0;
  const long base = (long)col * a.ldc;
  int k = l;
  for (; k + 56 < a.nlev; k += 64) {
    double x[8], w[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      x[j] = term(base + k + 8 * j);
      w[j] = a.dp[base + k + 8 * j];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) p = fma(x[j], w[j] * kInvGravity, p);
  }
  for (; k < a.nlev; k += 8) p = fma(term(base + k), a.dp[base + k] * kInvGravity, p);
  return p;
}

template <bool COLLEV, class Term>
__device__ void budget_column(const BudgetArgs &a, const Term &term, int grp, int b) {
  const int lane = threadIdx.x & (kBudgetBlock - 1);
  if (!COLLEV) {
    const int col = grp * kBudgetBlock + lane;
    if (col < a.ncol) a.out[(long)b * a.ncol + col] = budget_lane(a, term, col);
  } else {
    const int col = grp * (kBudgetBlock / 8) + (lane >> 3), l = lane & 7;
    const bool live = col < a.ncol;
    double p = live ? budget_octet(a, term, col, l) : 0.0;
    // every lane takes part in the butterfly
    p = p + __shfl_xor(p, 4, 64);
    p = p + __shfl_xor(p, 2, 64);
    p = p + __shfl_xor(p, 1, 64);
    if (live && l == 0) a.out[(long)b * a.ncol + col] = p;
  }
}

template <bool COLLEV>
__global__ void __launch_bounds__(kBudgetWaves * kBudgetBlock) column_budget_kernel(BudgetArgs a) {
  const int xcd = blockIdx.x & 7, i = blockIdx.x >> 3;
  const int grp = xcd + 8 * (i / kBudgetQuads);
  const int b = (i % kBudgetQuads) * kBudgetWaves + (int)(threadIdx.x / kBudgetBlock);
  if (grp >= a.ngroup || b >= kNBudget) return;  // whole waves
  if (b < NSTATE) {
    budget_column<COLLEV>(a, FieldTerm{a.tend[b]}, grp, b);
  } else {
    const int j = b - NSTATE;
    ClosureTerm t;
#pragma unroll
    for (int q = 0; q < kClosureN; ++q) {
      t.t[q] = a.tend[a.clo_f[j][q]];
      t.c[q] = a.clo_c[j][q];
    }
    budget_column<COLLEV>(a, t, grp, b);
  }
}

// Aerosol water uptake (water_uptake.hpp): one lane per cell.
struct WaterArgs {
  const double *qi[21];  // interstitial mass planes, state order
  const double *T, *p, *qv, *cld, *dga[4];
  double *dgw[4], *qw[4], *wd[4];
  const ChainConsts *K;
  int ncol, nlev, ldc, layout;
};

__global__ void __launch_bounds__(kBlock) water_uptake_kernel(WaterArgs a) {
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  if (c >= (long)a.ncol * a.nlev) return;
  long off;
  if (a.layout == MAM4X_LAYOUT_COL_LEV) {
    const int icol = (int)(c / a.nlev);
    off = (long)icol * a.ldc + (c - (long)icol * a.nlev);
  } else {
    const int k = (int)(c / a.ncol);
    off = (long)k * a.ldc + (c - (long)k * a.ncol);
  }
  double dga[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) dga[m] = a.dga[m][off];
  wu::WaterOut o;
  wu::water_uptake_cell([&](int m, int s) { return a.qi[MODE_OFF[m] + s][off]; }, NSPEC_MODE,
                        a.K->vol_from_dgn, a.T[off], a.p[off], a.qv[off], a.cld[off], dga, o);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    a.dgw[m][off] = o.dgncur_awet[m];
    a.qw[m][off] = o.qaerwat[m];
    if (a.wd[m]) a.wd[m][off] = o.wetdens[m];
  }
}

// The wet-size producer (compute_calcsize_and_water_uptake_dr,
// modal_aero_opt.hpp:510-533): the E3SM-layout calcsize on the state, then
// water uptake on the dry diameters it returns.  One lane per cell.
struct WetSizeArgs {
  const double *state[NSTATE];
  const double *T, *p, *qv, *cld;
  double *dgi[4], *dgc[4], *dgw[4], *qw[4], *wd[4];
  double *tend[NSTATE];  // calcsize tendencies (NULL: not written)
  const ChainConsts *K;
  double dt;
  int ncol, nlev, ldc, layout;
};

__global__ void __launch_bounds__(kBlock) wet_size_kernel(WetSizeArgs a) {
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  if (c >= (long)a.ncol * a.nlev) return;
  long off;
  if (a.layout == MAM4X_LAYOUT_COL_LEV) {
    const int icol = (int)(c / a.nlev);
    off = (long)icol * a.ldc + (c - (long)icol * a.nlev);
  } else {
    const int k = (int)(c / a.ncol);
    off = (long)k * a.ldc + (c - (long)k * a.ncol);
  }
  const ChainConsts &K = *a.K;
  double n_i[4], n_c[4], q_i[4][7], q_c[4][7];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    n_i[m] = a.state[f_ni(m)][off];
    n_c[m] = a.state[f_nc(m)][off];
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      q_i[m][s] = s < NSPEC_MODE[m] ? a.state[f_qi(m, s)][off] : 0.0;
      q_c[m][s] = s < NSPEC_MODE[m] ? a.state[f_qc(m, s)][off] : 0.0;
    }
  }
  dev::CalcOut<double> o;
  dev::calcsize(K, true, true, a.dt, n_i, n_c, q_i, q_c, o, true);
  wu::WaterOut w;
  wu::water_uptake_cell([&](int m, int s) { return q_i[m][s]; }, NSPEC_MODE, K.vol_from_dgn,
                        a.T[off], a.p[off], a.qv[off], a.cld[off], o.dgn_i, w);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    a.dgi[m][off] = o.dgn_i[m];
    if (a.dgc[m]) a.dgc[m][off] = o.dgn_c[m];
    a.dgw[m][off] = w.dgncur_awet[m];
    a.qw[m][off] = w.qaerwat[m];
    if (a.wd[m]) a.wd[m][off] = w.wetdens[m];
  }
  if (a.tend[0]) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      a.tend[f_ni(m)][off] = o.dn_i[m];
      a.tend[f_nc(m)][off] = o.dn_c[m];
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        a.tend[f_qi(m, s)][off] = o.dq_i[m][s];
        a.tend[f_qc(m, s)][off] = o.dq_c[m][s];
      }
    }
    a.tend[F_H2SO4][off] = 0.0;
    a.tend[F_SOAG][off] = 0.0;
  }
}

// ------------------------------------------------------------------ tracers
// E3SM tracer arrays <-> state planes (utils.hpp:183-426).  A block moves a
// tile of kTile cells that are consecutive in the planes' memory order.  The
// cells' tracer rows (tracer index fastest) pass through LDS, so the rows are
// read or written as contiguous runs and every plane as 64 consecutive
// doubles per wave.  Only tracers 11-39 of state_q (H2SO4 to the last
// number; 12 and 13 ride along) and 15-39 of qqcw are moved.
constexpr int kTile = 64;
constexpr int kQ0 = 11, kNQ = MAM4X_PCNST - kQ0;  // state_q 11..39
constexpr int kC0 = 15, kNC = MAM4X_PCNST - kC0;  // qqcw 15..39
constexpr int kRow = kNQ + kNC + 1;               // LDS row, odd: no bank conflicts

struct TracerArgs {
  double *q, *qqcw;
  long col_stride, lev_stride;
  double *field[NSTATE];
  int ncol, nlev, ldc, layout;
  int mode;
};

// tracer index (state_q, or qqcw for cloud-borne fields) of state field f:
// gas_pcnst slot + 9 (utils.hpp:36-48)
__host__ __device__ constexpr int tracer_of(int f) {
  return f < F_Q_I ? NUMPTR[f & 3] + 9
       : f >= F_H2SO4 ? (f == F_H2SO4 ? 11 : 14)
       : mass_slot(field_mode(f), field_spec(f)) + 9;
}
__host__ __device__ constexpr bool cloud_borne(int f) {
  return (f >= F_NUM_C && f < F_Q_I) || (f >= F_Q_C && f < F_H2SO4);
}
// column of field f in an LDS row
__host__ __device__ constexpr int lds_col(int f) {
  return cloud_borne(f) ? kNQ + tracer_of(f) - kC0 : tracer_of(f) - kQ0;
}

static_assert(tracer_of(0) == 22 && tracer_of(1) == 27 && tracer_of(2) == 35 &&
                  tracer_of(3) == 39 && tracer_of(F_NUM_C + 3) == 39,
              "number tracers follow each mode's species");
static_assert(tracer_of(F_Q_I) == 15 && tracer_of(F_Q_I + 7) == 23 &&
                  tracer_of(F_Q_C + 11) == 28 && tracer_of(F_Q_C + 20) == 38,
              "species tracers in mode_aero_species order from 15");

// cell c of the planes' memory order: tracer-row offset and plane offset
__device__ __forceinline__ void tracer_cell(const TracerArgs &a, long c, long &row, long &off) {
  const int fast = a.layout == MAM4X_LAYOUT_COL_LEV ? a.nlev : a.ncol;
  const long slow = c / fast;
  const int i = (int)(c - slow * fast);
  off = slow * a.ldc + i;
  row = a.layout == MAM4X_LAYOUT_COL_LEV ? slow * a.col_stride + (long)i * a.lev_stride
                                         : (long)i * a.col_stride + slow * a.lev_stride;
}

// Each thread issues all its loads of a phase before it consumes any, so a
// block keeps ~15 loads per lane in flight instead of one.
constexpr int kItQ = (kTile * kNQ + kBlock - 1) / kBlock;   // 8
constexpr int kItC = (kTile * kNC + kBlock - 1) / kBlock;   // 7
constexpr int kItF = NSTATE / (kBlock / 64);                // 13 fields per wave
static_assert(kItF * (kBlock / 64) == NSTATE, "every wave takes the same number of fields");

__global__ void __launch_bounds__(kBlock) tracers_to_fields_kernel(TracerArgs a) {
  __shared__ double tile[kTile * kRow];
  __shared__ long roff[kTile], poff[kTile];
  const long ncell = (long)a.ncol * a.nlev;
  const long c0 = (long)blockIdx.x * kTile;
  const int n = (int)min((long)kTile, ncell - c0);
  if (threadIdx.x < n) tracer_cell(a, c0 + threadIdx.x, roff[threadIdx.x], poff[threadIdx.x]);
  __syncthreads();
  // rows in
  double vq[kItQ], vc[kItC];
#pragma unroll
  for (int j = 0; j < kItQ; ++j) {
    const int e = threadIdx.x + j * kBlock, i = e / kNQ, t = e - i * kNQ;
    if (e < n * kNQ) vq[j] = a.q[roff[i] + kQ0 + t];
  }
  if (a.qqcw) {
#pragma unroll
    for (int j = 0; j < kItC; ++j) {
      const int e = threadIdx.x + j * kBlock, i = e / kNC, t = e - i * kNC;
      if (e < n * kNC) vc[j] = a.qqcw[roff[i] + kC0 + t];
    }
  }
#pragma unroll
  for (int j = 0; j < kItQ; ++j) {
    const int e = threadIdx.x + j * kBlock, i = e / kNQ, t = e - i * kNQ;
    if (e < n * kNQ) tile[i * kRow + t] = vq[j];
  }
  if (a.qqcw) {
#pragma unroll
    for (int j = 0; j < kItC; ++j) {
      const int e = threadIdx.x + j * kBlock, i = e / kNC, t = e - i * kNC;
      if (e < n * kNC) tile[i * kRow + kNQ + t] = vc[j];
    }
  }
  __syncthreads();
  // planes out: one field per wave and pass
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
  for (int j = 0; j < kItF; ++j) {
    const int f = wave + j * (kBlock / 64);
    double *p = a.field[f];
    if (!p || (cloud_borne(f) && !a.qqcw) || lane >= n) continue;
    p[poff[lane]] = tile[lane * kRow + lds_col(f)];
  }
}

__global__ void __launch_bounds__(kBlock) fields_to_tracers_kernel(TracerArgs a) {
  __shared__ double tile[kTile * kRow];
  __shared__ long roff[kTile], poff[kTile];
  __shared__ unsigned char have[kRow];  // the row column has a source
  const long ncell = (long)a.ncol * a.nlev;
  const long c0 = (long)blockIdx.x * kTile;
  const int n = (int)min((long)kTile, ncell - c0);
  const bool tend = a.mode == MAM4X_PACK_TENDENCY;
  if (threadIdx.x < n) tracer_cell(a, c0 + threadIdx.x, roff[threadIdx.x], poff[threadIdx.x]);
  if (threadIdx.x < kRow) {
    // SO2 and DMS: zero tendency in tendency mode, else left alone
    const int t = threadIdx.x;
    have[t] = tend && (t == 12 - kQ0 || t == 13 - kQ0);
  }
  __syncthreads();
  if (threadIdx.x < NSTATE) {
    const int f = threadIdx.x;
    if (a.field[f] && !(cloud_borne(f) && !a.qqcw)) have[lds_col(f)] = 1;
  }
  // planes in: one field per wave and pass
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double v[kItF];
#pragma unroll
  for (int j = 0; j < kItF; ++j) {
    const int f = wave + j * (kBlock / 64);
    const double *p = a.field[f];
    if (p && !(cloud_borne(f) && !a.qqcw) && lane < n) v[j] = p[poff[lane]];
  }
#pragma unroll
  for (int j = 0; j < kItF; ++j) {
    const int f = wave + j * (kBlock / 64);
    if (a.field[f] && !(cloud_borne(f) && !a.qqcw) && lane < n)
      tile[lane * kRow + lds_col(f)] = v[j];
  }
  if (tend && threadIdx.x < n) {
    tile[threadIdx.x * kRow + 12 - kQ0] = 0.0;
    tile[threadIdx.x * kRow + 13 - kQ0] = 0.0;
  }
  __syncthreads();
  // rows out
#pragma unroll
  for (int j = 0; j < kItQ; ++j) {
    const int e = threadIdx.x + j * kBlock, i = e / kNQ, t = e - i * kNQ;
    if (e < n * kNQ && have[t]) a.q[roff[i] + kQ0 + t] = tile[i * kRow + t];
  }
  if (a.qqcw) {
#pragma unroll
    for (int j = 0; j < kItC; ++j) {
      const int e = threadIdx.x + j * kBlock, i = e / kNC, t = e - i * kNC;
      if (e < n * kNC && have[kNQ + t]) a.qqcw[roff[i] + kC0 + t] = tile[i * kRow + kNQ + t];
    }
  }
  if (tend)  // O3 and H2O2
    for (int e = threadIdx.x; e < n * 2; e += kBlock) a.q[roff[e >> 1] + 9 + (e & 1)] = 0.0;
}

// ------------------------------------------------------------------ setsox
// Aqueous sulfur chemistry, one lane per cell (setsox.hpp).
struct SoxArgs {
  const double *met[6];  // pressure, temperature, liquid water, cloud fraction,
                         // droplet number, air number density
  double *qin[4];        // O3 (read only), H2O2, H2SO4, SO2
  double *so4c[3];
  const double *numc[3];
  double *aq[MAM4X_GAS_PCNST], *h2[MAM4X_GAS_PCNST];
  int ncol, nlev, ldc, layout;
  double dt;
};

__global__ void __launch_bounds__(kBlock) setsox_kernel(SoxArgs a) {
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  if (c >= (long)a.ncol * a.nlev) return;
  const int fast = a.layout == MAM4X_LAYOUT_COL_LEV ? a.nlev : a.ncol;
  const long slow = c / fast;
  const long off = slow * a.ldc + (c - slow * fast);
  sox::Met m{a.met[0][off], a.met[1][off], a.met[2][off],
             a.met[3][off], a.met[4][off], a.met[5][off]};
  sox::Tracers q;
  q.o3 = a.qin[0][off];
  q.h2o2 = a.qin[1][off];
  q.h2so4 = a.qin[2][off];
  q.so2 = a.qin[3][off];
#pragma unroll
  for (int k = 0; k < 3; ++k) q.so4c[k] = a.so4c[k][off];
#pragma unroll
  for (int k = 0; k < 3; ++k) q.numc[k] = a.numc[k][off];
  q.numc[3] = 0.0;  // the primary-carbon mode has no cloud-borne sulfate slot
  sox::Rates r;
  sox::single_level(a.dt, m, q, r);
  a.qin[1][off] = q.h2o2;
  a.qin[2][off] = q.h2so4;
  a.qin[3][off] = q.so2;
#pragma unroll
  for (int k = 0; k < 3; ++k) a.so4c[k][off] = q.so4c[k];
  constexpr int so4[3] = {6, 14, 21};
  for (int i = 0; i < MAM4X_GAS_PCNST; ++i) {
    const int k = i == so4[0] ? 0 : i == so4[1] ? 1 : i == so4[2] ? 2 : -1;
    if (a.aq[i]) a.aq[i][off] = k >= 0 ? r.aqso4[k] : 0.0;
    if (a.h2[i]) a.h2[i][off] = k >= 0 ? r.aqh2so4[k] : 0.0;
  }
}

// ------------------------------------------------------------------ gas chem
// Gas-phase chemistry, one lane per cell (gas_chem.hpp).
struct GasChemArgs {
  const double *T, *photo;
  const double *ext[gc::kExt];
  const double *inv[5];  // M, H2O, OH, NO3, HO2
  const double *het[MAM4X_GAS_PCNST];
  double *qq[MAM4X_GAS_PCNST];
  int ncol, nlev, ldc, layout;
  double dt;
};

__global__ void __launch_bounds__(kBlock) gas_chem_kernel(GasChemArgs a) {
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  if (c >= (long)a.ncol * a.nlev) return;
  const int fast = a.layout == MAM4X_LAYOUT_COL_LEV ? a.nlev : a.ncol;
  const long slow = c / fast;
  const long off = slow * a.ldc + (c - slow * fast);
  double inv[gc::kInv] = {a.inv[0][off], 0.0, 0.0, a.inv[1][off],
                          a.inv[2][off], a.inv[3][off], a.inv[4][off], 0.0};
  double ext[gc::kExt], het[31], y[31], rxt[gc::kRxn];
#pragma unroll
  for (int i = 0; i < gc::kExt; ++i) ext[i] = (a.ext[i] ? a.ext[i][off] : 0.0) / inv[0];
  het[0] = y[0] = 0.0;
#pragma unroll
  for (int i = 1; i < 31; ++i) {
    het[i] = a.het[i] ? a.het[i][off] : 0.0;
    y[i] = a.qq[i][off];
  }
  gc::rates(a.T[off], a.photo[off], inv, rxt);
  (void)gc::imp_sol(y, rxt, het, ext, a.dt);
#pragma unroll
  for (int i = 1; i < 31; ++i) a.qq[i][off] = y[i];
}

// ------------------------------------------------------------------ setinv
// Invariant densities, one lane per cell (linoz.hpp).
struct SetinvArgs {
  const double *T, *p, *qv;
  const double *offline[4];
  double *inv[MAM4X_NFS];
  int ncol, nlev, ldc, layout;
};

__global__ void __launch_bounds__(kBlock) setinv_kernel(SetinvArgs a) {
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  if (c >= (long)a.ncol * a.nlev) return;
  const int fast = a.layout == MAM4X_LAYOUT_COL_LEV ? a.nlev : a.ncol;
  const long slow = c / fast;
  const long off = slow * a.ldc + (c - slow * fast);
  double co[4], v[MAM4X_NFS];
#pragma unroll
  for (int i = 0; i < 4; ++i) co[i] = a.offline[i] ? a.offline[i][off] : 0.0;
  lnz::invariants(a.T[off], a.p[off], a.qv[off], co, v);
#pragma unroll
  for (int i = 0; i < MAM4X_NFS; ++i)
    if (a.inv[i]) a.inv[i][off] = v[i];
}

// ------------------------------------------------------------------ linoz
// Stratospheric ozone (linoz.hpp).  A lane per column walking all its levels
// leaves a 21,600-column domain with 340 waves whose every level waits on
// memory.  The block kernel instead gives each 64-column block eight waves,
// each over an eighth of the levels: the waves form the levels' column
// slabs into LDS, one wave runs the columns' running sums there in level
// order (the reference's summation order, so the overhead column is
// unchanged to the bit), then all eight solve their levels, and one wave
// applies the surface sink.  Columns taller than kLinozLdsLevels use the
// lane-per-column kernel.
constexpr int kLinozBlock = 64;
constexpr int kLinozWaves = 8;  // 2 and 4 waves measured slower
constexpr int kLinozLdsLevels = 128;  // 64 KiB of slabs per block
struct LinozArgs {
  const double *T, *p, *pdel;
  const double *clim[lnz::kNClim];
  const double *sza, *lat, *col0;
  const int *ltrop;
  double *o3, *o3col, *do3, *do3_psc, *ss, *sink;
  lnz::Conf cf;
  int ncol, nlev, ldc, layout;
  double dt;
};

__device__ __forceinline__ int clamp_ltrop(const LinozArgs &a, int j) {
  const int lt = a.ltrop[j];
  return lt < 0 ? 0 : lt > a.nlev ? a.nlev : lt;
}

__global__ void __launch_bounds__(kLinozBlock) linoz_kernel(LinozArgs a) {
  const int j = blockIdx.x * kLinozBlock + threadIdx.x;
  if (j >= a.ncol) return;
  const bool lc = a.layout == MAM4X_LAYOUT_LEV_COL;
  const long base = lc ? (long)j : (long)j * a.ldc;
  const long ks = lc ? (long)a.ldc : 1L;
  const double *clim[lnz::kNClim];
#pragma unroll
  for (int i = 0; i < lnz::kNClim; ++i) clim[i] = a.clim[i] + base;
  double sink = 0.0;
  lnz::column(a.nlev, ks, a.dt, a.sza[j], a.lat[j], clamp_ltrop(a, j), a.col0 ? a.col0[j] : 0.0,
              a.cf, a.T + base, a.p + base, a.pdel + base, clim, a.o3 + base,
              a.o3col ? a.o3col + base : nullptr, a.do3 ? a.do3 + base : nullptr,
              a.do3_psc ? a.do3_psc + base : nullptr, a.ss ? a.ss + base : nullptr, &sink);
  if (a.sink) a.sink[j] = sink;
}

__global__ void __launch_bounds__(kLinozBlock *kLinozWaves) linoz_block_kernel(LinozArgs a) {
  extern __shared__ double slab[];  // [nlev][64]: level slabs, then overhead columns
  const int lane = threadIdx.x % kLinozBlock, w = threadIdx.x / kLinozBlock;
  const int j = blockIdx.x * kLinozBlock + lane;
  const bool live = j < a.ncol;
  const bool lc = a.layout == MAM4X_LAYOUT_LEV_COL;
  const long base = live ? (lc ? (long)j : (long)j * a.ldc) : 0L;
  const long ks = lc ? (long)a.ldc : 1L;
  const int per = (a.nlev + kLinozWaves - 1) / kLinozWaves;
  const int k0 = w * per, k1 = k0 + per < a.nlev ? k0 + per : a.nlev;
  if (live)
    for (int k = k0; k < k1; ++k) {
      const long i = base + k * ks;
      slab[k * kLinozBlock + lane] = lnz::kColumnFactor * a.pdel[i] * a.o3[i];
    }
  __syncthreads();
  if (w == 0 && live) {  // the running sums, top down
    const double col0 = a.col0 ? a.col0[j] : 0.0;
    double above = 0.0;
    for (int k = 0; k < a.nlev; ++k) {
      const double delta = slab[k * kLinozBlock + lane];
      slab[k * kLinozBlock + lane] = col0 + above + 0.5 * delta;
      above += delta;
    }
  }
  __syncthreads();
  if (live) {
    const int lt = clamp_ltrop(a, j);
    const double sza = a.sza[j], lats = a.lat[j] * lnz::kRadToDeg;
    for (int k = k0; k < k1; ++k) {
      const long i = base + k * ks;
      const double o3col = slab[k * kLinozBlock + lane];
      if (a.o3col) a.o3col[i] = o3col;
      lnz::LevelOut lo{0.0, 0.0, 0.0};
      if (k < lt) {
        double c[lnz::kNClim];
#pragma unroll
        for (int m = 0; m < lnz::kNClim; ++m) c[m] = a.clim[m][i];
        a.o3[i] = lnz::solve_level(o3col, a.T[i], sza, a.p[i], a.dt, lats, c, a.cf, a.o3[i], lo);
      }
      if (a.do3) a.do3[i] = lo.do3;
      if (a.do3_psc) a.do3_psc[i] = lo.do3_psc;
      if (a.ss) a.ss[i] = lo.ss_o3;
    }
  }
  __syncthreads();  // the sink reads ozone other waves wrote
  if (w == 0 && live) {
    const double sink = lnz::surface_sink(a.nlev, ks, a.dt, a.cf, a.pdel + base, a.o3 + base);
    if (a.sink) a.sink[j] = sink;
  }
}

// ------------------------------------------------------------------ sethet
// Washout rates, one lane per column (sethet.hpp); blocks of one wave
// spread the columns over the CUs.
struct SethetArgs {
  const double *press, *zmid, *T, *cmfdqr, *nrain, *nevapr, *inv, *h2o2, *lat, *phis;
  double *het[MAM4X_GAS_PCNST];
  double xkgm, dt;
  int ncol, nlev, ldc, layout;
};

__global__ void __launch_bounds__(kLinozBlock) sethet_kernel(SethetArgs a) {
  const int j = blockIdx.x * kLinozBlock + threadIdx.x;
  if (j >= a.ncol) return;
  const bool lc = a.layout == MAM4X_LAYOUT_LEV_COL;
  const long base = lc ? (long)j : (long)j * a.ldc;
  const long ks = lc ? (long)a.ldc : 1L;
  sh::column(a.nlev, ks, a.lat[j], a.phis[j], a.dt, a.xkgm, a.press + base, a.zmid + base,
             a.T + base, a.cmfdqr + base, a.nrain + base, a.nevapr + base, a.inv + base,
             a.h2o2 + base, a.het[1] + base, a.het[2] + base, a.het[3] + base);
  for (int m = 0; m < MAM4X_GAS_PCNST; ++m) {
    if (m >= 1 && m <= 3) continue;
    double *p = a.het[m];
    if (p)
      for (int k = 0; k < a.nlev; ++k) p[base + k * ks] = 0.0;
  }
}

// ------------------------------------------------------------------ setext
// External forcing, one lane per cell: each forced species' sectors summed
// in order (extfrc_set, mo_setext.hpp:22-54).
struct SetextArgs {
  const double *sec[MAM4X_EXTCNT][MAM4X_MAX_SECTORS];
  double *out[MAM4X_EXTCNT];  // by species mm: its destination plane
  int nsec[MAM4X_EXTCNT], alt[MAM4X_EXTCNT];
  int ncol, nlev, ldc, layout;
};

__global__ void __launch_bounds__(kBlock) setext_kernel(SetextArgs a) {
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  if (c >= (long)a.ncol * a.nlev) return;
  const bool cl = a.layout == MAM4X_LAYOUT_COL_LEV;
  const int fast = cl ? a.nlev : a.ncol;
  const long slow = c / fast;
  const int f = (int)(c - slow * fast);
  const int k = cl ? f : (int)slow, j = cl ? (int)slow : f;
  const long off = cl ? (long)j * a.ldc + k : (long)k * a.ldc + j;
  const int kf = a.nlev - 1 - k;  // the level read from bottom-up data
  const long off_alt = cl ? (long)j * a.ldc + kf : (long)kf * a.ldc + j;
  for (int m = 0; m < MAM4X_EXTCNT; ++m) {
    if (!a.out[m]) continue;
    const long o = a.alt[m] ? off_alt : off;
    double v = 0.0;
    for (int s = 0; s < a.nsec[m]; ++s) v += a.sec[m][s][o];
    a.out[m][off] = v;
  }
}

// ------------------------------------------------------------------ photolysis
// table_photo (photo.hpp): the column pass forms each daylight column's
// effective albedo and cloud multiplier into the two work planes, the level
// pass the rates.
struct PhotoArgs {
  pht::Table t;
  const double *pmid, *pdel, *T, *colo3, *lwc, *cld, *zen, *alb;
  double *photo, *eff, *mult;  // eff, mult: the work planes
  double esfact;
  int ncol, nlev, ldc, layout;
};

constexpr int kPhotoColBlock = 64;
// The column pass, blocked: a 256-lane block takes cpb columns (32, 16 or 8:
// three LDS slabs of nlev x cpb layer values within 64 KiB).  All lanes form
// the layers' optical depths, weighted depths and thick-cloud terms over a
// share of the levels (the pow and the divisions of every layer at once);
// then lanes 0..cpb-1 of wave 0 walk down the slabs for the multipliers
// while those of wave 1 walk up for the albedos, the running sums in level
// order.  Columns deeper than the slabs allow take the lane-per-column walk.
constexpr int kPhotoSlabBytes = 64 * 1024;
__global__ void __launch_bounds__(256) photo_column_block_kernel(PhotoArgs a, int cpb) {
  extern __shared__ double slab[];  // [3][nlev][cpb]: tau, ctau, f1c
  const int t = threadIdx.x, groups = 256 / cpb;
  const bool lc = a.layout == MAM4X_LAYOUT_LEV_COL;
  const int cc = lc ? t % cpb : t / groups, kk = lc ? t / cpb : t % groups;
  const int j = blockIdx.x * cpb + cc;
  const long nl = a.nlev;
  if (j < a.ncol) {
    const double zen = a.zen[j];
    if (pht::daylight(pht::sza_deg(zen))) {
      const double coschi = pht::coschi_of(zen);
      for (int k = kk; k < a.nlev; k += groups) {
        const long i = lc ? (long)k * a.ldc + j : (long)j * a.ldc + k;
        const pht::Layer l = pht::layer(coschi, a.cld[i], a.lwc[i], a.pdel[i]);
        slab[k * cpb + cc] = l.tau;
        slab[(nl + k) * cpb + cc] = l.ctau;
        slab[(2 * nl + k) * cpb + cc] = l.f1c;
      }
    }
  }
  __syncthreads();
  const int w = t / 64, lane = t % 64;
  if (w >= 2 || lane >= cpb) return;
  const int jw = blockIdx.x * cpb + lane;
  if (jw >= a.ncol) return;
  const double zen = a.zen[jw];
  if (!pht::daylight(pht::sza_deg(zen))) return;
  const long base = lc ? (long)jw : (long)jw * a.ldc;
  const long ks = lc ? (long)a.ldc : 1L;
  pht::Side s{0.0, 0.0, 0.0};
  if (w == 0) {  // top down: the multipliers
    const double coschi = pht::coschi_of(zen);
    for (int k = 0; k < a.nlev; ++k) {
      double a_tau = 0.0, a_cld = 0.0;
      if (k > 0) pht::side_cloud(s, a_tau, a_cld);
      a.mult[base + k * ks] =
          pht::multiplier(coschi, slab[(2 * nl + k) * cpb + lane], a_tau, a_cld);
      pht::side_add(s, slab[k * cpb + lane], slab[(nl + k) * cpb + lane]);
    }
  } else {  // bottom up: the albedos
    const double srf = a.alb[jw];
    for (int k = a.nlev - 1; k >= 0; --k) {
      double b_tau = 0.0, b_cld = 0.0;
      if (k < a.nlev - 1) pht::side_cloud(s, b_tau, b_cld);
      a.eff[base + k * ks] = pht::albedo(srf, b_tau, b_cld);
      pht::side_add(s, slab[k * cpb + lane], slab[(nl + k) * cpb + lane]);
    }
  }
}

// columns per block of photo_column_block_kernel for nlev levels (0: too deep)
static int photo_cols_per_block(int nlev) {
  for (int cpb = 32; cpb >= 8; cpb /= 2)
    if ((long)3 * nlev * cpb * sizeof(double) <= kPhotoSlabBytes) return cpb;
  return 0;
}

__global__ void __launch_bounds__(kPhotoColBlock) photo_column_kernel(PhotoArgs a) {
  const int j = blockIdx.x * kPhotoColBlock + threadIdx.x;
  if (j >= a.ncol) return;
  const double zen = a.zen[j];
  if (!pht::daylight(pht::sza_deg(zen))) return;
  const bool lc = a.layout == MAM4X_LAYOUT_LEV_COL;
  const long base = lc ? (long)j : (long)j * a.ldc;
  const long ks = lc ? (long)a.ldc : 1L;
  pht::cloud_column(a.nlev, ks, zen, a.cld + base, a.lwc + base, a.pdel + base, a.alb[j],
                    a.eff + base, a.mult + base);
}

// One lane per cell, its wavelengths summed in order.  The sixteen table
// values a wavelength needs are gathered per lane: neighbouring columns in
// different zenith-angle, albedo and ozone cells touch different lines, so
// the pass is bound by gathering ~8.6 KB of table per cell from L2 / the
// infinity cache.  (Sixteen lanes per cell over consecutive wavelengths of
// a wavelength-fastest copy, with the terms added in order through
// cross-lane reads, measured 1.28 ms against this kernel's 0.83 ms.)
__global__ void __launch_bounds__(kBlock) photo_level_kernel(PhotoArgs a) {
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  if (c >= (long)a.ncol * a.nlev) return;
  const bool cl = a.layout == MAM4X_LAYOUT_COL_LEV;
  const int fast = cl ? a.nlev : a.ncol;
  const long slow = c / fast;
  const int f = (int)(c - slow * fast);
  const int j = cl ? (int)slow : f;
  const long off = slow * a.ldc + f;
  const double sza = pht::sza_deg(a.zen[j]);
  if (!pht::daylight(sza)) return;
  const pht::SzaPos s = pht::sza_pos(a.t, sza);
  const double jr =
      pht::level_rate(a.t, s, a.eff[off], a.pmid[off] * pht::kPaToMb, a.T[off], a.colo3[off]);
  a.photo[off] = pht::update(a.t, a.photo[off], a.mult[off], a.esfact, jr);
}

}  // namespace

enum { kPre = 0, kPost = 1, kFused = 2 };

template <class P, class Src>
static void launch_fused(bool pre, bool multi, unsigned grid, hipStream_t s, const Src &src,
                         const Work &w) {
  if (multi) {  // ntsubstep > 1
    if (pre)
      hipLaunchKernelGGL((chain_fused_kernel<true, P, Src, true>), dim3(grid), dim3(kBlock), 0, s,
                         src, w);
    else
      hipLaunchKernelGGL((chain_fused_kernel<false, P, Src, true>), dim3(grid), dim3(kBlock), 0,
                         s, src, w);
  } else if (pre) {
    hipLaunchKernelGGL((chain_fused_kernel<true, P, Src, false>), dim3(grid), dim3(kBlock), 0, s,
                       src, w);
  } else {
    hipLaunchKernelGGL((chain_fused_kernel<false, P, Src, false>), dim3(grid), dim3(kBlock), 0,
                       s, src, w);
  }
}

// launches one of the per-cell kernels for policy P (FP64 wide type)
template <class P, class Src>
static void launch_cells(int kind, bool pre, unsigned grid, hipStream_t s, const Src &src,
                         const Work &w, bool multi = false) {
  if (kind == kPre) {
    if (pre)
      hipLaunchKernelGGL((chain_pre_kernel<true, P, Src>), dim3(grid), dim3(kBlock), 0, s, src, w);
    else
      hipLaunchKernelGGL((chain_pre_kernel<false, P, Src>), dim3(grid), dim3(kBlock), 0, s, src, w);
  } else if (kind == kPost) {
    if (pre)
      hipLaunchKernelGGL((chain_post_kernel<true, P, Src>), dim3(grid), dim3(kBlock), 0, s, src, w);
    else
      hipLaunchKernelGGL((chain_post_kernel<false, P, Src>), dim3(grid), dim3(kBlock), 0, s, src, w);
  } else {
    launch_fused<P>(pre, multi, grid, s, src, w);
  }
}

template <class P, bool MULTI = true>
static void launch_order(bool pre, unsigned grid, hipStream_t s, const OneDomain &src,
                         const Work &w, int pass) {
  if (pre)
    hipLaunchKernelGGL((chain_fused_order_kernel<true, P, MULTI>), dim3(grid), dim3(kBlock), 0, s,
                       src, w, pass);
  else
    hipLaunchKernelGGL((chain_fused_order_kernel<false, P, MULTI>), dim3(grid), dim3(kBlock), 0,
                       s, src, w, pass);
}

// Device memory of the allocation-free step (mam4x_compute_tendencies_noalloc):
// the work-queue workspace for up to kScratchCells cells -- W_NPLANE planes of
// doubles, the queue, the cost buckets -- and the queue counters.  Every GPU the
// library runs on holds one copy (117 B per cell: 61.3 MB) until the process
// ends; every call on a GPU shares it, so calls that overlap on one GPU race.
static constexpr long kScratchCells = MAM4X_NOALLOC_MAX_CELLS;
static constexpr size_t kScratchWsBytes = sizeof(double) * W_NPLANE * kScratchCells;
static constexpr size_t kScratchQueueOff = kScratchWsBytes;
static constexpr size_t kScratchBucketOff = kScratchQueueOff + sizeof(int) * kScratchCells;
static constexpr size_t kScratchCtlOff =
    (kScratchBucketOff + kScratchCells + 255) / 256 * 256;
static constexpr size_t kScratchBytes = kScratchCtlOff + sizeof(QueueCtl);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static_assert(kScratchBytes < 62 * 1000 * 1000, "scratch stays at 61.3 MB");

// The address of `scratch` on the current device.  In host code the array's
// name is only a stand-in; each device holds its own copy at its own address,
// resolved once per device and kept.
static int scratch_address(int device, unsigned char **out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static unsigned char *addr[kMaxDevices] = {};
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return MAM4X_ERR_ARG;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return MAM4X_ERR_HIP;
    }
    addr[device] = static_cast<unsigned char *>(p);
  }
  *out = addr[device];
  return MAM4X_OK;
}

struct mam4x_handle_s {
  mam4x_config cfg;
  int device = 0;
  ChainConsts *d_consts = nullptr;
  unsigned long long *d_count = nullptr;  // validate counter
  unsigned long long *d_stats = nullptr;  // SOA sub-steps since the last read
  bool stats = false;
  // work-queue schedule: workspace for up to `cap` cells (grown on demand)
  long cap = 0;
  double *ws = nullptr;
  unsigned char *bucket = nullptr;
  int *queue = nullptr;
  QueueCtl *ctl = nullptr;
  unsigned soa_grid = 0;  // resident blocks of soa_queue_kernel on the device
  int ncu = 0, soa_per_cu = 0;  // CUs, and the most SOA blocks one CU holds
  int nbatch = 0;               // live batches made from this handle
  unsigned log_grid = 0;  // resident blocks of soa_log_kernel on the device
  // the device's allocation-free workspace (`scratch`), NULL if unresolved
  unsigned char *scratch = nullptr;
  // adjoint: the phases' planes (kAdjPlanes x adj_cap doubles), and the SOA
  // log: log_cap chunks of values and tags, their back links, each wave's
  // end and the pool counter (grown on demand)
  double *adjws = nullptr;
  long adj_cap = 0;
  double *subws = nullptr;  // ntsubstep > 1: the sub-step checkpoints
  size_t sub_cap = 0;       // doubles
  double *log_rec = nullptr;
  int *log_tag = nullptr, *log_prev = nullptr;
  int2 *log_wend = nullptr;
  unsigned *log_next = nullptr;
  unsigned log_cap = 0;
  // phase timing (stats enabled): 4 events per call, recycled through a pool
  std::vector<std::array<hipEvent_t, 4>> pending, pool;
  std::string err;
};

// cells below which the fused one-lane-per-cell kernel is used by
// MAM4X_SCHED_AUTO: small domains do not fill the chip with persistent waves
// and pay four launches for little divergence saving
static constexpr long kFusedBelow = 16384;

static void destroy_events(mam4x_handle h) {
  for (auto *v : {&h->pending, &h->pool})
    for (auto &q : *v)
      for (hipEvent_t e : q) (void)hipEventDestroy(e);
  h->pending.clear();
  h->pool.clear();
}

static void free_work(mam4x_handle h) {
  if (h->ws) (void)hipFree(h->ws);
  if (h->bucket) (void)hipFree(h->bucket);
  if (h->queue) (void)hipFree(h->queue);
  h->ws = nullptr;
  h->bucket = nullptr;
  h->queue = nullptr;
  h->cap = 0;
}

static int fail(mam4x_handle h, int code, const std::string &msg) {
  if (h) h->err = msg;
  return code;
}

#define HIP_TRY(h, expr)                                                     \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail((h), MAM4X_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Workspace for ncell cells.  Growing frees the old buffers only after the
// stream has drained, since earlier launches may still use them.
static int reserve_work(mam4x_handle h, long ncell, hipStream_t s) {
  if (ncell <= h->cap) return MAM4X_OK;
  if (h->cap) HIP_TRY(h, hipStreamSynchronize(s));
  free_work(h);
  if (hipMalloc(&h->ws, sizeof(double) * W_NPLANE * ncell) != hipSuccess ||
      hipMalloc(&h->bucket, (size_t)ncell) != hipSuccess ||
      hipMalloc(&h->queue, sizeof(int) * ncell) != hipSuccess) {
    (void)hipGetLastError();
    free_work(h);
    return fail(h, MAM4X_ERR_NOMEM, "workspace allocation failed (" +
                std::to_string(ncell) + " cells)");
  }
  h->cap = ncell;
  return MAM4X_OK;
}

static int check_inputs(mam4x_handle h, const mam4x_inputs *in) {
  if (!in) return fail(h, MAM4X_ERR_ARG, "inputs is NULL");
  if (in->ncol < 0 || in->nlev < 0) return fail(h, MAM4X_ERR_ARG, "negative ncol or nlev");
  if (in->layout != MAM4X_LAYOUT_LEV_COL && in->layout != MAM4X_LAYOUT_COL_LEV)
    return fail(h, MAM4X_ERR_ARG, "unknown layout");
  const bool lc = in->layout == MAM4X_LAYOUT_LEV_COL;
  const int fast = lc ? in->ncol : in->nlev, slow = lc ? in->nlev : in->ncol;
  const int ld = in->ldc ? in->ldc : fast;
  if (ld < fast) return fail(h, MAM4X_ERR_ARG, "ldc smaller than the fastest extent");
  if ((long long)ld * slow > 0x7fffffffLL)
    return fail(h, MAM4X_ERR_ARG, "domain too large (ldc x slow extent must fit in int32)");
  for (int f = 0; f < MAM4X_NSTATE; ++f)
    if (!in->state[f]) return fail(h, MAM4X_ERR_ARG, "state plane " + std::to_string(f) + " is NULL");
  if (!in->temperature || !in->pressure || !in->height || !in->vapor_mixing_ratio ||
      !in->cloud_fraction || !in->pblh)
    return fail(h, MAM4X_ERR_ARG, "an atmosphere plane is NULL");
  for (int m = 0; m < 4; ++m)
    if (!in->dgncur_a[m] || !in->dgncur_awet[m] || !in->wetdens[m])
      return fail(h, MAM4X_ERR_ARG, "a modal diameter/density plane is NULL");
  return MAM4X_OK;
}

static void fill_args(mam4x_handle h, const mam4x_inputs *in, ChainArgs &a) {
  std::memset(&a, 0, sizeof(a));
  for (int f = 0; f < NSTATE; ++f) {
    a.state[f] = in->state[f];
    a.precld[f] = in->precldchem[f];
  }
  a.T = in->temperature;
  a.p = in->pressure;
  a.z = in->height;
  a.qv = in->vapor_mixing_ratio;
  a.cld = in->cloud_fraction;
  a.pblh = in->pblh;
  for (int m = 0; m < 4; ++m) {
    a.dga[m] = in->dgncur_a[m];
    a.dgw[m] = in->dgncur_awet[m];
    a.wd[m] = in->wetdens[m];
  }
  a.pre_h2so4 = in->pregaschem_h2so4;
  a.pre_soag = in->pregaschem_soag;
  a.K = h->d_consts;
  a.ncol = in->ncol;
  a.nlev = in->nlev;
  a.layout = in->layout;
  a.ldc = in->ldc ? in->ldc : (in->layout == MAM4X_LAYOUT_COL_LEV ? in->nlev : in->ncol);
  const mam4x_config &c = h->cfg;
  a.cfg.do_cond = c.do_cond;
  a.cfg.do_rename = c.do_rename;
  a.cfg.do_newnuc = c.do_newnuc;
  a.cfg.do_coag = c.do_coag;
  // QUIRK (mam4_amicphys.hpp:1129 vs :1914): the two optaa switches arrive
  // swapped in mam_amicphys_1subarea, so the H2SO4 production treatment
  // follows newnuc_h2so4_conc_optaa.
  a.cfg.gaexch_h2so4_uptake_optaa = c.newnuc_h2so4_conc_optaa;
  a.cfg.nmono = c.n_so4_monolayers_pcage;
  a.cfg.ntsubstep = c.ntsubstep;
  // level-major tiles only where they cover the domain exactly
  const int lpw = c.levels_per_wave > 1 ? c.levels_per_wave : 1;
  a.lpw = (lpw > 1 && a.layout == MAM4X_LAYOUT_LEV_COL && a.ncol % (64 / lpw) == 0 &&
           a.nlev % lpw == 0) ? lpw : 1;
  a.do_calcsize = c.do_calcsize;
  a.do_xfer = c.calcsize_do_aitacc_transfer;
  a.do_adjust = c.calcsize_do_adjust;
}

extern "C" {

int mam4x_abi_version(void) { return MAM4X_ABI_VERSION; }

int mam4x_default_config(mam4x_config *out) {
  if (!out) return MAM4X_ERR_ARG;
  out->do_cond = 1;
  out->do_rename = 1;
  out->do_newnuc = 1;
  out->do_coag = 1;
  out->gaexch_h2so4_uptake_optaa = 2;
  out->newnuc_h2so4_conc_optaa = 2;
  out->n_so4_monolayers_pcage = 8;
  out->do_calcsize = 1;
  out->calcsize_do_aitacc_transfer = 1;
  out->calcsize_do_adjust = 1;
  out->precision = MAM4X_PREC_FP64;
  out->schedule = MAM4X_SCHED_AUTO;
  out->soa_blocks_per_cu = 0;
  out->ntsubstep = 1;
  out->soa_chunk = 0;
  out->soa_short_chunk = 0;
  out->soa_short_bucket = 0;
  out->levels_per_wave = 1;
  return MAM4X_OK;
}

// the checks mam4x_create and mam4x_init apply to a configuration
static bool config_ok(const mam4x_config &c) {
  return (c.precision == MAM4X_PREC_FP64 || c.precision == MAM4X_PREC_MIXED) &&
         c.schedule >= MAM4X_SCHED_AUTO && c.schedule <= MAM4X_SCHED_QUEUE &&
         c.soa_blocks_per_cu >= 0 && c.ntsubstep >= 0 && c.ntsubstep <= MAM4X_MAX_NTSUBSTEP &&
         c.soa_chunk >= 0 && c.soa_chunk <= (1 << 20) && c.soa_short_chunk >= 0 &&
         c.soa_short_chunk <= (1 << 20) && c.soa_short_bucket >= 0 &&
         c.soa_short_bucket <= kNBucket &&
         (c.levels_per_wave == 0 || c.levels_per_wave == 1 || c.levels_per_wave == 2 ||
          c.levels_per_wave == 4 || c.levels_per_wave == 8);
}

// Resident blocks of the persistent SOA kernel per CU when the
// configuration leaves it at 0.  Measured on 21,600 x 72, chained steps from
// the seeded state: 3 blocks (12 waves per CU) 2.842 ms per step against
// 2.896 with as many as fit (4) over steps 1-10, and 2.611 against 2.605
// over steps 11-20.
constexpr int kSoaBlocksDefault = 3;

// Resident blocks per CU of the reverse sweep's SOA log kernel (its replay
// kernel runs one wave per log wave).  Measured on the 21,600 x 72 sweep of
// step 1: 2 blocks 12.03 ms per sweep, as many as fit 12.75, 1 block 13.34.
constexpr int kLogBlocks = 2;

// resident blocks of the persistent SOA kernel under the handle's cap
static unsigned soa_grid(mam4x_handle h) {
  const int cap = h->cfg.soa_blocks_per_cu > 0 ? h->cfg.soa_blocks_per_cu : kSoaBlocksDefault;
  const int per_cu = cap < h->soa_per_cu ? cap : h->soa_per_cu;
  return (unsigned)(h->ncu * per_cu);
}

int mam4x_create(const mam4x_config *config, mam4x_handle *out) {
  if (!out) return MAM4X_ERR_ARG;
  *out = nullptr;
  mam4x_handle h = new (std::nothrow) mam4x_handle_s();
  if (!h) return MAM4X_ERR_NOMEM;
  if (config)
    h->cfg = *config;
  else
    mam4x_default_config(&h->cfg);
  if (!config_ok(h->cfg)) {
    delete h;
    return MAM4X_ERR_ARG;
  }
  int ncu = 0, per_cu = 0;
  if (hipGetDevice(&h->device) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device) !=
          hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, soa_queue_kernel, kBlock, 0) !=
          hipSuccess) {
    delete h;
    return MAM4X_ERR_HIP;
  }
  h->ncu = ncu;
  h->soa_per_cu = per_cu < 1 ? 1 : per_cu;
  h->soa_grid = soa_grid(h);
  int log_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&log_cu, soa_log_kernel, kBlock, 0) !=
      hipSuccess) {
    delete h;
    return MAM4X_ERR_HIP;
  }
  if (log_cu > kLogBlocks) log_cu = kLogBlocks;
  h->log_grid = (unsigned)(ncu * (log_cu > 0 ? log_cu : 1));
  // only the allocation-free step uses it; that entry point reports a failure
  if (scratch_address(h->device, &h->scratch) != MAM4X_OK) h->scratch = nullptr;
  ChainConsts k;
  chain_consts_init(&k);
  if (hipMalloc(&h->d_consts, sizeof(ChainConsts)) != hipSuccess ||
      hipMalloc(&h->d_count, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&h->d_stats, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&h->ctl, sizeof(QueueCtl)) != hipSuccess) {
    (void)hipGetLastError();
    mam4x_destroy(h);
    return MAM4X_ERR_NOMEM;
  }
  if (hipMemcpy(h->d_consts, &k, sizeof(k), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(h->d_stats, 0, sizeof(unsigned long long)) != hipSuccess) {
    mam4x_destroy(h);
    return MAM4X_ERR_HIP;
  }
  *out = h;
  return MAM4X_OK;
}

int mam4x_destroy(mam4x_handle h) {
  if (!h) return MAM4X_ERR_ARG;
  if (h->d_consts) (void)hipFree(h->d_consts);
  if (h->d_count) (void)hipFree(h->d_count);
  if (h->d_stats) (void)hipFree(h->d_stats);
  if (h->ctl) (void)hipFree(h->ctl);
  for (void *q : {(void *)h->adjws, (void *)h->subws, (void *)h->log_rec, (void *)h->log_tag,
                  (void *)h->log_prev, (void *)h->log_wend, (void *)h->log_next})
    if (q) (void)hipFree(q);
  free_work(h);
  destroy_events(h);
  delete h;
  return MAM4X_OK;
}

int mam4x_get_config(mam4x_handle h, mam4x_config *out) {
  if (!h || !out) return MAM4X_ERR_ARG;
  *out = h->cfg;
  return MAM4X_OK;
}

int mam4x_init(mam4x_handle h, const mam4x_config *config) {
  if (!h) return MAM4X_ERR_ARG;
  mam4x_config c;
  if (config)
    c = *config;
  else
    mam4x_default_config(&c);
  if (!config_ok(c)) return fail(h, MAM4X_ERR_ARG, "invalid configuration");
  if (h->nbatch > 0)
    return fail(h, MAM4X_ERR_ARG, "the handle has " + std::to_string(h->nbatch) +
                " live batch(es); destroy them before re-initializing");
  h->cfg = c;
  h->soa_grid = soa_grid(h);
  return MAM4X_OK;
}

int mam4x_set_precision(mam4x_handle h, int precision) {
  if (!h) return MAM4X_ERR_ARG;
  if (precision != MAM4X_PREC_FP64 && precision != MAM4X_PREC_MIXED)
    return fail(h, MAM4X_ERR_ARG, "unknown precision");
  h->cfg.precision = precision;
  return MAM4X_OK;
}

const char *mam4x_last_error(mam4x_handle h) { return h ? h->err.c_str() : "NULL handle"; }

int mam4x_enable_stats(mam4x_handle h, int enable) {
  if (!h) return MAM4X_ERR_ARG;
  if (enable && !h->stats) {
    HIP_TRY(h, hipMemset(h->d_stats, 0, sizeof(unsigned long long)));
    for (auto &q : h->pending) h->pool.push_back(q);
    h->pending.clear();
  }
  h->stats = enable != 0;
  return MAM4X_OK;
}

int mam4x_last_stats(mam4x_handle h, long long *soa_iterations, void *stream) {
  if (!h || !soa_iterations) return MAM4X_ERR_ARG;
  if (!h->stats) return fail(h, MAM4X_ERR_UNSUPPORTED, "stats not enabled");
  unsigned long long v = 0;
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(h, hipMemcpyAsync(&v, h->d_stats, sizeof(v), hipMemcpyDeviceToHost, s));
  HIP_TRY(h, hipMemsetAsync(h->d_stats, 0, sizeof(v), s));
  HIP_TRY(h, hipStreamSynchronize(s));
  *soa_iterations = (long long)v;
  return MAM4X_OK;
}

int mam4x_last_phase_times(mam4x_handle h, double ms[MAM4X_NPHASE], long long *calls,
                           void *stream) {
  if (!h || !ms || !calls) return MAM4X_ERR_ARG;
  if (!h->stats) return fail(h, MAM4X_ERR_UNSUPPORTED, "stats not enabled");
  HIP_TRY(h, hipStreamSynchronize((hipStream_t)stream));
  for (int i = 0; i < MAM4X_NPHASE; ++i) ms[i] = 0.0;
  for (auto &q : h->pending) {
    for (int i = 0; i < MAM4X_NPHASE; ++i) {
      float t = 0.0f;
      HIP_TRY(h, hipEventElapsedTime(&t, q[i], q[i + 1]));
      ms[i] += t;
    }
    h->pool.push_back(q);
  }
  *calls = (long long)h->pending.size();
  h->pending.clear();
  return MAM4X_OK;
}

// what mam4x_validate reads: the domain, the state, T, p and qv
static int check_validate_inputs(mam4x_handle h, const mam4x_inputs *in) {
  if (!in) return fail(h, MAM4X_ERR_ARG, "inputs is NULL");
  if (in->ncol < 0 || in->nlev < 0) return fail(h, MAM4X_ERR_ARG, "negative ncol or nlev");
  if (in->layout != MAM4X_LAYOUT_LEV_COL && in->layout != MAM4X_LAYOUT_COL_LEV)
    return fail(h, MAM4X_ERR_ARG, "unknown layout");
  const bool lc = in->layout == MAM4X_LAYOUT_LEV_COL;
  const int fast = lc ? in->ncol : in->nlev, slow = lc ? in->nlev : in->ncol;
  const int ld = in->ldc ? in->ldc : fast;
  if (ld < fast) return fail(h, MAM4X_ERR_ARG, "ldc smaller than the fastest extent");
  if ((long long)ld * slow > 0x7fffffffLL)
    return fail(h, MAM4X_ERR_ARG, "domain too large (ldc x slow extent must fit in int32)");
  for (int f = 0; f < MAM4X_NSTATE; ++f)
    if (!in->state[f]) return fail(h, MAM4X_ERR_ARG, "state plane " + std::to_string(f) + " is NULL");
  if (!in->temperature || !in->pressure || !in->vapor_mixing_ratio)
    return fail(h, MAM4X_ERR_ARG, "temperature, pressure or vapor_mixing_ratio is NULL");
  return MAM4X_OK;
}

int mam4x_validate(mam4x_handle h, const mam4x_inputs *in, long long *violations,
                   void *stream) {
  if (!h || !violations) return MAM4X_ERR_ARG;
  int rc = check_validate_inputs(h, in);
  if (rc) return rc;
  ChainArgs a;
  fill_args(h, in, a);
  const long ncell = (long)in->ncol * in->nlev;
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(h, hipMemsetAsync(h->d_count, 0, sizeof(unsigned long long), s));
  if (ncell > 0) {
    const unsigned grid = (unsigned)((ncell + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(validate_kernel, dim3(grid), dim3(kBlock), 0, s, a, h->d_count);
    HIP_TRY(h, hipGetLastError());
  }
  unsigned long long v = 0;
  HIP_TRY(h, hipMemcpyAsync(&v, h->d_count, sizeof(v), hipMemcpyDeviceToHost, s));
  HIP_TRY(h, hipStreamSynchronize(s));
  *violations = (long long)v;
  return MAM4X_OK;
}

// outputs of a step into the argument block
static int fill_outputs(mam4x_handle h, const mam4x_outputs *out, ChainArgs &a) {
  if (!out) return fail(h, MAM4X_ERR_ARG, "outputs is NULL");
  for (int f = 0; f < NSTATE; ++f) {
    if (!out->tendencies[f])
      return fail(h, MAM4X_ERR_ARG, "tendency plane " + std::to_string(f) + " is NULL");
    a.tend[f] = out->tendencies[f];
    a.sout[f] = out->state_out[f];
  }
  int nsout = 0;
  for (int f = 0; f < NSTATE; ++f) nsout += out->state_out[f] != nullptr;
  if (nsout != 0 && nsout != NSTATE)
    return fail(h, MAM4X_ERR_ARG, "state_out must be all NULL or all set");
  a.has_sout = nsout == NSTATE;
  a.soacap = out->soa_capped;
  a.has_dgn = 0;
  for (int m = 0; m < 4; ++m) {
    a.dgi[m] = out->dgncur_i[m];
    a.dgc[m] = out->dgncur_c[m];
    a.has_dgn |= (a.dgi[m] != nullptr) || (a.dgc[m] != nullptr);
  }
  bool proc = false;
  for (int k = 0; k < MAM4X_NPROCESS; ++k) {
    a.proc[k] = out->process_tendencies[k];
    proc = proc || a.proc[k];
  }
  a.pstride = 0;
  if (proc) {
    // plane extent: slow extent x ldc
    const long extent = (long)(a.layout == MAM4X_LAYOUT_COL_LEV ? a.ncol : a.nlev) * a.ldc;
    if (out->process_stride < 0 || (out->process_stride && out->process_stride < extent))
      return fail(h, MAM4X_ERR_ARG, "process_stride smaller than the plane extent");
    a.pstride = out->process_stride ? (long)out->process_stride : extent;
  }
  return MAM4X_OK;
}

static bool has_proc(const ChainArgs &a) {
  for (int k = 0; k < MAM4X_NPROCESS; ++k)
    if (a.proc[k]) return true;
  return false;
}

static bool has_pre(const mam4x_inputs *in) {
  bool pre = in->pregaschem_h2so4 || in->pregaschem_soag;
  for (int f = 0; f < NSTATE; ++f) pre = pre || in->precldchem[f];
  return pre;
}

int mam4x_compute_tangent(mam4x_handle h, double t, double dt, const mam4x_inputs *in,
                          const mam4x_tangent *dir, const mam4x_outputs *out, void *stream) {
  (void)t;
  if (!h) return MAM4X_ERR_ARG;
  int rc = check_inputs(h, in);
  if (rc) return rc;
  if (!dir) return fail(h, MAM4X_ERR_ARG, "tangent is NULL");
  if (!(dt > 0.0)) return fail(h, MAM4X_ERR_ARG, "dt must be > 0");
  ChainArgs a;
  fill_args(h, in, a);
  rc = fill_outputs(h, out, a);
  if (rc) return rc;
  if (has_proc(a))
    return fail(h, MAM4X_ERR_UNSUPPORTED, "process tendencies are not formed by the tangent step");
  if (out->column_budgets)
    return fail(h, MAM4X_ERR_UNSUPPORTED, "column budgets are not formed by the tangent step");
  for (int f = 0; f < NSTATE; ++f) {
    if (!dir->tendencies[f])
      return fail(h, MAM4X_ERR_ARG, "tangent tendency plane " + std::to_string(f) + " is NULL");
    a.dstate[f] = dir->state[f];
    a.dtend[f] = dir->tendencies[f];
  }
  a.dT = dir->temperature;
  a.dp = dir->pressure;
  a.dqv = dir->vapor_mixing_ratio;
  a.dt = dt;
  hipStream_t s = (hipStream_t)stream;
  const long ncell = (long)in->ncol * in->nlev;
  if (ncell == 0) return MAM4X_OK;
  const unsigned grid = (unsigned)((ncell + kBlock - 1) / kBlock);
  Work w;
  std::memset(&w, 0, sizeof(w));
  w.qs = queue_shape(h->cfg);
  w.ncell = ncell;
  if (h->stats) w.stats = h->d_stats;
  const bool pre = has_pre(in), multi = dev::nsub(a.cfg) > 1;
  const OneDomain src{a};
  if (a.cfg.do_cond) {
    // cells in the work queue's order (the FP64 set-up kernel predicts each
    // cell's SOA cost): a wave's lanes then run loops of similar length
    rc = reserve_work(h, ncell, s);
    if (rc) return rc;
    w.ws = h->ws;
    w.bucket = h->bucket;
    w.queue = h->queue;
    w.ctl = h->ctl;
    w.dtfull = dev::dt_sub(a.cfg, dt);
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, s, w.ctl);
    launch_cells<Fp64>(kPre, pre, grid, s, src, w);
    hipLaunchKernelGGL(queue_scatter_kernel, dim3(grid), dim3(kBlock), 0, s, w);
    for (int pass = 0; pass < 2; ++pass) {
      if (multi)
        launch_order<Tangent, true>(pre, grid, s, src, w, pass);
      else
        launch_order<Tangent, false>(pre, grid, s, src, w, pass);
    }
  } else {
    launch_fused<Tangent>(pre, multi, grid, s, src, w);
  }
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

// Planes of the reverse sweep for ncell cells, and its log's per-wave ends
// and pool counter.  Growing frees the old planes only after the stream has
// drained.
static int reserve_adjoint(mam4x_handle h, long ncell, hipStream_t s) {
  if (!h->log_wend) {
    if (hipMalloc(&h->log_wend, sizeof(int2) * h->log_grid * (kBlock / 64)) != hipSuccess ||
        hipMalloc(&h->log_next, sizeof(unsigned)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(h, MAM4X_ERR_NOMEM, "adjoint log allocation failed");
    }
  }
  if (ncell <= h->adj_cap) return MAM4X_OK;
  if (h->adj_cap) HIP_TRY(h, hipStreamSynchronize(s));
  if (h->adjws) (void)hipFree(h->adjws);
  h->adjws = nullptr;
  h->adj_cap = 0;
  if (hipMalloc(&h->adjws, sizeof(double) * kAdjPlanes * (size_t)ncell) != hipSuccess) {
    (void)hipGetLastError();
    return fail(h, MAM4X_ERR_NOMEM, "adjoint workspace allocation failed");
  }
  h->adj_cap = ncell;
  return MAM4X_OK;
}

// The SOA log's pool: at least nchunk chunks (the stream is idle).
static int reserve_log(mam4x_handle h, unsigned long long nchunk) {
  if (nchunk <= h->log_cap) return MAM4X_OK;
  if (nchunk > 0x7fffffffULL)
    return fail(h, MAM4X_ERR_NOMEM, "adjoint log too large (" + std::to_string(nchunk) +
                " chunks)");
  for (void *q : {(void *)h->log_rec, (void *)h->log_tag, (void *)h->log_prev})
    if (q) (void)hipFree(q);
  h->log_rec = nullptr;
  h->log_tag = h->log_prev = nullptr;
  h->log_cap = 0;
  const size_t rows = (size_t)nchunk * kLogRows;
  if (hipMalloc(&h->log_rec, sizeof(double) * kLogVals * 64 * rows) != hipSuccess ||
      hipMalloc(&h->log_tag, sizeof(int) * 64 * rows) != hipSuccess ||
      hipMalloc(&h->log_prev, sizeof(int) * (size_t)nchunk) != hipSuccess) {
    (void)hipGetLastError();
    return fail(h, MAM4X_ERR_NOMEM, "adjoint log allocation failed (" +
                std::to_string(nchunk) + " chunks)");
  }
  h->log_cap = (unsigned)nchunk;
  return MAM4X_OK;
}

int mam4x_compute_adjoint(mam4x_handle h, double t, double dt, const mam4x_inputs *in,
                          const mam4x_adjoint *adj, void *stream) {
  (void)t;
  if (!h) return MAM4X_ERR_ARG;
  int rc = check_inputs(h, in);
  if (rc) return rc;
  if (!adj) return fail(h, MAM4X_ERR_ARG, "adjoint is NULL");
  if (!(dt > 0.0)) return fail(h, MAM4X_ERR_ARG, "dt must be > 0");
  ChainArgs a;
  fill_args(h, in, a);  // no primal outputs: tend/sout/dgn stay NULL
  for (int f = 0; f < NSTATE; ++f) a.tbar[f] = adj->tendencies[f];
  a.dt = dt;
  hipStream_t s = (hipStream_t)stream;
  const long ncell = (long)in->ncol * in->nlev;
  if (ncell == 0) return MAM4X_OK;
  if (ncell > 0x7fffffffL) return fail(h, MAM4X_ERR_ARG, "domain too large for the adjoint");
  const unsigned grid = (unsigned)((ncell + kBlock - 1) / kBlock);
  const bool pre = has_pre(in);
  AdjArgs o;
  std::memset(&o, 0, sizeof(o));
  for (int f = 0; f < NSTATE; ++f) o.sbar[f] = adj->state[f];
  o.Tbar = adj->temperature;
  o.pbar = adj->pressure;
  o.qvbar = adj->vapor_mixing_ratio;
  if (dev::nsub(a.cfg) > 1) {
    const size_t need = (size_t)dev::nsub(a.cfg) * kSubVals * (size_t)ncell;
    if (need > h->sub_cap) {
      if (h->sub_cap) HIP_TRY(h, hipStreamSynchronize(s));
      if (h->subws) (void)hipFree(h->subws);
      h->subws = nullptr;
      h->sub_cap = 0;
      if (hipMalloc(&h->subws, sizeof(double) * need) != hipSuccess) {
        (void)hipGetLastError();
        return fail(h, MAM4X_ERR_NOMEM, "adjoint sub-step checkpoint allocation failed");
      }
      h->sub_cap = need;
    }
    const int *queue = nullptr;
    const QueueCtl *ctl = nullptr;
    if (a.cfg.do_cond) {
      // the first sub-step's predicted SOA cost orders the cells
      rc = reserve_work(h, ncell, s);
      if (rc) return rc;
      Work w;
      std::memset(&w, 0, sizeof(w));
      w.qs = queue_shape(h->cfg);
      w.ncell = ncell;
      w.ctl = h->ctl;
      w.dtfull = dev::dt_sub(a.cfg, dt);
      w.ws = h->ws;
      w.bucket = h->bucket;
      w.queue = h->queue;
      hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, s, h->ctl);
      launch_cells<Fp64>(kPre, pre, grid, s, OneDomain{a}, w);
      hipLaunchKernelGGL(queue_scatter_kernel, dim3(grid), dim3(kBlock), 0, s, w);
      queue = h->queue;
      ctl = h->ctl;
      o.bucket = h->bucket;
      if (pre)
        hipLaunchKernelGGL((chain_adjoint_sub_kernel<true>), dim3(grid), dim3(kBlock), 0, s, a,
                           o, h->subws, queue, ctl, 0);
      else
        hipLaunchKernelGGL((chain_adjoint_sub_kernel<false>), dim3(grid), dim3(kBlock), 0, s, a,
                           o, h->subws, queue, ctl, 0);
    }
    if (pre)
      hipLaunchKernelGGL((chain_adjoint_sub_kernel<true>), dim3(grid), dim3(kBlock), 0, s, a, o,
                         h->subws, queue, ctl, 1);
    else
      hipLaunchKernelGGL((chain_adjoint_sub_kernel<false>), dim3(grid), dim3(kBlock), 0, s, a,
                         o, h->subws, queue, ctl, 1);
    HIP_TRY(h, hipGetLastError());
    return MAM4X_OK;
  }
  if (a.cfg.do_cond) {
    rc = reserve_work(h, ncell, s);
    if (rc) return rc;
    rc = reserve_adjoint(h, ncell, s);
    if (rc) return rc;
    if (!h->log_cap) {
      // a first pool for 64 sub-steps per cell (at most 6 GiB); grown below
      // when short
      const unsigned long long chunk_bytes = (unsigned long long)kLogRows * 64 *
                                             (kLogVals * sizeof(double) + sizeof(int));
      unsigned long long first = (unsigned long long)ncell * 64 / (64 * kLogRows) + 64;
      first = std::min(first, (6ULL << 30) / chunk_bytes);
      rc = reserve_log(h, first);
      if (rc) return rc;
    }
    // 1. set-up and queue
    Work w;
    std::memset(&w, 0, sizeof(w));
    w.qs = queue_shape(h->cfg);
    w.ncell = ncell;
    w.ctl = h->ctl;
    w.dtfull = (double)(int)(dt / 1);  // QUIRK (mam4_amicphys.hpp:1319): int sub-step
    w.ws = h->ws;
    w.bucket = h->bucket;
    w.queue = h->queue;
    const OneDomain src{a};
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, s, h->ctl);
    launch_cells<Fp64>(kPre, pre, grid, s, src, w);
    hipLaunchKernelGGL(queue_scatter_kernel, dim3(grid), dim3(kBlock), 0, s, w);
    HIP_TRY(h, hipGetLastError());
    AdjWork g;
    g.qs = queue_shape(h->cfg);
    g.ws = h->ws;
    g.res = h->adjws;
    g.pb = h->adjws + (size_t)R_NPLANE * ncell;
    g.inb = h->adjws + (size_t)(R_NPLANE + kPbPlanes) * ncell;
    g.wend = h->log_wend;
    g.next = h->log_next;
    g.nwave = h->log_grid * (kBlock / 64);
    g.bucket = h->bucket;
    g.queue = h->queue;
    g.ctl = h->ctl;
    g.ncell = ncell;
    g.dtfull = w.dtfull;
    // 2. the loop, logged; once more with a larger pool if it ran out
    for (int attempt = 0;; ++attempt) {
      g.rec = h->log_rec;
      g.tag = h->log_tag;
      g.prev = h->log_prev;
      g.cap = h->log_cap;
      hipLaunchKernelGGL(queue_rewind_kernel, dim3(1), dim3(64), 0, s, h->ctl, h->log_next);
      hipLaunchKernelGGL(soa_log_kernel, dim3(h->log_grid), dim3(kBlock), 0, s, g);
      HIP_TRY(h, hipGetLastError());
      unsigned used = 0;
      HIP_TRY(h, hipMemcpyAsync(&used, h->log_next, sizeof(used), hipMemcpyDeviceToHost, s));
      HIP_TRY(h, hipStreamSynchronize(s));
      if (used <= h->log_cap) break;
      if (attempt > 0) return fail(h, MAM4X_ERR_HIP, "adjoint log: pool sizing failed");
      rc = reserve_log(h, (unsigned long long)used + used / 4 + 64);
      if (rc) return rc;
    }
    // 3. reverse of cell_finish and clear_post
    if (pre)
      hipLaunchKernelGGL((adj_post_kernel<true>), dim3(grid), dim3(kBlock), 0, s, a, g);
    else
      hipLaunchKernelGGL((adj_post_kernel<false>), dim3(grid), dim3(kBlock), 0, s, a, g);
    // 4. reverse of the loops, wave by wave
    hipLaunchKernelGGL(soa_replay_kernel, dim3(h->log_grid), dim3(kBlock), 0, s, g);
    // 5. reverse of clear_pre, clear_build and cell_setup
    if (pre)
      hipLaunchKernelGGL((adj_pre_kernel<true>), dim3(grid), dim3(kBlock), 0, s, a, g, o);
    else
      hipLaunchKernelGGL((adj_pre_kernel<false>), dim3(grid), dim3(kBlock), 0, s, a, g, o);
    o.bucket = h->bucket;  // the rest: overcast cells
  }
  if (pre)
    hipLaunchKernelGGL((chain_adjoint_kernel<true>), dim3(grid), dim3(kBlock), 0, s, a, o);
  else
    hipLaunchKernelGGL((chain_adjoint_kernel<false>), dim3(grid), dim3(kBlock), 0, s, a, o);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

int mam4x_compute_water_uptake(mam4x_handle h, const mam4x_inputs *in,
                               const mam4x_water_outputs *out, void *stream) {
  if (!h) return MAM4X_ERR_ARG;
  if (!in || !out) return fail(h, MAM4X_ERR_ARG, "inputs or outputs is NULL");
  if (in->ncol < 0 || in->nlev < 0) return fail(h, MAM4X_ERR_ARG, "negative ncol or nlev");
  if (in->layout != MAM4X_LAYOUT_LEV_COL && in->layout != MAM4X_LAYOUT_COL_LEV)
    return fail(h, MAM4X_ERR_ARG, "unknown layout");
  const bool lc = in->layout == MAM4X_LAYOUT_LEV_COL;
  const int fast = lc ? in->ncol : in->nlev, slow = lc ? in->nlev : in->ncol;
  const int ld = in->ldc ? in->ldc : fast;
  if (ld < fast) return fail(h, MAM4X_ERR_ARG, "ldc smaller than the fastest extent");
  if ((long long)ld * slow > 0x7fffffffLL)
    return fail(h, MAM4X_ERR_ARG, "domain too large (ldc x slow extent must fit in int32)");
  WaterArgs a;
  std::memset(&a, 0, sizeof(a));
  for (int r = 0; r < 21; ++r) {
    a.qi[r] = in->state[8 + r];
    if (!a.qi[r]) return fail(h, MAM4X_ERR_ARG, "interstitial mass plane " + std::to_string(8 + r) + " is NULL");
  }
  a.T = in->temperature;
  a.p = in->pressure;
  a.qv = in->vapor_mixing_ratio;
  a.cld = in->cloud_fraction;
  if (!a.T || !a.p || !a.qv || !a.cld)
    return fail(h, MAM4X_ERR_ARG, "temperature, pressure, vapour or cloud fraction plane is NULL");
  for (int m = 0; m < 4; ++m) {
    a.dga[m] = in->dgncur_a[m];
    a.dgw[m] = out->dgncur_awet[m];
    a.qw[m] = out->qaerwat[m];
    a.wd[m] = out->wetdens[m];
    if (!a.dga[m] || !a.dgw[m] || !a.qw[m])
      return fail(h, MAM4X_ERR_ARG, "a dgncur_a, dgncur_awet or qaerwat plane is NULL");
  }
  a.K = h->d_consts;
  a.ncol = in->ncol;
  a.nlev = in->nlev;
  a.ldc = ld;
  a.layout = in->layout;
  const long ncell = (long)in->ncol * in->nlev;
  if (ncell == 0) return MAM4X_OK;
  hipLaunchKernelGGL(water_uptake_kernel, dim3((unsigned)((ncell + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, (hipStream_t)stream, a);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

int mam4x_compute_wet_size(mam4x_handle h, double dt, const mam4x_inputs *in,
                           const mam4x_wetsize_outputs *out, void *stream) {
  if (!h) return MAM4X_ERR_ARG;
  if (!in || !out) return fail(h, MAM4X_ERR_ARG, "inputs or outputs is NULL");
  if (!(dt > 0.0)) return fail(h, MAM4X_ERR_ARG, "dt must be > 0");
  if (in->ncol < 0 || in->nlev < 0) return fail(h, MAM4X_ERR_ARG, "negative ncol or nlev");
  if (in->layout != MAM4X_LAYOUT_LEV_COL && in->layout != MAM4X_LAYOUT_COL_LEV)
    return fail(h, MAM4X_ERR_ARG, "unknown layout");
  const bool lc = in->layout == MAM4X_LAYOUT_LEV_COL;
  const int fast = lc ? in->ncol : in->nlev, slow = lc ? in->nlev : in->ncol;
  const int ld = in->ldc ? in->ldc : fast;
  if (ld < fast) return fail(h, MAM4X_ERR_ARG, "ldc smaller than the fastest extent");
  if ((long long)ld * slow > 0x7fffffffLL)
    return fail(h, MAM4X_ERR_ARG, "domain too large (ldc x slow extent must fit in int32)");
  WetSizeArgs a;
  std::memset(&a, 0, sizeof(a));
  for (int f = 0; f < F_H2SO4; ++f) {
    a.state[f] = in->state[f];
    if (!a.state[f]) return fail(h, MAM4X_ERR_ARG, "state plane " + std::to_string(f) + " is NULL");
  }
  a.T = in->temperature;
  a.p = in->pressure;
  a.qv = in->vapor_mixing_ratio;
  a.cld = in->cloud_fraction;
  if (!a.T || !a.p || !a.qv || !a.cld)
    return fail(h, MAM4X_ERR_ARG, "temperature, pressure, vapour or cloud fraction plane is NULL");
  for (int m = 0; m < 4; ++m) {
    a.dgi[m] = out->dgncur_i[m];
    a.dgc[m] = out->dgncur_c[m];
    a.dgw[m] = out->dgncur_awet[m];
    a.qw[m] = out->qaerwat[m];
    a.wd[m] = out->wetdens[m];
    if (!a.dgi[m] || !a.dgw[m] || !a.qw[m])
      return fail(h, MAM4X_ERR_ARG, "a dgncur_i, dgncur_awet or qaerwat plane is NULL");
  }
  int ntend = 0;
  for (int f = 0; f < NSTATE; ++f) {
    a.tend[f] = out->tendencies[f];
    ntend += a.tend[f] != nullptr;
  }
  if (ntend != 0 && ntend != NSTATE)
    return fail(h, MAM4X_ERR_ARG, "tendencies must be all NULL or all set");
  a.K = h->d_consts;
  a.dt = dt;
  a.ncol = in->ncol;
  a.nlev = in->nlev;
  a.ldc = ld;
  a.layout = in->layout;
  const long ncell = (long)in->ncol * in->nlev;
  if (ncell == 0) return MAM4X_OK;
  hipLaunchKernelGGL(wet_size_kernel, dim3((unsigned)((ncell + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, (hipStream_t)stream, a);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

// closures of the column budgets: fields in state order, mol/kg factors
// (sulfur: 1000 / molecular weight; organic: the gas-aerosol exchange's
// basis, mam4_amicphys.hpp:131-154)
// the mode table gives SO4 and SOA to the accumulation, Aitken and coarse
// modes, interstitial and cloud-borne: six fields and the gas each
constexpr int closure_fields(int aid) {
  int n = 1;
  for (int f = F_Q_I; f < F_H2SO4; ++f) n += mode_spec(field_mode(f), field_spec(f)) == aid;
  return n;
}
static_assert(closure_fields(A_SO4) == kClosureN && closure_fields(A_SOA) == kClosureN,
              "closure field count");

static void budget_closures(BudgetArgs &b) {
  // no allocation: the allocation-free step forms these too
  int ns = 0, no = 0;
  for (int f = F_Q_I; f < F_H2SO4; ++f) {
    const int m = field_mode(f), sp = field_spec(f), aid = mode_spec(m, sp);
    if (aid == A_SO4 && ns < kClosureN - 1) {
      b.clo_f[0][ns] = f;
      b.clo_c[0][ns++] = 1000.0 / ADV_MASS[mass_slot(m, sp)];
    } else if (aid == A_SOA && no < kClosureN - 1) {
      b.clo_f[1][no] = f;
      b.clo_c[1][no++] = 1000.0 * fcvt_aer(A_SOA) / ADV_MASS[mass_slot(m, sp)];
    }
  }
  b.clo_f[0][kClosureN - 1] = F_H2SO4;
  b.clo_c[0][kClosureN - 1] = 1000.0 / 98.078400;
  b.clo_f[1][kClosureN - 1] = F_SOAG;
  b.clo_c[1][kClosureN - 1] = 1000.0 * FCVT_SOAG / 12.011000;
}

// the column budgets' argument block of one domain (in->ncol > 0)
static void budget_args(const mam4x_inputs *in, const mam4x_outputs *out, BudgetArgs &b) {
  std::memset(&b, 0, sizeof(b));
  for (int f = 0; f < NSTATE; ++f) b.tend[f] = out->tendencies[f];
  b.dp = in->hydrostatic_dp;
  b.out = out->column_budgets;
  b.ncol = in->ncol;
  b.nlev = in->nlev;
  const bool collev = in->layout == MAM4X_LAYOUT_COL_LEV;
  b.ldc = in->ldc ? in->ldc : (collev ? in->nlev : in->ncol);
  const int per = collev ? kBudgetBlock / 8 : kBudgetBlock;
  b.ngroup = (in->ncol + per - 1) / per;
  b.collev = collev ? 1 : 0;
  budget_closures(b);
}

static int launch_budget_args(mam4x_handle h, const BudgetArgs &b, hipStream_t s) {
  const unsigned nblk = 8u * kBudgetQuads * (unsigned)((b.ngroup + 7) / 8);
  const dim3 blk(kBudgetWaves * kBudgetBlock);
  if (b.collev)
    hipLaunchKernelGGL(column_budget_kernel<true>, dim3(nblk), blk, 0, s, b);
  else
    hipLaunchKernelGGL(column_budget_kernel<false>, dim3(nblk), blk, 0, s, b);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

static int launch_budgets(mam4x_handle h, const mam4x_inputs *in, const mam4x_outputs *out,
                          hipStream_t s) {
  if (in->ncol == 0) return MAM4X_OK;
  BudgetArgs b;
  budget_args(in, out, b);
  return launch_budget_args(h, b, s);
}

// One step on one domain.  With noalloc the work-queue workspace and its
// counters live in the device's `scratch` array and no phase events are
// recorded: the call allocates nothing and only launches kernels.
static int compute_step(mam4x_handle h, double dt, const mam4x_inputs *in,
                        const mam4x_outputs *out, void *stream, bool noalloc) {
  int rc = check_inputs(h, in);
  if (rc) return rc;
  if (noalloc) {
    if (!h->scratch)
      return fail(h, MAM4X_ERR_HIP, "the device's scratch workspace could not be resolved");
    if ((long)in->ncol * in->nlev > kScratchCells)
      return fail(h, MAM4X_ERR_ARG, "domain of " + std::to_string((long)in->ncol * in->nlev) +
                  " cells exceeds the allocation-free limit of " +
                  std::to_string(kScratchCells) + " cells");
  }
  if (!(dt > 0.0)) return fail(h, MAM4X_ERR_ARG, "dt must be > 0");
  const bool mixed = h->cfg.precision == MAM4X_PREC_MIXED;
  ChainArgs a;
  fill_args(h, in, a);
  rc = fill_outputs(h, out, a);
  if (rc) return rc;
  const bool proc = has_proc(a);
  if (proc && mixed)
    return fail(h, MAM4X_ERR_UNSUPPORTED, "process tendencies need FP64 precision");
  if (out->column_budgets && !in->hydrostatic_dp)
    return fail(h, MAM4X_ERR_ARG, "column budgets need inputs.hydrostatic_dp");
  a.dt = dt;
  hipStream_t s = (hipStream_t)stream;
  const bool pre = has_pre(in);
  const long ncell = (long)in->ncol * in->nlev;
  if (ncell == 0) return out->column_budgets ? launch_budgets(h, in, out, s) : MAM4X_OK;
  const unsigned grid = (unsigned)((ncell + kBlock - 1) / kBlock);
  bool fused = h->cfg.schedule == MAM4X_SCHED_FUSED ||
               (h->cfg.schedule == MAM4X_SCHED_AUTO && ncell < kFusedBelow);
  // ntsubstep > 1: one lane runs a cell's whole step (the split phases hold
  // one sub-step), in the queue's order unless the fused schedule is chosen
  const bool multi = dev::nsub(a.cfg) > 1;
  Work w;
  std::memset(&w, 0, sizeof(w));
  w.qs = queue_shape(h->cfg);
  w.ncell = ncell;
  w.ctl = noalloc ? reinterpret_cast<QueueCtl *>(h->scratch + kScratchCtlOff) : h->ctl;
  w.dtfull = dev::dt_sub(a.cfg, dt);  // QUIRK (mam4_amicphys.hpp:1319): int sub-step
  if (h->stats) w.stats = h->d_stats;
  const OneDomain src{a};
  const bool queue = !fused && a.cfg.do_cond;
  if (queue && noalloc) {
    // W_NPLANE planes of ncell doubles fit the kScratchCells-sized region
    w.ws = reinterpret_cast<double *>(h->scratch);
    w.queue = reinterpret_cast<int *>(h->scratch + kScratchQueueOff);
    w.bucket = h->scratch + kScratchBucketOff;
  } else if (queue) {
    rc = reserve_work(h, ncell, s);
    if (rc) return rc;
    w.ws = h->ws;
    w.bucket = h->bucket;
    w.queue = h->queue;
  }
  hipEvent_t *ev = nullptr;
  if (h->stats && !noalloc) {
    if (h->pool.empty()) {
      std::array<hipEvent_t, 4> q{};
      for (auto &e : q) HIP_TRY(h, hipEventCreate(&e));
      h->pool.push_back(q);
    }
    h->pending.push_back(h->pool.back());
    h->pool.pop_back();
    ev = h->pending.back().data();
    HIP_TRY(h, hipEventRecord(ev[0], s));
  }
  if (queue) {
    // the counters start at zero: cleared by a kernel (no memset node when captured)
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, s, w.ctl);
    if (mixed)
      launch_cells<Mixed>(kPre, pre, grid, s, src, w);
    else
      launch_cells<Fp64>(kPre, pre, grid, s, src, w);
    hipLaunchKernelGGL(queue_scatter_kernel, dim3(grid), dim3(kBlock), 0, s, w);
  }
  if (ev) HIP_TRY(h, hipEventRecord(ev[1], s));
  if (queue && !multi) {
    hipLaunchKernelGGL(soa_queue_kernel, dim3(h->soa_grid), dim3(kBlock), 0, s, w, w.stats);
    hipLaunchKernelGGL(soa_ieee_kernel, dim3(grid), dim3(kBlock), 0, s, w);
  }
  if (ev) HIP_TRY(h, hipEventRecord(ev[2], s));
  if (multi && queue) {
    for (int pass = 0; pass < 2; ++pass) {
      if (mixed)
        launch_order<Mixed>(pre, grid, s, src, w, pass);
      else if (proc)
        launch_order<Fp64Diag>(pre, grid, s, src, w, pass);
      else
        launch_order<Fp64>(pre, grid, s, src, w, pass);
    }
  } else {
    const int kind = fused || multi ? kFused : kPost;
    if (mixed)
      launch_cells<Mixed>(kind, pre, grid, s, src, w, multi);
    else if (proc)  // the same step, also forming the per-process diagnostics
      launch_cells<Fp64Diag>(kind, pre, grid, s, src, w, multi);
    else
      launch_cells<Fp64>(kind, pre, grid, s, src, w, multi);
  }
  HIP_TRY(h, hipGetLastError());
  if (ev) HIP_TRY(h, hipEventRecord(ev[3], s));
  if (out->column_budgets) return launch_budgets(h, in, out, s);
  return MAM4X_OK;
}

int mam4x_compute_tendencies(mam4x_handle h, double t, double dt, const mam4x_inputs *in,
                             const mam4x_outputs *out, void *stream) {
  (void)t;
  if (!h) return MAM4X_ERR_ARG;
  return compute_step(h, dt, in, out, stream, false);
}

int mam4x_compute_tendencies_noalloc(mam4x_handle h, double t, double dt,
                                     const mam4x_inputs *in, const mam4x_outputs *out,
                                     void *stream) {
  (void)t;
  if (!h) return MAM4X_ERR_ARG;
  return compute_step(h, dt, in, out, stream, true);
}

static int tracer_args(mam4x_handle h, const mam4x_tracers *tr, const mam4x_fields *fl,
                       TracerArgs &a) {
  if (!tr || !fl) return fail(h, MAM4X_ERR_ARG, "tracers or fields is NULL");
  if (!tr->state_q) return fail(h, MAM4X_ERR_ARG, "state_q is NULL");
  if (fl->ncol < 0 || fl->nlev < 0) return fail(h, MAM4X_ERR_ARG, "negative ncol or nlev");
  if (fl->layout != MAM4X_LAYOUT_LEV_COL && fl->layout != MAM4X_LAYOUT_COL_LEV)
    return fail(h, MAM4X_ERR_ARG, "unknown layout");
  const bool lc = fl->layout == MAM4X_LAYOUT_LEV_COL;
  const int fast = lc ? fl->ncol : fl->nlev, slow = lc ? fl->nlev : fl->ncol;
  const int ld = fl->ldc ? fl->ldc : fast;
  if (ld < fast) return fail(h, MAM4X_ERR_ARG, "ldc smaller than the fastest extent");
  const long long ls = tr->lev_stride ? tr->lev_stride : MAM4X_PCNST;
  const long long cs = tr->col_stride ? tr->col_stride : (long long)fl->nlev * ls;
  // rows of MAM4X_PCNST tracers, stacked level-major within a column or
  // column-major within a level, must not overlap
  const bool by_col = ls >= MAM4X_PCNST && cs >= (long long)fl->nlev * ls;
  const bool by_lev = cs >= MAM4X_PCNST && ls >= (long long)fl->ncol * cs;
  if (!by_col && !by_lev)
    return fail(h, MAM4X_ERR_ARG, "tracer rows overlap (col_stride / lev_stride too small)");
  std::memset(&a, 0, sizeof(a));
  a.q = tr->state_q;
  a.qqcw = tr->qqcw;
  a.col_stride = (long)cs;
  a.lev_stride = (long)ls;
  for (int f = 0; f < NSTATE; ++f) a.field[f] = fl->field[f];
  a.ncol = fl->ncol;
  a.nlev = fl->nlev;
  a.ldc = ld;
  a.layout = fl->layout;
  return MAM4X_OK;
}

int mam4x_tracers_to_fields(mam4x_handle h, const mam4x_tracers *src, const mam4x_fields *dst,
                            void *stream) {
  if (!h) return MAM4X_ERR_ARG;
  TracerArgs a;
  int rc = tracer_args(h, src, dst, a);
  if (rc) return rc;
  const long ncell = (long)a.ncol * a.nlev;
  if (ncell == 0) return MAM4X_OK;
  hipLaunchKernelGGL(tracers_to_fields_kernel, dim3((unsigned)((ncell + kTile - 1) / kTile)),
                     dim3(kBlock), 0, (hipStream_t)stream, a);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

int mam4x_fields_to_tracers(mam4x_handle h, const mam4x_fields *src, const mam4x_tracers *dst,
                            int mode, void *stream) {
  if (!h) return MAM4X_ERR_ARG;
  if (mode != MAM4X_PACK_STATE && mode != MAM4X_PACK_TENDENCY)
    return fail(h, MAM4X_ERR_ARG, "unknown pack mode");
  TracerArgs a;
  int rc = tracer_args(h, dst, src, a);
  if (rc) return rc;
  a.mode = mode;
  const long ncell = (long)a.ncol * a.nlev;
  if (ncell == 0) return MAM4X_OK;
  hipLaunchKernelGGL(fields_to_tracers_kernel, dim3((unsigned)((ncell + kTile - 1) / kTile)),
                     dim3(kBlock), 0, (hipStream_t)stream, a);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

int mam4x_compute_gas_chemistry(mam4x_handle h, double dt, const mam4x_gas_chem_args *in,
                                void *stream) {
  if (!h) return MAM4X_ERR_ARG;
  if (!in) return fail(h, MAM4X_ERR_ARG, "args is NULL");
  if (!(dt > 0.0)) return fail(h, MAM4X_ERR_ARG, "dt must be > 0");
  if (in->ncol < 0 || in->nlev < 0) return fail(h, MAM4X_ERR_ARG, "negative ncol or nlev");
  if (in->layout != MAM4X_LAYOUT_LEV_COL && in->layout != MAM4X_LAYOUT_COL_LEV)
    return fail(h, MAM4X_ERR_ARG, "unknown layout");
  const bool lc = in->layout == MAM4X_LAYOUT_LEV_COL;
  const int fast = lc ? in->ncol : in->nlev, slow = lc ? in->nlev : in->ncol;
  const int ld = in->ldc ? in->ldc : fast;
  if (ld < fast) return fail(h, MAM4X_ERR_ARG, "ldc smaller than the fastest extent");
  if ((long long)ld * slow > 0x7fffffffLL)
    return fail(h, MAM4X_ERR_ARG, "domain too large (ldc x slow extent must fit in int32)");
  GasChemArgs a;
  std::memset(&a, 0, sizeof(a));
  a.T = in->temperature;
  a.photo = in->photolysis_rate;
  if (!a.T || !a.photo) return fail(h, MAM4X_ERR_ARG, "temperature or photolysis plane is NULL");
  for (int i = 0; i < gc::kExt; ++i) a.ext[i] = in->extfrc[i];
  const int used[5] = {0, 3, 4, 5, 6};
  for (int k = 0; k < 5; ++k) {
    a.inv[k] = in->invariants[used[k]];
    if (!a.inv[k])
      return fail(h, MAM4X_ERR_ARG, "invariant plane " + std::to_string(used[k]) + " is NULL");
  }
  for (int i = 1; i < MAM4X_GAS_PCNST; ++i) {
    a.het[i] = in->het_rates[i];
    a.qq[i] = in->qq[i];
    if (!a.qq[i]) return fail(h, MAM4X_ERR_ARG, "qq plane " + std::to_string(i) + " is NULL");
  }
  a.ncol = in->ncol;
  a.nlev = in->nlev;
  a.ldc = ld;
  a.layout = in->layout;
  a.dt = dt;
  const long ncell = (long)in->ncol * in->nlev;
  if (ncell == 0) return MAM4X_OK;
  hipLaunchKernelGGL(gas_chem_kernel, dim3((unsigned)((ncell + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, (hipStream_t)stream, a);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

// extents, layout and leading dimension of a companion call's planes
static int plane_extents(mam4x_handle h, int ncol, int nlev, int layout, int ldc, int &ld) {
  if (ncol < 0 || nlev < 0) return fail(h, MAM4X_ERR_ARG, "negative ncol or nlev");
  if (layout != MAM4X_LAYOUT_LEV_COL && layout != MAM4X_LAYOUT_COL_LEV)
    return fail(h, MAM4X_ERR_ARG, "unknown layout");
  const bool lc = layout == MAM4X_LAYOUT_LEV_COL;
  const int fast = lc ? ncol : nlev, slow = lc ? nlev : ncol;
  ld = ldc ? ldc : fast;
  if (ld < fast) return fail(h, MAM4X_ERR_ARG, "ldc smaller than the fastest extent");
  if ((long long)ld * slow > 0x7fffffffLL)
    return fail(h, MAM4X_ERR_ARG, "domain too large (ldc x slow extent must fit in int32)");
  return MAM4X_OK;
}

int mam4x_compute_invariants(mam4x_handle h, const mam4x_setinv_args *in, void *stream) {
  if (!h) return MAM4X_ERR_ARG;
  if (!in) return fail(h, MAM4X_ERR_ARG, "args is NULL");
  int ld = 0;
  if (int rc = plane_extents(h, in->ncol, in->nlev, in->layout, in->ldc, ld)) return rc;
  SetinvArgs a;
  std::memset(&a, 0, sizeof(a));
  a.T = in->temperature;
  a.p = in->pressure;
  a.qv = in->vapor_mixing_ratio;
  if (!a.T || !a.p || !a.qv)
    return fail(h, MAM4X_ERR_ARG, "temperature, pressure or vapor plane is NULL");
  for (int i = 0; i < 4; ++i) a.offline[i] = in->cnst_offline[i];
  for (int i = 0; i < MAM4X_NFS; ++i) a.inv[i] = in->invariants[i];
  a.ncol = in->ncol;
  a.nlev = in->nlev;
  a.ldc = ld;
  a.layout = in->layout;
  const long ncell = (long)in->ncol * in->nlev;
  if (ncell == 0) return MAM4X_OK;
  hipLaunchKernelGGL(setinv_kernel, dim3((unsigned)((ncell + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, (hipStream_t)stream, a);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

int mam4x_compute_linoz(mam4x_handle h, double dt, const mam4x_linoz_args *in, void *stream) {
  if (!h) return MAM4X_ERR_ARG;
  if (!in) return fail(h, MAM4X_ERR_ARG, "args is NULL");
  if (!(dt > 0.0)) return fail(h, MAM4X_ERR_ARG, "dt must be > 0");
  int ld = 0;
  if (int rc = plane_extents(h, in->ncol, in->nlev, in->layout, in->ldc, ld)) return rc;
  if (in->o3_lbl < 0 || in->o3_lbl > in->nlev)
    return fail(h, MAM4X_ERR_ARG, "o3_lbl outside [0, nlev]");
  if (in->o3_lbl > 0 && !(in->o3_tau > 0.0))
    return fail(h, MAM4X_ERR_ARG, "o3_tau must be > 0 with a surface sink");
  LinozArgs a;
  std::memset(&a, 0, sizeof(a));
  a.T = in->temperature;
  a.p = in->pressure;
  a.pdel = in->pdel;
  a.o3 = in->o3;
  if (!a.T || !a.p || !a.pdel || !a.o3)
    return fail(h, MAM4X_ERR_ARG, "temperature, pressure, pdel or o3 plane is NULL");
  for (int i = 0; i < lnz::kNClim; ++i) {
    a.clim[i] = in->climatology[i];
    if (!a.clim[i])
      return fail(h, MAM4X_ERR_ARG, "climatology plane " + std::to_string(i) + " is NULL");
  }
  a.sza = in->zenith_angle;
  a.lat = in->latitude;
  a.ltrop = in->ltrop;
  if (!a.sza || !a.lat || !a.ltrop)
    return fail(h, MAM4X_ERR_ARG, "zenith_angle, latitude or ltrop is NULL");
  a.col0 = in->column_above_top;
  a.o3col = in->o3_column;
  a.do3 = in->do3_linoz;
  a.do3_psc = in->do3_linoz_psc;
  a.ss = in->ss_o3;
  a.sink = in->o3_sfcsink;
  a.cf = lnz::Conf{in->chlorine_loading, in->psc_T, in->o3_sfc, in->o3_tau, in->o3_lbl};
  a.ncol = in->ncol;
  a.nlev = in->nlev;
  a.ldc = ld;
  a.layout = in->layout;
  a.dt = dt;
  if (in->ncol == 0) return MAM4X_OK;
  const unsigned grid = (unsigned)((in->ncol + kLinozBlock - 1) / kLinozBlock);
  if (in->nlev <= kLinozLdsLevels)
    hipLaunchKernelGGL(linoz_block_kernel, dim3(grid), dim3(kLinozBlock * kLinozWaves),
                       (size_t)in->nlev * kLinozBlock * sizeof(double), (hipStream_t)stream, a);
  else
    hipLaunchKernelGGL(linoz_kernel, dim3(grid), dim3(kLinozBlock), 0, (hipStream_t)stream, a);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

int mam4x_compute_washout(mam4x_handle h, double dt, const mam4x_sethet_args *in, void *stream) {
  if (!h) return MAM4X_ERR_ARG;
  if (!in) return fail(h, MAM4X_ERR_ARG, "args is NULL");
  if (!(dt > 0.0)) return fail(h, MAM4X_ERR_ARG, "dt must be > 0");
  int ld = 0;
  if (int rc = plane_extents(h, in->ncol, in->nlev, in->layout, in->ldc, ld)) return rc;
  SethetArgs a;
  std::memset(&a, 0, sizeof(a));
  const double *need[10] = {in->pressure, in->zmid,        in->temperature, in->cmfdqr,
                            in->nrain,    in->nevapr,      in->air_density, in->h2o2,
                            in->latitude, in->phis};
  for (int i = 0; i < 10; ++i)
    if (!need[i]) return fail(h, MAM4X_ERR_ARG, "an input plane or column array is NULL");
  a.press = in->pressure;
  a.zmid = in->zmid;
  a.T = in->temperature;
  a.cmfdqr = in->cmfdqr;
  a.nrain = in->nrain;
  a.nevapr = in->nevapr;
  a.inv = in->air_density;
  a.h2o2 = in->h2o2;
  a.lat = in->latitude;
  a.phis = in->phis;
  for (int m = 0; m < MAM4X_GAS_PCNST; ++m) a.het[m] = in->het_rates[m];
  if (!a.het[1] || !a.het[2] || !a.het[3])
    return fail(h, MAM4X_ERR_ARG, "het_rates planes 1, 2 and 3 are required");
  // the drops' mass-transfer coefficient (mo_sethet.hpp:330-331), once on
  // the host
  const double xrm = .189, xum = 748.0, xvv = 6.18e-2, xdg = .112;
  a.xkgm = xdg / xrm * 2.0 +
           xdg / xrm * .6 * std::sqrt(xrm * xum / xvv) * std::pow((xvv / xdg), (1.0 / 3.0));
  a.dt = dt;
  a.ncol = in->ncol;
  a.nlev = in->nlev;
  a.ldc = ld;
  a.layout = in->layout;
  if (in->ncol == 0) return MAM4X_OK;
  hipLaunchKernelGGL(sethet_kernel, dim3((unsigned)((in->ncol + kLinozBlock - 1) / kLinozBlock)),
                     dim3(kLinozBlock), 0, (hipStream_t)stream, a);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

int mam4x_compute_forcing(mam4x_handle h, const mam4x_setext_args *in, void *stream) {
  if (!h) return MAM4X_ERR_ARG;
  if (!in) return fail(h, MAM4X_ERR_ARG, "args is NULL");
  int ld = 0;
  if (int rc = plane_extents(h, in->ncol, in->nlev, in->layout, in->ldc, ld)) return rc;
  SetextArgs a;
  std::memset(&a, 0, sizeof(a));
  bool used[MAM4X_EXTCNT] = {};
  for (int m = 0; m < MAM4X_EXTCNT; ++m) {
    const int nn = in->frc_ndx[m] - 1;
    if (nn < 0 || nn >= MAM4X_EXTCNT)
      return fail(h, MAM4X_ERR_ARG, "frc_ndx outside 1.." + std::to_string(MAM4X_EXTCNT));
    if (used[nn]) return fail(h, MAM4X_ERR_ARG, "two forcings write one extfrc entry");
    used[nn] = true;
    if (in->nsectors[m] < 0 || in->nsectors[m] > MAM4X_MAX_SECTORS)
      return fail(h, MAM4X_ERR_ARG, "nsectors outside 0..4");
    a.nsec[m] = in->nsectors[m];
    a.alt[m] = in->file_alt_data[m] != 0;
    for (int s = 0; s < a.nsec[m]; ++s) {
      a.sec[m][s] = in->sectors[m][s];
      if (!a.sec[m][s]) return fail(h, MAM4X_ERR_ARG, "a sector plane is NULL");
    }
    a.out[m] = in->extfrc[nn];
  }
  a.ncol = in->ncol;
  a.nlev = in->nlev;
  a.ldc = ld;
  a.layout = in->layout;
  const long ncell = (long)in->ncol * in->nlev;
  if (ncell == 0) return MAM4X_OK;
  hipLaunchKernelGGL(setext_kernel, dim3((unsigned)((ncell + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, (hipStream_t)stream, a);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

long long mam4x_photo_work_len(const mam4x_photo_table *tb, int ncol, int nlev, int ldc,
                               int layout) {
  if (!tb || ncol < 0 || nlev < 0) return -1;
  if (layout != MAM4X_LAYOUT_LEV_COL && layout != MAM4X_LAYOUT_COL_LEV) return -1;
  const bool lc = layout == MAM4X_LAYOUT_LEV_COL;
  const long long fast = lc ? ncol : nlev, slow = lc ? nlev : ncol;
  const long long ld = ldc ? ldc : fast;
  if (ld < fast) return -1;
  return 2 * slow * ld;
}

int mam4x_compute_photolysis(mam4x_handle h, const mam4x_photo_table *tb,
                             const mam4x_photo_args *in, void *stream) {
  if (!h) return MAM4X_ERR_ARG;
  if (!tb || !in) return fail(h, MAM4X_ERR_ARG, "table or args is NULL");
  int ld = 0;
  if (int rc = plane_extents(h, in->ncol, in->nlev, in->layout, in->ldc, ld)) return rc;
  if (tb->nw < 1 || tb->numj < 1 || tb->nt < pht::kNt || tb->np_xs < 2 || tb->nump < 2 ||
      tb->numsza < 2 || tb->numcolo3 < 2 || tb->numalb < 2)
    return fail(h, MAM4X_ERR_ARG,
                "table dimensions: nw, numj >= 1, nt >= 201, np_xs, nump, numsza, numcolo3, "
                "numalb >= 2");
  if (tb->lng_index < -1 || tb->lng_index >= tb->numj)
    return fail(h, MAM4X_ERR_ARG, "lng_index outside [-1, numj)");
  const double *tp[14] = {tb->xsqy,  tb->sza,   tb->del_sza, tb->alb,       tb->del_alb,
                          tb->press, tb->del_p, tb->colo3,   tb->o3rat,     tb->del_o3rat,
                          tb->etfphot, tb->prs, tb->dprs,    tb->rsf_tab};
  for (int i = 0; i < 14; ++i)
    if (!tp[i]) return fail(h, MAM4X_ERR_ARG, "a table array is NULL");
  PhotoArgs a;
  std::memset(&a, 0, sizeof(a));
  a.t = pht::Table{tb->nw,    tb->nt,      tb->np_xs,   tb->numj,     tb->nump,
                   tb->numsza, tb->numcolo3, tb->numalb, tb->xsqy,     tb->sza,
                   tb->del_sza, tb->alb,     tb->del_alb, tb->press,    tb->del_p,
                   tb->colo3,  tb->o3rat,   tb->del_o3rat, tb->etfphot, tb->prs,
                   tb->dprs,   tb->rsf_tab, tb->pht_alias_mult, tb->lng_index};
  a.pmid = in->pressure;
  a.pdel = in->pdel;
  a.T = in->temperature;
  a.colo3 = in->o3_column;
  a.lwc = in->liquid_water;
  a.cld = in->cloud_fraction;
  a.zen = in->zenith_angle;
  a.alb = in->surface_albedo;
  a.photo = in->photo;
  if (!a.pmid || !a.pdel || !a.T || !a.colo3 || !a.lwc || !a.cld || !a.zen || !a.alb ||
      !a.photo || !in->work)
    return fail(h, MAM4X_ERR_ARG, "a plane, zenith_angle, surface_albedo or work is NULL");
  const bool lc = in->layout == MAM4X_LAYOUT_LEV_COL;
  const long slow = lc ? in->nlev : in->ncol;
  a.eff = in->work;
  a.mult = in->work + slow * ld;
  a.esfact = in->esfact;
  a.ncol = in->ncol;
  a.nlev = in->nlev;
  a.ldc = ld;
  a.layout = in->layout;
  const long ncell = (long)in->ncol * in->nlev;
  if (ncell == 0 || tb->lng_index < 0) return MAM4X_OK;
  if (const int cpb = photo_cols_per_block(in->nlev))
    hipLaunchKernelGGL(photo_column_block_kernel, dim3((unsigned)((in->ncol + cpb - 1) / cpb)),
                       dim3(256), (size_t)3 * in->nlev * cpb * sizeof(double),
                       (hipStream_t)stream, a, cpb);
  else
    hipLaunchKernelGGL(photo_column_kernel,
                       dim3((unsigned)((in->ncol + kPhotoColBlock - 1) / kPhotoColBlock)),
                       dim3(kPhotoColBlock), 0, (hipStream_t)stream, a);
  hipLaunchKernelGGL(photo_level_kernel, dim3((unsigned)((ncell + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, (hipStream_t)stream, a);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

int mam4x_compute_setsox(mam4x_handle h, double dt, const mam4x_setsox_args *in, void *stream) {
  if (!h) return MAM4X_ERR_ARG;
  if (!in) return fail(h, MAM4X_ERR_ARG, "args is NULL");
  if (!(dt > 0.0)) return fail(h, MAM4X_ERR_ARG, "dt must be > 0");
  if (in->ncol < 0 || in->nlev < 0) return fail(h, MAM4X_ERR_ARG, "negative ncol or nlev");
  if (in->layout != MAM4X_LAYOUT_LEV_COL && in->layout != MAM4X_LAYOUT_COL_LEV)
    return fail(h, MAM4X_ERR_ARG, "unknown layout");
  const bool lc = in->layout == MAM4X_LAYOUT_LEV_COL;
  const int fast = lc ? in->ncol : in->nlev, slow = lc ? in->nlev : in->ncol;
  const int ld = in->ldc ? in->ldc : fast;
  if (ld < fast) return fail(h, MAM4X_ERR_ARG, "ldc smaller than the fastest extent");
  if ((long long)ld * slow > 0x7fffffffLL)
    return fail(h, MAM4X_ERR_ARG, "domain too large (ldc x slow extent must fit in int32)");
  SoxArgs a;
  std::memset(&a, 0, sizeof(a));
  const double *met[6] = {in->pressure, in->temperature, in->liquid_water, in->cloud_fraction,
                          in->droplet_number, in->air_number_density};
  for (int i = 0; i < 6; ++i) {
    if (!met[i]) return fail(h, MAM4X_ERR_ARG, "an atmosphere plane is NULL");
    a.met[i] = met[i];
  }
  for (int i = 0; i < 4; ++i) {
    a.qin[i] = in->qin[i];
    if (!a.qin[i]) return fail(h, MAM4X_ERR_ARG, "qin plane " + std::to_string(i) + " is NULL");
  }
  const int so4[3] = {6, 14, 21}, num[3] = {13, 18, 26};
  for (int k = 0; k < 3; ++k) {
    a.so4c[k] = in->qcw[so4[k]];
    if (!a.so4c[k]) return fail(h, MAM4X_ERR_ARG, "qcw plane " + std::to_string(so4[k]) + " is NULL");
  }
  for (int k = 0; k < 3; ++k) {
    a.numc[k] = in->qcw[num[k]];
    if (!a.numc[k]) return fail(h, MAM4X_ERR_ARG, "qcw plane " + std::to_string(num[k]) + " is NULL");
  }
  for (int i = 0; i < MAM4X_GAS_PCNST; ++i) {
    a.aq[i] = in->dqdt_aqso4[i];
    a.h2[i] = in->dqdt_aqh2so4[i];
  }
  a.ncol = in->ncol;
  a.nlev = in->nlev;
  a.ldc = ld;
  a.layout = in->layout;
  a.dt = dt;
  const long ncell = (long)in->ncol * in->nlev;
  if (ncell == 0) return MAM4X_OK;
  hipLaunchKernelGGL(setsox_kernel, dim3((unsigned)((ncell + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, (hipStream_t)stream, a);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

// A batch owns its argument blocks and its work-queue workspace (SOA loop
// planes, cost buckets, queue, queue counters), allocated once at creation.
// A run captured in a hipGraph keeps these pointers, so nothing a later call
// on the handle does (a larger step growing the handle's own workspace,
// another batch) can move them.
struct mam4x_batch_s {
  mam4x_handle h = nullptr;
  int nscene = 0;
  double dt = 0.0;
  bool pre = false, fused = false, queue = false;
  ChainArgs *d_args = nullptr;
  unsigned *d_blk0 = nullptr;
  unsigned nblock = 0;
  long ncell = 0;  // padded to whole blocks per scene
  double *ws = nullptr;
  unsigned char *bucket = nullptr;
  int *wqueue = nullptr;
  QueueCtl *ctl = nullptr;
  bool counted = false;  // included in h->nbatch
  std::vector<BudgetArgs> budgets;  // scenes that form column budgets
};

int mam4x_batch_destroy(mam4x_batch b) {
  if (!b) return MAM4X_ERR_ARG;
  if (b->counted) --b->h->nbatch;
  if (b->d_args) (void)hipFree(b->d_args);
  if (b->d_blk0) (void)hipFree(b->d_blk0);
  if (b->ws) (void)hipFree(b->ws);
  if (b->bucket) (void)hipFree(b->bucket);
  if (b->wqueue) (void)hipFree(b->wqueue);
  if (b->ctl) (void)hipFree(b->ctl);
  delete b;
  return MAM4X_OK;
}

int mam4x_batch_create(mam4x_handle h, int nscene, const mam4x_inputs *in,
                       const mam4x_outputs *out, double dt, mam4x_batch *batch) {
  if (!h || !batch) return MAM4X_ERR_ARG;
  *batch = nullptr;
  if (nscene <= 0 || !in || !out) return fail(h, MAM4X_ERR_ARG, "empty batch");
  if (!(dt > 0.0)) return fail(h, MAM4X_ERR_ARG, "dt must be > 0");
  std::vector<ChainArgs> args(nscene);
  std::vector<unsigned> blk0(nscene + 1, 0);
  bool pre = false;
  for (int i = 0; i < nscene; ++i) {
    int rc = check_inputs(h, &in[i]);
    if (rc) return rc;
    fill_args(h, &in[i], args[i]);
    rc = fill_outputs(h, &out[i], args[i]);
    if (rc) return rc;
    if (has_proc(args[i]))
      return fail(h, MAM4X_ERR_UNSUPPORTED, "process tendencies need a single-domain call");
    if (out[i].column_budgets && !in[i].hydrostatic_dp)
      return fail(h, MAM4X_ERR_ARG, "column budgets need inputs.hydrostatic_dp (scene " +
                  std::to_string(i) + ")");
    args[i].dt = dt;
    pre = pre || has_pre(&in[i]);
    const long nc = (long)in[i].ncol * in[i].nlev;
    const long nb = (nc + kBlock - 1) / kBlock;
    if ((long)blk0[i] + nb > 0x7fffffffL / kBlock)
      return fail(h, MAM4X_ERR_ARG, "batch too large (cells must fit in int32)");
    blk0[i + 1] = blk0[i] + (unsigned)nb;
  }
  mam4x_batch b = new (std::nothrow) mam4x_batch_s();
  if (!b) return MAM4X_ERR_NOMEM;
  b->h = h;
  b->nscene = nscene;
  b->dt = dt;
  b->pre = pre;
  b->nblock = blk0[nscene];
  b->ncell = (long)b->nblock * kBlock;
  for (int i = 0; i < nscene; ++i)
    if (out[i].column_budgets && in[i].ncol > 0) {
      b->budgets.emplace_back();
      budget_args(&in[i], &out[i], b->budgets.back());
    }
  b->fused = h->cfg.schedule == MAM4X_SCHED_FUSED ||
             (h->cfg.schedule == MAM4X_SCHED_AUTO && b->ncell < kFusedBelow) ||
             h->cfg.ntsubstep > 1;  // the split phases hold one sub-step
  b->queue = !b->fused && args[0].cfg.do_cond != 0 && b->ncell > 0;
  if (hipMalloc(&b->d_args, sizeof(ChainArgs) * nscene) != hipSuccess ||
      hipMalloc(&b->d_blk0, sizeof(unsigned) * (nscene + 1)) != hipSuccess ||
      (b->queue && (hipMalloc(&b->ws, sizeof(double) * W_NPLANE * b->ncell) != hipSuccess ||
                    hipMalloc(&b->bucket, (size_t)b->ncell) != hipSuccess ||
                    hipMalloc(&b->wqueue, sizeof(int) * b->ncell) != hipSuccess ||
                    hipMalloc(&b->ctl, sizeof(QueueCtl)) != hipSuccess))) {
    (void)hipGetLastError();
    mam4x_batch_destroy(b);
    return fail(h, MAM4X_ERR_NOMEM, "batch argument blocks or workspace");
  }
  if (hipMemcpy(b->d_args, args.data(), sizeof(ChainArgs) * nscene, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->d_blk0, blk0.data(), sizeof(unsigned) * (nscene + 1),
                hipMemcpyHostToDevice) != hipSuccess) {
    mam4x_batch_destroy(b);
    return fail(h, MAM4X_ERR_HIP, "batch upload failed");
  }
  b->counted = true;
  ++h->nbatch;
  *batch = b;
  return MAM4X_OK;
}

int mam4x_batch_run(mam4x_batch b, double t, double dt, void *stream) {
  (void)t;
  if (!b) return MAM4X_ERR_ARG;
  mam4x_handle h = b->h;
  if (dt != b->dt) return fail(h, MAM4X_ERR_ARG, "batch planned for another dt");
  hipStream_t s = (hipStream_t)stream;
  if (b->ncell == 0) {  // scenes without levels: their budgets are zeros
    for (const BudgetArgs &ba : b->budgets) {
      const int rc = launch_budget_args(h, ba, s);
      if (rc) return rc;
    }
    return MAM4X_OK;
  }
  const bool mixed = h->cfg.precision == MAM4X_PREC_MIXED;
  const SceneBatch src{b->d_args, b->d_blk0, b->nscene};
  Work w;
  std::memset(&w, 0, sizeof(w));
  w.qs = queue_shape(h->cfg);
  w.ncell = b->ncell;
  w.ctl = b->ctl;
  w.dtfull = (double)(int)(dt / 1);  // QUIRK (mam4_amicphys.hpp:1319): int sub-step
  if (h->stats) w.stats = h->d_stats;
  const bool queue = b->queue;
  const bool fused = b->fused;
  if (queue) {
    w.ws = b->ws;
    w.bucket = b->bucket;
    w.queue = b->wqueue;
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, s, b->ctl);
    if (mixed)
      launch_cells<Mixed>(kPre, b->pre, b->nblock, s, src, w);
    else
      launch_cells<Fp64>(kPre, b->pre, b->nblock, s, src, w);
    hipLaunchKernelGGL(queue_scatter_kernel, dim3(b->nblock), dim3(kBlock), 0, s, w);
    hipLaunchKernelGGL(soa_queue_kernel, dim3(h->soa_grid), dim3(kBlock), 0, s, w, w.stats);
    hipLaunchKernelGGL(soa_ieee_kernel, dim3(b->nblock), dim3(kBlock), 0, s, w);
  }
  const int kind = fused ? kFused : kPost;
  const bool multi = h->cfg.ntsubstep > 1;
  if (mixed)
    launch_cells<Mixed>(kind, b->pre, b->nblock, s, src, w, multi);
  else
    launch_cells<Fp64>(kind, b->pre, b->nblock, s, src, w, multi);
  HIP_TRY(h, hipGetLastError());
  for (const BudgetArgs &ba : b->budgets) {
    const int rc = launch_budget_args(h, ba, s);
    if (rc) return rc;
  }
  return MAM4X_OK;
}

}  // extern "C"
