// HykSort rounds with every step on the device (gs_vrank_sort, HykSort):
// splitter selection, bucket bounds, the virtual exchange and the k-way
// merge of every round are queued without waiting on the host, except one
// check per batch of selection iterations.
//
// Splitter selection restates SelectSplitters / SelectSplittersHelper
// (include/helpers/hyksort_helper.h:335-450) with regular draws in place of
// rand(); the rule, kept bit for bit with the oracle:
//   tol = 0.01 / (k-1), beta = (int)(cbrt(1/tol) * 3), exp_rank[i] =
//   (i+1) * N / k, loop while max_err > (int64)(N * tol), at most 64 times;
//   iteration 1 draws s regular samples run[(j+1)*n/(s+1)] of every rank;
//   later iterations draw, per splitter, c = min(beta, tot) samples of its
//   windows laid end to end in rank order at u_j = (2j+1)*tot / (2c);
//   a sample's global rank is the sum of its lower_bounds in the group's
//   runs; splitter i takes the sample with the largest global rank <=
//   exp_rank[i] (the smallest sample if there is none) when its error is
//   below the previous iteration's max_err, which becomes the largest
//   accepted error; an accepted splitter's window in every rank shrinks to
//   [lower_bound(that sample), lower_bound(the first sample above the
//   target)), keeping a bound when there is no sample on that side.
// All groups of a round select at once: one workgroup per group samples and
// picks, one lane per (sample, rank) ranks.
//
// Exchange and merge: bucket i of group-local rank r goes to group-local
// rank blk*i + r % blk (hyksort_helper.h:558, :584); plan kernels turn the
// bucket bounds into new rank offsets and the destinations' run lists with
// the k-way merge's sample / partition bases (kKwayMax runs per merge, more
// runs over several passes), so the merge runs from device tables with grids
// sized by upper bounds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gs_device.h"
#include "gs_internal.h"

namespace gs {

namespace {

constexpr uint32_t kHykMaxIters = 64;
constexpr uint32_t kHykThreads = 256;
constexpr uint32_t kNoUb = 0xFFFFFFFFu;

struct HykGroup {
  int64_t n, cond, max_err;
  uint32_t done, iters, nsamp, pad;
};

// Per (group, splitter) result of the last pick.
struct HykPick {
  uint32_t lb, ub;     // sample indexes (ub: kNoUb when no sample lies above the target)
  uint32_t accepted;   // 1: the splitter and its windows take this iteration's pick
  uint32_t lb_at_or_below;  // the lb sample's global rank is <= the target
};

template <typename U, typename V, bool LEX>
struct HElem {
  U k;
  V v;
  __device__ bool lt(const HElem& o) const {
    if constexpr (LEX) return k < o.k || (k == o.k && v < o.v);
    return k < o.k;
  }
};

template <typename U, typename V, bool LEX>
__device__ inline HElem<U, V, LEX> load_elem(const U* keys, const V* vals, uint64_t a, const Xform<U>& xf,
                                             const Xform<V>& vxf) {
  HElem<U, V, LEX> e;
  e.k = to_ordered(keys[a], xf);
  e.v = LEX ? to_ordered(vals[a], vxf) : V(0);
  return e;
}

// Elements of run [lo, hi) that precede x.
template <typename U, typename V, bool LEX>
__device__ inline uint64_t run_lower_bound(const U* keys, const V* vals, uint64_t lo, uint64_t hi,
                                           const HElem<U, V, LEX>& x, const Xform<U>& xf, const Xform<V>& vxf) {
  const uint64_t base = lo;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (load_elem<U, V, LEX>(keys, vals, mid, xf, vxf).lt(x))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - base;
}

struct HykDims {
  uint32_t p, n_p, groups, kway, kk, s, beta, cap;
  double tol;
};

// Workspace views (all device pointers).
template <typename U, typename V>
struct HykWs {
  HykGroup* grp;       // [groups]
  uint64_t* wstart;    // [p][kway] window start (element index within the run)
  uint64_t* wend;      // [p][kway]
  uint64_t* cumw;      // [groups][kway][n_p + 1] window size prefix sums
  uint64_t* cnt;       // [groups][kway + 1] sample offsets per splitter
  U* sk;               // [groups][cap] sample images
  V* sv;
  unsigned long long* glb;  // [groups][cap] global ranks
  HykPick* pick;       // [groups][kway]
  U* spl_k;            // [groups][kway] splitter images
  V* spl_v;
  uint32_t* summary;   // [2]: all groups done, largest iteration count
  uint64_t* loc;       // [groups][cap][n_p] lower_bound of every sample in every rank, or null
};

// Block-wide exclusive scan of one value per thread (kHykThreads threads);
// returns the thread's exclusive prefix and the block total in *total.
__device__ inline uint64_t block_excl_scan(uint64_t x, uint64_t* s_tmp, uint64_t* total) {
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint64_t incl = wave_incl_scan64(x);
  if (lane == 63) s_tmp[wave] = incl;
  __syncthreads();
  uint64_t add = 0, tot = 0;
  for (uint32_t w = 0; w < kHykThreads / 64; ++w) {
    if (w < wave) add += s_tmp[w];
    tot += s_tmp[w];
  }
  __syncthreads();
  *total = tot;
  return incl - x + add;
}

__global__ __launch_bounds__(kHykThreads) void hyk_init_kernel(const uint64_t* __restrict__ off, HykDims D,
                                                               HykGroup* __restrict__ grp,
                                                               uint64_t* __restrict__ wstart,
                                                               uint64_t* __restrict__ wend) {
  const uint32_t g = blockIdx.x;
  const uint32_t r0 = g * D.n_p;
  const int64_t n = (int64_t)(off[r0 + D.n_p] - off[r0]);
  if (threadIdx.x == 0) {
    HykGroup s{};
    s.n = n;
    s.cond = (int64_t)((double)n * D.tol);
    s.max_err = n;
    s.done = n == 0 ? 1u : 0u;
    grp[g] = s;
  }
  for (uint32_t x = threadIdx.x; x < D.n_p * D.kway; x += kHykThreads) {
    const uint32_t lr = x / D.kway;
    const uint64_t idx = (uint64_t)(r0 + lr) * D.kway + x % D.kway;
    wstart[idx] = 0;
    wend[idx] = off[r0 + lr + 1] - off[r0 + lr];
  }
}

// Draws one iteration's samples of every group that is not done.
template <typename U, typename V, bool LEX>
__global__ __launch_bounds__(kHykThreads) void hyk_sample_kernel(const U* __restrict__ keys,
                                                                 const V* __restrict__ vals,
                                                                 const uint64_t* __restrict__ off, HykDims D,
                                                                 HykWs<U, V> W, Xform<U> xf, Xform<V> vxf) {
  __shared__ uint64_t s_tmp[kHykThreads / 64];
  const uint32_t g = blockIdx.x, tid = threadIdx.x;
  const uint32_t r0 = g * D.n_p;
  HykGroup& st = W.grp[g];
  const uint32_t done = st.done, it = st.iters;
  if (done) {
    if (tid == 0) st.nsamp = 0;
    return;
  }
  U* sk = W.sk + (uint64_t)g * D.cap;
  V* sv = W.sv + (uint64_t)g * D.cap;
  unsigned long long* glb = W.glb + (uint64_t)g * D.cap;
  uint32_t nsamp = 0;
  if (it == 0) {
    // s regular samples of every non-empty rank, compacted in rank order.
    uint64_t carry = 0;
    for (uint32_t c0 = 0; c0 < D.n_p; c0 += kHykThreads) {
      const uint32_t lr = c0 + tid;
      const uint64_t len = lr < D.n_p ? off[r0 + lr + 1] - off[r0 + lr] : 0;
      const uint64_t cnt = len ? D.s : 0;
      uint64_t tot;
      const uint64_t base = carry + block_excl_scan(cnt, s_tmp, &tot);
      for (uint32_t j = 0; j < cnt; ++j) {
        const uint64_t a = off[r0 + lr] + (uint64_t)(j + 1) * len / (D.s + 1);
        const HElem<U, V, LEX> e = load_elem<U, V, LEX>(keys, vals, a, xf, vxf);
        sk[base + j] = e.k;
        if constexpr (LEX) sv[base + j] = e.v;
        glb[base + j] = 0;
      }
      carry += tot;
    }
    nsamp = (uint32_t)carry;
  } else {
    // Window size prefix sums of every splitter, then c_i = min(beta, tot_i)
    // draws per splitter at u_j = (2j+1) * tot / (2c).
    uint64_t* cumw = W.cumw + (uint64_t)g * D.kway * (D.n_p + 1);
    uint64_t* cnt = W.cnt + (uint64_t)g * (D.kway + 1);
    for (uint32_t i = 0; i < D.kway; ++i) {
      uint64_t carry = 0;
      for (uint32_t c0 = 0; c0 < D.n_p; c0 += kHykThreads) {
        const uint32_t lr = c0 + tid;
        const uint64_t idx = (uint64_t)(r0 + lr) * D.kway + i;
        const uint64_t w = lr < D.n_p ? W.wend[idx] - W.wstart[idx] : 0;
        uint64_t tot;
        const uint64_t ex = carry + block_excl_scan(w, s_tmp, &tot);
        if (lr < D.n_p) cumw[(uint64_t)i * (D.n_p + 1) + lr] = ex;
        carry += tot;
      }
      if (tid == 0) cumw[(uint64_t)i * (D.n_p + 1) + D.n_p] = carry;
    }
    __syncthreads();
    if (tid == 0) {
      uint64_t o = 0;
      for (uint32_t i = 0; i < D.kway; ++i) {
        cnt[i] = o;
        const uint64_t tot = cumw[(uint64_t)i * (D.n_p + 1) + D.n_p];
        o += tot < D.beta ? tot : D.beta;
      }
      cnt[D.kway] = o;
    }
    __syncthreads();
    nsamp = (uint32_t)cnt[D.kway];
    for (uint32_t x = tid; x < nsamp; x += kHykThreads) {
      // splitter of draw x: the last i with cnt[i] <= x
      uint32_t lo = 0, hi = D.kway;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cnt[mid] <= x)
          lo = mid;
        else
          hi = mid;
      }
      const uint32_t i = lo;
      const uint64_t* cw = cumw + (uint64_t)i * (D.n_p + 1);
      const uint64_t tot = cw[D.n_p];
      const uint64_t c = cnt[i + 1] - cnt[i];
      const uint64_t j = x - cnt[i];
      const uint64_t u = (2 * j + 1) * tot / (2 * c);
      // rank of u: the last lr with cw[lr] <= u and a non-empty window
      uint32_t a = 0, b = D.n_p;
      while (b - a > 1) {
        const uint32_t mid = (a + b) >> 1;
        if (cw[mid] <= u)
          a = mid;
        else
          b = mid;
      }
      const uint64_t idx = (uint64_t)(r0 + a) * D.kway + i;
      const uint64_t pos = off[r0 + a] + W.wstart[idx] + (u - cw[a]);
      const HElem<U, V, LEX> e = load_elem<U, V, LEX>(keys, vals, pos, xf, vxf);
      sk[x] = e.k;
      if constexpr (LEX) sv[x] = e.v;
      glb[x] = 0;
    }
  }
  if (tid == 0) st.nsamp = nsamp;
}

// Global rank of every sample: one lane per (sample, rank of the group).
template <typename U, typename V, bool LEX>
__global__ __launch_bounds__(kHykThreads) void hyk_rank_kernel(const U* __restrict__ keys, const V* __restrict__ vals,
                                                               const uint64_t* __restrict__ off, HykDims D,
                                                               HykWs<U, V> W, Xform<U> xf, Xform<V> vxf) {
  const uint32_t g = blockIdx.y;
  const uint64_t t = (uint64_t)blockIdx.x * kHykThreads + threadIdx.x;
  const uint32_t nsamp = W.grp[g].nsamp;
  const uint64_t x = t / D.n_p;
  if (x >= nsamp) return;
  const uint32_t lr = (uint32_t)(t % D.n_p);
  const uint32_t r = g * D.n_p + lr;
  HElem<U, V, LEX> e;
  e.k = W.sk[(uint64_t)g * D.cap + x];
  e.v = LEX ? W.sv[(uint64_t)g * D.cap + x] : V(0);
  const uint64_t lb = run_lower_bound<U, V, LEX>(keys, vals, off[r], off[r + 1], e, xf, vxf);
  if (W.loc) W.loc[((uint64_t)g * D.cap + x) * D.n_p + lr] = lb;
  if (lb) atomicAdd(&W.glb[(uint64_t)g * D.cap + x], (unsigned long long)lb);
}

// Picks every splitter of a group from the ranked samples (one workgroup per
// group) and advances the group's loop state.
template <typename U, typename V>
__global__ __launch_bounds__(kHykThreads) void hyk_pick_kernel(HykDims D, HykWs<U, V> W) {
  __shared__ unsigned long long s_red[3][kHykThreads / 64];
  __shared__ int64_t s_newmax;
  const uint32_t g = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  HykGroup& st = W.grp[g];
  HykPick* pick = W.pick + (uint64_t)g * D.kway;
  if (st.done) {
    for (uint32_t i = tid; i < D.kway; i += kHykThreads) pick[i].accepted = 0;
    return;
  }
  const uint32_t nsamp = st.nsamp;
  if (nsamp == 0) {  // nothing drawn: the loop ends
    for (uint32_t i = tid; i < D.kway; i += kHykThreads) pick[i].accepted = 0;
    __syncthreads();
    if (tid == 0) {
      st.iters += 1;
      st.done = 1;
    }
    return;
  }
  const unsigned long long* glb = W.glb + (uint64_t)g * D.cap;
  const int64_t n = st.n, max_err = st.max_err;
  if (tid == 0) s_newmax = 0;
  // Candidates packed as (global rank << 24 | sample), samples < 2^24.
  constexpr unsigned long long kNone = ~0ull;
  for (uint32_t i = 0; i < D.kway; ++i) {
    const unsigned long long exp_rank = (unsigned long long)(((int64_t)(i + 1) * n) / D.kk);
    unsigned long long best_lb = 0, has_lb = 0, best_ub = kNone, best_min = kNone;
    for (uint32_t x = tid; x < nsamp; x += kHykThreads) {
      const unsigned long long q = (glb[x] << 24) | x;
      if (glb[x] <= exp_rank) {
        best_lb = q > best_lb ? q : best_lb;
        has_lb = 1;
      } else {
        best_ub = q < best_ub ? q : best_ub;
      }
      best_min = q < best_min ? q : best_min;
    }
    // lb: max over lanes (has_lb folded into the top bit), ub / min: min.
    unsigned long long a = has_lb ? (best_lb | (1ull << 63)) : 0ull, b = best_ub, c = best_min;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long a2 = __shfl_xor(a, o, 64), b2 = __shfl_xor(b, o, 64), c2 = __shfl_xor(c, o, 64);
      a = a2 > a ? a2 : a;
      b = b2 < b ? b2 : b;
      c = c2 < c ? c2 : c;
    }
    if (lane == 0) {
      s_red[0][wave] = a;
      s_red[1][wave] = b;
      s_red[2][wave] = c;
    }
    __syncthreads();
    if (tid == 0) {
      for (uint32_t w = 1; w < kHykThreads / 64; ++w) {
        a = s_red[0][w] > a ? s_red[0][w] : a;
        b = s_red[1][w] < b ? s_red[1][w] : b;
        c = s_red[2][w] < c ? s_red[2][w] : c;
      }
      const bool found = (a >> 63) != 0;
      const unsigned long long lbq = found ? (a & ~(1ull << 63)) : c;
      const uint32_t lb = (uint32_t)(lbq & 0xFFFFFFull);
      const uint32_t ub = b == kNone ? kNoUb : (uint32_t)(b & 0xFFFFFFull);
      const int64_t lbr = (int64_t)glb[lb];
      const int64_t err = lbr > (int64_t)exp_rank ? lbr - (int64_t)exp_rank : (int64_t)exp_rank - lbr;
      HykPick pk{lb, ub, 0u, lbr <= (int64_t)exp_rank ? 1u : 0u};
      if (err < max_err) {
        pk.accepted = 1;
        W.spl_k[(uint64_t)g * D.kway + i] = W.sk[(uint64_t)g * D.cap + lb];
        W.spl_v[(uint64_t)g * D.kway + i] = W.sv[(uint64_t)g * D.cap + lb];
        if (err > s_newmax) s_newmax = err;
      }
      pick[i] = pk;
    }
    __syncthreads();
  }
  if (tid == 0) {
    st.max_err = s_newmax;
    st.iters += 1;
    if (s_newmax <= st.cond || st.iters >= kHykMaxIters) st.done = 1;
  }
}

// Window updates of the accepted splitters: one lane per (group, splitter,
// rank), two binary searches over the rank's run.
template <typename U, typename V, bool LEX>
__global__ __launch_bounds__(kHykThreads) void hyk_window_kernel(const U* __restrict__ keys,
                                                                 const V* __restrict__ vals,
                                                                 const uint64_t* __restrict__ off, HykDims D,
                                                                 HykWs<U, V> W, Xform<U> xf, Xform<V> vxf) {
  const uint64_t t = (uint64_t)blockIdx.x * kHykThreads + threadIdx.x;
  if (t >= (uint64_t)D.groups * D.kway * D.n_p) return;
  const uint32_t lr = (uint32_t)(t % D.n_p);
  const uint32_t i = (uint32_t)((t / D.n_p) % D.kway);
  const uint32_t g = (uint32_t)(t / ((uint64_t)D.n_p * D.kway));
  const HykPick pk = W.pick[(uint64_t)g * D.kway + i];
  if (!pk.accepted) return;
  const uint32_t r = g * D.n_p + lr;
  const uint64_t idx = (uint64_t)r * D.kway + i;
  const uint64_t base = (uint64_t)g * D.cap;
  if (W.loc) {  // the rank kernel kept every sample's bound in every rank
    if (pk.lb_at_or_below) W.wstart[idx] = W.loc[(base + pk.lb) * D.n_p + lr];
    if (pk.ub != kNoUb) W.wend[idx] = W.loc[(base + pk.ub) * D.n_p + lr];
    return;
  }
  if (pk.lb_at_or_below) {
    HElem<U, V, LEX> e;
    e.k = W.sk[base + pk.lb];
    e.v = LEX ? W.sv[base + pk.lb] : V(0);
    W.wstart[idx] = run_lower_bound<U, V, LEX>(keys, vals, off[r], off[r + 1], e, xf, vxf);
  }
  if (pk.ub != kNoUb) {
    HElem<U, V, LEX> e;
    e.k = W.sk[base + pk.ub];
    e.v = LEX ? W.sv[base + pk.ub] : V(0);
    W.wend[idx] = run_lower_bound<U, V, LEX>(keys, vals, off[r], off[r + 1], e, xf, vxf);
  }
}

__global__ __launch_bounds__(kHykThreads) void hyk_summary_kernel(HykDims D, const HykGroup* __restrict__ grp,
                                                                  uint32_t* __restrict__ summary) {
  __shared__ uint32_t s_all, s_max;
  if (threadIdx.x == 0) {
    s_all = 1;
    s_max = 0;
  }
  __syncthreads();
  for (uint32_t g = threadIdx.x; g < D.groups; g += kHykThreads) {
    if (!grp[g].done) atomicAnd(&s_all, 0u);
    atomicMax(&s_max, grp[g].iters);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    summary[0] = s_all;
    summary[1] = s_max;
  }
}

// Splitters as raw key / payload values, and each rank's first splitter.
template <typename U, typename V>
__global__ __launch_bounds__(kHykThreads) void hyk_finish_kernel(HykDims D, HykWs<U, V> W, Xform<U> xf,
                                                                 Xform<V> vxf, U* __restrict__ out_k,
                                                                 V* __restrict__ out_v,
                                                                 uint32_t* __restrict__ spl_first) {
  const uint32_t t = blockIdx.x * kHykThreads + threadIdx.x;
  if (t < D.groups * D.kway) {
    out_k[t] = from_ordered(W.spl_k[t], xf);
    out_v[t] = from_ordered(W.spl_v[t], vxf);
  }
  if (t < D.p) spl_first[t] = (t / D.n_p) * D.kway;
}

// New rank sizes of a round and their offsets: destination d (group-local
// dl) receives bucket i = dl / blk of the group's ranks dl % blk + q * blk,
// q < kk (hyksort_helper.h:558, :584).  One workgroup.
__global__ __launch_bounds__(kHykThreads) void hyk_sizes_kernel(const uint64_t* __restrict__ bnd, HykDims D,
                                                                uint32_t blk, uint64_t* __restrict__ off_next) {
  __shared__ uint64_t s_tmp[kHykThreads / 64];
  const uint32_t tid = threadIdx.x;
  const uint32_t per = D.kway + 2;  // bounds per rank: start, kway cuts, end
  uint64_t carry = 0;
  for (uint32_t c0 = 0; c0 < D.p; c0 += kHykThreads) {
    const uint32_t d = c0 + tid;
    uint64_t size = 0;
    if (d < D.p) {
      const uint32_t g = d / D.n_p, dl = d % D.n_p, i = dl / blk;
      for (uint32_t q = 0; q < D.kk; ++q) {
        const uint32_t r = g * D.n_p + dl % blk + q * blk;
        size += bnd[(uint64_t)r * per + i + 1] - bnd[(uint64_t)r * per + i];
      }
    }
    uint64_t tot;
    const uint64_t ex = carry + block_excl_scan(size, s_tmp, &tot);
    if (d < D.p) off_next[d] = ex;
    carry += tot;
  }
  if (tid == 0) off_next[D.p] = carry;
}

// K-way merge tables of one pass of a round.  A destination's kk received
// runs (in source order) are merged kKwayMax at a time: pass t merges groups
// of kKwayMax runs of span = kKwayMax^t original runs each (pass 0: the
// buckets where their senders hold them; later passes: the previous pass's
// results, consecutive in the destination's output range), `chunks` entries
// per destination.  Empty runs are dropped; an entry with one run copies it.
// One workgroup scans the entries' samples, boundaries and partitions.
__global__ __launch_bounds__(kHykThreads) void hyk_plan_kernel(const uint64_t* __restrict__ bnd, HykDims D,
                                                               uint32_t blk, uint32_t span, uint32_t chunks,
                                                               const uint64_t* __restrict__ off_next,
                                                               KwayDest* __restrict__ dests,
                                                               KwayTotals* __restrict__ tot) {
  __shared__ uint64_t s_tmp[kHykThreads / 64];
  const uint32_t tid = threadIdx.x;
  const uint32_t per = D.kway + 2;
  const uint32_t nent = D.p * chunks;
  uint64_t c_samp = 0, c_bnd = 0, c_part = 0;
  for (uint32_t c0 = 0; c0 < nent; c0 += kHykThreads) {
    const uint32_t e = c0 + tid;
    KwayDest K{};
    uint64_t ns = 0;
    if (e < nent) {
      const uint32_t d = e / chunks, c = e % chunks;
      const uint32_t g = d / D.n_p, dl = d % D.n_p, i = dl / blk;
      auto bucket = [&](uint32_t q, uint64_t* at, uint64_t* len) {
        const uint32_t r = g * D.n_p + dl % blk + q * blk;
        *at = bnd[(uint64_t)r * per + i];
        *len = bnd[(uint64_t)r * per + i + 1] - *at;
      };
      // Output position of original run q within the destination.
      uint64_t pos = off_next[d];
      const uint32_t q0 = c * kKwayMax * span;
      for (uint32_t q = 0; q < q0 && q < D.kk; ++q) {
        uint64_t at, len;
        bucket(q, &at, &len);
        pos += len;
      }
      K.out_off = pos;
      uint32_t m = 0;
      for (uint32_t j = 0; j < kKwayMax; ++j) {
        const uint32_t qa = q0 + j * span;
        if (qa >= D.kk) break;
        uint64_t at = 0, len = 0;
        if (span == 1) {
          bucket(qa, &at, &len);
        } else {
          at = pos;
          for (uint32_t q = qa; q < qa + span && q < D.kk; ++q) {
            uint64_t a2, l2;
            bucket(q, &a2, &l2);
            len += l2;
          }
        }
        pos += len;
        if (len) {
          K.run_off[m] = at;
          K.run_len[m] = len;
          ++m;
          ns += len / kKwaySampleStride;
        }
      }
      K.m = m;
      K.nparts = (uint32_t)(ns / kKwaySamplesPerPart) + 1;
    }
    uint64_t t_samp, t_bnd, t_part;
    const uint64_t e_samp = block_excl_scan(ns, s_tmp, &t_samp);
    const uint64_t e_bnd = block_excl_scan(e < nent ? K.nparts + 1 : 0, s_tmp, &t_bnd);
    const uint64_t e_part = block_excl_scan(e < nent ? K.nparts : 0, s_tmp, &t_part);
    if (e < nent) {
      K.samp_base = (uint32_t)(c_samp + e_samp);
      K.bnd_base = (uint32_t)(c_bnd + e_bnd);
      K.part_base = (uint32_t)(c_part + e_part);
      dests[e] = K;
    }
    c_samp += t_samp;
    c_bnd += t_bnd;
    c_part += t_part;
  }
  if (tid == 0) {
    KwayTotals t;
    t.ndests = nent;
    t.nsamples = (uint32_t)c_samp;
    t.nbounds = (uint32_t)c_bnd;
    t.nparts = (uint32_t)c_part;
    *tot = t;
  }
}

// Merge passes of a round with kk received runs per destination, and the
// entries per destination of pass t.
uint32_t merge_passes(uint32_t kk) {
  uint32_t t = 0;
  for (uint64_t span = 1; span < kk; span *= kKwayMax) ++t;
  return t;
}
uint32_t pass_chunks(uint32_t kk, uint32_t t) {
  uint64_t span = 1;
  for (uint32_t i = 0; i <= t; ++i) span *= kKwayMax;
  return (uint32_t)((kk + span - 1) / span);
}

// Layout of the device rounds' workspace.
struct HykLayout {
  size_t off[2], grp, wstart, wend, cumw, cnt, sk, sv, glb, pick, spl_k, spl_v, summary, spl_raw_k, spl_raw_v,
      spl_first, bnd, iters, dests, totals, kscratch, loc, total;
  size_t report;  // bytes of the leading region read back at the end (summary .. spl_raw_v)
  uint32_t rounds, max_cap, max_kway, max_chunks;
  bool keep_loc;  // room for every sample's bound in every rank
};

struct RoundShape {
  uint32_t n_p, kk, blk, groups, kway, beta, cap;
  double tol;
};

std::vector<RoundShape> round_shapes(uint32_t p, uint32_t k_way, uint32_t s) {
  std::vector<RoundShape> out;
  uint32_t n_p = p;
  while (n_p > 1) {
    RoundShape R{};
    R.n_p = n_p;
    R.kk = std::min(k_way, n_p);
    R.blk = n_p / R.kk;
    R.groups = p / n_p;
    R.kway = R.kk - 1;
    R.tol = 1e-2 / R.kway;
    R.beta = (uint32_t)(std::pow(1.0 / R.tol, 1.0 / 3.0) * 3.0);
    R.cap = std::max<uint32_t>(n_p * s, R.kway * R.beta);
    out.push_back(R);
    n_p = R.blk;
  }
  return out;
}

HykLayout hyk_layout(uint32_t p, uint32_t k_way, uint32_t s, uint64_t n, int kb, int vb) {
  HykLayout L{};
  const std::vector<RoundShape> shapes = round_shapes(p, k_way, s);
  L.rounds = (uint32_t)shapes.size();
  uint64_t sel_cap = 0, sel_cumw = 0, sel_cnt = 0, loc_words = 0;
  L.max_chunks = 1;
  for (const RoundShape& R : shapes) {
    L.max_chunks = std::max(L.max_chunks, pass_chunks(R.kk, 0));
    L.max_kway = std::max(L.max_kway, R.kway);
    L.max_cap = std::max(L.max_cap, R.cap);
    sel_cap = std::max<uint64_t>(sel_cap, (uint64_t)R.groups * R.cap);
    loc_words = std::max<uint64_t>(loc_words, (uint64_t)R.groups * R.cap * R.n_p);
    sel_cumw = std::max<uint64_t>(sel_cumw, (uint64_t)R.groups * R.kway * (R.n_p + 1));
    sel_cnt = std::max<uint64_t>(sel_cnt, (uint64_t)R.groups * (R.kway + 1));
  }
  const int vbb = vb ? vb : 4;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align_up(o + (bytes ? bytes : 8), 256);
    return at;
  };
  // What the host reads back comes first, in one region.
  L.summary = take(16);
  L.off[0] = take((size_t)(p + 1) * 8);
  L.off[1] = take((size_t)(p + 1) * 8);
  L.bnd = take((size_t)L.rounds * p * (L.max_kway + 2) * 8);
  L.spl_raw_k = take((size_t)L.rounds * p * L.max_kway * kb);
  L.spl_raw_v = take((size_t)L.rounds * p * L.max_kway * vbb);
  L.report = o;
  L.grp = take((size_t)p * sizeof(HykGroup));
  L.wstart = take((size_t)p * L.max_kway * 8);
  L.wend = take((size_t)p * L.max_kway * 8);
  L.cumw = take(sel_cumw * 8);
  L.cnt = take(sel_cnt * 8);
  L.sk = take(sel_cap * kb);
  L.sv = take(sel_cap * vbb);
  L.glb = take(sel_cap * 8);
  L.pick = take((size_t)p * L.max_kway * sizeof(HykPick));
  L.spl_k = take((size_t)p * L.max_kway * kb);
  L.spl_v = take((size_t)p * L.max_kway * vbb);
  L.spl_first = take((size_t)p * 4);
  L.iters = take((size_t)L.rounds * 4);
  L.dests = take((size_t)p * L.max_chunks * sizeof(KwayDest));
  L.totals = take(sizeof(KwayTotals));
  L.kscratch = take(kway_scratch_bytes(kway_upper_bounds(n, p * L.max_chunks), kb, vb));
  // Every sample's lower_bound in every rank of its group (groups * cap *
  // n_p = p * cap words), kept when it takes at most 256 MiB: the window
  // updates then read them instead of searching the runs again.
  L.keep_loc = loc_words * 8 <= (256ull << 20);
  L.loc = L.keep_loc ? take(loc_words * 8) : 0;
  L.total = o;
  return L;
}

template <int VB>
struct HValT {
  using T = uint32_t;
};
template <>
struct HValT<8> {
  using T = uint64_t;
};

template <typename U, int VB, bool LEX>
hipError_t hyk_rounds_t(HykRun* run, const uint64_t* h_off, char* ws, const HykLayout& L, hipStream_t stream,
                        LaunchTimer* timer, HykResult* res) {
  using V = typename HValT<VB>::T;
  const Xform<U> xf{(U)run->kx.float_mask, (U)run->kx.sign_flip, (U)run->kx.desc_mask};
  const Xform<V> vxf = LEX ? Xform<V>{(V)run->vx.float_mask, (V)run->vx.sign_flip, (V)run->vx.desc_mask}
                           : Xform<V>{0, 0, 0};
  const uint32_t p = run->p;
  uint64_t* off[2] = {reinterpret_cast<uint64_t*>(ws + L.off[0]), reinterpret_cast<uint64_t*>(ws + L.off[1])};
  int oc = 0;
  // the rank offsets go up through the pinned buffer (at their own offset
  // in the report region, which only the final read-back overwrites)
  std::memcpy(reinterpret_cast<char*>(run->host) + L.off[0], h_off, (size_t)(p + 1) * 8);
  hipError_t e = hipMemcpyAsync(off[0], reinterpret_cast<char*>(run->host) + L.off[0], (size_t)(p + 1) * 8,
                                hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) return e;
  HykWs<U, V> W;
  W.grp = reinterpret_cast<HykGroup*>(ws + L.grp);
  W.wstart = reinterpret_cast<uint64_t*>(ws + L.wstart);
  W.wend = reinterpret_cast<uint64_t*>(ws + L.wend);
  W.cumw = reinterpret_cast<uint64_t*>(ws + L.cumw);
  W.cnt = reinterpret_cast<uint64_t*>(ws + L.cnt);
  W.sk = reinterpret_cast<U*>(ws + L.sk);
  W.sv = reinterpret_cast<V*>(ws + L.sv);
  W.glb = reinterpret_cast<unsigned long long*>(ws + L.glb);
  W.pick = reinterpret_cast<HykPick*>(ws + L.pick);
  W.spl_k = reinterpret_cast<U*>(ws + L.spl_k);
  W.spl_v = reinterpret_cast<V*>(ws + L.spl_v);
  W.summary = reinterpret_cast<uint32_t*>(ws + L.summary);
  W.loc = L.keep_loc ? reinterpret_cast<uint64_t*>(ws + L.loc) : nullptr;
  uint32_t* spl_first = reinterpret_cast<uint32_t*>(ws + L.spl_first);
  KwayDest* dests = reinterpret_cast<KwayDest*>(ws + L.dests);
  KwayTotals* totals = reinterpret_cast<KwayTotals*>(ws + L.totals);
  const std::vector<RoundShape> shapes = round_shapes(p, run->k_way, run->s);
  res->rounds = (uint32_t)shapes.size();
  res->iterations.assign(shapes.size(), 0);
  res->kway.clear();
  res->n_p.clear();
  for (const RoundShape& R : shapes) {
    res->kway.push_back(R.kway);
    res->n_p.push_back(R.n_p);
  }
  const U* keys[2] = {reinterpret_cast<const U*>(run->kbuf[0]), reinterpret_cast<const U*>(run->kbuf[1])};
  const V* vals[2] = {reinterpret_cast<const V*>(run->vbuf[0]), reinterpret_cast<const V*>(run->vbuf[1])};
  for (size_t ri = 0; ri < shapes.size(); ++ri) {
    const RoundShape& R = shapes[ri];
    HykDims D{p, R.n_p, R.groups, R.kway, R.kk, run->s, R.beta, R.cap, R.tol};
    const U* kc = keys[run->cur];
    const V* vc = vals[run->cur];
    // Splitter selection: batches of iterations, one host check per batch.
    if (timer) timer->begin("hyk_select", 0, stream);
    hipLaunchKernelGGL(hyk_init_kernel, dim3(R.groups), dim3(kHykThreads), 0, stream, off[oc], D, W.grp, W.wstart,
                       W.wend);
    const uint32_t rank_blocks = (uint32_t)(((uint64_t)R.cap * R.n_p + kHykThreads - 1) / kHykThreads);
    const uint32_t win_blocks =
        (uint32_t)(((uint64_t)R.groups * R.kway * R.n_p + kHykThreads - 1) / kHykThreads);
    uint32_t summary[2] = {0, 0};
    for (uint32_t launched = 0; launched < kHykMaxIters;) {
      // Uniform keys converge in about 3 iterations; later batches are
      // shorter so that a finished loop wastes few launches.
      const uint32_t batch = launched == 0 ? 3u : 2u;
      for (uint32_t b = 0; b < batch; ++b) {
        hipLaunchKernelGGL((hyk_sample_kernel<U, V, LEX>), dim3(R.groups), dim3(kHykThreads), 0, stream, kc, vc,
                           off[oc], D, W, xf, vxf);
        hipLaunchKernelGGL((hyk_rank_kernel<U, V, LEX>), dim3(rank_blocks, R.groups), dim3(kHykThreads), 0, stream,
                           kc, vc, off[oc], D, W, xf, vxf);
        hipLaunchKernelGGL((hyk_pick_kernel<U, V>), dim3(R.groups), dim3(kHykThreads), 0, stream, D, W);
        hipLaunchKernelGGL((hyk_window_kernel<U, V, LEX>), dim3(win_blocks), dim3(kHykThreads), 0, stream, kc, vc,
                           off[oc], D, W, xf, vxf);
      }
      launched += batch;
      hipLaunchKernelGGL(hyk_summary_kernel, dim3(1), dim3(kHykThreads), 0, stream, D, W.grp, W.summary);
      e = hipMemcpyAsync(run->host, W.summary, sizeof(summary), hipMemcpyDeviceToHost, stream);
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
      if (e != hipSuccess) return e;
      std::memcpy(summary, run->host, sizeof(summary));
      if (summary[0]) break;
    }
    res->iterations[ri] = summary[1];
    U* spl_raw_k = reinterpret_cast<U*>(ws + L.spl_raw_k) + (size_t)ri * p * L.max_kway;
    V* spl_raw_v = reinterpret_cast<V*>(ws + L.spl_raw_v) + (size_t)ri * p * L.max_kway;
    hipLaunchKernelGGL((hyk_finish_kernel<U, V>), dim3((std::max(R.groups * R.kway, p) + kHykThreads - 1) / kHykThreads),
                       dim3(kHykThreads), 0, stream, D, W, xf, vxf, spl_raw_k, spl_raw_v, spl_first);
    if (timer) timer->end(stream);
    // Bucket bounds of every run (lower_bound of the group's splitters).
    uint64_t* bnd = reinterpret_cast<uint64_t*>(ws + L.bnd) + (size_t)ri * p * (L.max_kway + 2);
    if (timer) timer->begin("bucket_bounds", 0, stream);
    e = bucket_bounds((int)sizeof(U), kc, off[oc], p, spl_raw_k, spl_first, R.kway, 0, run->kx, bnd, stream,
                      LEX ? VB : 0, LEX ? vc : nullptr, LEX ? spl_raw_v : nullptr, LEX ? &run->vx : nullptr);
    if (timer) timer->end(stream);
    if (e != hipSuccess) return e;
    // Exchange: new rank offsets, then the merge passes (each into the
    // other buffer).
    hipLaunchKernelGGL(hyk_sizes_kernel, dim3(1), dim3(kHykThreads), 0, stream, bnd, D, R.blk, off[oc ^ 1]);
    const uint32_t passes = merge_passes(R.kk);
    uint32_t span = 1;
    for (uint32_t t = 0; t < passes; ++t, span *= kKwayMax) {
      const uint32_t chunks = pass_chunks(R.kk, t);
      hipLaunchKernelGGL(hyk_plan_kernel, dim3(1), dim3(kHykThreads), 0, stream, bnd, D, R.blk, span, chunks,
                         off[oc ^ 1], dests, totals);
      const KwaySizes ub = kway_upper_bounds(run->n, p * chunks);
      const KwayDevice dev{dests, totals, ws + L.kscratch};
      const int c2 = run->cur;
      e = kway_merge((int)sizeof(U), VB, run->kbuf[c2], VB ? run->vbuf[c2] : nullptr, run->kbuf[c2 ^ 1],
                     VB ? run->vbuf[c2 ^ 1] : nullptr, ub, dev, run->kx, LEX ? &run->vx : nullptr, stream, timer,
                     run->n);
      if (e != hipSuccess) return e;
      run->cur ^= 1;
    }
    oc ^= 1;
  }
  // Results: final offsets, every round's bounds and splitters.
  res->final_off.assign(p + 1, 0);
  res->bounds.assign((size_t)shapes.size() * p * (L.max_kway + 2), 0);
  res->spl_k.assign((size_t)shapes.size() * p * L.max_kway * sizeof(U), 0);
  res->spl_v.assign((size_t)shapes.size() * p * L.max_kway * sizeof(V), 0);
  res->bnd_stride = L.max_kway + 2;
  res->spl_stride = L.max_kway;
  e = hipMemcpyAsync(run->host, ws, L.report, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return e;
  const char* h = reinterpret_cast<const char*>(run->host);
  std::memcpy(res->final_off.data(), h + L.off[oc], res->final_off.size() * 8);
  std::memcpy(res->bounds.data(), h + L.bnd, res->bounds.size() * 8);
  std::memcpy(res->spl_k.data(), h + L.spl_raw_k, res->spl_k.size());
  if (LEX) std::memcpy(res->spl_v.data(), h + L.spl_raw_v, res->spl_v.size());
  return hipSuccess;
}

}  // namespace

size_t hyk_device_workspace_bytes(uint32_t p, uint32_t k_way, uint32_t s, uint64_t n, int kb, int vb) {
  return hyk_layout(p, k_way, s, n, kb, vb).total;
}

size_t hyk_host_bytes(uint32_t p, uint32_t k_way, uint32_t s, int kb, int vb) {
  return hyk_layout(p, k_way, s, 0, kb, vb).report;
}

bool hyk_device_supported(uint32_t p, uint32_t k_way, uint32_t s) {
  if (p < 2 || k_way < 2 || s == 0) return false;
  for (const RoundShape& R : round_shapes(p, k_way, s))
    if (R.groups > 65535 || R.cap >= (1u << 24) || (uint64_t)R.cap * R.n_p >= (1ull << 31)) return false;
  return true;
}

hipError_t hyk_rounds_device(HykRun* run, const uint64_t* h_off, void* workspace, size_t workspace_bytes,
                             hipStream_t stream, LaunchTimer* timer, HykResult* res) {
  const HykLayout L = hyk_layout(run->p, run->k_way, run->s, run->n, run->kb, run->vb);
  if (workspace_bytes < L.total || !hyk_device_supported(run->p, run->k_way, run->s)) return hipErrorInvalidValue;
  char* ws = reinterpret_cast<char*>(workspace);
#define GS_HYK_CASE(UT, VBV)                                                                  \
  return run->lex ? hyk_rounds_t<UT, VBV, true>(run, h_off, ws, L, stream, timer, res)       \
                  : hyk_rounds_t<UT, VBV, false>(run, h_off, ws, L, stream, timer, res)
  if (run->kb == 4) {
    if (run->vb == 0) return hyk_rounds_t<uint32_t, 0, false>(run, h_off, ws, L, stream, timer, res);
    if (run->vb == 4) GS_HYK_CASE(uint32_t, 4);
    if (run->vb == 8) GS_HYK_CASE(uint32_t, 8);
  } else if (run->kb == 8) {
    if (run->vb == 0) return hyk_rounds_t<uint64_t, 0, false>(run, h_off, ws, L, stream, timer, res);
    if (run->vb == 4) GS_HYK_CASE(uint64_t, 4);
    if (run->vb == 8) GS_HYK_CASE(uint64_t, 8);
  }
#undef GS_HYK_CASE
  return hipErrorInvalidValue;
}

}  // namespace gs
