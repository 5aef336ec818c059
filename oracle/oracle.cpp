// ORACLE -- test infrastructure only.
//
// CPU restatement of the reference's algorithms on the sort path, used as the
// checker of the HIP kernels by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py.  Nothing in the product links, loads or calls
// this library.
//
// Restated (citations are files of monotone-RK/gpusort):
//   omp_merge          OmpUtils::Merge          include/omp/omp_utils.h:54-145
//   omp_merge_sort     OmpUtils::MergeSort      include/omp/omp_utils.h:147-207
//   omp_scan           OmpUtils::Scan           include/omp/omp_utils.h:209-245
//   lifted comparator  ConvertComp              include/common.h:156-164
//   KV CPU local sort  HykSortHelper::SortByKey include/helpers/hyksort_helper.h:266-291
//   KV GPU local sort  thrust::sort_by_key      src/cuda_utils.cu:477 (stable by key)
//   samplesort rule    SampleSortHelper::SelectSplitters / TransferData
//                      include/helpers/samplesort_helper.h:390-417, :504-565
//   hyksort rounds     HykSortHelper::SortAmongProcesses / TransferData / Merge
//                      include/helpers/hyksort_helper.h:453-536, :546-633, :307-332
//   hyperquicksort     HyperQuickSortHelper::SelectSplitters / TransferData / Merge
//                      include/helpers/hyperquicksort_helper.h:204-259, :348-382, :183-202
//   rebalance          ParUtils::KeepNoProcessHasEmptyData include/par/par_utils.h:103-175
//
// The reference picks HykSort and HyperQuickSort splitters from rand()
// samples seeded by rank (hyksort_helper.h:382-396, hyperquicksort_helper.h:
// 215-226), which nothing but the reference's own run could reproduce.  The
// vrank_* functions restate the round structure with a deterministic sample
// rule instead (regular samples, the same rule as the GPU path): the rule is
// the GPU path's own, the structure (groups, partners, lower/upper bound
// buckets, merge of received runs, rebalance) is the reference's.
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <tuple>
#include <utility>
#include <vector>

namespace orc {

static int g_threads = 0;  // 0: omp_get_max_threads()

static int max_threads() { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

// ---------------------------------------------------------------- OmpUtils::Merge
template <typename T, typename Comp>
void omp_merge(Comp comp, const T* a_first, const T* a_last, const T* b_first, const T* b_last, T* out) {
  const int p = max_threads();
  const int64_t n1 = a_last - a_first, n2 = b_last - b_first;
  if (n1 == 0 && n2 == 0) return;
  if (n1 == 0 || n2 == 0) {
    const T* a = (n1 == 0) ? b_first : a_first;
    const int64_t n = (n1 == 0) ? n2 : n1;
#pragma omp parallel for num_threads(p)
    for (int i = 0; i < p; i++) {
      const int64_t i1 = (i * n) / p, i2 = ((i + 1) * n) / p;
      std::copy(a + i1, a + i2, out + i1);
    }
    return;
  }
  const int ns = 10;  // samples per thread and input
  std::vector<T> split(p * ns * 2);
  std::vector<int64_t> split_size(p * ns * 2);
#pragma omp parallel for num_threads(p)
  for (int i = 0; i < p; i++) {
    for (int j = 0; j < ns; j++) {
      int idx = i * ns + j;
      int64_t i1 = ((int64_t)idx * n1) / (p * ns);
      split[idx] = a_first[i1];
      split_size[idx] = i1 + (std::lower_bound(b_first, b_last, split[idx], comp) - b_first);
      i1 = ((int64_t)idx * n2) / (p * ns);
      idx += p * ns;
      split[idx] = b_first[i1];
      split_size[idx] = i1 + (std::lower_bound(a_first, a_last, split[idx], comp) - a_first);
    }
  }
  std::vector<int64_t> sa(p + 1), sb(p + 1);
  sa[0] = sb[0] = 0;
  sa[p] = n1;
  sb[p] = n2;
#pragma omp parallel for num_threads(p)
  for (int i = 1; i < p; i++) {
    const int64_t req = ((int64_t)i * (n1 + n2)) / p;
    int j = (int)(std::lower_bound(&split_size[0], &split_size[0] + p * ns, req) - &split_size[0]);
    if (j >= p * ns) j = p * ns - 1;
    T s1 = split[j];
    int64_t s1_size = split_size[j];
    j = p * ns + (int)(std::lower_bound(&split_size[0] + p * ns, &split_size[0] + 2 * p * ns, req) -
                       (&split_size[0] + p * ns));
    if (j >= 2 * p * ns) j = 2 * p * ns - 1;
    if (std::llabs(split_size[j] - req) < std::llabs(s1_size - req)) {
      s1 = split[j];
      s1_size = split_size[j];
    }
    sa[i] = std::lower_bound(a_first, a_last, s1, comp) - a_first;
    sb[i] = std::lower_bound(b_first, b_last, s1, comp) - b_first;
  }
#pragma omp parallel for num_threads(p)
  for (int i = 0; i < p; i++) {
    std::merge(a_first + sa[i], a_first + sa[i + 1], b_first + sb[i], b_first + sb[i + 1], out + sa[i] + sb[i], comp);
  }
}

// ---------------------------------------------------------------- OmpUtils::MergeSort
template <typename T, typename Comp>
void omp_merge_sort(Comp comp, T* first, int64_t n) {
  const int p = max_threads();
  if (n < 2 * p || p == 1) {
    std::sort(first, first + n, comp);
    return;
  }
  std::vector<int64_t> split(p + 1);
  split[p] = n;
  for (int id = 0; id < p; id++) split[id] = ((int64_t)id * n) / p;
#pragma omp parallel for num_threads(p)
  for (int id = 0; id < p; id++) std::sort(first + split[id], first + split[id + 1], comp);
  std::vector<T> b(n);
  T* a = first;
  T* bb = b.data();
  for (int j = 1; j < p; j *= 2) {
    for (int i = 0; i < p; i += 2 * j) {
      if (i + j < p) {
        omp_merge(comp, a + split[i], a + split[i + j], a + split[i + j], a + split[(i + 2 * j <= p) ? i + 2 * j : p],
                  bb + split[i]);
      } else {
        std::copy(a + split[i], a + split[p], bb + split[i]);
      }
    }
    std::swap(a, bb);
  }
  if (a != first) std::copy(a, a + n, first);
}

// ---------------------------------------------------------------- OmpUtils::Scan
// b[0] is preset by the caller; b[i] = b[i-1] + a[i-1].
template <typename T>
void omp_scan(int64_t cnt, const T* a, T* b) {
  const int p = max_threads();
  if (cnt < 100 * (int64_t)p) {
    for (int64_t i = 1; i < cnt; i++) b[i] = b[i - 1] + a[i - 1];
    return;
  }
  const int64_t step = cnt / p;
#pragma omp parallel for num_threads(p)
  for (int i = 0; i < p; i++) {
    const int64_t s = i * step;
    const int64_t e = (i == p - 1) ? cnt : s + step;
    if (i != 0) b[s] = 0;
    for (int64_t j = s + 1; j < e; j++) b[j] = b[j - 1] + a[j - 1];
  }
  std::vector<T> sum(p);
  sum[0] = 0;
  for (int i = 1; i < p; i++) sum[i] = sum[i - 1] + b[i * step - 1] + a[i * step - 1];
#pragma omp parallel for num_threads(p)
  for (int i = 1; i < p; i++) {
    const int64_t s = i * step;
    const int64_t e = (i == p - 1) ? cnt : s + step;
    for (int64_t j = s; j < e; j++) b[j] += sum[i];
  }
}

// ---------------------------------------------------------------- element views
// An element is a key alone or a (key, value) pair ordered by the lifted
// comparator: std::less<pair> / std::greater<pair> (lexicographic).
template <typename K, typename V, bool KV>
struct Elem;
template <typename K, typename V>
struct Elem<K, V, false> {
  using T = K;
  static T make(const K* k, const V*, int64_t i) { return k[i]; }
  static void store(const T& e, K* k, V*, int64_t i) { k[i] = e; }
  static const K& key(const T& e) { return e; }
};
template <typename K, typename V>
struct Elem<K, V, true> {
  using T = std::pair<K, V>;
  static T make(const K* k, const V* v, int64_t i) { return T(k[i], v[i]); }
  static void store(const T& e, K* k, V* v, int64_t i) {
    k[i] = e.first;
    v[i] = e.second;
  }
  static const K& key(const T& e) { return e.first; }
};

// ---------------------------------------------------------------- KeepNoProcessHasEmptyData
// Final per-rank sizes after the rebalance of empty ranks; data keeps its
// global order, only rank boundaries move.  `nelems` is the gathered size
// vector and is not updated inside the loop, as in the reference.
static void keep_no_empty(std::vector<int64_t>& sizes) {
  const int n_p = (int)sizes.size();
  const std::vector<int64_t> nelems = sizes;
  for (int i = 0; i < n_p; i++) {
    if (nelems[i] == 0) {
      int partner = i + 1;
      int empty_cnt = 1;
      bool use_prev = false;
      if (i == n_p - 1) {
        partner = n_p - 2;
        use_prev = true;
      } else {
        while (partner < n_p && nelems[partner] == 0) {
          partner++;
          empty_cnt++;
        }
        if (partner == n_p) {
          partner = i - 1;
          use_prev = true;
        }
      }
      if (partner < 0) break;
      const int64_t size_sync = nelems[partner] / (empty_cnt + 1);
      const int last = use_prev ? n_p : partner;
      for (int j = i; j < last; j++) sizes[j] = size_sync;
      sizes[partner] -= (int64_t)empty_cnt * size_sync;
      if (use_prev) i = n_p;
      if (i < n_p - 1) i = partner + 1;
    }
  }
}

// ---------------------------------------------------------------- virtual ranks
// Selection iterations of the last run_vrank on this thread.
static thread_local int g_last_splitter_iters = 0;

struct Report {
  std::vector<int64_t> bucket_sizes;  // per round, per source rank, per bucket
  std::vector<int64_t> final_sizes;
  int rounds = 0;
  int splitter_iters = 0;  // HykSort: selection iterations, summed over rounds
  double max_err = 0.0;
};

template <typename T, typename Comp>
static std::vector<T> merge_runs(const std::vector<std::vector<T>>& runs, Comp comp) {
  // Pairwise adjacent merges (ties from the earlier run), as the device
  // merge tree does.
  std::vector<std::vector<T>> cur = runs;
  if (cur.empty()) return {};
  while (cur.size() > 1) {
    std::vector<std::vector<T>> nxt;
    for (size_t i = 0; i < cur.size(); i += 2) {
      if (i + 1 == cur.size()) {
        nxt.push_back(std::move(cur[i]));
        continue;
      }
      std::vector<T> m(cur[i].size() + cur[i + 1].size());
      std::merge(cur[i].begin(), cur[i].end(), cur[i + 1].begin(), cur[i + 1].end(), m.begin(), comp);
      nxt.push_back(std::move(m));
    }
    cur.swap(nxt);
  }
  return cur[0];
}

template <typename T, typename Comp>
static std::vector<T> regular_samples(const std::vector<T>& run, int s) {
  std::vector<T> out;
  const int64_t n = (int64_t)run.size();
  if (n == 0) return out;
  for (int j = 0; j < s; j++) out.push_back(run[(int64_t)(j + 1) * n / (s + 1)]);
  return out;
}

// HykSort splitter selection: SelectSplitters / SelectSplittersHelper
// (hyksort_helper.h:335-450) with the random draws replaced by regular ones.
// Kept from the reference: tol = 0.01 / (k-1) and beta = (int)(cbrt(1/tol)
// * 3) (:340-344); targets exp_rank[i] = (i+1) * N / k (:348-349); the loop
// while max_err > (int64)(N * tol) (:351-356); global ranks as sums of
// per-rank lower_bounds of every sample, over the samples sorted (:403-423);
// for each splitter the sample just below its target, lb = upper_bound(
// glb_rank, exp) - 1 clamped to 0, and ub = upper_bound (:427-448), accepted
// only when its error is below the previous iteration's max_err, which then
// becomes the largest accepted error (:431, :446-449); an accepted
// splitter's window in every rank becomes [lower_bound(lb), lower_bound(ub))
// (:433-443).
// Deterministic draws: the first iteration takes s regular samples
// run[(j+1)*n/(s+1)] of every rank's whole run.  Later iterations take, for
// each splitter, c = min(beta, tot) samples of its windows laid end to end
// in rank order (tot elements): positions u_j = (2j+1)*tot / (2c), j < c,
// each mapped to its rank and offset in that rank's window.  (Regular draws
// at the same fractions of every rank's window would pick nearly equal
// values from ranks of alike data; laid end to end the ranks' fractions
// differ.)  Windows only shrink: where the reference sets start to 0 (no
// sample at or below the target) or end to the run end (no sample above
// it), the window keeps that bound.  An iteration that draws no sample ends
// the loop, and so does the 64th.  Returns the number of iterations run.
constexpr int kHykMaxIters = 64;

template <typename T, typename Comp>
static int hyk_select_splitters(const std::vector<std::vector<T>>& R, int g0, int n_p, int kk, int s, Comp comp,
                                std::vector<T>* out) {
  const int kway = kk - 1;
  int64_t N = 0;
  for (int r = g0; r < g0 + n_p; r++) N += (int64_t)R[r].size();
  const double tol = 1e-2 / kway;
  const int beta = (int)(std::pow(1.0 / tol, 1.0 / 3.0) * 3.0);
  std::vector<int64_t> exp_rank(kway);
  for (int i = 0; i < kway; i++) exp_rank[i] = ((int64_t)(i + 1) * N) / kk;
  std::vector<std::vector<int64_t>> start(n_p, std::vector<int64_t>(kway, 0)), end(n_p, std::vector<int64_t>(kway));
  for (int lr = 0; lr < n_p; lr++)
    for (int i = 0; i < kway; i++) end[lr][i] = (int64_t)R[g0 + lr].size();
  out->assign(kway, T());
  int64_t max_err = N;
  const int64_t cond = (int64_t)(N * tol);
  int it = 0;
  while (max_err > cond && it < kHykMaxIters) {
    std::vector<T> S;
    if (it == 0) {
      for (int lr = 0; lr < n_p; lr++) {
        auto smp = regular_samples<T, Comp>(R[g0 + lr], s);
        S.insert(S.end(), smp.begin(), smp.end());
      }
    } else {
      for (int i = 0; i < kway; i++) {
        int64_t tot = 0;
        for (int lr = 0; lr < n_p; lr++) tot += end[lr][i] - start[lr][i];
        if (tot == 0) continue;
        const int64_t c = std::min<int64_t>(beta, tot);
        int lr = 0;
        int64_t cum = 0;  // window elements of the ranks before lr
        for (int64_t j = 0; j < c; j++) {
          const int64_t u = (2 * j + 1) * tot / (2 * c);
          while (u >= cum + end[lr][i] - start[lr][i]) {
            cum += end[lr][i] - start[lr][i];
            lr++;
          }
          S.push_back(R[g0 + lr][start[lr][i] + u - cum]);
        }
      }
    }
    ++it;
    if (S.empty()) break;
    std::sort(S.begin(), S.end(), comp);
    const int64_t m = (int64_t)S.size();
    std::vector<std::vector<int64_t>> loc_rank(n_p, std::vector<int64_t>(m));
    std::vector<int64_t> glb(m, 0);
    for (int lr = 0; lr < n_p; lr++) {
      const auto& run = R[g0 + lr];
      for (int64_t x = 0; x < m; x++) {
        loc_rank[lr][x] = std::lower_bound(run.begin(), run.end(), S[x], comp) - run.begin();
        glb[x] += loc_rank[lr][x];
      }
    }
    int64_t new_max = 0;
    for (int i = 0; i < kway; i++) {
      const int64_t ub = std::upper_bound(glb.begin(), glb.end(), exp_rank[i]) - glb.begin();
      const int64_t lb = ub - 1 < 0 ? 0 : ub - 1;
      const int64_t err = std::llabs(glb[lb] - exp_rank[i]);
      if (err < max_err) {
        for (int lr = 0; lr < n_p; lr++) {
          if (glb[lb] <= exp_rank[i]) start[lr][i] = loc_rank[lr][lb];
          if (ub < m) end[lr][i] = loc_rank[lr][ub];
        }
        (*out)[i] = S[lb];
        if (new_max < err) new_max = err;
      }
    }
    max_err = new_max;
  }
  return it;
}

// HykSort rounds (hyksort_helper.h:453-536): splitters by
// hyk_select_splitters over the group; buckets by lower_bound
// (hyksort_helper.h:563-568); bucket i of group-local rank r goes to
// group-local rank blk_size*i + r % blk_size (:558, :584); received runs are
// merged in source-rank order; groups split by blk_id = r / blk_size (:466,
// :515).  rep.splitter_iters adds up, over the rounds, the largest
// selection iteration count of the round's groups.
template <typename T, typename Comp>
static void vrank_hyksort(std::vector<std::vector<T>>& R, int k_way, int s, Comp comp, Report& rep,
                          std::vector<T>& spl_out, int64_t total) {
  const int p = (int)R.size();
  int n_p = p;
  while (n_p > 1) {
    const int kk = std::min(k_way, n_p);
    const int blk = n_p / kk;
    std::vector<std::vector<T>> NR(p);
    int round_iters = 0;
    for (int g0 = 0; g0 < p; g0 += n_p) {
      std::vector<T> spl;
      int64_t gsz = 0;
      for (int r = g0; r < g0 + n_p; r++) gsz += (int64_t)R[r].size();
      if (gsz > 0) round_iters = std::max(round_iters, hyk_select_splitters(R, g0, n_p, kk, s, comp, &spl));
      int64_t gtot = 0;
      for (int r = g0; r < g0 + n_p; r++) gtot += (int64_t)R[r].size();
      if ((int)spl.size() != kk - 1) {
        // Empty group: nothing moves; splitters undefined.
        for (int r = g0; r < g0 + n_p; r++)
          for (int i = 0; i < kk; i++) rep.bucket_sizes.push_back(0);
        continue;
      }
      spl_out.insert(spl_out.end(), spl.begin(), spl.end());
      // Quality: |global rank of splitter i - (i+1)*N_group/k| / N_total.
      for (int i = 0; i < kk - 1; i++) {
        int64_t grank = 0;
        for (int r = g0; r < g0 + n_p; r++)
          grank += std::lower_bound(R[r].begin(), R[r].end(), spl[i], comp) - R[r].begin();
        const double err = std::fabs((double)grank - (double)(i + 1) * gtot / kk) / (double)std::max<int64_t>(total, 1);
        rep.max_err = std::max(rep.max_err, err);
      }
      std::vector<std::vector<int64_t>> bnd(n_p, std::vector<int64_t>(kk + 1));
      for (int lr = 0; lr < n_p; lr++) {
        const auto& run = R[g0 + lr];
        bnd[lr][0] = 0;
        bnd[lr][kk] = (int64_t)run.size();
        for (int i = 0; i < kk - 1; i++) bnd[lr][i + 1] = std::lower_bound(run.begin(), run.end(), spl[i], comp) - run.begin();
        for (int i = 0; i < kk; i++) rep.bucket_sizes.push_back(bnd[lr][i + 1] - bnd[lr][i]);
      }
      for (int dl = 0; dl < n_p; dl++) {
        const int i = dl / blk;
        std::vector<std::vector<T>> runs;
        for (int lr = dl % blk; lr < n_p; lr += blk) {
          const auto& run = R[g0 + lr];
          runs.emplace_back(run.begin() + bnd[lr][i], run.begin() + bnd[lr][i + 1]);
        }
        NR[g0 + dl] = merge_runs(runs, comp);
      }
    }
    R.swap(NR);
    n_p = blk;
    rep.rounds++;
    rep.splitter_iters += round_iters;
  }
}

// SampleSort (samplesort_helper.h:420-501): p-1 regular samples per rank at
// arr[(i+1)*n/p]; samples sorted; ascending: splitter r = S[(r+1)(p-1)-1]
// (largest sample of each sorted block but the last), descending:
// S[(r+1)(p-1)] (first sample of each block but the first); buckets by
// upper_bound; bucket i to rank i; received buffer sorted.
template <typename T, typename Comp>
static void vrank_samplesort(std::vector<std::vector<T>>& R, Comp comp, bool increasing, Report& rep,
                             std::vector<T>& spl_out) {
  const int p = (int)R.size();
  if (p == 1) return;
  std::vector<T> S;
  for (int r = 0; r < p; r++) {
    const int64_t n = (int64_t)R[r].size();
    for (int i = 0; i < p - 1; i++) S.push_back(R[r][(int64_t)(i + 1) * n / p]);
  }
  std::sort(S.begin(), S.end(), comp);
  std::vector<T> spl;
  for (int r = 0; r < p - 1; r++) spl.push_back(increasing ? S[(int64_t)(r + 1) * (p - 1) - 1] : S[(int64_t)(r + 1) * (p - 1)]);
  spl_out.insert(spl_out.end(), spl.begin(), spl.end());
  std::vector<std::vector<int64_t>> bnd(p, std::vector<int64_t>(p + 1));
  for (int r = 0; r < p; r++) {
    const auto& run = R[r];
    bnd[r][0] = 0;
    bnd[r][p] = (int64_t)run.size();
    for (int i = 0; i < p - 1; i++) bnd[r][i + 1] = std::upper_bound(run.begin(), run.end(), spl[i], comp) - run.begin();
    for (int i = 0; i < p; i++) rep.bucket_sizes.push_back(bnd[r][i + 1] - bnd[r][i]);
  }
  std::vector<std::vector<T>> NR(p);
  for (int d = 0; d < p; d++) {
    std::vector<std::vector<T>> runs;
    for (int r = 0; r < p; r++) runs.emplace_back(R[r].begin() + bnd[r][d], R[r].begin() + bnd[r][d + 1]);
    NR[d] = merge_runs(runs, comp);
  }
  R.swap(NR);
  rep.rounds = 1;
}

// SampleSort's small-input path (samplesort_helper.h:429-452): below 5 p^2
// keys in total the ranks run a bitonic merge-split network instead of
// sampling.  Restated from BitonicSort (:192-235), BitonicSortBinary
// (:236-258), BitonicSortBinaryHelper (:260-283), BitonicMerge (:170-190),
// Merge (:285-303) and MergeLists (:305-387); communicator splits as
// ParUtilsHelper::SplitCommBinary / SplitCommBinaryNoFlip
// (src/par_utils_helper.cc:237-329).  `comm` lists the global ranks of a
// communicator in its rank order.  Every round of merge-splits reads the
// lists as they were before the round, as the pairwise exchanges do.
static int prev_pow2(unsigned v) {
  v--;
  v |= v >> 1;
  v |= v >> 2;
  v |= v >> 4;
  v |= v >> 8;
  v |= v >> 16;
  v++;
  return (int)(v >> 1);
}

// MergeLists: list_a (own) merged with list_b (the partner's); keeps the
// low or high part by the reference's loop conditions, element operators
// and std::max / std::min as written there.
template <typename T>
static std::vector<T> merge_lists_ref(const std::vector<T>& list_a, const std::vector<T>& list_b, bool keep_low,
                                      bool inc) {
  const T low = inc ? std::max(list_a[0], list_b[0]) : std::max(list_a.back(), list_b.back());
  const T high = inc ? std::min(list_a.back(), list_b.back()) : std::min(list_a[0], list_b[0]);
  const int64_t n = (int64_t)(list_a.size() + list_b.size());
  std::vector<T> s(n);
  size_t i1 = 0, i2 = 0;
  for (int64_t i = 0; i < n; i++) {
    const bool take_a = i1 < list_a.size() &&
                        (i2 >= list_b.size() || (inc ? list_a[i1] <= list_b[i2] : list_a[i1] > list_b[i2]));
    s[i] = take_a ? list_a[i1++] : list_b[i2++];
  }
  const int64_t half = n / 2;
  if (keep_low) {
    if (inc) {
      int64_t ii = 0;
      while (ii < n && (s[ii] < low || ii < half) && s[ii] <= high) ii++;
      return std::vector<T>(s.begin(), s.begin() + ii);
    }
    int64_t ii = n - 1;
    while (ii >= 0 && ((s[ii] <= high && ii >= half) || s[ii] < low)) ii--;
    return std::vector<T>(s.begin() + ii + 1, s.end());
  }
  if (inc) {
    int64_t ii = n - 1;
    while (ii >= 0 && ((ii >= half && s[ii] >= low) || s[ii] > high)) ii--;
    return std::vector<T>(s.begin() + ii + 1, s.end());
  }
  int64_t ii = 0;
  while (ii < n && ((ii < half && s[ii] >= low) || s[ii] > high)) ii++;
  return std::vector<T>(s.begin(), s.begin() + ii);
}

template <typename T>
struct BitonicNet {
  std::vector<std::vector<T>>& R;
  bool inc;
  int rounds = 0;

  // One round: (rank, partner, keep_low) for every participating rank.
  void round(const std::vector<std::tuple<int, int, bool>>& ops) {
    std::vector<std::vector<T>> next;
    for (const auto& op : ops) next.push_back(merge_lists_ref(R[std::get<0>(op)], R[std::get<1>(op)], std::get<2>(op), inc));
    for (size_t i = 0; i < ops.size(); i++) R[std::get<0>(ops[i])] = std::move(next[i]);
    rounds++;
  }
  void sort_binary(const std::vector<int>& comm) {
    const int npes = (int)comm.size();
    unsigned and_bit = 2;
    for (int pss = 2; pss <= npes; pss *= 2, and_bit <<= 1) {
      int dim = 0;
      for (int x = pss; x > 1; x >>= 1) dim++;
      int eor_bit = 1 << (dim - 1);
      for (int stage = 0; stage < dim; stage++) {
        std::vector<std::tuple<int, int, bool>> ops;
        for (int r = 0; r < npes; r++) {
          const bool incr = ((unsigned)r & and_bit) == 0 ? inc : !inc;
          const int partner = r ^ eor_bit;
          const bool cond = incr ? r < partner : r > partner;
          ops.emplace_back(comm[r], comm[partner], cond);
        }
        round(ops);
        eor_bit >>= 1;
      }
    }
  }
  void sort(const std::vector<int>& comm) {
    const int npes = (int)comm.size();
    if (npes == 1) return;
    if ((npes & (npes - 1)) == 0) {
      sort_binary(comm);
      return;
    }
    const int splitter = prev_pow2((unsigned)npes);
    std::vector<int> A(comm.begin(), comm.begin() + splitter), Bflip;
    for (int j = 0; j < npes - splitter; j++) Bflip.push_back(comm[npes - 1 - j]);
    sort_binary(A);
    sort(Bflip);
    const int num_left = prev_pow2((unsigned)npes), num_right = npes - num_left;
    std::vector<std::tuple<int, int, bool>> ops;
    for (int r = num_left - num_right; r < num_left; r++) ops.emplace_back(comm[r], comm[r + num_right], inc);
    for (int r = num_left; r < npes; r++) ops.emplace_back(comm[r], comm[r - num_right], !inc);
    round(ops);
    std::vector<int> B(comm.begin() + splitter, comm.end());
    sort_binary(A);
    sort(B);
  }
};

// HyperQuickSort (hyperquicksort_helper.h:261-346) with the pivot chosen among
// s regular samples per rank (gathered in rank order): the first sample whose
// global lower_bound rank is closest to N_group/2 (:251-256).  Ranks
// <= (n_p-1)/2 keep the low side of lower_bound(pivot) (:366-372), partner =
// r + cmp_p0 - new_p0 (:358-362), kept run merged first with the received run
// (:197-198), groups halved (:325).
template <typename T, typename Comp>
static void vrank_hyperquicksort(std::vector<std::vector<T>>& R, int s, Comp comp, Report& rep,
                                 std::vector<T>& spl_out) {
  const int p = (int)R.size();
  int n_p = p;
  while (n_p > 1) {
    std::vector<std::vector<T>> NR(p);
    for (int g0 = 0; g0 < p; g0 += n_p) {
      int64_t tot = 0;
      for (int r = g0; r < g0 + n_p; r++) tot += (int64_t)R[r].size();
      if (tot == 0) {
        for (int r = g0; r < g0 + n_p; r++) {
          NR[r] = R[r];
          rep.bucket_sizes.push_back(0);
          rep.bucket_sizes.push_back(0);
        }
        continue;
      }
      std::vector<T> S;
      for (int r = g0; r < g0 + n_p; r++) {
        auto smp = regular_samples<T, Comp>(R[r], s);
        S.insert(S.end(), smp.begin(), smp.end());
      }
      size_t best = 0;
      int64_t best_rank = -1;
      for (size_t x = 0; x < S.size(); x++) {
        int64_t gr = 0;
        for (int r = g0; r < g0 + n_p; r++) gr += std::lower_bound(R[r].begin(), R[r].end(), S[x], comp) - R[r].begin();
        if (x == 0 || std::llabs(gr - tot / 2) < std::llabs(best_rank - tot / 2)) {
          best = x;
          best_rank = gr;
        }
      }
      const T pivot = S[best];
      spl_out.push_back(pivot);
      const int split_id = (n_p - 1) / 2;
      std::vector<int64_t> sidx(n_p);
      for (int lr = 0; lr < n_p; lr++) {
        const auto& run = R[g0 + lr];
        sidx[lr] = std::lower_bound(run.begin(), run.end(), pivot, comp) - run.begin();
        rep.bucket_sizes.push_back(sidx[lr]);
        rep.bucket_sizes.push_back((int64_t)run.size() - sidx[lr]);
      }
      for (int lr = 0; lr < n_p; lr++) {
        const bool low = lr <= split_id;
        const int new_p0 = low ? 0 : split_id + 1;
        const int cmp_p0 = low ? split_id + 1 : 0;
        int partner = lr + cmp_p0 - new_p0;
        if (partner >= n_p) partner = n_p - 1;
        const auto& mine = R[g0 + lr];
        const auto& theirs = R[g0 + partner];
        std::vector<T> kept = low ? std::vector<T>(mine.begin(), mine.begin() + sidx[lr])
                                  : std::vector<T>(mine.begin() + sidx[lr], mine.end());
        // The partner sends the side it does not keep (:369-370): a high
        // partner its low side, a low partner its high side.
        std::vector<T> recv = low ? std::vector<T>(theirs.begin(), theirs.begin() + sidx[partner])
                                  : std::vector<T>(theirs.begin() + sidx[partner], theirs.end());
        std::vector<T> m(kept.size() + recv.size());
        std::merge(kept.begin(), kept.end(), recv.begin(), recv.end(), m.begin(), comp);
        NR[g0 + lr] = std::move(m);
      }
    }
    R.swap(NR);
    n_p = n_p / 2;
    rep.rounds++;
  }
}

// Full per-rank pipeline on the CPU.  Local sort: keys-only -> the unique
// sorted run; with payload -> lexicographic (key, value) under the lifted
// comparator, the reference CPU path's order.
template <typename K, typename V, bool KV, typename Comp>
static int run_vrank(int algo, K* keys, V* vals, const int64_t* rank_off, int p, int k_way, int s, Comp comp,
                     bool increasing, K* spl_keys, V* spl_vals, int64_t spl_cap, int64_t* n_spl, int64_t* bucket_sizes,
                     int64_t bucket_cap, int64_t* n_bucket, int64_t* final_sizes, double* max_err, int* rounds) {
  using E = Elem<K, V, KV>;
  using T = typename E::T;
  auto lcomp = [&](const T& a, const T& b) {
    if constexpr (KV) {
      if (comp(a.first, b.first)) return true;
      if (comp(b.first, a.first)) return false;
      if (increasing) return a.second < b.second;
      return a.second > b.second;
    } else {
      return comp(a, b);
    }
  };
  std::vector<std::vector<T>> R(p);
  int64_t total = 0;
  for (int r = 0; r < p; r++) {
    for (int64_t i = rank_off[r]; i < rank_off[r + 1]; i++) R[r].push_back(E::make(keys, vals, i));
    std::sort(R[r].begin(), R[r].end(), lcomp);
    total += (int64_t)R[r].size();
  }
  Report rep;
  std::vector<T> spl;
  const bool pow2p = (p & (p - 1)) == 0, pow2k = k_way > 0 && (k_way & (k_way - 1)) == 0;
  if (algo == 0 && pow2p && pow2k) {
    vrank_hyksort(R, k_way, s, lcomp, rep, spl, total);
  } else if (algo == 2 && pow2p) {
    vrank_hyperquicksort(R, s, lcomp, rep, spl);
  } else if (p > 1 && total < 5ll * p * p) {
    BitonicNet<T> net{R, increasing};
    std::vector<int> comm(p);
    for (int r = 0; r < p; r++) comm[r] = r;
    net.sort(comm);
    rep.rounds = net.rounds;
  } else {
    vrank_samplesort(R, lcomp, increasing, rep, spl);
  }
  std::vector<int64_t> sizes(p);
  for (int r = 0; r < p; r++) sizes[r] = (int64_t)R[r].size();
  if (p > 1) keep_no_empty(sizes);
  int64_t o = 0;
  for (int r = 0; r < p; r++)
    for (const T& e : R[r]) E::store(e, keys, vals, o++);
  for (int r = 0; r < p; r++) final_sizes[r] = sizes[r];
  *n_spl = (int64_t)spl.size();
  for (int64_t i = 0; i < (int64_t)spl.size() && i < spl_cap; i++) {
    if constexpr (KV) {
      spl_keys[i] = spl[i].first;
      if (spl_vals) spl_vals[i] = spl[i].second;
    } else {
      spl_keys[i] = spl[i];
    }
  }
  *n_bucket = (int64_t)rep.bucket_sizes.size();
  for (int64_t i = 0; i < (int64_t)rep.bucket_sizes.size() && i < bucket_cap; i++) bucket_sizes[i] = rep.bucket_sizes[i];
  *max_err = rep.max_err;
  *rounds = rep.rounds;
  g_last_splitter_iters = rep.splitter_iters;
  return 0;
}

}  // namespace orc

// ====================================================================== C API
using namespace orc;

#define ORC_KEY_TYPES(X) X(i32, int32_t) X(u32, uint32_t) X(f32, float) X(i64, int64_t) X(u64, uint64_t) X(f64, double)
#define ORC_VAL_TYPES(X, kn, K) X(kn, K, u32, uint32_t) X(kn, K, u64, uint64_t) X(kn, K, i64, int64_t)

extern "C" {

void orc_set_threads(int n) { g_threads = n; }
int orc_last_splitter_iterations(void) { return g_last_splitter_iters; }
int orc_get_threads(void) { return max_threads(); }

#define DEF_KEYS(kn, K)                                                                          \
  void orc_merge_sort_##kn(K* a, int64_t n, int desc) {                                          \
    if (desc)                                                                                    \
      omp_merge_sort(std::greater<K>(), a, n);                                                   \
    else                                                                                         \
      omp_merge_sort(std::less<K>(), a, n);                                                      \
  }                                                                                              \
  void orc_std_sort_##kn(K* a, int64_t n, int desc) {                                            \
    if (desc)                                                                                    \
      std::sort(a, a + n, std::greater<K>());                                                    \
    else                                                                                         \
      std::sort(a, a + n, std::less<K>());                                                       \
  }                                                                                              \
  /* Batched form: every segment [off[s], off[s+1]) sorted on its own (the  \
     reference sorts one array per CudaUtils::Sort call). */                 \
  void orc_seg_sort_##kn(K* a, const int64_t* off, int64_t nseg, int desc) {                     \
    _Pragma("omp parallel for schedule(dynamic, 64)") for (int64_t s = 0; s < nseg; s++) {       \
      if (desc)                                                                                  \
        std::sort(a + off[s], a + off[s + 1], std::greater<K>());                                \
      else                                                                                       \
        std::sort(a + off[s], a + off[s + 1], std::less<K>());                                   \
    }                                                                                            \
  }                                                                                              \
  /* Bucket partition of unsorted keys: bucket = splitters <= key (upper 0, */ \
  /* the lower_bound cut of hyksort_helper.h:563-568) or < key (upper 1,     */ \
  /* samplesort_helper.h:525-537); stable within a bucket.                   */ \
  void orc_multisplit_##kn(const K* keys, int64_t n, const K* spl, int64_t nspl, int upper,      \
                           int desc, int64_t* out_index, int64_t* offsets) {                     \
    std::vector<int64_t> b(n), cnt(nspl + 2, 0);                                                 \
    for (int64_t i = 0; i < n; i++) {                                                            \
      const K* p = desc ? (upper ? std::lower_bound(spl, spl + nspl, keys[i], std::greater<K>())  \
                                 : std::upper_bound(spl, spl + nspl, keys[i], std::greater<K>())) \
                        : (upper ? std::lower_bound(spl, spl + nspl, keys[i], std::less<K>())     \
                                 : std::upper_bound(spl, spl + nspl, keys[i], std::less<K>()));   \
      b[i] = p - spl;                                                                            \
      cnt[b[i] + 1]++;                                                                           \
    }                                                                                            \
    for (int64_t j = 0; j <= nspl; j++) cnt[j + 1] += cnt[j];                                    \
    for (int64_t j = 0; j <= nspl + 1; j++) offsets[j] = cnt[j];                                 \
    for (int64_t i = 0; i < n; i++) out_index[cnt[b[i]]++] = i;                                  \
  }                                                                                              \
  void orc_merge_##kn(const K* a, int64_t na, const K* b, int64_t nb, K* out, int desc) {        \
    if (desc)                                                                                    \
      omp_merge(std::greater<K>(), a, a + na, b, b + nb, out);                                   \
    else                                                                                         \
      omp_merge(std::less<K>(), a, a + na, b, b + nb, out);                                      \
  }                                                                                              \
  void orc_bounds_##kn(const K* run, int64_t n, const K* spl, int64_t nspl, int upper, int desc, \
                       int64_t* out) {                                                           \
    for (int64_t j = 0; j < nspl; j++) {                                                         \
      if (desc)                                                                                  \
        out[j] = upper ? std::upper_bound(run, run + n, spl[j], std::greater<K>()) - run         \
                       : std::lower_bound(run, run + n, spl[j], std::greater<K>()) - run;        \
      else                                                                                       \
        out[j] = upper ? std::upper_bound(run, run + n, spl[j], std::less<K>()) - run            \
                       : std::lower_bound(run, run + n, spl[j], std::less<K>()) - run;           \
    }                                                                                            \
  }                                                                                              \
  int orc_vrank_##kn(int algo, K* keys, const int64_t* rank_off, int p, int k_way, int s,        \
                     int desc, K* spl, int64_t spl_cap, int64_t* n_spl, int64_t* bucket_sizes,   \
                     int64_t bucket_cap, int64_t* n_bucket, int64_t* final_sizes, double* err,   \
                     int* rounds) {                                                              \
    if (desc)                                                                                    \
      return run_vrank<K, uint32_t, false>(algo, keys, nullptr, rank_off, p, k_way, s,           \
                                           std::greater<K>(), false, spl, nullptr, spl_cap,      \
                                           n_spl, bucket_sizes, bucket_cap, n_bucket,            \
                                           final_sizes, err, rounds);                            \
    return run_vrank<K, uint32_t, false>(algo, keys, nullptr, rank_off, p, k_way, s,             \
                                         std::less<K>(), true, spl, nullptr, spl_cap, n_spl,     \
                                         bucket_sizes, bucket_cap, n_bucket, final_sizes, err,   \
                                         rounds);                                                \
  }                                                                                              \
  ORC_VAL_TYPES(DEF_PAIRS, kn, K)

#define DEF_PAIRS(kn, K, vn, V)                                                                        \
  void orc_merge_sort_pairs_##kn##_##vn(K* k, V* v, int64_t n, int desc) {                             \
    std::vector<std::pair<K, V>> items(n);                                                             \
    for (int64_t i = 0; i < n; i++) items[i] = std::make_pair(k[i], v[i]);                             \
    if (desc)                                                                                          \
      omp_merge_sort(std::greater<std::pair<K, V>>(), items.data(), n);                                \
    else                                                                                               \
      omp_merge_sort(std::less<std::pair<K, V>>(), items.data(), n);                                   \
    for (int64_t i = 0; i < n; i++) {                                                                  \
      k[i] = items[i].first;                                                                           \
      v[i] = items[i].second;                                                                          \
    }                                                                                                  \
  }                                                                                                    \
  void orc_stable_sort_pairs_##kn##_##vn(K* k, V* v, int64_t n, int desc) {                            \
    std::vector<int64_t> idx(n);                                                                       \
    for (int64_t i = 0; i < n; i++) idx[i] = i;                                                        \
    if (desc)                                                                                          \
      std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return k[a] > k[b]; });     \
    else                                                                                               \
      std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return k[a] < k[b]; });     \
    std::vector<K> kk(n);                                                                              \
    std::vector<V> vv(n);                                                                              \
    for (int64_t i = 0; i < n; i++) {                                                                  \
      kk[i] = k[idx[i]];                                                                               \
      vv[i] = v[idx[i]];                                                                               \
    }                                                                                                  \
    std::copy(kk.begin(), kk.end(), k);                                                                \
    std::copy(vv.begin(), vv.end(), v);                                                                \
  }                                                                                                    \
  /* Batched form of the stable KV sort, one segment at a time. */                                     \
  void orc_seg_stable_sort_pairs_##kn##_##vn(K* k, V* v, const int64_t* off, int64_t nseg, int desc) { \
    _Pragma("omp parallel for schedule(dynamic, 64)") for (int64_t s = 0; s < nseg; s++)               \
        orc_stable_sort_pairs_##kn##_##vn(k + off[s], v + off[s], off[s + 1] - off[s], desc);         \
  }                                                                                                    \
  int orc_vrank_pairs_##kn##_##vn(int algo, K* keys, V* vals, const int64_t* rank_off, int p,          \
                                  int k_way, int s, int desc, K* spl_k, V* spl_v, int64_t spl_cap,     \
                                  int64_t* n_spl, int64_t* bucket_sizes, int64_t bucket_cap,           \
                                  int64_t* n_bucket, int64_t* final_sizes, double* err, int* rounds) { \
    if (desc)                                                                                          \
      return run_vrank<K, V, true>(algo, keys, vals, rank_off, p, k_way, s, std::greater<K>(), false,  \
                                   spl_k, spl_v, spl_cap, n_spl, bucket_sizes, bucket_cap, n_bucket,   \
                                   final_sizes, err, rounds);                                          \
    return run_vrank<K, V, true>(algo, keys, vals, rank_off, p, k_way, s, std::less<K>(), true, spl_k, \
                                 spl_v, spl_cap, n_spl, bucket_sizes, bucket_cap, n_bucket,            \
                                 final_sizes, err, rounds);                                            \
  }

ORC_KEY_TYPES(DEF_KEYS)

void orc_scan_i64(int64_t cnt, const int64_t* a, int64_t* b) { omp_scan(cnt, a, b); }

// MergeLists of the small-input bitonic network on int64 lists (a own, b
// partner's); returns the kept list's length, written to out.
int64_t orc_merge_lists_i64(const int64_t* a, int64_t na, const int64_t* b, int64_t nb, int keep_low, int inc,
                            int64_t* out) {
  const std::vector<int64_t> va(a, a + na), vb(b, b + nb);
  const std::vector<int64_t> r = merge_lists_ref(va, vb, keep_low != 0, inc != 0);
  std::copy(r.begin(), r.end(), out);
  return (int64_t)r.size();
}

void orc_keep_no_empty(int64_t* sizes, int p) {
  std::vector<int64_t> s(sizes, sizes + p);
  keep_no_empty(s);
  std::copy(s.begin(), s.end(), sizes);
}

}  // extern "C"
