// Per-rank pipeline over p virtual ranks held as segments of one device array
// (gs_vrank_sort).
//
// The reference runs one MPI process per rank (SortAmongProcesses:
// include/helpers/hyksort_helper.h:453-536, samplesort_helper.h:420-501,
// hyperquicksort_helper.h:261-346).  Here every rank is a device segment, the
// MPI exchange is addressing (a received bucket is read where its sender
// holds it), and all ranks of a round are processed by one launch per step:
//
//   local sort   every rank's run sorted by one segmented sort over all ranks
//   samples      regular samples of every rank gathered on the device and
//                sorted per group (tile sort / radix sort)
//   splitters    picked from the sorted samples by the rule below
//   multisplit   bucket bounds of every run by binary search
//                (lower_bound for HykSort / HyperQuickSort, upper_bound for
//                SampleSort)
//   exchange     bucket sizes -> new rank offsets (host scan of p*k counts)
//   merge        one k-way pass over each destination's received runs
//                (merge_kway.hip), reading them in place in the senders' runs
//
// Order: with a payload, elements compare as (key, payload) pairs under the
// lifted comparator (std::less/greater of std::pair, include/common.h:
// 156-164), the order the reference's rounds use.  The local sort is made
// lexicographic by a stable payload sort first; it is skipped when the
// payloads are already non-decreasing in that order (e.g. original indexes).
//
// Splitter rules (the reference draws HykSort / HyperQuickSort samples with
// rand(), which no re-run can reproduce; these rules are deterministic and
// the oracle re-runs them exactly):
//   HykSort         the iterative selection of SelectSplitters with regular
//                   draws, every round on the device (hyk_device.hip)
//   SampleSort      p-1 samples run[(i+1)*n/p] per rank (samplesort_helper.h:
//                   402-405); ascending: splitter r = S[(r+1)(p-1)-1],
//                   descending: S[(r+1)(p-1)] (:411-416)
//   HyperQuickSort  s regular samples per rank; pivot = the first sample whose
//                   global lower_bound rank is closest to N_group/2 (:251-256)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/gpusort_hip.h"
#include "gs_capi_impl.h"
#include "gs_internal.h"

using namespace gs;

namespace {

struct Pipe {
  gs_handle* h;
  gs_dtype kdt, vdt;
  int kb, vb;
  bool desc, kv;
  XformParams kx, vx;
  char* kbuf[2];
  char* vbuf[2];
  int cur;
  uint64_t n;
  // host-side aux staging (device copies live in the aux arena slot)
  std::vector<char> stage;
  // host sources of queued uploads, kept until the call's final wait
  std::vector<std::vector<unsigned char>> keep;
};

bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }

// Sorted-order sort of a device segment: keys-only, or (key, payload)
// lexicographic when lex_values is set, else stable by key.
gs_status sort_segment(Pipe& P, char* k, char* v, uint64_t len, bool lex_values) {
  if (len <= 1) return GS_OK;
  if (P.kv) {
    if (lex_values) {
      gs_status st = sort_device(P.h, P.vdt, v, P.kdt, k, len, P.desc);
      if (st != GS_OK) return st;
    }
    return sort_device(P.h, P.kdt, k, P.vdt, v, len, P.desc);
  }
  return sort_device(P.h, P.kdt, k, GS_NONE, nullptr, len, P.desc);
}

// Sorts every run [off[r], off[r+1]) of the current buffer pair in place with
// the segmented sort (one launch sequence for all runs), lexicographically
// (payload first, then stable by key) when lex_values is set.
gs_status sort_runs(Pipe& P, char* k, char* v, const std::vector<uint64_t>& off, bool lex_values) {
  const uint32_t nseg = (uint32_t)off.size() - 1;
  std::vector<unsigned char> meta;
  std::vector<uint32_t> huge;
  auto pass = [&](gs_dtype kdt, char* keys, gs_dtype vdt, char* vals, int kb, int vb) -> gs_status {
    gs_status st = sort_segments_async(P.h, kdt, keys, vdt, vals, off.data(), nseg, P.desc, &meta, &huge);
    if (st != GS_OK) return st;
    for (uint32_t r : huge) {
      st = sort_device(P.h, kdt, keys + off[r] * kb, vdt, vals ? vals + off[r] * vb : nullptr, off[r + 1] - off[r],
                       P.desc);
      if (st != GS_OK) return st;
    }
    // `meta` is the source of an upload queued above: kept until the end.
    P.keep.push_back(std::move(meta));
    return GS_OK;
  };
  if (!P.kv) return pass(P.kdt, k, GS_NONE, nullptr, P.kb, 0);
  if (lex_values) {
    gs_status st = pass(P.vdt, v, P.kdt, k, P.vb, P.kb);
    if (st != GS_OK) return st;
  }
  return pass(P.kdt, k, P.vdt, v, P.kb, P.vb);
}

// Device aux area: a list of (host source, bytes) blocks uploaded together.
struct AuxLayout {
  std::vector<std::pair<const void*, size_t>> blocks;
  std::vector<size_t> offs;
  size_t total = 0;
  size_t add(const void* src, size_t bytes) {
    offs.push_back(total);
    blocks.push_back({src, bytes});
    const size_t o = total;
    total = align_up(total + (bytes ? bytes : 8), 256);
    return o;
  }
};

// Uploads the layout's blocks; `scratch` more device bytes follow them
// (at base + L.total), not uploaded.
gs_status upload(Pipe& P, AuxLayout& L, char** base, size_t scratch = 0) {
  void* ws = nullptr;
  hipError_t e = P.h->arena.get(Arena::kPipeAux, L.total + scratch, P.h->stream, &ws);
  if (e != hipSuccess) return from_hip(e, "pipeline workspace");
  *base = reinterpret_cast<char*>(ws);
  P.stage.assign(L.total, 0);
  for (size_t i = 0; i < L.blocks.size(); ++i)
    if (L.blocks[i].first && L.blocks[i].second) memcpy(P.stage.data() + L.offs[i], L.blocks[i].first, L.blocks[i].second);
  e = hipMemcpyAsync(ws, P.stage.data(), L.total, hipMemcpyHostToDevice, P.h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(P.h->stream);  // stage buffer is reused by the next upload
  return from_hip(e, "pipeline upload");
}

// Gathers element positions `idx` of the current runs (keys and payloads),
// sorts consecutive groups (group g = [gofs[g], gofs[g+1]) of the gathered
// list) in the pipeline order, and returns them on the host.
gs_status gather_sorted(Pipe& P, const std::vector<uint64_t>& idx, const std::vector<uint64_t>& gofs,
                        std::vector<char>* hk, std::vector<char>* hv) {
  const uint32_t m = (uint32_t)idx.size();
  hk->assign((size_t)m * P.kb, 0);
  hv->assign(P.kv ? (size_t)m * P.vb : 0, 0);
  if (m == 0) return GS_OK;
  AuxLayout L;
  const size_t o_idx = L.add(idx.data(), m * 8);
  const size_t o_k = L.add(nullptr, (size_t)m * P.kb);
  const size_t o_v = L.add(nullptr, P.kv ? (size_t)m * P.vb : 0);
  char* base = nullptr;
  gs_status st = upload(P, L, &base);
  if (st != GS_OK) return st;
  const uint64_t* d_idx = reinterpret_cast<const uint64_t*>(base + o_idx);
  hipError_t e = gather_elems(P.kb, P.kbuf[P.cur], d_idx, base + o_k, m, P.h->stream);
  if (e == hipSuccess && P.kv) e = gather_elems(P.vb, P.vbuf[P.cur], d_idx, base + o_v, m, P.h->stream);
  if (e != hipSuccess) return from_hip(e, "gather samples");
  for (size_t g = 0; g + 1 < gofs.size(); ++g) {
    const uint64_t a = gofs[g], b = gofs[g + 1];
    st = sort_segment(P, base + o_k + a * P.kb, P.kv ? base + o_v + a * P.vb : nullptr, b - a, P.kv);
    if (st != GS_OK) return st;
  }
  e = hipMemcpyAsync(hk->data(), base + o_k, (size_t)m * P.kb, hipMemcpyDeviceToHost, P.h->stream);
  if (e == hipSuccess && P.kv)
    e = hipMemcpyAsync(hv->data(), base + o_v, (size_t)m * P.vb, hipMemcpyDeviceToHost, P.h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(P.h->stream);
  return from_hip(e, "read samples");
}

// Bucket bounds of every run of the current buffer: run r cut by splitters
// spl[first[r] .. first[r]+nspl).  Returns bounds[r*(nspl+2) + j].
gs_status run_bounds(Pipe& P, const std::vector<uint64_t>& off, const std::vector<char>& spl_k,
                     const std::vector<char>& spl_v, const std::vector<uint32_t>& first, uint32_t nspl, bool upper,
                     std::vector<uint64_t>* bounds) {
  const uint32_t nruns = (uint32_t)off.size() - 1;
  bounds->assign((size_t)nruns * (nspl + 2), 0);
  AuxLayout L;
  const size_t o_off = L.add(off.data(), off.size() * 8);
  const size_t o_sk = L.add(spl_k.data(), spl_k.size());
  const size_t o_sv = L.add(P.kv ? spl_v.data() : nullptr, P.kv ? spl_v.size() : 0);
  const size_t o_first = L.add(first.data(), first.size() * 4);
  const size_t o_b = L.add(nullptr, bounds->size() * 8);
  char* base = nullptr;
  gs_status st = upload(P, L, &base);
  if (st != GS_OK) return st;
  hipError_t e = bucket_bounds(P.kb, P.kbuf[P.cur], reinterpret_cast<const uint64_t*>(base + o_off), nruns, base + o_sk,
                               reinterpret_cast<const uint32_t*>(base + o_first), nspl, upper ? 1 : 0, P.kx,
                               reinterpret_cast<uint64_t*>(base + o_b), P.h->stream, P.kv ? P.vb : 0,
                               P.kv ? P.vbuf[P.cur] : nullptr, P.kv ? base + o_sv : nullptr, P.kv ? &P.vx : nullptr);
  if (e == hipSuccess)
    e = hipMemcpyAsync(bounds->data(), base + o_b, bounds->size() * 8, hipMemcpyDeviceToHost, P.h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(P.h->stream);
  return from_hip(e, "bucket bounds");
}

struct Run {
  uint64_t off, len;
};

// Runs 2-way merge jobs from the current buffer into the other one (one
// launch for all jobs) and makes that buffer current.
gs_status run_merge_jobs(Pipe& P, const std::vector<MergeJob>& jobs) {
  std::vector<MergeJobDev> dj(jobs.size());
  std::vector<uint32_t> tile_job;
  const uint32_t ntiles = merge_plan(jobs.data(), (uint32_t)jobs.size(), dj.data(), &tile_job);
  if (ntiles) {
    AuxLayout L;
    const size_t o_j = L.add(dj.data(), dj.size() * sizeof(MergeJobDev));
    const size_t o_t = L.add(tile_job.data(), tile_job.size() * 4);
    const size_t o_s = L.add(nullptr, (size_t)ntiles * 8);
    char* base = nullptr;
    gs_status st = upload(P, L, &base);
    if (st != GS_OK) return st;
    const int src = P.cur, dst = P.cur ^ 1;
    hipError_t e = merge_tiles(P.kb, P.kv ? P.vb : 0, P.kbuf[src], P.kbuf[src], P.kv ? P.vbuf[src] : nullptr,
                               P.kv ? P.vbuf[src] : nullptr, P.kbuf[dst], P.kv ? P.vbuf[dst] : nullptr,
                               reinterpret_cast<const MergeJobDev*>(base + o_j),
                               reinterpret_cast<const uint32_t*>(base + o_t), ntiles, P.kx,
                               reinterpret_cast<uint64_t*>(base + o_s), P.h->stream, P.kv ? &P.vx : nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(P.h->stream);
    if (e != hipSuccess) return from_hip(e, "merge level");
  }
  P.cur ^= 1;
  return GS_OK;
}

// Merges, for every destination, its list of runs (read from the current
// buffer, in list order, ties to the earlier run) into one run placed at
// dest_off[d] of the result: one k-way pass for lists of up to kKwayMax
// runs; each pass ping-pongs between the two buffers and the result buffer
// becomes current.
gs_status merge_lists(Pipe& P, std::vector<std::vector<Run>> lists, const std::vector<uint64_t>& dest_off) {
  for (;;) {
    // Done once every destination holds at most one non-empty run, placed
    // at its offset.
    bool need = false;
    for (size_t d = 0; d < lists.size(); ++d) {
      size_t live = 0;
      uint64_t at = dest_off[d];
      for (const Run& r : lists[d])
        if (r.len) {
          ++live;
          at = r.off;
        }
      if (live > 1 || at != dest_off[d]) need = true;
    }
    if (!need) break;
    // One pass over every destination: up to kKwayMax runs merged in one
    // k-way step (a single run is copied, so the pass writes the whole
    // buffer it owns); longer lists merge consecutive groups of kKwayMax
    // runs, in list order, and go round again.
    KwayPlan plan;
    std::vector<std::vector<Run>> next(lists.size());
    for (size_t d = 0; d < lists.size(); ++d) {
      std::vector<Run> live;
      for (const Run& r : lists[d])
        if (r.len) live.push_back(r);
      uint64_t pos = dest_off[d];
      for (size_t j = 0; j < live.size(); j += kKwayMax) {
        KwayDest D{};
        const size_t m = std::min<size_t>(kKwayMax, live.size() - j);
        uint64_t len = 0;
        for (size_t r = 0; r < m; ++r) {
          D.run_off[r] = live[j + r].off;
          D.run_len[r] = live[j + r].len;
          len += live[j + r].len;
        }
        D.m = (uint32_t)m;
        D.out_off = pos;
        plan.dests.push_back(D);
        next[d].push_back(Run{pos, len});
        pos += len;
      }
    }
    kway_plan(&plan);
    AuxLayout L;
    const size_t o_d = L.add(plan.dests.data(), plan.dests.size() * sizeof(KwayDest));
    const size_t o_t = L.add(&plan.totals, sizeof(KwayTotals));
    char* base = nullptr;
    const int vb = P.kv ? P.vb : 0;
    gs_status st = upload(P, L, &base, kway_scratch_bytes(plan.totals, P.kb, vb));
    if (st != GS_OK) return st;
    const KwayDevice dev{reinterpret_cast<const KwayDest*>(base + o_d), reinterpret_cast<const KwayTotals*>(base + o_t),
                         base + L.total};
    const int src = P.cur, dst = P.cur ^ 1;
    hipError_t e = kway_merge(P.kb, vb, P.kbuf[src], P.kv ? P.vbuf[src] : nullptr, P.kbuf[dst],
                              P.kv ? P.vbuf[dst] : nullptr, plan.totals, dev, P.kx, P.kv ? &P.vx : nullptr,
                              P.h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(P.h->stream);
    if (e != hipSuccess) return from_hip(e, "k-way merge");
    P.cur ^= 1;
    lists.swap(next);
  }
  return GS_OK;
}

// ParUtils::KeepNoProcessHasEmptyData (include/par/par_utils.h:103-175) on
// the size vector: data keeps its global order, rank boundaries move.  The
// gathered sizes are not updated inside the loop and the rank after a
// handled partner is skipped, as in the reference.
void keep_no_empty(std::vector<uint64_t>& sizes) {
  const int n_p = (int)sizes.size();
  const std::vector<uint64_t> nelems = sizes;
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
      const uint64_t size_sync = nelems[partner] / (uint64_t)(empty_cnt + 1);
      const int last = use_prev ? n_p : partner;
      for (int j = i; j < last; j++) sizes[j] = size_sync;
      sizes[partner] -= (uint64_t)empty_cnt * size_sync;
      if (use_prev) i = n_p;
      if (i < n_p - 1) i = partner + 1;
    }
  }
}

std::vector<uint64_t> offsets_of(const std::vector<uint64_t>& sizes) {
  std::vector<uint64_t> off(sizes.size() + 1, 0);
  for (size_t i = 0; i < sizes.size(); ++i) off[i + 1] = off[i] + sizes[i];
  return off;
}

struct Report {
  gs_vrank_report* r;
  int kb;
  void add_splitters(const std::vector<char>& k) {
    if (!r) return;
    const uint32_t cnt = (uint32_t)(k.size() / kb);
    for (uint32_t i = 0; i < cnt; ++i) {
      if (r->splitters && r->n_splitters < r->max_splitters)
        memcpy(reinterpret_cast<char*>(r->splitters) + (size_t)r->n_splitters * kb, k.data() + (size_t)i * kb, kb);
      r->n_splitters++;
    }
  }
  void add_bucket(uint64_t s) {
    if (!r) return;
    if (r->bucket_sizes && r->n_bucket_sizes < r->max_bucket_sizes) r->bucket_sizes[r->n_bucket_sizes] = s;
    r->n_bucket_sizes++;
  }
};

// Regular sample positions of run [o, o+len): o + (j+1)*len/(s+1).
void regular_samples(uint64_t o, uint64_t len, uint32_t s, std::vector<uint64_t>* idx) {
  if (len == 0) return;
  for (uint32_t j = 0; j < s; ++j) idx->push_back(o + (uint64_t)(j + 1) * len / (s + 1));
}

// The handle's pinned host buffer, grown to at least `bytes` (every call
// that uses it synchronises before returning, so it is free to replace).
hipError_t host_pinned(gs_handle* h, size_t bytes, void** out) {
  if (h->pinned_bytes < bytes) {
    if (h->pinned) (void)hipHostFree(h->pinned);
    h->pinned = nullptr;
    h->pinned_bytes = 0;
    hipError_t e = hipHostMalloc(&h->pinned, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
      h->pinned = nullptr;
      return e;
    }
    h->pinned_bytes = bytes;
  }
  *out = h->pinned;
  return hipSuccess;
}

// HykSort's rounds, every step on the device (hyk_device.hip); the report
// is filled from the bounds and splitters the rounds leave behind.
gs_status hyksort_rounds(Pipe& P, std::vector<uint64_t>& sizes, uint32_t k_way, uint32_t s, Report& rep,
                         double* max_err, uint32_t* rounds, uint32_t* iterations) {
  const uint32_t p = (uint32_t)sizes.size();
  if (!hyk_device_supported(p, k_way, s))
    return fail(GS_UNSUPPORTED, "HykSort with %u ranks, k = %u and %u samples per rank exceeds the device "
                "selection's limits", p, k_way, s);
  const size_t ws_bytes = hyk_device_workspace_bytes(p, k_way, s, P.n, P.kb, P.kv ? P.vb : 0);
  void* ws = nullptr;
  hipError_t e = P.h->arena.get(Arena::kHykWork, ws_bytes, P.h->stream, &ws);
  if (e != hipSuccess) return from_hip(e, "HykSort workspace");
  void* host = nullptr;
  e = host_pinned(P.h, hyk_host_bytes(p, k_way, s, P.kb, P.kv ? P.vb : 0), &host);
  if (e != hipSuccess) return from_hip(e, "HykSort host buffer");
  const std::vector<uint64_t> off = offsets_of(sizes);
  HykRun run{};
  run.host = host;
  run.kb = P.kb;
  run.vb = P.kv ? P.vb : 0;
  run.lex = P.kv;
  run.kx = P.kx;
  run.vx = P.vx;
  run.kbuf[0] = P.kbuf[0];
  run.kbuf[1] = P.kbuf[1];
  run.vbuf[0] = P.vbuf[0];
  run.vbuf[1] = P.vbuf[1];
  run.cur = P.cur;
  run.p = p;
  run.k_way = k_way;
  run.s = s;
  run.n = P.n;
  HykResult res;
  e = hyk_rounds_device(&run, off.data(), ws, ws_bytes, P.h->stream, P.h->timer(), &res);
  if (e != hipSuccess) return from_hip(e, "HykSort rounds");
  P.cur = run.cur;
  uint64_t total = 0;
  for (auto z : sizes) total += z;
  for (uint32_t ri = 0; ri < res.rounds; ++ri) {
    const uint32_t kway = res.kway[ri], n_p = res.n_p[ri], kk = kway + 1, groups = p / n_p;
    const uint64_t* bb = res.bounds.data() + (size_t)ri * p * res.bnd_stride;
    auto B = [&](uint32_t r, uint32_t j) { return bb[(size_t)r * (kway + 2) + j]; };
    const char* sk = res.spl_k.data() + (size_t)ri * p * res.spl_stride * P.kb;
    for (uint32_t g = 0; g < groups; ++g) {
      uint64_t gtot = 0;
      for (uint32_t r = g * n_p; r < (g + 1) * n_p; ++r) gtot += B(r, kway + 1) - B(r, 0);
      if (gtot == 0) continue;  // an empty group cuts nothing
      rep.add_splitters(std::vector<char>(sk + (size_t)g * kway * P.kb, sk + (size_t)(g + 1) * kway * P.kb));
      // splitter quality: |global rank - (i+1) * N_group / k| / N
      for (uint32_t i = 0; i < kway; ++i) {
        uint64_t grank = 0;
        for (uint32_t r = g * n_p; r < (g + 1) * n_p; ++r) grank += B(r, i + 1) - B(r, 0);
        const double err =
            std::fabs((double)grank - (double)(i + 1) * (double)gtot / kk) / (double)std::max<uint64_t>(total, 1);
        *max_err = std::max(*max_err, err);
      }
    }
    for (uint32_t r = 0; r < p; ++r)
      for (uint32_t i = 0; i < kk; ++i) rep.add_bucket(B(r, i + 1) - B(r, i));
    *iterations += res.iterations[ri];
  }
  for (uint32_t r = 0; r < p; ++r) sizes[r] = res.final_off[r + 1] - res.final_off[r];
  *rounds = res.rounds;
  return GS_OK;
}

// ---------------------------------------------- SampleSort below 5 p^2 keys
//
// The reference sorts small inputs with a bitonic merge-split network across
// the ranks (samplesort_helper.h:429-452; BitonicSort :192-235,
// BitonicSortBinary :236-258, BitonicSortBinaryHelper :260-283, BitonicMerge
// :170-190): partners swap lists and each keeps a prefix or suffix of the
// merge (MergeLists :305-387), so rank sizes follow the data.  The network
// is generated here as the reference's communicator splits produce it
// (ParUtilsHelper::SplitCommBinary flips the upper group's rank order,
// SplitCommBinaryNoFlip does not; src/par_utils_helper.cc:237-329); every
// round is one cut kernel (kept lengths and merge-path splits, read back)
// and one batched merge of the kept parts.

struct NetOp {
  uint32_t rank, partner;
  bool keep_low;
};

uint32_t prev_pow2(uint32_t v) {  // GetPrevHighestPowerOfTwo (include/common.h:185-194)
  v--;
  v |= v >> 1;
  v |= v >> 2;
  v |= v >> 4;
  v |= v >> 8;
  v |= v >> 16;
  v++;
  return v >> 1;
}

void net_binary(const std::vector<uint32_t>& comm, bool inc, std::vector<std::vector<NetOp>>* rounds) {
  const uint32_t npes = (uint32_t)comm.size();
  uint32_t and_bit = 2;
  for (uint32_t pss = 2; pss <= npes; pss *= 2, and_bit <<= 1) {
    uint32_t dim = 0;
    for (uint32_t x = pss; x > 1; x >>= 1) ++dim;
    uint32_t eor_bit = 1u << (dim - 1);
    for (uint32_t stage = 0; stage < dim; ++stage, eor_bit >>= 1) {
      std::vector<NetOp> ops;
      for (uint32_t r = 0; r < npes; ++r) {
        const bool incr = (r & and_bit) == 0 ? inc : !inc;
        const uint32_t partner = r ^ eor_bit;
        ops.push_back(NetOp{comm[r], comm[partner], incr ? r < partner : r > partner});
      }
      rounds->push_back(ops);
    }
  }
}

void net_bitonic(const std::vector<uint32_t>& comm, bool inc, std::vector<std::vector<NetOp>>* rounds) {
  const uint32_t npes = (uint32_t)comm.size();
  if (npes == 1) return;
  if (is_pow2(npes)) {
    net_binary(comm, inc, rounds);
    return;
  }
  const uint32_t splitter = prev_pow2(npes);
  const std::vector<uint32_t> A(comm.begin(), comm.begin() + splitter);
  std::vector<uint32_t> b_flip;
  for (uint32_t j = 0; j < npes - splitter; ++j) b_flip.push_back(comm[npes - 1 - j]);
  net_binary(A, inc, rounds);
  net_bitonic(b_flip, inc, rounds);
  const uint32_t num_left = prev_pow2(npes), num_right = npes - num_left;
  std::vector<NetOp> ops;
  for (uint32_t r = num_left - num_right; r < num_left; ++r) ops.push_back(NetOp{comm[r], comm[r + num_right], inc});
  for (uint32_t r = num_left; r < npes; ++r) ops.push_back(NetOp{comm[r], comm[r - num_right], !inc});
  rounds->push_back(ops);
  net_binary(A, inc, rounds);
  net_bitonic(std::vector<uint32_t>(comm.begin() + splitter, comm.end()), inc, rounds);
}

gs_status samplesort_small(Pipe& P, std::vector<uint64_t>& sizes, uint32_t* rounds) {
  const uint32_t p = (uint32_t)sizes.size();
  std::vector<uint32_t> comm(p);
  for (uint32_t r = 0; r < p; ++r) comm[r] = r;
  std::vector<std::vector<NetOp>> net;
  net_bitonic(comm, !P.desc, &net);
  for (const std::vector<NetOp>& round : net) {
    std::vector<uint64_t> off(p + 1, 0);
    for (uint32_t r = 0; r < p; ++r) off[r + 1] = off[r] + sizes[r];
    const uint32_t nops = (uint32_t)round.size();
    std::vector<NetCut> ops(nops);
    for (uint32_t i = 0; i < nops; ++i) {
      const NetOp& o = round[i];
      ops[i] = NetCut{off[o.rank], sizes[o.rank], off[o.partner], sizes[o.partner], o.keep_low ? 1u : 0u, 0u};
    }
    AuxLayout L;
    const size_t o_ops = L.add(ops.data(), nops * sizeof(NetCut));
    const size_t o_out = L.add(nullptr, nops * sizeof(NetCutResult));
    char* base = nullptr;
    gs_status st = upload(P, L, &base);
    if (st != GS_OK) return st;
    std::vector<NetCutResult> res(nops);
    hipError_t e = net_cuts(P.kb, P.kv ? P.vb : 0, P.kbuf[P.cur], P.kv ? P.vbuf[P.cur] : nullptr,
                            reinterpret_cast<const NetCut*>(base + o_ops), nops, P.kx, P.kv ? &P.vx : nullptr,
                            P.desc ? 0 : 1, reinterpret_cast<NetCutResult*>(base + o_out), P.h->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(res.data(), base + o_out, nops * sizeof(NetCutResult), hipMemcpyDeviceToHost, P.h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(P.h->stream);
    if (e != hipSuccess) return from_hip(e, "bitonic cuts");
    std::vector<uint64_t> ns = sizes;
    std::vector<bool> busy(p, false);
    for (uint32_t i = 0; i < nops; ++i) {
      const NetOp& o = round[i];
      const NetCutResult& c = res[i];
      const uint64_t na = sizes[o.rank], nb = sizes[o.partner];
      const uint64_t taken = c.a_split + c.b_split;
      if (c.a_split > na || c.b_split > nb || c.keep == 0 || (c.prefix ? taken != c.keep : taken + c.keep != na + nb))
        return fail(GS_DEVICE_ERROR, "inconsistent bitonic cut at rank %u", o.rank);
      ns[o.rank] = c.keep;
      busy[o.rank] = true;
    }
    std::vector<uint64_t> noff(p + 1, 0);
    for (uint32_t r = 0; r < p; ++r) noff[r + 1] = noff[r] + ns[r];
    std::vector<MergeJob> jobs;
    for (uint32_t i = 0; i < nops; ++i) {
      const NetOp& o = round[i];
      const NetCutResult& c = res[i];
      if (c.prefix)
        jobs.push_back(MergeJob{off[o.rank], c.a_split, off[o.partner], c.b_split, noff[o.rank]});
      else
        jobs.push_back(MergeJob{off[o.rank] + c.a_split, sizes[o.rank] - c.a_split, off[o.partner] + c.b_split,
                                sizes[o.partner] - c.b_split, noff[o.rank]});
    }
    for (uint32_t r = 0; r < p; ++r)
      if (!busy[r]) jobs.push_back(MergeJob{off[r], sizes[r], off[r], 0, noff[r]});
    st = run_merge_jobs(P, jobs);
    if (st != GS_OK) return st;
    sizes.swap(ns);
  }
  *rounds = (uint32_t)net.size();
  return GS_OK;
}

gs_status samplesort_round(Pipe& P, std::vector<uint64_t>& sizes, Report& rep, uint32_t* rounds) {
  const uint32_t p = (uint32_t)sizes.size();
  if (p == 1) return GS_OK;
  std::vector<uint64_t> off = offsets_of(sizes);
  std::vector<uint64_t> idx;
  for (uint32_t r = 0; r < p; ++r)
    for (uint32_t i = 0; i + 1 < p; ++i) idx.push_back(off[r] + (uint64_t)(i + 1) * sizes[r] / p);
  std::vector<uint64_t> gofs{0, idx.size()};
  std::vector<char> sk, sv;
  gs_status st = gather_sorted(P, idx, gofs, &sk, &sv);
  if (st != GS_OK) return st;
  std::vector<char> spl_k((size_t)(p - 1) * P.kb), spl_v((size_t)(p - 1) * P.vb);
  for (uint32_t r = 0; r + 1 < p; ++r) {
    const uint64_t src = P.desc ? (uint64_t)(r + 1) * (p - 1) : (uint64_t)(r + 1) * (p - 1) - 1;
    memcpy(spl_k.data() + (size_t)r * P.kb, sk.data() + src * P.kb, P.kb);
    if (P.kv) memcpy(spl_v.data() + (size_t)r * P.vb, sv.data() + src * P.vb, P.vb);
  }
  rep.add_splitters(spl_k);
  std::vector<uint32_t> first(p, 0);
  std::vector<uint64_t> bnd;
  st = run_bounds(P, off, spl_k, spl_v, first, p - 1, true, &bnd);
  if (st != GS_OK) return st;
  auto B = [&](uint32_t r, uint32_t i) { return bnd[(size_t)r * (p + 1) + i]; };
  std::vector<uint64_t> nsizes(p, 0);
  std::vector<std::vector<Run>> lists(p);
  for (uint32_t r = 0; r < p; ++r)
    for (uint32_t d = 0; d < p; ++d) {
      const uint64_t a = B(r, d), b = B(r, d + 1);
      rep.add_bucket(b - a);
      if (b > a) lists[d].push_back(Run{a, b - a});
      nsizes[d] += b - a;
    }
  st = merge_lists(P, lists, offsets_of(nsizes));
  if (st != GS_OK) return st;
  sizes = nsizes;
  *rounds = 1;
  return GS_OK;
}

gs_status hyperquicksort_rounds(Pipe& P, std::vector<uint64_t>& sizes, uint32_t s, Report& rep, uint32_t* rounds) {
  const uint32_t p = (uint32_t)sizes.size();
  uint32_t n_p = p;
  while (n_p > 1) {
    const uint32_t groups = p / n_p;
    std::vector<uint64_t> off = offsets_of(sizes);
    std::vector<uint64_t> gtot(groups, 0);
    for (uint32_t r = 0; r < p; ++r) gtot[r / n_p] += sizes[r];
    // samples in rank order (unsorted: the pivot rule scans them in gather order)
    std::vector<uint64_t> idx, gofs{0};
    for (uint32_t g = 0; g < groups; ++g) {
      if (gtot[g])
        for (uint32_t r = g * n_p; r < (g + 1) * n_p; ++r) regular_samples(off[r], sizes[r], s, &idx);
      gofs.push_back(idx.size());
    }
    std::vector<char> sk, sv;
    {
      std::vector<uint64_t> one{0, 0};  // gather only, no sort
      gs_status st = gather_sorted(P, idx, one, &sk, &sv);
      if (st != GS_OK) return st;
    }
    // global lower_bound rank of every sample within its group
    std::vector<uint32_t> first(p);
    uint32_t maxm = 0;
    for (uint32_t g = 0; g < groups; ++g) maxm = std::max<uint32_t>(maxm, (uint32_t)(gofs[g + 1] - gofs[g]));
    std::vector<char> spl_k, spl_v;
    std::vector<uint64_t> pivots(groups, 0);
    std::vector<bool> active(groups, false);
    if (maxm) {
      // pad every group's sample list to maxm so one bounds launch serves all
      spl_k.assign((size_t)groups * maxm * P.kb, 0);
      spl_v.assign((size_t)groups * maxm * P.vb, 0);
      for (uint32_t g = 0; g < groups; ++g) {
        const uint64_t m = gofs[g + 1] - gofs[g];
        for (uint64_t x = 0; x < maxm; ++x) {
          const uint64_t srcx = gofs[g] + (m ? std::min<uint64_t>(x, m - 1) : 0);
          if (m == 0) continue;
          memcpy(spl_k.data() + ((size_t)g * maxm + x) * P.kb, sk.data() + srcx * P.kb, P.kb);
          if (P.kv) memcpy(spl_v.data() + ((size_t)g * maxm + x) * P.vb, sv.data() + srcx * P.vb, P.vb);
        }
      }
      for (uint32_t r = 0; r < p; ++r) first[r] = (r / n_p) * maxm;
      std::vector<uint64_t> bnd;
      gs_status st = run_bounds(P, off, spl_k, spl_v, first, maxm, false, &bnd);
      if (st != GS_OK) return st;
      for (uint32_t g = 0; g < groups; ++g) {
        const uint64_t m = gofs[g + 1] - gofs[g];
        if (m == 0 || gtot[g] == 0) continue;
        active[g] = true;
        const int64_t half = (int64_t)(gtot[g] / 2);
        uint64_t best = 0;
        int64_t best_rank = -1;
        for (uint64_t x = 0; x < m; ++x) {
          int64_t gr = 0;
          for (uint32_t r = g * n_p; r < (g + 1) * n_p; ++r)
            gr += (int64_t)(bnd[(size_t)r * (maxm + 2) + x + 1] - bnd[(size_t)r * (maxm + 2)]);
          if (x == 0 || std::llabs(gr - half) < std::llabs(best_rank - half)) {
            best = x;
            best_rank = gr;
          }
        }
        pivots[g] = gofs[g] + best;
      }
    }
    // pivot bounds per run
    std::vector<char> pk((size_t)groups * P.kb, 0), pv((size_t)groups * P.vb, 0);
    std::vector<char> rep_k;
    for (uint32_t g = 0; g < groups; ++g) {
      if (!active[g]) continue;
      memcpy(pk.data() + (size_t)g * P.kb, sk.data() + pivots[g] * P.kb, P.kb);
      if (P.kv) memcpy(pv.data() + (size_t)g * P.vb, sv.data() + pivots[g] * P.vb, P.vb);
      rep_k.insert(rep_k.end(), pk.data() + (size_t)g * P.kb, pk.data() + (size_t)(g + 1) * P.kb);
    }
    rep.add_splitters(rep_k);
    for (uint32_t r = 0; r < p; ++r) first[r] = r / n_p;
    std::vector<uint64_t> pb;
    gs_status st = run_bounds(P, off, pk, pv, first, 1, false, &pb);
    if (st != GS_OK) return st;
    std::vector<uint64_t> split(p);
    for (uint32_t r = 0; r < p; ++r) {
      split[r] = active[r / n_p] ? pb[(size_t)r * 3 + 1] : off[r + 1];
      rep.add_bucket(split[r] - off[r]);
      rep.add_bucket(off[r + 1] - split[r]);
    }
    // low half keeps [start, split), high half keeps [split, end); partner =
    // lr + cmp_p0 - new_p0 (hyperquicksort_helper.h:358-362); kept run first
    const uint32_t split_id = (n_p - 1) / 2;
    std::vector<uint64_t> nsizes(p, 0);
    std::vector<std::vector<Run>> lists(p);
    for (uint32_t g = 0; g < groups; ++g) {
      const uint32_t g0 = g * n_p;
      for (uint32_t lr = 0; lr < n_p; ++lr) {
        const uint32_t r = g0 + lr;
        if (!active[g]) {
          if (sizes[r]) lists[r].push_back(Run{off[r], sizes[r]});
          nsizes[r] = sizes[r];
          continue;
        }
        const bool low = lr <= split_id;
        const uint32_t new_p0 = low ? 0 : split_id + 1;
        const uint32_t cmp_p0 = low ? split_id + 1 : 0;
        uint32_t partner = lr + cmp_p0 - new_p0;
        if (partner >= n_p) partner = n_p - 1;
        const uint32_t q = g0 + partner;
        const Run kept = low ? Run{off[r], split[r] - off[r]} : Run{split[r], off[r + 1] - split[r]};
        const Run recv = low ? Run{off[q], split[q] - off[q]} : Run{split[q], off[q + 1] - split[q]};
        if (kept.len) lists[r].push_back(kept);
        if (recv.len) lists[r].push_back(recv);
        nsizes[r] = kept.len + recv.len;
      }
    }
    st = merge_lists(P, lists, offsets_of(nsizes));
    if (st != GS_OK) return st;
    sizes = nsizes;
    n_p /= 2;
    (*rounds)++;
  }
  return GS_OK;
}

}  // namespace

extern "C" gs_status gs_vrank_sort(gs_handle* h, gs_algorithm algo, gs_dtype kdt, void* d_keys, gs_dtype vdt,
                                   void* d_vals, const uint64_t* h_rank_off, uint32_t p, uint32_t k_way,
                                   uint32_t samples_per_rank, int descending, gs_vrank_report* report) {
  if (!h) return fail(GS_INVALID_ARG, "null handle");
  if (!dtype_ok(kdt)) return fail(GS_INVALID_ARG, "unsupported key type %d", (int)kdt);
  if (vdt != GS_NONE && !dtype_ok(vdt)) return fail(GS_INVALID_ARG, "unsupported value type %d", (int)vdt);
  if (p == 0 || !h_rank_off) return fail(GS_INVALID_ARG, "need p >= 1 rank offsets");
  if (algo != GS_HYKSORT && algo != GS_SAMPLESORT && algo != GS_HYPERQUICKSORT)
    return fail(GS_INVALID_ARG, "unknown algorithm %d", (int)algo);
  if (algo == GS_HYKSORT && k_way < 2) return fail(GS_INVALID_ARG, "k_way must be at least 2");
  if ((algo == GS_HYKSORT || algo == GS_HYPERQUICKSORT) && samples_per_rank == 0)
    return fail(GS_INVALID_ARG, "samples_per_rank must be at least 1");
  std::vector<uint64_t> sizes(p);
  for (uint32_t r = 0; r < p; ++r) {
    if (h_rank_off[r + 1] < h_rank_off[r]) return fail(GS_INVALID_ARG, "rank offsets decrease at rank %u", r);
    sizes[r] = h_rank_off[r + 1] - h_rank_off[r];
    // Every rank must hold data (hyksort_helper.h:105, samplesort_helper.h:69).
    if (sizes[r] == 0) return fail(GS_EMPTY_INPUT, "No data to sort.");
  }
  if (h_rank_off[0] != 0) return fail(GS_INVALID_ARG, "rank offsets must start at 0");
  const uint64_t n = h_rank_off[p];
  if (!d_keys || (vdt != GS_NONE && !d_vals)) return fail(GS_INVALID_ARG, "null pointer");
  // Power-of-two checks fall back to SampleSort, as HYKSORT_CHECK and
  // HYPERQUICKSORT_CHECK do (hyksort_common_vector.inl:31-43,
  // hyperquicksort_common_vector.inl:30-41).
  if (algo == GS_HYKSORT && (!is_pow2(p) || !is_pow2(k_way))) algo = GS_SAMPLESORT;
  if (algo == GS_HYPERQUICKSORT && !is_pow2(p)) algo = GS_SAMPLESORT;
  const bool small_samplesort = algo == GS_SAMPLESORT && p > 1 && n < 5ull * p * p;

  DeviceScope ds(h->device);
  Pipe P;
  P.h = h;
  P.kdt = kdt;
  P.vdt = vdt;
  P.kb = dtype_bytes(kdt);
  P.vb = val_bytes_of(vdt);
  P.kv = vdt != GS_NONE;
  P.desc = descending != 0;
  P.kx = make_xform((int)kdt, P.desc);
  P.vx = P.kv ? make_xform((int)vdt, P.desc) : XformParams{0, 0, 0};
  P.n = n;
  P.cur = 0;
  P.kbuf[0] = reinterpret_cast<char*>(d_keys);
  P.vbuf[0] = reinterpret_cast<char*>(d_vals);
  void *k1 = nullptr, *v1 = nullptr;
  hipError_t e = h->arena.get(Arena::kPipeKeys, n * (size_t)P.kb, h->stream, &k1);
  if (e == hipSuccess && P.kv) e = h->arena.get(Arena::kPipeVals, n * (size_t)P.vb, h->stream, &v1);
  if (e != hipSuccess) return fail(GS_OUT_OF_MEMORY, "Couldn't allocate device vector");
  P.kbuf[1] = reinterpret_cast<char*>(k1);
  P.vbuf[1] = reinterpret_cast<char*>(v1);

  if (report) {
    report->rounds = 0;
    report->n_splitters = 0;
    report->n_bucket_sizes = 0;
    report->max_rank_error = 0.0;
    report->splitter_iterations = 0;
  }
  Report rep{report, P.kb};

  // Local sort of every rank; lexicographic with a payload unless the
  // payloads are already non-decreasing in the requested order.
  bool lex_values = false;
  if (P.kv) {
    void* ws = nullptr;
    e = h->arena.get(Arena::kPipeAux, 256, h->stream, &ws);
    if (e != hipSuccess) return from_hip(e, "pipeline workspace");
    void* host = nullptr;
    e = host_pinned(h, 256, &host);
    if (e == hipSuccess) e = fill_u32(reinterpret_cast<uint32_t*>(ws), 1, 0u, h->stream);
    if (e == hipSuccess && h->timer()) h->timer()->begin("payload_order_check", n * (uint64_t)P.vb, h->stream);
    if (e == hipSuccess) e = unsorted_flag(P.vb, d_vals, n, P.vx, reinterpret_cast<uint32_t*>(ws), h->stream);
    if (e == hipSuccess && h->timer()) h->timer()->end(h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(host, ws, 4, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return from_hip(e, "payload order check");
    lex_values = *reinterpret_cast<const uint32_t*>(host) != 0;
  }
  {
    gs_status st = sort_runs(P, P.kbuf[0], P.kv ? P.vbuf[0] : nullptr,
                             std::vector<uint64_t>(h_rank_off, h_rank_off + p + 1), lex_values);
    if (st != GS_OK) return st;
  }

  uint32_t rounds = 0, iterations = 0;
  double max_err = 0.0;
  gs_status st = GS_OK;
  if (p > 1) {
    if (small_samplesort) {
      // No rebalance on this path: no rank ever runs empty (samplesort_helper.h:429-452).
      st = samplesort_small(P, sizes, &rounds);
      if (st != GS_OK) return st;
    } else {
      if (algo == GS_HYKSORT)
        st = hyksort_rounds(P, sizes, k_way, samples_per_rank, rep, &max_err, &rounds, &iterations);
      else if (algo == GS_SAMPLESORT)
        st = samplesort_round(P, sizes, rep, &rounds);
      else
        st = hyperquicksort_rounds(P, sizes, samples_per_rank, rep, &rounds);
      if (st != GS_OK) return st;
      keep_no_empty(sizes);
    }
  }
  if (P.cur != 0) {
    e = copy_bytes(P.kbuf[0], P.kbuf[1], n * (size_t)P.kb, h->stream);
    if (e == hipSuccess && P.kv) e = copy_bytes(P.vbuf[0], P.vbuf[1], n * (size_t)P.vb, h->stream);
    if (e != hipSuccess) return from_hip(e, "copy result");
  }
  e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return from_hip(e, "pipeline");
  if (report) {
    report->rounds = rounds;
    report->max_rank_error = max_err;
    report->splitter_iterations = iterations;
    if (report->final_rank_sizes)
      for (uint32_t r = 0; r < p; ++r) report->final_rank_sizes[r] = sizes[r];
  }
  return GS_OK;
}

extern "C" gs_status gs_vrank_sort_host(gs_handle* h, gs_algorithm algo, gs_dtype kdt, void* h_keys, gs_dtype vdt,
                                        void* h_vals, const uint64_t* h_rank_off, uint32_t p, uint32_t k_way,
                                        uint32_t samples_per_rank, int descending, gs_vrank_report* report) {
  if (!h) return fail(GS_INVALID_ARG, "null handle");
  if (!dtype_ok(kdt) || (vdt != GS_NONE && !dtype_ok(vdt))) return fail(GS_INVALID_ARG, "unsupported type");
  if (p == 0 || !h_rank_off) return fail(GS_INVALID_ARG, "need p >= 1 rank offsets");
  const uint64_t n = h_rank_off[p];
  if (n > 0 && (!h_keys || (vdt != GS_NONE && !h_vals))) return fail(GS_INVALID_ARG, "null pointer");
  DeviceScope ds(h->device);
  const size_t kbytes = n * (size_t)dtype_bytes(kdt), vbytes = n * (size_t)val_bytes_of(vdt);
  void *dk = nullptr, *dv = nullptr;
  if (h->arena.get(Arena::kStageKeys, kbytes, h->stream, &dk) != hipSuccess ||
      (vbytes && h->arena.get(Arena::kStageVals, vbytes, h->stream, &dv) != hipSuccess))
    return fail(GS_OUT_OF_MEMORY, "Couldn't allocate device vector");
  hipError_t e = hipMemcpyAsync(dk, h_keys, kbytes, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess && vbytes) e = hipMemcpyAsync(dv, h_vals, vbytes, hipMemcpyHostToDevice, h->stream);
  if (e != hipSuccess) return from_hip(e, "copy to device");
  gs_status st = gs_vrank_sort(h, algo, kdt, dk, vdt, dv, h_rank_off, p, k_way, samples_per_rank, descending, report);
  if (st != GS_OK) return st;
  e = hipMemcpyAsync(h_keys, dk, kbytes, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess && vbytes) e = hipMemcpyAsync(h_vals, dv, vbytes, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  return from_hip(e, "copy to host");
}

extern "C" gs_status gs_gather_host(gs_handle* h, int elem_bytes, const void* h_src, uint64_t n_src,
                                    const uint64_t* h_idx, void* h_dst, uint64_t n) {
  if (!h) return fail(GS_INVALID_ARG, "null handle");
  if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
    return fail(GS_INVALID_ARG, "element size %d not in {1,2,4,8}", elem_bytes);
  if (n == 0) return GS_OK;
  if (!h_src || !h_idx || !h_dst) return fail(GS_INVALID_ARG, "null pointer");
  for (uint64_t i = 0; i < n; ++i)
    if (h_idx[i] >= n_src) return fail(GS_INVALID_ARG, "index %llu out of range", (unsigned long long)h_idx[i]);
  DeviceScope ds(h->device);
  void *ds_src = nullptr, *ds_idx = nullptr, *ds_dst = nullptr;
  if (h->arena.get(Arena::kStageKeys, n_src * (size_t)elem_bytes, h->stream, &ds_src) != hipSuccess ||
      h->arena.get(Arena::kStagePairs, n * 8, h->stream, &ds_idx) != hipSuccess ||
      h->arena.get(Arena::kStageVals, n * (size_t)elem_bytes, h->stream, &ds_dst) != hipSuccess)
    return fail(GS_OUT_OF_MEMORY, "Couldn't allocate device vector");
  hipError_t e = hipMemcpyAsync(ds_src, h_src, n_src * (size_t)elem_bytes, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(ds_idx, h_idx, n * 8, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = gather_values(elem_bytes, ds_src, reinterpret_cast<const uint64_t*>(ds_idx), ds_dst, n, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(h_dst, ds_dst, n * (size_t)elem_bytes, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  return from_hip(e, "gather");
}

extern "C" gs_status gs_rebalance_sizes(uint64_t* sizes, uint32_t p) {
  if (p && !sizes) return fail(GS_INVALID_ARG, "null sizes");
  std::vector<uint64_t> s(sizes, sizes + p);
  keep_no_empty(s);
  for (uint32_t i = 0; i < p; ++i) sizes[i] = s[i];
  return GS_OK;
}
