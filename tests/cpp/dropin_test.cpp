// C++ drop-in test: calls the reference's API names (HykSort, SampleSort,
// HyperQuickSort, *WithGlobalIndexes, CudaUtils, ParUtils) from the headers
// under include/gpusort and checks every result against std::sort /
// std::stable_sort on the host.
//
//   dropin_test nodevice   no GPU visible: the reference's "no device"
//                          exceptions are thrown and nothing is sorted
//   dropin_test            GPU visible: every case below must pass
//
// Prints one line per case; exit status 0 when all pass.
#include <gpusort/cuda/cuda_utils.h>
#include <gpusort/hyksort.h>
#include <gpusort/hyperquicksort.h>
#include <gpusort/io/rank_files.h>
#include <gpusort/par/par_utils.h>
#include <gpusort/samplesort.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <numeric>
#include <string>
#include <utility>
#include <vector>

using gpusort::MpiComm;

namespace {

int g_failures = 0;

void Report(const char* name, bool ok, const std::string& why = "") {
  std::printf("%s %s%s%s\n", ok ? "PASS" : "FAIL", name, why.empty() ? "" : ": ", why.c_str());
  std::fflush(stdout);
  if (!ok) ++g_failures;
}

// splitmix64 draw i of stream `seed` (the project's data generator).
uint64_t Draw(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <class T>
std::vector<T> Keys(uint64_t seed, size_t n, uint64_t modulus = 0) {
  std::vector<T> v(n);
  for (size_t i = 0; i < n; ++i) {
    uint64_t d = Draw(seed, i);
    if (modulus) d %= modulus;
    if (std::is_floating_point<T>::value) {
      // finite, no zeros: ties of -0.0 / +0.0 are unordered for std::sort
      v[i] = static_cast<T>((static_cast<double>(d >> 11) - 4503599627370496.0) / 1024.0 + 0.5);
    } else {
      T t;
      memcpy(&t, &d, sizeof(T));
      v[i] = t;
    }
  }
  return v;
}

template <class T>
bool SameBits(const std::vector<T>& a, const std::vector<T>& b) {
  return a.size() == b.size() && (a.empty() || memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

std::vector<uint64_t> Sizes(size_t p, size_t n, uint64_t seed) {
  // ragged rank sizes that add up to n, every rank non-empty
  std::vector<uint64_t> s(p, 1);
  for (size_t i = p; i < n; ++i) s[Draw(seed, i) % p]++;
  return s;
}

bool Partitioned(const MpiComm& comm, size_t n) {
  uint64_t t = 0;
  for (uint64_t s : comm.Sizes()) {
    if (s == 0) return false;
    t += s;
  }
  return t == n;
}

template <class T, class Comp>
void CaseCudaUtilsSort(const char* name, Comp comp) {
  std::vector<T> v = Keys<T>(11, 100003);
  std::vector<T> want = v;
  std::sort(want.begin(), want.end(), comp);
  gpusort::CudaUtils::Sort(comp, 0, v.begin(), v.end());
  Report(name, SameBits(v, want));
}

void CaseSortByKey() {
  const size_t n = 70001;
  std::vector<uint32_t> k = Keys<uint32_t>(3, n, 1000);
  std::vector<uint64_t> v(n);
  std::iota(v.begin(), v.end(), 0);
  std::vector<std::pair<uint32_t, uint64_t>> want(n);
  for (size_t i = 0; i < n; ++i) want[i] = {k[i], v[i]};
  std::stable_sort(want.begin(), want.end(),
                   [](const std::pair<uint32_t, uint64_t>& a, const std::pair<uint32_t, uint64_t>& b) {
                     return a.first > b.first;
                   });
  std::vector<std::pair<uint32_t, uint64_t>> out(n);
  gpusort::CudaUtils::SortByKey(std::greater<uint32_t>(), 0, k.begin(), k.end(), v.begin(), out.data());
  bool ok = out == want;
  for (size_t i = 0; ok && i < n; ++i) ok = k[i] == want[i].first && v[i] == want[i].second;
  Report("cudautils_sort_by_key_desc_stable_pairs_out", ok);
}

void CaseHykSortKeys() {
  const size_t p = 16, n = 1 << 18;
  std::vector<uint64_t> k = Keys<uint64_t>(5, n);
  std::vector<uint64_t> want = k;
  std::sort(want.begin(), want.end());
  MpiComm comm = MpiComm::Virtual(Sizes(p, n, 6));
  gpusort::HykSort(comm, 4, true, &k);
  Report("hyksort_u64_p16_k4", SameBits(k, want) && Partitioned(comm, n) && comm.LastRounds() == 2,
         "rounds=" + std::to_string(comm.LastRounds()));
}

void CaseHykSortPairsLexicographic() {
  const size_t p = 8, n = 200000;
  std::vector<uint32_t> k = Keys<uint32_t>(7, n, 5000);  // many ties
  std::vector<int32_t> v = Keys<int32_t>(8, n);
  std::vector<std::pair<uint32_t, int32_t>> want(n);
  for (size_t i = 0; i < n; ++i) want[i] = {k[i], v[i]};
  std::sort(want.begin(), want.end(), std::greater<std::pair<uint32_t, int32_t>>());
  MpiComm comm = MpiComm::Virtual(Sizes(p, n, 9));
  gpusort::HykSort(comm, std::greater<uint32_t>(), 2, true, &k, &v);
  bool ok = Partitioned(comm, n);
  for (size_t i = 0; ok && i < n; ++i) ok = k[i] == want[i].first && v[i] == want[i].second;
  Report("hyksort_pairs_desc_lexicographic_p8_k2", ok);
}

// One rank, tied keys whose payloads run against the order: with is_using_gpu the
// reference's HykSort runs CudaUtils::SortByKey (stable by key: equal keys
// keep their input order, include/helpers/hyksort_helper.h:247-261); without
// it OmpUtils::MergeSort on (key, value) pairs (lexicographic, :262-270).
// SampleSort never takes the GPU branch, so it stays lexicographic.
template <class Comp>
void CaseOneRankPairOrder(const char* name, Comp comp, bool gpu) {
  const size_t n = 60000;
  std::vector<int32_t> k = Keys<int32_t>(25, n, 700);
  // Payloads run against the direction (decreasing for ascending keys,
  // increasing for descending), so equal keys come out in different orders
  // under the two rules.
  const bool asc = std::is_same<Comp, std::less<int32_t>>::value;
  std::vector<uint32_t> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = static_cast<uint32_t>(asc ? n - i : i + 1);
  std::vector<std::pair<int32_t, uint32_t>> stable(n), lexi(n);
  for (size_t i = 0; i < n; ++i) stable[i] = lexi[i] = {k[i], v[i]};
  std::stable_sort(stable.begin(), stable.end(), [&](const std::pair<int32_t, uint32_t>& a,
                                                     const std::pair<int32_t, uint32_t>& b) { return comp(a.first, b.first); });
  if (std::is_same<Comp, std::less<int32_t>>::value)
    std::sort(lexi.begin(), lexi.end(), std::less<std::pair<int32_t, uint32_t>>());
  else
    std::sort(lexi.begin(), lexi.end(), std::greater<std::pair<int32_t, uint32_t>>());
  const std::vector<std::pair<int32_t, uint32_t>>& want = gpu ? stable : lexi;
  std::vector<int32_t> k2 = k;
  std::vector<uint32_t> v2 = v;
  MpiComm comm;  // one rank
  gpusort::HykSort(comm, comp, 4, gpu, &k, &v);
  bool ok = stable != lexi;  // the two orders really differ on this input
  for (size_t i = 0; ok && i < n; ++i) ok = k[i] == want[i].first && v[i] == want[i].second;
  Report(name, ok);
  if (!gpu) return;
  gpusort::SampleSort(MpiComm(), comp, &k2, &v2);  // one rank, CPU branch only in the reference
  std::vector<std::pair<int32_t, uint32_t>> lex(n);
  for (size_t i = 0; i < n; ++i) lex[i] = {k2[i], v2[i]};
  const bool sorted = std::is_same<Comp, std::less<int32_t>>::value
                          ? std::is_sorted(lex.begin(), lex.end(), std::less<std::pair<int32_t, uint32_t>>())
                          : std::is_sorted(lex.begin(), lex.end(), std::greater<std::pair<int32_t, uint32_t>>());
  Report((std::string(name) + "_samplesort_lexicographic").c_str(), sorted);
}

struct Record {
  double weight;
  int32_t id;
  char tag[12];
};

void CaseHykSortVariadic() {
  const size_t p = 4, n = 50000;
  std::vector<float> k = Keys<float>(12, n);
  for (size_t i = 0; i < n; i += 7) k[i] = k[(i * 31) % n];  // ties
  std::vector<double> v1(n);
  std::vector<Record> v2(n);
  for (size_t i = 0; i < n; ++i) {
    v1[i] = static_cast<double>(i) * 0.5;
    v2[i].weight = -static_cast<double>(i);
    v2[i].id = static_cast<int32_t>(i);
    std::snprintf(v2[i].tag, sizeof(v2[i].tag), "r%zu", i);
  }
  std::vector<size_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return k[a] < k[b]; });
  const std::vector<float> k0 = k;
  MpiComm comm = MpiComm::Virtual(Sizes(p, n, 13));
  gpusort::HykSort(comm, 2, true, &k, &v1, &v2);
  bool ok = Partitioned(comm, n);
  for (size_t i = 0; ok && i < n; ++i) {
    const size_t j = order[i];
    char tag[12];
    std::snprintf(tag, sizeof(tag), "r%zu", j);
    ok = k[i] == k0[j] && v1[i] == static_cast<double>(j) * 0.5 && v2[i].id == static_cast<int32_t>(j) &&
         strcmp(v2[i].tag, tag) == 0;
  }
  Report("hyksort_variadic_two_payload_vectors_stable", ok);
}

void CaseHykSortKeyFunc() {
  const size_t n = 30000;
  std::vector<Record> list(n);
  for (size_t i = 0; i < n; ++i) {
    list[i].weight = static_cast<double>(Draw(14, i) % 997) - 500.0;
    list[i].id = static_cast<int32_t>(i);
    list[i].tag[0] = 0;
  }
  std::vector<Record> want = list;
  std::stable_sort(want.begin(), want.end(), [](const Record& a, const Record& b) { return a.weight < b.weight; });
  MpiComm comm;  // one rank
  gpusort::HykSort<double>(comm, 8, true, [](const Record& r) { return r.weight; }, &list);
  bool ok = true;
  for (size_t i = 0; ok && i < n; ++i) ok = list[i].id == want[i].id && list[i].weight == want[i].weight;
  Report("hyksort_keyfunc_user_type_one_rank", ok);
}

// A user element type over several ranks: the key function's values are
// sorted through the virtual ranks' rounds with their global index, so equal
// keys keep input order across ranks as well.
void CaseHykSortKeyFuncRanks(const char* name, size_t p, unsigned k_way) {
  const size_t n = 40000;
  std::vector<Record> list(n);
  for (size_t i = 0; i < n; ++i) {
    list[i].weight = static_cast<double>(Draw(21, i) % 1511) - 700.0;
    list[i].id = static_cast<int32_t>(i);
    std::snprintf(list[i].tag, sizeof(list[i].tag), "e%zu", i);
  }
  std::vector<Record> want = list;
  std::stable_sort(want.begin(), want.end(), [](const Record& a, const Record& b) { return a.weight < b.weight; });
  MpiComm comm = MpiComm::Virtual(Sizes(p, n, 22));
  gpusort::HykSort<double>(comm, k_way, true, [](const Record& r) { return r.weight; }, &list);
  bool ok = Partitioned(comm, n) && comm.LastRounds() >= 1;
  for (size_t i = 0; ok && i < n; ++i)
    ok = list[i].id == want[i].id && list[i].weight == want[i].weight && strcmp(list[i].tag, want[i].tag) == 0;
  Report(name, ok);
}

void CaseNonPow2FallsBackToSampleSort() {
  const size_t p = 6, n = 60000;
  std::vector<int64_t> k = Keys<int64_t>(15, n);
  std::vector<int64_t> want = k;
  std::sort(want.begin(), want.end());
  MpiComm comm = MpiComm::Virtual(Sizes(p, n, 16));
  gpusort::HykSort(comm, 4, true, &k);
  Report("hyksort_p6_runs_samplesort", SameBits(k, want) && Partitioned(comm, n) && comm.LastRounds() == 1);
}

void CaseSampleSortPairs() {
  const size_t p = 8, n = 100000;
  std::vector<double> k = Keys<double>(17, n);
  std::vector<uint32_t> v(n);
  std::iota(v.begin(), v.end(), 0u);
  std::vector<std::pair<double, uint32_t>> want(n);
  for (size_t i = 0; i < n; ++i) want[i] = {k[i], v[i]};
  std::sort(want.begin(), want.end());
  MpiComm comm = MpiComm::Virtual(Sizes(p, n, 18));
  gpusort::SampleSort(comm, &k, &v);
  bool ok = Partitioned(comm, n) && comm.LastSplitters().size() == p - 1;
  for (size_t i = 0; ok && i < n; ++i) ok = k[i] == want[i].first && v[i] == want[i].second;
  Report("samplesort_f64_pairs_p8", ok);
}

void CaseSampleSortSmallInput() {
  // Fewer than 5 p^2 keys: the bitonic merge-split network, rank sizes set by the data.
  const size_t p = 8, n = 200;
  std::vector<uint32_t> k = Keys<uint32_t>(23, n, 50);
  std::vector<uint32_t> v(n);
  std::iota(v.begin(), v.end(), 0u);
  std::vector<std::pair<uint32_t, uint32_t>> want(n);
  for (size_t i = 0; i < n; ++i) want[i] = {k[i], v[i]};
  std::sort(want.begin(), want.end());
  MpiComm comm = MpiComm::Virtual(Sizes(p, n, 24));
  gpusort::SampleSort(comm, &k, &v);
  bool ok = Partitioned(comm, n) && comm.LastRounds() > 0;
  for (size_t i = 0; ok && i < n; ++i) ok = k[i] == want[i].first && v[i] == want[i].second;
  Report("samplesort_small_input_bitonic_p8", ok);
}

void CaseHyperQuickSort() {
  const size_t p = 8, n = 120000;
  std::vector<int32_t> k = Keys<int32_t>(19, n);
  std::vector<int32_t> want = k;
  std::sort(want.begin(), want.end(), std::greater<int32_t>());
  MpiComm comm = MpiComm::Virtual(Sizes(p, n, 20));
  gpusort::HyperQuickSort(comm, std::greater<int32_t>(), &k);
  Report("hyperquicksort_i32_desc_p8", SameBits(k, want) && Partitioned(comm, n) && comm.LastRounds() == 3);
}

void CaseGlobalIndexesAndArrange() {
  const size_t p = 4, n = 40000;
  std::vector<uint32_t> k = Keys<uint32_t>(21, n, 300);
  const std::vector<uint32_t> k0 = k;
  MpiComm comm = MpiComm::Virtual(Sizes(p, n, 22));
  std::vector<int64_t> g;
  std::vector<int64_t> old_sizes;
  gpusort::ParUtils::GatherNElems(comm, static_cast<int64_t>(n), &old_sizes);
  gpusort::HykSortWithGlobalIndexes(comm, std::less<uint32_t>(), 2, true, old_sizes, &k, &g);
  bool ok = g.size() == n;
  for (size_t i = 0; ok && i < n; ++i) ok = k[i] == k0[g[i]] && (i == 0 || k[i - 1] < k[i] || g[i - 1] < g[i]);
  // ArrangeValues known answer: values[i] = old[g[i]]
  std::vector<uint16_t> vals(n);
  for (size_t i = 0; i < n; ++i) vals[i] = static_cast<uint16_t>(i * 7);
  gpusort::ParUtils::ArrangeValues(comm, g, old_sizes, &vals);
  for (size_t i = 0; ok && i < n; ++i) ok = vals[i] == static_cast<uint16_t>(g[i] * 7);
  Report("global_indexes_and_arrange_values", ok);
}

void CaseKeepNoProcessHasEmptyData() {
  // Known answers of the reference's rule (include/par/par_utils.h:103-175):
  // rank 0 takes half of rank 1; the loop then skips past rank 2 and rank 3
  // (last) takes nelems[2] / 2 = 0 of the sizes gathered before the call.
  std::vector<uint32_t> k(10, 7u);
  MpiComm a = MpiComm::Virtual({0, 10, 0, 0});
  gpusort::ParUtils::KeepNoProcessHasEmptyData(a, &k);
  MpiComm b = MpiComm::Virtual({10, 0, 0, 0});
  gpusort::ParUtils::KeepNoProcessHasEmptyData(b, &k);
  MpiComm c = MpiComm::Virtual({3, 0, 4, 3});
  gpusort::ParUtils::KeepNoProcessHasEmptyData(c, &k);
  const bool ok = a.Sizes() == std::vector<uint64_t>({5, 5, 0, 0}) &&
                  b.Sizes() == std::vector<uint64_t>({4, 2, 2, 2}) &&
                  c.Sizes() == std::vector<uint64_t>({3, 2, 2, 3});
  Report("keep_no_process_has_empty_data", ok);
}

template <class Exc, class F>
bool Throws(F f, const char* what) {
  try {
    f();
  } catch (const Exc& e) {
    return std::string(e.what()) == what;
  } catch (...) {
    return false;
  }
  return false;
}

void CaseEmptyInput() {
  std::vector<uint32_t> k;
  bool ok = Throws<gpusort::ThrustException>([&] { gpusort::HykSort(MpiComm(), 2, true, &k); }, "No data to sort.");
  ok = ok && Throws<gpusort::MpiException>([&] { gpusort::SampleSort(MpiComm(), &k); }, "No data to sort.");
  ok = ok && Throws<gpusort::MpiException>([&] { gpusort::HyperQuickSort(MpiComm(), &k); }, "No data to sort.");
  Report("empty_input_throws_reference_exceptions", ok);
}

// Rank files (no device needed): SaveData's block rule, ReadRanks and
// SaveRanks round trips; the directory is then read by the Python reader.
int RankFiles(const std::string& dir) {
  namespace io = gpusort::io;
  std::vector<uint32_t> a(10);
  std::iota(a.begin(), a.end(), 100u);
  io::SaveData(4, dir, ".dat", a.begin(), a.end());  // blocks of 3: 3, 3, 3, 1
  std::vector<uint32_t> back;
  MpiComm c = io::ReadRanks(4, dir, ".dat", &back);
  Report("rankfiles_save_data_blocks", back == a && c.Sizes() == std::vector<uint64_t>({3, 3, 3, 1}));
  std::vector<double> d = {0.5, -1.0, 2.25, 8.0, 3.5};
  MpiComm vc = MpiComm::Virtual({2, 0, 3});
  io::SaveRanks(vc, dir, ".v1", d);
  std::vector<double> d2;
  MpiComm c2 = io::ReadRanks(3, dir, ".v1", &d2);
  Report("rankfiles_save_ranks_round_trip", d2 == d && c2.Sizes() == vc.Sizes());
  std::vector<int64_t> b9(9, 7);
  io::SaveData(4, dir, ".nine", b9.begin(), b9.end());  // the reference's rule writes 2 per rank
  std::vector<int64_t> r9;
  MpiComm c9 = io::ReadRanks(4, dir, ".nine", &r9);
  Report("rankfiles_block_rule_tail", r9.size() == 8 && io::CalcNumElemsEachBlock(9, 4) == 2);
  bool threw = false;
  try {
    io::ReadData(9, dir, ".dat", &back);
  } catch (const gpusort::MpiException&) {
    threw = true;
  }
  Report("rankfiles_missing_file_throws", threw);
  return g_failures ? 1 : 0;
}

int NoDevice() {
  const char* msg = "No CUDA-capable devices were detected";
  Report("no_device_count", gpusort::CudaUtils::GetDeviceCount() == 0);
  Report("no_device_is_having_devices_throws",
         Throws<gpusort::ThrustException>([] { gpusort::CudaUtils::IsHavingDevices(); }, msg));
  std::vector<uint32_t> k = {3, 1, 2};
  Report("no_device_hyksort_throws", Throws<gpusort::ThrustException>([&] { gpusort::HykSort(MpiComm(), 2, true, &k); }, msg));
  Report("no_device_input_untouched", k[0] == 3 && k[1] == 1 && k[2] == 2);
  return g_failures ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc > 1 && std::string(argv[1]) == "nodevice") return NoDevice();
  if (argc > 2 && std::string(argv[1]) == "rankfiles") return RankFiles(argv[2]);
  CaseCudaUtilsSort<uint32_t>("cudautils_sort_u32_less", std::less<uint32_t>());
  CaseCudaUtilsSort<int32_t>("cudautils_sort_i32_greater", std::greater<int32_t>());
  CaseCudaUtilsSort<float>("cudautils_sort_f32_less", std::less<float>());
  CaseCudaUtilsSort<double>("cudautils_sort_f64_greater", std::greater<double>());
  CaseCudaUtilsSort<int64_t>("cudautils_sort_i64_less", std::less<int64_t>());
  CaseCudaUtilsSort<uint64_t>("cudautils_sort_u64_greater", std::greater<uint64_t>());
  CaseSortByKey();
  CaseHykSortKeys();
  CaseHykSortPairsLexicographic();
  CaseOneRankPairOrder("hyksort_one_rank_gpu_stable_by_key_asc", std::less<int32_t>(), true);
  CaseOneRankPairOrder("hyksort_one_rank_gpu_stable_by_key_desc", std::greater<int32_t>(), true);
  CaseOneRankPairOrder("hyksort_one_rank_cpu_lexicographic_asc", std::less<int32_t>(), false);
  CaseOneRankPairOrder("hyksort_one_rank_cpu_lexicographic_desc", std::greater<int32_t>(), false);
  CaseHykSortVariadic();
  CaseHykSortKeyFunc();
  CaseHykSortKeyFuncRanks("hyksort_keyfunc_user_type_p8_k2", 8, 2);
  CaseHykSortKeyFuncRanks("hyksort_keyfunc_user_type_p16_k4", 16, 4);
  CaseNonPow2FallsBackToSampleSort();
  CaseSampleSortPairs();
  CaseHyperQuickSort();
  CaseSampleSortSmallInput();
  CaseGlobalIndexesAndArrange();
  CaseKeepNoProcessHasEmptyData();
  CaseEmptyInput();
  std::printf("%d failure(s)\n", g_failures);
  return g_failures ? 1 : 0;
}
