// Per-rank data files in the reference's on-disk format
// (src/test/utils.h:610-661, src/test/ifstream.h:23-76,
// src/test/ofstream.h:23-66): file <parent>/<rank><ext> holds an int64
// element count followed by the raw elements.  Keys-only data uses ".dat";
// key-value data stores keys in ".key" and value vector i (1-based) in ".v<i>"
// (src/test/utils.h:79-81).  SaveData cuts the input into np blocks of
// CalcNumElemsEachBlock(total, np) elements (include/common.h:146-154); for
// some totals that rule leaves the tail unwritten, as in the reference.
// ReadRanks / SaveRanks map the files of np ranks to and from one vector
// with a virtual communicator (mpi/mpi_common.h).
#ifndef GPUSORT_AMD_IO_RANK_FILES_H_
#define GPUSORT_AMD_IO_RANK_FILES_H_

#include <stdint.h>
#include <stdio.h>

#include <algorithm>
#include <iterator>
#include <string>
#include <vector>

#include "../exceptions/mpi_exception.h"
#include "../mpi/mpi_common.h"

namespace gpusort {
namespace io {

const std::string kKeyFileExt = ".key";
const std::string kValueFileExtPrefix = ".v";

inline int64_t CalcNumElemsEachBlock(int64_t total, int p) {
  const int64_t r = total % p;
  if (r == 0) return total / p;
  const int64_t tmp = (total + (p - r)) / p;
  return (tmp * (p - 1) >= total) ? tmp - 1 : tmp;
}

inline std::string GetFilePath(const std::string& parent_path, int rank, const std::string& ext = ".dat") {
  return parent_path + "/" + std::to_string(rank) + ext;
}

template <typename T>
void ReadData(int rank, const std::string& parent_path, const std::string& ext, std::vector<T>* out) {
  const std::string path = GetFilePath(parent_path, rank, ext);
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) throw MpiException("cannot open " + path);
  int64_t size = 0;
  bool ok = fread(&size, sizeof(size), 1, f) == 1 && size >= 0;
  if (ok) {
    out->resize((size_t)size);
    ok = size == 0 || fread(out->data(), sizeof(T), (size_t)size, f) == (size_t)size;
  }
  fclose(f);
  if (!ok) throw MpiException("truncated data file " + path);
}

template <typename RandomAccessIterator>
void SaveMyData(int rank, const std::string& parent_path, const std::string& ext, RandomAccessIterator first,
                RandomAccessIterator last) {
  typedef typename std::iterator_traits<RandomAccessIterator>::value_type T;
  const std::string path = GetFilePath(parent_path, rank, ext);
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) throw MpiException("cannot create " + path);
  const int64_t n = (int64_t)(last - first);
  bool ok = fwrite(&n, sizeof(n), 1, f) == 1;
  if (ok && n > 0) ok = fwrite(&*first, sizeof(T), (size_t)n, f) == (size_t)n;
  ok = (fclose(f) == 0) && ok;
  if (!ok) throw MpiException("cannot write " + path);
}

template <typename RandomAccessIterator>
void SaveData(int np, const std::string& parent_path, const std::string& ext, RandomAccessIterator first,
              RandomAccessIterator last) {
  const int64_t total = (int64_t)(last - first);
  const int64_t each = CalcNumElemsEachBlock(total, np);
  for (int i = 0; i < np; ++i) {
    const int64_t start = each * i;
    const int64_t end = std::min(start + each - 1, total - 1);
    const int64_t size = end - start + 1;
    SaveMyData(i, parent_path, ext, first + start, first + start + std::max<int64_t>(size, 0));
  }
}

// Reads the files of ranks 0..np-1 into one vector; the communicator holds
// the per-rank sizes.
template <typename T>
MpiComm ReadRanks(int np, const std::string& parent_path, const std::string& ext, std::vector<T>* all) {
  std::vector<uint64_t> sizes(np);
  all->clear();
  std::vector<T> part;
  for (int r = 0; r < np; ++r) {
    ReadData(r, parent_path, ext, &part);
    sizes[r] = part.size();
    all->insert(all->end(), part.begin(), part.end());
  }
  return MpiComm::Virtual(sizes);
}

// Writes every virtual rank's slice of `all` to its own file.
template <typename T>
void SaveRanks(const MpiComm& comm, const std::string& parent_path, const std::string& ext,
               const std::vector<T>& all) {
  const std::vector<uint64_t> off = comm.RankOffsets(all.size());
  for (int r = 0; r + 1 < (int)off.size(); ++r)
    SaveMyData(r, parent_path, ext, all.begin() + off[r], all.begin() + off[r + 1]);
}

}  // namespace io
}  // namespace gpusort

#endif  // GPUSORT_AMD_IO_RANK_FILES_H_
