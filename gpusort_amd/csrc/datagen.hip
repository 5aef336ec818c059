// Seeded synthetic inputs generated in HBM (gs_generate).
//
// The workload counterpart of the reference's data generators
// (src/test/utils.h:425-444, :586-595), which fill per-rank files; here the
// arrays are produced directly in device memory so that the large workloads
// (2^28 keys) need no host generation or transfer.  Draw i of stream `seed`
// is splitmix64's output for state seed + (i+1)*0x9E3779B97F4A7C15, the same
// counter-based stream as gpusort_amd/datagen.py, so the integer
// distributions are bit-identical to the host generator.
#include <hip/hip_runtime.h>

#include "../../include/gpusort_hip.h"
#include "gs_capi_impl.h"
#include "gs_internal.h"

namespace gs {

__device__ inline uint64_t draw64(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}

enum Dist : int {
  kUniformU32 = 0,
  kUniformU64 = 1,
  kZipfU32 = 2,
  kSortedU32 = 3,
  kEqualU32 = 4,
  kNormalF32 = 5,
  kIotaU32 = 6,
  kIotaU64 = 7
};

__global__ __launch_bounds__(256) void generate_kernel(int dist, uint64_t seed, void* out, uint64_t n, uint64_t base,
                                                       double param, const double* __restrict__ cdf,
                                                       uint32_t cdf_n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    const uint64_t g = base + i;
    switch (dist) {
      case kUniformU32:
        reinterpret_cast<uint32_t*>(out)[i] = (uint32_t)(draw64(seed, g) >> 32);
        break;
      case kUniformU64:
        reinterpret_cast<uint64_t*>(out)[i] = draw64(seed, g);
        break;
      case kZipfU32: {
        const double u = (double)(draw64(seed, g) >> 11) * (1.0 / 9007199254740992.0);
        // searchsorted(cdf, u, side="right"): first index with cdf[idx] > u.
        uint32_t lo = 0, hi = cdf_n;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (cdf[mid] <= u)
            lo = mid + 1;
          else
            hi = mid;
        }
        uint32_t rank = lo + 1;
        if (rank > cdf_n) rank = cdf_n;
        reinterpret_cast<uint32_t*>(out)[i] = mix32(rank);
        break;
      }
      case kSortedU32: {
        const uint64_t step = (uint64_t)param;
        reinterpret_cast<uint32_t*>(out)[i] = (uint32_t)(g * step + draw64(seed, g) % step);
        break;
      }
      case kEqualU32:
        reinterpret_cast<uint32_t*>(out)[i] = (uint32_t)(draw64(seed, 0) >> 32);
        break;
      case kNormalF32: {
        const uint64_t d1 = draw64(seed, 2 * g), d2 = draw64(seed, 2 * g + 1);
        const double u1 = ((double)(d1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
        const double u2 = (double)(d2 >> 11) * (1.0 / 9007199254740992.0);
        float f = (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2) * param);
        if (f == 0.0f) f = 0.0f;
        reinterpret_cast<float*>(out)[i] = f;
        break;
      }
      case kIotaU32:
        reinterpret_cast<uint32_t*>(out)[i] = (uint32_t)g;
        break;
      case kIotaU64:
        reinterpret_cast<uint64_t*>(out)[i] = g;
        break;
    }
  }
}

}  // namespace gs

using namespace gs;

extern "C" gs_status gs_generate(gs_handle* h, int dist, uint64_t seed, void* d_out, uint64_t n, uint64_t index_base,
                                 double param, const double* h_cdf, uint32_t cdf_n) {
  if (!h) return fail(GS_INVALID_ARG, "null handle");
  if (dist < kUniformU32 || dist > kIotaU64) return fail(GS_INVALID_ARG, "unknown distribution %d", dist);
  if (n == 0) return GS_OK;
  if (!d_out) return fail(GS_INVALID_ARG, "null output");
  if (dist == kZipfU32 && (!h_cdf || cdf_n == 0)) return fail(GS_INVALID_ARG, "zipf needs a CDF table");
  if (dist == kSortedU32 && param < 1.0) return fail(GS_INVALID_ARG, "sorted needs a step >= 1");
  DeviceScope ds(h->device);
  const double* d_cdf = nullptr;
  if (dist == kZipfU32) {
    void* ws = nullptr;
    hipError_t e = h->arena.get(Arena::kSmall, cdf_n * sizeof(double), h->stream, &ws);
    if (e != hipSuccess) return from_hip(e, "cdf table");
    e = hipMemcpyAsync(ws, h_cdf, cdf_n * sizeof(double), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return from_hip(e, "cdf upload");
    d_cdf = reinterpret_cast<const double*>(ws);
  }
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(generate_kernel, dim3((uint32_t)blocks), dim3(256), 0, h->stream, dist, seed, d_out, n,
                     index_base, param, d_cdf, cdf_n);
  return from_hip(hipGetLastError(), "generate");
}
