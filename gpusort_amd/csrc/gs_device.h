// Device-side building blocks shared by every gfx950 kernel of the sort path:
// order-preserving key transforms, wave64 primitives (ballot match, scans) and
// agent-scope status-word access for decoupled look-back.
//
// All kernels work on "ordered" unsigned images of the keys: an unsigned
// integer whose unsigned order equals the requested order (less or greater)
// of the original key type.  This is how one integer kernel serves the six
// key types and two comparators the reference instantiates
// (src/cuda_utils.cu:249-318).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace gs {

constexpr int kWave = 64;

// --------------------------------------------------------------------------
// Key transform.
//
//   ordered(x) = x ^ (((sext(x) & float_mask) | sign_flip) ^ desc_mask)
//
//   unsigned : float_mask = 0,  sign_flip = 0
//   signed   : float_mask = 0,  sign_flip = MSB        (two's complement)
//   float    : float_mask = ~0, sign_flip = MSB        (IEEE-754 sign/magnitude)
//   greater  : desc_mask = ~0 (bitwise NOT reverses the unsigned order)
//
// sext(x) is x arithmetically shifted right by (bits - 1): all ones for a
// negative value, zero otherwise.  The map is a bijection, so sorting the
// images with an unsigned radix sort and mapping back gives the sorted keys.
// --------------------------------------------------------------------------
template <typename U>
struct Xform {
  U float_mask;
  U sign_flip;
  U desc_mask;
};

template <typename U>
__host__ __device__ inline U sext_msb(U x) {
  using S = typename std::conditional<sizeof(U) == 4, int32_t, int64_t>::type;
  return (U)((S)x >> (sizeof(U) * 8 - 1));
}

template <typename U>
__host__ __device__ inline U to_ordered(U x, const Xform<U>& t) {
  return x ^ (((sext_msb(x) & t.float_mask) | t.sign_flip) ^ t.desc_mask);
}

template <typename U>
__host__ __device__ inline U from_ordered(U u, const Xform<U>& t) {
  U y = u ^ t.desc_mask;
  return y ^ (t.sign_flip | (t.float_mask & ~sext_msb(y)));
}

template <typename U>
__device__ inline uint32_t digit_of(U ordered, int shift) {
  return (uint32_t)(ordered >> shift) & 0xFFu;
}

// --------------------------------------------------------------------------
// Wave64 primitives.
// --------------------------------------------------------------------------
__device__ inline uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ inline uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Number of set bits of m in lanes below the calling lane.
__device__ inline uint32_t count_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Mask of the lanes whose BITS-bit label equals the calling lane's label.
// Every lane of the wave must be active.  One ballot per label bit; the
// per-lane mask is accumulated as OR of (ballot XOR own-bit-broadcast), which
// is the set of lanes that differ in at least one bit, and then inverted.
template <int BITS>
__device__ inline uint64_t match_any(uint32_t label) {
  // Sign-extended bit b (all ones if set): one v_bfe_i32 per bit.  All
  // ballots are taken before they are consumed, and each half of the mask
  // folds in with one 3-input bit op: diff | (ballot ^ own) is truth table
  // 0xF6 with operands (diff, ballot, own).
  uint32_t own[BITS];
  uint64_t bal[BITS];
#pragma unroll
  for (int b = 0; b < BITS; ++b) own[b] = (uint32_t)((int32_t)(label << (31 - b)) >> 31);
#pragma unroll
  for (int b = 0; b < BITS; ++b) bal[b] = ballot((int32_t)own[b] < 0);
  uint32_t diff_lo = 0, diff_hi = 0;
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    diff_lo = __builtin_amdgcn_bitop3_b32(diff_lo, (uint32_t)bal[b], own[b], 0xF6);
    diff_hi = __builtin_amdgcn_bitop3_b32(diff_hi, (uint32_t)(bal[b] >> 32), own[b], 0xF6);
  }
  return ~(((uint64_t)diff_hi << 32) | diff_lo);
}

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ inline uint32_t wave_incl_scan(uint32_t x) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t y = __shfl_up(x, o, kWave);
    if (lane >= (uint32_t)o) x += y;
  }
  return x;
}

__device__ inline uint64_t wave_incl_scan64(uint64_t x) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint64_t y = __shfl_up(x, o, kWave);
    if (lane >= (uint32_t)o) x += y;
  }
  return x;
}

__device__ inline uint32_t wave_sum(uint32_t x) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
  return x;
}

// --------------------------------------------------------------------------
// Agent-scope (whole-GPU, across all XCD L2s) relaxed load/store of look-back
// status words.  A status word carries its payload itself, so relaxed order is
// enough: no other memory is published through it.
// --------------------------------------------------------------------------
template <typename S>
__device__ inline S status_load(const S* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename S>
__device__ inline void status_store(S* p, S v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Cache-allowed read of a status word.  A status word only ever moves
// 0 -> aggregate -> inclusive prefix, and every non-zero value it has held is
// correct, so a stale cached copy is usable unless it is 0; callers re-read a
// 0 with status_load.  Workgroup scope keeps it an atomic access (never
// hoisted out of a polling loop) that the L1/L2 may serve.
template <typename S>
__device__ inline S status_load_cached(const S* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Workgroup barrier that orders LDS only.  __syncthreads also waits for every
// outstanding global store of the calling wave to complete; the sort kernels
// share nothing through global memory inside a workgroup, so their barriers
// skip that wait and store latency overlaps the next phase.
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Fill kernel used instead of hipMemsetAsync so that every clear is a kernel
// node when a caller captures the stream in a graph.
__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v);

}  // namespace gs
