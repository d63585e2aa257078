// Deterministic FP64 force reduction used by the step kernels.
//
// Replaces the reference's single-workgroup FP32 reduce pass (src/shaders/aux.ts:144-172,
// src/shaders/aux3d.ts:155-183) with a fused two-level reduction: every wave of the group sums its
// cells' momentum-exchange forces in double over a fixed tree of lane permutes (wave_sum) and
// publishes one partial, and the last block to finish (completion counter) sums the partials in
// wave order and appends the sample to the history ring. When one block holds every node of the sample, the
// partials meet in LDS instead and that block appends the sample, with the same summation order.
// The summation order is fixed, so a sample is bit-reproducible run to run.
//
// Publication uses device-scope (agent) atomic stores and loads, which are coherent across the
// XCDs' separate L2 caches, and a wait for the stores' completion (s_waitcnt vmcnt(0)) before the
// counter increment. This ordering rests on the hardware's wait, not on a release/acquire pair of
// the HIP memory model: a device-scope release fence would say it in the model's terms but writes
// back every dirty L2 line of the issuing XCD, i.e. much of what the step just stored, once per
// block. tests/test_gpu_parity.py::test_force_reduction_across_many_blocks runs the multi-block path
// over 13 to 54 blocks and checks every sample against the oracle and for identical bits in two
// runs, so a partial read before it landed would show there.
#pragma once
#include "lbm_device.hpp"

namespace lbm {

// v + (v of lane CTRL's DPP source): the two 32-bit halves of the double moved by data-parallel
// primitive lane permutes, which stay in the vector ALU (a __shfl_xor is an LDS permute round trip
// per half).
template <int CTRL>
__device__ __forceinline__ double dpp_add(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = uint32_t(__builtin_amdgcn_update_dpp(0, int(uint32_t(b)), CTRL, 0xF, 0xF, false));
  const uint32_t hi = uint32_t(__builtin_amdgcn_update_dpp(0, int(uint32_t(b >> 32)), CTRL, 0xF, 0xF, false));
  return v + __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}
__device__ __forceinline__ double lane_value(double v, int lane) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b)), lane));
  const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b >> 32)), lane));
  return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// Sum over the 64 lanes of the wave, the same tree on every call: lanes l and l ^ 1, then the two
// pairs of a quad, then lanes i and 7 - i of each half row (two quads), lanes i and 15 - i of each
// row of 16 (two half rows), then the four row sums in row order, (r0 + r1) + (r2 + r3). Every
// lane receives the sum. All 64 lanes must be active.
template <int DIM>
__device__ __forceinline__ void wave_sum(double (&v)[DIM]) {
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    double s = dpp_add<0xB1>(v[d]);  // quad_perm [1, 0, 3, 2]
    s = dpp_add<0x4E>(s);            // quad_perm [2, 3, 0, 1]
    s = dpp_add<0x141>(s);           // row_half_mirror
    s = dpp_add<0x140>(s);           // row_mirror
    v[d] = (lane_value(s, 0) + lane_value(s, 16)) + (lane_value(s, 32) + lane_value(s, 48));
  }
}

// Every thread of every block of the group must call this (no early return before it). A group is
// a contiguous run of blocks whose threads hold the cells of the sample in order (the body-link
// nodes of one tunnel): block is this block's index inside the group, nblocks the group's size.
// Partials are per wave (64 consecutive nodes), and the last block sums them in wave order with a
// fixed lane assignment, so the sample does not depend on the block size.
// st0: the ring's sample count read by the caller before its own loads (only the appending thread of
// the sample writes it, so it cannot change within the launch), or ~0u to read it here.
template <int DIM>
__device__ void force_sample(const float (&force)[DIM], const ForceSink& fs, uint32_t block, uint32_t nblocks,
                             uint32_t st0 = ~0u) {
  __shared__ int is_last;
  __shared__ double part[16][DIM];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, wpb = blockDim.x >> 6;
  double v[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) v[d] = double(force[d]);
  wave_sum<DIM>(v);
  if (nblocks == 1) {
    // One block holds every node of the sample: its waves' partials meet in LDS and its first wave
    // sums them exactly as the last block does below (lane l takes partial l, then wave_sum),
    // without the round trips through device memory and the completion counter.
    if (lane == 0)
#pragma unroll
      for (int d = 0; d < DIM; d++) part[wave][d] = v[d];
    __syncthreads();
    if (wave != 0) return;
    double acc[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) acc[d] = 0.0 + (lane < wpb ? part[lane][d] : 0.0);
    wave_sum<DIM>(acc);
    if (lane == 0) {
      const uint32_t st = st0 != ~0u ? st0 : __hip_atomic_load(fs.state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t slot = st % HISTORY_LEN;
#pragma unroll
      for (int d = 0; d < DIM; d++) fs.ring[size_t(slot) * DIM + d] = acc[d];
      note_divergence<DIM>(acc, fs.state, fs.step);
      __hip_atomic_store(fs.state, st + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return;
  }
  if (lane == 0) {
    const size_t part = size_t(block) * wpb + wave;
#pragma unroll
    for (int d = 0; d < DIM; d++)
      __hip_atomic_store(fs.partials + part * DIM + d, v[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's partial has landed before the block reports
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t prev = __hip_atomic_fetch_add(fs.counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    is_last = (prev == nblocks - 1);
  }
  __syncthreads();
  if (!is_last || wave != 0) return;
  const uint32_t nparts = nblocks * wpb;
  double acc[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) acc[d] = 0.0;
  for (uint32_t k = lane; k < nparts; k += 64u)
#pragma unroll
    for (int d = 0; d < DIM; d++)
      acc[d] += __hip_atomic_load(fs.partials + size_t(k) * DIM + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  wave_sum<DIM>(acc);
  if (lane == 0) {
    const uint32_t st = __hip_atomic_load(fs.state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t slot = st % HISTORY_LEN;
#pragma unroll
    for (int d = 0; d < DIM; d++) fs.ring[size_t(slot) * DIM + d] = acc[d];
    note_divergence<DIM>(acc, fs.state, fs.step);
    __hip_atomic_store(fs.state, st + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(fs.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace lbm
