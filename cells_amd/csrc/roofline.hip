// Copy roofline: the memory traffic of one step with no arithmetic (the method of src/bench.ts:61-95:
// read every population of a cell and its flag word, write every population). Q structure-of-arrays
// streams of the given element size, ping-pong buffers, one thread per cell, 50 untimed launches
// before the timed ones as the reference's harness does.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/lbm.h"
#include "solver_internal.hpp"

namespace lbm {

// Fill of the copy buffers with the kind of data a step moves: f - w of a uniform stream at
// u = (0.05, 0, 0) plus a hashed perturbation of 1e-4 per value (FP16 elements in the step's x 2^15
// scale). The memory system moves zeros measurably faster than such data on MI355X (a pull copy over
// zeroed buffers runs ~20% faster), so a ceiling for the step is measured on filled buffers; zeroed
// buffers (fill 0) keep the reference page's definition (fresh WebGPU buffers are zeroed).
template <class T, int Q>
__global__ __launch_bounds__(256) void fill_flow_kernel(T* __restrict__ a, uint32_t N) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const float u = 0.05f;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const float w = Q == 9 ? W2[i] : W3[i];
    const int cx = Q == 9 ? CX2[i] : CX3[i];
    const float cu = 3.0f * float(cx) * u;
    const uint32_t h = lowbias32(c * 19u + uint32_t(i));
    const float v = w * (cu + 0.5f * cu * cu - 1.5f * u * u) + 1e-4f * (float(h >> 8) * (1.0f / 16777216.0f) - 0.5f);
    a[size_t(i) * N + c] = sizeof(T) == 2 ? T(v * 32768.0f) : T(v);
  }
}

template <class T, int Q>
static int fill_flow(T* a, T* b, uint32_t N, bool fill, size_t bytes) {
  if (!fill) {
    HIP_TRY(hipMemset(a, 0, bytes));
    HIP_TRY(hipMemset(b, 0, bytes));
    return LBM_OK;
  }
  for (T* p : {a, b}) fill_flow_kernel<T, Q><<<(N + 255u) / 256u, 256>>>(p, N);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return LBM_OK;
}

template <class T, int Q>
__global__ __launch_bounds__(256) void copy_kernel(const T* __restrict__ src, T* __restrict__ dst,
                                                   const uint32_t* __restrict__ flags, uint32_t N) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  const T k = T(flags[idx]);
  T v[Q];
#pragma unroll
  for (int i = 0; i < Q; i++) v[i] = src[size_t(i) * N + idx];
#pragma unroll
  for (int i = 0; i < Q; i++) dst[size_t(i) * N + idx] = v[i] + k;
}

template <class T, int Q>
static int run_copy(int64_t n, int block, bool fill, int iters, double* ms_out) {
  const uint32_t N = uint32_t(n);
  T *a = nullptr, *b = nullptr;
  uint32_t* fl = nullptr;
  const size_t bytes = sizeof(T) * size_t(Q) * N;
  HIP_TRY(hipMalloc(&a, bytes));
  HIP_TRY(hipMalloc(&b, bytes));
  HIP_TRY(hipMalloc(&fl, sizeof(uint32_t) * N));
  if (int rc = fill_flow<T, Q>(a, b, N, fill, bytes)) return rc;
  HIP_TRY(hipMemset(fl, 0, sizeof(uint32_t) * N));
  hipStream_t st;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  const uint32_t grid = (N + uint32_t(block) - 1) / uint32_t(block);
  for (int k = 0; k < 50; k++) copy_kernel<T, Q><<<grid, block, 0, st>>>(k & 1 ? b : a, k & 1 ? a : b, fl, N);
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, st));
  for (int k = 0; k < iters; k++) copy_kernel<T, Q><<<grid, block, 0, st>>>(k & 1 ? b : a, k & 1 ? a : b, fl, N);
  HIP_TRY(hipEventRecord(e1, st));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  *ms_out = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(fl);
  return LBM_OK;
}

// Pull copy: the same traffic with the D3Q19 step's access pattern, every population read from the
// upstream node x - c_i (the shifted, partly unaligned streams of a pull step), over the rows with a
// row above and below in y and z (so every address lies in the buffer), one block per row.
template <class T>
__global__ __launch_bounds__(1024) void pull_copy_kernel(const T* __restrict__ src, T* __restrict__ dst,
                                                         const uint32_t* __restrict__ flags, uint32_t W,
                                                         uint32_t H, uint32_t N) {
  const uint32_t x = threadIdx.x;
  if (x >= W) return;
  const uint32_t y = 1u + blockIdx.x, z = 1u + blockIdx.y;
  const uint32_t idx = x + W * (y + H * z);
  const int64_t WH = int64_t(W) * H;
  const T k = T(flags[idx]);
  T v[19];
#pragma unroll
  for (int i = 0; i < 19; i++) v[i] = src[int64_t(i) * N + idx - (CX3[i] + int64_t(W) * CY3[i] + WH * CZ3[i])];
#pragma unroll
  for (int i = 0; i < 19; i++) dst[size_t(i) * N + idx] = v[i] + k;
}

template <class T>
static int run_pull_copy(int W, int H, int D, bool fill, int iters, double* ms_out, int64_t* cells) {
  const uint32_t N = uint32_t(W) * uint32_t(H) * uint32_t(D);
  T *a = nullptr, *b = nullptr;
  uint32_t* fl = nullptr;
  const size_t bytes = sizeof(T) * 19 * size_t(N);
  HIP_TRY(hipMalloc(&a, bytes));
  HIP_TRY(hipMalloc(&b, bytes));
  HIP_TRY(hipMalloc(&fl, sizeof(uint32_t) * N));
  if (int rc = fill_flow<T, 19>(a, b, N, fill, bytes)) return rc;
  HIP_TRY(hipMemset(fl, 0, sizeof(uint32_t) * N));
  hipStream_t st;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  const dim3 grid(uint32_t(H - 2), uint32_t(D - 2));
  const uint32_t block = (uint32_t(W) + 63u) / 64u * 64u;
  for (int k = 0; k < 50; k++) pull_copy_kernel<T><<<grid, block, 0, st>>>(k & 1 ? b : a, k & 1 ? a : b, fl, W, H, N);
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, st));
  for (int k = 0; k < iters; k++) pull_copy_kernel<T><<<grid, block, 0, st>>>(k & 1 ? b : a, k & 1 ? a : b, fl, W, H, N);
  HIP_TRY(hipEventRecord(e1, st));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  *ms_out = ms;
  *cells = int64_t(W) * (H - 2) * (D - 2);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(fl);
  return LBM_OK;
}

}  // namespace lbm

extern "C" int lbm_pull_copy_roofline(int device, int32_t width, int32_t height, int32_t depth, int32_t elem_bytes,
                                      int32_t fill, int32_t iters, double* ms_total, int64_t* cells) {
  using namespace lbm;
  if (width < 1 || width > 1024 || height < 3 || depth < 3 || iters <= 0 || !ms_total || !cells ||
      (fill != 0 && fill != 1) || 19ull * uint64_t(width) * uint64_t(height) * uint64_t(depth) >= (1ull << 32))
    return fail(LBM_ERR_INVALID, "bad pull-copy roofline arguments (width 1-1024, height and depth >= 3, fill 0 or 1)");
  HIP_TRY(hipSetDevice(device));
  if (elem_bytes == 4) return run_pull_copy<float>(width, height, depth, fill != 0, iters, ms_total, cells);
  if (elem_bytes == 2) return run_pull_copy<_Float16>(width, height, depth, fill != 0, iters, ms_total, cells);
  return fail(LBM_ERR_UNSUPPORTED, "pull-copy roofline: elem_bytes must be 4 or 2");
}

extern "C" int lbm_copy_roofline_filled(int device, int64_t n_cells, int32_t q, int32_t elem_bytes, int32_t workgroup,
                                        int32_t fill, int32_t iters, double* ms_total) {
  using namespace lbm;
  if (n_cells <= 0 || iters <= 0 || !ms_total || (fill != 0 && fill != 1) ||
      uint64_t(n_cells) * uint64_t(q) >= (1ull << 32))
    return fail(LBM_ERR_INVALID, "bad copy roofline arguments");
  const int block = workgroup ? workgroup : 256;
  if (block != 64 && block != 128 && block != 256)
    return fail(LBM_ERR_INVALID, "copy roofline workgroup must be 64, 128 or 256 (0: 256)");
  HIP_TRY(hipSetDevice(device));
  if (q == 19 && elem_bytes == 4) return run_copy<float, 19>(n_cells, block, fill != 0, iters, ms_total);
  if (q == 19 && elem_bytes == 2) return run_copy<_Float16, 19>(n_cells, block, fill != 0, iters, ms_total);
  if (q == 9 && elem_bytes == 4) return run_copy<float, 9>(n_cells, block, fill != 0, iters, ms_total);
  return fail(LBM_ERR_UNSUPPORTED, "copy roofline: q must be 19 (4 or 2 bytes) or 9 (4 bytes)");
}

extern "C" int lbm_copy_roofline(int device, int64_t n_cells, int32_t q, int32_t elem_bytes, int32_t workgroup,
                                 int32_t iters, double* ms_total) {
  return lbm_copy_roofline_filled(device, n_cells, q, elem_bytes, workgroup, 0, iters, ms_total);
}
