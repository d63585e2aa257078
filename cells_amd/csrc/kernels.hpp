// Host-side launchers of the gfx950 kernels (defined in kernels2d.hip / kernels3d.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_device.hpp"

namespace lbm {

// Boundary-list tables of one tunnel: list[0, n_link) are link nodes whose pulls are single gathers
// gtab[(i - 1) * n_link + k] (plus, in 2D, a moving-lid term for the directions set in gmov[k]);
// list[n_link, n_link + n_body) are body nodes, which run the dynamic pull rules and carry force.
struct ListTables {
  const uint32_t* list;
  const uint32_t* gtab;
  const uint32_t* gmov;
  uint32_t n_link, n_body;
};

// ---- D3Q19; prec is LBM_FP32 / LBM_FP16 / LBM_BF16 ----
// One time step: the bulk kernel (plain fluid and plain inlet/outlet nodes, one thread per cell)
// followed by the boundary-list kernel (every other non-solid node, plus the force sample).
void launch_step3(int prec, const void* src, void* dst, const uint32_t* flags, const float* sdf,
                  const ListTables& lt, const Params3& p, const float* feq_in, const float4* xtab,
                  const ForceSink& fs, int bulk_block, hipStream_t st, hipEvent_t mid = nullptr);
// Per-x relaxation table (float4 per column) for the bulk kernel; rebuilt whenever parameters change.
void launch_xtab3(const Params3& p, float4* tab, hipStream_t st);
uint32_t bulk_block_x(int W, int requested);
void launch_macro3(int prec, const void* f, const uint32_t* flags, float4* mac, const Params3& p, hipStream_t st);
void launch_init3(int prec, void* fa, void* fb, const uint32_t* flags, const Params3& p, hipStream_t st);
void launch_flags3(int prec, const float* sdf, uint32_t* flags, void* fa, void* fb, const float4* mac,
                   const Params3& p, hipStream_t st);
// Blocks of 256 for the link nodes and for the body nodes (at least one body block when a force
// sample is due, so the sample is always written).
uint32_t link_blocks(uint32_t n_link);
uint32_t body_blocks(uint32_t n_body, bool sample);

// ---- D2Q9 (FP32 storage) ----
void launch_step2(const float* src, float* dst, const uint32_t* flags, const float* sdf, const ListTables& lt,
                  const Params2& p, const float4* xtab, const ForceSink& fs, int bulk_block, hipStream_t st,
                  hipEvent_t mid = nullptr);
void launch_xtab2(const Params2& p, float4* tab, hipStream_t st);
void launch_macro2(const float* f, const uint32_t* flags, float4* mac, const Params2& p, hipStream_t st);
void launch_init2(float* fa, float* fb, const float4* mac, const Params2& p, hipStream_t st);
void launch_flags2(const float* sdf, uint32_t* flags, uint8_t* touches, float* fa, float* fb, const float4* mac,
                   const Params2& p, hipStream_t st);

// ---- shared ----
void launch_reset_state(uint32_t* state, hipStream_t st);

}  // namespace lbm
