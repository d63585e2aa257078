// Host-side launchers of the gfx950 kernels (defined in kernels2d.hip / kernels3d.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_device.hpp"

namespace lbm {

// A Bouzidi link of the boundary list: its fluid fraction q (src/shaders/step.ts:140-142, fixed per
// geometry) and the gather index of the interpolation's second operand: f_j at x + c_i for q < 1/2
// (NO_OPERAND when that node is off the grid or solid, which falls back to f_j), f_i at x for q >= 1/2.
struct BzLink {
  float q;
  uint32_t gb;
};
constexpr uint32_t NO_OPERAND = 0xffffffffu;

// Boundary-list tables of one tunnel. list[0, n_link) are nodes without body links, list[n_link,
// n_link + n_body) nodes with at least one (they carry the force). For node k and direction i:
//   gtab[(i - 1) * (n_link + n_body) + k]  gather index of the pull (population * N + cell), which for
//                                          a Bouzidi or slip-through-solid direction is f_j at the node;
//   bmask[k] bit i                         Bouzidi link, record links[lbase[k] + rank of i in bmask[k]];
//   xmask[k] bit i                         3D: slip face mirroring a solid node (bounce-back, force 2 f_j);
//                                          2D: moving lid (adds 6 w_i c_ix uLid).
struct ListTables {
  const uint32_t* list;
  const uint32_t* gtab;
  const uint32_t* bmask;
  const uint32_t* xmask;
  const uint32_t* lbase;
  const BzLink* links;
  uint32_t n_link, n_body;
};

// What a recording step writes for the discrete adjoint: per cell, the moments (drho, ux, uy, uz) of
// the populations the collision takes (after the boundary rules), as a float4, or as four halves
// (Moments16) when the populations are stored in 16 bits; per Bouzidi link (record order of
// ListTables::links), the derivative of the interpolated population with respect to q.
struct Record {
  void* mom;
  float* dq;
};

struct alignas(8) Moments16 {
  _Float16 drho, ux, uy, uz;
};

// The record of one cell in the layout of the storage width (2: Moments16, rounded to nearest).
template <int ELEM_BYTES>
__device__ __forceinline__ void put_moments(void* mom, uint32_t idx, float drho, float ux, float uy, float uz) {
  if constexpr (ELEM_BYTES == 2)
    static_cast<Moments16*>(mom)[idx] = Moments16{_Float16(drho), _Float16(ux), _Float16(uy), _Float16(uz)};
  else
    static_cast<float4*>(mom)[idx] = make_float4(drho, ux, uy, uz);
}

// ---- D3Q19; prec is LBM_FP32 / LBM_FP16 / LBM_BF16 / LBM_FP16_SR ----
// One time step in one launch: bulk rows (plain fluid, slip-face rows and plain inlet/outlet nodes,
// one thread per cell) and boundary-list blocks (every other non-solid node, plus the force sample).
// descend: walk the rows top-down this step (alternated step by step for cache reuse). merge: the
// inlet and outlet nodes share their wave's stores (same values either way). early: the pulls are
// issued before the node's kind is known (same values; not for recording or subgrid-model steps).
// row_clean: one byte per row (y + H z), 1 where the row's node kinds follow from x alone.
void launch_step3(int prec, const void* src, void* dst, const uint32_t* flags, const ListTables& lt,
                  const Params3& p, const float* feq_in, const float4* xtab, const ForceSink& fs, int bulk_block,
                  bool descend, bool merge, bool early, const uint8_t* row_clean, hipStream_t st,
                  const Record* rec = nullptr);
// Per-x relaxation table (float4 per column) for the step kernels; rebuilt whenever parameters change.
void launch_xtab3(const Params3& p, float4* tab, hipStream_t st);
uint32_t bulk_block_x(int W, int requested);
void launch_macro3(int prec, const void* f, const uint32_t* flags, float4* mac, const Params3& p, hipStream_t st);
void launch_init3(int prec, void* fa, void* fb, const uint32_t* flags, const Params3& p, hipStream_t st);
void launch_flags3(int prec, const float* sdf, uint32_t* flags, void* fa, void* fb, const float4* mac,
                   const Params3& p, hipStream_t st);
// Blocks of bx threads for the list nodes without and with body links (at least one body block when
// a force sample is due, so the sample is always written).
uint32_t list_blocks(uint32_t n, uint32_t bx);
uint32_t body_list_blocks(uint32_t n_body, uint32_t bx, bool sample);

// ---- D3Q27 (prec LBM_FP32 / LBM_FP16 / LBM_BF16), same parameter block and relaxation table ----
// One launch per step (plus a one-wave reduction when a force sample is due); feq_in: the 27 inlet
// equilibria feq(i, 1, (uIn, 0, 0)) of the step (inlet_feq27); row_clean: one byte per row (y + H z),
// 1 where the row holds no solid or FLAG_SLOW27 node. Partials: 3 doubles per
// wave index, up to step27_partials(W, H, D) waves; force_waves lists, ascending, the n_force_waves
// wave indices holding a FLAG_SLOW27 node (wave = (block (z H + y) gx + x / bx) (bx / 64) + x % bx / 64,
// bx = step27_block_x(W)). descend: rows walked from the top (z and y reversed) in this step.
void launch_step27(int prec, const void* src, void* dst, const uint32_t* flags, const uint8_t* row_clean,
                   const float* sdf, const float4* xtab, const Params3& p, const float* feq_in, const ForceSink& fs,
                   const uint32_t* force_waves, uint32_t n_force_waves, bool descend, hipStream_t st);
uint32_t step27_block_x(int W);
size_t step27_partials(int W, int H, int D);
void launch_macro27(int prec, const void* f, const uint32_t* flags, float4* mac, const Params3& p, hipStream_t st);
void launch_init27(int prec, void* fa, void* fb, const uint32_t* flags, const Params3& p, hipStream_t st);
void launch_flags27(int prec, const float* sdf, uint32_t* flags, void* fa, void* fb, const float4* mac,
                    const Params3& p, hipStream_t st);

// ---- D2Q9 (FP32 storage) ----
// descend: bulk rows walked from the top (y reversed) in this step.
void launch_step2(const float* src, float* dst, const uint32_t* flags, const ListTables& lt,
                  const Params2& p, const float4* xtab, const ForceSink& fs, int bulk_block, bool descend,
                  hipStream_t st, const Record* rec = nullptr);
void launch_xtab2(const Params2& p, float4* tab, hipStream_t st);
void launch_macro2(const float* f, const uint32_t* flags, float4* mac, const Params2& p, hipStream_t st);
void launch_init2(float* fa, float* fb, const float4* mac, const Params2& p, hipStream_t st);
void launch_flags2(const float* sdf, uint32_t* flags, uint8_t* touches, float* fa, float* fb, const float4* mac,
                   const Params2& p, hipStream_t st);

// ---- shared ----
void launch_reset_state(uint32_t* state, hipStream_t st);

}  // namespace lbm
