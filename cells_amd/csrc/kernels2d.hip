// D2Q9 TRT kernels for gfx950: fused pull-stream + boundaries + collide, flags, init, macro.
//
// Restates src/shaders/step.ts:9-229 and src/shaders/aux.ts:14-141 for HIP with the same bulk /
// boundary-list split as the D3Q19 kernels (kernels3d.hip): the bulk rows stream plain fluid
// nodes and plain Zou-He inlet / outlet nodes with coalesced loads and stores; every other
// non-solid node (walls, periodic seams, body links) runs the reference's pullSlow rules from a
// list, with the momentum-exchange force reduced in FP64 in the same launch.
#include "d2q9.hpp"
#include "kernels.hpp"
#include "reduce.hpp"

namespace lbm {

// ============================================================================ step kernel
// One time step in one launch. Grid (ceil(W / blockDim.x), yl + H): the first yl rows of blocks are
// boundary-list blocks (dispatched first, so their dependent gathers overlap the streaming): list
// blocks [0, nlb) take the nodes without body links, the nbb after them the nodes with body links,
// which carry the force sample. The remaining rows are the bulk rows, where x and y come from the
// launch coordinates and per direction the source address is a wave-uniform base
// (src + i*N - delta_i) plus the lane's byte offset. With descend set the bulk rows are walked from
// the top: alternating the direction every step makes a step begin on the rows the previous one wrote
// last, which are still in the 256 MB Infinity Cache when the two population buffers do not fit in it.
template <bool REC>
__global__ __launch_bounds__(1024) void step2_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                     const uint32_t* __restrict__ flags,
                                                     const float4* __restrict__ xtab, Params2 p, ListTables lt,
                                                     uint32_t yl, uint32_t nlb, uint32_t nbb, ForceSink fs,
                                                     Record rc, uint32_t descend) {
  if (blockIdx.y < yl) {
    const uint32_t lb = blockIdx.y * gridDim.x + blockIdx.x;
    float force[2] = {0.0f, 0.0f};
    if (lb < nlb) {
      const uint32_t k = lb * blockDim.x + threadIdx.x;
      if (k < lt.n_link) list_cell2<false, REC>(src, dst, flags, xtab, lt, k, p, rc, force);
      return;
    }
    if (lb >= nlb + nbb) return;
    // a sample summed in this one block: its ring count is read now, beside the node's own loads
    const uint32_t st0 = (fs.sample && nbb == 1 && threadIdx.x == 0)
                             ? __hip_atomic_load(fs.state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                             : ~0u;
    const uint32_t kb = (lb - nlb) * blockDim.x + threadIdx.x;
    if (kb < lt.n_body) list_cell2<true, REC>(src, dst, flags, xtab, lt, lt.n_link + kb, p, rc, force);
    if (fs.sample) force_sample<2>(force, fs, lb - nlb, nbb, st0);
    return;
  }
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= uint32_t(p.W)) return;
  const uint32_t y = descend ? uint32_t(p.H) - 1u - (blockIdx.y - yl) : blockIdx.y - yl;
  bulk_cell2<REC>(src, dst, flags, xtab, p, x, int(y), x + uint32_t(p.W) * y, rc);
}

__global__ __launch_bounds__(256) void xtab2_kernel(Params2 p, float4* __restrict__ tab) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= p.W) return;
  const Relax2 r = relax2(x, p);
  tab[x] = make_float4(r.omBase, r.omM, r.sigma, r.t0);
}

// ============================================================================ aux kernels
// src/shaders/aux.ts:128-139: (rho, ux, uy, solid) with the half-step body-force correction.
__global__ __launch_bounds__(256) void macro2_kernel(const float* __restrict__ f, const uint32_t* __restrict__ flags,
                                                     float4* __restrict__ mac, Params2 p) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= p.N) return;
  if (flags[idx] & FLAG_SOLID2) {
    mac[idx] = make_float4(1.0f, 0.0f, 0.0f, 1.0f);
    return;
  }
  float v[9];
#pragma unroll
  for (int i = 0; i < 9; i++) v[i] = f[size_t(i) * p.NS + idx + p.NA * (idx & ~63u)];
  const float rho = 1.0f + v[0] + v[1] + v[2] + v[3] + v[4] + v[5] + v[6] + v[7] + v[8];
  float jx = v[1] - v[3] + v[5] - v[6] - v[7] + v[8];
  float jy = v[2] - v[4] + v[5] + v[6] - v[7] - v[8];
  if (p.forcing != 0) {
    jx -= 0.5f * p.fx;
    jy -= 0.5f * p.fy;
  }
  mac[idx] = make_float4(rho, jx / rho, jy / rho, 0.0f);
}

// src/shaders/aux.ts:100-110
__global__ __launch_bounds__(256) void init2_kernel(float* __restrict__ fa, float* __restrict__ fb,
                                                    const float4* __restrict__ mac, Params2 p) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= p.N) return;
  const float4 m = mac[idx];
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const float v = feq2(i, m.x, m.y, m.z);
    fa[size_t(i) * p.NS + idx + p.NA * (idx & ~63u)] = v;
    fb[size_t(i) * p.NS + idx + p.NA * (idx & ~63u)] = v;
  }
}

// src/shaders/aux.ts:33-87. touches[idx] = 1 for a fluid node next to the body; the host numbers
// those nodes in cell order and stores the slot in the flag bits >= 12, as the reference layout has.
__global__ __launch_bounds__(256) void flags2_kernel(const float* __restrict__ sdf, uint32_t* __restrict__ flags,
                                                     uint8_t* __restrict__ touches, float* __restrict__ fa,
                                                     float* __restrict__ fb, const float4* __restrict__ mac,
                                                     Params2 p) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= p.N) return;
  const int W = p.W, H = p.H;
  const int y = int(idx / uint32_t(W));
  const int x = int(idx - uint32_t(y) * uint32_t(W));
  auto solid_at = [&](int qx, int qy) {
    const bool open = (p.left == 2 && qx == 0) || (p.right == 2 && qx == W - 1);
    return !open && sdf[uint32_t(qy * W + qx)] < 0.0f;
  };
  auto wrap = [&](int& qx, int& qy) {
    if (qx < 0 && p.left == 0) qx += W;
    if (qx >= W && p.right == 0) qx -= W;
    if (qy < 0 && p.bottom == 0) qy += H;
    if (qy >= H && p.top == 0) qy -= H;
  };
  const uint32_t old = flags[idx];
  touches[idx] = 0;
  if (solid_at(x, y)) {
    flags[idx] = FLAG_SOLID2;
    return;
  }
  uint32_t bits = 0u;
  if (p.left == 2 && x == 0) bits = FLAG_INLET2;
  if (p.right == 2 && x == W - 1) bits = FLAG_OUTLET2;
  bool touch = false;
#pragma unroll
  for (int i = 1; i < 9; i++) {
    int sx = x - CX2[i], sy = y - CY2[i];
    wrap(sx, sy);
    const bool inside = sx >= 0 && sy >= 0 && sx < W && sy < H;
    if (inside && solid_at(sx, sy)) {
      bits |= 1u << i;
      touch = true;
    } else if (!inside || sx != x - CX2[i] || sy != y - CY2[i]) {
      bits |= 1u << i;
    }
  }
  flags[idx] = bits;
  touches[idx] = touch ? 1 : 0;
  if (old & FLAG_SOLID2) {
    float rho = 0.0f, ux = 0.0f, uy = 0.0f, n = 0.0f;
#pragma unroll
    for (int i = 1; i < 9; i++) {
      const int sx = x + CX2[i], sy = y + CY2[i];
      if (sx < 0 || sy < 0 || sx >= W || sy >= H) continue;
      const float4 m = mac[uint32_t(sy * W + sx)];
      if (m.w == 0.0f) {
        rho += m.x;
        ux += m.y;
        uy += m.z;
        n += 1.0f;
      }
    }
    if (n > 0.0f) {
      rho /= n;
      ux /= n;
      uy /= n;
    } else {
      rho = 1.0f;
    }
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const float v = feq2(i, rho, ux, uy);
      fa[size_t(i) * p.NS + idx + p.NA * (idx & ~63u)] = v;
      fb[size_t(i) * p.NS + idx + p.NA * (idx & ~63u)] = v;
    }
  }
}

// ============================================================================ launchers
static inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

void launch_step2(const float* src, float* dst, const uint32_t* flags, const ListTables& lt, const Params2& p,
                  const float4* xtab, const ForceSink& fs, int bb, bool descend, hipStream_t st, const Record* rec) {
  const uint32_t bx = bulk_block_x(p.W, bb);
  const uint32_t gx = (uint32_t(p.W) + bx - 1) / bx;
  const uint32_t nlb = list_blocks(lt.n_link, bx);
  const uint32_t nbb = body_list_blocks(lt.n_body, bx, fs.sample != 0);
  const uint32_t yl = (nlb + nbb + gx - 1) / gx;
  const dim3 grid(gx, yl + uint32_t(p.H));
  if (rec)
    step2_kernel<true><<<grid, bx, 0, st>>>(src, dst, flags, xtab, p, lt, yl, nlb, nbb, fs, *rec, descend ? 1u : 0u);
  else
    step2_kernel<false><<<grid, bx, 0, st>>>(src, dst, flags, xtab, p, lt, yl, nlb, nbb, fs, Record{nullptr, nullptr},
                                             descend ? 1u : 0u);
}
void launch_xtab2(const Params2& p, float4* tab, hipStream_t st) {
  xtab2_kernel<<<blocks_for(uint32_t(p.W), 256), 256, 0, st>>>(p, tab);
}
void launch_macro2(const float* f, const uint32_t* flags, float4* mac, const Params2& p, hipStream_t st) {
  macro2_kernel<<<blocks_for(p.N, 256), 256, 0, st>>>(f, flags, mac, p);
}
void launch_init2(float* fa, float* fb, const float4* mac, const Params2& p, hipStream_t st) {
  init2_kernel<<<blocks_for(p.N, 256), 256, 0, st>>>(fa, fb, mac, p);
}
void launch_flags2(const float* sdf, uint32_t* flags, uint8_t* touches, float* fa, float* fb, const float4* mac,
                   const Params2& p, hipStream_t st) {
  flags2_kernel<<<blocks_for(p.N, 256), 256, 0, st>>>(sdf, flags, touches, fa, fb, mac, p);
}

}  // namespace lbm
