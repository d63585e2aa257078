// D3Q27 kernels for gfx950: the velocity-set switch of the reference's 3D roadmap
// (docs/research/3d-lattice-boltzmann-webgpu.md:74: "the kernel generator should take the velocity
// set as a parameter ... TRT carries over to both sets unchanged"), with the boundary rules of
// src/shaders/step3d.ts:119-223 and src/shaders/aux3d.ts:20-152 applied to the 27 directions.
//
// One thread per cell over the grid in cell order, one launch per step. A node whose pulls meet no
// solid node (all but the few next to a body) forms its 27 source indices by arithmetic alone, the
// direct pull, the outlet copy or the slip-face mirror, chosen from its flag bits and its row's
// position (uniform over a wave, since a wave lies in one grid row), and issues the 27 loads
// together; the nodes the flag kernel marks FLAG_SLOW27 resolve each flagged direction in turn
// (solid checks, Bouzidi, forces). Slip edges reflect a direction through both faces (the node
// x - c_x on the same row, c_y and c_z negated): for D3Q19 that is the reference's bounce-back,
// whose edge directions all have c_x = 0, and D3Q27's corner directions need it for a slip edge to
// keep a uniform flow uniform. Force samples: only FLAG_SLOW27 nodes carry momentum-exchange force,
// so only the waves holding one write an FP64 partial, and a one-wave kernel sums those partials in
// wave order (a list the host keeps per geometry) into the history ring, the same order every run.
// Storage is structure-of-arrays f[i * NS + cell] of (f_i - w_i) in FP32, FP16 x 2^15 or BF16 (NS the
// population stride, Params3::NS).
#include "kernels.hpp"
#include "reduce.hpp"

namespace lbm {
namespace {

__host__ __device__ constexpr int opp27(int i) { return i == 0 ? 0 : (i % 2 ? i + 1 : i - 1); }

__device__ __forceinline__ float feq27(int i, float rho, float ux, float uy, float uz) {
#pragma clang fp contract(off)
  const float cu = 3.0f * (float(CX27[i]) * ux + float(CY27[i]) * uy + float(CZ27[i]) * uz);
  return W27[i] * (rho - 1.0f + rho * (cu + 0.5f * cu * cu - 1.5f * (ux * ux + uy * uy + uz * uz)));
}

struct Mom27 {
  float drho, rho, invRho, ux, uy, uz, usq;
};

// Moments in index order (src/shaders/step3d.ts:38-45 over 27 directions).
__device__ __forceinline__ Mom27 moments27(const float (&f)[27]) {
  Mom27 m;
  float d = f[0], jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
  for (int i = 1; i < 27; i++) {
    d = d + f[i];
    if (CX27[i] > 0) jx = jx + f[i];
    if (CX27[i] < 0) jx = jx - f[i];
    if (CY27[i] > 0) jy = jy + f[i];
    if (CY27[i] < 0) jy = jy - f[i];
    if (CZ27[i] > 0) jz = jz + f[i];
    if (CZ27[i] < 0) jz = jz - f[i];
  }
  m.drho = d;
  m.rho = 1.0f + d;
  m.invRho = 1.0f / m.rho;
  m.ux = jx * m.invRho;
  m.uy = jy * m.invRho;
  m.uz = jz * m.invRho;
  m.usq = 1.5f * (m.ux * m.ux + m.uy * m.uy + m.uz * m.uz);
  return m;
}

// TRT collision with sponge, Smagorinsky and the outlet absorbing layer (src/shaders/step3d.ts:47-102)
// over the 13 opposite pairs (a, a + 1). t = (omega_base, omega_minus, sigma, 1 / omega_base) of the
// node's column.
template <bool SMAG>
__device__ __forceinline__ void collide27(const float (&f)[27], const Mom27& m, float4 t, const Params3& p,
                                          float (&out)[27]) {
  float cu[27], eqS[27], eqA[27];
#pragma unroll
  for (int a = 1; a < 27; a += 2) {
    float acc = 0.0f;
    if (CX27[a] > 0) acc = acc + m.ux;
    if (CX27[a] < 0) acc = acc - m.ux;
    if (CY27[a] > 0) acc = acc + m.uy;
    if (CY27[a] < 0) acc = acc - m.uy;
    if (CZ27[a] > 0) acc = acc + m.uz;
    if (CZ27[a] < 0) acc = acc - m.uz;
    cu[a] = 3.0f * acc;
    eqS[a] = W27[a] * (m.drho + m.rho * (0.5f * cu[a] * cu[a] - m.usq));
    eqA[a] = W27[a] * m.rho * cu[a];
  }
  float omP = t.x, omM = t.y;
  const float sigma = t.z;
  if constexpr (SMAG) {
    float pxx = 0.0f, pyy = 0.0f, pzz = 0.0f, pxy = 0.0f, pxz = 0.0f, pyz = 0.0f;
#pragma unroll
    for (int a = 1; a < 27; a += 2) {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int i = a + h;
        const float n = f[i] - (h == 0 ? eqS[a] + eqA[a] : eqS[a] - eqA[a]);
        if (CX27[i] != 0) pxx = pxx + n;
        if (CY27[i] != 0) pyy = pyy + n;
        if (CZ27[i] != 0) pzz = pzz + n;
        if (CX27[i] * CY27[i] > 0) pxy = pxy + n;
        if (CX27[i] * CY27[i] < 0) pxy = pxy - n;
        if (CX27[i] * CZ27[i] > 0) pxz = pxz + n;
        if (CX27[i] * CZ27[i] < 0) pxz = pxz - n;
        if (CY27[i] * CZ27[i] > 0) pyz = pyz + n;
        if (CY27[i] * CZ27[i] < 0) pyz = pyz - n;
      }
    }
    const float pi = sqrtf(pxx * pxx + pyy * pyy + pzz * pzz + 2.0f * (pxy * pxy + pxz * pxz + pyz * pyz));
    const float t0 = t.w;
    const float tEff = 0.5f * (t0 + sqrtf(t0 * t0 + 25.455844122715714f * p.smagC2 * pi * m.invRho));
    omP = 1.0f / tEff;
    omM = 1.0f / (0.5f + p.lambda / (1.0f / omP - 0.5f));
  }
  float s0 = 0.0f;
  float sS[27], sA[27];
#pragma unroll
  for (int a = 1; a < 27; a += 2) sS[a] = sA[a] = 0.0f;
  if (sigma > 0.0f) {
    const float ut = p.uIn;
    const float dq = 1.5f * m.rho * (ut * ut - m.ux * m.ux - m.uy * m.uy - m.uz * m.uz);
    s0 += sigma * W27[0] * -dq;
#pragma unroll
    for (int a = 1; a < 27; a += 2) {
      const float cut = 3.0f * float(CX27[a]) * ut;
      sS[a] += sigma * W27[a] * (0.5f * m.rho * (cut * cut - cu[a] * cu[a]) - dq);
      sA[a] += sigma * W27[a] * m.rho * (cut - cu[a]);
    }
  }
  out[0] = f[0] - omP * (f[0] - W27[0] * (m.drho - m.rho * m.usq)) + s0;
#pragma unroll
  for (int a = 1; a < 27; a += 2) {
    const int b = a + 1;
    const float sym = omP * (0.5f * (f[a] + f[b]) - eqS[a]);
    const float asym = omM * (0.5f * (f[a] - f[b]) - eqA[a]);
    out[a] = f[a] - sym - asym + sS[a] + sA[a];
    out[b] = f[b] - sym + asym + sS[a] - sA[a];
  }
}

// One flagged direction i of node (x, y, z): src/shaders/step3d.ts:119-168 over 27 directions, the
// slip edge reflected through both faces (see the file comment). Every index it forms lies on the
// grid: inlet nodes do not pull, an outlet node's directions leaving the grid in x take the copy
// rule, and every other node has 1 <= x <= W - 2.
template <class S>
__device__ __forceinline__ float pull27(const typename S::elem* __restrict__ src, const uint32_t* __restrict__ flags,
                                        const float* __restrict__ sdf, const Params3& p, int i, int x, int y, int z,
                                        uint32_t idx, bool outlet, float (&force)[3]) {
  const uint32_t N = p.NS;  // population stride
  const int W = p.W, H = p.H, D = p.D;
  const int j = opp27(i);
  auto at = [&](int ax, int ay, int az) { return uint32_t(ax) + uint32_t(W) * (uint32_t(ay) + uint32_t(H) * uint32_t(az)); };
  auto ld = [&](int dir, uint32_t cell) { return S::load(src + size_t(dir) * N, cell + p.NA * (cell & ~63u)); };
  const int sx = x - CX27[i], sy = y - CY27[i], sz = z - CZ27[i];
  const bool outX = sx < 0 || sx >= W, outY = sy < 0 || sy >= H, outZ = sz < 0 || sz >= D;
  if (outX && outlet) return ld(i, at(W - 2, y, z));
  if (outY || outZ) {
    const uint32_t rs = at(sx, outY ? y : sy, outZ ? z : sz);
    const int m = outY && outZ ? mirror_y27(mirror_z27(i)) : outY ? mirror_y27(i) : mirror_z27(i);
    if (flags[rs] & FLAG_SOLID27) {
      const float fj = ld(j, idx);
      force[0] += float(CX27[j]) * (2.0f * fj);
      force[1] += float(CY27[j]) * (2.0f * fj);
      force[2] += float(CZ27[j]) * (2.0f * fj);
      return fj;
    }
    return ld(m, rs);
  }
  if (outX) return ld(j, idx);
  const uint32_t sc = at(sx, sy, sz);
  if (!(flags[sc] & FLAG_SOLID27)) return ld(i, sc);
  const float fj = ld(j, idx);
  const float dF = sdf[idx], dS = sdf[sc];
  const float q = wclamp(dF / fmaxf(dF - dS, 1e-6f), 1e-3f, 1.0f);
  float fi = fj;
  if (q < 0.5f) {
    const int nx = x + CX27[i], ny = y + CY27[i], nz = z + CZ27[i];
    if (nx >= 0 && nx < W && ny >= 0 && ny < H && nz >= 0 && nz < D) {
      const uint32_t n = at(nx, ny, nz);
      if (!(flags[n] & FLAG_SOLID27)) fi = 2.0f * q * fj + (1.0f - 2.0f * q) * ld(j, n);
    }
  } else {
    fi = fj / (2.0f * q) + (2.0f * q - 1.0f) / (2.0f * q) * ld(i, idx);
  }
  force[0] += float(CX27[j]) * (fj + fi);
  force[1] += float(CY27[j]) * (fj + fi);
  force[2] += float(CZ27[j]) * (fj + fi);
  return fi;
}

// Inlet equilibrium feq(i, 1, (uIn, 0, 0)), uniform over the inlet plane: computed on the host per
// step and passed by value.
struct FeqIn27 {
  float v[27];
};

// src/shaders/step3d.ts:170-223 for node (x, y, z) over 27 directions, leaving the values to store in
// out: the inlet's equilibrium (no pulls, no collision), the outlet's equilibrium at rho = 1 with its
// own velocity, or the collision. All nodes of a wave then store together (27 stores per wave,
// whole lines, also for the waves holding the inlet or outlet node). Returns whether the node is a
// FLAG_SLOW27 node (the only kind that adds to force).
template <class S, bool SMAG>
__device__ __forceinline__ bool node27(const typename S::elem* __restrict__ src, const uint32_t* __restrict__ flags,
                                       const float* __restrict__ sdf, const float4* __restrict__ xtab, const Params3& p,
                                       const FeqIn27& fin, int x, int y, int z, uint32_t idx, uint32_t fl,
                                       float (&force)[3], float (&out)[27]) {
  const uint32_t N = p.NS;  // population stride
  const uint32_t o0 = idx + p.NA * (idx & ~63u);  // the node's element within each population
  const bool slow = (fl & FLAG_SLOW27) != 0u;
  if (fl & FLAG_INLET27) {  // src/shaders/step3d.ts:190-193: equilibrium at (1, uIn, 0, 0)
#pragma unroll
    for (int i = 0; i < 27; i++) out[i] = fin.v[i];
    return slow;
  }
  float f[27];
  const bool outlet = (fl & FLAG_OUTLET27) != 0u;
  if (!slow) {
    // Source of each population by arithmetic. On these nodes a direction leaves the grid only
    // across a slip face or edge of the row (the mirrored population at the reflected node) or, at
    // the outlet node, across x (copy from (W-2, y, z), the previous node of the row); no solid is
    // involved. So the source is a row-uniform population and cell offset (scalar registers) plus
    // the lane's x, with the outlet node's nine c_x < 0 directions replaced.
    // The element of population m at cell idx - dyz - c_x is a wave-uniform base (population m, the
    // row shift dyz, a multiple of 64 cells in tiles since W % 64 == 0 then) plus this node's element,
    // or the one before / after it (crossing into the previous / next tile at the wave's first / last
    // lane). Fast-path nodes have 1 <= x <= W - 1 (the inlet node pulled nothing above), and at
    // x = W - 1 the outlet copy rule replaces the c_x < 0 pulls.
    const uint32_t op = x == 0 ? o0 : o0 - 1u - ((idx & 63u) == 0u ? 64u * p.NA : 0u);
    const uint32_t om = x == p.W - 1 ? o0 : o0 + 1u + ((idx & 63u) == 63u ? 64u * p.NA : 0u);
#pragma unroll
    for (int i = 0; i < 27; i++) {
      const int sy = y - CY27[i], sz = z - CZ27[i];
      const bool outY = sy < 0 || sy >= p.H, outZ = sz < 0 || sz >= p.D;
      const int m = outY && outZ ? mirror_y27(mirror_z27(i)) : outY ? mirror_y27(i) : outZ ? mirror_z27(i) : i;
      const int dyz = (outY ? 0 : CY27[i] * p.W) + (outZ ? 0 : CZ27[i] * int(p.WH));
      const typename S::elem* base = src + ptrdiff_t(m) * ptrdiff_t(N) - ptrdiff_t(1u + p.NA) * dyz;
      uint32_t off = CX27[i] > 0 ? op : CX27[i] < 0 ? om : o0;
      if (CX27[i] < 0 && outlet) {
        base = src + size_t(i) * N;
        off = o0 - 1u;
      }
      f[i] = S::load(base, off);
    }
  } else {
    f[0] = S::load(src, o0);
#pragma unroll
    for (int i = 1; i < 27; i++) {
      if (fl & (1u << i)) {
        f[i] = pull27<S>(src, flags, sdf, p, i, x, y, z, idx, outlet, force);
      } else {
        const int delta = CZ27[i] * int(p.WH) + CY27[i] * p.W + CX27[i];
        const uint32_t cell = uint32_t(int(idx) - delta);
        f[i] = S::load(src + size_t(i) * N, cell + p.NA * (cell & ~63u));
      }
    }
  }
  const Mom27 m = moments27(f);
  if (outlet) {
    // src/shaders/step3d.ts:205-209: equilibrium at rho = 1 with the node's velocity, per opposite
    // pair (c_b = -c_a): w (+-cu + cu^2 / 2 - 3 u^2 / 2) with the shared terms once (the outlet node
    // shares its wave with 63 colliding nodes, so this branch costs the whole wave)
    out[0] = W27[0] * -m.usq;
#pragma unroll
    for (int a = 1; a < 27; a += 2) {
      float cu = 0.0f;
      if (CX27[a] > 0) cu = cu + m.ux;
      if (CX27[a] < 0) cu = cu - m.ux;
      if (CY27[a] > 0) cu = cu + m.uy;
      if (CY27[a] < 0) cu = cu - m.uy;
      if (CZ27[a] > 0) cu = cu + m.uz;
      if (CZ27[a] < 0) cu = cu - m.uz;
      cu = 3.0f * cu;
      const float t = 0.5f * cu * cu - m.usq;
      out[a] = W27[a] * (t + cu);
      out[a + 1] = W27[a] * (t - cu);
    }
  } else {
    collide27<SMAG>(f, m, xtab[x], p, out);
  }
  return slow;
}

// Grid (ceil(W / blockDim.x), H, D): x from the block column and thread, y and z from the launch,
// walked from the top (z and y reversed) when descend is set. Alternating the direction every step
// makes a step begin on the rows the previous one wrote last, which are still in the 256 MB Infinity
// Cache when the two population buffers do not fit in it together (FP32: 2 x 191 MB at 192x96x96).
// With a force sample due, every wave holding a FLAG_SLOW27 node writes its FP64 partial at its wave
// index (block-major over the grid in cell order, whichever direction the rows are walked in).
template <class S, bool SMAG>
__global__ __launch_bounds__(256) void step27_kernel(const typename S::elem* __restrict__ src,
                                                     typename S::elem* __restrict__ dst,
                                                     const uint32_t* __restrict__ flags,
                                                     const uint8_t* __restrict__ row_clean, const float* __restrict__ sdf,
                                                     const float4* __restrict__ xtab, Params3 p, FeqIn27 fin,
                                                     double* __restrict__ partials, int sample, uint32_t descend) {
  const int x = int(blockIdx.x * blockDim.x + threadIdx.x);
  const int y = descend ? p.H - 1 - int(blockIdx.y) : int(blockIdx.y);
  const int z = descend ? p.D - 1 - int(blockIdx.z) : int(blockIdx.z);
  float force[3] = {0.0f, 0.0f, 0.0f};
  bool slow = false;
  if (x < p.W) {
    const uint32_t idx = uint32_t(x) + uint32_t(p.W) * (uint32_t(y) + uint32_t(p.H) * uint32_t(z));
    // In a clean row (no solid or FLAG_SLOW27 node) only the inlet and outlet bits matter, and they
    // follow from x, so the loads need not wait for the flag word.
    const uint32_t fl = row_clean[uint32_t(y) + uint32_t(p.H) * uint32_t(z)]
                            ? (x == 0 ? FLAG_INLET27 : x == p.W - 1 ? FLAG_OUTLET27 : 0u)
                            : flags[idx];
    if (!(fl & FLAG_SOLID27)) {
      float out[27];
      slow = node27<S, SMAG>(src, flags, sdf, xtab, p, fin, x, y, z, idx, fl, force, out);
#pragma unroll
      for (int i = 0; i < 27; i++) S::store(dst + size_t(i) * p.NS, idx + p.NA * (idx & ~63u), out[i]);
    }
  }
  if (!sample || __ballot(slow) == 0ull) return;  // uniform over the wave
  double v[3] = {double(force[0]), double(force[1]), double(force[2])};
  wave_sum<3>(v);
  if ((threadIdx.x & 63u) == 0u) {
    const size_t block = (size_t(z) * gridDim.y + size_t(y)) * gridDim.x + blockIdx.x;
    const size_t part = block * (blockDim.x >> 6) + (threadIdx.x >> 6);
#pragma unroll
    for (int d = 0; d < 3; d++) partials[part * 3 + d] = v[d];
  }
}

// One wave: the partials of the listed waves (ascending wave index) summed lane-strided, then across
// the wave, appended to the history ring (src/shaders/aux3d.ts:155-183's ring semantics). Eight of a
// lane's list entries are loaded at a time (their indices, then their partials) so that a chunk
// costs two memory round trips instead of two per entry; each lane adds in list order as before.
__global__ __launch_bounds__(64) void reduce27_kernel(const double* __restrict__ partials,
                                                      const uint32_t* __restrict__ waves, uint32_t nwaves,
                                                      double* __restrict__ ring, uint32_t* __restrict__ state,
                                                      uint64_t step) {
  constexpr int U = 8;
  double acc[3] = {0.0, 0.0, 0.0};
  for (uint32_t k0 = threadIdx.x; k0 < nwaves; k0 += 64u * U) {
    uint32_t w[U];
#pragma unroll
    for (int u = 0; u < U; u++) w[u] = k0 + 64u * u < nwaves ? waves[k0 + 64u * u] : ~0u;
    double v[U][3];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int d = 0; d < 3; d++) v[u][d] = w[u] != ~0u ? partials[size_t(w[u]) * 3 + d] : 0.0;
#pragma unroll
    for (int u = 0; u < U; u++)
      if (w[u] != ~0u)
#pragma unroll
        for (int d = 0; d < 3; d++) acc[d] += v[u][d];
  }
  wave_sum<3>(acc);
  if (threadIdx.x == 0) {
    const uint32_t st = state[0];
    const uint32_t slot = st % HISTORY_LEN;
#pragma unroll
    for (int d = 0; d < 3; d++) ring[size_t(slot) * 3 + d] = acc[d];
    note_divergence<3>(acc, state, step);
    state[0] = st + 1u;
  }
}

// src/shaders/aux3d.ts:137-150 over 27 directions: (ux, uy, uz, rho), zero for solids.
template <class S>
__global__ __launch_bounds__(256) void macro27_kernel(const typename S::elem* __restrict__ f,
                                                      const uint32_t* __restrict__ flags, float4* __restrict__ mac,
                                                      Params3 p) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= p.N) return;
  if (flags[idx] & FLAG_SOLID27) {
    mac[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  float drho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
  for (int i = 0; i < 27; i++) {
    const float v = S::load(f + size_t(i) * p.NS, idx + p.NA * (idx & ~63u));
    drho += v;
    jx += float(CX27[i]) * v;
    jy += float(CY27[i]) * v;
    jz += float(CZ27[i]) * v;
  }
  const float rho = 1.0f + drho;
  mac[idx] = make_float4(jx / rho, jy / rho, jz / rho, rho);
}

// src/shaders/aux3d.ts:113-123: inflow equilibrium in fluid, zero in solids.
template <class S>
__global__ __launch_bounds__(256) void init27_kernel(typename S::elem* __restrict__ fa, typename S::elem* __restrict__ fb,
                                                     const uint32_t* __restrict__ flags, Params3 p) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= p.N) return;
  const bool solid = (flags[idx] & FLAG_SOLID27) != 0u;
#pragma unroll
  for (int i = 0; i < 27; i++) {
    const float v = solid ? 0.0f : feq27(i, 1.0f, p.uIn, 0.0f, 0.0f);
    S::store(fa + size_t(i) * p.NS, idx + p.NA * (idx & ~63u), v);
    S::store(fb + size_t(i) * p.NS, idx + p.NA * (idx & ~63u), v);
  }
}

// src/shaders/aux3d.ts:32-100 over 27 directions: flags from the SDF (inlet and outlet planes never
// solid) and refill of the cells the body uncovered with the equilibrium at the mean density and
// velocity of their fluid neighbours (rest at the inflow velocity when there is none).
template <class S>
__global__ __launch_bounds__(256) void flags27_kernel(const float* __restrict__ sdf, uint32_t* __restrict__ flags,
                                                      typename S::elem* __restrict__ fa, typename S::elem* __restrict__ fb,
                                                      const float4* __restrict__ mac, Params3 p) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= p.N) return;
  const int W = p.W, H = p.H, D = p.D;
  const int z = int(idx / p.WH);
  const int y = int((idx - uint32_t(z) * p.WH) / uint32_t(W));
  const int x = int(idx - uint32_t(z) * p.WH - uint32_t(y) * uint32_t(W));
  auto in_range = [&](int qx, int qy, int qz) { return qx >= 0 && qy >= 0 && qz >= 0 && qx < W && qy < H && qz < D; };
  auto solid_at = [&](int qx, int qy, int qz) {
    if (qx == 0 || qx == W - 1) return false;
    return sdf[uint32_t(qx + W * (qy + H * qz))] < 0.0f;
  };
  const uint32_t old = flags[idx];
  if (solid_at(x, y, z)) {
    flags[idx] = FLAG_SOLID27;
    return;
  }
  uint32_t bits = 0u;
  if (x == 0) bits = FLAG_INLET27;
  if (x == W - 1) bits = FLAG_OUTLET27;
  bool slow = false;
#pragma unroll
  for (int i = 1; i < 27; i++) {
    const int sx = x - CX27[i], sy = y - CY27[i], sz = z - CZ27[i];
    if (in_range(sx, sy, sz)) {
      if (solid_at(sx, sy, sz)) {
        bits |= 1u << i;
        slow = true;
      }
      continue;
    }
    bits |= 1u << i;
    const bool outX = sx < 0 || sx >= W, outY = sy < 0 || sy >= H, outZ = sz < 0 || sz >= D;
    if (!outX && (outY || outZ) && solid_at(sx, outY ? y : sy, outZ ? z : sz)) slow = true;
  }
  flags[idx] = bits | (slow ? FLAG_SLOW27 : 0u);
  if (old & FLAG_SOLID27) {
    float rho = 0.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f, n = 0.0f;
#pragma unroll
    for (int i = 1; i < 27; i++) {
      const int qx = x + CX27[i], qy = y + CY27[i], qz = z + CZ27[i];
      if (!in_range(qx, qy, qz)) continue;
      const float4 m = mac[uint32_t(qx + W * (qy + H * qz))];
      if (m.w != 0.0f) {
        rho += m.w;
        ux += m.x;
        uy += m.y;
        uz += m.z;
        n += 1.0f;
      }
    }
    if (n > 0.0f) {
      rho /= n;
      ux /= n;
      uy /= n;
      uz /= n;
    } else {
      rho = 1.0f;
      ux = p.uIn;
      uy = 0.0f;
      uz = 0.0f;
    }
#pragma unroll
    for (int i = 0; i < 27; i++) {
      const float v = feq27(i, rho, ux, uy, uz);
      S::store(fa + size_t(i) * p.NS, idx + p.NA * (idx & ~63u), v);
      S::store(fb + size_t(i) * p.NS, idx + p.NA * (idx & ~63u), v);
    }
  }
}

inline uint32_t blocks_for27(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

template <class S>
void step27_t(const void* src, void* dst, const uint32_t* flags, const uint8_t* row_clean, const float* sdf,
              const float4* xtab, const Params3& p, const float* feq_in, const ForceSink& fs, const uint32_t* force_waves, uint32_t n_force_waves,
              bool descend, hipStream_t st) {
  FeqIn27 fin;
  for (int i = 0; i < 27; i++) fin.v[i] = feq_in[i];
  using E = typename S::elem;
  const uint32_t bx = step27_block_x(p.W);
  const dim3 grid(blocks_for27(uint32_t(p.W), bx), uint32_t(p.H), uint32_t(p.D));
  const E* s0 = static_cast<const E*>(src);
  E* d0 = static_cast<E*>(dst);
  if (p.smagC2 > 0.0f)
    step27_kernel<S, true><<<grid, bx, 0, st>>>(s0, d0, flags, row_clean, sdf, xtab, p, fin, fs.partials, fs.sample, descend);
  else
    step27_kernel<S, false><<<grid, bx, 0, st>>>(s0, d0, flags, row_clean, sdf, xtab, p, fin, fs.partials, fs.sample, descend);
  if (fs.sample) reduce27_kernel<<<1, 64, 0, st>>>(fs.partials, force_waves, n_force_waves, fs.ring, fs.state, fs.step);
}

}  // namespace

uint32_t step27_block_x(int W) { return W >= 256 ? 256u : uint32_t((W + 63) / 64 * 64); }

size_t step27_partials(int W, int H, int D) {
  const uint32_t bx = step27_block_x(W);
  return size_t((uint32_t(W) + bx - 1) / bx) * size_t(H) * size_t(D) * (bx / 64u);
}

void launch_step27(int prec, const void* src, void* dst, const uint32_t* flags, const uint8_t* row_clean,
                   const float* sdf, const float4* xtab, const Params3& p, const float* feq_in, const ForceSink& fs,
                   const uint32_t* force_waves, uint32_t n_force_waves, bool descend, hipStream_t st) {
  switch (prec) {
    case 1: step27_t<StoreF16>(src, dst, flags, row_clean, sdf, xtab, p, feq_in, fs, force_waves, n_force_waves, descend, st); break;
    case 2: step27_t<StoreBF16>(src, dst, flags, row_clean, sdf, xtab, p, feq_in, fs, force_waves, n_force_waves, descend, st); break;
    default: step27_t<StoreF32>(src, dst, flags, row_clean, sdf, xtab, p, feq_in, fs, force_waves, n_force_waves, descend, st); break;
  }
}

template <class S>
static void macro27_t(const void* f, const uint32_t* flags, float4* mac, const Params3& p, hipStream_t st) {
  macro27_kernel<S><<<blocks_for27(p.N, 256), 256, 0, st>>>(static_cast<const typename S::elem*>(f), flags, mac, p);
}
void launch_macro27(int prec, const void* f, const uint32_t* flags, float4* mac, const Params3& p, hipStream_t st) {
  switch (prec) {
    case 1: macro27_t<StoreF16>(f, flags, mac, p, st); break;
    case 2: macro27_t<StoreBF16>(f, flags, mac, p, st); break;
    default: macro27_t<StoreF32>(f, flags, mac, p, st); break;
  }
}

template <class S>
static void init27_t(void* fa, void* fb, const uint32_t* flags, const Params3& p, hipStream_t st) {
  using E = typename S::elem;
  init27_kernel<S><<<blocks_for27(p.N, 256), 256, 0, st>>>(static_cast<E*>(fa), static_cast<E*>(fb), flags, p);
}
void launch_init27(int prec, void* fa, void* fb, const uint32_t* flags, const Params3& p, hipStream_t st) {
  switch (prec) {
    case 1: init27_t<StoreF16>(fa, fb, flags, p, st); break;
    case 2: init27_t<StoreBF16>(fa, fb, flags, p, st); break;
    default: init27_t<StoreF32>(fa, fb, flags, p, st); break;
  }
}

template <class S>
static void flags27_t(const float* sdf, uint32_t* flags, void* fa, void* fb, const float4* mac, const Params3& p,
                      hipStream_t st) {
  using E = typename S::elem;
  flags27_kernel<S><<<blocks_for27(p.N, 256), 256, 0, st>>>(sdf, flags, static_cast<E*>(fa), static_cast<E*>(fb), mac, p);
}
void launch_flags27(int prec, const float* sdf, uint32_t* flags, void* fa, void* fb, const float4* mac,
                    const Params3& p, hipStream_t st) {
  switch (prec) {
    case 1: flags27_t<StoreF16>(sdf, flags, fa, fb, mac, p, st); break;
    case 2: flags27_t<StoreBF16>(sdf, flags, fa, fb, mac, p, st); break;
    default: flags27_t<StoreF32>(sdf, flags, fa, fb, mac, p, st); break;
  }
}

}  // namespace lbm
