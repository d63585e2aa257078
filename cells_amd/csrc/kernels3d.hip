// D3Q19 TRT kernels for gfx950: fused pull-stream + boundaries + collide, flags, init, macro.
//
// Restates src/shaders/step3d.ts:24-225 and src/shaders/aux3d.ts:20-152 for HIP, re-organised for
// wave64 hardware: instead of one kernel in which the x-boundary and face nodes of every row take
// a divergent slow path, a step is one launch (step3_kernel) of two kinds of blocks:
//   * bulk rows: one thread per cell, one block row per grid row. They handle plain fluid nodes,
//     plain inlet / outlet nodes (only the off-grid x bits set), which sit in the first and last wave
//     of each x row, and whole slip-face rows, whose pull pattern is uniform along the row, so every
//     wave streams fully coalesced loads and stores. Other nodes (solid, next to a body, irregular
//     face nodes) are skipped.
//   * boundary-list blocks, dispatched first: one thread per node of a list built once per geometry
//     change (every non-solid node the bulk rows skip, grouped by kind, in cell order within a kind).
//     The reference's per-direction boundary rules (pullSlow3) are resolved on the host into gather
//     indices and Bouzidi records, so these threads issue independent loads only; the blocks of nodes
//     with body links carry the momentum-exchange force reduction in FP64 (reduce.hpp).
// Storage is structure-of-arrays f[i * N + cell] of (f_i - w_i) in FP32, FP16 x 2^15 or BF16;
// arithmetic is FP32 throughout, as in the reference.
#include <type_traits>

#include "kernels.hpp"
#include "reduce.hpp"

namespace lbm {

__device__ __forceinline__ float feq3(int i, float rho, float ux, float uy, float uz) {
  // src/shaders/common3d.ts:78-81. Evaluated without contraction, so that every kernel that inlines it
  // (whatever code surrounds it) rounds the same way.
#pragma clang fp contract(off)
  const float cu = 3.0f * (float(CX3[i]) * ux + float(CY3[i]) * uy + float(CZ3[i]) * uz);
  return W3[i] * (rho - 1.0f + rho * (cu + 0.5f * cu * cu - 1.5f * (ux * ux + uy * uy + uz * uz)));
}

// Zou-He velocity (x = 0, u = (uIn, 0, 0)) and pressure (x = W-1, rho = 1, no transverse velocity)
// faces after Hecht and Harting, J. Stat. Mech. (2010) P01018 (arXiv:0811.4593), in the f - w
// variables. The populations entering from outside the grid (c_x = +1 at the inlet, -1 at the
// outlet) are rebuilt from their opposites,
//   f_i = f_opp(i) + 6 w_i rho (c_i . u) - c_iy N_y - c_iz N_z,  N_a = 1/2 sum_{c_x = 0} c_a f,
// with rho (1 - uIn) = 1 + sum_{c_x = 0} f + 2 sum_{c_x = -1} f at the inlet and
// u_x = sum_{c_x = 0} f + 2 sum_{c_x = +1} f at the outlet, so that the node's density and
// momentum are exactly the prescribed ones; the node then collides. On slip-face rows the mirrored
// populations count as known, so the planes' edges and corners take the same rule.
template <bool IN>
__device__ __forceinline__ void zou_he_face3(float (&f)[19], float uIn) {
  constexpr int in = IN ? 1 : -1;
  float s0 = 0.0f, sk = 0.0f, ny = 0.0f, nz = 0.0f;
#pragma unroll
  for (int i = 0; i < 19; i++) {
    if (CX3[i] == 0) {
      s0 = s0 + f[i];
      if (CY3[i]) ny = CY3[i] > 0 ? ny + f[i] : ny - f[i];
      if (CZ3[i]) nz = CZ3[i] > 0 ? nz + f[i] : nz - f[i];
    } else if (CX3[i] == -in) {
      sk = sk + f[i];
    }
  }
  ny = 0.5f * ny;
  nz = 0.5f * nz;
  float ru;
  if (IN) ru = (1.0f + s0 + 2.0f * sk) / (1.0f - uIn) * uIn;
  else ru = s0 + 2.0f * sk;
#pragma unroll
  for (int i = 1; i < 19; i++) {
    if (CX3[i] != in) continue;
    const float k = (CY3[i] || CZ3[i]) ? 1.0f / 6.0f : 1.0f / 3.0f;  // 6 w_i
    float v = IN ? f[opp3(i)] + k * ru : f[opp3(i)] - k * ru;
    if (CY3[i]) v = CY3[i] > 0 ? v - ny : v + ny;
    if (CZ3[i]) v = CZ3[i] > 0 ? v - nz : v + nz;
    f[i] = v;
  }
}

// Stores of the time step: with a stochastic codec every value gets its random word (node word nw =
// lowbias32(idx ^ seed), population i); otherwise round to nearest.
template <class S>
__device__ __forceinline__ void put_at(char* base, uint32_t off, float v, uint32_t nw, int i) {
  if constexpr (S::stochastic) S::store_at_r(base, off, v, step_random(nw, i));
  else S::store_at(base, off, v);
}
template <class S>
__device__ __forceinline__ void put(typename S::elem* p, uint32_t k, float v, uint32_t nw, int i) {
  if constexpr (S::stochastic) S::store_r(p, k, v, step_random(nw, i));
  else S::store(p, k, v);
}

struct Mom3 {
  float drho, rho, invRho, ux, uy, uz, usq;
  float s[9], d[9];  // f_a + f_b and f_a - f_b of the TRT pairs (a, b) = (2k + 1, 2k + 2)
};

// Moments of src/shaders/step3d.ts:38-45, summed through the pair sums and differences, which the
// collision reuses: drho = f_0 + sum_k s_k, j = sum_k c_(2k+1) d_k.
__device__ __forceinline__ Mom3 moments3(const float (&f)[19]) {
  Mom3 m;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    m.s[k] = f[2 * k + 1] + f[2 * k + 2];
    m.d[k] = f[2 * k + 1] - f[2 * k + 2];
  }
  float d = f[0];
#pragma unroll
  for (int k = 0; k < 9; k++) d = d + m.s[k];
  float jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const int a = 2 * k + 1;
    if (CX3[a] > 0) jx = jx + m.d[k];
    if (CX3[a] < 0) jx = jx - m.d[k];
    if (CY3[a] > 0) jy = jy + m.d[k];
    if (CY3[a] < 0) jy = jy - m.d[k];
    if (CZ3[a] > 0) jz = jz + m.d[k];
    if (CZ3[a] < 0) jz = jz - m.d[k];
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

// Relaxation coefficients that depend only on x (the sponge ramps tau toward tauSponge over the
// last fraction of the tunnel, src/shaders/step3d.ts:213-219): omega_base, omega_minus for the
// no-Smagorinsky case, the absorbing-layer strength sigma, and 1 / omega_base.
struct Relax {
  float omBase, omM, sigma, t0;
};

__device__ __forceinline__ Relax relax3(float x, const Params3& p) {
  Relax r;
  float omegaBase = 1.0f / p.tau;
  float sOut = 0.0f;
  const float wm1 = float(p.W - 1);
  if (p.spongeStart < wm1) {
    sOut = smoothstep(p.spongeStart, wm1, x);
    omegaBase = 1.0f / wmix(p.tau, p.tauSponge, sOut);
  }
  r.omBase = omegaBase;
  r.sigma = p.absorb * sOut * sOut;
  r.t0 = 1.0f / omegaBase;
  r.omM = 1.0f / (0.5f + p.lambda / (r.t0 - 0.5f));
  return r;
}

// TRT collision with sponge, Smagorinsky and the outlet absorbing layer: src/shaders/step3d.ts:47-102,
// 210-222, term by term as the reference writes it. Writes the post-collision populations to out.
__device__ __forceinline__ void collide3_general(const float (&f)[19], const Mom3& m, const Relax& r, const Params3& p,
                                                 float (&out)[19]) {
  float cu[19], eqS[19], eqA[19];
#pragma unroll
  for (int a = 1; a < 19; a += 2) {
    float acc = 0.0f;
    if (CX3[a] > 0) acc = acc + m.ux;
    if (CX3[a] < 0) acc = acc - m.ux;
    if (CY3[a] > 0) acc = acc + m.uy;
    if (CY3[a] < 0) acc = acc - m.uy;
    if (CZ3[a] > 0) acc = acc + m.uz;
    if (CZ3[a] < 0) acc = acc - m.uz;
    cu[a] = 3.0f * acc;
    eqS[a] = W3[a] * (m.drho + m.rho * (0.5f * cu[a] * cu[a] - m.usq));
    eqA[a] = W3[a] * m.rho * cu[a];
  }
  const float omegaBase = r.omBase;
  const float sigma = r.sigma;
  float omP = omegaBase;
  if (p.smagC2 > 0.0f) {
    float n[19];
    n[0] = 0.0f;
#pragma unroll
    for (int a = 1; a < 19; a += 2) {
      n[a] = f[a] - (eqS[a] + eqA[a]);
      n[a + 1] = f[a + 1] - (eqS[a] - eqA[a]);
    }
    float pxx = 0.0f, pyy = 0.0f, pzz = 0.0f, pxy = 0.0f, pxz = 0.0f, pyz = 0.0f;
#pragma unroll
    for (int i = 1; i < 19; i++) {
      if (CX3[i] != 0) pxx = pxx + n[i];
      if (CY3[i] != 0) pyy = pyy + n[i];
      if (CZ3[i] != 0) pzz = pzz + n[i];
      if (CX3[i] * CY3[i] > 0) pxy = pxy + n[i];
      if (CX3[i] * CY3[i] < 0) pxy = pxy - n[i];
      if (CX3[i] * CZ3[i] > 0) pxz = pxz + n[i];
      if (CX3[i] * CZ3[i] < 0) pxz = pxz - n[i];
      if (CY3[i] * CZ3[i] > 0) pyz = pyz + n[i];
      if (CY3[i] * CZ3[i] < 0) pyz = pyz - n[i];
    }
    const float pi = sqrtf(pxx * pxx + pyy * pyy + pzz * pzz + 2.0f * (pxy * pxy + pxz * pxz + pyz * pyz));
    const float t0 = r.t0;
    const float tEff = 0.5f * (t0 + sqrtf(t0 * t0 + 25.455844122715714f * p.smagC2 * pi * m.invRho));
    omP = 1.0f / tEff;
  }
  const float omM = p.smagC2 > 0.0f ? 1.0f / (0.5f + p.lambda / (1.0f / omP - 0.5f)) : r.omM;
  float s0 = 0.0f;
  float sS[19], sA[19];
#pragma unroll
  for (int a = 1; a < 19; a += 2) sS[a] = sA[a] = 0.0f;
  if (sigma > 0.0f) {
    const float ut = p.uIn;
    const float dq = 1.5f * m.rho * (ut * ut - m.ux * m.ux - m.uy * m.uy - m.uz * m.uz);
    s0 += sigma * W3[0] * -dq;
#pragma unroll
    for (int a = 1; a < 19; a += 2) {
      const float cut = 3.0f * float(CX3[a]) * ut;
      sS[a] += sigma * W3[a] * (0.5f * m.rho * (cut * cut - cu[a] * cu[a]) - dq);
      sA[a] += sigma * W3[a] * m.rho * (cut - cu[a]);
    }
  }
  out[0] = f[0] - omP * (f[0] - W3[0] * (m.drho - m.rho * m.usq)) + s0;
#pragma unroll
  for (int a = 1; a < 19; a += 2) {
    const int b = a + 1;
    const float sym = omP * (0.5f * (f[a] + f[b]) - eqS[a]);
    const float asym = omM * (0.5f * (f[a] - f[b]) - eqA[a]);
    out[a] = f[a] - sym - asym + sS[a] + sA[a];
    out[b] = f[b] - sym + asym + sS[a] - sA[a];
  }
}

// The same collision without the subgrid model, regrouped per pair around the pair sums s and
// differences d of the moments (f_a = (s + d) / 2, f_b = (s - d) / 2). With e = omega+ eqS and
// m = omega- eqA (src/shaders/step3d.ts:93-101):
//   out_a = f_a - sym - asym = A + B,  out_b = f_b - sym + asym = A - B,
//   A = (1/2 - omega+/2) s + e,  B = (1/2 - omega-/2) d + m,
//   e = P0 + P2 cu^2, P0 = omega+ w (drho - rho usq), P2 = omega+ w rho / 2,  m = Q cu, Q = omega- w rho,
// the coefficients depending only on the pair's weight class. The absorbing-layer sources are added
// where sigma > 0 with collide3_general's formulas. With the subgrid model (SMAG, chosen on the host
// from smagC2 > 0, so that kernels without it do not carry its registers) collide3_general runs.
template <bool SMAG>
__device__ __forceinline__ void collide3(const float (&f)[19], const Mom3& m, const Relax& r, const Params3& p,
                                         float (&out)[19]) {
  if constexpr (SMAG) {
    collide3_general(f, m, r, p, out);
    return;
  }
  const float omP = r.omBase, omM = r.omM;
  const float base = m.drho - m.rho * m.usq;
  const float cs = 0.5f - 0.5f * omP, cd = 0.5f - 0.5f * omM;
  const float P0a = omP * W3[1] * base, P0d = omP * W3[7] * base;
  const float P2a = 0.5f * omP * W3[1] * m.rho, P2d = 0.5f * omP * W3[7] * m.rho;
  const float Qa = omM * W3[1] * m.rho, Qd = omM * W3[7] * m.rho;
  const float u3x = 3.0f * m.ux, u3y = 3.0f * m.uy, u3z = 3.0f * m.uz;
  out[0] = f[0] - omP * (f[0] - W3[0] * base);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const int a = 2 * k + 1;
    float c = 0.0f;
    if (CX3[a] > 0) c = c + u3x;
    if (CX3[a] < 0) c = c - u3x;
    if (CY3[a] > 0) c = c + u3y;
    if (CY3[a] < 0) c = c - u3y;
    if (CZ3[a] > 0) c = c + u3z;
    if (CZ3[a] < 0) c = c - u3z;
    const bool axis = k < 3;
    const float A = cs * m.s[k] + ((axis ? P2a : P2d) * (c * c) + (axis ? P0a : P0d));
    const float B = cd * m.d[k] + (axis ? Qa : Qd) * c;
    out[a] = A + B;
    out[a + 1] = A - B;
  }
  if (r.sigma > 0.0f) {
    const float sigma = r.sigma, ut = p.uIn;
    const float dq = 1.5f * m.rho * (ut * ut - m.ux * m.ux - m.uy * m.uy - m.uz * m.uz);
    out[0] += sigma * W3[0] * -dq;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int a = 2 * k + 1;
      const float cu = 3.0f * (float(CX3[a]) * m.ux + float(CY3[a]) * m.uy + float(CZ3[a]) * m.uz);
      const float cut = 3.0f * float(CX3[a]) * ut;
      const float sS = sigma * W3[a] * (0.5f * m.rho * (cut * cut - cu * cu) - dq);
      const float sA = sigma * W3[a] * m.rho * (cut - cu);
      out[a] += sS + sA;
      out[a + 1] += sS - sA;
    }
  }
}

// The outlet plane's equilibrium at rho = 1 with the node's own velocity (src/shaders/step3d.ts:205-209,
// feq of common3d.ts:78-81 at rho = 1) for all 19 directions, in the pair form the collision uses:
// feq_a = w (S + cu), feq_b = w (S - cu), S = cu^2 / 2 - usq, for the pair (a, b) with cu = 3 c_a . u.
// Algebraically feq3(i, 1, u) (its rounding differs by a few ulps) with fewer live values.
__device__ __forceinline__ void outlet_feq3(const Mom3& m, float (&out)[19]) {
  const float u3x = 3.0f * m.ux, u3y = 3.0f * m.uy, u3z = 3.0f * m.uz;
  out[0] = W3[0] * -m.usq;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const int a = 2 * k + 1;
    float c = 0.0f;
    if (CX3[a] > 0) c = c + u3x;
    if (CX3[a] < 0) c = c - u3x;
    if (CY3[a] > 0) c = c + u3y;
    if (CY3[a] < 0) c = c - u3y;
    if (CZ3[a] > 0) c = c + u3z;
    if (CZ3[a] < 0) c = c - u3z;
    const float S = 0.5f * c * c - m.usq;
    out[a] = W3[a] * (S + c);
    out[a + 1] = W3[a] * (S - c);
  }
}

// Inlet equilibrium feq(i, 1, (uIn, 0, 0)) for all i, computed on the host per step (it is uniform
// over the inlet plane) and passed by value.
struct FeqIn {
  float v[19];
};

// ============================================================================ bulk rows
// Grid (ceil(W / blockDim.x), H, D): one block row of threads per x row, so x, y, z come from the
// launch coordinates. Per direction the source address is a wave-uniform base (src + d_i*N - delta_i,
// kept in SGPRs) plus the lane's byte offset, shared by all 19 loads; stores likewise.
//
// A row on a slip face (y = 0 or H-1, z = 0 or D-1) follows one pull pattern along its whole
// length (src/shaders/step3d.ts:132-145): a direction entering across one face reads the mirrored
// population at the node reflected through that face, one entering across two faces bounces back.
// The pattern is uniform per row, so it only changes the per-direction base (d_i, delta_i); the
// node's flag word must be exactly the face bits of its row (the host moves nodes whose pulls
// differ, e.g. where the mirror node is solid, to the boundary list and marks them FLAG_LIST3).
// Outlet nodes of face rows go to the boundary list as well.
template <class S, bool REC, bool SMAG, bool ZH>
__device__ __forceinline__ void list_part3(const typename S::elem* __restrict__ src, typename S::elem* __restrict__ dst,
                                           const uint32_t* __restrict__ flags, const ListTables& lt, uint32_t nlb,
                                           uint32_t nbb, const Params3& p, const FeqIn& fin,
                                           const float4* __restrict__ xtab, const ForceSink& fs, const Record& rc,
                                           uint32_t lb);

// One time step in one launch. Grid (ceil(W / blockDim.x), H, zl + D): the first zl z-slices are
// boundary-list blocks (dispatched first, so their dependent gathers overlap the streaming), the
// rest are the bulk rows. REC additionally writes the adjoint's per-step record.
//
// descend walks the rows in reverse order (z and y from the top). Alternating the direction every
// step makes a step start on the rows the previous step wrote last, which are still in the 256 MB
// Infinity Cache when the two population buffers do not fit in it together.
//
// MERGE selects how the inlet and outlet nodes of a row store: false, each stores its own values
// in a branch of its own and the wave's other nodes store as soon as their collision is done (the
// stores overlap the arithmetic); true, their values are chosen first and every node of the wave
// shares one store per population (whole lines per store instead of a line with a one-node hole
// plus a one-node store, which costs read-modify-writes where the lines go to HBM rather than stay
// in cache). Both give the same values; the autotuner picks per grid.
//
// EARLY issues the 19 pulls before the node's kind is known: the row's clean byte (and, on a row
// that is not clean, the node's flag word, a second dependent load) is then waited for while the
// pulls are in flight instead of ahead of them. Every node pulls, also one that turns out to be
// handled elsewhere (boundary list, solid) or to be the reference's inlet plane, which use none of
// the values. Each node computes what it does in the other order (bit for bit with per-node stores;
// with MERGE the compiler may fuse different multiply-adds); which order is faster depends on the
// grid (measured both ways), so the autotuner picks.
template <class S, bool REC, bool SMAG, bool MERGE, bool ZH, bool EARLY = false>
__global__ __launch_bounds__(1024) void step3_kernel(const typename S::elem* __restrict__ src,
                                                     typename S::elem* __restrict__ dst,
                                                     const uint32_t* __restrict__ flags,
                                                     const float4* __restrict__ xtab, Params3 p, FeqIn fin,
                                                     ListTables lt, uint32_t zl, uint32_t nlb, uint32_t nbb,
                                                     ForceSink fs, Record rc, uint32_t descend,
                                                     const uint8_t* __restrict__ row_clean) {
  using E = typename S::elem;
  if (blockIdx.z < zl) {
    const uint32_t lb = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    if (lb < nlb + nbb) list_part3<S, REC, SMAG, ZH>(src, dst, flags, lt, nlb, nbb, p, fin, xtab, fs, rc, lb);
    return;
  }
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= uint32_t(p.W)) return;
  const uint32_t y = descend ? uint32_t(p.H) - 1u - blockIdx.y : blockIdx.y;
  const uint32_t z = descend ? uint32_t(p.D) - 1u - (blockIdx.z - zl) : blockIdx.z - zl;
  const bool loY = y == 0u, hiY = y == uint32_t(p.H - 1), loZ = z == 0u, hiZ = z == uint32_t(p.D - 1);
  const bool face = loY || hiY || loZ || hiZ;
  uint32_t M = 0u;  // face bits of this row
#pragma unroll
  for (int i = 1; i < 19; i++) {
    const bool outY = (loY && CY3[i] > 0) || (hiY && CY3[i] < 0);
    const bool outZ = (loZ && CZ3[i] > 0) || (hiZ && CZ3[i] < 0);
    if (outY || outZ) M |= 1u << i;
  }
  const uint32_t idx = x + uint32_t(p.W) * (y + uint32_t(p.H) * z);
  const size_t N = p.NS;  // population stride
  constexpr uint32_t EB = uint32_t(sizeof(E));
  // Element of population i at cell c: c + NA (c & ~63) + NS i (NA = 0: one array per population;
  // NA = 18, NS = 64: tiles of 64 cells, whose rows start tiles because W % 64 == 0 there, so a
  // wave is one tile). off is this node's byte offset within a population; a pull from c - 1 or
  // c + 1 (c_x = +1 / -1) reads at offp / offm, which on a tiled layout cross into the previous /
  // next tile at the wave's first / last lane. At x = 0 and x = W - 1 those pulls read the node's
  // own element instead: the inlet node never uses them, and the x = W - 1 node either takes the
  // outlet copy rule or is a boundary-list node.
  const uint32_t off = (idx + p.NA * (idx & ~63u)) * EB;
  const uint32_t tile_step = 64u * p.NA * EB;
  const uint32_t offp = x == 0u ? off : off - EB - ((idx & 63u) == 0u ? tile_step : 0u);
  const uint32_t offm = x == uint32_t(p.W - 1) ? off : off + EB + ((idx & 63u) == 63u ? tile_step : 0u);
  const char* const sb = reinterpret_cast<const char*>(src);
  float f[19];
  // The 19 pulls; at_outlet selects the outlet copy rule for the directions entering from x = W.
  // Per direction the source is population sd at cell idx - dyz - dx: a wave-uniform base (the
  // population and the y / z shift dyz, a multiple of 64 cells on a tiled layout) plus off, offp or
  // offm for dx = 0, +1, -1.
  auto pull = [&](bool at_outlet) {
#pragma unroll
    for (int i = 0; i < 19; i++) {
      const bool outY = (loY && CY3[i] > 0) || (hiY && CY3[i] < 0);
      const bool outZ = (loZ && CZ3[i] > 0) || (hiZ && CZ3[i] < 0);
      int sd = i;
      ptrdiff_t dyz = ptrdiff_t(CZ3[i]) * p.WH + ptrdiff_t(CY3[i]) * p.W;
      if (outY && outZ) {
        sd = opp3(i);  // bounce-back: the node's own cell
        dyz = 0;
      } else if (outY) {
        sd = mirror_y3(i);
        dyz = ptrdiff_t(CZ3[i]) * p.WH;
      } else if (outZ) {
        sd = mirror_z3(i);
        dyz = ptrdiff_t(CY3[i]) * p.W;
      }
      const char* base = sb + (ptrdiff_t(sd) * ptrdiff_t(N) - ptrdiff_t(1u + p.NA) * dyz) * ptrdiff_t(EB);
      uint32_t o = CX3[i] == 0 || (outY && outZ) ? off : CX3[i] > 0 ? offp : offm;
      // Outlet copy rule (step3d.ts:130): the population entering from x = W comes from (W-2, y, z),
      // i.e. population i at cell idx - 1, in the same tile (rows off the faces only, where base is
      // the plain one).
      if (CX3[i] < 0 && at_outlet) o = off - EB + uint32_t(ptrdiff_t(1u + p.NA) * dyz) * EB;
      f[i] = S::load_at(base, o);
    }
  };
  const uint8_t* const clean_byte = row_clean + (y + uint32_t(p.H) * z);
  uint8_t clean = 0;
  if constexpr (EARLY) {
    // Only x = W-1 off the faces can hold an outlet node (face-row outlet nodes are on the boundary
    // list). Every address lies inside the population buffer: a pull that would leave the grid
    // through x = 0 or W-1 reads the node's own element (offp / offm above), and the y / z shifts
    // of a row stay on the grid (the slip faces' patterns).
    clean = *clean_byte;
    pull(!face && x == uint32_t(p.W - 1));
  } else {
    clean = *clean_byte;
  }
  // A clean row (no solid or boundary-list node besides the outlet node of a face row) is known
  // from its position alone, so its flag words are not read.
  uint32_t fl;
  if (clean)
    fl = x == 0u ? (FLAG_INLET3 | INLET_DIRS3 | M)
                 : x == uint32_t(p.W - 1) ? (face ? FLAG_LIST3 : (FLAG_OUTLET3 | OUTLET_DIRS3)) : M;
  else
    fl = flags[idx];
  const bool inlet = fl == (FLAG_INLET3 | INLET_DIRS3 | M);
  const bool outlet = !face && fl == (FLAG_OUTLET3 | OUTLET_DIRS3);
  if (fl != M && !inlet && !outlet) return;  // handled by the boundary list (or solid)
  char* const d = reinterpret_cast<char*>(dst);
  const uint32_t nw = S::stochastic ? lowbias32(idx ^ p.seed) : 0u;
  constexpr bool zh = ZH;  // Zou-He faces (p.faces != 0), compiled apart so the default kernels carry none of it
  if (!MERGE && inlet && !zh) {
    // src/shaders/step3d.ts:190-193: equilibrium at (1, uIn, 0, 0), no collision, no force.
#pragma unroll
    for (int i = 0; i < 19; i++) put_at<S>(d + i * N * sizeof(E), off, fin.v[i], nw, i);
    return;
  }
  // (MERGE or Zou-He faces: the inlet node's c_x = +1 pulls reach x = -1, the previous row's last
  // node: in bounds, replaced by the Zou-He rule or unused.)
  if constexpr (!EARLY) pull(outlet);
  if (zh && inlet) zou_he_face3<true>(f, p.uIn);
  else if (zh && outlet) zou_he_face3<false>(f, p.uIn);
  const Mom3 m = moments3(f);
  if (REC && !inlet && rc.mom) put_moments<sizeof(E)>(rc.mom, idx, m.drho, m.ux, m.uy, m.uz);
  if (!MERGE && outlet && !zh) {
    // src/shaders/step3d.ts:205-209: equilibrium at rho = 1 with the node's own velocity.
    float eq[19];
    outlet_feq3(m, eq);
#pragma unroll
    for (int i = 0; i < 19; i++) put_at<S>(d + i * N * sizeof(E), off, eq[i], nw, i);
    return;
  }
  const float4 t = xtab[x];
  Relax r;
  r.omBase = t.x;
  r.omM = t.y;
  r.sigma = t.z;
  r.t0 = t.w;
  float out[19];
  collide3<SMAG>(f, m, r, p, out);
  if (MERGE && outlet && !zh) {
    outlet_feq3(m, out);
  } else if (MERGE && inlet && !zh) {
#pragma unroll
    for (int i = 0; i < 19; i++) out[i] = fin.v[i];
  }
#pragma unroll
  for (int i = 0; i < 19; i++) put_at<S>(d + i * N * sizeof(E), off, out[i], nw, i);
}

__global__ __launch_bounds__(256) void xtab3_kernel(Params3 p, float4* __restrict__ tab) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= p.W) return;
  const Relax r = relax3(float(x), p);
  tab[x] = make_float4(r.omBase, r.omM, r.sigma, r.t0);
}

// ============================================================================ boundary-list blocks
// One node of the boundary list (k is its position in the list). Every pull is a gather through the
// table the host resolved from pullSlow3's rules (src/shaders/step3d.ts:119-168), so the 19 loads are
// independent. Nodes with body links (BODY) then apply them: a Bouzidi link (step3d.ts:147-166)
// replaces the gathered f_j with the interpolation against its record's second operand and adds
// c_j (f_j + f_i) to the force; a slip face mirroring a solid node (step3d.ts:139-146) keeps f_j and
// adds c_j 2 f_j.
// The body links of a list node for directions LO .. LO + 8: Bouzidi interpolation of the gathered
// f_j against the record's second operand (src/shaders/step3d.ts:147-166) and the momentum
// exchange c_j (f_j + f_i) into force; a slip face mirroring a solid node (step3d.ts:139-146) keeps
// f_j and adds c_j 2 f_j. Directions are taken in increasing order, as one loop over all would.
template <class S, bool REC, int LO>
__device__ __forceinline__ void link_half3(const typename S::elem* __restrict__ src, const ListTables& lt, uint32_t k,
                                            uint32_t bm, uint32_t xm, const BzLink* __restrict__ rec,
                                            const Record& rc, float (&f)[19], float (&force)[3],
                                            std::integral_constant<int, LO>) {
  float q[9], b[9];
  uint32_t gb[9];
#pragma unroll
  for (int h = 0; h < 9; h++) {
    const int i = LO + h;
    q[h] = 1.0f;
    gb[h] = NO_OPERAND;
    if (bm & (1u << i)) {
      const BzLink r = rec[__popc(bm & ((1u << i) - 1u))];
      q[h] = r.q;
      gb[h] = r.gb;
    }
  }
#pragma unroll
  for (int h = 0; h < 9; h++) b[h] = gb[h] != NO_OPERAND ? S::load(src, gb[h]) : 0.0f;
#pragma unroll
  for (int h = 0; h < 9; h++) {
    const int i = LO + h;
    const int j = opp3(i);
    const float fj = f[i];
    if (bm & (1u << i)) {
      const float qi = q[h];
      float fi = fj;
      if (qi < 0.5f) {
        if (gb[h] != NO_OPERAND) fi = 2.0f * qi * fj + (1.0f - 2.0f * qi) * b[h];
      } else {
        fi = fj / (2.0f * qi) + (2.0f * qi - 1.0f) / (2.0f * qi) * b[h];
      }
      force[0] += float(CX3[j]) * (fj + fi);
      force[1] += float(CY3[j]) * (fj + fi);
      force[2] += float(CZ3[j]) * (fj + fi);
      if (REC && rc.dq) rc.dq[lt.lbase[k] + __popc(bm & ((1u << i) - 1u))] = bouzidi_dq(qi, fj, b[h], gb[h] != NO_OPERAND);
      f[i] = fi;
    } else if (xm & (1u << i)) {
      force[0] += float(CX3[j]) * (2.0f * fj);
      force[1] += float(CY3[j]) * (2.0f * fj);
      force[2] += float(CZ3[j]) * (2.0f * fj);
    }
  }
}

template <class S, bool BODY, bool REC, bool SMAG, bool ZH>
__device__ __forceinline__ void list_node3(const typename S::elem* __restrict__ src, typename S::elem* __restrict__ dst,
                                           const uint32_t* __restrict__ flags, const ListTables& lt, uint32_t k,
                                           const Params3& p, const FeqIn& fin, const float4* __restrict__ xtab,
                                           const Record& rc, float (&force)[3]) {
  const uint32_t N = p.NS;  // population stride
  const uint32_t nl = lt.n_link + lt.n_body;
  const uint32_t idx = lt.list[k];
  const uint32_t ce = idx + p.NA * (idx & ~63u);  // the node's element within each population
  const uint32_t fl = flags[idx];
  const uint32_t nw = S::stochastic ? lowbias32(idx ^ p.seed) : 0u;
  constexpr bool zh = ZH;
  if ((fl & FLAG_INLET3) && !zh) {
#pragma unroll
    for (int i = 0; i < 19; i++) put<S>(dst + size_t(i) * N, ce, fin.v[i], nw, i);
    return;
  }
  float f[19];
  f[0] = S::load(src, ce);
#pragma unroll
  for (int i = 1; i < 19; i++) f[i] = S::load(src, lt.gtab[size_t(i - 1) * nl + k]);
  if (BODY) {
    const uint32_t bm = lt.bmask[k];
    const uint32_t xm = lt.xmask[k];
    const BzLink* __restrict__ rec = lt.links + lt.lbase[k];
    // The links in two halves of the directions (1-9, 10-18), each with its records and second
    // operands loaded together: a body node waits for two rounds of those loads instead of one and
    // keeps half the operands in registers. The kernel's registers are sized for its heaviest path,
    // and this one would otherwise set them for the bulk rows too (with 16-bit storage 7 instead of 8
    // waves per SIMD). Body nodes are a few thousand of millions and their blocks run first,
    // alongside the bulk rows.
    link_half3<S, REC>(src, lt, k, bm, xm, rec, rc, f, force, std::integral_constant<int, 1>{});
    asm volatile("" ::: "memory");
    link_half3<S, REC>(src, lt, k, bm, xm, rec, rc, f, force, std::integral_constant<int, 10>{});
  }
  if (zh && (fl & FLAG_INLET3)) zou_he_face3<true>(f, p.uIn);
  else if (zh && (fl & FLAG_OUTLET3)) zou_he_face3<false>(f, p.uIn);
  const Mom3 m = moments3(f);
  if (REC && rc.mom) put_moments<sizeof(typename S::elem)>(rc.mom, idx, m.drho, m.ux, m.uy, m.uz);
  if ((fl & FLAG_OUTLET3) && !zh) {
    float eq[19];
    outlet_feq3(m, eq);
#pragma unroll
    for (int i = 0; i < 19; i++) put<S>(dst + size_t(i) * N, ce, eq[i], nw, i);
    return;
  }
  const float4 t = xtab[idx % uint32_t(p.W)];
  Relax r;
  r.omBase = t.x;
  r.omM = t.y;
  r.sigma = t.z;
  r.t0 = t.w;
  float out[19];
  collide3<SMAG>(f, m, r, p, out);
#pragma unroll
  for (int i = 0; i < 19; i++) put<S>(dst + size_t(i) * N, ce, out[i], nw, i);
}

// Boundary-list part of the step: list block lb of the launch. List blocks [0, nlb) take the nodes
// without body links; the nbb blocks after them take the nodes with body links and carry the force
// sample (an FP64 reduction over those blocks).
template <class S, bool REC, bool SMAG, bool ZH>
__device__ __forceinline__ void list_part3(const typename S::elem* __restrict__ src, typename S::elem* __restrict__ dst,
                                           const uint32_t* __restrict__ flags, const ListTables& lt, uint32_t nlb,
                                           uint32_t nbb, const Params3& p, const FeqIn& fin,
                                           const float4* __restrict__ xtab, const ForceSink& fs, const Record& rc,
                                           uint32_t lb) {
  float force[3] = {0.0f, 0.0f, 0.0f};
  if (lb < nlb) {
    const uint32_t k = lb * blockDim.x + threadIdx.x;
    if (k < lt.n_link) list_node3<S, false, REC, SMAG, ZH>(src, dst, flags, lt, k, p, fin, xtab, rc, force);
    return;
  }
  const uint32_t kb = (lb - nlb) * blockDim.x + threadIdx.x;
  if (kb < lt.n_body) list_node3<S, true, REC, SMAG, ZH>(src, dst, flags, lt, lt.n_link + kb, p, fin, xtab, rc, force);
  if (fs.sample) force_sample<3>(force, fs, lb - nlb, nbb);
}

// ============================================================================ aux kernels
// src/shaders/aux3d.ts:137-150: (ux, uy, uz, rho), zero for solids.
template <class S>
__global__ __launch_bounds__(256) void macro3_kernel(const typename S::elem* __restrict__ f,
                                                     const uint32_t* __restrict__ flags, float4* __restrict__ mac,
                                                     Params3 p) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= p.N) return;
  if (flags[idx] & FLAG_SOLID3) {
    mac[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  float drho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
  for (int i = 0; i < 19; i++) {
    const float v = S::load(f + size_t(i) * p.NS, idx + p.NA * (idx & ~63u));
    drho += v;
    jx += float(CX3[i]) * v;
    jy += float(CY3[i]) * v;
    jz += float(CZ3[i]) * v;
  }
  const float rho = 1.0f + drho;
  mac[idx] = make_float4(jx / rho, jy / rho, jz / rho, rho);
}

// src/shaders/aux3d.ts:113-123
template <class S>
__global__ __launch_bounds__(256) void init3_kernel(typename S::elem* __restrict__ fa, typename S::elem* __restrict__ fb,
                                                    const uint32_t* __restrict__ flags, Params3 p) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= p.N) return;
  const bool solid = (flags[idx] & FLAG_SOLID3) != 0u;
#pragma unroll
  for (int i = 0; i < 19; i++) {
    const float v = solid ? 0.0f : feq3(i, 1.0f, p.uIn, 0.0f, 0.0f);
    S::store(fa + size_t(i) * p.NS, idx + p.NA * (idx & ~63u), v);
    S::store(fb + size_t(i) * p.NS, idx + p.NA * (idx & ~63u), v);
  }
}

// src/shaders/aux3d.ts:32-100: flags from the SDF and refill of cells the body uncovered. Force
// slots are not needed here (the boundary-list blocks reduce per block), so none are assigned.
template <class S>
__global__ __launch_bounds__(256) void flags3_kernel(const float* __restrict__ sdf, uint32_t* __restrict__ flags,
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
    flags[idx] = FLAG_SOLID3;
    return;
  }
  uint32_t bits = 0u;
  if (x == 0) bits = FLAG_INLET3;
  if (x == W - 1) bits = FLAG_OUTLET3;
#pragma unroll
  for (int i = 1; i < 19; i++) {
    const int sx = x - CX3[i], sy = y - CY3[i], sz = z - CZ3[i];
    if (in_range(sx, sy, sz)) {
      if (solid_at(sx, sy, sz)) bits |= 1u << i;
      continue;
    }
    bits |= 1u << i;
  }
  flags[idx] = bits;
  if (old & FLAG_SOLID3) {
    float rho = 0.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f, n = 0.0f;
#pragma unroll
    for (int i = 1; i < 19; i++) {
      const int qx = x + CX3[i], qy = y + CY3[i], qz = z + CZ3[i];
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
    for (int i = 0; i < 19; i++) {
      const float v = feq3(i, rho, ux, uy, uz);
      S::store(fa + size_t(i) * p.NS, idx + p.NA * (idx & ~63u), v);
      S::store(fb + size_t(i) * p.NS, idx + p.NA * (idx & ~63u), v);
    }
  }
}

__global__ void reset_state_kernel(uint32_t* state) {
  for (int k = 0; k < STATE_WORDS; k++) state[k] = 0u;
}

// ============================================================================ launchers
static inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1) / b; }
uint32_t list_blocks(uint32_t n, uint32_t bx) { return blocks_for(n, bx); }
uint32_t body_list_blocks(uint32_t n_body, uint32_t bx, bool sample) {
  return n_body ? blocks_for(n_body, bx) : (sample ? 1u : 0u);
}

// Threads per block along x: the widest multiple of 64 not above the requested size that divides W
// (no idle lanes), else one block of ceil(W / 64) waves per row.
uint32_t bulk_block_x(int W, int requested) {
  const int req = requested > 0 ? requested : 256;
  for (int b = (req / 64) * 64; b >= 64; b -= 64)
    if (W % b == 0) return uint32_t(b);
  const int whole = ((W + 63) / 64) * 64;
  return uint32_t(whole <= 1024 ? whole : 256);
}

template <class S>
static void step3_t(const void* src, void* dst, const uint32_t* flags, const ListTables& lt, const Params3& p,
                    const float* feq_in, const float4* xtab, const ForceSink& fs, int bb, bool descend, bool merge,
                    bool early, const uint8_t* row_clean, hipStream_t st, const Record* rec) {
  using E = typename S::elem;
  FeqIn fin;
  for (int i = 0; i < 19; i++) fin.v[i] = feq_in[i];
  const uint32_t bx = bulk_block_x(p.W, bb);
  const uint32_t gx = (uint32_t(p.W) + bx - 1) / bx;
  const uint32_t nlb = list_blocks(lt.n_link, bx);
  const uint32_t nbb = body_list_blocks(lt.n_body, bx, fs.sample != 0);
  const uint32_t per_slice = gx * uint32_t(p.H);
  const uint32_t zl = (nlb + nbb + per_slice - 1) / per_slice;
  const dim3 grid(gx, uint32_t(p.H), zl + uint32_t(p.D));
  const uint32_t desc = descend ? 1u : 0u;
  const E* s0 = static_cast<const E*>(src);
  E* d0 = static_cast<E*>(dst);
  const Record none{nullptr, nullptr};
#define LBM_STEP3(...) \
  __VA_ARGS__<<<grid, bx, 0, st>>>(s0, d0, flags, xtab, p, fin, lt, zl, nlb, nbb, fs, rec ? *rec : none, desc, row_clean)
  // Recording steps (the adjoint) never use the subgrid model or the Zou-He faces; the early-pull
  // order is compiled for the kernels without recording and without the subgrid model.
  if (rec && merge)
    LBM_STEP3(step3_kernel<S, true, false, true, false>);
  else if (rec)
    LBM_STEP3(step3_kernel<S, true, false, false, false>);
  else if (p.smagC2 > 0.0f && p.faces)
    LBM_STEP3(step3_kernel<S, false, true, false, true>);
  else if (p.smagC2 > 0.0f)
    LBM_STEP3(step3_kernel<S, false, true, false, false>);
  else if (p.faces && merge) {
    if (early)
      LBM_STEP3(step3_kernel<S, false, false, true, true, true>);
    else
      LBM_STEP3(step3_kernel<S, false, false, true, true>);
  } else if (p.faces) {
    if (early)
      LBM_STEP3(step3_kernel<S, false, false, false, true, true>);
    else
      LBM_STEP3(step3_kernel<S, false, false, false, true>);
  } else if (merge) {
    if (early)
      LBM_STEP3(step3_kernel<S, false, false, true, false, true>);
    else
      LBM_STEP3(step3_kernel<S, false, false, true, false>);
  } else {
    if (early)
      LBM_STEP3(step3_kernel<S, false, false, false, false, true>);
    else
      LBM_STEP3(step3_kernel<S, false, false, false, false>);
  }
#undef LBM_STEP3
}

void launch_step3(int prec, const void* src, void* dst, const uint32_t* flags, const ListTables& lt,
                  const Params3& p, const float* feq_in, const float4* xtab, const ForceSink& fs, int bulk_block,
                  bool descend, bool merge, bool early, const uint8_t* row_clean, hipStream_t st,
                  const Record* rec) {
  const int bb = bulk_block;
  switch (prec) {
    case 1: step3_t<StoreF16>(src, dst, flags, lt, p, feq_in, xtab, fs, bb, descend, merge, early, row_clean, st, rec); break;
    case 2: step3_t<StoreBF16>(src, dst, flags, lt, p, feq_in, xtab, fs, bb, descend, merge, early, row_clean, st, rec); break;
    case 3: step3_t<StoreF16SR>(src, dst, flags, lt, p, feq_in, xtab, fs, bb, descend, merge, early, row_clean, st, rec); break;
    default: step3_t<StoreF32>(src, dst, flags, lt, p, feq_in, xtab, fs, bb, descend, merge, early, row_clean, st, rec); break;
  }
}

void launch_xtab3(const Params3& p, float4* tab, hipStream_t st) {
  xtab3_kernel<<<blocks_for(uint32_t(p.W), 256), 256, 0, st>>>(p, tab);
}

template <class S>
static void macro3_t(const void* f, const uint32_t* flags, float4* mac, const Params3& p, hipStream_t st) {
  macro3_kernel<S><<<blocks_for(p.N, 256), 256, 0, st>>>(static_cast<const typename S::elem*>(f), flags, mac, p);
}
void launch_macro3(int prec, const void* f, const uint32_t* flags, float4* mac, const Params3& p, hipStream_t st) {
  switch (prec) {
    case 1:  // FP16 with stochastic rounding: same format, initial and refilled values round to nearest
    case 3: macro3_t<StoreF16>(f, flags, mac, p, st); break;
    case 2: macro3_t<StoreBF16>(f, flags, mac, p, st); break;
    default: macro3_t<StoreF32>(f, flags, mac, p, st); break;
  }
}

template <class S>
static void init3_t(void* fa, void* fb, const uint32_t* flags, const Params3& p, hipStream_t st) {
  using E = typename S::elem;
  init3_kernel<S><<<blocks_for(p.N, 256), 256, 0, st>>>(static_cast<E*>(fa), static_cast<E*>(fb), flags, p);
}
void launch_init3(int prec, void* fa, void* fb, const uint32_t* flags, const Params3& p, hipStream_t st) {
  switch (prec) {
    case 1:  // FP16 with stochastic rounding: same format, initial and refilled values round to nearest
    case 3: init3_t<StoreF16>(fa, fb, flags, p, st); break;
    case 2: init3_t<StoreBF16>(fa, fb, flags, p, st); break;
    default: init3_t<StoreF32>(fa, fb, flags, p, st); break;
  }
}

template <class S>
static void flags3_t(const float* sdf, uint32_t* flags, void* fa, void* fb, const float4* mac, const Params3& p,
                     hipStream_t st) {
  using E = typename S::elem;
  flags3_kernel<S><<<blocks_for(p.N, 256), 256, 0, st>>>(sdf, flags, static_cast<E*>(fa), static_cast<E*>(fb), mac, p);
}
void launch_flags3(int prec, const float* sdf, uint32_t* flags, void* fa, void* fb, const float4* mac,
                   const Params3& p, hipStream_t st) {
  switch (prec) {
    case 1:  // FP16 with stochastic rounding: same format, initial and refilled values round to nearest
    case 3: flags3_t<StoreF16>(sdf, flags, fa, fb, mac, p, st); break;
    case 2: flags3_t<StoreBF16>(sdf, flags, fa, fb, mac, p, st); break;
    default: flags3_t<StoreF32>(sdf, flags, fa, fb, mac, p, st); break;
  }
}

void launch_reset_state(uint32_t* state, hipStream_t st) { reset_state_kernel<<<1, 1, 0, st>>>(state); }

}  // namespace lbm
