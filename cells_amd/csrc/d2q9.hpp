// D2Q9 per-cell device code shared by the single-tunnel kernels (kernels2d.hip) and the batched
// kernel (batch2d.hip): equilibrium, inlet profile, Zou-He, the x-only relaxation coefficients,
// the TRT collision and the per-direction boundary pulls. Restates src/shaders/step.ts:9-229 and
// src/shaders/common.ts:84-102.
#pragma once
#include "kernels.hpp"

namespace lbm {

// src/shaders/aux.ts:15-18
__device__ __forceinline__ float feq2(int i, float rho, float ux, float uy) {
  const float cu = 3.0f * (float(CX2[i]) * ux + float(CY2[i]) * uy);
  return W2[i] * (rho - 1.0f + rho * (cu + 0.5f * cu * cu - 1.5f * (ux * ux + uy * uy)));
}

// src/shaders/common.ts:84-90
__device__ __forceinline__ float inlet_velocity(int y, const Params2& p) {
  if (p.parabolic == 0) return p.uIn;
  const float h = float(p.H);
  const float s = float(y) + 0.5f;
  return 4.0f * p.uIn * s * (h - s) / (h * h);
}

// src/shaders/step.ts:197-209: Zou-He reconstruction of the populations entering from outside.
__device__ __forceinline__ void zou_he_inlet(float (&f)[9], int y, const Params2& p) {
  const float u = inlet_velocity(y, p);
  const float r = (1.0f + f[0] + f[2] + f[4] + 2.0f * (f[3] + f[6] + f[7])) / (1.0f - u);
  f[1] = f[3] + 0.6666666666666666f * r * u;
  f[5] = f[7] - 0.5f * (f[2] - f[4]) + 0.16666666666666666f * r * u;
  f[8] = f[6] + 0.5f * (f[2] - f[4]) + 0.16666666666666666f * r * u;
}
__device__ __forceinline__ void zou_he_outlet(float (&f)[9]) {
  const float u = f[0] + f[2] + f[4] + 2.0f * (f[1] + f[5] + f[8]);
  f[3] = f[1] - 0.6666666666666666f * u;
  f[7] = f[5] + 0.5f * (f[2] - f[4]) - 0.16666666666666666f * u;
  f[6] = f[8] - 0.5f * (f[2] - f[4]) - 0.16666666666666666f * u;
}

// The adjoint's per-cell record: moments of the populations the collision takes (collide2's drho,
// ux, uy without the body-force shift, which the adjoint does not support).
__device__ __forceinline__ void record2(const float (&f)[9], const Record& rc, uint32_t idx) {
  const float drho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
  const float invRho = 1.0f / (1.0f + drho);
  static_cast<float4*>(rc.mom)[idx] = make_float4(drho, (f[1] - f[3] + f[5] - f[6] - f[7] + f[8]) * invRho,
                            (f[2] - f[4] + f[5] + f[6] - f[7] - f[8]) * invRho, 0.0f);
}

// Relaxation coefficients that depend only on x: omega_base (sponge ramp of tau), omega_minus for
// the no-Smagorinsky case, the absorbing-layer strength sigma (outlet sponge and inlet layer) and
// 1 / omega_base (src/shaders/step.ts:215-223, 48-63).
struct Relax2 {
  float omBase, omM, sigma, t0;
};

__device__ __forceinline__ Relax2 relax2(int x, const Params2& p) {
  Relax2 r;
  float omegaBase = 1.0f / p.tau;
  float sOut = 0.0f;
  const float wm1 = float(p.W - 1);
  if (p.spongeStart < wm1) {
    sOut = smoothstep(p.spongeStart, wm1, float(x));
    omegaBase = 1.0f / wmix(p.tau, p.tauSponge, sOut);
  }
  float sIn = 0.0f;
  if (p.inletLayer > 0.0f) sIn = 1.0f - smoothstep(0.0f, p.inletLayer, float(x));
  r.omBase = omegaBase;
  r.sigma = p.absorb * fmaxf(sOut * sOut, sIn * sIn);
  r.t0 = 1.0f / omegaBase;
  r.omM = 1.0f / (0.5f + p.lambda / (r.t0 - 0.5f));
  return r;
}

// Moments, TRT collision with sponge, inlet layer, Smagorinsky, absorbing layers and Guo forcing:
// src/shaders/step.ts:22-103, 211-226.
__device__ __forceinline__ void collide2(const float (&f)[9], const Relax2& r, int y, const Params2& p, float (&out)[9]) {
  const float drho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
  const float rho = 1.0f + drho;
  const float invRho = 1.0f / rho;
  float ux = (f[1] - f[3] + f[5] - f[6] - f[7] + f[8]) * invRho;
  float uy = (f[2] - f[4] + f[5] + f[6] - f[7] - f[8]) * invRho;
  if (p.forcing != 0) {
    ux += 0.5f * p.fx * invRho;
    uy += 0.5f * p.fy * invRho;
  }
  const float usq = 1.5f * (ux * ux + uy * uy);
  float cu[9], eqS[9], eqA[9];
  cu[1] = 3.0f * (ux);
  cu[2] = 3.0f * (uy);
  cu[5] = 3.0f * (ux + uy);
  cu[6] = 3.0f * (-ux + uy);
  constexpr int PA[4] = {1, 2, 5, 6};
  constexpr int PB[4] = {3, 4, 7, 8};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int a = PA[k];
    eqS[a] = W2[a] * (drho + rho * (0.5f * cu[a] * cu[a] - usq));
    eqA[a] = W2[a] * rho * cu[a];
  }
  const float omegaBase = r.omBase;
  const float sigma = r.sigma;
  float omP = omegaBase;
  if (p.smagC2 > 0.0f) {
    float n[9];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int a = PA[k], b = PB[k];
      n[a] = f[a] - (eqS[a] + eqA[a]);
      n[b] = f[b] - (eqS[a] - eqA[a]);
    }
    const float pxx = n[1] + n[3] + n[5] + n[6] + n[7] + n[8];
    const float pyy = n[2] + n[4] + n[5] + n[6] + n[7] + n[8];
    const float pxy = n[5] - n[6] + n[7] - n[8];
    const float pi = sqrtf(pxx * pxx + pyy * pyy + 2.0f * pxy * pxy);
    const float t0 = r.t0;
    const float tEff = 0.5f * (t0 + sqrtf(t0 * t0 + 25.455844122715714f * p.smagC2 * pi * invRho));
    omP = 1.0f / tEff;
  }
  const float omM = p.smagC2 > 0.0f ? 1.0f / (0.5f + p.lambda / (1.0f / omP - 0.5f)) : r.omM;
  float s0 = 0.0f;
  float sS[9], sA[9];
#pragma unroll
  for (int k = 0; k < 4; k++) sS[PA[k]] = sA[PA[k]] = 0.0f;
  if (sigma > 0.0f) {
    const float ut = inlet_velocity(y, p);
    const float dq = 1.5f * rho * (ut * ut - ux * ux - uy * uy);
    s0 += sigma * W2[0] * -dq;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int a = PA[k];
      const float cut = 3.0f * (float(CX2[a]) * ut);
      sS[a] += sigma * W2[a] * (0.5f * rho * (cut * cut - cu[a] * cu[a]) - dq);
      sA[a] += sigma * W2[a] * rho * (cut - cu[a]);
    }
  }
  if (p.forcing != 0) {
    const float uF = ux * p.fx + uy * p.fy;
    s0 += (1.0f - 0.5f * omP) * W2[0] * (-3.0f * uF);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int a = PA[k];
      const float cF = (float(CX2[a]) * p.fx + float(CY2[a]) * p.fy);
      sS[a] += (1.0f - 0.5f * omP) * W2[a] * (-3.0f * uF + 9.0f * (cu[a] / 3.0f) * cF);
      sA[a] += (1.0f - 0.5f * omM) * W2[a] * 3.0f * cF;
    }
  }
  out[0] = f[0] - omP * (f[0] - W2[0] * (drho - rho * usq)) + s0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int a = PA[k], b = PB[k];
    const float sym = omP * (0.5f * (f[a] + f[b]) - eqS[a]);
    const float asym = omM * (0.5f * (f[a] - f[b]) - eqA[a]);
    out[a] = f[a] - sym - asym + sS[a] + sA[a];
    out[b] = f[b] - sym + asym + sS[a] - sA[a];
  }
}

// One plain fluid / plain Zou-He node of the bulk pass (src/shaders/step.ts:177-227 for flags that
// are 0 or only the off-grid x bits of an open column). Per direction the source address is a
// wave-uniform base (population i, shifted by whole rows) plus the lane's byte offset.
template <bool REC>
__device__ __forceinline__ void bulk_cell2(const float* __restrict__ src, float* __restrict__ dst,
                                           const uint32_t* __restrict__ flags, const float4* __restrict__ xtab,
                                           const Params2& p, uint32_t x, int y, uint32_t idx, const Record& rc) {
  // On a row with a row above and below, every plain pull address lies in the buffer (at x = 0 and
  // x = W - 1 they reach into the neighbouring rows), so the loads go out before the flag word is
  // known and their latency overlaps its; the boundary rows read the flag word first.
  const bool interior = y > 0 && y + 1 < p.H;
  // The column's relaxation coefficients go out first (x < W): waiting for them after the node's kind
  // is known would add a second memory round trip to every wave, which small grids (one wave per
  // SIMD, nothing else to overlap) pay in full.
  const float4 t = xtab[x];
  uint32_t fl = 0u;
  bool inlet = false, outlet = false;
  if (!interior) {
    fl = flags[idx];
    inlet = fl == (FLAG_INLET2 | INLET_DIRS2);
    outlet = fl == (FLAG_OUTLET2 | OUTLET_DIRS2);
    if (fl != 0u && !inlet && !outlet) return;
  }
  const size_t N = p.NS;  // population stride
  // Element of population i at cell c: c + NA (c & ~63) + NS i (one array per population, or tiles of
  // 64 cells when W % 64 == 0, so that a wave is one tile): the y shift moves whole tiles and stays in
  // the wave-uniform base, an x shift crosses into the previous / next tile only at the wave's first /
  // last lane (offp, offm). At x = 0 and x = W - 1 those pulls read the node's own element: the
  // populations entering from off the grid are placeholders (step.ts:173-174) that Zou-He overwrites,
  // and a bulk node there is always a plain inlet or outlet node.
  const uint32_t off = (idx + p.NA * (idx & ~63u)) * 4u;
  const uint32_t tile_step = 64u * p.NA * 4u;
  const uint32_t offp = x == 0u ? off : off - 4u - ((idx & 63u) == 0u ? tile_step : 0u);
  const uint32_t offm = x == uint32_t(p.W - 1) ? off : off + 4u + ((idx & 63u) == 63u ? tile_step : 0u);
  const char* const sb = reinterpret_cast<const char*>(src);
  float f[9];
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const ptrdiff_t dy = ptrdiff_t(CY2[i]) * p.W;
    const char* base = sb + (ptrdiff_t(i) * ptrdiff_t(N) - ptrdiff_t(1u + p.NA) * dy) * 4;
    uint32_t o = CX2[i] > 0 ? offp : CX2[i] < 0 ? offm : off;
    // On the boundary rows the off-grid populations of an inlet / outlet node read the node's own
    // element, so the address stays on the grid (on interior rows any address above is in the buffer).
    if ((inlet && CX2[i] > 0) || (outlet && CX2[i] < 0)) o = off + uint32_t(ptrdiff_t(1u + p.NA) * dy) * 4u;
    f[i] = *reinterpret_cast<const float*>(base + o);
  }
  if (interior) {
    fl = flags[idx];
    inlet = fl == (FLAG_INLET2 | INLET_DIRS2);
    outlet = fl == (FLAG_OUTLET2 | OUTLET_DIRS2);
    if (fl != 0u && !inlet && !outlet) return;
  }
  if (inlet) zou_he_inlet(f, y, p);
  if (outlet) zou_he_outlet(f);
  if (REC && rc.mom) record2(f, rc, idx);
  Relax2 r;
  r.omBase = t.x;
  r.omM = t.y;
  r.sigma = t.z;
  r.t0 = t.w;
  float out[9];
  collide2(f, r, y, p, out);
  char* const d = reinterpret_cast<char*>(dst);
#pragma unroll
  for (int i = 0; i < 9; i++) *reinterpret_cast<float*>(d + i * N * 4 + off) = out[i];
}

// One node of the boundary list (k is its position in the list). Every pull is a gather through the
// table the host resolved from pullSlow's rules (src/shaders/step.ts:119-175), so the 8 loads are
// independent; a moving-lid pull adds 6 w_i c_ix uLid (step.ts:167). Nodes with body links (BODY)
// then apply their Bouzidi links (step.ts:138-160): the gathered f_j becomes the interpolation
// against the record's second operand, and c_j (f_j + f_i) goes into the force. Zou-He and the
// collision follow as in the bulk pass.
template <bool BODY, bool REC>
__device__ __forceinline__ void list_cell2(const float* __restrict__ src, float* __restrict__ dst,
                                           const uint32_t* __restrict__ flags, const float4* __restrict__ xtab,
                                           const ListTables& lt, uint32_t k, const Params2& p, const Record& rc,
                                           float (&force)[2]) {
  const uint32_t N = p.NS;  // population stride
  const uint32_t nl = lt.n_link + lt.n_body;
  const uint32_t idx = lt.list[k];
  const uint32_t fl = flags[idx];
  const int y = int(idx / uint32_t(p.W));
  const int x = int(idx - uint32_t(y) * uint32_t(p.W));
  float f[9];
  const uint32_t ce = idx + p.NA * (idx & ~63u);  // the node's element within each population
  f[0] = src[ce];
#pragma unroll
  for (int i = 1; i < 9; i++) f[i] = src[lt.gtab[size_t(i - 1) * nl + k]];
  if (BODY) {
    const uint32_t bm = lt.bmask[k];
    const BzLink* __restrict__ rec = lt.links + lt.lbase[k];
    float q[9], b[9];
    uint32_t gb[9];
#pragma unroll
    for (int i = 1; i < 9; i++) {
      q[i] = 1.0f;
      gb[i] = NO_OPERAND;
      if (bm & (1u << i)) {
        const BzLink r = rec[__popc(bm & ((1u << i) - 1u))];
        q[i] = r.q;
        gb[i] = r.gb;
      }
    }
#pragma unroll
    for (int i = 1; i < 9; i++) b[i] = gb[i] != NO_OPERAND ? src[gb[i]] : 0.0f;
#pragma unroll
    for (int i = 1; i < 9; i++) {
      if (bm & (1u << i)) {
        const int j = OPP2[i];
        const float fj = f[i];
        const float qi = q[i];
        float fi = fj;
        if (qi < 0.5f) {
          if (gb[i] != NO_OPERAND) fi = 2.0f * qi * fj + (1.0f - 2.0f * qi) * b[i];
        } else {
          fi = fj / (2.0f * qi) + (2.0f * qi - 1.0f) / (2.0f * qi) * b[i];
        }
        force[0] += float(CX2[j]) * (fj + fi);
        force[1] += float(CY2[j]) * (fj + fi);
        if (REC && rc.dq) rc.dq[lt.lbase[k] + __popc(bm & ((1u << i) - 1u))] = bouzidi_dq(qi, fj, b[i], gb[i] != NO_OPERAND);
        f[i] = fi;
      }
    }
  }
  const uint32_t mv = lt.xmask[k];
  if (mv) {
#pragma unroll
    for (int i = 1; i < 9; i++)
      if (mv & (1u << i)) f[i] = f[i] + 6.0f * W2[i] * float(CX2[i]) * p.uLid;
  }
  if (fl & FLAG_INLET2) zou_he_inlet(f, y, p);
  else if (fl & FLAG_OUTLET2) zou_he_outlet(f);
  if (REC && rc.mom) record2(f, rc, idx);
  const float4 t = xtab[x];
  Relax2 r;
  r.omBase = t.x;
  r.omM = t.y;
  r.sigma = t.z;
  r.t0 = t.w;
  float out[9];
  collide2(f, r, y, p, out);
#pragma unroll
  for (int i = 0; i < 9; i++) dst[size_t(i) * N + ce] = out[i];
}

}  // namespace lbm
