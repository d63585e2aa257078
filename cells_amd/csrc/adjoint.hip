// Discrete adjoint of the tunnel step (no counterpart in the reference): reverse-time kernels that
// carry dJ/d(populations) back through the stream-collide step, with the derivatives of the Bouzidi
// interpolation and of the momentum-exchange force with respect to each link fraction q, and the host
// driver that recomputes the forward pass from checkpoints.
//
// Notation. F_n are the stored populations entering step n (n = 0 .. T-1), P the pull (every rule of
// src/shaders/step.ts:119-175 / step3d.ts:119-168, linear in F_n for fixed q), B the boundary
// reconstruction (2D Zou-He columns, step.ts:197-209; linear), C the collision (TRT with sponge and
// absorbing layer, step.ts:22-103 / step3d.ts:38-102). A fluid node x gets
// F_{n+1}(x) = C(B(P F_n)(x)), the 3D inlet plane gets the inlet equilibrium and the 3D outlet plane
// feq(1, u(P F_n)) (step3d.ts:190-209). The force sample of step n is linear in F_n:
// sum over links of c_j (f_j + f_i) (step.ts:138-158), and J = sum_n w_n (d . F_sample(n)).
//
// The adjoint state is mu_n = dJ/d(P F_n), stored like the populations (Q x N, FP32). One reverse
// step, for every fluid node y:
//   lambda(y) = (P^T mu_{n+1})(y) + w_{n+1} * (force terms of step n+1 that read F_{n+1}(y)),
//   mu_n(y)   = B^T C'(m_n(y))^T lambda(y),
// where m_n(y) = (drho, u) are the moments of the collision input that the forward step recorded.
// C' depends on the populations only through these moments, so the record is 16 bytes per cell
// (8 with 16-bit population storage, whose moments are kept as halves) instead of the Q populations. For a Bouzidi link (x, i) with operands f_j, b:
//   dJ/dq += (d f_i / d q)_n * (mu_n[i](x) + w_n c_j . d),
// accumulated in FP64 per link; the host chains dq/dsdf at the link's two end nodes.
//
// P^T is a gather, not a scatter: a node whose every population is read exactly once, by the plain
// pull of its downstream neighbour, takes mu_{n+1}[a](y + c_a) (the adjoint bulk, a shifted stream
// like the forward one); every other fluid node (boundaries, body links, mirrored faces, the inlet
// and outlet planes) gets a host-built list of (source, weight, force weight) entries.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lbm.h"
#include "d2q9.hpp"
#include "solver_internal.hpp"

namespace lbm {

struct AdjEntry {
  uint32_t src;  // index into mu_{n+1} (population * N + cell)
  uint32_t a;    // population of the target node that receives it
  float w;       // weight of mu_{n+1}[src]
  float v;       // weight of w_{n+1} (force direction folded in)
};

// Adjoint-list node kinds.
constexpr uint32_t AK_PLAIN = 0, AK_ZOU_IN = 1, AK_ZOU_OUT = 2, AK_INLET3 = 3, AK_OUTLET3 = 4;

struct AdjTables {
  const uint8_t* aflag;   // per cell: 0 = adjoint bulk node
  const uint32_t* aidx;   // adjoint-list cells
  const uint32_t* akind;  // their kinds
  const uint32_t* ae0;    // node k has ae0[k + 1] - ae0[k] entries
  const AdjEntry* ent;    // entry j of node k at ent[j * n_list + k]
  const uint32_t* alk;    // forward list position of the node when it has Bouzidi links, else NO_OPERAND
  const uint32_t* bmask;  // forward tables (Bouzidi directions, first record)
  const uint32_t* lbase;
  uint32_t n_list;
};

struct AdjStep {
  const float* mu_next;  // mu_{n+1}, population a of cell u at a * Ns + u
  uint32_t Ns;           // population stride of the adjoint state (mu_stride)
  float* mu;             // mu_n
  const void* mom;       // record of step n: moments per cell (float4, or Moments16 when half_mom)
  uint32_t half_mom;
  const float* dq;       // record of step n: d f_i / d q per link
  double* acc;           // dJ/dq per link
  float w_next, w_cur;   // objective weights of the force samples of steps n+1 and n
  float ut;              // inlet speed of step n
  float dx, dy, dz;      // force direction of the objective
};

// The record of cell idx as float4 (drho, ux, uy, uz).
__device__ __forceinline__ float4 get_moments(const AdjStep& as, uint32_t idx) {
  if (as.half_mom) {
    const Moments16 h = static_cast<const Moments16*>(as.mom)[idx];
    return make_float4(float(h.drho), float(h.ux), float(h.uy), float(h.uz));
  }
  return static_cast<const float4*>(as.mom)[idx];
}

// ---------------------------------------------------------------------------- collision VJPs
// mu = C'^T lambda for the TRT collision at moments (drho, u) with relaxation rates omP (symmetric),
// omM (antisymmetric), absorbing strength sigma and target inflow speed ut (x component only).
// Direct part: the rest population relaxes with omP, each pair (a, b = opposite) through its
// symmetric and antisymmetric halves. Moment part: the scalar Phi(drho, u) = sum over outputs of
// lambda . (equilibrium and source terms) is differentiated in drho and u, then chained through
// drho = sum f and u = j / rho: d/df_i = dPhi/ddrho + (c_i - u) . dPhi/du / rho.
// Lattice accessors by population count (D2Q9 / D3Q19), folded at compile time in unrolled loops.
template <int Q>
__device__ __forceinline__ constexpr int cvec(int d, int i) {
  return Q == 19 ? (d == 0 ? CX3[i] : d == 1 ? CY3[i] : CZ3[i]) : (d == 0 ? CX2[i] : d == 1 ? CY2[i] : 0);
}
template <int Q>
__device__ __forceinline__ constexpr float wgt(int i) {
  return Q == 19 ? W3[i] : W2[i];
}
// TRT pairs (a, b = opposite of a): D3Q19 (2k+1, 2k+2); D2Q9 (1,3), (2,4), (5,7), (6,8).
template <int Q>
__device__ __forceinline__ constexpr int pair_a(int k) {
  return Q == 19 ? 2 * k + 1 : (k < 2 ? k + 1 : k + 3);
}
template <int Q>
__device__ __forceinline__ constexpr int pair_b(int k) {
  return Q == 19 ? 2 * k + 2 : (k < 2 ? k + 3 : k + 5);
}
template <int Q>
__device__ __forceinline__ constexpr int opposite(int i) {
  return Q == 19 ? opp3(i) : OPP2[i];
}

template <int Q>
__device__ __forceinline__ void trt_vjp(const float (&lam)[Q], float drho, const float (&u)[3], float omP, float omM,
                                        float sigma, float ut, float (&mu)[Q]) {
  constexpr int NP = (Q - 1) / 2;
  const float rho = 1.0f + drho;
  const float invRho = 1.0f / rho;
  const float u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const float usq = 1.5f * u2;
  const float dut = 1.5f * (ut * ut - u2);
  mu[0] = (1.0f - omP) * lam[0];
  float gd = omP * lam[0] * wgt<Q>(0) * (1.0f - usq) - lam[0] * sigma * wgt<Q>(0) * dut;
  const float E = lam[0] * wgt<Q>(0) * (sigma - omP);
  float P = 0.0f;
  float g[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int k = 0; k < NP; k++) {
    const int a = pair_a<Q>(k), b = pair_b<Q>(k);
    const float Ls = lam[a] + lam[b], La = lam[a] - lam[b];
    mu[a] = lam[a] - 0.5f * omP * Ls - 0.5f * omM * La;
    mu[b] = lam[b] - 0.5f * omP * Ls + 0.5f * omM * La;
    const float cu = 3.0f * (float(cvec<Q>(0, a)) * u[0] + float(cvec<Q>(1, a)) * u[1] + float(cvec<Q>(2, a)) * u[2]);
    const float cut = 3.0f * float(cvec<Q>(0, a)) * ut;
    const float wa = wgt<Q>(a);
    const float A = omP * Ls * wa, B = omM * La * wa, Cs = sigma * Ls * wa, Ca = sigma * La * wa;
    gd += A * (1.0f + 0.5f * cu * cu - usq) + B * cu + Cs * (0.5f * (cut * cut - cu * cu) - dut) + Ca * (cut - cu);
    P += A - Cs;
    const float e = (A - Cs) * cu + B - Ca;
#pragma unroll
    for (int d = 0; d < 3; d++) g[d] += float(cvec<Q>(d, a)) * e;
  }
  float gu[3];
#pragma unroll
  for (int d = 0; d < 3; d++) gu[d] = 3.0f * rho * (u[d] * (E - P) + g[d]);
  const float ug = u[0] * gu[0] + u[1] * gu[1] + u[2] * gu[2];
#pragma unroll
  for (int i = 0; i < Q; i++)
    mu[i] += gd + (float(cvec<Q>(0, i)) * gu[0] + float(cvec<Q>(1, i)) * gu[1] + float(cvec<Q>(2, i)) * gu[2] - ug) *
                      invRho;
}

// mu = (d feq(1, u(f)) / df)^T lambda for the 3D outlet plane (step3d.ts:205-209).
__device__ __forceinline__ void outlet3_vjp(const float (&lam)[19], float drho, const float (&u)[3], float (&mu)[19]) {
  float gu[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < 19; i++) {
    const float cu = 3.0f * (float(CX3[i]) * u[0] + float(CY3[i]) * u[1] + float(CZ3[i]) * u[2]);
    const float t = lam[i] * W3[i] * 3.0f;
    gu[0] += t * (float(CX3[i]) * (1.0f + cu) - u[0]);
    gu[1] += t * (float(CY3[i]) * (1.0f + cu) - u[1]);
    gu[2] += t * (float(CZ3[i]) * (1.0f + cu) - u[2]);
  }
  const float invRho = 1.0f / (1.0f + drho);
  const float ug = u[0] * gu[0] + u[1] * gu[1] + u[2] * gu[2];
#pragma unroll
  for (int i = 0; i < 19; i++)
    mu[i] = (float(CX3[i]) * gu[0] + float(CY3[i]) * gu[1] + float(CZ3[i]) * gu[2] - ug) * invRho;
}

// Transposes of the Zou-He reconstructions (d2q9.hpp zou_he_inlet / zou_he_outlet): nu is the
// adjoint of the reconstructed populations, the result that of the pulled ones.
__device__ __forceinline__ void zou_he_inlet_vjp(float (&nu)[9], float u) {
  const float c = (nu[1] * (2.0f / 3.0f) * u + (nu[5] + nu[8]) * u * (1.0f / 6.0f)) / (1.0f - u);
  const float m0 = nu[0] + c, m2 = nu[2] - 0.5f * nu[5] + 0.5f * nu[8] + c, m4 = nu[4] + 0.5f * nu[5] - 0.5f * nu[8] + c;
  const float m3 = nu[3] + nu[1] + 2.0f * c, m6 = nu[6] + nu[8] + 2.0f * c, m7 = nu[7] + nu[5] + 2.0f * c;
  nu[0] = m0;
  nu[2] = m2;
  nu[4] = m4;
  nu[3] = m3;
  nu[6] = m6;
  nu[7] = m7;
  nu[1] = nu[5] = nu[8] = 0.0f;
}
__device__ __forceinline__ void zou_he_outlet_vjp(float (&nu)[9]) {
  const float c = -(2.0f / 3.0f) * nu[3] - (nu[7] + nu[6]) * (1.0f / 6.0f);
  const float m0 = nu[0] + c, m2 = nu[2] + 0.5f * nu[7] - 0.5f * nu[6] + c, m4 = nu[4] - 0.5f * nu[7] + 0.5f * nu[6] + c;
  const float m1 = nu[1] + nu[3] + 2.0f * c, m5 = nu[5] + nu[7] + 2.0f * c, m8 = nu[8] + nu[6] + 2.0f * c;
  nu[0] = m0;
  nu[2] = m2;
  nu[4] = m4;
  nu[1] = m1;
  nu[5] = m5;
  nu[8] = m8;
  nu[3] = nu[6] = nu[7] = 0.0f;
}

// ---------------------------------------------------------------------------- adjoint list node
// kind decides the local transpose; ut is the inlet speed seen by this node (the 2D parabolic profile
// is evaluated by the caller).
template <int Q>
__device__ __forceinline__ void adj_list_node(const AdjTables& at, const AdjStep& as, const float4* __restrict__ xtab,
                                              uint32_t W, uint32_t N, uint32_t k, float ut) {
  const uint32_t idx = at.aidx[k];
  const uint32_t kind = at.akind[k];
  float lam[Q];
#pragma unroll
  for (int a = 0; a < Q; a++) lam[a] = 0.0f;
  const uint32_t e1 = at.ae0[k + 1] - at.ae0[k];  // entry count; entry j is at.ent[j * n_list + k]
  // Entries four at a time: their table loads, then their gathers, are issued together instead of
  // one dependent pair per entry; the sums still run in entry order.
  for (uint32_t e = 0; e < e1; e += 4) {
    AdjEntry en[4];
    float mv[4];
#pragma unroll
    for (int t = 0; t < 4; t++)
      if (e + t < e1) en[t] = at.ent[size_t(e + t) * at.n_list + k];
#pragma unroll
    for (int t = 0; t < 4; t++) mv[t] = e + t < e1 ? as.mu_next[en[t].src] : 0.0f;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      if (e + t >= e1) break;
      const float val = en[t].w * mv[t] + as.w_next * en[t].v;
#pragma unroll
      for (int a = 0; a < Q; a++)
        if (en[t].a == uint32_t(a)) lam[a] += val;
    }
  }
  float mu[Q];
  const float4 m = get_moments(as, idx);
  const float u[3] = {m.y, m.z, Q == 19 ? m.w : 0.0f};
  if (kind == AK_INLET3) {
#pragma unroll
    for (int i = 0; i < Q; i++) mu[i] = 0.0f;
  } else if (kind == AK_OUTLET3) {
    if constexpr (Q == 19) outlet3_vjp(lam, m.x, u, mu);
  } else {
    const float4 t = xtab[idx % W];
    trt_vjp<Q>(lam, m.x, u, t.x, t.y, t.z, ut, mu);
    if constexpr (Q == 9) {
      if (kind == AK_ZOU_IN) zou_he_inlet_vjp(mu, ut);
      else if (kind == AK_ZOU_OUT) zou_he_outlet_vjp(mu);
    }
  }
#pragma unroll
  for (int i = 0; i < Q; i++) as.mu[size_t(i) * N + idx] = mu[i];
  const uint32_t lk = at.alk[k];
  if (lk != NO_OPERAND) {
    const uint32_t bm = at.bmask[lk];
    const uint32_t base = at.lbase[lk];
#pragma unroll
    for (int i = 1; i < Q; i++) {
      if (bm & (1u << i)) {
        const int j = opposite<Q>(i);
        const float cd = float(cvec<Q>(0, j)) * as.dx + float(cvec<Q>(1, j)) * as.dy + float(cvec<Q>(2, j)) * as.dz;
        const uint32_t l = base + __popc(bm & ((1u << i) - 1u));
        as.acc[l] += double(as.dq[l]) * double(mu[i] + as.w_cur * cd);
      }
    }
  }
}

// ---------------------------------------------------------------------------- adjoint step kernels
// Grid (gx, H, zl + D) in 3D, (gx, yl + H) in 2D: list blocks first, then the bulk rows. In 3D the
// rows are walked top-down on every other reverse step (descend), so that a step starts on the rows
// the previous one wrote last, as the forward step does.
__global__ __launch_bounds__(1024) void adj3_kernel(AdjTables at, AdjStep as, const float4* __restrict__ xtab,
                                                    Params3 p, uint32_t zl, uint32_t descend) {
  const uint32_t N = as.Ns;  // the adjoint state's population stride
  if (blockIdx.z < zl) {
    const uint32_t k = ((blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < at.n_list) adj_list_node<19>(at, as, xtab, uint32_t(p.W), N, k, as.ut);
    return;
  }
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= uint32_t(p.W)) return;
  const uint32_t y = descend ? uint32_t(p.H) - 1u - blockIdx.y : blockIdx.y;
  const uint32_t z = descend ? uint32_t(p.D) - 1u - (blockIdx.z - zl) : blockIdx.z - zl;
  const uint32_t idx = x + uint32_t(p.W) * (y + uint32_t(p.H) * z);
  // Rows off the y and z faces: every node's 19 downstream neighbours lie in the buffer (x = 0 and
  // x = W - 1 reach into the adjacent rows), so the loads are issued before the node's flag is
  // known and its latency overlaps theirs. Face rows test the flag first.
  const bool inner = y > 0u && y + 1u < uint32_t(p.H) && z > 0u && z + 1u < uint32_t(p.D);
  if (!inner && at.aflag[idx]) return;
  const uint32_t off = idx * 4u;
  const char* const mb = reinterpret_cast<const char*>(as.mu_next);
  // On a y / z face row a population that would stream off the grid is read back from the node and
  // direction the forward step's mirrored (one face) or bounce-back (two faces) pull takes it to;
  // the pattern is uniform per row (the host admits only nodes that follow it).
  const bool loY = y == 0u, hiY = y + 1u == uint32_t(p.H), loZ = z == 0u, hiZ = z + 1u == uint32_t(p.D);
  float lam[19];
#pragma unroll
  for (int a = 0; a < 19; a++) {
    const bool oY = (loY && CY3[a] < 0) || (hiY && CY3[a] > 0);
    const bool oZ = (loZ && CZ3[a] < 0) || (hiZ && CZ3[a] > 0);
    int pa = a;
    ptrdiff_t delta = ptrdiff_t(CZ3[a]) * p.WH + ptrdiff_t(CY3[a]) * p.W + CX3[a];
    if (oY && oZ) {
      pa = opp3(a);
      delta = 0;
    } else if (oY) {
      pa = mirror_y3(a);
      delta = ptrdiff_t(CZ3[a]) * p.WH + CX3[a];
    } else if (oZ) {
      pa = mirror_z3(a);
      delta = ptrdiff_t(CY3[a]) * p.W + CX3[a];
    }
    const char* base = mb + (ptrdiff_t(pa) * ptrdiff_t(N) + delta) * 4;
    lam[a] = *reinterpret_cast<const float*>(base + off);
  }
  // The x planes next to the open ends (bulk_readers3 on the host admits a node only if its
  // populations are read exactly so): at x = 1 nothing reads the populations with c_x < 0 (the
  // inlet plane pulls nothing); at x = W - 1 nothing reads those with c_x > 0 (off the grid); at
  // x = W - 2 the outlet node of the same row also reads those with c_x < 0 (its copy rule,
  // src/shaders/step3d.ts:130), in the same direction.
  const uint32_t W = uint32_t(p.W);
#pragma unroll
  for (int a = 1; a < 19; a++)
    if ((x == 1u && CX3[a] < 0) || (x == W - 1u && CX3[a] > 0)) lam[a] = 0.0f;
  if (x == W - 2u) {
#pragma unroll
    for (int a = 1; a < 19; a++)
      if (CX3[a] < 0) lam[a] += as.mu_next[size_t(a) * N + idx + 1u];
  }
  const float4 m = get_moments(as, idx);
  const float4 t = xtab[x];
  if (inner && at.aflag[idx]) return;
  const float u[3] = {m.y, m.z, m.w};
  float mu[19];
  if (x == 0u) {  // the inlet plane stores the inflow equilibrium whatever it pulled: mu = 0
#pragma unroll
    for (int i = 0; i < 19; i++) mu[i] = 0.0f;
  } else if (x == W - 1u) {
    outlet3_vjp(lam, m.x, u, mu);
  } else {
    trt_vjp<19>(lam, m.x, u, t.x, t.y, t.z, as.ut, mu);
  }
  char* const d = reinterpret_cast<char*>(as.mu);
#pragma unroll
  for (int i = 0; i < 19; i++) *reinterpret_cast<float*>(d + size_t(i) * N * 4 + off) = mu[i];
}

__device__ __forceinline__ float inlet_profile2(int y, float u, const Params2& p) {
  if (p.parabolic == 0) return u;
  const float h = float(p.H);
  const float s = float(y) + 0.5f;
  return 4.0f * u * s * (h - s) / (h * h);
}

__global__ __launch_bounds__(1024) void adj2_kernel(AdjTables at, AdjStep as, const float4* __restrict__ xtab,
                                                    Params2 p, uint32_t yl) {
  const uint32_t N = as.Ns;  // the adjoint state's population stride
  if (blockIdx.y < yl) {
    const uint32_t k = (blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < at.n_list) {
      const int y = int(at.aidx[k] / uint32_t(p.W));
      adj_list_node<9>(at, as, xtab, uint32_t(p.W), N, k, inlet_profile2(y, as.ut, p));
    }
    return;
  }
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= uint32_t(p.W)) return;
  const uint32_t y = blockIdx.y - yl;
  const uint32_t idx = x + uint32_t(p.W) * y;
  if (at.aflag[idx]) return;
  float lam[9];
#pragma unroll
  for (int a = 0; a < 9; a++) {
    const int64_t nb = int64_t(idx) + int64_t(CY2[a]) * p.W + CX2[a];  // the downstream neighbour (in the grid)
    lam[a] = as.mu_next[size_t(a) * N + size_t(nb)];
  }
  const float4 m = get_moments(as, idx);
  const float u[3] = {m.y, m.z, 0.0f};
  const float4 t = xtab[x];
  float mu[9];
  trt_vjp<9>(lam, m.x, u, t.x, t.y, t.z, inlet_profile2(int(y), as.ut, p), mu);
#pragma unroll
  for (int i = 0; i < 9; i++) as.mu[size_t(i) * N + idx] = mu[i];
}

}  // namespace lbm

// ============================================================================ host side
using namespace lbm;

namespace {
// Bytes of one cell's record: the moments are kept in the width of the population storage (a
// 16-bit store holds them as halves, rounded to nearest; FP32 storage as floats).
size_t mom_bytes(const lbm_solver* s) { return s->elem == 2 ? sizeof(Moments16) : sizeof(float4); }
}  // namespace

struct lbm_adjoint {
  lbm_adjoint_config cfg{};
  bool have_forward = false;
  uint64_t forward_gen = 0;  // geometry (list tables generation) of the forward run
  int64_t T = 0, K = 0, nseg = 0, nstored = 0;  // segments [nseg - nstored, nseg) were recorded by the forward pass
  uint32_t Ns = 0;  // population stride of the adjoint state mu (mu_stride)
  // The forward pass also records the segment before them into the recomputation slot (idle until the
  // backward pass needs it); valid until a backward pass recomputes another segment there.
  bool scratch_valid = false;
  double J = 0;
  std::vector<double> w;  // objective weight of the force sample of step n (n = 0 .. T), w[T] = 0
  // device memory
  std::vector<void*> ckpt;          // populations entering step seg * K
  void* final_state = nullptr;      // populations after step T (restored after the backward pass)
  std::vector<void*> rec_mom;       // per segment slot: K x N moments (mom_bytes each)
  std::vector<float*> rec_dq;       // per segment slot: K x n_links
  float* mu[2] = {nullptr, nullptr};
  double* acc = nullptr;
  // tables
  uint64_t tables_gen = ~0ull;
  uint8_t* aflag = nullptr;
  uint32_t *aidx = nullptr, *akind = nullptr, *ae0 = nullptr, *alk = nullptr;
  AdjEntry* ent = nullptr;
  uint32_t n_list = 0;
  size_t n_ent = 0;
  int64_t n_regular = 0;
  std::vector<int64_t> layout;  // sizes the device buffers were allocated for
};

namespace lbm {

namespace {

void free_tables(lbm_adjoint* a) {
  for (void* p : {(void*)a->aflag, (void*)a->aidx, (void*)a->akind, (void*)a->ae0, (void*)a->alk, (void*)a->ent})
    if (p) (void)hipFree(p);
  a->aflag = nullptr;
  a->aidx = a->akind = a->ae0 = a->alk = nullptr;
  a->ent = nullptr;
  a->tables_gen = ~0ull;
}

void free_buffers(lbm_adjoint* a) {
  for (void* p : a->ckpt)
    if (p) (void)hipFree(p);
  a->ckpt.clear();
  for (void* p : a->rec_mom)
    if (p) (void)hipFree(p);
  a->rec_mom.clear();
  for (float* p : a->rec_dq)
    if (p) (void)hipFree(p);
  a->rec_dq.clear();
  for (void* p : {a->final_state, (void*)a->mu[0], (void*)a->mu[1], (void*)a->acc})
    if (p) (void)hipFree(p);
  a->final_state = nullptr;
  a->mu[0] = a->mu[1] = nullptr;
  a->acc = nullptr;
  a->have_forward = false;
  a->layout.clear();
}

// One forward pull source: population index, weight and force weight (direction folded in).
struct Src {
  uint32_t src;
  float w, v;
};

// The sources of pull (x, i) of fluid node idx as elements of the solver's population layout (the
// forward tables' indices); returns how many (0 when the reconstruction discards the pull). pos is
// the node's forward list position or NO_OPERAND for bulk nodes.
int pull_sources_at_stride(const lbm_solver* s, uint32_t idx, uint32_t pos, int i, const double* dir, Src* out) {
  const bool d2 = s->nd == 2;
  const uint32_t fl = s->h_flags[idx];
  if (d2) {
    if ((fl & FLAG_INLET2) && (INLET_DIRS2 & (1u << i))) return 0;
    if ((fl & FLAG_OUTLET2) && (OUTLET_DIRS2 & (1u << i))) return 0;
  }
  if (i == 0) {
    out[0] = {pelem(s, 0, idx), 1.0f, 0.0f};
    return 1;
  }
  if (pos == NO_OPERAND) {
    if (d2) {
      out[0] = {pelem(s, i, idx - uint32_t(CY2[i] * s->W + CX2[i])), 1.0f, 0.0f};
    } else {
      const uint32_t WH = uint32_t(s->W) * uint32_t(s->H);
      const Row3 r = row3(s, (idx % WH) / uint32_t(s->W), idx / WH);
      out[0] = {row_gather3(s, r, idx, (fl & FLAG_OUTLET3) != 0u, i), 1.0f, 0.0f};
    }
    return 1;
  }
  const uint32_t nl = uint32_t(s->h_list.size());
  const uint32_t g = s->h_gtab[size_t(i - 1) * nl + pos];
  const uint32_t bm = s->h_bmask[pos], xm = s->h_xmask[pos];
  const int j = d2 ? OPP2[i] : opp3(i);
  const double cd = d2 ? CX2[j] * dir[0] + CY2[j] * dir[1] : CX3[j] * dir[0] + CY3[j] * dir[1] + CZ3[j] * dir[2];
  if (bm & (1u << i)) {
    const BzLink& L = s->h_links[s->h_lbase[pos] + __builtin_popcount(bm & ((1u << i) - 1u))];
    const float q = L.q;
    float al, be;
    if (q < 0.5f) {
      al = L.gb != NO_OPERAND ? 2.0f * q : 1.0f;
      be = L.gb != NO_OPERAND ? 1.0f - 2.0f * q : 0.0f;
    } else {
      al = 1.0f / (2.0f * q);
      be = (2.0f * q - 1.0f) / (2.0f * q);
    }
    out[0] = {g, al, float((1.0 + al) * cd)};
    if (L.gb != NO_OPERAND && be != 0.0f) {
      out[1] = {L.gb, be, float(be * cd)};
      return 2;
    }
    return 1;
  }
  if (!d2 && (xm & (1u << i))) {
    out[0] = {g, 1.0f, float(2.0 * cd)};
    return 1;
  }
  out[0] = {g, 1.0f, 0.0f};
  return 1;
}

// pull_sources with every source as population * N + cell: the forward tables carry the solver's
// population layout, the adjoint tables index by N (and re-stride to their own state's stride).
int pull_sources(const lbm_solver* s, uint32_t idx, uint32_t pos, int i, const double* dir, Src* out) {
  const int n = pull_sources_at_stride(s, idx, pos, i, dir, out);
  if (s->NA != 0 || s->NS != s->N) {
    for (int e = 0; e < n; e++) {
      int pi;
      uint32_t c;
      pelem_decode(s->NA, s->NS, s->Q, out[e].src, &pi, &c);
      out[e].src = uint32_t(pi) * s->N + c;
    }
  }
  return n;
}

// Where the reverse step's bulk rows take lambda_a(u) from: the node and direction whose pull reads
// the post-collision f_a(u). Off the faces that is the plain pull of u + c_a in direction a. On a 3D
// slip-face row (y = 0 or H-1, z = 0 or D-1) a population that would stream off the grid across one
// face is read by the forward step's mirrored pull (src/shaders/step3d.ts:132-145) of the node
// u + c_a reflected back onto the row, in the mirrored direction; across two faces, by u's own
// bounce-back pull in the opposite direction. The same rule, per row, drives adj3_kernel; the host
// admits a node to the bulk rows only if every population of it is read exactly this way.
void expected_reader(const lbm_solver* s, uint32_t u, int a, uint32_t* x, int* i) {
  const int64_t W = s->W, H = s->H, D = s->D;
  const int64_t ux = u % W, uy = (u / W) % H, uz = u / (W * H);
  if (s->nd == 2) {
    const int64_t vx = ux + CX2[a], vy = uy + CY2[a];
    if (vx < 0 || vx >= W || vy < 0 || vy >= H) return;
    *x = uint32_t(vx + W * vy);
    *i = a;
    return;
  }
  const bool oY = (uy == 0 && CY3[a] < 0) || (uy == H - 1 && CY3[a] > 0);
  const bool oZ = (uz == 0 && CZ3[a] < 0) || (uz == D - 1 && CZ3[a] > 0);
  int64_t vx = ux + CX3[a], vy = uy + CY3[a], vz = uz + CZ3[a];
  int ia = a;
  if (oY && oZ) {
    vx = ux, vy = uy, vz = uz;
    ia = opp3(a);
  } else if (oY) {
    vy = uy;
    ia = mirror_y3(a);
  } else if (oZ) {
    vz = uz;
    ia = mirror_z3(a);
  }
  if (vx < 0 || vx >= W || vy < 0 || vy >= H || vz < 0 || vz >= D) return;
  *x = uint32_t(vx + W * (vy + H * vz));
  *i = ia;
}

// The readers adj3_kernel's streamed rows assume for population a of node u (x = 1 .. W - 1): the
// expected_reader, except at x = 1 for c_x < 0 (the inlet plane pulls nothing) and at x = W - 1 for
// c_x > 0 (off the grid); at x = W - 2 for c_x < 0 also the outlet node of the same row, whose copy
// rule reads u in direction a. Returns their number (at most 2).
int bulk_readers3(const lbm_solver* s, uint32_t u, int a, uint32_t* xs, int* is) {
  const uint32_t W = uint32_t(s->W), ux = u % W;
  int n = 0;
  uint32_t ex = NO_OPERAND;
  int ei = -1;
  expected_reader(s, u, a, &ex, &ei);
  if (ex != NO_OPERAND && !(ux == 1u && CX3[a] < 0) && !(ux == W - 1u && CX3[a] > 0)) {
    xs[n] = ex;
    is[n++] = ei;
  }
  if (ux + 2u == W && CX3[a] < 0) {
    xs[n] = u + 1u;
    is[n++] = a;
  }
  return n;
}

// The index adj3_kernel's streamed rows load for population a of node u (the row pattern of its
// face bits, read backwards, in the adjoint state's stride Ns), whether or not the node is streamed.

// Population stride of the adjoint state: with a stride of N a power of two times a small factor
// (2^22 at 256 x 128 x 128) the Q streams a node touches fall on the same memory channels; a stride
// rounded to 1024 elements plus 17 x 4096 spreads them (the 19-stream copy kernel measured 5.80 TB/s
// at the stride 1,769,472 = 27 x 2^16 and 6.36 TB/s at that stride + 69,632). Small grids keep N.
uint32_t mu_stride(uint32_t N, int Q) {
  if (N < 65536u) return N;
  const uint64_t ns = (uint64_t(N) + 1023u) / 1024u * 1024u + 17u * 4096u;
  return uint64_t(Q) * ns < (uint64_t(1) << 32) ? uint32_t(ns) : N;
}
uint64_t adj3_row_source(const lbm_solver* s, uint32_t Ns, uint32_t u, int a) {
  const int64_t W = s->W, H = s->H, D = s->D, WH = W * H;
  const int64_t y = (u / W) % H, z = u / WH;
  const bool oY = (y == 0 && CY3[a] < 0) || (y == H - 1 && CY3[a] > 0);
  const bool oZ = (z == 0 && CZ3[a] < 0) || (z == D - 1 && CZ3[a] > 0);
  int pa = a;
  int64_t delta = CZ3[a] * WH + CY3[a] * W + CX3[a];
  if (oY && oZ) {
    pa = opp3(a);
    delta = 0;
  } else if (oY) {
    pa = mirror_y3(a);
    delta = CZ3[a] * WH + CX3[a];
  } else if (oZ) {
    pa = mirror_z3(a);
    delta = CY3[a] * W + CX3[a];
  }
  const int64_t v = int64_t(pa) * int64_t(Ns) + int64_t(u) + delta;
  return v < 0 ? ~uint64_t(0) : uint64_t(v);
}

// Builds the adjoint tables for the solver's current geometry (after every rebuild of the list).
int build_tables(lbm_solver* s, lbm_adjoint* a, const double* dir) {
  const uint32_t N = s->N;
  const int Q = s->Q;
  const uint32_t Ns = mu_stride(N, Q);
  const bool d2 = s->nd == 2;
  const uint32_t solid = d2 ? FLAG_SOLID2 : FLAG_SOLID3;
  std::vector<uint32_t> pos(N, NO_OPERAND);
  for (uint32_t k = 0; k < s->h_list.size(); k++) pos[s->h_list[k]] = k;
  auto pulls = [&](uint32_t x) { return !(s->h_flags[x] & solid) && !(!d2 && (s->h_flags[x] & FLAG_INLET3)); };
  // pass 1: per target population, how many sources read it and whether the only one is the read
  // the reverse step's bulk rows assume (bit 7 marks any other kind): the plain pull of the
  // downstream neighbour, or in 3D on a y / z face row the forward step's row pattern read backwards
  // (expected_reader)
  std::vector<uint8_t> cnt(size_t(Q) * N, 0);
  Src src[2];
  for (uint32_t x = 0; x < N; x++) {
    if (!pulls(x)) continue;
    for (int i = 0; i < Q; i++) {
      const int ns = pull_sources(s, x, pos[x], i, dir, src);
      for (int e = 0; e < ns; e++) {
        uint8_t& c = cnt[src[e].src];
        if ((c & 0x7f) < 0x7f) c++;
        const uint32_t a = src[e].src / N, u = src[e].src % N;
        uint32_t rx[2] = {NO_OPERAND, NO_OPERAND};
        int ri[2] = {-1, -1};
        const int nr = d2 ? 0 : bulk_readers3(s, u, int(a), rx, ri);
        if (d2) expected_reader(s, u, int(a), &rx[0], &ri[0]);
        bool expected = false;
        for (int r = 0; r < (d2 ? 1 : nr); r++) expected = expected || (rx[r] == x && ri[r] == i);
        if (!(expected && src[e].w == 1.0f && src[e].v == 0.0f)) c |= 0x80;
      }
    }
  }
  // adjoint bulk nodes: every population read exactly by the readers the streamed rows assume,
  // plainly, and (2D) a plain collision or (3D) the collision, outlet or inlet rule of the node's x
  // plane; nodes with Bouzidi links stay in the list, which accumulates their dJ/dq
  std::vector<uint8_t> aflag(N, 1);
  std::vector<uint32_t> alist;
  a->n_regular = 0;
  const uint32_t Wu = uint32_t(s->W);
  for (uint32_t y = 0; y < N; y++) {
    const uint32_t fl = s->h_flags[y];
    if (fl & solid) continue;
    const bool links = pos[y] != NO_OPERAND && s->h_bmask[pos[y]] != 0u;
    bool reg;
    if (d2) {
      reg = !(fl & (FLAG_INLET2 | FLAG_OUTLET2)) && !links;
      for (int i = 0; i < Q && reg; i++) reg = cnt[size_t(i) * N + y] == 1;
    } else {
      const uint32_t yx = y % Wu;
      const uint32_t kind = (fl & FLAG_INLET3) ? AK_INLET3 : (fl & FLAG_OUTLET3) ? AK_OUTLET3 : AK_PLAIN;
      reg = !links && kind == (yx == 0u ? AK_INLET3 : yx + 1u == Wu ? AK_OUTLET3 : AK_PLAIN);
      if (yx != 0u) {  // the inlet plane's mu is zero whoever reads it
        for (int i = 0; i < Q && reg; i++) {
          uint32_t rx[2];
          int ri[2];
          const int nr = bulk_readers3(s, y, i, rx, ri);
          const uint8_t c = cnt[size_t(i) * N + y];
          reg = (c & 0x80) == 0 && int(c) == nr;
        }
      }
    }
    if (reg) {
      aflag[y] = 0;
      a->n_regular++;
    } else {
      alist.push_back(y);
    }
  }
  // pass 2: the entries of every adjoint-list node, grouped by node in list order
  std::vector<uint32_t> apos(N, NO_OPERAND);
  for (uint32_t k = 0; k < alist.size(); k++) apos[alist[k]] = k;
  std::vector<uint32_t> ne(alist.size() + 1, 0);
  for (uint32_t x = 0; x < N; x++) {
    if (!pulls(x)) continue;
    for (int i = 0; i < Q; i++) {
      const int ns = pull_sources(s, x, pos[x], i, dir, src);
      for (int e = 0; e < ns; e++) {
        const uint32_t k = apos[src[e].src % N];
        if (k != NO_OPERAND) ne[k + 1]++;
      }
    }
  }
  for (size_t k = 0; k < alist.size(); k++) ne[k + 1] += ne[k];
  std::vector<AdjEntry> ent(ne.back());
  std::vector<uint32_t> fill(ne.begin(), ne.end() - 1);
  for (uint32_t x = 0; x < N; x++) {
    if (!pulls(x)) continue;
    for (int i = 0; i < Q; i++) {
      const int ns = pull_sources(s, x, pos[x], i, dir, src);
      for (int e = 0; e < ns; e++) {
        const uint32_t k = apos[src[e].src % N];
        if (k == NO_OPERAND) continue;
        ent[fill[k]++] = AdjEntry{uint32_t(i) * N + x, src[e].src / N, src[e].w, src[e].v};
      }
    }
  }
  std::vector<uint32_t> kind(alist.size()), lk(alist.size());
  for (size_t k = 0; k < alist.size(); k++) {
    const uint32_t y = alist[k], fl = s->h_flags[y];
    if (d2) kind[k] = (fl & FLAG_INLET2) ? AK_ZOU_IN : (fl & FLAG_OUTLET2) ? AK_ZOU_OUT : AK_PLAIN;
    else kind[k] = (fl & FLAG_INLET3) ? AK_INLET3 : (fl & FLAG_OUTLET3) ? AK_OUTLET3 : AK_PLAIN;
    lk[k] = (pos[y] != NO_OPERAND && s->h_bmask[pos[y]]) ? pos[y] : NO_OPERAND;
  }
  // Every index a launch will use must address the grid: the streamed nodes' downstream neighbours
  // and every entry's source and target.
  for (uint32_t y = 0; y < N; y++) {
    if (aflag[y]) continue;
    for (int i = 0; i < Q; i++) {
      uint32_t ex = NO_OPERAND;
      int ei = -1;
      if (d2) {
        expected_reader(s, y, i, &ex, &ei);
        if (ex == NO_OPERAND) return fail(LBM_ERR_DEVICE, "adjoint tables: a streamed node has a reader off the grid");
      } else if (adj3_row_source(s, Ns, y, i) >= uint64_t(Q) * Ns) {
        return fail(LBM_ERR_DEVICE, "adjoint tables: a streamed node's load leaves the buffer");
      }
    }
  }
  for (const AdjEntry& en : ent)
    if (en.src >= uint64_t(Q) * N || en.a >= uint32_t(Q)) return fail(LBM_ERR_DEVICE, "adjoint tables: entry out of range");
  if (a->Ns && a->Ns != Ns) return fail(LBM_ERR_DEVICE, "adjoint tables: state stride changed");
  a->Ns = Ns;
  for (uint32_t v : lk)
    if (v != NO_OPERAND && v >= s->h_list.size()) return fail(LBM_ERR_DEVICE, "adjoint tables: link node out of range");
  free_tables(a);
  const size_t nl = alist.size();
  // Entries column-major ("ELL"): entry j of list node k at j * n_list + k, so the threads of a wave
  // (consecutive list nodes) read their j-th entries from consecutive addresses.
  uint32_t max_e = 0;
  for (size_t k = 0; k < nl; k++) max_e = std::max(max_e, ne[k + 1] - ne[k]);
  std::vector<AdjEntry> ell(size_t(max_e) * nl, AdjEntry{0u, 0u, 0.0f, 0.0f});
  for (size_t k = 0; k < nl; k++)
    for (uint32_t j = 0; j < ne[k + 1] - ne[k]; j++) {
      AdjEntry en = ent[ne[k] + j];
      en.src = (en.src / N) * Ns + en.src % N;  // into the adjoint state's stride
      ell[size_t(j) * nl + k] = en;
    }
  HIP_TRY(hipMalloc(&a->aflag, N));
  HIP_TRY(hipMalloc(&a->aidx, 4 * std::max<size_t>(1, nl)));
  HIP_TRY(hipMalloc(&a->akind, 4 * std::max<size_t>(1, nl)));
  HIP_TRY(hipMalloc(&a->alk, 4 * std::max<size_t>(1, nl)));
  HIP_TRY(hipMalloc(&a->ae0, 4 * (nl + 1)));
  HIP_TRY(hipMalloc(&a->ent, sizeof(AdjEntry) * std::max<size_t>(1, ell.size())));
  HIP_TRY(hipMemcpy(a->aflag, aflag.data(), N, hipMemcpyHostToDevice));
  if (nl) {
    HIP_TRY(hipMemcpy(a->aidx, alist.data(), 4 * nl, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(a->akind, kind.data(), 4 * nl, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(a->alk, lk.data(), 4 * nl, hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemcpy(a->ae0, ne.data(), 4 * (nl + 1), hipMemcpyHostToDevice));
  if (!ell.empty()) HIP_TRY(hipMemcpy(a->ent, ell.data(), sizeof(AdjEntry) * ell.size(), hipMemcpyHostToDevice));
  a->n_list = uint32_t(nl);
  a->n_ent = ent.size();
  a->tables_gen = s->tables_gen;
  return LBM_OK;
}


// One reverse step n on the solver's stream.
int adjoint_step(lbm_solver* s, lbm_adjoint* a, int64_t n, int slot, const float* mu_next, float* mu_cur) {
  AdjTables at{a->aflag, a->aidx, a->akind, a->ae0, a->ent, a->alk, s->bmask, s->lbase, a->n_list};
  AdjStep as;
  as.mu_next = mu_next;
  as.mu = mu_cur;
  as.Ns = a->Ns;
  as.mom = static_cast<const char*>(a->rec_mom[size_t(slot)]) + size_t(n % a->K) * s->N * mom_bytes(s);
  as.half_mom = mom_bytes(s) == sizeof(Moments16);
  as.dq = a->rec_dq[size_t(slot)] + size_t(n % a->K) * std::max<uint32_t>(s->n_links, 1);
  as.acc = a->acc;
  as.w_next = float(a->w[size_t(n + 1)]);
  as.w_cur = float(a->w[size_t(n)]);
  as.ut = float(s->ramp.kind ? s->ramp.at(n) : (s->nd == 2 ? s->cfg2.u_in : s->cfg3.u_in));
  as.dx = float(a->cfg.direction[0]);
  as.dy = float(a->cfg.direction[1]);
  as.dz = float(a->cfg.direction[2]);
  // 3D: whole rows per block (up to 1024 threads), the shape the reverse step runs fastest in
  const uint32_t bx = bulk_block_x(s->W, s->nd == 3 ? 1024 : s->bulk_block);
  const uint32_t gx = (uint32_t(s->W) + bx - 1) / bx;
  const uint32_t lblocks = (a->n_list + bx - 1) / bx;
  if (s->nd == 3) {
    const uint32_t per = gx * uint32_t(s->H);
    const uint32_t zl = (lblocks + per - 1) / per;
    adj3_kernel<<<dim3(gx, uint32_t(s->H), zl + uint32_t(s->D)), bx, 0, s->stream>>>(at, as, s->xtab, s->p3, zl,
                                                                                      uint32_t(n & 1));
  } else {
    const uint32_t yl = (lblocks + gx - 1) / gx;
    adj2_kernel<<<dim3(gx, yl + uint32_t(s->H)), bx, 0, s->stream>>>(at, as, s->xtab, s->p2, yl);
  }
  HIP_TRY(hipGetLastError());
  return LBM_OK;
}

int check_supported(const lbm_solver* s) {
  if (s->nd == 2) {
    if (s->cfg2.smagorinsky > 0) return fail(LBM_ERR_UNSUPPORTED, "adjoint: Smagorinsky is not supported");
    if (s->cfg2.body_force[0] != 0 || s->cfg2.body_force[1] != 0)
      return fail(LBM_ERR_UNSUPPORTED, "adjoint: body force is not supported");
  } else if (s->cfg3.smagorinsky > 0) {
    return fail(LBM_ERR_UNSUPPORTED, "adjoint: Smagorinsky is not supported");
  } else if (s->Q != 19) {
    return fail(LBM_ERR_UNSUPPORTED, "adjoint: D3Q19 only");
  } else if (s->faces != LBM_FACES_EQUILIBRIUM) {
    return fail(LBM_ERR_UNSUPPORTED, "adjoint: only the equilibrium inlet and outlet planes are supported");
  }
  return LBM_OK;
}

}  // namespace

void adjoint_free(lbm_solver* s) {
  if (!s->adj) return;
  free_buffers(s->adj);
  free_tables(s->adj);
  delete s->adj;
  s->adj = nullptr;
}

}  // namespace lbm

extern "C" {

int lbm_adjoint_forward(lbm_solver* s, const lbm_adjoint_config* cfg, double* J) {
  if (!s || !cfg || !J) return fail(LBM_ERR_INVALID, "bad arguments");
  if (cfg->steps <= 0 || cfg->window_start < 0 || cfg->window_start >= cfg->steps || cfg->checkpoint_every <= 0)
    return fail(LBM_ERR_INVALID, "adjoint: need 0 <= window_start < steps and checkpoint_every > 0");
  if (int rc = check_usable(s)) return rc;
  if (int rc = check_supported(s)) return rc;
  if (!s->adj) s->adj = new lbm_adjoint();
  lbm_adjoint* a = s->adj;
  const int64_t T = cfg->steps, K = cfg->checkpoint_every, nseg = (T + K - 1) / K;
  const size_t fbytes = pop_elems(s) * s->elem;  // whole population buffers, in the solver's layout
  const size_t nlk = std::max<uint32_t>(s->n_links, 1);
  const size_t seg_bytes = size_t(K) * s->N * mom_bytes(s) + size_t(K) * nlk * 4;
  const int64_t nstored = std::min<int64_t>(nseg, std::max<int64_t>(0, cfg->record_budget_bytes) / int64_t(seg_bytes));
  // The device buffers are kept from one forward run to the next while their layout is unchanged
  // (the usual case of finite differences or an optimisation loop over one tunnel).
  const std::vector<int64_t> layout{T, K, nstored, int64_t(s->N), int64_t(fbytes), int64_t(nlk)};
  a->have_forward = false;
  if (layout != a->layout) {
    free_buffers(a);
    a->ckpt.assign(size_t(nseg), nullptr);
    for (int64_t g = 0; g < nseg - nstored; g++) HIP_TRY(hipMalloc(&a->ckpt[size_t(g)], fbytes));
    // slots [0, nstored) hold the recorded segments for good; one more slot takes the recomputed ones
    const int64_t nslots = nstored + (nstored < nseg ? 1 : 0);
    a->rec_mom.assign(size_t(nslots), nullptr);
    a->rec_dq.assign(size_t(nslots), nullptr);
    for (int64_t g = 0; g < nslots; g++) {
      HIP_TRY(hipMalloc(&a->rec_mom[size_t(g)], size_t(K) * s->N * mom_bytes(s)));
      HIP_TRY(hipMalloc(&a->rec_dq[size_t(g)], size_t(K) * nlk * 4));
    }
    HIP_TRY(hipMalloc(&a->final_state, fbytes));
    a->layout = layout;
  }
  a->cfg = *cfg;
  a->T = T;
  a->K = K;
  a->nseg = nseg;
  a->nstored = nstored;
  // the run: from the initial field, with the configured inlet ramp and a force sample every step
  if (int rc = lbm_set_inlet_ramp(s, cfg->ramp_kind, cfg->ramp_u, cfg->ramp_steps)) return rc;
  // The 3D initial field is the equilibrium at the inlet speed of the first step (rest under a ramp
  // from zero), whatever speed earlier steps left in the parameters.
  if (s->nd == 3) s->p3.uIn = float(s->ramp.kind ? s->ramp.at(0) : s->cfg3.u_in);
  if (int rc = lbm_init_field(s, nullptr, 0)) return rc;
  const int64_t count = a->T - cfg->window_start;
  a->w.assign(size_t(a->T + 1), 0.0);
  for (int64_t n = cfg->window_start; n < a->T; n++) a->w[size_t(n)] = cfg->scale / double(count);
  std::vector<int64_t> st(LBM_HISTORY_LEN);
  std::vector<double> vals(size_t(LBM_HISTORY_LEN) * 3);
  double sum = 0;
  auto drain = [&]() -> int {
    int64_t got = 0;
    if (int rc = lbm_read_forces(s, st.data(), vals.data(), LBM_HISTORY_LEN, &got)) return rc;
    const int dim = s->dim();
    for (int64_t k = 0; k < got; k++) {
      if (st[size_t(k)] <= cfg->window_start) continue;  // sample k is taken in step st - 1
      double dot = 0;
      for (int d = 0; d < dim; d++) dot += cfg->direction[d] * vals[size_t(k) * dim + d];
      sum += dot;
    }
    return LBM_OK;
  };
  const int64_t first_stored = a->nseg - a->nstored;
  a->scratch_valid = false;
  for (int64_t n = 0; n < a->T; n++) {
    const int64_t g = n / a->K;
    if (n % a->K == 0 && g < first_stored)
      HIP_TRY(hipMemcpyAsync(a->ckpt[size_t(g)], s->f[s->parity], fbytes, hipMemcpyDeviceToDevice, s->stream));
    if (g + 1 >= first_stored) {
      // the segment before the recorded ones goes to the recomputation slot (index nstored)
      const size_t slot = g >= first_stored ? size_t(g - first_stored) : size_t(a->nstored);
      Record rec{static_cast<char*>(a->rec_mom[slot]) + size_t(n % a->K) * s->N * mom_bytes(s),
                 a->rec_dq[slot] + size_t(n % a->K) * std::max<uint32_t>(s->n_links, 1)};
      if (int rc = enqueue_step(s, &rec, 1)) return rc;
    } else {
      // The recording kernel with nothing to record: the same compiled arithmetic as the steps that
      // record (here and when the backward pass recomputes this segment), so J and the recomputed
      // trajectory are bit for bit those of the recorded plan.
      const Record none{nullptr, nullptr};
      if (int rc = enqueue_step(s, &none, 1)) return rc;
    }
    HIP_TRY(hipGetLastError());
    if ((n + 1) % 4096 == 0)
      if (int rc = drain()) return rc;
  }
  if (int rc = drain()) return rc;
  a->scratch_valid = first_stored > 0;
  HIP_TRY(hipMemcpyAsync(a->final_state, s->f[s->parity], fbytes, hipMemcpyDeviceToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  a->J = cfg->scale * sum / double(count);
  a->have_forward = true;
  a->forward_gen = s->tables_gen;
  *J = a->J;
  return LBM_OK;
}

int lbm_adjoint_backward(lbm_solver* s, double seed, double* dJ_dsdf, size_t n) {
  if (!s || !dJ_dsdf || n < s->N) return fail(LBM_ERR_INVALID, "dJ_dsdf must hold N doubles");
  if (int rc = check_usable(s)) return rc;
  lbm_adjoint* a = s->adj;
  if (!a || !a->have_forward) return fail(LBM_ERR_INVALID, "adjoint: call lbm_adjoint_forward first");
  if (a->forward_gen != s->tables_gen)
    return fail(LBM_ERR_INVALID, "adjoint: the geometry changed since the forward run; run lbm_adjoint_forward again");
  if (a->tables_gen != s->tables_gen)
    if (int rc = build_tables(s, a, a->cfg.direction)) return rc;
  const uint32_t N = s->N;
  const size_t fbytes = pop_elems(s) * s->elem;
  const size_t nlk = std::max<uint32_t>(s->n_links, 1);
  for (float*& m : a->mu) {
    if (!m) HIP_TRY(hipMalloc(&m, size_t(s->Q) * a->Ns * 4));
    HIP_TRY(hipMemsetAsync(m, 0, size_t(s->Q) * a->Ns * 4, s->stream));
  }
  if (!a->acc) HIP_TRY(hipMalloc(&a->acc, nlk * 8));
  HIP_TRY(hipMemsetAsync(a->acc, 0, nlk * 8, s->stream));
  // weights scaled by the upstream gradient
  const std::vector<double> w0 = a->w;
  for (double& x : a->w) x *= seed;
  const int64_t first_stored = a->nseg - a->nstored;
  int cur = 0;  // mu[cur] holds mu_{n+1}
  int rc = LBM_OK;
  for (int64_t g = a->nseg - 1; g >= 0 && rc == LBM_OK; g--) {
    const int64_t n0 = g * a->K, n1 = std::min(a->T, n0 + a->K);
    int slot;
    if (g >= first_stored) {
      slot = int(g - first_stored);
    } else if (g + 1 == first_stored && a->scratch_valid) {
      slot = int(a->nstored);  // recorded by the forward pass
    } else {
      // recompute the segment from its checkpoint into the scratch slot
      slot = int(a->nstored);
      a->scratch_valid = false;
      if (hipMemcpyAsync(s->f[0], a->ckpt[size_t(g)], fbytes, hipMemcpyDeviceToDevice, s->stream) != hipSuccess) {
        rc = fail(LBM_ERR_DEVICE, "adjoint: checkpoint restore failed");
        break;
      }
      s->parity = 0;
      s->step = n0;
      for (int64_t k = n0; k < n1 && rc == LBM_OK; k++) {
        Record rec{static_cast<char*>(a->rec_mom[size_t(slot)]) + size_t(k - n0) * N * mom_bytes(s),
                   a->rec_dq[size_t(slot)] + size_t(k - n0) * nlk};
        rc = enqueue_step(s, &rec, 0);
      }
    }
    for (int64_t k = n1 - 1; k >= n0 && rc == LBM_OK; k--) {
      rc = adjoint_step(s, a, k, slot, a->mu[cur], a->mu[cur ^ 1]);
      cur ^= 1;
    }
  }
  a->w = w0;
  // restore the state after the forward run (the recomputation moved the inlet parameter too)
  if (s->nd == 2) s->p2.uIn = float(s->ramp.kind ? s->ramp.at(a->T - 1) : s->cfg2.u_in);
  else s->p3.uIn = float(s->ramp.kind ? s->ramp.at(a->T - 1) : s->cfg3.u_in);
  if (hipMemcpyAsync(s->f[0], a->final_state, fbytes, hipMemcpyDeviceToDevice, s->stream) != hipSuccess && rc == LBM_OK)
    rc = fail(LBM_ERR_DEVICE, "adjoint: state restore failed");
  s->parity = 0;
  s->step = a->T;
  if (rc != LBM_OK) return rc;
  std::vector<double> acc(nlk);
  HIP_TRY(hipMemcpyAsync(acc.data(), a->acc, nlk * 8, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  // chain dJ/dq through q = clamp(dF / max(dF - dS, 1e-6), 1e-3, 1) at the link's two end nodes
  std::fill(dJ_dsdf, dJ_dsdf + N, 0.0);
  const uint32_t W = uint32_t(s->W), H = uint32_t(s->H), WH = W * H;
  for (uint32_t l = 0; l < s->n_links; l++) {
    const uint32_t x = s->h_link_node[l];
    const int i = s->h_link_dir[l];
    const int cx = s->nd == 2 ? CX2[i] : CX3[i], cy = s->nd == 2 ? CY2[i] : CY3[i], cz = s->nd == 2 ? 0 : CZ3[i];
    int64_t sx = int64_t(x % W) - cx, sy = int64_t((x / W) % H) - cy;
    const int64_t sz = int64_t(x / WH) - cz;
    if (s->nd == 2) {  // periodic wrap of the solid neighbour as in resolve2
      if (sx < 0 && s->cfg2.left == 0) sx += W;
      if (sx >= int64_t(W) && s->cfg2.right == 0) sx -= W;
      if (sy < 0 && s->cfg2.bottom == 0) sy += H;
      if (sy >= int64_t(H) && s->cfg2.top == 0) sy -= H;
    }
    const uint32_t sc = uint32_t(sx + int64_t(W) * (sy + int64_t(H) * sz));
    const double dF = s->sdf_host[x], dS = s->sdf_host[sc];
    const double den = dF - dS;
    if (den < 1e-6) continue;
    const double r = dF / den;
    if (r < 1e-3 || r > 1.0) continue;
    dJ_dsdf[x] += acc[l] * (-dS / (den * den));
    dJ_dsdf[sc] += acc[l] * (dF / (den * den));
  }
  return LBM_OK;
}

int lbm_adjoint_info(lbm_solver* s, int64_t* regular_nodes, int64_t* list_nodes, int64_t* entries,
                     int64_t* stored_segments, int64_t* segments) {
  if (!s || !s->adj) return fail(LBM_ERR_INVALID, "adjoint: no adjoint state");
  lbm_adjoint* a = s->adj;
  if (a->tables_gen != s->tables_gen)
    if (int rc = build_tables(s, a, a->cfg.direction)) return rc;
  if (regular_nodes) *regular_nodes = a->n_regular;
  if (list_nodes) *list_nodes = a->n_list;
  if (entries) *entries = int64_t(a->n_ent);
  if (stored_segments) *stored_segments = a->nstored;
  if (segments) *segments = a->nseg;
  return LBM_OK;
}

int lbm_adjoint_release(lbm_solver* s) {
  if (!s) return fail(LBM_ERR_INVALID, "bad arguments");
  if (s->stream) HIP_TRY(hipStreamSynchronize(s->stream));
  adjoint_free(s);
  return LBM_OK;
}

}  // extern "C"
