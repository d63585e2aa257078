// CPU restatement of the reference's D2Q9 and D3Q19 solvers (test infrastructure only).
//
// This is the oracle the HIP kernels are checked against. Only tests, the smoke check and the
// CPU-baseline leg of the benchmark load it; the product never does. It restates, operation
// for operation and in the same order, the WGSL kernels of the reference:
//   2D step      src/shaders/step.ts:9-229      (fast pull 11-15, pullSlow 119-175, main 177-227)
//   2D aux       src/shaders/aux.ts:14-172      (feq 14-19, flags 22-89, init 92-112, macro 115-141,
//                                                reduce 144-172)
//   3D step      src/shaders/step3d.ts:24-225   (pullSlow3 119-168, main 170-223)
//   3D aux       src/shaders/aux3d.ts:20-183
//   host         src/solver.ts:180-335, src/solver3d.ts:197-315 (param derivations, step loop,
//                force ring bookkeeping)
// Real is the compute type (float reproduces the reference's f32 arithmetic, double is the FP64
// accuracy reference); Store is the storage codec of the distribution buffers (codec.hpp).
// Deviations, all deliberate: force slots are numbered in cell order (the reference's atomicAdd
// order is nondeterministic) and force samples are summed in double.
#pragma once
#include <cmath>
#include <cstdint>
#include <deque>
#include <string>
#include <vector>

#include "../include/lbm.h"
#include "codec.hpp"
#include "lattice.hpp"

namespace orc {

struct Ramp {
  int kind = 0;  // 0 off, 1 linear, 2 smoothstep
  double u_target = 0;
  double steps = 0;
  // Inlet speed for the step that advances the solver from step k to k + 1.
  double at(int64_t k) const {
    double r = steps > 0 ? std::min(1.0, double(k) / steps) : 1.0;
    if (kind == 2) r = r * r * (3 - 2 * r);
    return u_target * r;
  }
};

// Host-side force ring bookkeeping of src/solver.ts:236-241, 279-284, 315-335.
struct ForceBook {
  int64_t samples_read = 0;
  std::deque<int64_t> steps;
  int64_t generation = 0;
  void reset() {
    samples_read = 0;
    steps.clear();
    generation++;
  }
  // ring: HISTORY_LEN * dim doubles; written: device sample counter.
  int64_t consume(const double* ring, int dim, int64_t written, int64_t* out_steps, double* out_vals,
                  int64_t cap) {
    if (steps.empty()) return 0;
    int64_t first = std::max(samples_read, written - int64_t(LBM_HISTORY_LEN));
    int64_t n = 0;
    for (int64_t k = first; k < written && n < cap; k++, n++) {
      int64_t slot = k % LBM_HISTORY_LEN;
      out_steps[n] = steps[size_t(k - samples_read)];
      for (int d = 0; d < dim; d++) out_vals[n * dim + d] = ring[slot * dim + d];
    }
    int64_t drop = std::min<int64_t>(written - samples_read, int64_t(steps.size()));
    steps.erase(steps.begin(), steps.begin() + drop);
    samples_read = written;
    return n;
  }
};

// ============================================================================================
// D2Q9
// ============================================================================================
template <class T, class S>
struct Solver2D {
  using E = typename S::elem;
  int W, H;
  int64_t N;
  lbm_config2d cfg;
  // Params uniform values (src/shaders/common.ts:55-63), derived by writeParams (solver.ts:180-197).
  struct Params {
    int left, right, bottom, top, parabolic, forcing;
    T tau, lambda, smagC2, uIn, spongeStart, tauSponge, uLid, fx, fy, absorb, inletLayer;
  } P;
  std::vector<E> f[2];
  std::vector<uint32_t> flags;
  std::vector<float> sdf;
  std::vector<T> mac;        // N * 4
  std::vector<T> cellForce;  // N * 2
  uint32_t counter = 0;
  std::vector<double> history;  // HISTORY_LEN * 2
  uint32_t state = 0;
  int64_t step = 0;
  int parity = 0;
  ForceBook book;
  Ramp ramp;

  explicit Solver2D(const lbm_config2d& c) : W(c.width), H(c.height), N(int64_t(c.width) * c.height), cfg(c) {
    f[0].assign(size_t(9 * N), S::template store<T>(T(0)));
    f[1] = f[0];
    flags.assign(size_t(N), 0u);
    sdf.assign(size_t(N), 1e6f);
    mac.assign(size_t(N * 4), T(0));
    cellForce.assign(size_t(N * 2), T(0));
    history.assign(LBM_HISTORY_LEN * 2, 0.0);
    write_params();
    rebuild_flags();
  }

  void write_params() {
    const lbm_config2d& c = cfg;
    P.left = c.left;
    P.right = c.right;
    P.bottom = c.bottom;
    P.top = c.top;
    P.tau = T(c.tau);
    P.lambda = T(c.lambda);
    P.smagC2 = T(c.smagorinsky * c.smagorinsky);
    P.uIn = T(c.u_in);
    P.parabolic = c.parabolic ? 1 : 0;
    P.spongeStart = T(c.sponge_fraction > 0 ? (W - 1) * (1 - c.sponge_fraction) : 1e9);
    P.tauSponge = T(std::max(c.sponge_tau, c.tau));
    P.uLid = T(c.u_lid);
    P.fx = T(c.body_force[0]);
    P.fy = T(c.body_force[1]);
    P.forcing = (c.body_force[0] != 0 || c.body_force[1] != 0) ? 1 : 0;
    P.absorb = T(c.absorb);
    P.inletLayer = T(c.inlet_layer_fraction * (W - 1));
  }

  T ld(const std::vector<E>& b, int64_t k) const { return S::template load<T>(b[size_t(k)]); }

  // src/shaders/aux.ts:15-18
  static T feq(int i, T rho, T ux, T uy) {
    T cu = T(3) * (T(CX2[i]) * ux + T(CY2[i]) * uy);
    return T(W2[i]) * (rho - T(1) + rho * (cu + T(0.5) * cu * cu - T(1.5) * (ux * ux + uy * uy)));
  }
  // src/shaders/common.ts:84-90
  T inlet_velocity(int y) const {
    if (P.parabolic == 0) return P.uIn;
    T h = T(H);
    T s = T(y) + T(0.5);
    return T(4) * P.uIn * s * (h - s) / (h * h);
  }
  // src/shaders/common.ts:93-102
  void wrap(int& qx, int& qy) const {
    if (qx < 0 && P.left == X_PERIODIC) qx += W;
    if (qx >= W && P.right == X_PERIODIC) qx -= W;
    if (qy < 0 && P.bottom == Y_PERIODIC) qy += H;
    if (qy >= H && P.top == Y_PERIODIC) qy -= H;
  }
  bool is_solid(int64_t s) const { return (flags[size_t(s)] & FLAG_SOLID2) != 0; }

  // src/shaders/step.ts:119-175
  T pull_slow(const std::vector<E>& src, int i, int x, int y, int64_t idx, uint32_t fl, T force[2]) const {
    const int j = OPP2[i];
    int sx = x - CX2[i];
    int sy = y - CY2[i];
    if ((fl & (1u << i)) == 0) return ld(src, i * N + int64_t(sy) * W + sx);
    wrap(sx, sy);
    const bool inX = sx >= 0 && sx < W;
    const bool inY = sy >= 0 && sy < H;
    if (inX && inY) {
      int64_t s = int64_t(sy) * W + sx;
      if (!is_solid(s)) return ld(src, i * N + s);
      T fj = ld(src, j * N + idx);
      T dF = T(sdf[size_t(idx)]);
      T dS = T(sdf[size_t(s)]);
      T q = wclamp<T>(dF / std::max<T>(dF - dS, T(1e-6)), T(1e-3), T(1));
      T fi = fj;
      if (q < T(0.5)) {
        int nx = x + CX2[i], ny = y + CY2[i];
        wrap(nx, ny);
        if (nx >= 0 && nx < W && ny >= 0 && ny < H) {
          int64_t n = int64_t(ny) * W + nx;
          if (!is_solid(n)) fi = T(2) * q * fj + (T(1) - T(2) * q) * ld(src, j * N + n);
        }
      } else {
        fi = fj / (T(2) * q) + (T(2) * q - T(1)) / (T(2) * q) * ld(src, i * N + idx);
      }
      force[0] += T(CX2[j]) * (fj + fi);
      force[1] += T(CY2[j]) * (fj + fi);
      return fi;
    }
    if (!inY) {
      int mode = sy < 0 ? P.bottom : P.top;
      if (mode == Y_SLIP) {
        int rx = std::min(std::max(sx, 0), W - 1);
        return ld(src, MIRROR_Y2[i] * N + int64_t(y) * W + rx);
      }
      if (mode == Y_MOVING) return ld(src, j * N + idx) + T(6) * T(W2[i]) * T(CX2[i]) * P.uLid;
      return ld(src, j * N + idx);
    }
    return ld(src, j * N + idx);
  }

  // src/shaders/step.ts:177-227 for one cell.
  void step_cell(const std::vector<E>& src, std::vector<E>& dst, int64_t idx) {
    const uint32_t fl = flags[size_t(idx)];
    const int y = int(idx / W);
    const int x = int(idx - int64_t(y) * W);
    T f[9];
    if (fl == 0) {
      f[0] = ld(src, idx);
      for (int i = 1; i < 9; i++) f[i] = ld(src, i * N + idx - int64_t(CY2[i]) * W - CX2[i]);
    } else {
      if (fl & FLAG_SOLID2) return;
      T force[2] = {T(0), T(0)};
      f[0] = ld(src, idx);
      for (int i = 1; i < 9; i++) f[i] = pull_slow(src, i, x, y, idx, fl, force);
      uint32_t bidx = fl >> BIDX_SHIFT2;
      if (bidx != 0) {
        cellForce[size_t(bidx - 1) * 2] = force[0];
        cellForce[size_t(bidx - 1) * 2 + 1] = force[1];
      }
      if (fl & FLAG_INLET2) {
        T u = inlet_velocity(y);
        T r = (T(1) + f[0] + f[2] + f[4] + T(2) * (f[3] + f[6] + f[7])) / (T(1) - u);
        f[1] = f[3] + T(2.0 / 3.0) * r * u;
        f[5] = f[7] - T(0.5) * (f[2] - f[4]) + T(1.0 / 6.0) * r * u;
        f[8] = f[6] + T(0.5) * (f[2] - f[4]) + T(1.0 / 6.0) * r * u;
      } else if (fl & FLAG_OUTLET2) {
        T u = f[0] + f[2] + f[4] + T(2) * (f[1] + f[5] + f[8]);
        f[3] = f[1] - T(2.0 / 3.0) * u;
        f[7] = f[5] + T(0.5) * (f[2] - f[4]) - T(1.0 / 6.0) * u;
        f[6] = f[8] - T(0.5) * (f[2] - f[4]) - T(1.0 / 6.0) * u;
      }
    }
    // moments (step.ts:22-32)
    T drho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    T rho = T(1) + drho;
    T invRho = T(1) / rho;
    T ux = (f[1] - f[3] + f[5] - f[6] - f[7] + f[8]) * invRho;
    T uy = (f[2] - f[4] + f[5] + f[6] - f[7] - f[8]) * invRho;
    if (P.forcing != 0) {
      ux += T(0.5) * P.fx * invRho;
      uy += T(0.5) * P.fy * invRho;
    }
    T usq = T(1.5) * (ux * ux + uy * uy);
    // equilibrium pair halves (step.ts:34-46)
    T cu[9];
    cu[1] = T(3) * (ux);
    cu[2] = T(3) * (uy);
    cu[5] = T(3) * (ux + uy);
    cu[6] = T(3) * (-ux + uy);
    T eqS[9], eqA[9];
    for (auto& pr : PAIRS2) {
      int a = pr[0];
      eqS[a] = T(W2[a]) * (drho + rho * (T(0.5) * cu[a] * cu[a] - usq));
      eqA[a] = T(W2[a]) * rho * cu[a];
    }
    // sponge and absorbing layers (step.ts:215-223)
    T omegaBase = T(1) / P.tau;
    T sOut = T(0);
    if (P.spongeStart < T(W - 1)) {
      sOut = smoothstep<T>(P.spongeStart, T(W - 1), T(x));
      omegaBase = T(1) / mix<T>(P.tau, P.tauSponge, sOut);
    }
    T sIn = T(0);
    if (P.inletLayer > T(0)) sIn = T(1) - smoothstep<T>(T(0), P.inletLayer, T(x));
    T sigma = P.absorb * std::max(sOut * sOut, sIn * sIn);
    // Smagorinsky and omega minus (step.ts:48-63)
    T omP = omegaBase;
    if (P.smagC2 > T(0)) {
      T n[9];
      for (auto& pr : PAIRS2) {
        int a = pr[0], b = pr[1];
        n[a] = f[a] - (eqS[a] + eqA[a]);
        n[b] = f[b] - (eqS[a] - eqA[a]);
      }
      T pxx = n[1] + n[3] + n[5] + n[6] + n[7] + n[8];
      T pyy = n[2] + n[4] + n[5] + n[6] + n[7] + n[8];
      T pxy = n[5] - n[6] + n[7] - n[8];
      T pi = std::sqrt(pxx * pxx + pyy * pyy + T(2) * pxy * pxy);
      T t0 = T(1) / omegaBase;
      T tEff = T(0.5) * (t0 + std::sqrt(t0 * t0 + T(18 * M_SQRT2) * P.smagC2 * pi * invRho));
      omP = T(1) / tEff;
    }
    T omM = T(1) / (T(0.5) + P.lambda / (T(1) / omP - T(0.5)));
    // sources (step.ts:65-91)
    T s0 = T(0);
    T sS[9] = {}, sA[9] = {};
    if (sigma > T(0)) {
      T ut = inlet_velocity(y);
      T dq = T(1.5) * rho * (ut * ut - ux * ux - uy * uy);
      s0 += sigma * T(W2[0]) * -dq;
      for (auto& pr : PAIRS2) {
        int a = pr[0];
        T cut = T(3) * (T(CX2[a]) * ut);
        sS[a] += sigma * T(W2[a]) * (T(0.5) * rho * (cut * cut - cu[a] * cu[a]) - dq);
        sA[a] += sigma * T(W2[a]) * rho * (cut - cu[a]);
      }
    }
    if (P.forcing != 0) {
      T uF = ux * P.fx + uy * P.fy;
      s0 += (T(1) - T(0.5) * omP) * T(W2[0]) * (T(-3) * uF);
      for (auto& pr : PAIRS2) {
        int a = pr[0];
        T cF = (T(CX2[a]) * P.fx + T(CY2[a]) * P.fy);
        sS[a] += (T(1) - T(0.5) * omP) * T(W2[a]) * (T(-3) * uF + T(9) * (cu[a] / T(3)) * cF);
        sA[a] += (T(1) - T(0.5) * omM) * T(W2[a]) * T(3) * cF;
      }
    }
    // collide (step.ts:93-103)
    dst[size_t(idx)] = S::template store<T>(f[0] - omP * (f[0] - T(W2[0]) * (drho - rho * usq)) + s0);
    for (auto& pr : PAIRS2) {
      int a = pr[0], b = pr[1];
      T sym = omP * (T(0.5) * (f[a] + f[b]) - eqS[a]);
      T asym = omM * (T(0.5) * (f[a] - f[b]) - eqA[a]);
      dst[size_t(a * N + idx)] = S::template store<T>(f[a] - sym - asym + sS[a] + sA[a]);
      dst[size_t(b * N + idx)] = S::template store<T>(f[b] - sym + asym + sS[a] - sA[a]);
    }
  }

  void one_step() {
    if (ramp.kind) P.uIn = T(ramp.at(step));
    const std::vector<E>& src = f[parity];
    std::vector<E>& dst = f[parity ^ 1];
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < N; idx++) step_cell(src, dst, idx);
    parity ^= 1;
    step++;
    if (cfg.force_every > 0 && step % cfg.force_every == 0) reduce();
  }

  // src/shaders/aux.ts:154-170, summed serially in double.
  void reduce() {
    double fx = 0, fy = 0;
    for (uint32_t k = 0; k < counter; k++) {
      fx += double(cellForce[size_t(k) * 2]);
      fy += double(cellForce[size_t(k) * 2 + 1]);
    }
    history[size_t(state % LBM_HISTORY_LEN) * 2] = fx;
    history[size_t(state % LBM_HISTORY_LEN) * 2 + 1] = fy;
    state++;
    book.steps.push_back(step);
  }

  // src/shaders/aux.ts:128-139
  void macro() {
    const std::vector<E>& b = f[parity];
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < N; idx++) {
      T* m = &mac[size_t(idx * 4)];
      if (flags[size_t(idx)] & FLAG_SOLID2) {
        m[0] = T(1), m[1] = T(0), m[2] = T(0), m[3] = T(1);
        continue;
      }
      T v[9];
      for (int i = 0; i < 9; i++) v[i] = ld(b, i * N + idx);
      T rho = T(1) + v[0] + v[1] + v[2] + v[3] + v[4] + v[5] + v[6] + v[7] + v[8];
      T jx = v[1] - v[3] + v[5] - v[6] - v[7] + v[8];
      T jy = v[2] - v[4] + v[5] + v[6] - v[7] - v[8];
      if (P.forcing != 0) {
        jx -= T(0.5) * P.fx;
        jy -= T(0.5) * P.fy;
      }
      m[0] = rho, m[1] = jx / rho, m[2] = jy / rho, m[3] = T(0);
    }
  }

  // src/shaders/aux.ts:34-37
  bool solid_at(int x, int y) const {
    bool open = (P.left == X_OPEN && x == 0) || (P.right == X_OPEN && x == W - 1);
    return !open && sdf[size_t(int64_t(y) * W + x)] < 0.0f;
  }

  // src/solver.ts:249-268: the disc's box grown by 2 cells, clipped to the grid (an empty box returns
  // without a rebuild); src/shaders/aux.ts:176-196 over it: d = |(x, y) - (cx, cy)| - r in float with
  // cx, cy, r rounded to float, union sdf = min(sdf, d), erase sdf = max(sdf, -d).
  void brush(double cx, double cy, double r, bool erase, bool rebuild) {
    const double pad = 2;
    const double x0 = std::max(0.0, std::floor(cx - r - pad)), y0 = std::max(0.0, std::floor(cy - r - pad));
    const double x1 = std::min(double(W - 1), std::ceil(cx + r + pad));
    const double y1 = std::min(double(H - 1), std::ceil(cy + r + pad));
    if (x1 < x0 || y1 < y0) return;
    const float fx = float(cx), fy = float(cy), fr = float(r);
    for (int y = int(y0); y <= int(y1); y++) {
      for (int x = int(x0); x <= int(x1); x++) {
        const float dx = float(x) - fx, dy = float(y) - fy;
        const float d = std::sqrt(dx * dx + dy * dy) - fr;
        float& v = sdf[size_t(int64_t(y) * W + x)];
        v = erase ? std::max(v, -d) : std::min(v, d);
      }
    }
    if (rebuild) rebuild_flags();
  }

  // src/solver.ts:205-217 + src/shaders/aux.ts:39-87; slots in cell order.
  void rebuild_flags() {
    macro();
    counter = 0;
    std::vector<uint32_t> old = flags;
    for (int64_t idx = 0; idx < N; idx++) {
      int y = int(idx / W), x = int(idx - int64_t(y) * W);
      if (solid_at(x, y)) {
        flags[size_t(idx)] = FLAG_SOLID2;
        continue;
      }
      uint32_t bits = 0;
      if (P.left == X_OPEN && x == 0) bits = FLAG_INLET2;
      if (P.right == X_OPEN && x == W - 1) bits = FLAG_OUTLET2;
      bool touches = false;
      for (int i = 1; i < 9; i++) {
        int sx = x - CX2[i], sy = y - CY2[i];
        wrap(sx, sy);
        bool inside = sx >= 0 && sy >= 0 && sx < W && sy < H;
        if (inside && solid_at(sx, sy)) {
          bits |= 1u << i;
          touches = true;
        } else if (!inside || sx != x - CX2[i] || sy != y - CY2[i]) {
          bits |= 1u << i;
        }
      }
      if (touches) {
        uint32_t slot = ++counter;
        if (slot < (1u << (32 - BIDX_SHIFT2))) bits |= slot << BIDX_SHIFT2;
      }
      flags[size_t(idx)] = bits;
      if (old[size_t(idx)] & FLAG_SOLID2) {
        T rho = T(0), ux = T(0), uy = T(0), n = T(0);
        for (int i = 1; i < 9; i++) {
          int sx = x + CX2[i], sy = y + CY2[i];
          if (sx < 0 || sy < 0 || sx >= W || sy >= H) continue;
          const T* m = &mac[size_t((int64_t(sy) * W + sx) * 4)];
          if (m[3] == T(0)) {
            rho += m[0];
            ux += m[1];
            uy += m[2];
            n += T(1);
          }
        }
        if (n > T(0)) {
          rho /= n;
          ux /= n;
          uy /= n;
        } else {
          rho = T(1);
        }
        for (int i = 0; i < 9; i++) {
          E v = S::template store<T>(feq(i, rho, ux, uy));
          f[0][size_t(i * N + idx)] = v;
          f[1][size_t(i * N + idx)] = v;
        }
      }
    }
    if (counter >= (1u << (32 - BIDX_SHIFT2))) counter = (1u << (32 - BIDX_SHIFT2)) - 1;
  }

  // src/solver.ts:220-234 + src/shaders/aux.ts:100-110
  void init_field(const float* m) {
    for (int64_t idx = 0; idx < N; idx++) {
      T rho = m ? T(m[idx * 4]) : T(1);
      T ux = m ? T(m[idx * 4 + 1]) : T(0);
      T uy = m ? T(m[idx * 4 + 2]) : T(0);
      for (int i = 0; i < 9; i++) {
        E v = S::template store<T>(feq(i, rho, ux, uy));
        f[0][size_t(i * N + idx)] = v;
        f[1][size_t(i * N + idx)] = v;
      }
    }
    parity = 0;
    step = 0;
    reset_forces();
  }
  void reset_forces() {
    state = 0;
    book.reset();
  }
};

// ============================================================================================
// D3Q19
// ============================================================================================
template <class T, class S, int Q = 19>
struct Solver3D {
  using L = Lattice3<Q>;
  using E = typename S::elem;
  int W, H, D;
  int64_t N, WH;
  lbm_config3d cfg;
  // Params3 uniform values (src/shaders/common3d.ts:49-56), derived by writeParams (solver3d.ts:197-208).
  struct Params {
    T tau, lambda, smagC2, uIn, spongeStart, tauSponge, absorb;
  } P;
  std::vector<E> f[2];
  std::vector<uint32_t> flags, slot;
  std::vector<float> sdf;
  std::vector<T> mac;        // N * 4
  std::vector<T> cellForce;  // N * 4
  uint32_t counter = 0;
  std::vector<double> history;  // HISTORY_LEN * 3
  uint32_t state = 0;
  int64_t step = 0;
  int parity = 0;
  ForceBook book;
  Ramp ramp;
  int OPP[Q], MY[Q], MZ[Q];
  // Open x planes: 0 = the reference's equilibrium inlet and outlet (src/shaders/step3d.ts:190-209),
  // 1 = Zou-He faces after Hecht and Harting (zou_he_face).
  int faces = 0;

  explicit Solver3D(const lbm_config3d& c)
      : W(c.width), H(c.height), D(c.depth), N(int64_t(c.width) * c.height * c.depth),
        WH(int64_t(c.width) * c.height), cfg(c) {
    for (int i = 0; i < Q; i++) OPP[i] = L::opp(i), MY[i] = L::mirror_y(i), MZ[i] = L::mirror_z(i);
    f[0].assign(size_t(Q * N), S::template store<T>(T(0)));
    f[1] = f[0];
    flags.assign(size_t(N), 0u);
    slot.assign(size_t(N), 0u);
    sdf.assign(size_t(N), 1e6f);
    mac.assign(size_t(N * 4), T(0));
    cellForce.assign(size_t(N * 4), T(0));
    history.assign(LBM_HISTORY_LEN * 3, 0.0);
    write_params();
    rebuild_flags();
  }

  void write_params() {
    const lbm_config3d& c = cfg;
    P.tau = T(c.tau);
    P.lambda = T(std::isnan(c.lambda) ? trt_lambda3(c.tau) : c.lambda);
    P.smagC2 = T(c.smagorinsky * c.smagorinsky);
    P.uIn = T(c.u_in);
    P.spongeStart = T(c.sponge_fraction > 0 ? (W - 1) * (1 - c.sponge_fraction) : 1e9);
    P.tauSponge = T(std::max(c.sponge_tau, c.tau));
    P.absorb = T(c.absorb);
  }

  T ld(const std::vector<E>& b, int64_t k) const { return S::template load<T>(b[size_t(k)]); }
  int64_t at(int x, int y, int z) const { return int64_t(x) + int64_t(W) * (int64_t(y) + int64_t(H) * z); }
  bool is_solid(int64_t s) const { return (flags[size_t(s)] & L::SOLID) != 0; }

  // src/shaders/common3d.ts:78-81
  static T feq(int i, T rho, T ux, T uy, T uz) {
    T cu = T(3) * (T(L::CX[i]) * ux + T(L::CY[i]) * uy + T(L::CZ[i]) * uz);
    return T(L::W[i]) * (rho - T(1) + rho * (cu + T(0.5) * cu * cu - T(1.5) * (ux * ux + uy * uy + uz * uz)));
  }

  // src/shaders/step3d.ts:119-168
  T pull_slow(const std::vector<E>& src, int i, int x, int y, int z, int64_t idx, uint32_t fl, bool outlet,
              T force[3]) const {
    const int j = OPP[i];
    const int sx = x - L::CX[i], sy = y - L::CY[i], sz = z - L::CZ[i];
    if ((fl & (1u << i)) == 0) return ld(src, i * N + at(sx, sy, sz));
    const bool outX = sx < 0 || sx >= W;
    const bool outY = sy < 0 || sy >= H;
    const bool outZ = sz < 0 || sz >= D;
    if (outX && outlet) return ld(src, i * N + at(W - 2, y, z));
    if (outY && outZ) {
      // Reflected through both slip faces: c_y and c_z negated, read at the node (x - c_x, y, z). In
      // D3Q19 every such direction has c_x = 0, so this is the node itself and the opposite
      // direction, the reference's bounce-back (step3d.ts:132); D3Q27's corner directions need the
      // double reflection for a slip edge to keep a uniform flow uniform. A solid reflected node
      // bounces back with force, as on one face.
      const int64_t rs = at(sx, y, z);
      if (is_solid(rs)) {
        T fj = ld(src, j * N + idx);
        force[0] += T(L::CX[j]) * (T(2) * fj);
        force[1] += T(L::CY[j]) * (T(2) * fj);
        force[2] += T(L::CZ[j]) * (T(2) * fj);
        return fj;
      }
      return ld(src, MY[MZ[i]] * N + rs);
    }
    if (outY || outZ) {
      int rx = sx, ry = sy, rz = sz;
      int m = MY[i];
      if (outY) {
        ry = y;
      } else {
        rz = z;
        m = MZ[i];
      }
      int64_t rs = at(rx, ry, rz);
      if (is_solid(rs)) {
        T fj = ld(src, j * N + idx);
        force[0] += T(L::CX[j]) * (T(2) * fj);
        force[1] += T(L::CY[j]) * (T(2) * fj);
        force[2] += T(L::CZ[j]) * (T(2) * fj);
        return fj;
      }
      return ld(src, m * N + rs);
    }
    if (outX) return ld(src, j * N + idx);
    int64_t s = at(sx, sy, sz);
    if (!is_solid(s)) return ld(src, i * N + s);
    T fj = ld(src, j * N + idx);
    T dF = T(sdf[size_t(idx)]);
    T dS = T(sdf[size_t(s)]);
    T q = wclamp<T>(dF / std::max<T>(dF - dS, T(1e-6)), T(1e-3), T(1));
    T fi = fj;
    if (q < T(0.5)) {
      int nx = x + L::CX[i], ny = y + L::CY[i], nz = z + L::CZ[i];
      if (nx >= 0 && nx < W && ny >= 0 && ny < H && nz >= 0 && nz < D) {
        int64_t n = at(nx, ny, nz);
        if (!is_solid(n)) fi = T(2) * q * fj + (T(1) - T(2) * q) * ld(src, j * N + n);
      }
    } else {
      fi = fj / (T(2) * q) + (T(2) * q - T(1)) / (T(2) * q) * ld(src, i * N + idx);
    }
    force[0] += T(L::CX[j]) * (fj + fi);
    force[1] += T(L::CY[j]) * (fj + fi);
    force[2] += T(L::CZ[j]) * (fj + fi);
    return fi;
  }

  // signed index-order sum of f_i weighted by c (step3d.ts:9-12): first term added to nothing.
  static T csum(const int* c, const T* f) {
    bool first = true;
    T acc = T(0);
    for (int i = 0; i < Q; i++) {
      if (c[i] == 0) continue;
      if (first) {
        acc = c[i] > 0 ? f[i] : -f[i];
        first = false;
      } else {
        acc = c[i] > 0 ? acc + f[i] : acc - f[i];
      }
    }
    return acc;
  }

  // A value the time step stores (population i of node idx): the stochastic codec (LBM_FP16_SR) rounds
  // it with the random word of (idx, step, i); the others round to nearest.
  E put(T v, int64_t idx, int i) const {
    if constexpr (S::stochastic)
      return S::template store_r<T>(v, step_random(uint32_t(idx), uint32_t(step), i));
    else
      return S::template store<T>(v);
  }

  // Zou-He velocity and pressure faces of Hecht and Harting (J. Stat. Mech. (2010) P01018,
  // arXiv:0811.4593), on the x = 0 plane (u = (uIn, 0, 0)) and the x = W-1 plane (rho = 1, no
  // transverse velocity), in the f - w variables (the weights of a plane and twice those entering
  // it sum to 1). The populations entering from outside the grid (c_x = +1 at the inlet, -1 at the
  // outlet) are rebuilt from their opposites:
  //   f_i = f_opp(i) + 6 w_i rho (c_i . u) - c_iy N_y - c_iz N_z,  N_a = 1/2 sum_{c_x = 0} c_a f,
  // with rho (1 - uIn) = 1 + sum_{c_x = 0} f + 2 sum_{c_x = -1} f at the inlet and
  // u_x = sum_{c_x = 0} f + 2 sum_{c_x = +1} f at the outlet, so that the node's moments are exactly
  // the prescribed ones. Slip-face rows use the mirrored populations as known ones, so edges and
  // corners of the planes take the same rule.
  void zou_he_face(T f[Q], bool inlet) const {
    const int in = inlet ? 1 : -1;
    T s0 = T(0), sk = T(0), ny = T(0), nz = T(0);
    for (int i = 0; i < Q; i++) {
      if (L::CX[i] == 0) {
        s0 = s0 + f[i];
        if (L::CY[i]) ny = L::CY[i] > 0 ? ny + f[i] : ny - f[i];
        if (L::CZ[i]) nz = L::CZ[i] > 0 ? nz + f[i] : nz - f[i];
      } else if (L::CX[i] == -in) {
        sk = sk + f[i];
      }
    }
    ny = T(0.5) * ny;
    nz = T(0.5) * nz;
    T rho = T(1), ux;
    if (inlet) {
      ux = P.uIn;
      rho = (T(1) + s0 + T(2) * sk) / (T(1) - ux);
    } else {
      ux = s0 + T(2) * sk;
    }
    const T ru = rho * ux;
    for (int i = 1; i < Q; i++) {
      if (L::CX[i] != in) continue;
      const T k = (L::CY[i] || L::CZ[i]) ? T(1) / T(6) : T(1) / T(3);  // 6 w_i
      T v = in > 0 ? f[OPP[i]] + k * ru : f[OPP[i]] - k * ru;
      if (L::CY[i]) v = L::CY[i] > 0 ? v - ny : v + ny;
      if (L::CZ[i]) v = L::CZ[i] > 0 ? v - nz : v + nz;
      f[i] = v;
    }
  }

  // src/shaders/step3d.ts:170-223 for one cell.
  void step_cell(const std::vector<E>& src, std::vector<E>& dst, int64_t idx) {
    const uint32_t fl = flags[size_t(idx)];
    const int z = int(idx / WH);
    const int64_t rem = idx - int64_t(z) * WH;
    const int y = int(rem / W);
    const int x = int(rem - int64_t(y) * W);
    T f[Q];
    bool outlet = false;
    if (fl == 0) {
      f[0] = ld(src, idx);
      for (int i = 1; i < Q; i++) f[i] = ld(src, i * N + idx - L::CZ[i] * WH - int64_t(L::CY[i]) * W - L::CX[i]);
    } else {
      if (fl & L::SOLID) return;
      const bool inlet = (fl & L::INLET) != 0;
      if (inlet && !faces) {
        for (int i = 0; i < Q; i++) dst[size_t(i * N + idx)] = put(feq(i, T(1), P.uIn, T(0), T(0)), idx, i);
        return;
      }
      outlet = (fl & L::OUTLET) != 0;
      T force[3] = {T(0), T(0), T(0)};
      f[0] = ld(src, idx);
      // (a Zou-He inlet rebuilds the populations entering from x = -1 instead of pulling them)
      for (int i = 1; i < Q; i++) f[i] = inlet && L::CX[i] > 0 ? T(0) : pull_slow(src, i, x, y, z, idx, fl, outlet, force);
      uint32_t sl = slot[size_t(idx)];
      if (sl != 0) {
        T* cf = &cellForce[size_t(sl - 1) * 4];
        cf[0] = force[0], cf[1] = force[1], cf[2] = force[2], cf[3] = T(0);
      }
      if (faces && (inlet || outlet)) {
        zou_he_face(f, inlet);
        outlet = false;  // the face node collides
      }
    }
    // moments (step3d.ts:38-45)
    T drho = f[0];
    for (int i = 1; i < Q; i++) drho = drho + f[i];
    T rho = T(1) + drho;
    T invRho = T(1) / rho;
    T ux = csum(L::CX, f) * invRho;
    T uy = csum(L::CY, f) * invRho;
    T uz = csum(L::CZ, f) * invRho;
    T usq = T(1.5) * (ux * ux + uy * uy + uz * uz);
    if (outlet) {
      for (int i = 0; i < Q; i++) dst[size_t(i * N + idx)] = put(feq(i, T(1), ux, uy, uz), idx, i);
      return;
    }
    // c_i . u in x, y, z order (step3d.ts:14-17, 47)
    T cu[Q], eqS[Q], eqA[Q];
    for (int a = 1; a < Q; a += 2) {
      bool first = true;
      T acc = T(0);
      const T u[3] = {ux, uy, uz};
      const int c[3] = {L::CX[a], L::CY[a], L::CZ[a]};
      for (int k = 0; k < 3; k++) {
        if (c[k] == 0) continue;
        if (first) {
          acc = c[k] > 0 ? u[k] : -u[k];
          first = false;
        } else {
          acc = c[k] > 0 ? acc + u[k] : acc - u[k];
        }
      }
      cu[a] = T(3) * acc;
      eqS[a] = T(L::W[a]) * (drho + rho * (T(0.5) * cu[a] * cu[a] - usq));
      eqA[a] = T(L::W[a]) * rho * cu[a];
    }
    // sponge (step3d.ts:213-219)
    T omegaBase = T(1) / P.tau;
    T sOut = T(0);
    if (P.spongeStart < T(W - 1)) {
      sOut = smoothstep<T>(P.spongeStart, T(W - 1), T(x));
      omegaBase = T(1) / mix<T>(P.tau, P.tauSponge, sOut);
    }
    T sigma = P.absorb * sOut * sOut;
    // Smagorinsky (step3d.ts:59-74)
    T omP = omegaBase;
    if (P.smagC2 > T(0)) {
      T n[Q];
      n[0] = T(0);
      for (int a = 1; a < Q; a += 2) {
        n[a] = f[a] - (eqS[a] + eqA[a]);
        n[a + 1] = f[a + 1] - (eqS[a] - eqA[a]);
      }
      int cxx[Q], cyy[Q], czz[Q], cxy[Q], cxz[Q], cyz[Q];
      for (int i = 0; i < Q; i++) {
        cxx[i] = L::CX[i] * L::CX[i], cyy[i] = L::CY[i] * L::CY[i], czz[i] = L::CZ[i] * L::CZ[i];
        cxy[i] = L::CX[i] * L::CY[i], cxz[i] = L::CX[i] * L::CZ[i], cyz[i] = L::CY[i] * L::CZ[i];
      }
      T pxx = csum(cxx, n), pyy = csum(cyy, n), pzz = csum(czz, n);
      T pxy = csum(cxy, n), pxz = csum(cxz, n), pyz = csum(cyz, n);
      T pi = std::sqrt(pxx * pxx + pyy * pyy + pzz * pzz + T(2) * (pxy * pxy + pxz * pxz + pyz * pyz));
      T t0 = T(1) / omegaBase;
      T tEff = T(0.5) * (t0 + std::sqrt(t0 * t0 + T(18 * M_SQRT2) * P.smagC2 * pi * invRho));
      omP = T(1) / tEff;
    }
    T omM = T(1) / (T(0.5) + P.lambda / (T(1) / omP - T(0.5)));
    // absorbing layer (step3d.ts:76-90)
    T s0 = T(0);
    T sS[Q] = {}, sA[Q] = {};
    if (sigma > T(0)) {
      T ut = P.uIn;
      T dq = T(1.5) * rho * (ut * ut - ux * ux - uy * uy - uz * uz);
      s0 += sigma * T(L::W[0]) * -dq;
      for (int a = 1; a < Q; a += 2) {
        T cut = T(3) * T(L::CX[a]) * ut;
        sS[a] += sigma * T(L::W[a]) * (T(0.5) * rho * (cut * cut - cu[a] * cu[a]) - dq);
        sA[a] += sigma * T(L::W[a]) * rho * (cut - cu[a]);
      }
    }
    // collide (step3d.ts:92-102)
    dst[size_t(idx)] = put(f[0] - omP * (f[0] - T(L::W[0]) * (drho - rho * usq)) + s0, idx, 0);
    for (int a = 1; a < Q; a += 2) {
      int b = a + 1;
      T sym = omP * (T(0.5) * (f[a] + f[b]) - eqS[a]);
      T asym = omM * (T(0.5) * (f[a] - f[b]) - eqA[a]);
      dst[size_t(a * N + idx)] = put(f[a] - sym - asym + sS[a] + sA[a], idx, a);
      dst[size_t(b * N + idx)] = put(f[b] - sym + asym + sS[a] - sA[a], idx, b);
    }
  }

  void one_step() {
    if (ramp.kind) P.uIn = T(ramp.at(step));
    const std::vector<E>& src = f[parity];
    std::vector<E>& dst = f[parity ^ 1];
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < N; idx++) step_cell(src, dst, idx);
    parity ^= 1;
    step++;
    if (cfg.force_every > 0 && step % cfg.force_every == 0) reduce();
  }

  void reduce() {
    double s[3] = {0, 0, 0};
    for (uint32_t k = 0; k < counter; k++)
      for (int d = 0; d < 3; d++) s[d] += double(cellForce[size_t(k) * 4 + d]);
    for (int d = 0; d < 3; d++) history[size_t(state % LBM_HISTORY_LEN) * 3 + d] = s[d];
    state++;
    book.steps.push_back(step);
  }

  // src/shaders/aux3d.ts:137-150
  void macro() {
    const std::vector<E>& b = f[parity];
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < N; idx++) {
      T* m = &mac[size_t(idx * 4)];
      if (flags[size_t(idx)] & L::SOLID) {
        m[0] = m[1] = m[2] = m[3] = T(0);
        continue;
      }
      T drho = T(0), jx = T(0), jy = T(0), jz = T(0);
      for (int i = 0; i < Q; i++) {
        T v = ld(b, i * N + idx);
        drho += v;
        jx += T(L::CX[i]) * v;
        jy += T(L::CY[i]) * v;
        jz += T(L::CZ[i]) * v;
      }
      T rho = T(1) + drho;
      m[0] = jx / rho, m[1] = jy / rho, m[2] = jz / rho, m[3] = rho;
    }
  }

  // src/shaders/aux3d.ts:37-40
  bool solid_at(int x, int y, int z) const {
    if (x == 0 || x == W - 1) return false;
    return sdf[size_t(at(x, y, z))] < 0.0f;
  }
  bool in_range(int x, int y, int z) const { return x >= 0 && y >= 0 && z >= 0 && x < W && y < H && z < D; }

  // src/solver3d.ts:216-228 + src/shaders/aux3d.ts:42-100; slots in cell order.
  void rebuild_flags() {
    macro();
    counter = 0;
    std::vector<uint32_t> old = flags;
    for (int64_t idx = 0; idx < N; idx++) {
      int z = int(idx / WH);
      int y = int((idx - int64_t(z) * WH) / W);
      int x = int(idx - int64_t(z) * WH - int64_t(y) * W);
      if (solid_at(x, y, z)) {
        flags[size_t(idx)] = L::SOLID;
        slot[size_t(idx)] = 0;
        continue;
      }
      uint32_t bits = 0;
      if (x == 0) bits = L::INLET;
      if (x == W - 1) bits = L::OUTLET;
      bool touches = false;
      for (int i = 1; i < Q; i++) {
        int sx = x - L::CX[i], sy = y - L::CY[i], sz = z - L::CZ[i];
        if (in_range(sx, sy, sz)) {
          if (solid_at(sx, sy, sz)) {
            bits |= 1u << i;
            touches = true;
          }
          continue;
        }
        bits |= 1u << i;
        bool outX = sx < 0 || sx >= W, outY = sy < 0 || sy >= H, outZ = sz < 0 || sz >= D;
        if (!outX && (outY || outZ)) {
          int rx = sx, ry = outY ? y : sy, rz = outZ ? z : sz;
          if (solid_at(rx, ry, rz)) touches = true;
        }
      }
      flags[size_t(idx)] = bits;
      slot[size_t(idx)] = touches ? ++counter : 0u;
      if (old[size_t(idx)] & L::SOLID) {
        T rho = T(0), ux = T(0), uy = T(0), uz = T(0), n = T(0);
        for (int i = 1; i < Q; i++) {
          int qx = x + L::CX[i], qy = y + L::CY[i], qz = z + L::CZ[i];
          if (!in_range(qx, qy, qz)) continue;
          const T* m = &mac[size_t(at(qx, qy, qz) * 4)];
          if (m[3] != T(0)) {
            rho += m[3];
            ux += m[0];
            uy += m[1];
            uz += m[2];
            n += T(1);
          }
        }
        if (n > T(0)) {
          rho /= n;
          ux /= n;
          uy /= n;
          uz /= n;
        } else {
          rho = T(1);
          ux = P.uIn;
          uy = T(0);
          uz = T(0);
        }
        for (int i = 0; i < Q; i++) {
          E v = S::template store<T>(feq(i, rho, ux, uy, uz));
          f[0][size_t(i * N + idx)] = v;
          f[1][size_t(i * N + idx)] = v;
        }
      }
    }
  }

  // src/solver3d.ts:231-242 + src/shaders/aux3d.ts:113-123
  void init_field() {
    for (int64_t idx = 0; idx < N; idx++) {
      bool solid = (flags[size_t(idx)] & L::SOLID) != 0;
      for (int i = 0; i < Q; i++) {
        E v = S::template store<T>(solid ? T(0) : feq(i, T(1), P.uIn, T(0), T(0)));
        f[0][size_t(i * N + idx)] = v;
        f[1][size_t(i * N + idx)] = v;
      }
    }
    parity = 0;
    step = 0;
    reset_forces();
  }
  void reset_forces() {
    state = 0;
    book.reset();
  }
};

}  // namespace orc
