// Lattice tables and WGSL built-ins restated for the oracle (test infrastructure only).
//
// D2Q9 tables: src/lattice.ts:2-36. D3Q19 tables: src/lattice3d.ts:4-31.
// WGSL built-ins follow the WGSL specification: smoothstep, mix, clamp.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>

namespace orc {

// ---- D2Q9 (src/lattice.ts:2-15, 30-36) ----
constexpr int Q2 = 9;
constexpr int CX2[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY2[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr double W2[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
constexpr int OPP2[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
constexpr int MIRROR_Y2[9] = {0, 1, 4, 3, 2, 8, 7, 6, 5};
constexpr int PAIRS2[4][2] = {{1, 3}, {2, 4}, {5, 7}, {6, 8}};
constexpr uint32_t FLAG_SOLID2 = 1u << 9;
constexpr uint32_t FLAG_INLET2 = 1u << 10;
constexpr uint32_t FLAG_OUTLET2 = 1u << 11;
constexpr uint32_t BIDX_SHIFT2 = 12;
constexpr double MAGIC_LAMBDA = 3.0 / 16.0;

enum XMode { X_PERIODIC = 0, X_WALL = 1, X_OPEN = 2 };
enum YMode { Y_PERIODIC = 0, Y_NOSLIP = 1, Y_SLIP = 2, Y_MOVING = 3 };

// ---- D3Q19 (src/lattice3d.ts:4-26) ----
constexpr int Q3 = 19;
constexpr int CX3[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
constexpr int CY3[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
constexpr int CZ3[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
constexpr double W3[19] = {1.0 / 3,  1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                           1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                           1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
constexpr uint32_t FLAG_SOLID3 = 1u << 19;
constexpr uint32_t FLAG_INLET3 = 1u << 20;
constexpr uint32_t FLAG_OUTLET3 = 1u << 21;

inline int opp3(int i) { return i == 0 ? 0 : (i % 2 ? i + 1 : i - 1); }
inline int find3(int cx, int cy, int cz) {
  for (int i = 0; i < Q3; i++)
    if (CX3[i] == cx && CY3[i] == cy && CZ3[i] == cz) return i;
  return -1;
}
// src/lattice3d.ts:17-19: direction with one component negated.
inline int mirror_y3(int i) { return find3(CX3[i], -CY3[i], CZ3[i]); }
inline int mirror_z3(int i) { return find3(CX3[i], CY3[i], -CZ3[i]); }

// ---- D3Q27 (the reference's roadmap switch: docs/research/3d-lattice-boltzmann-webgpu.md:74) ----
// The D3Q19 set in the same order, then the eight corners, so that odd i and i + 1 stay opposites
// and the TRT pairs are (1, 2), (3, 4), ..., (25, 26). Weights 8/27, 2/27, 1/54, 1/216 (the
// D3Q27 standard set). Flag words use bits 1..26 for directions, then SOLID, INLET, OUTLET.
constexpr int Q27 = 27;
constexpr int CX27[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
constexpr int CY27[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
constexpr int CZ27[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
constexpr double W27[27] = {8.0 / 27,  2.0 / 27,  2.0 / 27,  2.0 / 27,  2.0 / 27,  2.0 / 27,  2.0 / 27,
                            1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,
                            1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 216, 1.0 / 216,
                            1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216};

// The tables of a 3D velocity set by its size (19 or 27); flag bits SOLID, INLET, OUTLET follow the
// Q - 1 direction bits.
template <int Q>
struct Lattice3 {
  static constexpr const int* CX = Q == 19 ? CX3 : CX27;
  static constexpr const int* CY = Q == 19 ? CY3 : CY27;
  static constexpr const int* CZ = Q == 19 ? CZ3 : CZ27;
  static constexpr const double* W = Q == 19 ? W3 : W27;
  static constexpr uint32_t SOLID = 1u << Q, INLET = 1u << (Q + 1), OUTLET = 1u << (Q + 2);
  static int find(int cx, int cy, int cz) {
    for (int i = 0; i < Q; i++)
      if (CX[i] == cx && CY[i] == cy && CZ[i] == cz) return i;
    return -1;
  }
  static int opp(int i) { return i == 0 ? 0 : (i % 2 ? i + 1 : i - 1); }
  static int mirror_y(int i) { return find(CX[i], -CY[i], CZ[i]); }
  static int mirror_z(int i) { return find(CX[i], CY[i], -CZ[i]); }
};

// src/lattice3d.ts:29-31.
inline double trt_lambda3(double tau) { return std::min(MAGIC_LAMBDA, 50.0 * (tau - 0.5) * (tau - 0.5)); }

// ---- WGSL built-ins ----
template <class T> inline T wclamp(T e, T lo, T hi) { return std::min(std::max(e, lo), hi); }
template <class T> inline T smoothstep(T lo, T hi, T x) {
  T t = wclamp<T>((x - lo) / (hi - lo), T(0), T(1));
  return t * t * (T(3) - T(2) * t);
}
template <class T> inline T mix(T a, T b, T t) { return a * (T(1) - t) + b * t; }

}  // namespace orc
