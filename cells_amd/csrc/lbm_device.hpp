// Device-side lattice tables, storage codecs and kernel parameter blocks (gfx950).
//
// Tables: D2Q9 src/lattice.ts:2-36, D3Q19 src/lattice3d.ts:4-26. Codecs: FP32 and FP16 x 2^15
// follow src/shaders/common3d.ts:92-103; BF16 (round to nearest even of the float32 bit pattern,
// no scale, no clamp) has no counterpart in the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

// ---------------------------------------------------------------- D2Q9
constexpr int CX2[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY2[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP2[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
constexpr int MIRROR_Y2[9] = {0, 1, 4, 3, 2, 8, 7, 6, 5};
constexpr float W2[9] = {4.0f / 9, 1.0f / 9, 1.0f / 9, 1.0f / 9, 1.0f / 9, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36};
constexpr uint32_t FLAG_SOLID2 = 1u << 9;
constexpr uint32_t FLAG_INLET2 = 1u << 10;
constexpr uint32_t FLAG_OUTLET2 = 1u << 11;
constexpr uint32_t BIDX_SHIFT2 = 12;
// Direction bits a plain inlet / outlet column node carries: the upstream node is off the grid
// exactly for the directions with c_x = +1 (inlet) or c_x = -1 (outlet).
constexpr uint32_t INLET_DIRS2 = (1u << 1) | (1u << 5) | (1u << 8);
constexpr uint32_t OUTLET_DIRS2 = (1u << 3) | (1u << 6) | (1u << 7);

// ---------------------------------------------------------------- D3Q19
constexpr int CX3[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
constexpr int CY3[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
constexpr int CZ3[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
constexpr float W3[19] = {1.0f / 3,  1.0f / 18, 1.0f / 18, 1.0f / 18, 1.0f / 18, 1.0f / 18, 1.0f / 18,
                          1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36,
                          1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36};
constexpr uint32_t FLAG_SOLID3 = 1u << 19;
constexpr uint32_t FLAG_INLET3 = 1u << 20;
constexpr uint32_t FLAG_OUTLET3 = 1u << 21;
// Set by the host on nodes whose pulls the bulk rows' patterns do not cover; never reported.
constexpr uint32_t FLAG_LIST3 = 1u << 31;
constexpr uint32_t INLET_DIRS3 = (1u << 1) | (1u << 7) | (1u << 9) | (1u << 13) | (1u << 15);
constexpr uint32_t OUTLET_DIRS3 = (1u << 2) | (1u << 8) | (1u << 10) | (1u << 14) | (1u << 16);

__host__ __device__ constexpr int opp3(int i) { return i == 0 ? 0 : (i % 2 ? i + 1 : i - 1); }
__host__ __device__ constexpr int find3(int cx, int cy, int cz) {
  int r = -1;
  for (int i = 0; i < 19; i++)
    if (CX3[i] == cx && CY3[i] == cy && CZ3[i] == cz) r = i;
  return r;
}
__host__ __device__ constexpr int mirror_y3(int i) { return find3(CX3[i], -CY3[i], CZ3[i]); }
__host__ __device__ constexpr int mirror_z3(int i) { return find3(CX3[i], CY3[i], -CZ3[i]); }

// ---------------------------------------------------------------- D3Q27
// The velocity-set switch of the reference's 3D roadmap (docs/research/3d-lattice-boltzmann-webgpu.md:74):
// the D3Q19 directions in the same order, then the eight corners, so odd i and i + 1 stay opposites.
// Weights 8/27, 2/27, 1/54, 1/216. Flag words: direction bits 1..26, then SOLID, INLET, OUTLET.
constexpr int CX27[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
constexpr int CY27[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
constexpr int CZ27[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
constexpr float W27[27] = {8.0f / 27,  2.0f / 27,  2.0f / 27,  2.0f / 27,  2.0f / 27,  2.0f / 27,  2.0f / 27,
                           1.0f / 54,  1.0f / 54,  1.0f / 54,  1.0f / 54,  1.0f / 54,  1.0f / 54,  1.0f / 54,
                           1.0f / 54,  1.0f / 54,  1.0f / 54,  1.0f / 54,  1.0f / 54,  1.0f / 216, 1.0f / 216,
                           1.0f / 216, 1.0f / 216, 1.0f / 216, 1.0f / 216, 1.0f / 216, 1.0f / 216};
constexpr uint32_t FLAG_SOLID27 = 1u << 27;
constexpr uint32_t FLAG_INLET27 = 1u << 28;
constexpr uint32_t FLAG_OUTLET27 = 1u << 29;
// Set by the flag kernel on nodes whose pulls meet a solid node (a Bouzidi link, or a slip face or
// edge reflecting into a solid): the step resolves those per direction. Never reported.
constexpr uint32_t FLAG_SLOW27 = 1u << 30;
__host__ __device__ constexpr int find27(int cx, int cy, int cz) {
  int r = -1;
  for (int i = 0; i < 27; i++)
    if (CX27[i] == cx && CY27[i] == cy && CZ27[i] == cz) r = i;
  return r;
}
__host__ __device__ constexpr int mirror_y27(int i) { return find27(CX27[i], -CY27[i], CZ27[i]); }
__host__ __device__ constexpr int mirror_z27(int i) { return find27(CX27[i], CY27[i], -CZ27[i]); }

// ---------------------------------------------------------------- WGSL built-ins
__device__ __forceinline__ float wclamp(float e, float lo, float hi) { return fminf(fmaxf(e, lo), hi); }
__device__ __forceinline__ float smoothstep(float lo, float hi, float x) {
  float t = wclamp((x - lo) / (hi - lo), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
__device__ __forceinline__ float wmix(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// d f_i / d q of a Bouzidi link (src/shaders/step.ts:145-156) at its operands f_j (at the node) and
// b (f_j at x + c_i for q < 1/2, f_i at x otherwise): 2 (f_j - b), or 0 when the q < 1/2 operand is
// missing; (b - f_j) / (2 q^2) for q >= 1/2.
__device__ __forceinline__ float bouzidi_dq(float q, float fj, float b, bool has_operand) {
  if (q < 0.5f) return has_operand ? 2.0f * (fj - b) : 0.0f;
  return (b - fj) / (2.0f * q * q);
}

// ---------------------------------------------------------------- storage codecs
// decode / encode one stored element; load(p, k) and store(p, k, v) address element k of p, while
// load_at(base, off) and store_at(base, off, v) take a 32-bit byte offset from a wave-uniform base,
// which lets the compiler use the SGPR-base + VGPR-offset addressing of global loads and stores.
// A stochastic codec (FP16 with stochastic rounding) takes a 32-bit random word per stored value in
// the time-step kernels (store_r / store_at_r, see step_random); every other store rounds to nearest.
template <class Codec>
struct StoreOps {
  using elem = typename Codec::elem;
  static constexpr bool stochastic = Codec::stochastic;
  __device__ static __forceinline__ float load(const elem* p, uint32_t k) { return Codec::dec(p[k]); }
  __device__ static __forceinline__ void store(elem* p, uint32_t k, float v) { p[k] = Codec::enc(v); }
  __device__ static __forceinline__ void store_r(elem* p, uint32_t k, float v, uint32_t r) { p[k] = Codec::enc_r(v, r); }
  __device__ static __forceinline__ float load_at(const char* base, uint32_t off) {
    return Codec::dec(*reinterpret_cast<const elem*>(base + off));
  }
  __device__ static __forceinline__ void store_at(char* base, uint32_t off, float v) {
    *reinterpret_cast<elem*>(base + off) = Codec::enc(v);
  }
  __device__ static __forceinline__ void store_at_r(char* base, uint32_t off, float v, uint32_t r) {
    *reinterpret_cast<elem*>(base + off) = Codec::enc_r(v, r);
  }
};
struct CodecF32 {
  using elem = float;
  static constexpr bool stochastic = false;
  __device__ static __forceinline__ float dec(elem e) { return e; }
  __device__ static __forceinline__ elem enc(float v) { return v; }
  __device__ static __forceinline__ elem enc_r(float v, uint32_t) { return v; }
};
struct CodecF16 {
  using elem = _Float16;
  static constexpr bool stochastic = false;
  __device__ static __forceinline__ float dec(elem e) { return float(e) * (1.0f / 32768.0f); }
  // v_cvt_f16_f32 rounds to nearest even
  __device__ static __forceinline__ elem enc(float v) { return _Float16(wclamp(v, -1.99f, 1.99f) * 32768.0f); }
  __device__ static __forceinline__ elem enc_r(float v, uint32_t) { return enc(v); }
};
// The FP16 x 2^15 format of CodecF16 with stochastic rounding in the time step: a scaled value in the
// f16 normal range (|x| >= 2^-14) becomes the f16 value below |x| after adding r >> 19 to the 13
// mantissa bits f16 drops (v_cvt_sr_f16_f32; the rounding is up with probability equal to the dropped
// fraction, so it carries no bias); smaller values round to nearest even.
struct CodecF16SR {
  using elem = _Float16;
  static constexpr bool stochastic = true;
  typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
  __device__ static __forceinline__ float dec(elem e) { return CodecF16::dec(e); }
  __device__ static __forceinline__ elem enc(float v) { return CodecF16::enc(v); }
  __device__ static __forceinline__ elem enc_r(float v, uint32_t r) {
    const float x = wclamp(v, -1.99f, 1.99f) * 32768.0f;
    const half2_t h = __builtin_amdgcn_cvt_sr_f16_f32(half2_t{0, 0}, x, r, false);
    return fabsf(x) < 6.103515625e-05f ? _Float16(x) : h[0];
  }
};
struct CodecBF16 {
  using elem = uint16_t;
  static constexpr bool stochastic = false;
  __device__ static __forceinline__ float dec(elem e) { return __uint_as_float(uint32_t(e) << 16); }
  // v_cvt_pk_bf16_f32 (gfx950) rounds to nearest even
  __device__ static __forceinline__ elem enc(float v) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(v)); }
  __device__ static __forceinline__ elem enc_r(float v, uint32_t) { return enc(v); }
};
struct StoreF32 : StoreOps<CodecF32> {};
struct StoreF16 : StoreOps<CodecF16> {};
struct StoreF16SR : StoreOps<CodecF16SR> {};
struct StoreBF16 : StoreOps<CodecBF16> {};

// Random words of the stochastic codec: lowbias32 (a 32-bit integer hash: xor-shift, multiply,
// xor-shift, multiply, xor-shift) of the cell index xor the step's seed gives one word per node and
// step; population i uses that word plus i times 0x9E3779B9. The seed of step n is lowbias32(n).
__host__ __device__ __forceinline__ uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ uint32_t step_random(uint32_t node_word, int i) {
  return node_word + uint32_t(i) * 0x9E3779B9u;
}

// ---------------------------------------------------------------- parameter blocks
// Values of the reference's Params uniform (src/shaders/common.ts:55-63) after writeParams
// (src/solver.ts:180-197), passed by value as kernel arguments.
struct Params2 {
  int32_t W, H;
  uint32_t N;
  int32_t left, right, bottom, top, parabolic, forcing;
  float tau, lambda, smagC2, uIn, spongeStart, tauSponge, uLid, fx, fy, absorb, inletLayer;
  uint32_t NS, NA;  // population layout, as Params3::NS / NA (NA = 8 for tiles of 64 cells)
};
// Params3 (src/shaders/common3d.ts:49-56) after Solver3D.writeParams (src/solver3d.ts:197-208).
struct Params3 {
  int32_t W, H, D;
  uint32_t N, WH;
  float tau, lambda, smagC2, uIn, spongeStart, tauSponge, absorb;
  uint32_t seed;   // lowbias32(step index): the stochastic codec's per-step seed
  uint32_t faces;  // open x planes: 0 equilibrium (the reference's), 1 Zou-He faces (Hecht-Harting)
  // Population layout: population i of cell c is element c + NA (c & ~63) + NS i. NA = 0: one array
  // per population, NS elements apart (N, or N padded: the elements NS - N after each population
  // hold no population); NA = Q - 1, NS = 64 (rows of a multiple of 64 cells): tiles of 64 cells, each
  // holding its cells' Q populations one after another. N stays the cell count.
  uint32_t NS;
  uint32_t NA;
};

// Where the step kernels put a force sample: per-block FP64 partials, a completion counter,
// the FP64 history ring and its state words (STATE_WORDS u32: the sample count, then at word 2 the
// 64-bit number of the first step whose sample was not finite, 0 while none was). sample == 0
// skips the reduction; step is the number of the step the sample belongs to (1 for the first).
struct ForceSink {
  double* partials;
  uint32_t* counter;
  double* ring;
  uint32_t* state;
  int32_t sample;
  uint64_t step;
};

constexpr int HISTORY_LEN = 8192;
constexpr int STATE_WORDS = 4;

// The reference's blow-up rule (src/tunnel3d.ts:97-99): a force sample with a non-finite component
// means the flow has diverged. Records the step of the first such sample; called by the one thread
// that appends the sample, and a tunnel's samples are appended in stream order, so no atomics.
template <int DIM>
__device__ __forceinline__ void note_divergence(const double (&v)[DIM], uint32_t* state, uint64_t step) {
  bool finite = true;
#pragma unroll
  for (int d = 0; d < DIM; d++) finite = finite && __builtin_isfinite(v[d]);
  if (finite) return;
  uint64_t* first = reinterpret_cast<uint64_t*>(state + 2);
  if (__hip_atomic_load(first, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    __hip_atomic_store(first, step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace lbm
