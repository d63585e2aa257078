// Internal definition of a tunnel solver handle, shared by the C ABI (capi.hip) and the batched
// multi-tunnel driver (batch2d.hip). Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lbm.h"
#include "kernels.hpp"

namespace lbm {

int fail(int code, const std::string& msg);

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      return ::lbm::fail(LBM_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    }                                                                                     \
  } while (0)

// Inlet ramp evaluated per step (an extension; the reference's drivers ramp per chunk,
// src/cases.ts:52-54): kind 0 off, 1 linear, 2 smoothstep t^2 (3 - 2t). Evaluated in double, on the
// host for single tunnels and identically on the device for batches.
struct Ramp {
  int kind = 0;
  double u_target = 0, steps = 0;
  double at(int64_t k) const {
    double r = steps > 0 ? std::min(1.0, double(k) / steps) : 1.0;
    if (kind == 2) r = r * r * (3 - 2 * r);
    return u_target * r;
  }
};

struct Ticket {
  double* ring = nullptr;     // pinned, HISTORY_LEN * dim
  uint32_t* state = nullptr;  // pinned
  hipEvent_t ev = nullptr;
  int64_t generation = 0;
};

}  // namespace lbm

constexpr int SWEEP_MERGE = 4;
constexpr int SWEEP_EARLY = 8;

struct lbm_solver {
  int nd = 2;  // 2 or 3
  int device = 0;
  hipStream_t stream = nullptr;
  bool owns_stream = true;
  int W = 1, H = 1, D = 1;
  uint32_t N = 0;
  // Population layout of the buffers (Params3::NS / NA): population i of cell c is element
  // c + NA (c & ~63) + NS i. NA = 0, NS = N unless the autotuner or the caller chose a padded stride
  // (lbm_set_population_stride) or, for D3Q19 rows of a multiple of 64 cells, tiles of 64 cells
  // (NA = 18, NS = 64; lbm_set_population_layout).
  uint32_t NS = 0;
  uint32_t NA = 0;
  int Q = 9;
  int prec = LBM_FP32;
  size_t elem = 4;
  lbm_config2d cfg2{};
  lbm_config3d cfg3{};
  lbm::Params2 p2{};
  lbm::Params3 p3{};
  lbm::Ramp ramp;
  int faces = LBM_FACES_EQUILIBRIUM;  // 3D open x planes (lbm_set_open_faces)
  int bulk_block = 256;
  // 3D launch variant (autotuned). Bits 0-1, row order: 0 = always bottom-up; 1 / 2 = top-down on
  // the steps that read the buffer at the higher / lower address, bottom-up on the others. Bit 2
  // (SWEEP_MERGE): inlet and outlet nodes share their wave's stores. Bit 3 (SWEEP_EARLY): the pulls
  // are issued before the node's kind is known.
  int sweep = 1;

  void* f[2] = {nullptr, nullptr};
  uint32_t* flags = nullptr;
  uint8_t* touches = nullptr;
  float* sdf = nullptr;
  float4* mac = nullptr;
  uint32_t* list = nullptr;
  uint32_t nlist = 0;
  uint32_t n_link = 0, n_body = 0;  // the list holds n_link nodes without body links, then n_body with
  uint32_t* gtab = nullptr;         // gather table of every list node, direction-major
  uint32_t* bmask = nullptr;        // Bouzidi directions per list node
  uint32_t* xmask = nullptr;        // slip-through-solid (3D) / moving-lid (2D) directions per list node
  uint32_t* lbase = nullptr;        // first Bouzidi record of each list node
  lbm::BzLink* links = nullptr;     // Bouzidi records
  uint32_t n_links = 0;
  std::vector<float> sdf_host;      // the SDF as last set (q is computed from it on the host)
  uint8_t* row_clean = nullptr;     // 3D: per row (y + H z), 1 when no flag word of it needs reading
  // Host copies of the boundary-list tables and the flag words as last built, from which the
  // discrete adjoint builds its transposed tables; bumped generation on every rebuild.
  std::vector<uint32_t> h_flags, h_list, h_gtab, h_bmask, h_xmask, h_lbase;
  std::vector<lbm::BzLink> h_links;
  std::vector<uint32_t> h_link_node;  // cell of each Bouzidi record
  std::vector<uint8_t> h_link_dir;    // direction of each Bouzidi record
  uint64_t tables_gen = 0;
  // Non-empty after a geometry rebuild failed part way (the flag kernel had already rewritten the
  // device flags, so they no longer match the boundary tables): stepping is refused with this
  // message until a rebuild succeeds.
  std::string invalid;
  int debug_fail_rebuilds = 0;  // test hook (lbm_debug_fail_rebuilds)
  struct lbm_adjoint* adj = nullptr;  // discrete-adjoint state, created on first use
  int64_t n_bulk = 0, n_solid = 0;
  double* partials = nullptr;
  uint32_t* counter = nullptr;
  double* ring = nullptr;
  uint32_t* state = nullptr;
  float4* xtab = nullptr;  // per-x relaxation coefficients of the step kernels

  int64_t step = 0;
  int parity = 0;

  // readForces bookkeeping (src/solver.ts:57-61, 236-241, 315-335)
  int64_t samples_read = 0;
  // Step number of every force sample enqueued since the last reset, the newest HISTORY_LEN kept
  // (sample k at sample_steps[k % HISTORY_LEN]): readForces never returns a sample more than
  // HISTORY_LEN behind the newest, so nothing here grows while a run goes unread.
  std::vector<int64_t> sample_steps = std::vector<int64_t>(size_t(lbm::HISTORY_LEN), 0);
  int64_t samples_pushed = 0;
  void push_sample_step(int64_t st) {
    sample_steps[size_t(samples_pushed % lbm::HISTORY_LEN)] = st;
    samples_pushed++;
  }
  int64_t generation = 0;
  std::mutex mu;
  std::map<int64_t, lbm::Ticket> tickets;
  int64_t next_ticket = 1;

  int dim() const { return nd == 2 ? 2 : 3; }
  int force_every() const { return nd == 2 ? cfg2.force_every : cfg3.force_every; }
};

namespace lbm {
// Helpers of capi.hip used by the batch driver and the adjoint.
int create_2d_on(const lbm_config2d* cfg, int device, hipStream_t stream, lbm_solver** out);
int rebuild_flags(lbm_solver* s);
// LBM_OK when the solver's tables match its flags, else the error of the failed rebuild.
int check_usable(const lbm_solver* s);
void record_steps(lbm_solver* s, int64_t n);
// One step on the solver's stream: force sample as configured (sample < 0) or forced on / off
// (1 / 0); with rec, the step also writes the adjoint's record. partials / counter replace the
// handle's force-reduction scratch when given (lbm_enqueue_steps_noalloc).
int enqueue_step(lbm_solver* s, const Record* rec, int sample, double* partials = nullptr,
                 uint32_t* counter = nullptr);
// Force-reduction partials (3 doubles each) the next step of the solver writes, if it samples.
size_t force_partials_needed(const lbm_solver* s);

// The bulk rows' view of a D3Q19 node (step3_kernel): the face bits of its row, whether the flag
// word selects the bulk pass, and the gather index of each pull under the row's pattern.
struct Row3 {
  bool loY, hiY, loZ, hiZ;
  uint32_t mask;
};
Row3 row3(const lbm_solver* s, uint32_t y, uint32_t z);
// 0: not taken by the bulk rows, 1: fluid pulls, 2: inlet (stores the inlet equilibrium)
int bulk_kind3(const Row3& r, uint32_t v);
uint32_t row_gather3(const lbm_solver* s, const Row3& r, uint32_t idx, bool outlet, int i);

// Frees the adjoint state of a solver (adjoint.hip).
void adjoint_free(lbm_solver* s);

// Population layouts (capi.hip): move a solver's buffers and tables to layout (na, ns); whether tiles
// of 64 cells are allowed; copy the populations of buffer src in layout (na0, ns0) to dst in
// (na1, ns1) on stream st.
int set_layout(lbm_solver* s, uint32_t na, uint32_t ns);
bool tiles_ok(const lbm_solver* s);
int relayout(const lbm_solver* s, const void* src, void* dst, uint32_t na0, uint32_t ns0, uint32_t na1, uint32_t ns1,
             hipStream_t st);

// Element of population i at cell c in the solver's buffers, and the inverse; elements per buffer.
__host__ __device__ inline uint32_t pelem(uint32_t na, uint32_t ns, int i, uint32_t c) {
  return c + na * (c & ~63u) + ns * uint32_t(i);
}
inline uint32_t pelem(const lbm_solver* s, int i, uint32_t c) { return pelem(s->NA, s->NS, i, c); }
inline void pelem_decode(uint32_t na, uint32_t ns, int Q, uint32_t e, int* i, uint32_t* c) {
  if (na == 0) {
    *i = int(e / ns);
    *c = e % ns;
  } else {
    const uint32_t t = e / (64u * uint32_t(Q)), r = e % (64u * uint32_t(Q));
    *i = int(r / 64u);
    *c = t * 64u + r % 64u;
  }
}
inline size_t pop_elems(const lbm_solver* s) {
  return s->NA ? size_t(s->Q) * ((size_t(s->N) + 63u) / 64u * 64u) : size_t(s->Q) * s->NS;
}
}  // namespace lbm
