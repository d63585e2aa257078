// Batched D2Q9: many independent tunnels of uneven size and Reynolds number advanced by one launch
// per time step, with the step loop captured in hipGraphs.
//
// No counterpart in the reference (it runs one tunnel per page). Each tunnel is an ordinary
// single-tunnel solver (capi.hip) created on the batch's stream, so setSdf / initField / readMacro
// / readForces keep their per-tunnel semantics; only the stepping is shared. One launch covers every
// tunnel through a flat block table: blocks [0, n_bulk) are bulk tiles of 256 consecutive cells of
// one tunnel, then the link tiles and the body tiles of the tunnels' boundary lists. The tile width
// (64 to 512 cells or nodes, one thread each) and the order of the tunnels in the table are picked
// by timing (lbm_batch_autotune); neither changes a stored bit.
// Every tile kind runs the same per-cell code as the single-tunnel kernels (d2q9.hpp), and the body
// tiles of each tunnel reduce that tunnel's momentum-exchange force into its own FP64 ring in the
// same block order, so a batched tunnel computes exactly what it would alone. Per-tunnel parameters, the inlet ramp and the force cadence are read
// on the device from a descriptor table and a device step counter, so a captured graph stays valid
// while the host advances time.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "d2q9.hpp"
#include "reduce.hpp"
#include "solver_internal.hpp"

using namespace lbm;

namespace lbm {

struct TunnelDesc {
  Params2 p;
  float* f[2];
  const uint32_t* flags;
  const float4* xtab;
  ListTables lt;
  double* partials;
  uint32_t* counter;
  double* ring;
  uint32_t* state;
  uint32_t bulk_block0, link_block0, body_block0, n_body_blocks;
  // y = floor(idx / W) by multiplication: Granlund & Montgomery, "Division by invariant integers
  // using multiplication" (PLDI 1994), round-up variant: q = (mulhi(n, m) + n) >> l, exact for all
  // 32-bit n, with l = ceil(log2 W) and m = floor(2^32 (2^l - W) / W) + 1.
  uint32_t magic, shift;
  int32_t force_every;
  int32_t ramp_kind;
  double u_target, ramp_steps;
};

__device__ __forceinline__ uint32_t div_w(uint32_t n, uint32_t m, uint32_t l) {
  return uint32_t((uint64_t(__umulhi(n, m)) + n) >> l);
}

// Block table layout: bulk tiles of every tunnel, then link tiles of every tunnel, then body tiles
// of every tunnel (each tunnel has at least one body tile, which writes its force samples); tiles of
// TILE cells or list nodes, one thread each. With reverse set the bulk tiles are taken from the end
// of the table (the last tile first): alternating the direction every step makes a step begin on
// the cells the previous one wrote last, which are still in the 256 MB Infinity Cache when the
// batch's population buffers do not fit in it.
template <uint32_t TILE>
__global__ __launch_bounds__(TILE) void batch2_step(const TunnelDesc* __restrict__ T,
                                                   const uint16_t* __restrict__ block_tunnel, uint32_t n_bulk_blocks,
                                                   uint32_t link_end, int parity, const uint64_t* __restrict__ base_step,
                                                   uint32_t k, uint32_t reverse) {
  const uint32_t b = reverse && blockIdx.x < n_bulk_blocks ? n_bulk_blocks - 1u - blockIdx.x : blockIdx.x;
  const TunnelDesc& d = T[block_tunnel[b]];
  Params2 p = d.p;
  const uint64_t step = *base_step + k;
  if (d.ramp_kind) {
    // same double-precision evaluation as the single-tunnel host ramp
    double r = d.ramp_steps > 0 ? fmin(1.0, double(step) / d.ramp_steps) : 1.0;
    if (d.ramp_kind == 2) r = r * r * (3.0 - 2.0 * r);
    p.uIn = float(d.u_target * r);
  }
  const float* src = d.f[parity];
  float* dst = d.f[parity ^ 1];
  if (b < n_bulk_blocks) {
    const uint32_t idx = (b - d.bulk_block0) * TILE + threadIdx.x;
    if (idx >= p.N) return;
    const uint32_t y = div_w(idx, d.magic, d.shift);
    const uint32_t x = idx - y * uint32_t(p.W);
    bulk_cell2<false>(src, dst, d.flags, d.xtab, p, x, int(y), idx, Record{nullptr, nullptr});
    return;
  }
  if (b < link_end) {
    const uint32_t kl = (b - d.link_block0) * TILE + threadIdx.x;
    float none[2] = {0.0f, 0.0f};
    if (kl < d.lt.n_link) list_cell2<false, false>(src, dst, d.flags, d.xtab, d.lt, kl, p, Record{nullptr, nullptr}, none);
    return;
  }
  const uint32_t tile = b - d.body_block0;
  const uint32_t kk = tile * TILE + threadIdx.x;
  float force[2] = {0.0f, 0.0f};
  if (kk < d.lt.n_body) list_cell2<true, false>(src, dst, d.flags, d.xtab, d.lt, d.lt.n_link + kk, p, Record{nullptr, nullptr}, force);
  if (d.force_every > 0 && (step + 1) % uint64_t(d.force_every) == 0) {
    ForceSink fs{d.partials, d.counter, d.ring, d.state, 1, step + 1};
    force_sample<2>(force, fs, tile, d.n_body_blocks);
  }
}

__global__ void set_u64_kernel(uint64_t* p, uint64_t v) { *p = v; }

}  // namespace lbm

struct lbm_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<lbm_solver*> t;
  std::vector<TunnelDesc> desc;
  std::vector<uint16_t> block_tunnel;
  TunnelDesc* d_desc = nullptr;
  uint16_t* d_block_tunnel = nullptr;
  size_t block_cap = 0;
  uint64_t* d_step = nullptr;
  uint32_t n_bulk_blocks = 0, link_end = 0, n_blocks = 0;
  int64_t step = 0;
  int parity = 0;
  int ramp_kind = 0;
  double ramp_steps = 0;
  bool dirty = true;
  std::vector<uint64_t> seen_gen;  // each tunnel's boundary-table generation at the last refresh
  uint32_t tile = 256;  // cells (bulk) or list nodes per block
  int order = 0;        // bit 0: tunnels in creation order (0) or largest first (1); bit 1: bulk tiles
                        // walked from the end of the table on the steps of parity 1
  static constexpr int GRAPH_STEPS = 32;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
};

namespace {

void magic_for(uint32_t W, uint32_t* m, uint32_t* l) {
  uint32_t s = 0;
  while ((uint64_t(1) << s) < W) s++;
  *l = s;
  *m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << s) - W)) / W + 1);
}

void drop_graphs(lbm_batch* b) {
  for (auto& g : b->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// Whether a tunnel changed since the descriptors were built, also through a handle the caller kept
// instead of fetching it with lbm_batch_tunnel: a geometry rebuild or a population-stride change
// replaces its table and population buffers (and bumps tables_gen), a parameter update its Params2.
bool tunnels_changed(const lbm_batch* b) {
  if (b->desc.size() != b->t.size() || b->seen_gen.size() != b->t.size()) return true;
  for (size_t i = 0; i < b->t.size(); i++) {
    const lbm_solver* s = b->t[i];
    const TunnelDesc& d = b->desc[i];
    if (s->tables_gen != b->seen_gen[i] || std::memcmp(&s->p2, &d.p, sizeof(Params2)) != 0 ||
        d.force_every != s->cfg2.force_every || d.u_target != s->cfg2.u_in)
      return true;
  }
  return false;
}

// Rebuilds the descriptor and block tables (after geometry, parameter, ramp or launch-shape changes).
int refresh(lbm_batch* b) {
  if (!b->dirty && !tunnels_changed(b)) return LBM_OK;
  const size_t n = b->t.size();
  const uint32_t T = b->tile;
  std::vector<size_t> perm(n);
  for (size_t i = 0; i < n; i++) perm[i] = i;
  if (b->order & 1)
    std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t c) { return b->t[a]->N > b->t[c]->N; });
  b->desc.assign(n, TunnelDesc{});
  uint32_t blocks = 0;
  for (size_t i : perm) {
    lbm_solver* s = b->t[i];
    TunnelDesc& d = b->desc[i];
    d.p = s->p2;
    d.f[0] = static_cast<float*>(s->f[0]);
    d.f[1] = static_cast<float*>(s->f[1]);
    d.flags = s->flags;
    d.xtab = s->xtab;
    d.lt = ListTables{s->list, s->gtab, s->bmask, s->xmask, s->lbase, s->links, s->n_link, s->n_body};
    d.partials = s->partials;
    d.counter = s->counter;
    d.ring = s->ring;
    d.state = s->state;
    magic_for(uint32_t(s->W), &d.magic, &d.shift);
    d.force_every = s->cfg2.force_every;
    d.ramp_kind = b->ramp_kind;
    d.u_target = s->cfg2.u_in;
    d.ramp_steps = b->ramp_steps;
    d.bulk_block0 = blocks;
    blocks += (s->N + T - 1) / T;
  }
  if (blocks != b->n_bulk_blocks) drop_graphs(b);
  b->n_bulk_blocks = blocks;
  for (size_t i : perm) {
    TunnelDesc& d = b->desc[i];
    d.link_block0 = blocks;
    blocks += list_blocks(d.lt.n_link, T);
  }
  if (blocks != b->link_end) drop_graphs(b);
  b->link_end = blocks;
  for (size_t i : perm) {
    TunnelDesc& d = b->desc[i];
    d.body_block0 = blocks;
    d.n_body_blocks = body_list_blocks(d.lt.n_body, T, true);
    blocks += d.n_body_blocks;
  }
  if (blocks > 0x7fffffffu) return fail(LBM_ERR_INVALID, "batch too large for one launch");
  if (blocks != b->n_blocks) drop_graphs(b);  // the captured launches hold the old grid and table bounds
  b->n_blocks = blocks;
  b->block_tunnel.assign(blocks, 0);
  for (size_t i = 0; i < n; i++) {
    const TunnelDesc& d = b->desc[i];
    for (uint32_t k = d.bulk_block0; k < d.bulk_block0 + (b->t[i]->N + T - 1) / T; k++) b->block_tunnel[k] = uint16_t(i);
    for (uint32_t k = d.link_block0; k < d.link_block0 + list_blocks(d.lt.n_link, T); k++) b->block_tunnel[k] = uint16_t(i);
    for (uint32_t k = d.body_block0; k < d.body_block0 + d.n_body_blocks; k++) b->block_tunnel[k] = uint16_t(i);
  }
  if (blocks > b->block_cap) {
    if (b->d_block_tunnel) HIP_TRY(hipFree(b->d_block_tunnel));
    b->d_block_tunnel = nullptr;
    b->block_cap = 0;
    HIP_TRY(hipMalloc(&b->d_block_tunnel, sizeof(uint16_t) * blocks));
    b->block_cap = blocks;
    drop_graphs(b);  // the captured kernels hold the old table pointer
  }
  HIP_TRY(hipMemcpyAsync(b->d_desc, b->desc.data(), sizeof(TunnelDesc) * n, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->d_block_tunnel, b->block_tunnel.data(), sizeof(uint16_t) * blocks, hipMemcpyHostToDevice,
                         b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));  // host tables are reused on the next refresh
  b->seen_gen.resize(n);
  for (size_t i = 0; i < n; i++) b->seen_gen[i] = b->t[i]->tables_gen;
  b->dirty = false;
  return LBM_OK;
}

void launch_step(lbm_batch* b, int parity, uint32_t k) {
  const uint32_t reverse = (b->order & 2) && parity ? 1u : 0u;
#define LBM_BATCH_LAUNCH(T)                                                                                      \
  batch2_step<T><<<b->n_blocks, T, 0, b->stream>>>(b->d_desc, b->d_block_tunnel, b->n_bulk_blocks, b->link_end, \
                                                   parity, b->d_step, k, reverse)
  switch (b->tile) {
    case 64: LBM_BATCH_LAUNCH(64); break;
    case 128: LBM_BATCH_LAUNCH(128); break;
    case 512: LBM_BATCH_LAUNCH(512); break;
    default: LBM_BATCH_LAUNCH(256); break;
  }
#undef LBM_BATCH_LAUNCH
}

// Graph of GRAPH_STEPS steps starting at the given parity, reading the step from d_step + k.
int capture(lbm_batch* b, int parity) {
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < lbm_batch::GRAPH_STEPS; k++) launch_step(b, parity ^ (k & 1), uint32_t(k));
  HIP_TRY(hipStreamEndCapture(b->stream, &g));
  HIP_TRY(hipGraphInstantiate(&b->graph[parity], g, nullptr, nullptr, 0));
  HIP_TRY(hipGraphDestroy(g));
  return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_batch_create_2d(const lbm_config2d* cfgs, int32_t n, int device, lbm_batch** out) {
  if (!cfgs || n <= 0 || !out) return fail(LBM_ERR_INVALID, "bad arguments");
  if (n > 65535) return fail(LBM_ERR_INVALID, "at most 65535 tunnels per batch");
  *out = nullptr;
  lbm_batch* b = new lbm_batch();
  b->device = device;
  int rc = LBM_OK;
  auto cleanup = [&]() {
    for (lbm_solver* s : b->t) lbm_destroy(s);
    if (b->d_desc) (void)hipFree(b->d_desc);
    if (b->d_block_tunnel) (void)hipFree(b->d_block_tunnel);
    if (b->d_step) (void)hipFree(b->d_step);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
  };
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    cleanup();
    return fail(LBM_ERR_DEVICE, "batch: no device or stream");
  }
  for (int32_t i = 0; i < n; i++) {
    lbm_solver* s = nullptr;
    rc = create_2d_on(&cfgs[i], device, b->stream, &s);
    if (rc != LBM_OK) {
      std::string msg = lbm_last_error();
      cleanup();
      return fail(rc, "tunnel " + std::to_string(i) + ": " + msg);
    }
    b->t.push_back(s);
  }
  if (hipMalloc(&b->d_desc, sizeof(TunnelDesc) * size_t(n)) != hipSuccess ||
      hipMalloc(&b->d_step, sizeof(uint64_t)) != hipSuccess) {
    cleanup();
    return fail(LBM_ERR_ALLOC, "batch allocation failed");
  }
  rc = refresh(b);
  if (rc != LBM_OK) {
    std::string msg = lbm_last_error();
    cleanup();
    return fail(rc, msg);
  }
  *out = b;
  return LBM_OK;
}

void lbm_batch_destroy(lbm_batch* b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->stream);
  drop_graphs(b);
  for (lbm_solver* s : b->t) lbm_destroy(s);
  (void)hipFree(b->d_desc);
  (void)hipFree(b->d_block_tunnel);
  (void)hipFree(b->d_step);
  (void)hipStreamDestroy(b->stream);
  delete b;
}

int lbm_batch_size(const lbm_batch* b, int32_t* n, int64_t* total_cells) {
  if (!b) return fail(LBM_ERR_INVALID, "null batch");
  int64_t cells = 0;
  for (const lbm_solver* s : b->t) cells += s->N;
  if (n) *n = int32_t(b->t.size());
  if (total_cells) *total_cells = cells;
  return LBM_OK;
}

lbm_solver* lbm_batch_tunnel(lbm_batch* b, int32_t i) {
  if (!b || i < 0 || size_t(i) >= b->t.size()) {
    fail(LBM_ERR_INVALID, "tunnel index out of range");
    return nullptr;
  }
  b->dirty = true;  // the caller may change its geometry or parameters
  return b->t[size_t(i)];
}

int lbm_batch_set_inlet_ramp(lbm_batch* b, int kind, double steps) {
  if (!b || kind < 0 || kind > 2) return fail(LBM_ERR_INVALID, "bad ramp");
  b->ramp_kind = kind;
  b->ramp_steps = steps;
  b->dirty = true;
  return LBM_OK;
}

int lbm_batch_enqueue_steps(lbm_batch* b, int64_t n) {
  if (!b || n < 0) return fail(LBM_ERR_INVALID, "bad arguments");
  int rc = refresh(b);
  if (rc) return rc;
  for (lbm_solver* s : b->t)
    if (int rc2 = check_usable(s)) return rc2;
  for (lbm_solver* s : b->t)
    if (s->step != b->step || s->parity != b->parity)
      return fail(LBM_ERR_INVALID, "a tunnel of the batch was stepped or re-initialised on its own; call "
                                   "lbm_batch_init_all after initField");
  int64_t done = 0;
  while (done < n) {
    set_u64_kernel<<<1, 1, 0, b->stream>>>(b->d_step, uint64_t(b->step));
    const int64_t left = n - done;
    if (left >= lbm_batch::GRAPH_STEPS) {
      if (!b->graph[b->parity]) {
        rc = capture(b, b->parity);
        if (rc) return rc;
      }
      HIP_TRY(hipGraphLaunch(b->graph[b->parity], b->stream));
      b->step += lbm_batch::GRAPH_STEPS;
      done += lbm_batch::GRAPH_STEPS;
      // GRAPH_STEPS is even: the parity is unchanged
    } else {
      for (int64_t k = 0; k < left; k++) launch_step(b, b->parity ^ int(k & 1), uint32_t(k));
      b->step += left;
      b->parity ^= int(left & 1);
      done += left;
    }
  }
  HIP_TRY(hipGetLastError());
  for (lbm_solver* s : b->t) record_steps(s, n);
  return LBM_OK;
}

// After initField on the tunnels (which resets each to step 0, parity 0): realigns the batch clock.
int lbm_batch_reset_clock(lbm_batch* b) {
  if (!b) return fail(LBM_ERR_INVALID, "null batch");
  for (lbm_solver* s : b->t)
    if (s->step != 0 || s->parity != 0) return fail(LBM_ERR_INVALID, "every tunnel must be at step 0");
  b->step = 0;
  b->parity = 0;
  b->dirty = true;
  return LBM_OK;
}

int lbm_batch_set_launch(lbm_batch* b, int32_t tile, int32_t order) {
  if (!b) return fail(LBM_ERR_INVALID, "null batch");
  if (tile != 64 && tile != 128 && tile != 256 && tile != 512) return fail(LBM_ERR_INVALID, "tile must be 64, 128, 256 or 512");
  if (order < 0 || order > 3) return fail(LBM_ERR_INVALID, "order must be 0-3");
  if (uint32_t(tile) != b->tile || order != b->order) {
    b->tile = uint32_t(tile);
    b->order = order;
    b->dirty = true;
    drop_graphs(b);
  }
  return LBM_OK;
}

int lbm_batch_launch_info(lbm_batch* b, int32_t* tile, int32_t* order) {
  if (!b) return fail(LBM_ERR_INVALID, "null batch");
  if (tile) *tile = int32_t(b->tile);
  if (order) *order = b->order;
  return LBM_OK;
}

// Times every tile width with both tunnel orders, each with the bulk tiles walked one way or
// alternating, over trial_steps launches (after two untimed ones;
// every candidate three times, interleaved, keeping its fastest time) on the tunnels' own state, keeps the
// fastest, and restores every tunnel's populations and force ring (the batch's and tunnels' step
// counters are not advanced). Neither the tile width nor the order changes a result: cells are
// independent within a step and each force sample is summed per wave of body nodes in node order.
int lbm_batch_autotune(lbm_batch* b, int32_t trial_steps, int32_t* tile, int32_t* order) {
  if (!b || trial_steps <= 0) return fail(LBM_ERR_INVALID, "bad arguments");
  for (lbm_solver* s : b->t)
    if (int rc = check_usable(s)) return rc;
  struct Saved {
    void* f[2] = {nullptr, nullptr};
    void* ring = nullptr;
    void* state = nullptr;
    uint32_t na0 = 0, ns0 = 0;  // the tunnel's population layout when saved
  };
  struct Guard {
    lbm_batch* b;
    std::vector<Saved> sv;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    uint32_t tile0;
    int order0;
    bool saved = false, keep = false;
    int restore() {
      saved = false;
      for (size_t i = 0; i < b->t.size(); i++) {
        lbm_solver* s = b->t[i];
        const size_t fb = pop_elems(s) * s->elem;  // whole buffers, in the tunnel's layout
        for (int k = 0; k < 2; k++) {
          if (s->NA == sv[i].na0 && s->NS == sv[i].ns0) {
            HIP_TRY(hipMemcpyAsync(s->f[k], sv[i].f[k], fb, hipMemcpyDeviceToDevice, b->stream));
          } else if (int rc = relayout(s, sv[i].f[k], s->f[k], sv[i].na0, sv[i].ns0, s->NA, s->NS, b->stream)) {
            return rc;
          }
        }
        HIP_TRY(hipMemcpyAsync(s->ring, sv[i].ring, sizeof(double) * 3 * HISTORY_LEN, hipMemcpyDeviceToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(s->state, sv[i].state, 4 * STATE_WORDS, hipMemcpyDeviceToDevice, b->stream));
      }
      HIP_TRY(hipStreamSynchronize(b->stream));
      return LBM_OK;
    }
    ~Guard() {
      if (saved) (void)restore();
      if (!keep)
        for (size_t i = 0; i < b->t.size(); i++)
          if (b->t[i]->NA != sv[i].na0 || b->t[i]->NS != sv[i].ns0) (void)set_layout(b->t[i], sv[i].na0, sv[i].ns0);
      (void)hipStreamSynchronize(b->stream);
      for (Saved& x : sv)
        for (void* p : {x.f[0], x.f[1], x.ring, x.state})
          if (p) (void)hipFree(p);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (!keep && (b->tile != tile0 || b->order != order0)) {
        b->tile = tile0;
        b->order = order0;
        b->dirty = true;
        drop_graphs(b);
      }
    }
  } g{b, std::vector<Saved>(b->t.size())};
  g.tile0 = b->tile;
  g.order0 = b->order;
  for (size_t i = 0; i < b->t.size(); i++) {
    lbm_solver* s = b->t[i];
    const size_t fb = pop_elems(s) * s->elem;  // whole buffers, in the tunnel's layout
    g.sv[i].na0 = s->NA;
    g.sv[i].ns0 = s->NS;
    HIP_TRY(hipMalloc(&g.sv[i].f[0], fb));
    HIP_TRY(hipMalloc(&g.sv[i].f[1], fb));
    HIP_TRY(hipMalloc(&g.sv[i].ring, sizeof(double) * 3 * HISTORY_LEN));
    HIP_TRY(hipMalloc(&g.sv[i].state, 4 * STATE_WORDS));
  }
  HIP_TRY(hipEventCreate(&g.e0));
  HIP_TRY(hipEventCreate(&g.e1));
  for (size_t i = 0; i < b->t.size(); i++) {
    lbm_solver* s = b->t[i];
    const size_t fb = pop_elems(s) * s->elem;  // whole buffers, in the tunnel's layout
    HIP_TRY(hipMemcpyAsync(g.sv[i].f[0], s->f[0], fb, hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(g.sv[i].f[1], s->f[1], fb, hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(g.sv[i].ring, s->ring, sizeof(double) * 3 * HISTORY_LEN, hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(g.sv[i].state, s->state, 4 * STATE_WORDS, hipMemcpyDeviceToDevice, b->stream));
  }
  g.saved = true;
  std::vector<std::pair<uint32_t, int>> cand;
  for (int o : {0, 1, 2, 3})
    for (uint32_t t : {64u, 128u, 256u, 512u}) cand.push_back({t, o});
  const size_t nc = cand.size();
  std::vector<double> best_of(nc, 1e300);
  for (int rep = 0; rep < 3; rep++) {
    for (size_t c = 0; c < nc; c++) {
      b->tile = cand[c].first;
      b->order = cand[c].second;
      b->dirty = true;
      if (int rc = refresh(b)) return rc;
      set_u64_kernel<<<1, 1, 0, b->stream>>>(b->d_step, uint64_t(b->step));
      launch_step(b, b->parity, 0);
      launch_step(b, b->parity ^ 1, 1);
      HIP_TRY(hipEventRecord(g.e0, b->stream));
      for (int k = 0; k < trial_steps; k++) launch_step(b, b->parity ^ (k & 1), uint32_t(2 + k));
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipEventRecord(g.e1, b->stream));
      HIP_TRY(hipEventSynchronize(g.e1));
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, g.e0, g.e1));
      best_of[c] = std::min(best_of[c], double(ms));
    }
  }
  size_t best = 0;
  for (size_t c = 1; c < nc; c++)
    if (best_of[c] < best_of[best]) best = c;
  if (int rc = g.restore()) return rc;
  b->tile = cand[best].first;
  b->order = cand[best].second;
  b->dirty = true;
  drop_graphs(b);
  g.keep = true;
  if (int rc = refresh(b)) return rc;
  if (tile) *tile = int32_t(b->tile);
  if (order) *order = b->order;
  return LBM_OK;
}

int lbm_batch_diverged(lbm_batch* b, int64_t* first_steps, size_t n) {
  if (!b || !first_steps || n < b->t.size()) return fail(LBM_ERR_INVALID, "first_steps must hold one value per tunnel");
  std::vector<uint64_t> v(b->t.size(), 0);
  for (size_t i = 0; i < b->t.size(); i++)
    HIP_TRY(hipMemcpyAsync(&v[i], b->t[i]->state + 2, sizeof(uint64_t), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < b->t.size(); i++) first_steps[i] = int64_t(v[i]);
  return LBM_OK;
}

int lbm_batch_sync(lbm_batch* b) {
  if (!b) return fail(LBM_ERR_INVALID, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return LBM_OK;
}

int lbm_batch_time_steps(lbm_batch* b, int64_t n, double* elapsed_ms) {
  if (!b || !elapsed_ms) return fail(LBM_ERR_INVALID, "bad arguments");
  int rc = refresh(b);
  if (rc) return rc;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, b->stream));
  rc = lbm_batch_enqueue_steps(b, n);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(e1, b->stream));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  *elapsed_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return LBM_OK;
}

}  // extern "C"
