// Batched D2Q9: many independent tunnels of uneven size and Reynolds number advanced by one launch
// per time step, with the step loop captured in hipGraphs.
//
// No counterpart in the reference (it runs one tunnel per page). Each tunnel is an ordinary
// single-tunnel solver (capi.hip) created on the batch's stream, so setSdf / initField / readMacro
// / readForces keep their per-tunnel semantics; only the stepping is shared. One launch covers every
// tunnel through a flat block table: blocks [0, n_bulk) are bulk tiles of 256 consecutive cells of
// one tunnel, then the link tiles and the body tiles of the tunnels' boundary lists, 256 nodes each.
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
// of every tunnel (each tunnel has at least one body tile, which writes its force samples).
__global__ __launch_bounds__(256) void batch2_step(const TunnelDesc* __restrict__ T,
                                                   const uint16_t* __restrict__ block_tunnel, uint32_t n_bulk_blocks,
                                                   uint32_t link_end, int parity, const uint64_t* __restrict__ base_step,
                                                   uint32_t k) {
  const uint32_t b = blockIdx.x;
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
    const uint32_t idx = (b - d.bulk_block0) * 256u + threadIdx.x;
    if (idx >= p.N) return;
    const uint32_t y = div_w(idx, d.magic, d.shift);
    const uint32_t x = idx - y * uint32_t(p.W);
    bulk_cell2<false>(src, dst, d.flags, d.xtab, p, x, int(y), idx, Record{nullptr, nullptr});
    return;
  }
  if (b < link_end) {
    const uint32_t kl = (b - d.link_block0) * 256u + threadIdx.x;
    float none[2] = {0.0f, 0.0f};
    if (kl < d.lt.n_link) list_cell2<false, false>(src, dst, d.flags, d.xtab, d.lt, kl, p, Record{nullptr, nullptr}, none);
    return;
  }
  const uint32_t tile = b - d.body_block0;
  const uint32_t kk = tile * 256u + threadIdx.x;
  float force[2] = {0.0f, 0.0f};
  if (kk < d.lt.n_body) list_cell2<true, false>(src, dst, d.flags, d.xtab, d.lt, d.lt.n_link + kk, p, Record{nullptr, nullptr}, force);
  if (d.force_every > 0 && (step + 1) % uint64_t(d.force_every) == 0) {
    ForceSink fs{d.partials, d.counter, d.ring, d.state, 1};
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

// Rebuilds the descriptor and block tables (after geometry, parameter or ramp changes).
int refresh(lbm_batch* b) {
  if (!b->dirty) return LBM_OK;
  const size_t n = b->t.size();
  b->desc.assign(n, TunnelDesc{});
  uint32_t blocks = 0;
  for (size_t i = 0; i < n; i++) {
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
    blocks += (s->N + 255) / 256;
  }
  if (blocks != b->n_bulk_blocks) drop_graphs(b);
  b->n_bulk_blocks = blocks;
  for (size_t i = 0; i < n; i++) {
    TunnelDesc& d = b->desc[i];
    d.link_block0 = blocks;
    blocks += list_blocks(d.lt.n_link, 256);
  }
  if (blocks != b->link_end) drop_graphs(b);
  b->link_end = blocks;
  for (size_t i = 0; i < n; i++) {
    TunnelDesc& d = b->desc[i];
    d.body_block0 = blocks;
    d.n_body_blocks = body_list_blocks(d.lt.n_body, 256, true);
    blocks += d.n_body_blocks;
  }
  if (blocks > 0x7fffffffu) return fail(LBM_ERR_INVALID, "batch too large for one launch");
  if (blocks != b->n_blocks) drop_graphs(b);  // the captured launches hold the old grid and table bounds
  b->n_blocks = blocks;
  b->block_tunnel.assign(blocks, 0);
  for (size_t i = 0; i < n; i++) {
    const TunnelDesc& d = b->desc[i];
    for (uint32_t k = d.bulk_block0; k < d.bulk_block0 + (b->t[i]->N + 255) / 256; k++) b->block_tunnel[k] = uint16_t(i);
    for (uint32_t k = d.link_block0; k < d.link_block0 + list_blocks(d.lt.n_link, 256); k++) b->block_tunnel[k] = uint16_t(i);
    for (uint32_t k = d.body_block0; k < d.body_block0 + d.n_body_blocks; k++) b->block_tunnel[k] = uint16_t(i);
  }
  if (blocks > b->block_cap) {
    if (b->d_block_tunnel) HIP_TRY(hipFree(b->d_block_tunnel));
    HIP_TRY(hipMalloc(&b->d_block_tunnel, sizeof(uint16_t) * blocks));
    b->block_cap = blocks;
    drop_graphs(b);  // the captured kernels hold the old table pointer
  }
  HIP_TRY(hipMemcpyAsync(b->d_desc, b->desc.data(), sizeof(TunnelDesc) * n, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->d_block_tunnel, b->block_tunnel.data(), sizeof(uint16_t) * blocks, hipMemcpyHostToDevice,
                         b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));  // host tables are reused on the next refresh
  b->dirty = false;
  return LBM_OK;
}

void launch_step(lbm_batch* b, int parity, uint32_t k) {
  batch2_step<<<b->n_blocks, 256, 0, b->stream>>>(b->d_desc, b->d_block_tunnel, b->n_bulk_blocks, b->link_end, parity,
                                                  b->d_step, k);
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
