// C ABI of the solver (include/lbm.h): handles, device buffers, the step loop and readbacks.
//
// Stands in for the host halves of the reference's Solver (src/solver.ts:50-340) and Solver3D
// (src/solver3d.ts:42-319): parameter derivation (writeParams), geometry rebuild (setSdf /
// rebuildFlags), initField, resetForces, encodeSteps / encodeMacro / run, readMacro and the force
// ring bookkeeping of readForces. WebGPU buffers become hipMalloc allocations, the queue becomes
// one HIP stream per handle, and uniform-buffer writes become kernel arguments, so setters are
// stream-ordered exactly as queue.writeBuffer is.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lbm.h"
#include "kernels.hpp"
#include "solver_internal.hpp"

using namespace lbm;

namespace lbm {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

Row3 row3(const lbm_solver* s, uint32_t y, uint32_t z) {
  Row3 r{y == 0u, y == uint32_t(s->H - 1), z == 0u, z == uint32_t(s->D - 1), 0u};
  for (int i = 1; i < 19; i++) {
    const bool outY = (r.loY && CY3[i] > 0) || (r.hiY && CY3[i] < 0);
    const bool outZ = (r.loZ && CZ3[i] > 0) || (r.hiZ && CZ3[i] < 0);
    if (outY || outZ) r.mask |= 1u << i;
  }
  return r;
}
int bulk_kind3(const Row3& r, uint32_t v) {
  const bool face = r.loY || r.hiY || r.loZ || r.hiZ;
  if (v == (FLAG_INLET3 | INLET_DIRS3 | r.mask)) return 2;
  if (v == r.mask || (!face && v == (FLAG_OUTLET3 | OUTLET_DIRS3))) return 1;
  return 0;
}
uint32_t row_gather3(const lbm_solver* s, const Row3& r, uint32_t idx, bool outlet, int i) {
  const int64_t W = s->W, WH = int64_t(s->W) * s->H;
  const int64_t delta = CZ3[i] * WH + CY3[i] * W + CX3[i];
  const bool outY = (r.loY && CY3[i] > 0) || (r.hiY && CY3[i] < 0);
  const bool outZ = (r.loZ && CZ3[i] > 0) || (r.hiZ && CZ3[i] < 0);
  if (outlet && CX3[i] < 0) return pelem(s, i, idx - 1u);
  int sd = i;
  int64_t sdelta = delta;
  if (outY && outZ) {
    sd = opp3(i);
    sdelta = 0;
  } else if (outY) {
    sd = mirror_y3(i);
    sdelta = CZ3[i] * WH + CX3[i];
  } else if (outZ) {
    sd = mirror_z3(i);
    sdelta = CY3[i] * W + CX3[i];
  }
  return pelem(s, sd, uint32_t(int64_t(idx) - sdelta));
}

}  // namespace lbm

namespace {

constexpr double MAGIC_LAMBDA = 3.0 / 16.0;

void free_all(lbm_solver* s) {
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  adjoint_free(s);
  for (void* p : {s->f[0], s->f[1], (void*)s->flags, (void*)s->touches, (void*)s->sdf, (void*)s->mac,
                  (void*)s->list, (void*)s->partials, (void*)s->counter, (void*)s->ring, (void*)s->state,
                  (void*)s->xtab, (void*)s->gtab, (void*)s->bmask, (void*)s->xmask, (void*)s->lbase,
                  (void*)s->links, (void*)s->row_clean})
    if (p) (void)hipFree(p);
  for (auto& kv : s->tickets) {
    if (kv.second.ring) (void)hipHostFree(kv.second.ring);
    if (kv.second.state) (void)hipHostFree(kv.second.state);
    if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
  }
  s->tickets.clear();
  if (s->stream && s->owns_stream) (void)hipStreamDestroy(s->stream);
  s->stream = nullptr;
}

// Solver.writeParams (src/solver.ts:180-197).
void write_params2(lbm_solver* s) {
  const lbm_config2d& c = s->cfg2;
  Params2& p = s->p2;
  p.W = s->W;
  p.H = s->H;
  p.N = s->N;
  p.NS = s->NS;
  p.NA = s->NA;
  p.left = c.left;
  p.right = c.right;
  p.bottom = c.bottom;
  p.top = c.top;
  p.tau = float(c.tau);
  p.lambda = float(c.lambda);
  p.smagC2 = float(c.smagorinsky * c.smagorinsky);
  p.uIn = float(c.u_in);
  p.parabolic = c.parabolic ? 1 : 0;
  p.spongeStart = float(c.sponge_fraction > 0 ? (s->W - 1) * (1 - c.sponge_fraction) : 1e9);
  p.tauSponge = float(std::max(c.sponge_tau, c.tau));
  p.uLid = float(c.u_lid);
  p.fx = float(c.body_force[0]);
  p.fy = float(c.body_force[1]);
  p.forcing = (c.body_force[0] != 0 || c.body_force[1] != 0) ? 1 : 0;
  p.absorb = float(c.absorb);
  p.inletLayer = float(c.inlet_layer_fraction * (s->W - 1));
  if (s->xtab) launch_xtab2(p, s->xtab, s->stream);
}

// Solver3D.writeParams (src/solver3d.ts:197-208); lambda defaults to trtLambda3(tau) (src/lattice3d.ts:29-31).
void write_params3(lbm_solver* s) {
  const lbm_config3d& c = s->cfg3;
  Params3& p = s->p3;
  p.W = s->W;
  p.H = s->H;
  p.D = s->D;
  p.N = s->N;
  p.NS = s->NS;
  p.NA = s->NA;
  p.WH = uint32_t(s->W) * uint32_t(s->H);
  p.tau = float(c.tau);
  p.lambda = float(std::isnan(c.lambda) ? std::min(MAGIC_LAMBDA, 50.0 * (c.tau - 0.5) * (c.tau - 0.5)) : c.lambda);
  p.smagC2 = float(c.smagorinsky * c.smagorinsky);
  p.uIn = float(c.u_in);
  p.spongeStart = float(c.sponge_fraction > 0 ? (s->W - 1) * (1 - c.sponge_fraction) : 1e9);
  p.tauSponge = float(std::max(c.sponge_tau, c.tau));
  p.absorb = float(c.absorb);
  p.faces = uint32_t(s->faces);
  if (s->xtab) launch_xtab3(p, s->xtab, s->stream);
}

// feq(i, 1, (u, 0, 0)) in float, evaluated in the source's operation order without contraction
// (src/shaders/common3d.ts:78-81).
#pragma clang fp contract(off)
void inlet_feq3(float u, float* out) {
  static const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
  static const float Wt[19] = {1.0f / 3,  1.0f / 18, 1.0f / 18, 1.0f / 18, 1.0f / 18, 1.0f / 18, 1.0f / 18,
                               1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36,
                               1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36};
  const float rho = 1.0f, uy = 0.0f, uz = 0.0f;
  for (int i = 0; i < 19; i++) {
    float cu = 3.0f * (float(CX[i]) * u + 0.0f * uy + 0.0f * uz);
    out[i] = Wt[i] * (rho - 1.0f + rho * (cu + 0.5f * cu * cu - 1.5f * (u * u + uy * uy + uz * uz)));
  }
}

// The same for D3Q27 (kernels27.hip's feq27 in the same operation order).
void inlet_feq27(float u, float* out) {
  const float rho = 1.0f, uy = 0.0f, uz = 0.0f;
  for (int i = 0; i < 27; i++) {
    float cu = 3.0f * (float(CX27[i]) * u + float(CY27[i]) * uy + float(CZ27[i]) * uz);
    out[i] = W27[i] * (rho - 1.0f + rho * (cu + 0.5f * cu * cu - 1.5f * (u * u + uy * uy + uz * uz)));
  }
}

// Host resolution of one pull of a boundary node, following the rule order of pullSlow3
// (src/shaders/step3d.ts:119-168) and pullSlow (src/shaders/step.ts:119-175). Every pull becomes a
// gather index (population * N + cell); a Bouzidi pull also gets its record (q and the second
// operand), a 3D slip face whose mirror node is solid is marked (bounce-back carrying 2 f_j of
// force), and a 2D moving lid is marked (adds its lid term).
enum PullKind { PULL_GATHER = 0, PULL_BOUZIDI = 1, PULL_MARKED = 2 };

// q as the kernels computed it, in float: clamp(dF / max(dF - dS, 1e-6), 1e-3, 1).
float link_q(float dF, float dS) {
  return std::min(std::max(dF / std::max(dF - dS, 1e-6f), 1e-3f), 1.0f);
}

int resolve3(const lbm_solver* s, const uint32_t* fl, uint32_t idx, int i, uint32_t* g, BzLink* rec) {
  const int W = s->W, H = s->H, D = s->D;
  const uint32_t WH = uint32_t(W) * uint32_t(H);
  const int z = int(idx / WH), y = int((idx % WH) / uint32_t(W)), x = int(idx % uint32_t(W));
  const int j = opp3(i);
  const int sx = x - CX3[i], sy = y - CY3[i], sz = z - CZ3[i];
  auto at = [&](int ax, int ay, int az) { return uint32_t(ax) + uint32_t(W) * (uint32_t(ay) + uint32_t(H) * uint32_t(az)); };
  const uint32_t f = fl[idx];
  if ((f & (1u << i)) == 0u) {
    *g = pelem(s, i, at(sx, sy, sz));
    return PULL_GATHER;
  }
  const bool outX = sx < 0 || sx >= W, outY = sy < 0 || sy >= H, outZ = sz < 0 || sz >= D;
  const bool outlet = (f & FLAG_OUTLET3) != 0u;
  if (outX && (f & FLAG_INLET3)) {
    // entering from x = -1: a Zou-He inlet face rebuilds it (the equilibrium inlet pulls nothing)
    *g = pelem(s, j, idx);
  } else if (outX && outlet) {
    *g = pelem(s, i, at(W - 2, y, z));
  } else if (outY && outZ) {
    *g = pelem(s, j, idx);
  } else if (outY || outZ) {
    const uint32_t rs = outY ? at(sx, y, sz) : at(sx, sy, z);
    if (fl[rs] & FLAG_SOLID3) {
      *g = pelem(s, j, idx);
      return PULL_MARKED;
    }
    *g = pelem(s, outY ? mirror_y3(i) : mirror_z3(i), rs);
  } else if (outX) {
    *g = pelem(s, j, idx);
  } else {
    const uint32_t sc = at(sx, sy, sz);
    if (!(fl[sc] & FLAG_SOLID3)) {
      *g = pelem(s, i, sc);
      return PULL_GATHER;
    }
    *g = pelem(s, j, idx);
    rec->q = link_q(s->sdf_host[idx], s->sdf_host[sc]);
    if (rec->q < 0.5f) {
      const int nx = x + CX3[i], ny = y + CY3[i], nz = z + CZ3[i];
      const bool ok = nx >= 0 && nx < W && ny >= 0 && ny < H && nz >= 0 && nz < D && !(fl[at(nx, ny, nz)] & FLAG_SOLID3);
      rec->gb = ok ? pelem(s, j, at(nx, ny, nz)) : NO_OPERAND;
    } else {
      rec->gb = pelem(s, i, idx);
    }
    return PULL_BOUZIDI;
  }
  return PULL_GATHER;
}

int resolve2(const lbm_solver* s, const uint32_t* fl, uint32_t idx, int i, uint32_t* g, BzLink* rec) {
  const int W = s->W, H = s->H;
  const Params2& p = s->p2;
  const int y = int(idx / uint32_t(W)), x = int(idx % uint32_t(W));
  const int j = OPP2[i];
  int sx = x - CX2[i], sy = y - CY2[i];
  if ((fl[idx] & (1u << i)) == 0u) {
    *g = pelem(s, i, uint32_t(sy * W + sx));
    return PULL_GATHER;
  }
  auto wrap = [&](int& qx, int& qy) {
    if (qx < 0 && p.left == 0) qx += W;
    if (qx >= W && p.right == 0) qx -= W;
    if (qy < 0 && p.bottom == 0) qy += H;
    if (qy >= H && p.top == 0) qy -= H;
  };
  wrap(sx, sy);
  const bool inX = sx >= 0 && sx < W, inY = sy >= 0 && sy < H;
  if (inX && inY) {
    const uint32_t sc = uint32_t(sy * W + sx);
    if (!(fl[sc] & FLAG_SOLID2)) {
      *g = pelem(s, i, sc);
      return PULL_GATHER;
    }
    *g = pelem(s, j, idx);
    rec->q = link_q(s->sdf_host[idx], s->sdf_host[sc]);
    if (rec->q < 0.5f) {
      int nx = x + CX2[i], ny = y + CY2[i];
      wrap(nx, ny);
      const bool ok = nx >= 0 && nx < W && ny >= 0 && ny < H && !(fl[uint32_t(ny * W + nx)] & FLAG_SOLID2);
      rec->gb = ok ? pelem(s, j, uint32_t(ny * W + nx)) : NO_OPERAND;
    } else {
      rec->gb = pelem(s, i, idx);
    }
    return PULL_BOUZIDI;
  }
  if (!inY) {
    const int mode = sy < 0 ? p.bottom : p.top;
    if (mode == 2) {
      const int rx = std::min(std::max(sx, 0), W - 1);
      *g = pelem(s, MIRROR_Y2[i], uint32_t(y) * uint32_t(W) + uint32_t(rx));
      return PULL_GATHER;
    }
    *g = pelem(s, j, idx);
    return mode == 3 ? PULL_MARKED : PULL_GATHER;
  }
  *g = pelem(s, j, idx);
  return PULL_GATHER;
}

// The boundary list and its gather tables from the host copy of the flag words.
int build_links(lbm_solver* s, std::vector<uint32_t>& fl) {
  const uint32_t N = s->N;
  const bool d2 = s->nd == 2;
  const uint32_t solid = d2 ? FLAG_SOLID2 : FLAG_SOLID3;
  const uint32_t inlet_flag = d2 ? FLAG_INLET2 : FLAG_INLET3;
  const uint32_t outlet_flag = d2 ? FLAG_OUTLET2 : FLAG_OUTLET3;
  const uint32_t plain_in = d2 ? (FLAG_INLET2 | INLET_DIRS2) : (FLAG_INLET3 | INLET_DIRS3);
  const uint32_t plain_out = d2 ? (FLAG_OUTLET2 | OUTLET_DIRS2) : (FLAG_OUTLET3 | OUTLET_DIRS3);
  // Boundary list: every non-solid node the bulk rows of the step skip. Nodes without body links come first,
  // grouped inlet / outlet / other, then nodes with a body link (Bouzidi, or in 3D a slip face whose
  // mirror node is solid); cell order inside each group. Every pull is resolved here once.
  const int Q = s->Q;
  std::vector<uint32_t> grp[3], body;
  bool list_marks = false;
  int64_t n_bulk = 0, n_solid = 0;
  auto resolve = [&](uint32_t idx, int i, uint32_t* g, BzLink* rec) {
    return d2 ? resolve2(s, fl.data(), idx, i, g, rec) : resolve3(s, fl.data(), idx, i, g, rec);
  };
  for (uint32_t k = 0; k < N; k++) {
    const uint32_t v = fl[k];
    if (v & solid) {
      n_solid++;
      continue;
    }
    if (d2) {
      if (v == 0 || v == plain_in || v == plain_out) {
        n_bulk++;
        continue;
      }
    } else {
      // A node the bulk rows would take stays there only if every pull it makes under its row's
      // pattern is the one the reference's rules give; otherwise it is marked for the list.
      const uint32_t WH = uint32_t(s->W) * uint32_t(s->H);
      const Row3 r = row3(s, (k % WH) / uint32_t(s->W), k / WH);
      // (With Zou-He faces the inlet node pulls too; the directions entering from x = -1 are rebuilt,
      // so only the others are compared.)
      const int bk = bulk_kind3(r, v);
      bool ok = bk != 0;
      const bool check = bk == 1 || (bk == 2 && s->faces != LBM_FACES_EQUILIBRIUM);
      for (int i = 1; i < Q && check && ok; i++) {
        if (bk == 2 && CX3[i] > 0) continue;
        uint32_t gi;
        BzLink rec;
        ok = resolve(k, i, &gi, &rec) == PULL_GATHER && gi == row_gather3(s, r, k, (v & outlet_flag) != 0u, i);
      }
      if (ok) {
        n_bulk++;
        continue;
      }
      if (bk != 0) {
        fl[k] |= FLAG_LIST3;
        list_marks = true;
      }
    }
    const int kind = (v & inlet_flag) ? 0 : (v & outlet_flag) ? 1 : 2;
    bool has_body = false;
    const bool pulls = d2 || kind != 0 || s->faces != LBM_FACES_EQUILIBRIUM;  // the 3D equilibrium inlet takes none
    if (pulls) {
      for (int i = 1; i < Q && !has_body; i++) {
        uint32_t gi;
        BzLink rec;
        const int pk = resolve(k, i, &gi, &rec);
        has_body = pk == PULL_BOUZIDI || (!d2 && pk == PULL_MARKED);
      }
    }
    if (has_body) body.push_back(k);
    else grp[kind].push_back(k);
  }
  std::vector<uint32_t> lst;
  for (auto& l : grp) lst.insert(lst.end(), l.begin(), l.end());
  const uint32_t n_link = uint32_t(lst.size());
  lst.insert(lst.end(), body.begin(), body.end());
  const uint32_t nl = uint32_t(lst.size());
  std::vector<uint32_t> gt(size_t(Q - 1) * nl, 0u), bm(nl, 0u), xm(nl, 0u), lb(nl, 0u);
  std::vector<BzLink> recs;
  std::vector<uint32_t> rec_node;
  std::vector<uint8_t> rec_dir;
  for (uint32_t k = 0; k < nl; k++) {
    const uint32_t idx = lst[k];
    lb[k] = uint32_t(recs.size());
    if (!d2 && (fl[idx] & inlet_flag) && s->faces == LBM_FACES_EQUILIBRIUM) {
      for (int i = 1; i < Q; i++) gt[size_t(i - 1) * nl + k] = idx;  // never read
      continue;
    }
    for (int i = 1; i < Q; i++) {
      uint32_t gi = 0;
      BzLink rec{0.0f, NO_OPERAND};
      const int pk = resolve(idx, i, &gi, &rec);
      gt[size_t(i - 1) * nl + k] = gi;
      if (pk == PULL_BOUZIDI) {
        bm[k] |= 1u << i;
        recs.push_back(rec);
        rec_node.push_back(idx);
        rec_dir.push_back(uint8_t(i));
      } else if (pk == PULL_MARKED) {
        xm[k] |= 1u << i;
      }
    }
  }
  // Every index must address a population of the grid before anything is launched with it.
  const uint64_t lim = pop_elems(s);
  for (uint32_t v : gt)
    if (v >= lim) return lbm::fail(LBM_ERR_DEVICE, "boundary list: gather index out of range");
  for (const BzLink& r : recs)
    if (r.gb != NO_OPERAND && r.gb >= lim) return lbm::fail(LBM_ERR_DEVICE, "boundary list: operand out of range");
  // The new device tables are built beside the old ones and swapped in only once every allocation
  // and copy succeeded, so a failure leaves the handle's tables and counts as they were.
  struct Tables {
    uint32_t *gtab = nullptr, *bmask = nullptr, *xmask = nullptr, *lbase = nullptr;
    BzLink* links = nullptr;
    void release() {
      for (void* p : {(void*)gtab, (void*)bmask, (void*)xmask, (void*)lbase, (void*)links})
        if (p) (void)hipFree(p);
      *this = Tables{};
    }
  } fresh;
  auto upload = [&]() -> int {
    if (s->debug_fail_rebuilds > 0) {
      s->debug_fail_rebuilds--;
      return lbm::fail(LBM_ERR_ALLOC, "boundary tables: device allocation failed (test hook)");
    }
    if (!nl) return LBM_OK;
    HIP_TRY(hipMalloc(&fresh.gtab, sizeof(uint32_t) * gt.size()));
    HIP_TRY(hipMalloc(&fresh.bmask, sizeof(uint32_t) * nl));
    HIP_TRY(hipMalloc(&fresh.xmask, sizeof(uint32_t) * nl));
    HIP_TRY(hipMalloc(&fresh.lbase, sizeof(uint32_t) * nl));
    HIP_TRY(hipMalloc(&fresh.links, sizeof(BzLink) * std::max<size_t>(1, recs.size())));
    HIP_TRY(hipMemcpyAsync(fresh.gtab, gt.data(), sizeof(uint32_t) * gt.size(), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(fresh.bmask, bm.data(), sizeof(uint32_t) * nl, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(fresh.xmask, xm.data(), sizeof(uint32_t) * nl, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(fresh.lbase, lb.data(), sizeof(uint32_t) * nl, hipMemcpyHostToDevice, s->stream));
    if (!recs.empty())
      HIP_TRY(hipMemcpyAsync(fresh.links, recs.data(), sizeof(BzLink) * recs.size(), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return LBM_OK;
  };
  if (const int rc = upload()) {
    fresh.release();
    return rc;
  }
  // Swap: steps enqueued before this call may still read the old tables (and the list), so they
  // finish before the old buffers go and the list is overwritten.
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (void* p : {(void*)s->gtab, (void*)s->bmask, (void*)s->xmask, (void*)s->lbase, (void*)s->links})
    if (p) (void)hipFree(p);
  s->gtab = fresh.gtab;
  s->bmask = fresh.bmask;
  s->xmask = fresh.xmask;
  s->lbase = fresh.lbase;
  s->links = fresh.links;
  s->h_flags = fl;
  s->h_list = lst;
  s->h_gtab = gt;
  s->h_bmask = bm;
  s->h_xmask = xm;
  s->h_lbase = lb;
  s->h_links = recs;
  s->h_link_node = rec_node;
  s->h_link_dir = rec_dir;
  s->tables_gen++;
  s->nlist = nl;
  s->n_link = n_link;
  s->n_body = nl - n_link;
  s->n_links = uint32_t(recs.size());
  s->n_bulk = n_bulk;
  s->n_solid = n_solid;
  if (nl) HIP_TRY(hipMemcpyAsync(s->list, lst.data(), sizeof(uint32_t) * nl, hipMemcpyHostToDevice, s->stream));
  if (list_marks) HIP_TRY(hipMemcpyAsync(s->flags, fl.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice, s->stream));
  if (!d2) {
    // rows whose flag words are exactly what their position implies: inlet at x = 0, plain fluid
    // inside, the outlet at x = W - 1 (a bulk node off the faces, a list node on them)
    const uint32_t W = uint32_t(s->W), H = uint32_t(s->H), D = uint32_t(s->D);
    std::vector<uint8_t> clean(size_t(H) * D, 0);
    for (uint32_t z = 0; z < D; z++) {
      for (uint32_t y = 0; y < H; y++) {
        const Row3 r = row3(s, y, z);
        const bool face = r.loY || r.hiY || r.loZ || r.hiZ;
        const uint32_t* row = fl.data() + size_t(W) * (y + size_t(H) * z);
        bool ok = W >= 3 && row[0] == (FLAG_INLET3 | INLET_DIRS3 | r.mask);
        for (uint32_t x = 1; x + 1 < W && ok; x++) ok = row[x] == r.mask;
        if (ok) {
          const uint32_t last = row[W - 1];
          ok = face ? (bulk_kind3(r, last) == 0 && !(last & FLAG_SOLID3)) : last == (FLAG_OUTLET3 | OUTLET_DIRS3);
        }
        clean[y + size_t(H) * z] = ok ? 1 : 0;
      }
    }
    HIP_TRY(hipMemcpyAsync(s->row_clean, clean.data(), clean.size(), hipMemcpyHostToDevice, s->stream));
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  return LBM_OK;
}


// rebuildFlags (src/solver.ts:205-217, src/solver3d.ts:216-228): macro on the current buffer, flags
// and refill on the GPU, then the boundary list (and 2D force slots) from the flags, in cell order.
int rebuild_core(lbm_solver* s) {
  const uint32_t N = s->N;
  if (s->nd == 2) {
    launch_macro2(static_cast<const float*>(s->f[s->parity]), s->flags, s->mac, s->p2, s->stream);
    launch_flags2(s->sdf, s->flags, s->touches, static_cast<float*>(s->f[0]), static_cast<float*>(s->f[1]), s->mac,
                  s->p2, s->stream);
  } else if (s->Q == 27) {
    launch_macro27(s->prec, s->f[s->parity], s->flags, s->mac, s->p3, s->stream);
    launch_flags27(s->prec, s->sdf, s->flags, s->f[0], s->f[1], s->mac, s->p3, s->stream);
  } else {
    launch_macro3(s->prec, s->f[s->parity], s->flags, s->mac, s->p3, s->stream);
    launch_flags3(s->prec, s->sdf, s->flags, s->f[0], s->f[1], s->mac, s->p3, s->stream);
  }
  HIP_TRY(hipGetLastError());
  std::vector<uint32_t> fl(N);
  HIP_TRY(hipMemcpyAsync(fl.data(), s->flags, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, s->stream));
  std::vector<uint8_t> touch;
  if (s->nd == 2) {
    touch.resize(N);
    HIP_TRY(hipMemcpyAsync(touch.data(), s->touches, N, hipMemcpyDeviceToHost, s->stream));
  }
  HIP_TRY(hipStreamSynchronize(s->stream));

  if (s->Q == 27) {  // the D3Q27 step resolves its boundary rules in place: only the node counts
    s->n_bulk = s->n_solid = s->nlist = 0;
    const uint32_t W = uint32_t(s->W), bx = step27_block_x(s->W);
    const uint32_t gx = (W + bx - 1) / bx;
    std::vector<uint32_t> waves;  // ascending wave indices holding a FLAG_SLOW27 node
    for (uint32_t k = 0; k < N; k++) {
      if (fl[k] & FLAG_SOLID27) {
        s->n_solid++;
      } else if (fl[k] & FLAG_SLOW27) {
        s->nlist++;
        const uint32_t x = k % W, row = k / W;  // row = z H + y
        const uint32_t wave = (row * gx + x / bx) * (bx / 64u) + (x % bx) / 64u;
        if (waves.empty() || waves.back() != wave) waves.push_back(wave);
      } else {
        s->n_bulk++;
      }
    }
    if (!waves.empty())
      HIP_TRY(hipMemcpyAsync(s->list, waves.data(), 4 * waves.size(), hipMemcpyHostToDevice, s->stream));
    s->n_link = uint32_t(waves.size());
    // rows without a solid or FLAG_SLOW27 node: the step derives their flag words from x alone
    std::vector<uint8_t> clean(N / W, 1);
    for (uint32_t k = 0; k < N; k++)
      if (fl[k] & (FLAG_SOLID27 | FLAG_SLOW27)) clean[k / W] = 0;
    HIP_TRY(hipMemcpyAsync(s->row_clean, clean.data(), clean.size(), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->tables_gen++;
    return LBM_OK;
  }
  if (s->nd == 2) {
    // Force slots in cell order (the reference's atomicAdd order is arbitrary, src/shaders/aux.ts:63-66).
    uint32_t slot = 0;
    bool changed = false;
    for (uint32_t k = 0; k < N; k++) {
      if (!touch[k]) continue;
      ++slot;
      if (slot < (1u << (32 - BIDX_SHIFT2))) {
        fl[k] |= slot << BIDX_SHIFT2;
        changed = true;
      }
    }
    if (changed) HIP_TRY(hipMemcpyAsync(s->flags, fl.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice, s->stream));
  }
  return build_links(s, fl);
}

// Once the flag kernel ran, the device flags belong to the new geometry: a rebuild that fails after
// it leaves them out of step with the boundary tables, so the handle refuses to step until a later
// rebuild succeeds (check_usable).
int tables_done(lbm_solver* s, int rc) {
  if (rc == LBM_OK) s->invalid.clear();
  else s->invalid = g_err;
  return rc;
}

int rebuild_flags_impl(lbm_solver* s) { return tables_done(s, rebuild_core(s)); }

// Copies the Q populations of every cell from layout (na0, ns0) to layout (na1, ns1), element words of
// W bytes (the codecs' bits are moved, not converted). One thread per cell.
template <class Word>
__global__ __launch_bounds__(256) void relayout_kernel(const Word* __restrict__ src, Word* __restrict__ dst, uint32_t N,
                                                       int Q, uint32_t na0, uint32_t ns0, uint32_t na1, uint32_t ns1) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  for (int i = 0; i < Q; i++) dst[pelem(na1, ns1, i, c)] = src[pelem(na0, ns0, i, c)];
}

int launch_relayout(const lbm_solver* s, const void* src, void* dst, uint32_t na0, uint32_t ns0, uint32_t na1,
                    uint32_t ns1, hipStream_t st) {
  const uint32_t grid = (s->N + 255u) / 256u;
  if (s->elem == 2)
    relayout_kernel<uint16_t><<<grid, 256, 0, st>>>(static_cast<const uint16_t*>(src), static_cast<uint16_t*>(dst), s->N,
                                                    s->Q, na0, ns0, na1, ns1);
  else
    relayout_kernel<uint32_t><<<grid, 256, 0, st>>>(static_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst), s->N,
                                                    s->Q, na0, ns0, na1, ns1);
  HIP_TRY(hipGetLastError());
  return LBM_OK;
}

// Moves both population buffers to layout (na, ns) (new buffers, their unused elements zeroed) and
// re-encodes the boundary tables' element indices. Which nodes the bulk rows take does not depend
// on the layout (a bulk node's pulls are the same (population, cell) pairs in any layout), so only
// the gather indices and Bouzidi second operands change: each is decoded to (population, cell) and
// encoded again. A failure leaves the buffers, the layout and the tables as they were.
int set_layout_impl(lbm_solver* s, uint32_t na, uint32_t ns) {
  if (na == s->NA && ns == s->NS) return LBM_OK;
  const uint32_t na0 = s->NA, ns0 = s->NS;
  auto recode = [&](uint32_t g) {
    int i;
    uint32_t c;
    pelem_decode(na0, ns0, s->Q, g, &i, &c);
    return pelem(na, ns, i, c);
  };
  std::vector<uint32_t> gt(s->h_gtab.size());
  for (size_t k = 0; k < gt.size(); k++) gt[k] = recode(s->h_gtab[k]);
  std::vector<BzLink> lk = s->h_links;
  for (BzLink& r : lk)
    if (r.gb != NO_OPERAND) r.gb = recode(r.gb);
  const size_t bytes = (na ? size_t(s->Q) * ((size_t(s->N) + 63u) / 64u * 64u) : size_t(s->Q) * ns) * s->elem;
  void* nf[2] = {nullptr, nullptr};
  uint32_t* ngt = nullptr;
  BzLink* nlk = nullptr;
  auto drop = [&]() {
    (void)hipStreamSynchronize(s->stream);
    for (void* p : {nf[0], nf[1], (void*)ngt, (void*)nlk})
      if (p) (void)hipFree(p);
  };
  const int rc = [&]() -> int {
    for (int b = 0; b < 2; b++) {
      HIP_TRY(hipMalloc(&nf[b], bytes));
      HIP_TRY(hipMemsetAsync(nf[b], 0, bytes, s->stream));
      if (int r = launch_relayout(s, s->f[b], nf[b], na0, ns0, na, ns, s->stream)) return r;
    }
    if (s->gtab) {
      HIP_TRY(hipMalloc(&ngt, sizeof(uint32_t) * std::max<size_t>(1, gt.size())));
      if (!gt.empty())
        HIP_TRY(hipMemcpyAsync(ngt, gt.data(), sizeof(uint32_t) * gt.size(), hipMemcpyHostToDevice, s->stream));
    }
    if (s->links) {
      HIP_TRY(hipMalloc(&nlk, sizeof(BzLink) * std::max<size_t>(1, lk.size())));
      if (!lk.empty())
        HIP_TRY(hipMemcpyAsync(nlk, lk.data(), sizeof(BzLink) * lk.size(), hipMemcpyHostToDevice, s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return LBM_OK;
  }();
  if (rc) {
    drop();
    return rc;
  }
  // the stream is idle: no step still reads the old buffers
  for (int b = 0; b < 2; b++) {
    (void)hipFree(s->f[b]);
    s->f[b] = nf[b];
  }
  if (s->gtab) {
    (void)hipFree(s->gtab);
    s->gtab = ngt;
  }
  if (s->links) {
    (void)hipFree(s->links);
    s->links = nlk;
  }
  s->h_gtab = std::move(gt);
  s->h_links = std::move(lk);
  s->tables_gen++;
  s->NS = ns;
  s->NA = na;
  if (s->nd == 2) {
    s->p2.NS = ns;
    s->p2.NA = na;
  } else {
    s->p3.NS = ns;
    s->p3.NA = na;
  }
  return LBM_OK;
}

// Whether a solver can take tiles of 64 cells: rows of a multiple of 64 cells (so a wave of the bulk
// rows is one tile) and byte offsets within a buffer that fit 32 bits.
bool tiles_allowed(const lbm_solver* s) {
  return s->W % 64 == 0 && uint64_t(s->Q) * s->N * s->elem < (uint64_t(1) << 32);
}

// Population layouts the autotuner times for a 3D solver, (NA, NS) with the unpadded one first. The
// layout decides which memory channels and Infinity Cache sets the streams of one node share:
// measured on the D3Q19 step (MI355X, the 192 x 96 x 96 and 256 x 128 x 128 tunnels and six other
// grids), N rounded up to 1024 elements plus 69,632 (17 x 4096) gave 256 x 128 x 64 FP32 67.0 -> 59.0 us
// and 256 x 128 x 128 FP16 72.7 -> 64.8 us, plus 4096 was the better pad at 512 x 128 x 128 and
// 256 x 256 x 128 (FP16 128.9 -> 124.2 us) where 69,632 lost 6%, and neither helped at 192 x 96 x 96
// (FP32 45.3 us at N, 45.6 and 50.5 padded); on the D3Q27 step 256 x 128 x 64 FP32 109.5 -> 98.1 us
// and 256 x 128 x 128 FP16 117.3 -> 109.9 (+ 69,632), nothing at 192 x 96 x 96. Tiles of 64 cells
// (D3Q19) keep a tile's 19 populations in one 64 x 19-element block, which a copy of the step's
// shifted streams moved 4-14% faster than one array per population. So the choice is timed per
// grid. None while an adjoint state exists (its checkpoints hold buffers in the current layout) and
// none for small grids (N < 65,536). D2Q9 takes a padded stride by hand (lbm_set_population_stride)
// but is not timed here: on the D2Q9 grids measured (2048 x 1024, 3072 x 1024, 4096 x 2048, the
// 64-tunnel sweep) no pad was faster than N.
std::vector<std::pair<uint32_t, uint32_t>> layout_candidates(const lbm_solver* s) {
  std::vector<std::pair<uint32_t, uint32_t>> c{{0u, s->N}};
  if (s->adj || s->N < 65536u) return c;
  // D2Q9 takes a padded stride or tiles by hand but is not timed over them: on the D2Q9 step no pad
  // was faster than N, and tiles moved it by -2% to +5% (4096 x 2048 112.1 -> 109.6 us, 2048 x 1024
  // 24.7 -> 25.9, the 64-tunnel sweep within 0.3%), about the size of the trials' own noise.
  if (s->nd == 2) return c;
  const uint64_t base = (uint64_t(s->N) + 1023u) / 1024u * 1024u;
  for (uint64_t pad : {uint64_t(4096), uint64_t(17 * 4096)})
    if (uint64_t(s->Q) * (base + pad) < (uint64_t(1) << 32)) c.push_back({0u, uint32_t(base + pad)});
  if (tiles_allowed(s)) c.push_back({uint32_t(s->Q - 1), 64u});
  return c;
}


int alloc_common(lbm_solver* s) {
  const size_t N = s->N;
  HIP_TRY(hipSetDevice(s->device));
  if (!s->stream) {
    HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    s->owns_stream = true;
  }
  const size_t fbytes = pop_elems(s) * s->elem;
  HIP_TRY(hipMalloc(&s->f[0], fbytes));
  HIP_TRY(hipMalloc(&s->f[1], fbytes));
  HIP_TRY(hipMalloc(&s->flags, 4 * N));
  if (s->nd == 2) HIP_TRY(hipMalloc(&s->touches, N));
  HIP_TRY(hipMalloc(&s->sdf, 4 * N));
  HIP_TRY(hipMalloc(&s->mac, 16 * N));
  HIP_TRY(hipMalloc(&s->list, 4 * N));
  size_t maxwaves = (N + 63) / 64 + 16;  // one force partial per wave of body-link nodes
  if (s->Q == 27) maxwaves = std::max(maxwaves, step27_partials(s->W, s->H, s->D));  // per wave of the grid
  HIP_TRY(hipMalloc(&s->partials, sizeof(double) * 3 * maxwaves));
  HIP_TRY(hipMalloc(&s->counter, 4));
  HIP_TRY(hipMalloc(&s->ring, sizeof(double) * 3 * HISTORY_LEN));
  HIP_TRY(hipMalloc(&s->state, 4 * STATE_WORDS));
  HIP_TRY(hipMalloc(&s->xtab, sizeof(float4) * size_t(s->W)));
  if (s->nd == 3) {
    HIP_TRY(hipMalloc(&s->row_clean, size_t(s->H) * size_t(s->D)));
    HIP_TRY(hipMemsetAsync(s->row_clean, 0, size_t(s->H) * size_t(s->D), s->stream));
  }
  HIP_TRY(hipMemsetAsync(s->f[0], 0, fbytes, s->stream));
  HIP_TRY(hipMemsetAsync(s->f[1], 0, fbytes, s->stream));
  HIP_TRY(hipMemsetAsync(s->flags, 0, 4 * N, s->stream));
  HIP_TRY(hipMemsetAsync(s->mac, 0, 16 * N, s->stream));
  HIP_TRY(hipMemsetAsync(s->counter, 0, 4, s->stream));
  HIP_TRY(hipMemsetAsync(s->state, 0, 4 * STATE_WORDS, s->stream));
  HIP_TRY(hipMemsetAsync(s->ring, 0, sizeof(double) * 3 * HISTORY_LEN, s->stream));
  std::vector<float> far(N, 1e6f);  // src/solver.ts:133, src/solver3d.ts:142 (FAR)
  s->sdf_host = far;
  HIP_TRY(hipMemcpyAsync(s->sdf, far.data(), 4 * N, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return LBM_OK;
}

int bulk_block_for(int wg) {
  if (wg <= 0) return 256;
  if (wg % 64 != 0 || wg > 1024) return -1;
  return wg;
}

int enqueue_one(lbm_solver* s) { return enqueue_step(s, nullptr, -1); }

}  // namespace

namespace lbm {

int rebuild_flags(lbm_solver* s) { return rebuild_flags_impl(s); }
int set_layout(lbm_solver* s, uint32_t na, uint32_t ns) { return set_layout_impl(s, na, ns); }
bool tiles_ok(const lbm_solver* s) { return tiles_allowed(s); }
int relayout(const lbm_solver* s, const void* src, void* dst, uint32_t na0, uint32_t ns0, uint32_t na1, uint32_t ns1,
             hipStream_t st) {
  return launch_relayout(s, src, dst, na0, ns0, na1, ns1, st);
}

int check_usable(const lbm_solver* s) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  if (!s->invalid.empty())
    return fail(LBM_ERR_INVALID, "the last geometry rebuild failed (" + s->invalid +
                                     "); set the SDF or rebuild the flags again before stepping");
  return LBM_OK;
}

int enqueue_step(lbm_solver* s, const Record* rec, int sample, double* partials, uint32_t* counter) {
  const int every = s->force_every();
  ForceSink fs{partials ? partials : s->partials, counter ? counter : s->counter, s->ring, s->state, 0,
               uint64_t(s->step + 1)};
  fs.sample = sample >= 0 ? sample : ((every > 0 && (s->step + 1) % every == 0) ? 1 : 0);
  void* src = s->f[s->parity];
  void* dst = s->f[s->parity ^ 1];
  if (s->nd == 2) {
    if (s->ramp.kind) s->p2.uIn = float(s->ramp.at(s->step));
    const ListTables lt{s->list, s->gtab, s->bmask, s->xmask, s->lbase, s->links, s->n_link, s->n_body};
    const int order = s->sweep & 3;  // row order, as in 3D below
    const bool descend = order == 1 ? (src > dst) : order == 2 ? (src < dst) : false;
    launch_step2(static_cast<const float*>(src), static_cast<float*>(dst), s->flags, lt, s->p2, s->xtab, fs,
                 s->bulk_block, descend, s->stream, rec);
  } else if (s->Q == 27) {
    if (s->ramp.kind) s->p3.uIn = float(s->ramp.at(s->step));
    float fin[27];
    inlet_feq27(s->p3.uIn, fin);
    const int order = s->sweep & 3;  // row order, as for D3Q19 below
    const bool descend = order == 1 ? (src > dst) : order == 2 ? (src < dst) : false;
    launch_step27(s->prec, src, dst, s->flags, s->row_clean, s->sdf, s->xtab, s->p3, fin, fs, s->list, s->n_link,
                  descend, s->stream);
  } else {
    if (s->ramp.kind) s->p3.uIn = float(s->ramp.at(s->step));
    float fin[19];
    inlet_feq3(s->p3.uIn, fin);
    const ListTables lt{s->list, s->gtab, s->bmask, s->xmask, s->lbase, s->links, s->n_link, s->n_body};
    s->p3.seed = lowbias32(uint32_t(s->step));  // the stochastic codec's seed for this step
    // row order of this step: which buffer is walked top-down matters (measured), so both
    // alternations are candidates of the autotuner
    const int order = s->sweep & 3;
    const bool descend = order == 1 ? (src > dst) : order == 2 ? (src < dst) : false;
    launch_step3(s->prec, src, dst, s->flags, lt, s->p3, fin, s->xtab, fs, s->bulk_block, descend,
                 (s->sweep & SWEEP_MERGE) != 0, (s->sweep & SWEEP_EARLY) != 0, s->row_clean,
                 s->stream, rec);
  }
  s->parity ^= 1;
  s->step++;
  if (fs.sample) s->push_sample_step(s->step);
  return LBM_OK;
}

size_t force_partials_needed(const lbm_solver* s) {
  if (s->Q == 27) return step27_partials(s->W, s->H, s->D);
  const uint32_t bx = bulk_block_x(s->W, s->bulk_block);
  return size_t(body_list_blocks(s->n_body, bx, true)) * (bx / 64u);
}

// Advances the host bookkeeping of a tunnel by n steps that a batch launch performed.
void record_steps(lbm_solver* s, int64_t n) {
  const int every = s->force_every();
  std::lock_guard<std::mutex> lk(s->mu);
  for (int64_t k = 0; k < n; k++) {
    s->step++;
    if (every > 0 && s->step % every == 0) s->push_sample_step(s->step);
  }
  s->parity ^= int(n & 1);
}

}  // namespace lbm

extern "C" {

int lbm_abi_version(void) { return LBM_ABI_VERSION; }
const char* lbm_last_error(void) { return g_err.c_str(); }

int lbm_device_name(int device, char* out, size_t cap) {
  if (!out || cap == 0) return fail(LBM_ERR_INVALID, "lbm_device_name needs a buffer");
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (prop.name[0])
    std::snprintf(out, cap, "AMD %s %s (%d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  else
    std::snprintf(out, cap, "AMD %s (%d CUs)", prop.gcnArchName, prop.multiProcessorCount);
  return LBM_OK;
}

void lbm_config2d_defaults(lbm_config2d* c) {
  std::memset(c, 0, sizeof(*c));
  c->lambda = MAGIC_LAMBDA;
  c->sponge_tau = 1.0;
  c->workgroup_size = 128;
  c->force_every = 1;
}

void lbm_config3d_defaults(lbm_config3d* c) {
  std::memset(c, 0, sizeof(*c));
  c->lambda = std::nan("");
  c->sponge_tau = 1.0;
  c->workgroup_size = 128;
  c->force_every = 1;
  c->velocity_set = 19;
}

int lbm_create_2d(const lbm_config2d* cfg, int device, lbm_solver** out) {
  return lbm::create_2d_on(cfg, device, nullptr, out);
}

}  // extern "C"

int lbm::create_2d_on(const lbm_config2d* cfg, int device, hipStream_t stream, lbm_solver** out) {
  if (!cfg || !out) return fail(LBM_ERR_INVALID, "null argument");
  *out = nullptr;
  if (cfg->width < 1 || cfg->height < 1) return fail(LBM_ERR_INVALID, "width and height must be positive");
  if (cfg->height > 65535) return fail(LBM_ERR_INVALID, "height must be at most 65535 (one grid row per launch row)");
  const uint64_t N = uint64_t(cfg->width) * uint64_t(cfg->height);
  if (9 * N >= (1ull << 32)) return fail(LBM_ERR_INVALID, "grid too large: 9 * N must stay below 2^32");
  const int bb = bulk_block_for(cfg->workgroup_size);
  if (bb < 0) return fail(LBM_ERR_INVALID, "workgroup_size must be a multiple of 64 up to 1024");
  lbm_solver* s = new lbm_solver();
  s->nd = 2;
  s->device = device;
  if (stream) {
    s->stream = stream;
    s->owns_stream = false;
  }
  s->W = cfg->width;
  s->H = cfg->height;
  s->N = uint32_t(N);
  s->NS = s->N;
  s->Q = 9;
  s->elem = 4;
  s->cfg2 = *cfg;
  s->bulk_block = bb;
  int rc = alloc_common(s);
  if (rc == LBM_OK) {
    write_params2(s);
    rc = rebuild_flags_impl(s);
  }
  if (rc != LBM_OK) {
    std::string msg = g_err;
    free_all(s);
    delete s;
    return fail(rc == LBM_ERR_DEVICE ? LBM_ERR_ALLOC : rc, "Solver allocation failed: " + msg);
  }
  *out = s;
  return LBM_OK;
}

extern "C" {

int lbm_create_3d(const lbm_config3d* cfg, int device, lbm_solver** out) {
  if (!cfg || !out) return fail(LBM_ERR_INVALID, "null argument");
  *out = nullptr;
  if (cfg->precision < LBM_FP32 || cfg->precision > LBM_FP16_SR) return fail(LBM_ERR_UNSUPPORTED, "unknown precision");
  if (cfg->width < 3 || cfg->height < 1 || cfg->depth < 1)
    return fail(LBM_ERR_INVALID, "width must be at least 3, height and depth positive");
  if (cfg->height > 65535 || cfg->depth > 65535)
    return fail(LBM_ERR_INVALID, "height and depth must be at most 65535 (one grid row per launch row)");
  const int Q = cfg->velocity_set == 27 ? 27 : 19;
  if (cfg->velocity_set != 0 && cfg->velocity_set != 19 && cfg->velocity_set != 27)
    return fail(LBM_ERR_UNSUPPORTED, "velocity_set must be 19 or 27");
  if (Q == 27 && cfg->precision == LBM_FP16_SR)
    return fail(LBM_ERR_UNSUPPORTED, "D3Q27: precision fp32, fp16 or bf16");
  const uint64_t N = uint64_t(cfg->width) * uint64_t(cfg->height) * uint64_t(cfg->depth);
  const size_t elem = cfg->precision == LBM_FP32 ? 4 : 2;
  if (Q * N >= (1ull << 32)) {
    return fail(LBM_ERR_ALLOC, "Solver3D allocation failed: one distribution buffer needs " +
                                   std::to_string(Q * elem * N) + " bytes; this device allows " +
                                   std::to_string((1ull << 32) / Q * elem) + " per storage binding and " +
                                   std::to_string((1ull << 32) / Q * elem) + " per buffer");
  }
  const int bb = bulk_block_for(cfg->workgroup_size);
  if (bb < 0) return fail(LBM_ERR_INVALID, "workgroup_size must be a multiple of 64 up to 1024");
  lbm_solver* s = new lbm_solver();
  s->nd = 3;
  s->device = device;
  s->W = cfg->width;
  s->H = cfg->height;
  s->D = cfg->depth;
  s->N = uint32_t(N);
  s->NS = s->N;
  s->Q = Q;
  s->prec = cfg->precision;
  s->elem = elem;
  s->cfg3 = *cfg;
  s->bulk_block = bb;
  int rc = alloc_common(s);
  if (rc == LBM_OK) {
    write_params3(s);
    rc = rebuild_flags_impl(s);
  }
  if (rc != LBM_OK) {
    std::string msg = g_err;
    free_all(s);
    delete s;
    return fail(rc == LBM_ERR_DEVICE ? LBM_ERR_ALLOC : rc, "Solver3D allocation failed: " + msg);
  }
  *out = s;
  return LBM_OK;
}

void lbm_destroy(lbm_solver* s) {
  if (!s) return;
  free_all(s);
  delete s;
}

int lbm_update_2d(lbm_solver* s, const lbm_config2d* cfg) {
  if (!s || !cfg || s->nd != 2) return fail(LBM_ERR_INVALID, "lbm_update_2d needs a 2D solver and a config");
  if (cfg->width != s->W || cfg->height != s->H) return fail(LBM_ERR_INVALID, "width and height cannot change");
  const int bb = bulk_block_for(cfg->workgroup_size);
  if (bb < 0) return fail(LBM_ERR_INVALID, "workgroup_size must be a multiple of 64 up to 1024");
  const bool modes_changed = cfg->left != s->cfg2.left || cfg->right != s->cfg2.right ||
                             cfg->bottom != s->cfg2.bottom || cfg->top != s->cfg2.top;
  s->cfg2 = *cfg;
  s->bulk_block = bb;
  write_params2(s);
  if (modes_changed) {
    // The reference reads the modes in pullSlow every step with the flags it has; re-resolve the
    // gather tables against the current flags so the same rules apply.
    std::vector<uint32_t> fl(s->N);
    HIP_TRY(hipMemcpyAsync(fl.data(), s->flags, sizeof(uint32_t) * s->N, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return tables_done(s, build_links(s, fl));
  }
  return LBM_OK;
}

int lbm_update_3d(lbm_solver* s, const lbm_config3d* cfg) {
  if (!s || !cfg || s->nd != 3) return fail(LBM_ERR_INVALID, "lbm_update_3d needs a 3D solver and a config");
  if (cfg->width != s->W || cfg->height != s->H || cfg->depth != s->D || cfg->precision != s->prec ||
      (cfg->velocity_set == 27 ? 27 : 19) != s->Q)
    return fail(LBM_ERR_INVALID, "grid size, precision and velocity set cannot change");
  const int bb = bulk_block_for(cfg->workgroup_size);
  if (bb < 0) return fail(LBM_ERR_INVALID, "workgroup_size must be a multiple of 64 up to 1024");
  s->cfg3 = *cfg;
  s->bulk_block = bb;
  write_params3(s);
  return LBM_OK;
}

int lbm_set_inlet_ramp(lbm_solver* s, int kind, double u_target, double steps) {
  if (!s || kind < 0 || kind > 2) return fail(LBM_ERR_INVALID, "bad ramp");
  s->ramp.kind = kind;
  s->ramp.u_target = u_target;
  s->ramp.steps = steps;
  if (!kind) {
    if (s->nd == 2) write_params2(s);
    else write_params3(s);
  }
  return LBM_OK;
}

int lbm_set_sdf(lbm_solver* s, const float* sdf, size_t n) {
  if (!s || !sdf || n != s->N) return fail(LBM_ERR_INVALID, "sdf must hold N floats");
  s->sdf_host.assign(sdf, sdf + s->N);
  HIP_TRY(hipMemcpyAsync(s->sdf, sdf, 4 * size_t(s->N), hipMemcpyHostToDevice, s->stream));
  return rebuild_flags_impl(s);
}

int lbm_rebuild_flags(lbm_solver* s) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  return rebuild_flags_impl(s);
}

int lbm_set_open_faces(lbm_solver* s, int kind) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  if (s->nd != 3) return fail(LBM_ERR_INVALID, "open faces: 3D solvers only (the 2D inlet and outlet are Zou-He already)");
  if (s->Q == 27 && kind != LBM_FACES_EQUILIBRIUM) return fail(LBM_ERR_UNSUPPORTED, "open faces: Zou-He faces are D3Q19 only");
  if (kind != LBM_FACES_EQUILIBRIUM && kind != LBM_FACES_ZOU_HE) return fail(LBM_ERR_INVALID, "open faces: unknown kind");
  if (kind == s->faces) return LBM_OK;
  const int before = s->faces;
  s->faces = kind;
  write_params3(s);
  // which inlet nodes the bulk pass can take, and the inlet list nodes' gathers, depend on the kind
  const int rc = rebuild_flags_impl(s);
  if (rc != LBM_OK) {
    // Back to the previous kind: the flag kernel has rewritten the device flags already, so its
    // tables are rebuilt too (the geometry is unchanged, so nothing is refilled). Should that fail
    // as well, the handle refuses to step until a rebuild succeeds.
    const std::string msg = g_err;
    s->faces = before;
    write_params3(s);
    const int rc2 = rebuild_flags_impl(s);
    return fail(rc, msg + (rc2 == LBM_OK ? "; the previous faces are kept" : "; rebuilding the previous faces failed too: " + g_err));
  }
  return LBM_OK;
}

int lbm_debug_fail_rebuilds(lbm_solver* s, int n) {
  if (!s || n < 0) return fail(LBM_ERR_INVALID, "bad arguments");
  s->debug_fail_rebuilds = n;
  return LBM_OK;
}

namespace {

// src/solver.ts:249-268 (the disc's box grown by 2 cells and clipped to the grid) and
// src/shaders/aux.ts:176-196 over that box: d = |(x, y) - (cx, cy)| - r in float, with cx, cy and r
// rounded to float; union sdf = min(sdf, d), erase sdf = max(sdf, -d). The stamp goes into the host
// copy of the SDF, which the boundary tables are resolved from; the rows it touched widen [ylo, yhi].
// Returns false for an empty box (the reference then skips the rebuild too).
bool stamp_disc(lbm_solver* s, double cx, double cy, double r, bool erase, int& ylo, int& yhi) {
#pragma clang fp contract(off)
  const double pad = 2;
  const double x0 = std::max(0.0, std::floor(cx - r - pad)), y0 = std::max(0.0, std::floor(cy - r - pad));
  const double x1 = std::min(double(s->W) - 1, std::ceil(cx + r + pad));
  const double y1 = std::min(double(s->H) - 1, std::ceil(cy + r + pad));
  if (x1 < x0 || y1 < y0) return false;
  const float fx = float(cx), fy = float(cy), fr = float(r);
  for (int y = int(y0); y <= int(y1); y++) {
    float* row = s->sdf_host.data() + size_t(y) * s->W;
    for (int x = int(x0); x <= int(x1); x++) {
      const float dx = float(x) - fx, dy = float(y) - fy;
      const float d = std::sqrt(dx * dx + dy * dy) - fr;
      row[x] = erase ? std::max(row[x], -d) : std::min(row[x], d);
    }
  }
  ylo = std::min(ylo, int(y0));
  yhi = std::max(yhi, int(y1));
  return true;
}

enum class Rebuild { Never, IfStamped, Always };

int brush_impl(lbm_solver* s, const double* xy, size_t n, double r, int erase, Rebuild rebuild) {
  if (!s || s->nd != 2) return fail(LBM_ERR_INVALID, "the brush needs a 2D solver (src/solver.ts:244-268)");
  if (n && !xy) return fail(LBM_ERR_INVALID, "null points");
  if (!std::isfinite(r)) return fail(LBM_ERR_INVALID, "brush radius must be finite");
  for (size_t k = 0; k < 2 * n; k++)
    if (!std::isfinite(xy[k])) return fail(LBM_ERR_INVALID, "brush points must be finite");
  int ylo = s->H, yhi = -1;
  bool any = false;
  for (size_t k = 0; k < n; k++) any |= stamp_disc(s, xy[2 * k], xy[2 * k + 1], r, erase != 0, ylo, yhi);
  if (any) {
    const size_t off = size_t(ylo) * s->W, len = size_t(yhi - ylo + 1) * s->W;
    HIP_TRY(hipMemcpyAsync(s->sdf + off, s->sdf_host.data() + off, 4 * len, hipMemcpyHostToDevice, s->stream));
  }
  if (rebuild == Rebuild::Always || (rebuild == Rebuild::IfStamped && any)) return rebuild_flags_impl(s);
  if (any) HIP_TRY(hipStreamSynchronize(s->stream));  // the host copy may change again before the copy ran
  return LBM_OK;
}

}  // namespace

int lbm_brush(lbm_solver* s, double cx, double cy, double r, int erase, int rebuild) {
  // brush() with an empty box returns before its rebuild (src/solver.ts:255)
  const double xy[2] = {cx, cy};
  return brush_impl(s, xy, 1, r, erase, rebuild ? Rebuild::IfStamped : Rebuild::Never);
}

int lbm_brush_stroke(lbm_solver* s, const double* xy, size_t n_points, double r, int erase) {
  return brush_impl(s, xy, n_points, r, erase, Rebuild::Always);
}

int lbm_init_field(lbm_solver* s, const float* macro, size_t n) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  if (s->nd == 2) {
    std::vector<float> rest;
    if (!macro) {
      rest.assign(size_t(s->N) * 4, 0.0f);
      for (uint32_t k = 0; k < s->N; k++) rest[size_t(k) * 4] = 1.0f;
      macro = rest.data();
    } else if (n != size_t(s->N) * 4) {
      return fail(LBM_ERR_INVALID, "macro must hold N * 4 floats");
    }
    HIP_TRY(hipMemcpyAsync(s->mac, macro, 16 * size_t(s->N), hipMemcpyHostToDevice, s->stream));
    launch_init2(static_cast<float*>(s->f[0]), static_cast<float*>(s->f[1]), s->mac, s->p2, s->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
  } else {
    if (macro) return fail(LBM_ERR_INVALID, "the 3D initField takes no macro field");
    if (s->Q == 27) launch_init27(s->prec, s->f[0], s->f[1], s->flags, s->p3, s->stream);
    else launch_init3(s->prec, s->f[0], s->f[1], s->flags, s->p3, s->stream);
    HIP_TRY(hipGetLastError());
  }
  s->parity = 0;
  s->step = 0;
  return lbm_reset_forces(s);
}

int lbm_reset_forces(lbm_solver* s) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  launch_reset_state(s->state, s->stream);
  HIP_TRY(hipGetLastError());
  std::lock_guard<std::mutex> lk(s->mu);
  s->samples_read = 0;
  s->samples_pushed = 0;
  s->generation++;
  return LBM_OK;
}

int lbm_enqueue_steps(lbm_solver* s, int64_t n) {
  if (!s || n < 0) return fail(LBM_ERR_INVALID, "bad arguments");
  if (int rc = check_usable(s)) return rc;
  for (int64_t k = 0; k < n; k++) enqueue_one(s);
  HIP_TRY(hipGetLastError());
  return LBM_OK;
}

int lbm_enqueue_macro(lbm_solver* s) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  if (s->nd == 2) launch_macro2(static_cast<const float*>(s->f[s->parity]), s->flags, s->mac, s->p2, s->stream);
  else if (s->Q == 27) launch_macro27(s->prec, s->f[s->parity], s->flags, s->mac, s->p3, s->stream);
  else launch_macro3(s->prec, s->f[s->parity], s->flags, s->mac, s->p3, s->stream);
  HIP_TRY(hipGetLastError());
  return LBM_OK;
}

int lbm_run(lbm_solver* s, int64_t n, int64_t chunk) {
  if (!s || n < 0) return fail(LBM_ERR_INVALID, "bad arguments");
  if (chunk <= 0) chunk = 500;
  for (int64_t done = 0; done < n;) {
    const int64_t k = std::min(chunk, n - done);
    int rc = lbm_enqueue_steps(s, k);
    if (rc) return rc;
    rc = lbm_enqueue_macro(s);
    if (rc) return rc;
    done += k;
  }
  return lbm_sync(s);
}

int lbm_sync(lbm_solver* s) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  HIP_TRY(hipStreamSynchronize(s->stream));
  return LBM_OK;
}

int lbm_read_macro(lbm_solver* s, float* out, size_t n) {
  if (!s || !out || n < size_t(s->N) * 4) return fail(LBM_ERR_INVALID, "out must hold N * 4 floats");
  HIP_TRY(hipMemcpyAsync(out, s->mac, 16 * size_t(s->N), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return LBM_OK;
}

int lbm_read_forces_begin(lbm_solver* s, int64_t* ticket) {
  if (!s || !ticket) return fail(LBM_ERR_INVALID, "bad arguments");
  Ticket t;
  const size_t rb = sizeof(double) * size_t(s->dim()) * HISTORY_LEN;
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&t.ring), rb, hipHostMallocDefault));
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&t.state), 4, hipHostMallocDefault));
  HIP_TRY(hipEventCreateWithFlags(&t.ev, hipEventDisableTiming));
  // The device ring is HISTORY_LEN * 3 doubles wide; copy dim components per sample.
  if (s->dim() == 3) {
    HIP_TRY(hipMemcpyAsync(t.ring, s->ring, rb, hipMemcpyDeviceToHost, s->stream));
  } else {
    HIP_TRY(hipMemcpyAsync(t.ring, s->ring, rb, hipMemcpyDeviceToHost, s->stream));
  }
  HIP_TRY(hipMemcpyAsync(t.state, s->state, 4, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipEventRecord(t.ev, s->stream));
  std::lock_guard<std::mutex> lk(s->mu);
  t.generation = s->generation;
  *ticket = s->next_ticket++;
  s->tickets[*ticket] = t;
  return LBM_OK;
}

int lbm_read_forces_wait(lbm_solver* s, int64_t ticket) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  hipEvent_t ev;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(LBM_ERR_INVALID, "unknown ticket");
    ev = it->second.ev;
  }
  HIP_TRY(hipEventSynchronize(ev));
  return LBM_OK;
}

int lbm_read_forces_end(lbm_solver* s, int64_t ticket, int64_t* steps, double* values, int64_t cap, int64_t* count) {
  if (!s || !count) return fail(LBM_ERR_INVALID, "bad arguments");
  *count = 0;
  Ticket t;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(LBM_ERR_INVALID, "unknown ticket");
    t = it->second;
    s->tickets.erase(it);
  }
  HIP_TRY(hipEventSynchronize(t.ev));
  const int dim = s->dim();
  int rc = LBM_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    // A reset while the copy was in flight makes these samples stale (src/solver.ts:322-323).
    if (t.generation == s->generation && s->samples_pushed > s->samples_read) {
      const int64_t written = int64_t(*t.state);
      const int64_t first = std::max(s->samples_read, written - int64_t(HISTORY_LEN));
      const int64_t avail = written - first;
      if (avail > cap || (avail > 0 && (!steps || !values))) {
        rc = fail(LBM_ERR_INVALID, "output capacity too small for " + std::to_string(avail) + " samples");
      } else {
        int64_t n = 0;
        for (int64_t k = first; k < written; k++, n++) {
          const int64_t slot = k % HISTORY_LEN;
          const bool kept = k < s->samples_pushed && k >= s->samples_pushed - int64_t(HISTORY_LEN);
          steps[n] = kept ? s->sample_steps[size_t(k % HISTORY_LEN)] : -1;
          for (int d = 0; d < dim; d++) values[n * dim + d] = t.ring[slot * dim + d];
        }
        *count = n;
        s->samples_read = written;
      }
    }
  }
  (void)hipHostFree(t.ring);
  (void)hipHostFree(t.state);
  (void)hipEventDestroy(t.ev);
  return rc;
}

int lbm_diverged(lbm_solver* s, int64_t* first_step) {
  if (!s || !first_step) return fail(LBM_ERR_INVALID, "bad arguments");
  uint64_t v = 0;
  HIP_TRY(hipMemcpyAsync(&v, s->state + 2, sizeof(v), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  *first_step = int64_t(v);
  return LBM_OK;
}

int lbm_read_forces(lbm_solver* s, int64_t* steps, double* values, int64_t cap, int64_t* count) {
  int64_t t = 0;
  int rc = lbm_read_forces_begin(s, &t);
  if (rc) return rc;
  rc = lbm_read_forces_wait(s, t);
  if (rc) return rc;
  return lbm_read_forces_end(s, t, steps, values, cap, count);
}

int lbm_info(const lbm_solver* s, int32_t* dims, int64_t* n_cells, int64_t* step, int32_t* lattice_q,
             int32_t* precision) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  if (dims) dims[0] = s->W, dims[1] = s->H, dims[2] = s->D;
  if (n_cells) *n_cells = s->N;
  if (step) *step = s->step;
  if (lattice_q) *lattice_q = s->Q;
  if (precision) *precision = s->prec;
  return LBM_OK;
}

int lbm_read_populations(lbm_solver* s, void* out, size_t bytes) {
  const size_t need = size_t(s ? s->Q : 0) * (s ? s->N : 0) * (s ? s->elem : 0);
  if (!s || !out || bytes < need) return fail(LBM_ERR_INVALID, "out too small");
  // Q x N, population-major, whatever the device layout: a padded stride is copied row by row, tiles
  // are copied whole and taken apart on the host
  if (s->NA) {
    std::vector<uint8_t> raw(pop_elems(s) * s->elem);
    HIP_TRY(hipMemcpyAsync(raw.data(), s->f[s->parity], raw.size(), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    // population i of tile t: 64 consecutive elements at t * 64 Q + 64 i (N is a multiple of 64 here)
    uint8_t* o = static_cast<uint8_t*>(out);
    const size_t run = 64u * s->elem;
    for (uint32_t t = 0; t < s->N / 64u; t++)
      for (int i = 0; i < s->Q; i++)
        std::memcpy(o + (size_t(i) * s->N + size_t(t) * 64u) * s->elem,
                    raw.data() + size_t(pelem(s, i, t * 64u)) * s->elem, run);
    return LBM_OK;
  } else if (s->NS == s->N) {
    HIP_TRY(hipMemcpyAsync(out, s->f[s->parity], need, hipMemcpyDeviceToHost, s->stream));
  } else {
    HIP_TRY(hipMemcpy2DAsync(out, size_t(s->N) * s->elem, s->f[s->parity], size_t(s->NS) * s->elem,
                             size_t(s->N) * s->elem, size_t(s->Q), hipMemcpyDeviceToHost, s->stream));
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  return LBM_OK;
}

int lbm_read_flags(lbm_solver* s, uint32_t* out, size_t n) {
  if (!s || !out || n < s->N) return fail(LBM_ERR_INVALID, "out must hold N words");
  HIP_TRY(hipMemcpyAsync(out, s->flags, 4 * size_t(s->N), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (s->nd == 3)
    for (uint32_t k = 0; k < s->N; k++) out[k] &= s->Q == 27 ? ~FLAG_SLOW27 : ~FLAG_LIST3;
  return LBM_OK;
}

int lbm_time_steps(lbm_solver* s, int64_t n, double* elapsed_ms) {
  if (!s || n < 0 || !elapsed_ms) return fail(LBM_ERR_INVALID, "bad arguments");
  if (int rc = check_usable(s)) return rc;
  hipEvent_t a, b;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  HIP_TRY(hipEventRecord(a, s->stream));
  for (int64_t k = 0; k < n; k++) enqueue_one(s);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(b, s->stream));
  HIP_TRY(hipEventSynchronize(b));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, a, b));
  *elapsed_ms = ms;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return LBM_OK;
}

int lbm_profile_steps(lbm_solver* s, int64_t n, double* total_ms, double* step_ms, double* max_step_ms) {
  if (!s || n < 0 || !total_ms || !step_ms || !max_step_ms) return fail(LBM_ERR_INVALID, "bad arguments");
  if (int rc = check_usable(s)) return rc;
  std::vector<hipEvent_t> ev(size_t(n + 1));
  for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
  HIP_TRY(hipEventRecord(ev[0], s->stream));
  for (int64_t k = 0; k < n; k++) {
    enqueue_one(s);
    HIP_TRY(hipEventRecord(ev[size_t(k + 1)], s->stream));
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventSynchronize(ev.back()));
  double sum = 0, mx = 0;
  for (int64_t k = 0; k < n; k++) {
    float a = 0;
    HIP_TRY(hipEventElapsedTime(&a, ev[size_t(k)], ev[size_t(k + 1)]));
    sum += a;
    mx = std::max(mx, double(a));
  }
  float tot = 0;
  HIP_TRY(hipEventElapsedTime(&tot, ev[0], ev.back()));
  for (auto& e : ev) (void)hipEventDestroy(e);
  *total_ms = tot;
  *step_ms = sum;
  *max_step_ms = mx;
  return LBM_OK;
}

// Autotuner for the step kernel's launch shape: each distinct bulk block width the launcher can use
// for this W with the three row orders (and, in 3D, two store variants and two pull orders; D3Q27: its
// one block width in the three row orders) is timed over
// trial_steps steps (after eight untimed ones; three interleaved rounds, the fastest round counting)
// on a copy of the solver's own state, the fastest is kept, and the
// populations, step, parity and force ring are restored. Results are cached per (lattice, velocity
// set, precision, W, H, D, open-face kind, subgrid model on/off, force cadence) for the life of the
// process. Block width and row order do not change a result (cells are independent within a step
// and the force sum is ordered by node, not by block); the store variants, and the pull orders of
// the shared-store variant, are separate compilations and may differ in the last bits.
int lbm_autotune(lbm_solver* s, int32_t trial_steps, int32_t* chosen) {
  if (!s || trial_steps <= 0) return fail(LBM_ERR_INVALID, "bad arguments");
  if (int rc = check_usable(s)) return rc;
  static std::mutex cache_mu;
  struct Pick {
    int block, sweep;
    std::pair<uint32_t, uint32_t> layout;  // (NA, NS)
  };
  static std::map<std::string, Pick> cache;
  // Everything that selects a different kernel instantiation or changes the work of a step is part
  // of the key: lattice and velocity set, precision, grid, open-face kind, subgrid model on/off,
  // force cadence.
  const bool smag = (s->nd == 2 ? s->cfg2.smagorinsky : s->cfg3.smagorinsky) > 0;
  const std::string key = std::to_string(s->nd) + ":Q" + std::to_string(s->Q) + ":" + std::to_string(s->prec) + ":" +
                          std::to_string(s->W) + "x" + std::to_string(s->H) + "x" + std::to_string(s->D) + ":faces" +
                          std::to_string(s->faces) + ":smag" + std::to_string(int(smag)) + ":every" +
                          std::to_string(s->force_every());
  {
    std::lock_guard<std::mutex> lk(cache_mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      const Pick pk = it->second;
      const auto layouts = layout_candidates(s);
      if (std::find(layouts.begin(), layouts.end(), pk.layout) != layouts.end())
        if (int rc = set_layout_impl(s, pk.layout.first, pk.layout.second)) return rc;
      s->bulk_block = pk.block;
      s->sweep = pk.sweep;
      if (chosen) *chosen = s->Q == 27 ? int32_t(step27_block_x(s->W)) : s->bulk_block;
      return LBM_OK;
    }
  }
  // Trials run on the solver's own buffers and sample forces as it is configured to (the sampling
  // changes which launch shape wins). The populations, the force ring and its counter are saved
  // first; the guard restores them (into the buffers in whatever population layout is current), the
  // step, parity and pending sample steps, and frees the copies, whether the trials finish or a HIP
  // call fails part way; after a failure it also moves the populations back to the layout they had.
  const size_t fbytes = pop_elems(s) * s->elem;
  const size_t ring_bytes = sizeof(double) * 3 * HISTORY_LEN;
  struct Guard {
    lbm_solver* s;
    size_t fbytes, ring_bytes;
    void* save[2] = {nullptr, nullptr};
    void* ring = nullptr;
    void* state = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    int64_t step0;
    int parity0, block0, sweep0;
    uint32_t na0, ns0;
    int64_t pushed0;
    std::vector<int64_t> steps0;
    bool saved = false;  // every copy was enqueued: from here on the state must be restored
    bool keep_shape = false;
    int restore() {
      saved = false;
      for (int b = 0; b < 2; b++) {
        if (s->NA == na0 && s->NS == ns0) {
          HIP_TRY(hipMemcpyAsync(s->f[b], save[b], fbytes, hipMemcpyDeviceToDevice, s->stream));
        } else if (int rc = launch_relayout(s, save[b], s->f[b], na0, ns0, s->NA, s->NS, s->stream)) {
          return rc;
        }
      }
      HIP_TRY(hipMemcpyAsync(s->ring, ring, ring_bytes, hipMemcpyDeviceToDevice, s->stream));
      HIP_TRY(hipMemcpyAsync(s->state, state, 4 * STATE_WORDS, hipMemcpyDeviceToDevice, s->stream));
      HIP_TRY(hipStreamSynchronize(s->stream));
      return LBM_OK;
    }
    ~Guard() {
      if (saved) (void)restore();
      if (!keep_shape && (s->NA != na0 || s->NS != ns0)) (void)set_layout_impl(s, na0, ns0);
      (void)hipStreamSynchronize(s->stream);
      for (void* p : {save[0], save[1], ring, state})
        if (p) (void)hipFree(p);
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
      s->step = step0;
      s->parity = parity0;
      s->samples_pushed = pushed0;
      s->sample_steps = steps0;
      if (!keep_shape) {
        s->bulk_block = block0;
        s->sweep = sweep0;
      }
    }
  } g{s, fbytes, ring_bytes};
  g.step0 = s->step;
  g.parity0 = s->parity;
  g.block0 = s->bulk_block;
  g.sweep0 = s->sweep;
  g.na0 = s->NA;
  g.ns0 = s->NS;
  g.pushed0 = s->samples_pushed;
  g.steps0 = s->sample_steps;
  HIP_TRY(hipMalloc(&g.save[0], fbytes));
  HIP_TRY(hipMalloc(&g.save[1], fbytes));
  HIP_TRY(hipMalloc(&g.ring, ring_bytes));
  HIP_TRY(hipMalloc(&g.state, 4 * STATE_WORDS));
  HIP_TRY(hipEventCreate(&g.a));
  HIP_TRY(hipEventCreate(&g.b));
  HIP_TRY(hipMemcpyAsync(g.save[0], s->f[0], fbytes, hipMemcpyDeviceToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(g.save[1], s->f[1], fbytes, hipMemcpyDeviceToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(g.ring, s->ring, ring_bytes, hipMemcpyDeviceToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(g.state, s->state, 4 * STATE_WORDS, hipMemcpyDeviceToDevice, s->stream));
  g.saved = true;
  std::vector<int> tried;
  int best = s->bulk_block, best_sweep = s->sweep;
  double best_ms = 1e300;
  // every candidate is timed three times, interleaved with the others, and keeps its fastest time
  std::vector<std::pair<int, int>> cand;
  if (s->Q == 27)  // one block width (whole rows of at most 256 threads); the three row orders
    for (int sw : {1, 2, 0}) cand.push_back({s->bulk_block, sw});
  for (int req = 64; req <= 1024 && s->Q != 27; req += 64) {
    const int bx = int(bulk_block_x(s->W, req));
    if (std::find(tried.begin(), tried.end(), bx) != tried.end()) continue;
    tried.push_back(bx);
    if (s->nd == 3) {
      for (int sw : {1, 2, 0})
        for (int mg : {0, SWEEP_MERGE})
          for (int ea : {0, SWEEP_EARLY})
            if (!(ea && smag)) cand.push_back({bx, sw | mg | ea});  // the subgrid-model kernels have one pull order
    } else {
      for (int sw : {1, 2, 0}) cand.push_back({bx, sw});  // D2Q9: the three row orders
    }
  }
  auto trial = [&](size_t c, double* out) -> int {
    s->bulk_block = cand[c].first;
    s->sweep = cand[c].second;
    // untimed steps first: the previous candidate's row order leaves the Infinity Cache holding
    // other rows than this one's steady state does
    for (int k = 0; k < 8; k++) enqueue_one(s);
    HIP_TRY(hipEventRecord(g.a, s->stream));
    for (int k = 0; k < trial_steps; k++) enqueue_one(s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(g.b, s->stream));
    HIP_TRY(hipEventSynchronize(g.b));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, g.a, g.b));
    *out = std::min(*out, double(ms));
    return LBM_OK;
  };
  std::vector<double> best_of(cand.size(), 1e300);
  for (int rep = 0; rep < 3; rep++)
    for (size_t c = 0; c < cand.size(); c++)
      if (int rc = trial(c, &best_of[c])) return rc;
  for (size_t c = 0; c < cand.size(); c++) {
    if (best_of[c] < best_ms) {
      best_ms = best_of[c];
      best = cand[c].first;
      best_sweep = cand[c].second;
    }
  }
  // The population layout (3D): the three fastest shapes again in every candidate layout, the
  // layouts visited twice in opposite orders; the fastest (layout, shape) pair of these runs wins.
  // A layout change leaves the new buffers, just written by the copy, in the Infinity Cache: each
  // visit first runs untimed steps until the step's own traffic has replaced that state.
  std::pair<uint32_t, uint32_t> best_layout{s->NA, s->NS};
  const auto strides = layout_candidates(s);
  if (strides.size() > 1) {
    std::vector<size_t> top(cand.size());
    for (size_t c = 0; c < cand.size(); c++) top[c] = c;
    std::sort(top.begin(), top.end(), [&](size_t a, size_t b) { return best_of[a] < best_of[b]; });
    top.resize(std::min<size_t>(3, top.size()));
    std::vector<double> t2(strides.size() * top.size(), 1e300);
    for (int rep = 0; rep < 2; rep++) {
      for (size_t k = 0; k < strides.size(); k++) {
        const size_t si = rep == 0 ? k : strides.size() - 1 - k;
        if (int rc = set_layout_impl(s, strides[si].first, strides[si].second)) return rc;
        s->bulk_block = cand[top[0]].first;
        s->sweep = cand[top[0]].second;
        for (int k = 0; k < 2 * trial_steps; k++) enqueue_one(s);
        for (size_t t = 0; t < top.size(); t++)
          if (int rc = trial(top[t], &t2[si * top.size() + t])) return rc;
      }
    }
    double b2 = 1e300;
    for (size_t si = 0; si < strides.size(); si++) {
      for (size_t t = 0; t < top.size(); t++) {
        if (t2[si * top.size() + t] < b2) {
          b2 = t2[si * top.size() + t];
          best_layout = strides[si];
          best = cand[top[t]].first;
          best_sweep = cand[top[t]].second;
        }
      }
    }
  }
  if (int rc = set_layout_impl(s, best_layout.first, best_layout.second)) return rc;
  if (int rc = g.restore()) return rc;
  s->bulk_block = best;
  s->sweep = best_sweep;
  g.keep_shape = true;
  {
    std::lock_guard<std::mutex> lk(cache_mu);
    cache[key] = {best, best_sweep, best_layout};
  }
  if (chosen) *chosen = s->Q == 27 ? int32_t(step27_block_x(s->W)) : best;
  return LBM_OK;
}

int lbm_launch_info(lbm_solver* s, int32_t* block, int32_t* sweep) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  if (block) *block = int32_t(s->Q == 27 ? step27_block_x(s->W) : bulk_block_x(s->W, s->bulk_block));
  if (sweep) *sweep = s->sweep;
  return LBM_OK;
}

int lbm_set_launch(lbm_solver* s, int32_t block, int32_t sweep) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  const int bb = bulk_block_for(block);
  if (bb < 0 || block <= 0) return fail(LBM_ERR_INVALID, "block must be a multiple of 64 up to 1024");
  if (s->nd == 3 ? (sweep < 0 || sweep > (3 | SWEEP_MERGE | SWEEP_EARLY) || (sweep & 3) == 3) : (sweep < 0 || sweep > 2))
    return fail(LBM_ERR_INVALID, "variant must be 0-2, 4-6, 8-10 or 12-14 (3D), 0-2 (2D)");
  if (s->Q == 27) {  // one block width; the row order is the only variant
    if (block != int32_t(step27_block_x(s->W)) || (sweep & ~3) != 0)
      return fail(LBM_ERR_UNSUPPORTED, "D3Q27: block must be " + std::to_string(step27_block_x(s->W)) +
                                           " for this width and the variant a row order (0-2)");
    s->sweep = sweep;
    return LBM_OK;
  }
  s->bulk_block = bb;
  s->sweep = sweep;
  return LBM_OK;
}

int lbm_population_stride(lbm_solver* s, int64_t* stride) {
  if (!s || !stride) return fail(LBM_ERR_INVALID, "bad arguments");
  *stride = s->NS;
  return LBM_OK;
}

int lbm_population_layout(lbm_solver* s, int32_t* tiled, int64_t* stride) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  if (tiled) *tiled = s->NA ? 1 : 0;
  if (stride) *stride = s->NS;
  return LBM_OK;
}

int lbm_set_population_layout(lbm_solver* s, int32_t tiled, int64_t stride) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  if (int rc = check_usable(s)) return rc;
  uint32_t na = 0, ns = 0;
  if (tiled) {
    if (!tiles_allowed(s))
      return fail(LBM_ERR_UNSUPPORTED, "population layout: tiles of 64 cells need a row width that is a multiple of 64 "
                                       "and Q x N x element bytes below 2^32");
    if (stride != 0 && stride != 64) return fail(LBM_ERR_INVALID, "population layout: tiles take stride 0 or 64");
    na = uint32_t(s->Q - 1);
    ns = 64u;
  } else {
    if (stride == 0) stride = s->N;
    if (stride != int64_t(s->N) &&
        (stride < int64_t(s->N) || stride % 64 != 0 || uint64_t(s->Q) * uint64_t(stride) >= (uint64_t(1) << 32)))
      return fail(LBM_ERR_INVALID, "population stride: a multiple of 64 elements, at least N, with Q x stride below 2^32");
    ns = uint32_t(stride);
  }
  if (na == s->NA && ns == s->NS) return LBM_OK;
  if (s->adj)
    return fail(LBM_ERR_UNSUPPORTED, "population layout: release the adjoint state first (its checkpoints hold the buffers)");
  return set_layout_impl(s, na, ns);
}

int lbm_set_population_stride(lbm_solver* s, int64_t stride) { return lbm_set_population_layout(s, 0, stride); }

int lbm_node_counts(lbm_solver* s, int64_t* bulk, int64_t* list, int64_t* solid) {
  if (!s) return fail(LBM_ERR_INVALID, "null solver");
  if (bulk) *bulk = s->n_bulk;
  if (list) *list = s->nlist;
  if (solid) *solid = s->n_solid;
  return LBM_OK;
}

}  // extern "C"
