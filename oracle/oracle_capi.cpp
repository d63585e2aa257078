// C entry points of the oracle (test infrastructure only: loaded by tests, the smoke check and
// the CPU-baseline leg of the benchmark, never by the product).
//
// Mirrors the shape of include/lbm.h so that a test can drive the HIP library and the oracle with
// the same calls. real_bits selects float (32: the reference's f32 arithmetic, storage per
// cfg->precision in 3D) or double (64: the FP64 accuracy reference, double storage).
#include <omp.h>

#include <cstring>
#include <memory>

#include "solver.hpp"

namespace {

struct Base {
  virtual ~Base() = default;
  virtual int q() const = 0;
  virtual int dim() const = 0;
  virtual void dims(int32_t* d, int64_t* n) const = 0;
  virtual int64_t step() const = 0;
  virtual void update2(const lbm_config2d*) {}
  virtual void update3(const lbm_config3d*) {}
  virtual void set_sdf(const float* s) = 0;
  virtual void brush(double, double, double, bool, bool) {}
  virtual void rebuild() = 0;
  virtual void read_sdf(float* out) = 0;
  virtual void init_field(const float* m) = 0;
  virtual void reset_forces() = 0;
  virtual void set_ramp(int kind, double u, double steps) = 0;
  virtual void steps(int64_t n) = 0;
  virtual void macro(double* out) = 0;
  virtual int64_t read_forces(int64_t* st, double* v, int64_t cap) = 0;
  virtual void read_populations(double* out) = 0;
  virtual size_t raw_bytes() const = 0;
  virtual void read_raw(void* out) = 0;
  virtual void read_flags(uint32_t* out) = 0;
  virtual void read_slots(uint32_t* out) = 0;
  virtual void params(double* out) const = 0;
  virtual void set_faces(int) {}
};

template <class Sv>
struct Impl : Base {
  Sv s;
  template <class C> explicit Impl(const C& c) : s(c) {}
  int64_t step() const override { return s.step; }
  void set_sdf(const float* sd) override {
    std::memcpy(s.sdf.data(), sd, sizeof(float) * size_t(s.N));
    s.rebuild_flags();
  }
  void rebuild() override { s.rebuild_flags(); }
  void read_sdf(float* out) override { std::memcpy(out, s.sdf.data(), sizeof(float) * size_t(s.N)); }
  void reset_forces() override { s.reset_forces(); }
  void set_ramp(int kind, double u, double steps) override {
    s.ramp.kind = kind;
    s.ramp.u_target = u;
    s.ramp.steps = steps;
    if (!kind) s.write_params();
  }
  void steps(int64_t n) override {
    for (int64_t k = 0; k < n; k++) s.one_step();
  }
  void macro(double* out) override {
    s.macro();
    for (size_t k = 0; k < s.mac.size(); k++) out[k] = double(s.mac[k]);
  }
  int64_t read_forces(int64_t* st, double* v, int64_t cap) override {
    return s.book.consume(s.history.data(), dim(), int64_t(s.state), st, v, cap);
  }
  void read_populations(double* out) override {
    const auto& b = s.f[s.parity];
    for (size_t k = 0; k < b.size(); k++) out[k] = double(s.ld(b, int64_t(k)));
  }
  size_t raw_bytes() const override { return s.f[0].size() * sizeof(typename Sv::E); }
  void read_raw(void* out) override { std::memcpy(out, s.f[s.parity].data(), raw_bytes()); }
  void read_flags(uint32_t* out) override { std::memcpy(out, s.flags.data(), sizeof(uint32_t) * size_t(s.N)); }
};

template <class T, class S>
struct Impl2 : Impl<orc::Solver2D<T, S>> {
  using Impl<orc::Solver2D<T, S>>::Impl;
  int q() const override { return 9; }
  int dim() const override { return 2; }
  void dims(int32_t* d, int64_t* n) const override {
    d[0] = this->s.W, d[1] = this->s.H, d[2] = 1;
    *n = this->s.N;
  }
  void update2(const lbm_config2d* c) override {
    this->s.cfg = *c;
    this->s.write_params();
  }
  void init_field(const float* m) override { this->s.init_field(m); }
  void brush(double cx, double cy, double r, bool erase, bool rebuild) override {
    this->s.brush(cx, cy, r, erase, rebuild);
  }
  void read_slots(uint32_t* out) override {
    for (int64_t k = 0; k < this->s.N; k++) out[k] = this->s.flags[size_t(k)] >> orc::BIDX_SHIFT2;
  }
  // Params uniform order of src/shaders/common.ts:56-62 (floats only): tau, lambda, smagC2, uIn,
  // spongeStart, tauSponge, uLid, fx, fy, absorb, inletLayer, then parabolic and forcing.
  void params(double* o) const override {
    const auto& P = this->s.P;
    double v[13] = {double(P.tau), double(P.lambda), double(P.smagC2), double(P.uIn), double(P.spongeStart),
                    double(P.tauSponge), double(P.uLid), double(P.fx), double(P.fy), double(P.absorb),
                    double(P.inletLayer), double(P.parabolic), double(P.forcing)};
    for (int k = 0; k < 13; k++) o[k] = v[k];
  }
};

template <class T, class S, int Q>
struct Impl3 : Impl<orc::Solver3D<T, S, Q>> {
  using Impl<orc::Solver3D<T, S, Q>>::Impl;
  int q() const override { return Q; }
  int dim() const override { return 3; }
  void dims(int32_t* d, int64_t* n) const override {
    d[0] = this->s.W, d[1] = this->s.H, d[2] = this->s.D;
    *n = this->s.N;
  }
  void update3(const lbm_config3d* c) override {
    this->s.cfg = *c;
    this->s.write_params();
  }
  void init_field(const float*) override { this->s.init_field(); }
  void set_faces(int kind) override {
    if (Q == 19) this->s.faces = kind;  // the Zou-He faces are written for D3Q19
  }
  void read_slots(uint32_t* out) override { std::memcpy(out, this->s.slot.data(), sizeof(uint32_t) * size_t(this->s.N)); }
  // Params3 order of src/shaders/common3d.ts:53-54: tau, lambda, smagC2, uIn, spongeStart, tauSponge, absorb.
  void params(double* o) const override {
    const auto& P = this->s.P;
    double v[7] = {double(P.tau), double(P.lambda), double(P.smagC2), double(P.uIn), double(P.spongeStart),
                   double(P.tauSponge), double(P.absorb)};
    for (int k = 0; k < 7; k++) o[k] = v[k];
  }
};

}  // namespace

namespace {

template <int Q>
static void* create_3d(const lbm_config3d* c, int real_bits) {
  if (real_bits == 64) return static_cast<Base*>(new Impl3<double, orc::StoreF64, Q>(*c));
  switch (c->precision) {
    case LBM_FP16: return static_cast<Base*>(new Impl3<float, orc::StoreF16, Q>(*c));
    case LBM_FP16_SR: return static_cast<Base*>(new Impl3<float, orc::StoreF16SR, Q>(*c));
    case LBM_BF16: return static_cast<Base*>(new Impl3<float, orc::StoreBF16, Q>(*c));
    default: return static_cast<Base*>(new Impl3<float, orc::StoreF32, Q>(*c));
  }
}

}  // namespace

extern "C" {

void* orc_create_2d(const lbm_config2d* c, int real_bits) {
  if (!c || c->width < 1 || c->height < 1) return nullptr;
  if (real_bits == 64) return static_cast<Base*>(new Impl2<double, orc::StoreF64>(*c));
  return static_cast<Base*>(new Impl2<float, orc::StoreF32>(*c));
}

void* orc_create_3d(const lbm_config3d* c, int real_bits) {
  if (!c || c->width < 3 || c->height < 1 || c->depth < 1) return nullptr;
  if (c->velocity_set == 27) return create_3d<27>(c, real_bits);
  if (c->velocity_set != 0 && c->velocity_set != 19) return nullptr;
  return create_3d<19>(c, real_bits);
}

void orc_destroy(void* h) { delete static_cast<Base*>(h); }
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads(void) { return omp_get_max_threads(); }

void orc_update_2d(void* h, const lbm_config2d* c) { static_cast<Base*>(h)->update2(c); }
void orc_update_3d(void* h, const lbm_config3d* c) { static_cast<Base*>(h)->update3(c); }
void orc_set_sdf(void* h, const float* sdf) { static_cast<Base*>(h)->set_sdf(sdf); }
void orc_brush(void* h, double cx, double cy, double r, int erase, int rebuild) {
  static_cast<Base*>(h)->brush(cx, cy, r, erase != 0, rebuild != 0);
}
void orc_rebuild_flags(void* h) { static_cast<Base*>(h)->rebuild(); }
void orc_read_sdf(void* h, float* out) { static_cast<Base*>(h)->read_sdf(out); }
void orc_init_field(void* h, const float* macro) { static_cast<Base*>(h)->init_field(macro); }
void orc_reset_forces(void* h) { static_cast<Base*>(h)->reset_forces(); }
void orc_set_ramp(void* h, int kind, double u_target, double steps) {
  static_cast<Base*>(h)->set_ramp(kind, u_target, steps);
}
void orc_set_open_faces(void* h, int kind) { static_cast<Base*>(h)->set_faces(kind); }
void orc_steps(void* h, int64_t n) { static_cast<Base*>(h)->steps(n); }
void orc_macro(void* h, double* out) { static_cast<Base*>(h)->macro(out); }
int64_t orc_read_forces(void* h, int64_t* steps, double* values, int64_t cap) {
  return static_cast<Base*>(h)->read_forces(steps, values, cap);
}
void orc_read_populations(void* h, double* out) { static_cast<Base*>(h)->read_populations(out); }
size_t orc_raw_bytes(void* h) { return static_cast<Base*>(h)->raw_bytes(); }
void orc_read_raw(void* h, void* out) { static_cast<Base*>(h)->read_raw(out); }
void orc_read_flags(void* h, uint32_t* out) { static_cast<Base*>(h)->read_flags(out); }
void orc_read_slots(void* h, uint32_t* out) { static_cast<Base*>(h)->read_slots(out); }
void orc_info(void* h, int32_t* dims, int64_t* n, int64_t* step, int32_t* q) {
  Base* b = static_cast<Base*>(h);
  b->dims(dims, n);
  *step = b->step();
  *q = b->q();
}

void orc_params(void* h, double* out) { static_cast<Base*>(h)->params(out); }

// Tables and closed-form helpers, exported so tests can pin them against the reference's unit tests.
void orc_tables2(int* cx, int* cy, double* w, int* opp, int* mirror_y) {
  for (int i = 0; i < 9; i++) {
    cx[i] = orc::CX2[i], cy[i] = orc::CY2[i], w[i] = orc::W2[i], opp[i] = orc::OPP2[i], mirror_y[i] = orc::MIRROR_Y2[i];
  }
}
void orc_tables3(int* cx, int* cy, int* cz, double* w, int* opp, int* my, int* mz) {
  for (int i = 0; i < 19; i++) {
    cx[i] = orc::CX3[i], cy[i] = orc::CY3[i], cz[i] = orc::CZ3[i], w[i] = orc::W3[i];
    opp[i] = orc::opp3(i), my[i] = orc::mirror_y3(i), mz[i] = orc::mirror_z3(i);
  }
}
void orc_tables27(int* cx, int* cy, int* cz, double* w, int* opp, int* my, int* mz) {
  using L = orc::Lattice3<27>;
  for (int i = 0; i < 27; i++) {
    cx[i] = L::CX[i], cy[i] = L::CY[i], cz[i] = L::CZ[i], w[i] = L::W[i];
    opp[i] = L::opp(i), my[i] = L::mirror_y(i), mz[i] = L::mirror_z(i);
  }
}
uint32_t orc_flag_bits(int lattice, int which) {
  if (lattice == 2) return which == 0 ? orc::FLAG_SOLID2 : which == 1 ? orc::FLAG_INLET2 : orc::FLAG_OUTLET2;
  if (lattice == 27) return which == 0 ? orc::Lattice3<27>::SOLID : which == 1 ? orc::Lattice3<27>::INLET : orc::Lattice3<27>::OUTLET;
  return which == 0 ? orc::FLAG_SOLID3 : which == 1 ? orc::FLAG_INLET3 : orc::FLAG_OUTLET3;
}
double orc_trt_lambda3(double tau) { return orc::trt_lambda3(tau); }
// Shifted equilibria (f_eq - w) of src/shaders/aux.ts:15-18 and src/shaders/common3d.ts:78-81, in double.
double orc_feq2(int i, double rho, double ux, double uy) { return orc::Solver2D<double, orc::StoreF64>::feq(i, rho, ux, uy); }
double orc_feq3(int i, double rho, double ux, double uy, double uz) {
  return orc::Solver3D<double, orc::StoreF64>::feq(i, rho, ux, uy, uz);
}

// Codec helpers, exported so tests can pin them against closed forms.
uint16_t orc_f32_to_f16(float f) { return orc::f32_to_f16_rne(f); }
uint16_t orc_f32_to_f16_sr(float f, uint32_t r) { return orc::f32_to_f16_sr(f, r); }
uint32_t orc_step_random(uint32_t node, uint32_t step, int i) { return orc::step_random(node, step, i); }
float orc_f16_to_f32(uint16_t h) { return orc::f16_to_f32(h); }
uint16_t orc_f32_to_bf16(float f) { return orc::f32_to_bf16_rne(f); }
float orc_bf16_to_f32(uint16_t b) { return orc::bf16_to_f32(b); }

}  // extern "C"
