// abi.cpp - the g4hip C ABI: state objects, step operators, diagnostics.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace g4 {

namespace {
thread_local std::string g_last_error;
}
void context_child_added(Context* c);
void context_child_removed(Context* c);

void set_last_error(const std::string& m) { g_last_error = m; }

hipEvent_t Context::get_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  G4_HIP(hipEventCreate(&e));
  return e;
}

void Context::collect() {
  if (pending.empty()) return;
  G4_HIP(hipStreamSynchronize(stream));
  for (auto& p : pending) {
    float ms = 0.f;
    G4_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    prof[p.cls].launches += 1;
    prof[p.cls].ms += ms;
    event_pool.push_back(p.a);
    event_pool.push_back(p.b);
  }
  pending.clear();
}

void* device_alloc(Context* ctx, size_t bytes) {
  void* p = nullptr;
  if (!ctx) {
    G4_HIP(hipMalloc(&p, bytes));
    return p;
  }
  static std::mutex mu;
  static std::vector<int> pool_ready;
  {
    // keep freed blocks in the pool instead of returning them to the driver
    std::lock_guard<std::mutex> lk(mu);
    if (pool_ready.size() <= static_cast<size_t>(ctx->device))
      pool_ready.resize(static_cast<size_t>(ctx->device) + 1, 0);
    if (!pool_ready[static_cast<size_t>(ctx->device)]) {
      hipMemPool_t pool;
      G4_HIP(hipDeviceGetDefaultMemPool(&pool, ctx->device));
      uint64_t keep = ~0ull;
      G4_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
      pool_ready[static_cast<size_t>(ctx->device)] = 1;
    }
  }
  G4_HIP(hipMallocAsync(&p, bytes, ctx->stream));
  G4_HIP(hipStreamSynchronize(ctx->stream));
  return p;
}

void device_free(Context* ctx, void* p) {
  if (!p) return;
  if (ctx)
    (void)hipFreeAsync(p, ctx->stream);
  else
    (void)hipFree(p);
}

const void* Context::param_push(const void* data, size_t bytes) {
  require(bytes <= kParamBytes, "parameter block too large");
  if (capturing) {
    require(arena && arena->used < arena->slots, "parameter arena of the capture is full");
    char* h = arena->host + static_cast<size_t>(arena->used) * kParamBytes;
    char* d = arena->dev + static_cast<size_t>(arena->used) * kParamBytes;
    ++arena->used;
    std::memcpy(h, data, bytes);
    G4_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream));
    return d;
  }
  if (!param_host) {
    void* h = nullptr;
    G4_HIP(hipHostMalloc(&h, kParamSlots * kParamBytes, hipHostMallocDefault));
    param_host = static_cast<char*>(h);
    void* d = nullptr;
    G4_HIP(hipMalloc(&d, kParamSlots * kParamBytes));
    param_dev = static_cast<char*>(d);
    this->bytes += kParamSlots * kParamBytes;
    for (hipEvent_t& e : param_ev) G4_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  const int k = param_next;
  param_next = (k + 1) % kParamSlots;
  if (param_pending[k]) {
    G4_HIP(hipEventSynchronize(param_ev[k]));
    param_pending[k] = false;
  }
  char* h = param_host + static_cast<size_t>(k) * kParamBytes;
  char* d = param_dev + static_cast<size_t>(k) * kParamBytes;
  std::memcpy(h, data, bytes);
  G4_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream));
  param_last = k;
  return d;
}

void Context::param_commit() {
  if (capturing || param_last < 0) return;
  G4_HIP(hipEventRecord(param_ev[param_last], stream));
  param_pending[param_last] = true;
  param_last = -1;
}

void Context::param_release() {
  for (hipEvent_t e : param_ev)
    if (e) (void)hipEventDestroy(e);
  if (param_host) (void)hipHostFree(param_host);
  if (param_dev) (void)hipFree(param_dev);
  param_host = param_dev = nullptr;
}

// ---- parameter blocks ------------------------------------------------------

// TrackBeamCUDA::track strengths and maps (TrackBeamCUDA.cpp:163-188) and the
// chicane matrix (:93-148).
namespace {
struct M4 {
  double v[4][4];
};
M4 eye4() {
  M4 m{};
  for (int i = 0; i < 4; ++i) m.v[i][i] = 1.0;
  return m;
}
void lmul(M4& m, const M4& e) {
  M4 t{};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) t.v[i][j] += e.v[i][k] * m.v[k][j];
  m = t;
}
// Drift of length len in both planes.
M4 drift(double len) {
  M4 d = eye4();
  d.v[0][1] = d.v[2][3] = len;
  return d;
}
// Sector dipole of bend angle phi and radius rho: a rotation on the arc in
// the bending plane (x), a drift of the arc length in the other (y).  The
// chicane's dipoles bend alternately by +phi and -phi, each in its own frame,
// which gives this same matrix for all four.
M4 sector_bend(double phi, double rho) {
  M4 b = eye4();
  const double c = std::cos(phi), s = std::sin(phi);
  b.v[0][0] = b.v[1][1] = c;
  b.v[0][1] = rho * s;
  b.v[1][0] = -s / rho;
  b.v[2][3] = rho * phi;
  return b;
}
// Thin edge kick of a rectangular dipole face: strength k in x, -k in y.
M4 edge_kick(double k) {
  M4 e = eye4();
  e.v[1][0] = k;
  e.v[3][2] = -k;
  return e;
}

// Four-dipole chicane of total length lt, dipole length lb, outer drifts ld
// (the reference's TrackBeamCUDA.cpp:93-148): dipole, edge, drift, edge,
// dipole, centre drift, and the mirror image, followed by a drift of -lt
// because the transverse tracking still drifts over the chicane's length.
// The product is taken in that order so that it rounds as the reference's.
M4 chicane(double angle, double lb, double ld, double lt) {
  const double rho = lb / std::sin(angle);
  const M4 bend = sector_bend(angle, rho);
  const M4 edge = edge_kick(std::tan(angle) / rho);
  const M4 outer = drift(ld / std::cos(angle));  // dipole 1 to 2 and 3 to 4
  const M4 centre = drift(lt - 4 * lb - 2 * ld);
  const M4 back = drift(-lt);
  M4 m = eye4();
  for (const M4* e : {&bend, &edge, &outer, &edge, &bend, &centre, &bend, &edge, &outer, &edge,
                      &bend, &back})
    lmul(m, *e);
  return m;
}
void plane(double str, double off, int& mode, double& s, double& o) {
  s = str;
  o = 0.0;
  mode = 0;
  if (str == 0.0) return;
  o = off / str;
  mode = str > 0.0 ? 1 : 2;
}
}  // namespace

TransParams make_trans(const g4h_step& st, double gamma0, double delz, bool last) {
  TransParams t{};
  const double aw = st.aw;
  const double betpar0 = std::sqrt(1 - (1 + aw * aw) / gamma0 / gamma0);
  const double quad = st.qf * gamma0;
  const double qnatx = st.kx * aw * aw / gamma0 / betpar0;
  const double qnaty = st.ky * aw * aw / gamma0 / betpar0;
  plane(quad + qnatx, quad * st.dqx + qnatx * st.ax, t.xmode, t.xstr, t.xoff);
  plane(-quad + qnaty, -quad * st.dqy + qnaty * st.ay, t.ymode, t.ystr, t.yoff);
  t.delz = delz;
  t.aw2 = aw * aw;
  t.g2c = gamma0 * gamma0 - 1.0 - aw * aw;
  {
    // K = str/gammaz with gammaz >= gammaz(gamma0)/2 for gamma >= gamma0/2
    const double gz_min = 0.5 * std::sqrt(std::max(t.g2c, 1.0));
    auto series = [&](int mode, double str) {
      const double w = std::fabs(str) / gz_min * delz * delz;
      return mode == 0 || w > 0.5 ? 0 : (w <= 0.011 ? 2 : 1);
    };
    t.xpoly = series(t.xmode, t.xstr);
    t.ypoly = series(t.ymode, t.ystr);
  }
  t.g2chic = gamma0 * gamma0 - 1.0;
  if (last) {
    t.corrector = (st.cx != 0 || st.cy != 0) ? 1 : 0;
    t.cx = st.cx * gamma0;
    t.cy = st.cy * gamma0;
  } else if (st.chic_angle != 0) {
    t.chicane = 1;
    const M4 m = chicane(st.chic_angle, st.chic_lb, st.chic_ld, st.chic_lt);
    t.m00 = m.v[0][0];
    t.m01 = m.v[0][1];
    t.m10 = m.v[1][0];
    t.m11 = m.v[1][1];
    t.m22 = m.v[2][2];
    t.m23 = m.v[2][3];
    t.m32 = m.v[3][2];
    t.m33 = m.v[3][3];
  }
  return t;
}

// TrackBeamCUDA::applyR56 (TrackBeamCUDA.cpp:265-267); 0 when no chicane.
double make_r56(const g4h_step& st, double gamma0, double lambda0) {
  const double angle = st.chic_angle, lb = st.chic_lb, ld = st.chic_ld;
  if (angle == 0) return 0.0;
  double r56 = (4 * lb / std::sin(angle) * (1 - angle / std::tan(angle)) +
                2 * ld * std::tan(angle) / std::cos(angle)) *
               angle;
  return r56 * 4 * std::asin(1) / lambda0 / gamma0;
}

// BeamSolverCUDA::advance host part (BeamSolverCUDA.cpp:476-531) plus the
// FP64 constants of the mixed phase equation (see particle.hip).
void make_push(PushParams& p, const g4h_step& st, double gamma0, double delz,
               Field* const* fields, int nfield, const double* d_eloss) {
  require(nfield >= 1 && nfield <= kMaxFields, "1 <= nfield <= 8 fields");
  double xks_local = 1.0;
  for (int i = 0; i < nfield; ++i) {
    Field* f = fields[i];
    xks_local = f->xks / static_cast<double>(f->harm);
    GatherField& g = p.f[i];
    g.data = f->data();
    g.gridmax = f->gridmax;
    g.dgrid = f->dgrid;
    g.inv_dgrid = 1.0 / f->dgrid;
    g.rtmp = g4h_undulator_fc(st.aw, st.helical, f->harm) / f->xks;
    g.rharm = static_cast<double>(f->harm);
    g.first = f->first;
    g.ngrid = f->ngrid;
    g.nslice = f->nslice;
  }
  p.nfield = nfield;
  double xku = st.ku;
  if (xku == 0.0) xku = xks_local * 0.5 / gamma0 / gamma0;
  p.xks = xks_local;
  p.xku = xku;
  p.aw = st.aw;
  p.autophase = st.autophase;
  p.delz = delz;
  p.ax = st.ax;
  p.ay = st.ay;
  p.kx = st.kx;
  p.ky = st.ky;
  p.gradx = st.gradx;
  p.grady = st.grady;
  p.slice_eloss = d_eloss;
  const double b0 = 1.0 + st.aw * st.aw;
  const double eps0 = b0 / (gamma0 * gamma0);
  const double sq = std::sqrt(1.0 - eps0);
  const double f0 = eps0 / (sq * (1.0 + sq));  // 1/sqrt(1-eps0) - 1
  p.c0 = xks_local * f0 - xku;
  p.k1 = xks_local * eps0 * 0.5 * std::pow(1.0 - eps0, -1.5);
  p.k2 = xks_local * eps0 * eps0 * 0.375 * std::pow(1.0 - eps0, -2.5);
  p.inv_b0 = 1.0 / b0;
}

// Source switch and scale (FieldSolverADICUDA.cpp:697-713, :86-95).
bool make_dep(DepParams& d, Field* f, Beam* b, const g4h_step& st, double delz) {
  const bool on = st.in_undulator && f->enabled && (f->harm % 2) == 1;
  if (!on) return false;
  ensure_source(f);
  d.src = f->src();
  d.rect = f->src_rect.get();
  d.bbox = f->dep_bbox.get();
  d.cur_over_npar = b->cur_over_npar.get();
  const double fc = g4h_undulator_fc(st.aw, st.helical, f->harm);
  d.scl_factor = fc * G4H_VACIMP * f->xks * delz / (4.0 * G4H_EEV * f->dgrid * f->dgrid);
  d.gridmax = f->gridmax;
  d.dgrid = f->dgrid;
  d.inv_dgrid = 1.0 / f->dgrid;
  d.ax = st.ax;
  d.ay = st.ay;
  d.kx = st.kx;
  d.ky = st.ky;
  d.gradx = st.gradx;
  d.grady = st.grady;
  d.ngrid = f->ngrid;
  d.harm = f->harm;
  d.window_cap = 3072;  // 48 KiB of FP64 complex accumulators: 3 workgroups per CU
  return true;
}

// shared by g4h_apply_slippage and the fused run
int slippage_host(Field* f, double slippage, double sample, int timerun, int periodic) {
  if (!timerun) return 0;
  f->accuslip += slippage;
  const int n = f->nslice;
  int shifts = 0;
  int direction = 1;
  while (std::abs(f->accuslip) > (sample * 0.8)) {
    direction = f->accuslip < 0 ? -1 : 1;
    f->accuslip -= sample * direction;
    int last = (f->first + n - 1) % n;
    if (direction < 0) last = (last + 1) % n;
    if (!periodic) field_zero_plane(f, last);
    f->first = last;
    if (direction < 0) f->first = (last + 1) % n;
    ++shifts;
  }
  return shifts;
}

}  // namespace g4

using namespace g4;

static void release_context(g4h_context* ctx) {
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& p : ctx->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  ctx->tune_release();
  ctx->param_release();
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

namespace g4 {
void context_child_added(Context* c) { ++c->children; }
void context_child_removed(Context* c) {
  if (--c->children == 0 && c->destroy_pending) release_context(static_cast<g4h_context*>(c));
}
}  // namespace g4

extern "C" {

int g4h_abi_version(void) { return 1; }

int g4h_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* g4h_last_error(void) { return g_last_error.c_str(); }

g4h_status g4h_context_create(int device, g4h_context** out) {
  G4_API_BEGIN
  require(out != nullptr, "out is null");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) fail(G4H_ERR_NO_DEVICE, "no GPU visible");
  require(device >= 0 && device < n, "device index out of range");
  auto* c = new g4h_context();
  c->device = device;
  try {
    c->activate();
    G4_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    // device address of the fixed scratch array of g4h_diag_beam_raw_noalloc
    scratch_base(device);
  } catch (...) {
    delete c;
    throw;
  }
  *out = c;
  G4_API_END
}

void g4h_context_destroy(g4h_context* ctx) {
  if (!ctx) return;
  if (ctx->children > 0) {
    ctx->destroy_pending = true;
    return;
  }
  release_context(ctx);
}

g4h_status g4h_context_synchronize(g4h_context* ctx) {
  G4_API_BEGIN
  require(ctx, "ctx is null");
  ctx->activate();
  G4_HIP(hipStreamSynchronize(ctx->stream));
  G4_API_END
}

g4h_status g4h_context_memory(const g4h_context* ctx, uint64_t* bytes) {
  G4_API_BEGIN
  require(ctx && bytes, "null argument");
  *bytes = ctx->bytes;
  G4_API_END
}

// ---- beam -------------------------------------------------------------------

g4h_status g4h_beam_create(g4h_context* ctx, g4h_precision precision, int nslice,
                           const int32_t* off, const double* current, double gamma_ref,
                           g4h_beam** out) {
  G4_API_BEGIN
  require(ctx && off && out, "null argument");
  require(nslice > 0, "nslice must be positive");
  require(precision == G4H_FP64 || precision == G4H_MIXED, "unknown precision");
  require(gamma_ref > 1.0, "gamma_ref must exceed 1");
  *out = nullptr;
  require(off[0] == 0, "slice_offsets[0] must be 0");
  for (int s = 0; s < nslice; ++s) require(off[s + 1] >= off[s], "slice_offsets must be non-decreasing");
  ctx->activate();
  auto* b = new g4h_beam();
  try {
    b->ctx = ctx;
    b->prec = precision;
    b->nslice = nslice;
    b->np = off[nslice];
    b->gamma_ref = gamma_ref;
    b->off_h.assign(off, off + nslice + 1);
    b->current_h.assign(nslice, 0.0);
    if (current) b->current_h.assign(current, current + nslice);
    for (int s = 0; s < nslice; ++s) b->max_slice_np = std::max(b->max_slice_np, off[s + 1] - off[s]);
    b->off.alloc(ctx, nslice + 1);
    G4_HIP(hipMemcpy(b->off.get(), off, sizeof(int32_t) * (nslice + 1), hipMemcpyHostToDevice));
    b->cur_over_npar.alloc(ctx, nslice);
    const size_t n = static_cast<size_t>(std::max<int64_t>(b->np, 1));
    if (precision == G4H_FP64) {
      b->xd.alloc(ctx, n);
      b->yd.alloc(ctx, n);
      b->pxd.alloc(ctx, n);
      b->pyd.alloc(ctx, n);
      b->gd.alloc(ctx, n);
    } else {
      b->xf.alloc(ctx, n);
      b->yf.alloc(ctx, n);
      b->pxf.alloc(ctx, n);
      b->pyf.alloc(ctx, n);
      b->gf.alloc(ctx, n);
    }
    b->theta.alloc(ctx, n);
  } catch (...) {
    delete b;
    throw;
  }
  *out = b;
  context_child_added(ctx);
  if (g4h_status st = g4h_beam_set_current(b, b->current_h.data()); st != G4H_OK) {
    g4h_beam_destroy(b);
    *out = nullptr;
    return st;
  }
  G4_API_END
}

void g4h_beam_destroy(g4h_beam* b) {
  if (!b) return;
  g4::Context* ctx = b->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete b;
  context_child_removed(ctx);
}

g4h_status g4h_beam_set_current(g4h_beam* b, const double* current) {
  G4_API_BEGIN
  require(b && current, "null argument");
  b->ctx->activate();
  b->current_h.assign(current, current + b->nslice);
  std::vector<double> con(b->nslice);
  for (int s = 0; s < b->nslice; ++s) {
    const int np = b->off_h[s + 1] - b->off_h[s];
    con[s] = np > 0 ? current[s] / static_cast<double>(np) : 0.0;
  }
  G4_HIP(hipMemcpy(b->cur_over_npar.get(), con.data(), sizeof(double) * b->nslice,
                   hipMemcpyHostToDevice));
  G4_API_END
}

g4h_status g4h_beam_upload(g4h_beam* b, const double* x, const double* y, const double* px,
                           const double* py, const double* gamma, const double* theta) {
  G4_API_BEGIN
  require(b && x && y && px && py && gamma && theta, "null argument");
  b->ctx->activate();
  const size_t n = static_cast<size_t>(b->np);
  G4_HIP(hipStreamSynchronize(b->ctx->stream));
  if (n > 0) {
    if (b->prec == G4H_FP64) {
      G4_HIP(hipMemcpy(b->xd.get(), x, n * 8, hipMemcpyHostToDevice));
      G4_HIP(hipMemcpy(b->yd.get(), y, n * 8, hipMemcpyHostToDevice));
      G4_HIP(hipMemcpy(b->pxd.get(), px, n * 8, hipMemcpyHostToDevice));
      G4_HIP(hipMemcpy(b->pyd.get(), py, n * 8, hipMemcpyHostToDevice));
      G4_HIP(hipMemcpy(b->gd.get(), gamma, n * 8, hipMemcpyHostToDevice));
    } else {
      std::vector<float> tmp(n);
      auto up = [&](const double* src, float* dst, double shift) {
        for (size_t i = 0; i < n; ++i) tmp[i] = static_cast<float>(src[i] - shift);
        G4_HIP(hipMemcpy(dst, tmp.data(), n * 4, hipMemcpyHostToDevice));
      };
      up(x, b->xf.get(), 0.0);
      up(y, b->yf.get(), 0.0);
      up(px, b->pxf.get(), 0.0);
      up(py, b->pyf.get(), 0.0);
      up(gamma, b->gf.get(), b->gamma_ref);
    }
    G4_HIP(hipMemcpy(b->theta.get(), theta, n * 8, hipMemcpyHostToDevice));
  }
  b->uploaded = true;
  G4_API_END
}

g4h_status g4h_beam_download(g4h_beam* b, double* x, double* y, double* px, double* py,
                             double* gamma, double* theta) {
  G4_API_BEGIN
  require(b && x && y && px && py && gamma && theta, "null argument");
  require(b->uploaded, "beam was never uploaded");
  b->ctx->activate();
  G4_HIP(hipStreamSynchronize(b->ctx->stream));
  const size_t n = static_cast<size_t>(b->np);
  if (n > 0) {
    if (b->prec == G4H_FP64) {
      G4_HIP(hipMemcpy(x, b->xd.get(), n * 8, hipMemcpyDeviceToHost));
      G4_HIP(hipMemcpy(y, b->yd.get(), n * 8, hipMemcpyDeviceToHost));
      G4_HIP(hipMemcpy(px, b->pxd.get(), n * 8, hipMemcpyDeviceToHost));
      G4_HIP(hipMemcpy(py, b->pyd.get(), n * 8, hipMemcpyDeviceToHost));
      G4_HIP(hipMemcpy(gamma, b->gd.get(), n * 8, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> tmp(n);
      auto down = [&](const float* src, double* dst, double shift) {
        G4_HIP(hipMemcpy(tmp.data(), src, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) dst[i] = static_cast<double>(tmp[i]) + shift;
      };
      down(b->xf.get(), x, 0.0);
      down(b->yf.get(), y, 0.0);
      down(b->pxf.get(), px, 0.0);
      down(b->pyf.get(), py, 0.0);
      down(b->gf.get(), gamma, b->gamma_ref);
    }
    G4_HIP(hipMemcpy(theta, b->theta.get(), n * 8, hipMemcpyDeviceToHost));
  }
  G4_API_END
}

g4h_status g4h_beam_info(const g4h_beam* b, int* nslice, int64_t* np, int* precision) {
  G4_API_BEGIN
  require(b, "beam is null");
  if (nslice) *nslice = b->nslice;
  if (np) *np = b->np;
  if (precision) *precision = b->prec;
  G4_API_END
}

// ---- field --------------------------------------------------------------------

g4h_status g4h_field_create(g4h_context* ctx, g4h_precision precision, int nslice, int ngrid,
                            double gridmax, double xlambda0, double slicelength, int harm,
                            g4h_field** out) {
  G4_API_BEGIN
  require(ctx && out, "null argument");
  *out = nullptr;
  require(nslice > 0, "nslice must be positive");
  require(ngrid >= 3 && ngrid <= 1024, "ngrid must be in [3, 1024]");
  require(gridmax > 0 && xlambda0 > 0, "gridmax and xlambda0 must be positive");
  require(harm >= 1, "harm must be >= 1");
  require(precision == G4H_FP64 || precision == G4H_MIXED, "unknown precision");
  ctx->activate();
  auto* f = new g4h_field();
  try {
    f->ctx = ctx;
    f->prec = precision;
    f->nslice = nslice;
    f->ngrid = ngrid;
    f->harm = harm;
    f->gridmax = gridmax;
    f->dgrid = 2 * gridmax / static_cast<double>(ngrid - 1);  // Field.cpp:61
    f->xlambda = xlambda0 / static_cast<double>(harm);       // Field.cpp:57
    f->xks = 4. * std::asin(1) / f->xlambda;                  // Field.cpp:80
    f->slicelength = slicelength;
    const size_t cells = static_cast<size_t>(f->cells());
    if (precision == G4H_FP64) {
      f->ud.alloc(ctx, cells);
      G4_HIP(hipMemset(f->ud.get(), 0, cells * sizeof(double2)));
    } else {
      f->uf.alloc(ctx, cells);
      G4_HIP(hipMemset(f->uf.get(), 0, cells * sizeof(float2)));
    }
  } catch (...) {
    delete f;
    throw;
  }
  *out = f;
  context_child_added(ctx);
  G4_API_END
}

void g4h_field_destroy(g4h_field* f) {
  if (!f) return;
  g4::Context* ctx = f->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete f;
  context_child_removed(ctx);
}

g4h_status g4h_field_upload(g4h_field* f, const double* re, const double* im, int first) {
  G4_API_BEGIN
  require(f && re && im, "null argument");
  require(first >= 0 && first < f->nslice, "first out of range");
  f->ctx->activate();
  G4_HIP(hipStreamSynchronize(f->ctx->stream));
  const size_t cells = static_cast<size_t>(f->cells());
  if (f->prec == G4H_FP64) {
    std::vector<double2> tmp(cells);
    for (size_t i = 0; i < cells; ++i) tmp[i] = make_double2(re[i], im[i]);
    G4_HIP(hipMemcpy(f->ud.get(), tmp.data(), cells * sizeof(double2), hipMemcpyHostToDevice));
  } else {
    std::vector<float2> tmp(cells);
    for (size_t i = 0; i < cells; ++i) tmp[i] = make_float2(static_cast<float>(re[i]), static_cast<float>(im[i]));
    G4_HIP(hipMemcpy(f->uf.get(), tmp.data(), cells * sizeof(float2), hipMemcpyHostToDevice));
  }
  f->first = first;
  f->uploaded = true;
  G4_API_END
}

g4h_status g4h_field_download(g4h_field* f, double* re, double* im, int* first) {
  G4_API_BEGIN
  require(f && re && im, "null argument");
  f->ctx->activate();
  G4_HIP(hipStreamSynchronize(f->ctx->stream));
  const size_t cells = static_cast<size_t>(f->cells());
  if (f->prec == G4H_FP64) {
    std::vector<double2> tmp(cells);
    G4_HIP(hipMemcpy(tmp.data(), f->ud.get(), cells * sizeof(double2), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < cells; ++i) {
      re[i] = tmp[i].x;
      im[i] = tmp[i].y;
    }
  } else {
    std::vector<float2> tmp(cells);
    G4_HIP(hipMemcpy(tmp.data(), f->uf.get(), cells * sizeof(float2), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < cells; ++i) {
      re[i] = tmp[i].x;
      im[i] = tmp[i].y;
    }
  }
  if (first) *first = f->first;
  G4_API_END
}

g4h_status g4h_field_copy(g4h_context* ctx, g4h_field* dst, const g4h_field* src) {
  G4_API_BEGIN
  require(ctx && dst && src, "null argument");
  require(dst->prec == src->prec && dst->nslice == src->nslice && dst->ngrid == src->ngrid,
          "fields differ in precision or shape");
  ctx->activate();
  if (dst != src) field_copy(dst, src);
  G4_API_END
}

g4h_status g4h_field_get_ring(const g4h_field* f, int* first, double* accuslip) {
  G4_API_BEGIN
  require(f, "field is null");
  if (first) *first = f->first;
  if (accuslip) *accuslip = f->accuslip;
  G4_API_END
}

g4h_status g4h_field_set_ring(g4h_field* f, int first, double accuslip) {
  G4_API_BEGIN
  require(f, "field is null");
  require(first >= 0 && first < f->nslice, "first out of range");
  f->first = first;
  f->accuslip = accuslip;
  G4_API_END
}

g4h_status g4h_field_set_enabled(g4h_field* f, int enabled) {
  G4_API_BEGIN
  require(f, "field is null");
  f->enabled = enabled != 0;
  G4_API_END
}

g4h_status g4h_field_init_solver(g4h_field* f, g4h_solver_kind kind, int filter, double xc,
                                 double yc, double sig, int low_memory) {
  G4_API_BEGIN
  require(f, "field is null");
  require(kind == G4H_SOLVER_ADI || kind == G4H_SOLVER_FFT, "unknown solver kind");
  (void)low_memory;
  f->ctx->activate();
  f->kind = kind;
  f->xc = xc;
  f->yc = yc;
  f->sig = sig;
  // FieldSolverFFTCUDA::initSourceFilter (FieldSolverFFTCUDA.cpp:213-221)
  f->filter = kind == G4H_SOLVER_FFT && filter && !(sig <= 0 || xc <= 0 || yc <= 0);
  G4_HIP(hipStreamSynchronize(f->ctx->stream));  // tables may still be read
  f->fft_tw.release();  // tables are rebuilt on the next init
  f->fft_tw_conv.release();
  f->fft_sigmoid.release();
  f->delz_tables.clear();
  ++f->generation;  // device tables of the solver are gone
  f->adi_c = f->adi_cbet = f->adi_cwet = nullptr;
  f->fft_prop = nullptr;
  f->adi_delz = 0.0;
  f->fft_delz = 0.0;
  f->solver_ready = false;
  G4_API_END
}

// ---- operators -------------------------------------------------------------------

static void check_beam(g4h_beam* b) {
  require(b != nullptr, "beam is null");
  require(b->uploaded, "beam state was never uploaded");
}

g4h_status g4h_track_transverse(g4h_context* ctx, g4h_beam* b, const g4h_step* und, double delz,
                                int last_step) {
  G4_API_BEGIN
  require(ctx && und, "null argument");
  check_beam(b);
  ctx->activate();
  ParticleParams p{};
  p.gamma0 = b->gamma_ref;
  if (last_step) {
    p.stages = ST_T2;
    p.t2 = make_trans(*und, b->gamma_ref, delz, true);
  } else {
    p.stages = ST_T1;
    p.t1 = make_trans(*und, b->gamma_ref, delz, false);
  }
  launch_particle(ctx, b, p);
  G4_API_END
}

g4h_status g4h_apply_r56(g4h_context* ctx, g4h_beam* b, const g4h_step* und, double lambda0) {
  G4_API_BEGIN
  require(ctx && und, "null argument");
  require(lambda0 > 0, "lambda0 must be positive");
  check_beam(b);
  ctx->activate();
  if (und->chic_angle == 0) return G4H_OK;
  ParticleParams p{};
  p.gamma0 = b->gamma_ref;
  p.stages = ST_R56;
  p.r56 = make_r56(*und, b->gamma_ref, lambda0);
  launch_particle(ctx, b, p);
  G4_API_END
}

g4h_status g4h_beam_advance(g4h_context* ctx, g4h_beam* b, g4h_field* const* fields, int nfield,
                            const g4h_step* und, double delz, const double* slice_eloss) {
  G4_API_BEGIN
  require(ctx && fields && und, "null argument");
  check_beam(b);
  require(nfield >= 1 && nfield <= kMaxFields, "nfield must be in [1, 8]");
  for (int i = 0; i < nfield; ++i) {
    require(fields[i] != nullptr, "field is null");
    require(fields[i]->prec == b->prec, "beam and field precision differ");
    require(fields[i]->nslice == b->nslice, "field and beam slice counts differ");
  }
  require(!(slice_eloss && b->pez_on),
          "the beam holds a per-particle longitudinal field: fold the slice loss into it");
  ctx->activate();
  const double* d_eloss = nullptr;
  DevBuf<double>& eb = b->eloss;
  if (slice_eloss) {
    G4_HIP(hipStreamSynchronize(ctx->stream));
    if (eb.size() != static_cast<size_t>(b->nslice)) eb.alloc(ctx, b->nslice);
    G4_HIP(hipMemcpy(eb.get(), slice_eloss, sizeof(double) * b->nslice, hipMemcpyHostToDevice));
    d_eloss = eb.get();
  }
  ParticleParams p{};
  p.gamma0 = b->gamma_ref;
  p.stages = ST_PUSH;
  std::vector<Field*> fl(fields, fields + nfield);
  make_push(p.push, *und, b->gamma_ref, delz, fl.data(), nfield, d_eloss);
  p.push.particle_ez = b->pez_on ? b->pez.get() : nullptr;
  if (b->ef.nz > 0) {
    // short-range field from the current positions, the slice loss and a
    // per-particle term folded in (BeamSolverCUDA.cpp:523-547)
    const double g0 = b->gamma_ref;
    launch_space_charge(ctx, b, g0 * g0 / (1.0 + und->aw * und->aw), d_eloss, p.push.particle_ez);
    p.push.particle_ez = b->scez.get();
    p.push.slice_eloss = nullptr;
  }
  launch_particle(ctx, b, p);
  if (slice_eloss) G4_HIP(hipStreamSynchronize(ctx->stream));
  G4_API_END
}

g4h_status g4h_beam_init_efield(g4h_beam* b, const g4h_efield* ef, double lambda) {
  G4_API_BEGIN
  require(b != nullptr, "beam is null");
  require(ef != nullptr, "null argument");
  require(ef->nz >= 0 && ef->nphi >= 0 && ef->ngrid >= 0, "negative solver size");
  if (ef->nz > 0) {
    require(ef->ngrid >= 2 && ef->rmax > 0 && lambda > 0,
            "space charge needs ngrid >= 2, rmax > 0 and lambda > 0");
    require(!ef->longrange,
            "the long-range (current-profile) space-charge field is not built; pass it as the "
            "slice loss");
    require(sc_lds_bytes(*ef) <= 64 * 1024,
            "nz * (2 nphi + 1) * ngrid too large for the solver's LDS (at most 2730)");
  }
  b->ctx->activate();
  b->ef = *ef;
  b->ef_ks = ef->nz > 0 ? 2.0 * M_PI / lambda : 0.0;
  ++b->ef_generation;
  if (ef->nz > 0) {
    if (b->scez.size() != static_cast<size_t>(b->np)) b->scez.alloc(b->ctx, static_cast<size_t>(b->np));
    if (b->scrms.size() != static_cast<size_t>(b->nslice))
      b->scrms.alloc(b->ctx, static_cast<size_t>(b->nslice));
    G4_HIP(hipMemsetAsync(b->scrms.get(), 0, sizeof(double) * b->nslice, b->ctx->stream));
  }
  G4_API_END
}

g4h_status g4h_beam_sc_field(g4h_beam* b, double* field) {
  G4_API_BEGIN
  require(b != nullptr, "beam is null");
  require(field != nullptr, "null argument");
  b->ctx->activate();
  if (b->ef.nz <= 0) {
    for (int s = 0; s < b->nslice; ++s) field[s] = 0.0;
    return G4H_OK;
  }
  G4_HIP(hipStreamSynchronize(b->ctx->stream));
  G4_HIP(hipMemcpy(field, b->scrms.get(), sizeof(double) * b->nslice, hipMemcpyDeviceToHost));
  G4_API_END
}

g4h_status g4h_beam_add_energy(g4h_context* ctx, g4h_beam* b, const double* dgamma) {
  G4_API_BEGIN
  require(ctx && dgamma, "null argument");
  check_beam(b);
  if (b->nslice == 0 || b->np == 0) return G4H_OK;
  ctx->activate();
  DevBuf<double>& kb = b->kick;
  // the buffer may still be read by a kick queued earlier
  G4_HIP(hipStreamSynchronize(ctx->stream));
  if (kb.size() != static_cast<size_t>(b->nslice)) kb.alloc(ctx, b->nslice);
  G4_HIP(hipMemcpy(kb.get(), dgamma, sizeof(double) * b->nslice, hipMemcpyHostToDevice));
  launch_slice_kick(ctx, b, kb.get());
  G4_API_END
}

g4h_status g4h_beam_set_particle_ez(g4h_beam* b, const double* ez) {
  G4_API_BEGIN
  check_beam(b);
  Context* ctx = b->ctx;
  ctx->activate();
  if (!ez) {
    b->pez_on = false;
    return G4H_OK;
  }
  // the previous values may still be read by queued passes
  G4_HIP(hipStreamSynchronize(ctx->stream));
  if (b->pez.size() != static_cast<size_t>(b->np)) b->pez.alloc(ctx, static_cast<size_t>(b->np));
  if (b->np > 0)
    G4_HIP(hipMemcpy(b->pez.get(), ez, sizeof(double) * static_cast<size_t>(b->np),
                     hipMemcpyHostToDevice));
  b->pez_on = true;
  G4_API_END
}

g4h_status g4h_field_track(g4h_context* ctx, g4h_field* f, g4h_beam* b, const g4h_step* und,
                           double delz) {
  G4_API_BEGIN
  require(ctx && f && und, "null argument");
  check_beam(b);
  require(f->prec == b->prec, "beam and field precision differ");
  require(f->nslice == b->nslice, "field and beam slice counts differ");
  ctx->activate();
  field_prepare_solver(f, delz);
  ParticleParams p{};
  p.gamma0 = b->gamma_ref;
  p.stages = ST_DEP;
  const bool on = make_dep(p.dep, f, b, *und, delz) && b->np > 0;
  if (on) launch_particle(ctx, b, p);
  field_solve(f, on, false);
  G4_API_END
}

g4h_status g4h_apply_slippage(g4h_context* ctx, g4h_field* f, double slippage, double sample,
                              int timerun, int periodic, int* shifts) {
  G4_API_BEGIN
  require(ctx && f, "null argument");
  ctx->activate();
  const int n = slippage_host(f, slippage, sample, timerun, periodic);
  if (shifts) *shifts = n;
  G4_API_END
}

g4h_status g4h_beam_cell_index(g4h_context* ctx, g4h_beam* b, const g4h_field* f, int32_t* cell) {
  G4_API_BEGIN
  require(ctx && f && cell, "null argument");
  check_beam(b);
  ctx->activate();
  if (b->np == 0) return G4H_OK;
  DevBuf<int32_t> d;
  d.alloc(ctx, static_cast<size_t>(b->np));
  launch_cell_index(ctx, b, f, d.get());
  G4_HIP(hipMemcpyAsync(cell, d.get(), sizeof(int32_t) * b->np, hipMemcpyDeviceToHost, ctx->stream));
  G4_HIP(hipStreamSynchronize(ctx->stream));
  G4_API_END
}

g4h_status g4h_diag_beam_raw(g4h_context* ctx, g4h_beam* b, int nharm, double* raw) {
  G4_API_BEGIN
  require(ctx && raw, "null argument");
  check_beam(b);
  require(nharm >= 1 && nharm <= 8, "nharm must be in [1, 8]");
  ctx->activate();
  DevBuf<double> d;
  const size_t n = static_cast<size_t>(b->nslice) * G4H_BEAM_RAW(nharm);
  d.alloc(ctx, n);
  ParticleParams p{};
  p.gamma0 = b->gamma_ref;
  p.stages = ST_DIAG;
  p.diag.raw = d.get();
  p.diag.nharm = nharm;
  launch_particle(ctx, b, p);
  G4_HIP(hipMemcpyAsync(raw, d.get(), n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  G4_HIP(hipStreamSynchronize(ctx->stream));
  G4_API_END
}

g4h_status g4h_diag_beam_raw_noalloc(g4h_context* ctx, g4h_beam* b, int nharm, double* raw) {
  G4_API_BEGIN
  require(ctx && raw, "null argument");
  check_beam(b);
  require(nharm >= 1 && nharm <= 8, "nharm must be in [1, 8]");
  ctx->activate();
  if (b->nslice == 0) return G4H_OK;
  diag_beam_raw_noalloc(ctx, b, nharm, raw);
  G4_API_END
}

g4h_status g4h_diag_field_raw(g4h_context* ctx, g4h_field* f, double* raw) {
  G4_API_BEGIN
  require(ctx && f && raw, "null argument");
  ctx->activate();
  DevBuf<double> d;
  const size_t n = static_cast<size_t>(f->nslice) * G4H_FIELD_RAW;
  d.alloc(ctx, n);
  field_plane_stats(f);
  field_map_stats(f, d.get());
  G4_HIP(hipMemcpyAsync(raw, d.get(), n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  G4_HIP(hipStreamSynchronize(ctx->stream));
  G4_API_END
}

g4h_status g4h_profile_enable(g4h_context* ctx, int on) {
  G4_API_BEGIN
  require(ctx, "ctx is null");
  ctx->activate();
  ctx->collect();
  ctx->profile = on != 0;
  for (auto& p : ctx->prof) p = Context::Prof{};
  G4_API_END
}

g4h_status g4h_profile_read(g4h_context* ctx, int cls, int64_t* launches, double* ms) {
  G4_API_BEGIN
  require(ctx, "ctx is null");
  require(cls >= 0 && cls < G4H_K_COUNT, "kernel class out of range");
  ctx->activate();
  ctx->collect();
  if (launches) *launches = ctx->prof[cls].launches;
  if (ms) *ms = ctx->prof[cls].ms;
  G4_API_END
}

}  // extern "C"
