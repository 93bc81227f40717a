// run.cpp - the fused step loop (Gencore::run, src/Core/Gencore.cpp:161-254)
// and the host finalisation of the diagnostics (DiagnosticCUDA.cpp:509-790).
//
// Per lattice step the run enqueues, on the context's stream and without any
// host synchronisation:
//   1. one particle pass: half transverse (+chicane), gather+RK4, R56, half
//      transverse (+corrector), deposition into the source planes and, on
//      output steps, the per-slice beam sums      (Beam::track + source part of
//      FieldSolver*::advance + DiagnosticCUDA beam kernel);
//   2. the field solver passes (FFT rows+columns or ADI x+y sweeps), the last
//      one reducing the per-plane field sums on output steps;
//   3. the slippage ring shift (host integer update + zero fill of the
//      incoming plane);
//   4. on output steps, the map of per-plane field sums to logical slices.
// Diagnostics stay in device history buffers until g4h_run_download_diag.
//
// Ensembles (the same lattice and shapes run again with a new beam or seed:
// shot-noise statistics, scans) replay the run from a hipGraph: after one
// eager pass, g4h_run_capture records the initial record and every step of
// the run once; g4h_run_replay then enqueues the whole run as one graph
// launch on state the caller has re-initialised.
#include <algorithm>
#include <cmath>
#include <vector>

#include "kernels.hpp"

namespace g4 {
int slippage_host(Field* f, double slippage, double sample, int timerun, int periodic);
void context_child_added(Context* c);
void context_child_removed(Context* c);
}

using namespace g4;

struct g4h_run {
  Context* ctx = nullptr;
  Beam* b = nullptr;
  Field* f = nullptr;
  std::vector<g4h_step> steps;
  g4h_run_config cfg{};
  int done = 0;
  int nrec = 0;
  int cap = 0;
  DevBuf<double> beam_hist, field_hist;
  DevBuf<double> sc_hist;  // [cap][nslice] mean short-range field per record (zero at record 0)
  // field ring state when the run starts, and the captured run
  int first0 = 0;
  double acc0 = 0.0;
  int end_first = 0, end_nrec = 0;
  double end_acc = 0.0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int field_generation = 0;  // of the field when captured
  const double* pez = nullptr;  // per-particle longitudinal field read by the graph
  int ef_generation = 0;        // of the beam's space-charge settings when captured
  Context::ParamArena arena;

  ~g4h_run() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (arena.host) (void)hipHostFree(arena.host);
    if (arena.dev) (void)hipFree(arena.dev);
  }

  size_t beam_rec() const { return static_cast<size_t>(b->nslice) * G4H_BEAM_RAW(cfg.nharm); }
  size_t field_rec() const { return static_cast<size_t>(f->nslice) * G4H_FIELD_RAW; }

  void record_initial() {
    if (!cfg.record_diag) return;
    ParticleParams p{};
    p.gamma0 = b->gamma_ref;
    p.stages = ST_DIAG;
    p.diag.raw = beam_hist.get() + nrec * beam_rec();
    p.diag.nharm = cfg.nharm;
    launch_particle(ctx, b, p);
    field_plane_stats(f);
    field_map_stats(f, field_hist.get() + nrec * field_rec());
    ++nrec;
  }

  void step(const g4h_step& st) {
    const double delz = st.delz;
    const double g0 = b->gamma_ref;
    const bool out = cfg.record_diag && st.outstep && nrec < cap;
    field_prepare_solver(f, delz);
    ParticleParams p{};
    p.gamma0 = g0;
    p.stages = ST_T1 | ST_PUSH | ST_T2;
    p.t1 = make_trans(st, g0, delz * 0.5, false);
    p.t2 = make_trans(st, g0, delz * 0.5, true);
    Field* fl[1] = {f};
    make_push(p.push, st, g0, delz, fl, 1, nullptr);
    p.push.particle_ez = b->pez_on ? b->pez.get() : nullptr;
    p.r56 = make_r56(st, g0, cfg.lambda0);
    if (p.r56 != 0.0) p.stages |= ST_R56;
    const bool src_on = make_dep(p.dep, f, b, st, delz) && b->np > 0;
    if (src_on) p.stages |= ST_DEP;
    if (out) {
      p.stages |= ST_DIAG;
      p.diag.raw = beam_hist.get() + nrec * beam_rec();
      p.diag.nharm = cfg.nharm;
    }
    if (b->ef.nz > 0) {
      // BeamSolverCUDA::advance takes the short-range field from the
      // positions after the first transverse half step (Beam.cpp:123-125,
      // BeamSolverCUDA.cpp:523-547): that half step, the field, then the
      // rest of the pass with the field as its per-particle term
      ParticleParams p1{};
      p1.gamma0 = g0;
      p1.stages = ST_T1;
      p1.t1 = p.t1;
      launch_particle(ctx, b, p1);
      launch_space_charge(ctx, b, g0 * g0 / (1.0 + st.aw * st.aw), nullptr,
                          b->pez_on ? b->pez.get() : nullptr);
      p.stages &= ~ST_T1;
      p.push.particle_ez = b->scez.get();
    }
    launch_particle(ctx, b, p);
    field_solve(f, src_on, out);
    slippage_host(f, st.slippage, cfg.sample, cfg.timerun, cfg.periodic);
    if (out) {
      field_map_stats(f, field_hist.get() + nrec * field_rec());
      if (b->ef.nz > 0 && sc_hist.size() >= static_cast<size_t>(nrec + 1) * b->nslice)
        G4_HIP(hipMemcpyAsync(sc_hist.get() + static_cast<size_t>(nrec) * b->nslice,
                              b->scrms.get(), sizeof(double) * b->nslice,
                              hipMemcpyDeviceToDevice, ctx->stream));
      ++nrec;
    }
  }
};

extern "C" {

g4h_status g4h_run_create(g4h_context* ctx, g4h_beam* beam, g4h_field* field,
                          const g4h_step* steps, int nsteps, const g4h_run_config* cfg,
                          g4h_run** out) {
  G4_API_BEGIN
  require(ctx && beam && field && cfg && out, "null argument");
  *out = nullptr;
  require(nsteps >= 0 && (nsteps == 0 || steps), "bad step table");
  require(beam->uploaded, "beam state was never uploaded");
  require(beam->prec == field->prec, "beam and field precision differ");
  require(beam->nslice == field->nslice, "beam and field slice counts differ");
  require(cfg->nharm >= 1 && cfg->nharm <= 8, "nharm must be in [1, 8]");
  require(cfg->lambda0 > 0, "lambda0 must be positive");
  require(!cfg->timerun || cfg->sample > 0, "sample must be positive for time-dependent runs");
  for (int i = 0; i < nsteps; ++i) require(steps[i].delz > 0, "step length must be positive");
  ctx->activate();
  auto* r = new g4h_run();
  try {
    r->ctx = ctx;
    r->b = beam;
    r->f = field;
    r->steps.assign(steps, steps + nsteps);
    r->cfg = *cfg;
    r->first0 = field->first;
    r->acc0 = field->accuslip;
    int nout = 0;
    for (int i = 0; i < nsteps; ++i) nout += steps[i].outstep ? 1 : 0;
    r->cap = cfg->record_diag ? 1 + nout : 0;
    if (r->cap > 0) {
      r->beam_hist.alloc(ctx, r->cap * r->beam_rec());
      r->field_hist.alloc(ctx, r->cap * r->field_rec());
      r->sc_hist.alloc(ctx, static_cast<size_t>(r->cap) * beam->nslice);
      G4_HIP(hipMemsetAsync(r->sc_hist.get(), 0, sizeof(double) * r->cap * beam->nslice,
                            ctx->stream));
    }
    r->record_initial();
  } catch (...) {
    delete r;
    throw;
  }
  *out = r;
  context_child_added(ctx);
  G4_API_END
}

void g4h_run_destroy(g4h_run* r) {
  if (!r) return;
  g4::Context* ctx = r->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete r;
  context_child_removed(ctx);
}

g4h_status g4h_run_advance(g4h_run* r, int nsteps, int* done) {
  G4_API_BEGIN
  require(r, "run is null");
  require(nsteps >= 0, "nsteps must be >= 0");
  r->ctx->activate();
  const int end = std::min<int>(static_cast<int>(r->steps.size()), r->done + nsteps);
  for (; r->done < end; ++r->done) r->step(r->steps[r->done]);
  if (done) *done = r->done;
  G4_API_END
}

g4h_status g4h_run_capture(g4h_run* r) {
  G4_API_BEGIN
  require(r, "run is null");
  require(!r->exec, "the run is already captured");
  require(r->done == static_cast<int>(r->steps.size()),
          "advance the run to its end once before capturing it");
  Context* ctx = r->ctx;
  ctx->activate();
  // Every solver table the graph will point at is built before the capture:
  // a capture cannot allocate or upload tables, nor release those its earlier
  // nodes use.
  {
    std::vector<double> lens;
    for (const g4h_step& st : r->steps) lens.push_back(st.delz);
    std::sort(lens.begin(), lens.end());
    lens.erase(std::unique(lens.begin(), lens.end()), lens.end());
    require(lens.size() <= Field::kMaxDelzTables,
            "a captured run may use at most 64 distinct step lengths per field");
    size_t missing = 0;
    for (double d : lens) missing += r->f->delz_tables.count(d) ? 0 : 1;
    if (r->f->delz_tables.size() + missing > Field::kMaxDelzTables) r->f->evict_delz_tables();
    for (double d : lens) field_prepare_solver(r->f, d);
  }
  G4_HIP(hipStreamSynchronize(ctx->stream));
  ctx->tune_settle();
  if (r->b->prec == G4H_FP64) {  // one parameter block per particle launch
    // one parameter block per particle launch: two per step with space charge
    const int slots = static_cast<int>(r->steps.size()) * (r->b->ef.nz > 0 ? 2 : 1) + 1;
    void* h = nullptr;
    void* d = nullptr;
    G4_HIP(hipHostMalloc(&h, static_cast<size_t>(slots) * Context::kParamBytes, hipHostMallocDefault));
    r->arena.host = static_cast<char*>(h);
    G4_HIP(hipMalloc(&d, static_cast<size_t>(slots) * Context::kParamBytes));
    r->arena.dev = static_cast<char*>(d);
    r->arena.slots = slots;
  }
  r->end_first = r->f->first;
  r->end_acc = r->f->accuslip;
  r->end_nrec = r->nrec;
  r->done = 0;
  r->nrec = 0;
  r->f->first = r->first0;
  r->f->accuslip = r->acc0;
  ctx->capturing = true;
  ctx->arena = &r->arena;
  G4_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  try {
    r->record_initial();
    for (const g4h_step& st : r->steps) r->step(st);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(ctx->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    ctx->capturing = false;
    ctx->arena = nullptr;
    r->done = static_cast<int>(r->steps.size());
    r->nrec = r->end_nrec;
    r->f->first = r->end_first;
    r->f->accuslip = r->end_acc;
    throw;
  }
  ctx->capturing = false;
  ctx->arena = nullptr;
  G4_HIP(hipStreamEndCapture(ctx->stream, &r->graph));
  G4_HIP(hipGraphInstantiate(&r->exec, r->graph, nullptr, nullptr, 0));
  r->field_generation = r->f->generation;
  r->pez = r->b->pez_on ? r->b->pez.get() : nullptr;
  r->ef_generation = r->b->ef_generation;
  r->done = static_cast<int>(r->steps.size());
  require(r->nrec == r->end_nrec && r->f->first == r->end_first,
          "captured run differs from the eager run");
  G4_API_END
}

g4h_status g4h_run_replay(g4h_run* r) {
  G4_API_BEGIN
  require(r, "run is null");
  require(r->exec, "capture the run first (g4h_run_capture)");
  require(r->f->generation == r->field_generation,
          "the field's solver was re-initialised after the capture");
  require(r->f->first == r->first0,
          "re-initialise the field (upload or fill with the run's initial ring offset) before "
          "replaying");
  require((r->b->pez_on ? r->b->pez.get() : nullptr) == r->pez,
          "the beam's per-particle longitudinal field was set or cleared after the capture");
  require(r->b->ef_generation == r->ef_generation,
          "the beam's space-charge settings changed after the capture");
  r->ctx->activate();
  r->f->accuslip = r->acc0;
  G4_HIP(hipGraphLaunch(r->exec, r->ctx->stream));
  r->done = static_cast<int>(r->steps.size());
  r->nrec = r->end_nrec;
  r->f->first = r->end_first;
  r->f->accuslip = r->end_acc;
  G4_API_END
}

g4h_status g4h_run_records(const g4h_run* r, int* n) {
  G4_API_BEGIN
  require(r && n, "null argument");
  *n = r->nrec;
  G4_API_END
}

g4h_status g4h_run_download_diag(g4h_run* r, double* beam_raw, double* field_raw) {
  G4_API_BEGIN
  require(r, "run is null");
  r->ctx->activate();
  G4_HIP(hipStreamSynchronize(r->ctx->stream));
  if (r->nrec > 0 && beam_raw)
    G4_HIP(hipMemcpy(beam_raw, r->beam_hist.get(), sizeof(double) * r->nrec * r->beam_rec(),
                     hipMemcpyDeviceToHost));
  if (r->nrec > 0 && field_raw)
    G4_HIP(hipMemcpy(field_raw, r->field_hist.get(), sizeof(double) * r->nrec * r->field_rec(),
                     hipMemcpyDeviceToHost));
  G4_API_END
}

g4h_status g4h_run_download_sc(g4h_run* r, double* sc_field) {
  G4_API_BEGIN
  require(r && sc_field, "null argument");
  r->ctx->activate();
  G4_HIP(hipStreamSynchronize(r->ctx->stream));
  if (r->nrec > 0)
    G4_HIP(hipMemcpy(sc_field, r->sc_hist.get(), sizeof(double) * r->nrec * r->b->nslice,
                     hipMemcpyDeviceToHost));
  G4_API_END
}

// DiagnosticCUDA::calcBeam host part (DiagnosticCUDA.cpp:509-628), all output
// filters on, one rank.  Energy sums are relative to gamma_ref.
g4h_status g4h_diag_finalize_beam(int nslice, int nharm, const double* raw, const double* current,
                                  double gamma_ref, double* q, double* global) {
  G4_API_BEGIN
  require(raw && q && global, "null argument");
  require(nslice > 0 && nharm >= 1, "bad sizes");
  const int rs = G4H_BEAM_RAW(nharm), qs = G4H_BEAM_QN(nharm);
  double gc = 0, gg1 = 0, gg2 = 0, gx1 = 0, gx2 = 0, gy1 = 0, gy2 = 0;
  for (int is = 0; is < nslice; ++is) {
    const double* s = raw + static_cast<size_t>(is) * rs;
    double* o = q + static_cast<size_t>(is) * qs;
    const double cnt = s[G4H_B_COUNT];
    const double nrm = cnt > 0.0 ? 1.0 / cnt : 1.0;
    const double x1 = s[G4H_B_X1] * nrm, x2 = s[G4H_B_X2] * nrm;
    const double y1 = s[G4H_B_Y1] * nrm, y2 = s[G4H_B_Y2] * nrm;
    const double px1 = s[G4H_B_PX1] * nrm, px2 = s[G4H_B_PX2] * nrm;
    const double py1 = s[G4H_B_PY1] * nrm, py2 = s[G4H_B_PY2] * nrm;
    const double d1 = s[G4H_B_G1] * nrm, d2 = s[G4H_B_G2] * nrm;
    const double g1 = cnt > 0.0 ? gamma_ref + d1 : 0.0;
    const double xpx = s[G4H_B_XPX] * nrm, ypy = s[G4H_B_YPY] * nrm;
    const double vx = x2 - x1 * x1, vpx = px2 - px1 * px1, cx = xpx - x1 * px1;
    const double vy = y2 - y1 * y1, vpy = py2 - py1 * py1, cy = ypy - y1 * py1;
    o[G4H_Q_ENERGY] = g1;
    o[G4H_Q_ENERGYSPREAD] = std::sqrt(std::fabs(d2 - d1 * d1));
    o[G4H_Q_XPOSITION] = x1;
    o[G4H_Q_XSIZE] = std::sqrt(std::fabs(vx));
    o[G4H_Q_YPOSITION] = y1;
    o[G4H_Q_YSIZE] = std::sqrt(std::fabs(vy));
    o[G4H_Q_PXPOSITION] = px1;
    o[G4H_Q_PYPOSITION] = py1;
    const double cur = current ? current[is] : 0.0;
    o[G4H_Q_CURRENT] = cur;
    const double ex = std::sqrt(std::fabs(vx * vpx - cx * cx));
    const double ey = std::sqrt(std::fabs(vy * vpy - cy * cy));
    o[G4H_Q_EMITX] = ex;
    o[G4H_Q_EMITY] = ey;
    o[G4H_Q_BETAX] = ex > 0.0 ? vx / ex * g1 : 0.0;
    o[G4H_Q_BETAY] = ey > 0.0 ? vy / ey * g1 : 0.0;
    o[G4H_Q_ALPHAX] = ex > 0.0 ? -cx / ex : 0.0;
    o[G4H_Q_ALPHAY] = ey > 0.0 ? -cy / ey : 0.0;
    for (int k = 0; k < 10; ++k) o[G4H_Q_XMIN + k] = s[G4H_B_XMIN + k];
    for (int ih = 0; ih < nharm; ++ih) {
      const double bre = s[G4H_BEAM_STATS + 2 * ih] * nrm;
      const double bim = s[G4H_BEAM_STATS + 2 * ih + 1] * nrm;
      o[G4H_BEAM_Q + 2 * ih] = std::hypot(bre, bim);
      o[G4H_BEAM_Q + 2 * ih + 1] = std::atan2(bim, bre);
    }
    gc += cur;
    gg1 += cur * d1;
    gg2 += cur * d2;
    gx1 += cur * x1;
    gx2 += cur * x2;
    gy1 += cur * y1;
    gy2 += cur * y2;
  }
  const double nrm = gc > 0.0 ? 1.0 / gc : 1.0;
  gg1 *= nrm;
  gg2 *= nrm;
  gx1 *= nrm;
  gx2 *= nrm;
  gy1 *= nrm;
  gy2 *= nrm;
  global[G4H_BG_ENERGY] = gamma_ref + gg1;
  global[G4H_BG_ENERGYSPREAD] = std::sqrt(std::fabs(gg2 - gg1 * gg1));
  global[G4H_BG_XPOSITION] = gx1;
  global[G4H_BG_XSIZE] = std::sqrt(std::fabs(gx2 - gx1 * gx1));
  global[G4H_BG_YPOSITION] = gy1;
  global[G4H_BG_YSIZE] = std::sqrt(std::fabs(gy2 - gy1 * gy1));
  G4_API_END
}

// DiagnosticCUDA::calcField host part (DiagnosticCUDA.cpp:685-790), one rank.
g4h_status g4h_diag_finalize_field(int nslice, double dgrid, double xlambda, double slicelength,
                                   const double* raw, double* q, double* global) {
  G4_API_BEGIN
  require(raw && q && global, "null argument");
  require(nslice > 0 && dgrid > 0 && xlambda > 0, "bad sizes");
  const double ks = 4.0 * std::asin(1.0) / xlambda;
  const double scl = dgrid * G4H_EEV / ks;
  double gp = 0, gx1 = 0, gx2 = 0, gy1 = 0, gy2 = 0, gff = 0, gin = 0;
  for (int is = 0; is < nslice; ++is) {
    const double* s = raw + static_cast<size_t>(is) * G4H_FIELD_RAW;
    double* o = q + static_cast<size_t>(is) * G4H_FIELD_Q;
    const double pr = s[G4H_F_POWER];
    double x1 = s[G4H_F_X1], x2 = s[G4H_F_X2], y1 = s[G4H_F_Y1], y2 = s[G4H_F_Y2];
    gp += pr;
    gx1 += x1;
    gx2 += x2;
    gy1 += y1;
    gy2 += y2;
    if (pr > 0.0) {
      x1 /= pr;
      x2 /= pr;
      y1 /= pr;
      y2 /= pr;
    }
    x2 = std::sqrt(std::fabs(x2 - x1 * x1));
    y2 = std::sqrt(std::fabs(y2 - y1 * y1));
    const double cre = s[G4H_F_CENTER_RE], cim = s[G4H_F_CENTER_IM];
    double inten = cre * cre + cim * cim;
    double ff = s[G4H_F_FF_RE] * s[G4H_F_FF_RE] + s[G4H_F_FF_IM] * s[G4H_F_FF_IM];
    ff *= dgrid * dgrid;
    gff += ff;
    gin += inten;
    o[G4H_FQ_POWER] = pr * scl * scl / G4H_VACIMP;
    o[G4H_FQ_XPOSITION] = x1 * dgrid;
    o[G4H_FQ_XSIZE] = x2 * dgrid;
    o[G4H_FQ_YPOSITION] = y1 * dgrid;
    o[G4H_FQ_YSIZE] = y2 * dgrid;
    o[G4H_FQ_INTENSITY_NF] = inten * G4H_EEV * G4H_EEV / ks / ks / G4H_VACIMP;
    o[G4H_FQ_PHASE_NF] = std::atan2(cim, cre);
    o[G4H_FQ_INTENSITY_FF] = ff;
    o[G4H_FQ_PHASE_FF] = std::atan2(s[G4H_F_FF_IM], s[G4H_F_FF_RE]);
  }
  const double nrm = gp > 0.0 ? 1.0 / gp : 1.0;
  gx1 *= nrm;
  gx2 *= nrm;
  gy1 *= nrm;
  gy2 *= nrm;
  global[G4H_FG_ENERGY] = gp * scl * scl / G4H_VACIMP * slicelength / G4H_CLIGHT;
  global[G4H_FG_XPOSITION] = gx1 * dgrid;
  global[G4H_FG_XSIZE] = std::sqrt(std::fabs(gx2 - gx1 * gx1)) * dgrid;
  global[G4H_FG_YPOSITION] = gy1 * dgrid;
  global[G4H_FG_YSIZE] = std::sqrt(std::fabs(gy2 - gy1 * gy1)) * dgrid;
  const double avg = 1.0 / static_cast<double>(nslice);
  global[G4H_FG_INTENSITY_NF] = gin * avg * G4H_EEV * G4H_EEV / ks / ks / G4H_VACIMP;
  global[G4H_FG_INTENSITY_FF] = gff * avg;
  G4_API_END
}

}  // extern "C"
