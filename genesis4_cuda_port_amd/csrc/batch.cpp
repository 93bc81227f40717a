// batch.cpp - many independent steady-state runs advanced together.
//
// A taper scan or a stream of small requests is many one-slice runs that
// share a lattice but differ in particle count, grid and undulator taper.
// Launched one by one they leave the GPU idle (one workgroup per run); here
// runs with the same grid (ngrid, gridmax) and reference energy form a group
// that is one Beam (one slice per run, uneven particle counts through the
// slice offsets) and one Field (one ring plane per run), so each step of the
// whole group is one particle launch, one solver step and, on output steps,
// one field-sum map.  Per-run step parameters (the taper changes aw, and with
// it the focusing, the coupling fc(aw) and the phase-equation constants)
// reach the particle pass through a per-slice parameter table, built on the
// host from the same make_trans / make_push / make_dep as every other path,
// for every step at creation, and kept on the device ([step][run]).
//
// ADI batches on grids up to 256 (the segmented sweeps) run each step of
// EVERY run as one launch per kernel on the batch's stream (G4H_TUNE_BATCH =
// 0, the default): the particle passes (one workgroup per run, its group's
// beam slice and parameter entry), the x and the y sweep (one workgroup per
// 64-line block of every run's plane, its group's tables of the step), and on
// output steps one kernel writing every run's field record.  Three or four
// launches per step, whatever the number of grids: per-group launches on
// per-group streams (G4H_TUNE_BATCH = 2, and what other batches run) spent
// ~16 host launches per step on the 96-run taper scan, whose host time then
// bounded the step.  Groups keep their own contexts (their own streams) for
// creation, records and downloads; g4h_batch_advance orders the batch's
// stream after them and them after it.  ADI batches can also run every step
// of every run in ONE launch (G4H_TUNE_BATCH = 1; launch_batch_persistent:
// one workgroup per run loops over the steps with the single-chain sweeps);
// on the taper scan that launch is slower than the per-step launches (the
// sweeps' recurrences of a plane run on one CU with their memory latency
// exposed), so it is not the default.
//
// Steady state only: one slice per run, no slippage (Gencore.cpp:229 is a
// no-op there, ControlCUDA.cpp:150).
#include <algorithm>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "kernels.hpp"

namespace g4 {
void context_child_added(Context* c);
void context_child_removed(Context* c);
void set_last_error(const std::string& m);
}  // namespace g4

using namespace g4;

namespace {

void ok(g4h_status s) {
  if (s != G4H_OK) throw Error(s, g4h_last_error());
}

struct Group {
  g4h_context* gctx = nullptr;  // own stream: groups run concurrently
  int G = 0;
  double gridmax = 0, gamma0 = 0;
  std::vector<int> runs;
  g4h_beam* beam = nullptr;
  g4h_field* field = nullptr;
  // parameter tables of every step ([step][run]), built once at creation
  DevBuf<ParticleParams> dtab;
  std::vector<ParticleParams> head;  // host copy of each step's first entry
  std::vector<uint8_t> src_on;       // per step: deposition on
  hipEvent_t tail = nullptr;         // end of the group's work so far
  DevBuf<double> beam_hist, field_hist;
};

}  // namespace

struct g4h_batch {
  Context* ctx = nullptr;
  g4h_precision prec = G4H_FP64;
  g4h_solver_kind solver = G4H_SOLVER_ADI;
  double lambda0 = 0;
  int nharm = 1;
  std::vector<g4h_step> steps;
  std::vector<std::vector<double>> aw;  // per run, per step (empty: the step table's)
  std::vector<Group> groups;
  std::vector<std::pair<int, int>> where;  // run -> (group, slice in group)
  int done = 0, nrec = 0, cap = 0;
  std::vector<int32_t> rec_of_step;  // record index written by each step, or -1
  // one-launch path, and the per-step launches of every run (grid)
  bool persistent = false, grid = false;
  int win_cells = 0, gmax = 0;
  DevBuf<BatchSegDesc> dblocks;
  int nblocks = 0;
  std::vector<uint8_t> src_all;  // per step: the source is on in every group
  std::vector<BatchRunDesc> hdesc;  // host copy of druns
  // every run's parameter block per step, [step][run] (the per-step
  // particle launch reads its runs' blocks through one kernel argument)
  DevBuf<ParticleParams> steptab;
  bool any_split = false;            // some run's particle pass takes K > 1 parts
  DevBuf<int2> dparts;              // (run, part) of each particle workgroup
  int nparts = 0;
  DevBuf<double> sred;              // split scratch of every run
  DevBuf<int4> sfb;
  DevBuf<BatchRunDesc> druns;
  DevBuf<int32_t> drec;
  std::vector<DevBuf<const double2*>> dadi;  // per group: factors of each step
  std::vector<DevBuf<double>> dcs;           // per group: coefficient of each step
  hipEvent_t pdone = nullptr;

  ~g4h_batch() {
    dblocks.release();
    dparts.release();
    sred.release();
    sfb.release();
    druns.release();
    drec.release();
    steptab.release();
    dadi.clear();
    dcs.clear();
    if (pdone) (void)hipEventDestroy(pdone);
    for (Group& g : groups) {
      if (g.field) g4h_field_destroy(g.field);
      if (g.beam) g4h_beam_destroy(g.beam);
      if (g.tail) (void)hipEventDestroy(g.tail);
      g.dtab.release();
      g.beam_hist.release();
      g.field_hist.release();
      if (g.gctx) g4h_context_destroy(g.gctx);
    }
  }

  size_t beam_rec(const Group& g) const {
    return g.runs.size() * static_cast<size_t>(G4H_BEAM_RAW(nharm));
  }
  size_t field_rec(const Group& g) const { return g.runs.size() * G4H_FIELD_RAW; }

  void record_initial() {
    for (Group& g : groups) {
      ParticleParams p{};
      p.gamma0 = g.gamma0;
      p.stages = ST_DIAG;
      p.diag.raw = g.beam_hist.get();
      p.diag.nharm = nharm;
      launch_particle(g.gctx, g.beam, p);
      field_plane_stats(g.field);
      field_map_stats(g.field, g.field_hist.get());
    }
    nrec = 1;
  }

  // Every step's parameter table of every group, built on the host with the
  // same make_trans / make_push / make_dep as every other path and uploaded
  // once, so that a step enqueues launches only.
  void build_tables() {
    const int ns = static_cast<int>(steps.size());
    std::vector<int32_t>& rec = rec_of_step;
    rec.assign(static_cast<size_t>(ns), -1);
    int r = 1;
    for (int k = 0; k < ns; ++k)
      if (steps[static_cast<size_t>(k)].outstep && r < cap) rec[static_cast<size_t>(k)] = r++;
    for (Group& g : groups) {
      const int n = static_cast<int>(g.runs.size());
      std::vector<ParticleParams> tab(static_cast<size_t>(ns) * n);
      g.head.resize(static_cast<size_t>(ns));
      g.src_on.assign(static_cast<size_t>(ns), 0);
      for (int k = 0; k < ns; ++k) {
        const g4h_step& st0 = steps[static_cast<size_t>(k)];
        const double delz = st0.delz;
        bool src_on = false;
        for (int i = 0; i < n; ++i) {
          g4h_step st = st0;
          const auto& a = aw[static_cast<size_t>(g.runs[static_cast<size_t>(i)])];
          if (!a.empty()) st.aw = a[static_cast<size_t>(k)];
          ParticleParams p{};
          p.gamma0 = g.gamma0;
          p.stages = ST_T1 | ST_PUSH | ST_T2;
          p.t1 = make_trans(st, g.gamma0, delz * 0.5, false);
          p.t2 = make_trans(st, g.gamma0, delz * 0.5, true);
          Field* fl[1] = {g.field};
          make_push(p.push, st, g.gamma0, delz, fl, 1, nullptr);
          p.r56 = make_r56(st, g.gamma0, lambda0);
          if (p.r56 != 0.0) p.stages |= ST_R56;
          const bool on = make_dep(p.dep, g.field, g.beam, st, delz) && g.beam->np > 0;
          require(i == 0 || on == src_on, "source switched on for part of a batch group");
          src_on = on;
          if (on) p.stages |= ST_DEP;
          const int rk = rec[static_cast<size_t>(k)];
          if (rk >= 0) {
            p.stages |= ST_DIAG;
            p.diag.raw = g.beam_hist.get() + static_cast<size_t>(rk) * beam_rec(g);
            p.diag.nharm = nharm;
          }
          tab[static_cast<size_t>(k) * n + i] = p;
        }
        g.head[static_cast<size_t>(k)] = tab[static_cast<size_t>(k) * n];
        g.src_on[static_cast<size_t>(k)] = src_on ? 1 : 0;
      }
      g.dtab.alloc(g.gctx, tab.size());
      G4_HIP(hipMemcpy(g.dtab.get(), tab.data(), sizeof(ParticleParams) * tab.size(),
                       hipMemcpyHostToDevice));
    }
  }

  // The one-launch path: per-group per-step factor tables (every distinct
  // step length's table stays resident: at most Field::kMaxDelzTables), run
  // descriptors and record indices on the device.  Returns false (and the
  // batch keeps the per-step launches) when the kernel does not fit.
  bool setup_persistent(Context* c) {
    if (!lengths_fit()) return false;
    if (!batch_persistent_fits(c, prec, nharm, win_cells, gmax)) return false;
    build_run_descs(c);
    G4_HIP(hipEventCreateWithFlags(&pdone, hipEventDisableTiming));
    persistent = true;
    return true;
  }

  // Whether every distinct step length's solver tables can stay resident
  // (at most Field::kMaxDelzTables); sets the largest grid and window.
  bool lengths_fit() {
    if (solver != G4H_SOLVER_ADI) return false;
    std::vector<double> lens;
    for (const g4h_step& st : steps) lens.push_back(st.delz);
    std::sort(lens.begin(), lens.end());
    lens.erase(std::unique(lens.begin(), lens.end()), lens.end());
    if (lens.size() > Field::kMaxDelzTables) return false;
    int wc = 0, gm = 0;
    for (const Group& g : groups) {
      gm = std::max(gm, g.G);
      for (const ParticleParams& p : g.head) wc = std::max(wc, p.dep.window_cap);
    }
    win_cells = wc;
    gmax = gm;
    return true;
  }

  // The per-step launches of every run: run descriptors as for the one-launch
  // path, plus each run's x-sweep output and partials, and one descriptor per
  // 64-line block of every plane.  Returns false (the batch keeps the
  // per-group launches) for grids above the segmented sweeps' limit or when
  // the source is on in some groups and off in others at some step.
  bool setup_grid(Context* c) {
    if (!lengths_fit() || gmax > 256) return false;
    const int ns = static_cast<int>(steps.size());
    src_all.assign(static_cast<size_t>(ns), 0);
    for (int k = 0; k < ns; ++k) {
      const uint8_t a = groups[0].src_on[static_cast<size_t>(k)];
      for (const Group& g : groups)
        if (g.src_on[static_cast<size_t>(k)] != a) return false;
      src_all[static_cast<size_t>(k)] = a;
    }
    build_run_descs(c);
    {
      const size_t nr = hdesc.size();
      steptab.alloc(c, static_cast<size_t>(ns) * nr);
      for (size_t r = 0; r < nr; ++r) {
        const BatchRunDesc& d = hdesc[r];
        G4_HIP(hipMemcpy2DAsync(steptab.get() + r, sizeof(ParticleParams) * nr, d.tab + d.slice,
                                sizeof(ParticleParams) * static_cast<size_t>(d.nrun),
                                sizeof(ParticleParams), static_cast<size_t>(ns),
                                hipMemcpyDeviceToDevice, c->stream));
      }
    }
    // Particle workgroups: a run of n particles takes K = ceil(n / 2048)
    // parts (at most 16), eight iterations of the 256-thread loop each, so
    // that a batch of runs of a few thousand particles fills the GPU.  Each
    // part adds its deposition window to the source plane with atomics.
    std::vector<int2> parts;
    std::vector<int> kof(hdesc.size());
    size_t ktot = 0;
    any_split = false;
    for (size_t r = 0; r < hdesc.size(); ++r) {
      const auto [gi, si] = where[r];
      const Beam* bm = groups[static_cast<size_t>(gi)].beam;
      const int np = bm->off_h[static_cast<size_t>(si) + 1] - bm->off_h[static_cast<size_t>(si)];
      kof[r] = std::max(1, std::min(16, (np + 2047) / 2048));
      any_split = any_split || kof[r] > 1;
      for (int q = 0; q < kof[r]; ++q) parts.push_back(make_int2(static_cast<int>(r), q));
      ktot += static_cast<size_t>(kof[r]);
    }
    // split scratch: each part adds its window to the source plane with
    // atomics (no window copies); fb holds each run's two boxes, empty
    const size_t red = batch_split_red_stride();
    sred.alloc(c, ktot * red);
    sfb.alloc(c, ktot);
    const std::vector<int4> empty(ktot, make_int4(0x7fffffff, 0x7fffffff, -1, -1));
    G4_HIP(hipMemcpyAsync(sfb.get(), empty.data(), sizeof(int4) * ktot, hipMemcpyHostToDevice,
                          c->stream));
    size_t at = 0;
    for (size_t r = 0; r < hdesc.size(); ++r) {
      BatchRunDesc& d = hdesc[r];
      d.K = kof[r];
      d.swin = nullptr;
      d.sred = sred.get() + at * red;
      d.sfb = sfb.get() + at;
      at += static_cast<size_t>(kof[r]);
    }
    G4_HIP(hipMemcpyAsync(druns.get(), hdesc.data(), sizeof(BatchRunDesc) * hdesc.size(),
                          hipMemcpyHostToDevice, c->stream));
    nparts = static_cast<int>(parts.size());
    dparts.alloc(c, parts.size());
    G4_HIP(hipMemcpyAsync(dparts.get(), parts.data(), sizeof(int2) * parts.size(),
                          hipMemcpyHostToDevice, c->stream));
    std::vector<BatchSegDesc> blocks;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      Group& g = groups[gi];
      Field* f = g.field;
      const int n = static_cast<int>(g.runs.size());
      const int nblk = (g.G + 63) / 64;
      for (int i = 0; i < n; ++i)
        for (int b = 0; b < nblk; ++b) {
          BatchSegDesc d{};
          d.u = f->data();
          d.ustar = field_adi_ustar(f);
          d.src = f->src();
          d.rect = f->src_rect.get();
          d.partials = f->col_partials.get();
          d.adi = dadi[gi].get();
          d.cs = dcs[gi].get();
          d.G = g.G;
          d.nblk = nblk;
          d.nrun = n;
          d.slice = i;
          d.blk = b;
          blocks.push_back(d);
        }
    }
    nblocks = static_cast<int>(blocks.size());
    dblocks.alloc(c, blocks.size());
    G4_HIP(hipMemcpyAsync(dblocks.get(), blocks.data(), sizeof(BatchSegDesc) * blocks.size(),
                          hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    G4_HIP(hipEventCreateWithFlags(&pdone, hipEventDisableTiming));
    grid = true;
    return true;
  }

  // Run descriptors on the device: each run's beam arrays, parameter table,
  // field planes, x-sweep output, partials, source, per-step factor tables
  // and coefficients (every distinct step length's tables stay resident) and
  // record arrays; record indices of the steps.
  void build_run_descs(Context* c) {
    const int ns = static_cast<int>(steps.size());
    std::vector<BatchRunDesc> desc(where.size());
    dadi.resize(groups.size());
    dcs.resize(groups.size());
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      Group& g = groups[gi];
      Field* f = g.field;
      std::vector<const double2*> fac(static_cast<size_t>(ns));
      std::vector<double> cs(static_cast<size_t>(ns));
      for (int k = 0; k < ns; ++k) {
        const double dz = steps[static_cast<size_t>(k)].delz;
        field_prepare_solver(f, dz);
        fac[static_cast<size_t>(k)] = f->adi_c;  // c, cbet, cwet contiguous
        cs[static_cast<size_t>(k)] = 0.25 * dz / (f->xks * f->dgrid * f->dgrid);
      }
      if (f->adi_y.size() < static_cast<size_t>(f->cells())) f->adi_y.alloc(f->ctx, f->cells());
      ensure_source(f);
      (void)field_adi_ustar(f);
      const int nblk = (g.G + 63) / 64;
      const size_t need = g.runs.size() * static_cast<size_t>(nblk) * 8 + 2 * g.runs.size();
      if (f->col_partials.size() < need) f->col_partials.alloc(f->ctx, need);
      dadi[gi].alloc(g.gctx, static_cast<size_t>(ns));
      dcs[gi].alloc(g.gctx, static_cast<size_t>(ns));
      G4_HIP(hipMemcpyAsync(dadi[gi].get(), fac.data(), sizeof(const double2*) * ns,
                            hipMemcpyHostToDevice, g.gctx->stream));
      G4_HIP(hipMemcpyAsync(dcs[gi].get(), cs.data(), sizeof(double) * ns, hipMemcpyHostToDevice,
                            g.gctx->stream));
      G4_HIP(hipStreamSynchronize(g.gctx->stream));
      Beam* bm = g.beam;
      for (size_t i = 0; i < g.runs.size(); ++i) {
        BatchRunDesc d{};
        if (prec == G4H_FP64) {
          d.x = bm->xd.get();
          d.y = bm->yd.get();
          d.px = bm->pxd.get();
          d.py = bm->pyd.get();
          d.g = bm->gd.get();
        } else {
          d.x = bm->xf.get();
          d.y = bm->yf.get();
          d.px = bm->pxf.get();
          d.py = bm->pyf.get();
          d.g = bm->gf.get();
        }
        d.th = bm->theta.get();
        d.off = bm->off.get();
        d.tab = g.dtab.get();
        d.u = f->data();
        d.ys = f->adi_y.get();
        d.src = f->src();
        d.rect = f->src_rect.get();
        d.adi = dadi[gi].get();
        d.cs = dcs[gi].get();
        d.field_hist = g.field_hist.get();
        d.partials = f->col_partials.get();
        d.slice = static_cast<int>(i);
        d.nrun = static_cast<int>(g.runs.size());
        d.G = g.G;
        d.K = 1;
        desc[static_cast<size_t>(g.runs[i])] = d;
      }
    }
    druns.alloc(c, desc.size());
    drec.alloc(c, rec_of_step.size());
    G4_HIP(hipMemcpyAsync(druns.get(), desc.data(), sizeof(BatchRunDesc) * desc.size(),
                          hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipMemcpyAsync(drec.get(), rec_of_step.data(), sizeof(int32_t) * rec_of_step.size(),
                          hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    hdesc = std::move(desc);
  }

  // Steps [k0, k1) with three or four launches per step on the batch's
  // stream, after and before the groups' own streams.
  void advance_grid(int k0, int k1) {
    for (Group& g : groups) {
      G4_HIP(hipEventRecord(g.tail, g.gctx->stream));
      G4_HIP(hipStreamWaitEvent(ctx->stream, g.tail, 0));
    }
    const int nruns = static_cast<int>(druns.size());
    for (int k = k0; k < k1; ++k) {
      const int r = rec_of_step[static_cast<size_t>(k)];
      const bool src = src_all[static_cast<size_t>(k)] != 0;
      launch_batch_step_particles(ctx, prec, nharm, druns.get(), nruns, dparts.get(), nparts,
                                  steptab.get() + static_cast<size_t>(k) * nruns, k, win_cells,
                                  r >= 0, any_split);
      launch_batch_adi_step(ctx, prec, dblocks.get(), nblocks, k, gmax, src, r >= 0, src);
      if (r >= 0) {
        launch_batch_field_records(ctx, druns.get(), nruns, r);
        ++nrec;
      }
    }
    G4_HIP(hipEventRecord(pdone, ctx->stream));
    for (Group& g : groups) G4_HIP(hipStreamWaitEvent(g.gctx->stream, pdone, 0));
  }

  // Steps [k0, k1) of every run in one launch on the batch's stream, after
  // and before the groups' own streams.
  void advance_persistent(int k0, int k1) {
    for (Group& g : groups) {
      G4_HIP(hipEventRecord(g.tail, g.gctx->stream));
      G4_HIP(hipStreamWaitEvent(ctx->stream, g.tail, 0));
    }
    launch_batch_persistent(ctx, prec, nharm, druns.get(), static_cast<int>(druns.size()),
                            drec.get(), k0, k1, win_cells, gmax);
    G4_HIP(hipEventRecord(pdone, ctx->stream));
    for (Group& g : groups) G4_HIP(hipStreamWaitEvent(g.gctx->stream, pdone, 0));
    for (int k = k0; k < k1; ++k)
      if (rec_of_step[static_cast<size_t>(k)] >= 0) ++nrec;
  }

  void step(int istep) {
    const g4h_step& st0 = steps[static_cast<size_t>(istep)];
    const bool out = st0.outstep && nrec < cap;
    for (Group& g : groups) {
      const size_t n = g.runs.size();
      field_prepare_solver(g.field, st0.delz);
      launch_particle(g.gctx, g.beam, g.head[static_cast<size_t>(istep)],
                      g.dtab.get() + static_cast<size_t>(istep) * n);
      field_solve(g.field, g.src_on[static_cast<size_t>(istep)] != 0, out);
      if (out) field_map_stats(g.field, g.field_hist.get() + static_cast<size_t>(nrec) * field_rec(g));
    }
    if (out) ++nrec;
  }
};

extern "C" {

g4h_status g4h_batch_create(g4h_context* ctx, g4h_precision precision, g4h_solver_kind solver,
                            int nruns, const g4h_batch_run* runs, const g4h_step* steps,
                            int nsteps, double lambda0, int nharm, g4h_batch** out) {
  G4_API_BEGIN
  require(ctx && runs && steps && out, "null argument");
  *out = nullptr;
  require(nruns >= 1 && nsteps >= 1, "need at least one run and one step");
  require(nharm >= 1 && nharm <= 8, "nharm must be in [1, 8]");
  require(lambda0 > 0, "lambda0 must be positive");
  require(precision == G4H_FP64 || precision == G4H_MIXED, "bad precision");
  require(solver == G4H_SOLVER_ADI || solver == G4H_SOLVER_FFT, "bad solver");
  for (int i = 0; i < nsteps; ++i) require(steps[i].delz > 0, "step length must be positive");
  ctx->activate();
  auto b = std::make_unique<g4h_batch>();
  b->ctx = ctx;
  b->prec = precision;
  b->solver = solver;
  b->lambda0 = lambda0;
  b->nharm = nharm;
  b->steps.assign(steps, steps + nsteps);
  int nout = 0;
  for (int i = 0; i < nsteps; ++i) nout += steps[i].outstep ? 1 : 0;
  b->cap = 1 + nout;
  // group by (ngrid, gridmax, gamma0), runs in request order within a group
  std::map<std::tuple<int, double, double>, int> key;
  b->where.resize(static_cast<size_t>(nruns));
  b->aw.resize(static_cast<size_t>(nruns));
  for (int r = 0; r < nruns; ++r) {
    const g4h_batch_run& q = runs[r];
    require(q.npart >= 1 && q.nbins >= 1 && q.npart % q.nbins == 0,
            "npart must be a positive multiple of nbins");
    require(q.ngrid >= 3 && q.ngrid <= 1024 && q.gridmax > 0, "bad grid");
    require(q.gamma0 > 1 && q.current >= 0, "bad beam");
    const auto k = std::make_tuple(q.ngrid, q.gridmax, q.gamma0);
    auto it = key.find(k);
    if (it == key.end()) {
      it = key.emplace(k, static_cast<int>(b->groups.size())).first;
      Group g;
      g.G = q.ngrid;
      g.gridmax = q.gridmax;
      g.gamma0 = q.gamma0;
      b->groups.push_back(std::move(g));
    }
    Group& g = b->groups[static_cast<size_t>(it->second)];
    b->where[static_cast<size_t>(r)] = {it->second, static_cast<int>(g.runs.size())};
    g.runs.push_back(r);
    if (q.aw) b->aw[static_cast<size_t>(r)].assign(q.aw, q.aw + nsteps);
  }
  for (Group& g : b->groups) {
    const int n = static_cast<int>(g.runs.size());
    std::vector<int32_t> off(static_cast<size_t>(n) + 1, 0);
    std::vector<double> cur(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
      const g4h_batch_run& q = runs[g.runs[static_cast<size_t>(i)]];
      off[static_cast<size_t>(i) + 1] = off[static_cast<size_t>(i)] + q.npart;
      cur[static_cast<size_t>(i)] = q.current;
    }
    const size_t np = static_cast<size_t>(off.back());
    std::vector<double> x(np), y(np), px(np), py(np), ga(np), th(np);
    std::vector<double> re(static_cast<size_t>(n) * g.G * g.G), im(re.size());
    for (int i = 0; i < n; ++i) {
      const g4h_batch_run& q = runs[g.runs[static_cast<size_t>(i)]];
      g4h_beam_spec s{};
      s.gamma0 = q.gamma0;
      s.espread_rel = q.espread_rel;
      s.emitn_x = q.emitn_x;
      s.emitn_y = q.emitn_y;
      s.betax = q.betax;
      s.betay = q.betay;
      s.alphax = q.alphax;
      s.alphay = q.alphay;
      s.slicelength = lambda0;
      s.nslice = 1;
      s.npart = q.npart;
      s.nbins = q.nbins;
      s.shotnoise = q.shotnoise;
      s.seed = q.seed;
      const size_t o = static_cast<size_t>(off[static_cast<size_t>(i)]);
      ok(g4h_load_beam(&s, &q.current, x.data() + o, y.data() + o, px.data() + o, py.data() + o,
                       ga.data() + o, th.data() + o));
      if (q.seed_power > 0) {
        const size_t po = static_cast<size_t>(i) * g.G * g.G;
        ok(g4h_seed_gaussian(1, g.G, g.gridmax, lambda0, q.seed_power, q.seed_w0, re.data() + po,
                             im.data() + po));
      }
    }
    ok(g4h_context_create(ctx->device, &g.gctx));
    ok(g4h_beam_create(g.gctx, precision, n, off.data(), cur.data(), g.gamma0, &g.beam));
    ok(g4h_beam_upload(g.beam, x.data(), y.data(), px.data(), py.data(), ga.data(), th.data()));
    ok(g4h_field_create(g.gctx, precision, n, g.G, g.gridmax, lambda0, lambda0, 1, &g.field));
    ok(g4h_field_upload(g.field, re.data(), im.data(), 0));
    ok(g4h_field_init_solver(g.field, solver, 0, 1.0, 1.0, 1.0, 0));
    G4_HIP(hipEventCreateWithFlags(&g.tail, hipEventDisableTiming));
    g.beam_hist.alloc(g.gctx, static_cast<size_t>(b->cap) * b->beam_rec(g));
    g.field_hist.alloc(g.gctx, static_cast<size_t>(b->cap) * b->field_rec(g));
  }
  b->build_tables();
  b->record_initial();
  const int mode = ctx->tuner.forced[TUNE_BATCH];
  if (mode == 1)
    b->setup_persistent(ctx);
  else if (mode <= 0)
    b->setup_grid(ctx);
  *out = b.release();
  context_child_added(ctx);
  G4_API_END
}

void g4h_batch_destroy(g4h_batch* b) {
  if (!b) return;
  Context* ctx = b->ctx;
  (void)hipSetDevice(ctx->device);
  for (const Group& g : b->groups)
    if (g.gctx) (void)hipStreamSynchronize(g.gctx->stream);
  delete b;
  context_child_removed(ctx);
}

g4h_status g4h_batch_advance(g4h_batch* b, int nsteps, int* done) {
  G4_API_BEGIN
  require(b, "batch is null");
  require(nsteps >= 0, "nsteps must be >= 0");
  b->ctx->activate();
  const int end = std::min<int>(static_cast<int>(b->steps.size()), b->done + nsteps);
  if (b->persistent) {
    if (end > b->done) b->advance_persistent(b->done, end);
    b->done = end;
  }
  if (b->grid) {
    if (end > b->done) b->advance_grid(b->done, end);
    b->done = end;
  }
  for (; b->done < end; ++b->done) b->step(b->done);
  // the context's stream waits for every group, so work enqueued on it later
  // (or a synchronisation of it) comes after this advance
  for (Group& g : b->groups) {
    G4_HIP(hipEventRecord(g.tail, g.gctx->stream));
    G4_HIP(hipStreamWaitEvent(b->ctx->stream, g.tail, 0));
  }
  if (done) *done = b->done;
  G4_API_END
}

g4h_status g4h_batch_records(const g4h_batch* b, int* nrecords) {
  G4_API_BEGIN
  require(b && nrecords, "null argument");
  *nrecords = b->nrec;
  G4_API_END
}

g4h_status g4h_batch_download(g4h_batch* b, int run, double* beam_raw, double* field_raw) {
  G4_API_BEGIN
  require(b, "batch is null");
  require(run >= 0 && run < static_cast<int>(b->where.size()), "run out of range");
  b->ctx->activate();
  const auto [gi, si] = b->where[static_cast<size_t>(run)];
  const Group& g = b->groups[static_cast<size_t>(gi)];
  G4_HIP(hipStreamSynchronize(g.gctx->stream));
  const size_t br = G4H_BEAM_RAW(b->nharm);
  for (int r = 0; r < b->nrec; ++r) {
    if (beam_raw)
      G4_HIP(hipMemcpy(beam_raw + static_cast<size_t>(r) * br,
                       g.beam_hist.get() + static_cast<size_t>(r) * b->beam_rec(g) + si * br,
                       sizeof(double) * br, hipMemcpyDeviceToHost));
    if (field_raw)
      G4_HIP(hipMemcpy(field_raw + static_cast<size_t>(r) * G4H_FIELD_RAW,
                       g.field_hist.get() + static_cast<size_t>(r) * b->field_rec(g) +
                           static_cast<size_t>(si) * G4H_FIELD_RAW,
                       sizeof(double) * G4H_FIELD_RAW, hipMemcpyDeviceToHost));
  }
  G4_API_END
}

}  // extern "C"
