// genesis_gpu.hpp - C++ host layer over the g4hip C ABI with the shape of the
// reference's classes, so existing GENESIS 1.3 call sites keep their form.
//
//   reference (mulingLHY/Genesis4_CUDA_port)        this header
//   Genesis4BeamSoA + Beam::pack/unpack           -> g4gpu::Beam
//   Genesis4FieldSoA + Field::pack/unpack/track   -> g4gpu::Field
//   FieldSolverADICUDA / FieldSolverFFTCUDA       -> g4gpu::FieldSolverADIGPU / FieldSolverFFTGPU
//   BeamSolverCUDA (advance, track, applyR56)     -> g4gpu::BeamSolverGPU
//   TrackBeamCUDA (track, applyR56, ...)          -> g4gpu::TrackBeamGPU
//   ControlCUDA::applySlippage                    -> g4gpu::ControlGPU::applySlippage
//   DiagnosticCUDA::calc                          -> g4gpu::DiagnosticGPU::calc
//   Gencore::run                                  -> g4gpu::Gencore::run
//   upstream Undulator step accessors             -> g4gpu::Undulator (step table)
//
// Errors: every ABI status other than G4H_OK throws g4gpu::Error, the way the
// reference's genesis4_cuda::check throws std::runtime_error
// (include/Genesis4CudaRuntime.h:13-23); the bool-returning calls (applySlippage,
// calc) return false instead, which makes the reference's CPU fallback in
// Gencore.cpp:81-101 fire.
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/g4hip.h"

namespace g4gpu {

struct Error : std::runtime_error {
  g4h_status status;
  Error(g4h_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(g4h_status s) {
  if (s != G4H_OK) throw Error(s, g4h_last_error());
}

class Context {
 public:
  explicit Context(int device = 0) { check(g4h_context_create(device, &h_)); }
  ~Context() { g4h_context_destroy(h_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  g4h_context* get() const { return h_; }
  void synchronize() { check(g4h_context_synchronize(h_)); }

 private:
  g4h_context* h_ = nullptr;
};

// Per-step lattice table with the accessors the reference calls on the
// upstream Undulator (BeamSolverCUDA.cpp:503-531, TrackBeamCUDA.cpp:163-167,
// FieldSolverADICUDA.cpp:697-730, Gencore.cpp:172-174,229,247).
class Undulator {
 public:
  Undulator(std::vector<g4h_step> steps, double gamma_ref)
      : steps_(std::move(steps)), gamma_ref_(gamma_ref) {}
  bool advance() { return ++istep_ < static_cast<int>(steps_.size()); }
  int getStep() const { return istep_; }
  const g4h_step& step() const { return steps_.at(static_cast<size_t>(istep_)); }
  double steplength() const { return step().delz; }
  double getz() const { return istep_ < 0 ? 0.0 : step().z; }
  double getaw() const { return step().aw; }
  double getku() const { return step().ku; }
  double autophase() const { return step().autophase; }
  double slippage() const { return step().slippage; }
  bool inUndulator() const { return step().in_undulator != 0; }
  bool outstep() const { return step().outstep != 0; }
  int getMarker() const { return step().marker; }
  double getGammaRef() const { return gamma_ref_; }
  double fc(int harm) const { return g4h_undulator_fc(step().aw, step().helical, harm); }
  const std::vector<g4h_step>& table() const { return steps_; }
  void reset() { istep_ = -1; }

 private:
  std::vector<g4h_step> steps_;
  double gamma_ref_;
  int istep_ = -1;
};

// Device-resident beam.  Members named as in the reference's Beam.
class Beam {
 public:
  Beam(Context& ctx, g4h_precision prec, std::vector<int32_t> slice_offsets,
       std::vector<double> current_in, double gamma_ref, double reflength_in,
       double slicelength_in)
      : offsets(std::move(slice_offsets)), current(std::move(current_in)),
        reflength(reflength_in), slicelength(slicelength_in), gamma_ref(gamma_ref), ctx_(ctx) {
    check(g4h_beam_create(ctx.get(), prec, static_cast<int>(offsets.size()) - 1, offsets.data(),
                          current.data(), gamma_ref, &soa_));
  }
  ~Beam() { g4h_beam_destroy(soa_); }
  Beam(const Beam&) = delete;
  Beam& operator=(const Beam&) = delete;

  // Beam::pack_beam_to_soa / unpack_soa_to_beam (Beam.cpp:392-525)
  void pack(const std::vector<double>& x, const std::vector<double>& y,
            const std::vector<double>& px, const std::vector<double>& py,
            const std::vector<double>& gamma, const std::vector<double>& theta) {
    check(g4h_beam_upload(soa_, x.data(), y.data(), px.data(), py.data(), gamma.data(),
                          theta.data()));
  }
  void unpack(std::vector<double>& x, std::vector<double>& y, std::vector<double>& px,
              std::vector<double>& py, std::vector<double>& gamma, std::vector<double>& theta) {
    const size_t n = static_cast<size_t>(offsets.back());
    for (auto* v : {&x, &y, &px, &py, &gamma, &theta}) v->resize(n);
    check(g4h_beam_download(soa_, x.data(), y.data(), px.data(), py.data(), gamma.data(),
                            theta.data()));
  }
  // Beam::initEField (include/Beam.h:126-131): short-range space charge of
  // this beam, computed on the device before every push
  void initEField(double rmax, int ngrid, int nz, int nphi, double lambda, bool longr) {
    const g4h_efield ef{rmax, ngrid, nz, nphi, longr ? 1 : 0};
    check(g4h_beam_init_efield(soa_, &ef, lambda));
  }
  // Beam::getSCField (include/Beam.h:139-144): RMS short-range E_z [V/m] over
  // slice islice at the last push
  double getSCField(int islice) const {
    std::vector<double> v(static_cast<size_t>(nslice()));
    check(g4h_beam_sc_field(soa_, v.data()));
    return v.at(static_cast<size_t>(islice));
  }
  int nslice() const { return static_cast<int>(offsets.size()) - 1; }
  g4h_beam* soa() const { return soa_; }
  Context& ctx() const { return ctx_; }

  std::vector<int32_t> offsets;
  std::vector<double> current;
  double reflength, slicelength, gamma_ref;

 private:
  Context& ctx_;
  g4h_beam* soa_ = nullptr;
};

class Field;

// FieldSolver interface (upstream, overridden at include/FieldSolverADICUDA.h:23-25).
class FieldSolver {
 public:
  virtual ~FieldSolver() = default;
  virtual void init(double delz, double dgrid, double xks, unsigned ngrid) = 0;
  virtual void advance(double delz, Field* field, Beam* beam, Undulator* und) = 0;
  virtual void initSourceFilter(double xc, double yc, double sig, bool filter) = 0;
};

class Field {
 public:
  Field(Context& ctx, g4h_precision prec, int nslice, int ngrid, double gridmax,
        double xlambda0, double slicelength, int harm = 1)
      : ngrid(ngrid), harm(harm), gridmax(gridmax), dgrid(2 * gridmax / (ngrid - 1)),
        xlambda(xlambda0 / harm), xks(2.0 * 3.14159265358979323846 / (xlambda0 / harm)),
        slicelength(slicelength), ctx_(ctx), nslice_(nslice) {
    check(g4h_field_create(ctx.get(), prec, nslice, ngrid, gridmax, xlambda0, slicelength, harm,
                           &soa_));
  }
  ~Field() {
    delete solver_;
    g4h_field_destroy(soa_);
  }
  Field(const Field&) = delete;
  Field& operator=(const Field&) = delete;

  // Field::initSolver (Field.cpp:87-110); use_gpu=false is not served here.
  void initSolver(bool isFFT, bool filter, double xc, double yc, double sig, bool use_gpu = true,
                  bool low_memory_adisolver = false);
  // Field::track (Field.cpp:204-212)
  void track(double delz, Beam* beam, Undulator* und) {
    solver_->init(delz, dgrid, xks, static_cast<unsigned>(ngrid));
    solver_->advance(delz, this, beam, und);
  }
  void pack(const std::vector<double>& re, const std::vector<double>& im, int first_in = 0) {
    check(g4h_field_upload(soa_, re.data(), im.data(), first_in));
  }
  void unpack(std::vector<double>& re, std::vector<double>& im) {
    const size_t n = static_cast<size_t>(nslice_) * ngrid * ngrid;
    re.resize(n);
    im.resize(n);
    int f = 0;
    check(g4h_field_download(soa_, re.data(), im.data(), &f));
  }
  int first() const {
    int f = 0;
    check(g4h_field_get_ring(soa_, &f, nullptr));
    return f;
  }
  double accuslip() const {
    double a = 0;
    check(g4h_field_get_ring(soa_, nullptr, &a));
    return a;
  }
  int getHarm() const { return harm; }
  int nslice() const { return nslice_; }
  g4h_field* soa() const { return soa_; }
  Context& ctx() const { return ctx_; }

  int ngrid, harm;
  double gridmax, dgrid, xlambda, xks, slicelength;

 private:
  Context& ctx_;
  int nslice_;
  g4h_field* soa_ = nullptr;
  FieldSolver* solver_ = nullptr;
};

// FieldSolverADICUDA / FieldSolverFFTCUDA: the coefficient tables are cached
// inside the library per (delz, ngrid), so init() has nothing to do; advance()
// is Field::track's device work (deposition + one solver step).
class FieldSolverGPU : public FieldSolver {
 public:
  explicit FieldSolverGPU(g4h_solver_kind kind, bool low_memory = false)
      : kind_(kind), low_memory_(low_memory) {}
  void init(double, double, double, unsigned) override {}
  void advance(double delz, Field* field, Beam* beam, Undulator* und) override {
    configure(field);
    check(g4h_field_track(field->ctx().get(), field->soa(), beam->soa(), &und->step(), delz));
  }
  void initSourceFilter(double xc, double yc, double sig, bool filter) override {
    xc_ = xc;
    yc_ = yc;
    sig_ = sig;
    filter_ = filter;
    configured_ = nullptr;
  }
  // set the field's device solver now (Field::initSolver), so that the fused
  // loop of Gencore::run, which never calls advance(), runs this solver
  void bind(Field* f) { configure(f); }

 private:
  void configure(Field* f) {
    if (configured_ == f) return;
    check(g4h_field_init_solver(f->soa(), kind_, filter_ ? 1 : 0, xc_, yc_, sig_,
                                low_memory_ ? 1 : 0));
    configured_ = f;
  }
  g4h_solver_kind kind_;
  bool low_memory_;
  bool filter_ = false;
  double xc_ = 1, yc_ = 1, sig_ = 1;
  Field* configured_ = nullptr;
};

class FieldSolverADIGPU : public FieldSolverGPU {
 public:
  explicit FieldSolverADIGPU(bool low_memory = false) : FieldSolverGPU(G4H_SOLVER_ADI, low_memory) {}
};
class FieldSolverFFTGPU : public FieldSolverGPU {
 public:
  FieldSolverFFTGPU() : FieldSolverGPU(G4H_SOLVER_FFT) {}
};

inline void Field::initSolver(bool isFFT, bool filter, double xc, double yc, double sig,
                              bool use_gpu, bool low_memory_adisolver) {
  if (!use_gpu) throw Error(G4H_ERR_UNSUPPORTED, "use_gpu=false: use the upstream CPU solvers");
  delete solver_;
  solver_ = nullptr;
  FieldSolverGPU* s = nullptr;
  if (isFFT) {
    s = new FieldSolverFFTGPU();
    s->initSourceFilter(xc, yc, sig, filter);
  } else {
    s = new FieldSolverADIGPU(low_memory_adisolver);
  }
  solver_ = s;
  s->bind(this);
}

// TrackBeamCUDA (TrackBeamCUDA.cpp:156-278)
class TrackBeamGPU {
 public:
  void track(double delz, Beam* beam, Undulator* und, bool lastStep) {
    check(g4h_track_transverse(beam->ctx().get(), beam->soa(), &und->step(), delz, lastStep ? 1 : 0));
  }
  void applyR56(Beam* beam, Undulator* und, double lambda0) {
    check(g4h_apply_r56(beam->ctx().get(), beam->soa(), &und->step(), lambda0));
  }
};

// BeamSolverCUDA (BeamSolverCUDA.cpp:449-687, include/BeamSolverCUDA.h:19-44)
class BeamSolverGPU {
 public:
  // BeamSolverCUDA::initEField (include/BeamSolverCUDA.h:42-44): settings of
  // the short-range space charge, given to the beam at the next advance (the
  // device beam computes the field itself before its push)
  void initEField(double rmax, int ngrid, int nz, int nphi, double lambda, bool longr) {
    ef_ = g4h_efield{rmax, ngrid, nz, nphi, longr ? 1 : 0};
    ef_lambda_ = lambda;
    ef_pending_ = true;
  }
  // BeamSolverCUDA::checkAllocation (BeamSolverCUDA.cpp:685-687): room for
  // the short-range field of nslice slices
  void checkAllocation(unsigned long nslice) { sc_.assign(nslice, 0.0); }
  // BeamSolverCUDA::getSCField (include/BeamSolverCUDA.h:39): RMS
  // short-range E_z [V/m] over slice islice at the last advance (0 when
  // space charge is off or the slice is beyond the allocation)
  double getSCField(int islice) {
    if (sc_stale_ && beam_) {
      sc_.resize(static_cast<size_t>(beam_->nslice()));
      check(g4h_beam_sc_field(beam_->soa(), sc_.data()));
      sc_stale_ = false;
    }
    return islice >= 0 && static_cast<size_t>(islice) < sc_.size() ? sc_[static_cast<size_t>(islice)]
                                                                    : 0.0;
  }
  // slice_eloss: -longESC/511000 per slice (:555); particle_ez: the
  // short-range field per particle plus that loss, beam order (:523-547),
  // which replaces it when given
  void advance(double delz, Beam* beam, std::vector<Field*>* fields, Undulator* und,
               const double* slice_eloss = nullptr, const double* particle_ez = nullptr) {
    std::vector<g4h_field*> fl;
    for (Field* f : *fields)
      if (f) fl.push_back(f->soa());
    if (ef_pending_) {
      check(g4h_beam_init_efield(beam->soa(), &ef_, ef_lambda_));
      ef_pending_ = false;
    }
    check(g4h_beam_set_particle_ez(beam->soa(), particle_ez));
    check(g4h_beam_advance(beam->ctx().get(), beam->soa(), fl.data(), static_cast<int>(fl.size()),
                           &und->step(), delz, particle_ez ? nullptr : slice_eloss));
    beam_ = beam;
    sc_stale_ = true;
  }
  // the energy change per slice of Collective::apply / Incoherent::apply
  // (Beam.cpp:127-128), applied to the device beam
  void applyEnergyChange(Beam* beam, const double* dgamma) {
    check(g4h_beam_add_energy(beam->ctx().get(), beam->soa(), dgamma));
  }
  void track(double delz, Beam* beam, Undulator* und, bool last) { tracker_.track(delz, beam, und, last); }
  void applyR56(Beam* beam, Undulator* und, double reflen) { tracker_.applyR56(beam, und, reflen); }

 private:
  TrackBeamGPU tracker_;
  g4h_efield ef_{};
  double ef_lambda_ = 0.0;
  bool ef_pending_ = false;
  Beam* beam_ = nullptr;
  std::vector<double> sc_;
  bool sc_stale_ = false;
};

// ControlCUDA (ControlCUDA.cpp:122-252), one rank
class ControlGPU {
 public:
  void init(int rank, int size, bool isTime, bool periodic, double sample) {
    rank_ = rank;
    size_ = size;
    time_ = isTime;
    periodic_ = periodic;
    sample_ = sample;
  }
  // Returns false (reference fallback) instead of throwing.
  bool applySlippage(double slippage, Field* field) {
    if (!field || size_ != 1) return false;
    return g4h_apply_slippage(field->ctx().get(), field->soa(), slippage, sample_, time_ ? 1 : 0,
                              periodic_ ? 1 : 0, nullptr) == G4H_OK;
  }

 private:
  int rank_ = 0, size_ = 1;
  bool time_ = false, periodic_ = false;
  double sample_ = 1.0;
};

// One output record of DiagnosticCUDA::calc: the named per-slice values of
// DiagnosticCUDA.cpp:453-507 and :665-683 (layouts in g4hip_types.h).
struct DiagnosticRecord {
  double z = 0;
  std::vector<double> beam, beam_global;                 // nslice x G4H_BEAM_QN, 6
  std::vector<double> sscfield;                          // nslice (DiagnosticCUDA.cpp:570)
  std::vector<std::vector<double>> field, field_global;  // per field: nslice x 9, 7
};

class DiagnosticGPU {
 public:
  void init(int nharm) { nharm_ = nharm; }
  bool calc(Beam* beam, std::vector<Field*>* fields, double z, DiagnosticRecord& out) {
    const int S = beam->nslice();
    std::vector<double> raw(static_cast<size_t>(S) * G4H_BEAM_RAW(nharm_));
    if (g4h_diag_beam_raw(beam->ctx().get(), beam->soa(), nharm_, raw.data()) != G4H_OK) return false;
    out.z = z;
    out.beam.resize(static_cast<size_t>(S) * G4H_BEAM_QN(nharm_));
    out.beam_global.resize(G4H_BEAM_GLOBAL);
    check(g4h_diag_finalize_beam(S, nharm_, raw.data(), beam->current.data(), beam->gamma_ref,
                                 out.beam.data(), out.beam_global.data()));
    out.sscfield.resize(static_cast<size_t>(S));
    check(g4h_beam_sc_field(beam->soa(), out.sscfield.data()));
    out.field.assign(fields->size(), {});
    out.field_global.assign(fields->size(), {});
    for (size_t i = 0; i < fields->size(); ++i) {
      Field* f = (*fields)[i];
      std::vector<double> fr(static_cast<size_t>(f->nslice()) * G4H_FIELD_RAW);
      if (g4h_diag_field_raw(f->ctx().get(), f->soa(), fr.data()) != G4H_OK) return false;
      out.field[i].resize(static_cast<size_t>(f->nslice()) * G4H_FIELD_Q);
      out.field_global[i].resize(G4H_FIELD_GLOBAL);
      check(g4h_diag_finalize_field(f->nslice(), f->dgrid, f->xlambda, f->slicelength, fr.data(),
                                    out.field[i].data(), out.field_global[i].data()));
    }
    return true;
  }

 private:
  int nharm_ = 1;
};

// Gencore::run (Gencore.cpp:161-254): with one field (and `fused`) the whole
// loop runs as the fused device loop (g4h_run); otherwise it is driven
// operator by operator exactly in the reference's order.
class Gencore {
 public:
  static std::vector<DiagnosticRecord> run(Beam* beam, std::vector<Field*>* fields,
                                           Undulator* und, bool isTime, bool periodic,
                                           int nharm = 1, bool fused = true) {
    std::vector<DiagnosticRecord> hist;
    const double sample = beam->slicelength / beam->reflength;
    if (fused && fields->size() == 1) {
      Field* f = (*fields)[0];
      g4h_run_config cfg{beam->reflength, sample, isTime ? 1 : 0, periodic ? 1 : 0, nharm, 1};
      g4h_run* r = nullptr;
      check(g4h_run_create(beam->ctx().get(), beam->soa(), f->soa(), und->table().data(),
                           static_cast<int>(und->table().size()), &cfg, &r));
      check(g4h_run_advance(r, static_cast<int>(und->table().size()), nullptr));
      int n = 0;
      check(g4h_run_records(r, &n));
      const int S = beam->nslice();
      std::vector<double> braw(static_cast<size_t>(n) * S * G4H_BEAM_RAW(nharm));
      std::vector<double> fraw(static_cast<size_t>(n) * S * G4H_FIELD_RAW);
      check(g4h_run_download_diag(r, braw.data(), fraw.data()));
      g4h_run_destroy(r);
      std::vector<double> zs{0.0};
      for (const auto& st : und->table())
        if (st.outstep) zs.push_back(st.z);
      for (int k = 0; k < n; ++k) {
        DiagnosticRecord rec;
        rec.z = zs[static_cast<size_t>(k)];
        rec.beam.resize(static_cast<size_t>(S) * G4H_BEAM_QN(nharm));
        rec.beam_global.resize(G4H_BEAM_GLOBAL);
        check(g4h_diag_finalize_beam(S, nharm, braw.data() + static_cast<size_t>(k) * S * G4H_BEAM_RAW(nharm),
                                     beam->current.data(), beam->gamma_ref, rec.beam.data(),
                                     rec.beam_global.data()));
        rec.field.assign(1, std::vector<double>(static_cast<size_t>(S) * G4H_FIELD_Q));
        rec.field_global.assign(1, std::vector<double>(G4H_FIELD_GLOBAL));
        check(g4h_diag_finalize_field(S, f->dgrid, f->xlambda, f->slicelength,
                                      fraw.data() + static_cast<size_t>(k) * S * G4H_FIELD_RAW,
                                      rec.field[0].data(), rec.field_global[0].data()));
        hist.push_back(std::move(rec));
      }
      return hist;
    }
    ControlGPU control;
    control.init(0, 1, isTime, periodic, sample);
    DiagnosticGPU diag;
    diag.init(nharm);
    BeamSolverGPU solver;
    DiagnosticRecord rec;
    if (diag.calc(beam, fields, und->getz(), rec)) hist.push_back(rec);
    und->reset();
    while (und->advance()) {
      const double delz = und->steplength();
      solver.track(delz * 0.5, beam, und, false);  // Beam::track (Beam.cpp:123-132)
      solver.advance(delz, beam, fields, und);
      solver.applyR56(beam, und, beam->reflength);
      solver.track(delz * 0.5, beam, und, true);
      for (Field* f : *fields) f->track(delz, beam, und);
      for (Field* f : *fields) control.applySlippage(und->slippage(), f);
      if (und->outstep() && diag.calc(beam, fields, und->getz(), rec)) hist.push_back(rec);
    }
    return hist;
  }
};

}  // namespace g4gpu
