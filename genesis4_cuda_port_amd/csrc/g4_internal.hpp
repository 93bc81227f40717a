// g4_internal.hpp - internal state behind the g4hip C ABI.
//
// Device layout (all in HBM, one allocation per array):
//   beam  FP64 : x,y,px,py,gamma,theta  double[N]
//   beam  MIXED: x,y,px,py,dgamma float[N] (dgamma = gamma - gamma_ref), theta double[N]
//   field       : interleaved complex (double2 / float2) [nslice*ngrid*ngrid],
//                 plane p = ring slot p, cell ix + iy*ngrid
//   source      : same element type, indexed by BEAM slice (not ring slot), with a
//                 per-slice half-open rectangle outside which it is zero
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/g4hip.h"

namespace g4 {

// Error carried to the ABI boundary and turned into a status code there.
struct Error : std::runtime_error {
  g4h_status code;
  Error(g4h_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(g4h_status c, const std::string& m) { throw Error(c, m); }

inline void require(bool ok, const char* what) {
  if (!ok) fail(G4H_ERR_INVALID_ARGUMENT, what);
}

inline void hip_check(hipError_t e, const char* where) {
  if (e != hipSuccess) {
    const g4h_status c = (e == hipErrorOutOfMemory) ? G4H_ERR_OUT_OF_MEMORY : G4H_ERR_DEVICE;
    fail(c, std::string(where) + ": " + hipGetErrorString(e));
  }
}
#define G4_HIP(expr) ::g4::hip_check((expr), #expr)

void set_last_error(const std::string& m);

// Body of every extern "C" entry point: no exception crosses the ABI; the
// status is returned and the message kept for g4h_last_error().
#define G4_API_BEGIN try {
#define G4_API_END                                \
  return G4H_OK;                                  \
  }                                               \
  catch (const g4::Error& e) {                    \
    g4::set_last_error(e.what());                 \
    return e.code;                                \
  }                                               \
  catch (const std::bad_alloc&) {                 \
    g4::set_last_error("host allocation failed"); \
    return G4H_ERR_OUT_OF_MEMORY;                 \
  }                                               \
  catch (const std::exception& e) {               \
    g4::set_last_error(e.what());                 \
    return G4H_ERR_DEVICE;                        \
  }

// Per-context byte accounting so a caller (or a serving loop) can admit work
// against an HBM budget.
struct Context;

// RAII device allocation of `n` elements of T.
template <class T>
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept { swap(o); }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      swap(o);
    }
    return *this;
  }
  ~DevBuf() { release(); }

  void alloc(Context* ctx, size_t n);
  void release();
  T* get() const { return p_; }
  size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }

 private:
  void swap(DevBuf& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    std::swap(ctx_, o.ctx_);
  }
  T* p_ = nullptr;
  size_t n_ = 0;
  Context* ctx_ = nullptr;
};

// Kernel families with interchangeable variants, chosen per shape by the
// tuner of the context.  Variants of a family produce bit-identical results,
// so the choice changes only the time a step takes.
enum TuneFamily : int {
  TUNE_ADI = 0,    // 0: workgroup per plane (LDS-staged), 1: one wave per 64 lines
  TUNE_SPLIT = 1,  // workgroups per slice of the particle pass: not tuned, chosen
                   // from the slice count; pinning sets it (>= 1)
  TUNE_BATCH = 2,  // ADI batches: 0 one launch per kernel and step for every run
                   // (default), 1 one launch for all steps, 2 per-grid launches
  TUNE_COUNT
};

// Autotuner: for each (family, shape) the variants are tried in turn on live
// launches (kReps times each), timed with events that the host waits for
// once, after the last trial, and the fastest is kept for the rest of the
// context's life.  A family can be pinned to one variant.
struct Tuner {
  static constexpr int kReps = 3;
  struct Trial {
    int variant;
    hipEvent_t a, b;
  };
  struct Entry {
    int nvar = 0;
    int choice = -1;  // decided variant, -1 while measuring
    int issued = 0;   // trials launched so far
    std::vector<Trial> trials;
  };
  std::map<std::vector<int>, Entry> table;
  int forced[TUNE_COUNT] = {-1, -1, -1};
};

struct Context {
  int device = 0;
  int num_cus = 0;  // compute units of the device (queried on first use)
  hipStream_t stream = nullptr;
  uint64_t bytes = 0;
  // Objects (beams, fields, runs) created from this context.  Destroying the
  // context while some are alive defers the release until the last one goes.
  int children = 0;
  bool destroy_pending = false;
  // stream capture into a hipGraph in progress on this context's stream:
  // no profiling events, no tuner trials, no allocations, parameter blocks
  // from the capture arena
  bool capturing = false;
  // profiling
  bool profile = false;
  struct Prof {
    int64_t launches = 0;
    double ms = 0.0;
  } prof[G4H_K_COUNT];
  struct Pending {
    hipEvent_t a, b;
    int cls;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;

  void activate() const { G4_HIP(hipSetDevice(device)); }
  hipEvent_t get_event();
  // Bracket a launch with events when profiling is on.
  template <class F>
  void launch(int cls, F&& f) {
    if (!profile || capturing) {
      f();
      G4_HIP(hipGetLastError());
      return;
    }
    Pending p{get_event(), get_event(), cls};
    G4_HIP(hipEventRecord(p.a, stream));
    f();
    G4_HIP(hipGetLastError());
    G4_HIP(hipEventRecord(p.b, stream));
    pending.push_back(p);
    if (pending.size() > 4096) collect();
  }
  void collect();  // synchronises and folds pending events into prof[]

  // Parameter blocks that kernels read through a pointer (rather than as
  // by-value arguments, which the compiler keeps live in scalar registers
  // and spills): staged through pinned host memory into a device ring.
  // param_push copies a block into the next slot and returns its device
  // address; param_commit marks the slot in use until the work enqueued so
  // far on the stream has run.
  static constexpr int kParamSlots = 64;
  static constexpr size_t kParamBytes = 4096;
  char* param_host = nullptr;
  char* param_dev = nullptr;
  hipEvent_t param_ev[kParamSlots] = {};
  bool param_pending[kParamSlots] = {};
  int param_next = 0, param_last = -1;
  const void* param_push(const void* data, size_t bytes);
  void param_commit();
  void param_release();
  // While capturing, blocks go to an arena that lives as long as the graph
  // (a replay reads them again): pinned host and device slots, filled in
  // order.
  struct ParamArena {
    char* host = nullptr;
    char* dev = nullptr;
    int slots = 0, used = 0;
  };
  ParamArena* arena = nullptr;

  Tuner tuner;
  // Variant of `family` for `shape` to launch now; while measuring, *a and *b
  // are events to record around the launch (else null).
  int tune_pick(int family, const std::vector<int>& shape, int nvar, hipEvent_t* a,
                hipEvent_t* b);
  void tune_release();  // destroys the events of undecided entries
  void tune_settle();   // decides every entry whose trials have completed
};

// Records the tuner's events around the launches of one family call.
class TuneScope {
 public:
  TuneScope(Context* ctx, int family, std::vector<int> shape, int nvar) : ctx_(ctx) {
    variant_ = ctx->tune_pick(family, shape, nvar, &a_, &b_);
    if (a_) G4_HIP(hipEventRecord(a_, ctx->stream));
  }
  ~TuneScope() {
    if (b_) (void)hipEventRecord(b_, ctx_->stream);
  }
  int variant() const { return variant_; }

 private:
  Context* ctx_;
  int variant_ = 0;
  hipEvent_t a_ = nullptr, b_ = nullptr;
};

// Device memory of a context comes from the device's stream-ordered pool:
// hipFreeAsync returns a buffer after the work already enqueued on the
// context's stream, without the device-wide synchronisation of hipFree (a
// request loop releasing one lane's memory would otherwise stall every other
// lane).  An allocation waits for its stream, so the buffer is free of any
// earlier work before other streams or blocking copies touch it.
void* device_alloc(Context* ctx, size_t bytes);
void device_free(Context* ctx, void* p);

template <class T>
void DevBuf<T>::alloc(Context* ctx, size_t n) {
  release();
  if (n == 0) return;
  p_ = static_cast<T*>(device_alloc(ctx, n * sizeof(T)));
  n_ = n;
  ctx_ = ctx;
  if (ctx_) ctx_->bytes += n * sizeof(T);
}

template <class T>
void DevBuf<T>::release() {
  if (p_) {
    device_free(ctx_, p_);
    if (ctx_) ctx_->bytes -= n_ * sizeof(T);
  }
  p_ = nullptr;
  n_ = 0;
  ctx_ = nullptr;
}

struct Beam {
  Context* ctx = nullptr;
  g4h_precision prec = G4H_FP64;
  int nslice = 0;
  int64_t np = 0;
  double gamma_ref = 0.0;
  bool uploaded = false;
  std::vector<int32_t> off_h;
  std::vector<double> current_h;
  DevBuf<int32_t> off;           // [nslice+1]
  DevBuf<double> cur_over_npar;  // [nslice]  current/npar (0 if npar == 0)
  // state
  DevBuf<double> xd, yd, pxd, pyd, gd;  // FP64
  DevBuf<float> xf, yf, pxf, pyf, gf;   // MIXED (gf = gamma - gamma_ref)
  DevBuf<double> theta;                 // both
  DevBuf<double> eloss;                 // per-slice longitudinal field (optional)
  DevBuf<double> pez;                   // per-particle longitudinal field (optional)
  DevBuf<double> kick;                  // per-slice energy change (g4h_beam_add_energy)
  bool pez_on = false;
  // scratch of particle passes split over several workgroups per slice
  DevBuf<double2> split_win;
  DevBuf<double> split_red;
  DevBuf<int4> split_fb;
  int max_slice_np = 0;
  // short-range space charge (g4h_beam_init_efield); ef.nz == 0: off
  g4h_efield ef{};
  double ef_ks = 0.0;        // 2 pi / lambda of the solver
  int ef_generation = 0;     // bumped by every g4h_beam_init_efield
  DevBuf<double> scez;       // [np] push term of the last space-charge pass
  DevBuf<double> scrms;      // [nslice] RMS E_z [V/m] over each slice, last pass
};

struct Field {
  Context* ctx = nullptr;
  g4h_precision prec = G4H_FP64;
  int nslice = 0, ngrid = 0, harm = 1;
  double gridmax = 0, dgrid = 0, xlambda = 0, xks = 0, slicelength = 0;
  int first = 0;
  double accuslip = 0.0;
  bool enabled = true;
  bool uploaded = false;
  // storage
  DevBuf<double2> ud;  // FP64 planes
  DevBuf<float2> uf;   // MIXED planes
  // source (beam-slice order) + bookkeeping for the windowed deposition
  DevBuf<double2> srcd;
  DevBuf<float2> srcf;
  DevBuf<int4> src_rect;  // [nslice] half-open (x0,y0,x1,y1) of possibly nonzero cells
  DevBuf<int4> dep_bbox;  // [nslice] (minx,miny,maxx,maxy) of the cells deposited last step
  // solver
  g4h_solver_kind kind = G4H_SOLVER_ADI;
  bool filter = false;
  double xc = 1, yc = 1, sig = 1;
  bool solver_ready = false;
  // Step-length dependent solver tables, one per distinct delz of the
  // lattice: built once, uploaded asynchronously (the host copy lives with
  // the entry), selected per step without a host synchronisation.
  struct DelzTable {
    std::vector<double2> host;  // [3G] ADI c, cbet, cwet, or [G] FFT factor ([M] convolution)
    DevBuf<double2> dev;
  };
  std::map<double, DelzTable> delz_tables;
  static constexpr size_t kMaxDelzTables = 64;
  int generation = 0;  // bumped when device tables a captured run may use are released
  // Many distinct step lengths: release every table and start over.  Graphs
  // captured earlier hold pointers into these tables, so the generation is
  // bumped and their replay refused; a capture in progress cannot release
  // them (its earlier nodes use them) and fails instead.
  void evict_delz_tables() {
    require(!ctx->capturing,
            "a captured run may use at most 64 distinct step lengths per field");
    G4_HIP(hipStreamSynchronize(ctx->stream));
    delz_tables.clear();
    adi_c = adi_cbet = adi_cwet = nullptr;
    fft_prop = nullptr;
    ++generation;
  }
  // ADI
  double adi_delz = 0.0;
  const double2* adi_c = nullptr;  // current tables [ngrid] (views into delz_tables)
  const double2* adi_cbet = nullptr;
  const double2* adi_cwet = nullptr;
  DevBuf<double2> adi_y;  // FP64 intermediates [nslice*ngrid^2]
  DevBuf<double2> adi_sd;  // x-sweep output of the wave-per-block sweeps, FP64 planes
  DevBuf<float2> adi_sf;   // same, MIXED planes
  // FFT
  double fft_delz = 0.0;
  DevBuf<double2> fft_tw;            // four-step twiddles / DFT twiddles, [G]
  DevBuf<double2> fft_tw_conv;       // four-step twiddles of the M-point convolution transforms
  const double2* fft_prop = nullptr;  // current separable propagator factor [G], or its
                                      // convolution spectrum [M] (conv_points)
  DevBuf<double> fft_sigmoid;  // [G*G] (filter only)
  DevBuf<double2> fft_work;    // filtered-source work planes (filter only)
  // diagnostics
  DevBuf<double> plane_stats;  // [nslice * G4H_FIELD_RAW] per ring slot (last solve)
  DevBuf<double> col_partials; // FFT column-tile partial sums
  DevBuf<int32_t> zeroed;      // [nslice] ring slots zeroed by slippage since last map
  bool stats_valid = false;

  int64_t cells() const { return static_cast<int64_t>(nslice) * ngrid * ngrid; }
  void* data() const { return prec == G4H_FP64 ? static_cast<void*>(ud.get()) : static_cast<void*>(uf.get()); }
  void* src() const { return prec == G4H_FP64 ? static_cast<void*>(srcd.get()) : static_cast<void*>(srcf.get()); }
};

}  // namespace g4

struct g4h_context : g4::Context {};
struct g4h_beam : g4::Beam {};
struct g4h_field : g4::Field {};
