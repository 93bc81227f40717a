// kernels.hpp - parameter blocks and host launchers of the gfx950 kernels.
#pragma once

#include "g4_internal.hpp"

namespace g4 {

// ---- streaming access ----------------------------------------------------
// Data that a pass reads or writes exactly once (particle state, field
// planes in the FFT passes, source planes) goes through nontemporal loads
// and stores, which keep it from displacing the reused data in the caches:
// a plain 2 GiB device copy reaches ~5.0 TB/s with ordinary and ~5.8 TB/s
// with nontemporal accesses on MI355X.
typedef float nt_f2 __attribute__((ext_vector_type(2)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));
// The streaming accesses go through global-address-space pointers: a pointer
// that the compiler cannot trace to a kernel argument (one read from a run
// descriptor in device memory) would otherwise give flat accesses, which
// also count on the LDS / scalar-load counter, so that every LDS wait would
// wait for them too.
#define G4_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ float ld_stream(const float* p) {
  return __builtin_nontemporal_load((G4_GLOBAL const float*)(p));
}
__device__ __forceinline__ double ld_stream(const double* p) {
  return __builtin_nontemporal_load((G4_GLOBAL const double*)(p));
}
__device__ __forceinline__ float2 ld_stream(const float2* p) {
  const nt_f2 v = __builtin_nontemporal_load((G4_GLOBAL const nt_f2*)(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ double2 ld_stream(const double2* p) {
  const nt_d2 v = __builtin_nontemporal_load((G4_GLOBAL const nt_d2*)(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_stream(float* p, float v) {
  __builtin_nontemporal_store(v, (G4_GLOBAL float*)(p));
}
__device__ __forceinline__ void st_stream(double* p, double v) {
  __builtin_nontemporal_store(v, (G4_GLOBAL double*)(p));
}
__device__ __forceinline__ void st_stream(float2* p, float2 v) {
  nt_f2 t;
  t.x = v.x;
  t.y = v.y;
  __builtin_nontemporal_store(t, (G4_GLOBAL nt_f2*)(p));
}
__device__ __forceinline__ void st_stream(double2* p, double2 v) {
  nt_d2 t;
  t.x = v.x;
  t.y = v.y;
  __builtin_nontemporal_store(t, (G4_GLOBAL nt_d2*)(p));
}

// Loads through pointers that address global memory but that the compiler
// cannot prove to (read from a parameter block in device memory): as flat
// loads they would also count on the counter of LDS and scalar loads, so
// every wait for a scalar load would wait for them too.
typedef G4_GLOBAL const nt_f2* gl_f2;
typedef G4_GLOBAL const nt_d2* gl_d2;
typedef G4_GLOBAL const double* gl_d;
__device__ __forceinline__ float2 ld_global(const float2* p) {
  const nt_f2 v = *(gl_f2)(p);
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ double2 ld_global(const double2* p) {
  const nt_d2 v = *(gl_d2)(p);
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ double ld_global(const double* p) { return *(gl_d)(p); }

// ---- particle pass -------------------------------------------------------
enum Stage : int {
  ST_T1 = 1,     // first transverse half step (with chicane)
  ST_PUSH = 2,   // gather + RK4
  ST_R56 = 4,    // chicane R56 phase shift
  ST_T2 = 8,     // last transverse half step (with corrector)
  ST_DEP = 16,   // source deposition
  ST_DIAG = 32   // per-slice beam sums
};

// One transverse half step: TrackBeamCUDA::track (TrackBeamCUDA.cpp:156-204).
struct TransParams {
  double m00, m01, m10, m11, m22, m23, m32, m33;  // chicane matrix
  double cx, cy;                                   // corrector kick * gamma0
  double xstr, xoff, ystr, yoff;                   // plane maps (offset already / strength)
  double delz;                                     // half step
  double aw2;                                      // aw^2 of the step
  double g2c;      // gamma0^2 - 1 - aw^2   (mixed gammaz)
  double g2chic;   // gamma0^2 - 1          (mixed chicane gammaz)
  int xmode, ymode;  // 0 drift, 1 focusing, 2 defocusing
  int chicane, corrector;
  // plane maps from their power series in w = K dz^2, bounded over every
  // particle with gamma >= gamma0/2: 0 no (trigonometric form), 1 |w| <= 1/2
  // (long series), 2 |w| <= 0.011 (short series)
  int xpoly, ypoly;
};

struct GatherField {
  const void* data;  // double2* or float2*, ring storage
  double gridmax, dgrid, inv_dgrid, rtmp, rharm;
  int first, ngrid, nslice, pad_;
};

constexpr int kMaxFields = 8;

struct PushParams {
  double xks, xku, aw, autophase, delz, ax, ay, kx, ky, gradx, grady;
  const double* slice_eloss;  // device [nslice] or null
  const double* particle_ez;  // device [np] in beam order, or null; replaces slice_eloss
  // mixed-precision phase equation: theta' = -(c0 + eta*(k1 + k2*eta))
  double c0, k1, k2, inv_b0;
  int nfield, pad_;
  GatherField f[kMaxFields];
};

struct DepParams {
  void* src;                    // beam-slice order, double2*/float2*
  int4* rect;                   // [nslice]
  int4* bbox;                   // [nslice]
  const double* cur_over_npar;  // [nslice]
  double scl_factor;            // vacimp*xks*fc*delz/(4*eev*dgrid^2)
  double gridmax, dgrid, inv_dgrid;
  double ax, ay, kx, ky, gradx, grady;
  int ngrid, harm;
  int window_cap;  // cells of the LDS window
  int pad_;
};

struct DiagParams {
  double* raw;  // [nslice * G4H_BEAM_RAW(nharm)]
  int nharm, pad_;
};

struct ParticleParams {
  int stages, pad_;
  double gamma0;  // beam reference energy (mixed encoding, R56)
  double r56;     // R56*2pi/lambda0/gamma0
  TransParams t1, t2;
  PushParams push;
  DepParams dep;
  DiagParams diag;
};

// With d_table (device, one entry per slice) each slice's workgroup takes its
// parameters from the table; p still selects the kernel variant (stages,
// field count, harmonics, window size), which must be common to all entries.
void launch_particle(Context* ctx, Beam* b, const ParticleParams& p,
                     const ParticleParams* d_table = nullptr);
void launch_cell_index(Context* ctx, Beam* b, const Field* f, int32_t* d_cell);
// One run of a persistent batch launch (launch_batch_persistent): where its
// slice lives in its group's beam and field, and its group's per-step tables.
struct BatchRunDesc {
  void* x;  // group beam arrays (double* in FP64, float* in MIXED)
  void* y;
  void* px;
  void* py;
  void* g;
  double* th;
  const int32_t* off;          // group slice offsets
  const ParticleParams* tab;   // group parameter table [step][nrun]
  void* u;                     // group field planes (double2* / float2*)
  double2* ys;                 // group FP64 forward-value planes
  const void* src;             // group source planes
  const int4* rect;            // group source rectangles
  const double2* const* adi;   // per step: Thomas factors c, cbet, cwet [3G]
  const double* cs;            // per step: the sweeps' i*rtmp coefficient
  double* field_hist;          // group field records [record][nrun][G4H_FIELD_RAW]
  double* partials;            // group field-sum partials of the segmented y sweep
  // the run's own split scratch when its particle pass takes K > 1 parts
  double2* swin;               // [K][window cells]
  double* sred;                // [K][sums]
  int4* sfb;                   // [K]
  int slice;                   // the run's slice (= plane) in its group
  int nrun;                    // runs in the group
  int G;
  int K;                       // parts of the run's particle pass (per-step launches)
};

// One 64-line block of one run's plane for the segmented sweeps of a whole
// batch in one launch (launch_batch_adi_step).
struct BatchSegDesc {
  void* u;                     // group field planes (double2* / float2*)
  void* ustar;                 // group x-sweep output planes
  const void* src;             // group source planes
  const int4* rect;            // group source rectangles
  double* partials;            // group partials [nrun][nblk][8], then 2 centre values per run
  const double2* const* adi;   // per step: the group's c, cbet, cwet, beta, delta [5G]
  const double* cs;            // per step: the sweeps' i*rtmp coefficient
  int G, nblk, nrun, slice, blk, pad_;
};

// One step k of every run of a batch, each kernel in one launch: the fused
// particle passes (one workgroup per run, the run's parameter entry of the
// step), the two segmented sweeps of every run's plane (one workgroup per
// 64-line block, the run's group tables of the step; G <= 256), and on
// output steps every run's field record from its sweep partials (record r).
// d_parts: (run, part) of each workgroup, every run's K parts in order.  A
// split run's parts deposit their out-of-window particles with atomics onto
// its source plane, which launch_batch_adi_step clears behind its y sweep
// (clear_src) so that each step's deposition starts from zero.  d_tabk: the
// runs' parameter blocks of step k, [nruns].  With any_split, a second
// launch (one workgroup per run of d_runs[nruns]) publishes the split runs'
// source rectangles, footprints and sums.
void launch_batch_step_particles(Context* ctx, g4h_precision prec, int nharm,
                                 const BatchRunDesc* d_runs, int nruns, const int2* d_parts,
                                 int nparts, const ParticleParams* d_tabk, int k, int win_cells,
                                 bool sums, bool any_split);
void launch_batch_adi_step(Context* ctx, g4h_precision prec, const BatchSegDesc* d_blocks,
                           int nblocks, int k, int gmax, bool add_src, bool stats,
                           bool clear_src);
// split scratch of a batch run: window cells, sums per part
size_t batch_split_red_stride();
void launch_batch_field_records(Context* ctx, const BatchRunDesc* d_runs, int nruns, int r);

// Steps [k0, k1) of every run of a batch of steady-state ADI runs in ONE
// launch: one workgroup per run loops over the steps (particle pass, both
// sweeps, records on steps with rec[k] >= 0), with workgroup barriers only.
// lds_bytes: window and factor space (max window cells + 3 max G) * 16.
void launch_batch_persistent(Context* ctx, g4h_precision prec, int nharm,
                             const BatchRunDesc* d_runs, int nruns, const int32_t* d_rec,
                             int k0, int k1, int win_cells, int gmax);
// Whether the batch kernel can run grids up to gmax with a window of
// win_cells on the context's GPU (ngrid <= 512, LDS budget).
bool batch_persistent_fits(Context* ctx, g4h_precision prec, int nharm, int win_cells, int gmax);

// Short-range space charge of every slice (space_charge.hip): writes
// b->scez[np] (the push's per-particle longitudinal term: E_z/511000 plus the
// slice loss d_slice_eloss[s] and d_base[i] when given) and b->scrms[nslice]
// (RMS E_z [V/m] over each slice's particles).  gz2 = gamma_ref^2/(1+aw^2) of the step.
struct ScParams {
  double ks, gz2, rmax, vacimp_over_2pi;
  int ngrid, nz, nphi, pad_;
};
void launch_space_charge(Context* ctx, Beam* b, double gz2, const double* d_slice_eloss,
                         const double* d_base);
size_t sc_lds_bytes(const g4h_efield& ef);

// Beam sums into the library's fixed device array (no allocation): blocking,
// raw[nslice * G4H_BEAM_RAW(nharm)] on the host; fails with
// G4H_ERR_INVALID_ARGUMENT when the pass needs more than kNoallocScratchBytes.
constexpr size_t kNoallocScratchBytes = size_t(8) << 20;
void diag_beam_raw_noalloc(Context* ctx, Beam* b, int nharm, double* raw);
// Device address of that array on `device` (current), looked up once.
void* scratch_base(int device);
// gamma += d_dgamma[slice] (device, nslice values) for every particle
void launch_slice_kick(Context* ctx, Beam* b, const double* d_dgamma);

// Host construction of parameter blocks from a lattice step.
TransParams make_trans(const g4h_step& st, double gamma0, double delz, bool last);
double make_r56(const g4h_step& st, double gamma0, double lambda0);
void make_push(PushParams& p, const g4h_step& st, double gamma0, double delz,
               Field* const* fields, int nfield, const double* d_eloss);
bool make_dep(DepParams& d, Field* f, Beam* b, const g4h_step& st, double delz);

// ---- field solvers -------------------------------------------------------
void field_prepare_solver(Field* f, double delz);  // (re)build coefficient tables
void* field_adi_ustar(Field* f);  // the ADI x sweep's output planes (allocated on first use)
// One field step with the source (if source_on).  When stats != 0 the last
// pass also writes per-ring-slot field sums into f->plane_stats.
void field_solve(Field* f, bool source_on, bool stats);
void field_plane_stats(Field* f);                      // standalone field sums
// Per-tile partial field sums partials[(p*ntile + t)*8 + k] (k < 7) and the
// centre cell at partials[nslice*ntile*8 + 2p] -> plane_stats, tiles in order.
void field_reduce_partials(Field* f, const double* partials, int ntile);
void field_zero_plane(Field* f, int slot);             // slippage zero fill
// plane_stats (ring order) -> raw[nslice*9] (logical order), honouring zeroed slots.
void field_map_stats(Field* f, double* d_raw);
// source bookkeeping reset (new beam or new source allocation)
void field_reset_source(Field* f);
void field_copy(Field* dst, const Field* src);
void ensure_source(Field* f);

}  // namespace g4
