/*
 * g4hip.h - C ABI of the MI355X-native GENESIS 1.3 v4 time-dependent step.
 *
 * This is the drop-in boundary.  Each entry point replaces one call site of
 * the reference CUDA overlay (mulingLHY/Genesis4_CUDA_port, paths relative to
 * its root) and is cited next to its declaration.  All arguments are plain
 * pointers, sizes and PODs; no exception crosses this boundary: every call
 * returns a g4h_status and leaves a message in g4h_last_error().
 *
 * Threading: a g4h_context owns one HIP stream.  A context and every object
 * created from it are used from one host thread at a time.  Calls enqueue
 * work on the context's stream and return without synchronising unless they
 * hand data back to the host.
 */
#ifndef G4HIP_H_
#define G4HIP_H_

#include <stddef.h>
#include <stdint.h>

#include "g4hip_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct g4h_context g4h_context;
typedef struct g4h_beam g4h_beam;
typedef struct g4h_field g4h_field;
typedef struct g4h_run g4h_run;
typedef struct g4h_batch g4h_batch;
typedef struct g4h_server g4h_server;

/* ------------------------------------------------------------------------ */
/* Library / context (replaces genesis4_cuda::initialize/check/stream/        */
/* synchronize, include/Genesis4CudaRuntime.h:13-50)                          */
/* ------------------------------------------------------------------------ */

/* ABI version; bumps on any incompatible change. */
int g4h_abi_version(void);

/* Number of visible GPUs (0 when none); never fails. */
int g4h_device_count(void);

/* Message of the last failed call on this host thread ("" if none). */
const char* g4h_last_error(void);

/* Create a context on `device` with its own stream (default priority).
 * Destroying a context while beams, fields or runs created from it are alive
 * defers the release until the last of them is destroyed. */
g4h_status g4h_context_create(int device, g4h_context** out);
void g4h_context_destroy(g4h_context* ctx);
/* Wait for all work enqueued on the context's stream. */
g4h_status g4h_context_synchronize(g4h_context* ctx);
/* Device bytes currently held by objects of this context. */
g4h_status g4h_context_memory(const g4h_context* ctx, uint64_t* bytes_in_use);

/* Kernel variants and the autotuner.  Some kernel families have variants
 * that compute bit-identical results with different work decompositions;
 * per shape the context tries each on its first launches and keeps the
 * fastest.  No counterpart in the reference (its kernels have one launch
 * shape each).  Families: */
#define G4H_TUNE_ADI 0   /* 0: workgroup per plane (ngrid <= 360, else as 1),
                            1 / 2: wave per 64 lines, loads 8 / 16 elements
                            ahead; bit-identical results.  3: segmented
                            sweeps (16 waves per 64 lines, no forward-value
                            plane; ngrid <= 256, else as 1), results agree
                            with 0-2 to round-off.  Unpinned, 3 runs for
                            ngrid <= 256 and 0-2 are tuned above */
#define G4H_TUNE_SPLIT 1 /* workgroups per slice of the particle pass (>= 1; by
                            default chosen from the slice count so that few
                            slices still fill the GPU; results agree to
                            round-off, not bit for bit) */
#define G4H_TUNE_BATCH 2 /* batches of steady-state ADI runs: 0 (the default)
                            each step of every run as one launch per kernel
                            (grids <= 256, else as 2), 1 every step of every
                            run in one launch (one workgroup per run looping
                            over the steps; grids <= 512), 2 one launch per
                            kernel, step and grid on per-grid streams;
                            results agree to round-off */
/* Pin `family` to `variant` on this context (-1: autotune, the default). */
g4h_status g4h_context_set_variant(g4h_context* ctx, int family, int variant);
/* Decided variant of `family` for a shape ({ngrid, nslice, precision} for
 * G4H_TUNE_ADI), -1 while still measuring or never launched. */
g4h_status g4h_context_get_variant(g4h_context* ctx, int family, const int* shape, int nshape,
                                   int* variant);

/* ------------------------------------------------------------------------ */
/* Beam: device-resident SoA (replaces Genesis4BeamSoA, include/              */
/* Genesis4BeamSoA.h:7-21, and Beam::init_beamSoA/pack_beam_to_soa/          */
/* unpack_soa_to_beam, src/Core/Beam.cpp:342-525)                             */
/* ------------------------------------------------------------------------ */

/* slice_offsets[nslice+1]: particles of slice s are [off[s], off[s+1]).
 * gamma_ref is the reference energy (Undulator::getGammaRef); in G4H_MIXED
 * the energy is stored as FP32 gamma - gamma_ref.  current[nslice] [A] is the
 * slice current used for the source scale (Beam::current). */
g4h_status g4h_beam_create(g4h_context* ctx, g4h_precision precision, int nslice,
                           const int32_t* slice_offsets, const double* current,
                           double gamma_ref, g4h_beam** out);
void g4h_beam_destroy(g4h_beam* beam);
/* Host FP64 arrays of total_particles entries each (Beam::pack_beam_to_soa). */
g4h_status g4h_beam_upload(g4h_beam* beam, const double* x, const double* y,
                           const double* px, const double* py, const double* gamma,
                           const double* theta);
/* Blocking copy back to host FP64 arrays (Beam::unpack_soa_to_beam). */
g4h_status g4h_beam_download(g4h_beam* beam, double* x, double* y, double* px,
                             double* py, double* gamma, double* theta);
/* Beam::current per slice (A; include/Beam.h:69, recomputed from the
 * particle count after a sort at src/Core/Beam.cpp:100-104): the source
 * scale (FieldSolverADICUDA.cpp:75-96) and the diagnostics read it. */
g4h_status g4h_beam_set_current(g4h_beam* beam, const double* current);
/* Slice count, particle count and precision of a beam. */
g4h_status g4h_beam_info(const g4h_beam* beam, int* nslice, int64_t* total_particles,
                         int* precision);

/* ------------------------------------------------------------------------ */
/* Field: device-resident complex planes in a ring of slices (replaces        */
/* Genesis4FieldSoA, include/Genesis4FieldSoA.h:7-17, and Field::init_fieldSoA/*/
/* pack_field_to_soa/unpack_soa_to_field, src/Core/Field.cpp:582-674)         */
/* ------------------------------------------------------------------------ */

/* gridmax is the half width (Field::gridmax); dgrid = 2*gridmax/(ngrid-1)
 * (Field.cpp:60-61).  xlambda0 is the fundamental wavelength; the field's
 * wavelength is xlambda0/harm (Field.cpp:57). */
g4h_status g4h_field_create(g4h_context* ctx, g4h_precision precision, int nslice,
                            int ngrid, double gridmax, double xlambda0, double slicelength,
                            int harm, g4h_field** out);
void g4h_field_destroy(g4h_field* field);
/* Storage-order planes (plane p holds ring slot p), cell ix + iy*ngrid;
 * `first` is the ring offset (Field::first). */
g4h_status g4h_field_upload(g4h_field* field, const double* re, const double* im,
                            int first);
/* Every ring plane <- one plane (re, im of ngrid^2 cells), ring offset
 * `first`: a seed field that is the same in every slice (upstream's
 * LoadField with a Gaussian mode) without a host array of the whole field.
 * No counterpart in the reference (its field is built on the host). */
g4h_status g4h_field_fill_planes(g4h_field* field, const double* re, const double* im, int first);
g4h_status g4h_field_download(g4h_field* field, double* re, double* im, int* first);
/* Device-to-device copy of the planes and ring state of `src` into `dst`
 * (same precision, nslice and ngrid): a state snapshot / clone with one
 * streaming copy kernel.  No counterpart in the reference, where the only
 * copy of the field is the host-side pack/unpack (Field.cpp:598-674). */
g4h_status g4h_field_copy(g4h_context* ctx, g4h_field* dst, const g4h_field* src);
/* Ring state used by the slippage (Field::first, Field::accuslip). */
g4h_status g4h_field_get_ring(const g4h_field* field, int* first, double* accuslip);
g4h_status g4h_field_set_ring(g4h_field* field, int first, double accuslip);
/* Field::disable / isEnabled (include/Field.h:117-130). */
g4h_status g4h_field_set_enabled(g4h_field* field, int enabled);

/* Field::initSolver (src/Core/Field.cpp:87-110): solver kind, FFT source
 * filter (FieldSolverFFTCUDA::initSourceFilter, FieldSolverFFTCUDA.cpp:
 * 213-221).  The ADI low-memory switch (low_memory_adisolver,
 * src/Main/Track.cpp:19-20, which deposits into the right-hand sides
 * instead of a source grid) is accepted for deck compatibility and has no
 * effect: both settings keep one source plane per slice, of which the
 * deposition writes and the solvers read only the slice's footprint
 * rectangle, and give the same results. */
g4h_status g4h_field_init_solver(g4h_field* field, g4h_solver_kind kind, int source_filter,
                                 double xcut, double ycut, double sigmoid,
                                 int low_memory_adisolver);

/* ------------------------------------------------------------------------ */
/* Step operators (the hot path)                                              */
/* ------------------------------------------------------------------------ */

/* TrackBeamCUDA::track (src/Core/TrackBeamCUDA.cpp:156-204) including
 * applyChicane on the first half step and applyCorrector on the last
 * (:181-185, :206-255).  delz is the half step. */
g4h_status g4h_track_transverse(g4h_context* ctx, g4h_beam* beam, const g4h_step* und,
                                double delz, int last_step);

/* TrackBeamCUDA::applyR56 (src/Core/TrackBeamCUDA.cpp:257-278); lambda0 is
 * Beam::reflength (Beam.cpp:130). */
g4h_status g4h_apply_r56(g4h_context* ctx, g4h_beam* beam, const g4h_step* und,
                         double lambda0);

/* BeamSolverCUDA::advance (src/Core/BeamSolverCUDA.cpp:449-683): bilinear
 * gather of up to 8 fields and the RK4 (gamma, theta) push.  slice_eloss
 * (host, nslice values, may be NULL) is -longESC/511000 per slice (:555). */
g4h_status g4h_beam_advance(g4h_context* ctx, g4h_beam* beam, g4h_field* const* fields,
                            int nfield, const g4h_step* und, double delz,
                            const double* slice_eloss);

/* Per-particle longitudinal field of the push, the UseParticleEz kernels of
 * BeamSolverCUDA::advance (src/Core/BeamSolverCUDA.cpp:183-191, uploaded
 * there at :523-547 as efield.getEField(ip) - longESC[is]/511000).  ez: host,
 * beam->np values in beam order (slice offsets as at creation), in the units
 * of slice_eloss (subtracted from dgamma/dz); copied to the device and used
 * by every later g4h_beam_advance and g4h_run_advance of this beam in place
 * of slice_eloss (which g4h_beam_advance then rejects) until cleared with
 * ez = NULL.  Set it before g4h_run_capture to have a graph read it (new
 * values set later are read by the replays; clearing it is an error there). */
g4h_status g4h_beam_set_particle_ez(g4h_beam* beam, const double* ez);

/* Energy change per slice applied to the device beam: gamma += dgamma[is]
 * for every particle of slice is (host, nslice values).  Beam::track applies
 * the collective effects (wakes, Collective::apply) and incoherent radiation
 * (Incoherent::apply) between the push and the R56 (src/Core/Beam.cpp:
 * 127-128), on the host particle arrays, which the CUDA path no longer
 * keeps current; with this call the energy change they compute reaches the
 * device state instead.  Enqueued on the context stream. */
g4h_status g4h_beam_add_energy(g4h_context* ctx, g4h_beam* beam, const double* dgamma);

/* BeamSolverCUDA::initEField (include/BeamSolverCUDA.h:42-44), called by
 * Beam::initEField (include/Beam.h:126-131) from the deck's &efield: the
 * short-range space charge of the beam (g4h_efield, g4hip_types.h).  The
 * reference computes it on the host from the host particle arrays
 * (EFieldSolver::shortRange, BeamSolverCUDA.cpp:523-547); here it is computed
 * on the device from the device beam before every push, by g4h_beam_advance
 * (from the positions it is given) and by g4h_run_advance (after the first
 * transverse half step, where BeamSolverCUDA::advance computes it), and added
 * to the slice loss and to a per-particle field set with
 * g4h_beam_set_particle_ez.  lambda is the radiation wavelength (ks = 2
 * pi/lambda).  nz = 0 switches it off; longrange != 0 is rejected (the
 * long-range field enters as the slice loss). */
g4h_status g4h_beam_init_efield(g4h_beam* beam, const g4h_efield* efield, double lambda);

/* BeamSolverCUDA::getSCField (include/BeamSolverCUDA.h:39, read for the
 * sscfield output at DiagnosticCUDA.cpp:570): per slice, the RMS short-range
 * E_z [V/m] over its particles from the last push (its mean vanishes to first
 * order; zeros when space charge is off).  field[nslice] on the host;
 * blocking. */
g4h_status g4h_beam_sc_field(g4h_beam* beam, double* field);

/* Field::track = solver->init + solver->advance (src/Core/Field.cpp:204-212;
 * FieldSolverADICUDA::advance :638-818, FieldSolverFFTCUDA::advance
 * :325-483): source deposition from the beam, then one ADI or FFT step. */
g4h_status g4h_field_track(g4h_context* ctx, g4h_field* field, g4h_beam* beam,
                           const g4h_step* und, double delz);

/* ControlCUDA::applySlippage (src/Core/ControlCUDA.cpp:148-252), one rank.
 * sample = slicelength/reflength (Gencore.cpp:72).  *shifts (may be NULL)
 * receives the number of ring shifts performed. */
g4h_status g4h_apply_slippage(g4h_context* ctx, g4h_field* field, double slippage,
                              double sample, int timerun, int periodic, int* shifts);

/* Particle-to-cell index of every particle on `field`'s grid: ix + iy*ngrid
 * of the lower-left cell, or -1 off the grid, computed by the same device
 * function the gather and the deposition use (BeamSolverCUDA.cpp:99-116,
 * FieldSolverADICUDA.cpp:126-143).  cell[total_particles] host; blocking. */
g4h_status g4h_beam_cell_index(g4h_context* ctx, g4h_beam* beam, const g4h_field* field,
                               int32_t* cell);

/* DiagnosticCUDA::calc raw sums (src/Core/DiagnosticCUDA.cpp:188-363,
 * 414-447, 633-659).  beam_raw: nslice*G4H_BEAM_RAW(nharm) doubles, energy
 * sums relative to the beam's gamma_ref; field_raw: nslice*G4H_FIELD_RAW in
 * logical slice order.  Blocking. */
g4h_status g4h_diag_beam_raw(g4h_context* ctx, g4h_beam* beam, int nharm, double* beam_raw);
g4h_status g4h_diag_field_raw(g4h_context* ctx, g4h_field* field, double* field_raw);

/* g4h_diag_beam_raw without any device allocation, for callers that cannot
 * allocate (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same results as
 * g4h_diag_beam_raw, bit for bit.  Its device memory (the FP64 parameter
 * block, the per-slice sums before the copy to the host, the split pass's
 * per-part partials and boxes) is one fixed 8 MiB array per GPU,
 * held until the process ends: nslice * (23 + 2*nharm) doubles plus, for
 * beams of few slices, ~60 doubles per workgroup of the split pass must fit,
 * about 40 000 slices at nharm = 1; larger beams get G4H_ERR_INVALID_ARGUMENT
 * before any work is enqueued.  Every call on a GPU shares that GPU's array,
 * which keeps the last call's contents; the library serialises the calls
 * made through it (blocking, one at a time per GPU), and work of any other
 * process or library on the same memory would race. */
g4h_status g4h_diag_beam_raw_noalloc(g4h_context* ctx, g4h_beam* beam, int nharm,
                                     double* beam_raw);

/* Host finalisation (DiagnosticCUDA.cpp:509-628 and 685-790): raw sums to
 * the named per-slice quantities (layouts in g4hip_types.h) and globals. */
g4h_status g4h_diag_finalize_beam(int nslice, int nharm, const double* beam_raw,
                                  const double* current, double gamma_ref, double* q,
                                  double* global);
g4h_status g4h_diag_finalize_field(int nslice, double dgrid, double xlambda,
                                   double slicelength, const double* field_raw, double* q,
                                   double* global);

/* ------------------------------------------------------------------------ */
/* Fused run: the Gencore::run loop (src/Core/Gencore.cpp:161-254) with the  */
/* beam push, deposition and beam diagnostics fused into one particle pass   */
/* and the field diagnostics fused into the last field-solver pass.          */
/* ------------------------------------------------------------------------ */

typedef struct g4h_run_config {
  double lambda0;      /* Beam::reflength [m]                              */
  double sample;       /* slicelength/reflength                            */
  int32_t timerun;     /* time-dependent (slippage on)                     */
  int32_t periodic;    /* periodic time window                             */
  int32_t nharm;       /* bunching harmonics recorded (&track bunchharm)   */
  int32_t record_diag; /* 1: record diagnostics on outsteps                */
} g4h_run_config;

/* steps[nsteps] is copied.  The run records diagnostics before the first
 * step and after every step with outstep != 0 into device history buffers.
 * The beam's and field's precision must match. */
g4h_status g4h_run_create(g4h_context* ctx, g4h_beam* beam, g4h_field* field,
                          const g4h_step* steps, int nsteps, const g4h_run_config* cfg,
                          g4h_run** out);
void g4h_run_destroy(g4h_run* run);
/* Execute up to `nsteps` further lattice steps (asynchronous). *done
 * (may be NULL) receives the number of steps executed so far. */
g4h_status g4h_run_advance(g4h_run* run, int nsteps, int* done);
/* Ensembles: after the run has been advanced to its end once, capture its
 * whole sequence (initial record and every step) into a hipGraph; then, after
 * re-initialising the beam (g4h_beam_upload / g4h_beam_generate) and the
 * field (g4h_field_upload / g4h_field_fill_planes with the run's initial ring
 * offset), g4h_run_replay enqueues the whole run again as one graph launch
 * and the records are read with g4h_run_download_diag as before.  No
 * counterpart in the reference (one run per process). */
g4h_status g4h_run_capture(g4h_run* run);
g4h_status g4h_run_replay(g4h_run* run);
/* Number of diagnostic records written so far. */
g4h_status g4h_run_records(const g4h_run* run, int* nrecords);
/* Per record and slice, the RMS short-range space-charge field [V/m] of the
 * record's step (the sscfield output, DiagnosticCUDA.cpp:570, from
 * Beam::getSCField); zeros for record 0 and when space charge is off.
 * sc_field[nrecords * nslice]; blocking. */
g4h_status g4h_run_download_sc(g4h_run* run, double* sc_field);
/* Blocking copy of records [0, nrecords) (layouts as g4h_diag_*_raw). */
g4h_status g4h_run_download_diag(g4h_run* run, double* beam_raw, double* field_raw);

/* ---- batches of independent steady-state runs -------------------------
 * No counterpart class in the reference (which runs one steady-state
 * simulation per process, Gencore.cpp:161-254); serves taper scans and
 * request streams of small runs.  Runs sharing (ngrid, gridmax, gamma0) are
 * advanced together: one particle launch, one solver step per group and
 * step, each run with its own step parameters (aw taper).  Every run is one
 * slice (steady state: no slippage). */
typedef struct g4h_batch_run {
  int32_t npart, nbins, ngrid, shotnoise;
  uint64_t seed;
  double gridmax, current, gamma0, espread_rel;
  double emitn_x, emitn_y, betax, betay, alphax, alphay;
  double seed_power, seed_w0; /* Gaussian seed at its waist [W], [m]        */
  const double* aw;           /* aw per step [nsteps] (taper), or NULL      */
} g4h_batch_run;

/* steps[nsteps] is the shared lattice (its aw is replaced per run when the
 * run gives a taper); records are taken before the first step and after
 * every step with outstep != 0, as for g4h_run. */
g4h_status g4h_batch_create(g4h_context* ctx, g4h_precision precision, g4h_solver_kind solver,
                            int nruns, const g4h_batch_run* runs, const g4h_step* steps,
                            int nsteps, double lambda0, int nharm, g4h_batch** out);
void g4h_batch_destroy(g4h_batch* batch);
g4h_status g4h_batch_advance(g4h_batch* batch, int nsteps, int* done);
g4h_status g4h_batch_records(const g4h_batch* batch, int* nrecords);
/* Records of run `run`: beam_raw[nrecords * G4H_BEAM_RAW(nharm)],
 * field_raw[nrecords * G4H_FIELD_RAW] (layouts of g4h_diag_*_raw). */
g4h_status g4h_batch_download(g4h_batch* batch, int run, double* beam_raw, double* field_raw);

/* ---- request loop ------------------------------------------------------
 * No counterpart in the reference (one simulation per process).  A server
 * keeps several runs resident on one GPU: requests are admitted while their
 * estimated HBM footprint fits `hbm_budget` and one of `nstreams` stream
 * lanes is free; every admitted request (or batch of steady-state requests
 * sharing lattice, precision and solver) runs asynchronously on its lane,
 * so work of different requests overlaps on the device; finished requests
 * free their memory for the next ones.  Latency runs from submission until
 * the request's records are on the host. */
typedef struct g4h_request {
  int32_t kind; /* 0 steady state (one slice, batched), 1 time-dependent */
  g4h_precision precision;
  g4h_solver_kind solver;
  int32_t nslice, npart, nbins, ngrid, shotnoise, nharm, periodic, nsteps;
  uint64_t seed;
  double gridmax, current, gamma0, espread_rel, chirp_rel, emitn, beta;
  double seed_power, seed_w0; /* Gaussian seed at its waist               */
  double lambda0, sample;
  const g4h_step* steps; /* [nsteps], copied at submission               */
  const double* aw;      /* per-step aw (taper) or NULL, copied          */
} g4h_request;

typedef struct g4h_result_info {
  int32_t status, nrecords;
  double latency_s; /* submission -> records on the host               */
  double queued_s;  /* submission -> admission                         */
} g4h_result_info;

typedef struct g4h_server_stats_t {
  int64_t completed;
  double p50_s, p99_s, mean_s;       /* latency over completed requests  */
  double particle_slice_steps_per_s; /* completed work / (last end - first submission) */
  uint64_t hbm_in_use;
} g4h_server_stats_t;

g4h_status g4h_server_create(int device, uint64_t hbm_budget, int nstreams, g4h_server** out);
/* Waits for the queue to drain, then releases everything. */
void g4h_server_destroy(g4h_server* server);
g4h_status g4h_server_submit(g4h_server* server, const g4h_request* request, int64_t* id);
/* Blocks until request `id` finished (timeout_s < 0: no limit); returns
 * the request's own status. */
g4h_status g4h_server_wait(g4h_server* server, int64_t id, double timeout_s,
                           g4h_result_info* info);
/* Records of a finished request: beam_raw[nrecords * nslice * G4H_BEAM_RAW(nharm)],
 * field_raw[nrecords * nslice * G4H_FIELD_RAW]. */
g4h_status g4h_server_result(g4h_server* server, int64_t id, double* beam_raw, double* field_raw);
g4h_status g4h_server_release(g4h_server* server, int64_t id);
g4h_status g4h_server_stats(g4h_server* server, g4h_server_stats_t* stats);

/* Per-kernel timing with HIP events on the context's stream.  While on,
 * every launch is bracketed by events; g4h_profile_read returns for kernel
 * class k (g4h_kernel_class) the number of launches and total milliseconds. */
typedef enum g4h_kernel_class {
  G4H_K_PARTICLE = 0,  /* fused particle pass (push, deposit, beam diag)   */
  G4H_K_FFT_ROWS = 1,
  G4H_K_FFT_COLS = 2,
  G4H_K_ADI_X = 3,
  G4H_K_ADI_Y = 4,
  G4H_K_OTHER = 5,
  G4H_K_COUNT = 6
} g4h_kernel_class;
g4h_status g4h_profile_enable(g4h_context* ctx, int on);
g4h_status g4h_profile_read(g4h_context* ctx, int kernel_class, int64_t* launches,
                            double* total_ms);

/* ------------------------------------------------------------------------ */
/* Input front-end (host only, no GPU): the pieces of the upstream loaders   */
/* the configurations need.  None of them is part of the reference; they are */
/* restated from the published methods named in each comment.                */
/* ------------------------------------------------------------------------ */

/* Undulator coupling fc(h) = aw*[JJ]_h (planar) or aw (helical, h=1). */
double g4h_undulator_fc(double aw, int helical, int harm);

/* Unscrambled Sobol points (Joe-Kuo direction numbers), dims <= 8, points
 * index0 .. index0+n-1, written row-major out[n*dims], values in [0,1). */
g4h_status g4h_sobol(int dims, int64_t index0, int64_t n, double* out);

/* Inverse standard normal CDF (Wichura AS241). */
double g4h_inverse_normal(double p);

/* Quiet-start beam loader for one run.  Fills x,y,px,py,gamma,theta for
 * nslice slices of npart particles (npart a multiple of nbins): beamlets of
 * nbins particles share (x,px,y,py,gamma) from a seeded (digitally shifted)
 * Sobol sequence mapped through the inverse normal CDF, and their phases are
 * mirrored at 2*pi*k/nbins.  gamma gets gamma0*(1+espread*g) plus
 * chirp_rel*gamma0*(s/(nslice-1)-0.5).  If shotnoise != 0, phases get the
 * Fawley random perturbation for the electron count of each beamlet
 * (current[s]*slicelength/G4H_CE/nbeamlets); with shotnoise == 0 they stay
 * mirrored.  theta0 offset per slice is 0.  Returns G4H_ERR_INVALID_ARGUMENT
 * on bad sizes. */
typedef struct g4h_beam_spec {
  double gamma0, espread_rel, chirp_rel;
  double emitn_x, emitn_y, betax, betay, alphax, alphay;
  double xcen, ycen, pxcen, pycen;
  double slicelength;
  int32_t nslice, npart, nbins, shotnoise;
  uint64_t seed;
} g4h_beam_spec;
g4h_status g4h_load_beam(const g4h_beam_spec* spec, const double* current, double* x,
                         double* y, double* px, double* py, double* gamma, double* theta);

/* Generate the quiet-start beam of `spec` (the particles g4h_load_beam
 * returns, to round-off) directly into the beam's device arrays; current:
 * per-slice current (needed for shot noise).  Every slice of the beam must
 * hold spec->npart particles.  Replaces the host loading and upload of
 * Beam::init_beamSoA (src/Core/Beam.cpp:342-390) for generated beams. */
g4h_status g4h_beam_generate(g4h_beam* beam, const g4h_beam_spec* spec, const double* current);

/* Gaussian seed field at its waist: u = u0*exp(-r^2/w0^2) normalised to
 * `power` [W] per slice (power = sum|u|^2 (dgrid eev/ks)^2/vacimp), written
 * to storage planes re/im[nslice*ngrid^2]. */
g4h_status g4h_seed_gaussian(int nslice, int ngrid, double gridmax, double xlambda,
                             double power, double w0, double* re, double* im);

/* A planar undulator beamline of nmod modules x nper periods (period
 * lambdau, rms aw_mod[nmod]), one step per period, with FODO quadrupoles
 * superimposed every quad_spacing_periods periods (length quad_len_periods,
 * k1 alternating +k1/-k1, strength averaged over the step overlap), natural
 * focusing ky = ku^2 (planar), slippage one reference length per step, an
 * output every output_step steps.  Writes min(capacity, needed) steps and
 * returns the count needed in *nsteps. */
g4h_status g4h_lattice_fodo(double lambdau, const double* aw_mod, int nmod, int nper,
                            int quad_spacing_periods, int quad_len_periods, double k1,
                            int output_step, g4h_step* steps, int capacity, int* nsteps);

/* ---- input deck and lattice file --------------------------------------
 * The reference's driver reads a namelist deck (&setup, &lattice, &time,
 * &field, &beam, &track; src/Main/Track.cpp:62-134, example test/fel.in) and
 * a lattice file (test/fel.lat) through upstream GENESIS, which is not part
 * of the reference.  These entry points serve that front-end.              */

/* Lattice file text -> per-step table of beamline `beamline` (UNDULATOR,
 * DRIFT, QUADRUPOLE, CORRECTOR, CHICANE, PHASESHIFTER, MARKER, LINE with
 * n*label repeats): undulators cut into steps of about delz, other elements
 * one step each, slippage in reference wavelengths, outstep every
 * output_step steps, truncated at zstop.  Writes min(capacity, needed) steps
 * and returns the count needed in *nsteps. */
g4h_status g4h_lattice_parse(const char* lattice_text, const char* beamline, double delz,
                             double lambda0, double gamma0, int output_step, double zstop,
                             g4h_step* steps, int capacity, int* nsteps);

/* Periodic (matched) Twiss parameters of the section [0, zmatch] of a step
 * table (&lattice zmatch): twiss = {betax, alphax, betay, alphay}. */
g4h_status g4h_lattice_match(const g4h_step* steps, int nsteps, double gamma0, double zmatch,
                             double* twiss);

/* Namelist deck text -> JSON [{"name": ..., "keys": [[key, value], ...]}, ...]
 * (lower-case names and keys, raw values, file order).  Writes at most
 * capacity bytes (NUL-terminated) and the size needed in *needed. */
g4h_status g4h_deck_parse(const char* deck_text, char* json, int capacity, int* needed);

/* Runs a deck end to end on GPU `device` (the driver of src/Main/Track.cpp
 * with &track use_gpu, alias use_cuda; use_gpu = false returns
 * G4H_ERR_UNSUPPORTED since the CPU path is upstream GENESIS) and writes the
 * records as .npy arrays named like the reference's output groups
 * (out_dir/Beam/energy.npy, out_dir/Field/power.npy, ..., shape
 * [records, slices]) plus out_dir/meta.json.  *nrecords: records written. */
g4h_status g4h_deck_run(const char* deck_path, const char* out_dir, int device, int* nrecords);

#ifdef __cplusplus
}
#endif

#endif /* G4HIP_H_ */
