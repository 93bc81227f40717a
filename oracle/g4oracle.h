/*
 * g4oracle.h - FP64 CPU restatement of the GENESIS 1.3 v4 time-dependent
 * step as the reference CUDA overlay (mulingLHY/Genesis4_CUDA_port) computes
 * it.  TEST INFRASTRUCTURE ONLY: this library is the checker for the HIP
 * path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it; the product library never links or calls it.
 *
 * Every function names the reference file:line it follows.  All arrays are
 * caller-owned host memory in FP64.
 *
 * How each function is pinned (tests/test_oracle_pins.py), since the
 * reference's CPU path (upstream GENESIS 1.3 v4.6.14) is not on disk and its
 * CUDA path cannot be built here, and the reference holds no golden vectors:
 *   orc_fc                 Bessel closed form vs scipy.special.jv
 *   orc_track_transverse   quad map vs scipy solve_ivp; natural focusing
 *                          wavenumber aw*ku/gamma; det = 1 (symplectic);
 *                          chicane -> identity at zero angle
 *   orc_apply_r56          small-angle R56 = 2 theta^2 (ld + 2 lb / 3)
 *   orc_beam_advance       zero-field closed form; resonant particle at rest;
 *                          4th-order convergence of RK4 in a constant field
 *   orc_beam_advance_ez    zero-field closed form per particle (gamma falls
 *                          by ez[i] delz); equals orc_beam_advance when
 *                          ez[i] is its slice's loss
 *   orc_deposit            single-particle known answers; partition of unity
 *   orc_cell_index         brute-force floor rule incl. grid-edge inputs
 *   orc_adi_*              Thomas vs dense complex solve (numpy); factors
 *                          reproduce the tridiagonal matrix; power
 *                          conservation; Gaussian-beam diffraction
 *   orc_fft_advance, orc_dft  numpy FFT; 2S identity; odd grids equal the
 *                          textbook Fresnel propagator and numpy-layout filter;
 *                          even grids equal the textbook step plus a
 *                          shift-theorem translation of pi dz/(G dgrid^2 ks)
 *                          cells and a constant phase (filter at the same
 *                          half-bin frequencies), and a Gaussian mode's
 *                          centroid drifts by that much per step;
 *                          Gaussian-beam diffraction; ADI ~ FFT on a smooth beam
 *   orc_apply_slippage     integer ring model
 *   orc_diag_*, orc_finalize_*  quiet-start moments and zero bunching;
 *                          shot-noise statistics; Gaussian-mode power, size
 *   orc_run                beam-to-field energy balance (first order in delz)
 *   orc_sc_field           uniform modulated disk: potential and field equal
 *                          the Bessel-function solution, second order in the
 *                          shell width; leading particles of a bunch gain
 *                          energy; linear in current; zero for a quiet start;
 *                          azimuthal modes vanish for a symmetric beam
 * Parity unpinned: the physical constants (vacimp, eev, ce) and fc(h) follow upstream
 * definitions that are absent from the reference.
 */
#ifndef G4ORACLE_H_
#define G4ORACLE_H_

#include <stdint.h>
#include "../include/g4hip_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One radiation field: storage-order planes (slice s of the ring lives in
 * plane (s+first)%nslice), cell index ix + iy*ngrid (Field.cpp:187-201). */
typedef struct orc_field {
  int32_t nslice, ngrid, first, harm;
  int32_t enabled, pad_;
  double gridmax;      /* half width of the grid [m]            */
  double dgrid;        /* 2*gridmax/(ngrid-1)                    */
  double xks;          /* 2*pi/xlambda                           */
  double xlambda;      /* wavelength of this harmonic [m]        */
  double slicelength;  /* [m]                                    */
  double accuslip;     /* accumulated slippage (Control state)   */
  double* re;          /* nslice*ngrid*ngrid                     */
  double* im;
} orc_field;

int orc_version(void);
void orc_set_threads(int nthreads);

/* Undulator coupling factor fc(h) (upstream Undulator::fc, called at
 * BeamSolverCUDA.cpp:503 and FieldSolverADICUDA.cpp:707). */
double orc_fc(double aw, int helical, int harm);

/* TrackBeamCUDA::track (TrackBeamCUDA.cpp:156-204) including the chicane on
 * the first half step and the corrector on the last (TrackBeamCUDA.cpp:181-185). */
void orc_track_transverse(int64_t n, double* x, double* y, double* px, double* py,
                          const double* gamma, const g4h_step* st, double gamma0,
                          double delz, int last_step);

/* TrackBeamCUDA::applyR56 (TrackBeamCUDA.cpp:257-278). */
void orc_apply_r56(int64_t n, const double* gamma, double* theta, const g4h_step* st,
                   double gamma0, double lambda0);

/* BeamSolverCUDA::advance (BeamSolverCUDA.cpp:449-683): gather + RK4. */
void orc_beam_advance(int nslice, const int32_t* off, const double* x, const double* y,
                      const double* px, const double* py, double* gamma, double* theta,
                      const orc_field* fields, int nfield, const g4h_step* st,
                      double gamma0, double delz, const double* slice_eloss);

/* The same with a longitudinal field per particle (beam order), the
 * UseParticleEz kernels (BeamSolverCUDA.cpp:183-191,523-547). */
void orc_beam_advance_ez(int nslice, const int32_t* off, const double* x, const double* y,
                         const double* px, const double* py, double* gamma, double* theta,
                         const orc_field* fields, int nfield, const g4h_step* st,
                         double gamma0, double delz, const double* particle_ez);

/* Source construction (FieldSolverADICUDA.cpp:75-174, FFT twin
 * FieldSolverFFTCUDA.cpp:77-155,383-395).  Writes full source planes in
 * storage order (zeroed first).  Returns 1 if the source is on this step. */
int orc_deposit(int nslice, const int32_t* off, const double* x, const double* y,
                const double* gamma, const double* theta, const orc_field* f,
                const g4h_step* st, const double* current, double delz,
                double* src_re, double* src_im);

/* Particle-to-cell indices (lower-left cell, or -1 when the particle is
 * off the grid) exactly as the gather and the deposition compute them
 * (BeamSolverCUDA.cpp:99-121, FieldSolverADICUDA.cpp:126-155). */
void orc_cell_index(int64_t n, const double* x, const double* y, double gridmax,
                    double dgrid, int ngrid, int32_t* cell);

/* FieldSolverADICUDA::init Thomas factors (FieldSolverADICUDA.cpp:821-866). */
void orc_adi_coefficients(int ngrid, double rtmp, double* c_re, double* c_im,
                          double* cbet_re, double* cbet_im, double* cwet_re,
                          double* cwet_im);

/* FieldSolverADICUDA::advance field part (FieldSolverADICUDA.cpp:789-816)
 * with Thomas sweeps (:222-277, :325-380).  src may be NULL. */
void orc_adi_advance(orc_field* f, const double* src_re, const double* src_im,
                     double delz);

/* FieldSolverFFTCUDA::advance field part (FieldSolverFFTCUDA.cpp:413-481)
 * with the k2/sigmoid tables of init (:273-323).  src may be NULL. */
void orc_fft_advance(orc_field* f, const double* src_re, const double* src_im,
                     double delz, int do_filter, double xc, double yc, double sig);

/* Plain 1-D DFT used by the FFT solver restatement (forward = exp(-i...)). */
void orc_dft(int n, const double* in_re, const double* in_im, double* out_re,
             double* out_im, int inverse);

/* ControlCUDA::applySlippage, single rank (ControlCUDA.cpp:148-252).
 * Returns the number of ring shifts performed. */
int orc_apply_slippage(orc_field* f, double slippage, double sample, int timerun,
                       int periodic);

/* Raw per-slice beam sums (DiagnosticCUDA.cpp:188-303); layout of
 * g4hip_types.h, G4H_BEAM_RAW(nharm) doubles per slice.  gamma sums are taken
 * relative to gamma_shift (see the .cpp). */
void orc_diag_beam_raw(int nslice, const int32_t* off, const double* x, const double* y,
                       const double* px, const double* py, const double* gamma,
                       const double* theta, int nharm, double gamma_shift, double* raw);

/* Raw per-slice field sums in logical slice order (DiagnosticCUDA.cpp:305-363). */
void orc_diag_field_raw(const orc_field* f, double* raw);

/* Host finalisation (DiagnosticCUDA.cpp:509-628 beam, :685-790 field). */
void orc_finalize_beam(int nslice, int nharm, const double* raw, const double* current,
                       double gamma_shift, double* q, double* global);
void orc_finalize_field(int nslice, int ngrid, double dgrid, double xlambda,
                        double slicelength, const double* raw, double* q,
                        double* global);

/* Short-range longitudinal space-charge field of every particle, the
 * EFieldSolver::shortRange the reference runs on the host before each push
 * (BeamSolverCUDA.cpp:523-547; upstream GENESIS 1.3, not in the reference),
 * restated from its published method (see the .cpp).  ks: radiation
 * wavenumber of the fundamental; gz2 = gamma0^2/(1+aw^2) (:527).  ez[np]
 * receives E_z/511000 (the push's units, :532), slice_rms[nslice] (may be
 * NULL) the RMS E_z [V/m] over each slice's particles. */
void orc_sc_field(int nslice, const int32_t* off, const double* x, const double* y,
                  const double* theta, const double* current, double ks, double gz2,
                  const g4h_efield* ef, double* ez, double* slice_rms);

/* The Gencore::run step loop (Gencore.cpp:161-254, Beam::track Beam.cpp:
 * 111-134, Field::track Field.cpp:204-212) for one field.  Records raw
 * diagnostics before the loop and after every outstep.  beam_raw capacity:
 * (1+#outsteps)*nslice*G4H_BEAM_RAW(nharm); field_raw: (1+#outsteps)*nslice*9.
 * Returns the number of records written. */
int orc_run(int nslice, const int32_t* off, double* x, double* y, double* px,
            double* py, double* gamma, double* theta, double gamma0, double lambda0,
            const double* current, const double* slice_eloss, orc_field* field,
            int solver, int do_filter, double xc, double yc, double sig,
            const g4h_step* steps, int nsteps, double sample, int timerun,
            int periodic, int nharm, double* beam_raw, double* field_raw);

/* orc_run with the short-range space-charge field of orc_sc_field computed
 * after the first transverse half step of every step (where
 * BeamSolverCUDA::advance computes it) and added to slice_eloss in the push;
 * ef may be NULL (= orc_run). */
int orc_run_sc(int nslice, const int32_t* off, double* x, double* y, double* px,
               double* py, double* gamma, double* theta, double gamma0, double lambda0,
               const double* current, const double* slice_eloss, orc_field* field,
               int solver, int do_filter, double xc, double yc, double sig,
               const g4h_step* steps, int nsteps, double sample, int timerun,
               int periodic, int nharm, double* beam_raw, double* field_raw,
               const g4h_efield* ef);

#ifdef __cplusplus
}
#endif

#endif /* G4ORACLE_H_ */
