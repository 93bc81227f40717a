/*
 * g4hip_types.h - plain-old-data types shared by the g4hip C ABI.
 *
 * Everything here is C (no C++ types, no HIP types) so that the ABI can be
 * bound from ctypes, cgo or a C++ host without dragging in a GPU runtime.
 *
 * Units and meanings follow GENESIS 1.3 v4 as used by the reference CUDA
 * overlay (mulingLHY/Genesis4_CUDA_port):
 *   - transverse coordinates x, y in metres, momenta px, py = gamma*beta_x,y;
 *   - gamma is the Lorentz factor, theta the ponderomotive phase [rad];
 *   - the field u is the dimensionless complex envelope whose power is
 *     sum|u|^2 * (dgrid*eev/ks)^2 / vacimp (DiagnosticCUDA.cpp:685-735).
 */
#ifndef G4HIP_TYPES_H_
#define G4HIP_TYPES_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physical constants with the values GENESIS 1.3 uses (declared extern at
 * include/Field.h:25-26 and include/Beam.h:20 of the reference, defined
 * upstream). */
#define G4H_VACIMP 376.73            /* vacuum impedance [Ohm] */
#define G4H_EEV 510999.06            /* electron rest energy [eV] */
#define G4H_CE 4.8032e-11            /* e*c [A m] */
#define G4H_CLIGHT 299792458.0       /* speed of light [m/s] */

/* Short-range space-charge solver (BeamSolverCUDA::initEField,
 * include/BeamSolverCUDA.h:42-44; the deck's &efield namelist): per slice,
 * nz longitudinal harmonics of the phase (0: off) times 2*nphi+1 azimuthal
 * modes around the slice centroid, on ngrid radial shells out to rmax times
 * the slice's largest radial offset. */
typedef struct g4h_efield {
  double rmax;
  int32_t ngrid, nz, nphi;
  int32_t longrange; /* the long-range (current-profile) field: not built */
} g4h_efield;

/* Status codes: every C-ABI call returns one; no exception crosses the ABI. */
typedef enum g4h_status {
  G4H_OK = 0,
  G4H_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad size, inconsistent shapes */
  G4H_ERR_NOT_INITIALIZED = 2,  /* state used before upload / solver init     */
  G4H_ERR_OUT_OF_MEMORY = 3,    /* device allocation failed or over budget     */
  G4H_ERR_DEVICE = 4,           /* a HIP runtime call failed                   */
  G4H_ERR_UNSUPPORTED = 5,      /* valid request this build does not handle    */
  G4H_ERR_NO_DEVICE = 6         /* no usable GPU                               */
} g4h_status;

/* Arithmetic mode of a beam / field / run. */
typedef enum g4h_precision {
  G4H_FP64 = 0,  /* everything in FP64 (the reference's arithmetic)           */
  G4H_MIXED = 1  /* FP32 particle and field storage and push; FP64 phase,
                    Thomas sweeps and per-slice reductions                    */
} g4h_precision;

/* Field solver selection (Field::initSolver, Field.cpp:87-110). */
typedef enum g4h_solver_kind {
  G4H_SOLVER_ADI = 0,
  G4H_SOLVER_FFT = 1
} g4h_solver_kind;

/*
 * One integration step of the lattice: the per-step values the reference
 * reads from the upstream Undulator object (call sites BeamSolverCUDA.cpp:
 * 503,514-531,607,625-630; TrackBeamCUDA.cpp:163-167,261-262;
 * FieldSolverADICUDA.cpp:697,707,722-730; Gencore.cpp:172-174,229,247).
 * 22 doubles followed by 4 int32: 192 bytes, no padding.
 */
typedef struct g4h_step {
  double delz;       /* steplength [m]                                      */
  double z;          /* position at the end of the step [m]                 */
  double aw;         /* rms undulator parameter (0 in drifts)              */
  double ku;         /* undulator wavenumber 2*pi/lambdau (0 in drifts)    */
  double kx, ky;     /* natural focusing, already scaled with ku^2          */
  double gradx, grady; /* transverse gradient of the undulator field        */
  double ax, ay;     /* undulator offsets [m]                               */
  double qf;         /* quadrupole strength k1 [1/m^2] (0 if none)          */
  double dqx, dqy;   /* quadrupole offsets [m]                              */
  double cx, cy;     /* corrector kicks [rad]                               */
  double chic_angle; /* chicane bending angle [rad] (0 if none)            */
  double chic_lb, chic_ld, chic_lt; /* dipole, inner drift, total length   */
  double autophase;  /* phase added to every particle before the push      */
  double slippage;   /* slippage of this step in units of the ref. length  */
  double reserved;   /* keeps the struct at 22 doubles; must be 0          */
  int32_t helical;   /* 1 = helical undulator, 0 = planar                  */
  int32_t in_undulator; /* 1 if the field is driven by the beam this step   */
  int32_t marker;    /* marker bits (1 dump field, 2 dump beam, 4 sort, 8 stop) */
  int32_t outstep;   /* 1 if diagnostics are recorded after this step       */
} g4h_step;

/* Number of raw per-slice beam sums before the bunching sums
 * (DiagnosticCUDA.cpp:102-125). */
#define G4H_BEAM_STATS 23
enum {
  G4H_B_G1 = 0, G4H_B_G2, G4H_B_X1, G4H_B_X2, G4H_B_Y1, G4H_B_Y2,
  G4H_B_PX1, G4H_B_PX2, G4H_B_PY1, G4H_B_PY2, G4H_B_XPX, G4H_B_YPY,
  G4H_B_COUNT, G4H_B_XMIN, G4H_B_XMAX, G4H_B_PXMIN, G4H_B_PXMAX,
  G4H_B_YMIN, G4H_B_YMAX, G4H_B_PYMIN, G4H_B_PYMAX, G4H_B_GMIN, G4H_B_GMAX
};
/* Raw beam record per slice: G4H_BEAM_STATS sums, then for h=1..nharm the
 * pair (sum cos h*theta, sum sin h*theta). */
#define G4H_BEAM_RAW(nharm) (G4H_BEAM_STATS + 2 * (nharm))

/* Raw per-slice field sums (DiagnosticCUDA.cpp:127-134) plus the centre
 * cell (re, im) (DiagnosticCUDA.cpp:358-362): 9 doubles per slice, stored in
 * logical slice order. */
#define G4H_FIELD_RAW 9
enum {
  G4H_F_POWER = 0, G4H_F_X1, G4H_F_X2, G4H_F_Y1, G4H_F_Y2, G4H_F_FF_RE,
  G4H_F_FF_IM, G4H_F_CENTER_RE, G4H_F_CENTER_IM
};

/* Finalised per-slice beam quantities (names as written by
 * DiagnosticCUDA::calcBeam, DiagnosticCUDA.cpp:453-597), then per harmonic
 * h=1..nharm the pair (bunching, bunchingphase). */
#define G4H_BEAM_Q 25
enum {
  G4H_Q_ENERGY = 0, G4H_Q_ENERGYSPREAD, G4H_Q_XPOSITION, G4H_Q_XSIZE,
  G4H_Q_YPOSITION, G4H_Q_YSIZE, G4H_Q_PXPOSITION, G4H_Q_PYPOSITION,
  G4H_Q_CURRENT, G4H_Q_EMITX, G4H_Q_EMITY, G4H_Q_BETAX, G4H_Q_BETAY,
  G4H_Q_ALPHAX, G4H_Q_ALPHAY, G4H_Q_XMIN, G4H_Q_XMAX, G4H_Q_PXMIN,
  G4H_Q_PXMAX, G4H_Q_YMIN, G4H_Q_YMAX, G4H_Q_PYMIN, G4H_Q_PYMAX,
  G4H_Q_EMIN, G4H_Q_EMAX
};
#define G4H_BEAM_QN(nharm) (G4H_BEAM_Q + 2 * (nharm))

/* Finalised per-slice field quantities (DiagnosticCUDA.cpp:665-755). */
#define G4H_FIELD_Q 9
enum {
  G4H_FQ_POWER = 0, G4H_FQ_XPOSITION, G4H_FQ_XSIZE, G4H_FQ_YPOSITION,
  G4H_FQ_YSIZE, G4H_FQ_INTENSITY_NF, G4H_FQ_PHASE_NF, G4H_FQ_INTENSITY_FF,
  G4H_FQ_PHASE_FF
};

/* Global (per output step) values: beam (DiagnosticCUDA.cpp:611-628) and
 * field (DiagnosticCUDA.cpp:758-790). */
#define G4H_BEAM_GLOBAL 6
enum {
  G4H_BG_ENERGY = 0, G4H_BG_ENERGYSPREAD, G4H_BG_XPOSITION, G4H_BG_XSIZE,
  G4H_BG_YPOSITION, G4H_BG_YSIZE
};
#define G4H_FIELD_GLOBAL 7
enum {
  G4H_FG_ENERGY = 0, G4H_FG_XPOSITION, G4H_FG_XSIZE, G4H_FG_YPOSITION,
  G4H_FG_YSIZE, G4H_FG_INTENSITY_NF, G4H_FG_INTENSITY_FF
};

#ifdef __cplusplus
}
#endif

#endif /* G4HIP_TYPES_H_ */
