// g4oracle.cpp - FP64 CPU restatement of the reference's time-dependent step.
//
// TEST INFRASTRUCTURE ONLY (the checker).  The product library never links,
// loads or calls this file.  Each function follows the reference text it
// cites (paths relative to the reference repository root).
//
// Pinning (where each function is checked against something other than
// itself) is listed in tests/test_oracle_*.py; see the table in the design
// notes of the repository.
//
// Deliberate, documented deviation from the reference text: the beam energy
// sums are accumulated relative to a caller-supplied shift (the reference
// energy gamma0).  In exact arithmetic sum(g) and sum(g^2) then give the same
// energy and energy spread as DiagnosticCUDA.cpp:231-232/539; in FP64 the
// shifted form avoids the 1e-8 relative cancellation of <g^2>-<g>^2 at
// gamma ~ 1e4, so two FP64 implementations that only differ in summation
// order agree to ~1e-14 instead of ~1e-8.

#include "g4oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cplx = std::complex<double>;

constexpr double kVacimp = G4H_VACIMP;
constexpr double kEev = G4H_EEV;
constexpr double kC = G4H_CLIGHT;

// ---------------------------------------------------------------------------
// Transverse maps (TrackBeamCUDA.cpp:11-69)
// ---------------------------------------------------------------------------
enum class Mode { Drift, Focusing, Defocusing };
struct PlaneMap {
  Mode mode;
  double strength;
  double offset;
};

// make_plane_map (TrackBeamCUDA.cpp:61-69)
PlaneMap plane_map(double strength, double offset) {
  PlaneMap m{Mode::Drift, strength, 0.0};
  if (strength == 0.0) return m;
  m.offset = offset / strength;
  m.mode = strength > 0.0 ? Mode::Focusing : Mode::Defocusing;
  return m;
}

// advance_focusing_plane (TrackBeamCUDA.cpp:27-59)
inline void advance_plane(const PlaneMap& m, double delz, double& coord,
                          double& mom, double gammaz) {
  if (m.mode == Mode::Drift) {
    coord += mom * delz / gammaz;
    return;
  }
  if (m.mode == Mode::Focusing) {
    const double foc = std::sqrt(m.strength / gammaz);
    const double omg = foc * delz;
    const double a1 = std::cos(omg);
    const double a2 = std::sin(omg) / foc;
    const double a3 = -a2 * foc * foc;
    const double xtmp = coord - m.offset;
    coord = a1 * xtmp + a2 * mom / gammaz + m.offset;
    mom = a3 * xtmp * gammaz + a1 * mom;
    return;
  }
  const double foc = std::sqrt(-m.strength / gammaz);
  const double omg = foc * delz;
  const double a1 = std::cosh(omg);
  const double a2 = std::sinh(omg) / foc;
  const double a3 = a2 * foc * foc;
  const double xtmp = coord - m.offset;
  coord = a1 * xtmp + a2 * mom / gammaz + m.offset;
  mom = a3 * xtmp * gammaz + a1 * mom;
}

struct M4 {
  double v[4][4];
};
M4 eye4() {
  M4 m{};
  for (int i = 0; i < 4; ++i) m.v[i][i] = 1.0;
  return m;
}
// left_multiply (TrackBeamCUDA.cpp:80-91): m <- e * m
void lmul(M4& m, const M4& e) {
  M4 t{};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) t.v[i][j] += e.v[i][k] * m.v[k][j];
  m = t;
}

// build_chicane_matrix (TrackBeamCUDA.cpp:93-148)
M4 chicane_matrix(double angle, double lb, double ld, double lt) {
  M4 m = eye4(), d1 = eye4(), d2 = eye4(), d3 = eye4(), bp = eye4(), bn = eye4(),
     ep = eye4(), en = eye4();
  d1.v[0][1] = ld / std::cos(angle);
  d1.v[2][3] = ld / std::cos(angle);
  d2.v[0][1] = lt - 4 * lb - 2 * ld;
  d2.v[2][3] = lt - 4 * lb - 2 * ld;
  d3.v[0][1] = -lt;
  d3.v[2][3] = -lt;
  const double radius = lb / std::sin(angle);
  const double path = radius * angle;
  bp.v[2][3] = path;
  bp.v[0][0] = std::cos(angle);
  bp.v[0][1] = radius * std::sin(angle);
  bp.v[1][0] = -std::sin(angle) / radius;
  bp.v[1][1] = std::cos(angle);
  bn.v[2][3] = path;
  bn.v[0][0] = std::cos(-angle);
  bn.v[0][1] = radius * std::sin(-angle) * -1;
  bn.v[1][0] = -std::sin(-angle) / radius * -1;
  bn.v[1][1] = std::cos(-angle);
  const double edge = std::tan(angle) / radius;
  ep.v[1][0] = edge;
  ep.v[3][2] = -edge;
  en.v[1][0] = -edge * -1;
  en.v[3][2] = edge * -1;
  lmul(m, bp);
  lmul(m, ep);
  lmul(m, d1);
  lmul(m, en);
  lmul(m, bn);
  lmul(m, d2);
  lmul(m, bn);
  lmul(m, en);
  lmul(m, d1);
  lmul(m, ep);
  lmul(m, bp);
  lmul(m, d3);
  return m;
}

// ---------------------------------------------------------------------------
// Gather + ODE (BeamSolverCUDA.cpp:76-223)
// ---------------------------------------------------------------------------
struct FieldView {
  int first, ngrid, nslice;
  double gridmax, dgrid, rtmp, rharm;
  const double* re;
  const double* im;
};

// sample_single_field (BeamSolverCUDA.cpp:87-143)
inline void sample_field(const FieldView& f, int slice, double x, double y, double awloc,
                         double& rr, double& ri) {
  rr = 0.0;
  ri = 0.0;
  const int ng = f.ngrid;
  const double gm = f.gridmax;
  if ((x <= -gm) || (x >= gm) || (y <= -gm) || (y >= gm)) return;
  double wx = (x + gm) / f.dgrid;
  double wy = (y + gm) / f.dgrid;
  const int ix = static_cast<int>(std::floor(wx));
  const int iy = static_cast<int>(std::floor(wy));
  if (ix < 0 || ix >= ng - 1 || iy < 0 || iy >= ng - 1) return;
  wx = 1.0 + static_cast<double>(ix) - wx;
  wy = 1.0 + static_cast<double>(iy) - wy;
  const int islice = (slice + f.first) % f.nslice;
  const int64_t base = static_cast<int64_t>(islice) * ng * ng;
  int64_t idx = ix + static_cast<int64_t>(iy) * ng;
  const double w00 = wx * wy;
  const double w10 = (1.0 - wx) * wy;
  const double w01 = wx * (1.0 - wy);
  const double w11 = (1.0 - wx) * (1.0 - wy);
  double cre = f.re[base + idx] * w00;
  double cim = f.im[base + idx] * w00;
  ++idx;
  cre += f.re[base + idx] * w10;
  cim += f.im[base + idx] * w10;
  idx += ng - 1;
  cre += f.re[base + idx] * w01;
  cim += f.im[base + idx] * w01;
  ++idx;
  cre += f.re[base + idx] * w11;
  cim += f.im[base + idx] * w11;
  const double scale = f.rtmp * awloc;
  rr = scale * cre;
  ri = -scale * cim;
}

// evaluate_longitudinal_ode_multi_field (BeamSolverCUDA.cpp:145-180); for one
// field it produces the same values as the single-field form (:197-223).
inline void ode(double tgam, double tthet, double btpar, double ez, double xks, double xku,
                int nfield, const double* rharm, const double* rr, const double* ri,
                double& k2gg, double& k2pp) {
  const double ztemp1 = -2.0 / xks;
  double cre = 0.0, cim = 0.0;
  for (int i = 0; i < nfield; ++i) {
    const double arg = rharm[i] * tthet;
    const double c = std::cos(arg);
    const double s = std::sin(arg);
    cre += rr[i] * c + ri[i] * s;
    cim += ri[i] * c - rr[i] * s;
  }
  const double btper0 = btpar + ztemp1 * cre;
  const double btpar0 = std::sqrt(1.0 - btper0 / (tgam * tgam));
  k2pp += xks * (1.0 - 1.0 / btpar0) + xku;
  k2gg += cim / btpar0 / tgam - ez;
}

bool power_of_two(int n) { return n > 0 && (n & (n - 1)) == 0; }

// In-place 1-D transform along a strided line (forward = exp(-i)), the
// arithmetic cuFFT's Z2Z performs (FieldSolverFFTCUDA.cpp:444-474): an
// iterative radix-2 DIT for powers of two, a direct DFT otherwise.  Twiddles
// exp(-2*pi*i*m/n) are evaluated with the exponent reduced mod n.
void fft_line(cplx* a, int n, int64_t stride, bool inverse, const std::vector<cplx>& w,
              std::vector<cplx>& tmp) {
  tmp.resize(n);
  for (int i = 0; i < n; ++i) tmp[i] = a[i * stride];
  if (power_of_two(n)) {
    // bit reversal
    for (int i = 1, j = 0; i < n; ++i) {
      int bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) std::swap(tmp[i], tmp[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
      const int step = n / len;
      for (int i = 0; i < n; i += len) {
        for (int k = 0; k < len / 2; ++k) {
          cplx t = w[k * step];
          if (inverse) t = std::conj(t);
          const cplx u = tmp[i + k];
          const cplx v = tmp[i + k + len / 2] * t;
          tmp[i + k] = u + v;
          tmp[i + k + len / 2] = u - v;
        }
      }
    }
    for (int i = 0; i < n; ++i) a[i * stride] = tmp[i];
  } else {
    for (int k = 0; k < n; ++k) {
      cplx acc(0.0, 0.0);
      for (int m = 0; m < n; ++m) {
        cplx t = w[(static_cast<int64_t>(k) * m) % n];
        if (inverse) t = std::conj(t);
        acc += tmp[m] * t;
      }
      a[k * stride] = acc;
    }
  }
}

std::vector<cplx> twiddles(int n) {
  std::vector<cplx> w(n);
  const double pi2 = 4.0 * std::asin(1.0);
  for (int m = 0; m < n; ++m) {
    const double ang = -pi2 * static_cast<double>(m) / static_cast<double>(n);
    w[m] = cplx(std::cos(ang), std::sin(ang));
  }
  return w;
}

// 2-D transform of one plane: rows (x) then columns (y).
void fft2_plane(cplx* p, int n, bool inverse, const std::vector<cplx>& w,
                std::vector<cplx>& tmp) {
  for (int iy = 0; iy < n; ++iy) fft_line(p + static_cast<int64_t>(iy) * n, n, 1, inverse, w, tmp);
  for (int ix = 0; ix < n; ++ix) fft_line(p + ix, n, n, inverse, w, tmp);
}

}  // namespace

extern "C" {

int orc_version(void) { return 1; }

void orc_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

// fc(h): planar coupling aw*[JJ]_h with [JJ]_h = (-1)^((h-1)/2) *
// (J_{(h-1)/2}(h xi) - J_{(h+1)/2}(h xi)), xi = aw^2/(2(1+aw^2)) (aw rms);
// helical: aw for h=1, 0 otherwise.  Even harmonics couple only off axis and
// get 0.  Restated from the published planar/helical undulator coupling
// (upstream Undulator::fc is not part of the reference).
double orc_fc(double aw, int helical, int harm) {
  if (aw == 0.0 || harm < 1) return 0.0;
  if (helical) return harm == 1 ? aw : 0.0;
  if ((harm % 2) == 0) return 0.0;
  const double xi = static_cast<double>(harm) * aw * aw / (2.0 * (1.0 + aw * aw));
  const int n = (harm - 1) / 2;
  const double sign = (n % 2) ? -1.0 : 1.0;
  return aw * sign * (::jn(n, xi) - ::jn(n + 1, xi));
}

// TrackBeamCUDA::track (TrackBeamCUDA.cpp:156-204), applyCorrector (:206-218),
// applyChicane (:220-255).
void orc_track_transverse(int64_t n, double* x, double* y, double* px, double* py,
                          const double* gamma, const g4h_step* st, double gamma0,
                          double delz, int last_step) {
  const double aw = st->aw;
  const double betpar0 = std::sqrt(1 - (1 + aw * aw) / gamma0 / gamma0);
  const double quad = st->qf * gamma0;
  const double qnatx = st->kx * aw * aw / gamma0 / betpar0;
  const double qnaty = st->ky * aw * aw / gamma0 / betpar0;
  const double qx = quad + qnatx;
  const double qy = -quad + qnaty;
  const double xoff = quad * st->dqx + qnatx * st->ax;
  const double yoff = -quad * st->dqy + qnaty * st->ay;

  if (last_step) {
    if (st->cx != 0 || st->cy != 0) {
      const double cx = st->cx * gamma0, cy = st->cy * gamma0;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) {
        px[i] += cx;
        py[i] += cy;
      }
    }
  } else if (st->chic_angle != 0) {
    const M4 m = chicane_matrix(st->chic_angle, st->chic_lb, st->chic_ld, st->chic_lt);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const double gz = std::sqrt(gamma[i] * gamma[i] - 1.0 - px[i] * px[i] - py[i] * py[i]);
      double tmp = x[i];
      x[i] = m.v[0][0] * tmp + m.v[0][1] * px[i] / gz;
      px[i] = m.v[1][0] * tmp * gz + m.v[1][1] * px[i];
      tmp = y[i];
      y[i] = m.v[2][2] * tmp + m.v[2][3] * py[i] / gz;
      py[i] = m.v[3][2] * tmp * gz + m.v[3][3] * py[i];
    }
  }

  const PlaneMap xm = plane_map(qx, xoff);
  const PlaneMap ym = plane_map(qy, yoff);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const double gz =
        std::sqrt(gamma[i] * gamma[i] - 1.0 - aw * aw - px[i] * px[i] - py[i] * py[i]);
    advance_plane(xm, delz, x[i], px[i], gz);
    advance_plane(ym, delz, y[i], py[i], gz);
  }
}

// TrackBeamCUDA::applyR56 (TrackBeamCUDA.cpp:257-278)
void orc_apply_r56(int64_t n, const double* gamma, double* theta, const g4h_step* st,
                   double gamma0, double lambda0) {
  const double angle = st->chic_angle, lb = st->chic_lb, ld = st->chic_ld;
  if (angle == 0) return;
  double r56 = (4 * lb / std::sin(angle) * (1 - angle / std::tan(angle)) +
                2 * ld * std::tan(angle) / std::cos(angle)) *
               angle;
  r56 = r56 * 4 * std::asin(1) / lambda0 / gamma0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) theta[i] += r56 * (gamma[i] - gamma0);
}

// BeamSolverCUDA::advance (BeamSolverCUDA.cpp:449-683) with the RK4
// two-register scheme of push_longitudinal_particles_* (:225-409).  The
// longitudinal field of a particle is particle_ez[ip] when given (the
// UseParticleEz kernels, :183-191 and :523-547), else slice_eloss[is].
static void beam_advance(int nslice, const int32_t* off, const double* x, const double* y,
                         const double* px, const double* py, double* gamma, double* theta,
                         const orc_field* fields, int nfield, const g4h_step* st,
                         double gamma0, double delz, const double* slice_eloss,
                         const double* particle_ez) {
  if (nslice <= 0 || nfield <= 0 || nfield > 8) return;
  FieldView fv[8];
  double rharm[8];
  double xks_local = 1.0;
  for (int i = 0; i < nfield; ++i) {
    const orc_field& f = fields[i];
    xks_local = f.xks / static_cast<double>(f.harm);  // :497
    fv[i] = FieldView{f.first, f.ngrid, f.nslice, f.gridmax, f.dgrid,
                      orc_fc(st->aw, st->helical, f.harm) / f.xks,  // :503
                      static_cast<double>(f.harm), f.re, f.im};
    rharm[i] = fv[i].rharm;
  }
  double xku = st->ku;  // :514-518
  if (xku == 0.0) xku = xks_local * 0.5 / gamma0 / gamma0;
  const double aw = st->aw;
  const double autophase = st->autophase;
  const double xks = xks_local;

#pragma omp parallel for schedule(dynamic, 4)
  for (int is = 0; is < nslice; ++is) {
    const double ez_slice = slice_eloss ? slice_eloss[is] : 0.0;  // :555
    for (int64_t ip = off[is]; ip < off[is + 1]; ++ip) {
      const double ez = particle_ez ? particle_ez[ip] : ez_slice;
      double g = gamma[ip];
      double t = theta[ip] + autophase;
      const double xp = x[ip], yp = y[ip], pxp = px[ip], pyp = py[ip];
      const double dx = xp - st->ax, dy = yp - st->ay;
      const double awloc =
          1.0 + 0.5 * (st->kx * dx * dx + st->ky * dy * dy) + st->gradx * dx + st->grady * dy;
      const double btpar = 1.0 + pxp * pxp + pyp * pyp + aw * aw * awloc * awloc;
      double rr[8], ri[8];
      for (int k = 0; k < nfield; ++k) sample_field(fv[k], is, xp, yp, awloc, rr[k], ri[k]);

      double k2gg = 0.0, k2pp = 0.0, k3gg = 0.0, k3pp = 0.0;
      ode(g, t, btpar, ez, xks, xku, nfield, rharm, rr, ri, k2gg, k2pp);
      double stpz = 0.5 * delz;
      g += stpz * k2gg;
      t += stpz * k2pp;
      k3gg = k2gg;
      k3pp = k2pp;
      k2gg = 0.0;
      k2pp = 0.0;
      ode(g, t, btpar, ez, xks, xku, nfield, rharm, rr, ri, k2gg, k2pp);
      g += stpz * (k2gg - k3gg);
      t += stpz * (k2pp - k3pp);
      k3gg /= 6.0;
      k3pp /= 6.0;
      k2gg *= -0.5;
      k2pp *= -0.5;
      ode(g, t, btpar, ez, xks, xku, nfield, rharm, rr, ri, k2gg, k2pp);
      stpz = delz;
      g += stpz * k2gg;
      t += stpz * k2pp;
      k3gg -= k2gg;
      k3pp -= k2pp;
      k2gg *= 2.0;
      k2pp *= 2.0;
      ode(g, t, btpar, ez, xks, xku, nfield, rharm, rr, ri, k2gg, k2pp);
      g += stpz * (k3gg + k2gg / 6.0);
      t += stpz * (k3pp + k2pp / 6.0);
      gamma[ip] = g;
      theta[ip] = t;
    }
  }
}

void orc_beam_advance(int nslice, const int32_t* off, const double* x, const double* y,
                      const double* px, const double* py, double* gamma, double* theta,
                      const orc_field* fields, int nfield, const g4h_step* st,
                      double gamma0, double delz, const double* slice_eloss) {
  beam_advance(nslice, off, x, y, px, py, gamma, theta, fields, nfield, st, gamma0, delz,
               slice_eloss, nullptr);
}

void orc_beam_advance_ez(int nslice, const int32_t* off, const double* x, const double* y,
                         const double* px, const double* py, double* gamma, double* theta,
                         const orc_field* fields, int nfield, const g4h_step* st,
                         double gamma0, double delz, const double* particle_ez) {
  beam_advance(nslice, off, x, y, px, py, gamma, theta, fields, nfield, st, gamma0, delz,
               nullptr, particle_ez);
}

void orc_cell_index(int64_t n, const double* x, const double* y, double gridmax,
                    double dgrid, int ngrid, int32_t* cell) {
  for (int64_t i = 0; i < n; ++i) {
    cell[i] = -1;
    const double xp = x[i], yp = y[i];
    if (!(xp > -gridmax && xp < gridmax && yp > -gridmax && yp < gridmax)) continue;
    const int ix = static_cast<int>(std::floor((xp + gridmax) / dgrid));
    const int iy = static_cast<int>(std::floor((yp + gridmax) / dgrid));
    if (ix < 0 || ix >= ngrid - 1 || iy < 0 || iy >= ngrid - 1) continue;
    cell[i] = ix + iy * ngrid;
  }
}

// adi_compute_slice_scl (FieldSolverADICUDA.cpp:75-96) and
// adi_deposit_source_atomic (:109-174); the source switch is :697.
int orc_deposit(int nslice, const int32_t* off, const double* x, const double* y,
                const double* gamma, const double* theta, const orc_field* f,
                const g4h_step* st, const double* current, double delz,
                double* src_re, double* src_im) {
  const int ng = f->ngrid;
  const int64_t plane = static_cast<int64_t>(ng) * ng;
  std::memset(src_re, 0, sizeof(double) * plane * f->nslice);
  std::memset(src_im, 0, sizeof(double) * plane * f->nslice);
  const int harm = f->harm;
  const bool do_source = st->in_undulator && f->enabled && ((harm % 2) == 1);
  if (!do_source) return 0;
  const double fch = orc_fc(st->aw, st->helical, harm);
  const double gm = f->gridmax, dg = f->dgrid;

#pragma omp parallel for schedule(dynamic, 4)
  for (int is = 0; is < nslice; ++is) {
    double scl = 0.0;
    const int npar = off[is + 1] - off[is];
    if (npar > 0) {
      scl = fch * kVacimp * current[is] * f->xks * delz;
      scl /= 4.0 * kEev * static_cast<double>(npar) * dg * dg;
    }
    if (scl == 0.0) continue;
    const int target = (is + f->first) % f->nslice;
    double* tre = src_re + target * plane;
    double* tim = src_im + target * plane;
    for (int64_t ip = off[is]; ip < off[is + 1]; ++ip) {
      const double xp = x[ip], yp = y[ip];
      if (!(xp > -gm && xp < gm && yp > -gm && yp < gm)) continue;
      const double wxr = (xp + gm) / dg;
      const double wyr = (yp + gm) / dg;
      const int ix = static_cast<int>(std::floor(wxr));
      const int iy = static_cast<int>(std::floor(wyr));
      if (ix < 0 || ix >= ng - 1 || iy < 0 || iy >= ng - 1) continue;
      const double wx = 1.0 + static_cast<double>(ix) - wxr;
      const double wy = 1.0 + static_cast<double>(iy) - wyr;
      const double dx = xp - st->ax, dy = yp - st->ay;
      const double faw2 =
          1.0 + st->kx * dx * dx + st->ky * dy * dy + 2.0 * (st->gradx * dx + st->grady * dy);
      const double part = std::sqrt(faw2) * scl / gamma[ip];
      const double phase = static_cast<double>(harm) * theta[ip];
      const double cre = std::sin(phase) * part;
      const double cim = std::cos(phase) * part;
      const int64_t base = ix + static_cast<int64_t>(iy) * ng;
      const double w00 = wx * wy;
      const double w10 = (1.0 - wx) * wy;
      const double w01 = wx * (1.0 - wy);
      const double w11 = (1.0 - wx) * (1.0 - wy);
      tre[base] += cre * w00;
      tim[base] += cim * w00;
      tre[base + 1] += cre * w10;
      tim[base + 1] += cim * w10;
      tre[base + ng] += cre * w01;
      tim[base + ng] += cim * w01;
      tre[base + ng + 1] += cre * w11;
      tim[base + ng + 1] += cim * w11;
    }
  }
  return 1;
}

// FieldSolverADICUDA::init (FieldSolverADICUDA.cpp:821-866)
void orc_adi_coefficients(int ngrid, double rtmp, double* c_re, double* c_im,
                          double* cbet_re, double* cbet_im, double* cwet_re,
                          double* cwet_im) {
  std::vector<double> mupp(ngrid), mmid(ngrid), mlow(ngrid);
  mupp[0] = rtmp;
  mmid[0] = -2 * rtmp;
  mlow[0] = 0;
  for (int i = 1; i < ngrid - 1; i++) {
    mupp[i] = rtmp;
    mmid[i] = -2 * rtmp;
    mlow[i] = rtmp;
  }
  mupp[ngrid - 1] = 0;
  mmid[ngrid - 1] = -2 * rtmp;
  mlow[ngrid - 1] = rtmp;
  std::vector<cplx> w1(ngrid), w2(ngrid), c(ngrid), cbet(ngrid), cwet(ngrid);
  for (int i = 0; i < ngrid; i++) {
    w1[i] = cplx(0.0, -mupp[i]);
    w2[i] = cplx(1.0, -mmid[i]);
    c[i] = cplx(0.0, -mlow[i]);
  }
  cbet[0] = cplx(1.0, 0.0) / w2[0];
  cwet[0] = cplx(0.0, 0.0);
  for (int i = 1; i < ngrid; i++) {
    cwet[i] = w1[i - 1] * cbet[i - 1];
    cbet[i] = cplx(1.0, 0.0) / (w2[i] - c[i] * cwet[i]);
  }
  for (int i = 0; i < ngrid; ++i) {
    c_re[i] = c[i].real();
    c_im[i] = c[i].imag();
    cbet_re[i] = cbet[i].real();
    cbet_im[i] = cbet[i].imag();
    cwet_re[i] = cwet[i].real();
    cwet_im[i] = cwet[i].imag();
  }
}

// FieldSolverADICUDA::advance (FieldSolverADICUDA.cpp:789-816): RHS with the
// y Laplacian (:176-220) -> Thomas along x (:222-277) -> RHS with the x
// Laplacian (:279-323) -> Thomas along y (:325-380).  The source enters both
// half steps.
void orc_adi_advance(orc_field* f, const double* src_re, const double* src_im,
                     double delz) {
  const int ng = f->ngrid;
  const int64_t plane = static_cast<int64_t>(ng) * ng;
  const double rtmp = 0.25 * delz / (f->xks * f->dgrid * f->dgrid);
  std::vector<double> cre(ng), cim(ng), bre(ng), bim(ng), wre(ng), wim(ng);
  orc_adi_coefficients(ng, rtmp, cre.data(), cim.data(), bre.data(), bim.data(), wre.data(),
                       wim.data());
  std::vector<cplx> c(ng), cbet(ng), cwet(ng);
  for (int i = 0; i < ng; ++i) {
    c[i] = cplx(cre[i], cim[i]);
    cbet[i] = cplx(bre[i], bim[i]);
    cwet[i] = cplx(wre[i], wim[i]);
  }
  const double cs = rtmp;  // cstep.imag() (:790, :830-831)

#pragma omp parallel for schedule(dynamic, 1)
  for (int s = 0; s < f->nslice; ++s) {
    double* fre = f->re + s * plane;
    double* fim = f->im + s * plane;
    const double* sre = src_re ? src_re + s * plane : nullptr;
    const double* sim = src_im ? src_im + s * plane : nullptr;
    std::vector<cplx> u(plane), r(plane);
    for (int64_t g = 0; g < plane; ++g) u[g] = cplx(fre[g], fim[g]);
    // step A: rhs with the y Laplacian
    for (int iy = 0; iy < ng; ++iy) {
      for (int ix = 0; ix < ng; ++ix) {
        const int64_t g = ix + static_cast<int64_t>(iy) * ng;
        const cplx center = u[g];
        cplx lap;
        if (iy == 0)
          lap = u[g + ng] - center * 2.0;
        else if (iy == ng - 1)
          lap = u[g - ng] - center * 2.0;
        else
          lap = u[g + ng] + u[g - ng] - center * 2.0;
        cplx rhs = center + cplx(-cs * lap.imag(), cs * lap.real());
        if (sre) rhs += cplx(sre[g], sim[g]);
        r[g] = rhs;
      }
    }
    // Thomas along x for every row
    for (int iy = 0; iy < ng; ++iy) {
      cplx* row = u.data() + static_cast<int64_t>(iy) * ng;
      const cplx* rr = r.data() + static_cast<int64_t>(iy) * ng;
      row[0] = rr[0] * cbet[0];
      for (int k = 1; k < ng; ++k) row[k] = (rr[k] - c[k] * row[k - 1]) * cbet[k];
      for (int k = ng - 2; k >= 0; --k) row[k] = row[k] - cwet[k + 1] * row[k + 1];
    }
    // step B: rhs with the x Laplacian
    for (int iy = 0; iy < ng; ++iy) {
      for (int ix = 0; ix < ng; ++ix) {
        const int64_t g = ix + static_cast<int64_t>(iy) * ng;
        const cplx center = u[g];
        cplx lap;
        if (ix == 0)
          lap = u[g + 1] - center * 2.0;
        else if (ix == ng - 1)
          lap = u[g - 1] - center * 2.0;
        else
          lap = u[g + 1] + u[g - 1] - center * 2.0;
        cplx rhs = center + cplx(-cs * lap.imag(), cs * lap.real());
        if (sre) rhs += cplx(sre[g], sim[g]);
        r[g] = rhs;
      }
    }
    // Thomas along y for every column
    for (int ix = 0; ix < ng; ++ix) {
      u[ix] = r[ix] * cbet[0];
      for (int k = 1; k < ng; ++k) {
        const int64_t g = ix + static_cast<int64_t>(k) * ng;
        u[g] = (r[g] - c[k] * u[g - ng]) * cbet[k];
      }
      for (int k = ng - 2; k >= 0; --k) {
        const int64_t g = ix + static_cast<int64_t>(k) * ng;
        u[g] = u[g] - cwet[k + 1] * u[g + ng];
      }
    }
    for (int64_t g = 0; g < plane; ++g) {
      fre[g] = u[g].real();
      fim[g] = u[g].imag();
    }
  }
}

void orc_dft(int n, const double* in_re, const double* in_im, double* out_re,
             double* out_im, int inverse) {
  std::vector<cplx> a(n), tmp;
  for (int i = 0; i < n; ++i) a[i] = cplx(in_re[i], in_im[i]);
  const std::vector<cplx> w = twiddles(n);
  fft_line(a.data(), n, 1, inverse != 0, w, tmp);
  for (int i = 0; i < n; ++i) {
    out_re[i] = a[i].real();
    out_im[i] = a[i].imag();
  }
}

// FieldSolverFFTCUDA::init tables (FieldSolverFFTCUDA.cpp:273-323) and
// ::advance (:413-481): u <- IFFT2[FFT2(u) e^{i k2 delz} + 2 filt FFT2(S)] / G^2.
void orc_fft_advance(orc_field* f, const double* src_re, const double* src_im,
                     double delz, int do_filter, double xc, double yc, double sig) {
  const int ng = f->ngrid;
  const int64_t plane = static_cast<int64_t>(ng) * ng;
  const double dk = 4. * std::asin(1.) / (static_cast<double>(ng) * f->dgrid);
  const bool shape_ok = (xc > 0.0 && yc > 0.0 && sig > 0.0);
  const bool filt_on = do_filter && shape_ok;
  std::vector<double> k2(plane), sigm(plane);
  const double shift = -0.5 * static_cast<double>(ng - 1);
  for (int iy = 0; iy < ng; iy++) {
    const double dy = static_cast<double>(iy) + shift;
    const double yv = shape_ok ? dy / static_cast<double>(ng) / yc : 0.0;
    for (int ix = 0; ix < ng; ix++) {
      const double dx = static_cast<double>(ix) + shift;
      const double xv = shape_ok ? dx / static_cast<double>(ng) / xc : 0.0;
      const int iiy = (iy + (ng + 1) / 2) % ng;
      const int iix = (ix + (ng + 1) / 2) % ng;
      const int ii = iiy * ng + iix;
      k2[ii] = -(dx * dx + dy * dy) * dk * dk / 2. / f->xks;
      sigm[ii] = shape_ok ? 1. / (1 + std::exp((std::sqrt(xv * xv + yv * yv) - 1) / sig)) : 1.0;
    }
  }
  const std::vector<cplx> w = twiddles(ng);
  const double norm = 1. / static_cast<double>(plane);

#pragma omp parallel for schedule(dynamic, 1)
  for (int s = 0; s < f->nslice; ++s) {
    std::vector<cplx> u(plane), sf(plane, cplx(0.0, 0.0)), tmp;
    double* fre = f->re + s * plane;
    double* fim = f->im + s * plane;
    for (int64_t g = 0; g < plane; ++g) u[g] = cplx(fre[g], fim[g]);
    fft2_plane(u.data(), ng, false, w, tmp);
    if (src_re) {
      for (int64_t g = 0; g < plane; ++g) sf[g] = cplx(src_re[s * plane + g], src_im[s * plane + g]);
      fft2_plane(sf.data(), ng, false, w, tmp);
    }
    for (int64_t g = 0; g < plane; ++g) {
      const double ph = k2[g] * delz;
      cplx out = u[g] * cplx(std::cos(ph), std::sin(ph));
      double sre = sf[g].real(), sim = sf[g].imag();
      if (filt_on) {
        sre *= sigm[g];
        sim *= sigm[g];
      }
      out += cplx(2.0 * sre, 2.0 * sim);
      u[g] = out;
    }
    fft2_plane(u.data(), ng, true, w, tmp);
    for (int64_t g = 0; g < plane; ++g) {
      fre[g] = u[g].real() * norm;
      fim[g] = u[g].imag() * norm;
    }
  }
}

// ControlCUDA::applySlippage for a single rank (ControlCUDA.cpp:148-252)
int orc_apply_slippage(orc_field* f, double slippage, double sample, int timerun,
                       int periodic) {
  if (!timerun) return 0;
  f->accuslip += slippage;
  const int n = f->nslice;
  const int64_t plane = static_cast<int64_t>(f->ngrid) * f->ngrid;
  int shifts = 0;
  int direction = 1;
  while (std::abs(f->accuslip) > (sample * 0.8)) {
    direction = f->accuslip < 0 ? -1 : 1;
    f->accuslip -= sample * direction;
    int last = (f->first + n - 1) % n;
    if (direction < 0) last = (last + 1) % n;
    if (!periodic) {
      // rank 0 == size-1 on one rank: the incoming plane is zeroed either way
      std::memset(f->re + last * plane, 0, sizeof(double) * plane);
      std::memset(f->im + last * plane, 0, sizeof(double) * plane);
    }
    f->first = last;
    if (direction < 0) f->first = (last + 1) % n;
    ++shifts;
  }
  return shifts;
}

// g4_beam_slice_reduce_kernel (DiagnosticCUDA.cpp:188-303), summed in
// particle order; gamma sums relative to gamma_shift (see file header).
void orc_diag_beam_raw(int nslice, const int32_t* off, const double* x, const double* y,
                       const double* px, const double* py, const double* gamma,
                       const double* theta, int nharm, double gamma_shift, double* raw) {
  const int stride = G4H_BEAM_RAW(nharm);
#pragma omp parallel for schedule(dynamic, 4)
  for (int is = 0; is < nslice; ++is) {
    double s[G4H_BEAM_STATS] = {0};
    s[G4H_B_XMIN] = 1.0e5;
    s[G4H_B_XMAX] = -1.0e5;
    s[G4H_B_PXMIN] = 1.0e5;
    s[G4H_B_PXMAX] = -1.0e5;
    s[G4H_B_YMIN] = 1.0e5;
    s[G4H_B_YMAX] = -1.0e5;
    s[G4H_B_PYMIN] = 1.0e5;
    s[G4H_B_PYMAX] = -1.0e5;
    s[G4H_B_GMIN] = 1.0e7;
    s[G4H_B_GMAX] = 1.0;
    for (int64_t ip = off[is]; ip < off[is + 1]; ++ip) {
      const double xv = x[ip], yv = y[ip], pxv = px[ip], pyv = py[ip], gv = gamma[ip];
      const double dg = gv - gamma_shift;
      s[G4H_B_G1] += dg;
      s[G4H_B_G2] += dg * dg;
      s[G4H_B_X1] += xv;
      s[G4H_B_X2] += xv * xv;
      s[G4H_B_Y1] += yv;
      s[G4H_B_Y2] += yv * yv;
      s[G4H_B_PX1] += pxv;
      s[G4H_B_PX2] += pxv * pxv;
      s[G4H_B_PY1] += pyv;
      s[G4H_B_PY2] += pyv * pyv;
      s[G4H_B_XPX] += xv * pxv;
      s[G4H_B_YPY] += yv * pyv;
      s[G4H_B_COUNT] += 1.0;
      s[G4H_B_XMIN] = std::fmin(s[G4H_B_XMIN], xv);
      s[G4H_B_XMAX] = std::fmax(s[G4H_B_XMAX], xv);
      s[G4H_B_PXMIN] = std::fmin(s[G4H_B_PXMIN], pxv);
      s[G4H_B_PXMAX] = std::fmax(s[G4H_B_PXMAX], pxv);
      s[G4H_B_YMIN] = std::fmin(s[G4H_B_YMIN], yv);
      s[G4H_B_YMAX] = std::fmax(s[G4H_B_YMAX], yv);
      s[G4H_B_PYMIN] = std::fmin(s[G4H_B_PYMIN], pyv);
      s[G4H_B_PYMAX] = std::fmax(s[G4H_B_PYMAX], pyv);
      s[G4H_B_GMIN] = std::fmin(s[G4H_B_GMIN], gv);
      s[G4H_B_GMAX] = std::fmax(s[G4H_B_GMAX], gv);
    }
    double* out = raw + static_cast<int64_t>(is) * stride;
    for (int k = 0; k < G4H_BEAM_STATS; ++k) out[k] = s[k];
    for (int ih = 0; ih < nharm; ++ih) {
      double br = 0.0, bi = 0.0;
      const double h = static_cast<double>(ih + 1);
      for (int64_t ip = off[is]; ip < off[is + 1]; ++ip) {
        br += std::cos(h * theta[ip]);
        bi += std::sin(h * theta[ip]);
      }
      out[G4H_BEAM_STATS + 2 * ih] = br;
      out[G4H_BEAM_STATS + 2 * ih + 1] = bi;
    }
  }
}

// g4_field_slice_reduce_kernel (DiagnosticCUDA.cpp:305-363)
void orc_diag_field_raw(const orc_field* f, double* raw) {
  const int ng = f->ngrid, n = f->nslice;
  const int64_t plane = static_cast<int64_t>(ng) * ng;
  const double shift = -0.5 * static_cast<double>(ng - 1);
#pragma omp parallel for schedule(dynamic, 1)
  for (int is0 = 0; is0 < n; ++is0) {
    const int is = (n + is0 - f->first) % n;
    const double* re = f->re + is0 * plane;
    const double* im = f->im + is0 * plane;
    double p = 0, x1 = 0, x2 = 0, y1 = 0, y2 = 0, fr = 0, fi = 0;
    for (int64_t ixy = 0; ixy < plane; ++ixy) {
      const int iy = static_cast<int>(ixy / ng);
      const int ix = static_cast<int>(ixy - static_cast<int64_t>(iy) * ng);
      const double wei = re[ixy] * re[ixy] + im[ixy] * im[ixy];
      const double dx = static_cast<double>(ix) + shift;
      const double dy = static_cast<double>(iy) + shift;
      p += wei;
      x1 += dx * wei;
      x2 += dx * dx * wei;
      y1 += dy * wei;
      y2 += dy * dy * wei;
      fr += re[ixy];
      fi += im[ixy];
    }
    double* o = raw + static_cast<int64_t>(is) * G4H_FIELD_RAW;
    o[G4H_F_POWER] = p;
    o[G4H_F_X1] = x1;
    o[G4H_F_X2] = x2;
    o[G4H_F_Y1] = y1;
    o[G4H_F_Y2] = y2;
    o[G4H_F_FF_RE] = fr;
    o[G4H_F_FF_IM] = fi;
    const int64_t ic = (plane - 1) / 2;
    o[G4H_F_CENTER_RE] = re[ic];
    o[G4H_F_CENTER_IM] = im[ic];
  }
}

// DiagnosticCUDA::calcBeam host part (DiagnosticCUDA.cpp:509-628) with all
// output filters on; q is nslice x G4H_BEAM_QN(nharm), global 6 values.
void orc_finalize_beam(int nslice, int nharm, const double* raw, const double* current,
                       double gamma_shift, double* q, double* global) {
  const int rs = G4H_BEAM_RAW(nharm), qs = G4H_BEAM_QN(nharm);
  double gc = 0, gg1 = 0, gg2 = 0, gx1 = 0, gx2 = 0, gy1 = 0, gy2 = 0;
  for (int is = 0; is < nslice; ++is) {
    const double* s = raw + static_cast<int64_t>(is) * rs;
    double* o = q + static_cast<int64_t>(is) * qs;
    const double count = s[G4H_B_COUNT];
    const double norm = count > 0.0 ? 1.0 / count : 1.0;
    const double x1 = s[G4H_B_X1] * norm, x2 = s[G4H_B_X2] * norm;
    const double y1 = s[G4H_B_Y1] * norm, y2 = s[G4H_B_Y2] * norm;
    const double px1 = s[G4H_B_PX1] * norm, px2 = s[G4H_B_PX2] * norm;
    const double py1 = s[G4H_B_PY1] * norm, py2 = s[G4H_B_PY2] * norm;
    const double d1 = s[G4H_B_G1] * norm, d2 = s[G4H_B_G2] * norm;  // shifted moments
    const double g1 = count > 0.0 ? gamma_shift + d1 : 0.0;
    const double xpx = s[G4H_B_XPX] * norm, ypy = s[G4H_B_YPY] * norm;
    o[G4H_Q_ENERGY] = g1;
    o[G4H_Q_ENERGYSPREAD] = std::sqrt(std::fabs(d2 - d1 * d1));
    o[G4H_Q_XPOSITION] = x1;
    o[G4H_Q_XSIZE] = std::sqrt(std::fabs(x2 - x1 * x1));
    o[G4H_Q_YPOSITION] = y1;
    o[G4H_Q_YSIZE] = std::sqrt(std::fabs(y2 - y1 * y1));
    o[G4H_Q_PXPOSITION] = px1;
    o[G4H_Q_PYPOSITION] = py1;
    const double cur = current ? current[is] : 0.0;
    o[G4H_Q_CURRENT] = cur;
    const double ex =
        std::sqrt(std::fabs((x2 - x1 * x1) * (px2 - px1 * px1) - (xpx - x1 * px1) * (xpx - x1 * px1)));
    const double ey =
        std::sqrt(std::fabs((y2 - y1 * y1) * (py2 - py1 * py1) - (ypy - y1 * py1) * (ypy - y1 * py1)));
    o[G4H_Q_EMITX] = ex;
    o[G4H_Q_EMITY] = ey;
    o[G4H_Q_BETAX] = ex > 0.0 ? (x2 - x1 * x1) / ex * g1 : 0.0;
    o[G4H_Q_BETAY] = ey > 0.0 ? (y2 - y1 * y1) / ey * g1 : 0.0;
    o[G4H_Q_ALPHAX] = ex > 0.0 ? -(xpx - x1 * px1) / ex : 0.0;
    o[G4H_Q_ALPHAY] = ey > 0.0 ? -(ypy - y1 * py1) / ey : 0.0;
    o[G4H_Q_XMIN] = s[G4H_B_XMIN];
    o[G4H_Q_XMAX] = s[G4H_B_XMAX];
    o[G4H_Q_PXMIN] = s[G4H_B_PXMIN];
    o[G4H_Q_PXMAX] = s[G4H_B_PXMAX];
    o[G4H_Q_YMIN] = s[G4H_B_YMIN];
    o[G4H_Q_YMAX] = s[G4H_B_YMAX];
    o[G4H_Q_PYMIN] = s[G4H_B_PYMIN];
    o[G4H_Q_PYMAX] = s[G4H_B_PYMAX];
    o[G4H_Q_EMIN] = s[G4H_B_GMIN];
    o[G4H_Q_EMAX] = s[G4H_B_GMAX];
    for (int ih = 0; ih < nharm; ++ih) {
      const double bre = s[G4H_BEAM_STATS + 2 * ih] * norm;
      const double bim = s[G4H_BEAM_STATS + 2 * ih + 1] * norm;
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
  const double norm = gc > 0.0 ? 1.0 / gc : 1.0;
  gg1 *= norm;
  gg2 *= norm;
  gx1 *= norm;
  gx2 *= norm;
  gy1 *= norm;
  gy2 *= norm;
  global[G4H_BG_ENERGY] = gamma_shift + gg1;
  global[G4H_BG_ENERGYSPREAD] = std::sqrt(std::fabs(gg2 - gg1 * gg1));
  global[G4H_BG_XPOSITION] = gx1;
  global[G4H_BG_XSIZE] = std::sqrt(std::fabs(gx2 - gx1 * gx1));
  global[G4H_BG_YPOSITION] = gy1;
  global[G4H_BG_YSIZE] = std::sqrt(std::fabs(gy2 - gy1 * gy1));
}

// DiagnosticCUDA::calcField host part (DiagnosticCUDA.cpp:685-790), one rank.
void orc_finalize_field(int nslice, int ngrid, double dgrid, double xlambda,
                        double slicelength, const double* raw, double* q,
                        double* global) {
  (void)ngrid;
  const double ks = 4.0 * std::asin(1.0) / xlambda;
  const double scl = dgrid * kEev / ks;
  double gp = 0, gx1 = 0, gx2 = 0, gy1 = 0, gy2 = 0, gff = 0, gin = 0;
  for (int is = 0; is < nslice; ++is) {
    const double* s = raw + static_cast<int64_t>(is) * G4H_FIELD_RAW;
    double* o = q + static_cast<int64_t>(is) * G4H_FIELD_Q;
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
    const double iphi = std::atan2(cim, cre);
    double ff = s[G4H_F_FF_RE] * s[G4H_F_FF_RE] + s[G4H_F_FF_IM] * s[G4H_F_FF_IM];
    const double ffphi = std::atan2(s[G4H_F_FF_IM], s[G4H_F_FF_RE]);
    ff *= dgrid * dgrid;
    gff += ff;
    gin += inten;
    o[G4H_FQ_POWER] = pr * scl * scl / kVacimp;
    o[G4H_FQ_XPOSITION] = x1 * dgrid;
    o[G4H_FQ_XSIZE] = x2 * dgrid;
    o[G4H_FQ_YPOSITION] = y1 * dgrid;
    o[G4H_FQ_YSIZE] = y2 * dgrid;
    o[G4H_FQ_INTENSITY_NF] = inten * kEev * kEev / ks / ks / kVacimp;
    o[G4H_FQ_PHASE_NF] = iphi;
    o[G4H_FQ_INTENSITY_FF] = ff;
    o[G4H_FQ_PHASE_FF] = ffphi;
  }
  const double norm = gp > 0.0 ? 1.0 / gp : 1.0;
  gx1 *= norm;
  gx2 *= norm;
  gy1 *= norm;
  gy2 *= norm;
  global[G4H_FG_ENERGY] = gp * scl * scl / kVacimp * slicelength / kC;
  global[G4H_FG_XPOSITION] = gx1 * dgrid;
  global[G4H_FG_XSIZE] = std::sqrt(std::fabs(gx2 - gx1 * gx1)) * dgrid;
  global[G4H_FG_YPOSITION] = gy1 * dgrid;
  global[G4H_FG_YSIZE] = std::sqrt(std::fabs(gy2 - gy1 * gy1)) * dgrid;
  const double avg = 1.0 / static_cast<double>(nslice);
  global[G4H_FG_INTENSITY_NF] = gin * avg * kEev * kEev / ks / ks / kVacimp;
  global[G4H_FG_INTENSITY_FF] = gff * avg;
}

// Gencore::run loop body (Gencore.cpp:172-254) for a single field.
// Short-range space charge (EFieldSolver::shortRange, called by the
// reference at BeamSolverCUDA.cpp:523-547; upstream GENESIS 1.3 owns it and
// it is absent from the reference tree), restated from the published method.
// Per slice, around the centroid (xc, yc) of its particles: radial shells
// n = 0..N-1 of width dr = rmax * max|r - rc| / N; the charge density of the
// slice (charge I*slicelength/c in macro particles of equal charge, the phase
// structure periodic in the radiation wavelength) is expanded in harmonics
// l = 1..nz of the phase theta and azimuthal modes m = -nphi..nphi:
//   S_lmn = sum over particles in shell n of exp(-i (l theta + m phi)).
// The mode's potential solves (1/r)(r u')' - (m^2/r^2 + kappa^2) u =
// -rho/eps0 with kappa = l ks / gz (the Lorentz factor of the longitudinal
// motion, gz2 = gamma0^2/(1+aw^2)), u = 0 at r = N dr, discretised at the
// shell centres r_n = (n+1/2) dr with fluxes through the shell faces n dr and
// (n+1) dr:
//   n u_{n-1} - [2n+1 + m^2/(n+1/2) + (n+1/2)(kappa dr)^2] u_n + (n+1) u_{n+1}
//     = (I vacimp / (2 pi Np)) S_lmn                      (charge -e)
// (Thomas), and the field of harmonic l in a relativistic beam is
// E_z = -i (l ks / gz2) u e^{i(l theta + m phi)} + c.c.  The push takes
// E_z/511000 (BeamSolverCUDA.cpp:532) as its longitudinal term.
void orc_sc_field(int nslice, const int32_t* off, const double* x, const double* y,
                  const double* theta, const double* current, double ks, double gz2,
                  const g4h_efield* ef, double* ez, double* slice_rms) {
  const int N = ef ? ef->ngrid : 0;
  const int nz = ef ? ef->nz : 0;
  const int nphi = ef ? ef->nphi : 0;
  const int nm = 2 * nphi + 1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int is = 0; is < nslice; ++is) {
    const int64_t i0 = off[is], i1 = off[is + 1];
    const int64_t np = i1 - i0;
    for (int64_t i = i0; i < i1; ++i) ez[i] = 0.0;
    if (slice_rms) slice_rms[is] = 0.0;
    if (np <= 0 || nz <= 0 || N <= 0 || current[is] == 0.0) continue;
    double sx = 0.0, sy = 0.0;
    for (int64_t i = i0; i < i1; ++i) {
      sx += x[i];
      sy += y[i];
    }
    const double xc = sx / static_cast<double>(np), yc = sy / static_cast<double>(np);
    double r2 = 0.0;
    for (int64_t i = i0; i < i1; ++i) {
      const double dx = x[i] - xc, dy = y[i] - yc;
      r2 = std::max(r2, dx * dx + dy * dy);
    }
    const double dr = ef->rmax * std::sqrt(r2) / N;
    if (!(dr > 0.0)) continue;
    std::vector<int> shell(static_cast<size_t>(np));
    std::vector<std::complex<double>> eth(static_cast<size_t>(np)), eph(static_cast<size_t>(np));
    std::vector<std::complex<double>> S(static_cast<size_t>(nz) * nm * N, 0.0);
    for (int64_t i = i0; i < i1; ++i) {
      const double dx = x[i] - xc, dy = y[i] - yc;
      const double r = std::sqrt(dx * dx + dy * dy);
      const size_t k = static_cast<size_t>(i - i0);
      shell[k] = std::min(static_cast<int>(r / dr), N - 1);
      eph[k] = r > 0.0 ? std::complex<double>(dx / r, dy / r) : std::complex<double>(1.0, 0.0);
      eth[k] = std::complex<double>(std::cos(theta[i]), std::sin(theta[i]));
      std::complex<double> el = 1.0;
      for (int l = 1; l <= nz; ++l) {
        el *= eth[k];
        // em = e^{i m phi} for m = -nphi..nphi; the source term is its
        // product with e^{i l theta}, conjugated
        std::complex<double> em = 1.0;
        for (int q = 0; q < nphi; ++q) em *= std::conj(eph[k]);  // e^{-i nphi phi}
        for (int m = -nphi; m <= nphi; ++m) {
          S[(static_cast<size_t>(l - 1) * nm + (m + nphi)) * N + shell[k]] += std::conj(el * em);
          em *= eph[k];
        }
      }
    }
    const double d0 = current[is] * kVacimp / (2.0 * M_PI * static_cast<double>(np));
    std::vector<double> cp(static_cast<size_t>(N));
    std::vector<std::complex<double>> u(static_cast<size_t>(N));
    for (int l = 1; l <= nz; ++l) {
      const double kd = l * ks * dr;
      const double kd2 = kd * kd / gz2;
      for (int m = -nphi; m <= nphi; ++m) {
        std::complex<double>* s = &S[(static_cast<size_t>(l - 1) * nm + (m + nphi)) * N];
        // Thomas: a_n = n, b_n, c_n = n + 1, u_N = 0
        for (int n = 0; n < N; ++n) {
          const double h = n + 0.5;
          const double b = -(2.0 * n + 1.0) - static_cast<double>(m * m) / h - h * kd2;
          const double a = static_cast<double>(n);
          const double den = n == 0 ? b : b - a * cp[static_cast<size_t>(n - 1)];
          cp[static_cast<size_t>(n)] = (n + 1.0) / den;
          const std::complex<double> d = d0 * s[n];
          u[static_cast<size_t>(n)] = n == 0 ? d / den : (d - a * u[static_cast<size_t>(n - 1)]) / den;
        }
        for (int n = N - 2; n >= 0; --n)
          u[static_cast<size_t>(n)] -= cp[static_cast<size_t>(n)] * u[static_cast<size_t>(n + 1)];
        for (int n = 0; n < N; ++n) s[n] = u[static_cast<size_t>(n)];
      }
    }
    double esum = 0.0;
    for (int64_t i = i0; i < i1; ++i) {
      const size_t k = static_cast<size_t>(i - i0);
      double e = 0.0;
      std::complex<double> el = 1.0;
      for (int l = 1; l <= nz; ++l) {
        el *= eth[k];
        std::complex<double> em = 1.0;
        for (int q = 0; q < nphi; ++q) em *= std::conj(eph[k]);
        for (int m = -nphi; m <= nphi; ++m) {
          const std::complex<double> uu =
              S[(static_cast<size_t>(l - 1) * nm + (m + nphi)) * N + shell[k]];
          // 2 Re[-i (l ks/gz2) u e^{i(l theta + m phi)}] = 2 (l ks/gz2) Im[u e^{..}]
          e += 2.0 * (l * ks / gz2) * std::imag(uu * el * em);
          em *= eph[k];
        }
      }
      ez[i] = e / 511000.0;
      esum += e * e;
    }
    if (slice_rms) slice_rms[is] = std::sqrt(esum / static_cast<double>(np));
  }
}

static int run_impl(int nslice, const int32_t* off, double* x, double* y, double* px,
                    double* py, double* gamma, double* theta, double gamma0, double lambda0,
                    const double* current, const double* slice_eloss, orc_field* field,
                    int solver, int do_filter, double xc, double yc, double sig,
                    const g4h_step* steps, int nsteps, double sample, int timerun,
                    int periodic, int nharm, double* beam_raw, double* field_raw,
                    const g4h_efield* ef) {
  const int64_t np = off[nslice];
  const int64_t plane = static_cast<int64_t>(field->ngrid) * field->ngrid;
  const int rs = G4H_BEAM_RAW(nharm);
  std::vector<double> sre(plane * field->nslice), sim(plane * field->nslice);
  std::vector<double> ez(ef && ef->nz > 0 ? static_cast<size_t>(np) : 0);
  int rec = 0;
  auto record = [&]() {
    if (beam_raw)
      orc_diag_beam_raw(nslice, off, x, y, px, py, gamma, theta, nharm, gamma0,
                        beam_raw + static_cast<int64_t>(rec) * nslice * rs);
    if (field_raw)
      orc_diag_field_raw(field, field_raw + static_cast<int64_t>(rec) * nslice * G4H_FIELD_RAW);
    ++rec;
  };
  record();
  for (int i = 0; i < nsteps; ++i) {
    const g4h_step* st = steps + i;
    const double delz = st->delz;
    // Beam::track (Beam.cpp:111-134)
    orc_track_transverse(np, x, y, px, py, gamma, st, gamma0, delz * 0.5, 0);
    if (ef && ef->nz > 0) {
      // BeamSolverCUDA::advance: short-range field + slice loss per particle
      const double gz2 = gamma0 * gamma0 / (1.0 + st->aw * st->aw);
      orc_sc_field(nslice, off, x, y, theta, current, field->xks, gz2, ef, ez.data(), nullptr);
      if (slice_eloss)
        for (int is = 0; is < nslice; ++is)
          for (int64_t ip = off[is]; ip < off[is + 1]; ++ip) ez[ip] += slice_eloss[is];
      orc_beam_advance_ez(nslice, off, x, y, px, py, gamma, theta, field, 1, st, gamma0, delz,
                          ez.data());
    } else {
      orc_beam_advance(nslice, off, x, y, px, py, gamma, theta, field, 1, st, gamma0, delz,
                       slice_eloss);
    }
    orc_apply_r56(np, gamma, theta, st, gamma0, lambda0);
    orc_track_transverse(np, x, y, px, py, gamma, st, gamma0, delz * 0.5, 1);
    // Field::track (Field.cpp:204-212)
    const int on = orc_deposit(nslice, off, x, y, gamma, theta, field, st, current, delz,
                               sre.data(), sim.data());
    if (solver == G4H_SOLVER_FFT)
      orc_fft_advance(field, on ? sre.data() : nullptr, on ? sim.data() : nullptr, delz,
                      do_filter, xc, yc, sig);
    else
      orc_adi_advance(field, on ? sre.data() : nullptr, on ? sim.data() : nullptr, delz);
    // Control::applySlippage (Gencore.cpp:228-230)
    orc_apply_slippage(field, st->slippage, sample, timerun, periodic);
    if (st->outstep) record();
  }
  return rec;
}

int orc_run(int nslice, const int32_t* off, double* x, double* y, double* px,
            double* py, double* gamma, double* theta, double gamma0, double lambda0,
            const double* current, const double* slice_eloss, orc_field* field,
            int solver, int do_filter, double xc, double yc, double sig,
            const g4h_step* steps, int nsteps, double sample, int timerun,
            int periodic, int nharm, double* beam_raw, double* field_raw) {
  return run_impl(nslice, off, x, y, px, py, gamma, theta, gamma0, lambda0, current, slice_eloss,
                  field, solver, do_filter, xc, yc, sig, steps, nsteps, sample, timerun, periodic,
                  nharm, beam_raw, field_raw, nullptr);
}

int orc_run_sc(int nslice, const int32_t* off, double* x, double* y, double* px,
               double* py, double* gamma, double* theta, double gamma0, double lambda0,
               const double* current, const double* slice_eloss, orc_field* field,
               int solver, int do_filter, double xc, double yc, double sig,
               const g4h_step* steps, int nsteps, double sample, int timerun,
               int periodic, int nharm, double* beam_raw, double* field_raw,
               const g4h_efield* ef) {
  return run_impl(nslice, off, x, y, px, py, gamma, theta, gamma0, lambda0, current, slice_eloss,
                  field, solver, do_filter, xc, yc, sig, steps, nsteps, sample, timerun, periodic,
                  nharm, beam_raw, field_raw, ef);
}

}  // extern "C"
