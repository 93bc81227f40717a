// particle.hip - the fused particle pass for gfx950.
//
// One workgroup owns one beam slice (particles [off[s], off[s+1]) are
// contiguous) and runs, per particle and as selected by ParticleParams.stages:
//   first transverse half step  (TrackBeamCUDA.cpp:156-204, chicane :220-255)
//   bilinear gather + RK4 push  (BeamSolverCUDA.cpp:87-304, multi-field :306-409)
//   R56                          (TrackBeamCUDA.cpp:257-278)
//   last transverse half step   (corrector :206-218)
//   source deposition           (FieldSolverADICUDA.cpp:109-174)
//   per-slice beam sums          (DiagnosticCUDA.cpp:188-303)
// so the particle state crosses HBM once per step (read + write).
//
// Deposition: the workgroup accumulates into an LDS window (FP64 complex
// accumulators, ds_add_f64) placed on the slice's deposition footprint of the
// previous step plus a 2-cell margin; cells outside the window go straight to
// the source plane with global atomics.  After the slice the window is
// written with plain stores (the workgroup owns the slice's source plane), so
// each touched cell costs one global write instead of one atomic per
// particle.  FP64 accumulation of the per-particle terms is the compensated
// sum: it carries 29 more bits than an FP32 sum, which is what a Kahan
// compensation term would add, at the cost of one LDS atomic.
//
// Few slices (steady-state batches, short time windows) would leave most
// CUs idle, so such a pass is split: K workgroups per slice (gridDim.y) each
// take a contiguous part of its particles with the same window placement,
// write their window and sums to scratch, and the last to arrive (an atomic
// counter per slice) adds the K windows in part order and writes the plane
// and the slice's sums.  The previous source rectangle is then zeroed by a
// pass before, so the fallback atomics of every part land on cleared cells.
//
// MIXED arithmetic: x, y, px, py and dgamma = gamma - gamma_ref in FP32,
// theta in FP64.  The phase equation theta' = xks(1 - 1/beta_par) + xku is
// evaluated around the reference particle,
//   theta' = -(c0 + k1*eta + k2*eta^2),  eta = eps/eps0 - 1,
//   eps = btper0/gamma^2, eps0 = (1+aw^2)/gamma_ref^2,
// with c0, k1, k2 from FP64 on the host; eta is small and well conditioned in
// FP32, so FP32 keeps ~1e-9 rad/m absolute accuracy where the direct form
// would cancel 240 - 240 in FP32.
#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "adi_lines.hpp"
#include "kernels.hpp"

namespace g4 {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxHarm = 8;
constexpr int kNMinMax = 10;

template <class P>
struct V2;
template <>
struct V2<double> {
  using T = double2;
};
template <>
struct V2<float> {
  using T = float2;
};

template <class P>
struct BeamPtr {
  P *x, *y, *px, *py, *g;  // g: gamma (FP64) or gamma - gamma_ref (MIXED)
  double* th;
};

// MIXED: an FP64 phase reduced to [-pi, pi] in FP64 (two-term 2*pi) and
// rounded to FP32.
__device__ __forceinline__ float reduce_phase(double a) {
  const double n = rint(a * 0.15915494309189535);
  double r = fma(-n, 6.283185307179586, a);
  r = fma(-n, 2.4492935982947064e-16, r);
  return static_cast<float>(r);
}

// MIXED sine and cosine of an FP64 phase: the reduced phase through the
// hardware v_sin_f32 / v_cos_f32 (absolute error ~1e-6, far inside the 1e-3
// mixed-precision budget).
__device__ __forceinline__ void fsincos(double a, float& s, float& c) {
  const float rf = reduce_phase(a);
  s = __sinf(rf);
  c = __cosf(rf);
}

// ---- transverse ---------------------------------------------------------
// advance_focusing_plane (TrackBeamCUDA.cpp:27-59) in the reference's
// trigonometric form; FP64 takes it where the power series below does not
// apply (a strong quadrupole over a long step).
template <class P>
__device__ __forceinline__ void plane_step(int mode, P str, P off, P delz, P& coord, P& mom,
                                           P gz) {
  if (mode == 0) {
    coord += mom * delz / gz;
    return;
  }
  if (mode == 1) {
    const P foc = sqrt(str / gz);
    const P omg = foc * delz;
    const P a1 = cos(omg);
    const P a2 = sin(omg) / foc;
    const P a3 = -a2 * foc * foc;
    const P xt = coord - off;
    coord = a1 * xt + a2 * mom / gz + off;
    mom = a3 * xt * gz + a1 * mom;
    return;
  }
  const P foc = sqrt(-str / gz);
  const P omg = foc * delz;
  const P a1 = cosh(omg);
  const P a2 = sinh(omg) / foc;
  const P a3 = a2 * foc * foc;
  const P xt = coord - off;
  coord = a1 * xt + a2 * mom / gz + off;
  mom = a3 * xt * gz + a1 * mom;
}

// MIXED overload: fast reciprocal / square root and hardware sin/cos (the
// focusing phase advance per half step is ~1e-2 rad).  With `poly` > 0 the
// map is taken from its power series in w = K dz^2 (K = str/gammaz, either
// sign: cos(sqrt(w)) and sin(sqrt(w))/sqrt(w) are entire in w and cover
// focusing and defocusing alike) and no transcendental is evaluated: to w^4
// for |w| <= 1/2 (poly 1), to w^2 for |w| <= 0.011 (poly 2).  The series
// alternate with decreasing terms, so the remainder is below the first
// omitted term: w^5/10! < 1e-8 and w^3/6! < 2e-9, under FP32 rounding.
__device__ __forceinline__ void plane_step(int mode, float str, float off, float delz,
                                           float& coord, float& mom, float gz, float igz,
                                           int poly) {
  if (mode == 0) {
    coord += mom * delz * igz;
    return;
  }
  if (poly) {
    const float k = str * igz;
    const float w = k * delz * delz;
    float c, sn;
    if (poly == 2) {
      c = fmaf(fmaf(w, 1.0f / 24.0f, -0.5f), w, 1.0f);  // cos(sqrt(w)) = 1 - w/2 + w^2/24
      sn = fmaf(fmaf(w, 1.0f / 120.0f, -1.0f / 6.0f), w, 1.0f);  // 1 - w/6 + w^2/120
    } else {
      c = 1.0f / 40320.0f;  // cos(sqrt(w)) = 1 - w/2 + w^2/24 - ...
      c = fmaf(c, w, -1.0f / 720.0f);
      c = fmaf(c, w, 1.0f / 24.0f);
      c = fmaf(c, w, -0.5f);
      c = fmaf(c, w, 1.0f);
      sn = 1.0f / 362880.0f;  // sin(sqrt(w))/sqrt(w) = 1 - w/6 + w^2/120 - ...
      sn = fmaf(sn, w, -1.0f / 5040.0f);
      sn = fmaf(sn, w, 1.0f / 120.0f);
      sn = fmaf(sn, w, -1.0f / 6.0f);
      sn = fmaf(sn, w, 1.0f);
    }
    const float a2 = delz * sn;  // sin(omega)/foc
    const float xt = coord - off;
    coord = fmaf(c, xt, a2 * mom * igz) + off;
    mom = fmaf(-k * a2 * gz, xt, c * mom);
    return;
  }
  if (mode == 1) {
    const float foc = __builtin_amdgcn_sqrtf(str * igz);
    const float omg = foc * delz;
    const float a1 = __cosf(omg);
    const float sn = __sinf(omg);
    const float a2 = sn * __builtin_amdgcn_rcpf(foc);
    const float a3 = -sn * foc;
    const float xt = coord - off;
    coord = a1 * xt + a2 * mom * igz + off;
    mom = a3 * xt * gz + a1 * mom;
    return;
  }
  const float foc = __builtin_amdgcn_sqrtf(-str * igz);
  const float omg = foc * delz;
  const float a1 = coshf(omg);
  const float sh = sinhf(omg);
  const float a2 = sh * __builtin_amdgcn_rcpf(foc);
  const float a3 = sh * foc;
  const float xt = coord - off;
  coord = a1 * xt + a2 * mom * igz + off;
  mom = a3 * xt * gz + a1 * mom;
}

// An FP64 constant in a scalar register pair, materialized where it is used.
// gfx950 has no 64-bit literal operands; left to itself the compiler builds
// each polynomial coefficient in a vector register pair (two v_mov_b32) and
// accumulates with v_fmac_f64: three vector instructions per Horner step.
// Passed through a scalar move that cannot be hoisted, the coefficient is the
// scalar operand of one v_fma_f64, and its moves run on the scalar unit.
__device__ __forceinline__ double sconst(double c) {
  double r;
  asm volatile("s_mov_b64 %0, %1" : "=s"(r) : "s"(c));
  return r;
}

// FP64 with `poly` > 0: the power series of the MIXED overload below, to
// w^8 for |w| <= 1/2 (poly 1; remainder w^9/18! < 4e-19) or to w^4 for
// |w| <= 0.011 (poly 2; w^5/10! < 5e-17), either under FP64 rounding, with
// one reciprocal of gammaz for both planes; otherwise the reference's
// trigonometric form.
__device__ __forceinline__ void plane_step(int mode, double str, double off, double delz,
                                           double& coord, double& mom, double gz, double igz,
                                           int poly) {
  if (mode == 0) {
    coord += mom * delz * igz;
    return;
  }
  if (!poly) {
    plane_step<double>(mode, str, off, delz, coord, mom, gz);
    return;
  }
  const double k = str * igz;
  const double w = k * delz * delz;
  double c, sn;
  if (poly == 2) {
    c = fma(sconst(1.0 / 40320.0), w, sconst(-1.0 / 720.0));  // (-w)^n / (2n)!, n <= 4
    c = fma(c, w, sconst(1.0 / 24.0));
    c = fma(c, w, -0.5);
    c = fma(c, w, 1.0);
    sn = fma(sconst(1.0 / 362880.0), w, sconst(-1.0 / 5040.0));  // (-w)^n / (2n+1)!
    sn = fma(sn, w, sconst(1.0 / 120.0));
    sn = fma(sn, w, sconst(-1.0 / 6.0));
    sn = fma(sn, w, 1.0);
  } else {
    c = sconst(1.0 / 20922789888000.0);  // cos(sqrt(w)): (-w)^n / (2n)!, n <= 8
    c = fma(c, w, sconst(-1.0 / 87178291200.0));
    c = fma(c, w, sconst(1.0 / 479001600.0));
    c = fma(c, w, sconst(-1.0 / 3628800.0));
    c = fma(c, w, sconst(1.0 / 40320.0));
    c = fma(c, w, sconst(-1.0 / 720.0));
    c = fma(c, w, sconst(1.0 / 24.0));
    c = fma(c, w, -0.5);
    c = fma(c, w, 1.0);
    sn = sconst(1.0 / 355687428096000.0);  // sin(sqrt(w))/sqrt(w): (-w)^n / (2n+1)!
    sn = fma(sn, w, sconst(-1.0 / 1307674368000.0));
    sn = fma(sn, w, sconst(1.0 / 6227020800.0));
    sn = fma(sn, w, sconst(-1.0 / 39916800.0));
    sn = fma(sn, w, sconst(1.0 / 362880.0));
    sn = fma(sn, w, sconst(-1.0 / 5040.0));
    sn = fma(sn, w, sconst(1.0 / 120.0));
    sn = fma(sn, w, sconst(-1.0 / 6.0));
    sn = fma(sn, w, 1.0);
  }
  const double a2 = delz * sn;  // sin(omega)/foc
  const double xt = coord - off;
  coord = fma(c, xt, a2 * mom * igz) + off;
  mom = fma(-k * a2 * gz, xt, c * mom);
}

__device__ __forceinline__ double psqrt(double v) { return sqrt(v); }
__device__ __forceinline__ float psqrt(float v) { return __builtin_amdgcn_sqrtf(v); }

// ---- FP64 reciprocals and square roots ------------------------------------
// 1/v and 1/sqrt(v) from the hardware estimates and two Newton steps, to
// within an ulp: for quantities that no cancellation amplifies (transverse
// maps, deposition weights, dgamma/dz), and as the approximate reciprocals
// of the correctly rounded forms below.
__device__ __forceinline__ double rcp_nr(double v) {
  double y = __builtin_amdgcn_rcp(v);
  double e = fma(-v, y, 1.0);
  y = fma(y, e, y);
  e = fma(-v, y, 1.0);
  return fma(y, e, y);
}
__device__ __forceinline__ double rsq_nr(double v) {
  double y = __builtin_amdgcn_rsq(v);
#pragma unroll
  for (int it = 0; it < 2; ++it) y = fma(0.5 * y, fma(-v * y, y, 1.0), y);
  return y;
}

// The reference's q = 1 / fl(sqrt(fl(1 - fl(a / b)))) with every operation
// correctly rounded (BeamSolverCUDA.cpp:145-223): theta' = xks (1 - q) + xku
// cancels ~8 digits, so q must round exactly as there.  For b = gamma^2 the
// square root is of u = 1 - d with d ~ aw^2/gamma^2 tiny, and both it and
// the reciprocal lie next to one.  Each result is an estimate within a few
// ulp (a * rb with rb ~ 1/b; the series 1 - d/2 - d^2/8 and 1 + e + e^2)
// followed by one Newton correction whose residual is exact in an FMA
// (a - q b, u - r^2, 1 - r q) and whose one rounding is then the correctly
// rounded result: the error left before it is ~(estimate error)^2, far
// below the distance to a rounding boundary except for a fraction ~2^-50
// of inputs.  Outside 0 <= d <= 2^-16 the division, sqrt and reciprocal
// are taken as such.
__device__ __forceinline__ double inv_btpar(double a, double b, double rb) {
  const double q1 = a * rb;
  const double x = fma(fma(-q1, b, a), rb, q1);  // fl(a / b)
  const double u = 1.0 - x;
  const double d = 1.0 - u;  // exact for u in [1/2, 2]
  if (d >= 0.0 && d <= 0x1p-16) {
    const double r0 = 1.0 - d * fma(d, 0.125, 0.5);
    const double r = fma(fma(-r0, r0, u), fma(0.25, d, 0.5), r0);  // fl(sqrt(u))
    const double e = 1.0 - r;                                      // exact
    const double q0 = 1.0 + fma(e, e, e);
    return fma(fma(-r, q0, 1.0), q0, q0);  // fl(1 / r)
  }
  return 1.0 / sqrt(1.0 - a / b);
}

// gamma^2 - 1 - extra (extra = aw^2 + px^2 + py^2 or px^2 + py^2)
__device__ __forceinline__ double gz2(double g, double g0, double g2c, double aw2, bool with_aw) {
  (void)g0;
  (void)g2c;
  return with_aw ? g * g - 1.0 - aw2 : g * g - 1.0;
}
__device__ __forceinline__ float gz2(float dg, double g0, double g2c, double aw2, bool with_aw) {
  (void)aw2;
  (void)with_aw;
  return static_cast<float>(g2c) + dg * (2.0f * static_cast<float>(g0) + dg);
}

template <class P>
__device__ __forceinline__ void transverse(const TransParams& t, double g0, P g, P& x, P& y, P& px,
                                           P& py, bool first_half) {
  if (first_half && t.chicane) {
    const P gz = psqrt(gz2(g, g0, t.g2chic, 0.0, false) - px * px - py * py);
    P tmp = x;
    x = P(t.m00) * tmp + P(t.m01) * px / gz;
    px = P(t.m10) * tmp * gz + P(t.m11) * px;
    tmp = y;
    y = P(t.m22) * tmp + P(t.m23) * py / gz;
    py = P(t.m32) * tmp * gz + P(t.m33) * py;
  }
  if (!first_half && t.corrector) {
    px += P(t.cx);
    py += P(t.cy);
  }
  if constexpr (sizeof(P) == 8) {
    const double gzz = gz2(g, g0, t.g2c, t.aw2, true) - px * px - py * py;
    const double igz = rsq_nr(gzz);  // 1/gammaz
    const double gz = gzz * igz;
    plane_step(t.xmode, t.xstr, t.xoff, t.delz, x, px, gz, igz, t.xpoly);
    plane_step(t.ymode, t.ystr, t.yoff, t.delz, y, py, gz, igz, t.ypoly);
  } else {
    const float gzz = gz2(g, g0, t.g2c, t.aw2, true) - px * px - py * py;
    const float igz = __builtin_amdgcn_rsqf(gzz);  // 1/gammaz
    const float gz = gzz * igz;
    plane_step(t.xmode, float(t.xstr), float(t.xoff), float(t.delz), x, px, gz, igz, t.xpoly);
    plane_step(t.ymode, float(t.ystr), float(t.yoff), float(t.delz), y, py, gz, igz, t.ypoly);
  }
}

// ---- particle -> cell (BeamSolverCUDA.cpp:99-116, FieldSolverADICUDA.cpp:
// 126-143): lower-left cell (ix, iy) and the weight wx, wy of that cell's
// column/row.  The gather, the deposition and g4h_beam_cell_index share it.
// FP64 with EXACT (the cell index) evaluates the reference's (x + gm)/dg;
// the gather and the deposition take (x + gm) * (1/dg), within an ulp of it
// (a particle that this moves across a cell edge has weight 0 or 1 on the
// edge, which both cells agree on).  FP32 uses gm/dg = (ng-1)/2
// exactly and the correctly rounded 1/dg: the cell coordinate is one FMA and
// the in-cell fraction a second FMA against the exact half-integer
// (ng-1)/2 - ix, so the weights keep FP32 precision even where the
// coordinate itself is ~ng/2 (a direct FP32 (x+gm)/dg would leave only
// ~2^-24 * ng/2 of a cell, 3e-5 at ng = 1024).
template <class P, bool EXACT = false>
__device__ __forceinline__ bool cell_of(P x, P y, P gm, P dg, P idg, int ng, int& ix, int& iy,
                                        P& wx, P& wy) {
  if (!(x > -gm && x < gm && y > -gm && y < gm)) return false;  // NaN: off the grid
  if constexpr (sizeof(P) == 8) {
    const P wxr = EXACT ? (x + gm) / dg : (x + gm) * idg;
    const P wyr = EXACT ? (y + gm) / dg : (y + gm) * idg;
    ix = static_cast<int>(floor(wxr));
    iy = static_cast<int>(floor(wyr));
    if (!EXACT) {
      // |x| < gm puts the exact coordinate below ng - 1; the product may
      // round up onto it
      ix = min(ix, ng - 2);
      iy = min(iy, ng - 2);
    }
    if (ix < 0 || ix >= ng - 1 || iy < 0 || iy >= ng - 1) return false;
    wx = P(1) + static_cast<P>(ix) - wxr;
    wy = P(1) + static_cast<P>(iy) - wyr;
  } else {
    (void)dg;
    const float half = 0.5f * static_cast<float>(ng - 1);
    ix = static_cast<int>(floorf(fmaf(x, idg, half)));
    iy = static_cast<int>(floorf(fmaf(y, idg, half)));
    if (ix < 0 || ix >= ng - 1 || iy < 0 || iy >= ng - 1) return false;
    wx = 1.0f - fmaf(x, idg, half - static_cast<float>(ix));
    wy = 1.0f - fmaf(y, idg, half - static_cast<float>(iy));
  }
  return true;
}

// ---- gather (BeamSolverCUDA.cpp:87-143) ---------------------------------
// Split in two so that the caller can issue other loads between the cell
// loads and their use: gather_load reads the four cells around the particle
// (zeros off the grid), gather_sum weights them.
template <class P>
struct Gathered {
  typename V2<P>::T a, b, c, d;
  P wx, wy;
};

template <class P>
__device__ __forceinline__ void gather_load(const GatherField& f, int slice, P x, P y,
                                            Gathered<P>& q) {
  using T2 = typename V2<P>::T;
  const int ng = f.ngrid;
  int ix, iy;
  q.a.x = q.a.y = q.b.x = q.b.y = q.c.x = q.c.y = q.d.x = q.d.y = P(0);
  q.wx = q.wy = P(0);
  if (!cell_of<P>(x, y, P(f.gridmax), P(f.dgrid), P(f.inv_dgrid), ng, ix, iy, q.wx, q.wy)) return;
  const int islice = (slice + f.first) % f.nslice;
  const T2* u = static_cast<const T2*>(f.data) + static_cast<int64_t>(islice) * ng * ng +
                ix + static_cast<int64_t>(iy) * ng;
  q.a = ld_global(&u[0]);
  q.b = ld_global(&u[1]);
  q.c = ld_global(&u[ng]);
  q.d = ld_global(&u[ng + 1]);
}

template <class P>
__device__ __forceinline__ void gather_sum(const GatherField& f, const Gathered<P>& q, P awloc,
                                           P& rr, P& ri) {
  const P wx = q.wx, wy = q.wy;
  const P w00 = wx * wy;
  const P w10 = (P(1) - wx) * wy;
  const P w01 = wx * (P(1) - wy);
  const P w11 = (P(1) - wx) * (P(1) - wy);
  P cre = q.a.x * w00;
  P cim = q.a.y * w00;
  cre += q.b.x * w10;
  cim += q.b.y * w10;
  cre += q.c.x * w01;
  cim += q.c.y * w01;
  cre += q.d.x * w11;
  cim += q.d.y * w11;
  const P scale = P(f.rtmp) * awloc;
  rr = scale * cre;
  ri = -scale * cim;
}

// ---- FP64 phase near a known one ----------------------------------------
// sin and cos of a0 + d from those of a0: the RK4 stages and the final
// phase of a step differ from the step's initial phase by |d| << 1, so one
// full sincos per step and field serves all of them.  Taylor series of sin
// (to d^13) and cos (to d^14) in Horner form; for |d| <= 1/2 the remainders
// are below 3e-17, i.e. the result is the direct sincos(a0 + d) to a few
// ulp.  When every lane of the wave has |d| <= 1/16 (a phase advance of a
// step well inside the FEL bandwidth) the series stop at d^9 and d^8: the
// first omitted terms, d^11/11! < 2e-21 and d^10/10! < 3e-19, are as far
// below FP64 rounding.  Larger offsets take the full routine.
__device__ __forceinline__ void sincos_near(double d, double a, double s0, double c0, double& s,
                                            double& c) {
  const double d2 = d * d;
  double sd, cd;
  if (__builtin_amdgcn_ballot_w64(!(fabs(d) <= 0.0625)) == 0) {
    double ps = sconst(1.0 / 362880.0);  // 1/9!
    ps = fma(ps, d2, sconst(-1.0 / 5040.0));
    ps = fma(ps, d2, sconst(1.0 / 120.0));
    ps = fma(ps, d2, sconst(-1.0 / 6.0));
    sd = fma(ps * d2, d, d);  // d - d^3/3! + ...
    double pc = sconst(1.0 / 40320.0);  // 1/8!
    pc = fma(pc, d2, sconst(-1.0 / 720.0));
    pc = fma(pc, d2, sconst(1.0 / 24.0));
    pc = fma(pc, d2, -0.5);
    cd = fma(pc, d2, 1.0);  // 1 - d^2/2! + ...
  } else {
    if (fabs(d) > 0.5) {
      sincos(a, &s, &c);
      return;
    }
    double ps = sconst(-1.0 / 6227020800.0);  // -1/13!
    ps = fma(ps, d2, sconst(1.0 / 39916800.0));
    ps = fma(ps, d2, sconst(-1.0 / 362880.0));
    ps = fma(ps, d2, sconst(1.0 / 5040.0));
    ps = fma(ps, d2, sconst(-1.0 / 120.0));
    ps = fma(ps, d2, sconst(1.0 / 6.0));
    sd = fma(-ps * d2, d, d);  // d - d^3/3! + ...
    double pc = sconst(1.0 / 87178291200.0);  // 1/14!
    pc = fma(pc, d2, sconst(-1.0 / 479001600.0));
    pc = fma(pc, d2, sconst(1.0 / 3628800.0));
    pc = fma(pc, d2, sconst(-1.0 / 40320.0));
    pc = fma(pc, d2, sconst(1.0 / 720.0));
    pc = fma(pc, d2, sconst(-1.0 / 24.0));
    pc = fma(pc, d2, 0.5);
    cd = fma(-pc, d2, 1.0);  // 1 - d^2/2! + ...
  }
  s = fma(s0, cd, c0 * sd);
  c = fma(c0, cd, -s0 * sd);
}

// ---- ODE ----------------------------------------------------------------
// FP64: evaluate_longitudinal_ode_* (BeamSolverCUDA.cpp:145-223); the phase
// factors exp(-i h theta) come from the step's initial ones (bs, bc at
// theta0 = tthet - dth) through sincos_near, and 1/beta_par through
// inv_btpar, rounded as the reference rounds it.
template <bool ACC, int NF, bool FIRST = false>
__device__ __forceinline__ void ode64(const PushParams& p, int nf, double tgam, double tthet,
                                      double dth, const double* bs, const double* bc, double btpar,
                                      double ez, const double* rr, const double* ri,
                                      double& k2gg, double& k2pp) {
  const double ztemp1 = -2.0 / p.xks;
  // the field sums start from the first field's term, not from 0 + term
  // (an FP64 add the compiler may not drop: 0 + -0 is +0)
  double cre, cim;
#pragma unroll
  for (int i = 0; i < (NF > 0 ? NF : kMaxFields); ++i) {
    if (NF == 0 && i >= nf) break;
    double s, c;
    if (FIRST) {  // the first stage is at the initial phase itself
      s = bs[i];
      c = bc[i];
    } else {
      sincos_near(p.f[i].rharm * dth, p.f[i].rharm * tthet, bs[i], bc[i], s, c);
    }
    if (i == 0) {
      cre = rr[i] * c + ri[i] * s;
      cim = ri[i] * c - rr[i] * s;
    } else {
      cre += rr[i] * c + ri[i] * s;
      cim += ri[i] * c - rr[i] * s;
    }
  }
  const double btper0 = btpar + ztemp1 * cre;
  const double itg = rcp_nr(tgam);
  const double ib = inv_btpar(btper0, tgam * tgam, itg * itg);  // 1/btpar0
  // ACC: add to the stage sums; else they are zero before the call and take
  // the values themselves
  const double dpp = p.xks * (1.0 - ib) + p.xku;
  const double dgg = cim * ib * itg - ez;
  if (ACC) {
    k2pp += dpp;
    k2gg += dgg;
  } else {
    k2pp = dpp;
    k2gg = dgg;
  }
}

// fl(x / 6) for the RK4 weights, correctly rounded as the division is:
// q0 = x * fl(1/3) is within an ulp of x/3, the residual x - 3 q0 is exact in
// an FMA, and q0 + r * fl(1/3) rounds to fl(x / 3) (Markstein's theorem: a
// correctly rounded reciprocal and a quotient within one ulp; checked
// against x / 6.0 on 4.2e8 values on the host); halving is exact.
__device__ __forceinline__ double div6(double x) {
  const double third = 1.0 / 3.0;
  const double q0 = x * third;
  const double r = fma(-q0, 3.0, x);
  return 0.5 * fma(r, third, q0);
}

// MIXED: dgamma (FP32) and theta (FP64) with the expansion of the header.
struct MixedK {
  float ztemp1, c0, k1, k2, inv_b0, inv_g0;
};

// The stage phase is the step's initial phase plus dth (the phase advance
// within the step, |dth| << 1): harmonic h takes sin and cos of its reduced
// initial phase r0[i] plus h * dth in FP32.
template <bool ACC, int NF>
__device__ __forceinline__ void ode32(const PushParams& p, const MixedK& m, int nf, float dg,
                                      const float* r0, float dth, float db0, float ez,
                                      const float* rr, const float* ri, float& k2gg,
                                      float& k2pp) {
  float cre, cim;
#pragma unroll
  for (int i = 0; i < (NF > 0 ? NF : kMaxFields); ++i) {
    if (NF == 0 && i >= nf) break;
    const float a = r0[i] + float(p.f[i].rharm) * dth;
    const float s = __sinf(a), c = __cosf(a);
    if (i == 0) {
      cre = rr[i] * c + ri[i] * s;
      cim = ri[i] * c - rr[i] * s;
    } else {
      cre += rr[i] * c + ri[i] * s;
      cim += ri[i] * c - rr[i] * s;
    }
  }
  const float db = db0 + m.ztemp1 * cre;
  const float d = dg * m.inv_g0;
  const float inv = __builtin_amdgcn_rcpf(1.0f + d);
  const float eta = (db * m.inv_b0 - d * (2.0f + d)) * inv * inv;
  const float dpp = -(m.c0 + eta * (m.k1 + m.k2 * eta));
  const float dgg = cim * m.inv_g0 * inv - ez;
  if (ACC) {
    k2pp += dpp;
    k2gg += dgg;
  } else {
    k2pp = dpp;
    k2gg = dgg;
  }
}

// RK4 two-register scheme (BeamSolverCUDA.cpp:258-299), kept in order.
template <int NF, class PF>
__device__ __forceinline__ void push64(const PushParams& p, int slice, int64_t ip, double x,
                                       double y, double px, double py, double& gamma,
                                       double& theta, double& base_t, double& base_s,
                                       double& base_c, bool on, PF&& pf) {
  const int nf = NF > 0 ? NF : p.nfield;
  // The field cells first, then the caller's loads (the next particle): the
  // memory counter retires in issue order, so waiting for the cells leaves
  // the caller's loads in flight.  pf runs here whether or not the push is
  // on (one call site: the loaded registers need no copies).
  Gathered<double> q[NF > 0 ? NF : kMaxFields];
  double ez = 0.0;
  if (on) {
    // load_longitudinal_ez (BeamSolverCUDA.cpp:183-191), before the cells so
    // that its wait is theirs
    ez = p.particle_ez ? ld_global(&p.particle_ez[ip])
                       : (p.slice_eloss ? ld_global(&p.slice_eloss[slice]) : 0.0);
#pragma unroll
    for (int i = 0; i < (NF > 0 ? NF : kMaxFields); ++i) {
      if (NF == 0 && i >= nf) break;
      gather_load<double>(p.f[i], slice, x, y, q[i]);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  pf();
  __builtin_amdgcn_sched_barrier(0);
  if (!on) return;
  double g = gamma;
  double t = theta + p.autophase;
  const double dx = x - p.ax, dy = y - p.ay;
  const double awloc = 1.0 + 0.5 * (p.kx * dx * dx + p.ky * dy * dy) + p.gradx * dx + p.grady * dy;
  const double btpar = 1.0 + px * px + py * py + p.aw * p.aw * awloc * awloc;
  double rr[NF > 0 ? NF : kMaxFields], ri[NF > 0 ? NF : kMaxFields];
  {
#pragma unroll
    for (int i = 0; i < (NF > 0 ? NF : kMaxFields); ++i) {
      if (NF == 0 && i >= nf) break;
      gather_sum<double>(p.f[i], q[i], awloc, rr[i], ri[i]);
    }
  }
  // phase factors at the step's initial phase, one full sincos per field
  const double t0 = t;
  double bs[NF > 0 ? NF : kMaxFields], bc[NF > 0 ? NF : kMaxFields];
#pragma unroll
  for (int i = 0; i < (NF > 0 ? NF : kMaxFields); ++i) {
    if (NF == 0 && i >= nf) break;
    sincos(p.f[i].rharm * t0, &bs[i], &bc[i]);
  }
  double k2gg, k2pp, k3gg, k3pp;
  ode64<false, NF, true>(p, nf, g, t, 0.0, bs, bc, btpar, ez, rr, ri, k2gg, k2pp);
  double stpz = 0.5 * p.delz;
  g += stpz * k2gg;
  t += stpz * k2pp;
  k3gg = k2gg;
  k3pp = k2pp;
  ode64<false, NF>(p, nf, g, t, t - t0, bs, bc, btpar, ez, rr, ri, k2gg, k2pp);
  g += stpz * (k2gg - k3gg);
  t += stpz * (k2pp - k3pp);
  k3gg = div6(k3gg);
  k3pp = div6(k3pp);
  k2gg *= -0.5;
  k2pp *= -0.5;
  ode64<true, NF>(p, nf, g, t, t - t0, bs, bc, btpar, ez, rr, ri, k2gg, k2pp);
  stpz = p.delz;
  g += stpz * k2gg;
  t += stpz * k2pp;
  k3gg -= k2gg;
  k3pp -= k2pp;
  k2gg *= 2.0;
  k2pp *= 2.0;
  ode64<true, NF>(p, nf, g, t, t - t0, bs, bc, btpar, ez, rr, ri, k2gg, k2pp);
  g += stpz * (k3gg + div6(k2gg));
  t += stpz * (k3pp + div6(k2pp));
  gamma = g;
  theta = t;
  // hand the first field's initial phase factors on to the deposition and
  // the bunching sums (which need the phase after this step)
  base_t = t0;
  base_s = bs[0];
  base_c = bc[0];
}

// MIXED RK4: the same two-register scheme with the phase advance of the
// step accumulated in FP32 relative to the FP64 initial phase, which is
// reduced once per field; theta itself stays FP64 (theta0 + advance).
// Outputs the reduced initial phase of the first field and the FP64 initial
// phase for the sincos of the final phase.
template <int NF, class PF>
__device__ __forceinline__ void push32(const PushParams& p, const MixedK& m, int slice, int64_t ip,
                                       float x, float y, float px, float py, float& dgamma,
                                       double& theta, double& base_t, float& base_r, bool on,
                                       PF&& pf) {
  const int nf = NF > 0 ? NF : p.nfield;
  Gathered<float> q[NF > 0 ? NF : kMaxFields];  // ez and cells, then pf, as in push64
  float ez = 0.0f;
  if (on) {
    ez = p.particle_ez ? float(ld_global(&p.particle_ez[ip]))
                       : (p.slice_eloss ? float(ld_global(&p.slice_eloss[slice])) : 0.0f);
#pragma unroll
    for (int i = 0; i < (NF > 0 ? NF : kMaxFields); ++i) {
      if (NF == 0 && i >= nf) break;
      gather_load<float>(p.f[i], slice, x, y, q[i]);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  pf();
  __builtin_amdgcn_sched_barrier(0);
  if (!on) return;
  float g = dgamma;
  const double t0 = theta + p.autophase;
  const float dx = x - float(p.ax), dy = y - float(p.ay);
  const float awm1 = 0.5f * (float(p.kx) * dx * dx + float(p.ky) * dy * dy) + float(p.gradx) * dx +
                     float(p.grady) * dy;
  const float awloc = 1.0f + awm1;
  const float aw2 = float(p.aw * p.aw);
  const float db0 = px * px + py * py + aw2 * (awm1 * (2.0f + awm1));
  float rr[NF > 0 ? NF : kMaxFields], ri[NF > 0 ? NF : kMaxFields];
  {
#pragma unroll
    for (int i = 0; i < (NF > 0 ? NF : kMaxFields); ++i) {
      if (NF == 0 && i >= nf) break;
      gather_sum<float>(p.f[i], q[i], awloc, rr[i], ri[i]);
    }
  }
  float r0[NF > 0 ? NF : kMaxFields];
#pragma unroll
  for (int i = 0; i < (NF > 0 ? NF : kMaxFields); ++i) {
    if (NF == 0 && i >= nf) break;
    r0[i] = reduce_phase(p.f[i].rharm * t0);
  }
  const float h = float(p.delz);
  float t = 0.0f;  // phase advance within the step
  float k2gg, k2pp, k3gg, k3pp;
  ode32<false, NF>(p, m, nf, g, r0, t, db0, ez, rr, ri, k2gg, k2pp);
  float stpz = 0.5f * h;
  g += stpz * k2gg;
  t += stpz * k2pp;
  k3gg = k2gg;
  k3pp = k2pp;
  ode32<false, NF>(p, m, nf, g, r0, t, db0, ez, rr, ri, k2gg, k2pp);
  g += stpz * (k2gg - k3gg);
  t += stpz * (k2pp - k3pp);
  k3gg *= (1.0f / 6.0f);
  k3pp *= (1.0f / 6.0f);
  k2gg *= -0.5f;
  k2pp *= -0.5f;
  ode32<true, NF>(p, m, nf, g, r0, t, db0, ez, rr, ri, k2gg, k2pp);
  stpz = h;
  g += stpz * k2gg;
  t += stpz * k2pp;
  k3gg -= k2gg;
  k3pp -= k2pp;
  k2gg *= 2.0f;
  k2pp *= 2.0f;
  ode32<true, NF>(p, m, nf, g, r0, t, db0, ez, rr, ri, k2gg, k2pp);
  g += stpz * (k3gg + k2gg * (1.0f / 6.0f));
  t += stpz * (k3pp + k2pp * (1.0f / 6.0f));
  dgamma = g;
  theta = t0 + double(t);
  base_t = t0;
  base_r = r0[0];
}

__device__ __forceinline__ double gshift(double g, double g0) { return g - g0; }
__device__ __forceinline__ double gshift(float dg, double) { return double(dg); }

__device__ __forceinline__ void lds_add(double* a, double v) {
  __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// Source-plane atomics as global (not flat) operations: a pending flat
// operation forces every later wait of the wave to drain all its memory
// counters, the next particle's loads included.
__device__ __forceinline__ void glob_add(double* a, double v) {
  __hip_atomic_fetch_add((G4_GLOBAL double*)(a), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void glob_add(float* a, float v) {
  __hip_atomic_fetch_add((G4_GLOBAL float*)(a), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Per-thread partial sums of the beam diagnostics.  The energy and bunching
// sums are FP64 in both modes.  The ten transverse moments are FP64 in FP64
// mode and, in MIXED mode, FP32 over the thread's own particles (a slice's
// particles / kThreads of them); every reduction across threads, waves and
// slices is FP64.
// quad_perm DPP: lane i takes lane i ^ 1 / i ^ 2 of its quad.  Callers run
// them with whole quads active; bound_ctrl (zero from an inactive source)
// leaves the old value unused, so the compiler needs no copy of v for it.
__device__ __forceinline__ int quad_xor1(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xf, 0xf, true);
}
__device__ __forceinline__ int quad_xor2(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xf, 0xf, true);
}
__device__ __forceinline__ float quad_xor1(float v) {
  return __builtin_bit_cast(float, quad_xor1(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float quad_xor2(float v) {
  return __builtin_bit_cast(float, quad_xor2(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double quad_xor1(double v) {
  int2 a = __builtin_bit_cast(int2, v);
  a.x = quad_xor1(a.x);
  a.y = quad_xor1(a.y);
  return __builtin_bit_cast(double, a);
}
__device__ __forceinline__ double quad_xor2(double v) {
  int2 a = __builtin_bit_cast(int2, v);
  a.x = quad_xor2(a.x);
  a.y = quad_xor2(a.y);
  return __builtin_bit_cast(double, a);
}

// One particle's source terms into the 2x2 cells from (ix, iy): LDS window
// cells with ds_add_f64, cells outside the window with global atomics (and
// the box of such cells for the next footprint).
template <class P>
__device__ __forceinline__ void deposit_one(P cre, P cim, const P* w, int ix, int iy, int wx0,
                                            int wy0, int ww, int wh, int ng,
                                            double2* win, typename V2<P>::T* src, int* s_fb) {
  const int lx0 = ix - wx0, ly0 = iy - wy0;
  if (lx0 >= 0 && lx0 + 1 < ww && ly0 >= 0 && ly0 + 1 < wh) {
    // the whole 2x2 block lies in the window (the common case)
    double* a = &win[lx0 + ly0 * ww].x;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double* ak = a + 2 * ((k & 1) + (k >> 1) * ww);
      lds_add(ak, static_cast<double>(cre * w[k]));
      lds_add(ak + 1, static_cast<double>(cim * w[k]));
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int cx = ix + (k & 1), cy = iy + (k >> 1);
      const int lx = cx - wx0, ly = cy - wy0;
      if (lx >= 0 && lx < ww && ly >= 0 && ly < wh) {
        double* a = &win[lx + ly * ww].x;
        lds_add(a, static_cast<double>(cre * w[k]));
        lds_add(a + 1, static_cast<double>(cim * w[k]));
      } else {
        P* a = reinterpret_cast<P*>(&src[cx + static_cast<int64_t>(cy) * ng]);
        glob_add(a, cre * w[k]);
        glob_add(a + 1, cim * w[k]);
        atomicMin(&s_fb[0], cx);
        atomicMin(&s_fb[1], cy);
        atomicMax(&s_fb[2], cx);
        atomicMax(&s_fb[3], cy);
      }
    }
  }
}

template <class P, int NH>
struct PassState {
  double e[NH > 0 ? 2 + 2 * NH : 1];  // dgamma, dgamma^2, then (cos, sin) per harmonic
  P m[NH > 0 ? 10 : 1];               // x, x^2, y, y^2, px, px^2, py, py^2, x px, y py
  P mm[NH > 0 ? 10 : 1];              // xmin, xmax, pxmin, pxmax, ymin, ymax, pymin, pymax, gmin, gmax
};

// Scratch of a slice split over K workgroups (gridDim.y = K > 1): each
// part's deposition window, its fallback-cell box and its diagnostic sums,
// combined by a following launch.
constexpr int kSumMax = 12 + 2 * kMaxHarm;
constexpr int kRedStride = kSumMax + kNMinMax;
struct SplitScratch {
  double2* win;  // [nslice][K][window_cap]
  double* red;   // [nslice][K][kRedStride]
  int4* fb;      // [nslice][K]
  int own;       // nonzero: the arrays are this slice's own ([K][..])
  // nonzero: each part adds its window to the (cleared) source plane with
  // atomics and folds its boxes into fb[0] (non-zero cells) and fb[1] (the
  // written region), both empty between launches (no K-window combine; win
  // is unused)
  int atomic_win;
};

// The deposition window of a slice: the cells it touched last step (bb)
// plus a 2-cell margin, clipped to the grid and to cap cells (centred);
// empty without a footprint or without current.
__device__ __forceinline__ void place_window(int4 bb, bool on, int ng, int cap, int& wx0,
                                             int& wy0, int& ww, int& wh) {
  wx0 = wy0 = ww = wh = 0;
  if (!(bb.x <= bb.z && on)) return;
  int x0 = max(0, bb.x - 2), y0 = max(0, bb.y - 2);
  const int x1 = min(ng, bb.z + 3), y1 = min(ng, bb.w + 3);
  int w = x1 - x0, h = y1 - y0;
  if (w * h > cap) {
    int side = 1;
    while ((side + 1) * (side + 1) <= cap) ++side;
    const int nw = min(w, max(side, cap / h));
    const int nh = min(h, cap / nw);
    x0 += (w - nw) / 2;
    y0 += (h - nh) / 2;
    w = nw;
    h = nh;
  }
  wx0 = x0;
  wy0 = y0;
  ww = w;
  wh = h;
}

// A slice's source rectangle (the written window and the cells reached
// through the global fallback) and footprint (its non-zero cells, bb, and
// those fallback cells, fb) for the field step and the next step's window.
__device__ __forceinline__ void publish_boxes(const ParticleParams& prm, int s, int wx0, int wy0,
                                              int ww, int wh, const int* bb, const int* fb) {
  int4 r = ww > 0 && wh > 0 ? make_int4(wx0, wy0, wx0 + ww, wy0 + wh)
                            : make_int4(0x7fffffff, 0x7fffffff, -1, -1);
  int4 t = make_int4(bb[0], bb[1], bb[2], bb[3]);
  if (fb[0] <= fb[2]) {
    r = make_int4(min(r.x, fb[0]), min(r.y, fb[1]), max(r.z, fb[2] + 1), max(r.w, fb[3] + 1));
    t = make_int4(min(t.x, fb[0]), min(t.y, fb[1]), max(t.z, fb[2]), max(t.w, fb[3]));
  }
  if (r.x > r.z) r = make_int4(0, 0, 0, 0);
  prm.dep.rect[s] = r;
  prm.dep.bbox[s] = t;
}

// pf issues the loads of the caller's next particle, inside the push right
// after the field cells' loads (also when the push is off).
template <class P, int NF, bool DEP, int NH, class PF>
__device__ __forceinline__ void process(bool diag_on, const ParticleParams& prm, const MixedK& mk, int s,
                                        int64_t i, BeamPtr<P> b, P x, P y, P px, P py, P g,
                                        double th, P scl, typename V2<P>::T* src, double2* win,
                                        int wx0, int wy0, int ww, int wh, int* s_fb,
                                        PassState<P, NH>& ps, PF&& pf) {
  const int st = prm.stages;
  const double g0 = prm.gamma0;
  const int ng = prm.dep.ngrid;
  if (st & ST_T1) transverse<P>(prm.t1, g0, g, x, y, px, py, true);
  // FP64: sine and cosine of a phase near which the final phase lies
  // (base_t), known after the push
  double base_t = 0.0, base_s = 0.0, base_c = 1.0;
  float base_r = 0.0f;  // MIXED: reduced base_t
  bool have_base = false;
  {
    const bool on = (st & ST_PUSH) != 0;  // pf runs inside either way
    if constexpr (sizeof(P) == 8) {
      double gg = g;
      push64<NF>(prm.push, s, i, x, y, px, py, gg, th, base_t, base_s, base_c, on, pf);
      g = gg;
    } else {
      float gg = g;
      push32<NF>(prm.push, mk, s, i, x, y, px, py, gg, th, base_t, base_r, on, pf);
      g = gg;
    }
    have_base = on && prm.push.f[0].rharm == 1.0;
  }
  if (st & ST_R56) th += prm.r56 * gshift(g, g0);
  if (st & ST_T2) transverse<P>(prm.t2, g0, g, x, y, px, py, false);
  if (st & (ST_T1 | ST_T2)) {
    st_stream(&b.x[i], x);
    st_stream(&b.y[i], y);
    st_stream(&b.px[i], px);
    st_stream(&b.py[i], py);
  }
  if (st & ST_PUSH) st_stream(&b.g[i], g);
  if (st & (ST_PUSH | ST_R56)) st_stream(&b.th[i], th);

  // sine and cosine of the final phase, shared by the deposition (harm 1)
  // and the h = 1 bunching sum
  P sn1 = P(0), cs1 = P(1);
  bool have1 = false;
  auto phase1 = [&]() {
    if (!have1) {
      if constexpr (sizeof(P) == 8) {
        double a, c;
        if (have_base)
          sincos_near(th - base_t, th, base_s, base_c, a, c);
        else
          sincos(th, &a, &c);
        sn1 = a;
        cs1 = c;
      } else if (have_base) {
        const float a = base_r + static_cast<float>(th - base_t);
        sn1 = __sinf(a);
        cs1 = __cosf(a);
      } else {
        fsincos(th, sn1, cs1);
      }
      have1 = true;
    }
  };

  if constexpr (DEP) {
    if (scl != P(0)) {
      // Every lane computes its contribution (zero weight when the
      // particle is off the grid), so that the four lanes of a quad (the
      // nbins = 4 mirrored particles of a beamlet, which share their
      // transverse position unless the beam has been disturbed) can combine
      // theirs when they fall on the same 2x2 block of the window: one lane
      // then adds the quad's sums, a quarter of the LDS atomics and none of
      // the same-address conflicts of four lanes adding to one cell.
      const DepParams& d = prm.dep;
      int ix = 0, iy = 0;
      P wx = 0, wy = 0;
      const bool in = cell_of<P>(x, y, P(d.gridmax), P(d.dgrid), P(d.inv_dgrid), ng, ix, iy, wx, wy);
      const P dx = x - P(d.ax), dy = y - P(d.ay);
      const P faw2 = P(1) + P(d.kx) * dx * dx + P(d.ky) * dy * dy +
                     P(2) * (P(d.gradx) * dx + P(d.grady) * dy);
      P sn, cs;
      if (d.harm == 1) {
        phase1();
        sn = sn1;
        cs = cs1;
      } else if constexpr (sizeof(P) == 8) {
        sincos(static_cast<double>(d.harm) * th, &sn, &cs);
      } else {
        fsincos(static_cast<double>(d.harm) * th, sn, cs);
      }
      P part = P(0);
      if (in) {
        if constexpr (sizeof(P) == 8)
          part = faw2 * rsq_nr(faw2) * scl * rcp_nr(g);
        else
          part = psqrt(faw2) * scl * __builtin_amdgcn_rcpf(float(g0) + g);
      }
      const P cre = sn * part;
      const P cim = cs * part;
      const P w[4] = {wx * wy, (P(1) - wx) * wy, wx * (P(1) - wy), (P(1) - wx) * (P(1) - wy)};
      const int lx0 = ix - wx0, ly0 = iy - wy0;
      const bool fast = in && lx0 >= 0 && lx0 + 1 < ww && ly0 >= 0 && ly0 + 1 < wh;
      bool done = false;
      if (__builtin_amdgcn_read_exec() == ~0ull) {  // whole wave here: quads are complete
        const int key = fast ? lx0 + ly0 * ww : -1 - static_cast<int>(threadIdx.x & 63);
        int same = (key == quad_xor1(key) ? 1 : 0) & (key == quad_xor2(key) ? 1 : 0);
        same &= quad_xor1(same) & quad_xor2(same);
        if (same) {
          P v[8];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            v[2 * k] = cre * w[k];
            v[2 * k + 1] = cim * w[k];
          }
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            v[k] += quad_xor1(v[k]);
            v[k] += quad_xor2(v[k]);
          }
          if ((threadIdx.x & 3) == 0) {
            double* a = &win[lx0 + ly0 * ww].x;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              double* ak = a + 2 * ((k & 1) + (k >> 1) * ww);
              lds_add(ak, static_cast<double>(v[2 * k]));
              lds_add(ak + 1, static_cast<double>(v[2 * k + 1]));
            }
          }
          done = true;
        }
      }
      if (!done && in) deposit_one<P>(cre, cim, w, ix, iy, wx0, wy0, ww, wh, ng, win, src, s_fb);
    }
  }
  if constexpr (NH > 0) {
    if (!diag_on) return;
    const double dg = gshift(g, g0);
    ps.e[0] += dg;
    ps.e[1] += dg * dg;
    P* m = ps.m;
    m[0] += x;
    m[1] += x * x;
    m[2] += y;
    m[3] += y * y;
    m[4] += px;
    m[5] += px * px;
    m[6] += py;
    m[7] += py * py;
    m[8] += x * px;
    m[9] += y * py;
    // min/max in the storage type (exact); MIXED energy as gamma - gamma_ref
    P* mm = ps.mm;
    mm[0] = fmin(mm[0], x);
    mm[1] = fmax(mm[1], x);
    mm[2] = fmin(mm[2], px);
    mm[3] = fmax(mm[3], px);
    mm[4] = fmin(mm[4], y);
    mm[5] = fmax(mm[5], y);
    mm[6] = fmin(mm[6], py);
    mm[7] = fmax(mm[7], py);
    mm[8] = fmin(mm[8], g);
    mm[9] = fmax(mm[9], g);
    phase1();
    ps.e[2] += cs1;
    ps.e[3] += sn1;
#pragma unroll
    for (int h = 1; h < NH; ++h) {
      if (h < prm.diag.nharm) {
        if constexpr (sizeof(P) == 8) {
          double sn, cs;
          sincos(static_cast<double>(h + 1) * th, &sn, &cs);
          ps.e[2 + 2 * h] += cs;
          ps.e[3 + 2 * h] += sn;
        } else {
          float sn, cs;
          fsincos(static_cast<double>(h + 1) * th, sn, cs);
          ps.e[2 + 2 * h] += cs;
          ps.e[3 + 2 * h] += sn;
        }
      }
    }
  }
}

// Parameters of this workgroup's slice.  With TAB the slice's entry of a
// device table (batches of independent runs whose per-step parameters differ
// from run to run); else one block for all slices: by value in MIXED, and in
// FP64 a device copy read through a pointer, whose fields are then loaded
// where they are used instead of being held (and spilled) in scalar
// registers across the particle loop of the larger FP64 kernel.
template <bool TAB, class P>
__device__ __forceinline__ const ParticleParams& slice_params(const ParticleParams& arg,
                                                             const ParticleParams* tab) {
  if constexpr (TAB)
    return tab[blockIdx.x];
  else if constexpr (sizeof(P) == 8)
    return tab[0];
  else
    return arg;
}

// The particle pass over part kp of K of slice s, by one workgroup of NT
// threads (NT / 64 waves) whose dynamic LDS holds the deposition window win.
//
// With RT the deposition and the beam sums run only when prm.stages asks for
// them (ST_DEP, ST_DIAG; uniform over the workgroup), so that one copy of the
// pass serves every step of a persistent launch; otherwise DEP and NH > 0
// decide at compile time.
template <class P, int NF, bool DEP, int NH, int NT, bool RT = false>
__device__ __forceinline__ void particle_slice(BeamPtr<P> b, const int32_t* off,
                                               const ParticleParams& prm, SplitScratch sp, int s,
                                               int K, int kp, double2* win) {
  constexpr int NW = NT / 64;
  const bool dep_on = DEP && (!RT || (prm.stages & ST_DEP) != 0);
  const bool diag_on = NH > 0 && (!RT || (prm.stages & ST_DIAG) != 0);
  using T2 = typename V2<P>::T;
  constexpr int NSUM = NH > 0 ? 12 + 2 * NH : 1;  // sums other than the count
  __shared__ int s_bb[4];  // touched-cell bbox (min x, min y, max x, max y), inclusive
  __shared__ int s_fb[4];  // same, for cells reached through the global fallback
  __shared__ double s_red[NW][NSUM + kNMinMax];

  const int tid = threadIdx.x;
  // With K > 1 the slice's particles are split into K contiguous parts, one
  // per workgroup; the last part to finish combines them.
  const int64_t i0s = off[s], i1s = off[s + 1];
  int64_t i0 = i0s, i1 = i1s;
  if (K > 1) {
    const int64_t chunk = (i1s - i0s + K - 1) / K;
    i0 = min(i1s, i0s + kp * chunk);
    i1 = min(i1s, i0 + chunk);
  }
  const double g0 = prm.gamma0;

  // ---- deposition set-up -------------------------------------------------
  // The LDS window covers the cells the slice touched last step (dep.bbox)
  // plus a 2-cell margin, clipped to the grid and to window_cap cells.
  const int ng = prm.dep.ngrid;
  int wx0 = 0, wy0 = 0, ww = 0, wh = 0;
  P scl = P(0);
  T2* src = nullptr;
  if constexpr (DEP) if (dep_on) {
    src = static_cast<T2*>(prm.dep.src) + static_cast<int64_t>(s) * ng * ng;
    scl = static_cast<P>(prm.dep.cur_over_npar[s] * prm.dep.scl_factor);
    int4 bb = prm.dep.bbox[s];
    if (K == 1 && bb.x > bb.z && scl != P(0)) {
      // No footprint from a previous step (first deposition of a run): take
      // it from the current positions, which move by far less than the
      // window margin within one step.  A split slice (K > 1) has it taken
      // by split_prepass_kernel before any part moves a particle, so that
      // all parts place the same window.
      if (tid == 0) {
        s_bb[0] = s_bb[1] = 0x7fffffff;
        s_bb[2] = s_bb[3] = -1;
      }
      __syncthreads();
      int mnx = 0x7fffffff, mny = 0x7fffffff, mxx = -1, mxy = -1;
      for (int64_t k = i0s + tid; k < i1s; k += NT) {
        int ix, iy;
        P wx, wy;
        if (cell_of<P>(b.x[k], b.y[k], P(prm.dep.gridmax), P(prm.dep.dgrid),
                       P(prm.dep.inv_dgrid), ng, ix, iy, wx, wy)) {
          mnx = min(mnx, ix);
          mny = min(mny, iy);
          mxx = max(mxx, ix + 1);
          mxy = max(mxy, iy + 1);
        }
      }
      atomicMin(&s_bb[0], mnx);
      atomicMin(&s_bb[1], mny);
      atomicMax(&s_bb[2], mxx);
      atomicMax(&s_bb[3], mxy);
      __syncthreads();
      bb = make_int4(s_bb[0], s_bb[1], s_bb[2], s_bb[3]);
      __syncthreads();
    }
    place_window(bb, scl != P(0), ng, prm.dep.window_cap, wx0, wy0, ww, wh);
    // zero what the previous deposition left in this slice's source plane
    // (a split slice has that done by a pass before this one)
    const int4 old = prm.dep.rect[s];
    const int ow = old.z - old.x, oh = old.w - old.y;
    if (K == 1 && ow > 0 && oh > 0) {
      for (int k = tid; k < ow * oh; k += NT) {
        const int cy = old.y + k / ow, cx = old.x + k % ow;
        T2 z;
        z.x = 0;
        z.y = 0;
        src[cx + static_cast<int64_t>(cy) * ng] = z;
      }
    }
    for (int k = tid; k < ww * wh; k += NT) win[k] = make_double2(0.0, 0.0);
    if (tid == 0) {
      s_bb[0] = s_bb[1] = 0x7fffffff;
      s_bb[2] = s_bb[3] = -1;
      s_fb[0] = s_fb[1] = 0x7fffffff;
      s_fb[2] = s_fb[3] = -1;
    }
    __syncthreads();
  }

  PassState<P, NH> ps;
  if constexpr (NH > 0) if (diag_on) {
#pragma unroll
    for (int k = 0; k < 2 + 2 * NH; ++k) ps.e[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 10; ++k) ps.m[k] = P(0);
    const P big = P(1.0e5);
    ps.mm[0] = big;  ps.mm[1] = -big;  // x
    ps.mm[2] = big;  ps.mm[3] = -big;  // px
    ps.mm[4] = big;  ps.mm[5] = -big;  // y
    ps.mm[6] = big;  ps.mm[7] = -big;  // py
    // gamma: 1e7 and 1.0 (DiagnosticCUDA.cpp:222), shifted in MIXED
    ps.mm[8] = sizeof(P) == 8 ? P(1.0e7) : P(1.0e7 - g0);
    ps.mm[9] = sizeof(P) == 8 ? P(1.0) : P(1.0 - g0);
  }

  MixedK mk;
  if constexpr (sizeof(P) == 4) {
    mk.ztemp1 = float(-2.0 / prm.push.xks);
    mk.c0 = float(prm.push.c0);
    mk.k1 = float(prm.push.k1);
    mk.k2 = float(prm.push.k2);
    mk.inv_b0 = float(prm.push.inv_b0);
    mk.inv_g0 = float(1.0 / g0);
  }

  // Software-pipelined particle loop: the next particle's state is loaded
  // while the current one is processed (issued after the current particle's
  // field cells, whose wait would otherwise cover these loads too: memory
  // counters retire in issue order).
  int64_t i = i0 + tid;
  bool have = i < i1;
  P nx = P(0), ny = P(0), npx = P(0), npy = P(0), ngm = P(0);
  double nth = 0.0;
  if (have) {
    nx = ld_stream(&b.x[i]);
    ny = ld_stream(&b.y[i]);
    npx = ld_stream(&b.px[i]);
    npy = ld_stream(&b.py[i]);
    ngm = ld_stream(&b.g[i]);
    nth = ld_stream(&b.th[i]);
  }
  while (have) {
    const P x = nx, y = ny, px = npx, py = npy, g = ngm;
    const double th = nth;
    const int64_t inext = i + NT;
    const bool hnext = inext < i1;
    // unconditional (the last particle reloads itself): a load under a
    // branch would make the counter wait for the cells conservative
    const int64_t ipf = hnext ? inext : i;
    auto prefetch = [&]() {
      nx = ld_stream(&b.x[ipf]);
      ny = ld_stream(&b.y[ipf]);
      npx = ld_stream(&b.px[ipf]);
      npy = ld_stream(&b.py[ipf]);
      ngm = ld_stream(&b.g[ipf]);
      nth = ld_stream(&b.th[ipf]);
    };
    process<P, NF, DEP, NH>(diag_on, prm, mk, s, i, b, x, y, px, py, g, th, scl, src, win, wx0, wy0, ww,
                            wh, s_fb, ps, prefetch);
    i = inext;
    have = hnext;
  }

  const int nharm = NH > 0 ? prm.diag.nharm : 0;
  const int nsum = 12 + 2 * nharm;
  if constexpr (NH > 0) if (diag_on) {
    // per-wave sums and extremes
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int k = 0; k < NSUM; ++k) {
      if (k < nsum) {
        const double a = k < 2 ? ps.e[k] : (k < 12 ? static_cast<double>(ps.m[k - 2]) : ps.e[k - 10]);
        const double v = wave_sum(a);
        if (lane == 0) s_red[wave][k] = v;
      }
    }
#pragma unroll
    for (int k = 0; k < kNMinMax; ++k) {
      const double m = static_cast<double>(ps.mm[k]);
      const double v = (k & 1) ? wave_max(m) : wave_min(m);
      if (lane == 0) s_red[wave][NSUM + k] = v;
    }
  }
  if (K > 1) {
    // this part's window (or its atomic adds), fallback box and sums to
    // scratch; a following launch combines the K parts (split_combine_kernel,
    // batch_publish_kernel), so that no part waits on another: an arrival
    // count would need device-scope release and acquire fences, which write
    // back and invalidate the L2 of every part's CU
    __syncthreads();
    const int64_t ss = sp.own ? 0 : s;  // the slice's slot in the scratch arrays
    const int64_t part = ss * K + kp;
    if constexpr (DEP) if (dep_on) {
      if (sp.atomic_win) {
        int mnx = 0x7fffffff, mny = 0x7fffffff, mxx = -1, mxy = -1;
        for (int k = tid; k < ww * wh; k += NT) {
          const double2 v = win[k];
          if (v.x != 0.0 || v.y != 0.0) {
            const int cy = wy0 + k / ww, cx = wx0 + k % ww;
            P* a = reinterpret_cast<P*>(&src[cx + static_cast<int64_t>(cy) * ng]);
            glob_add(a, static_cast<P>(v.x));
            glob_add(a + 1, static_cast<P>(v.y));
            mnx = min(mnx, cx);
            mny = min(mny, cy);
            mxx = max(mxx, cx);
            mxy = max(mxy, cy);
          }
        }
        atomicMin(&s_bb[0], mnx);
        atomicMin(&s_bb[1], mny);
        atomicMax(&s_bb[2], mxx);
        atomicMax(&s_bb[3], mxy);
        __syncthreads();
        if (tid == 0) {
          int4 t = make_int4(s_bb[0], s_bb[1], s_bb[2], s_bb[3]);
          int4 r = ww > 0 && wh > 0 ? make_int4(wx0, wy0, wx0 + ww, wy0 + wh)
                                    : make_int4(0x7fffffff, 0x7fffffff, -1, -1);
          if (s_fb[0] <= s_fb[2]) {
            r = make_int4(min(r.x, s_fb[0]), min(r.y, s_fb[1]), max(r.z, s_fb[2] + 1),
                          max(r.w, s_fb[3] + 1));
            t = make_int4(min(t.x, s_fb[0]), min(t.y, s_fb[1]), max(t.z, s_fb[2]),
                          max(t.w, s_fb[3]));
          }
          int4* bt = &sp.fb[ss * K];
          int4* br = &sp.fb[ss * K + 1];
          atomicMin(&bt->x, t.x);
          atomicMin(&bt->y, t.y);
          atomicMax(&bt->z, t.z);
          atomicMax(&bt->w, t.w);
          atomicMin(&br->x, r.x);
          atomicMin(&br->y, r.y);
          atomicMax(&br->z, r.z);
          atomicMax(&br->w, r.w);
        }
      } else {
        double2* w = sp.win + part * prm.dep.window_cap;
        for (int k = tid; k < ww * wh; k += NT) w[k] = win[k];
        if (tid == 0) sp.fb[part] = make_int4(s_fb[0], s_fb[1], s_fb[2], s_fb[3]);
      }
    }
    if constexpr (NH > 0) if (diag_on) {
      double* r = sp.red + part * kRedStride;
      if (tid < nsum) {
        double v = 0.0;
        for (int w = 0; w < NW; ++w) v += s_red[w][tid];
        r[tid] = v;
      } else if (tid >= 128 && tid < 128 + kNMinMax) {
        const int k = tid - 128;
        double v = s_red[0][NSUM + k];
        for (int w = 1; w < NW; ++w)
          v = (k & 1) ? fmax(v, s_red[w][NSUM + k]) : fmin(v, s_red[w][NSUM + k]);
        r[kSumMax + k] = v;
      }
    }
    return;
  }

  if constexpr (DEP) if (dep_on) {
    __syncthreads();
    // write the window with plain stores; its non-zero cells give the
    // footprint for the next step's window
    int mnx = 0x7fffffff, mny = 0x7fffffff, mxx = -1, mxy = -1;
    for (int k = tid; k < ww * wh; k += NT) {
      const int cy = wy0 + k / ww, cx = wx0 + k % ww;
      const double2 v = win[k];
      T2 o;
      o.x = static_cast<P>(v.x);
      o.y = static_cast<P>(v.y);
      src[cx + static_cast<int64_t>(cy) * ng] = o;
      if (v.x != 0.0 || v.y != 0.0) {
        mnx = min(mnx, cx);
        mny = min(mny, cy);
        mxx = max(mxx, cx);
        mxy = max(mxy, cy);
      }
    }
    atomicMin(&s_bb[0], mnx);
    atomicMin(&s_bb[1], mny);
    atomicMax(&s_bb[2], mxx);
    atomicMax(&s_bb[3], mxy);
    __syncthreads();
    if (tid == 0) publish_boxes(prm, s, wx0, wy0, ww, wh, s_bb, s_fb);
  }

  if constexpr (NH > 0) if (diag_on) {
    __syncthreads();
    double* out = prm.diag.raw + static_cast<int64_t>(s) * G4H_BEAM_RAW(nharm);
    // raw layout: 12 sums (G4H_B_G1..G4H_B_YPY), the count, 10 min/max, then
    // (cos, sin) sums per harmonic
    if (tid < nsum) {
      double v = 0.0;
      for (int w = 0; w < NW; ++w) v += s_red[w][tid];
      out[tid < 12 ? tid : G4H_BEAM_STATS + (tid - 12)] = v;
    } else if (tid == 64) {
      out[G4H_B_COUNT] = static_cast<double>(i1s - i0s);
    } else if (tid >= 128 && tid < 128 + kNMinMax) {
      const int k = tid - 128;
      double v = s_red[0][NSUM + k];
      for (int w = 1; w < NW; ++w)
        v = (k & 1) ? fmax(v, s_red[w][NSUM + k]) : fmin(v, s_red[w][NSUM + k]);
      if (sizeof(P) == 4 && k >= 8) v += g0;  // MIXED energy extremes were shifted
      // order: xmin,xmax,pxmin,pxmax,ymin,ymax,pymin,pymax,gmin,gmax
      out[G4H_B_XMIN + k] = v;
    }
  }
}

template <class P, int NF, bool DEP, int NH, bool TAB>
__global__ __launch_bounds__(kThreads) void particle_kernel(BeamPtr<P> b, const int32_t* off,
                                                            ParticleParams prm_arg,
                                                            const ParticleParams* __restrict__ tab,
                                                            SplitScratch sp) {
  extern __shared__ double2 win[];
  particle_slice<P, NF, DEP, NH, kThreads>(b, off, slice_params<TAB, P>(prm_arg, tab), sp,
                                           blockIdx.x, gridDim.y, blockIdx.y, win);
}

// The combine of a split particle pass (K parts per slice; one workgroup per
// slice, after the pass): the slice's window as the sum of the K parts'
// windows in part order, written to the source plane, its rectangle and
// footprint, and the beam sums of the K parts in part order.  Parts that
// added their windows with atomics (the batch's per-step launch) publish
// through batch_publish_kernel instead.
template <class P, bool TAB>
__global__ __launch_bounds__(kThreads) void split_combine_kernel(
    const int32_t* off, ParticleParams prm_arg, const ParticleParams* __restrict__ tab,
    SplitScratch sp, int K) {
  using T2 = typename V2<P>::T;
  const ParticleParams& prm = slice_params<TAB, P>(prm_arg, tab);
  const int s = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t ss = sp.own ? 0 : s;
  __shared__ int s_bb[4];
  __shared__ int s_fb[4];
  if ((prm.stages & ST_DEP) && !sp.atomic_win) {
    const int ng = prm.dep.ngrid;
    const int cap = prm.dep.window_cap;
    T2* src = static_cast<T2*>(prm.dep.src) + static_cast<int64_t>(s) * ng * ng;
    const P scl = static_cast<P>(prm.dep.cur_over_npar[s] * prm.dep.scl_factor);
    int wx0, wy0, ww, wh;
    place_window(prm.dep.bbox[s], scl != P(0), ng, cap, wx0, wy0, ww, wh);
    if (tid == 0) {
      s_bb[0] = s_bb[1] = 0x7fffffff;
      s_bb[2] = s_bb[3] = -1;
      s_fb[0] = s_fb[1] = 0x7fffffff;
      s_fb[2] = s_fb[3] = -1;
      for (int q = 0; q < K; ++q) {
        const int4 f = sp.fb[ss * K + q];
        s_fb[0] = min(s_fb[0], f.x);
        s_fb[1] = min(s_fb[1], f.y);
        s_fb[2] = max(s_fb[2], f.z);
        s_fb[3] = max(s_fb[3], f.w);
      }
    }
    __syncthreads();
    const double2* parts = sp.win + ss * K * cap;
    int mnx = 0x7fffffff, mny = 0x7fffffff, mxx = -1, mxy = -1;
    for (int k = tid; k < ww * wh; k += kThreads) {
      const int cy = wy0 + k / ww, cx = wx0 + k % ww;
      double2 v = parts[k];
      for (int q = 1; q < K; ++q) {
        const double2 a = parts[static_cast<int64_t>(q) * cap + k];
        v.x += a.x;
        v.y += a.y;
      }
      T2 o;
      o.x = static_cast<P>(v.x);
      o.y = static_cast<P>(v.y);
      src[cx + static_cast<int64_t>(cy) * ng] = o;
      if (v.x != 0.0 || v.y != 0.0) {
        mnx = min(mnx, cx);
        mny = min(mny, cy);
        mxx = max(mxx, cx);
        mxy = max(mxy, cy);
      }
    }
    atomicMin(&s_bb[0], mnx);
    atomicMin(&s_bb[1], mny);
    atomicMax(&s_bb[2], mxx);
    atomicMax(&s_bb[3], mxy);
    __syncthreads();
    if (tid == 0) publish_boxes(prm, s, wx0, wy0, ww, wh, s_bb, s_fb);
  }
  if (!(prm.stages & ST_DIAG)) return;
  const int nharm = prm.diag.nharm;
  const int nsum = 12 + 2 * nharm;
  const double* parts = sp.red + ss * K * kRedStride;
  double* out = prm.diag.raw + static_cast<int64_t>(s) * G4H_BEAM_RAW(nharm);
  if (tid < nsum) {
    double v = 0.0;
    for (int q = 0; q < K; ++q) v += parts[q * kRedStride + tid];
    out[tid < 12 ? tid : G4H_BEAM_STATS + (tid - 12)] = v;
  } else if (tid == 64) {
    out[G4H_B_COUNT] = static_cast<double>(off[s + 1] - off[s]);
  } else if (tid >= 128 && tid < 128 + kNMinMax) {
    const int m = tid - 128;
    double v = parts[kSumMax + m];
    for (int q = 1; q < K; ++q)
      v = (m & 1) ? fmax(v, parts[q * kRedStride + kSumMax + m])
                  : fmin(v, parts[q * kRedStride + kSumMax + m]);
    if (sizeof(P) == 4 && m >= 8) v += prm.gamma0;  // MIXED energy extremes were shifted
    out[G4H_B_XMIN + m] = v;
  }
}

template <class P>
__global__ void cell_index_kernel(BeamPtr<P> b, int64_t n, double gm, double dg, int ng,
                                  int32_t* cell) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  int ix, iy;
  P wx, wy;
  cell[i] = cell_of<P, true>(b.x[i], b.y[i], P(gm), P(dg), P(1.0 / dg), ng, ix, iy, wx, wy)
                ? ix + iy * ng
                : -1;
}

// gamma += dgamma[slice] for every particle of the slice (one workgroup
// per slice; MIXED adds to gamma - gamma_ref).
template <class P>
__global__ void slice_kick_kernel(P* g, const int32_t* off, const double* dgamma) {
  const int s = blockIdx.x;
  const P d = static_cast<P>(dgamma[s]);
  for (int64_t i = off[s] + threadIdx.x; i < off[s + 1]; i += blockDim.x) g[i] += d;
}

// Set-up of a split pass (K workgroups per slice), one workgroup per slice,
// run before particle_kernel on the same stream: zero the cells the previous
// deposition left in the slice's source plane (the source rectangle), and on a
// run's first deposition (no footprint yet) take the footprint of the
// current positions into dep.bbox.  Doing the latter here rather than in each
// part keeps every part's window identical: inside particle_kernel the parts
// of a slice run concurrently and overwrite their positions while a sibling
// would still be scanning them.
template <class P, bool TAB>
__global__ void split_prepass_kernel(BeamPtr<P> b, const int32_t* off, ParticleParams prm_arg,
                                     const ParticleParams* __restrict__ tab) {
  const ParticleParams& prm = slice_params<TAB, P>(prm_arg, tab);
  using T2 = typename V2<P>::T;
  const int s = blockIdx.x;
  const int tid = threadIdx.x;
  const int ng = prm.dep.ngrid;
  T2* src = static_cast<T2*>(prm.dep.src) + static_cast<int64_t>(s) * ng * ng;
  const int4 old = prm.dep.rect[s];
  const int ow = old.z - old.x, oh = old.w - old.y;
  if (ow > 0 && oh > 0) {
    for (int k = tid; k < ow * oh; k += blockDim.x) {
      const int cy = old.y + k / ow, cx = old.x + k % ow;
      T2 z;
      z.x = 0;
      z.y = 0;
      src[cx + static_cast<int64_t>(cy) * ng] = z;
    }
  }
  const int4 bb = prm.dep.bbox[s];
  const P scl = static_cast<P>(prm.dep.cur_over_npar[s] * prm.dep.scl_factor);
  if (!(bb.x > bb.z && scl != P(0))) return;
  __shared__ int s_bb[4];
  if (tid == 0) {
    s_bb[0] = s_bb[1] = 0x7fffffff;
    s_bb[2] = s_bb[3] = -1;
  }
  __syncthreads();
  int mnx = 0x7fffffff, mny = 0x7fffffff, mxx = -1, mxy = -1;
  for (int64_t k = off[s] + tid; k < off[s + 1]; k += blockDim.x) {
    int ix, iy;
    P wx, wy;
    if (cell_of<P>(b.x[k], b.y[k], P(prm.dep.gridmax), P(prm.dep.dgrid), P(prm.dep.inv_dgrid), ng,
                   ix, iy, wx, wy)) {
      mnx = min(mnx, ix);
      mny = min(mny, iy);
      mxx = max(mxx, ix + 1);
      mxy = max(mxy, iy + 1);
    }
  }
  atomicMin(&s_bb[0], mnx);
  atomicMin(&s_bb[1], mny);
  atomicMax(&s_bb[2], mxx);
  atomicMax(&s_bb[3], mxy);
  __syncthreads();
  // stays empty when no particle is on the grid: every part then deposits
  // through the global fallback
  if (tid == 0 && s_bb[0] <= s_bb[2]) prm.dep.bbox[s] = make_int4(s_bb[0], s_bb[1], s_bb[2], s_bb[3]);
}

// Workgroups per slice: a pass over few slices is split so that about two
// workgroups per CU are in flight, keeping at least 8 iterations of the
// particle loop per workgroup (or as pinned on the context).
int split_of(Context* ctx, const Beam* b) {
  const int forced = ctx->tuner.forced[TUNE_SPLIT];
  if (forced >= 1) return std::min(forced, 64);
  if (ctx->num_cus == 0) {
    int n = 0;
    G4_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, ctx->device));
    ctx->num_cus = n > 0 ? n : 1;
  }
  const int target = 2 * ctx->num_cus;
  if (b->nslice >= target) return 1;
  int k = (target + b->nslice - 1) / b->nslice;
  k = std::min(k, b->max_slice_np / (8 * kThreads));
  return std::max(1, std::min(k, 64));
}

template <class P, int NF, bool DEP, int NH, bool TAB>
void launch_t(Context* ctx, Beam* b, const ParticleParams& p, BeamPtr<P> bp,
              const ParticleParams* tab) {
  const size_t lds = DEP ? static_cast<size_t>(p.dep.window_cap) * sizeof(double2) : 0;
  const int K = split_of(ctx, b);
  if (!tab && sizeof(P) == 8)
    tab = static_cast<const ParticleParams*>(ctx->param_push(&p, sizeof(p)));
  SplitScratch sp{nullptr, nullptr, nullptr, 0, 0};
  if (K > 1) {
    const size_t parts = static_cast<size_t>(b->nslice) * K;
    const size_t nwin = DEP ? parts * static_cast<size_t>(p.dep.window_cap) : 0;
    if (b->split_win.size() < nwin) b->split_win.alloc(ctx, nwin);
    if (b->split_red.size() < parts * kRedStride) b->split_red.alloc(ctx, parts * kRedStride);
    if (b->split_fb.size() < parts) b->split_fb.alloc(ctx, parts);
    sp = SplitScratch{b->split_win.get(), b->split_red.get(), b->split_fb.get(), 0, 0};
    if (DEP) {
      ctx->launch(G4H_K_OTHER, [&] {
        hipLaunchKernelGGL((split_prepass_kernel<P, TAB>), dim3(b->nslice), dim3(256), 0,
                           ctx->stream, bp, b->off.get(), p, tab);
      });
    }
  }
  ctx->launch(G4H_K_PARTICLE, [&] {
    hipLaunchKernelGGL((particle_kernel<P, NF, DEP, NH, TAB>), dim3(b->nslice, K),
                       dim3(kThreads), lds, ctx->stream, bp, b->off.get(), p, tab, sp);
  });
  if (K > 1 && (DEP || NH > 0)) {
    ctx->launch(G4H_K_OTHER, [&] {
      hipLaunchKernelGGL((split_combine_kernel<P, TAB>), dim3(b->nslice), dim3(kThreads), 0,
                         ctx->stream, b->off.get(), p, tab, sp, K);
    });
  }
  ctx->param_commit();
}

template <class P, int NF, bool DEP, bool TAB>
void dispatch_nh(Context* ctx, Beam* b, const ParticleParams& p, BeamPtr<P> bp,
                 const ParticleParams* tab) {
  if (!(p.stages & ST_DIAG)) return launch_t<P, NF, DEP, 0, TAB>(ctx, b, p, bp, tab);
  const int nh = p.diag.nharm;
  if (nh == 1) return launch_t<P, NF, DEP, 1, TAB>(ctx, b, p, bp, tab);
  if (nh == 2) return launch_t<P, NF, DEP, 2, TAB>(ctx, b, p, bp, tab);
  if (nh <= 4) return launch_t<P, NF, DEP, 4, TAB>(ctx, b, p, bp, tab);
  return launch_t<P, NF, DEP, 8, TAB>(ctx, b, p, bp, tab);
}

template <class P>
void dispatch(Context* ctx, Beam* b, const ParticleParams& p, BeamPtr<P> bp,
              const ParticleParams* tab) {
  const bool dep = (p.stages & ST_DEP) != 0;
  const bool single = !(p.stages & ST_PUSH) || p.push.nfield == 1;
  if (tab) {
    require(single, "per-slice parameter tables serve single-field passes");
    if (dep) return dispatch_nh<P, 1, true, true>(ctx, b, p, bp, tab);
    return dispatch_nh<P, 1, false, true>(ctx, b, p, bp, tab);
  }
  if (single) {
    if (dep) return dispatch_nh<P, 1, true, false>(ctx, b, p, bp, nullptr);
    return dispatch_nh<P, 1, false, false>(ctx, b, p, bp, nullptr);
  }
  if (dep) return dispatch_nh<P, 0, true, false>(ctx, b, p, bp, nullptr);
  return dispatch_nh<P, 0, false, false>(ctx, b, p, bp, nullptr);
}

// ---- batches of steady-state runs in one launch ----------------------------
// One workgroup of kBatchThreads owns one run (one slice, one ring plane) and
// loops over the steps: the fused particle pass over the run's particles
// (the same particle_slice as the grid launch, unsplit), then the two ADI
// sweeps of its plane with the wave-per-64-lines code of adi_lines.hpp
// (waves take the plane's 64-line blocks in turn; a barrier between the
// forward and the backward pass of each sweep makes the in-place update
// safe), and on output steps the plane's field sums, reduced in block order
// as the grid launch reduces them.  No grid-wide synchronisation: runs are
// independent.  Dynamic LDS: the deposition window, then the factors.
constexpr int kBatchThreads = 512;
constexpr int kBatchWaves = kBatchThreads / 64;
constexpr int kBatchChunk = 4;

template <class P, int NH>
__device__ __forceinline__ void batch_particles(const BatchRunDesc& R, const ParticleParams& prm,
                                                double2* win) {
  BeamPtr<P> b{static_cast<P*>(R.x),  static_cast<P*>(R.y), static_cast<P*>(R.px),
               static_cast<P*>(R.py), static_cast<P*>(R.g), R.th};
  const SplitScratch none{nullptr, nullptr, nullptr, 0, 0};
  // one copy of the pass; deposition and sums as the step's stages ask
  particle_slice<P, 1, true, NH, kBatchThreads, true>(b, R.off, prm, none, R.slice, 1, 0, win);
}

template <class P, int NH>
__global__ __launch_bounds__(kBatchThreads) void batch_persistent_kernel(
    const BatchRunDesc* __restrict__ runs, const int32_t* __restrict__ rec, int k0, int k1,
    int win_cells) {
  using T2 = typename V2<P>::T;
  extern __shared__ double2 lds[];
  __shared__ double s_part[kBatchWaves][8];
  __shared__ double s_centre[2];
  const BatchRunDesc R = runs[blockIdx.x];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int G = R.G;
  const int nblk = (G + adi::kLane - 1) / adi::kLane;
  const int64_t plane = static_cast<int64_t>(G) * G;
  T2* u = static_cast<T2*>(R.u) + R.slice * plane;
  double2* ys = R.ys + R.slice * plane;
  const T2* srcp = static_cast<const T2*>(R.src) + R.slice * plane;
  double2* win = lds;
  double2* coef = lds + win_cells;
  for (int k = k0; k < k1; ++k) {
    const ParticleParams& prm = R.tab[static_cast<int64_t>(k) * R.nrun + R.slice];
    batch_particles<P, NH>(R, prm, win);
    const bool dep = (prm.stages & ST_DEP) != 0;
    const double2* fac = R.adi[k];
    const double cs = R.cs[k];
    const int r = rec[k];
    __syncthreads();  // source plane and rectangle written, window LDS free
    const adi::Factors f = adi::stage_factors(coef, fac, fac + G, fac + 2 * G, G, tid,
                                              kBatchThreads);
    const T2* sp = dep ? srcp : nullptr;
    const int4 rc = dep ? R.rect[R.slice] : make_int4(0, 0, 0, 0);
    __syncthreads();
    for (int blk = wave; blk < nblk; blk += kBatchWaves)
      adi::x_forward<P, kBatchChunk>(u, ys, G, cs, f, sp, rc, blk, lane);
    __syncthreads();  // every block has read its neighbours' rows
    for (int blk = wave; blk < nblk; blk += kBatchWaves)
      adi::x_backward<P, kBatchChunk>(ys, u, G, f, blk, lane);
    __syncthreads();
    for (int blk = wave; blk < nblk; blk += kBatchWaves)
      adi::y_forward<P, kBatchChunk>(u, ys, G, cs, f, sp, rc, blk, lane);
    __syncthreads();
    if (r >= 0) {
      for (int blk = wave; blk < nblk; blk += kBatchWaves) {
        adi::PlaneSums sums;
        sums.clear();
        adi::y_backward<P, kBatchChunk, true>(ys, u, G, f, blk, lane, sums, s_centre);
        // one wave per block here (nblk <= kBatchWaves for G <= 512)
#pragma unroll
        for (int q = 0; q < 7; ++q) {
          const double v = adi::wave_sum64(sums.acc[q]);
          if (lane == 0) s_part[blk][q] = v;
        }
      }
      __syncthreads();
      if (tid < G4H_FIELD_RAW) {
        double v;
        if (tid < 7) {
          v = 0.0;
          for (int q = 0; q < nblk; ++q) v += s_part[q][tid];
        } else {
          v = s_centre[tid - 7];
        }
        R.field_hist[(static_cast<int64_t>(r) * R.nrun + R.slice) * G4H_FIELD_RAW + tid] = v;
      }
    } else {
      adi::PlaneSums none;
      for (int blk = wave; blk < nblk; blk += kBatchWaves)
        adi::y_backward<P, kBatchChunk, false>(ys, u, G, f, blk, lane, none, nullptr);
    }
    __syncthreads();  // the plane is final before the next step gathers from it
  }
}

// One step of every run: workgroup r runs the fused particle pass of run r
// (unsplit, its parameter entry of step k; deposition and sums as the step's
// stages ask).
// Workgroups of kThreads as the grid launch's, and the beam sums compiled
// in only for output steps (NH > 0): a register count that keeps several
// workgroups per CU.
template <class P, int NH>
__global__ __launch_bounds__(kThreads) void batch_step_particle_kernel(
    const BatchRunDesc* __restrict__ runs, const int2* __restrict__ parts,
    const ParticleParams* __restrict__ tabk) {
  extern __shared__ double2 lds[];
  const int2 pr = parts[blockIdx.x];  // (run, part)
  const BatchRunDesc& R = runs[pr.x];
  // through a restrict kernel argument the parameters are scalar loads; read
  // through the run descriptor's table pointer they would be vector loads
  // inside the particle loop (the compiler cannot show that the loop's
  // stores leave that table alone), each waited for with everything else
  const ParticleParams& prm = tabk[pr.x];
  BeamPtr<P> b{static_cast<P*>(R.x),  static_cast<P*>(R.y), static_cast<P*>(R.px),
               static_cast<P*>(R.py), static_cast<P*>(R.g), R.th};
  const SplitScratch sp{R.swin, R.sred, R.sfb, 1, 1};
  particle_slice<P, 1, true, NH, kThreads, true>(b, R.off, prm, sp, R.slice, R.K, pr.y, lds);
}

// After a per-step batch launch whose split runs defer publication: per
// split run (one workgroup each), the slice's source rectangle and footprint
// from the boxes its parts folded with atomics (then emptied for the next
// step) and, with sums, the slice's beam sums from the K parts' partial sums
// in part order: what the last part of a counting launch does.
template <class P>
__global__ __launch_bounds__(256) void batch_publish_kernel(const BatchRunDesc* __restrict__ runs,
                                                            int k, int sums) {
  const BatchRunDesc& R = runs[blockIdx.x];
  const int K = R.K;
  if (K <= 1) return;
  const ParticleParams& prm = R.tab[static_cast<int64_t>(k) * R.nrun + R.slice];
  const int s = R.slice;
  const int tid = threadIdx.x;
  if ((prm.stages & ST_DEP) && tid == 0) {
    const int4 t = R.sfb[0], r = R.sfb[1];
    prm.dep.rect[s] = r.x > r.z ? make_int4(0, 0, 0, 0) : r;
    prm.dep.bbox[s] = t;
    R.sfb[0] = R.sfb[1] = make_int4(0x7fffffff, 0x7fffffff, -1, -1);
  }
  if (!sums || !(prm.stages & ST_DIAG)) return;
  const int nharm = prm.diag.nharm;
  const int nsum = 12 + 2 * nharm;
  const double* parts = R.sred;
  double* out = prm.diag.raw + static_cast<int64_t>(s) * G4H_BEAM_RAW(nharm);
  if (tid < nsum) {
    double v = 0.0;
    for (int q = 0; q < K; ++q) v += parts[q * kRedStride + tid];
    out[tid < 12 ? tid : G4H_BEAM_STATS + (tid - 12)] = v;
  } else if (tid == 64) {
    out[G4H_B_COUNT] = static_cast<double>(R.off[s + 1] - R.off[s]);
  } else if (tid >= 128 && tid < 128 + kNMinMax) {
    const int m = tid - 128;
    double v = parts[kSumMax + m];
    for (int q = 1; q < K; ++q)
      v = (m & 1) ? fmax(v, parts[q * kRedStride + kSumMax + m])
                  : fmin(v, parts[q * kRedStride + kSumMax + m]);
    if (sizeof(P) == 4 && m >= 8) v += prm.gamma0;  // MIXED energy extremes were shifted
    out[G4H_B_XMIN + m] = v;
  }
}

// Beam-sums pass (ST_DIAG only) whose scratch the caller provides: d_prm is
// the FP64 parameter block in device memory (unused in MIXED, which takes
// it by value), sp the split pass's partials and counters for K > 1.
template <class P, int NH>
void launch_diag_given(Context* ctx, Beam* b, const ParticleParams& p, BeamPtr<P> bp,
                       const ParticleParams* d_prm, SplitScratch sp, int K) {
  ctx->launch(G4H_K_PARTICLE, [&] {
    hipLaunchKernelGGL((particle_kernel<P, 1, false, NH, false>), dim3(b->nslice, K),
                       dim3(kThreads), 0, ctx->stream, bp, b->off.get(), p, d_prm, sp);
  });
  if (K > 1) {
    ctx->launch(G4H_K_OTHER, [&] {
      hipLaunchKernelGGL((split_combine_kernel<P, false>), dim3(b->nslice), dim3(kThreads), 0,
                         ctx->stream, b->off.get(), p, d_prm, sp, K);
    });
  }
}

template <class P>
void diag_given(Context* ctx, Beam* b, const ParticleParams& p, BeamPtr<P> bp,
                const ParticleParams* d_prm, SplitScratch sp, int K) {
  const int nh = p.diag.nharm;
  if (nh == 1) return launch_diag_given<P, 1>(ctx, b, p, bp, d_prm, sp, K);
  if (nh == 2) return launch_diag_given<P, 2>(ctx, b, p, bp, d_prm, sp, K);
  if (nh <= 4) return launch_diag_given<P, 4>(ctx, b, p, bp, d_prm, sp, K);
  return launch_diag_given<P, 8>(ctx, b, p, bp, d_prm, sp, K);
}

}  // namespace

void launch_cell_index(Context* ctx, Beam* b, const Field* f, int32_t* d_cell) {
  if (b->np <= 0) return;
  const unsigned blocks = static_cast<unsigned>((b->np + 255) / 256);
  ctx->launch(G4H_K_OTHER, [&] {
    if (b->prec == G4H_FP64) {
      BeamPtr<double> bp{b->xd.get(), b->yd.get(), b->pxd.get(), b->pyd.get(), b->gd.get(),
                         b->theta.get()};
      hipLaunchKernelGGL(cell_index_kernel<double>, dim3(blocks), dim3(256), 0, ctx->stream, bp,
                         b->np, f->gridmax, f->dgrid, f->ngrid, d_cell);
    } else {
      BeamPtr<float> bp{b->xf.get(), b->yf.get(), b->pxf.get(), b->pyf.get(), b->gf.get(),
                        b->theta.get()};
      hipLaunchKernelGGL(cell_index_kernel<float>, dim3(blocks), dim3(256), 0, ctx->stream, bp,
                         b->np, f->gridmax, f->dgrid, f->ngrid, d_cell);
    }
  });
}

void launch_slice_kick(Context* ctx, Beam* b, const double* d_dgamma) {
  if (b->nslice <= 0 || b->np <= 0) return;
  ctx->launch(G4H_K_OTHER, [&] {
    if (b->prec == G4H_FP64)
      hipLaunchKernelGGL(slice_kick_kernel<double>, dim3(b->nslice), dim3(256), 0, ctx->stream,
                         b->gd.get(), b->off.get(), d_dgamma);
    else
      hipLaunchKernelGGL(slice_kick_kernel<float>, dim3(b->nslice), dim3(256), 0, ctx->stream,
                         b->gf.get(), b->off.get(), d_dgamma);
  });
}

void launch_particle(Context* ctx, Beam* b, const ParticleParams& p,
                     const ParticleParams* d_table) {
  if (b->nslice <= 0) return;
  require(p.dep.window_cap * sizeof(double2) <= 96 * 1024 || !(p.stages & ST_DEP),
          "deposition window exceeds the LDS budget");
  require(!(p.stages & ST_DIAG) || (p.diag.nharm >= 1 && p.diag.nharm <= kMaxHarm),
          "nharm must be in [1, 8]");
  if (b->prec == G4H_FP64) {
    BeamPtr<double> bp{b->xd.get(), b->yd.get(), b->pxd.get(), b->pyd.get(), b->gd.get(),
                       b->theta.get()};
    dispatch<double>(ctx, b, p, bp, d_table);
  } else {
    BeamPtr<float> bp{b->xf.get(), b->yf.get(), b->pxf.get(), b->pyf.get(), b->gf.get(),
                      b->theta.get()};
    dispatch<float>(ctx, b, p, bp, d_table);
  }
}

namespace {
template <class F>
void batch_kernel_of(g4h_precision prec, int nharm, F&& use) {
  if (prec == G4H_FP64) {
    if (nharm == 1)
      use(batch_persistent_kernel<double, 1>);
    else
      use(batch_persistent_kernel<double, 8>);
  } else {
    if (nharm == 1)
      use(batch_persistent_kernel<float, 1>);
    else
      use(batch_persistent_kernel<float, 8>);
  }
}

size_t batch_lds_bytes(int win_cells, int gmax) {
  return sizeof(double2) * (static_cast<size_t>(win_cells) + 3 * static_cast<size_t>(gmax));
}
}  // namespace

bool batch_persistent_fits(Context* ctx, g4h_precision prec, int nharm, int win_cells, int gmax) {
  if (gmax > kBatchWaves * adi::kLane || nharm < 1 || nharm > kMaxHarm) return false;
  int limit = 0;
  G4_HIP(hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
  size_t statics = 0;
  batch_kernel_of(prec, nharm, [&](auto kern) {
    hipFuncAttributes attr{};
    G4_HIP(hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(kern)));
    statics = attr.sharedSizeBytes;
  });
  return batch_lds_bytes(win_cells, gmax) + statics <= static_cast<size_t>(limit);
}

void launch_batch_persistent(Context* ctx, g4h_precision prec, int nharm,
                             const BatchRunDesc* d_runs, int nruns, const int32_t* d_rec, int k0,
                             int k1, int win_cells, int gmax) {
  if (nruns <= 0 || k1 <= k0) return;
  require(batch_persistent_fits(ctx, prec, nharm, win_cells, gmax),
          "batch kernel: grid above 512 or LDS budget exceeded");
  const size_t lds = batch_lds_bytes(win_cells, gmax);
  batch_kernel_of(prec, nharm, [&](auto kern) {
    ctx->launch(G4H_K_PARTICLE, [&] {
      hipLaunchKernelGGL(kern, dim3(nruns), dim3(kBatchThreads), lds, ctx->stream, d_runs, d_rec,
                         k0, k1, win_cells);
    });
  });
}

size_t batch_split_red_stride() { return kRedStride; }

void launch_batch_step_particles(Context* ctx, g4h_precision prec, int nharm,
                                 const BatchRunDesc* d_runs, int nruns, const int2* d_parts,
                                 int nparts, const ParticleParams* d_tabk, int k, int win_cells,
                                 bool sums, bool any_split) {
  if (nparts <= 0) return;
  require(nharm >= 1 && nharm <= kMaxHarm, "nharm must be in [1, 8]");
  const size_t lds = sizeof(double2) * static_cast<size_t>(win_cells);
  auto go = [&](auto kern) {
    ctx->launch(G4H_K_PARTICLE, [&] {
      hipLaunchKernelGGL(kern, dim3(nparts), dim3(kThreads), lds, ctx->stream, d_runs, d_parts,
                         d_tabk);
    });
  };
  if (prec == G4H_FP64) {
    if (!sums)
      go(batch_step_particle_kernel<double, 0>);
    else if (nharm == 1)
      go(batch_step_particle_kernel<double, 1>);
    else
      go(batch_step_particle_kernel<double, 8>);
  } else {
    if (!sums)
      go(batch_step_particle_kernel<float, 0>);
    else if (nharm == 1)
      go(batch_step_particle_kernel<float, 1>);
    else
      go(batch_step_particle_kernel<float, 8>);
  }
  if (!any_split) return;
  ctx->launch(G4H_K_OTHER, [&] {
    if (prec == G4H_FP64)
      hipLaunchKernelGGL(batch_publish_kernel<double>, dim3(nruns), dim3(256), 0, ctx->stream,
                         d_runs, k, sums ? 1 : 0);
    else
      hipLaunchKernelGGL(batch_publish_kernel<float>, dim3(nruns), dim3(256), 0, ctx->stream,
                         d_runs, k, sums ? 1 : 0);
  });
}

// ---- beam sums without allocation ----------------------------------------
// Device memory for callers that cannot allocate (code under graph capture,
// tight serving loops): the FP64 parameter block, the per-slice sums before
// their copy to the host, and the split pass's per-part partials and
// per-slice arrival counters all live in this array.  Each GPU the library
// runs on holds its own copy until the process ends.  Host code must use the
// address from hipGetSymbolAddress (scratch_base), never the array's name.
static __device__ __align__(256) double scratch[kNoallocScratchBytes / sizeof(double)];

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_dev;          // per device, from hipGetSymbolAddress
std::vector<std::unique_ptr<std::mutex>> g_scratch_use;  // one caller at a time per device

size_t align256(size_t n) { return (n + 255) & ~static_cast<size_t>(255); }
}  // namespace

void* scratch_base(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev.size() <= static_cast<size_t>(device)) {
    g_scratch_dev.resize(static_cast<size_t>(device) + 1, nullptr);
    g_scratch_use.resize(static_cast<size_t>(device) + 1);
  }
  if (!g_scratch_dev[static_cast<size_t>(device)]) {
    // the caller has made `device` current
    void* ptr = nullptr;
    G4_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[static_cast<size_t>(device)] = ptr;
    g_scratch_use[static_cast<size_t>(device)] = std::make_unique<std::mutex>();
  }
  return g_scratch_dev[static_cast<size_t>(device)];
}

size_t diag_beam_noalloc_bytes(Context* ctx, const Beam* b, int nharm, int* split) {
  const int K = split_of(ctx, b);
  if (split) *split = K;
  size_t n = align256(sizeof(ParticleParams));
  n += align256(static_cast<size_t>(b->nslice) * G4H_BEAM_RAW(nharm) * sizeof(double));
  if (K > 1) {
    const size_t parts = static_cast<size_t>(b->nslice) * K;
    n += align256(parts * kRedStride * sizeof(double));
    n += align256(parts * sizeof(int4));
  }
  return n;
}

void diag_beam_raw_noalloc(Context* ctx, Beam* b, int nharm, double* raw) {
  int K = 1;
  const size_t need = diag_beam_noalloc_bytes(ctx, b, nharm, &K);
  if (need > kNoallocScratchBytes)
    fail(G4H_ERR_INVALID_ARGUMENT,
         "g4h_diag_beam_raw_noalloc: " + std::to_string(b->nslice) + " slices with " +
             std::to_string(nharm) + " harmonics need " + std::to_string(need) +
             " bytes of scratch, more than its fixed " + std::to_string(kNoallocScratchBytes) +
             " (use g4h_diag_beam_raw)");
  char* base = static_cast<char*>(scratch_base(ctx->device));
  std::mutex* use;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    use = g_scratch_use[static_cast<size_t>(ctx->device)].get();
  }
  std::lock_guard<std::mutex> hold(*use);
  size_t at = 0;
  auto* d_prm = reinterpret_cast<ParticleParams*>(base + at);
  at += align256(sizeof(ParticleParams));
  auto* d_raw = reinterpret_cast<double*>(base + at);
  const size_t nraw = static_cast<size_t>(b->nslice) * G4H_BEAM_RAW(nharm);
  at += align256(nraw * sizeof(double));
  SplitScratch sp{nullptr, nullptr, nullptr, 0, 0};
  if (K > 1) {
    const size_t parts = static_cast<size_t>(b->nslice) * K;
    sp.red = reinterpret_cast<double*>(base + at);
    at += align256(parts * kRedStride * sizeof(double));
    sp.fb = reinterpret_cast<int4*>(base + at);
    at += align256(parts * sizeof(int4));
  }
  ParticleParams p{};
  p.gamma0 = b->gamma_ref;
  p.stages = ST_DIAG;
  p.diag.raw = d_raw;
  p.diag.nharm = nharm;
  if (b->prec == G4H_FP64) {
    G4_HIP(hipMemcpyAsync(d_prm, &p, sizeof(p), hipMemcpyHostToDevice, ctx->stream));
    BeamPtr<double> bp{b->xd.get(), b->yd.get(), b->pxd.get(), b->pyd.get(), b->gd.get(),
                       b->theta.get()};
    diag_given<double>(ctx, b, p, bp, d_prm, sp, K);
  } else {
    BeamPtr<float> bp{b->xf.get(), b->yf.get(), b->pxf.get(), b->pyf.get(), b->gf.get(),
                      b->theta.get()};
    diag_given<float>(ctx, b, p, bp, nullptr, sp, K);
  }
  G4_HIP(hipMemcpyAsync(raw, d_raw, nraw * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  G4_HIP(hipStreamSynchronize(ctx->stream));
}

}  // namespace g4
