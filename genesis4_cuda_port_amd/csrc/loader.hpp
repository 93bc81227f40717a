// loader.hpp - the quiet-start beam of one beamlet, shared by the host
// loader (g4h_load_beam, frontend.cpp) and the device generator
// (g4h_beam_generate, loader.hip), so both produce the same particles.
//
// Beamlet j of slice s takes Sobol point 1 + s*nb + j (6 dimensions, Gray-code
// order, digitally shifted by seeded words): five coordinates become
// Gaussian x, px, y, py, gamma through the inverse normal CDF, the sixth
// the phase of the first of nbins mirrored particles.  Shot noise (Fawley,
// PRST-AB 5, 070701, 2002) perturbs the nbins phases by
// sum_h a_h cos(h th) + b_h sin(h th) with a_h, b_h ~ N(0, 2/(h^2 Ne)); the
// normal deviates come from a counter-based generator (splitmix64 of the
// slice seed and the (beamlet, harmonic, component) counter, Box-Muller), so
// every beamlet is independent of every other one and of the thread count.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/g4hip.h"

namespace g4 {
namespace loader {

constexpr double kPi = 3.14159265358979323846;
constexpr int kDims = 6;

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// seed of slice s's shot-noise stream
__host__ __device__ inline uint64_t slice_seed(uint64_t seed, int s) {
  return splitmix64(seed ^ (0x5851F42D4C957F2Dull * static_cast<uint64_t>(s + 1)));
}

// uniform in (0, 1) from counter k of a stream
__host__ __device__ inline double uniform(uint64_t stream, uint64_t k) {
  const uint64_t z = splitmix64(stream ^ splitmix64(k));
  return (static_cast<double>(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Box-Muller normal deviate number k of a stream
__host__ __device__ inline double normal(uint64_t stream, uint64_t k) {
  const double u1 = uniform(stream, 2 * k), u2 = uniform(stream, 2 * k + 1);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
}

// Inverse standard normal CDF: Halley iteration on Phi(x) - p with
// Phi(x) = erfc(-x/sqrt2)/2, started from the Tukey-lambda approximation
// x0 = 4.91 (p^0.14 - (1-p)^0.14).  The upper half uses Phi^-1(p) =
// -Phi^-1(1-p) (1-p is exact there) so both tails are solved where erfc is
// accurate.
__host__ __device__ inline double inverse_normal(double p) {
  double sign = 1.0;
  if (p > 0.5) {
    p = 1.0 - p;
    sign = -1.0;
  }
  double x = 4.91 * (pow(p, 0.14) - pow(1.0 - p, 0.14));
  for (int it = 0; it < 60; ++it) {
    const double phi = 0.5 * erfc(-x / sqrt(2.0));
    const double pdf = exp(-0.5 * x * x) / sqrt(2.0 * kPi);
    const double f = phi - p;
    if (pdf <= 0.0) break;
    const double r = f / pdf;
    const double dx = r / (1.0 + 0.5 * x * r);
    x -= dx;
    if (fabs(dx) <= 1e-16 * fmax(1.0, fabs(x))) break;
  }
  return sign * x;
}

struct Spec {
  uint32_t v[kDims][32];  // Sobol direction integers
  uint32_t shift[kDims];  // digital shift
  double gamma0, espread_rel, chirp_rel;
  double sx, sxp, sy, syp, alphax, alphay;
  double xcen, ycen, pxcen, pycen;
  double slicelength;
  uint64_t seed;
  int nslice, npart, nbins, shotnoise;
};

// Sobol point n (Gray-code order) of the six dimensions
__host__ __device__ inline void sobol_point(const Spec& sp, uint64_t n, uint32_t* out) {
  const uint64_t g = n ^ (n >> 1);
  for (int d = 0; d < kDims; ++d) {
    uint32_t x = 0;
    for (int i = 0; i < 32; ++i)
      if ((g >> i) & 1) x ^= sp.v[d][i];
    out[d] = x;
  }
}

// Beamlet j of slice s: coordinates shared by its nbins particles and the
// phases of each (th[0..nbins), nbins <= 64).  current: the slice current
// (shot noise amplitude); ce: electron charge times c.
__host__ __device__ inline void beamlet(const Spec& sp, int s, int j, double current, double ce,
                                        double& x, double& px, double& y, double& py,
                                        double& gamma, double* th) {
  const int nb = sp.npart / sp.nbins;
  uint32_t pt[kDims];
  sobol_point(sp, 1 + static_cast<uint64_t>(s) * nb + j, pt);
  double u[kDims];
  for (int d = 0; d < kDims; ++d)
    u[d] = (static_cast<double>(pt[d] ^ sp.shift[d]) + 0.5) * (1.0 / 4294967296.0);
  const double gx = inverse_normal(u[0]);
  const double gpx = inverse_normal(u[1]);
  const double gy = inverse_normal(u[2]);
  const double gpy = inverse_normal(u[3]);
  const double gg = inverse_normal(u[4]);
  const double chirp =
      sp.nslice > 1 ? sp.chirp_rel * sp.gamma0 * (static_cast<double>(s) / (sp.nslice - 1) - 0.5) : 0.0;
  x = sp.xcen + sp.sx * gx;
  px = sp.pxcen + sp.gamma0 * sp.sxp * (gpx - sp.alphax * gx);
  y = sp.ycen + sp.sy * gy;
  py = sp.pycen + sp.gamma0 * sp.syp * (gpy - sp.alphay * gy);
  gamma = sp.gamma0 * (1.0 + sp.espread_rel * gg) + chirp;
  const double twopi = 2.0 * kPi;
  const double base = u[5] * twopi / sp.nbins;
  for (int k = 0; k < sp.nbins; ++k) th[k] = base + twopi * k / sp.nbins;
  const double ne = sp.shotnoise ? current * sp.slicelength / ce / nb : 0.0;
  if (sp.shotnoise && ne > 0.0) {
    const uint64_t stream = slice_seed(sp.seed, s);
    double t0[64];
    for (int k = 0; k < sp.nbins; ++k) t0[k] = th[k];
    for (int h = 1; h <= sp.nbins / 2; ++h) {
      const bool nyq = (sp.nbins % 2 == 0) && (2 * h == sp.nbins);
      const double sig = sqrt((nyq ? 1.0 : 2.0) / (static_cast<double>(h) * h * ne));
      const uint64_t k0 = (static_cast<uint64_t>(j) * 64 + static_cast<uint64_t>(h)) * 2;
      const double a = sig * normal(stream, k0), b = sig * normal(stream, k0 + 1);
      for (int k = 0; k < sp.nbins; ++k) th[k] += a * cos(h * t0[k]) + b * sin(h * t0[k]);
    }
  }
}

}  // namespace loader

// Loader parameters of a beam specification; false if it is invalid.
bool loader_spec(const g4h_beam_spec* sp, loader::Spec& out);

}  // namespace g4
