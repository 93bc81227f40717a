// space_charge.hip - short-range longitudinal space charge on the device.
//
// The reference computes this field on the host (EFieldSolver::shortRange of
// upstream GENESIS 1.3, called per slice at BeamSolverCUDA.cpp:523-547 on the
// host particle arrays, which its CUDA path no longer keeps current) and
// uploads it per particle for the UseParticleEz push.  Here one workgroup per
// slice computes it from the device beam, in FP64 in both precisions, and
// writes the per-particle term the push reads:
//   1. centroid of the slice and its largest radial offset (block sums);
//   2. source of each harmonic l = 1..nz and azimuthal mode m = -nphi..nphi
//      on N radial shells of width rmax * max|r - rc| / N, summed in LDS:
//      S_lmn = sum over particles in shell n of exp(-i (l theta + m phi));
//   3. one thread per mode solves
//      n u_{n-1} - [2n+1 + m^2/(n+1/2) + (n+1/2)(l ks dr)^2/gz2] u_n
//        + (n+1) u_{n+1} = (I vacimp / (2 pi Np)) S_lmn,   u_N = 0
//      (the cell-centred radial Poisson equation of the mode with the
//      relativistic screening l ks / gz, Thomas in LDS);
//   4. every particle gets E_z = sum 2 (l ks / gz2) Im[u_lm(n) e^{i(l theta +
//      m phi)}] and the push term E_z / 511000 (BeamSolverCUDA.cpp:532), plus
//      the slice's loss and a caller's per-particle term when given; the
//      slice's RMS E_z over its particles is kept for g4h_beam_sc_field (its
//      mean vanishes: the field does no net work on a slice to first order).
// The LDS sums of step 2 are unordered, so results agree with the FP64
// oracle to round-off, not bit for bit.
#include "kernels.hpp"

namespace g4 {
namespace {

constexpr int kScThreads = 256;
constexpr int kScWaves = kScThreads / 64;

__device__ __forceinline__ double sc_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double sc_wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Sum (or maximum) over the workgroup, the same value returned to all.
template <bool MAX>
__device__ __forceinline__ double sc_block(double v, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = MAX ? sc_wave_max(v) : sc_wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < kScWaves; ++w) r = MAX ? fmax(r, red[w]) : r + red[w];
  return r;
}

template <class P>
__global__ __launch_bounds__(kScThreads) void sc_field_kernel(
    const P* __restrict__ x, const P* __restrict__ y, const double* __restrict__ th,
    const int32_t* __restrict__ off, const double* __restrict__ cur_over_npar, ScParams sp,
    const double* __restrict__ slice_eloss, const double* __restrict__ base, double* __restrict__ ez,
    double* __restrict__ slice_rms) {
  extern __shared__ double2 lds[];
  __shared__ double red[kScWaves];
  const int s = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t i0 = off[s], i1 = off[s + 1];
  const int64_t np = i1 - i0;
  const int N = sp.ngrid, nz = sp.nz, nphi = sp.nphi;
  const int nm = 2 * nphi + 1;
  const int nmodes = nz * nm;
  double2* S = lds;                                              // [nmodes][N]
  double* cp = reinterpret_cast<double*>(lds + nmodes * N);      // [nmodes][N]
  const double extra = slice_eloss ? slice_eloss[s] : 0.0;
  const double d0 = cur_over_npar[s] * sp.vacimp_over_2pi;  // I vacimp / (2 pi Np)
  if (np <= 0) {
    if (tid == 0) slice_rms[s] = 0.0;
    return;
  }
  // 1. centroid and radius
  double ax = 0.0, ay = 0.0;
  for (int64_t i = i0 + tid; i < i1; i += kScThreads) {
    ax += static_cast<double>(x[i]);
    ay += static_cast<double>(y[i]);
  }
  const double xc = sc_block<false>(ax, red) / static_cast<double>(np);
  const double yc = sc_block<false>(ay, red) / static_cast<double>(np);
  double r2 = 0.0;
  for (int64_t i = i0 + tid; i < i1; i += kScThreads) {
    const double dx = static_cast<double>(x[i]) - xc, dy = static_cast<double>(y[i]) - yc;
    r2 = fmax(r2, dx * dx + dy * dy);
  }
  r2 = sc_block<true>(r2, red);
  const double dr = sp.rmax * sqrt(r2) / N;
  const bool on = d0 != 0.0 && dr > 0.0;
  if (on) {
    // 2. mode sources
    for (int k = tid; k < nmodes * N; k += kScThreads) S[k] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int64_t i = i0 + tid; i < i1; i += kScThreads) {
      const double dx = static_cast<double>(x[i]) - xc, dy = static_cast<double>(y[i]) - yc;
      const double r = sqrt(dx * dx + dy * dy);
      const int n = min(static_cast<int>(r / dr), N - 1);
      const double pc = r > 0.0 ? dx / r : 1.0, ps = r > 0.0 ? dy / r : 0.0;
      double ts, tc;
      sincos(th[i], &ts, &tc);
      double lc = 1.0, ls = 0.0;  // e^{i l theta}
      for (int l = 1; l <= nz; ++l) {
        const double nc = lc * tc - ls * ts, ns = lc * ts + ls * tc;
        lc = nc;
        ls = ns;
        double mc = 1.0, ms = 0.0;  // e^{i m phi}, from m = -nphi
        for (int q = 0; q < nphi; ++q) {
          const double c2 = mc * pc + ms * ps, s2 = ms * pc - mc * ps;
          mc = c2;
          ms = s2;
        }
        for (int m = -nphi; m <= nphi; ++m) {
          // conj(e^{i l theta} e^{i m phi})
          const double re = lc * mc - ls * ms, im = -(lc * ms + ls * mc);
          double2* t = &S[((l - 1) * nm + (m + nphi)) * N + n];
          atomicAdd(&t->x, re);
          atomicAdd(&t->y, im);
          const double c2 = mc * pc - ms * ps, s2 = mc * ps + ms * pc;
          mc = c2;
          ms = s2;
        }
      }
    }
    __syncthreads();
    // 3. radial solves, one thread per mode
    if (tid < nmodes) {
      const int l = tid / nm + 1, m = tid % nm - nphi;
      const double kd = l * sp.ks * dr;
      const double kd2 = kd * kd / sp.gz2;
      double2* u = S + tid * N;
      double* c = cp + tid * N;
      double2 prev = make_double2(0.0, 0.0);
      for (int n = 0; n < N; ++n) {
        const double h = n + 0.5;
        const double b = -(2.0 * n + 1.0) - static_cast<double>(m * m) / h - h * kd2;
        const double a = static_cast<double>(n);
        const double den = n == 0 ? b : b - a * c[n - 1];
        c[n] = (n + 1.0) / den;
        const double2 d = make_double2(d0 * u[n].x, d0 * u[n].y);
        prev = n == 0 ? make_double2(d.x / den, d.y / den)
                      : make_double2((d.x - a * prev.x) / den, (d.y - a * prev.y) / den);
        u[n] = prev;
      }
      for (int n = N - 2; n >= 0; --n) {
        u[n].x -= c[n] * u[n + 1].x;
        u[n].y -= c[n] * u[n + 1].y;
      }
    }
    __syncthreads();
  }
  // 4. field at every particle
  double esum = 0.0;
  for (int64_t i = i0 + tid; i < i1; i += kScThreads) {
    double e = 0.0;
    if (on) {
      const double dx = static_cast<double>(x[i]) - xc, dy = static_cast<double>(y[i]) - yc;
      const double r = sqrt(dx * dx + dy * dy);
      const int n = min(static_cast<int>(r / dr), N - 1);
      const double pc = r > 0.0 ? dx / r : 1.0, ps = r > 0.0 ? dy / r : 0.0;
      double ts, tc;
      sincos(th[i], &ts, &tc);
      double lc = 1.0, ls = 0.0;
      for (int l = 1; l <= nz; ++l) {
        const double nc = lc * tc - ls * ts, ns = lc * ts + ls * tc;
        lc = nc;
        ls = ns;
        double mc = 1.0, ms = 0.0;
        for (int q = 0; q < nphi; ++q) {
          const double c2 = mc * pc + ms * ps, s2 = ms * pc - mc * ps;
          mc = c2;
          ms = s2;
        }
        const double f = 2.0 * l * sp.ks / sp.gz2;
        for (int m = -nphi; m <= nphi; ++m) {
          const double2 uu = S[((l - 1) * nm + (m + nphi)) * N + n];
          const double er = lc * mc - ls * ms, ei = lc * ms + ls * mc;  // e^{i(l theta + m phi)}
          e += f * (uu.x * ei + uu.y * er);  // f Im[u e^{..}]
          const double c2 = mc * pc - ms * ps, s2 = mc * ps + ms * pc;
          mc = c2;
          ms = s2;
        }
      }
    }
    ez[i] = e / 511000.0 + extra + (base ? base[i] : 0.0);
    esum += e * e;
  }
  esum = sc_block<false>(esum, red);
  if (tid == 0) slice_rms[s] = sqrt(esum / static_cast<double>(np));
}

}  // namespace

size_t sc_lds_bytes(const g4h_efield& ef) {
  const size_t modes = static_cast<size_t>(ef.nz) * (2 * ef.nphi + 1);
  return modes * ef.ngrid * (sizeof(double2) + sizeof(double));
}

void launch_space_charge(Context* ctx, Beam* b, double gz2, const double* d_slice_eloss,
                         const double* d_base) {
  if (b->nslice <= 0) return;
  ScParams sp{};
  sp.ks = b->ef_ks;
  sp.gz2 = gz2;
  sp.rmax = b->ef.rmax;
  sp.vacimp_over_2pi = G4H_VACIMP / (2.0 * M_PI);
  sp.ngrid = b->ef.ngrid;
  sp.nz = b->ef.nz;
  sp.nphi = b->ef.nphi;
  const size_t lds = sc_lds_bytes(b->ef);
  ctx->launch(G4H_K_OTHER, [&] {
    if (b->prec == G4H_FP64)
      hipLaunchKernelGGL(sc_field_kernel<double>, dim3(b->nslice), dim3(kScThreads), lds,
                         ctx->stream, b->xd.get(), b->yd.get(), b->theta.get(), b->off.get(),
                         b->cur_over_npar.get(), sp, d_slice_eloss, d_base, b->scez.get(),
                         b->scrms.get());
    else
      hipLaunchKernelGGL(sc_field_kernel<float>, dim3(b->nslice), dim3(kScThreads), lds,
                         ctx->stream, b->xf.get(), b->yf.get(), b->theta.get(), b->off.get(),
                         b->cur_over_npar.get(), sp, d_slice_eloss, d_base, b->scez.get(),
                         b->scrms.get());
  });
}

}  // namespace g4
