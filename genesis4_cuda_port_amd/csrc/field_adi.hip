// field_adi.hip - ADI paraxial step (FieldSolverADICUDA.cpp:176-380,
// 789-816, 821-866) as two Thomas sweeps in FP64 for both precisions.
//
// Step A: r = u + i rtmp d2y(u) + S, solve (1 - i rtmp d2x) u* = r along rows.
// Step B: r = u* + i rtmp d2x(u*) + S, solve (1 - i rtmp d2y) u = r along
// columns.  Missing neighbours at the grid edge are zero (the reference's
// edge rule).  Thomas coefficients c, cbet, cwet are the reference's
// (FieldSolverADICUDA.cpp:821-866), built once per (delz, ngrid) on the host.
//
// One workgroup owns one ring plane, so the sweeps update the plane in place.
//  - x sweep: blocks of kB columns of all rows are staged through LDS with
//    coalesced row segments; each thread runs the recurrence of one row; the
//    forward values go to an FP64 scratch plane and come back for the
//    backward substitution.
//  - y sweep: each thread runs the recurrence of one column; rows are read
//    coalesced in chunks of kR rows (one memory latency per chunk), the x
//    neighbours come from an LDS row buffer.  The backward pass of the y
//    sweep produces the final field and reduces the field diagnostics of the
//    plane when asked.
// The RHS, the source add and the Thomas recurrences are FP64 in both
// precisions; the stored field is FP32 in MIXED.
#include "adi_lines.hpp"

#include <complex>
#include <vector>

namespace g4 {
namespace {

using adi::A2;
using adi::kLane;
using adi::rhs_of;
using adi::to_d;
using adi::zadd;
using adi::zmul;
using adi::zsub;

constexpr int kB = 8;   // columns per x-sweep block
constexpr int kR = 8;   // rows per y-sweep chunk

// Floating-point contraction is off in this file: the fused multiply-adds
// below are the only ones, so every kernel variant rounds identically.
#pragma clang fp contract(off)

template <class P>
__global__ __launch_bounds__(1024) void adi_x_kernel(typename A2<P>::T* u, double2* ys, int nslice,
                                                     int G, double cs, const double2* c_d,
                                                     const double2* cbet_d, const double2* cwet_d,
                                                     const typename A2<P>::T* src, const int4* rect,
                                                     int first, int add_src) {
  using T2 = typename A2<P>::T;
  extern __shared__ double2 sm[];
  double2* cl = sm;
  double2* cb = sm + G;
  double2* cw = sm + 2 * G;
  double2* tile = sm + 3 * G;  // (G+2) rows x kB
  const int p = blockIdx.x;
  const int t = threadIdx.x;
  const int nt = blockDim.x;
  const int64_t plane = static_cast<int64_t>(G) * G;
  T2* up = u + p * plane;
  double2* yp = ys + p * plane;
  for (int k = t; k < G; k += nt) {
    cl[k] = c_d[k];
    cb[k] = cbet_d[k];
    cw[k] = cwet_d[k];
  }
  for (int k = t; k < kB; k += nt) {
    tile[k] = make_double2(0.0, 0.0);
    tile[(G + 1) * kB + k] = make_double2(0.0, 0.0);
  }
  const int s = ((p - first) % nslice + nslice) % nslice;
  int4 rc = make_int4(0, 0, 0, 0);
  const T2* sp = nullptr;
  if (add_src) {
    rc = rect[s];
    sp = src + s * plane;
  }
  const int nblk = (G + kB - 1) / kB;
  double2 yprev = make_double2(0.0, 0.0);
  for (int b = 0; b < nblk; ++b) {
    const int c0 = b * kB, nb = min(kB, G - c0);
    __syncthreads();
    for (int e = t; e < G * nb; e += nt) {
      const int iy = e / nb, c = e - iy * nb;
      tile[(iy + 1) * kB + c] = to_d(up[static_cast<int64_t>(iy) * G + c0 + c]);
    }
    __syncthreads();
    double2 yv[kB];
    if (t < G) {
#pragma unroll
      for (int c = 0; c < kB; ++c) {
        if (c < nb) {
          const int k = c0 + c;
          const double2 center = tile[(t + 1) * kB + c];
          double2 r = rhs_of(center, tile[(t + 2) * kB + c], tile[t * kB + c], cs);
          if (sp && k >= rc.x && k < rc.z && t >= rc.y && t < rc.w)
            r = zadd(r, to_d(sp[static_cast<int64_t>(t) * G + k]));
          yprev = zmul(zsub(r, zmul(cl[k], yprev)), cb[k]);
          yv[c] = yprev;
        }
      }
    }
    __syncthreads();
    if (t < G) {
#pragma unroll
      for (int c = 0; c < kB; ++c)
        if (c < nb) tile[(t + 1) * kB + c] = yv[c];
    }
    __syncthreads();
    for (int e = t; e < G * nb; e += nt) {
      const int iy = e / nb, c = e - iy * nb;
      yp[static_cast<int64_t>(iy) * G + c0 + c] = tile[(iy + 1) * kB + c];
    }
  }
  // backward substitution, right to left
  double2 unext = make_double2(0.0, 0.0);
  for (int b = nblk - 1; b >= 0; --b) {
    const int c0 = b * kB, nb = min(kB, G - c0);
    __syncthreads();
    for (int e = t; e < G * nb; e += nt) {
      const int iy = e / nb, c = e - iy * nb;
      tile[(iy + 1) * kB + c] = yp[static_cast<int64_t>(iy) * G + c0 + c];
    }
    __syncthreads();
    if (t < G) {
#pragma unroll
      for (int c = kB - 1; c >= 0; --c) {
        if (c < nb) {
          const int k = c0 + c;
          const double2 y = tile[(t + 1) * kB + c];
          const double2 uu = (k == G - 1) ? y : zsub(y, zmul(cw[k + 1], unext));
          tile[(t + 1) * kB + c] = uu;
          unext = uu;
        }
      }
    }
    __syncthreads();
    for (int e = t; e < G * nb; e += nt) {
      const int iy = e / nb, c = e - iy * nb;
      const double2 v = tile[(iy + 1) * kB + c];
      T2 o;
      o.x = static_cast<P>(v.x);
      o.y = static_cast<P>(v.y);
      up[static_cast<int64_t>(iy) * G + c0 + c] = o;
    }
  }
}

template <class P>
__global__ __launch_bounds__(1024) void adi_y_kernel(typename A2<P>::T* u, double2* ys, int nslice,
                                                     int G, double cs, const double2* c_d,
                                                     const double2* cbet_d, const double2* cwet_d,
                                                     const typename A2<P>::T* src, const int4* rect,
                                                     int first, int add_src, double* stats) {
  using T2 = typename A2<P>::T;
  extern __shared__ double2 sm[];
  double2* cl = sm;
  double2* cb = sm + G;
  double2* cw = sm + 2 * G;
  double2* rb = sm + 3 * G;  // 2 x (G+2) row buffers
  __shared__ double red[16][8];
  const int p = blockIdx.x;
  const int t = threadIdx.x;
  const int nt = blockDim.x;
  const int64_t plane = static_cast<int64_t>(G) * G;
  T2* up = u + p * plane;
  double2* yp = ys + p * plane;
  for (int k = t; k < G; k += nt) {
    cl[k] = c_d[k];
    cb[k] = cbet_d[k];
    cw[k] = cwet_d[k];
  }
  if (t < 2) {
    rb[t * (G + 2)] = make_double2(0.0, 0.0);
    rb[t * (G + 2) + G + 1] = make_double2(0.0, 0.0);
  }
  const int s = ((p - first) % nslice + nslice) % nslice;
  int4 rc = make_int4(0, 0, 0, 0);
  const T2* sp = nullptr;
  if (add_src) {
    rc = rect[s];
    sp = src + s * plane;
  }
  const bool col = t < G;
  const bool src_col = sp && t >= rc.x && t < rc.z;
  double2 yprev = make_double2(0.0, 0.0);
  for (int k0 = 0; k0 < G; k0 += kR) {
    double2 v[kR];
    double2 sv[kR];
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      const int k = k0 + r;
      v[r] = make_double2(0.0, 0.0);
      sv[r] = make_double2(0.0, 0.0);
      if (col && k < G) {
        v[r] = to_d(ld_stream(&up[static_cast<int64_t>(k) * G + t]));
        if (src_col && k >= rc.y && k < rc.w) sv[r] = to_d(ld_stream(&sp[static_cast<int64_t>(k) * G + t]));
      }
    }
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      const int k = k0 + r;
      if (k < G) {  // uniform
        double2* buf = rb + (k & 1) * (G + 2);
        if (col) buf[t + 1] = v[r];
        __syncthreads();
        if (col) {
          double2 rr = rhs_of(v[r], buf[t + 2], buf[t], cs);
          if (src_col && k >= rc.y && k < rc.w) rr = zadd(rr, sv[r]);
          yprev = zmul(zsub(rr, zmul(cl[k], yprev)), cb[k]);
          st_stream(&yp[static_cast<int64_t>(k) * G + t], yprev);
        }
      }
    }
  }
  // backward substitution, bottom to top; field diagnostics of the plane
  const double shift = -0.5 * static_cast<double>(G - 1);
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const int64_t icenter = (plane - 1) / 2;
  double2 unext = make_double2(0.0, 0.0);
  const int nchunk = (G + kR - 1) / kR;
  for (int ch = nchunk - 1; ch >= 0; --ch) {
    const int k0 = ch * kR;
    double2 y[kR];
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      const int k = k0 + r;
      y[r] = (col && k < G) ? ld_stream(&yp[static_cast<int64_t>(k) * G + t]) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int r = kR - 1; r >= 0; --r) {
      const int k = k0 + r;
      if (col && k < G) {
        const double2 uu = (k == G - 1) ? y[r] : zsub(y[r], zmul(cw[k + 1], unext));
        unext = uu;
        T2 o;
        o.x = static_cast<P>(uu.x);
        o.y = static_cast<P>(uu.y);
        const int64_t gi = static_cast<int64_t>(k) * G + t;
        st_stream(&up[gi], o);
        if (stats) {
          const double re = o.x, im = o.y;
          const double wei = re * re + im * im;
          const double dx = t + shift, dy = k + shift;
          acc[0] += wei;
          acc[1] += dx * wei;
          acc[2] += dx * dx * wei;
          acc[3] += dy * wei;
          acc[4] += dy * dy * wei;
          acc[5] += re;
          acc[6] += im;
          if (gi == icenter) {
            stats[p * G4H_FIELD_RAW + 7] = re;
            stats[p * G4H_FIELD_RAW + 8] = im;
          }
        }
      }
    }
  }
  if (stats) {
    const int lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      double v = acc[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (t < 7) {
      double v = 0.0;
      for (int w = 0; w < nt / 64; ++w) v += red[w][t];
      stats[p * G4H_FIELD_RAW + t] = v;
    }
  }
}

// ---- lane variant ---------------------------------------------------------
// For few planes (steady-state batches, short time windows) the workgroup
// per plane above leaves most of the GPU idle and pays a barrier per row.
// Here one 64-lane wave owns 64 rows (x sweep) or 64 columns (y sweep) of a
// plane (adi_lines.hpp), so a plane of G = 255 is four independent waves.
// The blocks of a plane may run in any order and read each other's edge
// lines, so the x sweep writes its result to a second plane set (ustar)
// instead of in place, and the y sweep reads it from there.  Every element
// goes through the same operations in the same order as in the workgroup
// kernels, so both variants give bit-identical fields and, with the
// per-block partial sums reduced in block order, bit-identical stats.

template <class P, int kCh>
__global__ __launch_bounds__(kLane) void adi_x_lane_kernel(const typename A2<P>::T* u, double2* ys,
                                                           typename A2<P>::T* ustar, int nslice,
                                                           int G, double cs,
                                                           const double2* __restrict__ c_d,
                                                           const double2* __restrict__ cbet_d,
                                                           const double2* __restrict__ cwet_d,
                                                           const typename A2<P>::T* src,
                                                           const int4* rect, int first,
                                                           int add_src) {
  using T2 = typename A2<P>::T;
  const int p = blockIdx.x;
  const int t = threadIdx.x;
  // the Thomas factors in LDS: read in the recurrence at one element per
  // step, from memory they would put a load latency on every step
  extern __shared__ double2 coef[];
  const adi::Factors f = adi::stage_factors(coef, c_d, cbet_d, cwet_d, G, t, kLane);
  __syncthreads();
  const int64_t plane = static_cast<int64_t>(G) * G;
  const int s = ((p - first) % nslice + nslice) % nslice;
  int4 rc = make_int4(0, 0, 0, 0);
  const T2* sp = nullptr;
  if (add_src) {
    rc = rect[s];
    sp = src + s * plane;
  }
  adi::x_forward<P, kCh>(u + p * plane, ys + p * plane, G, cs, f, sp, rc, blockIdx.y, t);
  adi::x_backward<P, kCh>(ys + p * plane, ustar + p * plane, G, f, blockIdx.y, t);
}

// y sweep of ustar into u; the back substitution writes the block's partial
// field sums (partials[(p * nblk + b) * 8 + k], centre cell at the tail),
// reduced in block order afterwards.
template <class P, int kCh>
__global__ __launch_bounds__(kLane) void adi_y_lane_kernel(typename A2<P>::T* u, double2* ys,
                                                           const typename A2<P>::T* ustar,
                                                           int nslice, int G, double cs,
                                                           const double2* __restrict__ c_d,
                                                           const double2* __restrict__ cbet_d,
                                                           const double2* __restrict__ cwet_d,
                                                           const typename A2<P>::T* src,
                                                           const int4* rect, int first,
                                                           int add_src, double* partials) {
  using T2 = typename A2<P>::T;
  const int p = blockIdx.x;
  const int t = threadIdx.x;
  const int nblk = gridDim.y;
  extern __shared__ double2 coef[];
  const adi::Factors f = adi::stage_factors(coef, c_d, cbet_d, cwet_d, G, t, kLane);
  __syncthreads();
  const int64_t plane = static_cast<int64_t>(G) * G;
  const int s = ((p - first) % nslice + nslice) % nslice;
  int4 rc = make_int4(0, 0, 0, 0);
  const T2* sp = nullptr;
  if (add_src) {
    rc = rect[s];
    sp = src + s * plane;
  }
  adi::y_forward<P, kCh>(ustar + p * plane, ys + p * plane, G, cs, f, sp, rc, blockIdx.y, t);
  adi::PlaneSums sums;
  sums.clear();
  if (partials) {
    double* centre = partials + static_cast<int64_t>(nslice) * nblk * 8 + 2 * p;
    adi::y_backward<P, kCh, true>(ys + p * plane, u + p * plane, G, f, blockIdx.y, t, sums,
                                  centre);
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const double v = adi::wave_sum64(sums.acc[k]);
      if (t == 0) partials[(static_cast<int64_t>(p) * nblk + blockIdx.y) * 8 + k] = v;
    }
  } else {
    adi::y_backward<P, kCh, false>(ys + p * plane, u + p * plane, G, f, blockIdx.y, t, sums,
                                   nullptr);
  }
}

// Segmented sweeps (adi_lines.hpp): a workgroup of kSegs waves per 64 lines
// of a plane, every element in registers between its load and its store,
// the tables of the step (c, cbet, cwet, beta, delta) staged in LDS.  For
// G <= kSegs * kSegMax.
template <class P, int LPW>
__global__ __launch_bounds__(adi::kSegs * LPW, 4) void adi_x_seg_kernel(
    const typename A2<P>::T* u, typename A2<P>::T* ustar, int nslice, int G, int m, double cs,
    const double2* __restrict__ tab, const typename A2<P>::T* src, const int4* rect, int first,
    int add_src) {
  using T2 = typename A2<P>::T;
  extern __shared__ double2 lds[];
  for (int k = threadIdx.x; k < 5 * G; k += blockDim.x) lds[k] = tab[k];
  __syncthreads();
  const int p = blockIdx.x;
  const int64_t plane = static_cast<int64_t>(G) * G;
  const int s = ((p - first) % nslice + nslice) % nslice;
  int4 rc = make_int4(0, 0, 0, 0);
  const T2* sp = nullptr;
  if (add_src) {
    rc = rect[s];
    sp = src + s * plane;
  }
  const adi::SegTables tb{lds, lds + G, lds + 2 * G, lds + 3 * G, lds + 4 * G};
  adi::seg_sweep<P, LPW, true, false>(u + p * plane, ustar + p * plane, G, m, cs, tb, sp, rc,
                                 blockIdx.y, lds + 5 * G, nullptr, nullptr, nullptr);
}

// ... and the y sweep, with the block's field sums in the layout of the
// wave-per-block variants (partials[(p * nblk + b) * 8 + k], centres at the
// tail).
template <class P, int LPW>
__global__ __launch_bounds__(adi::kSegs * LPW, 4) void adi_y_seg_kernel(
    typename A2<P>::T* u, const typename A2<P>::T* ustar, int nslice, int G, int m, double cs,
    const double2* __restrict__ tab, const typename A2<P>::T* src, const int4* rect, int first,
    int add_src, double* partials) {
  using T2 = typename A2<P>::T;
  extern __shared__ double2 lds[];
  __shared__ double red[adi::kSegs * LPW / kLane * 8];
  for (int k = threadIdx.x; k < 5 * G; k += blockDim.x) lds[k] = tab[k];
  __syncthreads();
  const int p = blockIdx.x;
  const int nblk = gridDim.y;
  const int64_t plane = static_cast<int64_t>(G) * G;
  const int s = ((p - first) % nslice + nslice) % nslice;
  int4 rc = make_int4(0, 0, 0, 0);
  const T2* sp = nullptr;
  if (add_src) {
    rc = rect[s];
    sp = src + s * plane;
  }
  const adi::SegTables tb{lds, lds + G, lds + 2 * G, lds + 3 * G, lds + 4 * G};
  if (partials) {
    double* part = partials + (static_cast<int64_t>(p) * nblk + blockIdx.y) * 8;
    double* centre = partials + static_cast<int64_t>(nslice) * nblk * 8 + 2 * p;
    adi::seg_sweep<P, LPW, false, true>(ustar + p * plane, u + p * plane, G, m, cs, tb, sp, rc,
                                   blockIdx.y, lds + 5 * G, red, part, centre);
  } else {
    adi::seg_sweep<P, LPW, false, false>(ustar + p * plane, u + p * plane, G, m, cs, tb, sp, rc,
                                    blockIdx.y, lds + 5 * G, nullptr, nullptr, nullptr);
  }
}

// The segmented sweeps of every run of a batch in one launch: workgroup b
// takes block b of the descriptor list (one run's 64-line block), the run's
// group tables and coefficient of step k.
template <class P>
__global__ __launch_bounds__(adi::kSegs * kLane, 4) void adi_x_seg_batch_kernel(
    const BatchSegDesc* __restrict__ blocks, int k, int add_src) {
  using T2 = typename A2<P>::T;
  extern __shared__ double2 lds[];
  const BatchSegDesc D = blocks[blockIdx.x];
  const int G = D.G;
  const double2* tab = D.adi[k];
  for (int i = threadIdx.x; i < 5 * G; i += blockDim.x) lds[i] = tab[i];
  __syncthreads();
  const int64_t plane = static_cast<int64_t>(G) * G;
  int4 rc = make_int4(0, 0, 0, 0);
  const T2* sp = nullptr;
  if (add_src) {
    rc = D.rect[D.slice];
    sp = static_cast<const T2*>(D.src) + D.slice * plane;
  }
  const adi::SegTables tb{lds, lds + G, lds + 2 * G, lds + 3 * G, lds + 4 * G};
  adi::seg_sweep<P, kLane, true, false>(static_cast<const T2*>(D.u) + D.slice * plane,
                                        static_cast<T2*>(D.ustar) + D.slice * plane, G,
                                        (G + adi::kSegs - 1) / adi::kSegs, D.cs[k], tb, sp, rc,
                                        D.blk, lds + 5 * G, nullptr, nullptr, nullptr);
}

template <class P, bool STATS>
__global__ __launch_bounds__(adi::kSegs * kLane, 4) void adi_y_seg_batch_kernel(
    const BatchSegDesc* __restrict__ blocks, int k, int add_src, int clear_src) {
  using T2 = typename A2<P>::T;
  extern __shared__ double2 lds[];
  __shared__ double red[adi::kSegs * 8];
  const BatchSegDesc D = blocks[blockIdx.x];
  const int G = D.G;
  const double2* tab = D.adi[k];
  for (int i = threadIdx.x; i < 5 * G; i += blockDim.x) lds[i] = tab[i];
  __syncthreads();
  const int64_t plane = static_cast<int64_t>(G) * G;
  int4 rc = make_int4(0, 0, 0, 0);
  const T2* sp = nullptr;
  if (add_src) {
    rc = D.rect[D.slice];
    sp = static_cast<const T2*>(D.src) + D.slice * plane;
  }
  const adi::SegTables tb{lds, lds + G, lds + 2 * G, lds + 3 * G, lds + 4 * G};
  const int m = (G + adi::kSegs - 1) / adi::kSegs;
  const T2* in = static_cast<const T2*>(D.ustar) + D.slice * plane;
  T2* out = static_cast<T2*>(D.u) + D.slice * plane;
  T2* clr = clear_src && sp ? const_cast<T2*>(sp) : nullptr;
  if constexpr (STATS) {
    double* part = D.partials + (static_cast<int64_t>(D.slice) * D.nblk + D.blk) * 8;
    double* centre = D.partials + static_cast<int64_t>(D.nrun) * D.nblk * 8 + 2 * D.slice;
    adi::seg_sweep<P, kLane, false, true>(in, out, G, m, D.cs[k], tb, sp, rc, D.blk,
                                          lds + 5 * G, red, part, centre, clr);
  } else {
    adi::seg_sweep<P, kLane, false, false>(in, out, G, m, D.cs[k], tb, sp, rc, D.blk,
                                           lds + 5 * G, nullptr, nullptr, nullptr, clr);
  }
}

// Field record r of every run from its block partials, summed in block order
// (as reduce_partials_kernel sums them for a single field).
__global__ void batch_field_record_kernel(const BatchRunDesc* __restrict__ runs, int nruns,
                                          int r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nruns) return;
  const BatchRunDesc R = runs[i];
  const int nblk = (R.G + kLane - 1) / kLane;
  double* o = R.field_hist + (static_cast<int64_t>(r) * R.nrun + R.slice) * G4H_FIELD_RAW;
  for (int q = 0; q < 7; ++q) {
    double v = 0.0;
    for (int t = 0; t < nblk; ++t) v += R.partials[(static_cast<int64_t>(R.slice) * nblk + t) * 8 + q];
    o[q] = v;
  }
  const double* c = R.partials + static_cast<int64_t>(R.nrun) * nblk * 8 + 2 * R.slice;
  o[7] = c[0];
  o[8] = c[1];
}

// The x sweep's output of the wave-per-block variants (the y sweep's input),
// one plane per ring slot in the field's storage type.
void* adi_ustar(Field* f) {
  const size_t n = static_cast<size_t>(f->cells());
  if (f->prec == G4H_FP64) {
    if (f->adi_sd.size() < n) f->adi_sd.alloc(f->ctx, n);
    return f->adi_sd.get();
  }
  if (f->adi_sf.size() < n) f->adi_sf.alloc(f->ctx, n);
  return f->adi_sf.get();
}

template <class P>
void adi_step(Field* f, bool source_on, bool stats) {
  using T2 = typename A2<P>::T;
  const int G = f->ngrid;
  const int nt = ((G + 63) / 64) * 64;
  const double cs = 0.25 * f->adi_delz / (f->xks * f->dgrid * f->dgrid);
  T2* u = static_cast<T2*>(f->data());
  const T2* src = static_cast<const T2*>(f->src());
  const size_t ldsx = sizeof(double2) * (3 * static_cast<size_t>(G) + (G + 2) * static_cast<size_t>(kB));
  const size_t ldsy = sizeof(double2) * (3 * static_cast<size_t>(G) + 2 * static_cast<size_t>(G + 2));
  const int add = source_on ? 1 : 0;
  // variant 0: workgroup per plane (its x sweep stages G + 2 rows of kB
  // columns plus three coefficient tables in LDS, so G <= 360; above that
  // it runs as variant 1); variants 1 and 2: wave per 64 lines, any G, with
  // the loads 8 or 16 elements ahead of the recurrence (16 holds more loads
  // in flight, which is what a sweep over few planes waits on, for more
  // registers).  Tuned per shape; all three give bit-identical fields.
  //
  // Variant 3, the segmented sweeps, holds every element in registers (no
  // FP64 forward-value plane) and spreads a line over kSegs waves; it rounds
  // differently from the single-chain variants, so it is not a tuning
  // candidate: it is what runs for G <= kSegs * kSegMax unless a variant is
  // pinned, and the tuner chooses among variants 0-2 for larger grids only.
  const bool segable = G <= adi::kSegs * adi::kSegMax;
  const int forced = f->ctx->tuner.forced[TUNE_ADI];
  if (segable && (forced < 0 || forced == 3)) {
    auto seg = [&](auto lpw) {
      constexpr int LPW = decltype(lpw)::value;
      const int nblk = (G + LPW - 1) / LPW;
      const int m = (G + adi::kSegs - 1) / adi::kSegs;
      double* partials = nullptr;
      if (stats) {
        const size_t need = static_cast<size_t>(f->nslice) * nblk * 8 + 2 * static_cast<size_t>(f->nslice);
        if (f->col_partials.size() < need) f->col_partials.alloc(f->ctx, need);
        partials = f->col_partials.get();
      }
      const size_t lds = sizeof(double2) * (5 * static_cast<size_t>(G) + adi::kSegs * LPW);
      T2* ustar = static_cast<T2*>(adi_ustar(f));
      f->ctx->launch(G4H_K_ADI_X, [&] {
        hipLaunchKernelGGL((adi_x_seg_kernel<P, LPW>), dim3(f->nslice, nblk),
                           dim3(adi::kSegs * LPW), lds, f->ctx->stream, u, ustar, f->nslice, G,
                           m, cs, f->adi_c, src, f->src_rect.get(), f->first, add);
      });
      f->ctx->launch(G4H_K_ADI_Y, [&] {
        hipLaunchKernelGGL((adi_y_seg_kernel<P, LPW>), dim3(f->nslice, nblk),
                           dim3(adi::kSegs * LPW), lds, f->ctx->stream, u, ustar, f->nslice, G,
                           m, cs, f->adi_c, src, f->src_rect.get(), f->first, add, partials);
      });
      if (stats) field_reduce_partials(f, partials, nblk);
    };
    // 64 lines per workgroup (one wave per segment): with 32 (two segments
    // per wave) the segment bounds are no longer wave-uniform, the element
    // loops diverge and spill, and the time-dependent G = 151 sweeps ran
    // 1.5x slower (0.52 against 0.35 ms per x sweep)
    seg(std::integral_constant<int, 64>{});
    return;
  }
  if (f->adi_y.size() < static_cast<size_t>(f->cells())) f->adi_y.alloc(f->ctx, f->cells());
  TuneScope tune(f->ctx, TUNE_ADI, {G, f->nslice, static_cast<int>(f->prec)}, 3);
  const int picked = forced == 3 ? 1 : tune.variant();  // 3 pinned on a grid it cannot take
  const int variant = (picked == 0 && G > 360) ? 1 : picked;
  if (variant >= 1) {
    const int nblk = (G + kLane - 1) / kLane;
    double* partials = nullptr;
    if (stats) {
      const size_t need = static_cast<size_t>(f->nslice) * nblk * 8 + 2 * static_cast<size_t>(f->nslice);
      if (f->col_partials.size() < need) f->col_partials.alloc(f->ctx, need);
      partials = f->col_partials.get();
    }
    const size_t ldsc = sizeof(double2) * 3 * static_cast<size_t>(G);
    T2* ustar = static_cast<T2*>(adi_ustar(f));
    auto sweeps = [&](auto ch) {
      constexpr int CH = decltype(ch)::value;
      f->ctx->launch(G4H_K_ADI_X, [&] {
        hipLaunchKernelGGL((adi_x_lane_kernel<P, CH>), dim3(f->nslice, nblk), dim3(kLane), ldsc,
                           f->ctx->stream, u, f->adi_y.get(), ustar, f->nslice, G, cs, f->adi_c,
                           f->adi_cbet, f->adi_cwet, src, f->src_rect.get(), f->first, add);
      });
      f->ctx->launch(G4H_K_ADI_Y, [&] {
        hipLaunchKernelGGL((adi_y_lane_kernel<P, CH>), dim3(f->nslice, nblk), dim3(kLane), ldsc,
                           f->ctx->stream, u, f->adi_y.get(), ustar, f->nslice, G, cs, f->adi_c,
                           f->adi_cbet, f->adi_cwet, src, f->src_rect.get(), f->first, add,
                           partials);
      });
    };
    if (variant == 1)
      sweeps(std::integral_constant<int, 8>{});
    else
      sweeps(std::integral_constant<int, 16>{});
    if (stats) field_reduce_partials(f, partials, nblk);
    return;
  }
  f->ctx->launch(G4H_K_ADI_X, [&] {
    hipLaunchKernelGGL(adi_x_kernel<P>, dim3(f->nslice), dim3(nt), ldsx, f->ctx->stream, u,
                       f->adi_y.get(), f->nslice, G, cs, f->adi_c, f->adi_cbet,
                       f->adi_cwet, src, f->src_rect.get(), f->first, add);
  });
  f->ctx->launch(G4H_K_ADI_Y, [&] {
    hipLaunchKernelGGL(adi_y_kernel<P>, dim3(f->nslice), dim3(nt), ldsy, f->ctx->stream, u,
                       f->adi_y.get(), f->nslice, G, cs, f->adi_c, f->adi_cbet,
                       f->adi_cwet, src, f->src_rect.get(), f->first, add,
                       stats ? f->plane_stats.get() : nullptr);
  });
}

}  // namespace

void* field_adi_ustar(Field* f) { return adi_ustar(f); }

void launch_batch_adi_step(Context* ctx, g4h_precision prec, const BatchSegDesc* d_blocks,
                           int nblocks, int k, int gmax, bool add_src, bool stats,
                           bool clear_src) {
  if (nblocks <= 0) return;
  require(gmax <= adi::kSegs * adi::kSegMax, "segmented sweeps: ngrid above 256");
  const size_t lds = sizeof(double2) * (5 * static_cast<size_t>(gmax) + adi::kSegs * kLane);
  const int src = add_src ? 1 : 0, clr = clear_src ? 1 : 0;
  ctx->launch(G4H_K_ADI_X, [&] {
    if (prec == G4H_FP64)
      hipLaunchKernelGGL(adi_x_seg_batch_kernel<double>, dim3(nblocks), dim3(adi::kSegs * kLane),
                         lds, ctx->stream, d_blocks, k, src);
    else
      hipLaunchKernelGGL(adi_x_seg_batch_kernel<float>, dim3(nblocks), dim3(adi::kSegs * kLane),
                         lds, ctx->stream, d_blocks, k, src);
  });
  auto y = [&](auto kern) {
    ctx->launch(G4H_K_ADI_Y, [&] {
      hipLaunchKernelGGL(kern, dim3(nblocks), dim3(adi::kSegs * kLane), lds, ctx->stream,
                         d_blocks, k, src, clr);
    });
  };
  if (prec == G4H_FP64) {
    if (stats)
      y(adi_y_seg_batch_kernel<double, true>);
    else
      y(adi_y_seg_batch_kernel<double, false>);
  } else {
    if (stats)
      y(adi_y_seg_batch_kernel<float, true>);
    else
      y(adi_y_seg_batch_kernel<float, false>);
  }
}

void launch_batch_field_records(Context* ctx, const BatchRunDesc* d_runs, int nruns, int r) {
  if (nruns <= 0) return;
  ctx->launch(G4H_K_OTHER, [&] {
    hipLaunchKernelGGL(batch_field_record_kernel, dim3((nruns + 127) / 128), dim3(128), 0,
                       ctx->stream, d_runs, nruns, r);
  });
}

// Thomas factors of FieldSolverADICUDA::init (FieldSolverADICUDA.cpp:821-866).
void adi_prepare(Field* f, double delz) {
  const int G = f->ngrid;
  if (f->adi_c && delz == f->adi_delz) return;
  const double key = delz;
  auto it = f->delz_tables.find(key);
  if (it == f->delz_tables.end()) {
    if (f->delz_tables.size() >= Field::kMaxDelzTables) f->evict_delz_tables();
    const double rtmp = 0.25 * delz / (f->xks * f->dgrid * f->dgrid);
    using cd = std::complex<double>;
    std::vector<double> mupp(G), mmid(G), mlow(G);
    mupp[0] = rtmp;
    mmid[0] = -2 * rtmp;
    mlow[0] = 0;
    for (int i = 1; i < G - 1; i++) {
      mupp[i] = rtmp;
      mmid[i] = -2 * rtmp;
      mlow[i] = rtmp;
    }
    mupp[G - 1] = 0;
    mmid[G - 1] = -2 * rtmp;
    mlow[G - 1] = rtmp;
    std::vector<cd> w1(G), w2(G), c(G), cbet(G), cwet(G);
    for (int i = 0; i < G; i++) {
      w1[i] = cd(0.0, -mupp[i]);
      w2[i] = cd(1.0, -mmid[i]);
      c[i] = cd(0.0, -mlow[i]);
    }
    cbet[0] = cd(1.0, 0.0) / w2[0];
    cwet[0] = cd(0.0, 0.0);
    for (int i = 1; i < G; i++) {
      cwet[i] = w1[i - 1] * cbet[i - 1];
      cbet[i] = cd(1.0, 0.0) / (w2[i] - c[i] * cwet[i]);
    }
    // the segmented sweeps' products over each segment of m elements:
    // beta_k = prod_(i=s..k) (-c_i cbet_i), delta_k = prod_(i=k+1..e) (-cwet_i),
    // delta = 0 in the segment that ends the line
    const int m = (G + adi::kSegs - 1) / adi::kSegs;
    std::vector<cd> beta(G), delta(G);
    for (int s0 = 0; s0 < G; s0 += m) {
      const int e0 = std::min(G, s0 + m);
      cd b(1.0, 0.0);
      for (int k = s0; k < e0; ++k) {
        b = -c[k] * cbet[k] * b;
        beta[k] = b;
      }
      cd d(e0 < G ? -cwet[e0] : cd(0.0, 0.0));
      for (int k = e0 - 1; k >= s0; --k) {
        delta[k] = d;
        if (k > s0) d = -cwet[k] * d;
      }
    }
    Field::DelzTable t;
    t.host.resize(5 * static_cast<size_t>(G));
    for (int i = 0; i < G; ++i) {
      t.host[i] = make_double2(c[i].real(), c[i].imag());
      t.host[G + i] = make_double2(cbet[i].real(), cbet[i].imag());
      t.host[2 * G + i] = make_double2(cwet[i].real(), cwet[i].imag());
      t.host[3 * G + i] = make_double2(beta[i].real(), beta[i].imag());
      t.host[4 * G + i] = make_double2(delta[i].real(), delta[i].imag());
    }
    t.dev.alloc(f->ctx, 5 * static_cast<size_t>(G));
    G4_HIP(hipMemcpyAsync(t.dev.get(), t.host.data(), sizeof(double2) * 5 * G,
                          hipMemcpyHostToDevice, f->ctx->stream));
    it = f->delz_tables.emplace(key, std::move(t)).first;
  }
  const double2* d = it->second.dev.get();
  f->adi_c = d;
  f->adi_cbet = d + G;
  f->adi_cwet = d + 2 * G;
  f->adi_delz = delz;
}

void adi_solve(Field* f, bool source_on, bool stats) {
  if (f->prec == G4H_FP64)
    adi_step<double>(f, source_on, stats);
  else
    adi_step<float>(f, source_on, stats);
}

}  // namespace g4
