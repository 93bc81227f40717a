// adi_lines.hpp - device building blocks of the ADI sweeps (FieldSolverADICUDA.cpp:
// 176-380): one 64-lane wave runs the Thomas recurrences of 64 rows (x sweep)
// or 64 columns (y sweep) of one plane.
//
// Step A: r = u + i cs d2y(u) + S, solve (1 - i cs d2x) u* = r along rows.
// Step B: r = u* + i cs d2x(u*) + S, solve (1 - i cs d2y) u = r along columns.
// Missing neighbours at the grid edge are zero (the reference's edge rule).
// The neighbours across lanes come from DPP wave shifts; the two edge lanes of
// the wave read the line beyond the block themselves.  A sweep's input plane
// is read-only while the sweep runs (the blocks of a plane may run in any
// order, and every block reads its neighbours' edge lines); the forward
// values go to an FP64 plane of the caller's and come back, reversed, for the
// back substitution, which writes the output plane.
//
// Every element goes through the same operations in the same order in every
// caller (the workgroup-per-plane kernels, the wave-per-block kernels and the
// batch kernel), and floating-point contraction is off in each function body
// here, so all of them give bit-identical fields.
#pragma once

#include "kernels.hpp"

namespace g4 {
namespace adi {

template <class P>
struct A2;
template <>
struct A2<double> {
  using T = double2;
};
template <>
struct A2<float> {
  using T = float2;
};

constexpr int kLane = 64;

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
#pragma clang fp contract(off)
  return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 zsub(double2 a, double2 b) {
#pragma clang fp contract(off)
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 zadd(double2 a, double2 b) {
#pragma clang fp contract(off)
  return make_double2(a.x + b.x, a.y + b.y);
}

// rhs = center + i*cs*((a + b) - 2 center)
__device__ __forceinline__ double2 rhs_of(double2 center, double2 a, double2 b, double cs) {
#pragma clang fp contract(off)
  const double2 lap = zsub(zadd(a, b), make_double2(center.x * 2.0, center.y * 2.0));
  return make_double2(fma(-cs, lap.y, center.x), fma(cs, lap.x, center.y));
}

template <class T2>
__device__ __forceinline__ double2 to_d(T2 v) {
  return make_double2(static_cast<double>(v.x), static_cast<double>(v.y));
}

template <class P>
__device__ __forceinline__ typename A2<P>::T from_d(double2 v) {
  typename A2<P>::T o;
  o.x = static_cast<P>(v.x);
  o.y = static_cast<P>(v.y);
  return o;
}

// lane i <- lane i-1 (wave_shr:1); lane 0 keeps `edge`
__device__ __forceinline__ double2 from_lane_below(double2 v, double2 edge) {
  int r[4];
  const int* a = reinterpret_cast<const int*>(&v);
  const int* e = reinterpret_cast<const int*>(&edge);
#pragma unroll
  for (int k = 0; k < 4; ++k) r[k] = __builtin_amdgcn_update_dpp(e[k], a[k], 0x138, 0xf, 0xf, false);
  double2 o;
  __builtin_memcpy(&o, r, sizeof(o));
  return o;
}
// lane i <- lane i+1 (wave_shl:1); lane 63 keeps `edge`
__device__ __forceinline__ double2 from_lane_above(double2 v, double2 edge) {
  int r[4];
  const int* a = reinterpret_cast<const int*>(&v);
  const int* e = reinterpret_cast<const int*>(&edge);
#pragma unroll
  for (int k = 0; k < 4; ++k) r[k] = __builtin_amdgcn_update_dpp(e[k], a[k], 0x130, 0xf, 0xf, false);
  double2 o;
  __builtin_memcpy(&o, r, sizeof(o));
  return o;
}

// Thomas factors of one plane size in LDS: cl, cb, cw = coef, coef+G, coef+2G.
struct Factors {
  const double2* cl;
  const double2* cb;
  const double2* cw;
};

// Copy c, cbet, cwet ([G] each) into LDS coef[3G] with threads t0 .. of
// stride nt.  The caller synchronises before reading them.
__device__ __forceinline__ Factors stage_factors(double2* coef, const double2* c_d,
                                                 const double2* cbet_d, const double2* cwet_d,
                                                 int G, int t0, int nt) {
  for (int k = t0; k < G; k += nt) {
    coef[k] = c_d[k];
    coef[G + k] = cbet_d[k];
    coef[2 * G + k] = cwet_d[k];
  }
  return Factors{coef, coef + G, coef + 2 * G};
}

// ---- range-checked buffer access ------------------------------------------
// The sweeps address a plane through a buffer resource of the plane's size:
// an access at an offset beyond it reads zero and writes nothing.  Lanes off
// the grid, edge lanes without a neighbour line and cells outside the source
// rectangle take the offset kOut, so the chunk loops carry no divergent
// branches and the loads of the next chunk stay in flight (counted exactly by
// the wait before their use) while the current chunk's recurrence runs.
using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr uint32_t kOut = 0x80000000u;  // past any plane; offset + 16 B does not wrap
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 raw buffer, no format conversion
constexpr int kNT = 2;                  // cache policy: nontemporal

__device__ __forceinline__ Rsrc plane_rsrc(const void* p, int G, int elem_bytes) {
  const uint32_t n = p ? static_cast<uint32_t>(G) * static_cast<uint32_t>(G) * elem_bytes : 0u;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, static_cast<int>(n), kRsrcWord3);
}
// byte offset of cell (row, col), or kOut when `ok` is false
__device__ __forceinline__ uint32_t cell_off(bool ok, int row, int col, int G, int elem_bytes) {
  return ok ? (static_cast<uint32_t>(row) * G + static_cast<uint32_t>(col)) * elem_bytes : kOut;
}

template <int AUX>
__device__ __forceinline__ float2 bload(Rsrc r, uint32_t off, float2*) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX);
  float2 o;
  __builtin_memcpy(&o, &v, sizeof(o));
  return o;
}
template <int AUX>
__device__ __forceinline__ double2 bload(Rsrc r, uint32_t off, double2*) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX);
  double2 o;
  __builtin_memcpy(&o, &v, sizeof(o));
  return o;
}
template <int AUX, class T2>
__device__ __forceinline__ T2 bload(Rsrc r, uint32_t off) {
  return bload<AUX>(r, off, static_cast<T2*>(nullptr));
}
template <int AUX>
__device__ __forceinline__ void bstore(Rsrc r, uint32_t off, float2 v) {
  __attribute__((ext_vector_type(2))) int d;
  __builtin_memcpy(&d, &v, sizeof(d));
  __builtin_amdgcn_raw_buffer_store_b64(d, r, off, 0, AUX);
}
template <int AUX>
__device__ __forceinline__ void bstore(Rsrc r, uint32_t off, double2 v) {
  __attribute__((ext_vector_type(4))) int d;
  __builtin_memcpy(&d, &v, sizeof(d));
  __builtin_amdgcn_raw_buffer_store_b128(d, r, off, 0, AUX);
}

// x sweep, forward elimination of rows 64 blk + t (t = lane) of one plane:
// uin the plane before the sweep, ys the plane's FP64 forward values,
// sp the slice's source plane inside rc (nullptr: no source).
template <class P, int kCh>
__device__ __forceinline__ void x_forward(const typename A2<P>::T* uin, double2* ys, int G,
                                          double cs, Factors f, const typename A2<P>::T* sp,
                                          int4 rc, int blk, int t) {
#pragma clang fp contract(off)
  using T2 = typename A2<P>::T;
  constexpr int E = sizeof(T2);
  const int iy = blk * kLane + t;
  const bool act = iy < G;
  const int ey = t == 0 ? iy - 1 : (t == kLane - 1 ? iy + 1 : -1);
  const bool eact = ey >= 0 && ey < G;
  const bool srow = sp && act && iy >= rc.y && iy < rc.w;
  const Rsrc ru = plane_rsrc(uin, G, E);
  const Rsrc rs = plane_rsrc(sp, G, E);
  const Rsrc ry = plane_rsrc(ys, G, 16);
  // chunk loads run one chunk ahead of the recurrence
  T2 nc[kCh], ne[kCh], ns[kCh];
  auto load = [&](int k0) {
#pragma unroll
    for (int c = 0; c < kCh; ++c) {
      const int k = k0 + c;
      const bool kin = k < G;
      nc[c] = bload<0, T2>(ru, cell_off(act && kin, iy, k, G, E));
      ne[c] = bload<0, T2>(ru, cell_off(eact && kin, ey, k, G, E));
      ns[c] = bload<0, T2>(rs, cell_off(srow && kin && k >= rc.x && k < rc.z, iy, k, G, E));
    }
  };
  load(0);
  double2 yprev = make_double2(0.0, 0.0);
  for (int k0 = 0; k0 < G; k0 += kCh) {
    double2 cen[kCh], edg[kCh], sv[kCh];
#pragma unroll
    for (int c = 0; c < kCh; ++c) {
      cen[c] = to_d(nc[c]);
      edg[c] = to_d(ne[c]);
      sv[c] = to_d(ns[c]);
    }
    if (k0 + kCh < G) load(k0 + kCh);
#pragma unroll
    for (int c = 0; c < kCh; ++c) {
      const int k = k0 + c;
      if (k < G) {  // uniform
        const double2 above = from_lane_below(cen[c], edg[c]);  // row iy - 1
        const double2 below = from_lane_above(cen[c], edg[c]);  // row iy + 1
        double2 r = rhs_of(cen[c], below, above, cs);
        const bool sin_ = srow && k >= rc.x && k < rc.z;
        const double2 rs_ = zadd(r, sv[c]);
        r.x = sin_ ? rs_.x : r.x;
        r.y = sin_ ? rs_.y : r.y;
        yprev = zmul(zsub(r, zmul(f.cl[k], yprev)), f.cb[k]);
        bstore<0>(ry, cell_off(act, iy, k, G, 16), yprev);
      }
    }
  }
}

// x sweep, back substitution of the same rows, right to left, into uout.
template <class P, int kCh>
__device__ __forceinline__ void x_backward(const double2* ys, typename A2<P>::T* uout, int G,
                                           Factors f, int blk, int t) {
#pragma clang fp contract(off)
  using T2 = typename A2<P>::T;
  constexpr int E = sizeof(T2);
  const int iy = blk * kLane + t;
  const bool act = iy < G;
  const Rsrc ry = plane_rsrc(ys, G, 16);
  const Rsrc ro = plane_rsrc(uout, G, E);
  double2 unext = make_double2(0.0, 0.0);
  double2 ny[kCh];
  auto load_y = [&](int k0) {
#pragma unroll
    for (int c = 0; c < kCh; ++c) {
      const int k = k0 + c;
      ny[c] = bload<0, double2>(ry, cell_off(act && k < G, iy, k, G, 16));
    }
  };
  load_y(((G - 1) / kCh) * kCh);
  for (int k0 = ((G - 1) / kCh) * kCh; k0 >= 0; k0 -= kCh) {
    double2 y[kCh];
#pragma unroll
    for (int c = 0; c < kCh; ++c) y[c] = ny[c];
    if (k0 > 0) load_y(k0 - kCh);
#pragma unroll
    for (int c = kCh - 1; c >= 0; --c) {
      const int k = k0 + c;
      if (k < G) {  // uniform
        const double2 uu = (k == G - 1) ? y[c] : zsub(y[c], zmul(f.cw[k + 1], unext));
        unext = uu;
        bstore<0>(ro, cell_off(act, iy, k, G, E), from_d<P>(uu));
      }
    }
  }
}

// y sweep, forward elimination of columns 64 blk + t of one plane.
template <class P, int kCh>
__device__ __forceinline__ void y_forward(const typename A2<P>::T* uin, double2* ys, int G,
                                          double cs, Factors f, const typename A2<P>::T* sp,
                                          int4 rc, int blk, int t) {
#pragma clang fp contract(off)
  using T2 = typename A2<P>::T;
  constexpr int E = sizeof(T2);
  const int ix = blk * kLane + t;
  const bool act = ix < G;
  const int ex = t == 0 ? ix - 1 : (t == kLane - 1 ? ix + 1 : -1);
  const bool eact = ex >= 0 && ex < G;
  const bool scol = sp && act && ix >= rc.x && ix < rc.z;
  const Rsrc ru = plane_rsrc(uin, G, E);
  const Rsrc rs = plane_rsrc(sp, G, E);
  const Rsrc ry = plane_rsrc(ys, G, 16);
  T2 nv[kCh], ne[kCh], ns[kCh];
  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < kCh; ++r) {
      const int k = k0 + r;
      const bool kin = k < G;
      nv[r] = bload<kNT, T2>(ru, cell_off(act && kin, k, ix, G, E));
      ne[r] = bload<0, T2>(ru, cell_off(eact && kin, k, ex, G, E));
      ns[r] = bload<0, T2>(rs, cell_off(scol && kin && k >= rc.y && k < rc.w, k, ix, G, E));
    }
  };
  load(0);
  double2 yprev = make_double2(0.0, 0.0);
  for (int k0 = 0; k0 < G; k0 += kCh) {
    double2 v[kCh], e[kCh], sv[kCh];
#pragma unroll
    for (int r = 0; r < kCh; ++r) {
      v[r] = to_d(nv[r]);
      e[r] = to_d(ne[r]);
      sv[r] = to_d(ns[r]);
    }
    if (k0 + kCh < G) load(k0 + kCh);
#pragma unroll
    for (int r = 0; r < kCh; ++r) {
      const int k = k0 + r;
      if (k < G) {  // uniform
        const double2 left = from_lane_below(v[r], e[r]);   // column ix - 1
        const double2 right = from_lane_above(v[r], e[r]);  // column ix + 1
        double2 rr = rhs_of(v[r], right, left, cs);
        const bool sin_ = scol && k >= rc.y && k < rc.w;
        const double2 rs_ = zadd(rr, sv[r]);
        rr.x = sin_ ? rs_.x : rr.x;
        rr.y = sin_ ? rs_.y : rr.y;
        yprev = zmul(zsub(rr, zmul(f.cl[k], yprev)), f.cb[k]);
        bstore<kNT>(ry, cell_off(act, k, ix, G, 16), yprev);
      }
    }
  }
}

// Field sums of a plane (DiagnosticCUDA.cpp:305-363): sum |u|^2, dx |u|^2,
// dx^2 |u|^2, dy |u|^2, dy^2 |u|^2, Re u, Im u with dx = ix - (G-1)/2.
struct PlaneSums {
  double acc[7];
  __device__ void clear() {
#pragma unroll
    for (int k = 0; k < 7; ++k) acc[k] = 0.0;
  }
};

// y sweep, back substitution of the same columns, bottom to top, into uout;
// with STATS each lane adds its cells to `sums` and the lane holding the
// centre cell (G^2 - 1)/2 writes it to centre[0..1].
template <class P, int kCh, bool STATS>
__device__ __forceinline__ void y_backward(const double2* ys, typename A2<P>::T* uout, int G,
                                           Factors f, int blk, int t, PlaneSums& sums,
                                           double* centre) {
#pragma clang fp contract(off)
  using T2 = typename A2<P>::T;
  constexpr int E = sizeof(T2);
  const int ix = blk * kLane + t;
  const bool act = ix < G;
  const double shift = -0.5 * static_cast<double>(G - 1);
  const int64_t icenter = (static_cast<int64_t>(G) * G - 1) / 2;
  const int crow = static_cast<int>(icenter / G), ccol = static_cast<int>(icenter % G);
  const Rsrc ry = plane_rsrc(ys, G, 16);
  const Rsrc ro = plane_rsrc(uout, G, E);
  double2 unext = make_double2(0.0, 0.0);
  double2 cval = make_double2(0.0, 0.0);
  double2 ny[kCh];
  auto load_y = [&](int k0) {
#pragma unroll
    for (int r = 0; r < kCh; ++r) {
      const int k = k0 + r;
      ny[r] = bload<kNT, double2>(ry, cell_off(act && k < G, k, ix, G, 16));
    }
  };
  load_y(((G - 1) / kCh) * kCh);
  for (int k0 = ((G - 1) / kCh) * kCh; k0 >= 0; k0 -= kCh) {
    double2 y[kCh];
#pragma unroll
    for (int r = 0; r < kCh; ++r) y[r] = ny[r];
    if (k0 > 0) load_y(k0 - kCh);
#pragma unroll
    for (int r = kCh - 1; r >= 0; --r) {
      const int k = k0 + r;
      if (k < G) {  // uniform
        const double2 uu = (k == G - 1) ? y[r] : zsub(y[r], zmul(f.cw[k + 1], unext));
        unext = uu;
        const T2 o = from_d<P>(uu);
        bstore<kNT>(ro, cell_off(act, k, ix, G, E), o);
        if constexpr (STATS) {
          const double re = o.x, im = o.y;
          const double wei = re * re + im * im;
          const double dx = ix + shift, dy = k + shift;
          const double a[7] = {wei, dx * wei, dx * dx * wei, dy * wei, dy * dy * wei, re, im};
#pragma unroll
          for (int q = 0; q < 7; ++q) {
            const double n = sums.acc[q] + a[q];
            sums.acc[q] = act ? n : sums.acc[q];
          }
          const bool isc = k == crow && ix == ccol;
          cval.x = isc ? re : cval.x;
          cval.y = isc ? im : cval.y;
        }
      }
    }
  }
  if constexpr (STATS) {
    if (ix == ccol) {
      centre[0] = cval.x;
      centre[1] = cval.y;
    }
  }
}

// Sum of v over the 64 lanes (butterfly), the same order in every caller.
__device__ __forceinline__ double wave_sum64(double v) {
#pragma clang fp contract(off)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}


// ---- segmented sweeps -------------------------------------------------------
// The sweeps above run each line's recurrence in one lane from end to end:
// a plane of G lines is ceil(G / 64) waves, each waiting on its chain of G
// elements, and the forward values cross HBM (an FP64 plane written and read
// back).  Here a workgroup of kSegs waves owns 64 lines of a plane and wave j
// the elements [j m, j m + m) of each (m = ceil(G / kSegs) <= kSegMax), kept
// in registers from the load to the store.  Both Thomas recurrences are
// affine in the value entering a segment,
//   forward   y_k = alpha_k + beta_k  y_(s-1),   alpha: y_(s-1) = 0,
//   backward  u_k = gamma_k + delta_k u_e,       gamma: u_e = 0,
// where beta_k = prod_(i=s..k) (-c_i cbet_i) and delta_k = prod_(i=k+1..e)
// (-cwet_i) do not depend on the line (host tables, zero past the grid), so
// each wave runs its segment with nothing entering, the values entering
// every segment follow from the segments' end values in one short pass
// through LDS (j steps), and one more pass over the registers corrects them.
// The factors satisfy |c cbet|, |cwet| < 1 (the system is diagonally
// dominant), so the products decay and the corrections add no error growth;
// the result differs from the single-chain sweep by rounding only.
constexpr int kSegs = 16;    // waves per workgroup = segments per line
constexpr int kSegMax = 16;  // longest segment: G <= kSegs * kSegMax

struct SegTables {
  const double2* cl;  // c
  const double2* cb;  // cbet
  const double2* cw;  // cwet
  const double2* be;  // beta, per element of its segment
  const double2* de;  // delta, per element of its segment
};

// Segments of the workgroup's LPW lines (blk) of one plane, kSegs * LPW
// threads (thread j * LPW + l: segment j of line l): XDIR sweeps along rows
// (step A, lines are rows), else along columns (step B).  sA: LDS
// [kSegs][LPW]; with STATS the field sums of the block go to part[0..6] and
// the centre cell to centre[0..1] (red: LDS [waves][8]).
// sclear: the source plane sp again, whose cells this block read are then
// set to zero (the step's last reader of the source clears it for the next
// deposition), or nullptr.
template <class P, int LPW, bool XDIR, bool STATS>
__device__ __forceinline__ void seg_sweep(const typename A2<P>::T* uin,
                                          typename A2<P>::T* uout, int G, int m, double cs,
                                          SegTables tb, const typename A2<P>::T* sp, int4 rc,
                                          int blk, double2* sA, double* red, double* part,
                                          double* centre,
                                          typename A2<P>::T* sclear = nullptr) {
#pragma clang fp contract(off)
  using T2 = typename A2<P>::T;
  constexpr int E = sizeof(T2);
  constexpr int M = kSegMax;
  constexpr int NW = kSegs * LPW / kLane;  // waves of the workgroup
  // edge and source loads ahead (fewer in the y sweep, whose field sums and
  // source clearing hold their own registers: 128 per lane fit four waves
  // per SIMD)
  constexpr int CH = (sizeof(P) == 8 ? 2 : 4) / (XDIR ? 1 : 2);
  constexpr int AIN = XDIR ? 0 : kNT;  // the y sweep reads its lines once
  const int t = threadIdx.x & (kLane - 1);
  const int l = threadIdx.x % LPW, j = threadIdx.x / LPW;
  const int L = blk * LPW + l;
  const bool act = L < G;
  const int eL = l == 0 ? L - 1 : (l == LPW - 1 ? L + 1 : -1);
  const bool eact = eL >= 0 && eL < G;
  const bool sline =
      sp != nullptr && act && (XDIR ? (L >= rc.y && L < rc.w) : (L >= rc.x && L < rc.z));
  const int klo = XDIR ? rc.x : rc.y, khi = XDIR ? rc.z : rc.w;
  const int s = j * m;
  const int n = max(0, min(G, s + m) - s);  // elements of this segment (uniform)
  const Rsrc ru = plane_rsrc(uin, G, E);
  const Rsrc rs = plane_rsrc(sp, G, E);
  const Rsrc ro = plane_rsrc(uout, G, E);
  auto off = [&](bool ok, int line, int k) {
    return XDIR ? cell_off(ok, line, k, G, E) : cell_off(ok, k, line, G, E);
  };
  // 1. the lane's own segment, all of it in flight at once
  double2 y[M];
#pragma unroll
  for (int c = 0; c < M; ++c) y[c] = to_d(bload<AIN, T2>(ru, off(act && c < n, L, s + c)));
  // the edge lanes' neighbour lines and the source, CH elements ahead
  T2 ne[CH], ns[CH];
  auto load_es = [&](int c0) {
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      const int c = c0 + q, k = s + c;
      const bool in = c < n;
      ne[q] = bload<0, T2>(ru, off(eact && in, eL, k));
      ns[q] = bload<0, T2>(rs, off(sline && in && k >= klo && k < khi, L, k));
    }
  };
  load_es(0);
  double2 a = make_double2(0.0, 0.0);  // alpha of the previous element
#pragma unroll
  for (int c0 = 0; c0 < M; c0 += CH) {
    double2 e[CH], sv[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      e[q] = to_d(ne[q]);
      sv[q] = to_d(ns[q]);
    }
    if (c0 + CH < M) load_es(c0 + CH);
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      const int c = c0 + q, k = s + c;
      if (c < n) {  // uniform
        const double2 cen = y[c];
        double2 lo = from_lane_below(cen, e[q]);  // line L - 1
        double2 hi = from_lane_above(cen, e[q]);  // line L + 1
        if constexpr (LPW < kLane) {  // the wave holds several segments' lines
          lo = l == 0 ? e[q] : lo;
          hi = l == LPW - 1 ? e[q] : hi;
        }
        double2 r = rhs_of(cen, hi, lo, cs);
        const bool sin_ = sline && k >= klo && k < khi;
        const double2 rs_ = zadd(r, sv[q]);
        r.x = sin_ ? rs_.x : r.x;
        r.y = sin_ ? rs_.y : r.y;
        a = zmul(zsub(r, zmul(tb.cl[k], a)), tb.cb[k]);
        y[c] = a;
      }
    }
  }
  // 2. the forward value entering the segment, from the end values of the
  //    segments before it
  sA[j * LPW + l] = a;
  __syncthreads();
  double2 yin = make_double2(0.0, 0.0);
  for (int i = 0; i < j; ++i) {
    const int si = i * m, ei = min(G, si + m);
    if (ei > si) yin = zadd(sA[i * LPW + l], zmul(tb.be[ei - 1], yin));
  }
  // 3. forward values, then the back substitution inside the segment with
  //    zero entering from the right
#pragma unroll
  for (int c = 0; c < M; ++c)
    if (c < n) y[c] = zadd(y[c], zmul(tb.be[s + c], yin));
  double2 g = make_double2(0.0, 0.0);
#pragma unroll
  for (int c = M - 1; c >= 0; --c) {
    if (c < n) {
      if (c < n - 1) y[c] = zsub(y[c], zmul(tb.cw[s + c + 1], g));
      g = y[c];
    }
  }
  // 4. the value entering from the right, from the segments after it
  __syncthreads();  // every wave has read the forward end values
  sA[j * LPW + l] = g;
  __syncthreads();
  double2 uin_r = make_double2(0.0, 0.0);
  for (int i = kSegs - 1; i > j; --i) {
    const int si = i * m;
    if (si < G) uin_r = zadd(sA[i * LPW + l], zmul(tb.de[si], uin_r));
  }
  // 5. final values
  PlaneSums sums;
  sums.clear();
  const double shift = -0.5 * static_cast<double>(G - 1);
  const int64_t icenter = (static_cast<int64_t>(G) * G - 1) / 2;
  const int crow = static_cast<int>(icenter / G), ccol = static_cast<int>(icenter % G);
#pragma unroll
  for (int c = 0; c < M; ++c) {
    if (c < n) {
      const int k = s + c;
      const double2 uu = zadd(y[c], zmul(tb.de[k], uin_r));
      const T2 o = from_d<P>(uu);
      bstore<AIN>(ro, off(act, L, k), o);
      if constexpr (STATS) {
        // the lane's column L is fixed: its x moments are taken from the sum
        // of |u|^2 after the loop (acc[1], acc[2] hold the y ones meanwhile)
        const double re = o.x, im = o.y;
        const double wei = re * re + im * im;
        const double dy = k + shift;  // row k
        const double v[5] = {wei, dy * wei, dy * dy * wei, re, im};
        constexpr int slot[5] = {0, 3, 4, 5, 6};
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          const double nq = sums.acc[slot[q]] + v[q];
          sums.acc[slot[q]] = act ? nq : sums.acc[slot[q]];
        }
        if (k == crow && L == ccol) {
          centre[0] = re;
          centre[1] = im;
        }
      }
    }
  }
  if (sclear) {  // after the stores, when the segment's registers are free
    const Rsrc rw = plane_rsrc(sclear, G, E);
    T2 z;
    z.x = 0;
    z.y = 0;
#pragma unroll
    for (int c = 0; c < M; ++c) {
      const int k = s + c;
      bstore<0>(rw, off(sline && c < n && k >= klo && k < khi, L, k), z);
    }
  }
  if constexpr (STATS) {
    const double dx = L + shift;  // column L
    sums.acc[1] = dx * sums.acc[0];
    sums.acc[2] = dx * dx * sums.acc[0];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      const double v = wave_sum64(sums.acc[q]);
      if (t == 0) red[(threadIdx.x / kLane) * 8 + q] = v;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
      double v = 0.0;
      for (int i = 0; i < NW; ++i) v += red[i * 8 + threadIdx.x];
      part[threadIdx.x] = v;
    }
  }
}

}  // namespace adi
}  // namespace g4
