// field_fft.hip - spectral paraxial step (FieldSolverFFTCUDA.cpp:157-188,
// 273-323, 413-481) as two hand-written passes over each plane.
//
// The reference computes u' = IFFT2[FFT2(u) e^{i k2 delz} + 2 filt FFT2(S)]/G^2
// with three batched cuFFT transforms and four elementwise passes.  Because
// k2 = -(dx^2+dy^2) dk^2/(2 xks) the propagator factorises, e^{i k2 delz} =
// Px(mx) Py(my), and with the filter off the source term is exactly 2S, so
//   u' = Cy( Cx(u) ) + 2S,   Cx = IFFT_x Px FFT_x / G  (rows),
//                            Cy = IFFT_y Py FFT_y / G  (columns).
// The row pass holds whole rows in the registers of a wave and does
// FFT -> Px -> IFFT there; the column pass stages a tile of columns through
// LDS (coalesced row segments), transforms each column in registers, adds 2S
// inside the slice's source rectangle and, on output steps, reduces the field
// diagnostics of the tile.  Each plane crosses HBM twice per step (read+write
// per pass).
//
// Transforms: power-of-two sizes use register-resident radix-2 transforms
// (forward DIF, inverse DIT, multipliers in bit-reversed order); other sizes
// use a direct DFT in LDS.  With the source filter on (non-separable), the filtered
// source T = 2 IFFT2(filt FFT2(S))/G^2 is built in a work plane first
// (FWD rows, FWD*filt*INV columns, INV rows) and added in place of 2S.
#include "kernels.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

namespace g4 {
namespace {

constexpr int kMaxEpt = 16;  // complex elements held per thread

template <class P>
struct C2;
template <>
struct C2<double> {
  using T = double2;
};
template <>
struct C2<float> {
  using T = float2;
};

template <class T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  T2 r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <class T2>
__device__ __forceinline__ T2 cadd(T2 a, T2 b) {
  T2 r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <class T2>
__device__ __forceinline__ T2 csub(T2 a, T2 b) {
  T2 r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}
template <class T2>
__device__ __forceinline__ T2 conjv(T2 a) {
  a.y = -a.y;
  return a;
}

// Lines stored in LDS: element e of line l at l*lstride + e*estride.
struct LineLayout {
  int nlines, G, lstride, estride;
  bool line_fast;  // consecutive threads take consecutive lines (column tiles)
};

// Direct DFT of all lines, for sizes that are not powers of two.
template <class T2>
__device__ void dft_lines(T2* lds, const LineLayout& L, const T2* tw, bool inverse, int nthreads) {
  const int G = L.G;
  const int total = L.nlines * G;
  T2 v[kMaxEpt];
  int dst[kMaxEpt];
#pragma unroll
  for (int c = 0; c < kMaxEpt; ++c) {
    const int q = threadIdx.x + c * nthreads;
    if (q < total) {
      int line, k;
      if (L.line_fast) {
        line = q % L.nlines;
        k = q / L.nlines;
      } else {
        k = q % G;
        line = q / G;
      }
      const T2* b = lds + line * L.lstride;
      T2 acc;
      acc.x = 0;
      acc.y = 0;
      int idx = 0;  // (n*k) mod G
      for (int n = 0; n < G; ++n) {
        T2 w = tw[idx];
        if (inverse) w = conjv(w);
        acc = cadd(acc, cmul(b[n * L.estride], w));
        idx += k;
        if (idx >= G) idx -= G;
      }
      v[c] = acc;
      dst[c] = line * L.lstride + k * L.estride;
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < kMaxEpt; ++c) {
    const int q = threadIdx.x + c * nthreads;
    if (q < total) lds[dst[c]] = v[c];
  }
  __syncthreads();
}

enum RowMode { ROW_PROP = 0, ROW_FWD = 1, ROW_INV = 2 };
enum ColMode { COL_PROP = 0, COL_FILTER = 1 };

// Row pass for sizes that are not powers of two: `nl` rows per workgroup,
// transformed in LDS by direct DFT (natural order throughout).
template <class P>
__global__ __launch_bounds__(256) void fft_rows_dft_kernel(typename C2<P>::T* u, int64_t nrows,
                                                           int G, int nl, const double2* tw_d,
                                                           const double2* prop_d, int mode,
                                                           double scale) {
  using T2 = typename C2<P>::T;
  extern __shared__ unsigned char smem[];
  T2* tw = reinterpret_cast<T2*>(smem);
  T2* lds = tw + G;
  const int pitch = G + 1;
  const int nt = blockDim.x;
  for (int k = threadIdx.x; k < G; k += nt) {
    tw[k].x = static_cast<P>(tw_d[k].x);
    tw[k].y = static_cast<P>(tw_d[k].y);
  }
  const int64_t row0 = static_cast<int64_t>(blockIdx.x) * nl;
  const int nvalid = static_cast<int>(min<int64_t>(nl, nrows - row0));
  T2* g = u + row0 * G;
  for (int e = threadIdx.x; e < nl * G; e += nt) {
    const int l = e / G, c = e - l * G;
    T2 z;
    z.x = 0;
    z.y = 0;
    if (l < nvalid) z = g[e];
    lds[l * pitch + c] = z;
  }
  __syncthreads();
  const LineLayout L{nl, G, pitch, 1, false};
  if (mode == ROW_PROP || mode == ROW_FWD) dft_lines(lds, L, tw, false, nt);
  if (mode == ROW_PROP) {
    for (int e = threadIdx.x; e < nl * G; e += nt) {
      const int l = e / G, c = e - l * G;
      T2 pr;
      pr.x = static_cast<P>(prop_d[c].x);
      pr.y = static_cast<P>(prop_d[c].y);
      lds[l * pitch + c] = cmul(lds[l * pitch + c], pr);
    }
    __syncthreads();
  }
  if (mode == ROW_PROP || mode == ROW_INV) dft_lines(lds, L, tw, true, nt);
  const P sc = static_cast<P>(scale);
  for (int e = threadIdx.x; e < nvalid * G; e += nt) {
    const int l = e / G, c = e - l * G;
    T2 v = lds[l * pitch + c];
    v.x *= sc;
    v.y *= sc;
    g[e] = v;
  }
}

// Column pass for sizes that are not powers of two: tile of tc columns x G
// rows of one ring plane, direct DFT in LDS.
template <class P>
__global__ __launch_bounds__(512) void fft_cols_dft_kernel(
    typename C2<P>::T* u, int nslice, int G, int tc, const double2* tw_d,
    const double2* prop_d, const double* sigmoid, int mode, double scale,
    const typename C2<P>::T* src, const int4* rect, int first, int add_src,
    const double2* work, double* partials) {
  using T2 = typename C2<P>::T;
  extern __shared__ unsigned char smem[];
  T2* tw = reinterpret_cast<T2*>(smem);
  T2* lds = tw + G;
  __shared__ double red[16][8];
  const int pitch = tc + 1;  // row pitch (columns are the lines)
  const int nt = blockDim.x;
  const int ntile = (G + tc - 1) / tc;
  const int p = blockIdx.x / ntile;
  const int tile = blockIdx.x - p * ntile;
  const int c0 = tile * tc;
  const int ncol = min(tc, G - c0);
  for (int k = threadIdx.x; k < G; k += nt) {
    tw[k].x = static_cast<P>(tw_d[k].x);
    tw[k].y = static_cast<P>(tw_d[k].y);
  }
  T2* plane = u + static_cast<int64_t>(p) * G * G;
  for (int e = threadIdx.x; e < tc * G; e += nt) {
    const int r = e / tc, c = e - r * tc;
    T2 z;
    z.x = 0;
    z.y = 0;
    if (c < ncol) z = plane[static_cast<int64_t>(r) * G + c0 + c];
    lds[r * pitch + c] = z;
  }
  __syncthreads();
  const LineLayout L{tc, G, 1, pitch, true};
  dft_lines(lds, L, tw, false, nt);
  for (int e = threadIdx.x; e < tc * G; e += nt) {
    const int r = e / tc, c = e - r * tc;
    T2 f;
    if (mode == COL_PROP) {
      f.x = static_cast<P>(prop_d[r].x);
      f.y = static_cast<P>(prop_d[r].y);
    } else {
      const double sg = (c < ncol) ? sigmoid[static_cast<int64_t>(r) * G + c0 + c] : 0.0;
      f.x = static_cast<P>(sg);
      f.y = 0;
    }
    lds[r * pitch + c] = cmul(lds[r * pitch + c], f);
  }
  __syncthreads();
  dft_lines(lds, L, tw, true, nt);

  const P sc = static_cast<P>(scale);
  const int s = ((p - first) % nslice + nslice) % nslice;  // beam slice of this ring plane
  int4 rc = make_int4(0, 0, 0, 0);
  if (add_src == 1) rc = rect[s];
  const T2* sp = (add_src == 1) ? src + static_cast<int64_t>(s) * G * G : nullptr;
  const double2* wp = (add_src == 2) ? work + static_cast<int64_t>(p) * G * G : nullptr;
  const double shift = -0.5 * static_cast<double>(G - 1);
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const int64_t icenter = (static_cast<int64_t>(G) * G - 1) / 2;
  for (int e = threadIdx.x; e < ncol * G; e += nt) {
    const int r = e / ncol, c = e - r * ncol;
    const int ix = c0 + c;
    T2 v = lds[r * pitch + c];
    v.x *= sc;
    v.y *= sc;
    if (sp && ix >= rc.x && ix < rc.z && r >= rc.y && r < rc.w) {
      const T2 sv = sp[static_cast<int64_t>(r) * G + ix];
      v.x += P(2) * sv.x;
      v.y += P(2) * sv.y;
    } else if (wp) {
      const double2 wv = wp[static_cast<int64_t>(r) * G + ix];
      v.x += static_cast<P>(wv.x);
      v.y += static_cast<P>(wv.y);
    }
    const int64_t gi = static_cast<int64_t>(r) * G + ix;
    plane[gi] = v;
    if (partials) {
      const double re = v.x, im = v.y;
      const double wei = re * re + im * im;
      const double dx = ix + shift, dy = r + shift;
      acc[0] += wei;
      acc[1] += dx * wei;
      acc[2] += dx * dx * wei;
      acc[3] += dy * wei;
      acc[4] += dy * dy * wei;
      acc[5] += re;
      acc[6] += im;
      if (gi == icenter) {
        partials[static_cast<int64_t>(nslice) * ntile * 8 + 2 * p] = re;
        partials[static_cast<int64_t>(nslice) * ntile * 8 + 2 * p + 1] = im;
      }
    }
  }
  if (partials) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      double v = acc[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
      double v = 0.0;
      for (int w = 0; w < nt / 64; ++w) v += red[w][threadIdx.x];
      partials[(static_cast<int64_t>(p) * ntile + tile) * 8 + threadIdx.x] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Register-resident transforms for power-of-two G (2..1024).
//
// A line of G points is spread over L = G/R lanes of a wave, R = max(1, G/64)
// points per lane: lane t holds points e = t + L*r, r < R.  Butterfly spans
// s >= L pair registers of one lane; spans s < L pair lanes t and t^s and cost
// one cross-lane shuffle per value.  The forward transform is decimation in
// frequency (natural order in, bit-reversed out) and the inverse decimation
// in time (bit-reversed in, natural out), so FFT -> multiply -> IFFT needs no
// reordering: the multiplier is applied in bit-reversed order.  No LDS and no
// barrier is needed for the transform itself.
// ---------------------------------------------------------------------------
template <class T2>
__device__ __forceinline__ T2 shfl_xor2(T2 v, int m) {
  T2 r;
  r.x = __shfl_xor(v.x, m, 64);
  r.y = __shfl_xor(v.y, m, 64);
  return r;
}

template <int G>
struct RegShape {
  static constexpr int R = G >= 64 ? G / 64 : 1;  // points per lane
  static constexpr int L = G / R;                 // lanes per line
  static constexpr int LPW = 64 / L;              // lines per wave
};

// In-register butterflies of span s = RS*L (pairs r, r+RS of one lane).
template <class T2, int G, int RS, bool INV>
__device__ __forceinline__ void reg_stage(T2 (&v)[RegShape<G>::R], int t, const T2* tw) {
  constexpr int R = RegShape<G>::R, L = RegShape<G>::L;
  constexpr int s = RS * L;
  constexpr int tstep = G / (2 * s);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if ((r & RS) == 0) {
      const int k = (t + L * r) & (2 * s - 1);
      if (!INV) {
        const T2 a = v[r], b = v[r + RS];
        v[r] = cadd(a, b);
        v[r + RS] = cmul(csub(a, b), tw[k * tstep]);
      } else {
        const T2 a = v[r], b = cmul(v[r + RS], conjv(tw[k * tstep]));
        v[r] = cadd(a, b);
        v[r + RS] = csub(a, b);
      }
    }
  }
}

// Cross-lane butterflies of span S < L (lanes t and t^S).
template <class T2, int G, int S, bool INV>
__device__ __forceinline__ void lane_stage(T2 (&v)[RegShape<G>::R], int t, const T2* tw) {
  constexpr int R = RegShape<G>::R;
  constexpr int tstep = G / (2 * S);
  const bool lower = (t & S) == 0;
  T2 w = tw[(t & (S - 1)) * tstep];
  if (INV) w = conjv(w);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (!INV) {
      const T2 q = shfl_xor2(v[r], S);
      v[r] = lower ? cadd(v[r], q) : cmul(csub(q, v[r]), w);
    } else {
      const T2 tv = lower ? v[r] : cmul(v[r], w);
      const T2 q = shfl_xor2(tv, S);
      v[r] = lower ? cadd(tv, q) : csub(q, tv);
    }
  }
}

template <class T2, int G, int RS>
__device__ __forceinline__ void reg_stages_fwd(T2 (&v)[RegShape<G>::R], int t, const T2* tw) {
  if constexpr (RS >= 1) {
    reg_stage<T2, G, RS, false>(v, t, tw);
    reg_stages_fwd<T2, G, RS / 2>(v, t, tw);
  }
}
template <class T2, int G, int RS>
__device__ __forceinline__ void reg_stages_inv(T2 (&v)[RegShape<G>::R], int t, const T2* tw) {
  if constexpr (RS >= 1 && RS < RegShape<G>::R) {
    reg_stage<T2, G, RS, true>(v, t, tw);
    reg_stages_inv<T2, G, RS * 2>(v, t, tw);
  }
}
template <class T2, int G, int S>
__device__ __forceinline__ void lane_stages_fwd(T2 (&v)[RegShape<G>::R], int t, const T2* tw) {
  if constexpr (S >= 1) {
    lane_stage<T2, G, S, false>(v, t, tw);
    lane_stages_fwd<T2, G, S / 2>(v, t, tw);
  }
}
template <class T2, int G, int S>
__device__ __forceinline__ void lane_stages_inv(T2 (&v)[RegShape<G>::R], int t, const T2* tw) {
  if constexpr (S >= 1 && S < RegShape<G>::L) {
    lane_stage<T2, G, S, true>(v, t, tw);
    lane_stages_inv<T2, G, S * 2>(v, t, tw);
  }
}

// Forward DIF, spans G/2 .. 1: natural order in, bit-reversed order out.
template <class T2, int G>
__device__ __forceinline__ void reg_fft_fwd(T2 (&v)[RegShape<G>::R], int t, const T2* tw) {
  reg_stages_fwd<T2, G, RegShape<G>::R / 2>(v, t, tw);
  lane_stages_fwd<T2, G, RegShape<G>::L / 2>(v, t, tw);
}

// Inverse DIT with conjugate twiddles, spans 1 .. G/2 (unscaled):
// bit-reversed order in, natural order out.
template <class T2, int G>
__device__ __forceinline__ void reg_fft_inv(T2 (&v)[RegShape<G>::R], int t, const T2* tw) {
  lane_stages_inv<T2, G, 1>(v, t, tw);
  reg_stages_inv<T2, G, 1>(v, t, tw);
}

// Loads G/2 twiddles and, when wanted, the G bit-reversed multipliers to LDS.
template <class T2>
__device__ __forceinline__ void load_tables(T2* tw, T2* mu, int G, const double2* tw_d,
                                            const double2* mult_d, bool want_mult) {
  for (int k = threadIdx.x; k < G / 2; k += blockDim.x) {
    tw[k].x = static_cast<decltype(tw[k].x)>(tw_d[k].x);
    tw[k].y = static_cast<decltype(tw[k].x)>(tw_d[k].y);
  }
  if (want_mult)
    for (int k = threadIdx.x; k < G; k += blockDim.x) {
      mu[k].x = static_cast<decltype(tw[k].x)>(mult_d[k].x);
      mu[k].y = static_cast<decltype(tw[k].x)>(mult_d[k].y);
    }
}

// Row pass: each wave transforms LPW whole rows; waves stride over the rows.
template <class P, int G>
__global__ __launch_bounds__(256) void fft_rows_reg_kernel(typename C2<P>::T* u, int64_t nrows,
                                                           const double2* tw_d,
                                                           const double2* mult_d, int mode,
                                                           double scale) {
  using T2 = typename C2<P>::T;
  constexpr int R = RegShape<G>::R, L = RegShape<G>::L, LPW = RegShape<G>::LPW;
  __shared__ T2 tw[G / 2 > 0 ? G / 2 : 1];
  __shared__ T2 mu[G];
  load_tables(tw, mu, G, tw_d, mult_d, mode == ROW_PROP);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int t = lane % L;
  const int64_t wave = blockIdx.x * static_cast<int64_t>(blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (blockDim.x >> 6);
  const P sc = static_cast<P>(scale);
  for (int64_t grp = wave; grp * LPW < nrows; grp += nwaves) {
    const int64_t line = grp * LPW + lane / L;
    if (line >= nrows) continue;  // whole lane groups drop out together
    T2* g = u + line * G;
    T2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = g[t + L * r];
    if (mode == ROW_PROP || mode == ROW_FWD) reg_fft_fwd<T2, G>(v, t, tw);
    if (mode == ROW_PROP) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = cmul(v[r], mu[t + L * r]);
    }
    if (mode == ROW_PROP || mode == ROW_INV) reg_fft_inv<T2, G>(v, t, tw);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      T2 o = v[r];
      o.x *= sc;
      o.y *= sc;
      g[t + L * r] = o;
    }
  }
}

// Column pass: a tile of tc columns x G rows of one ring plane is staged in
// LDS with coalesced row segments; each wave transforms whole columns in
// registers; the tile is stored back with the source added and the field
// sums of the tile reduced.
template <class P, int G>
__global__ __launch_bounds__(256) void fft_cols_reg_kernel(
    typename C2<P>::T* u, int nslice, int tc, const double2* tw_d, const double2* mult_d,
    const double* sigmoid, int mode, double scale, const typename C2<P>::T* src, const int4* rect,
    int first, int add_src, const double2* work, double* partials) {
  using T2 = typename C2<P>::T;
  constexpr int R = RegShape<G>::R, L = RegShape<G>::L, LPW = RegShape<G>::LPW;
  extern __shared__ unsigned char smem[];
  T2* tw = reinterpret_cast<T2*>(smem);
  T2* mu = tw + (G / 2 > 0 ? G / 2 : 1);
  T2* lds = mu + G;
  __shared__ double red[4][8];
  const int pitch = tc + 1;
  const int nt = blockDim.x;
  const int ntile = (G + tc - 1) / tc;
  const int p = blockIdx.x / ntile;
  const int tile = blockIdx.x - p * ntile;
  const int c0 = tile * tc;
  const int ncol = min(tc, G - c0);
  load_tables(tw, mu, G, tw_d, mult_d, mode == COL_PROP);
  T2* plane = u + static_cast<int64_t>(p) * G * G;
  for (int e = threadIdx.x; e < tc * G; e += nt) {
    const int r = e / tc, c = e - r * tc;
    T2 z;
    z.x = 0;
    z.y = 0;
    if (c < ncol) z = plane[static_cast<int64_t>(r) * G + c0 + c];
    lds[r * pitch + c] = z;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int t = lane % L;
  const int wave = threadIdx.x >> 6, nwave = nt >> 6;
  int lg = 0;
  while ((1 << lg) < G) ++lg;
  for (int cg = wave; cg * LPW < ncol; cg += nwave) {
    const int c = cg * LPW + lane / L;
    if (c >= ncol) continue;
    T2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = lds[(t + L * r) * pitch + c];
    reg_fft_fwd<T2, G>(v, t, tw);
    if (mode == COL_PROP) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = cmul(v[r], mu[t + L * r]);
    } else {
      // filter: position ey holds my = bitrev(ey), column c holds mx = bitrev(c0+c)
      const int mx = static_cast<int>(__brev(static_cast<unsigned>(c0 + c)) >> (32 - lg));
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int my = static_cast<int>(__brev(static_cast<unsigned>(t + L * r)) >> (32 - lg));
        const P f = static_cast<P>(sigmoid[static_cast<int64_t>(my) * G + mx]);
        v[r].x *= f;
        v[r].y *= f;
      }
    }
    reg_fft_inv<T2, G>(v, t, tw);
#pragma unroll
    for (int r = 0; r < R; ++r) lds[(t + L * r) * pitch + c] = v[r];
  }
  __syncthreads();

  const P sc = static_cast<P>(scale);
  const int s = ((p - first) % nslice + nslice) % nslice;  // beam slice of this ring plane
  int4 rc = make_int4(0, 0, 0, 0);
  if (add_src == 1) rc = rect[s];
  const T2* sp = (add_src == 1) ? src + static_cast<int64_t>(s) * G * G : nullptr;
  const double2* wp = (add_src == 2) ? work + static_cast<int64_t>(p) * G * G : nullptr;
  const double shift = -0.5 * static_cast<double>(G - 1);
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const int64_t icenter = (static_cast<int64_t>(G) * G - 1) / 2;
  for (int e = threadIdx.x; e < ncol * G; e += nt) {
    const int r = e / ncol, c = e - r * ncol;
    const int ix = c0 + c;
    T2 v = lds[r * pitch + c];
    v.x *= sc;
    v.y *= sc;
    if (sp && ix >= rc.x && ix < rc.z && r >= rc.y && r < rc.w) {
      const T2 sv = sp[static_cast<int64_t>(r) * G + ix];
      v.x += P(2) * sv.x;
      v.y += P(2) * sv.y;
    } else if (wp) {
      const double2 wv = wp[static_cast<int64_t>(r) * G + ix];
      v.x += static_cast<P>(wv.x);
      v.y += static_cast<P>(wv.y);
    }
    const int64_t gi = static_cast<int64_t>(r) * G + ix;
    plane[gi] = v;
    if (partials) {
      const double re = v.x, im = v.y;
      const double wei = re * re + im * im;
      const double dx = ix + shift, dy = r + shift;
      acc[0] += wei;
      acc[1] += dx * wei;
      acc[2] += dx * dx * wei;
      acc[3] += dy * wei;
      acc[4] += dy * dy * wei;
      acc[5] += re;
      acc[6] += im;
      if (gi == icenter) {
        partials[static_cast<int64_t>(nslice) * ntile * 8 + 2 * p] = re;
        partials[static_cast<int64_t>(nslice) * ntile * 8 + 2 * p + 1] = im;
      }
    }
  }
  if (partials) {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      double v = acc[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
      double v = 0.0;
      for (int w = 0; w < nwave; ++w) v += red[w][threadIdx.x];
      partials[(static_cast<int64_t>(p) * ntile + tile) * 8 + threadIdx.x] = v;
    }
  }
}

// partials -> plane_stats[p*9 + k] (k < 7 sums, 7..8 centre)
__global__ void reduce_partials_kernel(const double* partials, int nslice, int ntile,
                                       double* plane_stats) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nslice) return;
  double* o = plane_stats + static_cast<int64_t>(p) * G4H_FIELD_RAW;
  for (int k = 0; k < 7; ++k) {
    double v = 0.0;
    for (int t = 0; t < ntile; ++t) v += partials[(static_cast<int64_t>(p) * ntile + t) * 8 + k];
    o[k] = v;
  }
  o[7] = partials[static_cast<int64_t>(nslice) * ntile * 8 + 2 * p];
  o[8] = partials[static_cast<int64_t>(nslice) * ntile * 8 + 2 * p + 1];
}

// Gather the (rectangle-bounded) source into ring-ordered dense work planes.
template <class P>
__global__ void gather_source_kernel(const typename C2<P>::T* src, const int4* rect, int nslice,
                                     int G, int first, double2* work) {
  const int p = blockIdx.y;
  const int s = ((p - first) % nslice + nslice) % nslice;
  const int4 rc = rect[s];
  const int64_t plane = static_cast<int64_t>(G) * G;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < plane;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int iy = static_cast<int>(e / G), ix = static_cast<int>(e - static_cast<int64_t>(iy) * G);
    double2 v = make_double2(0.0, 0.0);
    if (ix >= rc.x && ix < rc.z && iy >= rc.y && iy < rc.w) {
      const auto sv = src[s * plane + e];
      v = make_double2(sv.x, sv.y);
    }
    work[p * plane + e] = v;
  }
}

bool is_pow2(int g) { return g >= 2 && (g & (g - 1)) == 0; }

// Launch shape of the DFT kernels (sizes that are not powers of two).
struct DftCfg {
  int row_lines, tc, col_threads;
};

DftCfg dft_cfg(int G, size_t esz) {
  DftCfg c;
  c.row_lines = std::max(1, 2048 / G);
  while (c.row_lines * G > 256 * kMaxEpt) --c.row_lines;
  while (c.row_lines > 1 && (static_cast<size_t>(c.row_lines) * (G + 1) + G) * esz > 60 * 1024)
    --c.row_lines;
  c.tc = esz == 8 ? 16 : 8;
  while (c.tc > 1 && (static_cast<size_t>(G) * (c.tc + 1) + G) * esz > 60 * 1024) c.tc /= 2;
  c.col_threads = 256;
  while (c.tc * G > c.col_threads * kMaxEpt && c.col_threads < 512) c.col_threads *= 2;
  while (c.tc * G > c.col_threads * kMaxEpt && c.tc > 1) c.tc /= 2;
  return c;
}

// Column tile width of the register kernels: 16 (float2) or 8 (double2)
// columns, fewer when the tile would not fit 60 KiB of LDS.
int reg_tc(int G, size_t esz) {
  int tc = esz == 8 ? 16 : 8;
  while (tc > 1 && (static_cast<size_t>(G) * (tc + 1) + G + G / 2) * esz > 60 * 1024) tc /= 2;
  return std::min(tc, G);
}

int col_tile(int G, size_t esz) { return is_pow2(G) ? reg_tc(G, esz) : dft_cfg(G, esz).tc; }

template <class P, int G>
void rows_reg(Field* f, typename C2<P>::T* data, int mode, double scale) {
  constexpr int LPW = RegShape<G>::LPW;
  const int64_t nrows = static_cast<int64_t>(f->nslice) * G;
  const int64_t ngroups = (nrows + LPW - 1) / LPW;
  const int64_t nblk = std::min<int64_t>((ngroups + 3) / 4, 8192);
  hipLaunchKernelGGL((fft_rows_reg_kernel<P, G>), dim3(static_cast<unsigned>(nblk)), dim3(256), 0,
                     f->ctx->stream, data, nrows, f->fft_tw.get(), f->fft_prop.get(), mode, scale);
}

template <class P, int G>
void cols_reg(Field* f, typename C2<P>::T* data, int mode, double scale, int add_src,
              double* partials) {
  using T2 = typename C2<P>::T;
  const int tc = reg_tc(G, sizeof(T2));
  const int ntile = (G + tc - 1) / tc;
  const size_t lds = sizeof(T2) * (std::max(G / 2, 1) + G + static_cast<size_t>(G) * (tc + 1));
  hipLaunchKernelGGL((fft_cols_reg_kernel<P, G>), dim3(f->nslice * ntile), dim3(256), lds,
                     f->ctx->stream, data, f->nslice, tc, f->fft_tw.get(), f->fft_prop.get(),
                     f->fft_sigmoid.get(), mode, scale, static_cast<const T2*>(f->src()),
                     f->src_rect.get(), f->first, add_src, f->fft_work.get(), partials);
}

// Calls fn with the compile-time size G (powers of two 2..1024).
template <class Fn>
void with_pow2(int G, Fn&& fn) {
  switch (G) {
    case 2: fn(std::integral_constant<int, 2>{}); break;
    case 4: fn(std::integral_constant<int, 4>{}); break;
    case 8: fn(std::integral_constant<int, 8>{}); break;
    case 16: fn(std::integral_constant<int, 16>{}); break;
    case 32: fn(std::integral_constant<int, 32>{}); break;
    case 64: fn(std::integral_constant<int, 64>{}); break;
    case 128: fn(std::integral_constant<int, 128>{}); break;
    case 256: fn(std::integral_constant<int, 256>{}); break;
    case 512: fn(std::integral_constant<int, 512>{}); break;
    case 1024: fn(std::integral_constant<int, 1024>{}); break;
    default: throw Error(G4H_ERR_UNSUPPORTED, "FFT size is not a power of two in [2, 1024]");
  }
}

template <class P>
void run_rows(Field* f, typename C2<P>::T* data, int mode, double scale) {
  using T2 = typename C2<P>::T;
  const int G = f->ngrid;
  f->ctx->launch(G4H_K_FFT_ROWS, [&] {
    if (is_pow2(G)) {
      with_pow2(G, [&](auto g) { rows_reg<P, decltype(g)::value>(f, data, mode, scale); });
      return;
    }
    const DftCfg c = dft_cfg(G, sizeof(T2));
    const int64_t nrows = static_cast<int64_t>(f->nslice) * G;
    const int64_t nblk = (nrows + c.row_lines - 1) / c.row_lines;
    const size_t lds = sizeof(T2) * (G + static_cast<size_t>(c.row_lines) * (G + 1));
    hipLaunchKernelGGL(fft_rows_dft_kernel<P>, dim3(static_cast<unsigned>(nblk)), dim3(256), lds,
                       f->ctx->stream, data, nrows, G, c.row_lines, f->fft_tw.get(),
                       f->fft_prop.get(), mode, scale);
  });
}

template <class P>
void run_cols(Field* f, typename C2<P>::T* data, int mode, double scale, int add_src,
              double* partials) {
  using T2 = typename C2<P>::T;
  const int G = f->ngrid;
  f->ctx->launch(G4H_K_FFT_COLS, [&] {
    if (is_pow2(G)) {
      with_pow2(G, [&](auto g) {
        cols_reg<P, decltype(g)::value>(f, data, mode, scale, add_src, partials);
      });
      return;
    }
    const DftCfg c = dft_cfg(G, sizeof(T2));
    const int ntile = (G + c.tc - 1) / c.tc;
    const size_t lds = sizeof(T2) * (G + static_cast<size_t>(G) * (c.tc + 1));
    hipLaunchKernelGGL(fft_cols_dft_kernel<P>, dim3(f->nslice * ntile), dim3(c.col_threads), lds,
                       f->ctx->stream, data, f->nslice, G, c.tc, f->fft_tw.get(),
                       f->fft_prop.get(), f->fft_sigmoid.get(), mode, scale,
                       static_cast<const T2*>(f->src()), f->src_rect.get(), f->first, add_src,
                       f->fft_work.get(), partials);
  });
}

template <class P>
void fft_step(Field* f, bool source_on, bool stats) {
  using T2 = typename C2<P>::T;
  const int G = f->ngrid;
  T2* data = static_cast<T2*>(f->data());
  const double invg = 1.0 / static_cast<double>(G);
  const int ntile = (G + col_tile(G, sizeof(T2)) - 1) / col_tile(G, sizeof(T2));
  double* partials = nullptr;
  if (stats) {
    const size_t need = static_cast<size_t>(f->nslice) * ntile * 8 + 2 * static_cast<size_t>(f->nslice);
    if (f->col_partials.size() < need) f->col_partials.alloc(f->ctx, need);
    partials = f->col_partials.get();
  }
  int add = source_on ? 1 : 0;
  if (source_on && f->filter) {
    // T = 2 IFFT2(filt FFT2(S)) / G^2 into fft_work (ring order)
    const int64_t cells = f->cells();
    if (f->fft_work.size() < static_cast<size_t>(cells)) f->fft_work.alloc(f->ctx, cells);
    f->ctx->launch(G4H_K_OTHER, [&] {
      hipLaunchKernelGGL(gather_source_kernel<P>, dim3(64, f->nslice), dim3(256), 0, f->ctx->stream,
                         static_cast<const T2*>(f->src()), f->src_rect.get(), f->nslice, G, f->first,
                         f->fft_work.get());
    });
    run_rows<double>(f, f->fft_work.get(), ROW_FWD, 1.0);
    run_cols<double>(f, f->fft_work.get(), COL_FILTER, 1.0, 0, nullptr);
    run_rows<double>(f, f->fft_work.get(), ROW_INV, 2.0 * invg * invg);
    add = 2;
  }
  run_rows<P>(f, data, ROW_PROP, invg);
  run_cols<P>(f, data, COL_PROP, invg, add, partials);
  if (stats) {
    f->ctx->launch(G4H_K_OTHER, [&] {
      hipLaunchKernelGGL(reduce_partials_kernel, dim3((f->nslice + 255) / 256), dim3(256), 0,
                         f->ctx->stream, partials, f->nslice, ntile, f->plane_stats.get());
    });
  }
}

}  // namespace

void fft_prepare(Field* f, double delz) {
  const int G = f->ngrid;
  if (f->fft_tw.size() != static_cast<size_t>(G)) {
    std::vector<double2> tw(G);
    const double pi2 = 4.0 * std::asin(1.0);
    for (int k = 0; k < G; ++k) {
      const double a = -pi2 * static_cast<double>(k) / static_cast<double>(G);
      tw[k] = make_double2(std::cos(a), std::sin(a));
    }
    f->fft_tw.alloc(f->ctx, G);
    G4_HIP(hipMemcpy(f->fft_tw.get(), tw.data(), sizeof(double2) * G, hipMemcpyHostToDevice));
    f->fft_delz = 0.0;
    f->fft_prop.alloc(f->ctx, G);
    if (f->filter) {
      // sigmoid filter at shifted index, FieldSolverFFTCUDA.cpp:297-316
      const double shift = -0.5 * static_cast<double>(G - 1);
      std::vector<double> sg(static_cast<size_t>(G) * G);
      for (int iy = 0; iy < G; ++iy) {
        const double dy = iy + shift;
        const double yv = dy / static_cast<double>(G) / f->yc;
        for (int ix = 0; ix < G; ++ix) {
          const double dx = ix + shift;
          const double xv = dx / static_cast<double>(G) / f->xc;
          const int iiy = (iy + (G + 1) / 2) % G, iix = (ix + (G + 1) / 2) % G;
          sg[static_cast<size_t>(iiy) * G + iix] =
              1. / (1 + std::exp((std::sqrt(xv * xv + yv * yv) - 1) / f->sig));
        }
      }
      f->fft_sigmoid.alloc(f->ctx, sg.size());
      G4_HIP(hipMemcpy(f->fft_sigmoid.get(), sg.data(), sizeof(double) * sg.size(),
                       hipMemcpyHostToDevice));
    }
  }
  if (delz != f->fft_delz) {
    // k2 at frequency index m: dx(m) = ((m - (G+1)/2) mod G) - (G-1)/2 and
    // k2 = -(dx^2 + dy^2) dk^2/(2 xks) (FieldSolverFFTCUDA.cpp:297-309);
    // the x (or y) factor is exp(-i dx^2 dk^2 delz / (2 xks)).
    // Power-of-two sizes use the register transforms, whose spectrum is in
    // bit-reversed order: entry e then holds the factor of frequency bitrev(e).
    const double dk = 4. * std::asin(1.) / (static_cast<double>(G) * f->dgrid);
    int lg = 0;
    while ((1 << lg) < G) ++lg;
    std::vector<double2> pr(G);
    for (int e = 0; e < G; ++e) {
      int m = e;
      if (is_pow2(G)) {
        m = 0;
        for (int b = 0; b < lg; ++b) m |= ((e >> b) & 1) << (lg - 1 - b);
      }
      const int ix = ((m - (G + 1) / 2) % G + G) % G;
      const double dx = static_cast<double>(ix) - 0.5 * static_cast<double>(G - 1);
      const double ph = -dx * dx * dk * dk / 2. / f->xks * delz;
      pr[e] = make_double2(std::cos(ph), std::sin(ph));
    }
    G4_HIP(hipMemcpyAsync(f->fft_prop.get(), pr.data(), sizeof(double2) * G, hipMemcpyHostToDevice,
                          f->ctx->stream));
    G4_HIP(hipStreamSynchronize(f->ctx->stream));
    f->fft_delz = delz;
  }
}

void fft_solve(Field* f, bool source_on, bool stats) {
  if (f->prec == G4H_FP64)
    fft_step<double>(f, source_on, stats);
  else
    fft_step<float>(f, source_on, stats);
}

bool fft_supported(int G) { return G >= 2 && G <= 1024; }

}  // namespace g4
