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
// FFT -> Px -> IFFT there; the column pass loads a tile of columns as
// coalesced row segments, transforms each column in registers (its lanes in
// several waves, transposed through LDS), adds 2S
// inside the slice's source rectangle and, on output steps, reduces the field
// diagnostics of the tile.  Each plane crosses HBM twice per step (read+write
// per pass).
//
// Transforms: power-of-two sizes use register-resident radix-2 transforms
// (forward DIF, inverse DIT, multipliers in bit-reversed order); other sizes
// use a direct DFT in LDS.  With the source filter on (non-separable), the filtered
// source T = 2 IFFT2(filt FFT2(S))/G^2 is built in a work plane first
// (FWD rows, FWD*filt*INV columns, INV rows) and added in place of 2S.
#include "fft_regs.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

namespace g4 {
namespace {

constexpr int kMaxEpt = 16;  // complex elements held per thread

// v of the lane that DPP control CTRL selects (both 32-bit halves moved)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}
__device__ __forceinline__ double lane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}
// Sum of v over the 64 lanes of the wave without LDS: within each row of 16
// lanes by DPP (lane xor 1, xor 2, then the mirrored lane of the 8 and of the
// 16, each step pairing lanes whose partial sums cover disjoint halves),
// then the four row sums read from lanes 0, 16, 32, 48.
__device__ __forceinline__ double wave_total(double v) {
  v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x141>(v);  // row_half_mirror
  v += dpp_f64<0x140>(v);  // row_mirror
  return (lane_f64(v, 0) + lane_f64(v, 16)) + (lane_f64(v, 32) + lane_f64(v, 48));
}

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
// Power-of-two sizes: register-resident four-step transforms (fft_regs.hpp).
// ---------------------------------------------------------------------------
// Waves per workgroup: as many as fit the transpose buffers in 48 KiB (FP32)
// or 72 KiB (FP64: at G = 256 four waves instead of two, and 256-byte
// instead of 128-byte tile rows in the column pass, measured 15 % faster
// there although a CU then holds two workgroups instead of four), at most 4.
template <class P, int G>
struct RegCfg {
  using S = fftr::Shape<G>;
  static constexpr int kWaveBytes = S::LPW * S::BUF * 2 * static_cast<int>(sizeof(P));
  static constexpr int kLds = sizeof(P) == 8 ? 72 * 1024 : 48 * 1024;
  static constexpr int kWaves = kWaveBytes * 4 <= kLds ? 4 : (kWaveBytes * 2 <= kLds ? 2 : 1);
  static constexpr int kThreads = 64 * kWaves;
  // column tile: one column per line slot of the workgroup
  static constexpr int kTile = kWaves * S::LPW < G ? kWaves * S::LPW : G;
  // Row-segment tile of the column pass, when every line slot holds a
  // column (kTile = kThreads / L) and a tile row is at least 64 bytes:
  // thread (t = tid / kTile, column tid % kTile) loads and stores the rows
  // t + L r of its column, so each load or store of a wave covers
  // 64 / kTile whole row segments of the tile, and the registers are already
  // in the natural layout of the line transforms.  A column's L lanes then sit in different
  // waves: its transposes go through LDS buffers interleaved by column (the
  // lanes of one row segment hit adjacent words) with workgroup barriers.
  static constexpr bool kDirect = S::L > 1 && kWaves * S::LPW <= G &&
                                  kTile * 2 * static_cast<int>(sizeof(P)) >= 64;
  static constexpr int kBuf = kWaves * S::LPW * S::BUF;
  // FP64 tiles at G = 128 and 256 transpose real and imaginary parts in
  // turn through half the buffer (fftr::xpose), so that with at most 168
  // registers a CU holds three workgroups instead of two.
  static constexpr bool kSplit = kDirect && sizeof(P) == 8 && G >= 128 && G <= 256;
  static constexpr int kMinWaves = kSplit ? 3 : 1;  // per SIMD
};

template <class T2>
__device__ __forceinline__ void load_table(T2* dst, const double2* src, int n) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    dst[k].x = static_cast<decltype(dst[k].x)>(src[k].x);
    dst[k].y = static_cast<decltype(dst[k].x)>(src[k].y);
  }
}

// Row pass: each wave transforms LPW rows at a time; waves stride over rows.
template <class P, int G>
__global__ __launch_bounds__((RegCfg<P, G>::kThreads)) void fft_rows_reg_kernel(
    typename C2<P>::T* u, int64_t nrows, const double2* tw_d, const double2* mult_d, int mode,
    double scale) {
  using T2 = typename C2<P>::T;
  using S = fftr::Shape<G>;
  using C = RegCfg<P, G>;
  constexpr int R = S::R, L = S::L;
  __shared__ T2 tw[G];
  __shared__ T2 mu[G];
  __shared__ T2 buf[C::kWaves * S::LPW * S::BUF];
  load_table(tw, tw_d, G);
  if (mode == ROW_PROP) load_table(mu, mult_d, G);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane / L, t = lane % L;
  T2* lbuf = buf + (wave * S::LPW + g) * S::BUF;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * C::kWaves;
  const P sc = static_cast<P>(scale);
  // software pipelined: the next line group is loaded before this one is
  // transformed (invalid line groups compute on zeros)
  auto load = [&](int64_t grp, T2 (&d)[R]) {
    const int64_t line = grp * S::LPW + g;
    const bool ok = line < nrows;
    const T2* row = u + line * G;
    fftr::static_for<0, R>([&](auto ic) {
      constexpr int r = decltype(ic)::value;
      d[r].x = 0;
      d[r].y = 0;
      if (ok) d[r] = ld_stream(&row[t + L * r]);
    });
  };
  int64_t grp = blockIdx.x * static_cast<int64_t>(C::kWaves) + wave;
  T2 vn[R];
  if (grp * S::LPW < nrows) load(grp, vn);
  for (; grp * S::LPW < nrows; grp += nwaves) {
    const int64_t line = grp * S::LPW + g;
    const bool valid = line < nrows;
    T2* row = u + line * G;
    T2 v[R];
    fftr::static_for<0, R>([&](auto ic) {
      constexpr int r = decltype(ic)::value;
      v[r] = vn[r];
    });
    if ((grp + nwaves) * S::LPW < nrows) load(grp + nwaves, vn);
    if (mode == ROW_PROP || mode == ROW_FWD) fftr::forward<T2, G>(v, t, tw, lbuf);
    if (mode == ROW_PROP) {
      fftr::static_for<0, R>([&](auto ic) {
        constexpr int q = decltype(ic)::value;
        v[q] = fftr::cmul(v[q], mu[q * L + t]);
      });
    }
    if (mode == ROW_PROP || mode == ROW_INV) fftr::inverse<T2, G>(v, t, tw, lbuf);
    if (valid) {
      fftr::static_for<0, R>([&](auto ic) {
        constexpr int r = decltype(ic)::value;
        T2 o = v[r];
        o.x *= sc;
        o.y *= sc;
        st_stream(&row[t + L * r], o);
      });
    }
  }
}

// Column pass: a workgroup owns kTile adjacent columns of one ring plane,
// loaded as whole row segments (RegCfg::kDirect; for tiles narrower than
// 64 bytes each wave's line slots load their columns), transformed in
// registers and stored back with the source added and the field sums of the
// tile reduced.
template <class P, int G>
__global__ __launch_bounds__((RegCfg<P, G>::kThreads),
                             (RegCfg<P, G>::kMinWaves)) void fft_cols_reg_kernel(
    typename C2<P>::T* u, int nslice, const double2* tw_d, const double2* mult_d,
    const double* sigmoid, int mode, double scale, const typename C2<P>::T* src, const int4* rect,
    int first, int add_src, const double2* work, double* partials) {
  using T2 = typename C2<P>::T;
  using S = fftr::Shape<G>;
  using C = RegCfg<P, G>;
  constexpr int R = S::R, L = S::L;
  constexpr int ntile = G / C::kTile;
  __shared__ T2 tw[G];
  __shared__ T2 mu[G];
  // line transpose buffers: complex, or scalars for the two-pass transposes
  using B = std::conditional_t<C::kSplit, P, T2>;
  __shared__ B buf[C::kBuf];
  __shared__ double red[C::kWaves][8];
  const int p = blockIdx.x / ntile;
  const int tile = blockIdx.x - p * ntile;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr bool kDirect = C::kDirect;
  constexpr int TW = C::kTile;
  // line slot (column of the tile) and lane of the line
  const int t = kDirect ? static_cast<int>(threadIdx.x) / TW : lane % L;
  const int cslot = kDirect ? static_cast<int>(threadIdx.x) % TW : wave * S::LPW + lane / L;
  const int c = tile * C::kTile + cslot;  // column of this line slot
  const bool valid = cslot < C::kTile;
  B* lbuf = kDirect ? buf + cslot : buf + cslot * S::BUF;
  constexpr int kStride = kDirect ? TW : 1;
  T2* plane = u + static_cast<int64_t>(p) * G * G;
  T2 v[R];
  if constexpr (kDirect) {
    fftr::static_for<0, R>([&](auto ic) {
      constexpr int r = decltype(ic)::value;
      v[r] = ld_stream(&plane[static_cast<int64_t>(t + L * r) * G + c]);
    });
  }
  load_table(tw, tw_d, G);
  if (mode == COL_PROP) load_table(mu, mult_d, G);
  __syncthreads();
  if constexpr (!kDirect) {
    fftr::static_for<0, R>([&](auto ic) {
      constexpr int r = decltype(ic)::value;
      v[r].x = 0;
      v[r].y = 0;
      if (valid) v[r] = plane[static_cast<int64_t>(t + L * r) * G + c];
    });
  }
  fftr::forward<T2, G, kStride, kDirect>(v, t, tw, lbuf);
  if (mode == COL_PROP) {
    fftr::static_for<0, R>([&](auto ic) {
      constexpr int q = decltype(ic)::value;
      v[q] = fftr::cmul(v[q], mu[q * L + t]);
    });
  } else if (valid) {
    // filter: this column holds x frequency pi(c), register q of lane t the
    // y frequency pi(t + L q)
    const int mx = fftr::spectrum_index(c, G);
    fftr::static_for<0, R>([&](auto ic) {
      constexpr int q = decltype(ic)::value;
      const int my = fftr::spectrum_index(t + L * q, G);
      const P f = static_cast<P>(sigmoid[static_cast<int64_t>(my) * G + mx]);
      v[q].x *= f;
      v[q].y *= f;
    });
  }
  fftr::inverse<T2, G, kStride, kDirect>(v, t, tw, lbuf);

  const P sc = static_cast<P>(scale);
  const int s = ((p - first) % nslice + nslice) % nslice;  // beam slice of this ring plane
  int4 rc = make_int4(0, 0, 0, 0);
  if (add_src == 1) rc = rect[s];
  const T2* sp = (add_src == 1) ? src + static_cast<int64_t>(s) * G * G : nullptr;
  const double2* wp = (add_src == 2) ? work + static_cast<int64_t>(p) * G * G : nullptr;
  const double shift = -0.5 * static_cast<double>(G - 1);
  const int64_t icenter = (static_cast<int64_t>(G) * G - 1) / 2;
  const bool src_col = sp && c >= rc.x && c < rc.z;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  if constexpr (kDirect) {
    const double dx = c + shift;
    // the stored value of cell (row, this thread's column)
    auto out = [&](int row, int64_t gi, T2 o) {
      o.x *= sc;
      o.y *= sc;
      if (src_col && row >= rc.y && row < rc.w) {
        const T2 sv = ld_stream(&sp[gi]);
        o.x += P(2) * sv.x;
        o.y += P(2) * sv.y;
      } else if (wp) {
        const double2 wv = ld_stream(&wp[gi]);
        o.x += static_cast<P>(wv.x);
        o.y += static_cast<P>(wv.y);
      }
      return o;
    };
    // The thread's column is fixed, so its x moments are dx and dx^2 times
    // its sum of |u|^2, taken once after the loop.  MIXED sums its R
    // cells in FP32 (as the particle pass sums its moments per thread); the
    // sums across threads and tiles are FP64.
    using A = std::conditional_t<sizeof(P) == 4, float, double>;
    A sw = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0;
    fftr::static_for<0, R>([&](auto ic) {
      constexpr int r = decltype(ic)::value;
      const int row = t + L * r;
      const int64_t gi = static_cast<int64_t>(row) * G + c;
      const T2 o = out(row, gi, v[r]);
      st_stream(&plane[gi], o);
      if (partials) {
        const A re = o.x, im = o.y;
        const A wei = re * re + im * im;
        const A dy = static_cast<A>(row + shift);
        sw += wei;
        s3 += dy * wei;
        s4 += dy * dy * wei;
        s5 += re;
        s6 += im;
      }
    });
    if (partials) {
      acc[0] = sw;
      acc[1] = dx * static_cast<double>(sw);
      acc[2] = dx * dx * static_cast<double>(sw);
      acc[3] = s3;
      acc[4] = s4;
      acc[5] = s5;
      acc[6] = s6;
      // the centre cell, recomputed as stored by the thread that holds it
      constexpr int crow = (G * G - 1) / 2 / G, ccol = (G * G - 1) / 2 % G;
      if (c == ccol && t == crow % L) {
        const T2 o = out(crow, icenter, v[crow / L]);
        partials[static_cast<int64_t>(nslice) * ntile * 8 + 2 * p] = o.x;
        partials[static_cast<int64_t>(nslice) * ntile * 8 + 2 * p + 1] = o.y;
      }
    }
  } else if (valid) {
    const double dx = c + shift;
    fftr::static_for<0, R>([&](auto ic) {
      constexpr int r = decltype(ic)::value;
      const int row = t + L * r;
      const int64_t gi = static_cast<int64_t>(row) * G + c;
      T2 o = v[r];
      o.x *= sc;
      o.y *= sc;
      if (src_col && row >= rc.y && row < rc.w) {
        const T2 sv = sp[gi];
        o.x += P(2) * sv.x;
        o.y += P(2) * sv.y;
      } else if (wp) {
        const double2 wv = wp[gi];
        o.x += static_cast<P>(wv.x);
        o.y += static_cast<P>(wv.y);
      }
      plane[gi] = o;
      if (partials) {
        const double re = o.x, im = o.y;
        const double wei = re * re + im * im;
        const double dy = row + shift;
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
    });
  }
  if (partials) {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const double a = wave_total(acc[k]);
      if (lane == 0) red[wave][k] = a;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
      double a = 0.0;
      for (int w = 0; w < C::kWaves; ++w) a += red[w][threadIdx.x];
      partials[(static_cast<int64_t>(p) * ntile + tile) * 8 + threadIdx.x] = a;
    }
  }
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

// ---------------------------------------------------------------------------
// Other sizes N with the propagator: convolution through power-of-two
// transforms.  One line's step IDFT_N(P . DFT_N(u)) is the circular
// convolution of u with h = IDFT_N(P); with g_j = h_((j - N + 1) mod N) for
// j <= 2N - 2 it is the linear convolution (u * g) at N - 1 ... 2N - 2, and a
// circular convolution of M >= 2N - 1 points gives those entries without
// wrap-around.  So each line is padded with zeros to M points, transformed
// forward, multiplied by H = DFT_M(g) / M (the host's table, in pi order) and
// transformed back, and points N - 1 ... 2N - 2 are its result: two M-point
// register transforms instead of two N-point direct DFTs (O(M log M) instead
// of O(N^2) per line).
// ---------------------------------------------------------------------------

// Launch shape of the convolution passes: as many waves (up to four) as fit
// their transpose buffers in 72 KiB (48 KiB for MIXED lines of 1024 points,
// where four waves measured slower than two), the padded lines taking twice
// the LDS of N-point ones; TW columns per column-pass tile (a power of two).
template <class P, int M>
struct ConvCfg {
  using S = fftr::Shape<M>;
  static constexpr int kWaveBytes = S::LPW * S::BUF * 2 * static_cast<int>(sizeof(P));
  static constexpr int kLds = sizeof(P) == 8 || M <= 512 ? 72 * 1024 : 48 * 1024;
  static constexpr int kWaves = kWaveBytes * 4 <= kLds ? 4 : (kWaveBytes * 2 <= kLds ? 2 : 1);
  static constexpr int kThreads = 64 * kWaves;
  // columns per column-pass tile: one line slot (L lanes) per column
  static constexpr int TW = kThreads / S::L;
  static constexpr int kBuf = kWaves * S::LPW * S::BUF;
};

// Row pass: rows of N points, each held on the lanes of a line slot as in
// fft_rows_reg_kernel (without its prefetch: the padded lines already take
// twice the registers).
template <class P, int M>
__global__ __launch_bounds__((ConvCfg<P, M>::kThreads)) void fft_rows_conv_kernel(
    typename C2<P>::T* u, int64_t nrows, int N, const double2* tw_d, const double2* mult_d) {
  using T2 = typename C2<P>::T;
  using S = fftr::Shape<M>;
  using C = ConvCfg<P, M>;
  constexpr int R = S::R, L = S::L;
  __shared__ T2 tw[M];
  __shared__ T2 mu[M];
  __shared__ T2 buf[C::kWaves * S::LPW * S::BUF];
  load_table(tw, tw_d, M);
  load_table(mu, mult_d, M);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane / L, t = lane % L;
  T2* lbuf = buf + (wave * S::LPW + g) * S::BUF;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * C::kWaves;
  for (int64_t grp = blockIdx.x * static_cast<int64_t>(C::kWaves) + wave; grp * S::LPW < nrows;
       grp += nwaves) {
    const int64_t line = grp * S::LPW + g;
    const bool valid = line < nrows;
    T2* row = u + line * N;
    T2 v[R];
    fftr::static_for<0, R>([&](auto ic) {
      constexpr int r = decltype(ic)::value;
      const int j = t + L * r;
      v[r].x = 0;
      v[r].y = 0;
      if (valid && j < N) v[r] = ld_stream(&row[j]);
    });
    fftr::forward<T2, M>(v, t, tw, lbuf);
    fftr::static_for<0, R>([&](auto ic) {
      constexpr int q = decltype(ic)::value;
      v[q] = fftr::cmul(v[q], mu[q * L + t]);
    });
    fftr::inverse<T2, M>(v, t, tw, lbuf);
    if (valid) {
      fftr::static_for<0, R>([&](auto ic) {
        constexpr int r = decltype(ic)::value;
        const int j = t + L * r - (N - 1);
        if (j >= 0 && j < N) st_stream(&row[j], v[r]);
      });
    }
  }
}

// Column pass: a workgroup owns TW adjacent columns of one ring plane (the
// last tile clipped at N), loaded and stored as row segments in the line
// transforms' own layout as in fft_cols_reg_kernel (thread (t, column
// tid % TW) holds rows t + L r, a column's lanes in several waves, the
// transposes through column-interleaved LDS), transformed as above, and
// written back with 2S added inside the slice's source rectangle (or the
// filtered source of the work plane) and the field sums of the tile.

template <class P, int M>
__global__ __launch_bounds__((ConvCfg<P, M>::kThreads)) void fft_cols_conv_kernel(
    typename C2<P>::T* u, int nslice, int N, const double2* tw_d, const double2* mult_d,
    const typename C2<P>::T* src, const int4* rect, int first, int add_src, const double2* work,
    double* partials) {
  using T2 = typename C2<P>::T;
  using S = fftr::Shape<M>;
  using C = ConvCfg<P, M>;
  constexpr int R = S::R, L = S::L, TW = C::TW;
  static_assert(C::kThreads / TW == L, "one line slot per column, L lanes each");
  __shared__ T2 tw[M];
  __shared__ T2 mu[M];
  __shared__ T2 buf[C::kBuf];
  __shared__ double red[C::kWaves][8];
  const int ntile = (N + TW - 1) / TW;
  const int p = blockIdx.x / ntile;
  const int tile = blockIdx.x - p * ntile;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = static_cast<int>(threadIdx.x) / TW;
  const int cslot = static_cast<int>(threadIdx.x) % TW;
  const int col = tile * TW + cslot;
  const bool colok = col < N;
  T2* plane = u + static_cast<int64_t>(p) * N * N;
  T2 v[R];
  fftr::static_for<0, R>([&](auto ic) {
    constexpr int r = decltype(ic)::value;
    const int j = t + L * r;
    v[r].x = 0;
    v[r].y = 0;
    if (colok && j < N) v[r] = ld_stream(&plane[static_cast<int64_t>(j) * N + col]);
  });
  load_table(tw, tw_d, M);
  load_table(mu, mult_d, M);
  __syncthreads();
  T2* lbuf = buf + cslot;
  fftr::forward<T2, M, TW, true>(v, t, tw, lbuf);
  fftr::static_for<0, R>([&](auto ic) {
    constexpr int q = decltype(ic)::value;
    v[q] = fftr::cmul(v[q], mu[q * L + t]);
  });
  fftr::inverse<T2, M, TW, true>(v, t, tw, lbuf);

  const int s = ((p - first) % nslice + nslice) % nslice;  // beam slice of this ring plane
  int4 rc = make_int4(0, 0, 0, 0);
  if (add_src == 1) rc = rect[s];
  const T2* sp = (add_src == 1) ? src + static_cast<int64_t>(s) * N * N : nullptr;
  const double2* wp = (add_src == 2) ? work + static_cast<int64_t>(p) * N * N : nullptr;
  const bool scol = sp && col >= rc.x && col < rc.z;
  const double shift = -0.5 * static_cast<double>(N - 1);
  const int64_t icenter = (static_cast<int64_t>(N) * N - 1) / 2;
  const double dx = col + shift;
  double sw = 0.0, acc[7] = {0, 0, 0, 0, 0, 0, 0};
  if (colok) {
    // register r holds output row t + L r - (N - 1)
    fftr::static_for<0, R>([&](auto ic) {
      constexpr int r = decltype(ic)::value;
      const int row = t + L * r - (N - 1);
      if (row >= 0 && row < N) {
        const int64_t gi = static_cast<int64_t>(row) * N + col;
        T2 o = v[r];
        if (scol && row >= rc.y && row < rc.w) {
          const T2 sv = ld_stream(&sp[gi]);
          o.x += P(2) * sv.x;
          o.y += P(2) * sv.y;
        } else if (wp) {
          const double2 wv = ld_stream(&wp[gi]);
          o.x += static_cast<P>(wv.x);
          o.y += static_cast<P>(wv.y);
        }
        st_stream(&plane[gi], o);
        if (partials) {
          const double re = o.x, im = o.y;
          const double wei = re * re + im * im;
          const double dy = row + shift;
          sw += wei;
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
    });
  }
  if (partials) {
    acc[0] = sw;
    acc[1] = dx * sw;
    acc[2] = dx * dx * sw;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const double a = wave_total(acc[k]);
      if (lane == 0) red[wave][k] = a;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
      double a = 0.0;
      for (int w = 0; w < C::kWaves; ++w) a += red[w][threadIdx.x];
      partials[(static_cast<int64_t>(p) * ntile + tile) * 8 + threadIdx.x] = a;
    }
  }
}

bool is_pow2(int g) { return g >= 2 && (g & (g - 1)) == 0; }

// Points M of the power-of-two transforms that step a line of G points by
// convolution (sizes that are not powers of two, up to 512), else 0.
int conv_points(int G) {
  if (G < 3 || G > 512 || is_pow2(G)) return 0;
  int m = 2;
  while (m < 2 * G - 1) m *= 2;
  return m;
}

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

template <class P, int G>
void rows_reg(Field* f, typename C2<P>::T* data, int mode, double scale) {
  using C = RegCfg<P, G>;
  constexpr int LPW = fftr::Shape<G>::LPW;
  const int64_t nrows = static_cast<int64_t>(f->nslice) * G;
  const int64_t ngroups = (nrows + LPW - 1) / LPW;
  const int64_t nblk = std::min<int64_t>((ngroups + C::kWaves - 1) / C::kWaves, 8192);
  hipLaunchKernelGGL((fft_rows_reg_kernel<P, G>), dim3(static_cast<unsigned>(nblk)),
                     dim3(C::kThreads), 0, f->ctx->stream, data, nrows, f->fft_tw.get(),
                     f->fft_prop, mode, scale);
}

template <class P, int G>
void cols_reg(Field* f, typename C2<P>::T* data, int mode, double scale, int add_src,
              double* partials) {
  using T2 = typename C2<P>::T;
  using C = RegCfg<P, G>;
  constexpr int ntile = G / C::kTile;
  hipLaunchKernelGGL((fft_cols_reg_kernel<P, G>), dim3(f->nslice * ntile), dim3(C::kThreads), 0,
                     f->ctx->stream, data, f->nslice, f->fft_tw.get(), f->fft_prop,
                     f->fft_sigmoid.get(), mode, scale, static_cast<const T2*>(f->src()),
                     f->src_rect.get(), f->first, add_src, f->fft_work.get(), partials);
}

template <class P, int M>
void rows_conv(Field* f, typename C2<P>::T* data) {
  using C = ConvCfg<P, M>;
  constexpr int LPW = fftr::Shape<M>::LPW;
  const int N = f->ngrid;
  const int64_t nrows = static_cast<int64_t>(f->nslice) * N;
  const int64_t ngroups = (nrows + LPW - 1) / LPW;
  const int64_t nblk = std::min<int64_t>((ngroups + C::kWaves - 1) / C::kWaves, 8192);
  hipLaunchKernelGGL((fft_rows_conv_kernel<P, M>), dim3(static_cast<unsigned>(nblk)),
                     dim3(C::kThreads), 0, f->ctx->stream, data, nrows, N, f->fft_tw_conv.get(),
                     f->fft_prop);
}

template <class P, int M>
void cols_conv(Field* f, typename C2<P>::T* data, int add_src, double* partials) {
  using T2 = typename C2<P>::T;
  using C = ConvCfg<P, M>;
  const int N = f->ngrid;
  const int ntile = (N + ConvCfg<P, M>::TW - 1) / ConvCfg<P, M>::TW;
  hipLaunchKernelGGL((fft_cols_conv_kernel<P, M>), dim3(f->nslice * ntile), dim3(C::kThreads), 0,
                     f->ctx->stream, data, f->nslice, N, f->fft_tw_conv.get(), f->fft_prop,
                     static_cast<const T2*>(f->src()), f->src_rect.get(), f->first, add_src,
                     f->fft_work.get(), partials);
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

// Columns per workgroup tile of the column pass.
int col_tile(int G, size_t esz) {
  if (const int M = conv_points(G)) {
    int tc = 0;
    with_pow2(M, [&](auto m) {
      constexpr int mv = decltype(m)::value;
      tc = esz == sizeof(double2) ? ConvCfg<double, mv>::TW : ConvCfg<float, mv>::TW;
    });
    return tc;
  }
  if (!is_pow2(G)) return dft_cfg(G, esz).tc;
  int tc = 0;
  with_pow2(G, [&](auto g) {
    constexpr int gv = decltype(g)::value;
    tc = esz == sizeof(double2) ? RegCfg<double, gv>::kTile : RegCfg<float, gv>::kTile;
  });
  return tc;
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
    if (mode == ROW_PROP && conv_points(G)) {  // the table holds the normalisation
      with_pow2(conv_points(G), [&](auto m) { rows_conv<P, decltype(m)::value>(f, data); });
      return;
    }
    const DftCfg c = dft_cfg(G, sizeof(T2));
    const int64_t nrows = static_cast<int64_t>(f->nslice) * G;
    const int64_t nblk = (nrows + c.row_lines - 1) / c.row_lines;
    const size_t lds = sizeof(T2) * (G + static_cast<size_t>(c.row_lines) * (G + 1));
    hipLaunchKernelGGL(fft_rows_dft_kernel<P>, dim3(static_cast<unsigned>(nblk)), dim3(256), lds,
                       f->ctx->stream, data, nrows, G, c.row_lines, f->fft_tw.get(),
                       f->fft_prop, mode, scale);
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
    if (mode == COL_PROP && conv_points(G)) {
      with_pow2(conv_points(G), [&](auto m) {
        cols_conv<P, decltype(m)::value>(f, data, add_src, partials);
      });
      return;
    }
    const DftCfg c = dft_cfg(G, sizeof(T2));
    const int ntile = (G + c.tc - 1) / c.tc;
    const size_t lds = sizeof(T2) * (G + static_cast<size_t>(G) * (c.tc + 1));
    hipLaunchKernelGGL(fft_cols_dft_kernel<P>, dim3(f->nslice * ntile), dim3(c.col_threads), lds,
                       f->ctx->stream, data, f->nslice, G, c.tc, f->fft_tw.get(),
                       f->fft_prop, f->fft_sigmoid.get(), mode, scale,
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
  if (stats) field_reduce_partials(f, partials, ntile);
}

// The M-point convolution table of a line step with factors P (natural
// frequency order, N points): h = IDFT_N(P), g_j = h_((j - N + 1) mod N) for
// j <= 2N - 2 (zero beyond), H = DFT_M(g) / M, stored in the pi order of the
// M-point register transforms.  Sums in long double over exact twiddle
// indices, so the table is correctly rounded to a few ulp.
std::vector<double2> conv_spectrum(const std::vector<double2>& P, int M) {
  using LD = long double;
  const int N = static_cast<int>(P.size());
  const LD pi2 = 8 * std::atan(static_cast<LD>(1));
  std::vector<LD> cn(N), sn(N), cm(M), sm(M);
  for (int k = 0; k < N; ++k) {
    cn[k] = std::cos(pi2 * k / N);
    sn[k] = std::sin(pi2 * k / N);
  }
  for (int k = 0; k < M; ++k) {
    cm[k] = std::cos(pi2 * k / M);
    sm[k] = std::sin(pi2 * k / M);
  }
  std::vector<LD> hr(N), hi(N);
  for (int n = 0; n < N; ++n) {  // h_n = (1/N) sum_m P_m e^{+2 pi i m n / N}
    LD ar = 0, ai = 0;
    for (int m = 0; m < N; ++m) {
      const int q = static_cast<int>((static_cast<int64_t>(m) * n) % N);
      ar += P[m].x * cn[q] - P[m].y * sn[q];
      ai += P[m].x * sn[q] + P[m].y * cn[q];
    }
    hr[n] = ar / N;
    hi[n] = ai / N;
  }
  std::vector<double2> out(M);
  for (int k = 0; k < M; ++k) {  // H_k = (1/M) sum_j g_j e^{-2 pi i j k / M}
    LD ar = 0, ai = 0;
    for (int j = 0; j <= 2 * N - 2; ++j) {
      const int n = (j - N + 1 + N) % N;
      const int q = static_cast<int>((static_cast<int64_t>(j) * k) % M);
      ar += hr[n] * cm[q] + hi[n] * sm[q];
      ai += hi[n] * cm[q] - hr[n] * sm[q];
    }
    out[k] = make_double2(static_cast<double>(ar / M), static_cast<double>(ai / M));
  }
  std::vector<double2> perm(M);
  for (int e = 0; e < M; ++e) perm[e] = out[fftr::spectrum_index(e, M)];
  return perm;
}

}  // namespace

void fft_prepare(Field* f, double delz) {
  const int G = f->ngrid;
  if (f->fft_tw.size() != static_cast<size_t>(G)) {
    // Power-of-two sizes: the four-step middle twiddles W_G^(t*k2) at
    // [k2*L + t]; other sizes: W_G^k at [k] for the direct DFT.
    std::vector<double2> tw(G);
    const double pi2 = 4.0 * std::asin(1.0);
    for (int k = 0; k < G; ++k) {
      int64_t e = k;
      if (is_pow2(G)) {
        const int L = fftr::lanes_per_line(G);
        e = static_cast<int64_t>(k % L) * (k / L);
      }
      const double a = -pi2 * static_cast<double>(e % G) / static_cast<double>(G);
      tw[k] = make_double2(std::cos(a), std::sin(a));
    }
    f->fft_tw.alloc(f->ctx, G);
    G4_HIP(hipMemcpy(f->fft_tw.get(), tw.data(), sizeof(double2) * G, hipMemcpyHostToDevice));
    if (const int M = conv_points(G)) {
      // the M-point convolution transforms' four-step twiddles
      const int L = fftr::lanes_per_line(M);
      std::vector<double2> twm(M);
      for (int k = 0; k < M; ++k) {
        const int64_t e = static_cast<int64_t>(k % L) * (k / L);
        const double a = -pi2 * static_cast<double>(e % M) / static_cast<double>(M);
        twm[k] = make_double2(std::cos(a), std::sin(a));
      }
      f->fft_tw_conv.alloc(f->ctx, M);
      G4_HIP(hipMemcpy(f->fft_tw_conv.get(), twm.data(), sizeof(double2) * M,
                       hipMemcpyHostToDevice));
    }
    f->fft_delz = 0.0;
    f->fft_prop = nullptr;
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
  if (!f->fft_prop || delz != f->fft_delz) {
    auto it = f->delz_tables.find(delz);
    if (it == f->delz_tables.end()) {
      if (f->delz_tables.size() >= Field::kMaxDelzTables) f->evict_delz_tables();
      // k2 at frequency index m: dx(m) = ((m - (G+1)/2) mod G) - (G-1)/2 and
      // k2 = -(dx^2 + dy^2) dk^2/(2 xks) (FieldSolverFFTCUDA.cpp:297-309);
      // the x (or y) factor is exp(-i dx^2 dk^2 delz / (2 xks)).
      // Power-of-two sizes use the four-step register transforms, whose
      // spectrum sits at permuted positions: entry e holds the factor of
      // frequency spectrum_index(e).
      const double dk = 4. * std::asin(1.) / (static_cast<double>(G) * f->dgrid);
      Field::DelzTable t;
      t.host.resize(G);
      for (int e = 0; e < G; ++e) {
        const int m = is_pow2(G) ? fftr::spectrum_index(e, G) : e;
        const int ix = ((m - (G + 1) / 2) % G + G) % G;
        const double dx = static_cast<double>(ix) - 0.5 * static_cast<double>(G - 1);
        const double ph = -dx * dx * dk * dk / 2. / f->xks * delz;
        t.host[e] = make_double2(std::cos(ph), std::sin(ph));
      }
      if (const int M = conv_points(G)) t.host = conv_spectrum(t.host, M);
      t.dev.alloc(f->ctx, t.host.size());
      G4_HIP(hipMemcpyAsync(t.dev.get(), t.host.data(), sizeof(double2) * t.host.size(),
                            hipMemcpyHostToDevice, f->ctx->stream));
      it = f->delz_tables.emplace(delz, std::move(t)).first;
    }
    f->fft_prop = it->second.dev.get();
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
