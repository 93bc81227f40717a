// field_fft.hip - spectral paraxial step (FieldSolverFFTCUDA.cpp:157-188,
// 273-323, 413-481) as two hand-written passes over each plane.
//
// The reference computes u' = IFFT2[FFT2(u) e^{i k2 delz} + 2 filt FFT2(S)]/G^2
// with three batched cuFFT transforms and four elementwise passes.  Because
// k2 = -(dx^2+dy^2) dk^2/(2 xks) the propagator factorises, e^{i k2 delz} =
// Px(mx) Py(my), and with the filter off the source term is exactly 2S, so
//   u' = Cy( Cx(u) ) + 2S,   Cx = IFFT_x Px FFT_x / G  (rows),
//                            Cy = IFFT_y Py FFT_y / G  (columns).
// The row pass keeps a block of rows in LDS and does FFT -> Px -> IFFT
// there; the column pass stages a tile of columns through LDS (coalesced row
// segments), does FFT -> Py -> IFFT, adds 2S inside the slice's source
// rectangle and, on output steps, reduces the field diagnostics of the tile.
// Each plane crosses HBM twice per step (read+write per pass).
//
// Transforms: in-place Stockham (self-sorting) radix-4 with one radix-2 stage
// for odd log2(G); sizes that are not powers of two in [32, 1024] use a
// direct DFT in LDS.  With the source filter on (non-separable), the filtered
// source T = 2 IFFT2(filt FFT2(S))/G^2 is built in a work plane first
// (FWD rows, FWD*filt*INV columns, INV rows) and added in place of 2S.
#include "kernels.hpp"

#include <cmath>
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

__device__ __forceinline__ void bf_index(const LineLayout& L, int q, int nbf, int& line, int& j) {
  if (L.line_fast) {
    line = q % L.nlines;
    j = q / L.nlines;
  } else {
    j = q % nbf;
    line = q / nbf;
  }
}

// In-place Stockham transform of all lines; inverse uses conjugated
// twiddles and the conjugated radix-4 kernel.  tw[k] = exp(-2 pi i k/G).
// Each thread holds at most kMaxEpt complex values between the load and the
// store half of a stage; the loops are unrolled so they stay in registers.
template <class T2>
__device__ void stockham(T2* lds, const LineLayout& L, int log2g, const T2* tw, bool inverse,
                         int nthreads) {
  const int G = L.G;
  int ns = 1;
  if (log2g & 1) {  // one radix-2 stage first (ns = 1: no twiddles)
    const int nbf = G / 2;
    const int total = L.nlines * nbf;
    constexpr int MB = kMaxEpt / 2;
    T2 v[MB][2];
    int dst[MB];
#pragma unroll
    for (int c = 0; c < MB; ++c) {
      const int q = threadIdx.x + c * nthreads;
      if (q < total) {
        int line, j;
        bf_index(L, q, nbf, line, j);
        T2* b = lds + line * L.lstride;
        const T2 a0 = b[j * L.estride];
        const T2 a1 = b[(j + nbf) * L.estride];
        v[c][0] = cadd(a0, a1);
        v[c][1] = csub(a0, a1);
        dst[c] = line * L.lstride + (2 * j) * L.estride;
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MB; ++c) {
      const int q = threadIdx.x + c * nthreads;
      if (q < total) {
        lds[dst[c]] = v[c][0];
        lds[dst[c] + L.estride] = v[c][1];
      }
    }
    __syncthreads();
    ns = 2;
  }
  for (; ns < G; ns *= 4) {
    const int nbf = G / 4;
    const int total = L.nlines * nbf;
    const int twstep = G / (ns * 4);
    constexpr int MB = kMaxEpt / 4;
    T2 v[MB][4];
    int dst[MB];
#pragma unroll
    for (int c = 0; c < MB; ++c) {
      const int q = threadIdx.x + c * nthreads;
      if (q < total) {
        int line, j;
        bf_index(L, q, nbf, line, j);
        T2* b = lds + line * L.lstride;
        T2 a0 = b[j * L.estride];
        T2 a1 = b[(j + nbf) * L.estride];
        T2 a2 = b[(j + 2 * nbf) * L.estride];
        T2 a3 = b[(j + 3 * nbf) * L.estride];
        const int k = j % ns;
        if (ns > 1) {
          T2 w1 = tw[k * twstep], w2 = tw[2 * k * twstep], w3 = tw[3 * k * twstep];
          if (inverse) {
            w1 = conjv(w1);
            w2 = conjv(w2);
            w3 = conjv(w3);
          }
          a1 = cmul(a1, w1);
          a2 = cmul(a2, w2);
          a3 = cmul(a3, w3);
        }
        const T2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3);
        const T2 t3 = csub(a1, a3);
        // forward: t3 * (-i); inverse: t3 * (+i)
        T2 r3;
        if (inverse) {
          r3.x = -t3.y;
          r3.y = t3.x;
        } else {
          r3.x = t3.y;
          r3.y = -t3.x;
        }
        v[c][0] = cadd(t0, t2);
        v[c][1] = cadd(t1, r3);
        v[c][2] = csub(t0, t2);
        v[c][3] = csub(t1, r3);
        dst[c] = line * L.lstride + ((j - k) * 4 + k) * L.estride;
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MB; ++c) {
      const int q = threadIdx.x + c * nthreads;
      if (q < total) {
#pragma unroll
        for (int r = 0; r < 4; ++r) lds[dst[c] + r * ns * L.estride] = v[c][r];
      }
    }
    __syncthreads();
  }
}

// Direct DFT for sizes that are not powers of two in [32, 1024].
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

template <class T2>
__device__ __forceinline__ void transform(T2* lds, const LineLayout& L, int log2g, const T2* tw,
                                          bool inverse, int nthreads) {
  if (log2g >= 0)
    stockham(lds, L, log2g, tw, inverse, nthreads);
  else
    dft_lines(lds, L, tw, inverse, nthreads);
}

enum RowMode { ROW_PROP = 0, ROW_FWD = 1, ROW_INV = 2 };
enum ColMode { COL_PROP = 0, COL_FILTER = 1 };

// Row pass over all nslice*G rows, `nl` rows per workgroup.
template <class P>
__global__ __launch_bounds__(256) void fft_rows_kernel(typename C2<P>::T* u, int64_t nrows, int G,
                                                       int log2g, int nl, const double2* tw_d,
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
    lds[l * pitch + c] = (l < nvalid) ? g[e] : z;
  }
  __syncthreads();
  const LineLayout L{nl, G, pitch, 1, false};
  if (mode == ROW_PROP || mode == ROW_FWD) transform(lds, L, log2g, tw, false, nt);
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
  if (mode == ROW_PROP || mode == ROW_INV) transform(lds, L, log2g, tw, true, nt);
  const P sc = static_cast<P>(scale);
  for (int e = threadIdx.x; e < nvalid * G; e += nt) {
    const int l = e / G, c = e - l * G;
    T2 v = lds[l * pitch + c];
    v.x *= sc;
    v.y *= sc;
    g[e] = v;
  }
}

// Column pass: tile of tc columns x G rows of one ring plane.
template <class P>
__global__ __launch_bounds__(512) void fft_cols_kernel(
    typename C2<P>::T* u, int nslice, int G, int log2g, int tc, const double2* tw_d,
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
    lds[r * pitch + c] = (c < ncol) ? plane[static_cast<int64_t>(r) * G + c0 + c] : z;
  }
  __syncthreads();
  const LineLayout L{tc, G, 1, pitch, true};
  transform(lds, L, log2g, tw, false, nt);
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
  transform(lds, L, log2g, tw, true, nt);

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

int ilog2_pow2(int g) {
  if (g < 32 || g > 1024 || (g & (g - 1))) return -1;
  int l = 0;
  while ((1 << l) < g) ++l;
  return l;
}

struct FftCfg {
  int log2g, row_lines, row_threads, tc, col_threads;
};

FftCfg fft_cfg(int G, size_t esz) {
  FftCfg c;
  c.log2g = ilog2_pow2(G);
  c.row_threads = 256;
  c.row_lines = std::max(1, 2048 / G);
  while (c.row_lines * G > c.row_threads * kMaxEpt) --c.row_lines;
  while (c.row_lines > 1 && (static_cast<size_t>(c.row_lines) * (G + 1) + G) * esz > 60 * 1024)
    --c.row_lines;
  c.tc = esz == 8 ? 16 : 8;
  // tile (padded rows) + twiddles within 60 KiB of LDS
  while (c.tc > 1 && (static_cast<size_t>(G) * (c.tc + 1) + G) * esz > 60 * 1024) c.tc /= 2;
  c.col_threads = 256;
  while (c.tc * G > c.col_threads * kMaxEpt && c.col_threads < 512) c.col_threads *= 2;
  while (c.tc * G > c.col_threads * kMaxEpt && c.tc > 1) c.tc /= 2;
  return c;
}

template <class P>
void run_rows(Field* f, typename C2<P>::T* data, int mode, double scale, const FftCfg& c) {
  using T2 = typename C2<P>::T;
  const int G = f->ngrid;
  const int64_t nrows = static_cast<int64_t>(f->nslice) * G;
  const int64_t nblk = (nrows + c.row_lines - 1) / c.row_lines;
  const size_t lds = sizeof(T2) * (G + static_cast<size_t>(c.row_lines) * (G + 1));
  f->ctx->launch(G4H_K_FFT_ROWS, [&] {
    hipLaunchKernelGGL(fft_rows_kernel<P>, dim3(static_cast<unsigned>(nblk)), dim3(c.row_threads),
                       lds, f->ctx->stream, data, nrows, G, c.log2g, c.row_lines,
                       f->fft_tw.get(), f->fft_prop.get(), mode, scale);
  });
}

template <class P>
void run_cols(Field* f, typename C2<P>::T* data, int mode, double scale, int add_src,
              double* partials, const FftCfg& c) {
  using T2 = typename C2<P>::T;
  const int G = f->ngrid;
  const int ntile = (G + c.tc - 1) / c.tc;
  const size_t lds = sizeof(T2) * (G + static_cast<size_t>(G) * (c.tc + 1));
  const T2* src = static_cast<const T2*>(f->src());
  f->ctx->launch(G4H_K_FFT_COLS, [&] {
    hipLaunchKernelGGL(fft_cols_kernel<P>, dim3(f->nslice * ntile), dim3(c.col_threads), lds,
                       f->ctx->stream, data, f->nslice, G, c.log2g, c.tc, f->fft_tw.get(),
                       f->fft_prop.get(), f->fft_sigmoid.get(), mode, scale, src,
                       f->src_rect.get(), f->first, add_src, f->fft_work.get(), partials);
  });
}

template <class P>
void fft_step(Field* f, bool source_on, bool stats) {
  using T2 = typename C2<P>::T;
  const int G = f->ngrid;
  const FftCfg c = fft_cfg(G, sizeof(T2));
  T2* data = static_cast<T2*>(f->data());
  const double invg = 1.0 / static_cast<double>(G);
  const int ntile = (G + c.tc - 1) / c.tc;
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
    run_rows<double>(f, f->fft_work.get(), ROW_FWD, 1.0, fft_cfg(G, sizeof(double2)));
    const FftCfg cd = fft_cfg(G, sizeof(double2));
    run_cols<double>(f, f->fft_work.get(), COL_FILTER, 1.0, 0, nullptr, cd);
    run_rows<double>(f, f->fft_work.get(), ROW_INV, 2.0 * invg * invg, cd);
    add = 2;
  }
  run_rows<P>(f, data, ROW_PROP, invg, c);
  run_cols<P>(f, data, COL_PROP, invg, add, partials, c);
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
    const double dk = 4. * std::asin(1.) / (static_cast<double>(G) * f->dgrid);
    std::vector<double2> pr(G);
    for (int m = 0; m < G; ++m) {
      const int ix = ((m - (G + 1) / 2) % G + G) % G;
      const double dx = static_cast<double>(ix) - 0.5 * static_cast<double>(G - 1);
      const double ph = -dx * dx * dk * dk / 2. / f->xks * delz;
      pr[m] = make_double2(std::cos(ph), std::sin(ph));
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

bool fft_supported(int G) {
  if (G < 2 || G > 1024) return false;
  return true;
}

}  // namespace g4
