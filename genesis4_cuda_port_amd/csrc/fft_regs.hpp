// fft_regs.hpp - register-resident power-of-two line transforms for gfx950.
//
// A line of G points is split G = R x L.  In the natural layout lane t of the
// line's L lanes holds the points n = t + L*r in register r (r < R); a wave
// holds LPW = 64/L lines.  The forward transform is the four-step
// factorisation
//   X[k2 + R*k1] = sum_t W_L^(t*k1) W_G^(t*k2) sum_r x[t + L*r] W_R^(r*k2)
// done as: an R-point DFT over the registers of each lane, the twiddle
// W_G^(t*k2), a transpose through a per-line LDS buffer, and L-point DFTs
// over the registers again.  Both register DFTs have compile-time twiddles
// (trivial ones cost nothing).  After the transpose lane t' holds, in
// register q = j*L + k1 (j < R/L), the frequency
//   m = pi(t' + L*q) = t' + L*j + R*k1,
// i.e. the spectrum is stored in the natural layout's positions permuted by
// pi (identity when R == L).  The inverse runs the same steps backwards with
// conjugate twiddles and returns to the natural layout, so
// forward -> multiply (table in pi order) -> inverse needs no reordering.
// Lines with G <= 32 sit in one lane (L = 1) and need no transpose.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace g4 {
namespace fftr {

// ---- complex helpers ------------------------------------------------------
template <class T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  T2 r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <class T2>
__device__ __forceinline__ T2 cmulc(T2 a, T2 b) {  // a * conj(b)
  T2 r;
  r.x = a.x * b.x + a.y * b.y;
  r.y = a.y * b.x - a.x * b.y;
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

// ---- compile-time trigonometry (cos and sin of 2*pi*k/n) --------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double sin_series(double x) {  // |x| <= pi/4
  double x2 = x * x, term = x, sum = x;
  for (int n = 1; n < 12; ++n) {
    term *= -x2 / static_cast<double>((2 * n) * (2 * n + 1));
    sum += term;
  }
  return sum;
}
constexpr double cos_series(double x) {  // |x| <= pi/4
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int n = 1; n < 12; ++n) {
    term *= -x2 / static_cast<double>((2 * n - 1) * (2 * n));
    sum += term;
  }
  return sum;
}
// quadrant reduction in integers, then the octant with the smaller angle
constexpr double trig2pi(int k, int n, bool want_sin) {
  k = ((k % n) + n) % n;
  const int q = (4 * k) / n;
  const int rem = 4 * k - q * n;  // angle in quadrant = (pi/2) rem/n
  double c = 0.0, s = 0.0;
  if (2 * rem <= n) {
    const double a = 0.5 * kPi * static_cast<double>(rem) / static_cast<double>(n);
    c = cos_series(a);
    s = sin_series(a);
  } else {
    const double b = 0.5 * kPi * static_cast<double>(n - rem) / static_cast<double>(n);
    c = sin_series(b);
    s = cos_series(b);
  }
  double cq = c, sq = s;
  if (q == 1) {
    cq = -s;
    sq = c;
  } else if (q == 2) {
    cq = -c;
    sq = -s;
  } else if (q == 3) {
    cq = s;
    sq = -c;
  }
  return want_sin ? sq : cq;
}

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

constexpr int bitrev(int k, int n) {
  int r = 0;
  for (int b = 1; b < n; b <<= 1) {
    r = (r << 1) | (k & 1);
    k >>= 1;
  }
  return r;
}

// z * W_M^J with W_M = exp(-2 pi i / M) (forward) or its conjugate (inverse)
template <int M, int J, bool INV, class T2>
__device__ __forceinline__ T2 twmul(T2 z) {
  using P = decltype(z.x);
  constexpr int j = J % M;
  T2 r;
  if constexpr (j == 0) {
    return z;
  } else if constexpr (2 * j == M) {
    r.x = -z.x;
    r.y = -z.y;
  } else if constexpr (4 * j == M) {  // -i (forward), +i (inverse)
    if constexpr (!INV) {
      r.x = z.y;
      r.y = -z.x;
    } else {
      r.x = -z.y;
      r.y = z.x;
    }
  } else if constexpr (4 * j == 3 * M) {  // +i (forward), -i (inverse)
    if constexpr (!INV) {
      r.x = -z.y;
      r.y = z.x;
    } else {
      r.x = z.y;
      r.y = -z.x;
    }
  } else {
    constexpr double c = trig2pi(j, M, false);
    constexpr double s = INV ? trig2pi(j, M, true) : -trig2pi(j, M, true);
    r.x = z.x * static_cast<P>(c) - z.y * static_cast<P>(s);
    r.y = z.x * static_cast<P>(s) + z.y * static_cast<P>(c);
  }
  return r;
}

// radix-2 decimation-in-frequency stages of an N-point DFT on v[OFF .. OFF+N)
template <class T2, int R, int N, int S, int OFF, bool INV>
__device__ __forceinline__ void dif_stages(T2 (&v)[R]) {
  if constexpr (S >= 1) {
    static_for<0, N>([&](auto ic) {
      constexpr int r = decltype(ic)::value;
      if constexpr ((r & S) == 0) {
        const T2 a = v[OFF + r], b = v[OFF + r + S];
        v[OFF + r] = cadd(a, b);
        v[OFF + r + S] = twmul<2 * S, (r & (S - 1)), INV>(csub(a, b));
      }
    });
    dif_stages<T2, R, N, S / 2, OFF, INV>(v);
  }
}

// N-point DFT (unnormalised) of v[OFF .. OFF+N), natural order in and out;
// the bit-reversal is a compile-time renaming of registers.
template <class T2, int R, int N, int OFF, bool INV>
__device__ __forceinline__ void dft_regs(T2 (&v)[R]) {
  if constexpr (N > 1) {
    dif_stages<T2, R, N, N / 2, OFF, INV>(v);
    T2 t[N];
    static_for<0, N>([&](auto ic) {
      constexpr int k = decltype(ic)::value;
      t[k] = v[OFF + bitrev(k, N)];
    });
    static_for<0, N>([&](auto ic) {
      constexpr int k = decltype(ic)::value;
      v[OFF + k] = t[k];
    });
  }
}

// Lanes per line L of a G-point line (R = G/L registers per lane, R >= L).
__host__ __device__ constexpr int lanes_per_line(int G) {
  return G <= 32 ? 1 : (G <= 128 ? 8 : (G <= 512 ? 16 : 32));
}

template <int G>
struct Shape {
  static constexpr int L = lanes_per_line(G);
  static constexpr int R = G / L;          // registers per lane
  static constexpr int LPW = 64 / L;       // lines per wave
  static constexpr int SUB = R / L;        // L-point vectors per lane after the transpose
  static constexpr int LP = L + 1;         // padded transpose row
  static constexpr int BUF = L > 1 ? R * LP : 1;  // transpose buffer per line (complex)
};

// Ordering fence between the LDS writes and reads of one wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Ordering between the writes and reads of a line's transpose buffer: the
// line's lanes are in one wave, or (BLOCK) spread over the workgroup's waves.
template <bool BLOCK>
__device__ __forceinline__ void line_sync() {
  if constexpr (BLOCK) {
    __syncthreads();
  } else {
    wave_lds_sync();
  }
}

// One transpose of a line: v[k] goes to buf[wi(k) * STRIDE], then v[k] is read
// back from buf[ri(k) * STRIDE].  A buffer of scalars (B = the component
// type) takes the real parts, then the imaginary ones (half the LDS, two
// more barriers); the reads of each pass overwrite only the component that
// pass has written.
template <class T2, int R, int STRIDE, bool BLOCK, class B, class WI, class RI>
__device__ __forceinline__ void xpose(T2 (&v)[R], B* buf, WI&& wi, RI&& ri) {
  if constexpr (!std::is_same_v<B, T2>) {
    static_for<0, R>([&](auto ic) { buf[wi(ic) * STRIDE] = v[decltype(ic)::value].x; });
    line_sync<BLOCK>();
    static_for<0, R>([&](auto ic) { v[decltype(ic)::value].x = buf[ri(ic) * STRIDE]; });
    line_sync<BLOCK>();
    static_for<0, R>([&](auto ic) { buf[wi(ic) * STRIDE] = v[decltype(ic)::value].y; });
    line_sync<BLOCK>();
    static_for<0, R>([&](auto ic) { v[decltype(ic)::value].y = buf[ri(ic) * STRIDE]; });
  } else {
    static_for<0, R>([&](auto ic) { buf[wi(ic) * STRIDE] = v[decltype(ic)::value]; });
    line_sync<BLOCK>();
    static_for<0, R>([&](auto ic) { v[decltype(ic)::value] = buf[ri(ic) * STRIDE]; });
  }
}

// Forward transform: natural layout in, spectrum in pi order out.
// tw[k2*L + t] = W_G^(t*k2); buf is this line's transpose buffer, whose
// element e sits at buf[e * STRIDE] (STRIDE > 1: the buffers of STRIDE lines
// interleaved, so that lanes of different lines at one e hit adjacent LDS
// words); BLOCK: the line's lanes are in different waves of the workgroup;
// B: complex, or scalar for a transpose in two passes (xpose).
template <class T2, int G, int STRIDE = 1, bool BLOCK = false, class B = T2>
__device__ __forceinline__ void forward(T2 (&v)[Shape<G>::R], int t, const T2* tw, B* buf) {
  using S = Shape<G>;
  constexpr int R = S::R, L = S::L;
  dft_regs<T2, R, R, 0, false>(v);
  if constexpr (L > 1) {
    static_for<1, R>([&](auto ic) {
      constexpr int k2 = decltype(ic)::value;
      v[k2] = cmul(v[k2], tw[k2 * L + t]);
    });
    // v[k2] to row k2 of the buffer, register q = j*L + n1 from row
    // t + L*j, column n1
    xpose<T2, R, STRIDE, BLOCK>(
        v, buf, [&](auto k2) { return decltype(k2)::value * S::LP + t; },
        [&](auto qc) {
          constexpr int q = decltype(qc)::value;
          return (t + L * (q / L)) * S::LP + (q % L);
        });
    line_sync<BLOCK>();
    static_for<0, S::SUB>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      dft_regs<T2, R, L, j * L, false>(v);
    });
  }
}

// Inverse (unnormalised): spectrum in pi order in, natural layout out.
template <class T2, int G, int STRIDE = 1, bool BLOCK = false, class B = T2>
__device__ __forceinline__ void inverse(T2 (&v)[Shape<G>::R], int t, const T2* tw, B* buf) {
  using S = Shape<G>;
  constexpr int R = S::R, L = S::L;
  if constexpr (L > 1) {
    static_for<0, S::SUB>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      dft_regs<T2, R, L, j * L, true>(v);
    });
    xpose<T2, R, STRIDE, BLOCK>(
        v, buf,
        [&](auto qc) {
          constexpr int q = decltype(qc)::value;
          return (t + L * (q / L)) * S::LP + (q % L);
        },
        [&](auto k2) { return decltype(k2)::value * S::LP + t; });
    // a wave's next line reuses its buffer at once; a workgroup's callers
    // synchronise before they reuse theirs
    if constexpr (!BLOCK) wave_lds_sync();
    static_for<1, R>([&](auto ic) {
      constexpr int k2 = decltype(ic)::value;
      v[k2] = cmulc(v[k2], tw[k2 * L + t]);
    });
  }
  dft_regs<T2, R, R, 0, true>(v);
}

// Frequency held by position e of the natural layout after forward().
__host__ __device__ constexpr int spectrum_index(int e, int G) {
  const int L = lanes_per_line(G);
  const int R = G / L;
  const int t = e % L, q = e / L;
  return t + L * (q / L) + R * (q % L);
}

}  // namespace fftr
}  // namespace g4
