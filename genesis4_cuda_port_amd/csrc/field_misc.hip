// field_misc.hip - solver dispatch, slippage plane fill and field sums.
#include "kernels.hpp"

namespace g4 {

void adi_prepare(Field* f, double delz);
void adi_solve(Field* f, bool source_on, bool stats);
void fft_prepare(Field* f, double delz);
void fft_solve(Field* f, bool source_on, bool stats);

namespace {

// Zero ring slot `slot` (ControlCUDA.cpp:237-244 / zero_field_slice_gpu) and
// flag it so the diagnostics of this step report it as empty.
template <class T2>
__global__ void zero_plane_kernel(T2* u, int64_t plane, int slot, int32_t* zeroed) {
  T2* p = u + slot * plane;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < plane;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    T2 z;
    z.x = 0;
    z.y = 0;
    p[e] = z;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) zeroed[slot] = 1;
}

// raw[is] = plane_stats[(is+first)%n] unless that slot was zeroed since the
// stats were taken (DiagnosticCUDA.cpp:320 maps storage slot -> logical slice).
__global__ void map_stats_kernel(const double* ps, const int32_t* zeroed, int n, int first,
                                 double* raw) {
  const int is = blockIdx.x * blockDim.x + threadIdx.x;
  if (is >= n) return;
  const int p = (is + first) % n;
  const bool z = zeroed[p] != 0;
  for (int k = 0; k < G4H_FIELD_RAW; ++k)
    raw[static_cast<int64_t>(is) * G4H_FIELD_RAW + k] = z ? 0.0 : ps[static_cast<int64_t>(p) * G4H_FIELD_RAW + k];
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

__global__ void clear_flags_kernel(int32_t* zeroed, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) zeroed[i] = 0;
}

// Standalone per-plane field sums (g4_field_slice_reduce_kernel,
// DiagnosticCUDA.cpp:305-363): one workgroup per ring slot, wave shuffles.
template <class T2>
__global__ __launch_bounds__(256) void plane_stats_kernel(const T2* u, int G, double* stats) {
  __shared__ double red[4][8];
  const int p = blockIdx.x;
  const int64_t plane = static_cast<int64_t>(G) * G;
  const T2* b = u + p * plane;
  const double shift = -0.5 * static_cast<double>(G - 1);
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int64_t e = threadIdx.x; e < plane; e += blockDim.x) {
    const int iy = static_cast<int>(e / G);
    const int ix = static_cast<int>(e - static_cast<int64_t>(iy) * G);
    const double re = b[e].x, im = b[e].y;
    const double wei = re * re + im * im;
    const double dx = ix + shift, dy = iy + shift;
    acc[0] += wei;
    acc[1] += dx * wei;
    acc[2] += dx * dx * wei;
    acc[3] += dy * wei;
    acc[4] += dy * dy * wei;
    acc[5] += re;
    acc[6] += im;
  }
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
    for (int w = 0; w < 4; ++w) v += red[w][threadIdx.x];
    stats[static_cast<int64_t>(p) * G4H_FIELD_RAW + threadIdx.x] = v;
  }
  if (threadIdx.x == 0) {
    const int64_t ic = (plane - 1) / 2;
    stats[static_cast<int64_t>(p) * G4H_FIELD_RAW + 7] = b[ic].x;
    stats[static_cast<int64_t>(p) * G4H_FIELD_RAW + 8] = b[ic].y;
  }
}

// Streaming copy, 16 bytes per lane per iteration.
__global__ void copy16_kernel(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

__global__ void reset_rect_kernel(int4* rect, int4* bbox, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    rect[i] = make_int4(0, 0, 0, 0);
    bbox[i] = make_int4(0x7fffffff, 0x7fffffff, -1, -1);
  }
}

}  // namespace

void field_reduce_partials(Field* f, const double* partials, int ntile) {
  f->ctx->launch(G4H_K_OTHER, [&] {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((f->nslice + 255) / 256), dim3(256), 0,
                       f->ctx->stream, partials, f->nslice, ntile, f->plane_stats.get());
  });
}

void ensure_source(Field* f) {
  const int64_t cells = f->cells();
  bool fresh = false;
  if (f->prec == G4H_FP64) {
    if (f->srcd.size() != static_cast<size_t>(cells)) {
      f->srcd.alloc(f->ctx, cells);
      G4_HIP(hipMemsetAsync(f->srcd.get(), 0, sizeof(double2) * cells, f->ctx->stream));
      fresh = true;
    }
  } else {
    if (f->srcf.size() != static_cast<size_t>(cells)) {
      f->srcf.alloc(f->ctx, cells);
      G4_HIP(hipMemsetAsync(f->srcf.get(), 0, sizeof(float2) * cells, f->ctx->stream));
      fresh = true;
    }
  }
  if (f->src_rect.size() != static_cast<size_t>(f->nslice)) {
    f->src_rect.alloc(f->ctx, f->nslice);
    f->dep_bbox.alloc(f->ctx, f->nslice);
    fresh = true;
  }
  if (fresh) field_reset_source(f);
}

void field_reset_source(Field* f) {
  if (f->src_rect.empty()) return;
  f->ctx->launch(G4H_K_OTHER, [&] {
    hipLaunchKernelGGL(reset_rect_kernel, dim3((f->nslice + 255) / 256), dim3(256), 0,
                       f->ctx->stream, f->src_rect.get(), f->dep_bbox.get(), f->nslice);
  });
  // The source planes may hold stale values: clear them once.
  if (f->prec == G4H_FP64 && !f->srcd.empty())
    G4_HIP(hipMemsetAsync(f->srcd.get(), 0, sizeof(double2) * f->srcd.size(), f->ctx->stream));
  if (f->prec == G4H_MIXED && !f->srcf.empty())
    G4_HIP(hipMemsetAsync(f->srcf.get(), 0, sizeof(float2) * f->srcf.size(), f->ctx->stream));
}

static void ensure_stats(Field* f) {
  if (f->plane_stats.size() != static_cast<size_t>(f->nslice) * G4H_FIELD_RAW) {
    f->plane_stats.alloc(f->ctx, static_cast<size_t>(f->nslice) * G4H_FIELD_RAW);
    f->zeroed.alloc(f->ctx, f->nslice);
    G4_HIP(hipMemsetAsync(f->zeroed.get(), 0, sizeof(int32_t) * f->nslice, f->ctx->stream));
  }
}

void field_prepare_solver(Field* f, double delz) {
  if (f->kind == G4H_SOLVER_FFT)
    fft_prepare(f, delz);
  else
    adi_prepare(f, delz);
  f->solver_ready = true;
}

void field_solve(Field* f, bool source_on, bool stats) {
  ensure_source(f);
  if (stats) {
    ensure_stats(f);
    f->ctx->launch(G4H_K_OTHER, [&] {
      hipLaunchKernelGGL(clear_flags_kernel, dim3((f->nslice + 255) / 256), dim3(256), 0,
                         f->ctx->stream, f->zeroed.get(), f->nslice);
    });
  }
  if (f->kind == G4H_SOLVER_FFT)
    fft_solve(f, source_on, stats);
  else
    adi_solve(f, source_on, stats);
  f->stats_valid = stats;
}

void field_plane_stats(Field* f) {
  ensure_stats(f);
  f->ctx->launch(G4H_K_OTHER, [&] {
    hipLaunchKernelGGL(clear_flags_kernel, dim3((f->nslice + 255) / 256), dim3(256), 0,
                       f->ctx->stream, f->zeroed.get(), f->nslice);
  });
  f->ctx->launch(G4H_K_OTHER, [&] {
    if (f->prec == G4H_FP64)
      hipLaunchKernelGGL(plane_stats_kernel<double2>, dim3(f->nslice), dim3(256), 0, f->ctx->stream,
                         f->ud.get(), f->ngrid, f->plane_stats.get());
    else
      hipLaunchKernelGGL(plane_stats_kernel<float2>, dim3(f->nslice), dim3(256), 0, f->ctx->stream,
                         f->uf.get(), f->ngrid, f->plane_stats.get());
  });
  f->stats_valid = true;
}

void field_zero_plane(Field* f, int slot) {
  ensure_stats(f);
  const int64_t plane = static_cast<int64_t>(f->ngrid) * f->ngrid;
  const int blocks = static_cast<int>(std::min<int64_t>(1024, (plane + 255) / 256));
  f->ctx->launch(G4H_K_OTHER, [&] {
    if (f->prec == G4H_FP64)
      hipLaunchKernelGGL(zero_plane_kernel<double2>, dim3(blocks), dim3(256), 0, f->ctx->stream,
                         f->ud.get(), plane, slot, f->zeroed.get());
    else
      hipLaunchKernelGGL(zero_plane_kernel<float2>, dim3(blocks), dim3(256), 0, f->ctx->stream,
                         f->uf.get(), plane, slot, f->zeroed.get());
  });
}

void field_copy(Field* dst, const Field* src) {
  const size_t bytes = static_cast<size_t>(src->cells()) * (src->prec == G4H_FP64 ? 16 : 8);
  const int64_t n16 = static_cast<int64_t>(bytes / 16);
  const int blocks = static_cast<int>(std::min<int64_t>(8192, (n16 + 255) / 256));
  dst->ctx->launch(G4H_K_OTHER, [&] {
    hipLaunchKernelGGL(copy16_kernel, dim3(std::max(blocks, 1)), dim3(256), 0, dst->ctx->stream,
                       static_cast<const float4*>(src->data()), static_cast<float4*>(dst->data()),
                       n16);
  });
  dst->first = src->first;
  dst->accuslip = src->accuslip;
  dst->uploaded = true;
}

void field_map_stats(Field* f, double* d_raw) {
  require(f->stats_valid, "field statistics requested before they were computed");
  f->ctx->launch(G4H_K_OTHER, [&] {
    hipLaunchKernelGGL(map_stats_kernel, dim3((f->nslice + 255) / 256), dim3(256), 0,
                       f->ctx->stream, f->plane_stats.get(), f->zeroed.get(), f->nslice, f->first,
                       d_raw);
  });
}

}  // namespace g4
