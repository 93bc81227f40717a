// loader.hip - initial state generated on the device: the quiet-start beam
// and a seed field that is the same plane in every slice.
//
// g4h_beam_generate fills a created beam with exactly the beamlets of
// g4h_load_beam (loader.hpp holds the one routine both use), one thread
// per beamlet writing its nbins particles in place, so a run needs no host
// arrays and no upload: generating 33 M particles is a kernel of a few ms
// where the host loader takes ~0.7 s on 16 cores.  FP64 math of the device
// (inverse normal, Box-Muller) and of the host library round alike to the
// last bit or two, so the two loaders agree to round-off.
#include <algorithm>
#include <vector>

#include "kernels.hpp"
#include "loader.hpp"

namespace g4 {
namespace {

template <class P>
__global__ void beam_generate_kernel(loader::Spec sp, const double* current, double gamma_ref,
                                     P* x, P* y, P* px, P* py, P* g, double* th) {
  const int nb = sp.npart / sp.nbins;
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(sp.nslice) * nb) return;
  const int s = static_cast<int>(t / nb), j = static_cast<int>(t - static_cast<int64_t>(s) * nb);
  double xv, pxv, yv, pyv, gv, tk[64];
  loader::beamlet(sp, s, j, sp.shotnoise ? current[s] : 0.0, G4H_CE, xv, pxv, yv, pyv, gv, tk);
  const int64_t i0 = static_cast<int64_t>(sp.npart) * s + static_cast<int64_t>(j) * sp.nbins;
  // MIXED stores gamma - gamma_ref (g4h_beam_upload's convention)
  const P gs = static_cast<P>(sizeof(P) == 4 ? gv - gamma_ref : gv);
  for (int k = 0; k < sp.nbins; ++k) {
    x[i0 + k] = static_cast<P>(xv);
    px[i0 + k] = static_cast<P>(pxv);
    y[i0 + k] = static_cast<P>(yv);
    py[i0 + k] = static_cast<P>(pyv);
    g[i0 + k] = gs;
    th[i0 + k] = tk[k];
  }
}

// every ring plane <- the plane `src`
template <class T2>
__global__ void fill_planes_kernel(const T2* src, int64_t plane, int nslice, T2* u) {
  const int64_t n = plane * nslice;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x)
    u[e] = src[e % plane];
}

}  // namespace
}  // namespace g4

using namespace g4;

extern "C" {

g4h_status g4h_beam_generate(g4h_beam* b, const g4h_beam_spec* spec, const double* current) {
  G4_API_BEGIN
  require(b && spec, "null argument");
  loader::Spec sp;
  require(loader_spec(spec, sp), "invalid beam specification");
  require(!spec->shotnoise || current, "shot noise needs the slice currents");
  require(spec->nslice == b->nslice, "beam and specification differ in slice count");
  for (int s = 0; s <= b->nslice; ++s)
    require(b->off_h[static_cast<size_t>(s)] == static_cast<int64_t>(s) * spec->npart,
            "beam slices must hold npart particles each");
  b->ctx->activate();
  DevBuf<double> cur;
  const double* dcur = nullptr;
  if (spec->shotnoise) {
    cur.alloc(b->ctx, static_cast<size_t>(b->nslice));
    G4_HIP(hipMemcpyAsync(cur.get(), current, sizeof(double) * b->nslice, hipMemcpyHostToDevice,
                          b->ctx->stream));
    dcur = cur.get();
  }
  const int64_t nthreads = static_cast<int64_t>(spec->nslice) * (spec->npart / spec->nbins);
  const unsigned blocks = static_cast<unsigned>((nthreads + 255) / 256);
  b->ctx->launch(G4H_K_OTHER, [&] {
    if (b->prec == G4H_FP64)
      hipLaunchKernelGGL(beam_generate_kernel<double>, dim3(blocks), dim3(256), 0, b->ctx->stream,
                         sp, dcur, b->gamma_ref, b->xd.get(), b->yd.get(), b->pxd.get(),
                         b->pyd.get(), b->gd.get(), b->theta.get());
    else
      hipLaunchKernelGGL(beam_generate_kernel<float>, dim3(blocks), dim3(256), 0, b->ctx->stream,
                         sp, dcur, b->gamma_ref, b->xf.get(), b->yf.get(), b->pxf.get(),
                         b->pyf.get(), b->gf.get(), b->theta.get());
  });
  b->uploaded = true;
  // the slice currents buffer is released after the kernel (stream-ordered)
  G4_API_END
}

g4h_status g4h_field_fill_planes(g4h_field* f, const double* re, const double* im, int first) {
  G4_API_BEGIN
  require(f && re && im, "null argument");
  require(first >= 0 && first < f->nslice, "first out of range");
  f->ctx->activate();
  const int64_t plane = static_cast<int64_t>(f->ngrid) * f->ngrid;
  const unsigned blocks = static_cast<unsigned>(std::min<int64_t>((plane * f->nslice + 255) / 256, 8192));
  if (f->prec == G4H_FP64) {
    std::vector<double2> h(static_cast<size_t>(plane));
    for (int64_t i = 0; i < plane; ++i) h[static_cast<size_t>(i)] = make_double2(re[i], im[i]);
    DevBuf<double2> d;
    d.alloc(f->ctx, h.size());
    G4_HIP(hipMemcpy(d.get(), h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice));
    f->ctx->launch(G4H_K_OTHER, [&] {
      hipLaunchKernelGGL(fill_planes_kernel<double2>, dim3(blocks), dim3(256), 0, f->ctx->stream,
                         d.get(), plane, f->nslice, f->ud.get());
    });
  } else {
    std::vector<float2> h(static_cast<size_t>(plane));
    for (int64_t i = 0; i < plane; ++i)
      h[static_cast<size_t>(i)] = make_float2(static_cast<float>(re[i]), static_cast<float>(im[i]));
    DevBuf<float2> d;
    d.alloc(f->ctx, h.size());
    G4_HIP(hipMemcpy(d.get(), h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice));
    f->ctx->launch(G4H_K_OTHER, [&] {
      hipLaunchKernelGGL(fill_planes_kernel<float2>, dim3(blocks), dim3(256), 0, f->ctx->stream,
                         d.get(), plane, f->nslice, f->uf.get());
    });
  }
  f->first = first;
  f->uploaded = true;
  G4_API_END
}

}  // extern "C"
