// frontend.cpp - host-only input front-end: undulator coupling, quiet-start
// loader, seed field and a FODO lattice table.  The reference takes all of
// these from upstream GENESIS 1.3 (Undulator, QuietLoading, ShotNoise,
// LoadField, Lattice), which is not part of it; each piece below is restated
// from the published method named in its comment.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/g4hip.h"
#include "loader.hpp"

namespace g4 {
void set_last_error(const std::string& m);
}

namespace {

constexpr double kPi = 3.14159265358979323846;

// Joe & Kuo (2008) direction numbers, file new-joe-kuo-6.21201, dims 2..8:
// degree s, polynomial coefficient a, initial m_1..m_s.
struct SobolDim {
  int s, a;
  uint32_t m[5];
};
const SobolDim kSobol[7] = {
    {1, 0, {1}},          {2, 1, {1, 3}},          {3, 1, {1, 3, 1}},        {3, 2, {1, 1, 1}},
    {4, 1, {1, 1, 3, 3}}, {4, 4, {1, 3, 5, 13}},   {5, 2, {1, 1, 5, 5, 17}},
};

// v[d][i] (i = 0..31): direction integers scaled to 32 bits.
void sobol_directions(int dims, std::vector<uint32_t>& v) {
  v.assign(static_cast<size_t>(dims) * 32, 0u);
  for (int i = 0; i < 32; ++i) v[i] = 1u << (31 - i);  // dim 1: van der Corput
  for (int d = 1; d < dims; ++d) {
    const SobolDim& sd = kSobol[d - 1];
    uint32_t* vd = v.data() + static_cast<size_t>(d) * 32;
    for (int i = 0; i < 32; ++i) {
      if (i < sd.s) {
        vd[i] = sd.m[i] << (31 - i);
      } else {
        uint32_t x = vd[i - sd.s] ^ (vd[i - sd.s] >> sd.s);
        for (int k = 1; k < sd.s; ++k)
          if ((sd.a >> (sd.s - 1 - k)) & 1) x ^= vd[i - k];
        vd[i] = x;
      }
    }
  }
}

// Sobol point n (gray-code order as in Bratley & Fox / scipy) as 32-bit ints.
void sobol_point(int dims, uint64_t n, const std::vector<uint32_t>& v, uint32_t* out) {
  const uint64_t g = n ^ (n >> 1);
  for (int d = 0; d < dims; ++d) {
    uint32_t x = 0;
    for (int i = 0; i < 32; ++i)
      if ((g >> i) & 1) x ^= v[static_cast<size_t>(d) * 32 + i];
    out[d] = x;
  }
}

}  // namespace

namespace g4 {
// Loader parameters of a beam specification (shared with g4h_beam_generate).
bool loader_spec(const g4h_beam_spec* sp, loader::Spec& out) {
  if (!sp || sp->nslice < 1 || sp->npart < 1 || sp->nbins < 1 || sp->nbins > 64 ||
      (sp->npart % sp->nbins) != 0 || sp->gamma0 <= 1.0 || sp->betax <= 0 || sp->betay <= 0)
    return false;
  std::vector<uint32_t> v;
  sobol_directions(loader::kDims, v);
  for (int d = 0; d < loader::kDims; ++d)
    for (int i = 0; i < 32; ++i) out.v[d][i] = v[static_cast<size_t>(d) * 32 + i];
  // digital shift (XOR with seeded random words): a randomised Sobol net
  std::mt19937_64 rs(sp->seed);
  for (int d = 0; d < loader::kDims; ++d) out.shift[d] = static_cast<uint32_t>(rs() >> 32);
  const double ex = sp->emitn_x / sp->gamma0, ey = sp->emitn_y / sp->gamma0;
  out.sx = std::sqrt(ex * sp->betax);
  out.sxp = std::sqrt(ex / sp->betax);
  out.sy = std::sqrt(ey * sp->betay);
  out.syp = std::sqrt(ey / sp->betay);
  out.alphax = sp->alphax;
  out.alphay = sp->alphay;
  out.gamma0 = sp->gamma0;
  out.espread_rel = sp->espread_rel;
  out.chirp_rel = sp->chirp_rel;
  out.xcen = sp->xcen;
  out.ycen = sp->ycen;
  out.pxcen = sp->pxcen;
  out.pycen = sp->pycen;
  out.slicelength = sp->slicelength;
  out.seed = sp->seed;
  out.nslice = sp->nslice;
  out.npart = sp->npart;
  out.nbins = sp->nbins;
  out.shotnoise = sp->shotnoise;
  return true;
}
}  // namespace g4

extern "C" {

// fc(h): planar coupling aw*[JJ]_h, [JJ]_h = (-1)^((h-1)/2) (J_{(h-1)/2}(h xi) -
// J_{(h+1)/2}(h xi)), xi = aw^2/(2(1+aw^2)) with aw the rms parameter; helical:
// aw for h = 1, else 0; even harmonics: 0 (no on-axis coupling).
double g4h_undulator_fc(double aw, int helical, int harm) {
  if (aw == 0.0 || harm < 1) return 0.0;
  if (helical) return harm == 1 ? aw : 0.0;
  if ((harm % 2) == 0) return 0.0;
  const double xi = static_cast<double>(harm) * aw * aw / (2.0 * (1.0 + aw * aw));
  const int n = (harm - 1) / 2;
  const double sign = (n % 2) ? -1.0 : 1.0;
  return aw * sign * (::jn(n, xi) - ::jn(n + 1, xi));
}

g4h_status g4h_sobol(int dims, int64_t index0, int64_t n, double* out) {
  if (dims < 1 || dims > 8 || index0 < 0 || n < 0 || (n > 0 && !out)) {
    g4::set_last_error("g4h_sobol: dims must be in [1,8], index0, n >= 0");
    return G4H_ERR_INVALID_ARGUMENT;
  }
  std::vector<uint32_t> v;
  sobol_directions(dims, v);
  uint32_t pt[8];
  for (int64_t i = 0; i < n; ++i) {
    sobol_point(dims, static_cast<uint64_t>(index0 + i), v, pt);
    for (int d = 0; d < dims; ++d) out[i * dims + d] = static_cast<double>(pt[d]) * (1.0 / 4294967296.0);
  }
  return G4H_OK;
}

// Inverse standard normal CDF (the loader's Halley iteration, loader.hpp).
double g4h_inverse_normal(double p) {
  if (!(p > 0.0 && p < 1.0)) {
    if (p == 0.0) return -HUGE_VAL;
    if (p == 1.0) return HUGE_VAL;
    return NAN;
  }
  return g4::loader::inverse_normal(p);
}


g4h_status g4h_load_beam(const g4h_beam_spec* sp, const double* current, double* x, double* y,
                         double* px, double* py, double* gamma, double* theta) {
  g4::loader::Spec ls;
  if (!x || !y || !px || !py || !gamma || !theta || (sp && sp->shotnoise && !current) ||
      !g4::loader_spec(sp, ls)) {
    g4::set_last_error("g4h_load_beam: invalid specification");
    return G4H_ERR_INVALID_ARGUMENT;
  }
  const int nb = sp->npart / sp->nbins;
  // Slices are independent, so they are loaded by a pool of host threads;
  // the result does not depend on the thread count.
  auto load_slices = [&](int s0, int s1) {
    double th[64];
    for (int is = s0; is < s1; ++is) {
      const double cur = sp->shotnoise ? current[is] : 0.0;
      for (int j = 0; j < nb; ++j) {
        double xv, pxv, yv, pyv, gv;
        g4::loader::beamlet(ls, is, j, cur, G4H_CE, xv, pxv, yv, pyv, gv, th);
        const int64_t i0 = static_cast<int64_t>(sp->npart) * is + static_cast<int64_t>(j) * sp->nbins;
        for (int k = 0; k < sp->nbins; ++k) {
          x[i0 + k] = xv;
          px[i0 + k] = pxv;
          y[i0 + k] = yv;
          py[i0 + k] = pyv;
          gamma[i0 + k] = gv;
          theta[i0 + k] = th[k];
        }
      }
    }
  };
  int nthr = static_cast<int>(std::thread::hardware_concurrency());
  if (const char* e = std::getenv("OMP_NUM_THREADS")) nthr = std::atoi(e);
  nthr = std::max(1, std::min({nthr, 64, sp->nslice}));
  if (static_cast<int64_t>(sp->nslice) * sp->npart < (1 << 16)) nthr = 1;
  std::vector<std::thread> pool;
  for (int t = 0; t < nthr; ++t) {
    const int s0 = static_cast<int>(static_cast<int64_t>(sp->nslice) * t / nthr);
    const int s1 = static_cast<int>(static_cast<int64_t>(sp->nslice) * (t + 1) / nthr);
    pool.emplace_back(load_slices, s0, s1);
  }
  for (auto& th : pool) th.join();
  return G4H_OK;
}

g4h_status g4h_seed_gaussian(int nslice, int ngrid, double gridmax, double xlambda, double power,
                             double w0, double* re, double* im) {
  if (nslice < 1 || ngrid < 3 || gridmax <= 0 || xlambda <= 0 || power < 0 || w0 <= 0 || !re || !im) {
    g4::set_last_error("g4h_seed_gaussian: invalid arguments");
    return G4H_ERR_INVALID_ARGUMENT;
  }
  const double dgrid = 2 * gridmax / (ngrid - 1);
  const double ks = 2.0 * kPi / xlambda;
  const double scl = dgrid * G4H_EEV / ks;
  const size_t plane = static_cast<size_t>(ngrid) * ngrid;
  std::vector<double> shape(plane);
  double sum = 0.0;
  for (int iy = 0; iy < ngrid; ++iy) {
    const double yv = -gridmax + iy * dgrid;
    for (int ix = 0; ix < ngrid; ++ix) {
      const double xv = -gridmax + ix * dgrid;
      const double a = std::exp(-(xv * xv + yv * yv) / (w0 * w0));
      shape[static_cast<size_t>(iy) * ngrid + ix] = a;
      sum += a * a;
    }
  }
  // power = u0^2 sum(shape^2) scl^2 / vacimp  (DiagnosticCUDA.cpp:735)
  const double u0 = sum > 0 ? std::sqrt(power * G4H_VACIMP / (scl * scl * sum)) : 0.0;
  for (int s = 0; s < nslice; ++s)
    for (size_t e = 0; e < plane; ++e) {
      re[s * plane + e] = u0 * shape[e];
      im[s * plane + e] = 0.0;
    }
  return G4H_OK;
}

g4h_status g4h_lattice_fodo(double lambdau, const double* aw_mod, int nmod, int nper,
                            int quad_spacing, int quad_len, double k1, int output_step,
                            g4h_step* steps, int capacity, int* nsteps) {
  if (lambdau <= 0 || !aw_mod || nmod < 1 || nper < 1 || quad_len < 0 ||
      (quad_len > 0 && quad_spacing <= quad_len) || output_step < 1 || !nsteps ||
      (capacity > 0 && !steps)) {
    g4::set_last_error("g4h_lattice_fodo: invalid arguments");
    return G4H_ERR_INVALID_ARGUMENT;
  }
  const int n = nmod * nper;
  *nsteps = n;
  const double ku = 2.0 * kPi / lambdau;
  const int qoff = quad_len > 0 ? (quad_spacing - quad_len) / 2 : 0;
  for (int i = 0; i < n && i < capacity; ++i) {
    g4h_step st;
    std::memset(&st, 0, sizeof(st));
    st.delz = lambdau;
    st.z = (i + 1) * lambdau;
    st.aw = aw_mod[i / nper];
    st.ku = ku;
    st.ky = ku * ku;
    st.in_undulator = st.aw > 0 ? 1 : 0;
    if (quad_len > 0) {
      const int q = i / quad_spacing, r = i % quad_spacing;
      if (r >= qoff && r < qoff + quad_len) st.qf = (q % 2 == 0) ? k1 : -k1;
    }
    st.slippage = 1.0;
    st.outstep = ((i + 1) % output_step) == 0 ? 1 : 0;
    steps[i] = st;
  }
  return G4H_OK;
}

}  // extern "C"
