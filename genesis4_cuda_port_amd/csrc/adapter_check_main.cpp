// g4adapter_check - drives the C++ adapter layer (genesis_gpu.hpp) the way
// the reference's Gencore::run drives its CUDA classes, on a small
// time-dependent case, and prints the per-slice power of the first field and
// the bunching of every output record:
//
//   g4adapter_check <fp64|mixed> <adi|fft> <nfield 1|2> <fused 0|1> [sc]
//
// one line per record: z, then nslice powers, then nslice bunching factors,
// and with `sc` (short-range space charge through Beam::initEField: rmax 2,
// 40 shells, 2 harmonics, 3 azimuthal modes) the nslice getSCField values
// of the operator-by-operator loop.
// The case: 6 slices x 512 particles (4 mirrored bins, shot noise, seed
// 4242), gamma0 11350, 3 kA, emittance 0.4 um, beta 15 m, a planar undulator
// of 12 periods of 26 mm (aw 1.7 rms) with 1.5 m^-2 FODO quadrupoles, a 1e5 W
// Gaussian seed of 50 um waist on a grid of 33 (ADI) or 32 (FFT) points over
// +-0.3 mm; with two fields the second is the third harmonic, unseeded.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <vector>

#include "genesis_gpu.hpp"

int main(int argc, char** argv) {
  if (argc != 5 && !(argc == 6 && std::strcmp(argv[5], "sc") == 0)) {
    std::fprintf(stderr, "usage: %s <fp64|mixed> <adi|fft> <nfield 1|2> <fused 0|1> [sc]\n",
                 argv[0]);
    return 2;
  }
  const bool sc = argc == 6;
  const g4h_precision prec = std::strcmp(argv[1], "mixed") == 0 ? G4H_MIXED : G4H_FP64;
  const bool fft = std::strcmp(argv[2], "fft") == 0;
  const int nfield = std::atoi(argv[3]) == 2 ? 2 : 1;
  const bool fused = std::atoi(argv[4]) != 0;
  try {
    const int nslice = 6, npart = 512, ngrid = fft ? 32 : 33;
    const double lambdau = 0.026, aw = 1.7, gamma0 = 11350.0, gridmax = 0.3e-3;
    const double lambda0 = lambdau * (1 + aw * aw) / (2 * gamma0 * gamma0);

    g4h_beam_spec spec{};
    spec.gamma0 = gamma0;
    spec.espread_rel = 1e-4;
    spec.emitn_x = spec.emitn_y = 0.4e-6;
    spec.betax = spec.betay = 15.0;
    spec.slicelength = lambda0;
    spec.nslice = nslice;
    spec.npart = npart;
    spec.nbins = 4;
    spec.shotnoise = 1;
    spec.seed = 4242;
    std::vector<double> current(nslice, 3000.0);
    const size_t np = static_cast<size_t>(nslice) * npart;
    std::vector<double> x(np), y(np), px(np), py(np), g(np), th(np);
    g4gpu::check(g4h_load_beam(&spec, current.data(), x.data(), y.data(), px.data(), py.data(),
                               g.data(), th.data()));

    int nsteps = 0;
    const double awm[1] = {aw};
    g4gpu::check(g4h_lattice_fodo(lambdau, awm, 1, 12, 46, 3, 1.5, 1, nullptr, 0, &nsteps));
    std::vector<g4h_step> steps(static_cast<size_t>(nsteps));
    g4gpu::check(g4h_lattice_fodo(lambdau, awm, 1, 12, 46, 3, 1.5, 1, steps.data(), nsteps,
                                  &nsteps));

    const size_t cells = static_cast<size_t>(nslice) * ngrid * ngrid;
    std::vector<double> re(cells), im(cells);
    g4gpu::check(g4h_seed_gaussian(nslice, ngrid, gridmax, lambda0, 1e5, 50e-6, re.data(),
                                   im.data()));

    g4gpu::Context ctx(0);
    std::vector<int32_t> off(static_cast<size_t>(nslice) + 1);
    for (int s = 0; s <= nslice; ++s) off[static_cast<size_t>(s)] = s * npart;
    g4gpu::Beam beam(ctx, prec, off, current, gamma0, lambda0, lambda0);
    beam.pack(x, y, px, py, g, th);
    if (sc) beam.initEField(2.0, 40, 2, 1, lambda0, false);
    g4gpu::Field f1(ctx, prec, nslice, ngrid, gridmax, lambda0, lambda0, 1);
    f1.pack(re, im, 0);
    f1.initSolver(fft, false, 1.0, 1.0, 1.0);
    std::vector<g4gpu::Field*> fields{&f1};
    g4gpu::Field f3(ctx, prec, nslice, ngrid, gridmax, lambda0, lambda0, 3);
    if (nfield == 2) {
      std::vector<double> zero(cells, 0.0);
      f3.pack(zero, zero, 0);
      f3.initSolver(fft, false, 1.0, 1.0, 1.0);
      fields.push_back(&f3);
    }
    g4gpu::Undulator und(steps, gamma0);
    const auto hist = g4gpu::Gencore::run(&beam, &fields, &und, true, false, 1, fused);
    for (const g4gpu::DiagnosticRecord& r : hist) {
      std::printf("%.17g", r.z);
      for (int s = 0; s < nslice; ++s)
        std::printf(" %.17g", r.field[0][static_cast<size_t>(s) * G4H_FIELD_Q + G4H_FQ_POWER]);
      for (int s = 0; s < nslice; ++s)
        std::printf(" %.17g", r.beam[static_cast<size_t>(s) * G4H_BEAM_QN(1) + G4H_BEAM_Q]);
      if (sc && !fused)
        for (int s = 0; s < nslice; ++s) std::printf(" %.17g", r.sscfield[static_cast<size_t>(s)]);
      std::printf("\n");
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "g4adapter_check: %s\n", e.what());
    return 1;
  }
  return 0;
}
