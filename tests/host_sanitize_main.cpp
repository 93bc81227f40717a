// Host-code driver for the address and undefined-behaviour sanitizers (no
// GPU code): the deck and lattice parsers on the example deck and on
// malformed inputs, the host beam loader, Sobol sequence, seed field and
// lattice tables, and the oracle's run loop with both field solvers and two
// harmonics.  Built and run by tests/test_host_sanitizers.py:
//
//   host_sanitize_main <deck.in> <deck.lat>
//
// Prints "done" and exits 0 when every call returned normally; a sanitizer
// report ends the process with a nonzero status.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <vector>

#include "deck.hpp"
#include "g4oracle.h"

// The library's error slot lives in its ABI layer, which this driver does
// not link.
namespace g4 {
void set_last_error(const std::string& m) { std::fprintf(stderr, "error: %s\n", m.c_str()); }
}  // namespace g4

namespace {

std::string slurp(const char* path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error(std::string("cannot read ") + path);
  std::stringstream s;
  s << f.rdbuf();
  return s.str();
}

void ok(g4h_status s, const char* what) {
  if (s != G4H_OK) {
    std::fprintf(stderr, "%s: status %d\n", what, static_cast<int>(s));
    std::exit(1);
  }
}

void parsers(const std::string& deck, const std::string& lattice) {
  using namespace g4::deck;
  const auto nl = parse_namelists(deck);
  const auto lat = parse_lattice(lattice);
  const auto line = expand(lat, "FEL");
  StepOptions o;
  o.delz = 0.24;
  o.lambda0 = 0.15e-9;
  o.gamma0 = 27984.34;
  const auto steps = build_steps(line, o);
  const Twiss tw = match_section(steps, o.gamma0, 5.72);
  std::printf("namelists %zu steps %zu betax %.6g betay %.6g\n", nl.size(), steps.size(),
              tw.betax, tw.betay);
  const char* bad_decks[] = {"&setup\nrootname=\n", "&setup\n=3\n&end\n", "&a\n&b\n&end",
                             "key=1\n", "&x\nk=1,,=2;\n&end\n", "", "&", "&end\n"};
  for (const char* b : bad_decks) {
    try {
      parse_namelists(b);
    } catch (const std::exception&) {
    }
  }
  const char* bad_lattices[] = {"a: UNDULATOR = {lambdau=0.015, nwig=",
                                "a: LINE = {a};",
                                "b: LINE = {3*};",
                                "q: QUADRUPOLE = {l=0.1, k1=};",
                                "x: FOO = {};",
                                ": DRIFT = {l=1};",
                                "d: DRIFT = {l=1}; L: LINE = {2*d, e};",
                                "d: DRIFT = {l=1}; L: LINE = {-1*d};",
                                "c: CHICANE = {l=1, lb=0.2, ld=0.3, delay=1e-3};",
                                "u: UNDULATOR = {lambdau=0.02, nwig=0, aw=1};",
                                ":",
                                "{};"};
  for (const char* b : bad_lattices) {
    try {
      const auto m = parse_lattice(b);
      for (const auto& kv : m) {
        try {
          build_steps(expand(m, kv.first), o);
        } catch (const std::exception&) {
        }
      }
    } catch (const std::exception&) {
    }
  }
}

void loaders_and_oracle() {
  std::vector<double> sob(3 * 100);
  ok(g4h_sobol(3, 7, 100, sob.data()), "g4h_sobol");
  const int nslice = 4, npart = 256;
  const double lambdau = 0.026, aw = 1.7, gamma0 = 11350.0;
  const double lambda0 = lambdau * (1 + aw * aw) / (2 * gamma0 * gamma0);
  g4h_beam_spec spec{};
  spec.gamma0 = gamma0;
  spec.espread_rel = 1e-4;
  spec.emitn_x = spec.emitn_y = 0.4e-6;
  spec.betax = spec.betay = 15;
  spec.slicelength = lambda0;
  spec.nslice = nslice;
  spec.npart = npart;
  spec.nbins = 4;
  spec.shotnoise = 1;
  spec.seed = 99;
  std::vector<double> cur(nslice, 3000.0);
  const size_t np = static_cast<size_t>(nslice) * npart;
  std::vector<double> x(np), y(np), px(np), py(np), g(np), th(np);
  ok(g4h_load_beam(&spec, cur.data(), x.data(), y.data(), px.data(), py.data(), g.data(),
                   th.data()),
     "g4h_load_beam");
  int nsteps = 0;
  const double awm[1] = {aw};
  ok(g4h_lattice_fodo(lambdau, awm, 1, 12, 46, 3, 1.5, 1, nullptr, 0, &nsteps), "fodo size");
  std::vector<g4h_step> st(static_cast<size_t>(nsteps));
  ok(g4h_lattice_fodo(lambdau, awm, 1, 12, 46, 3, 1.5, 1, st.data(), nsteps, &nsteps), "fodo");
  std::printf("fc(h=3) %.12g\n", g4h_undulator_fc(aw, 0, 3));
  std::vector<int32_t> off(nslice + 1);
  for (int s = 0; s <= nslice; ++s) off[static_cast<size_t>(s)] = s * npart;
  for (int solver = 0; solver < 2; ++solver) {
    const int G = solver ? 32 : 33;
    const size_t cells = static_cast<size_t>(nslice) * G * G;
    std::vector<double> re(cells), im(cells);
    ok(g4h_seed_gaussian(nslice, G, 0.3e-3, lambda0, 1e5, 50e-6, re.data(), im.data()),
       "g4h_seed_gaussian");
    orc_field f{};
    f.nslice = nslice;
    f.ngrid = G;
    f.harm = 1;
    f.enabled = 1;
    f.gridmax = 0.3e-3;
    f.dgrid = 2 * f.gridmax / (G - 1);
    f.xks = 2 * M_PI / lambda0;
    f.xlambda = lambda0;
    f.slicelength = lambda0;
    f.re = re.data();
    f.im = im.data();
    auto bx = x, by = y, bpx = px, bpy = py, bg = g, bth = th;
    const int nharm = 2;
    std::vector<double> braw(static_cast<size_t>(1 + nsteps) * nslice * G4H_BEAM_RAW(nharm));
    std::vector<double> fraw(static_cast<size_t>(1 + nsteps) * nslice * 9);
    const int nrec = orc_run(nslice, off.data(), bx.data(), by.data(), bpx.data(), bpy.data(),
                             bg.data(), bth.data(), gamma0, lambda0, cur.data(), nullptr, &f,
                             solver, solver, 1.0, 1.0, 1.0, st.data(), nsteps, 1.0, 1, 0, nharm,
                             braw.data(), fraw.data());
    std::vector<double> pez(np, 1e-4);
    orc_beam_advance_ez(nslice, off.data(), bx.data(), by.data(), bpx.data(), bpy.data(),
                        bg.data(), bth.data(), &f, 1, &st[0], gamma0, st[0].delz, pez.data());
    std::printf("solver %d records %d\n", solver, nrec);
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s deck.in deck.lat\n", argv[0]);
    return 2;
  }
  try {
    parsers(slurp(argv[1]), slurp(argv[2]));
    loaders_and_oracle();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "unexpected exception: %s\n", e.what());
    return 1;
  }
  std::printf("done\n");
  return 0;
}
