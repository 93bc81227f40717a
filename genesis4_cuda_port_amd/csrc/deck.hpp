// deck.hpp - input deck and lattice file front-end (host only).
//
// Reads the namelist input deck the reference's driver takes
// (src/Main/Track.cpp:62-134 for &track, example test/fel.in) and the
// lattice file format of test/fel.lat, and turns them into the per-step
// table (g4h_step) the device loop consumes.  Upstream GENESIS 1.3, which
// owns this code in the reference, is not part of it; the semantics here are
// stated in each function.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "../../include/g4hip.h"

namespace g4 {
namespace deck {

struct Namelist {
  std::string name;                                       // lower case, without '&'
  std::vector<std::pair<std::string, std::string>> keys;  // lower-case key, raw value
  int line = 0;
};

// "&name" opens a namelist, "&end" closes it; "key = value" pairs one per
// line (several separated by ',' or ';' also accepted); '#' and '!' start
// comments.  Throws std::runtime_error with the line number on bad input.
std::vector<Namelist> parse_namelists(const std::string& text);

struct Element {
  std::string type;  // upper case: UNDULATOR, DRIFT, QUADRUPOLE, CORRECTOR, CHICANE,
                     // PHASESHIFTER, MARKER, LINE
  std::map<std::string, std::string> par;  // lower-case key -> raw value
  std::vector<std::pair<int, std::string>> line;  // LINE: (repeat, label)
};

// "label: TYPE = { key = value, ... };" and "label: LINE = { a, 3*b, ... };"
std::map<std::string, Element> parse_lattice(const std::string& text);

// The flat element sequence of a beamline (LINE labels expanded, repeats
// unrolled); throws on unknown labels or recursive lines.
std::vector<const Element*> expand(const std::map<std::string, Element>& lat,
                                   const std::string& beamline);

struct StepOptions {
  double delz = 0;       // integration step inside undulators [m]
  double lambda0 = 0;    // reference wavelength [m]
  double gamma0 = 0;     // reference energy
  int output_step = 1;   // record every output_step-th step
  double zstop = 1e9;    // truncate the lattice after this z [m]
};

// Per-step table of a flat beamline:
//  * UNDULATOR (lambdau, nwig, aw, helical, kx, ky, ax, ay, gradx, grady):
//    length nwig*lambdau cut into round(length/delz) (at least 1) equal
//    steps; ku = 2 pi/lambdau; natural focusing kx*ku^2, ky*ku^2 with
//    defaults (0, 1) planar and (0.5, 0.5) helical;
//  * DRIFT (l), QUADRUPOLE (l, k1, dx, dy): one step of length l, aw = 0;
//  * CORRECTOR (l, cx, cy): kicks on the last half of its step (l > 0) or of
//    the previous step (l = 0);
//  * CHICANE (l, lb, ld, delay): one step; the bending angle solves
//    delay = 2 lb (theta/sin(theta) - 1) + ld (1/cos(theta) - 1) * 2;
//  * PHASESHIFTER (l, phi): one step with the phase advance phi;
//  * MARKER (dumpfield, dumpbeam, sort, stop): flags of the step ending at it.
// Slippage of a step in reference wavelengths: dz (1 + aw^2)/(2 gamma0^2
// lambda0) (one per period in a resonant undulator), plus delay/lambda0 in
// a chicane.
std::vector<g4h_step> build_steps(const std::vector<const Element*>& line,
                                  const StepOptions& opt);

struct Twiss {
  double betax = 15, betay = 15, alphax = 0, alphay = 0;
};

// Periodic Twiss parameters of the section [0, zmatch] of a step table: the
// 2x2 transfer matrices of the reference particle's (x, x') and (y, y')
// motion under the per-step strengths of TrackBeamCUDA.cpp:169-179
// (x'' = -q/gammaz x, q = qf gamma0 + k aw^2/(gamma0 beta_par)), then
// beta = M12/sin(mu), alpha = (M11 - M22)/(2 sin(mu)).  Throws when the
// section has no periodic solution (|cos mu| >= 1).
Twiss match_section(const std::vector<g4h_step>& steps, double gamma0, double zmatch);

// Parse helpers (throw std::runtime_error naming the key).
double to_double(const std::string& key, const std::string& v);
int to_int(const std::string& key, const std::string& v);
bool to_bool(const std::string& key, const std::string& v);

}  // namespace deck
}  // namespace g4
