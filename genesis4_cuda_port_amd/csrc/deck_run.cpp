// deck_run.cpp - runs an input deck end to end on the GPU and writes the
// per-step records as .npy arrays under the names of the reference's output
// (DiagnosticCUDA.cpp:453-507 beam, :665-683 field).
//
// Namelists are executed in file order like the reference's driver: &setup
// must come first; &lattice, &time, &field and &beam set up state; every
// &track runs the lattice on the current beam and field (a second &track
// continues from where the first stopped).  Keys a namelist does not know
// are an error, as in Track::init (src/Main/Track.cpp:100-106).
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "deck.hpp"

namespace g4 {
void set_last_error(const std::string& m);
}

namespace {

using g4::deck::Namelist;
using g4::deck::to_bool;
using g4::deck::to_double;
using g4::deck::to_int;

constexpr double kPi = 3.14159265358979323846;

std::string read_file(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot read '" + path + "'");
  std::ostringstream s;
  s << in.rdbuf();
  return s.str();
}

std::string dir_of(const std::string& path) {
  const size_t k = path.find_last_of('/');
  return k == std::string::npos ? std::string(".") : path.substr(0, k);
}

void make_dirs(const std::string& path) {
  std::string cur;
  std::istringstream in(path);
  std::string part;
  if (!path.empty() && path[0] == '/') cur = "/";
  while (std::getline(in, part, '/')) {
    if (part.empty()) continue;
    cur += part + "/";
    if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
      throw std::runtime_error("cannot create directory '" + cur + "'");
  }
}

// NumPy .npy v1.0, little-endian float64, C order.
void write_npy(const std::string& path, const std::vector<double>& v, int rows, int cols) {
  std::string hdr = "{'descr': '<f8', 'fortran_order': False, 'shape': (" +
                    std::to_string(rows) + ", " + std::to_string(cols) + "), }";
  const size_t total = 10 + hdr.size() + 1;
  hdr.append((64 - total % 64) % 64, ' ');
  hdr += '\n';
  std::ofstream out(path, std::ios::binary);
  if (!out) throw std::runtime_error("cannot write '" + path + "'");
  const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  out.write(reinterpret_cast<const char*>(magic), 8);
  const uint16_t hl = static_cast<uint16_t>(hdr.size());
  out.write(reinterpret_cast<const char*>(&hl), 2);
  out.write(hdr.data(), static_cast<std::streamsize>(hdr.size()));
  out.write(reinterpret_cast<const char*>(v.data()),
            static_cast<std::streamsize>(v.size() * sizeof(double)));
  if (!out) throw std::runtime_error("cannot write '" + path + "'");
}

// An ABI call that failed: keeps its status for g4h_deck_run's return.
struct StatusError : std::runtime_error {
  g4h_status status;
  StatusError(g4h_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};
// Deck features this build does not serve.
struct Unsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void check(g4h_status st) {
  if (st != G4H_OK) throw StatusError(st, g4h_last_error());
}

struct Keys {
  const Namelist& nl;
  std::set<std::string> used;
  explicit Keys(const Namelist& n) : nl(n) {}
  const std::string* get(const char* k) {
    const std::string* r = nullptr;
    for (const auto& kv : nl.keys)
      if (kv.first == k) r = &kv.second;  // the last assignment wins
    if (r) used.insert(k);
    return r;
  }
  double d(const char* k, double dflt) {
    const std::string* v = get(k);
    return v ? to_double(k, *v) : dflt;
  }
  int i(const char* k, int dflt) {
    const std::string* v = get(k);
    return v ? to_int(k, *v) : dflt;
  }
  bool b(const char* k, bool dflt) {
    const std::string* v = get(k);
    return v ? to_bool(k, *v) : dflt;
  }
  std::string s(const char* k, const std::string& dflt) {
    const std::string* v = get(k);
    if (!v) return dflt;
    std::string t = *v;
    if (t.size() >= 2 && (t[0] == '"' || t[0] == '\'')) t = t.substr(1, t.size() - 2);
    return t;
  }
  // Unknown keys are an error (Track.cpp:100-106).
  void finish() {
    for (const auto& kv : nl.keys)
      if (!used.count(kv.first))
        throw std::runtime_error("unknown key '" + kv.first + "' in &" + nl.name);
  }
};

struct Setup {
  std::string rootname = "g4run", lattice, beamline;
  double gamma0 = 0, lambda0 = 0, delz = 0;
  uint64_t seed = 123456789;
  int npart = 8192, nbins = 4;
  bool shotnoise = true;
};

struct RunState {
  g4h_context* ctx = nullptr;
  g4h_beam* beam = nullptr;
  g4h_field* field = nullptr;
  ~RunState() {
    if (field) g4h_field_destroy(field);
    if (beam) g4h_beam_destroy(beam);
    if (ctx) g4h_context_destroy(ctx);
  }
};

struct Deck {
  Setup setup;
  g4::deck::Twiss twiss;
  double zmatch = 0;
  bool time = false;
  double slen = 0, s0 = 0;
  int sample = 1;
  // field
  bool have_field = false;
  double power = 0, phase = 0, waist_pos = 0, waist_size = 0, gridmax = 0;
  int ngrid = 151, harm = 1;
  // beam
  bool have_beam = false;
  double gamma = 0, delgam = 0, current = 0, ex = 0, ey = 0;
  double xc = 0, yc = 0, pxc = 0, pyc = 0;
  g4::deck::Twiss btw;
  g4h_precision prec = G4H_FP64;
  int nslice = 1;
  std::vector<int32_t> offsets;
  std::vector<double> cur;
  double zdone = 0;  // lattice position reached by earlier &track namelists
  int stepsdone = 0;  // integration steps of earlier &track namelists
  int nrecords = 0;
  int ndumps = 0;
  g4h_efield ef{};  // &efield: short-range space charge
};

void do_setup(Deck& d, const Namelist& nl) {
  Keys k(nl);
  Setup& s = d.setup;
  s.rootname = k.s("rootname", s.rootname);
  s.lattice = k.s("lattice", s.lattice);
  s.beamline = k.s("beamline", s.beamline);
  s.gamma0 = k.d("gamma0", s.gamma0);
  s.lambda0 = k.d("lambda0", s.lambda0);
  s.delz = k.d("delz", s.delz);
  s.seed = static_cast<uint64_t>(k.d("seed", static_cast<double>(s.seed)));
  s.npart = k.i("npart", s.npart);
  s.nbins = k.i("nbins", s.nbins);
  s.shotnoise = k.b("shotnoise", s.shotnoise);
  if (k.b("one4one", false)) throw Unsupported("&setup one4one is not supported");
  k.b("beam_global_stat", true);   // global records are always written
  k.b("field_global_stat", true);
  k.finish();
  if (s.gamma0 <= 1 || s.lambda0 <= 0 || s.delz <= 0)
    throw std::runtime_error("&setup needs gamma0 > 1, lambda0 > 0 and delz > 0");
  if (s.npart < 1 || s.nbins < 1 || s.npart % s.nbins)
    throw std::runtime_error("&setup: npart must be a positive multiple of nbins");
}

std::vector<g4h_step> lattice_steps(const Deck& d, const std::string& base, int output_step,
                                    double zstop) {
  if (d.setup.lattice.empty() || d.setup.beamline.empty())
    throw std::runtime_error("&setup needs lattice and beamline");
  std::string path = d.setup.lattice;
  if (path[0] != '/') path = base + "/" + path;
  const auto lat = g4::deck::parse_lattice(read_file(path));
  g4::deck::StepOptions o;
  o.delz = d.setup.delz;
  o.lambda0 = d.setup.lambda0;
  o.gamma0 = d.setup.gamma0;
  o.output_step = output_step;
  o.zstop = zstop;
  return g4::deck::build_steps(g4::deck::expand(lat, d.setup.beamline), o);
}

void do_field(Deck& d, const Namelist& nl, RunState& rs) {
  Keys k(nl);
  d.power = k.d("power", 0.0);
  d.phase = k.d("phase", 0.0);
  d.waist_pos = k.d("waist_position", 0.0);
  d.waist_size = k.d("waist_size", 100e-6);
  d.gridmax = k.d("dgrid", 1e-3);
  d.ngrid = k.i("ngrid", 151);
  d.harm = k.i("harm", 1);
  if (k.i("nx", 0) != 0 || k.i("ny", 0) != 0)
    throw Unsupported("&field: only the fundamental Gaussian mode (nx = ny = 0)");
  if (k.b("accumulate", false)) throw Unsupported("&field accumulate is not supported");
  k.finish();
  if (d.harm != 1) throw Unsupported("&field: harmonic fields are not supported");
  if (d.ngrid < 3 || d.gridmax <= 0) throw std::runtime_error("&field: bad ngrid or dgrid");
  d.have_field = true;
  (void)rs;
}

void do_beam(Deck& d, const Namelist& nl) {
  Keys k(nl);
  d.gamma = k.d("gamma", d.setup.gamma0);
  d.delgam = k.d("delgam", 0.0);
  d.current = k.d("current", 1000.0);
  d.ex = k.d("ex", 0.3e-6);
  d.ey = k.d("ey", 0.3e-6);
  d.btw.betax = k.d("betax", d.twiss.betax);
  d.btw.betay = k.d("betay", d.twiss.betay);
  d.btw.alphax = k.d("alphax", d.twiss.alphax);
  d.btw.alphay = k.d("alphay", d.twiss.alphay);
  d.xc = k.d("xcenter", 0.0);
  d.yc = k.d("ycenter", 0.0);
  d.pxc = k.d("pxcenter", 0.0);
  d.pyc = k.d("pycenter", 0.0);
  if (k.d("bunch", 0.0) != 0 || k.d("emod", 0.0) != 0)
    throw Unsupported("&beam: bunch and emod are not supported");
  k.finish();
  if (d.gamma <= 1 || d.current < 0 || d.btw.betax <= 0 || d.btw.betay <= 0)
    throw std::runtime_error("&beam: bad gamma, current or beta");
  d.have_beam = true;
}

// Gaussian seed with its waist at z = waist_position, evaluated at z = 0:
// u = q0/q exp(i ks r^2/(2 q)), q = -waist_position - i zR (the paraxial
// field equation 2 i ks du/dz + lap u = 0), normalised to `power` per slice.
void seed_field(const Deck& d, std::vector<double>& re, std::vector<double>& im) {
  const int G = d.ngrid, S = d.nslice;
  const double xlambda = d.setup.lambda0 / d.harm;
  const double ks = 2 * kPi / xlambda;
  const double dg = 2 * d.gridmax / (G - 1);
  re.assign(static_cast<size_t>(S) * G * G, 0.0);
  im.assign(re.size(), 0.0);
  if (d.power <= 0) return;
  const double w0 = d.waist_size;
  const double zr = 0.5 * ks * w0 * w0;
  const std::complex<double> q0(0, -zr), q(-d.waist_pos, -zr);
  std::vector<std::complex<double>> u(static_cast<size_t>(G) * G);
  double p = 0;
  for (int iy = 0; iy < G; ++iy)
    for (int ix = 0; ix < G; ++ix) {
      const double x = -d.gridmax + ix * dg, y = -d.gridmax + iy * dg;
      const std::complex<double> v =
          q0 / q * std::exp(std::complex<double>(0, ks * (x * x + y * y)) / (2.0 * q)) *
          std::polar(1.0, d.phase);
      u[static_cast<size_t>(iy) * G + ix] = v;
      p += std::norm(v);
    }
  const double eev = 510999.06, vacimp = 376.73;
  const double scale = std::sqrt(d.power / (p * std::pow(dg * eev / ks, 2) / vacimp));
  for (int s = 0; s < S; ++s)
    for (size_t c = 0; c < u.size(); ++c) {
      re[static_cast<size_t>(s) * G * G + c] = scale * u[c].real();
      im[static_cast<size_t>(s) * G * G + c] = scale * u[c].imag();
    }
}

void create_state(Deck& d, RunState& rs, int device) {
  if (!d.have_beam || !d.have_field)
    throw std::runtime_error("&track needs &field and &beam before it");
  if (!rs.ctx) check(g4h_context_create(device, &rs.ctx));
  const double slicelength = d.setup.lambda0 * d.sample;
  d.nslice = 1;
  if (d.time) {
    d.nslice = static_cast<int>(std::lround(d.slen / slicelength));
    if (d.nslice < 1) throw std::runtime_error("&time: slen shorter than one slice");
  }
  const int P = d.setup.npart;
  d.offsets.resize(static_cast<size_t>(d.nslice) + 1);
  for (int s = 0; s <= d.nslice; ++s) d.offsets[static_cast<size_t>(s)] = s * P;
  d.cur.assign(static_cast<size_t>(d.nslice), d.current);
  g4h_beam_spec spec;
  std::memset(&spec, 0, sizeof(spec));
  spec.gamma0 = d.gamma;
  spec.espread_rel = d.delgam / d.gamma;
  spec.emitn_x = d.ex;
  spec.emitn_y = d.ey;
  spec.betax = d.btw.betax;
  spec.betay = d.btw.betay;
  spec.alphax = d.btw.alphax;
  spec.alphay = d.btw.alphay;
  spec.xcen = d.xc;
  spec.ycen = d.yc;
  spec.pxcen = d.pxc;
  spec.pycen = d.pyc;
  spec.slicelength = slicelength;
  spec.nslice = d.nslice;
  spec.npart = P;
  spec.nbins = d.setup.nbins;
  spec.shotnoise = d.setup.shotnoise ? 1 : 0;
  spec.seed = d.setup.seed;
  const size_t n = static_cast<size_t>(d.nslice) * P;
  std::vector<double> x(n), y(n), px(n), py(n), g(n), th(n);
  check(g4h_load_beam(&spec, d.cur.data(), x.data(), y.data(), px.data(), py.data(), g.data(),
                      th.data()));
  check(g4h_beam_create(rs.ctx, d.prec, d.nslice, d.offsets.data(), d.cur.data(),
                        d.setup.gamma0, &rs.beam));
  check(g4h_beam_upload(rs.beam, x.data(), y.data(), px.data(), py.data(), g.data(), th.data()));
  if (d.ef.nz > 0) check(g4h_beam_init_efield(rs.beam, &d.ef, d.setup.lambda0));
  std::vector<double> re, im;
  seed_field(d, re, im);
  check(g4h_field_create(rs.ctx, d.prec, d.nslice, d.ngrid, d.gridmax, d.setup.lambda0,
                         slicelength, d.harm, &rs.field));
  check(g4h_field_upload(rs.field, re.data(), im.data(), 0));
}

// State after integration step `gstep` (lattice position z) in
// out/Dump/<gstep>/: the field in logical slice order (rows: slices,
// columns: cells ix + iy*ngrid) as field_re.npy / field_im.npy, and the
// particles (rows: slices, columns: particles of the slice) as
// beam_<x|y|px|py|gamma|theta>.npy, with z and the kinds in dump.json.
void write_dump(Deck& d, RunState& rs, const std::string& out, int gstep, double z, bool field,
                bool beam) {
  const std::string dir = out + "/Dump/" + std::to_string(gstep);
  make_dirs(dir);
  const int S = d.nslice;
  if (field) {
    const size_t plane = static_cast<size_t>(d.ngrid) * d.ngrid;
    std::vector<double> re(plane * S), im(plane * S), lre(plane * S), lim(plane * S);
    int first = 0;
    check(g4h_field_download(rs.field, re.data(), im.data(), &first));
    for (int is = 0; is < S; ++is) {
      const size_t p = static_cast<size_t>((is + first) % S) * plane;
      std::copy(re.begin() + p, re.begin() + p + plane, lre.begin() + static_cast<size_t>(is) * plane);
      std::copy(im.begin() + p, im.begin() + p + plane, lim.begin() + static_cast<size_t>(is) * plane);
    }
    write_npy(dir + "/field_re.npy", lre, S, static_cast<int>(plane));
    write_npy(dir + "/field_im.npy", lim, S, static_cast<int>(plane));
  }
  if (beam) {
    const size_t np = static_cast<size_t>(d.setup.npart) * S;
    std::vector<double> a[6];
    for (auto& v : a) v.resize(np);
    check(g4h_beam_download(rs.beam, a[0].data(), a[1].data(), a[2].data(), a[3].data(),
                            a[4].data(), a[5].data()));
    const char* names[6] = {"x", "y", "px", "py", "gamma", "theta"};
    for (int c = 0; c < 6; ++c)
      write_npy(dir + "/beam_" + names[c] + ".npy", a[c], S, d.setup.npart);
  }
  std::ofstream js(dir + "/dump.json");
  js << "{\"step\": " << gstep << ", \"z\": " << std::setprecision(17) << z
     << ", \"field\": " << (field ? "true" : "false") << ", \"beam\": " << (beam ? "true" : "false")
     << ", \"nslice\": " << S << ", \"ngrid\": " << d.ngrid << ", \"npart\": " << d.setup.npart
     << "}\n";
  ++d.ndumps;
}

void write_records(const Deck& d, const std::string& out, const std::vector<double>& zs,
                   const std::vector<double>& braw, const std::vector<double>& fraw, int nrec,
                   int nharm, bool exclharm, const std::vector<double>& sc,
                   const std::vector<g4h_step>& steps) {
  const int S = d.nslice;
  const int bq = G4H_BEAM_QN(nharm);
  std::vector<double> q(static_cast<size_t>(nrec) * S * bq), bg(static_cast<size_t>(nrec) * G4H_BEAM_GLOBAL);
  std::vector<double> fq(static_cast<size_t>(nrec) * S * G4H_FIELD_Q),
      fg(static_cast<size_t>(nrec) * G4H_FIELD_GLOBAL);
  const double dg = 2 * d.gridmax / (d.ngrid - 1);
  for (int r = 0; r < nrec; ++r) {
    check(g4h_diag_finalize_beam(S, nharm, braw.data() + static_cast<size_t>(r) * S * G4H_BEAM_RAW(nharm),
                                 d.cur.data(), d.setup.gamma0, q.data() + static_cast<size_t>(r) * S * bq,
                                 bg.data() + static_cast<size_t>(r) * G4H_BEAM_GLOBAL));
    check(g4h_diag_finalize_field(S, dg, d.setup.lambda0 / d.harm, d.setup.lambda0 * d.sample,
                                  fraw.data() + static_cast<size_t>(r) * S * G4H_FIELD_RAW,
                                  fq.data() + static_cast<size_t>(r) * S * G4H_FIELD_Q,
                                  fg.data() + static_cast<size_t>(r) * G4H_FIELD_GLOBAL));
  }
  static const char* bnames[G4H_BEAM_Q] = {
      "energy", "energyspread", "xposition", "xsize",  "yposition", "ysize", "pxposition",
      "pyposition", "current", "emitx", "emity", "betax", "betay", "alphax", "alphay", "xmin",
      "xmax", "pxmin", "pxmax", "ymin", "ymax", "pymin", "pymax", "emin", "emax"};
  static const char* fnames[G4H_FIELD_Q] = {"power", "xposition", "xsize", "yposition", "ysize",
                                             "intensity-nearfield", "phase-nearfield",
                                             "intensity-farfield", "phase-farfield"};
  static const char* bgn[G4H_BEAM_GLOBAL] = {"energy", "energyspread", "xposition",
                                             "xsize", "yposition", "ysize"};
  static const char* fgn[G4H_FIELD_GLOBAL] = {"energy", "xposition", "xsize", "yposition",
                                              "ysize", "intensity-nearfield",
                                              "intensity-farfield"};
  make_dirs(out + "/Beam/Global");
  make_dirs(out + "/Field/Global");
  make_dirs(out + "/Lattice");
  auto column = [&](const std::vector<double>& a, int width, int col) {
    std::vector<double> v(static_cast<size_t>(nrec) * S);
    for (int r = 0; r < nrec; ++r)
      for (int s = 0; s < S; ++s)
        v[static_cast<size_t>(r) * S + s] = a[(static_cast<size_t>(r) * S + s) * width + col];
    return v;
  };
  for (int c = 0; c < G4H_BEAM_Q; ++c)
    write_npy(out + "/Beam/" + bnames[c] + ".npy", column(q, bq, c), nrec, S);
  // &track exclusive_harmonics: the fundamental and the highest harmonic
  // only (DiagnosticCUDA.cpp:495-507)
  for (int h = 0; h < nharm; ++h) {
    if (exclharm && h > 0 && h < nharm - 1) continue;
    const std::string suf = h == 0 ? "" : std::to_string(h + 1);
    write_npy(out + "/Beam/bunching" + suf + ".npy", column(q, bq, G4H_BEAM_Q + 2 * h), nrec, S);
    write_npy(out + "/Beam/bunchingphase" + suf + ".npy", column(q, bq, G4H_BEAM_Q + 2 * h + 1),
              nrec, S);
  }
  // longitudinal fields of the auxiliary output (DiagnosticCUDA.cpp:558-570):
  // wakes (eloss) and the long-range space charge (longESC) are not modelled
  // by this runner and are zero; sscfield is the device's short-range field
  {
    const std::vector<double> zero(static_cast<size_t>(nrec) * S, 0.0);
    write_npy(out + "/Beam/efield.npy", zero, nrec, S);
    write_npy(out + "/Beam/wakefield.npy", zero, nrec, S);
    write_npy(out + "/Beam/lscfield.npy", zero, nrec, S);
    write_npy(out + "/Beam/sscfield.npy", sc, nrec, S);
  }
  for (int c = 0; c < G4H_FIELD_Q; ++c)
    write_npy(out + "/Field/" + fnames[c] + ".npy", column(fq, G4H_FIELD_Q, c), nrec, S);
  auto gcol = [&](const std::vector<double>& a, int width, int col) {
    std::vector<double> v(static_cast<size_t>(nrec));
    for (int r = 0; r < nrec; ++r) v[static_cast<size_t>(r)] = a[static_cast<size_t>(r) * width + col];
    return v;
  };
  for (int c = 0; c < G4H_BEAM_GLOBAL; ++c)
    write_npy(out + "/Beam/Global/" + bgn[c] + ".npy", gcol(bg, G4H_BEAM_GLOBAL, c), nrec, 1);
  for (int c = 0; c < G4H_FIELD_GLOBAL; ++c)
    write_npy(out + "/Field/Global/" + fgn[c] + ".npy", gcol(fg, G4H_FIELD_GLOBAL, c), nrec, 1);
  write_npy(out + "/Global_zplot.npy", zs, nrec, 1);
  const int ns = static_cast<int>(steps.size());
  std::vector<double> lz(steps.size()), law(steps.size()), lqf(steps.size()), lslip(steps.size());
  for (int i = 0; i < ns; ++i) {
    lz[static_cast<size_t>(i)] = steps[static_cast<size_t>(i)].z;
    law[static_cast<size_t>(i)] = steps[static_cast<size_t>(i)].aw;
    lqf[static_cast<size_t>(i)] = steps[static_cast<size_t>(i)].qf;
    lslip[static_cast<size_t>(i)] = steps[static_cast<size_t>(i)].slippage;
  }
  write_npy(out + "/Lattice/z.npy", lz, ns, 1);
  write_npy(out + "/Lattice/aw.npy", law, ns, 1);
  write_npy(out + "/Lattice/qf.npy", lqf, ns, 1);
  write_npy(out + "/Lattice/slippage.npy", lslip, ns, 1);
  std::ofstream meta(out + "/meta.json");
  meta.precision(17);
  meta << "{\"rootname\": \"" << d.setup.rootname << "\", \"nslice\": " << S
       << ", \"nrecords\": " << nrec << ", \"ngrid\": " << d.ngrid
       << ", \"gridmax\": " << d.gridmax << ", \"lambda0\": " << d.setup.lambda0
       << ", \"sample\": " << d.sample << ", \"gamma0\": " << d.setup.gamma0
       << ", \"betax\": " << d.btw.betax << ", \"betay\": " << d.btw.betay
       << ", \"alphax\": " << d.btw.alphax << ", \"alphay\": " << d.btw.alphay
       << ", \"precision\": \"" << (d.prec == G4H_FP64 ? "fp64" : "mixed") << "\"}\n";
}

void do_track(Deck& d, const Namelist& nl, const std::string& base, const std::string& out,
              RunState& rs, int device) {
  Keys k(nl);
  const double zstop = k.d("zstop", 1e9);
  const int output_step = std::max(1, k.i("output_step", 1));
  // use_gpu, with use_cuda accepted as its alias (Track.cpp:19-20,79-80)
  bool use_gpu = true;
  if (k.get("use_cuda")) use_gpu = k.b("use_cuda", true);
  if (k.get("use_gpu")) use_gpu = k.b("use_gpu", true);
  const bool low_mem = k.b("low_memory_adisolver", false);
  const bool fft = k.b("fft_fieldsolver", false);
  const bool filt = k.b("source_filter", false);
  const double xc = k.d("xcut", 1.0), yc = k.d("ycut", 1.0), sig = k.d("sigmoid", 1.0);
  const bool periodic = k.b("periodic", false);
  const int nharm = std::max(1, k.i("bunchharm", 1));
  const std::string prec = k.s("precision", d.prec == G4H_FP64 ? "fp64" : "mixed");
  const int field_dump_step = k.i("field_dump_step", 0);
  const int beam_dump_step = k.i("beam_dump_step", 0);
  const bool dump_at_exit = k.b("field_dump_at_undexit", false);
  if (field_dump_step < 0 || beam_dump_step < 0)
    throw std::runtime_error("&track dump steps must be >= 0");
  if (k.i("sort_step", 0) != 0) throw Unsupported("&track: sorting is not supported");
  const bool exclharm = k.b("exclusive_harmonics", false);
  k.get("s0");
  k.get("slen");
  k.finish();
  if (!use_gpu)
    throw Unsupported("&track use_gpu = false: the CPU path is upstream GENESIS 1.3");
  if (nharm > 8) throw std::runtime_error("&track bunchharm must be at most 8");
  const g4h_precision want = prec == "fp64" ? G4H_FP64
                             : prec == "mixed" ? G4H_MIXED
                                               : throw std::runtime_error("&track precision is fp64 or mixed");
  if (rs.beam && want != d.prec)
    throw std::runtime_error("&track: precision cannot change between &track namelists");
  d.prec = want;
  if (!rs.beam) create_state(d, rs, device);
  check(g4h_field_init_solver(rs.field, fft ? G4H_SOLVER_FFT : G4H_SOLVER_ADI, filt ? 1 : 0, xc,
                              yc, sig, low_mem ? 1 : 0));
  std::vector<g4h_step> all = lattice_steps(d, base, output_step, zstop);
  std::vector<g4h_step> steps;
  for (const g4h_step& st : all)
    if (st.z > d.zdone + 1e-12) steps.push_back(st);
  if (steps.empty()) throw std::runtime_error("&track: no lattice steps left to track");
  g4h_run_config cfg;
  cfg.lambda0 = d.setup.lambda0;
  cfg.sample = d.sample;
  cfg.timerun = d.time ? 1 : 0;
  cfg.periodic = periodic ? 1 : 0;
  cfg.nharm = nharm;
  cfg.record_diag = 1;
  g4h_run* run = nullptr;
  check(g4h_run_create(rs.ctx, rs.beam, rs.field, steps.data(), static_cast<int>(steps.size()),
                       &cfg, &run));
  std::unique_ptr<g4h_run, void (*)(g4h_run*)> guard(run, g4h_run_destroy);
  // Dumps of the field and / or beam after a step: lattice markers (bit 1
  // field, bit 2 beam; Gencore.cpp:176-181 applies a step's marker before
  // the next step, and the last one after the loop, :276), every
  // field_dump_step / beam_dump_step integration steps, and the field at
  // each undulator exit.  The run advances to each such step, then the state
  // is copied out.  Sorting (marker bit 4) is not built.
  int done = 0, at = 0;
  const int ns = static_cast<int>(steps.size());
  for (int i = 0; i < ns; ++i) {
    const g4h_step& st = steps[static_cast<size_t>(i)];
    if (st.marker & 4) throw Unsupported("lattice marker sort = 1: sorting is not supported");
    const int gstep = d.stepsdone + i + 1;
    const bool exit_und =
        st.in_undulator && (i + 1 == ns || !steps[static_cast<size_t>(i) + 1].in_undulator);
    const bool fd = (st.marker & 1) || (field_dump_step > 0 && gstep % field_dump_step == 0) ||
                    (dump_at_exit && exit_und);
    const bool bd = (st.marker & 2) || (beam_dump_step > 0 && gstep % beam_dump_step == 0);
    if (!fd && !bd) continue;
    check(g4h_run_advance(run, i + 1 - at, &done));
    at = i + 1;
    write_dump(d, rs, out, gstep, st.z, fd, bd);
  }
  check(g4h_run_advance(run, ns - at, &done));
  int nrec = 0;
  check(g4h_run_records(run, &nrec));
  const int S = d.nslice;
  std::vector<double> braw(static_cast<size_t>(nrec) * S * G4H_BEAM_RAW(nharm));
  std::vector<double> fraw(static_cast<size_t>(nrec) * S * G4H_FIELD_RAW);
  check(g4h_run_download_diag(run, braw.data(), fraw.data()));
  std::vector<double> sc(static_cast<size_t>(nrec) * S);
  check(g4h_run_download_sc(run, sc.data()));
  std::vector<double> zs{d.zdone};
  for (const g4h_step& st : steps)
    if (st.outstep) zs.push_back(st.z);
  zs.resize(static_cast<size_t>(nrec));
  d.zdone = steps.back().z;
  d.stepsdone += ns;
  const std::string dir = out + (d.nrecords > 0 ? "/track" + std::to_string(d.nrecords) : "");
  write_records(d, dir, zs, braw, fraw, nrec, nharm, exclharm, sc, steps);
  d.nrecords += nrec;
}

}  // namespace

extern "C" {

g4h_status g4h_lattice_parse(const char* lattice_text, const char* beamline, double delz,
                             double lambda0, double gamma0, int output_step, double zstop,
                             g4h_step* steps, int capacity, int* nsteps) {
  try {
    if (!lattice_text || !beamline || !nsteps || (capacity > 0 && !steps))
      throw std::invalid_argument("g4h_lattice_parse: null argument");
    const auto lat = g4::deck::parse_lattice(lattice_text);
    g4::deck::StepOptions o;
    o.delz = delz;
    o.lambda0 = lambda0;
    o.gamma0 = gamma0;
    o.output_step = output_step;
    o.zstop = zstop;
    const auto v = g4::deck::build_steps(g4::deck::expand(lat, beamline), o);
    *nsteps = static_cast<int>(v.size());
    for (int i = 0; i < capacity && i < *nsteps; ++i) steps[i] = v[static_cast<size_t>(i)];
    return G4H_OK;
  } catch (const std::exception& e) {
    g4::set_last_error(e.what());
    return G4H_ERR_INVALID_ARGUMENT;
  }
}

g4h_status g4h_lattice_match(const g4h_step* steps, int nsteps, double gamma0, double zmatch,
                             double* twiss) {
  try {
    if (!steps || nsteps < 1 || !twiss || gamma0 <= 1 || zmatch <= 0)
      throw std::invalid_argument("g4h_lattice_match: invalid arguments");
    const g4::deck::Twiss t = g4::deck::match_section(
        std::vector<g4h_step>(steps, steps + nsteps), gamma0, zmatch);
    twiss[0] = t.betax;
    twiss[1] = t.alphax;
    twiss[2] = t.betay;
    twiss[3] = t.alphay;
    return G4H_OK;
  } catch (const std::exception& e) {
    g4::set_last_error(e.what());
    return G4H_ERR_INVALID_ARGUMENT;
  }
}

g4h_status g4h_deck_parse(const char* deck_text, char* json, int capacity, int* needed) {
  try {
    if (!deck_text || !needed || (capacity > 0 && !json))
      throw std::invalid_argument("g4h_deck_parse: null argument");
    const auto nls = g4::deck::parse_namelists(deck_text);
    std::string s = "[";
    auto esc = [](const std::string& v) {
      std::string o;
      for (char c : v) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
      }
      return o;
    };
    for (size_t i = 0; i < nls.size(); ++i) {
      s += (i ? ", " : "") + std::string("{\"name\": \"") + esc(nls[i].name) + "\", \"keys\": [";
      for (size_t j = 0; j < nls[i].keys.size(); ++j)
        s += (j ? ", " : "") + std::string("[\"") + esc(nls[i].keys[j].first) + "\", \"" +
             esc(nls[i].keys[j].second) + "\"]";
      s += "]}";
    }
    s += "]";
    *needed = static_cast<int>(s.size()) + 1;
    if (capacity > 0) {
      const size_t n = std::min(static_cast<size_t>(capacity) - 1, s.size());
      std::memcpy(json, s.data(), n);
      json[n] = 0;
    }
    return G4H_OK;
  } catch (const std::exception& e) {
    g4::set_last_error(e.what());
    return G4H_ERR_INVALID_ARGUMENT;
  }
}

g4h_status g4h_deck_run(const char* deck_path, const char* out_dir, int device, int* nrecords) {
  try {
    if (!deck_path || !out_dir) throw std::invalid_argument("g4h_deck_run: null argument");
    const std::string base = dir_of(deck_path);
    const auto nls = g4::deck::parse_namelists(read_file(deck_path));
    Deck d;
    RunState rs;
    bool have_setup = false;
    for (const Namelist& nl : nls) {
      if (nl.name == "setup") {
        if (have_setup) throw std::runtime_error("&setup given twice");
        do_setup(d, nl);
        have_setup = true;
        continue;
      }
      if (!have_setup) throw std::runtime_error("&setup must be the first namelist");
      if (nl.name == "lattice") {
        Keys k(nl);
        d.zmatch = k.d("zmatch", 0.0);
        k.finish();
        if (d.zmatch > 0)
          d.twiss = g4::deck::match_section(lattice_steps(d, base, 1, 1e9), d.setup.gamma0, d.zmatch);
      } else if (nl.name == "time") {
        Keys k(nl);
        d.s0 = k.d("s0", 0.0);
        d.slen = k.d("slen", 0.0);
        d.sample = k.i("sample", 1);
        d.time = k.b("time", true);
        k.finish();
        if (d.sample < 1) throw std::runtime_error("&time sample must be >= 1");
      } else if (nl.name == "field") {
        if (rs.beam) throw Unsupported("&field after &track is not supported");
        do_field(d, nl, rs);
      } else if (nl.name == "beam") {
        if (rs.beam) throw Unsupported("&beam after &track is not supported");
        do_beam(d, nl);
      } else if (nl.name == "efield") {
        // &efield (upstream GENESIS 1.3; Beam::initEField, include/Beam.h:126-131)
        Keys k(nl);
        d.ef.rmax = k.d("rmax", 0.0);
        d.ef.ngrid = k.i("ngrid", 100);
        d.ef.nz = k.i("nz", 0);
        d.ef.nphi = k.i("nphi", 0);
        d.ef.longrange = k.b("longrange", false) ? 1 : 0;
        k.finish();
        if (d.ef.longrange)
          throw Unsupported("&efield longrange: the long-range space-charge field is not built");
        if (rs.beam) check(g4h_beam_init_efield(rs.beam, &d.ef, d.setup.lambda0));
      } else if (nl.name == "track") {
        do_track(d, nl, base, out_dir, rs, device);
      } else {
        throw Unsupported("namelist &" + nl.name + " is not supported");
      }
    }
    if (nrecords) *nrecords = d.nrecords;
    return G4H_OK;
  } catch (const StatusError& e) {
    g4::set_last_error(std::string("g4h_deck_run: ") + e.what());
    return e.status;
  } catch (const Unsupported& e) {
    g4::set_last_error(std::string("g4h_deck_run: ") + e.what());
    return G4H_ERR_UNSUPPORTED;
  } catch (const std::exception& e) {
    g4::set_last_error(std::string("g4h_deck_run: ") + e.what());
    return G4H_ERR_INVALID_ARGUMENT;
  }
}

}  // extern "C"
