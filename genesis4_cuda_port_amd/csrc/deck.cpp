// deck.cpp - input deck and lattice file front-end (see deck.hpp).
#include "deck.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <functional>
#include <set>
#include <sstream>
#include <stdexcept>

namespace g4 {
namespace deck {
namespace {

constexpr double kPi = 3.14159265358979323846;

std::string lower(std::string s) {
  for (auto& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return s;
}
std::string upper(std::string s) {
  for (auto& c : s) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
  return s;
}
std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
  while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
  return s.substr(a, b - a);
}
std::string unquote(const std::string& v) {
  if (v.size() >= 2 && ((v.front() == '"' && v.back() == '"') ||
                        (v.front() == '\'' && v.back() == '\'')))
    return v.substr(1, v.size() - 2);
  return v;
}
std::string strip_comment(const std::string& line) {
  bool inq = false;
  char q = 0;
  for (size_t i = 0; i < line.size(); ++i) {
    const char c = line[i];
    if (inq) {
      if (c == q) inq = false;
    } else if (c == '"' || c == '\'') {
      inq = true;
      q = c;
    } else if (c == '#' || c == '!') {
      return line.substr(0, i);
    }
  }
  return line;
}
// split on separators outside quotes and braces
std::vector<std::string> split(const std::string& s, const char* seps) {
  std::vector<std::string> out;
  std::string cur;
  int depth = 0;
  bool inq = false;
  char q = 0;
  for (char c : s) {
    if (inq) {
      cur += c;
      if (c == q) inq = false;
      continue;
    }
    if (c == '"' || c == '\'') {
      inq = true;
      q = c;
      cur += c;
      continue;
    }
    if (c == '{') ++depth;
    if (c == '}') --depth;
    if (depth == 0 && std::strchr(seps, c)) {
      out.push_back(cur);
      cur.clear();
      continue;
    }
    cur += c;
  }
  out.push_back(cur);
  return out;
}

[[noreturn]] void fail(int line, const std::string& msg) {
  std::ostringstream o;
  if (line > 0) o << "line " << line << ": ";
  o << msg;
  throw std::runtime_error(o.str());
}

}  // namespace

double to_double(const std::string& key, const std::string& v) {
  const std::string t = trim(unquote(trim(v)));
  char* end = nullptr;
  const double d = std::strtod(t.c_str(), &end);
  if (t.empty() || end != t.c_str() + t.size())
    throw std::runtime_error("'" + key + "' expects a number, got '" + v + "'");
  return d;
}
int to_int(const std::string& key, const std::string& v) {
  const double d = to_double(key, v);
  if (d != std::floor(d) || std::fabs(d) > 2e9)
    throw std::runtime_error("'" + key + "' expects an integer, got '" + v + "'");
  return static_cast<int>(d);
}
bool to_bool(const std::string& key, const std::string& v) {
  const std::string t = lower(trim(unquote(trim(v))));
  if (t == "1" || t == "true" || t == "t" || t == ".true.") return true;
  if (t == "0" || t == "false" || t == "f" || t == ".false.") return false;
  throw std::runtime_error("'" + key + "' expects a boolean, got '" + v + "'");
}

std::vector<Namelist> parse_namelists(const std::string& text) {
  std::vector<Namelist> out;
  std::istringstream in(text);
  std::string raw;
  int ln = 0;
  bool open = false;
  while (std::getline(in, raw)) {
    ++ln;
    const std::string line = trim(strip_comment(raw));
    if (line.empty()) continue;
    if (line[0] == '&') {
      const std::string name = lower(trim(line.substr(1)));
      if (name == "end") {
        if (!open) fail(ln, "&end without an open namelist");
        open = false;
        continue;
      }
      if (open) fail(ln, "namelist &" + out.back().name + " not closed before &" + name);
      if (name.empty() || name.find_first_of(" \t=") != std::string::npos)
        fail(ln, "bad namelist header '" + line + "'");
      Namelist nl;
      nl.name = name;
      nl.line = ln;
      out.push_back(nl);
      open = true;
      continue;
    }
    if (!open) fail(ln, "'" + line + "' outside a namelist");
    for (const std::string& item : split(line, ",;")) {
      const std::string it = trim(item);
      if (it.empty()) continue;
      const size_t eq = it.find('=');
      if (eq == std::string::npos) fail(ln, "expected key = value, got '" + it + "'");
      const std::string key = lower(trim(it.substr(0, eq)));
      const std::string val = trim(it.substr(eq + 1));
      if (key.empty() || val.empty()) fail(ln, "expected key = value, got '" + it + "'");
      out.back().keys.emplace_back(key, val);
    }
  }
  if (open) fail(ln, "namelist &" + out.back().name + " not closed");
  return out;
}

std::map<std::string, Element> parse_lattice(const std::string& text) {
  // statements end with ';'
  std::string body;
  {
    std::istringstream in(text);
    std::string raw;
    while (std::getline(in, raw)) body += strip_comment(raw) + "\n";
  }
  std::map<std::string, Element> lat;
  for (const std::string& stmt0 : split(body, ";")) {
    const std::string stmt = trim(stmt0);
    if (stmt.empty()) continue;
    const size_t colon = stmt.find(':');
    const size_t eq = stmt.find('=');
    if (colon == std::string::npos || eq == std::string::npos || eq < colon)
      throw std::runtime_error("lattice: expected 'label: TYPE = {...}', got '" + stmt + "'");
    const std::string label = upper(trim(stmt.substr(0, colon)));
    std::string type = upper(trim(stmt.substr(colon + 1, eq - colon - 1)));
    std::string rest = trim(stmt.substr(eq + 1));
    if (label.empty()) throw std::runtime_error("lattice: element without a label in '" + stmt + "'");
    if (rest.size() < 2 || rest.front() != '{' || rest.back() != '}')
      throw std::runtime_error("lattice: element '" + label + "' needs a {...} body");
    rest = rest.substr(1, rest.size() - 2);
    if (type.size() > 4 && type.compare(0, 4, "UNDU") == 0) type = "UNDULATOR";
    if (type.size() > 4 && type.compare(0, 4, "QUAD") == 0) type = "QUADRUPOLE";
    if (type.size() > 4 && type.compare(0, 4, "DRIF") == 0) type = "DRIFT";
    if (type.size() > 4 && type.compare(0, 4, "CORR") == 0) type = "CORRECTOR";
    if (type.size() > 4 && type.compare(0, 4, "CHIC") == 0) type = "CHICANE";
    if (type.size() > 4 && type.compare(0, 4, "PHAS") == 0) type = "PHASESHIFTER";
    if (type.size() > 4 && type.compare(0, 4, "MARK") == 0) type = "MARKER";
    static const std::set<std::string> known = {"UNDULATOR", "DRIFT", "QUADRUPOLE", "CORRECTOR",
                                                "CHICANE", "PHASESHIFTER", "MARKER", "LINE"};
    if (!known.count(type))
      throw std::runtime_error("lattice: element '" + label + "' has unknown type '" + type + "'");
    Element el;
    el.type = type;
    for (const std::string& item0 : split(rest, ",")) {
      const std::string item = trim(item0);
      if (item.empty()) continue;
      if (type == "LINE") {
        int rep = 1;
        std::string name = item;
        const size_t star = item.find('*');
        if (star != std::string::npos) {
          rep = to_int("repeat of " + label, item.substr(0, star));
          name = trim(item.substr(star + 1));
          if (rep < 0) throw std::runtime_error("lattice: negative repeat in line '" + label + "'");
        }
        if (name.empty())
          throw std::runtime_error("lattice: empty element name in line '" + label + "'");
        el.line.emplace_back(rep, upper(name));
      } else {
        const size_t e = item.find('=');
        if (e == std::string::npos)
          throw std::runtime_error("lattice: expected key = value in '" + label + "'");
        el.par[lower(trim(item.substr(0, e)))] = trim(item.substr(e + 1));
      }
    }
    if (lat.count(label)) throw std::runtime_error("lattice: label '" + label + "' defined twice");
    lat[label] = el;
  }
  return lat;
}

std::vector<const Element*> expand(const std::map<std::string, Element>& lat,
                                   const std::string& beamline) {
  std::vector<const Element*> out;
  std::vector<std::string> stack;
  std::function<void(const std::string&)> walk = [&](const std::string& name) {
    const auto it = lat.find(name);
    if (it == lat.end()) throw std::runtime_error("lattice: unknown label '" + name + "'");
    if (it->second.type != "LINE") {
      out.push_back(&it->second);
      return;
    }
    if (std::find(stack.begin(), stack.end(), name) != stack.end())
      throw std::runtime_error("lattice: line '" + name + "' contains itself");
    stack.push_back(name);
    for (const auto& [rep, sub] : it->second.line)
      for (int k = 0; k < rep; ++k) walk(sub);
    stack.pop_back();
  };
  walk(upper(trim(beamline)));
  return out;
}

namespace {

double par(const Element& e, const char* key, double dflt) {
  const auto it = e.par.find(key);
  return it == e.par.end() ? dflt : to_double(key, it->second);
}
bool parb(const Element& e, const char* key, bool dflt) {
  const auto it = e.par.find(key);
  return it == e.par.end() ? dflt : to_bool(key, it->second);
}

// bending angle of a four-dipole chicane with path-length delay `delay`
double chicane_angle(double lb, double ld, double delay) {
  if (delay <= 0) return 0.0;
  if (lb <= 0) throw std::runtime_error("lattice: chicane with delay needs lb > 0");
  auto f = [&](double t) {
    return 2 * lb * (t / std::sin(t) - 1) + 2 * ld * (1 / std::cos(t) - 1) - delay;
  };
  double lo = 0.0, hi = 1.5;
  if (f(hi) < 0) throw std::runtime_error("lattice: chicane delay too large");
  for (int it = 0; it < 200; ++it) {
    const double m = 0.5 * (lo + hi);
    (f(m) > 0 ? hi : lo) = m;
  }
  return 0.5 * (lo + hi);
}

}  // namespace

std::vector<g4h_step> build_steps(const std::vector<const Element*>& line,
                                  const StepOptions& opt) {
  if (opt.delz <= 0 || opt.lambda0 <= 0 || opt.gamma0 <= 1 || opt.output_step < 1)
    throw std::runtime_error("lattice: delz, lambda0, gamma0 and output_step must be positive");
  const double slip_per_m = 1.0 / (2.0 * opt.gamma0 * opt.gamma0 * opt.lambda0);
  std::vector<g4h_step> steps;
  double z = 0.0;
  auto push = [&](g4h_step st) {
    z += st.delz;
    st.z = z;
    steps.push_back(st);
  };
  auto blank = [](double dz) {
    g4h_step st;
    std::memset(&st, 0, sizeof(st));
    st.delz = dz;
    return st;
  };
  bool stopped = false;
  for (const Element* e : line) {
    if (stopped || z >= opt.zstop) break;
    const std::string& t = e->type;
    if (t == "UNDULATOR") {
      const double lu = par(*e, "lambdau", 0.0);
      const int nwig = static_cast<int>(par(*e, "nwig", 0.0));
      const double aw = par(*e, "aw", 0.0);
      const bool hel = parb(*e, "helical", false);
      if (lu <= 0 || nwig < 0) throw std::runtime_error("lattice: undulator needs lambdau > 0");
      const double len = lu * nwig;
      if (len <= 0) continue;
      const int n = std::max(1, static_cast<int>(std::lround(len / opt.delz)));
      const double dz = len / n;
      const double ku = 2 * kPi / lu;
      const double kx = par(*e, "kx", hel ? 0.5 : 0.0), ky = par(*e, "ky", hel ? 0.5 : 1.0);
      for (int k = 0; k < n; ++k) {
        g4h_step st = blank(dz);
        st.aw = aw;
        st.ku = ku;
        st.kx = kx * ku * ku;
        st.ky = ky * ku * ku;
        st.ax = par(*e, "ax", 0.0);
        st.ay = par(*e, "ay", 0.0);
        st.gradx = par(*e, "gradx", 0.0);
        st.grady = par(*e, "grady", 0.0);
        st.helical = hel ? 1 : 0;
        st.in_undulator = aw > 0 ? 1 : 0;
        st.slippage = dz * (1 + aw * aw) * slip_per_m;
        push(st);
      }
    } else if (t == "DRIFT" || t == "QUADRUPOLE" || t == "PHASESHIFTER") {
      const double l = par(*e, "l", 0.0);
      if (l < 0) throw std::runtime_error("lattice: negative element length");
      if (l == 0 && t != "PHASESHIFTER") continue;
      g4h_step st = blank(l);
      if (t == "QUADRUPOLE") {
        st.qf = par(*e, "k1", 0.0);
        st.dqx = par(*e, "dx", 0.0);
        st.dqy = par(*e, "dy", 0.0);
      }
      if (t == "PHASESHIFTER") st.autophase = par(*e, "phi", 0.0);
      st.slippage = l * slip_per_m;
      if (l == 0 && !steps.empty()) {
        steps.back().autophase += st.autophase;
        continue;
      }
      push(st);
    } else if (t == "CORRECTOR") {
      const double l = par(*e, "l", 0.0);
      const double cx = par(*e, "cx", 0.0), cy = par(*e, "cy", 0.0);
      if (l > 0) {
        g4h_step st = blank(l);
        st.cx = cx;
        st.cy = cy;
        st.slippage = l * slip_per_m;
        push(st);
      } else if (!steps.empty()) {
        steps.back().cx += cx;
        steps.back().cy += cy;
      }
    } else if (t == "CHICANE") {
      const double l = par(*e, "l", 0.0), lb = par(*e, "lb", 0.0), ld = par(*e, "ld", 0.0);
      const double delay = par(*e, "delay", 0.0);
      if (l <= 0 || l < 4 * lb + 2 * ld)
        throw std::runtime_error("lattice: chicane needs l >= 4 lb + 2 ld > 0");
      g4h_step st = blank(l);
      st.chic_angle = chicane_angle(lb, ld, delay);
      st.chic_lb = lb;
      st.chic_ld = ld;
      st.chic_lt = l;
      st.slippage = l * slip_per_m + delay / opt.lambda0;
      push(st);
    } else if (t == "MARKER") {
      if (steps.empty()) continue;
      int m = 0;
      if (par(*e, "dumpfield", 0.0) != 0) m |= 1;
      if (par(*e, "dumpbeam", 0.0) != 0) m |= 2;
      if (par(*e, "sort", 0.0) != 0) m |= 4;
      if (par(*e, "stop", 0.0) != 0) {
        m |= 8;
        stopped = true;
      }
      steps.back().marker |= m;
    }
  }
  for (size_t i = 0; i < steps.size(); ++i)
    steps[i].outstep = ((i + 1) % static_cast<size_t>(opt.output_step)) == 0 ? 1 : 0;
  if (!steps.empty()) steps.back().outstep = 1;
  return steps;
}

Twiss match_section(const std::vector<g4h_step>& steps, double gamma0, double zmatch) {
  double mx[4] = {1, 0, 0, 1}, my[4] = {1, 0, 0, 1};
  auto mul = [](double* m, double a, double b, double c, double d) {
    const double r0 = a * m[0] + b * m[2], r1 = a * m[1] + b * m[3];
    const double r2 = c * m[0] + d * m[2], r3 = c * m[1] + d * m[3];
    m[0] = r0;
    m[1] = r1;
    m[2] = r2;
    m[3] = r3;
  };
  auto plane = [&](double* m, double kq, double dz) {
    if (kq > 0) {
      const double f = std::sqrt(kq), w = f * dz;
      mul(m, std::cos(w), std::sin(w) / f, -f * std::sin(w), std::cos(w));
    } else if (kq < 0) {
      const double f = std::sqrt(-kq), w = f * dz;
      mul(m, std::cosh(w), std::sinh(w) / f, f * std::sinh(w), std::cosh(w));
    } else {
      mul(m, 1, dz, 0, 1);
    }
  };
  for (const g4h_step& st : steps) {
    if (st.z > zmatch + 1e-12) break;
    const double aw2 = st.aw * st.aw;
    const double gz = std::sqrt(gamma0 * gamma0 - 1 - aw2);
    const double bpar = std::sqrt(1 - (1 + aw2) / (gamma0 * gamma0));
    const double qx = st.qf * gamma0 + st.kx * aw2 / (gamma0 * bpar);
    const double qy = -st.qf * gamma0 + st.ky * aw2 / (gamma0 * bpar);
    plane(mx, qx / gz, st.delz);
    plane(my, qy / gz, st.delz);
  }
  const double cx = 0.5 * (mx[0] + mx[3]), cy = 0.5 * (my[0] + my[3]);
  if (std::fabs(cx) >= 1 || std::fabs(cy) >= 1)
    throw std::runtime_error("zmatch: the section has no periodic solution");
  const double sx = std::copysign(std::sqrt(1 - cx * cx), mx[1]);
  const double sy = std::copysign(std::sqrt(1 - cy * cy), my[1]);
  Twiss t;
  t.betax = mx[1] / sx;
  t.betay = my[1] / sy;
  t.alphax = (mx[0] - mx[3]) / (2 * sx);
  t.alphay = (my[0] - my[3]) / (2 * sy);
  return t;
}

}  // namespace deck
}  // namespace g4
