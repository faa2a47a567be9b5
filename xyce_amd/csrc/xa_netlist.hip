// xa_netlist.hip -- SPICE-netlist subset reader and circuit builder (host
// code).  The reader takes the element and control lines a device-load
// circuit needs -- M (BSIM4 level 14/54, MOSFET level 1), D, Q, R, C and V
// elements, .MODEL cards, .TRAN, .IC, .STEP and .END -- and the builder turns
// a parsed netlist into the library's device groups on one GPU: one group per
// device type and instance parameter set, node LIDs, store/state slots and
// the CSR structure of dF/dx and dQ/dx, ready for the loads.  With .STEP it
// can build every step's copy of the circuit into the same groups, so that
// one load evaluates all of them (the batched form).
//
// Reference behaviour kept: the first line is the title; `*` starts a
// comment line and `;` an end-of-line comment; `+` continues the previous
// line; names are case-insensitive; numbers take the SPICE scale suffixes
// (T G MEG K M MIL U N P F) and ignore the letters after them
// (N_UTL_Misc.C:197-236 Value, :239 isValue); node 0 is ground; the title
// line is read first (N_IO_CircuitBlock.C:483-511).  Device types follow the reference's
// registrations: "m" level 1 MOSFET1 (N_DEV_MOSFET1.C:4984-4996), levels 14
// and 54 BSIM4 (N_DEV_MOSFET_B4.C:17270-17280), "d" (N_DEV_Diode.C:2179),
// "q" NPN/PNP level 1 (N_DEV_BJT.C:4368), "r", "c", "v"
// (N_DEV_Resistor.C:1198, N_DEV_Capacitor.C:2152, N_DEV_Vsrc.C:1090).
// Independent source waveforms follow the Vsrc parameter table
// (N_DEV_Vsrc.C:74-210): SIN(V0 VA FREQ TD THETA PHASE), EXP(V1 V2 TD1
// TAU1 TD2 TAU2), PULSE(V1 V2 TD TR TF PW PER); DC value DCV0.
//
// Layout of a built circuit (the convention the tests reproduce): the
// netlist's nodes in order of first appearance take LIDs 0..n-1 (ground is
// -1); the groups are created in order of the first appearance of their
// device type and parameter set (the reference's DeviceMgr creates device
// types as the parser meets them, N_DEV_DeviceMgr.C:1206-1229, 1446-1470); each group's
// instances keep netlist order, and their internal nodes take fresh LIDs
// after the external ones, group by group, instance by instance; state and
// store slots likewise.  A batched .STEP build repeats the whole layout per
// step with every LID offset by the copy's unknown count.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>
#include "xa_internal.h"

namespace xa {
namespace nl {

static std::string upper(std::string s) {
  for (char& c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

// SPICE number: mantissa, optional scale suffix, trailing letters ignored
static bool parse_number(const std::string& tok, double* out) {
  const char* s = tok.c_str();
  char* end = nullptr;
  const double v = std::strtod(s, &end);
  if (end == s) return false;
  std::string rest = upper(std::string(end));
  double scale = 1.0;
  if (rest.compare(0, 3, "MEG") == 0) scale = 1e6;
  else if (rest.compare(0, 3, "MIL") == 0) scale = 25.4e-6;
  else if (!rest.empty()) {
    switch (rest[0]) {
      case 'T': scale = 1e12; break;
      case 'G': scale = 1e9; break;
      case 'K': scale = 1e3; break;
      case 'M': scale = 1e-3; break;
      case 'U': scale = 1e-6; break;
      case 'N': scale = 1e-9; break;
      case 'P': scale = 1e-12; break;
      case 'F': scale = 1e-15; break;
      default:
        if (!std::isalpha((unsigned char)rest[0])) return false;
    }
  }
  *out = v * scale;
  return true;
}

struct Model {
  std::string name;
  int type = 0;       // XA_* model type
  int level = 1;
  std::vector<std::pair<std::string, double>> params;
};

struct Device {
  char letter = 0;
  std::string name;
  std::vector<int> nodes;      // node ids, -1 ground
  std::string model_name;
  int model = -1;
  std::vector<std::pair<std::string, double>> params;
  int line = 0;
};

struct Step {
  std::string target, param;   // target: a model or device name
  std::vector<double> values;
};

struct Netlist {
  std::string title;
  std::vector<std::string> node_names;
  std::map<std::string, int> node_id;
  std::vector<Model> models;
  std::map<std::string, int> model_id;
  std::set<std::string> model_names;   // every .MODEL name, collected before the elements
  std::vector<Device> devices;
  bool tran = false;
  double tstep = 0, tstop = 0, tstart = 0, tmax = 0;
  std::vector<std::pair<int, double>> ics;
  std::vector<Step> steps;
  // parts: devices grouped by (type, instance parameter set), in order of
  // first appearance
  struct Part {
    int type = 0;
    std::vector<int> models;      // netlist model ids, in part order
    std::vector<int> devices;     // netlist device ids, netlist order
    std::vector<std::string> inames;
    std::vector<std::string> mnames_flat;   // string storage for the ABI
  };
  std::vector<Part> parts;
};

static int perr(int line, const std::string& msg) {
  return fail(XA_ERR_ARG, "netlist line " + std::to_string(line) + ": " + msg);
}

static int node_of(Netlist& n, const std::string& tok) {
  const std::string u = upper(tok);
  if (u == "0" || u == "GND") return -1;
  auto it = n.node_id.find(u);
  if (it != n.node_id.end()) return it->second;
  const int id = (int)n.node_names.size();
  n.node_names.push_back(u);
  n.node_id[u] = id;
  return id;
}

// tokens of one logical line: whitespace separated; '(' ')' ',' '=' are
// separate tokens
static std::vector<std::string> tokenize(const std::string& line) {
  std::vector<std::string> t;
  std::string cur;
  auto flush = [&]() { if (!cur.empty()) { t.push_back(cur); cur.clear(); } };
  for (char c : line) {
    if (std::isspace((unsigned char)c) || c == ',') { flush(); continue; }
    if (c == '(' || c == ')' || c == '=') { flush(); t.push_back(std::string(1, c)); continue; }
    cur += c;
  }
  flush();
  return t;
}

// name=value pairs from tokens[k..]
static int read_pairs(const std::vector<std::string>& t, size_t k, int line,
                      std::vector<std::pair<std::string, double>>& out) {
  while (k < t.size()) {
    if (t[k] == "(" || t[k] == ")") { ++k; continue; }
    if (k + 2 >= t.size() || t[k + 1] != "=") return perr(line, "expected name=value at " + t[k]);
    double v;
    if (!parse_number(t[k + 2], &v)) return perr(line, "bad number " + t[k + 2]);
    out.emplace_back(upper(t[k]), v);
    k += 3;
  }
  return XA_OK;
}

// positional tokens from t[k..]: up to the first name=value pair
static std::vector<std::string> positional(const std::vector<std::string>& t, size_t& k) {
  std::vector<std::string> p;
  while (k < t.size() && !(k + 1 < t.size() && t[k + 1] == "=")) p.push_back(t[k++]);
  return p;
}

static int parse_model(Netlist& n, const std::vector<std::string>& t, int line) {
  if (t.size() < 3) return perr(line, ".MODEL needs a name and a type");
  Model m;
  m.name = upper(t[1]);
  const std::string ty = upper(t[2]);
  std::vector<std::pair<std::string, double>> ps;
  int rc = read_pairs(t, 3, line, ps);
  if (rc != XA_OK) return rc;
  int level = 1;
  for (auto& p : ps)
    if (p.first == "LEVEL") level = (int)p.second;
  if (ty == "NMOS" || ty == "PMOS") {
    if (level == 14 || level == 54) m.type = XA_BSIM4;
    else if (level == 1) m.type = XA_MOSFET1;
    else return perr(line, "MOSFET level " + std::to_string(level) + " is not a device-load model");
    m.params.emplace_back("TYPE", ty == "NMOS" ? 1.0 : -1.0);
  } else if (ty == "D") {
    m.type = XA_DIODE;
  } else if (ty == "NPN" || ty == "PNP") {
    m.type = XA_BJT;
    m.params.emplace_back("TYPE", ty == "NPN" ? 1.0 : -1.0);
  } else if (ty == "R") {
    m.type = XA_RESISTOR;
  } else if (ty == "C") {
    m.type = XA_CAPACITOR;
  } else {
    return perr(line, "model type " + ty + " is not a device-load model");
  }
  m.level = level;
  for (auto& p : ps)
    if (p.first != "LEVEL") m.params.push_back(p);
  if (n.model_id.count(m.name)) return perr(line, "model " + m.name + " defined twice");
  n.model_id[m.name] = (int)n.models.size();
  n.models.push_back(m);
  return XA_OK;
}

// V source value specification: [DC] v, SIN(...), PULSE(...), EXP(...)
static int parse_vsource(Device& d, const std::vector<std::string>& t, size_t k, int line) {
  static const char* sin_names[] = {"V0", "VA", "FREQ", "TD", "THETA", "PHASE"};
  static const char* exp_names[] = {"V1", "V2", "TD1", "TAU1", "TD2", "TAU2"};
  static const char* pulse_names[] = {"V1", "V2", "TD", "TR", "TF", "PW", "PER"};
  while (k < t.size()) {
    const std::string u = upper(t[k]);
    double v;
    if (u == "DC") {
      if (k + 1 < t.size() && t[k + 1] == "=") ++k;   // DC=v
      if (k + 1 >= t.size() || !parse_number(t[k + 1], &v)) return perr(line, "DC needs a value");
      d.params.emplace_back("DCV0", v);
      k += 2;
    } else if (u == "SIN" || u == "EXP" || u == "PULSE") {
      const char* const* names = u == "SIN" ? sin_names : u == "EXP" ? exp_names : pulse_names;
      const int maxn = u == "PULSE" ? 7 : 6;
      const double code = u == "SIN" ? 0.0 : u == "EXP" ? 1.0 : 2.0;
      ++k;
      if (k >= t.size() || t[k] != "(") return perr(line, u + " needs (arguments)");
      ++k;
      int a = 0;
      while (k < t.size() && t[k] != ")") {
        if (a >= maxn) return perr(line, "too many " + u + " arguments");
        if (!parse_number(t[k], &v)) return perr(line, "bad number " + t[k]);
        d.params.emplace_back(names[a++], v);
        ++k;
      }
      if (k >= t.size()) return perr(line, "unclosed " + u + "(");
      ++k;
      d.params.emplace_back("TRANSIENTSOURCETYPE", code);
    } else if (parse_number(t[k], &v)) {
      d.params.emplace_back("DCV0", v);
      ++k;
    } else {
      return perr(line, "unexpected source token " + t[k]);
    }
  }
  return XA_OK;
}

// A subcircuit instance being expanded: its name prefix ("X1:", "X1:X2:")
// and its ports (definition name -> the node name in the caller).
struct Scope {
  std::string prefix;
  std::map<std::string, std::string> ports;
};

// a node token of a line inside `sc`: ground stays ground, a port is the
// caller's node, any other node is the instance's own ("X1:" + name)
static std::string scoped_node(const Scope* sc, const std::string& tok) {
  const std::string u = upper(tok);
  if (!sc || u == "0" || u == "GND") return u;
  auto it = sc->ports.find(u);
  return it != sc->ports.end() ? it->second : sc->prefix + u;
}

static int parse_element(Netlist& n, const std::vector<std::string>& t, int line,
                         const Scope* sc = nullptr) {
  Device d;
  d.line = line;
  d.name = (sc ? sc->prefix : std::string()) + upper(t[0]);
  d.letter = (char)std::toupper((unsigned char)t[0][0]);
  size_t k = 1;
  double v;
  auto is_model = [&](const std::string& s) { return n.model_names.count(upper(s)) > 0; };
  auto nodes = [&](int cnt) {
    for (int j = 0; j < cnt; ++j) d.nodes.push_back(node_of(n, scoped_node(sc, t[k++])));
  };
  auto value = [&](const std::string& s, const char* name) {
    if (!parse_number(s, &v)) return perr(line, std::string("bad ") + name + " " + s);
    d.params.emplace_back(name, v);
    return XA_OK;
  };
  int rc = XA_OK;
  switch (d.letter) {
    case 'M': {
      if (t.size() < 6) return perr(line, "M needs d g s b model");
      nodes(4);
      std::vector<std::string> p = positional(t, k);
      if (p.size() != 1) return perr(line, "M takes d g s b model and name=value pairs");
      d.model_name = upper(p[0]);
      break;
    }
    case 'D': {
      if (t.size() < 4) return perr(line, "D needs anode cathode model");
      nodes(2);
      std::vector<std::string> p = positional(t, k);
      if (p.empty() || p.size() > 2) return perr(line, "D takes anode cathode model [area]");
      d.model_name = upper(p[0]);
      if (p.size() == 2) rc = value(p[1], "AREA");
      break;
    }
    case 'Q': {
      // c b e [s] model [area]: the token after e is the model when a
      // .MODEL card names it, else the substrate node
      if (t.size() < 5) return perr(line, "Q needs c b e [s] model");
      nodes(3);
      std::vector<std::string> p = positional(t, k);
      size_t j = 0;
      if (p.empty()) return perr(line, "Q needs a model");
      if (!is_model(p[0])) {
        if (p.size() < 2) return perr(line, "Q: no .MODEL card names " + p[0]);
        d.nodes.push_back(node_of(n, scoped_node(sc, p[0])));
        j = 1;
      }
      d.model_name = upper(p[j++]);
      if (j < p.size()) rc = value(p[j++], "AREA");
      if (rc == XA_OK && j < p.size()) return perr(line, "Q takes c b e [s] model [area]");
      break;
    }
    case 'R': case 'C': {
      // n1 n2 [model] [value]
      const char* vname = d.letter == 'R' ? "R" : "C";
      if (t.size() < 4) return perr(line, std::string(1, d.letter) + " needs two nodes and a value");
      nodes(2);
      std::vector<std::string> p = positional(t, k);
      size_t j = 0;
      if (j < p.size() && is_model(p[j])) d.model_name = upper(p[j++]);
      if (j < p.size()) rc = value(p[j++], vname);
      if (rc == XA_OK && j < p.size()) return perr(line, std::string(1, d.letter) + " takes n1 n2 [model] [value]");
      break;
    }
    case 'V': {
      if (t.size() < 3) return perr(line, "V needs two nodes");
      nodes(2);
      rc = parse_vsource(d, t, k, line);
      if (rc != XA_OK) return rc;
      n.devices.push_back(d);
      return XA_OK;
    }
    default:
      return perr(line, "element " + d.name + " is not a device-load element (M D Q R C V, X)");
  }
  if (rc == XA_OK) rc = read_pairs(t, k, line, d.params);
  if (rc != XA_OK) return rc;
  n.devices.push_back(d);
  return XA_OK;
}

static int parse_control(Netlist& n, const std::vector<std::string>& t, int line, bool* end) {
  const std::string c = upper(t[0]);
  double v;
  if (c == ".MODEL") return parse_model(n, t, line);
  if (c == ".END") { *end = true; return XA_OK; }
  if (c == ".TRAN") {
    std::vector<double> a;
    for (size_t k = 1; k < t.size(); ++k) {
      if (upper(t[k]) == "UIC") continue;
      if (!parse_number(t[k], &v)) return perr(line, "bad .TRAN value " + t[k]);
      a.push_back(v);
    }
    if (a.size() < 2 || a.size() > 4) return perr(line, ".TRAN needs tstep tstop [tstart [tmax]]");
    n.tran = true;
    n.tstep = a[0];
    n.tstop = a[1];
    n.tstart = a.size() > 2 ? a[2] : 0.0;
    n.tmax = a.size() > 3 ? a[3] : 0.0;
    return XA_OK;
  }
  if (c == ".IC") {
    // V(node)=value or node=value, repeated
    size_t k = 1;
    while (k < t.size()) {
      std::string node, val;
      if (k + 5 < t.size() && upper(t[k]) == "V" && t[k + 1] == "(" && t[k + 3] == ")" &&
          t[k + 4] == "=") {
        node = t[k + 2];
        val = t[k + 5];
        k += 6;
      } else if (k + 2 < t.size() && t[k + 1] == "=") {
        node = t[k];
        val = t[k + 2];
        k += 3;
      } else {
        return perr(line, ".IC expects V(node)=value");
      }
      if (!parse_number(val, &v)) return perr(line, "bad .IC value " + val);
      const int id = node_of(n, node);
      if (id < 0) return perr(line, ".IC on ground");
      n.ics.emplace_back(id, v);
    }
    return XA_OK;
  }
  if (c == ".STEP") {
    // .STEP [LIN] target:param start stop step  |  .STEP target:param LIST v ...
    size_t k = 1;
    if (k < t.size() && upper(t[k]) == "LIN") ++k;
    if (k >= t.size()) return perr(line, ".STEP needs a parameter");
    const std::string name = upper(t[k++]);
    const size_t colon = name.find(':');
    if (colon == std::string::npos)
      return perr(line, ".STEP takes MODEL:PARAM or DEVICE:PARAM (global parameters are not read)");
    Step s;
    s.target = name.substr(0, colon);
    s.param = name.substr(colon + 1);
    if (k < t.size() && upper(t[k]) == "LIST") {
      for (++k; k < t.size(); ++k) {
        if (!parse_number(t[k], &v)) return perr(line, "bad .STEP value " + t[k]);
        s.values.push_back(v);
      }
    } else {
      double a[3];
      if (k + 3 != t.size()) return perr(line, ".STEP needs start stop step");
      for (int j = 0; j < 3; ++j)
        if (!parse_number(t[k + j], &a[j])) return perr(line, "bad .STEP value " + t[k + j]);
      if (a[2] == 0.0 || (a[1] - a[0]) / a[2] < 0) return perr(line, ".STEP step has the wrong sign");
      // the reference's count of linear steps: floor((stop-start)/step + 1 + tiny)
      const int m = (int)std::floor((a[1] - a[0]) / a[2] + 1.0 + 1e-12);
      for (int j = 0; j < m; ++j) s.values.push_back(a[0] + j * a[2]);
    }
    if (s.values.empty()) return perr(line, ".STEP without values");
    if (!n.steps.empty() && n.steps[0].values.size() != s.values.size())
      return perr(line, "every .STEP must have the same number of values (they step together)");
    n.steps.push_back(s);
    return XA_OK;
  }
  if (c == ".OPTIONS" || c == ".OPTION" || c == ".PRINT" || c == ".MEASURE" || c == ".MEAS")
    return XA_OK;   // output and solver controls: not the device loads' business
  return perr(line, "control line " + c + " is not read");
}

// models resolved, parts formed
static int resolve(Netlist& n) {
  // a .STEP target exists; a stepped device parameter joins the device's
  // parameters (its value outside the sweep is the first step's)
  for (auto& s : n.steps) {
    bool found = n.model_id.count(s.target) > 0;
    for (auto& d : n.devices) {
      if (d.name != s.target) continue;
      found = true;
      bool has = false;
      for (auto& p : d.params) has |= p.first == s.param;
      if (!has) d.params.emplace_back(s.param, s.values[0]);
    }
    if (!found) return fail(XA_ERR_ARG, "netlist: .STEP target " + s.target + " not found");
  }
  std::map<std::pair<int, std::string>, int> part_of;   // (type, sorted param names)
  for (size_t i = 0; i < n.devices.size(); ++i) {
    Device& d = n.devices[i];
    int type = 0;
    if (!d.model_name.empty()) {
      auto it = n.model_id.find(d.model_name);
      if (it == n.model_id.end()) return perr(d.line, "model " + d.model_name + " not defined");
      d.model = it->second;
      type = n.models[d.model].type;
      const char want = type == XA_BSIM4 || type == XA_MOSFET1 ? 'M' : type == XA_DIODE ? 'D'
                        : type == XA_BJT ? 'Q' : type == XA_RESISTOR ? 'R' : 'C';
      if (want != d.letter) return perr(d.line, "model " + d.model_name + " does not fit " + d.name);
    } else {
      if (d.letter == 'V') type = XA_VSOURCE;
      else if (d.letter == 'R') type = XA_RESISTOR;
      else if (d.letter == 'C') type = XA_CAPACITOR;
      else return perr(d.line, d.name + " needs a model");
    }
    if (type == XA_BJT && d.nodes.size() == 3) d.nodes.push_back(-1);   // substrate at ground
    std::vector<std::string> names;
    for (auto& p : d.params) names.push_back(p.first);
    std::sort(names.begin(), names.end());
    names.erase(std::unique(names.begin(), names.end()), names.end());
    std::string key;
    for (auto& s : names) key += s + ",";
    auto pk = std::make_pair(type, key);
    auto it = part_of.find(pk);
    int part;
    if (it == part_of.end()) {
      part = (int)n.parts.size();
      part_of[pk] = part;
      Netlist::Part P;
      P.type = type;
      P.inames = names;
      n.parts.push_back(P);
    } else {
      part = it->second;
    }
    Netlist::Part& P = n.parts[part];
    if (d.model >= 0 && std::find(P.models.begin(), P.models.end(), d.model) == P.models.end())
      P.models.push_back(d.model);
    P.devices.push_back((int)i);
  }
  // a part's models in .MODEL order
  for (auto& P : n.parts) std::sort(P.models.begin(), P.models.end());
  return XA_OK;
}

struct Subckt {
  std::vector<std::string> ports;
  std::vector<std::pair<std::string, int>> body;   // logical lines, line numbers
};

// X<name> n1 .. nk <subckt>: the subcircuit's lines with its nodes renamed
// and its elements named "<X name>:<element>" (flattened, as the reference's
// parser expands subcircuits before the devices are made); nested instances
// recurse.
static int expand(Netlist& n, const std::map<std::string, Subckt>& defs,
                  const std::vector<std::string>& t, int line, const Scope* sc, int depth) {
  if (depth > 32) return perr(line, "subcircuits nested deeper than 32 (recursive definition?)");
  std::vector<std::string> p;
  for (size_t k = 1; k < t.size() && t[k] != "=" && upper(t[k]) != "PARAMS:"; ++k) p.push_back(t[k]);
  if (!p.empty() && t.size() > p.size() + 1 && t[p.size() + 1] == "=") p.pop_back();
  if (p.empty()) return perr(line, upper(t[0]) + " names no subcircuit");
  const std::string sub = upper(p.back());
  auto it = defs.find(sub);
  if (it == defs.end()) return perr(line, "subcircuit " + sub + " not defined");
  const Subckt& S = it->second;
  if (p.size() - 1 != S.ports.size())
    return perr(line, upper(t[0]) + " connects " + std::to_string(p.size() - 1) + " nodes, " +
                          sub + " has " + std::to_string(S.ports.size()) + " ports");
  if (p.size() + 1 < t.size()) return perr(line, "subcircuit parameters are not read");
  Scope inner;
  inner.prefix = (sc ? sc->prefix : std::string()) + upper(t[0]) + ":";
  for (size_t j = 0; j < S.ports.size(); ++j) inner.ports[S.ports[j]] = scoped_node(sc, p[j]);
  for (const auto& l : S.body) {
    const std::vector<std::string> b = tokenize(l.first);
    if (b.empty()) continue;
    int rc;
    if (b[0][0] == '.') rc = perr(l.second, "control line " + upper(b[0]) + " inside a subcircuit");
    else if (std::toupper((unsigned char)b[0][0]) == 'X') rc = expand(n, defs, b, l.second, &inner, depth + 1);
    else rc = parse_element(n, b, l.second, &inner);
    if (rc != XA_OK) return rc;
  }
  return XA_OK;
}

static int parse(const char* text, Netlist& n) {
  std::istringstream in(text ? text : "");
  std::string raw;
  std::vector<std::pair<std::string, int>> lines;   // logical lines and their first line number
  int lineno = 0;
  bool first = true;
  while (std::getline(in, raw)) {
    ++lineno;
    if (first) { n.title = raw; first = false; continue; }
    const size_t sc = raw.find(';');
    if (sc != std::string::npos) raw = raw.substr(0, sc);
    size_t b = raw.find_first_not_of(" \t\r");
    if (b == std::string::npos) continue;
    if (raw[b] == '*') continue;
    if (raw[b] == '+') {
      if (lines.empty()) return perr(lineno, "continuation with no line to continue");
      lines.back().first += " " + raw.substr(b + 1);
      continue;
    }
    lines.emplace_back(raw.substr(b), lineno);
  }
  for (auto& l : lines) {
    const std::vector<std::string> t = tokenize(l.first);
    if (t.size() >= 2 && upper(t[0]) == ".MODEL") n.model_names.insert(upper(t[1]));
  }
  // .SUBCKT name ports ... .ENDS blocks out of the main line list (a .MODEL
  // inside one is read as a global card)
  std::map<std::string, Subckt> defs;
  std::vector<std::pair<std::string, int>> main_lines;
  Subckt* open = nullptr;
  for (auto& l : lines) {
    const std::vector<std::string> t = tokenize(l.first);
    const std::string c = t.empty() ? "" : upper(t[0]);
    if (c == ".SUBCKT") {
      if (open) return perr(l.second, "nested .SUBCKT definitions are not read");
      if (t.size() < 2) return perr(l.second, ".SUBCKT needs a name");
      const std::string name = upper(t[1]);
      if (defs.count(name)) return perr(l.second, "subcircuit " + name + " defined twice");
      open = &defs[name];
      for (size_t k = 2; k < t.size(); ++k) {
        if (t[k] == "=" || upper(t[k]) == "PARAMS:") return perr(l.second, "subcircuit parameters are not read");
        open->ports.push_back(upper(t[k]));
      }
    } else if (c == ".ENDS") {
      if (!open) return perr(l.second, ".ENDS without .SUBCKT");
      open = nullptr;
    } else if (open && c != ".MODEL") {
      open->body.push_back(l);
    } else {
      main_lines.push_back(l);
    }
  }
  if (open) return perr(lineno, ".SUBCKT without .ENDS");
  bool end = false;
  for (auto& l : main_lines) {
    if (end) break;
    const std::vector<std::string> t = tokenize(l.first);
    if (t.empty()) continue;
    int rc;
    if (t[0][0] == '.') rc = parse_control(n, t, l.second, &end);
    else if (std::toupper((unsigned char)t[0][0]) == 'X') rc = expand(n, defs, t, l.second, nullptr, 0);
    else rc = parse_element(n, t, l.second);
    if (rc != XA_OK) return rc;
  }
  return resolve(n);
}

// value of `param` of model `m` or device `d` at step s (the stepped value if
// a .STEP names it)
static void apply_step(const Netlist& n, int s, const std::string& target, bool is_model,
                       std::vector<std::pair<std::string, double>>& ps) {
  if (s < 0) return;
  for (const Step& st : n.steps) {
    if (st.target != target) continue;
    bool set = false;
    for (auto& p : ps)
      if (p.first == st.param) { p.second = st.values[s]; set = true; }
    if (!set) ps.emplace_back(st.param, st.values[s]);
  }
  (void)is_model;
}

}  // namespace nl

// ---- circuit -------------------------------------------------------------
struct Circuit {
  Ctx* ctx = nullptr;
  std::vector<xa_group> groups;
  std::vector<int> types;
  int n_copies = 1;
  int n_unknowns = 0, n_sta = 0, n_sto = 0;
  std::vector<int32_t> row_ptr, col_idx;
  std::vector<std::vector<int32_t>> node_lids;   // per group, instance-major
  ~Circuit() {
    for (xa_group g : groups) xa_group_destroy(g);
  }
};

}  // namespace xa

using namespace xa;

struct xa_netlist_s { nl::Netlist n; };
struct xa_circuit_s { Circuit c; };

extern "C" {

int xa_netlist_parse(const char* text, xa_netlist* out) {
  if (!out) return fail(XA_ERR_ARG, "xa_netlist_parse: null out");
  std::unique_ptr<xa_netlist_s> h(new xa_netlist_s());
  const int rc = nl::parse(text, h->n);
  if (rc != XA_OK) return rc;
  *out = h.release();
  return XA_OK;
}

int xa_netlist_destroy(xa_netlist h) {
  delete h;
  return XA_OK;
}

int xa_netlist_sizes(xa_netlist h, int* n_nodes, int* n_parts, int* n_steps, int* n_ic) {
  if (!h) return fail(XA_ERR_ARG, "null netlist");
  if (n_nodes) *n_nodes = (int)h->n.node_names.size();
  if (n_parts) *n_parts = (int)h->n.parts.size();
  if (n_steps) *n_steps = h->n.steps.empty() ? 0 : (int)h->n.steps[0].values.size();
  if (n_ic) *n_ic = (int)h->n.ics.size();
  return XA_OK;
}

int xa_netlist_node_name(xa_netlist h, int node, char* buf, int len) {
  if (!h || node < 0 || node >= (int)h->n.node_names.size() || !buf || len <= 0)
    return fail(XA_ERR_ARG, "xa_netlist_node_name: bad node or buffer");
  std::snprintf(buf, (size_t)len, "%s", h->n.node_names[node].c_str());
  return XA_OK;
}

int xa_netlist_tran(xa_netlist h, double* out4) {
  if (!h || !out4) return fail(XA_ERR_ARG, "xa_netlist_tran: null argument");
  if (!h->n.tran) return fail(XA_ERR_STATE, "netlist has no .TRAN");
  out4[0] = h->n.tstep;
  out4[1] = h->n.tstop;
  out4[2] = h->n.tstart;
  out4[3] = h->n.tmax;
  return XA_OK;
}

int xa_netlist_ic(xa_netlist h, int32_t* nodes, double* values) {
  if (!h) return fail(XA_ERR_ARG, "null netlist");
  for (size_t k = 0; k < h->n.ics.size(); ++k) {
    if (nodes) nodes[k] = h->n.ics[k].first;
    if (values) values[k] = h->n.ics[k].second;
  }
  return XA_OK;
}

int xa_netlist_part(xa_netlist h, int part, int* model_type, int* n_models, int* n_inst,
                    int* n_ext, int* n_iparams) {
  if (!h || part < 0 || part >= (int)h->n.parts.size()) return fail(XA_ERR_ARG, "bad part");
  const nl::Netlist::Part& P = h->n.parts[part];
  if (model_type) *model_type = P.type;
  if (n_models) *n_models = (int)std::max<size_t>(P.models.size(), 1);
  if (n_inst) *n_inst = (int)P.devices.size();
  if (n_ext) {
    int e = 0;
    for (int d : P.devices) e += (int)h->n.devices[d].nodes.size();
    *n_ext = e;
  }
  if (n_iparams) *n_iparams = (int)P.inames.size();
  return XA_OK;
}

// model k of a part at step `step` (-1: the netlist's values): level and its
// name=value pairs; with names == NULL, *n_params is the count.  Parts of
// model-less elements (R, C, V with no model) have one empty model.
int xa_netlist_part_model(xa_netlist h, int part, int k, int step, int* level, int* n_params,
                          const char** names, double* values) {
  if (!h || part < 0 || part >= (int)h->n.parts.size()) return fail(XA_ERR_ARG, "bad part");
  nl::Netlist::Part& P = h->n.parts[part];
  std::vector<std::pair<std::string, double>> ps;
  int lev = 1;
  if (!P.models.empty()) {
    if (k < 0 || k >= (int)P.models.size()) return fail(XA_ERR_ARG, "bad model index");
    const nl::Model& m = h->n.models[P.models[k]];
    ps = m.params;
    lev = m.level;
    nl::apply_step(h->n, step, m.name, true, ps);
  } else if (k != 0) {
    return fail(XA_ERR_ARG, "bad model index");
  }
  if (level) *level = lev;
  if (n_params) *n_params = (int)ps.size();
  if (names) {
    P.mnames_flat.clear();
    for (auto& p : ps) P.mnames_flat.push_back(p.first);
    for (size_t j = 0; j < ps.size(); ++j) {
      names[j] = P.mnames_flat[j].c_str();
      if (values) values[j] = ps[j].second;
    }
  }
  return XA_OK;
}

// instances of a part at step `step`: model index within the part, the
// external node ids (instance-major), and the instance parameters as
// columns values[p * n + i] in the part's parameter order (names[p]).
int xa_netlist_part_instances(xa_netlist h, int part, int step, int32_t* model_index,
                              int32_t* ext_nodes, const char** names, double* values) {
  if (!h || part < 0 || part >= (int)h->n.parts.size()) return fail(XA_ERR_ARG, "bad part");
  const nl::Netlist& N = h->n;
  const nl::Netlist::Part& P = N.parts[part];
  const int n = (int)P.devices.size();
  int e = 0;
  for (int i = 0; i < n; ++i) {
    const nl::Device& d = N.devices[P.devices[i]];
    if (model_index) {
      int mi = 0;
      if (d.model >= 0)
        mi = (int)(std::find(P.models.begin(), P.models.end(), d.model) - P.models.begin());
      model_index[i] = mi;
    }
    for (int x : d.nodes) {
      if (ext_nodes) ext_nodes[e] = x;
      ++e;
    }
    std::vector<std::pair<std::string, double>> ps = d.params;
    nl::apply_step(N, step, d.name, false, ps);
    for (size_t p = 0; p < P.inames.size(); ++p) {
      double v = NAN;
      for (auto& q : ps)
        if (q.first == P.inames[p]) v = q.second;
      if (values) values[p * (size_t)n + i] = v;
    }
  }
  if (names)
    for (size_t p = 0; p < P.inames.size(); ++p) names[p] = P.inames[p].c_str();
  return XA_OK;
}

int xa_circuit_build(xa_netlist h, xa_ctx ctx, const xa_device_options* opt, int step,
                     xa_circuit* out) {
  if (!h || !ctx || !out) return fail(XA_ERR_ARG, "xa_circuit_build: null argument");
  const nl::Netlist& N = h->n;
  const int n_steps = N.steps.empty() ? 0 : (int)N.steps[0].values.size();
  // step >= 0: that step's circuit; -1: the netlist's values; -2: every
  // step's copy in one build (batched)
  const bool ok = step == -1 || (step == -2 && n_steps > 0) || (step >= 0 && step < n_steps);
  if (!ok)
    return fail(XA_ERR_ARG, "xa_circuit_build: step " + std::to_string(step) + " out of range (" +
                                std::to_string(n_steps) + " .STEP points)");
  std::unique_ptr<xa_circuit_s> hc(new xa_circuit_s());
  Circuit& C = hc->c;
  C.ctx = reinterpret_cast<Ctx*>(ctx);
  const int copies = step == -2 ? n_steps : 1;
  C.n_copies = copies;
  const int n_nodes = (int)N.node_names.size();
  struct Built {
    std::vector<int> n_ext, n_int, n_sta, n_sto;
  };
  std::vector<Built> built(N.parts.size());
  int rc;
  for (size_t p = 0; p < N.parts.size(); ++p) {
    const nl::Netlist::Part& P = N.parts[p];
    xa_group g = nullptr;
    if ((rc = xa_group_create(ctx, P.type, opt, &g)) != XA_OK) return rc;
    C.groups.push_back(g);
    C.types.push_back(P.type);
    const int nm = (int)std::max<size_t>(P.models.size(), 1);
    const int ni = (int)P.devices.size();
    // a part whose models no .STEP touches shares one set of cards among
    // the copies (fewer cards to stage; the same values)
    bool stepped = false;
    for (const nl::Step& st : N.steps)
      for (int m : P.models) stepped |= N.models[m].name == st.target;
    const int model_copies = stepped ? copies : 1;
    for (int c = 0; c < model_copies; ++c) {
      const int s = step == -2 ? c : step;
      for (int k = 0; k < nm; ++k) {
        int level, np;
        if ((rc = xa_netlist_part_model(h, (int)p, k, s, &level, &np, nullptr, nullptr)) != XA_OK) return rc;
        std::vector<const char*> names(np);
        std::vector<double> vals(np);
        if ((rc = xa_netlist_part_model(h, (int)p, k, s, &level, &np, names.data(), vals.data())) != XA_OK)
          return rc;
        std::vector<std::string> keep(names.begin(), names.end());
        std::vector<const char*> cn;
        for (auto& x : keep) cn.push_back(x.c_str());
        int idx;
        if ((rc = xa_group_add_model(g, level, np, cn.data(), vals.data(), &idx)) != XA_OK) return rc;
      }
    }
    // instances of every copy in one batch: copy c's instances use copy c's models
    const int np = (int)P.inames.size();
    std::vector<int32_t> mi_all;
    std::vector<std::vector<double>> cols(np);
    for (int c = 0; c < copies; ++c) {
      const int s = step == -2 ? c : step;
      std::vector<int32_t> mi(ni);
      std::vector<double> vals((size_t)np * ni);
      if ((rc = xa_netlist_part_instances(h, (int)p, s, mi.data(), nullptr, nullptr, vals.data())) != XA_OK)
        return rc;
      for (int i = 0; i < ni; ++i) mi_all.push_back(mi[i] + (stepped ? c * nm : 0));
      for (int q = 0; q < np; ++q) cols[q].insert(cols[q].end(), vals.begin() + (size_t)q * ni,
                                                   vals.begin() + (size_t)(q + 1) * ni);
    }
    std::vector<const char*> cn;
    for (auto& x : P.inames) cn.push_back(x.c_str());
    std::vector<const double*> cp;
    for (auto& col : cols) cp.push_back(col.data());
    if ((rc = xa_group_add_instances(g, ni * copies, mi_all.data(), np, cn.data(), cp.data())) != XA_OK)
      return rc;
    if ((rc = xa_group_setup(g)) != XA_OK) return rc;
    Built& B = built[p];
    const int nt = ni * copies;
    B.n_ext.resize(nt); B.n_int.resize(nt); B.n_sta.resize(nt); B.n_sto.resize(nt);
    if ((rc = xa_group_counts(g, B.n_ext.data(), B.n_int.data(), B.n_sta.data(), B.n_sto.data())) != XA_OK)
      return rc;
  }
  // LIDs: per copy, the netlist nodes, then every part's internal nodes in
  // part and instance order; states and stores likewise
  int per_copy_int = 0;
  for (size_t p = 0; p < N.parts.size(); ++p)
    for (size_t i = 0; i < N.parts[p].devices.size(); ++i) per_copy_int += built[p].n_int[i];
  const int per_copy = n_nodes + per_copy_int;
  C.n_unknowns = per_copy * copies;
  std::vector<std::vector<int32_t>> node_lids(N.parts.size()), sta(N.parts.size()), sto(N.parts.size());
  int sta0 = 0, sto0 = 0;
  for (int c = 0; c < copies; ++c) {
    int nxt = c * per_copy + n_nodes;
    for (size_t p = 0; p < N.parts.size(); ++p) {
      const nl::Netlist::Part& P = N.parts[p];
      const int ni = (int)P.devices.size();
      for (int i = 0; i < ni; ++i) {
        const int gi = c * ni + i;
        const nl::Device& d = N.devices[P.devices[i]];
        if ((int)d.nodes.size() != built[p].n_ext[gi])
          return fail(XA_ERR_ARG, "netlist: " + d.name + " has " + std::to_string(d.nodes.size()) +
                                      " nodes, the model takes " + std::to_string(built[p].n_ext[gi]));
        for (int x : d.nodes) node_lids[p].push_back(x < 0 ? -1 : x + c * per_copy);
        for (int k = 0; k < built[p].n_int[gi]; ++k) node_lids[p].push_back(nxt++);
      }
    }
  }
  // states and stores: copy-major like the nodes, then part and instance
  // order (copy c's instances are c*ni .. c*ni+ni-1 of each group, so every
  // group's list stays in its instance order)
  for (int c = 0; c < copies; ++c)
    for (size_t p = 0; p < N.parts.size(); ++p) {
      const int ni = (int)N.parts[p].devices.size();
      for (int gi = c * ni; gi < (c + 1) * ni; ++gi) {
        for (int k = 0; k < built[p].n_sta[gi]; ++k) sta[p].push_back(sta0++);
        for (int k = 0; k < built[p].n_sto[gi]; ++k) sto[p].push_back(sto0++);
      }
    }
  C.n_sta = sta0;
  C.n_sto = sto0;
  std::vector<int64_t> keys;
  for (size_t p = 0; p < N.parts.size(); ++p) {
    xa_group g = C.groups[p];
    if ((rc = xa_group_register_lids(g, node_lids[p].data(), sta[p].empty() ? nullptr : sta[p].data(),
                                     sto[p].empty() ? nullptr : sto[p].data())) != XA_OK)
      return rc;
    int64_t cnt = 0;
    if ((rc = xa_group_jac_pairs(g, &cnt, nullptr, nullptr)) != XA_OK) return rc;
    std::vector<int32_t> r(cnt), q(cnt);
    if ((rc = xa_group_jac_pairs(g, &cnt, r.data(), q.data())) != XA_OK) return rc;
    for (int64_t j = 0; j < cnt; ++j) keys.push_back((int64_t)r[j] * C.n_unknowns + q[j]);
  }
  C.node_lids = node_lids;
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  C.row_ptr.assign((size_t)C.n_unknowns + 1, 0);
  C.col_idx.resize(keys.size());
  for (size_t j = 0; j < keys.size(); ++j) {
    C.row_ptr[(size_t)(keys[j] / C.n_unknowns) + 1]++;
    C.col_idx[j] = (int32_t)(keys[j] % C.n_unknowns);
  }
  for (int r = 0; r < C.n_unknowns; ++r) C.row_ptr[r + 1] += C.row_ptr[r];
  for (xa_group g : C.groups) {
    if ((rc = xa_group_register_csr(g, C.n_unknowns, C.row_ptr.data(), C.col_idx.data())) != XA_OK) return rc;
    if ((rc = xa_group_finalize(g)) != XA_OK) return rc;
  }
  *out = hc.release();
  return XA_OK;
}

int xa_circuit_destroy(xa_circuit h) {
  delete h;
  return XA_OK;
}

int xa_circuit_sizes(xa_circuit h, int* n_unknowns, int* n_sta, int* n_sto, int64_t* nnz,
                     int* n_groups, int* n_copies) {
  if (!h) return fail(XA_ERR_ARG, "null circuit");
  const Circuit& C = h->c;
  if (n_unknowns) *n_unknowns = C.n_unknowns;
  if (n_sta) *n_sta = C.n_sta;
  if (n_sto) *n_sto = C.n_sto;
  if (nnz) *nnz = (int64_t)C.col_idx.size();
  if (n_groups) *n_groups = (int)C.groups.size();
  if (n_copies) *n_copies = C.n_copies;
  return XA_OK;
}

int xa_circuit_csr(xa_circuit h, int32_t* row_ptr, int32_t* col_idx) {
  if (!h) return fail(XA_ERR_ARG, "null circuit");
  const Circuit& C = h->c;
  if (row_ptr) std::copy(C.row_ptr.begin(), C.row_ptr.end(), row_ptr);
  if (col_idx) std::copy(C.col_idx.begin(), C.col_idx.end(), col_idx);
  return XA_OK;
}

int xa_circuit_group(xa_circuit h, int k, xa_group* g, int* model_type) {
  if (!h || k < 0 || k >= (int)h->c.groups.size()) return fail(XA_ERR_ARG, "bad group index");
  if (g) *g = h->c.groups[k];
  if (model_type) *model_type = h->c.types[k];
  return XA_OK;
}

// node LIDs of group k (instance-major: each instance's external then
// internal nodes; ground -1); with node_lids NULL, *count is the length
int xa_circuit_group_lids(xa_circuit h, int k, int64_t* count, int32_t* node_lids) {
  if (!h || k < 0 || k >= (int)h->c.groups.size() || !count)
    return fail(XA_ERR_ARG, "bad group index or null count");
  const std::vector<int32_t>& v = h->c.node_lids[k];
  *count = (int64_t)v.size();
  if (node_lids) std::copy(v.begin(), v.end(), node_lids);
  return XA_OK;
}

// DeviceMgr's load sequence over the circuit's device types
// (N_DEV_DeviceMgr.C:2053-2502): every group's four calls, groups in
// creation order; each group adds onto the same vectors.
int xa_circuit_load(xa_circuit h, const xa_solver_state* ss, const xa_vectors* v) {
  if (!h) return fail(XA_ERR_ARG, "null circuit");
  int rc;
  for (xa_group g : h->c.groups)
    if ((rc = xa_load_all(g, ss, v)) != XA_OK) return rc;
  return XA_OK;
}

}  // extern "C"
