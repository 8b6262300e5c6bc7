// xnb_operators.cpp -- host side mirroring exaNBody's operator interface for the short-range
// path: operators registered by their YAML names, slots resolved by name from the deck and
// from a global slot store, batches with condition / loop / rebind, and a YAML-subset reader
// for the deck files.  Every particle operation goes through the C ABI (include/xnb_amd.h).
//
// Operator names and slot names follow the reference (data/config/numerical-scheme.msp,
// update-particles.msp, compute-loop.msp; operator sources cited per operator below).
#include <algorithm>
#include <array>
#include <cctype>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <random>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/xnb_amd.h"

namespace xnbdeck {

// ------------------------------------------------------------------ YAML subset
struct Node {
  enum Kind { Null, Scalar, Map, Seq } kind = Null;
  std::string s;
  std::vector<std::pair<std::string, Node>> m;
  std::vector<Node> q;

  const Node* get(const std::string& k) const {
    for (auto& kv : m)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
  Node* get(const std::string& k) {
    for (auto& kv : m)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
  void set(const std::string& k, Node v) {
    if (Node* p = get(k)) *p = std::move(v);
    else m.emplace_back(k, std::move(v));
  }
  static Node scalar(const std::string& v) {
    Node n;
    n.kind = Scalar;
    n.s = v;
    return n;
  }
};

struct DeckError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

static std::string trim(const std::string& x) {
  size_t a = 0, b = x.size();
  while (a < b && std::isspace((unsigned char)x[a])) a++;
  while (b > a && std::isspace((unsigned char)x[b - 1])) b--;
  return x.substr(a, b - a);
}

static std::string strip_comment(const std::string& line) {
  bool sq = false, dq = false;
  for (size_t i = 0; i < line.size(); i++) {
    const char c = line[i];
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (c == '#' && !sq && !dq && (i == 0 || std::isspace((unsigned char)line[i - 1]))) return line.substr(0, i);
  }
  return line;
}

// flow-style value: {a: 1, b: [x, y]}, [1, 2], scalar
struct Flow {
  const std::string& t;
  size_t i = 0;
  explicit Flow(const std::string& s) : t(s) {}
  void ws() {
    while (i < t.size() && std::isspace((unsigned char)t[i])) i++;
  }
  Node value() {
    ws();
    if (i < t.size() && t[i] == '{') return map();
    if (i < t.size() && t[i] == '[') return seq();
    return Node::scalar(word());
  }
  std::string word() {
    ws();
    if (i < t.size() && (t[i] == '"' || t[i] == '\'')) {
      const char qc = t[i++];
      std::string w;
      while (i < t.size() && t[i] != qc) w += t[i++];
      i++;
      return w;
    }
    std::string w;
    int depth = 0;
    while (i < t.size()) {
      const char c = t[i];
      if (depth == 0 && (c == ',' || c == '}' || c == ']')) break;
      if (c == '[' || c == '{') depth++;
      if (c == ']' || c == '}') depth--;
      w += c;
      i++;
    }
    return trim(w);
  }
  std::string key() {
    ws();
    std::string k;
    while (i < t.size() && t[i] != ':') k += t[i++];
    if (i >= t.size()) throw DeckError("flow map: missing ':' in " + t);
    i++;
    return trim(k);
  }
  Node map() {
    Node n;
    n.kind = Node::Map;
    i++;
    ws();
    if (i < t.size() && t[i] == '}') {
      i++;
      return n;
    }
    while (i < t.size()) {
      std::string k = key();
      n.m.emplace_back(k, value());
      ws();
      if (i < t.size() && t[i] == ',') {
        i++;
        continue;
      }
      if (i < t.size() && t[i] == '}') {
        i++;
        break;
      }
      throw DeckError("flow map: expected ',' or '}' in " + t);
    }
    return n;
  }
  Node seq() {
    Node n;
    n.kind = Node::Seq;
    i++;
    ws();
    if (i < t.size() && t[i] == ']') {
      i++;
      return n;
    }
    while (i < t.size()) {
      n.q.push_back(value());
      ws();
      if (i < t.size() && t[i] == ',') {
        i++;
        continue;
      }
      if (i < t.size() && t[i] == ']') {
        i++;
        break;
      }
      throw DeckError("flow seq: expected ',' or ']' in " + t);
    }
    return n;
  }
};

static Node parse_inline(const std::string& v) {
  Flow f(v);
  return f.value();
}

struct Line {
  int indent;
  std::string text;
};

// block-structure parser over pre-split lines
struct Block {
  std::vector<Line> L;
  size_t p = 0;

  Node parse(int indent) {
    if (p >= L.size()) return Node();
    if (L[p].text.rfind("- ", 0) == 0 || L[p].text == "-") return seq(L[p].indent);
    return map(L[p].indent);
  }

  Node value_after(const std::string& rest, int parent_indent) {
    if (!rest.empty()) return parse_inline(rest);
    if (p < L.size() && L[p].indent > parent_indent) return parse(L[p].indent);
    if (p < L.size() && L[p].indent == parent_indent && L[p].text.rfind("- ", 0) == 0) return seq(parent_indent);
    return Node();
  }

  Node map(int indent) {
    Node n;
    n.kind = Node::Map;
    while (p < L.size() && L[p].indent == indent && L[p].text.rfind("- ", 0) != 0) {
      const std::string& t = L[p].text;
      const size_t c = find_colon(t);
      if (c == std::string::npos) throw DeckError("expected 'key:' at: " + t);
      std::string k = trim(t.substr(0, c));
      std::string rest = trim(t.substr(c + 1));
      p++;
      n.m.emplace_back(k, value_after(rest, indent));
    }
    return n;
  }

  Node seq(int indent) {
    Node n;
    n.kind = Node::Seq;
    while (p < L.size() && L[p].indent == indent && (L[p].text.rfind("- ", 0) == 0 || L[p].text == "-")) {
      std::string rest = L[p].text == "-" ? "" : trim(L[p].text.substr(2));
      const int item_indent = indent + 2;
      p++;
      if (rest.empty()) {
        n.q.push_back(p < L.size() && L[p].indent > indent ? parse(L[p].indent) : Node());
      } else if (rest[0] == '{' || rest[0] == '[' || find_colon(rest) == std::string::npos) {
        n.q.push_back(parse_inline(rest));
      } else {
        // "- key: value" possibly continued by more keys at item_indent
        L.insert(L.begin() + (long)p, Line{item_indent, rest});
        n.q.push_back(map(item_indent));
      }
    }
    return n;
  }

  static size_t find_colon(const std::string& t) {
    int depth = 0;
    bool sq = false, dq = false;
    for (size_t i = 0; i < t.size(); i++) {
      const char c = t[i];
      if (c == '\'' && !dq) sq = !sq;
      if (c == '"' && !sq) dq = !dq;
      if (sq || dq) continue;
      if (c == '{' || c == '[') depth++;
      if (c == '}' || c == ']') depth--;
      if (c == ':' && depth == 0 && (i + 1 == t.size() || std::isspace((unsigned char)t[i + 1]))) return i;
    }
    return std::string::npos;
  }
};

Node parse_yaml(const std::string& text) {
  Block b;
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    std::string s = strip_comment(line);
    if (trim(s).empty()) continue;
    int ind = 0;
    while (ind < (int)s.size() && s[ind] == ' ') ind++;
    b.L.push_back(Line{ind, trim(s)});
  }
  if (b.L.empty()) {
    Node n;
    n.kind = Node::Map;
    return n;
  }
  return b.parse(b.L[0].indent);
}

// ------------------------------------------------------------------ values
// Quantities: a number, optionally followed by a unit expression ("3.48 ang", "2e-3 ps",
// "0.3729 eV", "58.693 Da", "ang/ps^2"), converted to the internal units of the reference's
// physics library: length angstrom, time picosecond, mass Dalton, energy Da ang^2 / ps^2,
// charge e, temperature K.  A bare number is taken as already internal (reduced-unit decks).
// Constants: those of the reference's unit system, fixed by its golden check_values_lj_Ni.dat
// (elementary charge 1.6021892e-19 C, Dalton 1.660539040e-27 kg).
static constexpr double kElemCharge = 1.6021892e-19;
static constexpr double kDaltonKg = 1.660539040e-27;
static constexpr double kEnergyJ = kDaltonKg * 1.0e-20 / 1.0e-24;  // internal energy unit in J

static bool unit_factor(const std::string& u, double& f) {
  static const std::map<std::string, double> table = {
      {"m", 1e10}, {"cm", 1e8}, {"mm", 1e7}, {"um", 1e4}, {"nm", 10.0}, {"ang", 1.0}, {"angstrom", 1.0},
      {"s", 1e12}, {"ms", 1e9}, {"us", 1e6}, {"ns", 1e3}, {"ps", 1.0}, {"fs", 1e-3},
      {"kg", 1.0 / kDaltonKg}, {"g", 1e-3 / kDaltonKg}, {"Da", 1.0}, {"Dalton", 1.0}, {"amu", 1.0},
      {"J", 1.0 / kEnergyJ}, {"eV", kElemCharge / kEnergyJ},
      {"C", 1.0 / kElemCharge}, {"elementary_charge", 1.0},
      {"K", 1.0}, {"rad", 1.0}, {"radian", 1.0}, {"degree", 3.14159265358979323846 / 180.0}};
  auto it = table.find(u);
  if (it == table.end()) return false;
  f = it->second;
  return true;
}

// "a.b/c^2": products ('.', '*') and quotients ('/') of units with integer powers ('^')
static bool unit_expression(const std::string& e, double& f) {
  f = 1.0;
  size_t i = 0;
  bool divide = false;
  while (i < e.size()) {
    size_t j = i;
    while (j < e.size() && e[j] != '.' && e[j] != '*' && e[j] != '/' && e[j] != '^') j++;
    double uf;
    if (!unit_factor(e.substr(i, j - i), uf)) return false;
    int pw = 1;
    if (j < e.size() && e[j] == '^') {
      char* end = nullptr;
      pw = (int)std::strtol(e.c_str() + j + 1, &end, 10);
      j = (size_t)(end - e.c_str());
    }
    const double x = std::pow(uf, pw);
    f = divide ? f / x : f * x;
    if (j < e.size()) divide = e[j] == '/';
    i = j + 1;
  }
  return true;
}

static double as_double(const Node& n, const std::string& what) {
  if (n.kind != Node::Scalar) throw DeckError(what + ": expected a number");
  char* end = nullptr;
  const double v = std::strtod(n.s.c_str(), &end);
  if (end == n.s.c_str()) throw DeckError(what + ": not a number: " + n.s);
  const std::string unit = trim(std::string(end));
  if (unit.empty()) return v;
  double f;
  if (!unit_expression(unit, f)) throw DeckError(what + ": unknown unit '" + unit + "'");
  return v * f;
}
static bool as_bool(const Node& n, const std::string& what) {
  if (n.kind != Node::Scalar) throw DeckError(what + ": expected a boolean");
  if (n.s == "true" || n.s == "True" || n.s == "1") return true;
  if (n.s == "false" || n.s == "False" || n.s == "0") return false;
  throw DeckError(what + ": not a boolean: " + n.s);
}
static std::string fmt(double v) {
  char b[64];
  std::snprintf(b, sizeof b, "%.17g", v);
  return b;
}

// ------------------------------------------------------------------ runner
struct Runner;
using OpFn = std::function<void(Runner&, const Node& params)>;

struct Runner {
  Node deck;                           // merged deck definitions
  std::map<std::string, Node> slots;   // global slot store
  std::vector<std::map<std::string, std::string>> rebinds;
  std::map<std::string, OpFn> ops;
  xnb_ctx* ctx = nullptr;
  std::ostringstream log;
  std::vector<double> in_r, in_v;
  std::vector<uint64_t> in_id;
  std::vector<uint8_t> in_type;
  int depth = 0;

  std::string resolve(const std::string& name) const {
    for (auto it = rebinds.rbegin(); it != rebinds.rend(); ++it) {
      auto f = it->find(name);
      if (f != it->end()) return f->second;
    }
    return name;
  }
  // slot lookup: operator parameters first, then the (rebound) global slot
  const Node* slot(const Node& params, const std::string& name) const {
    if (params.kind == Node::Map)
      if (const Node* v = params.get(name)) return v;
    auto it = slots.find(resolve(name));
    return it == slots.end() ? nullptr : &it->second;
  }
  double num(const Node& params, const std::string& name, double dflt, bool required = false) const {
    const Node* v = slot(params, name);
    if (!v) {
      if (required) throw DeckError("required slot '" + name + "' is not set");
      return dflt;
    }
    return as_double(*v, name);
  }
  bool flag(const Node& params, const std::string& name, bool dflt) const {
    const Node* v = slot(params, name);
    return v ? as_bool(*v, name) : dflt;
  }
  void out(const std::string& name, const Node& v) { slots[resolve(name)] = v; }
  void out_num(const std::string& name, double v) { out(name, Node::scalar(fmt(v))); }
  void out_bool(const std::string& name, bool v) { out(name, Node::scalar(v ? "true" : "false")); }

  void check(int rc, const char* what) {
    if (rc != XNB_OK)
      throw DeckError(std::string(what) + " failed (" + xnb_error_string(rc) + "): " + (ctx ? xnb_last_error(ctx) : xnb_last_error(nullptr)));
  }
  xnb_ctx* need_ctx(const char* what) {
    if (!ctx) throw DeckError(std::string(what) + ": no domain/grid (run 'domain' first)");
    return ctx;
  }

  bool eval_condition(const Node& c) {
    if (c.kind != Node::Scalar) throw DeckError("condition must be a slot name");
    std::string s = trim(c.s);
    bool neg = false;
    if (s.rfind("not ", 0) == 0) {
      neg = true;
      s = trim(s.substr(4));
    }
    if (s == "true") return !neg;
    if (s == "false") return neg;
    auto it = slots.find(resolve(s));
    const bool v = it != slots.end() && as_bool(it->second, s);
    return neg ? !v : v;
  }

  // run a named entry: deck definition, or a registered operator with the deck's parameters
  void run_name(const std::string& name, const Node* inline_params) {
    if (++depth > 64) throw DeckError("operator nesting too deep at '" + name + "'");
    const Node* def = deck.get(name);
    auto op = ops.find(name);
    if (def && def->kind == Node::Scalar && def->s != name) {
      if (def->s == "nop") {
        depth--;
        return;
      }
      run_name(def->s, inline_params);
    } else if (def && (def->kind == Node::Seq || (def->kind == Node::Map && def->get("body")))) {
      run_block(*def);
    } else if (op != ops.end()) {
      Node params;
      params.kind = Node::Map;
      if (def && def->kind == Node::Map) params = *def;
      if (inline_params && inline_params->kind == Node::Map)
        for (auto& kv : inline_params->m) params.set(kv.first, kv.second);
      else if (inline_params && inline_params->kind == Node::Scalar)
        params.set("__scalar", *inline_params);  // "- op: value" sets the operator's main slot
      op->second(*this, params);
    } else if (name == "nop") {
    } else {
      throw DeckError("unknown operator '" + name + "'");
    }
    depth--;
  }

  void run_item(const Node& item) {
    if (item.kind == Node::Scalar) run_name(item.s, nullptr);
    else if (item.kind == Node::Map && item.m.size() == 1 && !item.get("body")) run_name(item.m[0].first, &item.m[0].second);
    else if (item.kind == Node::Map) run_block(item);
    else if (item.kind == Node::Seq) run_block(item);
  }

  void run_block(const Node& b) {
    if (b.kind == Node::Seq) {
      for (auto& it : b.q) run_item(it);
      return;
    }
    const Node* body = b.get("body");
    const Node* cond = b.get("condition");
    const Node* loop = b.get("loop");
    const Node* rb = b.get("rebind");
    std::map<std::string, std::string> alias;
    if (rb && rb->kind == Node::Map)
      for (auto& kv : rb->m) alias[kv.first] = resolve(kv.second.s);
    rebinds.push_back(alias);
    const bool is_loop = loop && as_bool(*loop, "loop");
    long guard = 0;
    while (true) {
      if (cond && !eval_condition(*cond)) break;
      if (body) {
        if (body->kind == Node::Seq) run_block(*body);
        else run_item(*body);
      }
      if (!is_loop) break;
      if (++guard > 100000000) throw DeckError("loop does not terminate");
    }
    rebinds.pop_back();
  }
};

// ------------------------------------------------------------------ input and checks
// Operators of the reference's sample decks around the path: particle types and their
// properties, the lattice generator, deterministic position noise, the mass division of the
// forces and the golden-file check.  Single rank, identity xform.

// particle_type_map slot -> type index of a name (particle_types, particle_type_map.h)
static int type_index(Runner& r, const std::string& name) {
  auto it = r.slots.find("particle_type_map");
  if (it != r.slots.end() && it->second.kind == Node::Map) {
    if (const Node* v = it->second.get(name)) return (int)as_double(*v, "particle_type_map");
    throw DeckError("particle type '" + name + "' is not in particle_type_map");
  }
  Node& order = r.slots["__type_order"];  // no map: types numbered in order of first use
  order.kind = Node::Seq;
  for (size_t q = 0; q < order.q.size(); q++)
    if (order.q[q].s == name) return (int)q;
  order.q.push_back(Node::scalar(name));
  return (int)order.q.size() - 1;
}

struct DomainBox {
  double lo[3], hi[3], cs;
  int64_t dims[3];
};
static DomainBox domain_box(Runner& r, const char* what) {
  auto it = r.slots.find("domain");
  if (it == r.slots.end()) throw DeckError(std::string(what) + ": no domain");
  const Node& p = it->second;
  DomainBox b{};
  b.cs = as_double(*p.get("cell_size"), "cell_size");
  for (int a = 0; a < 3; a++) {
    b.lo[a] = as_double(p.get("bounds")->q.at(0).q.at(a), "bounds");
    b.hi[a] = as_double(p.get("bounds")->q.at(1).q.at(a), "bounds");
    b.dims[a] = (int64_t)as_double(p.get("grid_dims")->q.at(a), "grid_dims");
  }
  return b;
}

static void register_input_ops(Runner& R) {
  // particle_types (src/core/particle_type_map.cpp): name -> index map
  R.ops["particle_types"] = [](Runner& r, const Node& p) {
    const Node* m = p.get("particle_type_map");
    if (!m || m->kind != Node::Map) throw DeckError("particle_types: particle_type_map {name: index} is required");
    r.slots["particle_type_map"] = *m;
  };
  // particle_type_add_properties: per type name, scalar properties (mass, charge, ...)
  R.ops["particle_type_add_properties"] = [](Runner& r, const Node& p) {
    Node& props = r.slots["particle_type_properties"];
    props.kind = Node::Map;
    for (auto& kv : p.m) {
      if (kv.second.kind != Node::Map) continue;
      Node* t = props.get(kv.first);
      if (!t) {
        Node e;
        e.kind = Node::Map;
        props.set(kv.first, e);
        t = props.get(kv.first);
      }
      for (auto& q : kv.second.m) t->set(q.first, Node::scalar(fmt(as_double(q.second, q.first))));
    }
  };
  // lattice (lattice_generator.h:100-245, generate_particle_lattice.h:157, 262-383 with
  // deterministic numbering): one thread walks k, j, i from floor(lo / a) - 1 to ceil(hi / a) + 1
  // and the unit cell's sites in order; a site inside the domain bounds (ends included) joins
  // the cell floor((x - lo) / cell_size) if that is an inner cell; ids start after the largest
  // existing id and run through the cells in index order, within a cell in insertion order.
  R.ops["lattice"] = [](Runner& r, const Node& p) {
    xnb_ctx* c = r.need_ctx("lattice");
    size_t n0 = 0, g0 = 0;
    r.check(xnb_num_particles(c, &n0, &g0), "lattice");
    if (n0) throw DeckError("lattice: the grid already holds particles (only an empty grid is filled)");
    const Node* st = p.get("structure");
    const Node* ty = p.get("types");
    const Node* sz = p.get("size");
    if (!st || !ty || !sz || ty->kind != Node::Seq || sz->kind != Node::Seq) throw DeckError("lattice: structure, types and size are required");
    std::vector<std::array<double, 3>> sites;
    if (st->s == "SC") sites = {{0., 0., 0.}};
    else if (st->s == "BCC") sites = {{0., 0., 0.}, {.5, .5, .5}};
    else if (st->s == "FCC") sites = {{0., 0., 0.}, {0., .5, .5}, {.5, 0., .5}, {.5, .5, 0.}};
    else if (st->s == "2BCT") sites = {{0., 0., 0.}, {0., .5, .25}, {.5, .5, .5}, {.5, 0., .75}};
    else throw DeckError("lattice: structure '" + st->s + "' is not supported (SC, BCC, FCC, 2BCT)");
    if (ty->q.size() != sites.size()) throw DeckError("lattice: structure " + st->s + " needs " + std::to_string(sites.size()) + " types");
    std::vector<int> tix;
    for (auto& t : ty->q) tix.push_back(type_index(r, t.s));
    double a[3], shift[3] = {0., 0., 0.};
    for (int d = 0; d < 3; d++) a[d] = as_double(sz->q.at(d), "size");
    if (const Node* sh = p.get("shift"))
      for (int d = 0; d < 3; d++) shift[d] = as_double(sh->q.at(d), "shift");
    const DomainBox b = domain_box(r, "lattice");
    int64_t lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = (int64_t)std::floor(b.lo[d] / a[d]) - 1;
      hi[d] = (int64_t)std::ceil(b.hi[d] / a[d]) + 1;
    }
    const int64_t nc = b.dims[0] * b.dims[1] * b.dims[2];
    std::vector<std::vector<std::array<double, 4>>> cells((size_t)nc);
    for (int64_t k = lo[2]; k <= hi[2]; k++)
      for (int64_t j = lo[1]; j <= hi[1]; j++)
        for (int64_t i = lo[0]; i <= hi[0]; i++)
          for (size_t l = 0; l < sites.size(); l++) {
            const double x[3] = {(i + sites[l][0]) * a[0] + shift[0], (j + sites[l][1]) * a[1] + shift[1],
                                 (k + sites[l][2]) * a[2] + shift[2]};
            int64_t ci[3];
            bool in = true;
            for (int d = 0; d < 3; d++) {
              in = in && x[d] >= b.lo[d] && x[d] <= b.hi[d];
              ci[d] = (int64_t)std::floor((x[d] - b.lo[d]) / b.cs);
              in = in && ci[d] >= 0 && ci[d] < b.dims[d];
            }
            if (!in) continue;
            cells[(size_t)(ci[0] + b.dims[0] * (ci[1] + b.dims[1] * ci[2]))].push_back({x[0], x[1], x[2], (double)tix[l]});
          }
    std::vector<double> rr, vv;
    std::vector<uint64_t> ids;
    std::vector<uint8_t> tt;
    uint64_t next = 1;
    for (auto& cl : cells)
      for (auto& s4 : cl) {
        rr.insert(rr.end(), {s4[0], s4[1], s4[2]});
        vv.insert(vv.end(), {0., 0., 0.});
        ids.push_back(next++);
        tt.push_back((uint8_t)s4[3]);
      }
    r.check(xnb_set_particles(c, ids.size(), rr.data(), vv.data(), ids.data(), tt.data()), "lattice");
    r.log << "lattice " << st->s << ": " << ids.size() << " particles\n";
  };
  // gaussian_noise_r (gaussian_noise.h:62-79, 122-165, gaussian_noise.cpp:32): normal deviates
  // of standard deviation sigma added to rx, ry, rz of every inner particle, cell by cell in
  // particle order; with deterministic_noise each cell seeds its own std::mt19937_64 with
  // (domain cell index) * 1023 and uses a fresh std::normal_distribution.  The particles are
  // then re-binned by their new positions.
  R.ops["gaussian_noise_r"] = [](Runner& r, const Node& p) {
    xnb_ctx* c = r.need_ctx("gaussian_noise_r");
    const double sigma = r.num(p, "sigma", 1.0);
    if (p.get("sigma_cut") || p.get("region")) throw DeckError("gaussian_noise_r: sigma_cut and region are not supported");
    const bool det = r.flag(p, "deterministic_noise", false);
    size_t n = 0, ng = 0;
    r.check(xnb_num_particles(c, &n, &ng), "gaussian_noise_r");
    int64_t dims[3];
    int gl = 0;
    size_t ncells = 0;
    r.check(xnb_grid_info(c, dims, &gl, &ncells), "gaussian_noise_r");
    std::vector<uint32_t> cell(n);
    std::vector<double> rr(3 * n), vv(3 * n);
    std::vector<uint64_t> ids(n);
    std::vector<uint8_t> tt(n);
    r.check(xnb_get_particles(c, 0, cell.data(), rr.data(), vv.data(), nullptr, nullptr, ids.data(), tt.data()), "gaussian_noise_r");
    std::random_device rd;
    std::mt19937_64 free_re(rd());
    std::mt19937_64 det_re;
    const int64_t Di = dims[0] - 2 * gl, Dj = dims[1] - 2 * gl;
    size_t q = 0;
    while (q < n) {
      size_t e = q;
      while (e < n && cell[e] == cell[q]) e++;
      const int64_t ci = (int64_t)cell[q] % dims[0], cj = ((int64_t)cell[q] / dims[0]) % dims[1],
                    ck = (int64_t)cell[q] / (dims[0] * dims[1]);
      const uint64_t dom_idx = (uint64_t)((ci - gl) + Di * ((cj - gl) + Dj * (ck - gl)));
      std::mt19937_64& re = det ? det_re : free_re;
      if (det) det_re.seed(dom_idx * 1023);
      std::normal_distribution<double> gaussian(0.0, sigma);
      for (size_t k = q; k < e; k++)
        for (int d = 0; d < 3; d++) rr[3 * k + d] += gaussian(re);
      q = e;
    }
    r.check(xnb_set_particles(c, n, rr.data(), vv.data(), ids.data(), tt.data()), "gaussian_noise_r");
  };
  // replicate_domain (src/grid_cell_particles/replicate_domain.cpp): only repeat [1, 1, 1]
  R.ops["replicate_domain"] = [](Runner&, const Node& p) {
    if (const Node* rep = p.get("repeat"))
      for (auto& x : rep->q)
        if (as_double(x, "repeat") != 1.0) throw DeckError("replicate_domain: only repeat [1, 1, 1] is supported");
  };
  // divide_force_by_type_scalar (vec3_typescalar_op.cu:52-120, 161): property slot (the scalar
  // form "- divide_force_by_type_scalar: mass" sets it), values from particle_type_properties
  R.ops["divide_force_by_type_scalar"] = [](Runner& r, const Node& p) {
    const Node* pr = p.get("property");
    if (!pr) pr = p.get("__scalar");
    if (!pr) throw DeckError("divide_force_by_type_scalar: property is required");
    auto it = r.slots.find("particle_type_properties");
    if (it == r.slots.end()) throw DeckError("divide_force_by_type_scalar: no particle_type_properties");
    std::vector<double> vals;
    for (auto& kv : it->second.m) {
      const Node* v = kv.second.get(pr->s);
      if (!v) throw DeckError("particle type property '" + pr->s + "' does not exist for type " + kv.first);
      const int t = type_index(r, kv.first);
      if (t < 0 || t > 255) throw DeckError("divide_force_by_type_scalar: type index out of range");
      if ((int)vals.size() <= t) vals.resize((size_t)t + 1, 1.0);
      vals[(size_t)t] = as_double(*v, pr->s);
    }
    if (vals.empty()) throw DeckError("divide_force_by_type_scalar: no types");
    r.check(xnb_divide_force_by_type_scalar(r.need_ctx("divide_force_by_type_scalar"), (int)vals.size(), vals.data(), 0),
            "divide_force_by_type_scalar");
  };
  // check_values (src/debug/check_values.cpp:48-400): compare sampled particles (id, position,
  // acceleration = force field, velocity; hex floats) with a reference file; positions are
  // compared through the periodic image closest to the particle.  A missing file is created
  // (samples particles at evenly spaced indices of the id-sorted set, the reference picks
  // them at random).  Thresholds: pos / acc default `threshold` (1e-16), vel unchecked.
  R.ops["check_values"] = [](Runner& r, const Node& p) {
    xnb_ctx* c = r.need_ctx("check_values");
    const Node* fnode = p.get("file");
    if (!fnode) throw DeckError("check_values: file is required");
    const double thr = r.num(p, "threshold", 1e-16);
    const double pos_thr = p.get("pos_threshold") ? r.num(p, "pos_threshold", thr) : thr;
    const double acc_thr = p.get("acc_threshold") ? r.num(p, "acc_threshold", thr) : thr;
    const double vel_thr = p.get("vel_threshold") ? r.num(p, "vel_threshold", thr) : 1.0e40;
    const bool fatal = r.flag(p, "fatal", true);
    size_t n = 0, ng = 0;
    r.check(xnb_num_particles(c, &n, &ng), "check_values");
    std::vector<double> rr(3 * n), vv(3 * n), ff(3 * n);
    std::vector<uint64_t> ids(n);
    r.check(xnb_get_particles(c, 0, nullptr, rr.data(), vv.data(), ff.data(), nullptr, ids.data(), nullptr), "check_values");
    std::map<uint64_t, size_t> at;
    for (size_t q = 0; q < n; q++) at[ids[q]] = q;
    const DomainBox b = domain_box(r, "check_values");
    std::ifstream fin(fnode->s);
    if (!fin) {
      const long want = std::min<long>((long)r.num(p, "samples", 128), (long)n);
      std::ofstream fout(fnode->s);
      if (!fout) throw DeckError("check_values: cannot create " + fnode->s);
      fout << "length_unit: 1.0 ang\nvalues:\n";
      std::vector<uint64_t> sorted(ids);
      std::sort(sorted.begin(), sorted.end());
      for (long s = 0; s < want; s++) {
        const size_t q = at[sorted[(size_t)((double)s * (double)n / (double)want)]];
        char line[1024];
        std::snprintf(line, sizeof line, "- [%llu, %a, %a, %a, %a, %a, %a, %a, %a, %a]\n", (unsigned long long)ids[q], rr[3 * q],
                      rr[3 * q + 1], rr[3 * q + 2], ff[3 * q], ff[3 * q + 1], ff[3 * q + 2], vv[3 * q], vv[3 * q + 1], vv[3 * q + 2]);
        fout << line;
      }
      r.log << "check_values: created " << fnode->s << " with " << want << " samples\n";
      return;
    }
    std::stringstream ss;
    ss << fin.rdbuf();
    const Node doc = parse_yaml(ss.str());
    const Node* vals = doc.get("values");
    if (!vals || vals->kind != Node::Seq) throw DeckError("check_values: no values in " + fnode->s);
    const double ls = doc.get("length_unit") ? as_double(*doc.get("length_unit"), "length_unit") : 1.0;
    double mr = 0., ma = 0., mv = 0.;
    long bad = 0;
    std::string first_bad;
    for (const Node& row : vals->q) {
      if (row.kind != Node::Seq || row.q.size() < 7) throw DeckError("check_values: bad value row");
      const uint64_t id = (uint64_t)std::strtoull(row.q[0].s.c_str(), nullptr, 10);
      auto f = at.find(id);
      if (f == at.end()) throw DeckError("check_values: particle #" + row.q[0].s + " is not in the grid");
      const size_t q = f->second;
      double er = 0., ea = 0., ev = 0.;
      for (int d = 0; d < 3; d++) {
        const double L = b.hi[d] - b.lo[d];
        double dr = as_double(row.q[1 + d], "r") * ls - rr[3 * q + d];
        dr -= L * std::nearbyint(dr / L);
        er += dr * dr;
        const double da = as_double(row.q[4 + d], "a") * ls - ff[3 * q + d];
        ea += da * da;
        if (row.q.size() >= 10) {
          const double dv = as_double(row.q[7 + d], "v") - vv[3 * q + d];
          ev += dv * dv;
        }
      }
      er = std::sqrt(er);
      ea = std::sqrt(ea);
      ev = std::sqrt(ev);
      mr = std::max(mr, er);
      ma = std::max(ma, ea);
      mv = std::max(mv, ev);
      if (er > pos_thr || ea > acc_thr || ev > vel_thr) {
        if (bad++ == 0)
          first_bad = "particle #" + row.q[0].s + " : re=" + fmt(er) + "/" + fmt(pos_thr) + " ae=" + fmt(ea) + "/" + fmt(acc_thr) +
                      " ve=" + fmt(ev) + "/" + fmt(vel_thr);
      }
    }
    r.out_num("check_values_max_dr", mr);
    r.out_num("check_values_max_da", ma);
    r.out_num("check_values_max_dv", mv);
    r.log << "check_values: " << vals->q.size() << " particles, max |dr| " << fmt(mr) << ", |da| " << fmt(ma) << ", |dv| " << fmt(mv)
          << (bad ? ", FAILED" : ", ok") << "\n";
    if (bad && fatal) throw DeckError("check_values: " + std::to_string(bad) + " particles beyond the thresholds, first " + first_bad);
  };
  // single-rank / host-side no-ops of the reference's default_simulation (main-config.msp:33-49);
  // update_force_from_ghost adds nothing here: forces are computed for inner particles from
  // full lists, ghost particles never receive any
  for (const char* nm : {"logo_banner", "mpi_comm_world", "init_cuda", "update_ghost_config", "unit_system", "print_domain",
                         "memory_stats", "finalize_cuda", "update_force_from_ghost", "trigger_dump_data", "trigger_print_log",
                         "dump_data_if_triggered", "print_log_if_triggered", "global"})
    R.ops[nm] = [](Runner&, const Node&) {};
}

// ------------------------------------------------------------------ operators
static void register_ops(Runner& R) {
  // domain (src/io/domain_init.cpp; Domain, src/core/include/exanb/core/domain.h:36-129) + init_rcb_grid
  R.ops["domain"] = [](Runner& r, const Node& p) {
    xnb_domain d{};
    const Node* cs = p.get("cell_size");
    const Node* gd = p.get("grid_dims");
    const Node* bd = p.get("bounds");
    const Node* pr = p.get("periodic");
    if (!cs || !gd || !bd) throw DeckError("domain: cell_size, grid_dims and bounds are required");
    if (const Node* xf = p.get("xform"))  // only the identity transform
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          if (as_double(xf->q.at(a).q.at(b), "xform") != (a == b ? 1.0 : 0.0)) throw DeckError("domain: only the identity xform is supported");
    d.cell_size = as_double(*cs, "cell_size");
    for (int a = 0; a < 3; a++) {
      d.grid_dims[a] = (int64_t)as_double(gd->q.at(a), "grid_dims");
      d.bounds_min[a] = as_double(bd->q.at(0).q.at(a), "bounds");
      d.bounds_max[a] = as_double(bd->q.at(1).q.at(a), "bounds");
      d.periodic[a] = pr ? as_bool(pr->q.at(a), "periodic") : 1;
    }
    if (r.ctx) xnb_destroy(r.ctx);
    r.ctx = nullptr;
    r.check(xnb_create(&r.ctx, &d, 0), "domain");
    r.slots["domain"] = p;
  };
  R.ops["init_rcb_grid"] = [](Runner& r, const Node&) { r.need_ctx("init_rcb_grid"); };
  // nbh_dist (src/particle_neighbors/nbh_dist.cpp:45-72)
  R.ops["nbh_dist"] = [](Runner& r, const Node& p) {
    double lab, md, gd;
    r.check(xnb_nbh_dist(r.need_ctx("nbh_dist"), r.num(p, "rcut_max", 0.0), r.num(p, "rcut_inc", 0.0), &lab, &md, &gd), "nbh_dist");
    r.out_num("nbh_dist_lab", lab);
    r.out_num("nbh_dist", lab);
    r.out_num("max_displ", md);
    r.out_num("max_displ_lab", md);
    r.out_num("ghost_dist", gd);
  };
  // particles handed over through xnb_deck_set_particles (stands for a reader / lattice operator)
  R.ops["particle_input"] = [](Runner& r, const Node&) {
    const size_t n = r.in_id.size();
    r.check(xnb_set_particles(r.need_ctx("particle_input"), n, r.in_r.data(), r.in_v.empty() ? nullptr : r.in_v.data(),
                              r.in_id.data(), r.in_type.empty() ? nullptr : r.in_type.data()),
            "particle_input");
  };
  // seeded FCC lattice of this library (xnb_generate_fcc)
  R.ops["fcc_lattice"] = [](Runner& r, const Node& p) {
    const int64_t nc = (int64_t)r.num(p, "cells", 0, true);
    const int64_t n = 4 * nc * nc * nc;
    r.in_r.assign(3 * (size_t)n, 0.);
    r.in_v.assign(3 * (size_t)n, 0.);
    r.in_id.assign((size_t)n, 0);
    double L = 0.;
    r.check(xnb_generate_fcc(nc, r.num(p, "density", 0, true), n, r.num(p, "jitter", 0.0), r.num(p, "temperature", 0.0),
                             (uint64_t)r.num(p, "seed", 1.0), r.in_r.data(), r.in_v.data(), r.in_id.data(), &L),
            "fcc_lattice");
    r.in_type.clear();
    r.out_num("box_length", L);
  };
  // move_particles (src/grid_cell_particles/move_particles.cpp:39-47)
  R.ops["move_particles"] = [](Runner& r, const Node&) { r.check(xnb_move_particles(r.need_ctx("move_particles")), "move_particles"); };
  // backup_r (src/io/backup_r.cpp:35-116)
  R.ops["backup_r"] = [](Runner& r, const Node&) { r.check(xnb_backup_r(r.need_ctx("backup_r")), "backup_r"); };
  // ghost_comm_scheme is built together with the ghost copy in ghost_update_all on one rank
  R.ops["ghost_comm_scheme"] = [](Runner& r, const Node&) { r.need_ctx("ghost_comm_scheme"); };
  R.ops["ghost_update_all"] = [](Runner& r, const Node&) {
    r.check(xnb_ghost_update_all(r.need_ctx("ghost_update_all")), "ghost_update_all");
  };
  R.ops["ghost_update_r"] = [](Runner& r, const Node&) { r.check(xnb_ghost_update_r(r.need_ctx("ghost_update_r")), "ghost_update_r"); };
  // particle_displ_over (src/mpi/particle_displ_over.cu:115-179): result = count > 0
  R.ops["particle_displ_over"] = [](Runner& r, const Node& p) {
    uint64_t cnt = 0;
    r.check(xnb_particle_displ_over(r.need_ctx("particle_displ_over"), r.num(p, "threshold", 0.0), &cnt), "particle_displ_over");
    r.out_bool("result", cnt > 0);
  };
  // chunk_neighbors (src/particle_neighbors/chunk_neighbors.cpp:47-90), config slot as in
  // chunk_neighbors_config.h:61-83
  R.ops["chunk_neighbors"] = [](Runner& r, const Node& p) {
    xnb_chunk_neighbors_config c{0, 0, 0, 0, 0};
    if (const Node* cf = r.slot(p, "config")) {
      if (const Node* v = cf->get("build_particle_offset")) c.build_particle_offset = as_bool(*v, "build_particle_offset");
      if (const Node* v = cf->get("dual_particle_offset")) c.dual_particle_offset = as_bool(*v, "dual_particle_offset");
      if (const Node* v = cf->get("random_access")) c.random_access = as_bool(*v, "random_access");
      if (const Node* v = cf->get("half_symmetric")) c.half_symmetric = as_bool(*v, "half_symmetric");
      if (const Node* v = cf->get("skip_ghosts")) c.skip_ghosts = as_bool(*v, "skip_ghosts");
    }
    r.check(xnb_chunk_neighbors(r.need_ctx("chunk_neighbors"), r.num(p, "nbh_dist_lab", 0, true), &c, 0), "chunk_neighbors");
  };
  // single-rank no-ops of update_particles_full / parallel_update_particles
  // rebuild_amr (src/amr/rebuild_amr.cpp:34-69): slot defaults 5.0 and 1 as there
  R.ops["rebuild_amr"] = [](Runner& r, const Node& p) {
    if (r.flag(p, "amr_z_curve", false)) throw DeckError("rebuild_amr: amr_z_curve is not implemented");
    r.check(xnb_rebuild_amr(r.need_ctx("rebuild_amr"), r.num(p, "sub_grid_density", 5.0), (int)r.num(p, "enforced_ordering", 1.0)),
            "rebuild_amr");
  };
  for (const char* nm : {"amr_grid_pairs", "resize_particle_locks", "migrate_cell_particles", "extend_domain",
                         "simple_cost_model", "load_balance_rcb", "grid_stats", "chunk_neighbors_stats", "print_grid"})
    R.ops[nm] = [](Runner&, const Node&) {};
  R.ops["trigger_load_balance"] = [](Runner& r, const Node&) { r.out_bool("trigger_load_balance", false); };
  // zero_particle_force (src/compute/zero_particle_force.cu:19-55)
  R.ops["zero_particle_force"] = [](Runner& r, const Node& p) {
    r.check(xnb_zero_particle_force(r.need_ctx("zero_particle_force"), r.flag(p, "ghost", false)), "zero_particle_force");
  };
  // lennard_jones_force (contribs/md/lennard_jones/lennard_jones.cu:171-223); slot 'mixed' selects
  // the FP32-pair arithmetic
  R.ops["lennard_jones_force"] = [](Runner& r, const Node& p) {
    const Node* cf = r.slot(p, "config");
    if (!cf || cf->kind != Node::Map) throw DeckError("lennard_jones_force: config {epsilon, sigma} is required");
    xnb_lj_params lj{as_double(*cf->get("epsilon"), "epsilon"), as_double(*cf->get("sigma"), "sigma")};
    const double rcut = r.num(p, "rcut", 0.0);
    r.out_num("rcut_max", std::max(rcut, r.num(p, "rcut_max", 0.0)));
    r.check(xnb_lennard_jones_force(r.need_ctx("lennard_jones_force"), &lj, rcut,
                                    r.flag(p, "mixed", false) ? XNB_PRECISION_MIXED : XNB_PRECISION_FP64, r.flag(p, "ghost", false)),
            "lennard_jones_force");
  };
  // push_f_v_r / push_f_v (src/defbox/include/exanb/defbox/push_vec3_2nd_order.h:85-117,
  // push_vec3_1st_order.h:79-106), identity xform
  R.ops["push_f_v_r"] = [](Runner& r, const Node& p) {
    r.check(xnb_push_f_v_r(r.need_ctx("push_f_v_r"), r.num(p, "dt", 0, true), r.num(p, "dt_scale", 0, true)), "push_f_v_r");
  };
  R.ops["push_f_v"] = [](Runner& r, const Node& p) {
    r.check(xnb_push_f_v(r.need_ctx("push_f_v"), r.num(p, "dt", 0, true), r.num(p, "dt_scale", 0, true)), "push_f_v");
  };
  // control operators of compute-loop.msp
  R.ops["nth_timestep"] = [](Runner& r, const Node& p) {
    const long ts = (long)r.num(p, "timestep", 0.0);
    const long fq = (long)r.num(p, "freq", 1.0);
    const bool first = r.flag(p, "first", true);
    r.out_bool("result", fq > 0 && ts % fq == 0 && (first || ts > 0));
  };
  // loop continues while timestep <= end_at: with compute-loop.msp's order (first_iteration
  // advances the timestep to 1, each iteration advances it after its step) end_at steps are
  // done, which the reference's golden check_values_lj_Ni.dat (100 steps) fixes
  R.ops["sim_continue"] = [](Runner& r, const Node& p) {
    r.out_bool("result", r.num(p, "timestep", 0.0) <= r.num(p, "end_at", 0, true));
  };
  R.ops["next_time_step"] = [](Runner& r, const Node& p) {
    r.out_num("timestep", r.num(p, "timestep", 0.0) + 1.0);
    r.out_num("physical_time", r.num(p, "physical_time", 0.0) + r.num(p, "dt", 0.0));
  };
  R.ops["simulation_stats"] = [](Runner& r, const Node&) {
    double ke = 0., pe = 0.;
    r.check(xnb_simulation_energies(r.need_ctx("simulation_stats"), &ke, &pe), "simulation_stats");
    r.out_num("kinetic_energy", ke);
    r.out_num("potential_energy", pe);
  };
  R.ops["print_simulation_stats"] = [](Runner& r, const Node& p) {
    r.log << "step " << (long)r.num(p, "timestep", 0.0) << " ke " << fmt(r.num(p, "kinetic_energy", 0.0)) << " pe "
          << fmt(r.num(p, "potential_energy", 0.0)) << "\n";
  };
  R.ops["message"] = [](Runner& r, const Node& p) {
    if (const Node* m = p.get("text")) r.log << m->s << "\n";
  };
  register_input_ops(R);
}

// merge a parsed deck into the runner: "name:" replaces, "+name:" appends to a sequence
static void merge_deck(Runner& R, const Node& doc) {
  if (doc.kind != Node::Map) throw DeckError("deck must be a mapping at top level");
  for (auto& kv : doc.m) {
    if (!kv.first.empty() && kv.first[0] == '+') {
      const std::string nm = kv.first.substr(1);
      Node* cur = R.deck.get(nm);
      if (!cur) {
        R.deck.set(nm, kv.second);
        continue;
      }
      if (cur->kind != Node::Seq) {
        Node s;
        s.kind = Node::Seq;
        s.q.push_back(Node::scalar(nm + "__base"));
        R.deck.set(nm + "__base", *cur);
        *cur = s;
      }
      if (kv.second.kind == Node::Seq)
        for (auto& x : kv.second.q) cur->q.push_back(x);
      else
        cur->q.push_back(kv.second);
    } else if (kv.first == "global" && kv.second.kind == Node::Map) {
      for (auto& g : kv.second.m) R.slots[g.first] = g.second;
      R.deck.set(kv.first, kv.second);
    } else {
      R.deck.set(kv.first, kv.second);
    }
  }
}

struct Deck {
  Runner R;
  Deck() {
    R.deck.kind = Node::Map;
    register_ops(R);
  }
  ~Deck() {
    if (R.ctx) xnb_destroy(R.ctx);
  }
};

}  // namespace xnbdeck

// ------------------------------------------------------------------ C ABI
using xnbdeck::Deck;

struct xnb_deck {
  Deck d;
  std::string err;
  std::string log;
};

extern "C" {

int xnb_deck_create(xnb_deck** out) {
  if (!out) return XNB_E_ARG;
  *out = new xnb_deck();
  return XNB_OK;
}

int xnb_deck_destroy(xnb_deck* k) {
  delete k;
  return XNB_OK;
}

int xnb_deck_load(xnb_deck* k, const char* yaml_text) {
  if (!k || !yaml_text) return XNB_E_ARG;
  try {
    xnbdeck::merge_deck(k->d.R, xnbdeck::parse_yaml(yaml_text));
  } catch (const std::exception& e) {
    k->err = e.what();
    return XNB_E_ARG;
  }
  return XNB_OK;
}

int xnb_deck_set_particles(xnb_deck* k, size_t n, const double* r, const double* v, const uint64_t* id, const uint8_t* type) {
  if (!k || (n && !r)) return XNB_E_ARG;
  auto& R = k->d.R;
  R.in_r.assign(r, r + 3 * n);
  if (v) R.in_v.assign(v, v + 3 * n);
  else R.in_v.clear();
  R.in_id.resize(n);
  for (size_t q = 0; q < n; q++) R.in_id[q] = id ? id[q] : (uint64_t)q;
  if (type) R.in_type.assign(type, type + n);
  else R.in_type.clear();
  return XNB_OK;
}

int xnb_deck_run(xnb_deck* k, const char* entry) {
  if (!k) return XNB_E_ARG;
  try {
    k->d.R.run_name(entry ? entry : "simulation", nullptr);
  } catch (const std::exception& e) {
    k->err = e.what();
    k->log = k->d.R.log.str();
    return XNB_E_STATE;
  }
  k->log = k->d.R.log.str();
  return XNB_OK;
}

xnb_ctx* xnb_deck_context(xnb_deck* k) { return k ? k->d.R.ctx : nullptr; }

const char* xnb_deck_error(const xnb_deck* k) { return k ? k->err.c_str() : ""; }

const char* xnb_deck_log(const xnb_deck* k) { return k ? k->log.c_str() : ""; }

int xnb_deck_slot(xnb_deck* k, const char* name, char* out, size_t cap) {
  if (!k || !name || !out || cap == 0) return XNB_E_ARG;
  auto it = k->d.R.slots.find(name);
  if (it == k->d.R.slots.end() || it->second.kind != xnbdeck::Node::Scalar) return XNB_E_STATE;
  std::snprintf(out, cap, "%s", it->second.s.c_str());
  return XNB_OK;
}

}  // extern "C"
