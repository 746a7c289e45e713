// Host-side mesh construction and hp-refinement. See mesh.hpp.
#include "mesh.hpp"
#include "parallel.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>

#include "json.hpp"

namespace fem2d {

namespace {
[[noreturn]] void panic(const std::string& msg) { throw MeshPanic(msg); }
}  // namespace

// ---------------------------------------------------------------------------
// FloatRep / Point (space.rs:214-381)
// ---------------------------------------------------------------------------

FloatRep FloatRep::from(double value) {
  double a = std::fabs(value);
  double integer_part = std::trunc(a);
  double frac = a - integer_part;
  double fractional_rounded = std::round(frac / POINT_UNIQUENESS_ACCURACY) * POINT_UNIQUENESS_ACCURACY;
  double total = integer_part + fractional_rounded;
  FloatRep r;
  r.sign_positive = !std::signbit(value);
  std::memcpy(&r.bits, &total, sizeof(double));
  return r;
}

int FloatRep::cmp(const FloatRep& o) const {
  auto c = [](uint64_t a, uint64_t b) { return a < b ? -1 : (a > b ? 1 : 0); };
  if (sign_positive && o.sign_positive) return c(bits, o.bits);
  if (!sign_positive && o.sign_positive) return -1;
  if (sign_positive && !o.sign_positive) return 1;
  return -c(bits, o.bits);
}

Point Point::make(double x, double y) {
  Point p;
  p.x = x;
  p.y = y;
  p.x_cmp = FloatRep::from(x);
  p.y_cmp = FloatRep::from(y);
  return p;
}

Point Point::between(const Point& a, const Point& b) { return make((a.x + b.x) / 2.0, (a.y + b.y) / 2.0); }

ParaDir Point::orientation_with(const Point& other) const {
  if (!(x_cmp != other.x_cmp || y_cmp != other.y_cmp))
    panic("Cannot compute the orientation between two Points at the same location");
  double dx = std::fabs(other.x - x);
  double dy = std::fabs(other.y - y);
  double theta = std::atan(dy / dx);
  return theta < (M_PI / 4.0) ? ParaDir::U : ParaDir::V;
}

double Point::dist(const Point& other) const {
  double dx = std::fabs(other.x - x);
  double dy = std::fabs(other.y - y);
  return std::sqrt(dx * dx + dy * dy);
}

// ---------------------------------------------------------------------------
// Element (element.rs)
// ---------------------------------------------------------------------------

static double map_range(double val, double in_min, double in_max, double out_min, double out_max) {
  return (val - in_min) * (out_max - out_min) / (in_max - in_min) + out_min;
}

std::array<double, 2> Element::parametric_mapping(const Range2& range) const {
  double real_x_min = map_range(range[0][0], -1.0, 1.0, points[0].x, points[3].x);
  double real_x_max = map_range(range[0][1], -1.0, 1.0, points[0].x, points[3].x);
  double real_y_min = map_range(range[1][0], -1.0, 1.0, points[0].y, points[3].y);
  double real_y_max = map_range(range[1][1], -1.0, 1.0, points[0].y, points[3].y);
  return {(real_x_max - real_x_min) / 2.0, (real_y_max - real_y_min) / 2.0};
}

int Element::order_points(const Point& p0, const Point& p1) const {
  return p0.orientation_with(p1) == ParaDir::U ? p0.x_order(p1) : p0.y_order(p1);
}

// ---------------------------------------------------------------------------
// h-refinement types (h_refinement.rs)
// ---------------------------------------------------------------------------

Range2 sub_range(HRefLoc loc, const Range2& r) {
  double mu = (r[0][0] + r[0][1]) / 2.0;
  double mv = (r[1][0] + r[1][1]) / 2.0;
  std::array<double, 2> u = r[0], v = r[1];
  switch (loc) {
    case HRefLoc::SW: u = {r[0][0], mu}; v = {r[1][0], mv}; break;
    case HRefLoc::SE: u = {mu, r[0][1]}; v = {r[1][0], mv}; break;
    case HRefLoc::NW: u = {r[0][0], mu}; v = {mv, r[1][1]}; break;
    case HRefLoc::NE: u = {mu, r[0][1]}; v = {mv, r[1][1]}; break;
    case HRefLoc::W: u = {r[0][0], mu}; break;
    case HRefLoc::E: u = {mu, r[0][1]}; break;
    case HRefLoc::S: v = {r[1][0], mv}; break;
    case HRefLoc::N: v = {mv, r[1][1]}; break;
  }
  return {u, v};
}

HRefLoc HRef::loc(int idx) const {
  switch (kind) {
    case Kind::T: {
      static const HRefLoc t[4] = {HRefLoc::SW, HRefLoc::SE, HRefLoc::NW, HRefLoc::NE};
      return t[idx];
    }
    case Kind::U: return idx == 0 ? HRefLoc::W : HRefLoc::E;
    case Kind::V: return idx == 0 ? HRefLoc::S : HRefLoc::N;
  }
  return HRefLoc::SW;
}

HRef& HRef::operator+=(const HRef& rhs) {
  if (kind == Kind::T) return *this;
  const Kind other = kind == Kind::U ? Kind::V : Kind::U;
  if (rhs.kind == Kind::T || rhs.kind == other) {
    *this = HRef::t();
  } else if (ext >= 0 && rhs.ext >= 0 && rhs.ext != ext) {
    // two extended refinements of the same direction on different children
    *this = HRef::t();
  }
  return *this;
}

HLevels HLevels::refined(const HRef& r) const {
  HLevels h = *this;
  if (r.kind == HRef::Kind::T || r.kind == HRef::Kind::U) h.u = uint8_t(h.u + 1);
  if (r.kind == HRef::Kind::T || r.kind == HRef::Kind::V) h.v = uint8_t(h.v + 1);
  return h;
}

// ---------------------------------------------------------------------------
// p-refinement types (p_refinement.rs)
// ---------------------------------------------------------------------------

PRef PRef::constrained_to(const std::array<std::array<int, 2>, 2>& b) const {
  PRef r = *this;
  auto clamp = [](int d, const std::array<int, 2>& lim) { return d < lim[0] ? lim[0] : (d > lim[1] ? lim[1] : d); };
  r.di = clamp(di, b[0]);
  r.dj = clamp(dj, b[1]);
  return r;
}

bool PRef::falls_within(const std::array<std::array<int, 2>, 2>& b) const {
  return di >= b[0][0] && di <= b[0][1] && dj >= b[1][0] && dj <= b[1][1];
}

static int refine_order(int n, int delta) {
  if (delta > 0) {
    if (n + delta > MAX_POLYNOMIAL_ORDER)
      throw PRefError(PRefErrorKind::ExceededMaxExpansion, 0,
                      "Positive p-Refinement results in expansion order over maximum; Cannot p-Refine!");
    return n + delta;
  }
  if (delta < 0) {
    if (-delta >= n)
      throw PRefError(PRefErrorKind::NegExpansion, 0,
                      "Negative p-Refinement results in 0 or negative expansion order; Cannot p-Refine!");
    return n + delta;
  }
  return n;
}

void PolyOrders::refine(const PRef& r) {
  ni = uint8_t(refine_order(ni, r.di));
  nj = uint8_t(refine_order(nj, r.dj));
}

void PolyOrders::set(int i, int j) {
  if (i > MAX_POLYNOMIAL_ORDER || j > MAX_POLYNOMIAL_ORDER)
    throw PRefError(PRefErrorKind::ExceededMaxExpansion, 0, "expansion order over maximum");
  if (i < 1 || j < 1) throw PRefError(PRefErrorKind::NegExpansion, 0, "expansion order below 1");
  ni = uint8_t(i);
  nj = uint8_t(j);
}

Orders Orders::make(int i, int j) {
  if (i > MAX_POLYNOMIAL_ORDER || j > MAX_POLYNOMIAL_ORDER || i < 1 || j < 1)
    throw PRefError(PRefErrorKind::InvalidExpansionOrders, 0,
                    "Invalid expansion orders: (" + std::to_string(i) + ", " + std::to_string(j) + ")");
  return {i, j};
}

// ---------------------------------------------------------------------------
// Elem (elem.rs)
// ---------------------------------------------------------------------------

Range2 Elem::relative_parametric_range(int from_ancestor) const {
  size_t start = ancestors.size();
  for (size_t k = 0; k < ancestors.size(); ++k)
    if (ancestors[k].first == from_ancestor) {
      start = k;
      break;
    }
  if (start == ancestors.size())
    panic("Elem " + std::to_string(from_ancestor) + " is not an ancestor of Elem " + std::to_string(id) +
          "; cannot compute relative parametric range!");
  Range2 acc{{{-1.0, 1.0}, {-1.0, 1.0}}};
  for (size_t k = start; k < ancestors.size(); ++k) acc = sub_range(ancestors[k].second, acc);
  return acc;
}

Range2 Elem::parametric_range() const {
  Range2 acc{{{-1.0, 1.0}, {-1.0, 1.0}}};
  for (const auto& a : ancestors) acc = sub_range(a.second, acc);
  return acc;
}

// ---------------------------------------------------------------------------
// Edge / Node (edge.rs, node.rs)
// ---------------------------------------------------------------------------

static Edge make_edge(int id, const Node& n0, const Node& n1, bool boundary) {
  Edge e;
  e.id = id;
  e.nodes = {n0.id, n1.id};
  e.boundary = boundary;
  e.dir = n0.coords.orientation_with(n1.coords);
  e.length = n0.coords.dist(n1.coords);
  return e;
}

void Edge::connect_elem(const Elem& elem) {
  int index_of_self = -1;
  for (int k = 0; k < 4; ++k)
    if (elem.edges[k] == id) {
      index_of_self = k;
      break;
    }
  if (index_of_self < 0)
    panic("Elem " + std::to_string(elem.id) + " is not connected to Edge " + std::to_string(id) +
          "; cannot reciprocate connection!");
  uint16_t address = elem.h_levels.edge_ranking(dir);
  int side = (index_of_self == 0 || index_of_self == 2) ? 1 : 0;
  auto it = elems[side].find(address);
  if (it != elems[side].end()) {
    if (it->second != elem.id)
      panic("Edge " + std::to_string(id) + " is already connected to Elem " + std::to_string(it->second) +
            "; cannot connect to Elem " + std::to_string(elem.id));
  } else {
    elems[side][address] = elem.id;
    if (top[side] < 0 || address > top_key[side]) {
      top[side] = elem.id;
      top_key[side] = address;
    }
  }
}

bool Edge::set_activation() {
  if (top[0] >= 0 && top[1] >= 0) {
    active = top;
    has_active = true;
    return true;
  }
  has_active = false;
  return false;
}

int Edge::other_active_elem_id(int elem_id) const {
  if (!has_active) return -1;
  if (active[0] == elem_id) return active[1];
  if (active[1] == elem_id) return active[0];
  panic(std::to_string(elem_id) + " isn't an active Elem on Edge " + std::to_string(id) +
        "; Cannot retrieve other active Elem!");
}

void Node::connect_elem(const Elem& elem) {
  int index_of_self = -1;
  for (int k = 0; k < 4; ++k)
    if (elem.nodes[k] == id) {
      index_of_self = k;
      break;
    }
  if (index_of_self < 0)
    panic("Elem " + std::to_string(elem.id) + " is not connected to Node " + std::to_string(id) +
          "; cannot reciprocate connection!");
  uint16_t address = elem.h_levels.node_ranking();
  auto it = elems[index_of_self].find(address);
  if (it != elems[index_of_self].end()) {
    if (it->second != elem.id)
      panic("Node " + std::to_string(id) + " is already connected to Elem " + std::to_string(it->second) +
            "; cannot connect to Elem " + std::to_string(elem.id));
  } else {
    elems[index_of_self][address] = elem.id;
  }
}

// ---------------------------------------------------------------------------
// Mesh construction
// ---------------------------------------------------------------------------

Mesh Mesh::blank() { return Mesh(); }

Mesh Mesh::unit() {
  Mesh m;
  std::array<Point, 4> pts = {Point::make(-1.0, -1.0), Point::make(1.0, -1.0), Point::make(-1.0, 1.0),
                              Point::make(1.0, 1.0)};
  Element el;
  el.id = 0;
  el.points = pts;
  m.elements.push_back(el);
  Elem e;
  e.id = 0;
  e.nodes = {0, 1, 2, 3};
  e.edges = {0, 1, 2, 3};
  e.element = 0;
  m.elems.push_back(e);
  for (int k = 0; k < 4; ++k) {
    Node n;
    n.id = k;
    n.coords = pts[k];
    n.boundary = true;
    m.nodes.push_back(n);
  }
  // The unit mesh does not connect its Elem to its Edges or Nodes (mesh.rs:59-88).
  m.edges.push_back(make_edge(0, m.nodes[0], m.nodes[1], true));
  m.edges.push_back(make_edge(1, m.nodes[2], m.nodes[3], true));
  m.edges.push_back(make_edge(2, m.nodes[0], m.nodes[2], true));
  m.edges.push_back(make_edge(3, m.nodes[1], m.nodes[3], true));
  return m;
}

Mesh Mesh::from_parts(const std::vector<Point>& points, const std::vector<ElementSpec>& specs) {
  const int num_nodes = int(points.size());
  for (const auto& s : specs) {
    for (int k = 0; k < 4; ++k)
      if (s.node_ids[k] < 0 || s.node_ids[k] >= num_nodes)
        panic("node_ids must be smaller than the total number of nodes!");
    for (int a = 0; a < 4; ++a)
      for (int b = a + 1; b < 4; ++b)
        if (s.node_ids[a] == s.node_ids[b]) panic("Element's node_ids should have 4 unique values!");
  }
  {
    // has_duplicates over the points (mesh.rs:1777-1797), by sorting the comparison keys
    std::vector<size_t> order(points.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = k;
    auto key_less = [&](size_t a, size_t b) {
      int cx = points[a].x_cmp.cmp(points[b].x_cmp);
      if (cx != 0) return cx < 0;
      return points[a].y_cmp.cmp(points[b].y_cmp) < 0;
    };
    std::sort(order.begin(), order.end(), key_less);
    for (size_t k = 1; k < order.size(); ++k)
      if (points[order[k - 1]] == points[order[k]]) panic("All Nodes must be at unique locations!");
  }

  Mesh m;
  for (size_t id = 0; id < specs.size(); ++id) {
    Element el;
    el.id = int(id);
    for (int k = 0; k < 4; ++k) el.points[k] = points[specs[id].node_ids[k]];
    el.materials = specs[id].materials;
    m.elements.push_back(el);
  }

  std::vector<int> counts(points.size(), 0);
  for (const auto& s : specs)
    for (int k = 0; k < 4; ++k) counts[s.node_ids[k]] += 1;
  for (int id = 0; id < num_nodes; ++id) {
    if (counts[id] > 4) panic("Nodes can only be shared by a maximum of 4 Elements; Cannot construct Mesh from file!");
    Node n;
    n.id = id;
    n.coords = points[id];
    n.boundary = counts[id] < 4;
    m.nodes.push_back(n);
  }

  // {[node_0, node_1] -> [element below/left, element above/right]} ordered by
  // the node-id pair; the rank in this order is the edge id (mesh.rs:200-258).
  static const int kEdgeDefs[4][3] = {{0, 1, 1}, {2, 3, 0}, {0, 2, 1}, {1, 3, 0}};
  std::map<std::array<int, 2>, std::array<int, 2>> pairs;
  for (size_t eid = 0; eid < specs.size(); ++eid) {
    for (const auto& d : kEdgeDefs) {
      std::array<int, 2> key = {specs[eid].node_ids[d[0]], specs[eid].node_ids[d[1]]};
      int side = d[2];
      auto it = pairs.find(key);
      if (it != pairs.end()) {
        if (it->second[side] != -1)
          panic("Edge side of Elem has already been set; cannot set to " + std::to_string(eid));
        it->second[side] = int(eid);
      } else {
        std::array<int, 2> v = {-1, -1};
        v[side] = int(eid);
        pairs[key] = v;
      }
    }
  }

  std::vector<std::array<int, 4>> elem_edges(specs.size(), {-1, -1, -1, -1});
  int edge_id = 0;
  for (const auto& kv : pairs) {
    int nadj = (kv.second[0] >= 0) + (kv.second[1] >= 0);
    if (nadj == 0) panic("Edge has no adjacent Elements; cannot construct Mesh from file!");
    m.edges.push_back(make_edge(edge_id, m.nodes[kv.first[0]], m.nodes[kv.first[1]], nadj == 1));
    const ParaDir dir = m.edges.back().dir;
    for (int side = 0; side < 2; ++side) {
      int el = kv.second[side];
      if (el < 0) continue;
      int edge_idx = side == 0 ? (dir == ParaDir::U ? 1 : 3) : (dir == ParaDir::U ? 0 : 2);
      if (elem_edges[el][edge_idx] != -1)
        panic("Edge (" + std::to_string(edge_idx) + ") of Elem " + std::to_string(el) +
              " has already been set; cannot set to " + std::to_string(edge_id) + "!");
      elem_edges[el][edge_idx] = edge_id;
    }
    ++edge_id;
  }

  for (size_t eid = 0; eid < specs.size(); ++eid) {
    Elem e;
    e.id = int(eid);
    e.nodes = specs[eid].node_ids;
    for (int k = 0; k < 4; ++k)
      if (elem_edges[eid][k] < 0) panic("Elem " + std::to_string(eid) + " is missing an Edge");
    e.edges = elem_edges[eid];
    e.element = int(eid);
    for (int k = 0; k < 4; ++k) m.edges[e.edges[k]].connect_elem(e);
    for (int k = 0; k < 4; ++k) m.nodes[e.nodes[k]].connect_elem(e);
    m.elems.push_back(e);
  }

  m.set_edge_activation();
  return m;
}

Mesh Mesh::from_json_text(const std::string& text) {
  json::Value doc = json::parse(text);
  const json::Value* jel = doc.get("Elements");
  const json::Value* jnd = doc.get("Nodes");
  if (!jel || !jel->is_array()) panic("Elements must be an Array!");
  if (!jnd || !jnd->is_array()) panic("Nodes must be an Array!");
  std::vector<Point> points;
  for (const auto& n : jnd->items) {
    if (!n.is_array()) panic("nodes must be arrays!");
    if (n.items.size() != 2) panic("nodes must be arrays of length 2!");
    if (!n.items[0].is_number() || !n.items[1].is_number())
      panic("nodes must be composed of numerical values!");
    points.push_back(Point::make(n.items[0].number, n.items[1].number));
  }
  std::vector<ElementSpec> specs;
  for (const auto& e : jel->items) {
    const json::Value* ids = e.get("node_ids");
    const json::Value* mats = e.get("materials");
    if (!ids || !ids->is_array()) panic("Elements must have an Array of node_ids!");
    if (ids->items.size() != 4) panic("Elements Array of node_ids must have a length of 4!");
    if (!mats || !mats->is_array()) panic("Elements must have an Array of materials!");
    if (mats->items.size() != 4) panic("Elements Array of materials must have a length of 4!");
    ElementSpec s;
    for (int k = 0; k < 4; ++k) {
      const auto& v = ids->items[k];
      if (!v.is_number() || !v.is_integer || v.number < 0) panic("node_ids must be positive integers!");
      s.node_ids[k] = int(v.number);
    }
    double mp[4];
    for (int k = 0; k < 4; ++k) {
      if (!mats->items[k].is_number()) panic("Element materials must be numerical values");
      mp[k] = mats->items[k].number;
    }
    s.materials = {mp[0], mp[1], mp[2], mp[3]};
    specs.push_back(s);
  }
  return from_parts(points, specs);
}

Mesh Mesh::from_file(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw IoError("unable to open mesh file: " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return from_json_text(ss.str());
}

// ---------------------------------------------------------------------------
// Queries
// ---------------------------------------------------------------------------

std::array<const Point*, 4> Mesh::elem_points(int elem_id) const {
  if (elem_id < 0 || elem_id >= int(elems.size()))
    throw MeshAccessError("Attempt to access non-existent elem: " + std::to_string(elem_id) + "!");
  const Elem& e = elems[elem_id];
  return {&nodes[e.nodes[0]].coords, &nodes[e.nodes[1]].coords, &nodes[e.nodes[2]].coords,
          &nodes[e.nodes[3]].coords};
}

std::array<const Point*, 2> Mesh::edge_points(int edge_id) const {
  if (edge_id < 0 || edge_id >= int(edges.size()))
    throw MeshAccessError("Attempt to access non-existent edge: " + std::to_string(edge_id) + "!");
  return {&nodes[edges[edge_id].nodes[0]].coords, &nodes[edges[edge_id].nodes[1]].coords};
}

std::vector<int> Mesh::descendant_elems(int elem_id, bool include_start) const {
  if (elem_id < 0 || elem_id >= int(elems.size()))
    throw MeshAccessError("Attempt to access non-existent elem: " + std::to_string(elem_id) + "!");
  std::vector<int> out;
  // pre-order DFS, children in stored order
  std::vector<std::pair<int, bool>> stack = {{elem_id, include_start}};
  while (!stack.empty()) {
    auto [id, inc] = stack.back();
    stack.pop_back();
    if (inc) out.push_back(id);
    const auto& ch = elems[id].children;
    for (auto it = ch.rbegin(); it != ch.rend(); ++it) stack.push_back({*it, true});
  }
  return out;
}

std::vector<int> Mesh::ancestor_elems(int elem_id, bool include_start) const {
  if (elem_id < 0 || elem_id >= int(elems.size()))
    throw MeshAccessError("Attempt to access non-existent elem: " + std::to_string(elem_id) + "!");
  std::vector<int> out;
  if (include_start) out.push_back(elem_id);
  int p = elems[elem_id].parent_id();
  while (p >= 0) {
    out.push_back(p);
    p = elems[p].parent_id();
  }
  return out;
}

std::vector<int> Mesh::descendant_edges(int edge_id, bool include_start) const {
  if (edge_id < 0 || edge_id >= int(edges.size()))
    throw MeshAccessError("Attempt to access non-existent edge: " + std::to_string(edge_id) + "!");
  std::vector<int> out;
  std::vector<std::pair<int, bool>> stack = {{edge_id, include_start}};
  while (!stack.empty()) {
    auto [id, inc] = stack.back();
    stack.pop_back();
    if (inc) out.push_back(id);
    if (edges[id].has_children()) {
      stack.push_back({edges[id].children[1], true});
      stack.push_back({edges[id].children[0], true});
    }
  }
  return out;
}

std::array<int, 2> Mesh::max_expansion_orders() const {
  std::array<int, 2> acc = {0, 0};
  for (const auto& e : elems) {
    acc[0] = std::max<int>(acc[0], e.poly_orders.ni);
    acc[1] = std::max<int>(acc[1], e.poly_orders.nj);
  }
  return acc;
}

bool Mesh::elem_is_h_refineable(int elem_id) const {
  if (elem_id < 0 || elem_id >= int(elems.size()))
    throw MeshAccessError("Attempt to access non-existent elem: " + std::to_string(elem_id) + "!");
  const Elem& e = elems[elem_id];
  if (e.has_children()) return false;
  for (int k = 0; k < 4; ++k)
    if (!(edges[e.edges[k]].length > MIN_EDGE_LENGTH)) return false;
  return true;
}

std::array<std::array<int, 2>, 2> Mesh::elem_p_refinement_window(int elem_id) const {
  if (elem_id < 0 || elem_id >= int(elems.size()))
    throw MeshAccessError("Attempt to access non-existent elem: " + std::to_string(elem_id) + "!");
  const auto& po = elems[elem_id].poly_orders;
  return {{{-(po.ni - 1), MAX_POLYNOMIAL_ORDER - po.ni}, {-(po.nj - 1), MAX_POLYNOMIAL_ORDER - po.nj}}};
}

// ---------------------------------------------------------------------------
// h-refinement (mesh.rs:693-1178)
// ---------------------------------------------------------------------------

struct Mesh::ElemUninit {
  int id = 0;
  std::array<int, 4> nodes{{-1, -1, -1, -1}};
  std::array<int, 4> edges{{-1, -1, -1, -1}};
  int element = 0;
  std::vector<std::pair<int, HRefLoc>> ancestors;
  HLevels h_levels;
  PolyOrders poly_orders;

  void set_node(int idx, int node_id) {
    if (nodes[idx] >= 0) {
      if (nodes[idx] != node_id)
        panic("Node (" + std::to_string(idx) + ") has already been set on ElemUninit " + std::to_string(id));
    } else {
      nodes[idx] = node_id;
    }
  }
  void set_edge(int idx, int edge_id) {
    if (edges[idx] >= 0)
      panic("Edge (" + std::to_string(idx) + ") has already been set on ElemUninit " + std::to_string(id));
    edges[idx] = edge_id;
  }
  Elem into_elem() const {
    for (int k = 0; k < 4; ++k)
      if (nodes[k] < 0 || edges[k] < 0)
        throw HRefError(HRefErrorKind::UninitializedElem, id,
                        "ElemUninit " + std::to_string(id) + " was not fully initialized by the conclusion of h-refinement");
    Elem e;
    e.id = id;
    e.nodes = nodes;
    e.edges = edges;
    e.element = element;
    e.h_levels = h_levels;
    e.poly_orders = poly_orders;
    e.ancestors = ancestors;
    return e;
  }
};

void Mesh::global_h_refinement(const HRef& r) {
  std::vector<std::pair<int, HRef>> refs;
  for (const auto& e : elems)
    if (elem_is_h_refineable(e.id)) refs.push_back({e.id, r});
  execute_h_refinements(refs);
}

void Mesh::h_refine_elems(const std::vector<int>& ids, const HRef& r) {
  // any repeated id fails the batch (mesh.rs:754-760 compares every pair; a
  // sorted copy finds the same repeats in n log n)
  std::vector<int> sorted(ids);
  std::sort(sorted.begin(), sorted.end());
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
    throw HRefError(HRefErrorKind::DuplicateElemIds, -1, "Duplicate element ids in h-Refinement; Cannot h-Refine!");
  std::vector<std::pair<int, HRef>> refs;
  for (int id : ids) refs.push_back({id, r});
  execute_h_refinements(refs);
}

// How one refinement creates its Nodes and Edges: appended as they are made
// (the reference's sequential loop), or at the ids planned for the whole batch
struct Mesh::RefineOps {
  virtual ~RefineOps() = default;
  virtual int center_node(const Point& p) = 0;
  // the child Edges of an Edge and the Node between them (split now if needed)
  virtual std::pair<std::array<int, 2>, int> split(int edge_id) = 0;
  virtual int inner_edge(std::array<int, 2> node_ids, int parent_elem) = 0;
};

struct Mesh::AppendOps : Mesh::RefineOps {
  Mesh& m;
  explicit AppendOps(Mesh& mesh) : m(mesh) {}
  int center_node(const Point& p) override {
    Node n;
    n.id = int(m.nodes.size());
    n.coords = p;
    n.boundary = false;
    m.nodes.push_back(n);
    return n.id;
  }
  std::pair<std::array<int, 2>, int> split(int edge_id) override { return m.h_refine_edge_if_needed(edge_id); }
  int inner_edge(std::array<int, 2> node_ids, int parent_elem) override {
    return m.new_edge_between_nodes(node_ids, parent_elem);
  }
};

std::vector<Mesh::ElemUninit> Mesh::uninit_children(int elem_id, const HRef& r, int first_id) {
  std::vector<ElemUninit> ne;
  const int n = r.num_children();
  for (int idx = 0; idx < n; ++idx) {
    ElemUninit u;
    u.id = first_id + idx;
    u.element = elems[elem_id].element;
    u.ancestors = elems[elem_id].ancestors;
    u.ancestors.push_back({elem_id, r.loc(idx)});
    u.h_levels = elems[elem_id].h_levels.refined(r);
    u.poly_orders = elems[elem_id].poly_orders;
    ne.push_back(u);
    elems[elem_id].children.push_back(u.id);
  }
  return ne;
}

void Mesh::execute_h_refinements(const std::vector<std::pair<int, HRef>>& refinements, int threads) {
  if (threads <= 0) threads = default_threads();
  std::map<int, HRef> refmap;
  for (const auto& [elem_id, h] : refinements) {
    if (elem_id < 0 || elem_id >= int(elems.size()))
      throw HRefError(HRefErrorKind::ElemDoesNotExist, elem_id,
                      "Elem " + std::to_string(elem_id) + " does not exist; cannot apply h-Refinement!");
    if (!elem_is_h_refineable(elem_id))
      throw HRefError(HRefErrorKind::ElemNotRefineable, elem_id,
                      "Elem " + std::to_string(elem_id) +
                          " cannot be h-refined; it is either too small or has already been refined!");
    auto it = refmap.find(elem_id);
    if (it == refmap.end())
      refmap[elem_id] = h;
    else
      it->second += h;
  }

  std::vector<std::pair<int, HRef>> extensions;
  const std::vector<std::pair<int, HRef>> batch(refmap.begin(), refmap.end());
  if (threads == 1 || batch.size() < kParallelRefinements ||
      !execute_h_refinements_planned(batch, threads, extensions)) {
    int elem_id_tracker = int(elems.size());
    AppendOps ops(*this);
    for (const auto& [elem_id, r] : batch) {
      // Elem::h_refine (elem.rs:128-155)
      if (elems[elem_id].has_children())
        throw HRefError(HRefErrorKind::ElemHasChildren, elem_id,
                        "Elem " + std::to_string(elem_id) + " already has children; Cannot h-refine!");
      std::vector<ElemUninit> ne = uninit_children(elem_id, r, elem_id_tracker);
      elem_id_tracker += r.num_children();
      if (r.kind == HRef::Kind::T) {
        execute_t_refinement(ne, elem_id, ops);
      } else if (r.kind == HRef::Kind::U) {
        execute_u_refinement(ne, elem_id, ops);
        if (r.ext >= 0) extensions.push_back({ne[size_t(r.ext)].id, HRef::v()});
      } else {
        execute_v_refinement(ne, elem_id, ops);
        if (r.ext >= 0) extensions.push_back({ne[size_t(r.ext)].id, HRef::u()});
      }
      for (auto& e : upgrade_uninit_elems(ne)) elems.push_back(std::move(e));
    }
  }

  if (!extensions.empty()) execute_h_refinements(extensions, threads);
  set_edge_activation(threads);
}

// The outer Edge slots a refinement splits (if not yet split), in the order the
// sequential loop visits them
static int refine_slots(HRef::Kind kind, int (&slot)[4]) {
  if (kind == HRef::Kind::T) {
    for (int k = 0; k < 4; ++k) slot[k] = k;
    return 4;
  }
  if (kind == HRef::Kind::U) {
    slot[0] = 0;
    slot[1] = 1;
  } else {
    slot[0] = 2;
    slot[1] = 3;
  }
  return 2;
}

// Ids of everything a batch creates, assigned in the sequential loop's order
// from per-refinement counts: refinement k creates its children, then (T) a
// centre Node, and for each outer Edge slot in order the split of that Edge
// (two Edges and a Node) when k is the batch's first refinement to meet the
// Edge unsplit, and (T) the inner Edge of the slot; U and V create their one
// inner Edge after both slots.
struct Mesh::RefinePlan {
  std::vector<int> first_child, center, inner;  // inner: 4 per refinement (-1 unused)
  std::vector<int> split_c0, split_c1, split_node;  // per Edge split by the batch, else -1
  std::vector<int> split_owner;                     // the refinement that splits it, else -1
  int new_elems = 0, new_nodes = 0, new_edges = 0;
};

struct Mesh::PlannedOps : Mesh::RefineOps {
  Mesh& m;
  const RefinePlan& plan;
  int k, used = 0;
  PlannedOps(Mesh& mesh, const RefinePlan& p, int refinement) : m(mesh), plan(p), k(refinement) {}
  int center_node(const Point& p) override {
    Node n;
    n.id = plan.center[size_t(k)];
    n.coords = p;
    n.boundary = false;
    m.nodes[size_t(n.id)] = n;
    return n.id;
  }
  std::pair<std::array<int, 2>, int> split(int edge_id) override {
    const size_t e = size_t(edge_id);
    if (plan.split_owner[e] >= 0) return {{plan.split_c0[e], plan.split_c1[e]}, plan.split_node[e]};
    return {m.edges[e].children, m.edges[e].child_node};  // split before the batch
  }
  int inner_edge(std::array<int, 2> node_ids, int parent_elem) override {
    const int id = plan.inner[size_t(4 * k + used++)];
    const Node& n0 = m.nodes[size_t(node_ids[0])];
    const Node& n1 = m.nodes[size_t(node_ids[1])];
    const int ord = m.elements[size_t(m.elems[size_t(parent_elem)].element)].order_points(n0.coords, n1.coords);
    m.edges[size_t(id)] = ord < 0 ? make_edge(id, n0, n1, false) : make_edge(id, n1, n0, false);
    return id;
  }
};

bool Mesh::execute_h_refinements_planned(const std::vector<std::pair<int, HRef>>& batch, int threads,
                                         std::vector<std::pair<int, HRef>>& extensions) {
  const size_t K = batch.size(), n_edges0 = edges.size();
  RefinePlan P;
  P.first_child.resize(K);
  P.center.assign(K, -1);
  P.inner.assign(4 * K, -1);
  P.split_c0.assign(n_edges0, -1);
  P.split_c1.assign(n_edges0, -1);
  P.split_node.assign(n_edges0, -1);
  P.split_owner.assign(n_edges0, -1);
  // the plan, in the batch's order (integer work only), with the checks that
  // would make the sequential loop throw: then nothing is changed here
  int elem_next = int(elems.size()), node_next = int(nodes.size()), edge_next = int(n_edges0);
  for (size_t k = 0; k < K; ++k) {
    const int elem_id = batch[k].first;
    const HRef& r = batch[k].second;
    if (elems[size_t(elem_id)].has_children()) return false;
    P.first_child[k] = elem_next;
    elem_next += r.num_children();
    const bool t = r.kind == HRef::Kind::T;
    auto pp = elem_points(elem_id);
    const Point centre = Point::between(*pp[0], *pp[3]);
    if (t) P.center[k] = node_next++;
    int slot[4];
    const int n_slots = refine_slots(r.kind, slot);
    Point shared[4];
    for (int s = 0; s < n_slots; ++s) {
      const int e = elems[size_t(elem_id)].edges[size_t(slot[s])];
      const Edge& ed = edges[size_t(e)];
      if (ed.has_children()) {
        shared[s] = nodes[size_t(ed.child_node)].coords;
      } else {
        if (P.split_owner[size_t(e)] < 0) {
          if (ed.length / 2.0 < MIN_EDGE_LENGTH) return false;
          P.split_owner[size_t(e)] = int(k);
          P.split_c0[size_t(e)] = edge_next;
          P.split_c1[size_t(e)] = edge_next + 1;
          P.split_node[size_t(e)] = node_next++;
          edge_next += 2;
        }
        shared[s] = Point::between(nodes[size_t(ed.nodes[0])].coords, nodes[size_t(ed.nodes[1])].coords);
      }
      if (t) {
        P.inner[4 * k + size_t(s)] = edge_next++;
        if (elements[size_t(elems[size_t(elem_id)].element)].order_points(shared[s], centre) == 0) return false;
      }
    }
    if (!t) {
      P.inner[4 * k] = edge_next++;
      if (elements[size_t(elems[size_t(elem_id)].element)].order_points(shared[0], shared[1]) == 0) return false;
    }
  }
  P.new_elems = elem_next - int(elems.size());
  P.new_nodes = node_next - int(nodes.size());
  P.new_edges = edge_next - int(n_edges0);

  // storage for everything new, then three parallel passes: the splits (their
  // Edges and Nodes, which other refinements of the batch share), the
  // refinements themselves, the Edges' and Nodes' Elem maps
  const size_t elems0 = elems.size();
  elems.resize(elems0 + size_t(P.new_elems));
  nodes.resize(nodes.size() + size_t(P.new_nodes));
  edges.resize(n_edges0 + size_t(P.new_edges));
  parallel_ranges(K, threads, 64, [&](size_t b, size_t e) {
    for (size_t k = b; k < e; ++k) {
      const int elem_id = batch[k].first;
      int slot[4];
      const int n_slots = refine_slots(batch[k].second.kind, slot);
      for (int s = 0; s < n_slots; ++s) {
        const int eid = elems[size_t(elem_id)].edges[size_t(slot[s])];
        if (P.split_owner[size_t(eid)] != int(k)) continue;
        // Edge::h_refine (edge.rs:126-170), at the planned ids
        Edge& parent = edges[size_t(eid)];
        const int new0 = P.split_c0[size_t(eid)], new1 = P.split_c1[size_t(eid)], node = P.split_node[size_t(eid)];
        parent.has_children_ = true;
        parent.children = {new0, new1};
        parent.child_node = node;
        Edge c0, c1;
        c0.id = new0;
        c0.nodes = {parent.nodes[0], node};
        c1.id = new1;
        c1.nodes = {node, parent.nodes[1]};
        for (Edge* c : {&c0, &c1}) {
          c->boundary = parent.boundary;
          c->dir = parent.dir;
          c->length = parent.length / 2.0;
          c->parent = eid;
        }
        edges[size_t(new0)] = c0;
        edges[size_t(new1)] = c1;
        Node n;
        n.id = node;
        n.coords = Point::between(nodes[size_t(parent.nodes[0])].coords, nodes[size_t(parent.nodes[1])].coords);
        n.boundary = parent.boundary;
        nodes[size_t(node)] = n;
      }
    }
  });
  std::vector<std::vector<std::pair<int, HRef>>> ext(K);
  parallel_ranges(K, threads, 64, [&](size_t b, size_t e) {
    for (size_t k = b; k < e; ++k) {
      const int elem_id = batch[k].first;
      const HRef& r = batch[k].second;
      std::vector<ElemUninit> ne = uninit_children(elem_id, r, P.first_child[k]);
      PlannedOps ops(*this, P, int(k));
      if (r.kind == HRef::Kind::T) {
        execute_t_refinement(ne, elem_id, ops);
      } else if (r.kind == HRef::Kind::U) {
        execute_u_refinement(ne, elem_id, ops);
        if (r.ext >= 0) ext[k].push_back({ne[size_t(r.ext)].id, HRef::v()});
      } else {
        execute_v_refinement(ne, elem_id, ops);
        if (r.ext >= 0) ext[k].push_back({ne[size_t(r.ext)].id, HRef::u()});
      }
      for (const ElemUninit& u : ne) elems[size_t(u.id)] = u.into_elem();
    }
  });
  for (auto& v : ext) extensions.insert(extensions.end(), v.begin(), v.end());
  // an Elem's connection to its Edges and Nodes (maps keyed by h-level, one
  // entry per Elem: the order of insertion does not change them); striped locks
  constexpr size_t kStripes = 1024;
  std::vector<std::mutex> edge_lock(kStripes), node_lock(kStripes);
  parallel_ranges(size_t(P.new_elems), threads, 256, [&](size_t b, size_t e) {
    for (size_t x = elems0 + b; x < elems0 + e; ++x) {
      const Elem& el = elems[x];
      for (int k = 0; k < 4; ++k) {
        const size_t eid = size_t(el.edges[size_t(k)]);
        std::lock_guard<std::mutex> g(edge_lock[eid % kStripes]);
        edges[eid].connect_elem(el);
      }
      for (int k = 0; k < 4; ++k) {
        const size_t nid = size_t(el.nodes[size_t(k)]);
        std::lock_guard<std::mutex> g(node_lock[nid % kStripes]);
        nodes[nid].connect_elem(el);
      }
    }
  });
  return true;
}

void Mesh::execute_t_refinement(std::vector<ElemUninit>& ne, int parent, RefineOps& ops) {
  if (ne.size() != 4) panic("T refinement needs 4 children");
  auto pp = elem_points(parent);
  const int center_id = ops.center_node(Point::between(*pp[0], *pp[3]));

  for (int idx = 0; idx < 4; ++idx) {
    ne[idx].set_node(3 - idx, center_id);
    ne[idx].set_node(idx, elems[parent].nodes[idx]);
  }

  // (parent edge index, adjacent children, shared node slots, internal edge slots)
  static const int kT[4][7] = {
      {0, 0, 1, 1, 0, 3, 2},
      {1, 2, 3, 3, 2, 3, 2},
      {2, 0, 2, 2, 0, 1, 0},
      {3, 1, 3, 3, 1, 1, 0},
  };
  for (const auto& t : kT) {
    auto [child_edges, shared] = ops.split(elems[parent].edges[t[0]]);
    ne[t[1]].set_edge(t[0], child_edges[0]);
    ne[t[2]].set_edge(t[0], child_edges[1]);
    ne[t[1]].set_node(t[3], shared);
    ne[t[2]].set_node(t[4], shared);
    int new_edge = ops.inner_edge({shared, center_id}, parent);
    ne[t[1]].set_edge(t[5], new_edge);
    ne[t[2]].set_edge(t[6], new_edge);
  }
}

void Mesh::execute_u_refinement(std::vector<ElemUninit>& ne, int parent, RefineOps& ops) {
  if (ne.size() != 2) panic("U refinement needs 2 children");
  std::array<int, 2> outer = {0, 0};
  static const int kU[2][5] = {{0, 1, 0, 0, 1}, {1, 3, 2, 2, 3}};
  for (const auto& t : kU) {
    auto [child_edges, shared] = ops.split(elems[parent].edges[t[0]]);
    outer[t[0]] = shared;
    ne[0].set_edge(t[0], child_edges[0]);
    ne[1].set_edge(t[0], child_edges[1]);
    ne[0].set_node(t[1], shared);
    ne[1].set_node(t[2], shared);
    ne[0].set_node(t[3], elems[parent].nodes[t[3]]);
    ne[1].set_node(t[4], elems[parent].nodes[t[4]]);
  }
  int new_edge = ops.inner_edge(outer, parent);
  ne[0].set_edge(3, new_edge);
  ne[1].set_edge(2, new_edge);
  ne[0].set_edge(2, elems[parent].edges[2]);
  ne[1].set_edge(3, elems[parent].edges[3]);
}

void Mesh::execute_v_refinement(std::vector<ElemUninit>& ne, int parent, RefineOps& ops) {
  if (ne.size() != 2) panic("V refinement needs 2 children");
  std::array<int, 2> outer = {0, 0};
  static const int kV[2][5] = {{2, 2, 0, 0, 2}, {3, 3, 1, 1, 3}};
  for (const auto& t : kV) {
    auto [child_edges, shared] = ops.split(elems[parent].edges[t[0]]);
    outer[t[0] - 2] = shared;
    ne[0].set_edge(t[0], child_edges[0]);
    ne[1].set_edge(t[0], child_edges[1]);
    ne[0].set_node(t[1], shared);
    ne[1].set_node(t[2], shared);
    ne[0].set_node(t[3], elems[parent].nodes[t[3]]);
    ne[1].set_node(t[4], elems[parent].nodes[t[4]]);
  }
  int new_edge = ops.inner_edge(outer, parent);
  ne[0].set_edge(1, new_edge);
  ne[1].set_edge(0, new_edge);
  ne[0].set_edge(0, elems[parent].edges[0]);
  ne[1].set_edge(1, elems[parent].edges[1]);
}

std::pair<std::array<int, 2>, int> Mesh::h_refine_edge_if_needed(int edge_id) {
  if (edges[edge_id].has_children()) return {edges[edge_id].children, edges[edge_id].child_node};

  const int new0 = int(edges.size()), new1 = new0 + 1;
  const int new_node = int(nodes.size());
  Edge& parent = edges[edge_id];
  // Edge::h_refine (edge.rs:126-170)
  const double child_len = parent.length / 2.0;
  if (child_len < MIN_EDGE_LENGTH)
    throw HRefError(HRefErrorKind::MinEdgeLength, edge_id,
                    "h-refinement will result in Edge length below minimum; Cannot h-refine Edge " +
                        std::to_string(edge_id) + "!");
  parent.has_children_ = true;
  parent.children = {new0, new1};
  parent.child_node = new_node;
  Edge c0, c1;
  c0.id = new0;
  c0.nodes = {parent.nodes[0], new_node};
  c1.id = new1;
  c1.nodes = {new_node, parent.nodes[1]};
  for (Edge* c : {&c0, &c1}) {
    c->boundary = parent.boundary;
    c->dir = parent.dir;
    c->length = child_len;
    c->parent = edge_id;
  }
  const bool boundary = parent.boundary;
  const Point mid = Point::between(nodes[parent.nodes[0]].coords, nodes[parent.nodes[1]].coords);
  edges.push_back(c0);
  edges.push_back(c1);
  Node n;
  n.id = new_node;
  n.coords = mid;
  n.boundary = boundary;
  nodes.push_back(n);
  return {{new0, new1}, new_node};
}

int Mesh::new_edge_between_nodes(std::array<int, 2> node_ids, int parent_elem) {
  if (node_ids[0] == node_ids[1]) panic("new edge between identical node ids");
  const int new_id = int(edges.size());
  const Node& n0 = nodes[node_ids[0]];
  const Node& n1 = nodes[node_ids[1]];
  int ord = elements[elems[parent_elem].element].order_points(n0.coords, n1.coords);
  if (ord == 0)
    throw HRefError(HRefErrorKind::EdgeOnEqualPoints, parent_elem,
                    "Attempt to generate a child-Edge between two identical points over Elem " +
                        std::to_string(parent_elem) + "!");
  Edge e = ord < 0 ? make_edge(new_id, n0, n1, false) : make_edge(new_id, n1, n0, false);
  edges.push_back(e);
  return new_id;
}

std::vector<Elem> Mesh::upgrade_uninit_elems(std::vector<ElemUninit>& ne) {
  std::vector<Elem> out;
  for (const auto& u : ne) out.push_back(u.into_elem());
  for (const auto& e : out) {
    for (int k = 0; k < 4; ++k) edges[e.edges[k]].connect_elem(e);
    for (int k = 0; k < 4; ++k) nodes[e.nodes[k]].connect_elem(e);
  }
  return out;
}

void Mesh::set_edge_activation(int threads) {
  if (threads <= 0) threads = default_threads();
  for (auto& e : edges) e.reset_activation();
  std::vector<int> base;
  for (const auto& e : edges)
    if (e.parent < 0 && !e.boundary) base.push_back(e.id);
  // each base Edge's tree holds Edges of its own, so the trees are independent;
  // the first failing base Edge in id order is the one reported
  parallel_ranges(base.size(), threads, 256, [&](size_t b, size_t e) {
    for (size_t k = b; k < e; ++k)
      if (!rec_set_edge_activation(base[k]))
        panic("Unable to find active Edge pair over Edge " + std::to_string(base[k]) +
              "; Something must be wrong with the mesh!");
  });
}

Mesh Mesh::clone(int threads) const {
  if (threads <= 0) threads = default_threads();
  Mesh m;
  m.elements = elements;
  parallel_copy(m.elems, elems, threads);
  parallel_copy(m.nodes, nodes, threads);
  parallel_copy(m.edges, edges, threads);
  return m;
}

bool Mesh::rec_set_edge_activation(int edge_id) {
  if (!edges[edge_id].set_activation()) return false;
  if (edges[edge_id].has_children()) {
    bool a = rec_set_edge_activation(edges[edge_id].children[0]);
    bool b = rec_set_edge_activation(edges[edge_id].children[1]);
    if (a && b)
      edges[edge_id].reset_activation();
    else if (a != b)
      panic("Children of Edge " + std::to_string(edge_id) +
            " do not have consistent support for Basis Functions; Cannot set activation states!");
  }
  return true;
}

// ---------------------------------------------------------------------------
// p-refinement (mesh.rs:1265-1665)
// ---------------------------------------------------------------------------

void Mesh::global_p_refinement(const PRef& r) {
  std::vector<std::pair<int, PRef>> refs;
  for (const auto& e : elems) refs.push_back({e.id, r.constrained_to(elem_p_refinement_window(e.id))});
  execute_p_refinements(refs);
}

void Mesh::p_refine_elems(const std::vector<int>& ids, const PRef& r) {
  std::vector<int> sorted(ids);  // any repeated id fails the batch, as in h_refine_elems
  std::sort(sorted.begin(), sorted.end());
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
    throw PRefError(PRefErrorKind::DuplicateElemIds, -1,
                    "Duplicate element ids in p-Refinement; Cannot apply p-Refinement!");
  std::vector<std::pair<int, PRef>> refs;
  for (int id : ids) refs.push_back({id, r});
  execute_p_refinements(refs);
}

void Mesh::execute_p_refinements(const std::vector<std::pair<int, PRef>>& refinements) {
  std::map<int, PRef> refmap;
  for (const auto& [id, r] : refinements) {
    if (id < 0 || id >= int(elems.size()))
      throw PRefError(PRefErrorKind::ElemDoesNotExist, id,
                      "Elem " + std::to_string(id) + " does not exist; Cannot apply p-Refinement!");
    auto it = refmap.find(id);
    if (it == refmap.end())
      refmap[id] = r;
    else
      it->second += r;
  }
  for (const auto& [id, r] : refmap)
    if (!r.falls_within(elem_p_refinement_window(id)))
      throw PRefError(PRefErrorKind::RefinementOutOfBounds, id,
                      "Refinement out of bounds for Elem " + std::to_string(id) + "; Cannot apply p-Refinement!");
  for (const auto& [id, r] : refmap) elems[id].poly_orders.refine(r);
}

void Mesh::set_global_expansion_orders(const Orders& o) {
  std::vector<std::pair<int, Orders>> v;
  for (const auto& e : elems) v.push_back({e.id, o});
  set_expansion_orders(v);
}

void Mesh::set_expansion_on_elems(const std::vector<int>& ids, const Orders& o) {
  std::vector<std::pair<int, Orders>> v;
  for (int id : ids) v.push_back({id, o});
  set_expansion_orders(v);
}

void Mesh::set_expansion_orders(const std::vector<std::pair<int, Orders>>& orders) {
  std::map<int, Orders> m;
  for (const auto& [id, o] : orders) {
    if (id < 0 || id >= int(elems.size()))
      throw PRefError(PRefErrorKind::ElemDoesNotExist, id,
                      "Elem " + std::to_string(id) + " does not exist; Cannot apply p-Refinement!");
    if (m.count(id))
      throw PRefError(PRefErrorKind::DuplicateElemIds, -1,
                      "Duplicate element ids in p-Refinement; Cannot apply p-Refinement!");
    m[id] = o;
  }
  for (const auto& [id, o] : m) elems[id].poly_orders.set(o.i, o.j);
}

}  // namespace fem2d
