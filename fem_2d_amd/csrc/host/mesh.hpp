// Host-side mesh with refinement-by-superposition hp-refinement.
//
// This is the input builder for the Galerkin assembly: it restates the
// reference's Mesh/Elem/Edge/Node/Element types and their refinement rules
// (reference src/fem_domain/domain/mesh.rs and mesh/*.rs) in C++, keeping the
// exact id assignment, edge ordering and edge activation the reference uses,
// because DoF numbering (and therefore the sparsity pattern) follows from them.
// Where the reference panics (assert!/panic!/unwrap of an Err), this code
// throws MeshPanic; recoverable errors throw HRefError / PRefError /
// MeshAccessError, mirroring the reference's error enums.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace fem2d {

// mesh.rs:36, mesh.rs:42, space.rs:214
constexpr double MIN_EDGE_LENGTH = 3.0518e-5;
constexpr int MAX_POLYNOMIAL_ORDER = 20;
constexpr double POINT_UNIQUENESS_ACCURACY = 1e-12;

// Raised where the reference would panic.
struct MeshPanic : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// h_refinement.rs:284-296
enum class HRefErrorKind : int {
  MinEdgeLength = 1,
  ElemHasChildren,
  EdgeHasChildren,
  UninitializedElem,
  EdgeOnEqualPoints,
  BisectionIdxExceeded,
  ElemDoesNotExist,
  ElemNotRefineable,
  DuplicateElemIds,
};
struct HRefError : std::runtime_error {
  HRefErrorKind kind;
  long long id;
  HRefError(HRefErrorKind k, long long i, const std::string& msg) : std::runtime_error(msg), kind(k), id(i) {}
};

// p_refinement.rs:305-317
enum class PRefErrorKind : int {
  NegExpansion = 1,
  ExceededMaxExpansion,
  DuplicateElemIds,
  ElemDoesNotExist,
  RefinementOutOfBounds,
  InvalidExpansionOrders,
};
struct PRefError : std::runtime_error {
  PRefErrorKind kind;
  long long id;
  PRefError(PRefErrorKind k, long long i, const std::string& msg) : std::runtime_error(msg), kind(k), id(i) {}
};

// A mesh file that cannot be read (the io::Error of Mesh::from_file).
struct IoError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// mesh.rs:1799-1804
struct MeshAccessError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

enum class ParaDir : uint8_t { U = 0, V = 1 };

// Point comparison key: integer part plus the fraction rounded to 1e-12, with
// the sign kept apart (space.rs:343-381).
struct FloatRep {
  bool sign_positive = true;
  uint64_t bits = 0;
  static FloatRep from(double value);
  int cmp(const FloatRep& o) const;
  bool operator==(const FloatRep& o) const { return sign_positive == o.sign_positive && bits == o.bits; }
  bool operator!=(const FloatRep& o) const { return !(*this == o); }
};

struct Point {
  double x = 0.0, y = 0.0;
  FloatRep x_cmp, y_cmp;
  static Point make(double x, double y);
  static Point between(const Point& a, const Point& b);
  ParaDir orientation_with(const Point& other) const;  // space.rs:250-266
  double dist(const Point& other) const;
  int x_order(const Point& o) const { return x_cmp.cmp(o.x_cmp); }
  int y_order(const Point& o) const { return y_cmp.cmp(o.y_cmp); }
  bool operator==(const Point& o) const { return x_cmp == o.x_cmp && y_cmp == o.y_cmp; }
};

struct Materials {
  double eps_re = 1.0, eps_im = 0.0, mu_re = 1.0, mu_im = 0.0;
};

// A real-space element (element.rs:13-80).
struct Element {
  int id = 0;
  std::array<Point, 4> points;
  Materials materials;
  // [dx/du, dy/dv] of the affine map restricted to a parametric range
  // (element.rs:33-50). The Jacobian is diag(dx/du, dy/dv).
  std::array<double, 2> parametric_mapping(const std::array<std::array<double, 2>, 2>& range) const;
  int order_points(const Point& p0, const Point& p1) const;  // element.rs:56-61
};

using Range2 = std::array<std::array<double, 2>, 2>;

// h_refinement.rs:210-280
enum class HRefLoc : uint8_t { SW = 0, SE = 1, NW = 2, NE = 3, W = 4, E = 5, S = 6, N = 7 };
Range2 sub_range(HRefLoc loc, const Range2& r);

// h_refinement.rs:60-195
struct HRef {
  enum class Kind : uint8_t { T = 0, U = 1, V = 2 };
  Kind kind = Kind::T;
  int ext = -1;  // -1: no extension, 0/1: index of the child refined in the other direction
  static HRef t() { return {Kind::T, -1}; }
  static HRef u(int ext = -1) { return {Kind::U, ext}; }
  static HRef v(int ext = -1) { return {Kind::V, ext}; }
  int num_children() const { return kind == Kind::T ? 4 : 2; }
  HRefLoc loc(int idx) const;
  HRef& operator+=(const HRef& rhs);
  bool operator==(const HRef& o) const { return kind == o.kind && ext == o.ext; }
};

struct HLevels {
  uint8_t u = 0, v = 0;
  HLevels refined(const HRef& r) const;
  // Keys compare lexicographically as [a, b] pairs (h_refinement.rs:31-41).
  uint16_t edge_ranking(ParaDir dir) const {
    return dir == ParaDir::U ? uint16_t((v << 8) | u) : uint16_t((u << 8) | v);
  }
  uint16_t node_ranking() const { return uint16_t((u << 8) | v); }
};

// p-refinement delta (p_refinement.rs:96-260): a signed change per direction.
struct PRef {
  int di = 0, dj = 0;
  static PRef from(int di, int dj) { return {di, dj}; }
  PRef constrained_to(const std::array<std::array<int, 2>, 2>& bounds) const;
  bool falls_within(const std::array<std::array<int, 2>, 2>& bounds) const;
  PRef& operator+=(const PRef& r) {
    di += r.di;
    dj += r.dj;
    return *this;
  }
};

struct PolyOrders {
  uint8_t ni = 1, nj = 1;
  void refine(const PRef& r);  // throws PRefError
  void set(int ni, int nj);    // throws PRefError
};

// Orders::new / Orders::try_new (p_refinement.rs:266-297)
struct Orders {
  int i = 1, j = 1;
  static Orders make(int i, int j);  // throws PRefError::InvalidExpansionOrders
};

struct Elem {
  int id = 0;
  std::array<int, 4> nodes{};
  std::array<int, 4> edges{};
  int element = 0;  // index into Mesh::elements
  HLevels h_levels;
  PolyOrders poly_orders;
  std::vector<int> children;                      // empty: no children
  std::vector<std::pair<int, HRefLoc>> ancestors;  // (ancestor id, loc of the next elem down)

  bool has_children() const { return !children.empty(); }
  int parent_id() const { return ancestors.empty() ? -1 : ancestors.back().first; }
  Range2 relative_parametric_range(int from_ancestor) const;  // elem.rs:170-188
  Range2 parametric_range() const;                            // elem.rs:191-197
};

struct Edge {
  int id = 0;
  std::array<int, 2> nodes{};
  bool boundary = false;
  ParaDir dir = ParaDir::U;
  double length = 0.0;
  bool has_children_ = false;
  std::array<int, 2> children{{-1, -1}};
  int parent = -1;
  std::array<std::map<uint16_t, int>, 2> elems;  // side -> (h-level key -> elem id)
  // per side, the elem of the highest h-level key (the map's last entry), -1
  // for none; kept by connect_elem, the maps' only writer
  std::array<int, 2> top{{-1, -1}};
  std::array<uint16_t, 2> top_key{{0, 0}};
  bool has_active = false;
  std::array<int, 2> active{{-1, -1}};
  int child_node = -1;

  bool has_children() const { return has_children_; }
  void connect_elem(const Elem& elem);
  bool set_activation();  // edge.rs:203-214
  void reset_activation() { has_active = false; }
  int other_active_elem_id(int elem_id) const;  // -1 when no active pair
};

struct Node {
  int id = 0;
  Point coords;
  bool boundary = false;
  std::array<std::map<uint16_t, int>, 4> elems;
  void connect_elem(const Elem& elem);
};

struct ElementSpec {
  std::array<int, 4> node_ids;
  Materials materials;
};

class Mesh {
 public:
  std::vector<Element> elements;
  std::vector<Elem> elems;
  std::vector<Node> nodes;
  std::vector<Edge> edges;

  static Mesh unit();   // mesh.rs:59-88
  static Mesh blank();  // mesh.rs:91-98
  // Construction from node coordinates and element node-id lists; the body of
  // Mesh::from_file after JSON parsing (mesh.rs:142-327).
  static Mesh from_parts(const std::vector<Point>& points, const std::vector<ElementSpec>& elements);
  static Mesh from_json_text(const std::string& text);
  static Mesh from_file(const std::string& path);

  std::array<const Point*, 4> elem_points(int elem_id) const;
  std::array<const Point*, 2> edge_points(int edge_id) const;
  std::vector<int> descendant_elems(int elem_id, bool include_start) const;  // mesh.rs:470-493
  std::vector<int> ancestor_elems(int elem_id, bool include_start) const;    // mesh.rs:520-541
  std::vector<int> descendant_edges(int edge_id, bool include_start) const;  // mesh.rs:577-600
  std::array<int, 2> max_expansion_orders() const;                          // mesh.rs:626-630
  bool elem_is_h_refineable(int elem_id) const;                             // mesh.rs:641-653
  std::array<std::array<int, 2>, 2> elem_p_refinement_window(int elem_id) const;

  // h-refinement (mesh.rs:713-914)
  void global_h_refinement(const HRef& r);
  void h_refine_elems(const std::vector<int>& ids, const HRef& r);
  // The batch of refinements, with the ids the reference's sequential loop
  // gives every new Elem, Node and Edge. Batches of at least
  // kParallelRefinements Elems run data-parallel on up to `threads` threads
  // (0: the machine's cores, at most 16; 1: the sequential loop).
  void execute_h_refinements(const std::vector<std::pair<int, HRef>>& refinements, int threads = 0);
  static constexpr size_t kParallelRefinements = 256;

  // p-refinement and expansion orders (mesh.rs:1265-1665)
  void global_p_refinement(const PRef& r);
  void p_refine_elems(const std::vector<int>& ids, const PRef& r);
  void execute_p_refinements(const std::vector<std::pair<int, PRef>>& refinements);
  void set_global_expansion_orders(const Orders& o);
  void set_expansion_on_elems(const std::vector<int>& ids, const Orders& o);
  void set_expansion_orders(const std::vector<std::pair<int, Orders>>& orders);

  // mesh.rs:1180-1215; the Edge trees under distinct base Edges are set on up
  // to `threads` threads (0: the machine's cores, at most 16)
  void set_edge_activation(int threads = 1);
  // a deep copy made on up to `threads` threads
  Mesh clone(int threads) const;

 private:
  struct ElemUninit;
  struct RefineOps;
  struct AppendOps;
  struct PlannedOps;
  struct RefinePlan;
  // the children of one refinement: their Nodes and Edges set through ops,
  // which creates what is new
  void execute_t_refinement(std::vector<ElemUninit>& ne, int parent, RefineOps& ops);
  void execute_u_refinement(std::vector<ElemUninit>& ne, int parent, RefineOps& ops);
  void execute_v_refinement(std::vector<ElemUninit>& ne, int parent, RefineOps& ops);
  std::pair<std::array<int, 2>, int> h_refine_edge_if_needed(int edge_id);
  int new_edge_between_nodes(std::array<int, 2> node_ids, int parent_elem);
  std::vector<Elem> upgrade_uninit_elems(std::vector<ElemUninit>& ne);
  // the children (uninitialised) of refinement r of parent elem_id, first id first_id
  std::vector<ElemUninit> uninit_children(int elem_id, const HRef& r, int first_id);
  // the batch run data-parallel; false (nothing changed) when a check of the
  // sequential loop would fail, which then runs instead and reports it
  bool execute_h_refinements_planned(const std::vector<std::pair<int, HRef>>& batch, int threads,
                                     std::vector<std::pair<int, HRef>>& extensions);
  bool rec_set_edge_activation(int edge_id);
};

}  // namespace fem2d
