// C ABI of the host input builder (Mesh, Domain), error reporting and
// Gauss-Legendre nodes. Declarations and reference citations: include/fem2d.h.
#include <atomic>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../../include/fem2d.h"
#include "capi_internal.hpp"
#include "domain.hpp"
#include "mesh.hpp"
#include "parallel.hpp"

using namespace fem2d;

struct fem2d_mesh {
  Mesh m;
};

struct fem2d_domain {
  Domain d;
  // flattened view storage (fem2d_domain_view)
  std::vector<int32_t> parent;
  std::vector<uint8_t> loc;
  std::vector<double> element_box;
  std::vector<double> materials;
  std::vector<uint8_t> orders;
  std::vector<int64_t> bs_offsets;
  // sized unwritten and filled by parallel loops
  fem2d::uninit_vector<uint8_t> bs_dir, bs_i, bs_j;
  fem2d::uninit_vector<int32_t> bs_dof;
  fem2d::uninit_vector<int32_t> dof_bs;
};

namespace fem2d {
namespace capi {

static thread_local std::string g_err;
static thread_local int g_detail = 0;

void set_error(const std::string& msg, int detail) {
  g_err = msg;
  g_detail = detail;
}

int translate_exception() {
  try {
    throw;
  } catch (const HRefError& e) {
    set_error(e.what(), int(e.kind));
    return FEM2D_ERR_HREF;
  } catch (const PRefError& e) {
    set_error(e.what(), int(e.kind));
    return FEM2D_ERR_PREF;
  } catch (const MeshAccessError& e) {
    set_error(e.what(), 0);
    return FEM2D_ERR_MESH_ACCESS;
  } catch (const IoError& e) {
    set_error(e.what(), 0);
    return FEM2D_ERR_IO;
  } catch (const MeshPanic& e) {
    set_error(e.what(), 0);
    return FEM2D_ERR_PANIC;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed", 0);
    return FEM2D_ERR_OOM;
  } catch (const std::exception& e) {
    // JSON syntax errors land here: the reference panics on them (mesh.rs:146)
    set_error(e.what(), 0);
    return FEM2D_ERR_PANIC;
  } catch (...) {
    set_error("unknown error", 0);
    return FEM2D_ERR_PANIC;
  }
}

}  // namespace capi
}  // namespace fem2d

using fem2d::capi::set_error;

#define FEM2D_GUARD(...)                          \
  try {                                           \
    __VA_ARGS__;                                  \
    return FEM2D_OK;                              \
  } catch (...) {                                 \
    return fem2d::capi::translate_exception();    \
  }

#define FEM2D_REQUIRE(cond, msg)                  \
  do {                                            \
    if (!(cond)) {                                \
      set_error(msg, 0);                          \
      return FEM2D_ERR_INVALID_ARG;               \
    }                                             \
  } while (0)

static int check_elem(const Mesh& m, int32_t id) {
  if (id < 0 || id >= int32_t(m.elems.size())) {
    set_error("Attempt to access non-existent elem: " + std::to_string(id) + "!", 0);
    return FEM2D_ERR_MESH_ACCESS;
  }
  return FEM2D_OK;
}

static HRef make_href(int32_t kind, int32_t ext) {
  if (ext != -1 && ext != 0 && ext != 1)
    throw HRefError(HRefErrorKind::BisectionIdxExceeded, -1, "Extended refinement index must be 0 or 1!");
  switch (kind) {
    case FEM2D_HREF_T: return HRef::t();
    case FEM2D_HREF_U: return HRef::u(ext);
    case FEM2D_HREF_V: return HRef::v(ext);
    default: throw MeshPanic("unknown h-refinement kind");
  }
}

// The Mesh as the flat arrays of the device Domain builder (every index checked
// here, so the builder's kernels need not). The Edge side arrays hold the Elem
// Edge::set_activation would pick on each side (edge.rs:203-214: the last entry
// of the side's h-level map, Edge::top).
static void flatten_mesh(const Mesh& mesh, int threads, capi::MeshFlat& f) {
  const size_t ne = mesh.elems.size(), ned = mesh.edges.size();
  if (ne > size_t(INT32_MAX) || ned > size_t(INT32_MAX)) throw MeshPanic("too many Elems or Edges for 32-bit ids");
  f.n_elems = int32_t(ne);
  f.n_edges = int32_t(ned);
  f.parent.resize(ne);
  f.loc.resize(ne);
  f.box.resize(4 * ne);
  f.mat.resize(2 * ne);
  f.edges.resize(4 * ne);
  f.meta.resize(ne);
  f.side.resize(2 * ned);
  f.edge_parent.resize(ned);
  f.child.resize(2 * ned);
  f.open_base.resize(ned);
  std::atomic<int64_t> uv{0}, all{0};
  std::atomic<int> max_i{0}, max_j{0};
  // Elems [0, ne) and Edges [ne, ne + ned) in one pass of contiguous ranges
  auto elem_range = [&](size_t b, size_t e) {
    int64_t uv_part = 0, all_part = 0;
    int mi = 0, mj = 0;
    for (size_t x = b; x < e; ++x) {
      const Elem& el = mesh.elems[x];
      if (el.element < 0 || size_t(el.element) >= mesh.elements.size())
        throw MeshPanic("Elem " + std::to_string(x) + " refers to a non-existent Element");
      const Element& elt = mesh.elements[size_t(el.element)];
      const int p = el.parent_id();
      if (p < -1 || p >= int(ne)) throw MeshPanic("Elem " + std::to_string(x) + " has a non-existent parent");
      f.parent[x] = p;
      f.loc[x] = el.ancestors.empty() ? 0 : uint8_t(el.ancestors.back().second);
      f.box[4 * x + 0] = elt.points[0].x;
      f.box[4 * x + 1] = elt.points[0].y;
      f.box[4 * x + 2] = elt.points[3].x;
      f.box[4 * x + 3] = elt.points[3].y;
      f.mat[2 * x + 0] = elt.materials.eps_re;
      f.mat[2 * x + 1] = elt.materials.mu_re;
      for (int k = 0; k < 4; ++k) {
        const int g = el.edges[size_t(k)];
        if (g < 0 || g >= int(ned)) throw MeshPanic("Elem " + std::to_string(x) + " refers to a non-existent Edge");
        for (int q = 0; q < k; ++q)
          if (el.edges[size_t(q)] == g)
            throw MeshPanic("Elem " + std::to_string(x) + " lists Edge " + std::to_string(g) + " twice");
        f.edges[4 * x + size_t(k)] = g;
      }
      const int ni = el.poly_orders.ni, nj = el.poly_orders.nj;
      if (ni < 1 || nj < 1) throw MeshPanic("Elem " + std::to_string(x) + " has an expansion order below 1");
      f.meta[x] = uint32_t(ni) | uint32_t(nj) << 8 | uint32_t(el.has_children() ? 0 : 1) << 16;
      uv_part += int64_t(ni) * (nj + 1) + int64_t(ni + 1) * nj;
      all_part += int64_t(ni) * (nj + 1) + int64_t(ni + 1) * nj + int64_t(ni + 1) * (nj + 1);
      mi = std::max(mi, ni);
      mj = std::max(mj, nj);
    }
    uv += uv_part;
    all += all_part;
    for (int cur = max_i.load(); mi > cur && !max_i.compare_exchange_weak(cur, mi);) {
    }
    for (int cur = max_j.load(); mj > cur && !max_j.compare_exchange_weak(cur, mj);) {
    }
  };
  auto edge_range = [&](size_t b, size_t e) {
    for (size_t g = b; g < e; ++g) {
      const Edge& ed = mesh.edges[g];
      for (int s = 0; s < 2; ++s) {
        const int x = ed.top[size_t(s)];
        if (x < -1 || x >= int(ne)) throw MeshPanic("Edge " + std::to_string(g) + " refers to a non-existent Elem");
        f.side[2 * g + size_t(s)] = x;
      }
      if (ed.parent < -1 || ed.parent >= int(ned))
        throw MeshPanic("Edge " + std::to_string(g) + " has a non-existent parent");
      f.edge_parent[g] = ed.parent;
      for (int s = 0; s < 2; ++s) {
        const int c = ed.has_children() ? ed.children[size_t(s)] : -1;
        if (ed.has_children() && (c < 0 || c >= int(ned)))
          throw MeshPanic("Edge " + std::to_string(g) + " has a non-existent child");
        f.child[2 * g + size_t(s)] = c;
      }
      f.open_base[g] = uint8_t(ed.parent < 0 && !ed.boundary);
    }
  };
  parallel_ranges(ne + ned, threads, 4096, [&](size_t b, size_t e) {
    if (b < ne) elem_range(b, std::min(e, ne));
    if (e > ne) edge_range(std::max(b, ne) - ne, e - ne);
  });
  f.uv_specs = uv.load();
  f.all_specs = all.load();
  f.max_i = max_i.load();
  f.max_j = max_j.load();
}

static int write_ids(const std::vector<int>& v, int32_t* out, int64_t cap, int64_t* n) {
  if (n) *n = int64_t(v.size());
  if (out)
    for (int64_t k = 0; k < cap && k < int64_t(v.size()); ++k) out[k] = v[k];
  return FEM2D_OK;
}

extern "C" {

const char* fem2d_last_error(void) { return fem2d::capi::g_err.c_str(); }
int fem2d_last_error_detail(void) { return fem2d::capi::g_detail; }

// ---------------------------------------------------------------------------
// Mesh
// ---------------------------------------------------------------------------

int fem2d_mesh_from_file(const char* path, fem2d_mesh** out) {
  FEM2D_REQUIRE(path && out, "null argument");
  FEM2D_GUARD(*out = new fem2d_mesh{Mesh::from_file(path)})
}

int fem2d_mesh_from_json(const char* text, fem2d_mesh** out) {
  FEM2D_REQUIRE(text && out, "null argument");
  FEM2D_GUARD(*out = new fem2d_mesh{Mesh::from_json_text(text)})
}

int fem2d_mesh_from_parts(int64_t n_nodes, const double* xy, int64_t n_elements, const int32_t* node_ids,
                          const double* materials, fem2d_mesh** out) {
  FEM2D_REQUIRE(out && n_nodes >= 0 && n_elements >= 0, "invalid argument");
  FEM2D_REQUIRE((n_nodes == 0 || xy) && (n_elements == 0 || (node_ids && materials)), "null array");
  FEM2D_GUARD({
    std::vector<Point> pts;
    pts.reserve(size_t(n_nodes));
    for (int64_t k = 0; k < n_nodes; ++k) pts.push_back(Point::make(xy[2 * k], xy[2 * k + 1]));
    std::vector<ElementSpec> specs(static_cast<size_t>(n_elements));
    for (int64_t e = 0; e < n_elements; ++e) {
      for (int k = 0; k < 4; ++k) specs[e].node_ids[k] = node_ids[4 * e + k];
      specs[e].materials = {materials[4 * e], materials[4 * e + 1], materials[4 * e + 2], materials[4 * e + 3]};
    }
    *out = new fem2d_mesh{Mesh::from_parts(pts, specs)};
  })
}

int fem2d_mesh_unit(fem2d_mesh** out) {
  FEM2D_REQUIRE(out, "null argument");
  FEM2D_GUARD(*out = new fem2d_mesh{Mesh::unit()})
}

int fem2d_mesh_blank(fem2d_mesh** out) {
  FEM2D_REQUIRE(out, "null argument");
  FEM2D_GUARD(*out = new fem2d_mesh{Mesh::blank()})
}

int fem2d_mesh_clone(const fem2d_mesh* m, fem2d_mesh** out) {
  FEM2D_REQUIRE(m && out, "null argument");
  FEM2D_GUARD(*out = new fem2d_mesh{m->m})
}

void fem2d_mesh_free(fem2d_mesh* m) { delete m; }

int fem2d_mesh_counts(const fem2d_mesh* m, int64_t* n_elems, int64_t* n_nodes, int64_t* n_edges,
                      int64_t* n_elements) {
  FEM2D_REQUIRE(m, "null mesh");
  if (n_elems) *n_elems = int64_t(m->m.elems.size());
  if (n_nodes) *n_nodes = int64_t(m->m.nodes.size());
  if (n_edges) *n_edges = int64_t(m->m.edges.size());
  if (n_elements) *n_elements = int64_t(m->m.elements.size());
  return FEM2D_OK;
}

int fem2d_mesh_elem_info(const fem2d_mesh* m, int32_t elem_id, fem2d_elem_info* out) {
  FEM2D_REQUIRE(m && out, "null argument");
  if (int rc = check_elem(m->m, elem_id)) return rc;
  const Elem& e = m->m.elems[elem_id];
  out->id = e.id;
  for (int k = 0; k < 4; ++k) {
    out->nodes[k] = e.nodes[k];
    out->edges[k] = e.edges[k];
    out->children[k] = -1;
  }
  out->element = e.element;
  out->parent = e.parent_id();
  out->loc = e.ancestors.empty() ? -1 : int32_t(e.ancestors.back().second);
  out->h_levels[0] = e.h_levels.u;
  out->h_levels[1] = e.h_levels.v;
  out->orders[0] = e.poly_orders.ni;
  out->orders[1] = e.poly_orders.nj;
  out->n_children = int32_t(e.children.size());
  for (size_t k = 0; k < e.children.size() && k < 4; ++k) out->children[k] = e.children[k];
  out->depth = int32_t(e.ancestors.size());
  return FEM2D_OK;
}

int fem2d_mesh_edge_info(const fem2d_mesh* m, int32_t edge_id, fem2d_edge_info* out) {
  FEM2D_REQUIRE(m && out, "null argument");
  if (edge_id < 0 || edge_id >= int32_t(m->m.edges.size())) {
    set_error("Attempt to access non-existent edge: " + std::to_string(edge_id) + "!", 0);
    return FEM2D_ERR_MESH_ACCESS;
  }
  const Edge& e = m->m.edges[edge_id];
  out->id = e.id;
  out->nodes[0] = e.nodes[0];
  out->nodes[1] = e.nodes[1];
  out->boundary = e.boundary;
  out->dir = int32_t(e.dir);
  out->length = e.length;
  out->parent = e.parent;
  out->children[0] = e.has_children() ? e.children[0] : -1;
  out->children[1] = e.has_children() ? e.children[1] : -1;
  out->child_node = e.child_node;
  out->active[0] = e.has_active ? e.active[0] : -1;
  out->active[1] = e.has_active ? e.active[1] : -1;
  return FEM2D_OK;
}

int fem2d_mesh_node_info(const fem2d_mesh* m, int32_t node_id, double* x, double* y, int32_t* boundary) {
  FEM2D_REQUIRE(m, "null mesh");
  if (node_id < 0 || node_id >= int32_t(m->m.nodes.size())) {
    set_error("Attempt to access non-existent node: " + std::to_string(node_id) + "!", 0);
    return FEM2D_ERR_MESH_ACCESS;
  }
  const Node& n = m->m.nodes[node_id];
  if (x) *x = n.coords.x;
  if (y) *y = n.coords.y;
  if (boundary) *boundary = n.boundary;
  return FEM2D_OK;
}

int fem2d_mesh_element_info(const fem2d_mesh* m, int32_t element_id, double* points8, double* materials4) {
  FEM2D_REQUIRE(m, "null mesh");
  if (element_id < 0 || element_id >= int32_t(m->m.elements.size())) {
    set_error("Attempt to access non-existent element: " + std::to_string(element_id) + "!", 0);
    return FEM2D_ERR_MESH_ACCESS;
  }
  const Element& el = m->m.elements[element_id];
  if (points8)
    for (int k = 0; k < 4; ++k) {
      points8[2 * k] = el.points[k].x;
      points8[2 * k + 1] = el.points[k].y;
    }
  if (materials4) {
    materials4[0] = el.materials.eps_re;
    materials4[1] = el.materials.eps_im;
    materials4[2] = el.materials.mu_re;
    materials4[3] = el.materials.mu_im;
  }
  return FEM2D_OK;
}

int fem2d_mesh_descendant_elems(const fem2d_mesh* m, int32_t elem_id, int include_start, int32_t* out, int64_t cap,
                                int64_t* n) {
  FEM2D_REQUIRE(m, "null mesh");
  FEM2D_GUARD(write_ids(m->m.descendant_elems(elem_id, include_start != 0), out, cap, n))
}

int fem2d_mesh_ancestor_elems(const fem2d_mesh* m, int32_t elem_id, int include_start, int32_t* out, int64_t cap,
                              int64_t* n) {
  FEM2D_REQUIRE(m, "null mesh");
  FEM2D_GUARD(write_ids(m->m.ancestor_elems(elem_id, include_start != 0), out, cap, n))
}

int fem2d_mesh_descendant_edges(const fem2d_mesh* m, int32_t edge_id, int include_start, int32_t* out, int64_t cap,
                                int64_t* n) {
  FEM2D_REQUIRE(m, "null mesh");
  FEM2D_GUARD(write_ids(m->m.descendant_edges(edge_id, include_start != 0), out, cap, n))
}

int fem2d_mesh_max_expansion_orders(const fem2d_mesh* m, int32_t* out2) {
  FEM2D_REQUIRE(m && out2, "null argument");
  auto o = m->m.max_expansion_orders();
  out2[0] = o[0];
  out2[1] = o[1];
  return FEM2D_OK;
}

int fem2d_mesh_elem_is_h_refineable(const fem2d_mesh* m, int32_t elem_id, int32_t* out) {
  FEM2D_REQUIRE(m && out, "null argument");
  FEM2D_GUARD(*out = m->m.elem_is_h_refineable(elem_id) ? 1 : 0)
}

int fem2d_mesh_elem_p_refinement_window(const fem2d_mesh* m, int32_t elem_id, int32_t* out4) {
  FEM2D_REQUIRE(m && out4, "null argument");
  FEM2D_GUARD({
    auto w = m->m.elem_p_refinement_window(elem_id);
    out4[0] = w[0][0];
    out4[1] = w[0][1];
    out4[2] = w[1][0];
    out4[3] = w[1][1];
  })
}

int fem2d_mesh_elem_parametric_range(const fem2d_mesh* m, int32_t elem_id, int32_t from_ancestor, double* out4) {
  FEM2D_REQUIRE(m && out4, "null argument");
  if (int rc = check_elem(m->m, elem_id)) return rc;
  FEM2D_GUARD({
    const Elem& e = m->m.elems[elem_id];
    Range2 r = from_ancestor < 0 ? e.parametric_range() : e.relative_parametric_range(from_ancestor);
    out4[0] = r[0][0];
    out4[1] = r[0][1];
    out4[2] = r[1][0];
    out4[3] = r[1][1];
  })
}

int fem2d_mesh_global_h_refinement(fem2d_mesh* m, int32_t kind, int32_t ext) {
  FEM2D_REQUIRE(m, "null mesh");
  FEM2D_GUARD(m->m.global_h_refinement(make_href(kind, ext)))
}

int fem2d_mesh_h_refine_elems(fem2d_mesh* m, const int32_t* ids, int64_t n, int32_t kind, int32_t ext) {
  FEM2D_REQUIRE(m && (n == 0 || ids) && n >= 0, "invalid argument");
  FEM2D_GUARD(m->m.h_refine_elems(std::vector<int>(ids, ids + n), make_href(kind, ext)))
}

int fem2d_mesh_execute_h_refinements_threads(fem2d_mesh* m, const int32_t* ids, const int32_t* kinds,
                                             const int32_t* exts, int64_t n, int32_t n_threads) {
  FEM2D_REQUIRE(m && n >= 0 && (n == 0 || (ids && kinds && exts)), "invalid argument");
  FEM2D_REQUIRE(n_threads >= 0, "negative thread count");
  FEM2D_GUARD({
    std::vector<std::pair<int, HRef>> refs;
    for (int64_t k = 0; k < n; ++k) refs.push_back({ids[k], make_href(kinds[k], exts[k])});
    m->m.execute_h_refinements(refs, n_threads);
  })
}

int fem2d_mesh_execute_h_refinements(fem2d_mesh* m, const int32_t* ids, const int32_t* kinds, const int32_t* exts,
                                     int64_t n) {
  return fem2d_mesh_execute_h_refinements_threads(m, ids, kinds, exts, n, 0);
}

int fem2d_mesh_global_p_refinement(fem2d_mesh* m, int32_t di, int32_t dj) {
  FEM2D_REQUIRE(m, "null mesh");
  FEM2D_GUARD(m->m.global_p_refinement(PRef::from(di, dj)))
}

int fem2d_mesh_p_refine_elems(fem2d_mesh* m, const int32_t* ids, int64_t n, int32_t di, int32_t dj) {
  FEM2D_REQUIRE(m && n >= 0 && (n == 0 || ids), "invalid argument");
  FEM2D_GUARD(m->m.p_refine_elems(std::vector<int>(ids, ids + n), PRef::from(di, dj)))
}

int fem2d_mesh_execute_p_refinements(fem2d_mesh* m, const int32_t* ids, const int32_t* dis, const int32_t* djs,
                                     int64_t n) {
  FEM2D_REQUIRE(m && n >= 0 && (n == 0 || (ids && dis && djs)), "invalid argument");
  FEM2D_GUARD({
    std::vector<std::pair<int, PRef>> refs;
    for (int64_t k = 0; k < n; ++k) refs.push_back({ids[k], PRef::from(dis[k], djs[k])});
    m->m.execute_p_refinements(refs);
  })
}

int fem2d_mesh_set_global_expansion_orders(fem2d_mesh* m, int32_t i, int32_t j) {
  FEM2D_REQUIRE(m, "null mesh");
  FEM2D_GUARD(m->m.set_global_expansion_orders(Orders::make(i, j)))
}

int fem2d_mesh_set_expansion_orders(fem2d_mesh* m, const int32_t* ids, const int32_t* is, const int32_t* js,
                                    int64_t n) {
  FEM2D_REQUIRE(m && n >= 0 && (n == 0 || (ids && is && js)), "invalid argument");
  FEM2D_GUARD({
    std::vector<std::pair<int, Orders>> v;
    for (int64_t k = 0; k < n; ++k) v.push_back({ids[k], Orders::make(is[k], js[k])});
    m->m.set_expansion_orders(v);
  })
}

// ---------------------------------------------------------------------------
// Domain
// ---------------------------------------------------------------------------

static void flatten(fem2d_domain* dom, int threads) {
  const Domain& d = dom->d;
  const Mesh& m = d.mesh;
  const size_t ne = m.elems.size();
  dom->parent.resize(ne);
  dom->loc.resize(ne);
  dom->element_box.resize(4 * ne);
  dom->materials.resize(2 * ne);
  dom->orders.resize(2 * ne);
  dom->bs_offsets.assign(ne + 1, 0);
  for (size_t e = 0; e < ne; ++e) dom->bs_offsets[e + 1] = dom->bs_offsets[e] + int64_t(d.basis_specs[e].size());
  const size_t nbs = size_t(dom->bs_offsets[ne]);
  dom->bs_dir.resize(nbs);
  dom->bs_i.resize(nbs);
  dom->bs_j.resize(nbs);
  dom->bs_dof.resize(nbs);
  parallel_ranges(ne, threads, 512, [&](size_t b, size_t en) {
    for (size_t e = b; e < en; ++e) {
      const Elem& el = m.elems[e];
      const Element& elt = m.elements[el.element];
      dom->parent[e] = el.parent_id();
      dom->loc[e] = el.ancestors.empty() ? 0 : uint8_t(el.ancestors.back().second);
      dom->element_box[4 * e + 0] = elt.points[0].x;
      dom->element_box[4 * e + 1] = elt.points[0].y;
      dom->element_box[4 * e + 2] = elt.points[3].x;
      dom->element_box[4 * e + 3] = elt.points[3].y;
      dom->materials[2 * e + 0] = elt.materials.eps_re;
      dom->materials[2 * e + 1] = elt.materials.mu_re;
      dom->orders[2 * e + 0] = el.poly_orders.ni;
      dom->orders[2 * e + 1] = el.poly_orders.nj;
      size_t t = size_t(dom->bs_offsets[e]);
      for (const BasisSpec& bs : d.basis_specs[e]) {
        dom->bs_dir[t] = uint8_t(bs.dir);
        dom->bs_i[t] = bs.i;
        dom->bs_j[t] = bs.j;
        dom->bs_dof[t] = bs.dof_id;
        ++t;
      }
    }
  });
  // each DoF's basis-spec addresses as flat spec indices
  dom->dof_bs.resize(2 * d.dofs.size());
  parallel_ranges(d.dofs.size(), threads, 8192, [&](size_t b, size_t e) {
    for (size_t f = b; f < e; ++f) {
      const DoF& dof = d.dofs[f];
      for (size_t k = 0; k < 2; ++k) {
        const BSAddress* a = k < dof.basis_specs.size() ? &dof.basis_specs[k] : nullptr;
        dom->dof_bs[2 * f + k] = a ? int32_t(dom->bs_offsets[size_t(a->elem_id)] + a->elem_idx) : -1;
      }
    }
  });
}

int fem2d_domain_from_mesh_threads(const fem2d_mesh* m, int32_t cc, int32_t n_threads, fem2d_domain** out) {
  FEM2D_REQUIRE(m && out, "null argument");
  FEM2D_REQUIRE(cc >= 0 && cc <= 2, "unknown continuity condition");
  FEM2D_REQUIRE(n_threads >= 0, "negative thread count");
  FEM2D_GUARD({
    const int threads = n_threads > 0 ? int(n_threads) : default_threads();
    auto* dom = new fem2d_domain();
    dom->d = Domain::from_mesh(m->m.clone(threads), ContinuityCondition(cc), threads);
    flatten(dom, threads);
    *out = dom;
  })
}

int fem2d_domain_from_mesh(const fem2d_mesh* m, int32_t cc, fem2d_domain** out) {
  return fem2d_domain_from_mesh_threads(m, cc, 0, out);
}

int fem2d_domain_from_mesh_serial(const fem2d_mesh* m, int32_t cc, fem2d_domain** out) {
  FEM2D_REQUIRE(m && out, "null argument");
  FEM2D_REQUIRE(cc >= 0 && cc <= 2, "unknown continuity condition");
  FEM2D_GUARD({
    auto* dom = new fem2d_domain();
    dom->d = Domain::from_mesh_serial(m->m, ContinuityCondition(cc));
    flatten(dom, 1);
    *out = dom;
  })
}

int fem2d_domain_blank(int32_t cc, fem2d_domain** out) {
  FEM2D_REQUIRE(out, "null argument");
  FEM2D_REQUIRE(cc >= 0 && cc <= 2, "unknown continuity condition");
  FEM2D_GUARD({
    auto* dom = new fem2d_domain();
    dom->d = Domain::blank(ContinuityCondition(cc));
    flatten(dom, 1);
    *out = dom;
  })
}

void fem2d_domain_free(fem2d_domain* d) { delete d; }

int fem2d_domain_num_dofs(const fem2d_domain* d, int64_t* n) {
  FEM2D_REQUIRE(d && n, "null argument");
  *n = int64_t(d->d.dofs.size());
  return FEM2D_OK;
}

int fem2d_domain_mesh(const fem2d_domain* d, const fem2d_mesh** out) {
  FEM2D_REQUIRE(d && out, "null argument");
  // fem2d_mesh holds exactly one Mesh, so the Domain's Mesh can be viewed as one
  *out = reinterpret_cast<const fem2d_mesh*>(&d->d.mesh);
  return FEM2D_OK;
}

int fem2d_domain_get_view(const fem2d_domain* d, fem2d_domain_view* out) {
  FEM2D_REQUIRE(d && out, "null argument");
  out->cc = int32_t(d->d.cc);
  out->n_elems = int32_t(d->d.mesh.elems.size());
  out->n_dofs = int64_t(d->d.dofs.size());
  out->parent = d->parent.data();
  out->loc = d->loc.data();
  out->element_box = d->element_box.data();
  out->materials = d->materials.data();
  out->orders = d->orders.data();
  out->bs_offsets = d->bs_offsets.data();
  out->bs_dir = d->bs_dir.data();
  out->bs_i = d->bs_i.data();
  out->bs_j = d->bs_j.data();
  out->bs_dof = d->bs_dof.data();
  out->dof_bs = d->dof_bs.data();
  return FEM2D_OK;
}

int fem2d_domain_dof_info(const fem2d_domain* d, int64_t dof, int32_t* n_bs, int32_t* elem_ids4, int32_t* elem_idx4) {
  FEM2D_REQUIRE(d, "null domain");
  FEM2D_REQUIRE(dof >= 0 && dof < int64_t(d->d.dofs.size()), "dof out of range");
  const DoF& f = d->d.dofs[size_t(dof)];
  if (n_bs) *n_bs = int32_t(f.basis_specs.size());
  for (size_t k = 0; k < f.basis_specs.size() && k < 4; ++k) {
    if (elem_ids4) elem_ids4[k] = f.basis_specs[k].elem_id;
    if (elem_idx4) elem_idx4[k] = f.basis_specs[k].elem_idx;
  }
  return FEM2D_OK;
}

int fem2d_device_domain_from_mesh(fem2d_ctx* ctx, const fem2d_mesh* m, int32_t cc, fem2d_device_domain** out) {
  FEM2D_REQUIRE(ctx && m && out, "null argument");
  *out = nullptr;
  FEM2D_REQUIRE(cc >= 0 && cc <= 2, "unknown continuity condition");
  FEM2D_GUARD({
    const Mesh& mesh = m->m;
    // Domain::from_mesh sets the edge activation before it generates the
    // first BasisSpec, so an activation panic comes before this one
    if (cc != FEM2D_CC_HCURL && !mesh.elems.empty()) {
      mesh.clone(1).set_edge_activation(1);
      throw MeshPanic("not implemented: only H(Curl) BasisSpecs are supported");
    }
    capi::MeshFlat f;
    flatten_mesh(mesh, default_threads(), f);
    if (f.all_specs > INT32_MAX) {  // Domain::from_mesh's id limit, after its activation
      mesh.clone(1).set_edge_activation(1);
      throw MeshPanic("too many BasisSpecs for 32-bit ids");
    }
    int rc = capi::device_domain_from_flat(ctx, f, cc, out);
    if (rc == capi::kActivationFailed) {
      // the device found an Edge without an active pair where the reference
      // panics; the sequential activation names the Edge it meets first
      mesh.clone(1).set_edge_activation(1);
      set_error("edge activation failed on the device, but not in the sequential activation", 0);
      rc = FEM2D_ERR_PANIC;
    }
    if (rc) return rc;
  })
}

// ---------------------------------------------------------------------------
// Gauss-Legendre nodes by Newton iteration on the three-term recurrence
// ---------------------------------------------------------------------------

int fem2d_gauss_legendre(int32_t n, double* points, double* weights) {
  FEM2D_REQUIRE(n >= 1 && points && weights, "invalid argument");
  fem2d::gauss_legendre_nodes(n, points, weights);
  return FEM2D_OK;
}

}  // extern "C"

namespace fem2d {

void gauss_legendre_nodes(int n, double* x, double* w) {
  for (int i = 0; i < n; ++i) {
    // i-th root in descending order, Tricomi's initial estimate
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p0 = 1.0;
      // P_n'(z) from P_n and P_{n-1}
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) <= 1e-16 * (1.0 + std::fabs(z))) break;
    }
    {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p0 = 1.0;
      dp = n * (z * p1 - p0) / (z * z - 1.0);
    }
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

}  // namespace fem2d
