// Host-side Domain construction. See domain.hpp.
#include <functional>

#include "domain.hpp"

namespace fem2d {

void for_each_permutation(const PolyOrders& po, BasisDir dir, const std::function<void(uint8_t, uint8_t)>& f) {
  const int iu = dir == BasisDir::U ? po.ni - 1 : po.ni;  // inclusive upper bounds
  const int ju = dir == BasisDir::V ? po.nj - 1 : po.nj;
  for (int i = 0; i <= iu; ++i)
    for (int j = 0; j <= ju; ++j) f(uint8_t(i), uint8_t(j));
}

BasisSpec BasisSpec::make(int id, uint8_t i, uint8_t j, BasisDir dir, const Elem& elem, ContinuityCondition cc) {
  if (cc != ContinuityCondition::HCurl) throw MeshPanic("not implemented: only H(Curl) BasisSpecs are supported");
  auto edge = [&](int idx) {
    BasisLoc l;
    l.kind = BasisLoc::Kind::Edge;
    l.idx = uint8_t(idx);
    l.id = elem.edges[idx];
    return l;
  };
  auto node = [&](int idx) {
    BasisLoc l;
    l.kind = BasisLoc::Kind::Node;
    l.idx = uint8_t(idx);
    l.id = elem.nodes[idx];
    return l;
  };
  BasisLoc loc;  // ElemBs by default
  if (i >= 2 && j >= 2) {
    // elem type
  } else if (dir == BasisDir::U && j <= 1) {
    loc = edge(j);
  } else if (dir == BasisDir::V && i <= 1) {
    loc = edge(i + 2);
  } else if (dir == BasisDir::W) {
    const bool ilo = i < 2, jlo = j < 2;
    if (ilo && !jlo) loc = edge(i + 2);
    else if (!ilo && jlo) loc = edge(j);
    else if (ilo && jlo) loc = node(i + 2 * j);
  }
  BasisSpec bs;
  bs.id = id;
  bs.i = i;
  bs.j = j;
  bs.dir = dir;
  bs.elem_id = elem.id;
  bs.loc = loc;
  return bs;
}

bool BasisSpec::matches_with_edge(const BasisSpec& o) const {
  if (loc.kind != BasisLoc::Kind::Edge || o.loc.kind != BasisLoc::Kind::Edge)
    throw MeshPanic("Cannot test for edge-type BasisSpec match with non-edge-type BasisSpecs!");
  if (loc.id != o.loc.id)
    throw MeshPanic("Cannot attempt to match Edge-Type BasisSpecs associated with different Edges!");
  const int idx_sum = loc.idx + o.loc.idx;
  if (dir == BasisDir::U && o.dir == BasisDir::U) return i == o.i && j + o.j == 1 && idx_sum == 1;
  if (dir == BasisDir::V && o.dir == BasisDir::V) return j == o.j && i + o.i == 1 && idx_sum == 5;
  if (dir == BasisDir::W && o.dir == BasisDir::W) {
    if (i >= 2 && o.i >= 2) return i == o.i && j + o.j == 1 && idx_sum == 1;
    if (j >= 2 && o.j >= 2) return j == o.j && i + o.i == 1 && idx_sum == 5;
    return false;
  }
  return false;
}

Domain Domain::blank(ContinuityCondition cc) {
  Domain d;
  d.cc = cc;
  return d;
}

static BSAddress push_basis_spec(std::vector<std::vector<BasisSpec>>& lists, BasisSpec bs, int dof_id) {
  const int elem_id = bs.elem_id;
  const int elem_idx = int(lists[elem_id].size());
  if (bs.dof_id >= 0 || bs.elem_idx >= 0) throw MeshPanic("DoF ID was already set on BasisSpec");
  bs.dof_id = dof_id;
  bs.elem_idx = elem_idx;
  lists[elem_id].push_back(bs);
  return {elem_id, elem_idx};
}

Domain Domain::from_mesh(Mesh mesh, ContinuityCondition cc) {
  mesh.set_edge_activation();

  // gen_basis_specs (domain.rs:201-235): per elem, U then V then W permutations,
  // grouped by location (elem / edge / node), each map ordered by id.
  std::map<int, std::vector<BasisSpec>> elem_bs, edge_bs, node_bs;
  int bs_id = 0;
  for (const Elem& elem : mesh.elems) {
    for (BasisDir dir : {BasisDir::U, BasisDir::V, BasisDir::W}) {
      for_each_permutation(elem.poly_orders, dir, [&](uint8_t i, uint8_t j) {
        BasisSpec bs = BasisSpec::make(bs_id++, i, j, dir, elem, cc);
        switch (bs.loc.kind) {
          case BasisLoc::Kind::Elem: elem_bs[elem.id].push_back(bs); break;
          case BasisLoc::Kind::Edge: edge_bs[bs.loc.id].push_back(bs); break;
          case BasisLoc::Kind::Node: node_bs[bs.loc.id].push_back(bs); break;
        }
      });
    }
  }

  Domain d;
  d.cc = cc;
  d.basis_specs.assign(mesh.elems.size(), {});
  int dof_id = 0;

  // elem-type DoFs on leaf Elems (domain.rs:83-96)
  for (auto& [elem_id, list] : elem_bs) {
    if (mesh.elems[elem_id].has_children()) continue;
    for (const BasisSpec& bs : list) {
      if (bs.dir != BasisDir::U && bs.dir != BasisDir::V) continue;
      const int id = dof_id++;
      BSAddress a = push_basis_spec(d.basis_specs, bs, id);
      d.dofs.push_back({id, {a}});
    }
  }

  // edge-type DoFs from matched BasisSpecs on each Edge's active Elem pair
  // (domain.rs:99-149)
  for (auto& [edge_id, list] : edge_bs) {
    const Edge& edge = mesh.edges[edge_id];
    if (!edge.has_active) continue;
    std::vector<BasisSpec> rel;
    for (const BasisSpec& bs : list)
      if ((bs.dir == BasisDir::U || bs.dir == BasisDir::V) &&
          (bs.elem_id == edge.active[0] || bs.elem_id == edge.active[1]))
        rel.push_back(bs);
    std::vector<std::array<size_t, 2>> pairs;
    for (size_t a = 0; a < rel.size(); ++a)
      for (size_t b = a + 1; b < rel.size(); ++b)
        if (rel[a].matches_with_edge(rel[b])) {
          pairs.push_back({a, b});
          break;
        }
    for (const auto& p : pairs) {
      const int id = dof_id++;
      BSAddress a0 = push_basis_spec(d.basis_specs, rel[p[0]], id);
      BSAddress a1 = push_basis_spec(d.basis_specs, rel[p[1]], id);
      d.dofs.push_back({id, {a0, a1}});
    }
  }

  d.mesh = std::move(mesh);
  return d;
}

const std::vector<BasisSpec>& Domain::local_basis_specs(int elem_id) const {
  if (elem_id < 0 || elem_id >= int(mesh.elems.size()))
    throw MeshAccessError("Attempt to access non-existent elem: " + std::to_string(elem_id) + "!");
  return basis_specs[elem_id];
}

std::vector<std::pair<int, const std::vector<BasisSpec>*>> Domain::descendant_basis_specs(int elem_id) const {
  std::vector<std::pair<int, const std::vector<BasisSpec>*>> out;
  for (int id : mesh.descendant_elems(elem_id, false)) out.push_back({id, &basis_specs[id]});
  return out;
}

std::vector<std::pair<int, const std::vector<BasisSpec>*>> Domain::ancestor_basis_specs(int elem_id) const {
  std::vector<std::pair<int, const std::vector<BasisSpec>*>> out;
  for (int id : mesh.ancestor_elems(elem_id, false)) out.push_back({id, &basis_specs[id]});
  return out;
}

}  // namespace fem2d
