// Host-side Domain construction. See domain.hpp.
#include <algorithm>
#include <array>
#include <functional>

#include "domain.hpp"
#include "parallel.hpp"

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

Domain Domain::from_mesh_serial(Mesh mesh, ContinuityCondition cc) {
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

namespace {

// BasisSpecs generated per Elem: U, V and W permutations
int64_t specs_per_elem(const PolyOrders& po) {
  return int64_t(po.ni) * (po.nj + 1) + int64_t(po.ni + 1) * po.nj + int64_t(po.ni + 1) * (po.nj + 1);
}
// of which U and V (the only directions of H(curl) DoFs)
int64_t uv_specs_per_elem(const PolyOrders& po) { return int64_t(po.ni) * (po.nj + 1) + int64_t(po.ni + 1) * po.nj; }

}  // namespace

// The sequential algorithm visits the Elems in id order to generate the specs,
// numbers the elem-type DoFs of the leaves in Elem order and then the matched
// edge-type DoFs in Edge order, appending each DoF's specs to the lists of
// their Elems as it goes. Every one of these orders is a prefix sum over
// independent per-Elem or per-Edge counts, so the same Domain comes out of:
//  1. per Elem: its U and V specs in generation order into one flat array (W
//     specs never become H(curl) DoFs; the ids still count them);
//  2. per leaf Elem: the number of its elem-type specs -> elem DoF ids;
//  3. per active Edge: the U/V specs on it of its active pair (Elem-id order,
//     generation order) and their first-match pairs -> edge DoF ids after a
//     prefix sum over Edges;
//  4. per Elem: its list = its elem-type DoF specs, then for each of its Edges
//     in id order its spec of each of that Edge's pairs.
Domain Domain::from_mesh(Mesh mesh, ContinuityCondition cc, int threads) {
  if (threads <= 0) threads = default_threads();
  mesh.set_edge_activation(threads);
  if (cc != ContinuityCondition::HCurl && !mesh.elems.empty())
    throw MeshPanic("not implemented: only H(Curl) BasisSpecs are supported");
  const size_t ne = mesh.elems.size(), ned = mesh.edges.size();

  // 1. U/V specs of every Elem
  std::vector<int64_t> spec_base(ne + 1, 0), uv_base(ne + 1, 0);
  for (size_t x = 0; x < ne; ++x) {
    spec_base[x + 1] = spec_base[x] + specs_per_elem(mesh.elems[x].poly_orders);
    uv_base[x + 1] = uv_base[x] + uv_specs_per_elem(mesh.elems[x].poly_orders);
  }
  if (spec_base[ne] > INT32_MAX) throw MeshPanic("too many BasisSpecs for 32-bit ids");
  // written once by the loop below, not by the allocating thread
  union SpecSlot {
    SpecSlot() {}
    BasisSpec s;
  };
  uninit_vector<SpecSlot> slots(static_cast<size_t>(uv_base[ne]));
  BasisSpec* const uv = &slots.data()->s;
  static_assert(sizeof(SpecSlot) == sizeof(BasisSpec), "flat spec array");
  std::vector<int> n_elem_type(ne, 0);  // elem-type specs of each leaf
  parallel_ranges(ne, threads, 256, [&](size_t b, size_t e) {
    for (size_t x = b; x < e; ++x) {
      const Elem& elem = mesh.elems[x];
      int id = int(spec_base[x]);
      BasisSpec* out = uv + uv_base[x];
      int ce = 0;
      for (BasisDir dir : {BasisDir::U, BasisDir::V})
        for_each_permutation(elem.poly_orders, dir, [&](uint8_t i, uint8_t j) {
          ::new (static_cast<void*>(out)) BasisSpec(BasisSpec::make(id++, i, j, dir, elem, cc));
          ce += out->loc.kind == BasisLoc::Kind::Elem;
          ++out;
        });
      n_elem_type[x] = elem.has_children() ? 0 : ce;
    }
  });

  // 2. elem-type DoF ids of the leaves
  std::vector<int64_t> elem_base(ne + 1, 0);
  for (size_t x = 0; x < ne; ++x) elem_base[x + 1] = elem_base[x] + n_elem_type[x];

  // 3. first-match pairs on each active Edge, as indices into `uv`; Edges are
  // taken in fixed blocks, each block's pairs in a vector of its own
  constexpr size_t kEdgeBlock = 2048;
  const size_t n_blocks = (ned + kEdgeBlock - 1) / kEdgeBlock;
  std::vector<std::vector<std::array<int64_t, 2>>> block_pairs(n_blocks);
  std::vector<int> pair_count(ned, 0), pair_start(ned, 0);  // start within the Edge's block
  parallel_ranges(n_blocks, threads, 1, [&](size_t bb, size_t be) {
    std::vector<int64_t> rel;
    for (size_t blk = bb; blk < be; ++blk) {
      auto& pairs = block_pairs[blk];
      for (size_t id = blk * kEdgeBlock; id < std::min(ned, (blk + 1) * kEdgeBlock); ++id) {
        pair_start[id] = int(pairs.size());
        const Edge& edge = mesh.edges[id];
        if (!edge.has_active) continue;
        int act[2] = {edge.active[0], edge.active[1]};
        if (act[1] < act[0]) std::swap(act[0], act[1]);
        rel.clear();
        for (int s = 0; s < 2; ++s) {
          const int a = act[s];
          if (a < 0 || a >= int(ne) || (s == 1 && a == act[0])) continue;
          for (int64_t t = uv_base[size_t(a)]; t < uv_base[size_t(a) + 1]; ++t)
            if (uv[size_t(t)].loc.kind == BasisLoc::Kind::Edge && uv[size_t(t)].loc.id == int(id)) rel.push_back(t);
        }
        for (size_t p = 0; p < rel.size(); ++p)
          for (size_t q = p + 1; q < rel.size(); ++q)
            if (uv[size_t(rel[p])].matches_with_edge(uv[size_t(rel[q])])) {
              pairs.push_back({rel[p], rel[q]});
              break;
            }
        pair_count[id] = int(pairs.size()) - pair_start[id];
      }
    }
  });
  std::vector<int64_t> edge_base(ned + 1, elem_base[ne]);
  for (size_t id = 0; id < ned; ++id) edge_base[id + 1] = edge_base[id] + pair_count[id];
  const int64_t n_dofs = edge_base[ned];
  if (n_dofs > INT32_MAX) throw MeshPanic("too many DoFs for 32-bit ids");
  auto edge_pairs = [&](size_t id) { return block_pairs[id / kEdgeBlock].data() + pair_start[id]; };

  // 4. DoF records (sized first, so the two Elems of an edge-type DoF fill
  // their own address concurrently) and per-Elem lists
  Domain d;
  d.cc = cc;
  d.basis_specs.assign(ne, {});
  d.dofs.resize(size_t(n_dofs));  // unwritten until here
  parallel_ranges(size_t(n_dofs), threads, 4096, [&](size_t b, size_t e) {
    for (size_t k = b; k < e; ++k) d.dofs[k] = DoF{int(k), int64_t(k) < elem_base[ne] ? BSAddressList{{0, 0}}
                                                                                      : BSAddressList{{0, 0}, {0, 0}}};
  });
  parallel_ranges(ne, threads, 256, [&](size_t b, size_t e) {
    for (size_t x = b; x < e; ++x) {
      // the Edges of this Elem in id order
      std::array<int, 4> eds = mesh.elems[x].edges;
      std::sort(eds.begin(), eds.end());
      size_t n_list = size_t(n_elem_type[x]);
      for (size_t k = 0; k < 4; ++k) {
        const int id = eds[k];
        if (id < 0 || id >= int(ned) || (k > 0 && eds[k - 1] == id)) continue;
        const auto* P = edge_pairs(size_t(id));
        for (int pi = 0; pi < pair_count[size_t(id)]; ++pi)
          n_list += size_t(uv[size_t(P[pi][0])].elem_id == int(x)) + size_t(uv[size_t(P[pi][1])].elem_id == int(x));
      }
      std::vector<BasisSpec>& list = d.basis_specs[x];
      list.reserve(n_list);
      auto push = [&](const BasisSpec& s0, int64_t dof_id, size_t side) {
        BasisSpec bs = s0;
        bs.elem_idx = int(list.size());
        bs.dof_id = int(dof_id);
        d.dofs[size_t(dof_id)].basis_specs[side] = BSAddress{int(x), bs.elem_idx};
        list.push_back(bs);
      };
      if (n_elem_type[x] > 0) {
        int64_t dof_id = elem_base[x];
        for (int64_t t = uv_base[x]; t < uv_base[x + 1]; ++t)
          if (uv[size_t(t)].loc.kind == BasisLoc::Kind::Elem) push(uv[size_t(t)], dof_id++, 0);
      }
      for (size_t k = 0; k < 4; ++k) {
        const int id = eds[k];
        if (id < 0 || id >= int(ned) || (k > 0 && eds[k - 1] == id)) continue;
        const auto* P = edge_pairs(size_t(id));
        for (int pi = 0; pi < pair_count[size_t(id)]; ++pi)
          for (size_t side = 0; side < 2; ++side)
            if (uv[size_t(P[pi][side])].elem_id == int(x)) push(uv[size_t(P[pi][side])], edge_base[size_t(id)] + pi, side);
      }
    }
  });

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
