// Host-side Domain: basis specifications, edge matching and DoF numbering.
//
// Restates Domain::from_mesh (reference src/fem_domain/domain.rs:69-159),
// BasisSpec (domain/dof/basis_spec.rs) and DoF (domain/dof.rs). The DoF ids it
// assigns define the row/column numbering of the assembled matrices: every
// elem-type DoF of every leaf Elem first (ascending elem id, U directions then
// V directions, i outer and j inner), then edge-type DoFs by ascending edge id.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "mesh.hpp"

namespace fem2d {

// domain.rs:18-23
enum class ContinuityCondition : int { HCurl = 0, HDiv = 1, Discontinuous = 2 };

// basis_spec.rs:178-185
enum class BasisDir : uint8_t { U = 0, V = 1, W = 2 };

// basis_spec.rs:187-199
struct BasisLoc {
  enum class Kind : uint8_t { Elem = 0, Edge = 1, Node = 2 };
  Kind kind = Kind::Elem;
  uint8_t idx = 0;  // local edge/node index
  int id = -1;      // edge/node id
  bool operator==(const BasisLoc& o) const { return kind == o.kind && idx == o.idx && id == o.id; }
};

struct BasisSpec {
  int id = 0;
  uint8_t i = 0, j = 0;
  BasisDir dir = BasisDir::U;
  int elem_id = 0;
  int elem_idx = -1;
  int dof_id = -1;
  BasisLoc loc;

  // basis_spec.rs:41-75 (HCurl location rule; other conditions are not implemented)
  static BasisSpec make(int id, uint8_t i, uint8_t j, BasisDir dir, const Elem& elem, ContinuityCondition cc);
  // basis_spec.rs:80-110
  bool matches_with_edge(const BasisSpec& other) const;
};

struct BSAddress {
  int elem_id = 0;
  int elem_idx = 0;
};

struct DoF {
  int id = 0;
  std::vector<BSAddress> basis_specs;  // 1 (elem), 2 (edge) or 4 (node) addresses
};

class Domain {
 public:
  Mesh mesh;
  std::vector<DoF> dofs;
  std::vector<std::vector<BasisSpec>> basis_specs;
  ContinuityCondition cc = ContinuityCondition::HCurl;

  static Domain blank(ContinuityCondition cc);
  static Domain from_mesh(Mesh mesh, ContinuityCondition cc);

  const std::vector<BasisSpec>& local_basis_specs(int elem_id) const;
  std::vector<std::pair<int, const std::vector<BasisSpec>*>> descendant_basis_specs(int elem_id) const;
  std::vector<std::pair<int, const std::vector<BasisSpec>*>> ancestor_basis_specs(int elem_id) const;
};

// Elem-type BasisSpec permutations of one Elem for a direction
// (PolyOrders::permutations, p_refinement.rs:49-64).
void for_each_permutation(const PolyOrders& po, BasisDir dir, const std::function<void(uint8_t, uint8_t)>& f);

}  // namespace fem2d
