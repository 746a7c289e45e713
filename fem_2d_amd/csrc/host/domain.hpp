// Host-side Domain: basis specifications, edge matching and DoF numbering.
//
// Restates Domain::from_mesh (reference src/fem_domain/domain.rs:69-159),
// BasisSpec (domain/dof/basis_spec.rs) and DoF (domain/dof.rs). The DoF ids it
// assigns define the row/column numbering of the assembled matrices: every
// elem-type DoF of every leaf Elem first (ascending elem id, U directions then
// V directions, i outer and j inner), then edge-type DoFs by ascending edge id.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <initializer_list>
#include <vector>

#include "mesh.hpp"
#include "parallel.hpp"

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

// BSAddress{} is (0, 0); a default-initialised one is unset (so vectors of
// DoFs can be sized without writing them)
struct BSAddress {
  int elem_id;
  int elem_idx;
};

// A DoF's BasisSpec addresses, held inline: 1 for an elem-type DoF, 2 for an
// edge-type one, 4 for a node-type one (dof.rs:11-28)
class BSAddressList {
 public:
  // BSAddressList{} is empty; a default-initialised list is unset until resize()
  BSAddressList() = default;
  BSAddressList(std::initializer_list<BSAddress> l) : a_{}, n_(0) {
    for (const BSAddress& x : l) push_back(x);
  }
  size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  const BSAddress& operator[](size_t k) const { return a_[k]; }
  BSAddress& operator[](size_t k) { return a_[k]; }
  const BSAddress* begin() const { return a_.data(); }
  const BSAddress* end() const { return a_.data() + n_; }
  void push_back(const BSAddress& x) {
    if (n_ == a_.size()) throw MeshPanic("a DoF holds at most 4 BasisSpecs");
    a_[n_++] = x;
  }
  void resize(size_t n) {
    if (n > a_.size()) throw MeshPanic("a DoF holds at most 4 BasisSpecs");
    n_ = uint8_t(n);
  }

 private:
  std::array<BSAddress, 4> a_;
  uint8_t n_;
};

// DoF{} is (0, no specs); a default-initialised DoF is unset
struct DoF {
  int id;
  BSAddressList basis_specs;
};

class Domain {
 public:
  Mesh mesh;
  uninit_vector<DoF> dofs;
  std::vector<std::vector<BasisSpec>> basis_specs;
  ContinuityCondition cc = ContinuityCondition::HCurl;

  static Domain blank(ContinuityCondition cc);
  // Domain::from_mesh, data-parallel on up to `threads` host threads (0: the
  // machine's cores, at most 16). Same Domain as from_mesh_serial, bit for bit.
  static Domain from_mesh(Mesh mesh, ContinuityCondition cc, int threads = 0);
  // the reference's sequential algorithm, step by step (domain.rs:69-159)
  static Domain from_mesh_serial(Mesh mesh, ContinuityCondition cc);

  const std::vector<BasisSpec>& local_basis_specs(int elem_id) const;
  std::vector<std::pair<int, const std::vector<BasisSpec>*>> descendant_basis_specs(int elem_id) const;
  std::vector<std::pair<int, const std::vector<BasisSpec>*>> ancestor_basis_specs(int elem_id) const;
};

// Elem-type BasisSpec permutations of one Elem for a direction
// (PolyOrders::permutations, p_refinement.rs:49-64).
void for_each_permutation(const PolyOrders& po, BasisDir dir, const std::function<void(uint8_t, uint8_t)>& f);

}  // namespace fem2d
