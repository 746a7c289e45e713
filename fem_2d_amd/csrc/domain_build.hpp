// Domain construction on the GPU from a flattened Mesh: the edge activation of
// Mesh::set_edge_activation (mesh.rs:1180-1215, edge.rs:203-214) and the DoF
// numbering and per-Elem BasisSpec lists of Domain::from_mesh
// (domain.rs:69-159), written straight into the device Domain's arrays.
#pragma once

#include <hip/hip_runtime.h>

namespace fem2d {
namespace build {

// failures the kernels record (Args::status->flags)
constexpr unsigned kFlagNoPair = 1u;        // a base Edge without an active Elem pair
constexpr unsigned kFlagInconsistent = 2u;  // one child of an active Edge has a pair and the other not
constexpr unsigned kFlagTooDeep = 4u;       // an Edge tree deeper than kMaxEdgeDepth
constexpr unsigned kFlagCapacity = 8u;      // a spec or DoF index beyond the arrays' capacity
constexpr int kMaxEdgeDepth = 64;

// Written by the kernels, read back by the host (starts zeroed).
struct Status {
  long long elem_dofs;   // elem-type DoFs (the leaves' interior specs)
  long long edge_dofs;   // edge-type DoFs (matched pairs over the active Edges)
  long long specs;       // entries of the per-Elem BasisSpec lists
  unsigned long long groups;  // row groups: runs of DoFs whose specs lie on the same Elems
  unsigned flags;
  int err;               // the Domain preparation kernels' error word
};

struct Args {
  int n_elems = 0, n_edges = 0;
  long long spec_cap = 0;  // capacity of the spec arrays, and of the DoF arrays
  // the flattened Mesh
  const int4* elem_edges = nullptr;        // [n_elems]
  const unsigned* elem_meta = nullptr;     // [n_elems] ni | nj << 8 | (no children) << 16
  const int2* side = nullptr;              // [n_edges] candidate active Elem of each side, -1 for none
  const int* edge_parent = nullptr;        // [n_edges]
  const int2* child = nullptr;             // [n_edges] -1 when not refined
  const unsigned char* open_base = nullptr;  // [n_edges] base Edge off the boundary
  // work
  int2* act = nullptr;         // [n_edges] active pair (smaller id first), (-1, -1) for none
  int* n_pairs = nullptr;      // [n_edges] matched BasisSpec pairs = edge-type DoFs of the Edge
  int* n_elem = nullptr;       // [n_elems] elem-type DoFs of the Elem
  int* list_len = nullptr;     // [n_elems] length of the Elem's BasisSpec list
  long long* elem_base = nullptr;  // [n_elems + 1] first elem-type DoF of each Elem
  long long* edge_base = nullptr;  // [n_edges + 1] first edge-type DoF of each Edge, after the elem-type ones
  long long* tile_sums = nullptr;  // [scan_tiles()]
  Status* status = nullptr;
  // the device Domain's arrays
  long long* bs_off = nullptr;  // [n_elems + 1]
  unsigned char *bs_dir = nullptr, *bs_i = nullptr, *bs_j = nullptr;  // [spec_cap]
  int* bs_dof = nullptr;        // [spec_cap]
  int* dof_bs = nullptr;        // [2 * spec_cap]
};

// entries of Args::tile_sums the counts need
long long scan_tiles(int n_elems, int n_edges);
// edge activation, per-Elem and per-Edge counts and their prefix sums; the
// totals land in Args::status
hipError_t launch_counts(const Args& a, hipStream_t st);
// the per-Elem BasisSpec lists and each DoF's spec addresses (after the counts)
hipError_t launch_lists(const Args& a, hipStream_t st);
// row groups of the DoFs into status->groups (bs_elem: each spec's Elem)
hipError_t launch_group_count(const int* dof_bs, const int* bs_elem, long long n_dofs, long long n_bs, Status* status,
                              hipStream_t st);

}  // namespace build
}  // namespace fem2d
