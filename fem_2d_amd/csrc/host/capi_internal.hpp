// Helpers shared by the host and device halves of the C ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

struct fem2d_ctx;
struct fem2d_gep_device;
struct fem2d_eigen_stats;
struct fem2d_device_domain;

namespace fem2d {

// Gauss-Legendre nodes (ascending) and weights on (-1, 1).
void gauss_legendre_nodes(int n, double* x, double* w);

namespace capi {
void set_error(const std::string& msg, int detail);
// Maps the in-flight exception to a status code and records its message.
int translate_exception();
// HIP device of a context
int ctx_device(fem2d_ctx* ctx);
// the context's eigensolver handle (null until the first solve)
void** ctx_solver_slot(fem2d_ctx* ctx);
// destroys a handle held in that slot
void release_solver(void* handle);
// the context's statistics of its last sparse eigensolve
fem2d_eigen_stats* ctx_eigen_stats(fem2d_ctx* ctx);
// FEM2D_OK for a device GEP handle that may be read on this context: one the
// caller built over arrays of its own (generation 0), or the latest assembly
// of this context; FEM2D_ERR_INVALID_ARG (message set) for a stale handle
int check_gep_handle(fem2d_ctx* ctx, const fem2d_gep_device* gep);

// A Mesh as flat arrays, the input of the device Domain builder. Every index
// is checked against its array when the arrays are made, and each Elem's four
// Edges are distinct.
struct MeshFlat {
  int32_t n_elems = 0, n_edges = 0;
  int max_i = 0, max_j = 0;       // Mesh::max_expansion_orders
  int64_t uv_specs = 0;           // U and V BasisSpecs over every Elem: bounds the Domain's specs and DoFs
  int64_t all_specs = 0;          // BasisSpecs of every direction (the reference's spec ids count them all)
  // per Elem
  std::vector<int32_t> parent;    // -1 on the base layer
  std::vector<uint8_t> loc;       // HRefLoc inside the parent (0 on the base layer)
  std::vector<double> box;        // [4 * n_elems] x0, y0, x3, y3 of the Elem's Element
  std::vector<double> mat;        // [2 * n_elems] eps_rel.re, mu_rel.re
  std::vector<int32_t> edges;     // [4 * n_elems]
  std::vector<uint32_t> meta;     // ni | nj << 8 | (no children) << 16
  // per Edge
  std::vector<int32_t> side;      // [2 * n_edges] Elem of the highest h-level key on each side, -1 for none
  std::vector<int32_t> edge_parent;
  std::vector<int32_t> child;     // [2 * n_edges], -1 when not refined
  std::vector<uint8_t> open_base; // 1 for a base Edge (no parent) off the boundary
};

// Builds the device Domain of a flattened Mesh on the context's GPU (edge
// activation, DoF numbering and the per-Elem BasisSpec lists of
// Domain::from_mesh). Returns kActivationFailed, with no message set, when an
// Edge has no active pair where Mesh::set_edge_activation would panic; the
// caller reports that panic.
constexpr int kActivationFailed = -1000;
int device_domain_from_flat(fem2d_ctx* ctx, const MeshFlat& m, int32_t cc, fem2d_device_domain** out);
}  // namespace capi

}  // namespace fem2d
