// Helpers shared by the host and device halves of the C ABI.
#pragma once

#include <string>

struct fem2d_ctx;
struct fem2d_gep_device;
struct fem2d_eigen_stats;

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
}  // namespace capi

}  // namespace fem2d
