// Helpers shared by the host and device halves of the C ABI.
#pragma once

#include <string>

struct fem2d_ctx;

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
}  // namespace capi

}  // namespace fem2d
