/*
 * fem2d.h - C ABI of the MI355X Galerkin GEP assembly library (libfem2d_amd.so).
 *
 * The drop-in boundary is fem2d_galerkin_sample_gep_hcurl(), which replaces
 *   galerkin_sample_gep_hcurl::<BSpace, AI, BI>(&domain, glq_grid_dim)
 * of the reference (src/fem_problem/galerkin.rs:33-187). The reference's
 * Domain crosses the boundary as a flattened, caller-owned fem2d_domain_view
 * (plain arrays, no ownership transfer); the GEP comes back as upper-triangular
 * CSR holding A and B on one shared sparsity pattern.
 *
 * The remaining entry points are the host-side input builder (Mesh and Domain,
 * reference src/fem_domain/domain.rs and domain/mesh.rs), the device-resident
 * variant of the assembly used by callers that keep the Domain in HBM across
 * calls, and error/timing queries.
 *
 * All functions return FEM2D_OK (0) or a status code; they never abort the
 * process. fem2d_last_error() returns the message of the calling thread's last
 * failure.
 */
#ifndef FEM2D_H
#define FEM2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Status codes
 * ------------------------------------------------------------------------- */
enum fem2d_status {
  FEM2D_OK = 0,
  /* GalerkinSamplingError (galerkin.rs:191-195), checked before any work */
  FEM2D_ERR_WRONG_CC = 1,    /* WrongContinuityCondition(HCurl, got) */
  FEM2D_ERR_EMPTY_DOFS = 2,  /* EmptyDOFSet */
  FEM2D_ERR_INVALID_GLQ = 3, /* InvalidGLQSettings: fewer than FEM2D_MIN_GLQ_ORDER points */
  /* input-builder errors; fem2d_last_error_detail() gives the variant */
  FEM2D_ERR_HREF = 10,        /* HRefError (h_refinement.rs:284-296) */
  FEM2D_ERR_PREF = 11,        /* PRefError (p_refinement.rs:305-317) */
  FEM2D_ERR_MESH_ACCESS = 12, /* MeshAccessError (mesh.rs:1799-1804) */
  FEM2D_ERR_PANIC = 13,       /* a condition on which the reference panics */
  FEM2D_ERR_IO = 14,          /* mesh file missing or unreadable */
  /* eigensolver errors: NalgebraGEPError (nalgebra_solve.rs:52-58) and
   * SlepcGEPError::FailedToConverge (slepc_solve.rs:70-81) */
  FEM2D_ERR_FAILED_TO_INVERT_B = 20,   /* B is not positive definite (its Cholesky factorisation fails) */
  FEM2D_ERR_SPURIOUSLY_CONVERGED = 21, /* every eigenvalue is below 1e-12 in magnitude */
  FEM2D_ERR_PROBLEM_TOO_LARGE = 22,    /* more DoFs than the solver takes */
  FEM2D_ERR_NOT_CONVERGED = 23,        /* the eigensolver's iteration did not converge */
  /* UniformFieldError::MismatchedSolutionSize (fields.rs:407-411) */
  FEM2D_ERR_MISMATCHED_SOLUTION_SIZE = 24,
  /* library errors */
  FEM2D_ERR_INVALID_ARG = 100,
  FEM2D_ERR_INVALID_VIEW = 101, /* domain view violates the Domain invariants */
  FEM2D_ERR_UNSUPPORTED = 102,  /* a size or option beyond this build's limits */
  FEM2D_ERR_HIP = 110,          /* HIP runtime failure (no device, launch failure, ...) */
  FEM2D_ERR_OOM = 111
};

/* galerkin.rs:13 */
#define FEM2D_MIN_GLQ_ORDER 4

/* ContinuityCondition (domain.rs:18-23) */
enum fem2d_cc { FEM2D_CC_HCURL = 0, FEM2D_CC_HDIV = 1, FEM2D_CC_DISCONTINUOUS = 2 };

/* HierCurlBasisFnSpace implementations: HierPoly (hierarchical_basis_fns.rs:15-198)
 * and KOLShapeFn, the Legendre-based maximally-orthogonal space modelled on
 * HierMaxOrtho (hierarchical_basis_fns.rs:201-647) with the order-10 coefficient
 * row corrected so every T_n (n >= 2) vanishes at both endpoints. */
enum fem2d_basis { FEM2D_BASIS_HIERPOLY = 0, FEM2D_BASIS_KOL = 1 };

/* HierCurlIntegral implementations (integrals.rs): CurlCurl and L2Inner */
enum fem2d_integral { FEM2D_INTEGRAL_CURLCURL = 0, FEM2D_INTEGRAL_L2INNER = 1 };

/* Precision of the assembled values. FP64: the reference's. The mixed modes keep
 * the same FP64 arithmetic and round values once, at the store; the pattern is
 * the same.
 * FP32: A and B stored as FP32, 12 bytes of HBM per stored entry instead of 20
 *   (relative entry error <= 2^-24; the Rayleigh quotients of the lowest modes
 *   move by up to ~1e-7).
 * FP32R: A and B stored as FP32 plus a 16-bit residual each, 16 bytes per entry:
 *   x = hi + q * 2^(e - 38) with hi = fl32(x), e the binary exponent of hi and q
 *   a signed 16-bit integer, so the relative entry error is <= 2^-39 and hi alone
 *   is the FP32 matrix.
 * The separable (sum-factorised) contraction leaves the assembly bound by these
 * writes. Host results (fem2d_gep_csr) hold the stored values widened to double
 * (FP32R: hi plus its residual). */
enum fem2d_precision { FEM2D_PRECISION_FP64 = 0, FEM2D_PRECISION_FP32 = 1, FEM2D_PRECISION_FP32R = 2 };

/* BasisDir (basis_spec.rs:259-268) */
enum fem2d_basis_dir { FEM2D_DIR_U = 0, FEM2D_DIR_V = 1, FEM2D_DIR_W = 2 };

/* HRefLoc (h_refinement.rs:210-229) */
enum fem2d_href_loc {
  FEM2D_LOC_SW = 0, FEM2D_LOC_SE = 1, FEM2D_LOC_NW = 2, FEM2D_LOC_NE = 3,
  FEM2D_LOC_W = 4, FEM2D_LOC_E = 5, FEM2D_LOC_S = 6, FEM2D_LOC_N = 7
};

/* HRef variants (h_refinement.rs:60-77): kind T/U/V plus the optional index of
 * the child that is refined again in the other direction (-1 for none). */
enum fem2d_href_kind { FEM2D_HREF_T = 0, FEM2D_HREF_U = 1, FEM2D_HREF_V = 2 };

/* ---------------------------------------------------------------------------
 * Drop-in boundary: flattened Domain in, upper-triangular CSR GEP out
 * ------------------------------------------------------------------------- */

/* Flattened view of a reference Domain (domain.rs:42-50). Arrays are owned by
 * the caller and only read during the call. Elem ids are array indices. */
typedef struct fem2d_domain_view {
  int32_t cc;        /* Domain::cc */
  int32_t n_elems;   /* mesh.elems.len() */
  int64_t n_dofs;    /* dofs.len() */
  const int32_t* parent;      /* [n_elems] Elem::parent_id(), -1 on the base layer */
  const uint8_t* loc;         /* [n_elems] HRefLoc of the Elem inside its parent */
  const double* element_box;  /* [4*n_elems] x0, y0, x3, y3: points[0] and points[3] of the Elem's Element */
  const double* materials;    /* [2*n_elems] eps_rel.re, mu_rel.re of the Elem's Element */
  const uint8_t* orders;      /* [2*n_elems] poly_orders ni, nj */
  const int64_t* bs_offsets;  /* [n_elems+1] basis_specs[e] occupies [bs_offsets[e], bs_offsets[e+1]) */
  const uint8_t* bs_dir;      /* [n_bs] BasisSpec::dir */
  const uint8_t* bs_i;        /* [n_bs] BasisSpec::i */
  const uint8_t* bs_j;        /* [n_bs] BasisSpec::j */
  const int32_t* bs_dof;      /* [n_bs] BasisSpec::dof_id */
  const int32_t* dof_bs;      /* [2*n_dofs] DoF::basis_specs (dof.rs:11-28): the flat index
                                 bs_offsets[elem_id] + elem_idx of each DoF's first and second
                                 BSAddress, -1 for a DoF with one (H1 node DoFs, with four,
                                 list their first two; the H(curl) assembly rejects them) */
} fem2d_domain_view;

/* GEP { a, b } (linalg.rs:29-34) as upper-triangular CSR: one pattern shared by
 * A and B, columns ascending within a row and >= the row index, explicit zeros
 * kept. nnz counts the stored upper-triangular entries (the reference's
 * SparseMatrix::num_entries() is 2*nnz - number of diagonal entries). */
typedef struct fem2d_gep_csr {
  int64_t n;
  int64_t nnz;
  int64_t* row_ptr; /* [n+1] */
  int32_t* col;     /* [nnz] */
  double* a_val;    /* [nnz] */
  double* b_val;    /* [nnz] */
} fem2d_gep_csr;

/* GPU context: owns a HIP stream, device work buffers and per-stage timers.
 * One context per host thread; calls on one context are not re-entrant. */
typedef struct fem2d_ctx fem2d_ctx;

int fem2d_ctx_create(int device, fem2d_ctx** out);
void fem2d_ctx_destroy(fem2d_ctx* ctx);
/* hipStream_t the context launches on */
void* fem2d_ctx_stream(fem2d_ctx* ctx);
/* device-wide synchronisation of the context's device */
int fem2d_ctx_synchronize(fem2d_ctx* ctx);
/* Whether assemblies on this context are recorded once as a HIP graph and
 * replayed (default 1) or launched kernel by kernel (0). Results are the same;
 * graphs remove the per-kernel launch gaps of the ~20-kernel pipeline. */
int fem2d_ctx_set_graphs(fem2d_ctx* ctx, int enable);
/* Whether the first assembly of a Domain in a precision times the entry fill at
 * each candidate chunk size (128, 256, 512, 1024 CSR entries per work item) and
 * library placement (an LDS copy per block when it fits, or global memory) and
 * keeps the fastest for later assemblies of that Domain (default 1). The tuned
 * shape changes timing only, never results. */
int fem2d_ctx_set_autotune(fem2d_ctx* ctx, int enable);
/* Forces the streaming fill's chunk size (CSR entries per wavefront work item,
 * a power of two from 64 to 4096) for every later assembly on the context, with
 * no tuning while set; 0 returns to the tuned or default size. Results do not
 * depend on it. */
int fem2d_ctx_set_fill_chunk(fem2d_ctx* ctx, int chunk);

/* Replaces galerkin_sample_gep_hcurl (galerkin.rs:33-187).
 * basis: fem2d_basis; a_integral / b_integral: fem2d_integral (the AI / BI
 * generic arguments); glq_u, glq_v: Gauss-Legendre points per direction, both 0
 * for the default of 4*max_order rounded up to a power of two (basis.rs:172-177);
 * precision: fem2d_precision. Errors are checked in the reference's order:
 * continuity condition, empty DoF set, GLQ settings. On success *out holds
 * library-allocated host arrays; release them with fem2d_gep_csr_free. */
int fem2d_galerkin_sample_gep_hcurl(fem2d_ctx* ctx, const fem2d_domain_view* domain, int basis, int a_integral,
                                    int b_integral, int glq_u, int glq_v, int precision, fem2d_gep_csr* out);
void fem2d_gep_csr_free(fem2d_gep_csr* gep);

/* ---------------------------------------------------------------------------
 * Device-resident assembly (same computation, Domain and GEP stay in HBM)
 * ------------------------------------------------------------------------- */

typedef struct fem2d_device_domain fem2d_device_domain;
/* Mesh (mesh.rs) of the host input builder below */
typedef struct fem2d_mesh fem2d_mesh;

/* Device pointers owned by the context; valid until the next assembly on that
 * context or its destruction. fem2d_assemble_device stamps the handle with the
 * assembly's generation (unique in the process); every function that takes a
 * handle rejects one whose generation is not the context's latest assembly
 * with FEM2D_ERR_INVALID_ARG (a stale handle, or one from another context).
 * Generation 0 marks a handle the caller built over device arrays of its own;
 * those are not checked. */
typedef struct fem2d_gep_device {
  int64_t n;
  int64_t nnz;
  int64_t* row_ptr;
  int32_t* col;
  double* a_val;     /* FEM2D_PRECISION_FP64 values, else NULL */
  double* b_val;
  int32_t precision; /* fem2d_precision of the values */
  float* a_val32;    /* FEM2D_PRECISION_FP32 and _FP32R values (FP32R: the FP32 parts), else NULL */
  float* b_val32;
  uint64_t generation; /* assembly that produced the handle; 0 for caller-owned arrays */
  uint32_t* res16;   /* FEM2D_PRECISION_FP32R: per entry, A's 16-bit residual in the low half and B's
                        in the high half (two's complement), else NULL */
} fem2d_gep_device;

int fem2d_device_domain_upload(fem2d_ctx* ctx, const fem2d_domain_view* domain, fem2d_device_domain** out);
/* Domain::from_mesh (domain.rs:69-159) built on the GPU, straight into a device
 * Domain: the Mesh is flattened on the host and copied up once, then the edge
 * activation (Mesh::set_edge_activation, mesh.rs:1180-1215), the DoF numbering
 * and the per-Elem BasisSpec lists are computed by kernels. The result is the
 * device Domain fem2d_device_domain_upload makes of the view of
 * fem2d_domain_from_mesh(m, cc): the same arrays, bit for bit. Panics of the
 * reference's construction (no active Edge pair, a non-H(curl) condition on a
 * non-empty Mesh) return FEM2D_ERR_PANIC with its message. */
int fem2d_device_domain_from_mesh(fem2d_ctx* ctx, const fem2d_mesh* m, int32_t cc, fem2d_device_domain** out);
void fem2d_device_domain_free(fem2d_device_domain* dd);
/* sizes of a device Domain (any pointer may be NULL) */
int fem2d_device_domain_info(const fem2d_device_domain* dd, int32_t* n_elems, int64_t* n_dofs, int64_t* n_bs);
/* copies a device Domain's view arrays to host arrays sized as fem2d_domain_view
 * describes them (orders are not kept on the device); NULL arrays are skipped */
int fem2d_device_domain_download_view(fem2d_ctx* ctx, const fem2d_device_domain* dd, int32_t* parent, uint8_t* loc,
                                      double* element_box, double* materials, int64_t* bs_offsets, uint8_t* bs_dir,
                                      uint8_t* bs_i, uint8_t* bs_j, int32_t* bs_dof, int32_t* dof_bs);
int fem2d_assemble_device(fem2d_ctx* ctx, const fem2d_device_domain* dd, int basis, int a_integral, int b_integral,
                          int glq_u, int glq_v, int precision, fem2d_gep_device* out);
int fem2d_gep_device_download(fem2d_ctx* ctx, const fem2d_gep_device* src, fem2d_gep_csr* out);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory during the call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the assembly of fem2d_assemble_device, same
 * kernels and results, with no hipMalloc, hipFree or graph instantiation in the
 * call. The caller passes the outputs: row_ptr[n_dofs + 1] (int64), col[capacity]
 * (int32), a_val / b_val[capacity] (double for FEM2D_PRECISION_FP64, float for
 * FEM2D_PRECISION_FP32 and _FP32R) and, for FP32R only, res16[capacity] (NULL
 * otherwise). The work memory is a static device array of
 * FEM2D_NOALLOC_SCRATCH_BYTES per device, held until the process ends; a Domain
 * whose assembly needs more (about 7.3 KB per row group plus ~100 B per DoF)
 * is rejected with FEM2D_ERR_UNSUPPORTED before any device work. *nnz receives
 * the GEP's stored entries; when they exceed capacity the call returns
 * FEM2D_ERR_INVALID_ARG after writing row_ptr and nothing else, so the caller can
 * size its buffers and call again. Every call on one device shares that device's
 * array: calls that overlap on one device (from several contexts or threads)
 * race and must be serialised by the caller. The kernels run on the context's
 * stream; the call returns after they finish. */
#define FEM2D_NOALLOC_SCRATCH_BYTES (64ll << 20)
int fem2d_assemble_device_noalloc(fem2d_ctx* ctx, const fem2d_device_domain* dd, int basis, int a_integral,
                                  int b_integral, int glq_u, int glq_v, int precision, int64_t capacity,
                                  int64_t* row_ptr, int32_t* col, void* a_val, void* b_val, uint32_t* res16,
                                  int64_t* nnz);

/* PETSc AIJ binary files of a GEP, replacing GEP::print_to_petsc_binary_files
 * (linalg.rs:44-52) with AIJMatrixBinary (sparse_matrix.rs:184-265): writes
 * "{dir}/tmp/{prefix}_a.dat" and "{dir}/tmp/{prefix}_b.dat", each the full
 * symmetric matrix (both triangles, rows in order, columns ascending) as
 * big-endian u32 1211216, rows, columns, nnz, row lengths, column indices and
 * f64 values. The transpose and the byte order are produced on the device.
 * The device variant takes a GEP as fem2d_assemble_device returns it; the host
 * variant checks the CSR and uploads it first. FEM2D_ERR_IO if a file cannot
 * be written (the reference's std::io::Error). */
int fem2d_gep_device_print_to_petsc_binary_files(fem2d_ctx* ctx, const fem2d_gep_device* gep, const char* dir,
                                                 const char* prefix);
int fem2d_gep_csr_print_to_petsc_binary_files(fem2d_ctx* ctx, const fem2d_gep_csr* gep, const char* dir,
                                              const char* prefix);

/* ---------------------------------------------------------------------------
 * Dense generalized eigensolves of an assembled GEP (rocSOLVER on the
 * context's stream; the matrices are expanded to dense column-major storage
 * in HBM by the library's own kernel)
 * ------------------------------------------------------------------------- */

/* EigenPair (linalg.rs:84-97) */
typedef struct fem2d_eigenpair {
  double value;
  int64_t n;
  double* vector; /* n values in host memory owned by the library; fem2d_eigenpair_free */
} fem2d_eigenpair;
void fem2d_eigenpair_free(fem2d_eigenpair* pair);

/* nalgebra_solve_gep (nalgebra_solve.rs:14-50), its arithmetic kept: at most
 * FEM2D_NALGEBRA_MAX_DENSE_SIZE DoFs (FEM2D_ERR_PROBLEM_TOO_LARGE above);
 * B = L L^T (FEM2D_ERR_FAILED_TO_INVERT_B if that fails); C = B^-1 A; the
 * symmetric eigenproblem of C's lower triangle (C itself is not symmetric,
 * and this is what the reference's SymmetricEigen reads); FEM2D_ERR_SPURIOUSLY_CONVERGED
 * if every eigenvalue is below 1e-12 in magnitude; else the eigenpair whose
 * value is nearest target_eigenvalue (the lowest such value on a tie), with a
 * unit-2-norm vector. */
#define FEM2D_NALGEBRA_MAX_DENSE_SIZE 1000
int fem2d_nalgebra_solve_gep(fem2d_ctx* ctx, const fem2d_gep_csr* gep, double target_eigenvalue,
                             fem2d_eigenpair* out);
int fem2d_gep_device_nalgebra_solve(fem2d_ctx* ctx, const fem2d_gep_device* gep, double target_eigenvalue,
                                    fem2d_eigenpair* out);

/* The symmetric-definite problem A x = lambda B x itself, standing in for
 * slepc_solve_gep (slepc_solve.rs:14-68) up to FEM2D_DENSE_GEP_MAX_SIZE DoFs:
 * the eigenpair whose value is nearest target_eigenvalue (the lowest on a tie),
 * the vector B-normalised (x^T B x = 1). Shift-invert inverse iteration on the
 * LU factors of A - target B; the whole spectrum (dsygvd) when that is
 * inconclusive. Errors: FEM2D_ERR_PROBLEM_TOO_LARGE,
 * FEM2D_ERR_FAILED_TO_INVERT_B, FEM2D_ERR_SPURIOUSLY_CONVERGED (as above),
 * FEM2D_ERR_NOT_CONVERGED. */
#define FEM2D_DENSE_GEP_MAX_SIZE 32768
int fem2d_solve_gep_dense(fem2d_ctx* ctx, const fem2d_gep_csr* gep, double target_eigenvalue, fem2d_eigenpair* out);
int fem2d_gep_device_solve_dense(fem2d_ctx* ctx, const fem2d_gep_device* gep, double target_eigenvalue,
                                 fem2d_eigenpair* out);

/* The same problem for GEPs of any size whose factor fits the device memory,
 * standing in for slepc_solve_gep (slepc_solve.rs:14-68), which runs SLEPc's
 * shift-and-invert around the target: the eigenpair nearest target_eigenvalue,
 * the vector B-normalised. The matrix graph is ordered by breadth-first level
 * sets (host, from the pattern), which makes A - target B block tridiagonal;
 * the SPD matrix A + tau B (tau = max(|target|, 1)) is block-Cholesky factored
 * on the device (rocSOLVER / rocBLAS, dense Schur-complement blocks) and
 * preconditions GMRES solves with A - target B inside a shift-invert inverse
 * iteration that runs against the sparse matrices in place. It converges when the relative residual
 * |A x - lambda B x| / max(|A x|, |target| |B x|) reaches 1e-12 (1e-10 once the
 * Rayleigh quotient stops moving). Device memory: 16 bytes times the sum of the
 * squared block widths (about 41 GB for the 445 380-DoF benchmark GEP).
 * Errors: FEM2D_ERR_PROBLEM_TOO_LARGE (the blocks do not fit the free device
 * memory), FEM2D_ERR_FAILED_TO_INVERT_B (A + tau B is not positive definite),
 * FEM2D_ERR_SPURIOUSLY_CONVERGED (no DoFs), FEM2D_ERR_NOT_CONVERGED. */
int fem2d_solve_gep_sparse(fem2d_ctx* ctx, const fem2d_gep_csr* gep, double target_eigenvalue,
                           fem2d_eigenpair* out);
/* Statistics of the last sparse eigensolve on a context (also after a failure). */
typedef struct fem2d_eigen_stats {
  int64_t n;                 /* DoFs */
  int32_t blocks;            /* blocks of the level ordering */
  int32_t max_block;         /* widest block */
  double factor_bytes;       /* device bytes of the kept inverse Cholesky factors and their transposes */
  int32_t iterations;        /* outer (shift-invert) iterations */
  int32_t inner_iterations;  /* preconditioned GMRES steps over all outer iterations */
  double residual;           /* final relative residual */
  double max_inner_per_solve; /* most GMRES steps of one inner solve */
  double symbolic_ms;        /* host ordering and entry lists */
  double factor_ms;          /* block factorisation on the device */
  double iterate_ms;         /* the iteration */
  double min_pivot;          /* smallest squared diagonal entry of the Cholesky factors */
  double precond_error;      /* |M (A + tau B) z - z| / |z| for a fixed z: the factorisation's accuracy */
  int32_t bad_block;         /* first block that is not positive definite, -1 for none */
} fem2d_eigen_stats;
int fem2d_ctx_last_eigen_stats(fem2d_ctx* ctx, fem2d_eigen_stats* out);
int fem2d_gep_device_solve_sparse(fem2d_ctx* ctx, const fem2d_gep_device* gep, double target_eigenvalue,
                                  fem2d_eigenpair* out);

/* ---------------------------------------------------------------------------
 * Field evaluation: UniformFieldSpace::xy_fields (fields.rs:63-127)
 * ------------------------------------------------------------------------- */

/* The x and y fields of a solution vector (one value per DoF) on a
 * density x density grid of points uniform_range(-1, 1, density)
 * (fields.rs:402-405) over each leaf Elem (an Elem without children), leaves
 * in ascending id: out_x / out_y[(k * density + m) * density + n] for leaf k,
 * u point m and v point n. The value at a point is the sum, over the leaf and
 * its ancestors, of every basis spec's f_u or f_v (basis.rs:225-233) times its
 * DoF's value. Errors, in the reference's order:
 * FEM2D_ERR_MISMATCHED_SOLUTION_SIZE if n_solution != the Domain's DoFs;
 * FEM2D_ERR_PANIC if density_u != density_v (the reference sizes its grids
 * [density_v][density_u] and indexes them [m][n], which fails unless they are
 * equal). Densities above 4096 are rejected (FEM2D_ERR_INVALID_ARG).
 * The device variant takes device arrays (solution and outputs of
 * num_leaves * density^2 values); the host variant uploads the view and the
 * solution, and downloads the fields. */
int fem2d_device_domain_num_leaves(const fem2d_device_domain* dd, int64_t* out);
/* Device memory on the context's device for callers without an allocator of
 * their own (hipMalloc / hipFree / synchronous copies). */
int fem2d_device_malloc(fem2d_ctx* ctx, int64_t bytes, void** out);
int fem2d_device_free(fem2d_ctx* ctx, void* p);
int fem2d_memcpy_to_device(fem2d_ctx* ctx, void* dst, const void* src, int64_t bytes);
int fem2d_memcpy_to_host(fem2d_ctx* ctx, void* dst, const void* src, int64_t bytes);
int fem2d_device_xy_fields(fem2d_ctx* ctx, const fem2d_device_domain* dd, int basis, int density_u, int density_v,
                           const double* solution, int64_t n_solution, double* out_x, double* out_y);
int fem2d_xy_fields(fem2d_ctx* ctx, const fem2d_domain_view* domain, int basis, int density_u, int density_v,
                    const double* solution, int64_t n_solution, double* out_x, double* out_y);

/* Per-stage device time of the last assembly on this context, measured with
 * HIP events recorded on the context's stream (milliseconds). An assembly that
 * replays its recorded HIP graph (the default, fem2d_ctx_set_graphs) is timed
 * end to end only: total_ms; the stage and kernel times are those of
 * kernel-by-kernel assemblies and read 0 otherwise. */
typedef struct fem2d_timing {
  double total_ms;    /* first to last event of the assembly */
  double prep_ms;     /* elem geometry, DoF signatures, row groups, Gram tables */
  double symbolic_ms; /* per-group column sets and row lengths, row_ptr scan */
  double fill_ms;     /* column, A and B values of every stored entry */
  int64_t n_groups;   /* row groups (runs of DoFs whose basis specs share their Elems) */
  double count_kernel_ms; /* the wave-per-group row-length kernel alone */
  double fill_kernel_ms;  /* the wave-per-group entry-writing kernel alone */
  int64_t nnz;
  int64_t n_large;    /* row groups assembled one block per group (large column sets) */
  int64_t fill_chunk; /* CSR entries per work item of the entry fill (default, tuned or forced) */
  int64_t fill_lib_lds; /* where the entry fill read the Gram library: 0 global memory, 1 an LDS copy,
                           2 an LDS copy in 512-thread blocks (default or tuned) */
} fem2d_timing;
int fem2d_ctx_last_timing(fem2d_ctx* ctx, fem2d_timing* out);

/* ---------------------------------------------------------------------------
 * Errors
 * ------------------------------------------------------------------------- */
const char* fem2d_last_error(void);
/* variant of the last HRefError / PRefError (1-based, in the reference's
 * declaration order), 0 otherwise */
int fem2d_last_error_detail(void);

/* ---------------------------------------------------------------------------
 * Host input builder: Mesh (mesh.rs) and Domain (domain.rs)
 * ------------------------------------------------------------------------- */

typedef struct fem2d_domain fem2d_domain;

typedef struct fem2d_elem_info {
  int32_t id;
  int32_t nodes[4];
  int32_t edges[4];
  int32_t element;
  int32_t parent;     /* -1 on the base layer */
  int32_t loc;        /* HRefLoc inside the parent, -1 on the base layer */
  int32_t h_levels[2];
  int32_t orders[2];
  int32_t n_children;
  int32_t children[4];
  int32_t depth;      /* number of ancestors */
} fem2d_elem_info;

typedef struct fem2d_edge_info {
  int32_t id;
  int32_t nodes[2];
  int32_t boundary;
  int32_t dir;        /* ParaDir: 0 U, 1 V */
  double length;
  int32_t parent;
  int32_t children[2]; /* -1 when not refined */
  int32_t child_node;
  int32_t active[2];   /* active Elem pair, -1 when none */
} fem2d_edge_info;

int fem2d_mesh_from_file(const char* path, fem2d_mesh** out);       /* Mesh::from_file (mesh.rs:142) */
int fem2d_mesh_from_json(const char* json_text, fem2d_mesh** out);  /* same parser on a string */
/* Mesh from node coordinates ([2*n_nodes]) and element node ids ([4*n_elements],
 * order SW, SE, NW, NE) with materials ([4*n_elements]: eps re/im, mu re/im);
 * identical to reading the equivalent JSON file. */
int fem2d_mesh_from_parts(int64_t n_nodes, const double* xy, int64_t n_elements, const int32_t* node_ids,
                          const double* materials, fem2d_mesh** out);
int fem2d_mesh_unit(fem2d_mesh** out);   /* Mesh::unit (mesh.rs:59) */
int fem2d_mesh_blank(fem2d_mesh** out);  /* Mesh::blank (mesh.rs:91) */
int fem2d_mesh_clone(const fem2d_mesh* m, fem2d_mesh** out);
void fem2d_mesh_free(fem2d_mesh* m);

int fem2d_mesh_counts(const fem2d_mesh* m, int64_t* n_elems, int64_t* n_nodes, int64_t* n_edges,
                      int64_t* n_elements);
int fem2d_mesh_elem_info(const fem2d_mesh* m, int32_t elem_id, fem2d_elem_info* out);
int fem2d_mesh_edge_info(const fem2d_mesh* m, int32_t edge_id, fem2d_edge_info* out);
int fem2d_mesh_node_info(const fem2d_mesh* m, int32_t node_id, double* x, double* y, int32_t* boundary);
int fem2d_mesh_element_info(const fem2d_mesh* m, int32_t element_id, double* points8, double* materials4);
/* id lists: writes at most cap ids to out, *n receives the full count */
int fem2d_mesh_descendant_elems(const fem2d_mesh* m, int32_t elem_id, int include_start, int32_t* out, int64_t cap,
                                int64_t* n);
int fem2d_mesh_ancestor_elems(const fem2d_mesh* m, int32_t elem_id, int include_start, int32_t* out, int64_t cap,
                              int64_t* n);
int fem2d_mesh_descendant_edges(const fem2d_mesh* m, int32_t edge_id, int include_start, int32_t* out, int64_t cap,
                                int64_t* n);
int fem2d_mesh_max_expansion_orders(const fem2d_mesh* m, int32_t* out2);
int fem2d_mesh_elem_is_h_refineable(const fem2d_mesh* m, int32_t elem_id, int32_t* out);
int fem2d_mesh_elem_p_refinement_window(const fem2d_mesh* m, int32_t elem_id, int32_t* out4);
int fem2d_mesh_elem_parametric_range(const fem2d_mesh* m, int32_t elem_id, int32_t from_ancestor, double* out4);

int fem2d_mesh_global_h_refinement(fem2d_mesh* m, int32_t kind, int32_t ext);
int fem2d_mesh_h_refine_elems(fem2d_mesh* m, const int32_t* ids, int64_t n, int32_t kind, int32_t ext);
int fem2d_mesh_execute_h_refinements(fem2d_mesh* m, const int32_t* ids, const int32_t* kinds, const int32_t* exts,
                                     int64_t n);
/* the same batch on n_threads host threads (0: the machine's cores, at most 16;
   1: the reference's sequential loop, mesh.rs:844-914); batches of at least 256
   Elems run data-parallel with the sequential loop's ids, the same Mesh */
int fem2d_mesh_execute_h_refinements_threads(fem2d_mesh* m, const int32_t* ids, const int32_t* kinds,
                                             const int32_t* exts, int64_t n, int32_t n_threads);
int fem2d_mesh_global_p_refinement(fem2d_mesh* m, int32_t di, int32_t dj);
int fem2d_mesh_p_refine_elems(fem2d_mesh* m, const int32_t* ids, int64_t n, int32_t di, int32_t dj);
int fem2d_mesh_execute_p_refinements(fem2d_mesh* m, const int32_t* ids, const int32_t* dis, const int32_t* djs,
                                     int64_t n);
int fem2d_mesh_set_global_expansion_orders(fem2d_mesh* m, int32_t i, int32_t j);
int fem2d_mesh_set_expansion_orders(fem2d_mesh* m, const int32_t* ids, const int32_t* is, const int32_t* js,
                                    int64_t n);

/* Domain::from_mesh (domain.rs:69-159); the mesh is copied. Built data-parallel
   on the host (up to 16 threads); the same Domain, bit for bit, as the
   sequential fem2d_domain_from_mesh_serial. */
int fem2d_domain_from_mesh(const fem2d_mesh* m, int32_t cc, fem2d_domain** out);
/* fem2d_domain_from_mesh on n_threads host threads (0: the machine's cores, at most 16) */
int fem2d_domain_from_mesh_threads(const fem2d_mesh* m, int32_t cc, int32_t n_threads, fem2d_domain** out);
/* Domain::from_mesh in the reference's sequential order (domain.rs:69-159) */
int fem2d_domain_from_mesh_serial(const fem2d_mesh* m, int32_t cc, fem2d_domain** out);
int fem2d_domain_blank(int32_t cc, fem2d_domain** out);
void fem2d_domain_free(fem2d_domain* d);
int fem2d_domain_num_dofs(const fem2d_domain* d, int64_t* n);
/* borrowed pointer to the Domain's mesh */
int fem2d_domain_mesh(const fem2d_domain* d, const fem2d_mesh** out);
/* flattened view; arrays stay valid while the Domain lives */
int fem2d_domain_get_view(const fem2d_domain* d, fem2d_domain_view* out);
/* DoF -> (elem id, index in that elem's basis_specs) of its 1, 2 or 4 BasisSpecs */
int fem2d_domain_dof_info(const fem2d_domain* d, int64_t dof, int32_t* n_bs, int32_t* elem_ids4, int32_t* elem_idx4);

/* Gauss-Legendre nodes (ascending) and weights on (-1, 1), n >= 1 */
int fem2d_gauss_legendre(int32_t n, double* points, double* weights);

#ifdef __cplusplus
}
#endif

#endif /* FEM2D_H */
