/*
 * ddgx -- MI355X (gfx950) SIPG Poisson/Helmholtz operator and CG solver.
 *
 * Plain C ABI: opaque handles, plain pointers and sizes, int status codes.
 * Device buffers are caller-owned (e.g. torch tensors on the HIP device);
 * the library allocates device memory only inside handles, which are freed
 * by the matching *_destroy call.  Every call that launches work takes a
 * hipStream_t (passed as void*) and is ordered on it.
 *
 * Vector layout: element-major, ``x[k*Np + i]`` -- the reference's flat
 * ``u.T.reshape(-1)`` of an (Np, K) field (src/ddg/dg2d/sipg.py:204, 864).
 * Face data (Neumann values) are ``q[k*3*Nfp + f*Nfp + p]`` -- the
 * reference's (3*Nfp, K) array flattened in F order (sipg.py:211-214).
 *
 * The reference is pure Python/JAX and has no FFI of its own; each entry
 * point below names the reference function it stands in for.  The Python
 * binding that calls them is ddg_amd/_lib.py (ctypes).
 */
#ifndef DDGX_H
#define DDGX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DDGX_OK = 0,
  DDGX_ERR_INVALID = 1,      /* bad argument, shape or handle            */
  DDGX_ERR_UNSUPPORTED = 2,  /* polynomial order / dtype not built       */
  DDGX_ERR_HIP = 3,          /* a HIP runtime call failed                */
  DDGX_ERR_NOMEM = 4,        /* device allocation failed                 */
  DDGX_ERR_BUSY = 5          /* shared buffers held by a concurrent call */
};

enum { DDGX_F64 = 0, DDGX_F32 = 1 };

/* Number of words in one element's geometry record and table length for
 * order N (float64 on the host side). */
#define DDGX_REC_WORDS 16
int ddgx_table_size(int N);

typedef struct ddgx_op_s* ddgx_op;
typedef struct ddgx_pc_s* ddgx_pc;

typedef struct {
  double tol;       /* relative tolerance on ||r|| / ||b||   (IterativeConfig.tol) */
  double atol;      /* absolute tolerance                     (IterativeConfig.atol) */
  int maxiter;      /* iteration cap                          (IterativeConfig.maxiter) */
  int check_every;  /* iterations between host polls of the device stop flag (>=1) */
} ddgx_cg_cfg;

typedef struct {
  int iterations;    /* CG iterations executed                              */
  int converged;     /* 1 if the stopping test was met before maxiter        */
  double residual2;  /* final value tested: <r,z> without PC, <r,r> with PC  */
  double atol2;      /* threshold max(tol^2 <b,b>, atol^2)                   */
} ddgx_cg_info;

/* Last error message of the calling thread (static storage). */
const char* ddgx_last_error(void);

/* Operator handle: one SIPG operator (mesh geometry + BC tags + penalty +
 * optional per-element coefficient) at order N in precision dtype.
 * Replaces the per-call setup of sipg_apply_2d / _bc_face_arrays /
 * _local_mass / _face_1d_mass_matrices (src/ddg/dg2d/sipg.py:78-139,
 * 304-337).  Host inputs (float64 / int32):
 *   tables  ddgx_table_size(N) reference-element constants
 *   rec     K x 16 per-element geometry/penalty record
 *   nbr     K x 4  face codes: (k'<<3)|(f'<<1)|reversed, -1 Dirichlet,
 *                  -2 no coupling (Neumann)
 *   fsj     K x 4  face Jacobian sJ on Neumann faces, 0 elsewhere (used by
 *                  the Neumann load only)
 *   order   K      optional (NULL = identity) processing order of the
 *                  elements; the apply works on tiles of 64 consecutive
 *                  entries, each with at most 64 out-of-tile neighbours, so
 *                  a spatially compact order is required for large meshes
 *                  whose own numbering is not (ddg_amd.operator computes a
 *                  Morton order when needed).  Vectors keep the caller's
 *                  element numbering.                                   */
int ddgx_op_create(int N, int64_t K, int dtype, const double* tables,
                   const double* rec, const int32_t* nbr, const double* fsj,
                   const int32_t* order, ddgx_op* out);
/* General operator for curved (non-affine) elements: the reference's SIPG
 * form with per-node metrics and, with use_cub, the cubature branch
 * (sipg.py:142-301 with cub: volume 190-201, faces 281-299; the per-mesh data
 * of cubature.py:107-221).  Host inputs (float64 / int32), element k's data
 * contiguous, elements in the caller's order:
 *   reftab  Dr, Ds, M (Np x Np, row-major), then m1d (3 x Nfp x Nfp)
 *   geo     K x 5 x Np    rx, sx, ry, sy, J at the nodes
 *   fnorm   K x 3 x Nfp x 2  outward normal (nx, ny) at the face nodes
 *   fpar    K x 3 x 3     per element side: w (0.5 interior, 1 Dirichlet,
 *                         0 Neumann), w tau_scale (N+1)^2 Fscale_0, sJ_0
 *   codes   K x 3         (k'<<3)|(f'<<1)|reversed, -1 Dirichlet, -2 Neumann
 *   use_cub = 1: fmass (K x 3 x Nfp x Nfp face cubature mass of each side),
 *                estiff, emass (K x Np(Np+1)/2, upper triangles row after
 *                row): each element's volume cubature stiffness
 *                sum_q w_q J_q grad L_i . grad L_j and mass sum_q w_q J_q L_i L_j;
 *   use_cub = 0: the nodal quadrature (fmass, estiff, emass unused).
 * ddgx_sipg_apply, ddgx_sipg_rhs, ddgx_pcg_solve (without a preconditioner),
 * ddgx_dot and the solver contexts accept the handle; one apply at a time per
 * handle (its trace buffers are the handle's).  Block-Jacobi, multigrid and
 * the parameter gradient need affine operators (DDGX_ERR_UNSUPPORTED). */
int ddgx_op_create_curved(int N, int64_t K, int dtype, int use_cub, const double* reftab,
                          const double* geo, const double* fnorm, const double* fpar,
                          const int32_t* codes, const double* fmass, const double* estiff,
                          const double* emass, ddgx_op* out);
int ddgx_op_destroy(ddgx_op op);
int ddgx_op_info(ddgx_op op, int* N, int64_t* K, int* dtype);

/* y = (A + alpha*M + shift*I) u  with Dirichlet ghost values u_dir on
 * Dirichlet faces (u_dir may be NULL: homogeneous).  Stands in for
 * sipg_apply_2d (sipg.py:304-337) and the Helmholtz/multigrid applies
 * (sipg.py:589-601, multigrid.py:53-74).  u, y, u_dir: device, op dtype. */
int ddgx_sipg_apply(ddgx_op op, const void* u, void* y, const void* u_dir,
                    double alpha, double shift, void* stream);

/* rhs = J M f - L_D(u_dir) + L_N(q_neu)   (sipg.py:849-864, 398-453).
 * f, u_dir (NULL = none), q_neu (NULL = none), rhs: device, op dtype. */
int ddgx_sipg_rhs(ddgx_op op, const void* f, const void* u_dir,
                  const void* q_neu, void* rhs, void* stream);

/* Block-Jacobi preconditioner of A + alpha*M: per-element diagonal blocks
 * built on the device from a float64 operator, inverted in float64, and
 * stored in ``dtype`` (the precision of the vectors it is applied to).
 * Replaces make_sipg_block_jacobi_preconditioner(matfree=True) /
 * make_helmholtz_block_jacobi_preconditioner (sipg.py:671-805). */
int ddgx_block_jacobi_create(ddgx_op op64, double alpha, int dtype, void* stream,
                             ddgx_pc* out);
/* Copy the (K, Np, Np) diagonal blocks (before inversion) of a block-Jacobi
 * preconditioner to a device buffer. */
int ddgx_block_jacobi_blocks(ddgx_pc pc, void* blocks, void* stream);
/* z = M^{-1} r  (the preconditioner callable: sipg.py:711-713 for
 * block-Jacobi, multigrid.py:246 for a V-cycle). */
int ddgx_pc_apply(ddgx_pc pc, const void* r, void* z, void* stream);
int ddgx_pc_destroy(ddgx_pc pc);

/* Multigrid V-cycle settings (make_sipg_mg_preconditioner keywords,
 * src/ddg/dg2d/multigrid.py:253-264). */
typedef struct {
  int nu_smooth;   /* block-Jacobi sweeps before and after the coarse correction */
  double omega;    /* smoother damping                                           */
  double alpha;    /* every level applies A + alpha*M                            */
  int regularize;  /* add eps*max|diag| to the coarsest matrix                   */
  double eps;
} ddgx_mg_cfg;

/* Multigrid V-cycle preconditioner over n_levels operators, fine to coarse
 * (_make_level / _assemble_vcycle / make_sipg_mg_preconditioner /
 * make_sipg_hpmg_preconditioner, multigrid.py:87-99, 211-367).
 *   ops[l]     level operator in the precision of the cycle (one dtype for
 *              all levels; the preconditioner applies to vectors of it)
 *   ops64[l]   float64 operator of the same level (the smoother blocks and
 *              the coarsest matrix are built from it; may equal ops[l])
 *   transfer_kind[l], l < n_levels-1: 0 = p (same mesh, order N_{l+1} <=
 *              N_l), 1 = h (same order, element 4k+c of level l is child c
 *              of element k of level l+1)
 *   transfer[l] host float64, row-major: p: (Np_l x Np_{l+1}) interpolation
 *              matrix; h: the (4, Np, Np) child interpolation matrices
 * The coarsest level is solved with a dense float64 inverse (at most 8192
 * unknowns).  The level operators must outlive the preconditioner. */
int ddgx_mg_create(int n_levels, const ddgx_op* ops, const ddgx_op* ops64,
                   const int32_t* transfer_kind, const double* const* transfer,
                   const ddgx_mg_cfg* cfg, void* stream, ddgx_pc* out);

/* Preconditioned CG on (A + alpha*M + shift*I) x = b with the stopping rule
 * of jax.scipy.sparse.linalg.cg as called at sipg.py:603-609, 909-915.
 * x holds x0 on entry and the solution on exit.  pc may be NULL; it has the
 * operator's dtype, except that a float32 multigrid may precondition a
 * float64 solve (the V-cycle runs in float32 on a scaled copy of the
 * residual; everything else stays float64 -- an extension of the reference,
 * whose solves are float64 only).
 * Iterations run without host synchronisation; the host polls a device
 * stop flag every cfg->check_every iterations. */
int ddgx_pcg_solve(ddgx_op op, double alpha, double shift, ddgx_pc pc,
                   const void* b, void* x, const ddgx_cg_cfg* cfg,
                   ddgx_cg_info* info, void* stream);

/* ddgx_pcg_solve without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops that must not call hipMalloc, harnesses that forbid it).  Same
 * arguments, results and kernels as ddgx_pcg_solve, but the Krylov state
 * (four vectors of the operator's dtype, two float32 vectors when a float32
 * multigrid preconditions a float64 solve, reduction partials, scalars) lives
 * in one fixed 96 MiB device array per GPU instead of the operator's default
 * workspace: at most about 3 million float64 unknowns (K*Np), 2.4 million
 * with a float32 multigrid, twice as many in float32.  Larger solves are
 * refused with DDGX_ERR_INVALID before any device work.  Every call on a GPU
 * shares that GPU's array, and a multigrid preconditioner's cycle buffers are
 * shared with ddgx_pcg_solve and ddgx_pc_apply: a call made while another
 * solve holds either is refused with DDGX_ERR_BUSY; a ddgx_pc_apply cycle
 * still running on another stream is not seen and races.  No graph capture
 * of CG chunks on this path. */
int ddgx_pcg_solve_noalloc(ddgx_op op, double alpha, double shift, ddgx_pc pc,
                           const void* b, void* x, const ddgx_cg_cfg* cfg,
                           ddgx_cg_info* info, void* stream);

/* Parameter gradient of the SIPG form a(u, lam; u_dir) with respect to the
 * per-element diffusion coefficient kappa_k and penalty scale tau_scale_k
 * (sipg.py:263 with a (K,) tau_scale; kappa is this library's extension,
 * kappa = 1 being the reference operator):
 *   g_kappa[k] = d a / d kappa_k,   g_tau[k] = d a / d tau_scale_k.
 * With lam the adjoint solution (A lam = dL/du) the loss gradient is
 * -g (the adjoint of the reference's jax.grad through
 * lax.custom_linear_solve, sipg.py:604-609, 910-915).  kappa and tau_scale
 * must be the values the float64 operator was built with (kappa NULL = 1).
 * All pointers: device float64, caller element order (K or K*Np). */
int ddgx_sipg_param_grad(ddgx_op op, const double* u, const double* lam,
                         const double* u_dir, const double* kappa,
                         const double* tau_scale, double* g_kappa, double* g_tau,
                         void* stream);

/* Mixed-precision iterative refinement (no counterpart in the reference,
 * whose solves are float64 only; it solves the same systems as
 * ddgx_pcg_solve), with a flexible-CG outer loop:
 *   r = b - A x (float64); stop when <r,r> <= max(tol^2 <b,b>, atol^2)
 *   d = PCG_float32(A, r/|r|, pc32) to inner_tol;  z = |r| d
 *   p = z - (<z, A p_old> / <p_old, A p_old>) p_old
 *   x += (<p, r> / <p, A p>) p;  r = b - A x               (all float64)
 * The loop also ends, unconverged, at the float64 residual's round-off
 * floor: when two steps in a row fail to reduce ||r|| by 10%
 * (info->at_floor = 1).
 * op64/op32: the same operator in float64 and float32; pc32: float32
 * preconditioner (block-Jacobi or multigrid) or NULL.  b, x: float64. */
typedef struct {
  double tol, atol;   /* outer stopping rule on the float64 residual          */
  int max_outer;      /* refinement steps cap                                 */
  double inner_tol;   /* relative tolerance of each float32 PCG solve         */
  int inner_maxiter;  /* iteration cap of each float32 PCG solve              */
  int check_every;    /* host polls of the inner solves (as ddgx_cg_cfg)      */
} ddgx_ir_cfg;

typedef struct {
  int outer;          /* refinement steps taken                               */
  int inner_total;    /* float32 PCG iterations over all steps                */
  int converged;      /* outer stopping rule met                              */
  double residual2;   /* final <r,r> of the float64 residual                  */
  double atol2;
  int at_floor;       /* ended at the float64 residual's round-off floor      */
} ddgx_ir_info;

int ddgx_ir_solve(ddgx_op op64, ddgx_op op32, double alpha, double shift, ddgx_pc pc32,
                  const void* b, void* x, const ddgx_ir_cfg* cfg, ddgx_ir_info* info,
                  void* stream);

/* Deterministic dot product in float64 of two device vectors of the op's
 * dtype and length K*Np; result written to *out (host). */
int ddgx_dot(ddgx_op op, const void* a, const void* b, double* out, void* stream);

/* Solver context: private Krylov state (and multigrid cycle buffers) for
 * solves with one operator and preconditioner.  No counterpart in the
 * reference, whose JAX solves are functional and hold no state; it exists so
 * that several solves on the same resident operators and preconditioner can
 * run at once on different streams (the request loop in ddg_amd/serve.py).
 * ddgx_pcg_solve / ddgx_ir_solve / ddgx_dot use one default context per
 * handle instead (and a multigrid's default cycle buffers), held until the
 * call returns: a concurrent call on the same handle is refused with
 * DDGX_ERR_BUSY.  ddgx_pc_apply holds the cycle buffers only while it
 * enqueues the cycle, which runs after it returns, so its device work on one
 * stream still races with a solve or another ddgx_pc_apply on another.
 *   op32 == NULL: a PCG solver on op (pc of op's dtype, a float32 multigrid
 *                 for a float64 op, or NULL);
 *   op32 != NULL: a refinement solver (op float64, op32 its float32 twin, pc
 *                 float32 or NULL).
 * The operators and preconditioner must outlive the solver; one solver runs
 * one solve at a time (a second concurrent call is refused with
 * DDGX_ERR_BUSY).  On a non-NULL stream a solver replays captured
 * chunks of check_every CG iterations as HIP graphs; an odd check_every is
 * raised by one there (a chunk starts on an even iteration). */
typedef struct ddgx_solver_s* ddgx_solver;
int ddgx_solver_create(ddgx_op op, ddgx_op op32, ddgx_pc pc, ddgx_solver* out);
/* Device bytes the solver holds (its share of a request's HBM footprint). */
int ddgx_solver_bytes(ddgx_solver solver, int64_t* bytes);
/* ddgx_pcg_solve on the solver's own state. */
int ddgx_solver_pcg(ddgx_solver solver, double alpha, double shift, const void* b, void* x,
                    const ddgx_cg_cfg* cfg, ddgx_cg_info* info, void* stream);
/* ddgx_ir_solve on the solver's own state. */
int ddgx_solver_ir(ddgx_solver solver, double alpha, double shift, const void* b, void* x,
                   const ddgx_ir_cfg* cfg, ddgx_ir_info* info, void* stream);
int ddgx_solver_destroy(ddgx_solver solver);

/* Batched solves: nseg requests of one problem class (one mesh, order and
 * preconditioner; each request its own right-hand side and Helmholtz
 * coefficient) advanced together, one launch per kernel for all of them.
 * Like the solver contexts this has no counterpart in the reference (each
 * solve_sipg_helmholtz_2d call, sipg.py:536-612, solves one problem); it is
 * the batched launch of the request loop in ddg_amd/serve.py.
 *   ddgx_batch_create: op float64 (affine, a multiple of 64 elements), pc a
 *     multigrid of op (float32 or float64) or NULL.  The batch holds device
 *     copies of both for nseg (1..64) problems; op and pc may be destroyed
 *     afterwards.
 *   ddgx_batch_pcg: alpha[nseg] (host) Helmholtz coefficients, b and x
 *     device float64 vectors of nseg * K * Np, request s at offset s*K*Np; x
 *     holds the initial guesses.  info[nseg] receives each request's CG
 *     information.  Each request follows ddgx_pcg_solve's iteration with the
 *     same stopping rule and the same arithmetic: its x is bitwise the x that
 *     ddgx_pcg_solve (op, alpha[s], the same pc and cfg) returns for it alone,
 *     whatever the other requests of the batch.  On a non-NULL stream chunks
 *     of check_every iterations replay as HIP graphs (check_every raised to
 *     even).  One solve at a time per batch. */
typedef struct ddgx_batch_s* ddgx_batch;
int ddgx_batch_create(ddgx_op op, ddgx_pc pc, int nseg, void* stream, ddgx_batch* out);
int ddgx_batch_bytes(ddgx_batch batch, int64_t* bytes);
int ddgx_batch_pcg(ddgx_batch batch, const double* alpha, const void* b, void* x,
                   const ddgx_cg_cfg* cfg, ddgx_cg_info* info, void* stream);
int ddgx_batch_destroy(ddgx_batch batch);

#ifdef __cplusplus
}
#endif

#endif /* DDGX_H */
