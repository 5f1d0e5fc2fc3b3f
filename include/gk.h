/*
 * gk.h -- C ABI of libgkhip.so, the MI355X (gfx950) implicit co-kriging engine.
 *
 * The reference (gempy) has no C ABI: every numeric step of gp.compute_model is
 * a Python call into the external engine (gempy/API/compute_API.py:68-73).
 * Each entry point below replaces one stage of that engine, as cited, and is
 * reached from the package's `hip` backend (gempy_amd/_lib.py) through ctypes.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers owned by the caller, row-major,
 *    unless stated otherwise; the library never frees caller memory.
 *  - Calls are stream-ordered on `stream` (a hipStream_t; NULL = default).
 *    Calls that return scalars to the host (info, iteration counts) synchronise
 *    that stream before returning.
 *  - Every function returns GK_OK (0) or a negative GK_ERR_* code; the message
 *    of the last error on the calling thread is in gk_last_error().
 *  - A handle owns a growable device workspace; one handle must not be used by
 *    two host threads at once.
 */
#ifndef GK_H_
#define GK_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gk_context* gk_handle;
typedef void* gk_stream;

enum {
  GK_OK = 0,
  GK_ERR_ARG = -1,
  GK_ERR_HIP = -2,
  GK_ERR_NOT_SPD = -3,
  GK_ERR_SINGULAR = -4,
  GK_ERR_NOT_CONVERGED = -5,
  GK_ERR_UNSUPPORTED = -6
};

enum { GK_F32 = 0, GK_F64 = 1 };
enum { GK_POINTS_EXPLICIT = 0, GK_POINTS_LATTICE = 1 };
enum { GK_SOLVE_FP64_LU = 0, GK_SOLVE_FP32_REFINE = 1, GK_SOLVE_FP64_CHOL = 2 };
enum { GK_REL_ERODE = 1, GK_REL_ONLAP = 2 };   /* stack relations (values of the reference's StackRelationType) */
#define GK_MAX_THRESHOLDS 64
#define GK_SPAI_K 16          /* neighbours per data point of the learned preconditioner */
#define GK_SPAI_FEATURES 8    /* pair features fed to its network */
#define GK_SPAI_HIDDEN 32     /* width of its two hidden layers */
#define GK_BLOCK 256   /* factorization block size; lda of factored matrices is a multiple */
#define GK_CHOL_PLANES 5     /* planes of GK_BLOCK-wide blocks left by gk_cholesky_* */
enum { GK_SPAI_WITH_FACTOR = 1, GK_SPAI_NO_FALLBACK = 2, GK_SPAI_SANDWICH = 4 };   /* gk_solve_spai flags */

/* One stack's co-kriging data in the engine's layout
 * (SurfacePoints / Orientations records built at
 * gempy/modules/data_manipulation/_engine_factory.py:27-37, split into the
 * reference point (first point) and rest points of every surface).
 * Unknown order: [w_G: x comps of all orientations, then y, then z | w_I per
 * rest point | universal drift (0, 3 or 9) | one per fault column]. */
typedef struct {
  int64_t n_ori;
  int64_t n_rest;
  int64_t n_surf;
  int32_t uni_degree;          /* 0, 1 or 2 (kernel_options.uni_degree) */
  int32_t n_fault;             /* fault-drift columns */
  double range;                /* kernel_options.range */
  double c_o;                  /* kernel_options.c_o */
  double i_res;                /* kernel_options.i_res (4) */
  double gi_res;               /* kernel_options.gi_res (2) */
  const double* ori_xyz;       /* [n_ori*3] */
  const double* ori_grad;      /* [n_ori*3] */
  const double* ori_nugget;    /* [n_ori]   nugget_effect_grad */
  const double* rest_xyz;      /* [n_rest*3] */
  const int32_t* rest_surf;    /* [n_rest]  surface index 0..n_surf-1 */
  const double* rest_nugget;   /* [n_rest]  nu_i + nu_ref(i) */
  const double* ref_xyz;       /* [n_surf*3] */
  const double* fault_rest;    /* [n_rest*n_fault] F(p_i)-F(p_ref(i)); NULL if n_fault==0 */
  double center[3];            /* origin of the device's normalised frame (x-center)/range */
} gk_system;

/* Where and what to evaluate (the engine's exported fields: scalar field,
 * its gradient when compute_scalar_gradient (2-layers.approved.txt:174), and
 * the ids block from the surface thresholds). */
typedef struct {
  int32_t kind;                /* GK_POINTS_EXPLICIT or GK_POINTS_LATTICE */
  int32_t precision;           /* GK_F32 or GK_F64 arithmetic of the pair sums */
  int64_t n_points;            /* explicit: rows of xyz; lattice: filled by the library */
  const double* xyz;           /* explicit points [n_points*3] */
  double origin[3];            /* lattice: first cell centre */
  double spacing[3];           /* lattice: cell size */
  int64_t res[3];              /* lattice: nx, ny, nz (z fastest, regular_grid.py:58-71) */
  const double* fault_values;  /* [n_points*n_fault] fault-block values at the points, or NULL */
  double* scalar;              /* out [n_points] (required) */
  double* gradient;            /* out [n_points*3] or NULL */
  int32_t* ids;                /* out [n_points] or NULL */
  int32_t n_thresholds;        /* host-side thresholds (<= GK_MAX_THRESHOLDS) for ids */
  double thresholds[GK_MAX_THRESHOLDS];
} gk_eval_desc;

typedef struct {
  int32_t form;                /* GK_SOLVE_* actually used (after any fallback) */
  int32_t iterations;          /* Krylov iterations of the refinement */
  int32_t fallback;            /* 1 if the iterative form fell back to the FP64 Cholesky form */
  int32_t info;                /* factorization info (0 = ok, k>0 = failed pivot) */
  double relres;               /* final ||b - A x|| / ||b|| (FP64, matrix-free) */
  double t_assemble_ms, t_factor_ms, t_refine_ms;
  double shift;                /* diagonal shift of the FP32-factored block (unit diagonal) */
} gk_solve_info;

/* ---- library ---------------------------------------------------------- */
int gk_version(void);
const char* gk_last_error(void);
int gk_create(int device, gk_handle* out);
void gk_destroy(gk_handle h);
int64_t gk_system_n(const gk_system* sys);
int64_t gk_system_padded(const gk_system* sys);  /* N rounded up to GK_BLOCK */

/* ---- (a) kernel_constructor: assembly -----------------------------------
 * Replaces the engine's covariance/drift assembly behind
 * compute_API.py:68-73 (kernel_options at 2-layers.approved.txt:150-162).
 * dtype GK_F64: full saddle-point matrix N x N into A (lda >= N), rows/cols
 *   N..lda-1 set to the identity (lda a multiple of GK_BLOCK).
 * dscale != NULL (required for GK_F32, optional for GK_F64): covariance
 *   block only (Nc = 3 n_ori + n_rest), symmetrically equilibrated
 *   H = D A_cc D with D = diag(A_cc)^-1/2 written to dscale[Nc]; padding
 *   rows/cols to lda are the identity.
 * rhs (nullable) receives b = [G_x | G_y | G_z | 0 ...] (length N). */
int gk_assemble(gk_handle h, const gk_system* sys, int32_t dtype, void* A, int64_t lda,
                double* rhs, double* dscale, gk_stream stream);

/* Drift block U (Nc x (L + n_fault)) in FP64, row-major. */
int gk_assemble_drift(gk_handle h, const gk_system* sys, double* U, gk_stream stream);

/* y = A x for the full system, matrix-free (the kernel tiles are recomputed;
 * the reference builds A densely).  precision GK_F64 or GK_F32 pair sums. */
int gk_matvec(gk_handle h, const gk_system* sys, int32_t precision, const double* x, double* y,
              gk_stream stream);

/* ---- (b) solver (kernel_solver, 2-layers.approved.txt:158) -------------- */
/* In-place lower Cholesky of an lda x lda FP32 matrix (lda multiple of GK_BLOCK;
 * rows/cols >= n are reset to the identity).  dinv [GK_CHOL_PLANES*lda*GK_BLOCK]
 * receives five planes of GK_BLOCK-square blocks (block k at row k*GK_BLOCK
 * of each plane): 0: the inverses Dinv_k of the diagonal blocks (lower
 * triangular, with zeros stored above the diagonal: the solves read whole
 * blocks); 1, 2: Dinv_k L_{k,k-d} for d = 2, 1; 3, 4: L_{k+d,k} Dinv_k for
 * d = 2, 1 (used by the solves).  info = 0 or the first
 * failing block+1. */
int gk_cholesky_f32(gk_handle h, float* A, int64_t n, int64_t lda, float* dinv, int32_t* info,
                    gk_stream stream);
/* Solve (L L^T) x = b in place with the output of gk_cholesky_f32 (FP32 vector of
 * length lda); one launch per triangle, calls on one handle are stream-ordered. */
int gk_cholesky_solve_f32(gk_handle h, const float* L, const float* dinv, int64_t lda, float* b,
                          gk_stream stream);
/* FP64 counterparts (FP64 MFMA trailing updates). */
int gk_cholesky_f64(gk_handle h, double* A, int64_t n, int64_t lda, double* dinv, int32_t* info,
                    gk_stream stream);
int gk_cholesky_solve_f64(gk_handle h, const double* L, const double* dinv, int64_t lda, double* b,
                          gk_stream stream);
/* In-place LU without pivoting (the co-kriging matrix has an SPD leading
 * block) of an lda x lda FP64 matrix; dinv [2*lda*GK_BLOCK] receives the
 * inverses of the unit-lower and upper diagonal blocks. */
int gk_lu_f64(gk_handle h, double* A, int64_t n, int64_t lda, double* dinv, int32_t* info,
              gk_stream stream);
int gk_lu_solve_f64(gk_handle h, const double* LU, const double* dinv, int64_t lda, double* b,
                    gk_stream stream);
/* C = beta C - A op(B): MFMA GEMM building block (exported for testing).
 * trans_b = 1: B is n x k (C -= A B^T); 0: B is k x n.  lower_only (m == n,
 * SYRK): update the lower triangle and leave the 256-blocks above the diagonal
 * untouched; entries above the diagonal inside a diagonal 256-block may be
 * updated too (callers read the lower triangle only). */
int gk_gemm_f32(gk_handle h, int64_t m, int64_t n, int64_t k, const float* A, int64_t lda,
                const float* B, int64_t ldb, int32_t trans_b, float beta, float* C, int64_t ldc,
                int32_t lower_only, gk_stream stream);
int gk_gemm_f64(gk_handle h, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda,
                const double* B, int64_t ldb, int32_t trans_b, double beta, double* C, int64_t ldc,
                int32_t lower_only, gk_stream stream);

/* Whole solve of one stack.
 *  GK_SOLVE_FP64_LU    : FP64 full-system assembly + blocked LU (the reference's
 *                        dense direct solve, numpy.linalg.solve class).
 *  GK_SOLVE_FP64_CHOL  : FP64 equilibrated Cholesky of the SPD covariance block +
 *                        Schur complement on the drift block + FP64 refinement steps.
 *  GK_SOLVE_FP32_REFINE: FP32 Cholesky of the equilibrated covariance block with a
 *                        diagonal shift (4 u32 sqrt(n), x4 per breakdown), combined
 *                        with the exact local sparse approximate inverse M
 *                        (y = M u; y += P (u - H y)) as the preconditioner of FP64
 *                        right-preconditioned flexible GMRES on the matrix-free FP64
 *                        operator; if the factor breaks down 4 times or the iteration
 *                        stagnates above 1e-6, falls back to GK_SOLVE_FP64_CHOL. */
int gk_solve(gk_handle h, const gk_system* sys, int32_t form, double tol, int32_t max_iter,
             double* x, gk_solve_info* info, gk_stream stream);

/* gk_solve with an initial guess x0 (device, length N; must not alias x): the
 * weights of a previous solve of a related system, as the reference passes the
 * previous solution's weights into the next solve
 * (gempy/modules/data_manipulation/_engine_factory.py:45-48).  The FP32 form
 * starts its FGMRES refinement from x0; the direct FP64 forms do not use it. */
int gk_solve_x0(gk_handle h, const gk_system* sys, int32_t form, double tol, int32_t max_iter,
                const double* x0, double* x, gk_solve_info* info, gk_stream stream);

/* Solve A x = rhs with the factorization kept by the last gk_solve on this
 * handle (FP64_CHOL or FP32_REFINE forms; the same system). */
int gk_solve_rhs(gk_handle h, const gk_system* sys, const double* rhs, double* x, gk_solve_info* info,
                 gk_stream stream);

/* Condition number of the co-kriging matrix in the 2-norm, sigma_max /
 * sigma_min (the reference's kernel_options.compute_condition_number ->
 * kernel_options.condition_number, read at
 * gempy/modules/optimize_nuggets/_ops.py:40 and examples/examples/real/Hecho.py:219).
 * Lanczos with full reorthogonalisation on A (the matrix-free FP64 operator)
 * and on A^-1 (the factorization kept by the last gk_solve of this system,
 * FP64_CHOL or FP32_REFINE), each to a relative change below tol over three
 * steps or max_steps (<= 300) steps.  sigma (nullable) [2]: sigma_max, sigma_min;
 * steps (nullable) [2]: Lanczos steps on A and on A^-1. */
int gk_condition_number(gk_handle h, const gk_system* sys, int32_t max_steps, double tol, double* cond, double* sigma,
                        int32_t* steps, gk_stream stream);

/* ---- differentiation (the reference tapes torch autograd through the
 * engine: compute_API.py:49,66; gradients w.r.t. sp_coords and the
 * orientation gradients, test_ad_I.py:41-51) ---------------------------------
 * Adjoint of  L(S(obs)) with S = field of the solution w of A w = b(G):
 * e[n_obs] = dL/dS at the observation points.  Returns dL/dp for the rest
 * points then the reference points (d_points [(n_rest+n_surf)*3]) and dL/dG
 * (d_grad [n_ori*3]); lam_out (nullable, [N]) the adjoint vector.  Uses the
 * factorization kept by the last gk_solve of `sys`.  With fault columns,
 * obs_fault [n_obs*n_fault] holds the fault-block values at the observations
 * (the block is piecewise constant: it enters the adjoint right-hand side
 * only). */
int gk_adjoint(gk_handle h, const gk_system* sys, const double* w, int64_t n_obs, const double* obs_xyz,
               const double* obs_fault, const double* e, double* d_points, double* d_grad, double* lam_out,
               gk_stream stream);

/* gk_adjoint plus dL/d(orientation positions), d_ori_xyz [n_ori*3] (nullable;
 * the reference's AD target dip_positions, test/test_modules/test_ad/
 * test_ad_I.py:48-50): gi_res Hess E(q) w_q - (Hess S_w(q) lam_q +
 * Hess S_lam(q) w_q) / gi_res with E the observation field, S_v the field of
 * weights v and w_q, lam_q the orientation's gradient weights. */
int gk_adjoint_full(gk_handle h, const gk_system* sys, const double* w, int64_t n_obs, const double* obs_xyz,
                    const double* obs_fault, const double* e, double* d_points, double* d_grad, double* d_ori_xyz,
                    double* lam_out, gk_stream stream);

/* ---- (c) evaluation: scalar field, gradient, ids ------------------------ */
int gk_eval(gk_handle h, const gk_system* sys, const double* weights, gk_eval_desc* desc,
            gk_stream stream);
/* gk_eval without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The packed sources and the
 * split partial sums live in a fixed file-scope device array of
 * GK_EVAL_SCRATCH_BYTES per GPU instead of the handle's growable workspace:
 * systems whose packed sources (8 n_surf + 32 (n_rest + n_surf) + 64 n_ori
 * bytes, FP64) exceed it are rejected with GK_ERR_ARG; explicit point sets
 * use fewer source splits when their partials do not fit.  Results equal
 * gk_eval's whenever the same split is used.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (other streams, other threads)
 * race and must be serialised by the caller. */
#define GK_EVAL_SCRATCH_BYTES (32u << 20)
int gk_eval_noalloc(gk_handle h, const gk_system* sys, const double* weights, gk_eval_desc* desc,
                    gk_stream stream);

/* F(x) = sum_s w_s C(|x - s|) and grad F (nullable) at explicit points in FP64,
 * with explicitly given sources (range, c_o, frame centre as in gk_system). */
int gk_eval_sources(gk_handle h, double range, double c_o, const double* center, int64_t n_src,
                    const double* src_xyz, const double* src_w, int64_t n_points, const double* xyz,
                    double* scalar, double* gradient, gk_stream stream);

/* Stack combination (the engine's erosion masking; relations at
 * gempy/core/data/structural_frame.py:243-273): ids [n_stacks][n] of the
 * non-fault stacks, youngest first, with n_surf[s] surfaces each (host array);
 * lith[p] = global element id of the youngest stack whose domain (above its
 * last surface) holds p, basement = sum(n_surf) + 1 otherwise. */
int gk_combine_lith(gk_handle h, int32_t n_stacks, const int32_t* ids, int64_t n, const int32_t* n_surf,
                    int32_t* lith, gk_stream stream);
/* The same with each stack's relation to the older ones (host array,
 * GK_REL_ERODE or GK_REL_ONLAP; StackRelationType of
 * gempy/core/data/structural_frame.py:243-273): an onlapping stack holds a
 * point only above the first surface of the next older stack (whose top is
 * preserved).  masks (nullable, [n_stacks][n] uint8): 1 where that stack holds
 * the point, the oldest stack holding the basement too (the engine's squeezed
 * mask matrix; marching_cubes.py:44-46). */
int gk_combine_stacks(gk_handle h, int32_t n_stacks, const int32_t* ids, int64_t n, const int32_t* n_surf,
                      const int32_t* relation, int32_t* lith, uint8_t* masks, gk_stream stream);

/* Smooth ids of one stack at n points (the engine's differentiable `block`,
 * `values_block`; gempy/API/gp2_gp3_compatibility/gp3_to_gp2_output.py:97,124):
 * block = 1 + sum_k logistic(sigmoid_slope (t_k - S)) for decreasing
 * thresholds (host array), with (S - t_k) when they increase; it tends to the
 * integer ids of gk_eval as the slope grows. */
int gk_block_values(gk_handle h, const double* scalar, int64_t n, int32_t n_thresholds, const double* thresholds,
                    double sigmoid_slope, double* block, gk_stream stream);

/* ---- meshes: isosurface of a lattice field (marching cubes; the reference
 * extracts one mesh per surface at its threshold with skimage's marching cubes,
 * gempy/modules/mesh_extranction/marching_cubes.py:59-101) --------------------
 * field [res0*res1*res2] (z fastest), mask (nullable, same layout: a cube is
 * meshed when its far corner (+1, +1, +1) is set), vertices at origin + (index + t e_axis)
 * * spacing on the lattice edges that cross `level` (one per edge, shared),
 * faces as vertex-index triples wound so that normals point towards
 * increasing values; the mesh is watertight away from the lattice boundary and
 * the mask.  With verts or faces NULL only *n_verts / *n_faces are returned;
 * otherwise the outputs must hold max_verts / max_faces entries.  Synchronises
 * the stream. */
int gk_marching_cubes(gk_handle h, const double* field, const int64_t* res, const uint8_t* mask, double level,
                      const double* origin, const double* spacing, double* verts, int64_t max_verts, int32_t* faces,
                      int64_t max_faces, int64_t* n_verts, int64_t* n_faces, gk_stream stream);

/* ---- finite faults (the reference's FiniteFaultData: implicit function,
 * its transform and pivot; gempy/core/data/__init__.py:20, the ellipsoid of
 * gempy/modules/custom_implicit_functions/ellipsoid_implicit_function.py:21-35).
 * values[i] = (values[i] - 1) f(R^T (x_i - pivot) + pivot) in place, with
 * f(x) = 1 - sigmoid(k1 k2 k3 (sum_d ((x_d - c_d) / (r_d k_d))^2 - 1)): the fault
 * block's drift values fade out outside the ellipsoid.  Points: xyz [n*3], or
 * (xyz NULL) the lattice given as lattice[9] = origin[3], spacing[3], res[3]
 * (z fastest); rotation [9] row-major; all small arrays on the host. */
int gk_finite_fault_values(gk_handle h, int64_t n, const double* xyz, const double* lattice, const double* rotation,
                           const double* pivot, const double* center, const double* radii, const double* k,
                           double* values, gk_stream stream);

/* ---- gravity on a centred grid (geophysics input tz + densities ->
 * Solutions.gravity; gempy/API/grid_API.py:99-108, test_gravity.py:67-89) ----
 * gk_gravity_tz: vertical attraction (unit density, G given: 6.674e-3 for mGal
 * with m and g/cm^3) at the device of n prisms, cell centres relative to the
 * device and half-sizes towards lower / higher coordinates [n*3] each.
 * gk_gravity: out[d] = sum_k tz[k] densities[ids[d*n_kernel + k] - 1] (ids are
 * lith ids 1..n_densities; others contribute 0; densities a host array). */
int gk_gravity_tz(gk_handle h, int64_t n, const double* centres, const double* half_lo, const double* half_hi,
                  double G, double* tz, gk_stream stream);
int gk_gravity(gk_handle h, int64_t n_devices, int64_t n_kernel, const double* tz, const int32_t* ids,
               int32_t n_densities, const double* densities, double* out, gk_stream stream);

/* ---- octree evaluation (the engine's octree levels: _number_octree_levels,
 * fault.approved.txt:210-214; resolution = base * 2^(levels-1),
 * gempy/core/data/grid_modules/regular_grid.py:36-43) ------------------------
 * Probe points of `n` voxels (linear indices (ix*ny+iy)*nz+iz at resolution
 * res[3] of the box extent[6]): the centre and the 8 corners of every voxel
 * (xyz [n*9*3], centre first) or, with centres_only, the centres [n*3]. */
int gk_octree_points(gk_handle h, const double* extent, const int64_t* res, const int64_t* voxels, int64_t n,
                     int32_t centres_only, double* xyz, gk_stream stream);
/* ids [n_fields][n*9] (int32) at the probe points; refine[v] = 1 when any field
 * differs among voxel v's 9 points; children [8*n] receives the 8 children (at
 * resolution 2*res) of every refined voxel in voxel order; *n_children their
 * count (host, the call synchronises the stream). */
int gk_octree_refine(gk_handle h, int32_t n_fields, const int32_t* ids, int64_t n, const int64_t* voxels,
                     const int64_t* res, uint8_t* refine, int64_t* children, int64_t* n_children,
                     gk_stream stream);

/* ---- learned preconditioner for the matrix-free solve (config 4) -------------
 * No counterpart in the reference (its iterative solver is plain CG/KeOps,
 * log_2024-05.md:10-50; warm start at _engine_factory.py:45-48).  Data points
 * = orientations then rest points; each owns a 3-dim unknown block.
 * gk_knn: the GK_SPAI_K nearest data points of every data point (itself first),
 *   idx [n*K] int32, hk [n] distance to the K-th (range units).
 * gk_spai_features: pair features [nq*K*GK_SPAI_FEATURES] of points pts[nq]
 *   (NULL = all points).
 * gk_spai_local_blocks: equilibrated covariance among the unknowns of each
 *   neighbourhood, Hloc [nq][3K][3K] (FP32; dummy slots identity).
 * gk_mlp_forward / gk_mlp_backward: the network (F -> H -> H -> 9, tanh) on n
 *   samples; params [gk_spai_nparams()], acts [n*2H] saved for the backward;
 *   grad receives d(sum dy.y)/dparams.
 * gk_spai_loss: per point ||M_P Hloc - E_P||^2 and dY = dloss/dY.
 * gk_adam: one Adam step (beta 0.9/0.999) with grad scaled by grad_scale.
 * gk_spai_local_inverse: the exact local sparse approximate inverse, Y[p][j] =
 *   the rows of H_loc(p)^-1 belonging to p's unknowns (FP64 inversion), in the
 *   layout the network predicts (what the network is trained towards).
 * gk_spai_apply: y = M u on the covariance unknowns (FP64 vectors, FP32 blocks).
 * gk_solve_spai: right-preconditioned restarted flexible GMRES on the
 *   matrix-free operator (Arnoldi matvec in `krylov_precision`, true residual
 *   in FP64) and the exact Schur complement of the drift.  Without
 *   GK_SPAI_WITH_FACTOR in `flags` the covariance block is preconditioned by M
 *   alone (idx, Y required); with it, by the shifted FP32 Cholesky factor P
 *   and, when idx and Y are given, by the combination y = M u; y += P (u - H y)
 *   (M corrects the local modes the shift leaves), or with GK_SPAI_SANDWICH
 *   by y = P u; y += M (u - H y); y += P (u - H y) (suits an inexact M).  Falls back to GK_SOLVE_FP64_CHOL on
 *   stagnation unless GK_SPAI_NO_FALLBACK is set (then x is the last iterate
 *   and info->relres its residual). */
int gk_spai_nparams(void);
int gk_knn(gk_handle h, const gk_system* sys, int32_t* idx, float* hk, gk_stream stream);
int gk_spai_features(gk_handle h, const gk_system* sys, int64_t nq, const int32_t* pts, const int32_t* idx,
                     const float* hk, float* feat, gk_stream stream);
int gk_spai_local_blocks(gk_handle h, const gk_system* sys, int64_t nq, const int32_t* pts, const int32_t* idx,
                         const double* dscale, float* Hloc, gk_stream stream);
int gk_mlp_forward(gk_handle h, int64_t n, const float* params, const float* x, float* y, float* acts,
                   gk_stream stream);
int gk_mlp_backward(gk_handle h, int64_t n, const float* params, const float* x, const float* acts,
                    const float* dy, float* grad, gk_stream stream);
int gk_spai_loss(gk_handle h, const gk_system* sys, int64_t nq, const int32_t* pts, const int32_t* idx,
                 const float* Hloc, const float* Y, float* dY, float* loss, gk_stream stream);
int gk_adam(gk_handle h, float* params, const float* grad, float* m, float* v, int32_t n, float lr, int32_t step,
            float grad_scale, gk_stream stream);
int gk_spai_apply(gk_handle h, const gk_system* sys, const int32_t* idx, const float* Y, const double* u, double* y,
                  gk_stream stream);
int gk_diag_scale(gk_handle h, const gk_system* sys, double* dscale, gk_stream stream);
int gk_spai_local_inverse(gk_handle h, const gk_system* sys, const int32_t* idx, const double* dscale, float* Y,
                          gk_stream stream);
int gk_solve_spai(gk_handle h, const gk_system* sys, const int32_t* idx, const float* Y, double tol,
                  int32_t max_iter, int32_t krylov_precision, int32_t flags, double* x, gk_solve_info* info,
                  gk_stream stream);

/* ---- measurement -------------------------------------------------------
 * gk_profile(h, 1) starts recording HIP events around every launch of the
 * dense kernels (process-wide, cleared on enable; 0 stops).  gk_profile_read
 * sums the recorded launches of one kind: 0..7 = MFMA GEMM with
 * kind = 4*(FP64) + 2*(B transposed) + (lower tiles only), 8 / 9 = the FP32 /
 * FP64 triangular solves (one record per L^-T L^-1 b), 10 / 11 = the
 * split-fp16 rank-256 update of the FP32 factorization (full / lower tiles,
 * three fp16 MFMA products per product); `work` = algorithmic flops (GEMM:
 * 2mnk, lower: m(m+1)k) or factor bytes read (solves). */
int gk_profile(gk_handle h, int32_t enable);
int gk_profile_read(gk_handle h, int32_t kind, int64_t* launches, double* ms, double* work);

/* ---- testing ------------------------------------------------------------
 * No counterpart in the reference.  Stores `code` (1 = spin limit reached,
 * 2 = ticket out of range) into the error word of the sync-free triangular
 * solves of `h`, as a failing readiness wait would: the next solves skip their
 * waits, and the entry points that check the word (gk_cholesky_solve_*, the
 * gk_solve forms, gk_solve_rhs) return GK_ERR_HIP and reset it.  Needs one
 * triangular solve to have run on the handle. */
int gk_test_trsv_set_error(gk_handle h, int32_t code, gk_stream stream);
/* Testing: the split-fp16 rank-256 update the FP32 factorization uses,
 * C[m x n] -= A B^T with A (m x 256) and B (n x 256) dense FP32 rows (entries
 * within [-16, 16]), m, n multiples of 128; lower: tiles on or below the
 * diagonal only (m == n).  With A2, B2 (same shapes; both or neither) also
 * -= A2 B2^T in the same pass: the rank-512 update of paired panels.  No
 * counterpart in the reference. */
/* Testing: the grid search for the FP32 form's local-inverse neighbourhoods,
 * idx[p * GK_LOCAL_NEIGHBOURS + j] = the j-th nearest data point of data point
 * p (itself first; data points = orientations then rest points), hk[p] = the
 * distance of the last (in units of the range).  No counterpart in the
 * reference. */
#define GK_LOCAL_NEIGHBOURS 32
int gk_test_knn_local(gk_handle h, const gk_system* sys, int32_t* idx, float* hk, gk_stream stream);
/* Testing: the FP32 factorization with split-fp16 updates that the FP32 solve
 * form uses (paired rank-512 trailing updates, blocked panels), with the
 * arguments and outputs of gk_cholesky_f32 (L in A's lower triangle, the
 * chain planes in dinv); A must be equilibrated (unit diagonal).  No
 * counterpart in the reference. */
int gk_test_cholesky_f32_split16(gk_handle h, float* A, int64_t n, int64_t lda, float* dinv, int32_t* info,
                                 gk_stream stream);
int gk_test_syrk_split16(gk_handle h, int64_t m, int64_t n, const float* A, const float* B, const float* A2,
                         const float* B2, float* C, int64_t ldc, int32_t lower, gk_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* GK_H_ */
