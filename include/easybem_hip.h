/* easybem-hip: C-ABI of the MI355X (gfx950) matrix-free layer-potential
 * apply, its Krylov solvers and the solve server.
 *
 * Plain C types only: host/device pointers, sizes, opaque handles and int
 * status codes (0 = success, < 0 = error; ebh_last_error() explains).
 *
 * Which reference interface each entry point replaces:
 *   ebh_mesh_create  initGPU's mesh upload and createMeshTopology
 *                    (include/BEM3dLaplaceGPU.h:8-10; src/BEM3dLaplaceGPU.cu:830-921;
 *                    src/meshTopology.cpp:196-261)
 *   ebh_op_create    initGPU's quadrature upload (src/BEM3dLaplaceGPU.cu:867-917)
 *                    plus the near-field Sauter-Schwab integrals that
 *                    applyGPULaplace3d1Layer/2Layer recompute on every call
 *                    (src/BEM3dLaplaceGPU.cu:997-1222,1340-1460)
 *   ebh_op_apply     applyGPULaplace3d1Layer (ansatz 0/1) and
 *                    applyGPULaplace3d2Layer (include/BEM3dLaplaceGPU.h:14-16;
 *                    src/BEM3dLaplaceGPU.cu:985-1551); the Kelvin single layer
 *                    of getElasticity3d (src/BEM3dElasticity.cpp:838-846)
 *                    applied matrix-free
 *   ebh_op_destroy / ebh_mesh_destroy   closeGPU (src/BEM3dLaplaceGPU.cu:924-957)
 *   ebh_mass_apply   getLaplace3dMass (src/BEM3dLaplace.cpp:448-486), P0 x P1,
 *                    and the P1 x P1 mass matrix
 *   ebh_solve*       the scipy Krylov loop of
 *                    python/examples/laplace3dDirichletGPU.py:60-83
 *                    (GMRES instead of CG; the reference has no solver)
 *   ebh_server_*     no reference counterpart (solve service)
 */
#ifndef EASYBEM_HIP_H
#define EASYBEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  EBH_OK = 0,
  EBH_ERR_ARG = -1,         /* invalid argument or size                 */
  EBH_ERR_MESH = -2,        /* mesh rejected (open, degenerate, ...)    */
  EBH_ERR_HIP = -3,         /* HIP runtime error                        */
  EBH_ERR_UNSUPPORTED = -4, /* operator/basis/rule combination          */
  EBH_ERR_NOMEM = -5,       /* device allocation failed / over budget   */
  EBH_ERR_NOCONV = -6,      /* solver hit its iteration limit           */
  EBH_ERR_STATE = -7,       /* handle in the wrong state                */
  EBH_ERR_NONFINITE = -8    /* NaN/Inf in an input vector               */
};

/* EBH_LAPLACE_HS: the hypersingular operator on P1 through surface curls,
 * D = sum_c C_c^T V C_c (+ stab m m^T), as getLaplace3d's dense D
 * (src/BEM3dLaplace.cpp:652,769) and the stub applyGPULaplace3dHyperSing
 * (src/BEM3dLaplaceGPU.cu:1554-1555); test and trial basis must be P1. */
/* EBH_LAPLACE_ADL: the adjoint double layer, kernel (y - x).n_x / (4 pi
 * r^3), whose Galerkin matrix is the transpose of the double layer's
 * (getLaplace3d's K, src/BEM3dLaplace.cpp:386-437,647); it gives the
 * Neumann right-hand side (M^T/2 - K^T) t.  P1 test, P0 or P1 trial; in
 * FP32 on the pair-symmetric schedule only (far rules of at most 7 points,
 * max_rhs 1), whose hi/lo centroid differences keep (y-x).n accurate. */
/* EBH_KELVIN_DL, EBH_KELVIN_HS: the elasticity double layer (P0 test x P1
 * trial, 3 components each) and hypersingular operator (P1 x P1) of
 * getElasticity3d (src/BEM3dElasticity.cpp:760-943), applied through the
 * Laplace single and double layer and the Kelvin single layer (E = 1,
 * Poisson ratio nu). */
/* EBH_SPARSE_INVERSE: the kind of the operators ebh_precond_spai returns
 * (a sparse matrix applied on the device); ebh_op_create does not take it. */
enum { EBH_LAPLACE_SL = 0, EBH_LAPLACE_DL = 1, EBH_KELVIN_SL = 2,
       EBH_LAPLACE_HS = 3, EBH_LAPLACE_ADL = 4, EBH_KELVIN_DL = 5,
       EBH_KELVIN_HS = 6, EBH_SPARSE_INVERSE = 7 };
enum { EBH_P0 = 0, EBH_P1 = 1 };
enum { EBH_FP64 = 0, EBH_FP32 = 1 };

typedef struct ebh_mesh ebh_mesh;
typedef struct ebh_op ebh_op;
typedef struct ebh_solver ebh_solver;
typedef struct ebh_server ebh_server;

/* Operator descriptor.  line_rule / tri_rule are indices into the
 * reference's Gauss-Legendre and triangle-rule tables
 * (src/quadrature.cpp:286-287,1047-1048); the reference's single `order`
 * argument corresponds to line_rule == tri_rule == order. */
typedef struct {
  int kernel;        /* EBH_LAPLACE_SL/_DL/_HS/_ADL, EBH_KELVIN_SL/_DL/_HS */
  int test_basis;    /* EBH_P0 (one DOF per triangle) or EBH_P1 (nodes)   */
  int trial_basis;   /* DL requires EBH_P1; Kelvin requires test==trial   */
  int far_precision; /* EBH_FP64 or EBH_FP32 (near field is always FP64) */
  int line_rule;     /* near-field Gauss index, 0..18                     */
  int tri_rule;      /* far-field triangle rule index, 0..19 (1 to 79
                        points; 16+ points: one lane per target point)   */
  double nu;         /* Poisson ratio for Kelvin (E = 1)                  */
  int max_rhs;       /* largest nrhs passed to ebh_op_apply (>= 1)        */
  double stab;       /* EBH_LAPLACE_HS: adds stab m (m.x), m_n the lumped
                        node areas (laplace3dNeumannGPU.py:36-49); else 0  */
} ebh_op_desc;

const char *ebh_last_error(void);
int ebh_version(int *major, int *minor);

/* ---- device memory helpers (so callers need no HIP headers) ---------- */
int ebh_malloc(void **ptr, size_t bytes);
int ebh_free(void *ptr);
int ebh_memcpy_h2d(void *dst, const void *src, size_t bytes);
int ebh_memcpy_d2h(void *dst, const void *src, size_t bytes);
int ebh_memset(void *dst, int value, size_t bytes);
int ebh_synchronize(void);
int ebh_device_info(int *n_devices, int *n_cu, size_t *free_bytes,
                    size_t *total_bytes);

/* ---- mesh ------------------------------------------------------------- */
/* nodes: np x 3 FP64 row-major; elems: ne x 3 int32, 1-based, outward
 * orientation (P2-P1)x(P3-P1).  The mesh must be a closed, consistently
 * oriented 2-manifold; otherwise EBH_ERR_MESH. */
int ebh_mesh_create(int np, const double *nodes, int ne, const int *elems,
                    ebh_mesh **out);
/* info[0..7] = np, ne, coincident, edge-adjacent, vertex-adjacent and
 * disjoint pair counts, max touching triangles per triangle, bytes held. */
int ebh_mesh_info(const ebh_mesh *mesh, int64_t *info);
/* Host-only validation and pair classification (no device needed): fills
 * info[0..6] like ebh_mesh_info and, when nnz is non-NULL, nnz[0..3] = the
 * near-field CSR entries of the P0xP0, P0xP1, P1xP0 and P1xP1 patterns.
 * Returns EBH_ERR_MESH with the reason in ebh_last_error() for a rejected
 * mesh. */
int ebh_mesh_check(int np, const double *nodes, int ne, const int *elems,
                   int64_t *info, int64_t *nnz);
void ebh_mesh_destroy(ebh_mesh *mesh);

/* ---- operator --------------------------------------------------------- */
int ebh_op_create(ebh_mesh *mesh, const ebh_op_desc *desc, ebh_op **out);
/* Per-RHS vector lengths including Kelvin's three components:
 * rows = test DOFs, cols = trial DOFs. */
int ebh_op_dims(const ebh_op *op, int *nrows, int *ncols);
/* y = A x on device memory; x holds nrhs vectors of ncols entries
 * back-to-back, y receives nrhs vectors of nrows entries.  Kelvin vectors
 * are component-blocked (all x components, then y, then z), like the
 * reference's 3ne x 3ne layout.  stream: a hipStream_t or NULL. */
int ebh_op_apply(ebh_op *op, const double *x, double *y, int nrhs,
                 void *stream);
/* Same on host memory (copies in and out, synchronous). */
int ebh_op_apply_host(ebh_op *op, const double *x, double *y, int nrhs);
/* Far-field launch configuration: block threads, source tile (triangles),
 * source chunks.  Zero keeps the current value.  Selects the chunked
 * schedule. */
int ebh_op_set_launch(ebh_op *op, int block, int tile, int chunks);
int ebh_op_get_launch(const ebh_op *op, int *block, int *tile, int *chunks);
/* Far-field schedule.  EBH_SCHED_CHUNKED: every target triangle against
 * source chunks (the shape set by ebh_op_set_launch).  EBH_SCHED_SYMMETRIC:
 * each unordered pair of 64-triangle groups once, one rsqrt per point pair
 * serving both directions; for a median-symmetric far rule (triangle rule
 * indices 0-4) with EBH_LAPLACE_SL or EBH_KELVIN_SL on equal test and trial
 * bases, and EBH_LAPLACE_DL (P1 x P1) or EBH_LAPLACE_ADL (P1 x P0);
 * EBH_ERR_UNSUPPORTED otherwise.  split: the symmetric schedule's step
 * ranges per group pair (workgroups per pair; a power of 2 up to 64, 0 for
 * a default from the mesh size), ignored for EBH_SCHED_CHUNKED.  Both
 * schedules compute the same sums in a different order.  No counterpart in
 * the reference (its far-field loop, src/BEM3dLaplaceGPU.cu:445-488, visits
 * every ordered pair). */
#define EBH_SCHED_CHUNKED 0
#define EBH_SCHED_SYMMETRIC 1
int ebh_op_set_schedule(ebh_op *op, int schedule, int split);
int ebh_op_get_schedule(const ebh_op *op, int *schedule, int *split);
/* How the symmetric schedule's group partial sums reach y.  fused = 0 (the
 * default): a reduction kernel, then the near-field pass gathering the sums;
 * fused = 1: one near-field pass sums each test DOF's partials (which the
 * far kernel then stores at the DOF's node-incidence positions) and adds
 * the near CSR row.  Both give bitwise the same y; on MI355X the separate
 * passes measure as fast or faster (the fused pass saves a launch, but its
 * scattered stores cost the far kernel 0.1-1%).  get returns whether the
 * fused pass runs (requested, symmetric schedule, one component, every node
 * row block within the pass's LDS stage).  No counterpart in the
 * reference. */
int ebh_op_set_reduction(ebh_op *op, int fused);
int ebh_op_get_reduction(const ebh_op *op, int *fused);
/* Bound on the symmetric schedule's partial sums (bytes; 0 restores the
 * default, 4 GiB).  In one pass they take ceil(ne/64) x split x (outputs
 * per triangle) x ne FP64 values per right-hand side, quadratic in the
 * mesh; above the bound the schedule runs in passes of block rows (4
 * groups of 64 target triangles each), whose partial buffers stay within
 * it (one block row at least), each pass followed by a reduction into 8
 * slice sums per output (O(ne) memory).  y is bitwise that of one pass.
 * get: the bound, the passes of the current schedule (0 for the chunked
 * schedule) and the bytes of the partial buffer.  No counterpart in the
 * reference. */
int ebh_op_set_partial_limit(ebh_op *op, int64_t bytes);
int ebh_op_get_partial(const ebh_op *op, int64_t *limit, int *passes, int64_t *bytes);
/* Schedule of a multi-RHS apply (max_rhs > 1) of a P0 single layer with a
 * far rule of at most 7 points.  EBH_MRHS_TARGET: each target triangle
 * against every source triangle, every triangle pair's point sums formed in
 * both orders.  EBH_MRHS_PAIRS: the point sums of each unordered triangle
 * pair formed once, into 64 x 64 tiles of group pairs, then applied to the
 * right-hand sides in both directions as FP64 tile products on the matrix
 * cores; in passes of blocks of 32 tile rows whose tiles and item sums stay
 * within the partial limit (ebh_op_set_partial_limit; one block at least),
 * y bitwise the same for every limit.  Both schedules sum in a fixed order
 * (bitwise repeatable) and agree to rounding.
 * EBH_MRHS_AUTO (the default) takes EBH_MRHS_PAIRS where it applies.  get:
 * the schedule in effect (EBH_MRHS_TARGET for operators without it) and its
 * passes (0 for EBH_MRHS_TARGET).  No counterpart in the reference. */
#define EBH_MRHS_AUTO 0
#define EBH_MRHS_TARGET 1
#define EBH_MRHS_PAIRS 2
int ebh_op_set_mrhs_schedule(ebh_op *op, int schedule);
int ebh_op_get_mrhs_schedule(const ebh_op *op, int *schedule, int *passes);
/* Resident tiles of the EBH_MRHS_PAIRS schedule: when every tile and item
 * sum of the operator fits in max_bytes, the tiles are formed once, here
 * (synchronously, on the op's stream), and kept in device memory; every
 * later apply runs only the tile products and their reduction, in one pass,
 * and y is bitwise that of the apply without them.  The far field is then
 * a stored operator, not a matrix-free one: 8 (FP32) or 16 (FP64) bytes per
 * unordered triangle pair, 13.4 / 26.9 GB for 81920 triangles.  0 turns it
 * off (the default); a limit the tiles do not fit in leaves the passes of
 * the partial limit.  Fails with EBH_ERR_UNSUPPORTED for an operator
 * without that schedule.  get: whether the tiles are resident and their
 * bytes.  No counterpart in the reference, which is matrix-free. */
int ebh_op_set_resident_tiles(ebh_op *op, int64_t max_bytes);
int ebh_op_get_resident_tiles(const ebh_op *op, int *resident, int64_t *bytes);
/* Averages over the last `ebh_op_time` calls: milliseconds per apply of
 * [0] fold, [1] far field, [2] reduce + near field, [3] whole apply;
 * measured with HIP events on the op's stream over `reps` applies of the
 * given device vectors. */
int ebh_op_time(ebh_op *op, const double *x, double *y, int nrhs, int reps,
                double *ms);
/* Far-kernel timing over a region of ordinary applies: enable=1 starts
 * recording HIP events around every far-field launch (up to max_launches);
 * each call first returns the total milliseconds and launch count recorded
 * since the previous call (events on the op's stream), then resets. */
int ebh_op_profile(ebh_op *op, int enable, int max_launches, double *far_ms,
                   int64_t *launches);
/* The near-field pass (near CSR times x plus the far sums gathered to the
 * test DOFs, one kernel per single-RHS apply) over the same region:
 * total milliseconds and launches recorded since profiling was enabled; no
 * reset.  No counterpart in the reference. */
int ebh_op_profile_near(ebh_op *op, double *near_ms, int64_t *launches);
/* Surface curls of the P1 hat functions, getLaplace3dCurl
 * (src/BEM3dLaplace.cpp:489-558): curl[(3i + c)*3 + k] is component c of
 * the curl of triangle i's local hat function k.  elems are 1-based.
 * Host only (no device work); rejects out-of-range and degenerate
 * triangles. */
int ebh_surface_curls(int np, const double *nodes, int ne, const int *elems,
                      double *curl);

/* Dense Galerkin matrices, getLaplace3d (src/BEM3dLaplace.cpp:561-781), on
 * the device at rule index `order` (line and triangle rule): V (ne x ne,
 * row-major [test][trial]), K (ne x np, P0 test x P1 trial), D = sum_c
 * C_c^T V C_c (np x np), M (ne x np).  Host outputs, each may be NULL.
 * Entries equal the matrix-free operators' (same far quadrature data, same
 * near-field values).  Memory: 8 (ne^2 + 2 ne np + np^2) bytes on the
 * device. */
int ebh_assemble_laplace(ebh_mesh *mesh, int order, double *V, double *K,
                         double *D, double *M);

/* Dense elasticity matrices, getElasticity3d (src/BEM3dElasticity.cpp:
 * 604-952), E = 1: V (3ne x 3ne), K (3ne x 3np), D (3np x 3np), M
 * (3ne x 3np), component-blocked (all x rows/columns, then y, then z),
 * row-major host outputs, each may be NULL.  V, K and D are built column
 * by column from the matrix-free operators (EBH_KELVIN_SL/_DL/_HS). */
int ebh_assemble_elasticity(ebh_mesh *mesh, double nu, int order, double *V,
                            double *K, double *D, double *M);

/* ---- off-surface evaluation ------------------------------------------- */
/* The representation formula at points off the surface, evaluateLaplace3d
 * (src/BEM3dLaplace.cpp:784-843):
 *   u(x) = sum_j J_j/(4 pi) sum_k w_k [t_j/|x - y_k|
 *                                      - u(y_k) (x - y_k).n_j/|x - y_k|^3]
 * with t P0 (ne entries), u P1 (np entries) and the triangle rule
 * tri_rule (the reference's `order`).  points: npts x 3, result: npts.
 * Device pointers on `stream` (returns after the stream has finished); the
 * _host variant copies in and out.  No singular quadrature: the points
 * must lie off the surface, as in the reference. */
int ebh_evaluate_laplace(ebh_mesh *mesh, int tri_rule, int far_precision,
                         int npts, const double *points, const double *t,
                         const double *u, double *result, void *stream);
int ebh_evaluate_laplace_host(ebh_mesh *mesh, int tri_rule, int far_precision,
                              int npts, const double *points, const double *t,
                              const double *u, double *result);
/* Allocation-free variant of ebh_evaluate_laplace.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels, with the source records and partial sums
 * in a 64 MiB workspace that the library holds on each GPU until the
 * process ends (set up by ebh_mesh_create); the call allocates nothing and
 * does not synchronise (the result is ready when `stream` is).  Limit: the
 * source records (48 B per far-rule point in FP32, 64 B in FP64; ne x nt
 * points) plus 8 B per evaluation point must fit in 64 MiB, else
 * EBH_ERR_ARG; within that, fewer source chunks are used where the
 * partial sums would not fit (the sum order then differs from
 * ebh_evaluate_laplace's).  Every call on one GPU shares that GPU's
 * workspace, so calls that overlap on one GPU race: issue them on one
 * stream or synchronise between them. */
int ebh_evaluate_laplace_noalloc(ebh_mesh *mesh, int tri_rule, int far_precision,
                                 int npts, const double *points, const double *t,
                                 const double *u, double *result, void *stream);

/* ---- multi-mesh batches ------------------------------------------------ */
/* No counterpart in the reference, which keeps one mesh at a time
 * (initGPU, src/BEM3dLaplaceGPU.cu:830-921).  Operators of one kernel,
 * far-field precision and rule size on different meshes, applied together:
 * per-operator fold kernels, one persistent far-field launch per schedule,
 * then the per-operator reductions (one launch each for all members).
 * Symmetric members: their workgroups are items of one atomic queue.
 * Chunked members: (operator, 256-triangle target block, source chunk)
 * items from 8 XCD-affine atomic queues.  Workgroups take items until the
 * queues are empty.  Creating a batch (at most 64 distinct operators) sets
 * every member's schedule: the symmetric one with one step range where it
 * applies, else 256-triangle target blocks, 64-triangle tiles and one
 * source-chunk size for all chunked members, sized to about 8 items per
 * resident workgroup; each y_k then equals ebh_op_apply(ops[k], x_k) bit
 * for bit.  The operators must not be applied elsewhere while a batch apply
 * runs; a schedule changed later is picked up at the next apply. */
typedef struct ebh_batch ebh_batch;
int ebh_batch_create(ebh_op *const *ops, int nops, ebh_batch **out);
/* x[k], y[k]: device vectors of ops[k]; stream: hipStream_t or NULL (the
 * first operator's stream). */
int ebh_batch_apply(ebh_batch *b, const double *const *x, double *const *y,
                    void *stream);
/* [0] operators, [1] work items, [2] workgroups, [3] LDS bytes */
int ebh_batch_info(const ebh_batch *b, int64_t *info);
void ebh_batch_destroy(ebh_batch *b);

/* ---- launch-shape autotuner ----------------------------------------- */
/* No counterpart in the reference, whose initGPU takes nblocks/nthreads by
 * hand (include/BEM3dLaplaceGPU.h:8-10).  Times whole applies with HIP
 * events on the op's stream over far-field launch shapes -- workgroup size
 * 64/128/256 target triangles, LDS source tile 16/32/64 triangles, 1024 to
 * 32768 workgroups -- by coordinate descent from the current shape, then,
 * where it applies, the symmetric schedule with 1 to 32 step ranges; keeps
 * the fastest and records it in a process-wide table under the operator's
 * key (device, kernel, precision, rule size, test basis, RHS block, mesh
 * size class in half-octaves); operators created later under the same key
 * start from it.  Multi-RHS operators also tune the workgroup count of the
 * multi-RHS kernel.  budget_ms bounds the GPU time spent timing (the
 * starting shape is always timed).  The op must not be used by another
 * thread meanwhile. */
typedef struct {
  int block, tile, chunks, grid;  /* chosen single-RHS shape */
  int mgrid;                      /* chosen multi-RHS workgroup count, 0 if none */
  double ms, ms_default;          /* apply ms: chosen / starting shape */
  double mms, mms_default;        /* multi-RHS apply ms: chosen / starting */
  int tried;                      /* shapes timed */
  int schedule, split;            /* chosen EBH_SCHED_* and symmetric split */
} ebh_tune_result;
int ebh_op_autotune(ebh_op *op, double budget_ms, ebh_tune_result *out);
/* Tuning table as text, one line per key: "key block tile grid mgrid ms
 * sched", sched 0 for the chunked schedule or the split of the symmetric
 * one (a missing last column reads as 0).  Loading merges into the table
 * (later lines win). */
int ebh_tune_save(const char *path);
int ebh_tune_load(const char *path);
int ebh_tune_clear(void);

/* Work counts: [0] far interactions per RHS, [1] near nnz (entries),
 * [2] near rows, [3] near cols. */
int ebh_op_work(const ebh_op *op, int64_t *work);
void ebh_op_destroy(ebh_op *op);

/* ---- mass matrices ---------------------------------------------------- */
/* y = M u with M the P0 x P1 (test_basis 0) or P1 x P1 (test_basis 1) mass
 * matrix; u has np entries.  Host pointers. */
int ebh_mass_apply(const ebh_mesh *mesh, int test_basis, const double *u,
                   double *y);

/* ---- Krylov solver ---------------------------------------------------- */
/* GMRES(restart) with FP64 Arnoldi (classical Gram-Schmidt, twice) and
 * FP64 Givens rotations, on device vectors.  With `mixed` set and an FP32
 * far-field operator a32 given, the solve is mixed-precision iterative
 * refinement: the residual r = b - A64 x is formed with the FP64 operator,
 * each correction solves A32 d = r by GMRES to inner_tol, x += d, until
 * ||b - A64 x|| / ||b|| <= tol. */
typedef struct {
  int restart;      /* GMRES restart length (default 50)                  */
  int max_iters;    /* cap on Arnoldi steps over the whole solve          */
  double tol;       /* target FP64 relative residual                      */
  int mixed;        /* 1: iterative refinement with the FP32 operator     */
  double inner_tol; /* relative tolerance of each inner solve (mixed)     */
  int max_outer;    /* cap on refinement steps (mixed)                    */
} ebh_solve_opts;

typedef struct {
  int iters;        /* Arnoldi steps (operator applies inside GMRES)      */
  int outer;        /* refinement steps (mixed) or restarts               */
  int matvec64;     /* FP64 operator applies                              */
  int matvec32;     /* FP32 operator applies                              */
  double relres;    /* final ||b - A64 x|| / ||b|| (true residual, FP64)  */
  double ms;        /* wall time of the solve on the stream (HIP events)  */
  int converged;
  int precond;      /* right-preconditioner applies                       */
  int graph_steps;  /* Arnoldi (or block) steps replayed from a hipGraph  */
} ebh_solve_stats;

/* a64: FP64 operator (square: rows == cols).  a32: optional FP32-far
 * operator of the same shape (needed when opts->mixed).  Allocates its
 * Krylov basis once; one solver handle serves one solve at a time. */
int ebh_solver_create(ebh_op *a64, ebh_op *a32, int restart, ebh_solver **out);
/* Batched variant: up to max_batch systems with the same operator solved
 * together (one Krylov space per system, one multi-RHS apply per Arnoldi
 * step); max_batch must not exceed the operators' max_rhs. */
int ebh_solver_create_batched(ebh_op *a64, ebh_op *a32, int restart,
                              int max_batch, ebh_solver **out);
/* use_graphs: capture each Arnoldi step as a hipGraph and replay it
 * (default 1); check_every: read convergence flags every k steps (default:
 * every step for meshes of 4096+ triangles, else every 4th).  A solver runs
 * on its FP64 operator's stream.  A capture holds a process-wide gate that
 * the library's own device allocations, frees and uploads also take, so
 * solvers on several host threads (the solve server's workers) capture
 * safely; a capture the runtime invalidates anyway (another component's
 * allocation) falls back to eager launches for that solver. */
int ebh_solver_set(ebh_solver *s, int use_graphs, int check_every);
/* Right preconditioning (no counterpart in the reference): GMRES runs on
 * A P y = b with x = P y, so the residual it minimises and reports stays
 * ||b - A x||.  p64 goes with the FP64 operator, p32 with the FP32 one
 * (the inner solves of the mixed refinement); either may be NULL.  The
 * preconditioners are square operators of the system's size whose max_rhs
 * covers the solver's batch. */
int ebh_solver_set_precond(ebh_solver *s, ebh_op *p64, ebh_op *p32);
/* Calderon preconditioner for the P1 single layer (no counterpart in the
 * reference): P = Q (D + m m^T / (4 |Gamma|)) Q with D the hypersingular
 * operator at the given rules and far-field precision, m the lumped node
 * areas and Q five Chebyshev steps for the consistent P1 mass matrix M,
 * preconditioned by M_L = diag(m) (M_L^-1 M has its eigenvalues in
 * [1/4, 1]), a fixed polynomial, so P is linear.  Since D V = 1/4 - K^2
 * (Calderon), P V is well conditioned independently of the mesh size, and
 * a coarse far rule (tri_rule 0, one point per triangle) suffices: FP64
 * GMRES(50) on the L5 icosphere takes 4 iterations (19 with M_L^-1 in
 * place of Q, 61 without a preconditioner). */
int ebh_precond_calderon(ebh_mesh *mesh, int far_precision, int line_rule,
                         int tri_rule, int max_rhs, ebh_op **out);
/* Near-field sparse approximate inverse of a single layer V (P0 x P0 or
 * P1 x P1, either precision) or of the hypersingular operator D + stab m m^T
 * (EBH_LAPLACE_HS without the Calderon scaling; its entries come from the
 * curls and the wrapped P0 single layer); no counterpart in the reference.
 * For V it beats the Calderon preconditioner; for D it helps on coarse
 * meshes only (D's small eigenvalues belong to smooth modes, which a local
 * inverse cannot reach), and ebh_precond_calderon_hs is the better choice
 * from about 20000 triangles up.
 * Row c of M is row c of the inverse of the operator restricted to the
 * patch of DOFs within `rings` (1..4) adjacency steps of DOF c -- P1 nodes
 * are adjacent when they share a triangle, P0 triangles when they share a
 * vertex -- with the operator's entries on the patch computed exactly
 * (near-field CSR plus far quadrature).  Returns a square
 * EBH_SPARSE_INVERSE operator applying y = M x in FP64, for
 * ebh_solver_set_precond (as both the FP64 and the FP32 preconditioner).
 * Build cost: one quadrature pass over the pairs within 2 rings steps on
 * the device and one dense solve per DOF on host threads. */
int ebh_precond_spai(ebh_op *V, int rings, int max_rhs, ebh_op **out);
/* The dual Calderon preconditioner, for the (stabilised) hypersingular
 * operator on P1: P = Q V Q with V the P1 single layer at the given rules
 * and Q five Chebyshev steps for the consistent P1 mass matrix M,
 * preconditioned by the lumped mass M_L (whose ratio to M has eigenvalues
 * in [1/4, 1]): a fixed polynomial, so P is linear.  Rules 3 and 1 (4-point
 * Gauss near field, 3-point far rule) keep FP64 GMRES(50) on D + m m^T at
 * 6 iterations on the L4 and L5 serving meshes, where M_L^-1 V M_L^-1
 * with a 1-point far rule took 31 and 33 (no counterpart in the
 * reference). */
int ebh_precond_calderon_hs(ebh_mesh *mesh, int far_precision, int line_rule,
                            int tri_rule, ebh_op **out);
/* Block GMRES for batched solves (no counterpart in the reference): with
 * block set, a batched solve of nb >= 2 systems builds one block Krylov
 * space of nb columns per restart cycle -- each step one nb-column apply of
 * A P, block classical Gram-Schmidt twice (rocBLAS GEMMs), a column
 * Gram-Schmidt QR of the new block and the block Hessenberg least-squares
 * problem reduced by Householder reflectors on the host -- instead of nb
 * independent Krylov spaces in lockstep; the restart length counts block
 * steps.  Mixed-precision refinement uses it for its inner solves.  Each
 * step's device work is captured as a hipGraph when graphs are on.  The
 * solver needs max_batch >= 2. */
int ebh_solver_set_block(ebh_solver *s, int block);
/* nb systems: b and x hold nb vectors of rows entries back to back; stats
 * (if non-NULL) receives nb entries. */
int ebh_solver_solve_batched(ebh_solver *s, int nb, const double *b,
                             double *x, const ebh_solve_opts *opts,
                             ebh_solve_stats *stats);
/* b, x: device vectors of length rows; x holds the initial guess on entry.
 * Returns EBH_ERR_NOCONV (with x the last iterate) if the tolerance is not
 * reached within the caps. */
int ebh_solver_solve(ebh_solver *s, const double *b, double *x,
                     const ebh_solve_opts *opts, ebh_solve_stats *stats);
void ebh_solver_destroy(ebh_solver *s);

/* ---- solve server (no reference counterpart) -------------------------- */
/* Interior Dirichlet Laplace solves served as they arrive.  Each request
 * brings a mesh and an exterior point source x0; the server forms
 * u = 1/|x - x0| at the nodes, b = (1/2 M + K) u with the P0-test x
 * P1-trial double layer and mass matrix, and solves V t = b for the P0
 * Neumann trace t -- the formulation of the reference example
 * python/examples/laplace3dDirichletGPU.py:60-81 -- by GMRES with
 * mixed-precision refinement.
 * Continuous batching: up to max_slots requests are on the device at once,
 * each on its own HIP streams, so their work overlaps; a queued request is
 * admitted as soon as a slot is free and its estimated device footprint
 * (mesh, near-field CSR, far-field workspace, Krylov basis) fits in the
 * remaining budget.  Meshes submitted with the same non-zero mesh_key stay
 * resident (with their near-field blocks) and are reused; a request whose
 * key is resident for a different mesh (other sizes, nodes or elements)
 * fails with EBH_ERR_ARG and leaves the resident mesh in place.  Idle
 * resident meshes are evicted, least recently used first, only when that
 * lets the head of the queue in.  Every device allocation the library makes
 * for a request is counted (ebh_request_result.device_peak, and the
 * server's totals in ebh_server_stats). */
typedef struct {
  int max_slots;       /* concurrent requests on the device (default 4)     */
  size_t hbm_budget;   /* bytes the server may hold; 0 = 80% of free HBM     */
  int restart;         /* GMRES restart (default 50)                         */
  double tol;          /* FP64 relative residual target (default 1e-8)       */
  int mixed;           /* 1 = mixed-precision refinement (default), 0 = FP64 */
  int line_rule;       /* near-field rule index (default 4)                  */
  int tri_rule;        /* far-field rule index (default 4)                   */
  int max_iters;       /* per request (default 2000)                         */
  int dirichlet_p1;    /* 1: Dirichlet flux in P1 (np entries); 0: P0 flux
                          (ne entries), as the reference's GPU example       */
  int no_precond;      /* 1: no preconditioners (by default P1 Dirichlet
                          solves use the Calderon preconditioner of V, P0
                          ones its sparse approximate inverse, Neumann
                          solves the dual Calderon preconditioner)           */
  int no_graphs;       /* 1: launch the Arnoldi steps eagerly; 0 (default):
                          capture each step of a resident mesh's solver as a
                          hipGraph once and replay it in later solves        */
  size_t partial_limit; /* far-field partial sums per operator, bytes
                          (ebh_op_set_partial_limit; 0 = 512 MiB): larger
                          meshes run the symmetric schedule in passes       */
  double inner_tol;    /* mixed solves: relative tolerance of the first FP32
                          pass; below 0.1 tol the first pass solves to it,
                          so that one FP64 residual can certify the solve
                          (0: 5e-9 for Dirichlet, 1e-5 for Neumann)         */
} ebh_server_opts;

typedef struct {
  int status;          /* EBH_OK or the request's error code                 */
  int iters;           /* Arnoldi steps                                      */
  double relres;       /* final FP64 relative residual                       */
  double queue_ms;     /* submit -> admitted                                 */
  double setup_ms;     /* mesh / operators / right-hand side                 */
  double solve_ms;     /* the Krylov solve                                   */
  double total_ms;     /* submit -> done                                     */
  int ne;              /* unknowns: triangles (Dirichlet) or nodes (Neumann) */
  size_t footprint;    /* estimated device bytes while admitted              */
  int cache_hit;       /* 1 when the mesh was already resident               */
  size_t device_peak;  /* measured: peak device bytes the library held for
                          this request (its set-up and solve allocations; a
                          resident mesh it reuses is counted by the request
                          that built it)                                     */
  int graph_steps;     /* solver steps replayed from captured hipGraphs      */
} ebh_request_result;

int ebh_server_create(const ebh_server_opts *opts, ebh_server **out);
/* Boundary value problems for u = 1/|x - x0| inside the mesh (x0 outside):
 * EBH_DIRICHLET solves V t = (M/2 + K) u for the P0 flux t
 * (python/examples/laplace3dDirichletGPU.py:60-83); EBH_NEUMANN solves
 * (D + m m^T) u = (M^T/2 - K'^T) t for the P1 trace u with the stabilised
 * hypersingular operator (python/examples/laplace3dNeumannGPU.py:36-66). */
enum { EBH_DIRICHLET = 0, EBH_NEUMANN = 1 };
/* Copies the inputs and returns at once with a request id (Dirichlet). */
int ebh_server_submit(ebh_server *s, int np, const double *nodes, int ne,
                      const int *elems, const double *x0, uint64_t mesh_key,
                      int64_t *request_id);
/* Same for either problem; resident meshes are cached per (mesh_key,
 * problem). */
int ebh_server_submit_bvp(ebh_server *s, int problem, int np, const double *nodes,
                          int ne, const int *elems, const double *x0,
                          uint64_t mesh_key, int64_t *request_id);
/* Blocks until the request finishes (timeout_s <= 0: no limit); t (may be
 * NULL) receives the unknown trace: the Dirichlet flux (ne entries, or np
 * with dirichlet_p1) or the Neumann trace (np entries).  Each id can be
 * waited on once. */
int ebh_server_wait(ebh_server *s, int64_t request_id, double timeout_s,
                    ebh_request_result *res, double *t);
/* stats[0..9]: [0] submitted, [1] completed, [2] failed, [3] active,
 * [4] queued, [5] estimated bytes admitted now, [6] peak estimated bytes
 * admitted, [7] cache hits, [8] measured device bytes held now (requests
 * and resident meshes), [9] measured peak device bytes. */
int ebh_server_stats(ebh_server *s, int64_t *stats);
/* Finishes queued work, then stops the workers. */
void ebh_server_destroy(ebh_server *s);

#ifdef __cplusplus
}
#endif
#endif
