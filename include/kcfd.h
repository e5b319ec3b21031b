/*
 * kcfd.h -- C-ABI of libkcfd.so, the MI355X (gfx950) kernels for K-CFD's
 * per-iteration hot path: FVM coefficient assembly, Green-Gauss gradients,
 * MUSCL/TVD deferred correction, Jacobi / ILU(0) preconditioners and
 * BiCGSTAB (FP64, and FP32 inner iterations inside FP64 refinement), for one
 * system or a batch of uneven systems.
 *
 * Conventions
 *  - Every entry point returns int: KCFD_OK (0) or a negative KCFD_E* code;
 *    kcfd_last_error() gives a thread-local message for the last failure.
 *  - Handles are opaque and owned by the library; *_destroy frees them.
 *  - Arrays marked (dev) are device pointers (from kcfd_buf_alloc or any
 *    hipMalloc'd memory on the context's device); arrays marked (host) are
 *    host pointers.  Sizes are element counts.
 *  - All work is issued on the context's stream; entry points that return
 *    results to the host synchronise that stream.
 *  - Calls on one context are serialised by a mutex inside the library.
 *
 * Each entry point cites the reference (senily0/K-CFD) interface it replaces.
 */
#ifndef KCFD_H
#define KCFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KCFD_OK 0
#define KCFD_EARG -1      /* bad argument (null pointer, size, unknown kind) */
#define KCFD_EHIP -2      /* HIP runtime error */
#define KCFD_EOOM -3      /* device allocation failed */
#define KCFD_ESTATE -4    /* object not in a usable state (e.g. no factors) */
#define KCFD_EPIVOT -5    /* ILU(0) zero / missing pivot */
#define KCFD_EBUSY -6     /* admission refused: not enough device memory */

/* preconditioner kinds (core/preconditioner.py:18-76) */
#define KCFD_PC_NONE 0
#define KCFD_PC_JACOBI 1
#define KCFD_PC_ILU0 2

/* precision modes */
#define KCFD_FP64 0       /* everything in FP64 */
#define KCFD_MIXED 1      /* FP32 SpMV/preconditioner/vectors inside FP64 refinement */

/* solve status (scipy bicgstab info convention, iterative.py:157-302) */
#define KCFD_INFO_OK 0
#define KCFD_INFO_RHO_BREAKDOWN -10
#define KCFD_INFO_OMEGA_BREAKDOWN -11
#define KCFD_INFO_PQ_BREAKDOWN -12    /* CG: p.q == 0 (scipy would divide by zero) */

/* Krylov method (kcfd_solve_opts.method) */
#define KCFD_METHOD_BICGSTAB 0  /* scipy bicgstab: solve_linear_system (core/linear_solver.py:58-62) */
#define KCFD_METHOD_CG 1        /* scipy cg: solve_pressure_correction (core/linear_solver.py:88-93);
                                   FP64 only */
#define KCFD_METHOD_GMRES 2     /* scipy gmres, restarted, left-preconditioned:
                                   solve_linear_system(method='gmres') (core/linear_solver.py:64-68);
                                   FP64, kcfd_solve only; maxiter counts restart cycles */

/* boundary-condition kinds (core/fvm_operators.py:178-225) */
#define KCFD_BC_ZERO_GRADIENT 0
#define KCFD_BC_DIRICHLET 1
#define KCFD_BC_NEUMANN 2
#define KCFD_BC_NONE 3          /* unknown type string: no contribution (as the reference) */

/* TVD limiters (core/interpolation.py:17-44) */
#define KCFD_LIM_MINMOD 0
#define KCFD_LIM_SUPERBEE 1
#define KCFD_LIM_VAN_ALBADA 2
#define KCFD_LIM_VAN_LEER 3

typedef struct kcfd_ctx kcfd_ctx;
typedef struct kcfd_mat kcfd_mat;
typedef struct kcfd_mesh kcfd_mesh;
typedef struct kcfd_batch kcfd_batch;

/* ---------------------------------------------------------------- context */

/* Replaces the CuPy device set-up of core/gpu_solver.py:26-52
 * (detect_gpu_backend).  One context = one device + one stream. */
int kcfd_init(int device, kcfd_ctx **out);
int kcfd_destroy(kcfd_ctx *ctx);
const char *kcfd_last_error(void);
/* name (host, len bytes), compute units, total / free device bytes */
int kcfd_device_info(kcfd_ctx *ctx, char *name, int len, int *cus,
                     long long *total_bytes, long long *free_bytes);
int kcfd_sync(kcfd_ctx *ctx);
/* STREAM-copy ceiling: mean time of a device-to-device copy of `bytes`
 * (read + write) over `reps` launches.  Measurement hook with no reference
 * counterpart; calibrates the roofline and the FETCH/WRITE counters. */
int kcfd_stream_copy(kcfd_ctx *ctx, size_t bytes, int reps, double *avg_ms);
/* Read ceiling: mean time to stream `bytes` from HBM (loads only, one small
 * result per thread).  Measurement hook with no reference counterpart: the
 * solver kernels are read-mostly, so this is their practical ceiling. */
int kcfd_stream_read(kcfd_ctx *ctx, size_t bytes, int reps, double *avg_ms);

/* ------------------------------------------------------- device buffers */

/* Replaces the per-call host->device uploads of core/gpu_solver.py:113-115
 * (cp.asarray) and the download at :155 (cp.asnumpy). */
int kcfd_buf_alloc(kcfd_ctx *ctx, size_t bytes, void **dev);
int kcfd_buf_free(kcfd_ctx *ctx, void *dev);
int kcfd_buf_upload(kcfd_ctx *ctx, void *dev, const void *host, size_t bytes);
int kcfd_buf_download(kcfd_ctx *ctx, void *host, const void *dev, size_t bytes);
int kcfd_buf_fill(kcfd_ctx *ctx, double *dev, double value, size_t n);

/* ---------------------------------------------------------------- matrix */

/* Device-resident canonical CSR (int32 indices, sorted columns, explicit
 * zeros kept), with an assembled right-hand side and the diagonal tracker of
 * FVMSystem.  Replaces FVMSystem.to_sparse (core/fvm_operators.py:43-48)
 * and cp_sparse.csr_matrix (core/gpu_solver.py:72,113). (host) structure. */
int kcfd_mat_create(kcfd_ctx *ctx, int n, int nnz, const int *indptr,
                    const int *indices, kcfd_mat **out);
int kcfd_mat_destroy(kcfd_mat *m);
int kcfd_mat_set_values(kcfd_mat *m, const double *vals /*host nnz*/);
int kcfd_mat_get_values(kcfd_mat *m, double *vals /*host nnz*/);
/* device pointers of the value array, the rhs and the diag tracker */
int kcfd_mat_dev_ptrs(kcfd_mat *m, double **vals, double **rhs, double **diag);
/* y = A x (dev); replaces gpu_spmv, core/gpu_solver.py:55-79 */
int kcfd_mat_spmv(kcfd_mat *m, const double *x, double *y);
/* device bytes held by the matrix (structure, values, preconditioner) */
int kcfd_mat_bytes(kcfd_mat *m, long long *bytes);

/* Preconditioner set-up on the device: Jacobi dinv = where(|a_ii|>1e-30,
 * 1/a_ii, 0) (core/preconditioner.py:43-47) or structural ILU(0) on A's
 * pattern, level-scheduled (replaces the spilu of :49-54).  With
 * precision KCFD_MIXED an FP32 copy of values and factors is kept too. */
int kcfd_precond_setup(kcfd_mat *m, int kind, int precision);
/* ILU(0) factors (host nnz) and the number of levels of the L and U solves.
 * The factors are stored on the pattern of P A P^T (see kcfd_ilu0_order):
 * with the natural ordering that is A's own pattern. */
int kcfd_ilu0_get(kcfd_mat *m, double *lu, int *levels_l, int *levels_u);
/* Row ordering of the ILU(0) factorisation (no reference counterpart; the
 * reference's spilu chooses its own COLAMD ordering, core/preconditioner.py:
 * 52).  KCFD_ORDER_NATURAL factors A itself.  KCFD_ORDER_COLOR factors
 * P A P^T, P sorting rows by a first-fit colouring of the graph of A + A^T
 * (rows of one colour are never coupled), so each triangular solve has as
 * many dependency levels as colours (2 on the 5- and 7-point grids).
 * KCFD_ORDER_SEGMENT colours segments instead of rows: runs of up to 4
 * consecutive rows each coupled to the one before (pieces of grid lines on a
 * structured mesh in its natural numbering); groups of 64 segments of one
 * colour are numbered step-major (first rows, then second rows, ...).  The
 * factors keep the couplings inside a segment, which costs the solves up to
 * 4 levels per colour and saves Krylov iterations on grids.  Takes effect at
 * the next kcfd_precond_setup. */
#define KCFD_ORDER_NATURAL 0
#define KCFD_ORDER_COLOR 1
#define KCFD_ORDER_SEGMENT 2
int kcfd_precond_set_order(kcfd_mat *m, int order);
/* The ordering of the current ILU(0) factors: perm (host n, may be null)
 * receives the row of A at each position of P A P^T, ncolors the number of
 * colours (0 for the natural ordering). */
int kcfd_ilu0_order(kcfd_mat *m, int *perm, int *ncolors);
/* Give dst (same structure as src) src's ILU(0) ordering and level schedule,
 * so that its kcfd_precond_setup skips the structural analysis (no
 * reference counterpart: a server solving many systems on one mesh). */
int kcfd_mat_copy_order(kcfd_mat *dst, const kcfd_mat *src);
/* y = M^{-1} x (dev) with the preconditioner set up on m (identity for
 * none): the LinearOperator matvec that create_preconditioner returns
 * (core/preconditioner.py:45-46, :52-53). */
int kcfd_precond_apply(kcfd_mat *m, const double *x, double *y);

/* ---------------------------------------------------------------- solve */

typedef struct kcfd_solve_opts {
    double rtol;        /* stop when ||r|| < max(atol, rtol*||b||) */
    double atol;
    int maxiter;        /* <=0: 10*n, as scipy */
    int precond;        /* KCFD_PC_* */
    int precision;      /* KCFD_FP64 / KCFD_MIXED */
    double inner_rtol;  /* MIXED: reliable-update trigger -- the FP64 residual is
                           recomputed once the FP32 residual has dropped by this
                           factor since the last update (default 0.1) */
    int max_refine;     /* MIXED: cap on reliable updates (default 1000) */
    int check_every;    /* iterations per captured graph between host checks */
    int grid_per_cu;    /* launch shape: persistent workgroups (256 lanes) per CU of
                           the SpMV / vector kernels, 1..32 (0: 8; above 8 the
                           workgroups beyond the resident ones start as others end) */
    int trsv_per_cu;    /* launch shape: workgroups per CU of the ILU(0) solves,
                           1..4 (0: 4; the solve needs its grid resident) */
    int tile_walk;      /* launch shape: how a persistent workgroup walks the row
                           tiles of its XCD's range -- 1 interleaved (tiles j,
                           j+J, ...), 2 one contiguous run per workgroup
                           (0: library default) */
    int method;         /* KCFD_METHOD_* (0: BiCGSTAB) */
    int restart;        /* GMRES: Krylov vectors per cycle, 1..64 (0: 20, scipy's) */
    int rows_per_thread; /* launch shape: rows each thread of the SpMV / vector kernels
                            takes per tile, 1, 2 or 4 (a wave covers 64x that many rows
                            of a tile of 256x that many); 0: 2 */
    int fold_colour0;   /* BiCGSTAB with a two-colour ILU(0) (KCFD_ORDER_COLOR on a
                           bipartite pattern): the colour-0 rows of the U factor hold A's
                           own off-diagonal entries, so the colour-0 half of each
                           preconditioner apply also computes those rows of A M^-1 z and
                           the product kernel streams the colour-1 rows only.  In
                           KCFD_MIXED the folded rows use the FP64 diagonal
                           a_ii + sum_j (a_ij - fl32(a_ij)), which keeps A's row sums as the
                           difference-form FP32 copy does.  0: on in KCFD_MIXED, off in
                           KCFD_FP64 (whose products follow scipy's CSR row sums);
                           1: on; -1: off */
} kcfd_solve_opts;

typedef struct kcfd_solve_result {
    int info;           /* 0 converged, >0 maxiter, -10/-11 breakdown */
    int iters;          /* completed full BiCGSTAB / CG iterations (scipy callback count);
                           GMRES: inner iterations of all cycles */
    int half_step;      /* 1 if it stopped at the half-step test on ||s|| */
    int refinements;    /* MIXED: outer FP64 refinement steps; GMRES: restart cycles */
    double rel_resid;   /* FP64 true relative residual ||b-Ax||/||b|| */
    double bytes;       /* algorithmic HBM bytes of the solve (byte model) */
    double ms;          /* device time of the solve */
} kcfd_solve_result;

/* Replaces solve_linear_system(method='bicgstab') (core/linear_solver.py:
 * 14-62) + scipy bicgstab (iterative.py:157-302) and gpu_bicgstab
 * (core/gpu_solver.py:82-155); with opts->method, scipy cg
 * (solve_pressure_correction, core/linear_solver.py:80-93) and scipy gmres
 * (core/linear_solver.py:64-68).  b, x (dev, n); x holds x0 on entry. */
int kcfd_solve(kcfd_mat *m, const double *b, double *x,
               const kcfd_solve_opts *opts, kcfd_solve_result *res);

/* kcfd_solve without any device allocation: no reference counterpart.  For
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * BiCGSTAB / CG kernels and arguments as kcfd_solve (GMRES is not offered),
 * iterations launched directly, results equal to kcfd_solve's bit for bit;
 * the matrix must be set up (kcfd_precond_setup) beforehand.  The whole
 * workspace lives in one static 64 MiB device array per GPU, held until the
 * process ends: a system whose workspace does not fit (about 1.2M rows in
 * FP64 with Jacobi) is refused with KCFD_EARG.  Calls on one GPU share that
 * array and are serialised by a per-device lock (overlapping calls would
 * race on it); each call rewrites what it uses, and the array keeps the
 * last call's contents. */
int kcfd_solve_noalloc(kcfd_mat *m, const double *b, double *x,
                       const kcfd_solve_opts *opts, kcfd_solve_result *res);

/* Batched: many uneven systems advanced by the same kernels (one launch per
 * phase covers every active system; a hipGraph holds the iterations).
 * No reference counterpart: the reference solves one system per call.
 * bs[i], xs[i] (dev) belong to system i. */
int kcfd_batch_create(kcfd_ctx *ctx, int nsys, kcfd_mat *const *mats,
                      const double *const *bs, double *const *xs,
                      const kcfd_solve_opts *opts, kcfd_batch **out);
int kcfd_batch_run(kcfd_batch *b, kcfd_solve_result *results /*host nsys*/);
int kcfd_batch_destroy(kcfd_batch *b);
/* device bytes of the batch workspace */
int kcfd_batch_bytes(kcfd_batch *b, long long *bytes);

/* Dynamic batches, for serving (no reference counterpart).  Systems attach
 * to free slots and detach between rounds; each kcfd_batch_step initialises
 * the systems attached since the last step, replays one captured graph of
 * opts->check_every iterations over every member, computes the FP64 true
 * residual of the systems that finished, and lists their slots (host array
 * of `capacity` ints).  A system's arithmetic does not depend on the other
 * members, so its result is the same alone and in any mix.  The matrix of
 * an attached system (set up for opts' preconditioner and precision) and its
 * b / x buffers must stay alive until it is detached; one matrix may back
 * several members (one per right-hand side).  kcfd_batch_attach returns
 * KCFD_EBUSY when every slot is taken.  A detached system's vector block
 * stays with the batch and is reused by the next attach that needs a block
 * of the same size (no device allocation or free, each of which would
 * synchronise the device); kcfd_batch_trim frees the blocks kept this way
 * (kcfd_batch_bytes counts them).  The graph launches the segment-ordered
 * ILU(0) apply (KCFD_ORDER_SEGMENT) once the first such member attaches. */
int kcfd_batch_create_dynamic(kcfd_ctx *ctx, int capacity, const kcfd_solve_opts *opts,
                              kcfd_batch **out);
int kcfd_batch_attach(kcfd_batch *b, kcfd_mat *m, const double *bvec, double *x, int *slot);
int kcfd_batch_step(kcfd_batch *b, int *finished, int *nfinished);
int kcfd_batch_detach(kcfd_batch *b, int slot, kcfd_solve_result *res);
int kcfd_batch_trim(kcfd_batch *b);
int kcfd_batch_members(kcfd_batch *b, int *members);
/* the single-system workspace kcfd_solve cached on a matrix */
int kcfd_mat_solver(kcfd_mat *m, kcfd_batch **out);
/* Measurement hooks (no reference counterpart; the reference times whole
 * solves with time.time, core/gpu_solver.py:202-250).  kcfd_batch_profile
 * runs `reps` solver iterations (after one warm-up) launched directly on the
 * context stream, with HIP events in the dispatch packets of one phase's
 * kernels in each (hipExtLaunchKernelGGL: the kernels' own start and end),
 * and returns that phase's mean duration: phase 1 = v = A M p with the fused
 * rt.v reduction, 2 = fused x/r update, 3 = ILU(0) apply to p (BiCGSTAB
 * with a two-colour ordering: its colour-1 launch; CG: both launches;
 * otherwise the sync-free lower triangular solve), 4 = p update, 5 = the
 * colour-0 launch of the two-colour apply to p (BiCGSTAB), 6 = the whole
 * apply to p, one event pair from its first launch to its last
 * (BiCGSTAB).  The solver
 * state is restored afterwards; the vectors (x included) are overwritten.
 * kcfd_batch_kernel_bytes gives that phase's algorithmic HBM bytes per
 * launch (summed over the batch). */
int kcfd_batch_profile(kcfd_batch *b, int phase, int reps, double *avg_ms);
int kcfd_batch_kernel_bytes(kcfd_batch *b, int phase, double *bytes);

/* ------------------------------------------------------ mesh builder */

/* The face arrays of a cell-centred FV mesh, built on the device from its
 * cells: build_fvmesh_from_arrays (mesh/mesh_reader.py:141-242, 2D: the
 * edges k, k+1 of each cell) and _build_fvmesh_3d (:245-323, 3D: each
 * cell's face node lists).  Faces are numbered in the order the reference's
 * cell walk first meets them; the first cell to meet a face owns it and the
 * last other cell is its neighbour (-1 on the boundary); normals point out
 * of the owner.  Geometry: face centre = mean of its nodes, Newell normal
 * and area (3D) or the edge normal and length (2D); cell centre = mean of
 * its nodes; volume by the divergence theorem over the cell's face list (3D,
 * floored at 1e-30) or the shoelace area (2D).
 * Inputs (host): nodes n_nodes x 3 (2D coordinates padded with z = 0);
 * ndim 2 selects the edge path, 3 the face lists; cell_nodes n_cells x cell_width
 * (3D: -1 pads a cell with fewer nodes); 3D only: entry_nodes n_entries x
 * face_width (3 or 4; -1 pads a triangle) listing each cell's faces in
 * order, entry_ptr (n_cells + 1) the offsets of each cell's faces (1..8
 * faces a cell).  Boundary lists: n_bfaces rows of face_width (2 in 2D)
 * node ids, bface_patch[b] in 0..n_patches-1 the patch it names; a face
 * named by several rows takes the last one's patch, an unnamed boundary
 * face gets n_patches (the reference's "default"). */
typedef struct kcfd_mesh_faces kcfd_mesh_faces;
int kcfd_mesh_build(kcfd_ctx *ctx, int ndim, int n_nodes, const double *nodes, int n_cells, int cell_width,
                    const int *cell_nodes, int n_entries, int face_width, const int *entry_nodes,
                    const int *entry_ptr, int n_bfaces, const int *bface_nodes, const int *bface_patch,
                    int n_patches, kcfd_mesh_faces **out);
/* counts of a built mesh (face_width: 2 in 2D) and the device time of the
 * build (uploads included) */
int kcfd_mesh_faces_info(const kcfd_mesh_faces *m, int *n_faces, int *n_internal, int *face_width,
                         double *build_ms);
/* copies into caller arrays (any may be null): owner, neighbour (nf),
 * face_nodes (nf x face_width, first-meeting cell's node order), face_patch
 * (nf: -1 internal), area (nf), normal, centre (nf x 3), cell_centre
 * (n x 3), cell_volume (n) */
int kcfd_mesh_faces_get(const kcfd_mesh_faces *m, int *owner, int *neighbour, int *face_nodes, int *face_patch,
                        double *area, double *normal, double *centre, double *cell_centre, double *cell_volume);
int kcfd_mesh_faces_free(kcfd_mesh_faces *m);
/* The structured configuration meshes generated on the device and built as
 * kcfd_mesh_build does, with no host arrays: ndim 2, the quads of
 * _make_structured_quad_mesh (mesh/mesh_generator.py:13-93; nodes
 * (ox + i dx, oy + j dy) row-major, cells counter-clockwise, sides bottom,
 * top, left, right); ndim 3, the hexes of generate_3d_channel_mesh
 * (mesh/mesh_generator_3d.py:12-159; nodes (i dx, j dy, k dz), cells i
 * fastest, faces z-, z+, y-, y+, x-, x+; sides x_min, x_max, y_min, y_max,
 * z_min, z_max; ox, oy ignored).  side_patch[4 or 6]: the patch index each
 * side's boundary faces are listed under, -1 to list them under none. */
int kcfd_mesh_generate(kcfd_ctx *ctx, int ndim, int nx, int ny, int nz, double ox, double oy, double dx, double dy,
                       double dz, const int *side_patch, int n_patches, kcfd_mesh_faces **out);
/* node and cell counts and the node coordinates (n_nodes x 3) of a built
 * mesh; any pointer may be null */
int kcfd_mesh_faces_nodes(const kcfd_mesh_faces *m, int *n_nodes, int *n_cells, double *nodes);

/* ------------------------------------------------------ mesh + assembly */

/* SoA mesh on the device (replaces the AoS FVMesh consumed by the operators,
 * mesh/mesh_reader.py:13-63).  All arrays (host).  Face-major geometry:
 * normal/fcenter are nf*3, ccenter n*3.  bpatch[f] = patch index of a
 * boundary face (-1 internal), bidx[f] = its position in the concatenated
 * per-patch boundary value arrays.  cf_ptr/cf_face/cf_side: cell->face CSR
 * in face-id order (side +1 owner, -1 neighbour).  The matrix structure is
 * the canonical CSR of diag U {(O,N),(N,O)}; slot_on/slot_no give, per
 * face, the CSR positions of (O,N) and (N,O) (-1 on boundary faces or faces
 * skipped for d_PN < 1e-30) and diag_slot[c] the position of (c,c). */
typedef struct kcfd_mesh_desc {
    int n_cells, n_faces, ndim, n_patches, nnz;
    const int *owner, *neighbour;
    const double *area, *normal, *fcenter, *ccenter, *volume;
    const double *d_pn, *d_pf, *gweight;
    const int *bpatch, *bidx;
    const int *cf_ptr, *cf_face;
    const signed char *cf_side;
    const int *indptr, *indices, *slot_on, *slot_no, *diag_slot;
} kcfd_mesh_desc;

int kcfd_mesh_create(kcfd_ctx *ctx, const kcfd_mesh_desc *desc, kcfd_mesh **out);
int kcfd_mesh_destroy(kcfd_mesh *mesh);
/* a new matrix with the mesh's CSR structure, values / rhs / tracker zero */
int kcfd_mesh_new_system(kcfd_mesh *mesh, kcfd_mat **out);

/* FVMSystem.reset (core/fvm_operators.py:50-56) */
int kcfd_asm_reset(kcfd_mat *m);
/* diffusion_operator (core/fvm_operators.py:59-100); gamma (dev, n) */
int kcfd_asm_diffusion(kcfd_mesh *mesh, kcfd_mat *m, const double *gamma);
/* convection_operator_upwind (core/fvm_operators.py:103-129); F (dev, nf) */
int kcfd_asm_upwind(kcfd_mesh *mesh, kcfd_mat *m, const double *F);
/* apply_boundary_conditions (core/fvm_operators.py:178-225).  gamma (dev,n),
 * F (dev,nf), phib (dev, total boundary faces, concatenated per patch);
 * kinds/values (host, n_patches): KCFD_BC_* and the Neumann gradient. */
int kcfd_asm_boundary(kcfd_mesh *mesh, kcfd_mat *m, const double *gamma,
                      const double *F, const double *phib,
                      const int *kinds, const double *values);
/* temporal_operator (core/fvm_operators.py:132-148); phi_old (dev, n) */
int kcfd_asm_temporal(kcfd_mesh *mesh, kcfd_mat *m, double rho, double dt,
                      const double *phi_old);
/* source_term (core/fvm_operators.py:151-161); S (dev, n) */
int kcfd_asm_source(kcfd_mesh *mesh, kcfd_mat *m, const double *S);
/* linearized_source (core/fvm_operators.py:164-175); Sp, Su (dev, n) */
int kcfd_asm_linearized(kcfd_mesh *mesh, kcfd_mat *m, const double *Sp,
                        const double *Su);
/* diag += kd*V, rhs += ks*V per cell (the implicit drag of
 * models/two_fluid.py:592-596: kd = K, ks = K*u_other); kd, ks (dev, n) */
int kcfd_asm_cell_terms(kcfd_mesh *mesh, kcfd_mat *m, const double *kd,
                        const double *ks);
/* under_relax (core/fvm_operators.py:228-240); phi (dev, n) */
int kcfd_asm_under_relax(kcfd_mat *m, const double *phi, double alpha);
/* SIMPLE pressure-correction coefficients (models/single_phase.py:288-325):
 * df = rho*0.5*(dP_O+dP_N)*A/max(|d.n|, 0.1 d) on internal faces, and
 * rho*dP_O*A/d_Pf on faces of patches with pdirichlet[p] != 0.  dP (dev n),
 * pdirichlet (host, n_patches).  Writes diag, off-diagonals; rhs untouched. */
int kcfd_asm_pressure(kcfd_mesh *mesh, kcfd_mat *m, const double *dP,
                      double rho, const int *pdirichlet);
/* The rest of the SIMPLE pressure-correction step (models/single_phase.py:
 * 274-365) around the solve, on device arrays:
 * kcfd_simple_dp: dP = V / max((aP_0 + aP_1 [+ aP_2]) / ndim, 1e-30) from the
 *   momentum diagonals (:279-283; ap2 ignored in 2D);
 * kcfd_simple_rhs: rhs += the mass imbalance of the face fluxes mf (dev nf):
 *   b_O -= F, b_N += F, boundary b_O -= F (:305-311);
 * kcfd_simple_fix_reference: row 0 := e_0, b_0 = 0, for a system without a
 *   Dirichlet p' patch (:327-343);
 * kcfd_simple_correct: grad p' by Green-Gauss with p' = 0 on Dirichlet
 *   patches (pdirichlet, host, n_patches) and the owner's p' on the other
 *   boundary faces, then U -= dP grad p' (U dev n*ndim row-major) and
 *   p += alpha_p p' (:345-363). */
int kcfd_simple_dp(kcfd_mesh *mesh, const double *ap0, const double *ap1,
                   const double *ap2, double *dP);
int kcfd_simple_rhs(kcfd_mesh *mesh, kcfd_mat *m, const double *mf);
int kcfd_simple_fix_reference(kcfd_mat *m);
int kcfd_simple_correct(kcfd_mesh *mesh, const double *pp, const double *dP,
                        const int *pdirichlet, double alpha_p, double *U,
                        double *p);

/* The two-fluid pressure-correction step, models/two_fluid.py:641-725
 * (_solve_pressure_correction_two_phase) around its solve.  Adds to a reset
 * matrix of the mesh: the face coefficients (rho_l al_f V_f / aPl_f +
 * rho_g ag_f V_f / aPg_f) A / d_PN, the mass imbalance of
 * mf_l al_O + mf_g ag_O per face, and 1e10 on diagonal 0 (:645-690).
 * alpha_*, aP_* (dev, n); mf_* (dev, n_faces). */
int kcfd_twofluid_pressure(kcfd_mesh *mesh, kcfd_mat *m, const double *alpha_l,
                           const double *alpha_g, const double *aP_l, const double *aP_g,
                           double rho_l, double rho_g, const double *mf_l, const double *mf_g);
/* :699-723: g = sum of +-(p'_N - p'_O) n A over internal faces / V;
 * U_k += alpha_p (-alpha_k V / max(aP_k, 1e-20)) g for both phases;
 * p += alpha_p p'.  U_l, U_g (dev, n*ndim row-major), p (dev, n), updated. */
int kcfd_twofluid_correct(kcfd_mesh *mesh, const double *pp, const double *alpha_l,
                          const double *alpha_g, const double *aP_l, const double *aP_g,
                          double alpha_p, double *U_l, double *U_g, double *p);
/* green_gauss_gradient (core/gradient.py:12-48); phi (dev n), phib (dev,
 * boundary faces), grad (dev, n*ndim, row-major) */
int kcfd_grad_gg(kcfd_mesh *mesh, const double *phi, const double *phib,
                 double *grad);
/* least_squares_gradient (core/gradient.py:51-99); arguments as
 * kcfd_grad_gg; a cell whose normal-equation matrix is singular gets 0 */
int kcfd_grad_lsq(kcfd_mesh *mesh, const double *phi, const double *phib,
                  double *grad);
/* Face values, core/interpolation.py:55-161: mode 0 compute_face_values_upwind
 * (upwind_face_value), 1 compute_face_values_central (central_face_value),
 * 2 tvd_face_value for every face.  phi (dev, n), phib (dev, boundary faces in
 * patch order), F (dev, n_faces; modes 0, 2), grad (dev, n*ndim; mode 2),
 * phif (dev, n_faces). */
int kcfd_face_values(kcfd_mesh *mesh, int mode, const double *phi,
                     const double *phib, const double *F, const double *grad,
                     double *phif);
/* The TVD limiters of core/interpolation.py:17-44 (limiter_minmod ...
 * limiter_van_leer, KCFD_LIM_*) on n values: psi = limiter(r), dev arrays. */
int kcfd_limiter(kcfd_ctx *ctx, int kind, const double *r, double *psi, long long n);
/* schiller_naumann_drag (models/closure.py:13-34): cd = C_D(max(Re, 1e-10)),
 * 24/Re (1 + 0.15 Re^0.687) below Re 1000, else 0.44; re, cd (dev, n). */
int kcfd_schiller_naumann(kcfd_ctx *ctx, const double *re, double *cd, long long n);
/* drag_coefficient_implicit (models/closure.py:72-92), the interphase drag
 * coefficient the two-fluid momentum assembly puts on the diagonal
 * (models/two_fluid.py:592-596, then kcfd_asm_cell_terms with kd = K):
 * K = 0.75 C_D(Re) alpha_g rho_l |u_g - u_l| / d_b, Re = rho_l |u_g - u_l|
 * d_b / mu_l, |u_g - u_l| floored at 1e-15.  alpha_g, K (dev, n); u_g, u_l
 * (dev, n*ndim row-major). */
int kcfd_drag_implicit(kcfd_ctx *ctx, int n, int ndim, const double *alpha_g, double rho_l,
                       const double *u_g, const double *u_l, double d_b, double mu_l, double *K);
/* muscl_deferred_correction (core/interpolation.py:227-303); corr (dev n) */
int kcfd_muscl(kcfd_mesh *mesh, const double *phi, const double *F,
               const double *grad, int limiter, double *corr);
/* compute_mass_flux (core/interpolation.py:163-182); U (dev n*ndim),
 * Ub (dev, boundary faces*ndim), F (dev nf) */
int kcfd_mass_flux(kcfd_mesh *mesh, const double *U, const double *Ub,
                   double rho, double *F);
/* device-time of the last kernel sequence issued by an asm/grad call (ms) */
int kcfd_last_kernel_ms(kcfd_ctx *ctx, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* KCFD_H */
