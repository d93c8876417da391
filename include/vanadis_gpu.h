/*
 * vanadis_gpu.h -- C ABI of libvanadis_gpu.so, the MI355X (gfx950) solve path
 * for Vanadis's element-by-element directional-residual (DR) HEX8 operator and
 * its Jacobi-preconditioned BiCG.
 *
 * Plain C types only: host pointers unless a name ends in _dev, sizes as int /
 * int64_t, every call returns an int status (0 = VGPU_OK) except the legacy
 * Fortran entry, which keeps its reference signature.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   kernel_dbcg_simple_  <- src_v2026_2_1/cudatest.cu:301-305 (atomic scatter)
 *                           src_v2026_2_1/cudatest_coloring_gtx1060_optimized.cu:434-438
 *                           (8-colour scatter); called from VANADIS.FOR:809-810
 *   vgpu_create / vgpu_set_coeffs <- the per-element data that crosses into
 *                           el3d (VANADIS.FOR:456-636) via data3d (398-419),
 *                           atmosfera (188-224) and licz_alfa (168-186)
 *   vgpu_apply           <- matvec_0 / mttvec_0 (VANADIS.FOR:2149-2245)
 *   vgpu_assemble_rhs    <- the RHS/diagonal part of przest3d (VANADIS.FOR:723-769,
 *                           799-805)
 *   vgpu_solve           <- dbcg_simple (VANADIS.FOR:977-1055) / kernel_dbcg_simple_
 *   vgpu_mass            <- integrate_mass_hex8 (VANADIS.FOR:2493-2561)
 */
#ifndef VANADIS_GPU_H
#define VANADIS_GPU_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define VGPU_OK 0
#define VGPU_ERR_ARG 1          /* bad argument (null, size, enum)            */
#define VGPU_ERR_HIP 2          /* a HIP runtime call failed                  */
#define VGPU_ERR_MESH 3         /* coords/ISS are not the structured HEX8 grid */
#define VGPU_ERR_STATE 4        /* call order (e.g. solve before set_coeffs)  */
#define VGPU_NOT_CONVERGED 5    /* itmax reached with err > rtol              */
#define VGPU_BREAKDOWN 6        /* BiCG breakdown: a zero/NaN scalar          */
#define VGPU_ERR_NOMEM 7

/* ---- enums ---- */
/* scatter modes of the element operator */
#define VGPU_SCATTER_OWNER 0    /* owner-computes tiles, deterministic, 1 pass */
#define VGPU_SCATTER_COLORED 1  /* 2 colours of z-slabs, deterministic        */
#define VGPU_SCATTER_ATOMIC 2   /* hardware FP64/FP32 atomics on tile faces   */
/* DR variants (gauss_point_hex8DR): V1 VANADIS.FOR:1868-2023,
 * V2 src_updates/gauss_point_hex8DR.for, V3 vanadis_full_DR_OpenMP_timing.for:1567-1716 */
#define VGPU_DR_V1 1
#define VGPU_DR_V2 2
#define VGPU_DR_V3 3
/* element-system schemes: A = cH*H + cC*C */
#define VGPU_SCHEME_THETA 0     /* A = 2H + (3/dt)C   (VANADIS.FOR:743-748)   */
#define VGPU_SCHEME_STEADY 1    /* A = H                                      */
#define VGPU_SCHEME_EULER 2     /* A = H + C/dt                               */
/* precisions */
#define VGPU_FP64 0
#define VGPU_FP32 1             /* FP32 element math, FP64 node accumulation  */
#define VGPU_MIXED 2            /* solve: FP32 inner BiCGSTAB + FP64 refinement */
/* Krylov methods of vgpu_solve (FP64 / FP32 precisions) */
#define VGPU_METHOD_BICG 0      /* Jacobi-BiCG, the reference's dbcg_simple    */
#define VGPU_METHOD_BICGSTAB 1  /* Jacobi-BiCGSTAB (right preconditioned)      */

typedef struct vgpu_ctx vgpu_ctx;

/* Last error message of the calling thread (static storage, never NULL). */
const char *vgpu_last_error(void);
/* Library version string. */
const char *vgpu_version(void);
/* Number of visible HIP devices (0 when none); never initialises a context. */
int vgpu_device_count(int *count);

/* ------------------------------------------------------------------------ *
 * Legacy drop-in (src_v2026_2_1/cudatest.cu:301-305).  Every argument by
 * reference, Fortran default INTEGER / REAL*8.  Semantics of the reference:
 * Jacobi-PBiCG on (I-D)A(I-D) with A the assembled 8x8 element matrices BB
 * (row-major per element, BB[64e+8i+j]), ISS column-major 1-based, stop when
 * ||r||/||b|| <= TOL on the recursive residual; X in/out, zeroed at masked
 * nodes.  Differences, by design: ITER/ERR are also written when x0 already
 * satisfies TOL (ITER=0), a HIP failure sets ERR=+Inf so the host takes
 * its "stopping criterion not satisfied" path, and B = 0 returns the exact
 * X = 0 with ITER=0, ERR=0 (the reference's loop divides 0 by 0 there and
 * leaves NaN in X).  Scatter mode: environment
 * VGPU_LEGACY_SCATTER=atomic|colored (default colored) or
 * vgpu_legacy_set_scatter().  Device buffers persist between calls with the
 * same sizes (one cache per process; not re-entrant, like the reference).
 * ------------------------------------------------------------------------ */
void kernel_dbcg_simple_(int *N, double *B, double *X, double *TOL, int *ITMAX,
                         int *ITER, double *ERR, double *BB, int *ISS,
                         double *DIAG, int *NAX, int *NAY, int *NAZ,
                         int *bc_dirichlet);
int vgpu_legacy_set_scatter(int scatter);
/* Release the legacy path's cached device buffers. */
int vgpu_legacy_release(void);
/* The colouring and per-colour element lists of the colour variant, as the
 * legacy path builds them (generate_hex8_colors / build_color_element_lists,
 * src_v2026_2_1/cudatest_coloring_gtx1060_optimized.cu:50-106): color[E],
 * elems[E] (0-based element ids, stable by colour), offsets[9], counts[8].
 * Host computation; no device work. */
int vgpu_legacy_colors(int nax, int nay, int naz, int *color, int *elems, int *offsets,
                       int *counts);
/* The element lists the legacy path holds on the device after a call
 * (read back), with their offsets and counts; capacity = length of elems
 * (at least E of the last call).  VGPU_ERR_STATE before the first call. */
int vgpu_legacy_get_color_lists(int64_t capacity, int *elems, int *offsets, int *counts);
/* Wall times (seconds) of the last kernel_dbcg_simple_ call: host-to-device
 * uploads of b, x, BB, ISS, DIAG and the mask (with the buffer set-up when
 * the size changed), the BiCG solve on the device (first kernel to the
 * convergence readback), and the download of X.  No counterpart in the
 * reference, whose host times the whole call (VANADIS.FOR:807-812). */
int vgpu_legacy_last_times(double *upload_s, double *solve_s, double *download_s);

/* Allocation-free form of the legacy drop-in.  No counterpart in the
 * reference (whose solver makes 19 device allocations per call,
 * cudatest.cu:408-450): for callers that cannot allocate device memory at
 * solve time (tight time-stepping or serving loops, harnesses that forbid
 * allocation).  Same algorithm, kernels and launch shapes as
 * kernel_dbcg_simple_ (bitwise the same iterates with the colored scatter),
 * but every array is a device pointer (b, x in/out, bb [64 E], iss [8 E]
 * column-major 1-based, diag, bc [N] int), scalars by value, iter / err on
 * the host, scatter VGPU_SCATTER_COLORED or VGPU_SCATTER_ATOMIC, stream a
 * hipStream_t (NULL: the null stream); the call returns when the solve is
 * done.  The BiCG work vectors, Dirichlet flags, colour lists, partials and
 * status live in one fixed device array of the library (about 14 MB per
 * device, held until the process ends): N may not exceed
 * vgpu_noalloc_max_nodes() (2^18; larger N: VGPU_ERR_ARG).  Calls on one
 * device share that array: calls that overlap on one device (other host
 * threads or streams) race.  Returns VGPU_OK (converged; b = 0: x = 0,
 * iter 0, err 0), VGPU_NOT_CONVERGED (itmax), VGPU_BREAKDOWN or
 * VGPU_ERR_ARG / VGPU_ERR_HIP. */
int kernel_dbcg_simple_noalloc(int n, const double *b_dev, double *x_dev, double tol,
                               int itmax, int *iter, double *err, const double *bb_dev,
                               const int *iss_dev, const double *diag_dev, int nax, int nay,
                               int naz, const int *bc_dev, int scatter, void *stream);
int vgpu_noalloc_max_nodes(int64_t *max_nodes);

/* ------------------------------------------------------------------------ *
 * Context API: matrix-free operator, device-resident across time steps.
 * ------------------------------------------------------------------------ */
/* Build a context on the current device from the host's node coordinates
 * xx/yy/zz[nax*nay*naz] (siatka order, VANADIS.FOR:226-244), the
 * connectivity iss[8*E] (w_el layout, VANADIS.FOR:246-276; NULL = implied)
 * and the Dirichlet mask bc[N] in {0,1}.  The coordinates must form a tensor
 * grid and iss must equal the structured connectivity (VGPU_ERR_MESH). */
int vgpu_create(int nax, int nay, int naz, const double *xx, const double *yy,
                const double *zz, const int *iss, const int *bc,
                vgpu_ctx **out);
int vgpu_destroy(vgpu_ctx *ctx);
int vgpu_get_sizes(const vgpu_ctx *ctx, int64_t *n_nodes, int64_t *n_elems);

/* Per-element coefficients (atmosfera hook): isotropic diffusivity lam[E],
 * velocity vel[3*E] (SoA: vx[0..E), vy[E..2E), vz[2E..3E)), decay P, Robin
 * alpha on the x-min face (face 3, VANADIS.FOR:579-627; 0 disables), DR
 * variant, scheme and dt.  tau is computed on the device per element. */
int vgpu_set_coeffs(vgpu_ctx *ctx, const double *lam, const double *vel,
                    double P, double alpha, int dr_variant, int scheme,
                    double dt);
/* Change only the scheme / dt (e.g. steady vs implicit Euler). */
int vgpu_set_scheme(vgpu_ctx *ctx, int scheme, double dt);
/* Options: scatter mode for the solver, z chunk and tile shape (-1 = keep).
 *   OWNER:   tiles of TX x TY element columns (tile 0 = 32x8, 1 = 16x16,
 *            2 = 64x4) overlapping by one column, each node written by one
 *            thread; zchunk 0 = the device's resident workgroups split the
 *            node planes evenly, n > 0 = one workgroup per (chunk of n node
 *            planes, tile) group, groups numbered chunk by chunk so that
 *            neighbouring tiles march the same planes together.  Results do
 *            not depend on the z chunk or the tile shape (bitwise).  Needs
 *            8 N and 40 E below 2^31 bytes (else VGPU_ERR_ARG; such systems
 *            default to the coloured scatter).
 *   COLORED: the same tiles over z-slabs of zchunk element layers (0 = the
 *            deepest slab of at most 8 layers whose colour launch still
 *            gives every compute unit two workgroups); slabs coloured by
 *            parity, two launches, each slab's face planes summed in colour
 *            order: deterministic for a fixed choice, <= 1e-13 of the
 *            oracle for every choice.
 *   ATOMIC:  disjoint bricks of the tile shape and zchunk element layers
 *            (0 = the deepest brick of at most 8 layers that still gives
 *            every compute unit two workgroups), brick faces added with
 *            hardware atomics (sum order not reproducible). */
int vgpu_set_options(vgpu_ctx *ctx, int scatter, int zchunk, int tile);
/* Krylov method of the FP64 and FP32 solves (and of the FP64 passes of a
 * mixed solve): VGPU_METHOD_BICG (default; the reference's method,
 * VANADIS.FOR:977-1055) or VGPU_METHOD_BICGSTAB (van der Vorst's BiCGSTAB
 * with the same Jacobi preconditioner and stopping rule, on the owner
 * kernel's launches whatever the scatter mode; needs 8 N and 40 E below
 * 2^31 bytes).  No counterpart in the reference: it adds a method with two
 * single operator applies per iteration and no transposed apply. */
int vgpu_set_method(vgpu_ctx *ctx, int method);

/* y = (I-D) A (I-D) x  (transpose=1: A^T), host vectors of length N.
 * precision VGPU_FP64 or VGPU_FP32 (x,y still FP64 on the host). */
int vgpu_apply(vgpu_ctx *ctx, int scatter, int precision, int transpose,
               const double *x, double *y);
/* Device-pointer variant on the context's stream (x_dev/y_dev FP64 [N]). */
int vgpu_apply_dev(vgpu_ctx *ctx, int scatter, int transpose,
                   const double *x_dev, double *y_dev);
/* Jacobi diagonal diag[N] = sum_e A_e(a,a), 1 on masked nodes
 * (VANADIS.FOR:761-769, 799-805). */
int vgpu_diag(vgpu_ctx *ctx, double *diag);
/* RHS of the element system, masked (VANADIS.FOR:723-741, 756-758, 799-805):
 * theta: b = 3F - (H - (3/dt)C) t_old;  steady: b = F;
 * Euler: b = F + (C/dt) t_old.  q_in[E] per-element source [mass/time];
 * t_old may be NULL (zero). */
int vgpu_assemble_rhs(vgpu_ctx *ctx, const double *q_in, const double *t_old,
                      double *b);
/* Solve A x = b.  x is in/out (warm start, zeroed at masked nodes).
 * precision VGPU_FP64, VGPU_FP32 (FP32 vectors and operator, FP64 scalars),
 * both with the context's method (vgpu_set_method), or VGPU_MIXED (inner
 * FP32 BiCGSTAB to inner_rtol, 0 = 1e-4, on r/|r| with FP64 true-residual
 * refinement x += |r| dx to rtol; FP32 BiCG inside for systems too large for
 * the owner launches; if a refinement step fails to halve the true residual
 * the solve finishes with FP64 BiCGSTAB passes, or the context's method for
 * systems too large for the owner launches).  iters = total
 * Krylov iterations, err = final relative residual (recursive for FP64 and
 * FP32, true for MIXED).  VGPU_MIXED checks itmax between refinement steps,
 * and an inner solve runs at most 1000 iterations, so iters can exceed itmax
 * by up to 1000.  b = 0: x = 0, the exact solution from any start, with
 * iters 0, err 0 and VGPU_OK (the relative residual is undefined there). */
int vgpu_solve(vgpu_ctx *ctx, int precision, const double *b, double *x,
               double rtol, int itmax, double inner_rtol, int *iters,
               double *err);
/* Same with b, x already on the device (FP64 [N]).  seconds (optional): wall
 * time of the solve between two stream synchronisations; NULL skips the timer
 * and the synchronisation before the solve (the solve queues behind the
 * caller's work on the context's stream). */
int vgpu_solve_dev(vgpu_ctx *ctx, int precision, const double *b_dev,
                   double *x_dev, double rtol, int itmax, double inner_rtol,
                   int *iters, double *err, double *seconds);
/* ||true residual||/||b|| for the current coefficients (FP64). */
int vgpu_true_residual(vgpu_ctx *ctx, const double *b, const double *x,
                       double *rel);
/* Mass integral (2x2x2 Gauss, exact for trilinear c on boxes). */
int vgpu_mass(vgpu_ctx *ctx, const double *c, double *mass);
/* Diagonal-dominance diagnostics of the element matrices of the current
 * scheme (diagonal_dominance_factor, VANADIS.FOR:2434-2489, and
 * is_diagonally_dominant, 2248-2303): ddf = max over nodes of
 * sum_e sum_{j != i} |A_e(i,j)| / sum_e |A_e(i,i)|; dominant = 1 when every
 * node's diagonal sum exceeds its off-diagonal sum (may be NULL). */
int vgpu_ddf(vgpu_ctx *ctx, double *ddf, int *dominant);
/* The host's T-file (VANADIS.FOR:835-875): one line per node,
 * x, y, z, c * 1e6 in Fortran e12.5 separated by two blanks.  Host arrays;
 * no device work. */
int vgpu_write_tfile(const char *path, int64_t n, const double *xx, const double *yy,
                     const double *zz, const double *c);

/* Device buffer helpers for callers that keep data resident (bench, batch). */
int vgpu_dev_alloc(vgpu_ctx *ctx, int64_t bytes, void **ptr);
int vgpu_dev_free(vgpu_ctx *ctx, void *ptr);
int vgpu_memcpy_h2d(vgpu_ctx *ctx, void *dst_dev, const void *src, int64_t bytes);
int vgpu_memcpy_d2h(vgpu_ctx *ctx, void *dst, const void *src_dev, int64_t bytes);

/* Timing of the dominant kernel with HIP events on the context's stream:
 * `reps` back-to-back operator launches; avg_ms per launch.
 * what: 0 = single apply (A x), 1 = dual apply (A p and A^T pp),
 *       2 = device copy of N doubles (bandwidth calibration),
 *       3 = the owner launch of a BiCG iteration of the solver: convergence
 *           test, beta and p = z + beta p, pp = zz + beta pp fused into the
 *           dual apply (owner scatter only),
 *       4 = device copy of 8 N doubles (the size of an iteration's working
 *           set; buffers allocated for the call),
 *       5 / 6 = the BiCGSTAB direction / S owner launch (FP64 or FP32, owner
 *           scatter; the context's BiCGSTAB buffers of that precision, a
 *           status that never stops),
 *       7 = device copy of 64 N doubles with 16-byte accesses (the practical
 *           HBM roof: buffers far above the 256 MB last-level cache,
 *           allocated for the call). */
int vgpu_time_kernel(vgpu_ctx *ctx, int what, int scatter, int precision,
                     int reps, double *avg_ms);

/* Autotuner (no counterpart in the reference, whose CUDA launches use fixed
 * 256-thread blocks, cudatest.cu:339-347): times the solver's owner launch
 * (vgpu_time_kernel what = 3) of `precision` (VGPU_MIXED: the FP32 BiCGSTAB
 * direction and S launches, what = 5 + 6, the work of a mixed solve's inner
 * iterations, whose best shape can differ from the FP64 one: 33- against
 * 22-plane chunks at 256x256x64)
 * for every owner tile shape and z chunk in {0, 4, 8, 16} and ceil(naz / k),
 * k = 1..8 (`reps` launches each), keeps the fastest on the context and
 * returns it.  Every owner node is summed by one thread in a fixed order, so
 * the choice does not change an operator apply's result bits; a solve's
 * dot-product partials, and so its iterates, follow the tile shape and z
 * chunk. */
int vgpu_autotune(vgpu_ctx *ctx, int precision, int reps, int *scatter,
                  int *tile, int *zchunk, double *best_ms);

/* Brick tuner of the reference's two solver flavours (colour-block sizes of
 * the colored scatter, coloring.cu:289-360; brick sizes of the atomic one,
 * cudatest.cu:184-227, whose launches use fixed 256-thread blocks): times the
 * dual apply (vgpu_time_kernel what = 1) of `scatter` (VGPU_SCATTER_COLORED
 * or VGPU_SCATTER_ATOMIC) for every brick shape (tile 0 = 32 x 8, 1 = 16 x 16,
 * 2 = 64 x 4 element columns) and brick depth in {library rule (0), 1, 2, 3,
 * 4, 6, 8, 12, 16} element layers, keeps the fastest on the context (scatter,
 * tile, zchunk: the solver then uses that scatter) and returns it.  Colored
 * results are bitwise reproducible for a fixed choice and agree with the
 * oracle to <= 1e-13 for every choice (the order of a brick-face node's sums
 * depends on the bricks). */
int vgpu_autotune_bricks(vgpu_ctx *ctx, int scatter, int precision, int reps, int *tile,
                         int *zchunk, double *best_ms);

/* Device-resident forms for time stepping: b_dev = right-hand side of the
 * scheme from element sources q_dev [E] and the previous solution t_old_dev
 * [N] (may be NULL), as vgpu_assemble_rhs; and a Gaussian source
 * exp(-|x - c|^2 / (2 sigma^2)) with centre c: at element centroids times the
 * element volume into q_elem_dev [E] and at the nodes into src_node_dev [N]
 * (either may be NULL).  Both run on the context's stream. */
int vgpu_assemble_rhs_dev(vgpu_ctx *ctx, const double *q_dev, const double *t_old_dev,
                          double *b_dev);
int vgpu_gaussian_dev(vgpu_ctx *ctx, double cx, double cy, double cz, double sigma,
                      double *q_elem_dev, double *src_node_dev);
/* Device vector helpers on the context's stream: copy; out = a x + b y (zero on
 * Dirichlet nodes; y may be NULL when b = 0); out = a x + b y + c z (zero on
 * Dirichlet nodes; e.g. the quadratic extrapolation 3 x1 - 3 x2 + x3 of a
 * time stepper's initial guess); |b - A x| / |b| (FP64 owner apply,
 * synchronous).  No counterpart in the reference, whose host keeps the
 * previous step in host memory (VANADIS.FOR:809-810). */
int vgpu_copy_dev(vgpu_ctx *ctx, void *dst, const void *src, int64_t bytes);
int vgpu_axpby_dev(vgpu_ctx *ctx, double a, const double *x, double b, const double *y,
                   double *out);
int vgpu_axpbypcz_dev(vgpu_ctx *ctx, double a, const double *x, double b, const double *y,
                      double c, const double *z, double *out);
int vgpu_residual_norm_dev(vgpu_ctx *ctx, const double *b, const double *x, double *rel);
/* `sweeps` damped Jacobi sweeps x += omega D^-1 (b - A x) on the device (D the
 * solver's Jacobi diagonal): smooths the rough part of an initial guess. */
int vgpu_jacobi_dev(vgpu_ctx *ctx, const double *b, double *x, int sweeps, double omega);

/* ---- Batched solves (no counterpart in the reference, whose host solves one
 * system per call, VANADIS.FOR:809-810) ----
 * A batch groups nm contexts on one device (each with its own mesh, mask,
 * coefficients and scheme; several right-hand sides of one operator are
 * several contexts).  Every BiCG iteration runs one owner launch over all
 * systems (convergence test, direction update and dual operator) and one
 * vector launch (update), captured in one hipGraph; the systems stop
 * independently.  The batch uses the owner tile shape and z chunk that
 * ctxs[0] has when the batch is created (later vgpu_set_options calls do not
 * change the batch's); with the same tile shape each system's iterates are
 * bitwise those of vgpu_solve on its context.  Contexts must outlive the
 * batch; coefficient, scheme and diagonal changes of any member are picked
 * up at the next solve. */
typedef struct vgpu_batch vgpu_batch;
int vgpu_batch_create(int nm, vgpu_ctx *const *ctxs, vgpu_batch **out);
int vgpu_batch_destroy(vgpu_batch *batch);
/* b[m], x[m]: host vectors of system m (x in/out).  precision VGPU_FP64,
 * VGPU_FP32 or VGPU_MIXED.  VGPU_MIXED: each member as vgpu_solve's
 * VGPU_MIXED (inner rtol 1e-4), with one FP64 owner launch per refinement for
 * every member's true residual and the members' inner FP32 BiCGSTAB solves
 * run together (two owner launches per iteration over the whole batch, one
 * hipGraph); with the same tile shape and a z chunk > 0 on every member,
 * each member's result is bitwise vgpu_solve's on its context; iters = inner
 * + FP64 iterations, errs = true relative residuals.  iters/errs/status
 * [nm]: status 1 converged, 2 itmax, 3 breakdown; a member with b = 0 gets
 * x = 0, status 1, iters 0, errs 0.  Returns VGPU_OK when every system
 * converged. */
int vgpu_batch_solve(vgpu_batch *batch, int precision, const double *const *b,
                     double *const *x, double rtol, int itmax, int *iters,
                     double *errs, int *status);
/* Fallback for FP64 batch solves (no counterpart in the reference): with
 * method VGPU_METHOD_BICGSTAB, every member that the batched BiCG leaves at
 * itmax or in breakdown, or stops on its recursive residual while its true
 * relative residual |b - A x| / |b| is above rtol, is finished by FP64
 * BiCGSTAB on its own context from its current iterate (from zero when that
 * iterate's true residual is above 1), to the same rtol, with at most itmax
 * iterations per pass and up to three passes while the true residual stays
 * above rtol (status 2 if it still does; iters count every solve); VGPU_METHOD_BICG switches it
 * off (the default: members keep the batched BiCG's result).  Members that
 * converge in the batch with a true residual within rtol are untouched (still
 * bitwise vgpu_solve's). */
int vgpu_batch_set_fallback(vgpu_batch *batch, int method, int itmax);
/* Device-resident form: b, x in the batch layout (system m at offsets[m]
 * of a vector of `total` doubles; padding between systems is ignored). */
int vgpu_batch_layout(vgpu_batch *batch, int64_t *offsets, int64_t *total);
int vgpu_batch_solve_dev(vgpu_batch *batch, int precision, const double *b,
                         double *x, double rtol, int itmax, int *iters,
                         double *errs, int *status, double *seconds);

/* ---- Learned initial guess (no counterpart in the reference, whose host
 * restarts BiCG from the previous step's solution, VANADIS.FOR:809-810) ----
 * A 5-layer 3x3x3 convolution network on the context's node grid: 16
 * channels, zero "same" padding, leaky ReLU (slope 0.1) between layers, BF16
 * operands on MFMA
 * with FP32 accumulation.  Inputs: x_{n-1}, x_{n-1} - x_{n-2}, a nodal source
 * and the nodal velocity (vx, vy, vz), each times in_scale[c]; the prediction is
 * guess = x_{n-1} + out_scale * (output channel 0).
 * Parameters: w [5][27][16][16] (layer, tap dx + 3 dy + 9 dz at offset d-1,
 * output channel, input channel), b [5][16], FP32 on the host. */
typedef struct vgpu_net vgpu_net;
int vgpu_net_create(vgpu_ctx *ctx, vgpu_net **out);
/* may follow vgpu_destroy of the network's context (it waits for the device
 * and frees the network's own buffers only); every other vgpu_net_* call
 * needs the context alive */
int vgpu_net_destroy(vgpu_net *net);
int vgpu_net_set_params(vgpu_net *net, const float *w, const float *b);
int vgpu_net_get_params(vgpu_net *net, float *w, float *b);
/* gradients of the last vgpu_net_train_step_dev (same layout as w, b) */
int vgpu_net_get_grads(vgpu_net *net, float *dw, float *db);
/* nodal velocity (host, [N] each), the channel scales and the base of the
 * prediction: guess = x1 + extrap (x1 - x2) + out_scale * net (extrap 0: the
 * previous step, 1: linear extrapolation) */
int vgpu_net_set_static(vgpu_net *net, const double *vx, const double *vy,
                        const double *vz, const float *in_scale, double out_scale,
                        double extrap);
/* device vectors [N]: guess = x1 + extrap (x1 - x2) + out_scale * net(...) */
int vgpu_net_predict_dev(vgpu_net *net, const double *x1, const double *x2,
                         const double *src, double *guess);
/* one training step on one sample: loss = mean over free nodes of
 * (net - (target - x1 - extrap (x1 - x2)) / out_scale)^2, gradients by the
 * backward kernels,
 * then Adam with learning rate lr (lr <= 0: gradients only). */
int vgpu_net_train_step_dev(vgpu_net *net, const double *x1, const double *x2,
                            const double *src, const double *target, float lr,
                            double *loss);
int vgpu_net_reset_optimizer(vgpu_net *net);
/* training objective: 0 = the increment loss above (default); 1 = the initial
 * residual BiCG starts from, |A (target - guess)|^2 / |A (target - base)|^2
 * with A the context's masked FP64 operator (gradient through A^T). */
int vgpu_net_set_loss(vgpu_net *net, int kind);
/* passes (0..8, default 0) of a [1 4 1]^3 / 216 low-pass filter on the network
 * output, in prediction and training alike: it damps the grid-scale modes
 * that BF16 rounding leaves in the correction, the slowest modes of a
 * mass-dominated operator for BiCG. */
int vgpu_net_set_smoothing(vgpu_net *net, int passes);
/* Node planes each wave of the layer kernels marches (0, the default: the
 * count that balances the waves over the device's resident slots; n > 0: n).
 * The outputs do not depend on it (bitwise). */
int vgpu_net_set_zchunk(vgpu_net *net, int zchunk);
/* Average GPU time (HIP events around a captured graph of reps launches on
 * the context's stream) of: 0 a hidden forward layer (16 -> 16 channels,
 * leaky ReLU), 1 a backward-data layer, 2 a layer's weight and bias
 * gradients, 3 the whole prediction network (input packing + 5 layers).
 * Timing only: runs on the network's own buffers. */
int vgpu_net_time_layers(vgpu_net *net, int what, int reps, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
