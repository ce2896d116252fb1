/*
 * fvdg.h -- C ABI of the MI355X explicit DG Euler solver (libfvdg.so).
 *
 * This is the only boundary between host code and the GPU.  Plain pointers,
 * sizes and status codes; no HIP or C++ types cross it.  Every call returns
 * FVDG_OK (0) or a negative status; fvdg_last_error() gives the message of
 * the last failure on the calling thread.
 *
 * Reference interface replaced (the reference is a C++ template library with no
 * FFI; these calls are what its ExplicitConvection path needs):
 *   fvdg_mesh_*              ComputingMesh(GeneralMesh) + OrthHexMesh + split_hex*_scan
 *                            (include/Mesh/ComputingMesh.h:98-109, src/Mesh/ComputingMesh.cpp:205-263,
 *                             include/Mesh/OrthHexMesh.h:7-106, include/Mesh/GeneralMesh.h:24-25)
 *   fvdg_residual            ExplicitConvection<Order,Flux>::eval / eval_cells / eval_internals /
 *                            eval_boundarys (include/DG/DG_Schemes/ExplicitConvection.h:32-43,
 *                             src/DG/DG_Schemes/ExplicitConvection.cpp:21-272)
 *   fvdg_mass_diagonal       the driver's diagonal mass b[cell](5i+k)
 *                            (examples/Vortex/123/VortexBatch.cpp:153-166)
 *   fvdg_ssprk3_step         the driver's explicit loop x += dt * (-1/b * eval(...))
 *                            (examples/Vortex/123/VortexBatch.cpp:183-208), as Shu-Osher SSP-RK3
 *   fvdg_project_analytic    DGBasisEvaluator<P>::func2coef of the vortex exact solution
 *                            (include/DG/DG_Basis/DG_Basis_Evaluator.h:33-59, examples/Vortex/problem.h:17-110)
 *
 * State layout at this boundary is the reference's LongVector<5N>: element e,
 * mode k, variable v (rho, rho u, rho v, rho w, rho E) at U[e*5N + 5k + v].
 */
#ifndef FVDG_H
#define FVDG_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fvdg_mesh fvdg_mesh; /* host mesh: topology + geometry */
typedef struct fvdg_ctx fvdg_ctx;   /* one GPU, one stream, one mesh, one scheme */

/* status codes */
#define FVDG_OK 0
#define FVDG_ERR_ARG (-1)
#define FVDG_ERR_HIP (-2)
#define FVDG_ERR_MESH (-3)
#define FVDG_ERR_STATE (-4)
#define FVDG_ERR_UNSUPPORTED (-5)
#define FVDG_ERR_NONFINITE (-6)

/* numerical flux */
#define FVDG_FLUX_LF 0  /* Lax-Friedrichs with lambda = max(|u|+c) (ExplicitConvection.cpp:232-240) */
#define FVDG_FLUX_ROE 1 /* Roe with Harten entropy fix (no counterpart in the reference) */

/* arithmetic of the volume/face contractions (state and pointwise flux stay FP64) */
#define FVDG_PREC_FP64 0
#define FVDG_PREC_FP32 1
#define FVDG_PREC_BF16X3 2

/* boundary types: the numbering of BoundaryType (include/Mesh/ComputingMesh.h) */
#define FVDG_BC_DIRICHLET 0
#define FVDG_BC_PSEUDO3DZ 5

/* residual parts (bit mask): eval_cells, eval_internals, eval_boundarys */
#define FVDG_PART_CELLS 1
#define FVDG_PART_INTERNALS 2
#define FVDG_PART_BOUNDARYS 4
#define FVDG_PART_ALL 7

/* built-in analytic fields for fvdg_project_analytic / fvdg_analytic_error */
#define FVDG_FIELD_VORTEX 0 /* params[8]: gamma, strength, x0, y0, u_inf, v_inf, Lx, Ly (periods, 0 = none) */
#define FVDG_FIELD_TGV 1    /* params[4]: gamma, rho0, V0, Mach */
#define FVDG_FIELD_CONST 2  /* params[5]: rho, rho u, rho v, rho w, rho E */

const char* fvdg_version(void);
const char* fvdg_last_error(void);
/* (P+1)(P+2)(P+3)/6 modes per variable (DGBasisEvaluator<P>::NumBasis); 0 for order < 0 */
int fvdg_num_basis(int order);
/* values val[N] and reference gradients grad[3N] (either may be NULL) of the order-`order`
 * modal basis at reference point (x,y,z) (DGBasisEvaluator<P>::eval_all / grad_all,
 * include/DG/DG_Basis/DG_Basis_Evaluator.h:14-31); returns N or a negative status */
int fvdg_basis_eval(int order, double x, double y, double z, double* val, double* grad);
/* points of the volume / face rule used at this order (reference AutoQuadSelector) */
int fvdg_quadrature(int order, int* n_vol, double* vol_pts, double* vol_w, int* n_face, double* face_pts,
                    double* face_w);

/* ------------------------------ host mesh ------------------------------ */
/* Box [lb,ub] of n[0]*n[1]*n[2] hexahedra split into tets.  split = 6 / 5: this library's
 * translation-invariant 6- / 5-tet patterns (periodic pairing possible); split = 16 / 15: the
 * reference's OrthHexMesh + split_hex6_scan / split_hex5_scan + rebuild_cell_topology
 * (include/Mesh/OrthHexMesh.h:3-106, src/Mesh/GeneralMesh.cpp:4-300), the same tets, vertex
 * orders and face numbering (no periodic pairing).
 * periodic_mask bit d pairs the faces on x_d = lb_d and x_d = ub_d.  jitter moves every
 * vertex class by U(-jitter*h, jitter*h) per coordinate (seeded, periodic images alike). */
int fvdg_mesh_create_box(fvdg_mesh** out, const double lb[3], const double ub[3], const int64_t n[3], int split,
                         int periodic_mask, double jitter, uint64_t seed);
/* From arrays.  face_cells[2f+1] = -1 marks a boundary face.  face_nodes_r (nullable) gives
 * the face nodes as numbered in face_cells[2f+1] (differs only across periodic boundaries). */
int fvdg_mesh_create(fvdg_mesh** out, int64_t n_points, const double* points, int64_t n_cells,
                     const int64_t* cell_nodes, int64_t n_faces, const int64_t* face_nodes,
                     const int64_t* face_cells, const int64_t* face_nodes_r);
/* read a Fluent ASCII mesh file (.msh, like the reference's data/case11.msh): cells other than
 * tetrahedra are split into tetrahedra around their centroids (quadrilateral faces cut along the
 * diagonal through their smallest node, so the result is conforming); boundary faces get the
 * BoundaryType of their zone's Fluent type (inlet -> Inflow, outlet -> Outflow, symmetry ->
 * Neumann, far field -> Dirichlet, others -> Wall).  No counterpart in the reference, whose
 * mesh reader is commented out (tests/a_new_test.cpp:16). */
int fvdg_mesh_read_fluent(fvdg_mesh** out, const char* path);
/* per face: the file zone of the face (-1: a face created inside a split cell or not from a file) */
int fvdg_mesh_face_zones(const fvdg_mesh* m, int32_t* zone);
/* zones of a mesh read from a file: ids, Fluent boundary-condition codes, names (name_len bytes each) */
int fvdg_mesh_zones(const fvdg_mesh* m, int* n, int32_t* ids, int32_t* bc_types, char* names, int name_len);
/* set the boundary type of every boundary face of a zone (e.g. FVDG_BC_DIRICHLET, FVDG_BC_PSEUDO3DZ) */
int fvdg_mesh_set_zone_boundary_type(fvdg_mesh* m, int32_t zone, uint8_t bc);
int fvdg_mesh_sizes(const fvdg_mesh* m, int64_t* n_points, int64_t* n_cells, int64_t* n_faces);
/* any output pointer may be NULL */
int fvdg_mesh_export(const fvdg_mesh* m, double* points, int64_t* cell_nodes, int64_t* face_nodes,
                     int64_t* face_cells, int64_t* face_nodes_r);
/* per cell: inverse Jacobian (A^{-1}, row-major 9), |det A|, h = V/sum(face areas);
 * per face: unit normal (cell 0 -> cell 1), area */
int fvdg_mesh_geometry(const fvdg_mesh* m, double* cell_inv_jac, double* cell_det, double* cell_h,
                       double* face_normal, double* face_area);
int fvdg_mesh_set_boundary_types(fvdg_mesh* m, const uint8_t* bc); /* n_faces entries */
int fvdg_mesh_get_boundary_types(const fvdg_mesh* m, uint8_t* bc);
/* the reference example's tagging (examples/Vortex/tools.h:22-31):
 * boundary face with |n_z| > 0.5 -> Pseudo3DZ, other boundary faces -> Dirichlet */
int fvdg_mesh_tag_boundaries_pseudo3dz(fvdg_mesh* m);
void fvdg_mesh_destroy(fvdg_mesh* m);

/* Diagonal L2 projection of a built-in field at time t onto the order-`order` basis,
 * written as U[e*5N + 5k + v] (host computation). */
int fvdg_project_analytic(const fvdg_mesh* m, int order, int field, const double* params, int n_params, double t,
                          double* U);
/* Relative L2 error per variable of U against the field at time t, on the volume rule
 * (sqrt(sum (U_h-U_s)^2 w|J| / sum U_s^2 w|J|), examples/Eular1D/Sod1D/Riemann1D.cpp:250-267). */
int fvdg_analytic_error(const fvdg_mesh* m, int order, int field, const double* params, int n_params, double t,
                        const double* U, double err[5]);

/* ---------------------------- device solver ---------------------------- */
/* number of GPUs the HIP runtime sees (0 without a driver or device); no reference counterpart:
 * lets callers choose a device without another GPU library */
int fvdg_device_count(int* n);
int fvdg_create(fvdg_ctx** out, int device);
void fvdg_destroy(fvdg_ctx* c);
/* order 0..5; flux FVDG_FLUX_*; gamma > 1; precision FVDG_PREC_FP64 (the drop-in),
 * FVDG_PREC_FP32 (volume and face-test contractions in FP32 with FP32 accumulation; face
 * traces in FP64 at P2-P4, where that kernel is the faster one, in FP32 elsewhere; face fluxes
 * stored as FP32 differences from an FP64 reference flux per face) or
 * FVDG_PREC_BF16X3 (volume contractions on the bf16 matrix cores: each FP32 operand split into
 * three bf16 parts, six part products accumulated in FP32; face contractions as FP32).  In both
 * mixed forms the pointwise and Riemann fluxes, stage update, state and conservation sums stay
 * FP64.  Changing only the precision keeps an uploaded state. */
int fvdg_set_scheme(fvdg_ctx* c, int order, int flux, double gamma, int precision);
/* Copies the mesh geometry to the device; the mesh object may be destroyed afterwards. */
int fvdg_set_mesh(fvdg_ctx* c, const fvdg_mesh* m);
int fvdg_state_upload(fvdg_ctx* c, const double* U);
int fvdg_state_download(fvdg_ctx* c, double* U);
/* Boundary data: number of boundary faces that take data (Dirichlet; WallTD and WallTN, used by
 * the implicit path only), and for each its face id and the physical positions of its face
 * quadrature points (xyz[(i*Qf+g)*3+d]); either pointer may be NULL. */
int fvdg_dirichlet_points(fvdg_ctx* c, int64_t* n_faces, int64_t* face_ids, double* xyz);
/* States U_D at those points, UD[(i*Qf+g)*5+v]; used by every later residual. */
int fvdg_set_dirichlet_states(fvdg_ctx* c, const double* UD);
/* R = sum of the selected parts of the residual of the current state, written as
 * R[e*5N+5k+v] when R is not NULL (the device copy is kept either way). */
int fvdg_residual(fvdg_ctx* c, double t, int parts, double* R);
/* Diagonal mass M[e*5N+5k+v] = |det A_e| * sum_q w_q phi_k(xi_q)^2. */
int fvdg_mass_diagonal(fvdg_ctx* c, double* M);
/* n_steps SSP-RK3 steps of dU/dt = -M^{-1} R(U) on the device state.  Returns
 * FVDG_ERR_NONFINITE when a stage wrote a non-finite coefficient or a non-positive mean density
 * (mode 0 of rho); the steps still run to the end and the state is left as computed.  (The
 * reference has no such check; its drivers stop on a non-positive time step,
 * examples/Sedov/SedovBatch.cpp:200.) */
int fvdg_ssprk3_step(fvdg_ctx* c, double dt, int n_steps);
/* stage 0, 1 or 2 of one Shu-Osher SSP-RK3 step on the device state (U^n -> U1 -> U2 -> U^{n+1});
 * calling 0, 1, 2 in order equals one fvdg_ssprk3_step.  Between stages the caller may
 * update the Dirichlet states to the stage times t, t + dt, t + dt/2. */
int fvdg_ssprk3_stage(fvdg_ctx* c, double dt, int stage);  /* same FVDG_ERR_NONFINITE check */
/* dt = cfl * h_min / ((2p+1) * lambda_max), h = 3 * volume / (sum of the cell's face areas),
 * lambda_max = max over face points of max_{L,R}(|u| + c) of the device state
 * (examples/Sedov/SedovBatch.cpp:174-199 form, with the tet inradius as h) */
int fvdg_stable_dt(fvdg_ctx* c, double cfl, double* dt);
/* fvdg_stable_dt without any device allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same result as fvdg_stable_dt.  Its reduction partials live in a
 * library-static device array of 4,096 + 1 doubles (32 KiB per GPU, held until the process ends;
 * the grid-stride reduction fits any mesh in it); calls that overlap on one GPU share that array
 * and race. */
int fvdg_stable_dt_noalloc(fvdg_ctx* c, double cfl, double* dt);
/* max over all face quadrature points of max(|u|+c) of the two traces; NaN when any trace is
 * not finite (fvdg_stable_dt then returns FVDG_ERR_NONFINITE) */
int fvdg_max_wave_speed(fvdg_ctx* c, double* out);
/* sum over elements of |det A_e| * U_{e,k=0,v} / 6 = integral of each conserved variable */
int fvdg_conservation_sums(fvdg_ctx* c, double out[5]);
/* Time n_stages SSP-RK3 stages with HIP events on the solver stream: ms[0] = wall time per
 * stage of a pass with no host synchronisation inside; then a second pass with events between
 * kernels gives ms[1] = volume, ms[2] = face flux, ms[3] = face test + stage update (or the fused
 * element kernel: volume + face test + update, with ms[1] = 0) (per-stage averages).  The state
 * advances by 2 n_stages / 3 steps. */
int fvdg_time_stages(fvdg_ctx* c, double dt, int n_stages, double* ms);
/* The kernels one stage launches on the current scheme, "volume=<name>;flux=<name>;test=<name>"
 * (NUL-terminated into buf[len]; an empty name: nothing is launched in that slot, e.g. the volume
 * term inside the fused element kernel).  Measurement aid; no reference counterpart. */
int fvdg_stage_kernels(fvdg_ctx* c, char* buf, int len);

/* ---------------- implicit path: ImplicitConvection<Order,Flux>::assemble
 * (include/DG/DG_Schemes/ImplicitConvection.h:10-54, src/DG/DG_Schemes/ImplicitConvection.cpp)
 * The Picard-linearised system R(U) ~ A(U_lin) U - rhs at the device state U_lin: per element a
 * diagonal block and one block per interior face (NB x NB, NB = 5N, row 5j+v, column 5i+u),
 * boundary faces linearised around the reference's ghost states (Dirichlet / WallTD / WallTN
 * with the data of fvdg_set_dirichlet_states, Pseudo3DX/Y/Z mirrors).  Kept on the device. */
int fvdg_implicit_assemble(fvdg_ctx* c, int parts);
/* host copies of the last assembly (any pointer may be NULL): diag[E][NB][NB] and off[E][4][NB][NB]
 * row-major, off_cols[E][4] the neighbour across local face l (-1: boundary face, zero block),
 * rhs[E][NB] */
int fvdg_implicit_blocks(fvdg_ctx* c, double* diag, double* off, int32_t* off_cols, double* rhs);
/* y = (A + shift M) x, M the diagonal mass (host vectors, n_cells x 5N) */
int fvdg_implicit_matvec(fvdg_ctx* c, double shift, const double* x, double* y);
/* solve (A + shift M) x = b (b NULL: the assembled rhs) by restarted GMRES(restart) with a
 * block-Jacobi preconditioner (each diagonal block + shift M inverted explicitly for NB <= 100,
 * LU-factored beyond), from the guess in x, to a
 * relative residual |b - (A + shift M) x| / |b| < tol or max_iters iterations; host vectors.
 * iters / rel_res (nullable): iterations done and the final true relative residual.
 * Returns FVDG_ERR_NONFINITE when a diagonal block is singular. */
int fvdg_implicit_solve(fvdg_ctx* c, double shift, const double* b, double* x, double tol, int restart,
                        int max_iters, int* iters, double* rel_res);
/* One backward-Euler step of the device state by Picard iteration, as the reference's implicit
 * drivers do (examples/Eular2D/Sedov/Sedov.cpp:137-226): dx = U; repeat { assemble at dx; solve
 * (M/dt + A) x = M/dt U + rhs from dx; limit x against dx if a stage limiter is set;
 * delta = |x - dx|; dx = x } until delta < picard_tol, delta / previous delta < picard_tol,
 * after 10 iterations delta / first delta > 0.95, or max_picard iterations; U = dx.
 * picard_iters / linear_iters / delta (nullable): iterations done, GMRES iterations in total,
 * last delta. */
int fvdg_implicit_euler_step(fvdg_ctx* c, double dt, int max_picard, double picard_tol, double tol, int restart,
                             int max_iters, int* picard_iters, int* linear_iters, double* delta);
/* Solve a general block-sparse system (the reference's BlockSparseMatrix + EigenSparseSolver::DGMRES,
 * include/Matrix/SparseMatrix.h, src/EigenSolver/EigenSparseSolver.cpp:163-195) on device `device`:
 * n_rows block rows of size block, CSR row_ptr[n_rows+1] / cols[nnz] / blocks[nnz][block][block]
 * row-major; same GMRES and stopping rule as fvdg_implicit_solve (shift 0). */
int fvdg_bsr_solve(int device, int64_t n_rows, int block, const int64_t* row_ptr, const int64_t* cols,
                   const double* blocks, const double* b, double* x, double tol, int restart, int max_iters,
                   int* iters, double* rel_res);

/* flow diagnostics of the device state, integrals at the volume quadrature points with the
 * weights w_q |det A| of the reference's err_integral (examples/Vortex/VortexBatch3.cpp:349-366):
 * out[0] volume, [1] mass (rho), [2] kinetic energy (rho |u|^2 / 2), [3] enstrophy
 * (rho |curl u|^2 / 2), [4] total energy (rho E); [5] min rho, [6] min p, [7] max Mach over the
 * points.  [2] and [3] divided by rho_0 * volume are the Taylor-Green vortex's E_k and epsilon
 * series (no counterpart in the reference). */
int fvdg_flow_diagnostics(fvdg_ctx* c, double out[8]);

/* ---------------- limiter: PositiveLimiter<Order, QuadC, OnlyNeigbAvg> of
 * include/DG/DG_Schemes/PositiveLimiter.h:12-261, sampling each element at the points of
 * its order's tetrahedron rule.  Cell averages (mode 0) are never changed. */
#define FVDG_LIMIT_BOUNDS 1     /* apply_1 (:135-170): point values into the neighbourhood bounds */
#define FVDG_LIMIT_POSITIVITY 2 /* apply_2 (:171-259): density, then pressure, kept positive */
/* constructMinMax (:62-119): bounds from U_prev (host, n_cells x 5N; NULL = the device
 * state), from the point values of each cell and its face neighbours, or from their
 * averages when neighbour_averages != 0 (OnlyNeigbAvg = true) */
int fvdg_limiter_bounds(fvdg_ctx* c, const double* U_prev, int neighbour_averages);
/* apply (parts = 3), apply_1 (1) or apply_2 (2) on U (host, n_cells x 5N, limited in
 * place; NULL = the device state); FVDG_LIMIT_BOUNDS needs fvdg_limiter_bounds first */
int fvdg_limiter_apply(fvdg_ctx* c, double* U, int parts);
/* limit every SSP-RK3 stage result (fvdg_ssprk3_step / _stage) with bounds from that
 * stage's input state; parts 0 switches it off (the default).  No counterpart in the
 * reference's explicit drivers, which call the limiter by hand after implicit solves
 * (examples/Eular2D/Sedov/Sedov.cpp:188-189). */
int fvdg_set_stage_limiter(fvdg_ctx* c, int parts, int neighbour_averages);

/* solution files in the reference's text format (examples/Vortex/VortexBatch3.cpp:402-415): one
 * line per (cell, volume quadrature point): x y z, then U_h U_s for rho, rho u, rho v, rho w,
 * rho E, 16 significant digits; U_s is the analytic field (field >= 0) or 0 (field < 0) */
int fvdg_solution_write(const fvdg_mesh* m, int order, const double* U, int field, const double* params,
                        int n_params, double t, const char* path);
/* read such a file back into coefficients (diagonal projection of the U_h point values,
 * examples/Vortex/tools.h:72-129) */
int fvdg_solution_read(const fvdg_mesh* m, int order, const char* path, double* U);

/* binary checkpoint of a state (restart without the projection error of the text files): a 56-byte
 * header -- "FVDGCKP1", version 1, order, cells, values = cells * 5N, time, step, FNV-1a hash of the
 * coefficient bytes -- then the raw FP64 coefficients in the LongVector layout; written to path.tmp
 * and renamed.  The reference restarts only from its text files (examples/Vortex/tools.h:73-129). */
int fvdg_checkpoint_write(const char* path, int order, int64_t n_cells, const double* U, double t, int64_t step);
/* reads one back: with U == NULL only the header fields (each pointer nullable); else also the
 * n_cells * 5N coefficients, checked against the hash.  FVDG_ERR_ARG: not a checkpoint;
 * FVDG_ERR_STATE: truncated, inconsistent or corrupt. */
int fvdg_checkpoint_read(const char* path, int* order, int64_t* n_cells, double* U, double* t, int64_t* step);

/* ---------------- ensembles: many independent problems of uneven size and order ----------------
 * (no counterpart in the reference, whose drivers run one problem at a time.)  Problems of one
 * order are stepped as one disjoint-union mesh with per-problem time steps; one SSP-RK3 step of
 * the whole ensemble can be captured into a HIP graph and replayed. */
typedef struct fvdg_ensemble fvdg_ensemble;
/* n problems: meshes[i] (caller keeps ownership; copied) at order orders[i] (0..5) */
int fvdg_ensemble_create(fvdg_ensemble** out, int device, int n, const fvdg_mesh* const* meshes, const int* orders,
                         int flux, double gamma, int precision);
void fvdg_ensemble_destroy(fvdg_ensemble* e);
int fvdg_ensemble_num_cells(fvdg_ensemble* e, int problem, int64_t* n_cells);
int fvdg_ensemble_num_groups(fvdg_ensemble* e, int* n_groups, int* orders);
/* state of one problem, LongVector layout (n_cells x 5N) */
int fvdg_ensemble_upload(fvdg_ensemble* e, int problem, const double* U);
int fvdg_ensemble_download(fvdg_ensemble* e, int problem, double* U);
/* per problem: max over face points of |u| + c of the device state (out[n]) */
int fvdg_ensemble_max_wave_speeds(fvdg_ensemble* e, double* out);
/* per-problem time steps dt[n] */
int fvdg_ensemble_set_dt(fvdg_ensemble* e, const double* dt);
/* n_steps SSP-RK3 steps of every problem; use_graph: replay a captured HIP graph of one step */
int fvdg_ensemble_step(fvdg_ensemble* e, int n_steps, int use_graph);
/* average ms per ensemble step over n_steps (after one untimed step) */
int fvdg_ensemble_time_steps(fvdg_ensemble* e, int n_steps, int use_graph, double* ms_per_step);
/* per order group, time the stage's kernel choices over trial_steps steps on the ensemble's own
 * state and keep the fastest (state restored): the face tiles (FP32 groups at P2-P4), the fused
 * element kernel or the volume + face-test kernels, and the face-flux kernel (row-broadcast in
 * 128-, 256- or 512-thread blocks at P2-P4, or LDS-staged); choices[2g] = 1 for the fused element
 * kernel, choices[2g+1] = the row-broadcast kernel's block size (0: LDS-staged, -1: face tiles) */
int fvdg_ensemble_autotune(fvdg_ensemble* e, int trial_steps, int* choices);
/* serial != 0: every group on the ensemble's stream one after another (for comparison); default:
 * each group on its own stream, forked from and joined back into the ensemble's stream per step */
int fvdg_ensemble_set_serial(fvdg_ensemble* e, int serial);

#ifdef __cplusplus
}
#endif
#endif /* FVDG_H */
