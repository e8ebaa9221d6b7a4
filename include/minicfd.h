/*
 * minicfd.h -- C ABI of the MI355X-native high-order residual and
 * time-stepping library (libminicfd.so).
 *
 * Every entry point replaces one piece of the reference's hot path behind
 * the mini::temporal::System contract (include/mini/temporal/ode.hpp:17-48)
 * and the mini::spatial::FiniteElement family (include/mini/spatial/fem.hpp).
 * The reference is a header-only C++ library with no foreign-function layer
 * of its own; these functions are what a C/ctypes/cgo binding of that
 * surface binds.  Each declaration names the reference interface it replaces.
 *
 * Conventions
 *   - every function returns 0 on success and a nonzero MCFD_E* code on
 *     failure; mcfd_last_error() describes the last failure of the calling
 *     thread (no exception crosses the ABI; the reference only asserts,
 *     include/mini/algebra/root.hpp:67 throws);
 *   - a solution/residual "column" has the reference's layout
 *     column[(cell * N + node) * 5 + k] with k = (rho, rho u, rho v, rho w, E)
 *     (include/mini/mesh/part.hpp:754-759, polynomial/expansion.hpp:83-104);
 *   - host pointers are caller-owned; the handle owns all device memory;
 *   - one handle per host thread; calls on a handle are stream-ordered on the
 *     handle's own HIP stream (the reference is single-threaded per rank).
 */
#ifndef MINICFD_H_
#define MINICFD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes */
#define MCFD_OK 0
#define MCFD_EINVAL 1   /* invalid argument or mesh */
#define MCFD_EHIP 2     /* HIP runtime error */
#define MCFD_ESTATE 3   /* call not valid in the handle's current state */
#define MCFD_ENOTSUP 4  /* configuration outside what the library implements */

/* ---- cell types (CGNS element types used by the reference demos) */
#define MCFD_HEXA_8 0
#define MCFD_TETRA_4 1

/* ---- spatial schemes
 * FR_GENERAL: nodal flux reconstruction, spatial/fr/general.hpp (solution
 *   points = tensor-product quadrature nodes; Gauss-Legendre or Lobatto).
 * DG_GENERAL: modal DG on the per-cell orthonormal basis,
 *   spatial/dg/general.hpp with polynomial/projection.hpp. */
#define MCFD_SCHEME_FR_GENERAL 0
#define MCFD_SCHEME_DG_GENERAL 1

#define MCFD_NODES_LEGENDRE 0
#define MCFD_NODES_LOBATTO 1

/* ---- Riemann solvers, riemann/euler/{exact,hllc}.hpp; Roe follows
 * python/riemann.py:405-451 extended to 3-D in the rotated frame. */
#define MCFD_RIEMANN_EXACT 0
#define MCFD_RIEMANN_HLLC 1
#define MCFD_RIEMANN_ROE 2

/* ---- precision forms */
#define MCFD_PRECISION_FP64 0
#define MCFD_PRECISION_MIXED 1

/* ---- creation flags
 * HOST_ONLY: build connectivity and tables on the host only (no device
 * memory, no HIP calls); compute entry points then return MCFD_ESTATE.
 * Used to inspect the face construction on machines without a GPU. */
#define MCFD_FLAG_HOST_ONLY 1

/* ---- limiters, applied whenever a solution column is set: after
 * SetSolutionColumn, Approximate and every fused RK stage
 * (spatial/with_limiter.hpp:56-66, limiter/reconstruct.hpp:22-47).
 * WENO_FLAGGED: weno::Lazy reconstruction (limiter/weno.hpp:97-195) of the
 *   cells that the troubled-cell indicator weno::Eigen::IsNotSmooth
 *   (weno.hpp:230-258) flags.
 * WENO_ALL: weno::Lazy on every cell (Lazy::IsNotSmooth is always true; the
 *   configuration of demo/euler/sourceless.hpp:80-81).
 * WENO_EIGEN: weno::Eigen reconstruction (weno.hpp:259-348), i.e. WENO in
 *   the characteristic variables of each adjacent face, of the flagged cells.
 * WENO_EIGEN_ALL: the same on every cell. */
#define MCFD_LIMITER_NONE 0
#define MCFD_LIMITER_WENO_FLAGGED 1
#define MCFD_LIMITER_WENO_ALL 2
#define MCFD_LIMITER_WENO_EIGEN 3
#define MCFD_LIMITER_WENO_EIGEN_ALL 4

/* ---- boundary conditions, spatial/fem.hpp:196-222.  The `state` of an
 * mcfd_bc_desc is the tag's constant given value; mcfd_set_boundary_function
 * replaces it by a function of the point and the time.
 * SUBSONIC_INLET: state = (p_total, u_cos, v_cos, w_cos, T_total), the total
 *   conditions and global flow direction of riemann/rotated/euler.hpp:181-246;
 * SUBSONIC_OUTLET: state[4] = the back pressure (euler.hpp:263-280). */
#define MCFD_BC_SMART 0               /* SetSmartBoundary(name, f) */
#define MCFD_BC_INVISCID_WALL 1       /* SetInviscidWall(name) */
#define MCFD_BC_SUPERSONIC_OUTLET 2   /* SetSupersonicOutlet(name) */
#define MCFD_BC_SUPERSONIC_INLET 3    /* SetSupersonicInlet(name, f) */
#define MCFD_BC_SUBSONIC_INLET 4      /* SetSubsonicInlet(name, f) */
#define MCFD_BC_SUBSONIC_OUTLET 5     /* SetSubsonicOutlet(name, f) */

typedef struct mcfd_solver mcfd_solver;

/* Mesh in the reference's input form (CGNS connectivity, one rank).
 * Replaces mesh::part::Part construction from CGNS + METIS partition files
 * (mesh/part.hpp:453-476).  Negative boundary tags mark periodic faces,
 * which are paired across the box by `period` (no reference counterpart). */
typedef struct {
  int32_t cell_type;            /* MCFD_HEXA_8 or MCFD_TETRA_4 */
  int64_t n_nodes;
  const double *coords;         /* [n_nodes][3] */
  int64_t n_cells;
  const int64_t *cell_nodes;    /* [n_cells][npe], CGNS order, 0-based */
  int64_t n_bfaces;
  const int64_t *bface_nodes;   /* [n_bfaces][face npe] */
  const int32_t *bface_tags;    /* [n_bfaces]: BC table index, < 0 periodic */
  double period[3];             /* box length of periodic directions, else 0 */
} mcfd_mesh_desc;

typedef struct {
  int32_t scheme;       /* MCFD_SCHEME_* */
  int32_t degree;       /* polynomial degree p */
  int32_t nodes;        /* MCFD_NODES_* (FR_GENERAL) */
  int32_t riemann;      /* MCFD_RIEMANN_* */
  int32_t precision;    /* MCFD_PRECISION_* */
  int32_t flags;        /* MCFD_FLAG_* */
  double c_next;        /* Vincent c_next: 1 = DG, p/(2p+1) = Huynh g2 */
  int32_t limiter;      /* MCFD_LIMITER_* (DG_GENERAL only) */
  int32_t reserved;
  double weno_w0;       /* linear weight of each neighbour (sourceless.cpp:114: 0.001) */
  double weno_eps;      /* smoothness offset (sourceless.cpp:114: 1e-6) */
} mcfd_scheme_desc;

typedef struct {
  int32_t kind;         /* MCFD_BC_* */
  int32_t reserved;
  double state[5];      /* conservative state for SMART / SUPERSONIC_INLET */
} mcfd_bc_desc;

/* Build a solver.  Replaces the constructors FiniteElement(Part*)
 * (spatial/fem.hpp:95-115), fr::General(Part*, c_next)
 * (spatial/fr/general.hpp:218-231) and dg::General(Part*)
 * (spatial/dg/general.hpp:38-40) plus the BC registration
 * (spatial/fem.hpp:196-222). */
int mcfd_create(const mcfd_mesh_desc *mesh, const mcfd_scheme_desc *scheme,
                const mcfd_bc_desc *bcs, int32_t n_bcs, mcfd_solver **out);
int mcfd_destroy(mcfd_solver *h);
const char *mcfd_last_error(void);

/* Sizes: cells, unknowns per cell per component (N), column length. */
int mcfd_sizes(const mcfd_solver *h, int64_t *n_cells, int32_t *n_per_cell,
               int64_t *column_size);

/* System::SetTime (temporal/ode.hpp:27, spatial/fem.hpp:225-227). */
int mcfd_set_time(mcfd_solver *h, double t);
/* System::SetSolutionColumn (ode.hpp:33, fem.hpp:228-236); host column. */
int mcfd_set_solution_column(mcfd_solver *h, const double *column);
/* System::GetSolutionColumn (ode.hpp:40, fem.hpp:237-247); host column. */
int mcfd_get_solution_column(mcfd_solver *h, double *column);
/* Part::ReadSolutions + ScatterSolutions, the restart path of
 * demo/euler/sourceless.cpp:167-171 (mesh/part.hpp:1070-1208): sets the
 * solution's coefficients as stored, without the limiter pass that
 * mcfd_set_solution_column runs on a handle with a limiter; host column. */
int mcfd_restore_solution_column(mcfd_solver *h, const double *column);
/* System::GetResidualColumn (ode.hpp:47, fem.hpp:248-258 with the FR
 * det(J) scaling of fr/general.hpp:249-268); host column. */
int mcfd_get_residual_column(mcfd_solver *h, double *column);

/* RungeKutta<3>::Update (temporal/rk.hpp:31-49) fused on the device: three
 * residual evaluations with the Shu-Osher combinations in the reference's
 * operation order; the state stays resident in HBM. */
int mcfd_rk3_update(mcfd_solver *h, double t, double dt);
/* One stage of the above: stage 1 U1 = U0 + dt R(U0); stage 2
 * U2 = (3 U0 + (U1 + dt R(U1))) / 4; stage 3 U0 <- (2 (U2 + dt R(U2)) + U0) / 3. */
int mcfd_rk3_stage(mcfd_solver *h, int32_t stage, double t, double dt);

/* FiniteElement::GetTimeStep (spatial/fem.hpp:417-435) with an explicit
 * CFL number instead of the cfl_ table (fem.hpp:26-37). */
int mcfd_get_time_step(mcfd_solver *h, double cfl, double dt_guess, double *dt);

/* Coordinates of the points where an initial condition is sampled:
 * FR_GENERAL: solution points ([n_cells][N][3], polynomial/hexahedron.hpp:646-652).
 * DG_GENERAL: cell quadrature points ([n_cells][Nq][3], projection.hpp:28-43). */
int mcfd_sample_points(const mcfd_solver *h, int32_t *n_per_cell, double *xyz);
/* polynomial().LocalToGlobalAndValue of every cell (the evaluation that
 * mesh/vtk.hpp:431-460 writes frames with): at n_local local points
 * ([n_local][3]; hexahedra (xi, eta, zeta) in [-1, 1]^3 with the CGNS corners
 * at +-1; tetrahedra with CGNS vertices 0..3 at (1,0,0), (0,1,0), (0,0,1),
 * (0,0,0)) the global points xyz [n_cells][n_local][3] and the solution's
 * values [n_cells][n_local][5].  `column` is a solution column in the layout
 * of mcfd_get_solution_column, or null for the handle's current solution
 * (MCFD_ESTATE on a host-only handle).  xyz or values may be null. */
int mcfd_evaluate(mcfd_solver *h, const double *column, int32_t n_local, const double *local, double *xyz,
                  double *values);
/* Polynomial::Approximate (fem.hpp:147-154): build the solution from values
 * at the sample points ([n_cells][n_per_cell][5]) and make it current. */
int mcfd_approximate(mcfd_solver *h, const double *values);

/* Connectivity export for bit-exact comparison with the reference's face
 * construction (mesh/part.hpp:932-975,1456-1549) and FR flux-point caches
 * (spatial/fr/general.hpp:85-211): per face holder, sharer (-1 on the
 * boundary), local face ids, and per face point the FluxPointCache::ijk of
 * the holder and of the sharer (FR_GENERAL handles only; DG_GENERAL handles
 * leave fp_ijk untouched, and any output pointer may be null). */
int mcfd_face_count(const mcfd_solver *h, int64_t *n_faces, int64_t *n_interior);
int mcfd_faces(const mcfd_solver *h, int64_t *holder, int64_t *sharer,
               int32_t *holder_face, int32_t *sharer_face, int32_t *fp_ijk);

/* Boundary values that vary in space and time: the function argument of
 * SetSmartBoundary / SetSupersonicInlet / SetSubsonicInlet /
 * SetSubsonicOutlet (spatial/fem.hpp:196-222), which the reference evaluates
 * at every boundary quadrature point at the system's time whenever it forms a
 * residual (dg/general.hpp:92-202, fr/general.hpp:435-640).  The library calls
 * f once per tag with all of the tag's points, states[n][5] = f(xyz[n][3], t),
 * on the calling thread, whenever a residual is formed at a new time: the
 * time of mcfd_set_time for mcfd_get_residual_column, and t, t + dt and
 * t + dt / 2 for the stages of a fused RK3 step (temporal/rk.hpp:31-49).
 * f = NULL restores the tag's constant state.  mcfd_boundary_points gives the
 * tag's points in the order f receives them ([n][3]; xyz may be null to query
 * n). */
typedef void (*mcfd_bc_function)(int64_t n, const double *xyz, double t, double *states, void *user);
int mcfd_boundary_points(const mcfd_solver *h, int32_t tag, int64_t *n_points, double *xyz);
int mcfd_set_boundary_function(mcfd_solver *h, int32_t tag, mcfd_bc_function f, void *user);
/* A boundary function of the point only (the f(x, t) of fem.hpp:196-222 that
 * ignores t), given as its values at the tag's points ([n][5] in the order of
 * mcfd_boundary_points): evaluated once instead of at every stage time.
 * Replaces the tag's function; uploaded with the next residual. */
int mcfd_set_boundary_values(mcfd_solver *h, int32_t tag, const double *states);

/* limiter::Reconstruct(part, limiter) (limiter/reconstruct.hpp:22-54): one
 * pass of the handle's limiter over the current solution, as the reference's
 * driver runs after Approximate (twice on tetrahedra,
 * demo/euler/sourceless.cpp:139-147).  MCFD_ESTATE without a limiter. */
int mcfd_limiter_reconstruct(mcfd_solver *h);

/* Start-up limiters of nodal (FR_GENERAL) solutions, applied to the current
 * solution:
 *   AVERAGE  limiter/average.hpp:11-28: every node of a cell takes the cell's
 *            average (the global-weight quadrature over its solution points);
 *   MAJORITY limiter/majority.hpp:14-42 (degree >= 2, else no change): every
 *            node takes, per component, the value that more than half of the
 *            cell's nodes hold; MCFD_EINVAL, and no change, where a component
 *            has none (the reference asserts).
 * MCFD_ENOTSUP for DG_GENERAL handles (the reference's start-up limiters need
 * nodal values). */
#define MCFD_STARTUP_AVERAGE 0
#define MCFD_STARTUP_MAJORITY 1
int mcfd_startup_limiter(mcfd_solver *h, int32_t kind);

/* Troubled-cell flags of the last limiter application (1 = reconstructed),
 * one byte per cell, and their count; limiter/reconstruct.hpp:29-34.  Either
 * output may be null. */
int mcfd_limiter_flags(mcfd_solver *h, uint8_t *flags, int64_t *n_troubled);

/* ---- face fluxes at independent points
 * The device code of the stage kernels evaluated on caller-given state
 * pairs (host arrays), for checking the solvers on known answers:
 *   kind UPWIND: rotated::Euler<...>::GetFluxUpwind (riemann/rotated/euler.hpp:
 *     153-159) with the solver `riemann`;
 *   kind INVISCID_WALL: GetFluxOnInviscidWall (euler.hpp:160-174) of `left`;
 *   kind SUPERSONIC_OUTLET: GetFluxOnSupersonicOutlet (euler.hpp:144-151) of `left`;
 *   kind SUBSONIC_INLET / SUBSONIC_OUTLET: GetFluxOnSubsonicInlet / Outlet
 *     (euler.hpp:181-280) of `left` with the given values in `right`.
 * form FR: the FR stage kernels' evaluation (Roe with one reciprocal of rho per
 * side); form DG: the modal-DG face pass's.  frames[n][9] holds the rows
 * (n, t, b) of each point's frame; left, right [n][5] conservative states
 * (right may be null for INVISCID_WALL and SUPERSONIC_OUTLET); flux [n][5] out.  exact_star[n]
 * (in/out, may be null = fresh solvers) is each point's exact-solver state:
 * 1 if the star velocity of its last ordinary solve was positive
 * (riemann/euler/exact.hpp:29,116), which the vacuum branches reuse for the
 * tangential velocities (exact.hpp:49-57). */
#define MCFD_FLUX_UPWIND 0
#define MCFD_FLUX_INVISCID_WALL 1
#define MCFD_FLUX_SUPERSONIC_OUTLET 2
#define MCFD_FLUX_SUBSONIC_INLET 3
#define MCFD_FLUX_SUBSONIC_OUTLET 4
#define MCFD_FLUX_FORM_FR 0
#define MCFD_FLUX_FORM_DG 1
int mcfd_face_fluxes(int32_t riemann, int32_t form, int32_t kind, int64_t n, const double *frames,
                     const double *left, const double *right, uint8_t *exact_star, double *flux);

/* mcfd_face_fluxes without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops and harnesses that forbid hipMalloc).  The points are staged through
 * a fixed device buffer that the library holds on each GPU until the process
 * ends (about 3 MiB), so a call takes at most MCFD_NOALLOC_MAX_POINTS points
 * and returns MCFD_EINVAL beyond that; it runs on the current device's
 * default stream and synchronizes it.  Calls that overlap on one GPU share
 * the buffer and race.  Same arguments and results as mcfd_face_fluxes. */
#define MCFD_NOALLOC_MAX_POINTS 16384
int mcfd_face_fluxes_noalloc(int32_t riemann, int32_t form, int32_t kind, int64_t n, const double *frames,
                             const double *left, const double *right, uint8_t *exact_star, double *flux);

/* Stream synchronisation and per-kernel timing of n RK3 steps with HIP
 * events on the handle's stream: total ms, summed ms of the stage kernels,
 * and (if stage_ms is not null) the ms of each of the 3 * n_steps stage
 * kernels in launch order. */
int mcfd_synchronize(mcfd_solver *h);
int mcfd_time_rk3(mcfd_solver *h, double dt, int32_t n_steps,
                  double *total_ms, double *stage_kernel_ms, double *stage_ms);


/* ---- batched form
 * Many independent cases (own mesh, own degree 1..3, own BC table) advanced
 * together: cases are grouped by degree, each SSP-RK3 stage is one
 * persistent launch over the blocks of every group, and the three stages of
 * a step are captured once in a hipGraph and replayed.  Each case keeps its
 * own time step (the reference advances one System per case through
 * RungeKutta<3>::Update, temporal/rk.hpp:31-49).  No reference counterpart
 * exists for the batching itself. */
typedef struct mcfd_batch mcfd_batch;
/* bcs[bc_offsets[c] .. bc_offsets[c+1]) is case c's BC table, indexed by its
 * mesh's boundary tags; scheme->degree is ignored in favour of degrees[c]. */
int mcfd_batch_create(int32_t n_cases, const mcfd_mesh_desc *meshes, const int32_t *degrees,
                      const mcfd_scheme_desc *scheme, const mcfd_bc_desc *bcs,
                      const int32_t *bc_offsets, mcfd_batch **out);
int mcfd_batch_destroy(mcfd_batch *b);
int mcfd_batch_case_sizes(const mcfd_batch *b, int32_t case_id, int64_t *n_cells, int32_t *n_per_cell);
/* Solution points of one case, [n_cells][N][3]. */
int mcfd_batch_sample_points(const mcfd_batch *b, int32_t case_id, double *xyz);
/* System::SetSolutionColumn / GetSolutionColumn of one case. */
int mcfd_batch_set_solution_column(mcfd_batch *b, int32_t case_id, const double *column);
int mcfd_batch_get_solution_column(mcfd_batch *b, int32_t case_id, double *column);
/* FiniteElement::GetTimeStep (fem.hpp:417-435) of every case at CFL `cfl`. */
int mcfd_batch_get_time_steps(mcfd_batch *b, double cfl, double *dt_per_case);
/* Time step of every case used by the following steps. */
int mcfd_batch_set_time_steps(mcfd_batch *b, const double *dt_per_case);
/* Autotuner: times every tile variant (cells per workgroup) of every degree
 * group in the persistent launch, then the persistent single launch against
 * one launch per degree group, keeps the fastest and re-captures the step
 * graph.  tiles[g] = 16 * Q + variant for each degree group in ascending
 * degree; *mode = 1 persistent, 0 per group (either output may be null). */
int mcfd_batch_autotune(mcfd_batch *b, int32_t repeats, int32_t *tiles, int32_t *mode);
/* Fixes the launch form instead of the autotuner's choice and re-captures
 * the step graph: mode 1 = one persistent launch per stage, 0 = one launch
 * per degree group; tiles[g] (0..2, degree groups in ascending degree) is the
 * tile variant of the persistent launch, or null to keep the current ones. */
int mcfd_batch_set_mode(mcfd_batch *b, int32_t mode, const int32_t *tiles);
/* n_steps replays of the captured SSP-RK3 step (stream-ordered). */
int mcfd_batch_rk3_steps(mcfd_batch *b, int32_t n_steps);
int mcfd_batch_synchronize(mcfd_batch *b);
int mcfd_batch_time_rk3(mcfd_batch *b, int32_t n_steps, double *total_ms);

#ifdef __cplusplus
}
#endif

#endif /* MINICFD_H_ */
