/*
 * jaxfluids_amd.h -- C ABI of the MI355X (gfx950) convective-flux engine.
 *
 * Plain C: pointers, sizes and two POD structs; no C++ or torch types cross the
 * boundary.  All array pointers are DEVICE pointers.  Field buffers use the
 * JAX-Fluids layout (nvar, Nx+2nh, Ny+2nh, Nz+2nh), C-contiguous, z fastest; an
 * inactive axis has extent 1 and no halos (initialization/helper_functions.py:49).
 * Face arrays along `axis` are (nvar, N_axis+1, interior extents of the other
 * axes) (solvers/space_solver.py:142-146).  `stream` is a hipStream_t (NULL =
 * default stream).  Every call is asynchronous on `stream`; none allocates
 * device memory; scratch comes from the caller.
 *
 * Status codes: 0 ok, JFA_EINVAL bad argument/shape, JFA_EUNSUPPORTED option not
 * built, JFA_ELAUNCH a HIP launch failed.  No C++ exception crosses the ABI.
 *
 * Citations are relative to the JAX-Fluids v0.2.1 tree (src/jaxfluids/...).
 */
#ifndef JAXFLUIDS_AMD_H
#define JAXFLUIDS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JFA_OK 0
#define JFA_EINVAL 1
#define JFA_EUNSUPPORTED 2
#define JFA_ELAUNCH 3

/* reconstruction_variable (solvers/__init__.py:9-12) */
#define JFA_PRIMITIVE 0
#define JFA_CONSERVATIVE 1
#define JFA_CHAR_PRIMITIVE 2
#define JFA_CHAR_CONSERVATIVE 3
/* reconstruction_stencil (stencils/reconstruction/shock_capturing/weno/__init__.py:26-41,
 * .../teno/__init__.py:15-25).  Every stencil but WENO7-JS reads at most the 6-cell window of a
 * face and runs on buffers with nh >= 3; WENO7-JS reads 8 cells and needs nh >= 4 (JFA_EINVAL
 * otherwise).  WENO7-JS is built for the Godunov flux (any Riemann solver), not for flux
 * splitting or the batched stage (JFA_EUNSUPPORTED). */
#define JFA_WENO5_JS 0
#define JFA_WENO5_Z 1
#define JFA_WENO1 2
#define JFA_WENO3_JS 3
#define JFA_WENO3_Z 4
#define JFA_TENO5 5
#define JFA_TENO5_A 6
#define JFA_WENO6_CU 7
#define JFA_TENO6 8
#define JFA_WENO6_CUM1 9
#define JFA_WENO6_CUM2 10
#define JFA_TENO6_A 11
#define JFA_WENO7_JS 12
/* riemann_solver (solvers/riemann_solvers/__init__.py:16-25) */
#define JFA_HLLC 0
#define JFA_RUSANOV 1
#define JFA_HLL 2
#define JFA_HLLC_LM 3
#define JFA_AUSMP 4
/* convective_solver FLUX-SPLITTING (solvers/convective_fluxes/flux_splitting_scheme.py:21-111) in
 * place of a Riemann solver, with the eigenvalues its flux_splitting names
 * (solvers/__init__.py:1-3; eigendecomposition.py:672-696): the window's conservatives and
 * physical fluxes are projected on the conservative eigenvectors of the frozen state
 * (frozen_state), split by those eigenvalues and reconstructed with `stencil` (the positive part
 * from the left, the negative from the right).  recon_var, signal_speed and the interpolation
 * limiter do not apply, as in the reference. */
#define JFA_FS_ROE 5
#define JFA_FS_CLLF 6
#define JFA_FS_LLF 7
/* signal_speed (solvers/riemann_solvers/__init__.py:27-34; DAVIS2 is not callable from HLLC there) */
#define JFA_SS_ARITHMETIC 0
#define JFA_SS_RUSANOV 1
#define JFA_SS_DAVIS 2
#define JFA_SS_EINFELDT 3
#define JFA_SS_TORO 4
#define JFA_SS_DAVIS2 5  /* plain function only (jfa_signal_speeds) */
/* frozen_state (solvers/__init__.py:5-7) */
#define JFA_FROZEN_ARITHMETIC 0
#define JFA_FROZEN_ROE 1
/* boundary types handled on the path (halos/outer/material.py:142-144) */
#define JFA_BC_PERIODIC 0
#define JFA_BC_ZEROGRADIENT 1
#define JFA_BC_SYMMETRY 2
/* storage / compute precision */
#define JFA_F64 0
#define JFA_F32 1
/* positivity.flux_limiter and flux_partition (solvers/positivity/__init__.py:1-7) */
#define JFA_FLUX_LIMITER_NONE 0
#define JFA_FLUX_LIMITER_SIMPLE 1
#define JFA_FLUX_LIMITER_NASA 2
#define JFA_FLUX_LIMITER_HAS 3
#define JFA_PARTITION_UNIFORM 0
#define JFA_PARTITION_CELLSIZE 1
#define JFA_PARTITION_WAVESPEED 2

/* Uniform grid of one domain (domain/domain_information.py:68-300). */
typedef struct jfa_grid {
    int32_t n[3];   /* interior cells per axis; 1 marks an inactive axis */
    int32_t nh;     /* halo cells on each side of an active axis (>= 3 for WENO5) */
    double dx[3];   /* cell size per axis, (b - a) / n (domain/mesh_creation/homogenous.py:13) */
} jfa_grid;

/* Laws of the transport properties (materials/single_materials/material.py:82-124). */
enum {
    JFA_LAW_POWER = 0,       /* ref (T / T_ref)^param; param = 0: the constant ref (CUSTOM value,
                                or a CUSTOM lambda that is a power law, e.g. "lambda T: 0.0028 * T**0.76") */
    JFA_LAW_SUTHERLAND = 1,  /* ref ((T_ref + param) / (T + param)) (T / T_ref)^(3/2) (SUTHERLAND, :94-96) */
    JFA_LAW_PRANDTL = 2      /* thermal conductivity only: cp mu / prandtl (PRANDTL, :115-116) */
};

/* Transport properties of the dissipative fluxes, second-order central (solvers/source_term_solver.py:
 * 188-350; flux assembly space_solver.py:565-587; dt limits time_integration/time_step_size.py:112-135):
 *   mu(T)    by mu_law    from (mu_ref, T_ref, mu_param)
 *   kappa(T) by kappa_law from (kappa_ref, kappa_T_ref, kappa_param), or cp mu / prandtl;
 *            cp = gamma R / (gamma - 1) (ideal_gas.py:32)
 *   T = p / (rho R)                      (ideal_gas.py:81-83)
 * The dissipative fluxes use central stencils across face-adjacent cells and their transverse
 * neighbours, halos and edge halos included on non-periodic domains. */
typedef struct jfa_transport {
    int32_t viscous;        /* 1: viscous stresses tau and work u.tau (active_physics.is_viscous_flux) */
    int32_t heat;           /* 1: heat flux q = -kappa dT/dx (active_physics.is_heat_flux) */
    double R;               /* specific gas constant */
    int32_t mu_law;         /* JFA_LAW_POWER / JFA_LAW_SUTHERLAND */
    int32_t kappa_law;      /* JFA_LAW_POWER / JFA_LAW_SUTHERLAND / JFA_LAW_PRANDTL */
    double mu_ref;
    double T_ref;
    double mu_param;        /* POWER: exponent; SUTHERLAND: constant C */
    double bulk_viscosity;  /* mu_b of tau's (mu_b - 2/3 mu) div u term (source_term_solver.py:527) */
    double prandtl;         /* JFA_LAW_PRANDTL */
    double kappa_ref;
    double kappa_T_ref;
    double kappa_param;
    int32_t halos;          /* jfa_compute_rhs only: 1 = the stencils read the buffer's face and edge
                               halos, which the caller filled (non-periodic boundaries; the edge halos
                               as jfa_edge_halo_update fills them); 0 = they wrap periodically.  The
                               fused stage decides this from its boundary types itself. */
} jfa_transport;

/* Numerical scheme of the high-order Godunov flux (data_types/numerical_setup/conservatives.py:35-42). */
typedef struct jfa_scheme {
    int32_t recon_var;     /* JFA_PRIMITIVE .. JFA_CHAR_CONSERVATIVE */
    int32_t stencil;       /* JFA_WENO5_JS ... JFA_TENO6_A */
    int32_t riemann;       /* JFA_HLLC / JFA_RUSANOV / JFA_HLL / JFA_HLLC_LM / JFA_AUSMP, or JFA_FS_* */
    int32_t signal_speed;  /* JFA_SS_* (HLLC only) */
    int32_t frozen_state;  /* JFA_FROZEN_* (characteristic variables only) */
    int32_t dtype;         /* JFA_F64 (float64 buffers) or JFA_F32 (float32 buffers) */
    double gamma;          /* ideal-gas specific heat ratio (materials/single_materials/ideal_gas.py:29) */
    double eps_stencil;    /* WENO epsilon, 1e-30 (config/precision.py:39,53) */
    const jfa_transport* transport;  /* dissipative fluxes of the fused stage / RHS; NULL: inviscid */
    /* Positivity limiter of the reconstructed face states (positivity.is_interpolation_limiter,
     * solvers/positivity/limiter_interpolation.py:86-208, single phase): a face state with density
     * below limiter_eps_density takes the upwind cell's density and pressure (all five primitives
     * with limit_velocity), then one with pressure below limiter_eps_pressure likewise; the face's
     * conservatives are recomputed from the limited primitives.  Epsilons: 1e-12 / 1e-10 in double
     * precision, 1e-6 / 1e-6 in single (config/precision.py:40,54). */
    int32_t interpolation_limiter;
    int32_t limit_velocity;
    double limiter_eps_density;
    double limiter_eps_pressure;
    /* Positivity flux limiter of the convective face flux (positivity.flux_limiter SIMPLE / NASA /
     * HAS, solvers/positivity/limiter_flux.py:147-316; applied by the stage and RHS entry points,
     * which know the step's dt, as SpaceSolver.compute_rhs_xi does, space_solver.py:530-541): where
     * the face's pseudo-updated cell states U_R + 2 lam F, U_L - 2 lam F (NASA: F minus each cell's
     * physical flux), lam = dt sigma / dx, have a density, then a pressure, below the epsilons, the
     * flux becomes the first-order HLLC-Einfeldt flux F_pos of the two cells (SIMPLE, NASA) or the
     * blend (1 - theta) F_pos + theta F whose linearised density, then pressure, of the offending
     * state is the epsilon (HAS, :200-251).  sigma: the number of active axes (UNIFORM),
     * sum(1/dx_a) / (1/dx) (CELLSIZE), or sum(m_a/dx_a) / (m/dx) with m_a the maximum of |u_a| + c
     * over the stage's input primitives, halos included (WAVESPEED, :717-726; the halo corners hold
     * the initialization's machine-epsilon state, so m_a >= eps + sqrt(gamma) on 2-D and 3-D
     * grids).  Epsilons 1e-12 / 1e-10 in double precision, 1e-6 / 1e-6 in single
     * (config/precision.py:41,55). */
    int32_t flux_limiter;    /* JFA_FLUX_LIMITER_* */
    int32_t flux_partition;  /* JFA_PARTITION_* */
    double flux_limiter_eps_density;
    double flux_limiter_eps_pressure;
} jfa_scheme;

/* Library version string; never NULL. */
const char* jfa_version(void);

/* Last HIP error string seen by this thread (for diagnostics after JFA_ELAUNCH). */
const char* jfa_last_error(void);

/* Cell-face reconstruction of every variable of `buffer` along `axis` with any JFA_* stencil:
 * j = 0 left state, j = 1 right state of each face.
 * Replaces WENO5JS.reconstruct_xi (stencils/reconstruction/shock_capturing/weno/weno5_js.py:73-115),
 * WENO5Z.reconstruct_xi (.../weno/weno5_z.py:138-173), WENO1 / WENO3JS / WENO3Z
 * (.../weno/weno1_js.py:25-30, weno3_js.py:15-42, weno3_z.py:21-41), TENO5 / TENO5A
 * (.../teno/teno5.py:32-73, teno5_a.py:42-99), WENO6CU / WENO6CUM1 / WENO6CUM2 / TENO6 / TENO6A
 * (.../weno/weno6_cu.py:24-61, weno6_cum1.py:28-67, weno6_cum2.py:29-72, teno/teno6.py:27-67, teno6_a.py:40-131; the
 * CUM stencils read the grid's cell size along `axis` as reconstruct_xi's dx) with the slice tables of WENO5Base / WENO3Base / WENO1
 * (.../weno5_base.py:30, weno3_base.py:29, weno/weno1_js.py:22) and SpatialStencil.array_slices
 * (stencils/spatial_stencil.py:45-112).
 * buffer: (nvar, buffer shape) of `dtype`; out: (nvar, face shape). */
int jfa_weno5_reconstruct_xi(const void* buffer, int32_t nvar, const jfa_grid* grid, int32_t axis,
                             int32_t j, int32_t stencil, double eps, int32_t dtype, void* out,
                             void* stream);

/* Riemann flux from reconstructed face states, all (5, nfaces) contiguous.
 * Replaces RiemannSolver.solve_riemann_problem_xi (solvers/riemann_solvers/riemann_solver.py:52-102)
 * for HLLC (HLLC.py:80-126) and Rusanov (Rusanov.py:25-47), single phase.  A flux-splitting
 * scheme (JFA_FS_*) has no face states: JFA_EUNSUPPORTED. */
int jfa_riemann_xi(const void* W_L, const void* W_R, const void* U_L, const void* U_R, int64_t nfaces,
                   int32_t axis, const jfa_scheme* scheme, void* flux, void* stream);

/* Signal-speed estimates (S_L, S_R) of n left/right state pairs, elementwise: the plain functions
 * signal_speed_Arithmetic, _Rusanov, _Davis, _Davis_2, _Einfeldt, _Toro
 * (solvers/riemann_solvers/signal_speeds.py:10-157; registry DICT_SIGNAL_SPEEDS,
 * solvers/riemann_solvers/__init__.py:27-34) that Riemann solvers call, with the same formulas the
 * face kernels evaluate.  kind: JFA_SS_*.  Every array holds n contiguous values of `dtype`:
 * u_L, u_R (normal velocities) and a_L, a_R (speeds of sound) always; rho_L, rho_R for EINFELDT,
 * TORO and DAVIS2; p_L, p_R for TORO; H_L, H_R (total enthalpies) for DAVIS2.  Inputs a kind does
 * not use may be NULL.  gamma: specific heat ratio (TORO, DAVIS2).  Outputs S_L, S_R: n values. */
int jfa_signal_speeds(int32_t kind, const void* u_L, const void* u_R, const void* a_L, const void* a_R,
                      const void* rho_L, const void* rho_R, const void* p_L, const void* p_R, const void* H_L,
                      const void* H_R, double gamma, int64_t n, int32_t dtype, void* S_L, void* S_R,
                      void* stream);

/* Numerical flux along `axis` on every face: reconstruction (4 variable modes), prim<->cons
 * conversion, the positivity limiter when the scheme enables it, and Riemann solve.
 * flux: (5, face shape).  limiter_count: NULL or one device int64 to which the number of limiter
 * replacements is added (the count_interpolation_limiter the reference returns).
 * Replaces HighOrderGodunov.compute_flux_xi (solvers/convective_fluxes/high_order_godunov.py:117-231)
 * with reconstruct_xi (:233-419), and for riemann = JFA_FS_* FluxSplittingScheme.compute_flux_xi
 * (solvers/convective_fluxes/flux_splitting_scheme.py:65-111; limiter_count is then unchanged). */
int jfa_compute_flux_xi(const void* W, const void* U, const jfa_grid* grid, const jfa_scheme* scheme,
                        int32_t axis, void* flux, int64_t* limiter_count, void* stream);

/* Right-hand side 0.0 + rhs_x + rhs_y + rhs_z, rhs_xi = (1/dx)(F[:-1]-F[1:]).
 * rhs: (5, Nx, Ny, Nz).  workspace: >= jfa_rhs_workspace_bytes(grid, dtype) bytes (FP64 partial RHS and
 * the WAVESPEED maxima).
 * Replaces SpaceSolver.compute_rhs (solvers/space_solver.py:151-453) and compute_rhs_xi (:456-674):
 * the convective flux (positivity-limited with the step's time step dt when
 * scheme->flux_limiter is set, as compute_rhs's physical_timestep_size; dt is unused otherwise),
 * plus the dissipative fluxes when scheme->transport is set (the caller's boundaries must then be
 * periodic). */
int jfa_compute_rhs(const void* W, const void* U, const jfa_grid* grid, const jfa_scheme* scheme, double dt,
                    void* rhs, void* workspace, size_t workspace_bytes, void* stream);
size_t jfa_rhs_workspace_bytes(const jfa_grid* grid, int32_t dtype);

/* One TVD-RK3 stage: for stage > 0, U = c0*U + c1*U_n (RK3.py:31-51); then
 * U[interior] += (dt*m_stage)*rhs (RK3.py:53-62, time_integrator.py:34-58); then
 * W[interior] = prim(U) (equation_manager.py:108-176, simulation_manager.py:943-944).
 * dt is read from device memory `dt_dev` when non-NULL, else `dt_host` is used.
 * Replaces TimeIntegrator.perform_stage_integration (time_integration/time_integrator.py:108-227). */
int jfa_rk3_stage(void* U, const void* U_n, const void* rhs, void* W, const jfa_grid* grid,
                  const jfa_scheme* scheme, int32_t stage, const double* dt_dev, double dt_host,
                  void* stream);

/* Face-halo fill of W and U = cons(W) in the halos for PERIODIC / ZEROGRADIENT / SYMMETRY;
 * bc[6] in the order east, west, north, south, top, bottom (JFA_BC_*).
 * Replaces HaloManager.perform_halo_update_material (halos/halo_manager.py:146-234) ->
 * BoundaryConditionMaterial.face_halo_update (halos/outer/material.py:94-287, miscellaneous :868-894). */
/* Edge halos of W, and U = cons(W) there, for viscous or heat fluxes on 2-D and 3-D grids
 * (halo_manager.py:125-129,183-186; BoundaryConditionMaterial.edge_halo_update,
 * halos/outer/material.py:289-380): on each active edge the first face (in the reference's edge
 * order, the lower axis first) that is PERIODIC or SYMMETRY decides -- the periodic image, or the
 * mirror with that axis' velocity negated, of the other face's halo -- else the mean of the two
 * face halos moved nh cells into the interior.  The face halos must be filled already.  Vertex
 * halos are left alone. */
int jfa_edge_halo_update(void* W, void* U, const jfa_grid* grid, const int32_t bc[6], double gamma,
                         int32_t dtype, void* stream);
int jfa_halo_update(void* W, void* U, const jfa_grid* grid, const int32_t bc[6], double gamma,
                    int32_t dtype, void* stream);

/* CFL time step dt = CFL * dx_min / (max_cells sum_active(|u_i| + c) + eps), written to dt_dev[0]
 * (a device double); with a transport that has viscous (heat) fluxes, the minimum with
 * 3/14 dx_min^2 / (max mu/rho + eps) (and 0.1 dx_min^2 / (max kappa/(rho cp) + eps)) before the
 * CFL factor.  transport may be NULL.  workspace: >= 32 bytes of device memory.
 * Replaces compute_time_step_size (time_integration/time_step_size.py:15-157). */
int jfa_compute_dt(const void* W, const jfa_grid* grid, double gamma, double cfl, double eps,
                   int32_t dtype, const jfa_transport* transport, double* dt_dev, void* workspace,
                   void* stream);

/* Whole-buffer conversions (every cell, halos included).
 * Replace EquationManager.get_conservatives_from_primitives (equation_manager.py:44-106) and
 * get_primitives_from_conservatives (:108-176), single phase. */
int jfa_cons_from_prim(const void* W, void* U, int64_t ncells, double gamma, int32_t dtype, void* stream);
int jfa_prim_from_cons(const void* U, void* W, int64_t ncells, double gamma, int32_t dtype, void* stream);

/* One fused RK3 stage of the whole hot path: face fluxes of every active axis, flux divergence
 * (accumulated in the reference's axis order), stage blend and update, W = prim(U), and the
 * face-halo fill of W and U for the boundary types bc[6] -- the work of jfa_compute_rhs +
 * jfa_rk3_stage + jfa_halo_update without face-flux arrays: one marching sweep per active
 * non-innermost axis and one innermost-axis sweep that finishes the stage.
 * If dt_next_dev is non-NULL, the CFL time step of the new state (as jfa_compute_dt, with
 * `cfl` and `eps`) is written to dt_next_dev[0]; it may alias dt_dev.
 * W_out/U_out must not alias W_in/U_in; U_out may alias U_n.
 * Replaces one pass of SimulationManager.do_runge_kutta_stages (simulation_manager.py:770-1047):
 * compute_rhs (space_solver.py:151-453) -> perform_stage_integration (time_integrator.py:108-227)
 * -> get_primitives_from_conservatives (:943-944) -> perform_halo_update_material (:961-970),
 * and at the last stage compute_time_step_size (time_step_size.py:15-157).
 * workspace: >= jfa_stage_workspace_bytes(grid, scheme) bytes. */
int jfa_rk3_stage_fused(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out,
                        const jfa_grid* grid, const jfa_scheme* scheme, const int32_t bc[6],
                        int32_t stage, const double* dt_dev, double dt_host, double cfl, double eps,
                        double* dt_next_dev, void* workspace, size_t workspace_bytes, void* stream);
size_t jfa_stage_workspace_bytes(const jfa_grid* grid, const jfa_scheme* scheme);

/* jfa_rk3_stage_fused without allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The stage's scratch (FP64 RHS and CFL partials) is a
 * library-owned device array of JFA_NOALLOC_SCRATCH_BYTES per GPU, held from the library's load
 * until the process ends; grids whose jfa_stage_workspace_bytes exceed it return JFA_EINVAL
 * (about 1.6 million interior cells).  Same arguments and results as jfa_rk3_stage_fused without
 * the workspace.  Calls share that GPU's array: calls that overlap on one GPU (other streams,
 * other threads) race; calls in one stream order are safe. */
#define JFA_NOALLOC_SCRATCH_BYTES (64u << 20)
int jfa_rk3_stage_fused_noalloc(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out,
                                const jfa_grid* grid, const jfa_scheme* scheme, const int32_t bc[6],
                                int32_t stage, const double* dt_dev, double dt_host, double cfl, double eps,
                                double* dt_next_dev, void* stream);

/* ---- Reverse mode of the RK3 stage at fixed dt (FP64; every reconstruction variable, the
 * characteristic ones with the ARITHMETIC or ROE frozen state; WENO5-JS/-Z, WENO1, WENO3-JS/-Z,
 * TENO5, TENO5-A;
 * HLLC, HLL or HLLC-LM with ARITHMETIC/RUSANOV/DAVIS/EINFELDT/TORO speeds, AUSM+, Rusanov or
 * the flux-splitting scheme (ROE / CLLF / LLF); with
 * scheme->transport, the
 * viscous and heat fluxes too, on periodic domains).  The reference obtains this
 * derivative by jax.grad through SimulationManager.do_runge_kutta_stages
 * (simulation_manager.py:770-1047, RK3.py:27-62); here it is a hand-written VJP.
 * For the forward stage U_out = c0 U_in + c1 U_n + m dt R(prim(U_in)) with W_in = prim(U_in)
 * (face halos filled), and lam_out = dL/dU_out (interior cells, buffer layout), it adds
 *   lam_in += c0 lam_out + (dW/dU)^T halo^T (dR/dW)^T (m dt lam_out)      (c0 = 1 at stage 0)
 *   lam_n  += c1 lam_out                                                  (stages 1, 2)
 * on interior cells; halos of lam_* are neither read nor written.  lam_in, lam_n and lam_out
 * must be distinct.  workspace: >= jfa_stage_vjp_workspace_bytes(grid, scheme) bytes (the face-halo
 * cotangents, plus 9 doubles per interior cell with a transport). */
size_t jfa_stage_vjp_workspace_bytes(const jfa_grid* grid, const jfa_scheme* scheme);
/* Tuning of jfa_rk3_stage_vjp (process-wide): an axis other than the innermost is swept by the
 * marching transpose kernel (coalesced columns, chunks of 16 cells) when that launch has at least
 * `waves` wavefronts, else by the lane-parallel one (one face per lane, strided loads).  Default
 * 2048; 0 forces marching, a huge value lane-parallel sweeps.  No reference counterpart (the
 * reference differentiates with jax.grad).  Returns JFA_EINVAL for waves < 0. */
int jfa_adjoint_set_march_min_waves(int64_t waves);
int jfa_rk3_stage_vjp(const void* W_in, const void* lam_out, void* lam_in, void* lam_n, const jfa_grid* grid,
                      const jfa_scheme* scheme, const int32_t bc[6], int32_t stage, double dt, void* workspace,
                      size_t workspace_bytes, void* stream);

/* ---- Ensembles of independent 2-D domains (z inactive), advanced by one launch per RK3
 * stage for all cases together.  Each case is a (5, Nx+2nh, Ny+2nh, 1) block inside pooled
 * buffers (W, U and their stage copies share one layout); cases may differ in size,
 * boundary types and time step.  This batching has no counterpart in the reference, which
 * advances one domain per SimulationManager (simulation_manager.py:670-1077); every case
 * follows that function's stage arithmetic exactly. */
typedef struct jfa_case2d {
    int32_t n[2];      /* interior cells along x and y */
    int32_t nh;        /* halo cells (>= 3) */
    int32_t bc[4];     /* east, west, north, south (JFA_BC_*) */
    int32_t reserved;
    double dx[2];      /* cell sizes along x and y */
    int64_t offset;    /* element offset of the case's block in the pooled buffers */
} jfa_case2d;

/* Workspace (device) for `ncases` cases tiled in x-chunks of `chunk` planes (1..64); 0 if the
 * table is invalid (nh < 3, fewer than 2 cells on an axis, a periodic or symmetric axis shorter
 * than nh, a case block of 4 GiB or more). */
size_t jfa_batch2d_workspace_bytes(const jfa_case2d* cases, int32_t ncases, int32_t chunk);

/* Writes the case table and the (case, x-chunk, 64-cell y-tile) work items into the workspace;
 * synchronises `stream` before returning (host tables are temporaries). */
int jfa_batch2d_plan(const jfa_case2d* cases, int32_t ncases, int32_t chunk, void* workspace,
                     size_t workspace_bytes, void* stream);

/* One RK3 stage of every case in one launch (the arithmetic of jfa_rk3_stage_fused per case):
 * dt_dev holds one time step per case; if dt_next_dev is non-NULL the next CFL time step of every
 * case is written there (may alias dt_dev).  The plan must have been made for the same `chunk`. */
int jfa_batch2d_stage(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out,
                      const jfa_case2d* cases, int32_t ncases, int32_t chunk, const jfa_scheme* scheme,
                      int32_t stage, const double* dt_dev, double cfl, double eps, double* dt_next_dev,
                      void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* JAXFLUIDS_AMD_H */
