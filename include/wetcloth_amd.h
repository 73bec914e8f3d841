/*
 * wetcloth_amd -- C ABI of the MI355X (gfx950) liquid-substep library.
 *
 * Plain pointers and sizes only.  Host pointers everywhere: the context owns
 * all device memory and keeps particles and grids resident in HBM between
 * calls; only residual/iteration counts and what the caller explicitly
 * downloads cross the host/device boundary.  One HIP stream per context; a
 * context is not thread-safe (like the reference's TwoDScene/PCGSolver).
 *
 * Status: 0 = OK, 1 = PCG not converged (PCGSolver::solve returned false),
 * < 0 = error (wc_last_error has the message).  Nothing throws across the ABI.
 *
 * Reference paths below are relative to libWetCloth/Core in the reference.
 */
#ifndef WETCLOTH_AMD_H
#define WETCLOTH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t wc_status;
#define WC_OK 0
#define WC_NOT_CONVERGED 1
#define WC_ERR_ARG (-1)
#define WC_ERR_HIP (-2)
#define WC_ERR_STATE (-3)
#define WC_ERR_NOMEM (-4)

typedef struct wc_ctx wc_ctx;

/* arithmetic of the substep */
#define WC_FP64 0  /* the reference's precision (scalar = double, MathDefs.h:24) */
#define WC_MIXED 1 /* no reference counterpart: FP32 particle state and
                      * transfers (P2G sums flushed into the FP64 grid), FP64
                      * grid, assembly and MIC(0) factor; the pressure PCG runs
                      * on FP32 vectors and rounded factor/matrix copies with
                      * FP64 inner products and alpha/beta, inside FP64
                      * iterative refinement on r = b - A p, and converges on
                      * the FP64 residual.  Systems too large for the pipelined
                      * sweeps are solved in FP64. */

/* Scene description: the LiquidInfo / bucket fields the liquid path reads
 * (App/TwoDSceneXMLParser.cpp:1849-1904, TwoDScene::setBucketInfo
 * TwoDScene.cpp:1614).  The container is the box [0, n*dx] per axis with
 * solid outside (a DFT_BOX solid with inside=1, DistanceFields.cpp:68-78). */
typedef struct wc_desc {
  int32_t ni, nj, nk;               /* container cells per axis */
  double dx;                        /* cell size = bucket size / 4 */
  double liquid_density;            /* kg/m^3 */
  double flip_coeff;                /* LiquidInfo::flip_coeff (default 0.99) */
  double liquid_boundary_friction;  /* LiquidInfo (default 1) */
  double gravity[3];                /* SimpleGravityForce */
  double particle_mass;             /* uniform m_fluid_m of every particle */
  double particle_volume;           /* uniform m_fluid_vol */
  double particle_radius;           /* uniform m_radius */
  int32_t solid_projection;         /* run solidProjection after advection */
  int32_t precision;                /* WC_FP64 or WC_MIXED */
  int64_t max_particles;            /* device capacity reserved at create */
} wc_desc;

/* Context lifetime. */
wc_status wc_create(const wc_desc* desc, wc_ctx** out);
void wc_destroy(wc_ctx* ctx);
const char* wc_last_error(const wc_ctx* ctx);

/* Particle state.  Replaces the TwoDScene members the transfers read and
 * write (TwoDScene.h:932-960): m_x (AoS, 4 doubles per particle),
 * m_fluid_v (AoS 4) and m_fB (Eigen 3N x 3, column-major).  Any pointer may
 * be NULL (that quantity is zeroed on set / skipped on get). */
wc_status wc_set_particles(wc_ctx* ctx, int64_t n, const double* x4,
                           const double* fluid_v4, const double* fB);
wc_status wc_get_particles(wc_ctx* ctx, double* x4, double* fluid_v4,
                           double* fB);
/* Same, structure-of-arrays: x[d*n+p], v[d*n+p], C[(r*3+c)*n+p] = fB(r,c). */
wc_status wc_set_particles_soa(wc_ctx* ctx, int64_t n, const double* x,
                               const double* v, const double* C);
wc_status wc_get_particles_soa(wc_ctx* ctx, double* x, double* v, double* C);
wc_status wc_num_particles(const wc_ctx* ctx, int64_t* n);
/* Synthetic dam-break block [lo, hi) at 8 particles per cell, jittered in
 * 2x2x2 sub-cells, optional smoothed Gaussian velocity of std sigma.  No
 * reference counterpart (the reference samples liquid from distance fields,
 * TwoDScene.cpp:5527-5608); seeded on the device so large scenes need no
 * host staging. */
wc_status wc_seed_block(wc_ctx* ctx, const double lo[3], const double hi[3],
                        uint64_t seed, double sigma, int64_t* n_out);

/* ---- a static porous solid in the liquid path (the mixture terms) ----
 * The LiquidInfo fields the mixture terms read (TwoDScene.h:35-95; the
 * parser's defaults, TwoDSceneXMLParser.cpp:1849-1904, are CGS values). */
typedef struct wc_mix_params {
  int32_t use_drag, use_nonlinear_drag;
  double yarn_diameter, viscosity, yazdchi_power;
  double pore_radius, surf_tension_coeff, rest_contact_angle;
} wc_mix_params;
/* The elasto particles (the cloth's vertices) of a scene whose solid is at
 * rest and not solved (LiquidInfo::solve_solid = 0): replaces the elasto
 * entries of TwoDScene's m_x, m_m, m_fluid_m, m_fluid_vol, m_rest_vol,
 * m_rest_volume_fraction, m_orientation and m_shape_factor (TwoDScene.h:
 * 932-960).  Arrays in structure-of-arrays form: x and orientation 3*ne
 * (x[d*ne+p]), the others ne.  Every following substep then takes the
 * solid's P2G (TwoDScene.cpp:6686-6736), the cell-centred saturation and
 * volume fraction with the pore pressure (:5933-5992), the drag between the
 * phases (LinearizedImplicitEuler.cpp:846-1107 with TwoDScene.cpp:
 * 4042-4097), the pore pressure gradient (Pressure.cpp:26-142) and the
 * volume-fraction-weighted assembly and pressure application
 * (Pressure.cpp:444-791, 1137-1314).  The particles must lie in the
 * container box.  ne = 0 removes the solid.  WC_ERR_STATE while decisions
 * are recorded or replayed (the adjoint covers the liquid-only substep). */
wc_status wc_set_elasto_particles(wc_ctx* ctx, int64_t ne, const double* x, const double* m,
                                  const double* fluid_m, const double* fluid_vol,
                                  const double* rest_vol, const double* rest_volume_fraction,
                                  const double* orientation, const double* shape_factor,
                                  const wc_mix_params* prm);

/* ---- on-disk formats of the reference's serializer (App/TwoDSceneSerializer.cpp) ----
 * The particles are written in the caller's order (the order of
 * wc_set_particles), as doubles; the files are written before the call
 * returns (the reference writes from a detached thread). */

/* TwoDSceneSerializer::serializePositionOnly (:161-170) with
 * serialize_pos_subprog (:23-37): an int byte count and m_x (4 doubles per
 * particle: x, y, z and theta = 0), then an int byte count and m_d_gauss,
 * which is empty for a liquid-only scene (count 0).  WC_ERR_ARG if the byte
 * count does not fit an int (the reference's int overflows there). */
wc_status wc_serialize_position_only(wc_ctx* ctx, const char* path);
/* TwoDSceneSerializer::loadPosOnly (:172-183): reads such a file back into
 * the particle positions (velocities and affine matrices are kept).  The
 * file must hold 4 doubles per particle of the context and an empty
 * m_d_gauss block; anything else is WC_ERR_ARG (the reference reads the
 * stored byte count into m_x unchecked). */
wc_status wc_load_position_only(wc_ctx* ctx, const char* path);
/* The fluid file of TwoDSceneSerializer::serializeScene (:135-159):
 * serialize_subprog's fluid loop (:39-47) over updateFluid's data
 * (:242-255), one line "v x y z r" per particle, each number as an
 * iostream inserts a double at precision 8 (printf "%.8g"), r = the
 * context's particle radius. */
wc_status wc_serialize_fluid(wc_ctx* ctx, const char* path);

/* ---- the substep, phase by phase, in Core/WetClothCore.cpp:77-300 order ---- */

/* updateParticleBoundingBox + rebucketizeParticles (TwoDScene.cpp:1588-1708):
 * sort particles by bucket on the device. */
wc_status wc_rebucketize(wc_ctx* ctx);
/* resampleNodes bucket activation (TwoDScene.cpp:5069-5120), updateLiquidPhi
 * (:3286), updateSolidPhi (:7940), updateSolidWeights (:1871) and
 * TwoDScene::mapParticleNodesAPIC (TwoDScene.h:487, TwoDScene.cpp:6609). */
wc_status wc_map_particle_nodes(wc_ctx* ctx);
/* LinearizedImplicitEuler::stepVelocity (LinearizedImplicitEuler.cpp:224). */
wc_status wc_step_velocity(wc_ctx* ctx, double dt);
/* LinearizedImplicitEuler::projectFine -> pressure::solveNodePressure
 * (Pressure.cpp:285-830) with the MIC(0) PCGSolver<double>::solve
 * (PCGSolver/PCGSolver.h:260) at the solve call site (Pressure.cpp:797), or
 * with the reference's AMGPCGSolveSparse there when the "solver" option is 1. */
wc_status wc_project(wc_ctx* ctx, double dt, double tol, int32_t maxit,
                     double* residual_out, int32_t* iters_out);
/* applyPressureDragFluid (LinearizedImplicitEuler.cpp:2435), acceptVelocity
 * (:2176) and TwoDScene::constrainLiquidVelocity (TwoDScene.cpp:1827). */
wc_status wc_apply_pressure(wc_ctx* ctx, double dt);
/* TwoDScene::mapNodeParticlesAPIC() (TwoDScene.h:493, TwoDScene.cpp:
 * 6936-7090): the fluid particles' velocity and affine matrix from the
 * projected grid velocity (no argument, like the reference call). */
wc_status wc_map_node_particles(wc_ctx* ctx);
/* SceneStepper::advectScene(scene, dt) (SceneStepper.cpp:17-41), fluid
 * branch: x += v dt. */
wc_status wc_advect_scene(wc_ctx* ctx, double dt);
/* TwoDScene::solidProjection(dt) (TwoDScene.cpp:1776-1822): a particle whose
 * solid distance would fall below 0 is moved back along the distance
 * gradient.  As in the reference the corner-grid weights belong to the
 * position the transfer weights were computed at (before wc_advect_scene)
 * and the gradient points from the current one. */
wc_status wc_solid_projection(wc_ctx* ctx, double dt);
/* The three calls above in one pass over the particles (the same results,
 * bit for bit): mapNodeParticlesAPIC, advectScene and, when the context's
 * solid_projection is set, solidProjection (WetClothCore.cpp:248-262). */
wc_status wc_map_node_particles_advect(wc_ctx* ctx, double dt);
/* TwoDScene::getMaxFluidVelocity (TwoDScene.cpp:4180-4188): the largest
 * fluid particle speed. */
wc_status wc_max_fluid_velocity(wc_ctx* ctx, double* vmax);
/* One substep: the phases above in WetClothCore.cpp:109-262 order with the
 * fused G2P (one iteration of stepSystem's substep loop). */
wc_status wc_substep(wc_ctx* ctx, double dt, double tol, int32_t maxit,
                     double* residual_out, int32_t* iters_out);
/* WetClothCore::stepSystem(dt) (WetClothCore.cpp:45-300): splits dt into
 * max(1, ceil(dt / max_dt)) equal substeps, max_dt = min(3 dx / v_max,
 * 1/30) with v_max from getMaxFluidVelocity (:52-62), and runs wc_substep
 * on each.  Every substep runs even when a solve does not converge (the
 * reference prints a warning and goes on); the status is then
 * WC_NOT_CONVERGED.  residual_out is the last substep's, iters_out the sum
 * over the substeps. */
wc_status wc_step(wc_ctx* ctx, double dt, double tol, int32_t maxit,
                  double* residual_out, int32_t* iters_out);
/* Substeps of the last wc_step.  No reference counterpart (stepSystem
 * prints them). */
wc_status wc_step_substeps(const wc_ctx* ctx, int32_t* n);

/* ---- the adjoint of the substep (no reference counterpart: the reference
 * has no differentiation) ----
 * The substep is differentiated with its discrete decisions held fixed: the
 * activated buckets, each cell's liquid-phi particle (the particle whose
 * distance sets the cell's phi, or the default or solid value), the liquid
 * cells and with them the pressure DOFs, the theta clamps at 0.01, the
 * faces with mass (M > 1e-20) and the particles the solid projection
 * moves.  A substep run in mode 1 records them into a slot of the context,
 * one in mode 2 replays a slot (so that a perturbed run is the same smooth
 * map, as central differences need).  WC_FP64 contexts only. */
/* The gravity vector of the following substeps (the scene's
 * SimpleGravityForce, SimpleGravityForce.cpp:26-54, fixed at wc_create
 * otherwise), e.g. to take differences in |g| on one context. */
wc_status wc_set_gravity(wc_ctx* ctx, const double g[3]);
#define WC_FROZEN_OFF 0
#define WC_FROZEN_RECORD 1
#define WC_FROZEN_REPLAY 2
/* Mode and slot (0 <= slot < 2^20) for the following substeps.  Replaying a
 * slot that holds no record of the same particle count is WC_ERR_STATE. */
wc_status wc_set_frozen(wc_ctx* ctx, int32_t mode, int32_t slot);
/* Decisions of replayed substeps that the run would have taken differently
 * from the record (a liquid cell's sign, a face's mass above 1e-20, a theta
 * clamp, a particle's solid projection), counted since the last
 * wc_set_frozen with mode 0 or 1 or the last wc_rollout_forward.  A
 * replayed face that has lost all its mass is taken as massless.  Zero
 * means the replay is the run's own smooth neighbourhood. */
wc_status wc_frozen_departures(wc_ctx* ctx, int64_t* count);
/* The particle state (x, v, C) copied to (restore = 0) or from (1) a
 * device-resident slot, in the caller's particle order. */
wc_status wc_checkpoint(wc_ctx* ctx, int32_t slot, int32_t restore);
/* The adjoint of the substep wc_substep has just run in mode 1 or 2
 * (WC_ERR_STATE otherwise).  xb, vb, Cb (SoA in the caller's order, as
 * wc_get_particles_soa: 3n, 3n, 9n; NULL = zero) hold dL/d(x, v, C after
 * the substep) on entry and dL/d(x, v, C before it) on return; *gb is
 * increased by dL/d|g| for gravity (0, -|g|, 0).  The adjoint pressure
 * solve is the forward's MIC(0)-PCG on the same symmetric matrix, to tol;
 * WC_NOT_CONVERGED if it stops at maxit. */
wc_status wc_substep_adjoint(wc_ctx* ctx, double dt, double tol, int32_t maxit, double* xb,
                             double* vb, double* Cb, double* gb, double* residual_out,
                             int32_t* iters_out);
/* A rollout of `steps` substeps of dt from the current particles, the
 * forward half of a forward/backward pair: with replay = 0 it records each
 * substep's decisions (slot k for substep k) and checkpoints the particle
 * state before every `every`-th substep; with replay = 1 it replays the
 * recorded decisions (the same `steps`), e.g. from perturbed particles.
 * iters_out: PCG iterations over all substeps. */
wc_status wc_rollout_forward(wc_ctx* ctx, int32_t steps, double dt, double tol,
                             int32_t maxit, int32_t every, int32_t replay,
                             int32_t* iters_out);
/* The backward half: xb, vb, Cb hold dL/d(final x, v, C) on entry and
 * dL/d(initial x, v, C) on return (layout as wc_substep_adjoint), *gb is
 * increased by dL/d|g|.  Each segment of `every` substeps is recomputed from
 * its checkpoint (every state of the segment kept on the device), then
 * differentiated substep by substep in reverse.  On return the particles
 * hold the rollout's initial state. */
wc_status wc_rollout_backward(wc_ctx* ctx, double dt, double tol, int32_t maxit, double* xb,
                              double* vb, double* Cb, double* gb, int32_t* iters_out);

/* ---- diagnostics (downloads for parity checks) ---- */
#define WC_FIELD_ACT 0       /* uint8 per bucket: activated */
#define WC_FIELD_LPHI 1      /* double per node: liquid phi at cell centres */
#define WC_FIELD_W 2         /* double per node and axis: face solid weight */
#define WC_FIELD_MASS 3      /* double per node and axis: P2G fluid mass */
#define WC_FIELD_VEL 4       /* double per node and axis: P2G fluid velocity */
#define WC_FIELD_VOL 5       /* double per node and axis: P2G fluid volume */
#define WC_FIELD_RHS 6       /* double per node and axis: M u + M g dt */
#define WC_FIELD_USTAR 7     /* double per node and axis: predicted velocity */
#define WC_FIELD_UNEW 8      /* double per node and axis: projected velocity */
#define WC_FIELD_DOF 9       /* int32 per node: pressure DOF or -1 */
#define WC_FIELD_PRESSURE 10 /* double per node */
#define WC_FIELD_VALID 11    /* uint8 per node and axis */
/* with a solid (wc_set_elasto_particles), double per node and axis: */
#define WC_FIELD_PSI 12      /* solid volume fraction of the face */
#define WC_FIELD_SAT 13      /* saturation of the face */
#define WC_FIELD_SMASS 14    /* solid mass (elasto mass + carried liquid) */
#define WC_FIELD_SVOL 15     /* solid volume (+ carried liquid volume) */
#define WC_FIELD_SF 16       /* shape factor */
#define WC_FIELD_ORI_X 17    /* orientation, normalised: x, y, z components */
#define WC_FIELD_ORI_Y 18
#define WC_FIELD_ORI_Z 19
#define WC_FIELD_Q 20        /* drag ratio hdV / (M_s + hdV) */
#define WC_FIELD_HDV 21      /* dt * drag coefficient * (V_s + V_f) */
/* and per cell (axis ignored): */
#define WC_FIELD_PSI_P 22    /* cell-centred volume fraction */
#define WC_FIELD_SAT_P 23    /* cell-centred saturation */
#define WC_FIELD_PORE_P 24   /* pore pressure */
wc_status wc_grid_info(const wc_ctx* ctx, int32_t nx[3], int64_t* nnode,
                       int64_t* nbucket);
wc_status wc_download_field(wc_ctx* ctx, int32_t field, int32_t axis,
                            void* host);
/* The assembled pressure system in the reference's DOF order as a compact
 * CSR with sorted rows (robertbridson::FixedSparseMatrix, SparseMatrix.h:222)
 * plus rhs and solution.  Call with NULL arrays to query ndof and nnz. */
wc_status wc_download_system(wc_ctx* ctx, int64_t* ndof, int64_t* nnz,
                             uint32_t* rowstart, uint32_t* col, double* val,
                             double* rhs, double* solution);
/* The context's last pressure system (after wc_project), with DOF-ordered
 * host vectors of length ndof, on the kernels wc_project used (pipelined or
 * level-scheduled): apply_preconditioner (PCGSolver.h:326-329) with the
 * MIC(0) factor wc_project computed, and multiply(FixedSparseMatrix, x, y)
 * (PCGSolver/SparseMatrix.h:298-322).  For parity checks. */
wc_status wc_apply_preconditioner(wc_ctx* ctx, const double* r, double* z);
wc_status wc_system_multiply(wc_ctx* ctx, const double* x, double* y);
/* Wall time of the last wc_step's phases on the context's stream (HIP
 * events), in ms: rebucketize, P2G, grid update, assemble, solve, apply,
 * G2P -- the reference's timing_buffer phases (App/ParticleSimulation.cpp:19). */
wc_status wc_phase_times(const wc_ctx* ctx, double ms[7]);

/* Runtime options of a context (no reference counterpart):
 *   "pipelined_sweeps" (default 1): apply MIC(0) with the persistent tiled
 *   wavefront kernels; 0 selects one launch per i+j+k level.  Both give
 *   bitwise the same result.
 *   "sweep_subtiles" (default 1): the persistent kernels work on 8x8
 *   quarters of the 16x16 tiles, one wave each; 0 selects one workgroup per
 *   tile.  Both give bitwise the same apply; the PCG's z . r is summed per
 *   quarter or per tile, so whole solves agree to rounding.
 *   "spmv_window" (default 1): the box SpMV walks each tile in runs of
 *   consecutive skewed steps with its neighbours in LDS; 0 selects one row
 *   per thread.  Both give bitwise the same product; the PCG's s . A s is
 *   summed per run or per block, so whole solves agree to rounding.
 *   "spmv_run" (default 12, 1..4096): the steps of one run.
 *   "debug_abort_sweeps" (default 0; a test hook): the pipelined sweeps'
 *   abort flag is set before each PCG loop, as a sweep whose bounded wait
 *   ran out sets it; the solve then fails with WC_ERR_HIP and a message
 *   naming the sweep wait, and the next solve without it is unaffected.
 *   "kernel_timing" (default 0): bracket the solver's sweep and SpMV launches
 *   with HIP events on the context's stream; n > 1 brackets those of every
 *   n-th PCG iteration only (each pair of events costs a little time on the
 *   stream).  Setting it clears the counters read by wc_kernel_stats.
 *   "mixed_digits" (default 6, 1..15): WC_MIXED contexts; each refinement
 *   pass iterates until its FP32 residual has dropped by 10^-digits (or
 *   meets the solve's tolerance).
 *   "mixed_refinements" (default 2): WC_MIXED contexts; refinement passes
 *   allowed after the first (the solve reports not converged if the FP64
 *   residual still misses the tolerance after them).
 *   "solver" (default 0): 0 solves the pressure system with MIC(0)-PCG
 *   (PCGSolver.h:260-324); 1 with AMG-preconditioned PCG in FP64 whatever
 *   the context's precision (AMGPCGSolveSparse, AlgebraicMultigrid.h:263-372,
 *   the reference's liquid solver at Pressure.cpp:797), on the DOFs' cells
 *   bucket handle * 4 + node in the bucket on the context's bucket grid
 *   (Pressure.cpp:316-318, 433-436).  With 1, wc_apply_preconditioner has
 *   no factor and fails.
 *   "deterministic" (default 1 for WC_MIXED contexts, 0 for WC_FP64): 1
 *   sums P2G's contributions in 64-bit fixed
 *   point (units of 2^-56 for weights and 2^-48 for weight x velocity, each
 *   contribution rounded to nearest; speeds below 1000 m/s), which does not
 *   depend on the order the particles reach a node in, so the substep
 *   repeats bit for bit; 0 sums them with FP64 atomics, whose order (and
 *   last bits) vary from run to run.  Every other reduction of the liquid
 *   substep runs in a fixed order either way.  The mixture's solid and cell
 *   sums (wc_set_elasto_particles) stay FP64 atomic sums. */
wc_status wc_set_option(wc_ctx* ctx, const char* name, int64_t value);

/* Refinement passes of the context's last mixed-precision solve (0 for a
 * WC_FP64 context).  No reference counterpart. */
wc_status wc_refinement_passes(const wc_ctx* ctx, int32_t* passes);

/* Accumulated device time of one kernel class since "kernel_timing" was
 * last set: total ms, number of launches, and units processed (the sum over
 * launches of the system size in DOFs for the solver kernels, of the
 * particle count for the transfers).  The transfers are timed at every
 * launch whatever the "kernel_timing" stride.  No reference counterpart. */
#define WC_KERNEL_SWEEP_FWD 0 /* MIC(0) forward sweep (solve_lower) */
#define WC_KERNEL_SWEEP_BWD 1 /* MIC(0) backward sweep (solve_lower_transpose) */
#define WC_KERNEL_SPMV 2      /* multiply(FixedSparseMatrix, s, z) */
#define WC_KERNEL_P2G 3       /* the P2G kernel of wc_map_particle_nodes */
#define WC_KERNEL_G2P 4       /* the G2P + advection kernel of wc_map_node_particles */
wc_status wc_kernel_stats(const wc_ctx* ctx, int32_t kernel, double* ms,
                          int64_t* launches, int64_t* units);

/* ---- robertbridson::PCGSolver<double> on a host CSR matrix ---- */

/* Literal drop-in for PCGSolver<double>::solve(matrix, rhs, result,
 * residual_out, iterations_out) (PCGSolver/PCGSolver.h:260-308) after
 * set_solver_parameters(tol, maxit) (:249).  The CSR is
 * FixedSparseMatrix::construct_from_matrix's layout (SparseMatrix.h:251):
 * rows with sorted column indices; the matrix must be symmetric.  Like
 * form_preconditioner (:322-324) the factor always uses tau = 0.97 and
 * min-diagonal ratio 0.25. */
wc_status wc_pcg_solve_csr(int64_t n, const uint32_t* rowstart,
                           const uint32_t* col, const double* val,
                           const double* rhs, double* result, double tol,
                           int32_t maxit, double* residual_out,
                           int32_t* iters_out);
/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): wc_pcg_solve_csr, with the same
 * arguments, results and kernels, taking all its device memory -- the
 * uploaded matrix and vectors, the MIC(0) factor, the level schedule, the
 * PCG vectors and reduction partials -- from one fixed array of
 * WC_NOALLOC_SCRATCH_BYTES per GPU that the library holds from the first
 * call on that GPU until the process ends.  No device or pinned host memory
 * is allocated (the first call on a GPU also creates the stream it uses).
 * The array holds n + 1 + 2 nnz 32-bit words, 2 nnz + 7 n doubles and 4 n
 * 32-bit words, each part rounded up to 256 bytes, plus 9 KB: about 520,000
 * rows of a 7-point system.  Larger systems are refused with WC_ERR_ARG.
 * Every call on one GPU uses that GPU's array and stream: calls that
 * overlap on one GPU (from several threads) race and must be serialised by
 * the caller. */
#define WC_NOALLOC_SCRATCH_BYTES (128u << 20)
wc_status wc_pcg_solve_csr_noalloc(int64_t n, const uint32_t* rowstart,
                                   const uint32_t* col, const double* val,
                                   const double* rhs, double* result, double tol,
                                   int32_t maxit, double* residual_out,
                                   int32_t* iters_out);
/* factor_modified_incomplete_cholesky0 (PCGSolver.h:97-205).  lv is
 * indexed like val: lv[e] = L(col[e], row) for entries above the diagonal. */
wc_status wc_mic0_factor_csr(int64_t n, const uint32_t* rowstart,
                             const uint32_t* col, const double* val,
                             double tau, double mdr, double* invdiag,
                             double* lv);
/* apply_preconditioner: solve_lower then solve_lower_transpose_in_place
 * (PCGSolver.h:211-239, :326-329). */
wc_status wc_mic0_apply_csr(int64_t n, const uint32_t* rowstart,
                            const uint32_t* col, const double* invdiag,
                            const double* lv, const double* r, double* z);
/* AMGPCGSolveSparse(matrix, rhs, result, Dof_ijk, tolerance_factor,
 * max_iterations, residual_out, iterations_out, ni, nj, nk)
 * (AlgebraicMultigrid.h:247-359), the solve Pressure.cpp:797 calls: PCG
 * preconditioned by one V-cycle of the aggregation AMG of
 * GeometricLevelGen.h:186-249 (2x2x2 blocks of the DOFs' cells, Galerkin
 * coarse operators 0.5 R A P, red-black Gauss-Seidel 4 sweeps per level and
 * 200 on the coarsest, coarsening while a level holds more than 4096 DOFs).
 * dof_ijk: 3 int32 per DOF, its cell (i, j, k) with 0 <= i < ni etc., no
 * two DOFs in one cell (WC_ERR_ARG otherwise: the reference accepts them,
 * these kernels keep one child per position of a 2x2x2 block).  The
 * rows must be sorted and every coarse row may hold at most 7 entries
 * (WC_ERR_HIP otherwise; a 7-point system always meets this).  Rows of one
 * colour (parity of i + j + k) must not couple: the reference smooths each
 * colour in a parallel loop (AlgebraicMultigrid.h:123, 150), whose result is
 * defined only then, and so do these kernels.  x = 0 start, |r|_inf <= tol |b|_inf;
 * WC_NOT_CONVERGED after max_iterations or when rho is 0 or NaN. */
wc_status wc_amgpcg_solve_csr(int64_t n, const uint32_t* rowstart, const uint32_t* col,
                              const double* val, const int32_t* dof_ijk, int32_t ni,
                              int32_t nj, int32_t nk, const double* rhs, double* result,
                              double tol, int32_t maxit, double* residual_out,
                              int32_t* iters_out);
/* amgPrecondCompressed (AlgebraicMultigrid.h:234-245): z = one V-cycle
 * applied to r from z = 0, on the hierarchy wc_amgpcg_solve_csr builds. */
wc_status wc_amg_apply_csr(int64_t n, const uint32_t* rowstart, const uint32_t* col,
                           const double* val, const int32_t* dof_ijk, int32_t ni, int32_t nj,
                           int32_t nk, const double* r, double* z);
/* multiply(FixedSparseMatrix, x, result) (SparseMatrix.h:298-322). */
wc_status wc_spmv_csr(int64_t n, const uint32_t* rowstart, const uint32_t* col,
                      const double* val, const double* x, double* y);

#ifdef __cplusplus
}
#endif
#endif /* WETCLOTH_AMD_H */
