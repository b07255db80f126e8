/*
 * zdm.h -- C ABI of the MI355X (gfx950) finite-volume time step for SME / VAM models.
 *
 * One shared library per model (libzdm_<model>.so), built from the model's generated
 * Model.H / Numerics.H exactly as the reference builds solver_cpu on them (Makefile:46-55).
 * Every entry point returns 0 on success and a negative code on failure; the message is
 * available from zdm_last_error().  All device work is ordered on the solver's own HIP
 * stream; a handle must not be used from two threads at once.
 *
 * State arrays at this boundary use the reference's layout: array-of-structs per cell,
 * x[c*n_q + i] (the PETSc Vec layout of dmQ, VirtualSolver.hpp:447-457), cells in the
 * mesh's solver order.  Device storage is structure-of-arrays and never leaves the handle.
 */
#ifndef ZDM_H
#define ZDM_H

#include <stddef.h>

#include "zdm_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct zdm_solver zdm_solver;

/* Solver options: the subset of SolverSettings (Settings.hpp:13-40) the step consumes. */
typedef struct zdm_options {
    int order;          /* reconstruction_order: 1 = PCM + forward Euler, >= 2 = MUSCL + SSP-RK */
    int ssp_stages;     /* s of the s-stage SSP-RK2 (TSSSP "rks2", SolverStrategies.hpp:33-35) */
    int use_limiters;   /* limiter != "none" -> Vacuum + TVD (MUSCLSolver.hpp:40-43) */
    int gradient;       /* 0 = Green-Gauss (reference, ModularSolver.hpp:62), 1 = least squares */
    int source;         /* the model's source: split implicit step (splitting, MUSCLSolver.hpp:179)
                           or the implicit part of the IMEX step */
    double cfl;         /* solver.cfl */
    double t_end;       /* solver.t_end */
    double min_dt;      /* solver.min_dt (first step only, MUSCLSolver.hpp:103) */
    int block_size;     /* workgroup size (0 = 256) */
    int well_balanced;  /* extension: add Model::hydrostatic_cell_source per face side (second-order
                           hydrostatic reconstruction); 0 = the reference scheme */
    int positivity;     /* solver.positivity: 0 "none", 1 "mood" (order 2: eta-WB reconstruction,
                           ModularSolver.hpp:46-58; any order: the local MOOD override of cells with
                           h < -1e-10, MUSCLSolver.hpp:130-171), 2 "zhang_shu" (order 2: eta-WB +
                           Zhang-Shu theta, CFL capped at 1/6, MUSCLSolver.hpp:31-38) */
    double wet_dry_eps; /* solver.wet_dry_eps (<= 0: the default 1e-2) */
    int time_integration; /* solver.time_integration: 0 "splitting" (SSP-RK / Euler transport, then
                           the split implicit source, SplittingStrategy), 1 "imex" (explicit
                           transport + implicit source in one additive Runge-Kutta step,
                           IMEXStrategy, SolverStrategies.hpp:62-84), 2 "implicit" (BDF2 on
                           transport + source, Newton-Krylov, FullyImplicitStrategy,
                           SolverStrategies.hpp:88-195) */
    int refresh_derivative_aux; /* solver.refresh_derivative_aux (Settings.hpp:24, default true): for a
                           model with the field-level update_aux_variables(Q, Qaux, dt, mesh), run it
                           (RefreshDerivativeAux, TransportStep.hpp:161-202) before every UpdateState;
                           0 skips it.  Splitting time integration without MOOD only */
} zdm_options;

/* Newton-Krylov settings of time_integration "implicit" (the SNES / KSP of the TS). */
typedef struct zdm_newton_options {
    double snes_atol;   /* 1e-4 (MUSCLSolver.hpp:108-110) */
    double snes_rtol;   /* 1e-3 */
    int snes_max_it;    /* 20 */
    double ksp_rtol;    /* PETSc default 1e-5 */
    int ksp_restart;    /* GMRES restart, PETSc default 30 */
    int ksp_max_it;     /* PETSc default 10000 */
} zdm_newton_options;

/* Model plug-in description (Model.H:31-44). */
int zdm_model_info(int* n_q, int* n_aux, int* n_params, int* dim);
const char* zdm_model_name(void);
int zdm_model_default_params(double* params);              /* Model::default_parameters() */
/* Model::get_boundary_tags(): writes up to cap names of 64 chars, returns the count. */
int zdm_model_boundary_tags(char* names64, int cap);
/* Model::parameter_names() (Model.H:43), same layout. */
int zdm_model_parameter_names(char* names64, int cap);
/* Model::initial_condition(X, p) at one point (Model.H:264-274; optional in a generated model,
 * VirtualSolver.hpp:33-38): n_q values into Q, or an error when the model emits none. */
int zdm_model_initial_condition(const double* X, const double* params, double* Q);

/* Create a solver on `mesh`.  face_bc[f] = model BC index of boundary face f, or -1 when its
 * tag maps to no model BC (such faces get no flux, TransportStep.hpp:498).  Replaces
 * VirtualSolver::SetupArchitecture (VirtualSolver.hpp:386-525) + InitializeComponents
 * (ModularSolver.hpp:39-76). */
int zdm_solver_create(const zdm_mesh* mesh, const int* face_bc, const double* params,
                      const zdm_options* opts, zdm_solver** out);
int zdm_solver_destroy(zdm_solver* s);

/* Ensemble of independent runs in one set of launches (the batching form).  `mesh` is the
 * disjoint union of the members (zdm_mesh_concat); member i owns cells [cell_offsets[i],
 * cell_offsets[i+1]) and whole workgroups, and keeps its own time, dt (local CFL over its
 * own cells), step count and t_end test, exactly as if it ran alone through
 * zdm_solver_create.  Every other entry point works on the union state; zdm_get_time reports
 * member 0 and zdm_batch_time any member.  Replaces running the reference's solver once
 * per member (one DMPlex + TS per run, VirtualSolver.hpp:386-525). */
int zdm_batch_create(const zdm_mesh* mesh, const int* face_bc, const double* params,
                     const zdm_options* opts, int nmembers, const int* cell_offsets,
                     zdm_solver** out);
int zdm_batch_time(zdm_solver* s, int member, double* time, double* dt, int* steps);

/* Host <-> device state transfer (AoS, n_q / n_aux per cell).  A may be NULL (aux = 0). */
int zdm_state_upload(zdm_solver* s, const double* X, const double* A);
int zdm_state_download(zdm_solver* s, double* X, double* A);

/* EnforcePhysicalConstraints(X) + the first dt = max(ComputeTimeStep(), min_dt)
 * (MUSCLSolver.hpp:54, 103).  Resets time and step count to 0. */
int zdm_initialize(zdm_solver* s);

/* Enqueue n full time steps of the MUSCLSolver loop (MUSCLSolver.hpp:140-188): clamp dt to
 * t_end, TSStep (SSP-RK / Euler), constraints, split source, constraints, CFL for the next
 * dt.  Steps past t_end are no-ops.  Asynchronous: no host synchronisation. */
int zdm_step(zdm_solver* s, int n);

/* As zdm_step, with the launches of one step captured once into a HIP graph and replayed
 * n times (same kernels and arguments, so the same results as zdm_step). */
int zdm_step_graph(zdm_solver* s, int n);

/* Run until t_end (or max_steps); synchronises every `check_every` steps. */
int zdm_run(zdm_solver* s, int max_steps, int check_every, int* steps_done);

/* Time, current/next dt and step count (synchronises). */
int zdm_get_time(zdm_solver* s, double* time, double* dt, int* steps);
int zdm_synchronize(zdm_solver* s);

/* One transport right-hand side F = L(X) at time t (TransportStep::FormRHS,
 * TransportStep.hpp:325-357) on host arrays; A is the scratch aux input (may be NULL). */
int zdm_transport_rhs(zdm_solver* s, double t, const double* X, const double* A, double* F);

/* The same right-hand side on AoS float64 device buffers (X: ncells*n_q, A: ncells*n_aux or
 * NULL, F: ncells*n_q), ordered on `stream` (a hipStream_t; the solver's own stream when it is
 * zdm_solver_stream's): the call's work starts after what `stream` holds and later work on
 * `stream` sees F.  It does not synchronise the host.  Bitwise zdm_transport_rhs's result.
 * Replaces TransportStep::FormRHS(time, X, F) (TransportStep.hpp:325) where X and F live on
 * the device; not with the derivative-aux refresh (that host loop synchronises). */
int zdm_transport_rhs_device(zdm_solver* s, double t, const double* dX, const double* dA, double* dF,
                             void* stream);

/* SourceStep::Solve(dt, X, A) on the resident state (SourceStep.hpp:72-122). */
int zdm_source_step(zdm_solver* s, double dt);

/* MUSCLSolver::EnforcePhysicalConstraints on the resident state (MUSCLSolver.hpp:234-254). */
int zdm_enforce_constraints(zdm_solver* s);

/* VirtualSolver::ComputeTimeStep on the resident state (VirtualSolver.hpp:285-357). */
int zdm_cfl_dt(zdm_solver* s, double* dt);

/* Device-pointer variants (AoS device buffers, ordered on the solver stream). */
int zdm_state_upload_device(zdm_solver* s, const double* dX, const double* dA);
int zdm_state_download_device(zdm_solver* s, double* dX, double* dA);

/* Per-kernel HIP-event timing of zdm_step launches while enabled (ms, averaged per launch;
 * events are resolved without synchronising inside zdm_step).  on = 0 off, 1 every step, k > 1
 * every k-th step of zdm_step (a sample that leaves the other steps without event markers).
 * Kernel ids: 0 = prep (gradient/bounds/reconstruction), 1 = flux + stage update, 2 = post-step,
 * 3 = the final stage's flux + stage update fused with the post-step (zdm_fuse_post). */
int zdm_enable_timing(zdm_solver* s, int on);
int zdm_kernel_time(zdm_solver* s, int kernel, double* avg_ms, int* launches);
/* on = 1 (the default): a step of order >= 2 runs its final SSP stage and the post-step
 * (constraints, split source, constraints, CFL) as one kernel, the new state staying in registers
 * between them, where the step has no MOOD override, no derivative-aux refresh and no
 * well-balanced term; 0: separate kernels.  Results are bitwise the same.  No counterpart in the
 * reference (its MUSCLSolver runs TSStep, then the source and the constraints, as separate
 * passes: MUSCLSolver.hpp:140-188). */
int zdm_fuse_post(zdm_solver* s, int on);
/* Cache policy of the transport kernels (one cell per thread): mode 1 loads and stores their
 * read-once and write-once streams non-temporally (in the flux launches the stage state, Q^n, the
 * output state and aux, the cell's own slot topology and volume; in the Green-Gauss
 * reconstruction the cell's own slot topology, slot geometry and volume), so that they do not
 * displace data that neighbours or the next launch read again; 0 uses ordinary accesses; -1 (the
 * default) picks 1 where one flux pass (48 n_q + 48 B per cell, 40 B per face) exceeds 256 MiB,
 * the last-level cache.  *active (may be null) receives the form now in use.  Results are bitwise
 * the same either way.  No counterpart in the reference (a cache policy of this build's kernels). */
int zdm_flux_nontemporal(zdm_solver* s, int mode, int* active);

/* Measured HBM bandwidth of a plain device copy of `bytes` (read + write counted), averaged
 * over `iters` launches: the achievable-peak reference and the traffic-counter calibration. */
int zdm_stream_copy(size_t bytes, int iters, double* gbs, double* avg_ms);

/* ---- Chorin pressure split for VAM models (ChorinVAMSolver, ChorinVAM.hpp:86-129).
 * Available in libraries built for a model whose Model.H defines ZOOMY_MODEL_IS_VAM and
 * whose directory provides ChorinOps.H; elsewhere these return an error, as the reference's
 * stub ChorinVAMSolver::Run does (ChorinVAM.hpp:26-37).  The preconditioner is always
 * point-block Jacobi (-pc_type pbjacobi) on the device-assembled 2x2-block Pmat. */
typedef struct zdm_krylov_options {
    int restart;     /* KSPGMRESSetRestart (ChorinVAM.hpp:403; default 40) */
    int max_it;      /* KSPSetTolerances max_it (ChorinVAM.hpp:449; default 400) */
    double rtol;     /* default 1e-9 */
    double atol;     /* default 1e-12 */
    double dtol;     /* PETSc default 1e5 */
    int op_mode;     /* 0 = matrix-free residual(P) - R0 (the MatShell Aop, ChorinVAM.hpp:577-584);
                        1 = the assembled Pmat as the operator */
} zdm_krylov_options;

/* Krylov settings of the pressure solve (SetupChorin + KSPSetFromOptions,
 * ChorinVAM.hpp:386-459); NULL restores the defaults.  The stencils and the block pattern are
 * built by zdm_solver_create. */
int zdm_chorin_setup(zdm_solver* s, const zdm_krylov_options* opts);

/* n Chorin steps of ChorinVAMSolver::Run's loop (ChorinVAM.hpp:102-125) on the resident
 * state, after zdm_initialize: dt = max(CFL, min_dt) clamped to t_end, predictor
 * (FormRHS(0) + forward Euler + source + constraints), frozen derivative aux, Pmat + PBJ,
 * GMRES on A P = -R0 from P = 0, corrector, constraints.  Reports the steps taken, the
 * total GMRES iterations and the last KSPConvergedReason (2 rtol, 3 atol, -3 max_it,
 * -4 dtol, -9 NaN).  Synchronises once per GMRES iteration. */
int zdm_chorin_step(zdm_solver* s, int n, int* steps_done, int* its_total, int* last_reason);

/* The last pressure solve: GMRES iterations, KSPConvergedReason and the preconditioned
 * residual norm at the end and at the start (rnorm / rnorm0 is the relative residual the
 * rtol test compares; a solve stopped by max_it reports how far it got). */
int zdm_chorin_stats(zdm_solver* s, int* its, int* reason, double* rnorm, double* rnorm0);

/* The last solved pressure, NP = 2 per cell (the dmP Vec layout). */
int zdm_chorin_pressure(zdm_solver* s, double* P);

/* Parts of one step on the resident state, exposed for parity tests and callers that drive
 * the split themselves:
 *   zdm_vam_frozen_aux: ComputeFrozenAux (FillCA(false), ChorinVAM.hpp:521-546); CA (may be
 *     NULL) receives all NCA slots per cell (pressure slots as last computed);
 *   zdm_vam_residual: PressureResidual (ChorinVAM.hpp:551-576) of P (NP per cell) with the
 *     frozen aux of the last zdm_vam_frozen_aux;
 *   zdm_vam_pmat: AssemblePmat (ChorinVAM.hpp:242-303) as 2x2-block CSR (row pointers
 *     ncells+1, block columns, 4 values per block row-major); returns the block count, and
 *     only that when cap is smaller. */
int zdm_vam_frozen_aux(zdm_solver* s, double* CA);
int zdm_vam_residual(zdm_solver* s, double dt, const double* P, double* R);
int zdm_vam_pmat(zdm_solver* s, double dt, int cap, int* rowptr, int* col, double* val);

/* GMRES(restart) + point-block Jacobi on a 2x2-block CSR system A x = b (host arrays; x holds
 * the initial guess): KSPGMRES with classical Gram-Schmidt, left preconditioning and the
 * preconditioned-residual test of KSPConvergedDefault, as the pressure solve configures it
 * (ChorinVAM.hpp:405-451).  Exported by every model library. */
int zdm_bsr2_gmres(int nb, const int* rowptr, const int* col, const double* val, const double* b,
                   double* x, const zdm_krylov_options* opts, int* its, int* reason);

/* zdm_bsr2_gmres for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the same kernels and the same results,
 * with rowptr, col, val, b and x (initial guess in, solution out) as DEVICE arrays and the work
 * ordered on `stream` (a hipStream_t; NULL = the null stream).  It calls no hipMalloc / hipFree:
 * its workspace (the point-block Jacobi inverse, the Krylov basis, the reduction partials and
 * the Hessenberg state) lives in a 64 MiB device array the library holds on each GPU until the
 * process ends, which bounds the size: about 560 * nb + 270 000 bytes for restart 30; larger
 * systems are rejected with an error.  Every call on one GPU shares that array (the contents
 * of the last call remain; each call clears what it needs), so calls that overlap on one GPU
 * race.  The structure is checked on the device (row pointers from 0, non-decreasing; columns
 * in [0, nb)).  Synchronises `stream` once per restart cycle.  No counterpart in the
 * reference: it exists for the callers named above, which cannot allocate. */
int zdm_bsr2_gmres_noalloc(int nb, const int* rowptr, const int* col, const double* val, const double* b,
                           double* x, const zdm_krylov_options* opts, void* stream, int* its, int* reason);

/* ---- Discrete adjoint of the time step (the differentiation form).  A calibration problem
 * of nsteps MUSCLSolver steps from (X0, A0) with the loss
 *     L = sum_{k>=1} sum_g (h(gauges[g]) after step k*obs_every - obs[(k-1)*ngauge + g])^2,
 * and its gradient with respect to the initial state X0 (bathymetry = slot 0) and the
 * initial aux A0 (pass-through aux such as a per-cell friction coefficient).  The gradient
 * is the exact transpose of the discrete step as computed, with the branches of clamps,
 * bounds, limiter and wet/dry switches and the time steps frozen at the forward run.
 * Limiter: exact_limiter = 0 (the default for calibration) holds the limiter value alpha of
 * every face fixed (the frozen-limiter linearisation: Barth-Jespersen's alpha is one minimum
 * over all components, so its exact derivative couples them through 1/delta and grows without
 * bound in quiet regions); exact_limiter = 1 differentiates alpha through its active branch
 * (the exact derivative of the discrete step, for finite-difference checks).
 * Memory (checkpoint_every):
 *   -2  a tape of every step's state, stage states, post-stage states and face states: the
 *       reverse sweep runs only transposed kernels (most memory);
 *   -1  the same without the face states (the reverse sweep recomputes them per stage);
 *    1  the state before every step (each step's stages are re-run in the reverse sweep);
 *    k  a checkpoint every k steps (each segment is re-run once more);
 *  -2-c (c >= 1) c state slots, checkpoints placed by the optimal binomial schedule
 *       (Griewank-Walther "revolve": the fewest forward re-steps for c slots);
 *       both with a stretch tape when it fits (zdm_adjoint_stretch_tape);
 *    0  the first of -2, -1 that fits in half of the free device memory, else 1 if the
 *       states fit in a quarter of it, else the binomial schedule on the slots that do.
 * Every layout gives bitwise the same gradient.  Order 1, or order 2 with
 * the Green-Gauss gradient; not for ensembles or the well-balanced extension.  No
 * counterpart in the reference, which has no differentiation path. */
int zdm_adjoint_setup(zdm_solver* s, int nsteps, int checkpoint_every, int obs_every, int ngauge,
                      const int* gauges, const double* obs, int exact_limiter);
/* The same calibration problem over ChorinVAMSolver steps (VAM models; ChorinVAM.hpp:102-125:
 * predictor, frozen derivative aux, pressure solve, corrector, constraints) with the Krylov
 * settings of zdm_chorin_setup.  The gradient transposes the split with the pressure taken as
 * the exact solution of its linear system: mu = A^-T (adjoint of P) by GMRES + point-block
 * Jacobi on the transposed Pmat, then the residual's derivative with respect to the
 * predicted state.  Layouts as above except the tapes (-1, -2): every layout re-runs each
 * step once in the reverse sweep.  No counterpart in the reference. */
int zdm_adjoint_setup_chorin(zdm_solver* s, int nsteps, int checkpoint_every, int obs_every, int ngauge,
                             const int* gauges, const double* obs, int exact_limiter);
/* GMRES iterations of the transposed pressure solves since zdm_adjoint_setup[_chorin]. */
int zdm_adjoint_krylov_its(zdm_solver* s, long long* its);
/* time_integration "implicit": Newton / Krylov settings (NULL restores the defaults above)
 * and the counters since creation (Newton and GMRES iterations, the last step's Newton
 * iterations and SNESConvergedReason: 2 |G| < atol, 3 |G| < rtol |G0|, -5 max_it, -4 NaN). */
int zdm_implicit_setup(zdm_solver* s, const zdm_newton_options* opts);
int zdm_implicit_stats(zdm_solver* s, long long* newton_its, long long* ksp_its, int* last_newton,
                       int* last_reason);
/* Launch shape of the transport kernels (k_prep, k_flux, k_post): workgroup size 64, 128 or
 * 256 and 1..8 cells per thread; ensembles keep whole workgroups per member.  Results are
 * bitwise the same for every shape.  Invalidates a step captured by zdm_step_graph.  No
 * counterpart in the reference. */
int zdm_set_launch(zdm_solver* s, int block_size, int cells_per_thread);
int zdm_get_launch(zdm_solver* s, int* block_size, int* cells_per_thread);
/* Autotuner: times `steps` steps of each candidate shape, {256, 128, 64} threads x {1, 2, 4}
 * cells per thread (in that order, cells-per-thread outer), on the resident state, restores
 * the state, aux and time, and keeps the fastest.  ms (may be NULL) receives the 9 ms/step.
 * Call it before running (steps past t_end do nothing). */
int zdm_autotune(zdm_solver* s, int steps, int* block_size, int* cells_per_thread, double* ms);
/* MOOD overrides so far (cells of h < -1e-10 replaced by their first-order update), summed
 * over ensemble members; reset by zdm_initialize.  The reference prints this count per step. */
int zdm_troubled_cells(zdm_solver* s, long long* count);
/* The layout zdm_adjoint_setup chose: checkpoint_every in its own encoding (1 with a tape,
 * -2-c for the binomial schedule on c slots) and the level of a full tape (0 none, 1 stage
 * states, 2 stage and face states; a stretch tape is reported by zdm_adjoint_stretch_tape). */
int zdm_adjoint_layout(zdm_solver* s, int* checkpoint_every, int* tape);
/* Stretch tape of the checkpointed layouts (checkpoint_every >= 2 or <= -3): a stretch of at
 * most `steps` steps (one segment for checkpoint_every >= 2) is re-run once in the reverse
 * sweep with its stage states, post-stage states and (level 2) face states recorded in place,
 * and its adjoints then run only transposed kernels, instead of each reversed step re-running
 * its own step; the forward run records the last stretch itself.  With two buffers the
 * re-runs run on a second stream of the solver: the next stretch is recorded while the
 * current one is reversed (level 2 only).  zdm_adjoint_setup picks ceil(nsteps / c) steps on
 * c binomial slots (one segment for checkpoint_every >= 2): two level-2 buffers if they fit in
 * half of the free device memory, else one at level 2, else level 1, else none.  steps >= 0
 * replaces it (0: none; buffers 0 automatic, 1 or 2), re-planning the schedule; steps < 0 only
 * queries.  steps_out / level_out / buffers_out (may be NULL) receive its length, level and
 * buffers (0: none).  The gradient is bitwise the same with or without it.  Not for the
 * Chorin split.  No counterpart in the reference. */
int zdm_adjoint_stretch_tape(zdm_solver* s, int steps, int buffers, int* steps_out, int* level_out,
                             int* buffers_out);
/* Forward run (X0, A0 in the zdm_state_upload layout, A0 may be NULL).  dts = NULL takes
 * dt = CFL each step and records it; otherwise the given sequence is used (a frozen-dt
 * re-run, e.g. for finite differences).  Returns the loss. */
int zdm_adjoint_forward(zdm_solver* s, const double* X0, const double* A0, const double* dts,
                        double* loss);
/* The forward map the frozen-limiter gradient (exact_limiter = 0) differentiates, for checking
 * it by differences: a forward run of the calibration problem whose face limiter values are
 * stored (record = 1: every step and stage; dts may be NULL for the CFL steps) or taken from the
 * last recording (record = 0; dts required, the recording's sequence).  The loss is returned as
 * zdm_adjoint_forward's; the run overwrites the forward a gradient would reverse (call
 * zdm_adjoint_forward again before zdm_adjoint_gradient).  Order 2 only.  No counterpart in the
 * reference (which has no adjoint). */
int zdm_adjoint_forward_frozen_alpha(zdm_solver* s, const double* X0, const double* A0, const double* dts,
                                     int record, double* loss);
/* Gradient of the loss of the last forward run: gX0 (ncells*n_q), gA0 (ncells*n_aux);
 * either may be NULL. */
int zdm_adjoint_gradient(zdm_solver* s, double* gX0, double* gA0);
/* Device-buffer variants (AoS float64 device buffers in the zdm_state_upload layout, read and
 * written in order on the solver stream; dts and loss stay host memory).  The forward returns
 * after the loss is known; the gradient returns with its copies enqueued (order later work
 * on the solver stream, or synchronise it). */
int zdm_adjoint_forward_device(zdm_solver* s, const double* dX0, const double* dA0, const double* dts,
                               double* loss);
int zdm_adjoint_gradient_device(zdm_solver* s, double* dgX0, double* dgA0);
/* The solver's stream (a non-blocking hipStream_t of default priority) for ordering
 * a caller's work against the solver's, e.g. through events. */
int zdm_solver_stream(zdm_solver* s, void** stream);
/* The dt sequence (nsteps) and the residuals h - obs (nobs*ngauge) of the last forward run. */
int zdm_adjoint_dts(zdm_solver* s, double* dts);
int zdm_adjoint_residuals(zdm_solver* s, double* r);

const char* zdm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
