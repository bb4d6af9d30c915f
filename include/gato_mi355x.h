/* gato_mi355x — batched SQP trajectory optimisation on one AMD Instinct MI355X (gfx950).
 *
 * C-ABI of the HIP library libgato_mi355x.so. Plain pointers and sizes only; device pointers
 * are HIP device allocations owned by the caller (e.g. torch tensors on cuda:0).
 * Every entry point returns 0 on success or a non-zero status (GATO_E_* or a hipError_t);
 * gato_last_error() describes the last failure of the calling thread. Nothing here exits
 * the process or throws across the ABI.
 *
 * Reference interface replaced (Mo0dy/GATO, python/bindings.cu): the pybind11 class
 * BSQP_{B}_float of module bsqpN{N}_{plant} (bindings.cu:212-266) and the solver it wraps,
 * BSQP<T, BatchSize> (gato/bsqp/bsqp.cuh:20-353). Knot count N and batch size B are run-time
 * arguments here instead of template parameters.
 */
#ifndef GATO_MI355X_H
#define GATO_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GATO_OK 0
#define GATO_E_ARG 1001     /* invalid argument (null handle, bad size, unknown plant) */
#define GATO_E_ALLOC 1002   /* device allocation failed */
#define GATO_E_LAUNCH 1003  /* kernel launch failed */

#define GATO_PLANT_IIWA14 0
#define GATO_PLANT_INDY7 1
/* indy7 with its links' centres of mass from the URDF, as the reference's MPC examples simulate it with pinocchio
 * (common.py:49-91); the solver's indy7 model (GRiD's) places them at the joint origins. Plant simulation only. */
#define GATO_PLANT_INDY7_URDF 2

/* execution paths */
#define GATO_MODE_PLAIN 0 /* one kernel per phase, host loop with a device->host sync per iteration */
#define GATO_MODE_FUSED 1 /* one persistent kernel: every SQP iteration of every problem on device */
#define GATO_MODE_GRAPH 2 /* the plain kernels for every iteration, gated by a device-side stop flag instead of
                             the host read, captured once into a hipGraph and replayed */

/* convergence rules */
#define GATO_EXIT_PCG0 0 /* reference: converged <=> PCG took 0 iterations (bsqp.cuh:149-163) */
#define GATO_EXIT_KKT 1  /* converged <=> max(|stationarity|_inf, |defect|_inf) <= kkt_tol */

/* Solver settings: the 15 constructor scalars of BSQP (bsqp.cuh:43-45, bindings.cu:30-44) plus
 * the MI355X build's extensions (marked). */
typedef struct gato_params {
    float dt;
    uint32_t max_sqp_iters;
    float kkt_tol;
    uint32_t max_pcg_iters;
    float pcg_tol;
    float solve_ratio;
    float mu;
    float q_cost, qd_cost, u_cost, N_cost, q_lim_cost, vel_lim_cost, ctrl_lim_cost;
    float rho;
    /* extensions */
    int32_t mode;          /* GATO_MODE_* (the Python layer defaults to GATO_MODE_GRAPH) */
    int32_t exit_rule;     /* GATO_EXIT_* (default GATO_EXIT_PCG0) */
    int32_t terminal_mode; /* 0: reference terminal block (x_{N-2}, q_cost); 1: consistent (x_{N-1}, N_cost) */
    int32_t hess_mode;     /* 0: reference rank-1 end-effector Hessian; 1: Gauss-Newton J^T J */
} gato_params;

/* Per-solve statistics, written to HOST memory by gato_solve. Arrays sized by the caller:
 * per-problem arrays [B]; per-iteration arrays [max_sqp_iters * B] (row = iteration). */
typedef struct gato_stats {
    double solve_time_us;   /* host wall clock around the device work (bsqp.cuh:109, 185) */
    int32_t n_iters;        /* SQP loop iterations run */
    int32_t ls_num_iters;   /* iterations that reached the line search */
    int32_t* sqp_iters;     /* [B] */
    int32_t* kkt_converged; /* [B] */
    float* initial_merit;   /* [B] */
    float* final_merit;     /* [B] */
    int32_t* pcg_iters;     /* [max_sqp_iters * B] */
    float* ls_min_merit;    /* [max_sqp_iters * B] */
    float* ls_step_size;    /* [max_sqp_iters * B]; -1 = line-search failure */
    float* kkt_residual;    /* [max_sqp_iters * B] */
    int32_t* status;        /* [B] GATO_STATUS_* bits, or NULL (no status pass is run) */
} gato_stats;

/* Per-problem failure status of a solve (gato_stats.status). The reference declares PCGStats::converged and never
 * sets it (gato/types.cuh:28), and reports failures only through its per-iteration arrays; these bits summarise
 * them per problem, and add a scan of the returned trajectory for non-finite values, on the device. */
#define GATO_STATUS_KKT_CONVERGED 1  /* kkt_converged: the problem met the stopping rule */
#define GATO_STATUS_ITER_CAP      2  /* ran max_sqp_iters iterations without meeting it */
#define GATO_STATUS_PCG_AT_CAP    4  /* in some iteration the PCG ran max_pcg_iters iterations */
#define GATO_STATUS_LS_FAILED     8  /* in some iteration no step length lowered the merit (step -1) */
#define GATO_STATUS_NONFINITE     16 /* a NaN or Inf in the returned trajectory or the initial or final merit */

typedef struct gato_solver gato_solver;

/* BSQP(...) constructor (bsqp.cuh:43-59; bindings.cu:27-56): allocates the workspace for
 * `batch` problems of `knots` knot points on the current device. */
int gato_create(gato_solver** out, int32_t plant, int32_t knots, int32_t batch, const gato_params* params);
/* ~BSQP() (bsqp.cuh:61; bindings.cu:58-66) */
int gato_destroy(gato_solver* s);

/* BSQP::solve (bsqp.cuh:103-197; bindings.cu:68-148). Device pointers:
 *   d_xu  [B * ((nx+nu)N - nu)] in/out trajectory, d_xs [B * nx] initial states,
 *   d_ref [B * 6N] end-effector reference (xyz used, rpy ignored).
 * stream: hipStream_t (NULL = default stream). Returns after the solve completed. */
int gato_solve(gato_solver* s, float* d_xu, const float* d_xs, const float* d_ref, float dt, void* stream,
               gato_stats* stats);

/* Allocation-free solve. No counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It runs gato_solve's
 * GATO_MODE_GRAPH kernels (params->mode is ignored), enqueued on `stream` directly, for a solver fresh from
 * gato_create: lambda = 0, rho = params->rho, drho = 1, mu = params->mu, PCG tolerance = params->pcg_tol, no
 * wrenches, the reference cost. Its workspace is carved from a static device buffer of GATO_NOALLOC_SCRATCH_BYTES
 * that the library holds on every GPU it runs on until the process ends; the call makes no hipMalloc, hipFree or
 * other allocating call. A problem set whose workspace exceeds the buffer is refused with GATO_E_ARG before anything
 * runs (iiwa14 at N = 32 and 10 SQP iterations: about 190 problems). Every call on one GPU uses that GPU's one
 * buffer and clears it first: nothing carries over from the previous call, and calls that overlap on one GPU (other
 * streams, other threads) race. stats: NULL for an asynchronous call; otherwise filled as by gato_solve, and the
 * call returns after the solve completed. Same arguments and layouts as gato_create + gato_solve. */
#define GATO_NOALLOC_SCRATCH_BYTES (32u << 20)
int gato_solve_noalloc(int32_t plant, int32_t knots, int32_t batch, const gato_params* params, float* d_xu,
                       const float* d_xs, const float* d_ref, float dt, void* stream, gato_stats* stats);

/* FP64 mode: the reference's USE_DOUBLES build (gato/settings.h:7-11), classes BSQP_{B}_double of bindings.cu:244-252.
 * The same SQP iteration in double precision on the device: the one-kernel-per-phase path (GATO_MODE_PLAIN; params->mode
 * is ignored) with every phase instantiated at double, library sin, cos, division and square root. All device arrays
 * are double with the layouts of the fp32 entry points; the statistics come back as gato_stats (merits rounded to
 * float; gato_last_merits_f64 gives them in double). For high-accuracy offline solves and oracle cross-checks. */
typedef struct gato_solver64 gato_solver64;
#define GATO_SET_RHO 0     /* set_rho_penalty_batch */
#define GATO_SET_DRHO 1    /* set_drho_batch */
#define GATO_SET_MU 2      /* set_mu_batch */
#define GATO_SET_PCG_TOL 3 /* set_pcg_tol_batch */
#define GATO_RESET_DUAL 0
#define GATO_RESET_RHO 1
/* BSQP<double, B>(...) (bsqp.cuh:43-59) */
int gato_create_f64(gato_solver64** out, int32_t plant, int32_t knots, int32_t batch, const gato_params* params);
int gato_destroy_f64(gato_solver64* s);
/* BSQP<double, B>::solve (bsqp.cuh:103-197): d_xu [B * traj], d_xs [B * nx], d_ref [B * 6N], device doubles */
int gato_solve_f64(gato_solver64* s, double* d_xu, const double* d_xs, const double* d_ref, double dt, void* stream,
                   gato_stats* stats);
/* initial and final merits of the last solve in double (host arrays [B]; either may be NULL) */
int gato_last_merits_f64(gato_solver64* s, double* h_initial, double* h_final);
/* set_joint_cost, set_f_ext_batch, the batched setters (GATO_SET_*, host arrays [B]), reset_dual / reset_rho
 * (GATO_RESET_*), set_rho_adaptation, get_state and sim_forward, as the fp32 entry points, in double */
int gato_set_joint_cost_f64(gato_solver64* s, const double* d_Qd, const double* d_Rd, const double* d_xref,
                            void* stream);
int gato_set_f_ext_batch_f64(gato_solver64* s, const double* d_f_ext, void* stream);
int gato_set_batch_f64(gato_solver64* s, int32_t what, const double* h_values, int32_t set_as_reset_default);
int gato_reset_f64(gato_solver64* s, int32_t what);
int gato_set_rho_adaptation_f64(gato_solver64* s, int32_t enabled);
int gato_get_state_f64(gato_solver64* s, double* d_lambda, double* h_rho, double* h_drho);
int gato_sim_forward_f64(gato_solver64* s, double* d_xnext, const double* d_x, const double* d_u, double dt,
                         void* stream);

/* Extension: joint-space tracking terms 1/2 sum_i Qd_i (x_i - xref_{k,i})^2 + 1/2 sum_i Rd_i u_i^2
 * replacing qd_cost / u_cost (NULL pointers restore the reference cost). Device pointers
 * d_Qd [B * nx], d_Rd [B * nu], d_xref [B * N * nx]; copied, so they may be freed afterwards. */
int gato_set_joint_cost(gato_solver* s, const float* d_Qd, const float* d_Rd, const float* d_xref, void* stream);

/* set_f_ext_batch (bsqp.cuh:63; bindings.cu:150-154): d_f_ext [B * 6] device pointer */
int gato_set_f_ext_batch(gato_solver* s, const float* d_f_ext, void* stream);
/* set_rho_penalty_batch (bsqp.cuh:66-70): host pointer [B] */
int gato_set_rho_penalty_batch(gato_solver* s, const float* h_rho, int32_t set_as_reset_default);
/* set_drho_batch (bsqp.cuh:72-76): host pointer [B] */
int gato_set_drho_batch(gato_solver* s, const float* h_drho, int32_t set_as_reset_default);
/* set_mu_batch (bsqp.cuh:78): host pointer [B] */
int gato_set_mu_batch(gato_solver* s, const float* h_mu);
/* set_pcg_tol_batch (bsqp.cuh:79): host pointer [B] */
int gato_set_pcg_tol_batch(gato_solver* s, const float* h_tol);
/* reset_dual (bsqp.cuh:81) */
int gato_reset_dual(gato_solver* s);
/* reset_rho (bsqp.cuh:83-87) */
int gato_reset_rho(gato_solver* s);
/* set_rho_adaptation (bsqp.cuh:89) */
int gato_set_rho_adaptation(gato_solver* s, int32_t enabled);
/* Extension: select GATO_MODE_* / GATO_EXIT_* after construction */
int gato_set_mode(gato_solver* s, int32_t mode, int32_t exit_rule);

/* Extension: persistent state for deterministic tests. d_lambda [B * N * nx], h_rho/h_drho [B];
 * any pointer may be NULL to skip it. */
int gato_get_state(gato_solver* s, float* d_lambda, float* h_rho, float* h_drho);
int gato_set_state(gato_solver* s, const float* d_lambda, const float* h_rho, const float* h_drho);

/* sim_forward (bsqp.cuh:91; sim.cuh:14-88; bindings.cu:180-194): one trapezoidal step of the
 * device state d_x [nx] under control d_u [nu] for each of the B wrenches set by
 * gato_set_f_ext_batch; d_xnext [B * nx]. */
int gato_sim_forward(gato_solver* s, float* d_xnext, const float* d_x, const float* d_u, float dt, void* stream);

/* Backward pass (no reference counterpart; the adjoint of the solve for differentiable MPC).
 * Differentiates the SQP step map z+ = z + dz(z; theta) at the solution d_xu (the trajectory a solve
 * returned, with the solver's current rho and lambda), theta = (x_s, reference xyz, diag Q, diag R), for the
 * loss gradient d_dL_dxu [B * traj]. The step is re-linearised at d_xu; the adjoint KKT system is solved with
 * that Schur complement and preconditioner by PCG (gato_set_backward_params). At a converged solution
 * (dz = 0) this is the implicit derivative of the solution under the solver's Hessian model.
 * Outputs (device): d_grad_xs [B * nx], d_grad_ref [B * 6N] (rpy entries 0), d_grad_Qd [B * nx],
 * d_grad_Rd [B * nu] - the gradients w.r.t. the diagonal weights of gato_set_joint_cost (or, without it,
 * of the weights the reference cost implies: Qd = [0, qd_cost], Rd = u_cost). h_adj_pcg_iters [B] (host,
 * may be NULL) receives the adjoint PCG iteration counts. Does not change lambda, rho or the inputs. */
int gato_backward(gato_solver* s, const float* d_xu, const float* d_xs, const float* d_ref, float dt,
                  const float* d_dL_dxu, float* d_grad_xs, float* d_grad_ref, float* d_grad_Qd, float* d_grad_Rd,
                  void* stream, int32_t* h_adj_pcg_iters);
/* Backward solves: PCG iteration cap per round (-1, the default: 1), relative tolerance on the preconditioned
 * residual r^T M^-1 r (default 1e-6) and rounds of iterative refinement of the step and adjoint solutions
 * (at most 8; -1, the default: 2 for N <= 64 and 4 beyond, where two rounds leave about 1e-2 of relative gradient
 * error against an FP64 solve and four about 2e-4). M is the twisted block-tridiagonal factorization of the Schur complement, equal to it up
 * to fp32 rounding, so one PCG iteration per round is a solve to the fp32 floor and the refinement rounds, whose
 * residuals come from the KKT blocks, correct it. */
int gato_set_backward_params(gato_solver* s, int32_t max_pcg_iters, float pcg_tol, int32_t refine);
/* The backward solves' settings in force (refine resolved to its count). */
int gato_get_backward_params(gato_solver* s, int32_t* max_pcg_iters, float* pcg_tol, int32_t* refine);

/* Backward, derivative of the solution (no reference counterpart). With rounds > 0 the backward returns the
 * implicit derivative of the solution z*(theta), defined by F(z, lambda; theta) = [g + C^T lambda; c] = 0, instead of
 * the derivative of one SQP step: the adjoint system then has the exact Jacobian of F (the Hessian of the Lagrangian,
 * dynamics curvature included, in place of the step's model G~). Each round is one iteration K w+ = [dL/dz; 0] +
 * (G~ - H) w, run as a refinement round of the adjoint solve, with H w from central differences of the stationarity
 * rows at z +- eps w, eps = fd_step / |w|_inf (default 1e-4), the two KKT set-ups evaluated in fp64; the rounds
 * contract at the SQP step's own local rate (about 6x per round where the solve converges). The result is the gradient of the solution where d_xu is one (a converged solve; with
 * the reference's terminal block, terminal_mode 0, the step has no fixed point); 0 rounds (the default) keeps the
 * one-step derivative. The first call with rounds > 0 allocates the rounds' workspace (about twice the KKT blocks). */
int gato_set_backward_implicit(gato_solver* s, int32_t rounds, float fd_step);

/* Extension: launch shape. `groups` 16-lane knot groups per work-group (a multiple of 4 in [4, 32], at most
 * the default 4*ceil(min(N,32)/4), with ceil(N / groups) <= 8); threads = 16 * groups. */
int gato_set_launch_shape(gato_solver* s, int32_t groups);
int gato_get_launch_shape(gato_solver* s, int32_t* groups, int32_t* threads);
/* Extension: launch-shape tuner. Times one whole solve per valid shape on these inputs (the trajectory is
 * solved in a scratch copy; lambda, rho and drho are restored), keeps the fastest, returns it. */
int gato_autotune(gato_solver* s, float* d_xu, const float* d_xs, const float* d_ref, float dt, void* stream,
                  int32_t* chosen_groups);

/* Uneven batching (no reference counterpart; the reference compiles one N and one plant per module).
 * A mixed solver holds `nclass` classes of problems, class i = (plants[i], knots[i], batches[i]), all with
 * `params` (params->mode is ignored: gato_mixed_set_mode chooses). By default gato_mixed_solve captures every
 * class's solve as a parallel branch of ONE hipGraph and launches it once; the exit rule applies per class. On the first solve each class's launch
 * shape is tuned on its inputs (gato_autotune) unless disabled. Per-class arrays: d_xu[i], d_xs[i], d_ref[i]
 * (device pointers, layouts as gato_solve); stats: NULL or an array of nclass gato_stats. */
typedef struct gato_mixed gato_mixed;
int gato_mixed_create(gato_mixed** out, int32_t nclass, const int32_t* plants, const int32_t* knots,
                      const int32_t* batches, const gato_params* params);
int gato_mixed_destroy(gato_mixed* m);
/* The class's solver handle, for the per-class setters (f_ext, rho, mu, joint costs, state). Owned by m. */
int gato_mixed_class(gato_mixed* m, int32_t i, gato_solver** out);
int gato_mixed_set_autotune(gato_mixed* m, int32_t enabled);
/* How gato_mixed_solve runs. GATO_MODE_GRAPH (the default): the hipGraph above. GATO_MODE_FUSED: ONE persistent
 * kernel for all classes; its work-groups (`groups` knot groups of 16 lanes each) take problems from a work list,
 * larger classes (N nx^2) first, and solve each to completion as GATO_MODE_FUSED does for one class, with the
 * class's own exit rule per problem. Every class's N / groups, rounded up, must be at most 8. groups = 0: 32,
 * and unless autotuning is off, the first solve times every valid count of 8, 16 and 32 on its inputs (state
 * restored) and keeps the fastest. The work list's counters belong to the mixed solver: its solves must not
 * overlap (gato_mixed_solve returns after the solve, so calls from one thread never do). */
int gato_mixed_set_mode(gato_mixed* m, int32_t mode, int32_t groups);
/* The current mode and the persistent kernel's knot groups per work-group (tuned or set). */
int gato_mixed_get_mode(gato_mixed* m, int32_t* mode, int32_t* groups);
int gato_mixed_solve(gato_mixed* m, float* const* d_xu, const float* const* d_xs, const float* const* d_ref, float dt,
                     void* stream, gato_stats* stats);
/* Device time of the last mixed launch (graph or persistent kernel; hipEvents on the caller's stream), microseconds. */
int gato_mixed_last_time(gato_mixed* m, double* us);

/* End-effector positions of a batch of trajectories d_xu [B * traj] (device): d_pos [B * N * 3] receives p(q_k), the
 * origin of the last joint frame as the cost uses it (iiwa14_grid.cuh:2596-2664), and d_jac [B * N * 3 * nq] (may be
 * NULL) its Jacobian dp/dq, row-major per knot. For losses on the solution's end-effector path (tracking error).
 * Asynchronous. */
int gato_ee_positions(gato_solver* s, const float* d_xu, float* d_pos, float* d_jac, void* stream);

/* MPC hypothesis selection (the caller's per-step host work in mpc_controller.py:294-309 and 240-242, done on
 * the device): one trapezoidal step of (d_x_last [nx], d_u_last [nu]) under each problem's wrench
 * (gato_set_f_ext_batch), err_b = |x_next_b - d_x_curr|_2 into d_errors [B] (device, may be NULL), best =
 * argmin in numpy.argmin's order (the first NaN if any, else the first minimum) into *h_best (host, may be
 * NULL); then every row of d_xu [B * traj] is
 * overwritten with row best (the next solve's warm start). Asynchronous unless h_best is given. */
int gato_select_hypothesis(gato_solver* s, const float* d_x_last, const float* d_u_last, const float* d_x_curr,
                           float dt, float* d_xu, float* d_errors, int32_t* h_best, void* stream);

/* Plant simulation for MPC loops (the caller's host integration, common.py:49-91, on the device): `steps` RK4
 * sub-steps of length h of every problem's state d_x [B * nx] (in place) under its control d_u [B * nu] and
 * wrench d_f_ext [B * 6] (device pointers; the wrench here is the "true" one, independent of the solver's).
 * Asynchronous. */
int gato_simulate_rk4(gato_solver* s, float* d_x, const float* d_u, const float* d_f_ext, float h, int32_t steps,
                      void* stream);
/* Plant advance of the MPC harness (mpc_controller.py:194-213, with rk4 of common.py:49-91) on the device for n
 * independent systems: sub-steps i = first_substep .. first_substep + steps - 1 of RK4 with step h, sub-step i applying
 * the control u_k of knot k = min(int(i / (knot_dt / h)), knots - 1) of the system's trajectory (the harness's offset
 * rule, evaluated in double). d_x [n * nx] states (in place); d_xu [n * ((nx+nu) knots - nu)] trajectories in the
 * solver's layout; d_f_ext [n * 6] wrench on the last link in its frame, or NULL for none. model: GATO_PLANT_IIWA14
 * (the URDF's model, which is also the solver's), GATO_PLANT_INDY7 (the solver's model) or GATO_PLANT_INDY7_URDF (the
 * plant the harness simulates). The harness's residual-time step (mpc_controller.py:206-213) is one more call with
 * steps = 1 and first_substep = the sub-step count. Asynchronous. */
int gato_simulate_plant(int32_t model, int32_t n, float* d_x, const float* d_xu, int32_t knots, const float* d_f_ext,
                        double h, int32_t first_substep, int32_t steps, double knot_dt, void* stream);
/* Reference window for every problem from a device-resident path d_path [npoints * 6]:
 * d_ref[b][k] = path[min(offset + k, npoints - 1)] (mpc_controller.py:226-229 tiles it on the host). Asynchronous. */
int gato_window_reference(gato_solver* s, float* d_ref, const float* d_path, int32_t npoints, int32_t offset,
                          void* stream);
/* The MPC examples' figure-8 end-effector path generated on the device (common.py:10-46, which builds it with
 * numpy on the host): points_per_cycle points of t in [0, 2 pi] (numpy.linspace spacing) on
 * (amp_x sin t, 0, amp_z sin(2t)/2 + amp_z/2) + offset[3], rotated by theta about z, zero rpy, repeated `cycles`
 * times, into d_path [points_per_cycle * cycles * 6] (evaluated in double, stored as float). Asynchronous. */
int gato_figure8_path(float* d_path, int32_t points_per_cycle, int32_t cycles, double amp_x, double amp_z,
                      const double* offset, double theta, void* stream);

/* Wrench-hypothesis estimator of the MPC examples (force_estimator.py:4-155 and its use in
 * mpc_controller.py:279-337) with its state on the device. Hypotheses: 0 the smoothed estimate, 1 zero, 2 the
 * estimate plus half its momentum, 3.. exploration around the estimate on a rotated Fibonacci sphere of the
 * current radius; world frame, [force; moment]. */
typedef struct gato_force_est gato_force_est;
#define GATO_FORCE_EST_STATE_LEN 30 /* estimate[6] momentum[6] smoothed[6] rotation[9] radius confidence updates */
/* batch > 3 (force_estimator.py:8); the radius starts at initial_radius and stays in [min_radius, max_radius];
 * seed drives the library's own uniforms for the exploration rotations (used when an update passes none). */
int gato_force_est_create(gato_force_est** e, int32_t batch, float initial_radius, float min_radius,
                          float max_radius, float smoothing_factor, uint64_t seed);
int gato_force_est_destroy(gato_force_est* e);
/* Back to the initial state (force_estimator.py:150-155) with the given radius. Synchronous. */
int gato_force_est_reset(gato_force_est* e, float radius, void* stream);
/* update_force_batch (mpc_controller.py:279-292) on the device: every hypothesis mapped to the solver's wrench
 * frame at joint positions d_q [nq] (device) as transform_force_to_gato_frame does (mpc_controller.py:311-337)
 * and written to the solver's per-problem wrenches (what gato_set_f_ext_batch sets); d_world [B * 6] receives
 * the world-frame hypotheses when not NULL. Batch sizes must match. Asynchronous. */
int gato_force_est_generate(gato_force_est* e, gato_solver* s, const float* d_q, float* d_world, void* stream);
/* ForceEstimator.update (force_estimator.py:99-136) on the device for the winning hypothesis *d_best (NULL: the
 * solver's last gato_select_hypothesis result) with that call's prediction errors d_errors [B]; h_uniforms: three
 * numbers in [0, 1] for the next exploration rotation (NULL: the estimator's own generator). An index outside
 * [0, B) leaves the estimator unchanged. Asynchronous. */
int gato_force_est_update(gato_force_est* e, gato_solver* s, const int32_t* d_best, const float* d_errors, float alpha,
                          float beta, const double* h_uniforms, void* stream);
/* Copy of the state (GATO_FORCE_EST_STATE_LEN doubles, layout above). Synchronous. */
int gato_force_est_state(gato_force_est* e, double* h_state, void* stream);

/* Plant dimensions: nq (= nu), nx = 2 nq. */
int gato_plant_dims(int32_t plant, int32_t* nq, int32_t* nx);

/* Extension: average device time per launch (microseconds) of the dominant kernel during the last
 * solve, measured with hipEvents on the solve's stream, and the launch count. */
int gato_last_kernel_time(gato_solver* s, double* avg_us, int32_t* launches);

/* Extension (phase-level tests, no reference counterpart): run one phase of the plain path on
 * the solver's workspace for all B problems, with the solver's current lambda / rho.
 * phase: 0 KKT setup, 1 Schur complement + preconditioner, 2 PCG, 3 dz + KKT residual,
 *        4 merit terms at the 8 step lengths (no line search),
 *        5 the PCG as the graph path runs it for 64 < N <= 128 (held on the CU, preconditioner in factored form). */
int gato_debug_phase(gato_solver* s, int32_t phase, float* d_xu, const float* d_xs, const float* d_ref, float dt,
                     void* stream);
/* Extension (phase-level tests): copy workspace array `name` ("A", "B", "c", "Qp", "Qv", "Rg", "q", "r",
 * "Qip", "Qiv", "Ri", "Sd", "So", "Pd", "Po", "gamma", "lambda", "dz", "merit_parts", "pcg_iters",
 * "kkt_residual", "f_ext", "Qd", "Rd") into the device buffer d_dst of capacity `cap` elements; *count = its length. */
int gato_debug_copy(gato_solver* s, const char* name, void* d_dst, int64_t cap, int64_t* count);

/* Description of the last failure on the calling thread ("" if none). */
const char* gato_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
