/*
 * igr.h — C ABI of the MI355X (gfx950) IGR-regularized Euler library.
 *
 * Replaces the hot path of the reference Julia packages
 * f-t-s/IGR_sensitivities (IGR_adjoints_1d/, IGR_adjoints_2d/): the nodal
 * Gauss–Lobatto DG step with LLF faces, the matrix-free SIP solve for the
 * entropic pressure Sigma (fixed-iteration PCG / damped Jacobi / Chebyshev),
 * SSP-RK3, the forward tangent (sensitivity) equations and the discrete
 * adjoint that the reference obtains from Enzyme.
 *
 * Conventions
 *  - All arrays are host pointers to contiguous FP64 data in the reference's
 *    column-major layout: 2D node (i,j,ex,ey) at i + np*(j + np*(ex+Nex*ey)),
 *    1D node (i,e) at i + np*e (np = p+1).  A "state" stores its conserved
 *    variables back to back: 2D (mx, my, rho, E), 1D (mu, rho, E), each of
 *    N = nodes entries (IGR_adjoints_2d/euler.jl:5-9).
 *  - Every function returns an IGR_* status.  Julia's ErrorException for an
 *    unknown solver/objective symbol maps to IGR_EINVAL; Julia's DomainError
 *    from sqrt of a negative pressure maps to IGR_EDOMAIN.
 *  - The library owns device memory inside the context; state lives in HBM
 *    between calls.  One HIP stream per context; a context is not
 *    thread-safe (the reference is serial), but distinct contexts may be used
 *    from distinct host threads at the same time.
 */
#ifndef IGR_H
#define IGR_H

#ifdef __cplusplus
extern "C" {
#endif

enum igr_status {
  IGR_OK = 0,
  IGR_EINVAL = 1,   /* bad argument, unknown solver or objective code */
  IGR_EDOMAIN = 2,  /* negative pressure under sqrt / non-finite state */
  IGR_EHIP = 3,     /* HIP runtime failure */
  IGR_ENOMEM = 4    /* device allocation failed */
};

/* Replaces the `solver::Symbol` argument (:pcg, :jacobi, :chebyshev) of
 * solve_elliptic! (IGR_adjoints_2d/elliptic.jl:413-427). */
enum igr_solver { IGR_SOLVER_PCG = 0, IGR_SOLVER_JACOBI = 1, IGR_SOLVER_CHEBYSHEV = 2 };

typedef struct igr_ctx igr_ctx;

/* Mesh + basis set-up: DGBasis(p) (IGR_adjoints_2d/basis.jl:56-68) and
 * CartesianMesh2D(nex, ney, lx, ly) (IGR_adjoints_2d/mesh.jl:40-57) for
 * dim = 2, or PeriodicMesh1D(nex, lx) (IGR_adjoints_1d/mesh.jl:21-33) for
 * dim = 1 (ney, ly ignored).  1 <= p <= 4. */
int igr_create(igr_ctx** ctx, int dim, int p, int nex, int ney, double lx, double ly);
int igr_destroy(igr_ctx* ctx);
const char* igr_last_error(const igr_ctx* ctx);
/* Number of nodes per field (N) and conserved variables (nv). */
int igr_sizes(const igr_ctx* ctx, long long* n_nodes, int* n_vars);

/* Options:
 *  IGR_OPT_GRAPHS   1 = capture each SSP-RK3 step in a hipGraph and replay it,
 *                   0 = launch kernels directly;
 *  IGR_OPT_TILE     2D elements per workgroup of the SIP / solver kernels,
 *                   tile_x*100 + tile_y (p = 3: 404, 804, 808, 202;
 *                   p = 4: 204, 404, 505, 805);
 *  IGR_OPT_HTILE    the same for the hyperbolic, tangent and objective kernels
 *                   (p = 3: 404, 804, 202);
 *  IGR_OPT_AUTOTUNE (value ignored) time the forward shapes on the current
 *                   state and keep the fastest of each family (2D, p = 3
 *                   and 4);
 *  IGR_OPT_RTILE    the same for the reverse SIP / solver kernels of the
 *                   gradient and the VJPs (p = 3: 404, 804, 808, 202);
 *  IGR_OPT_GROUPS   workgroups of the PCG x/r update kernel (2D): 0 = one per
 *                   SIP tile, n > 0 = at most n, each running every n-th
 *                   tile with the next tile's loads issued before the
 *                   current tile's stores (the default comes from the size
 *                   table; results change only in the dot products' order);
 *  IGR_OPT_WAVE1D   1D: 1 (default) = a plain run of N_e = 64 elements, or
 *                   128 up to p = 3, runs in one wavefront with no barrier
 *                   (elements in registers, faces by lane rotates), 0 = the
 *                   one-workgroup kernel; results agree to rounding;
 *  IGR_OPT_KEEP     2D gradient: the most whole steps whose solver iterates
 *                   are kept in HBM (-1, the default: as many as 85% of the
 *                   free memory holds; 0: none, the plain binomial schedule).
 *                   With every step checkpointed they are the last steps of
 *                   the forward sweep; with fewer checkpoint slots a segment
 *                   of up to that many steps + 1 is reversed after one
 *                   storing pass, without recomputation.  Gradients are
 *                   bit-identical for every value;
 *  IGR_OPT_RGROUPS  workgroups of the 2D gradient's reverse PCG iteration:
 *                   0 = one per reverse tile, n > 0 = at most n, each
 *                   running every n-th tile and storing one partial sum per
 *                   dot product (the default comes from the size table;
 *                   results change only in the dot products' order).
 * The defaults come from a size table measured on MI355X.  The shapes change
 * results only through the summation order of the solvers' dot products. */
enum igr_option {
  IGR_OPT_GRAPHS = 0, IGR_OPT_TILE = 1, IGR_OPT_HTILE = 2, IGR_OPT_AUTOTUNE = 3, IGR_OPT_RTILE = 4,
  IGR_OPT_GROUPS = 5, IGR_OPT_WAVE1D = 6, IGR_OPT_KEEP = 7, IGR_OPT_RGROUPS = 8
};
int igr_set_option(igr_ctx* ctx, int option, int value);
int igr_get_option(const igr_ctx* ctx, int option, int* value);

/* Upload / download the state q (nv*N) and Sigma (N; may be NULL). */
int igr_set_state(igr_ctx* ctx, const double* q, const double* sigma);
int igr_get_state(igr_ctx* ctx, double* q, double* sigma);

/* compute_hyperbolic_rhs! on host arrays (IGR_adjoints_2d/hyperbolic.jl:27;
 * IGR_adjoints_1d/hyperbolic.jl:23). */
int igr_hyperbolic_rhs(igr_ctx* ctx, double gamma, const double* q, const double* sigma, double* dq);
/* apply_sip!(y, x, rho, basis, mesh, alpha) (IGR_adjoints_2d/elliptic.jl:30;
 * IGR_adjoints_1d/elliptic.jl:32). */
int igr_apply_sip(igr_ctx* ctx, double alpha, const double* x, const double* rho, double* y);
/* compute_igr_rhs! (IGR_adjoints_2d/elliptic.jl:159) on a state q. */
int igr_igr_rhs(igr_ctx* ctx, double alpha, const double* q, double* b);
/* compute_sip_diagonal (IGR_adjoints_2d/elliptic.jl:205). */
int igr_sip_diagonal(igr_ctx* ctx, double alpha, const double* rho, double* diag);

/* solve_elliptic!(Sigma, q..., alpha, n_iter, solver) on the context state:
 * Sigma (context) is the warm start and is overwritten
 * (IGR_adjoints_2d/elliptic.jl:413-427; IGR_adjoints_1d/elliptic.jl:451). */
int igr_solve_elliptic(igr_ctx* ctx, double alpha, int n_iter, int solver);
/* PCG scalars of the last PCG solve: a_k and beta_k, k < n_iter. */
int igr_solver_trace(igr_ctx* ctx, int n_iter, double* a, double* beta);

/* compute_igr_euler_rhs! on the context state (IGR_adjoints_2d/forward.jl:22):
 * Sigma solve when alpha > 0 (else Sigma = 0), then the hyperbolic RHS,
 * written to dq (host, nv*N; may be NULL). */
int igr_euler_rhs(igr_ctx* ctx, double gamma, double alpha, int n_iter, int solver, double* dq);

/* run_forward!(q..., Sigma, dt, T, basis, mesh, gamma, alpha, n_iter, solver)
 * (IGR_adjoints_2d/forward.jl:41-75; IGR_adjoints_1d/forward.jl:50-87):
 * Sigma <- 0, n_steps = ceil(T/dt), dt = T/n_steps, SSP-RK3, final Sigma
 * solve when alpha > 0.  Advances the context state in place. */
int igr_run_forward(igr_ctx* ctx, double dt, double t_end, double gamma, double alpha, int n_iter,
                    int solver, int* n_steps_out);
/* The fixed-step loop inlined in the Enzyme objectives
 * (IGR_adjoints_2d/enzyme_adjoint.jl:50-70): Sigma <- 0, n_steps steps of
 * size dt, no final solve. */
int igr_run_steps(igr_ctx* ctx, int n_steps, double dt, double gamma, double alpha, int n_iter,
                  int solver);

/* Forward tangent (sensitivity) equations advanced together with the
 * primal: run_sensitivity (IGR_adjoints_2d/sensitivity.jl:353-413; 1D
 * IGR_adjoints_1d/sensitivity.jl:475-515).  The primal initial state is the
 * context state (on return it holds q(T), Sigma(T)); s0 holds n_tan tangent
 * states (nv*N each, back to back).  Returns the tangents s(T) and their
 * varsigma(T) (N each).  2D: n_steps = ceil(T/dt), dt = T/n_steps; 1D: the
 * reference's `while t < T - 1e-14; dt = min(dt, T - t)` loop.  form (1D
 * only): 1 = elliptic_rhs :form1, 2 = :form2; sgam: the gamma source
 * s_gamma rho e of the tangent pressure, a d/dgamma direction (1D:
 * sensitivity_flux, sensitivity.jl:36-50; 2D: the same term, which the
 * reference's 2D tangent does not have, sensitivity.jl:15-16; 0 gives the
 * reference's 2D tangent).  The dissipation of the tangent LLF flux uses the
 * primal wave speed. */
int igr_run_sensitivity(igr_ctx* ctx, int n_tan, const double* s0, double dt, double t_end, double gamma,
                        double alpha, int n_iter, int solver, int form, double sgam, double* s_out,
                        double* vs_out, int* n_steps_out);

/* One instance of a batched ensemble: the arguments of its own
 * igr_run_sensitivity call.  status and n_steps are outputs. */
typedef struct igr_batch_item {
  igr_ctx* ctx;          /* the instance's mesh; its state set with igr_set_state */
  double dt, t_end;      /* stepping as in igr_run_sensitivity */
  double gamma, alpha;
  double sgam;           /* gamma source of the tangent flux, as in igr_run_sensitivity */
  int n_tan;             /* 0..4 tangent directions */
  const double* s0;      /* n_tan*nv*N tangent initial states (host); NULL when n_tan = 0 */
  double* s_out;         /* optional: n_tan*nv*N tangents s(T) */
  double* vs_out;        /* optional: n_tan*N varsigma(T) */
  int n_steps;           /* out: SSP-RK3 steps taken */
  int status;            /* out: IGR_* of this instance */
} igr_batch_item;

/* Batched ensembles (no reference counterpart: the reference runs
 * run_sensitivity (IGR_adjoints_2d/sensitivity.jl:353; IGR_adjoints_1d/
 * sensitivity.jl:475) once per instance in a serial loop, e.g. the gamma and
 * shift sweeps of its paper scripts).  Runs every item's run_sensitivity at
 * once: the 1D items in one persistent launch driven by a work list, the 2D
 * items as batched launches over (instance, tile) work items with each step
 * replayed from a hipGraph.  Every item gets exactly the result its own
 * igr_run_sensitivity call would give.  n_iter, solver and form (1D) are
 * shared; each context may appear once.  Returns the first failing item's
 * status (each item's own status in items[k].status). */
int igr_batch_run_sensitivity(igr_batch_item* items, int n_items, int n_iter, int solver, int form);

/* Objectives: replace the `objective::Symbol` of enzyme_adjoint_ic
 * (IGR_adjoints_2d/enzyme_adjoint.jl:240-254).  0..4 are the reference's
 * :l2_density, :l2_pressure, :kinetic_energy (KE / mass),
 * :windowed_kinetic_energy (window prm = (xc, yc, sigma); NULL prm =
 * KE_WINDOW (0.55, 0.50, 0.15), enzyme_adjoint.jl:26) and
 * :weighted_momentum; 5 = int rho exp(-|x-c|^2/sigma^2) and 6 = kinetic
 * energy inside the disc |x-c| < R (prm = (xc, yc, R)) have no counterpart
 * in the reference. */
enum igr_objective {
  IGR_OBJ_L2_DENSITY = 0, IGR_OBJ_L2_PRESSURE = 1, IGR_OBJ_KINETIC_ENERGY = 2,
  IGR_OBJ_WINDOWED_KE = 3, IGR_OBJ_WEIGHTED_MOMENTUM = 4, IGR_OBJ_WINDOWED_DENSITY = 5,
  IGR_OBJ_DISC_KE = 6
};
/* Objective value on the context state (2D). */
int igr_objective(igr_ctx* ctx, int objective, const double* prm, double gamma, double* J);

/* Initial conditions, built on the device straight into the context state
 * (Sigma is left as it is).  Each replaces one of the reference's builders,
 * evaluated at the mesh nodes with the same expressions:
 *   2D, IGR_adjoints_2d/initial_conditions.jl:
 *     IGR_IC_CONSTANT           init_constant_state (:18-25)     prm = (rho0, ux0, uy0, p0)
 *     IGR_IC_ISENTROPIC_VORTEX  init_isentropic_vortex (:42-71)  prm = (beta, ux0, uy0)
 *     IGR_IC_TAYLOR_GREEN       build_taylor_green_ic! (:105-121), amplitude A
 *     IGR_IC_SEDOV_BLAST        build_sedov_blast_ic! (:136-151)  A; prm = (p_bg, sigma)
 *     IGR_IC_OSEEN_PAIR_BLAST   build_oseen_pair_blast_ic! (:175-232)  A; prm = (a_core,
 *                               M_star, blast_xc, blast_yc, sigma, then xc, yc, sign per
 *                               vortex); its enthalpy-defect table is built on the device
 *   1D, IGR_adjoints_1d/initial_conditions.jl:
 *     IGR_IC_CONSTANT           init_constant_state (:52-60)     prm = (rho0, u0, p0)
 *     IGR_IC_ENTROPY_WAVE_INIT  init_entropy_wave (:24-44)
 *     IGR_IC_ACOUSTIC_PULSE, IGR_IC_ENTROPY_WAVE, IGR_IC_DENSITY_PULSE,
 *     IGR_IC_PRESSURE_PULSE, IGR_IC_VELOCITY_PULSE   build_<kind>_ic! (:94-210), A
 * n_prm = 0 takes the reference's keyword defaults.  dq_out (host, nv*N,
 * optional) receives d q0 / d A in conservative variables: the tangent seed
 * of the amplitude (ic_dA_weights, 2D :246-258, 1D :227-244).  An IC of the
 * other dimension or an unknown code is IGR_EINVAL (the reference's
 * "Unknown IC type" error). */
enum igr_ic {
  IGR_IC_CONSTANT = 0, IGR_IC_ISENTROPIC_VORTEX = 1, IGR_IC_TAYLOR_GREEN = 2, IGR_IC_SEDOV_BLAST = 3,
  IGR_IC_OSEEN_PAIR_BLAST = 4, IGR_IC_ENTROPY_WAVE_INIT = 5, IGR_IC_ACOUSTIC_PULSE = 6, IGR_IC_ENTROPY_WAVE = 7,
  IGR_IC_DENSITY_PULSE = 8, IGR_IC_PRESSURE_PULSE = 9, IGR_IC_VELOCITY_PULSE = 10
};
int igr_build_ic(igr_ctx* ctx, int ic, double A, double gamma, const double* prm, int n_prm, double* dq_out);

/* Discrete adjoint: enzyme_adjoint_ic(q0..., n_steps, dt, basis, mesh,
 * gamma, alpha; n_iter, solver, objective) (IGR_adjoints_2d/enzyme_adjoint.jl:
 * 270-296; IGR_adjoints_1d/enzyme_adjoint.jl:255-314) extended to gamma and
 * alpha.  q0 is the context state; the map is the Enzyme objectives' loop
 * (Sigma <- 0, n_steps fixed steps of dt).  1D objectives: L2_DENSITY,
 * L2_PRESSURE, KINETIC_ENERGY, WEIGHTED_MOMENTUM (1D enzyme_adjoint.jl:65-71,
 * 110-117); the 1D gradient is returned in conservative variables (the
 * reference's primitive form, :306-311, is a pointwise map of it).
 * Returns J, dJ/dq0 (nv*N, conservative variables; may be NULL), dJ/dgamma
 * and dJ/dalpha.  ckpt_slots: number of (q, Sigma) checkpoints kept in HBM
 * (<= 0: as many as half the free HBM allows, up to n_steps); steps between
 * checkpoints are recomputed (optimal binomial schedule).  On return the
 * context state holds q(T).  Any Sigma solver: PCG is reversed through its
 * stored iterates and scalars, Jacobi and Chebyshev through their iterates. */
int igr_gradient(igr_ctx* ctx, int objective, const double* prm, int n_steps, double dt, double gamma,
                 double alpha, int n_iter, int solver, int ckpt_slots, double* J, double* qbar0,
                 double* gammabar, double* alphabar);

/* Reverse-mode building blocks (2D), each the exact VJP of the forward kernel:
 * compute_hyperbolic_rhs! (hyperbolic.jl:27): dq-bar -> q-bar (4N), Sigma-bar, gamma-bar. */
int igr_hyperbolic_vjp(igr_ctx* ctx, double gamma, const double* q, const double* sigma, const double* dqbar,
                       double* qbar, double* sigmabar, double* gammabar);
/* apply_sip! (elliptic.jl:30): y-bar -> x-bar = A^T y-bar, rho-bar, alpha-bar. */
int igr_apply_sip_vjp(igr_ctx* ctx, double alpha, const double* x, const double* rho, const double* ybar,
                      double* xbar, double* rhobar, double* alphabar);
/* solve_elliptic! (elliptic.jl:413, any solver) on the context state q and warm
 * start sigma0: Sigma-bar -> q-bar (4N), sigma0-bar, alpha-bar. */
int igr_solve_elliptic_vjp(igr_ctx* ctx, double alpha, int n_iter, int solver, const double* sigma0,
                           const double* sigmabar, double* qbar, double* sigma0bar, double* alphabar);

/* Continuous ("optimise-then-discretise") adjoint.
 * run_forward_store(q0..., dt, T, basis, mesh, gamma, alpha; n_iter, solver)
 * (IGR_adjoints_2d/forward.jl:111-158; IGR_adjoints_1d/forward.jl:137-187):
 * from the context state, Sigma <- 0, a Sigma solve (alpha > 0), then
 * n_steps = ceil(T/dt) SSP-RK3 steps of dt = T/n_steps, each followed by a
 * Sigma solve; the (state, Sigma) after the initial solve and after every
 * step are kept in HBM as records 0..n_steps, which replace the returned
 * `snapshots` vector.  On return the context state holds q(T). */
int igr_run_forward_store(igr_ctx* ctx, double dt, double t_end, double gamma, double alpha, int n_iter,
                          int solver, int* n_steps);
/* Number of steps stored (records 0..n_steps), -1 if none. */
int igr_snapshot_count(const igr_ctx* ctx, int* n_steps);
/* snapshots[k + 1] of the reference: state (nv*N) and Sigma (N) of record k. */
int igr_get_snapshot(igr_ctx* ctx, int k, double* q, double* sigma);
/* Replace the stored snapshots by n_steps + 1 host records of (nv + 1)*N
 * doubles each (state then Sigma). */
int igr_set_snapshots(igr_ctx* ctx, int n_steps, const double* records);
/* run_adjoint_conservative(snapshots, n_steps, dt, T, basis, mesh, gamma,
 * alpha; n_iter, objective, solver) (IGR_adjoints_2d/adjoint_pde.jl:522-611;
 * IGR_adjoints_1d/adjoint_pde.jl:466-552) over the stored snapshots: the
 * terminal condition from the last record, then n_steps backward SSP-RK3
 * steps of min(dt, T), step n using record n; writes the adjoint fields at
 * t = 0 (nv*N, conservative variables).  Objectives (enum igr_objective): in
 * 2D all of them (prm as for igr_objective; WINDOWED_DENSITY and DISC_KE have
 * no counterpart in the reference), in 1D L2_DENSITY, L2_PRESSURE,
 * KINETIC_ENERGY and WEIGHTED_MOMENTUM; anything else is IGR_EINVAL (the
 * reference's ErrorException). */
int igr_run_adjoint_conservative(igr_ctx* ctx, int objective, const double* prm, int n_steps, double dt,
                                 double t_end, double gamma, double alpha, int n_iter, int solver,
                                 double* a_out);

/* Frees the device memory the library keeps between contexts: the block that
 * holds the gradient's whole-step iterates (up to 85% of the free HBM) stays
 * allocated after its context is destroyed, because obtaining that much fresh
 * HBM again costs seconds; the next gradient on the device reuses it.  Blocks
 * in use by a live gradient workspace are not freed.  The library's own device
 * allocations free an idle cached block by themselves when the device is out
 * of memory; other allocators (PyTorch, another library) need this call.
 * Thread-safe.  No counterpart in the
 * reference (Julia's GC owns its tape). */
int igr_release_cache(void);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): solve_elliptic! on caller-owned DEVICE arrays with no
 * context.  mx, my, rho (N doubles each, N = (p+1)^2 nex ney, 2D layout as
 * above) are read; sigma (N) is the warm start and receives the result, as
 * in igr_solve_elliptic.  The solver's work vectors, partial sums and scalar
 * record live in a fixed device array of the library (16.1 MiB per GPU, kept
 * until the process ends); nothing is allocated and nothing is synchronized:
 * the work is enqueued on `stream` (a hipStream_t; NULL = the null stream)
 * and the call returns.  Limits: N <= 262144 (e.g. 128 x 128 elements at
 * p = 3), n_iter <= 1024; larger sizes, a bad solver code or a bad mesh
 * return IGR_EINVAL without enqueuing anything.  domain_flag (device int,
 * optional) is set to 0 and then to 1 if a density is non-positive (the
 * reference's DomainError); read it after the stream completes.  Calls on
 * one GPU share the array, so calls that overlap on one GPU (other streams,
 * other host threads) race; consecutive calls on one stream are safe.  The
 * first call on a device resolves the array's address and must not be made
 * under graph capture. */
int igr_solve_elliptic_noalloc(int p, int nex, int ney, double lx, double ly, double alpha, int n_iter,
                               int solver, const double* mx, const double* my, const double* rho, double* sigma,
                               int* domain_flag, void* stream);

/* Mean device time (ms) per launch of one hot kernel on the current state,
 * each launch stamped with start/end events by hipExtLaunchKernelGGL on the
 * context stream.  kernel 0 (fused SIP apply + PCG direction update) and 1
 * (PCG x/r update) are timed inside a PCG solve of `reps` iterations;
 * kernel 2 (hyperbolic RHS + RK3 stage update) over `reps` launches.
 * kernel 3: kernel 0 of iteration 1 launched `reps` times back to back with
 * events around the batch only (per-launch stamps add their own overhead):
 * the mean in-stream time per launch, gaps included; kernel 4: the same for
 * kernel 1; kernel 5: the same for kernel 2. */
int igr_time_kernel(igr_ctx* ctx, int kernel, double gamma, double alpha, int reps, double* ms_per_launch);

#ifdef __cplusplus
}
#endif

#endif /* IGR_H */
