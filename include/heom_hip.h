/*
 * heom_hip.h -- C ABI of the MI355X HEOM solver library (libheom_hip.so).
 *
 * This is the drop-in boundary for DM-HEOM's ODE plug-in point: the reference
 * instantiates population_dynamics_solver<heom::ode, num::meta_stepper,
 * heom::hierarchy_norm> (dm-heom/include/heom/population_dynamics_solver.hpp:31-32,
 * dm-heom/src/app_population_dynamics.cpp:59-60); everything that template drives on
 * the device -- the heom_ode RHS, the RK4 stepper with rk_weighted_add, the
 * hierarchy_norm kernel, mmult_left, the state double buffer and its read/write-back
 * -- sits behind the functions below.  Plain C types only; one handle owns all device
 * memory of one hierarchy; no pointer passed in is retained after a call returns.
 *
 * Data contract at the boundary is the reference's (dm-heom/include/heom/instance.hpp
 * :119-171): ADOs in the graph's uid order, each N x N complex matrix row-major with
 * interleaved (re, im) doubles; tables int32 [M * W], invalid edge = any negative value;
 * prefactors as produced by instance::precompute_baths_input (SI units).  Inside, the
 * library stores ADOs structure-of-arrays by matrix element in its own ADO order; the
 * layout never leaks through the ABI.
 *
 * Errors: every function returns HEOM_OK (0) or a negative HEOM_E_* code and sets a
 * thread-local message readable with heom_last_error().  No C++ exception crosses the
 * ABI (the reference's ocl::error_handler / config_error throw instead,
 * dm-heom/src/heom/ode.cpp:52, dm-heom/include/heom/config_error.hpp:17-23).
 *
 * Threading: one handle is used by one host thread at a time; calls on a handle are
 * ordered on its own HIP stream.  Calls that return data to the host synchronise.
 */
#ifndef HEOM_HIP_H
#define HEOM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HEOM_ABI_VERSION 2

enum heom_status {
    HEOM_OK = 0,
    HEOM_E_INVALID = -1,     /* bad argument or descriptor */
    HEOM_E_UNSUPPORTED = -2, /* valid but not supported by this build (e.g. N too large) */
    HEOM_E_DEVICE = -3,      /* HIP runtime error */
    HEOM_E_MEMORY = -4       /* host or device allocation failed */
};

enum heom_precision {
    HEOM_PRECISION_FP64 = 0, /* complex128 ADO storage, FP64 RHS and RK4 */
    HEOM_PRECISION_FP32 = 1  /* complex64 ADO storage and FP32 RHS; the top ADO and the tier-1
                                ADOs (tuple sum <= 1) are re-evaluated and accumulated in FP64,
                                and tops, populations and traces are read from FP64 */
};

/* Static description of one hierarchy (the reference's heom::instance plus the
 * generated #defines of solver_base::ocl_source_header,
 * dm-heom/include/heom/solver_base.hpp:179-195). */
typedef struct heom_desc {
    int32_t states;              /* N = NUM_STATES */
    int32_t matsubaras;          /* K = NUM_MATSUBARAS (modes per bath) */
    int32_t baths;               /* B */
    int32_t baths_max_per_state; /* BATHS_MAX_PER_STATE */
    int32_t width;               /* W = ADO_WIDTH = B * K */
    int32_t precision;           /* enum heom_precision */
    int64_t matrices;            /* M = NUM_MATRICES */
    const int32_t* coupling;     /* BATHS_COUPLING_LIST [N * baths_max_per_state], <0 = none */
    const double* hamiltonian;   /* [N * N * 2] */
    const int32_t* ado_tuples;   /* [M * W] */
    const int32_t* ado_plus;     /* [M * W] */
    const int32_t* ado_minus;    /* [M * W] */
    const double* pf_same_sum;   /* [2 * M] */
    const double* pf_same;       /* [2 * B] */
    const double* pf_circ;       /* [2 * B] */
    const double* cbk;           /* [2 * B * K] */
} heom_desc;

typedef struct heom_handle heom_handle;

/* Timing of the calls made so far on a handle (device time from HIP events). */
typedef struct heom_timing_info {
    int64_t rk4_steps;         /* RK4 steps taken */
    double rk4_seconds;        /* device time spent in RK4 stepping */
    int64_t rhs_launches;      /* stage kernel launches (4 per RK4 step, 1 per heom_rhs) */
    double last_call_seconds;  /* device time of the last stepping call */
} heom_timing_info;

/* ---- lifetime ----------------------------------------------------------------- */

/* Replaces heom::ode::ode + ode::initialise (dm-heom/src/heom/ode.cpp:35-95) and the
 * state buffers of population_dynamics_solver::initialize
 * (dm-heom/include/heom/population_dynamics_solver.hpp:122-159).  Copies every array
 * of desc to the device.  The initial state is all zeros and the mask all TRUE. */
int heom_create(const heom_desc* desc, heom_handle** out);
/* A batch of `count` hierarchies in one handle (no counterpart in the reference, whose
 * apps run one instance at a time; for ensembles such as static-disorder realisations
 * and parameter scans).  The instances share the precision; their Hamiltonians, depths,
 * bath prefactors (pf_same, pf_circ, cbk), coupling, baths, matsubaras and states may
 * differ.  They are stepped together: one kernel launch per RK4 stage for the whole
 * batch.  Per-instance arguments below ("[count]" arrays: top, hamiltonian, mmult_left
 * matrices, observables, recorded tops and populations) are indexed by instance; per-ADO
 * arrays (state, mask, norms, rhs) are the instances' arrays one after the other.
 * Every such array uses the batch's N = the largest `states` (and tuples the largest
 * `width`): an instance with fewer states has them first, in its own order, and rows
 * and columns past them that stay exactly zero (states without a bath and without
 * Hamiltonian couplings, which leave the instance's own block unchanged); an instance
 * with fewer modes has its tuples first and zeros after them.  heom_create is
 * heom_create_batch with count = 1. */
int heom_create_batch(const heom_desc* descs, int32_t count, heom_handle** out);
int heom_instances(heom_handle* handle, int32_t* count);
int heom_destroy(heom_handle* handle);
/* message of the last failing call on this thread ("" if none) */
const char* heom_last_error(void);
int heom_abi_version(void);

/* ---- inputs -------------------------------------------------------------------- */

/* ode::update_hamiltonian (dm-heom/src/heom/ode.cpp:197-203) */
int heom_set_hamiltonian(heom_handle* handle, const double* hamiltonian /* [count][N*N*2] */);
/* solver_base::update_hierarchy_mask (dm-heom/include/heom/solver_base.hpp:292-299):
 * values 0 = FALSE, 1 = TRUE, 2 = PROTECTED; as in heom_ode.cl:199,237,297,360 a term
 * is dropped when the NEIGHBOUR it reads is masked. */
int heom_set_mask(heom_handle* handle, const uint8_t* mask /* [M] */);
int heom_get_mask(heom_handle* handle, uint8_t* mask /* [M] */);
/* population_dynamics_solver::reset (population_dynamics_solver.hpp:161-175):
 * the whole state, reference layout. */
int heom_set_state(heom_handle* handle, const double* state /* [M*N*N*2] */);
/* solver_base::update_hierarchy_top (solver_base.hpp:283-289): overwrite ADO 0 only */
int heom_set_top(heom_handle* handle, const double* top /* [count][N*N*2] */);
/* instance::null_hierarchy (dm-heom/src/heom/instance.cpp:448-453) on the device:
 * every ADO zero (with heom_set_top: a fresh initial value without a host-side state) */
int heom_clear_state(heom_handle* handle);

/* ---- outputs ------------------------------------------------------------------- */

/* solver_base::hierarchy (solver_base.hpp:213-216) */
int heom_get_state(heom_handle* handle, double* state /* [M*N*N*2] */);
/* solver_base::hierarchy_top (solver_base.hpp:219-222) */
int heom_get_top(heom_handle* handle, double* top /* [count][N*N*2] */);
/* Tuple-keyed readback, so that no ADO order is part of the contract: the ADOs of
 * instance `instance` whose tuples (n_0 .. n_{W-1}, the reference's ado_tuples rows,
 * include/heom/instance.hpp:125-126) are given, out [n][N*N*2].  A tuple that is not an
 * ADO of the instance is an error; ADOs removed by heom_compact read as zero. */
int heom_get_ados(heom_handle* handle, int32_t instance, const int32_t* tuples /* [n][W] */, int64_t n,
                  double* out);
/* solver_base::update_hierarchy_norm + hierarchy_norm (solver_base.hpp:224-228,
 * 270-273; kernel cl/heom/hierarchy_norm.cl:15-45): norms[n] = sum_ij |sigma_n,ij|^2 in
 * uid order; ADOs with norm - threshold <= 0 that are not PROTECTED are masked
 * (a negative threshold never triggers). */
int heom_hierarchy_norm(heom_handle* handle, double threshold, double* norms /* [M] */);

/* ---- evaluation and stepping ----------------------------------------------------- */

/* One RHS evaluation f(current state) into out (reference layout): ode::solve's pure
 * evaluation (dm-heom/src/heom/ode.cpp:97-117, heom_ode.cl:430-434), as used by the
 * rate-kernel app (dm-heom/src/app_rate_kernel.cpp:91-93). The state is unchanged. */
int heom_rhs(heom_handle* handle, double* out /* [M*N*N*2] */);

/* ---- per-evaluation interface: the operator between device state buffers -------------
 *
 * The reference's ODE wrapper evaluates the operator between device buffers in four
 * forms (ode::solve, dm-heom/include/heom/ode.hpp:40-43, dm-heom/src/heom/ode.cpp:97-196,
 * kernel side dm-heom/cl/heom/heom_ode.cl:416-434) and calls its pre-evaluate actions first
 * (ode.hpp:53-55, ode.cpp:17-21); its steppers (noma/num, not vendored) compose steps from
 * those calls.  These entry points give a caller the same building blocks on the handle's
 * state buffers: buffer 0 is the state (heom_set_state, heom_get_state, heom_get_top),
 * buffers 1-3 are the work buffers of heom_step_rk4 and heom_propagate* (which overwrite
 * them), buffers 4.. are added with heom_add_buffers.  FP64 handles only (FP32 storage
 * keeps FP64 shadows of the tier-0/1 ADOs that these buffers do not carry); all calls are
 * ordered on the handle's stream.  heom_compact keeps buffer 0 and zeroes the others. */
#define HEOM_MAX_BUFFERS 16
#define HEOM_NO_BUFFER (-1)
/* Called by heom_evaluate before the evaluation, in the order added, with heom_evaluate's
 * time, step_size, in and out; may change the handle's Hamiltonian (heom_set_hamiltonian),
 * e.g. a time-dependent field.  A non-zero return fails the evaluation with that code's
 * message. */
typedef int (*heom_pre_evaluate_fn)(heom_handle* handle, double time, double step_size, int in, int out,
                                    void* user);
/* ode::add_pre_evaluate_action / clear_pre_evaluate_actions (ode.hpp:53-55). */
int heom_add_pre_evaluate(heom_handle* handle, heom_pre_evaluate_fn fn, void* user);
int heom_clear_pre_evaluate(heom_handle* handle);
/* f = L(buffer in), then
 *   out >= 0:  buffer out = (out_base >= 0 ? buffer out_base : 0) + out_coeff * f
 *   acc >= 0:  buffer acc = (acc_init ? buffer in : buffer acc) + acc_coeff * f
 * (out != acc; either may be the same buffer as in or out_base).  The reference's forms:
 *   ode::solve(t, h, in, out)                          -> (t, h, in, out, -1, 1, -1, 0, 0)
 *   ode::solve(t, h, in, out, acc, c, init)            -> (t, h, in, out, -1, 1, acc, c, init)
 *   ode::solve(t, h, in, out, oc, acc, c, init) with the sub-diagonal stage input
 *                                                      -> (t, h, in, out, base, oc, acc, c, init)
 * (the form ode::solve(t, h, in, out, c) scales only the flow tracking, which this build
 * does not carry, and is the pure evaluation here).  Asynchronous. */
int heom_evaluate(heom_handle* handle, double time, double step_size, int in, int out, int out_base,
                  double out_coeff, int acc, double acc_coeff, int acc_init);
/* buffer dst = sum_k coeff[k] * buffer src[k], 1 <= n <= 8; dst may be a source.  No
 * reference counterpart (noma/num's weighted adds); the building block of general
 * explicit Runge-Kutta tableaus.  Asynchronous. */
int heom_lincomb(heom_handle* handle, int dst, int n, const int* src, const double* coeff);
/* exchanges two buffers (their device addresses); buffer 0 stays "the state". */
int heom_swap_buffers(heom_handle* handle, int a, int b);
/* adds n zeroed state buffers (at most HEOM_MAX_BUFFERS in all); *count = buffers now
 * (count may be NULL). */
int heom_add_buffers(heom_handle* handle, int n, int* count);
int heom_buffer_count(heom_handle* handle, int* count);
/* a buffer from / into the reference layout, like heom_set_state / heom_get_state. */
int heom_set_buffer(heom_handle* handle, int b, const double* state /* [M*N*N*2] */);
int heom_get_buffer(heom_handle* handle, int b, double* state /* [M*N*N*2] */);

/* nsteps classical RK4 steps of size h (seconds): population_dynamics_solver::
 * step_forward (population_dynamics_solver.hpp:233-265) with the stepper's heom_ode and
 * rk_weighted_add launches fused into one kernel per stage.  Asynchronous. */
int heom_step_rk4(heom_handle* handle, double h, int64_t nsteps);
/* nsteps RK4 steps, recording the top ADO on the device every observe_steps steps
 * (first record = the state before stepping); rows = nsteps / observe_steps + 1 records
 * of N*N*2 doubles are written to tops at the end.  The reference reads the top back
 * synchronously after every step_forward(observe_steps) (app_population_dynamics.cpp:
 * 92-106); this does the same without a host round trip per observation. */
int heom_propagate(heom_handle* handle, double h, int64_t nsteps, int64_t observe_steps,
                   double* tops /* [nsteps/observe_steps + 1][count][N*N*2] */);
/* nsteps RK4 steps, observing every instance's top ADO on the device every
 * observe_steps steps (first record = before stepping): the populations
 * Re sigma_0[i][i] of dm-heom/include/heom/matrix_diagonal_observer.hpp:28-34 into
 * pops [rows][count][N] and, when ops [count][N*N*2] is given, the traces
 * Tr(ops_k sigma_0) of matrix_trace_observer.hpp:139-158 into traces [rows][count][2]
 * (rows = nsteps / observe_steps + 1).  The reductions run on the device in FP64 (with
 * FP32 storage, on the FP64 top), one wavefront per instance. */
int heom_propagate_observe(heom_handle* handle, double h, int64_t nsteps, int64_t observe_steps,
                           const double* ops, double* pops, double* traces);
/* heom_propagate_observe without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code around HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The observation
 * records and the operators live in a fixed device scratch array of
 * HEOM_OBSERVE_SCRATCH_BYTES per GPU, held until the process ends; a call whose
 * records do not fit (about rows * count * (8 N + 16) bytes plus 16 N^2 count for ops)
 * fails with HEOM_E_UNSUPPORTED and changes nothing, so a long run is split into calls of
 * fewer steps.  The stage kernels are launched one by one even when step graphs are
 * enabled (no graph is captured or instantiated).  Results equal heom_propagate_observe
 * bit for bit.  All calls on one GPU share its scratch: calls that overlap in time on
 * one GPU (other threads, other handles) race and must be serialised by the caller. */
#define HEOM_OBSERVE_SCRATCH_BYTES (8u << 20)
int heom_propagate_observe_noalloc(heom_handle* handle, double h, int64_t nsteps, int64_t observe_steps,
                                   const double* ops, double* pops, double* traces);
/* The observations of the current state, as heom_propagate_observe records them: the
 * populations Re sigma_0[i][i] (matrix_diagonal_observer.hpp:28-34) into pops
 * [count][N] and, when ops [count][N*N*2] is given, Tr(ops_k sigma_0)
 * (matrix_trace_observer.hpp:139-158) into traces [count][2]; FP64 reductions on the
 * device. */
int heom_observe(heom_handle* handle, const double* ops, double* pops, double* traces);
/* A time-dependent Hamiltonian: the reference's e_field pre-evaluate action
 * (dm-heom/include/heom/e_field.hpp:64-87, hooked in at include/heom/ode.hpp:57-59) sets
 * H = H0 - E(t) D- - conj(E(t)) D+ before every RHS evaluation, at that evaluation's
 * time.  heom_propagate_timedep is heom_propagate with those Hamiltonians given for the
 * RK4 stage times t_n, t_n + h/2, t_n + h/2, t_n + h of every step:
 * stage_hamiltonians [nsteps][4][count][N*N*2] (J).  Steps are not graph-captured. */
int heom_propagate_timedep(heom_handle* handle, double h, int64_t nsteps, int64_t observe_steps,
                           const double* stage_hamiltonians, double* tops /* [rows][count][N*N*2] */);
/* Thermal state search (dm-heom/include/heom/thermal_state_search_solver.hpp:124-300).
 * heom_create_tss builds a search handle for one instance: desc as for heom_create
 * (FP64; its Hamiltonian is not used, the search replaces it by the pseudo-Hamiltonian
 * with epsilon in every element, :199-209), c = C [N*N*2] (the transposed
 * eigenvectors of H, :160-161), eigenvalues [N*2] of H in J and epsilon [2].  One
 * heom_tss_step iteration is the reference's step_forward(1) (:267-299):
 *   k = f_TSS(y) (heom_ode.cl with HEOM_TSS_EQUATION),  k <- C k C^T,
 *   y_u = k_u / (i (E_i - E_j) / hbar + epsilon - pf_same_sum_u)  (thermal_state_search.cl:36-48),
 *   y <- C^T y C.
 * Load the start (the Boltzmann top, heom_host_tss_setup) with heom_set_top; state,
 * top, norms, mask and heom_rhs (= f_TSS) work as on any handle; RK4 propagation,
 * Hamiltonian updates, compaction and the adjoint are refused on a search handle. */
int heom_create_tss(const heom_desc* desc, const double* c, const double* eigenvalues, const double* epsilon,
                    heom_handle** out);
/* The search for `count` instances at once (as heom_create_batch: same states and baths,
 * any depth): c [count][N*N*2], eigenvalues [count][N*2], one epsilon.  Every
 * heom_tss_step iterates all of them; heom_create_tss is the batch of one. */
int heom_create_tss_batch(const heom_desc* descs, int32_t count, const double* c, const double* eigenvalues,
                          const double* epsilon, heom_handle** out);
int heom_tss_step(heom_handle* handle, int64_t n);
/* solver_base::hierarchy_mmult_left (solver_base.hpp:250-257; kernel
 * cl/heom/hierarchy_arithmetic_mmult_left.cl:15-50): sigma_n <- mx * sigma_n */
int heom_mmult_left(heom_handle* handle, const double* mx /* [count][N*N*2] */);
/* Step through a captured HIP graph of one RK4 step (its stage launches) instead of
 * launching the stage kernels one by one (no counterpart in the reference; it cuts
 * the host launch cost per step for small hierarchies and batches).  Off by default. */
int heom_use_graphs(heom_handle* handle, int enable);
/* wait for all work queued on the handle */
int heom_sync(heom_handle* handle);
int heom_timing(heom_handle* handle, heom_timing_info* info);

/* ---- introspection ------------------------------------------------------------- */

/* bytes of device memory held by the handle */
int heom_device_bytes(heom_handle* handle, int64_t* bytes);
/* algorithmic HBM bytes of one RK4 step of the whole hierarchy
 * (16 state passes x M x N^2 x sizeof(complex), complex128 or complex64) */
int heom_rk4_step_bytes(heom_handle* handle, int64_t* bytes);
/* which stage kernel the handle runs (no counterpart in the reference; for tests and
 * benchmarks): HEOM_STAGE_TILE = one workgroup per 64-ADO tile, HEOM_STAGE_PIPELINED =
 * persistent workgroups streaming tiles through LDS (N >= 6 and few bath modes),
 * HEOM_STAGE_WIDE = one wavefront per ADO for 8 < N <= 32.  A hierarchy of fewer
 * 64-ADO tiles than half the CUs runs the tile kernel with each tile split over 2 or 4
 * workgroups.  The environment variable HEOM_STAGE_KERNEL at heom_create overrides the
 * choice for N <= 8: tile (one workgroup per tile), split (the tile kernel, largest
 * split) or pipelined (where it applies). */
enum heom_stage_kind { HEOM_STAGE_TILE = 0, HEOM_STAGE_PIPELINED = 1, HEOM_STAGE_WIDE = 2 };
int heom_stage_kernel(heom_handle* handle, int* kind);

/* Hierarchy compaction (filtering with compaction; no counterpart in the reference,
 * which keeps computing masked ADOs): rebuilds the device layout with only the ADOs
 * whose mask is not FALSE (the mask of heom_set_mask, or of heom_hierarchy_norm with a
 * threshold).  A masked ADO is never read by its neighbours, so the other ADOs -- the
 * tops, populations and traces among them -- evolve exactly as before, at the cost of
 * the remaining ADOs only.  Removed ADOs read back as zero and stay removed (a later mask
 * cannot bring them back); the top ADO of every instance must be unmasked.  FP64
 * handles.  active receives the number of ADOs kept (may be NULL). */
int heom_compact(heom_handle* handle, int64_t* active);

/* The autotuner (no counterpart in the reference): times every valid configuration of
 * the stage kernels for this handle -- tile or pipelined kernel, tiles per XCD chunk of
 * the tile order (powers of two up to 64, 24, 48, 72, 96 and the default: 48 for the
 * pipelined kernel on large hierarchies, where multiples of 24 measured fastest) and
 * workgroups per tile of the tile kernel (N <= 8), ADOs per workgroup
 * (N > 8, HEOM_STAGE_WIDE) -- over `steps`
 * RK4 steps each on the handle's own state (restored afterwards) and keeps the fastest.
 * Writes min(count, capacity) results; count receives the number of candidates. */
typedef struct heom_tune_result {
    int32_t kernel;              /* enum heom_stage_kind */
    int32_t xcd_chunk;           /* tiles per XCD chunk of the tile order */
    int32_t workgroups;          /* pipelined kernel: persistent workgroups, else 0 */
    int32_t ados_per_workgroup;  /* 64, or the wide kernel's ADOs per workgroup */
    int32_t chosen;              /* 1 for the configuration kept */
    int32_t tile_split;          /* tile kernel: workgroups per 64-ADO tile (1, 2 or 4), else 1 */
    double ms_per_step;          /* device time of one RK4 step */
} heom_tune_result;
int heom_autotune(heom_handle* handle, int32_t steps, heom_tune_result* results, int32_t capacity,
                  int32_t* count);
/* Applies one configuration of the stage kernels without timing anything (no
 * counterpart in the reference): a result of heom_autotune on a handle of the same
 * shape, so a tuned configuration can be reused.  kernel, workgroups (pipelined kernel)
 * and ados_per_workgroup must be a candidate heom_autotune would time for this handle;
 * xcd_chunk lies in [1, tiles]; chosen and ms_per_step are ignored.  The tiles per XCD
 * chunk do not change a bit of the results; the kernels agree to rounding. */
int heom_set_stage_config(heom_handle* handle, const heom_tune_result* config);

/* ---- adjoint: gradients of observables ------------------------------------------ */
/* No counterpart in the reference, which has no differentiation; for fitting
 * Hamiltonians and bath parameters to measured dynamics and spectra.  FP64 handles with
 * an unfiltered hierarchy (no masked ADOs), N*N <= 64, and in a batch instances that share
 * states, modes, coupling and bath prefactors (the bath gradients are their sums).
 *
 * The objective is J = sum_r sum_k Re(w_rk X_rk) with X_rk = Tr(O_k sigma_0,k(t_r)), the
 * observable O_k on the top ADO of instance k at t_r = r * observe_steps * h,
 * r = 0 .. nsteps / observe_steps.  heom_adjoint_forward propagates from the current
 * state like heom_propagate_observe, records X and leaves the state at t = nsteps * h.
 * It keeps the state and the three RK4 stage vectors of every step up to the last
 * observation on the device when those 4 * L state copies (L = the last observation's
 * step) fit in half the free device memory; otherwise, or with HEOM_ADJOINT_RECOMPUTE=1
 * in the environment, the state every checkpoint_steps steps.  heom_adjoint_backward
 * takes the weights w (the cotangents of X) and integrates the adjoint of the RK4
 * trajectory backwards (from checkpoints it recomputes each segment and each step's
 * stage vectors; both modes give identical results), with the transposed operator in
 * hand-written kernels; it fills the gradients of J with respect to the coefficient
 * arrays of heom_desc, each complex entry as dJ/dRe + i dJ/dIm.  Any pointer may be NULL
 * to skip that output.  Setting the Hamiltonian or the mask drops the record.  Device
 * memory: 4 * L + 4 state copies kept, or (L / checkpoint_steps + checkpoint_steps + 7)
 * with checkpoints. */
typedef struct heom_gradients {
    double* hamiltonian;  /* [count][N*N*2] (per joule) */
    double* pf_same_sum;  /* [M*2] */
    double* pf_same;      /* [B*2] (Delta_b), summed over the instances */
    double* pf_circ;      /* [B*2] (Phi_b), summed over the instances */
    double* cbk;          /* [B*K*2] (c_bk), summed over the instances */
    double* state;        /* [M*N*N*2] the state at the start of the forward sweep */
} heom_gradients;

int heom_adjoint_forward(heom_handle* handle, double h, int64_t nsteps, int64_t observe_steps,
                         int64_t checkpoint_steps, const double* ops /* [count][N*N*2] */,
                         double* observed /* [rows][count][2] */);
int heom_adjoint_backward(heom_handle* handle, const double* weights /* [rows][count][2] */,
                          heom_gradients* grads);
/* one evaluation of the transposed operator on the current state, L^dag sigma, into out
 * [M*N*N*2] (the state is unchanged) */
int heom_rhs_adjoint(heom_handle* handle, double* out);

#ifdef __cplusplus
}
#endif

#endif /* HEOM_HIP_H */
