/*
 * heom_oracle.h -- CPU oracle for the HEOM population-dynamics / linear-absorption
 * path of DM-HEOM.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this code; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, as the
 * checker.  Every function restates a piece of the reference statement by statement
 * (FP64, no FMA contraction) and cites the reference file:line it follows.
 *
 * Layout is the reference's: a hierarchy state is M ADOs, each an N x N complex matrix
 * stored row-major with interleaved (re, im) doubles:
 *   sigma[(n * N * N + i * N + j) * 2 + {0, 1}]        (cl/heom/heom_ode.cl:53-59)
 * Tables are int32 [M * W] (tuples, plus, minus), invalid edges are -2
 * (src/heom/hierarchy_graph.cpp:24).
 */
#ifndef HEOM_ORACLE_H
#define HEOM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- hierarchy graph (src/heom/hierarchy_graph.cpp, src/heom/integer_partition.cpp) ---- */

/* number of partitions of n into at most s parts (integer_partition.cpp:20-38) */
int64_t oracle_partition_count(int n, int s);
/* partitions of n into at most s parts, in the reference generator's order, each row
 * padded with zeros to s entries; returns the count (integer_partition.cpp:41-105) */
int64_t oracle_partitions(int n, int s, int32_t* out);
/* M = C(W + D, W) (hierarchy_graph.cpp:34) */
int64_t oracle_graph_nodes(int width, int depth);
/* first ADO of the deepest layer, C(W + D - 1, W) (hierarchy_graph.cpp:40) */
int64_t oracle_graph_first_non_plus(int width, int depth);
/* fill tuples/plus/minus ([M * W] each); returns 0 on success */
int oracle_graph_build(int width, int depth, int32_t* tuples, int32_t* plus, int32_t* minus,
                       int64_t* nodes_per_depth /* [depth + 1] or NULL */);

/* ---- precompute (src/heom/instance.cpp:155-262) ---- */

/* Inputs in SI units after config::post_process (src/heom/config.cpp:182-190). */
void oracle_precompute(int baths, int matsubaras, const double* lambda, const double* invnu,
                       const double* omega, double temperature, int64_t matrices, int width,
                       const int32_t* tuples, double* pf_same_sum /* 2M */,
                       double* pf_same /* 2B */, double* pf_circ /* 2B */,
                       double* cbk /* 2BK */, double* cbk_sqrt /* BK */);

/* ---- the RHS, one evaluation (cl/heom/heom_ode.cl:24-451) ---- */

typedef struct oracle_instance {
    int states;                 /* N (NUM_STATES) */
    int matsubaras;             /* K (NUM_MATSUBARAS) */
    int baths_max_per_state;    /* BATHS_MAX_PER_STATE */
    int width;                  /* W (ADO_WIDTH) */
    int64_t matrices;           /* M (NUM_MATRICES) */
    int64_t first_non_plus;     /* first_non_plus_id */
    const int32_t* coupling;    /* [N * baths_max_per_state], negative = none (BATHS_COUPLING) */
    const double* hamiltonian;  /* [N * N * 2] */
    const int32_t* tuples;      /* [M * W] */
    const int32_t* plus;        /* [M * W] */
    const int32_t* minus;       /* [M * W] */
    const double* pf_same_sum;  /* [2M] */
    const double* pf_same;      /* [2B] */
    const double* pf_circ;      /* [2B] */
    const double* cbk;          /* [2BK] */
    const uint8_t* mask;        /* [M], NULL = all MASK_TRUE */
} oracle_instance;

/* out = f(in): pure evaluation (heom_ode.cl:430-434, neither ACCUMULATE nor SUBDIAGONAL) */
void oracle_rhs(const oracle_instance* inst, const double* in, double* out);
/* out = f(in) of the thermal state search (heom_ode.cl with HEOM_TSS_EQUATION,
 * :116-121 and :130-136): the commutator becomes the elementwise product
 * sigma_ij H_ij with the pseudo-Hamiltonian in inst->hamiltonian, and the S_n sigma
 * term is dropped */
void oracle_rhs_tss(const oracle_instance* inst, const double* in, double* out);

/* out = y_n + h * sum_{l < n, c_l != 0} c_l k_l (cl/heom/rk_weighted_add.cl:19-106) */
void oracle_rk_weighted_add(int64_t size, int n, double h, double* out, const double* y_n,
                            const double* c, const double* const* k);

/* One classical RK4 step, unfused: 4 RHS evaluations + 4 weighted adds.
 * y is advanced in place; work must hold 5 states (k1..k4, stage input). */
void oracle_rk4_step(const oracle_instance* inst, double h, double* y, double* work);

/* nsteps RK4 steps; writes the top ADO (N*N*2 doubles) after every observe_steps
 * steps into tops (first row = initial value); returns number of rows written. */
int64_t oracle_propagate(const oracle_instance* inst, double h, int64_t nsteps,
                         int64_t observe_steps, double* y, double* work, double* tops);

/* ---- observers and hierarchy arithmetic ---- */

/* norm_n = sum_ij |sigma_n[i,j]|^2 (cl/heom/hierarchy_norm.cl:15-45); the threshold
 * deactivates unprotected ADOs in mask when norm - threshold <= 0 (mask may be NULL). */
void oracle_hierarchy_norm(int states, int64_t matrices, const double* sigma, double threshold,
                           double* norms, uint8_t* mask);
/* sigma_n <- Mx * sigma_n for every ADO (cl/heom/hierarchy_arithmetic_mmult_left.cl:15-50) */
void oracle_mmult_left(int states, int64_t matrices, const double* mx, double* sigma);
/* Tr(D * top) (include/heom/matrix_trace_observer.hpp:139-158); D NULL = identity */
void oracle_trace(int states, const double* d, const double* top, double* trace2);

/* ---- time-dependent Hamiltonian (include/heom/e_field.hpp:64-87) ---- */
/* one RK4 step with stage Hamiltonians H4 [4][N*N*2] (stage times t, t+h/2, t+h/2, t+h) */
void oracle_rk4_step_timedep(const oracle_instance* inst, double h, double* y, double* work, const double* H4);
/* nsteps such steps, stage_h [nsteps][4][N*N*2]; tops as oracle_propagate */
int64_t oracle_propagate_timedep(const oracle_instance* inst, double h, int64_t nsteps, int64_t observe_steps,
                                 double* y, double* work, const double* stage_h, double* tops);
/* ---- adjoint (no reference counterpart: the mathematics of the discrete adjoint) ---- */
/* out = L^dag mu, the transpose of oracle_rhs as a linear map of the state */
void oracle_rhs_adjoint(const oracle_instance* inst, const double* mu, double* out);
/* adds w * d<mu, L y>/d(kappa) (complex entries as dRe + i dIm) for the coefficient
 * arrays kappa: hamiltonian [2NN], pf_same_sum [2M], pf_same [2B], pf_circ [2B], cbk [2BK] */
void oracle_adjoint_contract(const oracle_instance* inst, const double* mu, const double* y, double w,
                             double* g_h, double* g_pss, double* g_same, double* g_circ, double* g_cbk);
/* J = sum_r Re(w_r Tr(O sigma_0(t_r))) over nsteps RK4 steps from y0 (observations every
 * observe_steps steps, row 0 at t = 0): observed X_r [2 * rows] and the gradients of J
 * with respect to the coefficient arrays and (g_state, may be NULL) the initial state,
 * by the discrete adjoint with checkpoints every checkpoint_steps steps */
int oracle_adjoint_gradient(const oracle_instance* inst, double h, int64_t nsteps, int64_t observe_steps,
                            int64_t checkpoint_steps, const double* op, const double* weights, const double* y0,
                            double* observed, double* g_h, double* g_pss, double* g_same, double* g_circ,
                            double* g_cbk, double* g_state);
/* number of OpenMP threads the oracle runs on */
/* sigma <- sigma mx for every ADO (hierarchy_arithmetic_mmult_right.cl:15-50) */
void oracle_mmult_right(int states, int64_t matrices, const double* mx, double* sigma);
/* one thermal state search iteration (thermal_state_search_solver.hpp:267-299) */
void oracle_tss_step(const oracle_instance* inst, const double* c, const double* eigenvalues, double eps_r,
                     double eps_i, double* y, double* work);
int oracle_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif /* HEOM_ORACLE_H */
