/*
 * heom_host.h -- C ABI of the host side (libheom_host.so): the reference's
 * config-file instance API for population dynamics and linear absorption.
 *
 * heom_host_open reads a reference config file (INI with noma/typa literals) and builds
 * what dm-heom's apps build before they touch a device -- config (parse, check, SI
 * units: dm-heom/src/heom/config.cpp:37-190), hierarchy_graph
 * (dm-heom/src/heom/hierarchy_graph.cpp:30-181), instance with its sites-to-states
 * mapping and precomputed bath prefactors (dm-heom/src/heom/instance.cpp:25-69,155-262),
 * and the initial top ADO (rho_init, or |g><g| for linear absorption).  The result
 * feeds heom_create of include/heom_hip.h through heom_host_desc.  Errors return a
 * negative code and set heom_host_last_error(); config problems are
 * HEOM_HOST_E_CONFIG, mirroring heom::config_error.
 */
#ifndef HEOM_HOST_H
#define HEOM_HOST_H

#include <stdint.h>

#include "heom_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum heom_host_status { HEOM_HOST_OK = 0, HEOM_HOST_E_INVALID = -1, HEOM_HOST_E_CONFIG = -5 };

enum heom_host_task {
    HEOM_TASK_POPULATION_DYNAMICS = 0,
    HEOM_TASK_LINEAR_ABSORPTION = 1,
    HEOM_TASK_THERMAL_STATE_SEARCH = 2
};

typedef struct heom_host_instance heom_host_instance;

typedef struct heom_host_info {
    int64_t sites;
    int64_t states;          /* sites, or sites + 1 with a ground state */
    int64_t baths;
    int64_t matsubaras;
    int64_t width;
    int64_t depth;
    int64_t matrices;
    int64_t first_non_plus;
    double step_size;        /* s */
    int64_t steps;
    int64_t observe_steps;
    int64_t tensor_components; /* linear absorption: dipole tensor components */
    double tss_delta;          /* thermal state search: thermal_state_search.delta */
    int64_t tss_max_steps;     /* thermal state search: thermal_state_search.max_steps */
} heom_host_info;

/* task: enum heom_host_task; depth_override > 0 replaces system.ado_depth */
int heom_host_open(const char* config_path, int task, int64_t depth_override, heom_host_instance** out);
int heom_host_close(heom_host_instance* inst);
const char* heom_host_last_error(void);

int heom_host_get_info(const heom_host_instance* inst, heom_host_info* info);
/* descriptor for heom_create; its arrays belong to inst and live until heom_host_close */
int heom_host_desc(const heom_host_instance* inst, heom_desc* desc);
/* initial top ADO [N*N*2] */
int heom_host_top(const heom_host_instance* inst, double* top);
/* sqrt(|c_m|) per mode [B*K] (instance::cbk_sqrt, kept for the reference layout) */
int heom_host_cbk_sqrt(const heom_host_instance* inst, double* out);
/* the filtering mask of the config [M] (dm-heom/src/heom/hierarchy_mask.cpp:20-110) */
int heom_host_mask(const heom_host_instance* inst, uint8_t* mask);
/* linear absorption: D- (prefactor applied) and D+ of tensor component t, [N*N*2] each */
int heom_host_dipoles(const heom_host_instance* inst, int64_t t, double* d_minus, double* d_plus);
/* Parameters for fitting (no counterpart in the reference, which reads them from the
 * config only), all in cm^-1: the site Hamiltonian [sites*sites*2] (complex), and per
 * bath lambda_b, gamma_b = hbar / invnu_b as an energy, and Omega_b [B].  set replaces
 * the given ones (NULL keeps) and recomputes the prefactors; arrays from heom_host_desc
 * keep their addresses and hold the new values.  get reads them back. */
int heom_host_set_parameters(heom_host_instance* inst, const double* hamiltonian, const double* lambda,
                             const double* gamma, const double* omega);
int heom_host_get_parameters(const heom_host_instance* inst, double* hamiltonian, double* lambda, double* gamma,
                             double* omega);
/* The gradients of heom_adjoint_backward (of this instance's slice for a batch) as
 * gradients with respect to those parameters, all per cm^-1: d_hamiltonian
 * [sites*sites*2] (dJ/dRe H_ij + i dJ/dIm H_ij, entries independent), d_lambda,
 * d_gamma, d_omega [B].  Any output may be NULL. */
int heom_host_parameter_gradient(const heom_host_instance* inst, const heom_gradients* grads,
                                 double* d_hamiltonian, double* d_lambda, double* d_gamma, double* d_omega);
/* thermal_state_search_solver::initialize (dm-heom/include/heom/thermal_state_search_solver.hpp:124-209)
 * for the instance's current Hamiltonian, which must be real symmetric: c = C [N*N*2]
 * (the transposed eigenvectors), eigenvalues [N*2] (J), top [N*N*2] the Boltzmann state
 * the search starts from, epsilon [2] the pseudo-Hamiltonian's constant; outputs may be
 * NULL.  An instance opened with HEOM_TASK_THERMAL_STATE_SEARCH already holds this top. */
int heom_host_tss_setup(const heom_host_instance* inst, double* c, double* eigenvalues, double* top,
                        double* epsilon);
/* The reference's binary instance format (dm-heom/src/heom/instance.cpp:281-377): a
 * checkpoint of a hierarchy with every kernel-input array.  write stores this instance
 * with the state [M*N*N*2] (state NULL or top_only: the top ADO only) and a time;
 * read_instance_state checks that the file describes this instance (sizes, ADO tables)
 * and returns its hierarchy (zeros below the top for a top-only file) and time.  The
 * format holds no bath coupling list, so a file is read back against its config. */
int heom_host_write_instance(const heom_host_instance* inst, const char* path, const double* state, int top_only,
                             double time);
int heom_host_read_instance_state(const heom_host_instance* inst, const char* path, double* state, double* time);
/* eigen decomposition of a real symmetric n x n matrix a (row-major) by cyclic Jacobi
 * rotations, the routine heom_host_tss_setup uses (the reference calls Eigen's
 * ComplexEigenSolver, include/heom/eigen_decomposition.hpp:17-36): values [n] and the
 * orthonormal eigenvectors as the columns of vectors [n*n] (row-major) */
int heom_host_symmetric_eigen(int n, const double* a, double* values, double* vectors);
/* integer partitions of n into at most s parts in the reference order; writes
 * min(count, capacity) rows of s entries, returns count (negative on error) */
int64_t heom_host_partitions(int n, int s, int32_t* out, int64_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* HEOM_HOST_H */
