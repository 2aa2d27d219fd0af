// thermal_state.hpp -- set-up of the thermal state search
// (dm-heom/include/heom/thermal_state_search_solver.hpp:124-209): the eigen decomposition
// of the state Hamiltonian, the Boltzmann top ADO the search starts from, and the
// transform constants of heom_create_tss.
#pragma once

#include <vector>

#include "config.hpp"
#include "hierarchy.hpp"

namespace heom {

// Eigen decomposition of a real symmetric n x n matrix a (row-major) by cyclic Jacobi
// rotations (Golub and Van Loan, Matrix Computations, 4th ed., algorithm 8.5.3):
// values[k] and the orthonormal eigenvectors as the columns of vectors (row-major).
void symmetric_eigen(int n, std::vector<double> a, std::vector<double>& values, std::vector<double>& vectors);

struct thermal_setup {
    std::vector<double> c;            // [N*N*2] C = V^T, V the eigenvectors as columns
    std::vector<double> eigenvalues;  // [N*2] (J)
    std::vector<complex_t> top;       // [N*N] sigma_0 = C^T diag(exp(-beta E) / Z) C
    complex_t epsilon;                // the pseudo-Hamiltonian's constant
};

// thermal_state_search_solver::initialize for the instance's Hamiltonian (J), which must
// be real symmetric; a ground state (with_ground_state) gets the energy
// constants::tss_ground_state_energy so that it stays empty.
thermal_setup thermal_state_setup(const instance& inst, real_t temperature);

}  // namespace heom
