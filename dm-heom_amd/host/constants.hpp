// constants.hpp -- physical constants and Pade tables of the HEOM path.
//
// The values are the reference's own (dm-heom/include/heom/constants.hpp:21-33,
// dm-heom/cl/heom/common.cl:45): an older CODATA hbar and k_B that every prefactor
// depends on, so they are reproduced digit for digit.  The Pade tables are the
// bosonic [N-1/N] expansion coefficients of dm-heom/src/heom/constants.cpp:13-76
// (row K-1 serves K "matsubaras" per bath); they are data and are reproduced value for
// value in constants.cpp.
#pragma once

#include <cstddef>

namespace heom {
namespace constants {

constexpr double h_bar = 1.05457148e-34;
constexpr double k_b = 1.380650239684840458953781387048428305608616777679979316e-23;
constexpr double invcm_to_joule = 1.986456217327470903861498336971100097137908380913187891e-23;

// thermal state search: the ground-state energy that keeps it empty
// (dm-heom/include/heom/constants.hpp:27) and the constant of the pseudo-Hamiltonian
// (dm-heom/include/heom/thermal_state_search_solver.hpp:201-206)
constexpr double tss_ground_state_energy = 100000;
constexpr double tss_epsilon = 1.0e14;

constexpr std::size_t max_pade = 10;
extern const long double eta_pade[max_pade * max_pade];
extern const long double xi_pade[max_pade * max_pade];

}  // namespace constants
}  // namespace heom
