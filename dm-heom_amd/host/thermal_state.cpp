// thermal_state.cpp -- see thermal_state.hpp.
#include "thermal_state.hpp"

#include <algorithm>
#include <cmath>

#include "constants.hpp"

namespace heom {

void symmetric_eigen(int n, std::vector<double> a, std::vector<double>& values, std::vector<double>& vectors)
{
    vectors.assign((std::size_t)n * n, 0.0);
    for (int k = 0; k < n; ++k)
        vectors[(std::size_t)k * n + k] = 1.0;
    double frob = 0.0;
    for (double x : a)
        frob += x * x;
    auto at = [&](int i, int j) -> double& { return a[(std::size_t)i * n + j]; };
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q)
                off += 2.0 * at(p, q) * at(p, q);
        if (off <= 1e-32 * frob)
            break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = at(p, q);
                if (apq == 0.0)
                    continue;
                // the symmetric Schur rotation that zeroes a_pq
                const double tau = (at(q, q) - at(p, p)) / (2.0 * apq);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < n; ++k) {  // A <- A J
                    const double kp = at(k, p), kq = at(k, q);
                    at(k, p) = c * kp - s * kq;
                    at(k, q) = s * kp + c * kq;
                }
                for (int k = 0; k < n; ++k) {  // A <- J^T A
                    const double pk = at(p, k), qk = at(q, k);
                    at(p, k) = c * pk - s * qk;
                    at(q, k) = s * pk + c * qk;
                }
                for (int k = 0; k < n; ++k) {  // V <- V J
                    double& vp = vectors[(std::size_t)k * n + p];
                    double& vq = vectors[(std::size_t)k * n + q];
                    const double x = vp, y = vq;
                    vp = c * x - s * y;
                    vq = s * x + c * y;
                }
            }
    }
    values.resize((std::size_t)n);
    for (int k = 0; k < n; ++k)
        values[(std::size_t)k] = at(k, k);
}

thermal_setup thermal_state_setup(const instance& inst, real_t temperature)
{
    const int N = (int)inst.states();
    const auto& h = inst.hamiltonian();
    std::vector<double> a((std::size_t)N * N);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            if (h[2 * (i * N + j) + 1] != 0.0)
                throw config_error("the thermal state search needs a real symmetric Hamiltonian");
            a[(std::size_t)i * N + j] = h[2 * (i * N + j)];
        }
    if (inst.mode() == sites_to_states_mode::with_ground_state)
        a[0] = constants::tss_ground_state_energy;  // thermal_state_search_solver.hpp:135-137
    if (!(temperature > 0.0))
        throw config_error("the thermal state search needs a temperature > 0");
    std::vector<double> E, V;
    symmetric_eigen(N, a, E, V);
    // p_k = exp(-beta E_k) / Z (:180-187), with E shifted by its minimum so that no term
    // underflows (the ratio is unchanged)
    const double beta = 1.0 / (constants::k_b * temperature);
    const double e0 = *std::min_element(E.begin(), E.end());
    std::vector<double> p((std::size_t)N);
    double z = 0.0;
    for (int k = 0; k < N; ++k)
        z += (p[(std::size_t)k] = std::exp(-(E[(std::size_t)k] - e0) * beta));
    thermal_setup s;
    s.c.assign((std::size_t)N * N * 2, 0.0);
    s.eigenvalues.assign((std::size_t)N * 2, 0.0);
    s.top.assign((std::size_t)N * N, complex_t(0.0, 0.0));
    for (int i = 0; i < N; ++i) {
        s.eigenvalues[2 * (std::size_t)i] = E[(std::size_t)i];
        for (int j = 0; j < N; ++j)
            s.c[2 * ((std::size_t)i * N + j)] = V[(std::size_t)j * N + i];  // C = V^T (:160-161)
    }
    // sigma_0 = C^T diag(p) C = V diag(p) V^T (:189-191)
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double v = 0.0;
            for (int k = 0; k < N; ++k)
                v += V[(std::size_t)i * N + k] * (p[(std::size_t)k] / z) * V[(std::size_t)j * N + k];
            s.top[(std::size_t)i * N + j] = complex_t(v, 0.0);
        }
    s.epsilon = complex_t(constants::tss_epsilon, 0.0);
    return s;
}

}  // namespace heom
