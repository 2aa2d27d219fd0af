// host_capi.cpp -- the C ABI of include/heom_host.h over config / hierarchy / instance.
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "config.hpp"
#include "constants.hpp"
#include "heom_host.h"
#include "hierarchy.hpp"
#include "thermal_state.hpp"
#include "solver.hpp"

struct heom_host_instance {
    std::unique_ptr<heom::config> cfg;
    std::unique_ptr<heom::hierarchy_graph> graph;
    std::unique_ptr<heom::instance> inst;
};

namespace {

thread_local std::string g_host_error;

template <typename F>
int host_guard(F&& fn)
{
    try {
        fn();
        g_host_error.clear();
        return HEOM_HOST_OK;
    } catch (const heom::config_error& e) {
        g_host_error = e.what();
        return HEOM_HOST_E_CONFIG;
    } catch (const std::exception& e) {
        g_host_error = e.what();
        return HEOM_HOST_E_INVALID;
    }
}

void need(const heom_host_instance* p)
{
    if (!p)
        throw std::invalid_argument("instance is NULL");
}

}  // namespace

extern "C" {

const char* heom_host_last_error(void) { return g_host_error.c_str(); }

int heom_host_open(const char* path, int task, int64_t depth_override, heom_host_instance** out)
{
    if (!out) {
        g_host_error = "out is NULL";
        return HEOM_HOST_E_INVALID;
    }
    *out = nullptr;
    return host_guard([&] {
        if (!path)
            throw std::invalid_argument("config path is NULL");
        if (task != HEOM_TASK_POPULATION_DYNAMICS && task != HEOM_TASK_LINEAR_ABSORPTION &&
            task != HEOM_TASK_THERMAL_STATE_SEARCH)
            throw std::invalid_argument("unknown task");
        auto p = std::make_unique<heom_host_instance>();
        const auto t = task == HEOM_TASK_POPULATION_DYNAMICS  ? heom::program_task::population_dynamics
                       : task == HEOM_TASK_LINEAR_ABSORPTION ? heom::program_task::linear_absorption
                                                              : heom::program_task::thermal_state_search;
        p->cfg = std::make_unique<heom::config>(path, t);
        if (depth_override > 0)
            p->cfg->set_ado_depth(depth_override);
        p->graph = std::make_unique<heom::hierarchy_graph>(p->cfg->baths_number(), p->cfg->baths_matsubaras(),
                                                           p->cfg->system_ado_depth());
        // the search runs on the one-exciton states (app_thermal_state_search.cpp:60)
        const auto mode = task == HEOM_TASK_LINEAR_ABSORPTION ? heom::sites_to_states_mode::with_ground_state
                                                              : heom::sites_to_states_mode::identity;
        p->inst = std::make_unique<heom::instance>(*p->cfg, mode, *p->graph);
        if (task == HEOM_TASK_POPULATION_DYNAMICS)
            p->inst->set_hierarchy_top(p->cfg->population_dynamics_rho_init().data.data());
        else if (task == HEOM_TASK_LINEAR_ABSORPTION)
            p->inst->set_hierarchy_top();
        else
            p->inst->set_hierarchy_top(heom::thermal_state_setup(*p->inst, p->cfg->baths_temperature()).top.data());
        *out = p.release();
    });
}

int heom_host_close(heom_host_instance* inst)
{
    return host_guard([&] {
        need(inst);
        delete inst;
    });
}

int heom_host_get_info(const heom_host_instance* p, heom_host_info* info)
{
    return host_guard([&] {
        need(p);
        if (!info)
            throw std::invalid_argument("info is NULL");
        info->sites = p->cfg->system_sites();
        info->states = p->inst->states();
        info->baths = p->inst->baths();
        info->matsubaras = p->inst->matsubaras();
        info->width = p->inst->ado_width();
        info->depth = p->inst->ado_depth();
        info->matrices = p->inst->matrices();
        info->first_non_plus = p->inst->first_non_plus_id();
        info->step_size = p->cfg->solver_step_size();
        info->steps = p->cfg->solver_steps();
        info->observe_steps = p->cfg->program_observe_steps();
        info->tensor_components = (int64_t)p->cfg->dipole_tensor_prefactors().size();
        info->tss_delta = p->cfg->thermal_state_search_delta();
        info->tss_max_steps = p->cfg->thermal_state_search_max_steps();
    });
}

int heom_host_desc(const heom_host_instance* p, heom_desc* desc)
{
    return host_guard([&] {
        need(p);
        if (!desc)
            throw std::invalid_argument("desc is NULL");
        *desc = p->inst->desc();
    });
}

int heom_host_top(const heom_host_instance* p, double* top)
{
    return host_guard([&] {
        need(p);
        if (!top)
            throw std::invalid_argument("top is NULL");
        std::memcpy(top, p->inst->top().data(), p->inst->top().size() * sizeof(double));
    });
}

int heom_host_cbk_sqrt(const heom_host_instance* p, double* out)
{
    return host_guard([&] {
        need(p);
        if (!out)
            throw std::invalid_argument("out is NULL");
        std::memcpy(out, p->inst->cbk_sqrt().data(), p->inst->cbk_sqrt().size() * sizeof(double));
    });
}

int heom_host_mask(const heom_host_instance* p, uint8_t* mask)
{
    return host_guard([&] {
        need(p);
        if (!mask)
            throw std::invalid_argument("mask is NULL");
        const auto m = heom::create_hierarchy_mask(p->cfg->filtering_strategy(), *p->graph,
                                                   p->cfg->baths_matsubaras(), p->cfg->filtering_first_layer());
        std::memcpy(mask, m.data(), m.size());
    });
}

int heom_host_dipoles(const heom_host_instance* p, int64_t t, double* d_minus, double* d_plus)
{
    return host_guard([&] {
        need(p);
        if (!d_minus || !d_plus)
            throw std::invalid_argument("output is NULL");
        if (p->cfg->task() != heom::program_task::linear_absorption)
            throw std::invalid_argument("dipoles need a linear_absorption config");
        if (t < 0 || t >= (int64_t)p->cfg->dipole_tensor_prefactors().size())
            throw std::invalid_argument("tensor component out of range");
        std::vector<heom::complex_t> dm, dp;
        heom::dipole_matrices(*p->cfg, (std::size_t)t, dm, dp);
        std::memcpy(d_minus, dm.data(), dm.size() * sizeof(heom::complex_t));
        std::memcpy(d_plus, dp.data(), dp.size() * sizeof(heom::complex_t));
    });
}

int heom_host_set_parameters(heom_host_instance* p, const double* hamiltonian, const double* lambda,
                             const double* gamma, const double* omega)
{
    return host_guard([&] {
        need(p);
        heom::instance& in = *p->inst;
        const double cm = heom::constants::invcm_to_joule, hbar = heom::constants::h_bar;
        if (hamiltonian) {
            const heom::int_t S = p->cfg->system_sites();
            heom::complex_matrix H(S, S);
            for (std::size_t e = 0; e < H.data.size(); ++e)
                H.data[e] = heom::complex_t(hamiltonian[2 * e], hamiltonian[2 * e + 1]) * cm;
            const heom::complex_matrix Hs = heom::state_hamiltonian(H, in.mode());
            std::vector<double> flat(Hs.data.size() * 2);
            for (std::size_t e = 0; e < Hs.data.size(); ++e) {
                flat[2 * e] = Hs.data[e].real();
                flat[2 * e + 1] = Hs.data[e].imag();
            }
            in.set_hamiltonian(flat.data());
        }
        if (lambda || gamma || omega) {
            const std::size_t B = (std::size_t)in.baths();
            std::vector<double> l(in.lambda()), iv(in.invnu()), o(in.omega());
            for (std::size_t b = 0; b < B; ++b) {
                if (lambda)
                    l[b] = lambda[b] * cm;
                if (gamma) {
                    if (!(gamma[b] > 0.0))
                        throw std::invalid_argument("gamma must be > 0");
                    iv[b] = hbar / (gamma[b] * cm);
                }
                if (omega)
                    o[b] = omega[b] * cm / hbar;
            }
            in.set_baths(l.data(), iv.data(), o.data());
        }
    });
}

int heom_host_get_parameters(const heom_host_instance* p, double* hamiltonian, double* lambda, double* gamma,
                             double* omega)
{
    return host_guard([&] {
        need(p);
        const heom::instance& in = *p->inst;
        const double cm = heom::constants::invcm_to_joule, hbar = heom::constants::h_bar;
        if (hamiltonian) {
            const heom::int_t S = p->cfg->system_sites(), N = in.states(), off = N - S;
            for (heom::int_t i = 0; i < S; ++i)
                for (heom::int_t j = 0; j < S; ++j) {
                    const std::size_t e = (std::size_t)((i + off) * N + (j + off));
                    hamiltonian[2 * (i * S + j)] = in.hamiltonian()[2 * e] / cm;
                    hamiltonian[2 * (i * S + j) + 1] = in.hamiltonian()[2 * e + 1] / cm;
                }
        }
        for (std::size_t b = 0; b < (std::size_t)in.baths(); ++b) {
            if (lambda)
                lambda[b] = in.lambda()[b] / cm;
            if (gamma)
                gamma[b] = hbar / (in.invnu()[b] * cm);
            if (omega)
                omega[b] = in.omega()[b] * hbar / cm;
        }
    });
}

int heom_host_parameter_gradient(const heom_host_instance* p, const heom_gradients* g, double* d_hamiltonian,
                                 double* d_lambda, double* d_gamma, double* d_omega)
{
    return host_guard([&] {
        need(p);
        if (!g)
            throw std::invalid_argument("grads is NULL");
        const heom::instance& in = *p->inst;
        const double cm = heom::constants::invcm_to_joule, hbar = heom::constants::h_bar;
        if (d_hamiltonian) {
            if (!g->hamiltonian)
                throw std::invalid_argument("grads->hamiltonian is NULL");
            const heom::int_t S = p->cfg->system_sites(), N = in.states(), off = N - S;
            for (heom::int_t i = 0; i < S; ++i)
                for (heom::int_t j = 0; j < S; ++j) {
                    const std::size_t e = (std::size_t)((i + off) * N + (j + off));
                    d_hamiltonian[2 * (i * S + j)] = g->hamiltonian[2 * e] * cm;
                    d_hamiltonian[2 * (i * S + j) + 1] = g->hamiltonian[2 * e + 1] * cm;
                }
        }
        if (d_lambda || d_gamma || d_omega) {
            if (!g->pf_same_sum || !g->pf_same || !g->pf_circ || !g->cbk)
                throw std::invalid_argument("bath gradients need pf_same_sum, pf_same, pf_circ and cbk");
            const std::size_t B = (std::size_t)in.baths();
            std::vector<double> dl(B), dr(B), di(B);
            in.bath_gradient(g->pf_same_sum, g->pf_same, g->pf_circ, g->cbk, dl.data(), dr.data(), di.data());
            for (std::size_t b = 0; b < B; ++b) {
                if (d_lambda)
                    d_lambda[b] = dl[b] * cm;
                if (d_gamma)
                    d_gamma[b] = dr[b] * cm / hbar;
                if (d_omega)
                    d_omega[b] = di[b] * cm / hbar;
            }
        }
    });
}

int heom_host_tss_setup(const heom_host_instance* p, double* c, double* eigenvalues, double* top, double* epsilon)
{
    return host_guard([&] {
        need(p);
        const auto s = heom::thermal_state_setup(*p->inst, p->cfg->baths_temperature());
        if (c)
            std::copy(s.c.begin(), s.c.end(), c);
        if (eigenvalues)
            std::copy(s.eigenvalues.begin(), s.eigenvalues.end(), eigenvalues);
        if (top)
            for (std::size_t e = 0; e < s.top.size(); ++e) {
                top[2 * e] = s.top[e].real();
                top[2 * e + 1] = s.top[e].imag();
            }
        if (epsilon) {
            epsilon[0] = s.epsilon.real();
            epsilon[1] = s.epsilon.imag();
        }
    });
}

int heom_host_write_instance(const heom_host_instance* p, const char* path, const double* state, int top_only,
                             double time)
{
    return host_guard([&] {
        need(p);
        if (!path)
            throw std::invalid_argument("path is NULL");
        const double* s = state ? state : p->inst->top().data();
        p->inst->write_file(path, s, top_only != 0 || !state, time);
    });
}

int heom_host_read_instance_state(const heom_host_instance* p, const char* path, double* state, double* time)
{
    return host_guard([&] {
        need(p);
        if (!path || !state)
            throw std::invalid_argument("path and state must not be NULL");
        p->inst->read_state(path, state, time);
    });
}

int heom_host_symmetric_eigen(int n, const double* a, double* values, double* vectors)
{
    return host_guard([&] {
        if (n < 1 || !a || !values || !vectors)
            throw std::invalid_argument("n must be >= 1 and the arrays non-NULL");
        std::vector<double> m(a, a + (std::size_t)n * n), e, v;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j)
                if (m[(std::size_t)i * n + j] != m[(std::size_t)j * n + i])
                    throw std::invalid_argument("the matrix is not symmetric");
        heom::symmetric_eigen(n, m, e, v);
        std::copy(e.begin(), e.end(), values);
        std::copy(v.begin(), v.end(), vectors);
    });
}

int64_t heom_host_partitions(int n, int s, int32_t* out, int64_t capacity)
{
    int64_t count = -1;
    const int rc = host_guard([&] {
        if (n < 1 || s < 1)
            throw std::invalid_argument("n and s must be > 0");
        const auto parts = heom::integer_partitions(n, s);
        count = (int64_t)parts.size();
        for (int64_t r = 0; r < count && r < capacity && out; ++r)
            for (int c = 0; c < s; ++c)
                out[r * s + c] = parts[(std::size_t)r][(std::size_t)c];
    });
    return rc == HEOM_HOST_OK ? count : rc;
}

}  // extern "C"
