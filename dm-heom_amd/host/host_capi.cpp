// host_capi.cpp -- the C ABI of include/heom_host.h over config / hierarchy / instance.
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "config.hpp"
#include "heom_host.h"
#include "hierarchy.hpp"
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
        if (task != HEOM_TASK_POPULATION_DYNAMICS && task != HEOM_TASK_LINEAR_ABSORPTION)
            throw std::invalid_argument("unknown task");
        auto p = std::make_unique<heom_host_instance>();
        const auto t = task == HEOM_TASK_POPULATION_DYNAMICS ? heom::program_task::population_dynamics
                                                             : heom::program_task::linear_absorption;
        p->cfg = std::make_unique<heom::config>(path, t);
        if (depth_override > 0)
            p->cfg->set_ado_depth(depth_override);
        p->graph = std::make_unique<heom::hierarchy_graph>(p->cfg->baths_number(), p->cfg->baths_matsubaras(),
                                                           p->cfg->system_ado_depth());
        const auto mode = task == HEOM_TASK_POPULATION_DYNAMICS ? heom::sites_to_states_mode::identity
                                                                : heom::sites_to_states_mode::with_ground_state;
        p->inst = std::make_unique<heom::instance>(*p->cfg, mode, *p->graph);
        if (task == HEOM_TASK_POPULATION_DYNAMICS)
            p->inst->set_hierarchy_top(p->cfg->population_dynamics_rho_init().data.data());
        else
            p->inst->set_hierarchy_top();
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
