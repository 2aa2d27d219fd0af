// app_thermal_state_search.cpp -- the thermal (stationary) hierarchy of a reference config
// file (counterpart of dm-heom/src/app_thermal_state_search.cpp:37-128): parse, build the
// hierarchy graph and instance (identity sites-to-states mapping), start from the
// Boltzmann state of H, apply the filtering mask, then iterate the search on the GPU --
// at least solver.steps / observe_steps iterations of observe_steps steps, then on until
// the 1-norm of the top changes by less than thermal_state_search.delta across an
// iteration or max_steps / observe_steps iterations are done -- writing the configured
// observations after every iteration (time = steps x solver.step_size, which is 1).
#include <algorithm>
#include <chrono>
#include <iostream>
#include <memory>
#include <vector>

#include "app_common.hpp"
#include "config.hpp"
#include "hierarchy.hpp"
#include "solver.hpp"

int main(int argc, char** argv)
{
    const auto t0 = std::chrono::steady_clock::now();
    try {
        heom::command_line cl(argc, argv);
        std::cout << "Parsing HEOM configuration from file: " << cl.heom_config_filename << std::endl;
        heom::config cfg(cl.heom_config_filename, heom::program_task::thermal_state_search);

        heom::hierarchy_graph graph(cfg.baths_number(), cfg.baths_matsubaras(), cfg.system_ado_depth());
        heom::instance inst(cfg, heom::sites_to_states_mode::identity, graph);
        heom::thermal_state_search_solver solver(cfg, inst);
        const auto mask = heom::create_hierarchy_mask(cfg.filtering_strategy(), graph, cfg.baths_matsubaras(),
                                                      cfg.filtering_first_layer());
        solver.update_hierarchy_mask(mask.data());
        std::cout << "hierarchy: " << graph.nodes() << " ADOs, " << inst.states() << " states, width "
                  << graph.width() << ", depth " << graph.depth() << std::endl;

        std::vector<std::unique_ptr<heom::file_observer>> observers;
        bool norms = false;
        for (auto& o : cfg.observations()) {
            observers.emplace_back(new heom::file_observer(o.type, o.filename, inst));
            norms = norms || observers.back()->needs_norms();
        }
        const heom::int_t N = inst.states();
        std::vector<heom::complex_t> top((std::size_t)(N * N));
        std::vector<double> nv((std::size_t)graph.nodes());
        auto observe = [&](double time) {
            solver.hierarchy_top(top.data());
            if (norms)
                solver.hierarchy_norm(nv.data());
            for (auto& ob : observers)
                ob->observe(time, top.data(), norms ? &nv : nullptr);
        };
        observe(0.0);

        const heom::int_t per = cfg.program_observe_steps();
        const heom::int_t min_iterations = cfg.solver_steps() / per;
        const heom::int_t max_iterations = std::max(min_iterations, cfg.thermal_state_search_max_steps() / per);
        heom::int_t i = 0;
        bool converged = false;
        for (; i < max_iterations && !converged; ++i) {
            solver.step_forward(per);
            observe((double)((i + 1) * per) * cfg.solver_step_size());
            if (i >= min_iterations)
                converged = solver.test_convergence(cfg.thermal_state_search_delta());
            if (!cl.no_progress)
                std::cout << "\rCalculation thermal state search: " << (100 * (i + 1) / max_iterations) << " %"
                          << std::flush;
        }
        if (!cl.no_progress)
            std::cout << std::endl;
        std::cout << "thermal state search: " << i * per << " steps, "
                  << (converged ? "converged" : "not converged (max_steps reached)") << std::endl;
    } catch (const std::exception& e) {
        return heom::report_exception(e);
    }
    std::cout << "main(): application runtime: "
              << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() << " s" << std::endl;
    return 0;
}
