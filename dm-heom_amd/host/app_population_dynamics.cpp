// app_population_dynamics.cpp -- population dynamics from a reference config file
// (counterpart of dm-heom/src/app_population_dynamics.cpp:33-138): parse, build the
// hierarchy graph and instance (identity sites-to-states mapping, rho_init on the top
// ADO), apply the filtering mask, propagate with classical RK4 on the GPU and write the
// configured observations every program.observe_steps steps.
#include <chrono>
#include <cstdio>
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
        heom::config cfg(cl.heom_config_filename, heom::program_task::population_dynamics);

        heom::hierarchy_graph graph(cfg.baths_number(), cfg.baths_matsubaras(), cfg.system_ado_depth());
        heom::instance inst(cfg, heom::sites_to_states_mode::identity, graph);
        inst.set_hierarchy_top(cfg.population_dynamics_rho_init().data.data());
        heom::population_dynamics_solver solver(cfg, inst);
        const auto mask = heom::create_hierarchy_mask(cfg.filtering_strategy(), graph, cfg.baths_matsubaras(),
                                                      cfg.filtering_first_layer());
        solver.update_hierarchy_mask(mask.data());
        std::cout << "hierarchy: " << graph.nodes() << " ADOs, " << inst.states() << " states, width "
                  << graph.width() << ", depth " << graph.depth() << std::endl;

        std::vector<std::unique_ptr<heom::file_observer>> observers;
        bool norms = false, whole = false;
        for (auto& o : cfg.observations()) {
            observers.emplace_back(new heom::file_observer(o.type, o.filename, inst));
            norms = norms || observers.back()->needs_norms();
            whole = whole || observers.back()->needs_hierarchy();
        }

        const heom::int_t N = inst.states();
        const heom::int_t iterations = cfg.solver_steps() / cfg.program_observe_steps();
        std::vector<heom::complex_t> top((std::size_t)(N * N));
        if (!norms && !whole) {
            // every observation is a function of the top ADO: record it on the device
            const auto tops = solver.propagate_observe(cfg.solver_steps(), cfg.program_observe_steps());
            for (heom::int_t r = 0; r <= iterations; ++r) {
                const double time = (double)(r * cfg.program_observe_steps()) * cfg.solver_step_size();
                for (auto& ob : observers)
                    ob->observe(time, &tops[(std::size_t)(r * N * N)]);
            }
        } else {
            // observations that need the norms or the whole hierarchy read them back at
            // every observation
            std::vector<double> nv((std::size_t)graph.nodes());
            std::vector<heom::complex_t> hv(whole ? (std::size_t)(graph.nodes() * N * N) : 0);
            for (heom::int_t r = 0; r <= iterations; ++r) {
                if (r > 0)
                    solver.step_forward(cfg.program_observe_steps());
                solver.hierarchy_top(top.data());
                if (norms)
                    solver.hierarchy_norm(nv.data());
                if (whole)
                    solver.hierarchy(hv.data());
                const double time = (double)(r * cfg.program_observe_steps()) * cfg.solver_step_size();
                for (auto& ob : observers)
                    ob->observe(time, top.data(), &nv, &hv);
                if (!cl.no_progress)
                    std::cout << "\rCalculation population dynamics: " << (100 * r / iterations) << " %" << std::flush;
            }
            if (!cl.no_progress)
                std::cout << std::endl;
        }
        std::cout << "-------------------- Solver Runtime Summary ----------------" << std::endl;
        heom::write_runtime_summary(solver, std::cout);
        std::cout << "------------------------------------------------------------" << std::endl;
        std::cout << "-------------------- Memory Summary ------------------------" << std::endl;
        heom::write_memory_summary(inst, solver, std::cout);
        std::cout << "------------------------------------------------------------" << std::endl;
        const auto& st = solver.step_stats();
        std::cout << "solver: " << st.count << " steps (" << solver.stepper() << "), "
                  << (st.sum > 0 ? (double)graph.nodes() * (double)st.count / st.sum : 0.0) << " ADO-steps/s"
                  << std::endl;
    } catch (const std::exception& e) {
        return heom::report_exception(e);
    }
    std::cout << "main(): application runtime: "
              << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() << " s" << std::endl;
    return 0;
}
