// app_linear_absorption.cpp -- linear absorption from a reference config file
// (counterpart of dm-heom/src/app_linear_absorption.cpp:39-188): for every dipole tensor
// component, an instance with a ground state (sites_to_states with_ground_state),
// sigma_0(0) = D+ |g><g|, propagation with classical RK4 on the GPU and the trace
// C(t) = Tr(D- sigma_0(t)) every program.observe_steps steps; the traces are averaged
// over the components and written as (time, trace_real, trace_imag).
//
// --rotating-frame <E0 in cm^-1> (an addition): propagate with E0 subtracted from the
// excited-state block of H and multiply C(t) by exp(-i E0 t / hbar).  The coherences
// between ground and excited states then rotate at |E - E0| instead of |E|, which keeps
// classical RK4 stable at step sizes where the optical frequency alone would exceed its
// stability bound (omega dt > 2.8); the result is the same C(t).  Without the option
// the propagation is the reference's.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "app_common.hpp"
#include "config.hpp"
#include "constants.hpp"
#include "hierarchy.hpp"
#include "solver.hpp"

int main(int argc, char** argv)
{
    const auto t0 = std::chrono::steady_clock::now();
    try {
        heom::command_line cl(argc, argv, {"rotating-frame"});
        std::cout << "Parsing HEOM configuration from file: " << cl.heom_config_filename << std::endl;
        heom::config cfg(cl.heom_config_filename, heom::program_task::linear_absorption);
        const std::string* rf = cl.option("rotating-frame");
        const double e0 = rf ? std::strtod(rf->c_str(), nullptr) * heom::constants::invcm_to_joule : 0.0;

        heom::hierarchy_graph graph(cfg.baths_number(), cfg.baths_matsubaras(), cfg.system_ado_depth());
        const std::string& filename = cfg.observations().front().filename;
        std::ofstream out(filename);
        if (out.fail())
            throw std::runtime_error("Error: could not open output file: " + filename);

        const heom::int_t iterations = cfg.solver_steps() / cfg.program_observe_steps();
        std::vector<heom::complex_t> acc((std::size_t)iterations + 1, heom::complex_t(0.0, 0.0));
        const std::size_t ncomp = cfg.dipole_tensor_prefactors().size();
        double device_s = 0.0;
        heom::time_stats steps_total, ode_total;
        std::int64_t device_bytes = 0;
        double hierarchy_bytes = 0.0, instance_bytes = 0.0;
        std::string stepper;
        for (std::size_t t = 0; t < ncomp; ++t) {
            heom::instance inst(cfg, heom::sites_to_states_mode::with_ground_state, graph);
            inst.set_hierarchy_top();
            if (e0 != 0.0) {
                std::vector<double> h = inst.hamiltonian();
                const heom::int_t n = inst.states();
                for (heom::int_t i = 1; i < n; ++i)
                    h[(std::size_t)(2 * (i * n + i))] -= e0;
                inst.set_hamiltonian(h.data());
            }
            heom::population_dynamics_solver solver(cfg, inst);
            const auto mask = heom::create_hierarchy_mask(cfg.filtering_strategy(), graph, cfg.baths_matsubaras(),
                                                          cfg.filtering_first_layer());
            solver.update_hierarchy_mask(mask.data());
            std::vector<heom::complex_t> dm, dp;
            heom::dipole_matrices(cfg, t, dm, dp);
            solver.hierarchy_mmult_left(dp.data());
            const auto tops = solver.propagate_observe(cfg.solver_steps(), cfg.program_observe_steps());
            const heom::int_t n = inst.states();
            for (heom::int_t r = 0; r <= iterations; ++r) {
                heom::complex_t tr = heom::trace_product(dm, &tops[(std::size_t)(r * n * n)], n);
                if (e0 != 0.0) {
                    const double time = (double)(r * cfg.program_observe_steps()) * cfg.solver_step_size();
                    tr *= std::exp(heom::complex_t(0.0, -e0 * time / heom::constants::h_bar));
                }
                acc[(std::size_t)r] += tr;
            }
            const heom_timing_info ti = solver.timing();
            device_s += ti.rk4_seconds;
            steps_total.add(solver.step_stats().sum, solver.step_stats().count);
            ode_total.add(ti.rk4_seconds, ti.rhs_launches);
            device_bytes = solver.device_bytes();
            stepper = solver.stepper();
            hierarchy_bytes = (double)inst.matrices() * inst.states() * inst.states() * 16.0;
            instance_bytes = (double)inst.allocated_byte();
            if (!cl.no_progress)
                std::cout << "Calculation linear absorption: tensor component " << (t + 1) << " of " << ncomp
                          << std::endl;
        }
        out << "time" << heom::kDelimiter << "trace_real" << heom::kDelimiter << "trace_imag" << '\n';
        for (heom::int_t r = 0; r <= iterations; ++r) {
            const double time = (double)(r * cfg.program_observe_steps()) * cfg.solver_step_size();
            const heom::complex_t v = acc[(std::size_t)r] / (double)ncomp;
            out << heom::fmt_real(time) << heom::kDelimiter << heom::fmt_real(v.real()) << heom::kDelimiter
                << heom::fmt_real(v.imag()) << '\n';
        }
        // the runtime and memory summaries of dm-heom/src/app_linear_absorption.cpp (as in
        // app_population_dynamics.cpp:108-127), summed over the tensor components
        std::cout << "-------------------- Solver Runtime Summary ----------------" << std::endl;
        heom::write_time_statistics_summary(steps_total, "solver step (" + stepper + ")", std::cout);
        if (ode_total.sum > 0.0)
            heom::write_time_statistics_summary(ode_total, "heom_ode kernel", std::cout);
        std::cout << "------------------------------------------------------------" << std::endl;
        std::cout << "-------------------- Memory Summary ------------------------" << std::endl;
        char line[64];
        std::snprintf(line, sizeof line, "%11.2f MiB", hierarchy_bytes / 1024.0 / 1024.0);
        std::cout << "local hierarchy buffer size:            " << line << std::endl;
        std::snprintf(line, sizeof line, "%11.2f MiB", instance_bytes / 1024.0 / 1024.0);
        std::cout << "local instance size:                    " << line << std::endl;
        std::snprintf(line, sizeof line, "%11.2f MiB", (double)device_bytes / 1024.0 / 1024.0);
        std::cout << "device memory of the handle (HBM):      " << line << std::endl;
        std::cout << "------------------------------------------------------------" << std::endl;
        std::cout << "solver: " << device_s << " s device time over " << ncomp << " propagations" << std::endl;
    } catch (const std::exception& e) {
        return heom::report_exception(e);
    }
    std::cout << "time in main(): "
              << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() << " s" << std::endl;
    return 0;
}
