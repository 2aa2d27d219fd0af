// solver.hpp -- population_dynamics_solver over the HIP library, and the observers.
//
// Mirrors the reference's solver interface for this path
// (dm-heom/include/heom/population_dynamics_solver.hpp:31-312 and
// dm-heom/include/heom/solver_base.hpp:36-385): step_forward(n), hierarchy_top,
// hierarchy, hierarchy_norm, hierarchy_mmult_left, update_hierarchy_mask,
// update_hamiltonian, reset.  Instead of an OpenCL ODE_T / STEPPER_T pair, one
// heom_handle (include/heom_hip.h) owns the device state; errors from the library are
// rethrown as std::runtime_error, as the reference's ocl::error_handler throws.
// The observers produce the reference's observation files
// (matrix_{diagonal,complete}_observer.hpp, matrix_trace_observer.hpp,
// hierarchy_norm_observer.hpp): one row per observation, a header naming the columns.
// The text format of numbers (the un-vendored noma/num real_format) is not part of the
// contract; it is "%.17g" here, tab separated.
#pragma once

#include <chrono>
#include <complex>
#include <cstdio>
#include <fstream>
#include <ostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "config.hpp"
#include "heom_hip.h"
#include "hierarchy.hpp"
#include "stepper.hpp"
#include "thermal_state.hpp"

namespace heom {

inline void hip_call(int rc)
{
    if (rc != HEOM_OK)
        throw std::runtime_error(std::string("heom: ") + heom_last_error());
}

// seconds summed over `count` events (bmt::statistics of the reference, reduced to what
// its summary prints)
struct time_stats {
    double sum = 0.0;
    std::int64_t count = 0;
    void add(double seconds, std::int64_t n)
    {
        sum += seconds;
        count += n;
    }
};

inline double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

class population_dynamics_solver {
public:
    // heom_create from the instance, initial value = the instance's top ADO
    // solver.stepper_type: rk_rk4 runs the fused RK4 step (heom_step_rk4); any other
    // tableau of stepper.hpp runs on the per-evaluation interface
    population_dynamics_solver(const config& cfg, instance& inst)
        : inst_(inst), step_size_(cfg.solver_step_size())
    {
        const heom_desc d = inst.desc();
        hip_call(heom_create(&d, &h_));
        stepper_name_ = cfg.solver_stepper_type();
        if (cfg.solver_stepper_type() != "rk_rk4") {
            const rk_tableau* t = find_tableau(cfg.solver_stepper_type());
            if (!t)
                throw std::runtime_error("stepper type '" + cfg.solver_stepper_type() + "' is not one of " +
                                         stepper_names());
            stepper_ = std::make_unique<explicit_rk_stepper>(h_, *t);
        }
        reset();
    }
    ~population_dynamics_solver()
    {
        if (h_)
            heom_destroy(h_);
    }
    population_dynamics_solver(const population_dynamics_solver&) = delete;
    population_dynamics_solver& operator=(const population_dynamics_solver&) = delete;

    // population_dynamics_solver::reset (population_dynamics_solver.hpp:161-175)
    void reset()
    {
        hip_call(heom_clear_state(h_));
        hip_call(heom_set_top(h_, inst_.top().data()));
        step_count_ = 0;
        time_ = 0.0;
    }

    // population_dynamics_solver::step_forward (population_dynamics_solver.hpp:233-265)
    void step_forward(int_t n)
    {
        const auto t0 = std::chrono::steady_clock::now();
        if (stepper_) {
            for (int_t k = 0; k < n; ++k)
                stepper_->step(time_ + (double)k * step_size_, step_size_);
        } else {
            hip_call(heom_step_rk4(h_, step_size_, n));
        }
        hip_call(heom_sync(h_));
        step_stats_.add(seconds_since(t0), n);
        step_count_ += n;
        time_ += (double)n * step_size_;
    }

    // n steps, the top ADO recorded every observe_steps on the device; returns the
    // rows (n / observe_steps + 1 matrices, the first = the current state)
    std::vector<complex_t> propagate_observe(int_t n, int_t observe_steps)
    {
        const int_t N = inst_.states();
        std::vector<complex_t> tops((std::size_t)((n / observe_steps + 1) * N * N));
        const auto t0 = std::chrono::steady_clock::now();
        if (stepper_) {
            hierarchy_top(tops.data());
            for (int_t k = 1; k <= n; ++k) {
                stepper_->step(time_ + (double)(k - 1) * step_size_, step_size_);
                if (k % observe_steps == 0)
                    hierarchy_top(tops.data() + (std::size_t)(k / observe_steps * N * N));
            }
        } else {
            hip_call(heom_propagate(h_, step_size_, n, observe_steps, reinterpret_cast<double*>(tops.data())));
        }
        step_stats_.add(seconds_since(t0), n);
        step_count_ += n;
        time_ += (double)n * step_size_;
        return tops;
    }

    void hierarchy_top(complex_t* buffer) { hip_call(heom_get_top(h_, reinterpret_cast<double*>(buffer))); }
    void hierarchy(complex_t* buffer) { hip_call(heom_get_state(h_, reinterpret_cast<double*>(buffer))); }
    void hierarchy_norm(real_t* buffer, real_t threshold = -1.0)
    {
        const auto t0 = std::chrono::steady_clock::now();
        hip_call(heom_hierarchy_norm(h_, threshold, buffer));
        norm_stats_.add(seconds_since(t0), 1);
    }
    void update_hierarchy_mask(const std::uint8_t* mask) { hip_call(heom_set_mask(h_, mask)); }
    void hierarchy_mmult_left(const complex_t* m)
    {
        hip_call(heom_mmult_left(h_, reinterpret_cast<const double*>(m)));
    }
    void update_hamiltonian(const real_t* h)
    {
        inst_.set_hamiltonian(h);
        hip_call(heom_set_hamiltonian(h_, h));
    }

    real_t time() const { return time_; }
    int_t step_count() const { return step_count_; }
    instance& heom_instance() { return inst_; }
    heom_handle* handle() { return h_; }
    heom_timing_info timing()
    {
        heom_timing_info t{};
        hip_call(heom_timing(h_, &t));
        return t;
    }
    const time_stats& step_stats() const { return step_stats_; }
    const time_stats& norm_stats() const { return norm_stats_; }
    const std::string& stepper() const { return stepper_name_; }
    std::int64_t device_bytes()
    {
        std::int64_t b = 0;
        hip_call(heom_device_bytes(h_, &b));
        return b;
    }

private:
    instance& inst_;
    heom_handle* h_ = nullptr;
    real_t step_size_;
    int_t step_count_ = 0;
    real_t time_ = 0.0;
    std::unique_ptr<explicit_rk_stepper> stepper_;  // null: the fused RK4 step
    std::string stepper_name_;
    time_stats step_stats_, norm_stats_;
};

inline void write_time_statistics_summary(const time_stats& st, const std::string& name, std::ostream& os,
                                          std::size_t name_width = 25)
{
    char line[160];
    std::snprintf(line, sizeof line, "%11.2f s (count: %6lld, average: %10.2f ms)", st.sum,
                  (long long)st.count, st.count ? st.sum / (double)st.count * 1e3 : 0.0);
    os << "time in " << name << ": " << std::string(name_width > name.size() ? name_width - name.size() : 0, ' ')
       << line << '\n';
}

// solver_base::write_runtime_summary (dm-heom/include/heom/solver_base.hpp:376-383,
// format of dm-heom/src/heom/common.cpp:34-45): host time per RK4 step, device time of
// the stage kernels (4 per fused RK4 step) and the norm kernel
inline void write_runtime_summary(population_dynamics_solver& s, std::ostream& os)
{
    write_time_statistics_summary(s.step_stats(), "solver step (" + s.stepper() + ")", os);
    const heom_timing_info t = s.timing();
    time_stats ode;
    ode.add(t.rk4_seconds, t.rhs_launches);
    if (t.rk4_seconds > 0.0)
        write_time_statistics_summary(ode, "heom_ode kernel", os);
    else
        os << "heom_ode kernel launches: " << t.rhs_launches << " (timed within the solver steps)\n";
    write_time_statistics_summary(s.norm_stats(), "hierarchy_norm kernel", os);
}

// the memory summary of dm-heom/src/app_population_dynamics.cpp:112-127: the hierarchy
// in the reference layout, the host instance arrays, the device memory of the handle
inline void write_memory_summary(const instance& inst, population_dynamics_solver& s, std::ostream& os)
{
    auto mib = [](double b) {
        char line[32];
        std::snprintf(line, sizeof line, "%11.2f MiB", b / 1024.0 / 1024.0);
        return std::string(line);
    };
    const double hierarchy = (double)inst.matrices() * inst.states() * inst.states() * 16.0;
    os << "local hierarchy buffer size:            " << mib(hierarchy) << '\n';
    os << "local instance size:                    " << mib((double)inst.allocated_byte()) << '\n';
    os << "device memory of the handle (HBM):      " << mib((double)s.device_bytes()) << '\n';
}

// thermal_state_search_solver (dm-heom/include/heom/thermal_state_search_solver.hpp:31-376)
// over heom_create_tss: the set-up of initialize() (:124-209) on the host, step_forward
// (:267-299) and test_convergence (:302-325) over the device state.
class thermal_state_search_solver {
public:
    thermal_state_search_solver(const config& cfg, instance& inst)
        : inst_(inst), setup_(thermal_state_setup(inst, cfg.baths_temperature()))
    {
        const heom_desc d = inst.desc();
        const double eps[2] = {setup_.epsilon.real(), setup_.epsilon.imag()};
        hip_call(heom_create_tss(&d, setup_.c.data(), setup_.eigenvalues.data(), eps, &h_));
        inst.set_hierarchy_top(setup_.top.data());
        const std::size_t nn = (std::size_t)(inst.states() * inst.states());
        last_top_.resize(nn);
        current_top_.resize(nn);
        reset();
    }
    ~thermal_state_search_solver()
    {
        if (h_)
            heom_destroy(h_);
    }
    thermal_state_search_solver(const thermal_state_search_solver&) = delete;
    thermal_state_search_solver& operator=(const thermal_state_search_solver&) = delete;

    void reset()
    {
        hip_call(heom_clear_state(h_));
        hip_call(heom_set_top(h_, inst_.top().data()));
        step_count_ = 0;
    }
    // n iterations; the top before them is kept for test_convergence
    void step_forward(int_t n)
    {
        hierarchy_top(last_top_.data());
        hip_call(heom_tss_step(h_, n));
        step_count_ += n;
    }
    // |sum |last top| - sum |current top|| < delta (:302-325)
    bool test_convergence(real_t delta)
    {
        hierarchy_top(current_top_.data());
        auto norm1 = [](const std::vector<complex_t>& m) {
            real_t v = 0.0;
            for (auto& x : m)
                v += std::abs(x);
            return v;
        };
        return std::abs(norm1(last_top_) - norm1(current_top_)) < delta;
    }

    void hierarchy_top(complex_t* buffer) { hip_call(heom_get_top(h_, reinterpret_cast<double*>(buffer))); }
    void hierarchy(complex_t* buffer) { hip_call(heom_get_state(h_, reinterpret_cast<double*>(buffer))); }
    void hierarchy_norm(real_t* buffer, real_t threshold = -1.0)
    {
        hip_call(heom_hierarchy_norm(h_, threshold, buffer));
    }
    void update_hierarchy_mask(const std::uint8_t* mask) { hip_call(heom_set_mask(h_, mask)); }
    int_t step_count() const { return step_count_; }
    const thermal_setup& setup() const { return setup_; }
    heom_handle* handle() { return h_; }

private:
    instance& inst_;
    thermal_setup setup_;
    heom_handle* h_ = nullptr;
    int_t step_count_ = 0;
    std::vector<complex_t> last_top_, current_top_;
};

// ---- observation output ----------------------------------------------------------

inline std::string fmt_real(double v)
{
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

constexpr char kDelimiter = '\t';

// one observer: a header and a row writer over the top ADO (or the norms)
class file_observer {
public:
    file_observer(const std::string& type, const std::string& filename, const instance& inst)
        : type_(type), inst_(inst), out_(filename)
    {
        if (out_.fail())
            throw std::runtime_error("Error: could not open output file: " + filename);
        if (type != "matrix_diagonal" && type != "matrix_complete" && type != "matrix_trace_id" &&
            type != "hierarchy_norm" && type != "hierarchy_complete")
            throw std::runtime_error("observation type '" + type + "' is not supported by this application");
        write_header();
    }

    bool needs_norms() const { return type_ == "hierarchy_norm"; }
    bool needs_hierarchy() const { return type_ == "hierarchy_complete"; }

    // write one row from the top ADO (and the norms for hierarchy_norm, the whole
    // hierarchy in uid order for hierarchy_complete)
    void observe(double time, const complex_t* top, const std::vector<double>* norms = nullptr,
                 const std::vector<complex_t>* hierarchy = nullptr)
    {
        const int_t N = inst_.states();
        out_ << fmt_real(time);
        if (type_ == "matrix_diagonal") {  // matrix_diagonal_observer.hpp:28-34
            for (int_t i = 0; i < N; ++i)
                out_ << kDelimiter << fmt_real(top[i * N + i].real());
        } else if (type_ == "matrix_complete") {  // matrix_complete_observer.hpp
            for (int_t e = 0; e < N * N; ++e)
                out_ << kDelimiter << fmt_real(top[e].real()) << kDelimiter << fmt_real(top[e].imag());
        } else if (type_ == "matrix_trace_id") {  // matrix_trace_observer.hpp (identity)
            complex_t tr = 0.0;
            for (int_t i = 0; i < N; ++i)
                tr += top[i * N + i];
            out_ << kDelimiter << fmt_real(tr.real()) << kDelimiter << fmt_real(tr.imag());
        } else if (type_ == "hierarchy_complete") {  // hierarchy_complete_observer.hpp:26-32
            if (!hierarchy)
                throw std::runtime_error("hierarchy_complete observation without the hierarchy");
            for (const complex_t& v : *hierarchy)
                out_ << kDelimiter << fmt_real(v.real()) << kDelimiter << fmt_real(v.imag());
        } else if (norms) {  // hierarchy_norm_observer.hpp
            for (double v : *norms)
                out_ << kDelimiter << fmt_real(v);
        }
        out_ << '\n';
    }

private:
    void write_header()
    {
        const int_t N = inst_.states();
        out_ << "time";
        if (type_ == "matrix_diagonal") {
            for (int_t i = 0; i < N; ++i)
                out_ << kDelimiter << "site_" << i;
        } else if (type_ == "matrix_complete") {
            for (int_t i = 0; i < N; ++i)
                for (int_t j = 0; j < N; ++j)
                    out_ << kDelimiter << "elem_" << i << "_" << j << "_real" << kDelimiter << "elem_" << i << "_"
                         << j << "_imag";
        } else if (type_ == "matrix_trace_id") {
            out_ << kDelimiter << "trace_real" << kDelimiter << "trace_imag";
        } else if (type_ == "hierarchy_complete") {  // hierarchy_complete_observer.hpp:58-70
            for (int_t h = 0; h < inst_.matrices(); ++h)
                for (int_t i = 0; i < N; ++i)
                    for (int_t j = 0; j < N; ++j)
                        out_ << kDelimiter << "mat_" << h << "_elem_" << i << "_" << j << "_real" << kDelimiter
                             << "mat_" << h << "_elem_" << i << "_" << j << "_imag";
        } else {
            const auto& g = inst_.graph();
            for (int_t u = 0; u < g.nodes(); ++u) {
                out_ << kDelimiter << "{";
                for (int_t m = 0; m < g.width(); ++m)
                    out_ << (m ? ", " : "") << g.tuples()[(std::size_t)(u * g.width() + m)];
                out_ << "}";
            }
        }
        out_ << '\n';
    }

    std::string type_;
    const instance& inst_;
    std::ofstream out_;
};

// D- (scaled by the tensor prefactor) and D+ = (D-)^T of the linear-absorption pseudo
// pathway +{(-, L), (+, L)} with a ground state (dm-heom/include/heom/dipole_matrix.hpp:
// 30-74, dm-heom/include/heom/dipole_pseudo_pathway.hpp:20-36,
// dm-heom/src/heom/dipole_matrix.cpp:182-208): row 0 of D- holds the site dipoles
// mu_i = direction_i[component] * strength_i.
inline void dipole_matrices(const config& cfg, std::size_t tensor_index, std::vector<complex_t>& d_minus,
                            std::vector<complex_t>& d_plus)
{
    const int_t sites = cfg.system_sites(), n = sites + 1;
    const auto& comp = cfg.dipole_tensor_components().at(tensor_index);
    d_minus.assign((std::size_t)(n * n), complex_t(0.0, 0.0));
    d_plus.assign((std::size_t)(n * n), complex_t(0.0, 0.0));
    for (int_t i = 0; i < sites; ++i) {
        const double m0 = cfg.dipole_directions()[(std::size_t)i][(std::size_t)comp[0]] *
                          cfg.dipole_strengths()[(std::size_t)i];
        const double m1 = cfg.dipole_directions()[(std::size_t)i][(std::size_t)comp[1]] *
                          cfg.dipole_strengths()[(std::size_t)i];
        d_minus[(std::size_t)(i + 1)] = m0;           // row 0, column i + 1
        d_plus[(std::size_t)((i + 1) * n)] = m1;      // transposed: row i + 1, column 0
    }
    const complex_t pre = cfg.dipole_tensor_prefactors().at(tensor_index);
    for (auto& v : d_minus)
        v *= pre;
}

// Tr(D top) as matrix_trace_observer::compute_trace (matrix_trace_observer.hpp:139-158)
inline complex_t trace_product(const std::vector<complex_t>& d, const complex_t* top, int_t n)
{
    complex_t tr(0.0, 0.0);
    for (int_t i = 0; i < n; ++i) {
        complex_t v(0.0, 0.0);
        for (int_t k = 0; k < n; ++k)
            v += d[(std::size_t)(i * n + k)] * top[k * n + i];
        tr += v;
    }
    return tr;
}

}  // namespace heom
