// config.hpp -- the HEOM config-file instance API (population dynamics and linear
// absorption), read without Boost.
//
// File format as the reference reads it with Boost.program_options
// (dm-heom/src/heom/config_base.cpp:39-63): "[section]" headers, "key=value" lines,
// '#' comments, keys addressed as "section.key", every registered key required and
// unregistered keys rejected.  Values use the noma/typa literal forms: numbers, bare
// words, braced lists "{a, b}" (nested for matrices) and pairs "(a, b)"; a complex
// matrix entry is either a real number or a pair (re, im).
//
// Keys and checks follow config::config / config::check (dm-heom/src/heom/config.cpp:
// 37-175), population_dynamics_config (population_dynamics_config.cpp:25-42) and
// dipole_config_entries (dipole_config_entries.cpp:21-47); post_process converts to SI
// units (config.cpp:182-190).
#pragma once

#include <complex>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace heom {

using real_t = double;
using complex_t = std::complex<double>;
using int_t = std::int64_t;

// heom::config_error (dm-heom/include/heom/config_error.hpp:17-23)
struct config_error : std::runtime_error {
    explicit config_error(const std::string& m) : std::runtime_error(m) {}
};

template <typename T>
struct matrix {
    std::size_t rows = 0, cols = 0;
    std::vector<T> data;  // row-major
    matrix() = default;
    matrix(std::size_t r, std::size_t c, T v = T()) : rows(r), cols(c), data(r * c, v) {}
    T& at(std::size_t i, std::size_t j) { return data.at(i * cols + j); }
    const T& at(std::size_t i, std::size_t j) const { return data.at(i * cols + j); }
};

using complex_matrix = matrix<complex_t>;
using real_matrix = matrix<real_t>;
using int_matrix = matrix<int_t>;

// ---- literal grammar ------------------------------------------------------------

struct literal {
    enum kind_t { number, word, list, pair } kind = word;
    double num = 0.0;
    bool is_integer = false;
    std::string text;                // word, or the number's source text
    std::vector<literal> items;      // list elements or the two pair entries
};

literal parse_literal(const std::string& s);

// ---- config ---------------------------------------------------------------------

enum class program_task { population_dynamics, linear_absorption, thermal_state_search };

struct observation_entry {
    std::string type;
    std::string filename;
};

// observation types the reference accepts (dm-heom/include/heom/observation_type.hpp:27-38)
bool is_known_observation_type(const std::string& t);

class config {
public:
    // Parse, check and convert to SI.  Throws config_error.  The task in the file must
    // match `task` (population_dynamics_config::check_final,
    // population_dynamics_config.cpp:321-326).
    config(const std::string& path, program_task task);

    program_task task() const { return task_; }

    // [program]
    const std::vector<observation_entry>& observations() const { return observations_; }
    int_t program_observe_steps() const { return observe_steps_; }
    // [filtering]
    const std::string& filtering_strategy() const { return filtering_strategy_; }
    int_t filtering_first_layer() const { return filtering_first_layer_; }
    // [solver]
    const std::string& solver_stepper_type() const { return stepper_type_; }
    real_t solver_step_size() const { return step_size_; }
    int_t solver_steps() const { return steps_; }
    bool solver_track_flows() const { return track_flows_; }
    const std::string& solver_flow_filename() const { return flow_filename_; }
    // [system] (H in J after post_process)
    int_t system_sites() const { return sites_; }
    int_t system_ado_depth() const { return ado_depth_; }
    const complex_matrix& system_hamiltonian() const { return hamiltonian_; }
    // [baths] (lambda in J, invnu in s, Omega in rad/s after post_process)
    int_t baths_number() const { return baths_number_; }
    int_t baths_max_per_site() const { return baths_max_per_site_; }
    const int_matrix& baths_coupling() const { return baths_coupling_; }
    const std::vector<real_t>& baths_lambda() const { return lambda_; }
    const std::vector<real_t>& baths_invnu() const { return invnu_; }
    const std::vector<real_t>& baths_uppercase_omega() const { return omega_; }
    int_t baths_matsubaras() const { return matsubaras_; }
    real_t baths_temperature() const { return temperature_; }
    // [population_dynamics]
    const complex_matrix& population_dynamics_rho_init() const { return rho_init_; }
    // [thermal_state_search] (thermal_state_search_config.cpp:19-23)
    real_t thermal_state_search_delta() const { return tss_delta_; }
    int_t thermal_state_search_max_steps() const { return tss_max_steps_; }
    // [dipole]
    const std::vector<std::vector<real_t>>& dipole_directions() const { return dipole_directions_; }
    const std::vector<real_t>& dipole_strengths() const { return dipole_strengths_; }
    const std::vector<std::vector<real_t>>& dipole_centers() const { return dipole_centers_; }
    const std::vector<complex_t>& dipole_tensor_prefactors() const { return tensor_prefactors_; }
    const std::vector<std::vector<int_t>>& dipole_tensor_components() const { return tensor_components_; }

    // overrides used by benches and tests (the file stays the source of truth otherwise)
    void set_ado_depth(int_t d) { ado_depth_ = d; }

private:
    void check() const;
    void post_process();

    program_task task_;
    std::vector<observation_entry> observations_;
    int_t observe_steps_ = 0;
    std::string filtering_strategy_;
    int_t filtering_first_layer_ = 0;
    std::string stepper_type_;
    real_t step_size_ = 0.0;
    int_t steps_ = 0;
    bool track_flows_ = false;
    std::string flow_filename_;
    int_t sites_ = 0, ado_depth_ = 0;
    complex_matrix hamiltonian_;
    int_t baths_number_ = 0, baths_max_per_site_ = 0;
    int_matrix baths_coupling_;
    std::vector<real_t> lambda_, invnu_, omega_;
    int_t matsubaras_ = 0;
    real_t temperature_ = 0.0;
    complex_matrix rho_init_;
    real_t tss_delta_ = 0.0;
    int_t tss_max_steps_ = 0;
    std::vector<std::vector<real_t>> dipole_directions_, dipole_centers_;
    std::vector<real_t> dipole_strengths_;
    std::vector<complex_t> tensor_prefactors_;
    std::vector<std::vector<int_t>> tensor_components_;
};

// read "key=value" pairs of an INI file (section-prefixed keys); throws config_error
std::map<std::string, std::string> read_ini(const std::string& path);

}  // namespace heom
