// config.cpp -- see config.hpp.
#include "config.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <set>
#include <sstream>

#include "constants.hpp"
#include "stepper.hpp"

namespace heom {

namespace {

std::string trim(const std::string& s)
{
    std::size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a]))
        ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1]))
        --b;
    return s.substr(a, b - a);
}

class literal_parser {
public:
    explicit literal_parser(const std::string& s) : s_(s) {}

    literal parse_all()
    {
        literal v = parse_value();
        skip_ws();
        if (pos_ != s_.size())
            fail("trailing characters");
        return v;
    }

private:
    [[noreturn]] void fail(const std::string& why) const
    {
        throw config_error("cannot parse literal '" + s_ + "': " + why);
    }
    void skip_ws()
    {
        while (pos_ < s_.size() && std::isspace((unsigned char)s_[pos_]))
            ++pos_;
    }
    literal parse_value()
    {
        skip_ws();
        if (pos_ >= s_.size())
            fail("unexpected end");
        const char c = s_[pos_];
        if (c == '{')
            return parse_seq('{', '}', literal::list);
        if (c == '(') {
            literal p = parse_seq('(', ')', literal::pair);
            if (p.items.size() != 2)
                fail("a pair needs exactly two entries");
            return p;
        }
        return parse_atom();
    }
    literal parse_seq(char open, char close, literal::kind_t kind)
    {
        literal v;
        v.kind = kind;
        ++pos_;  // open
        (void)open;
        skip_ws();
        if (pos_ < s_.size() && s_[pos_] == close) {
            ++pos_;
            return v;
        }
        for (;;) {
            v.items.push_back(parse_value());
            skip_ws();
            if (pos_ >= s_.size())
                fail("missing closing bracket");
            if (s_[pos_] == ',') {
                ++pos_;
                continue;
            }
            if (s_[pos_] == close) {
                ++pos_;
                return v;
            }
            fail(std::string("unexpected '") + s_[pos_] + "'");
        }
    }
    literal parse_atom()
    {
        const std::size_t start = pos_;
        while (pos_ < s_.size() && s_[pos_] != ',' && s_[pos_] != '}' && s_[pos_] != ')' && s_[pos_] != '{' &&
               s_[pos_] != '(')
            ++pos_;
        literal v;
        v.text = trim(s_.substr(start, pos_ - start));
        if (v.text.empty())
            fail("empty entry");
        const char* b = v.text.c_str();
        char* e = nullptr;
        const double d = std::strtod(b, &e);
        if (e && *e == '\0' && e != b) {
            v.kind = literal::number;
            v.num = d;
            v.is_integer = v.text.find_first_of(".eEnN") == std::string::npos;
        } else {
            v.kind = literal::word;
        }
        return v;
    }

    const std::string& s_;
    std::size_t pos_ = 0;
};

const literal& require_kind(const literal& l, literal::kind_t k, const std::string& name)
{
    if (l.kind != k)
        throw config_error("faulty value for " + name);
    return l;
}

double to_real(const literal& l, const std::string& name)
{
    if (l.kind != literal::number)
        throw config_error("faulty value for " + name + ": '" + l.text + "' is not a number");
    return l.num;
}

int_t to_int(const literal& l, const std::string& name)
{
    if (l.kind != literal::number || !l.is_integer)
        throw config_error("faulty value for " + name + ": '" + l.text + "' is not an integer");
    return static_cast<int_t>(std::llround(l.num));
}

complex_t to_complex(const literal& l, const std::string& name)
{
    if (l.kind == literal::pair)
        return complex_t(to_real(l.items[0], name), to_real(l.items[1], name));
    return complex_t(to_real(l, name), 0.0);
}

std::vector<real_t> to_real_vector(const literal& l, const std::string& name)
{
    require_kind(l, literal::list, name);
    std::vector<real_t> v;
    for (auto& x : l.items)
        v.push_back(to_real(x, name));
    return v;
}

template <typename T, typename F>
matrix<T> to_matrix(const literal& l, const std::string& name, F conv)
{
    require_kind(l, literal::list, name);
    matrix<T> m;
    m.rows = l.items.size();
    m.cols = m.rows ? require_kind(l.items[0], literal::list, name).items.size() : 0;
    for (auto& row : l.items) {
        require_kind(row, literal::list, name);
        if (row.items.size() != m.cols)
            throw config_error("faulty value for " + name + ": rows of different length");
        for (auto& x : row.items)
            m.data.push_back(conv(x, name));
    }
    return m;
}

bool to_bool(const std::string& s, const std::string& name)
{
    std::string t;
    for (char c : trim(s))
        t += (char)std::tolower((unsigned char)c);
    if (t == "true" || t == "1" || t == "yes" || t == "on")
        return true;
    if (t == "false" || t == "0" || t == "no" || t == "off")
        return false;
    throw config_error("faulty value for " + name + ": '" + s + "'");
}

const char* kCommonKeys[] = {
    "program.task", "program.observations", "program.observe_steps",
    "filtering.strategy", "filtering.first_layer",
    "solver.stepper_type", "solver.step_size", "solver.steps", "solver.track_flows", "solver.flow_filename",
    "system.sites", "system.ado_depth", "system.hamiltonian",
    "baths.number", "baths.max_per_site", "baths.coupling", "baths.lambda", "baths.invnu", "baths.Omega",
    "baths.matsubaras", "baths.temperature",
};
const char* kPopulationKeys[] = {"population_dynamics.rho_init"};
const char* kThermalKeys[] = {"thermal_state_search.delta", "thermal_state_search.max_steps"};
const char* kDipoleKeys[] = {"dipole.directions", "dipole.strengths", "dipole.centers",
                             "dipole.tensor_prefactors", "dipole.tensor_components"};

}  // namespace

literal parse_literal(const std::string& s) { return literal_parser(s).parse_all(); }

bool is_known_observation_type(const std::string& t)
{
    static const std::set<std::string> names = {
        "matrix_diagonal", "matrix_complete", "matrix_complete_rate_kernel", "matrix_trace_id",
        "matrix_trace_linear_absorption", "matrix_trace_transient_absorption",
        "matrix_trace_static_fluorescence", "matrix_trace_two_dimensional_spectra", "hierarchy_norm",
        "hierarchy_complete"};
    return names.count(t) > 0;
}

std::map<std::string, std::string> read_ini(const std::string& path)
{
    std::ifstream f(path);
    if (f.fail())
        throw config_error("could not open config file: " + path);
    std::map<std::string, std::string> kv;
    std::string line, section;
    while (std::getline(f, line)) {
        const auto hash = line.find('#');
        if (hash != std::string::npos)
            line = line.substr(0, hash);
        line = trim(line);
        if (line.empty())
            continue;
        if (line.front() == '[' && line.back() == ']') {
            section = trim(line.substr(1, line.size() - 2));
            continue;
        }
        const auto eq = line.find('=');
        if (eq == std::string::npos)
            throw config_error("invalid line in config file: " + line);
        const std::string key = (section.empty() ? "" : section + ".") + trim(line.substr(0, eq));
        if (kv.count(key))
            throw config_error("multiple occurrences of option '" + key + "'");
        kv[key] = trim(line.substr(eq + 1));
    }
    return kv;
}

config::config(const std::string& path, program_task task) : task_(task)
{
    const auto kv = read_ini(path);
    auto it = kv.find("program.task");
    if (it == kv.end())
        throw config_error("the option 'program.task' is required but missing");
    const std::string want = task == program_task::population_dynamics    ? "population_dynamics"
                             : task == program_task::linear_absorption    ? "linear_absorption"
                                                                          : "thermal_state_search";
    if (it->second != want)
        throw config_error("wrong program_task for this application");

    std::set<std::string> allowed(std::begin(kCommonKeys), std::end(kCommonKeys));
    if (task == program_task::population_dynamics)
        allowed.insert(std::begin(kPopulationKeys), std::end(kPopulationKeys));
    else if (task == program_task::linear_absorption)
        allowed.insert(std::begin(kDipoleKeys), std::end(kDipoleKeys));
    else
        allowed.insert(std::begin(kThermalKeys), std::end(kThermalKeys));
    for (auto& [k, v] : kv)
        if (!allowed.count(k))
            throw config_error("unrecognised option '" + k + "'");
    for (auto& k : allowed)
        if (!kv.count(k))
            throw config_error("the option '" + k + "' is required but missing");

    auto lit = [&](const char* key) { return parse_literal(kv.at(key)); };

    // [program]
    {
        const literal obs = lit("program.observations");
        require_kind(obs, literal::list, "program.observations");
        for (auto& p : obs.items) {
            require_kind(p, literal::pair, "program.observations");
            observation_entry e{p.items[0].text, p.items[1].text};
            if (!is_known_observation_type(e.type))
                throw config_error("'" + e.type + "' is not a valid observation_type");
            observations_.push_back(e);
        }
    }
    observe_steps_ = to_int(lit("program.observe_steps"), "program.observe_steps");
    filtering_strategy_ = kv.at("filtering.strategy");
    if (filtering_strategy_ != "none" && filtering_strategy_ != "single_excitation")
        throw config_error("'" + filtering_strategy_ + "' is not a valid filtering_strategy");
    filtering_first_layer_ = to_int(lit("filtering.first_layer"), "filtering.first_layer");
    stepper_type_ = kv.at("solver.stepper_type");
    if (!find_tableau(stepper_type_))
        throw config_error("solver.stepper_type '" + stepper_type_ + "' is not supported (one of " +
                           stepper_names() + ")");
    step_size_ = to_real(lit("solver.step_size"), "solver.step_size");
    steps_ = to_int(lit("solver.steps"), "solver.steps");
    track_flows_ = to_bool(kv.at("solver.track_flows"), "solver.track_flows");
    flow_filename_ = kv.at("solver.flow_filename");
    sites_ = to_int(lit("system.sites"), "system.sites");
    ado_depth_ = to_int(lit("system.ado_depth"), "system.ado_depth");
    hamiltonian_ = to_matrix<complex_t>(lit("system.hamiltonian"), "system.hamiltonian", to_complex);
    baths_number_ = to_int(lit("baths.number"), "baths.number");
    baths_max_per_site_ = to_int(lit("baths.max_per_site"), "baths.max_per_site");
    baths_coupling_ = to_matrix<int_t>(lit("baths.coupling"), "baths.coupling", to_int);
    lambda_ = to_real_vector(lit("baths.lambda"), "baths.lambda");
    invnu_ = to_real_vector(lit("baths.invnu"), "baths.invnu");
    omega_ = to_real_vector(lit("baths.Omega"), "baths.Omega");
    matsubaras_ = to_int(lit("baths.matsubaras"), "baths.matsubaras");
    temperature_ = to_real(lit("baths.temperature"), "baths.temperature");

    if (task == program_task::population_dynamics) {
        rho_init_ = to_matrix<complex_t>(lit("population_dynamics.rho_init"), "population_dynamics.rho_init",
                                         to_complex);
    } else if (task == program_task::thermal_state_search) {
        tss_delta_ = to_real(lit("thermal_state_search.delta"), "thermal_state_search.delta");
        tss_max_steps_ = to_int(lit("thermal_state_search.max_steps"), "thermal_state_search.max_steps");
    } else {
        auto vecs = [&](const char* key) {
            const literal l = lit(key);
            require_kind(l, literal::list, key);
            std::vector<std::vector<real_t>> out;
            for (auto& x : l.items)
                out.push_back(to_real_vector(x, key));
            return out;
        };
        dipole_directions_ = vecs("dipole.directions");
        dipole_centers_ = vecs("dipole.centers");
        dipole_strengths_ = to_real_vector(lit("dipole.strengths"), "dipole.strengths");
        {
            const literal l = lit("dipole.tensor_prefactors");
            require_kind(l, literal::list, "dipole.tensor_prefactors");
            for (auto& x : l.items)
                tensor_prefactors_.push_back(to_complex(x, "dipole.tensor_prefactors"));
        }
        {
            const literal l = lit("dipole.tensor_components");
            require_kind(l, literal::list, "dipole.tensor_components");
            for (auto& x : l.items) {
                require_kind(x, literal::list, "dipole.tensor_components");
                std::vector<int_t> c;
                for (auto& y : x.items)
                    c.push_back(to_int(y, "dipole.tensor_components"));
                tensor_components_.push_back(c);
            }
        }
    }
    check();
    post_process();
}

// config::check (dm-heom/src/heom/config.cpp:73-175) and the task checks
void config::check() const
{
    if (observations_.empty())
        throw config_error("observations must at least contain one entry.");
    if (!(observe_steps_ > 0))
        throw config_error("program.observe_steps must be > 0");
    if (filtering_strategy_ == "none" && filtering_first_layer_ != -1)
        throw config_error("filtering.first_layer must be -1, because it has no effect for filtering_strategy = none");
    if (filtering_strategy_ == "single_excitation" && !(filtering_first_layer_ >= 0))
        throw config_error("filtering.first_layer must be greater or equal to zero for filtering.strategy = single excitation filter");
    if (ado_depth_ < filtering_first_layer_)
        throw config_error("filtering.first_layer cannot be larger than system_ado_depth");
    if (!(step_size_ > 0.0))
        throw config_error("solver.step_size must be > 0");
    if (!(steps_ > 0))
        throw config_error("solver.steps must be > 0");
    if (steps_ % observe_steps_ != 0)
        throw config_error("there's a remainder when dividing solver.steps by program.observe_steps");
    if (!(sites_ > 0))
        throw config_error("system.sites must be > 0");
    if (!(ado_depth_ > 0))
        throw config_error("system.ado_depth must be > 0");
    if (!((int_t)hamiltonian_.rows == sites_ && (int_t)hamiltonian_.cols == sites_))
        throw config_error("system.hamiltonien must be a system.sites by system.sites matrix");
    for (int_t i = 0; i < sites_; ++i)
        for (int_t j = 0; j < sites_; ++j)
            if (hamiltonian_.at(i, j) != std::conj(hamiltonian_.at(j, i)))
                throw config_error("system.hamiltonien must be hermitian");
    if (!(baths_number_ > 0))
        throw config_error("baths.number must be > 0");
    if (!(baths_max_per_site_ > 0))
        throw config_error("baths.max_per_site must be > 0");
    if (!(baths_max_per_site_ <= baths_number_))
        throw config_error("baths.max_per_site must be <= baths.number");
    if (!((int_t)baths_coupling_.rows == sites_ && (int_t)baths_coupling_.cols == baths_max_per_site_))
        throw config_error("baths.coupling must be a system.sites by baths_max_per_site_ matrix");
    for (auto c : baths_coupling_.data)
        if (!(c < baths_number_))
            throw config_error("baths.coupling values must between 0 and baths.number, or negative for non-coupling");
    if ((int_t)lambda_.size() != baths_number_)
        throw config_error("baths.lambda must have a length of baths.number");
    if ((int_t)invnu_.size() != baths_number_)
        throw config_error("baths.invnu must have a length of baths.number");
    for (auto v : invnu_)
        if (!(v > 0))
            throw config_error("baths.invnu must have values > 0");
    if ((int_t)omega_.size() != baths_number_)
        throw config_error("baths.Omega must have a length of baths.number");
    if (matsubaras_ < 1 || matsubaras_ > (int_t)constants::max_pade)
        throw config_error("baths.matsubaras must be between 1 and 10 (size of the Pade tables)");
    if (task_ == program_task::population_dynamics) {
        if (!((int_t)rho_init_.rows == sites_ && (int_t)rho_init_.cols == sites_))
            throw config_error("population_dynamics.rho_init must be a system.sites by system.sites matrix");
    } else if (task_ == program_task::thermal_state_search) {
        // thermal_state_search_config::check_final (thermal_state_search_config.cpp:31-39)
        if (step_size_ != 1.0)
            throw config_error("solver.step_size_ must equal 1.0");
    } else {
        if (dipole_directions_.size() != dipole_strengths_.size() || dipole_directions_.size() != dipole_centers_.size())
            throw config_error("dipole.directions, dipole.strengths and dipole.centers must have the same size");
        if (tensor_prefactors_.size() != tensor_components_.size())
            throw config_error("dipole.tensor_prefactors and dipole.tensor_components must have the same size");
        if ((int_t)dipole_directions_.size() < sites_)
            throw config_error("dipole.directions needs one entry per site");
        for (auto& d : dipole_directions_)
            if (d.size() != 3)
                throw config_error("dipole.directions entries must have 3 components");
        for (auto& c : tensor_components_)
            if (c.size() != 2 || c[0] < 0 || c[0] > 2 || c[1] < 0 || c[1] > 2)
                throw config_error("dipole.tensor_components entries must be pairs of 0, 1 or 2");
    }
}

// config::post_process (dm-heom/src/heom/config.cpp:182-190)
void config::post_process()
{
    for (auto& h : hamiltonian_.data)
        h *= constants::invcm_to_joule;
    for (auto& l : lambda_)
        l *= constants::invcm_to_joule;
    const double omega_scale = constants::invcm_to_joule / constants::h_bar;
    for (auto& o : omega_)
        o *= omega_scale;
    for (auto& v : invnu_)
        v *= 1.0E-15;
}

}  // namespace heom
