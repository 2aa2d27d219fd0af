// hierarchy.cpp -- see hierarchy.hpp.
#include "hierarchy.hpp"

#include <algorithm>
#include <cstring>
#include <fstream>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <unordered_map>

#include "constants.hpp"

namespace heom {

int_t binomial(int_t n, int_t k)
{
    if (k < 0 || k > n)
        return 0;
    k = std::min(k, n - k);
    __int128 r = 1;
    for (int_t t = 1; t <= k; ++t) {
        r = r * (n - k + t) / t;  // exact: r * (n-k+t) is divisible by t after t steps
        if (r > (__int128)INT64_MAX)
            throw std::overflow_error("binomial coefficient exceeds int64");
    }
    return (int_t)r;
}

namespace {

void partitions_rec(int remaining, int max_part, int slots, std::vector<int>& prefix,
                    std::vector<std::vector<int>>& out)
{
    if (remaining == 0) {
        std::vector<int> row(prefix);
        row.resize(slots, 0);
        out.push_back(std::move(row));
        return;
    }
    if ((int)prefix.size() == slots)
        return;  // no slot left for another part
    for (int p = 1; p <= std::min(remaining, max_part); ++p) {
        prefix.push_back(p);
        partitions_rec(remaining - p, p, slots, prefix, out);
        prefix.pop_back();
    }
}

struct tuple_key_hash {
    std::size_t operator()(const std::string& s) const { return std::hash<std::string>()(s); }
};

std::string key_of(const std::int32_t* t, int_t w)
{
    std::string k((std::size_t)w, '\0');
    for (int_t m = 0; m < w; ++m)
        k[(std::size_t)m] = (char)(unsigned char)t[m];
    return k;
}

}  // namespace

std::vector<std::vector<int>> integer_partitions(int n, int s)
{
    std::vector<std::vector<int>> out;
    std::vector<int> prefix;
    if (n > 0)
        partitions_rec(n, n, s, prefix, out);
    return out;
}

hierarchy_graph::hierarchy_graph(int_t width, int_t depth) : width_(width), depth_(depth)
{
    if (width < 1 || depth < 1)
        throw std::invalid_argument("hierarchy_graph: width and depth must be > 0");
    if (depth > 255)
        throw std::invalid_argument("hierarchy_graph: depth must be <= 255");
    nodes_ = binomial(width + depth, width);                   // hierarchy_graph.cpp:34
    first_non_plus_ = binomial(width + depth - 1, width);      // hierarchy_graph.cpp:40
    if (nodes_ > INT32_MAX)
        throw std::overflow_error("hierarchy_graph: more than 2^31 - 1 nodes");
    tuples_.reserve((std::size_t)(nodes_ * width));
    nodes_per_depth_.assign((std::size_t)depth + 1, 0);

    std::vector<std::int32_t> cur((std::size_t)width);
    for (int_t d = 0; d <= depth; ++d) {
        if (d == 0) {
            tuples_.insert(tuples_.end(), (std::size_t)width, 0);
            nodes_per_depth_[0] = 1;
            continue;
        }
        for (const auto& part : integer_partitions((int)d, (int)width)) {
            std::copy(part.begin(), part.end(), cur.begin());
            std::sort(cur.begin(), cur.end());
            do {
                tuples_.insert(tuples_.end(), cur.begin(), cur.end());
                ++nodes_per_depth_[(std::size_t)d];
            } while (std::next_permutation(cur.begin(), cur.end()));
        }
    }
    if ((int_t)(tuples_.size() / (std::size_t)width) != nodes_)
        throw std::logic_error("hierarchy_graph: node count mismatch");

    std::unordered_map<std::string, uid_t, tuple_key_hash> uid_of;
    uid_of.reserve((std::size_t)nodes_ * 2);
    for (int_t u = 0; u < nodes_; ++u)
        uid_of.emplace(key_of(&tuples_[(std::size_t)(u * width)], width), (uid_t)u);

    plus_.assign((std::size_t)(nodes_ * width), invalid_uid);
    minus_.assign((std::size_t)(nodes_ * width), invalid_uid);
    std::vector<std::int32_t> probe((std::size_t)width);
    for (int_t u = 0; u < nodes_; ++u) {
        const std::int32_t* t = &tuples_[(std::size_t)(u * width)];
        const int_t dep = std::accumulate(t, t + width, (int_t)0);
        std::copy(t, t + width, probe.begin());
        for (int_t m = 0; m < width; ++m) {
            if (dep < depth) {
                ++probe[(std::size_t)m];
                plus_[(std::size_t)(u * width + m)] = uid_of.at(key_of(probe.data(), width));
                --probe[(std::size_t)m];
            }
            if (t[m] > 0) {
                --probe[(std::size_t)m];
                minus_[(std::size_t)(u * width + m)] = uid_of.at(key_of(probe.data(), width));
                ++probe[(std::size_t)m];
            }
        }
    }
}

int_t sites_to_states(int_t sites, sites_to_states_mode mode)
{
    return mode == sites_to_states_mode::identity ? sites : sites + 1;
}

complex_matrix state_hamiltonian(const complex_matrix& h, sites_to_states_mode mode)
{
    if (mode == sites_to_states_mode::identity)
        return h;
    // ground state in row / column 0, zero energy (sites_to_states.cpp:44-56)
    complex_matrix r(h.rows + 1, h.cols + 1, complex_t(0.0, 0.0));
    for (std::size_t i = 0; i < h.rows; ++i)
        for (std::size_t j = 0; j < h.cols; ++j)
            r.at(i + 1, j + 1) = h.at(i, j);
    return r;
}

int_matrix state_baths_coupling(const int_matrix& coupling, sites_to_states_mode mode)
{
    if (mode == sites_to_states_mode::identity)
        return coupling;
    // a prepended ground-state row without baths (sites_to_states.cpp:131-143)
    int_matrix r(coupling.rows + 1, coupling.cols, -1);
    for (std::size_t i = 0; i < coupling.rows; ++i)
        for (std::size_t j = 0; j < coupling.cols; ++j)
            r.at(i + 1, j) = coupling.at(i, j);
    return r;
}

std::vector<std::uint8_t> create_hierarchy_mask(const std::string& strategy, const hierarchy_graph& graph,
                                                int_t matsubaras, int_t first_layer)
{
    const int_t M = graph.nodes(), W = graph.width();
    std::vector<std::uint8_t> mask((std::size_t)M, mask_true);
    if (strategy == "none")
        return mask;
    if (strategy != "single_excitation")
        throw std::runtime_error("create_hierarchy_mask(): error: unknown filtering strategy");
    // keep an ADO of depth >= first_layer only if all its excitations sit in the modes
    // of one bath (hierarchy_mask.cpp:45-110)
    const int_t baths = W / matsubaras;
    for (int_t u = 0; u < M; ++u) {
        const std::int32_t* t = &graph.tuples()[(std::size_t)(u * W)];
        const int_t sum = std::accumulate(t, t + W, (int_t)0);
        if (sum < first_layer)
            continue;
        bool found = false;
        for (int_t b = 0; b < baths && !found; ++b)
            found = std::accumulate(t + b * matsubaras, t + (b + 1) * matsubaras, (int_t)0) == sum;
        mask[(std::size_t)u] = found ? mask_true : mask_false;
    }
    return mask;
}

instance::instance(const config& cfg, sites_to_states_mode mode, const hierarchy_graph& graph)
    : mode_(mode), graph_(graph)
{
    states_ = sites_to_states(cfg.system_sites(), mode);
    matsubaras_ = cfg.baths_matsubaras();
    baths_per_state_ = cfg.baths_max_per_site();
    baths_ = cfg.baths_number();
    width_ = graph.width();
    depth_ = graph.depth();
    matrices_ = graph.nodes();
    first_non_plus_ = graph.first_non_plus_uid();
    if (width_ != baths_ * matsubaras_)
        throw std::invalid_argument("instance: graph width must equal baths.number * baths.matsubaras");

    const complex_matrix H = state_hamiltonian(cfg.system_hamiltonian(), mode);
    hamiltonian_.resize((std::size_t)(states_ * states_ * 2));
    for (std::size_t e = 0; e < H.data.size(); ++e) {
        hamiltonian_[2 * e] = H.data[e].real();
        hamiltonian_[2 * e + 1] = H.data[e].imag();
    }
    const int_matrix C = state_baths_coupling(cfg.baths_coupling(), mode);
    coupling_.resize(C.data.size());
    for (std::size_t t = 0; t < C.data.size(); ++t)
        coupling_[t] = (std::int32_t)C.data[t];
    top_.assign((std::size_t)(states_ * states_ * 2), 0.0);
    lambda_ = cfg.baths_lambda();
    invnu_ = cfg.baths_invnu();
    omega_ = cfg.baths_uppercase_omega();
    temperature_ = cfg.baths_temperature();
    precompute();
}

// instance::precompute_baths_input (dm-heom/src/heom/instance.cpp:155-262), SI units:
//   nu_b = 1/invnu_b + i Omega_b, beta = 1/(k_B T);
//   Phi_b = nu_b lambda_b / hbar;
//   c_b0 = 2 lambda_b/(beta hbar) (1 - sum_j 2 eta_j nu_b^2 / (xi_j^2/(beta hbar)^2 - nu_b^2));
//   c_bk = 4 lambda_b nu_b/(beta hbar)^2 eta_k xi_k / (xi_k^2/(beta hbar)^2 - nu_b^2), k >= 1,
//     with nu_bk = xi_k/(beta hbar);
//   Delta_b = 2 lambda_b/(beta hbar^2 nu_b) - sum_k c_bk/(hbar nu_bk);
//   S_n = -sum_{b,k} n_bk nu_bk (long double).
void instance::precompute()
{
    const double hbar = constants::h_bar;
    const double beta = 1.0 / (constants::k_b * temperature_);
    const std::size_t row = (std::size_t)(matsubaras_ - 1) * constants::max_pade;
    const int_t B = baths_, K = matsubaras_;
    pf_same_.assign((std::size_t)(2 * B), 0.0);
    pf_circ_.assign((std::size_t)(2 * B), 0.0);
    cbk_.assign((std::size_t)(2 * B * K), 0.0);
    cbk_sqrt_.assign((std::size_t)(B * K), 0.0);

    for (int_t b = 0; b < B; ++b) {
        const complex_t lambda_b{lambda_[(std::size_t)b]};
        const complex_t nu_b{1.0 / invnu_[(std::size_t)b], omega_[(std::size_t)b]};
        const complex_t circ = nu_b * lambda_b / hbar;
        complex_t delta = 2.0 * lambda_b / (beta * hbar * hbar * nu_b);
        for (int_t k = 0; k < K; ++k) {
            complex_t c, nubk;
            if (k == 0) {
                complex_t corr = 0.0;
                for (int_t j = 0; j < K; ++j) {
                    const double eta = (double)constants::eta_pade[row + (std::size_t)j];
                    const double xi = (double)constants::xi_pade[row + (std::size_t)j];
                    corr += (2.0 * eta * nu_b * nu_b) / (xi * xi / (beta * beta * hbar * hbar) - nu_b * nu_b);
                }
                c = 2.0 * lambda_b / (beta * hbar) * (1.0 - corr);
                nubk = nu_b;
            } else {
                const double eta = (double)constants::eta_pade[row + (std::size_t)k];
                const double xi = (double)constants::xi_pade[row + (std::size_t)k];
                c = 4.0 * lambda_b * nu_b / (beta * beta * hbar * hbar) * (eta * xi) /
                    (xi * xi / (beta * beta * hbar * hbar) - nu_b * nu_b);
                nubk = xi / (beta * hbar);
            }
            const std::size_t m = (std::size_t)(b * K + k);
            cbk_[2 * m] = c.real();
            cbk_[2 * m + 1] = c.imag();
            cbk_sqrt_[m] = std::sqrt(std::abs(c));
            delta -= c / (hbar * nubk);
        }
        pf_same_[2 * (std::size_t)b] = delta.real();
        pf_same_[2 * (std::size_t)b + 1] = delta.imag();
        pf_circ_[2 * (std::size_t)b] = circ.real();
        pf_circ_[2 * (std::size_t)b + 1] = circ.imag();
    }

    const int_t M = matrices_, W = width_;
    pf_same_sum_.assign((std::size_t)(2 * M), 0.0);
    std::vector<long double> nr((std::size_t)W), ni((std::size_t)W);
    for (int_t b = 0; b < B; ++b) {
        const complex_t nu_b{1.0 / invnu_[(std::size_t)b], omega_[(std::size_t)b]};
        for (int_t k = 0; k < K; ++k) {
            const std::size_t m = (std::size_t)(b * K + k);
            if (k == 0) {
                nr[m] = (long double)nu_b.real();
                ni[m] = (long double)nu_b.imag();
            } else {
                nr[m] = constants::xi_pade[row + (std::size_t)k] / (long double)(beta * hbar);
                ni[m] = 0.0L;
            }
        }
    }
    const auto& T = graph_.tuples();
    for (int_t u = 0; u < M; ++u) {
        long double zr = 0.0L, zi = 0.0L;
        for (int_t m = 0; m < W; ++m) {
            const long double n = (long double)T[(std::size_t)(u * W + m)];
            zr -= nr[(std::size_t)m] * n;
            zi -= ni[(std::size_t)m] * n;
        }
        pf_same_sum_[2 * (std::size_t)u] = (double)zr;
        pf_same_sum_[2 * (std::size_t)u + 1] = (double)zi;
    }
}

void instance::set_baths(const double* lambda, const double* invnu, const double* omega)
{
    for (int_t b = 0; b < baths_; ++b) {
        if (!(invnu[b] > 0.0))
            throw std::invalid_argument("set_baths: invnu must be > 0");
        lambda_[(std::size_t)b] = lambda[b];
        invnu_[(std::size_t)b] = invnu[b];
        omega_[(std::size_t)b] = omega[b];
    }
    precompute();
}

// Derivatives of the prefactors of precompute() (no counterpart in the reference).  All
// of Phi_b, c_bk and Delta_b are linear in lambda_b and holomorphic in nu_b, so with
// nu = nu_r + i nu_i, d/d nu_r = d/d nu and d/d nu_i = i d/d nu; for a real parameter
// theta, dJ/d theta = sum over the entries kappa of Re(conj(G_kappa) d kappa/d theta).
// With a_j = xi_j^2 / (beta hbar)^2:
//   d c_b0/d nu  = -2 lambda/(beta hbar) sum_j 4 eta_j nu a_j / (a_j - nu^2)^2
//   d c_bk/d nu  = 4 lambda/(beta hbar)^2 eta_k xi_k (a_k + nu^2) / (a_k - nu^2)^2, k >= 1
//   d Phi_b/d nu = lambda / hbar
//   d Delta_b/d nu = -2 lambda/(beta hbar^2 nu^2) - c_b0'/(hbar nu) + c_b0/(hbar nu^2)
//                    - sum_{k >= 1} c_bk'/(hbar nu_bk)
//   d S_n/d nu_b = -n_b0 (the Pade poles nu_bk, k >= 1, do not depend on nu_b)
void instance::bath_gradient(const double* g_pss, const double* g_same, const double* g_circ, const double* g_cbk,
                             double* d_lambda, double* d_nu_re, double* d_nu_im) const
{
    const double hbar = constants::h_bar;
    const double beta = 1.0 / (constants::k_b * temperature_);
    const std::size_t row = (std::size_t)(matsubaras_ - 1) * constants::max_pade;
    const int_t B = baths_, K = matsubaras_, M = matrices_, W = width_;
    const auto& T = graph_.tuples();
    auto G = [](const double* a, std::size_t i) { return complex_t(a[2 * i], a[2 * i + 1]); };
    // Re(conj(g) d)
    auto dot = [](complex_t g, complex_t d) { return g.real() * d.real() + g.imag() * d.imag(); };
    for (int_t b = 0; b < B; ++b) {
        const double lam = lambda_[(std::size_t)b];
        const complex_t nu{1.0 / invnu_[(std::size_t)b], omega_[(std::size_t)b]};
        complex_t dl_delta = 2.0 / (beta * hbar * hbar * nu);  // d Delta / d lambda
        complex_t dn_delta = -2.0 * lam / (beta * hbar * hbar * nu * nu);
        double dl = 0.0, dnr = 0.0, dni = 0.0;  // dJ/d lambda, dJ/d nu_r, dJ/d nu_i
        for (int_t k = 0; k < K; ++k) {
            const std::size_t m = (std::size_t)(b * K + k);
            complex_t dl_c, dn_c, c, nubk;
            if (k == 0) {
                complex_t corr = 0.0, dcorr = 0.0;
                for (int_t j = 0; j < K; ++j) {
                    const double eta = (double)constants::eta_pade[row + (std::size_t)j];
                    const double xi = (double)constants::xi_pade[row + (std::size_t)j];
                    const double a = xi * xi / (beta * beta * hbar * hbar);
                    corr += (2.0 * eta * nu * nu) / (a - nu * nu);
                    dcorr += 4.0 * eta * nu * a / ((a - nu * nu) * (a - nu * nu));
                }
                dl_c = 2.0 / (beta * hbar) * (1.0 - corr);
                c = lam * dl_c;
                dn_c = -2.0 * lam / (beta * hbar) * dcorr;
                nubk = nu;
                // Delta: - c_b0 / (hbar nu)
                dl_delta -= dl_c / (hbar * nu);
                dn_delta += -dn_c / (hbar * nu) + c / (hbar * nu * nu);
            } else {
                const double eta = (double)constants::eta_pade[row + (std::size_t)k];
                const double xi = (double)constants::xi_pade[row + (std::size_t)k];
                const double a = xi * xi / (beta * beta * hbar * hbar);
                dl_c = 4.0 * nu / (beta * beta * hbar * hbar) * (eta * xi) / (a - nu * nu);
                dn_c = 4.0 * lam / (beta * beta * hbar * hbar) * (eta * xi) * (a + nu * nu) / ((a - nu * nu) * (a - nu * nu));
                nubk = xi / (beta * hbar);
                dl_delta -= dl_c / (hbar * nubk);
                dn_delta -= dn_c / (hbar * nubk);
            }
            const complex_t gc = G(g_cbk, m);
            dl += dot(gc, dl_c);
            dnr += dot(gc, dn_c);
            dni += dot(gc, complex_t(0.0, 1.0) * dn_c);
        }
        const complex_t gs = G(g_same, (std::size_t)b), gp = G(g_circ, (std::size_t)b);
        dl += dot(gs, dl_delta) + dot(gp, nu / hbar);
        dnr += dot(gs, dn_delta) + dot(gp, complex_t(lam / hbar));
        dni += dot(gs, complex_t(0.0, 1.0) * dn_delta) + dot(gp, complex_t(0.0, lam / hbar));
        // S_n = -sum n_bk nu_bk: only the k = 0 mode of bath b moves with nu_b
        const std::size_t m0 = (std::size_t)(b * K);
        for (int_t u = 0; u < M; ++u) {
            const double n = (double)T[(std::size_t)(u * W) + m0];
            if (n != 0.0) {
                dnr -= n * g_pss[2 * (std::size_t)u];
                dni -= n * g_pss[2 * (std::size_t)u + 1];
            }
        }
        d_lambda[b] = dl;
        d_nu_re[b] = dnr;
        d_nu_im[b] = dni;
    }
}

void instance::set_hamiltonian(const double* h)
{
    std::copy(h, h + hamiltonian_.size(), hamiltonian_.begin());
}

void instance::set_hierarchy_top(const complex_t* top)
{
    for (std::size_t e = 0; e < (std::size_t)(states_ * states_); ++e) {
        top_[2 * e] = top[e].real();
        top_[2 * e + 1] = top[e].imag();
    }
}

void instance::set_hierarchy_top()
{
    std::fill(top_.begin(), top_.end(), 0.0);
    top_[0] = 1.0;
}

heom_desc instance::desc() const
{
    heom_desc d{};
    d.states = (int32_t)states_;
    d.matsubaras = (int32_t)matsubaras_;
    d.baths = (int32_t)baths_;
    d.baths_max_per_state = (int32_t)baths_per_state_;
    d.width = (int32_t)width_;
    d.precision = HEOM_PRECISION_FP64;
    d.matrices = matrices_;
    d.coupling = coupling_.data();
    d.hamiltonian = hamiltonian_.data();
    d.ado_tuples = graph_.tuples().data();
    d.ado_plus = graph_.plus().data();
    d.ado_minus = graph_.minus().data();
    d.pf_same_sum = pf_same_sum_.data();
    d.pf_same = pf_same_.data();
    d.pf_circ = pf_circ_.data();
    d.cbk = cbk_.data();
    return d;
}

namespace {

template <typename T>
void put(std::ofstream& f, const T& v)
{
    f.write(reinterpret_cast<const char*>(&v), sizeof(T));
}

template <typename T>
void put_array(std::ofstream& f, const T* p, std::size_t n)
{
    f.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(T)));
}

template <typename T>
T get(std::ifstream& f)
{
    T v{};
    f.read(reinterpret_cast<char*>(&v), sizeof(T));
    if (!f)
        throw config_error("instance file is truncated");
    return v;
}

template <typename T>
std::vector<T> get_array(std::ifstream& f, std::size_t n)
{
    std::vector<T> v(n);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)));
    if (!f)
        throw config_error("instance file is truncated");
    return v;
}

// enum values of the format: version_t::V_1, memory_layout_t::VANILLA,
// sites_to_states_mode_t {identity, with_ground_state, excited_state_absorption}
constexpr std::int32_t kVersion1 = 0, kVanilla = 0;

}  // namespace

void instance::write_file(const std::string& path, const double* state, bool top_only, double time) const
{
    std::ofstream f(path, std::ios::out | std::ios::binary | std::ios::trunc);
    if (!f)
        throw config_error("could not open instance file for writing: " + path);
    const std::size_t nn2 = (std::size_t)(states_ * states_ * 2), M = (std::size_t)matrices_,
                      W = (std::size_t)width_, B = (std::size_t)baths_, K = (std::size_t)matsubaras_;
    put<std::int32_t>(f, kVersion1);
    put<bool>(f, top_only);
    put<std::int32_t>(f, kVanilla);
    put<std::int32_t>(f, kVanilla);
    put<std::int32_t>(f, (std::int32_t)mode_);  // same enumerator order as the format
    put<double>(f, time);
    put<std::int32_t>(f, (std::int32_t)first_non_plus_);
    for (int_t v : {states_, matrices_, matsubaras_, baths_per_state_, baths_, width_, depth_})
        put<std::int64_t>(f, (std::int64_t)v);
    put_array(f, state, top_only ? nn2 : nn2 * M);
    put_array(f, hamiltonian_.data(), nn2);
    put_array(f, graph_.tuples().data(), M * W);
    put_array(f, graph_.plus().data(), M * W);
    put_array(f, graph_.minus().data(), M * W);
    put_array(f, pf_same_sum_.data(), 2 * M);
    put_array(f, pf_same_.data(), 2 * B);
    put_array(f, pf_circ_.data(), 2 * B);
    put_array(f, cbk_.data(), 2 * B * K);
    put_array(f, cbk_sqrt_.data(), B * K);
    if (!f)
        throw config_error("could not write instance file: " + path);
}

void instance::read_state(const std::string& path, double* state, double* time) const
{
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f)
        throw config_error("could not open instance file: " + path);
    if (get<std::int32_t>(f) != kVersion1)
        throw config_error("instance file: unknown format version");
    const bool top_only = get<bool>(f);
    if (get<std::int32_t>(f) != kVanilla || get<std::int32_t>(f) != kVanilla)
        throw config_error("instance file: unknown memory layout");
    const std::int32_t mode = get<std::int32_t>(f);
    const double t = get<double>(f);
    const std::int32_t fnp = get<std::int32_t>(f);
    std::int64_t sizes[7];
    for (auto& v : sizes)
        v = get<std::int64_t>(f);
    const std::int64_t want[7] = {states_, matrices_, matsubaras_, baths_per_state_, baths_, width_, depth_};
    for (int k = 0; k < 7; ++k)
        if (sizes[k] != want[k])
            throw config_error("instance file does not match this instance (sizes)");
    if (mode != (std::int32_t)mode_ || fnp != (std::int32_t)first_non_plus_)
        throw config_error("instance file does not match this instance (states mapping or graph)");
    const std::size_t nn2 = (std::size_t)(states_ * states_ * 2), M = (std::size_t)matrices_,
                      W = (std::size_t)width_;
    const std::vector<double> hier = get_array<double>(f, top_only ? nn2 : nn2 * M);
    (void)get_array<double>(f, nn2);  // the Hamiltonian at write time (it may have changed since)
    if (get_array<std::int32_t>(f, M * W) != graph_.tuples() || get_array<std::int32_t>(f, M * W) != graph_.plus() ||
        get_array<std::int32_t>(f, M * W) != graph_.minus())
        throw config_error("instance file does not match this instance (ADO tables)");
    // the prefactors of the file's parameters: read to check the length only, this
    // instance's own (its current parameters) stay in use
    const std::size_t B = (std::size_t)baths_, K = (std::size_t)matsubaras_;
    (void)get_array<double>(f, 2 * M + 4 * B + 3 * B * K);
    if (f.peek() != std::ifstream::traits_type::eof())
        throw config_error("instance file has trailing data");
    std::fill(state, state + nn2 * M, 0.0);
    std::memcpy(state, hier.data(), hier.size() * sizeof(double));
    if (time)
        *time = t;
}

}  // namespace heom
