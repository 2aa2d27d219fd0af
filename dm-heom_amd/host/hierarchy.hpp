// hierarchy.hpp -- hierarchy graph, sites-to-states mappings and the instance
// (kernel inputs) of the HEOM population-dynamics / linear-absorption path.
//
// hierarchy_graph reproduces the reference's ADO numbering exactly
// (dm-heom/src/heom/hierarchy_graph.cpp:30-181): layer by layer (depth 0..D); inside a
// layer, the integer partitions of the depth into at most W parts in increasing
// lexicographic order of their non-increasing part sequences (the order the
// reference's partition generator emits, dm-heom/src/heom/integer_partition.cpp:41-105),
// each padded with zeros, sorted ascending and expanded into all of its distinct
// permutations in lexicographic order.  The ADO order is part of the reference's data
// contract, so it is kept; the device library re-orders internally.
//
// instance computes the kernel inputs of dm-heom/src/heom/instance.cpp:25-69,155-262
// (same formulas and evaluation order, std::complex<double>, the per-ADO sum in long
// double).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "config.hpp"
#include "heom_hip.h"

namespace heom {

using uid_t = std::int32_t;
constexpr uid_t invalid_uid = -2;  // dm-heom/src/heom/hierarchy_graph.cpp:24

// exact binomial coefficient; throws std::overflow_error beyond int64
int_t binomial(int_t n, int_t k);

// partitions of n into at most s parts, increasing lexicographic order of the
// non-increasing part sequences, each padded with zeros to s entries
std::vector<std::vector<int>> integer_partitions(int n, int s);

class hierarchy_graph {
public:
    hierarchy_graph(int_t width, int_t depth);
    // reference constructor form: width = baths * matsubaras
    hierarchy_graph(int_t baths, int_t matsubaras, int_t depth) : hierarchy_graph(baths * matsubaras, depth) {}

    int_t width() const { return width_; }
    int_t depth() const { return depth_; }
    int_t nodes() const { return nodes_; }
    int_t first_non_plus_uid() const { return first_non_plus_; }
    const std::vector<int_t>& nodes_per_depth() const { return nodes_per_depth_; }
    // flat [nodes * width] arrays
    const std::vector<std::int32_t>& tuples() const { return tuples_; }
    const std::vector<std::int32_t>& plus() const { return plus_; }
    const std::vector<std::int32_t>& minus() const { return minus_; }

private:
    int_t width_, depth_, nodes_, first_non_plus_;
    std::vector<int_t> nodes_per_depth_;
    std::vector<std::int32_t> tuples_, plus_, minus_;
};

// dm-heom/include/heom/sites_to_states.hpp
enum class sites_to_states_mode { identity, with_ground_state };

int_t sites_to_states(int_t sites, sites_to_states_mode mode);
complex_matrix state_hamiltonian(const complex_matrix& h, sites_to_states_mode mode);
int_matrix state_baths_coupling(const int_matrix& coupling, sites_to_states_mode mode);

// mask values (dm-heom/cl/heom/common.cl:14-19)
enum mask_value : std::uint8_t { mask_false = 0, mask_true = 1, mask_protected = 2 };
// create_hierarchy_mask for strategies "none" and "single_excitation"
// (dm-heom/src/heom/hierarchy_mask.cpp:20-110)
std::vector<std::uint8_t> create_hierarchy_mask(const std::string& strategy, const hierarchy_graph& graph,
                                                int_t matsubaras, int_t first_layer);

class instance {
public:
    instance(const config& cfg, sites_to_states_mode mode, const hierarchy_graph& graph);

    int_t states() const { return states_; }
    int_t matsubaras() const { return matsubaras_; }
    int_t baths() const { return baths_; }
    int_t baths_per_state() const { return baths_per_state_; }
    int_t ado_width() const { return width_; }
    int_t ado_depth() const { return depth_; }
    int_t matrices() const { return matrices_; }
    int_t first_non_plus_id() const { return first_non_plus_; }
    sites_to_states_mode mode() const { return mode_; }

    const std::vector<double>& hamiltonian() const { return hamiltonian_; }   // [N*N*2]
    const std::vector<std::int32_t>& coupling() const { return coupling_; }   // [N*baths_per_state]
    const std::vector<double>& pf_same_sum() const { return pf_same_sum_; }   // [2M]
    const std::vector<double>& pf_same() const { return pf_same_; }           // [2B]
    const std::vector<double>& pf_circ() const { return pf_circ_; }           // [2B]
    const std::vector<double>& cbk() const { return cbk_; }                   // [2BK]
    const std::vector<double>& cbk_sqrt() const { return cbk_sqrt_; }         // [BK]
    const hierarchy_graph& graph() const { return graph_; }

    void set_hamiltonian(const double* h);
    // bath parameters in SI units after config::post_process (lambda in J, invnu in s,
    // Omega in rad/s); set_baths recomputes every prefactor
    const std::vector<double>& lambda() const { return lambda_; }
    const std::vector<double>& invnu() const { return invnu_; }
    const std::vector<double>& omega() const { return omega_; }
    void set_baths(const double* lambda, const double* invnu, const double* omega);
    // chain rule from the gradients of J with respect to the prefactor arrays (complex
    // entries as dJ/dRe + i dJ/dIm; pf_same_sum [2M], pf_same [2B], pf_circ [2B],
    // cbk [2BK]) to dJ/d lambda_b (per J) and dJ/d Re nu_b, dJ/d Im nu_b (per 1/s),
    // nu_b = 1/invnu_b + i Omega_b
    void bath_gradient(const double* g_pss, const double* g_same, const double* g_circ, const double* g_cbk,
                       double* d_lambda, double* d_nu_re, double* d_nu_im) const;
    // top ADO to load at the start (instance::set_hierarchy_top, instance.cpp:483-495)
    const std::vector<double>& top() const { return top_; }                   // [N*N*2]
    void set_hierarchy_top(const complex_t* top);
    void set_hierarchy_top();  // element [0][0] = 1 (linear absorption)

    // descriptor of the kernel inputs for heom_create (points into this instance)
    heom_desc desc() const;

    // The reference's binary instance format (dm-heom/src/heom/instance.cpp:281-377,
    // fields of dm-heom/include/heom/instance.hpp:182-211): header, then the hierarchy
    // (all ADOs, or the top only) and every kernel-input array.  write_file stores this
    // instance with `state` [M*N*N*2] (or its top ADO when top_only); read_state checks
    // that a file describes this instance (sizes, graph tables) and returns its
    // hierarchy (zeros below the top for a top-only file) and time.
    void write_file(const std::string& path, const double* state, bool top_only, double time) const;
    void read_state(const std::string& path, double* state, double* time) const;

    // bytes of the instance's host arrays and of its graph's tables (the reference's
    // instance::allocated_byte, as its memory summary prints it)
    std::size_t allocated_byte() const
    {
        std::size_t b = 0;
        for (const auto* v : {&hamiltonian_, &pf_same_sum_, &pf_same_, &pf_circ_, &cbk_, &cbk_sqrt_, &top_, &lambda_,
                              &invnu_, &omega_})
            b += v->size() * sizeof(double);
        b += coupling_.size() * sizeof(std::int32_t);
        b += (graph_.tuples().size() + graph_.plus().size() + graph_.minus().size()) * sizeof(std::int32_t);
        return b;
    }

private:
    void precompute();

    sites_to_states_mode mode_;
    const hierarchy_graph& graph_;
    int_t states_, matsubaras_, baths_, baths_per_state_, width_, depth_, matrices_, first_non_plus_;
    std::vector<double> hamiltonian_, pf_same_sum_, pf_same_, pf_circ_, cbk_, cbk_sqrt_, top_;
    std::vector<double> lambda_, invnu_, omega_;
    double temperature_ = 0.0;
    std::vector<std::int32_t> coupling_;
};

}  // namespace heom
