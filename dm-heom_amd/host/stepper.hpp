// stepper.hpp -- explicit Runge-Kutta steppers for the apps, on the per-evaluation
// interface of include/heom_hip.h (heom_evaluate, heom_lincomb, heom_swap_buffers).
//
// The reference picks its stepper by the config's solver.stepper_type
// (dm-heom/src/heom/config.cpp:32-34, :49) from noma/num, which is not vendored; its
// example configs use "rk_rk4", the classical method that heom_step_rk4 runs fused.  The
// other names are this build's, each a textbook Butcher tableau (the same table as
// heom_amd/steppers.py, whose tests pin every tableau by its order of accuracy):
//   rk_euler, rk_midpoint, rk_heun, rk_kutta3, rk_rk4, rk_rk4_38.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "heom_hip.h"

namespace heom {

struct rk_tableau {
    std::vector<std::vector<double>> a;  // a[i] = row of stage i + 1: a_{i+1,0} .. a_{i+1,i}
    std::vector<double> b, c;

    int stages() const { return (int)b.size(); }
    // the only non-zero entries are a_{i+1,i}: the fused accumulating form applies
    bool subdiagonal() const
    {
        for (const auto& row : a)
            for (std::size_t j = 0; j + 1 < row.size(); ++j)
                if (row[j] != 0.0)
                    return false;
        return true;
    }
};

// nullptr for an unknown name
inline const rk_tableau* find_tableau(const std::string& name)
{
    static const rk_tableau euler{{}, {1.0}, {0.0}};
    static const rk_tableau midpoint{{{0.5}}, {0.0, 1.0}, {0.0, 0.5}};
    static const rk_tableau heun{{{1.0}}, {0.5, 0.5}, {0.0, 1.0}};
    static const rk_tableau kutta3{{{0.5}, {-1.0, 2.0}}, {1.0 / 6, 2.0 / 3, 1.0 / 6}, {0.0, 0.5, 1.0}};
    static const rk_tableau rk4{{{0.5}, {0.0, 0.5}, {0.0, 0.0, 1.0}}, {1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6},
                                {0.0, 0.5, 0.5, 1.0}};
    static const rk_tableau rk4_38{{{1.0 / 3}, {-1.0 / 3, 1.0}, {1.0, -1.0, 1.0}}, {0.125, 0.375, 0.375, 0.125},
                                   {0.0, 1.0 / 3, 2.0 / 3, 1.0}};
    if (name == "rk_euler") return &euler;
    if (name == "rk_midpoint") return &midpoint;
    if (name == "rk_heun") return &heun;
    if (name == "rk_kutta3") return &kutta3;
    if (name == "rk_rk4") return &rk4;
    if (name == "rk_rk4_38") return &rk4_38;
    return nullptr;
}

inline const char* stepper_names() { return "rk_euler, rk_midpoint, rk_heun, rk_kutta3, rk_rk4, rk_rk4_38"; }

// one method on a handle's state (buffer 0); buffers: y, the accumulator and two stage
// inputs for a sub-diagonal tableau, y, the stage input and one buffer per k_i otherwise
class explicit_rk_stepper {
public:
    explicit_rk_stepper(heom_handle* h, const rk_tableau& t) : h_(h), t_(t)
    {
        const int need = t.subdiagonal() ? 4 : t.stages() + 2;
        int have = 0;
        check(heom_buffer_count(h, &have));
        if (have < need)
            check(heom_add_buffers(h, need - have, nullptr));
    }

    void step(double time, double dt)
    {
        const int s = t_.stages();
        if (t_.subdiagonal()) {
            int src = 0;
            for (int i = 0; i < s; ++i) {
                const bool more = i + 1 < s;
                const int nxt = more ? 2 + (i % 2) : HEOM_NO_BUFFER;
                check(heom_evaluate(h_, time + t_.c[i] * dt, dt, src, nxt, more ? 0 : HEOM_NO_BUFFER,
                                    more ? dt * t_.a[i][i] : 0.0, 1, dt * t_.b[i], i == 0));
                src = nxt;
            }
        } else {
            for (int i = 0; i < s; ++i) {
                int src = 0;
                if (i > 0) {
                    std::vector<int> ix{0};
                    std::vector<double> cf{1.0};
                    for (int j = 0; j < i; ++j)
                        if (t_.a[i - 1][j] != 0.0) {
                            ix.push_back(2 + j);
                            cf.push_back(dt * t_.a[i - 1][j]);
                        }
                    check(heom_lincomb(h_, 1, (int)ix.size(), ix.data(), cf.data()));
                    src = 1;
                }
                check(heom_evaluate(h_, time + t_.c[i] * dt, dt, src, 2 + i, HEOM_NO_BUFFER, 1.0, HEOM_NO_BUFFER,
                                    0.0, 0));
            }
            std::vector<int> ix{0};
            std::vector<double> cf{1.0};
            for (int i = 0; i < s; ++i)
                if (t_.b[i] != 0.0) {
                    ix.push_back(2 + i);
                    cf.push_back(dt * t_.b[i]);
                }
            check(heom_lincomb(h_, 1, (int)ix.size(), ix.data(), cf.data()));
        }
        check(heom_swap_buffers(h_, 0, 1));
    }

private:
    static void check(int rc)
    {
        if (rc != HEOM_OK)
            throw std::runtime_error(std::string("heom: ") + heom_last_error());
    }
    heom_handle* h_;
    rk_tableau t_;
};

}  // namespace heom
