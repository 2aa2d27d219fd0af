// evaluate.hip -- the per-evaluation interface of include/heom_hip.h: one evaluation of
// the HEOM operator between the handle's device state buffers, with the accumulating
// forms of the reference's ODE wrapper (dm-heom/include/heom/ode.hpp:40-43,
// dm-heom/src/heom/ode.cpp:97-196; the kernel side dm-heom/cl/heom/heom_ode.cl:416-434)
// and its pre-evaluate actions (ode.hpp:53-55, ode.cpp:17-21), plus the buffer algebra a
// caller needs to compose an explicit stepper of its own.  The reference takes its
// steppers from noma/num, which is not vendored; heom_amd/steppers.py builds explicit
// Runge-Kutta methods from Butcher tableaus on these calls.
//
// Buffers: 0 is the state (heom_set_state / heom_get_state / heom_get_top), 1-3 are the
// work buffers of the fused RK4 step, 4.. are added with heom_add_buffers.  All are
// complex128 planes in the handle's layout; ADO slots without an ADO (the zero slot and
// the padding of the last tile) hold zeros in every buffer, which every operation here
// keeps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <utility>

#include "handle.hpp"
#include "heom_device.hpp"
#include "heom_hip.h"

using namespace heom_hip;
using namespace heom_rt;

namespace {

constexpr int kMaxLinComb = 8;

// out = base + out_c f (base optional), acc = (acc_init ? in : acc) + acc_c f, per
// complex element; out and acc may alias in and base (each element is read before it is
// written by the same thread)
struct CombineArgs {
    const double2* f;
    const double2* in;
    const double2* base;
    double2* out;
    double out_c;
    double2* acc;
    double acc_c;
    int acc_init;
    int64_t n;
};

__global__ void combine_kernel(CombineArgs c)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 f = c.f[i];
        if (c.acc) {
            const double2 a = c.acc_init ? c.in[i] : c.acc[i];
            c.acc[i] = make_double2(fma(c.acc_c, f.x, a.x), fma(c.acc_c, f.y, a.y));
        }
        if (c.out) {
            const double2 b = c.base ? c.base[i] : make_double2(0.0, 0.0);
            c.out[i] = make_double2(fma(c.out_c, f.x, b.x), fma(c.out_c, f.y, b.y));
        }
    }
}

// dst = sum_k c_k src_k, summed in k order
struct LinCombArgs {
    const double2* src[kMaxLinComb];
    double c[kMaxLinComb];
    int nsrc;
    double2* dst;
    int64_t n;
};

__global__ void lincomb_kernel(LinCombArgs a)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * blockDim.x) {
        double2 s = make_double2(a.c[0] * a.src[0][i].x, a.c[0] * a.src[0][i].y);
        for (int k = 1; k < a.nsrc; ++k) {
            const double2 v = a.src[k][i];
            s.x = fma(a.c[k], v.x, s.x);
            s.y = fma(a.c[k], v.y, s.y);
        }
        a.dst[i] = s;
    }
}

unsigned grid_for(int64_t n)
{
    return (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
}

void need_fp64(heom_handle* h, const char* what)
{
    need(h);
    reject_tss(h, what);
    if (h->fp32())
        raise(HEOM_E_UNSUPPORTED, std::string(what) + " works on FP64 handles (FP32 storage keeps FP64 shadows "
                                                      "of the tier-0/1 ADOs that these buffers do not carry)");
}

int64_t plane_elements(const heom_handle* h) { return (int64_t)h->nel * h->Mp; }

}  // namespace

namespace heom_rt {

DevBuf<unsigned char>& state_buffer(heom_handle* h, int b)
{
    switch (b) {
    case 0: return h->y;
    case 1: return h->ya;
    case 2: return h->yb;
    case 3: return h->pbuf;
    default: break;
    }
    if (b >= 4 && b - 4 < (int)h->extra.size())
        return *h->extra[(size_t)(b - 4)];
    raise(HEOM_E_INVALID, "no state buffer " + std::to_string(b) + " (buffers 0.." +
                              std::to_string(3 + (int)h->extra.size()) + ")");
}

void reset_extra_buffers(heom_handle* h)
{
    h->fscratch.free();
    for (auto& b : h->extra) {
        b->alloc((size_t)h->nel * h->Mp * h->cbytes());
        HIPC(hipMemsetAsync(b->p, 0, b->bytes(), h->stream));
    }
}

}  // namespace heom_rt

extern "C" {

int heom_buffer_count(heom_handle* h, int* count)
{
    return guard([&] {
        need(h);
        if (!count)
            raise(HEOM_E_INVALID, "count is NULL");
        *count = 4 + (int)h->extra.size();
    });
}

int heom_add_buffers(heom_handle* h, int n, int* count)
{
    return guard([&] {
        need_fp64(h, "heom_add_buffers");
        if (n < 0 || 4 + (int)h->extra.size() + n > HEOM_MAX_BUFFERS)
            raise(HEOM_E_INVALID, "at most " + std::to_string(HEOM_MAX_BUFFERS) + " state buffers");
        for (int k = 0; k < n; ++k) {
            h->extra.emplace_back(new DevBuf<unsigned char>());
            auto& b = *h->extra.back();
            b.alloc((size_t)h->nel * h->Mp * h->cbytes());
            HIPC(hipMemsetAsync(b.p, 0, b.bytes(), h->stream));
        }
        HIPC(hipStreamSynchronize(h->stream));
        if (count)
            *count = 4 + (int)h->extra.size();
    });
}

int heom_set_buffer(heom_handle* h, int b, const double* state)
{
    return guard([&] {
        need_fp64(h, "heom_set_buffer");
        if (!state)
            raise(HEOM_E_INVALID, "state is NULL");
        upload_state(h, state, state_buffer(h, b).p);
    });
}

int heom_get_buffer(heom_handle* h, int b, double* state)
{
    return guard([&] {
        need_fp64(h, "heom_get_buffer");
        if (!state)
            raise(HEOM_E_INVALID, "state is NULL");
        download_state(h, state_buffer(h, b).p, nullptr, state);
    });
}

int heom_swap_buffers(heom_handle* h, int a, int b)
{
    return guard([&] {
        need_fp64(h, "heom_swap_buffers");
        auto& x = state_buffer(h, a);
        auto& y = state_buffer(h, b);
        std::swap(x.p, y.p);
        std::swap(x.n, y.n);
        drop_graph(h);  // a captured step holds the buffer addresses
    });
}

int heom_lincomb(heom_handle* h, int dst, int n, const int* src, const double* coeff)
{
    return guard([&] {
        need_fp64(h, "heom_lincomb");
        if (n < 1 || n > kMaxLinComb || !src || !coeff)
            raise(HEOM_E_INVALID, "heom_lincomb takes 1 to 8 sources and their coefficients");
        LinCombArgs a{};
        for (int k = 0; k < n; ++k) {
            a.src[k] = state_buffer(h, src[k]).as<double2>();
            a.c[k] = coeff[k];
        }
        a.nsrc = n;
        a.dst = state_buffer(h, dst).as<double2>();
        a.n = plane_elements(h);
        hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(a.n)), dim3(256), 0, h->stream, a);
        HIPC(hipGetLastError());
    });
}

int heom_add_pre_evaluate(heom_handle* h, heom_pre_evaluate_fn fn, void* user)
{
    return guard([&] {
        need(h);
        if (!fn)
            raise(HEOM_E_INVALID, "fn is NULL");
        h->pre_evaluate.emplace_back(fn, user);
    });
}

int heom_clear_pre_evaluate(heom_handle* h)
{
    return guard([&] {
        need(h);
        h->pre_evaluate.clear();
    });
}

int heom_evaluate(heom_handle* h, double time, double step_size, int in, int out, int out_base, double out_coeff,
                  int acc, double acc_coeff, int acc_init)
{
    return guard([&] {
        need_fp64(h, "heom_evaluate");
        auto& bin = state_buffer(h, in);
        unsigned char* pout = out >= 0 ? state_buffer(h, out).p : nullptr;
        unsigned char* pbase = out >= 0 && out_base >= 0 ? state_buffer(h, out_base).p : nullptr;
        unsigned char* pacc = acc >= 0 ? state_buffer(h, acc).p : nullptr;
        if (pout && pout == pacc)
            raise(HEOM_E_INVALID, "out and acc must be different buffers");
        // the reference's pre-evaluate actions, in the order they were added
        for (const auto& a : h->pre_evaluate) {
            const int rc = a.first(h, time, step_size, in, out, a.second);
            if (rc != 0)
                raise(HEOM_E_INVALID, "pre-evaluate action returned " + std::to_string(rc));
        }
        h->timing.rhs_launches += 1;
        if (pout && !pbase && out_coeff == 1.0 && !pacc && pout != bin.p) {
            // pure evaluation: the stage writes f into buffer out directly
            launch_stage_buffers(h, 0, bin.p, nullptr, nullptr, pout, 0.0, false, false);
            return;
        }
        const size_t bytes = (size_t)plane_elements(h) * sizeof(double2);
        if (h->fscratch.bytes() != bytes) {
            h->fscratch.alloc(bytes);
            HIPC(hipMemsetAsync(h->fscratch.p, 0, bytes, h->stream));
        }
        // f = L(buffer in) into the scratch planes (the pure-evaluation stage), then one
        // elementwise pass for the accumulating forms
        launch_stage_buffers(h, 0, bin.p, nullptr, nullptr, h->fscratch.p, 0.0, false, false);
        if (pout || pacc) {
            CombineArgs c{};
            c.f = h->fscratch.as<double2>();
            c.in = bin.as<double2>();
            c.base = reinterpret_cast<const double2*>(pbase);
            c.out = reinterpret_cast<double2*>(pout);
            c.out_c = out_coeff;
            c.acc = reinterpret_cast<double2*>(pacc);
            c.acc_c = acc_coeff;
            c.acc_init = acc_init ? 1 : 0;
            c.n = plane_elements(h);
            hipLaunchKernelGGL(combine_kernel, dim3(grid_for(c.n)), dim3(256), 0, h->stream, c);
            HIPC(hipGetLastError());
        }
    });
}

}  // extern "C"
