// tss.hip -- the thermal state search (heom_create_tss / heom_tss_step of include/heom_hip.h),
// after dm-heom/include/heom/thermal_state_search_solver.hpp:124-300 and
// dm-heom/cl/heom/thermal_state_search.cl:10-52.
//
// One iteration of the search maps the hierarchy y to
//   k   = f_TSS(y)                   the HEOM RHS with the commutator replaced by the
//                                    elementwise product with a pseudo-Hamiltonian that
//                                    holds eps in every element, and without the S_u term
//                                    (heom_ode.cl:116-121, 130-136), i.e. eps y + the rest
//   y_u = C^T [ (C k_u C^T) / D_u ] C    D_u(i, j) = i (E_i - E_j) / hbar + eps - S_u
// with C the transposed eigenvectors of H and E its eigenvalues; a fixed point solves
// L y = 0, the stationary (thermal) hierarchy.
//
// k needs no kernel of its own: eps y is the S_u y term of the ordinary RHS with S_u = eps
// for every ADO and H = 0, so a search handle is an ordinary handle built from such a
// descriptor, whose stage kernels (STAGE 0, one evaluation, y -> ya) compute f_TSS
// unchanged.  The transform runs per ADO in one kernel for N <= 8 (a tile of 64 ADOs in
// LDS, one wavefront per matrix row; C, C^T and (i/hbar)(E_i - E_j) are wave-uniform)
// and in four passes over out-of-place planes for 8 < N <= 32.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <vector>

#include "handle.hpp"
#include "heom_device.hpp"
#include "heom_hip.h"

using namespace heom_hip;
using namespace heom_rt;

namespace {

__device__ __forceinline__ double2 cdiv(double2 a, double2 d)  // a / d
{
    const double den = d.x * d.x + d.y * d.y;
    return make_double2((a.x * d.x + a.y * d.y) / den, (a.y * d.x - a.x * d.y) / den);
}

// The search's constants on the device (FP64): c = C, ct = C^T [nel]; w[i*N+j] =
// (i/hbar)(E_i - E_j); epsm[q] = eps - S_u of the ADO in slot q (eps in unused slots,
// which the transform writes as zero).
struct TssArgs {
    const double2* __restrict__ k;  // f_TSS(y), planes
    double2* __restrict__ y;        // result, planes
    const double2* __restrict__ c;
    const double2* __restrict__ ct;
    const double2* __restrict__ w;
    const double2* __restrict__ epsm;
    const int32_t* __restrict__ q2u;  // < 0: padding or the zero slot, kept at zero
    int64_t mp;
};

// y <- C^T [ (C k C^T) / D ] C for the 64 ADOs of a tile: wavefront i owns row i.
// C k needs all rows of k (LDS); (C k) C^T and the division act on row i alone; C^T B
// needs all rows of B (LDS again, after every wavefront has finished reading k); the
// final right product acts on row i.
// c * x (+ acc) with C's entries real when REALC (the eigenvectors of a real symmetric
// H): two multiplies per product instead of four
template <bool REALC>
__device__ __forceinline__ double2 cmac(double2 c, double2 x, double2 acc)
{
    if (REALC)
        return make_double2(fma(c.x, x.x, acc.x), fma(c.x, x.y, acc.y));
    return cfma(c, x, acc);
}

template <int N, bool REALC>
__global__ __launch_bounds__(64 * N) void tss_transform(TssArgs a)
{
    __shared__ double2 t[N * N * kTile];
    const int lane = threadIdx.x & 63;
    const int i = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t q = (int64_t)blockIdx.x * kTile + lane;
#pragma unroll
    for (int j = 0; j < N; ++j)
        t[(i * N + j) * kTile + lane] = a.k[(int64_t)(i * N + j) * a.mp + q];
    __syncthreads();
    double2 r[N], s[N];
#pragma unroll
    for (int j = 0; j < N; ++j)
        r[j] = make_double2(0.0, 0.0);
#pragma unroll
    for (int m = 0; m < N; ++m) {  // r = row i of C k
        const double2 cim = a.c[i * N + m];
#pragma unroll
        for (int j = 0; j < N; ++j)
            r[j] = cmac<REALC>(cim, t[(m * N + j) * kTile + lane], r[j]);
    }
    const double2 em = a.epsm[q];
    const bool pad = a.q2u[q] < 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {  // s = row i of (C k C^T) / D
        double2 v = make_double2(0.0, 0.0);
#pragma unroll
        for (int m = 0; m < N; ++m)
            v = cmac<REALC>(a.ct[m * N + j], r[m], v);
        const double2 wij = a.w[i * N + j];
        s[j] = cdiv(v, make_double2(wij.x + em.x, wij.y + em.y));
    }
    __syncthreads();  // every wavefront is done with k
#pragma unroll
    for (int j = 0; j < N; ++j)
        t[(i * N + j) * kTile + lane] = s[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < N; ++j)
        r[j] = make_double2(0.0, 0.0);
#pragma unroll
    for (int m = 0; m < N; ++m) {  // r = row i of C^T B
        const double2 ctim = a.ct[i * N + m];
#pragma unroll
        for (int j = 0; j < N; ++j)
            r[j] = cmac<REALC>(ctim, t[(m * N + j) * kTile + lane], r[j]);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {  // row i of C^T B C
        double2 v = make_double2(0.0, 0.0);
#pragma unroll
        for (int m = 0; m < N; ++m)
            v = cmac<REALC>(a.c[m * N + j], r[m], v);
        a.y[(int64_t)(i * N + j) * a.mp + q] = pad ? make_double2(0.0, 0.0) : v;
    }
}

// 8 < N <= 32, one pass, thread = (element e, slot q):
//   left:  out[e][q] = sum_m mx[i][m] in[m*N+j][q]
//   right: out[e][q] = sum_m in[i*N+m][q] mx[m][j]   (divided by D when w is given)
__global__ void tss_pass(const double2* __restrict__ in, double2* __restrict__ out, const double2* __restrict__ mx,
                         const double2* __restrict__ w, const double2* __restrict__ epsm,
                         const int32_t* __restrict__ q2u, int64_t mp, int N, int left)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nel = N * N;
    if (t >= (int64_t)nel * mp)
        return;
    const int e = (int)(t / mp);
    const int64_t q = t % mp;
    const int i = e / N, j = e % N;
    double2 v = make_double2(0.0, 0.0);
    if (left) {
        for (int m = 0; m < N; ++m)
            v = cfma(mx[i * N + m], in[(int64_t)(m * N + j) * mp + q], v);
    } else {
        for (int m = 0; m < N; ++m)
            v = cfma(in[(int64_t)(i * N + m) * mp + q], mx[m * N + j], v);
    }
    if (w) {
        const double2 em = epsm[q];
        v = cdiv(v, make_double2(w[e].x + em.x, w[e].y + em.y));
    }
    out[t] = q2u[q] < 0 ? make_double2(0.0, 0.0) : v;
}

template <int N>
void launch_transform(heom_handle* h, const TssArgs& a, bool real_c)
{
    if (real_c)
        hipLaunchKernelGGL((tss_transform<N, true>), dim3(h->ntiles), dim3(64 * N), 0, h->stream, a);
    else
        hipLaunchKernelGGL((tss_transform<N, false>), dim3(h->ntiles), dim3(64 * N), 0, h->stream, a);
}

void check_tss(heom_handle* h)
{
    if (!h->tss)
        raise(HEOM_E_INVALID, "not a thermal-state-search handle (heom_create_tss)");
}

}  // namespace

namespace heom_rt {

void reject_tss(heom_handle* h, const char* what)
{
    if (h->tss)
        raise(HEOM_E_UNSUPPORTED, std::string(what) + " is not available on a thermal-state-search handle");
}

}  // namespace heom_rt

extern "C" {

int heom_create_tss(const heom_desc* desc, const double* c, const double* eigenvalues, const double* epsilon,
                    heom_handle** out)
{
    if (!out) {
        g_error = "out is NULL";
        return HEOM_E_INVALID;
    }
    *out = nullptr;
    heom_handle* h = nullptr;
    const int rc = guard([&] {
        if (!desc || !c || !eigenvalues || !epsilon)
            raise(HEOM_E_INVALID, "desc, c, eigenvalues and epsilon must not be NULL");
        if (desc->precision != HEOM_PRECISION_FP64)
            raise(HEOM_E_UNSUPPORTED, "the thermal state search runs in FP64 only");
        const int N = desc->states;
        const int64_t M = desc->matrices;
        if (N < 1 || M < 1)
            raise(HEOM_E_INVALID, "states and matrices must be >= 1");
        if (N > 32)
            raise(HEOM_E_UNSUPPORTED, "more than 32 states");
        const std::complex<double> eps(epsilon[0], epsilon[1]);
        if (!std::isfinite(eps.real()) || !std::isfinite(eps.imag()) || eps == 0.0)
            raise(HEOM_E_INVALID, "epsilon must be finite and non-zero");
        for (int e = 0; e < 2 * N * N; ++e)
            if (!std::isfinite(c[e]))
                raise(HEOM_E_INVALID, "c must be finite");
        for (int e = 0; e < 2 * N; ++e)
            if (!std::isfinite(eigenvalues[e]))
                raise(HEOM_E_INVALID, "eigenvalues must be finite");
        // the f_TSS operator: H = 0 and S_u = eps (thermal_state_search_solver.hpp:199-209,
        // the pseudo-Hamiltonian eps in every element through heom_ode.cl:116-121)
        heom_desc d = *desc;
        std::vector<double> zero_h((size_t)(2 * N * N), 0.0), pss((size_t)(2 * M));
        for (int64_t u = 0; u < M; ++u) {
            pss[2 * u] = eps.real();
            pss[2 * u + 1] = eps.imag();
        }
        d.hamiltonian = zero_h.data();
        d.pf_same_sum = pss.data();
        h = new heom_handle();
        create_batch(h, &d, 1);
        auto T = std::make_unique<TssState>();
        const int nel = N * N;
        std::vector<double2> cv(nel), ctv(nel), wv(nel);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                cv[i * N + j] = make_double2(c[2 * (i * N + j)], c[2 * (i * N + j) + 1]);
                ctv[i * N + j] = make_double2(c[2 * (j * N + i)], c[2 * (j * N + i) + 1]);
                // (i / hbar)(E_i - E_j) as thermal_state_search.cl:45 forms it
                const double der = eigenvalues[2 * i] - eigenvalues[2 * j];
                const double dei = eigenvalues[2 * i + 1] - eigenvalues[2 * j + 1];
                wv[i * N + j] = make_double2(-dei / kHbar, der / kHbar);
            }
        // eps - S_u per slot, in the handle's ADO order
        std::vector<int32_t> q2u((size_t)h->Mp);
        HIPC(hipMemcpyAsync(q2u.data(), h->q2u.p, q2u.size() * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        std::vector<double2> em((size_t)h->Mp, make_double2(eps.real(), eps.imag()));
        for (int64_t q = 0; q < h->Mp; ++q) {
            const int32_t u = q2u[(size_t)q];
            if (u >= 0)
                em[(size_t)q] = make_double2(eps.real() - desc->pf_same_sum[2 * u], eps.imag() - desc->pf_same_sum[2 * u + 1]);
        }
        auto up = [&](DevBuf<double2>& b, const std::vector<double2>& v) {
            b.alloc(v.size());
            HIPC(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
        };
        T->real_c = true;
        for (int e = 0; e < nel; ++e)
            T->real_c = T->real_c && cv[e].y == 0.0;
        up(T->c, cv);
        up(T->ct, ctv);
        up(T->w, wv);
        up(T->epsm, em);
        HIPC(hipStreamSynchronize(h->stream));
        h->tss = std::move(T);
    });
    if (rc != HEOM_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return HEOM_OK;
}

int heom_tss_step(heom_handle* h, int64_t n)
{
    return guard([&] {
        need(h);
        check_tss(h);
        if (n < 0)
            raise(HEOM_E_INVALID, "n must be >= 0");
        const TssState& T = *h->tss;
        const int N = h->N;
        double2* y = h->y.as<double2>();
        double2* ya = h->ya.as<double2>();
        double2* yb = h->yb.as<double2>();
        TssArgs a{ya, y, T.c.p, T.ct.p, T.w.p, T.epsm.p, h->q2u.p, h->Mp};
        const int64_t total = (int64_t)h->nel * h->Mp;
        const dim3 pgrid(blocks_for(total, 256)), pblock(256);
        for (int64_t s = 0; s < n; ++s) {
            launch_stage_buffers(h, 0, h->y.p, h->y.p, h->pbuf.p, h->ya.p, 0.0, false, false);
            switch (N) {
            case 1: launch_transform<1>(h, a, T.real_c); break;
            case 2: launch_transform<2>(h, a, T.real_c); break;
            case 3: launch_transform<3>(h, a, T.real_c); break;
            case 4: launch_transform<4>(h, a, T.real_c); break;
            case 5: launch_transform<5>(h, a, T.real_c); break;
            case 6: launch_transform<6>(h, a, T.real_c); break;
            case 7: launch_transform<7>(h, a, T.real_c); break;
            case 8: launch_transform<8>(h, a, T.real_c); break;
            default:
                // C k -> yb, (. C^T) / D -> ya, C^T . -> yb, . C -> y, in the reference's order
                hipLaunchKernelGGL(tss_pass, pgrid, pblock, 0, h->stream, ya, yb, T.c.p, nullptr, nullptr, h->q2u.p, h->Mp, N, 1);
                hipLaunchKernelGGL(tss_pass, pgrid, pblock, 0, h->stream, yb, ya, T.ct.p, T.w.p, T.epsm.p, h->q2u.p, h->Mp, N, 0);
                hipLaunchKernelGGL(tss_pass, pgrid, pblock, 0, h->stream, ya, yb, T.ct.p, nullptr, nullptr, h->q2u.p, h->Mp, N, 1);
                hipLaunchKernelGGL(tss_pass, pgrid, pblock, 0, h->stream, yb, y, T.c.p, nullptr, nullptr, h->q2u.p, h->Mp, N, 0);
                break;
            }
            HIPC(hipGetLastError());
        }
    });
}

}  // extern "C"
