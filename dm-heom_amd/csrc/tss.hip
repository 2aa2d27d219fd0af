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

#include <algorithm>
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
    const double2* __restrict__ c;  // [count][nel] (c, ct, w): the instance of each tile
    const double2* __restrict__ ct;
    const double2* __restrict__ w;
    const double2* __restrict__ epsm;
    const int32_t* __restrict__ q2u;        // < 0: padding or the zero slot, kept at zero
    const int32_t* __restrict__ tile_inst;  // instance of each tile (null: one instance)
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
    const int64_t off = a.tile_inst ? (int64_t)a.tile_inst[blockIdx.x] * N * N : 0;
    const double2* __restrict__ C = a.c + off;
    const double2* __restrict__ CT = a.ct + off;
    const double2* __restrict__ Wd = a.w + off;
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
        const double2 cim = C[i * N + m];
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
            v = cmac<REALC>(CT[m * N + j], r[m], v);
        const double2 wij = Wd[i * N + j];
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
        const double2 ctim = CT[i * N + m];
#pragma unroll
        for (int j = 0; j < N; ++j)
            r[j] = cmac<REALC>(ctim, t[(m * N + j) * kTile + lane], r[j]);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {  // row i of C^T B C
        double2 v = make_double2(0.0, 0.0);
#pragma unroll
        for (int m = 0; m < N; ++m)
            v = cmac<REALC>(C[m * N + j], r[m], v);
        a.y[(int64_t)(i * N + j) * a.mp + q] = pad ? make_double2(0.0, 0.0) : v;
    }
}

// 8 < N <= 32, one pass, thread = (element e, slot q):
//   left:  out[e][q] = sum_m mx[i][m] in[m*N+j][q]
//   right: out[e][q] = sum_m in[i*N+m][q] mx[m][j]   (divided by D when w is given)
__global__ void tss_pass(const double2* __restrict__ in, double2* __restrict__ out, const double2* __restrict__ mx,
                         const double2* __restrict__ w, const double2* __restrict__ epsm,
                         const int32_t* __restrict__ q2u, const int32_t* __restrict__ tile_inst, int64_t mp, int N,
                         int left)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nel = N * N;
    if (t >= (int64_t)nel * mp)
        return;
    const int e = (int)(t / mp);
    const int64_t q = t % mp;
    const int i = e / N, j = e % N;
    if (tile_inst) {
        const int64_t off = (int64_t)tile_inst[q / kTile] * nel;
        mx += off;
        if (w)
            w += off;
    }
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

int heom_create_tss_batch(const heom_desc* descs, int32_t count, const double* c, const double* eigenvalues,
                          const double* epsilon, heom_handle** out)
{
    if (!out) {
        g_error = "out is NULL";
        return HEOM_E_INVALID;
    }
    *out = nullptr;
    heom_handle* h = nullptr;
    const int rc = guard([&] {
        if (!descs || !c || !eigenvalues || !epsilon)
            raise(HEOM_E_INVALID, "descs, c, eigenvalues and epsilon must not be NULL");
        if (count < 1)
            raise(HEOM_E_INVALID, "count must be >= 1");
        const int N = descs[0].states;
        if (N < 1)
            raise(HEOM_E_INVALID, "states must be >= 1");
        if (N > 32)
            raise(HEOM_E_UNSUPPORTED, "more than 32 states");
        for (int k = 0; k < count; ++k) {
            if (descs[k].precision != HEOM_PRECISION_FP64)
                raise(HEOM_E_UNSUPPORTED, "the thermal state search runs in FP64 only");
            if (descs[k].states != N || descs[k].matrices < 1)
                raise(HEOM_E_INVALID, "every instance needs the same states and matrices >= 1");
            // read below for eps - S_u; the descriptor handed on carries replacements
            if (!descs[k].pf_same_sum)
                raise(HEOM_E_INVALID, "desc has a NULL array (pf_same_sum)");
        }
        const std::complex<double> eps(epsilon[0], epsilon[1]);
        if (!std::isfinite(eps.real()) || !std::isfinite(eps.imag()) || eps == 0.0)
            raise(HEOM_E_INVALID, "epsilon must be finite and non-zero");
        const int nel = N * N;
        for (int64_t e = 0; e < (int64_t)count * 2 * nel; ++e)
            if (!std::isfinite(c[e]))
                raise(HEOM_E_INVALID, "c must be finite");
        for (int64_t e = 0; e < (int64_t)count * 2 * N; ++e)
            if (!std::isfinite(eigenvalues[e]))
                raise(HEOM_E_INVALID, "eigenvalues must be finite");
        // the f_TSS operator: H = 0 and S_u = eps (thermal_state_search_solver.hpp:199-209,
        // the pseudo-Hamiltonian eps in every element through heom_ode.cl:116-121)
        std::vector<heom_desc> d(descs, descs + count);
        const std::vector<double> zero_h((size_t)(2 * nel), 0.0);
        std::vector<std::vector<double>> pss((size_t)count);
        for (int k = 0; k < count; ++k) {
            pss[(size_t)k].resize((size_t)(2 * d[(size_t)k].matrices));
            for (int64_t u = 0; u < d[(size_t)k].matrices; ++u) {
                pss[(size_t)k][(size_t)(2 * u)] = eps.real();
                pss[(size_t)k][(size_t)(2 * u + 1)] = eps.imag();
            }
            d[(size_t)k].hamiltonian = zero_h.data();
            d[(size_t)k].pf_same_sum = pss[(size_t)k].data();
        }
        h = new heom_handle();
        create_batch(h, d.data(), count);
        auto T = std::make_unique<TssState>();
        std::vector<double2> cv((size_t)count * nel), ctv((size_t)count * nel), wv((size_t)count * nel);
        for (int k = 0; k < count; ++k) {
            const double* ck = c + (size_t)k * 2 * nel;
            const double* ek = eigenvalues + (size_t)k * 2 * N;
            double2* cvk = &cv[(size_t)k * nel];
            double2* ctk = &ctv[(size_t)k * nel];
            double2* wk = &wv[(size_t)k * nel];
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) {
                    cvk[i * N + j] = make_double2(ck[2 * (i * N + j)], ck[2 * (i * N + j) + 1]);
                    ctk[i * N + j] = make_double2(ck[2 * (j * N + i)], ck[2 * (j * N + i) + 1]);
                    // (i / hbar)(E_i - E_j) as thermal_state_search.cl:45 forms it
                    const double der = ek[2 * i] - ek[2 * j];
                    const double dei = ek[2 * i + 1] - ek[2 * j + 1];
                    wk[i * N + j] = make_double2(-dei / kHbar, der / kHbar);
                }
        }
        // eps - S_u per slot, in the handle's ADO order (q2u holds batch-global uids)
        std::vector<int32_t> q2u((size_t)h->Mp);
        HIPC(hipMemcpyAsync(q2u.data(), h->q2u.p, q2u.size() * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        std::vector<double2> em((size_t)h->Mp, make_double2(eps.real(), eps.imag()));
        for (int64_t q = 0; q < h->Mp; ++q) {
            const int32_t u = q2u[(size_t)q];
            if (u < 0)
                continue;
            const int k = (int)(std::upper_bound(h->inst_u.begin(), h->inst_u.end(), (int64_t)u) - h->inst_u.begin()) - 1;
            const int64_t lu = u - h->inst_u[(size_t)k];
            em[(size_t)q] = make_double2(eps.real() - descs[k].pf_same_sum[2 * lu],
                                         eps.imag() - descs[k].pf_same_sum[2 * lu + 1]);
        }
        auto up = [&](DevBuf<double2>& b, const std::vector<double2>& v) {
            b.alloc(v.size());
            HIPC(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
        };
        T->real_c = true;
        for (const double2& x : cv)
            T->real_c = T->real_c && x.y == 0.0;
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

int heom_create_tss(const heom_desc* desc, const double* c, const double* eigenvalues, const double* epsilon,
                    heom_handle** out)
{
    return heom_create_tss_batch(desc, 1, c, eigenvalues, epsilon, out);
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
        const int32_t* ti = h->count > 1 ? h->tile_inst.p : nullptr;
        TssArgs a{ya, y, T.c.p, T.ct.p, T.w.p, T.epsm.p, h->q2u.p, ti, h->Mp};
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
                hipLaunchKernelGGL(tss_pass, pgrid, pblock, 0, h->stream, ya, yb, T.c.p, nullptr, nullptr, h->q2u.p, ti, h->Mp, N, 1);
                hipLaunchKernelGGL(tss_pass, pgrid, pblock, 0, h->stream, yb, ya, T.ct.p, T.w.p, T.epsm.p, h->q2u.p, ti, h->Mp, N, 0);
                hipLaunchKernelGGL(tss_pass, pgrid, pblock, 0, h->stream, ya, yb, T.ct.p, nullptr, nullptr, h->q2u.p, ti, h->Mp, N, 1);
                hipLaunchKernelGGL(tss_pass, pgrid, pblock, 0, h->stream, yb, y, T.c.p, nullptr, nullptr, h->q2u.p, ti, h->Mp, N, 0);
                break;
            }
            HIPC(hipGetLastError());
        }
    });
}

}  // extern "C"
