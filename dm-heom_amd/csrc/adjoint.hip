// adjoint.hip -- gradients of observables of the RK4 trajectory by the discrete adjoint
// (heom_adjoint_forward / heom_adjoint_backward / heom_rhs_adjoint of include/heom_hip.h).
// The reference has no differentiation; the objective is
//   J = sum_r Re(w_r X_r),  X_r = Tr(O_k sigma_0(t_r)),  t_r = r * observe_steps * dt,
// summed over the instances k of the handle.
//
// One RK4 step y' = y + dt/6 (k1 + 2 k2 + 2 k3 + k4), k_s = L Y_s, is linear in y, and
// with lambda = dJ/dy' the adjoint stages
//   A1 = lambda, A2 = lambda + dt/2 L^dag A1, A3 = lambda + dt/2 L^dag A2, A4 = lambda + dt L^dag A3
// give dJ/dy = lambda + dt/6 L^dag (A1 + 2 A2 + 2 A3 + A4) -- the same four-stage schedule
// with L^dag in place of L (stage_kernel_adj.hip) -- and the step's contribution to the
// gradient with respect to any coefficient kappa of L,
//   dt/6 <A4, dL/dkappa Y1> + dt/3 <A3, dL/dkappa Y2> + dt/3 <A2, dL/dkappa Y3> + dt/6 <A1, dL/dkappa Y4>,
// with <a, b> = Re sum conj(a) b (the RHS is complex-linear, so its real transpose is
// L^dag).  When the state and the three forward stage vectors of every step fit in half
// the free device memory (12 GB for 1000 steps of 3003 ADOs of 8 x 8), the forward sweep
// keeps them all and the backward sweep recomputes nothing: per step its three adjoint
// stages, one contraction launch covering all four stage pairs, and the fourth stage.
// Otherwise (or with HEOM_ADJOINT_RECOMPUTE=1 in the environment) the forward sweep keeps
// the state every checkpoint_steps steps, and the backward sweep recomputes each
// segment's states from its checkpoint and each step's forward stage vectors from its
// state.  Both modes run the same kernels on the same inputs: their gradients agree to
// the last bit.
//
// The contraction kernel accumulates, per tile of 64 ADOs and in FP64, the sums the
// gradients of all coefficient arrays of heom_desc follow from:
//   T_ij  = sum_k Y_ki conj(mu_kj) - conj(mu_ik) Y_jk        (Hamiltonian)
//   D_ij  = conj(mu_ij) Y_ij                                 (the "same" constants)
//   P_e,t = i sqrt(n_m + 1) conj(mu_e) Y^{n+e_m}_e           (plus term t of element e)
//   Q_e,t = sqrt(n_m) conj(mu_e) Y^{n-e_m}_e                 (minus term t of element e)
// and per ADO sum_e conj(mu_e) Y_e (pf_same_sum); the host maps them onto the arrays
// through the term constants of build_coefficients (heom_hip.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "handle.hpp"
#include "heom_device.hpp"
#include "heom_hip.h"

using namespace heom_hip;
using namespace heom_rt;

namespace {

// The four (adjoint stage, forward stage) pairs of one RK4 step: pair p contracts
// mu = A_{p+1} with Y = Y_{4-p} with weight w_p (dt/6, dt/3, dt/3, dt/6).
constexpr int kPairs = 4;
struct ContractArgs {
    const double2* __restrict__ mu[kPairs];  // adjoint stage vectors (planes)
    const double2* __restrict__ y[kPairs];   // forward stage vectors (planes)
    double w[kPairs];
    const int2* __restrict__ nb;  // neighbour table
    double2* __restrict__ part;   // [ntiles][kPairs][nel * (2 + 2 NT)] per-tile sums
    double2* __restrict__ gs;     // [kPairs][Mp] per-ADO sums
    int64_t mp;
};

// v moved across lanes by one DPP pattern (both 32-bit halves); lanes of rows outside
// ROWS read 0
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}

// conj(a) * b
__device__ __forceinline__ double2 cdot(double2 a, double2 b)
{
    return make_double2(fma(a.x, b.x, a.y * b.y), fma(a.x, b.y, -a.y * b.x));
}

// Sums over the 64 lanes of P values at once (P a power of two, 16 <= P <= 64).  Four
// halving exchanges inside each 16-lane row read the lane XOR 1 and XOR 2 (quad_perm),
// then the lane 4 and 8 places along (row_ror): the lane read holds the same set of
// values and the opposite lane bit, each lane keeps the half its bit selects and adds
// the other lane's copy of it, so the work halves at every exchange instead of staying at
// P sums of six moves each.  (After the row_ror:4 exchange lane l holds x_l + x_{l+4} of
// its half and lane l + 8 holds x_{l+8} + x_{l+12} of the same half: the row_ror:8
// exchange adds those, each of the four lanes once.)  A lane then holds P / 16 values,
// each summed over its row; two LDS-crossbar shuffles add the rows.  Every lane l ends
// with the total of value bitrev4(l & 15) * P / 16 + u in x[u].
template <int CTRL, int H>
__device__ __forceinline__ void halve_exchange(double* x, bool upper)
{
#pragma unroll
    for (int t = 0; t < H; ++t) {
        const double keep = upper ? x[t + H] : x[t];
        const double send = upper ? x[t] : x[t + H];
        x[t] = keep + dpp_move<CTRL, 0xf>(send);
    }
}

template <int P>
__device__ __forceinline__ void transposed_sum(double (&x)[P])
{
    static_assert(P >= 16 && P <= 64 && (P & (P - 1)) == 0, "16, 32 or 64 values");
    const int lane = threadIdx.x & 63;
    halve_exchange<0xb1, P / 2>(x, lane & 1);           // quad_perm [1, 0, 3, 2]
    halve_exchange<0x4e, P / 4>(x, (lane >> 1) & 1);    // quad_perm [2, 3, 0, 1]
    halve_exchange<0x124, P / 8>(x, (lane >> 2) & 1);   // row_ror:4
    halve_exchange<0x128, P / 16>(x, (lane >> 3) & 1);  // row_ror:8
#pragma unroll
    for (int u = 0; u < P / 16; ++u) {
        x[u] += __shfl_xor(x[u], 16);
        x[u] += __shfl_xor(x[u], 32);
    }
}

// the value a lane holds after transposed_sum in x[u]
template <int P>
__device__ __forceinline__ int transposed_index(int lane, int u)
{
    const int t = lane & 15;
    const int rev = ((t & 1) << 3) | ((t & 2) << 1) | ((t & 4) >> 1) | ((t & 8) >> 3);
    return rev * (P / 16) + u;
}

constexpr int pow2_at_least_16(int v) { return v <= 16 ? 16 : v <= 32 ? 32 : 64; }

// Work split as the stage kernel: workgroup = tile of 64 ADOs (lane = ADO) and one
// stage pair (blockIdx.y), wavefront = 4 columns j of one row i; the tile of both
// vectors is staged in LDS for T.  Each (tile, pair, slot) has one writer.
template <int N, int NT>
__global__ __launch_bounds__(64 * stage_waves(N)) void adjoint_contract(ContractArgs a,
                                                                        const Term* __restrict__ terms)
{
    constexpr int CW = kStageColumns;
    constexpr int NJW = (N + CW - 1) / CW;
    constexpr int WAVES = N * NJW;
    constexpr int SLOTS = 2 + 2 * NT;
    constexpr int PV = pow2_at_least_16(2 * SLOTS);  // real values per element, padded
    // terms whose loads are in flight together: two once the element's sums take 32 or 64
    // registers' worth of values (four or eight terms), so that nothing spills at N = 8
    constexpr int TB = PV > 16 ? (NT < 2 ? NT : 2) : (NT < 4 ? NT : 4);
    __shared__ double2 sy[N * N * kTile];
    __shared__ double2 sm[N * N * kTile];
    __shared__ double2 red[WAVES * kTile];

    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = wv / NJW;
    const int jb = (wv % NJW) * CW;
    const int tile = blockIdx.x, pr = blockIdx.y;
    const int64_t mp = a.mp;
    const int64_t q = (int64_t)tile * kTile + lane;
    const double2* __restrict__ Y = a.y[pr];
    const double2* __restrict__ MU = a.mu[pr];
    const double w = a.w[pr];

    double2 ym[CW], mm[CW];
#pragma unroll
    for (int c = 0; c < CW; ++c) {
        const int j = jb + c;
        if (j < N) {
            ym[c] = Y[(int64_t)(i * N + j) * mp + q];
            mm[c] = MU[(int64_t)(i * N + j) * mp + q];
            sy[(i * N + j) * kTile + lane] = ym[c];
            sm[(i * N + j) * kTile + lane] = mm[c];
        }
    }
    __syncthreads();

    double2 dsum = make_double2(0.0, 0.0);
    double2* out = a.part + ((int64_t)tile * kPairs + pr) * (N * N * SLOTS);
    static_assert(2 * SLOTS <= 64, "at most 64 real sums per element");
#pragma unroll
    for (int c = 0; c < CW; ++c) {
        const int j = jb + c;
        if (j >= N)
            break;
        const int e = i * N + j;
        const double2* plane = Y + (int64_t)e * mp;
        // the element's SLOTS complex sums (consecutive in `out`) as 2 * SLOTS real values
        // (re, im per slot), summed over the lanes at once and added to `out` by the lanes
        // that hold them
        double vals[PV];
#pragma unroll
        for (int v = 2 * SLOTS; v < PV; ++v)
            vals[v] = 0.0;
        auto put = [&](int k, double re, double im) {
            vals[2 * k] = re;
            vals[2 * k + 1] = im;
        };
        // neighbour loads of the first term block go out before the LDS work
        int2 nbv[TB];
        double2 xp[TB], xm[TB];
#pragma unroll
        for (int k = 0; k < TB; ++k)
            nbv[k] = a.nb[(int64_t)terms[e * NT + k].mode * mp + q];
#pragma unroll
        for (int k = 0; k < TB; ++k) {
            xp[k] = plane[(uint32_t)nbv[k].x >> 6];
            xm[k] = plane[(uint32_t)nbv[k].y];
        }
        double2 tt = make_double2(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const double2 a1 = cdot(sm[(k * N + j) * kTile + lane], sy[(k * N + i) * kTile + lane]);
            const double2 a2 = cdot(sm[(i * N + k) * kTile + lane], sy[(j * N + k) * kTile + lane]);
            tt.x += a1.x - a2.x;
            tt.y += a1.y - a2.y;
        }
        const double2 d = cdot(mm[c], ym[c]);
        dsum.x += d.x;
        dsum.y += d.y;
        put(0, tt.x, tt.y);
        put(1, d.x, d.y);
#pragma unroll
        for (int k0 = 0; k0 < NT; k0 += TB) {
            if (k0 > 0) {
#pragma unroll
                for (int k = 0; k < TB; ++k)
                    nbv[k] = a.nb[(int64_t)terms[e * NT + k0 + k].mode * mp + q];
#pragma unroll
                for (int k = 0; k < TB; ++k) {
                    xp[k] = plane[(uint32_t)nbv[k].x >> 6];
                    xm[k] = plane[(uint32_t)nbv[k].y];
                }
            }
#pragma unroll
            for (int k = 0; k < TB; ++k) {
                const int v = nbv[k].x & 63;
                const double sp = sqrt((double)(v + 1)), sn = sqrt((double)v);
                const double2 pp = cdot(mm[c], xp[k]);  // times i sqrt(n+1)
                const double2 m = cdot(mm[c], xm[k]);
                put(2 + k0 + k, -sp * pp.y, sp * pp.x);
                put(2 + NT + k0 + k, sn * m.x, sn * m.y);
            }
        }
        transposed_sum<PV>(vals);
        if (lane < 16) {
            double* o = reinterpret_cast<double*>(out + e * SLOTS);
#pragma unroll
            for (int u = 0; u < PV / 16; ++u) {
                const int v = transposed_index<PV>(lane, u);
                if (v < 2 * SLOTS)
                    o[v] = fma(w, vals[u], o[v]);
            }
        }
    }
    red[wv * kTile + lane] = dsum;
    __syncthreads();
    if (wv == 0) {
        double2 t = make_double2(0.0, 0.0);
        for (int k = 0; k < WAVES; ++k) {
            t.x += red[k * kTile + lane].x;
            t.y += red[k * kTile + lane].y;
        }
        double2* g = a.gs + (int64_t)pr * mp + q;
        *g = make_double2(fma(w, t.x, g->x), fma(w, t.y, g->y));
    }
}

// out[k][s] = sum of part[r][s] over the rows r (tile r / per) of instance k (tile_inst;
// null = one instance)
__global__ void sum_tiles(const double2* __restrict__ part, const int32_t* __restrict__ tile_inst,
                          double2* __restrict__ out, int rows, int per, int slots, int count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * slots)
        return;
    const int k = t / slots, s = t % slots;
    double2 acc = make_double2(0.0, 0.0);
    for (int r = 0; r < rows; ++r)
        if (!tile_inst ? k == 0 : tile_inst[r / per] == k) {
            acc.x += part[(int64_t)r * slots + s].x;
            acc.y += part[(int64_t)r * slots + s].y;
        }
    out[t] = acc;
}

__global__ void conj_copy(const double2* __restrict__ in, double2* __restrict__ out, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
        out[t] = make_double2(in[t].x, -in[t].y);
}

// -H^dag / hbar per instance from H / hbar
__global__ void adjoint_hamiltonian(const double2* __restrict__ hs, double2* __restrict__ out, int count, int N)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nel = N * N;
    if (t >= count * nel)
        return;
    const int k = t / nel, e = t % nel, i = e / N, j = e % N;
    const double2 v = hs[(int64_t)k * nel + j * N + i];
    out[t] = make_double2(-v.x, v.y);
}

// lambda_top(k) += src[k] for every instance: thread = (k, e)
__global__ void add_tops(double2* __restrict__ lam, const double2* __restrict__ src, const int64_t* __restrict__ qs,
                         int count, int64_t mp, int nel)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * nel)
        return;
    const int k = t / nel, e = t % nel;
    double2* p = lam + (int64_t)e * mp + qs[k];
    *p = make_double2(p->x + src[t].x, p->y + src[t].y);
}

// populations and Tr(O sigma_0) of every instance straight from state planes (the top
// ADO's element e is plane e at slot qs[k]): observe_tops' sums in its order, in one
// launch instead of a gather and a reduction
__global__ __launch_bounds__(64) void observe_planes(const double2* __restrict__ y, const int64_t* __restrict__ qs,
                                                     int64_t mp, const double2* __restrict__ ops,
                                                     double* __restrict__ pops, double2* __restrict__ traces, int N)
{
    const int k = blockIdx.x, lane = threadIdx.x, nel = N * N;
    const int64_t qk = qs[k];
    double re = 0.0, im = 0.0;
    for (int e = lane; e < nel; e += 64) {
        const int i = e / N, j = e % N;
        const double2 s = y[(int64_t)e * mp + qk];
        if (i == j)
            pops[(int64_t)k * N + i] = s.x;
        // Tr(D sigma) = sum_ij D_ji sigma_ij
        const double2 d = ops[(int64_t)k * nel + j * N + i];
        re += fma(d.x, s.x, -d.y * s.y);
        im += fma(d.x, s.y, d.y * s.x);
    }
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_xor(re, off);
        im += __shfl_xor(im, off);
    }
    if (lane == 0)
        traces[k] = make_double2(re, im);
}

template <int N, int NT>
void launch_contract_t(const ContractArgs& a, const Term* terms, int ntiles, hipStream_t st)
{
    hipLaunchKernelGGL((adjoint_contract<N, NT>), dim3(ntiles, kPairs), dim3(64 * stage_waves(N)), 0, st, a, terms);
}

template <int N>
void launch_contract_n(int nt, const ContractArgs& a, const Term* terms, int ntiles, hipStream_t st)
{
    switch (nt) {
    case 2: launch_contract_t<N, 2>(a, terms, ntiles, st); break;
    case 4: launch_contract_t<N, 4>(a, terms, ntiles, st); break;
    default: launch_contract_t<N, 8>(a, terms, ntiles, st); break;
    }
}

// the gradient contributions of one RK4 step: pairs (A1, Y4), (A2, Y3), (A3, Y2), (A4, Y1)
void launch_contract(heom_handle* h, const void* const* mu, const void* const* y, double dt)
{
    AdjointState& A = *h->adj;
    ContractArgs a{};
    const double w[kPairs] = {dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0};
    for (int p = 0; p < kPairs; ++p) {
        a.mu[p] = reinterpret_cast<const double2*>(mu[p]);
        a.y[p] = reinterpret_cast<const double2*>(y[p]);
        a.w[p] = w[p];
    }
    a.nb = h->nb.p;
    a.part = A.part.p;
    a.gs = A.gs.p;
    a.mp = h->Mp;
    const int nt = h->term_slots, nb = h->ntiles;
    switch (h->N) {
    case 1: launch_contract_n<1>(nt, a, h->terms.p, nb, h->stream); break;
    case 2: launch_contract_n<2>(nt, a, h->terms.p, nb, h->stream); break;
    case 3: launch_contract_n<3>(nt, a, h->terms.p, nb, h->stream); break;
    case 4: launch_contract_n<4>(nt, a, h->terms.p, nb, h->stream); break;
    case 5: launch_contract_n<5>(nt, a, h->terms.p, nb, h->stream); break;
    case 6: launch_contract_n<6>(nt, a, h->terms.p, nb, h->stream); break;
    case 7: launch_contract_n<7>(nt, a, h->terms.p, nb, h->stream); break;
    default: launch_contract_n<8>(nt, a, h->terms.p, nb, h->stream); break;
    }
    HIPC(hipGetLastError());
}

AdjointState& adjoint_state(heom_handle* h)
{
    if (!h->adj)
        h->adj.reset(new AdjointState());
    return *h->adj;
}

void check_adjoint_supported(heom_handle* h)
{
    if (h->fp32())
        raise(HEOM_E_UNSUPPORTED, "the adjoint needs FP64 storage");
    // its bath gradients are sums over the instances of one set of bath parameters
    if (!h->uniform)
        raise(HEOM_E_UNSUPPORTED,
              "the adjoint needs a batch whose instances share states, modes, coupling and bath prefactors");
    if (h->nel > 64)
        raise(HEOM_E_UNSUPPORTED, "the adjoint needs N*N <= 64");
    // the transpose of a filtered hierarchy masks destinations instead of sources;
    // the adjoint kernels implement the unfiltered operator, so every ADO that is in the
    // layout must be unmasked (ADOs removed by heom_compact are not in it)
    std::vector<uint8_t> mask(h->M);
    std::vector<int32_t> q2u(h->Mp);
    HIPC(hipMemcpyAsync(mask.data(), h->mask.p, h->M, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipMemcpyAsync(q2u.data(), h->q2u.p, q2u.size() * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    for (int64_t q = 0; q < h->Mp; ++q)
        if (q2u[q] >= 0 && mask[q2u[q]] == 0)
            raise(HEOM_E_UNSUPPORTED,
                  "the adjoint needs an unfiltered hierarchy (mask ADOs only together with heom_compact)");
}

// element constants of L^dag: -H^dag / hbar, conj(gam), conj(pf_same_sum), and per term
// -plus_c and conj(minus_c) (the kernel exchanges the edges they multiply)
void prepare_operator(heom_handle* h)
{
    AdjointState& A = adjoint_state(h);
    if (A.ops_ready)
        return;
    const int n = h->count * h->nel;
    A.hs.alloc(n);
    hipLaunchKernelGGL(adjoint_hamiltonian, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->hs.p, A.hs.p,
                       h->count, h->N);
    A.gam.alloc(h->nel);
    hipLaunchKernelGGL(conj_copy, dim3(blocks_for(h->nel, 256)), dim3(256), 0, h->stream, h->gam.p, A.gam.p,
                       (int64_t)h->nel);
    A.pfss.alloc(h->Mp);
    hipLaunchKernelGGL(conj_copy, dim3(blocks_for(h->Mp, 256)), dim3(256), 0, h->stream, h->pfss.p, A.pfss.p,
                       h->Mp);
    HIPC(hipGetLastError());
    std::vector<Term> t((size_t)h->nel * h->term_slots);
    HIPC(hipMemcpyAsync(t.data(), h->terms.p, t.size() * sizeof(Term), hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    for (Term& x : t) {
        x.plus_c = -x.plus_c;
        x.minus_c.y = -x.minus_c.y;
    }
    A.terms.alloc(t.size());
    HIPC(hipMemcpyAsync(A.terms.p, t.data(), t.size() * sizeof(Term), hipMemcpyHostToDevice, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    A.real_h = h->real_h;
    A.ops_ready = true;
}

size_t plane_bytes(const heom_handle* h) { return h->y.bytes(); }

unsigned char* slot(DevBuf<unsigned char>& b, size_t bytes, int64_t k) { return b.p + (size_t)k * bytes; }

// dst planes = one RK4 step of src (FP64, explicit buffers); y2 / y3 / y4 receive the
// stage vectors Y2, Y3, Y4 (no fourth stage for dst = null)
void rk4_step_stages(heom_handle* h, const unsigned char* src, unsigned char* y2, unsigned char* y3,
                     unsigned char* y4, unsigned char* dst, double dt)
{
    launch_stage_buffers(h, 1, src, src, h->pbuf.p, y2, 0.5 * dt, false, false);
    launch_stage_buffers(h, 2, y2, src, h->pbuf.p, y3, 0.5 * dt, false, false);
    launch_stage_buffers(h, 3, y3, src, h->pbuf.p, y4, dt, false, false);
    if (dst)
        launch_stage_buffers(h, 4, y4, src, h->pbuf.p, dst, dt / 6.0, false, false);
}

// the same with the stage vectors in f2 / f3 / f4
void rk4_step_into(heom_handle* h, const unsigned char* src, unsigned char* dst, double dt)
{
    AdjointState& A = *h->adj;
    rk4_step_stages(h, src, A.f2.p, A.f3.p, A.f4.p, dst, dt);
}

// one backward RK4 step: lambda = dJ/dy_{n+1} becomes dJ/dy_n, with the step's gradient
// contributions added to the tile sums; yn, y2, y3, y4 = the step's forward stage vectors
void adjoint_step(heom_handle* h, const unsigned char* yn, const unsigned char* y2, const unsigned char* y3,
                  const unsigned char* y4, double dt, const double2* top_src)
{
    AdjointState& A = *h->adj;
    // adjoint stages A2 (a), A3 (b), A4 (c) from A1 = lambda, then all four stage pairs in
    // one contraction, then dJ/dy_n into lambda, with the source of the observation at
    // step n (top_src, or null) added on the top ADOs in the same launch
    launch_stage_buffers(h, 1, A.lam.p, A.lam.p, h->pbuf.p, A.a.p, 0.5 * dt, false, true);
    launch_stage_buffers(h, 2, A.a.p, A.lam.p, h->pbuf.p, A.b.p, 0.5 * dt, false, true);
    launch_stage_buffers(h, 3, A.b.p, A.lam.p, h->pbuf.p, A.c.p, dt, false, true);
    const void* mus[kPairs] = {A.lam.p, A.a.p, A.b.p, A.c.p};
    const void* ys[kPairs] = {y4, y3, y2, yn};
    launch_contract(h, mus, ys, dt);
    launch_stage_buffers(h, 4, A.c.p, A.lam.p, h->pbuf.p, A.lam.p, dt / 6.0, false, true, top_src);
}

// the adjoint source of observation row `row` (null for a row whose weights are all zero)
const double2* source_of(heom_handle* h, const std::vector<char>& nonzero, int64_t row)
{
    return nonzero[row] ? h->adj->src.p + (size_t)row * h->count * h->nel : nullptr;
}

// observation row `row` of the forward sweep from the state planes y
void observe_row(heom_handle* h, const unsigned char* y, const double2* ops, double* pops, double2* traces,
                 int64_t row)
{
    hipLaunchKernelGGL(observe_planes, dim3(h->count), dim3(64), 0, h->stream, reinterpret_cast<const double2*>(y),
                       h->q_top.p, h->Mp, ops, pops + row * h->count * h->N, traces + row * h->count, h->N);
    HIPC(hipGetLastError());
}

// whole-trajectory mode: the [last][4] stage planes fit in half the free device memory
// (the buffer already held counts as free) and HEOM_ADJOINT_RECOMPUTE is not 1
bool keep_trajectory(heom_handle* h, int64_t last, size_t pb)
{
    const char* env = std::getenv("HEOM_ADJOINT_RECOMPUTE");
    if (last < 1 || (env && std::strcmp(env, "1") == 0))
        return false;
    size_t free = 0, total = 0;
    HIPC(hipMemGetInfo(&free, &total));
    const double need = 4.0 * (double)pb * (double)last;
    return need <= 0.5 * ((double)free + (double)h->adj->traj.bytes());
}

// lambda_top += the adjoint source of observation row `row` (rows with all weights zero
// are skipped on the host)
void add_source(heom_handle* h, const std::vector<char>& nonzero, int64_t row)
{
    if (!nonzero[row])
        return;
    AdjointState& A = *h->adj;
    const int n = h->count * h->nel;
    hipLaunchKernelGGL(add_tops, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, A.lam.as<double2>(),
                       A.src.p + (size_t)row * n, h->q_top.p, h->count, h->Mp, h->nel);
    HIPC(hipGetLastError());
}

}  // namespace

namespace heom_rt {

void adjoint_invalidate(heom_handle* h)
{
    if (h->adj) {
        h->adj->ops_ready = false;
        h->adj->recorded = false;
    }
}

}  // namespace heom_rt

extern "C" {

int heom_rhs_adjoint(heom_handle* h, double* out)
{
    return guard([&] {
        need(h);
        reject_tss(h, "the adjoint");
        if (!out)
            raise(HEOM_E_INVALID, "out is NULL");
        check_adjoint_supported(h);
        prepare_operator(h);
        launch_stage_buffers(h, 0, h->y.p, h->y.p, h->pbuf.p, h->ya.p, 0.0, false, true);
        download_state(h, h->ya.p, nullptr, out);
    });
}

int heom_adjoint_forward(heom_handle* h, double dt, int64_t nsteps, int64_t observe_steps, int64_t checkpoint_steps,
                         const double* ops, double* observed)
{
    return guard([&] {
        need(h);
        reject_tss(h, "the adjoint");
        if (!ops || !observed)
            raise(HEOM_E_INVALID, "ops and observed must not be NULL");
        if (nsteps < 0 || observe_steps < 1 || checkpoint_steps < 1)
            raise(HEOM_E_INVALID, "nsteps must be >= 0, observe_steps and checkpoint_steps >= 1");
        if (!(dt > 0.0))
            raise(HEOM_E_INVALID, "step size must be > 0");
        check_adjoint_supported(h);
        AdjointState& A = adjoint_state(h);
        A.recorded = false;
        const int64_t rows = nsteps / observe_steps + 1;
        const int64_t last = (rows - 1) * observe_steps;
        const int64_t ncp = last / checkpoint_steps + 1;
        const size_t pb = plane_bytes(h);
        const bool keep = keep_trajectory(h, last, pb);
        if (keep) {
            A.cps.free();
            const size_t need = 4 * pb * (size_t)last;
            if (A.traj.bytes() != need) {
                A.traj.alloc(need);
                // padding slots of the planes must read as zero; the stage kernels keep them so
                HIPC(hipMemsetAsync(A.traj.p, 0, need, h->stream));
            }
        } else {
            A.traj.free();
            A.cps.alloc(pb * (size_t)ncp);
        }
        const int n = h->count * h->nel;
        A.op.assign(ops, ops + 2 * (size_t)n);
        DevBuf<double2> opsd, trec;
        DevBuf<double> prec;
        opsd.alloc(n);
        trec.alloc((size_t)(rows * h->count));
        prec.alloc((size_t)(rows * h->count * h->N));
        HIPC(hipMemcpyAsync(opsd.p, ops, n * sizeof(double2), hipMemcpyHostToDevice, h->stream));
        resolve_timing(h);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPC(hipEventCreate(&e0));
        HIPC(hipEventCreate(&e1));
        try {
            HIPC(hipEventRecord(e0, h->stream));
            observe_row(h, h->y.p, opsd.p, prec.p, trec.p, 0);
            if (keep) {
                // step n reads slot 4n and writes its stage vectors to 4n + 1 .. 4n + 3 and
                // y_{n+1} to slot 4n + 4 (the last one to the state)
                HIPC(hipMemcpyAsync(A.traj.p, h->y.p, pb, hipMemcpyDeviceToDevice, h->stream));
                for (int64_t n = 0; n < last; ++n) {
                    unsigned char* nxt = n + 1 < last ? slot(A.traj, pb, 4 * (n + 1)) : h->y.p;
                    rk4_step_stages(h, slot(A.traj, pb, 4 * n), slot(A.traj, pb, 4 * n + 1),
                                    slot(A.traj, pb, 4 * n + 2), slot(A.traj, pb, 4 * n + 3), nxt, dt);
                    if ((n + 1) % observe_steps == 0)
                        observe_row(h, nxt, opsd.p, prec.p, trec.p, (n + 1) / observe_steps);
                }
            } else {
                HIPC(hipMemcpyAsync(A.cps.p, h->y.p, pb, hipMemcpyDeviceToDevice, h->stream));
            }
            int64_t done = keep ? last : 0;
            while (done < last) {
                // advance to the next observation or checkpoint, whichever comes first
                const int64_t to_obs = observe_steps - done % observe_steps;
                const int64_t to_cp = checkpoint_steps - done % checkpoint_steps;
                const int64_t n_adv = std::min(std::min(to_obs, to_cp), last - done);
                rk4_steps(h, dt, n_adv);
                done += n_adv;
                if (done % observe_steps == 0)
                    observe_row(h, h->y.p, opsd.p, prec.p, trec.p, done / observe_steps);
                if (done % checkpoint_steps == 0)
                    HIPC(hipMemcpyAsync(slot(A.cps, pb, done / checkpoint_steps), h->y.p, pb,
                                        hipMemcpyDeviceToDevice, h->stream));
            }
            if (nsteps > last)
                rk4_steps(h, dt, nsteps - last);
            HIPC(hipEventRecord(e1, h->stream));
            HIPC(hipMemcpyAsync(observed, trec.p, trec.bytes(), hipMemcpyDeviceToHost, h->stream));
            HIPC(hipStreamSynchronize(h->stream));
            float ms = 0.f;
            HIPC(hipEventElapsedTime(&ms, e0, e1));
            h->timing.last_call_seconds = ms * 1e-3;
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        A.dt = dt;
        A.nsteps = nsteps;
        A.observe_steps = observe_steps;
        A.checkpoint_steps = checkpoint_steps;
        A.rows = rows;
        A.last = last;
        A.ncp = ncp;
        A.keep = keep;
        A.recorded = true;
    });
}

int heom_adjoint_backward(heom_handle* h, const double* weights, heom_gradients* g)
{
    return guard([&] {
        need(h);
        if (!weights || !g)
            raise(HEOM_E_INVALID, "weights and grads must not be NULL");
        if (!h->adj || !h->adj->recorded)
            raise(HEOM_E_INVALID, "no forward sweep on record (heom_adjoint_forward), or the operator changed since");
        check_adjoint_supported(h);
        prepare_operator(h);
        AdjointState& A = *h->adj;
        const int N = h->N, nel = h->nel, count = h->count, NT = h->term_slots;
        const int slots = nel * (2 + 2 * NT);
        const size_t pb = plane_bytes(h);
        const int64_t ck = A.checkpoint_steps;
        const double dt = A.dt;

        // adjoint sources conj(w_rk) O_k^dag, [rows][count][nel], resident on the device
        std::vector<double2> src((size_t)A.rows * count * nel);
        std::vector<char> nonzero(A.rows, 0);
        for (int64_t r = 0; r < A.rows; ++r)
            for (int k = 0; k < count; ++k) {
                const std::complex<double> w(weights[2 * (r * count + k)], weights[2 * (r * count + k) + 1]);
                if (w != 0.0)
                    nonzero[r] = 1;
                for (int i = 0; i < N; ++i)
                    for (int j = 0; j < N; ++j) {
                        const size_t o = ((size_t)k * nel + j * N + i) * 2;
                        const std::complex<double> v = std::conj(w * std::complex<double>(A.op[o], A.op[o + 1]));
                        src[((size_t)r * count + k) * nel + i * N + j] = make_double2(v.real(), v.imag());
                    }
            }

        if (A.keep) {
            for (auto* b : {&A.seg, &A.f2, &A.f3, &A.f4})
                b->free();
        } else {
            A.seg.alloc(pb * (size_t)std::min<int64_t>(ck, std::max<int64_t>(A.last, 1)));
            for (auto* b : {&A.f2, &A.f3, &A.f4})
                b->alloc(pb);
        }
        for (auto* b : {&A.lam, &A.a, &A.b, &A.c})
            b->alloc(pb);
        A.part.alloc((size_t)h->ntiles * kPairs * slots);
        A.gs.alloc((size_t)kPairs * h->Mp);
        A.src.alloc(src.size());
        HIPC(hipMemcpyAsync(A.src.p, src.data(), src.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
        // padding slots of the state planes must read as zero
        for (auto* b : {&A.f2, &A.f3, &A.f4, &A.lam, &A.a, &A.b, &A.c, &A.seg})
            if (b->p)
                HIPC(hipMemsetAsync(b->p, 0, b->bytes(), h->stream));
        HIPC(hipMemsetAsync(A.part.p, 0, A.part.bytes(), h->stream));
        HIPC(hipMemsetAsync(A.gs.p, 0, A.gs.bytes(), h->stream));

        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPC(hipEventCreate(&e0));
        HIPC(hipEventCreate(&e1));
        try {
            HIPC(hipEventRecord(e0, h->stream));
            add_source(h, nonzero, A.rows - 1);
            if (A.keep)
                for (int64_t n = A.last - 1; n >= 0; --n)
                    adjoint_step(h, slot(A.traj, pb, 4 * n), slot(A.traj, pb, 4 * n + 1), slot(A.traj, pb, 4 * n + 2),
                                 slot(A.traj, pb, 4 * n + 3), dt,
                                 n % A.observe_steps == 0 ? source_of(h, nonzero, n / A.observe_steps) : nullptr);
            for (int64_t c = A.keep ? -1 : A.last > 0 ? (A.last - 1) / ck : -1; c >= 0; --c) {
                const int64_t start = c * ck, end = std::min(start + ck, A.last);
                HIPC(hipMemcpyAsync(A.seg.p, slot(A.cps, pb, c), pb, hipMemcpyDeviceToDevice, h->stream));
                for (int64_t n = start + 1; n < end; ++n)
                    rk4_step_into(h, slot(A.seg, pb, n - start - 1), slot(A.seg, pb, n - start), dt);
                for (int64_t n = end - 1; n >= start; --n) {
                    unsigned char* yn = slot(A.seg, pb, n - start);
                    rk4_step_into(h, yn, nullptr, dt);  // Y2, Y3, Y4 of step n
                    adjoint_step(h, yn, A.f2.p, A.f3.p, A.f4.p, dt,
                                 n % A.observe_steps == 0 ? source_of(h, nonzero, n / A.observe_steps) : nullptr);
                }
            }
            HIPC(hipEventRecord(e1, h->stream));
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }

        // reduce the tiles per instance
        DevBuf<double2> red;
        red.alloc((size_t)count * slots);
        hipLaunchKernelGGL(sum_tiles, dim3(blocks_for((int64_t)count * slots, 256)), dim3(256), 0, h->stream,
                           A.part.p, count > 1 ? h->tile_inst.p : nullptr, red.p, h->ntiles * kPairs, kPairs,
                           slots, count);
        HIPC(hipGetLastError());
        std::vector<double2> sums((size_t)count * slots), gs4((size_t)kPairs * h->Mp), gs(h->Mp);
        std::vector<int32_t> q2u(h->Mp);
        std::vector<Term> terms((size_t)nel * NT);
        HIPC(hipMemcpyAsync(sums.data(), red.p, red.bytes(), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipMemcpyAsync(gs4.data(), A.gs.p, A.gs.bytes(), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipMemcpyAsync(q2u.data(), h->q2u.p, q2u.size() * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipMemcpyAsync(terms.data(), h->terms.p, terms.size() * sizeof(Term), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        h->timing.last_call_seconds = ms * 1e-3;
        for (int64_t q = 0; q < h->Mp; ++q) {
            double2 t = make_double2(0.0, 0.0);
            for (int p = 0; p < kPairs; ++p) {
                t.x += gs4[(size_t)p * h->Mp + q].x;
                t.y += gs4[(size_t)p * h->Mp + q].y;
            }
            gs[q] = t;
        }

        using cd = std::complex<double>;
        auto at = [&](int k, int e, int s) {
            const double2 v = sums[(size_t)k * slots + e * (2 + 2 * NT) + s];
            return cd(v.x, v.y);
        };
        // Hamiltonian: delta J = Re sum dH_ij (i/hbar) T_ij, so dJ/dRe + i dJ/dIm = conj(i T / hbar)
        if (g->hamiltonian)
            for (int k = 0; k < count; ++k)
                for (int e = 0; e < nel; ++e) {
                    const cd v = std::conj(cd(0.0, 1.0) * at(k, e, 0) / kHbar);
                    g->hamiltonian[2 * ((size_t)k * nel + e)] = v.real();
                    g->hamiltonian[2 * ((size_t)k * nel + e) + 1] = v.imag();
                }
        // pf_same_sum: S_n multiplies sigma_n, delta J = Re dS_n sum conj(mu) Y
        if (g->pf_same_sum) {
            std::fill(g->pf_same_sum, g->pf_same_sum + 2 * h->M, 0.0);
            for (int64_t q = 0; q < h->Mp; ++q)
                if (q2u[q] >= 0) {
                    g->pf_same_sum[2 * (size_t)q2u[q]] = gs[q].x;
                    g->pf_same_sum[2 * (size_t)q2u[q] + 1] = -gs[q].y;
                }
        }
        // bath arrays, summed over the instances: gam_ij = -sum Delta_b over the coupling
        // entries of i and j (i != j); plus_c = (ci - cj) sqrt(|c|/hbar) and
        // minus_c = ((ci - cj) i c / hbar + [k = 0] (ci + cj) Phi_b) sqrt(hbar / |c|)
        std::vector<cd> g_same(h->B), g_circ(h->B), g_cbk((size_t)h->B * h->K);
        std::vector<double> d_cr((size_t)h->B * h->K, 0.0), d_ci((size_t)h->B * h->K, 0.0), d_fr(h->B, 0.0),
            d_fi(h->B, 0.0);
        for (int e = 0; e < nel; ++e) {
            const int i = e / N, j = e % N;
            cd gd = 0.0;
            for (int k = 0; k < count; ++k)
                gd += at(k, e, 1);
            if (i != j)
                for (int b = 0; b < h->B; ++b) {
                    const int cnt = count_bath(h, i, b) + count_bath(h, j, b);
                    g_same[b] -= std::conj(gd) * (double)cnt;
                }
            const int nterms = terms[(size_t)e * NT].info & 0xff;
            for (int t = 0; t < nterms; ++t) {
                const Term& tm = terms[(size_t)e * NT + t];
                cd gp = 0.0, gm = 0.0;
                for (int k = 0; k < count; ++k) {
                    gp += at(k, e, 2 + t);
                    gm += at(k, e, 2 + NT + t);
                }
                const int m = tm.mode, b = m / h->K, kk = m % h->K;
                const int ci = count_bath(h, i, b), cj = count_bath(h, j, b);
                const double cr = h->cbk[2 * m], cim = h->cbk[2 * m + 1];
                const double a = std::hypot(cr, cim);
                const double s = std::sqrt(kHbar / a);
                const double D = ci - cj, Sg = ci + cj;
                cd X = D * cd(-cim / kHbar, cr / kHbar);
                if (kk == 0)
                    X += Sg * cd(h->pf_circ[2 * b], h->pf_circ[2 * b + 1]);
                const double dp = D / (2.0 * std::sqrt(a * kHbar)) / a;  // times cr or cim
                const double ds = -s / (2.0 * a * a);                     // times cr or cim
                const cd dm_cr = D * cd(0.0, 1.0 / kHbar) * s + X * ds * cr;
                const cd dm_ci = -D / kHbar * s + X * ds * cim;
                d_cr[m] += gp.real() * dp * cr + (gm * dm_cr).real();
                d_ci[m] += gp.real() * dp * cim + (gm * dm_ci).real();
                if (kk == 0) {
                    d_fr[b] += (gm * (Sg * s)).real();
                    d_fi[b] += (gm * cd(0.0, Sg * s)).real();
                }
            }
        }
        if (g->pf_same)
            for (int b = 0; b < h->B; ++b) {
                g->pf_same[2 * b] = g_same[b].real();
                g->pf_same[2 * b + 1] = g_same[b].imag();
            }
        if (g->pf_circ)
            for (int b = 0; b < h->B; ++b) {
                g->pf_circ[2 * b] = d_fr[b];
                g->pf_circ[2 * b + 1] = d_fi[b];
            }
        if (g->cbk)
            for (int m = 0; m < h->B * h->K; ++m) {
                g->cbk[2 * m] = d_cr[m];
                g->cbk[2 * m + 1] = d_ci[m];
            }
        if (g->state)
            download_state(h, A.lam.p, nullptr, g->state);
    });
}

}  // extern "C"
