// heom_hip.hip -- the C ABI of include/heom_hip.h: handle lifetime, the re-layout of
// the reference instances into the device layout (heom_device.hpp), the RK4 driver,
// the FP64 re-evaluation of the tier-0/1 ADOs of FP32-storage handles, and the small
// layout / reduction / hierarchy-arithmetic kernels.
//
// A handle holds a batch of one or more hierarchies with the same states and baths,
// concatenated into one device hierarchy (instance k owns the tiles from its tile
// offset on); everything per instance (top ADO, Hamiltonian, mmult_left matrix,
// recorded tops) is indexed by instance, everything per ADO (state, mask, norms) is the
// instances' reference-layout arrays one after the other.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <new>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "handle.hpp"
#include "heom_device.hpp"
#include "heom_hip.h"

using namespace heom_hip;
using namespace heom_rt;

// size of the device scratch of heom_propagate_observe_noalloc (one array per GPU)
constexpr size_t kObserveScratchBytes = HEOM_OBSERVE_SCRATCH_BYTES;

thread_local std::string heom_rt::g_error;

heom_handle::heom_handle() = default;
heom_handle::~heom_handle() { release(); }

namespace {


// ------------------------------------------------------------------------------------
// layout, reduction and hierarchy-arithmetic kernels (C = double2 or float2 planes)
// ------------------------------------------------------------------------------------

__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2(v.x, v.y); }
template <typename C>
__device__ __forceinline__ C from_d2(double2 v);
template <>
__device__ __forceinline__ double2 from_d2<double2>(double2 v)
{
    return v;
}
template <>
__device__ __forceinline__ float2 from_d2<float2>(double2 v)
{
    return make_float2((float)v.x, (float)v.y);
}

// Element e of ADO q in the planes: c64_index (heom_device.hpp); complex128 planes are
// never paired.
template <typename C>
__device__ __forceinline__ int64_t plane_index(bool paired, int N, int e, int64_t q, int64_t mp)
{
    return c64_index(sizeof(C) == sizeof(float2) && paired, N, e, q, mp);
}

// dst[e, q] = src[u(q)][e]   (reference AoS -> internal SoA; padding -> 0)
template <typename C>
__global__ void aos_to_soa(const double2* __restrict__ src, C* __restrict__ dst, const int32_t* __restrict__ q2u,
                           int64_t mp, int N, bool paired)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= mp)
        return;
    const int32_t u = q2u[q];
    const int nel = N * N;
    for (int e = 0; e < nel; ++e)
        dst[plane_index<C>(paired, N, e, q, mp)] =
            from_d2<C>(u >= 0 ? src[(int64_t)u * nel + e] : make_double2(0.0, 0.0));
}

// dst[u(q)][e] = src[e, q]   (internal SoA -> reference AoS)
template <typename C>
__global__ void soa_to_aos(const C* __restrict__ src, double2* __restrict__ dst, const int32_t* __restrict__ q2u,
                           int64_t mp, int N, bool paired)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= mp)
        return;
    const int32_t u = q2u[q];
    if (u < 0)
        return;
    const int nel = N * N;
    for (int e = 0; e < nel; ++e)
        dst[(int64_t)u * nel + e] = to_d2(src[plane_index<C>(paired, N, e, q, mp)]);
}

// dst[k][e] = src[e, q[k]]: one ADO per instance (the tops), thread = (k, e)
template <typename C>
__global__ void gather_ados(const C* __restrict__ src, double2* __restrict__ dst, const int64_t* __restrict__ qs,
                            int count, int64_t mp, int N, bool paired)
{
    const int nel = N * N;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * nel)
        return;
    const int k = t / nel, e = t % nel;
    dst[t] = to_d2(src[plane_index<C>(paired, N, e, qs[k], mp)]);
}

template <typename C>
__global__ void scatter_ados(const double2* __restrict__ src, C* __restrict__ dst, const int64_t* __restrict__ qs,
                             int count, int64_t mp, int N, bool paired)
{
    const int nel = N * N;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * nel)
        return;
    const int k = t / nel, e = t % nel;
    dst[plane_index<C>(paired, N, e, qs[k], mp)] = from_d2<C>(src[t]);
}

// rows of the FP64 shadow: dst[k][e] = shadow[idx[k]][e]
__global__ void gather_rows(const double2* __restrict__ shadow, double2* __restrict__ dst,
                            const int32_t* __restrict__ idx, int count, int nel)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * nel)
        return;
    const int k = t / nel, e = t % nel;
    dst[t] = shadow[(int64_t)idx[k] * nel + e];
}

__global__ void scatter_rows(const double2* __restrict__ src, double2* __restrict__ shadow,
                             const int32_t* __restrict__ idx, int count, int nel)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * nel)
        return;
    const int k = t / nel, e = t % nel;
    shadow[(int64_t)idx[k] * nel + e] = src[t];
}

// Populations Re sigma_0[i][i] (dm-heom/include/heom/matrix_diagonal_observer.hpp:28-34)
// and, with ops, traces Tr(ops_k sigma_0) (matrix_trace_observer.hpp:139-158) of every
// instance's top ADO: one wavefront per instance, lane = element, a shuffle reduction
// over the N*N products.
__global__ __launch_bounds__(64) void observe_tops(const double2* __restrict__ tops, const double2* __restrict__ ops,
                                                   double* __restrict__ pops, double2* __restrict__ traces, int N)
{
    const int k = blockIdx.x, lane = threadIdx.x, nel = N * N;
    double re = 0.0, im = 0.0;
    for (int e = lane; e < nel; e += 64) {
        const int i = e / N, j = e % N;
        const double2 s = tops[(int64_t)k * nel + e];
        if (i == j)
            pops[(int64_t)k * N + i] = s.x;
        if (ops) {
            // Tr(D sigma) = sum_ij D_ji sigma_ij
            const double2 d = ops[(int64_t)k * nel + j * N + i];
            re += fma(d.x, s.x, -d.y * s.y);
            im += fma(d.x, s.y, d.y * s.x);
        }
    }
    if (ops) {
        for (int off = 32; off > 0; off >>= 1) {
            re += __shfl_xor(re, off);
            im += __shfl_xor(im, off);
        }
        if (lane == 0)
            traces[k] = make_double2(re, im);
    }
}

// per-ADO squared Frobenius norm in uid order; threshold masking as in
// dm-heom/cl/heom/hierarchy_norm.cl:41-44 (mask in uid order)
// hierarchy_norm (dm-heom/cl/heom/hierarchy_norm.cl:15-45): one workgroup per 64-ADO tile,
// lane = ADO (so every load is a coalesced plane segment), the tile's physical planes
// dealt round-robin to kNormWaves wavefronts, whose partial sums wavefront 0 adds in a
// fixed order.  A 16-byte slot of the paired complex64 layout holds two elements; the
// squares of all its components are summed.  Splitting the planes keeps small hierarchies
// of large matrices (20,475 ADOs of 24 x 24 at 576 planes) from running one long serial
// loop per thread on a few waves per CU.
constexpr int kNormWaves = 8;

template <typename V>
__global__ __launch_bounds__(64 * kNormWaves) void hierarchy_norm_kernel(const V* __restrict__ planes,
                                                                        int nplanes, const int32_t* __restrict__ q2u,
                                                                        int64_t mp, double threshold,
                                                                        double* __restrict__ norms,
                                                                        uint8_t* __restrict__ mask)
{
    __shared__ double part[kNormWaves][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t q = (int64_t)blockIdx.x * 64 + lane;
    double acc = 0.0;
    if (q < mp) {
        for (int d = w; d < nplanes; d += kNormWaves) {
            const V v = planes[(int64_t)d * mp + q];
            if constexpr (sizeof(V) == 16 && std::is_same<V, double2>::value) {
                acc = fma(v.x, v.x, acc);
                acc = fma(v.y, v.y, acc);
            } else if constexpr (sizeof(V) == 16) {  // float4: two complex64 elements
                acc = fma((double)v.x, (double)v.x, acc);
                acc = fma((double)v.y, (double)v.y, acc);
                acc = fma((double)v.z, (double)v.z, acc);
                acc = fma((double)v.w, (double)v.w, acc);
            } else {  // float2
                acc = fma((double)v.x, (double)v.x, acc);
                acc = fma((double)v.y, (double)v.y, acc);
            }
        }
    }
    part[w][lane] = acc;
    __syncthreads();
    if (w != 0 || q >= mp)
        return;
    const int32_t u = q2u[q];
    if (u < 0)
        return;
    double sum = part[0][lane];
#pragma unroll
    for (int k = 1; k < kNormWaves; ++k)
        sum += part[k][lane];
    norms[u] = sum;
    if (sum - threshold <= 0.0 && mask[u] != 2)
        mask[u] = 0;
}

// Fold the uid-order mask into the packed neighbour table: an edge whose target is
// masked reads the zero slot (heom_ode.cl drops such a term with its mask checks).
__global__ void fold_mask(const int2* __restrict__ raw, int2* __restrict__ folded, const uint8_t* __restrict__ mask,
                          const int32_t* __restrict__ q2u, int64_t count, int32_t zero_slot)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count)
        return;
    const int2 r = raw[t];
    const int32_t p = (int32_t)((uint32_t)r.x >> 6);
    const int32_t n = r.x & 63;
    const int32_t pf = (p != zero_slot && mask[q2u[p]] != 0) ? p : zero_slot;
    const int32_t mf = (r.y != zero_slot && mask[q2u[r.y]] != 0) ? r.y : zero_slot;
    folded[t] = make_int2((int32_t)(((uint32_t)pf << 6) | (uint32_t)n), mf);
}

// new planes from old ones: dst[e * mp_new + q'] = src[e * mp_old + n2o[q']] (n2o < 0: zero)
__global__ void remap_planes(const double2* __restrict__ src, double2* __restrict__ dst,
                             const int32_t* __restrict__ n2o, int64_t mp_old, int64_t mp_new, int nel)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nel * mp_new)
        return;
    const int e = (int)(t / mp_new);
    const int64_t qn = t % mp_new;
    const int32_t qo = n2o[qn];
    dst[t] = qo >= 0 ? src[(int64_t)e * mp_old + qo] : make_double2(0.0, 0.0);
}

// sigma_q <- Mx_k sigma_q for every ADO q of instance k: wave = row i, lane = ADO,
// tile in LDS, arithmetic in FP64
template <typename C, int N>
__global__ __launch_bounds__(64 * N) void mmult_left_kernel(C* __restrict__ y, const double2* __restrict__ mx,
                                                            const int32_t* __restrict__ tile_inst, int64_t mp,
                                                            bool paired)
{
    __shared__ double2 t[N * N * kTile];
    const int lane = threadIdx.x & 63;
    const int i = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t q = (int64_t)blockIdx.x * kTile + lane;
    const double2* m_k = mx + (tile_inst ? (int64_t)tile_inst[blockIdx.x] * N * N : 0);
#pragma unroll
    for (int j = 0; j < N; ++j)
        t[(i * N + j) * kTile + lane] = to_d2(y[plane_index<C>(paired, N, i * N + j, q, mp)]);
    __syncthreads();
    double2 out[N];
#pragma unroll
    for (int j = 0; j < N; ++j)
        out[j] = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double2 m = m_k[i * N + k];
#pragma unroll
        for (int j = 0; j < N; ++j)
            out[j] = cfma(m, t[(k * N + j) * kTile + lane], out[j]);
    }
#pragma unroll
    for (int j = 0; j < N; ++j)
        y[plane_index<C>(paired, N, i * N + j, q, mp)] = from_d2<C>(out[j]);
}

// mmult_left for N > 8, out of place: out[e][q] = sum_k mx_ik y[k*N+j][q], thread = (e, q)
template <typename C>
__global__ void mmult_left_wide(const C* __restrict__ y, C* __restrict__ out, const double2* __restrict__ mx,
                                const int32_t* __restrict__ tile_inst, int64_t mp, int N)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nel = N * N;
    if (t >= (int64_t)nel * mp)
        return;
    const int e = (int)(t / mp);
    const int64_t q = t % mp;
    const int i = e / N, j = e % N;
    const double2* m = mx + (tile_inst ? (int64_t)tile_inst[q / kTile] * nel : 0);
    double2 acc = make_double2(0.0, 0.0);
    for (int k = 0; k < N; ++k)
        acc = cfma(m[i * N + k], to_d2(y[(int64_t)(k * N + j) * mp + q]), acc);
    out[t] = from_d2<C>(acc);
}

// mmult_left on the FP64 shadow rows: thread = (row r, element i*N+j)
__global__ void mmult_left_rows(const double2* __restrict__ in, double2* __restrict__ out,
                                const double2* __restrict__ mx, const int32_t* __restrict__ row_inst, int rows, int N)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nel = N * N;
    if (t >= rows * nel)
        return;
    const int r = t / nel, e = t % nel, i = e / N, j = e % N;
    const double2* m = mx + (int64_t)row_inst[r] * nel;
    double2 acc = make_double2(0.0, 0.0);
    for (int k = 0; k < N; ++k)
        acc = cfma(m[i * N + k], in[(int64_t)r * nel + k * N + j], acc);
    out[t] = acc;
}

// ------------------------------------------------------------------------------------
// FP32 storage: the tier-0/1 ADOs (tuple sum <= 1) in FP64
// ------------------------------------------------------------------------------------
//
// With complex64 planes, the stage kernels evaluate the RHS in FP32 for every ADO.
// For the top ADO and the W tier-1 ADOs this kernel then re-evaluates the RHS in FP64
// and carries their RK4 state (y_n, the stage values and P) in FP64 shadows, writing
// the rounded values into the planes for the neighbours to read.  Populations and
// traces come from the FP64 top.
struct TierArgs {
    const float2* yin;    // stage input planes (neighbours that are not tier-0/1)
    float2* yout;         // stage output planes (overwritten at the tier-0/1 ADOs)
    float2* pout;         // P planes (stage 2)
    const int2* nb;       // folded neighbour table
    const int64_t* q;     // [S] internal index
    const int32_t* inst;  // [S] instance
    const int2* sp_nb;    // [S][W] tier-0/1 index of the plus / minus neighbour, or -1
    const double2* in64;  // [S][nel] stage input
    double2* out64;       // [S][nel] stage output (stage 4: y_{n+1}, the y_n shadow itself)
    const double2* y64;   // [S][nel] y_n
    double2* p64;         // [S][nel] P
    const double2* hs;    // [instances][nel] H / hbar
    const double2* gam;   // [nel]
    const Term* terms;    // [nel][nt]
    const double2* pfss;  // [S] pf_same_sum
    const int32_t* inst_class;  // [instances] bath class, or null (one class)
    int32_t paired;       // the complex64 planes are in the paired layout
    int64_t mp;
    int32_t S, N, W, nt, zero;
    double c1;
};

// One workgroup per tier-0/1 ADO s: its FP64 stage input, H / hbar of its instance and
// its W neighbour-table entries (folded mask) go to LDS with coalesced loads; then each
// thread takes elements e and issues the loads of all of the element's neighbour terms
// (the FP64 shadow for a tier-0/1 neighbour, the complex64 plane otherwise) before the
// commutator, which reads LDS, so the gathers' latency overlaps it.  The launch is a
// few dependent round trips per workgroup over tens of thousands of single-wavefront
// workgroups: with two terms per element the kernel is held to 64 VGPRs (the commutator
// loop unrolled twice, not more), so that 8 wavefronts per SIMD are resident, not 4; with
// more terms it takes the registers it needs (no spills).
template <int STAGE, int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NT <= 2 ? 8 : 1))) void tier_fixup(TierArgs a)
{
    extern __shared__ double2 tier_lds[];  // [nel] sigma, [nel] H / hbar, then [W] edges
    const int N = a.N, nel = N * N, W = a.W;
    // the W + 1 tier-0/1 ADOs of an instance run on one XCD, whose L2 then serves the
    // tier-2 neighbours they share
    const int s = xcd_tile(blockIdx.x, a.S, W + 1);
    double2* sig = tier_lds;
    double2* H = tier_lds + nel;
    int4* edge = reinterpret_cast<int4*>(tier_lds + 2 * nel);  // {plus slot, minus slot, n, tier pair}
    int2* spn = reinterpret_cast<int2*>(edge + W);             // tier-0/1 index of plus / minus, or -1
    const int64_t q = a.q[s];
    const double2* in = a.in64 + (int64_t)s * nel;
    const double2* hs = a.hs + (int64_t)a.inst[s] * nel;
    for (int e = threadIdx.x; e < nel; e += blockDim.x) {
        sig[e] = in[e];
        H[e] = hs[e];
    }
    for (int m = threadIdx.x; m < W; m += blockDim.x) {
        const int2 v = a.nb[(int64_t)m * a.mp + q];
        edge[m] = make_int4((int32_t)((uint32_t)v.x >> 6), v.y, v.x & 63, 0);
        spn[m] = a.sp_nb[(int64_t)s * W + m];
    }
    __syncthreads();
    const double2 S = a.pfss[s];
    const int cls = a.inst_class ? a.inst_class[a.inst[s]] : 0;
    const Term* terms = a.terms + (int64_t)cls * nel * NT;
    const double2* gam = a.gam + (int64_t)cls * nel;
    for (int e = threadIdx.x; e < nel; e += blockDim.x) {
        const int i = e / N, j = e - i * N;
        double2 xp[NT], xm[NT];
        int nv[NT];
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const int mode = terms[e * NT + k].mode;
            xp[k] = xm[k] = make_double2(0.0, 0.0);
            nv[k] = 0;
            if (mode < W) {
                const int4 ed = edge[mode];
                const int2 sn = spn[mode];
                nv[k] = ed.z;
                if (ed.x != a.zero)
                    xp[k] = sn.x >= 0 ? a.in64[(int64_t)sn.x * nel + e]
                                      : to_d2(a.yin[c64_index(a.paired, N, e, ed.x, a.mp)]);
                if (ed.y != a.zero)
                    xm[k] = sn.y >= 0 ? a.in64[(int64_t)sn.y * nel + e]
                                      : to_d2(a.yin[c64_index(a.paired, N, e, ed.y, a.mp)]);
            }
        }
        double2 cm = make_double2(0.0, 0.0);
#pragma unroll 2
        for (int k = 0; k < N; ++k) {
            cm = cfma(sig[i * N + k], H[k * N + j], cm);
            const double2 hik = H[i * N + k];
            cm = cfma(make_double2(-hik.x, -hik.y), sig[k * N + j], cm);
        }
        const double2 g = gam[e];
        double2 f = cfma(make_double2(S.x + g.x, S.y + g.y), sig[e], make_double2(-cm.y, cm.x));
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const Term tm = terms[e * NT + k];
            const double cp = tm.plus_c * sqrt((double)(nv[k] + 1));
            const double sm = sqrt((double)nv[k]);
            f.x = fma(-cp, xp[k].y, f.x);
            f.y = fma(cp, xp[k].x, f.y);
            f = cfma(tm.minus_c, make_double2(sm * xm[k].x, sm * xm[k].y), f);
        }
        const int64_t row = (int64_t)s * nel + e;
        const int64_t pl = c64_index(a.paired, N, e, q, a.mp);
        double2 r0, r1 = make_double2(0.0, 0.0);
        const double2 ep = STAGE == 4 ? a.p64[row] : (STAGE == 2 || STAGE == 3) ? a.y64[row] : make_double2(0.0, 0.0);
        stage_values<STAGE, double>(a.c1, sig[e], f, ep, r0, r1);
        a.out64[row] = r0;
        a.yout[pl] = make_float2((float)r0.x, (float)r0.y);
        if (STAGE == 2) {
            a.p64[row] = r1;
            a.pout[pl] = make_float2((float)r1.x, (float)r1.y);
        }
    }
}


}  // namespace


namespace heom_rt {

template <typename T>
void upload(DevBuf<T>& d, const std::vector<T>& h, hipStream_t s)
{
    d.alloc(h.size());
    HIPC(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

void resolve_timing(heom_handle* h)
{
    if (!h->timing_pending)
        return;
    HIPC(hipEventSynchronize(h->ev_stop));
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, h->ev_start, h->ev_stop));
    h->timing.last_call_seconds = ms * 1e-3;
    h->timing.rk4_seconds += ms * 1e-3;
    h->timing.rk4_steps += h->pending_steps;
    h->timing_pending = false;
    h->pending_steps = 0;
}

template <typename R>
StageArgsT<R> stage_args(heom_handle* h, const void* yin, const void* y0, void* pbuf, void* yout, double c1,
                         bool reverse, const c2<R>* pfss)
{
    StageArgsT<R> a{};
    a.yin = reinterpret_cast<const c2<R>*>(yin);
    a.y0 = reinterpret_cast<const c2<R>*>(y0);
    a.pbuf = reinterpret_cast<c2<R>*>(pbuf);
    a.yout = reinterpret_cast<c2<R>*>(yout);
    a.nb = h->nb.p;
    a.pfss = pfss;
    a.tile_inst = h->count > 1 ? h->tile_inst.p : nullptr;
    a.tile_plus = h->tile_plus.p;
    a.tile_class = h->uniform ? nullptr : h->tile_class.p;
    a.mp = h->Mp;
    a.ntiles = h->ntiles;
    a.zero_slot = h->zero_slot;
    a.table_units = (h->W + 2) / 2;
    a.xcd_chunk = h->xcd_chunk;
    a.reverse = reverse ? 1 : 0;
    a.split = h->pipelined ? 1 : h->tile_split;
    a.c1 = (R)c1;
    return a;
}

template <typename R>
void launch_stage_r(heom_handle* h, int stage, const unsigned char* yin, unsigned char* yout, double c1, bool reverse,
                    const ElementArgsT<R>& el, const c2<R>* pfss)
{
    StageArgsT<R> a{};
    a.yin = reinterpret_cast<const c2<R>*>(yin);
    a.y0 = h->y.as<c2<R>>();
    a.pbuf = h->pbuf.as<c2<R>>();
    a.yout = reinterpret_cast<c2<R>*>(yout);
    a.nb = h->nb.p;
    a.pfss = pfss;
    a.tile_inst = h->count > 1 ? h->tile_inst.p : nullptr;
    a.tile_plus = h->tile_plus.p;
    a.tile_class = h->uniform ? nullptr : h->tile_class.p;
    a.mp = h->Mp;
    a.ntiles = h->ntiles;
    a.zero_slot = h->zero_slot;
    a.table_units = (h->W + 2) / 2;
    a.xcd_chunk = h->xcd_chunk;
    a.reverse = reverse ? 1 : 0;
    a.split = h->pipelined ? 1 : h->tile_split;
    a.c1 = (R)c1;
    if (h->N > kMaxStates) {
        heom_hip::launch_stage_wide(h->N, h->term_slots, h->real_h, stage, a, el, h->wide_ados, h->stream);
    } else if (h->pipelined) {
        if constexpr (std::is_same<R, float>::value)
            heom_hip::launch_stage_pair(h->N, h->pair_slots, h->real_h, stage, a, el, h->pterms.p, h->pipe_wgs,
                                        h->stream);
        else
            heom_hip::launch_stage_pipelined(h->N, h->term_slots, h->real_h, stage, a, el, h->pipe_wgs, h->stream);
    } else {
        heom_hip::launch_stage(h->N, h->term_slots, h->real_h, stage, a, el, h->stream);
    }
    HIPC(hipGetLastError());
}

void launch_stage_buffers(heom_handle* h, int stage, const void* yin, const void* y0, void* pbuf, void* yout,
                          double c1, bool reverse, bool adjoint, const double2* top_src)
{
    if (h->fp32())
        raise(HEOM_E_UNSUPPORTED, "explicit-buffer stages need FP64 storage");
    if (top_src && !(adjoint && stage == 4))
        raise(HEOM_E_INVALID, "a top source goes with the adjoint's fourth stage only");
    if (adjoint) {
        AdjointState& A = *h->adj;
        StageArgsT<double> a = stage_args<double>(h, yin, y0, pbuf, yout, c1, reverse, A.pfss.p);
        a.top_src = top_src;
        a.top_q = h->q_top.p;
        const ElementArgsT<double> el{A.hs.p, A.gam.p, A.terms.p};
        heom_hip::launch_stage_adjoint(h->N, h->term_slots, A.real_h, stage, a, el, h->stream);
    } else {
        const StageArgsT<double> a = stage_args<double>(h, yin, y0, pbuf, yout, c1, reverse, h->pfss.p);
        const ElementArgsT<double> el{h->hs.p, h->gam.p, h->terms.p};
        if (h->N > kMaxStates)
            heom_hip::launch_stage_wide(h->N, h->term_slots, h->real_h, stage, a, el, h->wide_ados, h->stream);
        else if (h->pipelined)
            heom_hip::launch_stage_pipelined(h->N, h->term_slots, h->real_h, stage, a, el, h->pipe_wgs, h->stream);
        else
            heom_hip::launch_stage(h->N, h->term_slots, h->real_h, stage, a, el, h->stream);
    }
    HIPC(hipGetLastError());
    h->timing.rhs_launches += 1;
}

template <int NT>
void tier_fixup_stage(int stage, dim3 grid, dim3 block, size_t lds, hipStream_t st, const TierArgs& a)
{
    switch (stage) {
    case 0: hipLaunchKernelGGL((tier_fixup<0, NT>), grid, block, lds, st, a); break;
    case 1: hipLaunchKernelGGL((tier_fixup<1, NT>), grid, block, lds, st, a); break;
    case 2: hipLaunchKernelGGL((tier_fixup<2, NT>), grid, block, lds, st, a); break;
    case 3: hipLaunchKernelGGL((tier_fixup<3, NT>), grid, block, lds, st, a); break;
    default: hipLaunchKernelGGL((tier_fixup<4, NT>), grid, block, lds, st, a); break;
    }
}

// the FP64 pass over the tier-0/1 ADOs after a complex64 stage kernel
void tier_fixup_launch(heom_handle* h, int stage, const unsigned char* yin, unsigned char* yout, double c1)
{
    TierArgs a{};
    a.yin = reinterpret_cast<const float2*>(yin);
    a.yout = reinterpret_cast<float2*>(yout);
    a.pout = h->pbuf.as<float2>();
    a.nb = h->nb.p;
    a.q = h->sp_q.p;
    a.inst = h->sp_inst.p;
    a.sp_nb = h->sp_nb.p;
    // shadows: stage 1 y -> A, stage 2 A -> B (+P), stage 3 B -> A, stage 4 A -> y;
    // stage 0 (one RHS evaluation) y -> A
    a.in64 = stage == 2 ? h->sa64.p : stage == 3 ? h->sb64.p : stage == 4 ? h->sa64.p : h->y64.p;
    a.out64 = stage == 2 ? h->sb64.p : stage == 4 ? h->y64.p : h->sa64.p;
    a.y64 = h->y64.p;
    a.p64 = h->p64.p;
    a.hs = h->hs_stage ? h->hs_stage : h->hs.p;
    a.gam = h->gam.p;
    a.terms = h->terms.p;
    a.pfss = h->sp_pfss.p;
    a.inst_class = h->uniform ? nullptr : h->inst_class_d.p;
    a.paired = h->paired ? 1 : 0;
    a.mp = h->Mp;
    a.S = h->S;
    a.N = h->N;
    a.W = h->W;
    a.nt = h->term_slots;
    a.zero = h->zero_slot;
    a.c1 = c1;
    // one workgroup per tier-0/1 ADO, a wavefront for N <= 8
    const dim3 grid(h->S), block(std::min(256, (h->nel + 63) / 64 * 64));
    const size_t lds = 2 * (size_t)h->nel * sizeof(double2) + (size_t)h->W * (sizeof(int4) + sizeof(int2));
    switch (h->term_slots) {
    case 2: tier_fixup_stage<2>(stage, grid, block, lds, h->stream, a); break;
    case 4: tier_fixup_stage<4>(stage, grid, block, lds, h->stream, a); break;
    default: tier_fixup_stage<8>(stage, grid, block, lds, h->stream, a); break;
    }
    HIPC(hipGetLastError());
}

void launch_stage(heom_handle* h, int stage, const unsigned char* yin, unsigned char* yout, double c1,
                  bool reverse = false)
{
    if (h->fp32()) {
        const ElementArgsT<float> el{h->hs32_stage ? h->hs32_stage : h->hs32.p, h->gam32.p, h->terms32.p};
        launch_stage_r<float>(h, stage, yin, yout, c1, reverse, el, h->pfss32.p);
        tier_fixup_launch(h, stage, yin, yout, c1);
    } else {
        const ElementArgsT<double> el{h->hs_stage ? h->hs_stage : h->hs.p, h->gam.p, h->terms.p};
        launch_stage_r<double>(h, stage, yin, yout, c1, reverse, el, h->pfss.p);
    }
    h->timing.rhs_launches += 1;
}

// one classical RK4 step in the 12-pass schedule of heom_device.hpp
void rk4_step_launches(heom_handle* h, double dt)
{
    // stages alternate their tile walk, so that each starts where the previous one
    // ended, on the part of its input that is still in the last-level cache
    launch_stage(h, 1, h->y.p, h->ya.p, 0.5 * dt, false);
    launch_stage(h, 2, h->ya.p, h->yb.p, 0.5 * dt, true);
    launch_stage(h, 3, h->yb.p, h->ya.p, dt, false);
    launch_stage(h, 4, h->ya.p, h->y.p, dt / 6.0, true);
}

void drop_graph(heom_handle* h)
{
    if (h->step_graph)
        HIPC(hipGraphExecDestroy(h->step_graph));
    h->step_graph = nullptr;
}

// The step graph holds the stage launches with their arguments (pointers to the
// handle's buffers, the stage coefficients of dt, the kernel variant for real or
// complex H); it is rebuilt when dt or the variant changes.  It contains kernel
// launches only.
void rk4_steps(heom_handle* h, double dt, int64_t nsteps)
{
    reject_tss(h, "RK4 propagation");
    if (!h->use_graphs || nsteps < 2) {
        for (int64_t s = 0; s < nsteps; ++s)
            rk4_step_launches(h, dt);
        return;
    }
    if (h->step_graph && (h->graph_dt != dt || h->graph_real_h != h->real_h))
        drop_graph(h);
    if (!h->step_graph) {
        hipGraph_t g = nullptr;
        HIPC(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        const int64_t launches = h->timing.rhs_launches;
        try {
            rk4_step_launches(h, dt);
        } catch (...) {
            (void)hipStreamEndCapture(h->stream, &g);
            if (g)
                (void)hipGraphDestroy(g);
            throw;
        }
        h->timing.rhs_launches = launches;
        HIPC(hipStreamEndCapture(h->stream, &g));
        const hipError_t e = hipGraphInstantiate(&h->step_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPC(e);
        h->graph_dt = dt;
        h->graph_real_h = h->real_h;
    }
    for (int64_t s = 0; s < nsteps; ++s)
        HIPC(hipGraphLaunch(h->step_graph, h->stream));
    h->timing.rhs_launches += 4 * nsteps;
}

void refold_tables(heom_handle* h)
{
    const int64_t count = (int64_t)(h->W + 2) * h->Mp;
    hipLaunchKernelGGL(fold_mask, dim3(blocks_for(count, 256)), dim3(256), 0, h->stream, h->nb_raw.p, h->nb.p,
                       h->mask.p, h->q2u.p, count, h->zero_slot);
    HIPC(hipGetLastError());
}

// Tile-homogeneous ADO order.  An ADO without plus neighbours (the deepest tier: two
// thirds of the ADOs of the 8-site, 16-mode, depth-8 hierarchy) needs no plus gathers,
// and a stage kernel skips them for a whole tile when none of its 64 ADOs has one.
// `seq` lists one instance's ADOs in locality order (lexicographic); they are dealt
// into two queues by class, and a queue that reaches 64 ADOs becomes the next tile.
// So every tile but the last one or two holds one class, and the tiles still follow
// `seq`: an ADO moves by about two tiles of the other class at most.
std::vector<int32_t> group_tiles(const std::vector<int32_t>& seq, const std::vector<uint8_t>& has_plus)
{
    std::vector<int32_t> out, qd, qs;
    out.reserve(seq.size());
    for (const int32_t u : seq) {
        std::vector<int32_t>& q = has_plus[(size_t)u] ? qs : qd;
        q.push_back(u);
        if ((int)q.size() == kTile) {
            out.insert(out.end(), q.begin(), q.end());
            q.clear();
        }
    }
    out.insert(out.end(), qs.begin(), qs.end());
    out.insert(out.end(), qd.begin(), qd.end());
    return out;
}

// [ntiles] 1 if some ADO of the tile has a plus neighbour in the unmasked table, else 0
void build_tile_plus(heom_handle* h, const std::vector<int2>& nb)
{
    std::vector<int32_t> f((size_t)h->ntiles, 0);
    for (int m = 0; m < h->W; ++m)
        for (int64_t q = 0; q < h->Mp; ++q)
            if ((int32_t)((uint32_t)nb[(size_t)m * h->Mp + q].x >> 6) != h->zero_slot)
                f[(size_t)(q / kTile)] = 1;
    upload(h->tile_plus, f, h->stream);
}

// H [count][N*N*2] in J -> H / hbar on the device
void set_hamiltonians(heom_handle* h, const double* H)
{
    reject_tss(h, "setting the Hamiltonian");
    const size_t n = (size_t)h->count * h->nel;
    std::vector<double2> hs(n);
    std::vector<float2> hs32(n);
    bool real = true;
    for (size_t e = 0; e < n; ++e) {
        hs[e] = make_double2(H[2 * e] / kHbar, H[2 * e + 1] / kHbar);
        hs32[e] = make_float2((float)hs[e].x, (float)hs[e].y);
        if (H[2 * e + 1] != 0.0)
            real = false;
    }
    if (h->hs.p == nullptr)
        h->hs.alloc(n);
    HIPC(hipMemcpyAsync(h->hs.p, hs.data(), n * sizeof(double2), hipMemcpyHostToDevice, h->stream));
    if (h->fp32()) {
        if (h->hs32.p == nullptr)
            h->hs32.alloc(n);
        HIPC(hipMemcpyAsync(h->hs32.p, hs32.data(), n * sizeof(float2), hipMemcpyHostToDevice, h->stream));
    }
    HIPC(hipStreamSynchronize(h->stream));
    h->real_h = real;
    drop_graph(h);
    adjoint_invalidate(h);
}

// coupling-list entries of state s equal to bath b (states past the class's own have none)
int count_bath(const BathClass& c, int s, int b)
{
    int n = 0;
    if (s < c.N)
        for (int t = 0; t < c.BM; ++t)
            if (c.coupling[(size_t)s * c.BM + t] == b)
                ++n;
    return n;
}

int count_bath(const heom_handle* h, int s, int b) { return count_bath(h->classes[0], s, b); }

// Neighbour terms per matrix element and the "same"-term constants.  For element
// (i, j) and mode m of bath b = m / K, heom_ode.cl adds, for every coupling-list entry
// of state i equal to b, +i sqrt((n_m+1)|c_m|/hbar) sigma^{n+m} and
// (+i c_m/hbar + [k=0] Phi_b) sqrt(n_m hbar/|c_m|) sigma^{n-m}; for every entry of
// state j equal to b the same with -i (heom_ode.cl:184-266, 280-405); the same term
// subtracts Delta_b for the baths of i and of j when i != j (heom_ode.cl:138-168).
// With c_i, c_j the number of entries equal to b in rows i and j, one Term per mode
// carries plus_c = (c_i - c_j) sqrt(|c_m|/hbar) and
// minus_c = (i c_m/hbar (c_i - c_j) + [k=0] Phi_b (c_i + c_j)) sqrt(hbar/|c_m|).
// The plane terms of the paired layout: per class and plane, the union of its two
// elements' modes with both elements' constants (zero where an element lacks the mode),
// padded with mode W.  Sets pair_slots to 0 when the layout does not apply (N outside
// 4..8, or more than 8 terms on a plane).
void build_pair_terms(heom_handle* h, const std::vector<std::vector<Term>>& per, int C)
{
    const int N = h->N, W = h->W, D = pair_planes(N);
    h->pair_slots = 0;
    if (N < 4 || N > kMaxStates)
        return;
    std::vector<std::vector<PairTerm>> planes((size_t)C * D);
    int most = 0;
    for (int cl = 0; cl < C; ++cl)
        for (int i = 0; i < N; ++i)
            for (int j = i; j < N; ++j) {
                // the plane's elements: (i, j) with (j, i), or the diagonals (i, i), (i+1, i+1)
                if (i == j && (i & 1))
                    continue;
                const bool diag = i == j;
                const int e0 = i * N + j, e1 = diag ? (i + 1 < N ? (i + 1) * N + i + 1 : -1) : j * N + i;
                const int d = pair_slot(N, i, j).x;
                std::vector<PairTerm>& out = planes[(size_t)cl * D + d];
                for (int half = 0; half < 2; ++half) {
                    const int e = half ? e1 : e0;
                    if (e < 0)
                        continue;
                    for (const Term& t : per[(size_t)cl * N * N + e]) {
                        auto it = std::find_if(out.begin(), out.end(),
                                               [&](const PairTerm& pt) { return pt.mode == t.mode; });
                        if (it == out.end()) {
                            PairTerm z{};
                            z.mode = t.mode;
                            out.push_back(z);
                            it = out.end() - 1;
                        }
                        it->plus_c[half] = (float)t.plus_c;
                        it->minus_c[half] = make_float2((float)t.minus_c.x, (float)t.minus_c.y);
                    }
                }
                most = std::max(most, (int)out.size());
            }
    if (most > 8)
        return;
    h->pair_slots = stage_term_slots(most);
    std::vector<PairTerm> flat((size_t)C * D * h->pair_slots);
    for (size_t pl = 0; pl < planes.size(); ++pl)
        for (int k = 0; k < h->pair_slots; ++k) {
            PairTerm pad{};
            pad.mode = W;
            PairTerm& dst = flat[pl * h->pair_slots + k];
            dst = k < (int)planes[pl].size() ? planes[pl][k] : pad;
            dst.info = (int32_t)planes[pl].size();
        }
    upload(h->pterms, flat, h->stream);
}

void build_coefficients(heom_handle* h)
{
    const int N = h->N, W = h->W, C = (int)h->classes.size();
    std::vector<std::vector<Term>> per((size_t)C * N * N);
    std::vector<double2> gam((size_t)C * N * N, make_double2(0.0, 0.0));
    int most = 0;
    for (int cl = 0; cl < C; ++cl) {
        const BathClass& bc = h->classes[cl];
        for (int i = 0; i < N; ++i) {
            for (int j = 0; j < N; ++j) {
                const int e = i * N + j;
                std::vector<Term>& pe = per[(size_t)cl * N * N + e];
                for (int m = 0; m < bc.W; ++m) {
                    const int b = m / bc.K, k = m % bc.K;
                    const int ci = count_bath(bc, i, b), cj = count_bath(bc, j, b);
                    const double cr = bc.cbk[2 * m], cim = bc.cbk[2 * m + 1];
                    const double a = std::hypot(cr, cim);
                    if (a == 0.0)
                        continue;  // a mode with c_m = 0 contributes nothing (c sqrt(1/|c|) -> 0)
                    const double cp = std::sqrt(a / kHbar), cs = std::sqrt(kHbar / a);
                    Term t{};
                    t.mode = m;
                    t.plus_c = (ci - cj) * cp;
                    double br = -cim / kHbar * (ci - cj), bi = cr / kHbar * (ci - cj);
                    if (k == 0) {
                        br += bc.pf_circ[2 * b] * (ci + cj);
                        bi += bc.pf_circ[2 * b + 1] * (ci + cj);
                    }
                    t.minus_c = make_double2(br * cs, bi * cs);
                    if (t.plus_c != 0.0 || t.minus_c.x != 0.0 || t.minus_c.y != 0.0)
                        pe.push_back(t);
                }
                most = std::max(most, (int)pe.size());
                if (i != j) {
                    double gr = 0.0, gi = 0.0;
                    for (int t = 0; t < bc.BM; ++t) {
                        const int bi = i < bc.N ? bc.coupling[(size_t)i * bc.BM + t] : -1;
                        if (bi >= 0) {
                            gr -= bc.pf_same[2 * bi];
                            gi -= bc.pf_same[2 * bi + 1];
                        }
                        const int bj = j < bc.N ? bc.coupling[(size_t)j * bc.BM + t] : -1;
                        if (bj >= 0) {
                            gr -= bc.pf_same[2 * bj];
                            gi -= bc.pf_same[2 * bj + 1];
                        }
                    }
                    gam[(size_t)cl * N * N + e] = make_double2(gr, gi);
                }
            }
        }
    }
    if (N <= kMaxStates ? !stage_supported(N, most) : most > 8)
        raise(HEOM_E_UNSUPPORTED, "more than 8 neighbour terms per matrix element");
    h->term_slots = stage_term_slots(most);
    // pad every element's list; padding terms read table column W, which always
    // holds the zero slot
    std::vector<Term> flat((size_t)C * N * N * h->term_slots);
    for (int e = 0; e < C * N * N; ++e)
        for (int k = 0; k < h->term_slots; ++k) {
            Term t{};
            t.mode = W;
            t.plus_c = 0.0;
            t.minus_c = make_double2(0.0, 0.0);
            Term& dst = flat[(size_t)e * h->term_slots + k];
            dst = k < (int)per[e].size() ? per[e][k] : t;
            dst.info = (int32_t)per[e].size() | (dst.plus_c != 0.0 ? kTermHasPlus : 0) |
                       (dst.minus_c.x != 0.0 || dst.minus_c.y != 0.0 ? kTermHasMinus : 0);
        }
    upload(h->terms, flat, h->stream);
    upload(h->gam, gam, h->stream);
    if (h->fp32()) {
        std::vector<TermT<float>> f32(flat.size());
        for (size_t t = 0; t < flat.size(); ++t) {
            f32[t].minus_c = make_float2((float)flat[t].minus_c.x, (float)flat[t].minus_c.y);
            f32[t].plus_c = (float)flat[t].plus_c;
            f32[t].mode = flat[t].mode;
            f32[t].info = flat[t].info;
        }
        std::vector<float2> g32(gam.size());
        for (size_t e = 0; e < gam.size(); ++e)
            g32[e] = make_float2((float)gam[e].x, (float)gam[e].y);
        upload(h->terms32, f32, h->stream);
        upload(h->gam32, g32, h->stream);
        build_pair_terms(h, per, C);
    }
}

void validate(const heom_desc* d)
{
    if (!d)
        raise(HEOM_E_INVALID, "desc is NULL");
    if (d->states < 1)
        raise(HEOM_E_INVALID, "states must be > 0");
    if (d->states > kMaxStates &&
        !wide_supported(d->states, d->precision == HEOM_PRECISION_FP32 ? 8 : 16, std::max(d->width, 0)))
        raise(HEOM_E_UNSUPPORTED, "states > 32 (or too many modes for the large-system kernel) not supported by this build");
    if (d->matsubaras < 1 || d->baths < 1 || d->baths_max_per_state < 1)
        raise(HEOM_E_INVALID, "matsubaras, baths and baths_max_per_state must be > 0");
    if (d->width != d->baths * d->matsubaras)
        raise(HEOM_E_INVALID, "width must equal baths * matsubaras");
    if (d->matrices < 1)
        raise(HEOM_E_INVALID, "matrices must be > 0");
    if (d->matrices > kMaxAdos - 2 * kTile)
        raise(HEOM_E_UNSUPPORTED, "more than 2^26 - 129 ADOs");
    if (d->precision != HEOM_PRECISION_FP64 && d->precision != HEOM_PRECISION_FP32)
        raise(HEOM_E_INVALID, "precision must be HEOM_PRECISION_FP64 or HEOM_PRECISION_FP32");
    if (!d->coupling || !d->hamiltonian || !d->ado_tuples || !d->ado_plus || !d->ado_minus || !d->pf_same_sum ||
        !d->pf_same || !d->pf_circ || !d->cbk)
        raise(HEOM_E_INVALID, "desc has a NULL array");
    for (int t = 0; t < d->states * d->baths_max_per_state; ++t)
        if (d->coupling[t] >= d->baths)
            raise(HEOM_E_INVALID, "coupling entry >= baths");
}

// the instances of a batch share states, baths, coupling and bath prefactors
// The instances of a batch may differ in states, modes, coupling and bath prefactors
// (they are padded to the largest N and W, and grouped into bath classes); they share
// the storage precision.
void validate_batch(const heom_desc* d, int count)
{
    for (int k = 1; k < count; ++k)
        if (d[k].precision != d[0].precision)
            raise(HEOM_E_INVALID, "batch instances differ in precision");
}

int guard(const std::function<void()>& fn)
{
    try {
        fn();
        g_error.clear();
        return HEOM_OK;
    } catch (const heom_error& e) {
        g_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_error = "host allocation failed";
        return HEOM_E_MEMORY;
    } catch (const std::exception& e) {
        g_error = e.what();
        return HEOM_E_INVALID;
    }
}

void need(heom_handle* h)
{
    if (!h)
        raise(HEOM_E_INVALID, "handle is NULL");
}

// The persistent pipelined stage kernel runs one workgroup per CU; it needs N >= 6 and
// the neighbour-table slice of a tile in LDS.  A hierarchy with fewer tiles than half the
// CUs would leave most of the GPU idle with one workgroup per tile: it takes the tile
// kernel with each tile split over 4 (or 2) workgroups where N allows.  In the
// environment, HEOM_STAGE_KERNEL=tile selects the one-workgroup-per-tile kernel,
// =split the tile kernel with the largest split, =pipelined the pipelined kernel where
// it applies.  The tile order (chunk_tile in heom_device.hpp) defaults to one
// contiguous range per XCD for the tile kernel and to XCD chunks of one tile per
// workgroup for the pipelined kernel; HEOM_XCD_CHUNK=<tiles> overrides it.
std::string stage_kernel_request()
{
    const char* env = std::getenv("HEOM_STAGE_KERNEL");
    const std::string want = env ? env : "";
    if (!want.empty() && want != "tile" && want != "split" && want != "pipelined")
        raise(HEOM_E_INVALID, "HEOM_STAGE_KERNEL must be tile, split or pipelined");
    return want;
}

int device_cus(heom_handle* h)
{
    int cus = 0;
    HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    return cus;
}

// FP32 storage with N <= 8 runs the pipelined kernel in the paired layout (stage_pair.hip)
// where it applies, and the tile kernel on one plane per element otherwise; the choice
// fixes the layout of the planes for the handle's lifetime.
bool use_paired_layout(heom_handle* h)
{
    if (!h->fp32() || h->N > kMaxStates || h->pair_slots == 0 || !pair_supported(h->N, h->W, h->Mp))
        return false;
    const std::string want = stage_kernel_request();
    return want == "pipelined" || (want.empty() && 2 * h->ntiles > device_cus(h));
}

void choose_stage_kernel(heom_handle* h)
{
    const std::string want = stage_kernel_request();
    const int cus = device_cus(h);
    const bool small = 2 * h->ntiles <= cus;
    h->tile_split = 1;
    if (h->N <= kMaxStates && !h->paired && (want == "split" || (want.empty() && small)))
        for (int sp : {4, 2})
            if (h->tile_split == 1 && tile_split_supported(h->N, sp))
                h->tile_split = sp;
    // complex128: the pipelined kernel where it applies; complex64: the layout decides
    // (its pipelined kernel on paired planes, the tile kernel on one plane per element)
    h->pipelined = h->fp32() ? h->paired
                             : want != "tile" && want != "split" && h->tile_split == 1 &&
                                   pipe_supported(h->N, h->W, h->Mp, (int)h->cbytes());
    if (h->pipelined) {
        const int wgs = std::max(1, std::min(cus, h->ntiles));
        h->pipe_wgs = (wgs + kNumXcd - 1) / kNumXcd * kNumXcd;
        // one and a half tiles per workgroup of the XCD per chunk (48 at 256 workgroups):
        // on the 735,471-ADO hierarchy chunks of 48, 72 and 96 tiles measured 1-2% faster
        // than 32 or 64 in both precisions (the autotuner times both kinds)
        const int per_xcd = h->pipe_wgs / kNumXcd;
        h->xcd_chunk = h->ntiles >= kNumXcd * 3 * per_xcd ? per_xcd * 3 / 2 : per_xcd;
    } else {
        h->xcd_chunk = std::max(1, h->ntiles / kNumXcd);
    }
    if (const char* c = std::getenv("HEOM_XCD_CHUNK")) {
        const long v = std::strtol(c, nullptr, 10);
        if (v < 1 || v > h->ntiles)
            raise(HEOM_E_INVALID, "HEOM_XCD_CHUNK must be in [1, number of 64-ADO tiles]");
        h->xcd_chunk = (int)v;
    }
}

// the tier-0/1 ADOs of every instance, their FP64 shadows and neighbour map
void setup_tiers(heom_handle* h, const heom_desc* d, const std::vector<std::vector<int32_t>>& u2q)
{
    const int W = h->W;
    std::vector<int64_t> sq;
    std::vector<int32_t> sinst;
    std::vector<double2> spf;
    std::vector<std::vector<int32_t>> local(h->count);  // local uid -> tier index, or -1
    h->sp_top.assign(h->count, 0);
    h->sp_uid.clear();
    for (int k = 0; k < h->count; ++k) {
        const int32_t* T = d[k].ado_tuples;
        const int Wk = d[k].width;
        local[k].assign(d[k].matrices, -1);
        for (int64_t u = 0; u < d[k].matrices; ++u) {
            int64_t sum = 0;
            for (int m = 0; m < Wk; ++m)
                sum += T[u * Wk + m];
            if (sum > 1)
                continue;
            local[k][u] = (int32_t)sq.size();
            if (u == 0)
                h->sp_top[k] = (int32_t)sq.size();
            sq.push_back(h->inst_q[k] + u2q[k][u]);
            sinst.push_back(k);
            spf.push_back(make_double2(d[k].pf_same_sum[2 * u], d[k].pf_same_sum[2 * u + 1]));
            h->sp_uid.push_back(h->inst_u[k] + u);
        }
    }
    h->S = (int)sq.size();
    std::vector<int2> snb((size_t)h->S * W, make_int2(-1, -1));
    for (int s = 0; s < h->S; ++s) {
        const int k = sinst[s];
        const int64_t u = h->sp_uid[s] - h->inst_u[k];
        const int Wk = d[k].width;
        for (int m = 0; m < Wk; ++m) {
            const int32_t p = d[k].ado_plus[u * Wk + m], mn = d[k].ado_minus[u * Wk + m];
            snb[(size_t)s * W + m] = make_int2(p >= 0 ? local[k][p] : -1, mn >= 0 ? local[k][mn] : -1);
        }
    }
    upload(h->sp_q, sq, h->stream);
    upload(h->sp_inst, sinst, h->stream);
    upload(h->sp_pfss, spf, h->stream);
    upload(h->sp_nb, snb, h->stream);
    upload(h->sp_top_d, h->sp_top, h->stream);
    for (auto* b : {&h->y64, &h->sa64, &h->sb64, &h->p64, &h->sp_tmp}) {
        b->alloc((size_t)h->S * h->nel);
        HIPC(hipMemsetAsync(b->p, 0, b->bytes(), h->stream));
    }
}

void create_batch(heom_handle* h, const heom_desc* d, int count)
{
    for (int k = 0; k < count; ++k)
        validate(&d[k]);
    validate_batch(d, count);
    HIPC(hipGetDevice(&h->device));
    HIPC(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPC(hipEventCreate(&h->ev_start));
    HIPC(hipEventCreate(&h->ev_stop));
    // the largest states and modes; bath classes of instances with equal bath data
    int N = 0, W = 0;
    h->classes.clear();
    h->inst_class.assign(count, 0);
    for (int k = 0; k < count; ++k) {
        N = std::max(N, d[k].states);
        W = std::max(W, d[k].width);
        BathClass bc;
        bc.N = d[k].states;
        bc.K = d[k].matsubaras;
        bc.B = d[k].baths;
        bc.BM = d[k].baths_max_per_state;
        bc.W = d[k].width;
        bc.coupling.assign(d[k].coupling, d[k].coupling + (size_t)bc.N * bc.BM);
        bc.cbk.assign(d[k].cbk, d[k].cbk + 2 * bc.W);
        bc.pf_same.assign(d[k].pf_same, d[k].pf_same + 2 * bc.B);
        bc.pf_circ.assign(d[k].pf_circ, d[k].pf_circ + 2 * bc.B);
        const auto it = std::find(h->classes.begin(), h->classes.end(), bc);
        h->inst_class[k] = (int32_t)(it - h->classes.begin());
        if (it == h->classes.end())
            h->classes.push_back(std::move(bc));
    }
    h->uniform = h->classes.size() == 1;
    if (N > kMaxStates && !wide_supported(N, d[0].precision == HEOM_PRECISION_FP32 ? 8 : 16, W))
        raise(HEOM_E_UNSUPPORTED, "states > 32 (or too many modes for the large-system kernel) not supported by this build");
    const BathClass& c0 = h->classes[0];
    h->N = N;
    h->K = c0.K;
    h->B = c0.B;
    h->BM = c0.BM;
    h->W = W;
    h->nel = N * N;
    h->prec = d[0].precision;
    h->count = count;
    h->coupling = c0.coupling;
    h->cbk = c0.cbk;
    h->pf_same = c0.pf_same;
    h->pf_circ = c0.pf_circ;

    // instance k: uids [inst_u, inst_u + M_k), internal slots from inst_q (a multiple of
    // 64); the zero slot follows the last instance
    h->inst_M.resize(count);
    h->inst_u.resize(count);
    h->inst_q.resize(count);
    h->inst_n.resize(count);
    int64_t u0 = 0, q0 = 0;
    for (int k = 0; k < count; ++k) {
        h->inst_M[k] = d[k].matrices;
        h->inst_u[k] = u0;
        h->inst_q[k] = q0;
        h->inst_n[k] = d[k].matrices;
        u0 += d[k].matrices;
        q0 += k + 1 < count ? (d[k].matrices + kTile - 1) / kTile * kTile : d[k].matrices;
    }
    if (q0 + 1 > kMaxAdos - kTile)
        raise(HEOM_E_UNSUPPORTED, "batch larger than 2^26 - 129 ADO slots");
    h->M = u0;
    h->active = u0;
    h->zero_slot = (int32_t)q0;
    h->Mp = (q0 + 1 + kTile - 1) / kTile * kTile;
    h->ntiles = (int)(h->Mp / kTile);
    const int64_t Mp = h->Mp;
    const int32_t zero = h->zero_slot;

    // rows 0..W-1 are the modes; rows W and W+1 hold the zero slot everywhere (row W
    // serves padding terms, and the pipelined kernel loads rows in pairs)
    std::vector<int2> nb((size_t)(W + 2) * Mp, make_int2((int32_t)((uint32_t)zero << 6), zero));
    std::vector<double2> pfss(Mp, make_double2(0.0, 0.0));
    std::vector<int32_t> q2u(Mp, -1), tinst(h->ntiles, count - 1);
    std::vector<int64_t> qtop(count);
    std::vector<std::vector<int32_t>> u2q_all(count);
    h->lex_tuples.assign((size_t)h->M * W, 0);
    h->lex_uid.assign((size_t)h->M, 0);
    h->u2q_host.assign((size_t)h->M, -1);
    for (int k = 0; k < count; ++k) {
        const int64_t M = d[k].matrices, qb = h->inst_q[k];
        const int32_t* T = d[k].ado_tuples;
        const int Wk = d[k].width;  // modes Wk..W-1 (padding) are unused by this instance
        for (int64_t t = 0; t < M * Wk; ++t)
            if (T[t] < 0 || T[t] > kMaxTupleValue)
                raise(HEOM_E_UNSUPPORTED, "tuple entries must be in [0, 62]");
        for (int m = 0; m < Wk; ++m)
            if (T[m] != 0)
                raise(HEOM_E_INVALID, "uid 0 must be the top ADO (all-zero tuple)");
        // lexicographic ADO order (mode 0 most significant)
        std::vector<int32_t> order(M);
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            return std::lexicographical_compare(T + (int64_t)a * Wk, T + (int64_t)a * Wk + Wk, T + (int64_t)b * Wk,
                                                T + (int64_t)b * Wk + Wk);
        });
        for (int64_t q = 0; q < M; ++q) {
            // the sorted tuples of heom_get_ados: lexicographic position -> uid
            const int64_t lx = h->inst_u[k] + q;
            for (int m = 0; m < Wk; ++m)
                h->lex_tuples[(size_t)lx * W + m] = (uint8_t)T[(int64_t)order[q] * Wk + m];
            h->lex_uid[(size_t)lx] = (int32_t)(h->inst_u[k] + order[q]);
        }
        // slots: the ADOs of tier <= 2 first (with FP32 storage the FP64 pass over the
        // tier-0/1 ADOs reads their tier-2 neighbours, which then share cache lines), then
        // the others, each part in lexicographic order, grouped into tiles with and
        // without plus edges
        std::vector<uint8_t> has_plus(M, 0), low(M, 0);
        for (int64_t u = 0; u < M; ++u) {
            int64_t tier = 0;
            for (int m = 0; m < Wk; ++m) {
                tier += T[u * Wk + m];
                if (d[k].ado_plus[u * Wk + m] >= 0)
                    has_plus[u] = 1;
            }
            low[u] = tier <= 2;
        }
        std::vector<int32_t> seq;
        seq.reserve(M);
        for (int pass = 1; pass >= 0; --pass)
            for (int64_t q = 0; q < M; ++q)
                if (low[order[q]] == pass)
                    seq.push_back(order[q]);
        const std::vector<int32_t> slots = group_tiles(seq, has_plus);
        std::vector<int32_t>& u2q = u2q_all[k];
        u2q.assign(M, 0);
        for (int64_t q = 0; q < M; ++q) {
            u2q[slots[q]] = (int32_t)q;
            q2u[qb + q] = (int32_t)(h->inst_u[k] + slots[q]);
            h->u2q_host[(size_t)(h->inst_u[k] + slots[q])] = (int32_t)(qb + q);
        }
        qtop[k] = qb + u2q[0];
        for (int64_t t = qb / kTile; t < (qb + M + kTile - 1) / kTile; ++t)
            tinst[t] = k;
        for (int64_t q = 0; q < M; ++q) {
            const int64_t u = slots[q];
            for (int m = 0; m < Wk; ++m) {
                const int32_t p = d[k].ado_plus[u * Wk + m];
                const int32_t mn = d[k].ado_minus[u * Wk + m];
                if (p >= M || mn >= M)
                    raise(HEOM_E_INVALID, "plus/minus entry out of range");
                const int32_t pq = p >= 0 ? (int32_t)(qb + u2q[p]) : zero;
                const int32_t mq = (mn >= 0 && T[u * Wk + m] > 0) ? (int32_t)(qb + u2q[mn]) : zero;
                nb[(size_t)m * Mp + qb + q] = make_int2((int32_t)(((uint32_t)pq << 6) | (uint32_t)T[u * Wk + m]), mq);
            }
            pfss[qb + q] = make_double2(d[k].pf_same_sum[2 * u], d[k].pf_same_sum[2 * u + 1]);
        }
    }
    upload(h->nb_raw, nb, h->stream);
    upload(h->nb, nb, h->stream);
    build_tile_plus(h, nb);
    upload(h->pfss, pfss, h->stream);
    upload(h->q2u, q2u, h->stream);
    upload(h->tile_inst, tinst, h->stream);
    upload(h->q_top, qtop, h->stream);
    if (!h->uniform) {
        std::vector<int32_t> tcls(h->ntiles, h->inst_class[count - 1]);
        for (int t = 0; t < h->ntiles; ++t)
            tcls[t] = h->inst_class[tinst[t]];
        upload(h->tile_class, tcls, h->stream);
        upload(h->inst_class_d, h->inst_class, h->stream);
    }
    if (h->fp32()) {
        std::vector<float2> p32(Mp);
        for (int64_t q = 0; q < Mp; ++q)
            p32[q] = make_float2((float)pfss[q].x, (float)pfss[q].y);
        upload(h->pfss32, p32, h->stream);
    }
    h->mask.alloc(h->M);
    HIPC(hipMemsetAsync(h->mask.p, 1, h->M, h->stream));
    h->norms.alloc(h->M);
    h->mx.alloc((size_t)count * h->nel);
    h->small.alloc((size_t)count * h->nel);
    build_coefficients(h);
    h->paired = use_paired_layout(h);
    const size_t plane = h->paired ? (size_t)pair_planes(N) * Mp * 2 * sizeof(float2)
                                   : (size_t)N * N * Mp * h->cbytes();
    for (auto* b : {&h->y, &h->pbuf, &h->ya, &h->yb}) {
        b->alloc(plane);
        HIPC(hipMemsetAsync(b->p, 0, b->bytes(), h->stream));
    }
    // H of each instance, padded to N x N
    std::vector<double> H((size_t)count * h->nel * 2, 0.0);
    for (int k = 0; k < count; ++k)
        for (int i = 0; i < d[k].states; ++i)
            for (int j = 0; j < d[k].states; ++j)
                for (int c = 0; c < 2; ++c)
                    H[((size_t)k * h->nel + (size_t)i * N + j) * 2 + c] =
                        d[k].hamiltonian[((size_t)i * d[k].states + j) * 2 + c];
    set_hamiltonians(h, H.data());
    if (h->fp32())
        setup_tiers(h, d, u2q_all);
    choose_stage_kernel(h);
    HIPC(hipStreamSynchronize(h->stream));
}

// state in reference layout (AoS, all instances) -> planes (and FP64 tier shadows)
void upload_state(heom_handle* h, const double* state, unsigned char* dst)
{
    DevBuf<double2> stage;
    stage.alloc((size_t)h->M * h->nel);
    HIPC(hipMemcpyAsync(stage.p, state, stage.bytes(), hipMemcpyHostToDevice, h->stream));
    if (h->fp32()) {
        hipLaunchKernelGGL(aos_to_soa<float2>, dim3(blocks_for(h->Mp, 256)), dim3(256), 0, h->stream, stage.p,
                           h->y.as<float2>(), h->q2u.p, h->Mp, h->N, h->paired);
        std::vector<double2> rows((size_t)h->S * h->nel);
        for (int s = 0; s < h->S; ++s)
            std::memcpy(&rows[(size_t)s * h->nel], state + (size_t)h->sp_uid[s] * h->nel * 2,
                        h->nel * sizeof(double2));
        HIPC(hipMemcpyAsync(h->y64.p, rows.data(), rows.size() * sizeof(double2), hipMemcpyHostToDevice,
                            h->stream));
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(h->stream));
    } else {
        hipLaunchKernelGGL(aos_to_soa<double2>, dim3(blocks_for(h->Mp, 256)), dim3(256), 0, h->stream, stage.p,
                           reinterpret_cast<double2*>(dst ? dst : h->y.p), h->q2u.p, h->Mp, h->N, false);
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(h->stream));
    }
}

// planes `src` -> reference layout; with FP32 storage the tier-0/1 ADOs come from the
// FP64 shadow `rows`
void download_state(heom_handle* h, const unsigned char* src, const double2* rows, double* state)
{
    DevBuf<double2> stage;
    stage.alloc((size_t)h->M * h->nel);
    if (h->active < h->M)  // ADOs removed by heom_compact read back as zero
        HIPC(hipMemsetAsync(stage.p, 0, stage.bytes(), h->stream));
    if (h->fp32())
        hipLaunchKernelGGL(soa_to_aos<float2>, dim3(blocks_for(h->Mp, 256)), dim3(256), 0, h->stream,
                           reinterpret_cast<const float2*>(src), stage.p, h->q2u.p, h->Mp, h->N, h->paired);
    else
        hipLaunchKernelGGL(soa_to_aos<double2>, dim3(blocks_for(h->Mp, 256)), dim3(256), 0, h->stream,
                           reinterpret_cast<const double2*>(src), stage.p, h->q2u.p, h->Mp, h->N, false);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(state, stage.p, stage.bytes(), hipMemcpyDeviceToHost, h->stream));
    std::vector<double2> sh;
    if (h->fp32()) {
        sh.resize((size_t)h->S * h->nel);
        HIPC(hipMemcpyAsync(sh.data(), rows, sh.size() * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
    }
    HIPC(hipStreamSynchronize(h->stream));
    if (h->fp32())
        for (int s = 0; s < h->S; ++s)
            std::memcpy(state + (size_t)h->sp_uid[s] * h->nel * 2, &sh[(size_t)s * h->nel],
                        h->nel * sizeof(double2));
}

// the tops of all instances [count][nel] into dst (device)
void gather_tops(heom_handle* h, double2* dst)
{
    const int n = h->count * h->nel;
    if (h->fp32())
        hipLaunchKernelGGL(gather_rows, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->y64.p, dst,
                           h->sp_top_d.p, h->count, h->nel);
    else
        hipLaunchKernelGGL(gather_ados<double2>, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream,
                           h->y.as<double2>(), dst, h->q_top.p, h->count, h->Mp, h->N, false);
    HIPC(hipGetLastError());
}

void observe_launch(heom_handle* h, const double2* ops, double* pops, double2* traces)
{
    gather_tops(h, h->small.p);
    hipLaunchKernelGGL(observe_tops, dim3(h->count), dim3(64), 0, h->stream, h->small.p, ops, pops, traces, h->N);
    HIPC(hipGetLastError());
}

template <typename C>
void mmult_left_planes(heom_handle* h)
{
    const dim3 grid(h->ntiles), block(64 * h->N);
    C* y = h->y.as<C>();
    const int32_t* ti = h->count > 1 ? h->tile_inst.p : nullptr;
    if (h->N > kMaxStates) {
        const int64_t n = (int64_t)h->nel * h->Mp;
        hipLaunchKernelGGL(mmult_left_wide<C>, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, y, h->ya.as<C>(),
                           h->mx.p, ti, h->Mp, h->N);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(y, h->ya.p, h->y.bytes(), hipMemcpyDeviceToDevice, h->stream));
        return;
    }
    switch (h->N) {
    case 1: hipLaunchKernelGGL((mmult_left_kernel<C, 1>), grid, block, 0, h->stream, y, h->mx.p, ti, h->Mp, h->paired); break;
    case 2: hipLaunchKernelGGL((mmult_left_kernel<C, 2>), grid, block, 0, h->stream, y, h->mx.p, ti, h->Mp, h->paired); break;
    case 3: hipLaunchKernelGGL((mmult_left_kernel<C, 3>), grid, block, 0, h->stream, y, h->mx.p, ti, h->Mp, h->paired); break;
    case 4: hipLaunchKernelGGL((mmult_left_kernel<C, 4>), grid, block, 0, h->stream, y, h->mx.p, ti, h->Mp, h->paired); break;
    case 5: hipLaunchKernelGGL((mmult_left_kernel<C, 5>), grid, block, 0, h->stream, y, h->mx.p, ti, h->Mp, h->paired); break;
    case 6: hipLaunchKernelGGL((mmult_left_kernel<C, 6>), grid, block, 0, h->stream, y, h->mx.p, ti, h->Mp, h->paired); break;
    case 7: hipLaunchKernelGGL((mmult_left_kernel<C, 7>), grid, block, 0, h->stream, y, h->mx.p, ti, h->Mp, h->paired); break;
    default: hipLaunchKernelGGL((mmult_left_kernel<C, 8>), grid, block, 0, h->stream, y, h->mx.p, ti, h->Mp, h->paired); break;
    }
    HIPC(hipGetLastError());
}

// One configuration of the stage kernels (the fields of heom_tune_result).
struct StageConfig {
    int kind, chunk, wgs, ados, split = 1;
};

// every configuration the autotuner times for this handle: for N <= 8 the tile kernel
// and, where it fits, the pipelined kernel (one workgroup per CU), each with 1, 2, 4, ...,
// 64 tiles per XCD chunk, 24, 48, 72 and 96 (multiples of 24 measured fastest with the
// pipelined kernel on large hierarchies) and the default chunk; for N > 8 the wide kernel
// with every power-of-two tile that fits
std::vector<StageConfig> stage_candidates(heom_handle* h)
{
    std::vector<StageConfig> cands;
    const int nt = h->ntiles;
    auto chunks = [&](int dflt) {
        std::vector<int> v{1, 2, 4, 8, 16, 24, 32, 48, 64, 72, 96, dflt};
        std::vector<int> out;
        for (int c : v)
            if (c >= 1 && c <= nt && std::find(out.begin(), out.end(), c) == out.end())
                out.push_back(c);
        return out;
    };
    if (h->N > kMaxStates) {
        for (int T = wide_max_ados(h->N, (int)h->cbytes(), h->W); T >= 1; T /= 2)
            cands.push_back({HEOM_STAGE_WIDE, 1, 0, T});
    } else {
        // complex64 planes: the layout fixes the kernel family (paired: pipelined only)
        const bool pipe_ok = h->fp32() ? h->paired : pipe_supported(h->N, h->W, h->Mp, (int)h->cbytes());
        if (!h->paired) {
            for (int c : chunks(std::max(1, nt / kNumXcd)))
                cands.push_back({HEOM_STAGE_TILE, c, 0, kTile});
            for (int sp : {2, 4})  // tiles split over workgroups, at the default chunk
                if (tile_split_supported(h->N, sp))
                    cands.push_back({HEOM_STAGE_TILE, std::max(1, nt / kNumXcd), 0, kTile, sp});
        }
        if (pipe_ok) {
            int cus = 0;
            HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
            const int wgs = (std::max(1, std::min(cus, nt)) + kNumXcd - 1) / kNumXcd * kNumXcd;
            for (int c : chunks(wgs / kNumXcd))
                cands.push_back({HEOM_STAGE_PIPELINED, c, wgs, kTile});
        }
    }
    return cands;
}

void apply_stage_config(heom_handle* h, const StageConfig& c)
{
    h->pipelined = c.kind == HEOM_STAGE_PIPELINED;
    h->pipe_wgs = c.kind == HEOM_STAGE_PIPELINED ? c.wgs : h->pipe_wgs;
    h->xcd_chunk = c.chunk;
    h->wide_ados = c.kind == HEOM_STAGE_WIDE ? c.ados : 0;
    h->tile_split = c.kind == HEOM_STAGE_TILE ? c.split : 1;
    drop_graph(h);
}

void check_observe_args(heom_handle* h, double dt, int64_t nsteps, int64_t observe_steps, const double* ops,
                        double* pops, double* traces)
{
    if (!pops)
        raise(HEOM_E_INVALID, "pops is NULL");
    if (ops && !traces)
        raise(HEOM_E_INVALID, "traces is NULL while ops is given");
    if (nsteps < 0 || observe_steps < 1)
        raise(HEOM_E_INVALID, "nsteps must be >= 0 and observe_steps >= 1");
    if (!(dt > 0.0))
        raise(HEOM_E_INVALID, "step size must be > 0");
}

// The propagation of heom_propagate_observe with caller-provided device buffers: opsd
// [count][nel] (when ops is given), prec [rows][count][N], trec [rows][count].  With
// graphs = false the stages are launched one by one even when step graphs are enabled,
// so that nothing is captured or instantiated.
void propagate_observe_into(heom_handle* h, double dt, int64_t nsteps, int64_t observe_steps, const double* ops,
                            double* pops, double* traces, double2* opsd, double* prec, double2* trec, bool graphs)
{
    const int64_t rows = nsteps / observe_steps + 1;
    if (ops)
        HIPC(hipMemcpyAsync(opsd, ops, (size_t)h->count * h->nel * sizeof(double2), hipMemcpyHostToDevice,
                            h->stream));
    auto observe = [&](int64_t row) {
        gather_tops(h, h->small.p);
        hipLaunchKernelGGL(observe_tops, dim3(h->count), dim3(64), 0, h->stream, h->small.p, ops ? opsd : nullptr,
                           prec + row * h->count * h->N, ops ? trec + row * h->count : nullptr, h->N);
        HIPC(hipGetLastError());
    };
    resolve_timing(h);
    HIPC(hipEventRecord(h->ev_start, h->stream));
    observe(0);
    int64_t row = 1, done = 0;
    while (done < nsteps) {
        const int64_t n = std::min(observe_steps, nsteps - done);
        if (graphs) {
            rk4_steps(h, dt, n);
        } else {
            reject_tss(h, "RK4 propagation");
            for (int64_t s = 0; s < n; ++s)
                rk4_step_launches(h, dt);
            h->timing.rhs_launches += 4 * n;
        }
        done += n;
        if (n == observe_steps)
            observe(row++);
    }
    HIPC(hipEventRecord(h->ev_stop, h->stream));
    h->timing_pending = true;
    h->pending_steps = nsteps;
    HIPC(hipMemcpyAsync(pops, prec, (size_t)(rows * h->count * h->N) * sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
    if (ops)
        HIPC(hipMemcpyAsync(traces, trec, (size_t)(rows * h->count) * sizeof(double2), hipMemcpyDeviceToHost,
                            h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    resolve_timing(h);
}

}  // namespace heom_rt

// Scratch of heom_propagate_observe_noalloc: a device array of fixed size, one copy per
// GPU for the life of the process.  Its device address comes from hipGetSymbolAddress,
// once per device with that device current; in host code the array's name is not a
// device address.  Every call on a GPU shares that GPU's copy, so calls that overlap on
// one GPU race; each call writes all of the scratch it reads.
static __device__ __align__(256) unsigned char scratch[kObserveScratchBytes];

namespace heom_rt {

namespace {
std::mutex g_scratch_mutex;
std::vector<unsigned char*> g_scratch_ptr;  // per device, null until first use
}  // namespace

unsigned char* observe_scratch(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if ((int)g_scratch_ptr.size() <= device)
        g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[(size_t)device]) {
        int cur = 0;
        HIPC(hipGetDevice(&cur));
        if (cur != device)
            HIPC(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device)
            HIPC(hipSetDevice(cur));
        HIPC(e);
        g_scratch_ptr[(size_t)device] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[(size_t)device];
}

}  // namespace heom_rt

extern "C" {

const char* heom_last_error(void) { return g_error.c_str(); }

int heom_abi_version(void) { return HEOM_ABI_VERSION; }

int heom_create_batch(const heom_desc* descs, int32_t count, heom_handle** out)
{
    if (!out) {
        g_error = "out is NULL";
        return HEOM_E_INVALID;
    }
    *out = nullptr;
    heom_handle* h = nullptr;
    const int rc = guard([&] {
        if (!descs)
            raise(HEOM_E_INVALID, "desc is NULL");
        if (count < 1)
            raise(HEOM_E_INVALID, "count must be >= 1");
        h = new heom_handle();
        create_batch(h, descs, count);
    });
    if (rc != HEOM_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return HEOM_OK;
}

int heom_create(const heom_desc* d, heom_handle** out) { return heom_create_batch(d, 1, out); }

int heom_destroy(heom_handle* h)
{
    return guard([&] {
        need(h);
        delete h;
    });
}

int heom_instances(heom_handle* h, int32_t* count)
{
    return guard([&] {
        need(h);
        if (!count)
            raise(HEOM_E_INVALID, "count is NULL");
        *count = h->count;
    });
}

int heom_set_hamiltonian(heom_handle* h, const double* H)
{
    return guard([&] {
        need(h);
        if (!H)
            raise(HEOM_E_INVALID, "hamiltonian is NULL");
        set_hamiltonians(h, H);
    });
}

int heom_set_mask(heom_handle* h, const uint8_t* mask)
{
    return guard([&] {
        need(h);
        if (!mask)
            raise(HEOM_E_INVALID, "mask is NULL");
        for (int64_t u = 0; u < h->M; ++u)
            if (mask[u] > 2)
                raise(HEOM_E_INVALID, "mask values must be 0, 1 or 2");
        HIPC(hipMemcpyAsync(h->mask.p, mask, h->M, hipMemcpyHostToDevice, h->stream));
        refold_tables(h);
        adjoint_invalidate(h);
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_get_mask(heom_handle* h, uint8_t* mask)
{
    return guard([&] {
        need(h);
        if (!mask)
            raise(HEOM_E_INVALID, "mask is NULL");
        HIPC(hipMemcpyAsync(mask, h->mask.p, h->M, hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_set_state(heom_handle* h, const double* state)
{
    return guard([&] {
        need(h);
        if (!state)
            raise(HEOM_E_INVALID, "state is NULL");
        upload_state(h, state);
    });
}

int heom_set_top(heom_handle* h, const double* top)
{
    return guard([&] {
        need(h);
        if (!top)
            raise(HEOM_E_INVALID, "top is NULL");
        const int n = h->count * h->nel;
        HIPC(hipMemcpyAsync(h->small.p, top, n * sizeof(double2), hipMemcpyHostToDevice, h->stream));
        if (h->fp32()) {
            hipLaunchKernelGGL(scatter_ados<float2>, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->small.p,
                               h->y.as<float2>(), h->q_top.p, h->count, h->Mp, h->N, h->paired);
            hipLaunchKernelGGL(scatter_rows, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->small.p,
                               h->y64.p, h->sp_top_d.p, h->count, h->nel);
        } else {
            hipLaunchKernelGGL(scatter_ados<double2>, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->small.p,
                               h->y.as<double2>(), h->q_top.p, h->count, h->Mp, h->N, false);
        }
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_clear_state(heom_handle* h)
{
    return guard([&] {
        need(h);
        HIPC(hipMemsetAsync(h->y.p, 0, h->y.bytes(), h->stream));
        if (h->fp32())
            HIPC(hipMemsetAsync(h->y64.p, 0, h->y64.bytes(), h->stream));
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_get_state(heom_handle* h, double* state)
{
    return guard([&] {
        need(h);
        if (!state)
            raise(HEOM_E_INVALID, "state is NULL");
        download_state(h, h->y.p, h->y64.p, state);
    });
}

int heom_get_top(heom_handle* h, double* top)
{
    return guard([&] {
        need(h);
        if (!top)
            raise(HEOM_E_INVALID, "top is NULL");
        gather_tops(h, h->small.p);
        HIPC(hipMemcpyAsync(top, h->small.p, (size_t)h->count * h->nel * sizeof(double2), hipMemcpyDeviceToHost,
                            h->stream));
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_get_ados(heom_handle* h, int32_t instance, const int32_t* tuples, int64_t n, double* out)
{
    return guard([&] {
        need(h);
        if (instance < 0 || instance >= h->count)
            raise(HEOM_E_INVALID, "instance out of range");
        if (n < 0 || (n > 0 && (!tuples || !out)))
            raise(HEOM_E_INVALID, "n must be >= 0, tuples and out non-NULL");
        const int W = h->W, nel = h->nel;
        const int64_t lo = h->inst_u[instance], hi = lo + h->inst_M[instance];
        std::vector<uint8_t> key((size_t)W);
        std::vector<int64_t> slot((size_t)n), uid((size_t)n);
        for (int64_t a = 0; a < n; ++a) {
            for (int m = 0; m < W; ++m) {
                const int32_t v = tuples[a * W + m];
                if (v < 0 || v > kMaxTupleValue)
                    raise(HEOM_E_INVALID, "tuple " + std::to_string(a) + " is not an ADO of the instance");
                key[(size_t)m] = (uint8_t)v;
            }
            // binary search over the instance's sorted tuples
            int64_t l = lo, r = hi;
            while (l < r) {
                const int64_t mid = l + (r - l) / 2;
                const uint8_t* t = &h->lex_tuples[(size_t)mid * W];
                if (std::lexicographical_compare(t, t + W, key.begin(), key.end()))
                    l = mid + 1;
                else
                    r = mid;
            }
            if (l == hi || !std::equal(key.begin(), key.end(), &h->lex_tuples[(size_t)l * W]))
                raise(HEOM_E_INVALID, "tuple " + std::to_string(a) + " is not an ADO of the instance");
            uid[(size_t)a] = h->lex_uid[(size_t)l];
            slot[(size_t)a] = h->u2q_host[(size_t)uid[(size_t)a]];  // -1: removed by heom_compact
        }
        std::fill(out, out + (size_t)n * nel * 2, 0.0);
        std::vector<int64_t> live, at;
        for (int64_t a = 0; a < n; ++a)
            if (slot[(size_t)a] >= 0) {
                live.push_back(slot[(size_t)a]);
                at.push_back(a);
            }
        constexpr int64_t kChunk = 1 << 20;  // ADOs per gather launch (nel * chunk < 2^31)
        DevBuf<int64_t> qs;
        DevBuf<double2> buf;
        std::vector<double2> host;
        for (int64_t c0 = 0; c0 < (int64_t)live.size(); c0 += kChunk) {
            const int cnt = (int)std::min<int64_t>(kChunk, (int64_t)live.size() - c0);
            if (qs.n < (size_t)cnt) {
                qs.alloc((size_t)cnt);
                buf.alloc((size_t)cnt * nel);
            }
            HIPC(hipMemcpyAsync(qs.p, live.data() + c0, (size_t)cnt * sizeof(int64_t), hipMemcpyHostToDevice,
                                h->stream));
            const dim3 grid(blocks_for((int64_t)cnt * nel, 256)), block(256);
            if (h->fp32())
                hipLaunchKernelGGL(gather_ados<float2>, grid, block, 0, h->stream, h->y.as<float2>(), buf.p, qs.p, cnt,
                                   h->Mp, h->N, h->paired);
            else
                hipLaunchKernelGGL(gather_ados<double2>, grid, block, 0, h->stream, h->y.as<double2>(), buf.p, qs.p,
                                   cnt, h->Mp, h->N, false);
            HIPC(hipGetLastError());
            host.resize((size_t)cnt * nel);
            HIPC(hipMemcpyAsync(host.data(), buf.p, host.size() * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
            HIPC(hipStreamSynchronize(h->stream));
            for (int b = 0; b < cnt; ++b)
                std::memcpy(out + (size_t)at[(size_t)(c0 + b)] * nel * 2, &host[(size_t)b * nel],
                            (size_t)nel * sizeof(double2));
        }
        if (h->fp32() && n > 0) {
            // tier-0/1 ADOs: their FP64 copies, as heom_get_state returns them
            std::vector<double2> sh((size_t)h->S * nel);
            HIPC(hipMemcpyAsync(sh.data(), h->y64.p, sh.size() * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
            HIPC(hipStreamSynchronize(h->stream));
            // uid -> tier-0/1 index through a sorted copy: O((S + n) log S)
            std::vector<std::pair<int64_t, int>> by_uid((size_t)h->S);
            for (int sp = 0; sp < h->S; ++sp)
                by_uid[(size_t)sp] = {h->sp_uid[(size_t)sp], sp};
            std::sort(by_uid.begin(), by_uid.end());
            for (int64_t a = 0; a < n; ++a) {
                const auto it = std::lower_bound(by_uid.begin(), by_uid.end(), std::make_pair(uid[(size_t)a], -1));
                if (it != by_uid.end() && it->first == uid[(size_t)a])
                    std::memcpy(out + (size_t)a * nel * 2, &sh[(size_t)it->second * nel], (size_t)nel * sizeof(double2));
            }
        }
    });
}

int heom_hierarchy_norm(heom_handle* h, double threshold, double* norms)
{
    return guard([&] {
        need(h);
        if (!norms)
            raise(HEOM_E_INVALID, "norms is NULL");
        if (h->active < h->M)  // ADOs removed by heom_compact have norm 0
            HIPC(hipMemsetAsync(h->norms.p, 0, h->norms.bytes(), h->stream));
        const dim3 grid((unsigned)(h->Mp / kTile)), block(64 * kNormWaves);
        if (h->fp32() && h->paired)
            hipLaunchKernelGGL(hierarchy_norm_kernel<float4>, grid, block, 0, h->stream, h->y.as<float4>(),
                               pair_planes(h->N), h->q2u.p, h->Mp, threshold, h->norms.p, h->mask.p);
        else if (h->fp32())
            hipLaunchKernelGGL(hierarchy_norm_kernel<float2>, grid, block, 0, h->stream, h->y.as<float2>(), h->nel,
                               h->q2u.p, h->Mp, threshold, h->norms.p, h->mask.p);
        else
            hipLaunchKernelGGL(hierarchy_norm_kernel<double2>, grid, block, 0, h->stream, h->y.as<double2>(), h->nel,
                               h->q2u.p, h->Mp, threshold, h->norms.p, h->mask.p);
        HIPC(hipGetLastError());
        if (threshold >= 0.0) {
            refold_tables(h);
            adjoint_invalidate(h);
        }
        HIPC(hipMemcpyAsync(norms, h->norms.p, h->M * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_rhs(heom_handle* h, double* out)
{
    return guard([&] {
        need(h);
        if (!out)
            raise(HEOM_E_INVALID, "out is NULL");
        launch_stage(h, 0, h->y.p, h->ya.p, 0.0);
        download_state(h, h->ya.p, h->sa64.p, out);
    });
}

int heom_step_rk4(heom_handle* h, double dt, int64_t nsteps)
{
    return guard([&] {
        need(h);
        if (nsteps < 0)
            raise(HEOM_E_INVALID, "nsteps must be >= 0");
        if (!(dt > 0.0))
            raise(HEOM_E_INVALID, "step size must be > 0");
        resolve_timing(h);
        HIPC(hipEventRecord(h->ev_start, h->stream));
        rk4_steps(h, dt, nsteps);
        HIPC(hipEventRecord(h->ev_stop, h->stream));
        h->timing_pending = true;
        h->pending_steps = nsteps;
    });
}

int heom_propagate(heom_handle* h, double dt, int64_t nsteps, int64_t observe_steps, double* tops)
{
    return guard([&] {
        need(h);
        if (!tops)
            raise(HEOM_E_INVALID, "tops is NULL");
        if (nsteps < 0 || observe_steps < 1)
            raise(HEOM_E_INVALID, "nsteps must be >= 0 and observe_steps >= 1");
        if (!(dt > 0.0))
            raise(HEOM_E_INVALID, "step size must be > 0");
        const int64_t rows = nsteps / observe_steps + 1;
        const int64_t per = (int64_t)h->count * h->nel;
        DevBuf<double2> rec;
        rec.alloc((size_t)(rows * per));
        resolve_timing(h);
        HIPC(hipEventRecord(h->ev_start, h->stream));
        gather_tops(h, rec.p);
        int64_t row = 1, done = 0;
        while (done < nsteps) {
            const int64_t n = std::min(observe_steps, nsteps - done);
            rk4_steps(h, dt, n);
            done += n;
            if (n == observe_steps) {
                gather_tops(h, rec.p + row * per);
                ++row;
            }
        }
        HIPC(hipEventRecord(h->ev_stop, h->stream));
        h->timing_pending = true;
        h->pending_steps = nsteps;
        HIPC(hipMemcpyAsync(tops, rec.p, rec.bytes(), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        resolve_timing(h);
    });
}

// The reference's e_field pre-evaluate action (dm-heom/include/heom/e_field.hpp:64-87)
// sets H before every RHS evaluation to its value at that evaluation's time; the caller
// passes those Hamiltonians for the RK4 stage times t, t + h/2, t + h/2, t + h.  They are
// uploaded in chunks of steps as H / hbar, and each stage launch reads its own.
int heom_propagate_timedep(heom_handle* h, double dt, int64_t nsteps, int64_t observe_steps,
                           const double* stage_h, double* tops)
{
    return guard([&] {
        need(h);
        reject_tss(h, "heom_propagate_timedep");
        if (!tops || !stage_h)
            raise(HEOM_E_INVALID, "tops and stage_hamiltonians must not be NULL");
        if (nsteps < 0 || observe_steps < 1)
            raise(HEOM_E_INVALID, "nsteps must be >= 0 and observe_steps >= 1");
        if (!(dt > 0.0))
            raise(HEOM_E_INVALID, "step size must be > 0");
        const int64_t rows = nsteps / observe_steps + 1;
        const int64_t per = (int64_t)h->count * h->nel;  // one stage's Hamiltonians
        const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nsteps, (64 << 20) / (4 * per * 16)));
        DevBuf<double2> rec, hsd;
        DevBuf<float2> hsd32;
        rec.alloc((size_t)(rows * per));
        hsd.alloc((size_t)(chunk * 4 * per));
        if (h->fp32())
            hsd32.alloc((size_t)(chunk * 4 * per));
        std::vector<double2> hb((size_t)(chunk * 4 * per));
        std::vector<float2> hb32(h->fp32() ? hb.size() : 0);
        bool real = true;
        for (int64_t t = 0; t < nsteps * 4 * per; ++t)
            if (stage_h[2 * t + 1] != 0.0)
                real = false;
        const bool real_kept = h->real_h;
        h->real_h = real;  // the kernels read the stage Hamiltonians only
        resolve_timing(h);
        try {
            gather_tops(h, rec.p);
            HIPC(hipEventRecord(h->ev_start, h->stream));
            int64_t row = 1;
            for (int64_t c0 = 0; c0 < nsteps; c0 += chunk) {
                const int64_t nc = std::min(chunk, nsteps - c0);
                HIPC(hipStreamSynchronize(h->stream));  // the previous chunk's Hamiltonians are consumed
                for (int64_t t = 0; t < nc * 4 * per; ++t) {
                    const double* src = stage_h + 2 * (c0 * 4 * per + t);
                    hb[t] = make_double2(src[0] / kHbar, src[1] / kHbar);
                    if (h->fp32())
                        hb32[t] = make_float2((float)hb[t].x, (float)hb[t].y);
                }
                HIPC(hipMemcpyAsync(hsd.p, hb.data(), (size_t)(nc * 4 * per) * sizeof(double2),
                                    hipMemcpyHostToDevice, h->stream));
                if (h->fp32())
                    HIPC(hipMemcpyAsync(hsd32.p, hb32.data(), (size_t)(nc * 4 * per) * sizeof(float2),
                                        hipMemcpyHostToDevice, h->stream));
                for (int64_t n = 0; n < nc; ++n) {
                    const double c1[4] = {0.5 * dt, 0.5 * dt, dt, dt / 6.0};
                    unsigned char* in[4] = {h->y.p, h->ya.p, h->yb.p, h->ya.p};
                    unsigned char* out[4] = {h->ya.p, h->yb.p, h->ya.p, h->y.p};
                    for (int s = 0; s < 4; ++s) {
                        h->hs_stage = hsd.p + (n * 4 + s) * per;
                        h->hs32_stage = h->fp32() ? hsd32.p + (n * 4 + s) * per : nullptr;
                        launch_stage(h, s + 1, in[s], out[s], c1[s], (s & 1) != 0);
                    }
                    h->hs_stage = nullptr;
                    h->hs32_stage = nullptr;
                    if ((c0 + n + 1) % observe_steps == 0)
                        gather_tops(h, rec.p + (row++) * per);
                }
            }
            HIPC(hipEventRecord(h->ev_stop, h->stream));
            h->timing_pending = true;
            h->pending_steps = nsteps;
            HIPC(hipMemcpyAsync(tops, rec.p, rec.bytes(), hipMemcpyDeviceToHost, h->stream));
            HIPC(hipStreamSynchronize(h->stream));
        } catch (...) {
            h->hs_stage = nullptr;
            h->hs32_stage = nullptr;
            h->real_h = real_kept;
            throw;
        }
        h->real_h = real_kept;
        resolve_timing(h);
    });
}

int heom_observe(heom_handle* h, const double* ops, double* pops, double* traces)
{
    return guard([&] {
        need(h);
        if (!pops)
            raise(HEOM_E_INVALID, "pops is NULL");
        if (ops && !traces)
            raise(HEOM_E_INVALID, "traces is NULL while ops is given");
        DevBuf<double> p;
        DevBuf<double2> t, o;
        p.alloc((size_t)h->count * h->N);
        if (ops) {
            t.alloc((size_t)h->count);
            o.alloc((size_t)h->count * h->nel);
            HIPC(hipMemcpyAsync(o.p, ops, o.bytes(), hipMemcpyHostToDevice, h->stream));
        }
        observe_launch(h, ops ? o.p : nullptr, p.p, ops ? t.p : nullptr);
        HIPC(hipMemcpyAsync(pops, p.p, p.bytes(), hipMemcpyDeviceToHost, h->stream));
        if (ops)
            HIPC(hipMemcpyAsync(traces, t.p, t.bytes(), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_propagate_observe(heom_handle* h, double dt, int64_t nsteps, int64_t observe_steps, const double* ops,
                           double* pops, double* traces)
{
    return guard([&] {
        need(h);
        check_observe_args(h, dt, nsteps, observe_steps, ops, pops, traces);
        const int64_t rows = nsteps / observe_steps + 1;
        DevBuf<double> prec;
        DevBuf<double2> trec, opsd;
        prec.alloc((size_t)(rows * h->count * h->N));
        if (ops) {
            trec.alloc((size_t)(rows * h->count));
            opsd.alloc((size_t)h->count * h->nel);
        }
        propagate_observe_into(h, dt, nsteps, observe_steps, ops, pops, traces, opsd.p, prec.p, trec.p, true);
    });
}

int heom_propagate_observe_noalloc(heom_handle* h, double dt, int64_t nsteps, int64_t observe_steps,
                                   const double* ops, double* pops, double* traces)
{
    return guard([&] {
        need(h);
        check_observe_args(h, dt, nsteps, observe_steps, ops, pops, traces);
        const int64_t rows = nsteps / observe_steps + 1;
        // scratch layout: ops [count][nel] double2, then the traces [rows][count] double2,
        // then the populations [rows][count][N] double (each part 256-byte aligned)
        auto up = [](int64_t b) { return (b + 255) / 256 * 256; };
        const int64_t ops_b = ops ? up((int64_t)h->count * h->nel * (int64_t)sizeof(double2)) : 0;
        const int64_t tr_b = ops ? up(rows * h->count * (int64_t)sizeof(double2)) : 0;
        const int64_t po_b = up(rows * h->count * h->N * (int64_t)sizeof(double));
        if (ops_b + tr_b + po_b > (int64_t)kObserveScratchBytes)
            raise(HEOM_E_UNSUPPORTED,
                  "heom_propagate_observe_noalloc: " + std::to_string(rows) + " observations of " +
                      std::to_string(h->count) + " instances need " + std::to_string(ops_b + tr_b + po_b) +
                      " bytes of scratch, more than its " + std::to_string(kObserveScratchBytes) +
                      " (split the run into calls with fewer steps)");
        unsigned char* base = observe_scratch(h->device);
        double2* opsd = ops ? reinterpret_cast<double2*>(base) : nullptr;
        double2* trec = ops ? reinterpret_cast<double2*>(base + ops_b) : nullptr;
        double* prec = reinterpret_cast<double*>(base + ops_b + tr_b);
        propagate_observe_into(h, dt, nsteps, observe_steps, ops, pops, traces, opsd, prec, trec, false);
    });
}

int heom_use_graphs(heom_handle* h, int enable)
{
    return guard([&] {
        need(h);
        h->use_graphs = enable != 0;
        if (!h->use_graphs)
            drop_graph(h);
    });
}

int heom_mmult_left(heom_handle* h, const double* mx)
{
    return guard([&] {
        need(h);
        if (!mx)
            raise(HEOM_E_INVALID, "mx is NULL");
        const int n = h->count * h->nel;
        HIPC(hipMemcpyAsync(h->mx.p, mx, n * sizeof(double2), hipMemcpyHostToDevice, h->stream));
        if (h->fp32()) {
            mmult_left_planes<float2>(h);
            const int m = h->S * h->nel;
            hipLaunchKernelGGL(mmult_left_rows, dim3(blocks_for(m, 256)), dim3(256), 0, h->stream, h->y64.p,
                               h->sp_tmp.p, h->mx.p, h->sp_inst.p, h->S, h->N);
            HIPC(hipMemcpyAsync(h->y64.p, h->sp_tmp.p, h->y64.bytes(), hipMemcpyDeviceToDevice, h->stream));
        } else {
            mmult_left_planes<double2>(h);
        }
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_sync(heom_handle* h)
{
    return guard([&] {
        need(h);
        HIPC(hipStreamSynchronize(h->stream));
        resolve_timing(h);
    });
}

int heom_timing(heom_handle* h, heom_timing_info* info)
{
    return guard([&] {
        need(h);
        if (!info)
            raise(HEOM_E_INVALID, "info is NULL");
        resolve_timing(h);
        *info = h->timing;
    });
}

int heom_device_bytes(heom_handle* h, int64_t* bytes)
{
    return guard([&] {
        need(h);
        if (!bytes)
            raise(HEOM_E_INVALID, "bytes is NULL");
        int64_t b = h->y.bytes() + h->pbuf.bytes() + h->ya.bytes() + h->yb.bytes();
        for (auto* x : {&h->pfss, &h->hs, &h->gam, &h->mx, &h->small, &h->y64, &h->sa64, &h->sb64, &h->p64,
                        &h->sp_pfss, &h->sp_tmp})
            b += x->bytes();
        b += h->nb_raw.bytes() + h->nb.bytes() + h->q2u.bytes() + h->tile_inst.bytes() + h->q_top.bytes() +
             h->mask.bytes() + h->terms.bytes() + h->norms.bytes() + h->pfss32.bytes() + h->hs32.bytes() +
             h->gam32.bytes() + h->terms32.bytes() + h->sp_q.bytes() + h->sp_inst.bytes() + h->sp_top_d.bytes() +
             h->sp_nb.bytes();
        *bytes = b;
    });
}

int heom_rk4_step_bytes(heom_handle* h, int64_t* bytes)
{
    return guard([&] {
        need(h);
        if (!bytes)
            raise(HEOM_E_INVALID, "bytes is NULL");
        *bytes = (int64_t)16 * h->active * h->nel * (int64_t)h->cbytes();
    });
}

int heom_stage_kernel(heom_handle* h, int* kind)
{
    return guard([&] {
        need(h);
        if (!kind)
            raise(HEOM_E_INVALID, "kind is NULL");
        *kind = h->N > kMaxStates ? HEOM_STAGE_WIDE : h->pipelined ? HEOM_STAGE_PIPELINED : HEOM_STAGE_TILE;
    });
}

// Hierarchy compaction: the device layout is rebuilt with only the ADOs whose mask is
// not FALSE.  A masked ADO is never read by its neighbours (heom_ode.cl:199,237,297,360),
// so the unmasked ADOs evolve exactly as before and every observable of them is
// unchanged; the masked ones no longer cost bytes.  Each instance keeps its ADOs in the
// same relative order, padded to 64 slots; edges to removed ADOs read the zero slot.
int heom_compact(heom_handle* h, int64_t* active)
{
    return guard([&] {
        need(h);
        reject_tss(h, "heom_compact");
        if (h->fp32())
            raise(HEOM_E_UNSUPPORTED, "compaction needs FP64 storage");
        const int W = h->W, nel = h->nel;
        const int64_t Mp = h->Mp;
        std::vector<uint8_t> mask(h->M);
        std::vector<int32_t> q2u(Mp);
        std::vector<int2> nb((size_t)(W + 2) * Mp);
        std::vector<double2> pfss(Mp);
        HIPC(hipMemcpyAsync(mask.data(), h->mask.p, h->M, hipMemcpyDeviceToHost, h->stream));
        HIPC(hipMemcpyAsync(q2u.data(), h->q2u.p, q2u.size() * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipMemcpyAsync(nb.data(), h->nb_raw.p, nb.size() * sizeof(int2), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipMemcpyAsync(pfss.data(), h->pfss.p, pfss.size() * sizeof(double2), hipMemcpyDeviceToHost,
                            h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        std::vector<int64_t> qtop(h->count);
        HIPC(hipMemcpy(qtop.data(), h->q_top.p, qtop.size() * sizeof(int64_t), hipMemcpyDeviceToHost));

        // new slots, instance by instance
        std::vector<int32_t> o2n(Mp, -1);
        std::vector<int64_t> new_q(h->count), new_n(h->count);
        int64_t q0 = 0, kept = 0;
        auto kept_slot = [&](int64_t q) { return q != h->zero_slot && q2u[q] >= 0 && mask[q2u[q]] != 0; };
        std::vector<uint8_t> has_plus(Mp, 0);
        for (int m = 0; m < W; ++m)
            for (int64_t q = 0; q < Mp; ++q)
                if (kept_slot((int32_t)((uint32_t)nb[(size_t)m * Mp + q].x >> 6)))
                    has_plus[q] = 1;
        for (int k = 0; k < h->count; ++k) {
            new_q[k] = q0;
            const int64_t lo = h->inst_q[k], hi = h->inst_q[k] + h->inst_n[k];
            std::vector<int32_t> seq;
            for (int64_t q = lo; q < hi; ++q)
                if (kept_slot(q))
                    seq.push_back((int32_t)q);
            const int64_t n = (int64_t)seq.size();
            const std::vector<int32_t> grouped = group_tiles(seq, has_plus);
            for (int64_t r = 0; r < n; ++r)
                o2n[grouped[r]] = (int32_t)(q0 + r);
            if (o2n[qtop[k]] < 0)
                raise(HEOM_E_INVALID, "the top ADO of every instance must stay unmasked");
            kept += n;
            new_n[k] = n;
            q0 += k + 1 < h->count ? (n + kTile - 1) / kTile * kTile : n;
        }
        const int32_t zero = (int32_t)q0;
        const int64_t mp = (q0 + 1 + kTile - 1) / kTile * kTile;
        const int ntiles = (int)(mp / kTile);
        std::vector<int32_t> n2o(mp, -1), q2u_n(mp, -1), tinst(ntiles, h->count - 1);
        std::vector<double2> pfss_n(mp, make_double2(0.0, 0.0));
        std::vector<int2> nb_n((size_t)(W + 2) * mp, make_int2((int32_t)((uint32_t)zero << 6), zero));
        for (int64_t q = 0; q < Mp; ++q) {
            const int32_t qn = o2n[q];
            if (qn < 0)
                continue;
            n2o[qn] = (int32_t)q;
            q2u_n[qn] = q2u[q];
            pfss_n[qn] = pfss[q];
            for (int m = 0; m < W; ++m) {
                const int2 r = nb[(size_t)m * Mp + q];
                const int32_t p = (int32_t)((uint32_t)r.x >> 6), nm = r.x & 63;
                const int32_t pn = p == h->zero_slot ? zero : (o2n[p] >= 0 ? o2n[p] : zero);
                const int32_t mn = r.y == h->zero_slot ? zero : (o2n[r.y] >= 0 ? o2n[r.y] : zero);
                nb_n[(size_t)m * mp + qn] = make_int2((int32_t)(((uint32_t)pn << 6) | (uint32_t)nm), mn);
            }
        }
        for (int k = 0; k < h->count; ++k) {
            for (int64_t t = new_q[k] / kTile; t < (new_q[k] + new_n[k] + kTile - 1) / kTile; ++t)
                tinst[t] = k;
            qtop[k] = o2n[qtop[k]];
        }

        // move the state and swap in the new layout
        DevBuf<int32_t> n2o_d;
        upload(n2o_d, n2o, h->stream);
        DevBuf<unsigned char> y_new;
        y_new.alloc((size_t)nel * mp * sizeof(double2));
        hipLaunchKernelGGL(remap_planes, dim3(blocks_for((int64_t)nel * mp, 256)), dim3(256), 0, h->stream,
                           h->y.as<double2>(), y_new.as<double2>(), n2o_d.p, Mp, mp, nel);
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(h->stream));
        std::swap(h->y.p, y_new.p);
        std::swap(h->y.n, y_new.n);
        for (auto* b : {&h->pbuf, &h->ya, &h->yb}) {
            b->alloc((size_t)nel * mp * sizeof(double2));
            HIPC(hipMemsetAsync(b->p, 0, b->bytes(), h->stream));
        }
        upload(h->nb_raw, nb_n, h->stream);
        upload(h->nb, nb_n, h->stream);
        upload(h->q2u, q2u_n, h->stream);
        upload(h->tile_inst, tinst, h->stream);
        upload(h->q_top, qtop, h->stream);
        h->Mp = mp;
        h->ntiles = ntiles;
        h->zero_slot = zero;
        h->active = kept;
        reset_extra_buffers(h);  // added state buffers follow the new layout, zeroed
        build_tile_plus(h, nb_n);
        upload(h->pfss, pfss_n, h->stream);
        h->inst_q = new_q;
        h->inst_n = new_n;
        std::fill(h->u2q_host.begin(), h->u2q_host.end(), -1);
        for (int64_t qn = 0; qn < mp; ++qn)
            if (q2u_n[qn] >= 0)
                h->u2q_host[(size_t)q2u_n[qn]] = (int32_t)qn;
        refold_tables(h);
        drop_graph(h);
        adjoint_invalidate(h);
        choose_stage_kernel(h);
        HIPC(hipStreamSynchronize(h->stream));
        if (active)
            *active = kept;
    });
}

// The autotuner: every valid stage-kernel configuration of the handle -- the kernel
// (tile, pipelined, or for N > 8 the wide kernel), the tiles per XCD chunk of the tile
// order, the ADOs per workgroup of the wide kernel -- is timed over `steps` RK4 steps
// (after one untimed step) on the handle's own state, which is restored afterwards; the
// fastest is kept.  The HIP graph of a step is rebuilt for it.
int heom_autotune(heom_handle* h, int32_t steps, heom_tune_result* results, int32_t capacity, int32_t* count)
{
    return guard([&] {
        need(h);
        reject_tss(h, "heom_autotune");
        if (steps < 1)
            raise(HEOM_E_INVALID, "steps must be >= 1");
        if (capacity > 0 && !results)
            raise(HEOM_E_INVALID, "results is NULL");
        const std::vector<StageConfig> cands = stage_candidates(h);
        // the state (and with FP32 storage its FP64 tier shadow) is restored after each trial
        DevBuf<unsigned char> save, save64;
        save.alloc(h->y.bytes());
        HIPC(hipMemcpyAsync(save.p, h->y.p, h->y.bytes(), hipMemcpyDeviceToDevice, h->stream));
        if (h->fp32()) {
            save64.alloc(h->y64.bytes());
            HIPC(hipMemcpyAsync(save64.p, h->y64.p, h->y64.bytes(), hipMemcpyDeviceToDevice, h->stream));
        }
        resolve_timing(h);
        const heom_timing_info timing_kept = h->timing;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPC(hipEventCreate(&e0));
        HIPC(hipEventCreate(&e1));
        std::vector<double> ms(cands.size());
        const double dt = 1e-15;
        auto apply = [&](const StageConfig& c) { apply_stage_config(h, c); };
        try {
            for (size_t k = 0; k < cands.size(); ++k) {
                apply(cands[k]);
                rk4_step_launches(h, dt);
                HIPC(hipEventRecord(e0, h->stream));
                for (int s = 0; s < steps; ++s)
                    rk4_step_launches(h, dt);
                HIPC(hipEventRecord(e1, h->stream));
                HIPC(hipEventSynchronize(e1));
                float t = 0.f;
                HIPC(hipEventElapsedTime(&t, e0, e1));
                ms[k] = t / steps;
                HIPC(hipMemcpyAsync(h->y.p, save.p, h->y.bytes(), hipMemcpyDeviceToDevice, h->stream));
                if (h->fp32())
                    HIPC(hipMemcpyAsync(h->y64.p, save64.p, h->y64.bytes(), hipMemcpyDeviceToDevice, h->stream));
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const size_t best = (size_t)(std::min_element(ms.begin(), ms.end()) - ms.begin());
        apply(cands[best]);
        HIPC(hipStreamSynchronize(h->stream));
        h->timing = timing_kept;
        if (count)
            *count = (int32_t)cands.size();
        for (size_t k = 0; k < cands.size() && (int32_t)k < capacity; ++k) {
            results[k].kernel = cands[k].kind;
            results[k].xcd_chunk = cands[k].chunk;
            results[k].workgroups = cands[k].wgs;
            results[k].ados_per_workgroup = cands[k].ados;
            results[k].tile_split = cands[k].split;
            results[k].chosen = k == best ? 1 : 0;
            results[k].ms_per_step = ms[k];
        }
    });
}

int heom_set_stage_config(heom_handle* h, const heom_tune_result* config)
{
    return guard([&] {
        need(h);
        reject_tss(h, "heom_set_stage_config");
        if (!config)
            raise(HEOM_E_INVALID, "config is NULL");
        const StageConfig c{config->kernel, config->xcd_chunk, config->workgroups, config->ados_per_workgroup,
                            config->tile_split < 1 ? 1 : config->tile_split};
        bool ok = false;
        for (const StageConfig& v : stage_candidates(h))
            if (v.kind == c.kind && v.ados == c.ados && v.split == c.split &&
                (c.kind != HEOM_STAGE_PIPELINED || v.wgs == c.wgs))
                ok = true;
        if (!ok)
            raise(HEOM_E_INVALID, "stage configuration (kernel " + std::to_string(c.kind) + ", " +
                                      std::to_string(c.ados) + " ADOs per workgroup, " + std::to_string(c.wgs) +
                                      " workgroups) is not valid for this handle");
        if (c.chunk < 1 || c.chunk > h->ntiles)
            raise(HEOM_E_INVALID, "xcd_chunk must lie in [1, " + std::to_string(h->ntiles) + "]");
        apply_stage_config(h, c);
    });
}

}  // extern "C"
