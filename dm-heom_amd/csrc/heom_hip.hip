// heom_hip.hip -- MI355X (gfx950) HEOM solver: device kernels and the C ABI of
// include/heom_hip.h.
//
// Device layout (one hierarchy, FP64):
//   * ADOs are renumbered into lexicographic order of their tuples (mode 0 most
//     significant).  Neighbours n +- e_m then sit close in index: at 735k ADOs four
//     fifths of all neighbour reads land within 30k ADOs, against one in twenty in the
//     reference's layer order, so most of them are served from L2 / MALL.
//   * State planes are structure-of-arrays by matrix element: Y[e * Mp + q] holds
//     element e = i*N + j of ADO q as a double2 (re, im).  Mp = M rounded up to a
//     multiple of the 64-ADO tile; padding ADOs have no neighbours and stay zero.
//   * Neighbour tables are SoA by mode: plus[m * Mp + q], minus[m * Mp + q] (int32,
//     -1 = no edge or neighbour masked), tup[m * Mp + q] (uint8, n_m).
//
// Stage kernel (rk4_stage): one workgroup = one tile of 64 ADOs, one wavefront per
// matrix row i (N wavefronts), lane = ADO.  Each lane loads row i of its ADO (one
// coalesced 1 KiB load per element plane), publishes it to an LDS copy of the tile,
// and computes row i of f(sigma): the commutator (row i of sigma from registers, the
// columns from LDS, H/hbar through scalar loads), the "same" term, and the neighbour
// terms as coalesced gathers from the element planes.  The classical RK4 combination
// of the stage is applied in the epilogue, so a step is four launches and 16 state
// passes (R y, W Y2, W A | R Y2, R y, R A, W Y3, W A | ... | R Y4, R A, W y).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <new>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "heom_hip.h"

namespace {

constexpr int kTile = 64;      // ADOs per workgroup (one per lane)
constexpr int kMaxStates = 8;  // N handled by the stage kernel
constexpr int kNumXcd = 8;     // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr double kHbar = 1.05457148e-34;  // cl/heom/common.cl:45

thread_local std::string g_error;

struct heom_error : std::runtime_error {
    int code;
    heom_error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void raise(int code, const std::string& msg) { throw heom_error(code, msg); }

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        raise(HEOM_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPC(x) hip_check((x), #x)

// ------------------------------------------------------------------------------------
// complex helpers (double2 = re, im)
// ------------------------------------------------------------------------------------

__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c)  // c + a*b
{
    c.x = fma(a.x, b.x, c.x);
    c.x = fma(-a.y, b.y, c.x);
    c.y = fma(a.x, b.y, c.y);
    c.y = fma(a.y, b.x, c.y);
    return c;
}

__device__ __forceinline__ double2 cfms(double2 a, double2 b, double2 c)  // c - a*b
{
    c.x = fma(-a.x, b.x, c.x);
    c.x = fma(a.y, b.y, c.x);
    c.y = fma(-a.x, b.y, c.y);
    c.y = fma(-a.y, b.x, c.y);
    return c;
}

// workgroup index -> tile index so that each XCD works on one contiguous range of
// tiles (neighbour reads then hit that XCD's L2)
__device__ __forceinline__ int xcd_tile(int b, int nt)
{
    const int per = nt / kNumXcd, rem = nt % kNumXcd;
    const int x = b % kNumXcd, k = b / kNumXcd;
    return x * per + min(x, rem) + k;
}

// ------------------------------------------------------------------------------------
// stage kernel
// ------------------------------------------------------------------------------------

struct StageArgs {
    const double2* __restrict__ yin;   // stage input Y_s (neighbours are read from here)
    const double2* __restrict__ y0;    // y_n
    double2* __restrict__ acc;         // RK accumulator A
    double2* __restrict__ yout;        // next stage input, or y_{n+1}, or f
    const int32_t* __restrict__ plus;  // [W][Mp]
    const int32_t* __restrict__ minus; // [W][Mp]
    const uint8_t* __restrict__ tup;   // [W][Mp]
    const double2* __restrict__ pfss;  // [Mp] pf_same_sum
    const double2* __restrict__ hs;    // [N*N] H / hbar
    const double2* __restrict__ gam;   // [N*N] -(sum of Delta_b over the baths of i and j), i != j
    const int32_t* __restrict__ rowm_cnt; // [N]      modes of the baths of state i
    const int32_t* __restrict__ rowm;     // [N][W]
    const int32_t* __restrict__ colm_cnt; // [N*N]    modes of the baths of j not in row i's list
    const int32_t* __restrict__ colm;     // [N*N][W]
    const double* __restrict__ pm;        // [N*N][W] count_i(m) - count_j(m)
    const double2* __restrict__ beta;     // [N*N][W] minus-term constant of (e, m)
    const double* __restrict__ pcoef;     // [W][nv] sqrt((v+1)|c_m|/hbar)
    const double* __restrict__ scoef;     // [W][nv] sqrt(v/|c_m| hbar)
    int64_t mp;
    int32_t width;
    int32_t nv;
    int32_t ntiles;
    double c_out;  // coefficient of f in yout
    double c_acc;  // coefficient of f in acc
};

// STAGE 0: yout = f(yin).  STAGE 1..4: classical RK4 stage s (see file header).
template <int N, int STAGE, bool REALH>
__global__ __launch_bounds__(64 * N) void rk4_stage(StageArgs a)
{
    __shared__ double2 tile_lds[N * N * kTile];
    const int lane = threadIdx.x & 63;
    const int i = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = xcd_tile(blockIdx.x, a.ntiles);
    const int64_t mp = a.mp;
    const int64_t q = (int64_t)tile * kTile + lane;
    const int W = a.width;

    double2 s[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        s[j] = a.yin[(int64_t)(i * N + j) * mp + q];
        tile_lds[(i * N + j) * kTile + lane] = s[j];
    }
    __syncthreads();

    // commutator c = sigma Hs - Hs sigma (row i); f = i c
    double2 c[N];
#pragma unroll
    for (int j = 0; j < N; ++j)
        c[j] = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < N; ++k) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const double2 h = a.hs[k * N + j];
            if (REALH) {
                c[j].x = fma(s[k].x, h.x, c[j].x);
                c[j].y = fma(s[k].y, h.x, c[j].y);
            } else {
                c[j] = cfma(s[k], h, c[j]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double2 h = a.hs[i * N + k];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const double2 skj = tile_lds[(k * N + j) * kTile + lane];
            if (REALH) {
                c[j].x = fma(-h.x, skj.x, c[j].x);
                c[j].y = fma(-h.x, skj.y, c[j].y);
            } else {
                c[j] = cfms(h, skj, c[j]);
            }
        }
    }
    double2 f[N];
#pragma unroll
    for (int j = 0; j < N; ++j)
        f[j] = make_double2(-c[j].y, c[j].x);

    // same term: (S_n + gamma_ij) sigma_ij
    const double2 S = a.pfss[q];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const double2 g = a.gam[i * N + j];
        const double2 t = make_double2(S.x + g.x, S.y + g.y);
        f[j] = cfma(t, s[j], f[j]);
    }

    // neighbour terms of the modes of state i's baths (present in every column j)
    const int nr = a.rowm_cnt[i];
    for (int r = 0; r < nr; ++r) {
        const int m = __builtin_amdgcn_readfirstlane(a.rowm[i * W + r]);
        const int32_t p = a.plus[(int64_t)m * mp + q];
        const int32_t mi = a.minus[(int64_t)m * mp + q];
        const int v = a.tup[(int64_t)m * mp + q];
        const double pf = a.pcoef[m * a.nv + v];
        const double sf = a.scoef[m * a.nv + v];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const int e = i * N + j;
            const double pmul = a.pm[e * W + m];
            const double2 bt = a.beta[e * W + m];
            if (pmul != 0.0 && p >= 0) {
                const double2 x = a.yin[(int64_t)e * mp + p];
                const double cf = pmul * pf;
                f[j].x = fma(-cf, x.y, f[j].x);
                f[j].y = fma(cf, x.x, f[j].y);
            }
            if ((bt.x != 0.0 || bt.y != 0.0) && mi >= 0) {
                const double2 x = a.yin[(int64_t)e * mp + mi];
                f[j] = cfma(make_double2(bt.x * sf, bt.y * sf), x, f[j]);
            }
        }
    }
    // neighbour terms of the modes of state j's baths that state i does not couple to
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int e = i * N + j;
        const int nc = a.colm_cnt[e];
        for (int cidx = 0; cidx < nc; ++cidx) {
            const int m = __builtin_amdgcn_readfirstlane(a.colm[e * W + cidx]);
            const double pmul = a.pm[e * W + m];
            const double2 bt = a.beta[e * W + m];
            const int v = a.tup[(int64_t)m * mp + q];
            if (pmul != 0.0) {
                const int32_t p = a.plus[(int64_t)m * mp + q];
                if (p >= 0) {
                    const double2 x = a.yin[(int64_t)e * mp + p];
                    const double cf = pmul * a.pcoef[m * a.nv + v];
                    f[j].x = fma(-cf, x.y, f[j].x);
                    f[j].y = fma(cf, x.x, f[j].y);
                }
            }
            if (bt.x != 0.0 || bt.y != 0.0) {
                const int32_t mi = a.minus[(int64_t)m * mp + q];
                if (mi >= 0) {
                    const double sf = a.scoef[m * a.nv + v];
                    const double2 x = a.yin[(int64_t)e * mp + mi];
                    f[j] = cfma(make_double2(bt.x * sf, bt.y * sf), x, f[j]);
                }
            }
        }
    }

    // RK4 epilogue
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int64_t idx = (int64_t)(i * N + j) * mp + q;
        if (STAGE == 0) {
            a.yout[idx] = f[j];
        } else if (STAGE == 1) {
            a.yout[idx] = make_double2(fma(a.c_out, f[j].x, s[j].x), fma(a.c_out, f[j].y, s[j].y));
            a.acc[idx] = make_double2(fma(a.c_acc, f[j].x, s[j].x), fma(a.c_acc, f[j].y, s[j].y));
        } else if (STAGE == 2 || STAGE == 3) {
            const double2 y = a.y0[idx];
            const double2 A = a.acc[idx];
            a.yout[idx] = make_double2(fma(a.c_out, f[j].x, y.x), fma(a.c_out, f[j].y, y.y));
            a.acc[idx] = make_double2(fma(a.c_acc, f[j].x, A.x), fma(a.c_acc, f[j].y, A.y));
        } else {
            const double2 A = a.acc[idx];
            a.yout[idx] = make_double2(fma(a.c_acc, f[j].x, A.x), fma(a.c_acc, f[j].y, A.y));
        }
    }
}

// ------------------------------------------------------------------------------------
// layout, reduction and hierarchy-arithmetic kernels
// ------------------------------------------------------------------------------------

// dst[e * mp + q] = src[u(q)][e]   (reference AoS -> internal SoA)
__global__ void aos_to_soa(const double2* __restrict__ src, double2* __restrict__ dst,
                           const int32_t* __restrict__ q2u, int64_t mp, int nel)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= mp)
        return;
    const int32_t u = q2u[q];
    for (int e = 0; e < nel; ++e)
        dst[(int64_t)e * mp + q] = u >= 0 ? src[(int64_t)u * nel + e] : make_double2(0.0, 0.0);
}

// dst[u(q)][e] = src[e * mp + q]   (internal SoA -> reference AoS)
__global__ void soa_to_aos(const double2* __restrict__ src, double2* __restrict__ dst,
                           const int32_t* __restrict__ q2u, int64_t mp, int nel)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= mp)
        return;
    const int32_t u = q2u[q];
    if (u < 0)
        return;
    for (int e = 0; e < nel; ++e)
        dst[(int64_t)u * nel + e] = src[(int64_t)e * mp + q];
}

// dst[e] = src[e * mp + q0] for one ADO (the top), e < nel
__global__ void copy_ado_out(const double2* __restrict__ src, double2* __restrict__ dst,
                             int64_t q0, int64_t mp, int nel)
{
    const int e = threadIdx.x;
    if (e < nel)
        dst[e] = src[(int64_t)e * mp + q0];
}

__global__ void copy_ado_in(const double2* __restrict__ src, double2* __restrict__ dst,
                            int64_t q0, int64_t mp, int nel)
{
    const int e = threadIdx.x;
    if (e < nel)
        dst[(int64_t)e * mp + q0] = src[e];
}

// per-ADO squared Frobenius norm in uid order; threshold masking as in
// cl/heom/hierarchy_norm.cl:41-44 (mask in uid order)
__global__ void hierarchy_norm_kernel(const double2* __restrict__ y, const int32_t* __restrict__ q2u,
                                      int64_t mp, int nel, double threshold,
                                      double* __restrict__ norms, uint8_t* __restrict__ mask)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= mp)
        return;
    const int32_t u = q2u[q];
    if (u < 0)
        return;
    double acc = 0.0;
    for (int e = 0; e < nel; ++e) {
        const double2 v = y[(int64_t)e * mp + q];
        acc = fma(v.x, v.x, acc);
        acc = fma(v.y, v.y, acc);
    }
    norms[u] = acc;
    if (acc - threshold <= 0.0 && mask[u] != 2)
        mask[u] = 0;
}

// fold the uid-order mask into the neighbour tables: an edge whose target is masked
// becomes -1 (the kernel then skips it, as heom_ode.cl does with its mask checks)
__global__ void fold_mask(const int32_t* __restrict__ raw, int32_t* __restrict__ folded,
                          const uint8_t* __restrict__ mask, const int32_t* __restrict__ q2u,
                          int64_t count)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count)
        return;
    const int32_t p = raw[t];
    folded[t] = (p >= 0 && mask[q2u[p]] != 0) ? p : -1;
}

// sigma_q <- Mx sigma_q for every ADO: wave = row i, lane = ADO, tile in LDS
template <int N>
__global__ __launch_bounds__(64 * N) void mmult_left_kernel(double2* __restrict__ y,
                                                            const double2* __restrict__ mx,
                                                            int64_t mp, int ntiles)
{
    __shared__ double2 t[N * N * kTile];
    const int lane = threadIdx.x & 63;
    const int i = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t q = (int64_t)blockIdx.x * kTile + lane;
#pragma unroll
    for (int j = 0; j < N; ++j)
        t[(i * N + j) * kTile + lane] = y[(int64_t)(i * N + j) * mp + q];
    __syncthreads();
    double2 out[N];
#pragma unroll
    for (int j = 0; j < N; ++j)
        out[j] = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double2 m = mx[i * N + k];
#pragma unroll
        for (int j = 0; j < N; ++j)
            out[j] = cfma(m, t[(k * N + j) * kTile + lane], out[j]);
    }
#pragma unroll
    for (int j = 0; j < N; ++j)
        y[(int64_t)(i * N + j) * mp + q] = out[j];
}

// ------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { free(); }
    void alloc(size_t count)
    {
        free();
        if (count == 0)
            count = 1;
        HIPC(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void free()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

}  // namespace

struct heom_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    int N = 0, K = 0, B = 0, BM = 0, W = 0, nel = 0, nv = 0;
    int64_t M = 0, Mp = 0;
    int ntiles = 0;
    bool real_h = false;
    int64_t q_top = 0;  // internal index of uid 0

    std::vector<int32_t> q2u_host;  // [Mp]
    std::vector<int32_t> coupling;  // [N * BM]
    std::vector<double> cbk;        // [2 * B * K]
    std::vector<double> pf_same;    // [2 * B]
    std::vector<double> pf_circ;    // [2 * B]

    DevBuf<double2> y, acc, ya, yb;
    DevBuf<int32_t> plus_raw, minus_raw, plus, minus, q2u;
    DevBuf<uint8_t> tup, mask;
    DevBuf<double2> pfss, hs, gam, beta, mx, small;
    DevBuf<int32_t> rowm_cnt, rowm, colm_cnt, colm;
    DevBuf<double> pm, pcoef, scoef, norms;

    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    bool timing_pending = false;
    int64_t pending_steps = 0;
    heom_timing_info timing{};

    ~heom_handle()
    {
        if (stream)
            (void)hipStreamSynchronize(stream);
        for (auto* b : {&y, &acc, &ya, &yb, &pfss, &hs, &gam, &beta, &mx, &small})
            b->free();
        for (auto* b : {&plus_raw, &minus_raw, &plus, &minus, &q2u, &rowm_cnt, &rowm, &colm_cnt, &colm})
            b->free();
        tup.free();
        mask.free();
        pm.free();
        pcoef.free();
        scoef.free();
        norms.free();
        if (ev_start)
            (void)hipEventDestroy(ev_start);
        if (ev_stop)
            (void)hipEventDestroy(ev_stop);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

namespace {

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

StageArgs stage_args(heom_handle* h)
{
    StageArgs a{};
    a.plus = h->plus.p;
    a.minus = h->minus.p;
    a.tup = h->tup.p;
    a.pfss = h->pfss.p;
    a.hs = h->hs.p;
    a.gam = h->gam.p;
    a.rowm_cnt = h->rowm_cnt.p;
    a.rowm = h->rowm.p;
    a.colm_cnt = h->colm_cnt.p;
    a.colm = h->colm.p;
    a.pm = h->pm.p;
    a.beta = h->beta.p;
    a.pcoef = h->pcoef.p;
    a.scoef = h->scoef.p;
    a.mp = h->Mp;
    a.width = h->W;
    a.nv = h->nv;
    a.ntiles = h->ntiles;
    return a;
}

template <int N, bool REALH>
void launch_stage_n(int stage, const StageArgs& a, int ntiles, hipStream_t s)
{
    dim3 grid(ntiles), block(64 * N);
    switch (stage) {
    case 0: hipLaunchKernelGGL((rk4_stage<N, 0, REALH>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((rk4_stage<N, 1, REALH>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((rk4_stage<N, 2, REALH>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((rk4_stage<N, 3, REALH>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((rk4_stage<N, 4, REALH>), grid, block, 0, s, a); break;
    }
}

template <bool REALH>
void launch_stage_r(int N, int stage, const StageArgs& a, int ntiles, hipStream_t s)
{
    switch (N) {
    case 1: launch_stage_n<1, REALH>(stage, a, ntiles, s); break;
    case 2: launch_stage_n<2, REALH>(stage, a, ntiles, s); break;
    case 3: launch_stage_n<3, REALH>(stage, a, ntiles, s); break;
    case 4: launch_stage_n<4, REALH>(stage, a, ntiles, s); break;
    case 5: launch_stage_n<5, REALH>(stage, a, ntiles, s); break;
    case 6: launch_stage_n<6, REALH>(stage, a, ntiles, s); break;
    case 7: launch_stage_n<7, REALH>(stage, a, ntiles, s); break;
    case 8: launch_stage_n<8, REALH>(stage, a, ntiles, s); break;
    default: raise(HEOM_E_UNSUPPORTED, "states > 8 not supported by this build");
    }
}

void launch_stage(heom_handle* h, int stage, const double2* yin, double2* yout, double c_out,
                  double c_acc)
{
    StageArgs a = stage_args(h);
    a.yin = yin;
    a.y0 = h->y.p;
    a.acc = h->acc.p;
    a.yout = yout;
    a.c_out = c_out;
    a.c_acc = c_acc;
    if (h->real_h)
        launch_stage_r<true>(h->N, stage, a, h->ntiles, h->stream);
    else
        launch_stage_r<false>(h->N, stage, a, h->ntiles, h->stream);
    HIPC(hipGetLastError());
    h->timing.rhs_launches += 1;
}

void rk4_steps(heom_handle* h, double dt, int64_t nsteps)
{
    for (int64_t s = 0; s < nsteps; ++s) {
        launch_stage(h, 1, h->y.p, h->ya.p, 0.5 * dt, dt / 6.0);
        launch_stage(h, 2, h->ya.p, h->yb.p, 0.5 * dt, dt / 3.0);
        launch_stage(h, 3, h->yb.p, h->ya.p, dt, dt / 3.0);
        launch_stage(h, 4, h->ya.p, h->y.p, 0.0, dt / 6.0);
    }
}

void refold_tables(heom_handle* h)
{
    const int64_t count = (int64_t)h->W * h->Mp;
    const int threads = 256;
    const int blocks = (int)((count + threads - 1) / threads);
    hipLaunchKernelGGL(fold_mask, dim3(blocks), dim3(threads), 0, h->stream, h->plus_raw.p,
                       h->plus.p, h->mask.p, h->q2u.p, count);
    hipLaunchKernelGGL(fold_mask, dim3(blocks), dim3(threads), 0, h->stream, h->minus_raw.p,
                       h->minus.p, h->mask.p, h->q2u.p, count);
    HIPC(hipGetLastError());
}

void set_hamiltonian(heom_handle* h, const double* H)
{
    const int N = h->N;
    std::vector<double2> hs(N * N);
    bool real = true;
    for (int e = 0; e < N * N; ++e) {
        hs[e] = make_double2(H[2 * e] / kHbar, H[2 * e + 1] / kHbar);
        if (H[2 * e + 1] != 0.0)
            real = false;
    }
    if (h->hs.p == nullptr)
        h->hs.alloc(N * N);
    HIPC(hipMemcpyAsync(h->hs.p, hs.data(), hs.size() * sizeof(double2), hipMemcpyHostToDevice,
                        h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->real_h = real;
}

// Per-element coefficient tables of the neighbour and "same" terms.  For element
// (i, j) and mode m of bath b = m / K, heom_ode.cl adds, for every coupling-list entry
// of state i equal to b, +i sqrt((n_m+1)|c_m|/hbar) sigma^{n+m} and
// (+i c_m/hbar + [k=0] Phi_b) sqrt(n_m hbar/|c_m|) sigma^{n-m}; for every entry of
// state j equal to b the same with -i (heom_ode.cl:184-266, 280-405).
void build_coefficients(heom_handle* h)
{
    const int N = h->N, W = h->W, K = h->K, BM = h->BM, B = h->B;
    auto count = [&](int s, int b) {
        int c = 0;
        for (int t = 0; t < BM; ++t)
            if (h->coupling[s * BM + t] == b)
                ++c;
        return c;
    };
    std::vector<int32_t> rowm_cnt(N, 0), rowm(N * W, 0), colm_cnt(N * N, 0), colm(N * N * W, 0);
    std::vector<double> pm(N * N * W, 0.0);
    std::vector<double2> beta(N * N * W, make_double2(0.0, 0.0)), gam(N * N, make_double2(0.0, 0.0));
    for (int i = 0; i < N; ++i) {
        for (int m = 0; m < W; ++m)
            if (count(i, m / K) > 0)
                rowm[i * W + rowm_cnt[i]++] = m;
        for (int j = 0; j < N; ++j) {
            const int e = i * N + j;
            for (int m = 0; m < W; ++m) {
                const int b = m / K, k = m % K;
                const int ci = count(i, b), cj = count(j, b);
                if (ci == 0 && cj > 0)
                    colm[e * W + colm_cnt[e]++] = m;
                pm[e * W + m] = (double)(ci - cj);
                const double cr = h->cbk[2 * m], cim = h->cbk[2 * m + 1];
                // i c_m / hbar (ci - cj)
                double br = -cim / kHbar * (ci - cj);
                double bi = cr / kHbar * (ci - cj);
                if (k == 0) {
                    br += h->pf_circ[2 * b] * (ci + cj);
                    bi += h->pf_circ[2 * b + 1] * (ci + cj);
                }
                beta[e * W + m] = make_double2(br, bi);
            }
            if (i != j) {
                double gr = 0.0, gi = 0.0;
                for (int t = 0; t < BM; ++t) {
                    const int bi = h->coupling[i * BM + t];
                    if (bi >= 0) {
                        gr -= h->pf_same[2 * bi];
                        gi -= h->pf_same[2 * bi + 1];
                    }
                    const int bj = h->coupling[j * BM + t];
                    if (bj >= 0) {
                        gr -= h->pf_same[2 * bj];
                        gi -= h->pf_same[2 * bj + 1];
                    }
                }
                gam[e] = make_double2(gr, gi);
            }
        }
    }
    (void)B;
    // sqrt factors exactly as heom_ode.cl:211,249 and :321,383 evaluate them
    std::vector<double> pcoef(W * h->nv, 0.0), scoef(W * h->nv, 0.0);
    for (int m = 0; m < W; ++m) {
        const double a = std::hypot(h->cbk[2 * m], h->cbk[2 * m + 1]);
        for (int v = 0; v < h->nv; ++v) {
            pcoef[m * h->nv + v] = std::sqrt((v + 1.0) * a / kHbar);
            scoef[m * h->nv + v] = v > 0 ? std::sqrt(v / a * kHbar) : 0.0;
        }
    }
    upload(h->rowm_cnt, rowm_cnt, h->stream);
    upload(h->rowm, rowm, h->stream);
    upload(h->colm_cnt, colm_cnt, h->stream);
    upload(h->colm, colm, h->stream);
    upload(h->pm, pm, h->stream);
    upload(h->beta, beta, h->stream);
    upload(h->gam, gam, h->stream);
    upload(h->pcoef, pcoef, h->stream);
    upload(h->scoef, scoef, h->stream);
}

void validate(const heom_desc* d)
{
    if (!d)
        raise(HEOM_E_INVALID, "desc is NULL");
    if (d->states < 1)
        raise(HEOM_E_INVALID, "states must be > 0");
    if (d->states > kMaxStates)
        raise(HEOM_E_UNSUPPORTED, "states > 8 not supported by this build");
    if (d->matsubaras < 1 || d->baths < 1 || d->baths_max_per_state < 1)
        raise(HEOM_E_INVALID, "matsubaras, baths and baths_max_per_state must be > 0");
    if (d->width != d->baths * d->matsubaras)
        raise(HEOM_E_INVALID, "width must equal baths * matsubaras");
    if (d->matrices < 1 || d->matrices > (int64_t)INT32_MAX - kTile)
        raise(HEOM_E_INVALID, "matrices out of range");
    if (d->precision != HEOM_PRECISION_FP64)
        raise(HEOM_E_UNSUPPORTED, "only HEOM_PRECISION_FP64 is supported by this build");
    if (!d->coupling || !d->hamiltonian || !d->ado_tuples || !d->ado_plus || !d->ado_minus ||
        !d->pf_same_sum || !d->pf_same || !d->pf_circ || !d->cbk)
        raise(HEOM_E_INVALID, "desc has a NULL array");
    for (int t = 0; t < d->states * d->baths_max_per_state; ++t)
        if (d->coupling[t] >= d->baths)
            raise(HEOM_E_INVALID, "coupling entry >= baths");
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
}  // namespace

extern "C" {

const char* heom_last_error(void) { return g_error.c_str(); }

int heom_abi_version(void) { return HEOM_ABI_VERSION; }

int heom_create(const heom_desc* d, heom_handle** out)
{
    if (!out) {
        g_error = "out is NULL";
        return HEOM_E_INVALID;
    }
    *out = nullptr;
    heom_handle* h = nullptr;
    int rc = guard([&] {
        validate(d);
        h = new heom_handle();
        HIPC(hipGetDevice(&h->device));
        HIPC(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        HIPC(hipEventCreate(&h->ev_start));
        HIPC(hipEventCreate(&h->ev_stop));
        const int N = d->states, W = d->width;
        const int64_t M = d->matrices;
        h->N = N;
        h->K = d->matsubaras;
        h->B = d->baths;
        h->BM = d->baths_max_per_state;
        h->W = W;
        h->nel = N * N;
        h->M = M;
        h->Mp = (M + kTile - 1) / kTile * kTile;
        h->ntiles = (int)(h->Mp / kTile);
        h->coupling.assign(d->coupling, d->coupling + N * h->BM);
        h->cbk.assign(d->cbk, d->cbk + 2 * h->B * h->K);
        h->pf_same.assign(d->pf_same, d->pf_same + 2 * h->B);
        h->pf_circ.assign(d->pf_circ, d->pf_circ + 2 * h->B);

        // lexicographic ADO order (mode 0 most significant)
        const int32_t* T = d->ado_tuples;
        int maxv = 0;
        for (int64_t t = 0; t < M * W; ++t) {
            if (T[t] < 0 || T[t] > 255)
                raise(HEOM_E_UNSUPPORTED, "tuple entries must be in [0, 255]");
            maxv = std::max(maxv, (int)T[t]);
        }
        h->nv = maxv + 2;
        std::vector<int32_t> order(M);
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            return std::lexicographical_compare(T + (int64_t)a * W, T + (int64_t)a * W + W,
                                                T + (int64_t)b * W, T + (int64_t)b * W + W);
        });
        const int64_t Mp = h->Mp;
        h->q2u_host.assign(Mp, -1);
        std::vector<int32_t> u2q(M);
        for (int64_t q = 0; q < M; ++q) {
            h->q2u_host[q] = order[q];
            u2q[order[q]] = (int32_t)q;
        }
        h->q_top = -1;
        for (int64_t u = 0; u < M; ++u) {
            bool zero = true;
            for (int m = 0; m < W; ++m)
                zero = zero && T[u * W + m] == 0;
            if (zero) {
                if (u != 0)
                    raise(HEOM_E_INVALID, "the all-zero tuple (top ADO) must be uid 0");
                h->q_top = u2q[u];
            }
        }
        if (h->q_top < 0)
            raise(HEOM_E_INVALID, "no all-zero tuple (top ADO) in ado_tuples");

        std::vector<int32_t> plus(W * Mp, -1), minus(W * Mp, -1);
        std::vector<uint8_t> tup(W * Mp, 0);
        std::vector<double2> pfss(Mp, make_double2(0.0, 0.0));
        for (int64_t q = 0; q < M; ++q) {
            const int64_t u = order[q];
            for (int m = 0; m < W; ++m) {
                const int32_t p = d->ado_plus[u * W + m];
                const int32_t mn = d->ado_minus[u * W + m];
                if (p >= M || mn >= M)
                    raise(HEOM_E_INVALID, "plus/minus entry out of range");
                plus[m * Mp + q] = p >= 0 ? u2q[p] : -1;
                minus[m * Mp + q] = (mn >= 0 && T[u * W + m] > 0) ? u2q[mn] : -1;
                tup[m * Mp + q] = (uint8_t)T[u * W + m];
            }
            pfss[q] = make_double2(d->pf_same_sum[2 * u], d->pf_same_sum[2 * u + 1]);
        }
        upload(h->plus_raw, plus, h->stream);
        upload(h->minus_raw, minus, h->stream);
        upload(h->plus, plus, h->stream);
        upload(h->minus, minus, h->stream);
        upload(h->tup, tup, h->stream);
        upload(h->pfss, pfss, h->stream);
        upload(h->q2u, h->q2u_host, h->stream);
        h->mask.alloc(M);
        HIPC(hipMemsetAsync(h->mask.p, 1, M, h->stream));
        h->norms.alloc(M);
        h->mx.alloc(N * N);
        h->small.alloc(N * N);
        const size_t plane = (size_t)N * N * Mp;
        h->y.alloc(plane);
        h->acc.alloc(plane);
        h->ya.alloc(plane);
        h->yb.alloc(plane);
        HIPC(hipMemsetAsync(h->y.p, 0, h->y.bytes(), h->stream));
        HIPC(hipMemsetAsync(h->acc.p, 0, h->acc.bytes(), h->stream));
        HIPC(hipMemsetAsync(h->ya.p, 0, h->ya.bytes(), h->stream));
        HIPC(hipMemsetAsync(h->yb.p, 0, h->yb.bytes(), h->stream));
        build_coefficients(h);
        set_hamiltonian(h, d->hamiltonian);
        HIPC(hipStreamSynchronize(h->stream));
    });
    if (rc != HEOM_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return HEOM_OK;
}

int heom_destroy(heom_handle* h)
{
    return guard([&] {
        need(h);
        delete h;
    });
}

int heom_set_hamiltonian(heom_handle* h, const double* H)
{
    return guard([&] {
        need(h);
        if (!H)
            raise(HEOM_E_INVALID, "hamiltonian is NULL");
        set_hamiltonian(h, H);
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
        const size_t bytes = (size_t)h->M * h->nel * sizeof(double2);
        HIPC(hipMemcpyAsync(h->ya.p, state, bytes, hipMemcpyHostToDevice, h->stream));
        const int threads = 256;
        hipLaunchKernelGGL(aos_to_soa, dim3((unsigned)((h->Mp + threads - 1) / threads)), dim3(threads), 0,
                           h->stream, h->ya.p, h->y.p, h->q2u.p, h->Mp, h->nel);
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_set_top(heom_handle* h, const double* top)
{
    return guard([&] {
        need(h);
        if (!top)
            raise(HEOM_E_INVALID, "top is NULL");
        HIPC(hipMemcpyAsync(h->small.p, top, h->nel * sizeof(double2), hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(copy_ado_in, dim3(1), dim3(((h->nel + 63) / 64) * 64), 0, h->stream, h->small.p,
                           h->y.p, h->q_top, h->Mp, h->nel);
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_get_state(heom_handle* h, double* state)
{
    return guard([&] {
        need(h);
        if (!state)
            raise(HEOM_E_INVALID, "state is NULL");
        const int threads = 256;
        hipLaunchKernelGGL(soa_to_aos, dim3((unsigned)((h->Mp + threads - 1) / threads)), dim3(threads), 0,
                           h->stream, h->y.p, h->yb.p, h->q2u.p, h->Mp, h->nel);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(state, h->yb.p, (size_t)h->M * h->nel * sizeof(double2),
                            hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_get_top(heom_handle* h, double* top)
{
    return guard([&] {
        need(h);
        if (!top)
            raise(HEOM_E_INVALID, "top is NULL");
        hipLaunchKernelGGL(copy_ado_out, dim3(1), dim3(((h->nel + 63) / 64) * 64), 0, h->stream, h->y.p,
                           h->small.p, h->q_top, h->Mp, h->nel);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(top, h->small.p, h->nel * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_hierarchy_norm(heom_handle* h, double threshold, double* norms)
{
    return guard([&] {
        need(h);
        if (!norms)
            raise(HEOM_E_INVALID, "norms is NULL");
        const int threads = 256;
        hipLaunchKernelGGL(hierarchy_norm_kernel, dim3((unsigned)((h->Mp + threads - 1) / threads)),
                           dim3(threads), 0, h->stream, h->y.p, h->q2u.p, h->Mp, h->nel, threshold,
                           h->norms.p, h->mask.p);
        HIPC(hipGetLastError());
        if (threshold >= 0.0)
            refold_tables(h);
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
        launch_stage(h, 0, h->y.p, h->ya.p, 0.0, 0.0);
        const int threads = 256;
        hipLaunchKernelGGL(soa_to_aos, dim3((unsigned)((h->Mp + threads - 1) / threads)), dim3(threads), 0,
                           h->stream, h->ya.p, h->yb.p, h->q2u.p, h->Mp, h->nel);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(out, h->yb.p, (size_t)h->M * h->nel * sizeof(double2), hipMemcpyDeviceToHost,
                            h->stream));
        HIPC(hipStreamSynchronize(h->stream));
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
        DevBuf<double2> rec;
        rec.alloc((size_t)rows * h->nel);
        const int tb = ((h->nel + 63) / 64) * 64;
        resolve_timing(h);
        HIPC(hipEventRecord(h->ev_start, h->stream));
        hipLaunchKernelGGL(copy_ado_out, dim3(1), dim3(tb), 0, h->stream, h->y.p, rec.p, h->q_top, h->Mp, h->nel);
        int64_t row = 1, done = 0;
        while (done < nsteps) {
            const int64_t n = std::min(observe_steps, nsteps - done);
            rk4_steps(h, dt, n);
            done += n;
            if (n == observe_steps) {
                hipLaunchKernelGGL(copy_ado_out, dim3(1), dim3(tb), 0, h->stream, h->y.p,
                                   rec.p + row * h->nel, h->q_top, h->Mp, h->nel);
                ++row;
            }
        }
        HIPC(hipGetLastError());
        HIPC(hipEventRecord(h->ev_stop, h->stream));
        h->timing_pending = true;
        h->pending_steps = nsteps;
        HIPC(hipMemcpyAsync(tops, rec.p, (size_t)rows * h->nel * sizeof(double2), hipMemcpyDeviceToHost,
                            h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        resolve_timing(h);
        rec.free();
    });
}

int heom_mmult_left(heom_handle* h, const double* mx)
{
    return guard([&] {
        need(h);
        if (!mx)
            raise(HEOM_E_INVALID, "mx is NULL");
        HIPC(hipMemcpyAsync(h->mx.p, mx, h->nel * sizeof(double2), hipMemcpyHostToDevice, h->stream));
        dim3 grid(h->ntiles), block(64 * h->N);
        switch (h->N) {
        case 1: hipLaunchKernelGGL(mmult_left_kernel<1>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp, h->ntiles); break;
        case 2: hipLaunchKernelGGL(mmult_left_kernel<2>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp, h->ntiles); break;
        case 3: hipLaunchKernelGGL(mmult_left_kernel<3>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp, h->ntiles); break;
        case 4: hipLaunchKernelGGL(mmult_left_kernel<4>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp, h->ntiles); break;
        case 5: hipLaunchKernelGGL(mmult_left_kernel<5>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp, h->ntiles); break;
        case 6: hipLaunchKernelGGL(mmult_left_kernel<6>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp, h->ntiles); break;
        case 7: hipLaunchKernelGGL(mmult_left_kernel<7>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp, h->ntiles); break;
        case 8: hipLaunchKernelGGL(mmult_left_kernel<8>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp, h->ntiles); break;
        default: raise(HEOM_E_UNSUPPORTED, "states > 8");
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
        int64_t b = 0;
        for (auto* x : {&h->y, &h->acc, &h->ya, &h->yb, &h->pfss, &h->hs, &h->gam, &h->beta, &h->mx, &h->small})
            b += x->bytes();
        for (auto* x : {&h->plus_raw, &h->minus_raw, &h->plus, &h->minus, &h->q2u, &h->rowm_cnt, &h->rowm,
                        &h->colm_cnt, &h->colm})
            b += x->bytes();
        b += h->tup.bytes() + h->mask.bytes() + h->pm.bytes() + h->pcoef.bytes() + h->scoef.bytes() +
             h->norms.bytes();
        *bytes = b;
    });
}

int heom_rk4_step_bytes(heom_handle* h, int64_t* bytes)
{
    return guard([&] {
        need(h);
        if (!bytes)
            raise(HEOM_E_INVALID, "bytes is NULL");
        *bytes = (int64_t)16 * h->M * h->nel * (int64_t)sizeof(double2);
    });
}

}  // extern "C"
