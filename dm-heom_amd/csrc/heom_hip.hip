// heom_hip.hip -- the C ABI of include/heom_hip.h: handle lifetime, the re-layout of
// the reference instance into the device layout (heom_device.hpp), the RK4 driver,
// and the small layout / reduction / hierarchy-arithmetic kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <new>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "heom_device.hpp"
#include "heom_hip.h"

using namespace heom_hip;

namespace {

constexpr double kHbar = 1.05457148e-34;  // dm-heom/cl/heom/common.cl:45

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
// layout, reduction and hierarchy-arithmetic kernels
// ------------------------------------------------------------------------------------

// dst[e * mp + q] = src[u(q)][e]   (reference AoS -> internal SoA; padding -> 0)
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
__global__ void copy_ado_out(const double2* __restrict__ src, double2* __restrict__ dst, int64_t q0,
                             int64_t mp, int nel)
{
    const int e = threadIdx.x;
    if (e < nel)
        dst[e] = src[(int64_t)e * mp + q0];
}

__global__ void copy_ado_in(const double2* __restrict__ src, double2* __restrict__ dst, int64_t q0,
                            int64_t mp, int nel)
{
    const int e = threadIdx.x;
    if (e < nel)
        dst[(int64_t)e * mp + q0] = src[e];
}

// per-ADO squared Frobenius norm in uid order; threshold masking as in
// dm-heom/cl/heom/hierarchy_norm.cl:41-44 (mask in uid order)
__global__ void hierarchy_norm_kernel(const double2* __restrict__ y, const int32_t* __restrict__ q2u,
                                      int64_t mp, int nel, double threshold, double* __restrict__ norms,
                                      uint8_t* __restrict__ mask)
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

// sigma_q <- Mx sigma_q for every ADO: wave = row i, lane = ADO, tile in LDS
template <int N>
__global__ __launch_bounds__(64 * N) void mmult_left_kernel(double2* __restrict__ y, const double2* __restrict__ mx,
                                                            int64_t mp)
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
// device buffers
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
    int N = 0, K = 0, B = 0, BM = 0, W = 0, nel = 0;
    int64_t M = 0, Mp = 0;
    int ntiles = 0;
    int32_t zero_slot = 0;
    bool real_h = false;
    int term_slots = 2;  // neighbour terms per element the stage kernel runs with
    bool pipelined = false;  // stage_pipelined.hip (persistent) instead of stage_kernel.hip
    int pipe_wgs = 0;        // its workgroups: one per CU, a multiple of the XCD count
    int xcd_chunk = 1;       // tiles per XCD chunk of the stage kernels' tile order
    int64_t q_top = 0;   // internal index of uid 0

    std::vector<int32_t> coupling;  // [N * BM]
    std::vector<double> cbk, pf_same, pf_circ;

    DevBuf<double2> y, pbuf, ya, yb;
    DevBuf<int2> nb_raw, nb;  // packed neighbour table, unmasked and mask-folded
    DevBuf<int32_t> q2u;
    DevBuf<uint8_t> mask;
    DevBuf<double2> pfss, hs, gam, mx, small;
    DevBuf<Term> terms;
    DevBuf<double> norms;

    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    bool timing_pending = false;
    int64_t pending_steps = 0;
    heom_timing_info timing{};

    ~heom_handle()
    {
        if (stream)
            (void)hipStreamSynchronize(stream);
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

void launch_stage(heom_handle* h, int stage, const double2* yin, double2* yout, double c1, bool reverse = false)
{
    StageArgs a{};
    a.yin = yin;
    a.y0 = h->y.p;
    a.pbuf = h->pbuf.p;
    a.yout = yout;
    a.nb = h->nb.p;
    a.pfss = h->pfss.p;
    a.mp = h->Mp;
    a.ntiles = h->ntiles;
    a.zero_slot = h->zero_slot;
    a.table_units = (h->W + 2) / 2;
    a.xcd_chunk = h->xcd_chunk;
    a.reverse = reverse ? 1 : 0;
    a.c1 = c1;
    ElementArgs el{h->hs.p, h->gam.p, h->terms.p};
    if (h->pipelined)
        heom_hip::launch_stage_pipelined(h->N, h->term_slots, h->real_h, stage, a, el, h->pipe_wgs, h->stream);
    else
        heom_hip::launch_stage(h->N, h->term_slots, h->real_h, stage, a, el, h->stream);
    HIPC(hipGetLastError());
    h->timing.rhs_launches += 1;
}

// classical RK4 in the 12-pass schedule of heom_device.hpp
void rk4_steps(heom_handle* h, double dt, int64_t nsteps)
{
    for (int64_t s = 0; s < nsteps; ++s) {
        // stages alternate their tile walk, so that each starts where the previous one
        // ended, on the part of its input that is still in the last-level cache
        launch_stage(h, 1, h->y.p, h->ya.p, 0.5 * dt, false);
        launch_stage(h, 2, h->ya.p, h->yb.p, 0.5 * dt, true);
        launch_stage(h, 3, h->yb.p, h->ya.p, dt, false);
        launch_stage(h, 4, h->ya.p, h->y.p, dt / 6.0, true);
    }
}

void refold_tables(heom_handle* h)
{
    const int64_t count = (int64_t)(h->W + 2) * h->Mp;
    const int threads = 256;
    const unsigned blocks = (unsigned)((count + threads - 1) / threads);
    hipLaunchKernelGGL(fold_mask, dim3(blocks), dim3(threads), 0, h->stream, h->nb_raw.p, h->nb.p, h->mask.p,
                       h->q2u.p, count, h->zero_slot);
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
    HIPC(hipMemcpyAsync(h->hs.p, hs.data(), hs.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->real_h = real;
}

int count_bath(const heom_handle* h, int s, int b)
{
    int c = 0;
    for (int t = 0; t < h->BM; ++t)
        if (h->coupling[s * h->BM + t] == b)
            ++c;
    return c;
}

// Neighbour terms per matrix element and the "same"-term constants.  For element
// (i, j) and mode m of bath b = m / K, heom_ode.cl adds, for every coupling-list entry
// of state i equal to b, +i sqrt((n_m+1)|c_m|/hbar) sigma^{n+m} and
// (+i c_m/hbar + [k=0] Phi_b) sqrt(n_m hbar/|c_m|) sigma^{n-m}; for every entry of
// state j equal to b the same with -i (heom_ode.cl:184-266, 280-405); the same term
// subtracts Delta_b for the baths of i and of j when i != j (heom_ode.cl:138-168).
// With c_i, c_j the number of entries equal to b in rows i and j, one Term per mode
// carries plus_c = (c_i - c_j) sqrt(|c_m|/hbar) and
// minus_c = (i c_m/hbar (c_i - c_j) + [k=0] Phi_b (c_i + c_j)) sqrt(hbar/|c_m|).
void build_coefficients(heom_handle* h)
{
    const int N = h->N, W = h->W, K = h->K, BM = h->BM;
    std::vector<std::vector<Term>> per(N * N);
    std::vector<double2> gam(N * N, make_double2(0.0, 0.0));
    int most = 0;
    for (int i = 0; i < N; ++i) {
        for (int j = 0; j < N; ++j) {
            const int e = i * N + j;
            for (int m = 0; m < W; ++m) {
                const int b = m / K, k = m % K;
                const int ci = count_bath(h, i, b), cj = count_bath(h, j, b);
                const double cr = h->cbk[2 * m], cim = h->cbk[2 * m + 1];
                const double a = std::hypot(cr, cim);
                if (a == 0.0)
                    continue;  // a mode with c_m = 0 contributes nothing (c sqrt(1/|c|) -> 0)
                const double cp = std::sqrt(a / kHbar), cs = std::sqrt(kHbar / a);
                Term t{};
                t.mode = m;
                t.plus_c = (ci - cj) * cp;
                double br = -cim / kHbar * (ci - cj), bi = cr / kHbar * (ci - cj);
                if (k == 0) {
                    br += h->pf_circ[2 * b] * (ci + cj);
                    bi += h->pf_circ[2 * b + 1] * (ci + cj);
                }
                t.minus_c = make_double2(br * cs, bi * cs);
                if (t.plus_c != 0.0 || t.minus_c.x != 0.0 || t.minus_c.y != 0.0)
                    per[e].push_back(t);
            }
            most = std::max(most, (int)per[e].size());
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
    if (!stage_supported(N, most))
        raise(HEOM_E_UNSUPPORTED, "more than 8 neighbour terms per matrix element");
    h->term_slots = stage_term_slots(most);
    // pad every element's list; padding terms read table column W, which always
    // holds the zero slot
    std::vector<Term> flat((size_t)N * N * h->term_slots);
    for (int e = 0; e < N * N; ++e)
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
    if (d->matrices < 1)
        raise(HEOM_E_INVALID, "matrices must be > 0");
    if (d->matrices > kMaxAdos - 2 * kTile)
        raise(HEOM_E_UNSUPPORTED, "more than 2^26 - 129 ADOs");
    if (d->precision != HEOM_PRECISION_FP64)
        raise(HEOM_E_UNSUPPORTED, "only HEOM_PRECISION_FP64 is supported by this build");
    if (!d->coupling || !d->hamiltonian || !d->ado_tuples || !d->ado_plus || !d->ado_minus || !d->pf_same_sum ||
        !d->pf_same || !d->pf_circ || !d->cbk)
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

// The persistent pipelined stage kernel runs one workgroup per CU; it needs N >= 6 and
// the neighbour-table slice of a tile in LDS.  HEOM_STAGE_KERNEL=tile in the
// environment selects the one-workgroup-per-tile kernel instead.  The tile order
// (chunk_tile in heom_device.hpp) defaults to one contiguous range per XCD for the
// tile kernel and to XCD chunks of one tile per workgroup for the pipelined kernel;
// HEOM_XCD_CHUNK=<tiles> overrides it.
void choose_stage_kernel(heom_handle* h)
{
    const char* env = std::getenv("HEOM_STAGE_KERNEL");
    const bool want_tile = env && std::string(env) == "tile";
    h->pipelined = !want_tile && pipe_supported(h->N, h->W, h->Mp);
    if (h->pipelined) {
        int cus = 0;
        HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
        const int wgs = std::max(1, std::min(cus, h->ntiles));
        h->pipe_wgs = (wgs + kNumXcd - 1) / kNumXcd * kNumXcd;
        h->xcd_chunk = h->pipe_wgs / kNumXcd;
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

unsigned blocks_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

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
    const int rc = guard([&] {
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
        h->Mp = (M + 1 + kTile - 1) / kTile * kTile;  // at least one zero slot
        h->ntiles = (int)(h->Mp / kTile);
        h->zero_slot = (int32_t)M;
        h->coupling.assign(d->coupling, d->coupling + N * h->BM);
        h->cbk.assign(d->cbk, d->cbk + 2 * h->B * h->K);
        h->pf_same.assign(d->pf_same, d->pf_same + 2 * h->B);
        h->pf_circ.assign(d->pf_circ, d->pf_circ + 2 * h->B);

        // lexicographic ADO order (mode 0 most significant)
        const int32_t* T = d->ado_tuples;
        for (int64_t t = 0; t < M * W; ++t)
            if (T[t] < 0 || T[t] > kMaxTupleValue)
                raise(HEOM_E_UNSUPPORTED, "tuple entries must be in [0, 62]");
        std::vector<int32_t> order(M);
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            return std::lexicographical_compare(T + (int64_t)a * W, T + (int64_t)a * W + W, T + (int64_t)b * W,
                                                T + (int64_t)b * W + W);
        });
        const int64_t Mp = h->Mp;
        std::vector<int32_t> q2u(Mp, -1), u2q(M);
        for (int64_t q = 0; q < M; ++q) {
            q2u[q] = order[q];
            u2q[order[q]] = (int32_t)q;
        }
        for (int m = 0; m < W; ++m)
            if (T[m] != 0)
                raise(HEOM_E_INVALID, "uid 0 must be the top ADO (all-zero tuple)");
        h->q_top = u2q[0];

        const int32_t zero = h->zero_slot;
        // rows 0..W-1 are the modes; rows W and W+1 hold the zero slot everywhere (row W
        // serves padding terms, and the pipelined kernel loads rows in pairs)
        std::vector<int2> nb((W + 2) * Mp, make_int2((int32_t)((uint32_t)zero << 6), zero));
        std::vector<double2> pfss(Mp, make_double2(0.0, 0.0));
        for (int64_t q = 0; q < M; ++q) {
            const int64_t u = order[q];
            for (int m = 0; m < W; ++m) {
                const int32_t p = d->ado_plus[u * W + m];
                const int32_t mn = d->ado_minus[u * W + m];
                if (p >= M || mn >= M)
                    raise(HEOM_E_INVALID, "plus/minus entry out of range");
                const int32_t pq = p >= 0 ? u2q[p] : zero;
                const int32_t mq = (mn >= 0 && T[u * W + m] > 0) ? u2q[mn] : zero;
                nb[m * Mp + q] = make_int2((int32_t)(((uint32_t)pq << 6) | (uint32_t)T[u * W + m]), mq);
            }
            pfss[q] = make_double2(d->pf_same_sum[2 * u], d->pf_same_sum[2 * u + 1]);
        }
        upload(h->nb_raw, nb, h->stream);
        upload(h->nb, nb, h->stream);
        upload(h->pfss, pfss, h->stream);
        upload(h->q2u, q2u, h->stream);
        h->mask.alloc(M);
        HIPC(hipMemsetAsync(h->mask.p, 1, M, h->stream));
        h->norms.alloc(M);
        h->mx.alloc(N * N);
        h->small.alloc(N * N);
        const size_t plane = (size_t)N * N * Mp;
        for (auto* b : {&h->y, &h->pbuf, &h->ya, &h->yb}) {
            b->alloc(plane);
            HIPC(hipMemsetAsync(b->p, 0, b->bytes(), h->stream));
        }
        build_coefficients(h);
        set_hamiltonian(h, d->hamiltonian);
        choose_stage_kernel(h);
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
        hipLaunchKernelGGL(aos_to_soa, dim3(blocks_for(h->Mp, 256)), dim3(256), 0, h->stream, h->ya.p, h->y.p,
                           h->q2u.p, h->Mp, h->nel);
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
        hipLaunchKernelGGL(copy_ado_in, dim3(1), dim3(((h->nel + 63) / 64) * 64), 0, h->stream, h->small.p, h->y.p,
                           h->q_top, h->Mp, h->nel);
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_clear_state(heom_handle* h)
{
    return guard([&] {
        need(h);
        HIPC(hipMemsetAsync(h->y.p, 0, h->y.bytes(), h->stream));
        HIPC(hipStreamSynchronize(h->stream));
    });
}

int heom_get_state(heom_handle* h, double* state)
{
    return guard([&] {
        need(h);
        if (!state)
            raise(HEOM_E_INVALID, "state is NULL");
        hipLaunchKernelGGL(soa_to_aos, dim3(blocks_for(h->Mp, 256)), dim3(256), 0, h->stream, h->y.p, h->yb.p,
                           h->q2u.p, h->Mp, h->nel);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(state, h->yb.p, (size_t)h->M * h->nel * sizeof(double2), hipMemcpyDeviceToHost,
                            h->stream));
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
        hipLaunchKernelGGL(hierarchy_norm_kernel, dim3(blocks_for(h->Mp, 256)), dim3(256), 0, h->stream, h->y.p,
                           h->q2u.p, h->Mp, h->nel, threshold, h->norms.p, h->mask.p);
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
        launch_stage(h, 0, h->y.p, h->ya.p, 0.0);
        hipLaunchKernelGGL(soa_to_aos, dim3(blocks_for(h->Mp, 256)), dim3(256), 0, h->stream, h->ya.p, h->yb.p,
                           h->q2u.p, h->Mp, h->nel);
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
                hipLaunchKernelGGL(copy_ado_out, dim3(1), dim3(tb), 0, h->stream, h->y.p, rec.p + row * h->nel,
                                   h->q_top, h->Mp, h->nel);
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
    });
}

int heom_mmult_left(heom_handle* h, const double* mx)
{
    return guard([&] {
        need(h);
        if (!mx)
            raise(HEOM_E_INVALID, "mx is NULL");
        HIPC(hipMemcpyAsync(h->mx.p, mx, h->nel * sizeof(double2), hipMemcpyHostToDevice, h->stream));
        const dim3 grid(h->ntiles), block(64 * h->N);
        switch (h->N) {
        case 1: hipLaunchKernelGGL(mmult_left_kernel<1>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp); break;
        case 2: hipLaunchKernelGGL(mmult_left_kernel<2>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp); break;
        case 3: hipLaunchKernelGGL(mmult_left_kernel<3>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp); break;
        case 4: hipLaunchKernelGGL(mmult_left_kernel<4>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp); break;
        case 5: hipLaunchKernelGGL(mmult_left_kernel<5>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp); break;
        case 6: hipLaunchKernelGGL(mmult_left_kernel<6>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp); break;
        case 7: hipLaunchKernelGGL(mmult_left_kernel<7>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp); break;
        default: hipLaunchKernelGGL(mmult_left_kernel<8>, grid, block, 0, h->stream, h->y.p, h->mx.p, h->Mp); break;
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
        for (auto* x : {&h->y, &h->pbuf, &h->ya, &h->yb, &h->pfss, &h->hs, &h->gam, &h->mx, &h->small})
            b += x->bytes();
        b += h->nb_raw.bytes() + h->nb.bytes() + h->q2u.bytes() + h->mask.bytes() + h->terms.bytes() +
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

int heom_stage_kernel(heom_handle* h, int* kind)
{
    return guard([&] {
        need(h);
        if (!kind)
            raise(HEOM_E_INVALID, "kind is NULL");
        *kind = h->pipelined ? HEOM_STAGE_PIPELINED : HEOM_STAGE_TILE;
    });
}

}  // extern "C"
