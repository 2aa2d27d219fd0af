// stage_wide.hip -- the HEOM RHS with the fused RK4 stage epilogue for large systems,
// 8 < N <= 32 states (dm-heom/cl/heom/heom_ode.cl:24-451 plus rk_weighted_add.cl:19-106),
// complex128 / FP64 and complex64 / FP32.
//
// The stage kernels for N <= 8 put one ADO on each lane and a 64-ADO tile of N^2
// elements in LDS; at N = 24 that tile would need 590 KB.  Here a workgroup owns a
// tile of T consecutive ADOs (a power of two up to 8 such that H and T matrices fit
// the 144 KB LDS budget -- the most that fit unless the autotuner chose fewer; T
// divides the 64-ADO instance granularity), and works in the three phases described
// at the kernel: the commutator per ADO wavefront from LDS, everything that touches
// global memory plane by plane across the tile's consecutive ADOs.
#include <type_traits>

#include "heom_device.hpp"

namespace heom_hip {

namespace {

constexpr int kWideLdsBytes = 147456;  // 144 KB of the 160 KB LDS
// a second instance with 48 KB, chosen when H and the tile fit: then up to three
// workgroups share a CU instead of one, as far as their registers allow
constexpr int kWideLdsSmallBytes = 49152;
constexpr int kMaxWideStates = 32;

// LDS row stride (complex elements) of one N x N matrix: one pad element breaks the
// bank pattern of the 8 rows a wavefront reads at once
__host__ __device__ constexpr int wide_row(int N) { return N + 1; }

// at most 8 wavefronts per workgroup, so that a lane may hold up to 256 VGPRs (the
// block's accumulators, operands and in-flight loads)
constexpr int kWideThreads = 512;

// (ADO, element) pairs a thread carries through one pass of phase C
constexpr int kWidePairsPerPass = 2;

// table rows a workgroup copies to LDS for W modes: rows 0..W (mode W is the zero row
// that padding terms name), rounded up to an even count
__host__ __device__ constexpr int wide_table_rows(int W) { return (W + 2) & ~1; }

// LDS bytes of a workgroup of T ADOs: H, the tile (later overwritten by the commutator-
// plus-same-term values) and the tile's slice of the neighbour table
__host__ __device__ constexpr int wide_lds_bytes(int N, int cbytes, int W, int T)
{
    return (T + 1) * N * wide_row(N) * cbytes + wide_table_rows(W) * T * (int)sizeof(int2);
}

// ADOs per workgroup for N states with `cbytes` per complex element and W modes, 0 if
// none fits
int wide_tile(int N, int cbytes, int W)
{
    for (int T = kWideThreads / 64; T >= 1; T /= 2)
        if (wide_lds_bytes(N, cbytes, W, T) <= kWideLdsBytes)
            return T;
    return 0;
}

// Three phases per workgroup (T ADOs):
//   A  the tile sigma and H / hbar into LDS, plane by plane (T consecutive ADOs of one
//      element per T consecutive threads);
//   B  wavefront = ADO: lane l owns the B x B block of elements (i, j) with
//      i = (l / 8) * B + r, j = (l % 8) * B + c (B = ceil(N / 8)); the commutator of the
//      block needs B rows of sigma and H and B columns of both per k, 4B LDS reads for
//      B^2 complex products; after a barrier, commutator plus same term overwrite the
//      tile in LDS;
//   C  back to the plane-by-plane mapping: each thread takes (ADO, element) pairs, adds
//      the element's neighbour terms (table entry and gathers) to the LDS value and runs
//      the RK4 epilogue with the own element re-read from global memory (still in L2),
//      so that loads of the stage operand and all stores move T consecutive ADOs of
//      one plane together.  LDS: H and one tile, so that several workgroups share a CU.
template <typename R, int B, int NT, int STAGE, bool REALH, int LDSB, int PP = kWidePairsPerPass>
__global__ __launch_bounds__(kWideThreads) void rk4_stage_wide(StageArgsT<R> a, const c2<R>* __restrict__ hs_all,
                                                               const c2<R>* __restrict__ gam,
                                                               const TermT<R>* __restrict__ terms, int N, int T)
{
    using C = c2<R>;
    __shared__ C lds[LDSB / sizeof(C)];
    __shared__ R sqrt_n[64];  // sqrt(n), n = 0..63
    const int nel = N * N, NP = wide_row(N), mat = N * NP;
    C* hl = lds;         // [N][NP] H / hbar
    C* sig = lds + mat;  // [T][N][NP] the tile, then commutator + same term
    int2* tabl = reinterpret_cast<int2*>(sig + T * mat);  // [rows][T] the tile's table slice
    const int trows = 2 * a.table_units;
    // index arithmetic without runtime divisions: T is a power of two, and
    // e / N = (e * ninv) >> 20 for every e < 8256 and N <= 32
    const int tsh = __builtin_ctz((unsigned)T);
    const uint32_t ninv = (1u << 20) / (uint32_t)N + 1u;
    auto div_n = [&](int e) { return (int)(((uint32_t)e * ninv) >> 20); };

    const int64_t mp = a.mp;
    // XCD-aware order: runs of 16 consecutive workgroups (16 T ADOs) on one XCD
    const int64_t q0 = (int64_t)xcd_tile(blockIdx.x, gridDim.x, 16) * T;
    const C* hs = hs_all + (a.tile_inst ? (int64_t)a.tile_inst[q0 / kTile] * nel : 0);
    const int cls = tile_bath_class(a, (int)(q0 / kTile));
    gam += cls * nel;
    terms += cls * nel * NT;
    if (threadIdx.x < 64)
        sqrt_n[threadIdx.x] = sqrt((R)threadIdx.x);
    for (int idx = threadIdx.x; idx < nel; idx += blockDim.x) {
        const int r = div_n(idx);
        hl[r * NP + idx - r * N] = hs[idx];
    }
    for (int idx = threadIdx.x; idx < T * nel; idx += blockDim.x) {
        const int ad = idx & (T - 1), e = idx >> tsh, r = div_n(e);
        sig[ad * mat + r * NP + e - r * N] = a.yin[(int64_t)e * mp + q0 + ad];
    }
    for (int idx = threadIdx.x; idx < trows * T; idx += blockDim.x)
        tabl[idx] = a.nb[(int64_t)(idx >> tsh) * mp + q0 + (idx & (T - 1))];
    __syncthreads();

    // phase B
    {
        const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int lane = threadIdx.x & 63;
        const int i0 = (lane >> 3) * B, j0 = (lane & 7) * B;
        const C S = a.pfss[q0 + wv];
        const C* s = sig + wv * mat;
        int ir[B], jc[B];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            ir[r] = i0 + r < N ? i0 + r : 0;  // rows beyond N read row 0 and are not kept
            jc[r] = j0 + r < N ? j0 + r : 0;
        }
        C cm[B][B];
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c)
                cm[r][c] = mk2<R>(R(0), R(0));
#pragma unroll 4
        for (int k = 0; k < N; ++k) {
            C sik[B], hik[B], hkj[B], skj[B];
#pragma unroll
            for (int r = 0; r < B; ++r) {
                sik[r] = s[ir[r] * NP + k];
                hik[r] = hl[ir[r] * NP + k];
                hkj[r] = hl[k * NP + jc[r]];
                skj[r] = s[k * NP + jc[r]];
            }
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int c = 0; c < B; ++c) {
                    if (REALH) {
                        cm[r][c].x = fma(sik[r].x, hkj[c].x, fma(-hik[r].x, skj[c].x, cm[r][c].x));
                        cm[r][c].y = fma(sik[r].y, hkj[c].x, fma(-hik[r].x, skj[c].y, cm[r][c].y));
                    } else {
                        cm[r][c] = cfma(sik[r], hkj[c], cm[r][c]);
                        cm[r][c] = cfma(mk2<R>(-hik[r].x, -hik[r].y), skj[c], cm[r][c]);
                    }
                }
        }
        C fv[B][B];
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const int i = ir[r], j = jc[c];
                const C g = gam[i * N + j];
                const C own = s[i * NP + j];
                const C fb = cfma(mk2<R>(S.x + g.x, S.y + g.y), own, mk2<R>(-cm[r][c].y, cm[r][c].x));
                // stages 1 and 4 take the element's own value linearly (s + c1 f, and
                // (ep + s) / 3 + c1 f): its part goes into the LDS value here, so that
                // phase C needs no second read of it
                if (STAGE == 1)
                    fv[r][c] = mk2<R>(fma(a.c1, fb.x, own.x), fma(a.c1, fb.y, own.y));
                else if (STAGE == 4)
                    fv[r][c] = mk2<R>(fma(own.x, R(1) / R(3), a.c1 * fb.x), fma(own.y, R(1) / R(3), a.c1 * fb.y));
                else
                    fv[r][c] = fb;
            }
        __syncthreads();  // every wavefront is done reading the tile
        C* fo = sig + wv * mat;
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c)
                if (i0 + r < N && j0 + c < N)
                    fo[(i0 + r) * NP + j0 + c] = fv[r][c];
    }
    __syncthreads();

    // phase C, PP (ADO, element) pairs per thread and pass so that their loads overlap
    // (table entries from LDS, then the gathers, stage operands and own values together);
    // PLUS = false (no ADO of the 64-ADO tile that holds these T has a plus neighbour)
    // leaves out the plus gathers
    auto phase_c = [&](auto plus_tag) {
        constexpr bool PLUS = decltype(plus_tag)::value;
        const int total = T * nel, stride = blockDim.x;
        for (int base = threadIdx.x; base < total; base += PP * stride) {
            int2 nbv[PP][NT];
            C ep[PP], own[PP], xp[PP][NT], xm[PP][NT];
            int64_t o[PP];
            int e2[PP], lo[PP];
            bool live[PP];
#pragma unroll
            for (int u = 0; u < PP; ++u) {
                const int idx = base + u * stride;
                live[u] = idx < total;
                const int id = live[u] ? idx : base;  // a dead pair repeats the first pair's loads
                const int ad = id & (T - 1), e = id >> tsh, r = div_n(e);
                e2[u] = e;
                lo[u] = ad * mat + r * NP + e - r * N;
                o[u] = (int64_t)e * mp + q0 + ad;
#pragma unroll
                for (int k = 0; k < NT; ++k)
                    nbv[u][k] = tabl[terms[e * NT + k].mode * T + ad];
                ep[u] = stage_operand<STAGE, R>(a, o[u]);
                if (STAGE == 2)  // stages 1 and 4 have it in the LDS value, 0 and 3 do not use it
                    own[u] = a.yin[o[u]];
            }
#pragma unroll
            for (int u = 0; u < PP; ++u) {
                const C* plane = a.yin + (int64_t)e2[u] * mp;
#pragma unroll
                for (int k = 0; k < NT; ++k) {
                    if constexpr (PLUS)
                        xp[u][k] = plane[(uint32_t)nbv[u][k].x >> 6];
                    xm[u][k] = plane[(uint32_t)nbv[u][k].y];
                }
            }
#pragma unroll
            for (int u = 0; u < PP; ++u) {
                if (!live[u])
                    continue;
                // stages 1 and 4: the LDS value already holds the own part and c1 times the
                // commutator and same terms; f gathers the neighbour terms alone
                const C lv = sig[lo[u]];
                C f = (STAGE == 1 || STAGE == 4) ? mk2<R>(R(0), R(0)) : lv;
#pragma unroll
                for (int k = 0; k < NT; ++k) {
                    const TermT<R> t = terms[e2[u] * NT + k];
                    const int v = nbv[u][k].x & 63;
                    if constexpr (PLUS) {
                        const R cp = t.plus_c * sqrt_n[v + 1];
                        f.x = fma(-cp, xp[u][k].y, f.x);
                        f.y = fma(cp, xp[u][k].x, f.y);
                    }
                    const R sm = sqrt_n[v];
                    f = cfma(t.minus_c, mk2<R>(sm * xm[u][k].x, sm * xm[u][k].y), f);
                }
                C r0, r1 = mk2<R>(R(0), R(0));
                if (STAGE == 1) {
                    r0 = mk2<R>(fma(a.c1, f.x, lv.x), fma(a.c1, f.y, lv.y));
                } else if (STAGE == 4) {
                    constexpr R third = R(1) / R(3);
                    r0 = mk2<R>(fma(ep[u].x, third, fma(a.c1, f.x, lv.x)), fma(ep[u].y, third, fma(a.c1, f.y, lv.y)));
                } else {
                    stage_values<STAGE, R>(a.c1, STAGE == 2 ? own[u] : mk2<R>(R(0), R(0)), f, ep[u], r0, r1);
                }
                stage_store<STAGE, R>(a, o[u], r0, r1);
            }
        }
    };
    if (tile_has_plus(a, (int)(q0 / kTile)))
        phase_c(std::true_type{});
    else
        phase_c(std::false_type{});
}

template <typename R, int B, int NT, bool REALH, int LDSB>
void launch_wl(int N, int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el, int T, hipStream_t st)
{
    const dim3 grid((unsigned)(a.mp / T)), block(64 * T);
    switch (stage) {
    case 0: hipLaunchKernelGGL((rk4_stage_wide<R, B, NT, 0, REALH, LDSB>), grid, block, 0, st, a, el.hs, el.gam, el.terms, N, T); break;
    case 1: hipLaunchKernelGGL((rk4_stage_wide<R, B, NT, 1, REALH, LDSB>), grid, block, 0, st, a, el.hs, el.gam, el.terms, N, T); break;
    case 2: hipLaunchKernelGGL((rk4_stage_wide<R, B, NT, 2, REALH, LDSB>), grid, block, 0, st, a, el.hs, el.gam, el.terms, N, T); break;
    case 3: hipLaunchKernelGGL((rk4_stage_wide<R, B, NT, 3, REALH, LDSB>), grid, block, 0, st, a, el.hs, el.gam, el.terms, N, T); break;
    default: hipLaunchKernelGGL((rk4_stage_wide<R, B, NT, 4, REALH, LDSB>), grid, block, 0, st, a, el.hs, el.gam, el.terms, N, T); break;
    }
}

// the LDS instance: H and T matrices of N x wide_row(N) complex elements
template <typename R, int B, int NT, bool REALH>
void launch_wb(int N, int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el, int T, hipStream_t st)
{
    if (wide_lds_bytes(N, (int)sizeof(c2<R>), 2 * a.table_units - 2, T) <= kWideLdsSmallBytes)
        launch_wl<R, B, NT, REALH, kWideLdsSmallBytes>(N, stage, a, el, T, st);
    else
        launch_wl<R, B, NT, REALH, kWideLdsBytes>(N, stage, a, el, T, st);
}

template <typename R, int B, bool REALH>
void launch_wn(int N, int nt, int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el, int T, hipStream_t st)
{
    switch (nt) {
    case 2: launch_wb<R, B, 2, REALH>(N, stage, a, el, T, st); break;
    case 4: launch_wb<R, B, 4, REALH>(N, stage, a, el, T, st); break;
    default: launch_wb<R, B, 8, REALH>(N, stage, a, el, T, st); break;
    }
}

template <typename R, bool REALH>
void launch_w(int N, int nt, int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el, int ados, hipStream_t st)
{
    const int fit = wide_tile(N, (int)sizeof(c2<R>), 2 * a.table_units - 2);
    const int T = ados > 0 && ados <= fit && kTile % ados == 0 ? ados : fit;
    const int B = (N + 7) / 8;
    if (B <= 2)
        launch_wn<R, 2, REALH>(N, nt, stage, a, el, T, st);
    else if (B == 3)
        launch_wn<R, 3, REALH>(N, nt, stage, a, el, T, st);
    else
        launch_wn<R, 4, REALH>(N, nt, stage, a, el, T, st);
}

}  // namespace

bool wide_supported(int N, int complex_bytes, int W)
{
    return N > kMaxStates && N <= kMaxWideStates && wide_tile(N, complex_bytes, W) > 0;
}

int wide_max_ados(int N, int complex_bytes, int W) { return wide_tile(N, complex_bytes, W); }

void launch_stage_wide(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<double>& a,
                       const ElementArgsT<double>& el, int ados, hipStream_t s)
{
    if (real_h)
        launch_w<double, true>(N, nterm_slots, stage, a, el, ados, s);
    else
        launch_w<double, false>(N, nterm_slots, stage, a, el, ados, s);
}

void launch_stage_wide(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<float>& a,
                       const ElementArgsT<float>& el, int ados, hipStream_t s)
{
    if (real_h)
        launch_w<float, true>(N, nterm_slots, stage, a, el, ados, s);
    else
        launch_w<float, false>(N, nterm_slots, stage, a, el, ados, s);
}

}  // namespace heom_hip
