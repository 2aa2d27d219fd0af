// stage_pair.hip -- the persistent, software-pipelined stage kernel for complex64 state
// in the paired layout (heom_device.hpp: the elements (i, j) and (j, i) share a plane of
// 16-byte slots, and so do pairs of diagonal elements).
//
// Same RHS and RK4 stage arithmetic as stage_kernel.hip (dm-heom/cl/heom/heom_ode.cl
// :24-451 with rk_weighted_add.cl:19-106 fused in), in FP32, organised like
// stage_pipelined.hip: one workgroup per CU walks a run of 64-ADO tiles, two LDS buffers
// hold a tile (all planes) and its slice of the neighbour table, and the next tile
// streams into one buffer with direct-to-LDS loads while the other is computed.
//
// A wavefront takes two planes (four matrix elements).  Both elements of an
// off-diagonal plane couple to the baths of i and of j, so each neighbour term is one
// 16-byte gather for two elements, half the gather instructions of a plane per element;
// the own loads and stores of a plane are 16 bytes per lane as well.
//
// Commutator.  The planes holding sigma_{r k} for every k hold sigma_{k r} too, so row r
// with a partner index p gives, from N 16-byte LDS reads,
//   U(r; p) = sum_k sigma_{r k} H_{k p},   V(r; p) = -sum_k H_{p k} sigma_{k r}.
// An off-diagonal plane (i, j) has [H, sigma]-terms U(i; j) + V(j; i) for (i, j) and
// U(j; i) + V(i; j) for (j, i); a diagonal plane (a, a+1) has U(a; a) + V(a; a) and
// U(a+1; a+1) + V(a+1; a+1).  The row is a template argument (a uniform switch picks
// it), so every LDS read is one base register with a constant offset and the half
// holding sigma_{r k} is known at compile time.
#include "heom_device.hpp"

namespace heom_hip {

namespace {

#define HEOM_LDS_PTR(p) ((__attribute__((address_space(3))) void*)(p))

// cache policy of the table loads: non-temporal (each table row is read once a stage)
constexpr int kPairTableCachePolicy = 2;
constexpr int kPairsPerWave = 2;

using f4 = float __attribute__((ext_vector_type(4)));

template <int N>
struct PairShape {
    static constexpr int D = pair_planes(N);                              // planes
    static constexpr int NWAVE = (D + kPairsPerWave - 1) / kPairsPerWave;  // wavefronts
    static constexpr int DS = NWAVE * kPairsPerWave;                       // LDS plane slots
    // table rows one LDS buffer holds: two state tiles (DS x 64 x 16 B) and two table
    // tiles (rows x 64 x 8 B) in 160 KiB with 4 KiB to spare, an even number, at most 64
    static constexpr int TC = ((163840 - 4096) / 2 - DS * 1024) / 512 < 64
                                  ? (((163840 - 4096) / 2 - DS * 1024) / 512) & ~1
                                  : 64;
};

// the two elements (r0, c0), (r1, c1) of plane d; valid1 = false for the lone diagonal of
// an odd N (its half 1 stays zero)
template <int N>
__device__ __forceinline__ void plane_elements(int d, int& r0, int& c0, int& r1, int& c1, bool& valid1)
{
    if (d < pair_count(N)) {
        // row i: the planes of rows 0 .. i-1 hold N-1, N-2, ... pairs
        int i = 0, base = 0;
#pragma unroll
        for (int a = 0; a < N - 1; ++a) {
            const int len = N - 1 - a;
            if (i == a && d >= base + len) {
                base += len;
                i = a + 1;
            }
        }
        const int j = i + 1 + (d - base);
        r0 = i, c0 = j, r1 = j, c1 = i;
        valid1 = true;
    } else {
        const int a = 2 * (d - pair_count(N));
        r0 = c0 = a;
        valid1 = a + 1 < N;
        r1 = c1 = valid1 ? a + 1 : a;
    }
}

// U(R; p) and V(R; p) of the comment at the top of this file; vb is this lane's address
// in the tile's LDS buffer, hcol = H + p (column p, stride N), hrow = H + p N (row p)
template <int N, int R, bool REALH>
__device__ __forceinline__ void row_terms(const char* vb, const float2* hcol, const float2* hrow, float2& u, float2& v)
{
    // the N reads in two batches of about N/2, each summed before the next is issued:
    // half the registers of one batch (eight-term planes then fit at N = 8)
    u = v = make_float2(0.f, 0.f);
    constexpr int HALF = (N + 1) / 2;
#pragma unroll
    for (int k0 = 0; k0 < N; k0 += HALF) {
        f4 A[N];
#pragma unroll
        for (int k = k0; k < k0 + HALF && k < N; ++k)
            A[k] = *reinterpret_cast<const f4*>(vb + pair_slot(N, R, k).x * (int)(sizeof(f4) * kTile));
#pragma unroll
        for (int k = k0; k < k0 + HALF && k < N; ++k) {
            const int h = pair_slot(N, R, k).y;  // half holding sigma_{R k}; sigma_{k R} is the other
            const int ht = k == R ? h : 1 - h;   // (the same one on the diagonal)
            const float2 s_rk = h ? make_float2(A[k].z, A[k].w) : make_float2(A[k].x, A[k].y);
            const float2 s_kr = ht ? make_float2(A[k].z, A[k].w) : make_float2(A[k].x, A[k].y);
            const float2 hkp = hcol[k * N], hpk = hrow[k];
            if (REALH) {
                u.x = fmaf(s_rk.x, hkp.x, u.x);
                u.y = fmaf(s_rk.y, hkp.x, u.y);
                v.x = fmaf(-hpk.x, s_kr.x, v.x);
                v.y = fmaf(-hpk.x, s_kr.y, v.y);
            } else {
                u = cfma(s_rk, hkp, u);
                v = cfma(make_float2(-hpk.x, -hpk.y), s_kr, v);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// row_terms for a wave-uniform row r
template <int N, bool REALH, int R = 0>
__device__ __forceinline__ void row_terms_of(int r, const char* vb, const float2* hcol, const float2* hrow, float2& u,
                                             float2& v)
{
    if constexpr (R < N - 1) {
        if (r == R)
            row_terms<N, R, REALH>(vb, hcol, hrow, u, v);
        else
            row_terms_of<N, REALH, R + 1>(r, vb, hcol, hrow, u, v);
    } else {
        row_terms<N, R, REALH>(vb, hcol, hrow, u, v);
    }
}

// issue the loads of tile `tile` into one LDS buffer: this wave's planes and its share of
// the table units (two modes x 64 ADOs per 16-byte-per-lane load)
template <int N>
__device__ __forceinline__ void prefetch_pair_tile(const StageArgsT<float>& a, int tile, int w, int lane, f4* sig,
                                                   int2* tab)
{
    using S = PairShape<N>;
    const int64_t mp = a.mp;
    const uint32_t q = (uint32_t)tile * kTile + lane;
    const f4* planes = reinterpret_cast<const f4*>(a.yin);
#pragma unroll
    for (int c = 0; c < kPairsPerWave; ++c) {
        const int d = w * kPairsPerWave + c;
        if (d < S::D)
            __builtin_amdgcn_global_load_lds((const void*)(planes + (int64_t)d * mp + q),
                                             HEOM_LDS_PTR(sig + d * kTile), 16, 0, 0);
    }
    const int2* nb = a.nb + (int64_t)tile * kTile + (lane >> 5) * mp + (lane & 31) * 2;
#pragma unroll
    for (int r = 0; r < (S::TC / 2 + S::NWAVE - 1) / S::NWAVE; ++r) {
        const int u = w + r * S::NWAVE;
        if (u < a.table_units)
            __builtin_amdgcn_global_load_lds((const void*)(nb + 2 * u * mp), HEOM_LDS_PTR(tab + 2 * u * kTile), 16, 0,
                                             kPairTableCachePolicy);
    }
}

template <int STAGE>
__device__ __forceinline__ f4 pair_operand(const StageArgsT<float>& a, int64_t idx)
{
    if (STAGE == 2 || STAGE == 3)
        return __builtin_nontemporal_load(reinterpret_cast<const f4*>(a.y0) + idx);
    if (STAGE == 4)
        return __builtin_nontemporal_load(reinterpret_cast<const f4*>(a.pbuf) + idx);
    return f4{0.f, 0.f, 0.f, 0.f};
}

// RHS of one tile's planes for this wave, results in r0 (and r1 for stage 2); `issue`
// runs right after the first plane's gathers are in flight (the previous tile's stores
// and the next tile's prefetch, as in stage_pipelined.hip).
template <int N, int NT, int STAGE, bool REALH, bool PLUS, typename Issue>
__device__ __forceinline__ void compute_pair_tile(const StageArgsT<float>& a, const float2* __restrict__ hs_all,
                                                  const float2* __restrict__ gam, const PairTerm* __restrict__ pterms,
                                                  int tile, int w, int lane, float sq_lane, const f4* sig,
                                                  const int2* tab, f4 (&r0)[kPairsPerWave],
                                                  f4 (&r1)[kPairsPerWave], Issue&& issue)
{
    using S = PairShape<N>;
    constexpr int CW = kPairsPerWave;
    const int64_t mp = a.mp;
    const uint32_t q = (uint32_t)tile * kTile + lane;
    auto opaque = [](const auto* p) {
        int z;
        asm volatile("s_mov_b32 %0, 0" : "=s"(z));
        return p + z;
    };
    const float2* hs = tile_hamiltonian<float, N>(hs_all, a.tile_inst, tile);
    const char* vb = reinterpret_cast<const char*>(sig + lane);
    const int cls = tile_bath_class(a, tile);
    gam += cls * (N * N);
    pterms += cls * (S::D * NT);
    const float2 S_ = a.pfss[q];

    // planes past the end (odd plane count) compute a copy of the last plane and are not stored
    auto plane = [&](int c) { return min(w * CW + c, S::D - 1); };
    // one buffer resource per plane, its range ending at the zero slot (the first slot
    // after every ADO): a gather of the zero slot is out of range and reads zeros
    const int nrec = a.zero_slot * (int)sizeof(f4);
    auto gather = [&](__amdgpu_buffer_rsrc_t rsrc, uint32_t idx) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, idx * (uint32_t)sizeof(f4), 0, 0);
        return f4{__int_as_float((int)v[0]), __int_as_float((int)v[1]), __int_as_float((int)v[2]),
                  __int_as_float((int)v[3])};
    };
    // one plane's gathers in flight at a time: the next plane's are issued once this
    // plane's terms are summed
    struct Loads {
        int2 nbv[NT];
        f4 xp[NT], xm[NT];
    } L;
    auto load = [&](int c) {
        const int d = plane(c);
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(reinterpret_cast<const f4*>(a.yin) + d * mp), (short)0, nrec, 0x00020000);
        const PairTerm* te = opaque(pterms) + d * NT;
#pragma unroll
        for (int k = 0; k < NT; ++k)
            L.nbv[k] = tab[te[k].mode * kTile + lane];
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            if constexpr (PLUS)
                L.xp[k] = gather(rsrc, (uint32_t)L.nbv[k].x >> 6);
            L.xm[k] = gather(rsrc, (uint32_t)L.nbv[k].y);
        }
    };

    f4 epv[CW];
#pragma unroll
    for (int c = 0; c < CW; ++c)
        epv[c] = pair_operand<STAGE>(a, (int64_t)plane(c) * mp + q);
    load(0);
    __builtin_amdgcn_sched_barrier(0);
    issue();
    __builtin_amdgcn_sched_barrier(0);

#pragma unroll
    for (int c = 0; c < CW; ++c) {
        const int d = plane(c);
        // with two terms the second plane's gathers go out before its commutator, which
        // covers their latency; with more, they would not fit the registers beside the
        // commutator's reads and go out after it (the other waves cover the latency)
        constexpr bool kEarly = NT <= 2;
        if (c > 0 && kEarly) {
            __builtin_amdgcn_sched_barrier(0);
            load(c);
        }
        int er0, ec0, er1, ec1;
        bool valid1;
        plane_elements<N>(d, er0, ec0, er1, ec1, valid1);
        const bool offdiag = d < pair_count(N);
        const PairTerm* te = opaque(pterms) + d * NT;
        const float2* hc = opaque(hs);

        // [H, sigma] of the plane's two elements from rows er0 and er1 (partners: the other
        // row of an off-diagonal plane, the row itself on the diagonal)
        const int p0 = offdiag ? er1 : er0, p1 = offdiag ? er0 : er1;
        float2 u0, v0, u1, v1;
#pragma unroll 1  // one copy of the row switch
        for (int t = 0; t < 2; ++t) {
            const int r = t ? er1 : er0, pp = t ? p1 : p0;
            float2 u, v;
            row_terms_of<N, REALH>(r, vb, hc + pp, hc + pp * N, u, v);
            if (t) {
                u1 = u;
                v1 = v;
            } else {
                u0 = u;
                v0 = v;
            }
        }
        const float2 w0 = offdiag ? v1 : v0, w1 = offdiag ? v0 : v1;
        const float2 cm0 = make_float2(u0.x + w0.x, u0.y + w0.y), cm1 = make_float2(u1.x + w1.x, u1.y + w1.y);
        if (c > 0 && !kEarly) {
            __builtin_amdgcn_sched_barrier(0);
            load(c);
        }
        const f4 own = sig[d * kTile + lane];
        const float2 o0 = make_float2(own.x, own.y), o1 = make_float2(own.z, own.w);
        const float2 g0 = opaque(gam)[er0 * N + ec0], g1 = opaque(gam)[er1 * N + ec1];
        float2 f0 = cfma(make_float2(S_.x + g0.x, S_.y + g0.y), o0, make_float2(-cm0.y, cm0.x));
        float2 f1 = cfma(make_float2(S_.x + g1.x, S_.y + g1.y), o1, make_float2(-cm1.y, cm1.x));

#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const int v = L.nbv[k].x & 63;
            if constexpr (PLUS) {
                const float sp = lane_value(sq_lane, v + 1);
                const float cp0 = te[k].plus_c[0] * sp, cp1 = te[k].plus_c[1] * sp;
                f0.x = fmaf(-cp0, L.xp[k].y, f0.x);
                f0.y = fmaf(cp0, L.xp[k].x, f0.y);
                f1.x = fmaf(-cp1, L.xp[k].w, f1.x);
                f1.y = fmaf(cp1, L.xp[k].z, f1.y);
            }
            const float sm = lane_value(sq_lane, v);
            f0 = cfma(te[k].minus_c[0], make_float2(sm * L.xm[k].x, sm * L.xm[k].y), f0);
            f1 = cfma(te[k].minus_c[1], make_float2(sm * L.xm[k].z, sm * L.xm[k].w), f1);
        }
        if (!valid1)
            f1 = make_float2(0.f, 0.f);
        float2 a0, b0, a1 = make_float2(0.f, 0.f), b1 = make_float2(0.f, 0.f);
        const f4 ep = epv[c];
        stage_values<STAGE, float>(a.c1, o0, f0, make_float2(ep.x, ep.y), a0, a1);
        stage_values<STAGE, float>(a.c1, o1, f1, make_float2(ep.z, ep.w), b0, b1);
        r0[c] = f4{a0.x, a0.y, b0.x, b0.y};
        r1[c] = f4{a1.x, a1.y, b1.x, b1.y};
    }
}

template <int N, int STAGE>
__device__ __forceinline__ void store_pair_tile(const StageArgsT<float>& a, int tile, int w, int lane,
                                                const f4 (&r0)[kPairsPerWave], const f4 (&r1)[kPairsPerWave])
{
    using S = PairShape<N>;
    const uint32_t q = (uint32_t)tile * kTile + lane;
#pragma unroll
    for (int c = 0; c < kPairsPerWave; ++c) {
        const int d = w * kPairsPerWave + c;
        if (d < S::D) {
            const int64_t idx = (int64_t)d * a.mp + q;
            __builtin_nontemporal_store(r0[c], reinterpret_cast<f4*>(a.yout) + idx);
            if (STAGE == 2)
                __builtin_nontemporal_store(r1[c], reinterpret_cast<f4*>(a.pbuf) + idx);
        }
    }
}

template <int N, int NT, int STAGE, bool REALH>
__global__ __launch_bounds__(64 * PairShape<N>::NWAVE) void rk4_stage_pair(StageArgsT<float> a,
                                                                          const float2* __restrict__ hs,
                                                                          const float2* __restrict__ gam,
                                                                          const PairTerm* __restrict__ pterms)
{
    using S = PairShape<N>;
    __shared__ __align__(16) f4 sigA[S::DS * kTile];
    __shared__ __align__(16) f4 sigB[S::DS * kTile];
    __shared__ __align__(16) int2 tabA[S::TC * kTile];
    __shared__ __align__(16) int2 tabB[S::TC * kTile];

    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    const int x = blockIdx.x % kNumXcd, g = blockIdx.x / kNumXcd, gx = gridDim.x / kNumXcd;
    const int chunk = a.xcd_chunk, nt = a.ntiles;
    int j = g;
    int t = chunk_tile(x, j, chunk);
    if (t >= nt)
        return;  // the whole workgroup leaves together
    auto phys = [&](int tt) { return a.reverse ? nt - 1 - tt : tt; };
    const float sq_lane = sqrtf((float)lane);

    f4 r0[kPairsPerWave], r1[kPairsPerWave];
#pragma unroll
    for (int c = 0; c < kPairsPerWave; ++c)
        r0[c] = r1[c] = f4{0.f, 0.f, 0.f, 0.f};

    prefetch_pair_tile<N>(a, phys(t), w, lane, sigA, tabA);
    __syncthreads();

    auto compute = [&](int tile, const f4* sig, const int2* tab, auto&& issue) {
        if (tile_has_plus(a, tile))
            compute_pair_tile<N, NT, STAGE, REALH, true>(a, hs, gam, pterms, tile, w, lane, sq_lane, sig, tab, r0,
                                                          r1, issue);
        else
            compute_pair_tile<N, NT, STAGE, REALH, false>(a, hs, gam, pterms, tile, w, lane, sq_lane, sig, tab, r0,
                                                           r1, issue);
    };

    // the buffers swap roles every tile (pointers, so the tile body is compiled once)
    f4* sig_cur = sigA;
    f4* sig_nxt = sigB;
    int2* tab_cur = tabA;
    int2* tab_nxt = tabB;
    int prev = -1;
    for (;;) {
        j += gx;
        const int tn = chunk_tile(x, j, chunk);
        compute(phys(t), sig_cur, tab_cur, [&] {
            if (prev >= 0)
                store_pair_tile<N, STAGE>(a, phys(prev), w, lane, r0, r1);
            if (tn < nt)
                prefetch_pair_tile<N>(a, phys(tn), w, lane, sig_nxt, tab_nxt);
        });
        prev = t;
        t = tn;
        __syncthreads();  // the next buffer is complete, and nobody reads this one any more
        if (t >= nt)
            break;
        f4* ts = sig_cur;
        sig_cur = sig_nxt;
        sig_nxt = ts;
        int2* tt = tab_cur;
        tab_cur = tab_nxt;
        tab_nxt = tt;
    }
    store_pair_tile<N, STAGE>(a, phys(prev), w, lane, r0, r1);
}

template <int N, int NT, bool REALH>
void launch_t(int stage, const StageArgsT<float>& a, const ElementArgsT<float>& el, const PairTerm* pt, int wgs,
              hipStream_t st)
{
    const dim3 grid(wgs), block(64 * PairShape<N>::NWAVE);
    switch (stage) {
    case 0: hipLaunchKernelGGL((rk4_stage_pair<N, NT, 0, REALH>), grid, block, 0, st, a, el.hs, el.gam, pt); break;
    case 1: hipLaunchKernelGGL((rk4_stage_pair<N, NT, 1, REALH>), grid, block, 0, st, a, el.hs, el.gam, pt); break;
    case 2: hipLaunchKernelGGL((rk4_stage_pair<N, NT, 2, REALH>), grid, block, 0, st, a, el.hs, el.gam, pt); break;
    case 3: hipLaunchKernelGGL((rk4_stage_pair<N, NT, 3, REALH>), grid, block, 0, st, a, el.hs, el.gam, pt); break;
    default: hipLaunchKernelGGL((rk4_stage_pair<N, NT, 4, REALH>), grid, block, 0, st, a, el.hs, el.gam, pt); break;
    }
}

template <int N, bool REALH>
void launch_n(int nt, int stage, const StageArgsT<float>& a, const ElementArgsT<float>& el, const PairTerm* pt,
              int wgs, hipStream_t st)
{
    switch (nt) {
    case 2: launch_t<N, 2, REALH>(stage, a, el, pt, wgs, st); break;
    case 4: launch_t<N, 4, REALH>(stage, a, el, pt, wgs, st); break;
    default: launch_t<N, 8, REALH>(stage, a, el, pt, wgs, st); break;
    }
}

template <bool REALH>
void launch_r(int N, int nt, int stage, const StageArgsT<float>& a, const ElementArgsT<float>& el,
              const PairTerm* pt, int wgs, hipStream_t st)
{
    switch (N) {
    case 4: launch_n<4, REALH>(nt, stage, a, el, pt, wgs, st); break;
    case 5: launch_n<5, REALH>(nt, stage, a, el, pt, wgs, st); break;
    case 6: launch_n<6, REALH>(nt, stage, a, el, pt, wgs, st); break;
    case 7: launch_n<7, REALH>(nt, stage, a, el, pt, wgs, st); break;
    default: launch_n<8, REALH>(nt, stage, a, el, pt, wgs, st); break;
    }
}

// table rows one LDS buffer of the paired kernel holds for N states
int pair_table_capacity(int N)
{
    switch (N) {
    case 4: return PairShape<4>::TC;
    case 5: return PairShape<5>::TC;
    case 6: return PairShape<6>::TC;
    case 7: return PairShape<7>::TC;
    default: return PairShape<8>::TC;
    }
}

}  // namespace

bool pair_supported(int N, int W, int64_t mp)
{
    // four to eight states (a workgroup of 2 to 16 wavefronts per CU) and a table slice
    // that fits the LDS buffer; each plane has its own buffer resource (a plane of at most
    // 2^26 slots of 16 bytes), so the state size is not limited beyond that
    (void)mp;
    return N >= 4 && N <= kMaxStates && ((W + 2) & ~1) <= pair_table_capacity(N);
}

void launch_stage_pair(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<float>& a,
                       const ElementArgsT<float>& el, const PairTerm* pterms, int workgroups, hipStream_t s)
{
    if (real_h)
        launch_r<true>(N, nterm_slots, stage, a, el, pterms, workgroups, s);
    else
        launch_r<false>(N, nterm_slots, stage, a, el, pterms, workgroups, s);
}

}  // namespace heom_hip
