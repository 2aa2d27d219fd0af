// stage_pipelined.hip -- persistent, software-pipelined form of the stage kernel.
//
// Same RHS and RK4 stage arithmetic as stage_kernel.hip (dm-heom/cl/heom/heom_ode.cl
// :24-451 with rk_weighted_add.cl:19-106 fused in), organised so that the streaming
// loads of one ADO tile overlap the neighbour gathers and arithmetic of the previous
// one:
//   * one workgroup per CU (N * ceil(N/4) wavefronts) walks a run of 64-ADO tiles: the
//     XCD's workgroups take turns through its sequence of tiles (chunk_tile), so they
//     move through the hierarchy together;
//   * two LDS buffers (A, B), each holding a tile's N*N element planes and its slice of
//     the mode-major neighbour table; while the tile in one buffer is computed, the
//     next tile streams into the other with direct-to-LDS loads (no registers);
//   * the neighbour indices come from the LDS table slice (modes 0..W, W being the
//     all-zero-slot row that padding terms name);
//   * a tile's results stay in registers across the barrier and are stored at the start
//     of the next tile, so no wave waits for its own stores to complete.
// The two buffers are separate LDS objects used with compile-time selection (the loop
// is unrolled by two); that lets the compiler see that loads into one buffer do not
// alias reads of the other, so it does not wait for the prefetch before computing.
// LDS element slots are laid out per matrix row with the row padded to the wavefronts'
// 4-column blocks (pipe_row_slots).  Instantiated for complex128 state; complex64 state
// runs the paired kernel (stage_pair.hip), whose 16-byte gathers serve two elements.
#include "heom_device.hpp"

namespace heom_hip {

namespace {

#define HEOM_LDS_PTR(p) ((__attribute__((address_space(3))) void*)(p))

// cache policy of the table loads: non-temporal (each table row is read once a stage)
constexpr int kTableCachePolicy = 2;


template <typename R, int N>
struct Shape {
    static constexpr int CW = kStageColumns;
    static constexpr int NJW = (N + CW - 1) / CW;  // wavefronts per row
    static constexpr int NWAVE = N * NJW;
    static constexpr int RS = pipe_row_slots(N);   // LDS element slots per matrix row
    static constexpr int TC = pipe_table_capacity(N, (int)sizeof(c2<R>));
};

// issue the loads of tile `tile` into one LDS buffer: this wave's own elements and its
// share of the table units (two modes x 64 ADOs per 16-byte-per-lane load)
template <typename R, int N>
__device__ __forceinline__ void prefetch_tile(const StageArgsT<R>& a, int tile, int w, int i, int jb, int lane,
                                              c2<R>* sig, int2* tab)
{
    using S = Shape<R, N>;
    const int64_t mp = a.mp;
    static_assert(sizeof(c2<R>) == 16, "complex128 state: one 16-byte element per lane");
    const uint32_t q = (uint32_t)tile * kTile + lane;
#pragma unroll
    for (int c = 0; c < S::CW; ++c) {
        const int j = jb + c;
        if (j < N)
            __builtin_amdgcn_global_load_lds((const void*)(a.yin + (int64_t)(i * N + j) * mp + q),
                                             HEOM_LDS_PTR(sig + (i * S::RS + j) * kTile), 16, 0, 0);
    }
    // table rows in pairs: lanes 0-31 load row 2u, lanes 32-63 row 2u+1, two ADOs per lane
    const int2* nb = a.nb + (int64_t)tile * kTile + (lane >> 5) * mp + (lane & 31) * 2;
#pragma unroll
    for (int r = 0; r < (S::TC / 2 + S::NWAVE - 1) / S::NWAVE; ++r) {
        const int u = w + r * S::NWAVE;
        if (u < a.table_units)
            __builtin_amdgcn_global_load_lds((const void*)(nb + 2 * u * mp), HEOM_LDS_PTR(tab + 2 * u * kTile), 16, 0,
                                             kTableCachePolicy);
    }
}

// RHS of one tile from LDS (own elements, table) plus neighbour gathers; the stage
// results of the wave's elements land in r0 (and r1 for stage 2).  `issue` runs right
// after the first column's gathers are in flight: the previous tile's stores and the
// next tile's prefetch go out there, behind those gathers, so that waiting for the
// gathers does not mean waiting for the prefetch (loads complete in issue order).
template <typename R, int N, int NT, int STAGE, bool REALH, bool PLUS, typename Issue>
__device__ __forceinline__ void compute_tile(const StageArgsT<R>& a, const c2<R>* __restrict__ hs_all,
                                             const c2<R>* __restrict__ gam, const TermT<R>* __restrict__ terms,
                                             int tile, int i, int jb, int lane, R sq_lane, const c2<R>* sig,
                                             const int2* tab, c2<R> (&r0)[kStageColumns],
                                             c2<R> (&r1)[kStageColumns], Issue&& issue)
{
    using C = c2<R>;
    using S = Shape<R, N>;
    constexpr int CW = kStageColumns;
    const int64_t mp = a.mp;
    const uint32_t q = (uint32_t)tile * kTile + lane;
    // The element constants are the same for every tile and column block.  Offsets the
    // compiler cannot see through keep their (scalar, cached) loads next to the column
    // that uses them instead of hoisting all of them into scalar registers at once.
    auto opaque = [](const auto* p) {
        int z;
        asm volatile("s_mov_b32 %0, 0" : "=s"(z));
        return p + z;
    };
    const C* hs = tile_hamiltonian<R, N>(hs_all, a.tile_inst, tile);
    const int cls = tile_bath_class(a, tile);
    gam += cls * (N * N);
    terms += cls * (N * N * NT);
    const C S_ = a.pfss[q];
    auto slot = [](int k, int j) { return (k * S::RS + j) * kTile; };

    // Column c's neighbour indices (from the LDS table; padding terms name mode W,
    // whose row is the zero slot) and its gathers (missing and masked neighbours read
    // the zero slot, so the code is branch-free).  Column 0's gathers go out before the
    // previous tile's stores and the next tile's prefetch (`issue`); each later column
    // issues its own at its start, and the commutator from LDS covers their latency.
    struct Loads {
        int2 nbv[NT];
        C xp[NT], xm[NT];
    } L[CW];
    // Columns past the row end (N not a multiple of 4) compute a copy of the last column
    // and are not stored: every wavefront runs the same straight-line code, which keeps
    // the per-column arrays in registers.
    auto col = [&](int c) { return min(jb + c, N - 1); };
    // Neighbour gathers as raw buffer loads, one buffer resource per element plane whose
    // range ends at the zero slot (the first slot after every ADO): an edge to the zero
    // slot (missing or masked neighbour; most minus lanes of the deepest tier) is out of
    // range, which returns zero without a memory request.
    const int nrec = a.zero_slot * (int)sizeof(C);
    auto gather = [&](__amdgpu_buffer_rsrc_t rsrc, uint32_t idx) {
        const uint32_t vo = idx * (uint32_t)sizeof(C);
        if constexpr (sizeof(C) == 16) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, vo, 0, 0);
            return mk2<R>(__hiloint2double((int)v[1], (int)v[0]), __hiloint2double((int)v[3], (int)v[2]));
        } else {
            const auto v = __builtin_amdgcn_raw_buffer_load_b64(rsrc, vo, 0, 0);
            return mk2<R>(__int_as_float((int)v[0]), __int_as_float((int)v[1]));
        }
    };
    auto load = [&](int c) {
        const int e = i * N + col(c);
        const __amdgpu_buffer_rsrc_t rsrc =
            __builtin_amdgcn_make_buffer_rsrc((void*)(a.yin + e * mp), (short)0, nrec, 0x00020000);
        const TermT<R>* te = opaque(terms) + e * NT;
#pragma unroll
        for (int k = 0; k < NT; ++k)
            L[c].nbv[k] = tab[te[k].mode * kTile + lane];
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            if constexpr (PLUS)
                L[c].xp[k] = gather(rsrc, (uint32_t)L[c].nbv[k].x >> 6);
            L[c].xm[k] = gather(rsrc, (uint32_t)L[c].nbv[k].y);
        }
    };

    // the epilogue operands (y_n or P) of all columns go out first: they come from HBM
    // and are needed only at the end of each column
    C epv[CW];
#pragma unroll
    for (int c = 0; c < CW; ++c)
        epv[c] = stage_operand<STAGE, R>(a, (int64_t)(i * N + col(c)) * mp + q);
    load(0);
    __builtin_amdgcn_sched_barrier(0);
    issue();
    __builtin_amdgcn_sched_barrier(0);

#pragma unroll
    for (int c = 0; c < CW; ++c) {
        const int j = col(c);
        const int e = i * N + j;
        const C ep = epv[c];
        if (c > 0)
            load(c);
        const TermT<R>* te = opaque(terms) + e * NT;
        const C* hc = opaque(hs);

        // commutator c = sum_k sigma_ik H_kj - H_ik sigma_kj, row i and column j from the tile
        C cm = mk2<R>(R(0), R(0));
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const C sik = sig[slot(i, k) + lane];
            const C skj = sig[slot(k, j) + lane];
            const C hkj = hc[k * N + j];
            const C hik = hc[i * N + k];
            if (REALH) {
                cm.x = fma(sik.x, hkj.x, fma(-hik.x, skj.x, cm.x));
                cm.y = fma(sik.y, hkj.x, fma(-hik.x, skj.y, cm.y));
            } else {
                cm = cfma(sik, hkj, cm);
                cm = cfma(mk2<R>(-hik.x, -hik.y), skj, cm);
            }
        }
        const C s = sig[slot(i, j) + lane];
        const C g = opaque(gam)[e];
        C f = cfma(mk2<R>(S_.x + g.x, S_.y + g.y), s, mk2<R>(-cm.y, cm.x));

#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const int v = L[c].nbv[k].x & 63;
            if constexpr (PLUS) {
                const R cp = te[k].plus_c * lane_value(sq_lane, v + 1);
                f.x = fma(-cp, L[c].xp[k].y, f.x);
                f.y = fma(cp, L[c].xp[k].x, f.y);
            }
            const R sm = lane_value(sq_lane, v);
            f = cfma(te[k].minus_c, mk2<R>(sm * L[c].xm[k].x, sm * L[c].xm[k].y), f);
        }
        stage_values<STAGE, R>(a.c1, s, f, ep, r0[c], r1[c]);
    }
}

template <typename R, int N, int STAGE>
__device__ __forceinline__ void store_tile(const StageArgsT<R>& a, int tile, int i, int jb, int lane,
                                           const c2<R> (&r0)[kStageColumns], const c2<R> (&r1)[kStageColumns])
{
    const uint32_t q = (uint32_t)tile * kTile + lane;
#pragma unroll
    for (int c = 0; c < kStageColumns; ++c)
        if (jb + c < N)
            stage_store<STAGE, R>(a, (int64_t)(i * N + jb + c) * a.mp + q, r0[c], r1[c]);
}

template <typename R, int N, int NT, int STAGE, bool REALH>
__global__ __launch_bounds__(64 * stage_waves(N)) void rk4_stage_pipe(StageArgsT<R> a,
                                                                       const c2<R>* __restrict__ hs,
                                                                       const c2<R>* __restrict__ gam,
                                                                       const TermT<R>* __restrict__ terms)
{
    using S = Shape<R, N>;
    using C = c2<R>;
    __shared__ __align__(16) C sigA[N * S::RS * kTile];
    __shared__ __align__(16) C sigB[N * S::RS * kTile];
    __shared__ __align__(16) int2 tabA[S::TC * kTile];
    __shared__ __align__(16) int2 tabB[S::TC * kTile];

    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = w / S::NJW;
    const int jb = (w % S::NJW) * S::CW;

    // tiles of this workgroup: the j-th tile of XCD x is chunk_tile(x, j, chunk); the gx
    // workgroups of the XCD take j = g, g + gx, ...  The sequence increases with j, so
    // the first tile past the end ends the walk.
    const int x = blockIdx.x % kNumXcd, g = blockIdx.x / kNumXcd, gx = gridDim.x / kNumXcd;
    const int chunk = a.xcd_chunk, nt = a.ntiles;
    int j = g;
    int t = chunk_tile(x, j, chunk);
    if (t >= nt)
        return;  // the whole workgroup leaves together
    auto phys = [&](int tt) { return a.reverse ? nt - 1 - tt : tt; };
    const R sq_lane = sqrt((R)lane);

    C r0[kStageColumns], r1[kStageColumns];
#pragma unroll
    for (int c = 0; c < kStageColumns; ++c)
        r0[c] = r1[c] = mk2<R>(R(0), R(0));

    prefetch_tile<R, N>(a, phys(t), w, i, jb, lane, sigA, tabA);
    __syncthreads();  // waits for the loads of this wave, then for every wave

    // a tile whose ADOs have no plus neighbours (tile_plus) runs the variant without plus
    // gathers; the branch is uniform over the workgroup
    auto compute = [&](int tile, const C* sig, const int2* tab, auto&& issue) {
        if (tile_has_plus(a, tile))
            compute_tile<R, N, NT, STAGE, REALH, true>(a, hs, gam, terms, tile, i, jb, lane, sq_lane, sig, tab, r0, r1,
                                                       issue);
        else
            compute_tile<R, N, NT, STAGE, REALH, false>(a, hs, gam, terms, tile, i, jb, lane, sq_lane, sig, tab, r0,
                                                        r1, issue);
    };

    int prev = -1;
    for (;;) {
        // tile t is in A; the next one goes to B
        j += gx;
        int tn = chunk_tile(x, j, chunk);
        compute(phys(t), sigA, tabA, [&] {
            if (prev >= 0)
                store_tile<R, N, STAGE>(a, phys(prev), i, jb, lane, r0, r1);
            if (tn < nt)
                prefetch_tile<R, N>(a, phys(tn), w, i, jb, lane, sigB, tabB);
        });
        prev = t;
        t = tn;
        __syncthreads();  // B is complete, and nobody reads A any more
        if (t >= nt)
            break;

        // tile t is in B; the next one goes to A
        j += gx;
        tn = chunk_tile(x, j, chunk);
        compute(phys(t), sigB, tabB, [&] {
            store_tile<R, N, STAGE>(a, phys(prev), i, jb, lane, r0, r1);
            if (tn < nt)
                prefetch_tile<R, N>(a, phys(tn), w, i, jb, lane, sigA, tabA);
        });
        prev = t;
        t = tn;
        __syncthreads();
        if (t >= nt)
            break;
    }
    store_tile<R, N, STAGE>(a, phys(prev), i, jb, lane, r0, r1);
}

template <typename R, int N, int NT, bool REALH>
void launch_t(int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el, int wgs, hipStream_t st)
{
    const dim3 grid(wgs), block(64 * stage_waves(N));
    switch (stage) {
    case 0: hipLaunchKernelGGL((rk4_stage_pipe<R, N, NT, 0, REALH>), grid, block, 0, st, a, el.hs, el.gam, el.terms); break;
    case 1: hipLaunchKernelGGL((rk4_stage_pipe<R, N, NT, 1, REALH>), grid, block, 0, st, a, el.hs, el.gam, el.terms); break;
    case 2: hipLaunchKernelGGL((rk4_stage_pipe<R, N, NT, 2, REALH>), grid, block, 0, st, a, el.hs, el.gam, el.terms); break;
    case 3: hipLaunchKernelGGL((rk4_stage_pipe<R, N, NT, 3, REALH>), grid, block, 0, st, a, el.hs, el.gam, el.terms); break;
    default: hipLaunchKernelGGL((rk4_stage_pipe<R, N, NT, 4, REALH>), grid, block, 0, st, a, el.hs, el.gam, el.terms); break;
    }
}

template <typename R, int N, bool REALH>
void launch_n(int nt, int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el, int wgs, hipStream_t st)
{
    switch (nt) {
    case 2: launch_t<R, N, 2, REALH>(stage, a, el, wgs, st); break;
    case 4: launch_t<R, N, 4, REALH>(stage, a, el, wgs, st); break;
    default: launch_t<R, N, 8, REALH>(stage, a, el, wgs, st); break;
    }
}

template <typename R, bool REALH>
void launch_r(int N, int nt, int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el, int wgs, hipStream_t st)
{
    switch (N) {
    case 6: launch_n<R, 6, REALH>(nt, stage, a, el, wgs, st); break;
    case 7: launch_n<R, 7, REALH>(nt, stage, a, el, wgs, st); break;
    default: launch_n<R, 8, REALH>(nt, stage, a, el, wgs, st); break;
    }
}

template <typename R>
void launch_any(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el,
                int workgroups, hipStream_t s)
{
    if (real_h)
        launch_r<R, true>(N, nterm_slots, stage, a, el, workgroups, s);
    else
        launch_r<R, false>(N, nterm_slots, stage, a, el, workgroups, s);
}

}  // namespace

bool pipe_supported(int N, int W, int64_t mp, int complex_bytes)
{
    // complex128 state (complex64 state runs the paired kernel, stage_pair.hip); wide
    // enough workgroups to carry a CU on their own and a table slice that fits the LDS
    // buffer.  Each element plane has its own buffer resource (at most 2^26 slots of 16
    // bytes), so the state size is not limited beyond that.
    (void)mp;
    const bool shape = complex_bytes == (int)sizeof(double2) && N >= 6 && N <= kMaxStates;
    return shape && ((W + 2) & ~1) <= pipe_table_capacity(N, complex_bytes);
}

void launch_stage_pipelined(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<double>& a,
                            const ElementArgsT<double>& el, int workgroups, hipStream_t s)
{
    launch_any<double>(N, nterm_slots, real_h, stage, a, el, workgroups, s);
}

}  // namespace heom_hip
