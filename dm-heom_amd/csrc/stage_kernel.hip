// stage_kernel.hip -- the HEOM RHS with the RK4 stage combination fused into its
// epilogue (dm-heom/cl/heom/heom_ode.cl:24-451 plus rk_weighted_add.cl:19-106).
//
// Work split: a workgroup owns a tile of 64 consecutive ADOs (lane = ADO); each
// wavefront computes a block of CW = 4 columns j of one matrix row i for those 64 ADOs
// (N * ceil(N/4) wavefronts per workgroup, 16 for N = 8).  Every global access of a
// wavefront is therefore one element plane at 64 consecutive (or, for neighbours,
// nearly consecutive) ADOs, and everything that depends only on the element --
// H / hbar, the "same" constant, the neighbour-term list -- is wave-uniform and comes
// through scalar loads.  Per wavefront:
//   * own elements sigma_ij (j in the block) -> registers and the LDS tile of the 64
//     ADOs x N^2 elements; barrier;
//   * row i of sigma from the tile (shared by the block's columns), then per column j:
//     commutator (i/hbar)(sigma H - H sigma)_ij with column j from the tile, the "same"
//     term (S_n + gamma_ij) sigma_ij, and the element's neighbour terms: for each
//     (mode m, plus constant, minus constant) triple built on the host from the
//     bath-coupling list, the plus / minus neighbour indices and n_m come from one
//     8-byte entry of the mode-major table and two gathers from the element plane give
//     f += i c_p sqrt(n_m+1) sigma^{n+e_m} + c_m sqrt(n_m) sigma^{n-e_m}
//     (missing or masked edges read the zero slot);
//   * the RK4 epilogue of heom_device.hpp.
// Any coupling (several baths per state, shared baths, uncoupled states) runs through
// the same code; the term loop is unrolled so a column's table loads and gathers are
// all in flight together.  The file is compiled three times: for complex128 state and
// FP64 arithmetic, with HEOM_STAGE_PRECISION defined (stage_kernel_f32.hip) for complex64
// state and FP32 arithmetic, and with HEOM_STAGE_ADJOINT defined (stage_kernel_adj.hip)
// for the adjoint operator L^dag in FP64.
//
// L^dag (ADJ = true) has the same neighbour structure with the roles of the two edges
// exchanged: the plus term i c_p sqrt(n_m+1) sigma^{n+e_m} of ADO n is, transposed, a
// contribution -i c_p sqrt(n_m) mu^{n-e_m} to ADO n, and the minus term
// c_m sqrt(n_m) sigma^{n-e_m} becomes conj(c_m) sqrt(n_m+1) mu^{n+e_m}.  The host passes
// -c_p and conj(c_m) in the term list, -H^dag / hbar for H / hbar and the conjugate
// "same" constants; the kernel exchanges which gather each constant multiplies.
#include "heom_device.hpp"

namespace heom_hip {

namespace {

// One tile, or with SPLIT > 1 row block rb of SPLIT of it (small hierarchies: SPLIT
// workgroups share a tile, each loading all of it into LDS for the commutator's columns
// and computing 1 / SPLIT of its wavefronts, so that more CUs share the tile's memory
// instructions); PLUS = false leaves out the plus gathers (no ADO of the tile has a plus
// neighbour, so they would all read the zero slot).
template <typename R, int N, int NT, int STAGE, bool REALH, bool ADJ, bool PLUS, int SPLIT>
__device__ __forceinline__ void stage_tile(const StageArgsT<R>& a, const c2<R>* __restrict__ hs_all,
                                           const c2<R>* __restrict__ gam, const TermT<R>* __restrict__ terms,
                                           int tile, int rb, c2<R>* sig)
{
    using C = c2<R>;
    constexpr int CW = kStageColumns;
    constexpr int NJW = (N + CW - 1) / CW;  // wavefronts per row
    constexpr int WPG = stage_waves(N) / SPLIT;  // wavefronts of this workgroup

    const int lane = threadIdx.x & 63;
    const int wl = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int w = rb * WPG + wl;  // this wavefront's place among the tile's
    const int i = w / NJW;
    const int jb = (w % NJW) * CW;
    const C* hs = tile_hamiltonian<R, N>(hs_all, a.tile_inst, tile);
    const int64_t mp = a.mp;
    const uint32_t q = (uint32_t)tile * kTile + lane;
    // the adjoint's fourth stage: the top ADO slot of the tile's instance, read up front
    // (a wave-uniform load that the epilogue would otherwise wait for), -1 without a source
    int64_t qtop = -1;
    if constexpr (ADJ && STAGE == 4)
        if (a.top_src)
            qtop = a.top_q[a.tile_inst ? a.tile_inst[tile] : 0];

    C s[CW];
#pragma unroll
    for (int c = 0; c < CW; ++c) {
        const int j = jb + c;
        if (j < N) {
            s[c] = a.yin[(int64_t)(i * N + j) * mp + q];
            sig[(i * N + j) * kTile + lane] = s[c];
        } else {
            s[c] = mk2<R>(R(0), R(0));
        }
    }
    // the elements of the other row blocks, for the commutator's columns
#pragma unroll
    for (int r = 1; r < SPLIT; ++r) {
        const int w2 = ((rb + r) % SPLIT) * WPG + wl;
        const int i2 = w2 / NJW, jb2 = (w2 % NJW) * CW;
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            const int j = jb2 + c;
            if (j < N)
                sig[(i2 * N + j) * kTile + lane] = a.yin[(int64_t)(i2 * N + j) * mp + q];
        }
    }
    // with few terms per element, the table entries and the epilogue operands of all
    // columns are requested before the barrier: they do not depend on the LDS tile, so
    // their latency overlaps the tile's own loads
    constexpr bool kHoist = NT <= 2;
    int2 nbh[kHoist ? CW : 1][kHoist ? NT : 1];
    C eph[kHoist ? CW : 1];
    if (kHoist) {
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            const int j = jb + c;
            if (j < N) {
#pragma unroll
                for (int k = 0; k < NT; ++k)
                    nbh[c][k] = a.nb[(int64_t)terms[(i * N + j) * NT + k].mode * mp + q];
                eph[c] = stage_operand<STAGE, R>(a, (int64_t)(i * N + j) * mp + q);
            }
        }
    }
    const C S = a.pfss[q];
    __syncthreads();

    C srow[N];
#pragma unroll
    for (int k = 0; k < N; ++k)
        srow[k] = sig[(i * N + k) * kTile + lane];
    const R sq_lane = sqrt((R)lane);  // sqrt(n), n = 0..63, one value per lane

#pragma unroll
    for (int c = 0; c < CW; ++c) {
        const int j = jb + c;
        if (j >= N)
            break;
        const int e = i * N + j;
        const C* plane = a.yin + (int64_t)e * mp;

        // neighbour terms of this element: tables and gathers first
        int2 nbv[NT];
#pragma unroll
        for (int k = 0; k < NT; ++k)
            nbv[k] = kHoist ? nbh[kHoist ? c : 0][kHoist ? k : 0] : a.nb[(int64_t)terms[e * NT + k].mode * mp + q];
        C xp[NT], xm[NT];
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            if constexpr (PLUS)
                xp[k] = plane[(uint32_t)nbv[k].x >> 6];
            xm[k] = plane[(uint32_t)nbv[k].y];
        }

        // commutator c = sum_k sigma_ik H_kj - H_ik sigma_kj
        C cm = mk2<R>(R(0), R(0));
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const C skj = sig[(k * N + j) * kTile + lane];
            const C hkj = hs[k * N + j];
            const C hik = hs[i * N + k];
            if (REALH) {
                cm.x = fma(srow[k].x, hkj.x, fma(-hik.x, skj.x, cm.x));
                cm.y = fma(srow[k].y, hkj.x, fma(-hik.x, skj.y, cm.y));
            } else {
                cm = cfma(srow[k], hkj, cm);
                cm = cfma(mk2<R>(-hik.x, -hik.y), skj, cm);
            }
        }
        const C g = gam[e];
        C f = cfma(mk2<R>(S.x + g.x, S.y + g.y), s[c], mk2<R>(-cm.y, cm.x));

#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const TermT<R> t = terms[e * NT + k];
            const int v = nbv[k].x & 63;
            if (ADJ) {
                const R cp = t.plus_c * lane_value(sq_lane, v);
                f.x = fma(-cp, xm[k].y, f.x);
                f.y = fma(cp, xm[k].x, f.y);
                if constexpr (PLUS) {
                    const R sp = lane_value(sq_lane, v + 1);
                    f = cfma(t.minus_c, mk2<R>(sp * xp[k].x, sp * xp[k].y), f);
                }
            } else {
                if constexpr (PLUS) {
                    const R cp = t.plus_c * lane_value(sq_lane, v + 1);
                    f.x = fma(-cp, xp[k].y, f.x);
                    f.y = fma(cp, xp[k].x, f.y);
                }
                const R sm = lane_value(sq_lane, v);
                f = cfma(t.minus_c, mk2<R>(sm * xm[k].x, sm * xm[k].y), f);
            }
        }

        if constexpr (ADJ && STAGE == 4) {
            // the observation source on the top ADO, added to the stored value as a launch
            // of its own would add it: the same bits, one launch less per step
            C r0, r1 = mk2<R>(R(0), R(0));
            stage_values<STAGE, R>(a.c1, s[c], f,
                                   kHoist ? eph[kHoist ? c : 0] : stage_operand<STAGE, R>(a, (int64_t)e * mp + q), r0,
                                   r1);
            if ((int64_t)q == qtop) {
                const C src = a.top_src[(int64_t)(a.tile_inst ? a.tile_inst[tile] : 0) * (N * N) + e];
                r0.x += src.x;
                r0.y += src.y;
            }
            stage_store<STAGE, R>(a, (int64_t)e * mp + q, r0, r1);
        } else if (kHoist) {
            C r0, r1 = mk2<R>(R(0), R(0));
            stage_values<STAGE, R>(a.c1, s[c], f, eph[kHoist ? c : 0], r0, r1);
            stage_store<STAGE, R>(a, (int64_t)e * mp + q, r0, r1);
        } else {
            stage_epilogue<STAGE, R>(a, (int64_t)e * mp + q, s[c], f);
        }
    }
}

template <typename R, int N, int NT, int STAGE, bool REALH, bool ADJ, int SPLIT>
__global__ __launch_bounds__(64 * stage_waves(N) / SPLIT) void rk4_stage(StageArgsT<R> a,
                                                                          const c2<R>* __restrict__ hs_all,
                                                                          const c2<R>* __restrict__ gam,
                                                                          const TermT<R>* __restrict__ terms)
{
    __shared__ c2<R> sig[N * N * kTile];
    const int t0 = xcd_tile(blockIdx.x / SPLIT, a.ntiles, a.xcd_chunk);
    const int rb = blockIdx.x % SPLIT;
    const int tile = a.reverse ? a.ntiles - 1 - t0 : t0;
    const int cls = tile_bath_class(a, tile);
    const c2<R>* gam_c = gam + cls * (N * N);
    const TermT<R>* terms_c = terms + cls * (N * N * NT);
    if (tile_has_plus(a, tile))
        stage_tile<R, N, NT, STAGE, REALH, ADJ, true, SPLIT>(a, hs_all, gam_c, terms_c, tile, rb, sig);
    else
        stage_tile<R, N, NT, STAGE, REALH, ADJ, false, SPLIT>(a, hs_all, gam_c, terms_c, tile, rb, sig);
}

template <typename R, int N, int NT, bool REALH, bool ADJ, int SPLIT>
void launch_s(int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el, hipStream_t st)
{
    const dim3 grid(a.ntiles * SPLIT), block(64 * stage_waves(N) / SPLIT);
    switch (stage) {
    case 0: hipLaunchKernelGGL((rk4_stage<R, N, NT, 0, REALH, ADJ, SPLIT>), grid, block, 0, st, a, el.hs, el.gam, el.terms); break;
    case 1: hipLaunchKernelGGL((rk4_stage<R, N, NT, 1, REALH, ADJ, SPLIT>), grid, block, 0, st, a, el.hs, el.gam, el.terms); break;
    case 2: hipLaunchKernelGGL((rk4_stage<R, N, NT, 2, REALH, ADJ, SPLIT>), grid, block, 0, st, a, el.hs, el.gam, el.terms); break;
    case 3: hipLaunchKernelGGL((rk4_stage<R, N, NT, 3, REALH, ADJ, SPLIT>), grid, block, 0, st, a, el.hs, el.gam, el.terms); break;
    default: hipLaunchKernelGGL((rk4_stage<R, N, NT, 4, REALH, ADJ, SPLIT>), grid, block, 0, st, a, el.hs, el.gam, el.terms); break;
    }
}

// row splits are compiled for N >= 5 (ten or more wavefronts per tile), where they divide
// the tile's wavefronts (tile_split_supported)
template <typename R, int N, int NT, bool REALH, bool ADJ>
void launch_t(int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el, hipStream_t st)
{
    if constexpr (N >= 5 && stage_waves(N) % 4 == 0) {
        if (a.split == 4)
            return launch_s<R, N, NT, REALH, ADJ, 4>(stage, a, el, st);
    }
    if constexpr (N >= 5 && stage_waves(N) % 2 == 0) {
        if (a.split == 2)
            return launch_s<R, N, NT, REALH, ADJ, 2>(stage, a, el, st);
    }
    launch_s<R, N, NT, REALH, ADJ, 1>(stage, a, el, st);
}

template <typename R, int N, bool REALH, bool ADJ>
void launch_n(int nt, int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el, hipStream_t st)
{
    switch (nt) {
    case 2: launch_t<R, N, 2, REALH, ADJ>(stage, a, el, st); break;
    case 4: launch_t<R, N, 4, REALH, ADJ>(stage, a, el, st); break;
    default: launch_t<R, N, 8, REALH, ADJ>(stage, a, el, st); break;
    }
}

template <typename R, bool REALH, bool ADJ>
void launch_r(int N, int nt, int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el, hipStream_t st)
{
    switch (N) {
    case 1: launch_n<R, 1, REALH, ADJ>(nt, stage, a, el, st); break;
    case 2: launch_n<R, 2, REALH, ADJ>(nt, stage, a, el, st); break;
    case 3: launch_n<R, 3, REALH, ADJ>(nt, stage, a, el, st); break;
    case 4: launch_n<R, 4, REALH, ADJ>(nt, stage, a, el, st); break;
    case 5: launch_n<R, 5, REALH, ADJ>(nt, stage, a, el, st); break;
    case 6: launch_n<R, 6, REALH, ADJ>(nt, stage, a, el, st); break;
    case 7: launch_n<R, 7, REALH, ADJ>(nt, stage, a, el, st); break;
    default: launch_n<R, 8, REALH, ADJ>(nt, stage, a, el, st); break;
    }
}

template <typename R, bool ADJ>
void launch_any(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el,
                hipStream_t s)
{
    if (real_h)
        launch_r<R, true, ADJ>(N, nterm_slots, stage, a, el, s);
    else
        launch_r<R, false, ADJ>(N, nterm_slots, stage, a, el, s);
}

}  // namespace

#if defined(HEOM_STAGE_ADJOINT)
void launch_stage_adjoint(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<double>& a,
                          const ElementArgsT<double>& el, hipStream_t s)
{
    launch_any<double, true>(N, nterm_slots, real_h, stage, a, el, s);
}
#elif !defined(HEOM_STAGE_PRECISION)
int stage_term_slots(int nterms) { return nterms <= 2 ? 2 : nterms <= 4 ? 4 : 8; }

bool stage_supported(int N, int nterms) { return N >= 1 && N <= kMaxStates && nterms >= 0 && nterms <= 8; }

bool tile_split_supported(int N, int split)
{
    return split == 1 || (N >= 5 && N <= kMaxStates && (split == 2 || split == 4) && stage_waves(N) % split == 0);
}

void launch_stage(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<double>& a,
                  const ElementArgsT<double>& el, hipStream_t s)
{
    launch_any<double, false>(N, nterm_slots, real_h, stage, a, el, s);
}
#else
void launch_stage(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<float>& a,
                  const ElementArgsT<float>& el, hipStream_t s)
{
    launch_any<float, false>(N, nterm_slots, real_h, stage, a, el, s);
}
#endif

}  // namespace heom_hip
