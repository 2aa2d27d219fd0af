// stage_wide.hip -- the HEOM RHS with the fused RK4 stage epilogue for large systems,
// 8 < N <= 32 states (dm-heom/cl/heom/heom_ode.cl:24-451 plus rk_weighted_add.cl:19-106),
// complex128 / FP64 and complex64 / FP32.
//
// The stage kernels for N <= 8 put one ADO on each lane and a 64-ADO tile of N^2
// elements in LDS; at N = 24 that tile would need 590 KB.  Here a workgroup owns a
// tile of T consecutive ADOs (a power of two up to 16 such that T + 1 matrices of N^2
// elements fit the 144 KB LDS budget -- the most that fit unless the autotuner chose
// fewer; T divides the 64-ADO instance granularity) and one wavefront per ADO:
//   * the workgroup loads the tile (plane by plane, T consecutive ADOs per plane
//     segment) and the instance's H / hbar into LDS; barrier;
//   * each lane of the ADO's wavefront takes elements e = lane, lane + 64, ... of the
//     ADO: the commutator (i/hbar)(sigma H - H sigma)_ij from LDS (lanes of one pass
//     share rows and read consecutive columns), the "same" term, and the element's
//     neighbour terms (table entry and gathers of the neighbours' element e);
//   * the RK4 epilogue of heom_device.hpp per element.
// The per-ADO constants (pf_same_sum) are wavefront-uniform; the per-element
// constants (gamma_ij, the term list) are per lane.
#include "heom_device.hpp"

namespace heom_hip {

namespace {

constexpr int kWideLdsBytes = 147456;  // 144 KB of the 160 KB LDS
constexpr int kMaxWideStates = 32;

// ADOs per workgroup for N states with `cbytes` per complex element, 0 if none fits
int wide_tile(int N, int cbytes)
{
    const int per = N * N * cbytes;
    for (int T = 16; T >= 1; T /= 2)
        if ((T + 1) * per <= kWideLdsBytes)
            return T;
    return 0;
}

template <typename R, int STAGE, bool REALH>
__global__ __launch_bounds__(1024) void rk4_stage_wide(StageArgsT<R> a, const c2<R>* __restrict__ hs_all,
                                                       const c2<R>* __restrict__ gam,
                                                       const TermT<R>* __restrict__ terms, int N, int nt, int T)
{
    using C = c2<R>;
    __shared__ C lds[kWideLdsBytes / sizeof(C)];
    const int nel = N * N;
    C* hl = lds;          // [nel] H / hbar
    C* sig = lds + nel;   // [T][nel]

    const int64_t mp = a.mp;
    const int64_t q0 = (int64_t)blockIdx.x * T;
    const C* hs = hs_all + (a.tile_inst ? (int64_t)a.tile_inst[q0 / kTile] * nel : 0);
    for (int idx = threadIdx.x; idx < nel; idx += blockDim.x)
        hl[idx] = hs[idx];
    for (int idx = threadIdx.x; idx < T * nel; idx += blockDim.x) {
        const int ad = idx % T, e = idx / T;
        sig[ad * nel + e] = a.yin[(int64_t)e * mp + q0 + ad];
    }
    __syncthreads();

    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t q = q0 + wv;
    const C S = a.pfss[q];
    const C* s = sig + wv * nel;
    for (int e = lane; e < nel; e += 64) {
        const int i = e / N, j = e - i * N;
        C cm = mk2<R>(R(0), R(0));
        for (int k = 0; k < N; ++k) {
            const C sik = s[i * N + k], skj = s[k * N + j];
            const C hkj = hl[k * N + j], hik = hl[i * N + k];
            if (REALH) {
                cm.x = fma(sik.x, hkj.x, fma(-hik.x, skj.x, cm.x));
                cm.y = fma(sik.y, hkj.x, fma(-hik.x, skj.y, cm.y));
            } else {
                cm = cfma(sik, hkj, cm);
                cm = cfma(mk2<R>(-hik.x, -hik.y), skj, cm);
            }
        }
        const C g = gam[e];
        const C se = s[e];
        C f = cfma(mk2<R>(S.x + g.x, S.y + g.y), se, mk2<R>(-cm.y, cm.x));
        const C* plane = a.yin + (int64_t)e * mp;
        for (int k = 0; k < nt; ++k) {
            const TermT<R> t = terms[e * nt + k];
            const int2 nbv = a.nb[(int64_t)t.mode * mp + q];
            const int v = nbv.x & 63;
            const C xp = plane[(uint32_t)nbv.x >> 6];
            const C xm = plane[(uint32_t)nbv.y];
            const R cp = t.plus_c * sqrt((R)(v + 1));
            const R sm = sqrt((R)v);
            f.x = fma(-cp, xp.y, f.x);
            f.y = fma(cp, xp.x, f.y);
            f = cfma(t.minus_c, mk2<R>(sm * xm.x, sm * xm.y), f);
        }
        stage_epilogue<STAGE, R>(a, (int64_t)e * mp + q, se, f);
    }
}

template <typename R, bool REALH>
void launch_w(int N, int nt, int stage, const StageArgsT<R>& a, const ElementArgsT<R>& el, int ados, hipStream_t st)
{
    const int fit = wide_tile(N, (int)sizeof(c2<R>));
    const int T = ados > 0 && ados <= fit && kTile % ados == 0 ? ados : fit;
    const dim3 grid((unsigned)(a.mp / T)), block(64 * T);
    switch (stage) {
    case 0: hipLaunchKernelGGL((rk4_stage_wide<R, 0, REALH>), grid, block, 0, st, a, el.hs, el.gam, el.terms, N, nt, T); break;
    case 1: hipLaunchKernelGGL((rk4_stage_wide<R, 1, REALH>), grid, block, 0, st, a, el.hs, el.gam, el.terms, N, nt, T); break;
    case 2: hipLaunchKernelGGL((rk4_stage_wide<R, 2, REALH>), grid, block, 0, st, a, el.hs, el.gam, el.terms, N, nt, T); break;
    case 3: hipLaunchKernelGGL((rk4_stage_wide<R, 3, REALH>), grid, block, 0, st, a, el.hs, el.gam, el.terms, N, nt, T); break;
    default: hipLaunchKernelGGL((rk4_stage_wide<R, 4, REALH>), grid, block, 0, st, a, el.hs, el.gam, el.terms, N, nt, T); break;
    }
}

}  // namespace

bool wide_supported(int N, int complex_bytes)
{
    return N > kMaxStates && N <= kMaxWideStates && wide_tile(N, complex_bytes) > 0;
}

int wide_max_ados(int N, int complex_bytes) { return wide_tile(N, complex_bytes); }

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
