// heom_device.hpp -- device-side data layout, stage arguments and helpers shared by
// the HIP translation units of libheom_hip.so.
//
// Device layout of one hierarchy (FP64):
//   * ADOs are renumbered into lexicographic order of their tuples (mode 0 most
//     significant).  Neighbours n +- e_m then sit close in index: at 735k ADOs four
//     fifths of all neighbour reads land within 30k ADOs, against one in twenty in the
//     reference's layer order, so most of them are served from L2 / MALL.
//   * State planes are structure-of-arrays by matrix element: Y[e * Mp + q] holds
//     element e = i*N + j of ADO q as a double2 (re, im).  Mp = M + 1 rounded up to a
//     multiple of 64.  Slot M (and every padding slot) is an ADO with no neighbours
//     that stays exactly zero; edges that do not exist, or whose target is masked,
//     point at slot M ("zero slot"), so gathers need no per-lane branches.
//   * Neighbour tables are mode-major, one int2 per (mode, ADO):
//     nb[m * Mp + q] = { plus index << 6 | n_m, minus index }, m = 0..W, so a term
//     costs one 8-byte load; the extra mode W holds the zero slot everywhere and serves
//     the padding entries of the per-element term lists.  Packing limits M to 2^26 - 1
//     ADOs and n_m to 62.
//
// RK4 schedule (classical RK4, 12 state passes per step instead of 16): with
// Y2 = y + h/2 k1, Y3 = y + h/2 k2, Y4 = y + h k3 and P = Y2 + 2 Y3 - y,
//   y_{n+1} = y + h/6 (k1 + 2 k2 + 2 k3 + k4) = (P + Y4) / 3 + h/6 k4.
//   stage 1: read y             write Y2                   (2 passes)
//   stage 2: read Y2, y         write Y3, P                (4 passes)
//   stage 3: read Y3, y         write Y4 (into Y2's plane) (3 passes)
//   stage 4: read Y4, P         write y_{n+1} (in place)   (3 passes)
// Each stage is one kernel: RHS f(Y_s) with the combination above in its epilogue.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace heom_hip {

constexpr int kTile = 64;           // ADO granularity of Mp and of the whole-ADO kernels
constexpr int kMaxStates = 8;       // N handled by the stage kernel
constexpr int kMaxTupleValue = 62;  // sqrt(n_m + 1) comes from a 64-entry table
constexpr int64_t kMaxAdos = (int64_t(1) << 26) - 1;  // plus index << 6 fits in 32 bits
constexpr int kNumXcd = 8;          // MI355X: 8 XCDs, workgroups dispatched round-robin

constexpr int kStageColumns = 4;    // matrix columns per wavefront of the stage kernel

// wavefronts per workgroup of the stage kernel: N rows x ceil(N / 4) column blocks
__host__ __device__ constexpr int stage_waves(int N) { return N * ((N + kStageColumns - 1) / kStageColumns); }

// One neighbour term of one matrix element: mode m, and the constants of
//   f_ij += i * plus_c * sqrt(n_m + 1) * sigma^{n+e_m}_ij + minus_c * sqrt(n_m) * sigma^{n-e_m}_ij
struct Term {
    double2 minus_c;
    double plus_c;
    int32_t mode;
    int32_t pad;
};

struct StageArgs {
    const double2* __restrict__ yin;   // stage input Y_s (neighbours are read from here)
    const double2* __restrict__ y0;    // y_n
    double2* __restrict__ pbuf;        // P = Y2 + 2 Y3 - y
    double2* __restrict__ yout;        // next stage input, y_{n+1}, or f (stage 0)
    const int2* __restrict__ nb;       // [W+1][Mp] packed neighbour table
    const double2* __restrict__ pfss;  // [Mp] pf_same_sum
    int64_t mp;
    int32_t ntiles;   // workgroups = Mp / 64
    double c1;        // stage coefficient (h/2, h/2, h, h/6)
};

// Element constants, read with wave-uniform indices; passed as separate restrict
// kernel arguments so that the compiler can use scalar loads for them.
struct ElementArgs {
    const double2* hs;   // [N*N] H / hbar
    const double2* gam;  // [N*N] -(Delta_b summed over the baths of i and of j), i != j
    const Term* terms;   // [N*N][NT] neighbour terms per element (padded)
};

// ---- helpers -------------------------------------------------------------------------

__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c)  // c + a*b
{
    c.x = fma(a.x, b.x, c.x);
    c.x = fma(-a.y, b.y, c.x);
    c.y = fma(a.x, b.y, c.y);
    c.y = fma(a.y, b.x, c.y);
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

// value of `tab` held by lane `idx` (0..63) of this wavefront; all lanes must be active
__device__ __forceinline__ double lane_value(double tab, int idx)
{
    const int addr = idx << 2;
    const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(tab));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(tab));
    return __hiloint2double(hi, lo);
}

// RK4 stage epilogue for one element (see the schedule in the file header)
template <int STAGE>
__device__ __forceinline__ void stage_epilogue(const StageArgs& a, int64_t idx, double2 s, double2 f)
{
    if (STAGE == 0) {
        a.yout[idx] = f;
    } else if (STAGE == 1) {
        a.yout[idx] = make_double2(fma(a.c1, f.x, s.x), fma(a.c1, f.y, s.y));
    } else if (STAGE == 2) {
        const double2 y = a.y0[idx];
        const double2 y3 = make_double2(fma(a.c1, f.x, y.x), fma(a.c1, f.y, y.y));
        a.yout[idx] = y3;
        a.pbuf[idx] = make_double2(fma(2.0, y3.x, s.x) - y.x, fma(2.0, y3.y, s.y) - y.y);
    } else if (STAGE == 3) {
        const double2 y = a.y0[idx];
        a.yout[idx] = make_double2(fma(a.c1, f.x, y.x), fma(a.c1, f.y, y.y));
    } else {
        const double2 p = a.pbuf[idx];
        constexpr double third = 1.0 / 3.0;
        a.yout[idx] = make_double2(fma(p.x + s.x, third, a.c1 * f.x), fma(p.y + s.y, third, a.c1 * f.y));
    }
}

// ---- launcher (stage_kernel.hip) -------------------------------------------------------

// true if a stage kernel exists for N states and `nterms` neighbour terms per element
bool stage_supported(int N, int nterms);
// the padded term count the kernel for `nterms` expects
int stage_term_slots(int nterms);
void launch_stage(int N, int nterm_slots, bool real_h, int stage, const StageArgs& a, const ElementArgs& el,
                  hipStream_t s);

}  // namespace heom_hip
