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
//     nb[m * Mp + q] = { plus index << 6 | n_m, minus index }, m = 0..W+1, so a term
//     costs one 8-byte load; the extra rows W and W+1 hold the zero slot everywhere:
//     row W serves the padding entries of the per-element term lists, and row W+1 lets
//     the pipelined kernel stage rows in pairs.  Packing limits M to 2^26 - 1
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
    int32_t info;  // bits 0-7: non-padding terms of this element (same in all its slots);
                   // bit 8: plus_c != 0; bit 9: minus_c != 0
};
constexpr int32_t kTermHasPlus = 1 << 8;
constexpr int32_t kTermHasMinus = 1 << 9;

struct StageArgs {
    const double2* __restrict__ yin;   // stage input Y_s (neighbours are read from here)
    const double2* __restrict__ y0;    // y_n
    double2* __restrict__ pbuf;        // P = Y2 + 2 Y3 - y
    double2* __restrict__ yout;        // next stage input, y_{n+1}, or f (stage 0)
    const int2* __restrict__ nb;       // [W+2][Mp] packed neighbour table
    const double2* __restrict__ pfss;  // [Mp] pf_same_sum
    int64_t mp;
    int32_t ntiles;       // Mp / 64
    int32_t zero_slot;    // index of the zero slot (M)
    int32_t table_units;  // (W + 2) / 2: pairs of table rows 0..W (pipelined kernel)
    int32_t xcd_chunk;    // tiles per XCD chunk (chunk_tile)
    int32_t reverse;      // walk the tiles from the last one (tile t -> ntiles - 1 - t)
    double c1;            // stage coefficient (h/2, h/2, h, h/6)
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

// Tile order across XCDs.  Workgroups are dispatched round-robin over the XCDs
// (workgroup b runs on XCD b % 8), and each XCD has its own L2.  The tiles are cut into
// chunks of `chunk` consecutive tiles; chunk k goes to XCD k % 8, so every XCD streams
// one chunk at a time and all XCDs advance through the hierarchy together.  Small
// chunks keep the XCDs' working sets close together (neighbour reads of far modes then
// find data another XCD just streamed in the shared last-level cache); a chunk of
// ntiles / 8 gives each XCD one contiguous range (neighbour reads of near modes stay in
// its own L2).
//   j-th tile of XCD x:
__device__ __forceinline__ int chunk_tile(int x, int j, int chunk)
{
    return (kNumXcd * (j / chunk) + x) * chunk + j % chunk;
}

// workgroup index -> tile index for one workgroup per tile: chunked over the largest
// prefix made of whole rounds of kNumXcd chunks, identity on the rest (a bijection)
__device__ __forceinline__ int xcd_tile(int b, int nt, int chunk)
{
    const int full = nt / (kNumXcd * chunk) * (kNumXcd * chunk);
    return b < full ? chunk_tile(b % kNumXcd, b / kNumXcd, chunk) : b;
}

// value of `tab` held by lane `idx` (0..63) of this wavefront; all lanes must be active
__device__ __forceinline__ double lane_value(double tab, int idx)
{
    const int addr = idx << 2;
    const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(tab));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(tab));
    return __hiloint2double(hi, lo);
}

// RK4 stage combination for one element (see the schedule in the file header):
// s = stage input, f = RHS at s, ep = y_n (stages 2, 3) or P (stage 4); r0 goes to the
// stage output, r1 (stage 2 only) to P.
template <int STAGE>
__device__ __forceinline__ void stage_values(double c1, double2 s, double2 f, double2 ep, double2& r0, double2& r1)
{
    if (STAGE == 0) {
        r0 = f;
    } else if (STAGE == 1) {
        r0 = make_double2(fma(c1, f.x, s.x), fma(c1, f.y, s.y));
    } else if (STAGE == 2) {
        const double2 y3 = make_double2(fma(c1, f.x, ep.x), fma(c1, f.y, ep.y));
        r0 = y3;
        r1 = make_double2(fma(2.0, y3.x, s.x) - ep.x, fma(2.0, y3.y, s.y) - ep.y);
    } else if (STAGE == 3) {
        r0 = make_double2(fma(c1, f.x, ep.x), fma(c1, f.y, ep.y));
    } else {
        constexpr double third = 1.0 / 3.0;
        r0 = make_double2(fma(ep.x + s.x, third, c1 * f.x), fma(ep.y + s.y, third, c1 * f.y));
    }
}

// the epilogue operand of a stage: y_n for stages 2 and 3, P for stage 4
// Streaming accesses with no reuse inside a stage (the stage operand y_n or P, and the
// stage outputs) carry the non-temporal hint, so that they do not push out of the
// caches the state lines that neighbour reads of other tiles will still ask for.
typedef double heom_d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 nt_load(const double2* p)
{
    const heom_d2v v = __builtin_nontemporal_load((const heom_d2v*)p);
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void nt_store(double2 v, double2* p)
{
    heom_d2v w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, (heom_d2v*)p);
}

template <int STAGE>
__device__ __forceinline__ double2 stage_operand(const StageArgs& a, int64_t idx)
{
    if (STAGE == 2 || STAGE == 3)
        return nt_load(a.y0 + idx);
    if (STAGE == 4)
        return nt_load(a.pbuf + idx);
    return make_double2(0.0, 0.0);
}

template <int STAGE>
__device__ __forceinline__ void stage_store(const StageArgs& a, int64_t idx, double2 r0, double2 r1)
{
    nt_store(r0, a.yout + idx);
    if (STAGE == 2)
        nt_store(r1, a.pbuf + idx);
}

// RK4 stage epilogue for one element
template <int STAGE>
__device__ __forceinline__ void stage_epilogue(const StageArgs& a, int64_t idx, double2 s, double2 f)
{
    double2 r0, r1 = make_double2(0.0, 0.0);
    stage_values<STAGE>(a.c1, s, f, stage_operand<STAGE>(a, idx), r0, r1);
    stage_store<STAGE>(a, idx, r0, r1);
}

// ---- launcher (stage_kernel.hip) -------------------------------------------------------

// true if a stage kernel exists for N states and `nterms` neighbour terms per element
bool stage_supported(int N, int nterms);
// the padded term count the kernel for `nterms` expects
int stage_term_slots(int nterms);
void launch_stage(int N, int nterm_slots, bool real_h, int stage, const StageArgs& a, const ElementArgs& el,
                  hipStream_t s);

// ---- pipelined launcher (stage_pipelined.hip) -----------------------------------------

// modes of the neighbour table one LDS buffer of the pipelined kernel holds: two state
// tiles (N*N x 64 complex128) and two table tiles (capacity x 64 int2) in 160 KiB,
// with 4 KiB to spare, an even number, at most 64
__host__ __device__ constexpr int pipe_table_capacity(int N)
{
    return (((163840 - 2048 * N * N) / 1024 - 4) < 64 ? ((163840 - 2048 * N * N) / 1024 - 4) : 64) & ~1;
}
// true if the pipelined kernel handles N states with W bath modes and Mp ADO slots
bool pipe_supported(int N, int W, int64_t mp);
// grid: `workgroups` persistent workgroups, a multiple of kNumXcd
void launch_stage_pipelined(int N, int nterm_slots, bool real_h, int stage, const StageArgs& a,
                            const ElementArgs& el, int workgroups, hipStream_t s);

}  // namespace heom_hip
