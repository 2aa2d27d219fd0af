// heom_device.hpp -- device-side data layout, stage arguments and helpers shared by
// the HIP translation units of libheom_hip.so.
//
// Device layout of one hierarchy (FP64; FP32 storage halves the element size):
//   * ADOs are renumbered into (nearly) lexicographic order of their tuples (mode 0 most
//     significant).  Neighbours n +- e_m then sit close in index: at 735k ADOs four
//     fifths of all neighbour reads land within 30k ADOs, against one in twenty in the
//     reference's layer order, so most of them are served from L2 / MALL.
//   * State planes are structure-of-arrays by matrix element: Y[e * Mp + q] holds
//     element e = i*N + j of ADO q as a double2 (re, im), or a float2.  Mp = M + 1 rounded up to a
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
//   * Within that order the ADOs are grouped into tiles of 64 that either all have or
//     all lack plus neighbours (the deepest tier has none); a stage kernel skips the
//     plus gathers of a tile whose flag says it has none (tile_plus).
//
//   * A batch of hierarchies is one concatenated hierarchy: instance k owns the tiles
//     from its tile offset on (its ADOs padded to a multiple of 64), its tables point
//     into its own range, a per-tile instance index selects its Hamiltonian and a
//     per-tile bath class its element constants.  Instances with fewer states or modes
//     than the largest are padded with states that have no bath (they stay exactly
//     zero) and with modes that no ADO uses.
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
constexpr int kMaxStates = 8;       // N handled by the tile and pipelined stage kernels
constexpr int kMaxTupleValue = 62;  // sqrt(n_m + 1) comes from a 64-entry table
constexpr int64_t kMaxAdos = (int64_t(1) << 26) - 1;  // plus index << 6 fits in 32 bits
constexpr int kNumXcd = 8;          // MI355X: 8 XCDs, workgroups dispatched round-robin

constexpr int kStageColumns = 4;    // matrix columns per wavefront of the stage kernel

// wavefronts per workgroup of the stage kernel: N rows x ceil(N / 4) column blocks
__host__ __device__ constexpr int stage_waves(int N) { return N * ((N + kStageColumns - 1) / kStageColumns); }

// Arithmetic and storage precision of the stage kernels: R = double (complex128
// state) or R = float (complex64 state, FP32 arithmetic; the RK state of the tier-0/1
// ADOs is kept in FP64 by a separate kernel, heom_hip.hip).
template <typename R>
struct Cplx;
template <>
struct Cplx<double> {
    using type = double2;
};
template <>
struct Cplx<float> {
    using type = float2;
};
template <typename R>
using c2 = typename Cplx<R>::type;

template <typename R>
__host__ __device__ __forceinline__ c2<R> mk2(R x, R y)
{
    c2<R> v;
    v.x = x;
    v.y = y;
    return v;
}

// One neighbour term of one matrix element: mode m, and the constants of
//   f_ij += i * plus_c * sqrt(n_m + 1) * sigma^{n+e_m}_ij + minus_c * sqrt(n_m) * sigma^{n-e_m}_ij
template <typename R>
struct TermT {
    c2<R> minus_c;
    R plus_c;
    int32_t mode;
    int32_t info;  // bits 0-7: non-padding terms of this element (same in all its slots);
                   // bit 8: plus_c != 0; bit 9: minus_c != 0
};
using Term = TermT<double>;
constexpr int32_t kTermHasPlus = 1 << 8;
constexpr int32_t kTermHasMinus = 1 << 9;

template <typename R>
struct StageArgsT {
    const c2<R>* __restrict__ yin;   // stage input Y_s (neighbours are read from here)
    const c2<R>* __restrict__ y0;    // y_n
    c2<R>* __restrict__ pbuf;        // P = Y2 + 2 Y3 - y
    c2<R>* __restrict__ yout;        // next stage input, y_{n+1}, or f (stage 0)
    const int2* __restrict__ nb;     // [W+2][Mp] packed neighbour table
    const c2<R>* __restrict__ pfss;  // [Mp] pf_same_sum
    const int32_t* __restrict__ tile_inst;  // [ntiles] instance of each tile, or null (one instance)
    const int32_t* __restrict__ tile_plus;  // [ntiles] 0 if no ADO of the tile has a plus edge, or null
    const int32_t* __restrict__ tile_class; // [ntiles] bath class of each tile (element constants), or null
    // adjoint, fourth stage: [count][N*N] added to the output at each instance's top ADO
    // (slot top_q[k]), or null
    const c2<R>* __restrict__ top_src;
    const int64_t* __restrict__ top_q;
    int64_t mp;
    int32_t ntiles;       // Mp / 64
    int32_t zero_slot;    // index of the zero slot (M)
    int32_t table_units;  // (W + 2) / 2: pairs of table rows 0..W (pipelined kernel)
    int32_t xcd_chunk;    // tiles per XCD chunk (chunk_tile)
    int32_t reverse;      // walk the tiles from the last one (tile t -> ntiles - 1 - t)
    int32_t split;        // tile kernel: workgroups per tile (row blocks; 1, 2 or 4)
    R c1;                 // stage coefficient (h/2, h/2, h, h/6)
};
using StageArgs = StageArgsT<double>;

// Element constants, read with wave-uniform indices; passed as separate restrict
// kernel arguments so that the compiler can use scalar loads for them.
template <typename R>
struct ElementArgsT {
    const c2<R>* hs;        // [instances][N*N] H / hbar
    const c2<R>* gam;       // [classes][N*N] -(Delta_b summed over the baths of i and of j), i != j
    const TermT<R>* terms;  // [classes][N*N][NT] neighbour terms per element (padded)
};
using ElementArgs = ElementArgsT<double>;

// ---- paired complex64 layout ---------------------------------------------------------
//
// With FP32 storage and the pipelined stage kernel (stage_pair.hip), the elements (i, j)
// and (j, i), i < j, share one plane of 16-byte slots (halves 0 and 1), and so do the
// diagonal elements (2a, 2a) and (2a + 1, 2a + 1) (an odd N leaves the last diagonal
// alone, half 1 stays zero).  Both elements of a pair couple to the same bath modes, so
// one 16-byte gather of a neighbour serves both.  A plane holds Mp slots.
__host__ __device__ constexpr int pair_count(int N) { return N * (N - 1) / 2; }
__host__ __device__ constexpr int pair_planes(int N) { return pair_count(N) + (N + 1) / 2; }
// plane (x) and half (y) of element (i, j)
__host__ __device__ __forceinline__ int2 pair_slot(int N, int i, int j)
{
    if (i == j)
        return make_int2(pair_count(N) + (i >> 1), i & 1);
    const int a = i < j ? i : j, b = i < j ? j : i;
    return make_int2(a * (2 * N - a - 1) / 2 + (b - a - 1), i < j ? 0 : 1);
}
// index, in complex64 units, of element e of ADO q: paired ((plane * mp + q) * 2 + half)
// or one plane per element (e * mp + q)
__host__ __device__ __forceinline__ int64_t c64_index(bool paired, int N, int e, int64_t q, int64_t mp)
{
    if (!paired)
        return (int64_t)e * mp + q;
    const int2 s = pair_slot(N, e / N, e % N);
    return ((int64_t)s.x * mp + q) * 2 + s.y;
}

// The neighbour terms of one plane of the paired layout: the union of its two elements'
// modes, each with both elements' constants (zero where an element lacks the mode).
struct PairTerm {
    float2 minus_c[2];
    float plus_c[2];
    int32_t mode;
    int32_t info;
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
__device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 c)
{
    c.x = fmaf(a.x, b.x, c.x);
    c.x = fmaf(-a.y, b.y, c.x);
    c.y = fmaf(a.x, b.y, c.y);
    c.y = fmaf(a.y, b.x, c.y);
    return c;
}

// false if no ADO of `tile` has a plus neighbour: its plus gathers would all read the
// zero slot, and the kernels leave them out (a wave-uniform branch per tile)
template <typename R>
__device__ __forceinline__ bool tile_has_plus(const StageArgsT<R>& a, int tile)
{
    return a.tile_plus == nullptr || a.tile_plus[tile] != 0;
}

// the element constants (terms [classes][nel][nt], gam [classes][nel]) of a tile's bath class
template <typename R>
__device__ __forceinline__ int tile_bath_class(const StageArgsT<R>& a, int tile)
{
    return a.tile_class ? a.tile_class[tile] : 0;
}

// H / hbar of the instance a tile belongs to
template <typename R, int N>
__device__ __forceinline__ const c2<R>* tile_hamiltonian(const c2<R>* hs, const int32_t* tile_inst, int tile)
{
    return tile_inst ? hs + (int64_t)tile_inst[tile] * (N * N) : hs;
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
__device__ __forceinline__ float lane_value(float tab, int idx)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(idx << 2, __float_as_int(tab)));
}

// RK4 stage combination for one element (see the schedule in the file header):
// s = stage input, f = RHS at s, ep = y_n (stages 2, 3) or P (stage 4); r0 goes to the
// stage output, r1 (stage 2 only) to P.
template <int STAGE, typename R>
__device__ __forceinline__ void stage_values(R c1, c2<R> s, c2<R> f, c2<R> ep, c2<R>& r0, c2<R>& r1)
{
    if (STAGE == 0) {
        r0 = f;
    } else if (STAGE == 1) {
        r0 = mk2<R>(fma(c1, f.x, s.x), fma(c1, f.y, s.y));
    } else if (STAGE == 2) {
        const c2<R> y3 = mk2<R>(fma(c1, f.x, ep.x), fma(c1, f.y, ep.y));
        r0 = y3;
        r1 = mk2<R>(fma(R(2), y3.x, s.x) - ep.x, fma(R(2), y3.y, s.y) - ep.y);
    } else if (STAGE == 3) {
        r0 = mk2<R>(fma(c1, f.x, ep.x), fma(c1, f.y, ep.y));
    } else {
        constexpr R third = R(1) / R(3);
        r0 = mk2<R>(fma(ep.x + s.x, third, c1 * f.x), fma(ep.y + s.y, third, c1 * f.y));
    }
}

// Streaming accesses with no reuse inside a stage (the stage operand y_n or P, and the
// stage outputs) carry the non-temporal hint, so that they do not push out of the
// caches the state lines that neighbour reads of other tiles will still ask for.
template <typename R>
using vec2_t = R __attribute__((ext_vector_type(2)));
template <typename R>
__device__ __forceinline__ c2<R> nt_load(const c2<R>* p)
{
    const vec2_t<R> v = __builtin_nontemporal_load((const vec2_t<R>*)p);
    return mk2<R>(v.x, v.y);
}
template <typename R>
__device__ __forceinline__ void nt_store(c2<R> v, c2<R>* p)
{
    vec2_t<R> w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, (vec2_t<R>*)p);
}

// the epilogue operand of a stage: y_n for stages 2 and 3, P for stage 4
template <int STAGE, typename R>
__device__ __forceinline__ c2<R> stage_operand(const StageArgsT<R>& a, int64_t idx)
{
    if (STAGE == 2 || STAGE == 3)
        return nt_load<R>(a.y0 + idx);
    if (STAGE == 4)
        return nt_load<R>(a.pbuf + idx);
    return mk2<R>(R(0), R(0));
}

template <int STAGE, typename R>
__device__ __forceinline__ void stage_store(const StageArgsT<R>& a, int64_t idx, c2<R> r0, c2<R> r1)
{
    nt_store<R>(r0, a.yout + idx);
    if (STAGE == 2)
        nt_store<R>(r1, a.pbuf + idx);
}

// RK4 stage epilogue for one element
template <int STAGE, typename R>
__device__ __forceinline__ void stage_epilogue(const StageArgsT<R>& a, int64_t idx, c2<R> s, c2<R> f)
{
    c2<R> r0, r1 = mk2<R>(R(0), R(0));
    stage_values<STAGE, R>(a.c1, s, f, stage_operand<STAGE, R>(a, idx), r0, r1);
    stage_store<STAGE, R>(a, idx, r0, r1);
}

// ---- launcher (stage_kernel.hip) -------------------------------------------------------

// true if a stage kernel exists for N states and `nterms` neighbour terms per element
bool stage_supported(int N, int nterms);
// the padded term count the kernel for `nterms` expects
int stage_term_slots(int nterms);
// true if the tile kernel for N states can split a tile over `split` workgroups
bool tile_split_supported(int N, int split);
void launch_stage(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<double>& a,
                  const ElementArgsT<double>& el, hipStream_t s);
void launch_stage(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<float>& a,
                  const ElementArgsT<float>& el, hipStream_t s);

// the same stage schedule with the adjoint operator L^dag (stage_kernel_adj.hip); el holds
// the adjoint element constants (-H^dag / hbar, conj(gam), terms with -plus_c and
// conj(minus_c)) and a.pfss conj(pf_same_sum)
void launch_stage_adjoint(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<double>& a,
                          const ElementArgsT<double>& el, hipStream_t s);

// ---- large systems (stage_wide.hip): 8 < N <= 32, one wavefront per ADO ---------------

// W: bath modes (the workgroup keeps its tile's slice of the neighbour table in LDS)
bool wide_supported(int N, int complex_bytes, int W);
// the largest number of ADOs per workgroup (a power of two <= 8) whose tiles fit in LDS
int wide_max_ados(int N, int complex_bytes, int W);
// ados: ADOs per workgroup, 0 = wide_max_ados
void launch_stage_wide(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<double>& a,
                       const ElementArgsT<double>& el, int ados, hipStream_t s);
void launch_stage_wide(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<float>& a,
                       const ElementArgsT<float>& el, int ados, hipStream_t s);

// ---- paired pipelined launcher (stage_pair.hip), complex64 state --------------------

// true if the paired kernel handles N states with W modes and Mp ADO slots
bool pair_supported(int N, int W, int64_t mp);
// el.terms is unused; pterms holds [classes][pair_planes(N)][nterm_slots] plane terms
void launch_stage_pair(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<float>& a,
                       const ElementArgsT<float>& el, const PairTerm* pterms, int workgroups, hipStream_t s);

// ---- pipelined launcher (stage_pipelined.hip) -----------------------------------------

// LDS row stride (element slots per matrix row) of the pipelined kernel's tile: the
// wavefronts' 4-column blocks, padded, so that 2-element loads never cross into a
// slot of another wavefront
__host__ __device__ constexpr int pipe_row_slots(int N) { return (N + kStageColumns - 1) / kStageColumns * kStageColumns; }
// modes of the neighbour table one LDS buffer of the pipelined kernel holds: two state
// tiles (N x row slots x 64 complex) and two table tiles (capacity x 64 int2) in
// 160 KiB, with 4 KiB to spare, an even number, at most 64
__host__ __device__ constexpr int pipe_table_capacity(int N, int complex_bytes)
{
    return (((163840 - 2 * N * pipe_row_slots(N) * 64 * complex_bytes) / 1024 - 4) < 64
                ? ((163840 - 2 * N * pipe_row_slots(N) * 64 * complex_bytes) / 1024 - 4)
                : 64) &
           ~1;
}
// true if the pipelined kernel handles N states with W bath modes and Mp ADO slots
bool pipe_supported(int N, int W, int64_t mp, int complex_bytes);
// grid: `workgroups` persistent workgroups, a multiple of kNumXcd
void launch_stage_pipelined(int N, int nterm_slots, bool real_h, int stage, const StageArgsT<double>& a,
                            const ElementArgsT<double>& el, int workgroups, hipStream_t s);

}  // namespace heom_hip
