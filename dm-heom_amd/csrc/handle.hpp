// handle.hpp -- the handle of include/heom_hip.h and the host-side helpers that the
// translation units of the C ABI share (heom_hip.hip: handles, state, RK4 driver;
// adjoint.hip: the adjoint sweep).  Host code only; no kernels.
#ifndef HEOM_HANDLE_HPP
#define HEOM_HANDLE_HPP

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "heom_device.hpp"
#include "heom_hip.h"

namespace heom_rt {

constexpr double kHbar = 1.05457148e-34;  // dm-heom/cl/heom/common.cl:45

// the message heom_last_error returns (defined in heom_hip.hip)
extern thread_local std::string g_error;

struct heom_error : std::runtime_error {
    int code;
    heom_error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void raise(int code, const std::string& msg) { throw heom_error(code, msg); }

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        raise(HEOM_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPC(x) ::heom_rt::hip_check((x), #x)

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
    template <typename U>
    U* as() const
    {
        return reinterpret_cast<U*>(p);
    }
};

// The adjoint (adjoint.hip): the recorded forward sweep (checkpoints), the element
// constants of L^dag, and the buffers of the backward sweep.
struct AdjointState {
    bool recorded = false;  // a forward sweep is on record for heom_adjoint_backward
    double dt = 0.0;
    int64_t nsteps = 0, observe_steps = 1, checkpoint_steps = 1;
    int64_t rows = 0;       // observations, t_r = r * observe_steps * dt
    int64_t last = 0;       // step of the last observation
    int64_t ncp = 0;        // checkpoints, at steps 0, k, 2k, ... <= last
    std::vector<double> op; // [nel * 2] the observable O
    DevBuf<unsigned char> cps;  // [ncp] state planes
    // whole-trajectory mode: [last][4] planes y_n, Y2, Y3, Y4 of every step n < last,
    // kept by the forward sweep when they fit (cps then holds nothing)
    bool keep = false;
    DevBuf<unsigned char> traj;

    bool ops_ready = false;  // the L^dag constants below match the handle's operator
    bool real_h = false;
    DevBuf<double2> hs, gam, pfss;  // -H^dag / hbar, conj(gam), conj(pf_same_sum)
    DevBuf<heom_hip::Term> terms;   // conj(minus_c) and -plus_c per term slot

    DevBuf<unsigned char> seg;                   // [checkpoint_steps] states y_n of a segment
    DevBuf<unsigned char> f2, f3, f4;            // forward stage vectors Y2, Y3, Y4
    DevBuf<unsigned char> lam, a, b, c;          // adjoint state and stage vectors A2, A3, A4
    DevBuf<double2> part;                        // [ntiles][4 stage pairs][slots] per-tile gradient sums
    DevBuf<double2> gs;                          // [4 stage pairs][Mp] per-ADO d J / d pf_same_sum
    DevBuf<double2> src;                         // [rows][count][nel] adjoint sources conj(w_r) O^dag
};

// The bath data of one class of instances: instances with the same states, coupling and
// bath prefactors share one set of element constants (neighbour terms and "same"
// constants) in a batch.
struct BathClass {
    int N = 0, K = 0, B = 0, BM = 0, W = 0;
    std::vector<int32_t> coupling;              // [N * BM], < 0 = none
    std::vector<double> cbk, pf_same, pf_circ;  // [2 B K], [2 B], [2 B]
    bool operator==(const BathClass& o) const
    {
        return N == o.N && K == o.K && B == o.B && BM == o.BM && coupling == o.coupling && cbk == o.cbk &&
               pf_same == o.pf_same && pf_circ == o.pf_circ;
    }
};

// The thermal state search (tss.hip): the constants of the per-ADO transform.
struct TssState {
    DevBuf<double2> c, ct;  // C and C^T [count][nel]
    DevBuf<double2> w;      // (i/hbar)(E_i - E_j) [count][nel]
    DevBuf<double2> epsm;   // eps - S_u per slot [Mp]
    bool real_c = false;    // C has no imaginary parts (real symmetric H)
};

}  // namespace heom_rt

struct heom_handle {
    template <typename T>
    using DevBuf = heom_rt::DevBuf<T>;
    using Term = heom_hip::Term;
    template <typename R>
    using TermT = heom_hip::TermT<R>;

    int device = 0;
    hipStream_t stream = nullptr;
    int N = 0, K = 0, B = 0, BM = 0, W = 0, nel = 0;
    int prec = HEOM_PRECISION_FP64;
    int count = 1;                  // instances in the batch
    std::vector<int64_t> inst_M;    // ADOs per instance
    std::vector<int64_t> inst_u;    // first global uid of each instance
    std::vector<int64_t> inst_q;    // first internal index (a multiple of 64)
    std::vector<int64_t> inst_n;    // internal slots holding ADOs (inst_M, or fewer after heom_compact)
    int64_t M = 0, Mp = 0;          // total ADOs (uid space), internal slots
    int64_t active = 0;             // ADOs with a slot (M, or fewer after heom_compact)
    int ntiles = 0;
    int32_t zero_slot = 0;
    bool real_h = false;
    int term_slots = 2;      // neighbour terms per element the stage kernel runs with
    bool pipelined = false;  // stage_pipelined.hip (persistent) instead of stage_kernel.hip
    int pipe_wgs = 0;        // its workgroups: one per CU, a multiple of the XCD count
    int xcd_chunk = 1;       // tiles per XCD chunk of the stage kernels' tile order
    int wide_ados = 0;       // N > 8: ADOs per workgroup of the wide kernel (0 = the most that fit)
    int tile_split = 1;      // tile kernel: workgroups per 64-ADO tile (1, 2 or 4)
    bool paired = false;     // complex64 planes in the paired layout (stage_pair.hip)
    int pair_slots = 0;      // neighbour terms per plane of the paired layout (0: not possible)

    // tuple-keyed readback (heom_get_ados): each instance's tuples in lexicographic
    // order [M][W] with the uid of each, and the current slot of every uid (-1: removed)
    std::vector<uint8_t> lex_tuples;
    std::vector<int32_t> lex_uid, u2q_host;

    std::vector<int32_t> coupling;  // [N * BM]  (bath class 0)
    std::vector<double> cbk, pf_same, pf_circ;
    // Batches of instances that differ in states, modes, coupling or bath prefactors:
    // every per-instance and per-ADO array uses the largest N and W (smaller instances
    // padded with states that have no bath and stay zero, and modes no ADO uses); each
    // instance belongs to a bath class whose element constants it uses.
    std::vector<heom_rt::BathClass> classes;
    std::vector<int32_t> inst_class;   // [count]
    DevBuf<int32_t> tile_class, inst_class_d;  // [ntiles], [count]; empty with one class
    bool uniform = true;               // one bath class and every instance of N states

    DevBuf<unsigned char> y, pbuf, ya, yb;  // planes, complex128 or complex64
    // the per-evaluation interface (evaluate.hip): state buffers 4.. added by the caller,
    // the planes of f for the accumulating forms, and the pre-evaluate actions
    std::vector<std::unique_ptr<DevBuf<unsigned char>>> extra;
    DevBuf<unsigned char> fscratch;
    std::vector<std::pair<heom_pre_evaluate_fn, void*>> pre_evaluate;
    DevBuf<int2> nb_raw, nb;                // packed neighbour table, unmasked and mask-folded
    DevBuf<int32_t> q2u, tile_inst;
    DevBuf<int32_t> tile_plus;              // [ntiles] 1 if an ADO of the tile has a plus edge
    DevBuf<int64_t> q_top;                  // [count] internal index of each top ADO
    DevBuf<uint8_t> mask;
    DevBuf<double2> pfss, hs, gam, mx, small;  // FP64 constants
    DevBuf<Term> terms;
    DevBuf<double> norms;
    // FP32 storage
    DevBuf<float2> pfss32, hs32, gam32;
    DevBuf<TermT<float>> terms32;
    DevBuf<heom_hip::PairTerm> pterms;  // [classes][pair_planes(N)][pair_slots]
    int S = 0;                    // tier-0/1 ADOs
    std::vector<int32_t> sp_top;  // [count] tier-0/1 index of each top
    std::vector<int64_t> sp_uid;  // [S] global uid
    DevBuf<int64_t> sp_q;
    DevBuf<int32_t> sp_inst, sp_top_d;
    DevBuf<int2> sp_nb;
    DevBuf<double2> y64, sa64, sb64, p64, sp_pfss, sp_tmp;

    // time-dependent Hamiltonian (heom_propagate_timedep): H / hbar of the current stage,
    // [count][nel], in place of hs / hs32 while set
    const double2* hs_stage = nullptr;
    const float2* hs32_stage = nullptr;

    // one RK4 step captured as a HIP graph (heom_use_graphs), replayed per step
    bool use_graphs = false;
    hipGraphExec_t step_graph = nullptr;
    double graph_dt = 0.0;
    bool graph_real_h = false;

    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    bool timing_pending = false;
    int64_t pending_steps = 0;
    heom_timing_info timing{};

    // the recorded forward sweep and buffers of the adjoint (adjoint.hip)
    std::unique_ptr<heom_rt::AdjointState> adj;
    // set on a thermal-state-search handle (heom_create_tss), whose operator is f_TSS
    std::unique_ptr<heom_rt::TssState> tss;

    bool fp32() const { return prec == HEOM_PRECISION_FP32; }
    size_t cbytes() const { return fp32() ? sizeof(float2) : sizeof(double2); }

    heom_handle();
    ~heom_handle();
    void release()
    {
        if (stream)
            (void)hipStreamSynchronize(stream);
        if (step_graph)
            (void)hipGraphExecDestroy(step_graph);
        if (ev_start)
            (void)hipEventDestroy(ev_start);
        if (ev_stop)
            (void)hipEventDestroy(ev_stop);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

namespace heom_rt {

// error boundary of every C ABI function: the status code, with g_error set
int guard(const std::function<void()>& fn);
void need(heom_handle* h);
inline unsigned blocks_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// one stage launch of the FP64 RK4 schedule with explicit buffers: yin the stage input
// (neighbours are read there), y0 = y_n (stages 2, 3), pbuf = P (written by stage 2, read
// by stage 4), yout the stage output; adjoint = true applies L^dag instead of L, and then
// top_src ([count][N*N], stage 4 only) is added to the output at each instance's top ADO
void launch_stage_buffers(heom_handle* h, int stage, const void* yin, const void* y0, void* pbuf, void* yout,
                          double c1, bool reverse, bool adjoint, const double2* top_src = nullptr);
// builds the device layout of count instances into a fresh handle (heom_create_batch)
void create_batch(heom_handle* h, const heom_desc* d, int count);
// raises HEOM_E_UNSUPPORTED for `what` on a thermal-state-search handle
void reject_tss(heom_handle* h, const char* what);
// nsteps RK4 steps of the handle's state (through the captured step graph when enabled)
void rk4_steps(heom_handle* h, double dt, int64_t nsteps);
// the tops of all instances [count][nel] into dst (device)
void gather_tops(heom_handle* h, double2* dst);
// Re sigma_0[i][i] and, with ops, Tr(ops_k sigma_0) of every instance's top into pops
// [count][N] and traces [count] (device pointers)
void observe_launch(heom_handle* h, const double2* ops, double* pops, double2* traces);
// planes in the internal layout -> reference layout (state [M*N*N*2], host)
void download_state(heom_handle* h, const unsigned char* src, const double2* rows, double* state);
// reference layout -> the state planes, or (FP64) the planes dst
void upload_state(heom_handle* h, const double* state, unsigned char* dst = nullptr);
// drops the captured RK4 step (its buffer addresses or coefficients changed)
void drop_graph(heom_handle* h);
// evaluate.hip: state buffer b (0 the state, 1-3 the RK4 work buffers, 4.. added ones)
DevBuf<unsigned char>& state_buffer(heom_handle* h, int b);
// evaluate.hip: the added buffers at the current layout size, zeroed (after a re-layout)
void reset_extra_buffers(heom_handle* h);
int count_bath(const heom_handle* h, int s, int b);
void resolve_timing(heom_handle* h);
// adjoint.hip: drops the recorded sweep when the operator changes
void adjoint_invalidate(heom_handle* h);

}  // namespace heom_rt

#endif  // HEOM_HANDLE_HPP
