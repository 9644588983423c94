// Shared definitions of the MI355X statevector engine: error handling, complex arithmetic,
// the per-handle state and the internal gate form every kernel consumes.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <memory>
#include <vector>

#include "qfvm.h"

namespace qfvm {

using cd = std::complex<double>;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define QFVM_HIP(expr)                                                                                  \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::qfvm::Error(e_ == hipErrorOutOfMemory ? QFVM_ERR_MEMORY : QFVM_ERR_DEVICE,          \
                                std::string(#expr) + ": " + hipGetErrorString(e_));                     \
    } while (0)

#define QFVM_LAUNCHED()                                                                                 \
    do {                                                                                                \
        hipError_t e_ = hipGetLastError();                                                              \
        if (e_ != hipSuccess)                                                                           \
            throw ::qfvm::Error(QFVM_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------------------------------------
// device complex helpers (interleaved re, im)
template <typename R>
struct Vec2;
template <>
struct Vec2<double> {
    using T = double2;
};
template <>
struct Vec2<float> {
    using T = float2;
};
template <typename R>
using V2 = typename Vec2<R>::T;

template <typename R>
__host__ __device__ inline V2<R> mk(R x, R y) {
    V2<R> v;
    v.x = x;
    v.y = y;
    return v;
}

// a*b
template <typename R>
__device__ inline V2<R> cmul(V2<R> a, V2<R> b) {
    return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a*b + c*d
template <typename R>
__device__ inline V2<R> cmul2(V2<R> a, V2<R> b, V2<R> c, V2<R> d) {
    return mk<R>(a.x * b.x - a.y * b.y + c.x * d.x - c.y * d.y, a.x * b.y + a.y * b.x + c.x * d.y + c.y * d.x);
}

// ------------------------------------------------------------------------------------------------
// internal gate form (host): one target, a control mask and a 2x2 in one of four shapes, or a
// dense k-target matrix.
enum GateKind : int {
    G_DENSE = 0,  // complex 2x2
    G_REAL = 1,   // real 2x2 (H, RY)
    G_DIAG = 2,   // diag(d0, d1)
    G_SWAP = 3,   // X
    G_MULTI = 4,  // dense 2^k x 2^k on k targets (row bit j <-> targets[j])
    G_GRAD = 5    // gradient marker of the two-state reverse sweep (changes no amplitude)
};

// generator of a gradient marker: the Pauli P of a rotation exp(-i theta P / 2) on `target`
enum GradGen : int { GEN_X = 1, GEN_Y = 2, GEN_Z = 3 };

struct Gate {
    int kind = G_DENSE;
    int target = 0;                // single-target kinds
    uint64_t cmask = 0;            // all control bits must be 1
    std::vector<int> ctrls;        // the controls (for index insertion)
    cd m[4] = {1.0, 0.0, 0.0, 1.0};  // DENSE / REAL: m00 m01 m10 m11; DIAG: m[0], m[1]
    std::vector<int> targets;      // G_MULTI
    std::vector<cd> mk;            // G_MULTI row-major 4^k
    int gen = 0;                   // G_GRAD: GradGen
    int slot = -1;                 // G_GRAD: output index; the marker yields Im <lam| P_ctrl |psi>
};

// per-kernel statistics (algorithmic bytes = what the kernel must read + write)
struct KStat {
    uint64_t launches = 0;
    double ms = 0.0;      // summed HIP-event durations (profiling mode only)
    double bytes = 0.0;   // summed algorithmic bytes
};

// Upload slots: a pinned host area and a device area per purpose. A slot's host area is reused
// once the previous copy out of it has finished (an event), so uploads need no stream sync; the
// device area is only rewritten by later copies on the same stream, after the kernels that read it.
enum StageSlot : int { STAGE_PLAN = 0, STAGE_PLAN2 = 1, STAGE_HAM = 2, STAGE_RESULT = 3, STAGE_COUNT = 4 };
struct Stage {
    unsigned char* host = nullptr;
    size_t host_bytes = 0;
    unsigned char* dev = nullptr;
    size_t dev_bytes = 0;
    hipEvent_t done = nullptr;
    unsigned char* cur_host = nullptr;  // host area of the latest stage_host (`host`, or the capture arena)
};

struct PlanCache;  // plan.cpp

// per-handle state
struct State {
    int n = 0;
    int dtype = QFVM_C128;
    uint64_t size = 1;
    void* psi = nullptr;        // device amplitudes (double2 / float2)
    hipStream_t stream = nullptr;
    int device = 0;
    // scratch (grown on demand, freed with the state)
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    double* host_pinned = nullptr;
    size_t host_pinned_bytes = 0;
    void* ops_buf = nullptr;  // fused-pass op arrays
    size_t ops_bytes = 0;
    void* aux_buf = nullptr;  // Hamiltonian terms and gradient outputs
    size_t aux_bytes = 0;
    void* psi2 = nullptr;     // second state of the reverse sweep (allocated on first use)
    Stage stages[STAGE_COUNT];
    // execution options (qfvm_set_option); tile_bits / low_bits of 0 pick the defaults
    int fusion = 1;
    int tile_bits = 0;
    int low_bits = 0;
    int max_groups = 12;
    int pass_cost = 0;  // fused-pass compute budget (quarter dense-gate units; 0 = unlimited)
    int stream_tiles = 0;   // persistent fused-pass kernel with next-tile prefetch (complex128)
    int split_rows = 1;     // fused-pass kernel with two lanes per 16-amplitude row (complex128)
    int rows8 = 1;          // complex128 one-state passes: one lane per 8-amplitude row, 3-qubit groups
    int rows8_c64 = 0;      // the same for complex64
    int layout = 0;         // layout-aware planning of row-of-8 one-state passes (option "layout")
    int stagger = 0;        // start delay of half the persistent workgroups, in 8128-cycle sleeps
    int stagger_shift = 0;  // which blockIdx bit picks the delayed half
    // adjoint checkpoints (option "checkpoints"): up to this many copies of the forward state,
    // restored into psi during the reverse sweep instead of continuing the uncompute past them
    int checkpoints = 0;
    int mfma = 1;  // dense gates on 4 to 8 targets run on the matrix cores (option "mfma")
    // runs of gates on at most dense_blocks qubits collapsed into one dense gate each, on the
    // matrix cores (options "dense_blocks", 0 = off, and "dense_min_gates"; states of 14 or more
    // qubits with "mfma" on)
    int dense_blocks = 5;
    int dense_min_gates = 24;
    int merge_1q = 1;  // runs of one-qubit gates on one qubit merged first (option "merge_1q")
    std::vector<void*> ckpt;
    // graph mode (option "graphs"): qfvm_adjoint captures its stream work into a hipGraph and
    // replays it through an executable graph cached per program structure (see capi.cpp)
    int graphs = 0;
    bool capturing = false;  // the stream is being captured: allocations and host waits throw NeedsAlloc
    uint64_t buf_gen = 0;    // bumped by every (re)allocation of a buffer that queued work may name
    uint64_t graph_gen = 0;  // buf_gen the cached graphs were captured under
    uint64_t graph_clock = 0;
    unsigned char* cap_host = nullptr;  // pinned capture arena: the host areas of captured uploads
    size_t cap_host_bytes = 0;
    size_t cap_off = 0;
    size_t cap_need = 0;  // arena size an abandoned capture asked for
    uint64_t cap_sig = 0;  // uploads of the running capture (slot and size of each), hashed
    struct GraphEntry {
        uint64_t key;
        uint64_t sig;  // shape of the graph the executable was made from
        uint64_t last_use;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graph_cache;
    // fused-pass plans per gate-list structure (option "plan_cache", default on)
    int plan_cache_on = 1;
    std::shared_ptr<PlanCache> plan_cache;
    // counters: always on; event timing only when `profiling`
    bool profiling = false;
    std::vector<std::pair<std::string, KStat>> kstats;
    struct PendingEvent {
        int stat;
        hipEvent_t a, b;
    };
    std::vector<PendingEvent> pending;

    size_t amp_bytes() const { return dtype == QFVM_C128 ? 16 : 8; }
    void* ensure_scratch(size_t bytes);
    double* ensure_pinned(size_t bytes);
    void* ensure_ops(size_t bytes);
    void* ensure_aux(size_t bytes);
    unsigned char* stage_host(int slot, size_t bytes);  // waits for the slot's previous copy
    unsigned char* stage_dev(int slot, size_t bytes);
    void stage_upload(int slot, size_t bytes);           // host area -> device area, async
    void* ensure_psi2();
    void* ensure_checkpoint(size_t k);  // state-sized buffer k (allocated on first use)
    int stat_index(const char* name);
    void resolve_events();  // fold finished event pairs into kstats (call after a stream sync)
    void before_alloc();    // called before any buffer (re)allocation
    void drop_graphs();
};

// Thrown when queued work would (re)allocate a buffer while the stream is being captured: the
// capture is abandoned and the call runs eagerly, which grows the buffer for later captures.
struct NeedsAlloc {};

// Brackets one kernel launch: counts it and, in profiling mode, records HIP events around it on
// the state's stream.
struct KernelScope {
    State& s;
    int idx;
    hipEvent_t a = nullptr, b = nullptr;
    KernelScope(State& st, const char* name, double bytes);
    ~KernelScope();
};

// Points the state's kernels at another buffer of the same shape for the scope's lifetime.
struct BufferScope {
    State& s;
    void* saved;
    BufferScope(State& st, void* buf) : s(st), saved(st.psi) { s.psi = buf; }
    ~BufferScope() { s.psi = saved; }
};

// Pauli-sum observable in kernel form. (P psi)[i] = i^{#Y} (-1)^popc(j & (y|z)) psi[j], j = i ^ flip,
// so every term is (flip, zy, c * i^{#Y}); terms with equal flip masks share one group and one read.
struct HamTerm {
    uint64_t zy;
    double cre, cim;
};
struct HamGroup {
    uint64_t flip;
    int32_t first, count;
};
struct Hamiltonian {
    std::vector<HamGroup> groups;
    std::vector<HamTerm> terms;
};

// ------------------------------------------------------------------------------------------------
// kernel launchers (defined in the .hip files)
void launch_gate(State& s, const Gate& g);
void launch_fill_basis(State& s, uint64_t index);
void launch_zero(State& s, void* p, size_t bytes);  // bytes: a multiple of 8
// raises a kernel's dynamic LDS limit to 160 KiB, once per (kernel, device); thread-safe
void allow_large_lds(const void* fn, int device);
void launch_h2d_convert(State& s, const double* host, uint64_t n);  // host c128 -> device
void launch_d2h_convert(State& s, double* host, uint64_t n);
double reduce_norm2(State& s);
void reduce_probabilities(State& s, double* host_out);
// Pauli-string expectations (expect_pauli's formula) of up to EXPECT_BATCH terms, queued on the
// stream: term t's value goes to out_dev[terms[t].slot]. Scratch is the caller's: `partials` holds at
// least EXPECT_PARTIALS doubles, `fallback` RED_PARTIALS. With account = false nothing is counted
// either, and the call allocates nothing (no device, pinned or host heap memory).
struct PauliHost {
    uint64_t flip, z;  // X / Y qubits; Z / Y qubits
    int flip_q, y;     // highest X / Y qubit; number of Y letters
    int slot;
};
constexpr int EXPECT_BATCH = 256;
constexpr size_t EXPECT_PARTIALS = 32 * 2048;
constexpr size_t RED_PARTIALS = 2048 + 1;
void queue_expectations(State& s, const PauliHost* terms, int n_terms, double* out_dev, double* partials,
                        double* fallback, bool account = true);
// queue_expectations on the library's fixed device scratch, results copied to host_out (n_terms
// doubles); allocates nothing. n_terms <= EXPECT_BATCH.
void expect_paulis_noalloc(State& s, const PauliHost* terms, int n_terms, double* host_out);
void reduce_marginal(State& s, const std::vector<int>& qbits, std::vector<double>& probs);
void launch_collapse(State& s, const std::vector<int>& qbits, uint64_t meas_state, double scale);
void launch_swap_states(State& s, const std::vector<int>& qbits, uint64_t a, uint64_t b);
// measure_samples' rule: per uniform, the first index whose sequential running sum of |amp|^2 exceeds it
void sample_sequential_rule(State& s, const double* host_u, int64_t shots, uint64_t* host_out);
// Re <psi|H|psi>; when `out` is not null, H psi is written there (a buffer shaped like the state).
// With dev_result, the value is left in that device double instead (no sync; returns 0).
double apply_hamiltonian(State& s, const Hamiltonian& H, void* out, double* dev_result = nullptr);
// grads_dev[g.slot] = Im <lam| P (controls projected) |psi> for one G_GRAD marker, over the whole state
void launch_grad_marker(State& s, const Gate& g, const void* lam, double* grads_dev);
// |K psi|^2 for a one-qubit operator K = [[m0, m1], [m2, m3]] on qubit q (one read of the state)
double reduce_op_norm2(State& s, int q, const cd m[4]);
// amplitudes at host-given indices (out: re, im pairs) and <a|b> of two states of one size
void gather_amplitudes(State& s, const uint64_t* host_idx, uint64_t n, double* host_out);
void reduce_inner(State& a, State& b, double* re, double* im);
// sums column c of a [cols x tiles] partial table into grads_dev[slot_of[c]] (fixed order)
void launch_column_sums(State& s, const double* partials, uint64_t tiles, int cols, const int* slot_of,
                        double* grads_dev);

inline unsigned grid_for(uint64_t work, unsigned block, uint64_t cap = (1ull << 20)) {
    uint64_t g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

}  // namespace qfvm
