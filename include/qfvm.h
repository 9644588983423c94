/*
 * qfvm.h -- C-ABI of the MI355X statevector engine (libqfvm_hip.so).
 *
 * This is the drop-in boundary for pyquafu's qfvm statevector path. Every entry point replaces one
 * piece of the reference's pybind module `qfvm` (src/qfvm/qfvm.cpp:263-288) or of the StateVector
 * engine it drives (src/qfvm/statevector.hpp). Plain C: no C++ or torch types cross it, so ctypes,
 * cffi, cgo or JNI can bind it directly.
 *
 * Conventions
 *   - Qubit q is bit q of the amplitude index (little-endian statevector), as in the reference.
 *   - Host state buffers are interleaved complex128 (re, im doubles), the layout of a numpy
 *     complex128 array. The device state is complex128 or complex64 (QFVM_C128 / QFVM_C64).
 *   - Every function returns a qfvm_status; qfvm_last_error() gives the message of the last
 *     failure on the calling thread. The reference throws C++ exceptions instead
 *     (src/qfvm/simulator.hpp:26, statevector.hpp:108-115); the Python layer maps the codes back
 *     to the reference's exception types.
 *   - A handle owns one device state and one HIP stream. Calls on one handle must not overlap.
 */
#ifndef QFVM_H
#define QFVM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qfvm_state* qfvm_handle;

typedef enum qfvm_status {
    QFVM_OK = 0,
    QFVM_ERR_VALUE = 1,   /* bad argument (ValueError) */
    QFVM_ERR_INDEX = 2,   /* qubit / cbit / size out of range (IndexError) */
    QFVM_ERR_MEMORY = 3,  /* device allocation failed (MemoryError) */
    QFVM_ERR_DEVICE = 4,  /* HIP runtime error or no device (RuntimeError) */
    QFVM_ERR_RUNTIME = 5  /* any other failure (RuntimeError) */
} qfvm_status;

typedef enum qfvm_dtype { QFVM_C64 = 1, QFVM_C128 = 2 } qfvm_dtype;

/*
 * Op codes. The named codes are the reference's native gate set (OPMAP, src/qfvm/qasm.hpp:17-46),
 * dispatched as in apply_op (src/qfvm/simulator.hpp:30-113). Every other gate reaches the engine
 * as GENERAL (one target, 2x2 matrix, any number of controls; apply_op_general,
 * simulator.hpp:12-28) or MULTI (k >= 2 targets, dense 2^k x 2^k matrix whose row-index bit j
 * belongs to the j-th target; statevector.hpp:945-995).
 */
typedef enum qfvm_opcode {
    QFVM_OP_GENERAL = 0,
    QFVM_OP_MULTI = 1,
    QFVM_OP_X = 2,
    QFVM_OP_Y = 3,
    QFVM_OP_Z = 4,
    QFVM_OP_H = 5,
    QFVM_OP_S = 6,
    QFVM_OP_SDG = 7,
    QFVM_OP_T = 8,
    QFVM_OP_TDG = 9,
    QFVM_OP_P = 10,
    QFVM_OP_RX = 11,
    QFVM_OP_RY = 12,
    QFVM_OP_RZ = 13,
    QFVM_OP_CX = 14, /* cx and cnot */
    QFVM_OP_CZ = 15,
    QFVM_OP_CP = 16,
    QFVM_OP_CCX = 17,
    QFVM_OP_RZZ = 18,
    QFVM_OP_MEASURE = 19, /* qubits -> cbits, projective, consumes one uniform */
    QFVM_OP_RESET = 20,   /* qubits -> |0>, consumes one uniform */
    QFVM_OP_CIF = 21,     /* if creg[cbits] (big-endian) == cond: run the next `body` ops */
    QFVM_OP_KRAUS = 22,   /* noise: `body` one-qubit Kraus matrices at m_off; one uniform */
    QFVM_OP_UCHAN = 23,   /* noise: `body` one-qubit unitaries at m_off, then their probabilities
                             (as re parts of `body` complex entries); one uniform */
    QFVM_OP_COUNT = 24
} qfvm_opcode;

/*
 * One instruction of a flat program (64 bytes). Variable-length data lives in the program's pools:
 *   qubits  idx[q_off .. q_off + n_qubits)   controls first, then targets (instructions.hpp:265)
 *   cbits   idx[c_off .. c_off + n_cbits)    MEASURE / CIF only
 *   matrix  mat[m_off .. m_off + 2*4^k)      row-major complex128, GENERAL (k=1) and MULTI only
 */
typedef struct qfvm_op {
    int32_t code;
    int32_t n_ctrl;
    int32_t n_qubits;
    int32_t n_cbits;
    int64_t q_off;
    int64_t c_off;
    int64_t m_off;
    double param; /* angle of P, RX, RY, RZ, CP, RZZ */
    int64_t cond; /* CIF condition */
    int64_t body; /* CIF: number of following ops in its body */
} qfvm_op;

typedef struct qfvm_program {
    const qfvm_op* ops;
    int64_t n_ops;
    const int32_t* idx;
    int64_t n_idx;
    const double* mat;
    int64_t n_mat;
} qfvm_program;

/* Last error message of the calling thread ("" if none). */
const char* qfvm_last_error(void);

/* Library / device information: HIP device count visible to this process. */
int qfvm_device_count(int* out);

/*
 * Allocate a device state of n_qubits (0..34) and initialise it to |0...0>.
 * Replaces StateVector(uint num) (statevector.hpp:148-152). Allocation failure -> QFVM_ERR_MEMORY.
 */
int qfvm_create(int32_t n_qubits, int32_t dtype, qfvm_handle* out);
int qfvm_destroy(qfvm_handle h);
int qfvm_num_qubits(qfvm_handle h, int32_t* out);
int qfvm_dtype_of(qfvm_handle h, int32_t* out);

/* |0...0> (statevector.hpp:149-152) and a computational basis state |index>. */
int qfvm_init_zero(qfvm_handle h);
int qfvm_init_basis(qfvm_handle h, uint64_t index);

/*
 * Copy a host complex128 state in / out (n_amps must be 2^n). load replaces adopting the caller's
 * numpy buffer (StateVector::load_data, statevector.hpp:44-54 / qfvm.cpp:87-89); store is the
 * device-to-host half of returning the state (qfvm.cpp:99-103). A complex64 device state is
 * widened / rounded to nearest on the way.
 */
int qfvm_load(qfvm_handle h, const double* host, uint64_t n_amps);
int qfvm_store(qfvm_handle h, double* host, uint64_t n_amps);

/*
 * Apply a unitary program (no MEASURE / RESET / CIF). Replaces simulate() over the gate list
 * (simulator.hpp:115-125) and applyop_statevec for a one-op program (qfvm.cpp:36-50). Gates may
 * be fused into multi-gate passes; the result equals the gate-by-gate product to rounding.
 */
int qfvm_apply(qfvm_handle h, const qfvm_program* prog);

/*
 * Run a program with mid-circuit semantics: MEASURE collapses (apply_measure,
 * statevector.hpp:1342-1356), RESET projects to |0> (apply_reset, :1358-1362), CIF tests the
 * classical register (check_cif, :186-193). Each MEASURE / RESET consumes the next entry of
 * `uniforms` (values in [0,1)) in place of the reference's std::random_device draw. KRAUS and
 * UCHAN (noise channels, NoiseSVSimulator, simulator.py:148-223) also take one uniform u each:
 * KRAUS applies K_k / |K_k psi| for the first k with u < sum_{j<=k} |K_j psi| / |psi| and leaves
 * the state unchanged when there is none (SVSimulator._apply_op, simulator.py:59-74; the ratio of
 * norms, not of squared norms, is the reference's rule); UCHAN applies unitary k, the first with
 * u < cdf[k], cdf = cumulative probabilities divided by their total (numpy's choice,
 * noise.py:64-68). When
 * skip_measure != 0, MEASURE ops are skipped (final-measure mode, simulator.hpp:119-122).
 * creg has n_creg entries (0/1), zeroed by the caller.
 */
int qfvm_run(qfvm_handle h, const qfvm_program* prog, const double* uniforms, int64_t n_uniforms,
             int32_t* creg, int32_t n_creg, int32_t skip_measure);

/*
 * Pauli-string expectations <psi|P_t|psi>, unscaled by coefficients. Replaces expect_statevec
 * (qfvm.cpp:246-261) -> expect_pauli (statevector.hpp:1250-1305, its formula per term). Term t uses
 * chars[off[t] .. off[t]+len[t]) ('I','X','Y','Z') on qubits pos[off[t] ..]. All terms are taken
 * in few passes over the state (the terms whose X / Y qubits fit one LDS tile share a pass) with
 * one host synchronisation per call, instead of the reference's pass per term.
 */
int qfvm_expect_paulis(qfvm_handle h, const char* chars, const int32_t* pos, const int64_t* off,
                       const int32_t* len, int64_t n_terms, double* out);

/*
 * qfvm_expect_paulis without any allocation. No counterpart in the reference: it is for callers that
 * cannot allocate device or pinned memory (code under HIP graph capture setup, tight serving loops,
 * harnesses that forbid hipMalloc). Same terms, same results (the same kernels, the same summation
 * order) as qfvm_expect_paulis, but every scratch byte lives in a fixed 530 KiB device array that
 * the library holds per GPU until the process ends: at most 256 terms per call (more ->
 * QFVM_ERR_VALUE). Calls on one GPU share that array, so calls that overlap on one GPU (two threads,
 * two handles) race; serialise them. The state must not be modified concurrently either.
 */
int qfvm_expect_paulis_noalloc(qfvm_handle h, const char* chars, const int32_t* pos, const int64_t* off,
                               const int32_t* len, int64_t n_terms, double* out);

/*
 * <psi|H|psi> for H = sum_t coef_t P_t (coef as re, im pairs; terms laid out as in
 * qfvm_expect_paulis, no qubit repeated within a term): the real part of the sum that
 * SVSimulator.run forms from expect_statevec and the coefficients (simulator.py:133-139), taken in
 * one pass over the state.
 */
int qfvm_expect_hamiltonian(qfvm_handle h, const char* chars, const int32_t* pos, const int64_t* off,
                            const int32_t* len, const double* coef, int64_t n_terms, double* out);

/*
 * Reverse-mode gradient of <psi|H|psi>, psi = U(prog)|state> (grad_adjoint, gradiant.py:144-173,
 * arXiv:2009.02823). prog holds unitary ops only. For each g < n_grads, op grad_ops[g] must be a
 * one-target rotation exp(-i theta P / 2) with P = grad_gen[g] ('X', 'Y' or 'Z'; controls allowed);
 * grads[g] = d<H>/d theta = 2 Re <lambda| dU |phi> with dU = -i/2 P U on the control-satisfied
 * subspace (grad_gate, gradiant.py:96-141), lambda = U_{>k}^dagger H psi and phi the state before
 * the op. *expectation (if not null) receives <psi|H|psi>. The sweep keeps a second state on the
 * device; on return the handle holds the state just before the first differentiated op.
 */
int qfvm_adjoint(qfvm_handle h, const qfvm_program* prog, const int64_t* grad_ops, const char* grad_gen,
                 int64_t n_grads, const char* chars, const int32_t* pos, const int64_t* off,
                 const int32_t* len, const double* coef, int64_t n_terms, double* grads,
                 double* expectation);

/*
 * Batched adjoint evaluation. No counterpart in the reference, which evaluates one circuit per call
 * (Estimator.run, quafu/algorithms/estimator.py:186-229; the per-row loop of
 * quafu/algorithms/interface/torch.py:76-79). A batch context keeps, per worker thread, one engine
 * handle per qubit count (each with its own HIP stream), created on first use with the options set
 * through qfvm_batch_set_option and kept between runs, so their plan and graph caches serve later
 * runs of the same circuits with new parameter values.
 */
typedef struct qfvm_batch_ctx* qfvm_batch;

typedef struct qfvm_batch_item {
    const qfvm_program* prog; /* unitary program on n_qubits, evaluated from |0...0> */
    int32_t n_qubits;
    int32_t reserved;
    const int64_t* grad_ops;  /* the differentiated ops and their generators, as qfvm_adjoint */
    const char* grad_gen;
    int64_t n_grads;
    const char* chars;        /* the observable, as qfvm_adjoint */
    const int32_t* pos;
    const int64_t* off;
    const int32_t* len;
    const double* coef;
    int64_t n_terms;
    double* grads;            /* out: n_grads values */
    double* expectation;      /* out: <H> */
    double* seconds;          /* out, may be null: wall time of this item's evaluation */
} qfvm_batch_item;

int qfvm_batch_create(int32_t dtype, int32_t workers, qfvm_batch* out);
int qfvm_batch_destroy(qfvm_batch b);
/*
 * qfvm_set_option on the batch's handles of n_qubits (0: all), those that exist and those created
 * later. One key belongs to the batch itself: "pin_qubits" (0..34, default 20), the qubit count
 * from which items all run on worker 0 one after another (circuits that fill the GPU alone only
 * stretch each other's latency side by side); 0 = no pinning.
 */
int qfvm_batch_set_option(qfvm_batch b, int32_t n_qubits, const char* key, int64_t value);
/*
 * Evaluates every item as qfvm_init_zero + qfvm_adjoint would on a handle of its own. Items of at
 * least pin_qubits qubits go to worker 0, the others largest first (cost 2^n x (ops + 8)) onto the
 * least-loaded worker, so one item list always gets the same assignment (and a circuit evaluated
 * again meets its cached plans and graphs).
 * All workers finish; the first failure's status is returned (its message in qfvm_last_error).
 */
int qfvm_batch_run(qfvm_batch b, const qfvm_batch_item* items, int64_t n_items);

/*
 * Amplitudes at the given indices (out: n re, im pairs, complex128). No counterpart in the
 * reference, where the caller holds the whole state in host memory; here it avoids copying a
 * state of up to 256 GiB to read a few amplitudes.
 */
int qfvm_gather(qfvm_handle h, const uint64_t* indices, int64_t n, double* out);

/*
 * Overlap <a|b> of two states with the same qubit count (dtypes may differ), as in the overlap
 * checks of the reference's circuit tests (tests/quafu/circuits/building_circuit_test.py:47-172,
 * np.vdot of two simulate() states), without copying either state to the host.
 */
int qfvm_inner(qfvm_handle a, qfvm_handle b, double* re, double* im);

/* |amp|^2 for every basis state (probabilities(), statevector.hpp:1023-1032). */
int qfvm_probabilities(qfvm_handle h, double* out, uint64_t n);

/*
 * Marginal probabilities of a set of distinct qubits: out[m] (2^n_qubits entries) is the sum of
 * |amp_i|^2 over the basis states i whose bit qubits[j] equals bit j of m. Replaces the host-side
 * reordering and partial trace of SimuResult.calc_probabilities (results.py:148-176, permutebits /
 * ptrace in default_simulator.py:50-76); the Python layer orders `qubits` to give the reference's
 * layout. Up to 12 qubits are reduced on the device (per-block histograms added in a fixed order);
 * larger sets go through the full probability vector on the host.
 */
int qfvm_marginals(qfvm_handle h, const int32_t* qubits, int32_t n_qubits, double* out);

/*
 * Sample basis-state indices, one per uniform u in [0,1), with measure_samples' own rule
 * (statevector.hpp:1380-1412, inner loop :1390-1396): the first index k whose running sum
 * S_k = S_{k-1} + p_k (one double rounding per add, in index order, p_k = fma(re, re, im*im) as
 * the reference's real(d * conj(d)) compiles under -mfma -ffast-math) satisfies u < S_k; 2^n when
 * none does (measured bits all zero, as in the reference). Bit-exact with that sequential scan for
 * every state and u, but the sums are evaluated once per call, in parallel (exact integer steps per
 * binade of S), and each shot is a binary search: O(2^n + shots log 2^n) instead of O(shots 2^n).
 */
int qfvm_sample(qfvm_handle h, const double* uniforms, int64_t shots, uint64_t* out_index);

/* Squared norm of the state (sum of |amp|^2). */
int qfvm_norm2(qfvm_handle h, double* out);

/* Wait for all work queued on the handle's stream. */
int qfvm_synchronize(qfvm_handle h);

/*
 * Kernel accounting (no counterpart in the reference, which has only TICK/TOCK prints,
 * src/qfvm/util.h:16-23). Every kernel launch is counted with its algorithmic bytes (the bytes it
 * must read and write); with profiling enabled each launch is also bracketed by HIP events on the
 * handle's stream. qfvm_profile_json writes {"kernel": {"launches": L, "ms": T, "bytes": B}, ...}
 * into buf (NUL-terminated, at most cap bytes, *len = full length) and, if reset != 0, clears the
 * counters.
 */
int qfvm_set_profiling(qfvm_handle h, int32_t enable);
int qfvm_profile_json(qfvm_handle h, char* buf, int64_t cap, int64_t* len, int32_t reset);

/*
 * Execution options of one handle (no counterpart in the reference). Keys:
 *   "fusion"      1 = plan gates into fused LDS-tiled passes (default), 0 = one kernel per gate
 *   "tile_bits"   qubits per fused tile b (4..14; 0 = default: 11 for complex128, 12 for complex64;
 *                 the adjoint's two-state sweep uses b - 1 of an explicit value, 11 / 12 by default)
 *   "low_bits"    low qubits always in a tile L (0 = default: 5 for one-state passes, 6 for two-state)
 *   "max_groups"  4-qubit register groups per fused pass (1..24, default 12)
 *   "pass_cost"   compute budget of one fused pass in quarter dense-gate units (0 = default)
 *   "split_rows"  1 = (default) complex128 one-state passes of 2^10..2^13-amplitude tiles and the
 *                 two-state passes of the adjoint (both precisions, 2^10 to 2^12 per state) run
 *                 with two lanes per 16-amplitude register row (twice the wavefronts per tile),
 *                 0 = one lane per row
 *   "plan_cache"  1 = (default) fused-pass plans are kept per handle (up to 64, least recently used
 *                 dropped first) and reused for any later gate list of the same structure -- the
 *                 same gate kinds on the same qubits, only parameter values changed -- with the new
 *                 coefficients; 0 = plan every call (and drop the kept plans)
 *   "rows8"       1 = (default) complex128 one-state passes of 2^10..2^13-amplitude tiles hold one
 *                 8-amplitude row (3-qubit groups) per lane instead of split rows, 0 = split rows
 *   "rows8_c64"   1 = complex64 one-state passes of such tiles do the same (with "rows8" on),
 *                 0 = (default) 16-amplitude rows
 *   "layout"      1 = layout-aware planning of those row-of-8 passes: a pass may store its tile with
 *                 its qubits in other places (the tile's addresses stay the same), so that the qubits
 *                 the next passes act on sit in the low positions every tile holds; fewer passes per
 *                 circuit, and passes without gates at the end of each call restore the qubit order,
 *                 so the state between calls is always in the standard order. 0 = (default) off
 *   "stream"      1 = persistent fused-pass kernel that prefetches the next tile (complex128,
 *                 one-state passes of 2^11..2^13-amplitude tiles, when split_rows does not apply),
 *                 0 = one workgroup per tile
 *   "stagger"     persistent kernel: start delay of half the workgroups, in 8128-cycle sleeps (0..64)
 *   "stagger_shift" the blockIdx bit that selects the delayed half (0..16)
 *   "mfma"        1 = (default) dense gates on 4 to 8 targets run on the matrix cores (FP64 MFMA at
 *                 complex128, FP32 MFMA at complex64), 0 = on the vector ALUs
 *   "dense_blocks" k (0 = off, or 4..8; default 5): qfvm_apply and qfvm_run (not qfvm_adjoint) on
 *                 states of 14 or more qubits with "mfma" on collapse runs of gates confined to at
 *                 most k qubits into one dense gate
 *                 on those qubits (the product of the run, formed on the host), run on the matrix
 *                 cores; a run is collapsed when it acts on 4 or more qubits and holds at least
 *                 "dense_min_gates" non-diagonal gates, and stands at the place of its last gate;
 *                 every other gate runs as before. Pays off for deep runs on few qubits, whose
 *                 gates would otherwise load the fused passes' arithmetic (brickwork circuits
 *                 spread over all qubits never form such runs)
 *   "merge_1q"    1 = (default) qfvm_apply and qfvm_run merge every run of uncontrolled one-qubit
 *                 gates on one qubit (no other gate on that qubit in between) into one gate, the
 *                 run's product, at the place of its last gate; 0 = off
 *   "dense_min_gates" the least number of non-diagonal gates a collapsed run holds (1..2^20,
 *                 default 24: on 28 qubits the collapse wins from about 18 such gates at k = 5 and
 *                 22 at k = 6; a 7- or 8-target gate costs 3 to 6.5 times a 5-target one)
 *   "checkpoints" K (0..64, default 0): qfvm_adjoint keeps up to K copies of the forward state at
 *                 op boundaries spread over the differentiated range and restores psi from them in
 *                 the reverse sweep instead of uncomputing through them; K shrinks to the copies that
 *                 free device memory holds (a reserve of 1/32 of the device memory is kept)
 *   "graphs"      1 = qfvm_adjoint captures its stream work (uploads from pinned slots, kernels,
 *                 the result read-back) into a hipGraph and launches it; executable graphs are
 *                 cached per program structure (up to 64 per handle) and updated in place by later
 *                 calls of the same structure with new angles, matrices or coefficients. Calls with
 *                 dense multi-target gates, calls that must grow a device buffer, and profiling mode
 *                 run eagerly. 0 = (default) eager launches. Results are identical either way.
 * Results do not depend on the options beyond floating-point rounding.
 */
int qfvm_set_option(qfvm_handle h, const char* key, int64_t value);

/*
 * The execution plan qfvm_apply would run for `prog` on an n-qubit state, as JSON (host only, no
 * device needed; no counterpart in the reference). Fused passes list their tile qubits (positions),
 * "out_bit" (layout-aware plans: where each tile bit is stored; empty = in place) and, per group,
 * the ops in execution order: {"kind", "g", "g2", "cmg", "cme", "run", "m": [re, im] x 4, or x 16
 * for a 4x4 on group bits g and g2 (kind 9)}; other items
 * carry the gate on the positions it runs at. Lets the planner's reordering be checked without a
 * GPU. Arguments of 0 pick the defaults, as for qfvm_set_option, except dense_blocks: 0 plans
 * without collapsing runs, 4..8 collapses them as option "dense_blocks" does (at any n), with
 * dense_min_gates (0 = 24); layout = 1 plans as option "layout" does. One-qubit runs are merged
 * as with option "merge_1q" on (its default). Items outside passes carry
 * their matrix ("m": the 2x2, or the 2^k x 2^k of a dense gate, row-major).
 */
int qfvm_plan_json(int32_t n, int32_t dtype, int32_t tile_bits, int32_t low_bits, int32_t max_groups,
                   int32_t pass_cost, int32_t layout, int32_t dense_blocks, int32_t dense_min_gates,
                   const qfvm_program* prog, char* buf, int64_t cap, int64_t* len);

#ifdef __cplusplus
}
#endif

#endif /* QFVM_H */
