// Fused multi-gate passes: host planner data and the launcher of the LDS-tiled pass kernel.
//
// A pass streams the state through LDS once. The state is cut into tiles of 2^b amplitudes: the
// tile's qubits Q are the L lowest qubits (contiguous in memory, for coalesced loads) plus b - L
// chosen higher qubits; the other n - b qubits number the tiles. Inside a tile the gates run in
// "groups": a group names 4 tile qubits; each thread gathers the 16 amplitudes that differ only in
// those 4 bits from LDS into registers, applies the group's gates there, and writes them back.
// Gate targets must lie in the group; controls and diagonal gates may sit on any qubit (in or out
// of the tile), because their effect only depends on index bits each thread knows.
#pragma once

#include <vector>

#include "common.h"

namespace qfvm {

enum FKind : int32_t {
    F_DENSE = 0,  // complex 2x2 on group bit g
    F_REAL = 1,   // real 2x2 on group bit g
    F_DIAG = 2,   // diag(d0, d1) on group bit g
    F_SWAP = 3,   // X on group bit g with a control among the group bits
    F_PHASE = 4,  // diag(d0, d1) on a qubit outside the group: g = that qubit's global position
    F_FLUSH = 5,  // apply the pending relabeling of group bit g to the data (mask in pxmask)
    F_GRAD = 6,   // gradient marker (stage groups only): g = GradGen, see FOpHost
    F_RELABEL = 7, // host only: an X folded into the relabeling of bit g (listed, never uploaded)
    F_DIAGRUN = 8,  // device only: a run of diagonal ops applied at once (see below)
    F_DENSE4 = 9    // dense 4x4 on group bits g (row/column bit 0) and target (bit 1), matrix in mx
                    //   (row-of-8 groups only)
};

// Diagonal runs. Diagonal gates commute, so consecutive diagonal ops of a group (relabelings may
// sit between them) act together as one factor per register member: member M of a row is
// multiplied by the product of f_S over the subsets S of M's bits, where each op contributes
// "terms" (S, lane condition, value). A diagonal on group bit j with group controls C gives
// (S = C, d0) and (S = C + j, d1 / d0); with a pending relabel of bit j the two values are chosen
// per lane by the relabel parity. A diagonal on a qubit outside the group gives (S = C, d0 or d1 by
// that qubit's bit). Controls outside the group make a term conditional on the row's index.
// Terms that are the same for every lane are multiplied into f_S on the host.
constexpr int DIAGRUN_MIN = 8;  // shortest run worth merging

// X gates without a control among the group bits move no data. An X on group bit j whose controls
// outside the group are one qubit c (or none) flips bit j of the row's relabeling exactly where bit
// c of the row's index is set, so the relabeling of bit j after a run of such X gates is the parity
// of (row index & M_j), plus a constant for uncontrolled X. The host folds those X gates into the
// masks M_j (F_RELABEL, not uploaded); register member P then holds logical member P ^ px with px_j
// = parity(G0 & M_j) ^ c_j, applied when the row is written back to LDS (the group's final masks
// are in the pass descriptor). A later gate whose target bit is pending uses swapped coefficients
// where px_j is set (flag FL_TARGET_PX, mask in pxmask, constant in FL_PX_CONST); a gate with a
// control on a pending bit is preceded by F_FLUSH, which moves the data.
enum FFlags : uint32_t { FL_TARGET_PX = 1, FL_PX_CONST = 2 };

constexpr int FUSE_GROUP_BITS = 4;
constexpr int FUSE_MAX_GROUPS = 24;
constexpr int FUSE_MAX_TILE_BITS = 14;

// one gate inside a group (host form; converted to the kernel's precision at upload)
struct FOpHost {
    int32_t kind;
    int32_t g;
    uint32_t cm_group;  // control bits among the 4 group bits
    uint32_t flags;     // FFlags
    uint64_t cm_ext;    // control qubits outside the group (global positions); F_GRAD: all controls
    cd m[4];            // DENSE/REAL: m00 m01 m10 m11; DIAG/PHASE: d0 = m[0], d1 = m[3]
    int32_t target;     // F_GRAD: the generator's qubit (global position)
    int32_t slot;       // F_GRAD: output index
    uint64_t pxmask;    // FL_TARGET_PX / F_FLUSH: relabel parity mask of the op's bit g
    int32_t run;        // 1 + index of the diagonal run the op belongs to in its group (0: none)
    int32_t src;        // index of the gate the op came from in the planned list (-1: none)
    std::vector<cd> mx; // F_DENSE4: the 4x4 matrix, row-major (index bit 0 <-> group bit g)
};

// A group is either a compute group (4 tile qubits and their gates) or, in the two-state reverse
// sweep, a stage: gradient markers evaluated on the tiles of both states between two compute groups.
struct FGroup {
    int S[FUSE_GROUP_BITS] = {-1, -1, -1, -1};  // global qubits of the group (ascending); unused by a stage
    bool stage = false;
    std::vector<FOpHost> ops;
    uint64_t pxm[FUSE_GROUP_BITS] = {0, 0, 0, 0};  // relabel masks at the end of the group
    uint32_t pxc = 0;                              // relabel constants (bit j for group bit j)
};

struct FPass {
    std::vector<int> Q;          // tile qubits, ascending; Q[0..L-1] = 0..L-1
    std::vector<FGroup> groups;
    int group_bits = FUSE_GROUP_BITS;  // qubits per compute group (3: rows of 8, 4: rows of 16)
    // layout-aware plans: the pass stores local bit k of its tile at local bit out_bit[k] (a
    // permutation of 0..b-1; empty = in place). The tile's addresses do not change, only which
    // qubit each of its address bits stands for.
    std::vector<int> out_bit;
};

struct FuseConfig {
    int tile_bits;   // b
    int low_bits;    // L
    int max_groups;
    int max_cost;    // compute budget of a pass in quarter dense-gate units (0: unlimited)
    int group_bits = FUSE_GROUP_BITS;  // qubits per compute group: 3 or 4
    bool layout = false;  // layout-aware planning (row-of-8 one-state passes only, see plan.cpp)
};

// Default tile shape per precision and state count (states = 2: the adjoint's two-state sweep, both
// tiles in LDS). complex128: 2^11-amplitude tiles, which let four workgroups share a CU (faster than
// 2^12 at 24 to 33 qubits by 1.5 to 3 %); complex64: 2^12 (one state: 41% of HBM peak per 30-qubit
// pass against 39% at 2^13, 33-qubit QFT 234 vs 249 ms). Contiguous runs of 32 amplitudes for one
// state and 64 for two.
inline int default_tile_bits(int dtype, int /*states*/) { return dtype == QFVM_C128 ? 11 : 12; }
inline int default_low_bits(int /*dtype*/, int states) { return states == 1 ? 5 : 6; }

// One item of an execution plan: a fused pass or a single gate for the per-gate kernels.
struct PlanItem {
    bool fused;
    FPass pass;
    Gate gate;     // single gates: the gate on the qubit positions of the layout it runs in
    int src = -1;  // single gates: index of the gate in the planned list
    std::vector<int> l2p;  // single gates in layout-aware plans: position of each qubit (empty = identity)
};

std::vector<PlanItem> plan_gates(const std::vector<Gate>& gates, int n, const FuseConfig& cfg);
// Runs of uncontrolled one-qubit gates on one qubit (no other gate on that qubit between them)
// merged into one gate each: diagonal if every factor is, real if every factor is real, else dense;
// it stands at the place of the run's last gate and every other gate keeps its place.
std::vector<Gate> merge_single_qubit_runs(const std::vector<Gate>& gates);
// Runs of gates confined to at most k qubits (4 <= k <= 8), collapsed into one dense k-target gate
// each (a G_MULTI on the run's qubits in ascending order, its matrix the product of the run's gates)
// when the run holds at least min_gates non-diagonal gates and acts on 4 or more qubits; the
// other gates keep their places (each collapsed gate stands at the place of its run's last gate).
std::vector<Gate> collapse_dense_blocks(const std::vector<Gate>& gates, int k, int min_gates);
// The plan of `gates` from the state's plan cache: a plan is reused for any gate list of the same
// structure (kinds, qubits, controls, and the gate classes the planner reads: dense / real /
// diagonal / exchange, zero diagonal entries), with its coefficients re-read from `gates`. Parameter
// values alone never trigger planning again.
const std::vector<PlanItem>& cached_plan(State& s, const std::vector<Gate>& gates, const FuseConfig& cfg);
// Runs a plan: fused passes on the LDS-tiled kernel, single gates on the per-gate kernels. With
// `lam` set, every gate acts on both s.psi and lam (the reverse sweep) and the plan's gradient
// markers write grads_dev[slot].
void execute_plan(State& s, const std::vector<PlanItem>& plan, void* lam = nullptr, double* grads_dev = nullptr);

}  // namespace qfvm
