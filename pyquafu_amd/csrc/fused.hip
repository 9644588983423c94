// LDS-tiled fused pass kernel (see fused.h for the tiling).
//
// One workgroup (2^(b-4) threads, at most 512: one 16-amplitude row each) owns one tile of 2^b
// amplitudes:
//   1. load the tile from HBM into LDS (lanes walk the L low qubits: runs of 2^L contiguous
//      amplitudes, coalesced);
//   2. for each group: every thread gathers the 16 amplitudes of one "row" (indices that differ only
//      in the group's 4 bits) into registers, applies the group's gates with compile-time register
//      indices, and scatters them back; a barrier separates groups;
//   3. store the tile back.
// Algorithmic bytes per launch: 2 * s * 2^n (each amplitude read and written once), s = 16 / 8.
// LDS addresses are XOR-swizzled (bits 0-2 ^= bits 4-6) so that a row's 16-amplitude stride and the
// load loop's unit stride both spread across banks.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "fused.h"
#include "sincos.h"

namespace qfvm {

template <typename R>
struct FOpDev {
    int32_t kind;
    int32_t g;
    uint32_t cm_group;
    uint32_t flags;
    uint64_t cm_ext;
    uint64_t pxmask;  // relabel parity mask of bit g (FL_TARGET_PX, F_FLUSH)
    R m[8];
};

struct PassDesc {
    int n, b, L;
    int q[FUSE_MAX_TILE_BITS];  // global qubit of local bit k (ascending; q[k] = k for k < L)
    int ngroups;
    uint32_t stage_mask;        // bit gi: group gi is a gradient stage (two-state passes)
    int gbits[FUSE_MAX_GROUPS][FUSE_GROUP_BITS];  // local bits of each group, ascending
    int gop[FUSE_MAX_GROUPS + 1];                  // op ranges in the op array
    uint64_t gpxm[FUSE_MAX_GROUPS][FUSE_GROUP_BITS];  // relabel masks at each group's end
    uint32_t gpxc[FUSE_MAX_GROUPS];                   // relabel constants at each group's end
    uint64_t top[8];     // global offsets of local index j << (b - 3): the tile's 3 highest local bits
    uint64_t top16[16];  // the same for j << (b - 4): the 4 highest
    // row-of-8 passes: byte offsets swz(member m's group bits) x amplitude bytes, the uniform
    // relabel offset of groups without lane parities, and per-group flags (bit gi): the group needs
    // the lane's global index (lane-dependent ops or relabel parities) / its relabeling is per lane
    int32_t goff[FUSE_MAX_GROUPS][8];
    int32_t gpxoff[FUSE_MAX_GROUPS];
    uint32_t g0_mask, pxlane_mask;
    // layout-aware row-of-8 passes (FPass::out_bit): output element e of the tile takes input
    // element Pinv(e) (Pinv the inverse bit permutation). Thread t stores output elements
    // rho(t) + j * NT, rho(t) = XOR of rbit[k] over the bits k of t (a reordering of the wavefront's
    // 64 elements per store, chosen so the LDS reads avoid bank conflicts), read from slot
    // swz(XOR of pbit[k] = Pinv(rbit[k])) ^ pst[j]; permuted = 0 stores every element in place
    uint32_t permuted;
    uint32_t has_d4;  // the pass holds 4x4 ops (host side: picks the kernel instance)
    int32_t pbit[FUSE_MAX_TILE_BITS];
    int32_t rbit[FUSE_MAX_TILE_BITS];
    int32_t pst[8];
};

// Diagonal-run record (fused.h): 16 host factors f0, then for each subset S its term range
// [tbeg[S], tbeg[S+1]) in the term array that follows the header.
template <typename R>
struct RunTerm {
    uint64_t cl;    // index bits outside the group that must all be set
    uint64_t pm;    // selector: v1 where parity(G0 & pm) ^ pc is 1, else v0
    uint32_t pc;
    int32_t angle;  // 1: unit-modulus values given as angles (v[0] = arg v0, v[1] = arg v1);
                    // 2: linear, the angle v[0] * (G0 & cl) (cl a set of index bits, no condition)
    R v[4];         // v0 (re, im), v1 (re, im)
};
template <typename R>
struct RunHead {
    R f0[32];
    int32_t tbeg[17];
    int32_t pad[3];
};

// bit j of the row's relabeling: parity(G0 & mask) ^ constant
__device__ __forceinline__ int px_bit(uint64_t G0, uint64_t mask, uint32_t c) {
    return (int)((__popcll(G0 & mask) ^ c) & 1u);
}

__device__ __forceinline__ int swz(int l) { return l ^ ((l >> 4) & 7); }

// c ? a : b on the two components (a ternary on the struct itself is lowered through memory)
template <typename R>
__device__ __forceinline__ V2<R> csel(bool c, V2<R> a, V2<R> b) {
    return mk<R>(c ? a.x : b.x, c ? a.y : b.y);
}

__device__ __forceinline__ int ins_bit(int x, int p) {
    const int low = (1 << p) - 1;
    return (x & low) | ((x & ~low) << 1);
}

// Register indices must be compile-time constants (a runtime index sends the array to scratch), so
// the 16 members of a row are walked with folds over integer sequences, never with loops. Every
// branch below is uniform across the wavefront (gate kind, target bit and control masks are the
// same for all lanes); per-lane facts (controls outside the group, the relabeling px) only select
// coefficients.
template <typename R, int KIND, int G, int M, bool CTRL>
__device__ __forceinline__ void pair_step(V2<R> (&v)[16], int cmg, V2<R> m00, V2<R> m01, V2<R> m10, V2<R> m11) {
    if constexpr ((M >> G) & 1) {
        return;
    } else {
        if constexpr (CTRL) {
            if ((M & cmg) != cmg) return;
        }
        constexpr int M1 = M | (1 << G);
        const V2<R> a = v[M], b = v[M1];
        if constexpr (KIND == F_SWAP) {
            v[M] = b;
            v[M1] = a;
        } else if constexpr (KIND == F_DIAG) {
            v[M] = cmul<R>(m00, a);
            v[M1] = cmul<R>(m11, b);
        } else if constexpr (KIND == F_REAL) {
            v[M] = mk<R>(m00.x * a.x + m01.x * b.x, m00.x * a.y + m01.x * b.y);
            v[M1] = mk<R>(m10.x * a.x + m11.x * b.x, m10.x * a.y + m11.x * b.y);
        } else {
            v[M] = cmul2<R>(m00, a, m01, b);
            v[M1] = cmul2<R>(m10, a, m11, b);
        }
    }
}

template <typename R, int KIND, int G, int... Ms>
__device__ __forceinline__ void pairs_seq(V2<R> (&v)[16], int cmg, V2<R> m00, V2<R> m01, V2<R> m10, V2<R> m11,
                                          std::integer_sequence<int, Ms...>) {
    if (cmg == 0)  // no control among the group bits: no per-pair tests
        (pair_step<R, KIND, G, Ms, false>(v, cmg, m00, m01, m10, m11), ...);
    else
        (pair_step<R, KIND, G, Ms, true>(v, cmg, m00, m01, m10, m11), ...);
}

// 2x2 kinds (dense, real, diagonal) and controlled X with group controls and outside controls.
template <typename R, int KIND, int G>
__device__ __forceinline__ void op_2x2(V2<R> (&v)[16], const FOpDev<R>& o, bool ok, uint64_t G0) {
    V2<R> m00 = mk<R>(o.m[0], o.m[1]), m01 = mk<R>(o.m[2], o.m[3]);
    V2<R> m10 = mk<R>(o.m[4], o.m[5]), m11 = mk<R>(o.m[6], o.m[7]);
    if (o.flags & FL_TARGET_PX) {
        // register pair (Q0, Q1) holds logical (L1, L0) where bit G of the row's relabeling is
        // set: the matrix acts as [[m11, m10], [m01, m00]] on the registers
        const bool sw = px_bit(G0, o.pxmask, (o.flags >> 1) & 1u);
        const V2<R> t00 = m00, t01 = m01, t10 = m10, t11 = m11;
        m00 = csel<R>(sw, t11, t00);
        m01 = csel<R>(sw, t10, t01);
        m10 = csel<R>(sw, t01, t10);
        m11 = csel<R>(sw, t00, t11);
    }
    if (o.cm_ext != 0) {  // lanes whose outside controls fail apply the identity
        const V2<R> one = mk<R>(1, 0), zero = mk<R>(0, 0);
        m00 = csel<R>(ok, m00, one);
        m01 = csel<R>(ok, m01, zero);
        m10 = csel<R>(ok, m10, zero);
        m11 = csel<R>(ok, m11, one);
    }
    pairs_seq<R, KIND, G>(v, (int)o.cm_group, m00, m01, m10, m11, std::make_integer_sequence<int, 16>{});
}

// Moves the data of pending relabel bit G: pairs swap where the row's relabel bit G is set (bitwise
// blend, so no select of registers is formed).
template <typename R, int G, int M>
__device__ __forceinline__ void flush_step(V2<R> (&v)[16], uint64_t msk) {
    if constexpr (!((M >> G) & 1)) {
        constexpr int M1 = M | (1 << G);
        static_assert(sizeof(V2<R>) % 8 == 0 || sizeof(V2<R>) == 8, "");
        if constexpr (sizeof(R) == 8) {
            const uint64_t ax = __builtin_bit_cast(uint64_t, v[M].x), ay = __builtin_bit_cast(uint64_t, v[M].y);
            const uint64_t bx = __builtin_bit_cast(uint64_t, v[M1].x), by = __builtin_bit_cast(uint64_t, v[M1].y);
            const uint64_t tx = (ax ^ bx) & msk, ty = (ay ^ by) & msk;
            v[M].x = __builtin_bit_cast(R, ax ^ tx);
            v[M].y = __builtin_bit_cast(R, ay ^ ty);
            v[M1].x = __builtin_bit_cast(R, bx ^ tx);
            v[M1].y = __builtin_bit_cast(R, by ^ ty);
        } else {
            const uint32_t m32 = (uint32_t)msk;
            const uint32_t ax = __builtin_bit_cast(uint32_t, v[M].x), ay = __builtin_bit_cast(uint32_t, v[M].y);
            const uint32_t bx = __builtin_bit_cast(uint32_t, v[M1].x), by = __builtin_bit_cast(uint32_t, v[M1].y);
            const uint32_t tx = (ax ^ bx) & m32, ty = (ay ^ by) & m32;
            v[M].x = __builtin_bit_cast(R, ax ^ tx);
            v[M].y = __builtin_bit_cast(R, ay ^ ty);
            v[M1].x = __builtin_bit_cast(R, bx ^ tx);
            v[M1].y = __builtin_bit_cast(R, by ^ ty);
        }
    }
}

template <typename R, int G, int... Ms>
__device__ __forceinline__ void flush_bit(V2<R> (&v)[16], int sw, std::integer_sequence<int, Ms...>) {
    const uint64_t msk = 0ull - (uint64_t)sw;
    (flush_step<R, G, Ms>(v, msk), ...);
}

template <typename R, int M>
__device__ __forceinline__ void phase_step(V2<R> (&v)[16], int cmg, V2<R> d) {
    if ((M & cmg) == cmg) v[M] = cmul<R>(d, v[M]);
}

template <typename R, int... Ms>
__device__ __forceinline__ void phase_seq(V2<R> (&v)[16], int cmg, V2<R> d, std::integer_sequence<int, Ms...>) {
    (phase_step<R, Ms>(v, cmg, d), ...);
}

// LDS slot of every member of a row, computed once per group and used by both gather and scatter.
template <int... Ms>
__device__ __forceinline__ void row_slots(int (&a)[16], int l0, int s0, int s1, int s2, int s3,
                                          std::integer_sequence<int, Ms...>) {
    ((a[Ms] = swz(l0 | ((Ms & 1) << s0) | (((Ms >> 1) & 1) << s1) | (((Ms >> 2) & 1) << s2) |
                  (((Ms >> 3) & 1) << s3))),
     ...);
}

template <typename R, int... Ms>
__device__ __forceinline__ void gather_seq(V2<R> (&v)[16], const V2<R>* tile, const int (&a)[16],
                                           std::integer_sequence<int, Ms...>) {
    ((v[Ms] = tile[a[Ms]]), ...);
}

// register member P goes back to the slot of logical member P ^ px; the swizzle is linear over
// XOR, so that slot is a[P] ^ swz(offset of px)
template <typename R, int... Ms>
__device__ __forceinline__ void scatter_seq(const V2<R> (&v)[16], V2<R>* tile, const int (&a)[16], int xo,
                                            std::integer_sequence<int, Ms...>) {
    ((tile[a[Ms] ^ xo] = v[Ms]), ...);
}

// global offset of local index bits (local bit k -> qubit q[k]); identity below L
__device__ __forceinline__ uint64_t local_to_global(int l, const PassDesc& pd) {
    uint64_t g = (uint64_t)(l & ((1 << pd.L) - 1));
    for (int k = pd.L; k < pd.b; k++)
        if ((l >> k) & 1) g |= 1ull << pd.q[k];
    return g;
}

// the same for an index below 2^(b-3) (a thread index of the split-row kernel): the loop stops at
// bit b - 4
__device__ __forceinline__ uint64_t local_to_global_low(int l, const PassDesc& pd) {
    uint64_t g = (uint64_t)(l & ((1 << pd.L) - 1));
    for (int k = pd.L; k < pd.b - 3; k++)
        if ((l >> k) & 1) g |= 1ull << pd.q[k];
    return g;
}

template <typename R, int KIND>
__device__ __forceinline__ void op_kind(V2<R> (&v)[16], const FOpDev<R>& o, bool ok, uint64_t G0) {
    switch (o.g) {
        case 0: op_2x2<R, KIND, 0>(v, o, ok, G0); break;
        case 1: op_2x2<R, KIND, 1>(v, o, ok, G0); break;
        case 2: op_2x2<R, KIND, 2>(v, o, ok, G0); break;
        default: op_2x2<R, KIND, 3>(v, o, ok, G0); break;
    }
}

template <typename R, int S, int M>
__device__ __forceinline__ void run_mul(V2<R> (&v)[16], V2<R> f) {
    if constexpr ((M & S) == S) v[M] = cmul<R>(f, v[M]);
}

// factor f_S of a diagonal run (host part times this lane's terms), applied to the members M ⊇ S
template <typename R, int S, int... Ms>
__device__ __forceinline__ void run_subset(V2<R> (&v)[16], const RunHead<R>* __restrict__ rh,
                                           const RunTerm<R>* __restrict__ terms, uint32_t active, uint32_t angles,
                                           uint64_t G0, std::integer_sequence<int, Ms...>) {
    if (!((active >> S) & 1u)) return;
    V2<R> f = mk<R>(rh->f0[2 * S], rh->f0[2 * S + 1]);
    R th = 0;  // unit-modulus terms add angles; one sincos per subset at the end
    const int t1 = rh->tbeg[S + 1];
    for (int t = rh->tbeg[S]; t < t1; t++) {
        const RunTerm<R>& tm = terms[t];
        if (tm.angle == 2) {
            th += tm.v[0] * (R)(G0 & tm.cl);
            continue;
        }
        const bool cond = (G0 & tm.cl) == tm.cl;
        const bool sel = tm.pm ? (bool)px_bit(G0, tm.pm, tm.pc) : (bool)(tm.pc & 1u);
        if (tm.angle) {
            th += cond ? (sel ? tm.v[1] : tm.v[0]) : (R)0;
        } else {
            V2<R> val = csel<R>(sel, mk<R>(tm.v[2], tm.v[3]), mk<R>(tm.v[0], tm.v[1]));
            val = csel<R>(cond, val, mk<R>(1, 0));
            f = cmul<R>(f, val);
        }
    }
    if ((angles >> S) & 1u) {
        R sn, cs;
        sin_cos(th, &sn, &cs);
        f = cmul<R>(f, mk<R>(cs, sn));
    }
    (run_mul<R, S, Ms>(v, f), ...);
}

template <typename R, int... Ss>
__device__ __forceinline__ void diag_run(V2<R> (&v)[16], const unsigned char* __restrict__ runs, const FOpDev<R>& o,
                                         uint64_t G0, std::integer_sequence<int, Ss...>) {
    const RunHead<R>* rh = reinterpret_cast<const RunHead<R>*>(runs + o.cm_ext);
    const RunTerm<R>* terms = reinterpret_cast<const RunTerm<R>*>(rh + 1);
    (run_subset<R, Ss>(v, rh, terms, (uint32_t)o.g, o.flags, G0, std::make_integer_sequence<int, 16>{}), ...);
}

template <typename R>
__device__ __forceinline__ void apply_op(V2<R> (&v)[16], const FOpDev<R>& o, uint64_t G0,
                                         const unsigned char* __restrict__ runs) {
    const bool ok = (G0 & o.cm_ext) == o.cm_ext;
    switch (o.kind) {
        case F_DENSE: op_kind<R, F_DENSE>(v, o, ok, G0); break;
        case F_REAL: op_kind<R, F_REAL>(v, o, ok, G0); break;
        case F_DIAG: op_kind<R, F_DIAG>(v, o, ok, G0); break;
        case F_SWAP:  // controls among the group bits (X without them is a relabeling, folded on the host)
            if (o.cm_ext == 0)
                op_kind<R, F_SWAP>(v, o, ok, G0);  // uniform register moves
            else
                op_kind<R, F_DENSE>(v, o, ok, G0);  // X as a matrix, identity where outside controls fail
            break;
        case F_FLUSH: {
            const int sw = px_bit(G0, o.pxmask, (o.flags >> 1) & 1u);
            switch (o.g) {
                case 0: flush_bit<R, 0>(v, sw, std::make_integer_sequence<int, 16>{}); break;
                case 1: flush_bit<R, 1>(v, sw, std::make_integer_sequence<int, 16>{}); break;
                case 2: flush_bit<R, 2>(v, sw, std::make_integer_sequence<int, 16>{}); break;
                default: flush_bit<R, 3>(v, sw, std::make_integer_sequence<int, 16>{}); break;
            }
            break;
        }
        case F_DIAGRUN: diag_run<R>(v, runs, o, G0, std::make_integer_sequence<int, 16>{}); break;
        default: {  // F_PHASE: one factor for the whole row, chosen by an index bit outside the group
            V2<R> d = csel<R>((G0 >> o.g) & 1ull, mk<R>(o.m[6], o.m[7]), mk<R>(o.m[0], o.m[1]));
            d = csel<R>(ok, d, mk<R>(1, 0));
            phase_seq<R>(v, (int)o.cm_group, d, std::make_integer_sequence<int, 16>{});
        }
    }
}

__device__ __forceinline__ double wave_total(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Gradient stage of a two-state pass: for each marker (kind F_GRAD; g = generator, cm_group = the
// generator's local tile bit for X / Y or its global qubit for Z, cm_ext = control mask, flags =
// column) the tile's share of Im(conj(lam) P psi) over control-satisfied amplitudes goes to
// partials[column * tiles + tile]. Markers are taken 8 at a time; each thread walks the elements it
// loads (l = lh + tid), then the wavefronts' sums are added in a fixed order.
template <typename R>
__device__ void grad_stage(const V2<R>* phi, const V2<R>* lam, int T, uint64_t g_tid, const PassDesc& pd,
                           const FOpDev<R>* __restrict__ ops, int cnt, double* red, double* __restrict__ partials,
                           uint64_t tiles) {
    const int tid = threadIdx.x, NT = blockDim.x, lane = tid & 63, wave = tid >> 6;
    for (int c0 = 0; c0 < cnt; c0 += 8) {
        const int cn = min(8, cnt - c0);
        double acc[8];
#pragma unroll
        for (int k = 0; k < 8; k++) acc[k] = 0.0;
        for (int lh = 0; lh < T; lh += NT) {
            const int e = lh + tid;
            if (e >= T) break;
            const uint64_t gi = g_tid | local_to_global(lh, pd);
            const V2<R> p = phi[swz(e)], l = lam[swz(e)];
            const double w = (double)l.x * p.y - (double)l.y * p.x;  // Im(conj(l) p)
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (k < cn) {
                    const FOpDev<R>& o = ops[c0 + k];
                    double v;
                    if (o.g == GEN_Z) {
                        v = ((gi >> o.cm_group) & 1ull) ? -w : w;
                    } else {
                        const int lb = (int)o.cm_group;
                        const V2<R> q = phi[swz(e ^ (1 << lb))];
                        if (o.g == GEN_X) {
                            v = (double)l.x * q.y - (double)l.y * q.x;
                        } else {  // Y: (Y psi)[e] = +i q where bit lb of e is set, -i q where clear
                            const double re = (double)l.x * q.x + (double)l.y * q.y;
                            v = ((e >> lb) & 1) ? re : -re;
                        }
                    }
                    acc[k] += ((gi & o.cm_ext) == o.cm_ext) ? v : 0.0;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const double t = wave_total(acc[k]);
            if (k < cn && lane == 0) red[k * 8 + wave] = t;
        }
        __syncthreads();
        if (tid < cn) {
            double sum = 0.0;
            for (int w = 0; w < (NT >> 6); w++) sum += red[tid * 8 + w];
            partials[(uint64_t)ops[c0 + tid].flags * tiles + blockIdx.x] = sum;
        }
        __syncthreads();
    }
}

// Sum over the wavefront of 8 values at once: a transposing butterfly (each exchange halves the
// values a lane keeps) takes 10 double shuffles instead of 8 x 6. Returns, in every lane, the
// total of value m = 4 * lane bit 5 + 2 * lane bit 4 + lane bit 3.
__device__ __forceinline__ double wave_total8(const double (&a)[8], int lane) {
    const bool b5 = (lane >> 5) & 1, b4 = (lane >> 4) & 1, b3 = (lane >> 3) & 1;
    double t4[4], t2[2];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double keep = b5 ? a[k + 4] : a[k], give = b5 ? a[k] : a[k + 4];
        t4[k] = keep + __shfl_xor(give, 32, 64);
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const double keep = b4 ? t4[k + 2] : t4[k], give = b4 ? t4[k] : t4[k + 2];
        t2[k] = keep + __shfl_xor(give, 16, 64);
    }
    double t = (b3 ? t2[1] : t2[0]) + __shfl_xor(b3 ? t2[0] : t2[1], 8, 64);
    t += __shfl_xor(t, 4, 64);
    t += __shfl_xor(t, 2, 64);
    t += __shfl_xor(t, 1, 64);
    return t;
}

// grad_stage for the split-row kernel, whose threads hold exactly 8 elements each (NT = T / 8):
// element j of thread tid is j * NT + tid, its global index g_tid | pd.top[j] (no per-element
// index walk). Im(conj(lam) phi) of the elements is computed once per stage and kept in registers
// across the markers (Z markers read nothing more), and the wavefront sums of 8 markers take one
// transposing butterfly. Same partial sums as grad_stage: the same per-thread order over
// elements, the same exchange partners over the wavefront, the same order over wavefronts.
template <typename R>
__device__ void grad_stage8(const V2<R>* phi, const V2<R>* lam, uint64_t g_tid, const PassDesc& pd,
                            const FOpDev<R>* __restrict__ ops, int cnt, double* red, double* __restrict__ partials,
                            uint64_t tiles) {
    const int tid = threadIdx.x, NT = blockDim.x, lane = tid & 63, wave = tid >> 6;
    double w[8];  // Im(conj(lam) phi) of the thread's elements
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int e = j * NT + tid;
        const V2<R> p = phi[swz(e)], l = lam[swz(e)];
        w[j] = (double)l.x * p.y - (double)l.y * p.x;
    }
    // An uncontrolled Z marker's sum over the thread's elements, in the element order of the
    // general loop below (so the same roundings): on one of the tile's three top bits (bit kk of
    // the element number j), wt[kk] = sum of (bit kk of j ? -w : w); on any other qubit the sign is
    // the lane's own index bit, the same for all 8 elements: +-wsum.
    double wsum = 0.0, wt[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 8; j++) {
        wsum += w[j];
#pragma unroll
        for (int kk = 0; kk < 3; kk++) wt[kk] += ((j >> kk) & 1) ? -w[j] : w[j];
    }
    const int top0 = pd.q[pd.b - 3], top1 = pd.q[pd.b - 2], top2 = pd.q[pd.b - 1];
    for (int c0 = 0; c0 < cnt; c0 += 8) {
        const int cn = min(8, cnt - c0);
        double acc[8];
#pragma unroll
        for (int k = 0; k < 8; k++) acc[k] = 0.0;
        for (int k = 0; k < cn; k++) {  // one marker at a time keeps few values live
            const FOpDev<R>& o = ops[c0 + k];
            // opaque copies: per-element indices and lam reads are redone per marker rather than
            // hoisted out of the loop into registers the gate groups need
            int tk = tid;
            uint64_t gk = g_tid;
            asm volatile("" : "+v"(tk), "+v"(gk));
            double a = 0.0;
            if (o.g == GEN_Z && o.cm_ext == 0) {
                const int q = (int)o.cm_group;
                a = q == top0 ? wt[0]
                    : q == top1 ? wt[1]
                    : q == top2 ? wt[2]
                    : (((gk >> q) & 1ull) ? 0.0 - wsum : wsum);  // 0 - x: +0 for x = +0, as the loop gives
            } else if (o.g == GEN_Z) {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const uint64_t gi = gk | pd.top[j];
                    const double v = ((gi >> o.cm_group) & 1ull) ? -w[j] : w[j];
                    a += ((gi & o.cm_ext) == o.cm_ext) ? v : 0.0;
                }
            } else if (o.cm_ext == 0) {  // uncontrolled X / Y: no index tests
                const int lb = (int)o.cm_group;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int e = j * NT + tk;
                    const V2<R> l = lam[swz(e)], q = phi[swz(e ^ (1 << lb))];
                    double v;
                    if (o.g == GEN_X) {
                        v = (double)l.x * q.y - (double)l.y * q.x;
                    } else {
                        const double re = (double)l.x * q.x + (double)l.y * q.y;
                        v = ((e >> lb) & 1) ? re : -re;
                    }
                    a += v;
                }
            } else {
                const int lb = (int)o.cm_group;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int e = j * NT + tk;
                    const uint64_t gi = gk | pd.top[j];
                    const V2<R> l = lam[swz(e)], q = phi[swz(e ^ (1 << lb))];
                    double v;
                    if (o.g == GEN_X) {
                        v = (double)l.x * q.y - (double)l.y * q.x;
                    } else {  // Y: (Y psi)[e] = +i q where bit lb of e is set, -i q where clear
                        const double re = (double)l.x * q.x + (double)l.y * q.y;
                        v = ((e >> lb) & 1) ? re : -re;
                    }
                    a += ((gi & o.cm_ext) == o.cm_ext) ? v : 0.0;
                }
            }
#pragma unroll
            for (int kk = 0; kk < 8; kk++) acc[kk] = kk == k ? a : acc[kk];
        }
        const double t = wave_total8(acc, lane);
        const int m = (((lane >> 5) & 1) << 2) | (((lane >> 4) & 1) << 1) | ((lane >> 3) & 1);
        if ((lane & 7) == 0 && m < cn) red[m * 8 + wave] = t;
        __syncthreads();
        if (tid < cn) {
            double sum = 0.0;
            for (int wv = 0; wv < (NT >> 6); wv++) sum += red[tid * 8 + wv];
            partials[(uint64_t)ops[c0 + tid].flags * tiles + blockIdx.x] = sum;
        }
        __syncthreads();
    }
}

// Streaming accesses: the state is far larger than L2 and the MALL, and a pass touches each line
// once, so tile loads and stores carry the non-temporal hint (the split-row kernel's direct
// HBM -> LDS loads set it in their cache-policy operand).
template <typename R>
__device__ __forceinline__ V2<R> load_nt(const V2<R>* p) {
    typedef R N2 __attribute__((ext_vector_type(2)));
    const N2 n = __builtin_nontemporal_load(reinterpret_cast<const N2*>(p));
    return mk<R>(n.x, n.y);
}
template <typename R>
__device__ __forceinline__ void store_nt(V2<R>* p, V2<R> v) {
    typedef R N2 __attribute__((ext_vector_type(2)));
    N2 n;
    n.x = v.x;
    n.y = v.y;
    __builtin_nontemporal_store(n, reinterpret_cast<N2*>(p));
}

// NS = 1: one state. NS = 2: the reverse sweep; every gate acts on psi and psi1 alike (their tiles
// sit side by side in LDS) and stage groups evaluate gradient markers between compute groups.
template <typename R, int NS>
__global__ void __launch_bounds__(512) k_fused_pass(V2<R>* __restrict__ psi, V2<R>* __restrict__ psi1,
                                                       const PassDesc pd, const FOpDev<R>* __restrict__ ops,
                                                       const unsigned char* __restrict__ runs,
                                                       double* __restrict__ partials, uint64_t tiles) {
    extern __shared__ unsigned char lds_raw[];
    V2<R>* tile = reinterpret_cast<V2<R>*>(lds_raw);
    const int b = pd.b;
    const int T = 1 << b;
    uint64_t base = blockIdx.x;
    for (int k = 0; k < b; k++) {
        const uint64_t low = (1ull << pd.q[k]) - 1;
        base = (base & low) | ((base & ~low) << 1);
    }
    // element l = l_hi + tid: the thread's bits (below log2 NT) and the loop's bits deposit into
    // disjoint global bits, so the loop part is uniform and stays in scalar registers
    const int tid = threadIdx.x;
    const int NT = blockDim.x;  // 64 .. 512, a power of two
    const uint64_t g_tid = base | local_to_global(tid, pd);
    // one row per thread (NT = T / 16): the loop runs over the tile's 4 highest local bits, whose
    // global offsets come precomputed (pd.top16), so the loads issue back to back
    const bool rows16 = NT == (T >> 4);
#pragma unroll
    for (int st = 0; st < NS; st++) {
        const V2<R>* src_state = st ? psi1 : psi;
        V2<R>* tl = tile + st * T;
        if (sizeof(R) == 8 && T >= NT && NT >= 128) {
            // direct HBM -> LDS loads (gfx950 global_load_lds_dwordx4): LDS slot s = lh + tid
            // receives element swz(s) = lh + swz(tid) (the swizzle only mixes bits 0-6, and lh is a
            // multiple of NT >= 128), so no VGPRs hold the tile in flight
            const uint64_t g_sw = base | local_to_global(swz(tid), pd);
            const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
            if (rows16) {
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    const void* src = (const void*)(src_state + (g_sw | pd.top16[j]));
                    auto* dst = (__attribute__((address_space(3))) void*)(tl + j * NT + wave * 64);
                    if constexpr (sizeof(R) == 8) __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
                }
            } else {
#pragma unroll 16
                for (int lh = 0; lh < T; lh += NT) {
                    const void* src = (const void*)(src_state + (g_sw | local_to_global(lh, pd)));
                    auto* dst = (__attribute__((address_space(3))) void*)(tl + lh + wave * 64);
                    if constexpr (sizeof(R) == 8) __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
                }
            }
        } else if (rows16) {
#pragma unroll
            for (int j = 0; j < 16; j++) tl[swz(j * NT + tid)] = load_nt<R>(&src_state[g_tid | pd.top16[j]]);
        } else {
#pragma unroll 16
            for (int lh = 0; lh < T; lh += NT)
                if (lh + tid < T) tl[swz(lh + tid)] = src_state[g_tid | local_to_global(lh, pd)];
        }
    }
    if (sizeof(R) == 8 && T >= NT && NT >= 128) __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    const int rows = T >> 4;
    for (int gi = 0; gi < pd.ngroups; gi++) {
        const int o_begin = pd.gop[gi], o_end = pd.gop[gi + 1];
        if constexpr (NS == 2) {
            if ((pd.stage_mask >> gi) & 1u) {
                double* red = reinterpret_cast<double*>(tile + 2 * T);
                grad_stage<R>(tile, tile + T, T, g_tid, pd, ops + o_begin, o_end - o_begin, red, partials, tiles);
                continue;  // grad_stage ends with a barrier
            }
        }
        const int s0 = pd.gbits[gi][0], s1 = pd.gbits[gi][1], s2 = pd.gbits[gi][2], s3 = pd.gbits[gi][3];
        for (int r = tid; r < rows; r += NT) {
            const int l0 = ins_bit(ins_bit(ins_bit(ins_bit(r, s0), s1), s2), s3);
            int a[16];
            row_slots(a, l0, s0, s1, s2, s3, std::make_integer_sequence<int, 16>{});
            const uint64_t G0 = base | local_to_global(l0, pd);
            for (int st = 0; st < NS; st++) {
                V2<R>* tl = tile + st * T;
                V2<R> v[16];
                gather_seq<R>(v, tl, a, std::make_integer_sequence<int, 16>{});
                for (int oi = o_begin; oi < o_end; oi++) apply_op<R>(v, ops[oi], G0, runs);
                const uint32_t pc = pd.gpxc[gi];
                const int xo = swz((px_bit(G0, pd.gpxm[gi][0], pc) << s0) | (px_bit(G0, pd.gpxm[gi][1], pc >> 1) << s1) |
                                   (px_bit(G0, pd.gpxm[gi][2], pc >> 2) << s2) |
                                   (px_bit(G0, pd.gpxm[gi][3], pc >> 3) << s3));
                scatter_seq<R>(v, tl, a, xo, std::make_integer_sequence<int, 16>{});
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int st = 0; st < NS; st++) {
        V2<R>* dst_state = st ? psi1 : psi;
        const V2<R>* tl = tile + st * T;
        if (rows16) {
#pragma unroll
            for (int j = 0; j < 16; j++) store_nt<R>(&dst_state[g_tid | pd.top16[j]], tl[swz(j * NT + tid)]);
        } else {
#pragma unroll 4
            for (int lh = 0; lh < T; lh += NT) {
                if (lh + tid < T) dst_state[g_tid | local_to_global(lh, pd)] = tl[swz(lh + tid)];
            }
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Persistent variant for one-state complex128 passes with one row per thread (2^11 .. 2^13
// amplitudes per tile). Each workgroup walks the tiles blockIdx.x, blockIdx.x + gridDim.x, ...; as
// soon as every thread has gathered its row of the last group, the next tile's HBM -> LDS load is
// issued, so it runs under the last group's arithmetic, and the last group stores its registers
// straight to HBM (register member P holds logical member P ^ px: global index G0 | bits(P ^ px)).
// `stagger` delays the workgroups with bit `stagger_shift` of blockIdx set by that many 8128-cycle
// sleeps before their first tile, so the load and compute phases of co-resident workgroups can be
// set apart. Algorithmic bytes per launch: 2 * 16 * 2^n, as k_fused_pass.
template <int... Ms>
__device__ __forceinline__ void member_offsets(uint64_t (&gm)[16], uint64_t q0, uint64_t q1, uint64_t q2, uint64_t q3,
                                               std::integer_sequence<int, Ms...>) {
    ((gm[Ms] = ((Ms & 1) ? q0 : 0) | ((Ms & 2) ? q1 : 0) | ((Ms & 4) ? q2 : 0) | ((Ms & 8) ? q3 : 0)), ...);
}

template <typename R, int... Ms>
__device__ __forceinline__ void store_seq(const V2<R> (&v)[16], V2<R>* __restrict__ psi, uint64_t g,
                                          const uint64_t (&gm)[16], std::integer_sequence<int, Ms...>) {
    ((psi[g ^ gm[Ms]] = v[Ms]), ...);
}

__device__ __forceinline__ uint64_t tile_base(uint64_t t, const PassDesc& pd) {
    for (int k = 0; k < pd.b; k++) {
        const uint64_t low = (1ull << pd.q[k]) - 1;
        t = (t & low) | ((t & ~low) << 1);
    }
    return t;
}

__global__ void __launch_bounds__(512) k_fused_stream(double2* __restrict__ psi, const PassDesc pd,
                                                       const FOpDev<double>* __restrict__ ops,
                                                       const unsigned char* __restrict__ runs, uint64_t tiles,
                                                       uint32_t stagger, uint32_t stagger_shift) {
    extern __shared__ unsigned char lds_raw[];
    double2* tile = reinterpret_cast<double2*>(lds_raw);
    const int T = 1 << pd.b;
    const int tid = threadIdx.x, NT = blockDim.x;  // NT = T / 16: one row per thread
    const int wave = tid >> 6;
    uint64_t t = blockIdx.x;
    if (t >= tiles) return;
    if (stagger != 0 && ((blockIdx.x >> stagger_shift) & 1u))
        for (uint32_t i = 0; i < stagger; i++) __builtin_amdgcn_s_sleep(127);
    const uint64_t l_sw = local_to_global(swz(tid), pd);
    auto issue_load = [&](uint64_t base) {
        // LDS slot lh + tid receives element lh + swz(tid) (see k_fused_pass)
#pragma unroll 16
        for (int lh = 0; lh < T; lh += NT) {
            const void* src = (const void*)(psi + (base | l_sw | local_to_global(lh, pd)));
            auto* dst = (__attribute__((address_space(3))) void*)(tile + lh + wave * 64);
            __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
        }
    };
    uint64_t base = tile_base(t, pd);
    issue_load(base);
    const int last = pd.ngroups - 1;
    // the last group's member offsets in global index space (uniform)
    const int ls0 = pd.gbits[last][0], ls1 = pd.gbits[last][1], ls2 = pd.gbits[last][2], ls3 = pd.gbits[last][3];
    uint64_t gm[16];
    member_offsets(gm, 1ull << pd.q[ls0], 1ull << pd.q[ls1], 1ull << pd.q[ls2], 1ull << pd.q[ls3],
                   std::make_integer_sequence<int, 16>{});
    const int l_last = ins_bit(ins_bit(ins_bit(ins_bit(tid, ls0), ls1), ls2), ls3);
    int a_last[16];
    row_slots(a_last, l_last, ls0, ls1, ls2, ls3, std::make_integer_sequence<int, 16>{});
    const uint64_t g_last = local_to_global(l_last, pd);
    for (;;) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int gi = 0; gi < last; gi++) {
            const int s0 = pd.gbits[gi][0], s1 = pd.gbits[gi][1], s2 = pd.gbits[gi][2], s3 = pd.gbits[gi][3];
            const int l0 = ins_bit(ins_bit(ins_bit(ins_bit(tid, s0), s1), s2), s3);
            int a[16];
            row_slots(a, l0, s0, s1, s2, s3, std::make_integer_sequence<int, 16>{});
            const uint64_t G0 = base | local_to_global(l0, pd);
            double2 v[16];
            gather_seq<double>(v, tile, a, std::make_integer_sequence<int, 16>{});
            for (int oi = pd.gop[gi]; oi < pd.gop[gi + 1]; oi++) apply_op<double>(v, ops[oi], G0, runs);
            const uint32_t pc = pd.gpxc[gi];
            const int xo = swz((px_bit(G0, pd.gpxm[gi][0], pc) << s0) | (px_bit(G0, pd.gpxm[gi][1], pc >> 1) << s1) |
                               (px_bit(G0, pd.gpxm[gi][2], pc >> 2) << s2) |
                               (px_bit(G0, pd.gpxm[gi][3], pc >> 3) << s3));
            scatter_seq<double>(v, tile, a, xo, std::make_integer_sequence<int, 16>{});
            __syncthreads();
        }
        const uint64_t G0 = base | g_last;
        double2 v[16];
        gather_seq<double>(v, tile, a_last, std::make_integer_sequence<int, 16>{});
        __syncthreads();  // every row of the tile is in registers: LDS is free for the next tile
        const uint64_t tn = t + gridDim.x;
        const uint64_t base_n = tn < tiles ? tile_base(tn, pd) : 0;
        if (tn < tiles) issue_load(base_n);
        for (int oi = pd.gop[last]; oi < pd.gop[last + 1]; oi++) apply_op<double>(v, ops[oi], G0, runs);
        const uint32_t pc = pd.gpxc[last];
        const uint64_t pxg = (px_bit(G0, pd.gpxm[last][0], pc) ? gm[1] : 0) |
                             (px_bit(G0, pd.gpxm[last][1], pc >> 1) ? gm[2] : 0) |
                             (px_bit(G0, pd.gpxm[last][2], pc >> 2) ? gm[4] : 0) |
                             (px_bit(G0, pd.gpxm[last][3], pc >> 3) ? gm[8] : 0);
        store_seq<double>(v, psi, G0 ^ pxg, gm, std::make_integer_sequence<int, 16>{});
        if (tn >= tiles) break;
        t = tn;
        base = base_n;
    }
}

// ------------------------------------------------------------------------------------------------
// Split-row variant (one-state complex128 passes of 2^11 .. 2^13 amplitudes): two lanes (lane,
// lane ^ 8) share each 16-member row; lane bit 3 = h holds the members with group bit 3 equal to h,
// register m <-> member (h << 3) | m. A tile then has 2^(b-3) threads, twice the waves of
// k_fused_pass, to stream it in and out, and a thread holds 8 amplitudes instead of 16. Gates on
// group bits 0..2 stay lane-local; a gate on group bit 3 pairs register m of the two lanes: the
// partner's value comes over a DPP row rotation by 8 and each lane computes its own output with
// per-lane coefficients (m00, m01 where h = 0; m11, m10 where h = 1). Controls on group bit 3 and
// the diagonal-run subsets that contain it become per-lane conditions.

// the partner lane's (lane ^ 8) copy of a value: DPP row_ror:8 (0x128) rotates each 16-lane row by
// 8; every lane of the wavefront must be active
__device__ __forceinline__ double xlane(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, 0x128, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), 0x128, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float xlane(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x128, 0xF, 0xF, false));
}
template <typename R>
__device__ __forceinline__ V2<R> xlane_v(V2<R> v) {
    return mk<R>(xlane(v.x), xlane(v.y));
}
// bit pattern of one component (relabel flushes blend registers bitwise)
template <typename R>
using UBits = std::conditional_t<sizeof(R) == 8, uint64_t, uint32_t>;

// 2x2 kinds on a lane-local group bit G < 3 (pairs m, m | 1 << G) with group controls cm7 on bits 0..2
// (CTRL = false: no control among bits 0..2, no per-pair tests)
template <typename R, int KIND, int G, int M, bool CTRL>
__device__ __forceinline__ void hpair_local(V2<R> (&v)[8], int cm7, V2<R> m00, V2<R> m01, V2<R> m10,
                                            V2<R> m11) {
    if constexpr ((M >> G) & 1) {
        return;
    } else {
        if constexpr (CTRL) {
            if ((M & cm7) != cm7) return;
        }
        constexpr int M1 = M | (1 << G);
        const V2<R> a = v[M], b = v[M1];
        if constexpr (KIND == F_SWAP) {
            v[M] = b;
            v[M1] = a;
        } else if constexpr (KIND == F_DIAG) {
            v[M] = cmul<R>(m00, a);
            v[M1] = cmul<R>(m11, b);
        } else if constexpr (KIND == F_REAL) {
            v[M] = mk<R>(m00.x * a.x + m01.x * b.x, m00.x * a.y + m01.x * b.y);
            v[M1] = mk<R>(m10.x * a.x + m11.x * b.x, m10.x * a.y + m11.x * b.y);
        } else {
            v[M] = cmul2<R>(m00, a, m01, b);
            v[M1] = cmul2<R>(m10, a, m11, b);
        }
    }
}

// 2x2 kinds on group bit 3 (across the lane pair): own coefficient c0, partner coefficient c1
template <typename R, int KIND, int M, bool CTRL>
__device__ __forceinline__ void hpair_cross(V2<R> (&v)[8], int cm7, V2<R> c0, V2<R> c1) {
    if constexpr (CTRL) {
        if ((M & cm7) != cm7) return;
    }
    if constexpr (KIND == F_DIAG) {
        v[M] = cmul<R>(c0, v[M]);
    } else if constexpr (KIND == F_SWAP) {
        v[M] = xlane_v<R>(v[M]);
    } else {
        const V2<R> o = xlane_v<R>(v[M]);
        if constexpr (KIND == F_REAL)
            v[M] = mk<R>(c0.x * v[M].x + c1.x * o.x, c0.x * v[M].y + c1.x * o.y);
        else
            v[M] = cmul2<R>(c0, v[M], c1, o);
    }
}

template <typename R, int KIND, int G, int... Ms>
__device__ __forceinline__ void hpairs(V2<R> (&v)[8], int cm7, V2<R> m00, V2<R> m01, V2<R> m10, V2<R> m11,
                                       std::integer_sequence<int, Ms...>) {
    if constexpr (G < 3) {
        if (cm7 == 0)
            (hpair_local<R, KIND, G, Ms, false>(v, cm7, m00, m01, m10, m11), ...);
        else
            (hpair_local<R, KIND, G, Ms, true>(v, cm7, m00, m01, m10, m11), ...);
    } else {  // m00 / m01 carry this lane's own / partner coefficients
        if (cm7 == 0)
            (hpair_cross<R, KIND, Ms, false>(v, cm7, m00, m01), ...);
        else
            (hpair_cross<R, KIND, Ms, true>(v, cm7, m00, m01), ...);
    }
}

template <typename R, int KIND, int G>
__device__ __forceinline__ void hop_2x2(V2<R> (&v)[8], const FOpDev<R>& o, bool ok, uint64_t G0, int h) {
    V2<R> m00 = mk<R>(o.m[0], o.m[1]), m01 = mk<R>(o.m[2], o.m[3]);
    V2<R> m10 = mk<R>(o.m[4], o.m[5]), m11 = mk<R>(o.m[6], o.m[7]);
    if (o.flags & FL_TARGET_PX) {  // registers hold the pair swapped where the relabel bit is set
        const bool sw = px_bit(G0, o.pxmask, (o.flags >> 1) & 1u);
        const V2<R> t00 = m00, t01 = m01, t10 = m10, t11 = m11;
        m00 = csel<R>(sw, t11, t00);
        m01 = csel<R>(sw, t10, t01);
        m10 = csel<R>(sw, t01, t10);
        m11 = csel<R>(sw, t00, t11);
    }
    // lanes whose outside controls fail, or (G < 3) whose half misses a control on group bit 3,
    // apply the identity
    const bool lane_ok = ok && (G == 3 || !(o.cm_group & 8u) || h);
    if (o.cm_ext != 0 || (G < 3 && (o.cm_group & 8u))) {
        const V2<R> one = mk<R>(1, 0), zero = mk<R>(0, 0);
        m00 = csel<R>(lane_ok, m00, one);
        m01 = csel<R>(lane_ok, m01, zero);
        m10 = csel<R>(lane_ok, m10, zero);
        m11 = csel<R>(lane_ok, m11, one);
    }
    if constexpr (G == 3) {
        const V2<R> c0 = csel<R>(h, m11, m00), c1 = csel<R>(h, m10, m01);
        hpairs<R, KIND, 3>(v, (int)(o.cm_group & 7u), c0, c1, c0, c1, std::make_integer_sequence<int, 8>{});
    } else {
        hpairs<R, KIND, G>(v, (int)(o.cm_group & 7u), m00, m01, m10, m11, std::make_integer_sequence<int, 8>{});
    }
}

template <typename R, int KIND>
__device__ __forceinline__ void hop_kind(V2<R> (&v)[8], const FOpDev<R>& o, bool ok, uint64_t G0, int h) {
    switch (o.g) {
        case 0: hop_2x2<R, KIND, 0>(v, o, ok, G0, h); break;
        case 1: hop_2x2<R, KIND, 1>(v, o, ok, G0, h); break;
        case 2: hop_2x2<R, KIND, 2>(v, o, ok, G0, h); break;
        default: hop_2x2<R, KIND, 3>(v, o, ok, G0, h); break;
    }
}

// X on group bit 3 with group controls cm7 (bits 0..2): registers take the partner's value where
// the lane condition holds
template <typename R, int... Ms>
__device__ __forceinline__ void hswap_cross(V2<R> (&v)[8], int cm7, bool ok, std::integer_sequence<int, Ms...>) {
    if (cm7 == 0)
        ((v[Ms] = csel<R>(ok, xlane_v<R>(v[Ms]), v[Ms])), ...);
    else
        ((((Ms & cm7) == cm7) ? (void)(v[Ms] = csel<R>(ok, xlane_v<R>(v[Ms]), v[Ms])) : (void)0), ...);
}

// relabel flush of group bit G: pairs swap where sw is set (bitwise blend)
template <typename R, int G, int M>
__device__ __forceinline__ void hflush_step(V2<R> (&v)[8], UBits<R> msk) {
    if constexpr (G == 3) {
        const V2<R> o = xlane_v<R>(v[M]);
        const UBits<R> ax = __builtin_bit_cast(UBits<R>, v[M].x), ay = __builtin_bit_cast(UBits<R>, v[M].y);
        const UBits<R> bx = __builtin_bit_cast(UBits<R>, o.x), by = __builtin_bit_cast(UBits<R>, o.y);
        v[M].x = __builtin_bit_cast(R, ax ^ ((ax ^ bx) & msk));
        v[M].y = __builtin_bit_cast(R, ay ^ ((ay ^ by) & msk));
    } else if constexpr (!((M >> G) & 1)) {
        constexpr int M1 = M | (1 << G);
        const UBits<R> ax = __builtin_bit_cast(UBits<R>, v[M].x), ay = __builtin_bit_cast(UBits<R>, v[M].y);
        const UBits<R> bx = __builtin_bit_cast(UBits<R>, v[M1].x), by = __builtin_bit_cast(UBits<R>, v[M1].y);
        const UBits<R> tx = (ax ^ bx) & msk, ty = (ay ^ by) & msk;
        v[M].x = __builtin_bit_cast(R, ax ^ tx);
        v[M].y = __builtin_bit_cast(R, ay ^ ty);
        v[M1].x = __builtin_bit_cast(R, bx ^ tx);
        v[M1].y = __builtin_bit_cast(R, by ^ ty);
    }
}

template <typename R, int G, int... Ms>
__device__ __forceinline__ void hflush(V2<R> (&v)[8], int sw, std::integer_sequence<int, Ms...>) {
    const UBits<R> msk = (UBits<R>)0 - (UBits<R>)sw;
    (hflush_step<R, G, Ms>(v, msk), ...);
}

template <typename R, int... Ms>
__device__ __forceinline__ void hphase(V2<R> (&v)[8], int cm7, V2<R> d, std::integer_sequence<int, Ms...>) {
    if (cm7 == 0)
        ((v[Ms] = cmul<R>(d, v[Ms])), ...);
    else
        ((((Ms & cm7) == cm7) ? (void)(v[Ms] = cmul<R>(d, v[Ms])) : (void)0), ...);
}

// factor f_S of a diagonal run applied to this lane's members (h << 3) | m that contain S
template <typename R, int S, int... Ms>
__device__ __forceinline__ void hrun_subset(V2<R> (&v)[8], const RunHead<R>* __restrict__ rh,
                                            const RunTerm<R>* __restrict__ terms, uint32_t active,
                                            uint32_t angles, uint64_t G0, int h, std::integer_sequence<int, Ms...>) {
    if (!((active >> S) & 1u)) return;
    V2<R> f = mk<R>(rh->f0[2 * S], rh->f0[2 * S + 1]);
    R th = 0;
    const int t1 = rh->tbeg[S + 1];
    for (int t = rh->tbeg[S]; t < t1; t++) {
        const RunTerm<R>& tm = terms[t];
        if (tm.angle == 2) {
            th += tm.v[0] * (R)(G0 & tm.cl);
            continue;
        }
        const bool cond = (G0 & tm.cl) == tm.cl;
        const bool sel = tm.pm ? (bool)px_bit(G0, tm.pm, tm.pc) : (bool)(tm.pc & 1u);
        if (tm.angle) {
            th += cond ? (sel ? tm.v[1] : tm.v[0]) : (R)0;
        } else {
            V2<R> val = csel<R>(sel, mk<R>(tm.v[2], tm.v[3]), mk<R>(tm.v[0], tm.v[1]));
            val = csel<R>(cond, val, mk<R>(1, 0));
            f = cmul<R>(f, val);
        }
    }
    if ((angles >> S) & 1u) {
        R sn, cs;
        sin_cos(th, &sn, &cs);
        f = cmul<R>(f, mk<R>(cs, sn));
    }
    if constexpr ((S & 8) != 0) f = csel<R>(h, f, mk<R>(1, 0));
    constexpr int S7 = S & 7;
    ((((Ms & S7) == S7) ? (void)(v[Ms] = cmul<R>(f, v[Ms])) : (void)0), ...);
}

template <typename R, int... Ss>
__device__ __forceinline__ void hdiag_run(V2<R> (&v)[8], const unsigned char* __restrict__ runs,
                                          const FOpDev<R>& o, uint64_t G0, int h, std::integer_sequence<int, Ss...>) {
    const RunHead<R>* rh = reinterpret_cast<const RunHead<R>*>(runs + o.cm_ext);
    const RunTerm<R>* terms = reinterpret_cast<const RunTerm<R>*>(rh + 1);
    (hrun_subset<R, Ss>(v, rh, terms, (uint32_t)o.g, o.flags, G0, h, std::make_integer_sequence<int, 8>{}), ...);
}

template <typename R>
__device__ __forceinline__ void apply_hop(V2<R> (&v)[8], const FOpDev<R>& o, uint64_t G0, int h,
                                          const unsigned char* __restrict__ runs) {
    const bool ok = (G0 & o.cm_ext) == o.cm_ext;
    switch (o.kind) {
        case F_DENSE: hop_kind<R, F_DENSE>(v, o, ok, G0, h); break;
        case F_REAL: hop_kind<R, F_REAL>(v, o, ok, G0, h); break;
        case F_DIAG: hop_kind<R, F_DIAG>(v, o, ok, G0, h); break;
        case F_SWAP:
            if (o.g == 3) {
                hswap_cross<R>(v, (int)(o.cm_group & 7u), ok, std::make_integer_sequence<int, 8>{});
            } else if (o.cm_ext == 0 && !(o.cm_group & 8u)) {
                hop_kind<R, F_SWAP>(v, o, ok, G0, h);  // uniform register moves
            } else {
                hop_kind<R, F_DENSE>(v, o, ok, G0, h);  // X as a matrix, identity where a condition fails
            }
            break;
        case F_FLUSH: {
            const int sw = px_bit(G0, o.pxmask, (o.flags >> 1) & 1u);
            switch (o.g) {
                case 0: hflush<R, 0>(v, sw, std::make_integer_sequence<int, 8>{}); break;
                case 1: hflush<R, 1>(v, sw, std::make_integer_sequence<int, 8>{}); break;
                case 2: hflush<R, 2>(v, sw, std::make_integer_sequence<int, 8>{}); break;
                default: hflush<R, 3>(v, sw, std::make_integer_sequence<int, 8>{}); break;
            }
            break;
        }
        case F_DIAGRUN: hdiag_run<R>(v, runs, o, G0, h, std::make_integer_sequence<int, 16>{}); break;
        default: {  // F_PHASE
            V2<R> d = csel<R>((G0 >> o.g) & 1ull, mk<R>(o.m[6], o.m[7]), mk<R>(o.m[0], o.m[1]));
            d = csel<R>(ok && (!(o.cm_group & 8u) || h), d, mk<R>(1, 0));
            hphase<R>(v, (int)(o.cm_group & 7u), d, std::make_integer_sequence<int, 8>{});
        }
    }
}

// LDS slots of a lane's 8 members: the swizzle is linear over XOR and the member bits are disjoint
// from the row's bits, so slot(m) = swz(row) ^ swz(member offset), the second term uniform
template <int... Ms>
__device__ __forceinline__ void hrow_slots(int (&a)[8], int l0, int s0, int s1, int s2,
                                           std::integer_sequence<int, Ms...>) {
    const int sl = swz(l0);
    ((a[Ms] = sl ^ swz(((Ms & 1) << s0) | (((Ms >> 1) & 1) << s1) | (((Ms >> 2) & 1) << s2))), ...);
}

template <typename R, int... Ms>
__device__ __forceinline__ void hgather(V2<R> (&v)[8], const V2<R>* tile, const int (&a)[8],
                                        std::integer_sequence<int, Ms...>) {
    ((v[Ms] = tile[a[Ms]]), ...);
}

// the slots are recomputed from the row's swizzled base (slot ^ relabel offset), so no address
// register has to stay live across the op loop
template <typename R, int... Ms>
__device__ __forceinline__ void hscatter(const V2<R> (&v)[8], V2<R>* tile, int sx, int s0, int s1, int s2,
                                         std::integer_sequence<int, Ms...>) {
    ((tile[sx ^ swz(((Ms & 1) << s0) | (((Ms >> 1) & 1) << s1) | (((Ms >> 2) & 1) << s2))] = v[Ms]), ...);
}

// NS = 2: the reverse sweep, both states' tiles side by side in LDS (as k_fused_pass), gradient
// stages evaluated from LDS between compute groups.
// (two-state tiles fit two workgroups per CU at most, so that instantiation may use 256 registers)
template <typename R, int NS>
__global__ void __launch_bounds__(NS == 1 ? 1024 : 512) k_fused_half(V2<R>* __restrict__ psi, V2<R>* __restrict__ psi1,
                                                     const PassDesc pd, const FOpDev<R>* __restrict__ ops,
                                                     const unsigned char* __restrict__ runs,
                                                     double* __restrict__ partials, uint64_t tiles) {
    extern __shared__ unsigned char lds_raw[];
    V2<R>* tile = reinterpret_cast<V2<R>*>(lds_raw);
    const int T = 1 << pd.b;
    const int tid = threadIdx.x, NT = blockDim.x;  // NT = T / 8: two lanes per row
    const uint64_t base = tile_base(blockIdx.x, pd);
    // after the tiles: the gradient stages' wavefront sums (NS = 2), then the global offsets of the
    // tile's high local bits (local index >> L)
    double* red = reinterpret_cast<double*>(tile + NS * T);
    uint64_t* hi_tab = reinterpret_cast<uint64_t*>(red + (NS == 2 ? 64 : 0));
    const int L = pd.L;
    {
        // element e of the tile goes to LDS slot swz(e) (the swizzle is an involution). complex128:
        // direct HBM -> LDS loads (global_load_lds_dwordx4), slot lh + tid receiving element
        // lh + swz(tid) (lh is a multiple of NT >= 128); complex64: 8-byte loads through registers
        // lh = j NT = j 2^(b-3) runs over the tile's 3 highest local bits, whose global offsets come
        // precomputed (pd.top), so the 8 loads issue back to back; the wavefront index is made
        // uniform so the LDS destination (M0) is scalar arithmetic
        // memory phases run at a raised wave priority, so a workgroup's loads and stores issue
        // ahead of the arithmetic of the co-resident workgroups (-1.8% per 30-qubit pass)
        __builtin_amdgcn_s_setprio(2);
        const uint64_t g_sw = base | local_to_global_low(sizeof(R) == 8 ? swz(tid) : tid, pd);
        const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
#pragma unroll
        for (int st = 0; st < NS; st++) {
            const V2<R>* src_state = st ? psi1 : psi;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int lh = j * NT;
                if constexpr (sizeof(R) == 8) {
                    const void* src = (const void*)(src_state + (g_sw | pd.top[j]));
                    auto* dst = (__attribute__((address_space(3))) void*)(tile + st * T + lh + wave * 64);
                    __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 2);  // aux 2: non-temporal
                } else {
                    tile[st * T + swz(lh + tid)] = load_nt<R>(&src_state[g_sw | pd.top[j]]);
                }
            }
        }
    }
    __builtin_amdgcn_s_setprio(0);
    // the high-bit table is built while the loads are in flight
    for (int i = tid; i < (T >> L); i += NT) hi_tab[i] = local_to_global(i << L, pd);
    if constexpr (sizeof(R) == 8) __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // the two lanes of a row are lane bit 3 apart, so the 8 lanes that one LDS cycle of a 16-byte
    // access serves (lanes 8k .. 8k+7) hold 8 consecutive rows, which the swizzle puts in distinct
    // banks (lane bit 0 pairing put both halves of 4 rows there: 2-way bank conflicts)
    const int h = (tid >> 3) & 1, r = (tid & 7) | ((tid >> 4) << 3);
    for (int gi = 0; gi < pd.ngroups; gi++) {
        if constexpr (NS == 2) {
            if ((pd.stage_mask >> gi) & 1u) {
                grad_stage8<R>(tile, tile + T, base | local_to_global(tid, pd), pd, ops + pd.gop[gi],
                               pd.gop[gi + 1] - pd.gop[gi], red, partials, tiles);
                continue;  // grad_stage8 ends with a barrier
            }
        }
        const int s0 = pd.gbits[gi][0], s1 = pd.gbits[gi][1], s2 = pd.gbits[gi][2], s3 = pd.gbits[gi][3];
        const int l0 = ins_bit(ins_bit(ins_bit(ins_bit(r, s0), s1), s2), s3);
        const int sl = swz(l0 | (h << s3));
        const uint64_t G0 = base | (uint64_t)(l0 & ((1 << L) - 1)) | hi_tab[l0 >> L];
        for (int st = 0; st < NS; st++) {
            V2<R>* tl = tile + st * T;
            V2<R> v[8];
            {
                int a[8];
                hrow_slots(a, l0 | (h << s3), s0, s1, s2, std::make_integer_sequence<int, 8>{});
                hgather<R>(v, tl, a, std::make_integer_sequence<int, 8>{});
            }
            for (int oi = pd.gop[gi]; oi < pd.gop[gi + 1]; oi++) apply_hop<R>(v, ops[oi], G0, h, runs);
            // the group's relabeling; bits without a parity mask are the same for every lane
            const uint32_t pc = pd.gpxc[gi];
            int px = 0;
#pragma unroll
            for (int j = 0; j < FUSE_GROUP_BITS; j++) {
                const uint64_t m = pd.gpxm[gi][j];
                const int bit = m ? px_bit(G0, m, pc >> j) : (int)((pc >> j) & 1u);
                px |= bit << pd.gbits[gi][j];
            }
            // a row's two lanes are in one wavefront, whose LDS reads all precede its writes
            hscatter<R>(v, tl, sl ^ swz(px), s0, s1, s2, std::make_integer_sequence<int, 8>{});
        }
        __syncthreads();
    }
    const uint64_t g_tid = base | local_to_global_low(tid, pd);
    __builtin_amdgcn_s_setprio(2);
#pragma unroll
    for (int st = 0; st < NS; st++) {
        V2<R>* dst_state = st ? psi1 : psi;
        const V2<R>* tl = tile + st * T;
#pragma unroll
        for (int j = 0; j < 8; j++) store_nt<R>(&dst_state[g_tid | pd.top[j]], tl[swz(j * NT + tid)]);
    }
}

// ------------------------------------------------------------------------------------------------
// Row-of-8 pass kernel (one-state passes of 2^10 .. 2^13-amplitude tiles, 3-qubit groups;
// complex128 by default, complex64 with option "rows8_c64"). One lane holds the 8 amplitudes that differ only in its group's 3 bits, so no gate
// needs another lane's data (the split-row kernel pairs lanes for its 4th group bit, and each of
// those gates pays 32 DPP moves); a 2^b tile still gets 2^(b-3) threads, and the load / store
// skeleton is the split-row kernel's.
// The op loop alternates two register sets (a -> o, then o -> a), so no op copies its results
// back, and the common ops -- a 2x2 on one group bit whose coefficients are the same for every
// lane -- keep their coefficients in scalar registers. The pass is bound by HBM only while the
// vector instructions per wave stay well under the load / store time, so the skeleton is kept
// lean: member slots are XORs of host-computed offsets and the lane's global index is built only
// in groups whose ops or relabeling depend on it. Device op kinds (FOpDev.kind):
enum K8Kind : int32_t {
    K8_U2 = 0,  // complex 2x2 on group bit g, the same for every lane; pairs outside the group
                //   controls cm_group keep their values
    K8_R2 = 1,  // the same with a real matrix
    K8_G2 = 2,  // complex 2x2 whose coefficients differ per lane: a pending relabel swaps them
                //   (FL_TARGET_PX), controls outside the group (cm_ext) make it the identity
    K8_DG = 3,  // diagonal: diag(d0, d1) on group bit g, or (g = -1 - q) on qubit q outside the
                //   group, with group / outside controls and relabel swaps
    K8_FL = 4,  // relabel flush of group bit g (pairs swap where the lane's parity is set)
    K8_DR = 5,  // diagonal run (the split-row kernel's records, subsets of 3 bits)
    K8_SW = 6,  // exchange of the pair members on group bit g where the group controls are set
                //   (X and controlled X inside the group): register moves only
    K8_D4 = 7   // dense 4x4 on group bits g & 15 (index bit 0) and g >> 4 (bit 1), the matrix (16
                //   complex, row-major) at byte offset pxmask of the run buffer; a group control on
                //   the third bit (cm) and controls outside the group (cm_ext) as for K8_G2
};

// An op record as the row-of-8 kernel reads it (wide scalar loads of the header, the outside-
// control and relabel masks and the 4 coefficients)
template <typename R>
struct Op8 {
    int32_t kind, g, cm, flags;
    uint64_t cm_ext, pxmask;
    R m[8];
};
template <typename R>
__device__ __forceinline__ Op8<R> load_op8(const FOpDev<R>* __restrict__ p) {
    static_assert(sizeof(FOpDev<R>) == 32 + 8 * sizeof(R), "op record layout");
    Op8<R> o;
    const int4 h = *reinterpret_cast<const int4*>(p);
    o.kind = h.x;
    o.g = h.y;
    o.cm = h.z;
    o.flags = h.w;
    const ulonglong2 mk2 = *reinterpret_cast<const ulonglong2*>(reinterpret_cast<const char*>(p) + 16);
    o.cm_ext = mk2.x;
    o.pxmask = mk2.y;
#pragma unroll
    for (int k = 0; k < 8; k++) o.m[k] = p->m[k];
    return o;
}

// Pair (M, M | 1 << G) of a gate on group bit G, for the pairs whose group-control bits CM are
// set (CM a compile-time subset of the other two bits); the other pairs are copied. OPK: 0 complex
// 2x2, 1 real 2x2, 2 exchange.
template <typename R, int G, int CM, int OPK, int M>
__device__ __forceinline__ void r8_pair(const V2<R> (&a)[8], V2<R> (&o)[8], V2<R> m00, V2<R> m01, V2<R> m10, V2<R> m11) {
    if constexpr (!((M >> G) & 1)) {
        constexpr int M1 = M | (1 << G);
        if constexpr ((M & CM) != CM) {
            o[M] = a[M];
            o[M1] = a[M1];
        } else if constexpr (OPK == 2) {
            o[M] = a[M1];
            o[M1] = a[M];
        } else {
            const V2<R> x = a[M], y = a[M1];
            if constexpr (OPK == 1) {
                o[M] = mk<R>(m00.x * x.x + m01.x * y.x, m00.x * x.y + m01.x * y.y);
                o[M1] = mk<R>(m10.x * x.x + m11.x * y.x, m10.x * x.y + m11.x * y.y);
            } else {
                o[M] = cmul2<R>(m00, x, m01, y);
                o[M1] = cmul2<R>(m10, x, m11, y);
            }
        }
    }
}

template <typename R, int G, int CM, int OPK, int... Ms>
__device__ __forceinline__ void r8_2x2(const V2<R> (&a)[8], V2<R> (&o)[8], V2<R> m00, V2<R> m01, V2<R> m10, V2<R> m11,
                                       std::integer_sequence<int, Ms...>) {
    (r8_pair<R, G, CM, OPK, Ms>(a, o, m00, m01, m10, m11), ...);
}

// the group-control masks of a gate on bit G: subsets of the other two bits
template <typename R, int G, int OPK>
__device__ __forceinline__ void r8_ctrl(const V2<R> (&a)[8], V2<R> (&o)[8], int cm, V2<R> m00, V2<R> m01, V2<R> m10,
                                        V2<R> m11) {
    constexpr int B0 = G == 0 ? 2 : 1, B1 = G == 2 ? 2 : 4;
    constexpr auto seq = std::make_integer_sequence<int, 8>{};
    if (cm == 0)
        r8_2x2<R, G, 0, OPK>(a, o, m00, m01, m10, m11, seq);
    else if (cm == B0)
        r8_2x2<R, G, B0, OPK>(a, o, m00, m01, m10, m11, seq);
    else if (cm == B1)
        r8_2x2<R, G, B1, OPK>(a, o, m00, m01, m10, m11, seq);
    else
        r8_2x2<R, G, B0 | B1, OPK>(a, o, m00, m01, m10, m11, seq);
}

template <typename R, int OPK>
__device__ __forceinline__ void r8_bit(const V2<R> (&a)[8], V2<R> (&o)[8], int g, int cm, V2<R> m00, V2<R> m01, V2<R> m10,
                                       V2<R> m11) {
    switch (g) {
        case 0: r8_ctrl<R, 0, OPK>(a, o, cm, m00, m01, m10, m11); break;
        case 1: r8_ctrl<R, 1, OPK>(a, o, cm, m00, m01, m10, m11); break;
        default: r8_ctrl<R, 2, OPK>(a, o, cm, m00, m01, m10, m11); break;
    }
}

// diagonal factors: member M gets f1 where its bit G is set (G = -1: f0 everywhere), where its
// group-control bits CM are set; the others are copied
template <typename R, int G, int CM, int M>
__device__ __forceinline__ void r8_scale(const V2<R> (&a)[8], V2<R> (&o)[8], V2<R> f0, V2<R> f1) {
    if constexpr ((M & CM) != CM) {
        o[M] = a[M];
    } else if constexpr (G >= 0 && ((M >> G) & 1)) {
        o[M] = cmul<R>(f1, a[M]);
    } else {
        o[M] = cmul<R>(f0, a[M]);
    }
}

template <typename R, int G, int CM, int... Ms>
__device__ __forceinline__ void r8_diag_cm(const V2<R> (&a)[8], V2<R> (&o)[8], V2<R> f0, V2<R> f1,
                                           std::integer_sequence<int, Ms...>) {
    (r8_scale<R, G, CM, Ms>(a, o, f0, f1), ...);
}

template <typename R, int G>
__device__ __forceinline__ void r8_diag(const V2<R> (&a)[8], V2<R> (&o)[8], int cm, V2<R> f0, V2<R> f1) {
    constexpr auto seq = std::make_integer_sequence<int, 8>{};
    switch (cm) {  // any subset of the 3 group bits (never the target's own bit)
        case 0: r8_diag_cm<R, G, 0>(a, o, f0, f1, seq); break;
        case 1: r8_diag_cm<R, G, 1>(a, o, f0, f1, seq); break;
        case 2: r8_diag_cm<R, G, 2>(a, o, f0, f1, seq); break;
        case 3: r8_diag_cm<R, G, 3>(a, o, f0, f1, seq); break;
        case 4: r8_diag_cm<R, G, 4>(a, o, f0, f1, seq); break;
        case 5: r8_diag_cm<R, G, 5>(a, o, f0, f1, seq); break;
        case 6: r8_diag_cm<R, G, 6>(a, o, f0, f1, seq); break;
        default: r8_diag_cm<R, G, 7>(a, o, f0, f1, seq); break;
    }
}

template <typename R, int G, int M>
__device__ __forceinline__ void r8_flush_pair(const V2<R> (&a)[8], V2<R> (&o)[8], bool sw) {
    if constexpr (!((M >> G) & 1)) {
        constexpr int M1 = M | (1 << G);
        o[M] = csel<R>(sw, a[M1], a[M]);
        o[M1] = csel<R>(sw, a[M], a[M1]);
    }
}

template <typename R, int G, int... Ms>
__device__ __forceinline__ void r8_flush(const V2<R> (&a)[8], V2<R> (&o)[8], bool sw, std::integer_sequence<int, Ms...>) {
    (r8_flush_pair<R, G, Ms>(a, o, sw), ...);
}

// factor f_S of a diagonal run applied to the members M that contain S (terms as in run_subset)
template <typename R, int S, int... Ms>
__device__ __forceinline__ void r8_run_subset(V2<R> (&v)[8], const RunHead<R>* __restrict__ rh,
                                              const RunTerm<R>* __restrict__ terms, uint32_t active, uint32_t angles,
                                              uint64_t G0, std::integer_sequence<int, Ms...>) {
    if (!((active >> S) & 1u)) return;
    V2<R> f = mk<R>(rh->f0[2 * S], rh->f0[2 * S + 1]);
    R th = 0;
    const int t1 = rh->tbeg[S + 1];
    for (int t = rh->tbeg[S]; t < t1; t++) {
        const RunTerm<R>& tm = terms[t];
        if (tm.angle == 2) {
            th += tm.v[0] * (R)(G0 & tm.cl);
            continue;
        }
        const bool cond = (G0 & tm.cl) == tm.cl;
        const bool sel = tm.pm ? (bool)px_bit(G0, tm.pm, tm.pc) : (bool)(tm.pc & 1u);
        if (tm.angle) {
            th += cond ? (sel ? tm.v[1] : tm.v[0]) : (R)0;
        } else {
            V2<R> val = csel<R>(sel, mk<R>(tm.v[2], tm.v[3]), mk<R>(tm.v[0], tm.v[1]));
            val = csel<R>(cond, val, mk<R>(1, 0));
            f = cmul<R>(f, val);
        }
    }
    if ((angles >> S) & 1u) {
        R sn, cs;
        sin_cos(th, &sn, &cs);
        f = cmul<R>(f, mk<R>(cs, sn));
    }
    ((((Ms & S) == S) ? (void)(v[Ms] = cmul<R>(f, v[Ms])) : (void)0), ...);
}

template <typename R, int... Ss>
__device__ __forceinline__ void r8_run(const V2<R> (&a)[8], V2<R> (&o)[8], const unsigned char* __restrict__ runs,
                                       const Op8<R>& op, uint64_t G0, std::integer_sequence<int, Ss...>) {
    const RunHead<R>* rh = reinterpret_cast<const RunHead<R>*>(runs + op.cm_ext);
    const RunTerm<R>* terms = reinterpret_cast<const RunTerm<R>*>(rh + 1);
#pragma unroll
    for (int m = 0; m < 8; m++) o[m] = a[m];
    (r8_run_subset<R, Ss>(o, rh, terms, (uint32_t)op.g, (uint32_t)op.flags, G0, std::make_integer_sequence<int, 8>{}),
     ...);
}

// A 4x4 on group bits T0 (index bit 0) and T1 (bit 1) of the row, for both values of the third bit
// (CM = 1: a group control on the third bit, so only its set half changes). Where the lane's outside
// controls do not hold (ok false), the members are copied. The matrix is read one row at a time
// (8 scalar registers of coefficients live at once; the compiler barrier keeps the loads of the
// next row from being hoisted).
template <typename R, int T0, int T1, int CM>
__device__ __forceinline__ void r8_d4(const V2<R> (&a)[8], V2<R> (&o)[8], const R* __restrict__ M, bool ok) {
    constexpr int TB = 3 - T0 - T1;
    constexpr int P0 = 0, P1 = 1 << T0, P2 = 1 << T1, P3 = (1 << T0) | (1 << T1), H = 1 << TB;
    V2<R> y0[4], y1[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        asm volatile("" ::: "memory");
        R m[8];
#pragma unroll
        for (int k = 0; k < 8; k++) m[k] = M[8 * r + k];
        auto dot = [&](V2<R> x0, V2<R> x1, V2<R> x2, V2<R> x3) {
            R re = m[0] * x0.x - m[1] * x0.y;
            R im = m[0] * x0.y + m[1] * x0.x;
            re = fma(m[2], x1.x, fma(-m[3], x1.y, re));
            im = fma(m[2], x1.y, fma(m[3], x1.x, im));
            re = fma(m[4], x2.x, fma(-m[5], x2.y, re));
            im = fma(m[4], x2.y, fma(m[5], x2.x, im));
            re = fma(m[6], x3.x, fma(-m[7], x3.y, re));
            im = fma(m[6], x3.y, fma(m[7], x3.x, im));
            return mk<R>(re, im);
        };
        y1[r] = dot(a[H | P0], a[H | P1], a[H | P2], a[H | P3]);
        if constexpr (CM == 0) y0[r] = dot(a[P0], a[P1], a[P2], a[P3]);
    }
    o[H | P0] = csel<R>(ok, y1[0], a[H | P0]);
    o[H | P1] = csel<R>(ok, y1[1], a[H | P1]);
    o[H | P2] = csel<R>(ok, y1[2], a[H | P2]);
    o[H | P3] = csel<R>(ok, y1[3], a[H | P3]);
    if constexpr (CM == 0) {
        o[P0] = csel<R>(ok, y0[0], a[P0]);
        o[P1] = csel<R>(ok, y0[1], a[P1]);
        o[P2] = csel<R>(ok, y0[2], a[P2]);
        o[P3] = csel<R>(ok, y0[3], a[P3]);
    } else {
        o[P0] = a[P0];
        o[P1] = a[P1];
        o[P2] = a[P2];
        o[P3] = a[P3];
    }
}

template <typename R, int T0, int T1>
__device__ __forceinline__ void r8_d4_cm(const V2<R> (&a)[8], V2<R> (&o)[8], int cm, const R* __restrict__ M, bool ok) {
    if (cm == 0)
        r8_d4<R, T0, T1, 0>(a, o, M, ok);
    else
        r8_d4<R, T0, T1, 1>(a, o, M, ok);
}

// D4: the pass holds 4x4 ops (K8_D4); kernels without them leave the case out, so the common
// op dispatch keeps its code size and register allocation
template <typename R, bool D4>
__device__ __forceinline__ void r8_op(const V2<R> (&a)[8], V2<R> (&o)[8], const Op8<R>& op, uint64_t G0,
                                      const unsigned char* __restrict__ runs) {
    switch (op.kind) {
        case K8_U2:
            r8_bit<R, 0>(a, o, op.g, op.cm, mk<R>(op.m[0], op.m[1]), mk<R>(op.m[2], op.m[3]), mk<R>(op.m[4], op.m[5]),
                         mk<R>(op.m[6], op.m[7]));
            break;
        case K8_R2:
            r8_bit<R, 1>(a, o, op.g, op.cm, mk<R>(op.m[0], 0), mk<R>(op.m[2], 0), mk<R>(op.m[4], 0), mk<R>(op.m[6], 0));
            break;
        case K8_SW: {
            const V2<R> z = mk<R>(0, 0);
            r8_bit<R, 2>(a, o, op.g, op.cm, z, z, z, z);
            break;
        }
        case K8_G2: {
            V2<R> m00 = mk<R>(op.m[0], op.m[1]), m01 = mk<R>(op.m[2], op.m[3]);
            V2<R> m10 = mk<R>(op.m[4], op.m[5]), m11 = mk<R>(op.m[6], op.m[7]);
            if (op.flags & FL_TARGET_PX) {  // registers hold the pair swapped where the relabel bit is set
                const bool sw = px_bit(G0, op.pxmask, ((uint32_t)op.flags >> 1) & 1u);
                const V2<R> t00 = m00, t01 = m01, t10 = m10, t11 = m11;
                m00 = csel<R>(sw, t11, t00);
                m01 = csel<R>(sw, t10, t01);
                m10 = csel<R>(sw, t01, t10);
                m11 = csel<R>(sw, t00, t11);
            }
            if (op.cm_ext) {
                const bool ok = (G0 & op.cm_ext) == op.cm_ext;
                m00 = csel<R>(ok, m00, mk<R>(1, 0));
                m01 = csel<R>(ok, m01, mk<R>(0, 0));
                m10 = csel<R>(ok, m10, mk<R>(0, 0));
                m11 = csel<R>(ok, m11, mk<R>(1, 0));
            }
            r8_bit<R, 0>(a, o, op.g, op.cm, m00, m01, m10, m11);
            break;
        }
        case K8_DG: {
            const V2<R> d0 = mk<R>(op.m[0], op.m[1]), d1 = mk<R>(op.m[6], op.m[7]);
            // an outside target picks the factor by the lane's index bit; a pending relabel of an
            // inside target swaps the two
            const bool outside = op.g < 0;
            const bool sel = outside ? (bool)((G0 >> (-1 - op.g)) & 1ull)
                                     : ((op.flags & FL_TARGET_PX)
                                            ? (bool)px_bit(G0, op.pxmask, ((uint32_t)op.flags >> 1) & 1u)
                                            : false);
            const bool ok = (G0 & op.cm_ext) == op.cm_ext;
            V2<R> f0 = csel<R>(sel, d1, d0), f1 = csel<R>(sel, d0, d1);
            f0 = csel<R>(ok, f0, mk<R>(1, 0));
            f1 = csel<R>(ok, f1, mk<R>(1, 0));
            switch (op.g) {
                case 0: r8_diag<R, 0>(a, o, op.cm, f0, f1); break;
                case 1: r8_diag<R, 1>(a, o, op.cm, f0, f1); break;
                case 2: r8_diag<R, 2>(a, o, op.cm, f0, f1); break;
                default: r8_diag<R, -1>(a, o, op.cm, f0, f1); break;
            }
            break;
        }
        case K8_FL: {
            const bool sw = px_bit(G0, op.pxmask, ((uint32_t)op.flags >> 1) & 1u);
            switch (op.g) {
                case 0: r8_flush<R, 0>(a, o, sw, std::make_integer_sequence<int, 8>{}); break;
                case 1: r8_flush<R, 1>(a, o, sw, std::make_integer_sequence<int, 8>{}); break;
                default: r8_flush<R, 2>(a, o, sw, std::make_integer_sequence<int, 8>{}); break;
            }
            break;
        }
        case K8_D4: {
            if constexpr (D4) {
                const R* M = reinterpret_cast<const R*>(runs + op.pxmask);
                const bool ok = (G0 & op.cm_ext) == op.cm_ext;
                switch (op.g) {
                    case 0x10: r8_d4_cm<R, 0, 1>(a, o, op.cm, M, ok); break;
                    case 0x01: r8_d4_cm<R, 1, 0>(a, o, op.cm, M, ok); break;
                    case 0x20: r8_d4_cm<R, 0, 2>(a, o, op.cm, M, ok); break;
                    case 0x02: r8_d4_cm<R, 2, 0>(a, o, op.cm, M, ok); break;
                    case 0x21: r8_d4_cm<R, 1, 2>(a, o, op.cm, M, ok); break;
                    default: r8_d4_cm<R, 2, 1>(a, o, op.cm, M, ok); break;
                }
                break;
            }
            // kernels without 4x4 ops: the label joins the default case (never taken)
            [[fallthrough]];
        }
        default:
            r8_run<R>(a, o, runs, op, G0, std::make_integer_sequence<int, 8>{});
            break;
    }
}

// global offset of a local index: identity below L, then bit k -> qubit q[k] (branch-free)
__device__ __forceinline__ uint64_t local_to_global_bits(int l, const PassDesc& pd) {
    uint64_t g = (uint64_t)(l & ((1 << pd.L) - 1));
    for (int k = pd.L; k < pd.b; k++) g |= (uint64_t)((l >> k) & 1) << pd.q[k];
    return g;
}

// Loads the tile at global offset `base` into LDS, element e at slot swz(e) (as k_fused_half):
// complex128 with direct HBM -> LDS loads (slot lh + tid receives element lh + swz(tid)), complex64
// through registers (8-byte loads); the 8 loads cover the tile's 3 highest local bits.
template <typename R>
__device__ __forceinline__ void pass8_load(const V2<R>* __restrict__ psi, V2<R>* tile, uint64_t base, int tid,
                                           const PassDesc& pd, int NT, int wave) {
    __builtin_amdgcn_s_setprio(2);
    if constexpr (sizeof(R) == 8) {
        const uint64_t g = base | local_to_global_low(swz(tid), pd);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const void* src = (const void*)(psi + (g | pd.top[j]));
            auto* dst = (__attribute__((address_space(3))) void*)(tile + j * NT + wave * 64);
            __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 2);  // aux 2: non-temporal
        }
    } else {
        const uint64_t g = base | local_to_global_low(tid, pd);
#pragma unroll
        for (int j = 0; j < 8; j++) tile[swz(j * NT + tid)] = load_nt<R>(&psi[g | pd.top[j]]);
    }
    __builtin_amdgcn_s_setprio(0);
}

// The pass's groups on the tile in LDS (at byte address `tb`): each group gathers its 8 members,
// runs its ops and scatters them back under its relabeling; every group ends at a barrier.
template <typename R, bool D4>
__device__ __forceinline__ void pass8_groups(unsigned char* tb, int tid, uint64_t base, const PassDesc& pd,
                                             const FOpDev<R>* __restrict__ ops, const unsigned char* __restrict__ runs) {
    for (int gi = 0; gi < pd.ngroups; gi++) {
        const int s0 = pd.gbits[gi][0], s1 = pd.gbits[gi][1], s2 = pd.gbits[gi][2];
        const int l0 = ins_bit(ins_bit(ins_bit(tid, s0), s1), s2);
        // slot of member m = swz(l0 | bits of m) = swz(l0) ^ swz(bits of m) (the swizzle is linear
        // over XOR): one XOR with a host-computed byte offset per member
        constexpr int ES = (int)sizeof(V2<R>);  // bytes per amplitude
        const int sl = swz(l0) * ES;
        V2<R> A[8], B[8];
#pragma unroll
        for (int m = 0; m < 8; m++) A[m] = *reinterpret_cast<const V2<R>*>(tb + (sl ^ pd.goff[gi][m]));
        const uint64_t G0 = ((pd.g0_mask >> gi) & 1u) ? base | local_to_global_bits(l0, pd) : 0;
        // two ops per iteration, a -> b and b -> a
        const int o_end = pd.gop[gi + 1];
        int oi = pd.gop[gi];
        if (oi < o_end) {
            for (;;) {
                r8_op<R, D4>(A, B, load_op8<R>(ops + oi), G0, runs);
                if (++oi >= o_end) {
#pragma unroll
                    for (int m = 0; m < 8; m++) A[m] = B[m];
                    break;
                }
                r8_op<R, D4>(B, A, load_op8<R>(ops + oi), G0, runs);
                if (++oi >= o_end) break;
            }
        }
        // the group's relabeling, applied by the scatter (register m goes to logical m ^ px)
        int sx = sl ^ pd.gpxoff[gi];
        if ((pd.pxlane_mask >> gi) & 1u) {
            const uint32_t pc = pd.gpxc[gi];
            int px = 0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const uint64_t msk = pd.gpxm[gi][j];
                const int bit = msk ? px_bit(G0, msk, pc >> j) : (int)((pc >> j) & 1u);
                px |= bit << pd.gbits[gi][j];
            }
            sx = sl ^ (swz(px) * ES);
        }
#pragma unroll
        for (int m = 0; m < 8; m++) *reinterpret_cast<V2<R>*>(tb + (sx ^ pd.goff[gi][m])) = A[m];
        __syncthreads();
    }
}

// One tile per workgroup. D4: the pass holds 4x4 ops.
template <typename R, bool D4>
__global__ void __launch_bounds__(1024) k_pass8(V2<R>* __restrict__ psi, const PassDesc pd,
                                                const FOpDev<R>* __restrict__ ops, const unsigned char* __restrict__ runs) {
    extern __shared__ unsigned char lds_raw[];
    V2<R>* tile = reinterpret_cast<V2<R>*>(lds_raw);
    const int tid = threadIdx.x, NT = blockDim.x;  // NT = 2^b / 8: one row of 8 per thread
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint64_t base = tile_base(blockIdx.x, pd);
    pass8_load<R>(psi, tile, base, tid, pd, NT, wave);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    pass8_groups<R, D4>(lds_raw, tid, base, pd, ops, runs);
    const uint64_t g_tid = base | local_to_global_low(tid, pd);
    if (pd.permuted) {  // the tile's qubits change places: element e is written from slot swz(Pinv(e))
        int pt = 0, rt = 0;
#pragma unroll
        for (int k = 0; k < FUSE_MAX_TILE_BITS - 3; k++)
            if (k < pd.b - 3 && ((tid >> k) & 1)) {
                pt ^= pd.pbit[k];
                rt ^= pd.rbit[k];
            }
        const int st = swz(pt);
        const uint64_t g_rt = base | local_to_global_low(rt, pd);
        __builtin_amdgcn_s_setprio(2);
#pragma unroll
        for (int j = 0; j < 8; j++) store_nt<R>(&psi[g_rt | pd.top[j]], tile[st ^ pd.pst[j]]);
        return;
    }
    __builtin_amdgcn_s_setprio(2);
#pragma unroll
    for (int j = 0; j < 8; j++) store_nt<R>(&psi[g_tid | pd.top[j]], tile[swz(j * NT + tid)]);
}

// ------------------------------------------------------------------------------------------------
// Appends the record of the diagonal run `run` of group G to `runs` (16-byte aligned) and returns
// the device op that applies it.
template <typename R>
static FOpDev<R> encode_run(const FGroup& G, int run, std::vector<unsigned char>& runs) {
    struct HT {
        int S;
        uint64_t cl, pm;
        uint32_t pc;
        cd v0, v1;
        double scale = 0.0;  // linear terms: the angle is scale * (G0 & cl)
    };
    cd f0[16];
    for (int S = 0; S < 16; S++) f0[S] = 1.0;
    std::vector<HT> terms;
    auto add = [&](int S, uint64_t cl, uint64_t pm, uint32_t pc, cd v0, cd v1) {
        if (cl == 0 && pm == 0)
            f0[S] *= (pc & 1u) ? v1 : v0;
        else if (!(v0 == cd(1.0, 0.0) && v1 == cd(1.0, 0.0)))
            terms.push_back(HT{S, cl, pm, pc & 1u, v0, v1});
    };
    for (const FOpHost& o : G.ops) {
        if (o.run != run) continue;
        const cd d0 = o.m[0], d1 = o.m[3];
        const int C = (int)o.cm_group;
        if (o.kind == F_PHASE) {
            add(C, o.cm_ext, 1ull << o.g, 0, d0, d1);
        } else if (o.flags & FL_TARGET_PX) {
            const uint32_t pc = (o.flags & FL_PX_CONST) ? 1u : 0u;
            add(C, o.cm_ext, o.pxmask, pc, d0, d1);
            add(C | (1 << o.g), o.cm_ext, o.pxmask, pc, d1 / d0, d0 / d1);
        } else {
            add(C, o.cm_ext, 0, 0, d0, d0);
            add(C | (1 << o.g), o.cm_ext, 0, 0, d1 / d0, d1 / d0);
        }
    }
    auto unit = [](cd z) { return std::abs(std::abs(z) - 1.0) < 1e-14; };
    // Unit-modulus terms of one subset that each depend on a single index bit k (no relabel
    // selector) with angle a_k = c * 2^k for a common c -- the controlled phases of a QFT -- add up
    // to c * (G0 & M), M the union of their bits: one linear term replaces them. c agrees to 1e-15
    // relative between the merged terms, so each angle moves by at most that much.
    {
        std::vector<HT> kept, linear;
        std::vector<char> used(terms.size(), 0);
        for (size_t a = 0; a < terms.size(); a++) {
            const HT& ta = terms[a];
            auto eligible = [&](const HT& t) {
                return t.pm == 0 && __builtin_popcountll(t.cl) == 1 && unit(t.v0) && unit(t.v1);
            };
            if (used[a] || !eligible(ta)) continue;
            auto coef = [&](const HT& t) {
                return std::arg(t.pc ? t.v1 : t.v0) / std::ldexp(1.0, __builtin_ctzll(t.cl));
            };
            const double c = coef(ta);
            uint64_t mask = 0;
            std::vector<size_t> members;
            for (size_t b = a; b < terms.size(); b++) {
                const HT& tb = terms[b];
                if (used[b] || tb.S != ta.S || !eligible(tb) || (mask & tb.cl)) continue;
                const double cb = coef(tb);
                if (std::abs(cb - c) <= 1e-15 * std::abs(c) || (c == 0.0 && cb == 0.0)) {
                    members.push_back(b);
                    mask |= tb.cl;
                }
            }
            if (members.size() < 2) continue;
            for (size_t b : members) used[b] = 1;
            HT lt = ta;
            lt.cl = mask;
            lt.scale = c;
            linear.push_back(lt);
        }
        for (size_t a = 0; a < terms.size(); a++)
            if (!used[a]) kept.push_back(terms[a]);
        kept.insert(kept.end(), linear.begin(), linear.end());
        terms.swap(kept);
    }
    std::stable_sort(terms.begin(), terms.end(), [](const HT& a, const HT& b) { return a.S < b.S; });
    RunHead<R> head{};
    uint32_t active = 0, angles = 0;
    for (int S = 0; S < 16; S++) {
        head.f0[2 * S] = (R)f0[S].real();
        head.f0[2 * S + 1] = (R)f0[S].imag();
        if (f0[S] != cd(1.0, 0.0)) active |= 1u << S;
    }
    size_t t = 0;
    for (int S = 0; S <= 16; S++) {
        head.tbeg[S] = (int32_t)t;
        while (S < 16 && t < terms.size() && terms[t].S == S) {
            active |= 1u << S;
            t++;
        }
    }
    const size_t off = (runs.size() + 15) / 16 * 16;
    runs.resize(off + sizeof(head) + terms.size() * sizeof(RunTerm<R>));
    std::memcpy(runs.data() + off, &head, sizeof(head));
    for (size_t k = 0; k < terms.size(); k++) {
        RunTerm<R> rt{};
        rt.cl = terms[k].cl;
        rt.pm = terms[k].pm;
        rt.pc = terms[k].pc;
        if (terms[k].scale != 0.0) {
            rt.angle = 2;
            rt.v[0] = (R)terms[k].scale;
            angles |= 1u << terms[k].S;
        } else if (unit(terms[k].v0) && unit(terms[k].v1)) {
            rt.angle = 1;
            rt.v[0] = (R)std::arg(terms[k].v0);
            rt.v[1] = (R)std::arg(terms[k].v1);
            angles |= 1u << terms[k].S;
        } else {
            rt.v[0] = (R)terms[k].v0.real();
            rt.v[1] = (R)terms[k].v0.imag();
            rt.v[2] = (R)terms[k].v1.real();
            rt.v[3] = (R)terms[k].v1.imag();
        }
        std::memcpy(runs.data() + off + sizeof(head) + k * sizeof(rt), &rt, sizeof(rt));
    }
    FOpDev<R> d{};
    d.kind = F_DIAGRUN;
    d.g = (int32_t)active;
    d.flags = angles;
    d.cm_ext = off;
    return d;
}

// Appends the pass's ops to h (diagonal-run records to runs) and, for two-state passes, the output
// slot of each marker column to slots (column c of the pass = its c-th marker).
template <typename R>
static PassDesc encode_pass(const State& s, const FPass& p, std::vector<FOpDev<R>>& h, std::vector<int>& slots,
                            std::vector<unsigned char>& runs) {
    const int b = (int)p.Q.size();
    int L = 0;
    while (L < b && p.Q[L] == L) L++;
    PassDesc pd{};
    pd.n = s.n;
    pd.b = b;
    pd.L = L;
    for (int k = 0; k < b; k++) pd.q[k] = p.Q[k];
    for (int j = 0; j < 8; j++) {
        pd.top[j] = 0;
        for (int i = 0; i < 3; i++)
            if ((j >> i) & 1) pd.top[j] |= 1ull << p.Q[b - 3 + i];
    }
    for (int j = 0; j < 16; j++) {
        pd.top16[j] = 0;
        for (int i = 0; i < 4; i++)
            if ((j >> i) & 1) pd.top16[j] |= 1ull << p.Q[b - 4 + i];
    }
    pd.ngroups = (int)p.groups.size();
    const int gbn = p.group_bits;
    if (pd.ngroups > FUSE_MAX_GROUPS || b > FUSE_MAX_TILE_BITS || b < gbn || b > s.n || gbn < 3 ||
        gbn > FUSE_GROUP_BITS)
        throw Error(QFVM_ERR_RUNTIME, "fused pass exceeds kernel limits");
    int local_of[64];
    for (int k = 0; k < 64; k++) local_of[k] = -1;
    for (int k = 0; k < b; k++) local_of[p.Q[k]] = k;
    const int first = (int)h.size();
    for (int gi = 0; gi < pd.ngroups; gi++) {
        const FGroup& G = p.groups[gi];
        if (G.stage) {
            pd.stage_mask |= 1u << gi;
            pd.gop[gi] = (int)h.size() - first;
            for (int j = 0; j < FUSE_GROUP_BITS; j++) pd.gbits[gi][j] = 0;
            for (const FOpHost& o : G.ops) {
                FOpDev<R> d{};
                d.kind = F_GRAD;
                d.g = o.g;
                if (o.g == GEN_Z) {
                    d.cm_group = (uint32_t)o.target;
                } else {
                    if (local_of[o.target] < 0) throw Error(QFVM_ERR_RUNTIME, "gradient marker outside the tile");
                    d.cm_group = (uint32_t)local_of[o.target];
                }
                d.cm_ext = o.cm_ext;
                d.flags = (uint32_t)slots.size();
                slots.push_back(o.slot);
                h.push_back(d);
            }
            continue;
        }
        for (int j = 0; j < gbn; j++) {
            if (G.S[j] < 0 || local_of[G.S[j]] < 0) throw Error(QFVM_ERR_RUNTIME, "group qubit outside the tile");
            pd.gbits[gi][j] = local_of[G.S[j]];
        }
        for (int j = gbn; j < FUSE_GROUP_BITS; j++) pd.gbits[gi][j] = 0;
        pd.gop[gi] = (int)h.size() - first;
        for (int j = 0; j < FUSE_GROUP_BITS; j++) pd.gpxm[gi][j] = j < gbn ? G.pxm[j] : 0;
        pd.gpxc[gi] = G.pxc;
        if (gbn == 3) {  // member slot offsets and the relabeling of the row-of-8 kernel
            const auto hswz = [](int l) { return l ^ ((l >> 4) & 7); };
            for (int m = 0; m < 8; m++) {
                int c = 0;
                for (int j = 0; j < 3; j++)
                    if ((m >> j) & 1) c |= 1 << pd.gbits[gi][j];
                pd.goff[gi][m] = hswz(c) * (int)sizeof(V2<R>);
            }
            int pxc = 0;
            bool lane_px = false;
            for (int j = 0; j < 3; j++) {
                if (G.pxm[j]) lane_px = true;
                if ((G.pxc >> j) & 1u) pxc |= 1 << pd.gbits[gi][j];
            }
            pd.gpxoff[gi] = hswz(pxc) * (int)sizeof(V2<R>);
            if (lane_px) {
                pd.pxlane_mask |= 1u << gi;
                pd.g0_mask |= 1u << gi;
            }
        }
        int done_run = 0;
        for (const FOpHost& o : G.ops) {
            if (o.kind == F_RELABEL) continue;  // folded into the parity masks
            if (o.run > 0) {                    // runs are numbered in op order
                if (o.run > done_run) {
                    h.push_back(encode_run<R>(G, o.run, runs));
                    if (gbn == 3) {
                        h.back().kind = K8_DR;
                        pd.g0_mask |= 1u << gi;
                    }
                    done_run = o.run;
                }
                continue;
            }
            FOpDev<R> d{};
            d.kind = o.kind;
            d.g = o.g;
            d.cm_group = o.cm_group;
            d.flags = o.flags;
            d.cm_ext = o.cm_ext;
            d.pxmask = o.pxmask;
            for (int k = 0; k < 4; k++) {
                d.m[2 * k] = (R)o.m[k].real();
                d.m[2 * k + 1] = (R)o.m[k].imag();
            }
            if (gbn == 3) {  // the row-of-8 kernel's kinds (K8Kind)
                const bool lane = (o.flags & FL_TARGET_PX) || o.cm_ext != 0;
                switch (o.kind) {
                    case F_DENSE: d.kind = lane ? K8_G2 : K8_U2; break;
                    case F_SWAP: d.kind = lane ? K8_G2 : K8_SW; break;
                    case F_REAL: d.kind = lane ? K8_G2 : K8_R2; break;
                    case F_DIAG: d.kind = K8_DG; break;
                    case F_PHASE:
                        d.kind = K8_DG;
                        d.g = -1 - o.g;
                        break;
                    case F_FLUSH: d.kind = K8_FL; break;
                    case F_DENSE4: {
                        // the 4x4 goes to the run buffer (16-byte aligned), its offset into pxmask
                        if (o.mx.size() != 16 || o.g < 0 || o.g > 2 || o.target < 0 || o.target > 2 || o.g == o.target ||
                            (o.cm_group & ((1u << o.g) | (1u << o.target))))
                            throw Error(QFVM_ERR_RUNTIME, "malformed 4x4 op");
                        d.kind = K8_D4;
                        d.g = o.g | (o.target << 4);
                        pd.has_d4 = 1;
                        const size_t off = (runs.size() + 15) / 16 * 16;
                        runs.resize(off + 32 * sizeof(R));
                        for (int k = 0; k < 16; k++) {
                            const R v[2] = {(R)o.mx[(size_t)k].real(), (R)o.mx[(size_t)k].imag()};
                            std::memcpy(runs.data() + off + (size_t)k * 2 * sizeof(R), v, sizeof(v));
                        }
                        d.pxmask = off;
                        break;
                    }
                    default: throw Error(QFVM_ERR_RUNTIME, "op kind not supported by the row-of-8 kernel");
                }
                // ops that read the lane's global index
                if (lane || o.kind == F_PHASE || o.kind == F_FLUSH) pd.g0_mask |= 1u << gi;
            } else if (o.kind == F_DENSE4) {
                throw Error(QFVM_ERR_RUNTIME, "4x4 ops run in row-of-8 passes only");
            }
            h.push_back(d);
        }
    }
    pd.gop[pd.ngroups] = (int)h.size() - first;
    if (!p.out_bit.empty()) {
        // out_bit[k]: where input local bit k goes; the store needs the inverse
        if ((int)p.out_bit.size() != b || gbn != 3) throw Error(QFVM_ERR_RUNTIME, "qubit permutation outside a row-of-8 pass");
        int from[FUSE_MAX_TILE_BITS];
        for (int k = 0; k < b; k++) from[k] = -1;
        for (int k = 0; k < b; k++) {
            const int o = p.out_bit[(size_t)k];
            if (o < 0 || o >= b || from[o] >= 0) throw Error(QFVM_ERR_RUNTIME, "tile qubit permutation is not a permutation");
            from[o] = k;
        }
        const auto hswz = [](int l) { return l ^ ((l >> 4) & 7); };
        pd.permuted = 1;
        // lane order of the stores: lane bits 0..3 take output bits (among 0..5, the wavefront's 64
        // elements) whose input bits land on distinct LDS slot bits 0..3 after the swizzle (input
        // bit f < 4 on slot bit f, 4..6 on f - 4), so 16 consecutive lanes read 16 different slots
        // modulo 16; the other bits keep their order
        int rho[FUSE_MAX_TILE_BITS];
        for (int k = 0; k < b - 3; k++) rho[k] = k;
        {
            const auto proj = [](int f) { return f < 4 ? f : (f < 7 ? f - 4 : -1); };
            bool used[6] = {false, false, false, false, false, false};
            int covered = 0;
            for (int i = 0; i < 4; i++) {
                int pick = -1;
                for (int t = 0; t < 6 && pick < 0; t++)
                    if (!used[t] && proj(from[t]) >= 0 && !((covered >> proj(from[t])) & 1)) pick = t;
                for (int t = 0; t < 6 && pick < 0; t++)
                    if (!used[t]) pick = t;
                rho[i] = pick;
                used[pick] = true;
                if (proj(from[pick]) >= 0) covered |= 1 << proj(from[pick]);
            }
            int i = 4;
            for (int t = 0; t < 6; t++)
                if (!used[t]) rho[i++] = t;
        }
        for (int t = 0; t < b - 3; t++) {
            pd.rbit[t] = 1 << rho[t];
            pd.pbit[t] = 1 << from[rho[t]];
        }
        for (int j = 0; j < 8; j++) {
            int v = 0;
            for (int i = 0; i < 3; i++)
                if ((j >> i) & 1) v |= 1 << from[b - 3 + i];
            pd.pst[j] = hswz(v);
        }
    }
    return pd;
}

// Raises a kernel's dynamic LDS limit to 160 KiB, once per (kernel, device); handles run on several
// host threads.
void allow_large_lds(const void* fn, int device) {
    static std::mutex lock;
    static std::vector<std::pair<const void*, int>> done;
    std::lock_guard<std::mutex> guard(lock);
    for (const auto& d : done)
        if (d.first == fn && d.second == device) return;
    QFVM_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    done.emplace_back(fn, device);
}

// Workgroups of a persistent k_fused_stream launch: as many as fit on the device at once.
static unsigned stream_grid(const State& s, const void* fn, size_t lds, unsigned threads, uint64_t tiles) {
    struct Known {
        const void* fn;
        uint64_t key;
        unsigned grid;
    };
    static std::vector<Known> known;  // (kernel, device, lds, threads) -> grid
    static std::mutex known_lock;     // handles run on several host threads
    std::lock_guard<std::mutex> guard(known_lock);
    const uint64_t key = ((uint64_t)s.device << 40) | ((uint64_t)lds << 12) | threads;
    unsigned grid = 0;
    for (const auto& kv : known)
        if (kv.fn == fn && kv.key == key) grid = kv.grid;
    if (grid == 0) {
        allow_large_lds(fn, s.device);
        int per_cu = 0, cus = 0;
        QFVM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)threads, lds));
        QFVM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s.device));
        grid = (unsigned)std::max(1, per_cu * cus);
        known.push_back({fn, key, grid});
    }
    return (unsigned)std::min<uint64_t>(grid, tiles);
}

template <typename R, int NS>
static void execute_typed(State& s, const std::vector<PlanItem>& plan, void* lam, double* grads_dev) {
    // encode every pass's ops (and marker slots) into one array, upload once, then launch the
    // items in order
    std::vector<FOpDev<R>> h;
    std::vector<PassDesc> descs;
    std::vector<size_t> first;
    std::vector<int> slots_all;
    std::vector<size_t> slot_first, slot_count;
    std::vector<unsigned char> runs;
    for (const PlanItem& it : plan) {
        if (!it.fused) continue;
        first.push_back(h.size());
        std::vector<int> slots;
        descs.push_back(encode_pass<R>(s, it.pass, h, slots, runs));
        if (NS == 1 && !slots.empty()) throw Error(QFVM_ERR_RUNTIME, "gradient marker in a one-state pass");
        slot_first.push_back(slots_all.size());
        slot_count.push_back(slots.size());
        slots_all.insert(slots_all.end(), slots.begin(), slots.end());
    }
    FOpDev<R>* dops = nullptr;
    int* dslots = nullptr;
    const unsigned char* druns = nullptr;
    if (!h.empty()) {
        const size_t obytes = h.size() * sizeof(FOpDev<R>);
        const size_t soff = (obytes + 255) / 256 * 256;
        const size_t roff = (soff + slots_all.size() * sizeof(int) + 255) / 256 * 256;
        const size_t bytes = roff + runs.size();
        // one-state plans and the reverse sweep upload through different slots, so the adjoint's
        // two plans can be in flight together
        const int slot = NS == 2 ? STAGE_PLAN2 : STAGE_PLAN;
        unsigned char* stage = s.stage_host(slot, bytes);
        unsigned char* dev = s.stage_dev(slot, bytes);
        std::memcpy(stage, h.data(), obytes);
        if (!slots_all.empty()) std::memcpy(stage + soff, slots_all.data(), slots_all.size() * sizeof(int));
        if (!runs.empty()) std::memcpy(stage + roff, runs.data(), runs.size());
        s.stage_upload(slot, bytes);
        dops = reinterpret_cast<FOpDev<R>*>(dev);
        dslots = reinterpret_cast<int*>(dev + soff);
        druns = dev + roff;
    }
    V2<R>* psi = reinterpret_cast<V2<R>*>(s.psi);
    V2<R>* psi1 = reinterpret_cast<V2<R>*>(lam);
    size_t k = 0;
    for (const PlanItem& it : plan) {
        if (!it.fused) {
            launch_gate(s, it.gate);
            if (NS == 2) {
                BufferScope other(s, lam);
                launch_gate(s, it.gate);
            }
            continue;
        }
        const PassDesc& pd = descs[k];
        const FOpDev<R>* ops = dops + first[k];
        const int cols = (int)slot_count[k];
        const int* slot_of = dslots + slot_first[k];
        k++;
        const size_t lds = NS * ((size_t)1 << pd.b) * sizeof(V2<R>) + (NS == 2 ? 64 * sizeof(double) : 0);
        if (lds > 160 * 1024) throw Error(QFVM_ERR_RUNTIME, "fused pass tiles exceed 160 KiB of LDS");
        if (lds > 65536) allow_large_lds(reinterpret_cast<const void*>(&k_fused_pass<R, NS>), s.device);
        const uint64_t tiles = 1ull << (s.n - pd.b);
        const unsigned threads = (unsigned)std::min(512, std::max(64, (1 << pd.b) >> 4));  // one row each
        if (it.pass.group_bits == 3) {
            if constexpr (NS == 1) {
                // the load swizzle needs at least 128 threads (slot bits 0-6 from the thread index)
                if (pd.b < 10 || pd.b > 13) throw Error(QFVM_ERR_RUNTIME, "row-of-8 tiles hold 2^10 to 2^13 amplitudes");
                const size_t lds8 = ((size_t)1 << pd.b) * sizeof(V2<R>);
                const unsigned threads8 = (1u << pd.b) >> 3;
                KernelScope ks(s, "fused_pass", 2.0 * (double)s.size * sizeof(V2<R>));
                if (lds8 > 160 * 1024) throw Error(QFVM_ERR_RUNTIME, "row-of-8 tile exceeds 160 KiB of LDS");
                const auto fn = pd.has_d4 ? k_pass8<R, true> : k_pass8<R, false>;
                if (lds8 > 65536) allow_large_lds(reinterpret_cast<const void*>(fn), s.device);
                hipLaunchKernelGGL(fn, dim3((unsigned)tiles), dim3(threads8), lds8, s.stream, psi, pd, ops, druns);
                QFVM_LAUNCHED();
                continue;
            } else {
                throw Error(QFVM_ERR_RUNTIME, "row-of-8 passes are one-state passes");
            }
        }
        // per-tile marker partials live in the scratch buffer (fetched per pass: a dense gate
        // between passes stages its matrix there)
        double* partials = cols > 0 ? reinterpret_cast<double*>(s.ensure_scratch(cols * tiles * sizeof(double))) : nullptr;
        // split rows run complex128 passes and complex64 two-state passes (complex64 one-state
        // passes measured slower on them: 30-qubit circuit at 2^12 tiles 822 vs 744 ms)
        if constexpr (sizeof(R) == 8 || NS == 2) {
            const int hb_max = NS == 1 ? 13 : 12;  // at most 1024 / 512 threads
            if (s.split_rows && pd.b >= 10 && pd.b <= hb_max) {
                const size_t lds_h = lds + ((size_t)1 << (pd.b - pd.L)) * sizeof(uint64_t);
                if (lds_h > 160 * 1024) throw Error(QFVM_ERR_RUNTIME, "split-row tile exceeds 160 KiB of LDS");
                if (lds_h > 65536) allow_large_lds(reinterpret_cast<const void*>(&k_fused_half<R, NS>), s.device);
                {
                    KernelScope ks(s, NS == 1 ? "fused_pass" : "fused_pass2", 2.0 * NS * (double)s.size * sizeof(V2<R>));
                    hipLaunchKernelGGL((k_fused_half<R, NS>), dim3((unsigned)tiles), dim3((1u << pd.b) >> 3), lds_h,
                                       s.stream, psi, psi1, pd, ops, druns, partials, tiles);
                    QFVM_LAUNCHED();
                }
                if (cols > 0) launch_column_sums(s, partials, tiles, cols, slot_of, grads_dev);
                continue;
            }
        }
        if constexpr (NS == 1 && sizeof(R) == 8) {
            if (s.stream_tiles && pd.b >= 11 && pd.b <= 13 && pd.ngroups >= 1 && pd.ngroups <= FUSE_MAX_GROUPS) {
                const unsigned grid = stream_grid(s, reinterpret_cast<const void*>(&k_fused_stream), lds, threads, tiles);
                KernelScope ks(s, "fused_pass", 2.0 * (double)s.size * sizeof(V2<R>));
                hipLaunchKernelGGL(k_fused_stream, dim3(grid), dim3(threads), lds, s.stream,
                                   reinterpret_cast<double2*>(psi), pd, reinterpret_cast<const FOpDev<double>*>(ops),
                                   druns, tiles, (uint32_t)s.stagger, (uint32_t)s.stagger_shift);
                QFVM_LAUNCHED();
                continue;
            }
        }
        {
            KernelScope ks(s, NS == 1 ? "fused_pass" : "fused_pass2", 2.0 * NS * (double)s.size * sizeof(V2<R>));
            hipLaunchKernelGGL((k_fused_pass<R, NS>), dim3((unsigned)tiles), dim3(threads), lds, s.stream, psi, psi1,
                               pd, ops, druns, partials, tiles);
            QFVM_LAUNCHED();
        }
        if (cols > 0) launch_column_sums(s, partials, tiles, cols, slot_of, grads_dev);
    }
}

void execute_plan(State& s, const std::vector<PlanItem>& plan, void* lam, double* grads_dev) {
    if (lam) {
        if (s.dtype == QFVM_C128)
            execute_typed<double, 2>(s, plan, lam, grads_dev);
        else
            execute_typed<float, 2>(s, plan, lam, grads_dev);
    } else {
        if (s.dtype == QFVM_C128)
            execute_typed<double, 1>(s, plan, nullptr, nullptr);
        else
            execute_typed<float, 1>(s, plan, nullptr, nullptr);
    }
}

}  // namespace qfvm
