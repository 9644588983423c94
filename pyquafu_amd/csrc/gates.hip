// Per-gate kernels: one launch applies one gate to the whole (control-satisfied) state.
//
// These are the unfused path and the fallback for gates the fused pass kernel does not take
// (dense k-target matrices). Work item j in [0, 2^(n - #positions)) is expanded into the pair
// (i0, i1 = i0 | 2^t) by inserting zero bits at the sorted target/control positions and OR-ing in
// the control mask -- the index map of the reference's apply_one_targe_gate_* kernels
// (src/qfvm/statevector.hpp:446-499), evaluated in 64-bit registers instead of through
// std::function. Algorithmic bytes per launch: 2 * s * 2^(n-c) (read + write of every amplitude
// in the control-satisfied subspace), s = 16 (complex128) or 8 (complex64); a diagonal with
// d0 == 1 touches only the half with the target bit set.
#include <algorithm>

#include "common.h"

namespace qfvm {

struct InsertSpec {
    int n;
    int pos[40];
};

__device__ __forceinline__ uint64_t insert_zeros(uint64_t j, const InsertSpec& sp) {
#pragma unroll 1
    for (int k = 0; k < sp.n; k++) {
        const uint64_t low = (1ull << sp.pos[k]) - 1;
        j = (j & low) | ((j & ~low) << 1);
    }
    return j;
}

template <typename R>
struct Coef {
    R m[8];  // m00.re m00.im m01.re m01.im m10.re m10.im m11.re m11.im
};

enum { K_DENSE = 0, K_REAL = 1, K_DIAG = 2, K_DIAG1 = 3, K_SWAP = 4 };

template <typename R, int KIND>
__global__ void __launch_bounds__(256) k_gate1(V2<R>* __restrict__ psi, uint64_t npairs, InsertSpec sp,
                                               uint64_t cmask, uint64_t off, Coef<R> c) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < npairs; j += stride) {
        const uint64_t i0 = insert_zeros(j, sp) | cmask;
        const uint64_t i1 = i0 | off;
        if constexpr (KIND == K_DIAG1) {
            // diag(1, d1): only the target-set half changes
            V2<R> b = psi[i1];
            psi[i1] = cmul<R>(mk<R>(c.m[6], c.m[7]), b);
        } else {
            V2<R> a = psi[i0];
            V2<R> b = psi[i1];
            if constexpr (KIND == K_SWAP) {
                psi[i0] = b;
                psi[i1] = a;
            } else if constexpr (KIND == K_DIAG) {
                psi[i0] = cmul<R>(mk<R>(c.m[0], c.m[1]), a);
                psi[i1] = cmul<R>(mk<R>(c.m[6], c.m[7]), b);
            } else if constexpr (KIND == K_REAL) {
                psi[i0] = mk<R>(c.m[0] * a.x + c.m[2] * b.x, c.m[0] * a.y + c.m[2] * b.y);
                psi[i1] = mk<R>(c.m[4] * a.x + c.m[6] * b.x, c.m[4] * a.y + c.m[6] * b.y);
            } else {
                psi[i0] = cmul2<R>(mk<R>(c.m[0], c.m[1]), a, mk<R>(c.m[2], c.m[3]), b);
                psi[i1] = cmul2<R>(mk<R>(c.m[4], c.m[5]), a, mk<R>(c.m[6], c.m[7]), b);
            }
        }
    }
}

// Dense k-target matrix, 2 <= K <= 5: the 2^K amplitudes of one group live in registers.
template <int K>
struct TMask {
    uint64_t m[1 << K];
};

template <typename R, int K>
__global__ void __launch_bounds__(256) k_multi_reg(V2<R>* __restrict__ psi, uint64_t ngroups, InsertSpec sp,
                                                   uint64_t cmask, TMask<K> tm, const V2<R>* __restrict__ mat) {
    constexpr int D = 1 << K;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < ngroups; j += stride) {
        const uint64_t base = insert_zeros(j, sp) | cmask;
        V2<R> v[D];
#pragma unroll
        for (int m = 0; m < D; m++) v[m] = psi[base | tm.m[m]];
#pragma unroll
        for (int r = 0; r < D; r++) {
            R ax = 0, ay = 0;
#pragma unroll
            for (int q = 0; q < D; q++) {
                const V2<R> w = mat[r * D + q];
                ax += w.x * v[q].x - w.y * v[q].y;
                ay += w.x * v[q].y + w.y * v[q].x;
            }
            psi[base | tm.m[r]] = mk<R>(ax, ay);
        }
    }
}

// Dense 2^K x 2^K complex matrix on the matrix cores (K = 4, 5, 6): complex128 on
// v_mfma_f64_16x16x4f64, complex64 on v_mfma_f32_16x16x4f32. The gate is a real GEMM: with X the
// tile's amplitudes as columns (rows 0..D-1 real parts, D..2D-1 imaginary parts) and
// A = [[Ur, -Ui], [Ui, Ur]], Y = A X. One workgroup (4 wavefronts) takes NG = 64 groups: gather
// into LDS, the wavefronts multiply (A held in registers, X from LDS), results back through LDS,
// scatter. A dense gate moves 2 x s B per amplitude and costs 8 x 2^K flops per amplitude: at
// complex128 4, 8 and 16 flop/B for K = 4, 5, 6 against the FP64 ridge of ~9.8 flop/B, so from
// K = 5 on the arithmetic decides and the matrix cores carry it.
// Operand layouts of both instructions (one 16x16 output block per wavefront):
//   A (16 x 4): lane l supplies A[l % 16][l / 16];  B (4 x 16): lane l supplies B[l / 16][l % 16];
//   D (16 x 16): lane l holds, in accumulator element v = 0..3, D[4 v + l / 16][l % 16] for the f64
//   instruction (measured on gfx950 with one-hot operands) and D[4 (l / 16) + v][l % 16] for f32.
template <typename R>
struct MfmaOp;
template <>
struct MfmaOp<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int v, int kq) { return 4 * v + kq; }
};
template <>
struct MfmaOp<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int v, int kq) { return 4 * kq + v; }
};
constexpr int MFMA_NG = 64;  // groups per workgroup tile
// LDS row pitch in elements: odd, so that the member-fast gather (consecutive lanes on consecutive
// rows) hits distinct banks; 79 keeps two K = 6 complex128 workgroups within the 160 KiB of a CU
constexpr int mfma_pitch(int K) { return K == 6 ? 79 : 81; }
// LDS bytes of one workgroup: the M x P tile, plus (complex64) the group-base and member-offset
// tables, which complex128 keeps in the row padding
template <typename R, int K>
constexpr size_t mfma_lds() {
    return (size_t)(2 << K) * mfma_pitch(K) * sizeof(R) + (sizeof(R) == 8 ? 0 : (MFMA_NG + (1 << K)) * sizeof(uint64_t));
}

// Persistent: each workgroup walks tiles blockIdx.x, blockIdx.x + gridDim.x, ... For K >= 5 wavefront
// w owns the row blocks w, w + 4, .. of A (M / 64 of them) and all four 16-column blocks; for K = 4
// (two row blocks) it owns both row blocks and column block w. Its A fragments stay in registers for
// the whole launch, so per tile only X comes from LDS (one fragment per column block and k step).
template <typename R, int K>
__global__ void __launch_bounds__(256, K == 6 ? 2 : 4) k_multi_mfma(V2<R>* __restrict__ psi, uint64_t ngroups,
                                                                    InsertSpec sp, uint64_t cmask, TMask<K> tm,
                                                                    const R* __restrict__ at, int member_fast,
                                                                    uint64_t tiles) {
    typedef typename MfmaOp<R>::v4 v4;
    constexpr int D = 1 << K, M = 2 * D, P = mfma_pitch(K), KS = M / 4;
    constexpr bool ROWSPLIT = M >= 64;
    constexpr int NR = ROWSPLIT ? M / 64 : M / 16, NC = ROWSPLIT ? 4 : 1;  // blocks per wavefront
    // complex128: M x P doubles; the pitch padding (PAD slots per row) of the first rows holds the
    // tile's 64 group bases, that of the next rows the D member offsets. complex64: the tables
    // follow the M x P floats.
    constexpr int PAD = P - MFMA_NG, GROWS = (MFMA_NG + PAD - 1) / PAD;
    static_assert(GROWS + (D + PAD - 1) / PAD <= M, "lookup tables exceed the padding");
    extern __shared__ unsigned char lds_raw[];
    R* const xs = reinterpret_cast<R*>(lds_raw);
    uint64_t* const xw = reinterpret_cast<uint64_t*>(lds_raw);
    uint64_t* const tw = reinterpret_cast<uint64_t*>(lds_raw + (size_t)M * P * sizeof(R));
    auto gbase = [&](int c) -> uint64_t& {
        if constexpr (sizeof(R) == 8)
            return xw[(c / PAD) * P + MFMA_NG + c % PAD];
        else
            return tw[c];
    };
    auto moff = [&](int m) -> uint64_t& {
        if constexpr (sizeof(R) == 8)
            return xw[(GROWS + m / PAD) * P + MFMA_NG + m % PAD];
        else
            return tw[MFMA_NG + m];
    };
    constexpr int ITER = D * MFMA_NG / 256;  // elements per thread and tile (4, 8 or 16)
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, kq = lane >> 4, l16 = lane & 15;
    auto rowb = [&](int j) { return ROWSPLIT ? wave + 4 * j : j; };
    auto colb = [&](int c) { return ROWSPLIT ? c : wave; };
    R a[NR][KS];  // A[rowb(j) 16 + l16][4 ks + kq]
#pragma unroll
    for (int j = 0; j < NR; j++)
#pragma unroll
        for (int ks = 0; ks < KS; ks++) a[j][ks] = at[(size_t)(4 * ks + kq) * M + rowb(j) * 16 + l16];
    // element e of a tile: member m, column c; member-fast order keeps consecutive lanes on
    // consecutive addresses when a target is among the lowest qubits, column-fast otherwise
    auto where = [&](int e, int& m, int& c) {
        if (member_fast) {
            m = e & (D - 1);
            c = e >> K;
        } else {
            c = e & (MFMA_NG - 1);
            m = e / MFMA_NG;
        }
    };
    if (tid < D) moff(tid) = tm.m[tid];  // read before the first tile's barrier
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint64_t g0 = t * MFMA_NG;
        if (tid < MFMA_NG) {
            const uint64_t g = g0 + (uint64_t)tid;
            gbase(tid) = g < ngroups ? (insert_zeros(g, sp) | cmask) : ~0ull;
        }
        __syncthreads();
        // loads issued CH at a time, then written to LDS (K = 6 holds 128 registers of A, so fewer)
        constexpr int CH = K == 6 ? 4 : (ITER < 8 ? ITER : 8);
#pragma unroll
        for (int i0 = 0; i0 < ITER; i0 += CH) {
            V2<R> v[CH];
#pragma unroll
            for (int i = 0; i < CH; i++) {
                int m, c;
                where(tid + 256 * (i0 + i), m, c);
                const uint64_t gb = gbase(c);
                v[i] = gb != ~0ull ? psi[gb | moff(m)] : mk<R>(0, 0);
            }
#pragma unroll
            for (int i = 0; i < CH; i++) {
                int m, c;
                where(tid + 256 * (i0 + i), m, c);
                xs[m * P + c] = v[i].x;
                xs[(m + D) * P + c] = v[i].y;
            }
        }
        __syncthreads();
        v4 acc[NR][NC];
#pragma unroll
        for (int j = 0; j < NR; j++)
#pragma unroll
            for (int ct = 0; ct < NC; ct++) acc[j][ct] = v4{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            R b[NC];
#pragma unroll
            for (int ct = 0; ct < NC; ct++) b[ct] = xs[(4 * ks + kq) * P + colb(ct) * 16 + l16];
#pragma unroll
            for (int j = 0; j < NR; j++)
#pragma unroll
                for (int ct = 0; ct < NC; ct++) acc[j][ct] = MfmaOp<R>::mma(a[j][ks], b[ct], acc[j][ct]);
        }
        __syncthreads();  // every wavefront has read X
#pragma unroll
        for (int j = 0; j < NR; j++)
#pragma unroll
            for (int ct = 0; ct < NC; ct++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    xs[(rowb(j) * 16 + MfmaOp<R>::row(v, kq)) * P + colb(ct) * 16 + l16] = acc[j][ct][v];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < ITER; i++) {
            int m, c;
            where(tid + 256 * i, m, c);
            const uint64_t gb = gbase(c);
            if (gb != ~0ull) psi[gb | moff(m)] = mk<R>(xs[m * P + c], xs[(m + D) * P + c]);
        }
        __syncthreads();  // the next tile rewrites xs and gbase
    }
}

struct TPos {
    int k;
    int t[16];
};

// Dense gates on K = 7 or 8 targets on the matrix cores. The real embedding A (M x M, M = 2^(K+1):
// 512 KiB / 2 MiB at complex128) no longer fits in registers or LDS, so it streams from memory (L2
// holds it: every workgroup reads the same matrix) while the workgroup's tile X of NG groups sits
// in LDS. Wavefront w owns the 16-row blocks w, w + 4, .. of Y = A X and all NG / 16 column blocks;
// per k step it loads its A fragments for the next step (one step ahead) and takes X from LDS.
// NG is chosen so a tile is ~68 KiB: two workgroups per CU.
template <typename R, int K>
constexpr int wide_ng() { return (sizeof(R) == 8 ? 32 : 64) >> (K - 7); }
template <typename R, int K>
constexpr size_t wide_lds() {
    return (size_t)(2 << K) * (wide_ng<R, K>() + 1) * sizeof(R) + (wide_ng<R, K>() + (1 << K)) * sizeof(uint64_t);
}

template <typename R, int K>
__global__ void __launch_bounds__(256, 2) k_multi_wide(V2<R>* __restrict__ psi, uint64_t ngroups, InsertSpec sp,
                                                       uint64_t cmask, TPos tp, const R* __restrict__ at,
                                                       int member_fast, uint64_t tiles) {
    typedef typename MfmaOp<R>::v4 v4;
    constexpr int D = 1 << K, M = 2 * D, NG = wide_ng<R, K>(), P = NG + 1, KS = M / 4;
    constexpr int NR = M / 64, NC = NG / 16;  // row blocks per wavefront, column blocks
    constexpr int ITER = D * NG / 256;         // elements per thread and tile
    extern __shared__ unsigned char lds_raw[];
    R* const xs = reinterpret_cast<R*>(lds_raw);
    uint64_t* const gbase = reinterpret_cast<uint64_t*>(lds_raw + (size_t)M * P * sizeof(R));
    uint64_t* const moff = gbase + NG;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, kq = lane >> 4, l16 = lane & 15;
    auto where = [&](int e, int& m, int& c) {
        if (member_fast) {
            m = e & (D - 1);
            c = e >> K;
        } else {
            c = e & (NG - 1);
            m = e / NG;
        }
    };
    for (int m = tid; m < D; m += 256) {
        uint64_t o = 0;
        for (int b = 0; b < K; b++) o |= (uint64_t)((m >> b) & 1) << tp.t[b];
        moff[m] = o;
    }
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint64_t g0 = t * NG;
        if (tid < NG) {
            const uint64_t g = g0 + (uint64_t)tid;
            gbase[tid] = g < ngroups ? (insert_zeros(g, sp) | cmask) : ~0ull;
        }
        __syncthreads();
        constexpr int CH = 8;
#pragma unroll
        for (int i0 = 0; i0 < ITER; i0 += CH) {
            V2<R> v[CH];
#pragma unroll
            for (int i = 0; i < CH; i++) {
                int m, c;
                where(tid + 256 * (i0 + i), m, c);
                const uint64_t gb = gbase[c];
                v[i] = gb != ~0ull ? psi[gb | moff[m]] : mk<R>(0, 0);
            }
#pragma unroll
            for (int i = 0; i < CH; i++) {
                int m, c;
                where(tid + 256 * (i0 + i), m, c);
                xs[m * P + c] = v[i].x;
                xs[(m + D) * P + c] = v[i].y;
            }
        }
        __syncthreads();
        v4 acc[NR][NC];
#pragma unroll
        for (int j = 0; j < NR; j++)
#pragma unroll
            for (int ct = 0; ct < NC; ct++) acc[j][ct] = v4{0, 0, 0, 0};
        // A[(wave + 4 j) 16 + l16][4 ks + kq] = at[(4 ks + kq) M + (wave + 4 j) 16 + l16]
        const R* ap = at + (size_t)kq * M + wave * 16 + l16;
        R a[NR], an[NR];
#pragma unroll
        for (int j = 0; j < NR; j++) a[j] = ap[j * 64];
#pragma unroll 2
        for (int ks = 0; ks < KS; ks++) {
            if (ks + 1 < KS) {
#pragma unroll
                for (int j = 0; j < NR; j++) an[j] = ap[(size_t)(ks + 1) * 4 * M + j * 64];
            }
            R b[NC];
#pragma unroll
            for (int ct = 0; ct < NC; ct++) b[ct] = xs[(4 * ks + kq) * P + ct * 16 + l16];
#pragma unroll
            for (int j = 0; j < NR; j++)
#pragma unroll
                for (int ct = 0; ct < NC; ct++) acc[j][ct] = MfmaOp<R>::mma(a[j], b[ct], acc[j][ct]);
#pragma unroll
            for (int j = 0; j < NR; j++) a[j] = an[j];
        }
        __syncthreads();  // every wavefront has read X
#pragma unroll
        for (int j = 0; j < NR; j++)
#pragma unroll
            for (int ct = 0; ct < NC; ct++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    xs[((wave + 4 * j) * 16 + MfmaOp<R>::row(v, kq)) * P + ct * 16 + l16] = acc[j][ct][v];
        __syncthreads();
#pragma unroll 4
        for (int i = 0; i < ITER; i++) {
            int m, c;
            where(tid + 256 * i, m, c);
            const uint64_t gb = gbase[c];
            if (gb != ~0ull) psi[gb | moff[m]] = mk<R>(xs[m * P + c], xs[(m + D) * P + c]);
        }
        __syncthreads();  // the next tile rewrites xs and gbase
    }
}

// Dense k-target matrix, K >= 6: one block per group; the group is staged in LDS.

template <typename R>
__global__ void __launch_bounds__(256) k_multi_lds(V2<R>* __restrict__ psi, uint64_t ngroups, InsertSpec sp,
                                                   uint64_t cmask, TPos tp, const V2<R>* __restrict__ mat) {
    extern __shared__ unsigned char lds_raw[];
    V2<R>* v = reinterpret_cast<V2<R>*>(lds_raw);
    const int D = 1 << tp.k;
    for (uint64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const uint64_t base = insert_zeros(g, sp) | cmask;
        for (int m = threadIdx.x; m < D; m += blockDim.x) {
            uint64_t off = 0;
            for (int b = 0; b < tp.k; b++) off |= (uint64_t)((m >> b) & 1) << tp.t[b];
            v[m] = psi[base | off];
        }
        __syncthreads();
        for (int r = threadIdx.x; r < D; r += blockDim.x) {
            R ax = 0, ay = 0;
            const V2<R>* row = mat + (uint64_t)r * D;
            for (int q = 0; q < D; q++) {
                const V2<R> w = row[q];
                ax += w.x * v[q].x - w.y * v[q].y;
                ay += w.x * v[q].y + w.y * v[q].x;
            }
            uint64_t off = 0;
            for (int b = 0; b < tp.k; b++) off |= (uint64_t)((r >> b) & 1) << tp.t[b];
            psi[base | off] = mk<R>(ax, ay);
        }
        __syncthreads();
    }
}

__global__ void k_fill_basis_c128(double2* psi, uint64_t size, uint64_t index) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride)
        psi[i] = make_double2(i == index ? 1.0 : 0.0, 0.0);
}

// zero fill of 8-byte words (a kernel rather than hipMemsetAsync, so that captured graphs hold
// only kernel and copy nodes)
__global__ void k_zero_words(uint64_t* p, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

__global__ void k_fill_basis_c64(float2* psi, uint64_t size, uint64_t index) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride)
        psi[i] = make_float2(i == index ? 1.0f : 0.0f, 0.0f);
}

__global__ void k_widen(const float2* __restrict__ src, double2* __restrict__ dst, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = make_double2((double)src[i].x, (double)src[i].y);
}

__global__ void k_narrow(const double2* __restrict__ src, float2* __restrict__ dst, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = make_float2((float)src[i].x, (float)src[i].y);
}

// ------------------------------------------------------------------------------------------------
static InsertSpec make_spec(std::vector<int> positions) {
    std::sort(positions.begin(), positions.end());
    InsertSpec sp{};
    if (positions.size() > 40) throw Error(QFVM_ERR_VALUE, "too many gate positions");
    sp.n = (int)positions.size();
    for (size_t k = 0; k < positions.size(); k++) sp.pos[k] = positions[k];
    return sp;
}

template <typename R>
static Coef<R> make_coef(const Gate& g) {
    Coef<R> c{};
    for (int k = 0; k < 4; k++) {
        c.m[2 * k] = (R)g.m[k].real();
        c.m[2 * k + 1] = (R)g.m[k].imag();
    }
    if (g.kind == G_DIAG) {  // diag stored in m[0], m[1] -> positions 00 and 11
        c.m[0] = (R)g.m[0].real();
        c.m[1] = (R)g.m[0].imag();
        c.m[6] = (R)g.m[1].real();
        c.m[7] = (R)g.m[1].imag();
        c.m[2] = c.m[3] = c.m[4] = c.m[5] = 0;
    }
    return c;
}

template <typename R>
static void gate1_typed(State& s, const Gate& g) {
    std::vector<int> pos = g.ctrls;
    pos.push_back(g.target);
    const InsertSpec sp = make_spec(pos);
    const uint64_t npairs = s.size >> pos.size();
    const uint64_t off = 1ull << g.target;
    const Coef<R> c = make_coef<R>(g);
    V2<R>* psi = reinterpret_cast<V2<R>*>(s.psi);
    const unsigned block = 256;
    const unsigned grid = grid_for(npairs, block);
    const bool phase_only = g.kind == G_DIAG && g.m[0] == cd(1.0, 0.0);
    static const char* names[] = {"gate1_dense", "gate1_real", "gate1_diag", "gate1_swap"};
    KernelScope ks(s, phase_only ? "gate1_phase" : names[g.kind],
                   (double)npairs * (phase_only ? 2.0 : 4.0) * sizeof(V2<R>));
    switch (g.kind) {
        case G_SWAP:
            hipLaunchKernelGGL((k_gate1<R, K_SWAP>), dim3(grid), dim3(block), 0, s.stream, psi, npairs, sp,
                               g.cmask, off, c);
            break;
        case G_DIAG:
            if (phase_only)
                hipLaunchKernelGGL((k_gate1<R, K_DIAG1>), dim3(grid), dim3(block), 0, s.stream, psi, npairs, sp,
                                   g.cmask, off, c);
            else
                hipLaunchKernelGGL((k_gate1<R, K_DIAG>), dim3(grid), dim3(block), 0, s.stream, psi, npairs, sp,
                                   g.cmask, off, c);
            break;
        case G_REAL:
            hipLaunchKernelGGL((k_gate1<R, K_REAL>), dim3(grid), dim3(block), 0, s.stream, psi, npairs, sp,
                               g.cmask, off, c);
            break;
        default:
            hipLaunchKernelGGL((k_gate1<R, K_DENSE>), dim3(grid), dim3(block), 0, s.stream, psi, npairs, sp,
                               g.cmask, off, c);
            break;
    }
    QFVM_LAUNCHED();
}

template <typename R, int K>
static void multi_reg_launch(State& s, const Gate& g, const InsertSpec& sp, uint64_t ngroups, const V2<R>* dmat) {
    TMask<K> tm{};
    for (int m = 0; m < (1 << K); m++) {
        uint64_t o = 0;
        for (int b = 0; b < K; b++)
            if ((m >> b) & 1) o |= 1ull << g.targets[b];
        tm.m[m] = o;
    }
    const unsigned block = 256;
    KernelScope ks(s, "multi_reg", (double)ngroups * (1 << K) * 2.0 * sizeof(V2<R>));
    hipLaunchKernelGGL((k_multi_reg<R, K>), dim3(grid_for(ngroups, block)), dim3(block), 0, s.stream,
                       reinterpret_cast<V2<R>*>(s.psi), ngroups, sp, g.cmask, tm, dmat);
    QFVM_LAUNCHED();
}

// A^T of the real embedding of the gate's matrix, at[k * M + r] = A[r][k] (A = [[Ur, -Ui], [Ui, Ur]]),
// in the state's precision
template <typename R>
static std::vector<R> real_embedding_t(const Gate& g, int D) {
    const size_t M = 2 * (size_t)D;
    std::vector<R> at(M * M);
    for (int r = 0; r < D; r++)
        for (int q = 0; q < D; q++) {
            const cd u = g.mk[(size_t)r * D + q];
            at[(size_t)q * M + r] = (R)u.real();
            at[(size_t)(q + D) * M + r] = (R)-u.imag();
            at[(size_t)q * M + r + D] = (R)u.imag();
            at[(size_t)(q + D) * M + r + D] = (R)u.real();
        }
    return at;
}

// gather order of the matrix-core kernels: member-fast when a target is below every other free
// qubit (consecutive members are then the closest addresses), column-fast otherwise
static int member_fast_order(const State& s, const Gate& g) {
    int lowest = 64;
    for (int t : g.targets) lowest = std::min(lowest, t);
    int lowest_free = 0;  // lowest qubit outside the gate
    while (lowest_free < s.n && (((g.cmask >> lowest_free) & 1) || std::find(g.targets.begin(), g.targets.end(), lowest_free) != g.targets.end()))
        lowest_free++;
    return lowest < lowest_free ? 1 : 0;
}

template <typename R, int K>
static void multi_wide_launch(State& s, const Gate& g, const InsertSpec& sp, uint64_t ngroups) {
    constexpr int D = 1 << K;
    TPos tp{};
    tp.k = K;
    for (int b = 0; b < K; b++) tp.t[b] = g.targets[b];
    const std::vector<R> at = real_embedding_t<R>(g, D);
    R* dat = reinterpret_cast<R*>(s.ensure_scratch(at.size() * sizeof(R)));
    QFVM_HIP(hipMemcpyAsync(dat, at.data(), at.size() * sizeof(R), hipMemcpyHostToDevice, s.stream));
    const size_t lds = wide_lds<R, K>();
    const void* fn = reinterpret_cast<const void*>(&k_multi_wide<R, K>);
    if (lds > 65536) allow_large_lds(fn, s.device);
    const uint64_t tiles = (ngroups + wide_ng<R, K>() - 1) / wide_ng<R, K>();
    int cus = 0, per_cu = 0;
    QFVM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s.device));
    QFVM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, lds));
    const unsigned grid = (unsigned)std::min<uint64_t>(tiles, (uint64_t)std::max(cus, 1) * (uint64_t)std::max(per_cu, 1));
    KernelScope ks(s, "multi_mfma", (double)ngroups * D * 2.0 * sizeof(V2<R>));
    hipLaunchKernelGGL((k_multi_wide<R, K>), dim3(grid), dim3(256), lds, s.stream, reinterpret_cast<V2<R>*>(s.psi),
                       ngroups, sp, g.cmask, tp, dat, member_fast_order(s, g), tiles);
    QFVM_LAUNCHED();
    QFVM_HIP(hipStreamSynchronize(s.stream));  // the host matrix copy must outlive the async copy
}

template <typename R, int K>
static void multi_mfma_launch(State& s, const Gate& g, const InsertSpec& sp, uint64_t ngroups) {
    constexpr int D = 1 << K;
    TMask<K> tm{};
    for (int m = 0; m < D; m++) {
        uint64_t o = 0;
        for (int b = 0; b < K; b++)
            if ((m >> b) & 1) o |= 1ull << g.targets[b];
        tm.m[m] = o;
    }
    // A^T of the real embedding, staged through scratch
    const std::vector<R> at = real_embedding_t<R>(g, D);
    R* dat = reinterpret_cast<R*>(s.ensure_scratch(at.size() * sizeof(R)));
    QFVM_HIP(hipMemcpyAsync(dat, at.data(), at.size() * sizeof(R), hipMemcpyHostToDevice, s.stream));
    const int member_fast = member_fast_order(s, g);
    const size_t lds = mfma_lds<R, K>();
    const void* fn = reinterpret_cast<const void*>(&k_multi_mfma<R, K>);
    if (lds > 65536) allow_large_lds(fn, s.device);
    const uint64_t tiles = (ngroups + MFMA_NG - 1) / MFMA_NG;
    int cus = 0, per_cu = 0;  // resident workgroups: registers and LDS both count
    QFVM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s.device));
    QFVM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, lds));
    const unsigned grid = (unsigned)std::min<uint64_t>(tiles, (uint64_t)std::max(cus, 1) * (uint64_t)std::max(per_cu, 1));
    KernelScope ks(s, "multi_mfma", (double)ngroups * D * 2.0 * sizeof(V2<R>));
    hipLaunchKernelGGL((k_multi_mfma<R, K>), dim3(grid), dim3(256), lds, s.stream,
                       reinterpret_cast<V2<R>*>(s.psi), ngroups, sp, g.cmask, tm, dat, member_fast, tiles);
    QFVM_LAUNCHED();
    QFVM_HIP(hipStreamSynchronize(s.stream));  // the host matrix copy must outlive the async copy
}

template <typename R>
static void multi_typed(State& s, const Gate& g) {
    const int k = (int)g.targets.size();
    const uint64_t dim = 1ull << k;
    const size_t lds = dim * sizeof(V2<R>);
    if (k > 16 || lds > 65536) throw Error(QFVM_ERR_VALUE, "dense gate on " + std::to_string(k) + " targets is too large");
    std::vector<int> pos = g.ctrls;
    pos.insert(pos.end(), g.targets.begin(), g.targets.end());
    const InsertSpec sp = make_spec(pos);
    const uint64_t ngroups = s.size >> pos.size();
    if (s.mfma && k >= 4 && k <= 8) {
        switch (k) {
            case 4: multi_mfma_launch<R, 4>(s, g, sp, ngroups); break;
            case 5: multi_mfma_launch<R, 5>(s, g, sp, ngroups); break;
            case 6: multi_mfma_launch<R, 6>(s, g, sp, ngroups); break;
            case 7: multi_wide_launch<R, 7>(s, g, sp, ngroups); break;
            default: multi_wide_launch<R, 8>(s, g, sp, ngroups); break;
        }
        return;
    }
    // matrix -> device (scratch), converted to the state's precision
    std::vector<V2<R>> hm(dim * dim);
    for (uint64_t q = 0; q < dim * dim; q++) hm[q] = mk<R>((R)g.mk[q].real(), (R)g.mk[q].imag());
    V2<R>* dmat = reinterpret_cast<V2<R>*>(s.ensure_scratch(hm.size() * sizeof(V2<R>)));
    QFVM_HIP(hipMemcpyAsync(dmat, hm.data(), hm.size() * sizeof(V2<R>), hipMemcpyHostToDevice, s.stream));
    switch (k) {
        case 2: multi_reg_launch<R, 2>(s, g, sp, ngroups, dmat); break;
        case 3: multi_reg_launch<R, 3>(s, g, sp, ngroups, dmat); break;
        case 4: multi_reg_launch<R, 4>(s, g, sp, ngroups, dmat); break;
        case 5: multi_reg_launch<R, 5>(s, g, sp, ngroups, dmat); break;
        default: {
            TPos tp{};
            tp.k = k;
            for (int b = 0; b < k; b++) tp.t[b] = g.targets[b];
            const unsigned grid = (unsigned)std::min<uint64_t>(ngroups, 1u << 16);
            KernelScope ks(s, "multi_lds", (double)ngroups * dim * 2.0 * sizeof(V2<R>));
            hipLaunchKernelGGL(k_multi_lds<R>, dim3(grid), dim3(256), lds, s.stream, reinterpret_cast<V2<R>*>(s.psi),
                               ngroups, sp, g.cmask, tp, dmat);
            QFVM_LAUNCHED();
        }
    }
    // the host matrix copy must outlive the async copy
    QFVM_HIP(hipStreamSynchronize(s.stream));
}

void launch_gate(State& s, const Gate& g) {
    if (g.kind == G_MULTI) {
        if (s.dtype == QFVM_C128)
            multi_typed<double>(s, g);
        else
            multi_typed<float>(s, g);
        return;
    }
    if (s.dtype == QFVM_C128)
        gate1_typed<double>(s, g);
    else
        gate1_typed<float>(s, g);
}

void launch_zero(State& s, void* p, size_t bytes) {
    if (bytes % 8 != 0) throw Error(QFVM_ERR_RUNTIME, "zero fill of a partial word");
    const uint64_t n = bytes / 8;
    if (n == 0) return;
    const unsigned grid = (unsigned)std::min<uint64_t>(grid_for(n, 256), 8192);
    hipLaunchKernelGGL(k_zero_words, dim3(grid), dim3(256), 0, s.stream, reinterpret_cast<uint64_t*>(p), n);
    QFVM_LAUNCHED();
}

void launch_fill_basis(State& s, uint64_t index) {
    const unsigned block = 256;
    if (s.dtype == QFVM_C128)
        hipLaunchKernelGGL(k_fill_basis_c128, dim3(grid_for(s.size, block)), dim3(block), 0, s.stream,
                           reinterpret_cast<double2*>(s.psi), s.size, index);
    else
        hipLaunchKernelGGL(k_fill_basis_c64, dim3(grid_for(s.size, block)), dim3(block), 0, s.stream,
                           reinterpret_cast<float2*>(s.psi), s.size, index);
    QFVM_LAUNCHED();
}

// host complex128 -> device state
void launch_h2d_convert(State& s, const double* host, uint64_t n) {
    if (s.dtype == QFVM_C128) {
        QFVM_HIP(hipMemcpyAsync(s.psi, host, n * 16, hipMemcpyHostToDevice, s.stream));
        QFVM_HIP(hipStreamSynchronize(s.stream));
        return;
    }
    const uint64_t chunk = 1ull << 22;  // amplitudes per staging round (64 MiB of complex128)
    double2* stage = reinterpret_cast<double2*>(s.ensure_scratch(std::min(n, chunk) * 16));
    for (uint64_t b = 0; b < n; b += chunk) {
        const uint64_t m = std::min(chunk, n - b);
        QFVM_HIP(hipMemcpyAsync(stage, host + 2 * b, m * 16, hipMemcpyHostToDevice, s.stream));
        hipLaunchKernelGGL(k_narrow, dim3(grid_for(m, 256)), dim3(256), 0, s.stream, stage,
                           reinterpret_cast<float2*>(s.psi) + b, m);
        QFVM_LAUNCHED();
        QFVM_HIP(hipStreamSynchronize(s.stream));
    }
}

void launch_d2h_convert(State& s, double* host, uint64_t n) {
    if (s.dtype == QFVM_C128) {
        QFVM_HIP(hipMemcpyAsync(host, s.psi, n * 16, hipMemcpyDeviceToHost, s.stream));
        QFVM_HIP(hipStreamSynchronize(s.stream));
        return;
    }
    const uint64_t chunk = 1ull << 22;
    double2* stage = reinterpret_cast<double2*>(s.ensure_scratch(std::min(n, chunk) * 16));
    for (uint64_t b = 0; b < n; b += chunk) {
        const uint64_t m = std::min(chunk, n - b);
        hipLaunchKernelGGL(k_widen, dim3(grid_for(m, 256)), dim3(256), 0, s.stream,
                           reinterpret_cast<const float2*>(s.psi) + b, stage, m);
        QFVM_LAUNCHED();
        QFVM_HIP(hipMemcpyAsync(host + 2 * b, stage, m * 16, hipMemcpyDeviceToHost, s.stream));
        QFVM_HIP(hipStreamSynchronize(s.stream));
    }
}

}  // namespace qfvm
