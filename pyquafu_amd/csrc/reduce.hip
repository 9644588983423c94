// Reductions over the state: squared norm, probabilities, Pauli-string expectations, marginal
// probabilities of measured qubits, Hamiltonian expectations and gradient markers; and the
// projective collapse / reset kernels.
//
// Sums are deterministic: a fixed grid (independent of the device), a fixed per-thread stride
// order, a fixed wavefront shuffle tree and a fixed block tree, then one block that adds the
// per-block partials in index order. Algorithmic bytes: one read of the state (s * 2^n) per
// reduction.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <optional>

#include <utility>

#include "common.h"

namespace qfvm {

constexpr unsigned RED_BLOCK = 256;
constexpr unsigned RED_GRID = 2048;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum of one value per thread over the block; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double wsum[RED_BLOCK / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wsum[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int k = 0; k < (int)(RED_BLOCK / 64); k++) r += wsum[k];
    }
    __syncthreads();
    return r;
}

template <typename R>
__device__ __forceinline__ double norm2_of(V2<R> a) {
    const double x = (double)a.x, y = (double)a.y;
    return x * x + y * y;
}

template <typename R>
__global__ void __launch_bounds__(RED_BLOCK) k_norm2(const V2<R>* __restrict__ psi, uint64_t size, double* partial) {
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) acc += norm2_of<R>(psi[i]);
    acc = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// expect_pauli (statevector.hpp:1250-1305): no flips -> sum |psi_j|^2 (-1)^popc(j & z);
// with flips -> sum over pairs i0 (bit flip_q clear), i1 = i0 ^ flip of 2 Re(psi_i0 conj(psi_i1) i^k),
// k = 2*popc(i0 & z) + #Y (mod 4).
template <typename R>
__global__ void __launch_bounds__(RED_BLOCK) k_pauli(const V2<R>* __restrict__ psi, uint64_t work, uint64_t flip,
                                                      uint64_t zmask, int flip_q, int y_num, double* partial) {
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    if (flip == 0) {
        for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < work; j += stride) {
            const double p = norm2_of<R>(psi[j]);
            acc += (__popcll(j & zmask) & 1) ? -p : p;
        }
    } else {
        const uint64_t low = (1ull << flip_q) - 1;
        for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < work; j += stride) {
            const uint64_t i0 = (j & low) | ((j & ~low) << 1);
            const uint64_t i1 = i0 ^ flip;
            const V2<R> a = psi[i0], b = psi[i1];
            const double re = (double)a.x * b.x + (double)a.y * b.y;  // Re(a conj(b))
            const double im = (double)a.y * b.x - (double)a.x * b.y;  // Im(a conj(b))
            const int k = (int)((2 * (__popcll(i0 & zmask) & 1) + y_num) & 3);
            const double v = (k == 0) ? re : (k == 1) ? -im : (k == 2) ? -re : im;
            acc += 2.0 * v;
        }
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void __launch_bounds__(RED_BLOCK) k_sum_partials(const double* __restrict__ partial, int n, double* out) {
    // fixed order: thread t adds t, t+256, ...; then the block tree
    double acc = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) acc += partial[k];
    acc = block_sum(acc);
    if (threadIdx.x == 0) out[0] = acc;
}

template <typename R>
__global__ void k_probs(const V2<R>* __restrict__ psi, uint64_t size, double* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) out[i] = norm2_of<R>(psi[i]);
}

struct QList {
    int n;
    int q[64];
};

__device__ __forceinline__ uint64_t gather_bits(uint64_t j, const QList& ql) {
    uint64_t m = 0;
    for (int b = 0; b < ql.n; b++) m |= ((j >> ql.q[b]) & 1ull) << b;
    return m;
}

// marginal probabilities of up to 12 qubits: per-block LDS histogram, per-block rows out
template <typename R>
__global__ void __launch_bounds__(RED_BLOCK) k_marginal(const V2<R>* __restrict__ psi, uint64_t size, QList ql,
                                                         double* __restrict__ rows) {
    extern __shared__ double hist[];
    const int bins = 1 << ql.n;
    for (int b = threadIdx.x; b < bins; b += blockDim.x) hist[b] = 0.0;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) {
        const double p = norm2_of<R>(psi[i]);
        if (p != 0.0) atomicAdd(&hist[gather_bits(i, ql)], p);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < bins; b += blockDim.x) rows[(uint64_t)blockIdx.x * bins + b] = hist[b];
}

template <typename R>
__global__ void k_collapse(V2<R>* __restrict__ psi, uint64_t size, QList ql, uint64_t meas, R scale) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) {
        V2<R> a = psi[i];
        if (gather_bits(i, ql) == meas) {
            if (scale != (R)1) a = mk<R>(a.x * scale, a.y * scale);
        } else {
            a = mk<R>(0, 0);
        }
        psi[i] = a;
    }
}

template <typename R>
__global__ void k_swap_states(V2<R>* __restrict__ psi, uint64_t size, uint64_t qmask, uint64_t fmask, uint64_t mmask) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) {
        if ((i & qmask) == fmask) {
            const uint64_t k = (i & ~qmask) | mmask;
            const V2<R> a = psi[i];
            psi[i] = psi[k];
            psi[k] = a;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Many Pauli-string expectations in few passes (expect_statevec, qfvm.cpp:246-261, term by term
// through expect_pauli, statevector.hpp:1250-1305). A pass streams the state through LDS in tiles
// of 2^b amplitudes whose qubits are the L lowest plus up to b - L chosen higher ones; every term
// whose X / Y qubits all lie in the tile is evaluated from LDS: its pairs (i0, i1 = i0 ^ flip, i0
// with the highest flip bit clear) are both in the tile, and its Z / Y mask only needs the global
// index. Terms are packed first-fit into passes, so a Hamiltonian of Z strings and single-qubit X
// terms on n qubits needs about (n - L) / (b - L) passes instead of one pass per term. Workgroups
// walk a fixed set of tiles; each thread adds its elements' contributions per term across them, the
// workgroup reduces once, and the per-workgroup sums are added in index order (deterministic).
// Algorithmic bytes per pass: one read of the state (s 2^n).
constexpr int PT_MAX = 32;        // terms per launch (kernel argument)
constexpr unsigned PT_GRID = 2048;
static_assert(EXPECT_PARTIALS >= (size_t)PT_MAX * PT_GRID, "expectation partials");
static_assert(RED_PARTIALS >= RED_GRID + 1, "reduction partials");

struct PTerm {
    uint64_t z;     // Z / Y mask, global bits
    uint32_t zl;    // the Z / Y bits inside the tile, as tile-local bits
    uint32_t f;     // flip mask in tile-local bits; 0 for a term of Z and I only
    int32_t fq;     // tile-local position of the highest flip bit
    int32_t y;      // number of Y letters
    int32_t slot;   // output index
};

struct PPass {
    int b, L, nterms;
    int q[16];  // tile qubits, ascending; q[k] = k for k < L
    PTerm t[PT_MAX];
};

template <typename F, int... Ks>
__device__ __forceinline__ void each_index(F&& f, std::integer_sequence<int, Ks...>) {
    (f(std::integral_constant<int, Ks>{}), ...);
}

// One thread holds up to EPT elements of each tile (EPT = 2^b / RED_BLOCK at the largest tile) and
// one accumulator per term across all of its workgroup's tiles; the workgroup reduces them once at
// the end. A term's Z / Y parity splits into the tile base's (uniform, per tile) and the local
// index's (32-bit, mask zl in local bits).
template <typename R, int EPT, int NT>
__global__ void __launch_bounds__(RED_BLOCK, NT <= 16 ? 4 : 1) k_expect_tile(const V2<R>* __restrict__ psi, const PPass pp,
                                                           uint64_t tiles, double* __restrict__ partials) {
    static_assert(EPT == 8 && RED_BLOCK == 256, "element i of a thread is local index tid + (i << 8)");
    extern __shared__ unsigned char lds_raw[];
    const int b = pp.b, L = pp.L, T = 1 << b;
    // Two layouts: with up to 16 terms (WREG) the Walsh sums stay in registers and the thread's own
    // elements are read back from the tile, so that 4 wavefronts fit per SIMD (118 VGPRs, 33.5 KiB
    // of LDS); with up to 32 the Walsh sums go to LDS and the elements stay in registers.
    constexpr bool WREG = NT <= 16;
    V2<R>* tile = reinterpret_cast<V2<R>*>(lds_raw);
    double* wz = reinterpret_cast<double*>(tile + 2048);  // !WREG: per thread, Walsh sums of its 8 probabilities
    uint64_t* hi = reinterpret_cast<uint64_t*>(wz + (WREG ? 0 : 8 * RED_BLOCK));  // global offsets of the high bits
    double* red = reinterpret_cast<double*>(hi + (1 << (b - L)));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t lowmask = (1ull << L) - 1;
    for (int i = tid; i < (1 << (b - L)); i += RED_BLOCK) {
        uint64_t g = 0;
        for (int k = L; k < b; k++)
            if ((i >> (k - L)) & 1) g |= 1ull << pp.q[k];
        hi[i] = g;
    }
    static_assert(NT <= PT_MAX, "terms per launch");
    double a[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) a[k] = 0.0;
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        uint64_t base = t;
        for (int k = 0; k < b; k++) {
            const uint64_t low = (1ull << pp.q[k]) - 1;
            base = (base & low) | ((base & ~low) << 1);
        }
        // parity of each term's Z / Y bits outside the tile (bit k for term k; uniform)
        uint32_t bpar = 0;
        for (int k = 0; k < pp.nterms; k++) bpar |= (uint32_t)(__popcll(base & pp.t[k].z) & 1) << k;
        __syncthreads();  // the previous tile is no longer read; hi is built
        V2<R> v[EPT];
#pragma unroll
        for (int i = 0; i < EPT; i++) {
            const int e = tid + i * (int)RED_BLOCK;
            if (e < T) v[i] = psi[base | ((uint64_t)e & lowmask) | hi[e >> L]];
        }
#pragma unroll
        for (int i = 0; i < EPT; i++) {
            const int e = tid + i * (int)RED_BLOCK;
            if (e < T) tile[e] = v[i];
        }
        __syncthreads();
        double pr[EPT];
#pragma unroll
        for (int i = 0; i < EPT; i++)
            pr[i] = tid + i * (int)RED_BLOCK < T ? (double)v[i].x * v[i].x + (double)v[i].y * v[i].y : 0.0;
        // Z strings: this thread's elements differ only in local bits 8..10 (= i), so a term's sum over
        // them is (-1)^(parity of tid's bits and the tile base's) times the Walsh sum
        // W[m] = sum_i (-1)^popc(i & m) pr[i] at m = the term's bits 8..10
        double w[8];
        {
#pragma unroll
            for (int m = 0; m < 8; m++) w[m] = pr[m];
#pragma unroll
            for (int h = 1; h < 8; h <<= 1)
#pragma unroll
                for (int m = 0; m < 8; m++)
                    if (!(m & h)) {
                        const double x = w[m], y = w[m | h];
                        w[m] = x + y;
                        w[m | h] = x - y;
                    }
            if constexpr (!WREG) {
#pragma unroll
                for (int m = 0; m < 8; m++) wz[m * RED_BLOCK + tid] = w[m];
            }
        }
        // the term records are read inside the tile loop (an opaque zero offset keeps the compiler
        // from holding all of them in scalar registers across it)
        int k0 = 0;
        asm volatile("" : "+s"(k0));
        // one term at a time, the term index a compile-time constant (the accumulators stay in
        // registers)
        auto term = [&](auto kc) {
            constexpr int k = decltype(kc)::value;
            if (k >= pp.nterms) return;
            const PTerm& tm = pp.t[k + k0];
            const uint32_t zl = tm.zl, f = tm.f;
            const int fq = tm.fq, ty = tm.y;
            const int bk = (int)(bpar >> k) & 1;
            if (f == 0) {
                const uint32_t zh = (zl >> 8) & 7u;
                double wv;
                if constexpr (WREG) {  // uniform: a select tree over the registers
                    const bool s0 = zh & 1u, s1 = zh & 2u, s2 = zh & 4u;
                    const double w01 = s0 ? w[1] : w[0], w23 = s0 ? w[3] : w[2];
                    const double w45 = s0 ? w[5] : w[4], w67 = s0 ? w[7] : w[6];
                    const double w03 = s1 ? w23 : w01, w47 = s1 ? w67 : w45;
                    wv = s2 ? w47 : w03;
                } else {
                    wv = wz[(int)zh * (int)RED_BLOCK + tid];
                }
                a[k] += ((__popc((uint32_t)tid & zl & 0xffu) ^ bk) & 1) ? -wv : wv;
                return;
            }
#pragma unroll
            for (int i = 0; i < EPT; i++) {
                const int e = tid + i * (int)RED_BLOCK;
                if (e >= T) continue;
                const int zpar = (__popc((uint32_t)e & zl) ^ bk) & 1;
                if (!((e >> fq) & 1)) {
                    const V2<R> x = WREG ? tile[e] : v[i], y = tile[e ^ (int)f];
                    const double re = (double)x.x * y.x + (double)x.y * y.y;  // Re(x conj(y))
                    const double im = (double)x.y * y.x - (double)x.x * y.y;  // Im(x conj(y))
                    const int kk = (2 * zpar + ty) & 3;  // phase i^kk (PHASE_YZ = 1, i, -1, -i)
                    const double val = kk == 0 ? re : kk == 1 ? -im : kk == 2 ? -re : im;
                    a[k] += 2.0 * val;
                }
            }
        };
        each_index(term, std::make_integer_sequence<int, NT>{});
    }
    // one reduction per workgroup: wavefront sums, then the waves in index order
#pragma unroll
    for (int k = 0; k < NT; k++) {
        if (k < pp.nterms) {
            const double w = wave_sum(a[k]);
            if (lane == 0) red[k * (RED_BLOCK / 64) + wave] = w;
        }
    }
    __syncthreads();
    if (tid < pp.nterms) {
        double v = 0.0;
        for (int w = 0; w < (int)(RED_BLOCK / 64); w++) v += red[tid * (RED_BLOCK / 64) + w];
        partials[(uint64_t)tid * gridDim.x + blockIdx.x] = v;
    }
}

// per-term sums of a launch's partials into out[term.slot]
__global__ void __launch_bounds__(RED_BLOCK) k_term_sums(const double* __restrict__ partials, unsigned cols,
                                                         const PPass pp, double* __restrict__ out) {
    const double* col = partials + (uint64_t)blockIdx.x * cols;
    double a = 0.0;
    for (unsigned k = threadIdx.x; k < cols; k += blockDim.x) a += col[k];
    a = block_sum(a);
    if (threadIdx.x == 0) out[pp.t[blockIdx.x].slot] = a;
}

// ------------------------------------------------------------------------------------------------
static double finish_sum(State& s, double* partial, unsigned n) {
    double* out = partial + n;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(RED_BLOCK), 0, s.stream, partial, (int)n, out);
    QFVM_LAUNCHED();
    double* h = s.ensure_pinned(sizeof(double));
    QFVM_HIP(hipMemcpyAsync(h, out, sizeof(double), hipMemcpyDeviceToHost, s.stream));
    QFVM_HIP(hipStreamSynchronize(s.stream));
    return h[0];
}

double reduce_norm2(State& s) {
    double* partial = reinterpret_cast<double*>(s.ensure_scratch((RED_GRID + 1) * sizeof(double)));
    const unsigned grid = std::min<unsigned>(RED_GRID, grid_for(s.size, RED_BLOCK));
    if (s.dtype == QFVM_C128)
        hipLaunchKernelGGL(k_norm2<double>, dim3(grid), dim3(RED_BLOCK), 0, s.stream,
                           reinterpret_cast<const double2*>(s.psi), s.size, partial);
    else
        hipLaunchKernelGGL(k_norm2<float>, dim3(grid), dim3(RED_BLOCK), 0, s.stream,
                           reinterpret_cast<const float2*>(s.psi), s.size, partial);
    QFVM_LAUNCHED();
    return finish_sum(s, partial, grid);
}

// Queues k_pauli for one term whose X / Y qubits do not fit one tile, leaving the value in *out.
static void queue_pauli_single(State& s, const PauliHost& t, double* partial, double* out, bool account) {
    const uint64_t work = t.flip ? (s.size >> 1) : s.size;
    const unsigned grid = std::min<unsigned>(RED_GRID, grid_for(work, RED_BLOCK));
    std::optional<KernelScope> ks;
    if (account) ks.emplace(s, "pauli_expect", (double)s.size * s.amp_bytes());
    if (s.dtype == QFVM_C128)
        hipLaunchKernelGGL(k_pauli<double>, dim3(grid), dim3(RED_BLOCK), 0, s.stream,
                           reinterpret_cast<const double2*>(s.psi), work, t.flip, t.z, t.flip_q, t.y, partial);
    else
        hipLaunchKernelGGL(k_pauli<float>, dim3(grid), dim3(RED_BLOCK), 0, s.stream,
                           reinterpret_cast<const float2*>(s.psi), work, t.flip, t.z, t.flip_q, t.y, partial);
    QFVM_LAUNCHED();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(RED_BLOCK), 0, s.stream, partial, (int)grid, out);
    QFVM_LAUNCHED();
}

void queue_expectations(State& s, const PauliHost* terms, int n_terms, double* out_dev, double* partials,
                        double* fallback, bool account) {
    if (n_terms > EXPECT_BATCH) throw Error(QFVM_ERR_VALUE, "too many terms in one expectation batch");
    const int n = s.n;
    const int b = std::min(n, 11);  // 2^11 amplitudes per tile: 8 per thread (k_expect_tile<R, 8>)
    const int L = b < n ? std::min(5, b - 1) : b;
    const uint64_t lowmask = (1ull << L) - 1;
    const uint64_t tiles = 1ull << (n - b);
    const unsigned grid = (unsigned)std::min<uint64_t>(tiles, PT_GRID);
    // Passes: the terms with flips above the L low qubits are packed greedily into passes whose
    // b - L chosen qubits hold all their high flips (a term too wide for any tile runs alone); the
    // other terms (Z strings, flips on low qubits only) fit every pass and go to the pass holding the
    // fewest terms, so that no pass is read for a handful of terms. Fixed-size arrays only: the
    // allocation-free entry point runs this too.
    int gid[EXPECT_BATCH];       // pass of each term (-1: runs alone)
    uint64_t gQh[EXPECT_BATCH];  // high flip qubits of each pass
    int gcnt[EXPECT_BATCH];
    int ng = 0;
    for (int i = 0; i < n_terms; i++) {
        gid[i] = -1;
        const uint64_t hb = terms[i].flip & ~lowmask;
        if (hb == 0) continue;
        if (__builtin_popcountll(hb) > b - L) {  // cannot share a tile
            queue_pauli_single(s, terms[i], fallback, out_dev + terms[i].slot, account);
            continue;
        }
        int g = 0;
        while (g < ng && __builtin_popcountll(gQh[g] | hb) > b - L) g++;
        if (g == ng) {
            gQh[ng] = 0;
            gcnt[ng++] = 0;
        }
        gQh[g] |= hb;
        gcnt[g]++;
        gid[i] = g;
    }
    for (int i = 0; i < n_terms; i++) {
        if ((terms[i].flip & ~lowmask) != 0) continue;
        if (ng == 0) {
            gQh[0] = 0;
            gcnt[ng++] = 0;
        }
        int g = 0;
        for (int c = 1; c < ng; c++)
            if (gcnt[c] < gcnt[g]) g = c;
        gcnt[g]++;
        gid[i] = g;
    }
    int order[EXPECT_BATCH];
    for (int grp = 0; grp < ng; grp++) {
        uint64_t Qh = gQh[grp];
        int cnt = 0;
        for (int i = 0; i < n_terms; i++)
            if (gid[i] == grp) order[cnt++] = i;
        for (int q = L; q < n && __builtin_popcountll(Qh) < b - L; q++) Qh |= 1ull << q;
        PPass pp{};
        pp.b = b;
        pp.L = L;
        int local_of[64];
        for (int k = 0; k < 64; k++) local_of[k] = -1;
        for (int k = 0; k < L; k++) {
            pp.q[k] = k;
            local_of[k] = k;
        }
        int kq = L;
        for (int q = L; q < n; q++)
            if ((Qh >> q) & 1) {
                pp.q[kq] = q;
                local_of[q] = kq++;
            }
        for (int c0 = 0; c0 < cnt; c0 += PT_MAX) {
            pp.nterms = std::min(PT_MAX, cnt - c0);
            for (int k = 0; k < pp.nterms; k++) {
                const PauliHost& t = terms[order[c0 + k]];
                PTerm& d = pp.t[k];
                d.z = t.z;
                d.zl = 0;
                for (int q = 0; q < n; q++)
                    if (((t.z >> q) & 1) && local_of[q] >= 0) d.zl |= 1u << local_of[q];
                d.f = 0;
                d.fq = 0;
                for (int q = 0; q < n; q++)
                    if ((t.flip >> q) & 1) {
                        d.f |= 1u << local_of[q];
                        d.fq = local_of[q];
                    }
                d.y = t.y;
                d.slot = t.slot;
            }
            // up to 16 terms: the instance with 16 accumulators and the Walsh sums in registers,
            // which fits four wavefronts per SIMD. LDS: a tile region of 2^11 amplitudes (smaller
            // tiles leave part of it unused), the threads' Walsh sums (more than 16 terms), the
            // high-bit offsets and the reduction slots
            const bool few = pp.nterms <= 16;
            const size_t lds = (size_t)2048 * s.amp_bytes() + (few ? 0 : (size_t)8 * RED_BLOCK * sizeof(double)) +
                               ((size_t)1 << (b - L)) * sizeof(uint64_t) + (size_t)PT_MAX * (RED_BLOCK / 64) * sizeof(double);
            std::optional<KernelScope> ks;
            if (account) ks.emplace(s, "expect_tile", (double)s.size * s.amp_bytes());
            if (s.dtype == QFVM_C128) {
                const auto fn = few ? k_expect_tile<double, 8, 16> : k_expect_tile<double, 8, PT_MAX>;
                hipLaunchKernelGGL(fn, dim3(grid), dim3(RED_BLOCK), lds, s.stream,
                                   reinterpret_cast<const double2*>(s.psi), pp, tiles, partials);
            } else {
                const auto fn = few ? k_expect_tile<float, 8, 16> : k_expect_tile<float, 8, PT_MAX>;
                hipLaunchKernelGGL(fn, dim3(grid), dim3(RED_BLOCK), lds, s.stream,
                                   reinterpret_cast<const float2*>(s.psi), pp, tiles, partials);
            }
            QFVM_LAUNCHED();
            hipLaunchKernelGGL(k_term_sums, dim3((unsigned)pp.nterms), dim3(RED_BLOCK), 0, s.stream, partials, grid, pp,
                               out_dev);
            QFVM_LAUNCHED();
        }
    }
}

// Allocation-free expectations (qfvm_expect_paulis_noalloc). Its scratch is this file-scope device
// array: per-workgroup partials, the wide-term fallback's partials and the results, 530 KiB. Every
// GPU the library runs on holds its own copy until the process ends, and every call on that GPU
// shares it (calls that overlap on one GPU race). Host code never names the array in a kernel
// argument or copy: it uses the address hipGetSymbolAddress returns, resolved once per device with
// that device current.
constexpr size_t NOALLOC_DOUBLES = EXPECT_PARTIALS + RED_PARTIALS + EXPECT_BATCH;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

static double* noalloc_scratch(int device) {
    static std::mutex lock;
    static double* ptr[64] = {};
    std::lock_guard<std::mutex> guard(lock);
    if (device < 0 || device >= 64) throw Error(QFVM_ERR_DEVICE, "device index out of range");
    if (!ptr[device]) {
        void* p = nullptr;
        QFVM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[device] = static_cast<double*>(p);
    }
    return ptr[device];
}

void expect_paulis_noalloc(State& s, const PauliHost* terms, int n_terms, double* host_out) {
    if (n_terms > EXPECT_BATCH) throw Error(QFVM_ERR_VALUE, "at most 256 terms per allocation-free call");
    double* base = noalloc_scratch(s.device);
    double* partials = base;
    double* fallback = base + EXPECT_PARTIALS;
    double* out_dev = fallback + RED_PARTIALS;
    queue_expectations(s, terms, n_terms, out_dev, partials, fallback, false);
    QFVM_HIP(hipMemcpyAsync(host_out, out_dev, (size_t)n_terms * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    QFVM_HIP(hipStreamSynchronize(s.stream));
}

void reduce_probabilities(State& s, double* host_out) {
    const uint64_t chunk = 1ull << 23;
    double* stage = reinterpret_cast<double*>(s.ensure_scratch(std::min(s.size, chunk) * sizeof(double)));
    for (uint64_t b = 0; b < s.size; b += chunk) {
        const uint64_t m = std::min(chunk, s.size - b);
        if (s.dtype == QFVM_C128)
            hipLaunchKernelGGL(k_probs<double>, dim3(grid_for(m, 256)), dim3(256), 0, s.stream,
                               reinterpret_cast<const double2*>(s.psi) + b, m, stage);
        else
            hipLaunchKernelGGL(k_probs<float>, dim3(grid_for(m, 256)), dim3(256), 0, s.stream,
                               reinterpret_cast<const float2*>(s.psi) + b, m, stage);
        QFVM_LAUNCHED();
        QFVM_HIP(hipMemcpyAsync(host_out + b, stage, m * sizeof(double), hipMemcpyDeviceToHost, s.stream));
        QFVM_HIP(hipStreamSynchronize(s.stream));
    }
}

static QList make_qlist(const std::vector<int>& q) {
    QList ql{};
    if (q.size() > 64) throw Error(QFVM_ERR_VALUE, "too many qubits");
    ql.n = (int)q.size();
    for (size_t b = 0; b < q.size(); b++) ql.q[b] = q[b];
    return ql;
}

void reduce_marginal(State& s, const std::vector<int>& qbits, std::vector<double>& probs) {
    const int k = (int)qbits.size();
    const uint64_t bins = 1ull << k;
    probs.assign(bins, 0.0);
    if (k > 12) {
        // many measured qubits: full probabilities on the host, then marginalise
        std::vector<double> p(s.size);
        reduce_probabilities(s, p.data());
        for (uint64_t i = 0; i < s.size; i++) {
            uint64_t m = 0;
            for (int b = 0; b < k; b++) m |= ((i >> qbits[b]) & 1ull) << b;
            probs[m] += p[i];
        }
        return;
    }
    const unsigned grid = std::min<unsigned>(512, grid_for(s.size, RED_BLOCK));
    double* rows = reinterpret_cast<double*>(s.ensure_scratch((size_t)grid * bins * sizeof(double)));
    const QList ql = make_qlist(qbits);
    const size_t lds = bins * sizeof(double);
    if (s.dtype == QFVM_C128)
        hipLaunchKernelGGL(k_marginal<double>, dim3(grid), dim3(RED_BLOCK), lds, s.stream,
                           reinterpret_cast<const double2*>(s.psi), s.size, ql, rows);
    else
        hipLaunchKernelGGL(k_marginal<float>, dim3(grid), dim3(RED_BLOCK), lds, s.stream,
                           reinterpret_cast<const float2*>(s.psi), s.size, ql, rows);
    QFVM_LAUNCHED();
    std::vector<double> h((size_t)grid * bins);
    QFVM_HIP(hipMemcpyAsync(h.data(), rows, h.size() * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    QFVM_HIP(hipStreamSynchronize(s.stream));
    for (unsigned g = 0; g < grid; g++)
        for (uint64_t b = 0; b < bins; b++) probs[b] += h[(size_t)g * bins + b];
}

void launch_collapse(State& s, const std::vector<int>& qbits, uint64_t meas_state, double scale) {
    const QList ql = make_qlist(qbits);
    const unsigned grid = grid_for(s.size, 256);
    if (s.dtype == QFVM_C128)
        hipLaunchKernelGGL(k_collapse<double>, dim3(grid), dim3(256), 0, s.stream, reinterpret_cast<double2*>(s.psi),
                           s.size, ql, meas_state, scale);
    else
        hipLaunchKernelGGL(k_collapse<float>, dim3(grid), dim3(256), 0, s.stream, reinterpret_cast<float2*>(s.psi),
                           s.size, ql, meas_state, (float)scale);
    QFVM_LAUNCHED();
}

// ------------------------------------------------------------------------------------------------
// Observables and gradients.
//
// k_hamiltonian: per amplitude i, (H psi)[i] = sum over flip groups of psi[i ^ flip] times the sum
// of the group's signed coefficients; adds Re(conj(psi[i]) (H psi)[i]) to the expectation and, when
// `out` is given, stores (H psi)[i]. Bytes: one write plus one read per flip group (neighbouring
// flips are served from cache).
template <typename R>
__global__ void __launch_bounds__(RED_BLOCK) k_hamiltonian(const V2<R>* __restrict__ psi, V2<R>* __restrict__ out,
                                                            uint64_t size, const HamGroup* __restrict__ groups,
                                                            int ngroups, const HamTerm* __restrict__ terms,
                                                            double* partial) {
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) {
        double sr = 0.0, si = 0.0;
        for (int g = 0; g < ngroups; g++) {
            const uint64_t j = i ^ groups[g].flip;
            double cr = 0.0, ci = 0.0;
            const int t1 = groups[g].first + groups[g].count;
            for (int t = groups[g].first; t < t1; t++) {
                const bool neg = __popcll(j & terms[t].zy) & 1;
                cr += neg ? -terms[t].cre : terms[t].cre;
                ci += neg ? -terms[t].cim : terms[t].cim;
            }
            const V2<R> p = psi[j];
            sr += cr * (double)p.x - ci * (double)p.y;
            si += cr * (double)p.y + ci * (double)p.x;
        }
        if (out) out[i] = mk<R>((R)sr, (R)si);
        const V2<R> a = psi[i];
        acc += (double)a.x * sr + (double)a.y * si;
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// One gradient marker over the whole state: sum over control-satisfied i of
// Im(conj(lam[i]) (P psi)[i]); X: (P psi)[i] = psi[i ^ t]; Y: -i psi[i | t] where bit t of i is clear,
// +i psi[i & ~t] where set; Z: (-1)^bit psi[i].
template <typename R>
__global__ void __launch_bounds__(RED_BLOCK) k_grad_marker(const V2<R>* __restrict__ psi, const V2<R>* __restrict__ lam,
                                                            uint64_t size, int gen, int q, uint64_t cmask,
                                                            double* partial) {
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t tb = 1ull << q;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) {
        if ((i & cmask) != cmask) continue;
        const V2<R> l = lam[i];
        const bool bit = (i & tb) != 0;
        double v;
        if (gen == GEN_Z) {
            const V2<R> p = psi[i];
            const double w = (double)l.x * p.y - (double)l.y * p.x;
            v = bit ? -w : w;
        } else {
            const V2<R> p = psi[i ^ tb];
            if (gen == GEN_X) {
                v = (double)l.x * p.y - (double)l.y * p.x;
            } else {
                const double re = (double)l.x * p.x + (double)l.y * p.y;
                v = bit ? re : -re;
            }
        }
        acc += v;
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// Column c of a [cols x tiles] table of per-tile partials -> grads[slot_of[c]], in a fixed order.
__global__ void __launch_bounds__(RED_BLOCK) k_column_sums(const double* __restrict__ partials, uint64_t tiles,
                                                            const int* __restrict__ slot_of, double* grads) {
    const double* col = partials + (uint64_t)blockIdx.x * tiles;
    double acc = 0.0;
    for (uint64_t k = threadIdx.x; k < tiles; k += blockDim.x) acc += col[k];
    acc = block_sum(acc);
    if (threadIdx.x == 0) grads[slot_of[blockIdx.x]] = acc;
}

double apply_hamiltonian(State& s, const Hamiltonian& H, void* out, double* dev_result) {
    const size_t gbytes = H.groups.size() * sizeof(HamGroup), tbytes = H.terms.size() * sizeof(HamTerm);
    const size_t toff = (gbytes + 255) / 256 * 256;
    unsigned char* stage = s.stage_host(STAGE_HAM, toff + tbytes + 256);
    unsigned char* dev = s.stage_dev(STAGE_HAM, toff + tbytes + 256);
    if (!H.groups.empty()) {
        std::memcpy(stage, H.groups.data(), gbytes);
        std::memcpy(stage + toff, H.terms.data(), tbytes);
        s.stage_upload(STAGE_HAM, toff + tbytes);
    }
    double* partial = reinterpret_cast<double*>(s.ensure_scratch((RED_GRID + 1) * sizeof(double)));
    const unsigned grid = std::min<unsigned>(RED_GRID, grid_for(s.size, RED_BLOCK));
    const auto* g = reinterpret_cast<const HamGroup*>(dev);
    const auto* t = reinterpret_cast<const HamTerm*>(dev + toff);
    KernelScope ks(s, "hamiltonian", (double)s.size * s.amp_bytes() * (1.0 + (out ? 1.0 : 0.0) + H.groups.size()));
    if (s.dtype == QFVM_C128)
        hipLaunchKernelGGL(k_hamiltonian<double>, dim3(grid), dim3(RED_BLOCK), 0, s.stream,
                           reinterpret_cast<const double2*>(s.psi), reinterpret_cast<double2*>(out), s.size, g,
                           (int)H.groups.size(), t, partial);
    else
        hipLaunchKernelGGL(k_hamiltonian<float>, dim3(grid), dim3(RED_BLOCK), 0, s.stream,
                           reinterpret_cast<const float2*>(s.psi), reinterpret_cast<float2*>(out), s.size, g,
                           (int)H.groups.size(), t, partial);
    QFVM_LAUNCHED();
    if (dev_result) {  // leave the sum on the device; the caller reads it after its own sync
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(RED_BLOCK), 0, s.stream, partial, (int)grid, dev_result);
        QFVM_LAUNCHED();
        return 0.0;
    }
    return finish_sum(s, partial, grid);
}

void launch_grad_marker(State& s, const Gate& g, const void* lam, double* grads_dev) {
    double* partial = reinterpret_cast<double*>(s.ensure_scratch((RED_GRID + 1) * sizeof(double)));
    const unsigned grid = std::min<unsigned>(RED_GRID, grid_for(s.size, RED_BLOCK));
    KernelScope ks(s, "grad_marker", 2.0 * (double)s.size * s.amp_bytes());
    if (s.dtype == QFVM_C128)
        hipLaunchKernelGGL(k_grad_marker<double>, dim3(grid), dim3(RED_BLOCK), 0, s.stream,
                           reinterpret_cast<const double2*>(s.psi), reinterpret_cast<const double2*>(lam), s.size,
                           g.gen, g.target, g.cmask, partial);
    else
        hipLaunchKernelGGL(k_grad_marker<float>, dim3(grid), dim3(RED_BLOCK), 0, s.stream,
                           reinterpret_cast<const float2*>(s.psi), reinterpret_cast<const float2*>(lam), s.size,
                           g.gen, g.target, g.cmask, partial);
    QFVM_LAUNCHED();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(RED_BLOCK), 0, s.stream, partial, (int)grid, grads_dev + g.slot);
    QFVM_LAUNCHED();
}

void launch_column_sums(State& s, const double* partials, uint64_t tiles, int cols, const int* slot_of,
                        double* grads_dev) {
    if (cols <= 0) return;
    hipLaunchKernelGGL(k_column_sums, dim3((unsigned)cols), dim3(RED_BLOCK), 0, s.stream, partials, tiles, slot_of,
                       grads_dev);
    QFVM_LAUNCHED();
}

// |K psi|^2 for a one-qubit operator K on qubit q: sum over pairs of |k00 a + k01 b|^2 + |k10 a + k11 b|^2
template <typename R>
__global__ void __launch_bounds__(RED_BLOCK) k_op_norm2(const V2<R>* __restrict__ psi, uint64_t pairs, int q,
                                                         double2 k00, double2 k01, double2 k10, double2 k11,
                                                         double* partial) {
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t low = (1ull << q) - 1;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < pairs; j += stride) {
        const uint64_t i0 = (j & low) | ((j & ~low) << 1);
        const V2<R> a = psi[i0], b = psi[i0 | (1ull << q)];
        const double ax = a.x, ay = a.y, bx = b.x, by = b.y;
        const double u = k00.x * ax - k00.y * ay + k01.x * bx - k01.y * by;
        const double v = k00.x * ay + k00.y * ax + k01.x * by + k01.y * bx;
        const double w = k10.x * ax - k10.y * ay + k11.x * bx - k11.y * by;
        const double z = k10.x * ay + k10.y * ax + k11.x * by + k11.y * bx;
        acc += u * u + v * v + w * w + z * z;
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

double reduce_op_norm2(State& s, int q, const cd m[4]) {
    double* partial = reinterpret_cast<double*>(s.ensure_scratch((RED_GRID + 1) * sizeof(double)));
    const uint64_t pairs = s.size >> 1;
    const unsigned grid = std::min<unsigned>(RED_GRID, grid_for(pairs, RED_BLOCK));
    const double2 k00 = make_double2(m[0].real(), m[0].imag()), k01 = make_double2(m[1].real(), m[1].imag());
    const double2 k10 = make_double2(m[2].real(), m[2].imag()), k11 = make_double2(m[3].real(), m[3].imag());
    if (s.dtype == QFVM_C128)
        hipLaunchKernelGGL(k_op_norm2<double>, dim3(grid), dim3(RED_BLOCK), 0, s.stream,
                           reinterpret_cast<const double2*>(s.psi), pairs, q, k00, k01, k10, k11, partial);
    else
        hipLaunchKernelGGL(k_op_norm2<float>, dim3(grid), dim3(RED_BLOCK), 0, s.stream,
                           reinterpret_cast<const float2*>(s.psi), pairs, q, k00, k01, k10, k11, partial);
    QFVM_LAUNCHED();
    return finish_sum(s, partial, grid);
}

// ------------------------------------------------------------------------------------------------
// Amplitudes at given indices, and the overlap <a|b> of two states of one size (any dtypes).
template <typename R>
__global__ void k_gather(const V2<R>* __restrict__ psi, const uint64_t* __restrict__ idx, uint64_t n,
                         double2* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const V2<R> a = psi[idx[k]];
        out[k] = make_double2((double)a.x, (double)a.y);
    }
}

template <typename RA, typename RB>
__global__ void __launch_bounds__(RED_BLOCK) k_inner(const V2<RA>* __restrict__ a, const V2<RB>* __restrict__ b,
                                                      uint64_t size, double* partial) {
    double re = 0.0, im = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) {
        const V2<RA> x = a[i];
        const V2<RB> y = b[i];
        re += (double)x.x * (double)y.x + (double)x.y * (double)y.y;  // conj(x) y
        im += (double)x.x * (double)y.y - (double)x.y * (double)y.x;
    }
    re = block_sum(re);
    im = block_sum(im);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = re;
        partial[gridDim.x + blockIdx.x] = im;
    }
}

void gather_amplitudes(State& s, const uint64_t* host_idx, uint64_t n, double* host_out) {
    const uint64_t chunk = 1ull << 22;
    for (uint64_t b = 0; b < n; b += chunk) {
        const uint64_t m = std::min(chunk, n - b);
        unsigned char* dev = reinterpret_cast<unsigned char*>(s.ensure_scratch(m * (sizeof(uint64_t) + 16)));
        uint64_t* didx = reinterpret_cast<uint64_t*>(dev);
        double2* dout = reinterpret_cast<double2*>(dev + m * sizeof(uint64_t));
        QFVM_HIP(hipMemcpyAsync(didx, host_idx + b, m * sizeof(uint64_t), hipMemcpyHostToDevice, s.stream));
        if (s.dtype == QFVM_C128)
            hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(m, 256)), dim3(256), 0, s.stream,
                               reinterpret_cast<const double2*>(s.psi), didx, m, dout);
        else
            hipLaunchKernelGGL(k_gather<float>, dim3(grid_for(m, 256)), dim3(256), 0, s.stream,
                               reinterpret_cast<const float2*>(s.psi), didx, m, dout);
        QFVM_LAUNCHED();
        QFVM_HIP(hipMemcpyAsync(host_out + 2 * b, dout, m * 16, hipMemcpyDeviceToHost, s.stream));
        QFVM_HIP(hipStreamSynchronize(s.stream));
    }
}

void reduce_inner(State& a, State& b, double* re, double* im) {
    double* partial = reinterpret_cast<double*>(a.ensure_scratch((2 * RED_GRID + 2) * sizeof(double)));
    const unsigned grid = std::min<unsigned>(RED_GRID, grid_for(a.size, RED_BLOCK));
    // b's stream finishes its work first; the kernel runs on a's stream
    QFVM_HIP(hipStreamSynchronize(b.stream));
    if (a.dtype == QFVM_C128 && b.dtype == QFVM_C128)
        hipLaunchKernelGGL((k_inner<double, double>), dim3(grid), dim3(RED_BLOCK), 0, a.stream,
                           reinterpret_cast<const double2*>(a.psi), reinterpret_cast<const double2*>(b.psi), a.size, partial);
    else if (a.dtype == QFVM_C128)
        hipLaunchKernelGGL((k_inner<double, float>), dim3(grid), dim3(RED_BLOCK), 0, a.stream,
                           reinterpret_cast<const double2*>(a.psi), reinterpret_cast<const float2*>(b.psi), a.size, partial);
    else if (b.dtype == QFVM_C128)
        hipLaunchKernelGGL((k_inner<float, double>), dim3(grid), dim3(RED_BLOCK), 0, a.stream,
                           reinterpret_cast<const float2*>(a.psi), reinterpret_cast<const double2*>(b.psi), a.size, partial);
    else
        hipLaunchKernelGGL((k_inner<float, float>), dim3(grid), dim3(RED_BLOCK), 0, a.stream,
                           reinterpret_cast<const float2*>(a.psi), reinterpret_cast<const float2*>(b.psi), a.size, partial);
    QFVM_LAUNCHED();
    double* out = partial + 2 * grid;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(RED_BLOCK), 0, a.stream, partial, (int)grid, out);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(RED_BLOCK), 0, a.stream, partial + grid, (int)grid, out + 1);
    QFVM_LAUNCHED();
    double* h = a.ensure_pinned(2 * sizeof(double));
    QFVM_HIP(hipMemcpyAsync(h, out, 2 * sizeof(double), hipMemcpyDeviceToHost, a.stream));
    QFVM_HIP(hipStreamSynchronize(a.stream));
    *re = h[0];
    *im = h[1];
}

void launch_swap_states(State& s, const std::vector<int>& qbits, uint64_t a, uint64_t b) {
    uint64_t qmask = 0, fmask = 0, mmask = 0;
    for (size_t k = 0; k < qbits.size(); k++) {
        qmask |= 1ull << qbits[k];
        if ((a >> k) & 1) fmask |= 1ull << qbits[k];
        if ((b >> k) & 1) mmask |= 1ull << qbits[k];
    }
    const unsigned grid = grid_for(s.size, 256);
    if (s.dtype == QFVM_C128)
        hipLaunchKernelGGL(k_swap_states<double>, dim3(grid), dim3(256), 0, s.stream,
                           reinterpret_cast<double2*>(s.psi), s.size, qmask, fmask, mmask);
    else
        hipLaunchKernelGGL(k_swap_states<float>, dim3(grid), dim3(256), 0, s.stream,
                           reinterpret_cast<float2*>(s.psi), s.size, qmask, fmask, mmask);
    QFVM_LAUNCHED();
}

}  // namespace qfvm
