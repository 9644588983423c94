// Shot sampler: pairwise sum tree over p_i = |amp_i|^2 plus a descent per uniform.
//
// Replaces measure_samples (src/qfvm/statevector.hpp:1380-1412), whose serial running-sum scan
// costs O(shots * 2^n). The tree makes the sampled index a fixed function of (state, u), so the
// CPU oracle reproduces it bit for bit:
//   p_i = fma(re, re, im*im) in double; S_0 = p; S_l[j] = S_{l-1}[2j] + S_{l-1}[2j+1];
//   u >= S_n[0] -> 2^n (all measured bits zero, as in the reference);
//   otherwise descend from the root with base = 0: a zero-sum child is never entered; go left when
//   u < base + L, else base += L and go right.
// Levels 0..LB-1 (LB = 5: 32-amplitude leaf blocks) are recomputed from the state during the
// descent; levels LB..n are stored (2^(n-LB+1) doubles, 1/32 of a complex128 state's size).
#include <algorithm>

#include "common.h"

namespace qfvm {

constexpr int LB_MAX = 5;

template <typename R>
__device__ __forceinline__ double prob_at(const V2<R>* __restrict__ psi, uint64_t i) {
    const V2<R> a = psi[i];
    const double re = (double)a.x, im = (double)a.y;
    return __fma_rn(re, re, __dmul_rn(im, im));
}

// exact pairwise tree sum of w <= 16 consecutive probabilities (w a power of two); padding with
// zeros leaves every partial sum unchanged, so one unrolled 16-wide tree serves all widths.
template <typename R>
__device__ __forceinline__ double tree_sum16(const V2<R>* __restrict__ psi, uint64_t a, int w) {
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = (i < w) ? prob_at<R>(psi, a + i) : 0.0;
#pragma unroll
    for (int h = 8; h >= 1; h >>= 1) {
#pragma unroll
        for (int i = 0; i < h; i++) v[i] = __dadd_rn(v[2 * i], v[2 * i + 1]);
    }
    return v[0];
}

template <typename R>
__global__ void __launch_bounds__(256) k_leaf_sums(const V2<R>* __restrict__ psi, uint64_t nblocks, int lb,
                                                   double* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const int w = 1 << lb;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nblocks; b += stride) {
        const uint64_t a = b << lb;
        double s;
        if (w == 32)
            s = __dadd_rn(tree_sum16<R>(psi, a, 16), tree_sum16<R>(psi, a + 16, 16));
        else
            s = tree_sum16<R>(psi, a, w);
        out[b] = s;
    }
}

__global__ void __launch_bounds__(256) k_level_up(const double* __restrict__ lo, double* __restrict__ hi, uint64_t nhi) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nhi; j += stride)
        hi[j] = __dadd_rn(lo[2 * j], lo[2 * j + 1]);
}

struct TreeSpec {
    int n, lb;
    uint64_t off[36];  // off[l] = start of level l (l = lb..n) in the tree buffer
};

template <typename R>
__global__ void __launch_bounds__(256) k_descend(const V2<R>* __restrict__ psi, const double* __restrict__ tree,
                                                 TreeSpec ts, const double* __restrict__ u, int64_t shots,
                                                 uint64_t* __restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= shots) return;
    const double x = u[s];
    const double root = tree[ts.off[ts.n]];
    if (!(x < root)) {
        out[s] = 1ull << ts.n;
        return;
    }
    double base = 0.0;
    uint64_t idx = 0;
    for (int l = ts.n; l > ts.lb; l--) {
        const double* lv = tree + ts.off[l - 1];
        const double L = lv[2 * idx], Rt = lv[2 * idx + 1];
        if (Rt == 0.0) {
            idx = 2 * idx;
        } else if (L == 0.0) {
            idx = 2 * idx + 1;
        } else {
            const double t = __dadd_rn(base, L);
            if (x < t) {
                idx = 2 * idx;
            } else {
                base = t;
                idx = 2 * idx + 1;
            }
        }
    }
    // inside the leaf block: child sums recomputed from the amplitudes
    uint64_t lo = idx << ts.lb;
    for (int w = 1 << ts.lb; w > 1; w >>= 1) {
        const int h = w >> 1;
        const double L = tree_sum16<R>(psi, lo, h);
        const double Rt = tree_sum16<R>(psi, lo + h, h);
        if (Rt == 0.0) {
            // left
        } else if (L == 0.0) {
            lo += h;
        } else {
            const double t = __dadd_rn(base, L);
            if (x < t) {
                // left
            } else {
                base = t;
                lo += h;
            }
        }
    }
    out[s] = lo;
}

template <typename R>
static void sample_typed(State& s, const double* host_u, int64_t shots, uint64_t* host_out) {
    const int n = s.n;
    const int lb = std::min(LB_MAX, n);
    TreeSpec ts{};
    ts.n = n;
    ts.lb = lb;
    uint64_t total = 0;
    for (int l = lb; l <= n; l++) {
        ts.off[l] = total;
        total += 1ull << (n - l);
    }
    const size_t tree_bytes = total * sizeof(double);
    const size_t u_bytes = (size_t)shots * sizeof(double);
    const size_t o_bytes = (size_t)shots * sizeof(uint64_t);
    auto align = [](size_t b) { return (b + 255) & ~size_t(255); };
    char* base = reinterpret_cast<char*>(s.ensure_scratch(align(tree_bytes) + align(u_bytes) + align(o_bytes)));
    double* tree = reinterpret_cast<double*>(base);
    double* du = reinterpret_cast<double*>(base + align(tree_bytes));
    uint64_t* dout = reinterpret_cast<uint64_t*>(base + align(tree_bytes) + align(u_bytes));
    const V2<R>* psi = reinterpret_cast<const V2<R>*>(s.psi);

    const uint64_t nblocks = 1ull << (n - lb);
    KernelScope ks(s, "sample_tree_build", (double)s.size * s.amp_bytes());
    hipLaunchKernelGGL(k_leaf_sums<R>, dim3(grid_for(nblocks, 256)), dim3(256), 0, s.stream, psi, nblocks, lb,
                       tree + ts.off[lb]);
    QFVM_LAUNCHED();
    for (int l = lb + 1; l <= n; l++) {
        const uint64_t nhi = 1ull << (n - l);
        hipLaunchKernelGGL(k_level_up, dim3(grid_for(nhi, 256)), dim3(256), 0, s.stream, tree + ts.off[l - 1],
                           tree + ts.off[l], nhi);
        QFVM_LAUNCHED();
    }
    QFVM_HIP(hipMemcpyAsync(du, host_u, u_bytes, hipMemcpyHostToDevice, s.stream));
    hipLaunchKernelGGL(k_descend<R>, dim3((unsigned)((shots + 255) / 256)), dim3(256), 0, s.stream, psi, tree, ts, du,
                       shots, dout);
    QFVM_LAUNCHED();
    QFVM_HIP(hipMemcpyAsync(host_out, dout, o_bytes, hipMemcpyDeviceToHost, s.stream));
    QFVM_HIP(hipStreamSynchronize(s.stream));
}

void sample_tree(State& s, const double* host_u, int64_t shots, uint64_t* host_out) {
    if (shots <= 0) return;
    if (s.dtype == QFVM_C128)
        sample_typed<double>(s, host_u, shots, host_out);
    else
        sample_typed<float>(s, host_u, shots, host_out);
}

}  // namespace qfvm
