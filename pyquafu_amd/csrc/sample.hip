// Shot sampler with qfvm's own rule, evaluated exactly and in parallel.
//
// measure_samples (src/qfvm/statevector.hpp:1380-1412) gives, for each uniform u, the first index k
// whose running sum S_k = fl(S_{k-1} + p_k), p_k = |amp_k|^2, exceeds u (2^n when none does), by
// rescanning the state from index 0 for every shot: O(shots * 2^n). Here S is evaluated once, at
// every 32-element block end, bit for bit equal to that sequential sum, and each shot is a binary
// search over the block ends plus at most 32 sequential additions inside its block.
//
// p_k is what g++ emits for the reference's std::real(d * std::conj(d)) under -O3 -mfma
// -ffast-math: fma(re, re, im * im) in double. The block ends come from integer arithmetic:
// * while S stays in one binade [2^e, 2^(e+1)), S = m 2^(e-52) for an integer m < 2^53 and
//   fl(S + p) = (m + rne(p / 2^(e-52))) 2^(e-52): an integer step of floor(p / ulp), plus one when
//   the fraction is above 1/2; at exactly 1/2 plus one only when that makes m even (ties to even),
//   so the step depends on the parity of the incoming m;
// * a block's effect is therefore a pair (w0, w1), the increment for an even / odd incoming m, and
//   pairs compose associatively: (f then g)_p = f_p + g_{(p + f_p) mod 2};
// * per chunk of blocks (growing geometrically), one kernel forms every block's pair and every
//   tile's pair (64 blocks), one workgroup scans the tiles, one kernel turns the prefixes into m at
//   each block end and reports the first block whose m could reach 2^53 - 1 (the sum may leave the
//   binade there); that block is replayed with double additions from its exact start, and the next
//   chunk starts in the new binade. Blocks before the first nonzero p end at 0.
// Algorithmic bytes of a sample call: about one read of the state (s 2^n) plus 8 bytes per 32
// amplitudes for the block ends, plus shots * (8 + 32 s) for the searches.
#include <algorithm>
#include <climits>
#include <cmath>

#include "common.h"

namespace qfvm {

constexpr int SB = 32;            // amplitudes per block
constexpr int TILE_BLOCKS = 64;   // blocks per tile of the scan (one 256-thread workgroup)
constexpr int64_t XSAT = 1ll << 60;
constexpr int64_t M_CROSS = (1ll << 53) - 1;
constexpr uint64_t NONE = ~0ull;

struct Xfer {
    int64_t w0, w1;  // increment of m for an even / odd incoming m (saturated at XSAT)
};

struct SeqStatus {
    uint64_t first;     // first index with p > 0
    uint64_t crossing;  // first block of the chunk whose m could leave the binade
    uint64_t next;      // first block not yet evaluated
    double s;           // exact S at the end of block next - 1
};

template <typename R>
__device__ __forceinline__ double prob_at(const V2<R>* __restrict__ psi, uint64_t i) {
    const V2<R> a = psi[i];
    const double re = (double)a.x, im = (double)a.y;
    return __fma_rn(re, re, __dmul_rn(im, im));
}

__device__ __forceinline__ Xfer xcompose(Xfer f, Xfer g) {
    Xfer r;
    r.w0 = min(f.w0 + ((f.w0 & 1) ? g.w1 : g.w0), XSAT);
    r.w1 = min(f.w1 + (((f.w1 + 1) & 1) ? g.w1 : g.w0), XSAT);
    return r;
}

// the step of one p in the binade of ulp 2^-sc (lim = the binade's upper end)
__device__ __forceinline__ Xfer xelem(double p, int sc, double lim) {
    if (!(p < lim)) return Xfer{XSAT, XSAT};
    const double r = ldexp(p, sc);  // exact: a power-of-two scaling, r < 2^53
    const double fl = floor(r);
    const double f = r - fl;
    const int64_t q = (int64_t)fl;
    if (f > 0.5) return Xfer{q + 1, q + 1};
    if (f < 0.5) return Xfer{q, q};
    return Xfer{q + (q & 1), q + ((q + 1) & 1)};
}

__global__ void k_seq_init(SeqStatus* st) {
    if (threadIdx.x == 0) {
        st->first = NONE;
        st->crossing = NONE;
        st->next = 0;
        st->s = 0.0;
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_first_nonzero(const V2<R>* __restrict__ psi, uint64_t size, SeqStatus* st) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t best = NONE;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) {
        if (prob_at<R>(psi, i) > 0.0) {
            best = i;
            break;  // later indices of this thread are larger
        }
    }
    if (best != NONE) atomicMin(reinterpret_cast<unsigned long long*>(&st->first), (unsigned long long)best);
}

// Pairs of blocks [blk0, blk0 + nb) (one per block, chunk-relative in bx) and of their tiles (tx).
// A wavefront covers two blocks per row: lane k of a 32-lane half reads element k of its block, so
// the loads are contiguous; the halves reduce their pairs in index order over five shuffles.
template <typename R>
__global__ void __launch_bounds__(256) k_seq_blocks(const V2<R>* __restrict__ psi, uint64_t blk0, uint64_t nb, int sc,
                                                    double lim, Xfer* __restrict__ bx, Xfer* __restrict__ tx) {
    __shared__ Xfer bl[TILE_BLOCKS];
    const int tid = threadIdx.x, k = tid & (SB - 1);
    for (int r = 0; r < TILE_BLOCKS / 8; r++) {
        const int tb = r * 8 + (tid >> 5);
        const uint64_t cb = (uint64_t)blockIdx.x * TILE_BLOCKS + tb;
        Xfer x{0, 0};
        if (cb < nb) x = xelem(prob_at<R>(psi, (blk0 + cb) * SB + k), sc, lim);
#pragma unroll
        for (int off = 1; off < SB; off <<= 1) {
            Xfer o;
            o.w0 = __shfl_down(x.w0, off, 64);
            o.w1 = __shfl_down(x.w1, off, 64);
            if ((k & (2 * off - 1)) == 0) x = xcompose(x, o);
        }
        if (k == 0) {
            bl[tb] = x;  // identity for blocks past the chunk
            if (cb < nb) bx[cb] = x;
        }
    }
    __syncthreads();
    if (tid == 0) {
        Xfer t{0, 0};
        for (int j = 0; j < TILE_BLOCKS; j++) t = xcompose(t, bl[j]);
        tx[blockIdx.x] = t;
    }
}

// exclusive prefix of the tile pairs, one workgroup of 1024 threads
__global__ void __launch_bounds__(1024) k_seq_tile_scan(const Xfer* __restrict__ tx, uint64_t ntiles,
                                                        Xfer* __restrict__ tpre) {
    __shared__ Xfer part[1024];
    const int tid = threadIdx.x;
    const uint64_t per = (ntiles + 1023) / 1024;
    const uint64_t a = (uint64_t)tid * per, b = min(a + per, ntiles);
    Xfer x{0, 0};
    for (uint64_t i = a; i < b; i++) x = xcompose(x, tx[i]);
    part[tid] = x;
    __syncthreads();
    if (tid == 0) {
        Xfer acc{0, 0};
        for (int j = 0; j < 1024; j++) {
            const Xfer v = part[j];
            part[j] = acc;
            acc = xcompose(acc, v);
        }
    }
    __syncthreads();
    Xfer pre = part[tid];
    for (uint64_t i = a; i < b; i++) {
        tpre[i] = pre;
        pre = xcompose(pre, tx[i]);
    }
}

// m at every block end of the chunk (one thread per tile), written to B as the exact double
// S = m 2^(e-52); the first block whose m reaches 2^53 - 1 is reported and nothing at or after it
// is written
__global__ void __launch_bounds__(64) k_seq_apply(const Xfer* __restrict__ bx, const Xfer* __restrict__ tpre,
                                                  uint64_t blk0, uint64_t nb, uint64_t ntiles, int64_t m0, int e52,
                                                  double* __restrict__ B, SeqStatus* st) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const Xfer pre = tpre[t];
    // (a tile that starts at or past 2^53 - 1 reports its first block; the atomic keeps the earliest)
    int64_t m = m0 + ((m0 & 1) ? pre.w1 : pre.w0);
    const uint64_t j1 = min((t + 1) * TILE_BLOCKS, nb);
    for (uint64_t j = t * TILE_BLOCKS; j < j1; j++) {
        const Xfer b = bx[j];
        m += (m & 1) ? b.w1 : b.w0;
        if (m >= M_CROSS) {
            atomicMin(reinterpret_cast<unsigned long long*>(&st->crossing), (unsigned long long)(blk0 + j));
            return;
        }
        B[blk0 + j] = ldexp((double)m, e52);
    }
}

// Closes a chunk: with no crossing, the chunk's last block end is the new start; otherwise (and for
// a block evaluated alone, seq_only) the block is replayed with double additions from its exact
// start. Writes the status the host reads.
template <typename R>
__global__ void k_seq_finish(const V2<R>* __restrict__ psi, uint64_t chunk_start, uint64_t chunk_end, double s_start,
                             int seq_only, int w, double* __restrict__ B, SeqStatus* st) {
    if (threadIdx.x != 0) return;
    const uint64_t c = seq_only ? chunk_start : st->crossing;
    if (c == NONE) {
        st->next = chunk_end;
        st->s = B[chunk_end - 1];
    } else {
        double s = c > chunk_start ? B[c - 1] : s_start;
        for (int k = 0; k < w; k++) s = __dadd_rn(s, prob_at<R>(psi, c * (uint64_t)w + k));
        B[c] = s;
        st->next = c + 1;
        st->s = s;
    }
    st->crossing = NONE;
}

__global__ void __launch_bounds__(256) k_fill_value(double* __restrict__ B, uint64_t a, uint64_t b, double v) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = a + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < b; j += stride) B[j] = v;
}

// one shot per thread: first block whose end exceeds u, then the sequential sum inside it
template <typename R>
__global__ void __launch_bounds__(256) k_seq_descend(const V2<R>* __restrict__ psi, const double* __restrict__ B,
                                                     uint64_t nblk, int w, int n, const double* __restrict__ u,
                                                     int64_t shots, uint64_t* __restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= shots) return;
    const double x = u[s];
    if (!(x < B[nblk - 1])) {
        out[s] = 1ull << n;
        return;
    }
    uint64_t lo = 0, hi = nblk - 1;  // B[hi] > x
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (B[mid] > x)
            hi = mid;
        else
            lo = mid + 1;
    }
    double acc = lo > 0 ? B[lo - 1] : 0.0;
    const uint64_t base = lo * (uint64_t)w;
    uint64_t res = base + (uint64_t)w - 1;
    for (int k = 0; k < w; k++) {
        acc = __dadd_rn(acc, prob_at<R>(psi, base + k));
        if (x < acc) {
            res = base + k;
            break;
        }
    }
    out[s] = res;
}

template <typename R>
static void sample_typed(State& s, const double* host_u, int64_t shots, uint64_t* host_out) {
    const int n = s.n;
    const uint64_t size = s.size;
    const int w = size >= (uint64_t)SB ? SB : (int)size;
    const uint64_t nblk = size / (uint64_t)w;
    const uint64_t chunk_max = std::min<uint64_t>(nblk, 1ull << 22);
    const uint64_t tiles_max = (chunk_max + TILE_BLOCKS - 1) / TILE_BLOCKS;
    auto align = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t b_bytes = align(nblk * sizeof(double));
    const size_t bx_bytes = align(chunk_max * sizeof(Xfer));
    const size_t tx_bytes = align(tiles_max * sizeof(Xfer));
    const size_t st_bytes = align(sizeof(SeqStatus));
    const size_t u_bytes = align((size_t)shots * sizeof(double));
    const size_t o_bytes = align((size_t)shots * sizeof(uint64_t));
    char* base = reinterpret_cast<char*>(
        s.ensure_scratch(b_bytes + bx_bytes + 2 * tx_bytes + st_bytes + u_bytes + o_bytes));
    double* B = reinterpret_cast<double*>(base);
    Xfer* bx = reinterpret_cast<Xfer*>(base + b_bytes);
    Xfer* tx = reinterpret_cast<Xfer*>(base + b_bytes + bx_bytes);
    Xfer* tpre = reinterpret_cast<Xfer*>(base + b_bytes + bx_bytes + tx_bytes);
    SeqStatus* st = reinterpret_cast<SeqStatus*>(base + b_bytes + bx_bytes + 2 * tx_bytes);
    double* du = reinterpret_cast<double*>(base + b_bytes + bx_bytes + 2 * tx_bytes + st_bytes);
    uint64_t* dout = reinterpret_cast<uint64_t*>(base + b_bytes + bx_bytes + 2 * tx_bytes + st_bytes + u_bytes);
    SeqStatus* hst = reinterpret_cast<SeqStatus*>(s.ensure_pinned(sizeof(SeqStatus)));
    const V2<R>* psi = reinterpret_cast<const V2<R>*>(s.psi);
    auto read_status = [&] {
        QFVM_HIP(hipMemcpyAsync(hst, st, sizeof(SeqStatus), hipMemcpyDeviceToHost, s.stream));
        QFVM_HIP(hipStreamSynchronize(s.stream));
    };

    {
        KernelScope ks(s, "sample_scan", (double)size * s.amp_bytes() + (double)nblk * sizeof(double));
        hipLaunchKernelGGL(k_seq_init, dim3(1), dim3(64), 0, s.stream, st);
        QFVM_LAUNCHED();
        hipLaunchKernelGGL(k_first_nonzero<R>, dim3(grid_for(size, 256, 1ull << 14)), dim3(256), 0, s.stream, psi, size,
                           st);
        QFVM_LAUNCHED();
        read_status();
        const uint64_t b0 = hst->first == NONE ? nblk : hst->first / (uint64_t)w;
        if (b0 > 0) launch_zero(s, B, b0 * sizeof(double));
        uint64_t next = nblk;
        double S = 0.0;
        auto finish = [&](uint64_t start, uint64_t end, double s_start, int seq_only) {
            hipLaunchKernelGGL(k_seq_finish<R>, dim3(1), dim3(64), 0, s.stream, psi, start, end, s_start, seq_only, w, B,
                               st);
            QFVM_LAUNCHED();
            read_status();
            next = hst->next;
            S = hst->s;
        };
        if (b0 < nblk) finish(b0, b0 + 1, 0.0, 1);
        const uint64_t min_chunk = 1ull << 12;
        while (next < nblk) {
            if (!std::isfinite(S)) {  // inf / NaN amplitudes: the sum stays where it is
                hipLaunchKernelGGL(k_fill_value, dim3(grid_for(nblk - next, 256)), dim3(256), 0, s.stream, B, next, nblk,
                                   S);
                QFVM_LAUNCHED();
                break;
            }
            if (!(S >= 2.2250738585072014e-308)) {  // subnormal start: one block at a time
                finish(next, next + 1, S, 1);
                continue;
            }
            const int e = std::ilogb(S);
            const int64_t m0 = (int64_t)std::ldexp(S, 52 - e);
            const uint64_t len = std::min(std::max(min_chunk, next - b0), chunk_max);
            const uint64_t end = std::min(nblk, next + len);
            const uint64_t nb = end - next;
            const uint64_t ntiles = (nb + TILE_BLOCKS - 1) / TILE_BLOCKS;
            hipLaunchKernelGGL(k_seq_blocks<R>, dim3((unsigned)ntiles), dim3(256), 0, s.stream, psi, next, nb, 52 - e,
                               std::ldexp(1.0, e + 1), bx, tx);
            QFVM_LAUNCHED();
            hipLaunchKernelGGL(k_seq_tile_scan, dim3(1), dim3(1024), 0, s.stream, tx, ntiles, tpre);
            QFVM_LAUNCHED();
            hipLaunchKernelGGL(k_seq_apply, dim3((unsigned)((ntiles + 63) / 64)), dim3(64), 0, s.stream, bx, tpre, next,
                               nb, ntiles, m0, e - 52, B, st);
            QFVM_LAUNCHED();
            finish(next, end, S, 0);
        }
    }
    if (shots <= 0) return;
    QFVM_HIP(hipMemcpyAsync(du, host_u, (size_t)shots * sizeof(double), hipMemcpyHostToDevice, s.stream));
    {
        KernelScope ks(s, "sample_search", (double)shots * (8.0 + SB * (double)s.amp_bytes()));
        hipLaunchKernelGGL(k_seq_descend<R>, dim3((unsigned)((shots + 255) / 256)), dim3(256), 0, s.stream, psi, B, nblk,
                           w, n, du, shots, dout);
        QFVM_LAUNCHED();
    }
    QFVM_HIP(hipMemcpyAsync(host_out, dout, (size_t)shots * sizeof(uint64_t), hipMemcpyDeviceToHost, s.stream));
    QFVM_HIP(hipStreamSynchronize(s.stream));
}

void sample_sequential_rule(State& s, const double* host_u, int64_t shots, uint64_t* host_out) {
    if (shots <= 0) return;
    if (s.dtype == QFVM_C128)
        sample_typed<double>(s, host_u, shots, host_out);
    else
        sample_typed<float>(s, host_u, shots, host_out);
}

}  // namespace qfvm
