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
    __shared__ uint64_t wmin[4];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t best = NONE;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) {
        if (prob_at<R>(psi, i) > 0.0) {
            best = i;
            break;  // later indices of this thread are larger
        }
    }
    // one atomic per workgroup: a dense state would otherwise send every thread to one address
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) best = min(best, (uint64_t)__shfl_xor((unsigned long long)best, o, 64));
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        best = min(min(wmin[0], wmin[1]), min(wmin[2], wmin[3]));
        if (best != NONE) atomicMin(reinterpret_cast<unsigned long long*>(&st->first), (unsigned long long)best);
    }
}

__device__ __forceinline__ Xfer shfl_down_x(Xfer x, int off) {
    return Xfer{__shfl_down(x.w0, off, 64), __shfl_down(x.w1, off, 64)};
}
__device__ __forceinline__ Xfer shfl_up_x(Xfer x, int off) {
    return Xfer{__shfl_up(x.w0, off, 64), __shfl_up(x.w1, off, 64)};
}

// Inclusive scan over the lanes of a wavefront in lane order (lane i gets lanes 0 .. i composed).
__device__ __forceinline__ Xfer wave_scan_x(Xfer x, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const Xfer o = shfl_up_x(x, off);
        if (lane >= off) x = xcompose(o, x);
    }
    return x;
}

// Composition of the pairs of `width` consecutive lanes in lane order, left in the first lane of each
// run of `width` (a power of two up to 64).
__device__ __forceinline__ Xfer lane_tree_x(Xfer x, int lane, int width) {
    for (int off = 1; off < width; off <<= 1) {
        const Xfer o = shfl_down_x(x, off);
        if ((lane & (2 * off - 1)) == 0) x = xcompose(x, o);
    }
    return x;
}

// Pairs of blocks [blk0, blk0 + nb) (one per block, chunk-relative in bx) and of their tiles (tx).
// The tile's 2048 probabilities are read with contiguous loads (element r 256 + tid per thread and
// row) into LDS; each thread then composes a run of 8 consecutive elements in order, four runs make
// a block (two shuffles) and the tile's 64 blocks are composed in order by wavefront 0 (six).
// Elements past the chunk have p = 0, the identity.
template <typename R>
__global__ void __launch_bounds__(256) k_seq_blocks(const V2<R>* __restrict__ psi, uint64_t blk0, uint64_t nb, int sc,
                                                    double lim, Xfer* __restrict__ bx, Xfer* __restrict__ tx) {
    constexpr int TE = TILE_BLOCKS * SB;  // elements per tile
    constexpr int RUN = 8;                // elements per thread's run
    constexpr int ROWS = TE / 256;
    static_assert(TE == 256 * RUN && SB == 4 * RUN, "one run per thread, four runs per block");
    __shared__ double pl[TE / RUN * (RUN + 1)];  // run-major, padded by one slot per run
    __shared__ Xfer bl[TILE_BLOCKS];
    const int tid = threadIdx.x;
    const uint64_t tb0 = (uint64_t)blockIdx.x * TILE_BLOCKS;  // first block of the tile in the chunk
    const uint64_t ne = min((uint64_t)TILE_BLOCKS, nb - tb0) * SB;
    const V2<R>* __restrict__ src = psi + (blk0 + tb0) * SB;
    double p[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        const int e = r * 256 + tid;
        p[r] = (uint64_t)e < ne ? prob_at<R>(src, e) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        const int e = r * 256 + tid;
        pl[(e / RUN) * (RUN + 1) + e % RUN] = p[r];
    }
    __syncthreads();
    Xfer x{0, 0};
#pragma unroll
    for (int j = 0; j < RUN; j++) x = xcompose(x, xelem(pl[tid * (RUN + 1) + j], sc, lim));
    x = lane_tree_x(x, tid & 63, SB / RUN);
    if ((tid & (SB / RUN - 1)) == 0) {
        const int b = tid / (SB / RUN);
        bl[b] = x;
        if (tb0 + b < nb) bx[tb0 + b] = x;
    }
    __syncthreads();
    if (tid < TILE_BLOCKS) {
        const Xfer t = lane_tree_x(bl[tid], tid, TILE_BLOCKS);
        if (tid == 0) tx[blockIdx.x] = t;
    }
}

// Pair of each group of SCAN_GROUP consecutive tiles (one 1024-thread workgroup per group).
constexpr int SCAN_GROUP = 1024;

__global__ void __launch_bounds__(SCAN_GROUP) k_seq_group_sums(const Xfer* __restrict__ tx, uint64_t ntiles,
                                                               Xfer* __restrict__ gx) {
    __shared__ Xfer wt[SCAN_GROUP / 64];
    const int tid = threadIdx.x, lane = tid & 63;
    const uint64_t t = (uint64_t)blockIdx.x * SCAN_GROUP + tid;
    Xfer x = t < ntiles ? tx[t] : Xfer{0, 0};
    x = lane_tree_x(x, lane, 64);
    if (lane == 0) wt[tid >> 6] = x;
    __syncthreads();
    if (tid < SCAN_GROUP / 64) {
        const Xfer g = lane_tree_x(wt[tid], tid, SCAN_GROUP / 64);
        if (tid == 0) gx[blockIdx.x] = g;
    }
}

// Exclusive prefix of every tile of the chunk: the group's prefix (its predecessors' pairs,
// composed by wavefront 0), then an inclusive scan of the group's tiles over shuffles and the 16
// wavefront totals.
__global__ void __launch_bounds__(SCAN_GROUP) k_seq_tile_prefix(const Xfer* __restrict__ tx, uint64_t ntiles,
                                                                const Xfer* __restrict__ gx, Xfer* __restrict__ tpre) {
    __shared__ Xfer wt[SCAN_GROUP / 64];
    __shared__ Xfer gpre;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint64_t g = blockIdx.x;
    if (wv == 0) {
        const uint64_t per = (g + 63) / 64;
        const uint64_t a = (uint64_t)lane * per, b = min(a + per, g);
        Xfer x{0, 0};
        for (uint64_t i = a; i < b; i++) x = xcompose(x, gx[i]);
        x = wave_scan_x(x, lane);
        if (lane == 63) gpre = x;
    }
    const uint64_t t = g * SCAN_GROUP + tid;
    const Xfer v = t < ntiles ? tx[t] : Xfer{0, 0};
    const Xfer inc = wave_scan_x(v, lane);
    Xfer ex = shfl_up_x(inc, 1);
    if (lane == 0) ex = Xfer{0, 0};
    if (lane == 63) wt[wv] = inc;
    __syncthreads();
    Xfer pre = gpre;
    for (int j = 0; j < wv; j++) pre = xcompose(pre, wt[j]);
    if (t < ntiles) tpre[t] = xcompose(pre, ex);
}

// m at every block end of the chunk, written to B as the exact double S = m 2^(e-52): one
// wavefront per tile, lane = block, an inclusive scan of the tile's block pairs from the tile's
// prefix. The first block whose m reaches 2^53 - 1 is reported (the atomic keeps the earliest over
// the tiles; a tile that starts at or past it reports its first block) and nothing at or after it
// is written.
__global__ void __launch_bounds__(256) k_seq_apply(const Xfer* __restrict__ bx, const Xfer* __restrict__ tpre,
                                                   uint64_t blk0, uint64_t nb, uint64_t ntiles, int64_t m0, int e52,
                                                   double* __restrict__ B, SeqStatus* st) {
    const int lane = threadIdx.x & 63;
    const uint64_t t = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= ntiles) return;  // whole wavefronts
    const Xfer pre = tpre[t];
    const int64_t mt = m0 + ((m0 & 1) ? pre.w1 : pre.w0);
    const uint64_t cb = t * TILE_BLOCKS + lane;
    const bool in = cb < nb;
    const Xfer inc = wave_scan_x(in ? bx[cb] : Xfer{0, 0}, lane);
    const int64_t m = mt + ((mt & 1) ? inc.w1 : inc.w0);  // nondecreasing over the lanes
    const uint64_t cross = __ballot(in && m >= M_CROSS);
    const int first = cross ? __ffsll((unsigned long long)cross) - 1 : 64;
    // (tiles past a reported crossing skip the atomic: one address would serialise them all)
    if (lane == first && blk0 + cb < *reinterpret_cast<volatile uint64_t*>(&st->crossing))
        atomicMin(reinterpret_cast<unsigned long long*>(&st->crossing), (unsigned long long)(blk0 + cb));
    if (in && lane < first) B[blk0 + cb] = ldexp((double)m, e52);
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
    const uint64_t chunk_max = std::min<uint64_t>(nblk, 1ull << 24);
    const uint64_t tiles_max = (chunk_max + TILE_BLOCKS - 1) / TILE_BLOCKS;
    const uint64_t groups_max = (tiles_max + SCAN_GROUP - 1) / SCAN_GROUP;
    auto align = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t b_bytes = align(nblk * sizeof(double));
    const size_t bx_bytes = align(chunk_max * sizeof(Xfer));
    const size_t tx_bytes = align(tiles_max * sizeof(Xfer));
    const size_t gx_bytes = align(groups_max * sizeof(Xfer));
    const size_t st_bytes = align(sizeof(SeqStatus));
    const size_t u_bytes = align((size_t)shots * sizeof(double));
    const size_t o_bytes = align((size_t)shots * sizeof(uint64_t));
    char* base = reinterpret_cast<char*>(
        s.ensure_scratch(b_bytes + bx_bytes + 2 * tx_bytes + gx_bytes + st_bytes + u_bytes + o_bytes));
    double* B = reinterpret_cast<double*>(base);
    Xfer* bx = reinterpret_cast<Xfer*>(base + b_bytes);
    Xfer* tx = reinterpret_cast<Xfer*>(base + b_bytes + bx_bytes);
    Xfer* tpre = reinterpret_cast<Xfer*>(base + b_bytes + bx_bytes + tx_bytes);
    Xfer* gx = reinterpret_cast<Xfer*>(base + b_bytes + bx_bytes + 2 * tx_bytes);
    char* tail = base + b_bytes + bx_bytes + 2 * tx_bytes + gx_bytes;
    SeqStatus* st = reinterpret_cast<SeqStatus*>(tail);
    double* du = reinterpret_cast<double*>(tail + st_bytes);
    uint64_t* dout = reinterpret_cast<uint64_t*>(tail + st_bytes + u_bytes);
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
        hipLaunchKernelGGL(k_first_nonzero<R>, dim3(grid_for(size, 256, 2048)), dim3(256), 0, s.stream, psi, size,
                           st);
        QFVM_LAUNCHED();
        read_status();
        const uint64_t b0 = hst->first == NONE ? nblk : hst->first / (uint64_t)w;
        if (b0 > 0) launch_zero(s, B, b0 * sizeof(double));
        uint64_t next = nblk;
        double S = 0.0;
        double rate = 0.0;  // growth of S per block over the last evaluated span
        auto finish = [&](uint64_t start, uint64_t end, double s_start, int seq_only) {
            hipLaunchKernelGGL(k_seq_finish<R>, dim3(1), dim3(64), 0, s.stream, psi, start, end, s_start, seq_only, w, B,
                               st);
            QFVM_LAUNCHED();
            read_status();
            next = hst->next;
            S = hst->s;
            rate = next > start ? (S - s_start) / (double)(next - start) : 0.0;
        };
        int singles = 0;  // consecutive blocks replayed alone because a crossing was expected there
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
            uint64_t len = std::min(std::max(min_chunk, next - b0), chunk_max);
            if (rate > 0.0 && std::isfinite(rate)) {
                // Blocks until S reaches its binade's top at the last span's rate (performance only:
                // any chunk gives exact results). A crossing expected in the first block is one
                // replayed block, not a chunk read in vain; a crossing expected well before the
                // chunk's end ends the chunk shortly after it.
                const double need = (std::ldexp(1.0, e + 1) - S) / rate;
                if (need < 1.0 && singles < 4) {
                    singles++;
                    finish(next, next + 1, S, 1);
                    continue;
                }
                const double cap = need * 1.25 + (double)TILE_BLOCKS;
                if (cap < (double)len) len = std::max<uint64_t>((uint64_t)cap, 1);
            }
            singles = 0;
            const uint64_t end = std::min(nblk, next + len);
            const uint64_t nb = end - next;
            const uint64_t ntiles = (nb + TILE_BLOCKS - 1) / TILE_BLOCKS;
            hipLaunchKernelGGL(k_seq_blocks<R>, dim3((unsigned)ntiles), dim3(256), 0, s.stream, psi, next, nb, 52 - e,
                               std::ldexp(1.0, e + 1), bx, tx);
            QFVM_LAUNCHED();
            const uint64_t ngroups = (ntiles + SCAN_GROUP - 1) / SCAN_GROUP;
            hipLaunchKernelGGL(k_seq_group_sums, dim3((unsigned)ngroups), dim3(SCAN_GROUP), 0, s.stream, tx, ntiles, gx);
            QFVM_LAUNCHED();
            hipLaunchKernelGGL(k_seq_tile_prefix, dim3((unsigned)ngroups), dim3(SCAN_GROUP), 0, s.stream, tx, ntiles, gx,
                               tpre);
            QFVM_LAUNCHED();
            hipLaunchKernelGGL(k_seq_apply, dim3((unsigned)((ntiles + 3) / 4)), dim3(256), 0, s.stream, bx, tpre, next,
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
