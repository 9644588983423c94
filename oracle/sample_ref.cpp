/*
 * ORACLE -- test infrastructure, not product code.
 *
 * qfvm_sample for the CPU oracle: measure_samples' rule (statevector.hpp:1380-1412) -- for a
 * uniform u, the first index k whose running sum S_k = fl(S_{k-1} + p_k) exceeds u, 2^n when none
 * does -- evaluated with one sequential pass and a binary search per shot instead of one scan per
 * shot. S never decreases (p_k >= 0 and rounding is monotone) and every shot of the reference
 * rebuilds the same S_0, S_1, ..., so "first k with u < S_k" is the same index either way.
 *
 *   p_k = fma(re_k, re_k, im_k * im_k)   -- what g++ emits for std::real(d * std::conj(d)) under the
 *                                           reference's -O3 -mavx2 -mfma -ffast-math
 *                                           (vmulsd + vfmadd132sd, then vaddsd into the sum)
 *   S_k = S_{k-1} + p_k                   -- one rounding per add, in index order
 *
 * Built WITHOUT -ffast-math and with -ffp-contract=off so each line rounds exactly as written.
 * tests/ pin it against qfvm_sample_sequential (the reference's loop itself, compiled with the
 * reference's flags) on random, dyadic, tie-heavy and degenerate states.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "qfvm.h"

extern "C" double* qfvm_oracle_data(qfvm_handle h);

extern "C" int qfvm_sample(qfvm_handle h, const double* uniforms, int64_t shots, uint64_t* out_index) {
    int32_t n = 0;
    int rc = qfvm_num_qubits(h, &n);
    if (rc != QFVM_OK) return rc;
    const double* d = qfvm_oracle_data(h);
    const uint64_t size = 1ull << n;
    std::vector<double> S(size);
    double acc = 0.0;
    for (uint64_t i = 0; i < size; i++) {
        const double re = d[2 * i], im = d[2 * i + 1];
        acc = acc + std::fma(re, re, im * im);
        S[i] = acc;
    }
    for (int64_t s = 0; s < shots; s++) {
        const double u = uniforms[s];
        // first k with u < S[k]
        out_index[s] = (uint64_t)(std::upper_bound(S.begin(), S.end(), u) - S.begin());
    }
    return QFVM_OK;
}
