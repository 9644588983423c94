/*
 * ORACLE -- test infrastructure, not product code.
 *
 * qfvm_sample for the CPU oracle: the pairwise sum-tree descent that defines the HIP sampler's
 * result bit for bit (include/qfvm.h, qfvm_sample). Built WITHOUT -ffast-math and with
 * -ffp-contract=off so every operation below rounds exactly as written:
 *
 *   p_i        = fma(re_i, re_i, im_i * im_i)                  (double)
 *   S_0[i]     = p_i;   S_l[j] = S_{l-1}[2j] + S_{l-1}[2j+1]     (level l has 2^(n-l) nodes)
 *   descent    : if !(u < S_n[0]) -> 2^n  (the reference's "no index reached", all-zero bits)
 *                node = root, base = 0; at each level with children L, R:
 *                  R == 0 -> left;  else L == 0 -> right;  else u < base + L -> left;
 *                  else base = base + L, right.
 *
 * In exact arithmetic this is the reference's rule (statevector.hpp:1390-1396: first index whose
 * running sum exceeds u); the two differ only for u within rounding of a cumulative boundary.
 * tests/ pin this function against the sequential rule and against exact dyadic cases.
 */
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

    // levels[l] holds S_l
    std::vector<std::vector<double>> levels(n + 1);
    levels[0].resize(size);
    for (uint64_t i = 0; i < size; i++) {
        const double re = d[2 * i], im = d[2 * i + 1];
        levels[0][i] = std::fma(re, re, im * im);
    }
    for (int l = 1; l <= n; l++) {
        const std::vector<double>& lo = levels[l - 1];
        std::vector<double>& hi = levels[l];
        hi.resize(lo.size() / 2);
        for (uint64_t j = 0; j < hi.size(); j++) hi[j] = lo[2 * j] + lo[2 * j + 1];
    }
    const double root = levels[n][0];
    for (int64_t s = 0; s < shots; s++) {
        const double u = uniforms[s];
        if (!(u < root)) {
            out_index[s] = size;
            continue;
        }
        double base = 0.0;
        uint64_t idx = 0;
        for (int l = n; l >= 1; l--) {
            const double L = levels[l - 1][2 * idx];
            const double R = levels[l - 1][2 * idx + 1];
            if (R == 0.0) {
                idx = 2 * idx;
            } else if (L == 0.0) {
                idx = 2 * idx + 1;
            } else if (u < base + L) {
                idx = 2 * idx;
            } else {
                base = base + L;
                idx = 2 * idx + 1;
            }
        }
        out_index[s] = idx;
    }
    return QFVM_OK;
}
