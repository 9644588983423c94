/*
 * ORACLE -- test infrastructure, not product code.
 *
 * A CPU restatement of the reference's qfvm statevector engine (src/qfvm/statevector.hpp,
 * simulator.hpp, qfvm.cpp), exporting the same C-ABI as the HIP library (include/qfvm.h) so a
 * test can run one program through both and compare. Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load it; the product never does.
 *
 * It follows the reference algorithm step for step: one OpenMP sweep per gate; the index maps
 * that insert zero bits at the target/control positions and OR in the control bits
 * (statevector.hpp:441-499, kept as std::function as the reference does); the three loop shapes
 * (target 0, control 0, unroll-by-2); named-gate dispatch (simulator.hpp:30-113); dense
 * multi-target matvec (statevector.hpp:945-995); expect_pauli (:1250-1305); mid-circuit
 * measure / reset (:1191-1377); and the per-shot sequential CDF sampler (:1380-1412).
 * Built with the reference's flags (-O3 -mavx2 -mfma -ffast-math -fopenmp, CMakeLists.txt:63-76).
 *
 * Deliberate differences (the reference's int-width bugs are not copied): all shifts and indices
 * are 64-bit (reference: `1 << pos` int shifts at statevector.hpp:208,251,311; uint popcount at
 * util.h:127-135; uint/int loops at :997-1006,1025,1058,1185); qubit and cbit indices are
 * range-checked and reported as status codes instead of undefined behaviour; uniforms are passed
 * in instead of drawn from std::random_device (util.h:35-46).
 *
 * The sum-tree sampler that the HIP library uses (qfvm_sample) lives in sample_tree.cpp, built
 * without -ffast-math so its floating-point order is exactly the documented one.
 */
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "qfvm.h"

using cd = std::complex<double>;
using std::size_t;

namespace {

thread_local std::string g_err;

struct QfvmError : std::runtime_error {
    int code;
    QfvmError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

const double kPi = 3.14159265358979323846;
const cd kI(0.0, 1.0);

}  // namespace

struct qfvm_state {
    int n = 0;
    size_t size = 1;
    std::vector<cd> data;

    // ---------------------------------------------------------------- index maps
    // statevector.hpp:446-499. ctrl_num 0, 1 and "2 = two or more controls".
    struct IndexMap {
        std::function<size_t(size_t)> near, far;
        size_t rsize = 0, offset = 0, targe = 0, control = 0;
        bool has_control = false;
    };

    IndexMap make_map(int ctrl_num, const std::vector<int>& posv) const {
        IndexMap m;
        if (ctrl_num == 0) {
            size_t targe = posv[0];
            size_t offset = 1ull << targe;
            m.targe = targe;
            m.offset = offset;
            m.rsize = size >> 1;
            m.near = [](size_t j) -> size_t { return 2 * j; };
            m.far = [offset, targe](size_t j) -> size_t {
                return (j & (offset - 1)) | (j >> targe << targe << 1);
            };
        } else if (ctrl_num == 1) {
            m.has_control = true;
            size_t control = posv[0];
            size_t targe = posv[1];
            size_t offset = 1ull << targe;
            size_t setbit = 1ull << control;
            if (control > targe) control--;  // insert the control bit first, below the target
            size_t poffset = 1ull << control;
            m.targe = targe;
            m.offset = offset;
            m.control = control;
            m.rsize = size >> 2;
            m.far = [=](size_t j) -> size_t {
                size_t i = (j >> control << (control + 1)) | (j & (poffset - 1));
                return (i >> targe << (targe + 1)) | (i & (offset - 1)) | setbit;
            };
            m.near = m.far;
        } else {
            m.has_control = true;
            m.control = (size_t)*std::min_element(posv.begin(), posv.end() - 1);
            m.targe = (size_t)posv.back();
            m.offset = 1ull << m.targe;
            std::vector<int> sorted = posv;
            std::sort(sorted.begin(), sorted.end());
            m.rsize = size >> posv.size();
            std::vector<int> ctrls(posv.begin(), posv.end() - 1);
            m.far = [sorted, ctrls](size_t j) -> size_t {
                size_t i = j;
                for (int p : sorted) i = (i & ((1ull << p) - 1)) | (i >> p << p << 1);
                for (int c : ctrls) i |= 1ull << c;
                return i;
            };
            m.near = m.far;
        }
        return m;
    }

    // ---------------------------------------------------------------- 1-target kernels
    // apply_one_targe_gate_general<c> (statevector.hpp:441-581)
    void gate_general(int ctrl_num, const std::vector<int>& posv, const cd* mat) {
        IndexMap m = make_map(ctrl_num, posv);
        const cd m00 = mat[0], m01 = mat[1], m10 = mat[2], m11 = mat[3];
        cd* d = data.data();
        const size_t off = m.offset;
        if (m.targe == 0) {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)m.rsize; j++) {
                size_t i = m.near(j);
                cd t = d[i];
                d[i] = m00 * d[i] + m01 * d[i + 1];
                d[i + 1] = m10 * t + m11 * d[i + 1];
            }
        } else if (m.has_control && m.control == 0) {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)m.rsize; j++) {
                size_t i = m.far(j);
                cd t = d[i];
                d[i] = m00 * d[i] + m01 * d[i + off];
                d[i + off] = m10 * t + m11 * d[i + off];
            }
        } else {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)m.rsize; j += 2) {
                size_t i = m.far(j);
                size_t i1 = i + 1;
                cd t = d[i], t1 = d[i1];
                d[i] = m00 * d[i] + m01 * d[i + off];
                d[i + off] = m10 * t + m11 * d[i + off];
                d[i1] = m00 * d[i1] + m01 * d[i1 + off];
                d[i1 + off] = m10 * t1 + m11 * d[i1 + off];
            }
        }
    }

    // apply_one_targe_gate_x<c> (statevector.hpp:583-689)
    void gate_x(int ctrl_num, const std::vector<int>& posv) {
        IndexMap m = make_map(ctrl_num, posv);
        cd* d = data.data();
        const size_t off = m.offset;
        if (m.targe == 0) {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)m.rsize; j++) {
                size_t i = m.near(j);
                std::swap(d[i], d[i + 1]);
            }
        } else if (m.has_control && m.control == 0) {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)m.rsize; j++) {
                size_t i = m.far(j);
                std::swap(d[i], d[i + off]);
            }
        } else {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)m.rsize; j += 2) {
                size_t i = m.far(j);
                std::swap(d[i], d[i + off]);
                std::swap(d[i + 1], d[i + 1 + off]);
            }
        }
    }

    // apply_one_targe_gate_real<c> (statevector.hpp:691-818): real 2x2 (H, RY)
    void gate_real(int ctrl_num, const std::vector<int>& posv, const cd* mat) {
        IndexMap m = make_map(ctrl_num, posv);
        const double m00 = mat[0].real(), m01 = mat[1].real(), m10 = mat[2].real(), m11 = mat[3].real();
        cd* d = data.data();
        const size_t off = m.offset;
        if (m.targe == 0) {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)m.rsize; j++) {
                size_t i = m.near(j);
                cd t = d[i];
                d[i] = m00 * d[i] + m01 * d[i + 1];
                d[i + 1] = m10 * t + m11 * d[i + 1];
            }
        } else if (m.has_control && m.control == 0) {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)m.rsize; j++) {
                size_t i = m.far(j);
                cd t = d[i];
                d[i] = m00 * d[i] + m01 * d[i + off];
                d[i + off] = m10 * t + m11 * d[i + off];
            }
        } else {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)m.rsize; j += 2) {
                size_t i = m.far(j);
                size_t i1 = i + 1;
                cd t = d[i], t1 = d[i1];
                d[i] = m00 * d[i] + m01 * d[i + off];
                d[i + off] = m10 * t + m11 * d[i + off];
                d[i1] = m00 * d[i1] + m01 * d[i1 + off];
                d[i1 + off] = m10 * t1 + m11 * d[i1 + off];
            }
        }
    }

    // apply_one_targe_gate_diag<c> (statevector.hpp:820-943): mat = {d0, d1}
    void gate_diag(int ctrl_num, const std::vector<int>& posv, const cd* mat) {
        IndexMap m = make_map(ctrl_num, posv);
        const cd d0 = mat[0], d1 = mat[1];
        cd* d = data.data();
        const size_t off = m.offset;
        if (m.targe == 0) {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)m.rsize; j++) {
                size_t i = m.near(j);
                d[i] *= d0;
                d[i + 1] *= d1;
            }
        } else if (m.has_control && m.control == 0) {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)m.rsize; j++) {
                size_t i = m.far(j);
                d[i] *= d0;
                d[i + off] *= d1;
            }
        } else {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)m.rsize; j += 2) {
                size_t i = m.far(j);
                d[i] *= d0;
                d[i + off] *= d1;
                d[i + 1] *= d0;
                d[i + 1 + off] *= d1;
            }
        }
    }

    // ---------------------------------------------------------------- named Paulis
    // apply_x / apply_y / apply_z (statevector.hpp:206-338), 64-bit offsets.
    void apply_x(int pos) {
        cd* d = data.data();
        const size_t offset = 1ull << pos;
        const size_t rsize = size >> 1;
        if (pos == 0) {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)size; j += 2) std::swap(d[j], d[j + 1]);
        } else {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)rsize; j += 2) {
                size_t i = (j & (offset - 1)) | ((size_t)j >> pos << pos << 1);
                std::swap(d[i], d[i + offset]);
                std::swap(d[i + 1], d[i + 1 + offset]);
            }
        }
    }

    void apply_y(int pos) {
        cd* d = data.data();
        const size_t offset = 1ull << pos;
        const size_t rsize = size >> 1;
        if (pos == 0) {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)size; j += 2) {
                cd t = d[j];
                d[j] = -kI * d[j + 1];
                d[j + 1] = kI * t;
            }
        } else {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)rsize; j += 2) {
                size_t i = (j & (offset - 1)) | ((size_t)j >> pos << pos << 1);
                for (size_t k = i; k < i + 2; k++) {
                    cd t = d[k];
                    d[k] = -kI * d[k + offset];
                    d[k + offset] = kI * t;
                }
            }
        }
    }

    void apply_z(int pos) {
        cd* d = data.data();
        const size_t offset = 1ull << pos;
        const size_t rsize = size >> 1;
        if (pos == 0) {
#pragma omp parallel for
            for (int64_t j = 1; j < (int64_t)size; j += 2) d[j] *= -1;
        } else {
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)rsize; j += 2) {
                size_t i = (j & (offset - 1)) | ((size_t)j >> pos << pos << 1);
                d[i + offset] *= -1;
                d[i + offset + 1] *= -1;
            }
        }
    }

    // ---------------------------------------------------------------- multi-target dense
    // apply_multi_targe_gate_general (statevector.hpp:945-995): row-index bit j <-> targs[j].
    void gate_multi(const std::vector<int>& posv, int control_num, const std::vector<cd>& mat) {
        std::vector<int> sorted = posv;
        std::sort(sorted.begin(), sorted.end());
        std::vector<int> targs(posv.begin() + control_num, posv.end());
        const size_t rsize = size >> posv.size();
        const size_t tn = targs.size();
        const size_t ms = 1ull << tn;
        std::vector<size_t> mask(ms, 0);
        for (size_t mm = 0; mm < ms; mm++)
            for (size_t j = 0; j < tn; j++)
                if ((mm >> j) & 1) mask[mm] |= 1ull << targs[j];
        cd* d = data.data();
#pragma omp parallel
        {
            std::vector<cd> v(ms), w(ms);
#pragma omp for
            for (int64_t j = 0; j < (int64_t)rsize; j++) {
                size_t i = j;
                for (int p : sorted) i = (i & ((1ull << p) - 1)) | (i >> p << p << 1);
                for (int k = 0; k < control_num; k++) i |= 1ull << posv[k];
                for (size_t mm = 0; mm < ms; mm++) v[mm] = d[i | mask[mm]];
                for (size_t r = 0; r < ms; r++) {
                    cd acc = 0.0;
                    for (size_t c = 0; c < ms; c++) acc += mat[r * ms + c] * v[c];
                    w[r] = acc;
                }
                for (size_t mm = 0; mm < ms; mm++) d[i | mask[mm]] = w[mm];
            }
        }
    }

    // ---------------------------------------------------------------- measurement helpers
    // sample_measure_probs (statevector.hpp:1307-1340): probs[m], bit i of m <-> qbits[i].
    std::vector<double> marginal_probs(const std::vector<int>& qbits) const {
        const size_t N = qbits.size();
        std::vector<double> probs(1ull << N, 0.0);
        for (size_t j = 0; j < size; j++) {
            size_t m = 0;
            for (size_t b = 0; b < N; b++) m |= ((j >> qbits[b]) & 1ull) << b;
            probs[m] += std::norm(data[j]);
        }
        return probs;
    }

    // std::discrete_distribution as implemented by libstdc++ (<bits/random.tcc>): normalise by
    // the sum, partial sums, last cumulative set to 1, first index with u <= cp[i].
    static size_t discrete_pick(const std::vector<double>& p, double u) {
        if (p.size() < 2) return 0;
        double sum = 0.0;
        for (double x : p) sum += x;
        std::vector<double> cp(p.size());
        double acc = 0.0;
        for (size_t i = 0; i < p.size(); i++) {
            acc += p[i] / sum;
            cp[i] = acc;
        }
        cp.back() = 1.0;
        return (size_t)(std::lower_bound(cp.begin(), cp.end(), u) - cp.begin());
    }

    // update (statevector.hpp:1191-1248): project onto meas_state, renormalise, then move the
    // surviving amplitudes to final_state (reset).
    void update(const std::vector<int>& qbits, size_t final_state, size_t meas_state, double prob) {
        const size_t N = qbits.size();
        const double scale = 1.0 / std::sqrt(prob);
        cd* d = data.data();
#pragma omp parallel for
        for (int64_t j = 0; j < (int64_t)size; j++) {
            size_t m = 0;
            for (size_t b = 0; b < N; b++) m |= (((size_t)j >> qbits[b]) & 1ull) << b;
            if (m == meas_state) {
                if (scale != 1.0) d[j] *= scale;
            } else {
                d[j] = 0.0;
            }
        }
        if (final_state != meas_state) {
            if (N == 1) {
                apply_x(qbits[0]);
            } else {
                size_t fmask = 0, mmask = 0;
                for (size_t b = 0; b < N; b++) {
                    if ((final_state >> b) & 1) fmask |= 1ull << qbits[b];
                    if ((meas_state >> b) & 1) mmask |= 1ull << qbits[b];
                }
                size_t qmask = 0;
                for (int q : qbits) qmask |= 1ull << q;
#pragma omp parallel for
                for (int64_t j = 0; j < (int64_t)size; j++) {
                    if (((size_t)j & qmask) == fmask) {
                        size_t k = ((size_t)j & ~qmask) | mmask;
                        std::swap(d[j], d[k]);
                    }
                }
            }
        }
    }

    // expect_pauli (statevector.hpp:1250-1305), 64-bit popcount.
    double expect_pauli(const char* s, const int32_t* pos, int len) const {
        size_t flip_mask = 0, z_mask = 0, y_num = 0;
        int flip_q = 0;
        for (int i = 0; i < len; i++) {
            int q = pos[i];
            switch (s[i]) {
                case 'X':
                    flip_mask += 1ull << q;
                    flip_q = std::max(flip_q, q);
                    break;
                case 'Y':
                    flip_mask += 1ull << q;
                    flip_q = std::max(flip_q, q);
                    y_num++;
                    z_mask += 1ull << q;
                    break;
                case 'Z':
                    z_mask += 1ull << q;
                    break;
                default:
                    break;
            }
        }
        const cd* d = data.data();
        double val = 0.0;
        if (!flip_mask) {
#pragma omp parallel for reduction(+ : val)
            for (int64_t j = 0; j < (int64_t)size; j++) {
                int par = __builtin_popcountll((size_t)j & z_mask) & 1;
                val += std::norm(d[j]) * (1 - 2 * par);
            }
        } else {
            const cd phases[4] = {cd(1, 0), cd(0, 1), cd(-1, 0), cd(0, -1)};
            const size_t rs = size >> 1;
#pragma omp parallel for reduction(+ : val)
            for (int64_t j = 0; j < (int64_t)rs; j++) {
                size_t i0 = ((size_t)j & ((1ull << flip_q) - 1)) | ((size_t)j >> flip_q << flip_q << 1);
                size_t i1 = i0 ^ flip_mask;
                size_t zp = __builtin_popcountll(i0 & z_mask) & 1;
                cd ph = phases[(zp * 2 + y_num) % 4];
                val += 2.0 * (d[i0] * std::conj(d[i1]) * ph).real();
            }
        }
        return val;
    }
};

// ==================================================================== program interpreter
namespace {

void check_qubits(const qfvm_state* s, const std::vector<int>& q) {
    for (size_t a = 0; a < q.size(); a++) {
        if (q[a] < 0 || q[a] >= s->n)
            throw QfvmError(QFVM_ERR_INDEX, "qubit index " + std::to_string(q[a]) + " out of range for " +
                                                std::to_string(s->n) + " qubits");
        for (size_t b = 0; b < a; b++)
            if (q[a] == q[b]) throw QfvmError(QFVM_ERR_VALUE, "repeated qubit " + std::to_string(q[a]));
    }
}

std::vector<int> op_list(const qfvm_program* p, int64_t off, int32_t n) {
    if (n < 0 || off < 0 || off + n > p->n_idx) throw QfvmError(QFVM_ERR_VALUE, "index pool range");
    return std::vector<int>(p->idx + off, p->idx + off + n);
}

std::vector<cd> op_matrix(const qfvm_program* p, const qfvm_op& op, size_t dim) {
    const int64_t need = (int64_t)(2 * dim * dim);
    if (op.m_off < 0 || op.m_off + need > p->n_mat) throw QfvmError(QFVM_ERR_VALUE, "matrix pool range");
    std::vector<cd> m(dim * dim);
    for (size_t k = 0; k < dim * dim; k++) m[k] = cd(p->mat[op.m_off + 2 * k], p->mat[op.m_off + 2 * k + 1]);
    return m;
}

// apply_op / apply_op_general (simulator.hpp:12-113) for one unitary op.
void apply_gate(qfvm_state* s, const qfvm_program* p, const qfvm_op& op) {
    std::vector<int> q = op_list(p, op.q_off, op.n_qubits);
    check_qubits(s, q);
    auto need = [&](int k) {
        if ((int)q.size() < k) throw QfvmError(QFVM_ERR_VALUE, "op needs " + std::to_string(k) + " qubits");
    };
    const double th = op.param;
    switch (op.code) {
        case QFVM_OP_GENERAL:
        case QFVM_OP_MULTI: {
            int nc = op.n_ctrl;
            int nt = (int)q.size() - nc;
            if (nc < 0 || nt < 1) throw QfvmError(QFVM_ERR_VALUE, "Invalid target number");
            if (nt == 1) {
                std::vector<cd> m = op_matrix(p, op, 2);
                s->gate_general(std::min(nc, 2), q, m.data());
            } else {
                if (nt > 16) throw QfvmError(QFVM_ERR_VALUE, "too many targets");
                std::vector<cd> m = op_matrix(p, op, 1ull << nt);
                s->gate_multi(q, nc, m);
            }
            break;
        }
        case QFVM_OP_X: need(1); s->apply_x(q[0]); break;
        case QFVM_OP_Y: need(1); s->apply_y(q[0]); break;
        case QFVM_OP_Z: need(1); s->apply_z(q[0]); break;
        case QFVM_OP_H: {  // statevector.hpp:340-345
            need(1);
            const double r = 1. / std::sqrt(2.);
            cd m[4] = {r, r, r, -r};
            s->gate_real(0, {q[0]}, m);
            break;
        }
        case QFVM_OP_S: { need(1); cd m[2] = {1., kI}; s->gate_diag(0, {q[0]}, m); break; }
        case QFVM_OP_SDG: { need(1); cd m[2] = {1., -kI}; s->gate_diag(0, {q[0]}, m); break; }
        case QFVM_OP_T: { need(1); cd m[2] = {1., std::exp(kI * kPi / 4.)}; s->gate_diag(0, {q[0]}, m); break; }
        case QFVM_OP_TDG: { need(1); cd m[2] = {1., std::exp(-kI * kPi / 4.)}; s->gate_diag(0, {q[0]}, m); break; }
        case QFVM_OP_P: { need(1); cd m[2] = {1., std::exp(kI * th)}; s->gate_diag(0, {q[0]}, m); break; }
        case QFVM_OP_RX: {  // statevector.hpp:379-385
            need(1);
            cd m[4] = {std::cos(th / 2), -kI * std::sin(th / 2), -kI * std::sin(th / 2), std::cos(th / 2)};
            s->gate_general(0, {q[0]}, m);
            break;
        }
        case QFVM_OP_RY: {
            need(1);
            cd m[4] = {std::cos(th / 2), -std::sin(th / 2), std::sin(th / 2), std::cos(th / 2)};
            s->gate_real(0, {q[0]}, m);
            break;
        }
        case QFVM_OP_RZ: {
            need(1);
            cd m[2] = {std::exp(-kI * th / 2.), std::exp(kI * th / 2.)};
            s->gate_diag(0, {q[0]}, m);
            break;
        }
        case QFVM_OP_CX: need(2); s->gate_x(1, {q[0], q[1]}); break;
        case QFVM_OP_CZ: { need(2); cd m[2] = {1., -1.}; s->gate_diag(1, {q[0], q[1]}, m); break; }
        case QFVM_OP_CP: { need(2); cd m[2] = {1., std::exp(kI * th)}; s->gate_diag(1, {q[0], q[1]}, m); break; }
        case QFVM_OP_CCX: need(3); s->gate_x(2, {q[0], q[1], q[2]}); break;
        case QFVM_OP_RZZ: {  // simulator.hpp:90-94: CX, RZ on the second qubit, CX
            need(2);
            s->gate_x(1, {q[0], q[1]});
            cd m[2] = {std::exp(-kI * th / 2.), std::exp(kI * th / 2.)};
            s->gate_diag(0, {q[1]}, m);
            s->gate_x(1, {q[0], q[1]});
            break;
        }
        default:
            throw QfvmError(QFVM_ERR_VALUE, "op code " + std::to_string(op.code) + " is not a unitary gate");
    }
}

struct Runner {
    qfvm_state* s;
    const qfvm_program* p;
    const double* u;
    int64_t nu;
    int64_t ui = 0;
    int32_t* creg;
    int32_t ncreg;
    bool skip_measure;

    double next_uniform() {
        if (ui >= nu) throw QfvmError(QFVM_ERR_VALUE, "not enough uniforms for measure/reset");
        return u[ui++];
    }

    // Executes ops [begin, end); returns nothing. CIF bodies are nested ranges.
    void exec(int64_t begin, int64_t end) {
        int64_t k = begin;
        while (k < end) {
            const qfvm_op& op = p->ops[k];
            if (op.code == QFVM_OP_MEASURE) {
                k++;
                if (skip_measure) continue;
                std::vector<int> q = op_list(p, op.q_off, op.n_qubits);
                std::vector<int> c = op_list(p, op.c_off, op.n_cbits);
                check_qubits(s, q);
                if (c.size() < q.size()) throw QfvmError(QFVM_ERR_VALUE, "measure needs one cbit per qubit");
                for (size_t j = 0; j < q.size(); j++)
                    if (c[j] < 0 || c[j] >= ncreg) throw QfvmError(QFVM_ERR_INDEX, "cbit out of range");
                std::vector<double> pr = s->marginal_probs(q);
                size_t m = qfvm_state::discrete_pick(pr, next_uniform());
                s->update(q, m, m, pr[m]);
                for (size_t j = 0; j < q.size(); j++) creg[c[j]] = (int32_t)((m >> j) & 1);
            } else if (op.code == QFVM_OP_RESET) {
                k++;
                std::vector<int> q = op_list(p, op.q_off, op.n_qubits);
                check_qubits(s, q);
                std::vector<double> pr = s->marginal_probs(q);
                size_t m = qfvm_state::discrete_pick(pr, next_uniform());
                s->update(q, 0, m, pr[m]);
            } else if (op.code == QFVM_OP_CIF) {
                std::vector<int> c = op_list(p, op.c_off, op.n_cbits);
                int64_t body_end = k + 1 + op.body;
                if (op.body < 0 || body_end > end) throw QfvmError(QFVM_ERR_VALUE, "cif body range");
                uint64_t out = 0;
                for (int cb : c) {
                    if (cb < 0 || cb >= ncreg) throw QfvmError(QFVM_ERR_INDEX, "cbit out of range");
                    out = out * 2 + (uint64_t)creg[cb];
                }
                if ((int64_t)out == op.cond) exec(k + 1, body_end);
                k = body_end;
            } else {
                apply_gate(s, p, op);
                k++;
            }
        }
    }
};

template <class F>
int guarded(F&& f) {
    try {
        f();
        return QFVM_OK;
    } catch (const QfvmError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return QFVM_ERR_MEMORY;
    } catch (const std::exception& e) {
        g_err = e.what();
        return QFVM_ERR_RUNTIME;
    }
}

void need_handle(qfvm_handle h) {
    if (!h) throw QfvmError(QFVM_ERR_VALUE, "null handle");
}

}  // namespace

extern "C" {

const char* qfvm_last_error(void) { return g_err.c_str(); }

int qfvm_device_count(int* out) {
    *out = 0;
    return QFVM_OK;
}

int qfvm_create(int32_t n_qubits, int32_t dtype, qfvm_handle* out) {
    return guarded([&] {
        if (n_qubits < 0 || n_qubits > 34) throw QfvmError(QFVM_ERR_VALUE, "n_qubits must be in [0, 34]");
        if (dtype != QFVM_C128) throw QfvmError(QFVM_ERR_VALUE, "the CPU oracle holds complex128 only");
        auto s = std::make_unique<qfvm_state>();
        s->n = n_qubits;
        s->size = 1ull << n_qubits;
        s->data.assign(s->size, cd(0.0, 0.0));
        s->data[0] = 1.0;
        *out = s.release();
    });
}

int qfvm_destroy(qfvm_handle h) {
    delete h;
    return QFVM_OK;
}

int qfvm_num_qubits(qfvm_handle h, int32_t* out) {
    return guarded([&] { need_handle(h); *out = h->n; });
}

int qfvm_dtype_of(qfvm_handle h, int32_t* out) {
    return guarded([&] { need_handle(h); *out = QFVM_C128; });
}

int qfvm_init_zero(qfvm_handle h) {
    return guarded([&] {
        need_handle(h);
        std::fill(h->data.begin(), h->data.end(), cd(0.0, 0.0));
        h->data[0] = 1.0;
    });
}

int qfvm_init_basis(qfvm_handle h, uint64_t index) {
    return guarded([&] {
        need_handle(h);
        if (index >= h->size) throw QfvmError(QFVM_ERR_INDEX, "basis index out of range");
        std::fill(h->data.begin(), h->data.end(), cd(0.0, 0.0));
        h->data[index] = 1.0;
    });
}

int qfvm_load(qfvm_handle h, const double* host, uint64_t n_amps) {
    return guarded([&] {
        need_handle(h);
        if (n_amps != h->size) throw QfvmError(QFVM_ERR_VALUE, "state length must be 2^n");
        std::memcpy(h->data.data(), host, n_amps * sizeof(cd));
    });
}

int qfvm_store(qfvm_handle h, double* host, uint64_t n_amps) {
    return guarded([&] {
        need_handle(h);
        if (n_amps != h->size) throw QfvmError(QFVM_ERR_VALUE, "state length must be 2^n");
        std::memcpy(host, h->data.data(), n_amps * sizeof(cd));
    });
}

int qfvm_apply(qfvm_handle h, const qfvm_program* prog) {
    return guarded([&] {
        need_handle(h);
        for (int64_t k = 0; k < prog->n_ops; k++) apply_gate(h, prog, prog->ops[k]);
    });
}

int qfvm_run(qfvm_handle h, const qfvm_program* prog, const double* uniforms, int64_t n_uniforms, int32_t* creg,
             int32_t n_creg, int32_t skip_measure) {
    return guarded([&] {
        need_handle(h);
        Runner r{h, prog, uniforms, n_uniforms, 0, creg, n_creg, skip_measure != 0};
        r.exec(0, prog->n_ops);
    });
}

int qfvm_expect_paulis(qfvm_handle h, const char* chars, const int32_t* pos, const int64_t* off,
                       const int32_t* len, int64_t n_terms, double* out) {
    return guarded([&] {
        need_handle(h);
        for (int64_t t = 0; t < n_terms; t++) {
            for (int32_t k = 0; k < len[t]; k++) {
                int q = pos[off[t] + k];
                if (q < 0 || q >= h->n) throw QfvmError(QFVM_ERR_INDEX, "pauli qubit out of range");
            }
            out[t] = h->expect_pauli(chars + off[t], pos + off[t], len[t]);
        }
    });
}

int qfvm_probabilities(qfvm_handle h, double* out, uint64_t n) {
    return guarded([&] {
        need_handle(h);
        if (n != h->size) throw QfvmError(QFVM_ERR_VALUE, "output length must be 2^n");
        const cd* d = h->data.data();
#pragma omp parallel for
        for (int64_t j = 0; j < (int64_t)n; j++) out[j] = std::real(d[j] * std::conj(d[j]));
    });
}

int qfvm_norm2(qfvm_handle h, double* out) {
    return guarded([&] {
        need_handle(h);
        double acc = 0.0;
        const cd* d = h->data.data();
#pragma omp parallel for reduction(+ : acc)
        for (int64_t j = 0; j < (int64_t)h->size; j++) acc += std::norm(d[j]);
        *out = acc;
    });
}

int qfvm_synchronize(qfvm_handle h) {
    (void)h;
    return QFVM_OK;
}

/*
 * grad_adjoint (quafu/algorithms/gradients/gradiant.py:144-173) on the oracle's kernels, step for
 * step, with the HIP library's argument layout (include/qfvm.h qfvm_adjoint):
 *   forward: every op of prog on the state (SVSimulator.run);
 *   phi = psi; lam = H psi (_apply_hamil, simulator.py:78-88: per term, a copy with the term's
 *   X / Y / Z gates applied one by one, times the coefficient, summed);
 *   for op i from the last down to the first differentiated one: phi <- U_i^dagger phi; for a
 *   differentiated op, mu = projCtrl(dU phi) with dU = -1/2 i P U on the target alone (grad_gate,
 *   gradiant.py:96-141) and grad = Re(2 <lam, mu>); then lam <- U_i^dagger lam.
 * U^dagger is the op's own inverse as op.dagger() builds it: the conjugate transpose of a matrix,
 * the negated angle of a rotation / phase, S <-> SDG, T <-> TDG, the Paulis, H, CX, CZ and CCX
 * themselves. On return the state holds phi just before the first differentiated op.
 */
int qfvm_adjoint(qfvm_handle h, const qfvm_program* prog, const int64_t* grad_ops, const char* grad_gen,
                 int64_t n_grads, const char* chars, const int32_t* pos, const int64_t* off,
                 const int32_t* len, const double* coef, int64_t n_terms, double* grads, double* expectation) {
    return guarded([&] {
        need_handle(h);
        if (!prog || prog->n_ops < 0) throw QfvmError(QFVM_ERR_VALUE, "null program");
        std::vector<int64_t> slot_of((size_t)prog->n_ops, -1);
        int64_t begin = prog->n_ops;
        for (int64_t g = 0; g < n_grads; g++) {
            const int64_t k = grad_ops[g];
            if (k < 0 || k >= prog->n_ops) throw QfvmError(QFVM_ERR_INDEX, "gradient op out of range");
            slot_of[(size_t)k] = g;
            begin = std::min(begin, k);
        }
        for (int64_t k = 0; k < prog->n_ops; k++) apply_gate(h, prog, prog->ops[k]);
        const std::vector<cd> psi = h->data;
        // lam = H psi
        std::vector<cd> lam(h->size, cd(0.0, 0.0));
        for (int64_t t = 0; t < n_terms; t++) {
            h->data = psi;
            for (int32_t k = 0; k < len[t]; k++) {
                const int q = pos[off[t] + k];
                if (q < 0 || q >= h->n) throw QfvmError(QFVM_ERR_INDEX, "pauli qubit out of range");
                switch (chars[off[t] + k]) {
                    case 'X': h->apply_x(q); break;
                    case 'Y': h->apply_y(q); break;
                    case 'Z': h->apply_z(q); break;
                    case 'I': break;
                    default: throw QfvmError(QFVM_ERR_VALUE, "Pauli letters are I, X, Y, Z");
                }
            }
            const cd c(coef[2 * t], coef[2 * t + 1]);
            const cd* v = h->data.data();
            cd* l = lam.data();
#pragma omp parallel for
            for (int64_t j = 0; j < (int64_t)h->size; j++) l[j] += v[j] * c;
        }
        if (expectation) {
            double e = 0.0;
            const cd* a = psi.data();
            const cd* l = lam.data();
#pragma omp parallel for reduction(+ : e)
            for (int64_t j = 0; j < (int64_t)h->size; j++) e += std::real(std::conj(a[j]) * l[j]);
            *expectation = e;
        }
        // one-op programs for U^dagger and dU
        std::vector<int32_t> idx;
        std::vector<double> mat;
        auto dagger_of = [&](const qfvm_op& op, qfvm_op& d) {
            d = op;
            idx.assign(prog->idx + op.q_off, prog->idx + op.q_off + op.n_qubits);
            d.q_off = 0;
            switch (op.code) {
                case QFVM_OP_GENERAL:
                case QFVM_OP_MULTI: {
                    const int nt = op.n_qubits - op.n_ctrl;
                    const size_t dim = 1ull << nt;
                    std::vector<cd> m = op_matrix(prog, op, dim);
                    mat.assign(2 * dim * dim, 0.0);
                    for (size_t r = 0; r < dim; r++)
                        for (size_t c = 0; c < dim; c++) {
                            const cd z = std::conj(m[c * dim + r]);
                            mat[2 * (r * dim + c)] = z.real();
                            mat[2 * (r * dim + c) + 1] = z.imag();
                        }
                    d.m_off = 0;
                    break;
                }
                case QFVM_OP_S: d.code = QFVM_OP_SDG; break;
                case QFVM_OP_SDG: d.code = QFVM_OP_S; break;
                case QFVM_OP_T: d.code = QFVM_OP_TDG; break;
                case QFVM_OP_TDG: d.code = QFVM_OP_T; break;
                case QFVM_OP_P:
                case QFVM_OP_RX:
                case QFVM_OP_RY:
                case QFVM_OP_RZ:
                case QFVM_OP_CP:
                case QFVM_OP_RZZ: d.param = -op.param; break;
                case QFVM_OP_X: case QFVM_OP_Y: case QFVM_OP_Z: case QFVM_OP_H:
                case QFVM_OP_CX: case QFVM_OP_CZ: case QFVM_OP_CCX: break;
                default: throw QfvmError(QFVM_ERR_VALUE, "qfvm_adjoint takes unitary ops only");
            }
        };
        auto apply_dagger = [&](const qfvm_op& op) {
            qfvm_op d{};
            dagger_of(op, d);
            qfvm_program one{&d, 1, idx.data(), (int64_t)idx.size(), mat.data(), (int64_t)mat.size()};
            apply_gate(h, &one, d);
        };
        std::vector<cd> phi = psi;
        for (int64_t i = prog->n_ops - 1; i >= begin; i--) {
            const qfvm_op& op = prog->ops[i];
            h->data.swap(phi);
            apply_dagger(op);
            h->data.swap(phi);
            const int64_t g = slot_of[(size_t)i];
            if (g >= 0) {
                std::vector<int> q = op_list(prog, op.q_off, op.n_qubits);
                const int nc = (op.code == QFVM_OP_GENERAL) ? op.n_ctrl : 0;
                if ((int)q.size() != nc + 1) throw QfvmError(QFVM_ERR_VALUE, "differentiated ops have one target");
                const int target = q[(size_t)nc];
                cd u[4];
                const double th = op.param;
                switch (op.code) {
                    case QFVM_OP_RX:
                        u[0] = std::cos(th / 2); u[1] = -kI * std::sin(th / 2);
                        u[2] = -kI * std::sin(th / 2); u[3] = std::cos(th / 2);
                        break;
                    case QFVM_OP_RY:
                        u[0] = std::cos(th / 2); u[1] = -std::sin(th / 2);
                        u[2] = std::sin(th / 2); u[3] = std::cos(th / 2);
                        break;
                    case QFVM_OP_RZ:
                        u[0] = std::exp(-kI * th / 2.); u[1] = 0.0; u[2] = 0.0; u[3] = std::exp(kI * th / 2.);
                        break;
                    case QFVM_OP_GENERAL: {
                        std::vector<cd> m = op_matrix(prog, op, 2);
                        for (int e = 0; e < 4; e++) u[e] = m[(size_t)e];
                        break;
                    }
                    default: throw QfvmError(QFVM_ERR_VALUE, "differentiated ops are rotations");
                }
                cd p[4];
                switch (grad_gen[g]) {
                    case 'X': p[0] = 0.0; p[1] = 1.0; p[2] = 1.0; p[3] = 0.0; break;
                    case 'Y': p[0] = 0.0; p[1] = -kI; p[2] = kI; p[3] = 0.0; break;
                    case 'Z': p[0] = 1.0; p[1] = 0.0; p[2] = 0.0; p[3] = -1.0; break;
                    default: throw QfvmError(QFVM_ERR_VALUE, "gradient generators are X, Y or Z");
                }
                cd du[4];
                for (int r = 0; r < 2; r++)
                    for (int c = 0; c < 2; c++)
                        du[2 * r + c] = -0.5 * kI * (p[2 * r] * u[c] + p[2 * r + 1] * u[2 + c]);
                h->data = phi;
                h->gate_general(0, {target}, du);
                uint64_t cm = 0;
                for (int c = 0; c < nc; c++) cm |= 1ull << q[(size_t)c];
                double acc = 0.0;
                const cd* mu = h->data.data();
                const cd* l = lam.data();
#pragma omp parallel for reduction(+ : acc)
                for (int64_t j = 0; j < (int64_t)h->size; j++)
                    if (((uint64_t)j & cm) == cm) acc += std::real(std::conj(l[j]) * mu[j]);
                grads[g] = 2.0 * acc;
            }
            h->data.swap(lam);
            apply_dagger(op);
            h->data.swap(lam);
        }
        h->data = phi;
    });
}

/*
 * Oracle-only: the reference's own sampler semantics (measure_samples, statevector.hpp:1380-1412):
 * for each uniform, the first index whose sequential running sum of |amp|^2 exceeds it, 2^n if none.
 * Serial over shots, as the reference's orphaned `omp for` is.
 */
int qfvm_sample_sequential(qfvm_handle h, const double* uniforms, int64_t shots, uint64_t* out_index) {
    return guarded([&] {
        need_handle(h);
        const cd* d = h->data.data();
        for (int64_t s = 0; s < shots; s++) {
            double r = uniforms[s];
            double p = 0.0;
            size_t sample;
            for (sample = 0; sample < h->size; ++sample) {
                p += std::real(d[sample] * std::conj(d[sample]));
                if (r < p) break;
            }
            out_index[s] = sample;
        }
    });
}

/* Oracle-only: raw pointer to the host state (for zero-copy views in tests). */
double* qfvm_oracle_data(qfvm_handle h) { return reinterpret_cast<double*>(h->data.data()); }

}  // extern "C"
