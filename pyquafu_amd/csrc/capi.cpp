// C-ABI of the engine (include/qfvm.h): handle lifecycle, state I/O, program lowering and the
// mid-circuit runner. Replaces the reference's pybind entry points (src/qfvm/qfvm.cpp:27-288)
// and the C++ dispatch they drive (src/qfvm/simulator.hpp:12-125).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "common.h"
#include "fused.h"
#include "plan.h"

using namespace qfvm;

namespace {

thread_local std::string g_err;

const double kPi = 3.14159265358979323846;  // types.hpp:33
const cd kI(0.0, 1.0);

template <class F>
int guarded(F&& f) {
    try {
        f();
        return QFVM_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return QFVM_ERR_MEMORY;
    } catch (const std::exception& e) {
        g_err = e.what();
        return QFVM_ERR_RUNTIME;
    } catch (...) {
        g_err = "unknown failure";
        return QFVM_ERR_RUNTIME;
    }
}

State* need(qfvm_handle h) {
    if (!h) throw Error(QFVM_ERR_VALUE, "null handle");
    State* s = reinterpret_cast<State*>(h);
    QFVM_HIP(hipSetDevice(s->device));
    return s;
}

// ------------------------------------------------------------------------------------------------
// program parsing and lowering to internal gates (the reference's apply_op dispatch,
// simulator.hpp:30-113, with the named-gate matrices of statevector.hpp:340-437)
std::vector<int> op_list(const qfvm_program* p, int64_t off, int32_t n) {
    if (n < 0 || off < 0 || off + n > p->n_idx || (n > 0 && !p->idx)) throw Error(QFVM_ERR_VALUE, "index pool range");
    return std::vector<int>(p->idx + off, p->idx + off + n);
}

void check_qubits(const State& s, const std::vector<int>& q) {
    for (size_t a = 0; a < q.size(); a++) {
        if (q[a] < 0 || q[a] >= s.n)
            throw Error(QFVM_ERR_INDEX,
                        "qubit index " + std::to_string(q[a]) + " out of range for " + std::to_string(s.n) + " qubits");
        for (size_t b = 0; b < a; b++)
            if (q[a] == q[b]) throw Error(QFVM_ERR_VALUE, "repeated qubit " + std::to_string(q[a]));
    }
}

std::vector<cd> op_matrix(const qfvm_program* p, const qfvm_op& op, size_t dim) {
    const int64_t need_d = (int64_t)(2 * dim * dim);
    if (op.m_off < 0 || op.m_off + need_d > p->n_mat || !p->mat) throw Error(QFVM_ERR_VALUE, "matrix pool range");
    std::vector<cd> m(dim * dim);
    for (size_t k = 0; k < dim * dim; k++) m[k] = cd(p->mat[op.m_off + 2 * k], p->mat[op.m_off + 2 * k + 1]);
    return m;
}

Gate one(int kind, int target, std::vector<int> ctrls, cd a, cd b, cd c = 0.0, cd d = 0.0) {
    Gate g;
    g.kind = kind;
    g.target = target;
    g.ctrls = std::move(ctrls);
    g.cmask = 0;
    for (int q : g.ctrls) g.cmask |= 1ull << q;
    g.m[0] = a;
    g.m[1] = b;
    g.m[2] = c;
    g.m[3] = d;
    return g;
}

void lower_op(const State& s, const qfvm_program* p, const qfvm_op& op, std::vector<Gate>& out) {
    std::vector<int> q = op_list(p, op.q_off, op.n_qubits);
    check_qubits(s, q);
    auto need_q = [&](size_t k) {
        if (q.size() < k) throw Error(QFVM_ERR_VALUE, "op needs " + std::to_string(k) + " qubits");
    };
    const double th = op.param;
    switch (op.code) {
        case QFVM_OP_GENERAL:
        case QFVM_OP_MULTI: {
            const int nc = op.n_ctrl;
            const int nt = (int)q.size() - nc;
            if (nc < 0 || nt < 1) throw Error(QFVM_ERR_VALUE, "Invalid target number");
            std::vector<int> ctrls(q.begin(), q.begin() + nc);
            if (nt == 1) {
                std::vector<cd> m = op_matrix(p, op, 2);
                out.push_back(one(G_DENSE, q[nc], ctrls, m[0], m[1], m[2], m[3]));
            } else {
                if (nt > 16) throw Error(QFVM_ERR_VALUE, "too many targets");
                Gate g;
                g.kind = G_MULTI;
                g.ctrls = ctrls;
                for (int c : ctrls) g.cmask |= 1ull << c;
                g.targets.assign(q.begin() + nc, q.end());
                g.mk = op_matrix(p, op, 1ull << nt);
                out.push_back(std::move(g));
            }
            return;
        }
        case QFVM_OP_X: need_q(1); out.push_back(one(G_SWAP, q[0], {}, 0.0, 1.0, 1.0, 0.0)); return;
        case QFVM_OP_Y: need_q(1); out.push_back(one(G_DENSE, q[0], {}, 0.0, -kI, kI, 0.0)); return;
        case QFVM_OP_Z: need_q(1); out.push_back(one(G_DIAG, q[0], {}, 1.0, -1.0)); return;
        case QFVM_OP_H: {
            need_q(1);
            const double r = 1. / std::sqrt(2.);
            out.push_back(one(G_REAL, q[0], {}, r, r, r, -r));
            return;
        }
        case QFVM_OP_S: need_q(1); out.push_back(one(G_DIAG, q[0], {}, 1.0, kI)); return;
        case QFVM_OP_SDG: need_q(1); out.push_back(one(G_DIAG, q[0], {}, 1.0, -kI)); return;
        case QFVM_OP_T: need_q(1); out.push_back(one(G_DIAG, q[0], {}, 1.0, std::exp(kI * kPi / 4.))); return;
        case QFVM_OP_TDG: need_q(1); out.push_back(one(G_DIAG, q[0], {}, 1.0, std::exp(-kI * kPi / 4.))); return;
        case QFVM_OP_P: need_q(1); out.push_back(one(G_DIAG, q[0], {}, 1.0, std::exp(kI * th))); return;
        case QFVM_OP_RX: {
            need_q(1);
            const cd c = std::cos(th / 2), sn = -kI * std::sin(th / 2);
            out.push_back(one(G_DENSE, q[0], {}, c, sn, sn, c));
            return;
        }
        case QFVM_OP_RY: {
            need_q(1);
            out.push_back(one(G_REAL, q[0], {}, std::cos(th / 2), -std::sin(th / 2), std::sin(th / 2), std::cos(th / 2)));
            return;
        }
        case QFVM_OP_RZ:
            need_q(1);
            out.push_back(one(G_DIAG, q[0], {}, std::exp(-kI * th / 2.), std::exp(kI * th / 2.)));
            return;
        case QFVM_OP_CX: need_q(2); out.push_back(one(G_SWAP, q[1], {q[0]}, 0.0, 1.0, 1.0, 0.0)); return;
        case QFVM_OP_CZ: need_q(2); out.push_back(one(G_DIAG, q[1], {q[0]}, 1.0, -1.0)); return;
        case QFVM_OP_CP: need_q(2); out.push_back(one(G_DIAG, q[1], {q[0]}, 1.0, std::exp(kI * th))); return;
        case QFVM_OP_CCX: need_q(3); out.push_back(one(G_SWAP, q[2], {q[0], q[1]}, 0.0, 1.0, 1.0, 0.0)); return;
        case QFVM_OP_RZZ: {  // simulator.hpp:90-94
            need_q(2);
            out.push_back(one(G_SWAP, q[1], {q[0]}, 0.0, 1.0, 1.0, 0.0));
            out.push_back(one(G_DIAG, q[1], {}, std::exp(-kI * th / 2.), std::exp(kI * th / 2.)));
            out.push_back(one(G_SWAP, q[1], {q[0]}, 0.0, 1.0, 1.0, 0.0));
            return;
        }
        default:
            throw Error(QFVM_ERR_VALUE, "op code " + std::to_string(op.code) + " is not a unitary gate");
    }
}

bool is_gate_code(int c) { return c >= QFVM_OP_GENERAL && c <= QFVM_OP_RZZ; }

// U^dagger of an internal gate (the reference builds op.dagger() per gate, gradiant.py:163-172)
Gate dagger(const Gate& g) {
    Gate d = g;
    switch (g.kind) {
        case G_DENSE:
            d.m[0] = std::conj(g.m[0]);
            d.m[1] = std::conj(g.m[2]);
            d.m[2] = std::conj(g.m[1]);
            d.m[3] = std::conj(g.m[3]);
            break;
        case G_REAL:
            d.m[1] = g.m[2];
            d.m[2] = g.m[1];
            break;
        case G_DIAG:
            d.m[0] = std::conj(g.m[0]);
            d.m[1] = std::conj(g.m[1]);
            break;
        case G_MULTI: {
            const size_t dim = (size_t)1 << g.targets.size();
            for (size_t r = 0; r < dim; r++)
                for (size_t c = 0; c < dim; c++) d.mk[r * dim + c] = std::conj(g.mk[c * dim + r]);
            break;
        }
        default: break;  // X is its own inverse
    }
    return d;
}

// Pauli-sum observable from the flat term arrays of qfvm_expect_paulis plus complex coefficients;
// terms grouped by flip mask in ascending order.
Hamiltonian parse_hamiltonian(const State& s, const char* chars, const int32_t* pos, const int64_t* off,
                              const int32_t* len, const double* coef, int64_t n_terms) {
    if (n_terms < 0) throw Error(QFVM_ERR_VALUE, "negative term count");
    if (n_terms > 0 && (!chars || !pos || !off || !len || !coef)) throw Error(QFVM_ERR_VALUE, "null term array");
    std::vector<std::pair<uint64_t, HamTerm>> all;
    for (int64_t t = 0; t < n_terms; t++) {
        uint64_t flip = 0, zy = 0, seen = 0;
        int ny = 0;
        for (int32_t k = 0; k < len[t]; k++) {
            const int q = pos[off[t] + k];
            if (q < 0 || q >= s.n) throw Error(QFVM_ERR_INDEX, "pauli qubit out of range");
            if ((seen >> q) & 1) throw Error(QFVM_ERR_VALUE, "repeated qubit in a Pauli term");
            seen |= 1ull << q;
            switch (chars[off[t] + k]) {
                case 'X': flip |= 1ull << q; break;
                case 'Y': flip |= 1ull << q; zy |= 1ull << q; ny++; break;
                case 'Z': zy |= 1ull << q; break;
                case 'I': break;
                default: throw Error(QFVM_ERR_VALUE, "Pauli letters are I, X, Y, Z");
            }
        }
        cd c(coef[2 * t], coef[2 * t + 1]);
        for (int k = 0; k < (ny & 3); k++) c *= kI;
        all.push_back({flip, HamTerm{zy, c.real(), c.imag()}});
    }
    std::stable_sort(all.begin(), all.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    Hamiltonian H;
    for (const auto& e : all) {
        if (H.groups.empty() || H.groups.back().flip != e.first)
            H.groups.push_back(HamGroup{e.first, (int32_t)H.terms.size(), 0});
        H.terms.push_back(e.second);
        H.groups.back().count++;
    }
    return H;
}

void check_program(const qfvm_program* p) {
    if (!p) throw Error(QFVM_ERR_VALUE, "null program");
    if (p->n_ops < 0 || (p->n_ops > 0 && !p->ops)) throw Error(QFVM_ERR_VALUE, "bad op array");
}

// std::discrete_distribution as implemented by libstdc++ (<bits/random.tcc>), which the reference's
// sample_measure_probs uses (statevector.hpp:1339-1340): normalise by the sum, partial sums, the
// last set to 1, first index with u <= cp[i].
size_t discrete_pick(const std::vector<double>& pr, double u) {
    if (pr.size() < 2) return 0;
    double sum = 0.0;
    for (double x : pr) sum += x;
    std::vector<double> cp(pr.size());
    double acc = 0.0;
    for (size_t i = 0; i < pr.size(); i++) {
        acc += pr[i] / sum;
        cp[i] = acc;
    }
    cp.back() = 1.0;
    return (size_t)(std::lower_bound(cp.begin(), cp.end(), u) - cp.begin());
}

struct Runner {
    State& s;
    const qfvm_program* p;
    const double* u;
    int64_t nu;
    int64_t ui;
    int32_t* creg;
    int32_t ncreg;
    bool skip_measure;
    std::vector<Gate> pending;

    double next_uniform() {
        if (ui >= nu) throw Error(QFVM_ERR_VALUE, "not enough uniforms for measure/reset");
        return u[ui++];
    }

    void flush() {
        if (!pending.empty()) {
            apply_circuit(s, pending);
            pending.clear();
        }
    }

    void measure_like(const std::vector<int>& q, bool reset, const std::vector<int>* c) {
        flush();
        std::vector<double> pr;
        reduce_marginal(s, q, pr);
        const size_t m = discrete_pick(pr, next_uniform());
        const double prob = pr[m];
        if (!(prob > 0.0)) throw Error(QFVM_ERR_RUNTIME, "measured an outcome of zero probability");
        launch_collapse(s, q, m, 1.0 / std::sqrt(prob));  // update(), statevector.hpp:1191-1199
        if (reset && m != 0) {
            if (q.size() == 1) {
                std::vector<Gate> x{one(G_SWAP, q[0], {}, 0.0, 1.0, 1.0, 0.0)};
                apply_gates(s, x);
            } else {
                launch_swap_states(s, q, 0, m);
            }
        }
        if (c)
            for (size_t j = 0; j < q.size(); j++) creg[(*c)[j]] = (int32_t)((m >> j) & 1);
    }

    // a noise channel on one qubit (include/qfvm.h: QFVM_OP_KRAUS / QFVM_OP_UCHAN)
    void channel(const qfvm_op& op) {
        std::vector<int> q = op_list(p, op.q_off, op.n_qubits);
        check_qubits(s, q);
        if (q.size() != 1) throw Error(QFVM_ERR_VALUE, "noise channels act on one qubit");
        const int64_t nk = op.body;
        if (nk < 1 || nk > 64) throw Error(QFVM_ERR_VALUE, "channel operator count out of range");
        const int64_t need_d = 2 * (4 * nk + (op.code == QFVM_OP_UCHAN ? nk : 0));
        if (op.m_off < 0 || op.m_off + need_d > p->n_mat || !p->mat) throw Error(QFVM_ERR_VALUE, "matrix pool range");
        auto mat = [&](int64_t k, cd m[4]) {
            for (int e = 0; e < 4; e++)
                m[e] = cd(p->mat[op.m_off + 8 * k + 2 * e], p->mat[op.m_off + 8 * k + 2 * e + 1]);
        };
        const double u = next_uniform();
        flush();
        cd m[4];
        if (op.code == QFVM_OP_UCHAN) {  // numpy's choice: cdf / cdf[-1], first k with u < cdf[k]
            std::vector<double> cdf((size_t)nk);
            double acc = 0.0;
            for (int64_t k = 0; k < nk; k++) {
                acc += p->mat[op.m_off + 8 * nk + 2 * k];
                cdf[(size_t)k] = acc;
            }
            int64_t pick = nk - 1;
            for (int64_t k = 0; k < nk; k++)
                if (u < cdf[(size_t)k] / cdf.back()) {
                    pick = k;
                    break;
                }
            mat(pick, m);
            pending.push_back(one(G_DENSE, q[0], {}, m[0], m[1], m[2], m[3]));
            return;
        }
        // Kraus: running sum of |K_k psi| / |psi| (simulator.py:59-74)
        const double norm0 = std::sqrt(reduce_norm2(s));
        double acc = 0.0;
        for (int64_t k = 0; k < nk; k++) {
            mat(k, m);
            const double norm1 = std::sqrt(reduce_op_norm2(s, q[0], m));
            acc += norm1 / norm0;
            if (u < acc) {
                const double inv = 1.0 / norm1;
                pending.push_back(one(G_DENSE, q[0], {}, m[0] * inv, m[1] * inv, m[2] * inv, m[3] * inv));
                return;
            }
        }
    }

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
                if (c.size() < q.size()) throw Error(QFVM_ERR_VALUE, "measure needs one cbit per qubit");
                for (size_t j = 0; j < q.size(); j++)
                    if (c[j] < 0 || c[j] >= ncreg) throw Error(QFVM_ERR_INDEX, "cbit out of range");
                measure_like(q, false, &c);
            } else if (op.code == QFVM_OP_RESET) {
                k++;
                std::vector<int> q = op_list(p, op.q_off, op.n_qubits);
                check_qubits(s, q);
                measure_like(q, true, nullptr);
            } else if (op.code == QFVM_OP_KRAUS || op.code == QFVM_OP_UCHAN) {
                k++;
                channel(op);
            } else if (op.code == QFVM_OP_CIF) {
                std::vector<int> c = op_list(p, op.c_off, op.n_cbits);
                const int64_t body_end = k + 1 + op.body;
                if (op.body < 0 || body_end > end) throw Error(QFVM_ERR_VALUE, "cif body range");
                uint64_t out = 0;
                for (int cb : c) {
                    if (cb < 0 || cb >= ncreg) throw Error(QFVM_ERR_INDEX, "cbit out of range");
                    out = out * 2 + (uint64_t)creg[cb];
                }
                if ((int64_t)out == op.cond) exec(k + 1, body_end);
                k = body_end;
            } else {
                lower_op(s, p, op, pending);
                k++;
            }
        }
    }
};

}  // namespace

// ------------------------------------------------------------------------------------------------
void* State::ensure_scratch(size_t bytes) {
    if (bytes > scratch_bytes) {
        before_alloc();
        if (scratch) {
            QFVM_HIP(hipStreamSynchronize(stream));
            QFVM_HIP(hipFree(scratch));
            scratch = nullptr;
            scratch_bytes = 0;
        }
        QFVM_HIP(hipMalloc(&scratch, bytes));
        scratch_bytes = bytes;
    }
    return scratch;
}

double* State::ensure_pinned(size_t bytes) {
    if (bytes > host_pinned_bytes) {
        before_alloc();
        if (host_pinned) {
            QFVM_HIP(hipStreamSynchronize(stream));
            QFVM_HIP(hipHostFree(host_pinned));
            host_pinned = nullptr;
            host_pinned_bytes = 0;
        }
        QFVM_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_pinned), bytes, hipHostMallocDefault));
        host_pinned_bytes = bytes;
    }
    return host_pinned;
}

void* State::ensure_ops(size_t bytes) {
    if (bytes > ops_bytes) {
        before_alloc();
        if (ops_buf) {
            QFVM_HIP(hipStreamSynchronize(stream));
            QFVM_HIP(hipFree(ops_buf));
            ops_buf = nullptr;
            ops_bytes = 0;
        }
        const size_t want = std::max<size_t>(bytes, 1 << 20);
        QFVM_HIP(hipMalloc(&ops_buf, want));
        ops_bytes = want;
    }
    return ops_buf;
}

void* State::ensure_aux(size_t bytes) {
    if (bytes > aux_bytes) {
        before_alloc();
        if (aux_buf) {
            QFVM_HIP(hipStreamSynchronize(stream));
            QFVM_HIP(hipFree(aux_buf));
            aux_buf = nullptr;
            aux_bytes = 0;
        }
        const size_t want = std::max<size_t>(bytes, 1 << 16);
        QFVM_HIP(hipMalloc(&aux_buf, want));
        aux_bytes = want;
    }
    return aux_buf;
}

unsigned char* State::stage_host(int k, size_t bytes) {
    Stage& st = stages[k];
    if (capturing) {
        // a captured copy reads its host area when the graph runs, so every upload of one capture
        // gets its own stretch of the capture arena (one slot may upload several times per call)
        const size_t off = (cap_off + 255) / 256 * 256;
        if (off + bytes > cap_host_bytes) {
            cap_need = std::max({cap_need, 2 * (off + bytes), (size_t)1 << 20});
            throw NeedsAlloc{};
        }
        cap_off = off + bytes;
        st.cur_host = cap_host + off;
        return st.cur_host;
    }
    if (st.done) QFVM_HIP(hipEventSynchronize(st.done));
    if (bytes > st.host_bytes) {
        before_alloc();
        if (st.host) QFVM_HIP(hipHostFree(st.host));
        st.host = nullptr;
        st.host_bytes = 0;
        const size_t want = std::max<size_t>(bytes, 1 << 16);
        QFVM_HIP(hipHostMalloc(reinterpret_cast<void**>(&st.host), want, hipHostMallocDefault));
        st.host_bytes = want;
    }
    st.cur_host = st.host;
    return st.host;
}

unsigned char* State::stage_dev(int k, size_t bytes) {
    Stage& st = stages[k];
    if (bytes > st.dev_bytes) {
        before_alloc();
        if (st.dev) {
            QFVM_HIP(hipStreamSynchronize(stream));  // queued kernels may still read the old area
            QFVM_HIP(hipFree(st.dev));
        }
        st.dev = nullptr;
        st.dev_bytes = 0;
        const size_t want = std::max<size_t>(bytes, 1 << 16);
        QFVM_HIP(hipMalloc(reinterpret_cast<void**>(&st.dev), want));
        st.dev_bytes = want;
    }
    return st.dev;
}

void State::stage_upload(int k, size_t bytes) {
    Stage& st = stages[k];
    QFVM_HIP(hipMemcpyAsync(st.dev, st.cur_host, bytes, hipMemcpyHostToDevice, stream));
    if (capturing) {  // the graph's copy node reads the arena when the graph runs
        const uint64_t rec[2] = {(uint64_t)k, (uint64_t)bytes};
        for (size_t b = 0; b < sizeof(rec); b++) cap_sig = (cap_sig ^ reinterpret_cast<const unsigned char*>(rec)[b]) * 1099511628211ull;
        return;
    }
    if (!st.done) QFVM_HIP(hipEventCreateWithFlags(&st.done, hipEventDisableTiming));
    QFVM_HIP(hipEventRecord(st.done, stream));
}

void* State::ensure_psi2() {
    if (!psi2) {
        before_alloc();
        QFVM_HIP(hipMalloc(&psi2, size * amp_bytes()));
    }
    return psi2;
}

void* State::ensure_checkpoint(size_t k) {
    while (ckpt.size() <= k) ckpt.push_back(nullptr);
    if (!ckpt[k]) {
        before_alloc();
        QFVM_HIP(hipMalloc(&ckpt[k], size * amp_bytes()));
    }
    return ckpt[k];
}

void State::before_alloc() {
    if (capturing) throw NeedsAlloc{};
    buf_gen++;  // cached graphs name the old buffers
}

void State::drop_graphs() {
    for (GraphEntry& e : graph_cache)
        if (e.exec) hipGraphExecDestroy(e.exec);
    graph_cache.clear();
}

static void destroy_state(State* s) {
    if (!s) return;
    hipSetDevice(s->device);
    if (s->stream) hipStreamSynchronize(s->stream);
    s->resolve_events();
    s->drop_graphs();
    if (s->cap_host) hipHostFree(s->cap_host);
    if (s->psi) hipFree(s->psi);
    if (s->psi2) hipFree(s->psi2);
    for (void* c : s->ckpt)
        if (c) hipFree(c);
    for (Stage& st : s->stages) {
        if (st.host) hipHostFree(st.host);
        if (st.dev) hipFree(st.dev);
        if (st.done) hipEventDestroy(st.done);
    }
    if (s->aux_buf) hipFree(s->aux_buf);
    if (s->ops_buf) hipFree(s->ops_buf);
    if (s->scratch) hipFree(s->scratch);
    if (s->host_pinned) hipHostFree(s->host_pinned);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
}

// ------------------------------------------------------------------------------------------------
// Graph mode. A qfvm_adjoint call queues 20 to 200 launches plus three uploads and one read-back;
// with option "graphs" the call captures that work on its stream (relaxed capture: the host plans
// and fills the pinned upload slots exactly as in an eager call) and launches it as one executable
// graph. Executable graphs are cached per program structure (opcodes, qubits, differentiated ops,
// Pauli strings; not angles, matrices or coefficients): a later call of the same structure updates
// the cached graph's node parameters (hipGraphExecUpdate) instead of instantiating a new one, so a
// training loop over fixed circuits with new parameters replays graphs. Captured graphs hold kernel
// and copy nodes only (zero fills are kernels). Work that would allocate during the capture
// abandons it and runs eagerly, which grows the buffer; a (re)allocation drops the handle's cached
// graphs, which name the old buffers.

constexpr size_t GRAPH_CACHE_MAX = 64;

static uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t k = 0; k < n; k++) h = (h ^ b[k]) * 1099511628211ull;
    return h;
}

// structure of one qfvm_adjoint call (everything that shapes the plan except parameter values)
static uint64_t adjoint_key(const State& s, const qfvm_program* prog, const int64_t* grad_ops, const char* grad_gen,
                            int64_t n_grads, const char* chars, const int32_t* pos, const int64_t* off,
                            const int32_t* len, int64_t n_terms) {
    uint64_t h = 1469598103934665603ull;
    const int32_t opts[] = {s.n, s.dtype, s.mfma, s.fusion, s.tile_bits, s.low_bits, s.max_groups, s.pass_cost, s.split_rows,
                            s.stream_tiles, s.stagger, s.stagger_shift, s.rows8, s.rows8_c64, s.layout};
    h = fnv(h, opts, sizeof(opts));
    h = fnv(h, &prog->n_ops, sizeof(prog->n_ops));
    for (int64_t k = 0; k < prog->n_ops; k++) {
        const qfvm_op& o = prog->ops[k];
        const int32_t f[4] = {o.code, o.n_ctrl, o.n_qubits, o.n_cbits};
        h = fnv(h, f, sizeof(f));
        if (o.n_qubits > 0) h = fnv(h, prog->idx + o.q_off, (size_t)o.n_qubits * sizeof(int32_t));
    }
    h = fnv(h, &n_grads, sizeof(n_grads));
    if (n_grads > 0) {
        h = fnv(h, grad_ops, (size_t)n_grads * sizeof(int64_t));
        h = fnv(h, grad_gen, (size_t)n_grads);
    }
    h = fnv(h, &n_terms, sizeof(n_terms));
    for (int64_t t = 0; t < n_terms; t++) {
        h = fnv(h, &len[t], sizeof(int32_t));
        if (len[t] > 0) {
            h = fnv(h, chars + off[t], (size_t)len[t]);
            h = fnv(h, pos + off[t], (size_t)len[t] * sizeof(int32_t));
        }
    }
    return h;
}

// Shape of a captured graph: node count and types, every kernel node's function, grid, block and
// dynamic LDS bytes, the edge count, and (cap_sig, accumulated by stage_upload during the capture)
// the size of every upload. The plan of a call depends on parameter values as well as on the
// program's structure (an RX(0) plans as a diagonal gate, diagonal runs merge terms by value), so
// two calls with one structure key may capture different graphs; a cached executable is updated
// in place only from a capture of the same shape, which is what hipGraphExecUpdate supports.
static uint64_t graph_signature(hipGraph_t g, uint64_t seed) {
    size_t count = 0;
    QFVM_HIP(hipGraphGetNodes(g, nullptr, &count));
    std::vector<hipGraphNode_t> nodes(count);
    if (count) QFVM_HIP(hipGraphGetNodes(g, nodes.data(), &count));
    uint64_t h = fnv(seed, &count, sizeof(count));
    for (hipGraphNode_t nd : nodes) {
        hipGraphNodeType t{};
        QFVM_HIP(hipGraphNodeGetType(nd, &t));
        const int32_t ti = (int32_t)t;
        h = fnv(h, &ti, sizeof(ti));
        if (t == hipGraphNodeTypeKernel) {
            hipKernelNodeParams kp{};
            QFVM_HIP(hipGraphKernelNodeGetParams(nd, &kp));
            const uint64_t f[8] = {(uint64_t)(uintptr_t)kp.func, kp.gridDim.x, kp.gridDim.y, kp.gridDim.z,
                                   kp.blockDim.x, kp.blockDim.y, kp.blockDim.z, kp.sharedMemBytes};
            h = fnv(h, f, sizeof(f));
        }
    }
    size_t edges = 0;
    QFVM_HIP(hipGraphGetEdges(g, nullptr, nullptr, &edges));
    return fnv(h, &edges, sizeof(edges));
}

// dense multi-target gates stage their matrix from pageable host memory: such calls run eagerly
static bool graph_safe(const std::vector<Gate>& gates) {
    for (const Gate& g : gates)
        if (g.kind == G_MULTI) return false;
    return true;
}

// Captures `enqueue` on the state's stream and launches it as a graph. Returns false, with
// nothing queued, when the work needs an allocation first.
template <typename F>
static bool run_graph(State& s, uint64_t key, F&& enqueue) {
    if (s.graph_gen != s.buf_gen) {  // buffers were reallocated since the cached graphs were made
        s.drop_graphs();
        s.graph_gen = s.buf_gen;
    }
    // an eager call may still be copying out of an upload slot, which the capture does not wait for
    for (Stage& st : s.stages)
        if (st.done) QFVM_HIP(hipEventSynchronize(st.done));
    s.cap_need = std::max(s.cap_need, (size_t)1 << 20);
    if (s.cap_need > s.cap_host_bytes) {  // first capture, or an earlier one ran out of arena
        s.before_alloc();
        s.drop_graphs();  // their copy nodes read the old arena
        s.graph_gen = s.buf_gen;
        if (s.cap_host) QFVM_HIP(hipHostFree(s.cap_host));
        s.cap_host = nullptr;
        s.cap_host_bytes = 0;
        QFVM_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.cap_host), s.cap_need, hipHostMallocDefault));
        s.cap_host_bytes = s.cap_need;
    }
    s.cap_off = 0;
    s.cap_sig = 1469598103934665603ull;
    hipGraph_t graph = nullptr;
    QFVM_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeRelaxed));
    s.capturing = true;
    try {
        enqueue();
    } catch (...) {
        s.capturing = false;
        hipGraph_t junk = nullptr;
        const hipError_t e = hipStreamEndCapture(s.stream, &junk);
        if (e == hipSuccess && junk) hipGraphDestroy(junk);
        (void)hipGetLastError();
        try {
            throw;
        } catch (const NeedsAlloc&) {
            return false;
        }
    }
    s.capturing = false;
    QFVM_HIP(hipStreamEndCapture(s.stream, &graph));
    struct GraphGuard {
        hipGraph_t g;
        ~GraphGuard() {
            if (g) hipGraphDestroy(g);
        }
    } guard{graph};
    const uint64_t sig = graph_signature(graph, s.cap_sig);
    State::GraphEntry* entry = nullptr;
    for (State::GraphEntry& e : s.graph_cache)
        if (e.key == key) entry = &e;
    bool ready = false;
    if (entry && entry->exec && entry->sig != sig) {  // same structure key, different plan: start afresh
        hipGraphExecDestroy(entry->exec);
        entry->exec = nullptr;
    }
    if (entry && entry->exec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(entry->exec, graph, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
            ready = true;
            s.kstats[(size_t)s.stat_index("graph_update")].second.launches++;
        } else {
            (void)hipGetLastError();
            hipGraphExecDestroy(entry->exec);
            entry->exec = nullptr;
        }
    } else if (!entry) {
        if (s.graph_cache.size() >= GRAPH_CACHE_MAX) {  // evict the least recently used graph
            size_t lru = 0;
            for (size_t k = 1; k < s.graph_cache.size(); k++)
                if (s.graph_cache[k].last_use < s.graph_cache[lru].last_use) lru = k;
            if (s.graph_cache[lru].exec) hipGraphExecDestroy(s.graph_cache[lru].exec);
            s.graph_cache.erase(s.graph_cache.begin() + (std::ptrdiff_t)lru);
        }
        s.graph_cache.push_back(State::GraphEntry{key, sig, 0, nullptr});
        entry = &s.graph_cache.back();
    }
    if (!ready) {
        entry->sig = sig;
        QFVM_HIP(hipGraphInstantiate(&entry->exec, graph, nullptr, nullptr, 0));
        s.kstats[(size_t)s.stat_index("graph_instantiate")].second.launches++;
    }
    entry->last_use = ++s.graph_clock;
    QFVM_HIP(hipGraphLaunch(entry->exec, s.stream));
    s.kstats[(size_t)s.stat_index("graph_launch")].second.launches++;
    return true;
}


extern "C" {

const char* qfvm_last_error(void) { return g_err.c_str(); }

int qfvm_device_count(int* out) {
    return guarded([&] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess) n = 0;
        *out = n;
    });
}

int qfvm_create(int32_t n_qubits, int32_t dtype, qfvm_handle* out) {
    return guarded([&] {
        if (!out) throw Error(QFVM_ERR_VALUE, "null output handle");
        *out = nullptr;
        if (n_qubits < 0 || n_qubits > 34) throw Error(QFVM_ERR_VALUE, "n_qubits must be in [0, 34]");
        if (dtype != QFVM_C64 && dtype != QFVM_C128) throw Error(QFVM_ERR_VALUE, "dtype must be QFVM_C64 or QFVM_C128");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw Error(QFVM_ERR_DEVICE, "no HIP device is visible to this process");
        std::unique_ptr<State, void (*)(State*)> s(new State(), destroy_state);
        QFVM_HIP(hipGetDevice(&s->device));
        s->n = n_qubits;
        s->dtype = dtype;
        s->size = 1ull << n_qubits;
        QFVM_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        QFVM_HIP(hipMalloc(&s->psi, s->size * s->amp_bytes()));
        launch_fill_basis(*s, 0);
        QFVM_HIP(hipStreamSynchronize(s->stream));
        *out = reinterpret_cast<qfvm_handle>(s.release());
    });
}

int qfvm_destroy(qfvm_handle h) {
    return guarded([&] { destroy_state(reinterpret_cast<State*>(h)); });
}

int qfvm_num_qubits(qfvm_handle h, int32_t* out) {
    return guarded([&] { *out = need(h)->n; });
}

int qfvm_dtype_of(qfvm_handle h, int32_t* out) {
    return guarded([&] { *out = need(h)->dtype; });
}

int qfvm_init_zero(qfvm_handle h) {
    return guarded([&] {
        State* s = need(h);
        launch_fill_basis(*s, 0);
        QFVM_HIP(hipStreamSynchronize(s->stream));
    });
}

int qfvm_init_basis(qfvm_handle h, uint64_t index) {
    return guarded([&] {
        State* s = need(h);
        if (index >= s->size) throw Error(QFVM_ERR_INDEX, "basis index out of range");
        launch_fill_basis(*s, index);
        QFVM_HIP(hipStreamSynchronize(s->stream));
    });
}

int qfvm_load(qfvm_handle h, const double* host, uint64_t n_amps) {
    return guarded([&] {
        State* s = need(h);
        if (n_amps != s->size) throw Error(QFVM_ERR_VALUE, "state length must be 2^n");
        if (!host) throw Error(QFVM_ERR_VALUE, "null host buffer");
        launch_h2d_convert(*s, host, n_amps);
    });
}

int qfvm_store(qfvm_handle h, double* host, uint64_t n_amps) {
    return guarded([&] {
        State* s = need(h);
        if (n_amps != s->size) throw Error(QFVM_ERR_VALUE, "state length must be 2^n");
        if (!host) throw Error(QFVM_ERR_VALUE, "null host buffer");
        launch_d2h_convert(*s, host, n_amps);
    });
}

int qfvm_apply(qfvm_handle h, const qfvm_program* prog) {
    return guarded([&] {
        State* s = need(h);
        check_program(prog);
        std::vector<Gate> gates;
        for (int64_t k = 0; k < prog->n_ops; k++) {
            if (!is_gate_code(prog->ops[k].code))
                throw Error(QFVM_ERR_VALUE, "qfvm_apply takes unitary ops only (use qfvm_run for measure/reset/cif)");
            lower_op(*s, prog, prog->ops[k], gates);
        }
        apply_circuit(*s, gates);
        QFVM_HIP(hipStreamSynchronize(s->stream));
    });
}

int qfvm_run(qfvm_handle h, const qfvm_program* prog, const double* uniforms, int64_t n_uniforms, int32_t* creg,
             int32_t n_creg, int32_t skip_measure) {
    return guarded([&] {
        State* s = need(h);
        check_program(prog);
        if (n_creg < 0 || (n_creg > 0 && !creg)) throw Error(QFVM_ERR_VALUE, "bad classical register");
        Runner r{*s, prog, uniforms, n_uniforms, 0, creg, n_creg, skip_measure != 0, {}};
        r.exec(0, prog->n_ops);
        r.flush();
        QFVM_HIP(hipStreamSynchronize(s->stream));
    });
}

// expect_pauli's masks (statevector.hpp:1252-1279): X and Y flip, Y and Z weigh by parity, the
// number of Y letters sets the phase. The reference adds the bits (`+= 1 << q`), so a qubit named
// twice carries into the next bit, possibly past the last qubit, and its scan then reads outside the
// state; such a term is rejected here.
static PauliHost parse_pauli(const State& s, const char* chars, const int32_t* pos, int64_t off, int32_t len,
                             int slot) {
    PauliHost t{};
    t.slot = slot;
    uint64_t seen = 0;
    for (int32_t k = 0; k < len; k++) {
        const int q = pos[off + k];
        if (q < 0 || q >= s.n) throw Error(QFVM_ERR_INDEX, "pauli qubit out of range");
        const char c = chars[off + k];
        if (c != 'X' && c != 'Y' && c != 'Z') continue;
        if ((seen >> q) & 1) throw Error(QFVM_ERR_VALUE, "repeated qubit in a Pauli term");
        seen |= 1ull << q;
        if (c == 'X' || c == 'Y') {
            t.flip += 1ull << q;
            t.flip_q = std::max(t.flip_q, q);
        }
        if (c == 'Y') t.y++;
        if (c == 'Y' || c == 'Z') t.z += 1ull << q;
    }
    return t;
}

int qfvm_expect_paulis(qfvm_handle h, const char* chars, const int32_t* pos, const int64_t* off, const int32_t* len,
                       int64_t n_terms, double* out) {
    return guarded([&] {
        State* s = need(h);
        if (n_terms < 0) throw Error(QFVM_ERR_VALUE, "negative term count");
        if (n_terms == 0) return;
        if (!chars || !pos || !off || !len || !out) throw Error(QFVM_ERR_VALUE, "null term array");
        std::vector<PauliHost> terms((size_t)n_terms);
        for (int64_t t = 0; t < n_terms; t++) terms[(size_t)t] = parse_pauli(*s, chars, pos, off[t], len[t], (int)t);
        double* out_dev = reinterpret_cast<double*>(s->ensure_aux((size_t)n_terms * sizeof(double)));
        double* partials = reinterpret_cast<double*>(s->ensure_scratch((EXPECT_PARTIALS + RED_PARTIALS) * sizeof(double)));
        for (int64_t t0 = 0; t0 < n_terms; t0 += EXPECT_BATCH) {
            const int cnt = (int)std::min<int64_t>(EXPECT_BATCH, n_terms - t0);
            queue_expectations(*s, terms.data() + t0, cnt, out_dev, partials, partials + EXPECT_PARTIALS);
        }
        double* hbuf = s->ensure_pinned((size_t)n_terms * sizeof(double));
        QFVM_HIP(hipMemcpyAsync(hbuf, out_dev, (size_t)n_terms * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        QFVM_HIP(hipStreamSynchronize(s->stream));
        std::memcpy(out, hbuf, (size_t)n_terms * sizeof(double));
    });
}

int qfvm_expect_paulis_noalloc(qfvm_handle h, const char* chars, const int32_t* pos, const int64_t* off,
                               const int32_t* len, int64_t n_terms, double* out) {
    return guarded([&] {
        State* s = need(h);
        if (n_terms < 0) throw Error(QFVM_ERR_VALUE, "negative term count");
        if (n_terms > EXPECT_BATCH)
            throw Error(QFVM_ERR_VALUE, "qfvm_expect_paulis_noalloc takes at most 256 terms (its scratch is fixed)");
        if (n_terms == 0) return;
        if (!chars || !pos || !off || !len || !out) throw Error(QFVM_ERR_VALUE, "null term array");
        PauliHost terms[EXPECT_BATCH];
        for (int64_t t = 0; t < n_terms; t++) terms[t] = parse_pauli(*s, chars, pos, off[t], len[t], (int)t);
        expect_paulis_noalloc(*s, terms, (int)n_terms, out);
    });
}

int qfvm_expect_hamiltonian(qfvm_handle h, const char* chars, const int32_t* pos, const int64_t* off,
                            const int32_t* len, const double* coef, int64_t n_terms, double* out) {
    return guarded([&] {
        State* s = need(h);
        if (!out) throw Error(QFVM_ERR_VALUE, "null output");
        const Hamiltonian H = parse_hamiltonian(*s, chars, pos, off, len, coef, n_terms);
        *out = H.groups.empty() ? 0.0 : apply_hamiltonian(*s, H, nullptr);
    });
}

int qfvm_adjoint(qfvm_handle h, const qfvm_program* prog, const int64_t* grad_ops, const char* grad_gen,
                 int64_t n_grads, const char* chars, const int32_t* pos, const int64_t* off, const int32_t* len,
                 const double* coef, int64_t n_terms, double* grads, double* expectation) {
    return guarded([&] {
        State* s = need(h);
        check_program(prog);
        if (n_grads < 0 || (n_grads > 0 && (!grad_ops || !grad_gen || !grads)))
            throw Error(QFVM_ERR_VALUE, "bad gradient arrays");
        const Hamiltonian H = parse_hamiltonian(*s, chars, pos, off, len, coef, n_terms);
        // forward gates, with the range of internal gates each op lowers to
        std::vector<Gate> fwd;
        std::vector<std::pair<size_t, size_t>> range((size_t)prog->n_ops);
        for (int64_t k = 0; k < prog->n_ops; k++) {
            if (!is_gate_code(prog->ops[k].code))
                throw Error(QFVM_ERR_VALUE, "qfvm_adjoint takes unitary ops only");
            const size_t a = fwd.size();
            lower_op(*s, prog, prog->ops[k], fwd);
            range[(size_t)k] = {a, fwd.size()};
        }
        // one marker per differentiated op: the generator on the op's single target, its controls
        // as a projector (grad_gate, gradiant.py:96-141)
        std::vector<int> marker_of((size_t)prog->n_ops, -1);
        int64_t begin = prog->n_ops;
        std::vector<Gate> markers((size_t)n_grads);
        for (int64_t g = 0; g < n_grads; g++) {
            const int64_t k = grad_ops[g];
            if (k < 0 || k >= prog->n_ops) throw Error(QFVM_ERR_INDEX, "gradient op index out of range");
            if (marker_of[(size_t)k] >= 0) throw Error(QFVM_ERR_VALUE, "op differentiated twice");
            const auto r = range[(size_t)k];
            if (r.second - r.first != 1 || fwd[r.first].kind == G_MULTI)
                throw Error(QFVM_ERR_VALUE, "only one-target rotations can be differentiated");
            Gate m;
            m.kind = G_GRAD;
            m.target = fwd[r.first].target;
            m.ctrls = fwd[r.first].ctrls;
            m.cmask = fwd[r.first].cmask;
            switch (grad_gen[g]) {
                case 'X': m.gen = GEN_X; break;
                case 'Y': m.gen = GEN_Y; break;
                case 'Z': m.gen = GEN_Z; break;
                default: throw Error(QFVM_ERR_VALUE, "gradient generators are X, Y or Z");
            }
            m.slot = (int)g;
            markers[(size_t)g] = m;
            marker_of[(size_t)k] = (int)g;
            begin = std::min(begin, k);
        }
        // everything below is queued on the stream; the host plans the reverse sweep while the
        // forward passes run, and waits once at the end. Results: grads[0 .. n), <H> at [n].
        const size_t rbytes = ((size_t)n_grads + 1) * sizeof(double);
        double* rdev = reinterpret_cast<double*>(s->ensure_aux(rbytes));
        double* hbuf = reinterpret_cast<double*>(s->stage_host(STAGE_RESULT, rbytes));
        // Checkpoints: K op boundaries b_0 < .. < b_{K-1} spread evenly over the differentiated range
        // [begin, n_ops); the forward sweep copies psi at each (the state before op b_j) and the
        // reverse sweep, having uncomputed op b_j, restores psi from that copy, so rounding from
        // U^dagger does not accumulate across segments. K shrinks to what free device memory holds
        // beyond the copies already allocated (keeping a reserve of 1/32 of the device memory).
        const size_t sbytes = s->size * s->amp_bytes();
        int64_t K = n_grads > 0 ? std::min<int64_t>(s->checkpoints, prog->n_ops - begin - 1) : 0;
        if (K > (int64_t)s->ckpt.size()) {
            size_t free_b = 0, total_b = 0;
            QFVM_HIP(hipMemGetInfo(&free_b, &total_b));
            const size_t reserve = total_b / 32;
            const int64_t fit = free_b > reserve ? (int64_t)((free_b - reserve) / sbytes) : 0;
            K = std::min<int64_t>(K, (int64_t)s->ckpt.size() + fit);
        }
        K = std::max<int64_t>(K, 0);
        std::vector<int64_t> bound((size_t)K);
        for (int64_t j = 0; j < K; j++) bound[(size_t)j] = begin + (j + 1) * (prog->n_ops - begin) / (K + 1);
        // reverse sweep (gradiant.py:160-172): the marker of op k sits before U_k^dagger, where both
        // states are taken just after U_k; one gate list per segment between checkpoints
        std::vector<std::vector<Gate>> rev((size_t)K + 1);
        {
            int64_t seg = K;
            for (int64_t k = prog->n_ops - 1; k >= begin; k--) {
                if (marker_of[(size_t)k] >= 0) rev[(size_t)seg].push_back(markers[(size_t)marker_of[(size_t)k]]);
                const auto r = range[(size_t)k];
                for (size_t j = r.second; j > r.first; j--) rev[(size_t)seg].push_back(dagger(fwd[j - 1]));
                if (seg > 0 && k == bound[(size_t)seg - 1]) seg--;
            }
        }
        auto enqueue = [&] {
            launch_zero(*s, rdev, rbytes);
            size_t g0 = 0;
            for (int64_t j = 0; j < K; j++) {
                const size_t g1 = range[(size_t)bound[(size_t)j]].first;
                apply_gates(*s, std::vector<Gate>(fwd.begin() + (std::ptrdiff_t)g0, fwd.begin() + (std::ptrdiff_t)g1));
                void* c = s->ensure_checkpoint((size_t)j);
                QFVM_HIP(hipMemcpyAsync(c, s->psi, sbytes, hipMemcpyDeviceToDevice, s->stream));
                s->kstats[(size_t)s->stat_index("checkpoint_copy")].second.launches++;
                g0 = g1;
            }
            apply_gates(*s, K > 0 ? std::vector<Gate>(fwd.begin() + (std::ptrdiff_t)g0, fwd.end()) : fwd);
            void* lam = s->ensure_psi2();
            if (H.groups.empty())
                launch_zero(*s, lam, sbytes);
            else
                apply_hamiltonian(*s, H, lam, rdev + n_grads);
            if (n_grads > 0) {
                for (int64_t seg = K; seg >= 0; seg--) {
                    apply_gates2(*s, rev[(size_t)seg], lam, rdev);
                    if (seg > 0) {
                        QFVM_HIP(hipMemcpyAsync(s->psi, s->ckpt[(size_t)seg - 1], sbytes, hipMemcpyDeviceToDevice,
                                                s->stream));
                        s->kstats[(size_t)s->stat_index("checkpoint_restore")].second.launches++;
                    }
                }
            }
            QFVM_HIP(hipMemcpyAsync(hbuf, rdev, rbytes, hipMemcpyDeviceToHost, s->stream));
        };
        bool queued = false;
        if (s->graphs && !s->profiling && graph_safe(fwd)) {
            uint64_t key = adjoint_key(*s, prog, grad_ops, grad_gen, n_grads, chars, pos, off, len, n_terms);
            key = fnv(key, &K, sizeof(K));
            queued = run_graph(*s, key, enqueue);
        }
        if (!queued) enqueue();
        QFVM_HIP(hipStreamSynchronize(s->stream));
        if (n_grads > 0) std::memcpy(grads, hbuf, (size_t)n_grads * sizeof(double));
        if (expectation) *expectation = hbuf[n_grads];
    });
}

int qfvm_probabilities(qfvm_handle h, double* out, uint64_t n) {
    return guarded([&] {
        State* s = need(h);
        if (n != s->size) throw Error(QFVM_ERR_VALUE, "output length must be 2^n");
        reduce_probabilities(*s, out);
    });
}

int qfvm_marginals(qfvm_handle h, const int32_t* qubits, int32_t n_qubits, double* out) {
    return guarded([&] {
        State* s = need(h);
        if (n_qubits < 0 || n_qubits > s->n || n_qubits > 30) throw Error(QFVM_ERR_VALUE, "bad qubit count");
        if (!out || (n_qubits > 0 && !qubits)) throw Error(QFVM_ERR_VALUE, "null buffer");
        std::vector<int> q(qubits, qubits + n_qubits);
        check_qubits(*s, q);
        std::vector<double> probs;
        reduce_marginal(*s, q, probs);
        std::memcpy(out, probs.data(), probs.size() * sizeof(double));
    });
}

int qfvm_sample(qfvm_handle h, const double* uniforms, int64_t shots, uint64_t* out_index) {
    return guarded([&] {
        State* s = need(h);
        if (shots < 0) throw Error(QFVM_ERR_VALUE, "negative shot count");
        if (shots > 0 && (!uniforms || !out_index)) throw Error(QFVM_ERR_VALUE, "null buffer");
        sample_sequential_rule(*s, uniforms, shots, out_index);
    });
}

int qfvm_gather(qfvm_handle h, const uint64_t* indices, int64_t n, double* out) {
    return guarded([&] {
        State* s = need(h);
        if (n < 0 || (n > 0 && (!indices || !out))) throw Error(QFVM_ERR_VALUE, "bad gather arrays");
        for (int64_t k = 0; k < n; k++)
            if (indices[k] >= s->size) throw Error(QFVM_ERR_INDEX, "amplitude index out of range");
        if (n > 0) gather_amplitudes(*s, indices, (uint64_t)n, out);
    });
}

int qfvm_inner(qfvm_handle a, qfvm_handle b, double* re, double* im) {
    return guarded([&] {
        State* sa = need(a);
        State* sb = need(b);
        if (!re || !im) throw Error(QFVM_ERR_VALUE, "null output");
        if (sa->n != sb->n) throw Error(QFVM_ERR_VALUE, "states differ in qubit count");
        if (sa->device != sb->device) throw Error(QFVM_ERR_VALUE, "states live on different devices");
        reduce_inner(*sa, *sb, re, im);
    });
}

int qfvm_norm2(qfvm_handle h, double* out) {
    return guarded([&] { *out = reduce_norm2(*need(h)); });
}

int qfvm_synchronize(qfvm_handle h) {
    return guarded([&] {
        State* s = need(h);
        QFVM_HIP(hipStreamSynchronize(s->stream));
        s->resolve_events();
    });
}

int qfvm_set_option(qfvm_handle h, const char* key, int64_t value) {
    return guarded([&] {
        State* s = need(h);
        if (!key) throw Error(QFVM_ERR_VALUE, "null option key");
        const std::string k(key);
        if (k == "fusion") {
            s->fusion = value != 0;
        } else if (k == "tile_bits") {
            if (value != 0 && (value < 4 || value > 14)) throw Error(QFVM_ERR_VALUE, "tile_bits must be 0 or 4..14");
            if (value != 0 && ((1ll << value) * (int64_t)s->amp_bytes() > 160 * 1024))
                throw Error(QFVM_ERR_VALUE, "tile of 2^tile_bits amplitudes does not fit in 160 KiB of LDS");
            s->tile_bits = (int)value;
        } else if (k == "low_bits") {
            if (value < 0 || value > 14) throw Error(QFVM_ERR_VALUE, "low_bits must be in 0..14");
            s->low_bits = (int)value;
        } else if (k == "max_groups") {
            if (value < 1 || value > 24) throw Error(QFVM_ERR_VALUE, "max_groups must be in 1..24");
            s->max_groups = (int)value;
        } else if (k == "pass_cost") {
            if (value < 0 || value > (1 << 20)) throw Error(QFVM_ERR_VALUE, "pass_cost must be in 0..2^20");
            s->pass_cost = (int)value;
        } else if (k == "split_rows") {
            s->split_rows = value != 0;
        } else if (k == "rows8") {
            s->rows8 = value != 0;
        } else if (k == "rows8_c64") {
            s->rows8_c64 = value != 0;
        } else if (k == "layout") {
            s->layout = value != 0;
        } else if (k == "plan_cache") {
            s->plan_cache_on = value != 0;
            if (!value) s->plan_cache.reset();
        } else if (k == "stream") {
            s->stream_tiles = value != 0;
        } else if (k == "stagger") {
            if (value < 0 || value > 64) throw Error(QFVM_ERR_VALUE, "stagger must be in 0..64");
            s->stagger = (int)value;
        } else if (k == "graphs") {
            s->graphs = value != 0;
        } else if (k == "mfma") {
            s->mfma = value != 0;
        } else if (k == "dense_blocks") {
            if (value != 0 && (value < 4 || value > 8)) throw Error(QFVM_ERR_VALUE, "dense_blocks must be 0 or 4..8");
            s->dense_blocks = (int)value;
        } else if (k == "merge_1q") {
            s->merge_1q = value != 0;
        } else if (k == "dense_min_gates") {
            if (value < 1 || value > (1 << 20)) throw Error(QFVM_ERR_VALUE, "dense_min_gates must be in 1..2^20");
            s->dense_min_gates = (int)value;
        } else if (k == "checkpoints") {
            if (value < 0 || value > 64) throw Error(QFVM_ERR_VALUE, "checkpoints must be in 0..64");
            s->checkpoints = (int)value;
        } else if (k == "stagger_shift") {
            if (value < 0 || value > 16) throw Error(QFVM_ERR_VALUE, "stagger_shift must be in 0..16");
            s->stagger_shift = (int)value;
        } else {
            throw Error(QFVM_ERR_VALUE, "unknown option '" + k + "'");
        }
    });
}

int qfvm_plan_json(int32_t n, int32_t dtype, int32_t tile_bits, int32_t low_bits, int32_t max_groups,
                   int32_t pass_cost, int32_t layout, int32_t dense_blocks, int32_t dense_min_gates,
                   const qfvm_program* prog, char* buf, int64_t cap, int64_t* len) {
    return guarded([&] {
        if (n < 0 || n > 34) throw Error(QFVM_ERR_VALUE, "n_qubits must be in [0, 34]");
        check_program(prog);
        State tmp;
        tmp.n = n;
        tmp.dtype = dtype == QFVM_C64 ? QFVM_C64 : QFVM_C128;
        tmp.size = 1ull << n;
        std::vector<Gate> gates;
        for (int64_t k = 0; k < prog->n_ops; k++) {
            if (!is_gate_code(prog->ops[k].code)) throw Error(QFVM_ERR_VALUE, "unitary ops only");
            lower_op(tmp, prog, prog->ops[k], gates);
        }
        if (dense_blocks < 0 || (dense_blocks != 0 && (dense_blocks < 4 || dense_blocks > 8)))
            throw Error(QFVM_ERR_VALUE, "dense_blocks must be 0 or 4..8");
        // as qfvm_apply: one-qubit runs merged (option merge_1q, on by default), then deep runs
        // collapsed when asked
        gates = merge_single_qubit_runs(gates);
        if (dense_blocks) gates = collapse_dense_blocks(gates, dense_blocks, dense_min_gates > 0 ? dense_min_gates : 24);
        FuseConfig cfg;
        cfg.tile_bits = tile_bits > 0 ? tile_bits : default_tile_bits(tmp.dtype, 1);
        cfg.low_bits = low_bits > 0 ? low_bits : default_low_bits(tmp.dtype, 1);
        cfg.max_groups = max_groups > 0 ? max_groups : 12;
        cfg.max_cost = pass_cost > 0 ? pass_cost : 0;
        // the group width qfvm_apply would use with default options (complex128: rows of 8)
        const int bb = std::min(cfg.tile_bits, n);
        if (tmp.dtype == QFVM_C128 && bb >= 10 && bb <= 13) cfg.group_bits = 3;
        // complex64 gate lists with dense 2-target gates take the row-of-8 kernel (as qfvm_apply)
        if (tmp.dtype == QFVM_C64 && bb >= 10 && bb <= 13)
            for (const Gate& g : gates)
                if (g.kind == G_MULTI && g.targets.size() == 2) cfg.group_bits = 3;
        cfg.layout = layout != 0 && cfg.group_bits == 3;
        std::vector<PlanItem> plan = plan_gates(gates, n, cfg);
        std::string j = "{\"items\": [";
        char t[512];
        auto cplx = [&](const cd& z) {
            std::snprintf(t, sizeof(t), "[%.17g, %.17g]", z.real(), z.imag());
            return std::string(t);
        };
        for (size_t i = 0; i < plan.size(); i++) {
            const PlanItem& it = plan[i];
            j += i ? ", " : "";
            if (it.fused) {
                j += "{\"fused\": true, \"group_bits\": " + std::to_string(it.pass.group_bits) + ", \"Q\": [";
                for (size_t k = 0; k < it.pass.Q.size(); k++) j += (k ? ", " : "") + std::to_string(it.pass.Q[k]);
                j += "], \"out_bit\": [";
                for (size_t k = 0; k < it.pass.out_bit.size(); k++) j += (k ? ", " : "") + std::to_string(it.pass.out_bit[k]);
                j += "], \"groups\": [";
                for (size_t g = 0; g < it.pass.groups.size(); g++) {
                    const FGroup& G = it.pass.groups[g];
                    j += g ? ", " : "";
                    std::snprintf(t, sizeof(t), "{\"S\": [%d, %d, %d, %d], \"ops\": [", G.S[0], G.S[1], G.S[2], G.S[3]);
                    j += t;
                    for (size_t o = 0; o < G.ops.size(); o++) {
                        const FOpHost& op = G.ops[o];
                        std::snprintf(t, sizeof(t),
                                      "%s{\"kind\": %d, \"g\": %d, \"g2\": %d, \"cmg\": %u, \"cme\": %llu, \"run\": %d, "
                                      "\"flags\": %u, \"m\": [",
                                      o ? ", " : "", op.kind, op.g, op.kind == F_DENSE4 ? op.target : -1, op.cm_group,
                                      (unsigned long long)op.cm_ext, op.run, op.flags);
                        j += t;
                        if (op.kind == F_DENSE4)  // the 4x4, row-major
                            for (size_t k = 0; k < op.mx.size(); k++) j += (k ? ", " : "") + cplx(op.mx[k]);
                        else
                            for (int k = 0; k < 4; k++) j += (k ? ", " : "") + cplx(op.m[k]);
                        j += "]}";
                    }
                    j += "]}";
                }
                j += "]}";
            } else {
                const Gate& g = it.gate;
                std::snprintf(t, sizeof(t), "{\"fused\": false, \"kind\": %d, \"target\": %d, \"ctrls\": [", g.kind, g.target);
                j += t;
                for (size_t k = 0; k < g.ctrls.size(); k++) j += (k ? ", " : "") + std::to_string(g.ctrls[k]);
                j += "], \"targets\": [";
                for (size_t k = 0; k < g.targets.size(); k++) j += (k ? ", " : "") + std::to_string(g.targets[k]);
                j += "], \"m\": [";
                if (g.kind == G_MULTI)
                    for (size_t k = 0; k < g.mk.size(); k++) j += (k ? ", " : "") + cplx(g.mk[k]);
                else
                    for (int k = 0; k < 4; k++) j += (k ? ", " : "") + cplx(g.m[k]);
                j += "]}";
            }
        }
        j += "]}";
        if (len) *len = (int64_t)j.size();
        if (buf && cap > 0) {
            const size_t m = std::min<size_t>((size_t)cap - 1, j.size());
            std::memcpy(buf, j.data(), m);
            buf[m] = 0;
        }
    });
}

int qfvm_set_profiling(qfvm_handle h, int32_t enable) {
    return guarded([&] {
        State* s = need(h);
        QFVM_HIP(hipStreamSynchronize(s->stream));
        s->resolve_events();
        s->profiling = enable != 0;
    });
}

int qfvm_profile_json(qfvm_handle h, char* buf, int64_t cap, int64_t* len, int32_t reset) {
    return guarded([&] {
        State* s = need(h);
        QFVM_HIP(hipStreamSynchronize(s->stream));
        s->resolve_events();
        std::string j = "{";
        char tmp[256];
        for (size_t k = 0; k < s->kstats.size(); k++) {
            const KStat& st = s->kstats[k].second;
            std::snprintf(tmp, sizeof(tmp), "%s\"%s\": {\"launches\": %llu, \"ms\": %.6f, \"bytes\": %.0f}",
                          k ? ", " : "", s->kstats[k].first.c_str(), (unsigned long long)st.launches, st.ms, st.bytes);
            j += tmp;
        }
        j += "}";
        if (len) *len = (int64_t)j.size();
        if (buf && cap > 0) {
            const size_t m = std::min<size_t>((size_t)cap - 1, j.size());
            std::memcpy(buf, j.data(), m);
            buf[m] = 0;
        }
        if (reset) s->kstats.clear();
    });
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// Batched evaluation: worker threads, each with its own handles (one per qubit count), run items
// through the same entry points a caller would use one by one.
struct qfvm_batch_ctx {
    int dtype = QFVM_C128;
    int workers = 1;
    int device = 0;
    std::vector<std::vector<std::pair<int, qfvm_handle>>> handles;  // per worker: (n, handle)
    struct Opt {
        int n;
        std::string key;
        int64_t value;
    };
    std::vector<Opt> options;  // in the order they were set
    int pin_qubits = 20;       // items of at least this many qubits all go to worker 0 (0: none)
};

namespace {

struct BatchFailure {
    int code = QFVM_OK;
    std::string msg;
};

// the worker's handle for n qubits (created with the batch's options on first use)
qfvm_handle batch_handle(qfvm_batch_ctx& b, int w, int n, BatchFailure& f) {
    for (const auto& kv : b.handles[(size_t)w])
        if (kv.first == n) return kv.second;
    qfvm_handle h = nullptr;
    int rc = qfvm_create(n, b.dtype, &h);
    for (const auto& o : b.options) {
        if (rc != QFVM_OK) break;
        if (o.n == 0 || o.n == n) rc = qfvm_set_option(h, o.key.c_str(), o.value);
    }
    if (rc != QFVM_OK) {
        f.code = rc;
        f.msg = qfvm_last_error();
        if (h) qfvm_destroy(h);
        return nullptr;
    }
    b.handles[(size_t)w].emplace_back(n, h);
    return h;
}

}  // namespace

extern "C" {

int qfvm_batch_create(int32_t dtype, int32_t workers, qfvm_batch* out) {
    return guarded([&] {
        if (!out) throw Error(QFVM_ERR_VALUE, "null output batch");
        *out = nullptr;
        if (dtype != QFVM_C64 && dtype != QFVM_C128) throw Error(QFVM_ERR_VALUE, "dtype must be QFVM_C64 or QFVM_C128");
        if (workers < 1 || workers > 64) throw Error(QFVM_ERR_VALUE, "workers must be in [1, 64]");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw Error(QFVM_ERR_DEVICE, "no HIP device is visible to this process");
        auto b = std::make_unique<qfvm_batch_ctx>();
        QFVM_HIP(hipGetDevice(&b->device));
        b->dtype = dtype;
        b->workers = workers;
        b->handles.resize((size_t)workers);
        *out = b.release();
    });
}

int qfvm_batch_destroy(qfvm_batch b) {
    return guarded([&] {
        if (!b) return;
        for (auto& per : b->handles)
            for (auto& kv : per) qfvm_destroy(kv.second);
        delete b;
    });
}

int qfvm_batch_set_option(qfvm_batch b, int32_t n_qubits, const char* key, int64_t value) {
    return guarded([&] {
        if (!b) throw Error(QFVM_ERR_VALUE, "null batch");
        if (!key) throw Error(QFVM_ERR_VALUE, "null option key");
        if (n_qubits < 0 || n_qubits > 34) throw Error(QFVM_ERR_VALUE, "n_qubits must be in [0, 34]");
        if (std::string(key) == "pin_qubits") {  // a scheduling option of the batch itself
            if (value < 0 || value > 34) throw Error(QFVM_ERR_VALUE, "pin_qubits must be in [0, 34]");
            b->pin_qubits = (int)value;
            return;
        }
        for (auto& per : b->handles)
            for (auto& kv : per)
                if (n_qubits == 0 || kv.first == n_qubits) {
                    const int rc = qfvm_set_option(kv.second, key, value);
                    if (rc != QFVM_OK) throw Error(rc, qfvm_last_error());
                }
        b->options.push_back({n_qubits, key, value});
    });
}

int qfvm_batch_run(qfvm_batch b, const qfvm_batch_item* items, int64_t n_items) {
    return guarded([&] {
        if (!b) throw Error(QFVM_ERR_VALUE, "null batch");
        if (n_items < 0 || (n_items > 0 && !items)) throw Error(QFVM_ERR_VALUE, "bad item array");
        for (int64_t k = 0; k < n_items; k++) {
            const qfvm_batch_item& it = items[k];
            if (!it.prog || !it.expectation || (it.n_grads > 0 && !it.grads))
                throw Error(QFVM_ERR_VALUE, "batch item without program or outputs");
            if (it.n_qubits < 0 || it.n_qubits > 34) throw Error(QFVM_ERR_VALUE, "n_qubits must be in [0, 34]");
        }
        // Circuits large enough to fill the GPU alone gain nothing from running beside each other
        // and would only stretch each other's latency: they all go to worker 0, largest first.
        // The rest go largest first (cost ~ 2^n x ops) onto the least-loaded worker.
        std::vector<int64_t> order((size_t)n_items);
        std::vector<double> cost((size_t)n_items);
        for (int64_t k = 0; k < n_items; k++) {
            order[(size_t)k] = k;
            cost[(size_t)k] = std::ldexp((double)(items[k].prog->n_ops + 8), items[k].n_qubits);
        }
        std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t c) { return cost[(size_t)a] > cost[(size_t)c]; });
        std::vector<std::vector<int64_t>> part((size_t)b->workers);
        std::vector<double> load((size_t)b->workers, 0.0);
        const bool pin = b->pin_qubits > 0 && b->workers > 1;
        for (int64_t k : order) {
            size_t w = 0;
            if (!(pin && items[k].n_qubits >= b->pin_qubits))
                for (size_t j = 1; j < load.size(); j++)
                    if (load[j] < load[w]) w = j;
            part[w].push_back(k);
            load[w] += cost[(size_t)k];
        }
        std::vector<BatchFailure> fail((size_t)b->workers);
        auto work = [&](int w) {
            BatchFailure& f = fail[(size_t)w];
            if (hipSetDevice(b->device) != hipSuccess) {
                f.code = QFVM_ERR_DEVICE;
                f.msg = "hipSetDevice failed in a batch worker";
                return;
            }
            for (int64_t k : part[(size_t)w]) {
                const qfvm_batch_item& it = items[k];
                const auto t0 = std::chrono::steady_clock::now();
                qfvm_handle h = batch_handle(*b, w, it.n_qubits, f);
                if (!h) return;
                int rc = qfvm_init_zero(h);
                if (rc == QFVM_OK)
                    rc = qfvm_adjoint(h, it.prog, it.grad_ops, it.grad_gen, it.n_grads, it.chars, it.pos, it.off, it.len,
                                      it.coef, it.n_terms, it.grads, it.expectation);
                if (rc != QFVM_OK) {
                    f.code = rc;
                    f.msg = qfvm_last_error();
                    return;
                }
                if (it.seconds)
                    *it.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            }
        };
        std::vector<std::thread> threads;
        for (int w = 1; w < b->workers; w++)
            if (!part[(size_t)w].empty()) threads.emplace_back(work, w);
        work(0);
        for (std::thread& t : threads) t.join();
        QFVM_HIP(hipSetDevice(b->device));
        for (const BatchFailure& f : fail)
            if (f.code != QFVM_OK) throw Error(f.code, f.msg);
    });
}

}  // extern "C"
