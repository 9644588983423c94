// Gate-run execution: the fusion planner and kernel accounting.
//
// The planner packs a gate list into fused passes (fused.h). It walks the gates in order and takes
// every gate it can move into the current pass without changing the result:
//   * a gate's non-diagonal qubits N (the target of a dense / real / X gate) must fit in the tile,
//     which grows from the L low qubits up to b qubits;
//   * a gate that cannot be taken is skipped, and every later gate that does not commute with a
//     skipped one is skipped too. Commutation is checked on qubit roles: two gates commute when every
//     qubit they share is only a control or a diagonal target for both;
//   * inside the pass the gate joins the latest group whose 4 qubits can hold N (a group may grow
//     until it has 4 qubits) and may move back past later groups it commutes with.
// Diagonal gates (RZ, CZ, CP, controlled phases, diagonal 2x2 matrices) have N empty, so they never
// widen a tile. Dense gates on two or more targets run on the per-gate kernels.
#include <algorithm>

#include "fused.h"
#include "plan.h"

namespace qfvm {

namespace {

inline int popc(uint64_t x) { return __builtin_popcountll(x); }

struct GInfo {
    uint64_t N = 0, D = 0;
    bool diag = false, multi = false, grad = false;
    bool dense2 = false;  // a dense 2-target gate fused into row-of-8 groups (see plan_gates)
    cd d0, d1;
};

GInfo info_of(const Gate& g) {
    GInfo r;
    r.D = g.cmask;
    if (g.kind == G_GRAD) {  // the generator's roles: Z is diagonal, X and Y need the target in a tile
        r.grad = true;
        if (g.gen == GEN_Z)
            r.D |= 1ull << g.target;
        else
            r.N = 1ull << g.target;
        return r;
    }
    if (g.kind == G_MULTI) {
        r.multi = true;
        for (int t : g.targets) r.N |= 1ull << t;
        return r;
    }
    const bool dense_diag = g.kind == G_DENSE && g.m[1] == cd(0.0, 0.0) && g.m[2] == cd(0.0, 0.0);
    if (g.kind == G_DIAG || dense_diag) {
        r.diag = true;
        r.D |= 1ull << g.target;
        r.d0 = g.m[0];
        r.d1 = g.kind == G_DIAG ? g.m[1] : g.m[3];
    } else {
        r.N = 1ull << g.target;
    }
    return r;
}

struct GroupB {
    uint64_t S = 0, N = 0, A = 0;
    bool stage = false;
    std::vector<int> gates;
};

// register work of one gate in a fused pass, in quarter dense-gate units (measured per-op costs:
// dense 2x2 = 4; real and diagonal about 3; an X is mostly a free relabeling)
int gate_cost(const Gate& g, const GInfo& gi) {
    if (gi.grad) return 2;
    if (gi.dense2) return 16;  // a 4x4 on a row: four dense 2x2 gates' arithmetic
    if (gi.diag) return 3;
    switch (g.kind) {
        case G_SWAP: return 1;
        case G_REAL: return 3;
        default: return 4;
    }
}

// a qubit mask / a gate moved to the positions l2p[q] of a layout (l2p empty: identity)
uint64_t move_mask(uint64_t m, const std::vector<int>& l2p) {
    if (l2p.empty()) return m;
    uint64_t r = 0;
    while (m) {
        const int q = __builtin_ctzll(m);
        m &= m - 1;
        r |= 1ull << l2p[(size_t)q];
    }
    return r;
}

Gate move_gate(const Gate& g, const std::vector<int>& l2p) {
    Gate t = g;
    if (l2p.empty()) return t;
    if (g.target >= 0 && g.target < (int)l2p.size()) t.target = l2p[(size_t)g.target];
    t.cmask = move_mask(g.cmask, l2p);
    for (int& c : t.ctrls) c = l2p[(size_t)c];
    for (int& q : t.targets) q = l2p[(size_t)q];
    return t;
}

// two gates (or a gate and a group) commute when no qubit is non-diagonal for one and touched by
// the other
inline bool commutes(const GInfo& g, const GroupB& G) { return !((g.N & G.A) || ((g.N | g.D) & G.N)); }

FOpHost convert(const Gate& g, const GInfo& gi, const int S[4]) {
    FOpHost o{};
    auto idx_in = [&](int q) {
        for (int j = 0; j < FUSE_GROUP_BITS; j++)
            if (S[j] == q) return j;
        return -1;
    };
    for (int c : g.ctrls) {
        const int j = idx_in(c);
        if (j >= 0)
            o.cm_group |= 1u << j;
        else
            o.cm_ext |= 1ull << c;
    }
    if (gi.dense2) {
        o.kind = F_DENSE4;
        o.g = idx_in(g.targets[0]);
        o.target = idx_in(g.targets[1]);
        o.mx = g.mk;
        return o;
    }
    if (gi.diag) {
        const int j = idx_in(g.target);
        o.kind = j >= 0 ? F_DIAG : F_PHASE;
        o.g = j >= 0 ? j : g.target;
        o.m[0] = gi.d0;
        o.m[1] = 0.0;
        o.m[2] = 0.0;
        o.m[3] = gi.d1;
        return o;
    }
    o.g = idx_in(g.target);
    o.kind = g.kind == G_SWAP ? F_SWAP : g.kind == G_REAL ? F_REAL : F_DENSE;
    for (int k = 0; k < 4; k++) o.m[k] = g.m[k];
    return o;
}

// Groups consecutive diagonal ops (relabelings may sit between them) into runs of at least
// DIAGRUN_MIN ops (fused.h). An op joins only if its terms can be formed: d0 != 0, and d1 != 0 when
// its target bit carries a pending relabeling.
void mark_diagonal_runs(FGroup& fg) {
    int next = 1;
    std::vector<int> cur;
    auto close = [&]() {
        if ((int)cur.size() >= DIAGRUN_MIN) {
            for (int k : cur) fg.ops[k].run = next;
            next++;
        }
        cur.clear();
    };
    for (int k = 0; k < (int)fg.ops.size(); k++) {
        const FOpHost& o = fg.ops[k];
        if (o.kind == F_RELABEL) continue;
        const bool diag = o.kind == F_DIAG || o.kind == F_PHASE;
        const bool ok = diag && o.m[0] != cd(0.0, 0.0) &&
                        !((o.flags & FL_TARGET_PX) && o.m[3] == cd(0.0, 0.0));
        if (ok)
            cur.push_back(k);
        else
            close();
    }
    close();
}

}  // namespace

namespace {

uint64_t support_of(const Gate& g) {
    uint64_t m = g.cmask;
    if (g.kind == G_MULTI)
        for (int t : g.targets) m |= 1ull << t;
    else
        m |= 1ull << g.target;
    return m;
}

// M <- g M for the 2^m x 2^m row-major M acting on qubits S (ascending; local bit j <-> S[j]): g is
// applied to every column of M as to an m-qubit state.
void apply_to_block(const Gate& g, const std::vector<int>& S, std::vector<cd>& M) {
    const size_t d = (size_t)1 << S.size();
    auto local = [&](int q) { return (size_t)1 << (std::find(S.begin(), S.end(), q) - S.begin()); };
    size_t cm = 0;
    for (int q = 0; q < 64; q++)
        if ((g.cmask >> q) & 1) cm |= local(q);
    if (g.kind == G_MULTI) {
        const size_t k = g.targets.size(), dk = (size_t)1 << k;
        std::vector<size_t> tb(k);
        size_t tmask = 0;
        for (size_t j = 0; j < k; j++) tmask |= tb[j] = local(g.targets[j]);
        std::vector<size_t> at(dk);
        std::vector<cd> in(dk);
        for (size_t base = 0; base < d; base++) {
            if ((base & tmask) || (base & cm) != cm) continue;
            for (size_t r = 0; r < dk; r++) {
                size_t i = base;
                for (size_t j = 0; j < k; j++)
                    if ((r >> j) & 1) i |= tb[j];
                at[r] = i;
            }
            for (size_t col = 0; col < d; col++) {
                for (size_t r = 0; r < dk; r++) in[r] = M[at[r] * d + col];
                for (size_t r = 0; r < dk; r++) {
                    cd acc = 0.0;
                    for (size_t c = 0; c < dk; c++) acc += g.mk[r * dk + c] * in[c];
                    M[at[r] * d + col] = acc;
                }
            }
        }
        return;
    }
    const size_t t = local(g.target);
    for (size_t i0 = 0; i0 < d; i0++) {
        if ((i0 & t) || (i0 & cm) != cm) continue;
        cd* r0 = &M[i0 * d];
        cd* r1 = &M[(i0 | t) * d];
        for (size_t col = 0; col < d; col++) {
            const cd a = r0[col], b = r1[col];
            switch (g.kind) {
                case G_DIAG:
                    r0[col] = g.m[0] * a;
                    r1[col] = g.m[1] * b;
                    break;
                case G_SWAP:
                    r0[col] = b;
                    r1[col] = a;
                    break;
                default:  // G_DENSE, G_REAL
                    r0[col] = g.m[0] * a + g.m[1] * b;
                    r1[col] = g.m[2] * a + g.m[3] * b;
                    break;
            }
        }
    }
}

}  // namespace

std::vector<Gate> merge_single_qubit_runs(const std::vector<Gate>& gates) {
    // A run is a sequence of uncontrolled one-target gates on one qubit with no other gate on that
    // qubit between them; gates between the run's members act on other qubits, so the run's product
    // can stand at the place of its last gate.
    auto single = [](const Gate& g) {
        return g.cmask == 0 && (g.kind == G_DENSE || g.kind == G_REAL || g.kind == G_DIAG || g.kind == G_SWAP);
    };
    std::vector<std::vector<int>> run(64);
    std::vector<int> owner(gates.size(), -1);
    std::vector<Gate> merged;
    std::vector<int> last;
    auto close = [&](int q) {
        std::vector<int>& r = run[(size_t)q];
        if (r.size() >= 2) {
            cd M[4] = {1.0, 0.0, 0.0, 1.0};
            bool all_diag = true, all_real = true;
            for (int i : r) {
                const Gate& g = gates[(size_t)i];
                cd G[4];
                if (g.kind == G_DIAG) {
                    G[0] = g.m[0], G[1] = 0.0, G[2] = 0.0, G[3] = g.m[1];
                } else if (g.kind == G_SWAP) {
                    G[0] = 0.0, G[1] = 1.0, G[2] = 1.0, G[3] = 0.0;
                } else {
                    for (int j = 0; j < 4; j++) G[j] = g.m[j];
                }
                all_diag = all_diag && g.kind == G_DIAG;
                all_real = all_real && (g.kind == G_REAL || g.kind == G_SWAP);
                const cd P[4] = {G[0] * M[0] + G[1] * M[2], G[0] * M[1] + G[1] * M[3], G[2] * M[0] + G[3] * M[2],
                                 G[2] * M[1] + G[3] * M[3]};
                for (int j = 0; j < 4; j++) M[j] = P[j];
            }
            Gate m;
            m.target = q;
            if (all_diag) {
                m.kind = G_DIAG;
                m.m[0] = M[0];
                m.m[1] = M[3];
                m.m[2] = m.m[3] = 0.0;
            } else {
                m.kind = all_real ? G_REAL : G_DENSE;
                for (int j = 0; j < 4; j++) m.m[j] = M[j];
            }
            for (int i : r) owner[(size_t)i] = (int)merged.size();
            last.push_back(r.back());
            merged.push_back(std::move(m));
        }
        r.clear();
    };
    for (int i = 0; i < (int)gates.size(); i++) {
        const Gate& g = gates[(size_t)i];
        if (single(g)) {
            run[(size_t)g.target].push_back(i);
            continue;
        }
        const uint64_t sup = support_of(g);
        for (int q = 0; q < 64; q++)
            if ((sup >> q) & 1) close(q);
    }
    for (int q = 0; q < 64; q++) close(q);
    if (merged.empty()) return gates;
    std::vector<Gate> out;
    out.reserve(gates.size());
    for (int i = 0; i < (int)gates.size(); i++) {
        const int o = owner[(size_t)i];
        if (o < 0)
            out.push_back(gates[(size_t)i]);
        else if (last[(size_t)o] == i)
            out.push_back(merged[(size_t)o]);
    }
    return out;
}

std::vector<Gate> collapse_dense_blocks(const std::vector<Gate>& gates, int k, int min_gates) {
    // Open blocks have pairwise disjoint qubit sets; a gate joins (and merges) the blocks it shares
    // qubits with when their union stays within k qubits, else those blocks are closed and the gate
    // opens a block of its own. A gate h outside a block B but between B's gates shares no qubit
    // with B's earlier gates (it would have joined B or closed it), so a collapsed block's product
    // can stand at the place of its last gate; every other gate keeps its place.
    struct Block {
        uint64_t sup;
        std::vector<int> idx;
    };
    std::vector<Block> open;
    std::vector<int> owner(gates.size(), -1);  // collapsed block of each gate
    std::vector<Gate> dense;                   // the collapsed blocks' gates
    std::vector<int> last;                     // index of each collapsed block's last gate
    auto close = [&](Block& b) {
        int nd = 0;
        for (int i : b.idx) nd += gates[i].kind != G_DIAG;
        if (__builtin_popcountll(b.sup) < 4 || nd < min_gates) return;
        std::sort(b.idx.begin(), b.idx.end());
        Gate d;
        d.kind = G_MULTI;
        for (int q = 0; q < 64; q++)
            if ((b.sup >> q) & 1) d.targets.push_back(q);
        const size_t dim = (size_t)1 << d.targets.size();
        d.mk.assign(dim * dim, cd(0.0));
        for (size_t i = 0; i < dim; i++) d.mk[i * dim + i] = 1.0;
        for (int i : b.idx) {
            apply_to_block(gates[i], d.targets, d.mk);
            owner[i] = (int)dense.size();
        }
        last.push_back(b.idx.back());
        dense.push_back(std::move(d));
    };
    for (int i = 0; i < (int)gates.size(); i++) {
        const Gate& g = gates[i];
        if (g.kind == G_GRAD) throw Error(QFVM_ERR_RUNTIME, "dense blocks take unitary gates only");
        const uint64_t gs = support_of(g);
        uint64_t u = gs;
        std::vector<size_t> touched;
        for (size_t b = 0; b < open.size(); b++)
            if (open[b].sup & gs) {
                touched.push_back(b);
                u |= open[b].sup;
            }
        Block nb{gs, {}};
        if (__builtin_popcountll(u) <= k) {
            nb.sup = u;
            for (size_t t : touched) nb.idx.insert(nb.idx.end(), open[t].idx.begin(), open[t].idx.end());
        } else {
            for (size_t t : touched) close(open[t]);
        }
        for (size_t t = touched.size(); t-- > 0;) open.erase(open.begin() + (std::ptrdiff_t)touched[t]);
        if (__builtin_popcountll(nb.sup) <= k) {
            nb.idx.push_back(i);
            open.push_back(std::move(nb));
        }
    }
    for (Block& b : open) close(b);
    if (dense.empty()) return gates;
    std::vector<Gate> out;
    out.reserve(gates.size());
    for (int i = 0; i < (int)gates.size(); i++) {
        if (owner[i] < 0)
            out.push_back(gates[i]);
        else if (last[owner[i]] == i)
            out.push_back(dense[owner[i]]);
    }
    return out;
}

std::vector<PlanItem> plan_gates(const std::vector<Gate>& gates, int n, const FuseConfig& cfg) {
    const int b = std::min(cfg.tile_bits, n);
    const int gb = cfg.group_bits;
    // a tile keeps at least one free slot when it does not hold every qubit, so the first waiting
    // gate always fits and every pass makes progress
    const int L = b < n ? std::min(cfg.low_bits, b - 1) : b;
    const uint64_t all = n >= 64 ? ~0ull : ((1ull << n) - 1);
    // Layout-aware plans (cfg.layout, row-of-8 passes): qubit q sits at position phys[q] (inv is the
    // inverse); gates are planned on positions. At the end of every pass the tile qubits with the
    // nearest next non-diagonal use move into the L low positions, which every later tile holds, so
    // the next pass needs fewer new positions; passes without gates restore the identity at the end.
    const bool lay = cfg.layout && b < n && gb == 3;
    std::vector<int> phys, inv;
    if (lay) {
        phys.resize((size_t)n);
        inv.resize((size_t)n);
        for (int q = 0; q < n; q++) phys[(size_t)q] = inv[(size_t)q] = q;
    }
    bool moved = false;  // phys differs from the identity
    static const std::vector<int> kIdentity;
    std::vector<GInfo> info(gates.size());
    for (size_t i = 0; i < gates.size(); i++) {
        info[i] = info_of(gates[i]);
        // row-of-8 groups hold a dense 2-target gate as one 4x4 op (both targets in one group)
        if (gb == 3 && info[i].multi && gates[i].targets.size() == 2) {
            info[i].multi = false;
            info[i].dense2 = true;
        }
    }
    std::vector<char> done(gates.size(), 0);
    std::vector<PlanItem> plan;
    size_t first = 0;
    while (true) {
        while (first < gates.size() && done[first]) first++;
        if (first >= gates.size()) break;
        const std::vector<int>& cur = moved ? phys : kIdentity;
        if (info[first].multi) {
            plan.push_back(PlanItem{false, FPass{}, move_gate(gates[first], cur), (int)first, cur});
            done[first] = 1;
            continue;
        }
        uint64_t Q = L >= 64 ? ~0ull : ((1ull << L) - 1);
        uint64_t blockN = 0, blockA = 0;
        std::vector<GroupB> groups;
        int cost = 0;
        for (size_t i = first; i < gates.size(); i++) {
            if (done[i]) continue;
            GInfo g = info[i];
            if (moved) {
                g.N = move_mask(g.N, phys);
                g.D = move_mask(g.D, phys);
            }
            const uint64_t A = g.N | g.D;
            bool take = !((g.N & blockA) || (g.D & blockN)) && !g.multi;
            uint64_t newQ = Q | g.N;
            if (take && popc(newQ) > b) take = false;
            int place = -1;
            if (take) {
                // latest group that can hold the gate (a stage holds markers, a compute group gates),
                // moving back only past groups it commutes with
                for (int j = (int)groups.size() - 1; j >= 0; j--) {
                    const GroupB& G = groups[j];
                    const bool fits = g.grad ? G.stage
                                             : (!G.stage && ((G.S & g.N) == g.N || popc(G.S | g.N) <= gb));
                    if (fits) {
                        place = j;
                        break;
                    }
                    if (!commutes(g, G)) break;  // cannot move before group j
                }
                if (place < 0) {
                    if ((int)groups.size() < cfg.max_groups) {
                        groups.emplace_back();
                        groups.back().stage = g.grad;
                        place = (int)groups.size() - 1;
                    } else {
                        take = false;
                    }
                }
            }
            if (take && cfg.max_cost > 0) {
                // estimated register work of the gate (a new group adds its gather / scatter)
                const int c = gate_cost(gates[i], g) + (place == (int)groups.size() - 1 && groups.back().gates.empty() ? 4 : 0);
                if (cost + c > cfg.max_cost && cost > 0) {
                    if (place == (int)groups.size() - 1 && groups.back().gates.empty()) groups.pop_back();
                    break;  // the pass is full
                }
                cost += c;
            }
            if (!take) {
                blockN |= g.N;
                blockA |= A;
                if ((blockN & all) == all) break;  // every later gate touches a blocked qubit
                continue;
            }
            GroupB& G = groups[place];
            if (!G.stage) G.S |= g.N;
            G.N |= g.N;
            G.A |= A;
            G.gates.push_back((int)i);
            Q = newQ;
            done[i] = 1;
        }
        // finalize: pad the tile to b qubits and every group to gb tile qubits
        for (int q = 0; q < n && popc(Q) < b; q++) Q |= 1ull << q;
        PlanItem it{true, FPass{}, Gate{}};
        it.pass.group_bits = gb;
        for (int q = 0; q < n; q++)
            if ((Q >> q) & 1) it.pass.Q.push_back(q);
        for (GroupB& G : groups) {
            if (G.stage) {
                FGroup fg;
                fg.stage = true;
                for (int j = 0; j < FUSE_GROUP_BITS; j++) fg.S[j] = -1;
                for (int gi : G.gates) {
                    const Gate tg = move_gate(gates[gi], cur);
                    FOpHost o{};
                    o.src = -1;
                    o.kind = F_GRAD;
                    o.g = tg.gen;
                    o.cm_ext = tg.cmask;
                    o.target = tg.target;
                    o.slot = tg.slot;
                    fg.ops.push_back(o);
                }
                it.pass.groups.push_back(std::move(fg));
                continue;
            }
            for (int q : it.pass.Q) {
                if (popc(G.S) >= gb) break;
                G.S |= 1ull << q;
            }
            FGroup fg;
            int j = 0;
            for (int q = 0; q < n; q++)
                if ((G.S >> q) & 1) fg.S[j++] = q;
            // X gates without group controls and with at most one outside control become parity
            // relabelings (fused.h): M[j] / C track the pending relabeling of group bit j
            uint64_t M[FUSE_GROUP_BITS] = {0, 0, 0, 0};
            uint32_t C = 0;
            auto pending = [&](int j) { return M[j] != 0 || ((C >> j) & 1); };
            for (int gi : G.gates) {
                FOpHost o = convert(move_gate(gates[gi], cur), info[gi], fg.S);
                o.src = gi;
                // a 4x4 reads its two target bits as they are in memory: their pending relabelings
                // are applied first, as for controls
                const uint32_t flush_bits = o.cm_group | (o.kind == F_DENSE4 ? (1u << o.g) | (1u << o.target) : 0u);
                for (int j = 0; j < FUSE_GROUP_BITS; j++) {
                    if (((flush_bits >> j) & 1) && pending(j)) {
                        FOpHost f{};
                        f.src = -1;
                        f.kind = F_FLUSH;
                        f.g = j;
                        f.pxmask = M[j];
                        f.flags = ((C >> j) & 1) ? FL_PX_CONST : 0;
                        fg.ops.push_back(f);
                        M[j] = 0;
                        C &= ~(1u << j);
                    }
                }
                if (o.kind == F_SWAP && o.cm_group == 0) {
                    if (popc(o.cm_ext) <= 1) {
                        M[o.g] ^= o.cm_ext;
                        if (o.cm_ext == 0) C ^= 1u << o.g;
                        o.kind = F_RELABEL;  // listed for inspection, never uploaded
                        fg.ops.push_back(o);
                        continue;
                    }
                    // several outside controls: a physical X where they all hold (X commutes with
                    // a pending relabeling of the same bit)
                    o.kind = F_DENSE;
                }
                if (o.kind != F_SWAP && o.kind != F_PHASE && o.kind != F_DENSE4 && pending(o.g)) {
                    o.flags |= FL_TARGET_PX | (((C >> o.g) & 1) ? FL_PX_CONST : 0);
                    o.pxmask = M[o.g];
                }
                fg.ops.push_back(o);
            }
            for (int j = 0; j < FUSE_GROUP_BITS; j++) fg.pxm[j] = M[j];
            fg.pxc = C;
            mark_diagonal_runs(fg);
            it.pass.groups.push_back(std::move(fg));
        }
        if (lay) {
            // the L tile qubits whose next non-diagonal use comes first go to the low positions
            const std::vector<int>& Qv = it.pass.Q;
            const int bb = (int)Qv.size();
            std::vector<size_t> nxt((size_t)bb, gates.size());
            int found = 0;
            for (size_t i = first; i < gates.size() && found < bb; i++) {
                if (done[i]) continue;
                for (int k = 0; k < bb; k++)
                    if (nxt[(size_t)k] == gates.size() && ((info[i].N >> inv[(size_t)Qv[(size_t)k]]) & 1)) {
                        nxt[(size_t)k] = i;
                        found++;
                    }
            }
            std::vector<int> order((size_t)bb);
            for (int k = 0; k < bb; k++) order[(size_t)k] = k;
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return nxt[(size_t)x] < nxt[(size_t)y]; });
            std::vector<char> chosen((size_t)bb, 0);
            for (int k = 0; k < L; k++) chosen[(size_t)order[(size_t)k]] = 1;
            std::vector<int> out((size_t)bb);
            for (int k = 0; k < bb; k++) out[(size_t)k] = k;
            int lo = 0;
            bool any = false;
            for (int k = L; k < bb; k++) {
                if (!chosen[(size_t)k]) continue;
                while (chosen[(size_t)lo]) lo++;  // a low slot whose qubit is not chosen (one exists per mover)
                out[(size_t)k] = lo;
                out[(size_t)lo] = k;
                lo++;
                any = true;
            }
            if (any) {
                std::vector<int> was((size_t)bb);
                for (int k = 0; k < bb; k++) was[(size_t)k] = inv[(size_t)Qv[(size_t)k]];
                for (int k = 0; k < bb; k++) {
                    const int q = was[(size_t)k], p = Qv[(size_t)out[(size_t)k]];
                    phys[(size_t)q] = p;
                    inv[(size_t)p] = q;
                }
                it.pass.out_bit = out;
                moved = true;
            }
        }
        plan.push_back(std::move(it));
    }
    // restore the identity layout: passes without gates, each holding the low positions and up to
    // b - L high positions along the permutation's cycles; a tile qubit whose own position is in the
    // tile goes there (at least one more qubit reaches its position per pass)
    for (int guard = 0; moved && guard <= n; guard++) {
        bool ident = true;
        for (int q = 0; q < n; q++) ident = ident && phys[(size_t)q] == q;
        if (ident) break;
        uint64_t Qm = L >= 64 ? ~0ull : ((1ull << L) - 1);
        for (int p = L; p < n; p++) {
            if (inv[(size_t)p] == p) continue;
            const uint64_t add = (1ull << p) | (phys[(size_t)p] >= L ? 1ull << phys[(size_t)p] : 0ull);
            if (popc(Qm | add) <= b) Qm |= add;
        }
        for (int q = 0; q < n && popc(Qm) < b; q++) Qm |= 1ull << q;
        PlanItem it{true, FPass{}, Gate{}};
        it.pass.group_bits = gb;
        std::vector<int> loc((size_t)n, -1);
        for (int q = 0; q < n; q++)
            if ((Qm >> q) & 1) {
                loc[(size_t)q] = (int)it.pass.Q.size();
                it.pass.Q.push_back(q);
            }
        const int bb = (int)it.pass.Q.size();
        std::vector<int> out((size_t)bb, -1), was((size_t)bb);
        std::vector<char> used((size_t)bb, 0);
        for (int k = 0; k < bb; k++) {
            was[(size_t)k] = inv[(size_t)it.pass.Q[(size_t)k]];
            const int home = loc[(size_t)was[(size_t)k]];
            if (home >= 0) {
                out[(size_t)k] = home;
                used[(size_t)home] = 1;
            }
        }
        int nx = 0;
        for (int k = 0; k < bb; k++) {
            if (out[(size_t)k] >= 0) continue;
            while (used[(size_t)nx]) nx++;
            out[(size_t)k] = nx;
            used[(size_t)nx] = 1;
        }
        for (int k = 0; k < bb; k++) {
            const int q = was[(size_t)k], p = it.pass.Q[(size_t)out[(size_t)k]];
            phys[(size_t)q] = p;
            inv[(size_t)p] = q;
        }
        it.pass.out_bit = out;
        plan.push_back(std::move(it));
    }
    if (moved)
        for (int q = 0; q < n; q++)
            if (phys[(size_t)q] != q) throw Error(QFVM_ERR_RUNTIME, "layout-aware plan did not restore the qubit order");
    return plan;
}

// ------------------------------------------------------------------------------------------------
// Plan cache. The planner reads, per gate, its kind, qubits, controls and class (info_of); the
// coefficients only reach the plan's ops through convert. A plan is therefore valid for every gate
// list with the same per-gate structure words below, once each op's coefficients are re-read from
// its source gate.
struct PlanCache {
    struct Entry {
        uint64_t hash;
        std::vector<uint64_t> sig;
        std::vector<PlanItem> plan;
        uint64_t last;
    };
    std::vector<Entry> entries;
    std::vector<uint64_t> sig;  // scratch
    uint64_t clock = 0;
};
constexpr size_t PLAN_CACHE_MAX = 64;  // plans kept per handle (least recently used dropped first)

namespace {

void plan_signature(const std::vector<Gate>& gates, int n, const FuseConfig& cfg, std::vector<uint64_t>& sig) {
    sig.clear();
    sig.reserve(gates.size() * 5 + 2);
    sig.push_back((uint64_t)n | ((uint64_t)cfg.tile_bits << 8) | ((uint64_t)cfg.low_bits << 16) |
                  ((uint64_t)cfg.max_groups << 24) | ((uint64_t)cfg.group_bits << 32));
    sig.push_back((uint64_t)(uint32_t)cfg.max_cost | ((uint64_t)cfg.layout << 32));
    for (const Gate& g : gates) {
        const GInfo gi = info_of(g);
        const uint64_t cls = (gi.diag ? 1u : 0u) | (gi.multi ? 2u : 0u) | (gi.grad ? 4u : 0u) |
                             (gi.diag && gi.d0 == cd(0.0, 0.0) ? 8u : 0u) | (gi.diag && gi.d1 == cd(0.0, 0.0) ? 16u : 0u);
        sig.push_back((uint64_t)(uint32_t)g.kind | ((uint64_t)(uint32_t)g.target << 8) | ((uint64_t)(uint32_t)g.gen << 16) |
                      (cls << 24) | ((uint64_t)(uint32_t)g.slot << 32));
        sig.push_back(g.cmask);
        sig.push_back(gi.N);
        sig.push_back(gi.D);
        if (gi.multi) {
            sig.push_back(g.targets.size() | ((uint64_t)g.ctrls.size() << 32));
            for (int t : g.targets) sig.push_back((uint64_t)(uint32_t)t);
        }
    }
}

// re-reads every op's coefficients (and every single gate) from `gates`
void refresh_plan(std::vector<PlanItem>& plan, const std::vector<Gate>& gates) {
    for (PlanItem& it : plan) {
        if (!it.fused) {
            it.gate = move_gate(gates[(size_t)it.src], it.l2p);
            continue;
        }
        for (FGroup& fg : it.pass.groups) {
            if (fg.stage) continue;
            for (FOpHost& o : fg.ops) {
                if (o.src < 0) continue;
                switch (o.kind) {
                    case F_DENSE:
                    case F_REAL:
                    case F_DIAG:
                    case F_PHASE:
                    case F_SWAP: {
                        const Gate& g = gates[(size_t)o.src];
                        const FOpHost c = convert(g, info_of(g), fg.S);
                        for (int k = 0; k < 4; k++) o.m[k] = c.m[k];
                        break;
                    }
                    case F_DENSE4: o.mx = gates[(size_t)o.src].mk; break;
                    default: break;  // relabelings, flushes: no coefficients
                }
            }
        }
    }
}

}  // namespace

const std::vector<PlanItem>& cached_plan(State& s, const std::vector<Gate>& gates, const FuseConfig& cfg) {
    if (!s.plan_cache) s.plan_cache = std::make_shared<PlanCache>();
    PlanCache& pc = *s.plan_cache;
    plan_signature(gates, s.n, cfg, pc.sig);
    uint64_t h = 1469598103934665603ull;
    for (uint64_t w : pc.sig) h = (h ^ w) * 1099511628211ull;
    pc.clock++;
    for (PlanCache::Entry& e : pc.entries) {
        if (e.hash == h && e.sig == pc.sig) {
            e.last = pc.clock;
            refresh_plan(e.plan, gates);
            s.kstats[(size_t)s.stat_index("plan_cache_hit")].second.launches++;
            return e.plan;
        }
    }
    if (pc.entries.size() >= PLAN_CACHE_MAX) {
        size_t old = 0;
        for (size_t k = 1; k < pc.entries.size(); k++)
            if (pc.entries[k].last < pc.entries[old].last) old = k;
        pc.entries.erase(pc.entries.begin() + (std::ptrdiff_t)old);
    }
    pc.entries.push_back(PlanCache::Entry{h, pc.sig, plan_gates(gates, s.n, cfg), pc.clock});
    s.kstats[(size_t)s.stat_index("plan_cache_miss")].second.launches++;
    return pc.entries.back().plan;
}

static FuseConfig config_of(const State& s, int states) {
    FuseConfig cfg;
    // the two-state sweep holds both tiles in LDS, so an explicit tile size is halved for it
    cfg.tile_bits = s.tile_bits > 0 ? s.tile_bits - (states - 1) : default_tile_bits(s.dtype, states);
    cfg.tile_bits = std::max(FUSE_GROUP_BITS, std::min(cfg.tile_bits, FUSE_MAX_TILE_BITS));
    cfg.low_bits = s.low_bits > 0 ? s.low_bits : default_low_bits(s.dtype, states);
    cfg.max_groups = std::max(1, std::min(s.max_groups, FUSE_MAX_GROUPS));
    cfg.max_cost = s.pass_cost;
    // one-state passes run on the row-of-8 kernel (3-qubit groups) when its tile fits: complex128,
    // and complex64 with option "rows8_c64"
    const int b = std::min(cfg.tile_bits, s.n);
    if (states == 1 && s.rows8 && (s.dtype == QFVM_C128 || s.rows8_c64) && b >= 10 && b <= 13) cfg.group_bits = 3;
    cfg.layout = s.layout && cfg.group_bits == 3;
    return cfg;
}

void apply_circuit(State& s, const std::vector<Gate>& gates_in) {
    if (gates_in.empty()) return;
    // runs of one-qubit gates on one qubit: one gate each
    std::vector<Gate> merged;
    if (s.merge_1q) merged = merge_single_qubit_runs(gates_in);
    const std::vector<Gate>& gates = s.merge_1q ? merged : gates_in;
    // deep runs on few qubits: one dense gate per run on the matrix cores (below 2^14 amplitudes a
    // launch costs more than the arithmetic it saves)
    if (s.dense_blocks && s.mfma && s.n >= 14)
        apply_gates(s, collapse_dense_blocks(gates, s.dense_blocks, s.dense_min_gates));
    else
        apply_gates(s, gates);
}

void apply_gates(State& s, const std::vector<Gate>& gates) {
    if (gates.empty()) return;
    if (!s.fusion || s.n < 6) {
        for (const Gate& g : gates) launch_gate(s, g);
        return;
    }
    FuseConfig cfg = config_of(s, 1);
    // complex64 runs rows of 16 by default; a gate list with dense 2-target gates takes the
    // row-of-8 kernel instead, whose passes hold those gates (each would otherwise cost a sweep)
    const int b = std::min(cfg.tile_bits, s.n);
    if (cfg.group_bits != 3 && s.rows8 && s.dtype == QFVM_C64 && b >= 10 && b <= 13) {
        for (const Gate& g : gates)
            if (g.kind == G_MULTI && g.targets.size() == 2) {
                cfg.group_bits = 3;
                cfg.layout = s.layout;
                break;
            }
    }
    if (s.plan_cache_on)
        execute_plan(s, cached_plan(s, gates, cfg));
    else
        execute_plan(s, plan_gates(gates, s.n, cfg));
}

void apply_gates2(State& s, const std::vector<Gate>& gates, void* lam, double* grads_dev) {
    if (gates.empty()) return;
    if (!s.fusion || s.n < 6) {
        for (const Gate& g : gates) {
            if (g.kind == G_GRAD) {
                launch_grad_marker(s, g, lam, grads_dev);
                continue;
            }
            launch_gate(s, g);
            BufferScope other(s, lam);
            launch_gate(s, g);
        }
        return;
    }
    const FuseConfig cfg = config_of(s, 2);
    if (s.plan_cache_on)
        execute_plan(s, cached_plan(s, gates, cfg), lam, grads_dev);
    else
        execute_plan(s, plan_gates(gates, s.n, cfg), lam, grads_dev);
}

// ------------------------------------------------------------------------------------------------
int State::stat_index(const char* name) {
    for (size_t k = 0; k < kstats.size(); k++)
        if (kstats[k].first == name) return (int)k;
    kstats.emplace_back(name, KStat{});
    return (int)kstats.size() - 1;
}

void State::resolve_events() {
    for (PendingEvent& p : pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) kstats[p.stat].second.ms += ms;
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    pending.clear();
}

KernelScope::KernelScope(State& st, const char* name, double bytes) : s(st), idx(st.stat_index(name)) {
    s.kstats[idx].second.launches++;
    s.kstats[idx].second.bytes += bytes;
    if (s.profiling) {
        QFVM_HIP(hipEventCreate(&a));
        QFVM_HIP(hipEventCreate(&b));
        QFVM_HIP(hipEventRecord(a, s.stream));
    }
}

KernelScope::~KernelScope() {
    if (s.profiling && a) {
        hipEventRecord(b, s.stream);
        s.pending.push_back({idx, a, b});
    }
}

}  // namespace qfvm
