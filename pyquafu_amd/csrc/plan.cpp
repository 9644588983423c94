// Gate-run execution. This first version launches one pair-update kernel per gate.
#include "plan.h"

namespace qfvm {

void apply_gates(State& s, const std::vector<Gate>& gates) {
    for (const Gate& g : gates) launch_gate(s, g);
}

}  // namespace qfvm

namespace qfvm {

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
