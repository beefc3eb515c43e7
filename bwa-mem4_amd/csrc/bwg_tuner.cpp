// Autotuner table and report (the timing runs live with the planner in bwg_batched.cpp).
#include "bwg_tuner.h"

#include "bwg_batched_kernels.h"

#include <cstdio>
#include <cstring>

namespace bwg {

namespace {
uint64_t mix(uint64_t h, uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    return h;
}
}  // namespace

uint64_t scoring_sig(const ExtScoring &sc) {
    uint64_t h = 0x45585400ull;  // "EXT"
    for (int k = 0; k < 25; ++k) h = mix(h, (uint64_t)(uint8_t)sc.mat[k]);
    for (int32_t v : {sc.o_del, sc.e_del, sc.o_ins, sc.e_ins, sc.w, sc.end_bonus, sc.zdrop})
        h = mix(h, (uint64_t)(uint32_t)v);
    return h;
}

uint64_t scoring_sig(const KswScoring &sc) {
    uint64_t h = 0x4b535700ull;  // "KSW"
    for (int k = 0; k < 25; ++k) h = mix(h, (uint64_t)(uint8_t)sc.mat[k]);
    for (int32_t v : {sc.o_del, sc.e_del, sc.o_ins, sc.e_ins, sc.minsc, sc.max_sc})
        h = mix(h, (uint64_t)(uint32_t)v);
    return h;
}

const char *mapping_name(int kind, int param, int width, char *buf, size_t len) {
    switch (kind) {
    case kExtThread: std::snprintf(buf, len, "ext_thread<%d,i%d>", param, width); break;
    case kExtGroup: std::snprintf(buf, len, "ext_group<%d,i%d>", param, width); break;
    case kExtGroup8: std::snprintf(buf, len, "ext_group8<%d,i%d>", param, width); break;
    // packed 8-lane bins are named by their columns per lane, twice the 16-lane K
    case kKswFwd:
        if (width == 16 && ksw_param_lanes(param) == 8)
            std::snprintf(buf, len, "ksw_pk8<%d,i16>", 2 * ksw_param_k(param));
        else
            std::snprintf(buf, len, width == 16 ? "ksw_pk<%d,i16>" : "ksw_group<%d,i32>", param);
        break;
    case kKswRev:
        if (width == 16 && ksw_param_lanes(param) == 8)
            std::snprintf(buf, len, "ksw_pk8<%d,i16,reverse>", 2 * ksw_param_k(param));
        else
            std::snprintf(buf, len, width == 16 ? "ksw_pk<%d,i16,reverse>" : "ksw_group<%d,i32,reverse>", param);
        break;
    case kKswTracker: std::snprintf(buf, len, "ksw_tracker"); break;
    case kExtUngapped: std::snprintf(buf, len, "ext_ungapped"); break;
    case kExtRetryFilter: std::snprintf(buf, len, "ext_retry_filter"); break;
    case kKswMateRc: std::snprintf(buf, len, "mate_revcomp"); break;
    case kKswWindows: std::snprintf(buf, len, "matesw_windows"); break;
    case kGlbGroup: std::snprintf(buf, len, width == 16 ? "glb_group<%d,packed>" : "glb_group<%d>", param); break;
    case kGlbTrace: std::snprintf(buf, len, "glb_traceback"); break;
    case kGlbPlain: std::snprintf(buf, len, "glb_plain"); break;
    case kExtPlain: std::snprintf(buf, len, "ext_plain"); break;
    case kKswPlain: std::snprintf(buf, len, "ksw_plain"); break;
    default: std::snprintf(buf, len, "?"); break;
    }
    return buf;
}

bool Tuner::lookup(uint64_t sig, int cls, TuneEntry *out) const {
    std::lock_guard<std::mutex> g(mu_);
    auto it = table_.find({sig, cls});
    if (it == table_.end()) return false;
    if (out) *out = it->second;
    return true;
}

void Tuner::store(uint64_t sig, int cls, const TuneEntry &e) {
    std::lock_guard<std::mutex> g(mu_);
    table_[{sig, cls}] = e;
}

void Tuner::clear() {
    std::lock_guard<std::mutex> g(mu_);
    table_.clear();
}

std::string Tuner::report() const {
    std::lock_guard<std::mutex> g(mu_);
    std::string s;
    char line[512], name[64];
    for (const auto &kv : table_) {
        const int cls = kv.first.second;
        const TuneEntry &e = kv.second;
        const bool resc = !e.cands.empty() && e.cands[0].kind == kKswFwd;
        int n;
        if (resc)
            n = std::snprintf(line, sizeof line, "rescue pad<=%d%s", 32 * (cls % 8 + 1),
                              cls < 8 ? " (16-bit exact)" : "");
        else if (cls < kExtShortClasses)
            n = std::snprintf(line, sizeof line, "extend qlen %d-%d", 4 * cls + 1, 4 * cls + 4);
        else
            n = std::snprintf(line, sizeof line, "extend window class %d", cls - kExtShortClasses);
        n += std::snprintf(line + n, sizeof line - n, " (sig %016llx, %u jobs):",
                           (unsigned long long)kv.first.first, e.sample);
        for (size_t k = 0; k < e.cands.size() && n < (int)sizeof line - 80; ++k)
            n += std::snprintf(line + n, sizeof line - n, " %s%s %.3f ns/job",
                               (int)k == e.choice ? "*" : "",
                               mapping_name(e.cands[k].kind, e.cands[k].param, e.cands[k].width, name, sizeof name),
                               e.cands[k].ns_per_job);
        s += line;
        s += '\n';
    }
    return s;
}

}  // namespace bwg
