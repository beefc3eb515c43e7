// Per-bin autotuner: for each length class of a batch it times the candidate kernel mappings on a
// sample of the batch's own jobs and keeps the fastest, per context and scoring.  Every candidate
// computes identical results, so the choice changes speed only.
#pragma once
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "bwg_internal.h"

namespace bwg {

// Kernel mappings a bin can be run with.
enum ItemKind { kExtGroup, kExtThread, kExtPlain, kKswFwd, kKswTracker, kKswRev, kKswPlain, kExtUngapped,
                kExtRetryFilter, kKswMateRc, kKswWindows, kGlbGroup, kGlbTrace, kExtGroup8, kGlbPlain };

// The extension's 16- and 8-lane group mappings.
inline bool is_ext_group(int kind) { return kind == kExtGroup || kind == kExtGroup8; }
// Bins drained by the extension DP kernels (group or thread mappings).
inline bool is_ext_dp(int kind) { return is_ext_group(kind) || kind == kExtThread; }

struct TuneCandidate {
    int kind = 0;
    int param = 0;            // K, or the thread kernel's column capacity
    int width = 32;           // bits of the stored DP state (16: narrow, 32: wide)
    double ns_per_job = 0.0;  // best of the timed repetitions
};

struct TuneEntry {
    int choice = 0;           // index into cands
    uint32_t sample = 0;      // jobs timed per candidate
    std::vector<TuneCandidate> cands;
};

class Tuner {
  public:
    bool lookup(uint64_t sig, int cls, TuneEntry *out) const;
    void store(uint64_t sig, int cls, const TuneEntry &e);
    void clear();
    // One line per tuned class: the class, the sample size and each candidate's time per job, the
    // chosen one marked with '*'.
    std::string report() const;
    std::mutex tune_mu;  // one tuning run at a time per context

  private:
    mutable std::mutex mu_;
    std::map<std::pair<uint64_t, int>, TuneEntry> table_;
};

// Extension length classes: queries up to 32 in buckets of 4 (classes 0..7), longer queries by the
// group kernel's minimal K for their widest window (classes 8..).  Only jobs whose scores fit the
// narrow state are classed and tuned; -2: wide state, untuned; -1: always the plain kernel.
constexpr int kExtShortClasses = 8;
// Rescue classes: the minimal K of ksw_group_kernel (classes 0..7 for K = 2..16).

uint64_t scoring_sig(const ExtScoring &sc);
uint64_t scoring_sig(const KswScoring &sc);
const char *mapping_name(int kind, int param, int width, char *buf, size_t len);

}  // namespace bwg
