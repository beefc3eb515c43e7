// Launch-shape parameters of the batched form, per device.
#pragma once
#include "bwg_runtime.h"

namespace bwg {

struct Tuning {
    int ext_thread_max_q16 = 16;   // queries up to this length run one job per lane, 16-column LDS
    int ext_thread_max_q32 = 32;   // ... 32-column LDS
    int ext_group_max_qcap = 256;  // longest query the 16-lane group kernel keeps in LDS
    int ext_group_oversub = 1;     // persistent grid = oversub * occupancy * CUs
    int ksw_group_oversub = 1;
};

const Tuning &tuning_for(bwg_ctx *c);

}  // namespace bwg
