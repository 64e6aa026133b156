// Kernel-side shot description shared by the plain and batched paths.
#pragma once
#include "fd_common.h"

namespace swit {

struct ShotGeom {
    int isx, isz;       // forward source cell, padded, after the free-surface push-down
                        // (fdcore.f90:70-72)
    int isx_i, isz_i;   // reconstruction source cell, interior, no push-down
                        // (fdcore.f90:216-218)
    int ng;             // receivers of this shot
    int pad[3];
};

// Strip map entry: field (0 = p, 1 = u, 2 = w) in the top 2 bits, padded offset below.
constexpr uint32_t kStripFieldShift = 30;

}  // namespace swit
