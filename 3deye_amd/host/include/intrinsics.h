// intrinsics.h — pinhole camera of the reference host API
// (common/inc/types/intrinsics.h; operator() as common/src/intrinsics.cpp:7-18).
#ifndef KFX_DROPIN_INTRINSICS
#define KFX_DROPIN_INTRINSICS

#include "kfx_types.h"

typedef struct intrinsics {
    float fx, fy, cx, cy;
    int cols, rows;
    // Camera of pyramid level `level`: focal lengths and principal point
    // divided by 2^level, image size shifted right by level.
    intrinsics operator()(int level) const {
        const int div = 1 << level;
        intrinsics r = {fx / div, fy / div, cx / div, cy / div, cols >> level, rows >> level};
        return r;
    }
} intrinsics;

static_assert(sizeof(intrinsics) == sizeof(kfx_intrinsics), "intrinsics layout");

#endif
