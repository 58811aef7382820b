// dim3.h — extent of a buffer (reference: common/inc/types/dim3.h).
#ifndef KFX_DROPIN_DIM3
#define KFX_DROPIN_DIM3

#include <stddef.h>

struct dim3 {
    union {
        size_t data[3];
        struct {
            size_t x, y, z;
        };
        struct {
            size_t width, height, depth;
        };
    };
    dim3(size_t nx = 1, size_t ny = 1, size_t nz = 1) : x(nx), y(ny), z(nz) {}
};

#endif
