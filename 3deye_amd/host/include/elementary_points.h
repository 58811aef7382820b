// elementary_points.h — small vector/matrix PODs of the reference host API
// (common/inc/elementary_points.h, types/elementary_point_template.h,
// types/elementary_matrix_template.h): same names, layouts and operations
// (multiply = row . vector), layout-identical to include/kfx_types.h.
#ifndef KFX_DROPIN_ELEMENTARY_POINTS
#define KFX_DROPIN_ELEMENTARY_POINTS

#include <math.h>

#include "kfx_types.h"

typedef struct point_3f {
    union {
        float data[3];
        struct {
            float x, y, z;
        };
    };
} point_3f;
typedef point_3f vector_3f;

typedef struct point_3i {
    union {
        int data[3];
        struct {
            int x, y, z;
        };
    };
} point_3i;

typedef struct point_2us {
    union {
        unsigned int data[2];
        struct {
            unsigned int x, y;
        };
    };
} point_2us;

typedef struct matrix_3f {
    point_3f data[3];
} matrix_3f;

inline point_3f make_point_3f(float x, float y, float z) {
    point_3f p;
    p.x = x; p.y = y; p.z = z;
    return p;
}
inline point_3i make_point_3i(int x, int y, int z) {
    point_3i p;
    p.x = x; p.y = y; p.z = z;
    return p;
}
inline float dot_3f(point_3f a, point_3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline point_3f add_3f(point_3f a, point_3f b) { return make_point_3f(a.x + b.x, a.y + b.y, a.z + b.z); }
inline point_3f subtract_3f(point_3f a, point_3f b) { return make_point_3f(a.x - b.x, a.y - b.y, a.z - b.z); }
inline point_3f multiply(const matrix_3f* m, point_3f* v) {
    return make_point_3f(dot_3f(m->data[0], *v), dot_3f(m->data[1], *v), dot_3f(m->data[2], *v));
}

static_assert(sizeof(point_3f) == sizeof(kfx_point_3f), "point_3f layout");
static_assert(sizeof(point_3i) == sizeof(kfx_point_3i), "point_3i layout");
static_assert(sizeof(point_2us) == sizeof(kfx_point_2us), "point_2us layout");
static_assert(sizeof(matrix_3f) == sizeof(kfx_matrix_3f), "matrix_3f layout");

#endif
