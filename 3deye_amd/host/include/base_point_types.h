// base_point_types.h — element types of the reference's buffers
// (common/inc/types/base_points.h, base_point_types.h): the Point_<T>
// wrapper holds one `data` member of the element struct, so every Point_*
// has the element's layout.
#ifndef KFX_DROPIN_BASE_POINT_TYPES
#define KFX_DROPIN_BASE_POINT_TYPES

#include "kfx_types.h"

struct Tsdf {
    short value;
    short weight;
};
struct Depth {
    unsigned short value;
};
struct Color {
    unsigned char r, g, b, w;
};
struct XYZ {
    float x, y, z;
};
typedef XYZ NXYZ;
struct XYZI {
    union {
        float value[4];
        struct {
            float x, y, z, w;
        };
    };
};

template <class T>
struct Point {
    T data;
};
typedef Point<Tsdf> Point_Tsdf;
typedef Point<Depth> Point_Depth;
typedef Point<Color> Point_Color;
typedef Point<XYZ> Point_XYZ;
typedef Point<NXYZ> Point_NXYZ;
typedef Point<XYZI> Point_XYZI;

static_assert(sizeof(Point_Tsdf) == sizeof(kfx_tsdf), "Tsdf layout");
static_assert(sizeof(Point_Depth) == 2, "Depth layout");
static_assert(sizeof(Point_Color) == sizeof(kfx_color), "Color layout");
static_assert(sizeof(Point_XYZ) == sizeof(kfx_xyz), "XYZ layout");
static_assert(sizeof(Point_XYZI) == sizeof(kfx_xyzi), "XYZI layout");

#endif
