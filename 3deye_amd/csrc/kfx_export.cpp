// kfx_export.cpp — files of the world model: the reference's matrix text
// format for extracted slices, PCD point clouds (what the reference's PCL
// conversion feeds to PCL), and a binary file of (x, y, z, tsdf, weight)
// slice records.  Host code; inputs are host arrays.
#include <stdio.h>
#include <string.h>

#include "../../include/kfx.h"

namespace {

// Default ostream formatting of a float (6 significant digits, %g rules),
// as the reference's operator<< writes it.
void put_float(FILE* f, float v) { fprintf(f, "%g", (double)v); }

int finish(FILE* f) { return fclose(f) == 0 ? KFX_OK : KFX_ERR_INVALID_ARG; }

}  // namespace

extern "C" {

// matrix<T>::Save_To_File (common/inc/impl/matrix.hpp:455-476) writes
// operator<< of the matrix (matrix.hpp:411-438): "width height depth type"
// and a newline, then each of the depth*height rows as its width elements
// each followed by a space, and a newline.  A slice is a 1-D matrix of n
// elements (n x 1 x 1).  XYZI elements print as x, y, z, w joined by 'x'
// (COMMON_DELIMITER; base_points.cpp:570-574), Color elements as r.g.b.w
// and a space (COLOR_DELIMITER '.'; base_points.cpp:278-282).  The type
// field is the element type's name (the reference prints the compiler's
// typeid name, which differs between compilers).
int kfx_save_points_text(const char* path, const kfx_xyzi* points, uint64_t n) {
    if (!path || (n > 0 && !points)) return KFX_ERR_INVALID_ARG;
    FILE* f = fopen(path, "w");
    if (!f) return KFX_ERR_INVALID_ARG;
    fprintf(f, "%llu 1 1 Point_XYZI\n", (unsigned long long)n);
    for (uint64_t i = 0; i < n; ++i) {
        put_float(f, points[i].x);
        fputc('x', f);
        put_float(f, points[i].y);
        fputc('x', f);
        put_float(f, points[i].z);
        fputc('x', f);
        put_float(f, points[i].w);
        fputc(' ', f);
    }
    fputc('\n', f);
    return finish(f);
}

int kfx_save_colors_text(const char* path, const kfx_color* colors, uint64_t n) {
    if (!path || (n > 0 && !colors)) return KFX_ERR_INVALID_ARG;
    FILE* f = fopen(path, "w");
    if (!f) return KFX_ERR_INVALID_ARG;
    fprintf(f, "%llu 1 1 Point_Color\n", (unsigned long long)n);
    for (uint64_t i = 0; i < n; ++i)
        fprintf(f, "%u.%u.%u.%u  ", (unsigned)colors[i].r, (unsigned)colors[i].g, (unsigned)colors[i].b,
                (unsigned)colors[i].w);
    fputc('\n', f);
    return finish(f);
}

// PCD v0.7, the file form of the clouds the reference builds for PCL
// (convert_to_point_cloud, common/src/math_utils.cpp:216-328): with colours,
// pcl::PointXYZRGB with the reference's packing (its r takes the colour's b
// and its b the colour's r: rgb = b << 16 | g << 8 | r, read as a float);
// without, pcl::PointXYZI with intensity = w.  binary != 0 writes
// "DATA binary" (little-endian records of 16 B), else "DATA ascii".
int kfx_save_pcd(const char* path, const kfx_xyzi* points, const kfx_color* colors, uint64_t n, int binary) {
    if (!path || (n > 0 && !points)) return KFX_ERR_INVALID_ARG;
    FILE* f = fopen(path, binary ? "wb" : "w");
    if (!f) return KFX_ERR_INVALID_ARG;
    const char* last = colors ? "rgb" : "intensity";
    fprintf(f,
            "# .PCD v0.7 - Point Cloud Data file format\nVERSION 0.7\nFIELDS x y z %s\nSIZE 4 4 4 4\n"
            "TYPE F F F F\nCOUNT 1 1 1 1\nWIDTH %llu\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS %llu\nDATA %s\n",
            last, (unsigned long long)n, (unsigned long long)n, binary ? "binary" : "ascii");
    for (uint64_t i = 0; i < n; ++i) {
        float rec[4] = {points[i].x, points[i].y, points[i].z, points[i].w};
        if (colors) {
            const uint32_t rgb = (uint32_t)colors[i].b << 16 | (uint32_t)colors[i].g << 8 | (uint32_t)colors[i].r;
            memcpy(&rec[3], &rgb, 4);
        }
        if (binary) {
            if (fwrite(rec, 4, 4, f) != 4) {
                fclose(f);
                return KFX_ERR_INVALID_ARG;
            }
        } else {
            // ASCII float fields with 9 significant digits (exact round trip);
            // the packed colour as its unsigned integer, as PCL writes rgb.
            fprintf(f, "%.9g %.9g %.9g ", (double)rec[0], (double)rec[1], (double)rec[2]);
            if (colors) {
                uint32_t u;
                memcpy(&u, &rec[3], 4);
                fprintf(f, "%u\n", u);
            } else {
                fprintf(f, "%.9g\n", (double)rec[3]);
            }
        }
    }
    return finish(f);
}

// Binary file of slice records: the 8-byte magic "KFXREC1\0", the record
// count as a little-endian uint64, then the kfx_slice_record array
// (16 B each: int32 x, y, z; int16 tsdf, weight).
int kfx_save_records(const char* path, const kfx_slice_record* records, uint64_t n) {
    if (!path || (n > 0 && !records)) return KFX_ERR_INVALID_ARG;
    FILE* f = fopen(path, "wb");
    if (!f) return KFX_ERR_INVALID_ARG;
    const char magic[8] = {'K', 'F', 'X', 'R', 'E', 'C', '1', '\0'};
    bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(&n, 8, 1, f) == 1;
    if (ok && n) ok = fwrite(records, sizeof(kfx_slice_record), n, f) == n;
    if (!ok) {
        fclose(f);
        return KFX_ERR_INVALID_ARG;
    }
    return finish(f);
}

}  // extern "C"
