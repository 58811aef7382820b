// dropin_check.cpp — drives the reference host API (opencl_wrappers.h) over
// libkfx the way the reference's absent fusion application would for two
// frames, and dumps every intermediate result for the pytest comparison with
// the oracle (tests/test_gpu_dropin.py).
//
//   dropin_check <dir>   reads <dir>/in.bin, writes <dir>/*.bin
//
// in.bin: int32 W, H, N; float fx, fy, cx, cy, volume_m; float R0[9], t0[3]
// (camera -> volume, frame 0); uint16 depth0[W*H], depth1[W*H]; Color rgb0[W*H].
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "opencl_wrappers.h"

static void dump(const std::string& dir, const char* name, const void* p, size_t bytes) {
    FILE* f = fopen((dir + "/" + name).c_str(), "wb");
    if (!f || fwrite(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "cannot write %s\n", name);
        exit(1);
    }
    fclose(f);
}

template <class T>
static std::vector<T> download(gpu_array_base* b) {
    std::vector<T> v(b->get_total_size());
    b->memcpy_to_host(v.data());
    return v;
}

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: %s <dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    FILE* f = fopen((dir + "/in.bin").c_str(), "rb");
    if (!f) return 2;
    int hdr[3];
    float cam[5], pose[12];
    if (fread(hdr, 4, 3, f) != 3 || fread(cam, 4, 5, f) != 5 || fread(pose, 4, 12, f) != 12) return 2;
    const int W = hdr[0], H = hdr[1], N = hdr[2];
    std::vector<uint16_t> d0((size_t)W * H), d1((size_t)W * H);
    std::vector<Point_Color> rgb0((size_t)W * H);
    if (fread(d0.data(), 2, d0.size(), f) != d0.size() || fread(d1.data(), 2, d1.size(), f) != d1.size() ||
        fread(rgb0.data(), 4, rgb0.size(), f) != rgb0.size())
        return 2;
    fclose(f);

    try {
        std::vector<std::string> names = {"compute_bilateral_filtering", "compute_depth_pyramid",
                                          "compute_vertex_map",          "compute_normal_map_eigen",
                                          "integrate_tsdf_volume",       "raycast",
                                          "resize_vertex_map",           "resize_normal_map",
                                          "estimate_combined",           "transform_estimation",
                                          "transform_map",               "clean_buff",
                                          "clean_tsdf_slice",            "integrate_color_volume",
                                          "clean_color_slice",           "extract_colors"};
        kernel_manager km("unused", names);
        intrinsics intr = {cam[0], cam[1], cam[2], cam[3], W, H};
        round_buffer rb(make_point_3i(N, N, N), make_point_3f(cam[4], cam[4], cam[4]));
        point_3f cell = make_point_3f(cam[4] / N, cam[4] / N, cam[4] / N);
        matrix_3f R0, R0inv;
        vector_3f t0;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                R0.data[r].data[c] = pose[3 * r + c];
                R0inv.data[r].data[c] = pose[3 * c + r];
            }
        for (int k = 0; k < 3; ++k) t0.data[k] = pose[9 + k];

        gpu_array<Point_Depth> raw(dim3(W, H));
        raw.memcpy_from_host(d0.data());
        gpu_array<float> scaled(dim3(W, H));
        gpu_array<Point_Tsdf> volume(dim3(N, N, N));
        volume.reset_gpu_part();

        // Frame 0: integrate, raycast (the model maps), their pyramid.
        Integrate_Tsdf_Data(&raw, &scaled, &volume, intr, R0inv, t0, DEFAULT_TRUNCATION_DISTANCE, cell, &rb, &km, 0);
        std::vector<gpu_array_base*> vp, np_;
        for (int l = 0; l < 3; ++l) {
            vp.push_back(new gpu_array<Point_XYZ>(dim3(W >> l, H >> l)));
            np_.push_back(new gpu_array<Point_NXYZ>(dim3(W >> l, H >> l)));
        }
        Raycast(&volume, vp[0], np_[0], &rb, t0, R0, DEFAULT_TRUNCATION_DISTANCE, intr, cell, km.Get_Kernel("raycast"));
        for (int l = 1; l < 3; ++l) {
            intrinsics li = intr(l - 1);
            Resize_Vertex_Map(vp[l - 1], vp[l], li, km.Get_Kernel("resize_vertex_map"));
            Resize_Normal_Map(np_[l - 1], np_[l], li, km.Get_Kernel("resize_normal_map"));
        }
        // Colours of frame 0 at its raycast, weight cap 1 (DEFAULT_MAX_WEIGHT).
        gpu_array<Point_Color> rgb(dim3(W, H));
        rgb.memcpy_from_host(rgb0.data());
        gpu_array<Point_Color> colors(dim3(N, N, N));
        colors.reset_gpu_part();
        Integrate_Colors(&colors, &rgb, vp[0], intr, R0inv, t0, DEFAULT_TRUNCATION_DISTANCE, DEFAULT_MAX_WEIGHT, cell,
                         &rb, km.Get_Kernel("integrate_color_volume"));
        auto col0 = download<Point_Color>(&colors);
        dump(dir, "colors0.bin", col0.data(), col0.size() * 4);
        auto vol0 = download<Point_Tsdf>(&volume);
        dump(dir, "volume0.bin", vol0.data(), vol0.size() * sizeof(Point_Tsdf));
        auto v0 = download<Point_XYZ>(vp[0]);
        auto n0 = download<Point_NXYZ>(np_[0]);
        dump(dir, "vmap0.bin", v0.data(), v0.size() * 12);
        dump(dir, "nmap0.bin", n0.data(), n0.size() * 12);

        // Frame 1: bilateral, pyramid, vertex and normal maps.
        raw.memcpy_from_host(d1.data());
        std::vector<gpu_array_base*> dp, vc, nc;
        for (int l = 0; l < 3; ++l) {
            dp.push_back(new gpu_array<Point_Depth>(dim3(W >> l, H >> l)));
            vc.push_back(new gpu_array<Point_XYZ>(dim3(W >> l, H >> l)));
            nc.push_back(new gpu_array<Point_NXYZ>(dim3(W >> l, H >> l)));
        }
        Apply_Billateral_Filtration(&raw, intr, DEFAULT_TRUNCATION_DISTANCE,
                                    km.Get_Kernel("compute_bilateral_filtering"), dp[0]);
        Create_Pyramid(dp, 3, intr, &km);
        Create_Vertex_Maps(dp, vc, intr, 3, &km);
        Create_Normal_Maps(vc, nc, intr, 3, &km);
        for (int l = 0; l < 3; ++l) {
            auto d = download<Point_Depth>(dp[l]);
            auto v = download<Point_XYZ>(vc[l]);
            char name[32];
            snprintf(name, sizeof(name), "depth1_l%d.bin", l);
            dump(dir, name, d.data(), d.size() * 2);
            snprintf(name, sizeof(name), "vmap1_l%d.bin", l);
            dump(dir, name, v.data(), v.size() * 12);
        }
        auto nn = download<Point_NXYZ>(nc[0]);
        dump(dir, "nmap1_l0.bin", nn.data(), nn.size() * 12);

        // One ICP step at level 0 from the previous pose (A, b).
        const int stride = DIVUP(W, 32) * DIVUP(H, 8);
        gpu_array<float> gbuf(dim3(27 * stride));
        gpu_array<float> mbuf(dim3(27));
        Clean_Buff(&gbuf, km.Get_Kernel("clean_buff"));
        double A[36], b[6];
        Estimate_Combined(R0, t0, vc[0], nc[0], R0inv, t0, intr, vp[0], np_[0], DEFAULT_DISTANCE_THRESHOLD,
                          (float)sin(20.0 * M_PI / 180.0), &gbuf, &mbuf, A, b, &km);
        dump(dir, "A.bin", A, sizeof(A));
        dump(dir, "b.bin", b, sizeof(b));

        // Transform the model maps to the camera of frame 0 (R0^T, -R0^T t0).
        vector_3f tinv;
        for (int r = 0; r < 3; ++r)
            tinv.data[r] = -(R0inv.data[r].x * t0.x + R0inv.data[r].y * t0.y + R0inv.data[r].z * t0.z);
        gpu_array<Point_XYZ> vt(dim3(W, H));
        gpu_array<Point_NXYZ> nt(dim3(W, H));
        Transform_Maps(vp[0], np_[0], &vt, &nt, R0inv, tinv, intr, km.Get_Kernel("transform_map"));
        auto vtv = download<Point_XYZ>(&vt);
        dump(dir, "vmap0_cam.bin", vtv.data(), vtv.size() * 12);

        // Shift by 16 voxels in x: extract the slab leaving the cube, clean it.
        point_3i mn = make_point_3i(rb.origin_GRID.x, 1 << 30, 1 << 30);
        point_3i mx = make_point_3i(rb.origin_GRID.x + 15, -1, -1);
        matrix<Point_XYZI> pts;
        matrix<Point_Color> cols;
        Extract_Slice(mn, mx, &rb, &volume, &colors, cell, &km, pts, cols);
        dump(dir, "slice_points.bin", pts.data, pts.Get_Num_Of_Elem() * 16 * (pts.data != nullptr));
        dump(dir, "slice_colors.bin", cols.data, cols.Get_Num_Of_Elem() * 4 * (cols.data != nullptr));
        pts.Save_To_File(dir + "/slice_points.tst");
        cols.Save_To_File(dir + "/slice_colors.tst");
        // The whole cube as a metric cloud, and the colours under it.
        matrix<Point_XYZI> cloud;
        matrix<Point_Color> cloud_colors;
        Extract_Tsdf_Volume(&rb, cell, &volume, &km, cloud);
        Extract_Color_Volume(&colors, &rb, cloud, cell, km.Get_Kernel("extract_colors"), cloud_colors);
        dump(dir, "volume_cloud.bin", cloud.data, cloud.Get_Num_Of_Elem() * 16 * (cloud.data != nullptr));
        dump(dir, "volume_cloud_colors.bin", cloud_colors.data,
             cloud_colors.Get_Num_Of_Elem() * 4 * (cloud_colors.data != nullptr));
        Clean_Tsdf_Slice(mn, mx, &rb, &volume, km.Get_Kernel("clean_tsdf_slice"));
        Clean_Color_Slice(mn, mx, &rb, &colors, km.Get_Kernel("clean_color_slice"));
        auto col1 = download<Point_Color>(&colors);
        dump(dir, "colors_cleaned.bin", col1.data(), col1.size() * 4);
        auto vol1 = download<Point_Tsdf>(&volume);
        dump(dir, "volume_cleaned.bin", vol1.data(), vol1.size() * sizeof(Point_Tsdf));
        for (auto* p : vp) delete p;
        for (auto* p : np_) delete p;
        for (auto* p : dp) delete p;
        for (auto* p : vc) delete p;
        for (auto* p : nc) delete p;
    } catch (const std::exception& e) {
        fprintf(stderr, "dropin_check: %s\n", e.what());
        return 1;
    }
    printf("dropin_check ok\n");
    return 0;
}
