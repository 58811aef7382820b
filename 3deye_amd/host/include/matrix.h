// matrix.h — host container of the reference (common/inc/matrix.h:13-79,
// impl/matrix.hpp): the subset the hot path uses (allocation, element access,
// extend_from for appending extracted slices, Save_To_File for writing them).
#ifndef KFX_DROPIN_MATRIX
#define KFX_DROPIN_MATRIX

#include <stdlib.h>
#include <string.h>

#include <stdexcept>
#include <string>
#include <typeinfo>

#include "base_point_types.h"
#include "dim3.h"
#include "kfx.h"

// Element writers of Save_To_File (the slices are matrices of these).
inline int kfx_save_matrix_text(const char* path, const Point_XYZI* d, size_t n) {
    return kfx_save_points_text(path, reinterpret_cast<const kfx_xyzi*>(d), n);
}
inline int kfx_save_matrix_text(const char* path, const Point_Color* d, size_t n) {
    return kfx_save_colors_text(path, reinterpret_cast<const kfx_color*>(d), n);
}

template <typename T>
class matrix {
    dim3 sizes;

  public:
    T* data = nullptr;

    matrix() : sizes(0, 1, 1) {}
    explicit matrix(size_t width, size_t height = 1, size_t depth = 1) : sizes(width, height, depth) {
        allocate();
    }
    explicit matrix(const dim3 dims) : sizes(dims) { allocate(); }
    matrix(const matrix<T>& o) : sizes(o.sizes) {
        allocate();
        if (o.data) memcpy(data, o.data, sizeof(T) * Get_Num_Of_Elem());
    }
    matrix& operator=(const matrix<T>& o) = delete;
    ~matrix() { free_buffer(); }

    dim3 Get_Dimensions() const { return sizes; }
    size_t Get_Width() const { return sizes.width; }
    size_t Get_Height() const { return sizes.height; }
    size_t Get_Depth() const { return sizes.depth; }
    size_t Get_Num_Of_Elem() const { return sizes.width * sizes.height * sizes.depth; }
    bool Is_Empty() const { return Get_Num_Of_Elem() == 0 || data == nullptr; }
    T& operator[](const size_t i) { return data[i]; }
    T& operator[](const size_t i) const { return data[i]; }

    void free_buffer() {
        free(data);
        data = nullptr;
    }

    // The compiler's name of the element storage type (matrix.hpp:443-447).
    std::string Get_Matrix_Type() const { return typeid(data).name(); }

    // matrix.hpp:455-476: text dump to custom_file_name, or to
    // "<w>x<h>x<d>-<type>.tst" when it is empty.  1-D matrices of points or
    // colours (the extracted slices); other shapes and element types have
    // no writer here.
    void Save_To_File(const std::string& custom_file_name = "") const {
        if (sizes.height * sizes.depth != 1) throw std::runtime_error("matrix::Save_To_File: 1-D matrices only");
        std::string name = custom_file_name;
        if (name.empty())
            name = std::to_string(sizes.width) + "x" + std::to_string(sizes.height) + "x" +
                   std::to_string(sizes.depth) + "-" + Get_Matrix_Type() + ".tst";
        if (kfx_save_matrix_text(name.c_str(), data, Get_Num_Of_Elem() * (data != nullptr)) != KFX_OK)
            throw std::runtime_error("matrix::Save_To_File: cannot write " + name);
    }

    // Append another matrix's elements (the result is 1-D).
    void extend_from(const matrix<T>& extra) {
        const size_t n0 = Get_Num_Of_Elem() * (data != nullptr), n1 = extra.Get_Num_Of_Elem();
        if (n1 == 0) return;
        T* p = static_cast<T*>(realloc(data, sizeof(T) * (n0 + n1)));
        if (!p) throw std::runtime_error("matrix::extend_from: out of memory");
        memcpy(p + n0, extra.data, sizeof(T) * n1);
        data = p;
        sizes = dim3(n0 + n1, 1, 1);
    }

  private:
    void allocate() {
        const size_t n = Get_Num_Of_Elem();
        data = n ? static_cast<T*>(calloc(n, sizeof(T))) : nullptr;
        if (n && !data) throw std::runtime_error("matrix: out of memory");
    }
};

#endif
