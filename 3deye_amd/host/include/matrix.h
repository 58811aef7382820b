// matrix.h — host container of the reference (common/inc/matrix.h:13-79,
// impl/matrix.hpp): the subset the hot path uses (allocation, element access,
// extend_from for appending extracted slices).
#ifndef KFX_DROPIN_MATRIX
#define KFX_DROPIN_MATRIX

#include <stdlib.h>
#include <string.h>

#include <stdexcept>

#include "dim3.h"

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
