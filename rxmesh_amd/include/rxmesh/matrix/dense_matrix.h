// DenseMatrix<T>: the dense matrix of the reference
// (include/rxmesh/matrix/dense_matrix.h), reduced to what app kernels use:
// num_rows x num_cols values in column-major order on the host and the
// device, element access by (row, col) on either side and by mesh handle on
// the device (row = the element's linear id, the row the reference's
// get_row_id gives), reset(), move() and release(). A 3-column matrix over
// the vertices is the library's SoA layout [3][V] over linear ids. Passed to
// kernels by value; copies share the arrays, which release() frees once.
#pragma once
#include <hip/hip_runtime.h>

#include <cassert>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "rxmesh/rxmesh_static.h"

namespace rxmesh {

template <typename T>
struct DenseMatrix
{
    using IndexT = uint32_t;

    DenseMatrix() = default;

    // a matrix whose rows are the mesh's vertices, edges or faces: handles
    // index its rows
    explicit DenseMatrix(const RXMeshStatic& rx, IndexT num_rows, IndexT num_cols, locationT location = LOCATION_ALL)
        : m_ctx(rx.get_context()), m_rows(num_rows), m_cols(num_cols)
    {
        allocate(location);
    }

    // a matrix not tied to the mesh: (row, col) access only
    explicit DenseMatrix(IndexT num_rows, IndexT num_cols, locationT location = LOCATION_ALL)
        : m_rows(num_rows), m_cols(num_cols)
    {
        allocate(location);
    }

    __host__ __device__ T& operator()(IndexT row, IndexT col = 0) const
    {
        assert(row < m_rows && col < m_cols);
#if defined(__HIP_DEVICE_COMPILE__)
        return m_d_val[size_t(col) * m_rows + row];
#else
        return m_h_val[size_t(col) * m_rows + row];
#endif
    }

    template <typename HandleT, typename = std::enable_if_t<!std::is_integral_v<HandleT>>>
    __device__ T& operator()(const HandleT& handle, IndexT col = 0) const
    {
        return (*this)(m_ctx.linear_id(handle), col);
    }

    void reset(T value, locationT location, hipStream_t stream = nullptr)
    {
        if ((location & HOST) && m_h_val)
            for (size_t i = 0; i < size(); ++i) m_h_val[i] = value;
        if ((location & DEVICE) && m_d_val) {
            std::vector<T> v(size(), value);
            check(hipMemcpyAsync(m_d_val, v.data(), sizeof(T) * size(), hipMemcpyHostToDevice, stream));
            check(hipStreamSynchronize(stream));
        }
    }

    void move(locationT from, locationT to, hipStream_t stream = nullptr)
    {
        if (from == to) return;
        allocate(to);
        if (from == DEVICE && to == HOST)
            check(hipMemcpyAsync(m_h_val, m_d_val, sizeof(T) * size(), hipMemcpyDeviceToHost, stream));
        else if (from == HOST && to == DEVICE)
            check(hipMemcpyAsync(m_d_val, m_h_val, sizeof(T) * size(), hipMemcpyHostToDevice, stream));
        check(hipStreamSynchronize(stream));
    }

    void release(locationT location = LOCATION_ALL)
    {
        if ((location & DEVICE) && m_d_val) {
            (void)hipFree(m_d_val);
            m_d_val = nullptr;
        }
        if ((location & HOST) && m_h_val) {
            delete[] m_h_val;
            m_h_val = nullptr;
        }
    }

    __host__ __device__ IndexT rows() const { return m_rows; }
    __host__ __device__ IndexT cols() const { return m_cols; }
    T*                         data(locationT location = DEVICE) const { return location == DEVICE ? m_d_val : m_h_val; }

   private:
    size_t size() const { return size_t(m_rows) * m_cols; }

    void allocate(locationT location)
    {
        if ((location & DEVICE) && !m_d_val) check(hipMalloc(&m_d_val, sizeof(T) * (size() ? size() : 1)));
        if ((location & HOST) && !m_h_val) m_h_val = new T[size() ? size() : 1]();
    }

    static void check(hipError_t e)
    {
        if (e != hipSuccess) throw std::runtime_error("DenseMatrix: HIP call failed");
    }

    Context m_ctx{};
    IndexT  m_rows = 0, m_cols = 0;
    T*      m_d_val = nullptr;
    T*      m_h_val = nullptr;
};

}  // namespace rxmesh
