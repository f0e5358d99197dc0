// SparseMatrix<T>: the vertex-by-vertex sparse matrix of the reference
// (include/rxmesh/matrix/sparse_matrix.h), reduced to what app kernels use:
// a CSR over linear vertex ids whose row v holds the diagonal, then v's VV
// neighbours in VV query order (the pattern the reference builds,
// sparse_matrix_kernels.cuh:75-140), a device operator()(row, col) that
// finds an entry by searching its row, reset(), move() and release().
// A lookup of an entry outside the pattern returns a sink element past the
// last entry (the reference asserts there) so a kernel never reads or
// writes outside the value array. The object is passed to kernels by value
// (it is trivially copyable: the host arrays sit behind a pointer); copies
// share the arrays, which release() frees once, as in the reference.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <vector>

#include "rxmesh/matrix/dense_matrix.h"
#include "rxmesh/rxmesh_static.h"

namespace rxmesh {

template <typename T>
struct SparseMatrix
{
    SparseMatrix() = default;

    explicit SparseMatrix(RXMeshStatic& rx) : m_ctx(rx.get_context()), m_n(rx.get_num_vertices())
    {
        // VV rows through the C-ABI, then [diagonal, row] per vertex
        rxm_mesh* mh = rx.get_c_handle();
        int64_t   nvv = 0;
        check(rxm_query_nnz(mh, RXM_OP_VV, &nvv));
        int32_t *d_rp = nullptr, *d_col = nullptr;
        check_hip(hipMalloc(&d_rp, sizeof(int32_t) * (m_n + 1)));
        check_hip(hipMalloc(&d_col, sizeof(int32_t) * (nvv > 0 ? nvv : 1)));
        check(rxm_query_csr(mh, RXM_OP_VV, 0, d_rp, d_col, nullptr));
        std::vector<int32_t> rp(m_n + 1), col(nvv);
        check_hip(hipMemcpy(rp.data(), d_rp, sizeof(int32_t) * (m_n + 1), hipMemcpyDeviceToHost));
        if (nvv) check_hip(hipMemcpy(col.data(), d_col, sizeof(int32_t) * nvv, hipMemcpyDeviceToHost));
        (void)hipFree(d_rp);
        (void)hipFree(d_col);
        m_nnz  = static_cast<uint32_t>(nvv) + m_n;
        m_host = new HostPart();
        auto& h = *m_host;
        h.row_ptr.resize(m_n + 1);
        h.col.resize(m_nnz);
        for (uint32_t v = 0, k = 0; v < m_n; ++v) {
            h.row_ptr[v] = static_cast<int32_t>(k);
            h.col[k++]   = static_cast<int32_t>(v);
            for (int32_t j = rp[v]; j < rp[v + 1]; ++j) h.col[k++] = col[j];
        }
        h.row_ptr[m_n] = static_cast<int32_t>(m_nnz);
        h.val.assign(m_nnz + 1, T(0));
        check_hip(hipMalloc(&m_d_row_ptr, sizeof(int32_t) * (m_n + 1)));
        check_hip(hipMalloc(&m_d_col, sizeof(int32_t) * m_nnz));
        check_hip(hipMalloc(&m_d_val, sizeof(T) * (m_nnz + 1)));
        check_hip(hipMemcpy(m_d_row_ptr, h.row_ptr.data(), sizeof(int32_t) * (m_n + 1), hipMemcpyHostToDevice));
        check_hip(hipMemcpy(m_d_col, h.col.data(), sizeof(int32_t) * m_nnz, hipMemcpyHostToDevice));
        check_hip(hipMemcpy(m_d_val, h.val.data(), sizeof(T) * (m_nnz + 1), hipMemcpyHostToDevice));
    }

    // entry (row, col) by owner handles; the sink element if it is not in the pattern
    __device__ T& operator()(const VertexHandle& row, const VertexHandle& col) const
    {
        const uint32_t r = m_ctx.linear_id(row), c = m_ctx.linear_id(col);
        for (int32_t k = m_d_row_ptr[r]; k < m_d_row_ptr[r + 1]; ++k)
            if (m_d_col[k] == static_cast<int32_t>(c)) return m_d_val[k];
        return m_d_val[m_nnz];
    }

    // host entry by linear ids (after move(DEVICE, HOST))
    T& host(uint32_t r, uint32_t c)
    {
        auto& h = *m_host;
        for (int32_t k = h.row_ptr[r]; k < h.row_ptr[r + 1]; ++k)
            if (h.col[k] == static_cast<int32_t>(c)) return h.val[k];
        return h.val[m_nnz];
    }

    void reset(T value, locationT location, hipStream_t stream = nullptr)
    {
        if (location & HOST) std::fill(m_host->val.begin(), m_host->val.end(), value);
        if (location & DEVICE) {
            std::vector<T> v(m_nnz + 1, value);
            check_hip(hipMemcpyAsync(m_d_val, v.data(), sizeof(T) * (m_nnz + 1), hipMemcpyHostToDevice, stream));
            check_hip(hipStreamSynchronize(stream));
        }
    }

    void move(locationT from, locationT to, hipStream_t stream = nullptr)
    {
        if (from == DEVICE && (to & HOST)) {
            check_hip(hipMemcpyAsync(m_host->val.data(), m_d_val, sizeof(T) * (m_nnz + 1), hipMemcpyDeviceToHost,
                                     stream));
            check_hip(hipStreamSynchronize(stream));
        } else if (from == HOST && (to & DEVICE)) {
            check_hip(hipMemcpyAsync(m_d_val, m_host->val.data(), sizeof(T) * (m_nnz + 1), hipMemcpyHostToDevice,
                                     stream));
            check_hip(hipStreamSynchronize(stream));
        }
    }

    void release(locationT location = LOCATION_ALL)
    {
        if (location & DEVICE) {
            (void)hipFree(m_d_row_ptr);
            (void)hipFree(m_d_col);
            (void)hipFree(m_d_val);
            m_d_row_ptr = m_d_col = nullptr;
            m_d_val                = nullptr;
        }
        if (location & HOST) {
            delete m_host;
            m_host = nullptr;
        }
    }

    __host__ __device__ uint32_t rows() const { return m_n; }
    __host__ __device__ uint32_t cols() const { return m_n; }
    __host__ __device__ uint32_t non_zeros() const { return m_nnz; }
    int32_t* row_ptr(locationT location = DEVICE) { return location == DEVICE ? m_d_row_ptr : m_host->row_ptr.data(); }
    int32_t* col_idx(locationT location = DEVICE) { return location == DEVICE ? m_d_col : m_host->col.data(); }
    T*       get_val(locationT location = DEVICE) { return location == DEVICE ? m_d_val : m_host->val.data(); }

   private:
    static void check(int rc)
    {
        if (rc != 0) throw std::runtime_error("SparseMatrix: rxm call failed");
    }
    static void check_hip(hipError_t e)
    {
        if (e != hipSuccess) throw std::runtime_error("SparseMatrix: HIP call failed");
    }

    struct HostPart
    {
        std::vector<int32_t> row_ptr, col;
        std::vector<T>       val;
    };
    Context   m_ctx{};
    uint32_t  m_n = 0, m_nnz = 0;
    int32_t * m_d_row_ptr = nullptr, *m_d_col = nullptr;
    T*        m_d_val  = nullptr;
    HostPart* m_host   = nullptr;
};

}  // namespace rxmesh
