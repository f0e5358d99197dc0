// Attribute<T, HandleT>: per-element values of a mesh (reference
// include/rxmesh/attribute.h:56-670). Values are stored once per element, at
// the owner's linear id (prefix[p] + local), attribute-major ("SoA":
// data[attr * n + linear id], which is also how AoSoA requests are stored)
// or element-major for AoS (data[linear id * num_attr + attr]). A handle to
// a ribbon (not-owned) copy resolves to the owner's slot through the patch's
// owner table, so every copy of an element reads and writes one value.
//
// An Attribute is a trivially copyable view (raw host and device pointers)
// passed by value into kernels; RXMeshStatic owns the memory and releases it.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "rxmesh/handle.h"
#include "rxmesh/kernels/for_each.h"
#include "rxmesh/types.h"

namespace rxmesh {

namespace detail {
// handle -> linear id of the owner's copy, for one element kind
struct AttrIndex
{
    uint32_t        n          = 0;        // number of elements
    const uint32_t* d_prefix   = nullptr;  // P + 1 owned-prefix (device)
    const uint32_t* d_slab_off = nullptr;  // P + 1 slab offsets (device)
    const uint32_t* d_lin      = nullptr;  // slab: owner linear id of every local element (device)
    const uint32_t* h_prefix   = nullptr;  // same, host
    const uint32_t* h_slab_off = nullptr;
    const uint32_t* h_lin      = nullptr;

    __host__ __device__ uint32_t linear(uint32_t p, uint16_t l) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        const uint32_t b = d_prefix[p];
        return (b + l < d_prefix[p + 1]) ? b + l : d_lin[d_slab_off[p] + l];
#else
        const uint32_t b = h_prefix[p];
        return (b + l < h_prefix[p + 1]) ? b + l : h_lin[h_slab_off[p] + l];
#endif
    }
};

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace detail

template <class T, class HandleT>
class Attribute
{
   public:
    using Type       = T;
    using HandleType = HandleT;

    __host__ __device__ Attribute() {}

    Attribute(const char* name, uint32_t num_attributes, locationT location, layoutT layout,
              const detail::AttrIndex& index)
        : m_index(index), m_num_attributes(num_attributes), m_layout(layout)
    {
        std::strncpy(m_name, name, sizeof(m_name) - 1);
        m_name[sizeof(m_name) - 1] = '\0';
        allocate(location);
    }

    const char* get_name() const { return m_name; }
    __host__ __device__ uint32_t  get_num_attributes() const { return m_num_attributes; }
    __host__ __device__ locationT get_allocated() const { return m_allocated; }
    __host__ __device__ layoutT   get_layout() const { return m_layout; }
    __host__ __device__ bool      is_device_allocated() const { return (m_allocated & DEVICE) == DEVICE; }
    __host__ __device__ bool      is_host_allocated() const { return (m_allocated & HOST) == HOST; }
    __host__ __device__ bool      is_empty() const { return m_index.n == 0; }
    __host__ __device__ T*        data(locationT location = DEVICE) const { return location == HOST ? m_h : m_d; }

    // elements x attributes, linear (host) view
    size_t   rows() const { return m_index.n; }
    size_t   cols() const { return m_num_attributes; }
    uint32_t size() const { return m_index.n; }
    size_t   get_total_bytes() const { return size_t(m_index.n) * m_num_attributes * sizeof(T); }

    // storage index of (linear id, attr)
    __host__ __device__ size_t slot(uint32_t lin, uint32_t attr) const
    {
        return (m_layout == AoS) ? size_t(lin) * m_num_attributes + attr : size_t(attr) * m_index.n + lin;
    }

    __host__ __device__ T& operator()(const uint32_t p_id, const uint16_t local_id, const uint32_t attr) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return m_d[slot(m_index.linear(p_id, local_id), attr)];
#else
        return m_h[slot(m_index.linear(p_id, local_id), attr)];
#endif
    }
    __host__ __device__ T& operator()(const HandleT handle, const uint32_t attr = 0) const
    {
        return (*this)(handle.patch_id(), handle.local_id(), attr);
    }
    // host linear accessor: row = linear id, column = attribute
    T& operator()(size_t i, size_t j = 0) const { return m_h[slot(static_cast<uint32_t>(i), static_cast<uint32_t>(j))]; }

    template <int N>
    __host__ __device__ vec<T, N> to_glm(const HandleT& handle) const
    {
        const uint32_t lin = m_index.linear(handle.patch_id(), handle.local_id());
        vec<T, N>      r;
#pragma unroll
        for (int i = 0; i < N; ++i) {
#if defined(__HIP_DEVICE_COMPILE__)
            r[i] = m_d[slot(lin, i)];
#else
            r[i] = m_h[slot(lin, i)];
#endif
        }
        return r;
    }
    template <int N>
    __host__ __device__ void from_glm(const HandleT& handle, const vec<T, N>& in) const
    {
        const uint32_t lin = m_index.linear(handle.patch_id(), handle.local_id());
#pragma unroll
        for (int i = 0; i < N; ++i) {
#if defined(__HIP_DEVICE_COMPILE__)
            m_d[slot(lin, i)] = in[i];
#else
            m_h[slot(lin, i)] = in[i];
#endif
        }
    }

    // set every value at `location` (reference attribute.h:271); the device
    // side is cleared with a fill kernel, so the call can sit in a captured graph
    void reset(const T value, locationT location, hipStream_t stream = nullptr)
    {
        const size_t n = size_t(m_index.n) * m_num_attributes;
        if ((location & HOST) && m_h)
            for (size_t i = 0; i < n; ++i) m_h[i] = value;
        if ((location & DEVICE) && m_d && n) {
            const uint32_t blocks = static_cast<uint32_t>(std::min<size_t>((n + 255) / 256, 8192));
            detail::fill_kernel<T><<<blocks, 256, 0, stream>>>(m_d, n, value);
            detail::hip_check(hipGetLastError(), "Attribute::reset");
        }
    }

    // copy the values from `source` to `target` (HOST <-> DEVICE), allocating
    // the target side if needed (reference attribute.h:283)
    void move(locationT source, locationT target, hipStream_t stream = nullptr)
    {
        if (source == target) return;
        if (!(m_allocated & source)) throw std::runtime_error("Attribute::move: source not allocated");
        if (!(m_allocated & target)) allocate(target);
        const size_t bytes = get_total_bytes();
        if (!bytes) return;
        if (source == HOST && target == DEVICE)
            detail::hip_check(hipMemcpyAsync(m_d, m_h, bytes, hipMemcpyHostToDevice, stream), "Attribute::move");
        else if (source == DEVICE && target == HOST)
            detail::hip_check(hipMemcpyAsync(m_h, m_d, bytes, hipMemcpyDeviceToHost, stream), "Attribute::move");
        detail::hip_check(hipStreamSynchronize(stream), "Attribute::move");
    }

    // copy values of another attribute of the same shape (reference attribute.h:302)
    void copy_from(Attribute<T, HandleT>& source, locationT source_flag, locationT dst_flag,
                   hipStream_t stream = nullptr)
    {
        if (source.m_index.n != m_index.n || source.m_num_attributes != m_num_attributes ||
            source.m_layout != m_layout)
            throw std::runtime_error("Attribute::copy_from: shape or layout mismatch");
        const size_t bytes = get_total_bytes();
        if (!bytes) return;
        T*         dst = (dst_flag == HOST) ? m_h : m_d;
        const T*   src = (source_flag == HOST) ? source.m_h : source.m_d;
        hipMemcpyKind kind = (source_flag == HOST) ? ((dst_flag == HOST) ? hipMemcpyHostToHost : hipMemcpyHostToDevice)
                                                   : ((dst_flag == HOST) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
        if (!dst || !src) throw std::runtime_error("Attribute::copy_from: side not allocated");
        detail::hip_check(hipMemcpyAsync(dst, src, bytes, kind, stream), "Attribute::copy_from");
        detail::hip_check(hipStreamSynchronize(stream), "Attribute::copy_from");
    }

    void release(locationT location = LOCATION_ALL)
    {
        if ((location & HOST) && m_h) {
            std::free(m_h);
            m_h = nullptr;
            m_allocated &= ~HOST;
        }
        if ((location & DEVICE) && m_d) {
            (void)hipFree(m_d);
            m_d = nullptr;
            m_allocated &= ~DEVICE;
        }
    }

    const detail::AttrIndex& index() const { return m_index; }

   private:
    void allocate(locationT location)
    {
        const size_t bytes = get_total_bytes();
        if ((location & HOST) && !m_h) {
            m_h = static_cast<T*>(std::calloc(bytes ? bytes : 1, 1));
            if (!m_h) throw std::runtime_error("Attribute: host allocation failed");
            m_allocated |= HOST;
        }
        if ((location & DEVICE) && !m_d) {
            detail::hip_check(hipMalloc(&m_d, bytes ? bytes : 16), "Attribute: hipMalloc");
            m_allocated |= DEVICE;
        }
    }

    detail::AttrIndex m_index{};
    T*                m_h              = nullptr;
    T*                m_d              = nullptr;
    uint32_t          m_num_attributes = 0;
    locationT         m_allocated      = LOCATION_NONE;
    layoutT           m_layout         = SoA;
    char              m_name[64]       = {0};
};

template <class T>
using VertexAttribute = Attribute<T, VertexHandle>;
template <class T>
using EdgeAttribute = Attribute<T, EdgeHandle>;
template <class T>
using FaceAttribute = Attribute<T, FaceHandle>;

}  // namespace rxmesh
