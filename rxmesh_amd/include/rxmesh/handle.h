// Element handles: (patch id << 32) | local id, as in the reference
// (include/rxmesh/handle.h:19-26). A handle always names the OWNER's copy of
// an element when produced by a query iterator.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "rxmesh/types.h"

namespace rxmesh {

namespace detail {
__host__ __device__ constexpr uint64_t unique_id(uint16_t local_id, uint32_t patch_id)
{
    return (uint64_t(patch_id) << 32) | uint64_t(local_id);
}
}  // namespace detail

#define RXM_LOCAL_TYPE(NAME)                                          \
    struct NAME                                                       \
    {                                                                 \
        __host__ __device__ constexpr NAME() : id(INVALID16) {}       \
        __host__ __device__ constexpr NAME(uint16_t i) : id(i) {}     \
        uint16_t id;                                                  \
    };
RXM_LOCAL_TYPE(LocalVertexT)
RXM_LOCAL_TYPE(LocalEdgeT)
RXM_LOCAL_TYPE(LocalFaceT)
#undef RXM_LOCAL_TYPE

#define RXM_HANDLE_TYPE(NAME, LOCAL)                                                                         \
    struct NAME                                                                                              \
    {                                                                                                        \
        using LocalT = LOCAL;                                                                                \
        using Handle = NAME;                                                                                 \
        __host__ __device__ constexpr NAME() : m_handle(INVALID64) {}                                        \
        __host__ __device__ constexpr explicit NAME(uint64_t h) : m_handle(h) {}                             \
        __host__ __device__ constexpr NAME(uint32_t patch_id, LOCAL local) :                                 \
            m_handle(detail::unique_id(local.id, patch_id)) {}                                               \
        __host__ __device__ constexpr bool operator==(const NAME& o) const { return m_handle == o.m_handle; } \
        __host__ __device__ constexpr bool operator!=(const NAME& o) const { return m_handle != o.m_handle; } \
        __host__ __device__ constexpr bool operator<(const NAME& o) const { return m_handle < o.m_handle; }   \
        __host__ __device__ constexpr bool is_valid() const { return m_handle != INVALID64; }               \
        __host__ __device__ constexpr uint64_t unique_id() const { return m_handle; }                        \
        __host__ __device__ std::pair<uint32_t, uint16_t> unpack() const                                     \
        {                                                                                                    \
            return {uint32_t(m_handle >> 32), uint16_t(m_handle & 0xFFFFu)};                                 \
        }                                                                                                    \
        __host__ __device__ constexpr uint32_t patch_id() const { return uint32_t(m_handle >> 32); }         \
        __host__ __device__ constexpr uint16_t local_id() const { return uint16_t(m_handle & 0xFFFFu); }     \
        uint64_t m_handle;                                                                                   \
    };
RXM_HANDLE_TYPE(VertexHandle, LocalVertexT)
RXM_HANDLE_TYPE(EdgeHandle, LocalEdgeT)
RXM_HANDLE_TYPE(FaceHandle, LocalFaceT)
#undef RXM_HANDLE_TYPE

}  // namespace rxmesh
