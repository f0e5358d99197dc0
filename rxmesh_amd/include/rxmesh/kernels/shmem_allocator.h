// Bump allocator over the kernel's dynamic LDS (the reference's
// ShmemAllocator, include/rxmesh/kernels/shmem_allocator.cuh:15-124).
// Allocations are 16-byte aligned and released in LIFO order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rxmesh {

struct ShmemAllocator
{
    __device__ ShmemAllocator() : m_ptr(nullptr), m_bytes(0)
    {
        extern __shared__ uint4 rxm_dyn_lds[];
        m_ptr = reinterpret_cast<char*>(rxm_dyn_lds);
    }
    template <typename T>
    __device__ T* alloc(uint32_t count)
    {
        m_bytes = (m_bytes + 15u) & ~15u;
        T* p    = reinterpret_cast<T*>(m_ptr + m_bytes);
        m_bytes += count * static_cast<uint32_t>(sizeof(T));
        return p;
    }
    __device__ char* alloc(uint32_t bytes) { return alloc<char>(bytes); }
    __device__ void dealloc(uint32_t bytes) { m_bytes = (bytes > m_bytes) ? 0u : m_bytes - bytes; }
    template <typename T>
    __device__ void dealloc(uint32_t count)
    {
        dealloc(count * static_cast<uint32_t>(sizeof(T)));
    }
    __device__ void reset_to(uint32_t bytes) { m_bytes = bytes; }
    __device__ uint32_t get_allocated_size_bytes() const { return m_bytes; }
    __device__ char* get_ptr() const { return m_ptr; }

   private:
    char*    m_ptr;
    uint32_t m_bytes;
};

}  // namespace rxmesh
