// Iterator over one source element's query result (reference
// include/rxmesh/iterator.cuh:51-194). operator[] returns the handle of the
// OWNER's copy of the element: owned locals (l < owned count) map directly;
// ribbon locals go through the patch's dense owner tables -- owner patch q
// and owner linear id g give the handle (q, g - prefix[q]) -- instead of the
// reference's cuckoo hash table and patch stash.
#pragma once
#include "rxmesh/handle.h"

namespace rxmesh {

template <typename HandleT>
struct Iterator
{
    using LocalT = typename HandleT::LocalT;
    using Handle = HandleT;

    __device__ Iterator() = default;
    // empty iterator of a non-participating source (reference Iterator(local_id, patch_id))
    __device__ Iterator(uint16_t, uint32_t patch) : m_patch(patch) {}
    __device__ Iterator(const uint16_t* row, uint16_t size, uint32_t patch, uint32_t n_owned,
                        const uint32_t* prefix, const uint32_t* lin, const uint32_t* own)
        : m_row(row), m_size(size), m_patch(patch), m_owned(n_owned), m_prefix(prefix), m_lin(lin), m_own(own)
    {
    }

    __device__ uint16_t size() const { return m_size; }

    __device__ HandleT operator[](const uint16_t i) const
    {
        if (i >= m_size) return HandleT();
        const uint16_t l = m_row[i];
        if (l == INVALID16) return HandleT();  // e.g. the missing side of an EVDiamond row
        if (l < m_owned) return HandleT(m_patch, LocalT(l));
        const uint32_t q = m_own[l];
        return HandleT(q, LocalT(static_cast<uint16_t>(m_lin[l] - m_prefix[q])));
    }
    // local id of entry i inside the querying patch (reference Iterator::local)
    __device__ uint16_t local(const uint16_t i) const { return (i < m_size) ? m_row[i] : INVALID16; }
    __device__ HandleT  front() const { return (*this)[0]; }
    __device__ HandleT  back() const { return (*this)[m_size - 1]; }

   private:
    const uint16_t* m_row    = nullptr;
    uint16_t        m_size   = 0;
    uint32_t        m_patch  = INVALID32;
    uint32_t        m_owned  = 0;
    const uint32_t* m_prefix = nullptr;
    const uint32_t* m_lin    = nullptr;
    const uint32_t* m_own    = nullptr;
};

using VertexIterator = Iterator<VertexHandle>;
using EdgeIterator   = Iterator<EdgeHandle>;
using FaceIterator   = Iterator<FaceHandle>;

}  // namespace rxmesh
