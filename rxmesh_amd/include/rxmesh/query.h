// Query<blockThreads>: the in-kernel query API of the reference
// (include/rxmesh/query.h:15-140, query.inl:1-260,
// kernels/query_dispatcher.cuh:27-266). One workgroup works on one patch:
// prologue<op>() builds the relation in LDS for the participating source
// elements (owned, or every local one with allow_not_owned, filtered by the
// user's active-set predicate), run_compute() calls the user lambda once per
// participant with an Iterator over its row, epilogue() releases the LDS.
//
// Rows are deterministic and identical to the C-ABI query (rxm_query_csr):
// VV / VE / VF / EF rows list targets in ascending patch-local source order,
// FF rows walk the face's edges in FE order, each edge's other faces in
// ascending order; oriented VV / VE rows are the canonical counter-clockwise
// one-ring. The reference's rows follow shared-memory atomic order instead.
#pragma once
#include <hip/hip_cooperative_groups.h>

#include <type_traits>

#include "lds_sort.hpp"
#include "orient.hpp"
#include "rxmesh/context.h"
#include "rxmesh/iterator.h"
#include "rxmesh/kernels/shmem_allocator.h"
#include "rxmesh/util/meta.h"

namespace rxmesh {
namespace detail {

// 0 = vertex, 1 = edge, 2 = face
__host__ __device__ constexpr int op_src_kind(Op op)
{
    return (op == Op::VV || op == Op::VE || op == Op::VF || op == Op::V) ? 0
           : (op == Op::EV || op == Op::EF || op == Op::EVDiamond || op == Op::E) ? 1
                                                                          : 2;
}
__host__ __device__ constexpr int op_dst_kind(Op op)
{
    return (op == Op::VV || op == Op::EV || op == Op::FV || op == Op::EVDiamond) ? 0
           : (op == Op::VE || op == Op::FE)                                        ? 1
                                                                                    : 2;
}
// entries per source of the fixed-width relations (reference fixed_offset,
// query_dispatcher.cuh:238-242); 0 = variable (CSR offsets in LDS)
__host__ __device__ constexpr uint32_t op_fixed_width(Op op)
{
    return (op == Op::EV) ? 2u : ((op == Op::FV || op == Op::FE) ? 3u : ((op == Op::EVDiamond) ? 4u : 0u));
}
__host__ __device__ constexpr bool op_supported(Op op)
{
    return op == Op::VV || op == Op::VE || op == Op::VF || op == Op::FV || op == Op::FE || op == Op::FF ||
           op == Op::EV || op == Op::EF || op == Op::EVDiamond;
}

template <typename HandleT>
struct HandleKind;
template <>
struct HandleKind<VertexHandle>
{
    static constexpr int value = 0;
};
template <>
struct HandleKind<EdgeHandle>
{
    static constexpr int value = 1;
};
template <>
struct HandleKind<FaceHandle>
{
    static constexpr int value = 2;
};

template <int K>
struct KindHandle;
template <>
struct KindHandle<0>
{
    using type = VertexHandle;
};
template <>
struct KindHandle<1>
{
    using type = EdgeHandle;
};
template <>
struct KindHandle<2>
{
    using type = FaceHandle;
};

// LDS carve of one query. The persistent part (participant mask, CSR
// offsets, u16 target ids) lives until epilogue(); the scratch part after it
// is only used while the relation is built. The host runs the same carve
// with a null base to size the launch.
struct QueryLDS
{
    uint32_t *mask, *off;
    uint16_t* val;
    uint32_t *ev, *slots, *wave, *gid, *eoff;
    uint16_t *fe, *hn;
    uint32_t  persist_bytes;
};

__host__ __device__ inline QueryLDS plan_query(rxm::Bump& b, Op op, uint32_t nv, uint32_t ne, uint32_t nf,
                                               bool oriented, uint32_t ef_max)
{
    QueryLDS       q{};
    const int      sk   = op_src_kind(op);
    const uint32_t nsrc = (sk == 0) ? nv : ((sk == 1) ? ne : nf);
    q.mask              = b.alloc<uint32_t>((nsrc + 31u) / 32u);
    if (op_fixed_width(op) == 0u) q.off = b.alloc<uint32_t>(nsrc + 1u);
    uint32_t nval = 0;
    switch (op) {
        case Op::VV:
        case Op::VE: nval = 2u * ne; break;
        case Op::EV: nval = 2u * rxm::r4(ne); break;  // staged as whole 16-byte chunks of EV
        case Op::FF: nval = 3u * nf * (ef_max > 2u ? ef_max - 1u : 1u); break;
        case Op::EVDiamond: nval = 4u * ne; break;
        default: nval = 3u * nf; break;
    }
    q.val           = b.alloc<uint16_t>(rxm::r4(nval));
    q.persist_bytes = b.bytes();
    switch (op) {
        case Op::VV:
        case Op::VE:
            q.ev = b.alloc<uint32_t>(rxm::r4(ne));
            q.slots = b.alloc<uint32_t>(2u * ne);
            if (oriented) {
                q.fe  = b.alloc<uint16_t>(rxm::r4(3u * nf));
                q.hn  = b.alloc<uint16_t>(2u * ne);
                q.gid = b.alloc<uint32_t>(nv);
            }
            break;
        case Op::VF:
            q.ev    = b.alloc<uint32_t>(rxm::r4(ne));
            q.fe    = b.alloc<uint16_t>(rxm::r4(3u * nf));
            q.slots = b.alloc<uint32_t>(3u * nf);
            break;
        case Op::FV:
            q.ev = b.alloc<uint32_t>(rxm::r4(ne));
            q.fe = b.alloc<uint16_t>(rxm::r4(3u * nf));
            break;
        case Op::EF:
            q.fe    = b.alloc<uint16_t>(rxm::r4(3u * nf));
            q.slots = b.alloc<uint32_t>(3u * nf);
            break;
        case Op::FF:
            q.fe    = b.alloc<uint16_t>(rxm::r4(3u * nf));
            q.eoff  = b.alloc<uint32_t>(ne + 1u);
            q.slots = b.alloc<uint32_t>(3u * nf);
            break;
        case Op::EVDiamond:
            q.ev    = b.alloc<uint32_t>(rxm::r4(ne));
            q.fe    = b.alloc<uint16_t>(rxm::r4(3u * nf));
            q.slots = b.alloc<uint32_t>(2u * ne);
            break;
        default: break;
    }
    q.wave = b.alloc<uint32_t>(20);
    return q;
}

// LDS bytes one query needs on the largest patch (plus 16 bytes of alignment slack)
inline size_t query_shmem_bytes(Op op, uint32_t nv, uint32_t ne, uint32_t nf, bool oriented, uint32_t ef_max)
{
    rxm::Bump b(nullptr);
    plan_query(b, op, nv, ne, nf, oriented, ef_max);
    return b.bytes() + 16u;
}

}  // namespace detail

template <uint32_t blockThreads>
struct Query
{
    static_assert(blockThreads % 64u == 0u && blockThreads <= 1024u && (blockThreads & (blockThreads - 1u)) == 0u,
                  "blockThreads must be a power of two multiple of the 64-wide wavefront, at most 1024");

    __device__ Query(const Query&)            = delete;
    __device__ Query& operator=(const Query&) = delete;

    __device__ Query(const Context& context, const uint32_t pid = blockIdx.x)
        : m_ctx(context), m_pid(pid < context.m.P ? pid : INVALID32)
    {
        if (m_pid != INVALID32) m_desc = rxm::load_desc(m_ctx.m.patch, m_pid);
    }

    __device__ int get_patch_id() const { return static_cast<int>(m_pid); }

    // per-patch counts (incl. ribbon) and owned counts
    __device__ const rxm::PatchDesc& get_patch_info() const { return m_desc; }

    // valence of every local vertex in LDS (reference query.inl:37-66)
    __device__ void compute_vertex_valence(cooperative_groups::thread_block& block, ShmemAllocator& shrd_alloc)
    {
        if (m_pid == INVALID32) return;
        m_s_valence = shrd_alloc.alloc<uint32_t>(m_desc.nv);
        for (uint32_t v = threadIdx.x; v < m_desc.nv; v += blockThreads) m_s_valence[v] = 0;
        block.sync();
        for (uint32_t e = threadIdx.x; e < m_desc.ne; e += blockThreads) {
            const uint32_t x = m_ctx.m.ev[m_desc.e_off + e];
            atomicAdd(m_s_valence + (x & 0xFFFFu), 1u);
            atomicAdd(m_s_valence + (x >> 16), 1u);
        }
        block.sync();
    }
    __device__ uint16_t vertex_valence(uint16_t v) const { return static_cast<uint16_t>(m_s_valence[v]); }
    __device__ uint16_t vertex_valence(VertexHandle vh) const
    {
        return static_cast<uint16_t>(m_s_valence[vh.local_id()]);
    }

    template <Op op, typename computeT>
    __device__ void dispatch(cooperative_groups::thread_block& block, ShmemAllocator& shrd_alloc, computeT compute_op,
                             const bool oriented = false)
    {
        using HandleT = typename detail::FunctionTraits<computeT>::template arg<0>::type;
        dispatch<op>(block, shrd_alloc, compute_op, [](HandleT) { return true; }, oriented);
    }

    template <Op op, typename computeT, typename activeSetT>
    __device__ void dispatch(cooperative_groups::thread_block& block, ShmemAllocator& shrd_alloc, computeT compute_op,
                             activeSetT compute_active_set, const bool oriented = false,
                             const bool allow_not_owned = false)
    {
        if (m_pid == INVALID32) return;
        using ComputeTraits = detail::FunctionTraits<computeT>;
        using HandleT       = typename ComputeTraits::template arg<0>::type;
        using ActiveHandleT = typename detail::FunctionTraits<activeSetT>::template arg<0>::type;
        static_assert(std::is_same_v<HandleT, ActiveHandleT>,
                      "the compute lambda and the active-set lambda must take the same handle type");
        static_assert(detail::HandleKind<HandleT>::value == detail::op_src_kind(op),
                      "the compute lambda's handle type does not match the query's source element");
        prologue<op>(block, shrd_alloc, compute_active_set, oriented, allow_not_owned);
        run_compute(block, compute_op);
        epilogue(block, shrd_alloc);
    }

    template <Op op>
    __device__ void prologue(cooperative_groups::thread_block& block, ShmemAllocator& shrd_alloc,
                             const bool oriented = false, const bool allow_not_owned = true)
    {
        using SrcHandleT = typename detail::KindHandle<detail::op_src_kind(op)>::type;
        prologue<op>(block, shrd_alloc, [](SrcHandleT) { return true; }, oriented, allow_not_owned);
    }

    template <Op op, typename activeSetT>
    __device__ void prologue(cooperative_groups::thread_block& block, ShmemAllocator& shrd_alloc,
                             activeSetT compute_active_set, const bool oriented = false,
                             const bool allow_not_owned = true)
    {
        static_assert(detail::op_supported(op), "unsupported query operation");
        if (m_pid == INVALID32) return;
        using SrcHandleT         = typename detail::KindHandle<detail::op_src_kind(op)>::type;
        constexpr uint32_t T     = blockThreads;
        const rxm::PatchDesc& d  = m_desc;
        const rxm::DevMesh&   m  = m_ctx.m;
        block.sync();  // LDS of a previous query may still be read
        m_op           = op;
        m_fixed        = detail::op_fixed_width(op);
        m_shmem_before = shrd_alloc.get_allocated_size_bytes();
        const uint32_t start = (m_shmem_before + 15u) & ~15u;
        rxm::Bump      b(shrd_alloc.get_ptr() + start);
        const detail::QueryLDS L =
            detail::plan_query(b, op, d.nv, d.ne, d.nf, oriented, m_ctx.max_edge_faces);
        shrd_alloc.reset_to(start + L.persist_bytes);
        m_mask = L.mask;
        m_off  = L.off;
        m_val  = L.val;

        constexpr int  sk   = detail::op_src_kind(op);
        const uint32_t nsrc = (sk == 0) ? d.nv : ((sk == 1) ? d.ne : d.nf);
        const uint32_t nown = (sk == 0) ? d.nvo : ((sk == 1) ? d.neo : d.nfo);
        const uint32_t rows = allow_not_owned ? nsrc : nown;

        // participants: (owned or allowed) and in the user's active set
        for (uint32_t w = threadIdx.x; w < (nsrc + 31u) / 32u; w += T) m_mask[w] = 0;
        block.sync();
        int any = 0;
        for (uint32_t s = threadIdx.x; s < rows; s += T) {
            if (compute_active_set(SrcHandleT(m_pid, static_cast<uint16_t>(s)))) {
                atomicOr(&m_mask[s >> 5], 1u << (s & 31u));
                any = 1;
            }
        }
        if (__syncthreads_or(any) == 0) {
            m_num_src = 0;
            return;
        }
        m_num_src = nsrc;

        if constexpr (op == Op::VV || op == Op::VE) {
            rxm::stage_u32(L.ev, m.ev + d.e_off, d.ne);
            auto key = [&](uint32_t i) {
                const uint32_t x = L.ev[i >> 1];
                return (i & 1u) ? rxm::hi16(x) : rxm::lo16(x);
            };
            if (oriented) {
                // half-edge successor table + edge columns, then each
                // participating row ordered into its canonical CCW ring
                rxm::stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
                for (uint32_t v = threadIdx.x; v < d.nv; v += T) L.gid[v] = m.ltog_v[d.v_off + v];
                for (uint32_t i = threadIdx.x; i < d.ne; i += T) reinterpret_cast<uint32_t*>(L.hn)[i] = 0xFFFFFFFFu;
                block.sync();
                rxm::build_half_next(L.fe, L.ev, d.nf, L.hn);
                rxm::fill_columns<T>(2u * d.ne, d.nv, key, [&](uint32_t i) { return i >> 1; }, m_off, L.slots,
                                     L.wave);
                for (uint32_t v = threadIdx.x; v < rows; v += T)
                    rxm::orient_ring<12, op == Op::VV>(v, m_off[v], m_off[v + 1], L.slots, m_val, L.ev, L.hn,
                                                       L.gid);
            } else {
                block.sync();
                rxm::fill_columns<T>(
                    2u * d.ne, d.nv, key,
                    [&](uint32_t i) {
                        const uint32_t x = L.ev[i >> 1];
                        return ((i >> 1) << 16) | ((i & 1u) ? rxm::lo16(x) : rxm::hi16(x));
                    },
                    m_off, L.slots, L.wave);
                for (uint32_t v = threadIdx.x; v < rows; v += T) {
                    const uint32_t o = m_off[v];
                    rxm::sort_row(L.slots + o, m_off[v + 1] - o, [&](uint32_t j, uint32_t s) {
                        m_val[o + j] = static_cast<uint16_t>(op == Op::VV ? rxm::lo16(s) : rxm::hi16(s));
                    });
                }
            }
        } else if constexpr (op == Op::VF) {
            rxm::stage_u32(L.ev, m.ev + d.e_off, d.ne);
            rxm::stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
            block.sync();
            rxm::fill_columns<T>(
                3u * d.nf, d.nv, [&](uint32_t i) { return rxm::face_vertex(L.fe, L.ev, i / 3u, i % 3u); },
                [&](uint32_t i) { return i / 3u; }, m_off, L.slots, L.wave);
            for (uint32_t v = threadIdx.x; v < rows; v += T) {
                const uint32_t o = m_off[v];
                rxm::sort_row(L.slots + o, m_off[v + 1] - o,
                              [&](uint32_t j, uint32_t s) { m_val[o + j] = static_cast<uint16_t>(s); });
            }
        } else if constexpr (op == Op::EV) {
            rxm::stage_u32(reinterpret_cast<uint32_t*>(m_val), m.ev + d.e_off, d.ne);
        } else if constexpr (op == Op::FV) {
            rxm::stage_u32(L.ev, m.ev + d.e_off, d.ne);
            rxm::stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
            block.sync();
            for (uint32_t i = threadIdx.x; i < 3u * d.nf; i += T)
                m_val[i] = static_cast<uint16_t>(rxm::face_vertex(L.fe, L.ev, i / 3u, i % 3u));
        } else if constexpr (op == Op::FE) {
            rxm::stage_fe(m_val, m.fe + 3u * d.f_off, d.nf);
            block.sync();
            for (uint32_t i = threadIdx.x; i < 3u * d.nf; i += T) m_val[i] = static_cast<uint16_t>(m_val[i] >> 1);
        } else if constexpr (op == Op::EF) {
            rxm::stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
            block.sync();
            rxm::fill_columns<T>(
                3u * d.nf, d.ne, [&](uint32_t i) { return static_cast<uint32_t>(L.fe[i] >> 1); },
                [&](uint32_t i) { return i / 3u; }, m_off, L.slots, L.wave);
            for (uint32_t e = threadIdx.x; e < rows; e += T) {
                const uint32_t o = m_off[e];
                rxm::sort_row(L.slots + o, m_off[e + 1] - o,
                              [&](uint32_t j, uint32_t s) { m_val[o + j] = static_cast<uint16_t>(s); });
            }
        } else if constexpr (op == Op::EVDiamond) {
            // [v0, opposite of v0 -> v1, v1, opposite of v1 -> v0] per local
            // edge, INVALID16 for a missing side; the opposite slots take the
            // u32 atomicMax of (face + 1) << 16 | vertex, so the later face
            // wins where two run an edge the same way (as rxm_query_csr)
            rxm::stage_u32(L.ev, m.ev + d.e_off, d.ne);
            rxm::stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
            for (uint32_t i = threadIdx.x; i < 2u * d.ne; i += T) L.slots[i] = 0u;
            block.sync();
            for (uint32_t i = threadIdx.x; i < 3u * d.nf; i += T) {
                const uint32_t f = i / 3u, k = i - 3u * f;
                const uint32_t y = L.fe[3u * f + (k == 2u ? 0u : k + 1u)];
                atomicMax(&L.slots[y], ((f + 1u) << 16) | rxm::face_vertex(L.fe, L.ev, f, k));
            }
            block.sync();
            for (uint32_t e = threadIdx.x; e < d.ne; e += T) {
                const uint32_t x = L.ev[e], a = L.slots[2u * e], c = L.slots[2u * e + 1u];
                m_val[4u * e]      = static_cast<uint16_t>(rxm::lo16(x));
                m_val[4u * e + 1u] = a ? static_cast<uint16_t>(a & 0xFFFFu) : INVALID16;
                m_val[4u * e + 2u] = static_cast<uint16_t>(rxm::hi16(x));
                m_val[4u * e + 3u] = c ? static_cast<uint16_t>(c & 0xFFFFu) : INVALID16;
            }
        } else if constexpr (op == Op::FF) {
            rxm::stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
            block.sync();
            rxm::fill_columns<T>(
                3u * d.nf, d.ne, [&](uint32_t i) { return static_cast<uint32_t>(L.fe[i] >> 1); },
                [&](uint32_t i) { return i / 3u; }, L.eoff, L.slots, L.wave);
            for (uint32_t e = threadIdx.x; e < d.ne; e += T) {
                const uint32_t o = L.eoff[e], k = L.eoff[e + 1] - o;
                if (k > 1u) rxm::sort_row(L.slots + o, k, [&](uint32_t j, uint32_t s) { L.slots[o + j] = s; });
            }
            block.sync();
            for (uint32_t f = threadIdx.x; f < d.nf; f += T) {
                uint32_t c = 0;
                if (f < rows)
                    for (uint32_t i = 0; i < 3u; ++i) {
                        const uint32_t e = L.fe[3u * f + i] >> 1;
                        c += L.eoff[e + 1] - L.eoff[e] - 1u;
                    }
                m_off[f] = c;
            }
            const uint32_t total = rxm::scan_excl<T>(m_off, d.nf, L.wave);
            if (threadIdx.x == 0) m_off[d.nf] = total;
            for (uint32_t f = threadIdx.x; f < rows; f += T) {
                uint32_t k = m_off[f];
                for (uint32_t i = 0; i < 3u; ++i) {
                    const uint32_t e = L.fe[3u * f + i] >> 1;
                    for (uint32_t j = L.eoff[e]; j < L.eoff[e + 1]; ++j) {
                        const uint32_t g = L.slots[j];
                        if (g != f) m_val[k++] = static_cast<uint16_t>(g);
                    }
                }
            }
        }
        block.sync();
    }

    template <typename computeT>
    __device__ void run_compute(cooperative_groups::thread_block& block, computeT compute_op)
    {
        if (m_pid == INVALID32) return;
        using ComputeTraits = detail::FunctionTraits<computeT>;
        using HandleT       = typename ComputeTraits::template arg<0>::type;
        using IteratorT     = typename ComputeTraits::template arg<1>::type;
        for (uint32_t s = threadIdx.x; s < m_num_src; s += blockThreads) {
            if (is_participant(s)) {
                HandleT   h(m_pid, static_cast<uint16_t>(s));
                IteratorT it = get_iterator<IteratorT>(static_cast<uint16_t>(s));
                compute_op(h, it);
            }
        }
    }

    template <typename IteratorT>
    __device__ IteratorT get_iterator(uint16_t local_id) const
    {
        using H = typename IteratorT::Handle;
        if (m_num_src == 0 || !is_participant(local_id)) return IteratorT(local_id, m_pid);
        const uint16_t* row;
        uint16_t        size;
        if (m_fixed) {
            row  = m_val + m_fixed * local_id;
            size = static_cast<uint16_t>(m_fixed);
        } else {
            row  = m_val + m_off[local_id];
            size = static_cast<uint16_t>(m_off[local_id + 1] - m_off[local_id]);
        }
        constexpr int k = detail::HandleKind<H>::value;
        const rxm::PatchDesc& d = m_desc;
        const rxm::DevMesh&   m = m_ctx.m;
        if constexpr (k == 0) return IteratorT(row, size, m_pid, d.nvo, m_ctx.v_prefix, m.v_lin + d.v_off, m.v_own + d.v_off);
        else if constexpr (k == 1) return IteratorT(row, size, m_pid, d.neo, m_ctx.e_prefix, m.e_lin + d.e_off, m.e_own + d.e_off);
        else return IteratorT(row, size, m_pid, d.nfo, m_ctx.f_prefix, m.f_lin + d.f_off, m.f_own + d.f_off);
    }

    __device__ void epilogue(cooperative_groups::thread_block& block, ShmemAllocator& shrd_alloc)
    {
        if (m_pid == INVALID32) return;
        block.sync();  // every lambda has finished reading the rows
        shrd_alloc.reset_to(m_shmem_before);
        m_num_src = 0;
        m_mask    = nullptr;
        m_off     = nullptr;
        m_val     = nullptr;
        m_op      = Op::INVALID;
    }

   private:
    __device__ bool is_participant(uint32_t s) const { return (m_mask[s >> 5] >> (s & 31u)) & 1u; }

    const Context  m_ctx;
    uint32_t       m_pid;
    rxm::PatchDesc m_desc{};
    Op             m_op           = Op::INVALID;
    uint32_t       m_fixed        = 0;
    uint32_t       m_shmem_before = 0;
    uint32_t       m_num_src      = 0;
    uint32_t*      m_mask         = nullptr;
    uint32_t*      m_off          = nullptr;
    uint16_t*      m_val          = nullptr;
    uint32_t*      m_s_valence    = nullptr;
};

}  // namespace rxmesh
