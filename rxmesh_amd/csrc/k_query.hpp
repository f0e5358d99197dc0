// Query kernels: materialise any of the 8 relations as a CSR over linear ids.
//
// One workgroup per patch. Reference behaviour: query<op>
// (include/rxmesh/kernels/rxmesh_queries.cuh:861-1053) and the iterator's
// owner resolution (include/rxmesh/iterator.cuh:117-143). Rows are the owned
// source elements of the patch (linear ids prefix[p] + local); entries are the
// target elements' linear ids, i.e. always the owner's copy.
#pragma once
#include "orient.hpp"

namespace rxm {

constexpr uint32_t kQueryThreads = 256;

// ---- LDS plans (same carve on host for sizing and on device) -------------
struct PlanVX
{
    uint32_t *ev, *off, *cur, *wave, *ef, *gid;
    uint16_t *val, *rowof, *fe;
};
__host__ __device__ inline PlanVX plan_vx(Bump& b, uint32_t nv, uint32_t ne, uint32_t nf, bool oriented)
{
    PlanVX p{};
    p.ev    = b.alloc<uint32_t>(r4(ne));
    p.off   = b.alloc<uint32_t>(nv + 1);
    p.cur   = b.alloc<uint32_t>(nv);
    p.wave  = b.alloc<uint32_t>(20);
    p.val   = b.alloc<uint16_t>(2 * ne);
    p.rowof = b.alloc<uint16_t>(2 * ne);
    if (oriented) {
        p.fe  = b.alloc<uint16_t>(r4(3 * nf));
        p.ef  = b.alloc<uint32_t>(2 * ne);
        p.gid = b.alloc<uint32_t>(nv);
    }
    return p;
}

struct PlanFX
{
    uint32_t* ev;
    uint16_t* fe;
};
__host__ __device__ inline PlanFX plan_fx(Bump& b, uint32_t ne, uint32_t nf)
{
    PlanFX p{};
    p.ev = b.alloc<uint32_t>(r4(ne));
    p.fe = b.alloc<uint16_t>(r4(3 * nf));
    return p;
}

// Shared prologue of VV / VE: stage EV, transpose to VE, optionally orient.
__device__ inline void vx_prologue(const DevMesh& m, const PatchDesc& d, const PlanVX& L, bool oriented)
{
    const uint32_t T = blockDim.x;
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    if (oriented) {
        stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
        for (uint32_t v = threadIdx.x; v < d.nv; v += T) L.gid[v] = m.ltog_v[d.v_off + v];
        for (uint32_t i = threadIdx.x; i < 2 * d.ne; i += T) L.ef[i] = 0xFFFFFFFFu;
    }
    __syncthreads();
    // the canonical ring does not depend on the row order: no row sort needed
    lds_transpose<2>(2 * d.ne, d.nv,
                     [&](uint32_t i) {
                         const uint32_t x = L.ev[i >> 1];
                         return (i & 1u) ? hi16(x) : lo16(x);
                     },
                     L.off, L.cur, L.val, L.wave, oriented ? 0u : d.nvo);
    if (oriented) {
        for (uint32_t i = threadIdx.x; i < 3 * d.nf; i += T) {
            const uint32_t e = L.fe[i] >> 1, f = i / 3;
            if (atomicCAS(&L.ef[2 * e], 0xFFFFFFFFu, f) != 0xFFFFFFFFu) L.ef[2 * e + 1] = f;
        }
        __syncthreads();
        for (uint32_t v = threadIdx.x; v < d.nvo; v += T)
            orient_row_fast<12>(v, L.off[v], L.off[v + 1], L.val, L.ev, L.fe, L.ef, L.gid);
        __syncthreads();
    }
}

// VV (IS_VV) or VE
template <bool IS_VV>
__global__ __launch_bounds__(kQueryThreads) void k_query_vx(DevMesh m, const uint32_t* __restrict__ base,
                                                            int32_t* __restrict__ row_ptr,
                                                            int32_t* __restrict__ col, int oriented)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = blockIdx.x;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanVX    L = plan_vx(b, d.nv, d.ne, d.nf, oriented != 0);
    vx_prologue(m, d, L, oriented != 0);
    const uint32_t T = blockDim.x;
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T)
        for (uint32_t k = L.off[v]; k < L.off[v + 1]; ++k) L.rowof[k] = static_cast<uint16_t>(v);
    __syncthreads();
    const uint32_t bse  = base[p];
    const uint32_t nent = L.off[d.nvo];
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) row_ptr[d.vp + v] = static_cast<int32_t>(bse + L.off[v]);
    if (threadIdx.x == 0) row_ptr[d.vp + d.nvo] = static_cast<int32_t>(bse + nent);
    for (uint32_t k = threadIdx.x; k < nent; k += T) {
        const uint32_t e = L.val[k];
        uint32_t       out;
        if constexpr (IS_VV) {
            const uint32_t x = L.ev[e], v = L.rowof[k];
            const uint32_t w = (lo16(x) == v) ? hi16(x) : lo16(x);
            out              = lin_of(w, d.nvo, d.vp, m.v_lin + d.v_off);
        } else {
            out = lin_of(e, d.neo, d.ep, m.e_lin + d.e_off);
        }
        col[bse + k] = static_cast<int32_t>(out);
    }
}

// Fixed-width relations EV (2), FV (3), FE (3); OP as in the reference enum.
template <int OP>
__global__ __launch_bounds__(kQueryThreads) void k_query_fixed(DevMesh m, int32_t* __restrict__ row_ptr,
                                                               int32_t* __restrict__ col)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = blockIdx.x;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanFX    L = plan_fx(b, d.ne, d.nf);
    const uint32_t  T = blockDim.x;
    if constexpr (OP == 9) {  // EV
        for (uint32_t e = threadIdx.x; e < d.neo; e += T) {
            const uint32_t x = m.ev[d.e_off + e], r = d.ep + e;
            row_ptr[r]       = static_cast<int32_t>(2 * r);
            col[2 * r]       = static_cast<int32_t>(lin_of(lo16(x), d.nvo, d.vp, m.v_lin + d.v_off));
            col[2 * r + 1]   = static_cast<int32_t>(lin_of(hi16(x), d.nvo, d.vp, m.v_lin + d.v_off));
        }
        if (p + 1 == m.P && threadIdx.x == 0) row_ptr[d.ep + d.neo] = static_cast<int32_t>(2 * (d.ep + d.neo));
    } else {
        if constexpr (OP == 6) stage_u32(L.ev, m.ev + d.e_off, d.ne);
        stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < 3 * d.nfo; i += T) {
            const uint32_t r = d.fp + i / 3;
            uint32_t       out;
            if constexpr (OP == 6) out = lin_of(face_vertex(L.fe, L.ev, i / 3, i % 3), d.nvo, d.vp, m.v_lin + d.v_off);
            else out = lin_of(L.fe[i] >> 1, d.neo, d.ep, m.e_lin + d.e_off);
            col[3 * d.fp + i] = static_cast<int32_t>(out);
            if (i % 3 == 0) row_ptr[r] = static_cast<int32_t>(3 * r);
        }
        if (p + 1 == m.P && threadIdx.x == 0) row_ptr[d.fp + d.nfo] = static_cast<int32_t>(3 * (d.fp + d.nfo));
    }
}

}  // namespace rxm
