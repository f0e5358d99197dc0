// Query kernels: materialise any of the 8 relations as a CSR over linear ids.
//
// One workgroup per patch. Reference behaviour: query<op>
// (include/rxmesh/kernels/rxmesh_queries.cuh:861-1053) and the iterator's
// owner resolution (include/rxmesh/iterator.cuh:117-143). Rows are the owned
// source elements of the patch (linear ids prefix[p] + local); entries are the
// target elements' linear ids, i.e. always the owner's copy.
#pragma once
#include "lds_sort.hpp"
#include "orient.hpp"

namespace rxm {

constexpr uint32_t kQueryThreads = 256;

// ---- LDS plans (same carve on host for sizing and on device) -------------
// Oriented VV / VE: EV and FE slabs, the owned columns of the vertex-edge
// transpose (nrow = entries of the patch's owned rows), the half-edge
// successor table and the global ids that fix each ring's start.
struct PlanVXO
{
    uint32_t *ev, *off, *wave, *slots, *gid;
    uint16_t *fe, *hn;
};
__host__ __device__ inline PlanVXO plan_vxo(Bump& b, uint32_t nv, uint32_t ne, uint32_t nf, uint32_t nvo,
                                            uint32_t nrow)
{
    PlanVXO p{};
    p.ev    = b.alloc<uint32_t>(r4(ne));
    p.off   = b.alloc<uint32_t>(nvo + 1);
    p.wave  = b.alloc<uint32_t>(20);
    p.slots = b.alloc<uint32_t>(nrow);
    p.fe    = b.alloc<uint16_t>(r4(3 * nf));
    p.hn    = b.alloc<uint16_t>(2 * ne);
    p.gid   = b.alloc<uint32_t>(nv);
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

// Oriented rows of the patch's owned vertices: stage EV / FE / global ids,
// build the half-edge successor table and the owned columns of the
// vertex-edge transpose, then order every row into its canonical CCW ring
// (neighbour vertices if NBR, else edges) in place in L.slots. On return
// L.off[v] is the start of row v and L.off[nvo] the total.
template <uint32_t T, bool NBR>
__device__ inline void oriented_rows(const DevMesh& m, const PatchDesc& d, const PlanVXO& L)
{
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
    for (uint32_t v = threadIdx.x; v < d.nv; v += T) L.gid[v] = m.ltog_v[d.v_off + v];
    for (uint32_t i = threadIdx.x; i < d.ne; i += T) reinterpret_cast<uint32_t*>(L.hn)[i] = 0xFFFFFFFFu;
    for (uint32_t c = threadIdx.x; c <= d.nvo; c += T) L.off[c] = 0;
    __syncthreads();
    // the successor table is complete before the rows are ordered: the
    // counting sort below contains block barriers
    build_half_next(L.fe, L.ev, d.nf, L.hn);
    fill_columns<T, true>(
        2 * d.ne, d.nvo,
        [&](uint32_t i) {
            const uint32_t x = L.ev[i >> 1];
            return (i & 1u) ? hi16(x) : lo16(x);
        },
        [&](uint32_t i) { return i >> 1; }, L.off, L.slots, L.wave);
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T)
        orient_ring<12, NBR>(v, L.off[v], L.off[v + 1], L.slots, L.slots, L.ev, L.hn, L.gid);
    __syncthreads();
}

// Oriented VV (IS_VV) or VE; rows leave as one contiguous range per patch.
template <uint32_t T, bool IS_VV>
__global__ __launch_bounds__(T) void k_query_vx_oriented(DevMesh m, const uint32_t* __restrict__ base,
                                                         int32_t* __restrict__ row_ptr, int32_t* __restrict__ col)
{
    extern __shared__ uint4 smem[];
    const uint32_t p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d   = load_desc(m.patch, p);
    const uint32_t  bse = base[p], nent = base[p + 1] - bse;
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanVXO   L = plan_vxo(b, d.nv, d.ne, d.nf, d.nvo, nent);
    oriented_rows<T, IS_VV>(m, d, L);
    // the host's row counts (global valences) match the patch's owned
    // columns on every manifold patch; anything else leaves it unwritten
    if (L.off[d.nvo] != nent) return;
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) row_ptr[d.vp + v] = static_cast<int32_t>(bse + L.off[v]);
    if (threadIdx.x == 0) row_ptr[d.vp + d.nvo] = static_cast<int32_t>(bse + nent);
    for (uint32_t k = threadIdx.x; k < nent; k += T) {
        const uint32_t x = L.slots[k];
        const uint32_t out = IS_VV ? lin_of(x, d.nvo, d.vp, m.v_lin + d.v_off)
                                   : lin_of(x, d.neo, d.ep, m.e_lin + d.e_off);
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

// EVDiamond (reference e_v_diamond, rxmesh_queries.cuh:198-265): per owned
// edge [v0, the vertex closing the face that runs v0 -> v1, v1, the vertex
// closing the face that runs v1 -> v0] as linear vertex ids (v0 = lo16 of EV,
// the larger global id), -1 where the edge has no such face. Every local face
// slot names its start vertex as the opposite of the face's next slot, which
// lands in position 1 + 2 dir of that slot's edge. Owned edges' slots take a
// u32 LDS atomicMax over (face + 1) << 16 | vertex: where two faces run an
// edge the same way (non-manifold input) the larger local face index wins,
// as in a serial pass over the faces, independent of scheduling.
struct PlanEVD
{
    uint32_t* ev;
    uint16_t* fe;
    uint32_t* opp;
};
__host__ __device__ inline PlanEVD plan_evd(Bump& b, uint32_t ne, uint32_t nf, uint32_t neo)
{
    PlanEVD p{};
    p.ev  = b.alloc<uint32_t>(r4(ne));
    p.fe  = b.alloc<uint16_t>(r4(3 * nf));
    p.opp = b.alloc<uint32_t>(2 * neo);
    return p;
}

template <uint32_t T>
__global__ __launch_bounds__(T) void k_query_evd(DevMesh m, int32_t* __restrict__ row_ptr, int32_t* __restrict__ col)
{
    extern __shared__ uint4 smem[];
    const uint32_t p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanEVD   L = plan_evd(b, d.ne, d.nf, d.neo);
    for (uint32_t i = threadIdx.x; i < 2 * d.neo; i += T) L.opp[i] = 0u;
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 3 * d.nf; i += T) {
        const uint32_t f = i / 3u, k = i - 3u * f;
        const uint32_t y = L.fe[3u * f + (k == 2u ? 0u : k + 1u)];
        if ((y >> 1) < d.neo) atomicMax(&L.opp[y], ((f + 1u) << 16) | face_vertex(L.fe, L.ev, f, k));
    }
    __syncthreads();
    const uint32_t* vl = m.v_lin + d.v_off;
    for (uint32_t e = threadIdx.x; e < d.neo; e += T) {
        const uint32_t x = L.ev[e], r = d.ep + e, a = L.opp[2 * e], c = L.opp[2 * e + 1];
        row_ptr[r]     = static_cast<int32_t>(4 * r);
        col[4 * r]     = static_cast<int32_t>(lin_of(lo16(x), d.nvo, d.vp, vl));
        col[4 * r + 1] = a ? static_cast<int32_t>(lin_of(a & 0xFFFFu, d.nvo, d.vp, vl)) : -1;
        col[4 * r + 2] = static_cast<int32_t>(lin_of(hi16(x), d.nvo, d.vp, vl));
        col[4 * r + 3] = c ? static_cast<int32_t>(lin_of(c & 0xFFFFu, d.nvo, d.vp, vl)) : -1;
    }
    if (p + 1 == m.P && threadIdx.x == 0) row_ptr[d.ep + d.neo] = static_cast<int32_t>(4 * (d.ep + d.neo));
}

}  // namespace rxm
