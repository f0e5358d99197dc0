// Per-vertex / per-face geometry on top of the patch query engine.
//
// Vertex attributes use the reference's SoA layout
// (include/rxmesh/attribute.h:406-472): component c of the vertex with linear
// id l lives at data[c * N + l], l = prefix[patch] + local. Each workgroup
// stages its patch's positions (owned: one contiguous run; ribbon: gathered
// through the owner table) into LDS once, and writes owned outputs only.
// Because every owned vertex's full one-ring lies inside its patch, each
// output is a plain gather in LDS: no global atomics, deterministic order.
#pragma once
#include "k_query.hpp"
#include "k_query2.hpp"

namespace rxm {

constexpr uint32_t kGeomThreads = 256;

__device__ __forceinline__ void stage_positions(float* sx, float* sy, float* sz, const float* __restrict__ pos,
                                                uint32_t N, const PatchDesc& d, const uint32_t* __restrict__ v_lin)
{
    for (uint32_t l = threadIdx.x; l < d.nv; l += blockDim.x) {
        const uint32_t g = lin_of(l, d.nvo, d.vp, v_lin + d.v_off);
        sx[l]            = pos[g];
        sy[l]            = pos[N + g];
        sz[l]            = pos[2 * N + g];
    }
}

// Same staging with all of a thread's loads issued before any LDS store:
// R rounds of T vertices are unrolled, so a thread's owner-table loads go out
// together, then its position gathers, instead of one dependent load chain
// per loop iteration. Vertices past R * T (larger patches) take the loop.
template <uint32_t T, uint32_t R, bool AOS4>
__device__ __forceinline__ void stage_positions_batched(float* sx, float* sy, float* sz, float4* s4,
                                                        const float* __restrict__ pos, uint32_t N,
                                                        const PatchDesc& d, const uint32_t* __restrict__ v_lin)
{
    uint32_t g[R];
    float    x[R], y[R], z[R];
    const uint32_t* lin = v_lin + d.v_off;
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        const uint32_t l = threadIdx.x + r * T;
        g[r]             = (l < d.nvo) ? d.vp + l : ((l < d.nv) ? lin[l] : 0u);
    }
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        if (threadIdx.x + r * T < d.nv) {
            x[r] = pos[g[r]];
            y[r] = pos[N + g[r]];
            z[r] = pos[2 * N + g[r]];
        }
    }
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        const uint32_t l = threadIdx.x + r * T;
        if (l < d.nv) {
            if constexpr (AOS4) {
                s4[l] = make_float4(x[r], y[r], z[r], 0.f);
            } else {
                sx[l] = x[r];
                sy[l] = y[r];
                sz[l] = z[r];
            }
        }
    }
    for (uint32_t l = threadIdx.x + R * T; l < d.nv; l += T) {
        const uint32_t gg = lin_of(l, d.nvo, d.vp, lin);
        if constexpr (AOS4) {
            s4[l] = make_float4(pos[gg], pos[N + gg], pos[2 * N + gg], 0.f);
        } else {
            sx[l] = pos[gg];
            sy[l] = pos[N + gg];
            sz[l] = pos[2 * N + gg];
        }
    }
}

// ---- vertex normals -------------------------------------------------------
struct PlanVN
{
    uint32_t *ev, *off, *cur, *wave;
    uint16_t *fe, *fvl, *val;
    float *   x, *y, *z, *fn, *cw;
};
__host__ __device__ inline PlanVN plan_vn(Bump& b, uint32_t nv, uint32_t ne, uint32_t nf)
{
    PlanVN p{};
    p.ev   = b.alloc<uint32_t>(r4(ne));
    p.off  = b.alloc<uint32_t>(nv + 1);
    p.cur  = b.alloc<uint32_t>(nv);
    p.wave = b.alloc<uint32_t>(20);
    p.fe   = b.alloc<uint16_t>(r4(3 * nf));
    p.fvl  = b.alloc<uint16_t>(3 * nf);
    p.val  = b.alloc<uint16_t>(3 * nf);
    p.x    = b.alloc<float>(nv);
    p.y    = b.alloc<float>(nv);
    p.z    = b.alloc<float>(nv);
    p.fn   = b.alloc<float>(3 * nf);
    p.cw   = b.alloc<float>(3 * nf);
    return p;
}

// mode 0: Max (1999) weights n_f / (|e_i|^2 + |e_{i+2}|^2), restating
// apps/VertexNormal/vertex_normal_kernel.cuh:16-36; mode 1: area weights n_f/2.
// Output is unnormalised, as in the reference.
__global__ __launch_bounds__(kGeomThreads) void k_vertex_normals(DevMesh m, int mode, const float* __restrict__ pos,
                                                                 float* __restrict__ nrm)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = blockIdx.x;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanVN    L = plan_vn(b, d.nv, d.ne, d.nf);
    const uint32_t  T = blockDim.x, N = m.nV;
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
    stage_positions(L.x, L.y, L.z, pos, N, d, m.v_lin);
    __syncthreads();
    for (uint32_t f = threadIdx.x; f < d.nf; f += T) {
        const uint32_t a = face_vertex(L.fe, L.ev, f, 0), bb = face_vertex(L.fe, L.ev, f, 1),
                       c = face_vertex(L.fe, L.ev, f, 2);
        L.fvl[3 * f] = static_cast<uint16_t>(a);
        L.fvl[3 * f + 1] = static_cast<uint16_t>(bb);
        L.fvl[3 * f + 2] = static_cast<uint16_t>(c);
        const float ux = L.x[bb] - L.x[a], uy = L.y[bb] - L.y[a], uz = L.z[bb] - L.z[a];
        const float wx = L.x[c] - L.x[a], wy = L.y[c] - L.y[a], wz = L.z[c] - L.z[a];
        L.fn[3 * f]     = uy * wz - uz * wy;
        L.fn[3 * f + 1] = uz * wx - ux * wz;
        L.fn[3 * f + 2] = ux * wy - uy * wx;
        if (mode == 0) {
            const float vx = L.x[c] - L.x[bb], vy = L.y[c] - L.y[bb], vz = L.z[c] - L.z[bb];
            const float l0 = ux * ux + uy * uy + uz * uz;  // a-b
            const float l1 = vx * vx + vy * vy + vz * vz;  // b-c
            const float l2 = wx * wx + wy * wy + wz * wz;  // c-a
            L.cw[3 * f]     = 1.0f / (l0 + l2);
            L.cw[3 * f + 1] = 1.0f / (l1 + l0);
            L.cw[3 * f + 2] = 1.0f / (l2 + l1);
        } else {
            L.cw[3 * f] = L.cw[3 * f + 1] = L.cw[3 * f + 2] = 0.5f;
        }
    }
    __syncthreads();
    lds_transpose<3>(3 * d.nf, d.nv, [&](uint32_t i) { return static_cast<uint32_t>(L.fvl[i]); }, L.off, L.cur,
                     L.val, L.wave, d.nvo);
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        float sx = 0.f, sy = 0.f, sz = 0.f;
        for (uint32_t k = L.off[v]; k < L.off[v + 1]; ++k) {
            const uint32_t f = L.val[k];
            const uint32_t c = (L.fvl[3 * f] == v) ? 0u : ((L.fvl[3 * f + 1] == v) ? 1u : 2u);
            const float    w = L.cw[3 * f + c];
            sx += L.fn[3 * f] * w;
            sy += L.fn[3 * f + 1] * w;
            sz += L.fn[3 * f + 2] * w;
        }
        const uint32_t g = d.vp + v;
        nrm[g]           = sx;
        nrm[N + g]       = sy;
        nrm[2 * N + g]   = sz;
    }
}

// ---- uniform smoothing over the patch-local VV plan --------------------------
// One workgroup per patch with a small LDS footprint (positions + the
// patch's plan rows), so many patches are in flight per CU. Rows are read
// with 16-byte loads; neighbours are summed in the plan's row order.
struct PlanSL
{
    uint16_t *ro, *nb;
    float *   x, *y, *z;
    float4*   p4;
};
// AOS4: positions staged as one float4 per vertex (one 16-byte LDS access
// per neighbour) instead of three float arrays
__host__ __device__ inline PlanSL plan_sl(Bump& b, uint32_t nv, uint32_t nvo, uint32_t nent, bool aos4 = false)
{
    PlanSL p{};
    p.ro = b.alloc<uint16_t>(((nvo + 1u) + 7u) & ~7u);
    p.nb = b.alloc<uint16_t>((nent + 7u) & ~7u);
    if (aos4) {
        p.p4 = b.alloc<float4>(nv);
    } else {
        p.x = b.alloc<float>(nv);
        p.y = b.alloc<float>(nv);
        p.z = b.alloc<float>(nv);
    }
    return p;
}

// copy n u16 from a 16-byte aligned, 8-element padded run into LDS
template <uint32_t T>
__device__ __forceinline__ void stage_u16x8(uint16_t* dst, const uint16_t* __restrict__ src, uint32_t n)
{
    const uint32_t n8 = (n + 7u) >> 3;
    const uint4*   s  = reinterpret_cast<const uint4*>(src);
    uint4*         d  = reinterpret_cast<uint4*>(dst);
    for (uint32_t i = threadIdx.x; i < n8; i += T) d[i] = s[i];
}

template <uint32_t T, bool AOS4>
__global__ __launch_bounds__(T) void k_smooth_local(DevMesh m, const uint32_t* __restrict__ rbase,
                                                    const uint32_t* __restrict__ nbase,
                                                    const uint16_t* __restrict__ roff,
                                                    const uint16_t* __restrict__ nbr, float lambda,
                                                    const float* __restrict__ pos, float* __restrict__ out)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d = load_desc(m.patch, p);
    const uint32_t  rb = rbase[p], nbb = nbase[p], nent = nbase[p + 1] - nbb;
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanSL    L = plan_sl(b, d.nv, d.nvo, nent, AOS4);
    const uint32_t  N = m.nV;
    stage_u16x8<T>(L.ro, roff + rb, d.nvo + 1);
    stage_u16x8<T>(L.nb, nbr + nbb, nent);
    stage_positions_batched<T, 512u / T, AOS4>(L.x, L.y, L.z, L.p4, pos, N, d, m.v_lin);
    __syncthreads();
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        const uint32_t k0 = L.ro[v], k1 = L.ro[v + 1];
        float          sx = 0.f, sy = 0.f, sz = 0.f;
        float          px, py, pz;
        if constexpr (AOS4) {
            for (uint32_t k = k0; k < k1; ++k) {
                const float4 q = L.p4[L.nb[k]];
                sx += q.x;
                sy += q.y;
                sz += q.z;
            }
            const float4 q = L.p4[v];
            px             = q.x;
            py             = q.y;
            pz             = q.z;
        } else {
            for (uint32_t k = k0; k < k1; ++k) {
                const uint32_t w = L.nb[k];
                sx += L.x[w];
                sy += L.y[w];
                sz += L.z[w];
            }
            px = L.x[v];
            py = L.y[v];
            pz = L.z[v];
        }
        const uint32_t g = d.vp + v;
        if (k1 > k0) {
            const float inv = 1.0f / static_cast<float>(k1 - k0);
            out[g]          = px + lambda * (sx * inv - px);
            out[N + g]      = py + lambda * (sy * inv - py);
            out[2 * N + g]  = pz + lambda * (sz * inv - pz);
        } else {
            out[g]         = px;
            out[N + g]     = py;
            out[2 * N + g] = pz;
        }
    }
}

// ---- two Jacobi steps per launch over the two-step plan ------------------------
// The workgroup stages the positions of its patch's local vertices and of
// their second ring (extra vertices, gathered through the plan's linear-id
// list), runs step 1 for every local vertex (owned and ribbon, each with
// its owner's neighbour order) into LDS, then step 2 for the owned vertices
// from those values. Each value equals the one two separate steps produce.
// AOS4 keeps both position sets as one float4 per vertex in LDS.
struct PlanS2
{
    uint16_t *ro, *nb;
    float *   x, *y, *z;     // nv + nx: input positions
    float *   x1, *y1, *z1;  // nv: after step 1
    float4 *  p4, *q4;       // AOS4 forms of the two sets
};
__host__ __device__ inline PlanS2 plan_s2(Bump& b, uint32_t nv, uint32_t nx, uint32_t nent, bool aos4 = false)
{
    PlanS2 p{};
    p.ro = b.alloc<uint16_t>(((nv + 1u) + 7u) & ~7u);
    p.nb = b.alloc<uint16_t>((nent + 7u) & ~7u);
    if (aos4) {
        p.p4 = b.alloc<float4>(nv + nx);
        p.q4 = b.alloc<float4>(nv);
    } else {
        p.x  = b.alloc<float>(nv + nx);
        p.y  = b.alloc<float>(nv + nx);
        p.z  = b.alloc<float>(nv + nx);
        p.x1 = b.alloc<float>(nv);
        p.y1 = b.alloc<float>(nv);
        p.z1 = b.alloc<float>(nv);
    }
    return p;
}

template <bool AOS4>
__device__ __forceinline__ float3 s2_get(const PlanS2& L, uint32_t w, bool second)
{
    if constexpr (AOS4) {
        const float4 q = second ? L.q4[w] : L.p4[w];
        return make_float3(q.x, q.y, q.z);
    } else {
        return second ? make_float3(L.x1[w], L.y1[w], L.z1[w]) : make_float3(L.x[w], L.y[w], L.z[w]);
    }
}

// one Jacobi update of vertex v from set `second` (false: input, true: step 1)
template <bool AOS4>
__device__ __forceinline__ float3 s2_update(const PlanS2& L, uint32_t v, bool second, float lambda)
{
    const uint32_t k0 = L.ro[v], k1 = L.ro[v + 1];
    float          sx = 0.f, sy = 0.f, sz = 0.f;
    for (uint32_t k = k0; k < k1; ++k) {
        const float3 q = s2_get<AOS4>(L, L.nb[k], second);
        sx += q.x;
        sy += q.y;
        sz += q.z;
    }
    const float3 pv = s2_get<AOS4>(L, v, second);
    if (k1 > k0) {
        const float inv = 1.0f / static_cast<float>(k1 - k0);
        return make_float3(pv.x + lambda * (sx * inv - pv.x), pv.y + lambda * (sy * inv - pv.y),
                           pv.z + lambda * (sz * inv - pv.z));
    }
    return pv;
}

// Staging of the two-step kernel in two dependent levels of global loads,
// each level's loads of a thread issued before it waits: (1) the plan rows
// (u16 x 8 vectors), the owner-table ids of the ribbon vertices and the
// extra vertices' linear ids; (2) the position gathers. Then the LDS stores.
// Counts beyond the register rounds take plain loops.
// GIN: positions in HBM as one float4 per vertex (iterates between launches)
// instead of SoA [3][N]; a gather then touches one cache line, not three.
template <uint32_t T, bool AOS4, bool GIN = false>
__device__ __forceinline__ void stage_s2(const PlanS2& L, const DevMesh& m, const PatchDesc& d,
                                         const uint16_t* __restrict__ ro_src, const uint16_t* __restrict__ nb_src,
                                         uint32_t nent, const uint32_t* __restrict__ xl, uint32_t nx,
                                         const float* __restrict__ pos)
{
    constexpr uint32_t R1 = 2, RN = 4;
    constexpr uint32_t RV = (512u / T) > 0u ? 512u / T : 1u;
    constexpr uint32_t RX = (256u / T) > 0u ? 256u / T : 1u;
    const uint32_t     N   = m.nV;
    const uint32_t     nro = (d.nv + 1u + 7u) >> 3, nnb = (nent + 7u) >> 3;
    const uint4*       rs  = reinterpret_cast<const uint4*>(ro_src);
    const uint4*       ns  = reinterpret_cast<const uint4*>(nb_src);
    uint4*             rd  = reinterpret_cast<uint4*>(L.ro);
    uint4*             nd  = reinterpret_cast<uint4*>(L.nb);
    const uint32_t*    vl  = m.v_lin + d.v_off;
    const uint32_t     t   = threadIdx.x;
    uint4              ra[R1], na[RN];
    uint32_t           g[RV], gx[RX];
    // level 1
#pragma unroll
    for (uint32_t r = 0; r < R1; ++r) ra[r] = (t + r * T < nro) ? rs[t + r * T] : uint4{};
#pragma unroll
    for (uint32_t r = 0; r < RN; ++r) na[r] = (t + r * T < nnb) ? ns[t + r * T] : uint4{};
#pragma unroll
    for (uint32_t r = 0; r < RV; ++r) {
        const uint32_t l = t + r * T;
        g[r]             = (l < d.nvo) ? d.vp + l : ((l < d.nv) ? vl[l] : 0u);
    }
#pragma unroll
    for (uint32_t r = 0; r < RX; ++r) gx[r] = (t + r * T < nx) ? xl[t + r * T] : 0u;
    // level 2
    const float4* p4 = reinterpret_cast<const float4*>(pos);
    auto          ld = [&](uint32_t gg, float& a, float& bb, float& c) {
        if constexpr (GIN) {
            const float4 q = p4[gg];
            a              = q.x;
            bb             = q.y;
            c              = q.z;
        } else {
            a  = pos[gg];
            bb = pos[N + gg];
            c  = pos[2 * N + gg];
        }
    };
    float x[RV], y[RV], z[RV], xx[RX], xy[RX], xz[RX];
#pragma unroll
    for (uint32_t r = 0; r < RV; ++r)
        if (t + r * T < d.nv) ld(g[r], x[r], y[r], z[r]);
#pragma unroll
    for (uint32_t r = 0; r < RX; ++r)
        if (t + r * T < nx) ld(gx[r], xx[r], xy[r], xz[r]);
    auto put = [&](uint32_t l, float a, float bb, float c) {
        if constexpr (AOS4) {
            L.p4[l] = make_float4(a, bb, c, 0.f);
        } else {
            L.x[l] = a;
            L.y[l] = bb;
            L.z[l] = c;
        }
    };
#pragma unroll
    for (uint32_t r = 0; r < R1; ++r)
        if (t + r * T < nro) rd[t + r * T] = ra[r];
#pragma unroll
    for (uint32_t r = 0; r < RN; ++r)
        if (t + r * T < nnb) nd[t + r * T] = na[r];
#pragma unroll
    for (uint32_t r = 0; r < RV; ++r)
        if (t + r * T < d.nv) put(t + r * T, x[r], y[r], z[r]);
#pragma unroll
    for (uint32_t r = 0; r < RX; ++r)
        if (t + r * T < nx) put(d.nv + t + r * T, xx[r], xy[r], xz[r]);
    // remainders of larger patches
    for (uint32_t i = t + R1 * T; i < nro; i += T) rd[i] = rs[i];
    for (uint32_t i = t + RN * T; i < nnb; i += T) nd[i] = ns[i];
    for (uint32_t l = t + RV * T; l < d.nv; l += T) {
        float a, bb, c;
        ld(lin_of(l, d.nvo, d.vp, vl), a, bb, c);
        put(l, a, bb, c);
    }
    for (uint32_t i = t + RX * T; i < nx; i += T) {
        float a, bb, c;
        ld(xl[i], a, bb, c);
        put(d.nv + i, a, bb, c);
    }
}

// GIN / GOUT: input / output positions as one float4 per vertex (the
// iterates between the launches of a multi-launch smoothing) instead of SoA.
template <uint32_t T, bool AOS4, bool GIN = false, bool GOUT = false>
__global__ __launch_bounds__(T) void k_smooth2(DevMesh m, const uint32_t* __restrict__ rbase,
                                               const uint32_t* __restrict__ nbase, const uint32_t* __restrict__ xbase,
                                               const uint32_t* __restrict__ xcnt, const uint16_t* __restrict__ roff,
                                               const uint16_t* __restrict__ nbr, const uint32_t* __restrict__ xlin,
                                               float lambda, const float* __restrict__ pos, float* __restrict__ out)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d  = load_desc(m.patch, p);
    const uint32_t  rb = rbase[p], nb0 = nbase[p], nent = nbase[p + 1] - nb0, xb = xbase[p], nx = xcnt[p];
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanS2    L = plan_s2(b, d.nv, nx, nent, AOS4);
    const uint32_t  N = m.nV;
    stage_s2<T, AOS4, GIN>(L, m, d, roff + rb, nbr + nb0, nent, xlin + xb, nx, pos);
    __syncthreads();
    // step 1: every local vertex
    for (uint32_t v = threadIdx.x; v < d.nv; v += T) {
        const float3 r = s2_update<AOS4>(L, v, false, lambda);
        if constexpr (AOS4) {
            L.q4[v] = make_float4(r.x, r.y, r.z, 0.f);
        } else {
            L.x1[v] = r.x;
            L.y1[v] = r.y;
            L.z1[v] = r.z;
        }
    }
    __syncthreads();
    // step 2: owned vertices (their neighbours are all local)
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        const float3   r = s2_update<AOS4>(L, v, true, lambda);
        const uint32_t g = d.vp + v;
        if constexpr (GOUT) {
            reinterpret_cast<float4*>(out)[g] = make_float4(r.x, r.y, r.z, 0.f);
        } else {
            out[g]         = r.x;
            out[N + g]     = r.y;
            out[2 * N + g] = r.z;
        }
    }
}

// ---- K-step smoothing over the level plan (temporal blocking) ---------------
// One workgroup per patch runs `steps` Jacobi steps of uniform smoothing on
// the patch's extended vertex space (owned vertices, then the rings around
// them by level): step i computes the prefix [0, cnt[steps - i]) from the
// previous step's values, the last step the owned vertices, which go to HBM.
// Each vertex update lists its neighbours in the owner's canonical VV order,
// as k_smooth_local does, so any split of the iterations into launches gives
// bit-identical results. lvl: 2K + 1 u32 per patch (cnt[0..K], ent[0..K-1]).
struct PlanSK
{
    uint16_t *ro, *nb;
    float *   ax, *ay, *az;  // cnt[steps]: input positions, later even steps
    float *   bx, *by, *bz;  // cnt[steps - 1]: odd steps
};
__host__ __device__ inline PlanSK plan_sk(Bump& b, uint32_t nin, uint32_t ncomp, uint32_t nent)
{
    PlanSK p{};
    p.ro = b.alloc<uint16_t>(((ncomp + 1u) + 7u) & ~7u);
    p.nb = b.alloc<uint16_t>((nent + 7u) & ~7u);
    p.ax = b.alloc<float>(nin);
    p.ay = b.alloc<float>(nin);
    p.az = b.alloc<float>(nin);
    p.bx = b.alloc<float>(ncomp);
    p.by = b.alloc<float>(ncomp);
    p.bz = b.alloc<float>(ncomp);
    return p;
}

template <uint32_t T>
__global__ __launch_bounds__(T) void k_smoothk(DevMesh m, uint32_t K, const uint32_t* __restrict__ lvl,
                                               const uint32_t* __restrict__ rbase,
                                               const uint32_t* __restrict__ nbase,
                                               const uint32_t* __restrict__ xbase,
                                               const uint16_t* __restrict__ roff, const uint16_t* __restrict__ nbr,
                                               const uint32_t* __restrict__ xlin, uint32_t steps, float lambda,
                                               const float* __restrict__ pos, float* __restrict__ out)
{
    extern __shared__ uint4 smem[];
    const uint32_t p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d     = load_desc(m.patch, p);
    const uint32_t* c     = lvl + size_t(p) * (2 * K + 1);
    const uint32_t  nin   = c[steps], ncomp = c[steps - 1], nent = c[K + steps];
    const uint32_t  N     = m.nV;
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanSK    L = plan_sk(b, nin, ncomp, nent);
    {
        // staging in two levels of global loads: (1) rows and the extended
        // vertices' linear ids, (2) the position gathers
        constexpr uint32_t R1 = 2, RN = 6, RV = 4;
        const uint32_t     t   = threadIdx.x;
        const uint32_t     nro = (ncomp + 1u + 7u) >> 3, nnb = (nent + 7u) >> 3;
        const uint4*       rs  = reinterpret_cast<const uint4*>(roff + rbase[p]);
        const uint4*       ns  = reinterpret_cast<const uint4*>(nbr + nbase[p]);
        const uint32_t*    xl  = xlin + xbase[p];
        uint4*             rd  = reinterpret_cast<uint4*>(L.ro);
        uint4*             nd  = reinterpret_cast<uint4*>(L.nb);
        uint4              ra[R1], na[RN];
        uint32_t           g[RV];
#pragma unroll
        for (uint32_t r = 0; r < R1; ++r) ra[r] = (t + r * T < nro) ? rs[t + r * T] : uint4{};
#pragma unroll
        for (uint32_t r = 0; r < RN; ++r) na[r] = (t + r * T < nnb) ? ns[t + r * T] : uint4{};
#pragma unroll
        for (uint32_t r = 0; r < RV; ++r) {
            const uint32_t l = t + r * T;
            g[r]             = (l < d.nvo) ? d.vp + l : ((l < nin) ? xl[l - d.nvo] : 0u);
        }
        float x[RV], y[RV], z[RV];
#pragma unroll
        for (uint32_t r = 0; r < RV; ++r) {
            if (t + r * T < nin) {
                x[r] = pos[g[r]];
                y[r] = pos[N + g[r]];
                z[r] = pos[2 * N + g[r]];
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < R1; ++r)
            if (t + r * T < nro) rd[t + r * T] = ra[r];
#pragma unroll
        for (uint32_t r = 0; r < RN; ++r)
            if (t + r * T < nnb) nd[t + r * T] = na[r];
#pragma unroll
        for (uint32_t r = 0; r < RV; ++r) {
            const uint32_t l = t + r * T;
            if (l < nin) {
                L.ax[l] = x[r];
                L.ay[l] = y[r];
                L.az[l] = z[r];
            }
        }
        for (uint32_t i = t + R1 * T; i < nro; i += T) rd[i] = rs[i];
        for (uint32_t i = t + RN * T; i < nnb; i += T) nd[i] = ns[i];
        for (uint32_t l = t + RV * T; l < nin; l += T) {
            const uint32_t gg = (l < d.nvo) ? d.vp + l : xl[l - d.nvo];
            L.ax[l]           = pos[gg];
            L.ay[l]           = pos[N + gg];
            L.az[l]           = pos[2 * N + gg];
        }
    }
    __syncthreads();
    const float *sx = L.ax, *sy = L.ay, *sz = L.az;
    float *      tx = L.bx, *ty = L.by, *tz = L.bz;
    for (uint32_t i = 1; i <= steps; ++i) {
        const uint32_t n    = c[steps - i];
        const bool     last = (i == steps);
        for (uint32_t v = threadIdx.x; v < n; v += T) {
            const uint32_t k0 = L.ro[v], k1 = L.ro[v + 1];
            float          ax = 0.f, ay = 0.f, az = 0.f;
            for (uint32_t k = k0; k < k1; ++k) {
                const uint32_t w = L.nb[k];
                ax += sx[w];
                ay += sy[w];
                az += sz[w];
            }
            const float px = sx[v], py = sy[v], pz = sz[v];
            float       rx = px, ry = py, rz = pz;
            if (k1 > k0) {
                const float inv = 1.0f / static_cast<float>(k1 - k0);
                rx              = px + lambda * (ax * inv - px);
                ry              = py + lambda * (ay * inv - py);
                rz              = pz + lambda * (az * inv - pz);
            }
            if (last) {
                const uint32_t gg = d.vp + v;
                out[gg]           = rx;
                out[N + gg]       = ry;
                out[2 * N + gg]   = rz;
            } else {
                tx[v] = rx;
                ty[v] = ry;
                tz[v] = rz;
            }
        }
        if (!last) {
            __syncthreads();
            // the buffer just read becomes the next target (its prefix is large enough)
            float* ox = const_cast<float*>(sx);
            float* oy = const_cast<float*>(sy);
            float* oz = const_cast<float*>(sz);
            sx = tx; sy = ty; sz = tz;
            tx = ox; ty = oy; tz = oz;
        }
    }
}

// ---- vertex normals over the patch-local face plan ----------------------------
// Same weights and summation order as k_vertex_normals (incident faces in
// ascending local order), but the VF rows come precomputed from the face
// plan and every owned vertex evaluates its incident faces' normal and
// corner weight itself, so LDS holds only the plan rows and positions.
struct PlanVNP
{
    uint16_t *fvl, *ro, *vf;
    float *   x, *y, *z;
};
__host__ __device__ inline PlanVNP plan_vnp(Bump& b, uint32_t nv, uint32_t nvo, uint32_t nf, uint32_t nent)
{
    PlanVNP p{};
    p.fvl = b.alloc<uint16_t>((3 * nf + 7u) & ~7u);
    p.ro  = b.alloc<uint16_t>((nvo + 1u + 7u) & ~7u);
    p.vf  = b.alloc<uint16_t>((nent + 7u) & ~7u);
    p.x   = b.alloc<float>(nv);
    p.y   = b.alloc<float>(nv);
    p.z   = b.alloc<float>(nv);
    return p;
}

template <uint32_t T>
__global__ __launch_bounds__(T) void k_vertex_normals_plan(DevMesh m, int mode, const uint32_t* __restrict__ fbase,
                                                           const uint32_t* __restrict__ rbase,
                                                           const uint32_t* __restrict__ nbase,
                                                           const uint16_t* __restrict__ fvl,
                                                           const uint16_t* __restrict__ roff,
                                                           const uint16_t* __restrict__ vf,
                                                           const float* __restrict__ pos, float* __restrict__ nrm)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d = load_desc(m.patch, p);
    const uint32_t  nb0 = nbase[p], nent = nbase[p + 1] - nb0;
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanVNP   L = plan_vnp(b, d.nv, d.nvo, d.nf, nent);
    const uint32_t  N = m.nV;
    stage_u16x8<T>(L.fvl, fvl + fbase[p], 3 * d.nf);
    stage_u16x8<T>(L.ro, roff + rbase[p], d.nvo + 1);
    stage_u16x8<T>(L.vf, vf + nb0, nent);
    stage_positions_batched<T, 512u / T, false>(L.x, L.y, L.z, nullptr, pos, N, d, m.v_lin);
    __syncthreads();
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        float sx = 0.f, sy = 0.f, sz = 0.f;
        for (uint32_t k = L.ro[v]; k < L.ro[v + 1]; ++k) {
            const uint32_t s = L.vf[k], f = s >> 2, c = s & 3u;
            const uint32_t a = L.fvl[3 * f], bb = L.fvl[3 * f + 1], cc = L.fvl[3 * f + 2];
            const float    ux = L.x[bb] - L.x[a], uy = L.y[bb] - L.y[a], uz = L.z[bb] - L.z[a];
            const float    wx = L.x[cc] - L.x[a], wy = L.y[cc] - L.y[a], wz = L.z[cc] - L.z[a];
            const float    nx = uy * wz - uz * wy, ny = uz * wx - ux * wz, nz = ux * wy - uy * wx;
            float          w = 0.5f;
            if (mode == 0) {
                const float vx = L.x[cc] - L.x[bb], vy = L.y[cc] - L.y[bb], vz = L.z[cc] - L.z[bb];
                const float l0 = ux * ux + uy * uy + uz * uz;  // a-b
                const float l1 = vx * vx + vy * vy + vz * vz;  // b-c
                const float l2 = wx * wx + wy * wy + wz * wz;  // c-a
                const float den = (c == 0u) ? (l0 + l2) : ((c == 1u) ? (l1 + l0) : (l2 + l1));
                w               = __builtin_amdgcn_rcpf(den);  // 1 ulp: well inside the weights' use
            }
            sx += nx * w;
            sy += ny * w;
            sz += nz * w;
        }
        const uint32_t g = d.vp + v;
        nrm[g]           = sx;
        nrm[N + g]       = sy;
        nrm[2 * N + g]   = sz;
    }
}

// ---- Gaussian curvature over the patch-local face plan -------------------------
// Restates apps/GaussianCurvature/gaussian_curvature_kernel.cuh:19-62: per
// face s = |(c1 - c0) x (c2 - c0)|, corner dots c[i], squared edge lengths
// l[i] = |v_i - v_{i+1}|^2, angles atan2(s, c[i]); the mixed area of a corner
// is s/4 (obtuse corner of an obtuse face), s/8 (other corners of an obtuse
// face) or (l[i+2] c[i+1] / s + l[i] c[i+2] / s) / 8; the corner's angle is
// subtracted from 2 pi (obtuse angle <=> negative corner dot product). The
// reference scatters these with float atomics; here
// each owned vertex sums its incident faces (VF row of the face plan,
// ascending local face order) itself, so the result is deterministic and
// needs no atomics. gc = defect / mixed area; amix (optional) the mixed area.
// Angle in [0, pi] of a triangle corner from s = |u x v| >= 0 and c = u . v
// (atan2(s, c), both scaled alike). atan of the ratio min/max on [0, 1] is an
// odd polynomial t P(t^2) whose coefficients were fitted here by iteratively
// reweighted least squares (max error 1.4e-7 rad in fp32 evaluation); the
// result is folded into its octant: pi/2 - a when s > |c|, pi - a when c < 0.
__device__ __forceinline__ float corner_angle(float s, float c)
{
    const float ac = fabsf(c), mx = fmaxf(s, ac), mn = fminf(s, ac);
    const float t  = (mx > 0.f) ? mn * __builtin_amdgcn_rcpf(mx) : 0.f;
    const float u  = t * t;
    float       p  = -0.004054558929055929f;
    p              = p * u + 0.021862929686903954f;
    p              = p * u - 0.05591228976845741f;
    p              = p * u + 0.09642194956541061f;
    p              = p * u - 0.1390862911939621f;
    p              = p * u + 0.19946566224098206f;
    p              = p * u - 0.33329859375953674f;
    p              = p * u + 0.9999993443489075f;
    float a        = p * t;
    if (s > ac) a = 1.57079632679489662f - a;
    if (c < 0.f) a = 3.14159265358979324f - a;
    return a;
}

// LDS carve of the Gaussian-curvature kernel: the face plan (FV table, VF
// rows), positions, and per face corner the angle and the mixed-area share
// (3 floats each per face).
struct PlanGC
{
    uint16_t *fvl, *ro, *vf;
    float *   x, *y, *z, *ang, *amx;
};
__host__ __device__ inline PlanGC plan_gc(Bump& b, uint32_t nv, uint32_t nvo, uint32_t nf, uint32_t nent)
{
    PlanGC p{};
    p.fvl = b.alloc<uint16_t>((3 * nf + 7u) & ~7u);
    p.ro  = b.alloc<uint16_t>((nvo + 1u + 7u) & ~7u);
    p.vf  = b.alloc<uint16_t>((nent + 7u) & ~7u);
    p.x   = b.alloc<float>(nv);
    p.y   = b.alloc<float>(nv);
    p.z   = b.alloc<float>(nv);
    p.ang = b.alloc<float>(3 * nf);
    p.amx = b.alloc<float>(3 * nf);
    return p;
}

// Face pass: every local face evaluates its triangle once (cross product,
// corner dots, squared edge lengths) and stores each corner's angle and
// mixed-area share; vertex pass: each owned vertex sums its VF row's corner
// terms in ascending local face order.
template <uint32_t T>
__global__ __launch_bounds__(T) void k_gaussian_curvature(DevMesh m, const uint32_t* __restrict__ fbase,
                                                          const uint32_t* __restrict__ rbase,
                                                          const uint32_t* __restrict__ nbase,
                                                          const uint16_t* __restrict__ fvl,
                                                          const uint16_t* __restrict__ roff,
                                                          const uint16_t* __restrict__ vf,
                                                          const float* __restrict__ pos, float* __restrict__ gc,
                                                          float* __restrict__ amix)
{
    extern __shared__ uint4 smem[];
    const uint32_t p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d   = load_desc(m.patch, p);
    const uint32_t  nb0 = nbase[p], nent = nbase[p + 1] - nb0;
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanGC    L = plan_gc(b, d.nv, d.nvo, d.nf, nent);
    const uint32_t  N = m.nV;
    stage_u16x8<T>(L.fvl, fvl + fbase[p], 3 * d.nf);
    stage_u16x8<T>(L.ro, roff + rbase[p], d.nvo + 1);
    stage_u16x8<T>(L.vf, vf + nb0, nent);
    stage_positions_batched<T, (512u + T - 1u) / T, false>(L.x, L.y, L.z, nullptr, pos, N, d, m.v_lin);
    __syncthreads();
    for (uint32_t f = threadIdx.x; f < d.nf; f += T) {
        float px[3], py[3], pz[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const uint32_t q = L.fvl[3 * f + j];
            px[j]            = L.x[q];
            py[j]            = L.y[q];
            pz[j]            = L.z[q];
        }
        const float e1x = px[1] - px[0], e1y = py[1] - py[0], e1z = pz[1] - pz[0];
        const float e2x = px[2] - px[0], e2y = py[2] - py[0], e2z = pz[2] - pz[0];
        const float cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
        const float s = sqrtf(cx * cx + cy * cy + cz * cz);
        float       c[3], l[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int   j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            const float ax = px[j1] - px[j], ay = py[j1] - py[j], az = pz[j1] - pz[j];
            const float bx = px[j2] - px[j], by = py[j2] - py[j], bz = pz[j2] - pz[j];
            c[j]           = ax * bx + ay * by + az * bz;
            l[j]           = ax * ax + ay * ay + az * az;
        }
        // an angle exceeds pi / 2 exactly when its corner dot product is negative
        const bool  ob = (c[0] < 0.f) || (c[1] < 0.f) || (c[2] < 0.f);
        const float rs = __builtin_amdgcn_rcpf(s);  // cot = dot / |cross| for the non-obtuse case
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            L.ang[3 * f + i] = corner_angle(s, c[i]);
            L.amx[3 * f + i] = ob ? ((c[i] < 0.f) ? 0.25f * s : 0.125f * s)
                                  : 0.125f * (l[i2] * (c[i1] * rs) + l[i] * (c[i2] * rs));
        }
    }
    __syncthreads();
    const float two_pi = 6.28318530717958647692f;
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        float g = two_pi, am = 0.f;
        for (uint32_t k = L.ro[v]; k < L.ro[v + 1]; ++k) {
            const uint32_t sl = L.vf[k], fc = 3u * (sl >> 2) + (sl & 3u);  // corner (sl & 3) of face sl >> 2
            am += L.amx[fc];
            g -= L.ang[fc];
        }
        const uint32_t gg = d.vp + v;
        gc[gg]            = g / am;
        if (amix) amix[gg] = am;
    }
}

// ---- permutations between global (input) order and linear order ----------
__global__ void k_gather_rows(uint32_t n, uint32_t ncomp, const uint32_t* __restrict__ map,
                              const float* __restrict__ src, float* __restrict__ dst)
{
    // dst[c * n + i] = src[c * n + map[i]]
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = map[i];
    for (uint32_t c = 0; c < ncomp; ++c) dst[c * n + i] = src[c * n + s];
}

}  // namespace rxm
