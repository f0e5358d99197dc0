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

// ---- uniform Laplacian smoothing (one Jacobi step) -------------------------
struct PlanSM
{
    uint32_t *ev, *off, *cur, *wave;
    uint16_t* val;
    float *   x, *y, *z;
};
__host__ __device__ inline PlanSM plan_sm(Bump& b, uint32_t nv, uint32_t ne)
{
    PlanSM p{};
    p.ev   = b.alloc<uint32_t>(r4(ne));
    p.off  = b.alloc<uint32_t>(nv + 1);
    p.cur  = b.alloc<uint32_t>(nv);
    p.wave = b.alloc<uint32_t>(20);
    p.val  = b.alloc<uint16_t>(2 * ne);
    p.x    = b.alloc<float>(nv);
    p.y    = b.alloc<float>(nv);
    p.z    = b.alloc<float>(nv);
    return p;
}

// x' = x + lambda * (mean of VV neighbours - x); neighbours summed in
// ascending local-edge order.
__global__ __launch_bounds__(kGeomThreads) void k_smooth_uniform(DevMesh m, float lambda, const float* __restrict__ pos,
                                                                 float* __restrict__ out)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = blockIdx.x;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanSM    L = plan_sm(b, d.nv, d.ne);
    const uint32_t  T = blockDim.x, N = m.nV;
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    stage_positions(L.x, L.y, L.z, pos, N, d, m.v_lin);
    __syncthreads();
    lds_transpose<2>(2 * d.ne, d.nv,
                     [&](uint32_t i) {
                         const uint32_t x = L.ev[i >> 1];
                         return (i & 1u) ? hi16(x) : lo16(x);
                     },
                     L.off, L.cur, L.val, L.wave, d.nvo);
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        float          sx = 0.f, sy = 0.f, sz = 0.f;
        const uint32_t k0 = L.off[v], k1 = L.off[v + 1];
        for (uint32_t k = k0; k < k1; ++k) {
            const uint32_t x = L.ev[L.val[k]];
            const uint32_t w = (lo16(x) == v) ? hi16(x) : lo16(x);
            sx += L.x[w];
            sy += L.y[w];
            sz += L.z[w];
        }
        const float    inv = (k1 > k0) ? 1.0f / static_cast<float>(k1 - k0) : 0.f;
        const uint32_t g   = d.vp + v;
        const float    px = L.x[v], py = L.y[v], pz = L.z[v];
        out[g]         = (k1 > k0) ? px + lambda * (sx * inv - px) : px;
        out[N + g]     = (k1 > k0) ? py + lambda * (sy * inv - py) : py;
        out[2 * N + g] = (k1 > k0) ? pz + lambda * (sz * inv - pz) : pz;
    }
}

// ---- persistent, prefetching pipeline ---------------------------------------
// A grid of (CUs x k) workgroups walks the patches; workgroup b sits on XCD
// b % 8 and takes patches from that XCD's contiguous eighth of the patch
// range, so neighbouring patches (which share ribbon vertices) share an L2.
// While one patch is transposed and reduced in LDS, the next patch's EV words,
// owned positions and ribbon owner ids are already in flight into registers.
struct XcdRange
{
    uint32_t first, end, step;
};
__device__ __forceinline__ XcdRange xcd_range(uint32_t P)
{
    const uint32_t G = gridDim.x, x = blockIdx.x & 7u, j = blockIdx.x >> 3;
    const uint32_t per = (G + 7u - x) >> 3;  // workgroups on XCD x
    const uint32_t b = static_cast<uint32_t>((uint64_t(P) * x) >> 3);
    const uint32_t e = static_cast<uint32_t>((uint64_t(P) * (x + 1u)) >> 3);
    return {b + j, e, per};
}

struct PlanSM2
{
    uint32_t *ev, *off, *cnt, *wave;
    uint16_t* val;
    float *   x, *y, *z;
};
__host__ __device__ inline PlanSM2 plan_sm2(Bump& b, uint32_t nv, uint32_t ne, uint32_t nw)
{
    PlanSM2 p{};
    p.ev   = b.alloc<uint32_t>(r4(ne));
    p.off  = b.alloc<uint32_t>(nv + 1);
    p.cnt  = b.alloc<uint32_t>(nw * nv);
    p.wave = b.alloc<uint32_t>(20);
    p.val  = b.alloc<uint16_t>(2 * ne);
    p.x    = b.alloc<float>(nv);
    p.y    = b.alloc<float>(nv);
    p.z    = b.alloc<float>(nv);
    return p;
}

// Register-staged copy of one patch: RE uint4 of EV and RV vertex slots per
// thread (host guarantees ne <= 4 * RE * T and nv <= RV * T).
template <uint32_t T, uint32_t RE, uint32_t RV>
struct SmoothStage
{
    PatchDesc d;
    uint4     ev[RE];
    float     x[RV], y[RV], z[RV];
    uint32_t  lin[RV];

    __device__ __forceinline__ void load(const DevMesh& m, const float* __restrict__ pos, uint32_t p)
    {
        d                 = load_desc(m.patch, p);
        const uint4*  s   = reinterpret_cast<const uint4*>(m.ev + d.e_off);
        const uint32_t n4 = (d.ne + 3u) >> 2;
#pragma unroll
        for (uint32_t r = 0; r < RE; ++r) {
            const uint32_t i = threadIdx.x + r * T;
            if (i < n4) ev[r] = s[i];
        }
        const uint32_t N = m.nV;
#pragma unroll
        for (uint32_t r = 0; r < RV; ++r) {
            const uint32_t l = threadIdx.x + r * T;
            if (l < d.nvo) {
                const uint32_t g = d.vp + l;
                x[r] = pos[g];
                y[r] = pos[N + g];
                z[r] = pos[2 * N + g];
            } else if (l < d.nv) {
                lin[r] = m.v_lin[d.v_off + l];
            }
        }
    }
};

template <uint32_t T, uint32_t RE, uint32_t RV>
__global__ __launch_bounds__(T) void k_smooth_persistent(DevMesh m, float lambda, const float* __restrict__ pos,
                                                         float* __restrict__ out)
{
    extern __shared__ uint4 smem[];
    const uint32_t          N  = m.nV;
    const XcdRange          xr = xcd_range(m.P);
    SmoothStage<T, RE, RV>  st;
    if (xr.first < xr.end) st.load(m, pos, xr.first);
    for (uint32_t p = xr.first; p < xr.end; p += xr.step) {
        const PatchDesc d = st.d;
        Bump            b(reinterpret_cast<char*>(smem));
        const PlanSM2   L = plan_sm2(b, d.nv, d.ne, T / 64u);
        const uint32_t  n4 = (d.ne + 3u) >> 2;
#pragma unroll
        for (uint32_t r = 0; r < RE; ++r) {
            const uint32_t i = threadIdx.x + r * T;
            if (i < n4) reinterpret_cast<uint4*>(L.ev)[i] = st.ev[r];
        }
        float gx[RV], gy[RV], gz[RV];
#pragma unroll
        for (uint32_t r = 0; r < RV; ++r) {
            const uint32_t l = threadIdx.x + r * T;
            if (l < d.nvo) {
                L.x[l] = st.x[r];
                L.y[l] = st.y[r];
                L.z[l] = st.z[r];
            } else if (l < d.nv) {  // ribbon: gather from the owner's slot
                const uint32_t g = st.lin[r];
                gx[r] = pos[g];
                gy[r] = pos[N + g];
                gz[r] = pos[2 * N + g];
            }
        }
        if (p + xr.step < xr.end) st.load(m, pos, p + xr.step);
        __syncthreads();
        lds_transpose_ordered<2>(2 * d.ne, d.nv, bits_for(d.nv),
                                 [&](uint32_t i) {
                                     const uint32_t x = L.ev[i >> 1];
                                     return (i & 1u) ? hi16(x) : lo16(x);
                                 },
                                 L.off, L.cnt, L.val, L.wave);
#pragma unroll
        for (uint32_t r = 0; r < RV; ++r) {
            const uint32_t l = threadIdx.x + r * T;
            if (l >= d.nvo && l < d.nv) {
                L.x[l] = gx[r];
                L.y[l] = gy[r];
                L.z[l] = gz[r];
            }
        }
        __syncthreads();
        for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
            float          sx = 0.f, sy = 0.f, sz = 0.f;
            const uint32_t k0 = L.off[v], k1 = L.off[v + 1];
            for (uint32_t k = k0; k < k1; ++k) {
                const uint32_t x = L.ev[L.val[k]];
                const uint32_t w = (lo16(x) == v) ? hi16(x) : lo16(x);
                sx += L.x[w];
                sy += L.y[w];
                sz += L.z[w];
            }
            const float    inv = (k1 > k0) ? 1.0f / static_cast<float>(k1 - k0) : 0.f;
            const uint32_t g   = d.vp + v;
            const float    px = L.x[v], py = L.y[v], pz = L.z[v];
            out[g]         = (k1 > k0) ? px + lambda * (sx * inv - px) : px;
            out[N + g]     = (k1 > k0) ? py + lambda * (sy * inv - py) : py;
            out[2 * N + g] = (k1 > k0) ? pz + lambda * (sz * inv - pz) : pz;
        }
        __syncthreads();
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
