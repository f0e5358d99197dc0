// Numerics on top of the query engine: cotangent CSR assembly over the
// oriented VV ring, SpMV-3 with the fused explicit mean-curvature-flow update,
// and the symmetric Dirichlet energy with its hand-written adjoint.
#pragma once
#include <cfloat>

#include "k_geometry.hpp"

namespace rxm {

struct f3
{
    float x, y, z;
};
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// reference: include/rxmesh/geometry_util.cuh:50-58 (tri_area)
__device__ __forceinline__ float tri_area(f3 p0, f3 p1, f3 p2)
{
    const f3 c = cross(sub(p1, p0), sub(p2, p0));
    return 0.5f * sqrtf(dot(c, c));
}
// reference: geometry_util.cuh:107-114 (clamp_cot, +-19.1)
__device__ __forceinline__ float clamp_cot(float v) { return (v < -19.1f) ? -19.1f : ((v > 19.1f) ? 19.1f : v); }

// reference: geometry_util.cuh:120-172 (partial_voronoi_area of p in CCW p->q->r)
__device__ __forceinline__ float partial_voronoi_area(f3 p, f3 q, f3 r)
{
    const f3    pq = sub(q, p), qr = sub(r, q), pr = sub(r, p);
    const float area = tri_area(p, q, r);
    if (area <= FLT_MIN) return -1.f;
    const float dotp = dot(pq, pr), dotq = -dot(qr, pq), dotr = dot(qr, pr);
    if (dotp < 0.f) return 0.25f * area;
    if (dotq < 0.f || dotr < 0.f) return 0.125f * area;
    const float cotq = clamp_cot(dotq / area), cotr = clamp_cot(dotr / area);
    return 0.125f * (dot(pr, pr) * cotq + dot(pq, pq) * cotr);
}

// reference: geometry_util.cuh:178-206 (edge_cotan_weight of edge p-r with
// opposite vertices q and s; "cot" = dot / area)
__device__ __forceinline__ float edge_cotan_weight(f3 p, f3 r, f3 q, f3 s)
{
    auto part = [&](f3 v) {
        const f3    d0 = sub(p, v), d1 = sub(r, v);
        const float area = tri_area(p, r, v);
        return (area > FLT_MIN) ? clamp_cot(dot(d0, d1) / area) : 0.f;
    };
    return part(q) + part(s);
}

// ---- cotangent CSR assembly ------------------------------------------------
// One wedge (p, a, b) of a vertex's ring: the clamped cotangents of the
// angles at a and at b (reference "cot" = dot / tri_area) and the wedge's
// partial Voronoi area at p (geometry_util.cuh:120-172, clamped at 0). Each
// wedge enters two edge weights and one area, so a ring of K edges costs K
// wedges (one reciprocal square root each) instead of 3K triangle areas.
// |cross|^2 <= FLT_MIN counts as degenerate (area 0): cotangents 0 and no
// Voronoi share, as the reference's area <= FLT_MIN tests give.
struct Wedge
{
    float ca, cb, vor;
};
__device__ __forceinline__ Wedge wedge(f3 p, f3 a, f3 b)
{
    const f3    pa = sub(a, p), pb = sub(b, p), ab = sub(b, a);
    const f3    c  = cross(pa, pb);
    const float cc = dot(c, c);
    Wedge       w{0.f, 0.f, 0.f};
    if (!(cc > FLT_MIN)) return w;
    const float rs   = __builtin_amdgcn_rsqf(cc);
    const float area = 0.5f * cc * rs;  // tri_area
    const float inv  = 2.0f * rs;       // 1 / tri_area
    const float dp = dot(pa, pb), da = -dot(pa, ab), db = dot(pb, ab);
    w.ca = clamp_cot(da * inv);
    w.cb = clamp_cot(db * inv);
    if (dp < 0.f) w.vor = 0.25f * area;
    else if (da < 0.f || db < 0.f) w.vor = 0.125f * area;
    else w.vor = fmaxf(0.125f * (dot(pb, pb) * w.ca + dot(pa, pa) * w.cb), 0.f);
    return w;
}

// Per-face terms of the reference weights (geometry_util.cuh:107-206) for
// face (a0, a1, a2): cot[j] = the clamped cotangent of the angle at a_j
// ("cot" = dot / tri_area), which enters the weight of the opposite edge
// (face slot j + 1, a_{j+1} -> a_{j+2}); vor[j] = the partial Voronoi area of
// a_j in the face (partial_voronoi_area(a_j, a_{j+1}, a_{j+2}) clamped at 0).
// |cross|^2 <= FLT_MIN counts as degenerate: all terms 0, as the
// reference's area <= FLT_MIN tests give.
struct FaceCot
{
    float cot[3], vor[3];
};
__device__ __forceinline__ FaceCot face_cot(f3 P0, f3 P1, f3 P2)
{
    const f3    E0 = sub(P1, P0), E1 = sub(P2, P1), E2 = sub(P0, P2);
    const f3    c  = cross(E0, E2);
    const float cc = dot(c, c);
    FaceCot     r{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    if (!(cc > FLT_MIN)) return r;
    const float rs = __builtin_amdgcn_rsqf(cc);
    const float area = 0.5f * cc * rs, inv = 2.0f * rs;
    const float D[3] = {-dot(E0, E2), -dot(E1, E0), -dot(E2, E1)};  // corner dots
    const float L[3] = {dot(E0, E0), dot(E1, E1), dot(E2, E2)};     // |a_{j+1} - a_j|^2
#pragma unroll
    for (int j = 0; j < 3; ++j) r.cot[j] = clamp_cot(D[j] * inv);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        if (D[j] < 0.f) r.vor[j] = 0.25f * area;
        else if (D[j1] < 0.f || D[j2] < 0.f) r.vor[j] = 0.125f * area;
        else r.vor[j] = fmaxf(0.125f * (L[j2] * r.cot[j1] + L[j] * r.cot[j2]), 0.f);
    }
    return r;
}

// LDS carve of the assembly: float4 positions and linear ids of the patch
// vertices, the EV slab, per-edge half-edge terms he[e] = {cot(2e), vor(2e),
// cot(2e + 1), vor(2e + 1)} (half-edge 2e + dir starts at the edge's
// (dir ? hi16 : lo16) endpoint; cot = the clamped cotangent of the angle
// opposite it in its face, vor = the partial Voronoi area of its start in
// that face; zeros where no face of the patch holds it), and the patch's
// pattern: row starts (nvo + 1) and entries (nent).
struct PlanCOT
{
    float4*   pos;
    float2 *  cot, *vor;  // per edge: the two half-edges' opposite cotangents / start Voronoi terms
    uint32_t *lin, *ev, *rs;
    uint16_t* ent;
};
__host__ __device__ inline PlanCOT plan_cot(Bump& b, uint32_t nv, uint32_t ne, uint32_t nvo, uint32_t nent)
{
    PlanCOT p{};
    p.pos = b.alloc<float4>(nv);
    p.cot = b.alloc<float2>(ne);
    p.vor = b.alloc<float2>(ne);
    p.lin = b.alloc<uint32_t>(nv);
    p.ev  = b.alloc<uint32_t>(r4(ne));
    p.rs  = b.alloc<uint32_t>(nvo + 1);
    p.ent = b.alloc<uint16_t>(nent);
    return p;
}

// Copy n u32 global -> LDS (any alignment), R loads of a thread in flight
// before its stores; dst[i] = src[i] - sub.
template <uint32_t T, uint32_t R>
__device__ __forceinline__ void stage_words(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                            uint32_t n, uint32_t sub = 0)
{
    for (uint32_t b0 = 0; b0 < n; b0 += R * T) {
        uint32_t x[R];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            const uint32_t i = b0 + r * T + threadIdx.x;
            x[r]             = (i < n) ? src[i] : 0u;
        }
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            const uint32_t i = b0 + r * T + threadIdx.x;
            if (i < n) dst[i] = x[r] - sub;
        }
    }
}

// Copy n u16 global -> LDS (any alignment), R loads of a thread in flight
// before its stores.
template <uint32_t T, uint32_t R>
__device__ __forceinline__ void stage_halves(uint16_t* __restrict__ dst, const uint16_t* __restrict__ src, uint32_t n)
{
    for (uint32_t b0 = 0; b0 < n; b0 += R * T) {
        uint16_t x[R];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            const uint32_t i = b0 + r * T + threadIdx.x;
            x[r]             = (i < n) ? src[i] : uint16_t(0);
        }
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            const uint32_t i = b0 + r * T + threadIdx.x;
            if (i < n) dst[i] = x[r];
        }
    }
}

// Stage, per patch vertex, its linear id (owned by prefix, ribbon from the
// owner table) and position; all global loads of a thread are issued before
// its LDS stores.
template <uint32_t T>
__device__ __forceinline__ void stage_vertices(const DevMesh& m, const PatchDesc& d, const float* __restrict__ pos,
                                               float4* s_pos, uint32_t* s_lin)
{
    constexpr uint32_t R  = 2;
    const uint32_t     N  = m.nV;
    const uint32_t*    vl = m.v_lin + d.v_off;
    uint32_t           g[R];
    float              x[R], y[R], z[R];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        const uint32_t l = threadIdx.x + r * T;
        g[r]             = (l < d.nvo) ? d.vp + l : ((l < d.nv) ? vl[l] : 0u);
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
            s_pos[l] = make_float4(x[r], y[r], z[r], 0.f);
            s_lin[l] = g[r];
        }
    }
    for (uint32_t l = threadIdx.x + R * T; l < d.nv; l += T) {
        const uint32_t gg = lin_of(l, d.nvo, d.vp, vl);
        s_pos[l]          = make_float4(pos[gg], pos[N + gg], pos[2 * N + gg], 0.f);
        s_lin[l]          = gg;
    }
}

// Rows of A = M + dt L over linear ids (apps/MCF/mcf_kernels.cuh:354-426,
// mcf_A_setup): the diagonal, then the neighbours in VV query order
// (ascending local id of the connecting edge), the column order SparseMatrix
// fills (sparse_matrix_kernels.cuh:75-140); mass[v] = 1 / v_weight.
// The pattern is static, like the reference's SparseMatrix: `ent` holds, in
// VV query order, each owned row's entries of the patch as (local edge) |
// (local row << 16), `rs` each linear vertex's first entry (patch-relative
// offsets come from vv_base, the VV query's per-patch base; the system's
// base is vv_base + the vertex prefix, one diagonal per owned row).
// Every face of the patch is evaluated once and stores, for each of its
// half-edges, the cotangent opposite it and the Voronoi term of its start
// (plain stores: a half-edge has one face on a consistently oriented
// manifold). An edge's weight is cot(2e) + cot(2e + 1) -- the reference's
// part(q) + part(s), in an order floating-point addition does not see --
// and a row's area sums the Voronoi terms of its outgoing half-edges. This
// equals the reference's walk over the oriented ring for every vertex whose
// ring is one closed fan; the rows of other vertices (boundary, several
// fans) take wedges across the ring's gaps, as the reference's cyclic walk
// does, and are rewritten afterwards by k_cot_ring_rows.
template <uint32_t T>
__global__ __launch_bounds__(T) void k_cot_assemble(DevMesh m, const uint32_t* __restrict__ vv_base,
                                                    const uint32_t* __restrict__ rs,
                                                    const uint16_t* __restrict__ ent, const float* __restrict__ pos,
                                                    float dt, int32_t* __restrict__ row_ptr,
                                                    int32_t* __restrict__ col, float* __restrict__ val,
                                                    float* __restrict__ mass)
{
    extern __shared__ uint4 smem[];
    const uint32_t p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d    = load_desc(m.patch, p);
    const uint32_t  vb   = vv_base[p], nent = vv_base[p + 1] - vb;
    const uint32_t  bse  = vb + d.vp;  // first entry of the patch's rows
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanCOT   L = plan_cot(b, d.nv, d.ne, d.nvo, nent);
    // the FE words of this thread's faces, loaded with the staging's loads
    constexpr uint32_t RF = (1024u + T - 1u) / T;
    const uint16_t*    fe = m.fe + 3u * d.f_off;
    uint32_t           fw[RF][3];
#pragma unroll
    for (uint32_t r = 0; r < RF; ++r)
#pragma unroll
        for (uint32_t j = 0; j < 3; ++j) fw[r][j] = (threadIdx.x + r * T < d.nf) ? fe[3 * (threadIdx.x + r * T) + j] : 0u;
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    stage_halves<T, 8>(L.ent, ent + vb, nent);
    stage_words<T, 2>(L.rs, rs + d.vp, d.nvo + 1, vb);
    stage_vertices<T>(m, d, pos, L.pos, L.lin);
    for (uint32_t i = threadIdx.x; i < d.ne; i += T) {
        L.cot[i] = make_float2(0.f, 0.f);
        L.vor[i] = make_float2(0.f, 0.f);
    }
    __syncthreads();
    // per face: the terms of its half-edges, stored at the half-edge's FE
    // word 2 e + dir (float half dir of edge e's pair)
    float* hc = reinterpret_cast<float*>(L.cot);
    float* hv = reinterpret_cast<float*>(L.vor);
    auto   face = [&](const uint32_t (&x)[3]) {
        uint32_t a[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const uint32_t e = L.ev[x[j] >> 1];
            a[j]             = (x[j] & 1u) ? hi16(e) : lo16(e);
        }
        const float4  q0 = L.pos[a[0]], q1 = L.pos[a[1]], q2 = L.pos[a[2]];
        const FaceCot c  = face_cot(f3{q0.x, q0.y, q0.z}, f3{q1.x, q1.y, q1.z}, f3{q2.x, q2.y, q2.z});
        // slot j runs a_j -> a_{j+1}: opposite corner a_{j+2}, start a_j
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            hc[x[j]] = c.cot[(j + 2) % 3];
            hv[x[j]] = c.vor[j];
        }
    };
#pragma unroll
    for (uint32_t r = 0; r < RF; ++r)
        if (threadIdx.x + r * T < d.nf) face(fw[r]);
    for (uint32_t f = threadIdx.x + RF * T; f < d.nf; f += T) {
        const uint32_t x[3] = {fe[3 * f], fe[3 * f + 1], fe[3 * f + 2]};
        face(x);
    }
    __syncthreads();
    // rows: diagonal and mass (entries in pattern order)
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        const uint32_t k0 = L.rs[v], k1 = L.rs[v + 1];
        float          sum_w = 0.f, vw = 0.f;
        for (uint32_t k = k0; k < k1; ++k) {
            const uint32_t u = L.ent[k];  // the half-edge leaving v
            const float2   c = L.cot[u >> 1];
            const float    w = c.x + c.y;
            sum_w += ((w >= 0.f) ? w : 0.f) * dt;
            vw += hv[u];
        }
        const float    ms = 1.0f / (0.5f / vw);
        const uint32_t o  = bse + k0 + v;
        row_ptr[d.vp + v] = static_cast<int32_t>(o);
        mass[d.vp + v]    = ms;
        col[o]            = static_cast<int32_t>(d.vp + v);
        val[o]            = ms + sum_w;
    }
    if (threadIdx.x == 0) row_ptr[d.vp + d.nvo] = static_cast<int32_t>(bse + nent + d.nvo);
    // off-diagonal entries, contiguous over the patch's output range
    for (uint32_t k = threadIdx.x; k < nent; k += T) {
        const uint32_t u = L.ent[k], e = u >> 1, x = L.ev[e], v = (u & 1u) ? hi16(x) : lo16(x);
        const float2   c = L.cot[e];
        const float    w = c.x + c.y;
        const uint32_t o = bse + k + v + 1;
        col[o]           = static_cast<int32_t>(L.lin[(u & 1u) ? lo16(x) : hi16(x)]);
        val[o]           = -(((w >= 0.f) ? w : 0.f) * dt);
    }
}

// LDS carve of the ring-row pass: EV, FE, half-edge successors, global ids,
// column offsets, the owned rows (sorted) and a scratch copy of one row's
// edges per row (ordered into the canonical ring in place).
struct PlanCOTR
{
    uint32_t *ev, *gid, *off, *wave, *slots, *ring;
    uint16_t *fe, *hn;
};
__host__ __device__ inline PlanCOTR plan_cot_ring(Bump& b, uint32_t nv, uint32_t ne, uint32_t nf, uint32_t nvo,
                                                  uint32_t nrow)
{
    PlanCOTR p{};
    p.ev    = b.alloc<uint32_t>(r4(ne));
    p.gid   = b.alloc<uint32_t>(nv);
    p.off   = b.alloc<uint32_t>(nvo + 1);
    p.wave  = b.alloc<uint32_t>(20);
    p.slots = b.alloc<uint32_t>(nrow);
    p.ring  = b.alloc<uint32_t>(nrow);
    p.fe    = b.alloc<uint16_t>(r4(3 * nf));
    p.hn    = b.alloc<uint16_t>(2 * ne);
    return p;
}

// Rows of the vertices whose ring is not one closed fan (flag bit clear in
// `closed`, per linear id), for the patches listed in `plist`: the
// reference's cyclic walk over the canonical oriented ring (wedge (p, r_i,
// r_{i+1}) for every consecutive pair, the last one wrapping to the first),
// each weight written to its neighbour's column of the row k_cot_assemble
// laid out (neighbours ascending by local id of the connecting edge). Positions come from global
// memory: these rows are few.
template <uint32_t T>
__global__ __launch_bounds__(T) void k_cot_ring_rows(DevMesh m, const uint32_t* __restrict__ base,
                                                     const uint32_t* __restrict__ plist,
                                                     const uint32_t* __restrict__ closed,
                                                     const float* __restrict__ pos, float dt,
                                                     float* __restrict__ val, float* __restrict__ mass)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p   = plist[blockIdx.x];
    const PatchDesc d   = load_desc(m.patch, p);
    const uint32_t  bse = base[p], nrow = base[p + 1] - bse;
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanCOTR  L = plan_cot_ring(b, d.nv, d.ne, d.nf, d.nvo, nrow);
    const uint32_t  N = m.nV;
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
    for (uint32_t v = threadIdx.x; v < d.nv; v += T) L.gid[v] = m.ltog_v[d.v_off + v];
    for (uint32_t i = threadIdx.x; i < d.ne; i += T) reinterpret_cast<uint32_t*>(L.hn)[i] = 0xFFFFFFFFu;
    __syncthreads();
    build_half_next(L.fe, L.ev, d.nf, L.hn);
    fill_columns<T>(
        2 * d.ne, d.nvo,
        [&](uint32_t i) {
            const uint32_t x = L.ev[i >> 1];
            return (i & 1u) ? hi16(x) : lo16(x);
        },
        [&](uint32_t i) {
            const uint32_t x = L.ev[i >> 1];
            return ((i >> 1) << 16) | ((i & 1u) ? lo16(x) : hi16(x));
        },
        L.off, L.slots, L.wave);
    if (L.off[d.nvo] + d.nvo != nrow) return;
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        const uint32_t lv = d.vp + v;
        if ((closed[lv >> 5] >> (lv & 31u)) & 1u) continue;
        const uint32_t k0 = L.off[v], K = L.off[v + 1] - k0;
        sort_row(L.slots + k0, K, [&](uint32_t j, uint32_t s) {
            L.slots[k0 + j] = s;
            L.ring[k0 + j]  = s >> 16;
        });
        orient_ring<12, true>(v, k0, k0 + K, L.ring, L.ring, L.ev, L.hn, L.gid);
        auto P3 = [&](uint32_t l) {
            const uint32_t g = lin_of(l, d.nvo, d.vp, m.v_lin + d.v_off);
            return f3{pos[g], pos[N + g], pos[2 * N + g]};
        };
        // column of neighbour r: its position in the sorted row
        auto rank = [&](uint32_t r) {
            uint32_t c = 0;
            for (uint32_t j = 0; j < K; ++j) c = ((L.slots[k0 + j] & 0xFFFFu) == r) ? j : c;
            return c;
        };
        const uint32_t o  = bse + k0 + v;
        const f3       pp = P3(v);
        float          sum_w = 0.f, vw = 0.f;
        if (K > 0) {
            // wedge i = (p, r_i, r_{i+1}); edge p-r_i takes the angle at
            // r_{i-1} of wedge i-1 and the angle at r_{i+1} of wedge i
            uint32_t    rc = L.ring[k0];
            f3          Pc = P3(rc);
            const Wedge wl = wedge(pp, P3(L.ring[k0 + K - 1]), Pc);
            float       ca_prev = wl.ca;
            vw                  = wl.vor;
            for (uint32_t i = 0; i < K; ++i) {
                uint32_t rn = 0;
                f3       Pn{};
                Wedge    w = wl;
                if (i + 1 < K) {
                    rn = L.ring[k0 + i + 1];
                    Pn = P3(rn);
                    w  = wedge(pp, Pc, Pn);
                    vw += w.vor;
                }
                float wt = ca_prev + w.cb;
                wt       = ((wt >= 0.f) ? wt : 0.f) * dt;
                sum_w += wt;
                val[o + 1 + rank(rc)] = -wt;
                ca_prev = w.ca;
                rc      = rn;
                Pc      = Pn;
            }
        }
        const float ms = 1.0f / (0.5f / vw);
        mass[lv]       = ms;
        val[o]         = ms + sum_w;
    }
}

// ---- SpMV over 3 SoA columns, optionally fused with the explicit MCF step --
// LANES consecutive lanes share a row (rows hold ~7 entries), so the col /
// val loads of a wave are contiguous; partial sums meet through a fixed
// xor-shuffle tree (deterministic).
//   FUSED == false: y = A x
//   FUSED == true : y = x + (M x - A x) / M
template <uint32_t LANES, bool FUSED>
__global__ __launch_bounds__(256) void k_spmv3(uint32_t n, const int32_t* __restrict__ row_ptr,
                                               const int32_t* __restrict__ col, const float* __restrict__ val,
                                               const float* __restrict__ mass, const float* __restrict__ x,
                                               float* __restrict__ y)
{
    const uint32_t gid  = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t row  = gid / LANES;
    const uint32_t lane = gid % LANES;
    float          ax = 0.f, ay = 0.f, az = 0.f;
    const bool     ok = row < n;
    if (ok) {
        const int32_t s = row_ptr[row], e = row_ptr[row + 1];
        for (int32_t k = s + static_cast<int32_t>(lane); k < e; k += LANES) {
            const uint32_t c = static_cast<uint32_t>(col[k]);
            const float    w = val[k];
            ax += w * x[c];
            ay += w * x[n + c];
            az += w * x[2 * n + c];
        }
    }
#pragma unroll
    for (uint32_t o = LANES / 2; o > 0; o >>= 1) {
        ax += __shfl_xor(ax, o, 64);
        ay += __shfl_xor(ay, o, 64);
        az += __shfl_xor(az, o, 64);
    }
    if (ok && lane == 0) {
        if constexpr (FUSED) {
            const float M = mass[row];
            const float px = x[row], py = x[n + row], pz = x[2 * n + row];
            y[row]         = px + (M * px - ax) / M;
            y[n + row]     = py + (M * py - ay) / M;
            y[2 * n + row] = pz + (M * pz - az) / M;
        } else {
            y[row]         = ax;
            y[n + row]     = ay;
            y[2 * n + row] = az;
        }
    }
}

// ---- explicit MCF step, patch-local ------------------------------------------
// y = x + (M x - A x) / M on the mesh's cotangent system (rxm_cot_csr_assemble
// layout: each patch's rows are one contiguous run, diagonal first, then the
// neighbours in VV query order). The pattern is the patch-local VV plan
// (u16 row offsets and local neighbour ids, the same order), so a workgroup
// stages its patch's positions (owned + ribbon), plan rows and the run of
// values in LDS and the matrix's column indices are never read: a row's
// neighbours come from LDS instead of one global gather per entry. Rows sum
// the diagonal term first, then the neighbours in row order.
template <uint32_t T>
__global__ __launch_bounds__(T) void k_mcf_patch(DevMesh m, const uint32_t* __restrict__ vv_base,
                                                 const uint32_t* __restrict__ rbase,
                                                 const uint32_t* __restrict__ nbase,
                                                 const uint16_t* __restrict__ roff, const uint16_t* __restrict__ nbr,
                                                 const float* __restrict__ val, const float* __restrict__ mass,
                                                 const float* __restrict__ x, float* __restrict__ y)
{
    extern __shared__ uint4 smem[];
    const uint32_t p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d    = load_desc(m.patch, p);
    // nbase runs are padded to 8 entries; the system's run is not: nent is
    // the patch's true VV entry count
    const uint32_t  rb   = rbase[p], nbb = nbase[p], nent = vv_base[p + 1] - vv_base[p];
    const uint32_t  bse  = vv_base[p] + d.vp;  // the patch's run in the system
    const uint32_t  N    = m.nV;
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanSL    L = plan_sl(b, d.nv, d.nvo, nent, false);
    float*          w = b.alloc<float>(nent + d.nvo);
    stage_u16x8<T>(L.ro, roff + rb, d.nvo + 1);
    stage_u16x8<T>(L.nb, nbr + nbb, nent);
    stage_words<T, 8>(reinterpret_cast<uint32_t*>(w), reinterpret_cast<const uint32_t*>(val) + bse, nent + d.nvo);
    stage_positions_batched<T, 512u / T, false>(L.x, L.y, L.z, nullptr, x, N, d, m.v_lin);
    __syncthreads();
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        const uint32_t k0 = L.ro[v], k1 = L.ro[v + 1];
        const float*   wr = w + k0 + v;  // diagonal, then the row's neighbours
        const float    px = L.x[v], py = L.y[v], pz = L.z[v];
        float          ax = wr[0] * px, ay = wr[0] * py, az = wr[0] * pz;
        for (uint32_t k = k0; k < k1; ++k) {
            const uint32_t q = L.nb[k];
            const float    c = wr[1 + k - k0];
            ax += c * L.x[q];
            ay += c * L.y[q];
            az += c * L.z[q];
        }
        const uint32_t g = d.vp + v;
        const float    M = mass[g];
        y[g]             = px + (M * px - ax) / M;
        y[N + g]         = py + (M * py - ay) / M;
        y[2 * N + g]     = pz + (M * pz - az) / M;
    }
}

// ---- implicit MCF: conjugate gradients on the cotangent system ---------------
// Restates the reference's CGSolver (include/rxmesh/matrix/cg_solver.h:37-125)
// on A X = B with B = M x (apps/MCF/mcf_kernels.cuh:302-352, mcf_B_setup) for
// the three coordinate columns at once: dot products and norms run over all
// 3V entries, as the reference's cuBLAS calls over the whole V x 3 matrix do.
// The loop runs on the device: the scalars live in CgState, every kernel
// returns at once when `done` is set, and the host launches max_iter
// iterations without reading anything back. Partial sums are per workgroup
// in a fixed order and reduced by one workgroup in a fixed order, so a solve
// is deterministic.
struct CgState
{
    float    delta, delta_old, alpha, beta, start;
    uint32_t iters, done, pad;
};

// y = A x (MODE 0, the set-up), or, for a CG iteration, p_new = r (MODE 1,
// the first) or r + beta p_old (MODE 2) staged for the patch's local
// vertices, y = A p_new, p_new written for the owned rows and <y, p_new>
// summed per workgroup into partial[p]. The pattern is the patch-local VV
// plan (the system's pattern; col is unused).
template <uint32_t T, int MODE>
__global__ __launch_bounds__(T) void k_cg_apply(DevMesh m, const uint32_t* __restrict__ vv_base,
                                                const uint32_t* __restrict__ rbase,
                                                const uint32_t* __restrict__ nbase,
                                                const uint16_t* __restrict__ roff, const uint16_t* __restrict__ nbr,
                                                const float* __restrict__ val, const float* __restrict__ x,
                                                const float* __restrict__ r, float* __restrict__ p_new,
                                                float* __restrict__ y, float* __restrict__ partial,
                                                const CgState* __restrict__ st)
{
    extern __shared__ uint4 smem[];
    const uint32_t p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    if (MODE != 0 && st->done) return;
    const PatchDesc d   = load_desc(m.patch, p);
    // true VV entry count (nbase runs are padded to 8; the system's run is not)
    const uint32_t  rb  = rbase[p], nbb = nbase[p], nent = vv_base[p + 1] - vv_base[p];
    const uint32_t  bse = vv_base[p] + d.vp;
    const uint32_t  N   = m.nV;
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanSL    L = plan_sl(b, d.nv, d.nvo, nent, false);
    float*          w = b.alloc<float>(nent + d.nvo);
    float*          red = b.alloc<float>(T / 64u);
    stage_u16x8<T>(L.ro, roff + rb, d.nvo + 1);
    stage_u16x8<T>(L.nb, nbr + nbb, nent);
    stage_words<T, 8>(reinterpret_cast<uint32_t*>(w), reinterpret_cast<const uint32_t*>(val) + bse, nent + d.nvo);
    if constexpr (MODE == 0) {
        stage_positions_batched<T, 512u / T, false>(L.x, L.y, L.z, nullptr, x, N, d, m.v_lin);
    } else {
        // x holds p_old here (unread in the first iteration)
        const float     beta = MODE == 2 ? st->beta : 0.f;
        const uint32_t* vl   = m.v_lin + d.v_off;
        for (uint32_t l = threadIdx.x; l < d.nv; l += T) {
            const uint32_t g = lin_of(l, d.nvo, d.vp, vl);
            if constexpr (MODE == 1) {
                L.x[l] = r[g];
                L.y[l] = r[N + g];
                L.z[l] = r[2 * N + g];
            } else {
                L.x[l] = r[g] + beta * x[g];
                L.y[l] = r[N + g] + beta * x[N + g];
                L.z[l] = r[2 * N + g] + beta * x[2 * N + g];
            }
        }
    }
    __syncthreads();
    float dotp = 0.f;
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        const uint32_t k0 = L.ro[v], k1 = L.ro[v + 1];
        const float*   wr = w + k0 + v;  // diagonal, then the row's neighbours
        const float    px = L.x[v], py = L.y[v], pz = L.z[v];
        float          ax = wr[0] * px, ay = wr[0] * py, az = wr[0] * pz;
        for (uint32_t k = k0; k < k1; ++k) {
            const uint32_t q = L.nb[k];
            const float    c = wr[1 + k - k0];
            ax += c * L.x[q];
            ay += c * L.y[q];
            az += c * L.z[q];
        }
        const uint32_t g = d.vp + v;
        y[g]             = ax;
        y[N + g]         = ay;
        y[2 * N + g]     = az;
        if constexpr (MODE != 0) {
            p_new[g]         = px;
            p_new[N + g]     = py;
            p_new[2 * N + g] = pz;
            dotp += ax * px + ay * py + az * pz;
        }
    }
    if constexpr (MODE != 0) {
        // fixed-order workgroup sum: wave shuffles, then wave totals in order
        for (uint32_t o = 32; o > 0; o >>= 1) dotp += __shfl_xor(dotp, o, 64);
        if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = dotp;
        __syncthreads();
        if (threadIdx.x == 0) {
            float t = 0.f;
            for (uint32_t i = 0; i < T / 64u; ++i) t += red[i];
            partial[p] = t;
        }
    }
}

// ---- matrix-free implicit MCF product -----------------------------------------
// LDS carve of the matrix-free product: the assembly's carve (positions and
// half-edge terms only for cotangent weights) plus the staged input vector.
struct PlanMF
{
    float4*   pos;
    float4*   he;
    uint32_t *lin, *ev, *rs;
    uint16_t* ent;
    float *   vx, *vy, *vz, *red;
};
__host__ __device__ inline PlanMF plan_mf(Bump& b, uint32_t nv, uint32_t ne, uint32_t nvo, uint32_t nent,
                                          bool uniform, uint32_t threads)
{
    PlanMF p{};
    if (!uniform) {
        p.pos = b.alloc<float4>(nv);
        p.he  = b.alloc<float4>(ne);
        p.lin = b.alloc<uint32_t>(nv);
    }
    p.ev  = b.alloc<uint32_t>(r4(ne));
    p.rs  = b.alloc<uint32_t>(nvo + 1);
    p.ent = b.alloc<uint16_t>(nent);
    p.vx  = b.alloc<float>(nv);
    p.vy  = b.alloc<float>(nv);
    p.vz  = b.alloc<float>(nv);
    p.red = b.alloc<float>(threads / 64u);
    return p;
}

// y = A v for the implicit MCF system A = M + dt L with the weights
// recomputed from the fixed coordinates in every product (the reference's
// matvec, apps/MCF/mcf_kernels.cuh:114-213): per owned row, in the order of
// the reference's ring loop, acc -= (dt w_e) v[r_e] over the row's edges,
// then y = acc + diag v[p] with diag = 1 / v_weight + sum_e dt w_e.
// Cotangent weights (UNIFORM false): w_e = max(0, cot + cot) of the edge and
// v_weight = 0.5 / (the vertex's Voronoi area), from one evaluation of every
// face of the patch (the terms k_cot_assemble computes); every vertex ring
// must be one closed fan (checked on the host). Uniform weights: w_e = 1,
// v_weight = 1 / valence, no orientation needed.
// MODE 0 (set-up): v = x; also writes mass[g] = 1 / v_weight (the
// reference's init_B scale: valence for uniform weights) and, if diag is not
// null, diag[g] for the preconditioner. MODE 1 / 2 (CG iteration): v = the
// new direction p = z (MODE 1) or z + beta p_old (MODE 2), z = r, or r / diag
// (Jacobi, the reference's precond_matvec, :219-299) when PRECOND; p is
// written for the owned rows and <y, p> summed per workgroup into partial[p].
template <uint32_t T, int MODE, bool UNIFORM, bool PRECOND>
__global__ __launch_bounds__(T) void k_mf_apply(DevMesh m, const uint32_t* __restrict__ vv_base,
                                                const uint32_t* __restrict__ rs, const uint16_t* __restrict__ ent,
                                                const float* __restrict__ coords, float dt,
                                                const float* __restrict__ x, const float* __restrict__ r,
                                                float* __restrict__ diag, float* __restrict__ mass,
                                                float* __restrict__ p_new, float* __restrict__ y,
                                                float* __restrict__ partial, const CgState* __restrict__ st)
{
    extern __shared__ uint4 smem[];
    const uint32_t p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    if (MODE != 0 && st->done) return;
    const PatchDesc d  = load_desc(m.patch, p);
    const uint32_t  vb = vv_base[p], nent = vv_base[p + 1] - vb;
    const uint32_t  N  = m.nV;
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanMF    L = plan_mf(b, d.nv, d.ne, d.nvo, nent, UNIFORM, T);
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    stage_halves<T, 8>(L.ent, ent + vb, nent);
    stage_words<T, 2>(L.rs, rs + d.vp, d.nvo + 1, vb);
    if constexpr (!UNIFORM) {
        stage_vertices<T>(m, d, coords, L.pos, L.lin);
        for (uint32_t i = threadIdx.x; i < d.ne; i += T) L.he[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    {
        const uint32_t* vl   = m.v_lin + d.v_off;
        const float     beta = MODE == 2 ? st->beta : 0.f;
        for (uint32_t l = threadIdx.x; l < d.nv; l += T) {
            const uint32_t g = lin_of(l, d.nvo, d.vp, vl);
            float          a[3];
#pragma unroll
            for (uint32_t c = 0; c < 3; ++c) {
                if constexpr (MODE == 0) {
                    a[c] = x[c * N + g];
                } else {
                    const float z = PRECOND ? r[c * N + g] / diag[g] : r[c * N + g];
                    a[c]          = (MODE == 2) ? z + beta * x[c * N + g] : z;  // x holds p_old
                }
            }
            L.vx[l] = a[0];
            L.vy[l] = a[1];
            L.vz[l] = a[2];
        }
    }
    __syncthreads();
    if constexpr (!UNIFORM) {
        const uint16_t* fe = m.fe + 3u * d.f_off;
        float2*         hh = reinterpret_cast<float2*>(L.he);  // hh[2e + dir]
        for (uint32_t f = threadIdx.x; f < d.nf; f += T) {
            uint32_t xs[3], a[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) xs[j] = fe[3 * f + j];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const uint32_t e = L.ev[xs[j] >> 1];
                a[j]             = (xs[j] & 1u) ? hi16(e) : lo16(e);
            }
            const float4  q0 = L.pos[a[0]], q1 = L.pos[a[1]], q2 = L.pos[a[2]];
            const FaceCot c  = face_cot(f3{q0.x, q0.y, q0.z}, f3{q1.x, q1.y, q1.z}, f3{q2.x, q2.y, q2.z});
#pragma unroll
            for (int j = 0; j < 3; ++j) hh[xs[j]] = make_float2(c.cot[(j + 2) % 3], c.vor[j]);
        }
        __syncthreads();
    }
    float dotp = 0.f;
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        const uint32_t k0 = L.rs[v], k1 = L.rs[v + 1];
        float          sum_w = 0.f, vw = 0.f, ax = 0.f, ay = 0.f, az = 0.f;
        for (uint32_t k = k0; k < k1; ++k) {
            const uint32_t e = L.ent[k] >> 1, xe = L.ev[e];
            const uint32_t q = (lo16(xe) == v) ? hi16(xe) : lo16(xe);
            float          w = 1.f;
            if constexpr (!UNIFORM) {
                const float4 h = L.he[e];
                w              = h.x + h.z;
                w              = (w >= 0.f) ? w : 0.f;
                vw += (lo16(xe) == v) ? h.y : h.w;
            }
            const float wd = w * dt;
            sum_w += wd;
            ax -= wd * L.vx[q];
            ay -= wd * L.vy[q];
            az -= wd * L.vz[q];
        }
        const float vwt = UNIFORM ? 1.0f / static_cast<float>(k1 - k0) : 0.5f / vw;
        const float dg  = 1.0f / vwt + sum_w;
        const float px = L.vx[v], py = L.vy[v], pz = L.vz[v];
        ax += dg * px;
        ay += dg * py;
        az += dg * pz;
        const uint32_t g = d.vp + v;
        y[g]             = ax;
        y[N + g]         = ay;
        y[2 * N + g]     = az;
        if constexpr (MODE == 0) {
            // the reference's init_B: x * valence (uniform) or x / v_weight
            mass[g] = UNIFORM ? static_cast<float>(k1 - k0) : 1.0f / vwt;
            if (diag) diag[g] = dg;
        } else {
            p_new[g]         = px;
            p_new[N + g]     = py;
            p_new[2 * N + g] = pz;
            dotp += ax * px + ay * py + az * pz;
        }
    }
    if constexpr (MODE != 0) {
        for (uint32_t o = 32; o > 0; o >>= 1) dotp += __shfl_xor(dotp, o, 64);
        if ((threadIdx.x & 63u) == 0) L.red[threadIdx.x >> 6] = dotp;
        __syncthreads();
        if (threadIdx.x == 0) {
            float t = 0.f;
            for (uint32_t i = 0; i < T / 64u; ++i) t += L.red[i];
            partial[p] = t;
        }
    }
}

// Per linear vertex (MODE 0, set-up): r = b - s with b = mass * x0, and the
// workgroup's partial <r, r> (<r, r / diag> with the Jacobi preconditioner,
// diag non-null). MODE 1 (iteration): x += alpha p, r -= alpha s and the
// partial <r, r> (<r, r / diag>).
template <int MODE>
__global__ __launch_bounds__(256) void k_cg_update(uint32_t n, const float* __restrict__ mass,
                                                   const float* __restrict__ x0, float* __restrict__ x,
                                                   float* __restrict__ r, const float* __restrict__ pvec,
                                                   const float* __restrict__ s, float* __restrict__ partial,
                                                   const CgState* __restrict__ st,
                                                   const float* __restrict__ diag = nullptr)
{
    __shared__ float red[4];
    if (MODE == 1 && st->done) return;
    const uint32_t i   = blockIdx.x * 256u + threadIdx.x;
    float          acc = 0.f;
    if (i < n) {
        if constexpr (MODE == 0) {
            // b = M x0 rounded on its own (the reference stores B, then forms
            // B - A x): no fused multiply-subtract here
#pragma clang fp contract(off)
            const float M = mass[i];
            const float D = diag ? diag[i] : 1.f;
#pragma unroll
            for (uint32_t c = 0; c < 3; ++c) {
                const float v = M * x0[c * n + i] - s[c * n + i];
                r[c * n + i]  = v;
                x[c * n + i]  = x0[c * n + i];
                acc += diag ? v * (v / D) : v * v;
            }
        } else {
            const float a = st->alpha;
            const float D = diag ? diag[i] : 1.f;
#pragma unroll
            for (uint32_t c = 0; c < 3; ++c) {
                x[c * n + i] += a * pvec[c * n + i];
                const float v = r[c * n + i] - a * s[c * n + i];
                r[c * n + i]  = v;
                acc += diag ? v * (v / D) : v * v;
            }
        }
    }
    for (uint32_t o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// One workgroup: fixed-order sum of the partials, then the scalar step.
//   STEP 0 (set-up): delta = start = sum; done if already converged.
//   STEP 1 (after the SpMV): alpha = delta / sum.
//   STEP 2 (after the update): delta_old = delta, delta = sum; done when
//          delta < abs_tol or delta / start < rel_tol, else beta =
//          delta / delta_old and one more iteration taken.
template <int STEP>
__global__ __launch_bounds__(1024) void k_cg_scalar(uint32_t n, const float* __restrict__ partial,
                                                    CgState* __restrict__ st, float abs_tol, float rel_tol)
{
    __shared__ double s[1024];
    if (STEP != 0 && st->done) return;
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) acc += static_cast<double>(partial[i]);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t k = 512; k > 0; k >>= 1) {
        if (threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const float sum = static_cast<float>(s[0]);
    if constexpr (STEP == 0) {
        st->delta = st->start = sum;
        st->delta_old = 0.f;
        st->alpha = st->beta = 0.f;
        st->iters = 0;
        st->done  = (sum < abs_tol || sum / sum < rel_tol) ? 1u : 0u;
    } else if constexpr (STEP == 1) {
        st->alpha = st->delta / sum;
    } else {
        st->delta_old = st->delta;
        st->delta     = sum;
        if (sum < abs_tol || sum / st->start < rel_tol) {
            st->done = 1u;
        } else {
            st->beta = sum / st->delta_old;
            st->iters += 1u;
        }
    }
}

// ---- uniform smoothing over a materialised VV CSR ---------------------------
// One Jacobi step x' = x + lambda (mean of row neighbours - x) over the CSR of
// a VV query (linear ids). LANES lanes share a row; fixed xor-tree reduction
// (LANES = 1: the row summed in order).
template <uint32_t LANES>
__global__ __launch_bounds__(256) void k_smooth_csr(uint32_t n, const int32_t* __restrict__ row_ptr,
                                                    const int32_t* __restrict__ col, float lambda,
                                                    const float* __restrict__ x, float* __restrict__ y)
{
    const uint32_t gid  = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t row  = gid / LANES;
    const uint32_t lane = gid % LANES;
    float          sx = 0.f, sy = 0.f, sz = 0.f;
    const bool     ok = row < n;
    int32_t        s = 0, e = 0;
    if (ok) {
        s = row_ptr[row];
        e = row_ptr[row + 1];
        for (int32_t k = s + static_cast<int32_t>(lane); k < e; k += LANES) {
            const uint32_t c = static_cast<uint32_t>(col[k]);
            sx += x[c];
            sy += x[n + c];
            sz += x[2 * n + c];
        }
    }
#pragma unroll
    for (uint32_t o = LANES / 2; o > 0; o >>= 1) {
        sx += __shfl_xor(sx, o, 64);
        sy += __shfl_xor(sy, o, 64);
        sz += __shfl_xor(sz, o, 64);
    }
    if (ok && lane == 0) {
        const float px = x[row], py = x[n + row], pz = x[2 * n + row];
        if (e > s) {
            const float inv = 1.0f / static_cast<float>(e - s);
            y[row]          = px + lambda * (sx * inv - px);
            y[n + row]      = py + lambda * (sy * inv - py);
            y[2 * n + row]  = pz + lambda * (sz * inv - pz);
        } else {
            y[row]         = px;
            y[n + row]     = py;
            y[2 * n + row] = pz;
        }
    }
}

// ---- symmetric Dirichlet ----------------------------------------------------
// Rest shape of a face: apps/Param/param.cu:113-141. Mr is stored SoA over
// linear face ids: Mr[c * F + f], c = (m00, m10, m01, m11), columns b-a, c-a
// in the face's local 2D frame.
struct PlanFV
{
    uint32_t* ev;
    uint16_t* fe;
};
__host__ __device__ inline PlanFV plan_fv(Bump& b, uint32_t ne, uint32_t nf)
{
    PlanFV p{};
    p.ev = b.alloc<uint32_t>(r4(ne));
    p.fe = b.alloc<uint16_t>(r4(3 * nf));
    return p;
}

__global__ __launch_bounds__(kGeomThreads) void k_symdir_rest(DevMesh m, const float* __restrict__ pos,
                                                              float* __restrict__ Mr)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = blockIdx.x;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanFV    L = plan_fv(b, d.ne, d.nf);
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
    __syncthreads();
    const uint32_t N = m.nV, F = m.nF;
    for (uint32_t f = threadIdx.x; f < d.nfo; f += blockDim.x) {
        f3 P[3];
        for (uint32_t i = 0; i < 3; ++i) {
            const uint32_t g = lin_of(face_vertex(L.fe, L.ev, f, i), d.nvo, d.vp, m.v_lin + d.v_off);
            P[i] = f3{pos[g], pos[N + g], pos[2 * N + g]};
        }
        const f3    e1 = sub(P[1], P[0]), e2 = sub(P[2], P[0]);
        const f3    n  = cross(e1, e2);
        const float l1 = sqrtf(dot(e1, e1));
        const f3    b1 = {e1.x / l1, e1.y / l1, e1.z / l1};
        const f3    t  = cross(n, b1);
        const float lt = sqrtf(dot(t, t));
        const f3    b2 = {t.x / lt, t.y / lt, t.z / lt};
        const uint32_t lf = d.fp + f;
        Mr[lf]         = dot(e1, b1);
        Mr[F + lf]     = 0.f;
        Mr[2 * F + lf] = dot(e2, b1);
        Mr[3 * F + lf] = dot(e2, b2);
    }
}

// Energy and (optionally) corner gradients of one face. uv corners a, b, c;
// R = Mr (m00, m10, m01, m11). Returns FLT_MAX (and zero gradient) for a
// face whose UV triangle is not positively oriented (param.cu:165-168).
// E = A (|J|^2 + |J^-1|^2), J = M R^-1, A = det(R) / 2; for 2x2 J,
// |J^-1|^2 = |J|^2 / det(J)^2. Gradient (closed form):
// dE/dJ = A (2 (1 + 1/d^2) J - 2 |J|^2 / d^3 cof(J)), dE/dM = dE/dJ R^-T.
// a*d - b*c with one rounding error at most a few ulps (Kahan's FMA
// algorithm), so near-degenerate faces keep their relative accuracy
__device__ __forceinline__ float det2(float a, float d, float b, float c)
{
    const float w = b * c;
    const float e = fmaf(-b, c, w);
    const float f = fmaf(a, d, -w);
    return f + e;
}

template <bool GRAD>
__device__ __forceinline__ float symdir_face(const float* a, const float* b, const float* c, const float* R,
                                             float* g /* 6: da, db, dc */)
{
    const float M00 = b[0] - a[0], M10 = b[1] - a[1], M01 = c[0] - a[0], M11 = c[1] - a[1];
    const float dm  = det2(M00, M11, M01, M10);
    if (dm <= 0.f) {
        if (GRAD)
            for (int i = 0; i < 6; ++i) g[i] = 0.f;
        return FLT_MAX;
    }
    const float dr = det2(R[0], R[3], R[2], R[1]);
    const float A  = 0.5f * dr;
    const float i00 = R[3] / dr, i01 = -R[2] / dr, i10 = -R[1] / dr, i11 = R[0] / dr;
    const float J00 = M00 * i00 + M01 * i10, J01 = M00 * i01 + M01 * i11;
    const float J10 = M10 * i00 + M11 * i10, J11 = M10 * i01 + M11 * i11;
    const float n2 = J00 * J00 + J01 * J01 + J10 * J10 + J11 * J11;
    const float dj = dm / dr;  // det J = det M / det R
    const float id2 = 1.0f / (dj * dj);
    if (GRAD) {
        const float s1 = 2.0f * (1.0f + id2), s2 = -2.0f * n2 * id2 / dj;
        const float G00 = A * (s1 * J00 + s2 * J11), G01 = A * (s1 * J01 - s2 * J10);
        const float G10 = A * (s1 * J10 - s2 * J01), G11 = A * (s1 * J11 + s2 * J00);
        const float D00 = G00 * i00 + G01 * i01, D01 = G00 * i10 + G01 * i11;
        const float D10 = G10 * i00 + G11 * i01, D11 = G10 * i10 + G11 * i11;
        g[2] = D00;
        g[3] = D10;
        g[4] = D01;
        g[5] = D11;
        g[0] = -(D00 + D01);
        g[1] = -(D10 + D11);
    }
    return A * (n2 + n2 * id2);
}

// Per owned face energy into energy[linear face]; each workgroup also writes
// its patch's sum (fixed-order tree) into partial[p].
__global__ __launch_bounds__(kGeomThreads) void k_symdir_energy(DevMesh m, const float* __restrict__ Mr,
                                                                const float* __restrict__ uv,
                                                                float* __restrict__ energy,
                                                                float* __restrict__ partial)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = blockIdx.x;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanFV    L  = plan_fv(b, d.ne, d.nf);
    float*          sr = b.alloc<float>(kGeomThreads);
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
    __syncthreads();
    const uint32_t N = m.nV, F = m.nF;
    float          acc = 0.f;
    for (uint32_t f = threadIdx.x; f < d.nfo; f += blockDim.x) {
        float U[3][2];
        for (uint32_t i = 0; i < 3; ++i) {
            const uint32_t g = lin_of(face_vertex(L.fe, L.ev, f, i), d.nvo, d.vp, m.v_lin + d.v_off);
            U[i][0] = uv[g];
            U[i][1] = uv[N + g];
        }
        const uint32_t lf   = d.fp + f;
        const float    R[4] = {Mr[lf], Mr[F + lf], Mr[2 * F + lf], Mr[3 * F + lf]};
        const float    e    = symdir_face<false>(U[0], U[1], U[2], R, nullptr);
        energy[lf]          = e;
        acc += e;
    }
    sr[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t s = kGeomThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sr[threadIdx.x] += sr[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[p] = sr[0];
}

// Fixed-order sum of n floats by one workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_reduce_sum(uint32_t n, const float* __restrict__ in, float* __restrict__ out)
{
    __shared__ float s[1024];
    float            acc = 0.f;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) acc += in[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t k = 512; k > 0; k >>= 1) {
        if (threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

struct PlanSG
{
    uint32_t *ev, *off, *cur, *wave;
    uint16_t *fe, *fvl, *val;
    float *   u, *v, *gc;
};
__host__ __device__ inline PlanSG plan_sg(Bump& b, uint32_t nv, uint32_t ne, uint32_t nf)
{
    PlanSG p{};
    p.ev   = b.alloc<uint32_t>(r4(ne));
    p.off  = b.alloc<uint32_t>(nv + 1);
    p.cur  = b.alloc<uint32_t>(nv);
    p.wave = b.alloc<uint32_t>(20);
    p.fe   = b.alloc<uint16_t>(r4(3 * nf));
    p.fvl  = b.alloc<uint16_t>(3 * nf);
    p.val  = b.alloc<uint16_t>(3 * nf);
    p.u    = b.alloc<float>(nv);
    p.v    = b.alloc<float>(nv);
    p.gc   = b.alloc<float>(6 * nf);
    return p;
}

// Adjoint of sum_f E_f w.r.t. UV: every local face (owned and ribbon) of the
// patch recomputes its Jacobian and corner gradients into LDS; each owned
// vertex then sums the corners of its VF row in ascending local-face order.
// No global atomics (the reference scatters with atomicAdd,
// include/rxmesh/diff/diff_query_kernel.cuh:275-282). grad = scale * dE/duv,
// scale read from device memory (1 if null).
__global__ __launch_bounds__(kGeomThreads) void k_symdir_grad(DevMesh m, const float* __restrict__ Mr,
                                                              const float* __restrict__ uv,
                                                              const float* __restrict__ scale,
                                                              float* __restrict__ grad)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = blockIdx.x;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanSG    L = plan_sg(b, d.nv, d.ne, d.nf);
    const uint32_t  T = blockDim.x, N = m.nV, F = m.nF;
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
    for (uint32_t l = threadIdx.x; l < d.nv; l += T) {
        const uint32_t g = lin_of(l, d.nvo, d.vp, m.v_lin + d.v_off);
        L.u[l]           = uv[g];
        L.v[l]           = uv[N + g];
    }
    __syncthreads();
    for (uint32_t f = threadIdx.x; f < d.nf; f += T) {
        float U[3][2];
        for (uint32_t i = 0; i < 3; ++i) {
            const uint32_t l = face_vertex(L.fe, L.ev, f, i);
            L.fvl[3 * f + i] = static_cast<uint16_t>(l);
            U[i][0]          = L.u[l];
            U[i][1]          = L.v[l];
        }
        const uint32_t lf   = lin_of(f, d.nfo, d.fp, m.f_lin + d.f_off);
        const float    R[4] = {Mr[lf], Mr[F + lf], Mr[2 * F + lf], Mr[3 * F + lf]};
        symdir_face<true>(U[0], U[1], U[2], R, L.gc + 6 * f);
    }
    __syncthreads();
    lds_transpose<3>(3 * d.nf, d.nv, [&](uint32_t i) { return static_cast<uint32_t>(L.fvl[i]); }, L.off, L.cur,
                     L.val, L.wave, d.nvo);
    const float sc = scale ? scale[0] : 1.0f;
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        float gu = 0.f, gv = 0.f;
        for (uint32_t k = L.off[v]; k < L.off[v + 1]; ++k) {
            const uint32_t f = L.val[k];
            const uint32_t c = (L.fvl[3 * f] == v) ? 0u : ((L.fvl[3 * f + 1] == v) ? 1u : 2u);
            gu += L.gc[6 * f + 2 * c];
            gv += L.gc[6 * f + 2 * c + 1];
        }
        grad[d.vp + v]     = sc * gu;
        grad[N + d.vp + v] = sc * gv;
    }
}

// ---- symmetric Dirichlet with 3-D positions as the variable ------------------
// For a closed surface (no UV domain) the current triangle is compared with
// its rest shape directly in 3-D: J = [b-a, c-a] Mr^{-1} is 3x2,
// |J^{-1}|^2 -> |J|^2 / G with G = det(J^T J) = |J0 x J1|^2, so
// E = A |J|^2 (1 + 1/G). Reduces to symdir_face for planar input.
// g receives the 9 corner gradients (da, db, dc); FLT_MAX when G <= 0.
__device__ __forceinline__ float symdir3_face(const f3 a, const f3 b, const f3 c, const float* R, float* g)
{
    const f3    e1 = sub(b, a), e2 = sub(c, a);
    const float dr = det2(R[0], R[3], R[2], R[1]);
    const float A  = 0.5f * dr;
    const float i00 = R[3] / dr, i01 = -R[2] / dr, i10 = -R[1] / dr, i11 = R[0] / dr;
    const f3    J0 = {e1.x * i00 + e2.x * i10, e1.y * i00 + e2.y * i10, e1.z * i00 + e2.z * i10};
    const f3    J1 = {e1.x * i01 + e2.x * i11, e1.y * i01 + e2.y * i11, e1.z * i01 + e2.z * i11};
    const float a00 = dot(J0, J0), a11 = dot(J1, J1), a01 = dot(J0, J1);
    const float n2 = a00 + a11, G = det2(a00, a11, a01, a01);
    if (G <= 0.f) {
        for (int i = 0; i < 9; ++i) g[i] = 0.f;
        return FLT_MAX;
    }
    const float iG = 1.0f / G;
    const float s1 = 2.0f * (1.0f + iG), s2 = -n2 * iG * iG;
    const float q0[3] = {J0.x, J0.y, J0.z}, q1[3] = {J1.x, J1.y, J1.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float dG0 = 2.0f * (a11 * q0[k] - a01 * q1[k]);
        const float dG1 = 2.0f * (a00 * q1[k] - a01 * q0[k]);
        const float E0 = A * (s1 * q0[k] + s2 * dG0), E1 = A * (s1 * q1[k] + s2 * dG1);
        const float D0 = E0 * i00 + E1 * i01, D1 = E0 * i10 + E1 * i11;
        g[3 + k] = D0;
        g[6 + k] = D1;
        g[k]     = -(D0 + D1);
    }
    return A * n2 * (1.0f + iG);
}

// Work-queue counter reset (a kernel, so graphs hold no memset nodes).
__global__ void k_reset_counter(uint32_t* c) { c[0] = 0u; }

// Gradient of corner c of one face only (3 values), recomputed per (vertex,
// face) pair from the face plan instead of stored per face; the corner is
// selected with conditional moves (no runtime-indexed local array).
__device__ __forceinline__ f3 symdir3_corner(const f3 a, const f3 b, const f3 c, const float* R, uint32_t corner)
{
    const f3    e1 = sub(b, a), e2 = sub(c, a);
    const float dr = det2(R[0], R[3], R[2], R[1]);
    const float A  = 0.5f * dr;
    const float i00 = R[3] / dr, i01 = -R[2] / dr, i10 = -R[1] / dr, i11 = R[0] / dr;
    const f3    J0 = {e1.x * i00 + e2.x * i10, e1.y * i00 + e2.y * i10, e1.z * i00 + e2.z * i10};
    const f3    J1 = {e1.x * i01 + e2.x * i11, e1.y * i01 + e2.y * i11, e1.z * i01 + e2.z * i11};
    const float a00 = dot(J0, J0), a11 = dot(J1, J1), a01 = dot(J0, J1);
    const float n2 = a00 + a11, G = det2(a00, a11, a01, a01);
    if (G <= 0.f) return f3{0.f, 0.f, 0.f};
    const float iG = 1.0f / G;
    const float s1 = 2.0f * (1.0f + iG), s2 = -n2 * iG * iG;
    auto d = [&](float j0, float j1, float& D0, float& D1) {
        const float E0 = A * (s1 * j0 + s2 * 2.0f * (a11 * j0 - a01 * j1));
        const float E1 = A * (s1 * j1 + s2 * 2.0f * (a00 * j1 - a01 * j0));
        D0 = E0 * i00 + E1 * i01;
        D1 = E0 * i10 + E1 * i11;
    };
    float bx, by, bz, cx, cy, cz;
    d(J0.x, J1.x, bx, cx);
    d(J0.y, J1.y, by, cy);
    d(J0.z, J1.z, bz, cz);
    if (corner == 1u) return f3{bx, by, bz};
    if (corner == 2u) return f3{cx, cy, cz};
    return f3{-(bx + cx), -(by + cy), -(bz + cz)};
}

struct PlanS3P
{
    uint16_t *fvl, *ro, *vf;
    float *   x, *y, *z, *mr;
};
__host__ __device__ inline PlanS3P plan_s3p(Bump& b, uint32_t nv, uint32_t nvo, uint32_t nf, uint32_t nent)
{
    PlanS3P p{};
    p.fvl = b.alloc<uint16_t>((3 * nf + 7u) & ~7u);
    p.ro  = b.alloc<uint16_t>((nvo + 1u + 7u) & ~7u);
    p.vf  = b.alloc<uint16_t>((nent + 7u) & ~7u);
    p.x   = b.alloc<float>(nv);
    p.y   = b.alloc<float>(nv);
    p.z   = b.alloc<float>(nv);
    p.mr  = b.alloc<float>(4 * nf);
    return p;
}

// x_out = x - eta * dE/dx over the patch-local face plan (GRAD: x_out =
// dE/dx itself, eta unused): positions, rest shapes and plan rows are staged
// into LDS; each owned vertex recomputes the corner gradient of every
// incident face in the plan's ascending order.
template <uint32_t T, bool GRAD = false>
__device__ __forceinline__ void symdir3_plan_patch(const DevMesh& m, uint32_t p, const uint32_t* __restrict__ fbase,
                                                   const uint32_t* __restrict__ rbase,
                                                   const uint32_t* __restrict__ nbase,
                                                   const uint16_t* __restrict__ fvl,
                                                   const uint16_t* __restrict__ roff,
                                                   const uint16_t* __restrict__ vf, const float* __restrict__ Mr,
                                                   float eta, const float* __restrict__ x, float* __restrict__ xo,
                                                   char* smem)
{
    const PatchDesc d = load_desc(m.patch, p);
    const uint32_t  nb0 = nbase[p], nent = nbase[p + 1] - nb0;
    Bump            b(smem);
    const PlanS3P   L = plan_s3p(b, d.nv, d.nvo, d.nf, nent);
    const uint32_t  N = m.nV, F = m.nF;
    stage_u16x8<T>(L.fvl, fvl + fbase[p], 3 * d.nf);
    stage_u16x8<T>(L.ro, roff + rbase[p], d.nvo + 1);
    stage_u16x8<T>(L.vf, vf + nb0, nent);
    stage_positions(L.x, L.y, L.z, x, N, d, m.v_lin);
    for (uint32_t f = threadIdx.x; f < d.nf; f += T) {
        const uint32_t lf = lin_of(f, d.nfo, d.fp, m.f_lin + d.f_off);
        L.mr[4 * f]       = Mr[lf];
        L.mr[4 * f + 1]   = Mr[F + lf];
        L.mr[4 * f + 2]   = Mr[2 * F + lf];
        L.mr[4 * f + 3]   = Mr[3 * F + lf];
    }
    __syncthreads();
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        float gx = 0.f, gy = 0.f, gz = 0.f;
        for (uint32_t k = L.ro[v]; k < L.ro[v + 1]; ++k) {
            const uint32_t s = L.vf[k], f = s >> 2, c = s & 3u;
            const uint32_t a = L.fvl[3 * f], bb = L.fvl[3 * f + 1], cc = L.fvl[3 * f + 2];
            const f3 g = symdir3_corner(f3{L.x[a], L.y[a], L.z[a]}, f3{L.x[bb], L.y[bb], L.z[bb]},
                                        f3{L.x[cc], L.y[cc], L.z[cc]}, L.mr + 4 * f, c);
            gx += g.x;
            gy += g.y;
            gz += g.z;
        }
        const uint32_t g = d.vp + v;
        if constexpr (GRAD) {
            xo[g]         = gx;
            xo[N + g]     = gy;
            xo[2 * N + g] = gz;
        } else {
            xo[g]         = L.x[v] - eta * gx;
            xo[N + g]     = L.y[v] - eta * gy;
            xo[2 * N + g] = L.z[v] - eta * gz;
        }
    }
}

template <uint32_t T>
__global__ __launch_bounds__(T) void k_symdir3_step_plan(DevMesh m, const uint32_t* __restrict__ fbase,
                                                         const uint32_t* __restrict__ rbase,
                                                         const uint32_t* __restrict__ nbase,
                                                         const uint16_t* __restrict__ fvl,
                                                         const uint16_t* __restrict__ roff,
                                                         const uint16_t* __restrict__ vf, const float* __restrict__ Mr,
                                                         float eta, const float* __restrict__ x, float* __restrict__ xo)
{
    extern __shared__ uint4 smem[];
    symdir3_plan_patch<T>(m, blockIdx.x, fbase, rbase, nbase, fvl, roff, vf, Mr, eta, x, xo,
                          reinterpret_cast<char*>(smem));
}

// grad = dE/dx (3 x V SoA) over the face plan, one workgroup per patch.
template <uint32_t T>
__global__ __launch_bounds__(T) void k_symdir3_grad_plan(DevMesh m, const uint32_t* __restrict__ fbase,
                                                         const uint32_t* __restrict__ rbase,
                                                         const uint32_t* __restrict__ nbase,
                                                         const uint16_t* __restrict__ fvl,
                                                         const uint16_t* __restrict__ roff,
                                                         const uint16_t* __restrict__ vf, const float* __restrict__ Mr,
                                                         const float* __restrict__ x, float* __restrict__ grad)
{
    extern __shared__ uint4 smem[];
    symdir3_plan_patch<T, true>(m, blockIdx.x, fbase, rbase, nbase, fvl, roff, vf, Mr, 0.f, x, grad,
                                reinterpret_cast<char*>(smem));
}

// Persistent form: a grid sized to the machine pulls patch ids from a global
// work queue (one atomic ticket per patch, counter cleared by k_reset_counter
// before the launch); every workgroup exits once the queue is drained.
template <uint32_t T>
__global__ __launch_bounds__(T) void k_symdir3_step_plan_queue(
    DevMesh m, uint32_t* __restrict__ queue, const uint32_t* __restrict__ fbase, const uint32_t* __restrict__ rbase,
    const uint32_t* __restrict__ nbase, const uint16_t* __restrict__ fvl, const uint16_t* __restrict__ roff,
    const uint16_t* __restrict__ vf, const float* __restrict__ Mr, float eta, const float* __restrict__ x,
    float* __restrict__ xo)
{
    extern __shared__ uint4 smem[];
    __shared__ uint32_t s_p;
    while (true) {
        if (threadIdx.x == 0) s_p = atomicAdd(queue, 1u);
        __syncthreads();
        const uint32_t p = s_p;
        __syncthreads();
        if (p >= m.P) break;
        symdir3_plan_patch<T>(m, p, fbase, rbase, nbase, fvl, roff, vf, Mr, eta, x, xo, reinterpret_cast<char*>(smem));
        __syncthreads();
    }
}

// Per-face 3-D symmetric Dirichlet energy of owned faces (FLT_MAX if degenerate).
__global__ __launch_bounds__(kGeomThreads) void k_symdir3_energy(DevMesh m, const float* __restrict__ Mr,
                                                                 const float* __restrict__ x,
                                                                 float* __restrict__ energy)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = blockIdx.x;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanFV    L = plan_fv(b, d.ne, d.nf);
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
    __syncthreads();
    const uint32_t N = m.nV, F = m.nF;
    for (uint32_t f = threadIdx.x; f < d.nfo; f += blockDim.x) {
        f3 P[3];
        for (uint32_t i = 0; i < 3; ++i) {
            const uint32_t g = lin_of(face_vertex(L.fe, L.ev, f, i), d.nvo, d.vp, m.v_lin + d.v_off);
            P[i] = f3{x[g], x[N + g], x[2 * N + g]};
        }
        const uint32_t lf   = d.fp + f;
        const float    R[4] = {Mr[lf], Mr[F + lf], Mr[2 * F + lf], Mr[3 * F + lf]};
        float          gdummy[9];
        energy[lf] = symdir3_face(P[0], P[1], P[2], R, gdummy);
    }
}

}  // namespace rxm
