// ARAP-style deformation terms (reference apps/DiffARAP): cotangent edge
// weights over the EVDiamond relation, the best rotation per vertex, and the
// EV energy with its hand-written gradient. One workgroup per patch; every
// per-vertex result is gathered over the owner's rows, without atomics.
#pragma once
#include "k_numerics.hpp"
#include "k_query.hpp"

namespace rxm {

// cotangent of the angle at c of triangle (a, b, c) as arap_kernels.h:41-45
// writes it: dot(a - c, b - c) / length(cross(a - c, b - c))
__device__ __forceinline__ float arap_cot(f3 a, f3 b, f3 c)
{
    const f3 u = sub(a, c), v = sub(b, c), x = cross(u, v);
    return dot(u, v) / sqrtf(dot(x, x));
}

// ---- edge weights ------------------------------------------------------------
struct PlanAW
{
    uint32_t* ev;
    uint16_t* fe;
    uint32_t* opp;
    float *   x, *y, *z;
};
__host__ __device__ inline PlanAW plan_aw(Bump& b, uint32_t nv, uint32_t ne, uint32_t nf, uint32_t neo)
{
    PlanAW p{};
    p.ev  = b.alloc<uint32_t>(r4(ne));
    p.fe  = b.alloc<uint16_t>(r4(3 * nf));
    p.opp = b.alloc<uint32_t>(2 * neo);
    p.x   = b.alloc<float>(nv);
    p.y   = b.alloc<float>(nv);
    p.z   = b.alloc<float>(nv);
    return p;
}

// Per owned edge p-r with EVDiamond opposite vertices q and s (the slots of
// k_query_evd): w = max(0, (cot_q + cot_s) / 2) (arap_kernels.h:8-54); 0 for
// an edge that lacks a side, which the reference skips, leaving the weight at
// its reset value 0. w is indexed by linear edge id.
template <uint32_t T>
__global__ __launch_bounds__(T) void k_arap_weights(DevMesh m, const float* __restrict__ pos, float* __restrict__ w)
{
    extern __shared__ uint4 smem[];
    const uint32_t p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanAW    L = plan_aw(b, d.nv, d.ne, d.nf, d.neo);
    for (uint32_t i = threadIdx.x; i < 2 * d.neo; i += T) L.opp[i] = 0u;
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    stage_fe(L.fe, m.fe + 3u * d.f_off, d.nf);
    stage_positions_batched<T, (512u + T - 1u) / T, false>(L.x, L.y, L.z, nullptr, pos, m.nV, d, m.v_lin);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 3 * d.nf; i += T) {
        const uint32_t f = i / 3u, k = i - 3u * f;
        const uint32_t y = L.fe[3u * f + (k == 2u ? 0u : k + 1u)];
        if ((y >> 1) < d.neo) atomicMax(&L.opp[y], ((f + 1u) << 16) | face_vertex(L.fe, L.ev, f, k));
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < d.neo; e += T) {
        const uint32_t qa = L.opp[2 * e], sa = L.opp[2 * e + 1];
        float          wt = 0.f;
        if (qa && sa) {
            const uint32_t x = L.ev[e], P = lo16(x), R = hi16(x), Q = qa & 0xFFFFu, S = sa & 0xFFFFu;
            const f3       pp{L.x[P], L.y[P], L.z[P]}, rr{L.x[R], L.y[R], L.z[R]};
            const f3       qq{L.x[Q], L.y[Q], L.z[Q]}, ss{L.x[S], L.y[S], L.z[S]};
            wt = (arap_cot(pp, rr, qq) + arap_cot(pp, rr, ss)) * 0.5f;
            wt = fmaxf(0.f, wt);  // also 0 for a NaN from a degenerate face, as std::max(0, NaN)
        }
        w[d.ep + e] = wt;
    }
}

// ---- best rotation -------------------------------------------------------------
// One Jacobi rotation of the symmetric 4x4 a in the (P, Q) plane (P < Q),
// zeroing a[P][Q]; v accumulates the eigenvectors as columns. Only the upper
// triangle of a is kept: a_PP -= t a_PQ, a_QQ += t a_PQ, a_PQ = 0, and the
// two other rows' (P, Q) pairs rotate. The rotation only has to be
// orthogonal, not exact (an angle off by an ulp leaves a tiny a[P][Q] that
// later sweeps remove): c = rsqrt(1 + t^2) and s = t c keep c^2 + s^2 = 1 to
// an ulp, so the hardware reciprocal and reciprocal square root serve.
template <int P, int Q>
__device__ __forceinline__ void jacobi4(float (&a)[4][4], float (&v)[4][4])
{
    static_assert(P < Q, "upper triangle");
    const float apq = a[P][Q];
    if (fabsf(apq) < 1e-30f) return;  // negligible in the unit-scaled matrix; keeps rcp away from denormals
    const float theta = 0.5f * (a[Q][Q] - a[P][P]) * __builtin_amdgcn_rcpf(apq);
    const float at    = fabsf(theta);
    // t = sign(theta) / (|theta| + sqrt(theta^2 + 1)); ~1 / (2 theta) once theta^2 would overflow
    const float den = (at < 1e18f) ? at + (at * at + 1.f) * __builtin_amdgcn_rsqf(at * at + 1.f) : 2.f * at;
    const float t   = copysignf(__builtin_amdgcn_rcpf(den), theta);
    const float c = __builtin_amdgcn_rsqf(t * t + 1.f), s = t * c;
    a[P][P] -= t * apq;
    a[Q][Q] += t * apq;
    a[P][Q] = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k == P || k == Q) continue;
        // the upper-triangle entries (k, P) and (k, Q)
        float& akp = (k < P) ? a[k][P] : a[P][k];
        float& akq = (k < Q) ? a[k][Q] : a[Q][k];
        const float xp = akp, xq = akq;
        akp              = c * xp - s * xq;
        akq              = s * xp + c * xq;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float vkp = v[k][P], vkq = v[k][Q];
        v[k][P]         = c * vkp - s * vkq;
        v[k][Q]         = s * vkp + c * vkq;
    }
}

// The rotation R maximising tr(R S) (row-major R, S[a][b] = sum w e_a e'_b
// of rest e and deformed e' edge vectors): the reference takes R = V U^T
// from S = U Sigma V^T with a sign fix for reflections
// (arap_kernels.h:89-110). Computed here by Horn's closed form (J. Opt. Soc.
// Am. A 4(4):629, 1987): R's unit quaternion is the eigenvector of the
// largest eigenvalue of the symmetric 4x4 N(S), found by cyclic Jacobi
// sweeps; the result is a proper rotation by construction. S = 0 gives I.
__device__ inline void best_rotation(const float (&S)[3][3], float (&R)[3][3])
{
    float a[4][4], v[4][4];
    a[0][0] = S[0][0] + S[1][1] + S[2][2];
    a[1][1] = S[0][0] - S[1][1] - S[2][2];
    a[2][2] = -S[0][0] + S[1][1] - S[2][2];
    a[3][3] = -S[0][0] - S[1][1] + S[2][2];
    a[0][1] = a[1][0] = S[1][2] - S[2][1];
    a[0][2] = a[2][0] = S[2][0] - S[0][2];
    a[0][3] = a[3][0] = S[0][1] - S[1][0];
    a[1][2] = a[2][1] = S[0][1] + S[1][0];
    a[1][3] = a[3][1] = S[2][0] + S[0][2];
    a[2][3] = a[3][2] = S[1][2] + S[2][1];
    float mx = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            mx      = fmaxf(mx, fabsf(a[i][j]));
            v[i][j] = (i == j) ? 1.f : 0.f;
        }
    if (mx > 0.f) {
        const float inv = 1.f / mx;  // scale: no under- or overflow in the sweeps
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) a[i][j] *= inv;
        for (int sweep = 0; sweep < 8; ++sweep) {
            // quadratic convergence: stop once the off-diagonal mass is below
            // fp32 resolution of the unit-scaled matrix (3-4 sweeps typically)
            const float off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[0][3] * a[0][3] + a[1][2] * a[1][2] +
                              a[1][3] * a[1][3] + a[2][3] * a[2][3];
            if (off < 1e-15f) break;
            jacobi4<0, 1>(a, v);
            jacobi4<0, 2>(a, v);
            jacobi4<0, 3>(a, v);
            jacobi4<1, 2>(a, v);
            jacobi4<1, 3>(a, v);
            jacobi4<2, 3>(a, v);
        }
    }
    // eigenvector of the largest eigenvalue (static indices keep a, v in registers)
    float lk = a[0][0], q0 = v[0][0], qx = v[1][0], qy = v[2][0], qz = v[3][0];
#pragma unroll
    for (int i = 1; i < 4; ++i)
        if (a[i][i] > lk) {
            lk = a[i][i];
            q0 = v[0][i];
            qx = v[1][i];
            qy = v[2][i];
            qz = v[3][i];
        }
    const float n = 1.f / sqrtf(q0 * q0 + qx * qx + qy * qy + qz * qz);
    q0 *= n;
    qx *= n;
    qy *= n;
    qz *= n;
    R[0][0] = q0 * q0 + qx * qx - qy * qy - qz * qz;
    R[0][1] = 2.f * (qx * qy - q0 * qz);
    R[0][2] = 2.f * (qx * qz + q0 * qy);
    R[1][0] = 2.f * (qy * qx + q0 * qz);
    R[1][1] = q0 * q0 - qx * qx + qy * qy - qz * qz;
    R[1][2] = 2.f * (qy * qz - q0 * qx);
    R[2][0] = 2.f * (qz * qx - q0 * qy);
    R[2][1] = 2.f * (qz * qy + q0 * qx);
    R[2][2] = q0 * q0 - qx * qx - qy * qy + qz * qz;
}

// Patch-local data of the rotation and energy kernels: EV, the owned rows'
// edges (the cached VV edge pattern, VV query order), rest and deformed
// positions and (energy kernel) rotations of every local vertex, weights of
// every local edge.
struct PlanAR
{
    uint32_t *ev, *rs;
    uint16_t* ent;
    float *   px, *py, *pz, *xx, *xy, *xz, *w, *rot;
};
__host__ __device__ inline PlanAR plan_ar(Bump& b, uint32_t nv, uint32_t ne, uint32_t nvo, uint32_t nent, bool rot)
{
    PlanAR p{};
    p.ev  = b.alloc<uint32_t>(r4(ne));
    p.rs  = b.alloc<uint32_t>(nvo + 1);
    p.ent = b.alloc<uint16_t>(nent);
    p.px  = b.alloc<float>(nv);
    p.py  = b.alloc<float>(nv);
    p.pz  = b.alloc<float>(nv);
    p.xx  = b.alloc<float>(nv);
    p.xy  = b.alloc<float>(nv);
    p.xz  = b.alloc<float>(nv);
    p.w   = b.alloc<float>(ne);
    p.rot = rot ? b.alloc<float>(9 * nv) : nullptr;
    return p;
}

template <uint32_t T>
__device__ __forceinline__ void stage_arap(const DevMesh& m, const PatchDesc& d, const PlanAR& L,
                                           const uint32_t* __restrict__ rs, const uint16_t* __restrict__ ent,
                                           uint32_t vb, uint32_t nent, const float* __restrict__ rest,
                                           const float* __restrict__ x, const float* __restrict__ w,
                                           const float* __restrict__ R)
{
    constexpr uint32_t RB = (512u + T - 1u) / T;
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    stage_halves<T, 8>(L.ent, ent + vb, nent);
    stage_words<T, 2>(L.rs, rs + d.vp, d.nvo + 1, vb);
    stage_positions_batched<T, RB, false>(L.px, L.py, L.pz, nullptr, rest, m.nV, d, m.v_lin);
    stage_positions_batched<T, RB, false>(L.xx, L.xy, L.xz, nullptr, x, m.nV, d, m.v_lin);
    const uint32_t* el = m.e_lin + d.e_off;
    for (uint32_t e = threadIdx.x; e < d.ne; e += T) L.w[e] = w[lin_of(e, d.neo, d.ep, el)];
    if (R) {
        const uint32_t* vl = m.v_lin + d.v_off;
        for (uint32_t l = threadIdx.x; l < d.nv; l += T) {
            const float* r = R + 9ull * lin_of(l, d.nvo, d.vp, vl);  // one 36-byte record per vertex
            float        t[9];
#pragma unroll
            for (uint32_t c = 0; c < 9; ++c) t[c] = r[c];
#pragma unroll
            for (uint32_t c = 0; c < 9; ++c) L.rot[9 * l + c] = t[c];
        }
    }
}

// Best rotation per owned vertex (arap_kernels.h:60-117): S = sum over the
// vertex's VV row of w (P_v - P_j)(X_v - X_j)^T, R from best_rotation. R
// holds one row-major 3 x 3 record per vertex, R[9 v + 3 i + j] = R_v(i, j)
// (the reference's rotations(v, i * 3 + j)): a ribbon vertex's rotation is
// one or two cache lines for the energy kernel's gather, not nine.
template <uint32_t T>
__global__ __launch_bounds__(T) void k_arap_rotations(DevMesh m, const uint32_t* __restrict__ vv_base,
                                                      const uint32_t* __restrict__ rs,
                                                      const uint16_t* __restrict__ ent,
                                                      const float* __restrict__ rest, const float* __restrict__ x,
                                                      const float* __restrict__ w, float* __restrict__ R)
{
    extern __shared__ uint4 smem[];
    const uint32_t p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d  = load_desc(m.patch, p);
    const uint32_t  vb = vv_base[p], nent = vv_base[p + 1] - vb;
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanAR    L = plan_ar(b, d.nv, d.ne, d.nvo, nent, false);
    stage_arap<T>(m, d, L, rs, ent, vb, nent, rest, x, w, nullptr);
    __syncthreads();
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        float S[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
        for (uint32_t k = L.rs[v]; k < L.rs[v + 1]; ++k) {
            const uint32_t e = L.ent[k] >> 1, y = L.ev[e];
            const uint32_t j = (lo16(y) == v) ? hi16(y) : lo16(y);
            const float    wt = L.w[e];
            const float    a[3] = {L.px[v] - L.px[j], L.py[v] - L.py[j], L.pz[v] - L.pz[j]};
            const float    c[3] = {L.xx[v] - L.xx[j], L.xy[v] - L.xy[j], L.xz[v] - L.xz[j]};
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int s = 0; s < 3; ++s) S[r][s] += wt * a[r] * c[s];
        }
        float Rv[3][3];
        best_rotation(S, Rv);
        float* out = R + 9ull * (d.vp + v);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int s = 0; s < 3; ++s) out[3 * r + s] = Rv[r][s];
    }
}

// edge residual d = (X_i - X_j) - R_i (P_i - P_j) for local endpoints i, j
__device__ __forceinline__ f3 arap_residual(const PlanAR& L, uint32_t i, uint32_t j)
{
    const float  e0 = L.px[i] - L.px[j], e1 = L.py[i] - L.py[j], e2 = L.pz[i] - L.pz[j];
    const float* r  = L.rot + 9 * i;
    return {(L.xx[i] - L.xx[j]) - (r[0] * e0 + r[1] * e1 + r[2] * e2),
            (L.xy[i] - L.xy[j]) - (r[3] * e0 + r[4] * e1 + r[5] * e2),
            (L.xz[i] - L.xz[j]) - (r[6] * e0 + r[7] * e1 + r[8] * e2)};
}

// Energy E = sum over owned edges, i = EV's first vertex (the larger global
// id, the reference's iter[0] of Op::EV), of w |(X_i - X_j) - R_i (P_i -
// P_j)|^2 (diff_arap.cu:87-118), one partial per patch; gradient with the
// rotations held fixed (the reference's term reads them as constants):
// dE/dX_v = sum over v's edges of +-2 w d, + where v is the edge's i.
template <uint32_t T>
__global__ __launch_bounds__(T) void k_arap_energy_grad(DevMesh m, const uint32_t* __restrict__ vv_base,
                                                        const uint32_t* __restrict__ rs,
                                                        const uint16_t* __restrict__ ent,
                                                        const float* __restrict__ rest, const float* __restrict__ x,
                                                        const float* __restrict__ w, const float* __restrict__ R,
                                                        float* __restrict__ partial, float* __restrict__ grad)
{
    extern __shared__ uint4 smem[];
    __shared__ float red[T / 64u];
    const uint32_t   p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d  = load_desc(m.patch, p);
    const uint32_t  vb = vv_base[p], nent = vv_base[p + 1] - vb;
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanAR    L = plan_ar(b, d.nv, d.ne, d.nvo, nent, true);
    stage_arap<T>(m, d, L, rs, ent, vb, nent, rest, x, w, R);
    __syncthreads();
    float acc = 0.f;
    for (uint32_t e = threadIdx.x; e < d.neo; e += T) {
        const uint32_t y = L.ev[e];
        const f3       r = arap_residual(L, lo16(y), hi16(y));
        acc += L.w[e] * dot(r, r);
    }
    const uint32_t N = m.nV;
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        float gx = 0.f, gy = 0.f, gz = 0.f;
        for (uint32_t k = L.rs[v]; k < L.rs[v + 1]; ++k) {
            const uint32_t e = L.ent[k] >> 1, y = L.ev[e];
            const f3       r = arap_residual(L, lo16(y), hi16(y));
            const float    s = (lo16(y) == v ? 2.f : -2.f) * L.w[e];
            gx += s * r.x;
            gy += s * r.y;
            gz += s * r.z;
        }
        const uint32_t g = d.vp + v;
        grad[g]          = gx;
        grad[N + g]      = gy;
        grad[2 * N + g]  = gz;
    }
    // fixed-order workgroup sum of the energy
    for (uint32_t o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (uint32_t i = 0; i < T / 64u; ++i) t += red[i];
        partial[p] = t;
    }
}

}  // namespace rxm
