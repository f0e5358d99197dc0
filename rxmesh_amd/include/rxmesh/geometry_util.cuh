// Triangle geometry that mesh app kernels call on vec3 positions: area,
// normal, the clamped cotangent, the partial Voronoi area of a corner and the
// cotangent weight of an edge. Interface of the reference's
// include/rxmesh/geometry_util.cuh:39-206 (same names, argument order and
// results, so app kernels such as apps/MCF/mcf_kernels.cuh compile against
// it); written over this library's glm-compatible vec3.
#pragma once

#include <cassert>
#include <cmath>
#include <limits>

#include "rxmesh/types.h"

namespace rxmesh {

// unit normal of the triangle (p0, p1, p2), counter-clockwise
template <typename T>
__host__ __device__ inline vec3<T> tri_normal(const vec3<T>& p0, const vec3<T>& p1, const vec3<T>& p2)
{
    return glm::normalize(glm::cross(p1 - p0, p2 - p0));
}

// area of the triangle (p0, p1, p2)
template <typename T>
__host__ __device__ inline T tri_area(const vec3<T>& p0, const vec3<T>& p1, const vec3<T>& p2)
{
    return T(0.5) * glm::length(glm::cross(p1 - p0, p2 - p0));
}

// cotangents clamped to those of angles in [1, 179] degrees (|cot| <= 19.1)
template <typename T>
__host__ __device__ __forceinline__ void clamp_cot(T& v)
{
    const T lim = T(19.1);
    if (v < -lim) v = -lim;
    if (v > lim) v = lim;
}

// Share of the triangle p -> q -> r (counter-clockwise, p the centre) in the
// mixed Voronoi area of p: a quarter of the area when the angle at p is
// obtuse, an eighth when another angle is, else the Voronoi region
// (|p r|^2 cot q + |p q|^2 cot r) / 8, with "cot" the corner's dot product
// over the triangle's area. -1 for a degenerate triangle.
template <typename T>
__host__ __device__ __forceinline__ T partial_voronoi_area(const vec3<T>& p, const vec3<T>& q, const vec3<T>& r)
{
    const T area = tri_area(p, q, r);
    if (area <= std::numeric_limits<T>::min()) return T(-1);
    const vec3<T> pq = q - p, qr = r - q, pr = r - p;
    const T       at_p = glm::dot(pq, pr);
    const T       at_q = -glm::dot(qr, pq);
    const T       at_r = glm::dot(qr, pr);
    if (at_p < T(0)) return T(0.25) * area;
    if (at_q < T(0) || at_r < T(0)) return T(0.125) * area;
    T cq = at_q / area, cr = at_r / area;
    clamp_cot(cq);
    clamp_cot(cr);
    return T(0.125) * (glm::length2(pr) * cq + glm::length2(pq) * cr);
}

// Weight of the edge p - r whose two opposite vertices are q and s: the sum
// of the clamped "cot" (dot product over the triangle's area) of the angles
// at q and at s; a degenerate triangle adds nothing.
template <typename T>
__host__ __device__ __forceinline__ T edge_cotan_weight(const vec3<T>& p, const vec3<T>& r, const vec3<T>& q,
                                                        const vec3<T>& s)
{
    auto opposite = [&](const vec3<T>& v) -> T {
        const T area = tri_area(p, r, v);
        if (!(area > std::numeric_limits<T>::min())) return T(0);
        T c = glm::dot(p - v, r - v) / area;
        clamp_cot(c);
        return c;
    };
    T w = T(0);
    w += opposite(q);
    w += opposite(s);
    assert(!isnan(w) && !isinf(w));
    return w;
}

}  // namespace rxmesh
