// Minimal vector library for per-element lambdas: the subset of glm's vec2 /
// vec3 / vec4 interface (component access, arithmetic, dot, cross, length,
// length2, distance, distance2, normalize) that mesh kernels use, in
// namespace glm so lambdas written against glm compile unchanged.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace glm {

template <int N, typename T>
struct vec;

template <typename T>
struct vec<2, T>
{
    T x, y;
    __host__ __device__ constexpr vec() : x(0), y(0) {}
    __host__ __device__ constexpr explicit vec(T s) : x(s), y(s) {}
    __host__ __device__ constexpr vec(T a, T b) : x(a), y(b) {}
    __host__ __device__ T&       operator[](int i) { return i == 0 ? x : y; }
    __host__ __device__ const T& operator[](int i) const { return i == 0 ? x : y; }
    static constexpr int         length() { return 2; }
};

template <typename T>
struct vec<3, T>
{
    T x, y, z;
    __host__ __device__ constexpr vec() : x(0), y(0), z(0) {}
    __host__ __device__ constexpr explicit vec(T s) : x(s), y(s), z(s) {}
    __host__ __device__ constexpr vec(T a, T b, T c) : x(a), y(b), z(c) {}
    __host__ __device__ T&       operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    __host__ __device__ const T& operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    static constexpr int         length() { return 3; }
};

template <typename T>
struct vec<4, T>
{
    T x, y, z, w;
    __host__ __device__ constexpr vec() : x(0), y(0), z(0), w(0) {}
    __host__ __device__ constexpr explicit vec(T s) : x(s), y(s), z(s), w(s) {}
    __host__ __device__ constexpr vec(T a, T b, T c, T d) : x(a), y(b), z(c), w(d) {}
    __host__ __device__ T&       operator[](int i) { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
    __host__ __device__ const T& operator[](int i) const { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
    static constexpr int         length() { return 4; }
};

template <typename T> using vec2 = vec<2, T>;
template <typename T> using vec3_t = vec<3, T>;
using vec3 = vec<3, float>;
using dvec3 = vec<3, double>;
using vec2f = vec<2, float>;

#define RXM_GLM_BINOP(OP)                                                                          \
    template <int N, typename T>                                                                   \
    __host__ __device__ inline vec<N, T> operator OP(const vec<N, T>& a, const vec<N, T>& b)       \
    {                                                                                              \
        vec<N, T> r;                                                                               \
        for (int i = 0; i < N; ++i) r[i] = a[i] OP b[i];                                           \
        return r;                                                                                  \
    }                                                                                              \
    template <int N, typename T>                                                                   \
    __host__ __device__ inline vec<N, T> operator OP(const vec<N, T>& a, T s)                      \
    {                                                                                              \
        vec<N, T> r;                                                                               \
        for (int i = 0; i < N; ++i) r[i] = a[i] OP s;                                              \
        return r;                                                                                  \
    }                                                                                              \
    template <int N, typename T>                                                                   \
    __host__ __device__ inline vec<N, T> operator OP(T s, const vec<N, T>& a)                      \
    {                                                                                              \
        vec<N, T> r;                                                                               \
        for (int i = 0; i < N; ++i) r[i] = s OP a[i];                                              \
        return r;                                                                                  \
    }                                                                                              \
    template <int N, typename T>                                                                   \
    __host__ __device__ inline vec<N, T>& operator OP##=(vec<N, T>& a, const vec<N, T>& b)         \
    {                                                                                              \
        for (int i = 0; i < N; ++i) a[i] = a[i] OP b[i];                                           \
        return a;                                                                                  \
    }                                                                                              \
    template <int N, typename T>                                                                   \
    __host__ __device__ inline vec<N, T>& operator OP##=(vec<N, T>& a, T s)                        \
    {                                                                                              \
        for (int i = 0; i < N; ++i) a[i] = a[i] OP s;                                              \
        return a;                                                                                  \
    }
RXM_GLM_BINOP(+)
RXM_GLM_BINOP(-)
RXM_GLM_BINOP(*)
RXM_GLM_BINOP(/)
#undef RXM_GLM_BINOP

template <int N, typename T>
__host__ __device__ inline vec<N, T> operator-(const vec<N, T>& a)
{
    vec<N, T> r;
    for (int i = 0; i < N; ++i) r[i] = -a[i];
    return r;
}

template <int N, typename T>
__host__ __device__ inline T dot(const vec<N, T>& a, const vec<N, T>& b)
{
    T s = T(0);
    for (int i = 0; i < N; ++i) s += a[i] * b[i];
    return s;
}
template <typename T>
__host__ __device__ inline vec<3, T> cross(const vec<3, T>& a, const vec<3, T>& b)
{
    return vec<3, T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <int N, typename T>
__host__ __device__ inline T length2(const vec<N, T>& a) { return dot(a, a); }
template <int N, typename T>
__host__ __device__ inline T length(const vec<N, T>& a) { return std::sqrt(dot(a, a)); }
template <int N, typename T>
__host__ __device__ inline T distance2(const vec<N, T>& a, const vec<N, T>& b) { return length2(a - b); }
template <int N, typename T>
__host__ __device__ inline T distance(const vec<N, T>& a, const vec<N, T>& b) { return length(a - b); }
template <int N, typename T>
__host__ __device__ inline vec<N, T> normalize(const vec<N, T>& a) { return a / length(a); }

}  // namespace glm
