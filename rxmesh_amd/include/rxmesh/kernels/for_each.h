// Kernels behind RXMeshStatic::for_each_* and for_each<Op, N>(lambda)
// (reference include/rxmesh/kernels/for_each.cuh:12-166 and
// kernels/query_kernel.cuh). One workgroup per patch; every owned element is
// visited exactly once, so a mesh-wide loop touches each element once.
#pragma once
#include "rxmesh/query.h"

namespace rxmesh {
namespace detail {

template <typename LambdaT>
__global__ void for_each_vertex(const uint32_t num_patches, const Context context, LambdaT apply)
{
    const uint32_t p = blockIdx.x;
    if (p >= num_patches) return;
    const rxm::PatchDesc d = rxm::load_desc(context.m.patch, p);
    for (uint32_t v = threadIdx.x; v < d.nvo; v += blockDim.x) apply(VertexHandle(p, static_cast<uint16_t>(v)));
}

template <typename LambdaT>
__global__ void for_each_edge(const uint32_t num_patches, const Context context, LambdaT apply)
{
    const uint32_t p = blockIdx.x;
    if (p >= num_patches) return;
    const rxm::PatchDesc d = rxm::load_desc(context.m.patch, p);
    for (uint32_t e = threadIdx.x; e < d.neo; e += blockDim.x) apply(EdgeHandle(p, static_cast<uint16_t>(e)));
}

template <typename LambdaT>
__global__ void for_each_face(const uint32_t num_patches, const Context context, LambdaT apply)
{
    const uint32_t p = blockIdx.x;
    if (p >= num_patches) return;
    const rxm::PatchDesc d = rxm::load_desc(context.m.patch, p);
    for (uint32_t f = threadIdx.x; f < d.nfo; f += blockDim.x) apply(FaceHandle(p, static_cast<uint16_t>(f)));
}

// for_each<Op, N>(lambda): one query per patch, the lambda per participant
template <uint32_t blockThreads, Op op, typename LambdaT>
__global__ __launch_bounds__(blockThreads) void query_kernel(const Context context, const bool oriented,
                                                             LambdaT user_lambda)
{
    auto                      block = cooperative_groups::this_thread_block();
    Query<blockThreads>       query(context);
    ShmemAllocator            shrd_alloc;
    query.template dispatch<op>(block, shrd_alloc, user_lambda, oriented);
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ dst, const size_t n, const T value)
{
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = value;
}

}  // namespace detail
}  // namespace rxmesh
