// Device-side view of an RXMeshStatic, passed by value to kernels (the role
// of the reference's Context, include/rxmesh/context.h:15-441): the patch
// slabs plus the linear-id prefix of each element kind.
#pragma once
#include "device_mesh.hpp"
#include "rxmesh/handle.h"

namespace rxmesh {

struct Context
{
    rxm::DevMesh    m{};
    const uint32_t* v_prefix       = nullptr;  // P + 1, device: first linear id of each patch's owned vertices
    const uint32_t* e_prefix       = nullptr;
    const uint32_t* f_prefix       = nullptr;
    uint32_t        max_edge_faces = 2;        // largest edge -> face degree (2 on a manifold)

    __host__ __device__ uint32_t get_num_patches() const { return m.P; }
    __host__ __device__ uint32_t get_num_vertices() const { return m.nV; }
    __host__ __device__ uint32_t get_num_edges() const { return m.nE; }
    __host__ __device__ uint32_t get_num_faces() const { return m.nF; }

    // linear id of an owned handle, prefix[p] + local (reference context.h:276-331)
    __device__ uint32_t linear_id(VertexHandle h) const { return v_prefix[h.patch_id()] + h.local_id(); }
    __device__ uint32_t linear_id(EdgeHandle h) const { return e_prefix[h.patch_id()] + h.local_id(); }
    __device__ uint32_t linear_id(FaceHandle h) const { return f_prefix[h.patch_id()] + h.local_id(); }
};

}  // namespace rxmesh
