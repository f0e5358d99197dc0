// TEST INFRASTRUCTURE. Runs the reference's own app kernels, compiled
// unchanged from the reference tree (apps/VertexNormal/vertex_normal_kernel.cuh
// and apps/MCF/mcf_kernels.cuh, found through the include path the recipe in
// oracle/Makefile sets), on this library's C++ API mirror (RXMeshStatic,
// Query, Attribute, SparseMatrix, DenseMatrix, geometry_util), and exports
// their results in the caller's (global) ids so the tests can compare them
// with the C-ABI kernels. Built into oracle/_ref/ only where the reference
// tree is present; nothing in the product loads it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "rxmesh/matrix/sparse_matrix.h"
#include "rxmesh/rxmesh_static.h"

#include "mcf_kernels.cuh"           // the reference's MCF app kernels
#include "vertex_normal_kernel.cuh"  // the reference's vertex-normal kernel

using namespace rxmesh;

namespace {

std::vector<std::vector<uint32_t>> faces(const uint32_t* fv, uint32_t nF)
{
    std::vector<std::vector<uint32_t>> r(nF);
    for (uint32_t f = 0; f < nF; ++f) r[f] = {fv[3 * f], fv[3 * f + 1], fv[3 * f + 2]};
    return r;
}

void set_coordinates(RXMeshStatic& rx, const float* xyz, uint32_t nV)
{
    std::vector<std::vector<float>> verts(nV, std::vector<float>(3));
    for (uint32_t v = 0; v < nV; ++v)
        for (int i = 0; i < 3; ++i) verts[v][i] = xyz[3 * v + i];
    rx.add_vertex_coordinates(verts);
}

template <typename F>
int guarded(const char* what, F body)
{
    try {
        body();
        return hipDeviceSynchronize() == hipSuccess ? 0 : 5;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s: %s\n", what, e.what());
        return 6;
    }
}

}  // namespace

extern "C" {

// compute_vertex_normal (apps/VertexNormal/vertex_normal_kernel.cuh) into
// out (3 V, global vertex order, AoS).
int ref_vertex_normals(const uint32_t* fv, uint32_t nF, const float* xyz, uint32_t nV, uint32_t patch_size,
                       float* out)
{
    return guarded("ref_vertex_normals", [&] {
        auto         f = faces(fv, nF);
        RXMeshStatic rx(f, "", patch_size);
        set_coordinates(rx, xyz, nV);
        auto coords  = rx.get_input_vertex_coordinates();
        auto normals = rx.add_vertex_attribute<float>("normals", 3, LOCATION_ALL);
        normals->reset(0.f, DEVICE);
        LaunchBox<256> lb;
        rx.prepare_launch_box({Op::FV}, lb, (void*)compute_vertex_normal<float, 256>);
        compute_vertex_normal<float, 256>
            <<<lb.blocks, lb.num_threads, lb.smem_bytes_dyn>>>(rx.get_context(), *coords, *normals);
        normals->move(DEVICE, HOST);
        rx.for_each_vertex(HOST, [&](const VertexHandle vh) {
            const uint32_t g = rx.map_to_global(vh);
            for (int i = 0; i < 3; ++i) out[3 * g + i] = (*normals)(vh, i);
        });
    });
}

// mcf_A_setup and mcf_B_setup (apps/MCF/mcf_kernels.cuh:302-426) into a
// SparseMatrix and a DenseMatrix: A as COO in global ids (rows, cols, vals:
// capacity cap entries, *nnz set), B as 3 V AoS in global order.
int ref_mcf_system(const uint32_t* fv, uint32_t nF, const float* xyz, uint32_t nV, uint32_t patch_size, float dt,
                   int uniform, int32_t* rows, int32_t* cols, float* vals, int64_t cap, int64_t* nnz, float* B_out)
{
    return guarded("ref_mcf_system", [&] {
        auto         f = faces(fv, nF);
        RXMeshStatic rx(f, "", patch_size);
        set_coordinates(rx, xyz, nV);
        auto                coords = rx.get_input_vertex_coordinates();
        SparseMatrix<float> A(rx);
        DenseMatrix<float>  B(rx, rx.get_num_vertices(), 3);
        LaunchBox<256>      lb;
        rx.prepare_launch_box({Op::VV}, lb, (void*)mcf_B_setup<float, 256>, !uniform);
        mcf_B_setup<float, 256><<<lb.blocks, lb.num_threads, lb.smem_bytes_dyn>>>(rx.get_context(), *coords, B,
                                                                                    uniform != 0);
        rx.prepare_launch_box({Op::VV}, lb, (void*)mcf_A_setup<float, 256>, !uniform);
        mcf_A_setup<float, 256><<<lb.blocks, lb.num_threads, lb.smem_bytes_dyn>>>(rx.get_context(), *coords, A,
                                                                                    uniform != 0, dt);
        if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("kernel failed");
        A.move(DEVICE, HOST);
        B.move(DEVICE, HOST);
        std::vector<uint32_t> l2g(rx.get_num_vertices());
        rx.for_each_vertex(HOST, [&](const VertexHandle vh) {
            l2g[rx.linear_id(vh)] = rx.map_to_global(vh);
            for (int i = 0; i < 3; ++i) B_out[3 * rx.map_to_global(vh) + i] = B(rx.linear_id(vh), i);
        });
        const int32_t* rp = A.row_ptr(HOST);
        const int32_t* ci = A.col_idx(HOST);
        const float*   va = A.get_val(HOST);
        int64_t        k  = 0;
        for (uint32_t r = 0; r < A.rows(); ++r)
            for (int32_t j = rp[r]; j < rp[r + 1]; ++j, ++k) {
                if (k >= cap) throw std::runtime_error("COO capacity");
                rows[k] = static_cast<int32_t>(l2g[r]);
                cols[k] = static_cast<int32_t>(l2g[ci[j]]);
                vals[k] = va[j];
            }
        *nnz = k;
        A.release();
        B.release();
    });
}

// One product of the matrix-free operator, matvec (apps/MCF/mcf_kernels.cuh:
// 114-213): out = A in, in and out 3 V AoS in global order.
int ref_mcf_matvec(const uint32_t* fv, uint32_t nF, const float* xyz, uint32_t nV, uint32_t patch_size, float dt,
                   int uniform, const float* in, float* out)
{
    return guarded("ref_mcf_matvec", [&] {
        auto         f = faces(fv, nF);
        RXMeshStatic rx(f, "", patch_size);
        set_coordinates(rx, xyz, nV);
        auto coords = rx.get_input_vertex_coordinates();
        auto vin    = rx.add_vertex_attribute<float>("in", 3, LOCATION_ALL);
        auto vout   = rx.add_vertex_attribute<float>("out", 3, LOCATION_ALL);
        rx.for_each_vertex(HOST, [&](const VertexHandle vh) {
            for (int i = 0; i < 3; ++i) (*vin)(vh, i) = in[3 * rx.map_to_global(vh) + i];
        });
        vin->move(HOST, DEVICE);
        LaunchBox<256> lb;
        rx.prepare_launch_box({Op::VV}, lb, (void*)matvec<float, 256>, !uniform);
        matvec<float, 256><<<lb.blocks, lb.num_threads, lb.smem_bytes_dyn>>>(rx.get_context(), *coords, *vin, *vout,
                                                                               uniform != 0, dt);
        vout->move(DEVICE, HOST);
        rx.for_each_vertex(HOST, [&](const VertexHandle vh) {
            for (int i = 0; i < 3; ++i) out[3 * rx.map_to_global(vh) + i] = (*vout)(vh, i);
        });
    });
}

}  // extern "C"
