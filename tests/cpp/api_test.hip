// Exercises the header-only C++ API (RXMeshStatic, Query<N>::dispatch,
// Iterator, VertexAttribute, for_each_*) the way the reference's apps and
// tests use it, and exports the results through extern "C" functions so the
// pytest suite can compare them with the C-ABI and the oracle.
//
// Kernels below follow the structure of the reference's own users:
// apps/VertexNormal/vertex_normal_kernel.cuh:10-42 (FV dispatch + atomics),
// tests/RXMesh_test/query_kernel.cuh (one query, iterator copied out).
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "rxmesh/rxmesh_static.h"
#include "rxmesh/matrix/sparse_matrix.h"

using namespace rxmesh;

namespace {

std::vector<std::vector<uint32_t>> to_fv(const uint32_t* fv, uint32_t nF)
{
    std::vector<std::vector<uint32_t>> r(nF);
    for (uint32_t f = 0; f < nF; ++f) r[f] = {fv[3 * f], fv[3 * f + 1], fv[3 * f + 2]};
    return r;
}

// spelled exactly as a reference app kernel: no `template` disambiguators
template <typename T, uint32_t blockThreads>
__global__ void compute_vertex_normal(const rxmesh::Context context, rxmesh::VertexAttribute<T> coords,
                                      rxmesh::VertexAttribute<T> normals)
{
    using namespace rxmesh;
    auto vn_lambda = [&](FaceHandle face_id, VertexIterator& fv) {
        (void)face_id;
        const vec3<T> c0 = coords.to_glm<3>(fv[0]);
        const vec3<T> c1 = coords.to_glm<3>(fv[1]);
        const vec3<T> c2 = coords.to_glm<3>(fv[2]);
        const vec3<T> n  = glm::cross(c1 - c0, c2 - c0);
        const T       l[3] = {glm::distance2(c0, c1), glm::distance2(c1, c2), glm::distance2(c2, c0)};
        for (uint32_t v = 0; v < 3; ++v) {
            const T w = T(1) / (l[v] + l[(v + 2) % 3]);
            for (uint32_t i = 0; i < 3; ++i) atomicAdd(&normals(fv[v], i), n[i] * w);
        }
    };
    auto                block = cooperative_groups::this_thread_block();
    Query<blockThreads> query(context);
    ShmemAllocator      shrd_alloc;
    query.dispatch<Op::FV>(block, shrd_alloc, vn_lambda);
}

// the DiffARAP edge-weight kernel as the reference app spells it
// (apps/DiffARAP/arap_kernels.h:8-54), writing one weight per linear edge id
// instead of into a SparseMatrix
template <typename T, uint32_t blockThreads>
__global__ void calc_edge_weights(const rxmesh::Context context, const rxmesh::VertexAttribute<T> coords,
                                  T* edge_weights)
{
    using namespace rxmesh;
    auto calc_weights = [&](EdgeHandle edge_id, VertexIterator& vv) {
        const VertexHandle p_id = vv[0];
        const VertexHandle r_id = vv[2];
        const VertexHandle q_id = vv[1];
        const VertexHandle s_id = vv[3];
        if (!p_id.is_valid() || !r_id.is_valid() || !q_id.is_valid() || !s_id.is_valid()) {
            return;
        }
        const vec3<T> p = coords.to_glm<3>(p_id);
        const vec3<T> r = coords.to_glm<3>(r_id);
        const vec3<T> q = coords.to_glm<3>(q_id);
        const vec3<T> s = coords.to_glm<3>(s_id);
        T weight = 0;
        weight += dot((p - q), (r - q)) / length(cross(p - q, r - q));
        weight += dot((p - s), (r - s)) / length(cross(p - s, r - s));
        weight /= 2;
        weight = std::max(T(0), weight);
        edge_weights[context.linear_id(edge_id)] = weight;
    };
    auto                block = cooperative_groups::this_thread_block();
    Query<blockThreads> query(context);
    ShmemAllocator      shrd_alloc;
    query.dispatch<Op::EVDiamond>(block, shrd_alloc, calc_weights);
}

// the same kernel with the reference's SparseMatrix output, unchanged from
// apps/DiffARAP/arap_kernels.h:8-54 but for the name
template <typename T, uint32_t blockThreads>
__global__ static void calc_edge_weights_mat(const rxmesh::Context            context,
                                             const rxmesh::VertexAttribute<T> coords,
                                             rxmesh::SparseMatrix<T>          edge_weights)
{
    using namespace rxmesh;
    auto calc_weights = [&](EdgeHandle edge_id, VertexIterator& vv) {
        const VertexHandle p_id = vv[0];
        const VertexHandle r_id = vv[2];
        const VertexHandle q_id = vv[1];
        const VertexHandle s_id = vv[3];
        if (!p_id.is_valid() || !r_id.is_valid() || !q_id.is_valid() || !s_id.is_valid()) {
            return;
        }
        const vec3<T> p = coords.to_glm<3>(p_id);
        const vec3<T> r = coords.to_glm<3>(r_id);
        const vec3<T> q = coords.to_glm<3>(q_id);
        const vec3<T> s = coords.to_glm<3>(s_id);
        T weight = 0;
        weight += dot((p - q), (r - q)) / length(cross(p - q, r - q));
        weight += dot((p - s), (r - s)) / length(cross(p - s, r - s));
        weight /= 2;
        weight = std::max(0.f, weight);
        edge_weights(p_id, r_id) = weight;
        edge_weights(r_id, p_id) = weight;
    };
    auto                block = cooperative_groups::this_thread_block();
    Query<blockThreads> query(context);
    ShmemAllocator      shrd_alloc;
    query.dispatch<Op::EVDiamond>(block, shrd_alloc, calc_weights);
}

template <uint32_t blockThreads>
__global__ void valence_kernel(const Context context, uint32_t* out)
{
    auto                block = cooperative_groups::this_thread_block();
    Query<blockThreads> query(context);
    ShmemAllocator      shrd_alloc;
    query.compute_vertex_valence(block, shrd_alloc);
    const rxm::PatchDesc& d = query.get_patch_info();
    for (uint32_t v = threadIdx.x; v < d.nvo; v += blockThreads)
        out[context.linear_id(VertexHandle(query.get_patch_id(), uint16_t(v)))] = query.vertex_valence(uint16_t(v));
}

template <typename F>
__global__ __launch_bounds__(256) void run_body(F body)
{
    body();
}

// Two query passes (row sizes, then entries) of `op` through
// RXMeshStatic::for_each<op, 256>, written as a CSR over linear ids.
template <Op op, typename SrcH, typename IterT>
int run_query(RXMeshStatic& rx, bool oriented, int active_mod, int32_t* h_row, int32_t* h_col, int64_t col_cap)
{
    const Context  ctx  = rx.get_context();
    const uint32_t nsrc = rx.get_num_elements<SrcH>();
    int32_t *      cnt = nullptr, *col = nullptr;
    if (hipMalloc(&cnt, (nsrc + 1) * sizeof(int32_t)) != hipSuccess) return 5;
    if (hipMemset(cnt, 0, (nsrc + 1) * sizeof(int32_t)) != hipSuccess) return 5;
    auto active = [=] __device__(SrcH h) { return active_mod <= 1 || ctx.linear_id(h) % active_mod == 0; };
    {
        LaunchBox<256> lb;
        auto           count = [=] __device__(SrcH h, IterT & it) { cnt[ctx.linear_id(h)] = it.size(); };
        rx.prepare_launch_box({op}, lb, nullptr, oriented);
        // the active-set form of dispatch, through a user kernel
        auto body = [=] __device__() {
            auto           block = cooperative_groups::this_thread_block();
            Query<256>     query(ctx);
            ShmemAllocator shrd_alloc;
            query.dispatch<op>(block, shrd_alloc, count, active, oriented);
        };
        if (active_mod > 1) {
            run_body<<<lb.blocks, 256, lb.smem_bytes_dyn>>>(body);
        } else {
            rx.for_each<op, 256>(count, oriented);
        }
    }
    std::vector<int32_t> c(nsrc + 1);
    if (hipMemcpy(c.data(), cnt, nsrc * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess) return 5;
    h_row[0] = 0;
    for (uint32_t i = 0; i < nsrc; ++i) h_row[i + 1] = h_row[i] + c[i];
    const int64_t nnz = h_row[nsrc];
    if (nnz > col_cap) return 1;
    if (hipMemcpy(cnt, h_row, (nsrc + 1) * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) return 5;
    if (hipMalloc(&col, (nnz + 1) * sizeof(int32_t)) != hipSuccess) return 5;
    const int32_t* rp   = cnt;
    auto           fill = [=] __device__(SrcH h, IterT & it) {
        const int32_t b = rp[ctx.linear_id(h)];
        for (uint16_t i = 0; i < it.size(); ++i) col[b + i] = it[i].is_valid() ? int32_t(ctx.linear_id(it[i])) : -1;
    };
    if (active_mod > 1) {
        LaunchBox<256> lb;
        rx.prepare_launch_box({op}, lb, nullptr, oriented);
        auto body = [=] __device__() {
            auto           block = cooperative_groups::this_thread_block();
            Query<256>     query(ctx);
            ShmemAllocator shrd_alloc;
            query.dispatch<op>(block, shrd_alloc, fill, active, oriented);
        };
        run_body<<<lb.blocks, 256, lb.smem_bytes_dyn>>>(body);
    } else {
        rx.for_each<op, 256>(fill, oriented);
    }
    int rc = 0;
    if (hipMemcpy(h_col, col, nnz * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess) rc = 5;
    (void)hipFree(cnt);
    (void)hipFree(col);
    return rc;
}

}  // namespace

extern "C" {

// CSR over linear ids of query `op` (reference Op values). active_mod > 1
// restricts the sources to linear id % active_mod == 0 through the
// active-set predicate. Returns 0, 1 (bad argument / capacity), 5 (HIP error)
// or 6 (exception).
int api_query(const uint32_t* fv, uint32_t nF, uint32_t patch_size, int op, int oriented, int active_mod,
              int32_t* row_ptr, int32_t* col, int64_t col_cap)
{
    try {
        auto         f = to_fv(fv, nF);
        RXMeshStatic rx(f, "", patch_size);
        const bool   o = oriented != 0;
        switch (op) {
            case 3: return run_query<Op::VV, VertexHandle, VertexIterator>(rx, o, active_mod, row_ptr, col, col_cap);
            case 4: return run_query<Op::VE, VertexHandle, EdgeIterator>(rx, o, active_mod, row_ptr, col, col_cap);
            case 5: return run_query<Op::VF, VertexHandle, FaceIterator>(rx, o, active_mod, row_ptr, col, col_cap);
            case 6: return run_query<Op::FV, FaceHandle, VertexIterator>(rx, o, active_mod, row_ptr, col, col_cap);
            case 7: return run_query<Op::FE, FaceHandle, EdgeIterator>(rx, o, active_mod, row_ptr, col, col_cap);
            case 8: return run_query<Op::FF, FaceHandle, FaceIterator>(rx, o, active_mod, row_ptr, col, col_cap);
            case 9: return run_query<Op::EV, EdgeHandle, VertexIterator>(rx, o, active_mod, row_ptr, col, col_cap);
            case 11: return run_query<Op::EF, EdgeHandle, FaceIterator>(rx, o, active_mod, row_ptr, col, col_cap);
            case 12:
                return run_query<Op::EVDiamond, EdgeHandle, VertexIterator>(rx, o, active_mod, row_ptr, col, col_cap);
            default: return 1;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "api_query: %s\n", e.what());
        return 6;
    }
}

// Max-weighted vertex normals (reference vertex normal app) of xyz (V x 3,
// input order) written in input order.
int api_vertex_normals(const uint32_t* fv, uint32_t nF, const float* xyz, uint32_t nV, uint32_t patch_size,
                       float* out)
{
    try {
        auto                            f = to_fv(fv, nF);
        RXMeshStatic                    rx(f, "", patch_size);
        std::vector<std::vector<float>> verts(nV, std::vector<float>(3));
        for (uint32_t v = 0; v < nV; ++v)
            for (int i = 0; i < 3; ++i) verts[v][i] = xyz[3 * v + i];
        rx.add_vertex_coordinates(verts);
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
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "api_vertex_normals: %s\n", e.what());
        return 6;
    }
}

// ARAP edge weights through the reference-style kernel above, per linear
// edge id (0 where the kernel skips the edge).
int api_arap_weights(const uint32_t* fv, uint32_t nF, const float* xyz, uint32_t nV, uint32_t patch_size,
                     float* out)
{
    try {
        auto                            f = to_fv(fv, nF);
        RXMeshStatic                    rx(f, "", patch_size);
        std::vector<std::vector<float>> verts(nV, std::vector<float>(3));
        for (uint32_t v = 0; v < nV; ++v)
            for (int i = 0; i < 3; ++i) verts[v][i] = xyz[3 * v + i];
        rx.add_vertex_coordinates(verts);
        auto           coords = rx.get_input_vertex_coordinates();
        const uint32_t nE     = rx.get_num_edges();
        float*         w      = nullptr;
        if (hipMalloc(&w, nE * sizeof(float)) != hipSuccess) return 5;
        if (hipMemset(w, 0, nE * sizeof(float)) != hipSuccess) return 5;
        LaunchBox<256> lb;
        rx.prepare_launch_box({Op::EVDiamond}, lb, (void*)calc_edge_weights<float, 256>);
        calc_edge_weights<float, 256><<<lb.blocks, lb.num_threads, lb.smem_bytes_dyn>>>(rx.get_context(), *coords, w);
        int rc = (hipMemcpy(out, w, nE * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess) ? 0 : 5;
        (void)hipFree(w);
        return rc;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "api_arap_weights: %s\n", e.what());
        return 6;
    }
}

// ARAP edge weights through the SparseMatrix form of the kernel: out[2 e]
// and out[2 e + 1] are the (v0, v1) and (v1, v0) entries of linear edge e,
// read back through SparseMatrix::operator() in an EV pass.
int api_arap_weights_mat(const uint32_t* fv, uint32_t nF, const float* xyz, uint32_t nV, uint32_t patch_size,
                         float* out)
{
    try {
        auto                            f = to_fv(fv, nF);
        RXMeshStatic                    rx(f, "", patch_size);
        std::vector<std::vector<float>> verts(nV, std::vector<float>(3));
        for (uint32_t v = 0; v < nV; ++v)
            for (int i = 0; i < 3; ++i) verts[v][i] = xyz[3 * v + i];
        rx.add_vertex_coordinates(verts);
        auto                 coords = rx.get_input_vertex_coordinates();
        SparseMatrix<float>  weight_matrix(rx);
        weight_matrix.reset(0.f, LOCATION_ALL);
        LaunchBox<256> lb;
        rx.prepare_launch_box({Op::EVDiamond}, lb, (void*)calc_edge_weights_mat<float, 256>);
        calc_edge_weights_mat<float, 256>
            <<<lb.blocks, lb.num_threads, lb.smem_bytes_dyn>>>(rx.get_context(), *coords, weight_matrix);
        const uint32_t nE  = rx.get_num_edges();
        float*         d   = nullptr;
        const Context  ctx = rx.get_context();
        if (hipMalloc(&d, 2 * nE * sizeof(float)) != hipSuccess) return 5;
        rx.for_each<Op::EV, 256>([=] __device__(EdgeHandle eh, VertexIterator & ev) {
            const uint32_t e = ctx.linear_id(eh);
            d[2 * e]         = weight_matrix(ev[0], ev[1]);
            d[2 * e + 1]     = weight_matrix(ev[1], ev[0]);
        });
        int rc = (hipMemcpy(out, d, 2 * nE * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess) ? 0 : 5;
        (void)hipFree(d);
        weight_matrix.release();
        return rc;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "api_arap_weights_mat: %s\n", e.what());
        return 6;
    }
}

// Checks of the host-side API on one mesh: for_each_vertex on DEVICE and
// HOST, attribute move / reset / copy_from, handle <-> global id round trips,
// vertex valence vs VV row sizes. Returns the number of failed checks (or
// a negative code on error); `report` receives one line per failure.
int api_misc(const uint32_t* fv, uint32_t nF, uint32_t patch_size, char* report, int report_len)
{
    int         fails = 0;
    std::string msg;
    auto        fail = [&](const char* what) {
        ++fails;
        msg += what;
        msg += "\n";
    };
    try {
        auto         f = to_fv(fv, nF);
        RXMeshStatic rx(f, "", patch_size);
        const uint32_t V = rx.get_num_vertices();
        // global -> handle -> global for every vertex, edge and face
        for (uint32_t g = 0; g < V; ++g)
            if (rx.map_to_global(rx.get_vertex_handle(g)) != g) { fail("vertex handle round trip"); break; }
        for (uint32_t g = 0; g < rx.get_num_edges(); ++g)
            if (rx.map_to_global(rx.get_edge_handle(g)) != g) { fail("edge handle round trip"); break; }
        for (uint32_t g = 0; g < rx.get_num_faces(); ++g)
            if (rx.map_to_global(rx.get_face_handle(g)) != g) { fail("face handle round trip"); break; }
        // device for_each writes the linear id; host for_each checks it
        auto a = rx.add_vertex_attribute<uint32_t>("lin", 1);
        a->reset(0xFFFFFFFFu, LOCATION_ALL);
        const Context ctx = rx.get_context();
        auto          av  = *a;
        rx.for_each_vertex(DEVICE, [=] __device__(const VertexHandle vh) { av(vh) = ctx.linear_id(vh); });
        a->move(DEVICE, HOST);
        uint32_t visited = 0;
        rx.for_each_vertex(HOST, [&](const VertexHandle vh) {
            ++visited;
            if ((*a)(vh) != rx.linear_id(vh)) fail("for_each_vertex DEVICE value");
        });
        if (visited != V) fail("for_each_vertex HOST count");
        // copy_from + reset
        auto b = rx.add_vertex_attribute_like("lin2", *a);
        b->copy_from(*a, DEVICE, DEVICE);
        b->move(DEVICE, HOST);
        for (uint32_t i = 0; i < V; ++i)
            if ((*b)(i) != (*a)(i)) { fail("copy_from"); break; }
        b->reset(7u, DEVICE);
        b->move(DEVICE, HOST);
        for (uint32_t i = 0; i < V; ++i)
            if ((*b)(i) != 7u) { fail("reset"); break; }
        rx.remove_attribute("lin2");
        if (rx.does_attribute_exist("lin2")) fail("remove_attribute");
        // edges and faces visited once each on the device
        auto ce = rx.add_edge_attribute<uint32_t>("ce", 1);
        auto cf = rx.add_face_attribute<uint32_t>("cf", 1);
        ce->reset(0u, DEVICE);
        cf->reset(0u, DEVICE);
        auto cev = *ce;
        auto cfv = *cf;
        rx.for_each_edge(DEVICE, [=] __device__(const EdgeHandle eh) { atomicAdd(&cev(eh), 1u); });
        rx.for_each_face(DEVICE, [=] __device__(const FaceHandle fh) { atomicAdd(&cfv(fh), 1u); });
        ce->move(DEVICE, HOST);
        cf->move(DEVICE, HOST);
        for (uint32_t i = 0; i < rx.get_num_edges(); ++i)
            if ((*ce)(i) != 1u) { fail("for_each_edge"); break; }
        for (uint32_t i = 0; i < rx.get_num_faces(); ++i)
            if ((*cf)(i) != 1u) { fail("for_each_face"); break; }
        // valence in LDS == VV row size
        uint32_t* d_val = nullptr;
        int32_t*  d_cnt = nullptr;
        if (hipMalloc(&d_val, V * 4) != hipSuccess || hipMalloc(&d_cnt, V * 4) != hipSuccess) return -5;
        LaunchBox<256> lb;
        rx.prepare_launch_box({}, lb, (void*)valence_kernel<256>, false, true);
        valence_kernel<256><<<lb.blocks, 256, lb.smem_bytes_dyn>>>(ctx, d_val);
        rx.for_each<Op::VV, 256>(
            [=] __device__(VertexHandle vh, VertexIterator & it) { d_cnt[ctx.linear_id(vh)] = it.size(); });
        std::vector<uint32_t> hv(V);
        std::vector<int32_t>  hc(V);
        if (hipMemcpy(hv.data(), d_val, V * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hc.data(), d_cnt, V * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return -5;
        for (uint32_t i = 0; i < V; ++i)
            if (hv[i] != uint32_t(hc[i])) { fail("vertex valence"); break; }
        (void)hipFree(d_val);
        (void)hipFree(d_cnt);
    } catch (const std::exception& e) {
        msg += e.what();
        std::strncpy(report, msg.c_str(), report_len - 1);
        report[report_len - 1] = 0;
        return -6;
    }
    std::strncpy(report, msg.c_str(), report_len - 1);
    report[report_len - 1] = 0;
    return fails;
}

}  // extern "C"
