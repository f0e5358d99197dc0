// RXMeshStatic: the host object of the C++ API (reference
// include/rxmesh/rxmesh_static.h:34-1100, rxmesh.h:29-660). It builds the
// patch mesh through the C-ABI (rxm_mesh_create, which runs the patcher and
// the connectivity build and uploads the slabs), keeps the device view as a
// Context, owns the attributes, and launches user kernels and lambdas with
// one workgroup per patch.
//
// Differences from the reference: errors throw std::runtime_error instead
// of exiting; capacity / patch-alloc / hashtable factors are accepted and
// ignored (a static mesh needs no spare capacity and no hash tables);
// with_omp is ignored (host loops are serial).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <functional>
#include <map>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "rxmesh/attribute.h"
#include "rxmesh/context.h"
#include "rxmesh/kernels/for_each.h"
#include "rxmesh/launch_box.h"
#include "rxmesh/query.h"
#include "rxmesh/util/import_obj.h"
#include "rxmesh_amd.h"

namespace rxmesh {

namespace detail {
// host_callable<F, H>::value: F(H) can be called from host code (false for a
// __device__ lambda); the call is checked in the context of a __host__ function
template <typename F, typename H>
__host__ auto host_probe(F& f, H h, int) -> decltype(f(h), std::true_type());
template <typename F, typename H>
__host__ std::false_type host_probe(F&, H, long);
template <typename F, typename H>
using host_callable = decltype(host_probe(std::declval<F&>(), std::declval<H>(), 0));
}  // namespace detail

class RXMeshStatic
{
   public:
    RXMeshStatic(const RXMeshStatic&)            = delete;
    RXMeshStatic& operator=(const RXMeshStatic&) = delete;

    explicit RXMeshStatic(const std::string file_path, const std::string patcher_file = "",
                          const uint32_t patch_size = 512, const float capacity_factor = 1.0,
                          const float patch_alloc_factor = 1.0, const float lp_hashtable_load_factor = 0.8)
    {
        (void)capacity_factor;
        (void)patch_alloc_factor;
        (void)lp_hashtable_load_factor;
        std::vector<std::vector<float>>    verts;
        std::vector<std::vector<uint32_t>> fv;
        if (!import_obj(file_path, verts, fv)) throw std::runtime_error("RXMeshStatic: cannot read " + file_path);
        init(fv, static_cast<uint32_t>(verts.size()), patcher_file, patch_size);
        add_vertex_coordinates(verts);
    }

    explicit RXMeshStatic(std::vector<std::vector<uint32_t>>& fv, const std::string patcher_file = "",
                          const uint32_t patch_size = 512, const float capacity_factor = 1.0,
                          const float patch_alloc_factor = 1.0, const float lp_hashtable_load_factor = 0.8)
    {
        (void)capacity_factor;
        (void)patch_alloc_factor;
        (void)lp_hashtable_load_factor;
        init(fv, 0, patcher_file, patch_size);
    }

    virtual ~RXMeshStatic()
    {
        for (auto it = m_attr_order.rbegin(); it != m_attr_order.rend(); ++it) m_releasers[*it]();
        m_releasers.clear();
        for (uint32_t* p : m_dev_arrays) (void)hipFree(p);
        if (m_mesh) rxm_mesh_destroy(m_mesh);
    }

    // ---- sizes and statistics (reference rxmesh.h:44-200) ---------------
    uint32_t get_num_vertices() const { return m_info.num_vertices; }
    uint32_t get_num_edges() const { return m_info.num_edges; }
    uint32_t get_num_faces() const { return m_info.num_faces; }
    uint32_t get_num_patches() const { return m_info.num_patches; }
    uint32_t get_patch_size() const { return m_info.patch_size; }
    uint32_t get_max_valence() const { return m_info.max_valence; }
    uint32_t get_max_edge_incident_faces() const { return m_info.max_edge_incident_faces; }
    uint32_t get_max_face_adjacent_faces() const { return m_info.max_face_adjacent_faces; }
    bool     is_closed() const { return m_info.is_closed != 0; }
    bool     is_edge_manifold() const { return m_info.is_edge_manifold != 0; }
    bool     is_consistently_oriented() const { return m_info.is_consistently_oriented != 0; }
    uint32_t get_num_vertices(const uint32_t p) const { return m_counts[6 * p + 0]; }
    uint32_t get_num_edges(const uint32_t p) const { return m_counts[6 * p + 1]; }
    uint32_t get_num_faces(const uint32_t p) const { return m_counts[6 * p + 2]; }
    uint32_t get_num_owned_vertices(const uint32_t p) const { return m_counts[6 * p + 3]; }
    uint32_t get_num_owned_edges(const uint32_t p) const { return m_counts[6 * p + 4]; }
    uint32_t get_num_owned_faces(const uint32_t p) const { return m_counts[6 * p + 5]; }
    template <typename HandleT>
    uint32_t get_num_elements() const
    {
        constexpr int k = detail::HandleKind<HandleT>::value;
        return k == 0 ? get_num_vertices() : (k == 1 ? get_num_edges() : get_num_faces());
    }
    const Context& get_context() const { return m_ctx; }
    rxm_mesh*      get_c_handle() const { return m_mesh; }
    uint64_t       get_device_bytes() const { return m_info.device_bytes; }

    // ---- handles <-> ids ------------------------------------------------
    // linear id of (the owner's copy of) a handle, in [0, num elements)
    template <typename HandleT>
    uint32_t linear_id(HandleT h) const
    {
        return m_index[detail::HandleKind<HandleT>::value].linear(h.patch_id(), h.local_id());
    }
    // id of the element in the input mesh (reference rxmesh_static.h:874-889)
    uint32_t map_to_global(const VertexHandle vh) const { return m_l2g[0][linear_id(vh)]; }
    uint32_t map_to_global(const EdgeHandle eh) const { return m_l2g[1][linear_id(eh)]; }
    uint32_t map_to_global(const FaceHandle fh) const { return m_l2g[2][linear_id(fh)]; }
    // owner handle of an input-mesh element
    VertexHandle get_vertex_handle(uint32_t global) const { return handle_of<VertexHandle>(m_g2l[0][global]); }
    EdgeHandle   get_edge_handle(uint32_t global) const { return handle_of<EdgeHandle>(m_g2l[1][global]); }
    FaceHandle   get_face_handle(uint32_t global) const { return handle_of<FaceHandle>(m_g2l[2][global]); }

    // ---- coordinates ----------------------------------------------------
    void add_vertex_coordinates(std::vector<std::vector<float>>& vertices, std::string mesh_name = "")
    {
        (void)mesh_name;
        if (vertices.size() != get_num_vertices())
            throw std::runtime_error("add_vertex_coordinates: vertex count does not match the mesh");
        if (does_attribute_exist("rx:vertices")) remove_attribute("rx:vertices");
        m_coords = add_vertex_attribute<float>(vertices, "rx:vertices", SoA);
    }
    std::shared_ptr<VertexAttribute<float>> get_input_vertex_coordinates()
    {
        if (!m_coords) throw std::runtime_error("get_input_vertex_coordinates: the mesh has no coordinates");
        return m_coords;
    }

    // ---- loops over owned elements (reference rxmesh_static.h:205-405) --
    template <typename LambdaT>
    void for_each_vertex(locationT location, LambdaT apply, hipStream_t stream = nullptr, bool with_omp = true)
    {
        for_each_kind<0>(location, apply, stream, with_omp);
    }
    template <typename LambdaT>
    void for_each_edge(locationT location, LambdaT apply, hipStream_t stream = nullptr, bool with_omp = true)
    {
        for_each_kind<1>(location, apply, stream, with_omp);
    }
    template <typename LambdaT>
    void for_each_face(locationT location, LambdaT apply, hipStream_t stream = nullptr, bool with_omp = true)
    {
        for_each_kind<2>(location, apply, stream, with_omp);
    }
    template <typename HandleT, typename LambdaT>
    void for_each(locationT location, LambdaT apply, hipStream_t stream = nullptr, bool with_omp = true)
    {
        for_each_kind<detail::HandleKind<HandleT>::value>(location, apply, stream, with_omp);
    }

    // ---- query lambdas (reference rxmesh_static.h:524-566) --------------
    template <Op op, uint32_t blockThreads, typename LambdaT>
    void for_each(const LambdaT user_lambda, const bool oriented = false, hipStream_t stream = nullptr) const
    {
        LaunchBox<blockThreads> lb;
        prepare_launch_box({op}, lb, (void*)detail::query_kernel<blockThreads, op, LambdaT>, oriented);
        for_each<op>(lb, user_lambda, oriented, stream);
    }
    template <Op op, uint32_t blockThreads, typename LambdaT>
    void for_each(LaunchBox<blockThreads> lb, const LambdaT user_lambda, const bool oriented = false,
                  hipStream_t stream = nullptr) const
    {
        if (!lb.blocks) return;
        detail::query_kernel<blockThreads, op>
            <<<lb.blocks, lb.num_threads, lb.smem_bytes_dyn, stream>>>(get_context(), oriented, user_lambda);
        detail::hip_check(hipGetLastError(), "RXMeshStatic::for_each");
    }

    // ---- user kernels (reference rxmesh_static.h:408-505) ----------------
    template <uint32_t blockThreads, typename KernelT, typename... ArgsT>
    void run_kernel(const LaunchBox<blockThreads>& lb, KernelT kernel, hipStream_t stream, ArgsT... args) const
    {
        if (!lb.blocks) return;
        kernel<<<lb.blocks, lb.num_threads, lb.smem_bytes_dyn, stream>>>(get_context(), args...);
        detail::hip_check(hipGetLastError(), "RXMeshStatic::run_kernel");
    }
    template <uint32_t blockThreads, typename KernelT, typename... ArgsT>
    void run_kernel(const LaunchBox<blockThreads>& lb, KernelT kernel, ArgsT... args) const
    {
        run_kernel<blockThreads>(lb, kernel, hipStream_t(nullptr), args...);
    }
    template <uint32_t blockThreads, typename KernelT, typename... ArgsT>
    void run_kernel(const std::vector<Op> op, KernelT kernel, ArgsT... args) const
    {
        run_kernel<blockThreads>(kernel, op, false, false, false,
                                 [](uint32_t, uint32_t, uint32_t) { return size_t(0); }, hipStream_t(nullptr),
                                 args...);
    }
    template <uint32_t blockThreads, typename KernelT, typename... ArgsT>
    void run_kernel(KernelT kernel, const std::vector<Op> op, const bool oriented, const bool with_vertex_valence,
                    const bool is_concurrent, std::function<size_t(uint32_t, uint32_t, uint32_t)> user_shmem,
                    hipStream_t stream, ArgsT... args) const
    {
        LaunchBox<blockThreads> lb;
        prepare_launch_box(op, lb, (void*)kernel, oriented, with_vertex_valence, is_concurrent, user_shmem);
        run_kernel<blockThreads>(lb, kernel, stream, args...);
    }

    // Grid = one workgroup per patch; dynamic LDS = the largest query carve
    // of `op` on the largest patch (summed when is_concurrent), plus the
    // vertex valence table and user_shmem(max V, max E, max F) bytes
    // (reference rxmesh_static.h:585-594, rxmesh_static.inl:443-841).
    template <uint32_t blockThreads>
    void prepare_launch_box(const std::vector<Op> op, LaunchBox<blockThreads>& launch_box, const void* kernel,
                            const bool oriented = false, const bool with_vertex_valence = false,
                            const bool is_concurrent = false,
                            std::function<size_t(uint32_t, uint32_t, uint32_t)> user_shmem =
                                [](uint32_t, uint32_t, uint32_t) { return size_t(0); }) const
    {
        const uint32_t nv = m_info.max_vertices_per_patch, ne = m_info.max_edges_per_patch,
                       nf = m_info.max_faces_per_patch;
        if (oriented && !(is_edge_manifold() && is_consistently_oriented()))
            throw std::runtime_error(
                "prepare_launch_box: oriented queries need an edge-manifold, consistently oriented mesh");
        size_t q = 0;
        for (Op o : op) {
            if (o == Op::V || o == Op::E || o == Op::F) continue;
            if (!detail::op_supported(o)) throw std::runtime_error("prepare_launch_box: unsupported query");
            const size_t b =
                detail::query_shmem_bytes(o, nv, ne, nf, oriented && (o == Op::VV || o == Op::VE),
                                          m_ctx.max_edge_faces);
            q = is_concurrent ? q + b : std::max(q, b);
        }
        if (with_vertex_valence) q += 4u * size_t(nv) + 16u;
        q += user_shmem(nv, ne, nf);
        launch_box.blocks         = get_num_patches();
        launch_box.smem_bytes_dyn = q;
        if (kernel) {
            hipFuncAttributes attr{};
            detail::hip_check(hipFuncGetAttributes(&attr, kernel), "prepare_launch_box: hipFuncGetAttributes");
            launch_box.num_registers_per_thread = attr.numRegs;
            launch_box.smem_bytes_static        = attr.sharedSizeBytes;
            launch_box.local_mem_per_thread     = attr.localSizeBytes;
            if (q + attr.sharedSizeBytes > kMaxLDS)
                throw std::runtime_error("prepare_launch_box: patches need more LDS than a workgroup has");
            if (q > 65536u)
                detail::hip_check(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                      static_cast<int>(q)),
                                  "prepare_launch_box: hipFuncSetAttribute");
        }
    }

    // ---- attributes (reference rxmesh_static.h:600-806) ------------------
    template <class T, class HandleT>
    std::shared_ptr<Attribute<T, HandleT>> add_attribute(const std::string& name, uint32_t num_attributes,
                                                          locationT location = LOCATION_ALL,
                                                          layoutT   layout   = AoSoA)
    {
        if (does_attribute_exist(name)) throw std::runtime_error("add_attribute: '" + name + "' exists");
        auto a = std::make_shared<Attribute<T, HandleT>>(name.c_str(), num_attributes, location, layout,
                                                         m_index[detail::HandleKind<HandleT>::value]);
        m_releasers[name] = [a]() mutable { a->release(); };
        m_attr_order.push_back(name);
        return a;
    }
    template <class T, class HandleT>
    std::shared_ptr<Attribute<T, HandleT>> add_attribute_like(const std::string& name,
                                                               const Attribute<T, HandleT>& other)
    {
        return add_attribute<T, HandleT>(name, other.get_num_attributes(), other.get_allocated(),
                                         other.get_layout());
    }
    template <class T>
    std::shared_ptr<VertexAttribute<T>> add_vertex_attribute(const std::string& name, uint32_t num_attributes,
                                                             locationT location = LOCATION_ALL,
                                                             layoutT   layout   = AoSoA)
    {
        return add_attribute<T, VertexHandle>(name, num_attributes, location, layout);
    }
    template <class T>
    std::shared_ptr<EdgeAttribute<T>> add_edge_attribute(const std::string& name, uint32_t num_attributes,
                                                         locationT location = LOCATION_ALL, layoutT layout = AoSoA)
    {
        return add_attribute<T, EdgeHandle>(name, num_attributes, location, layout);
    }
    template <class T>
    std::shared_ptr<FaceAttribute<T>> add_face_attribute(const std::string& name, uint32_t num_attributes,
                                                         locationT location = LOCATION_ALL, layoutT layout = AoSoA)
    {
        return add_attribute<T, FaceHandle>(name, num_attributes, location, layout);
    }
    template <class T>
    std::shared_ptr<VertexAttribute<T>> add_vertex_attribute_like(const std::string& name,
                                                                  const VertexAttribute<T>& other)
    {
        return add_attribute_like<T, VertexHandle>(name, other);
    }
    template <class T>
    std::shared_ptr<EdgeAttribute<T>> add_edge_attribute_like(const std::string& name, const EdgeAttribute<T>& other)
    {
        return add_attribute_like<T, EdgeHandle>(name, other);
    }
    template <class T>
    std::shared_ptr<FaceAttribute<T>> add_face_attribute_like(const std::string& name, const FaceAttribute<T>& other)
    {
        return add_attribute_like<T, FaceHandle>(name, other);
    }
    // values given per input-mesh element (row = global id), copied to both sides
    template <class T>
    std::shared_ptr<VertexAttribute<T>> add_vertex_attribute(const std::vector<std::vector<T>>& values,
                                                             const std::string& name, layoutT layout = AoSoA)
    {
        return add_attribute_from<T, VertexHandle>(values, name, layout);
    }
    template <class T>
    std::shared_ptr<VertexAttribute<T>> add_vertex_attribute(const std::vector<T>& values, const std::string& name,
                                                             layoutT layout = AoSoA)
    {
        return add_attribute_from<T, VertexHandle>(column(values), name, layout);
    }
    template <class T>
    std::shared_ptr<FaceAttribute<T>> add_face_attribute(const std::vector<std::vector<T>>& values,
                                                         const std::string& name, layoutT layout = AoSoA)
    {
        return add_attribute_from<T, FaceHandle>(values, name, layout);
    }
    template <class T>
    std::shared_ptr<FaceAttribute<T>> add_face_attribute(const std::vector<T>& values, const std::string& name,
                                                         layoutT layout = AoSoA)
    {
        return add_attribute_from<T, FaceHandle>(column(values), name, layout);
    }
    bool does_attribute_exist(const std::string& name) const { return m_releasers.count(name) != 0; }
    void remove_attribute(const std::string& name)
    {
        auto it = m_releasers.find(name);
        if (it == m_releasers.end()) return;
        it->second();
        m_releasers.erase(it);
        m_attr_order.erase(std::remove(m_attr_order.begin(), m_attr_order.end(), name), m_attr_order.end());
    }
    uint32_t get_num_attributes() const { return static_cast<uint32_t>(m_releasers.size()); }

    // ---- output -----------------------------------------------------------
    // write the mesh with the given coordinates (host side) as OBJ, in input order
    void export_obj(const std::string& filename, const VertexAttribute<float>& coords) const
    {
        FILE* f = std::fopen(filename.c_str(), "w");
        if (!f) throw std::runtime_error("export_obj: cannot open " + filename);
        for (uint32_t g = 0; g < get_num_vertices(); ++g) {
            const uint32_t lin = m_g2l[0][g];
            std::fprintf(f, "v %.9g %.9g %.9g\n", coords(lin, 0), coords(lin, 1), coords(lin, 2));
        }
        for (size_t i = 0; i < m_fv.size(); i += 3)
            std::fprintf(f, "f %u %u %u\n", m_fv[i] + 1, m_fv[i + 1] + 1, m_fv[i + 2] + 1);
        std::fclose(f);
    }
    // write the patch layout as a reference-compatible patch archive
    void save(const std::string& filename) const
    {
        check(rxm_mesh_save_patches(m_mesh, filename.c_str()), "save");
    }

   private:
    static constexpr size_t kMaxLDS = 160u * 1024u;

    static void check(int rc, const char* what)
    {
        if (rc != RXM_OK) throw std::runtime_error(std::string("RXMeshStatic::") + what + ": status " + std::to_string(rc));
    }

    template <class T>
    static std::vector<std::vector<T>> column(const std::vector<T>& v)
    {
        std::vector<std::vector<T>> r(v.size());
        for (size_t i = 0; i < v.size(); ++i) r[i] = {v[i]};
        return r;
    }

    template <class T, class HandleT>
    std::shared_ptr<Attribute<T, HandleT>> add_attribute_from(const std::vector<std::vector<T>>& values,
                                                               const std::string& name, layoutT layout)
    {
        constexpr int k = detail::HandleKind<HandleT>::value;
        if (values.size() != get_num_elements<HandleT>())
            throw std::runtime_error("add_attribute: one row per element expected");
        const uint32_t na = values.empty() ? 1u : static_cast<uint32_t>(values[0].size());
        auto           a  = add_attribute<T, HandleT>(name, na, LOCATION_ALL, layout);
        for (size_t g = 0; g < values.size(); ++g)
            for (uint32_t j = 0; j < na; ++j) (*a)(m_g2l[k][g], j) = values[g][j];
        a->move(HOST, DEVICE);
        return a;
    }

    template <typename HandleT>
    HandleT handle_of(uint32_t lin) const
    {
        constexpr int   k  = detail::HandleKind<HandleT>::value;
        const auto&     pr = m_prefix[k];
        const uint32_t  p  = static_cast<uint32_t>(std::upper_bound(pr.begin(), pr.end(), lin) - pr.begin()) - 1u;
        return HandleT(p, static_cast<uint16_t>(lin - pr[p]));
    }

    template <int K, typename LambdaT>
    void for_each_kind(locationT location, LambdaT apply, hipStream_t stream, bool)
    {
        using H = typename detail::KindHandle<K>::type;
        // a lambda runs on the host when it can be called there, and on the
        // device when it is a __device__ lambda (reference IS_D_LAMBDA,
        // util/macros.h:227-228); unannotated lambdas are host lambdas
        constexpr bool on_host = detail::host_callable<LambdaT, H>::value;
        if (location & HOST) {
#if !defined(__HIP_DEVICE_COMPILE__)
            if constexpr (on_host) {
                for (uint32_t p = 0; p < get_num_patches(); ++p)
                    for (uint32_t l = 0; l < m_counts[6 * p + 3 + K]; ++l) apply(H(p, static_cast<uint16_t>(l)));
            } else {
                throw std::runtime_error("for_each: a __device__ lambda cannot run on the HOST");
            }
#endif
        }
        if (location & DEVICE) {
            if constexpr (!on_host) {
                const uint32_t P = get_num_patches();
                if (!P) return;
                if constexpr (K == 0) detail::for_each_vertex<<<P, 256, 0, stream>>>(P, m_ctx, apply);
                else if constexpr (K == 1) detail::for_each_edge<<<P, 256, 0, stream>>>(P, m_ctx, apply);
                else detail::for_each_face<<<P, 256, 0, stream>>>(P, m_ctx, apply);
                detail::hip_check(hipGetLastError(), "RXMeshStatic::for_each");
            } else {
                throw std::runtime_error("for_each: the lambda must be annotated __device__ to run on the DEVICE");
            }
        }
    }

    template <typename T>
    std::vector<T> get_array(int which) const
    {
        uint64_t n = 0;
        check(rxm_mesh_array_size(m_mesh, which, &n), "get_array");
        std::vector<T> v(n);
        if (n) check(rxm_mesh_get_array(m_mesh, which, v.data(), n * sizeof(T)), "get_array");
        return v;
    }

    const uint32_t* upload(const std::vector<uint32_t>& v)
    {
        uint32_t* d = nullptr;
        detail::hip_check(hipMalloc(&d, std::max<size_t>(v.size(), 1) * sizeof(uint32_t)), "RXMeshStatic: hipMalloc");
        m_dev_arrays.push_back(d);
        if (!v.empty())
            detail::hip_check(hipMemcpy(d, v.data(), v.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
                              "RXMeshStatic: hipMemcpy");
        return d;
    }

    void init(std::vector<std::vector<uint32_t>>& fv, uint32_t num_vertices, const std::string& patcher_file,
              uint32_t patch_size)
    {
        m_fv.resize(3 * fv.size());
        for (size_t f = 0; f < fv.size(); ++f) {
            if (fv[f].size() != 3) throw std::runtime_error("RXMeshStatic: only triangle meshes are supported");
            for (int i = 0; i < 3; ++i) m_fv[3 * f + i] = fv[f][i];
        }
        check(rxm_mesh_create(m_fv.data(), static_cast<uint32_t>(fv.size()), num_vertices, patch_size,
                              patcher_file.empty() ? nullptr : patcher_file.c_str(), 0, &m_mesh),
              "rxm_mesh_create");
        check(rxm_mesh_info(m_mesh, &m_info), "rxm_mesh_info");
        check(rxm_mesh_device_view(m_mesh, &m_ctx.m, sizeof(m_ctx.m)), "rxm_mesh_device_view");
        m_ctx.max_edge_faces = m_info.max_edge_incident_faces;

        const uint32_t P = m_info.num_patches;
        m_counts         = get_array<uint32_t>(RXM_ARR_PATCH_COUNTS);
        const auto offs  = get_array<uint32_t>(RXM_ARR_PATCH_OFFSETS);
        const int ltog_ids[3] = {RXM_ARR_LTOG_V, RXM_ARR_LTOG_E, RXM_ARR_LTOG_F};
        const int l2g_ids[3]  = {RXM_ARR_V_L2G, RXM_ARR_E_L2G, RXM_ARR_F_L2G};
        const int g2l_ids[3]  = {RXM_ARR_V_G2L, RXM_ARR_E_G2L, RXM_ARR_F_G2L};
        const uint32_t* d_lin[3] = {m_ctx.m.v_lin, m_ctx.m.e_lin, m_ctx.m.f_lin};
        const uint32_t  nel[3]   = {m_info.num_vertices, m_info.num_edges, m_info.num_faces};
        const uint32_t* d_prefix[3];
        for (int k = 0; k < 3; ++k) {
            auto& pr = m_prefix[k];
            pr.assign(P + 1, 0);
            for (uint32_t p = 0; p < P; ++p) pr[p + 1] = pr[p] + m_counts[6 * p + 3 + k];
            m_slab_off[k].assign(offs.begin() + k * (P + 1), offs.begin() + (k + 1) * (P + 1));
            m_l2g[k]          = get_array<uint32_t>(l2g_ids[k]);
            m_g2l[k]          = get_array<uint32_t>(g2l_ids[k]);
            const auto ltog   = get_array<uint32_t>(ltog_ids[k]);
            m_lin[k].resize(ltog.size());
            // slab padding entries hold INVALID32
            for (size_t i = 0; i < ltog.size(); ++i)
                m_lin[k][i] = (ltog[i] < nel[k]) ? m_g2l[k][ltog[i]] : INVALID32;
            d_prefix[k] = upload(pr);
            detail::AttrIndex& ix = m_index[k];
            ix.n                  = nel[k];
            ix.d_prefix           = d_prefix[k];
            ix.d_slab_off         = upload(m_slab_off[k]);
            ix.d_lin              = d_lin[k];
            ix.h_prefix           = m_prefix[k].data();
            ix.h_slab_off         = m_slab_off[k].data();
            ix.h_lin              = m_lin[k].data();
        }
        m_ctx.v_prefix = d_prefix[0];
        m_ctx.e_prefix = d_prefix[1];
        m_ctx.f_prefix = d_prefix[2];
    }

    rxm_mesh*                                    m_mesh = nullptr;
    rxm_mesh_info_t                              m_info{};
    Context                                      m_ctx{};
    std::vector<uint32_t>                        m_fv, m_counts;
    std::vector<uint32_t>                        m_prefix[3], m_slab_off[3], m_lin[3], m_l2g[3], m_g2l[3];
    detail::AttrIndex                            m_index[3];
    std::vector<uint32_t*>                       m_dev_arrays;
    std::map<std::string, std::function<void()>> m_releasers;
    std::vector<std::string>                     m_attr_order;
    std::shared_ptr<VertexAttribute<float>>      m_coords;
};

}  // namespace rxmesh
