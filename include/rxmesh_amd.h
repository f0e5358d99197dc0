/*
 * rxmesh_amd: C-ABI of the MI355X (gfx950) patch-based static mesh query
 * engine. Plain pointers and sizes only; device pointers are caller-owned
 * unless stated; every function returns an rxm_status and never exits.
 *
 * Layout conventions
 *   - Element ids on input are the caller's ("global") ids: vertex = index in
 *     the vertex array, face = index in fv, edge = first-appearance order over
 *     half-edges (f, i) -> (fv[3f+i], fv[3f+(i+1)%3]).
 *   - Device attributes and query results use LINEAR ids: the owned elements
 *     of patch 0, then patch 1, ... (the reference's linear_id,
 *     include/rxmesh/context.h:299-331). rxm_mesh_get_array() exports the
 *     linear<->global maps; rxm_to_linear / rxm_to_global permute on device.
 *   - Multi-component vertex attributes are SoA over linear ids: component c
 *     of linear vertex l is at data[c * num_vertices + l] (the reference's SoA
 *     layout, include/rxmesh/attribute.h:406-472).
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *
 * Thread compatibility: a mesh may be used from one host thread at a time.
 * Streams: a mesh owns scratch buffers that several entry points share (the
 * energy partial sums of rxm_arap_energy_grad and rxm_symdir_energy, the CG
 * state of rxm_mcf_implicit_cg, the work-queue counter of the batched steps).
 * Nothing orders their use across streams, so issue all calls on one mesh on
 * one stream (a captured graph of a mesh's calls included): calls on one
 * mesh from two streams that overlap in time may overwrite each other's
 * partials.
 */
#ifndef RXMESH_AMD_H
#define RXMESH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int rxm_status;
enum {
    RXM_OK = 0,
    RXM_ERR_INVALID_ARG = 1,
    RXM_ERR_NON_MANIFOLD = 2, /* also: oriented paths on an inconsistently oriented mesh */
    RXM_ERR_PATCH_TOO_BIG = 3,
    RXM_ERR_IO = 4,
    RXM_ERR_DEVICE = 5,
    RXM_ERR_UNSUPPORTED = 6,
    RXM_ERR_OUT_OF_MEMORY = 7,
    RXM_ERR_SCRATCH_TOO_SMALL = 8 /* input beyond the fixed scratch of an allocation-free entry point */
};

/* Query operations; values follow the reference enum Op
 * (include/rxmesh/types.h:113-129). */
enum {
    RXM_OP_VV = 3, RXM_OP_VE = 4, RXM_OP_VF = 5, RXM_OP_FV = 6,
    RXM_OP_FE = 7, RXM_OP_FF = 8, RXM_OP_EV = 9, RXM_OP_EF = 11,
    RXM_OP_EVD = 12 /* EVDiamond (rxmesh_queries.cuh:198-265): 4 per edge [v0, opposite of v0->v1, v1,
                       opposite of v1->v0], -1 if absent */
};

/* rxm_mesh_create flags. The default patcher is a recursive proportional
 * bisection (about ceil(F / patch_size) connected patches, the fewest the cap
 * allows, with the smallest ribbons); RXM_PATCHER_KMEANS selects the k-means
 * patcher instead (patches ~0.7 of the cap). RXM_BUILD_SERIAL runs the host
 * side of the build on the calling thread only (for callers that build many
 * meshes from their own threads at once; the result is the same). */
enum { RXM_HOST_ONLY = 1, RXM_PATCHER_KMEANS = 2, RXM_BUILD_SERIAL = 4 };

/* vertex normal weighting */
enum { RXM_NORMALS_MAX = 0, RXM_NORMALS_AREA = 1 };

typedef struct rxm_mesh rxm_mesh;
typedef struct rxm_graph rxm_graph;

typedef struct rxm_mesh_info_t {
    uint32_t num_vertices, num_edges, num_faces, num_patches;
    uint32_t patch_size;
    uint32_t max_vertices_per_patch, max_edges_per_patch, max_faces_per_patch; /* incl. ribbon */
    uint32_t max_owned_vertices, max_owned_edges, max_owned_faces;
    uint32_t is_closed, is_edge_manifold;
    uint32_t max_valence, max_edge_incident_faces, max_face_adjacent_faces;
    uint32_t num_ribbon_faces;
    uint32_t on_device;
    uint64_t device_bytes;
    uint32_t is_consistently_oriented; /* each edge's faces traverse it in opposite directions */
} rxm_mesh_info_t;

/* Arrays exportable with rxm_mesh_get_array (host copies, uint32 unless noted). */
enum {
    RXM_ARR_EV = 0,            /* 2E global vertex ids, first-seen orientation */
    RXM_ARR_FACE_PATCH = 1,    /* F */
    RXM_ARR_VERTEX_PATCH = 2,  /* V owner patch */
    RXM_ARR_EDGE_PATCH = 3,    /* E owner patch */
    RXM_ARR_PATCHES_VAL = 4,   /* F faces grouped by patch */
    RXM_ARR_PATCHES_OFFSET = 5,/* P inclusive ends */
    RXM_ARR_RIBBON_VAL = 6,    /* ribbon faces per patch, discovery order */
    RXM_ARR_RIBBON_OFFSET = 7, /* P inclusive ends */
    RXM_ARR_V_L2G = 8, RXM_ARR_E_L2G = 9, RXM_ARR_F_L2G = 10,   /* linear -> global */
    RXM_ARR_V_G2L = 11, RXM_ARR_E_G2L = 12, RXM_ARR_F_G2L = 13, /* global -> linear */
    RXM_ARR_PATCH_COUNTS = 14, /* 6P: nv, ne, nf, nvo, neo, nfo per patch */
    RXM_ARR_PATCH_OFFSETS = 15,/* 3(P+1): v_off, e_off, f_off slab offsets */
    RXM_ARR_LTOG_V = 16, RXM_ARR_LTOG_E = 17, RXM_ARR_LTOG_F = 18, /* slabs */
    RXM_ARR_EV_LOCAL = 19,     /* e-slab u32: lo16 local(max gid), hi16 local(min gid) */
    RXM_ARR_FE_LOCAL = 20,     /* 3 * f-slab uint16: (local edge << 1) | dir */
    /* two-step smoothing plan (built on demand): per patch p, a row for every
     * local vertex in its owner's VV order, neighbours in p's extended local
     * numbering ([0, nv) local vertices, then the second-ring extras) */
    RXM_ARR_VV2_ROFF = 21,     /* uint16 row offsets, run of patch p at rbase[p] */
    RXM_ARR_VV2_NBR = 22,      /* uint16 neighbours, run of patch p at nbase[p] */
    RXM_ARR_VV2_XLIN = 23,     /* linear ids of the extras, run of patch p at xbase[p] */
    RXM_ARR_VV2_OFFSETS = 24   /* rbase (P+1), nbase (P+1), xbase (P+1), extras per patch (P) */
};

/* Build the patch mesh from F x 3 triangle indices and upload it.
 * Replaces RXMeshStatic(fv, patcher_file, patch_size, ...)
 * (include/rxmesh/rxmesh_static.h:61-100) -> RXMesh::init / build
 * (include/rxmesh/rxmesh.cpp:89-225, 290-448).
 * num_vertices = 0 infers max index + 1; every vertex must belong to a face
 * (RXM_ERR_INVALID_ARG otherwise, as for an empty or degenerate face list).
 * patch_file (may be NULL) is a
 * reference patch archive (cereal PortableBinary, patcher.h:154-182) whose
 * layout is replayed instead of running the patcher. */
rxm_status rxm_mesh_create(const uint32_t* fv, uint32_t num_faces, uint32_t num_vertices,
                           uint32_t patch_size, const char* patch_file, uint32_t flags,
                           rxm_mesh** out);
/* Same, with an explicit face -> patch assignment (faces grouped by patch in
 * ascending face order). */
rxm_status rxm_mesh_create_with_patches(const uint32_t* fv, uint32_t num_faces,
                                        uint32_t num_vertices, const uint32_t* face_patch,
                                        uint32_t num_patches, uint32_t flags, rxm_mesh** out);
/* Build `count` meshes and upload them as ONE patch mesh: their disjoint
 * union, whose patches are the concatenation of every mesh's patches, so any
 * entry point below runs on all meshes in a single launch. Mesh k owns the
 * contiguous id ranges reported by rxm_batch_offsets. No reference
 * counterpart (the reference builds one RXMeshStatic per mesh); meant for
 * batched workloads of many uneven meshes. num_vertices may be NULL. */
rxm_status rxm_mesh_create_batch(const uint32_t* const* fvs, const uint32_t* num_faces,
                                 const uint32_t* num_vertices, uint32_t count, uint32_t patch_size,
                                 uint32_t flags, rxm_mesh** out);
/* Same union built from meshes already created (their host data is
 * reused; the parts stay valid and independent). */
rxm_status rxm_mesh_batch_from(const rxm_mesh* const* meshes, uint32_t count, uint32_t flags,
                               rxm_mesh** out);
/* Number of meshes in a batch (1 for a mesh made by rxm_mesh_create). */
rxm_status rxm_batch_count(const rxm_mesh* mesh, uint32_t* count);
/* count + 1 offsets of mesh k's ids: kind 0 vertices, 1 edges, 2 faces
 * (global = linear ranges), 3 patches. */
rxm_status rxm_batch_offsets(const rxm_mesh* mesh, int kind, uint32_t* out, uint32_t count);
/* Frees host and device state (reference ~RXMesh, rxmesh.cpp:227-288). */
rxm_status rxm_mesh_destroy(rxm_mesh* mesh);
rxm_status rxm_mesh_info(const rxm_mesh* mesh, rxm_mesh_info_t* info);
/* Size in elements of an exportable array. */
rxm_status rxm_mesh_array_size(const rxm_mesh* mesh, int which, uint64_t* count);
/* Copy an exportable array into host memory of `bytes` bytes. */
rxm_status rxm_mesh_get_array(const rxm_mesh* mesh, int which, void* dst, uint64_t bytes);
/* Copy the device view of the mesh (struct rxm::DevMesh of
 * rxmesh_amd/csrc/device_mesh.hpp: raw device pointers + counts, the role of
 * the reference's Context, include/rxmesh/context.h:15-441) into `out`; used
 * by the header-only C++ API to build kernel arguments. */
rxm_status rxm_mesh_device_view(const rxm_mesh* mesh, void* out, uint64_t bytes);
/* Write the patch layout as a reference-compatible patch archive
 * (Patcher::save, include/rxmesh/patcher/patcher.h:154-159). */
rxm_status rxm_mesh_save_patches(const rxm_mesh* mesh, const char* path);

/* Number of entries of a query over all owned source elements. */
rxm_status rxm_query_nnz(rxm_mesh* mesh, int op, int64_t* nnz);
/* Materialise a query as CSR over linear ids: row_ptr[n_src + 1], col[nnz].
 * Replaces for_each<Op, N>(lambda, oriented) / Query::dispatch<Op>
 * (include/rxmesh/rxmesh_static.h:524-566, include/rxmesh/query.inl:108-156)
 * driven by the reference test lambda output(id, i) = iter[i]
 * (tests/RXMesh_test/query_kernel.cuh:13-45). Row order within a row:
 * ascending source-local order (unoriented), or the canonical CCW one-ring
 * (oriented, VV/VE only, edge-manifold and consistently oriented meshes). */
rxm_status rxm_query_csr(rxm_mesh* mesh, int op, int oriented, int32_t* row_ptr,
                         int32_t* col, void* stream);

/* VV and VF (unoriented) in one launch: the same two CSRs as two
 * rxm_query_csr calls (row_ptr V + 1, col sized by rxm_query_nnz), with the
 * patch topology staged once for both. Replaces a v_v and a v_f query of
 * the reference (include/rxmesh/kernels/rxmesh_queries.cuh:502-858) issued
 * back to back; no single reference call does both. */
rxm_status rxm_query_vv_vf(rxm_mesh* mesh, int32_t* vv_row_ptr, int32_t* vv_col, int32_t* vf_row_ptr,
                           int32_t* vf_col, void* stream);

/* Vertex normals, unnormalised. mode RXM_NORMALS_MAX restates the reference
 * VertexNormal app (apps/VertexNormal/vertex_normal_kernel.cuh:10-42). */
rxm_status rxm_vertex_normals(rxm_mesh* mesh, int mode, const float* pos, float* normals,
                              void* stream);

/* Gaussian curvature per vertex: gc = (2 pi - sum of the vertex's corner
 * angles) / mixed area, and optionally the mixed (Voronoi / obtuse) area
 * (`area_mixed` may be NULL). Replaces compute_gaussian_curvature
 * (apps/GaussianCurvature/gaussian_curvature_kernel.cuh:10-70) followed by
 * the for_each_vertex division (apps/GaussianCurvature/gaussian_curvature.cu
 * :55-58); the serial reference is gaussian_curvature_ref.h:8-118. Each owned
 * vertex sums its incident faces in a fixed order (no atomics). */
rxm_status rxm_gaussian_curvature(rxm_mesh* mesh, const float* pos, float* gc, float* area_mixed,
                                  void* stream);

/* ARAP-style deformation terms, replacing the pieces of the reference's
 * DiffARAP app (apps/DiffARAP). Positions are 3 x V SoA in linear order.
 *
 * Edge weights per linear edge id: w = max(0, (cot_q + cot_s) / 2) over the
 * EVDiamond opposite vertices q, s (calc_edge_weights_mat,
 * apps/DiffARAP/arap_kernels.h:8-54, which stores them in a symmetric
 * SparseMatrix); 0 for edges with one face. */
rxm_status rxm_arap_weights(rxm_mesh* mesh, const float* pos, float* w, void* stream);
/* Best rotation per vertex (calculate_rotation_matrix, arap_kernels.h:60-117):
 * the rotation maximising tr(R S), S = sum over the vertex's neighbours j of
 * w (P_v - P_j)(X_v - X_j)^T, i.e. V U^T from S = U Sigma V^T with the
 * reflection fix. R holds one row-major record per vertex (linear order):
 * R[9 v + 3 i + j] = R_v(i, j), the reference's rotations(v, 3 i + j). */
rxm_status rxm_arap_rotations(rxm_mesh* mesh, const float* rest, const float* x, const float* w, float* R,
                              void* stream);
/* Energy sum over edges (i, j) in EV order (i the larger global id) of
 * w |(X_i - X_j) - R_i (P_i - P_j)|^2 (the term of diff_arap.cu:87-118) into
 * *total (device float), and its gradient with respect to X with the
 * rotations held fixed, as the reference's term reads them, into grad
 * (3 x V). The gradient is gathered per vertex without atomics. */
rxm_status rxm_arap_energy_grad(rxm_mesh* mesh, const float* rest, const float* x, const float* w,
                                const float* R, float* total, float* grad, void* stream);

/* Connectivity build on the device, the edge step of the mesh build
 * (build_supporting_structures, include/rxmesh/rxmesh.cpp:518-681) as
 * sort-based deduplication: edge ids in order of first appearance over the
 * half-edges (f, i) -> (fv[3f+i], fv[3f+(i+1)%3]), EV in the first
 * half-edge's orientation, and every half-edge's edge id -- the same ids as
 * the host build. fv (3F), ev (capacity 2 x 3F) and fe (3F) are device
 * pointers; the call waits for the stream and sets *num_edges. Allocates its
 * work buffers. RXM_ERR_INVALID_ARG for a vertex id >= num_vertices or a
 * degenerate face. */
rxm_status rxm_build_edges_device(const uint32_t* fv, uint32_t num_faces, uint32_t num_vertices, uint32_t* ev,
                                  uint32_t* fe, uint32_t* num_edges, void* stream);

/* Build (once; later calls return at once) the mesh's patch-local VV plan:
 * the VV query evaluated by the query engine and kept in its native
 * patch-local u16 encoding for iterative consumers. rxm_laplacian_smooth
 * builds it on first use; call this to keep the build out of a timed loop. */
rxm_status rxm_vv_plan_build(rxm_mesh* mesh, void* stream);
/* Launch configuration knobs (used by the autotuner):
 *   kernel 0: smoothing workgroup size 64 / 128 / 256 / 512;
 *   kernel 1: symmetric-Dirichlet step workgroup size 128 / 256;
 *   kernel 2: symmetric-Dirichlet step launch, 0 = one workgroup per patch,
 *             1 = persistent grid over a global patch work queue;
 *   kernel 3: smoothing LDS layout of positions, 0 = three float arrays,
 *             1 = one float4 per vertex;
 *   kernel 4: smoothing, 1 = two Jacobi steps per launch over the two-step
 *             plan (RXM_ARR_VV2_*), 0 = one step per launch;
 *   kernel 5: SpMV-3 / explicit MCF lanes per CSR row, 1 / 2 / 4 / 8;
 *   kernel 6: oriented VV / VE workgroup size 256 / 512;
 *   kernel 7: cotangent assembly workgroup size 256 / 512;
 *   kernel 8: smoothing, 1-6 = that many Jacobi steps per launch over the
 *             K-level plan (patch + rings of neighbours kept in LDS; overrides
 *             kernels 3 and 4), 0 = off;
 *   kernel 9: smoothing with two steps per launch (kernel 4 = 1, float
 *             arrays in LDS) and >= 4 iterations: 1 = the iterates between
 *             launches are one float4 per vertex in two buffers the mesh
 *             allocates on first use (32 V bytes; a gather then reads one
 *             cache line instead of three), 0 = SoA between launches;
 *   kernel 10: unoriented VV / VE / VF / EF / FF queries (rxm_query_csr,
 *             rxm_query_vv_vf): 0 = LDS-atomic histogram and slot fill
 *             followed by a sort of every row; 1 = atomic-free stable
 *             counting sort (wave ballots + mbcnt ranks + one block scan;
 *             rows come out in order, no row sort); 2 (default) = one pass
 *             per launch for VV / VE / VF / EF / FF and for rxm_query_vv_vf
 *             (both relations in one counting sort): keys per edge or per
 *             face held in registers, rows sorted in registers, one
 *             coalesced write.
 * Settings of kernels 0-4 and 6-10 give bit-identical results; kernel 5 changes
 * the summation order of a row (fp32 results within rounding). */
rxm_status rxm_set_launch_threads(rxm_mesh* mesh, int kernel, int threads);

/* `iters` Jacobi steps of uniform Laplacian smoothing
 * x <- x + lambda (mean of VV neighbours - x); `out` receives the result,
 * `tmp` is a caller workspace of 3V floats (may be NULL when iters == 1).
 * Neighbours come from the patch-local VV plan (rxm_vv_plan_build).
 * No single reference function; nearest are apps/Smoothing/manual.h:85-103
 * and the uniform branch of apps/MCF/mcf_kernels.cuh:114-213. */
rxm_status rxm_laplacian_smooth(rxm_mesh* mesh, const float* pos, float* out, float* tmp,
                                float lambda, int iters, void* stream);

/* Same smoothing, consuming a materialised VV query (rxm_query_csr with
 * RXM_OP_VV, unoriented) instead of re-deriving the one-rings per launch:
 * for iterative use on a static mesh, query once and iterate on the CSR. */
rxm_status rxm_laplacian_smooth_csr(rxm_mesh* mesh, const int32_t* vv_row_ptr, const int32_t* vv_col,
                                    const float* pos, float* out, float* tmp, float lambda,
                                    int iters, void* stream);

/* Entries of the cotangent system (V + 2E on closed manifolds). */
rxm_status rxm_cot_nnz(rxm_mesh* mesh, int64_t* nnz);
/* Assemble A = M + dt L over the oriented VV ring as CSR over linear vertex
 * ids (row = [diagonal, CCW ring neighbours]) and the lumped mass M.
 * Replaces SparseMatrix<T>(rx, Op::VV) (include/rxmesh/matrix/sparse_matrix.h:
 * 133-242) filled by mcf_A_setup (apps/MCF/mcf_kernels.cuh:354-426); weights
 * restate edge_cotan_weight / partial_voronoi_area
 * (include/rxmesh/geometry_util.cuh:107-206). Needs an edge-manifold, consistently
 * oriented mesh. */
rxm_status rxm_cot_csr_assemble(rxm_mesh* mesh, const float* pos, float dt, int32_t* row_ptr,
                                int32_t* col, float* val, float* mass, void* stream);
/* y = A x for 3 SoA columns over the V vertex rows. Replaces
 * SparseMatrix::multiply (sparse_matrix.h:1317-1351, cuSPARSE SpMV). */
rxm_status rxm_spmv3(rxm_mesh* mesh, const int32_t* row_ptr, const int32_t* col, const float* val,
                     const float* x, float* y, void* stream);
/* `iters` explicit mean-curvature-flow steps x <- x + (M x - A x) / M with the
 * SpMV fused into the update; `tmp` is a 3V workspace when iters > 1. A must
 * be the mesh's cotangent system as rxm_cot_csr_assemble lays it out (values
 * may be any): the steps run patch-local over the mesh's VV plan, which is
 * that system's pattern, so `col` is not read. */
rxm_status rxm_mcf_explicit(rxm_mesh* mesh, const int32_t* row_ptr, const int32_t* col,
                            const float* val, const float* mass, const float* x, float* out,
                            float* tmp, int iters, void* stream);

/* Symmetric Dirichlet parameterisation energy (apps/Param/param.cu:113-184).
 * rxm_symdir_rest: per-face 2x2 rest shapes Mr (SoA over linear faces, 4F
 * floats: m00, m10, m01, m11) from 3D rest positions (param.cu:113-141). */
rxm_status rxm_symdir_rest(rxm_mesh* mesh, const float* rest_pos, float* Mr, void* stream);
/* Per-face energies (FLT_MAX for flipped faces) into energy[F] and their sum
 * into total[1] (fixed-order, deterministic). Replaces the passive
 * evaluation diff_scalar_kernel_passive (diff/diff_query_kernel.cuh:156-196)
 * plus ReduceHandle (reduce_handle.h:23-210). */
rxm_status rxm_symdir_energy(rxm_mesh* mesh, const float* Mr, const float* uv, float* energy,
                             float* total, void* stream);
/* Same results as rxm_symdir_energy (same kernels), allocating nothing: no
 * hipMalloc, hipFree or allocating library call on any call, the first
 * included. No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops). The per-patch
 * partial sums live in a fixed library-owned device array of 2^18 floats
 * (1 MiB per device, held until the process ends), so meshes with more than
 * 262,144 patches are refused with RXM_ERR_SCRATCH_TOO_SMALL. Every call on a
 * device shares that device's array: calls that overlap in time on one device
 * (two streams, two threads) race and may return wrong totals. The current
 * device must be the mesh's (else RXM_ERR_INVALID_ARG). */
rxm_status rxm_symdir_energy_noalloc(rxm_mesh* mesh, const float* Mr, const float* uv, float* energy,
                                     float* total, void* stream);
/* grad (2 x V SoA) = scale[0] * d(sum_f E_f)/d uv (scale may be NULL = 1):
 * hand-written adjoint recomputing each face's Jacobian, gathered per vertex
 * over VF without atomics. Replaces the forward-mode AD gradient of
 * DiffScalarProblem::eval_terms_grad_only (diff/diff_scalar_problem.h:327-335,
 * diff/diff_query_kernel.cuh:199-330). */
rxm_status rxm_symdir_grad(rxm_mesh* mesh, const float* Mr, const float* uv, const float* scale,
                           float* grad, void* stream);

/* Implicit mean-curvature-flow step: solves A X = M x with conjugate
 * gradients, restating the reference's MCF CG path (apps/MCF/mcf_cg.h:13-125,
 * CGSolver include/rxmesh/matrix/cg_solver.h:37-125, B from mcf_B_setup
 * apps/MCF/mcf_kernels.cuh:302-352): X starts at x, the three coordinate
 * columns share alpha and beta (dots and norms over all 3V entries), and the
 * solve stops when |r|^2 < abs_tol or |r|^2 / |r0|^2 < rel_tol, or after
 * max_iter iterations. A is the mesh's cotangent system (rxm_cot_csr_assemble
 * layout; col is not read). The loop runs on the device without host
 * synchronisation; when any of iters / start_residual / final_residual is
 * non-NULL the call waits for the stream and reports the iterations taken and
 * the squared residual norms, as the reference's solver does. Work vectors
 * (4 x 3V floats) are allocated by the mesh on first use. */
rxm_status rxm_mcf_implicit_cg(rxm_mesh* mesh, const int32_t* row_ptr, const int32_t* col,
                               const float* val, const float* mass, const float* x, float* out,
                               int max_iter, float abs_tol, float rel_tol, int* iters,
                               float* start_residual, float* final_residual, void* stream);

/* Implicit MCF step with the system never assembled: CG (precond = 0) or
 * Jacobi-preconditioned CG (precond = 1) on A X = B, where every product
 * recomputes the weights from the fixed input coordinates `coords` (3 x V
 * SoA): A = M + dt L with cotangent weights max(0, cot + cot) and Voronoi
 * mass (uniform = 0), or unit weights and mass = valence (uniform = 1); B =
 * M x (the reference's init_B). x is the initial X (3 x V SoA; the reference
 * starts from the input coordinates), the solution goes to `out`. Stopping
 * rules, residuals (|r|^2, or <r, r / diag> with the preconditioner) and
 * iteration counts as rxm_mcf_implicit_cg. Cotangent weights need an
 * edge-manifold, consistently oriented mesh whose every vertex ring is one
 * closed fan (else RXM_ERR_NON_MANIFOLD / RXM_ERR_UNSUPPORTED; the reference
 * app takes closed meshes only). Replaces mcf_cg_mat_free / mcf_pcg_mat_free
 * (apps/MCF/mcf_cg_mat_free.h:148-240) with the matvec / precond_matvec
 * kernels (apps/MCF/mcf_kernels.cuh:114-299) under CGMatFreeAttrSolver /
 * PCGMatFreeAttrSolver (include/rxmesh/matrix/cg_mat_free_attr_solver.h,
 * pcg_mat_free_attr_solver.h). */
rxm_status rxm_mcf_matfree_cg(rxm_mesh* mesh, const float* coords, const float* x, float* out,
                              float dt, int uniform, int precond, int max_iter, float abs_tol,
                              float rel_tol, int* iters, float* start_residual,
                              float* final_residual, void* stream);
/* Symmetric Dirichlet with 3-D positions as the variable (closed surfaces
 * have no UV domain): J = [b-a, c-a] Mr^{-1} is 3x2 and |J^{-1}|^2 becomes
 * |J|^2 / det(J^T J); equals the 2-D energy for planar input. Mr from
 * rxm_symdir_rest on the rest positions. Per-face energies into energy[F]. */
rxm_status rxm_symdir3_energy(rxm_mesh* mesh, const float* Mr, const float* x, float* energy,
                              void* stream);
/* x_out = x - eta * grad (sum_f E_f)(x), adjoint gathered per vertex over VF
 * without atomics (x, x_out: 3 x V SoA, distinct). */
rxm_status rxm_symdir3_grad_step(rxm_mesh* mesh, const float* Mr, const float* x, float eta,
                                 float* x_out, void* stream);
/* grad (3 x V SoA) = d(sum_f E_f)/dx of the 3-D energy above, the same
 * adjoint as rxm_symdir3_grad_step (per-face Jacobians recomputed, gathered
 * per vertex over VF without atomics) written out instead of applied.
 * Replaces the gradient DiffScalarProblem::eval_terms_grad_only returns
 * (include/rxmesh/diff/diff_scalar_problem.h:327-335). x and grad distinct. */
rxm_status rxm_symdir3_grad(rxm_mesh* mesh, const float* Mr, const float* x, float* grad, void* stream);
/* `steps` iterations of { x1 = one uniform smoothing step of x (lambda);
 * x = x1 - eta grad E(x1) } in place on x, `tmp` a 3V workspace. Two launches
 * per step over all patches (all meshes of a batch). */
rxm_status rxm_smooth_symdir_steps(rxm_mesh* mesh, const float* Mr, float* x, float* tmp,
                                   float lambda, float eta, int steps, void* stream);
/* The same loop as plain per-mesh launches over `count` separate meshes (the
 * baseline the batched form is measured against). */
rxm_status rxm_smooth_symdir_steps_each(rxm_mesh* const* meshes, const float* const* Mr,
                                        float* const* x, float* const* tmp, uint32_t count,
                                        float lambda, float eta, int steps, void* stream);
/* The same step loop with the VV and VF queries at the start of every step:
 * each step launches rxm_query_vv_vf into the caller's CSR buffers (sized as
 * for rxm_query_vv_vf), then the smoothing and gradient steps -- the
 * reference's per-step query dispatch (Query<N>::dispatch, query.h:21-75,
 * inside each kernel of a step) made explicit. Allocates nothing inside the
 * loop (plans and query bases are prepared first), so it can be captured. */
rxm_status rxm_query_smooth_symdir_steps(rxm_mesh* mesh, const float* Mr, float* x, float* tmp,
                                         float lambda, float eta, int steps, int32_t* vv_row_ptr,
                                         int32_t* vv_col, int32_t* vf_row_ptr, int32_t* vf_col,
                                         void* stream);
/* Plain per-mesh launches of the loop above over `count` separate meshes,
 * one set of CSR buffers per mesh. */
rxm_status rxm_query_smooth_symdir_steps_each(rxm_mesh* const* meshes, const float* const* Mr,
                                              float* const* x, float* const* tmp,
                                              int32_t* const* vv_row_ptr, int32_t* const* vv_col,
                                              int32_t* const* vf_row_ptr, int32_t* const* vf_col,
                                              uint32_t count, float lambda, float eta, int steps,
                                              void* stream);
/* Capture rxm_smooth_symdir_steps into a hipGraph (no memset nodes; the VV
 * plan and the work-queue counter are prepared before capture) and replay
 * it with rxm_graph_launch. Pointers are baked into the graph. The
 * reference's only graph use is a CUDA graph with a conditional node in its
 * Geodesic app (apps/Geodesic/geodesic_ptp_rxmesh_graph.h:100-175). */
rxm_status rxm_graph_capture_smooth_symdir(rxm_mesh* mesh, const float* Mr, float* x, float* tmp,
                                           float lambda, float eta, int steps, rxm_graph** out);
rxm_status rxm_graph_launch(rxm_graph* graph, void* stream);
rxm_status rxm_graph_destroy(rxm_graph* graph);

/* Permute a vertex/edge/face attribute (kind 0/1/2) of ncomp SoA components
 * between global and linear order on device (reference: attribute copy by
 * ltog in RXMeshStatic::add_vertex_attribute, rxmesh_static.inl:147-189, and
 * map_to_global). */
rxm_status rxm_to_linear(rxm_mesh* mesh, int kind, int ncomp, const float* global_data,
                         float* linear_data, void* stream);
rxm_status rxm_to_global(rxm_mesh* mesh, int kind, int ncomp, const float* linear_data,
                         float* global_data, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* RXMESH_AMD_H */
