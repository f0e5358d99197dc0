// C-ABI implementation: mesh lifetime, device upload, kernel launches.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <atomic>
#include <vector>

#include "../../include/rxmesh_amd.h"
#include "host_build.h"
#include "k_geometry.hpp"
#include "k_arap.hpp"
#include "k_numerics.hpp"
#include "k_query.hpp"
#include "k_query2.hpp"

using namespace rxm;

struct rxm_mesh
{
    HostMesh              h;
    bool                  on_device = false;
    int                   device    = -1;
    DevMesh               dm{};
    std::vector<void*>    allocs;
    uint64_t              device_bytes = 0;
    std::map<int, uint32_t*>            d_base;  // per-op patch entry bases (P+1)
    std::map<int, std::vector<uint32_t>> h_base;
    std::map<int, uint32_t>              h_maxrow;  // largest per-patch entry count of an op
    float*                              d_partial = nullptr;  // P per-patch partial sums
    // patch-local VV plan (u16 row offsets + u16 neighbour local ids)
    const uint32_t* d_vv_rbase = nullptr;
    const uint32_t* d_vv_nbase = nullptr;
    uint16_t*       d_vv_roff  = nullptr;
    uint16_t*       d_vv_nbr   = nullptr;
    uint32_t        vv_max_nent = 0;
    // two-step smoothing plan (host copy built on demand, device copy on first use)
    mutable Vv2Plan   vv2;
    mutable bool      vv2_built = false;
    const uint32_t*   d_vv2_rbase = nullptr;
    const uint32_t*   d_vv2_nbase = nullptr;
    const uint32_t*   d_vv2_xbase = nullptr;
    const uint32_t*   d_vv2_xcnt  = nullptr;
    const uint16_t*   d_vv2_roff  = nullptr;
    const uint16_t*   d_vv2_nbr   = nullptr;
    const uint32_t*   d_vv2_xlin  = nullptr;
    // K-level smoothing plan (temporal blocking), built on first use
    VvkPlan           vvk;
    const uint32_t*   d_vvk_lvl   = nullptr;
    const uint32_t*   d_vvk_rbase = nullptr;
    const uint32_t*   d_vvk_nbase = nullptr;
    const uint32_t*   d_vvk_xbase = nullptr;
    const uint16_t*   d_vvk_roff  = nullptr;
    const uint16_t*   d_vvk_nbr   = nullptr;
    const uint32_t*   d_vvk_xlin  = nullptr;
    uint32_t          smooth_steps = 0;  // > 0: Jacobi steps per launch over the K-level plan
    uint32_t          smooth_g4    = 1;  // two-step launches pass float4 iterates between them
    float*            d_cg         = nullptr;  // CG work vectors: r, p (two), s (3 V floats each), partials
    CgState*          d_cg_state   = nullptr;
    float*            d_mf         = nullptr;  // matrix-free CG: r, p (two), s (3 V each), mass, diag (V), partials
    float*            d_iter4      = nullptr;  // two float4 iterates (8 V floats), allocated on first use
    uint32_t          smooth_fused = 0;  // 1: two Jacobi steps per launch
    uint32_t          spmv_lanes   = 8;  // lanes per CSR row in the SpMV-3 kernels
    // patch-local face plan (FV table + VF rows of (face << 2 | corner))
    const uint32_t* d_fp_fbase = nullptr;
    const uint32_t* d_fp_rbase = nullptr;
    const uint32_t* d_fp_nbase = nullptr;
    uint16_t*       d_fp_fvl   = nullptr;
    uint16_t*       d_fp_roff  = nullptr;
    uint16_t*       d_fp_vf    = nullptr;
    uint32_t        fp_max_nent = 0;
    // cotangent assembly: closed-fan flags per linear vertex (bitmask) and the
    // patches owning a vertex whose ring is not one closed fan
    const uint32_t* d_cot_closed = nullptr;
    const uint32_t* d_cot_plist  = nullptr;
    uint32_t        cot_nplist   = 0;
    const uint16_t* d_cot_ent    = nullptr;  // patch-local VV edge pattern (build_vv_edge_pattern)
    const uint32_t* d_cot_rs     = nullptr;
    uint32_t        smooth_threads = 128;
    uint32_t        ring_threads   = 256;      // oriented VV, VE
    uint32_t        cot_threads    = 512;      // cotangent assembly
    uint32_t        smooth_aos4    = 0;        // 1: smoothing stages positions as float4
    uint32_t        grad_threads   = 256;
    uint32_t        grad_queue     = 0;        // 1: persistent work-queue launch
    uint32_t*       d_queue        = nullptr;  // work-queue ticket counter
    bool            system_runs_ok = false;    // check_system_runs passed
    uint32_t        query_stable   = 2;        // transposes: 0 LDS atomics + row sort, 1 atomic-free stable sort, 2 as 0 with the fused VV + VF launch in one pass
    uint32_t        vvf_nslots     = 0;        // largest VV + VF entry count of a patch (0: not computed yet)
    uint32_t        max_nrv = 0, max_nre = 0, max_nrf = 0;  // largest ribbon vertex / edge / face count of a patch
    int             num_cus        = 0;
    // batch (disjoint union of meshes): per-mesh id / patch offsets
    std::vector<uint32_t> b_v, b_e, b_f, b_p;
};

struct rxm_graph
{
    hipGraph_t     graph = nullptr;
    hipGraphExec_t exec  = nullptr;
};

namespace {

constexpr int kOpCot = 100;  // rows of the cotangent system: VV + diagonal

#define RXM_HIP(call)                                 \
    do {                                              \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return RXM_ERR_DEVICE;  \
    } while (0)

template <typename T>
int upload(rxm_mesh* m, const std::vector<T>& v, const T*& dst)
{
    T*     p     = nullptr;
    size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
    if (hipMalloc(&p, bytes) != hipSuccess) return RXM_ERR_OUT_OF_MEMORY;
    m->allocs.push_back(p);
    m->device_bytes += bytes;
    if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return RXM_ERR_DEVICE;
    dst = p;
    return RXM_OK;
}

int upload_mesh(rxm_mesh* m)
{
    const HostMesh& h = m->h;
    RXM_HIP(hipGetDevice(&m->device));
    std::vector<PatchDesc> pd(h.P);
    for (uint32_t p = 0; p < h.P; ++p)
        pd[p] = PatchDesc{h.nv[p], h.ne[p], h.nf[p], h.nvo[p], h.neo[p], h.nfo[p],
                          h.v_off[p], h.e_off[p], h.f_off[p], h.vp[p], h.ep[p], h.fp[p]};
    DevMesh& d = m->dm;
    d.nV = h.nV; d.nE = h.nE; d.nF = h.nF; d.P = h.P;
    d.max_nv = h.max_nv; d.max_ne = h.max_ne; d.max_nf = h.max_nf;
    int rc = RXM_OK;
    if ((rc = upload(m, pd, d.patch))) return rc;
    if ((rc = upload(m, h.ev_local, d.ev))) return rc;
    if ((rc = upload(m, h.fe_local, d.fe))) return rc;
    if ((rc = upload(m, h.v_lin, d.v_lin))) return rc;
    if ((rc = upload(m, h.e_lin, d.e_lin))) return rc;
    if ((rc = upload(m, h.f_lin, d.f_lin))) return rc;
    if ((rc = upload(m, h.v_own, d.v_own))) return rc;
    if ((rc = upload(m, h.e_own, d.e_own))) return rc;
    if ((rc = upload(m, h.f_own, d.f_own))) return rc;
    if ((rc = upload(m, h.ltog_v, d.ltog_v))) return rc;
    if ((rc = upload(m, h.ltog_e, d.ltog_e))) return rc;
    if ((rc = upload(m, h.ltog_f, d.ltog_f))) return rc;
    if ((rc = upload(m, h.v_l2g, d.v_l2g))) return rc;
    if ((rc = upload(m, h.e_l2g, d.e_l2g))) return rc;
    if ((rc = upload(m, h.f_l2g, d.f_l2g))) return rc;
    if ((rc = upload(m, h.v_g2l, d.v_g2l))) return rc;
    if ((rc = upload(m, h.e_g2l, d.e_g2l))) return rc;
    if ((rc = upload(m, h.f_g2l, d.f_g2l))) return rc;
    m->on_device = true;
    return RXM_OK;
}

int finish_create(rxm_mesh* m, uint32_t flags, rxm_mesh** out)
{
    int rc = build_patches(m->h);
    if (rc == RXM_OK && !(flags & RXM_HOST_ONLY)) rc = upload_mesh(m);
    if (rc != RXM_OK) {
        rxm_mesh_destroy(m);
        return rc;
    }
    *out = m;
    return RXM_OK;
}

// Per-patch entry base of a query (static mesh: counts are fixed), P + 1.
const std::vector<uint32_t>& op_base(rxm_mesh* m, int op)
{
    auto it = m->h_base.find(op);
    if (it != m->h_base.end()) return it->second;
    const HostMesh& h = m->h;
    std::vector<uint32_t> b(h.P + 1, 0);
    std::vector<uint32_t> vval(h.nV, 0);
    if (op == RXM_OP_VV || op == RXM_OP_VE || op == kOpCot)
        for (uint32_t e = 0; e < h.nE; ++e) {
            vval[h.ev[2ull * e]]++;
            vval[h.ev[2ull * e + 1]]++;
        }
    for (uint32_t p = 0; p < h.P; ++p) {
        uint64_t c = 0;
        switch (op) {
            case RXM_OP_VV:
            case RXM_OP_VE:
                for (uint32_t l = 0; l < h.nvo[p]; ++l) c += vval[h.ltog_v[h.v_off[p] + l]];
                break;
            case RXM_OP_VF:
                for (uint32_t l = 0; l < h.nvo[p]; ++l) {
                    uint32_t g = h.ltog_v[h.v_off[p] + l];
                    c += h.vf_off[g + 1] - h.vf_off[g];
                }
                break;
            case RXM_OP_EF:
                for (uint32_t l = 0; l < h.neo[p]; ++l) {
                    uint32_t g = h.ltog_e[h.e_off[p] + l];
                    c += h.ef_off[g + 1] - h.ef_off[g];
                }
                break;
            case RXM_OP_FF:
                for (uint32_t l = 0; l < h.nfo[p]; ++l) {
                    uint32_t g = h.ltog_f[h.f_off[p] + l];
                    for (int i = 0; i < 3; ++i) {
                        uint32_t e = h.fe_g[3ull * g + i];
                        c += h.ef_off[e + 1] - h.ef_off[e] - 1;
                    }
                }
                break;
            case RXM_OP_EV: c = 2ull * h.neo[p]; break;
            case RXM_OP_EVD: c = 4ull * h.neo[p]; break;
            case kOpCot:
                for (uint32_t l = 0; l < h.nvo[p]; ++l) c += vval[h.ltog_v[h.v_off[p] + l]] + 1;
                break;
            default: c = 3ull * h.nfo[p]; break;
        }
        b[p + 1] = b[p] + static_cast<uint32_t>(c);
    }
    return m->h_base[op] = std::move(b);
}

// Largest number of entries one patch contributes to an op.
uint32_t op_max_patch(rxm_mesh* m, int op)
{
    auto it = m->h_maxrow.find(op);
    if (it != m->h_maxrow.end()) return it->second;
    const auto& b  = op_base(m, op);
    uint32_t    mx = 0;
    for (uint32_t p = 0; p < m->h.P; ++p) mx = std::max(mx, b[p + 1] - b[p]);
    return m->h_maxrow[op] = mx;
}

// Sizes of the one-pass query carves, computed once per mesh: the largest
// VV + VF entry count of a patch and its largest ribbon vertex, edge and
// face counts.
void vvf_slots(rxm_mesh* m)
{
    if (m->vvf_nslots) return;
    const auto& vb = op_base(m, RXM_OP_VV);
    const auto& fb = op_base(m, RXM_OP_VF);
    for (uint32_t p = 0; p < m->h.P; ++p) {
        m->vvf_nslots = std::max(m->vvf_nslots, vb[p + 1] - vb[p] + fb[p + 1] - fb[p]);
        m->max_nrv    = std::max(m->max_nrv, m->h.nv[p] - m->h.nvo[p]);
        m->max_nre    = std::max(m->max_nre, m->h.ne[p] - m->h.neo[p]);
        m->max_nrf    = std::max(m->max_nrf, m->h.nf[p] - m->h.nfo[p]);
    }
    m->vvf_nslots = std::max(m->vvf_nslots, 1u);
}

// The patch-local MCF and CG kernels stage patch p's run of the cotangent
// system's values as [vv_base[p] + vp[p], vv_base[p + 1] + vp[p + 1]): the
// VV entries of its owned rows plus one diagonal per row. Every such run
// must start where the system's row layout puts the patch and end at or
// before rxm_cot_nnz, so no workgroup reads past the caller's value array.
// Checked once per mesh on the host.
int check_system_runs(rxm_mesh* m)
{
    if (m->system_runs_ok) return RXM_OK;
    const auto&     vb = op_base(m, RXM_OP_VV);
    const auto&     cb = op_base(m, kOpCot);
    const HostMesh& h  = m->h;
    const uint64_t  nnz = cb.back();
    for (uint32_t p = 0; p < h.P; ++p) {
        const uint64_t start = uint64_t(vb[p]) + h.vp[p];
        const uint64_t end   = uint64_t(vb[p + 1]) + h.vp[p] + h.nvo[p];
        if (start != cb[p] || end != cb[p + 1] || end > nnz) return RXM_ERR_INVALID_ARG;
    }
    m->system_runs_ok = true;
    return RXM_OK;
}

int device_base(rxm_mesh* m, int op, const uint32_t** out)
{
    auto it = m->d_base.find(op);
    if (it == m->d_base.end()) {
        const auto& b = op_base(m, op);
        const uint32_t* p = nullptr;
        int rc = upload(m, b, p);
        if (rc) return rc;
        it = m->d_base.emplace(op, const_cast<uint32_t*>(p)).first;
    }
    *out = it->second;
    return RXM_OK;
}

// Launch a one-workgroup-per-patch kernel with `lds` bytes of dynamic LDS.
constexpr uint32_t kMaxLdsBytes = 160u * 1024u;  // LDS of one gfx950 CU

template <typename K, typename... Args>
int launch_grid(K kernel, uint32_t grid, uint32_t threads, uint32_t lds, void* stream, Args... args)
{
    if (lds > kMaxLdsBytes) return RXM_ERR_PATCH_TOO_BIG;
    if (lds > 64u * 1024u)
        RXM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), lds, static_cast<hipStream_t>(stream), args...);
    RXM_HIP(hipGetLastError());
    return RXM_OK;
}

// one workgroup per patch, block index = patch
template <typename K, typename... Args>
int launch_patches(K kernel, uint32_t threads, uint32_t lds, void* stream, const rxm_mesh* m, Args... args)
{
    return launch_grid(kernel, m->h.P, threads, lds, stream, args...);
}

// one workgroup per patch, XCD-contiguous patch ranges (patch_of_block)
template <typename K, typename... Args>
int launch_patches_xcd(K kernel, uint32_t threads, uint32_t lds, void* stream, const rxm_mesh* m, Args... args)
{
    return launch_grid(kernel, patch_grid(m->h.P), threads, lds, stream, args...);
}

int ensure_vv_plan(rxm_mesh* m, void* stream)
{
    if (m->d_vv_nbr) return RXM_OK;
    const HostMesh& h = m->h;
    const auto&     vb = op_base(m, RXM_OP_VV);
    std::vector<uint32_t> rb(h.P + 1, 0), nb(h.P + 1, 0);
    uint32_t maxn = 0;
    for (uint32_t p = 0; p < h.P; ++p) {
        const uint32_t nent = vb[p + 1] - vb[p];
        if (nent >= 0xFFFFu) return RXM_ERR_PATCH_TOO_BIG;
        maxn = std::max(maxn, nent);
        rb[p + 1] = rb[p] + ((h.nvo[p] + 1u + 7u) & ~7u);
        nb[p + 1] = nb[p] + ((nent + 7u) & ~7u);
    }
    m->vv_max_nent = maxn;
    int rc = upload(m, rb, m->d_vv_rbase);
    if (!rc) rc = upload(m, nb, m->d_vv_nbase);
    if (rc) return rc;
    if (hipMalloc(&m->d_vv_roff, sizeof(uint16_t) * std::max<uint32_t>(rb[h.P], 8)) != hipSuccess)
        return RXM_ERR_OUT_OF_MEMORY;
    m->allocs.push_back(m->d_vv_roff);
    if (hipMalloc(&m->d_vv_nbr, sizeof(uint16_t) * std::max<uint32_t>(nb[h.P], 8)) != hipSuccess)
        return RXM_ERR_OUT_OF_MEMORY;
    m->allocs.push_back(m->d_vv_nbr);
    m->device_bytes += 2ull * (rb[h.P] + nb[h.P]);
    Bump b(nullptr);
    plan_q2(b, 0, h.max_nv, h.max_ne, h.max_nf);
    rc = launch_patches(k_query2_vv_local<kQueryThreads>, kQueryThreads, b.bytes(), stream, m, m->dm,
                        m->d_vv_rbase, m->d_vv_nbase, m->d_vv_roff, m->d_vv_nbr);
    if (rc) {
        m->d_vv_nbr = nullptr;  // freed with the mesh; rebuild on the next call
        return rc;
    }
    return RXM_OK;
}

int ensure_face_plan(rxm_mesh* m, void* stream)
{
    if (m->d_fp_vf) return RXM_OK;
    const HostMesh& h = m->h;
    const auto&     vb = op_base(m, RXM_OP_VF);
    std::vector<uint32_t> fb(h.P + 1, 0), rb(h.P + 1, 0), nb(h.P + 1, 0);
    uint32_t maxn = 0;
    for (uint32_t p = 0; p < h.P; ++p) {
        const uint32_t nent = vb[p + 1] - vb[p];
        if (nent >= 0xFFFFu || h.nf[p] >= 16384u) return RXM_ERR_PATCH_TOO_BIG;
        maxn = std::max(maxn, nent);
        fb[p + 1] = fb[p] + ((3u * h.nf[p] + 7u) & ~7u);
        rb[p + 1] = rb[p] + ((h.nvo[p] + 1u + 7u) & ~7u);
        nb[p + 1] = nb[p] + ((nent + 7u) & ~7u);
    }
    int rc = upload(m, fb, m->d_fp_fbase);
    if (!rc) rc = upload(m, rb, m->d_fp_rbase);
    if (!rc) rc = upload(m, nb, m->d_fp_nbase);
    if (rc) return rc;
    uint16_t* bufs[3] = {nullptr, nullptr, nullptr};
    const uint32_t sizes[3] = {fb[h.P], rb[h.P], nb[h.P]};
    for (int i = 0; i < 3; ++i) {
        if (hipMalloc(&bufs[i], sizeof(uint16_t) * std::max<uint32_t>(sizes[i], 8)) != hipSuccess)
            return RXM_ERR_OUT_OF_MEMORY;
        m->allocs.push_back(bufs[i]);
        m->device_bytes += 2ull * sizes[i];
    }
    Bump b(nullptr);
    plan_q2(b, 1, h.max_nv, h.max_ne, h.max_nf);
    rc = launch_patches(k_query2_vf_local<kQueryThreads>, kQueryThreads, b.bytes(), stream, m, m->dm, m->d_fp_fbase,
                        m->d_fp_rbase, m->d_fp_nbase, bufs[0], bufs[1], bufs[2]);
    if (rc) return rc;
    m->d_fp_fvl = bufs[0];
    m->d_fp_roff = bufs[1];
    m->d_fp_vf = bufs[2];
    m->fp_max_nent = maxn;
    return RXM_OK;
}

template <bool AOS4>
int smooth_once_t(rxm_mesh* m, float lambda, const float* src, float* dst, void* stream)
{
    Bump b(nullptr);
    plan_sl(b, m->h.max_nv, m->h.max_nvo, m->vv_max_nent, AOS4);
    const uint16_t* ro = m->d_vv_roff;
    const uint16_t* nb = m->d_vv_nbr;
    switch (m->smooth_threads) {
        case 64:
            return launch_patches_xcd(k_smooth_local<64, AOS4>, 64, b.bytes(), stream, m, m->dm, m->d_vv_rbase,
                                      m->d_vv_nbase, ro, nb, lambda, src, dst);
        case 256:
            return launch_patches_xcd(k_smooth_local<256, AOS4>, 256, b.bytes(), stream, m, m->dm, m->d_vv_rbase,
                                      m->d_vv_nbase, ro, nb, lambda, src, dst);
        case 512:
            return launch_patches_xcd(k_smooth_local<512, AOS4>, 512, b.bytes(), stream, m, m->dm, m->d_vv_rbase,
                                      m->d_vv_nbase, ro, nb, lambda, src, dst);
        default:
            return launch_patches_xcd(k_smooth_local<128, AOS4>, 128, b.bytes(), stream, m, m->dm, m->d_vv_rbase,
                                      m->d_vv_nbase, ro, nb, lambda, src, dst);
    }
}

int host_vv2_plan(const rxm_mesh* m)
{
    if (m->vv2_built) return RXM_OK;
    int rc = build_vv2_plan(m->h, m->vv2);
    if (rc == RXM_OK) m->vv2_built = true;
    return rc;
}

int ensure_vv2_plan(rxm_mesh* m)
{
    if (m->d_vv2_nbr) return RXM_OK;
    int rc = host_vv2_plan(m);
    const Vv2Plan& v = m->vv2;
    if (!rc) rc = upload(m, v.rbase, m->d_vv2_rbase);
    if (!rc) rc = upload(m, v.nbase, m->d_vv2_nbase);
    if (!rc) rc = upload(m, v.xbase, m->d_vv2_xbase);
    if (!rc) rc = upload(m, v.nx, m->d_vv2_xcnt);
    if (!rc) rc = upload(m, v.roff, m->d_vv2_roff);
    if (!rc) rc = upload(m, v.xlin, m->d_vv2_xlin);
    if (!rc) rc = upload(m, v.nbr, m->d_vv2_nbr);  // last: marks the plan complete
    if (rc) m->d_vv2_nbr = nullptr;
    return rc;
}

// Static part of the cotangent system, built once per mesh (the role of the
// reference's SparseMatrix constructor): the patch-local VV edge pattern, the
// closed-fan bitmask (per linear vertex id) and the list of patches with
// owned vertices outside it, for the ring-row pass.
int ensure_cot_plan(rxm_mesh* m)
{
    if (m->d_cot_closed) return RXM_OK;
    const HostMesh&       h = m->h;
    std::vector<uint16_t> ent;
    std::vector<uint32_t> rs;
    if (!build_vv_edge_pattern(h, op_base(m, RXM_OP_VV), ent, rs)) return RXM_ERR_PATCH_TOO_BIG;
    int rc0 = upload(m, ent, m->d_cot_ent);
    if (!rc0) rc0 = upload(m, rs, m->d_cot_rs);
    if (rc0) return rc0;
    std::vector<uint8_t> closed;
    closed_fan_vertices(h, closed);
    std::vector<uint32_t> bits((h.nV + 31) / 32, 0), plist;
    for (uint32_t p = 0; p < h.P; ++p) {
        bool ring = false;
        for (uint32_t l = 0; l < h.nvo[p]; ++l) {
            const uint32_t lin = h.vp[p] + l;
            if (closed[h.ltog_v[h.v_off[p] + l]]) bits[lin >> 5] |= 1u << (lin & 31u);
            else ring = true;
        }
        if (ring) plist.push_back(p);
    }
    const uint32_t* dl = nullptr;
    int             rc = upload(m, plist, dl);
    if (!rc) rc = upload(m, bits, m->d_cot_closed);  // last: marks the plan complete
    if (rc) {
        m->d_cot_closed = nullptr;
        return rc;
    }
    m->d_cot_plist = dl;
    m->cot_nplist  = static_cast<uint32_t>(plist.size());
    return RXM_OK;
}

constexpr uint32_t kSmoothLevels = 6;  // rings of the K-level plan: up to 6 steps per launch

int ensure_vvk_plan(rxm_mesh* m)
{
    if (m->d_vvk_nbr) return RXM_OK;
    int rc = build_vvk_plan(m->h, kSmoothLevels, m->vvk);
    const VvkPlan& v = m->vvk;
    if (!rc) rc = upload(m, v.lvl, m->d_vvk_lvl);
    if (!rc) rc = upload(m, v.rbase, m->d_vvk_rbase);
    if (!rc) rc = upload(m, v.nbase, m->d_vvk_nbase);
    if (!rc) rc = upload(m, v.xbase, m->d_vvk_xbase);
    if (!rc) rc = upload(m, v.roff, m->d_vvk_roff);
    if (!rc) rc = upload(m, v.xlin, m->d_vvk_xlin);
    if (!rc) rc = upload(m, v.nbr, m->d_vvk_nbr);  // last: marks the plan complete
    if (rc) m->d_vvk_nbr = nullptr;
    return rc;
}

// `steps` Jacobi steps in one launch over the K-level plan
int smooth_steps(rxm_mesh* m, uint32_t steps, float lambda, const float* src, float* dst, void* stream)
{
    const VvkPlan& v = m->vvk;
    Bump           b(nullptr);
    plan_sk(b, v.max_cnt[steps], v.max_cnt[steps - 1], v.max_ent[steps - 1]);
    auto go = [&](auto kernel, uint32_t T) {
        return launch_patches_xcd(kernel, T, b.bytes(), stream, m, m->dm, v.K, m->d_vvk_lvl, m->d_vvk_rbase,
                                  m->d_vvk_nbase, m->d_vvk_xbase, m->d_vvk_roff, m->d_vvk_nbr, m->d_vvk_xlin, steps,
                                  lambda, src, dst);
    };
    switch (m->smooth_threads) {
        case 64: return go(k_smoothk<64>, 64);
        case 256: return go(k_smoothk<256>, 256);
        case 512: return go(k_smoothk<512>, 512);
        default: return go(k_smoothk<128>, 128);
    }
}

// two Jacobi steps in one launch
template <bool AOS4>
int smooth_twice_t(rxm_mesh* m, float lambda, const float* src, float* dst, void* stream)
{
    Bump b(nullptr);
    plan_s2(b, m->h.max_nv, m->vv2.max_nx, m->vv2.max_nent, AOS4);
    auto go = [&](auto kernel, uint32_t T) {
        return launch_patches_xcd(kernel, T, b.bytes(), stream, m, m->dm, m->d_vv2_rbase, m->d_vv2_nbase,
                                  m->d_vv2_xbase, m->d_vv2_xcnt, m->d_vv2_roff, m->d_vv2_nbr, m->d_vv2_xlin, lambda,
                                  src, dst);
    };
    switch (m->smooth_threads) {
        case 64: return go(k_smooth2<64, AOS4>, 64);
        case 256: return go(k_smooth2<256, AOS4>, 256);
        case 512: return go(k_smooth2<512, AOS4>, 512);
        default: return go(k_smooth2<128, AOS4>, 128);
    }
}

// two Jacobi steps, SoA-float-array LDS layout, with the HBM layout of the
// input / output positions chosen per launch (float4 iterates in between)
template <bool GIN, bool GOUT>
int smooth_twice_g(rxm_mesh* m, float lambda, const float* src, float* dst, void* stream)
{
    Bump b(nullptr);
    plan_s2(b, m->h.max_nv, m->vv2.max_nx, m->vv2.max_nent, false);
    auto go = [&](auto kernel, uint32_t T) {
        return launch_patches_xcd(kernel, T, b.bytes(), stream, m, m->dm, m->d_vv2_rbase, m->d_vv2_nbase,
                                  m->d_vv2_xbase, m->d_vv2_xcnt, m->d_vv2_roff, m->d_vv2_nbr, m->d_vv2_xlin, lambda,
                                  src, dst);
    };
    switch (m->smooth_threads) {
        case 64: return go(k_smooth2<64, false, GIN, GOUT>, 64);
        case 256: return go(k_smooth2<256, false, GIN, GOUT>, 256);
        case 512: return go(k_smooth2<512, false, GIN, GOUT>, 512);
        default: return go(k_smooth2<128, false, GIN, GOUT>, 128);
    }
}

int smooth_twice(rxm_mesh* m, float lambda, const float* src, float* dst, void* stream)
{
    return m->smooth_aos4 ? smooth_twice_t<true>(m, lambda, src, dst, stream)
                          : smooth_twice_t<false>(m, lambda, src, dst, stream);
}

int smooth_once(rxm_mesh* m, float lambda, const float* src, float* dst, void* stream)
{
    return m->smooth_aos4 ? smooth_once_t<true>(m, lambda, src, dst, stream)
                          : smooth_once_t<false>(m, lambda, src, dst, stream);
}

// The connectivity step of the build: on the device (edges, EV and the VF /
// EF CSRs by stable radix sorts) unless the mesh is host-only; both give the
// same arrays.
int supporting(HostMesh& h, const uint32_t* fv, uint32_t nF, uint32_t nV_hint, uint32_t flags)
{
    if (flags & RXM_HOST_ONLY) return build_supporting(h, fv, nF, nV_hint);
    int rc = supporting_prepare(h, fv, nF, nV_hint);
    if (rc != RXM_OK) return rc;
    rc = build_supporting_device(h.fv.data(), nF, h.nV, h.ev, h.fe_g, h.vf_off, h.vf_val, h.ef_off, h.ef_val, h.nE);
    return rc != RXM_OK ? rc : supporting_finish(h);
}

bool valid_op(int op)
{
    return op == RXM_OP_VV || op == RXM_OP_VE || op == RXM_OP_VF || op == RXM_OP_FV ||
           op == RXM_OP_FE || op == RXM_OP_FF || op == RXM_OP_EV || op == RXM_OP_EF || op == RXM_OP_EVD;
}

}  // namespace

extern "C" {

rxm_status rxm_mesh_create(const uint32_t* fv, uint32_t num_faces, uint32_t num_vertices,
                           uint32_t patch_size, const char* patch_file, uint32_t flags, rxm_mesh** out)
{
    if (!out || !fv || num_faces == 0) return RXM_ERR_INVALID_ARG;
    *out = nullptr;
    rxm_mesh* m = new (std::nothrow) rxm_mesh();
    if (!m) return RXM_ERR_OUT_OF_MEMORY;
    ThreadLimit serial((flags & RXM_BUILD_SERIAL) ? 1u : 0u);  // 0: keep the calling thread's cap
    const auto  t0 = std::chrono::steady_clock::now();
    int rc = supporting(m->h, fv, num_faces, num_vertices, flags);
    const auto t1 = std::chrono::steady_clock::now();
    if (rc == RXM_OK) {
        if (patch_file && patch_file[0]) {
            std::vector<uint32_t> fp, pv, po;
            uint32_t              P = 0, nV = 0, nE = 0, nF = 0;
            rc = read_patch_file(patch_file, fp, pv, po, P, nV, nE, nF);
            if (rc == RXM_OK && (nF != m->h.nF || nE != m->h.nE || nV != m->h.nV)) rc = RXM_ERR_INVALID_ARG;
            if (rc == RXM_OK) rc = set_patches(m->h, fp.data(), P, pv.data(), po.data());
            m->h.patch_size = patch_size;
        } else {
            rc = run_patcher(m->h, patch_size, (flags & RXM_PATCHER_KMEANS) ? kPatcherKMeans : kPatcherBisect);
        }
    }
    if (rc != RXM_OK) {
        delete m;
        return rc;
    }
    const auto t2 = std::chrono::steady_clock::now();
    rc = finish_create(m, flags, out);
    if (std::getenv("RXM_TIMING")) {
        using sec = std::chrono::duration<double>;
        fprintf(stderr, "mesh_create: supporting %.3f s, patcher %.3f s, patches + upload %.3f s\n",
                sec(t1 - t0).count(), sec(t2 - t1).count(), sec(std::chrono::steady_clock::now() - t2).count());
    }
    return rc;
}

rxm_status rxm_mesh_create_with_patches(const uint32_t* fv, uint32_t num_faces, uint32_t num_vertices,
                                        const uint32_t* face_patch, uint32_t num_patches, uint32_t flags,
                                        rxm_mesh** out)
{
    if (!out || !fv || !face_patch || num_faces == 0) return RXM_ERR_INVALID_ARG;
    *out = nullptr;
    rxm_mesh* m = new (std::nothrow) rxm_mesh();
    if (!m) return RXM_ERR_OUT_OF_MEMORY;
    int rc = supporting(m->h, fv, num_faces, num_vertices, flags);
    if (rc == RXM_OK) rc = set_patches(m->h, face_patch, num_patches, nullptr, nullptr);
    if (rc != RXM_OK) {
        delete m;
        return rc;
    }
    m->h.patch_size = 0;
    return finish_create(m, flags, out);
}

rxm_status rxm_mesh_create_batch(const uint32_t* const* fvs, const uint32_t* num_faces,
                                 const uint32_t* num_vertices, uint32_t count, uint32_t patch_size, uint32_t flags,
                                 rxm_mesh** out)
{
    if (!out || !fvs || !num_faces || count == 0) return RXM_ERR_INVALID_ARG;
    *out = nullptr;
    std::vector<std::unique_ptr<HostMesh>> parts(count);
    auto t0 = std::chrono::steady_clock::now();
    // the meshes are built concurrently, largest first, each worker taking
    // the next one; the first error in mesh order is returned
    std::vector<uint32_t> order(count);
    for (uint32_t k = 0; k < count; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return num_faces[a] > num_faces[b]; });
    std::vector<int>      status(count, RXM_OK);
    // a mesh above an even share of the faces would alone outlast the
    // others' concurrent builds: such meshes are built first, one at a time,
    // with every host thread; the rest concurrently, one thread each
    const unsigned workers = std::max(1u, std::min<unsigned>(num_threads(), count));
    uint64_t       total   = 0;
    for (uint32_t k = 0; k < count; ++k) total += num_faces[k];
    uint32_t nbig = 0;
    while (nbig < count && workers > 1 && uint64_t(num_faces[order[nbig]]) * workers > total) ++nbig;
    std::atomic<uint32_t> next{nbig};
    auto build_one = [&](uint32_t k) {
        parts[k].reset(new (std::nothrow) HostMesh());
        if (!parts[k]) {
            status[k] = RXM_ERR_OUT_OF_MEMORY;
            return;
        }
        int rc = supporting(*parts[k], fvs[k], num_faces[k], num_vertices ? num_vertices[k] : 0, flags);
        if (rc == RXM_OK)
            rc = run_patcher(*parts[k], patch_size, (flags & RXM_PATCHER_KMEANS) ? kPatcherKMeans : kPatcherBisect);
        if (rc == RXM_OK) rc = build_patches(*parts[k]);
        status[k] = rc;
    };
    for (uint32_t i = 0; i < nbig; ++i) build_one(order[i]);
    auto work = [&] {
        ThreadLimit serial(1);  // one mesh per thread: its host build runs serially
        for (uint32_t i; (i = next.fetch_add(1)) < count;) {
            build_one(order[i]);
        }
    };
    {
        std::vector<std::thread> pool;
        for (unsigned w = 1; w < workers; ++w) pool.emplace_back(work);
        work();
        for (auto& th : pool) th.join();
    }
    int rc = RXM_OK;
    for (uint32_t k = 0; k < count && rc == RXM_OK; ++k) rc = status[k];
    if (rc != RXM_OK) return rc;
    rxm_mesh* m = new (std::nothrow) rxm_mesh();
    if (!m) return RXM_ERR_OUT_OF_MEMORY;
    std::vector<const HostMesh*> ptrs;
    for (auto& q : parts) ptrs.push_back(q.get());
    auto t1 = std::chrono::steady_clock::now();
    rc = concat_meshes(ptrs, m->h, m->b_v, m->b_e, m->b_f, m->b_p);
    auto t2 = std::chrono::steady_clock::now();
    parts.clear();
    auto t3 = std::chrono::steady_clock::now();
    if (std::getenv("RXM_TIMING"))
        fprintf(stderr, "batch: build %.3f s, concat %.3f s, free %.3f s\n",
                std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count(),
                std::chrono::duration<double>(t3 - t2).count());
    if (rc == RXM_OK && !(flags & RXM_HOST_ONLY)) rc = upload_mesh(m);
    if (rc != RXM_OK) {
        rxm_mesh_destroy(m);
        return rc;
    }
    *out = m;
    return RXM_OK;
}

rxm_status rxm_mesh_batch_from(const rxm_mesh* const* meshes, uint32_t count, uint32_t flags, rxm_mesh** out)
{
    if (!out || !meshes || count == 0) return RXM_ERR_INVALID_ARG;
    *out = nullptr;
    std::vector<const HostMesh*> ptrs;
    for (uint32_t k = 0; k < count; ++k) {
        if (!meshes[k] || !meshes[k]->b_v.empty()) return RXM_ERR_INVALID_ARG;  // no nested batches
        ptrs.push_back(&meshes[k]->h);
    }
    rxm_mesh* m = new (std::nothrow) rxm_mesh();
    if (!m) return RXM_ERR_OUT_OF_MEMORY;
    const auto t0 = std::chrono::steady_clock::now();
    int        rc = concat_meshes(ptrs, m->h, m->b_v, m->b_e, m->b_f, m->b_p);
    const auto t1 = std::chrono::steady_clock::now();
    if (rc == RXM_OK && !(flags & RXM_HOST_ONLY)) rc = upload_mesh(m);
    if (std::getenv("RXM_TIMING"))
        fprintf(stderr, "batch_from: concat %.3f s, upload %.3f s\n", std::chrono::duration<double>(t1 - t0).count(),
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count());
    if (rc != RXM_OK) {
        rxm_mesh_destroy(m);
        return rc;
    }
    *out = m;
    return RXM_OK;
}

rxm_status rxm_batch_offsets(const rxm_mesh* m, int kind, uint32_t* out, uint32_t count)
{
    if (!m || !out || kind < 0 || kind > 3) return RXM_ERR_INVALID_ARG;
    const std::vector<uint32_t>& v = (kind == 0) ? m->b_v : (kind == 1) ? m->b_e : (kind == 2) ? m->b_f : m->b_p;
    if (v.empty()) {  // a single mesh is a batch of one
        if (count < 2) return RXM_ERR_INVALID_ARG;
        out[0] = 0;
        out[1] = (kind == 0) ? m->h.nV : (kind == 1) ? m->h.nE : (kind == 2) ? m->h.nF : m->h.P;
        return RXM_OK;
    }
    if (count < v.size()) return RXM_ERR_INVALID_ARG;
    std::memcpy(out, v.data(), v.size() * sizeof(uint32_t));
    return RXM_OK;
}

rxm_status rxm_batch_count(const rxm_mesh* m, uint32_t* count)
{
    if (!m || !count) return RXM_ERR_INVALID_ARG;
    *count = m->b_v.empty() ? 1u : static_cast<uint32_t>(m->b_v.size() - 1);
    return RXM_OK;
}

rxm_status rxm_mesh_destroy(rxm_mesh* m)
{
    if (!m) return RXM_ERR_INVALID_ARG;
    int rc = RXM_OK;
    for (void* p : m->allocs)
        if (hipFree(p) != hipSuccess) rc = RXM_ERR_DEVICE;
    delete m;
    return rc;
}

rxm_status rxm_mesh_info(const rxm_mesh* m, rxm_mesh_info_t* info)
{
    if (!m || !info) return RXM_ERR_INVALID_ARG;
    const HostMesh& h = m->h;
    std::memset(info, 0, sizeof(*info));
    info->num_vertices = h.nV;
    info->num_edges = h.nE;
    info->num_faces = h.nF;
    info->num_patches = h.P;
    info->patch_size = h.patch_size;
    info->max_vertices_per_patch = h.max_nv;
    info->max_edges_per_patch = h.max_ne;
    info->max_faces_per_patch = h.max_nf;
    info->max_owned_vertices = h.max_nvo;
    info->max_owned_edges = h.max_neo;
    info->max_owned_faces = h.max_nfo;
    info->is_closed = h.closed;
    info->is_edge_manifold = h.manifold;
    info->is_consistently_oriented = h.oriented;
    info->max_valence = h.max_valence;
    info->max_edge_incident_faces = h.max_edge_faces;
    info->max_face_adjacent_faces = h.max_face_adj;
    info->num_ribbon_faces = static_cast<uint32_t>(h.ribbon_val.size());
    info->on_device = m->on_device;
    info->device_bytes = m->device_bytes;
    return RXM_OK;
}

static int array_ref(const rxm_mesh* m, int which, const void** ptr, uint64_t* count, uint32_t* esize)
{
    const HostMesh& h = m->h;
    static thread_local std::vector<uint32_t> tmp;
    *esize = 4;
    auto set = [&](const std::vector<uint32_t>& v) {
        *ptr   = v.data();
        *count = v.size();
    };
    switch (which) {
        case RXM_ARR_EV: set(h.ev); break;
        case RXM_ARR_FACE_PATCH: set(h.face_patch); break;
        case RXM_ARR_VERTEX_PATCH: set(h.vertex_patch); break;
        case RXM_ARR_EDGE_PATCH: set(h.edge_patch); break;
        case RXM_ARR_PATCHES_VAL: set(h.patches_val); break;
        case RXM_ARR_PATCHES_OFFSET: set(h.patches_offset); break;
        case RXM_ARR_RIBBON_VAL: set(h.ribbon_val); break;
        case RXM_ARR_RIBBON_OFFSET: set(h.ribbon_offset); break;
        case RXM_ARR_V_L2G: set(h.v_l2g); break;
        case RXM_ARR_E_L2G: set(h.e_l2g); break;
        case RXM_ARR_F_L2G: set(h.f_l2g); break;
        case RXM_ARR_V_G2L: set(h.v_g2l); break;
        case RXM_ARR_E_G2L: set(h.e_g2l); break;
        case RXM_ARR_F_G2L: set(h.f_g2l); break;
        case RXM_ARR_PATCH_COUNTS:
            tmp.assign(6ull * h.P, 0);
            for (uint32_t p = 0; p < h.P; ++p) {
                tmp[6 * p] = h.nv[p]; tmp[6 * p + 1] = h.ne[p]; tmp[6 * p + 2] = h.nf[p];
                tmp[6 * p + 3] = h.nvo[p]; tmp[6 * p + 4] = h.neo[p]; tmp[6 * p + 5] = h.nfo[p];
            }
            set(tmp);
            break;
        case RXM_ARR_PATCH_OFFSETS:
            tmp.clear();
            tmp.insert(tmp.end(), h.v_off.begin(), h.v_off.end());
            tmp.insert(tmp.end(), h.e_off.begin(), h.e_off.end());
            tmp.insert(tmp.end(), h.f_off.begin(), h.f_off.end());
            set(tmp);
            break;
        case RXM_ARR_LTOG_V: set(h.ltog_v); break;
        case RXM_ARR_LTOG_E: set(h.ltog_e); break;
        case RXM_ARR_LTOG_F: set(h.ltog_f); break;
        case RXM_ARR_EV_LOCAL: set(h.ev_local); break;
        case RXM_ARR_FE_LOCAL:
            *ptr   = h.fe_local.data();
            *count = h.fe_local.size();
            *esize = 2;
            break;
        case RXM_ARR_VV2_ROFF:
        case RXM_ARR_VV2_NBR:
        case RXM_ARR_VV2_XLIN:
        case RXM_ARR_VV2_OFFSETS: {
            const int rc = host_vv2_plan(m);
            if (rc) return rc;
            const Vv2Plan& v = m->vv2;
            if (which == RXM_ARR_VV2_ROFF || which == RXM_ARR_VV2_NBR) {
                const auto& a = (which == RXM_ARR_VV2_ROFF) ? v.roff : v.nbr;
                *ptr   = a.data();
                *count = a.size();
                *esize = 2;
            } else if (which == RXM_ARR_VV2_XLIN) {
                set(v.xlin);
            } else {
                tmp.clear();
                tmp.insert(tmp.end(), v.rbase.begin(), v.rbase.end());
                tmp.insert(tmp.end(), v.nbase.begin(), v.nbase.end());
                tmp.insert(tmp.end(), v.xbase.begin(), v.xbase.end());
                tmp.insert(tmp.end(), v.nx.begin(), v.nx.end());
                set(tmp);
            }
            break;
        }
        default: return RXM_ERR_INVALID_ARG;
    }
    return RXM_OK;
}

rxm_status rxm_mesh_array_size(const rxm_mesh* m, int which, uint64_t* count)
{
    if (!m || !count) return RXM_ERR_INVALID_ARG;
    const void* p;
    uint32_t    es;
    return array_ref(m, which, &p, count, &es);
}

rxm_status rxm_mesh_get_array(const rxm_mesh* m, int which, void* dst, uint64_t bytes)
{
    if (!m || !dst) return RXM_ERR_INVALID_ARG;
    const void* p;
    uint64_t    n;
    uint32_t    es;
    int         rc = array_ref(m, which, &p, &n, &es);
    if (rc) return rc;
    if (bytes < n * es) return RXM_ERR_INVALID_ARG;
    std::memcpy(dst, p, n * es);
    return RXM_OK;
}

rxm_status rxm_mesh_device_view(const rxm_mesh* m, void* out, uint64_t bytes)
{
    if (!m || !out || bytes < sizeof(DevMesh)) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    std::memcpy(out, &m->dm, sizeof(DevMesh));
    return RXM_OK;
}

rxm_status rxm_mesh_save_patches(const rxm_mesh* m, const char* path)
{
    if (!m || !path) return RXM_ERR_INVALID_ARG;
    return write_patch_file(path, m->h);
}

rxm_status rxm_query_nnz(rxm_mesh* m, int op, int64_t* nnz)
{
    if (!m || !nnz || !valid_op(op)) return RXM_ERR_INVALID_ARG;
    *nnz = op_base(m, op).back();
    return RXM_OK;
}

rxm_status rxm_query_csr(rxm_mesh* m, int op, int oriented, int32_t* row_ptr, int32_t* col, void* stream)
{
    if (!m || !row_ptr || !col || !valid_op(op)) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    if (oriented && op != RXM_OP_VV && op != RXM_OP_VE) return RXM_ERR_UNSUPPORTED;
    if (oriented && !(m->h.manifold && m->h.oriented)) return RXM_ERR_NON_MANIFOLD;
    const HostMesh& h = m->h;
    const uint32_t* base = nullptr;
    int rc = device_base(m, op, &base);
    if (rc) return rc;
    Bump b(nullptr);
    constexpr uint32_t T = kQueryThreads;
    if (oriented) {  // oriented VV / VE: owned transpose columns + canonical CCW walk
        plan_vxo(b, h.max_nv, h.max_ne, h.max_nf, h.max_nv, op_max_patch(m, op));
        auto go = [&](auto kernel, uint32_t threads) {
            return launch_patches_xcd(kernel, threads, b.bytes(), stream, m, m->dm, base, row_ptr, col);
        };
        if (m->ring_threads == 512)
            return (op == RXM_OP_VV) ? go(k_query_vx_oriented<512, true>, 512)
                                     : go(k_query_vx_oriented<512, false>, 512);
        return (op == RXM_OP_VV) ? go(k_query_vx_oriented<256, true>, 256) : go(k_query_vx_oriented<256, false>, 256);
    }
    // VV / VE / VF / EF in the one-pass form (knob 10 = 2, the default) where
    // the patches fit its carve (VV / VE slots are EV's u16 indices 2 e + 1 -
    // end, 16 bits)
    if (m->query_stable == 2 && (op == RXM_OP_VV || op == RXM_OP_VE || op == RXM_OP_VF || op == RXM_OP_EF)) {
        vvf_slots(m);
        const bool     ef     = op == RXM_OP_EF;
        const uint32_t nrid   = (op == RXM_OP_VV) ? m->max_nrv : ((op == RXM_OP_VE) ? m->max_nre : m->max_nrf);
        const uint32_t nslots = std::max(op_max_patch(m, op), 1u);
        Bump           bx(nullptr);
        plan_x1(bx, ef ? h.max_neo : h.max_nvo, ef ? 0u : h.max_ne, nslots, nrid);
        const bool entry16 = op == RXM_OP_VF || ef || 2u * h.max_ne <= 0x10000u;
        if (bx.bytes() <= kMaxLdsBytes && entry16) {
            auto go = [&](auto kernel) {
                return launch_patches_xcd(kernel, T, bx.bytes(), stream, m, m->dm, nslots, nrid, base, row_ptr, col);
            };
            switch (op) {
                case RXM_OP_VV: return go(k_query2_x1<T, kX1VV>);
                case RXM_OP_VE: return go(k_query2_x1<T, kX1VE>);
                case RXM_OP_VF: return go(k_query2_x1<T, kX1VF>);
                default: return go(k_query2_x1<T, kX1EF>);
            }
        }
    }
    if (m->query_stable == 2 && op == RXM_OP_FF) {  // FF in one pass
        vvf_slots(m);
        const uint32_t nout = std::max(op_max_patch(m, op), 1u);
        Bump           bf(nullptr);
        plan_ff1(bf, h.max_ne, h.max_nf, h.max_nfo, nout, m->max_nrf);
        if (bf.bytes() <= kMaxLdsBytes)
            return launch_patches_xcd(k_query2_ff1<T>, T, bf.bytes(), stream, m, m->dm, nout, m->max_nrf, base, row_ptr,
                                      col);
    }
    // transposed relations: atomic-free stable counting sort, or the LDS
    // atomic histogram + row sort (knob 10); identical rows. The stable
    // sort's (key, wave) table takes T / 64 times the histogram's LDS, so a
    // mesh whose largest patch does not fit that plan takes the atomic form.
    auto q2 = [&](int kind, auto k_atomic, auto k_stable) {
        Bump bs(nullptr), ba(nullptr);
        plan_q2(bs, kind, h.max_nv, h.max_ne, h.max_nf, T / 64u);
        plan_q2(ba, kind, h.max_nv, h.max_ne, h.max_nf, 1u);
        if (m->query_stable == 1 && bs.bytes() <= kMaxLdsBytes)
            return launch_patches_xcd(k_stable, T, bs.bytes(), stream, m, m->dm, base, row_ptr, col);
        return launch_patches_xcd(k_atomic, T, ba.bytes(), stream, m, m->dm, base, row_ptr, col);
    };
    switch (op) {
        case RXM_OP_VV: return q2(0, k_query2_vx<T, true, false>, k_query2_vx<T, true, true>);
        case RXM_OP_VE: return q2(3, k_query2_vx<T, false, false>, k_query2_vx<T, false, true>);
        case RXM_OP_VF: return q2(1, k_query2_vf<T, false>, k_query2_vf<T, true>);
        case RXM_OP_EF: return q2(2, k_query2_ef<T, false, false>, k_query2_ef<T, false, true>);
        case RXM_OP_FF: return q2(4, k_query2_ef<T, true, false>, k_query2_ef<T, true, true>);
        case RXM_OP_EV:
            plan_fx(b, h.max_ne, h.max_nf);
            return launch_patches(k_query_fixed<9>, T, b.bytes(), stream, m, m->dm, row_ptr, col);
        case RXM_OP_FV:
            plan_fx(b, h.max_ne, h.max_nf);
            return launch_patches(k_query_fixed<6>, T, b.bytes(), stream, m, m->dm, row_ptr, col);
        case RXM_OP_FE:
            plan_fx(b, h.max_ne, h.max_nf);
            return launch_patches(k_query_fixed<7>, T, b.bytes(), stream, m, m->dm, row_ptr, col);
        case RXM_OP_EVD:
            plan_evd(b, h.max_ne, h.max_nf, h.max_neo);
            return launch_patches_xcd(k_query_evd<T>, T, b.bytes(), stream, m, m->dm, row_ptr, col);
        default: return RXM_ERR_INVALID_ARG;
    }
}

}  // extern "C"

// The VV and VF CSRs from one launch (rxm_query_vv_vf without its checks).
static int query_vv_vf_launch(rxm_mesh* m, int32_t* vv_row_ptr, int32_t* vv_col, int32_t* vf_row_ptr,
                              int32_t* vf_col, void* stream)
{
    const uint32_t *vv_base = nullptr, *vf_base = nullptr;
    int             rc = device_base(m, RXM_OP_VV, &vv_base);
    if (!rc) rc = device_base(m, RXM_OP_VF, &vf_base);
    if (rc) return rc;
    constexpr uint32_t T = kQueryThreads;
    if (m->query_stable == 2) {  // one counting sort for both relations
        vvf_slots(m);
        Bump b1(nullptr);
        plan_vvf1(b1, m->h.max_nvo, m->h.max_ne, m->vvf_nslots, m->max_nrv, m->max_nrf);
        // VV slots hold EV u16 indices (2 e + 1 - end) in 16 bits
        if (b1.bytes() <= kMaxLdsBytes && 2u * m->h.max_ne <= 0x10000u)
            return launch_patches_xcd(k_query2_vv_vf1<T>, T, b1.bytes(), stream, m, m->dm, m->vvf_nslots, m->max_nrv,
                                      m->max_nrf, vv_base, vf_base, vv_row_ptr, vv_col, vf_row_ptr, vf_col);
    }
    Bump bs(nullptr), ba(nullptr);  // stable / atomic plans, as in rxm_query_csr
    plan_vvf(bs, m->h.max_nv, m->h.max_ne, m->h.max_nf, T / 64u);
    plan_vvf16(ba, m->h.max_nv, m->h.max_nvo, m->h.max_ne, m->h.max_nf);
    if (m->query_stable == 1 && bs.bytes() <= kMaxLdsBytes)
        return launch_patches_xcd(k_query2_vv_vf<T, true>, T, bs.bytes(), stream, m, m->dm, vv_base, vf_base,
                                  vv_row_ptr, vv_col, vf_row_ptr, vf_col);
    return launch_patches_xcd(k_query2_vv_vf<T, false>, T, ba.bytes(), stream, m, m->dm, vv_base, vf_base, vv_row_ptr,
                              vv_col, vf_row_ptr, vf_col);
}

extern "C" {

rxm_status rxm_query_vv_vf(rxm_mesh* m, int32_t* vv_row_ptr, int32_t* vv_col, int32_t* vf_row_ptr, int32_t* vf_col,
                           void* stream)
{
    if (!m || !vv_row_ptr || !vv_col || !vf_row_ptr || !vf_col) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    return query_vv_vf_launch(m, vv_row_ptr, vv_col, vf_row_ptr, vf_col, stream);
}

rxm_status rxm_vertex_normals(rxm_mesh* m, int mode, const float* pos, float* normals, void* stream)
{
    if (!m || !pos || !normals || (mode != RXM_NORMALS_MAX && mode != RXM_NORMALS_AREA)) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    // the face plan (built once per mesh) carries the VF rows; meshes whose
    // patches exceed its u16 encoding use the self-contained kernel
    const int prc = ensure_face_plan(m, stream);
    if (prc == RXM_OK) {
        Bump b(nullptr);
        plan_vnp(b, m->h.max_nv, m->h.max_nvo, m->h.max_nf, m->fp_max_nent);
        return launch_patches_xcd(k_vertex_normals_plan<256>, 256, b.bytes(), stream, m, m->dm, mode, m->d_fp_fbase,
                              m->d_fp_rbase, m->d_fp_nbase, m->d_fp_fvl, m->d_fp_roff, m->d_fp_vf, pos, normals);
    }
    if (prc != RXM_ERR_PATCH_TOO_BIG) return prc;
    Bump b(nullptr);
    plan_vn(b, m->h.max_nv, m->h.max_ne, m->h.max_nf);
    return launch_patches(k_vertex_normals, kGeomThreads, b.bytes(), stream, m, m->dm, mode, pos, normals);
}

rxm_status rxm_gaussian_curvature(rxm_mesh* m, const float* pos, float* gc, float* area_mixed, void* stream)
{
    if (!m || !pos || !gc) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    const int rc = ensure_face_plan(m, stream);
    if (rc) return rc;
    Bump b(nullptr);
    plan_gc(b, m->h.max_nv, m->h.max_nvo, m->h.max_nf, m->fp_max_nent);
    return launch_patches_xcd(k_gaussian_curvature<512>, 512, b.bytes(), stream, m, m->dm, m->d_fp_fbase,
                              m->d_fp_rbase, m->d_fp_nbase, m->d_fp_fvl, m->d_fp_roff, m->d_fp_vf, pos, gc,
                              area_mixed);
}

// ---- ARAP-style deformation -------------------------------------------------
rxm_status rxm_arap_weights(rxm_mesh* m, const float* pos, float* w, void* stream)
{
    if (!m || !pos || !w) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    Bump b(nullptr);
    plan_aw(b, m->h.max_nv, m->h.max_ne, m->h.max_nf, m->h.max_neo);
    return launch_patches_xcd(k_arap_weights<256>, 256, b.bytes(), stream, m, m->dm, pos, w);
}

namespace {
// the owned rows' edges (the VV edge pattern the cotangent assembly caches)
int arap_rows(rxm_mesh* m, const uint32_t** vv_base, uint32_t* lds, bool rot)
{
    int rc = ensure_cot_plan(m);
    if (!rc) rc = device_base(m, RXM_OP_VV, vv_base);
    if (rc) return rc;
    Bump b(nullptr);
    plan_ar(b, m->h.max_nv, m->h.max_ne, m->h.max_nvo, op_max_patch(m, RXM_OP_VV), rot);
    *lds = b.bytes();
    return RXM_OK;
}
}  // namespace

rxm_status rxm_arap_rotations(rxm_mesh* m, const float* rest, const float* x, const float* w, float* R,
                              void* stream)
{
    if (!m || !rest || !x || !w || !R) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    const uint32_t* vv_base = nullptr;
    uint32_t        lds     = 0;
    const int       rc      = arap_rows(m, &vv_base, &lds, false);
    if (rc) return rc;
    return launch_patches_xcd(k_arap_rotations<512>, 512, lds, stream, m, m->dm, vv_base,
                              static_cast<const uint32_t*>(m->d_cot_rs), m->d_cot_ent,
                              rest, x, w, R);
}

rxm_status rxm_arap_energy_grad(rxm_mesh* m, const float* rest, const float* x, const float* w, const float* R,
                                float* total, float* grad, void* stream)
{
    if (!m || !rest || !x || !w || !R || !total || !grad) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    const uint32_t* vv_base = nullptr;
    uint32_t        lds     = 0;
    int             rc      = arap_rows(m, &vv_base, &lds, true);
    if (rc) return rc;
    if (!m->d_partial) {
        if (hipMalloc(&m->d_partial, sizeof(float) * m->h.P) != hipSuccess) return RXM_ERR_OUT_OF_MEMORY;
        m->allocs.push_back(m->d_partial);
        m->device_bytes += sizeof(float) * m->h.P;
    }
    rc = launch_patches_xcd(k_arap_energy_grad<512>, 512, lds, stream, m, m->dm, vv_base,
                            static_cast<const uint32_t*>(m->d_cot_rs), m->d_cot_ent,
                            rest, x, w, R, m->d_partial, grad);
    if (rc) return rc;
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(1024), 0, static_cast<hipStream_t>(stream), m->h.P,
                       static_cast<const float*>(m->d_partial), total);
    RXM_HIP(hipGetLastError());
    return RXM_OK;
}

rxm_status rxm_vv_plan_build(rxm_mesh* m, void* stream)
{
    if (!m) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    return ensure_vv_plan(m, stream);
}

rxm_status rxm_laplacian_smooth(rxm_mesh* m, const float* pos, float* out, float* tmp, float lambda, int iters,
                                void* stream)
{
    if (!m || !pos || !out || iters < 1 || pos == out) return RXM_ERR_INVALID_ARG;
    if (iters > 1 && (!tmp || tmp == out || tmp == pos)) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    int rc = RXM_OK;
    if (m->smooth_steps > 0) {  // K-level plan: launches of up to smooth_steps steps
        if ((rc = ensure_vvk_plan(m))) return rc;
        const uint32_t S        = m->smooth_steps;
        const int      launches = static_cast<int>((iters + S - 1) / S);
        const float*   src      = pos;
        int            left     = iters;
        for (int i = 0; i < launches; ++i) {
            float*         dst = ((launches - 1 - i) % 2 == 0) ? out : tmp;
            const uint32_t s   = std::min<uint32_t>(S, static_cast<uint32_t>(left));
            if ((rc = smooth_steps(m, s, lambda, src, dst, stream))) return rc;
            left -= static_cast<int>(s);
            src = dst;
        }
        return RXM_OK;
    }
    rc = ensure_vv_plan(m, stream);
    if (rc) return rc;
    const bool fused = m->smooth_fused && iters > 1 && ensure_vv2_plan(m) == RXM_OK;
    if (fused && m->smooth_g4 && !m->smooth_aos4 && iters >= 4) {
        // an odd step first (SoA -> SoA), then pairs whose iterates between
        // launches are float4 rows in two mesh-owned buffers
        const size_t n = m->h.nV;
        if (!m->d_iter4) {
            float* p = nullptr;
            if (hipMalloc(&p, 8 * n * sizeof(float)) != hipSuccess) return RXM_ERR_OUT_OF_MEMORY;
            m->allocs.push_back(p);
            m->device_bytes += 8 * n * sizeof(float);
            m->d_iter4 = p;
        }
        const float* src = pos;
        if (iters % 2) {
            if ((rc = smooth_once(m, lambda, pos, tmp, stream))) return rc;
            src = tmp;
        }
        const int pairs = iters / 2;
        float*    buf[2] = {m->d_iter4, m->d_iter4 + 4 * n};
        for (int i = 0; i < pairs && !rc; ++i) {
            const bool first = (i == 0), last = (i == pairs - 1);
            float*     dst   = last ? out : buf[i & 1];
            if (first) rc = smooth_twice_g<false, true>(m, lambda, src, dst, stream);
            else if (last) rc = smooth_twice_g<true, false>(m, lambda, src, dst, stream);
            else rc = smooth_twice_g<true, true>(m, lambda, src, dst, stream);
            src = dst;
        }
        return rc;
    }
    // launches: pairs of steps (fused) then a single one; the last writes out
    const int    launches = fused ? iters / 2 + iters % 2 : iters;
    const float* src      = pos;
    for (int i = 0; i < launches; ++i) {
        float* dst = ((launches - 1 - i) % 2 == 0) ? out : tmp;
        const bool pair = fused && (i < iters / 2);
        rc = pair ? smooth_twice(m, lambda, src, dst, stream) : smooth_once(m, lambda, src, dst, stream);
        if (rc) return rc;
        src = dst;
    }
    return RXM_OK;
}

rxm_status rxm_set_launch_threads(rxm_mesh* m, int kernel, int threads)
{
    if (!m) return RXM_ERR_INVALID_ARG;
    if (kernel == 0 && (threads == 64 || threads == 128 || threads == 256 || threads == 512)) {
        m->smooth_threads = static_cast<uint32_t>(threads);
        return RXM_OK;
    }
    if (kernel == 1 && (threads == 128 || threads == 256)) {
        m->grad_threads = static_cast<uint32_t>(threads);
        return RXM_OK;
    }
    if (kernel == 2 && (threads == 0 || threads == 1)) {  // gradient step launch: 0 grid, 1 work queue
        m->grad_queue = static_cast<uint32_t>(threads);
        return RXM_OK;
    }
    if (kernel == 3 && (threads == 0 || threads == 1)) {  // smoothing LDS layout: 0 three float arrays, 1 float4
        m->smooth_aos4 = static_cast<uint32_t>(threads);
        return RXM_OK;
    }
    if (kernel == 4 && (threads == 0 || threads == 1)) {  // smoothing: 1 = two Jacobi steps per launch
        m->smooth_fused = static_cast<uint32_t>(threads);
        return RXM_OK;
    }
    if (kernel == 5 && (threads == 1 || threads == 2 || threads == 4 || threads == 8)) {  // SpMV-3 lanes per row
        m->spmv_lanes = static_cast<uint32_t>(threads);
        return RXM_OK;
    }
    if (kernel == 6 && (threads == 256 || threads == 512)) {
        m->ring_threads = static_cast<uint32_t>(threads);
        return RXM_OK;
    }
    // query transposes: 0 LDS-atomic counting sort + row sort, 1 atomic-free
    // stable sort, 2 (default) as 0 with the fused VV + VF launch in one
    // counting sort over both relations (k_query2_vv_vf1)
    if (kernel == 10 && threads >= 0 && threads <= 2) {
        m->query_stable = static_cast<uint32_t>(threads);
        return RXM_OK;
    }
    if (kernel == 7 && (threads == 256 || threads == 512)) {
        m->cot_threads = static_cast<uint32_t>(threads);
        return RXM_OK;
    }
    if (kernel == 9 && (threads == 0 || threads == 1)) {  // float4 iterates between two-step launches
        m->smooth_g4 = static_cast<uint32_t>(threads);
        return RXM_OK;
    }
    if (kernel == 8 && threads >= 0 && threads <= static_cast<int>(kSmoothLevels)) {
        m->smooth_steps = static_cast<uint32_t>(threads);
        return RXM_OK;
    }
    return RXM_ERR_INVALID_ARG;
}

rxm_status rxm_laplacian_smooth_csr(rxm_mesh* m, const int32_t* row_ptr, const int32_t* col, const float* pos,
                                    float* out, float* tmp, float lambda, int iters, void* stream)
{
    if (!m || !row_ptr || !col || !pos || !out || iters < 1 || pos == out) return RXM_ERR_INVALID_ARG;
    if (iters > 1 && (!tmp || tmp == out || tmp == pos)) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    // one lane per row: rows hold about six entries, so wider rows idle lanes
    // and pay a shuffle reduction; one lane also sums in row order, the
    // order the patch-local smoothing uses
    constexpr uint32_t LANES = 1;
    const uint32_t     n     = m->h.nV;
    const uint64_t     thr   = uint64_t(n) * LANES;
    const float*       src   = pos;
    for (int it = 0; it < iters; ++it) {
        float* dst = ((iters - 1 - it) % 2 == 0) ? out : tmp;
        hipLaunchKernelGGL((k_smooth_csr<LANES>), dim3(static_cast<uint32_t>((thr + 255) / 256)), dim3(256), 0,
                           static_cast<hipStream_t>(stream), n, row_ptr, col, lambda, src, dst);
        RXM_HIP(hipGetLastError());
        src = dst;
    }
    return RXM_OK;
}

static rxm_status permute(rxm_mesh* m, int kind, int ncomp, const float* src, float* dst, bool to_linear,
                          void* stream)
{
    if (!m || !src || !dst || src == dst || kind < 0 || kind > 2 || ncomp < 1) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    const DevMesh& d = m->dm;
    const uint32_t n = (kind == 0) ? d.nV : (kind == 1) ? d.nE : d.nF;
    // to linear: dst[l] = src[l2g[l]]; to global: dst[g] = src[g2l[g]]
    const uint32_t* map = to_linear ? ((kind == 0) ? d.v_l2g : (kind == 1) ? d.e_l2g : d.f_l2g)
                                    : ((kind == 0) ? d.v_g2l : (kind == 1) ? d.e_g2l : d.f_g2l);
    hipLaunchKernelGGL(k_gather_rows, dim3((n + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), n,
                       static_cast<uint32_t>(ncomp), map, src, dst);
    RXM_HIP(hipGetLastError());
    return RXM_OK;
}

rxm_status rxm_to_linear(rxm_mesh* m, int kind, int ncomp, const float* g, float* l, void* stream)
{
    return permute(m, kind, ncomp, g, l, true, stream);
}

rxm_status rxm_to_global(rxm_mesh* m, int kind, int ncomp, const float* l, float* g, void* stream)
{
    return permute(m, kind, ncomp, l, g, false, stream);
}

rxm_status rxm_cot_nnz(rxm_mesh* m, int64_t* nnz)
{
    if (!m || !nnz) return RXM_ERR_INVALID_ARG;
    *nnz = op_base(m, kOpCot).back();
    return RXM_OK;
}

rxm_status rxm_cot_csr_assemble(rxm_mesh* m, const float* pos, float dt, int32_t* row_ptr, int32_t* col,
                                float* val, float* mass, void* stream)
{
    if (!m || !pos || !row_ptr || !col || !val || !mass) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    if (!(m->h.manifold && m->h.oriented)) return RXM_ERR_NON_MANIFOLD;
    const uint32_t* base = nullptr;
    int rc = device_base(m, kOpCot, &base);
    if (rc) return rc;
    const uint32_t* vv_base = nullptr;
    if ((rc = device_base(m, RXM_OP_VV, &vv_base))) return rc;
    if ((rc = ensure_cot_plan(m))) return rc;
    const uint32_t maxrow = op_max_patch(m, kOpCot);
    Bump           b(nullptr);
    plan_cot(b, m->h.max_nv, m->h.max_ne, m->h.max_nvo, op_max_patch(m, RXM_OP_VV));
    auto go = [&](auto kernel, uint32_t threads) {
        return launch_patches_xcd(kernel, threads, b.bytes(), stream, m, m->dm, vv_base, m->d_cot_rs, m->d_cot_ent,
                                  pos, dt, row_ptr, col, val, mass);
    };
    rc = (m->cot_threads == 512) ? go(k_cot_assemble<512>, 512) : go(k_cot_assemble<256>, 256);
    if (rc || m->cot_nplist == 0) return rc;
    Bump br(nullptr);
    plan_cot_ring(br, m->h.max_nv, m->h.max_ne, m->h.max_nf, m->h.max_nv, maxrow);
    return launch_grid(k_cot_ring_rows<kQueryThreads>, m->cot_nplist, kQueryThreads, br.bytes(), stream, m->dm,
                       base, m->d_cot_plist, m->d_cot_closed, pos, dt, val, mass);
}

}  // extern "C"

// lanes per CSR row of the SpMV-3 kernels (autotuned: 1, 2, 4 or 8)
template <bool FUSED>
int launch_spmv3(rxm_mesh* m, const int32_t* row_ptr, const int32_t* col, const float* val, const float* mass,
                 const float* x, float* y, void* stream)
{
    const uint32_t n       = m->h.nV;
    const uint32_t lanes   = m->spmv_lanes;
    const uint64_t threads = uint64_t(n) * lanes;
    const dim3     grid(static_cast<uint32_t>((threads + 255) / 256)), block(256);
    hipStream_t    s = static_cast<hipStream_t>(stream);
    switch (lanes) {
        case 1: hipLaunchKernelGGL((k_spmv3<1, FUSED>), grid, block, 0, s, n, row_ptr, col, val, mass, x, y); break;
        case 2: hipLaunchKernelGGL((k_spmv3<2, FUSED>), grid, block, 0, s, n, row_ptr, col, val, mass, x, y); break;
        case 4: hipLaunchKernelGGL((k_spmv3<4, FUSED>), grid, block, 0, s, n, row_ptr, col, val, mass, x, y); break;
        default: hipLaunchKernelGGL((k_spmv3<8, FUSED>), grid, block, 0, s, n, row_ptr, col, val, mass, x, y); break;
    }
    RXM_HIP(hipGetLastError());
    return RXM_OK;
}

extern "C" {

rxm_status rxm_spmv3(rxm_mesh* m, const int32_t* row_ptr, const int32_t* col, const float* val, const float* x,
                     float* y, void* stream)
{
    if (!m || !row_ptr || !col || !val || !x || !y || x == y) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    return launch_spmv3<false>(m, row_ptr, col, val, nullptr, x, y, stream);
}

rxm_status rxm_mcf_explicit(rxm_mesh* m, const int32_t* row_ptr, const int32_t* col, const float* val,
                            const float* mass, const float* x, float* out, float* tmp, int iters, void* stream)
{
    if (!m || !row_ptr || !col || !val || !mass || !x || !out || iters < 1 || x == out) return RXM_ERR_INVALID_ARG;
    if (iters > 1 && (!tmp || tmp == out || tmp == x)) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    // patch-local steps over the VV plan (the system's pattern); ping-pong
    // between out and tmp so that the last step writes out
    int rc = check_system_runs(m);
    if (!rc) rc = ensure_vv_plan(m, stream);
    if (rc) return rc;
    const uint32_t* vv_base = nullptr;
    if ((rc = device_base(m, RXM_OP_VV, &vv_base))) return rc;
    Bump b(nullptr);
    plan_sl(b, m->h.max_nv, m->h.max_nvo, m->vv_max_nent, false);
    b.alloc<float>(m->vv_max_nent + m->h.max_nvo);
    const float* src = x;
    for (int it = 0; it < iters && !rc; ++it) {
        float* dst = ((iters - 1 - it) % 2 == 0) ? out : tmp;
        rc = launch_patches_xcd(k_mcf_patch<256>, 256, b.bytes(), stream, m, m->dm, vv_base, m->d_vv_rbase,
                                m->d_vv_nbase, m->d_vv_roff, m->d_vv_nbr, val, mass, src, dst);
        src = dst;
    }
    return rc;
}

rxm_status rxm_mcf_implicit_cg(rxm_mesh* m, const int32_t* row_ptr, const int32_t* col, const float* val,
                               const float* mass, const float* x, float* out, int max_iter, float abs_tol,
                               float rel_tol, int* iters, float* start_residual, float* final_residual,
                               void* stream)
{
    if (!m || !row_ptr || !col || !val || !mass || !x || !out || x == out || max_iter < 0)
        return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    int rc = check_system_runs(m);
    if (!rc) rc = ensure_vv_plan(m, stream);
    if (rc) return rc;
    const uint32_t* vv_base = nullptr;
    if ((rc = device_base(m, RXM_OP_VV, &vv_base))) return rc;
    const uint32_t n = m->h.nV, P = m->h.P, nb = (n + 255u) / 256u;
    const size_t   v3 = size_t(3) * n, npart = std::max<size_t>(P, nb);
    if (!m->d_cg) {
        float* p = nullptr;
        if (hipMalloc(&p, (4 * v3 + npart) * sizeof(float)) != hipSuccess) return RXM_ERR_OUT_OF_MEMORY;
        m->allocs.push_back(p);
        m->device_bytes += (4 * v3 + npart) * sizeof(float);
        CgState* st = nullptr;
        if (hipMalloc(&st, sizeof(CgState)) != hipSuccess) return RXM_ERR_OUT_OF_MEMORY;
        m->allocs.push_back(st);
        m->d_cg       = p;
        m->d_cg_state = st;
    }
    float *    r = m->d_cg, *pv[2] = {m->d_cg + v3, m->d_cg + 2 * v3}, *s = m->d_cg + 3 * v3, *part = m->d_cg + 4 * v3;
    CgState*   st = m->d_cg_state;
    hipStream_t hs = static_cast<hipStream_t>(stream);
    Bump b(nullptr);
    plan_sl(b, m->h.max_nv, m->h.max_nvo, m->vv_max_nent, false);
    b.alloc<float>(m->vv_max_nent + m->h.max_nvo);
    b.alloc<float>(256 / 64);
    const uint32_t lds = b.bytes();
    auto apply = [&](auto kernel, const float* xin, float* pnew, float* y) {
        return launch_patches_xcd(kernel, 256, lds, stream, m, m->dm, vv_base, m->d_vv_rbase, m->d_vv_nbase,
                                  m->d_vv_roff, m->d_vv_nbr, val, xin, static_cast<const float*>(r), pnew, y, part,
                                  static_cast<const CgState*>(st));
    };
    // set-up: s = A x0, r = M x0 - s, x = x0, delta = <r, r>
    if ((rc = apply(k_cg_apply<256, 0>, x, nullptr, s))) return rc;
    hipLaunchKernelGGL((k_cg_update<0>), dim3(nb), dim3(256), 0, hs, n, mass, x, out, r, nullptr, s, part, st);
    hipLaunchKernelGGL((k_cg_scalar<0>), dim3(1), dim3(1024), 0, hs, nb, part, st, abs_tol, rel_tol);
    for (int it = 0; it < max_iter; ++it) {
        float* pold = pv[it & 1];
        float* pnew = pv[(it + 1) & 1];
        // the first iteration stages p = r (p_old unread), later ones r + beta p_old
        rc = it == 0 ? apply(k_cg_apply<256, 1>, pold, pnew, s) : apply(k_cg_apply<256, 2>, pold, pnew, s);
        if (rc) return rc;
        hipLaunchKernelGGL((k_cg_scalar<1>), dim3(1), dim3(1024), 0, hs, P, part, st, abs_tol, rel_tol);
        hipLaunchKernelGGL((k_cg_update<1>), dim3(nb), dim3(256), 0, hs, n, mass, x, out, r, pnew, s, part, st);
        hipLaunchKernelGGL((k_cg_scalar<2>), dim3(1), dim3(1024), 0, hs, nb, part, st, abs_tol, rel_tol);
    }
    RXM_HIP(hipGetLastError());
    if (iters || start_residual || final_residual) {
        CgState h{};
        RXM_HIP(hipMemcpyAsync(&h, st, sizeof(CgState), hipMemcpyDeviceToHost, hs));
        RXM_HIP(hipStreamSynchronize(hs));
        if (iters) *iters = static_cast<int>(h.iters);
        if (start_residual) *start_residual = h.start;
        if (final_residual) *final_residual = h.delta;
    }
    return RXM_OK;
}

rxm_status rxm_mcf_matfree_cg(rxm_mesh* m, const float* coords, const float* x, float* out, float dt, int uniform,
                              int precond, int max_iter, float abs_tol, float rel_tol, int* iters,
                              float* start_residual, float* final_residual, void* stream)
{
    if (!m || !coords || !x || !out || x == out || coords == out || max_iter < 0) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    const bool uni = uniform != 0, pre = precond != 0;
    int        rc  = RXM_OK;
    if (!uni) {
        if (!(m->h.manifold && m->h.oriented)) return RXM_ERR_NON_MANIFOLD;
        if ((rc = ensure_cot_plan(m))) return rc;
        // the face pass gives a vertex its reference weights only when its
        // ring is one closed fan (a closed vertex-manifold mesh)
        if (m->cot_nplist != 0) return RXM_ERR_UNSUPPORTED;
    } else if ((rc = ensure_cot_plan(m))) {
        return rc;
    }
    const uint32_t* vv_base = nullptr;
    if ((rc = device_base(m, RXM_OP_VV, &vv_base))) return rc;
    const uint32_t n = m->h.nV, P = m->h.P, nb = (n + 255u) / 256u;
    const size_t   v3 = size_t(3) * n, npart = std::max<size_t>(P, nb);
    if (!m->d_mf) {
        float* p = nullptr;
        const size_t cnt = 4 * v3 + 2 * size_t(n) + npart;
        if (hipMalloc(&p, cnt * sizeof(float)) != hipSuccess) return RXM_ERR_OUT_OF_MEMORY;
        m->allocs.push_back(p);
        m->device_bytes += cnt * sizeof(float);
        m->d_mf = p;
    }
    if (!m->d_cg_state) {
        CgState* st = nullptr;
        if (hipMalloc(&st, sizeof(CgState)) != hipSuccess) return RXM_ERR_OUT_OF_MEMORY;
        m->allocs.push_back(st);
        m->d_cg_state = st;
    }
    float *r = m->d_mf, *pv[2] = {m->d_mf + v3, m->d_mf + 2 * v3}, *s = m->d_mf + 3 * v3;
    float *mass = m->d_mf + 4 * v3, *diag = mass + n, *part = diag + n;
    CgState*    st = m->d_cg_state;
    hipStream_t hs = static_cast<hipStream_t>(stream);
    constexpr uint32_t T = 256;
    Bump b(nullptr);
    plan_mf(b, m->h.max_nv, m->h.max_ne, m->h.max_nvo, op_max_patch(m, RXM_OP_VV), uni, T);
    const uint32_t lds = b.bytes();
    float* dg = pre ? diag : nullptr;
    auto apply = [&](auto kernel, const float* xin, float* pnew, float* y) {
        return launch_patches_xcd(kernel, T, lds, stream, m, m->dm, vv_base, m->d_cot_rs, m->d_cot_ent, coords, dt,
                                  xin, static_cast<const float*>(r), dg, mass, pnew, y, part,
                                  static_cast<const CgState*>(st));
    };
    // the kernel for (mode, weights, preconditioner)
    auto launch = [&](int mode, const float* xin, float* pnew, float* y) {
#define RXM_MF(MODE)                                                                         \
    (uni ? (pre ? apply(k_mf_apply<T, MODE, true, true>, xin, pnew, y)                       \
                : apply(k_mf_apply<T, MODE, true, false>, xin, pnew, y))                     \
         : (pre ? apply(k_mf_apply<T, MODE, false, true>, xin, pnew, y)                      \
                : apply(k_mf_apply<T, MODE, false, false>, xin, pnew, y)))
        return mode == 0 ? RXM_MF(0) : (mode == 1 ? RXM_MF(1) : RXM_MF(2));
#undef RXM_MF
    };
    // set-up: s = A x0, mass (and diag), r = B - s with B = mass * x0, x = x0, delta
    if ((rc = launch(0, x, nullptr, s))) return rc;
    hipLaunchKernelGGL((k_cg_update<0>), dim3(nb), dim3(256), 0, hs, n, static_cast<const float*>(mass), x, out, r,
                       nullptr, s, part, st, static_cast<const float*>(dg));
    hipLaunchKernelGGL((k_cg_scalar<0>), dim3(1), dim3(1024), 0, hs, nb, part, st, abs_tol, rel_tol);
    for (int it = 0; it < max_iter; ++it) {
        float* pold = pv[it & 1];
        float* pnew = pv[(it + 1) & 1];
        if ((rc = launch(it == 0 ? 1 : 2, pold, pnew, s))) return rc;
        hipLaunchKernelGGL((k_cg_scalar<1>), dim3(1), dim3(1024), 0, hs, P, part, st, abs_tol, rel_tol);
        hipLaunchKernelGGL((k_cg_update<1>), dim3(nb), dim3(256), 0, hs, n, static_cast<const float*>(mass), x, out,
                           r, pnew, s, part, st, static_cast<const float*>(dg));
        hipLaunchKernelGGL((k_cg_scalar<2>), dim3(1), dim3(1024), 0, hs, nb, part, st, abs_tol, rel_tol);
    }
    RXM_HIP(hipGetLastError());
    if (iters || start_residual || final_residual) {
        CgState h{};
        RXM_HIP(hipMemcpyAsync(&h, st, sizeof(CgState), hipMemcpyDeviceToHost, hs));
        RXM_HIP(hipStreamSynchronize(hs));
        if (iters) *iters = static_cast<int>(h.iters);
        if (start_residual) *start_residual = h.start;
        if (final_residual) *final_residual = h.delta;
    }
    return RXM_OK;
}

rxm_status rxm_symdir_rest(rxm_mesh* m, const float* rest_pos, float* Mr, void* stream)
{
    if (!m || !rest_pos || !Mr) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    Bump b(nullptr);
    plan_fv(b, m->h.max_ne, m->h.max_nf);
    return launch_patches(k_symdir_rest, kGeomThreads, b.bytes(), stream, m, m->dm, rest_pos, Mr);
}

}  // extern "C"

// Per-face energies, per-patch partials into `partial` (P floats, every
// entry written), then their fixed-order sum into total[0].
static int symdir_energy_launch(rxm_mesh* m, const float* Mr, const float* uv, float* energy, float* total,
                                float* partial, void* stream)
{
    Bump b(nullptr);
    plan_fv(b, m->h.max_ne, m->h.max_nf);
    b.alloc<float>(kGeomThreads);
    int rc = launch_patches(k_symdir_energy, kGeomThreads, b.bytes(), stream, m, m->dm, Mr, uv, energy, partial);
    if (rc) return rc;
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(1024), 0, static_cast<hipStream_t>(stream), m->h.P,
                       static_cast<const float*>(partial), total);
    RXM_HIP(hipGetLastError());
    return RXM_OK;
}

extern "C" {

rxm_status rxm_symdir_energy(rxm_mesh* m, const float* Mr, const float* uv, float* energy, float* total,
                             void* stream)
{
    if (!m || !Mr || !uv || !energy || !total) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    if (!m->d_partial) {
        if (hipMalloc(&m->d_partial, sizeof(float) * m->h.P) != hipSuccess) return RXM_ERR_OUT_OF_MEMORY;
        m->allocs.push_back(m->d_partial);
        m->device_bytes += sizeof(float) * m->h.P;
    }
    return symdir_energy_launch(m, Mr, uv, energy, total, m->d_partial, stream);
}

}  // extern "C"

// Scratch of rxm_symdir_energy_noalloc: the per-patch partial sums, one float
// per patch, every entry rewritten by each call (nothing needs clearing).
// Each device holds its own copy for the life of the process.
static constexpr uint32_t kNoAllocMaxPatches = 1u << 18;  // 1 MiB
static __device__ __align__(256) float scratch[kNoAllocMaxPatches];

namespace {
constexpr int kMaxDevices = 64;
std::mutex    g_scratch_mu;
float*        g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device, resolved once per
// device (the array's name is only a host-side stand-in).
int scratch_on_current_device(float** out)
{
    int dev = -1;
    RXM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return RXM_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        RXM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return RXM_ERR_DEVICE;
        g_scratch[dev] = static_cast<float*>(p);
    }
    *out = g_scratch[dev];
    return RXM_OK;
}
}  // namespace

extern "C" {

rxm_status rxm_symdir_energy_noalloc(rxm_mesh* m, const float* Mr, const float* uv, float* energy, float* total,
                                     void* stream)
{
    if (!m || !Mr || !uv || !energy || !total) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    if (m->h.P > kNoAllocMaxPatches) return RXM_ERR_SCRATCH_TOO_SMALL;
    int dev = -1;
    RXM_HIP(hipGetDevice(&dev));
    if (dev != m->device) return RXM_ERR_INVALID_ARG;  // the mesh lives on another device
    float* part = nullptr;
    const int rc = scratch_on_current_device(&part);
    if (rc) return rc;
    return symdir_energy_launch(m, Mr, uv, energy, total, part, stream);
}

rxm_status rxm_symdir_grad(rxm_mesh* m, const float* Mr, const float* uv, const float* scale, float* grad,
                           void* stream)
{
    if (!m || !Mr || !uv || !grad || uv == grad) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    Bump b(nullptr);
    plan_sg(b, m->h.max_nv, m->h.max_ne, m->h.max_nf);
    return launch_patches(k_symdir_grad, kGeomThreads, b.bytes(), stream, m, m->dm, Mr, uv, scale, grad);
}

rxm_status rxm_symdir3_energy(rxm_mesh* m, const float* Mr, const float* x, float* energy, void* stream)
{
    if (!m || !Mr || !x || !energy) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    Bump b(nullptr);
    plan_fv(b, m->h.max_ne, m->h.max_nf);
    return launch_patches(k_symdir3_energy, kGeomThreads, b.bytes(), stream, m, m->dm, Mr, x, energy);
}

static int launch_symdir3_step(rxm_mesh* m, const float* Mr, float eta, const float* x, float* xo, void* stream)
{
    Bump b(nullptr);
    plan_s3p(b, m->h.max_nv, m->h.max_nvo, m->h.max_nf, m->fp_max_nent);
    const uint32_t  lds = b.bytes();
    const uint16_t* fvl = m->d_fp_fvl;
    const uint16_t* ro  = m->d_fp_roff;
    const uint16_t* vf  = m->d_fp_vf;
    if (m->grad_queue) {
        hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream), m->d_queue);
        RXM_HIP(hipGetLastError());
        // enough workgroups to fill the machine; LDS decides residency
        const uint32_t per_cu = std::max(1u, std::min(8u, (160u * 1024u) / std::max(lds, 1u)));
        const uint32_t grid   = std::min<uint32_t>(m->h.P, static_cast<uint32_t>(m->num_cus) * per_cu);
        if (m->grad_threads == 128)
            return launch_grid(k_symdir3_step_plan_queue<128>, grid, 128, lds, stream, m->dm, m->d_queue,
                               m->d_fp_fbase, m->d_fp_rbase, m->d_fp_nbase, fvl, ro, vf, Mr, eta, x, xo);
        return launch_grid(k_symdir3_step_plan_queue<256>, grid, 256, lds, stream, m->dm, m->d_queue, m->d_fp_fbase,
                           m->d_fp_rbase, m->d_fp_nbase, fvl, ro, vf, Mr, eta, x, xo);
    }
    if (m->grad_threads == 128)
        return launch_patches(k_symdir3_step_plan<128>, 128, lds, stream, m, m->dm, m->d_fp_fbase, m->d_fp_rbase,
                              m->d_fp_nbase, fvl, ro, vf, Mr, eta, x, xo);
    return launch_patches(k_symdir3_step_plan<256>, 256, lds, stream, m, m->dm, m->d_fp_fbase, m->d_fp_rbase,
                          m->d_fp_nbase, fvl, ro, vf, Mr, eta, x, xo);
}

// Everything the step loop needs on the device before launching (or
// capturing): the VV and face plans, the work-queue counter, the CU count.
static int prepare_steps(rxm_mesh* m, void* stream)
{
    int rc = ensure_vv_plan(m, stream);
    if (!rc) rc = ensure_face_plan(m, stream);
    if (!rc && !m->d_queue) {
        if (hipMalloc(&m->d_queue, 256) != hipSuccess) return RXM_ERR_OUT_OF_MEMORY;
        m->allocs.push_back(m->d_queue);
    }
    if (!rc && m->num_cus == 0)
        RXM_HIP(hipDeviceGetAttribute(&m->num_cus, hipDeviceAttributeMultiprocessorCount, m->device));
    return rc;
}

rxm_status rxm_symdir3_grad_step(rxm_mesh* m, const float* Mr, const float* x, float eta, float* x_out,
                                 void* stream)
{
    if (!m || !Mr || !x || !x_out || x == x_out) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    int rc = prepare_steps(m, stream);
    if (rc) return rc;
    return launch_symdir3_step(m, Mr, eta, x, x_out, stream);
}

rxm_status rxm_symdir3_grad(rxm_mesh* m, const float* Mr, const float* x, float* grad, void* stream)
{
    if (!m || !Mr || !x || !grad || x == grad) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    int rc = ensure_face_plan(m, stream);
    if (rc) return rc;
    Bump b(nullptr);
    plan_s3p(b, m->h.max_nv, m->h.max_nvo, m->h.max_nf, m->fp_max_nent);
    return launch_patches(k_symdir3_grad_plan<256>, 256, b.bytes(), stream, m, m->dm, m->d_fp_fbase, m->d_fp_rbase,
                          m->d_fp_nbase, static_cast<const uint16_t*>(m->d_fp_fvl),
                          static_cast<const uint16_t*>(m->d_fp_roff), static_cast<const uint16_t*>(m->d_fp_vf), Mr,
                          x, grad);
}

// The batched workload step: x1 = smooth(x), x = x1 - eta grad E_symdir3(x1).
// Per-step query outputs of the step loop (all four null: no query).
struct StepQuery
{
    int32_t *vv_row_ptr = nullptr, *vv_col = nullptr, *vf_row_ptr = nullptr, *vf_col = nullptr;
    bool     on() const { return vv_row_ptr != nullptr; }
    bool     valid() const
    {
        return (vv_row_ptr && vv_col && vf_row_ptr && vf_col) || (!vv_row_ptr && !vv_col && !vf_row_ptr && !vf_col);
    }
};

static int smooth_symdir_launches(rxm_mesh* m, const float* Mr, float* x, float* tmp, float lambda, float eta,
                                  int steps, void* stream, const StepQuery& q = StepQuery{})
{
    for (int s = 0; s < steps; ++s) {
        int rc = q.on() ? query_vv_vf_launch(m, q.vv_row_ptr, q.vv_col, q.vf_row_ptr, q.vf_col, stream) : RXM_OK;
        if (!rc) rc = smooth_once(m, lambda, x, tmp, stream);
        if (!rc) rc = launch_symdir3_step(m, Mr, eta, tmp, x, stream);
        if (rc) return rc;
    }
    return RXM_OK;
}

// Everything the step loop's launches need that is built or allocated once
// per mesh (plans, query bases), so that the loop itself allocates nothing.
static int prepare_query_steps(rxm_mesh* m, void* stream)
{
    int rc = prepare_steps(m, stream);
    const uint32_t* b = nullptr;
    if (!rc) rc = device_base(m, RXM_OP_VV, &b);
    if (!rc) rc = device_base(m, RXM_OP_VF, &b);
    return rc;
}

rxm_status rxm_smooth_symdir_steps(rxm_mesh* m, const float* Mr, float* x, float* tmp, float lambda, float eta,
                                   int steps, void* stream)
{
    if (!m || !Mr || !x || !tmp || x == tmp || steps < 1) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    int rc = prepare_steps(m, stream);
    if (rc) return rc;
    return smooth_symdir_launches(m, Mr, x, tmp, lambda, eta, steps, stream);
}

rxm_status rxm_smooth_symdir_steps_each(rxm_mesh* const* meshes, const float* const* Mr, float* const* x,
                                        float* const* tmp, uint32_t count, float lambda, float eta, int steps,
                                        void* stream)
{
    if (!meshes || !Mr || !x || !tmp || count == 0 || steps < 1) return RXM_ERR_INVALID_ARG;
    for (uint32_t k = 0; k < count; ++k) {
        if (!meshes[k] || !meshes[k]->on_device) return RXM_ERR_INVALID_ARG;
        int rc = prepare_steps(meshes[k], stream);
        if (rc) return rc;
    }
    // plain launches: every mesh gets its own two launches per step
    for (int s = 0; s < steps; ++s)
        for (uint32_t k = 0; k < count; ++k) {
            int rc = smooth_symdir_launches(meshes[k], Mr[k], x[k], tmp[k], lambda, eta, 1, stream);
            if (rc) return rc;
        }
    return RXM_OK;
}

rxm_status rxm_query_smooth_symdir_steps(rxm_mesh* m, const float* Mr, float* x, float* tmp, float lambda,
                                         float eta, int steps, int32_t* vv_row_ptr, int32_t* vv_col,
                                         int32_t* vf_row_ptr, int32_t* vf_col, void* stream)
{
    const StepQuery q{vv_row_ptr, vv_col, vf_row_ptr, vf_col};
    if (!m || !Mr || !x || !tmp || x == tmp || steps < 1 || !q.on() || !q.valid()) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    int rc = prepare_query_steps(m, stream);
    if (rc) return rc;
    return smooth_symdir_launches(m, Mr, x, tmp, lambda, eta, steps, stream, q);
}

rxm_status rxm_query_smooth_symdir_steps_each(rxm_mesh* const* meshes, const float* const* Mr, float* const* x,
                                              float* const* tmp, int32_t* const* vv_row_ptr, int32_t* const* vv_col,
                                              int32_t* const* vf_row_ptr, int32_t* const* vf_col, uint32_t count,
                                              float lambda, float eta, int steps, void* stream)
{
    if (!meshes || !Mr || !x || !tmp || !vv_row_ptr || !vv_col || !vf_row_ptr || !vf_col || count == 0 ||
        steps < 1)
        return RXM_ERR_INVALID_ARG;
    for (uint32_t k = 0; k < count; ++k) {
        const StepQuery q{vv_row_ptr[k], vv_col[k], vf_row_ptr[k], vf_col[k]};
        if (!meshes[k] || !meshes[k]->on_device || !q.on() || !q.valid() || !Mr[k] || !x[k] || !tmp[k] ||
            x[k] == tmp[k])
            return RXM_ERR_INVALID_ARG;
        int rc = prepare_query_steps(meshes[k], stream);
        if (rc) return rc;
    }
    // plain launches: every mesh gets its own three launches per step
    for (int s = 0; s < steps; ++s)
        for (uint32_t k = 0; k < count; ++k) {
            const StepQuery q{vv_row_ptr[k], vv_col[k], vf_row_ptr[k], vf_col[k]};
            int rc = smooth_symdir_launches(meshes[k], Mr[k], x[k], tmp[k], lambda, eta, 1, stream, q);
            if (rc) return rc;
        }
    return RXM_OK;
}

rxm_status rxm_graph_capture_smooth_symdir(rxm_mesh* m, const float* Mr, float* x, float* tmp, float lambda,
                                           float eta, int steps, rxm_graph** out)
{
    if (!m || !Mr || !x || !tmp || !out || x == tmp || steps < 1) return RXM_ERR_INVALID_ARG;
    if (!m->on_device) return RXM_ERR_UNSUPPORTED;
    *out = nullptr;
    hipStream_t s;
    RXM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int rc = prepare_steps(m, s);  // allocations and plan builds stay outside the capture
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = RXM_ERR_DEVICE;
    rxm_graph* g = nullptr;
    if (!rc) {
        g = new (std::nothrow) rxm_graph();
        if (!g) rc = RXM_ERR_OUT_OF_MEMORY;
    }
    if (!rc) {
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = RXM_ERR_DEVICE;
        else {
            int lrc = smooth_symdir_launches(m, Mr, x, tmp, lambda, eta, steps, s);
            hipError_t e = hipStreamEndCapture(s, &g->graph);
            if (lrc) rc = lrc;
            else if (e != hipSuccess) rc = RXM_ERR_DEVICE;
            else if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) rc = RXM_ERR_DEVICE;
        }
    }
    if (hipStreamDestroy(s) != hipSuccess && !rc) rc = RXM_ERR_DEVICE;
    if (rc) {
        if (g) rxm_graph_destroy(g);
        return rc;
    }
    *out = g;
    return RXM_OK;
}

rxm_status rxm_graph_launch(rxm_graph* g, void* stream)
{
    if (!g || !g->exec) return RXM_ERR_INVALID_ARG;
    RXM_HIP(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
    return RXM_OK;
}

rxm_status rxm_graph_destroy(rxm_graph* g)
{
    if (!g) return RXM_ERR_INVALID_ARG;
    int rc = RXM_OK;
    if (g->exec && hipGraphExecDestroy(g->exec) != hipSuccess) rc = RXM_ERR_DEVICE;
    if (g->graph && hipGraphDestroy(g->graph) != hipSuccess) rc = RXM_ERR_DEVICE;
    delete g;
    return rc;
}

}  // extern "C"
