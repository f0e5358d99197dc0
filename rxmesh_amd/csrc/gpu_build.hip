// Edge build on the device: the connectivity step of the host build
// (reference include/rxmesh/rxmesh.cpp:518-681, build_supporting_structures)
// as sort-based edge deduplication. Half-edge h = (f, i) runs fv[h] ->
// fv[next(h)]; edge ids number the distinct undirected edges in order of
// first appearance over h, EV keeps the first half-edge's orientation, and
// every half-edge gets its edge id -- the host build's edges, EV and fe_g.
//
//   1. key(h) = min(a, b) * nV + max(a, b), value h;
//   2. a stable radix sort by key (rocPRIM): equal keys stay in h order, so
//      the head of each run is the edge's first half-edge (its leader);
//   3. leaders are flagged at their own index; an exclusive scan of the
//      flags in h order ranks them: edge id = rank of the leader;
//   4. an inclusive max-scan of head positions gives every sorted slot its
//      run head, hence its leader and edge id.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>  // rocPRIM's texture cache iterator uses memset
#include <vector>

#include <rocprim/rocprim.hpp>

#include "../../include/rxmesh_amd.h"

namespace {

constexpr uint32_t kT = 256;

__device__ __forceinline__ uint32_t he_next(uint32_t h) { return (h % 3u == 2u) ? h - 2u : h + 1u; }

__global__ __launch_bounds__(kT) void k_he_keys(const uint32_t* __restrict__ fv, uint32_t nH, uint32_t nV,
                                                uint64_t* __restrict__ key, uint32_t* __restrict__ idx,
                                                uint32_t* __restrict__ bad)
{
    const uint32_t h = blockIdx.x * kT + threadIdx.x;
    if (h >= nH) return;
    const uint32_t a = fv[h], b = fv[he_next(h)];
    if (a >= nV || b >= nV || a == b) atomicOr(bad, 1u);
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    key[h] = (lo < nV && hi < nV) ? static_cast<uint64_t>(lo) * nV + hi : 0ull;
    idx[h] = h;
}

__global__ __launch_bounds__(kT) void k_run_heads(const uint64_t* __restrict__ key, const uint32_t* __restrict__ idx,
                                                  uint32_t nH, uint32_t* __restrict__ leader_flag,
                                                  uint32_t* __restrict__ head_pos)
{
    const uint32_t j = blockIdx.x * kT + threadIdx.x;
    if (j >= nH) return;
    const bool head = (j == 0u) || key[j] != key[j - 1u];
    head_pos[j]     = head ? j : 0u;
    if (head) leader_flag[idx[j]] = 1u;
}

__global__ __launch_bounds__(kT) void k_assign_edges(const uint32_t* __restrict__ fv, const uint32_t* __restrict__ idx,
                                                     const uint32_t* __restrict__ run_head,
                                                     const uint32_t* __restrict__ rank, uint32_t nH,
                                                     uint32_t* __restrict__ ev, uint32_t* __restrict__ fe)
{
    const uint32_t j = blockIdx.x * kT + threadIdx.x;
    if (j >= nH) return;
    const uint32_t leader = idx[run_head[j]], h = idx[j], e = rank[leader];
    fe[h] = e;
    if (h == leader) {
        ev[2u * e]      = fv[h];
        ev[2u * e + 1u] = fv[he_next(h)];
    }
}

uint32_t bits_for(uint64_t x)
{
    uint32_t b = 0;
    while (b < 64 && (x >> b) != 0ull) ++b;
    return b;
}

// key[h] = the vertex of half-edge h (VF) or its edge id (EF), value h
__global__ __launch_bounds__(kT) void k_keys_of(const uint32_t* __restrict__ src, uint32_t nH,
                                                uint32_t* __restrict__ key, uint32_t* __restrict__ idx)
{
    const uint32_t h = blockIdx.x * kT + threadIdx.x;
    if (h >= nH) return;
    key[h] = src[h];
    idx[h] = h;
}

// CSR of a stable key sort: off[r] = first sorted position with key >= r
// (binary search, r in [0, nrows]), val[j] = face of the sorted half-edge j
__global__ __launch_bounds__(kT) void k_csr_offsets(const uint32_t* __restrict__ key, uint32_t n, uint32_t nrows,
                                                    uint32_t* __restrict__ off)
{
    const uint32_t r = blockIdx.x * kT + threadIdx.x;
    if (r > nrows) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (key[mid] < r) lo = mid + 1u;
        else hi = mid;
    }
    off[r] = lo;
}
__global__ __launch_bounds__(kT) void k_faces_of(const uint32_t* __restrict__ idx, uint32_t n,
                                                 uint32_t* __restrict__ val)
{
    const uint32_t j = blockIdx.x * kT + threadIdx.x;
    if (j < n) val[j] = idx[j] / 3u;
}

}  // namespace

namespace rxm {

// Edge ids, EV, per-half-edge edge ids and the VF / EF CSRs (rows in face
// order) of a face list on the device: the edge pass of
// rxm_build_edges_device, then one stable radix sort of the half-edges by
// vertex and one by edge. fv is host memory; outputs are host vectors.
int build_supporting_device(const uint32_t* fv, uint32_t nF, uint32_t nV, std::vector<uint32_t>& ev,
                            std::vector<uint32_t>& fe, std::vector<uint32_t>& vf_off, std::vector<uint32_t>& vf_val,
                            std::vector<uint32_t>& ef_off, std::vector<uint32_t>& ef_val, uint32_t& nE)
{
    const uint64_t nH64 = 3ull * nF;
    if (nF == 0 || nV < 3 || nH64 >= 0xFFFFFFFFull) return RXM_ERR_INVALID_ARG;
    const uint32_t nH = static_cast<uint32_t>(nH64);
    hipStream_t    s  = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return RXM_ERR_DEVICE;
    // device buffers: fv, EV (2 x 3F bound), FE, and the sort's key / value pairs
    // the sorts' temporary storage, sized for every key width they use
    size_t sort_bytes = 0;
    for (uint32_t bits : {std::max(1u, bits_for(nV)), std::max(1u, bits_for(nH)), 32u}) {
        size_t b = 0;
        if (rocprim::radix_sort_pairs(nullptr, b, static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                      static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr), nH, 0, bits,
                                      s) != hipSuccess) {
            (void)hipStreamDestroy(s);
            return RXM_ERR_DEVICE;
        }
        sort_bytes = std::max(sort_bytes, b);
    }
    auto         up   = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t n    = nH;
    const size_t o_ev = up(4 * n), o_fe = o_ev + up(8 * n), o_k0 = o_fe + up(4 * n), o_k1 = o_k0 + up(4 * n);
    const size_t o_v0 = o_k1 + up(4 * n), o_v1 = o_v0 + up(4 * n), o_off = o_v1 + up(4 * n);
    const size_t o_tmp = o_off + up(4 * (size_t(std::max(nV, nH)) + 1)), total = o_tmp + up(sort_bytes);
    char*        base = nullptr;
    if (hipMalloc(&base, total) != hipSuccess) {
        (void)hipStreamDestroy(s);
        return RXM_ERR_OUT_OF_MEMORY;
    }
    uint32_t* d_fv  = reinterpret_cast<uint32_t*>(base);
    uint32_t* d_ev  = reinterpret_cast<uint32_t*>(base + o_ev);
    uint32_t* d_fe  = reinterpret_cast<uint32_t*>(base + o_fe);
    uint32_t* k0    = reinterpret_cast<uint32_t*>(base + o_k0);
    uint32_t* k1    = reinterpret_cast<uint32_t*>(base + o_k1);
    uint32_t* v0    = reinterpret_cast<uint32_t*>(base + o_v0);
    uint32_t* v1    = reinterpret_cast<uint32_t*>(base + o_v1);
    uint32_t* d_off = reinterpret_cast<uint32_t*>(base + o_off);
    void*     tmp   = base + o_tmp;
    int       rc    = RXM_OK;
    auto      ok    = [&](hipError_t e) {
        if (e != hipSuccess && rc == RXM_OK) rc = RXM_ERR_DEVICE;
        return rc == RXM_OK;
    };
    const uint32_t grid = (nH + kT - 1) / kT;
    // one stable sort of the half-edges by `src` (vertex or edge id, below
    // nrows): off and val come back as a CSR
    auto csr = [&](const uint32_t* src, uint32_t nrows, std::vector<uint32_t>& off, std::vector<uint32_t>& val) {
        hipLaunchKernelGGL(k_keys_of, dim3(grid), dim3(kT), 0, s, src, nH, k0, v0);
        size_t b = sort_bytes;
        if (!ok(hipGetLastError()) ||
            !ok(rocprim::radix_sort_pairs(tmp, b, k0, k1, v0, v1, nH, 0, std::max(1u, bits_for(nrows)), s)))
            return;
        hipLaunchKernelGGL(k_csr_offsets, dim3((nrows + kT) / kT), dim3(kT), 0, s, k1, nH, nrows, d_off);
        hipLaunchKernelGGL(k_faces_of, dim3(grid), dim3(kT), 0, s, v1, nH, k0);
        if (!ok(hipGetLastError())) return;
        off.resize(size_t(nrows) + 1);
        val.resize(n);
        if (ok(hipMemcpyAsync(off.data(), d_off, 4 * (size_t(nrows) + 1), hipMemcpyDeviceToHost, s)))
            if (ok(hipMemcpyAsync(val.data(), k0, 4 * n, hipMemcpyDeviceToHost, s))) ok(hipStreamSynchronize(s));
    };
    if (ok(hipMemcpyAsync(d_fv, fv, 4 * n, hipMemcpyHostToDevice, s))) {
        csr(d_fv, nV, vf_off, vf_val);
        if (rc == RXM_OK) rc = rxm_build_edges_device(d_fv, nF, nV, d_ev, d_fe, &nE, s);
        if (rc == RXM_OK) {
            ev.resize(2ull * nE);
            fe.resize(n);
            if (ok(hipMemcpyAsync(ev.data(), d_ev, 8ull * nE, hipMemcpyDeviceToHost, s)) &&
                ok(hipMemcpyAsync(fe.data(), d_fe, 4 * n, hipMemcpyDeviceToHost, s)))
                csr(d_fe, nE, ef_off, ef_val);
        }
    }
    (void)hipStreamSynchronize(s);  // nothing may still use the buffers
    (void)hipFree(base);
    (void)hipStreamDestroy(s);
    return rc;
}

}  // namespace rxm

extern "C" rxm_status rxm_build_edges_device(const uint32_t* fv, uint32_t num_faces, uint32_t num_vertices,
                                             uint32_t* ev, uint32_t* fe, uint32_t* num_edges, void* stream)
{
    if (!fv || !ev || !fe || !num_edges || num_faces == 0 || num_vertices < 3) return RXM_ERR_INVALID_ARG;
    const uint64_t nH64 = 3ull * num_faces;
    if (nH64 >= 0xFFFFFFFFull) return RXM_ERR_INVALID_ARG;
    const uint32_t nH = static_cast<uint32_t>(nH64), nV = num_vertices;
    hipStream_t    s  = static_cast<hipStream_t>(stream);
    const uint32_t end_bit = bits_for(static_cast<uint64_t>(nV) * nV - 1ull);

    // work buffers: keys in/out, values in/out, leader flags, ranks, head
    // positions, run heads, the error flag and the library's temporary storage
    size_t sort_bytes = 0, scan_bytes = 0, max_bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<uint64_t*>(nullptr),
                                  static_cast<uint64_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                  static_cast<uint32_t*>(nullptr), nH, 0, end_bit, s) != hipSuccess ||
        rocprim::exclusive_scan(nullptr, scan_bytes, static_cast<uint32_t*>(nullptr),
                                static_cast<uint32_t*>(nullptr), 0u, nH, rocprim::plus<uint32_t>(),
                                s) != hipSuccess ||
        rocprim::inclusive_scan(nullptr, max_bytes, static_cast<uint32_t*>(nullptr),
                                static_cast<uint32_t*>(nullptr), nH, rocprim::maximum<uint32_t>(),
                                s) != hipSuccess)
        return RXM_ERR_DEVICE;
    const size_t tmp_bytes = std::max(sort_bytes, std::max(scan_bytes, max_bytes));
    const size_t n         = nH;
    // every buffer starts on a 256-byte boundary (hipMalloc's own alignment),
    // the alignment the library expects of its temporary storage
    auto         up        = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t o_k1 = up(8 * n), o_v0 = o_k1 + up(8 * n), o_v1 = o_v0 + up(4 * n), o_flag = o_v1 + up(4 * n);
    const size_t o_rank = o_flag + up(4 * n), o_hpos = o_rank + up(4 * n), o_head = o_hpos + up(4 * n);
    const size_t o_bad = o_head + up(4 * n), o_tmp = o_bad + 256, total = o_tmp + up(tmp_bytes);
    char*        base      = nullptr;
    if (hipMalloc(&base, total) != hipSuccess) return RXM_ERR_OUT_OF_MEMORY;
    uint64_t* k0   = reinterpret_cast<uint64_t*>(base);
    uint64_t* k1   = reinterpret_cast<uint64_t*>(base + o_k1);
    uint32_t* v0   = reinterpret_cast<uint32_t*>(base + o_v0);
    uint32_t* v1   = reinterpret_cast<uint32_t*>(base + o_v1);
    uint32_t* flag = reinterpret_cast<uint32_t*>(base + o_flag);
    uint32_t* rank = reinterpret_cast<uint32_t*>(base + o_rank);
    uint32_t* hpos = reinterpret_cast<uint32_t*>(base + o_hpos);
    uint32_t* head = reinterpret_cast<uint32_t*>(base + o_head);
    uint32_t* bad  = reinterpret_cast<uint32_t*>(base + o_bad);
    void*     tmp  = base + o_tmp;

    int            rc   = RXM_OK;
    const uint32_t grid = (nH + kT - 1) / kT;
    auto           ok   = [&](hipError_t e) {
        if (e != hipSuccess && rc == RXM_OK) rc = RXM_ERR_DEVICE;
        return rc == RXM_OK;
    };
    uint32_t h_bad = 0, last[2] = {0, 0};
    if (ok(hipMemsetAsync(bad, 0, sizeof(uint32_t), s)) && ok(hipMemsetAsync(flag, 0, 4 * n, s))) {
        hipLaunchKernelGGL(k_he_keys, dim3(grid), dim3(kT), 0, s, fv, nH, nV, k0, v0, bad);
        size_t b = tmp_bytes;
        if (ok(hipGetLastError()) && ok(rocprim::radix_sort_pairs(tmp, b, k0, k1, v0, v1, nH, 0, end_bit, s))) {
            hipLaunchKernelGGL(k_run_heads, dim3(grid), dim3(kT), 0, s, k1, v1, nH, flag, hpos);
            b = tmp_bytes;
            if (ok(hipGetLastError()) &&
                ok(rocprim::exclusive_scan(tmp, b, flag, rank, 0u, nH, rocprim::plus<uint32_t>(), s))) {
                b = tmp_bytes;
                if (ok(rocprim::inclusive_scan(tmp, b, hpos, head, nH, rocprim::maximum<uint32_t>(), s))) {
                    hipLaunchKernelGGL(k_assign_edges, dim3(grid), dim3(kT), 0, s, fv, v1, head, rank, nH, ev, fe);
                    if (ok(hipGetLastError()) &&
                        ok(hipMemcpyAsync(&h_bad, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s)) &&
                        ok(hipMemcpyAsync(last, rank + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s)) &&
                        ok(hipMemcpyAsync(last + 1, flag + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s)))
                        ok(hipStreamSynchronize(s));
                }
            }
        }
    }
    (void)hipStreamSynchronize(s);  // nothing may still use the work buffers
    (void)hipFree(base);
    if (rc != RXM_OK) return rc;
    if (h_bad) return RXM_ERR_INVALID_ARG;
    *num_edges = last[0] + last[1];
    return RXM_OK;
}
