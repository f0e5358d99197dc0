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

}  // namespace

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
