// Workgroup-level building blocks of the patch query engine (gfx950, wave64).
//
// Every query runs one workgroup per patch: the patch's EV / FE topology is
// staged into LDS with 16-byte / 8-byte loads, transposed relations are built
// with an LDS counting sort, and results are produced for owned elements
// only. The counting sort is made deterministic by sorting each (short) row
// after the slot fill, so every row lists its sources in ascending local
// order -- the order a serial execution of the reference's transpose
// (include/rxmesh/kernels/rxmesh_queries.cuh:16-107) produces.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "device_mesh.hpp"

namespace rxm {

constexpr uint16_t kInv16 = 0xFFFFu;

// element count rounded up to the 4-element slab padding
__host__ __device__ inline uint32_t r4(uint32_t n) { return (n + 3u) & ~3u; }

__device__ __forceinline__ uint32_t lo16(uint32_t x) { return x & 0xFFFFu; }
__device__ __forceinline__ uint32_t hi16(uint32_t x) { return x >> 16; }

// Patch handled by this workgroup. Workgroups are dispatched round-robin over
// the 8 XCDs (each with its own L2); block b runs patch (b % 8) * ceil(P/8) +
// b / 8 so every XCD sweeps one contiguous range of patches, and the ribbon
// gathers into neighbouring patches (adjacent ids) stay in that XCD's L2.
// The grid is 8 * ceil(P/8) blocks; blocks mapped past P return at once.
__device__ __forceinline__ uint32_t patch_of_block(uint32_t P)
{
    const uint32_t chunk = (P + 7u) >> 3;
    return (blockIdx.x & 7u) * chunk + (blockIdx.x >> 3);
}
__host__ __device__ inline uint32_t patch_grid(uint32_t P) { return ((P + 7u) >> 3) << 3; }

__device__ __forceinline__ PatchDesc load_desc(const PatchDesc* __restrict__ pd, uint32_t p)
{
    const uint4* q = reinterpret_cast<const uint4*>(pd + p);
    uint4        a = q[0], b = q[1], c = q[2];
    PatchDesc    d;
    d.nv = a.x; d.ne = a.y; d.nf = a.z; d.nvo = a.w;
    d.neo = b.x; d.nfo = b.y; d.v_off = b.z; d.e_off = b.w;
    d.f_off = c.x; d.vp = c.y; d.ep = c.z; d.fp = c.w;
    return d;
}

// Copy n vector elements global -> LDS, two rounds' loads issued before
// their stores.
template <typename V>
__device__ __forceinline__ void stage_vec2(V* __restrict__ d, const V* __restrict__ s, uint32_t n)
{
    const uint32_t T = blockDim.x;
    for (uint32_t base = 0; base < n; base += 2u * T) {
        const uint32_t i0 = base + threadIdx.x, i1 = i0 + T;
        V              a{}, b{};
        if (i0 < n) a = s[i0];
        if (i1 < n) b = s[i1];
        if (i0 < n) d[i0] = a;
        if (i1 < n) d[i1] = b;
    }
}

// Copy n u32 from a 16-byte aligned, 4-element padded slab into LDS.
__device__ __forceinline__ void stage_u32(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                          uint32_t n)
{
    stage_vec2(reinterpret_cast<uint4*>(dst), reinterpret_cast<const uint4*>(src), (n + 3u) >> 2);
}

// Copy 3*nf u16 of a face row (8-byte aligned, padded to 4 faces) into LDS.
__device__ __forceinline__ void stage_fe(uint16_t* __restrict__ dst, const uint16_t* __restrict__ src,
                                         uint32_t nf)
{
    stage_vec2(reinterpret_cast<uint2*>(dst), reinterpret_cast<const uint2*>(src), (3u * nf + 3u) >> 2);
}

// Exclusive scan of s[0..n) in place; returns the total. s_wave holds
// (blockDim/64 + 1) u32. Contains the barriers it needs (entry and exit).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t* s, uint32_t n, uint32_t* s_wave)
{
    const uint32_t T    = blockDim.x;
    const uint32_t tid  = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wid  = tid >> 6;
    const uint32_t nw   = T >> 6;
    const uint32_t per  = (n + T - 1) / T;
    const uint32_t b    = min(n, tid * per);
    const uint32_t e    = min(n, b + per);
    __syncthreads();
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; ++i) sum += s[i];
    uint32_t x = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= static_cast<uint32_t>(d)) x += y;
    }
    if (lane == 63u) s_wave[wid] = x;
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            uint32_t t = s_wave[w];
            s_wave[w]  = acc;
            acc += t;
        }
        s_wave[nw] = acc;
    }
    __syncthreads();
    uint32_t run = x - sum + s_wave[wid];
    for (uint32_t i = b; i < e; ++i) {
        uint32_t t = s[i];
        s[i]       = run;
        run += t;
    }
    const uint32_t total = s_wave[nw];
    __syncthreads();
    return total;
}

// Deterministic LDS transpose (counting sort).
//   n entries, entry i belongs to source row i / W and targets column key(i).
//   off[0..ncols] <- column offsets, val[off[c]..off[c+1]) <- source rows in
//   ascending order for every column c < sort_cols (other columns hold the
//   same sets in arbitrary order). cur is ncols u32 of scratch.
template <uint32_t W, typename KeyF>
__device__ __forceinline__ void lds_transpose(uint32_t n, uint32_t ncols, KeyF key, uint32_t* off,
                                              uint32_t* cur, uint16_t* val, uint32_t* s_wave,
                                              uint32_t sort_cols)
{
    const uint32_t T = blockDim.x;
    for (uint32_t c = threadIdx.x; c < ncols; c += T) off[c] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += T) atomicAdd(&off[key(i)], 1u);
    const uint32_t total = block_exclusive_scan(off, ncols, s_wave);
    if (threadIdx.x == 0) off[ncols] = total;
    for (uint32_t c = threadIdx.x; c < ncols; c += T) cur[c] = off[c];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += T) {
        const uint32_t slot = atomicAdd(&cur[key(i)], 1u);
        val[slot]           = static_cast<uint16_t>(i / W);
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < sort_cols; c += T) {
        const uint32_t b = off[c], e = off[c + 1];
        for (uint32_t i = b + 1; i < e; ++i) {
            const uint16_t x = val[i];
            uint32_t       j = i;
            while (j > b && val[j - 1] > x) {
                val[j] = val[j - 1];
                --j;
            }
            val[j] = x;
        }
    }
    __syncthreads();
}


// Local vertex i (0..2) of local face f from LDS FE / EV (reference f_v,
// rxmesh_queries.cuh:650-688: the direction bit selects the edge endpoint).
__device__ __forceinline__ uint32_t face_vertex(const uint16_t* s_fe, const uint32_t* s_ev, uint32_t f,
                                                uint32_t i)
{
    const uint32_t x = s_fe[3u * f + i];
    const uint32_t e = s_ev[x >> 1];
    return (x & 1u) ? hi16(e) : lo16(e);
}

// Linear id of local element l of the patch (owned ids map by prefix,
// ribbon ids through the dense owner table).
__device__ __forceinline__ uint32_t lin_of(uint32_t l, uint32_t n_owned, uint32_t prefix,
                                           const uint32_t* __restrict__ lin_slab)
{
    return (l < n_owned) ? prefix + l : lin_slab[l];
}

}  // namespace rxm
