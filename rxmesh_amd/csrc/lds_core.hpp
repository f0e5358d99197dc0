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

// Copy n u32 from a 16-byte aligned, 4-element padded slab into LDS.
__device__ __forceinline__ void stage_u32(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                          uint32_t n)
{
    const uint32_t n4 = (n + 3u) >> 2;
    const uint4*   s  = reinterpret_cast<const uint4*>(src);
    uint4*         d  = reinterpret_cast<uint4*>(dst);
    for (uint32_t i = threadIdx.x; i < n4; i += blockDim.x) d[i] = s[i];
}

// Copy 3*nf u16 of a face row (8-byte aligned, padded to 4 faces) into LDS.
__device__ __forceinline__ void stage_fe(uint16_t* __restrict__ dst, const uint16_t* __restrict__ src,
                                         uint32_t nf)
{
    const uint32_t n2 = (3u * nf + 3u) >> 2;  // uint2 = 4 x u16
    const uint2*   s  = reinterpret_cast<const uint2*>(src);
    uint2*         d  = reinterpret_cast<uint2*>(dst);
    for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) d[i] = s[i];
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

// Deterministic LDS transpose without fill atomics or row sorts.
//   n entries, entry i belongs to source row i / W and targets column key(i).
//   Wave w owns the contiguous entry range [w*span, (w+1)*span); per-wave
//   column counts (LDS atomics, order-free) are scanned column-major, so wave
//   w's slots of column c follow those of waves < w. Each wave then walks its
//   range in 64-entry chunks in order; equal keys inside a chunk are ranked by
//   lane through a ballot match over the key bits. Result: every column lists
//   its source rows in ascending order, the order of a serial fill.
//   cnt: nw * ncols u32 scratch (nw = blockDim / 64); key_bits >= bits(ncols-1).
template <uint32_t W, typename KeyF>
__device__ __forceinline__ void lds_transpose_ordered(uint32_t n, uint32_t ncols, uint32_t key_bits, KeyF key,
                                                      uint32_t* off, uint32_t* cnt, uint16_t* val,
                                                      uint32_t* s_wave)
{
    const uint32_t T = blockDim.x, nw = T >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t span = ((n + nw * 64u - 1u) / (nw * 64u)) * 64u;
    const uint32_t b0 = min(n, w * span), b1 = min(n, b0 + span);
    for (uint32_t c = threadIdx.x; c < ncols * nw; c += T) cnt[c] = 0;
    __syncthreads();
    for (uint32_t i = b0 + lane; i < b1; i += 64u) atomicAdd(&cnt[key(i) * nw + w], 1u);
    const uint32_t total = block_exclusive_scan(cnt, ncols * nw, s_wave);
    for (uint32_t c = threadIdx.x; c < ncols; c += T) off[c] = cnt[c * nw];
    if (threadIdx.x == 0) off[ncols] = total;
    __syncthreads();  // off[] must be read before the fill advances cnt[]
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    for (uint32_t base = b0; base < b1; base += 64u) {
        const uint32_t i  = base + lane;
        const bool     ok = i < b1;
        const uint32_t k  = ok ? key(i) : 0u;
        uint64_t       mk = __ballot(ok);
        for (uint32_t bit = 0; bit < key_bits; ++bit) {
            const bool     on = (k >> bit) & 1u;
            const uint64_t bb = __ballot(on);
            mk &= on ? bb : ~bb;
        }
        const uint32_t rank = __popcll(mk & lt), cntg = __popcll(mk);
        if (ok) {
            const uint32_t sb = cnt[k * nw + w];
            val[sb + rank]    = static_cast<uint16_t>(i / W);
            if (rank + 1u == cntg) cnt[k * nw + w] = sb + cntg;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ uint32_t bits_for(uint32_t n) { return (n <= 1u) ? 1u : 32u - __clz(n - 1u); }

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
