// LDS counting-sort primitives shared by the query kernels and the C++
// query API: compile-time-size block scan, register sorting network for
// short rows, and the histogram + scan + packed-slot fill.
#pragma once
#include "lds_core.hpp"

namespace rxm {

// Inclusive prefix sum over the 64 lanes of a wave with DPP moves: row
// shifts by 1, 2, 4, 8 within each 16-lane row (lanes without a source add
// 0), then the row totals broadcast from lanes 15 and 31 into the rows above.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
    x += __builtin_amdgcn_update_dpp(0u, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0u, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0u, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0u, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0u, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0u, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

template <uint32_t T>
__device__ __forceinline__ uint32_t scan_excl(uint32_t* s, uint32_t n, uint32_t* s_wave)
{
    constexpr uint32_t NW = T / 64u;
    const uint32_t     tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    const uint32_t     per = (n + T - 1u) / T;  // T is a power of two: a shift
    const uint32_t     b = min(n, tid * per), e = min(n, b + per);
    __syncthreads();
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; ++i) sum += s[i];
    const uint32_t x = wave_incl_scan(sum);
    if (lane == 63u) s_wave[wid] = x;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) {
        const uint32_t t = s_wave[w];
        before += (w < wid) ? t : 0u;
        total += t;
    }
    uint32_t run = before + x - sum;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t t = s[i];
        s[i]             = run;
        run += t;
    }
    __syncthreads();
    return total;
}

__device__ __forceinline__ void cswap(uint32_t& a, uint32_t& b)
{
    const uint32_t lo = min(a, b), hi = max(a, b);
    a = lo;
    b = hi;
}

// Sort the k packed slots of one row ascending. Rows of up to 8 entries are
// sorted in registers with Batcher's odd-even merge network (19
// comparators); longer rows fall back to an insertion sort in LDS. The
// callback out(j, slot) receives the sorted entries.
template <typename Out>
__device__ __forceinline__ void sort_row(uint32_t* slots, uint32_t k, Out out)
{
    if (k <= 8u) {
        uint32_t a[8];
#pragma unroll
        for (uint32_t j = 0; j < 8u; ++j) a[j] = (j < k) ? slots[j] : 0xFFFFFFFFu;
        cswap(a[0], a[1]); cswap(a[2], a[3]); cswap(a[4], a[5]); cswap(a[6], a[7]);
        cswap(a[0], a[2]); cswap(a[1], a[3]); cswap(a[4], a[6]); cswap(a[5], a[7]);
        cswap(a[1], a[2]); cswap(a[5], a[6]);
        cswap(a[0], a[4]); cswap(a[1], a[5]); cswap(a[2], a[6]); cswap(a[3], a[7]);
        cswap(a[2], a[4]); cswap(a[3], a[5]);
        cswap(a[1], a[2]); cswap(a[3], a[4]); cswap(a[5], a[6]);
#pragma unroll
        for (uint32_t j = 0; j < 8u; ++j)
            if (j < k) out(j, a[j]);
    } else {
        for (uint32_t i = 1; i < k; ++i) {
            const uint32_t x = slots[i];
            uint32_t       j = i;
            while (j > 0 && slots[j - 1] > x) {
                slots[j] = slots[j - 1];
                --j;
            }
            slots[j] = x;
        }
        for (uint32_t j = 0; j < k; ++j) out(j, slots[j]);
    }
}

// Histogram + scan + packed slot fill of n entries into ncols columns.
// key(i) -> column, pack(i) -> 32-bit slot value (sort key in the high bits).
// Entries whose column is >= ncols are dropped, so a caller that outputs
// only owned rows (owned ids are the low local ids) sorts only those.
// Counts of column c accumulate in off[c + 1]; the exclusive scan turns them
// into column starts, and the fill advances off[c + 1] from the start of c to
// the start of c + 1, so no cursor array is needed. On return off[c] is the
// start of column c and off[ncols] the total.
// When n <= R * T the keys and slot values of a thread's entries are
// computed once and held in registers between the two passes.
// ZEROED: the caller already cleared off[0..ncols] before a barrier (saves
// this function's clearing pass and its barrier).
template <uint32_t T, bool ZEROED = false, uint32_t R = 12, typename KeyF, typename PackF>
__device__ __forceinline__ void fill_columns(uint32_t n, uint32_t ncols, KeyF key, PackF pack, uint32_t* off,
                                             uint32_t* slots, uint32_t* s_wave)
{
    if constexpr (!ZEROED) {
        for (uint32_t c = threadIdx.x; c <= ncols; c += T) off[c] = 0;
        __syncthreads();
    }
    if (n <= R * T) {
        uint32_t kr[R], pr[R];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            const uint32_t i = threadIdx.x + r * T;
            kr[r]            = (i < n) ? key(i) : ncols;
            if (kr[r] < ncols) {
                pr[r] = pack(i);
                atomicAdd(&off[kr[r] + 1u], 1u);
            }
        }
        scan_excl<T>(off + 1, ncols, s_wave);
#pragma unroll
        for (uint32_t r = 0; r < R; ++r)
            if (kr[r] < ncols) slots[atomicAdd(&off[kr[r] + 1u], 1u)] = pr[r];
    } else {
        for (uint32_t i = threadIdx.x; i < n; i += T) {
            const uint32_t k = key(i);
            if (k < ncols) atomicAdd(&off[k + 1u], 1u);
        }
        scan_excl<T>(off + 1, ncols, s_wave);
        for (uint32_t i = threadIdx.x; i < n; i += T) {
            const uint32_t k = key(i);
            if (k < ncols) slots[atomicAdd(&off[k + 1u], 1u)] = pack(i);
        }
    }
    __syncthreads();
}

}  // namespace rxm
