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
template <typename SlotT, typename Out>
__device__ __forceinline__ void sort_row(SlotT* slots, uint32_t k, Out out)
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
            const SlotT x = slots[i];
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
// SlotT is uint32_t, or uint16_t where the slot value fits 16 bits (less LDS).
template <uint32_t T, bool ZEROED = false, uint32_t R = 12, typename KeyF, typename PackF, typename SlotT>
__device__ __forceinline__ void fill_columns(uint32_t n, uint32_t ncols, KeyF key, PackF pack, uint32_t* off,
                                             SlotT* slots, uint32_t* s_wave)
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
            if (kr[r] < ncols) slots[atomicAdd(&off[kr[r] + 1u], 1u)] = static_cast<SlotT>(pr[r]);
    } else {
        for (uint32_t i = threadIdx.x; i < n; i += T) {
            const uint32_t k = key(i);
            if (k < ncols) atomicAdd(&off[k + 1u], 1u);
        }
        scan_excl<T>(off + 1, ncols, s_wave);
        for (uint32_t i = threadIdx.x; i < n; i += T) {
            const uint32_t k = key(i);
            if (k < ncols) slots[atomicAdd(&off[k + 1u], 1u)] = static_cast<SlotT>(pack(i));
        }
    }
    __syncthreads();
}

// ---- atomic-free stable counting sort ------------------------------------
// Bits needed to tell keys 0 .. ncols - 1 apart.
__device__ __forceinline__ uint32_t key_bits(uint32_t ncols)
{
    return ncols > 1u ? 32u - __builtin_clz(ncols - 1u) : 0u;
}

// Lanes of this wave whose key equals this lane's key, among the lanes in
// `valid`: one ballot per key bit. Wave-uniform control flow required.
__device__ __forceinline__ uint64_t match_key(uint32_t k, uint64_t valid, uint32_t nbits)
{
    uint64_t m = valid;
    for (uint32_t b = 0; b < nbits; ++b) {
        const bool     bit = (k >> b) & 1u;
        const uint64_t bb  = __builtin_amdgcn_ballot_w64(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

// Number of set bits of m below this lane (mbcnt).
__device__ __forceinline__ uint32_t rank_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// Stable counting sort of n entries into ncols columns with no LDS atomics
// and no per-row sort: column c receives its entries in ascending entry
// index i. Wave w of the NW = T / 64 waves owns the contiguous entry block
// [w C, (w + 1) C) and walks it in rounds of 64; in each round the lanes
// holding the same key are matched with one ballot per key bit, the match's
// lowest lane adds the match size to hist[key * NW + w] (one writer per cell
// and round, so a plain read-modify-write), and every lane's rank inside its
// match comes from mbcnt. An exclusive scan of hist in key-major order then
// gives, for every (key, wave), where that wave's entries of the key start,
// and a second walk stores each entry's payload at start + rank, advancing
// the cell by the match size after each round. Entries whose key is >= ncols
// are dropped. hist holds ncols * NW u32 (cleared here); on return
// row_start(c) for c < ncols was passed to `rows(c, start)` right after the
// scan (before the cells advance), and the total is returned.
template <uint32_t T, uint32_t R = 12, typename KeyF, typename PayF, typename RowF>
__device__ __forceinline__ uint32_t sort_columns_stable(uint32_t n, uint32_t ncols, KeyF key, PayF payload,
                                                        RowF rows, uint32_t* hist, uint32_t* slots,
                                                        uint32_t* s_wave)
{
    constexpr uint32_t NW   = T / 64u;
    const uint32_t     lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t     C    = ((n + 64u * NW - 1u) / (64u * NW)) * 64u;  // entries per wave
    const uint32_t     b0 = min(n, w * C), b1 = min(n, b0 + C);
    const uint32_t     nr = (b1 - b0 + 63u) >> 6;  // rounds of this wave (wave-uniform)
    const uint32_t     nbits = key_bits(ncols);
    for (uint32_t c = threadIdx.x; c < ncols * NW; c += T) hist[c] = 0u;
    __syncthreads();
    uint64_t mk[R];
    uint32_t kk[R];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        if (r < nr) {
            const uint32_t i = b0 + (r << 6) + lane;
            const uint32_t k = (i < b1) ? key(i) : ncols;
            const uint64_t m = match_key(k, __builtin_amdgcn_ballot_w64(k < ncols), nbits);
            kk[r]            = k;
            mk[r]            = m;
            if (k < ncols && rank_below(m) == 0u) hist[k * NW + w] += static_cast<uint32_t>(__popcll(m));
        }
    }
    for (uint32_t r = R; r < nr; ++r) {
        const uint32_t i = b0 + (r << 6) + lane;
        const uint32_t k = (i < b1) ? key(i) : ncols;
        const uint64_t m = match_key(k, __builtin_amdgcn_ballot_w64(k < ncols), nbits);
        if (k < ncols && rank_below(m) == 0u) hist[k * NW + w] += static_cast<uint32_t>(__popcll(m));
    }
    const uint32_t total = scan_excl<T>(hist, ncols * NW, s_wave);
    for (uint32_t c = threadIdx.x; c < ncols; c += T) rows(c, hist[c * NW]);
    __syncthreads();  // row starts read before the cells advance
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        if (r < nr) {
            const uint32_t k = kk[r];
            const uint64_t m = mk[r];
            if (k < ncols) {
                const uint32_t rk = rank_below(m);
                const uint32_t at = hist[k * NW + w];
                slots[at + rk]    = payload(b0 + (r << 6) + lane);
                if (rk == 0u) hist[k * NW + w] = at + static_cast<uint32_t>(__popcll(m));
            }
        }
    }
    for (uint32_t r = R; r < nr; ++r) {
        const uint32_t i = b0 + (r << 6) + lane;
        const uint32_t k = (i < b1) ? key(i) : ncols;
        const uint64_t m = match_key(k, __builtin_amdgcn_ballot_w64(k < ncols), nbits);
        if (k < ncols) {
            const uint32_t rk = rank_below(m);
            const uint32_t at = hist[k * NW + w];
            slots[at + rk]    = payload(i);
            if (rk == 0u) hist[k * NW + w] = at + static_cast<uint32_t>(__popcll(m));
        }
    }
    __syncthreads();
    return total;
}

}  // namespace rxm
