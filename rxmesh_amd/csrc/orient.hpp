// Canonical counter-clockwise ordering of a vertex's edge row, used by the
// oriented VV / VE queries of the C-ABI, the cotangent assembly and the C++
// query API.
#pragma once
#include "lds_core.hpp"

namespace rxm {

// ---- half-edge successor table -------------------------------------------
// Face slot i of face f holds the half-edge fv[i] -> fv[i+1] on local edge
// e = fe >> 1, starting at the edge's (dir ? hi16 : lo16) endpoint, so the
// FE entry itself, 2e + dir, names the half-edge. hn[2e + dir] <- the third
// vertex of that face (fv[i+2], the start of slot i + 2), or kInv16 when no
// face of the patch holds the half-edge. On a consistently oriented
// edge-manifold mesh a half-edge has at most one face, so the stores of
// different slots never collide. hn must hold 2 * ne entries set to kInv16
// (and be visible to the block) before the call; the caller syncs after it.
__device__ __forceinline__ void build_half_next(const uint16_t* __restrict__ s_fe,
                                                const uint32_t* __restrict__ s_ev, uint32_t nf,
                                                uint16_t* __restrict__ hn)
{
    for (uint32_t i = threadIdx.x; i < 3u * nf; i += blockDim.x) {
        const uint32_t f = i / 3u, j = i - 3u * f;
        const uint32_t y  = s_fe[3u * f + ((j == 0u) ? 2u : j - 1u)];  // slot i + 2
        const uint32_t ey = s_ev[y >> 1];
        hn[s_fe[i]]       = static_cast<uint16_t>((y & 1u) ? hi16(ey) : lo16(ey));
    }
}

// ---- canonical CCW one-ring ------------------------------------------------
// The reference orders a vertex's edges by walking "next = third vertex of
// the face holding half-edge v->r" (rxmesh_queries.cuh:376-499) from a start
// that depends on atomic order. The start is fixed here: among the unvisited
// edges prefer a boundary edge from which the walk can proceed (its only
// face holds v->r), otherwise take the neighbour with the smallest global
// id; after a dead end the same rule restarts over the remaining edges. The
// result does not depend on the order the row's edges come in.
//
// src[k0..k1) holds the row's local edge ids; dst[k0..k1) receives the
// ordered row as neighbour vertices (NBR) or edges. dst may alias src.

// Rows longer than the register path: selection and walk over LDS.
template <bool NBR, typename ST, typename DT>
__device__ __forceinline__ void orient_ring_lds(uint32_t v, uint32_t k0, uint32_t k1, ST* src, DT* dst,
                                       const uint32_t* s_ev, const uint16_t* s_hn, const uint32_t* s_gid)
{
    auto other = [&](uint32_t e) {
        const uint32_t x = s_ev[e];
        return (lo16(x) == v) ? hi16(x) : lo16(x);
    };
    auto half = [&](uint32_t e) { return 2u * e + ((lo16(s_ev[e]) == v) ? 0u : 1u); };
    auto succ = [&](uint32_t e) { return static_cast<uint32_t>(s_hn[half(e)]); };
    auto startable = [&](uint32_t e) {
        const uint32_t h = half(e);
        return s_hn[h] != kInv16 && s_hn[h ^ 1u] == kInv16;
    };
    uint32_t pos = k0;
    while (pos < k1) {
        uint32_t best = pos;
        bool     bb   = startable(src[pos]);
        uint32_t bg   = s_gid[other(src[pos])];
        for (uint32_t j = pos + 1; j < k1; ++j) {
            const bool     b = startable(src[j]);
            const uint32_t g = s_gid[other(src[j])];
            if ((b && !bb) || (b == bb && g < bg)) {
                best = j;
                bb   = b;
                bg   = g;
            }
        }
        ST t      = src[pos];
        src[pos]  = src[best];
        src[best] = t;
        while (true) {
            const uint32_t q = succ(src[pos]);
            if (q == kInv16) break;
            uint32_t j = pos + 1;
            while (j < k1 && other(src[j]) != q) ++j;
            if (j >= k1) break;
            t            = src[pos + 1];
            src[pos + 1] = src[j];
            src[j]       = t;
            ++pos;
        }
        ++pos;
    }
    for (uint32_t k = k0; k < k1; ++k) {
        const uint32_t e = src[k];
        dst[k]           = static_cast<DT>(NBR ? other(e) : e);
    }
}

// Register path: each edge's neighbour, successor and startable flag are
// gathered with independent LDS reads (EV, then the two half-edges of the
// successor table), the first start is chosen during the gather, and the
// walk is compares over the register arrays.
template <uint32_t KMAX, bool NBR, typename ST, typename DT>
__device__ __forceinline__ void orient_ring(uint32_t v, uint32_t k0, uint32_t k1, ST* src, DT* dst, const uint32_t* s_ev,
                                   const uint16_t* s_hn, const uint32_t* s_gid)
{
    const uint32_t K = k1 - k0;
    if (K == 0u) return;
    if (K > KMAX) {
        orient_ring_lds<NBR>(v, k0, k1, src, dst, s_ev, s_hn, s_gid);
        return;
    }
    // er: output id (neighbour if NBR, else edge) in the low half, the other
    // one in the high half; sq: successor vertex (kInv16: none) | startable << 16
    uint32_t er[KMAX];
    uint32_t sq[KMAX];
    auto     nbr_of = [](uint32_t x) { return NBR ? (x & 0xFFFFu) : (x >> 16); };
    uint32_t cur  = 0;
    uint64_t best = ~0ull;
#pragma unroll
    for (uint32_t i = 0; i < KMAX; ++i) {
        er[i] = 0;
        sq[i] = kInv16;
        if (i < K) {
            const uint32_t e = src[k0 + i], x = s_ev[e];
            const bool     lo = lo16(x) == v;
            const uint32_t r  = lo ? hi16(x) : lo16(x);
            const uint32_t h  = 2u * e + (lo ? 0u : 1u);
            const uint32_t q = s_hn[h], back = s_hn[h ^ 1u];
            const uint32_t st = (q != kInv16 && back == kInv16) ? 1u : 0u;
            er[i]             = NBR ? (r | (e << 16)) : (e | (r << 16));
            sq[i]             = q | (st << 16);
            const uint64_t key = (uint64_t(st ^ 1u) << 32) | s_gid[r];
            if (key < best) {
                best = key;
                cur  = i;
            }
        }
    }
    uint32_t visited = 0, pos = 0;
    while (true) {
        while (true) {
            uint32_t ec = 0, qc = kInv16;
#pragma unroll
            for (uint32_t i = 0; i < KMAX; ++i) {
                if (i == cur) {
                    ec = er[i];
                    qc = sq[i] & 0xFFFFu;
                }
            }
            visited |= 1u << cur;
            dst[k0 + pos++] = static_cast<DT>(ec & 0xFFFFu);
            if (qc == kInv16) break;
            uint32_t nxt = KMAX;
#pragma unroll
            for (uint32_t i = 0; i < KMAX; ++i)
                if (i < K && !((visited >> i) & 1u) && nbr_of(er[i]) == qc) nxt = i;
            if (nxt == KMAX) break;
            cur = nxt;
        }
        if (pos >= K) break;
        // restart after a dead end: same rule over the unvisited edges
        best = ~0ull;
#pragma unroll
        for (uint32_t i = 0; i < KMAX; ++i) {
            if (i < K && !((visited >> i) & 1u)) {
                const uint64_t key = (uint64_t((sq[i] >> 16) ^ 1u) << 32) | s_gid[nbr_of(er[i])];
                if (key < best) {
                    best = key;
                    cur  = i;
                }
            }
        }
    }
}

}  // namespace rxm
