// Canonical counter-clockwise ordering of an owned vertex's edge row, used
// by the oriented VV / VE queries of the C-ABI and of the C++ query API.
#pragma once
#include "lds_core.hpp"

namespace rxm {

// ---- canonical CCW one-ring of an owned vertex --------------------------
// The reference orders a vertex's edges by walking "next = third vertex of
// the face holding half-edge v->r" (rxmesh_queries.cuh:376-499) from a start
// that depends on atomic order. The start is fixed here: among the unvisited
// edges prefer a boundary edge from which the walk can proceed, otherwise
// take the neighbour with the smallest global id; after a dead end the same
// rule restarts over the remaining edges. Reorders val[k0..k1) in place.
__device__ inline void orient_row(uint32_t v, uint32_t k0, uint32_t k1, uint16_t* val,
                                  const uint32_t* s_ev, const uint16_t* s_fe, const uint32_t* s_ef,
                                  const uint32_t* s_gid)
{
    auto other = [&](uint32_t e) {
        const uint32_t x = s_ev[e];
        return (lo16(x) == v) ? hi16(x) : lo16(x);
    };
    auto out_face = [&](uint32_t e) -> uint32_t {
        const uint32_t r = other(e);
        for (uint32_t s = 0; s < 2; ++s) {
            const uint32_t f = s_ef[2 * e + s];
            if (f == 0xFFFFFFFFu) continue;
            for (uint32_t i = 0; i < 3; ++i)
                if (face_vertex(s_fe, s_ev, f, i) == v &&
                    face_vertex(s_fe, s_ev, f, (i + 1) % 3) == r)
                    return f;
        }
        return 0xFFFFFFFFu;
    };
    auto startable = [&](uint32_t e) {
        return s_ef[2 * e + 1] == 0xFFFFFFFFu && out_face(e) != 0xFFFFFFFFu;
    };
    uint32_t pos = k0;
    while (pos < k1) {
        uint32_t best = pos;
        bool     bb   = startable(val[pos]);
        uint32_t bg   = s_gid[other(val[pos])];
        for (uint32_t j = pos + 1; j < k1; ++j) {
            const bool     b = startable(val[j]);
            const uint32_t g = s_gid[other(val[j])];
            if ((b && !bb) || (b == bb && g < bg)) {
                best = j;
                bb   = b;
                bg   = g;
            }
        }
        uint16_t t = val[pos];
        val[pos]   = val[best];
        val[best]  = t;
        while (true) {
            const uint32_t e = val[pos];
            const uint32_t f = out_face(e);
            if (f == 0xFFFFFFFFu) break;
            const uint32_t r = other(e);
            uint32_t       q = 0xFFFFFFFFu;
            for (uint32_t i = 0; i < 3; ++i) {
                const uint32_t w = face_vertex(s_fe, s_ev, f, i);
                if (w != v && w != r) q = w;
            }
            uint32_t j = pos + 1;
            while (j < k1 && other(val[j]) != q) ++j;
            if (j >= k1) break;
            t            = val[pos + 1];
            val[pos + 1] = val[j];
            val[j]       = t;
            ++pos;
        }
        ++pos;
    }
}

// Same ordering computed in registers: each edge's successor (the third
// vertex of the face holding half-edge v->r) and its startable flag are
// gathered once (independent LDS reads), then the start selection and the
// walk are compares over register arrays. The ordering does not depend on
// the order val[k0..k1) comes in. Rows longer than KMAX use orient_row.
template <uint32_t KMAX>
__device__ inline void orient_row_fast(uint32_t v, uint32_t k0, uint32_t k1, uint16_t* val, const uint32_t* s_ev,
                                       const uint16_t* s_fe, const uint32_t* s_ef, const uint32_t* s_gid)
{
    const uint32_t K = k1 - k0;
    if (K > KMAX) {
        orient_row(v, k0, k1, val, s_ev, s_fe, s_ef, s_gid);
        return;
    }
    uint32_t er[KMAX];  // edge | other endpoint << 16
    uint32_t sq[KMAX];  // successor vertex (0xFFFF: none) | startable << 16
#pragma unroll
    for (uint32_t i = 0; i < KMAX; ++i) {
        er[i] = 0;
        sq[i] = 0xFFFFu;
        if (i < K) {
            const uint32_t e = val[k0 + i], x = s_ev[e];
            const uint32_t r = (lo16(x) == v) ? hi16(x) : lo16(x);
            uint32_t       q = 0xFFFFu;
            for (uint32_t t = 0; t < 2; ++t) {
                const uint32_t f = s_ef[2 * e + t];
                if (f == 0xFFFFFFFFu || q != 0xFFFFu) continue;
                const uint32_t a = face_vertex(s_fe, s_ev, f, 0), b = face_vertex(s_fe, s_ev, f, 1),
                               c = face_vertex(s_fe, s_ev, f, 2);
                if (a == v && b == r) q = c;
                else if (b == v && c == r) q = a;
                else if (c == v && a == r) q = b;
            }
            const bool boundary = s_ef[2 * e + 1] == 0xFFFFFFFFu;
            er[i]               = e | (r << 16);
            sq[i]               = q | ((boundary && q != 0xFFFFu) ? 0x10000u : 0u);
        }
    }
    uint32_t visited = 0, pos = 0;
    while (pos < K) {
        // start: a startable edge first, then the smallest neighbour global id
        uint32_t cur  = 0;
        uint64_t best = ~0ull;
#pragma unroll
        for (uint32_t i = 0; i < KMAX; ++i) {
            if (i < K && !((visited >> i) & 1u)) {
                const uint64_t key = (uint64_t((sq[i] >> 16) ^ 1u) << 32) | s_gid[er[i] >> 16];
                if (key < best) {
                    best = key;
                    cur  = i;
                }
            }
        }
        while (true) {
            uint32_t ec = 0, qc = 0xFFFFu;
#pragma unroll
            for (uint32_t i = 0; i < KMAX; ++i) {
                if (i == cur) {
                    ec = er[i] & 0xFFFFu;
                    qc = sq[i] & 0xFFFFu;
                }
            }
            visited |= 1u << cur;
            val[k0 + pos++] = static_cast<uint16_t>(ec);
            if (qc == 0xFFFFu) break;
            uint32_t nxt = KMAX;
#pragma unroll
            for (uint32_t i = 0; i < KMAX; ++i)
                if (i < K && !((visited >> i) & 1u) && (er[i] >> 16) == qc) nxt = i;
            if (nxt == KMAX) break;
            cur = nxt;
        }
    }
}

}  // namespace rxm
