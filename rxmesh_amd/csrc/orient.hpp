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

}  // namespace rxm
