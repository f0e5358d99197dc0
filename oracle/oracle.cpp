// Oracle: CPU restatement of RXMesh's static patch query path, in FP64.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/, the smoke()
// entry and the cpu_baseline leg of bench.py load it. The product (rxmesh_amd)
// never links, loads or calls it.
//
// Each function cites the reference file:line whose behaviour it restates
// (paths relative to the RXMesh source tree). It is written for clarity, not
// speed, and deliberately shares no code with the product's host build.
//
// Parity pins (see tests/test_oracle.py):
//   edge ids / EV / FF ......... reference patch archives' num_edges, Euler, brute force
//   ribbons / ownership ........ reference archives input/{sphere3,torus}_patches (exact)
//   ltog / EV / FE encodings ... restated run_ltog_mapping_test invariants
//   queries .................... brute-force global sets (reference verifier semantics)
//   oriented VV ................ reference known answers (plane 90/180 deg, cube 45/90 deg)
//   normals .................... hand-computed single triangle / flat grid closed form
//   smoothing .................. constants are fixed points; linear fields preserved inside
//   cot CSR / MCF .............. row sums of L = 0, symmetry, flat-grid closed form
//   symmetric Dirichlet ........ J=I => E=4A, grad=0; J=diag(2,1) => 6.25A; FD;
//                                forward-mode dual restatement
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <numeric>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

namespace {

constexpr uint32_t INV32 = 0xFFFFFFFFu;
constexpr uint16_t INV16 = 0xFFFFu;

// reference: include/rxmesh/util/util.h:410-416 (edge_key returns (max,min))
inline std::pair<uint32_t, uint32_t> edge_key(uint32_t a, uint32_t b)
{
    return {std::max(a, b), std::min(a, b)};
}

struct Patch
{
    std::vector<uint32_t> ltog_v, ltog_e, ltog_f;
    uint32_t              nvo = 0, neo = 0, nfo = 0;
    std::vector<uint16_t> ev;  // 2 per local edge
    std::vector<uint16_t> fe;  // 3 per local face
};

struct Mesh
{
    uint32_t                                 nV = 0, nE = 0, nF = 0, P = 0;
    std::vector<std::array<uint32_t, 3>>     fv;
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> edges_map;
    std::vector<std::array<uint32_t, 2>>     ev;
    std::vector<uint32_t>                    ff_offset, ff_values;
    std::vector<std::vector<uint32_t>>       edge_faces;
    bool                                     closed = true, manifold = true;
    uint32_t max_valence = 0, max_edge_faces = 0, max_face_adj = 0;

    std::vector<uint32_t> face_patch, patches_val, patches_offset;
    std::vector<uint32_t> ribbon_val, ribbon_offset;
    std::vector<uint32_t> vertex_patch, edge_patch;
    std::vector<Patch>    patches;
    std::vector<uint32_t> v_prefix, e_prefix, f_prefix;
    // global id -> linear id (prefix[owner] + index in owner's owned part)
    std::vector<uint32_t> v_g2l, e_g2l, f_g2l;
    std::vector<uint32_t> v_l2g, e_l2g, f_l2g;

    uint32_t edge_id(uint32_t a, uint32_t b) const
    {
        return edges_map.at(edge_key(a, b));
    }
};

// reference: include/rxmesh/rxmesh.cpp:518-681 (build_supporting_structures)
void build_supporting(Mesh& m)
{
    m.nF = static_cast<uint32_t>(m.fv.size());
    m.nV = 0;
    m.nE = 0;
    m.ev.clear();
    m.edge_faces.clear();
    std::vector<uint32_t> ff_size(m.nF, 0);
    std::vector<uint32_t> vv_count;
    uint32_t              open_edges = 0;
    auto add_val = [&](uint32_t v) {
        if (v >= vv_count.size()) vv_count.resize(v + 1, 0);
        m.max_valence = std::max(m.max_valence, ++vv_count[v]);
    };
    for (uint32_t f = 0; f < m.nF; ++f) {
        for (uint32_t v = 0; v < 3; ++v) {
            uint32_t v0 = m.fv[f][v];
            uint32_t v1 = m.fv[f][(v + 1) % 3];
            m.nV        = std::max(m.nV, std::max(v0, v1));
            auto key    = edge_key(v0, v1);
            auto it     = m.edges_map.find(key);
            if (it == m.edges_map.end()) {
                m.edges_map.emplace(key, m.nE);
                ++m.nE;
                m.ev.push_back({v0, v1});
                m.edge_faces.push_back({f});
                ++open_edges;
                m.max_edge_faces = std::max(m.max_edge_faces, 1u);
                add_val(v0);
                add_val(v1);
            } else {
                uint32_t eid = it->second;
                uint32_t inc = static_cast<uint32_t>(m.edge_faces[eid].size());
                for (uint32_t k = 0; k < inc; ++k) ++ff_size[m.edge_faces[eid][k]];
                ff_size[f] += inc;
                m.edge_faces[eid].push_back(f);
                if (inc == 1) --open_edges;
                if (inc + 1 > 2) m.manifold = false;
                m.max_edge_faces = std::max(m.max_edge_faces, inc + 1);
            }
        }
    }
    ++m.nV;
    m.closed = (open_edges == 0);
    m.ff_offset.assign(m.nF + 1, 0);
    uint32_t c = 0;
    for (uint32_t f = 0; f < m.nF; ++f) {
        m.ff_offset[f] = c;
        c += ff_size[f];
        m.max_face_adj = std::max(m.max_face_adj, ff_size[f]);
    }
    m.ff_offset[m.nF] = c;
    m.ff_values.assign(c, 0);
    std::fill(ff_size.begin(), ff_size.end(), 0);
    for (uint32_t e = 0; e < m.nE; ++e) {
        const auto& ef = m.edge_faces[e];
        for (size_t i = 0; i < ef.size(); ++i)
            for (size_t j = i + 1; j < ef.size(); ++j) {
                uint32_t f0 = ef[i], f1 = ef[j];
                m.ff_values[m.ff_offset[f0] + ff_size[f0]++] = f1;
                m.ff_values[m.ff_offset[f1] + ff_size[f1]++] = f0;
            }
    }
}

// reference: include/rxmesh/patcher/patcher.cu:640-717 (extract_ribbons)
void extract_ribbons(Mesh& m)
{
    std::vector<std::vector<uint32_t>> vf(m.nV);
    for (uint32_t f = 0; f < m.nF; ++f)
        for (uint32_t i = 0; i < 3; ++i) vf[m.fv[f][i]].push_back(f);
    m.ribbon_val.clear();
    m.ribbon_offset.assign(m.P, 0);
    for (uint32_t p = 0; p < m.P; ++p) {
        uint32_t ps = (p == 0) ? 0 : m.patches_offset[p - 1];
        uint32_t pe = m.patches_offset[p];
        std::vector<uint32_t>        ribbon;
        std::unordered_set<uint32_t> seen;
        for (uint32_t k = ps; k < pe; ++k) {
            uint32_t face = m.patches_val[k];
            for (uint32_t i = 0; i < 3; ++i) {
                for (uint32_t n : vf[m.fv[face][i]]) {
                    if (m.face_patch[n] != p && seen.insert(n).second)
                        ribbon.push_back(n);
                }
            }
        }
        m.ribbon_val.insert(m.ribbon_val.end(), ribbon.begin(), ribbon.end());
        m.ribbon_offset[p] = static_cast<uint32_t>(m.ribbon_val.size());
    }
}

// reference: include/rxmesh/patcher/patcher.cu:719-757 (assign_patch)
void assign_patch(Mesh& m)
{
    m.vertex_patch.assign(m.nV, INV32);
    m.edge_patch.assign(m.nE, INV32);
    for (uint32_t p = 0; p < m.P; ++p) {
        uint32_t ps = (p == 0) ? 0 : m.patches_offset[p - 1];
        uint32_t pe = m.patches_offset[p];
        for (uint32_t k = ps; k < pe; ++k) {
            uint32_t face = m.patches_val[k];
            uint32_t v1   = m.fv[face][2];
            for (uint32_t v = 0; v < 3; ++v) {
                uint32_t v0 = m.fv[face][v];
                uint32_t e  = m.edge_id(v0, v1);
                if (m.vertex_patch[v0] == INV32) m.vertex_patch[v0] = p;
                if (m.edge_patch[e] == INV32) m.edge_patch[e] = p;
                v1 = v0;
            }
        }
    }
}

// Owned edges / vertices of every patch in ascending id order (bucketed once,
// so the safeguard loops of build_ltog visit exactly the elements the
// reference's full scans over all edges / vertices would accept, in the same
// order, without O(P (V + E)) work).
struct OwnedLists
{
    std::vector<uint32_t> e_off, e_val, v_off, v_val;
};
static void bucket(const std::vector<uint32_t>& owner, uint32_t P, std::vector<uint32_t>& off,
                   std::vector<uint32_t>& val)
{
    off.assign(P + 1, 0);
    for (uint32_t o : owner)
        if (o < P) off[o + 1]++;
    for (uint32_t p = 0; p < P; ++p) off[p + 1] += off[p];
    val.assign(off[P], 0);
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    for (uint32_t i = 0; i < owner.size(); ++i)
        if (owner[i] < P) val[cur[owner[i]]++] = i;
}

// reference: include/rxmesh/rxmesh.cpp:754-870 (build_single_patch_ltog)
void build_ltog(Mesh& m, uint32_t p, const OwnedLists& own)
{
    Patch&   P  = m.patches[p];
    uint32_t ps = (p == 0) ? 0 : m.patches_offset[p - 1];
    uint32_t pe = m.patches_offset[p];
    uint32_t rs = (p == 0) ? 0 : m.ribbon_offset[p - 1];
    uint32_t re = m.ribbon_offset[p];
    std::unordered_set<uint32_t> vadded, eadded;
    P.ltog_f.clear();
    P.ltog_v.clear();
    P.ltog_e.clear();
    auto add_face = [&](uint32_t f) {
        P.ltog_f.push_back(f);
        for (uint32_t v = 0; v < 3; ++v) {
            uint32_t v0 = m.fv[f][v], v1 = m.fv[f][(v + 1) % 3];
            uint32_t e  = m.edge_id(v0, v1);
            if (vadded.insert(v0).second) P.ltog_v.push_back(v0);
            if (eadded.insert(e).second) P.ltog_e.push_back(e);
        }
    };
    for (uint32_t k = ps; k < pe; ++k) add_face(m.patches_val[k]);
    for (uint32_t k = rs; k < re; ++k) add_face(m.ribbon_val[k]);
    // Safeguard loops of the reference (owned elements not reached through a
    // face). With the reference's ownership rule they never add anything.
    for (uint32_t k = own.e_off[p]; k < own.e_off[p + 1]; ++k) {
        const uint32_t e = own.e_val[k];  // edge_patch[e] == p, ascending e
        if (eadded.insert(e).second) {
            P.ltog_e.push_back(e);
            for (uint32_t i = 0; i < 2; ++i)
                if (!vadded.count(m.ev[e][i])) P.ltog_v.push_back(m.ev[e][i]);
        }
    }
    for (uint32_t k = own.v_off[p]; k < own.v_off[p + 1]; ++k) {
        const uint32_t v = own.v_val[k];  // vertex_patch[v] == p, ascending v
        if (vadded.insert(v).second) P.ltog_v.push_back(v);
    }

    auto uniq = [&](std::vector<uint32_t>& l, const std::vector<uint32_t>& owner) {
        std::sort(l.begin(), l.end());
        auto it = std::stable_partition(
            l.begin(), l.end(), [&](uint32_t i) { return owner[i] == p; });
        return static_cast<uint32_t>(it - l.begin());
    };
    P.nfo = uniq(P.ltog_f, m.face_patch);
    P.neo = uniq(P.ltog_e, m.edge_patch);
    P.nvo = uniq(P.ltog_v, m.vertex_patch);
}

// reference: include/rxmesh/rxmesh.cpp:872-996 (build_single_patch_topology)
void build_topology(Mesh& m, uint32_t p)
{
    Patch& P = m.patches[p];
    auto find_local = [&](uint32_t gid, uint32_t owner, uint32_t nown,
                          const std::vector<uint32_t>& ltog) -> uint16_t {
        size_t s = 0, e = nown;
        if (owner != p) {
            s = nown;
            e = ltog.size();
        }
        auto it = std::lower_bound(ltog.begin() + s, ltog.begin() + e, gid);
        if (it == ltog.begin() + e || *it != gid) return INV16;
        return static_cast<uint16_t>(it - ltog.begin());
    };
    P.ev.assign(2 * P.ltog_e.size(), INV16);
    P.fe.assign(3 * P.ltog_f.size(), INV16);
    for (uint32_t f : P.ltog_f) {
        uint16_t lf = find_local(f, m.face_patch[f], P.nfo, P.ltog_f);
        for (uint32_t v = 0; v < 3; ++v) {
            uint32_t g0 = m.fv[f][v], g1 = m.fv[f][(v + 1) % 3];
            auto     key = edge_key(g0, g1);
            uint32_t dir = (key.first == g0) ? 0u : 1u;
            uint32_t ge  = m.edges_map.at(key);
            uint16_t le  = find_local(ge, m.edge_patch[ge], P.neo, P.ltog_e);
            P.ev[2 * le]     = find_local(key.first, m.vertex_patch[key.first], P.nvo, P.ltog_v);
            P.ev[2 * le + 1] = find_local(key.second, m.vertex_patch[key.second], P.nvo, P.ltog_v);
            P.fe[3 * lf + v] = static_cast<uint16_t>((le << 1) | dir);
        }
    }
}

void build_linear(Mesh& m)
{
    m.v_prefix.assign(m.P + 1, 0);
    m.e_prefix.assign(m.P + 1, 0);
    m.f_prefix.assign(m.P + 1, 0);
    for (uint32_t p = 0; p < m.P; ++p) {
        m.v_prefix[p + 1] = m.v_prefix[p] + m.patches[p].nvo;
        m.e_prefix[p + 1] = m.e_prefix[p] + m.patches[p].neo;
        m.f_prefix[p + 1] = m.f_prefix[p] + m.patches[p].nfo;
    }
    m.v_g2l.assign(m.nV, INV32);
    m.e_g2l.assign(m.nE, INV32);
    m.f_g2l.assign(m.nF, INV32);
    m.v_l2g.assign(m.v_prefix[m.P], INV32);
    m.e_l2g.assign(m.e_prefix[m.P], INV32);
    m.f_l2g.assign(m.f_prefix[m.P], INV32);
    for (uint32_t p = 0; p < m.P; ++p) {
        const Patch& P = m.patches[p];
        for (uint32_t l = 0; l < P.nvo; ++l) {
            m.v_g2l[P.ltog_v[l]]      = m.v_prefix[p] + l;
            m.v_l2g[m.v_prefix[p] + l] = P.ltog_v[l];
        }
        for (uint32_t l = 0; l < P.neo; ++l) {
            m.e_g2l[P.ltog_e[l]]      = m.e_prefix[p] + l;
            m.e_l2g[m.e_prefix[p] + l] = P.ltog_e[l];
        }
        for (uint32_t l = 0; l < P.nfo; ++l) {
            m.f_g2l[P.ltog_f[l]]      = m.f_prefix[p] + l;
            m.f_l2g[m.f_prefix[p] + l] = P.ltog_f[l];
        }
    }
}

// ---------------------------------------------------------------------------
// Queries. Reference: include/rxmesh/kernels/rxmesh_queries.cuh:502-858.
// The reference fills transposed rows in atomic order; a serial schedule of
// the same fill gives rows in ascending source-local order, which is the
// order restated here (stable counting sort).
// Op numbering follows include/rxmesh/types.h:113-129.
enum OpId { OP_VV = 3, OP_VE = 4, OP_VF = 5, OP_FV = 6, OP_FE = 7, OP_FF = 8, OP_EV = 9, OP_EF = 11, OP_EVD = 12 };

// stable transpose of a fixed-width row matrix (num_rows x width) of column ids
void stable_transpose(uint32_t num_rows, uint32_t width, uint32_t num_cols,
                      const std::vector<uint16_t>& mat,
                      std::vector<uint32_t>& off, std::vector<uint16_t>& val)
{
    off.assign(num_cols + 1, 0);
    for (uint32_t i = 0; i < num_rows * width; ++i) off[mat[i] + 1]++;
    for (uint32_t c = 0; c < num_cols; ++c) off[c + 1] += off[c];
    val.assign(num_rows * width, 0);
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    for (uint32_t i = 0; i < num_rows * width; ++i)
        val[cur[mat[i]]++] = static_cast<uint16_t>(i / width);
}

struct PatchQuery
{
    std::vector<uint32_t> off;  // per owned source row (size n_owned_src + 1)
    std::vector<uint32_t> val;  // target LOCAL ids
};

// target local -> linear id
uint32_t to_linear(const Mesh& m, uint32_t p, int kind, uint32_t l)
{
    const Patch& P = m.patches[p];
    if (kind == 0) return m.v_g2l[P.ltog_v[l]];
    if (kind == 1) return m.e_g2l[P.ltog_e[l]];
    return m.f_g2l[P.ltog_f[l]];
}

int src_kind(int op)
{
    switch (op) {
        case OP_VV: case OP_VE: case OP_VF: return 0;
        case OP_EV: case OP_EF: case OP_EVD: return 1;
        default: return 2;
    }
}
int dst_kind(int op)
{
    switch (op) {
        case OP_VV: case OP_FV: case OP_EV: case OP_EVD: return 0;
        case OP_VE: case OP_FE: return 1;
        default: return 2;
    }
}

// Canonical CCW one-ring of owned local vertex v (reference
// rxmesh_queries.cuh:376-499 walks the same "next = third vertex of the face
// holding half-edge v->r"; its start is atomic-order dependent, so the start
// is fixed here: prefer a boundary edge from which the walk proceeds, else the
// neighbour with the smallest global id; restarts after a dead end use the
// same rule over the unvisited edges).
void orient_row(const Mesh& m, uint32_t p, uint32_t v, std::vector<uint32_t>& edges)
{
    const Patch& P   = m.patches[p];
    const uint32_t gv = P.ltog_v[v];
    auto other = [&](uint32_t le) {
        uint16_t a = P.ev[2 * le], b = P.ev[2 * le + 1];
        return (a == v) ? b : a;
    };
    auto ge_of = [&](uint32_t le) { return P.ltog_e[le]; };
    // face (global) with half-edge gv -> gr on edge
    auto out_face = [&](uint32_t le) -> int64_t {
        uint32_t gr = P.ltog_v[other(le)];
        for (uint32_t f : m.edge_faces[ge_of(le)]) {
            for (int i = 0; i < 3; ++i)
                if (m.fv[f][i] == gv && m.fv[f][(i + 1) % 3] == gr) return f;
        }
        return -1;
    };
    auto is_bnd = [&](uint32_t le) { return m.edge_faces[ge_of(le)].size() < 2; };
    std::vector<uint32_t> rest = edges, out;
    auto select = [&](const std::vector<uint32_t>& c) -> size_t {
        size_t best = SIZE_MAX;
        bool   best_b = false;
        for (size_t i = 0; i < c.size(); ++i) {
            bool b = is_bnd(c[i]) && out_face(c[i]) >= 0;
            uint32_t g = P.ltog_v[other(c[i])];
            if (best == SIZE_MAX || (b && !best_b) ||
                (b == best_b && g < P.ltog_v[other(c[best])])) {
                best = i;
                best_b = b;
            }
        }
        return best;
    };
    while (!rest.empty()) {
        size_t s   = select(rest);
        uint32_t cur = rest[s];
        rest.erase(rest.begin() + s);
        out.push_back(cur);
        while (true) {
            int64_t f = out_face(cur);
            if (f < 0) break;
            uint32_t gr = P.ltog_v[other(cur)];
            uint32_t gq = INV32;
            for (int i = 0; i < 3; ++i)
                if (m.fv[f][i] != gv && m.fv[f][i] != gr) gq = m.fv[f][i];
            size_t k = SIZE_MAX;
            for (size_t i = 0; i < rest.size(); ++i)
                if (P.ltog_v[other(rest[i])] == gq) { k = i; break; }
            if (k == SIZE_MAX) break;
            cur = rest[k];
            rest.erase(rest.begin() + k);
            out.push_back(cur);
        }
    }
    edges = out;
}

PatchQuery patch_query(const Mesh& m, uint32_t p, int op, bool oriented)
{
    const Patch& P  = m.patches[p];
    const uint32_t nv = static_cast<uint32_t>(P.ltog_v.size());
    const uint32_t ne = static_cast<uint32_t>(P.ltog_e.size());
    const uint32_t nf = static_cast<uint32_t>(P.ltog_f.size());
    PatchQuery q;
    // local FV: reference f_v (rxmesh_queries.cuh:650-688): vertex = ev[2e + dir]
    std::vector<uint16_t> fvl(3 * nf);
    for (uint32_t f = 0; f < nf; ++f)
        for (int i = 0; i < 3; ++i) {
            uint16_t e = P.fe[3 * f + i];
            fvl[3 * f + i] = P.ev[2 * (e >> 1) + (e & 1)];
        }
    std::vector<uint16_t> fe_e(3 * nf);
    for (uint32_t i = 0; i < 3 * nf; ++i) fe_e[i] = P.fe[i] >> 1;

    auto fixed = [&](uint32_t nsrc, uint32_t w, const std::vector<uint16_t>& mat) {
        q.off.resize(nsrc + 1);
        for (uint32_t s = 0; s <= nsrc; ++s) q.off[s] = s * w;
        q.val.assign(mat.begin(), mat.begin() + nsrc * w);
    };
    auto trans = [&](uint32_t nsrc, uint32_t rows, uint32_t w, uint32_t cols,
                     const std::vector<uint16_t>& mat) {
        std::vector<uint32_t> off;
        std::vector<uint16_t> val;
        stable_transpose(rows, w, cols, mat, off, val);
        q.off.assign(off.begin(), off.begin() + nsrc + 1);
        q.val.assign(val.begin(), val.begin() + off[nsrc]);
    };
    switch (op) {
        case OP_EV: fixed(P.neo, 2, P.ev); break;
        case OP_EVD: {
            // reference e_v_diamond (rxmesh_queries.cuh:198-265): per edge
            // [v0, opposite through the face running v0 -> v1, v1, opposite
            // through the face running v1 -> v0]; a face slot's start vertex
            // is opposite the face's next slot, and lands in position 1 + 2 dir
            // of that slot's edge. Missing sides stay INVALID (0xFFFF); the
            // serial face order makes the last face win where two faces run an
            // edge the same way.
            std::vector<uint16_t> out(4 * ne, 0xFFFF);
            for (uint32_t e = 0; e < ne; ++e) {
                out[4 * e]     = P.ev[2 * e];
                out[4 * e + 2] = P.ev[2 * e + 1];
            }
            for (uint32_t f = 0; f < nf; ++f)
                for (uint32_t k = 0; k < 3; ++k) {
                    const uint16_t x  = P.fe[3 * f + k];
                    const uint16_t vi = out[4 * (x >> 1) + 2 * (x & 1)];
                    const uint16_t y  = P.fe[3 * f + (k + 1) % 3];
                    out[4 * (y >> 1) + 1 + 2 * (y & 1)] = vi;
                }
            fixed(P.neo, 4, out);
            break;
        }
        case OP_FV: fixed(P.nfo, 3, fvl); break;
        case OP_FE: fixed(P.nfo, 3, fe_e); break;
        case OP_VE:
        case OP_VV: {
            trans(P.nvo, ne, 2, nv, P.ev);
            if (oriented) {
                for (uint32_t v = 0; v < P.nvo; ++v) {
                    std::vector<uint32_t> row(q.val.begin() + q.off[v], q.val.begin() + q.off[v + 1]);
                    orient_row(m, p, v, row);
                    std::copy(row.begin(), row.end(), q.val.begin() + q.off[v]);
                }
            }
            if (op == OP_VV) {
                for (uint32_t v = 0; v < P.nvo; ++v)
                    for (uint32_t k = q.off[v]; k < q.off[v + 1]; ++k) {
                        uint32_t e = q.val[k];
                        q.val[k] = (P.ev[2 * e] == v) ? P.ev[2 * e + 1] : P.ev[2 * e];
                    }
            }
            break;
        }
        case OP_VF: trans(P.nvo, nf, 3, nv, fvl); break;
        case OP_EF: trans(P.neo, nf, 3, ne, fe_e); break;
        case OP_FF: {
            std::vector<uint32_t> eoff;
            std::vector<uint16_t> eval;
            stable_transpose(nf, 3, ne, fe_e, eoff, eval);
            q.off.assign(P.nfo + 1, 0);
            q.val.clear();
            for (uint32_t f = 0; f < P.nfo; ++f) {
                for (int i = 0; i < 3; ++i) {
                    uint16_t e = fe_e[3 * f + i];
                    for (uint32_t k = eoff[e]; k < eoff[e + 1]; ++k)
                        if (eval[k] != f) q.val.push_back(eval[k]);
                }
                q.off[f + 1] = static_cast<uint32_t>(q.val.size());
            }
            break;
        }
        default: break;
    }
    return q;
}

// ---------------------------------------------------------------------------
// Forward-mode dual number over k variables: restates the arithmetic rules of
// include/rxmesh/diff/scalar.h (value + gradient; +, -, *, / and their
// product/quotient rules; Hessian omitted since only gradients are used).
template <int K>
struct Dual
{
    double v = 0;
    double g[K] = {};
    Dual() = default;
    Dual(double x) : v(x) {}
};
template <int K> Dual<K> operator+(const Dual<K>& a, const Dual<K>& b)
{ Dual<K> r; r.v = a.v + b.v; for (int i = 0; i < K; ++i) r.g[i] = a.g[i] + b.g[i]; return r; }
template <int K> Dual<K> operator-(const Dual<K>& a, const Dual<K>& b)
{ Dual<K> r; r.v = a.v - b.v; for (int i = 0; i < K; ++i) r.g[i] = a.g[i] - b.g[i]; return r; }
template <int K> Dual<K> operator-(const Dual<K>& a)
{ Dual<K> r; r.v = -a.v; for (int i = 0; i < K; ++i) r.g[i] = -a.g[i]; return r; }
template <int K> Dual<K> operator*(const Dual<K>& a, const Dual<K>& b)
{ Dual<K> r; r.v = a.v * b.v; for (int i = 0; i < K; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i]; return r; }
template <int K> Dual<K> operator/(const Dual<K>& a, const Dual<K>& b)
{ Dual<K> r; r.v = a.v / b.v; for (int i = 0; i < K; ++i) r.g[i] = (a.g[i] * b.v - a.v * b.g[i]) / (b.v * b.v); return r; }

// Rest frame of a face: reference apps/Param/param.cu:113-141.
// Returns Mr column-major {m00, m10, m01, m11} with columns (b-a, c-a) in 2D.
void rest_frame(const double* a, const double* b, const double* c, double* Mr)
{
    double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double n[3]  = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                    e1[0] * e2[1] - e1[1] * e2[0]};
    double l1    = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    double b1[3] = {e1[0] / l1, e1[1] / l1, e1[2] / l1};
    double t[3]  = {n[1] * b1[2] - n[2] * b1[1], n[2] * b1[0] - n[0] * b1[2],
                    n[0] * b1[1] - n[1] * b1[0]};
    double lt    = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    double b2[3] = {t[0] / lt, t[1] / lt, t[2] / lt};
    auto dot = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    // a -> (0,0), b -> (e1.b1, 0), c -> (e2.b1, e2.b2)
    Mr[0] = dot(e1, b1);
    Mr[1] = 0.0;
    Mr[2] = dot(e2, b1);
    Mr[3] = dot(e2, b2);
}

// Symmetric Dirichlet energy of one face with forward-mode duals over the six
// UV unknowns; restates the lambda of apps/Param/param.cu:145-184 with the
// 2x2 determinant and cofactor inverse written out.
template <typename S>
S symdir_face(const S* a, const S* b, const S* c, const double* Mr, bool& flipped)
{
    S m00 = b[0] - a[0], m10 = b[1] - a[1];
    S m01 = c[0] - a[0], m11 = c[1] - a[1];
    S det = m00 * m11 - m01 * m10;
    flipped = false;
    double dv;
    if constexpr (std::is_same_v<S, double>) dv = det; else dv = det.v;
    if (dv <= 0.0) {
        flipped = true;
        return S(static_cast<double>(FLT_MAX));
    }
    const double dr  = Mr[0] * Mr[3] - Mr[2] * Mr[1];
    const double A   = 0.5 * dr;
    // Mr^{-1} = [[m11, -m01], [-m10, m00]] / det
    const double i00 = Mr[3] / dr, i01 = -Mr[2] / dr, i10 = -Mr[1] / dr, i11 = Mr[0] / dr;
    S j00 = m00 * S(i00) + m01 * S(i10);
    S j01 = m00 * S(i01) + m01 * S(i11);
    S j10 = m10 * S(i00) + m11 * S(i10);
    S j11 = m10 * S(i01) + m11 * S(i11);
    S jn  = j00 * j00 + j01 * j01 + j10 * j10 + j11 * j11;
    S dj  = j00 * j11 - j01 * j10;
    S k00 = j11 / dj, k01 = -j01 / dj, k10 = -j10 / dj, k11 = j00 / dj;
    S kn  = k00 * k00 + k01 * k01 + k10 * k10 + k11 * k11;
    return S(A) * (jn + kn);
}

}  // namespace

// ===========================================================================
extern "C" {

typedef struct ora_mesh ora_mesh;

ora_mesh* ora_create(const uint32_t* fv, uint32_t nF)
{
    Mesh* m = new Mesh();
    m->fv.resize(nF);
    for (uint32_t f = 0; f < nF; ++f)
        for (int i = 0; i < 3; ++i) m->fv[f][i] = fv[3 * f + i];
    build_supporting(*m);
    return reinterpret_cast<ora_mesh*>(m);
}

void ora_destroy(ora_mesh* h) { delete reinterpret_cast<Mesh*>(h); }

// patches_offset: inclusive end of each patch in patches_val (reference layout)
int ora_set_patches(ora_mesh* h, const uint32_t* face_patch, const uint32_t* patches_val,
                    const uint32_t* patches_offset, uint32_t P)
{
    Mesh& m = *reinterpret_cast<Mesh*>(h);
    m.P = P;
    m.face_patch.assign(face_patch, face_patch + m.nF);
    m.patches_val.assign(patches_val, patches_val + m.nF);
    m.patches_offset.assign(patches_offset, patches_offset + P);
    extract_ribbons(m);
    assign_patch(m);
    m.patches.assign(P, Patch());
    OwnedLists own;
    bucket(m.edge_patch, P, own.e_off, own.e_val);
    bucket(m.vertex_patch, P, own.v_off, own.v_val);
    for (uint32_t p = 0; p < P; ++p) build_ltog(m, p, own);
    for (uint32_t p = 0; p < P; ++p) {
        if (m.patches[p].ltog_v.size() >= INV16 || m.patches[p].ltog_e.size() >= INV16 ||
            m.patches[p].ltog_f.size() >= INV16)
            return 1;
        build_topology(m, p);
    }
    build_linear(m);
    return 0;
}

// counts[0..5] = nV, nE, nF, P, ribbon size, ff size
void ora_counts(const ora_mesh* h, uint64_t* out)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    out[0] = m.nV;
    out[1] = m.nE;
    out[2] = m.nF;
    out[3] = m.P;
    out[4] = m.ribbon_val.size();
    out[5] = m.ff_values.size();
}

// stats[0..4] = closed, edge-manifold, max valence, max edge-incident faces, max face-adjacent faces
void ora_stats(const ora_mesh* h, uint32_t* out)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    out[0] = m.closed;
    out[1] = m.manifold;
    out[2] = m.max_valence;
    out[3] = m.max_edge_faces;
    out[4] = m.max_face_adj;
}

void ora_get_ev(const ora_mesh* h, uint32_t* ev)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    for (uint32_t e = 0; e < m.nE; ++e) {
        ev[2 * e]     = m.ev[e][0];
        ev[2 * e + 1] = m.ev[e][1];
    }
}

void ora_get_ff(const ora_mesh* h, uint32_t* off, uint32_t* val)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    std::copy(m.ff_offset.begin(), m.ff_offset.end(), off);
    std::copy(m.ff_values.begin(), m.ff_values.end(), val);
}

void ora_get_ownership(const ora_mesh* h, uint32_t* vertex_patch, uint32_t* edge_patch)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    std::copy(m.vertex_patch.begin(), m.vertex_patch.end(), vertex_patch);
    std::copy(m.edge_patch.begin(), m.edge_patch.end(), edge_patch);
}

void ora_get_ribbon(const ora_mesh* h, uint32_t* val, uint32_t* offset)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    std::copy(m.ribbon_val.begin(), m.ribbon_val.end(), val);
    std::copy(m.ribbon_offset.begin(), m.ribbon_offset.end(), offset);
}

// out[0..5] = nv, ne, nf (incl. ribbon), nvo, neo, nfo
void ora_patch_counts(const ora_mesh* h, uint32_t p, uint32_t* out)
{
    const Patch& P = reinterpret_cast<const Mesh*>(h)->patches[p];
    out[0] = P.ltog_v.size();
    out[1] = P.ltog_e.size();
    out[2] = P.ltog_f.size();
    out[3] = P.nvo;
    out[4] = P.neo;
    out[5] = P.nfo;
}

void ora_patch_ltog(const ora_mesh* h, uint32_t p, int kind, uint32_t* out)
{
    const Patch& P = reinterpret_cast<const Mesh*>(h)->patches[p];
    const auto& l = (kind == 0) ? P.ltog_v : (kind == 1) ? P.ltog_e : P.ltog_f;
    std::copy(l.begin(), l.end(), out);
}

void ora_patch_ev(const ora_mesh* h, uint32_t p, uint16_t* out)
{
    const Patch& P = reinterpret_cast<const Mesh*>(h)->patches[p];
    std::copy(P.ev.begin(), P.ev.end(), out);
}

void ora_patch_fe(const ora_mesh* h, uint32_t p, uint16_t* out)
{
    const Patch& P = reinterpret_cast<const Mesh*>(h)->patches[p];
    std::copy(P.fe.begin(), P.fe.end(), out);
}

// linear -> global id map for kind 0=V, 1=E, 2=F
void ora_linear_to_global(const ora_mesh* h, int kind, uint32_t* out)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    const auto& l = (kind == 0) ? m.v_l2g : (kind == 1) ? m.e_l2g : m.f_l2g;
    std::copy(l.begin(), l.end(), out);
}

// Number of entries of a query over all owned source elements.
int64_t ora_query_nnz(const ora_mesh* h, int op)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    int64_t n = 0;
    for (uint32_t p = 0; p < m.P; ++p) {
        PatchQuery q = patch_query(m, p, op, false);
        n += q.off.back();
    }
    return n;
}

// CSR over source linear ids; columns are target linear ids in the
// deterministic (stable / canonical-oriented) order.
int ora_query(const ora_mesh* h, int op, int oriented, int32_t* row_ptr, int32_t* col)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    const int sk = src_kind(op), dk = dst_kind(op);
    const auto& prefix = (sk == 0) ? m.v_prefix : (sk == 1) ? m.e_prefix : m.f_prefix;
    int64_t base = 0;
    for (uint32_t p = 0; p < m.P; ++p) {
        PatchQuery q = patch_query(m, p, op, oriented != 0);
        uint32_t nsrc = static_cast<uint32_t>(q.off.size() - 1);
        for (uint32_t s = 0; s < nsrc; ++s) {
            row_ptr[prefix[p] + s] = static_cast<int32_t>(base + q.off[s]);
            for (uint32_t k = q.off[s]; k < q.off[s + 1]; ++k)
                col[base + k] = (q.val[k] == 0xFFFFu) ? -1 : static_cast<int32_t>(to_linear(m, p, dk, q.val[k]));
        }
        base += q.off[nsrc];
    }
    row_ptr[prefix[m.P]] = static_cast<int32_t>(base);
    return 0;
}

// Brute-force global relations straight from fv and the edge map, in the
// style of tests/RXMesh_test/rxmesh_test.h:65-339 (ground truth for set
// checks). Rows are global source ids; order within a row is ascending.
int64_t ora_global_relation(const ora_mesh* h, int op, int32_t* row_ptr, int32_t* col)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    uint32_t nsrc = (src_kind(op) == 0) ? m.nV : (src_kind(op) == 1) ? m.nE : m.nF;
    std::vector<std::vector<uint32_t>> rel(nsrc);
    for (const auto& kv : m.edges_map) {
        uint32_t a = kv.first.first, b = kv.first.second, e = kv.second;
        if (op == OP_VV) { rel[a].push_back(b); rel[b].push_back(a); }
        if (op == OP_VE) { rel[a].push_back(e); rel[b].push_back(e); }
        if (op == OP_EV) { rel[e].push_back(a); rel[e].push_back(b); }
    }
    for (uint32_t f = 0; f < m.nF; ++f)
        for (int i = 0; i < 3; ++i) {
            uint32_t v = m.fv[f][i];
            uint32_t e = m.edge_id(m.fv[f][i], m.fv[f][(i + 1) % 3]);
            if (op == OP_VF) rel[v].push_back(f);
            if (op == OP_FV) rel[f].push_back(v);
            if (op == OP_FE) rel[f].push_back(e);
            if (op == OP_EF) rel[e].push_back(f);
        }
    if (op == OP_EVD) {
        // [a, c, b, d] per edge with a the larger global id: c closes the
        // face holding half-edge a -> b, d the face holding b -> a (-1 when
        // there is none; the last such face in face order otherwise)
        for (const auto& kv : m.edges_map)
            rel[kv.second] = {kv.first.first, INV32, kv.first.second, INV32};
        for (uint32_t f = 0; f < m.nF; ++f)
            for (int i = 0; i < 3; ++i) {
                const uint32_t x = m.fv[f][i], y = m.fv[f][(i + 1) % 3], z = m.fv[f][(i + 2) % 3];
                rel[m.edge_id(x, y)][x > y ? 1 : 3] = z;
            }
    }
    if (op == OP_FF)
        for (uint32_t e = 0; e < m.nE; ++e)
            for (uint32_t f0 : m.edge_faces[e])
                for (uint32_t f1 : m.edge_faces[e])
                    if (f0 != f1) rel[f0].push_back(f1);
    int64_t n = 0;
    for (uint32_t s = 0; s < nsrc; ++s) {
        if (op != OP_FV && op != OP_FE && op != OP_EV && op != OP_EVD)
            std::sort(rel[s].begin(), rel[s].end());
        if (row_ptr) row_ptr[s] = static_cast<int32_t>(n);
        for (uint32_t x : rel[s]) {
            if (col) col[n] = static_cast<int32_t>(x);
            ++n;
        }
    }
    if (row_ptr) row_ptr[nsrc] = static_cast<int32_t>(n);
    return n;
}

// ---------------------------------------------------------------------------
// Geometry (global ids, AoS doubles).

// Gaussian curvature, restating apps/GaussianCurvature/gaussian_curvature_ref.h
// (serial reference, lines 8-118; the GPU lambda is gaussian_curvature_kernel.cuh
// :19-62): per face, s = |(c1 - c0) x (c2 - c0)|, corner dots c[i], squared
// edge lengths l[i] = |v_i - v_{i+1}|^2, angles atan2(s, c[i]); an obtuse face
// gives its obtuse corner s/4 and the others s/8 of mixed area, otherwise
// corner i gets (l[i+2] c[i+1] / s + l[i] c[i+2] / s) / 8; every corner's
// angle is subtracted from 2 pi. gc = defect / mixed area.
void ora_gaussian_curvature(const ora_mesh* h, const double* pos, double* gc, double* amix)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    const double pi = 3.14159265358979323846;
    for (uint32_t v = 0; v < m.nV; ++v) {
        gc[v]   = 2.0 * pi;
        amix[v] = 0.0;
    }
    for (uint32_t f = 0; f < m.nF; ++f) {
        const double* P[3] = {pos + 3 * m.fv[f][0], pos + 3 * m.fv[f][1], pos + 3 * m.fv[f][2]};
        auto sub = [](const double* a, const double* b, double* o) {
            for (int k = 0; k < 3; ++k) o[k] = a[k] - b[k];
        };
        auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
        double e1[3], e2[3];
        sub(P[1], P[0], e1);
        sub(P[2], P[0], e2);
        const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                     cz = e1[0] * e2[1] - e1[1] * e2[0];
        const double s = std::sqrt(cx * cx + cy * cy + cz * cz);
        double       c[3], l[3], rads[3];
        bool         ob = false;
        for (int i = 0; i < 3; ++i) {
            double a[3], b[3], d[3];
            sub(P[(i + 1) % 3], P[i], a);
            sub(P[(i + 2) % 3], P[i], b);
            c[i] = dot(a, b);
            sub(P[i], P[(i + 1) % 3], d);
            l[i]    = dot(d, d);
            rads[i] = std::atan2(s, c[i]);
            if (rads[i] > 0.5 * pi) ob = true;
        }
        for (int i = 0; i < 3; ++i) {
            const int    i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            const uint32_t v = m.fv[f][i];
            if (ob) amix[v] += (rads[i] > 0.5 * pi) ? 0.25 * s : 0.125 * s;
            else amix[v] += 0.125 * (l[i2] * (c[i1] / s) + l[i] * (c[i2] / s));
            gc[v] -= rads[i];
        }
    }
    for (uint32_t v = 0; v < m.nV; ++v) gc[v] /= amix[v];
}

// Vertex normals. mode 0 = Max (1999) weights, restating
// apps/VertexNormal/vertex_normal_ref.h:6-85 (unnormalised sum of
// n_f / (|e_v|^2 + |e_{v+2}|^2)); mode 1 = area weights (sum of n_f / 2).
void ora_vertex_normals(const ora_mesh* h, const double* pos, int mode, double* out)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    std::fill(out, out + 3 * m.nV, 0.0);
    for (uint32_t f = 0; f < m.nF; ++f) {
        const uint32_t* v = m.fv[f].data();
        const double* p0 = pos + 3 * v[0];
        const double* p1 = pos + 3 * v[1];
        const double* p2 = pos + 3 * v[2];
        double u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        double w[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        auto d2 = [](const double* a, const double* b) {
            double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
            return x * x + y * y + z * z;
        };
        double l[3] = {d2(p0, p1), d2(p1, p2), d2(p2, p0)};
        for (int i = 0; i < 3; ++i) {
            double s = (mode == 0) ? 1.0 / (l[i] + l[(i + 2) % 3]) : 0.5;
            for (int c = 0; c < 3; ++c) out[3 * v[i] + c] += n[c] * s;
        }
    }
}

// Uniform Laplacian smoothing, Jacobi: x <- x + lambda * (mean(N(v)) - x).
void ora_smooth_uniform(const ora_mesh* h, const double* pos, double lambda, int iters, double* out)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    std::vector<double> cur(pos, pos + 3 * m.nV), nxt(3 * m.nV);
    std::vector<std::vector<uint32_t>> vv(m.nV);
    for (const auto& kv : m.edges_map) {
        vv[kv.first.first].push_back(kv.first.second);
        vv[kv.first.second].push_back(kv.first.first);
    }
    for (int it = 0; it < iters; ++it) {
        for (uint32_t v = 0; v < m.nV; ++v) {
            double s[3] = {0, 0, 0};
            for (uint32_t u : vv[v])
                for (int c = 0; c < 3; ++c) s[c] += cur[3 * u + c];
            double k = static_cast<double>(vv[v].size());
            for (int c = 0; c < 3; ++c)
                nxt[3 * v + c] = (k > 0) ? cur[3 * v + c] + lambda * (s[c] / k - cur[3 * v + c])
                                         : cur[3 * v + c];
        }
        std::swap(cur, nxt);
    }
    std::copy(cur.begin(), cur.end(), out);
}

// reference: include/rxmesh/geometry_util.cuh:107-206 restated in double
static void clamp_cot(double& v)
{
    const double bound = 19.1;
    v = (v < -bound) ? -bound : ((v > bound) ? bound : v);
}
static double tri_area3(const double* p0, const double* p1, const double* p2)
{
    double u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    double w[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
    return 0.5 * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
}
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double partial_voronoi(const double* p, const double* q, const double* r)
{
    double pq[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
    double qr[3] = {r[0] - q[0], r[1] - q[1], r[2] - q[2]};
    double pr[3] = {r[0] - p[0], r[1] - p[1], r[2] - p[2]};
    double area  = tri_area3(p, q, r);
    if (area <= static_cast<double>(FLT_MIN)) return -1;
    double dotp = dot3(pq, pr), dotq = -dot3(qr, pq), dotr = dot3(qr, pr);
    if (dotp < 0.0) return 0.25 * area;
    if (dotq < 0.0 || dotr < 0.0) return 0.125 * area;
    double cotq = dotq / area, cotr = dotr / area;
    clamp_cot(cotq);
    clamp_cot(cotr);
    return 0.125 * (dot3(pr, pr) * cotq + dot3(pq, pq) * cotr);
}
static double edge_cotan_weight(const double* p, const double* r, const double* q, const double* s)
{
    auto part = [&](const double* v) {
        double d0[3] = {p[0] - v[0], p[1] - v[1], p[2] - v[2]};
        double d1[3] = {r[0] - v[0], r[1] - v[1], r[2] - v[2]};
        double area  = tri_area3(p, r, v);
        if (area > static_cast<double>(FLT_MIN)) {
            double c = dot3(d0, d1) / area;
            clamp_cot(c);
            return c;
        }
        return 0.0;
    };
    return part(q) + part(s);
}

// Cotangent system of explicit mean-curvature flow, restating the lambda of
// apps/MCF/mcf_kernels.cuh:354-426 (mcf_A_setup): weights and areas are
// evaluated over the canonical oriented one-ring (cyclically, as the lambda
// does), and each weight lands in the column of its neighbour, as
// A_mat(p_id, iter[v]) does. The columns of a row are the diagonal first,
// then the neighbours in VV query order -- the order SparseMatrix fills them
// (sparse_matrix_kernels.cuh:75-140 walks the VV iterator after the
// diagonal). Output in GLOBAL vertex ids as COO (row, col, val);
// mass[v] = 1 / v_weight. Returns the number of entries.
int64_t ora_cot_assemble(const ora_mesh* h, const double* pos, double dt,
                         int32_t* rows, int32_t* cols, double* vals, double* mass)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    int64_t n = 0;
    for (uint32_t p = 0; p < m.P; ++p) {
        PatchQuery q  = patch_query(m, p, OP_VV, true);
        PatchQuery qu = patch_query(m, p, OP_VV, false);
        const Patch& P = m.patches[p];
        for (uint32_t v = 0; v < P.nvo; ++v) {
            uint32_t gp = P.ltog_v[v];
            std::vector<uint32_t> ring;
            for (uint32_t k = q.off[v]; k < q.off[v + 1]; ++k) ring.push_back(P.ltog_v[q.val[k]]);
            const size_t K = ring.size();
            double sum_w = 0, vw = 0;
            std::map<uint32_t, double> wcol;
            uint32_t gq = ring[K - 1];
            for (size_t i = 0; i < K; ++i) {
                uint32_t gr = ring[i];
                uint32_t gs = (i == K - 1) ? ring[0] : ring[i + 1];
                double w = edge_cotan_weight(pos + 3 * gp, pos + 3 * gr, pos + 3 * gq, pos + 3 * gs);
                w = (w >= 0.0) ? w : 0.0;
                w *= dt;
                sum_w += w;
                wcol[gr] = -w;
                double a = partial_voronoi(pos + 3 * gp, pos + 3 * gq, pos + 3 * gr);
                vw += (a > 0) ? a : 0;
                gq = gr;
            }
            vw = 0.5 / vw;
            rows[n] = gp; cols[n] = gp; vals[n] = 1.0 / vw + sum_w; ++n;
            mass[gp] = 1.0 / vw;
            for (uint32_t k = qu.off[v]; k < qu.off[v + 1]; ++k) {
                const uint32_t gr = P.ltog_v[qu.val[k]];
                rows[n] = gp; cols[n] = gr; vals[n] = wcol.at(gr); ++n;
            }
        }
    }
    return n;
}

// Explicit MCF steps on a COO matrix: x <- x + (M*x - A*x) / M, 3 columns (AoS).
void ora_mcf_explicit(int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals,
                      const double* mass, uint32_t nV, const double* x0, int iters, double* out)
{
    std::vector<double> x(x0, x0 + 3 * nV), y(3 * nV);
    for (int it = 0; it < iters; ++it) {
        std::fill(y.begin(), y.end(), 0.0);
        for (int64_t k = 0; k < nnz; ++k)
            for (int c = 0; c < 3; ++c) y[3 * rows[k] + c] += vals[k] * x[3 * cols[k] + c];
        for (uint32_t v = 0; v < nV; ++v)
            for (int c = 0; c < 3; ++c)
                x[3 * v + c] = x[3 * v + c] + (mass[v] * x[3 * v + c] - y[3 * v + c]) / mass[v];
    }
    std::copy(x.begin(), x.end(), out);
}

// Implicit MCF step with conjugate gradients, restating the reference's
// CGSolver (include/rxmesh/matrix/cg_solver.h:37-125) on A X = B, B = M x0
// (apps/MCF/mcf_kernels.cuh:302-352): dots and norms over all 3V entries,
// x starts at x0, stop when |r|^2 < abs_tol or |r|^2 / |r0|^2 < rel_tol.
// COO matrix (rows, cols, vals) as ora_cot_assemble emits it; positions AoS.
// Returns the iterations taken; res[0] = |r0|^2, res[1] = final |r|^2.
int ora_mcf_cg(int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals, const double* mass,
               uint32_t nV, const double* x0, int max_iter, double abs_tol, double rel_tol, double* out,
               double* res)
{
    const size_t n = 3ull * nV;
    std::vector<double> x(x0, x0 + n), r(n), p(n), s(n);
    auto matvec = [&](const std::vector<double>& a, std::vector<double>& y) {
        std::fill(y.begin(), y.end(), 0.0);
        for (int64_t k = 0; k < nnz; ++k)
            for (int c = 0; c < 3; ++c) y[3 * rows[k] + c] += vals[k] * a[3 * cols[k] + c];
    };
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
        double t = 0.0;
        for (size_t i = 0; i < n; ++i) t += a[i] * b[i];
        return t;
    };
    matvec(x, s);
    for (uint32_t v = 0; v < nV; ++v)
        for (int c = 0; c < 3; ++c) r[3 * v + c] = mass[v] * x0[3 * v + c] - s[3 * v + c];
    p = r;
    double delta = dot(r, r);
    const double start = delta;
    int it = 0;
    auto converged = [&](double d) { return d < abs_tol || d / start < rel_tol; };
    if (!converged(delta)) {
        while (it < max_iter) {
            matvec(p, s);
            const double alpha = delta / dot(s, p);
            for (size_t i = 0; i < n; ++i) x[i] += alpha * p[i];
            for (size_t i = 0; i < n; ++i) r[i] -= alpha * s[i];
            const double delta_old = delta;
            delta = dot(r, r);
            if (converged(delta)) break;
            const double beta = delta / delta_old;
            for (size_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
            ++it;
        }
    }
    std::copy(x.begin(), x.end(), out);
    res[0] = start;
    res[1] = delta;
    return it;
}

// The uniform-weight system of the MCF app (use_uniform_laplace, apps/MCF/
// mcf_kernels.cuh:354-426 and the matvec of :114-213): every edge weight is
// 1 (times dt), v_weight = 1 / valence, so A(p, p) = 1 / v_weight + dt *
// valence and A(p, r) = -dt; the right-hand side scale mass[p] = valence
// (init_B, :57-108). Same COO layout as ora_cot_assemble (diagonal, then the
// VV query row). Returns the number of entries.
int64_t ora_uniform_assemble(const ora_mesh* h, double dt, int32_t* rows, int32_t* cols, double* vals,
                             double* mass)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    int64_t n = 0;
    for (uint32_t p = 0; p < m.P; ++p) {
        PatchQuery   q = patch_query(m, p, OP_VV, false);
        const Patch& P = m.patches[p];
        for (uint32_t v = 0; v < P.nvo; ++v) {
            const uint32_t gp = P.ltog_v[v];
            const uint32_t K  = q.off[v + 1] - q.off[v];
            double sum_w = 0, vw = 0;
            for (uint32_t k = 0; k < K; ++k) {
                sum_w += 1.0 * dt;
                vw += 1.0;
            }
            vw = 1.0 / vw;
            rows[n] = gp; cols[n] = gp; vals[n] = 1.0 / vw + sum_w; ++n;
            mass[gp] = static_cast<double>(K);
            for (uint32_t k = q.off[v]; k < q.off[v + 1]; ++k) {
                rows[n] = gp; cols[n] = P.ltog_v[q.val[k]]; vals[n] = -dt; ++n;
            }
        }
    }
    return n;
}

// Jacobi-preconditioned CG, restating PCGMatFreeAttrSolver
// (include/rxmesh/matrix/pcg_mat_free_attr_solver.h:39-125) with the
// preconditioner of precond_matvec (apps/MCF/mcf_kernels.cuh:219-299): z =
// r / diag(A). B = M x0, x starts at x0; delta = <r, z>; stop when delta <
// abs_tol or delta / delta0 < rel_tol. A start residual already below the
// tolerance ends with zero iterations, as CGSolver does (the reference's
// matrix-free solvers always take one iteration, which divides 0 by 0 when
// the start residual is exactly zero). Same arguments and results as
// ora_mcf_cg.
int ora_mcf_pcg(int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals, const double* mass,
                uint32_t nV, const double* x0, int max_iter, double abs_tol, double rel_tol, double* out,
                double* res)
{
    const size_t n = 3ull * nV;
    std::vector<double> x(x0, x0 + n), r(n), p(n), s(n), z(n), diag(nV, 0.0);
    for (int64_t k = 0; k < nnz; ++k)
        if (rows[k] == cols[k]) diag[rows[k]] += vals[k];
    auto matvec = [&](const std::vector<double>& a, std::vector<double>& y) {
        std::fill(y.begin(), y.end(), 0.0);
        for (int64_t k = 0; k < nnz; ++k)
            for (int c = 0; c < 3; ++c) y[3 * rows[k] + c] += vals[k] * a[3 * cols[k] + c];
    };
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
        double t = 0.0;
        for (size_t i = 0; i < n; ++i) t += a[i] * b[i];
        return t;
    };
    auto precond = [&](const std::vector<double>& a, std::vector<double>& y) {
        for (size_t i = 0; i < n; ++i) y[i] = a[i] / diag[i / 3];
    };
    matvec(x, s);
    for (uint32_t v = 0; v < nV; ++v)
        for (int c = 0; c < 3; ++c) r[3 * v + c] = mass[v] * x0[3 * v + c] - s[3 * v + c];
    precond(r, p);
    double delta = std::abs(dot(r, p));
    const double start = delta;
    int it = 0;
    auto converged = [&](double d) { return d < abs_tol || d / start < rel_tol; };
    if (!converged(delta)) {
        while (it < max_iter) {
            matvec(p, s);
            const double alpha = delta / dot(s, p);
            for (size_t i = 0; i < n; ++i) x[i] += alpha * p[i];
            for (size_t i = 0; i < n; ++i) r[i] -= alpha * s[i];
            precond(r, z);
            const double delta_old = delta;
            delta = dot(r, z);
            if (converged(delta)) break;
            const double beta = delta / delta_old;
            for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
            ++it;
        }
    }
    std::copy(x.begin(), x.end(), out);
    res[0] = start;
    res[1] = delta;
    return it;
}

// Rest shapes Mr (4 per face, column-major) from 3D rest positions.
void ora_rest_shapes(const ora_mesh* h, const double* pos, double* Mr)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    for (uint32_t f = 0; f < m.nF; ++f)
        rest_frame(pos + 3 * m.fv[f][0], pos + 3 * m.fv[f][1], pos + 3 * m.fv[f][2], Mr + 4 * f);
}

// Per-face symmetric Dirichlet energy (FLT_MAX for flipped faces).
void ora_symdir_energy(const ora_mesh* h, const double* Mr, const double* uv, double* energy)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    for (uint32_t f = 0; f < m.nF; ++f) {
        bool fl;
        energy[f] = symdir_face<double>(uv + 2 * m.fv[f][0], uv + 2 * m.fv[f][1],
                                        uv + 2 * m.fv[f][2], Mr + 4 * f, fl);
    }
}

// Gradient of sum_f E_f w.r.t. UV (2 per vertex) by the forward-mode dual
// restatement (the reference differentiates the same lambda with
// Scalar<T,6> and scatters per-face gradients, diff_query_kernel.cuh:275-282).
// Flipped faces contribute a constant and hence zero gradient.
void ora_symdir_grad_dual(const ora_mesh* h, const double* Mr, const double* uv, double* grad)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    std::fill(grad, grad + 2 * m.nV, 0.0);
    for (uint32_t f = 0; f < m.nF; ++f) {
        Dual<6> x[3][2];
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 2; ++c) {
                x[i][c].v = uv[2 * m.fv[f][i] + c];
                x[i][c].g[2 * i + c] = 1.0;
            }
        bool fl;
        Dual<6> e = symdir_face<Dual<6>>(x[0], x[1], x[2], Mr + 4 * f, fl);
        if (fl) continue;
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 2; ++c) grad[2 * m.fv[f][i] + c] += e.g[2 * i + c];
    }
}

// Closed-form gradient: with E = A (|J|^2 + |J|^2 / det(J)^2),
// dE/dJ = A (2J (1 + 1/d^2) - 2|J|^2/d^3 cof(J)), dE/dM = dE/dJ Mr^{-T},
// d/db = col0, d/dc = col1, d/da = -(col0 + col1).
void ora_symdir_grad_closed(const ora_mesh* h, const double* Mr, const double* uv, double* grad)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    std::fill(grad, grad + 2 * m.nV, 0.0);
    for (uint32_t f = 0; f < m.nF; ++f) {
        const double* a = uv + 2 * m.fv[f][0];
        const double* b = uv + 2 * m.fv[f][1];
        const double* c = uv + 2 * m.fv[f][2];
        double M00 = b[0] - a[0], M10 = b[1] - a[1], M01 = c[0] - a[0], M11 = c[1] - a[1];
        if (M00 * M11 - M01 * M10 <= 0.0) continue;
        const double* R = Mr + 4 * f;
        double dr = R[0] * R[3] - R[2] * R[1];
        double A  = 0.5 * dr;
        double i00 = R[3] / dr, i01 = -R[2] / dr, i10 = -R[1] / dr, i11 = R[0] / dr;
        double J00 = M00 * i00 + M01 * i10, J01 = M00 * i01 + M01 * i11;
        double J10 = M10 * i00 + M11 * i10, J11 = M10 * i01 + M11 * i11;
        double n2 = J00 * J00 + J01 * J01 + J10 * J10 + J11 * J11;
        double d  = J00 * J11 - J01 * J10;
        double s1 = 2.0 * (1.0 + 1.0 / (d * d)), s2 = -2.0 * n2 / (d * d * d);
        double G00 = A * (s1 * J00 + s2 * J11), G01 = A * (s1 * J01 - s2 * J10);
        double G10 = A * (s1 * J10 - s2 * J01), G11 = A * (s1 * J11 + s2 * J00);
        // dE/dM = G Mr^{-T}: (G Minv^T)_{rc} = sum_k G_rk Minv_ck
        double D00 = G00 * i00 + G01 * i01, D01 = G00 * i10 + G01 * i11;
        double D10 = G10 * i00 + G11 * i01, D11 = G10 * i10 + G11 * i11;
        uint32_t va = m.fv[f][0], vb = m.fv[f][1], vc = m.fv[f][2];
        grad[2 * vb] += D00; grad[2 * vb + 1] += D10;
        grad[2 * vc] += D01; grad[2 * vc + 1] += D11;
        grad[2 * va] -= D00 + D01; grad[2 * va + 1] -= D10 + D11;
    }
}

// ---------------------------------------------------------------------------
// Symmetric Dirichlet with 3-D positions as the variable (closed surfaces have
// no UV domain): J = [b-a, c-a] Mr^{-1} is 3x2, |J^{-1}|^2 becomes
// |J|^2 / det(J^T J) with det(J^T J) = |J0 x J1|^2, E = A |J|^2 (1 + 1/G).
// FLT_MAX when G <= 0. This extends apps/Param/param.cu:145-184 (2-D UV) to a
// surface in 3-D; it reduces to the 2-D energy when the surface is planar.
static double symdir3_face(const double* a, const double* b, const double* c, const double* R, double* g)
{
    double M[3][2];
    for (int k = 0; k < 3; ++k) {
        M[k][0] = b[k] - a[k];
        M[k][1] = c[k] - a[k];
    }
    const double dr = R[0] * R[3] - R[2] * R[1];
    const double A  = 0.5 * dr;
    const double i00 = R[3] / dr, i01 = -R[2] / dr, i10 = -R[1] / dr, i11 = R[0] / dr;
    double J0[3], J1[3];
    for (int k = 0; k < 3; ++k) {
        J0[k] = M[k][0] * i00 + M[k][1] * i10;
        J1[k] = M[k][0] * i01 + M[k][1] * i11;
    }
    const double a00 = J0[0] * J0[0] + J0[1] * J0[1] + J0[2] * J0[2];
    const double a11 = J1[0] * J1[0] + J1[1] * J1[1] + J1[2] * J1[2];
    const double a01 = J0[0] * J1[0] + J0[1] * J1[1] + J0[2] * J1[2];
    const double n2 = a00 + a11, G = a00 * a11 - a01 * a01;
    if (G <= 0.0) {
        if (g)
            for (int i = 0; i < 9; ++i) g[i] = 0.0;
        return static_cast<double>(FLT_MAX);
    }
    if (g) {
        // dE/dJ = A [ 2 (1 + 1/G) J - n2/G^2 dG/dJ ]
        const double s1 = 2.0 * (1.0 + 1.0 / G), s2 = -n2 / (G * G);
        double D0[3], D1[3];
        for (int k = 0; k < 3; ++k) {
            const double dG0 = 2.0 * a11 * J0[k] - 2.0 * a01 * J1[k];
            const double dG1 = 2.0 * a00 * J1[k] - 2.0 * a01 * J0[k];
            const double E0 = A * (s1 * J0[k] + s2 * dG0), E1 = A * (s1 * J1[k] + s2 * dG1);
            // dE/dM = dE/dJ R^T with R = Mr^{-1}
            D0[k] = E0 * i00 + E1 * i01;
            D1[k] = E0 * i10 + E1 * i11;
        }
        for (int k = 0; k < 3; ++k) {
            g[3 + k] = D0[k];
            g[6 + k] = D1[k];
            g[k]     = -(D0[k] + D1[k]);
        }
    }
    return A * n2 * (1.0 + 1.0 / G);
}

void ora_symdir3_energy(const ora_mesh* h, const double* Mr, const double* x, double* energy)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    for (uint32_t f = 0; f < m.nF; ++f)
        energy[f] = symdir3_face(x + 3 * m.fv[f][0], x + 3 * m.fv[f][1], x + 3 * m.fv[f][2], Mr + 4 * f, nullptr);
}

void ora_symdir3_grad(const ora_mesh* h, const double* Mr, const double* x, double* grad)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    std::fill(grad, grad + 3 * m.nV, 0.0);
    for (uint32_t f = 0; f < m.nF; ++f) {
        double g[9];
        symdir3_face(x + 3 * m.fv[f][0], x + 3 * m.fv[f][1], x + 3 * m.fv[f][2], Mr + 4 * f, g);
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) grad[3 * m.fv[f][i] + k] += g[3 * i + k];
    }
}

// One step of the batched workload on one mesh: x1 = uniform smoothing of
// x (one Jacobi step, lambda), x2 = x1 - eta * grad E_symdir3(x1).
void ora_smooth_symdir_step(const ora_mesh* h, const double* Mr, const double* x, double lambda, double eta,
                            int steps, double* out)
{
    const Mesh& m = *reinterpret_cast<const Mesh*>(h);
    std::vector<double> cur(x, x + 3 * m.nV), sm(3 * m.nV), g(3 * m.nV);
    for (int it = 0; it < steps; ++it) {
        ora_smooth_uniform(h, cur.data(), lambda, 1, sm.data());
        ora_symdir3_grad(h, Mr, sm.data(), g.data());
        for (size_t i = 0; i < cur.size(); ++i) cur[i] = sm[i] - eta * g[i];
    }
    std::copy(cur.begin(), cur.end(), out);
}

}  // extern "C"
