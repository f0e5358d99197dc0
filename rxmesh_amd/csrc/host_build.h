// Host-side connectivity build of the static patch mesh.
//
// Produces the same patch-local layout the reference builds in
// include/rxmesh/rxmesh.cpp:290-448 (edge ids in first-appearance order,
// ribbons, first-patch ownership, owned-first sorted local numbering, EV/FE
// encodings) but with O(F) bucket passes and a thread pool instead of hash
// maps and per-patch O(V+E) scratch vectors, plus a deterministic patcher.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace rxm {

constexpr uint32_t kInvalid32 = 0xFFFFFFFFu;
constexpr uint16_t kInvalid16 = 0xFFFFu;
// per-patch slabs are padded to this many elements so that every patch's
// u16 / u32 rows start on an 8-byte / 16-byte boundary
constexpr uint32_t kSlabAlign = 4;

struct HostMesh
{
    uint32_t nV = 0, nE = 0, nF = 0, P = 0;
    uint32_t patch_size = 512;

    std::vector<uint32_t> fv;    // 3F
    std::vector<uint32_t> ev;    // 2E, first-seen orientation
    std::vector<uint32_t> fe_g;  // 3F, global edge of (fv[i], fv[i+1])
    std::vector<uint32_t> ef_off, ef_val;  // edge -> incident faces, face order
    std::vector<uint32_t> vf_off, vf_val;  // vertex -> incident faces, face order

    bool     closed = true, manifold = true;
    bool     oriented = true;  // every edge's faces traverse it in opposite directions
    uint32_t max_valence = 0, max_edge_faces = 0, max_face_adj = 0;

    // patch assignment in the reference's layout (patcher.h:162-182)
    std::vector<uint32_t> face_patch, patches_val, patches_offset;
    std::vector<uint32_t> ribbon_val, ribbon_offset;
    std::vector<uint32_t> vertex_patch, edge_patch;

    // per patch
    std::vector<uint32_t> nv, ne, nf, nvo, neo, nfo;  // P each
    std::vector<uint32_t> v_off, e_off, f_off;        // P+1, slab offsets
    std::vector<uint32_t> vp, ep, fp;                 // P+1, linear prefix
    uint32_t              max_nv = 0, max_ne = 0, max_nf = 0;
    uint32_t              max_nvo = 0, max_neo = 0, max_nfo = 0;

    // slabs indexed by off[p] + local
    std::vector<uint32_t> ltog_v, ltog_e, ltog_f;
    std::vector<uint32_t> ev_local;  // lo16 = local(max gid), hi16 = local(min gid)
    std::vector<uint16_t> fe_local;  // 3 per face at 3*f_off[p]: (local edge << 1) | dir
    std::vector<uint32_t> v_lin, e_lin, f_lin;  // linear id of the owner's copy
    std::vector<uint32_t> v_own, e_own, f_own;  // owner patch

    // linear <-> global
    std::vector<uint32_t> v_l2g, e_l2g, f_l2g;
    std::vector<uint32_t> v_g2l, e_g2l, f_g2l;
};

// status codes shared with the C-ABI
enum Status : int
{
    kOk              = 0,
    kErrInvalidArg   = 1,
    kErrNonManifold  = 2,
    kErrPatchTooBig  = 3,
    kErrIO           = 4,
    kErrDevice       = 5,
    kErrUnsupported  = 6,
    kErrOutOfMemory  = 7,
};

// Step a1: edges, incidence, statistics. nV_hint = 0 infers nV = max id + 1.
int build_supporting(HostMesh& m, const uint32_t* fv, uint32_t nF, uint32_t nV_hint);
// The two ends of build_supporting around the edge / VF / EF construction, for
// a caller that builds those elsewhere (on the device): prepare validates fv
// and sets fv, nF, nV; finish checks the arrays' sizes and derives the
// orientation flag and the statistics from ev, fe_g, ef_off and vf_off.
int supporting_prepare(HostMesh& m, const uint32_t* fv, uint32_t nF, uint32_t nV_hint);
int supporting_finish(HostMesh& m);
// The edge ids (first appearance), EV, fe_g and the VF / EF CSRs built on the
// current device by stable radix sorts (gpu_build.hip); fv is host memory.
int build_supporting_device(const uint32_t* fv, uint32_t nF, uint32_t nV, std::vector<uint32_t>& ev,
                            std::vector<uint32_t>& fe, std::vector<uint32_t>& vf_off, std::vector<uint32_t>& vf_val,
                            std::vector<uint32_t>& ef_off, std::vector<uint32_t>& ef_val, uint32_t& nE);

// Step a2: deterministic patcher (fills face_patch / patches_val / patches_offset).
//   kPatcherBisect: recursive proportional bisection into about
//     ceil(F / patch_size) connected patches per component (fewest patches,
//     least ribbon; subtrees run on worker threads);
//   kPatcherKMeans: k-means on the face graph with ceil(F / patch_size) seeds,
//     oversize patches halved (about 0.7 patch_size faces each).
enum PatcherMode : int
{
    kPatcherBisect = 0,
    kPatcherKMeans = 1,
};
int run_patcher(HostMesh& m, uint32_t patch_size, int mode = kPatcherBisect);

// Alternative to run_patcher: take a given assignment (e.g. a reference patch
// file). patches_val / patches_offset may be null (then grouped by face order).
int set_patches(HostMesh& m, const uint32_t* face_patch, uint32_t P,
                const uint32_t* patches_val, const uint32_t* patches_offset);

// Steps a3..a6: ribbons, ownership, local numbering, EV/FE, owner tables.
int build_patches(HostMesh& m);

// Reference patch archive (cereal PortableBinary) reader / writer.
int read_patch_file(const std::string& path, std::vector<uint32_t>& face_patch,
                    std::vector<uint32_t>& patches_val,
                    std::vector<uint32_t>& patches_offset, uint32_t& num_patches,
                    uint32_t& nV, uint32_t& nE, uint32_t& nF);
int write_patch_file(const std::string& path, const HostMesh& m);

// Disjoint union of already built meshes (a batch): element ids, patch ids,
// linear ids and slab offsets of mesh k are shifted by the totals of meshes
// < k, so the union is itself a valid patch mesh whose patches are the
// concatenation of all meshes' patches. v/e/f_base receive count + 1 linear
// (= global) id offsets per mesh, p_base the patch offsets.
int concat_meshes(const std::vector<const HostMesh*>& parts, HostMesh& out, std::vector<uint32_t>& v_base,
                  std::vector<uint32_t>& e_base, std::vector<uint32_t>& f_base, std::vector<uint32_t>& p_base);

// Two-step smoothing plan. For every patch p, a row for EVERY local vertex
// l in [0, nv) -- owned or ribbon -- listing its full one-ring in the
// canonical order of its owner patch's VV row (ascending owner-local edge
// id), with neighbours numbered in p's extended local space: [0, nv) are p's
// local vertices, [nv, nv + nx) the extra second-ring vertices, sorted by
// global id. One workgroup can then run two Jacobi steps of its patch: step 1
// on all local vertices, step 2 on the owned ones, with the same neighbour
// order (hence the same results) as two single steps.
struct Vv2Plan
{
    std::vector<uint32_t> rbase, nbase, xbase;  // P + 1 each: u16 / u16 / u32 offsets, runs padded to 8 / 8 / 4
    std::vector<uint32_t> nx;                   // P: extra vertices per patch
    std::vector<uint16_t> roff, nbr;            // per patch: nv + 1 row offsets; entries
    std::vector<uint32_t> xlin;                 // per patch: linear ids of the extra vertices
    uint32_t              max_nx = 0, max_nent = 0;
};
int build_vv2_plan(const HostMesh& m, Vv2Plan& out);

// K-level smoothing plan (temporal blocking). For every patch p an extended
// vertex space ordered by ring level around the patch's owned vertices:
// level 0 = the owned vertices (local order), level k = the vertices first
// reached in k VV steps, each level sorted by global id. cnt[k] = vertices of
// levels <= k, so a launch of s <= K Jacobi steps computes, at step i, the
// prefix [0, cnt[s - i]) -- exactly the vertices the later steps read -- and
// the last step the owned vertices. Rows (the full one-ring in the owner's
// canonical VV order, neighbours as extended ids) exist for [0, cnt[K - 1]).
// lvl holds 2K + 1 u32 per patch: cnt[0..K], then ent[0..K-1] with ent[k] =
// row entries of [0, cnt[k]).
struct VvkPlan
{
    uint32_t              K = 0;
    std::vector<uint32_t> lvl;                  // P * (2K + 1)
    std::vector<uint32_t> rbase, nbase, xbase;  // P + 1 each: u16 / u16 / u32 offsets, runs padded to 8 / 8 / 4
    std::vector<uint16_t> roff, nbr;            // per patch: cnt[K-1] + 1 row offsets; entries
    std::vector<uint32_t> xlin;                 // per patch: linear ids of extended vertices [nvo, cnt[K])
    std::vector<uint32_t> max_cnt, max_ent;     // K + 1 / K: largest cnt[k] / ent[k] over patches
};
int build_vvk_plan(const HostMesh& m, uint32_t K, VvkPlan& out);

// Vertices whose one-ring is one closed fan: the successor map of its faces
// (half-edge v -> a of face (v, a, b) is followed by b) is a single cycle
// through all its neighbours. The canonical ring of such a vertex has a real
// face between every two consecutive neighbours, so its cotangent row can be
// built from per-face quantities. Needs a consistently oriented mesh.
// closed[g] = 1 for such vertices (global ids).
void closed_fan_vertices(const HostMesh& m, std::vector<uint8_t>& closed);

// Patch-local pattern of the VV query: for every patch p and owned vertex l,
// the patch's local edges incident to l in ascending local edge id (the VV
// query's row order), as ent[vv_base[p] + k] = 2 edge + (l is the edge's
// second endpoint): the index, in FE's encoding, of the half-edge leaving l.
// rs[lin] is the first entry of linear vertex lin, rs[nV] the total. vv_base
// is the VV query's per-patch entry base (P + 1). Returns false when a patch
// has 2^15 edges or more (the entries hold 16 bits).
bool build_vv_edge_pattern(const HostMesh& m, const std::vector<uint32_t>& vv_base, std::vector<uint16_t>& ent,
                           std::vector<uint32_t>& rs);

// Host threads a build may use: RXM_NUM_THREADS or the hardware's count,
// capped on the calling thread by a ThreadLimit in scope (k = 0 keeps the
// current cap; threads a build spawns inherit the cap).
unsigned num_threads();
struct ThreadLimit
{
    explicit ThreadLimit(unsigned k);
    ~ThreadLimit();
    unsigned prev;
};

}  // namespace rxm
