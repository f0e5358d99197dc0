// Lean query kernels (compile-time workgroup size, register row sorts).
//
// Same results as the general transpose in lds_core.hpp -- every row lists
// its entries in ascending source-local order -- with far fewer vector
// instructions per entry:
//   1. degree histogram with LDS atomics (order-free), owned rows only;
//   2. exclusive scan (compile-time workgroup size, no runtime division);
//   3. slot fill with LDS atomics, each slot packing (sort key << 16 | payload);
//   4. one thread per owned row loads its slots into registers, sorts them
//      with a fixed compare-exchange network (rows up to 8 entries; longer
//      rows take an LDS insertion sort) and writes the row straight to HBM.
#pragma once
#include "lds_sort.hpp"

namespace rxm {

// ---- plans -----------------------------------------------------------------
struct PlanQ2
{
    uint32_t *ev, *off, *cur, *slots, *wave, *lin;
    uint16_t *fe;
};
// kind: 0 = VV, 3 = VE (EV only), 1 = VF (EV + FE), 2 = EF (FE only), 4 = FF
// (FE only, plus per owned face entry counts).
// Every array is sized for its kind, so a workgroup holds no more LDS than
// its relation needs (occupancy is LDS-bound).
// hist_mul = 1: offsets of the atomic sort (ncols + 1); hist_mul = waves per
// workgroup: the (key, wave) table of the stable sort (ncols * waves + 1).
__host__ __device__ inline PlanQ2 plan_q2(Bump& b, int kind, uint32_t nv, uint32_t ne, uint32_t nf,
                                          uint32_t hist_mul = 1)
{
    PlanQ2 p{};
    const bool     vx    = (kind == 0 || kind == 3);
    const uint32_t ncols = (kind == 2 || kind == 4) ? ne : nv;
    const uint32_t n     = vx ? 2 * ne : 3 * nf;
    if (kind != 2 && kind != 4) p.ev = b.alloc<uint32_t>(r4(ne));
    if (kind == 1 || kind == 2 || kind == 4) p.fe = b.alloc<uint16_t>(r4(3 * nf));
    p.off   = b.alloc<uint32_t>(ncols * hist_mul + 1);
    if (kind == 4) p.cur = b.alloc<uint32_t>(nf + 1);  // FF: per owned face entry counts
    p.slots = b.alloc<uint32_t>(n);
    p.wave  = b.alloc<uint32_t>(20);
    p.lin   = b.alloc<uint32_t>((kind == 0) ? nv : ((kind == 3) ? ne : nf));
    return p;
}

// One-round-trip staging of a query's inputs: the EV slab (16-byte chunks),
// the FE slab (8-byte chunks) and the target linear ids (owned: prefix + l;
// ribbon: owner table) -- every global load of a thread is issued before
// its first LDS store. A null destination skips that input; counts beyond
// the register rounds take plain loops.
template <uint32_t T>
__device__ __forceinline__ void stage_topology(uint32_t* ev_d, const uint32_t* __restrict__ ev_s, uint32_t ne,
                                               uint16_t* fe_d, const uint16_t* __restrict__ fe_s, uint32_t nf,
                                               uint32_t* lin_d, uint32_t n, uint32_t n_owned, uint32_t prefix,
                                               const uint32_t* __restrict__ slab)
{
    constexpr uint32_t RE = 2, RF = 4, RL = 5;
    const uint32_t     t   = threadIdx.x;
    const uint32_t     n4  = ev_d ? (ne + 3u) >> 2 : 0u;
    const uint32_t     n2  = fe_d ? (3u * nf + 3u) >> 2 : 0u;
    const uint32_t     nl  = lin_d ? n : 0u;
    const uint4*       es  = reinterpret_cast<const uint4*>(ev_s);
    const uint2*       fs  = reinterpret_cast<const uint2*>(fe_s);
    uint4*             ed  = reinterpret_cast<uint4*>(ev_d);
    uint2*             fd  = reinterpret_cast<uint2*>(fe_d);
    uint4              ea[RE];
    uint2              fa[RF];
    uint32_t           la[RL];
#pragma unroll
    for (uint32_t r = 0; r < RE; ++r) ea[r] = (t + r * T < n4) ? es[t + r * T] : uint4{};
#pragma unroll
    for (uint32_t r = 0; r < RF; ++r) fa[r] = (t + r * T < n2) ? fs[t + r * T] : uint2{};
#pragma unroll
    for (uint32_t r = 0; r < RL; ++r) {
        const uint32_t l = t + r * T;
        la[r]            = (l < n_owned) ? prefix + l : ((l < nl) ? slab[l] : 0u);
    }
#pragma unroll
    for (uint32_t r = 0; r < RE; ++r)
        if (t + r * T < n4) ed[t + r * T] = ea[r];
#pragma unroll
    for (uint32_t r = 0; r < RF; ++r)
        if (t + r * T < n2) fd[t + r * T] = fa[r];
#pragma unroll
    for (uint32_t r = 0; r < RL; ++r)
        if (t + r * T < nl) lin_d[t + r * T] = la[r];
    for (uint32_t i = t + RE * T; i < n4; i += T) ed[i] = es[i];
    for (uint32_t i = t + RF * T; i < n2; i += T) fd[i] = fs[i];
    for (uint32_t l = t + RL * T; l < nl; l += T) lin_d[l] = (l < n_owned) ? prefix + l : slab[l];
}

// Coalesced copy of n u32 from LDS to a global int32 run (thread i writes
// entries i, i + T, ...).
template <uint32_t T>
__device__ __forceinline__ void copy_out(int32_t* __restrict__ dst, const uint32_t* src, uint32_t n)
{
    for (uint32_t i = threadIdx.x; i < n; i += T) dst[i] = static_cast<int32_t>(src[i]);
}

// Key of VV / VE entry i = (edge i >> 1, endpoint i & 1): the endpoint's
// local vertex.
__device__ __forceinline__ uint32_t ev_end(const uint32_t* ev, uint32_t i)
{
    const uint32_t x = ev[i >> 1];
    return (i & 1u) ? hi16(x) : lo16(x);
}
__device__ __forceinline__ uint32_t ev_other(const uint32_t* ev, uint32_t i)
{
    const uint32_t x = ev[i >> 1];
    return (i & 1u) ? lo16(x) : hi16(x);
}

// VV (IS_VV) or VE over linear ids. STABLE: atomic-free counting sort
// (sort_columns_stable), rows in entry order with no row sort; otherwise the
// LDS-atomic histogram and slot fill followed by a sort of every row. Both
// give the same rows: ascending local edge id.
template <uint32_t T, bool IS_VV, bool STABLE = false>
__global__ __launch_bounds__(T) void k_query2_vx(DevMesh m, const uint32_t* __restrict__ base,
                                                 int32_t* __restrict__ row_ptr, int32_t* __restrict__ col)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanQ2    L = plan_q2(b, IS_VV ? 0 : 3, d.nv, d.ne, d.nf, STABLE ? T / 64u : 1u);
    if (IS_VV)
        stage_topology<T>(L.ev, m.ev + d.e_off, d.ne, nullptr, nullptr, 0, L.lin, d.nv, d.nvo, d.vp,
                          m.v_lin + d.v_off);
    else
        stage_topology<T>(L.ev, m.ev + d.e_off, d.ne, nullptr, nullptr, 0, L.lin, d.ne, d.neo, d.ep,
                          m.e_lin + d.e_off);
    if constexpr (STABLE) {
        const uint32_t bse = base[p];
        const uint32_t tot = sort_columns_stable<T>(
            2 * d.ne, d.nvo, [&](uint32_t i) { return ev_end(L.ev, i); },
            [&](uint32_t i) { return IS_VV ? L.lin[ev_other(L.ev, i)] : L.lin[i >> 1]; },
            [&](uint32_t c, uint32_t st) { row_ptr[d.vp + c] = static_cast<int32_t>(bse + st); }, L.off, L.slots,
            L.wave);
        if (threadIdx.x == 0) row_ptr[d.vp + d.nvo] = static_cast<int32_t>(bse + tot);
        copy_out<T>(col + bse, L.slots, tot);
        return;
    }
    for (uint32_t c = threadIdx.x; c <= d.nvo; c += T) L.off[c] = 0;
    __syncthreads();
    // entry i = (edge i >> 1, endpoint i & 1); slot = edge << 16 | other endpoint
    fill_columns<T, true>(
        2 * d.ne, d.nvo,
        [&](uint32_t i) {
            const uint32_t x = L.ev[i >> 1];
            return (i & 1u) ? hi16(x) : lo16(x);
        },
        [&](uint32_t i) {
            const uint32_t x = L.ev[i >> 1];
            return ((i >> 1) << 16) | ((i & 1u) ? lo16(x) : hi16(x));
        },
        L.off, L.slots, L.wave);
    const uint32_t bse = base[p];
    // sort each owned row in place, replacing slots by target linear ids;
    // owned rows are the prefix [0, off[nvo]) of the slot array
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        const uint32_t o = L.off[v], k = L.off[v + 1] - o;
        row_ptr[d.vp + v] = static_cast<int32_t>(bse + o);
        sort_row(L.slots + o, k, [&](uint32_t j, uint32_t s) {
            L.slots[o + j] = IS_VV ? L.lin[lo16(s)] : L.lin[hi16(s)];
        });
    }
    if (threadIdx.x == 0) row_ptr[d.vp + d.nvo] = static_cast<int32_t>(bse + L.off[d.nvo]);
    __syncthreads();
    copy_out<T>(col + bse, L.slots, L.off[d.nvo]);
}

// VV in patch-local form: per patch, u16 row offsets of the owned vertices
// (nvo + 1, relative to the patch's entry base) at roff + rbase[p], and u16
// local neighbour ids at nbr + nbase[p], rows in ascending local-edge order.
// This is the query engine's native (RXMesh-style) encoding of the relation;
// iterative kernels consume it with half the index bytes of an int32 CSR.
template <uint32_t T>
__global__ __launch_bounds__(T) void k_query2_vv_local(DevMesh m, const uint32_t* __restrict__ rbase,
                                                       const uint32_t* __restrict__ nbase,
                                                       uint16_t* __restrict__ roff, uint16_t* __restrict__ nbr)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = blockIdx.x;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanQ2    L = plan_q2(b, 0, d.nv, d.ne, d.nf);
    stage_u32(L.ev, m.ev + d.e_off, d.ne);
    for (uint32_t c = threadIdx.x; c <= d.nvo; c += T) L.off[c] = 0;
    __syncthreads();
    fill_columns<T, true>(
        2 * d.ne, d.nvo,
        [&](uint32_t i) {
            const uint32_t x = L.ev[i >> 1];
            return (i & 1u) ? hi16(x) : lo16(x);
        },
        [&](uint32_t i) {
            const uint32_t x = L.ev[i >> 1];
            return ((i >> 1) << 16) | ((i & 1u) ? lo16(x) : hi16(x));
        },
        L.off, L.slots, L.wave);
    uint16_t* ro = roff + rbase[p];
    uint16_t* nb = nbr + nbase[p];
    for (uint32_t v = threadIdx.x; v <= d.nvo; v += T) ro[v] = static_cast<uint16_t>(L.off[v]);
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        const uint32_t o = L.off[v], k = L.off[v + 1] - o;
        sort_row(L.slots + o, k, [&](uint32_t j, uint32_t s) { nb[o + j] = static_cast<uint16_t>(lo16(s)); });
    }
}

// Patch-local face plan: per patch, the FV table of every local face (3 u16
// per face at fvl + fbase[p]) and, for every owned vertex, its VF row of
// (local face << 2 | corner) slots in ascending order (u16 row offsets at
// roff + rbase[p], entries at vf + nbase[p]). Corner c of face f is the
// vertex at fvl[3 f + c]. Requires nf < 16384 (checked on the host).
template <uint32_t T>
__global__ __launch_bounds__(T) void k_query2_vf_local(DevMesh m, const uint32_t* __restrict__ fbase,
                                                       const uint32_t* __restrict__ rbase,
                                                       const uint32_t* __restrict__ nbase,
                                                       uint16_t* __restrict__ fvl, uint16_t* __restrict__ roff,
                                                       uint16_t* __restrict__ vf)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = blockIdx.x;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanQ2    L = plan_q2(b, 1, d.nv, d.ne, d.nf);
    stage_topology<T>(L.ev, m.ev + d.e_off, d.ne, L.fe, m.fe + 3u * d.f_off, d.nf, nullptr, 0, 0, 0, nullptr);
    __syncthreads();
    uint16_t* fo = fvl + fbase[p];
    for (uint32_t i = threadIdx.x; i < 3 * d.nf; i += T)
        fo[i] = static_cast<uint16_t>(face_vertex(L.fe, L.ev, i / 3u, i % 3u));
    fill_columns<T>(
        3 * d.nf, d.nvo, [&](uint32_t i) { return face_vertex(L.fe, L.ev, i / 3u, i % 3u); },
        [&](uint32_t i) { return ((i / 3u) << 2) | (i % 3u); }, L.off, L.slots, L.wave);
    uint16_t* ro = roff + rbase[p];
    uint16_t* nb = vf + nbase[p];
    for (uint32_t v = threadIdx.x; v <= d.nvo; v += T) ro[v] = static_cast<uint16_t>(L.off[v]);
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        const uint32_t o = L.off[v], k = L.off[v + 1] - o;
        sort_row(L.slots + o, k, [&](uint32_t j, uint32_t s) { nb[o + j] = static_cast<uint16_t>(s); });
    }
}

// VF: FV from FE + EV, transposed; slot = face (the sort key itself).
// STABLE as in k_query2_vx.
template <uint32_t T, bool STABLE = false>
__global__ __launch_bounds__(T) void k_query2_vf(DevMesh m, const uint32_t* __restrict__ base,
                                                 int32_t* __restrict__ row_ptr, int32_t* __restrict__ col)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanQ2    L = plan_q2(b, 1, d.nv, d.ne, d.nf, STABLE ? T / 64u : 1u);
    stage_topology<T>(L.ev, m.ev + d.e_off, d.ne, L.fe, m.fe + 3u * d.f_off, d.nf, L.lin, d.nf, d.nfo, d.fp,
                      m.f_lin + d.f_off);
    if constexpr (STABLE) {
        const uint32_t bse = base[p];
        const uint32_t tot = sort_columns_stable<T>(
            3 * d.nf, d.nvo, [&](uint32_t i) { return face_vertex(L.fe, L.ev, i / 3u, i % 3u); },
            [&](uint32_t i) { return L.lin[i / 3u]; },
            [&](uint32_t c, uint32_t st) { row_ptr[d.vp + c] = static_cast<int32_t>(bse + st); }, L.off, L.slots,
            L.wave);
        if (threadIdx.x == 0) row_ptr[d.vp + d.nvo] = static_cast<int32_t>(bse + tot);
        copy_out<T>(col + bse, L.slots, tot);
        return;
    }
    for (uint32_t c = threadIdx.x; c <= d.nvo; c += T) L.off[c] = 0;
    __syncthreads();
    fill_columns<T, true>(
        3 * d.nf, d.nvo, [&](uint32_t i) { return face_vertex(L.fe, L.ev, i / 3u, i % 3u); },
        [&](uint32_t i) { return i / 3u; }, L.off, L.slots, L.wave);
    const uint32_t bse = base[p];
    for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
        const uint32_t o = L.off[v], k = L.off[v + 1] - o;
        row_ptr[d.vp + v] = static_cast<int32_t>(bse + o);
        sort_row(L.slots + o, k, [&](uint32_t j, uint32_t s) { L.slots[o + j] = L.lin[s]; });
    }
    if (threadIdx.x == 0) row_ptr[d.vp + d.nvo] = static_cast<int32_t>(bse + L.off[d.nvo]);
    __syncthreads();
    copy_out<T>(col + bse, L.slots, L.off[d.nvo]);
}

// VV and VF over linear ids in one launch: the two relations of a vertex
// share their staged topology (EV, FE and both linear-id tables arrive in
// one round trip), then the VV transpose and the VF transpose run in turn
// in one slot array; each CSR is written exactly as k_query2_vx<T, true> and
// k_query2_vf<T> write it. The default form, k_query2_vv_vf1, sorts both
// relations in one counting sort instead.
struct PlanVVF
{
    uint32_t *ev, *off, *slots, *wave, *lin_v, *lin_f;
    uint16_t* fe;
};
__host__ __device__ inline PlanVVF plan_vvf(Bump& b, uint32_t nv, uint32_t ne, uint32_t nf, uint32_t hist_mul = 1)
{
    PlanVVF p{};
    p.ev    = b.alloc<uint32_t>(r4(ne));
    p.fe    = b.alloc<uint16_t>(r4(3 * nf));
    p.off   = b.alloc<uint32_t>(nv * hist_mul + 1);
    p.slots = b.alloc<uint32_t>(2 * ne > 3 * nf ? 2 * ne : 3 * nf);
    p.wave  = b.alloc<uint32_t>(20);
    p.lin_v = b.alloc<uint32_t>(nv);
    p.lin_f = b.alloc<uint32_t>(nf);
    return p;
}

// LDS carve of the fused launch's LDS-atomic form: u16 slots (an entry index
// for VV, a local face for VF: both below 2^16), the histogram over owned
// vertices only. Sorted rows are turned into linear ids as they are copied
// out, so no 32-bit staging is needed and more workgroups fit on a CU.
struct PlanVVF16
{
    uint32_t *ev, *off, *wave, *lin_v, *lin_f;
    uint16_t *fe, *slots;
};
__host__ __device__ inline PlanVVF16 plan_vvf16(Bump& b, uint32_t nv, uint32_t nvo, uint32_t ne, uint32_t nf)
{
    PlanVVF16 p{};
    p.ev    = b.alloc<uint32_t>(r4(ne));
    p.fe    = b.alloc<uint16_t>(r4(3 * nf));
    p.off   = b.alloc<uint32_t>(nvo + 1);
    p.slots = b.alloc<uint16_t>(r4(2 * ne > 3 * nf ? 2 * ne : 3 * nf));
    p.wave  = b.alloc<uint32_t>(20);
    p.lin_v = b.alloc<uint32_t>(nv);
    p.lin_f = b.alloc<uint32_t>(nf);
    return p;
}

// Staging of the fused launch: EV, FE, vertex and face linear ids, the face
// ids loaded ahead of the staging's LDS stores.
template <uint32_t T>
__device__ __forceinline__ void stage_vv_vf(const DevMesh& m, const PatchDesc& d, uint32_t* ev, uint16_t* fe,
                                            uint32_t* lin_v, uint32_t* lin_f)
{
    constexpr uint32_t RL = 3;
    const uint32_t*    fs = m.f_lin + d.f_off;
    uint32_t           fl[RL];
#pragma unroll
    for (uint32_t r = 0; r < RL; ++r) {
        const uint32_t l = threadIdx.x + r * T;
        fl[r]            = (l < d.nfo) ? d.fp + l : ((l < d.nf) ? fs[l] : 0u);
    }
    stage_topology<T>(ev, m.ev + d.e_off, d.ne, fe, m.fe + 3u * d.f_off, d.nf, lin_v, d.nv, d.nvo, d.vp,
                      m.v_lin + d.v_off);
#pragma unroll
    for (uint32_t r = 0; r < RL; ++r)
        if (threadIdx.x + r * T < d.nf) lin_f[threadIdx.x + r * T] = fl[r];
    for (uint32_t l = threadIdx.x + RL * T; l < d.nf; l += T) lin_f[l] = (l < d.nfo) ? d.fp + l : fs[l];
}

template <uint32_t T, bool STABLE = false>
__global__ __launch_bounds__(T) void k_query2_vv_vf(DevMesh m, const uint32_t* __restrict__ vv_base,
                                                    const uint32_t* __restrict__ vf_base,
                                                    int32_t* __restrict__ vv_row, int32_t* __restrict__ vv_col,
                                                    int32_t* __restrict__ vf_row, int32_t* __restrict__ vf_col)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    if constexpr (STABLE) {
        const PlanVVF L = plan_vvf(b, d.nv, d.ne, d.nf, T / 64u);
        stage_vv_vf<T>(m, d, L.ev, L.fe, L.lin_v, L.lin_f);
        {
            const uint32_t bse = vv_base[p];
            const uint32_t tot = sort_columns_stable<T>(
                2 * d.ne, d.nvo, [&](uint32_t i) { return ev_end(L.ev, i); },
                [&](uint32_t i) { return L.lin_v[ev_other(L.ev, i)]; },
                [&](uint32_t c, uint32_t st) { vv_row[d.vp + c] = static_cast<int32_t>(bse + st); }, L.off,
                L.slots, L.wave);
            if (threadIdx.x == 0) vv_row[d.vp + d.nvo] = static_cast<int32_t>(bse + tot);
            copy_out<T>(vv_col + bse, L.slots, tot);
        }
        __syncthreads();  // the VV slots are read out before VF reuses them
        {
            const uint32_t bse = vf_base[p];
            const uint32_t tot = sort_columns_stable<T>(
                3 * d.nf, d.nvo, [&](uint32_t i) { return face_vertex(L.fe, L.ev, i / 3u, i % 3u); },
                [&](uint32_t i) { return L.lin_f[i / 3u]; },
                [&](uint32_t c, uint32_t st) { vf_row[d.vp + c] = static_cast<int32_t>(bse + st); }, L.off,
                L.slots, L.wave);
            if (threadIdx.x == 0) vf_row[d.vp + d.nvo] = static_cast<int32_t>(bse + tot);
            copy_out<T>(vf_col + bse, L.slots, tot);
        }
        return;
    }
    const PlanVVF16 L = plan_vvf16(b, d.nv, d.nvo, d.ne, d.nf);
    stage_vv_vf<T>(m, d, L.ev, L.fe, L.lin_v, L.lin_f);
    for (uint32_t c = threadIdx.x; c <= d.nvo; c += T) L.off[c] = 0;
    __syncthreads();
    // VV: slot = entry index i (edge i >> 1, endpoint i & 1), so a row sorted
    // by slot lists its edges in ascending local id
    fill_columns<T, true>(
        2 * d.ne, d.nvo, [&](uint32_t i) { return ev_end(L.ev, i); }, [&](uint32_t i) { return i; }, L.off,
        L.slots, L.wave);
    {
        const uint32_t bse = vv_base[p];
        for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
            const uint32_t o = L.off[v], k = L.off[v + 1] - o;
            vv_row[d.vp + v] = static_cast<int32_t>(bse + o);
            sort_row(L.slots + o, k, [&](uint32_t j, uint32_t s) { L.slots[o + j] = static_cast<uint16_t>(s); });
        }
        if (threadIdx.x == 0) vv_row[d.vp + d.nvo] = static_cast<int32_t>(bse + L.off[d.nvo]);
        __syncthreads();
        int32_t* dst = vv_col + bse;
        for (uint32_t k = threadIdx.x; k < L.off[d.nvo]; k += T)
            dst[k] = static_cast<int32_t>(L.lin_v[ev_other(L.ev, L.slots[k])]);
    }
    __syncthreads();  // the VV slots and offsets are read out before VF reuses them
    for (uint32_t c = threadIdx.x; c <= d.nvo; c += T) L.off[c] = 0;
    __syncthreads();
    // VF: slot = local face
    fill_columns<T, true>(
        3 * d.nf, d.nvo, [&](uint32_t i) { return face_vertex(L.fe, L.ev, i / 3u, i % 3u); },
        [&](uint32_t i) { return i / 3u; }, L.off, L.slots, L.wave);
    {
        const uint32_t bse = vf_base[p];
        for (uint32_t v = threadIdx.x; v < d.nvo; v += T) {
            const uint32_t o = L.off[v], k = L.off[v + 1] - o;
            vf_row[d.vp + v] = static_cast<int32_t>(bse + o);
            sort_row(L.slots + o, k, [&](uint32_t j, uint32_t s) { L.slots[o + j] = static_cast<uint16_t>(s); });
        }
        if (threadIdx.x == 0) vf_row[d.vp + d.nvo] = static_cast<int32_t>(bse + L.off[d.nvo]);
        __syncthreads();
        int32_t* dst = vf_col + bse;
        for (uint32_t k = threadIdx.x; k < L.off[d.nvo]; k += T) dst[k] = static_cast<int32_t>(L.lin_f[L.slots[k]]);
    }
}

// ---- fused VV + VF in one counting sort ------------------------------------
// Sorting networks for rows of up to 4, 6 and 8 entries (5, 12 and 19
// comparators; each sorts every 0/1 input, which by the 0-1 principle makes
// it a sorting network).
template <uint32_t N>
__device__ __forceinline__ void net_sort(uint32_t (&a)[8])
{
    if constexpr (N == 4) {
        cswap(a[0], a[1]); cswap(a[2], a[3]);
        cswap(a[0], a[2]); cswap(a[1], a[3]);
        cswap(a[1], a[2]);
    } else if constexpr (N == 6) {
        cswap(a[0], a[5]); cswap(a[1], a[3]); cswap(a[2], a[4]);
        cswap(a[1], a[2]); cswap(a[3], a[4]);
        cswap(a[0], a[3]); cswap(a[2], a[5]);
        cswap(a[0], a[1]); cswap(a[2], a[3]); cswap(a[4], a[5]);
        cswap(a[1], a[2]); cswap(a[3], a[4]);
    } else {
        cswap(a[0], a[1]); cswap(a[2], a[3]); cswap(a[4], a[5]); cswap(a[6], a[7]);
        cswap(a[0], a[2]); cswap(a[1], a[3]); cswap(a[4], a[6]); cswap(a[5], a[7]);
        cswap(a[1], a[2]); cswap(a[5], a[6]);
        cswap(a[0], a[4]); cswap(a[1], a[5]); cswap(a[2], a[6]); cswap(a[3], a[7]);
        cswap(a[2], a[4]); cswap(a[3], a[5]);
        cswap(a[1], a[2]); cswap(a[3], a[4]); cswap(a[5], a[6]);
    }
}

// LDS carve of the one-pass form: EV, the column starts of both relations
// (VV rows are columns [0, nvo), VF rows [nvo, 2 nvo)), u16 slots for the
// owned rows' entries (nslots: the largest VV + VF entry count of a patch)
// and the linear ids of the ribbon (nrv vertices, nrf faces at most; owned
// ids follow from the prefixes). FE is not staged.
struct PlanVVF1
{
    uint32_t *ev, *off, *wave, *rv, *rf;
    uint16_t* slots;
};
__host__ __device__ inline PlanVVF1 plan_vvf1(Bump& b, uint32_t nvo, uint32_t ne, uint32_t nslots, uint32_t nrv,
                                              uint32_t nrf)
{
    PlanVVF1 p{};
    p.ev    = b.alloc<uint32_t>(r4(ne));
    p.off   = b.alloc<uint32_t>(2 * nvo + 1);
    p.slots = b.alloc<uint16_t>(nslots);
    p.wave  = b.alloc<uint32_t>(20);
    p.rv    = b.alloc<uint32_t>(nrv ? nrv : 1u);  // at least one: the column pass reads clamped indices
    p.rf    = b.alloc<uint32_t>(nrf ? nrf : 1u);
    return p;
}

// The fused VV + VF launch in one counting sort (the default form):
//  - keys per edge (both endpoints of one EV word) and per face (three
//    corners), not per entry; each thread's FE words come from HBM straight
//    into registers with the staging's loads, and its keys stay in registers
//    from the histogram to the fill;
//  - one histogram, one scan and one fill over the combined column space
//    (VV rows then VF rows), so the VF slots follow the VV slots;
//  - each owned row sorted in registers by the smallest network that covers
//    the longest row of its wave (a wave-uniform choice), written back to
//    LDS, then one coalesced pass writes both CSRs, turning entries into
//    linear ids (VV slot = EV's u16 index of the other endpoint, 2 e + 1 -
//    end: below 2^16 needs ne <= 2^15; VF slot = local face).
// Needs ne <= RE T and nf <= RF T for the register path; larger patches take
// plain loops. Same CSRs as k_query2_vv_vf.
template <uint32_t T>
__global__ __launch_bounds__(T) void k_query2_vv_vf1(DevMesh m, uint32_t nslots, uint32_t nrv, uint32_t nrf,
                                                     const uint32_t* __restrict__ vv_base,
                                                     const uint32_t* __restrict__ vf_base,
                                                     int32_t* __restrict__ vv_row, int32_t* __restrict__ vv_col,
                                                     int32_t* __restrict__ vf_row, int32_t* __restrict__ vf_col)
{
    constexpr uint32_t RE = (1536u + T - 1u) / T;  // edges held in registers
    constexpr uint32_t RF = (1024u + T - 1u) / T;  // faces held in registers
    extern __shared__ uint4 smem[];
    const uint32_t  p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanVVF1  L    = plan_vvf1(b, d.nvo, d.ne, nslots, nrv, nrf);
    const uint32_t  nvo  = d.nvo, t = threadIdx.x;
    const uint32_t* vl   = m.v_lin + d.v_off;
    const uint32_t* fl   = m.f_lin + d.f_off;
    const uint16_t* fs   = m.fe + 3u * d.f_off;
    const bool      held = d.ne <= RE * T && d.nf <= RF * T;
    // global loads first: this thread's FE words, the ribbon ids, EV
    uint32_t fw[RF][3];
    if (held) {
#pragma unroll
        for (uint32_t r = 0; r < RF; ++r) {
            const uint32_t f = t + r * T;
#pragma unroll
            for (uint32_t j = 0; j < 3; ++j) fw[r][j] = (f < d.nf) ? fs[3 * f + j] : 0u;
        }
    }
    const uint32_t nr = d.nv - nvo, mr = d.nf - d.nfo;
    const uint32_t x0 = (t < nr) ? vl[nvo + t] : 0u, y0 = (t < mr) ? fl[d.nfo + t] : 0u;
    stage_topology<T>(L.ev, m.ev + d.e_off, d.ne, nullptr, nullptr, 0, nullptr, 0, 0, 0, nullptr);
    if (t < nr) L.rv[t] = x0;
    if (t < mr) L.rf[t] = y0;
    for (uint32_t i = t + T; i < nr; i += T) L.rv[i] = vl[nvo + i];
    for (uint32_t i = t + T; i < mr; i += T) L.rf[i] = fl[d.nfo + i];
    for (uint32_t c = t; c <= 2 * nvo; c += T) L.off[c] = 0;
    __syncthreads();
    // VV entries of edge e are its endpoints (column = local vertex), VF
    // entries of face f its corners (column = nvo + local vertex); entries
    // of ribbon vertices (>= nvo) have no row and are dropped
    auto fe_end = [&](uint32_t x) {  // local vertex an FE word starts at
        const uint32_t e = L.ev[x >> 1];
        return (x & 1u) ? hi16(e) : lo16(e);
    };
    uint32_t xe[RE], kf[RF][3];
    if (held) {
#pragma unroll
        for (uint32_t r = 0; r < RE; ++r) {
            const uint32_t e = t + r * T;
            const uint32_t x = L.ev[min(e, d.ne - 1u)];
            xe[r]            = (e < d.ne) ? x : 0xFFFFFFFFu;
        }
#pragma unroll
        for (uint32_t r = 0; r < RF; ++r)
#pragma unroll
            for (uint32_t j = 0; j < 3; ++j)  // the FE word of a face past nf is 0: a valid word
                kf[r][j] = (t + r * T < d.nf) ? fe_end(fw[r][j]) : 0xFFFFu;
#pragma unroll
        for (uint32_t r = 0; r < RE; ++r) {
            if (lo16(xe[r]) < nvo) atomicAdd(&L.off[lo16(xe[r]) + 1u], 1u);
            if (hi16(xe[r]) < nvo) atomicAdd(&L.off[hi16(xe[r]) + 1u], 1u);
        }
#pragma unroll
        for (uint32_t r = 0; r < RF; ++r)
#pragma unroll
            for (uint32_t j = 0; j < 3; ++j)
                if (kf[r][j] < nvo) atomicAdd(&L.off[nvo + kf[r][j] + 1u], 1u);
    } else {
        for (uint32_t e = t; e < d.ne; e += T) {
            const uint32_t x = L.ev[e];
            if (lo16(x) < nvo) atomicAdd(&L.off[lo16(x) + 1u], 1u);
            if (hi16(x) < nvo) atomicAdd(&L.off[hi16(x) + 1u], 1u);
        }
        for (uint32_t f = t; f < d.nf; f += T)
            for (uint32_t j = 0; j < 3; ++j) {
                const uint32_t v = fe_end(fs[3 * f + j]);
                if (v < nvo) atomicAdd(&L.off[nvo + v + 1u], 1u);
            }
    }
    scan_excl<T>(L.off + 1, 2 * nvo, L.wave);
    // fill: off[c + 1] advances from the start of column c to that of c + 1;
    // a VV slot is the u16 index in EV of the row's other endpoint (2 e + 1
    // for the row of lo16, 2 e for that of hi16), which sorts like e
    if (held) {
#pragma unroll
        for (uint32_t r = 0; r < RE; ++r) {
            const uint32_t i = 2u * (t + r * T);
            if (lo16(xe[r]) < nvo) L.slots[atomicAdd(&L.off[lo16(xe[r]) + 1u], 1u)] = static_cast<uint16_t>(i + 1u);
            if (hi16(xe[r]) < nvo) L.slots[atomicAdd(&L.off[hi16(xe[r]) + 1u], 1u)] = static_cast<uint16_t>(i);
        }
#pragma unroll
        for (uint32_t r = 0; r < RF; ++r)
#pragma unroll
            for (uint32_t j = 0; j < 3; ++j)
                if (kf[r][j] < nvo)
                    L.slots[atomicAdd(&L.off[nvo + kf[r][j] + 1u], 1u)] = static_cast<uint16_t>(t + r * T);
    } else {
        for (uint32_t e = t; e < d.ne; e += T) {
            const uint32_t x = L.ev[e];
            if (lo16(x) < nvo) L.slots[atomicAdd(&L.off[lo16(x) + 1u], 1u)] = static_cast<uint16_t>(2u * e + 1u);
            if (hi16(x) < nvo) L.slots[atomicAdd(&L.off[hi16(x) + 1u], 1u)] = static_cast<uint16_t>(2u * e);
        }
        for (uint32_t f = t; f < d.nf; f += T)
            for (uint32_t j = 0; j < 3; ++j) {
                const uint32_t v = fe_end(fs[3 * f + j]);
                if (v < nvo) L.slots[atomicAdd(&L.off[nvo + v + 1u], 1u)] = static_cast<uint16_t>(f);
            }
    }
    __syncthreads();
    const uint32_t nvv = L.off[nvo], ntot = L.off[2 * nvo];
    const uint32_t bvv = vv_base[p], bvf = vf_base[p] - nvv;  // VF slots follow the VV slots (32-bit: may wrap)
    if (t == 0) {
        vv_row[d.vp + nvo] = static_cast<int32_t>(bvv + nvv);
        vf_row[d.vp + nvo] = static_cast<int32_t>(bvf + ntot);
    }
    for (uint32_t c = t; c < 2 * nvo; c += T) {
        const bool     isvv = c < nvo;
        const uint32_t o = L.off[c], k = L.off[c + 1] - o;
        (isvv ? vv_row : vf_row)[d.vp + (isvv ? c : c - nvo)] = static_cast<int32_t>((isvv ? bvv : bvf) + o);
        auto     put = [&](uint32_t j, uint32_t s) { L.slots[o + j] = static_cast<uint16_t>(s); };
        uint32_t a[8];
        if (!__builtin_amdgcn_ballot_w64(k > 4u)) {
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) a[j] = (j < k) ? L.slots[o + j] : 0xFFFFFFFFu;
            net_sort<4>(a);
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j)
                if (j < k) put(j, a[j]);
        } else if (!__builtin_amdgcn_ballot_w64(k > 6u)) {
#pragma unroll
            for (uint32_t j = 0; j < 6; ++j) a[j] = (j < k) ? L.slots[o + j] : 0xFFFFFFFFu;
            net_sort<6>(a);
#pragma unroll
            for (uint32_t j = 0; j < 6; ++j)
                if (j < k) put(j, a[j]);
        } else if (k <= 8u) {
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j) a[j] = (j < k) ? L.slots[o + j] : 0xFFFFFFFFu;
            net_sort<8>(a);
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j)
                if (j < k) put(j, a[j]);
        } else {
            sort_row(L.slots + o, k, put);
        }
    }
    __syncthreads();
    // columns: a ribbon id read at a clamped index (in range for every
    // entry), selected where the element is not owned; no branch per entry
    const uint16_t* ev16 = reinterpret_cast<const uint16_t*>(L.ev);
    for (uint32_t i = t; i < nvv; i += T) {
        const uint32_t w = ev16[L.slots[i]];
        const uint32_t r = L.rv[max(w, nvo) - nvo];
        vv_col[bvv + i]  = static_cast<int32_t>((w < nvo) ? d.vp + w : r);
    }
    for (uint32_t i = nvv + t; i < ntot; i += T) {  // bvf + i: 32-bit sum, back in range
        const uint32_t s = L.slots[i];
        const uint32_t r = L.rf[max(s, d.nfo) - d.nfo];
        vf_col[bvf + i]  = static_cast<int32_t>((s < d.nfo) ? d.fp + s : r);
    }
}

// ---- one relation in one counting sort ---------------------------------------
// VV, VE, VF or EF in the form of k_query2_vv_vf1: keys per edge (VV / VE:
// both endpoints of one EV word) or per face (VF: three corners; EF: three
// edges; FE words loaded into registers), held in registers from histogram
// to fill; only the ribbon's linear ids of the output's kind are staged
// (owned ids follow from the prefixes); each row sorted in registers by the
// smallest network that covers the longest row of its wave; one coalesced
// write. Columns are the owned vertices (VV / VE / VF) or edges (EF). EF
// needs no EV: an FE word names its edge. Same CSR as k_query2_vx /
// k_query2_vf / k_query2_ef.
enum : uint32_t { kX1VV = 0, kX1VE = 1, kX1VF = 2, kX1EF = 3 };

struct PlanX1
{
    uint32_t *ev, *off, *wave, *rid;
    uint16_t* slots;
};
// ncol: the largest column count of a patch (owned vertices, or owned edges
// for EF); ne: its largest edge count (0 for EF: no EV); nslots: the largest
// entry count; nrid: the largest ribbon count of the output's element kind
__host__ __device__ inline PlanX1 plan_x1(Bump& b, uint32_t ncol, uint32_t ne, uint32_t nslots, uint32_t nrid)
{
    PlanX1 p{};
    p.ev    = b.alloc<uint32_t>(r4(ne));
    p.off   = b.alloc<uint32_t>(ncol + 1);
    p.slots = b.alloc<uint16_t>(nslots);
    p.wave  = b.alloc<uint32_t>(20);
    p.rid   = b.alloc<uint32_t>(nrid ? nrid : 1u);  // at least one: the column pass reads clamped indices
    return p;
}

template <uint32_t T, uint32_t REL>
__global__ __launch_bounds__(T) void k_query2_x1(DevMesh m, uint32_t nslots, uint32_t nrid,
                                                 const uint32_t* __restrict__ base, int32_t* __restrict__ row_ptr,
                                                 int32_t* __restrict__ col)
{
    constexpr bool     EFK = REL == kX1EF;          // EF: keys are the faces' edges
    constexpr bool     FK  = REL == kX1VF || EFK;  // keys per face
    constexpr uint32_t RE = (1536u + T - 1u) / T;  // edges held in registers
    constexpr uint32_t RF = (1024u + T - 1u) / T;  // faces held in registers
    extern __shared__ uint4 smem[];
    const uint32_t  p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const uint32_t  ncol = EFK ? d.neo : d.nvo, rp0 = EFK ? d.ep : d.vp;  // columns, their linear-id prefix
    const PlanX1    L    = plan_x1(b, ncol, EFK ? 0u : d.ne, nslots, nrid);
    const uint32_t  nvo  = d.nvo, t = threadIdx.x;
    const uint16_t* fs   = m.fe + 3u * d.f_off;
    const bool      held = FK ? d.nf <= RF * T : d.ne <= RE * T;
    // global loads first: this thread's FE words (VF), the ribbon ids, EV
    uint32_t fw[RF][3];
    if constexpr (FK) {
        if (held) {
#pragma unroll
            for (uint32_t r = 0; r < RF; ++r) {
                const uint32_t f = t + r * T;
#pragma unroll
                for (uint32_t j = 0; j < 3; ++j) fw[r][j] = (f < d.nf) ? fs[3 * f + j] : 0u;
            }
        }
    }
    // ribbon ids of the output's elements: vertices (VV), edges (VE), faces (VF / EF)
    const uint32_t* rs = (REL == kX1VV) ? m.v_lin + d.v_off + nvo
                                        : ((REL == kX1VE) ? m.e_lin + d.e_off + d.neo : m.f_lin + d.f_off + d.nfo);
    const uint32_t  nr = (REL == kX1VV) ? d.nv - nvo : ((REL == kX1VE) ? d.ne - d.neo : d.nf - d.nfo);
    const uint32_t  x0 = (t < nr) ? rs[t] : 0u;
    if constexpr (!EFK)
        stage_topology<T>(L.ev, m.ev + d.e_off, d.ne, nullptr, nullptr, 0, nullptr, 0, 0, 0, nullptr);
    if (t < nr) L.rid[t] = x0;
    for (uint32_t i = t + T; i < nr; i += T) L.rid[i] = rs[i];
    for (uint32_t c = t; c <= ncol; c += T) L.off[c] = 0;
    __syncthreads();
    // the column of an FE word: the local vertex it starts at (VF), its edge (EF)
    auto fe_end = [&](uint32_t x) {
        if constexpr (EFK) {
            return x >> 1;
        } else {
            const uint32_t e = L.ev[x >> 1];
            return (x & 1u) ? hi16(e) : lo16(e);
        }
    };
    uint32_t xe[FK ? 1u : RE], kf[FK ? RF : 1u][3];
    if constexpr (FK) {
        if (held) {
#pragma unroll
            for (uint32_t r = 0; r < RF; ++r)
#pragma unroll
                for (uint32_t j = 0; j < 3; ++j)  // the FE word of a face past nf is 0: a valid word
                    kf[r][j] = (t + r * T < d.nf) ? fe_end(fw[r][j]) : 0xFFFFu;
#pragma unroll
            for (uint32_t r = 0; r < RF; ++r)
#pragma unroll
                for (uint32_t j = 0; j < 3; ++j)
                    if (kf[r][j] < ncol) atomicAdd(&L.off[kf[r][j] + 1u], 1u);
        } else {
            for (uint32_t f = t; f < d.nf; f += T)
                for (uint32_t j = 0; j < 3; ++j) {
                    const uint32_t v = fe_end(fs[3 * f + j]);
                    if (v < ncol) atomicAdd(&L.off[v + 1u], 1u);
                }
        }
    } else {
        if (held) {
#pragma unroll
            for (uint32_t r = 0; r < RE; ++r) {
                const uint32_t e = t + r * T;
                const uint32_t x = L.ev[min(e, d.ne - 1u)];
                xe[r]            = (e < d.ne) ? x : 0xFFFFFFFFu;
            }
#pragma unroll
            for (uint32_t r = 0; r < RE; ++r) {
                if (lo16(xe[r]) < nvo) atomicAdd(&L.off[lo16(xe[r]) + 1u], 1u);
                if (hi16(xe[r]) < nvo) atomicAdd(&L.off[hi16(xe[r]) + 1u], 1u);
            }
        } else {
            for (uint32_t e = t; e < d.ne; e += T) {
                const uint32_t x = L.ev[e];
                if (lo16(x) < nvo) atomicAdd(&L.off[lo16(x) + 1u], 1u);
                if (hi16(x) < nvo) atomicAdd(&L.off[hi16(x) + 1u], 1u);
            }
        }
    }
    scan_excl<T>(L.off + 1, ncol, L.wave);
    // fill: off[c + 1] advances from the start of column c to that of c + 1;
    // slot = EV's u16 index of the row's other endpoint, 2 e + 1 - end (VV /
    // VE: sorts like e), or local face (VF / EF)
    if constexpr (FK) {
        if (held) {
#pragma unroll
            for (uint32_t r = 0; r < RF; ++r)
#pragma unroll
                for (uint32_t j = 0; j < 3; ++j)
                    if (kf[r][j] < ncol)
                        L.slots[atomicAdd(&L.off[kf[r][j] + 1u], 1u)] = static_cast<uint16_t>(t + r * T);
        } else {
            for (uint32_t f = t; f < d.nf; f += T)
                for (uint32_t j = 0; j < 3; ++j) {
                    const uint32_t v = fe_end(fs[3 * f + j]);
                    if (v < ncol) L.slots[atomicAdd(&L.off[v + 1u], 1u)] = static_cast<uint16_t>(f);
                }
        }
    } else {
        if (held) {
#pragma unroll
            for (uint32_t r = 0; r < RE; ++r) {
                const uint32_t i = 2u * (t + r * T);
                if (lo16(xe[r]) < nvo)
                    L.slots[atomicAdd(&L.off[lo16(xe[r]) + 1u], 1u)] = static_cast<uint16_t>(i + 1u);
                if (hi16(xe[r]) < nvo) L.slots[atomicAdd(&L.off[hi16(xe[r]) + 1u], 1u)] = static_cast<uint16_t>(i);
            }
        } else {
            for (uint32_t e = t; e < d.ne; e += T) {
                const uint32_t x = L.ev[e];
                if (lo16(x) < nvo) L.slots[atomicAdd(&L.off[lo16(x) + 1u], 1u)] = static_cast<uint16_t>(2u * e + 1u);
                if (hi16(x) < nvo) L.slots[atomicAdd(&L.off[hi16(x) + 1u], 1u)] = static_cast<uint16_t>(2u * e);
            }
        }
    }
    __syncthreads();
    const uint32_t ntot = L.off[ncol], bse = base[p];
    if (t == 0) row_ptr[rp0 + ncol] = static_cast<int32_t>(bse + ntot);
    for (uint32_t v = t; v < ncol; v += T) {
        const uint32_t o = L.off[v], k = L.off[v + 1] - o;
        row_ptr[rp0 + v] = static_cast<int32_t>(bse + o);
        auto     put = [&](uint32_t j, uint32_t s) { L.slots[o + j] = static_cast<uint16_t>(s); };
        uint32_t a[8];
        if (!__builtin_amdgcn_ballot_w64(k > 4u)) {
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) a[j] = (j < k) ? L.slots[o + j] : 0xFFFFFFFFu;
            net_sort<4>(a);
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j)
                if (j < k) put(j, a[j]);
        } else if (!__builtin_amdgcn_ballot_w64(k > 6u)) {
#pragma unroll
            for (uint32_t j = 0; j < 6; ++j) a[j] = (j < k) ? L.slots[o + j] : 0xFFFFFFFFu;
            net_sort<6>(a);
#pragma unroll
            for (uint32_t j = 0; j < 6; ++j)
                if (j < k) put(j, a[j]);
        } else if (k <= 8u) {
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j) a[j] = (j < k) ? L.slots[o + j] : 0xFFFFFFFFu;
            net_sort<8>(a);
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j)
                if (j < k) put(j, a[j]);
        } else {
            sort_row(L.slots + o, k, put);
        }
    }
    __syncthreads();
    // columns: local element w of the output's kind, its ribbon id read at a
    // clamped index (in range for every entry) and selected where w is not
    // owned; no branch per entry
    const uint16_t* ev16 = reinterpret_cast<const uint16_t*>(L.ev);
    const uint32_t  no   = (REL == kX1VV) ? nvo : ((REL == kX1VE) ? d.neo : d.nfo);
    const uint32_t  pre  = (REL == kX1VV) ? d.vp : ((REL == kX1VE) ? d.ep : d.fp);
    for (uint32_t i = t; i < ntot; i += T) {
        const uint32_t s = L.slots[i];
        const uint32_t w = (REL == kX1VV) ? static_cast<uint32_t>(ev16[s]) : ((REL == kX1VE) ? s >> 1 : s);
        const uint32_t r = L.rid[max(w, no) - no];
        col[bse + i]     = static_cast<int32_t>((w < no) ? pre + w : r);
    }
}

// FF in one pass: the EF counting sort over every local edge (keys = the
// edges of the faces' FE words, held in registers), the edge rows sorted,
// then each owned face lists the faces of its three edges but itself (edges
// in FE order, faces ascending: the rows of k_query2_ef<T, true>) into LDS,
// and one coalesced pass writes the CSR, turning local faces into linear
// ids (ribbon faces' ids staged).
struct PlanFF1
{
    uint32_t *off, *wave, *cur, *rid;
    uint16_t *slots, *out;
};
// nout: the largest FF entry count of a patch; nrid: its largest ribbon face count
__host__ __device__ inline PlanFF1 plan_ff1(Bump& b, uint32_t ne, uint32_t nf, uint32_t nfo, uint32_t nout,
                                            uint32_t nrid)
{
    PlanFF1 p{};
    p.off   = b.alloc<uint32_t>(ne + 1);
    p.slots = b.alloc<uint16_t>(3 * nf);
    p.wave  = b.alloc<uint32_t>(20);
    p.cur   = b.alloc<uint32_t>(nfo + 1);
    p.out   = b.alloc<uint16_t>(nout);
    p.rid   = b.alloc<uint32_t>(nrid ? nrid : 1u);  // at least one: the column pass reads clamped indices
    return p;
}

template <uint32_t T>
__global__ __launch_bounds__(T) void k_query2_ff1(DevMesh m, uint32_t nout, uint32_t nrid,
                                                  const uint32_t* __restrict__ base, int32_t* __restrict__ row_ptr,
                                                  int32_t* __restrict__ col)
{
    constexpr uint32_t RF = (1024u + T - 1u) / T;  // faces held in registers
    extern __shared__ uint4 smem[];
    const uint32_t  p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanFF1   L    = plan_ff1(b, d.ne, d.nf, d.nfo, nout, nrid);
    const uint32_t  t    = threadIdx.x;
    const uint16_t* fs   = m.fe + 3u * d.f_off;
    const uint32_t* fl   = m.f_lin + d.f_off + d.nfo;
    const bool      held = d.nf <= RF * T;
    // global loads first: this thread's FE words, the ribbon faces' ids
    uint32_t fw[RF][3];
    if (held) {
#pragma unroll
        for (uint32_t r = 0; r < RF; ++r) {
            const uint32_t f = t + r * T;
#pragma unroll
            for (uint32_t j = 0; j < 3; ++j) fw[r][j] = (f < d.nf) ? fs[3 * f + j] : 0u;
        }
    }
    const uint32_t nr = d.nf - d.nfo;
    const uint32_t x0 = (t < nr) ? fl[t] : 0u;
    for (uint32_t c = t; c <= d.ne; c += T) L.off[c] = 0;
    if (t < nr) L.rid[t] = x0;
    for (uint32_t i = t + T; i < nr; i += T) L.rid[i] = fl[i];
    __syncthreads();
    // EF histogram over every local face's three edges (column = local edge)
    if (held) {
#pragma unroll
        for (uint32_t r = 0; r < RF; ++r)
            if (t + r * T < d.nf)
#pragma unroll
                for (uint32_t j = 0; j < 3; ++j) atomicAdd(&L.off[(fw[r][j] >> 1) + 1u], 1u);
    } else {
        for (uint32_t f = t; f < d.nf; f += T)
            for (uint32_t j = 0; j < 3; ++j) atomicAdd(&L.off[(fs[3 * f + j] >> 1) + 1u], 1u);
    }
    scan_excl<T>(L.off + 1, d.ne, L.wave);
    if (held) {
#pragma unroll
        for (uint32_t r = 0; r < RF; ++r)
            if (t + r * T < d.nf)
#pragma unroll
                for (uint32_t j = 0; j < 3; ++j)
                    L.slots[atomicAdd(&L.off[(fw[r][j] >> 1) + 1u], 1u)] = static_cast<uint16_t>(t + r * T);
    } else {
        for (uint32_t f = t; f < d.nf; f += T)
            for (uint32_t j = 0; j < 3; ++j)
                L.slots[atomicAdd(&L.off[(fs[3 * f + j] >> 1) + 1u], 1u)] = static_cast<uint16_t>(f);
    }
    __syncthreads();
    // every edge row ascending (two faces on a manifold edge: one compare)
    for (uint32_t e = t; e < d.ne; e += T) {
        const uint32_t o = L.off[e], k = L.off[e + 1] - o;
        if (k == 2u) {
            const uint16_t a0 = L.slots[o], a1 = L.slots[o + 1];
            if (a1 < a0) {
                L.slots[o]     = a1;
                L.slots[o + 1] = a0;
            }
        } else if (k > 2u) {
            sort_row(L.slots + o, k, [&](uint32_t j, uint32_t s) { L.slots[o + j] = static_cast<uint16_t>(s); });
        }
    }
    __syncthreads();
    // owned face f (edges e[0..2] in FE order): its FF row length, then its
    // row into LDS at its offset; the held faces' edges come from registers
    auto each_owned = [&](auto fn) {
#pragma unroll
        for (uint32_t r = 0; r < RF; ++r) {
            const uint32_t f = t + r * T;
            if (f < d.nfo) {
                uint32_t e[3];
#pragma unroll
                for (uint32_t j = 0; j < 3; ++j) e[j] = (held ? fw[r][j] : uint32_t(fs[3 * f + j])) >> 1;
                fn(f, e);
            }
        }
        for (uint32_t f = t + RF * T; f < d.nfo; f += T) {
            uint32_t e[3];
            for (uint32_t j = 0; j < 3; ++j) e[j] = uint32_t(fs[3 * f + j]) >> 1;
            fn(f, e);
        }
    };
    each_owned([&](uint32_t f, const uint32_t (&e)[3]) {
        uint32_t c = 0;
#pragma unroll
        for (uint32_t j = 0; j < 3; ++j) c += L.off[e[j] + 1] - L.off[e[j]] - 1u;
        L.cur[f] = c;
    });
    const uint32_t total = scan_excl<T>(L.cur, d.nfo, L.wave);
    const uint32_t bse   = base[p];
    if (t == 0) row_ptr[d.fp + d.nfo] = static_cast<int32_t>(bse + total);
    each_owned([&](uint32_t f, const uint32_t (&e)[3]) {
        uint32_t k        = L.cur[f];
        row_ptr[d.fp + f] = static_cast<int32_t>(bse + k);
#pragma unroll
        for (uint32_t j = 0; j < 3; ++j)
            for (uint32_t s = L.off[e[j]]; s < L.off[e[j] + 1]; ++s) {
                const uint32_t g = L.slots[s];
                if (g != f) L.out[k++] = static_cast<uint16_t>(g);
            }
    });
    __syncthreads();
    for (uint32_t i = t; i < total; i += T) {  // ribbon id at a clamped index, selected: no branch
        const uint32_t g = L.out[i];
        const uint32_t r = L.rid[max(g, d.nfo) - d.nfo];
        col[bse + i]     = static_cast<int32_t>((g < d.nfo) ? d.fp + g : r);
    }
}

// EF (IS_FF == false) or FF. STABLE as in k_query2_vx.
template <uint32_t T, bool IS_FF, bool STABLE = false>
__global__ __launch_bounds__(T) void k_query2_ef(DevMesh m, const uint32_t* __restrict__ base,
                                                 int32_t* __restrict__ row_ptr, int32_t* __restrict__ col)
{
    extern __shared__ uint4 smem[];
    const uint32_t  p = patch_of_block(m.P);  // XCD-contiguous patch ranges
    if (p >= m.P) return;
    const PatchDesc d = load_desc(m.patch, p);
    Bump            b(reinterpret_cast<char*>(smem));
    const PlanQ2    L = plan_q2(b, IS_FF ? 4 : 2, d.nv, d.ne, d.nf, STABLE ? T / 64u : 1u);
    stage_topology<T>(nullptr, nullptr, 0, L.fe, m.fe + 3u * d.f_off, d.nf, L.lin, d.nf, d.nfo, d.fp,
                      m.f_lin + d.f_off);
    if constexpr (STABLE) {
        constexpr uint32_t NW  = T / 64u;
        const uint32_t     bse = base[p];
        const uint32_t ncols = IS_FF ? d.ne : d.neo;
        const uint32_t tot   = sort_columns_stable<T>(
            3 * d.nf, ncols, [&](uint32_t i) { return static_cast<uint32_t>(L.fe[i] >> 1); },
            [&](uint32_t i) { return IS_FF ? i / 3u : L.lin[i / 3u]; },
            [&](uint32_t c, uint32_t st) {
                if (!IS_FF) row_ptr[d.ep + c] = static_cast<int32_t>(bse + st);
            },
            L.off, L.slots, L.wave);
        if constexpr (!IS_FF) {
            if (threadIdx.x == 0) row_ptr[d.ep + d.neo] = static_cast<int32_t>(bse + tot);
            copy_out<T>(col + bse, L.slots, tot);
        } else {
            // after the sort, table cell e * NW - 1 holds the start of edge e's row
            auto rs = [&](uint32_t e) { return e ? L.off[e * NW - 1u] : 0u; };
            for (uint32_t f = threadIdx.x; f < d.nfo; f += T) {
                uint32_t c = 0;
                for (uint32_t i = 0; i < 3; ++i) {
                    const uint32_t e = L.fe[3 * f + i] >> 1;
                    c += rs(e + 1) - rs(e) - 1;
                }
                L.cur[f] = c;
            }
            const uint32_t total = scan_excl<T>(L.cur, d.nfo, L.wave);
            for (uint32_t f = threadIdx.x; f < d.nfo; f += T) {
                uint32_t k        = bse + L.cur[f];
                row_ptr[d.fp + f] = static_cast<int32_t>(k);
                for (uint32_t i = 0; i < 3; ++i) {
                    const uint32_t e = L.fe[3 * f + i] >> 1;
                    for (uint32_t j = rs(e); j < rs(e + 1); ++j) {
                        const uint32_t g = L.slots[j];
                        if (g != f) col[k++] = static_cast<int32_t>(L.lin[g]);
                    }
                }
            }
            if (threadIdx.x == 0) row_ptr[d.fp + d.nfo] = static_cast<int32_t>(bse + total);
        }
        return;
    }
    for (uint32_t c = threadIdx.x; c <= (IS_FF ? d.ne : d.neo); c += T) L.off[c] = 0;
    __syncthreads();
    fill_columns<T, true>(
        3 * d.nf, IS_FF ? d.ne : d.neo, [&](uint32_t i) { return static_cast<uint32_t>(L.fe[i] >> 1); },
        [&](uint32_t i) { return i / 3u; }, L.off, L.slots, L.wave);
    const uint32_t bse = base[p];
    if constexpr (!IS_FF) {
        for (uint32_t e = threadIdx.x; e < d.neo; e += T) {
            const uint32_t o = L.off[e], k = L.off[e + 1] - o;
            row_ptr[d.ep + e] = static_cast<int32_t>(bse + o);
            sort_row(L.slots + o, k, [&](uint32_t j, uint32_t s) { L.slots[o + j] = L.lin[s]; });
        }
        if (threadIdx.x == 0) row_ptr[d.ep + d.neo] = static_cast<int32_t>(bse + L.off[d.neo]);
        __syncthreads();
        copy_out<T>(col + bse, L.slots, L.off[d.neo]);
    } else {
        // sort the EF rows of every edge touched by an owned face
        for (uint32_t e = threadIdx.x; e < d.ne; e += T) {
            const uint32_t o = L.off[e], k = L.off[e + 1] - o;
            if (k > 1) sort_row(L.slots + o, k, [&](uint32_t j, uint32_t s) { L.slots[o + j] = s; });
        }
        __syncthreads();
        // per owned face: (faces of its edges) - 1 each, scanned into L.cur
        for (uint32_t f = threadIdx.x; f < d.nfo; f += T) {
            uint32_t c = 0;
            for (uint32_t i = 0; i < 3; ++i) {
                const uint32_t e = L.fe[3 * f + i] >> 1;
                c += L.off[e + 1] - L.off[e] - 1;
            }
            L.cur[f] = c;
        }
        const uint32_t total = scan_excl<T>(L.cur, d.nfo, L.wave);
        for (uint32_t f = threadIdx.x; f < d.nfo; f += T) {
            uint32_t k = bse + L.cur[f];
            row_ptr[d.fp + f] = static_cast<int32_t>(k);
            for (uint32_t i = 0; i < 3; ++i) {
                const uint32_t e = L.fe[3 * f + i] >> 1;
                for (uint32_t j = L.off[e]; j < L.off[e + 1]; ++j) {
                    const uint32_t g = L.slots[j];
                    if (g != f) col[k++] = static_cast<int32_t>(L.lin[g]);
                }
            }
        }
        if (threadIdx.x == 0) row_ptr[d.fp + d.nfo] = static_cast<int32_t>(bse + total);
    }
}

}  // namespace rxm
