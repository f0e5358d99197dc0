// Device view of a patch mesh. A trivially copyable struct of raw device
// pointers passed by value to every kernel (the role of the reference's
// Context / PatchInfo, include/rxmesh/context.h:15-441 and
// include/rxmesh/patch_info.h:24-432), laid out as one slab per array with
// per-patch offsets instead of per-patch allocations.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace rxm {

// Per-patch scalars, 12 x u32 = 48 B, read with three 16-byte loads.
struct alignas(16) PatchDesc
{
    uint32_t nv, ne, nf;        // local counts incl. ribbon
    uint32_t nvo, neo, nfo;     // owned counts (owned ids are [0, n*o))
    uint32_t v_off, e_off, f_off;  // slab offsets (multiples of 4)
    uint32_t vp, ep, fp;        // linear-id prefix of the owned elements
};

struct DevMesh
{
    uint32_t nV, nE, nF, P;
    uint32_t max_nv, max_ne, max_nf;
    const PatchDesc* patch;   // P
    const uint32_t*  ev;      // e_off slab; lo16 = local(max gid), hi16 = local(min gid)
    const uint16_t*  fe;      // 3 * f_off slab; (local edge << 1) | dir
    const uint32_t*  v_lin;   // v_off slab; linear id of the owner's copy
    const uint32_t*  e_lin;
    const uint32_t*  f_lin;
    const uint32_t*  v_own;   // owner patch
    const uint32_t*  e_own;
    const uint32_t*  f_own;
    const uint32_t*  ltog_v;  // global id per local element
    const uint32_t*  ltog_e;
    const uint32_t*  ltog_f;
    const uint32_t*  v_l2g;   // linear -> global
    const uint32_t*  e_l2g;
    const uint32_t*  f_l2g;
    const uint32_t*  v_g2l;   // global -> linear
    const uint32_t*  e_g2l;
    const uint32_t*  f_g2l;
};

// 16-byte aligned bump allocator over dynamic LDS; the host runs the same
// carve with base = nullptr to size the launch.
struct Bump
{
    char*    base;
    uint32_t off;
    __host__ __device__ explicit Bump(char* b) : base(b), off(0) {}
    template <typename T>
    __host__ __device__ T* alloc(uint32_t n)
    {
        off = (off + 15u) & ~15u;
        T* p = reinterpret_cast<T*>(base + off);
        off += n * static_cast<uint32_t>(sizeof(T));
        return p;
    }
    __host__ __device__ uint32_t bytes() const { return (off + 15u) & ~15u; }
};

}  // namespace rxm
