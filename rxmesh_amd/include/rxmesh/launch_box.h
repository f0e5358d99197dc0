// Launch parameters computed by RXMeshStatic::prepare_launch_box (reference
// include/rxmesh/launch_box.h:13-20).
#pragma once
#include <cstddef>
#include <cstdint>

namespace rxmesh {

template <uint32_t blockThreads>
struct LaunchBox
{
    uint32_t       blocks = 0, num_registers_per_thread = 0;
    size_t         smem_bytes_dyn = 0, smem_bytes_static = 0;
    size_t         local_mem_per_thread = 0;
    const uint32_t num_threads = blockThreads;
};

}  // namespace rxmesh
