// Enumerations and small types of the RXMesh C++ API (names and values as in
// the reference's include/rxmesh/types.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rxmesh/glm_lite.h"
#include "rxmesh/util/template_names.h"

namespace rxmesh {

constexpr uint16_t INVALID16 = 0xFFFFu;
constexpr uint32_t INVALID32 = 0xFFFFFFFFu;
constexpr uint64_t INVALID64 = 0xFFFFFFFFFFFFFFFFull;

using locationT = uint32_t;
enum : locationT
{
    LOCATION_NONE = 0x00,
    HOST          = 0x01,
    DEVICE        = 0x02,
    LOCATION_ALL  = 0x0F,
};

// Attribute layouts of the reference. This implementation stores every
// attribute SoA over linear ids; AoS / AoSoA requests are accepted and
// stored the same way (the accessors hide the layout).
using layoutT = uint32_t;
enum : layoutT
{
    AoS   = 0x00,
    AoSoA = 0x01,
    SoA   = 0x02,
};

// reference include/rxmesh/types.h:113-129
enum class Op
{
    INVALID   = -1,
    V         = 0,
    E         = 1,
    F         = 2,
    VV        = 3,
    VE        = 4,
    VF        = 5,
    FV        = 6,
    FE        = 7,
    FF        = 8,
    EV        = 9,
    EE        = 10,
    EF        = 11,
    EVDiamond = 12,
};

template <typename T, int N>
using vec = glm::vec<N, T>;
template <typename T>
using vec1 = glm::vec<1, T>;
template <typename T>
using vec2 = glm::vec<2, T>;
template <typename T>
using vec3 = glm::vec<3, T>;
template <typename T>
using vec4 = glm::vec<4, T>;
using glm::cross;
using glm::dot;

using rx_coord_t = float;

}  // namespace rxmesh
