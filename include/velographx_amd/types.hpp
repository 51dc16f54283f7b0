// Basic types of the VeloGraphX interface (include/velographx/types.hpp).
#pragma once

#include <cstdint>

#include "vgx.h"

namespace velographx {

using VertexId = std::uint32_t;
using EdgeOffset = std::uint64_t;
inline constexpr std::uint32_t kUnreachable = VGX_UNREACHABLE;

}  // namespace velographx
