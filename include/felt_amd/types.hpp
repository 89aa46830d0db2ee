// Scalar and vector types of Felt's API (Common.hpp): positions are VecDi, real positions VecDf,
// distances float.  Shared by the host wrapper (Surface.hpp) and the device grid view
// (device_grid.hpp).
#pragma once

#include <array>
#include <cstddef>
#include <vector>

namespace Felt {

using Distance = float;
using LayerId = int;
using PosIdx = std::size_t;
using TupleIdx = int;
using ListIdx = std::size_t;
using PosIdxList = std::vector<PosIdx>;
template <int D>
using VecDi = std::array<int, D>;
template <int D>
using VecDf = std::array<float, D>;

}  // namespace Felt
