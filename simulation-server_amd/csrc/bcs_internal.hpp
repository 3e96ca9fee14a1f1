// bcs_internal.hpp -- declarations shared by the host sources of libbcs.
#pragma once

#include "../../include/bcs.h"
#include "../../include/bcs_scene.h"

#include <string>

namespace bcs {

// veinVertexMaxNeighbors, src/meta_factory/vein_factory.hpp:125
constexpr int kVeinSlots = 9;
// maxCudaStreams bound on the number of cell types, src/config/simulation.hpp:4
constexpr int kMaxTypes = 16;

extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);

}  // namespace bcs
