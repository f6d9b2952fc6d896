// zgl_bvh.h -- MTBVH construction on the device (zgl_bvh.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace zgl {

struct DeviceBvhStats {
    int levels = 0;      // rounds of the level-synchronous phase
    int64_t bigNodes = 0;
    int64_t chunks = 0;
    int smallTasks = 0;  // subtrees finished by one wave each
    int maxDepth = 0;
    int jumps = 0;       // pointer-doubling rounds of the threading
};

// Builds the reference's MTBVH (BVH::build + buildHitTable, src/accelerator/BVH.cpp:116-345)
// on the device from device-resident vertices (n_vertices x 3 floats) and indices
// (n_triangles x 3): node bounds ((2n-1) x 6 floats), the six hit tables (6 x (2n-1) x 3
// ints, optional) and the 32-byte traversal rows (6 x (2n-1) x 2 float4, optional),
// bit-identical to the host build.  Work memory is kept between builds.  Returns 0, or
// 1 (invalid input), 2 (HIP error), 3 (out of memory) with a message.  The stream is
// synchronised once per level of the big-node phase.
class DeviceBvh {
public:
    DeviceBvh() = default;
    DeviceBvh(const DeviceBvh&) = delete;
    DeviceBvh& operator=(const DeviceBvh&) = delete;
    ~DeviceBvh();
    int build(const float* dVerts, int64_t nV, const uint32_t* dIdx, int64_t nT, int subtree, float* dBounds,
              int32_t* dHit, void* dRows, hipStream_t st, std::string& err);
    void release();
    DeviceBvhStats stats;

private:
    void* work_ = nullptr;
    size_t workBytes_ = 0;
};

// Default size below which a subtree is finished by one wave.
constexpr int kDefaultBvhSubtree = 256;

}  // namespace zgl
