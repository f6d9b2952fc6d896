// zgl_build.h -- host-side producers of the scene buffers the kernels consume.
//
// These run once per scene on the CPU (not per sample) and produce the buffers in
// the reference's layout (src/core/Scene.h:12-27): MTBVH bounds + six threaded hit
// tables, light alias tables, the two-level environment alias table and the Sobol
// generator matrices.
#pragma once

#include <stdint.h>
#include <vector>

namespace zgl {

constexpr int32_t BvhLeafMask = (int32_t)0x80000000;

// Binned-SAH BVH (16 buckets, one triangle per leaf, DFS preorder) threaded six
// ways.  Follows BVH::build / quickBuild / buildHitTable
// (src/accelerator/BVH.cpp:116-144, :217-296, :298-345).
//   bounds:   (2n-1) nodes x {pMin.xyz, pMax.xyz}
//   hitTable: 6 tables x (2n-1) rows x {node, prim or -1, miss}
void buildMtbvh(const float* vertices, const uint32_t* indices, int64_t nTri,
                std::vector<float>& bounds, std::vector<int32_t>& hitTable);

// Vose/Walker alias table with two LIFO stacks (src/math/AliasTable.h:11-57).
void aliasBuild(const float* pdf, int32_t n, int32_t* alias, float* prob);

// One row (or the marginal column) of the environment alias table; leftover
// entries get probability 1 (src/core/EnvironmentMap.cpp:66-114).  Returns the
// un-normalised sum of the weights.
float envAliasRow(int32_t* alias, float* pdf, int32_t n, int32_t stride);

// Environment two-level table from an FP32 RGB image, (w+1) x h entries each;
// column w of row i holds the marginal of row i (src/core/EnvironmentMap.cpp:8-64).
// Returns the float sum; the reference hands the shaders its int truncation.
float envAliasBuild(const float* rgb, int32_t w, int32_t h, int32_t* alias, float* prob);

// 256 dims x 32 columns, expanded from the Joe-Kuo parameters.
const uint32_t* sobolMatrices();
uint32_t sobolSample(uint32_t index, int dim, uint32_t scramble);

// IEEE binary16 with round-to-nearest-even (the RGB16F upload of the env map).
uint16_t floatToHalf(float f);
float halfToFloat(uint16_t h);

}  // namespace zgl
