// zgl_bvh.hip -- MTBVH construction on the GPU, bit-identical to the host build.
//
// The host build (zgl_build.cpp, following src/accelerator/BVH.cpp:116-345) splits each
// node [l, r] of the primitive array with a 16-bucket binned SAH, partitions the range
// stably (left part in order, right part reversed), and lays the tree out in DFS preorder:
// node offset o, left child o + 1, right child o + 2 (split - l) + 2.  A node's result
// depends only on the primitives its range holds when it is processed, so nodes can be
// processed in any order that respects parent -> child; these kernels go level by level.
//
// Box reductions.  The host folds boxes with minG(x, y) = y < x ? y : x (and maxG alike),
// starting from {1e8, -1e8}: the result is the leftmost element attaining the minimum,
// which differs from any other only for +0 / -0.  Here each reduction is order-free:
//   * an integer min/max over a key in which -0 and +0 are both 0 and keys order like
//     the floats, giving the value;
//   * an integer min over (position << 1 | sign bit) of the elements equal to zero,
//     giving the sign of the leftmost zero when the extreme is zero.
// Bucket boxes are decoded exactly this way, and the prefix/suffix sweeps, surface
// areas and costs then run in the host's operation order (no FMA contraction).
//
// Phases.
//   1. k_prims: primitive bounds from the mesh and the root's centroid box.
//   2. Big nodes (at least `subtree` primitives), one level per round: k_map, k_bin
//      (per-chunk bucket accumulation, 2048 primitives per workgroup), k_split (one
//      wave per node: SAH choice, node bound), k_scan (left counts per chunk), k_scatter
//      (stable two-ended partition into the other primitive buffer plus the children's
//      centroid boxes), k_children (child tasks).  The host reads two counters per level.
//   3. k_small: one wave per remaining subtree, depth first, the same steps on 64 lanes.
//   4. Threading: the position of node k in table t is the sum, along its root path, of
//      1 (first child emitted) or 1 + size(first child) (second child); k_rel computes
//      the per-edge terms and k_jump sums them by pointer doubling in ceil(log2(depth+1))
//      rounds; k_table scatters the rows {node, prim, miss} and, optionally, the 32-byte
//      device rows the traversal kernels read.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <string>
#include <vector>

#include "zgl_bvh.h"

namespace zgl {

namespace {

constexpr int kBuckets = 16;
constexpr int kAccInts = 13;  // count, key min xyz, key max xyz, zero lo xyz, zero hi xyz
constexpr int kNodeAcc = kBuckets * kAccInts;
constexpr int kExtInts = 9;  // key min xyz, key max xyz, zero xyz (a point set)
constexpr int kChunk = 2048;
constexpr int kLeaf = (int)0x80000000;

// Error bits
constexpr int kErrIndex = 1;
constexpr int kErrNonFinite = 2;
constexpr int kErrSplit = 4;

// Counter slots
constexpr int cNextBig = 0, cNextChunks = 1, cSmall = 2, cError = 3, cMaxDepth = 4, cFetch = 5;
constexpr int kNumCounters = 8;

struct __align__(16) PrimRec {
    float4 a;  // lo.xyz, primitive index (int bits)
    float4 b;  // hi.xyz, unused
};

struct __align__(16) Task {
    int offset, l, r, dim;
    float axisMin, axisMax;
    int depth, buf;
    int chunk0, pad0, pad1, pad2;
};

struct __align__(16) SplitInfo {
    int bucket;    // split bucket, or -1 when the node is not split here
    int splitPos;  // last position of the left child
    int pad0, pad1;
};

__device__ inline float minG(float x, float y) { return (y < x) ? y : x; }
__device__ inline float maxG(float x, float y) { return (x < y) ? y : x; }

// Integer key ordered like the float, -0 folded into +0 (NaN excluded by validation).
__device__ inline int fkey(float v)
{
    const int i = __float_as_int(v == 0.0f ? 0.0f : v);
    return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ inline float unkey(int k) { return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff)); }
__device__ inline int zeroTag(int pos, float v) { return (pos << 1) | (int)((unsigned)__float_as_int(v) >> 31); }
__device__ inline float decode(int key, int ztag) { return key == 0 ? ((ztag & 1) ? -0.0f : 0.0f) : unkey(key); }

__device__ inline int accInit(int slot)
{
    if (slot == 0) return 0;
    if (slot < 4) return fkey(1e8f);
    if (slot < 7) return fkey(-1e8f);
    return INT_MAX;
}
__device__ inline int extInit(int slot)
{
    if (slot < 3) return fkey(1e8f);
    if (slot < 6) return fkey(-1e8f);
    return INT_MAX;
}

// Merge an LDS partial into a global accumulator (slot kinds as accInit / extInit).
__device__ inline void mergeAcc(int* g, int slot, int v)
{
    if (slot == 0) { if (v) atomicAdd(g, v); }
    else if (slot < 4) { if (v != fkey(1e8f)) atomicMin(g, v); }
    else if (slot < 7) { if (v != fkey(-1e8f)) atomicMax(g, v); }
    else if (v != INT_MAX) atomicMin(g, v);
}
__device__ inline void mergeExt(int* g, int slot, int v)
{
    if (slot < 3) { if (v != fkey(1e8f)) atomicMin(g, v); }
    else if (slot < 6) { if (v != fkey(-1e8f)) atomicMax(g, v); }
    else if (v != INT_MAX) atomicMin(g, v);
}

__device__ inline float comp(const float4& v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : v.z); }

__device__ inline float centroidOf(const PrimRec& p, int c) { return (comp(p.a, c) + comp(p.b, c)) * 0.5f; }

// BVH.cpp:256-257: int b = 16 * (c - axisMin) / (axisMax - axisMin), clamped; the
// float -> int conversion gives INT_MIN for NaN and out-of-range values (x86-64), which
// the clamp turns into bucket 0.
__device__ inline int bucketOf(float c, float mn, float mx)
{
    const float f = __fdiv_rn(16.0f * (c - mn), mx - mn);
    const int b = (f > -2147483904.0f && f < 2147483648.0f) ? (int)f : INT_MIN;
    return max(min(b, kBuckets - 1), 0);
}

// Accumulate a primitive's bound into a bucket accumulator (LDS).
__device__ inline void accAdd(int* a, const PrimRec& p, int pos)
{
    atomicAdd(a, 1);
    const float lo[3] = {p.a.x, p.a.y, p.a.z}, hi[3] = {p.b.x, p.b.y, p.b.z};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        atomicMin(a + 1 + c, fkey(lo[c]));
        atomicMax(a + 4 + c, fkey(hi[c]));
        if (lo[c] == 0.0f) atomicMin(a + 7 + c, zeroTag(pos, lo[c]));
        if (hi[c] == 0.0f) atomicMin(a + 10 + c, zeroTag(pos, hi[c]));
    }
}
// Accumulate a point (centroid) into an extent accumulator (LDS).
__device__ inline void extAdd(int* e, const float* c3, int pos)
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int k = fkey(c3[c]);
        atomicMin(e + c, k);
        atomicMax(e + 3 + c, k);
        if (c3[c] == 0.0f) atomicMin(e + 6 + c, zeroTag(pos, c3[c]));
    }
}

struct Box {
    float lo[3], hi[3];
};
__device__ inline Box emptyBox()
{
    Box b;
    for (int c = 0; c < 3; c++) { b.lo[c] = 1e8f; b.hi[c] = -1e8f; }
    return b;
}
__device__ inline void expand(Box& a, const Box& b)
{
    for (int c = 0; c < 3; c++) { a.lo[c] = minG(a.lo[c], b.lo[c]); a.hi[c] = maxG(a.hi[c], b.hi[c]); }
}
__device__ inline float surfaceArea(const Box& b)
{
    const float x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
    return 2.0f * (x * y + y * z + z * x);
}
__device__ inline int maxExtent(const Box& b)
{
    const float x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
    if (x > y) return x > z ? 0 : 2;
    return y > z ? 1 : 2;
}
__device__ inline Box decodeExt(const int* e)
{
    Box b;
    for (int c = 0; c < 3; c++) { b.lo[c] = decode(e[c], e[6 + c]); b.hi[c] = decode(e[3 + c], e[6 + c]); }
    return b;
}

// SAH choice of a node from its 16 bucket accumulators (BVH.cpp:259-282, host build
// zgl_build.cpp): the node bound, the split bucket, the number of primitives in buckets
// <= split, and the split position (or -1 when the left child would be empty, which
// valid scenes never produce).
struct SplitChoice {
    Box bound;
    int bucket;
    int leftCount;
    int splitPos;
};
__device__ inline Box bucketBox(const int* a)
{
    Box b;
    for (int c = 0; c < 3; c++) {
        b.lo[c] = decode(a[1 + c], a[7 + c]);
        b.hi[c] = decode(a[4 + c], a[10 + c]);
    }
    return b;
}
// One wave computes the choice: lane b (b < 16) decodes bucket b; the prefix boxes and
// counts are inclusive scans over lanes 0..b and the suffix ones over lanes 15..b.
// minG / maxG keep the leftmost extreme, which makes the fold associative, so a scan
// that always combines (earlier part, later part) in the host's order gives its bits:
// the prefix folds buckets in increasing order, the suffix in decreasing order
// (sbx[j] = sbx[j+1] expanded by bucket j).  The argmin then runs serially over the 15
// costs with the host's strict '<'.  All lanes return the same values.
__device__ inline Box shflBox(const Box& b, int src)
{
    Box o;
    for (int c = 0; c < 3; c++) { o.lo[c] = __shfl(b.lo[c], src, 64); o.hi[c] = __shfl(b.hi[c], src, 64); }
    return o;
}
__device__ inline void foldInto(Box& acc, const Box& left)
{
    // acc = left expanded by acc (left is the earlier part of the sequence)
    for (int c = 0; c < 3; c++) { acc.lo[c] = minG(left.lo[c], acc.lo[c]); acc.hi[c] = maxG(left.hi[c], acc.hi[c]); }
}
__device__ inline SplitChoice chooseSplitWave(const int* acc, int l, int r)
{
    const int lane = threadIdx.x & 63;
    const int b = lane & (kBuckets - 1);
    const Box own = bucketBox(acc + b * kAccInts);
    const int cnt = acc[b * kAccInts];
    Box pb = own, sb = own;
    int pc = cnt, sc = cnt;
#pragma unroll
    for (int off = 1; off < kBuckets; off <<= 1) {
        const Box lp = shflBox(pb, lane - off);
        const int lpc = __shfl(pc, lane - off, 64);
        const Box ls = shflBox(sb, lane + off);
        const int lsc = __shfl(sc, lane + off, 64);
        if (b >= off) { foldInto(pb, lp); pc += lpc; }
        if (b + off < kBuckets) { foldInto(sb, ls); sc += lsc; }
    }
    // cost of splitting after bucket b (b < 15): prefix b and suffix b + 1
    const Box sNext = shflBox(sb, lane + 1);
    const int scNext = __shfl(sc, lane + 1, 64);
    const float cost = (float)pc * surfaceArea(pb) + (float)scNext * surfaceArea(sNext);
    int split = 0;
    float minCost = __shfl(cost, 0, 64);
    for (int i = 1; i < kBuckets - 1; i++) {
        const float c = __shfl(cost, i, 64);
        if (c < minCost) { minCost = c; split = i; }
    }
    SplitChoice s;
    s.bound = shflBox(pb, kBuckets - 1);
    const int leftAtSplit = __shfl(pc, split, 64);
    const int n = r - l + 1;
    int hi = leftAtSplit;
    if (hi == n) hi--;
    s.bucket = split;
    s.leftCount = leftAtSplit;
    s.splitPos = hi == 0 ? -1 : l + hi - 1;
    return s;
}

__device__ inline void writeBound(float* bounds, int offset, const Box& b)
{
    float* o = bounds + (size_t)offset * 6;
    o[0] = b.lo[0]; o[1] = b.lo[1]; o[2] = b.lo[2];
    o[3] = b.hi[0]; o[4] = b.hi[1]; o[5] = b.hi[2];
}
__device__ inline Box primBox(const PrimRec& p)
{
    Box b;
    b.lo[0] = p.a.x; b.lo[1] = p.a.y; b.lo[2] = p.a.z;
    b.hi[0] = p.b.x; b.hi[1] = p.b.y; b.hi[2] = p.b.z;
    return b;
}

struct Bufs {
    PrimRec* prim[2];
    float* bounds;
    int* sizeIdx;
    int* parent;
    int* counters;
};

__device__ inline Task childTask(int offset, int l, int r, const Box& ext, int depth, int buf)
{
    Task t;
    t.offset = offset;
    t.l = l;
    t.r = r;
    t.dim = maxExtent(ext);
    t.axisMin = ext.lo[t.dim];
    t.axisMax = ext.hi[t.dim];
    t.depth = depth;
    t.buf = buf;
    t.chunk0 = 0;
    t.pad0 = t.pad1 = t.pad2 = 0;
    return t;
}

// Queue a task: big (>= subtree primitives) into the next level, else into the small list.
__device__ inline void pushTask(Task t, int subtree, Task* nextBig, Task* small, int* counters)
{
    const int n = t.r - t.l + 1;
    if (n >= subtree) {
        const int j = atomicAdd(counters + cNextBig, 1);
        t.chunk0 = atomicAdd(counters + cNextChunks, (n + kChunk - 1) / kChunk);
        nextBig[j] = t;
    } else {
        small[atomicAdd(counters + cSmall, 1)] = t;
    }
}

// ---------------------------------------------------------------- phase 1

__global__ __launch_bounds__(256) void k_prims(const float* __restrict__ verts, int64_t nV,
                                               const uint32_t* __restrict__ idx, int nT, PrimRec* __restrict__ prim,
                                               int* __restrict__ rootExt, int* __restrict__ counters)
{
    __shared__ int sExt[kExtInts];
    if (threadIdx.x < kExtInts) sExt[threadIdx.x] = extInit(threadIdx.x);
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nT) {
        uint32_t v[3];
        int err = 0;
        for (int k = 0; k < 3; k++) {
            v[k] = idx[(size_t)i * 3 + k];
            if ((int64_t)v[k] >= nV) { err |= kErrIndex; v[k] = 0; }
        }
        const float* a = verts + (size_t)v[0] * 3;
        const float* b = verts + (size_t)v[1] * 3;
        const float* c = verts + (size_t)v[2] * 3;
        float lo[3], hi[3], cen[3];
        for (int k = 0; k < 3; k++) {
            if (!isfinite(a[k]) || !isfinite(b[k]) || !isfinite(c[k])) err |= kErrNonFinite;
            lo[k] = minG(minG(a[k], b[k]), c[k]);
            hi[k] = maxG(maxG(a[k], b[k]), c[k]);
            cen[k] = (lo[k] + hi[k]) * 0.5f;
        }
        if (err) {
            atomicOr(counters + cError, err);
            for (int k = 0; k < 3; k++) lo[k] = hi[k] = cen[k] = 0.0f;
        }
        PrimRec p;
        p.a = make_float4(lo[0], lo[1], lo[2], __int_as_float(i));
        p.b = make_float4(hi[0], hi[1], hi[2], 0.0f);
        prim[i] = p;
        extAdd(sExt, cen, i);
    }
    __syncthreads();
    if (threadIdx.x < kExtInts) mergeExt(rootExt + threadIdx.x, threadIdx.x, sExt[threadIdx.x]);
}

__global__ void k_root(const int* __restrict__ rootExt, int nT, int subtree, Task* nextBig, Task* small,
                       int* parent, int* counters)
{
    Box ext;
    int e[kExtInts];
    for (int k = 0; k < kExtInts; k++) e[k] = rootExt[k];
    ext = decodeExt(e);
    parent[0] = -1;
    pushTask(childTask(0, 0, nT - 1, ext, 0, 0), subtree, nextBig, small, counters);
}

__global__ void k_reset(int* counters, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) counters[i] = 0;
}

__global__ void k_init_ext(int* rootExt)
{
    if (threadIdx.x < kExtInts) rootExt[threadIdx.x] = extInit(threadIdx.x);
}

// ---------------------------------------------------------------- phase 2 (big nodes)

// Per node: clear its bucket accumulators and map its chunks to it.
__global__ __launch_bounds__(256) void k_map(const Task* __restrict__ tasks, int* __restrict__ acc,
                                             int* __restrict__ chunkNode)
{
    const int j = blockIdx.x;
    const Task t = tasks[j];
    for (int i = threadIdx.x; i < kNodeAcc; i += blockDim.x) acc[(size_t)j * kNodeAcc + i] = accInit(i % kAccInts);
    const int nch = (t.r - t.l + 1 + kChunk - 1) / kChunk;
    for (int c = threadIdx.x; c < nch; c += blockDim.x) chunkNode[t.chunk0 + c] = j;
}

__global__ __launch_bounds__(256) void k_bin(const Task* __restrict__ tasks, const int* __restrict__ chunkNode,
                                             const PrimRec* __restrict__ p0, const PrimRec* __restrict__ p1,
                                             int* __restrict__ acc, int* __restrict__ chunkCount)
{
    __shared__ int sAcc[kNodeAcc];
    const int c = blockIdx.x;
    const int j = chunkNode[c];
    const Task t = tasks[j];
    for (int i = threadIdx.x; i < kNodeAcc; i += blockDim.x) sAcc[i] = accInit(i % kAccInts);
    __syncthreads();
    const PrimRec* P = t.buf ? p1 : p0;
    const int start = t.l + (c - t.chunk0) * kChunk;
    const int end = min(start + kChunk, t.r + 1);
    for (int pos = start + threadIdx.x; pos < end; pos += blockDim.x) {
        const PrimRec p = P[pos];
        const int b = bucketOf(centroidOf(p, t.dim), t.axisMin, t.axisMax);
        accAdd(sAcc + b * kAccInts, p, pos);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kNodeAcc; i += blockDim.x) {
        const int slot = i % kAccInts;
        mergeAcc(acc + (size_t)j * kNodeAcc + i, slot, sAcc[i]);
        if (slot == 0) chunkCount[(size_t)c * kBuckets + i / kAccInts] = sAcc[i];
    }
}

__global__ __launch_bounds__(64) void k_split(const Task* __restrict__ tasks, const int* __restrict__ acc,
                                              SplitInfo* __restrict__ split, int* __restrict__ ext, Bufs B)
{
    const int j = blockIdx.x;
    const int lane = threadIdx.x;
    const Task t = tasks[j];
    const SplitChoice s = chooseSplitWave(acc + (size_t)j * kNodeAcc, t.l, t.r);
    if (lane == 0) {
        writeBound(B.bounds, t.offset, s.bound);
        B.sizeIdx[t.offset] = (t.r - t.l) * 2 + 1;
        SplitInfo si;
        si.bucket = s.splitPos < 0 ? -1 : s.bucket;
        si.splitPos = s.splitPos;
        si.pad0 = s.leftCount;
        si.pad1 = 0;
        split[j] = si;
        if (s.splitPos < 0) atomicOr(B.counters + cError, kErrSplit);
    }
    if (lane < 2 * kExtInts) ext[(size_t)j * 2 * kExtInts + lane] = extInit(lane % kExtInts);
}

// One wave per node: per-chunk left counts -> exclusive prefix (chunk order).
__global__ __launch_bounds__(64) void k_scan(const Task* __restrict__ tasks, const SplitInfo* __restrict__ split,
                                             const int* __restrict__ chunkCount, int* __restrict__ chunkLeft)
{
    const int j = blockIdx.x;
    const int lane = threadIdx.x;
    const SplitInfo si = split[j];
    if (si.bucket < 0) return;
    const Task t = tasks[j];
    const int nch = (t.r - t.l + 1 + kChunk - 1) / kChunk;
    int running = 0;
    for (int base = 0; base < nch; base += 64) {
        const int c = base + lane;
        int cnt = 0;
        if (c < nch)
            for (int b = 0; b <= si.bucket; b++) cnt += chunkCount[(size_t)(t.chunk0 + c) * kBuckets + b];
        int incl = cnt;
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        if (c < nch) chunkLeft[t.chunk0 + c] = running + incl - cnt;
        running += __shfl(incl, 63, 64);
    }
}

__global__ __launch_bounds__(256) void k_scatter(const Task* __restrict__ tasks, const int* __restrict__ chunkNode,
                                                 const SplitInfo* __restrict__ split,
                                                 const int* __restrict__ chunkLeft, Bufs B, int* __restrict__ ext)
{
    __shared__ int sExt[2 * kExtInts];
    __shared__ int sL[4], sR[4];
    const int c = blockIdx.x;
    const int j = chunkNode[c];
    const SplitInfo si = split[j];
    if (si.bucket < 0) return;
    const Task t = tasks[j];
    if (threadIdx.x < 2 * kExtInts) sExt[threadIdx.x] = extInit(threadIdx.x % kExtInts);
    const PrimRec* P = B.prim[t.buf];
    PrimRec* Q = B.prim[t.buf ^ 1];
    const int start = t.l + (c - t.chunk0) * kChunk;
    const int end = min(start + kChunk, t.r + 1);
    const int leftBase = chunkLeft[c];
    const int rightBase = (start - t.l) - leftBase;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int leftRun = 0, rightRun = 0;
    __syncthreads();
    for (int r0 = start; r0 < end; r0 += 256) {
        const int pos = r0 + threadIdx.x;
        const bool valid = pos < end;
        PrimRec p;
        bool isLeft = false;
        if (valid) {
            p = P[pos];
            isLeft = bucketOf(centroidOf(p, t.dim), t.axisMin, t.axisMax) <= si.bucket;
        }
        const unsigned long long mL = __ballot(isLeft), mR = __ballot(valid && !isLeft);
        if (lane == 0) { sL[w] = __popcll(mL); sR[w] = __popcll(mR); }
        __syncthreads();
        int wl = 0, wr = 0, tl = 0, tr = 0;
        for (int k = 0; k < 4; k++) {
            if (k < w) { wl += sL[k]; wr += sR[k]; }
            tl += sL[k];
            tr += sR[k];
        }
        if (valid) {
            const int dest = isLeft ? t.l + leftBase + leftRun + wl + __popcll(mL & below)
                                    : t.r - (rightBase + rightRun + wr + __popcll(mR & below));
            Q[dest] = p;
            const float cen[3] = {centroidOf(p, 0), centroidOf(p, 1), centroidOf(p, 2)};
            extAdd(sExt + (dest <= si.splitPos ? 0 : kExtInts), cen, dest);
        }
        leftRun += tl;
        rightRun += tr;
        __syncthreads();
    }
    if (threadIdx.x < 2 * kExtInts)
        mergeExt(ext + (size_t)j * 2 * kExtInts + threadIdx.x, threadIdx.x % kExtInts, sExt[threadIdx.x]);
}

__global__ __launch_bounds__(64) void k_children(const Task* __restrict__ tasks, int nBig,
                                                 const SplitInfo* __restrict__ split, const int* __restrict__ ext,
                                                 int subtree, Task* nextBig, Task* small, Bufs B)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nBig) return;
    const SplitInfo si = split[j];
    if (si.bucket < 0) return;
    const Task t = tasks[j];
    const Box lc = decodeExt(ext + (size_t)j * 2 * kExtInts);
    const Box rc = decodeExt(ext + (size_t)j * 2 * kExtInts + kExtInts);
    const int sp = si.splitPos;
    const Task lt = childTask(t.offset + 1, t.l, sp, lc, t.depth + 1, t.buf ^ 1);
    const Task rt = childTask(t.offset + 2 * (sp - t.l) + 2, sp + 1, t.r, rc, t.depth + 1, t.buf ^ 1);
    B.parent[lt.offset] = t.offset;
    B.parent[rt.offset] = t.offset;
    atomicMax(B.counters + cMaxDepth, t.depth + 1);
    pushTask(lt, subtree, nextBig, small, B.counters);
    pushTask(rt, subtree, nextBig, small, B.counters);
}

// ---------------------------------------------------------------- phase 3 (small subtrees)

// Orders this wave's LDS accesses (a wave's LDS operations execute in order; this keeps
// the compiler from moving them).
__device__ inline void waveLdsSync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}
__device__ inline void wgFence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ inline void storeTask(Task* dst, const Task& t)
{
    const int* s = reinterpret_cast<const int*>(&t);
    int* d = reinterpret_cast<int*>(dst);
    for (int k = 0; k < 12; k++) __hip_atomic_store(d + k, s[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline Task loadTask(const Task* src)
{
    Task t;
    int* d = reinterpret_cast<int*>(&t);
    const int* s = reinterpret_cast<const int*>(src);
    for (int k = 0; k < 12; k++)
        d[k] = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(const_cast<int*>(s + k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
    return t;
}
__device__ inline PrimRec shflPrim(const PrimRec& p, int src)
{
    PrimRec q;
    q.a = make_float4(__shfl(p.a.x, src, 64), __shfl(p.a.y, src, 64), __shfl(p.a.z, src, 64),
                      __shfl(p.a.w, src, 64));
    q.b = make_float4(__shfl(p.b.x, src, 64), __shfl(p.b.y, src, 64), __shfl(p.b.z, src, 64),
                      __shfl(p.b.w, src, 64));
    return q;
}
__device__ inline void writeLeaf(Bufs& B, int offset, const PrimRec& p)
{
    writeBound(B.bounds, offset, primBox(p));
    B.sizeIdx[offset] = __float_as_int(p.a.w) | kLeaf;
}

// Processes one node with the whole wave.  Returns true with (left, right) children to
// continue with, false for a leaf, a two-primitive node (whose leaves it writes) or an
// invalid split.
__device__ bool smallNode(const Task& t, Bufs& B, int* sAcc, int* sExt, int& maxDepth, Task& left, Task& right)
{
    const int lane = threadIdx.x & 63;
    const int m = t.r - t.l + 1;
    const PrimRec* P = B.prim[t.buf];
    if (m < 1) {  // not produced by valid splits; kept out of the writes below
        if (lane == 0) atomicOr(B.counters + cError, kErrSplit);
        return false;
    }
    if (m <= 2) {
        PrimRec mine;
        if (lane < m) mine = P[t.l + lane];
        const PrimRec pl = shflPrim(mine, 0);
        if (m == 1) {
            if (lane == 0) writeLeaf(B, t.offset, pl);
            return false;
        }
        const PrimRec pr = shflPrim(mine, 1);
        if (lane == 0) {
            Box b = primBox(pl);
            expand(b, primBox(pr));
            writeBound(B.bounds, t.offset, b);
            B.sizeIdx[t.offset] = 3;
            const bool swap = centroidOf(pl, t.dim) > centroidOf(pr, t.dim);
            writeLeaf(B, t.offset + 1, swap ? pr : pl);
            writeLeaf(B, t.offset + 2, swap ? pl : pr);
            B.parent[t.offset + 1] = t.offset;
            B.parent[t.offset + 2] = t.offset;
        }
        maxDepth = max(maxDepth, t.depth + 1);
        return false;
    }
    for (int i = lane; i < kNodeAcc; i += 64) sAcc[i] = accInit(i % kAccInts);
    if (lane < 2 * kExtInts) sExt[lane] = extInit(lane % kExtInts);
    waveLdsSync();
    for (int base = t.l; base <= t.r; base += 64) {
        const int pos = base + lane;
        if (pos <= t.r) {
            const PrimRec p = P[pos];
            accAdd(sAcc + bucketOf(centroidOf(p, t.dim), t.axisMin, t.axisMax) * kAccInts, p, pos);
        }
    }
    waveLdsSync();
    const SplitChoice s = chooseSplitWave(sAcc, t.l, t.r);
    if (lane == 0) {
        writeBound(B.bounds, t.offset, s.bound);
        B.sizeIdx[t.offset] = (t.r - t.l) * 2 + 1;
        if (s.splitPos < 0) atomicOr(B.counters + cError, kErrSplit);
    }
    if (s.splitPos < 0) return false;
    PrimRec* Q = B.prim[t.buf ^ 1];
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int leftRun = 0, rightRun = 0;
    for (int base = t.l; base <= t.r; base += 64) {
        const int pos = base + lane;
        const bool valid = pos <= t.r;
        PrimRec p;
        bool isLeft = false;
        if (valid) {
            p = P[pos];
            isLeft = bucketOf(centroidOf(p, t.dim), t.axisMin, t.axisMax) <= s.bucket;
        }
        const unsigned long long mL = __ballot(isLeft), mR = __ballot(valid && !isLeft);
        if (valid) {
            const int dest = isLeft ? t.l + leftRun + __popcll(mL & below) : t.r - (rightRun + __popcll(mR & below));
            Q[dest] = p;
            const float cen[3] = {centroidOf(p, 0), centroidOf(p, 1), centroidOf(p, 2)};
            extAdd(sExt + (dest <= s.splitPos ? 0 : kExtInts), cen, dest);
        }
        leftRun += __popcll(mL);
        rightRun += __popcll(mR);
    }
    wgFence();
    int e[2 * kExtInts];
    for (int k = 0; k < 2 * kExtInts; k++) e[k] = sExt[k];
    const Box lc = decodeExt(e), rc = decodeExt(e + kExtInts);
    const int sp = s.splitPos;
    left = childTask(t.offset + 1, t.l, sp, lc, t.depth + 1, t.buf ^ 1);
    right = childTask(t.offset + 2 * (sp - t.l) + 2, sp + 1, t.r, rc, t.depth + 1, t.buf ^ 1);
    if (lane == 0) {
        B.parent[left.offset] = t.offset;
        B.parent[right.offset] = t.offset;
    }
    maxDepth = max(maxDepth, t.depth + 1);
    return true;
}

// One wave per small subtree, tasks fetched from a counter; depth first with the left
// child continued in registers and the right one on a per-subtree stack in global memory
// (capacity = primitives of the subtree, which bounds its depth).
__global__ __launch_bounds__(256) void k_small(const Task* __restrict__ small, int nSmall, Task* stack, Bufs B)
{
    __shared__ int sAccAll[4][kNodeAcc];
    __shared__ int sExtAll[4][2 * kExtInts];
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    int* sAcc = sAccAll[w];
    int* sExt = sExtAll[w];
    int maxDepth = 0;
    for (;;) {
        int ti = 0;
        if (lane == 0) ti = atomicAdd(B.counters + cFetch, 1);
        ti = __shfl(ti, 0, 64);
        if (ti >= nSmall) break;
        Task t = small[ti];
        Task* stk = stack + t.l;
        int top = 0;
        for (;;) {
            Task lt, rt;
            if (smallNode(t, B, sAcc, sExt, maxDepth, lt, rt)) {
                if (lane == 0) storeTask(stk + top, rt);
                top++;
                t = lt;
            } else {
                if (top == 0) break;
                top--;
                wgFence();
                t = loadTask(stk + top);
            }
        }
    }
    if (lane == 0) atomicMax(B.counters + cMaxDepth, maxDepth);
}

// ---------------------------------------------------------------- phase 4 (threading)

__device__ inline int nodeSize(const int* sizeIdx, int k)
{
    const int s = sizeIdx[k];
    return (s & kLeaf) ? 1 : s;
}
__device__ inline float centroidBound(const float* bounds, int k, int axis)
{
    return (bounds[(size_t)k * 6 + axis] + bounds[(size_t)k * 6 + 3 + axis]) * 0.5f;
}

// Per node and table: 1 if the node is the first child its parent emits, else
// 1 + size(first child) (buildHitTable, BVH.cpp:317-343 / zgl_build.cpp).
__global__ __launch_bounds__(256) void k_rel(const float* __restrict__ bounds, const int* __restrict__ sizeIdx,
                                             const int* __restrict__ parent, int S, int* __restrict__ val,
                                             int* __restrict__ anc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S) return;
    if (k == 0) {
        for (int o = 0; o < 6; o++) val[(size_t)o * S] = 0;
        anc[0] = -1;
        return;
    }
    const int p = parent[k];
    const int lch = p + 1;
    const int rch = lch + nodeSize(sizeIdx, lch);
    for (int o = 0; o < 6; o++) {
        const int axis = o >> 1;
        const float cl = centroidBound(bounds, lch, axis), cr = centroidBound(bounds, rch, axis);
        const bool first = (o & 1) == 0 ? (cl > cr) : (cl < cr);
        const int f = first ? lch : rch;
        val[(size_t)o * S + k] = (k == f) ? 1 : 1 + nodeSize(sizeIdx, f);
    }
    anc[k] = p;
}

__global__ __launch_bounds__(256) void k_jump(const int* __restrict__ valIn, const int* __restrict__ ancIn,
                                              int* __restrict__ valOut, int* __restrict__ ancOut, int S)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S) return;
    const int a = ancIn[k];
    for (int o = 0; o < 6; o++) {
        const size_t i = (size_t)o * S;
        valOut[i + k] = a >= 0 ? valIn[i + k] + valIn[i + a] : valIn[i + k];
    }
    ancOut[k] = a >= 0 ? ancIn[a] : -1;
}

__global__ __launch_bounds__(256) void k_table(const float* __restrict__ bounds, const int* __restrict__ sizeIdx,
                                               const int* __restrict__ val, int S, int* __restrict__ hit,
                                               float4* __restrict__ rows)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S) return;
    const int s = sizeIdx[k];
    const bool leaf = (s & kLeaf) != 0;
    const int ns = leaf ? 1 : s;
    const int prim = leaf ? (s ^ kLeaf) : -1;
    const float* b = bounds + (size_t)k * 6;
    for (int o = 0; o < 6; o++) {
        const int i = val[(size_t)o * S + k];
        if ((unsigned)i >= (unsigned)S) continue;  // positions are < S by construction
        if (hit) {
            int* h = hit + ((size_t)o * S + i) * 3;
            h[0] = k;
            h[1] = prim;
            h[2] = i + ns;
        }
        if (rows) {
            float4* r = rows + ((size_t)o * S + i) * 2;
            r[0] = make_float4(b[0], b[1], b[2], b[3]);
            r[1] = make_float4(b[4], b[5], __int_as_float(prim), __int_as_float(o * S + i + ns));
        }
    }
}

template <class T> T* carve(char*& p, size_t n)
{
    T* r = reinterpret_cast<T*>(p);
    p += (n * sizeof(T) + 255) & ~(size_t)255;
    return r;
}

}  // namespace

DeviceBvh::~DeviceBvh()
{
    if (work_) (void)hipFree(work_);
}

void DeviceBvh::release()
{
    if (work_) (void)hipFree(work_);
    work_ = nullptr;
    workBytes_ = 0;
}

int DeviceBvh::build(const float* dVerts, int64_t nV, const uint32_t* dIdx, int64_t nT, int subtree, float* dBounds,
                     int32_t* dHit, void* dRows, hipStream_t st, std::string& err)
{
    stats = DeviceBvhStats{};
    if (nT < 1 || nT > (int64_t)1 << 27 || nV < 1 || !dVerts || !dIdx || !dBounds) {
        err = "device MTBVH build: need 1 <= triangles <= 2^27, vertices and an output buffer";
        return 1;
    }
    if (subtree < 3) subtree = 3;
    const int n = (int)nT;
    const int S = 2 * n - 1;
    const size_t bigCap = (size_t)n / subtree + 2;
    const size_t chunkCap = (size_t)n / kChunk + bigCap + 2;
    auto need = [&](size_t elems, size_t sz) { return (elems * sz + 255) & ~(size_t)255; };
    const size_t bytes = 2 * need(n, sizeof(PrimRec)) + 2 * need(bigCap, sizeof(Task)) +
                         need(bigCap * kNodeAcc, 4) + need(bigCap, sizeof(SplitInfo)) +
                         need(bigCap * 2 * kExtInts, 4) + need(chunkCap, 4) * 2 + need(chunkCap * kBuckets, 4) +
                         need(n, sizeof(Task)) * 2 + need(S, 4) * 2 + need((size_t)S * 6, 4) * 2 + need(S, 4) * 2 +
                         need(kNumCounters, 4) + need(kExtInts, 4);
    if (workBytes_ < bytes) {
        release();
        const hipError_t e = hipMalloc(&work_, bytes);
        if (e != hipSuccess) {
            work_ = nullptr;
            err = std::string("device MTBVH build: hipMalloc: ") + hipGetErrorString(e);
            return e == hipErrorOutOfMemory ? 3 : 2;
        }
        workBytes_ = bytes;
    }
    char* p = static_cast<char*>(work_);
    Bufs B;
    B.prim[0] = carve<PrimRec>(p, n);
    B.prim[1] = carve<PrimRec>(p, n);
    Task* big[2] = {carve<Task>(p, bigCap), carve<Task>(p, bigCap)};
    int* acc = carve<int>(p, bigCap * kNodeAcc);
    SplitInfo* split = carve<SplitInfo>(p, bigCap);
    int* ext = carve<int>(p, bigCap * 2 * kExtInts);
    int* chunkNode = carve<int>(p, chunkCap);
    int* chunkLeft = carve<int>(p, chunkCap);
    int* chunkCount = carve<int>(p, chunkCap * kBuckets);
    Task* small = carve<Task>(p, n);
    Task* stack = carve<Task>(p, n);
    B.sizeIdx = carve<int>(p, S);
    B.parent = carve<int>(p, S);
    int* val[2] = {carve<int>(p, (size_t)S * 6), carve<int>(p, (size_t)S * 6)};
    int* anc[2] = {carve<int>(p, S), carve<int>(p, S)};
    B.counters = carve<int>(p, kNumCounters);
    int* rootExt = carve<int>(p, kExtInts);
    B.bounds = dBounds;

    auto check = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return true;
        err = std::string("device MTBVH build: ") + what + ": " + hipGetErrorString(e);
        return false;
    };
    int host[kNumCounters];
    auto readCounters = [&]() {
        return check(hipMemcpyAsync(host, B.counters, sizeof(host), hipMemcpyDeviceToHost, st), "read counters") &&
               check(hipStreamSynchronize(st), "synchronize");
    };

    k_reset<<<1, 64, 0, st>>>(B.counters, kNumCounters);
    k_init_ext<<<1, 64, 0, st>>>(rootExt);
    k_prims<<<(n + 255) / 256, 256, 0, st>>>(dVerts, nV, dIdx, n, B.prim[0], rootExt, B.counters);
    k_root<<<1, 1, 0, st>>>(rootExt, n, subtree, big[0], small, B.parent, B.counters);
    if (!check(hipGetLastError(), "launch") || !readCounters()) return 2;
    if (host[cError] & (kErrIndex | kErrNonFinite)) {
        err = (host[cError] & kErrIndex) ? "device MTBVH build: vertex index out of range"
                                         : "device MTBVH build: non-finite vertex coordinate";
        return 1;
    }
    int cur = 0;
    int nBig = host[cNextBig], nChunks = host[cNextChunks];
    while (nBig > 0) {
        stats.levels++;
        stats.bigNodes += nBig;
        stats.chunks += nChunks;
        const int next = cur ^ 1;
        k_map<<<nBig, 256, 0, st>>>(big[cur], acc, chunkNode);
        k_bin<<<nChunks, 256, 0, st>>>(big[cur], chunkNode, B.prim[0], B.prim[1], acc, chunkCount);
        k_split<<<nBig, 64, 0, st>>>(big[cur], acc, split, ext, B);
        k_scan<<<nBig, 64, 0, st>>>(big[cur], split, chunkCount, chunkLeft);
        k_scatter<<<nChunks, 256, 0, st>>>(big[cur], chunkNode, split, chunkLeft, B, ext);
        // the next level's counters start at zero; the small-list and error slots keep counting
        k_reset<<<1, 64, 0, st>>>(B.counters, 2);
        k_children<<<(nBig + 63) / 64, 64, 0, st>>>(big[cur], nBig, split, ext, subtree, big[next], small, B);
        if (!check(hipGetLastError(), "launch") || !readCounters()) return 2;
        if (host[cError] & kErrSplit) break;
        nBig = host[cNextBig];
        nChunks = host[cNextChunks];
        cur = next;
    }
    if (!(host[cError] & kErrSplit) && host[cSmall] > 0) {
        stats.smallTasks = host[cSmall];
        const int waves = host[cSmall];
        const int blocks = std::min((waves + 3) / 4, 256 * 8);
        k_small<<<blocks, 256, 0, st>>>(small, host[cSmall], stack, B);
        if (!check(hipGetLastError(), "launch") || !readCounters()) return 2;
    }
    if (host[cError] & kErrSplit) {
        err = "device MTBVH build: a split left a child empty (centroids at or beyond 1e8 in magnitude)";
        return 1;
    }
    stats.maxDepth = host[cMaxDepth];
    const int grid = (S + 255) / 256;
    k_rel<<<grid, 256, 0, st>>>(B.bounds, B.sizeIdx, B.parent, S, val[0], anc[0]);
    int src = 0;
    for (int reach = 1; reach < stats.maxDepth + 1; reach <<= 1) {
        k_jump<<<grid, 256, 0, st>>>(val[src], anc[src], val[src ^ 1], anc[src ^ 1], S);
        src ^= 1;
        stats.jumps++;
    }
    k_table<<<grid, 256, 0, st>>>(B.bounds, B.sizeIdx, val[src], S, dHit, static_cast<float4*>(dRows));
    if (!check(hipGetLastError(), "launch")) return 2;
    return 0;
}

}  // namespace zgl
