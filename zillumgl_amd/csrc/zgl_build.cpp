// zgl_build.cpp -- host producers: MTBVH, alias tables, env table, Sobol matrices.
#include "zgl_build.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <math.h>
#include <stdexcept>
#include <utility>

namespace zgl {

namespace {

inline float minG(float x, float y) { return (y < x) ? y : x; }
inline float maxG(float x, float y) { return (x < y) ? y : x; }

// float -> int conversion with the x86-64 cvttss2si result for NaN and out of
// range values (INT_MIN), which the bucket index of the reference relies on when a
// node's centroid extent is zero (src/accelerator/BVH.cpp:256-257).
inline int truncToIntX86(float f)
{
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return INT_MIN;
    return (int)f;
}

struct Box {
    float lo[3] = {1e8f, 1e8f, 1e8f};
    float hi[3] = {-1e8f, -1e8f, -1e8f};
    void expand(const Box& b)
    {
        for (int i = 0; i < 3; i++) { lo[i] = minG(lo[i], b.lo[i]); hi[i] = maxG(hi[i], b.hi[i]); }
    }
    void expandPoint(const float* p)
    {
        for (int i = 0; i < 3; i++) { lo[i] = minG(lo[i], p[i]); hi[i] = maxG(hi[i], p[i]); }
    }
    void centroid(float* c) const
    {
        for (int i = 0; i < 3; i++) c[i] = (lo[i] + hi[i]) * 0.5f;
    }
    float surfaceArea() const
    {
        float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
        return 2.0f * (x * y + y * z + z * x);
    }
    int maxExtent() const
    {
        float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
        if (x > y) return x > z ? 0 : 2;
        return y > z ? 1 : 2;
    }
    static Box point(const float* p)
    {
        Box b;
        for (int i = 0; i < 3; i++) { b.lo[i] = p[i]; b.hi[i] = p[i]; }
        return b;
    }
};

struct Prim {
    Box bound;
    float centroid[3];
    int index;
};

struct Bucket {
    int count = 0;
    Box box;
};

struct Task {
    int offset;
    Box extent;  // centroid bounds of [l, r]
    int splitDim;
    int l, r;
};

constexpr int NumBuckets = 16;

inline int bucketOf(float c, float axisMin, float axisMax)
{
    int b = truncToIntX86(16.0f * (c - axisMin) / (axisMax - axisMin));
    return std::max(std::min(b, NumBuckets - 1), 0);
}

}  // namespace

void buildMtbvh(const float* vertices, const uint32_t* indices, int64_t nTri,
                std::vector<float>& boundsOut, std::vector<int32_t>& hitTable)
{
    const int64_t treeSize = nTri * 2 - 1;
    std::vector<Prim> prims((size_t)nTri);
    std::vector<Box> bounds((size_t)treeSize);
    std::vector<int32_t> sizeIndices((size_t)treeSize);

    Box rootCent;
    for (int64_t i = 0; i < nTri; i++) {
        const float* a = vertices + 3 * (size_t)indices[3 * i + 0];
        const float* b = vertices + 3 * (size_t)indices[3 * i + 1];
        const float* c = vertices + 3 * (size_t)indices[3 * i + 2];
        Prim& p = prims[(size_t)i];
        for (int k = 0; k < 3; k++) {
            p.bound.lo[k] = minG(minG(a[k], b[k]), c[k]);
            p.bound.hi[k] = maxG(maxG(a[k], b[k]), c[k]);
        }
        p.bound.centroid(p.centroid);
        p.index = (int)i;
        rootCent.expandPoint(p.centroid);
    }

    std::vector<Task> stack;
    stack.push_back({0, rootCent, rootCent.maxExtent(), 0, (int)nTri - 1});
    std::vector<Prim> tmp;

    while (!stack.empty()) {
        Task t = stack.back();
        stack.pop_back();
        const int l = t.l, r = t.r;
        const int size = (r - l) * 2 + 1;
        sizeIndices[t.offset] = (size == 1) ? (prims[l].index | BvhLeafMask) : size;
        if (l == r) {
            bounds[t.offset] = prims[l].bound;
            continue;
        }
        const int n = r - l + 1;
        if (n == 2) {
            Box b = prims[l].bound;
            b.expand(prims[r].bound);
            bounds[t.offset] = b;
            if (prims[l].centroid[t.splitDim] > prims[r].centroid[t.splitDim]) std::swap(prims[l], prims[r]);
            stack.push_back({t.offset + 2, Box::point(prims[r].centroid), 0, r, r});
            stack.push_back({t.offset + 1, Box::point(prims[l].centroid), 0, l, l});
            continue;
        }
        const int dim = t.splitDim;
        const float axisMin = t.extent.lo[dim];
        const float axisMax = t.extent.hi[dim];

        Bucket buckets[NumBuckets], prefix[NumBuckets], suffix[NumBuckets];
        for (int i = l; i <= r; i++) {
            int b = bucketOf(prims[i].centroid[dim], axisMin, axisMax);
            buckets[b].count++;
            buckets[b].box.expand(prims[i].bound);
        }
        prefix[0] = buckets[0];
        suffix[NumBuckets - 1] = buckets[NumBuckets - 1];
        for (int i = 1; i < NumBuckets; i++) {
            prefix[i].count = prefix[i - 1].count + buckets[i].count;
            prefix[i].box = prefix[i - 1].box;
            prefix[i].box.expand(buckets[i].box);
            const int j = NumBuckets - 1 - i;
            suffix[j].count = suffix[j + 1].count + buckets[j].count;
            suffix[j].box = suffix[j + 1].box;
            suffix[j].box.expand(buckets[j].box);
        }
        bounds[t.offset] = prefix[NumBuckets - 1].box;

        int splitBucket = 0;
        float minCost = (float)prefix[0].count * prefix[0].box.surfaceArea() +
                        (float)suffix[1].count * suffix[1].box.surfaceArea();
        for (int i = 1; i < NumBuckets - 1; i++) {
            float cost = (float)prefix[i].count * prefix[i].box.surfaceArea() +
                         (float)suffix[i + 1].count * suffix[i + 1].box.surfaceArea();
            if (cost < minCost) { minCost = cost; splitBucket = i; }
        }

        // stable two-ended partition (src/accelerator/BVH.cpp:97-114)
        tmp.assign(prims.begin() + l, prims.begin() + r + 1);
        int lo = 0, hi = n;
        for (int i = 0; i < n; i++) {
            int b = bucketOf(tmp[i].centroid[dim], axisMin, axisMax);
            if (b <= splitBucket) prims[l + lo++] = tmp[i];
            else prims[l + --hi] = tmp[i];
        }
        if (hi == n) hi--;
        // An empty left child needs bucket 0 empty, i.e. no centroid at the node's
        // minimum: only centroids at or beyond the empty box's 1e8 do that.  The
        // reference would recurse on an empty range; the library rejects the scene.
        if (hi == 0) throw std::invalid_argument("MTBVH split with an empty child (centroids beyond 1e8)");
        const int split = l + hi - 1;

        Box lc, rc;
        for (int i = l; i <= split; i++) lc.expandPoint(prims[i].centroid);
        for (int i = split + 1; i <= r; i++) rc.expandPoint(prims[i].centroid);
        stack.push_back({t.offset + 2 * (split - l) + 2, rc, rc.maxExtent(), split + 1, r});
        stack.push_back({t.offset + 1, lc, lc.maxExtent(), l, split});
    }

    boundsOut.resize((size_t)treeSize * 6);
    for (int64_t i = 0; i < treeSize; i++)
        for (int k = 0; k < 3; k++) {
            boundsOut[(size_t)i * 6 + k] = bounds[(size_t)i].lo[k];
            boundsOut[(size_t)i * 6 + 3 + k] = bounds[(size_t)i].hi[k];
        }

    // Six threaded orders: X+, X-, Y+, Y-, Z+, Z-; the child whose centroid is larger
    // (+) or smaller (-) along the axis is emitted first.
    hitTable.assign((size_t)treeSize * 18, 0);
    std::vector<int64_t> dfs((size_t)treeSize);
    for (int order = 0; order < 6; order++) {
        const int axis = order / 2;
        const bool greater = (order % 2) == 0;
        int32_t* table = hitTable.data() + (size_t)treeSize * 3 * order;
        int64_t top = 0, index = 0;
        dfs[top++] = 0;
        while (top) {
            int64_t k = dfs[--top];
            const bool leaf = (sizeIndices[k] & BvhLeafMask) != 0;
            const int nodeSize = leaf ? 1 : sizeIndices[k];
            table[index * 3 + 0] = (int32_t)k;
            table[index * 3 + 1] = leaf ? (sizeIndices[k] ^ BvhLeafMask) : -1;
            table[index * 3 + 2] = (int32_t)(index + nodeSize);
            index++;
            if (leaf) continue;
            int lSize = sizeIndices[k + 1];
            if (lSize & BvhLeafMask) lSize = 1;
            int64_t lch = k + 1, rch = k + 1 + lSize;
            float cl[3], cr[3];
            bounds[(size_t)lch].centroid(cl);
            bounds[(size_t)rch].centroid(cr);
            const bool first = greater ? (cl[axis] > cr[axis]) : (cl[axis] < cr[axis]);
            if (!first) std::swap(lch, rch);
            dfs[top++] = rch;
            dfs[top++] = lch;
        }
    }
}

void aliasBuild(const float* pdf, int32_t n, int32_t* alias, float* prob)
{
    for (int i = 0; i < n; i++) prob[i] = pdf[i];
    float sum = 0.0f;
    for (int i = 0; i < n; i++) sum += prob[i];
    const float sumInv = (float)n / sum;
    std::vector<std::pair<int, float>> greater((size_t)n * 2 + 1), lesser((size_t)n * 2 + 1);
    int gTop = 0, lTop = 0;
    for (int i = 0; i < n; i++) {
        prob[i] *= sumInv;
        if (prob[i] < 1.0f) lesser[lTop++] = {i, prob[i]};
        else greater[gTop++] = {i, prob[i]};
    }
    while (gTop != 0 && lTop != 0) {
        auto [l, pl] = lesser[--lTop];
        auto [g, pg] = greater[--gTop];
        alias[l] = g;
        prob[l] = pl;
        pg += pl - 1.0f;
        if (pg < 1.0f) lesser[lTop++] = {g, pg};
        else greater[gTop++] = {g, pg};
    }
    while (gTop != 0) { auto [g, pg] = greater[--gTop]; alias[g] = g; prob[g] = pg; }
    while (lTop != 0) { auto [l, pl] = lesser[--lTop]; alias[l] = l; prob[l] = pl; }
}

float envAliasRow(int32_t* alias, float* pdf, int32_t n, int32_t stride)
{
    float sum = 0.0f;
    for (int i = 0, off = 0; i < n; i++, off += stride) sum += pdf[off];
    const float sumInv = (float)n / sum;
    std::vector<std::pair<int, float>> greater((size_t)n * 2 + 1), lesser((size_t)n * 2 + 1);
    int gTop = 0, lTop = 0;
    for (int i = 0, off = 0; i < n; i++, off += stride) {
        pdf[off] *= sumInv;
        if (pdf[off] < 1.0f) lesser[lTop++] = {i, pdf[off]};
        else greater[gTop++] = {i, pdf[off]};
    }
    while (gTop != 0 && lTop != 0) {
        auto [l, pl] = lesser[--lTop];
        auto [g, pg] = greater[--gTop];
        alias[(size_t)l * stride] = g;
        pdf[(size_t)l * stride] = pl;
        pg += pl - 1.0f;
        if (pg < 1.0f) lesser[lTop++] = {g, pg};
        else greater[gTop++] = {g, pg};
    }
    while (gTop != 0) { auto [g, pg] = greater[--gTop]; (void)pg; alias[(size_t)g * stride] = g; pdf[(size_t)g * stride] = 1.0f; }
    while (lTop != 0) { auto [l, pl] = lesser[--lTop]; (void)pl; alias[(size_t)l * stride] = l; pdf[(size_t)l * stride] = 1.0f; }
    return sum;
}

float envAliasBuild(const float* rgb, int32_t w, int32_t h, int32_t* alias, float* prob)
{
    const size_t stride = (size_t)w + 1;
    for (int i = 0; i < h; i++) {
        const float s = sinf((float)(i + 0.5f) / (float)h * 3.141592653589793f);
        for (int j = 0; j < w; j++) {
            const float* p = rgb + 3 * ((size_t)i * w + j);
            const float lum = 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
            prob[(size_t)i * stride + j] = lum * s;
        }
        alias[(size_t)i * stride + w] = 0;
    }
    for (int i = 0; i < h; i++)
        prob[(size_t)i * stride + w] = envAliasRow(alias + (size_t)i * stride, prob + (size_t)i * stride, w, 1);
    return envAliasRow(alias + w, prob + w, h, (int32_t)stride);
}

namespace {
struct JoeKuo { int s; int a; int m[16]; };
const JoeKuo kJoeKuo[255] = {
#include "sobol_joe_kuo.inc"
};

struct SobolTable {
    uint32_t v[256 * 32];
    SobolTable()
    {
        for (int k = 0; k < 32; k++) v[k] = 0x80000000u >> k;
        for (int d = 1; d < 256; d++) {
            const JoeKuo& p = kJoeKuo[d - 1];
            uint32_t* c = v + d * 32;
            for (int k = 0; k < 32; k++) {
                if (k < p.s) {
                    c[k] = (uint32_t)p.m[k] << (31 - k);
                } else {
                    uint32_t x = c[k - p.s] ^ (c[k - p.s] >> p.s);
                    for (int j = 1; j < p.s; j++)
                        if ((p.a >> (p.s - 1 - j)) & 1) x ^= c[k - j];
                    c[k] = x;
                }
            }
        }
    }
};
}  // namespace

const uint32_t* sobolMatrices()
{
    static const SobolTable table;
    return table.v;
}

uint32_t sobolSample(uint32_t index, int dim, uint32_t scramble)
{
    const uint32_t* m = sobolMatrices() + dim * 32;
    uint32_t r = scramble;
    for (int i = 0; index != 0; index >>= 1, i++)
        if (index & 1) r ^= m[i];
    return r;
}

uint16_t floatToHalf(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t absx = x & 0x7fffffffu;
    if (absx >= 0x7f800000u) return (uint16_t)(sign | (absx > 0x7f800000u ? 0x7e00u : 0x7c00u));
    if (absx >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds to >= 65520 -> inf
    if (absx < 0x38800000u) {                                      // half subnormal or zero
        if (absx < 0x33000000u) return (uint16_t)sign;             // < 2^-25 rounds to 0
        const uint32_t e = absx >> 23;
        const uint32_t mant = (absx & 0x7fffffu) | 0x800000u;
        const uint32_t shift = 126 - e;  // (113 - e) + 13
        uint32_t hm = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1);
        const uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (hm & 1))) hm++;
        return (uint16_t)(sign | hm);
    }
    uint32_t e = (absx >> 23) - 112;
    uint32_t mant = absx & 0x7fffffu;
    uint32_t hm = mant >> 13;
    const uint32_t rem = mant & 0x1fffu;
    uint32_t out = (e << 10) | hm;
    if (rem > 0x1000u || (rem == 0x1000u && (hm & 1))) out++;
    return (uint16_t)(sign | out);
}

float halfToFloat(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1f;
    uint32_t m = h & 0x3ffu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) x = sign;
        else {
            e = 113;
            while (!(m & 0x400u)) { m <<= 1; e--; }
            m &= 0x3ffu;
            x = sign | (e << 23) | (m << 13);
        }
    } else if (e == 31) {
        x = sign | 0x7f800000u | (m << 13);
    } else {
        x = sign | ((e + 112) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

}  // namespace zgl
