// Morton (Z-order) codes of points (MortonOrder.h:11-107 encodes cell coordinates the same way:
// bits interleaved with x lowest, 10 bits per axis in 3D, 16 in 2D).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace nuto_amd
{

// spread the low 10 bits of v to every third bit of 30 / the low 16 bits to every second of 32
inline uint64_t SpreadBits3(uint64_t v)
{
    v &= 0x3ff;
    v = (v | (v << 16)) & 0x30000ffULL;
    v = (v | (v << 8)) & 0x300f00fULL;
    v = (v | (v << 4)) & 0x30c30c3ULL;
    v = (v | (v << 2)) & 0x9249249ULL;
    return v;
}
inline uint64_t SpreadBits2(uint64_t v)
{
    v &= 0xffff;
    v = (v | (v << 8)) & 0x00ff00ffULL;
    v = (v | (v << 4)) & 0x0f0f0f0fULL;
    v = (v | (v << 2)) & 0x33333333ULL;
    v = (v | (v << 1)) & 0x55555555ULL;
    return v;
}

// inverses: gather every third bit (low 30) / every second bit (low 32)
inline uint32_t CompactBits3(uint64_t v)
{
    v &= 0x9249249ULL;
    v = (v | (v >> 2)) & 0x30c30c3ULL;
    v = (v | (v >> 4)) & 0x300f00fULL;
    v = (v | (v >> 8)) & 0x30000ffULL;
    v = (v | (v >> 16)) & 0x3ffULL;
    return (uint32_t)v;
}
inline uint32_t CompactBits2(uint64_t v)
{
    v &= 0x55555555ULL;
    v = (v | (v >> 1)) & 0x33333333ULL;
    v = (v | (v >> 2)) & 0x0f0f0f0fULL;
    v = (v | (v >> 4)) & 0x00ff00ffULL;
    v = (v | (v >> 8)) & 0xffffULL;
    return (uint32_t)v;
}

// Order of n points (dim 1..3, x[i * dim + d]) along a Morton curve over their bounding box,
// quantised to 2^10 (3D) or 2^16 (1D, 2D) steps per axis; ties keep the input order.
// valid: optional per-point flag, points without it sort first with code 0.
inline void MortonSort(int n, int dim, const double* x, std::vector<int>& order, const char* valid = nullptr)
{
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n; i++)
    {
        if (valid && !valid[i])
            continue;
        for (int d = 0; d < dim; d++)
        {
            lo[d] = std::min(lo[d], x[(size_t)i * dim + d]);
            hi[d] = std::max(hi[d], x[(size_t)i * dim + d]);
        }
    }
    const double steps = dim == 3 ? 1024.0 : 65536.0;
    std::vector<uint64_t> key(n);
    for (int i = 0; i < n; i++)
    {
        uint64_t m = 0;
        if (!valid || valid[i])
            for (int d = 0; d < dim; d++)
            {
                const double w = hi[d] - lo[d];
                const double t = w > 0.0 ? (x[(size_t)i * dim + d] - lo[d]) / w : 0.0;
                const uint64_t q = (uint64_t)std::min(steps - 1.0, std::max(0.0, t * steps));
                m |= (dim == 3 ? SpreadBits3(q) << d : SpreadBits2(q) << d);
            }
        key[i] = (m << 32) | (uint64_t)(uint32_t)i;
    }
    std::sort(key.begin(), key.end());
    order.resize(n);
    for (int i = 0; i < n; i++)
        order[i] = (int)(key[i] & 0xffffffffULL);
}

}  // namespace nuto_amd
