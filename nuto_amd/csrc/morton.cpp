// Morton (Z-order) codes and the Morton order of points behind the C ABI of include/nuto_amd.h
// (NuTo::MortonOrder, MortonOrder.h:11-107). Host code only.
#include <cstdint>
#include <vector>

#include "../../include/nuto_amd.h"
#include "morton.hpp"

using nuto_amd::CompactBits2;
using nuto_amd::CompactBits3;
using nuto_amd::SpreadBits2;
using nuto_amd::SpreadBits3;

extern "C" {

unsigned nuto_morton_encode3d(unsigned x, unsigned y, unsigned z)
{
    return (unsigned)(SpreadBits3(x) | (SpreadBits3(y) << 1) | (SpreadBits3(z) << 2));
}

unsigned nuto_morton_encode2d(unsigned x, unsigned y)
{
    return (unsigned)(SpreadBits2(x) | (SpreadBits2(y) << 1));
}

void nuto_morton_decode3d(unsigned code, unsigned* x, unsigned* y, unsigned* z)
{
    if (x)
        *x = CompactBits3(code);
    if (y)
        *y = CompactBits3(code >> 1);
    if (z)
        *z = CompactBits3(code >> 2);
}

void nuto_morton_decode2d(unsigned code, unsigned* x, unsigned* y)
{
    if (x)
        *x = CompactBits2(code);
    if (y)
        *y = CompactBits2(code >> 1);
}

unsigned nuto_morton_neighbor3d(unsigned key, int dx, int dy, int dz)
{
    const long long x = CompactBits3(key) + (long long)dx, y = CompactBits3(key >> 1) + (long long)dy,
                    z = CompactBits3(key >> 2) + (long long)dz;
    if (x < 0 || y < 0 || z < 0)
        return 0xffffffffu;
    return nuto_morton_encode3d((unsigned)x, (unsigned)y, (unsigned)z);
}

int nuto_morton_order(int n, int dim, const double* coords, int* order)
{
    if (n < 0 || dim < 1 || dim > 3 || (n > 0 && (!coords || !order)))
        return NUTO_ERR_ARGUMENT;
    std::vector<int> o;
    nuto_amd::MortonSort(n, dim, coords, o);
    for (int i = 0; i < n; i++)
        order[i] = o[i];
    return NUTO_OK;
}

}  // extern "C"
