// Batched assembly of many independent meshes.
//
// A batch groups its meshes by (coordinate element, dof element); the meshes of one group form a
// block-diagonal super-mesh (node, cell and dof numbers offset mesh by mesh), so its CSR is the
// concatenation of the meshes' CSRs and every mesh's rows and nonzeros are contiguous slices.
// The row-owner body (AssembleRowsBody) needs no change for this: the meshes share no node.
//
// kBatchRows is one launch over the tiles of all groups: block b assembles the rows of the nodes
// [first, end) of one group with the body of that group's element type (the body is chosen per
// block, uniformly, so no lane diverges on it). Tiles of the heaviest element types come first
// so the long blocks start early. kBatchRowsRange assembles one mesh (the plain per-mesh launch
// that the batch replaces).
#pragma once

#include <hip/hip_runtime.h>

#include "assemble_rows.hpp"

namespace nuto_amd
{

struct BatchTile
{
    int group;  // index into the RowArgs array
    int first;  // first node (group numbering)
    int end;    // one past the last node
    int type;   // BatchType of the group
};

// element types of the batched kernel: (coordinate element, dof element, IPs of the default rule)
enum BatchType
{
    kBtTet4 = 0,   // tet4 / tet4, 1 IP
    kBtTet10 = 1,  // tet4 / tet10, 4 IPs
    kBtHex8 = 2,   // hex8 / hex8, 2x2x2 Gauss
    kBtHex20 = 3,  // hex8 / hex20, 3x3x3 Gauss
};

// One body per element type, kept out of line so that each is register-allocated on its own
// (inlined together they would share one allocation and spill).
template <int KIND, int TYPE>
__device__ __noinline__ void BatchRowsType(const RowArgs* A, int p, int pEnd)
{
    if constexpr (TYPE == kBtTet4)
        AssembleRowsBody<KIND, 3, 4, 4, 1, 4, true, true>(*A, p, pEnd);
    else if constexpr (TYPE == kBtTet10)
        AssembleRowsBody<KIND, 3, 4, 10, 4, 5, true, true>(*A, p, pEnd);
    else if constexpr (TYPE == kBtHex8)
        AssembleRowsBody<KIND, 3, 8, 8, 8, 8, true, true>(*A, p, pEnd);
    else
        AssembleRowsBody<KIND, 3, 8, 20, 27, 5, true, true>(*A, p, pEnd);
}

template <int KIND>
__device__ inline void BatchRowsBody(const RowArgs* A, int type, int p, int pEnd)
{
    switch (type)
    {
    case kBtTet4:
        BatchRowsType<KIND, kBtTet4>(A, p, pEnd);
        break;
    case kBtTet10:
        BatchRowsType<KIND, kBtTet10>(A, p, pEnd);
        break;
    case kBtHex8:
        BatchRowsType<KIND, kBtHex8>(A, p, pEnd);
        break;
    default:
        BatchRowsType<KIND, kBtHex20>(A, p, pEnd);
        break;
    }
}

template <int KIND, int BS>
__global__ void __launch_bounds__(BS) kBatchRows(const RowArgs* __restrict__ args, const BatchTile* __restrict__ tiles)
{
    const BatchTile t = tiles[blockIdx.x];
    BatchRowsBody<KIND>(args + t.group, t.type, t.first + (int)threadIdx.x, t.end);
}

template <int KIND, int TYPE>
__global__ void __launch_bounds__(64) kBatchRowsRange(const RowArgs A, int first, int end)
{
    if constexpr (TYPE == kBtTet4)
        AssembleRowsBody<KIND, 3, 4, 4, 1, 4, true, true>(A, first + (int)(blockIdx.x * 64 + threadIdx.x), end);
    else if constexpr (TYPE == kBtTet10)
        AssembleRowsBody<KIND, 3, 4, 10, 4, 5, true, true>(A, first + (int)(blockIdx.x * 64 + threadIdx.x), end);
    else if constexpr (TYPE == kBtHex8)
        AssembleRowsBody<KIND, 3, 8, 8, 8, 8, true, true>(A, first + (int)(blockIdx.x * 64 + threadIdx.x), end);
    else
        AssembleRowsBody<KIND, 3, 8, 20, 27, 5, true, true>(A, first + (int)(blockIdx.x * 64 + threadIdx.x), end);
}

}  // namespace nuto_amd
