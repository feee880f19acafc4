// Tiled hex8 gradient-damage kernel plan (hot path). Placeholder: not usable yet, the generic
// row-owner kernel serves every combination.
#pragma once

#include <hip/hip_runtime.h>

#include "assemble_rows.hpp"
#include "mesh.hpp"

namespace nuto_amd
{

struct GdmHex8Plan
{
    bool ready = false;
    void Release()
    {
        ready = false;
    }
    bool Usable(int, int, int, int, bool) const
    {
        return false;
    }
    void Prepare(int, int, int, int, bool, const NodeGraph&, int, hipStream_t)
    {
    }
    void Launch(const RowArgs&, bool, bool, hipStream_t)
    {
    }
};

}  // namespace nuto_amd
