// The node graph (incidence, neighbourhoods, slots) built on several threads equals the one built
// on the calling thread alone, array for array, for structured hex8 / hex20 / tet4 / tet10 layers
// (host code only: nuto_amd/csrc/mesh.cpp).
#include <cstdio>
#include <string>
#include <vector>

#include "../../nuto_amd/csrc/elements.hpp"
#include "../../nuto_amd/csrc/mesh.hpp"

using namespace nuto_amd;

static bool Same(const NodeGraph& a, const NodeGraph& b)
{
    return a.nNodes == b.nNodes && a.nnD == b.nnD && a.incPtr == b.incPtr && a.incCell == b.incCell &&
           a.incLocal == b.incLocal && a.adjPtr == b.adjPtr && a.adj == b.adj && a.slot == b.slot;
}

int main()
{
    struct Case
    {
        int coord, dof, n;
    } cases[] = {{kHex8, kHex8, 40}, {kHex8, kHex20, 24}, {kTet4, kTet4, 32}, {kTet4, kTet10, 16}};
    for (const Case& c : cases)
    {
        std::vector<double> X;
        std::vector<int> C, ranks;
        std::string err;
        if (!StructuredMesh(c.coord, c.n, c.n, c.n, 1.0, 1.0, 1.0, 0.1, 7, X, C, err))
        {
            std::printf("mesh: %s\n", err.c_str());
            return 1;
        }
        const int nnc = c.coord == kHex8 ? 8 : 4;
        const int nCells = (int)(C.size() / nnc);
        int nn = 0;
        if (!DofLayer(c.coord, nCells, C.data(), c.dof, ranks, nn, err))
        {
            std::printf("dof layer: %s\n", err.c_str());
            return 1;
        }
        const int nnd = (int)(ranks.size() / nCells);
        NodeGraph one, many;
        if (!BuildNodeGraph(nCells, nnd, ranks.data(), nn, one, err, 1) ||
            !BuildNodeGraph(nCells, nnd, ranks.data(), nn, many, err, 8))
        {
            std::printf("graph: %s\n", err.c_str());
            return 1;
        }
        if (!Same(one, many))
        {
            std::printf("coordinate type %d, dof type %d: threaded graph differs\n", c.coord, c.dof);
            return 1;
        }
        std::printf("coordinate type %d, dof type %d: %d cells, %d nodes, %zu neighbours\n", c.coord, c.dof,
                    nCells, nn, one.adj.size());
    }
    std::printf("ok\n");
    return 0;
}
