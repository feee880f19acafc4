// C++ test of nuto_amd::host::MeshGmsh (the MeshGmsh mirror over the C ABI) on the reference's
// mesh test files (test/mechanics/mesh/MeshGmsh.cpp:7-125): physical groups by id and name, the
// binary import, node counts, the rejected inverted triangle. Host only, no GPU.
// usage: test_host_gmsh <directory of the .msh files>
#include <cstdio>
#include <string>

#include "../../nuto_amd/host/nuto_assembler.h"

using nuto_amd::host::Exception;
using nuto_amd::host::MeshGmsh;

static int gFailures = 0;
#define EXPECT(cond, what)                                                                                         \
    do                                                                                                             \
    {                                                                                                              \
        if (!(cond))                                                                                               \
        {                                                                                                          \
            std::printf("FAIL %s (%s:%d)\n", what, __FILE__, __LINE__);                                            \
            gFailures++;                                                                                           \
        }                                                                                                          \
    } while (0)

template <typename F>
static bool Throws(F f)
{
    try
    {
        f();
    }
    catch (const Exception&)
    {
        return true;
    }
    return false;
}

int main(int argc, char** argv)
{
    if (argc < 2)
    {
        std::printf("usage: %s <mesh directory>\n", argv[0]);
        return 2;
    }
    const std::string dir = std::string(argv[1]) + "/";

    MeshGmsh quad(dir + "quad.msh");
    EXPECT(!quad.GetPhysicalGroup(1).empty(), "group 1");
    EXPECT(Throws([&] { quad.GetPhysicalGroup(-42); }), "group -42 throws");
    EXPECT(quad.GetPhysicalGroup("domain") == quad.GetPhysicalGroup(1), "group by name");
    EXPECT(Throws([&] { quad.GetPhysicalGroup("are_you_there?"); }), "unknown name throws");
    EXPECT(!Throws([&] { MeshGmsh m(dir + "quadNoncontiguous.msh"); }), "non-contiguous node ids");

    const struct
    {
        const char* file;
        int nodes;
    } cases[] = {{"quad1.msh", 4},   {"quadSerendipity2.msh", 8}, {"triangle1.msh", 3}, {"triangle2.msh", 6},
                 {"line1.msh", 2},   {"line2.msh", 3},            {"tet1.msh", 4},      {"hex1.msh", 8},
                 {"prism1.msh", 6},  {"pyramid1.msh", 5},         {"quadL2.msh", 9},    {"brickL2.msh", 27}};
    for (const auto& c : cases)
    {
        MeshGmsh m(dir + c.file);
        EXPECT(m.NumNodes() == c.nodes, c.file);
        EXPECT((int)m.ElementNodes(0).size() == c.nodes, c.file);
    }

    MeshGmsh bin(dir + "binary.msh");
    EXPECT(bin.NumNodes() == 8 && bin.NumElements() == 7 && bin.ElementNodes(0).size() == 3, "binary import");
    EXPECT(Throws([&] { MeshGmsh m(dir + "invertedTriangle.msh"); }), "inverted triangle throws");

    MeshGmsh hex(dir + "hex1.msh");
    auto cells = hex.Cells(NUTO_HEX8);
    EXPECT(cells.NumCells() == 1 && cells.conn == hex.ElementNodes(0), "hex8 cell group");
    EXPECT(Throws([&] { quad.Cells(NUTO_HEX8); }), "3D cells of a 2D mesh throw");

    if (gFailures == 0)
        std::printf("ok\n");
    return gFailures ? 1 : 0;
}
