// Host-side set-up of the assembly path: structured meshes, first-touch dof layers and the
// node graph from which the CSR pattern and the kernels' gather maps are derived.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace nuto_amd
{

// Structured mesh of the box [0,lx]x[0,ly]x[0,lz] (2D: quad4, 3D: hex8, or tet4 = 6 tets per
// brick). Node coordinates are computed as double(i)/n and then scaled, like
// UnitMeshFem::CreateQuads/CreateBricks followed by Transform (UnitMeshFem.cpp:22-125).
// jitter > 0 displaces every node by a seeded uniform offset in [-jitter, jitter] * spacing.
bool StructuredMesh(int coordType, int nx, int ny, int nz, double lx, double ly, double lz, double jitter,
                    uint64_t seed, std::vector<double>& coords, std::vector<int>& conn, std::string& err);

// First-touch dof layer over a coordinate mesh (AddDofInterpolation, MeshFemDofConvert.cpp:164-210,
// numbered by DofNumbering::Build over MeshFem::NodesTotal, DofNumbering.cpp:7-58). Vertices are
// identified with coordinate nodes; the edge nodes of quadratic layers (tet10 over tet4, hex20 over
// hex8) by their two end vertices. Returns the per-cell node ranks in first-touch order.
bool DofLayer(int coordType, int nCells, const int* coordConn, int dofType, std::vector<int>& ranks, int& nNodes,
              std::string& err);

// Node graph of one dof layer: per node the incident cells in group order, the sorted
// neighbourhood (the block sparsity row), and for every (node, incident cell, local node b) the
// position of b's node in the neighbourhood with a "first contribution" flag in bit 15.
struct NodeGraph
{
    int nNodes = 0, nnD = 0;
    std::vector<int> incPtr, incCell;
    std::vector<unsigned char> incLocal;
    std::vector<int> adjPtr, adj;
    std::vector<unsigned short> slot;
};

// Large graphs are built on up to 16 threads (maxThreads > 0 caps that; 1: the calling thread only).
bool BuildNodeGraph(int nCells, int nnD, const int* ranks, int nNodes, NodeGraph& g, std::string& err,
                    int maxThreads = 0);

// Monolithic CSR row pointer over fields sharing one node graph: rows are field 0 nodes x ncomp0
// then field 1 nodes x ncomp1 (DofMatrixSparseConvertEigen.h:40-67 block order), every row of
// node p holds (sum of ncomp) * |adj(p)| columns.
bool BuildRowPtr(const NodeGraph& g, int nFields, const int* ncomp, std::vector<int>& rowPtr, long long& nnz,
                 std::string& err);

uint64_t SplitMix64(uint64_t seed, uint64_t counter);
double UniformAt(uint64_t seed, uint64_t counter);  // [0,1), top 53 bits

}  // namespace nuto_amd
