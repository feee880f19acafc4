// C ABI of the MI355X assembly path (declared in include/nuto_amd.h).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/nuto_amd.h"
#include "assemble_rows.hpp"
#include "elements.hpp"
#include "gdm_hex8.hpp"
#include "gdm_hex8_vjp.hpp"
#include "mesh.hpp"

using namespace nuto_amd;

namespace
{

struct Field
{
    bool set = false;
    int type = 0, ncomp = 0, nDofs = 0, nNodes = 0;
    std::vector<int> ranks;  // n_cells x nodes(type)
};

struct Failure
{
    int code;
    std::string msg;
};

}  // namespace

struct nuto_ctx
{
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    // cells
    bool cellsSet = false;
    int coordType = 0, dim = 0, nCells = 0, nCoordNodes = 0;
    DevBuf x, y, z, cconn, cellId;
    std::vector<int> hcellId;  // host copy of the cell ids (pattern-time tables)
    bool hasCellIds = false;
    int maxCellId = 0;
    Field f[2];
    DevBuf dconn;
    // integrand and rule
    bool integrandSet = false;
    nuto_params params{};
    bool ruleSet = false;
    int ruleKind = 0, ruleN = 0, nip = 0;
    DevBuf tab;
    bool tabValid = false;
    int historyLd = -1;
    DevBuf thE, thFt, thGf;
    bool hasTheta = false;
    // pattern
    bool patternBuilt = false;
    NodeGraph graph;
    DevBuf incPtr, incCell, incLocal, slot, adjPtr, adj, rowPtr, col;
    long long nnz = 0;
    int nRows = 0;
    // hex8 gradient-damage node kernel (prepared lazily from the node graph)
    GdmHex8Plan gdm;
    bool gdmTried = false;
};

namespace
{

void Check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw Failure{e == hipErrorOutOfMemory ? NUTO_ERR_MEMORY : NUTO_ERR_DEVICE,
                      std::string(what) + ": " + hipGetErrorString(e)};
}

void Require(bool ok, int code, const std::string& msg)
{
    if (!ok)
        throw Failure{code, msg};
}

template <class F>
int Guard(nuto_ctx* ctx, F&& fn)
{
    if (!ctx)
        return NUTO_ERR_ARGUMENT;
    try
    {
        Check(hipSetDevice(ctx->device), "hipSetDevice");
        fn();
        ctx->err.clear();
        return NUTO_OK;
    }
    catch (const Failure& f)
    {
        ctx->err = f.msg;
        return f.code;
    }
    catch (const std::bad_alloc&)
    {
        ctx->err = "host allocation failed";
        return NUTO_ERR_MEMORY;
    }
    catch (const std::exception& e)
    {
        ctx->err = e.what();
        return NUTO_ERR_ARGUMENT;
    }
}

// NUTO_AMD_SYNC_CHECK=1 synchronizes after every kernel and reports asynchronous errors at the
// call that caused them (debugging aid; off by default).
void SyncCheck(nuto_ctx* ctx, const char* what)
{
    static const bool on = [] {
        const char* v = std::getenv("NUTO_AMD_SYNC_CHECK");
        return v && v[0] == '1';
    }();
    if (on)
    {
        Check(hipStreamSynchronize(ctx->stream), what);
        Check(hipGetLastError(), what);
    }
}

// NUTO_AMD_GENERIC=1 routes every call through the generic row-owner kernel (used by the tests
// to check the specialised kernels against it).
bool ForceGeneric()
{
    const char* v = std::getenv("NUTO_AMD_GENERIC");
    return v && v[0] == '1';
}

template <class T>
void Upload(nuto_ctx* ctx, DevBuf& b, const T* src, size_t n, const char* what)
{
    Check(b.Alloc(n * sizeof(T)), what);
    if (n)
        Check(hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream), what);
    Check(hipStreamSynchronize(ctx->stream), what);
}

void InvalidatePattern(nuto_ctx* ctx)
{
    ctx->patternBuilt = false;
    ctx->gdm.Release();
    ctx->gdmTried = false;
}

// IP tables: per IP w, dNc[nnC][dim], N[nnD], dNd[nnD][dim] (shape tables of the coordinate
// element and of the slot-0 dof element, elements.hpp).
void BuildTables(nuto_ctx* ctx)
{
    Require(ctx->cellsSet && ctx->f[0].set && ctx->ruleSet, NUTO_ERR_STATE, "cells, dof type 0 and rule must be set");
    Rule r;
    Require(MakeRule(ctx->ruleKind, ctx->dim, ctx->ruleN, r), NUTO_ERR_UNSUPPORTED, "unsupported integration rule");
    const int nnC = NodesOf(ctx->coordType), nnD = NodesOf(ctx->f[0].type), dim = ctx->dim;
    std::vector<double> t;
    for (int ip = 0; ip < r.Num(); ip++)
    {
        const double* xi = &r.xi[3 * ip];
        double Nc[20], dNc[60], Nd[20], dNd[60];
        EvalShape(ctx->coordType, xi, Nc, dNc);
        EvalShape(ctx->f[0].type, xi, Nd, dNd);
        t.push_back(r.w[ip]);
        for (int k = 0; k < nnC; k++)
            for (int d = 0; d < dim; d++)
                t.push_back(dNc[3 * k + d]);
        for (int b = 0; b < nnD; b++)
            t.push_back(Nd[b]);
        for (int b = 0; b < nnD; b++)
            for (int d = 0; d < dim; d++)
                t.push_back(dNd[3 * b + d]);
    }
    ctx->nip = r.Num();
    Upload(ctx, ctx->tab, t.data(), t.size(), "ip tables");
    ctx->tabValid = true;
}

__global__ void kFillCols(const int* __restrict__ adjPtr, const int* __restrict__ adj, const int* __restrict__ rowPtr,
                          int nNodes, int nFields, int nc0, int nc1, int nU, int* __restrict__ col)
{
    // one block per dof node; all rows of node p share the column list
    const int p = blockIdx.x;
    if (p >= nNodes)
        return;
    const int a0 = adjPtr[p], nA = adjPtr[p + 1] - a0;
    const int W = nc0 + (nFields > 1 ? nc1 : 0);
    const int len = W * nA;
    const int nrows = nc0 + (nFields > 1 ? nc1 : 0);
    for (int rr = 0; rr < nrows; rr++)
    {
        const int row = rr < nc0 ? p * nc0 + rr : nU + p * nc1 + (rr - nc0);
        int* c = col + rowPtr[row];
        for (int k = threadIdx.x; k < len; k += blockDim.x)
        {
            int v;
            if (k < nc0 * nA)
                v = adj[a0 + k / nc0] * nc0 + k % nc0;
            else
            {
                const int kk = k - nc0 * nA;
                v = nU + adj[a0 + kk / nc1] * nc1 + kk % nc1;
            }
            c[k] = v;
        }
    }
}

RowArgs MakeRowArgs(nuto_ctx* ctx, const double* u, const double* e, const double* kappa, double* R, double* K)
{
    RowArgs a{};
    a.x = ctx->x.As<double>();
    a.y = ctx->y.As<double>();
    a.z = ctx->dim == 3 ? ctx->z.As<double>() : nullptr;
    a.cconn = ctx->cconn.As<int>();
    a.dconn = ctx->dconn.As<int>();
    a.cellId = ctx->hasCellIds ? ctx->cellId.As<int>() : nullptr;
    a.nNodes = ctx->f[0].nNodes;
    a.nCells = ctx->nCells;
    a.kappaLd = ctx->historyLd >= 0 ? ctx->historyLd : ctx->nCells;
    a.nU = ctx->f[0].nNodes * ctx->f[0].ncomp;
    a.incPtr = ctx->incPtr.As<int>();
    a.incCell = ctx->incCell.As<int>();
    a.incLocal = ctx->incLocal.As<unsigned char>();
    a.slot = ctx->slot.As<unsigned short>();
    a.adjPtr = ctx->adjPtr.As<int>();
    a.rowPtr = ctx->rowPtr.As<int>();
    a.u = u;
    a.e = e;
    a.kappa = kappa;
    if (ctx->hasTheta)
    {
        a.thE = ctx->thE.As<double>();
        a.thFt = ctx->thFt.As<double>();
        a.thGf = ctx->thGf.As<double>();
    }
    a.tab = ctx->tab.As<double>();
    const nuto_params& p = ctx->params;
    a.P = MakeDerived(LawParams{p.E, p.nu, p.k, p.kappa0, p.beta, p.alpha, p.c, p.damage_application,
                                p.plane_state, 0.0, 0.0, 0.0, 0.0});
    a.R = R;
    a.K = K;
    return a;
}

template <int KIND, int DIM, int NNC, int NND, int NIP, int BH>
void LaunchRowsT(nuto_ctx* ctx, const RowArgs& a, bool wr, bool wk)
{
    const int n = a.nNodes;
    if (n == 0)
        return;
    dim3 grid((n + 63) / 64), block(64);
    if (wr && wk)
        hipLaunchKernelGGL((kAssembleRows<KIND, DIM, NNC, NND, NIP, BH, true, true>), grid, block, 0, ctx->stream, a);
    else if (wr)
        hipLaunchKernelGGL((kAssembleRows<KIND, DIM, NNC, NND, NIP, BH, true, false>), grid, block, 0, ctx->stream, a);
    else
        hipLaunchKernelGGL((kAssembleRows<KIND, DIM, NNC, NND, NIP, BH, false, true>), grid, block, 0, ctx->stream, a);
    Check(hipGetLastError(), "assembly launch");
    SyncCheck(ctx, "assembly kernel");
}

template <int KIND, int DIM, int NNC, int NND, int NIP>
void LaunchHistoryT(nuto_ctx* ctx, const RowArgs& a, double* kappa)
{
    if constexpr (KIND != kLinearElastic && DIM == 3)
    {
        const long long n = (long long)a.nCells * NIP;
        if (n == 0)
            return;
        hipLaunchKernelGGL((kUpdateHistory<KIND, DIM, NNC, NND, NIP>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           ctx->stream, a, kappa);
        Check(hipGetLastError(), "history launch");
    }
}

// Dispatch over the instantiated (integrand, coordinate element, dof element, rule) combinations.
// mode 0: assembly, mode 1: history update
void Dispatch(nuto_ctx* ctx, const RowArgs& a, bool wr, bool wk, int mode, double* kappaOut)
{
    const int kind = ctx->params.kind, ct = ctx->coordType, dt = ctx->f[0].type, nip = ctx->nip;
#define NUTO_CASE(KIND, DIM, CT, NNC, DT, NND, NIP, BH)                       \
    if (kind == KIND && ct == CT && dt == DT && nip == NIP && ctx->dim == DIM) \
    {                                                                          \
        if (mode == 0)                                                         \
            LaunchRowsT<KIND, DIM, NNC, NND, NIP, BH>(ctx, a, wr, wk);         \
        else                                                                   \
            LaunchHistoryT<KIND, DIM, NNC, NND, NIP>(ctx, a, kappaOut);        \
        return;                                                                \
    }
    NUTO_CASE(kLinearElastic, 2, kQuad4, 4, kQuad4, 4, 4, 4)
    NUTO_CASE(kLinearElastic, 2, kQuad4, 4, kQuad4, 4, 1, 4)
    NUTO_CASE(kLinearElastic, 2, kQuad4, 4, kQuad4, 4, 9, 4)
    NUTO_CASE(kLinearElastic, 3, kHex8, 8, kHex8, 8, 8, 8)
    NUTO_CASE(kLinearElastic, 3, kTet4, 4, kTet4, 4, 1, 4)
    NUTO_CASE(kLinearElastic, 3, kTet4, 4, kTet10, 10, 4, 5)
    NUTO_CASE(kLinearElastic, 3, kHex8, 8, kHex20, 20, 27, 5)
    NUTO_CASE(kGradientDamage, 3, kHex8, 8, kHex8, 8, 8, 4)
    NUTO_CASE(kLocalDamage, 3, kHex8, 8, kHex8, 8, 8, 8)
    NUTO_CASE(kLocalDamage, 3, kTet4, 4, kTet4, 4, 1, 4)
    NUTO_CASE(kLocalDamage, 3, kTet4, 4, kTet10, 10, 4, 5)
    NUTO_CASE(kLocalDamage, 3, kHex8, 8, kHex20, 20, 27, 5)
#undef NUTO_CASE
    throw Failure{NUTO_ERR_UNSUPPORTED, "no kernel for this integrand / element / rule combination"};
}

void CheckReady(nuto_ctx* ctx, const double* u, const double* e, const double* kappa)
{
    // SimpleAssembler::ThrowOnZeroDofNumbering (SimpleAssembler.cpp:16-23)
    Require(ctx->cellsSet, NUTO_ERR_STATE, "no cells set");
    Require(ctx->f[0].set, NUTO_ERR_STATE, "You did not provide a dof numbering for dof type slot 0");
    Require(ctx->integrandSet, NUTO_ERR_STATE, "no integrand set");
    Require(ctx->ruleSet, NUTO_ERR_STATE, "no integration rule set");
    Require(ctx->patternBuilt, NUTO_ERR_STATE, "pattern not built (call nuto_build_pattern)");
    Require(u != nullptr, NUTO_ERR_ARGUMENT, "u is NULL");
    if (ctx->params.kind == kGradientDamage)
    {
        Require(ctx->f[1].set, NUTO_ERR_STATE, "You did not provide a dof numbering for dof type slot 1");
        Require(e != nullptr, NUTO_ERR_ARGUMENT, "nonlocal equivalent strain vector is NULL");
    }
    if (ctx->params.kind != kLinearElastic)
        Require(kappa != nullptr, NUTO_ERR_ARGUMENT, "history (kappa) is NULL");
    if (ctx->params.kind != kLinearElastic)
    {
        const int ld = ctx->historyLd >= 0 ? ctx->historyLd : ctx->nCells;
        Require(ld > ctx->maxCellId, NUTO_ERR_ARGUMENT,
                "history leading dimension must exceed the largest cell id (nuto_set_history_ld)");
    }
}

// Builds the hex8 node kernel's tables once per pattern (a no-op for other meshes / integrands).
void PrepareGdm(nuto_ctx* ctx)
{
    if (ctx->gdmTried)
        return;
    hipError_t err = hipSuccess;
    ctx->gdm.Prepare(ctx->params.kind, ctx->coordType, ctx->f[0].type, ctx->nip, ctx->graph, ctx->nCells,
                     ctx->hasCellIds ? ctx->hcellId.data() : nullptr, ctx->x.As<double>(), ctx->y.As<double>(),
                     ctx->z.As<double>(), ctx->cconn.As<int>(), ctx->tab.As<double>(), ctx->stream, err);
    Check(err, "hex8 node tables");
    ctx->gdmTried = true;
}

void Assemble(nuto_ctx* ctx, const double* u, const double* e, const double* kappa, double* R, double* K,
              bool needR, bool needK)
{
    CheckReady(ctx, u, e, kappa);
    Require(!needR || R != nullptr, NUTO_ERR_ARGUMENT, "R is NULL");
    Require(!needK || K != nullptr, NUTO_ERR_ARGUMENT, "K is NULL");
    if (!ctx->tabValid)
        BuildTables(ctx);
    const bool wr = R != nullptr, wk = K != nullptr;
    if (!wr && !wk)
        return;
    PrepareGdm(ctx);
    if (!ForceGeneric() && ctx->gdm.Usable(ctx->params.kind, ctx->coordType, ctx->f[0].type, ctx->nip))
    {
        RowArgs a = MakeRowArgs(ctx, u, e, kappa, R, K);
        ctx->gdm.Launch(a, wr, wk, ctx->stream);
        Check(hipGetLastError(), "gdm hex8 launch");
        SyncCheck(ctx, "gdm hex8 kernel");
        return;
    }
    RowArgs a = MakeRowArgs(ctx, u, e, kappa, R, K);
    Dispatch(ctx, a, wr, wk, 0, nullptr);
}

}  // namespace

// =================================================================================================
// C ABI
// =================================================================================================
extern "C" {

int nuto_mesh_structured(int coord_type, int nx, int ny, int nz, double lx, double ly, double lz, double jitter,
                         unsigned long long seed, int* n_nodes, int* n_cells, double* coords, int* conn)
{
    try
    {
        std::vector<double> X;
        std::vector<int> C;
        std::string err;
        if (!StructuredMesh(coord_type, nx, ny, nz, lx, ly, lz, jitter, seed, X, C, err))
            return NUTO_ERR_ARGUMENT;
        const int dim = coord_type == kQuad4 ? 2 : 3;
        if (n_nodes)
            *n_nodes = (int)(X.size() / dim);
        if (n_cells)
            *n_cells = (int)(C.size() / NodesOf(coord_type));
        if (coords)
            std::memcpy(coords, X.data(), X.size() * sizeof(double));
        if (conn)
            std::memcpy(conn, C.data(), C.size() * sizeof(int));
        return NUTO_OK;
    }
    catch (...)
    {
        return NUTO_ERR_MEMORY;
    }
}

int nuto_dof_layer(int coord_type, int n_cells, const int* coord_conn, int dof_type, int* cell_ranks,
                   int* n_dof_nodes)
{
    try
    {
        if (!coord_conn || n_cells < 0 || NodesOf(coord_type) == 0 || NodesOf(dof_type) == 0)
            return NUTO_ERR_ARGUMENT;
        std::vector<int> ranks;
        int nn = 0;
        std::string err;
        if (!DofLayer(coord_type, n_cells, coord_conn, dof_type, ranks, nn, err))
            return NUTO_ERR_ARGUMENT;
        if (n_dof_nodes)
            *n_dof_nodes = nn;
        if (cell_ranks)
            std::memcpy(cell_ranks, ranks.data(), ranks.size() * sizeof(int));
        return NUTO_OK;
    }
    catch (...)
    {
        return NUTO_ERR_MEMORY;
    }
}

int nuto_create(nuto_ctx** out, int device)
{
    if (!out)
        return NUTO_ERR_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return NUTO_ERR_DEVICE;
    if (device < 0 || device >= count)
        return NUTO_ERR_ARGUMENT;
    nuto_ctx* ctx = new (std::nothrow) nuto_ctx;
    if (!ctx)
        return NUTO_ERR_MEMORY;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking) != hipSuccess)
    {
        delete ctx;
        return NUTO_ERR_DEVICE;
    }
    ctx->stream = ctx->own;
    *out = ctx;
    return NUTO_OK;
}

int nuto_destroy(nuto_ctx* ctx)
{
    if (!ctx)
        return NUTO_ERR_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->own)
        (void)hipStreamDestroy(ctx->own);
    delete ctx;
    return NUTO_OK;
}

const char* nuto_last_error(const nuto_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

int nuto_set_stream(nuto_ctx* ctx, void* stream)
{
    // NULL is HIP's default (null) stream, which orders with every blocking stream of the device
    return Guard(ctx, [&] { ctx->stream = (hipStream_t)stream; });
}

int nuto_synchronize(nuto_ctx* ctx)
{
    return Guard(ctx, [&] { Check(hipStreamSynchronize(ctx->stream), "synchronize"); });
}

int nuto_set_cells(nuto_ctx* ctx, int coord_type, int n_cells, int n_coord_nodes, const double* coords,
                   const int* coord_conn, const int* cell_ids)
{
    return Guard(ctx, [&] {
        Require(coord_type == kQuad4 || coord_type == kHex8 || coord_type == kTet4, NUTO_ERR_ARGUMENT,
                "coordinate elements must be quad4, hex8 or tet4");
        Require(n_cells >= 0 && n_coord_nodes >= 0, NUTO_ERR_ARGUMENT, "negative sizes");
        Require(coords && coord_conn, NUTO_ERR_ARGUMENT, "coords / coord_conn is NULL");
        const int nn = NodesOf(coord_type), dim = DimOf(coord_type);
        for (size_t i = 0; i < (size_t)n_cells * nn; i++)
            Require(coord_conn[i] >= 0 && coord_conn[i] < n_coord_nodes, NUTO_ERR_ARGUMENT,
                    "coordinate connectivity out of range");
        std::vector<double> sx(n_coord_nodes), sy(n_coord_nodes), sz(dim == 3 ? n_coord_nodes : 0);
        for (int i = 0; i < n_coord_nodes; i++)
        {
            sx[i] = coords[(size_t)i * dim];
            sy[i] = coords[(size_t)i * dim + 1];
            if (dim == 3)
                sz[i] = coords[(size_t)i * dim + 2];
        }
        InvalidatePattern(ctx);
        ctx->f[0] = Field();
        ctx->f[1] = Field();
        ctx->tabValid = false;
        Upload(ctx, ctx->x, sx.data(), sx.size(), "coords");
        Upload(ctx, ctx->y, sy.data(), sy.size(), "coords");
        if (dim == 3)
            Upload(ctx, ctx->z, sz.data(), sz.size(), "coords");
        Upload(ctx, ctx->cconn, coord_conn, (size_t)n_cells * nn, "connectivity");
        ctx->hcellId.clear();
        if (cell_ids)
            ctx->hcellId.assign(cell_ids, cell_ids + n_cells);
        ctx->hasCellIds = cell_ids != nullptr;
        ctx->maxCellId = n_cells - 1;
        if (cell_ids)
        {
            ctx->maxCellId = -1;
            for (int c = 0; c < n_cells; c++)
            {
                Require(cell_ids[c] >= 0, NUTO_ERR_ARGUMENT, "negative cell id");
                ctx->maxCellId = std::max(ctx->maxCellId, cell_ids[c]);
            }
            Upload(ctx, ctx->cellId, cell_ids, (size_t)n_cells, "cell ids");
        }
        ctx->coordType = coord_type;
        ctx->dim = dim;
        ctx->nCells = n_cells;
        ctx->nCoordNodes = n_coord_nodes;
        ctx->cellsSet = true;
        ctx->hasTheta = false;
    });
}

int nuto_set_dof_type(nuto_ctx* ctx, int slot, int dof_type, int ncomp, int n_dofs, const int* cell_dofs)
{
    return Guard(ctx, [&] {
        Require(ctx->cellsSet, NUTO_ERR_STATE, "set the cells first");
        Require(slot == 0 || slot == 1, NUTO_ERR_ARGUMENT, "slot must be 0 or 1");
        Require(NodesOf(dof_type) > 0 && FamilyOf(dof_type) == FamilyOf(ctx->coordType), NUTO_ERR_ARGUMENT,
                "dof interpolation does not match the coordinate element shape");
        Require(ncomp >= 1 && ncomp <= 3 && n_dofs >= 0 && cell_dofs, NUTO_ERR_ARGUMENT, "bad dof type arguments");
        const int nn = NodesOf(dof_type);
        Field fd;
        fd.set = true;
        fd.type = dof_type;
        fd.ncomp = ncomp;
        fd.nDofs = n_dofs;
        fd.ranks.resize((size_t)ctx->nCells * nn);
        int maxRank = -1;
        for (int c = 0; c < ctx->nCells; c++)
            for (int a = 0; a < nn; a++)
            {
                const int* d = cell_dofs + ((size_t)c * nn + a) * ncomp;
                Require(d[0] >= 0 && d[0] + ncomp <= n_dofs, NUTO_ERR_ARGUMENT, "dof number out of range");
                for (int k = 1; k < ncomp; k++)
                    Require(d[k] == d[0] + k, NUTO_ERR_UNSUPPORTED,
                            "dof numbering is not node-blocked (components of a node must be consecutive)");
                Require(d[0] % ncomp == 0, NUTO_ERR_UNSUPPORTED, "dof numbering is not node-blocked");
                fd.ranks[(size_t)c * nn + a] = d[0] / ncomp;
                maxRank = std::max(maxRank, d[0] / ncomp);
            }
        fd.nNodes = std::max(maxRank + 1, n_dofs / ncomp);
        Require(fd.nNodes * ncomp == n_dofs || n_dofs == 0, NUTO_ERR_UNSUPPORTED,
                "dof count is not a multiple of the component count");
        InvalidatePattern(ctx);
        ctx->f[slot] = std::move(fd);
        ctx->tabValid = false;
        if (slot == 0)
            Upload(ctx, ctx->dconn, ctx->f[0].ranks.data(), ctx->f[0].ranks.size(), "dof connectivity");
    });
}

int nuto_set_integrand(nuto_ctx* ctx, const nuto_params* params)
{
    return Guard(ctx, [&] {
        Require(params != nullptr, NUTO_ERR_ARGUMENT, "params is NULL");
        Require(params->kind >= 0 && params->kind <= 2, NUTO_ERR_ARGUMENT, "unknown integrand kind");
        Require(params->kind == kLinearElastic || params->dim == 3, NUTO_ERR_UNSUPPORTED,
                "damage integrands are provided in 3D");
        Require(params->damage_application == 0 || params->damage_application == 1, NUTO_ERR_ARGUMENT,
                "damage application must be FULL (0) or UNILATERAL (1)");
        ctx->params = *params;
        ctx->integrandSet = true;
        ctx->gdm.Release();
        ctx->gdmTried = false;
    });
}

int nuto_set_rule(nuto_ctx* ctx, int kind, int n)
{
    return Guard(ctx, [&] {
        Require(kind == 0 || kind == 1, NUTO_ERR_ARGUMENT, "rule kind must be 0 (tensor Gauss) or 1 (tetrahedron)");
        Rule r;
        Require(MakeRule(kind, 3, n, r), NUTO_ERR_UNSUPPORTED, "unsupported integration rule order");
        ctx->ruleKind = kind;
        ctx->ruleN = n;
        ctx->ruleSet = true;
        ctx->tabValid = false;
        ctx->gdm.Release();
        ctx->gdmTried = false;
    });
}

int nuto_set_history_ld(nuto_ctx* ctx, int ld)
{
    return Guard(ctx, [&] {
        Require(ld >= 0, NUTO_ERR_ARGUMENT, "negative leading dimension");
        ctx->historyLd = ld;
    });
}

int nuto_set_material_fields(nuto_ctx* ctx, const double* E, const double* ft, const double* gf)
{
    return Guard(ctx, [&] {
        Require(ctx->cellsSet, NUTO_ERR_STATE, "set the cells first");
        if (!E && !ft && !gf)
        {
            ctx->hasTheta = false;
            return;
        }
        Require(E && ft && gf, NUTO_ERR_ARGUMENT, "E, ft and gf must all be given");
        Upload(ctx, ctx->thE, E, ctx->nCells, "material fields");
        Upload(ctx, ctx->thFt, ft, ctx->nCells, "material fields");
        Upload(ctx, ctx->thGf, gf, ctx->nCells, "material fields");
        ctx->hasTheta = true;
    });
}

int nuto_set_material_fields_device(nuto_ctx* ctx, const double* E, const double* ft, const double* gf)
{
    return Guard(ctx, [&] {
        Require(ctx->cellsSet, NUTO_ERR_STATE, "set the cells first");
        if (!E && !ft && !gf)
        {
            ctx->hasTheta = false;
            return;
        }
        Require(E && ft && gf, NUTO_ERR_ARGUMENT, "E, ft and gf must all be given");
        const size_t bytes = (size_t)ctx->nCells * sizeof(double);
        for (DevBuf* b : {&ctx->thE, &ctx->thFt, &ctx->thGf})
        {
            if (b->bytes != bytes || !b->p)
                Check(b->Alloc(bytes), "material fields");
        }
        if (bytes)
        {
            Check(hipMemcpyAsync(ctx->thE.p, E, bytes, hipMemcpyDeviceToDevice, ctx->stream), "material fields");
            Check(hipMemcpyAsync(ctx->thFt.p, ft, bytes, hipMemcpyDeviceToDevice, ctx->stream), "material fields");
            Check(hipMemcpyAsync(ctx->thGf.p, gf, bytes, hipMemcpyDeviceToDevice, ctx->stream), "material fields");
        }
        ctx->hasTheta = true;
    });
}

int nuto_build_pattern(nuto_ctx* ctx, long long* nnz, int* n_rows)
{
    return Guard(ctx, [&] {
        Require(ctx->cellsSet && ctx->f[0].set, NUTO_ERR_STATE, "cells and dof type slot 0 must be set");
        const bool two = ctx->f[1].set;
        if (two)
        {
            Require(ctx->f[1].type == ctx->f[0].type && ctx->f[1].ranks == ctx->f[0].ranks, NUTO_ERR_UNSUPPORTED,
                    "both dof types must share one interpolation and node numbering");
        }
        std::string err;
        NodeGraph& g = ctx->graph;
        const int nnD = NodesOf(ctx->f[0].type);
        Require(BuildNodeGraph(ctx->nCells, nnD, ctx->f[0].ranks.data(), ctx->f[0].nNodes, g, err),
                NUTO_ERR_ARGUMENT, err);
        const int nc[2] = {ctx->f[0].ncomp, two ? ctx->f[1].ncomp : 0};
        std::vector<int> rp;
        long long nz = 0;
        Require(BuildRowPtr(g, two ? 2 : 1, nc, rp, nz, err), NUTO_ERR_UNSUPPORTED, err);
        Upload(ctx, ctx->incPtr, g.incPtr.data(), g.incPtr.size(), "pattern");
        Upload(ctx, ctx->incCell, g.incCell.data(), g.incCell.size(), "pattern");
        Upload(ctx, ctx->incLocal, g.incLocal.data(), g.incLocal.size(), "pattern");
        Upload(ctx, ctx->slot, g.slot.data(), g.slot.size(), "pattern");
        Upload(ctx, ctx->adjPtr, g.adjPtr.data(), g.adjPtr.size(), "pattern");
        Upload(ctx, ctx->adj, g.adj.data(), g.adj.size(), "pattern");
        Upload(ctx, ctx->rowPtr, rp.data(), rp.size(), "pattern");
        Check(ctx->col.Alloc((size_t)nz * sizeof(int)), "pattern columns");
        const int nU = ctx->f[0].nNodes * ctx->f[0].ncomp;
        if (g.nNodes > 0)
            hipLaunchKernelGGL(kFillCols, dim3(g.nNodes), dim3(256), 0, ctx->stream, ctx->adjPtr.As<int>(),
                               ctx->adj.As<int>(), ctx->rowPtr.As<int>(), g.nNodes, two ? 2 : 1, nc[0], nc[1], nU,
                               ctx->col.As<int>());
        Check(hipGetLastError(), "pattern column fill");
        Check(hipStreamSynchronize(ctx->stream), "pattern column fill");
        ctx->nnz = nz;
        ctx->nRows = (int)(rp.size() - 1);
        ctx->patternBuilt = true;
        if (nnz)
            *nnz = nz;
        if (n_rows)
            *n_rows = ctx->nRows;
    });
}

int nuto_copy_pattern(nuto_ctx* ctx, int* row_ptr, int* col)
{
    return Guard(ctx, [&] {
        Require(ctx->patternBuilt, NUTO_ERR_STATE, "pattern not built");
        if (row_ptr)
            Check(hipMemcpy(row_ptr, ctx->rowPtr.p, (size_t)(ctx->nRows + 1) * sizeof(int), hipMemcpyDeviceToHost),
                  "copy row_ptr");
        if (col)
            Check(hipMemcpy(col, ctx->col.p, (size_t)ctx->nnz * sizeof(int), hipMemcpyDeviceToHost), "copy col");
    });
}

int nuto_pattern_device(nuto_ctx* ctx, const int** row_ptr, const int** col)
{
    return Guard(ctx, [&] {
        Require(ctx->patternBuilt, NUTO_ERR_STATE, "pattern not built");
        if (row_ptr)
            *row_ptr = ctx->rowPtr.As<int>();
        if (col)
            *col = ctx->col.As<int>();
    });
}

int nuto_build_vector(nuto_ctx* ctx, const double* u, const double* e, const double* kappa, double* R)
{
    return Guard(ctx, [&] {
        Assemble(ctx, u, e, kappa, R, nullptr, true, false);
    });
}

int nuto_build_matrix(nuto_ctx* ctx, const double* u, const double* e, const double* kappa, double* K)
{
    return Guard(ctx, [&] {
        Assemble(ctx, u, e, kappa, nullptr, K, false, true);
    });
}

int nuto_build_vector_matrix(nuto_ctx* ctx, const double* u, const double* e, const double* kappa, double* R,
                             double* K)
{
    return Guard(ctx, [&] {
        Assemble(ctx, u, e, kappa, R, K, true, true);
    });
}

int nuto_update_history(nuto_ctx* ctx, const double* u, const double* e, double* kappa)
{
    return Guard(ctx, [&] {
        Require(ctx->params.kind != kLinearElastic, NUTO_ERR_STATE, "linear elasticity has no history");
        CheckReady(ctx, u, e, kappa);
        if (!ctx->tabValid)
            BuildTables(ctx);
        RowArgs a = MakeRowArgs(ctx, u, e, kappa, nullptr, nullptr);
        Dispatch(ctx, a, false, false, 1, kappa);
    });
}

int nuto_vjp(nuto_ctx* ctx, const double* u, const double* e, const double* kappa, const double* lambda,
             double* d_state, double* d_theta)
{
    return Guard(ctx, [&] {
        CheckReady(ctx, u, e, kappa);
        Require(ctx->params.kind == kGradientDamage, NUTO_ERR_UNSUPPORTED,
                "the adjoint is provided for the gradient-damage integrand");
        Require(lambda != nullptr && d_state != nullptr, NUTO_ERR_ARGUMENT, "lambda / d_state is NULL");
        if (!ctx->tabValid)
            BuildTables(ctx);
        PrepareGdm(ctx);
        Require(ctx->gdm.Usable(ctx->params.kind, ctx->coordType, ctx->f[0].type, ctx->nip), NUTO_ERR_UNSUPPORTED,
                "the adjoint needs hex8 cells and dofs, 2x2x2 Gauss points and at most 8 cells and 32 "
                "neighbours per node");
        RowArgs a = MakeRowArgs(ctx, u, e, kappa, nullptr, nullptr);
        LaunchGdmHex8Vjp(ctx->gdm, a, lambda, d_state, d_theta, ctx->stream);
        Check(hipGetLastError(), "vjp launch");
        SyncCheck(ctx, "vjp kernel");
    });
}

int nuto_device_alloc(nuto_ctx* ctx, void** ptr, unsigned long long bytes)
{
    return Guard(ctx, [&] {
        Require(ptr != nullptr, NUTO_ERR_ARGUMENT, "ptr is NULL");
        Check(hipMalloc(ptr, bytes ? bytes : 16), "device alloc");
    });
}

int nuto_device_free(nuto_ctx* ctx, void* ptr)
{
    return Guard(ctx, [&] {
        if (ptr)
            Check(hipFree(ptr), "device free");
    });
}

int nuto_copy_to_device(nuto_ctx* ctx, void* dst, const void* src, unsigned long long bytes)
{
    return Guard(ctx, [&] { Check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "copy to device"); });
}

int nuto_copy_to_host(nuto_ctx* ctx, void* dst, const void* src, unsigned long long bytes)
{
    return Guard(ctx, [&] { Check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "copy to host"); });
}

}  // extern "C"
