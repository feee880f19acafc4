// C ABI of the MI355X assembly path (declared in include/nuto_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nuto_amd.h"
#include "assemble_rows.hpp"
#include "elements.hpp"
#include "gdm_hex8.hpp"
#include "gdm_hex8_vjp.hpp"
#include "batch.hpp"
#include "cell_fields.hpp"
#include "cell_gram.hpp"
#include "cell_mfma.hpp"
#include "cell_matrices.hpp"
#include "solve.hpp"
#include "mesh.hpp"

using namespace nuto_amd;

// Device scratch of nuto_build_vector_matrix_noalloc: the hex8 node kernel's pre-pass state records
// (gdm8::kSr doubles per cell and IP). Each device holds its own copy from the code object's load to
// the end of the process; host code reaches it only through the address that hipGetSymbolAddress
// returns for the current device (ScratchOf).
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
constexpr int kScratchMaxCells = (int)(kScratchDoubles / (8 * gdm8::kSr));  // 29127 cells of 8 IPs
static __device__ __align__(256) double scratch[kScratchDoubles];

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
    // the constrained solves' vectors and the pinned word their reductions are read through, kept
    // between solves (a Newton step solves once per iteration); grown when a solve needs more
    DevBuf solveWork;
    double* solveHost = nullptr;
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
    // constraint equations per dof type slot (Constraints::Add) and the unit constraint matrix
    // built from them for the monolithic system (C and C^T, CSR, device)
    struct Equations
    {
        std::vector<int> dep, termPtr{0}, termDof;
        std::vector<double> coef;
    } cons[2];
    bool consBuilt = false;
    int nInd = 0, nEq = 0;
    DevBuf cRow, cCol, cVal, ctRow, ctCol, ctVal, consDep;
    // per-block views of the monolithic matrix (nuto_block_pattern / nuto_block_values), built on
    // first use per (slot i, slot j): the block's CSR row pointers and, per row, where the block's
    // run starts in the monolithic row
    struct BlockLayout
    {
        bool valid = false;
        int rows = 0, row0 = 0;
        long long nnz = 0;
        DevBuf rowPtr, runBegin;
    } blocks[2][2];
    // staged cell matrices and vectors of the two-phase generic assembly
    DevBuf keStage, reStage;
    int maxAdj = -1;  // largest node neighbourhood (LDS rows of the wave gather), from the node graph
    int cellGram = -1;  // cell-matrix kernel of the two-phase assembly: 0 by columns, 1 / 2 Gram form with 1 / 2
                        // row nodes per lane, 3 Gram form on the matrix cores, -1 the element's default
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

// Device address of `scratch` per device, resolved once per device at context creation with that
// device current (ResolveScratch); nullptr until then.
std::mutex gScratchMu;
std::vector<double*> gScratch;

hipError_t ResolveScratch(int device)
{
    std::lock_guard<std::mutex> lock(gScratchMu);
    if (device < (int)gScratch.size() && gScratch[device])
        return hipSuccess;
    double* p = nullptr;
    const hipError_t err = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
    if (err != hipSuccess)
        return err;
    if (device >= (int)gScratch.size())
        gScratch.resize(device + 1, nullptr);
    gScratch[device] = p;
    return hipSuccess;
}

double* ScratchOf(int device)
{
    std::lock_guard<std::mutex> lock(gScratchMu);
    return device < (int)gScratch.size() ? gScratch[device] : nullptr;
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
    for (auto& row : ctx->blocks)
        for (auto& b : row)
        {
            b.valid = false;
            b.rowPtr.Release();
            b.runBegin.Release();
        }
    ctx->gdm.Release();
    ctx->gdmTried = false;
    ctx->consBuilt = false;
    ctx->maxAdj = -1;
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
                                p.plane_state, 0.0, 0.0, 0.0, 0.0, p.interaction, p.interaction_r,
                                p.interaction_eta});
    a.R = R;
    a.K = K;
    return a;
}

// NUTO_AMD_TWO_PHASE=0 keeps the row-owner kernel for every generic 3D matrix assembly
bool TwoPhaseOff()
{
    const char* v = std::getenv("NUTO_AMD_TWO_PHASE");
    return v && v[0] == '0';
}

// staging for the cell matrices of cells [0, nCells) (grown on demand, kept with the context)
// staging of `cells` cell matrices (nd x nd) and vectors; it only grows
void EnsureStage(nuto_ctx* ctx, int nd, long long cells)
{
    const size_t ke = (size_t)cells * nd * nd * sizeof(double), re = (size_t)cells * nd * sizeof(double);
    if (ctx->keStage.bytes < ke)
        Check(ctx->keStage.Alloc(ke), "cell matrix staging");
    if (ctx->reStage.bytes < re)
        Check(ctx->reStage.Alloc(re), "cell vector staging");
}

// the staging is there or could be allocated (a failed allocation clears the error state)
bool TryStage(nuto_ctx* ctx, int nd, long long cells)
{
    try
    {
        EnsureStage(ctx, nd, cells);
        return true;
    }
    catch (const Failure&)
    {
        (void)hipGetLastError();
        return false;
    }
}

// Two-phase assembly of the cells [c0, c1) and the rows of the dof nodes [p0, p1) (the rows'
// incident cells must lie in [c0, c1): a whole mesh of a batch, or the whole group).
// NUTO_AMD_CELL_GRAM=0 / 1 / 2 / 3 forces the cell-matrix form of every two-phase launch: by columns
// (kCellMatrices), Gram form with one or two row nodes per lane (kCellGram), Gram form on the
// matrix cores (kCellMfma); -1 when unset
int CellFormEnv()
{
    const char* v = std::getenv("NUTO_AMD_CELL_GRAM");
    return v && v[0] >= '0' && v[0] <= '3' ? v[0] - '0' : -1;
}

// kCellGram's shape per element and row nodes per lane (AP): cells per workgroup, threads, IPs per
// chunk, unroll of the IP loop of the blocks (3 only for hex20 with AP = 1, whose long chunk hides
// the LDS latency; the others lose occupancy)
template <int NND, int AP>
struct GramShape;
template <>
struct GramShape<4, 1>  // tet4: 8 lanes per cell
{
    static constexpr int CPW = 8, NT = 64, ICH = 1, UNR = 1;
};
template <>
struct GramShape<10, 1>  // tet10: 50 lanes per cell
{
    static constexpr int CPW = 2, NT = 128, ICH = 4, UNR = 1;
};
template <>
struct GramShape<8, 1>  // hex8: 32 lanes per cell
{
    static constexpr int CPW = 2, NT = 64, ICH = 4, UNR = 1;
};
template <>
struct GramShape<20, 1>  // hex20: 200 lanes per cell
{
    static constexpr int CPW = 1, NT = 256, ICH = 9, UNR = 3;
};
template <>
struct GramShape<4, 2>  // tet4: 4 lanes per cell
{
    static constexpr int CPW = 16, NT = 64, ICH = 1, UNR = 1;
};
template <>
struct GramShape<10, 2>  // tet10: 25 lanes per cell
{
    static constexpr int CPW = 2, NT = 64, ICH = 4, UNR = 1;
};
template <>
struct GramShape<8, 2>  // hex8: 16 lanes per cell
{
    static constexpr int CPW = 4, NT = 64, ICH = 4, UNR = 1;
};
template <>
struct GramShape<20, 2>  // hex20: 100 lanes per cell
{
    static constexpr int CPW = 2, NT = 256, ICH = 9, UNR = 1;
};

template <int KIND, int NNC, int NND, int NIP, int AP, bool TWO>
void LaunchGram(const RowArgs& a, int c0, int c1, double* ke, double* re, hipStream_t stream)
{
    using Sh = GramShape<NND, AP>;
    const dim3 grid((unsigned)((c1 - c0 + Sh::CPW - 1) / Sh::CPW));
    hipLaunchKernelGGL((kCellGram<KIND, NNC, NND, NIP, Sh::CPW, Sh::NT, Sh::ICH, Sh::UNR, AP, TWO>), grid,
                       dim3(Sh::NT), 0, stream, a, c0, c1, ke, re);
}

// form: 0 by columns (kCellMatrices), 1 Gram form with one row node per lane, 2 with two (four
// accumulated blocks per lane), 3 Gram form on the matrix cores (kCellMfma). UNILATERAL damage
// application, which needs a second Gram, runs forms 2 and 3 as 1.
constexpr int kCellForms = 4;
template <int KIND, int NNC, int NND, int NIP>
void LaunchCellMatrices(const RowArgs& a, int form, int c0, int c1, double* ke, double* re, hipStream_t stream)
{
    if (form == 0)
    {
        constexpr int CPW = 64 / (NND * 3);
        hipLaunchKernelGGL((kCellMatrices<KIND, NNC, NND, NIP>), dim3((unsigned)((c1 - c0 + CPW - 1) / CPW)), dim3(64),
                           0, stream, a, c0, c1, ke, re);
        return;
    }
    // one Gram unless UNILATERAL damage application makes c1 / c2 vary between IPs
    if (KIND == kLocalDamage && a.P.damage_application != 0)
        LaunchGram<KIND, NNC, NND, NIP, 1, true>(a, c0, c1, ke, re, stream);
    else if (form == 3)
        hipLaunchKernelGGL((kCellMfma<KIND, NNC, NND, NIP>), dim3((unsigned)(c1 - c0)), dim3(MfmaShape<NND>::NT), 0,
                           stream, a, c0, c1, ke, re);
    else if (form == 2)
        LaunchGram<KIND, NNC, NND, NIP, 2, false>(a, c0, c1, ke, re, stream);
    else
        LaunchGram<KIND, NNC, NND, NIP, 1, false>(a, c0, c1, ke, re, stream);
}

template <int KIND, int NNC, int NND, int NIP>
void LaunchTwoPhaseT(nuto_ctx* ctx, const RowArgs& a, bool wr, int c0, int c1, int p0, int p1, hipStream_t stream)
{
    constexpr int ND = NND * 3;
    if (c1 <= c0 || p1 <= p0)
        return;
    EnsureStage(ctx, ND, c1 - c0);  // cell c's matrix at staging position c - c0
    double* ke = ctx->keStage.As<double>();
    double* re = ctx->reStage.As<double>();
    // the Gram form wins on the hexahedra (32 and 200 lanes per cell); the column kernel is as fast or
    // faster on the tetrahedra, whose cell matrices are small (a batch's autotune measures every form)
    const int env = CellFormEnv();
    const int form = env >= 0 ? env : (ctx->cellGram >= 0 ? ctx->cellGram : (NND >= 8 ? 1 : 0));
    LaunchCellMatrices<KIND, NNC, NND, NIP>(a, form, c0, c1, ke, re, stream);
    if (ctx->maxAdj < 0)
    {
        int m = 0;
        for (int p = 0; p < ctx->graph.nNodes; p++)
            m = std::max(m, ctx->graph.adjPtr[p + 1] - ctx->graph.adjPtr[p]);
        ctx->maxAdj = m;
    }
    const size_t lds = (size_t)3 * 3 * ctx->maxAdj * sizeof(double);
    const char* lane = std::getenv("NUTO_AMD_GATHER_LANE");  // 1: the lane-per-node gather (tests)
    if (lds <= 48 * 1024 && !(lane && lane[0] == '1'))
    {
        // one wave per node, rows summed in LDS
        if (wr)
            hipLaunchKernelGGL((kGatherRowsWave<NND, true>), dim3((unsigned)(p1 - p0)), dim3(64), lds, stream, a, p0, p1,
                               c0, ke, re);
        else
            hipLaunchKernelGGL((kGatherRowsWave<NND, false>), dim3((unsigned)(p1 - p0)), dim3(64), lds, stream, a, p0,
                               p1, c0, ke, re);
    }
    else
    {
        const dim3 grid((unsigned)((p1 - p0 + 63) / 64));
        if (wr)
            hipLaunchKernelGGL((kGatherRows<NND, true>), grid, dim3(64), 0, stream, a, p0, p1, c0, ke, re);
        else
            hipLaunchKernelGGL((kGatherRows<NND, false>), grid, dim3(64), 0, stream, a, p0, p1, c0, ke, re);
    }
    Check(hipGetLastError(), "two-phase assembly launch");
}

template <int KIND, int DIM, int NNC, int NND, int NIP, int BH>
void LaunchRowsT(nuto_ctx* ctx, const RowArgs& a, bool wr, bool wk)
{
    const int n = a.nNodes;
    if (n == 0)
        return;
    // (the staging kernels keep every IP's state in LDS: rules up to 27 points)
    if constexpr (DIM == 3 && KIND != kGradientDamage && NND * 3 <= 64 && NIP <= 27)
    {
        // the single-pass kernel serves when the staging does not fit in device memory
        if (wk && !TwoPhaseOff() && TryStage(ctx, NND * 3, a.nCells))
        {
            LaunchTwoPhaseT<KIND, NNC, NND, NIP>(ctx, a, wr, 0, a.nCells, 0, n, ctx->stream);
            SyncCheck(ctx, "assembly kernel");
            return;
        }
    }
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

template <int DIM, int NNC, int NND, int NIP>
void LaunchLumpedT(nuto_ctx* ctx, const RowArgs& a, double rho, double* M)
{
    if (a.nNodes == 0)
        return;
    hipLaunchKernelGGL((kLumpedMass<DIM, NNC, NND, NIP>), dim3((a.nNodes + 63) / 64), dim3(64), 0, ctx->stream, a, rho,
                       M);
    Check(hipGetLastError(), "lumped mass launch");
    SyncCheck(ctx, "lumped mass kernel");
}

template <int KIND, int DIM, int NNC, int NND, int NIP>
void LaunchEvalT(nuto_ctx* ctx, const RowArgs& a, int field, double* out)
{
    const long long n = (long long)a.nCells * NIP;
    if (n == 0)
        return;
    hipLaunchKernelGGL((kEvalIp<KIND, DIM, NNC, NND, NIP>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       ctx->stream, a, field, out);
    Check(hipGetLastError(), "ip field launch");
    SyncCheck(ctx, "ip field kernel");
}

// Arguments of the per-cell outputs (modes 2 and 3 of Dispatch)
struct CellOutput
{
    double rho = 0.0;       // mode 2: density
    int field = 0;          // mode 3: IpField
    double* out = nullptr;  // mode 2: lumped mass per dof of slot 0; mode 3: field values
};

// Dispatch over the instantiated (integrand, coordinate element, dof element, rule) combinations.
// mode 0: assembly, 1: history update, 2: lumped mass, 3: IP field
void Dispatch(nuto_ctx* ctx, const RowArgs& a, bool wr, bool wk, int mode, double* kappaOut,
              const CellOutput* co = nullptr)
{
    const int kind = ctx->params.kind, ct = ctx->coordType, dt = ctx->f[0].type, nip = ctx->nip;
#define NUTO_CASE(KIND, DIM, CT, NNC, DT, NND, NIP, BH)                       \
    if (kind == KIND && ct == CT && dt == DT && nip == NIP && ctx->dim == DIM) \
    {                                                                          \
        if (mode == 0)                                                         \
            LaunchRowsT<KIND, DIM, NNC, NND, NIP, BH>(ctx, a, wr, wk);         \
        else if (mode == 1)                                                    \
            LaunchHistoryT<KIND, DIM, NNC, NND, NIP>(ctx, a, kappaOut);        \
        else if (mode == 2)                                                    \
            LaunchLumpedT<DIM, NNC, NND, NIP>(ctx, a, co->rho, co->out);       \
        else                                                                   \
            LaunchEvalT<KIND, DIM, NNC, NND, NIP>(ctx, a, co->field, co->out); \
        return;                                                                \
    }
    NUTO_CASE(kLinearElastic, 2, kQuad4, 4, kQuad4, 4, 4, 4)
    NUTO_CASE(kLinearElastic, 2, kQuad4, 4, kQuad4, 4, 1, 4)
    NUTO_CASE(kLinearElastic, 2, kQuad4, 4, kQuad4, 4, 9, 4)
    NUTO_CASE(kLinearElastic, 2, kQuad4, 4, kQuad4, 4, 16, 4)
    NUTO_CASE(kLinearElastic, 2, kQuad4, 4, kQuad4, 4, 25, 4)
    NUTO_CASE(kLinearElastic, 3, kHex8, 8, kHex8, 8, 8, 8)
    NUTO_CASE(kLinearElastic, 3, kHex8, 8, kHex8, 8, 27, 8)
    NUTO_CASE(kLinearElastic, 3, kHex8, 8, kHex8, 8, 64, 8)
    NUTO_CASE(kLinearElastic, 3, kTet4, 4, kTet4, 4, 1, 4)
    NUTO_CASE(kLinearElastic, 3, kTet4, 4, kTet10, 10, 4, 5)
    NUTO_CASE(kLinearElastic, 3, kTet10, 10, kTet10, 10, 4, 5)
    NUTO_CASE(kLinearElastic, 3, kHex8, 8, kHex20, 20, 27, 5)
    NUTO_CASE(kGradientDamage, 3, kHex8, 8, kHex8, 8, 8, 4)
    NUTO_CASE(kGradientDamage, 3, kHex8, 8, kHex8, 8, 27, 4)
    NUTO_CASE(kGradientDamage, 3, kHex8, 8, kHex8, 8, 64, 4)
    NUTO_CASE(kLocalDamage, 3, kHex8, 8, kHex8, 8, 8, 8)
    NUTO_CASE(kLocalDamage, 3, kHex8, 8, kHex8, 8, 27, 8)
    NUTO_CASE(kLocalDamage, 3, kHex8, 8, kHex8, 8, 64, 8)
    NUTO_CASE(kLocalDamage, 3, kTet4, 4, kTet4, 4, 1, 4)
    NUTO_CASE(kLocalDamage, 3, kTet4, 4, kTet10, 10, 4, 5)
    NUTO_CASE(kLocalDamage, 3, kTet10, 10, kTet10, 10, 4, 5)
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
                     ctx->hasCellIds ? ctx->hcellId.data() : nullptr, ctx->nCoordNodes, ctx->x.As<double>(),
                     ctx->y.As<double>(), ctx->z.As<double>(), ctx->cconn.As<int>(), ctx->tab.As<double>(),
                     ctx->rowPtr.As<int>(), ctx->nRows, ctx->stream, err);
    Check(err, "hex8 node tables");
    ctx->gdmTried = true;
}

// Block (slot i, slot j) of the monolithic matrix, row by row: the monolithic row of dof row0 + r
// holds slot 0's columns [0, colSplit) before slot 1's (columns ascending), so the block's entries
// are one run. count[r] = its length, runBegin[r] = its first monolithic entry.
__global__ void __launch_bounds__(256) kBlockRuns(const int* __restrict__ rowPtr, const int* __restrict__ col,
                                                  int row0, int rows, int colSplit, int slotJ,
                                                  int* __restrict__ count, long long* __restrict__ runBegin)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= rows)
        return;
    const int b = rowPtr[row0 + r], e = rowPtr[row0 + r + 1];
    int lo = b, hi = e;  // first entry with col >= colSplit
    while (lo < hi)
    {
        const int m = (lo + hi) >> 1;
        if (col[m] < colSplit)
            lo = m + 1;
        else
            hi = m;
    }
    count[r] = slotJ == 0 ? lo - b : e - lo;
    runBegin[r] = slotJ == 0 ? b : lo;
}

// block columns (colShift subtracted) or values, one wave per block row
template <class T>
__global__ void __launch_bounds__(256) kBlockGather(const int* __restrict__ blockRowPtr,
                                                    const long long* __restrict__ runBegin, int rows,
                                                    const T* __restrict__ src, T shift, T* __restrict__ dst)
{
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows)
        return;
    const int lane = threadIdx.x & 63;
    const int b = blockRowPtr[r], n = blockRowPtr[r + 1] - b;
    const T* s = src + runBegin[r];
    for (int k = lane; k < n; k += 64)
        dst[b + k] = s[k] - shift;
}

nuto_ctx::BlockLayout& BlockOf(nuto_ctx* ctx, int si, int sj)
{
    Require(ctx->patternBuilt, NUTO_ERR_STATE, "pattern not built");
    Require(si >= 0 && si < 2 && sj >= 0 && sj < 2, NUTO_ERR_ARGUMENT, "dof type slots are 0 and 1");
    Require(ctx->f[si].set && ctx->f[sj].set, NUTO_ERR_STATE, "no dof numbering for that slot");
    nuto_ctx::BlockLayout& L = ctx->blocks[si][sj];
    if (L.valid)
        return L;
    const int n0 = ctx->f[0].nDofs;
    L.row0 = si == 0 ? 0 : n0;
    L.rows = ctx->f[si].nDofs;
    DevBuf cnt;
    Check(cnt.Alloc((size_t)L.rows * sizeof(int)), "block counts");
    Check(L.runBegin.Alloc((size_t)L.rows * sizeof(long long)), "block runs");
    if (L.rows > 0)
    {
        hipLaunchKernelGGL(kBlockRuns, dim3((L.rows + 255) / 256), dim3(256), 0, ctx->stream, ctx->rowPtr.As<int>(),
                           ctx->col.As<int>(), L.row0, L.rows, n0, sj, cnt.As<int>(), L.runBegin.As<long long>());
        Check(hipGetLastError(), "block runs");
    }
    std::vector<int> c(L.rows), rp(L.rows + 1, 0);
    if (L.rows > 0)
        Check(hipMemcpyAsync(c.data(), cnt.p, c.size() * sizeof(int), hipMemcpyDeviceToHost, ctx->stream), "block counts");
    Check(hipStreamSynchronize(ctx->stream), "block counts");
    for (int r = 0; r < L.rows; r++)
        rp[r + 1] = rp[r] + c[r];  // a block has at most the monolithic nonzeros (int32 CSR)
    L.nnz = rp[L.rows];
    Check(L.rowPtr.Upload(rp.data(), rp.size() * sizeof(int), ctx->stream), "block row pointers");
    L.valid = true;
    return L;
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
    // the node kernel takes the Constant nonlocal interaction only: no tables for other cases
    const bool nodeKernel = !ForceGeneric() && ctx->params.interaction == 0;
    if (nodeKernel)
        PrepareGdm(ctx);
    if (nodeKernel && ctx->gdm.Usable(ctx->params.kind, ctx->coordType, ctx->f[0].type, ctx->nip))
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
    if (hipSetDevice(device) != hipSuccess || ResolveScratch(device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking) != hipSuccess)
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
    if (ctx->solveHost)
        (void)hipHostFree(ctx->solveHost);
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
        Require(coord_type == kQuad4 || coord_type == kHex8 || coord_type == kTet4 || coord_type == kTet10,
                NUTO_ERR_ARGUMENT, "coordinate elements must be quad4, hex8, tet4 or tet10");
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
        ctx->cons[slot] = nuto_ctx::Equations();  // the equations refer to the previous numbering
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
        Require(params->interaction == 0 || (params->interaction == 1 && params->kind == kGradientDamage &&
                                             params->interaction_eta != 0.0),
                NUTO_ERR_ARGUMENT, "interaction must be Constant (0) or, for gradient damage, Decreasing (1) with eta != 0");
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
        // gradient damage with the integrand and the rule already set: the IP tables and the node
        // kernel's tables now, so that assemblies (nuto_build_vector_matrix_noalloc) allocate nothing
        if (ctx->integrandSet && ctx->ruleSet && ctx->params.kind == kGradientDamage &&
            ctx->params.interaction == 0 && !ForceGeneric())
        {
            if (!ctx->tabValid)
                BuildTables(ctx);
            PrepareGdm(ctx);
        }
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

int nuto_block_pattern(nuto_ctx* ctx, int slot_i, int slot_j, long long* nnz, int* row_ptr, int* col,
                       long long* run_begin)
{
    return Guard(ctx, [&] {
        nuto_ctx::BlockLayout& L = BlockOf(ctx, slot_i, slot_j);
        if (nnz)
            *nnz = L.nnz;
        if (row_ptr)
            Check(hipMemcpy(row_ptr, L.rowPtr.p, (size_t)(L.rows + 1) * sizeof(int), hipMemcpyDeviceToHost),
                  "copy block row_ptr");
        if (run_begin && L.rows > 0)
            Check(hipMemcpy(run_begin, L.runBegin.p, (size_t)L.rows * sizeof(long long), hipMemcpyDeviceToHost),
                  "copy block runs");
        if (col && L.nnz > 0)
        {
            DevBuf c;
            Check(c.Alloc((size_t)L.nnz * sizeof(int)), "block columns");
            const int shift = slot_j == 0 ? 0 : ctx->f[0].nDofs;
            hipLaunchKernelGGL(kBlockGather<int>, dim3((L.rows + 3) / 4), dim3(256), 0, ctx->stream,
                               L.rowPtr.As<int>(), L.runBegin.As<long long>(), L.rows, ctx->col.As<int>(), shift,
                               c.As<int>());
            Check(hipGetLastError(), "block columns");
            Check(hipMemcpyAsync(col, c.p, (size_t)L.nnz * sizeof(int), hipMemcpyDeviceToHost, ctx->stream),
                  "copy block columns");
            Check(hipStreamSynchronize(ctx->stream), "copy block columns");
        }
    });
}

int nuto_block_values(nuto_ctx* ctx, int slot_i, int slot_j, const double* K, double* values)
{
    return Guard(ctx, [&] {
        Require(K != nullptr && values != nullptr, NUTO_ERR_ARGUMENT, "K or values is NULL");
        nuto_ctx::BlockLayout& L = BlockOf(ctx, slot_i, slot_j);
        if (L.rows > 0 && L.nnz > 0)
        {
            hipLaunchKernelGGL(kBlockGather<double>, dim3((L.rows + 3) / 4), dim3(256), 0, ctx->stream,
                               L.rowPtr.As<int>(), L.runBegin.As<long long>(), L.rows, K, 0.0, values);
            Check(hipGetLastError(), "block values");
        }
        SyncCheck(ctx, "block values");
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

int nuto_build_vector_matrix_noalloc(nuto_ctx* ctx, const double* u, const double* e, const double* kappa,
                                     double* R, double* K)
{
    return Guard(ctx, [&] {
        CheckReady(ctx, u, e, kappa);
        Require(R != nullptr && K != nullptr, NUTO_ERR_ARGUMENT, "R / K is NULL");
        Require(ctx->params.kind == kGradientDamage && ctx->params.interaction == 0 && !ForceGeneric(),
                NUTO_ERR_UNSUPPORTED, "the allocation-free assembly is provided for the hex8 gradient-damage node "
                                      "kernel (constant nonlocal interaction)");
        Require(ctx->tabValid && ctx->gdmTried, NUTO_ERR_STATE,
                "node tables not prepared: set the integrand and the rule before nuto_build_pattern, or call "
                "nuto_build_vector_matrix once");
        Require(ctx->gdm.Usable(ctx->params.kind, ctx->coordType, ctx->f[0].type, ctx->nip), NUTO_ERR_UNSUPPORTED,
                "the allocation-free assembly needs hex8 cells and dofs and 2x2x2 Gauss points");
        Require(ctx->gdm.StateDoubles() <= kScratchDoubles, NUTO_ERR_ARGUMENT,
                "the pre-pass records of " + std::to_string(ctx->nCells) + " cells exceed the device scratch of " +
                    std::to_string(kScratchBytes >> 20) + " MiB (at most " + std::to_string(kScratchMaxCells) +
                    " cells)");
        double* state = ScratchOf(ctx->device);
        Require(state != nullptr, NUTO_ERR_DEVICE, "device scratch address not resolved");
        RowArgs a = MakeRowArgs(ctx, u, e, kappa, R, K);
        ctx->gdm.Launch(a, true, true, ctx->stream, state);
        Check(hipGetLastError(), "gdm hex8 launch");
        SyncCheck(ctx, "gdm hex8 kernel");
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

int nuto_build_lumped_mass(nuto_ctx* ctx, double rho, double* M)
{
    return Guard(ctx, [&] {
        Require(ctx->cellsSet && ctx->f[0].set, NUTO_ERR_STATE, "cells and dof type slot 0 must be set");
        Require(ctx->integrandSet && ctx->ruleSet, NUTO_ERR_STATE, "integrand and rule must be set");
        Require(ctx->patternBuilt, NUTO_ERR_STATE, "pattern not built (call nuto_build_pattern)");
        Require(ctx->f[0].ncomp == ctx->dim, NUTO_ERR_UNSUPPORTED,
                "the lumped mass is defined for the displacement dof type (dim components) in slot 0");
        Require(M != nullptr, NUTO_ERR_ARGUMENT, "M is NULL");
        if (!ctx->tabValid)
            BuildTables(ctx);
        RowArgs a = MakeRowArgs(ctx, nullptr, nullptr, nullptr, nullptr, nullptr);
        CellOutput co;
        co.rho = rho;
        co.out = M;
        Dispatch(ctx, a, false, false, 2, nullptr, &co);
    });
}

int nuto_eval_ip(nuto_ctx* ctx, const double* u, const double* e, const double* kappa, int field, double* out)
{
    return Guard(ctx, [&] {
        Require(field >= NUTO_IP_STRAIN && field <= NUTO_IP_EQ_STRAIN, NUTO_ERR_ARGUMENT, "unknown field");
        Require(out != nullptr, NUTO_ERR_ARGUMENT, "out is NULL");
        const bool damageField = field == NUTO_IP_KAPPA || field == NUTO_IP_DAMAGE || field == NUTO_IP_EQ_STRAIN;
        Require(!(damageField && ctx->params.kind == kLinearElastic), NUTO_ERR_UNSUPPORTED,
                "kappa, damage and the equivalent strain need a damage integrand");
        CheckReady(ctx, u, e, kappa);
        if (!ctx->tabValid)
            BuildTables(ctx);
        RowArgs a = MakeRowArgs(ctx, u, e, kappa, nullptr, nullptr);
        CellOutput co;
        co.field = field;
        co.out = out;
        Dispatch(ctx, a, false, false, 3, nullptr, &co);
    });
}

int nuto_set_constraints(nuto_ctx* ctx, int slot, int n_eq, const int* dep, const int* term_ptr, const int* term_dof,
                         const double* coef)
{
    return Guard(ctx, [&] {
        Require(slot == 0 || slot == 1, NUTO_ERR_ARGUMENT, "slot must be 0 or 1");
        Require(ctx->f[slot].set, NUTO_ERR_STATE, "set the dof type of this slot first");
        Require(n_eq >= 0 && (n_eq == 0 || (dep && term_ptr)), NUTO_ERR_ARGUMENT, "bad constraint arrays");
        const int nd = ctx->f[slot].nDofs;
        nuto_ctx::Equations q;
        // Constraints::TermChecker (Constraints.h:112-129): a dependent dof appears in no other
        // equation, and no independent term is a dependent dof of another equation
        std::vector<char> isDep(nd, 0), isInd(nd, 0);
        for (int i = 0; i < n_eq; i++)
        {
            const int d = dep[i];
            // Constraints::GetJKNumbering (Constraints.cpp:62-71)
            Require(d >= 0, NUTO_ERR_ARGUMENT, "There is no dof numbering for a node in equation " + std::to_string(i));
            Require(d < nd, NUTO_ERR_ARGUMENT,
                    "The provided dof number of the dependent term exceeds the total number of dofs in equation " +
                        std::to_string(i));
            Require(!isDep[d] && !isInd[d], NUTO_ERR_ARGUMENT,
                    "equation " + std::to_string(i) + ": its dependent dof is already part of another equation");
            Require(term_ptr[i + 1] >= term_ptr[i] && term_ptr[0] == 0, NUTO_ERR_ARGUMENT, "bad term_ptr");
            for (int t = term_ptr[i]; t < term_ptr[i + 1]; t++)
            {
                const int g = term_dof[t];
                Require(g >= 0 && g < nd, NUTO_ERR_ARGUMENT,
                        "There is no dof numbering for a node in equation " + std::to_string(i));
                Require(!isDep[g] && g != d, NUTO_ERR_ARGUMENT,
                        "equation " + std::to_string(i) + ": a term is the dependent dof of another equation");
            }
            isDep[d] = 1;
            for (int t = term_ptr[i]; t < term_ptr[i + 1]; t++)
                isInd[term_dof[t]] = 1;
            q.dep.push_back(d);
            for (int t = term_ptr[i]; t < term_ptr[i + 1]; t++)
            {
                q.termDof.push_back(term_dof[t]);
                q.coef.push_back(coef[t]);
            }
            q.termPtr.push_back((int)q.termDof.size());
        }
        ctx->cons[slot] = std::move(q);
        ctx->consBuilt = false;
    });
}

}  // extern "C"

namespace
{

// The unit constraint matrix of the monolithic system: block diagonal over the slots, block s =
// Constraints::BuildUnitConstraintMatrix(dof_s, n_s) (Constraints.cpp:89-137): the JK numbering
// puts the independent dofs first (ascending), then the dependent ones (equation order); column i
// < nInd_s holds a 1 at independent dof jk(i), and the row of each dependent dof holds -coef at the
// independent column of each term with |coef| > 1e-18 (setFromTriplets: duplicates summed in order).
void BuildConstraints(nuto_ctx* ctx)
{
    if (ctx->consBuilt)
        return;
    const int nSlots = ctx->f[1].set ? 2 : 1;
    int n[2] = {ctx->f[0].nDofs, nSlots > 1 ? ctx->f[1].nDofs : 0};
    Require(n[0] + n[1] == ctx->nRows, NUTO_ERR_STATE, "dof counts do not match the pattern");
    std::vector<std::vector<std::pair<int, double>>> rows(ctx->nRows);
    std::vector<int> depAll;
    int rowOff = 0, colOff = 0;
    for (int s = 0; s < nSlots; s++)
    {
        const nuto_ctx::Equations& q = ctx->cons[s];
        std::vector<char> isDep(n[s], 0);
        for (int d : q.dep)
            isDep[d] = 1;
        std::vector<int> rev(n[s], -1);
        int nIndS = 0;
        for (int g = 0; g < n[s]; g++)
            if (!isDep[g])
            {
                rev[g] = nIndS;
                rows[rowOff + g].push_back({colOff + nIndS, 1.0});
                nIndS++;
            }
        for (size_t e = 0; e < q.dep.size(); e++)
        {
            depAll.push_back(rowOff + q.dep[e]);
            auto& row = rows[rowOff + q.dep[e]];
            for (int t = q.termPtr[e]; t < q.termPtr[e + 1]; t++)
            {
                if (!(std::fabs(q.coef[t]) > 1.e-18))
                    continue;
                const int c = colOff + rev[q.termDof[t]];
                bool merged = false;
                for (auto& kv : row)
                    if (kv.first == c)
                    {
                        kv.second += -q.coef[t];
                        merged = true;
                    }
                if (!merged)
                    row.push_back({c, -q.coef[t]});
            }
        }
        rowOff += n[s];
        colOff += nIndS;
    }
    const int nInd = colOff;
    std::vector<int> cr(ctx->nRows + 1, 0), cc;
    std::vector<double> cv;
    std::vector<std::vector<std::pair<int, double>>> cols(nInd);
    for (int r = 0; r < ctx->nRows; r++)
    {
        std::sort(rows[r].begin(), rows[r].end());
        for (auto& kv : rows[r])
        {
            cc.push_back(kv.first);
            cv.push_back(kv.second);
            cols[kv.first].push_back({r, kv.second});
        }
        cr[r + 1] = (int)cc.size();
    }
    std::vector<int> tr(nInd + 1, 0), tc;
    std::vector<double> tv;
    for (int j = 0; j < nInd; j++)
    {
        for (auto& kv : cols[j])
        {
            tc.push_back(kv.first);
            tv.push_back(kv.second);
        }
        tr[j + 1] = (int)tc.size();
    }
    Upload(ctx, ctx->cRow, cr.data(), cr.size(), "constraints");
    Upload(ctx, ctx->cCol, cc.data(), cc.size(), "constraints");
    Upload(ctx, ctx->cVal, cv.data(), cv.size(), "constraints");
    Upload(ctx, ctx->ctRow, tr.data(), tr.size(), "constraints");
    Upload(ctx, ctx->ctCol, tc.data(), tc.size(), "constraints");
    Upload(ctx, ctx->ctVal, tv.data(), tv.size(), "constraints");
    Upload(ctx, ctx->consDep, depAll.data(), depAll.size(), "constraints");
    ctx->nInd = nInd;
    ctx->nEq = (int)depAll.size();
    ctx->consBuilt = true;
}

// ConstrainedSystemSolver::Solve / SolveTrialState (Solve.cpp:8-78) with delta = NULL / given
// the node structure of the context's K (every row of a dof node holds the same columns)
NodeRows KNodeRows(const nuto_ctx* ctx)
{
    NodeRows r;
    r.nNodes = ctx->f[0].nNodes;
    r.c0 = ctx->f[0].ncomp;
    r.c1 = ctx->f[1].set ? ctx->f[1].ncomp : 0;
    r.nU = ctx->f[0].nNodes * ctx->f[0].ncomp;
    return r;
}

// The context's solve workspace, at least ReducedSolver::WorkDoubles(n_rows, n_independent) doubles
// (allocated once, or again when the constraints leave more independent dofs), and its pinned word.
// The size depends on the context only, so a solve inside a Newton step never reallocates it.
double* SolveWork(nuto_ctx* ctx)
{
    const size_t need = ReducedSolver::WorkDoubles(ctx->nRows, ctx->nInd) * sizeof(double);
    if (ctx->solveWork.bytes < need)
        Check(ctx->solveWork.Alloc(need), "solver buffers");
    if (!ctx->solveHost)
        Check(hipHostMalloc(&ctx->solveHost, 8 * sizeof(double), hipHostMallocDefault), "solver buffers");
    return ctx->solveWork.As<double>();
}

void SolveConstrained(nuto_ctx* ctx, const double* K, const double* f, const double* rhsOld, const double* rhsNew,
                      bool trial, double* x, const nuto_solver_options* opt, nuto_solve_info* info)
{
    Require(ctx->patternBuilt, NUTO_ERR_STATE, "pattern not built (call nuto_build_pattern)");
    Require(K && f && x, NUTO_ERR_ARGUMENT, "K, f and x must be device arrays");
    const int method = opt ? opt->method : NUTO_SOLVER_BICGSTAB;
    Require(method == NUTO_SOLVER_CG || method == NUTO_SOLVER_BICGSTAB, NUTO_ERR_ARGUMENT, "unknown solver");
    BuildConstraints(ctx);
    Require(!trial || ctx->nEq == 0 || rhsNew, NUTO_ERR_ARGUMENT, "rhs_new is NULL");
    const long long n = ctx->nRows;
    const int m = ctx->nInd;
    Csr Kc{ctx->nRows, ctx->rowPtr.As<int>(), ctx->col.As<int>(), K};
    Csr Cc{ctx->nRows, ctx->cRow.As<int>(), ctx->cCol.As<int>(), ctx->cVal.As<double>()};
    Csr Ct{m, ctx->ctRow.As<int>(), ctx->ctCol.As<int>(), ctx->ctVal.As<double>()};
    ReducedSolver rs(ctx->stream, Kc, Cc, Ct, n, m, KNodeRows(ctx), SolveWork(ctx), ctx->solveHost);
    double* delta = rs.Extra();
    double* g = delta + n;
    double* invd = g + n;
    double* b = invd + m;
    double* u = b + m;
    if (m > 0)
        hipLaunchKernelGGL(kDiagCtKC, dim3((m + 255) / 256), dim3(256), 0, ctx->stream, m, Ct.row, Ct.col, Ct.val,
                           Kc.row, Kc.col, Kc.val, invd);
    const double* rhs = f;
    if (trial)
    {
        // g = f + K delta, delta = rhs(new) - rhs(old) at the dependent dofs (GetSparseGlobalRhs)
        rs.Axpby(n, 0.0, nullptr, 0.0, delta);
        if (ctx->nEq > 0)
            hipLaunchKernelGGL(kScatterDelta, dim3((ctx->nEq + 255) / 256), dim3(256), 0, ctx->stream, ctx->nEq,
                               ctx->consDep.As<int>(), rhsOld, rhsNew, delta);
        rs.SpmvK(delta, g);
        rs.Axpby(n, 1.0, f, 1.0, g);
        rhs = g;
    }
    rs.Spmv(Ct, rhs, b);
    Check(hipGetLastError(), "solver setup");
    const double tol = opt && opt->tolerance > 0.0 ? opt->tolerance : 2.220446049250313e-16;
    const int maxIt = opt && opt->max_iterations > 0 ? opt->max_iterations : 2 * m;
    int iters = 0;
    double rel = 0.0;
    if (method == NUTO_SOLVER_CG)
        Check(rs.Cg(b, invd, u, tol, maxIt, iters, rel), "conjugate gradients");
    else
        Check(rs.Bicgstab(b, invd, u, tol, maxIt, iters, rel), "BiCGSTAB");
    rs.Spmv(Cc, u, x);  // x = C u
    if (trial)
        rs.Axpby(n, -1.0, delta, 1.0, x);  // the negative increment: C u - delta
    Check(hipGetLastError(), "solver result");
    Check(hipStreamSynchronize(ctx->stream), "solve");
    if (info)
    {
        info->iterations = iters;
        info->relative_residual = rel;
        info->independent_dofs = m;
    }
}

// QuasistaticSolver::DoStep (QuasistaticSolver.cpp:116-147) with NewtonRaphson::Solve
// (NewtonRaphson.h:62-94) and LineSearch / NoLineSearch (LineSearch.h:39-63,79-88). Every vector
// stays on the device; the host reads one norm per residual evaluation.
void QuasistaticStep(nuto_ctx* ctx, double* x, double* kappa, const double* rhsOld, const double* rhsNew,
                     const nuto_newton_options* o, nuto_newton_info* info)
{
    Require(ctx->patternBuilt, NUTO_ERR_STATE, "pattern not built (call nuto_build_pattern)");
    Require(x != nullptr, NUTO_ERR_ARGUMENT, "x is NULL");
    Require(ctx->params.kind == kLinearElastic || kappa != nullptr, NUTO_ERR_ARGUMENT,
            "this integrand keeps a history: kappa is NULL");
    const double tol = o && o->tolerance > 0.0 ? o->tolerance : 1e-10;
    const int maxIt = o && o->max_iterations > 0 ? o->max_iterations : 6;
    const int lsMax = !o || o->max_line_search_steps < 0 ? 6 : o->max_line_search_steps;
    const nuto_solver_options lin = o ? o->linear : nuto_solver_options{NUTO_SOLVER_BICGSTAB, 0.0, 0};
    BuildConstraints(ctx);
    const long long n = ctx->nRows;
    const int m = ctx->nInd;
    const long long nU = ctx->f[0].nDofs;
    const bool two = ctx->params.kind == kGradientDamage;
    double* hist = ctx->params.kind == kLinearElastic ? nullptr : kappa;

    DevBuf K, R, dx, xk, x0, ctr;
    Check(K.Alloc((size_t)ctx->nnz * 8), "Newton buffers");
    Check(R.Alloc((size_t)n * 8), "Newton buffers");
    Check(dx.Alloc((size_t)n * 8), "Newton buffers");
    Check(xk.Alloc((size_t)n * 8), "Newton buffers");
    Check(x0.Alloc((size_t)n * 8), "Newton buffers");
    Check(ctr.Alloc((size_t)std::max(m, 1) * 8), "Newton buffers");
    Csr Kc{ctx->nRows, ctx->rowPtr.As<int>(), ctx->col.As<int>(), K.As<double>()};
    Csr Cc{ctx->nRows, ctx->cRow.As<int>(), ctx->cCol.As<int>(), ctx->cVal.As<double>()};
    Csr Ct{m, ctx->ctRow.As<int>(), ctx->ctCol.As<int>(), ctx->ctVal.As<double>()};
    // the norms use the solve workspace's reduction slots between solves (all on one stream, each
    // read before the next solve starts)
    ReducedSolver rs(ctx->stream, Kc, Cc, Ct, n, m, KNodeRows(ctx), SolveWork(ctx), ctx->solveHost);

    auto E = [&](const double* v) -> const double* { return two ? v + nU : nullptr; };
    // QuasistaticSolver::Norm: |C^T v|
    auto Norm = [&](const double* v) {
        rs.Spmv(Ct, v, ctr.As<double>());
        hipError_t err = hipSuccess;
        const double d = m > 0 ? rs.Dot(m, ctr.As<double>(), ctr.As<double>(), err) : 0.0;
        Check(err, "residual norm");
        return std::sqrt(d);
    };
    int linIts = 0, evals = 0;
    auto LinearSolve = [&](const double* Kv, const double* f, const double* ro, const double* rn, bool trial) {
        nuto_solve_info li{};
        SolveConstrained(ctx, Kv, f, ro, rn, trial, dx.As<double>(), &lin, &li);
        linIts += li.iterations;
    };
    auto Residual = [&](const double* v) {
        Assemble(ctx, v, E(v), hist, R.As<double>(), nullptr, true, false);
        evals++;
        return Norm(R.As<double>());
    };

    // TrialState: hessian and gradient at the last converged state, x - SolveTrialState(...)
    Assemble(ctx, x, E(x), hist, R.As<double>(), K.As<double>(), true, true);
    LinearSolve(K.As<double>(), R.As<double>(), rhsOld, rhsNew, true);
    Check(hipMemcpyAsync(xk.p, x, (size_t)n * 8, hipMemcpyDeviceToDevice, ctx->stream), "trial state");
    rs.Axpby(n, -1.0, dx.As<double>(), 1.0, xk.As<double>());

    // NewtonRaphson::Solve
    double norm = Residual(xk.As<double>());
    int it = 0;
    bool converged = norm < tol;
    while (!converged && it < maxIt)
    {
        Assemble(ctx, xk.As<double>(), E(xk.As<double>()), hist, nullptr, K.As<double>(), false, true);
        LinearSolve(K.As<double>(), R.As<double>(), nullptr, nullptr, false);
        ++it;
        if (lsMax == 0)
        {
            rs.Axpby(n, -1.0, dx.As<double>(), 1.0, xk.As<double>());
            norm = Residual(xk.As<double>());
            converged = norm < tol;
            continue;
        }
        // LineSearch: the norm at x0 is the one just computed (the reference re-assembles it; the
        // assembly is deterministic, so the value is the same)
        const double previous = norm;
        Check(hipMemcpyAsync(x0.p, xk.p, (size_t)n * 8, hipMemcpyDeviceToDevice, ctx->stream), "line search");
        double alpha = 1.0;
        for (int step = 0; step < lsMax; step++)
        {
            Check(hipMemcpyAsync(xk.p, x0.p, (size_t)n * 8, hipMemcpyDeviceToDevice, ctx->stream), "line search");
            rs.Axpby(n, -alpha, dx.As<double>(), 1.0, xk.As<double>());
            norm = Residual(xk.As<double>());
            if (norm < tol)
            {
                converged = true;
                break;
            }
            alpha *= 0.5;
            if (norm < (1.0 - alpha) * previous)
                break;
        }
    }
    if (info)
    {
        info->iterations = it;
        info->linear_iterations = linIts;
        info->residual_evaluations = evals;
        info->residual_norm = norm;
    }
    Require(converged, NUTO_ERR_NO_CONVERGENCE, "No convergence after " + std::to_string(it) + " iterations.");
    Require(Norm(xk.As<double>()) <= 1e10, NUTO_ERR_NO_CONVERGENCE, "floating point exception");
    if (hist)
    {
        RowArgs a = MakeRowArgs(ctx, xk.As<double>(), E(xk.As<double>()), hist, nullptr, nullptr);
        Dispatch(ctx, a, false, false, 1, hist);  // UpdateHistory
    }
    Check(hipMemcpyAsync(x, xk.p, (size_t)n * 8, hipMemcpyDeviceToDevice, ctx->stream), "Newton result");
    Check(hipStreamSynchronize(ctx->stream), "Newton step");
}

}  // namespace

extern "C" {

int nuto_solve(nuto_ctx* ctx, const double* K, const double* f, double* x, const nuto_solver_options* options,
               nuto_solve_info* info)
{
    return Guard(ctx, [&] { SolveConstrained(ctx, K, f, nullptr, nullptr, false, x, options, info); });
}

int nuto_solve_trial_state(nuto_ctx* ctx, const double* K, const double* f, const double* rhs_old,
                           const double* rhs_new, double* x, const nuto_solver_options* options,
                           nuto_solve_info* info)
{
    return Guard(ctx, [&] { SolveConstrained(ctx, K, f, rhs_old, rhs_new, true, x, options, info); });
}

int nuto_quasistatic_step(nuto_ctx* ctx, double* x, double* kappa, const double* rhs_old, const double* rhs_new,
                          const nuto_newton_options* options, nuto_newton_info* info)
{
    return Guard(ctx, [&] { QuasistaticStep(ctx, x, kappa, rhs_old, rhs_new, options, info); });
}

int nuto_vjp(nuto_ctx* ctx, const double* u, const double* e, const double* kappa, const double* lambda,
             double* d_state, double* d_theta)
{
    return Guard(ctx, [&] {
        CheckReady(ctx, u, e, kappa);
        Require(ctx->params.kind == kGradientDamage, NUTO_ERR_UNSUPPORTED,
                "the adjoint is provided for the gradient-damage integrand");
        Require(lambda != nullptr && d_state != nullptr, NUTO_ERR_ARGUMENT, "lambda / d_state is NULL");
        Require(ctx->params.interaction == 0, NUTO_ERR_UNSUPPORTED,
                "the adjoint is provided for the constant nonlocal interaction");
        if (!ctx->tabValid)
            BuildTables(ctx);
        PrepareGdm(ctx);
        // nodes beyond the node kernel's bounds (more than 8 cells or 32 neighbours) get the list kernel
        Require(ctx->gdm.Usable(ctx->params.kind, ctx->coordType, ctx->f[0].type, ctx->nip), NUTO_ERR_UNSUPPORTED,
                "the adjoint needs hex8 cells and dofs and 2x2x2 Gauss points");
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

// ================================================================================================
// Batches of independent meshes (declared in include/nuto_amd.h, "batches")
// ================================================================================================

namespace
{

struct BatchMesh
{
    int ct = 0, dt = 0, nCells = 0, nCoord = 0, nDof = 0;
    std::vector<double> coords;
    std::vector<int> conn, ranks;
    int group = -1, nodeOff = 0, cellOff = 0, coordOff = 0;
};

struct BatchGroup
{
    int ct = 0, dt = 0, type = 0;
    nuto_ctx* ctx = nullptr;
    std::vector<int> meshes;
    int nNodes = 0, nCells = 0, nCoord = 0, nip = 0, nRows = 0;
    long long nnz = 0;
    long long uOff = 0, eOff = 0, ROff = 0, KOff = 0, hOff = 0, tOff = 0;  // offsets into the batch arrays
    long long stageCells = -1;  // two-phase chunk: cells whose matrices are staged at once (0: all; -1: default)
};

// default rule of each (coordinate, dof) element pair: tensor Gauss 2 / 3, tetrahedron order 1 / 2
bool BatchTypeOf(int ct, int dt, int& type, int& ruleKind, int& ruleN)
{
    if (ct == kTet4 && dt == kTet4)
        type = kBtTet4, ruleKind = 1, ruleN = 1;
    else if (ct == kTet4 && dt == kTet10)
        type = kBtTet10, ruleKind = 1, ruleN = 2;
    else if (ct == kHex8 && dt == kHex8)
        type = kBtHex8, ruleKind = 0, ruleN = 2;
    else if (ct == kHex8 && dt == kHex20)
        type = kBtHex20, ruleKind = 0, ruleN = 3;
    else
        return false;
    return true;
}

// Default dof nodes per group of a gradient-damage batch: the node kernel takes at most 32
// neighbours per node, each adding 16 nonzeros to the node's four rows, so a group of this many
// nodes keeps its CSR within int32 offsets (2^31 - 1 nonzeros).
constexpr long long kGdmGroupNodes = 2147483647LL / (16 * 32);

// 0..2 fused launch, 64/128/256 nodes per tile; 3 one launch per type; 4 one launch per mesh;
// 5 two-phase (cell matrices + row gather) per type; 6 two-phase per mesh
constexpr int kBatchModes = 7;
constexpr int kBatchBlock[3] = {64, 128, 256};

}  // namespace

struct nuto_batch
{
    int device = 0;
    hipStream_t own = nullptr, stream = nullptr;
    std::string err;
    nuto_params params{};
    bool integrandSet = false;
    std::vector<BatchMesh> meshes;
    std::vector<BatchGroup> groups;
    bool built = false;
    long long uLen = 0, eLen = 0, RLen = 0, KLen = 0, hLen = 0, tLen = 0;
    DevBuf args;  // RowArgs per group for the fused kernel
    const void* argPtrs[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    DevBuf tiles[3];
    int nTiles[3] = {0, 0, 0};
    int tuned = -1;  // launch mode chosen by nuto_batch_autotune (-1: not tuned)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graphKey = -1;  // mode the graph was captured for (+100 for forward + adjoint)
    long long groupNodes = 0;  // dof nodes per group (nuto_batch_set_group_limit; 0: the default)
    const void* graphPtrs[10] = {};
    DevBuf lambda;      // forward + adjoint: lambda = R - r_obs
};

namespace
{

template <class F>
int GuardB(nuto_batch* b, F&& fn)
{
    if (!b)
        return NUTO_ERR_ARGUMENT;
    try
    {
        Check(hipSetDevice(b->device), "hipSetDevice");
        fn();
        b->err.clear();
        return NUTO_OK;
    }
    catch (const Failure& f)
    {
        b->err = f.msg;
        return f.code;
    }
    catch (const std::bad_alloc&)
    {
        b->err = "host allocation failed";
        return NUTO_ERR_MEMORY;
    }
    catch (const std::exception& e)
    {
        b->err = e.what();
        return NUTO_ERR_ARGUMENT;
    }
}

// a failing call on a group context becomes a failure of the batch
void CheckCtx(nuto_ctx* ctx, int status)
{
    if (status != NUTO_OK)
        throw Failure{status, std::string("batch group: ") + ctx->err};
}

void DropGraph(nuto_batch* b)
{
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graphKey = -1;
}

void ReleaseGroups(nuto_batch* b)
{
    DropGraph(b);
    for (BatchGroup& g : b->groups)
        if (g.ctx)
            nuto_destroy(g.ctx);
    b->groups.clear();
    b->built = false;
    b->tuned = -1;
    for (auto& p : b->argPtrs)
        p = nullptr;
}

RowArgs GroupArgs(nuto_batch* b, const BatchGroup& g, const double* u, const double* e, const double* kappa,
                  double* R, double* K)
{
    return MakeRowArgs(g.ctx, u + g.uOff, e ? e + g.eOff : nullptr, kappa ? kappa + g.hOff : nullptr, R + g.ROff,
                       K + g.KOff);
}

// Device copy of the per-group arguments of the fused kernel, refreshed when the caller's arrays change.
void UploadArgs(nuto_batch* b, const double* u, const double* e, const double* kappa, double* R, double* K)
{
    const void* now[5] = {u, e, kappa, R, K};
    if (std::equal(now, now + 5, b->argPtrs) && b->args.p)
        return;
    std::vector<RowArgs> a;
    for (const BatchGroup& g : b->groups)
        a.push_back(GroupArgs(b, g, u, e, kappa, R, K));
    const size_t bytes = a.size() * sizeof(RowArgs);
    if (!b->args.p || b->args.bytes != bytes)
    {
        DropGraph(b);  // a captured graph refers to the old buffer
        Check(b->args.Alloc(bytes), "batch arguments");
    }
    // the buffer is rewritten in place: graphs captured with it stay valid
    Check(hipMemcpyAsync(b->args.p, a.data(), bytes, hipMemcpyHostToDevice, b->stream), "batch arguments");
    Check(hipStreamSynchronize(b->stream), "batch arguments");
    std::copy(now, now + 5, b->argPtrs);
}

template <int KIND>
void LaunchFused(nuto_batch* b, int which)
{
    const int n = b->nTiles[which];
    if (n == 0)
        return;
    const RowArgs* a = b->args.As<RowArgs>();
    const BatchTile* t = b->tiles[which].As<BatchTile>();
    if (which == 0)
        hipLaunchKernelGGL((kBatchRows<KIND, 64>), dim3(n), dim3(64), 0, b->stream, a, t);
    else if (which == 1)
        hipLaunchKernelGGL((kBatchRows<KIND, 128>), dim3(n), dim3(128), 0, b->stream, a, t);
    else
        hipLaunchKernelGGL((kBatchRows<KIND, 256>), dim3(n), dim3(256), 0, b->stream, a, t);
}

template <int KIND>
void LaunchRange(const RowArgs& a, int type, int first, int end, hipStream_t stream)
{
    if (end <= first)
        return;
    const dim3 grid((end - first + 63) / 64), block(64);
    if (type == kBtTet4)
        hipLaunchKernelGGL((kBatchRowsRange<KIND, kBtTet4>), grid, block, 0, stream, a, first, end);
    else if (type == kBtTet10)
        hipLaunchKernelGGL((kBatchRowsRange<KIND, kBtTet10>), grid, block, 0, stream, a, first, end);
    else if (type == kBtHex8)
        hipLaunchKernelGGL((kBatchRowsRange<KIND, kBtHex8>), grid, block, 0, stream, a, first, end);
    else
        hipLaunchKernelGGL((kBatchRowsRange<KIND, kBtHex20>), grid, block, 0, stream, a, first, end);
}

template <int KIND>
void LaunchTwoPhaseType(nuto_ctx* ctx, const RowArgs& a, int type, int c0, int c1, int p0, int p1, hipStream_t stream)
{
    if (type == kBtTet4)
        LaunchTwoPhaseT<KIND, 4, 4, 1>(ctx, a, true, c0, c1, p0, p1, stream);
    else if (type == kBtTet10)
        LaunchTwoPhaseT<KIND, 4, 10, 4>(ctx, a, true, c0, c1, p0, p1, stream);
    else if (type == kBtHex8)
        LaunchTwoPhaseT<KIND, 8, 8, 8>(ctx, a, true, c0, c1, p0, p1, stream);
    else
        LaunchTwoPhaseT<KIND, 8, 20, 27>(ctx, a, true, c0, c1, p0, p1, stream);
}

bool IsGdm(const nuto_batch* b)
{
    return b->params.kind == kGradientDamage;
}

// cells of a two-phase chunk of `bytes` of staging (matrix and vector per cell)
long long StageCellsOf(const BatchGroup& g, size_t bytes)
{
    const size_t nd = 3 * (size_t)NodesOf(g.dt);
    return std::max<long long>(1, (long long)(bytes / ((nd * nd + nd) * sizeof(double))));
}

// default chunk: the whole group when its staging can be allocated, else 1 GiB of staging
long long DefaultStageCells(BatchGroup& g)
{
    if (TryStage(g.ctx, 3 * NodesOf(g.dt), g.nCells))
        return 0;
    return StageCellsOf(g, size_t(1) << 30);
}

// cells of the largest chunk LaunchTwoPhaseGroup forms with the group's stageCells
long long MaxChunkCells(const nuto_batch* b, const BatchGroup& g)
{
    if (g.stageCells <= 0 || g.nCells <= g.stageCells)
        return g.nCells;
    long long best = 0;
    size_t i = 0;
    while (i < g.meshes.size())
    {
        long long cells = b->meshes[g.meshes[i]].nCells;
        size_t j = i + 1;
        while (j < g.meshes.size() && cells + b->meshes[g.meshes[j]].nCells <= g.stageCells)
            cells += b->meshes[g.meshes[j++]].nCells;
        best = std::max(best, cells);
        i = j;
    }
    return best;
}

// Staging of every group's two-phase launch (mode 5) allocated now, before a graph capture of the
// launch (which must not allocate); the chunk defaults where none is chosen yet
void PrepareStaging(nuto_batch* b)
{
    if (IsGdm(b))
        return;
    for (BatchGroup& g : b->groups)
    {
        if (g.stageCells < 0)
            g.stageCells = DefaultStageCells(g);
        EnsureStage(g.ctx, 3 * NodesOf(g.dt), MaxChunkCells(b, g));
    }
}

// Two-phase launches of a group in chunks of consecutive meshes whose cells fit stageCells (0: the
// whole group at once; a mesh larger than the chunk is a chunk of its own). The rows of a mesh only
// touch its own cells, so each chunk's staging is consumed by its gather before the next chunk
// rewrites it (same stream).
template <int KIND>
void LaunchTwoPhaseGroup(nuto_batch* b, BatchGroup& g, const RowArgs& a)
{
    if (g.stageCells < 0)
        g.stageCells = DefaultStageCells(g);
    if (g.stageCells == 0 || g.nCells <= g.stageCells)
    {
        LaunchTwoPhaseType<KIND>(g.ctx, a, g.type, 0, g.nCells, 0, g.nNodes, b->stream);
        return;
    }
    size_t i = 0;
    while (i < g.meshes.size())
    {
        const BatchMesh& first = b->meshes[g.meshes[i]];
        long long cells = first.nCells;
        size_t j = i + 1;
        while (j < g.meshes.size() && cells + b->meshes[g.meshes[j]].nCells <= g.stageCells)
            cells += b->meshes[g.meshes[j++]].nCells;
        const BatchMesh& last = b->meshes[g.meshes[j - 1]];
        LaunchTwoPhaseType<KIND>(g.ctx, a, g.type, first.cellOff, last.cellOff + last.nCells, first.nodeOff,
                                 last.nodeOff + last.nDof, b->stream);
        i = j;
    }
}

// The kernels of one batched R + K assembly in launch mode `mode` (no synchronisation).
void LaunchBatch(nuto_batch* b, int mode, const double* u, const double* e, const double* kappa, double* R, double* K)
{
    const int kind = b->params.kind;
    if (IsGdm(b) && mode < 3)
        mode = 3;
    if (IsGdm(b) && mode >= 5)
        mode -= 2;  // gradient damage has one kernel form: per group (3) or per mesh (4)
    if (mode < 3)
    {
        UploadArgs(b, u, e, kappa, R, K);
        if (kind == kLinearElastic)
            LaunchFused<kLinearElastic>(b, mode);
        else
            LaunchFused<kLocalDamage>(b, mode);
    }
    else
    {
        for (BatchGroup& g : b->groups)
        {
            RowArgs a = GroupArgs(b, g, u, e, kappa, R, K);
            if (mode >= 5 && !IsGdm(b))
            {
                // two-phase: cell matrices, then the row gather, per element type (5) or per mesh (6)
                if (mode == 5)
                {
                    if (kind == kLinearElastic)
                        LaunchTwoPhaseGroup<kLinearElastic>(b, g, a);
                    else
                        LaunchTwoPhaseGroup<kLocalDamage>(b, g, a);
                    continue;
                }
                for (int m : g.meshes)
                {
                    const BatchMesh& bm = b->meshes[m];
                    const int c0 = bm.cellOff, c1 = bm.cellOff + bm.nCells;
                    const int p0 = bm.nodeOff, p1 = bm.nodeOff + bm.nDof;
                    if (kind == kLinearElastic)
                        LaunchTwoPhaseType<kLinearElastic>(g.ctx, a, g.type, c0, c1, p0, p1, b->stream);
                    else
                        LaunchTwoPhaseType<kLocalDamage>(g.ctx, a, g.type, c0, c1, p0, p1, b->stream);
                }
                continue;
            }
            if (mode == 3)
            {
                if (IsGdm(b))
                    g.ctx->gdm.Launch(a, true, true, b->stream);
                else if (kind == kLinearElastic)
                    LaunchRange<kLinearElastic>(a, g.type, 0, g.nNodes, b->stream);
                else
                    LaunchRange<kLocalDamage>(a, g.type, 0, g.nNodes, b->stream);
                continue;
            }
            for (int m : g.meshes)
            {
                const BatchMesh& bm = b->meshes[m];
                const int first = bm.nodeOff, end = bm.nodeOff + bm.nDof;
                if (IsGdm(b))
                    g.ctx->gdm.LaunchRange(a, first, end, b->stream);
                else if (kind == kLinearElastic)
                    LaunchRange<kLinearElastic>(a, g.type, first, end, b->stream);
                else
                    LaunchRange<kLocalDamage>(a, g.type, first, end, b->stream);
            }
            if (IsGdm(b))
                g.ctx->gdm.LaunchExtras(a, true, true, b->stream);
        }
    }
    Check(hipGetLastError(), "batch launch");
}

__global__ void kResidualMinus(const double* __restrict__ R, const double* __restrict__ robs, double* __restrict__ lam,
                               long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        lam[i] = R[i] - robs[i];
}

void LaunchVjp(nuto_batch* b, const double* u, const double* e, const double* kappa, const double* lam,
               double* dstate, double* dtheta)
{
    for (BatchGroup& g : b->groups)
    {
        RowArgs a = MakeRowArgs(g.ctx, u + g.uOff, e + g.eOff, kappa + g.hOff, nullptr, nullptr);
        LaunchGdmHex8Vjp(g.ctx->gdm, a, lam + g.ROff, dstate + g.ROff, dtheta ? dtheta + g.tOff : nullptr, b->stream);
    }
    Check(hipGetLastError(), "batch vjp launch");
}

// Replays (capturing first if needed) a graph of `capture`, keyed by `key` and the pointers.
template <class F>
void GraphRun(nuto_batch* b, int key, const void* const* ptrs, int nptr, F&& capture)
{
    Require(b->stream != nullptr, NUTO_ERR_STATE, "graph replay needs a non-null stream (nuto_batch_set_stream)");
    bool same = b->exec && b->graphKey == key;
    for (int i = 0; same && i < nptr; i++)
        same = b->graphPtrs[i] == ptrs[i];
    if (!same)
    {
        DropGraph(b);
        Check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "graph capture");
        try
        {
            capture();
        }
        catch (...)
        {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(b->stream, &g);
            if (g)
                (void)hipGraphDestroy(g);
            throw;
        }
        Check(hipStreamEndCapture(b->stream, &b->graph), "graph capture");
        Check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
        b->graphKey = key;
        for (int i = 0; i < nptr; i++)
            b->graphPtrs[i] = ptrs[i];
    }
    Check(hipGraphLaunch(b->exec, b->stream), "graph launch");
}

void RequireBuilt(nuto_batch* b, const double* u, const double* e, const double* kappa)
{
    Require(b->built, NUTO_ERR_STATE, "batch not built (nuto_batch_build)");
    Require(u != nullptr, NUTO_ERR_ARGUMENT, "u is NULL");
    Require(!IsGdm(b) || e != nullptr, NUTO_ERR_ARGUMENT, "nonlocal equivalent strain vector is NULL");
    Require(b->params.kind == kLinearElastic || kappa != nullptr, NUTO_ERR_ARGUMENT, "history (kappa) is NULL");
}

}  // namespace

extern "C" {

int nuto_batch_create(nuto_batch** out, int device)
{
    if (!out)
        return NUTO_ERR_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return NUTO_ERR_DEVICE;
    if (device < 0 || device >= count)
        return NUTO_ERR_ARGUMENT;
    nuto_batch* b = new (std::nothrow) nuto_batch;
    if (!b)
        return NUTO_ERR_MEMORY;
    b->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking) != hipSuccess)
    {
        delete b;
        return NUTO_ERR_DEVICE;
    }
    b->stream = b->own;
    *out = b;
    return NUTO_OK;
}

int nuto_batch_destroy(nuto_batch* b)
{
    if (!b)
        return NUTO_ERR_ARGUMENT;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    ReleaseGroups(b);
    if (b->own)
        (void)hipStreamDestroy(b->own);
    delete b;
    return NUTO_OK;
}

const char* nuto_batch_last_error(const nuto_batch* b)
{
    return b ? b->err.c_str() : "null batch";
}

int nuto_batch_set_stream(nuto_batch* b, void* stream)
{
    return GuardB(b, [&] {
        DropGraph(b);
        b->stream = (hipStream_t)stream;
        for (BatchGroup& g : b->groups)
            CheckCtx(g.ctx, nuto_set_stream(g.ctx, stream));
    });
}

int nuto_batch_synchronize(nuto_batch* b)
{
    return GuardB(b, [&] { Check(hipStreamSynchronize(b->stream), "synchronize"); });
}

int nuto_batch_set_integrand(nuto_batch* b, const nuto_params* params)
{
    return GuardB(b, [&] {
        Require(params != nullptr, NUTO_ERR_ARGUMENT, "params is NULL");
        Require(params->dim == 3, NUTO_ERR_UNSUPPORTED, "batches hold 3D meshes");
        Require(params->interaction == 0, NUTO_ERR_UNSUPPORTED, "batches take the constant nonlocal interaction");
        Require(params->kind == kLinearElastic || params->kind == kLocalDamage || params->kind == kGradientDamage,
                NUTO_ERR_ARGUMENT, "unknown integrand kind");
        ReleaseGroups(b);
        b->params = *params;
        b->integrandSet = true;
    });
}

int nuto_batch_add_mesh(nuto_batch* b, int coord_type, int dof_type, int n_cells, int n_coord_nodes,
                        const double* coords, const int* coord_conn, int n_dof_nodes, const int* cell_ranks, int* mesh)
{
    return GuardB(b, [&] {
        int type, rk, rn;
        Require(BatchTypeOf(coord_type, dof_type, type, rk, rn), NUTO_ERR_UNSUPPORTED,
                "batches hold tet4/tet4, tet4/tet10, hex8/hex8 and hex8/hex20 meshes");
        Require(n_cells >= 0 && n_coord_nodes >= 0 && n_dof_nodes >= 0, NUTO_ERR_ARGUMENT, "negative sizes");
        Require(coords && coord_conn && cell_ranks, NUTO_ERR_ARGUMENT, "NULL mesh array");
        const int nnc = NodesOf(coord_type), nnd = NodesOf(dof_type);
        BatchMesh m;
        m.ct = coord_type;
        m.dt = dof_type;
        m.nCells = n_cells;
        m.nCoord = n_coord_nodes;
        m.nDof = n_dof_nodes;
        m.coords.assign(coords, coords + (size_t)n_coord_nodes * 3);
        m.conn.assign(coord_conn, coord_conn + (size_t)n_cells * nnc);
        m.ranks.assign(cell_ranks, cell_ranks + (size_t)n_cells * nnd);
        for (int v : m.conn)
            Require(v >= 0 && v < n_coord_nodes, NUTO_ERR_ARGUMENT, "coordinate connectivity out of range");
        for (int v : m.ranks)
            Require(v >= 0 && v < n_dof_nodes, NUTO_ERR_ARGUMENT, "dof node rank out of range");
        ReleaseGroups(b);
        b->meshes.push_back(std::move(m));
        if (mesh)
            *mesh = (int)b->meshes.size() - 1;
    });
}

int nuto_batch_set_group_limit(nuto_batch* b, long long max_dof_nodes)
{
    return GuardB(b, [&] {
        Require(max_dof_nodes >= 0, NUTO_ERR_ARGUMENT, "negative group limit");
        ReleaseGroups(b);
        b->groupNodes = max_dof_nodes;
    });
}

int nuto_batch_set_stage_bytes(nuto_batch* b, long long bytes)
{
    return GuardB(b, [&] {
        Require(b->built, NUTO_ERR_STATE, "batch not built (nuto_batch_build)");
        Require(bytes >= 0, NUTO_ERR_ARGUMENT, "negative staging budget");
        for (BatchGroup& g : b->groups)
            g.stageCells = bytes == 0 ? 0 : StageCellsOf(g, (size_t)bytes);
        DropGraph(b);
        PrepareStaging(b);
    });
}

int nuto_batch_build(nuto_batch* b)
{
    return GuardB(b, [&] {
        Require(b->integrandSet, NUTO_ERR_STATE, "no integrand set");
        ReleaseGroups(b);
        const bool gdm = IsGdm(b);
        // groups in the order hex20, tet10, hex8, tet4 (heaviest rows first); the meshes of one element
        // pair, in the order they were added, fill groups of at most `cap` dof nodes (a larger mesh
        // forms a group of its own)
        const long long cap = b->groupNodes > 0 ? b->groupNodes : (gdm ? kGdmGroupNodes : LLONG_MAX);
        // a group's CSR has int32 offsets: when the type's cells could exceed 2^31 - 1 nonzeros (each
        // cell adds at most (components x nodes)^2), the meshes' exact counts close groups in time
        const long long kMaxNnz = 2147483647LL;
        const int nc = gdm ? 4 : 3;  // components per dof node
        const int order[4][2] = {{kHex8, kHex20}, {kTet4, kTet10}, {kHex8, kHex8}, {kTet4, kTet4}};
        std::vector<BatchGroup> plan;
        for (const auto& ctdt : order)
        {
            BatchGroup g;
            g.ct = ctdt[0];
            g.dt = ctdt[1];
            int rk, rn;
            BatchTypeOf(g.ct, g.dt, g.type, rk, rn);
            const long long nd = (long long)nc * NodesOf(g.dt);
            long long bound = 0;
            for (const BatchMesh& bm : b->meshes)
                if (bm.ct == g.ct && bm.dt == g.dt)
                    bound += (long long)bm.nCells * nd * nd;
            const bool exact = bound > kMaxNnz;
            // exact nonzeros of the type's meshes, the meshes on several threads
            std::vector<long long> mnzOf(b->meshes.size(), 0);
            if (exact)
            {
                std::vector<int> todo;
                for (size_t m = 0; m < b->meshes.size(); m++)
                    if (b->meshes[m].ct == g.ct && b->meshes[m].dt == g.dt)
                        todo.push_back((int)m);
                std::vector<std::string> errs(todo.size());
                std::atomic<size_t> next{0};
                auto count = [&] {
                    for (size_t k; (k = next.fetch_add(1)) < todo.size();)
                    {
                        const BatchMesh& bm = b->meshes[todo[k]];
                        NodeGraph mg;
                        if (BuildNodeGraph(bm.nCells, NodesOf(bm.dt), bm.ranks.data(), bm.nDof, mg, errs[k], 1))
                            mnzOf[todo[k]] = (long long)nc * nc * (long long)mg.adj.size();
                        else if (errs[k].empty())
                            errs[k] = "node graph";
                    }
                };
                const unsigned hw = std::thread::hardware_concurrency();
                const int nt = (int)std::max(1u, std::min({hw ? hw : 1u, 16u, (unsigned)todo.size()}));
                std::vector<std::thread> pool;
                for (int t = 1; t < nt; t++)
                    pool.emplace_back(count);
                count();
                for (std::thread& th : pool)
                    th.join();
                for (const std::string& e : errs)
                    Require(e.empty(), NUTO_ERR_ARGUMENT, e);
            }
            long long nodes = 0, nnz = 0;
            for (size_t m = 0; m < b->meshes.size(); m++)
            {
                const BatchMesh& bm = b->meshes[m];
                if (bm.ct != g.ct || bm.dt != g.dt)
                    continue;
                Require(!gdm || (g.ct == kHex8 && g.dt == kHex8), NUTO_ERR_UNSUPPORTED,
                        "gradient-damage batches hold hex8 meshes only");
                const long long mnz = mnzOf[m];
                Require(mnz <= kMaxNnz, NUTO_ERR_UNSUPPORTED,
                        "a mesh of the batch exceeds 2^31 - 1 nonzeros (int32 CSR offsets)");
                if (!g.meshes.empty() && (nodes + bm.nDof > cap || nnz + mnz > kMaxNnz))
                {
                    plan.push_back(g);
                    g.meshes.clear();
                    nodes = 0;
                    nnz = 0;
                }
                g.meshes.push_back((int)m);
                nodes += bm.nDof;
                nnz += mnz;
            }
            if (!g.meshes.empty())
                plan.push_back(g);
        }
        for (BatchGroup& g : plan)
        {
            int rk, rn;
            BatchTypeOf(g.ct, g.dt, g.type, rk, rn);
            const int gi = (int)b->groups.size();
            std::vector<double> X;
            std::vector<int> C, Rk;
            for (int m : g.meshes)
            {
                BatchMesh& bm = b->meshes[m];
                bm.group = gi;
                bm.nodeOff = g.nNodes;
                bm.cellOff = g.nCells;
                bm.coordOff = g.nCoord;
                X.insert(X.end(), bm.coords.begin(), bm.coords.end());
                for (int v : bm.conn)
                    C.push_back(v + bm.coordOff);
                for (int v : bm.ranks)
                    Rk.push_back(v + bm.nodeOff);
                g.nNodes += bm.nDof;
                g.nCells += bm.nCells;
                g.nCoord += bm.nCoord;
            }
            int st = nuto_create(&g.ctx, b->device);
            if (st != NUTO_OK)
                throw Failure{st, "batch group context"};
            b->groups.push_back(g);
            BatchGroup& G = b->groups.back();
            nuto_ctx* c = G.ctx;
            CheckCtx(c, nuto_set_stream(c, b->stream));
            CheckCtx(c, nuto_set_cells(c, G.ct, G.nCells, G.nCoord, X.data(), C.data(), nullptr));
            std::vector<int> cd(Rk.size() * 3);
            for (size_t i = 0; i < Rk.size(); i++)
                for (int k = 0; k < 3; k++)
                    cd[i * 3 + k] = Rk[i] * 3 + k;
            CheckCtx(c, nuto_set_dof_type(c, 0, G.dt, 3, 3 * G.nNodes, cd.data()));
            if (gdm)
                CheckCtx(c, nuto_set_dof_type(c, 1, G.dt, 1, G.nNodes, Rk.data()));
            CheckCtx(c, nuto_set_integrand(c, &b->params));
            CheckCtx(c, nuto_set_rule(c, rk, rn));
            long long nnz = 0;
            int rows = 0;
            CheckCtx(c, nuto_build_pattern(c, &nnz, &rows));
            CheckCtx(c, Guard(c, [&] {
                BuildTables(c);
                if (gdm)
                {
                    PrepareGdm(c);
                    Require(c->gdm.Usable(c->params.kind, c->coordType, c->f[0].type, c->nip), NUTO_ERR_UNSUPPORTED,
                            "gradient-damage batches need the hex8 node kernel's tables");
                }
            }));
            G.nnz = nnz;
            G.nRows = rows;
            G.nip = c->nip;
        }
        // offsets of every group in the batch arrays
        b->uLen = b->eLen = b->RLen = b->KLen = b->hLen = b->tLen = 0;
        for (BatchGroup& g : b->groups)
        {
            g.uOff = b->uLen;
            g.eOff = b->eLen;
            g.ROff = b->RLen;
            g.KOff = b->KLen;
            g.hOff = b->hLen;
            g.tOff = b->tLen;
            b->uLen += 3LL * g.nNodes;
            b->eLen += gdm ? g.nNodes : 0;
            b->RLen += g.nRows;
            b->KLen += g.nnz;
            b->hLen += (long long)g.nCells * g.nip;
            b->tLen += 3LL * g.nCells;
        }
        // tiles of the fused kernel (groups are already heaviest first)
        for (int w = 0; w < 3; w++)
        {
            std::vector<BatchTile> t;
            for (size_t gi = 0; gi < b->groups.size() && !gdm; gi++)
            {
                const BatchGroup& g = b->groups[gi];
                for (int first = 0; first < g.nNodes; first += kBatchBlock[w])
                    t.push_back(BatchTile{(int)gi, first, std::min(first + kBatchBlock[w], g.nNodes), g.type});
            }
            b->nTiles[w] = (int)t.size();
            Check(b->tiles[w].Upload(t.data(), t.size() * sizeof(BatchTile), b->stream), "batch tiles");
        }
        b->built = true;
    });
}

int nuto_batch_sizes(nuto_batch* b, long long* sizes)
{
    return GuardB(b, [&] {
        Require(b->built, NUTO_ERR_STATE, "batch not built (nuto_batch_build)");
        Require(sizes != nullptr, NUTO_ERR_ARGUMENT, "sizes is NULL");
        sizes[0] = b->uLen;
        sizes[1] = b->eLen;
        sizes[2] = b->RLen;
        sizes[3] = b->KLen;
        sizes[4] = b->hLen;
        sizes[5] = b->tLen;
    });
}

int nuto_batch_group_info(nuto_batch* b, int group, long long* info)
{
    return GuardB(b, [&] {
        Require(b->built, NUTO_ERR_STATE, "batch not built (nuto_batch_build)");
        Require(group >= 0 && group < (int)b->groups.size() && info, NUTO_ERR_ARGUMENT, "bad group index / info");
        const BatchGroup& g = b->groups[group];
        info[0] = g.dt;
        info[1] = (long long)g.meshes.size();
        info[2] = g.nNodes;
        info[3] = g.nCells;
        info[4] = g.ctx->cellGram;
        info[5] = (long long)b->groups.size();
        info[6] = g.stageCells;
    });
}

int nuto_batch_mesh_info(nuto_batch* b, int mesh, long long* info)
{
    return GuardB(b, [&] {
        Require(b->built, NUTO_ERR_STATE, "batch not built (nuto_batch_build)");
        Require(mesh >= 0 && mesh < (int)b->meshes.size() && info, NUTO_ERR_ARGUMENT, "bad mesh index / info");
        const BatchMesh& m = b->meshes[mesh];
        const BatchGroup& g = b->groups[m.group];
        const int* rp = g.ctx->graph.adjPtr.data();  // neighbour counts give the row lengths
        const bool gdm = IsGdm(b);
        const int w = gdm ? 4 : 3;  // columns per neighbour
        // nonzeros before the mesh's first node and within the mesh, per row group
        long long before = 0, inside = 0;
        for (int p = 0; p < m.nodeOff; p++)
            before += (long long)(rp[p + 1] - rp[p]) * w;
        for (int p = m.nodeOff; p < m.nodeOff + m.nDof; p++)
            inside += (long long)(rp[p + 1] - rp[p]) * w;
        const long long nnzU = 3 * inside, nnzE = gdm ? inside : 0;
        info[0] = m.group;
        info[1] = m.nDof;
        info[2] = m.nCells;
        info[3] = g.nip;
        info[4] = g.uOff + 3LL * m.nodeOff;                     // u
        info[5] = gdm ? g.eOff + m.nodeOff : -1;                // eeq
        info[6] = g.ROff + 3LL * m.nodeOff;                     // R, displacement rows
        info[7] = gdm ? g.ROff + 3LL * g.nNodes + m.nodeOff : -1;  // R, eeq rows
        info[8] = (long long)m.nDof * (gdm ? 4 : 3);            // rows of the mesh
        info[9] = g.KOff + 3 * before;                          // K, displacement rows
        info[10] = nnzU;
        info[11] = gdm ? g.KOff + 3LL * (g.nnz / 4) + before : -1;  // K, eeq rows (after all displacement rows)
        info[12] = nnzE;
        info[13] = g.hOff + m.cellOff;                          // history of (cell 0, ip 0)
        info[14] = g.nCells;                                    // history leading dimension
        info[15] = g.tOff + m.cellOff;                          // d_theta of cell 0 (E row)
        info[16] = g.nCells;                                    // d_theta stride between E, f_t, G_f
    });
}

int nuto_batch_copy_pattern(nuto_batch* b, int mesh, int* row_ptr, int* col)
{
    return GuardB(b, [&] {
        Require(b->built, NUTO_ERR_STATE, "batch not built (nuto_batch_build)");
        Require(mesh >= 0 && mesh < (int)b->meshes.size() && row_ptr && col, NUTO_ERR_ARGUMENT, "bad arguments");
        const BatchMesh& m = b->meshes[mesh];
        const BatchGroup& g = b->groups[m.group];
        const NodeGraph& G = g.ctx->graph;
        const bool gdm = IsGdm(b);
        const int nU = 3 * m.nDof;
        // mesh-local monolithic CSR: rows [u ; e], columns [u, e], ascending (the group graph's
        // neighbours of a node are the mesh's nodes, offset by nodeOff)
        long long k = 0;
        int row = 0;
        auto emitRow = [&](int p) {
            row_ptr[row++] = (int)k;
            for (int a = G.adjPtr[p]; a < G.adjPtr[p + 1]; a++)
                for (int c = 0; c < 3; c++)
                    col[k++] = 3 * (G.adj[a] - m.nodeOff) + c;
            if (gdm)
                for (int a = G.adjPtr[p]; a < G.adjPtr[p + 1]; a++)
                    col[k++] = nU + (G.adj[a] - m.nodeOff);
        };
        for (int p = m.nodeOff; p < m.nodeOff + m.nDof; p++)
            for (int c = 0; c < 3; c++)
                emitRow(p);
        if (gdm)
            for (int p = m.nodeOff; p < m.nodeOff + m.nDof; p++)
                emitRow(p);
        row_ptr[row] = (int)k;
    });
}

int nuto_batch_set_material_fields(nuto_batch* b, const double* E, const double* ft, const double* gf)
{
    return GuardB(b, [&] {
        Require(b->built, NUTO_ERR_STATE, "batch not built (nuto_batch_build)");
        Require((E && ft && gf) || (!E && !ft && !gf), NUTO_ERR_ARGUMENT, "E, ft and gf must all be given");
        // host arrays over all cells, meshes in the order they were added
        DropGraph(b);  // captured launches carry the old field buffers
        std::vector<long long> base(b->meshes.size(), 0);
        for (size_t m = 1; m < b->meshes.size(); m++)
            base[m] = base[m - 1] + b->meshes[m - 1].nCells;
        for (BatchGroup& g : b->groups)
        {
            if (!E)
            {
                CheckCtx(g.ctx, nuto_set_material_fields(g.ctx, nullptr, nullptr, nullptr));
                continue;
            }
            std::vector<double> e(g.nCells), f(g.nCells), h(g.nCells);
            for (int m : g.meshes)
            {
                const BatchMesh& bm = b->meshes[m];
                for (int c = 0; c < bm.nCells; c++)
                {
                    e[bm.cellOff + c] = E[base[m] + c];
                    f[bm.cellOff + c] = ft[base[m] + c];
                    h[bm.cellOff + c] = gf[base[m] + c];
                }
            }
            CheckCtx(g.ctx, nuto_set_material_fields(g.ctx, e.data(), f.data(), h.data()));
        }
        b->argPtrs[0] = nullptr;  // the group arguments carry the field pointers
    });
}

int nuto_batch_assemble(nuto_batch* b, const double* u, const double* e, const double* kappa, double* R, double* K,
                        int mode)
{
    return GuardB(b, [&] {
        RequireBuilt(b, u, e, kappa);
        Require(R && K, NUTO_ERR_ARGUMENT, "R / K is NULL");
        Require(mode >= -1 && mode < kBatchModes, NUTO_ERR_ARGUMENT, "unknown launch mode");
        if (mode == -1)
        {
            // the tuned (or default fused) launch, replayed from a graph
            const int m = b->tuned >= 0 ? b->tuned : 3;
            if (m < 3)
                UploadArgs(b, u, e, kappa, R, K);  // outside the capture
            if (m == 5)
                PrepareStaging(b);  // outside the capture
            const void* p[5] = {u, e, kappa, R, K};
            GraphRun(b, m, p, 5, [&] { LaunchBatch(b, m, u, e, kappa, R, K); });
        }
        else
            LaunchBatch(b, mode, u, e, kappa, R, K);
    });
}

int nuto_batch_autotune(nuto_batch* b, const double* u, const double* e, const double* kappa, double* R, double* K,
                        double* ms, int* best)
{
    return GuardB(b, [&] {
        RequireBuilt(b, u, e, kappa);
        Require(R && K, NUTO_ERR_ARGUMENT, "R / K is NULL");
        hipEvent_t t0, t1;
        Check(hipEventCreate(&t0), "event");
        Check(hipEventCreate(&t1), "event");
        // median of 5 timed runs after one warm-up; a launch of more than 50 ms is timed once (large
        // batches: the spread is small against the run and the tuning stays within seconds)
        auto median = [&](auto&& launch) {
            launch();
            std::vector<float> rep;
            for (int it = 0; it < 5; it++)
            {
                if (it == 1 && rep[0] > 50.0f)
                    break;
                Check(hipEventRecord(t0, b->stream), "event");
                launch();
                Check(hipEventRecord(t1, b->stream), "event");
                Check(hipEventSynchronize(t1), "event");
                float x = 0.0f;
                Check(hipEventElapsedTime(&x, t0, t1), "event");
                rep.push_back(x);
            }
            std::sort(rep.begin(), rep.end());
            return (double)rep[rep.size() / 2];
        };
        // the two-phase modes (not for gradient damage): each group's cell-matrix kernel (by columns, Gram
        // form with one or two row nodes per lane or on the matrix cores) and staging chunk (the whole group when it fits, 1 GiB, 256 MiB, 64 MiB of staging),
        // timed on the group's two-phase launch
        const bool staged = !IsGdm(b);
        if (staged)
        {
            for (BatchGroup& g : b->groups)
            {
                const RowArgs a = GroupArgs(b, g, u, e, kappa, R, K);
                const int nd = 3 * NodesOf(g.dt);
                std::vector<long long> chunks;
                if (TryStage(g.ctx, nd, g.nCells))
                    chunks.push_back(0);
                for (size_t mib : {1024, 256, 64})
                {
                    const long long c = StageCellsOf(g, mib << 20);
                    if (c < g.nCells && TryStage(g.ctx, nd, c))
                        chunks.push_back(c);
                }
                Require(!chunks.empty(), NUTO_ERR_MEMORY, "no staging for the two-phase assembly");
                double bestT = 1e300;
                int bestGram = 1;
                long long bestChunk = chunks[0];
                for (int gram = 0; gram < kCellForms; gram++)
                    for (long long c : chunks)
                    {
                        g.ctx->cellGram = gram;
                        g.stageCells = c;
                        const double t = median([&] {
                            if (b->params.kind == kLinearElastic)
                                LaunchTwoPhaseGroup<kLinearElastic>(b, g, a);
                            else
                                LaunchTwoPhaseGroup<kLocalDamage>(b, g, a);
                        });
                        if (t < bestT)
                        {
                            bestT = t;
                            bestGram = gram;
                            bestChunk = c;
                        }
                    }
                g.ctx->cellGram = bestGram;
                g.stageCells = bestChunk;
                // staging of the chosen chunk only, allocated here: a graph capture of the tuned launch
                // must not allocate
                const long long keep = MaxChunkCells(b, g);
                if (g.ctx->keStage.bytes > (size_t)keep * nd * nd * sizeof(double))
                {
                    Check(hipStreamSynchronize(b->stream), "staging release");
                    g.ctx->keStage.Release();
                    g.ctx->reStage.Release();
                }
                EnsureStage(g.ctx, nd, keep);
            }
        }
        double bestMs = 1e300;
        int bestMode = 3;
        for (int mode = 0; mode < kBatchModes; mode++)
        {
            double v = -1.0;
            const bool usable = !(IsGdm(b) && (mode < 3 || mode >= 5)) && (mode < 5 || staged);
            if (usable)
            {
                v = median([&] { LaunchBatch(b, mode, u, e, kappa, R, K); });
                if (mode != 4 && mode != 6 && v < bestMs)  // per-mesh launches are the reference, not candidates
                {
                    bestMs = v;
                    bestMode = mode;
                }
            }
            if (ms)
                ms[mode] = v;
        }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        b->tuned = bestMode;
        DropGraph(b);
        if (best)
            *best = bestMode;
    });
}

int nuto_batch_vjp(nuto_batch* b, const double* u, const double* e, const double* kappa, const double* lambda,
                   double* d_state, double* d_theta)
{
    return GuardB(b, [&] {
        RequireBuilt(b, u, e, kappa);
        Require(IsGdm(b), NUTO_ERR_UNSUPPORTED, "the adjoint is provided for the gradient-damage integrand");
        Require(lambda && d_state, NUTO_ERR_ARGUMENT, "lambda / d_state is NULL");
        LaunchVjp(b, u, e, kappa, lambda, d_state, d_theta);
    });
}

int nuto_batch_residual_adjoint(nuto_batch* b, const double* u, const double* e, const double* kappa,
                                const double* r_obs, double* R, double* K, double* d_state, double* d_theta)
{
    return GuardB(b, [&] {
        RequireBuilt(b, u, e, kappa);
        Require(IsGdm(b), NUTO_ERR_UNSUPPORTED, "the adjoint is provided for the gradient-damage integrand");
        Require(r_obs && R && K && d_state, NUTO_ERR_ARGUMENT, "NULL array");
        if (b->lambda.bytes != (size_t)b->RLen * sizeof(double))
        {
            DropGraph(b);
            Check(b->lambda.Alloc((size_t)b->RLen * sizeof(double)), "lambda");
        }
        double* lam = b->lambda.As<double>();
        const void* p[9] = {u, e, kappa, r_obs, R, K, d_state, d_theta, lam};
        GraphRun(b, 100, p, 9, [&] {
            LaunchBatch(b, 3, u, e, kappa, R, K);
            if (b->RLen > 0)
                hipLaunchKernelGGL(kResidualMinus, dim3((unsigned)((b->RLen + 255) / 256)), dim3(256), 0, b->stream, R,
                                   r_obs, lam, b->RLen);
            LaunchVjp(b, u, e, kappa, lam, d_state, d_theta);
        });
    });
}

}  // extern "C"
