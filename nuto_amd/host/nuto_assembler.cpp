// C++ host mirror of SimpleAssembler over the nuto_amd C ABI (see nuto_assembler.h).
#include "nuto_assembler.h"

#include <algorithm>
#include <atomic>
#include <memory>

namespace nuto_amd
{
namespace host
{

namespace
{
std::atomic<int> gNextDofTypeId{0};  // UniqueId<DofType> (nuto/base/UniqueId.h:9-29): creation order

// device buffer owned through the C ABI allocator
struct DeviceArray
{
    nuto_ctx* ctx = nullptr;
    void* p = nullptr;
    DeviceArray(nuto_ctx* c, size_t bytes) : ctx(c)
    {
        if (nuto_device_alloc(ctx, &p, bytes) != NUTO_OK)
            throw Exception("nuto_amd::host", std::string("device allocation: ") + nuto_last_error(ctx));
    }
    ~DeviceArray()
    {
        if (p)
            nuto_device_free(ctx, p);
    }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    double* d() const { return static_cast<double*>(p); }
};
// GradientDamage<3, TInteraction> as the C ABI's parameter block
nuto_params GdmParams(const GradientDamage& f)
{
    nuto_params p{NUTO_GRADIENT_DAMAGE, 3, f.E, f.nu, 0, f.k, f.kappa0, f.beta, f.alpha, f.c, f.damageApplication,
                  f.interaction, f.interactionR, f.interactionEta};
    return p;
}

// The integration rule the group's cells use, as NuTo's tests construct it for each element pair:
// tetrahedra order 1, order 2 when an interpolation is the ten-node one; tensor-product Gauss with
// 2 points per direction, 3 when an interpolation is the twenty-node brick. Returns the IPs per cell.
int DefaultRule(const CellGroup& cells, const std::vector<DofType>& dofs, int& kind, int& n)
{
    bool quadratic = cells.coordType == NUTO_TET10;
    for (const DofType& d : dofs)
    {
        auto it = cells.dofNumbering.find(d.Id());
        if (it != cells.dofNumbering.end() && (it->second.first == NUTO_TET10 || it->second.first == NUTO_HEX20))
            quadratic = true;
    }
    const bool tet = cells.coordType == NUTO_TET4 || cells.coordType == NUTO_TET10;
    kind = tet ? 1 : 0;
    n = tet ? (quadratic ? 2 : 1) : (quadratic ? 3 : 2);
    if (tet)
        return quadratic ? 4 : 1;
    const int dim = cells.coordType == NUTO_QUAD4 ? 2 : 3;
    return dim == 2 ? n * n : n * n * n;
}
}  // namespace

DofType::DofType(std::string name, int num) : mName(std::move(name)), mNum(num), mId(gNextDofTypeId++)
{
}

int DofInfo::Total(const DofType& d) const
{
    auto a = numIndependent.find(d.Id());
    auto b = numDependent.find(d.Id());
    return (a == numIndependent.end() ? 0 : a->second) + (b == numDependent.end() ? 0 : b->second);
}

int CellGroup::NumCells() const
{
    const int nn = coordType == NUTO_QUAD4 ? 4 : coordType == NUTO_HEX8 ? 8 : 4;
    return conn.empty() ? 0 : (int)(conn.size() / nn);
}

SimpleAssembler::SimpleAssembler(DofInfo dofInfo, int device) : mDofInfo(std::move(dofInfo))
{
    const int s = nuto_create(&mCtx, device);
    if (s != NUTO_OK)
        throw Exception("SimpleAssembler", "no usable device (nuto_create status " + std::to_string(s) + ")");
}

SimpleAssembler::~SimpleAssembler()
{
    if (mCtx)
        nuto_destroy(mCtx);
}

void SimpleAssembler::Check(int status, const char* caller) const
{
    if (status != NUTO_OK)
        throw Exception(caller, nuto_last_error(mCtx));
}

void SimpleAssembler::Setup(const CellGroup& cells, const std::vector<DofType>& dofs, const nuto_params& p, int dim)
{
    // SimpleAssembler::ThrowOnZeroDofNumbering (SimpleAssembler.cpp:16-23)
    for (const DofType& d : dofs)
        if (mDofInfo.Total(d) == 0)
            throw Exception("SimpleAssembler", "You did not provide a dof numbering for DofType " + d.GetName() +
                                                   ". Please call SetDofInfo(...) first.");
    const int nCells = cells.NumCells();
    Check(nuto_set_cells(mCtx, cells.coordType, nCells, (int)(cells.coords.size() / dim), cells.coords.data(),
                         cells.conn.data(), cells.cellIds.empty() ? nullptr : cells.cellIds.data()),
          "SimpleAssembler::SetCells");
    for (size_t s = 0; s < dofs.size(); s++)
    {
        auto it = cells.dofNumbering.find(dofs[s].Id());
        if (it == cells.dofNumbering.end())
            throw Exception("SimpleAssembler", "the cell group has no numbering for DofType " + dofs[s].GetName());
        Check(nuto_set_dof_type(mCtx, (int)s, it->second.first, dofs[s].GetNum(), mDofInfo.Total(dofs[s]),
                                it->second.second.data()),
              "SimpleAssembler::SetDofType");
    }
    Check(nuto_set_integrand(mCtx, &p), "SimpleAssembler::SetIntegrand");
    int kind = 0, n = 0;
    mNip = DefaultRule(cells, dofs, kind, n);
    Check(nuto_set_rule(mCtx, kind, n), "SimpleAssembler::SetRule");
    Check(nuto_build_pattern(mCtx, &mNnz, &mRows), "SimpleAssembler::BuildPattern");
}

void SimpleAssembler::Run(const std::vector<DofType>& dofs, const DofVector<double>& state,
                          const std::vector<double>* kappas, bool wantR, bool wantK, DofVector<double>* R,
                          DofMatrixSparse<double>* K)
{
    std::vector<int> n;
    for (const DofType& d : dofs)
        n.push_back(mDofInfo.Total(d));
    DeviceArray u(mCtx, n[0] * sizeof(double));
    Check(nuto_copy_to_device(mCtx, u.p, state[dofs[0]].data(), n[0] * sizeof(double)), "BuildVector");
    std::unique_ptr<DeviceArray> e, k;
    if (dofs.size() > 1)
    {
        e.reset(new DeviceArray(mCtx, n[1] * sizeof(double)));
        Check(nuto_copy_to_device(mCtx, e->p, state[dofs[1]].data(), n[1] * sizeof(double)), "BuildVector");
    }
    if (kappas)
    {
        k.reset(new DeviceArray(mCtx, kappas->size() * sizeof(double)));
        Check(nuto_copy_to_device(mCtx, k->p, kappas->data(), kappas->size() * sizeof(double)), "BuildVector");
    }
    DeviceArray r(mCtx, (wantR ? mRows : 1) * sizeof(double));
    DeviceArray kv(mCtx, (wantK ? mNnz : 1) * sizeof(double));
    const double* ed = e ? e->d() : nullptr;
    const double* kd = k ? k->d() : nullptr;
    if (wantR && wantK)
        Check(nuto_build_vector_matrix(mCtx, u.d(), ed, kd, r.d(), kv.d()), "SimpleAssembler::BuildVectorMatrix");
    else if (wantR)
        Check(nuto_build_vector(mCtx, u.d(), ed, kd, r.d()), "SimpleAssembler::BuildVector");
    else
        Check(nuto_build_matrix(mCtx, u.d(), ed, kd, kv.d()), "SimpleAssembler::BuildMatrix");
    Check(nuto_synchronize(mCtx), "SimpleAssembler");
    if (wantR)
    {
        std::vector<double> all(mRows);
        Check(nuto_copy_to_host(mCtx, all.data(), r.p, mRows * sizeof(double)), "BuildVector");
        size_t off = 0;
        for (size_t s = 0; s < dofs.size(); s++)
        {
            (*R)[dofs[s]].assign(all.begin() + off, all.begin() + off + n[s]);
            off += n[s];
        }
    }
    if (wantK)
    {
        // BuildMatrix fills one sparse matrix per (dofI, dofJ) (SimpleAssembler.cpp:105-155): the
        // blocks of the device's monolithic matrix
        for (size_t si = 0; si < dofs.size(); si++)
            for (size_t sj = 0; sj < dofs.size(); sj++)
            {
                long long nb = 0;
                Check(nuto_block_pattern(mCtx, (int)si, (int)sj, &nb, nullptr, nullptr, nullptr), "BuildMatrix");
                SparseMatrix<double>& B = (*K)(dofs[si], dofs[sj]);
                B = SparseMatrix<double>(n[si], n[sj]);
                B.col.resize(nb);
                B.val.resize(nb);
                Check(nuto_block_pattern(mCtx, (int)si, (int)sj, &nb, B.rowPtr.data(), B.col.data(), nullptr),
                      "BuildMatrix");
                DeviceArray bv(mCtx, (size_t)std::max(nb, 1LL) * sizeof(double));
                Check(nuto_block_values(mCtx, (int)si, (int)sj, kv.d(), bv.d()), "BuildMatrix");
                Check(nuto_synchronize(mCtx), "BuildMatrix");
                Check(nuto_copy_to_host(mCtx, B.val.data(), bv.p, (size_t)nb * sizeof(double)), "BuildMatrix");
            }
    }
}

DofVector<double> SimpleAssembler::BuildVector(const CellGroup& cells, const GradientDamage& f,
                                               const DofVector<double>& state)
{
    const nuto_params p = GdmParams(f);
    Setup(cells, {f.disp, f.eeq}, p, 3);
    DofVector<double> R;
    Run({f.disp, f.eeq}, state, &f.kappas, true, false, &R, nullptr);
    return R;
}

DofMatrixSparse<double> SimpleAssembler::BuildMatrix(const CellGroup& cells, const GradientDamage& f,
                                                     const DofVector<double>& state)
{
    const nuto_params p = GdmParams(f);
    Setup(cells, {f.disp, f.eeq}, p, 3);
    DofMatrixSparse<double> K;
    Run({f.disp, f.eeq}, state, &f.kappas, false, true, nullptr, &K);
    return K;
}

DofVector<double> SimpleAssembler::BuildVector(const CellGroup& cells, const MomentumBalance& f,
                                               const DofVector<double>& state)
{
    const int dim = cells.coordType == NUTO_QUAD4 ? 2 : 3;
    const nuto_params p{NUTO_MOMENTUM_LINEAR_ELASTIC, dim, f.E, f.nu, f.planeState, 0, 0, 0, 0, 0, 0};
    Setup(cells, {f.disp}, p, dim);
    DofVector<double> R;
    Run({f.disp}, state, nullptr, true, false, &R, nullptr);
    return R;
}

DofMatrixSparse<double> SimpleAssembler::BuildMatrix(const CellGroup& cells, const MomentumBalance& f,
                                                     const DofVector<double>& state)
{
    const int dim = cells.coordType == NUTO_QUAD4 ? 2 : 3;
    const nuto_params p{NUTO_MOMENTUM_LINEAR_ELASTIC, dim, f.E, f.nu, f.planeState, 0, 0, 0, 0, 0, 0};
    Setup(cells, {f.disp}, p, dim);
    DofMatrixSparse<double> K;
    Run({f.disp}, state, nullptr, false, true, nullptr, &K);
    return K;
}

DofVector<double> SimpleAssembler::BuildDiagonallyLumpedMatrix(const CellGroup& cells,
                                                               const DynamicMomentumBalance& f)
{
    const int dim = cells.coordType == NUTO_QUAD4 ? 2 : 3;
    const nuto_params p{NUTO_MOMENTUM_LINEAR_ELASTIC, dim, f.E, f.nu, f.planeState, 0, 0, 0, 0, 0, 0};
    Setup(cells, {f.disp}, p, dim);
    const int n = mDofInfo.Total(f.disp);
    DeviceArray m(mCtx, n * sizeof(double));
    Check(nuto_build_lumped_mass(mCtx, f.rho, m.d()), "SimpleAssembler::BuildDiagonallyLumpedMatrix");
    Check(nuto_synchronize(mCtx), "SimpleAssembler::BuildDiagonallyLumpedMatrix");
    DofVector<double> M;
    M[f.disp].resize(n);
    Check(nuto_copy_to_host(mCtx, M[f.disp].data(), m.p, n * sizeof(double)),
          "SimpleAssembler::BuildDiagonallyLumpedMatrix");
    return M;
}

std::vector<double> SimpleAssembler::RunEval(const std::vector<DofType>& dofs, const DofVector<double>& state,
                                             const std::vector<double>* kappas, int nCells, int dim, IpField field)
{
    std::vector<int> n;
    for (const DofType& d : dofs)
        n.push_back(mDofInfo.Total(d));
    DeviceArray u(mCtx, n[0] * sizeof(double));
    Check(nuto_copy_to_device(mCtx, u.p, state[dofs[0]].data(), n[0] * sizeof(double)), "Eval");
    std::unique_ptr<DeviceArray> e, k;
    if (dofs.size() > 1)
    {
        e.reset(new DeviceArray(mCtx, n[1] * sizeof(double)));
        Check(nuto_copy_to_device(mCtx, e->p, state[dofs[1]].data(), n[1] * sizeof(double)), "Eval");
    }
    if (kappas)
    {
        k.reset(new DeviceArray(mCtx, kappas->size() * sizeof(double)));
        Check(nuto_copy_to_device(mCtx, k->p, kappas->data(), kappas->size() * sizeof(double)), "Eval");
    }
    const int nip = mNip;  // IPs of the rule Setup chose
    const int nc = (field == IpField::Strain || field == IpField::Stress) ? (dim == 2 ? 3 : 6) : 1;
    const size_t count = (size_t)nCells * nip * nc;
    DeviceArray o(mCtx, count * sizeof(double));
    Check(nuto_eval_ip(mCtx, u.d(), e ? e->d() : nullptr, k ? k->d() : nullptr, (int)field, o.d()),
          "SimpleAssembler::Eval");
    Check(nuto_synchronize(mCtx), "SimpleAssembler::Eval");
    std::vector<double> out(count);
    Check(nuto_copy_to_host(mCtx, out.data(), o.p, count * sizeof(double)), "SimpleAssembler::Eval");
    return out;
}

std::vector<double> SimpleAssembler::Eval(const CellGroup& cells, const GradientDamage& f,
                                          const DofVector<double>& state, IpField field)
{
    const nuto_params p = GdmParams(f);
    Setup(cells, {f.disp, f.eeq}, p, 3);
    return RunEval({f.disp, f.eeq}, state, &f.kappas, cells.NumCells(), 3, field);
}

void SimpleAssembler::UpdateHistory(const CellGroup& cells, GradientDamage& f, const DofVector<double>& state)
{
    const nuto_params p = GdmParams(f);
    Setup(cells, {f.disp, f.eeq}, p, 3);
    const int nu = mDofInfo.Total(f.disp), ne = mDofInfo.Total(f.eeq);
    const size_t nk = (size_t)cells.NumCells() * mNip;
    if (f.kappas.size() != nk)
        throw Exception("SimpleAssembler::UpdateHistory", "the history holds " + std::to_string(f.kappas.size()) +
                                                              " values, the cell group needs " + std::to_string(nk));
    DeviceArray u(mCtx, nu * sizeof(double)), e(mCtx, ne * sizeof(double)), k(mCtx, nk * sizeof(double));
    Check(nuto_copy_to_device(mCtx, u.p, state[f.disp].data(), nu * sizeof(double)), "UpdateHistory");
    Check(nuto_copy_to_device(mCtx, e.p, state[f.eeq].data(), ne * sizeof(double)), "UpdateHistory");
    Check(nuto_copy_to_device(mCtx, k.p, f.kappas.data(), nk * sizeof(double)), "UpdateHistory");
    Check(nuto_update_history(mCtx, u.d(), e.d(), k.d()), "SimpleAssembler::UpdateHistory");
    Check(nuto_synchronize(mCtx), "SimpleAssembler::UpdateHistory");
    Check(nuto_copy_to_host(mCtx, f.kappas.data(), k.p, nk * sizeof(double)), "UpdateHistory");
}

std::vector<double> SimpleAssembler::Eval(const CellGroup& cells, const MomentumBalance& f,
                                          const DofVector<double>& state, IpField field)
{
    const int dim = cells.coordType == NUTO_QUAD4 ? 2 : 3;
    const nuto_params p{NUTO_MOMENTUM_LINEAR_ELASTIC, dim, f.E, f.nu, f.planeState, 0, 0, 0, 0, 0, 0};
    Setup(cells, {f.disp}, p, dim);
    return RunEval({f.disp}, state, nullptr, cells.NumCells(), dim, field);
}

// ---- Constraints, QuasistaticSolver ----------------------------------------------------------------
const std::vector<Equation>& Constraints::Get(const DofType& dof) const
{
    static const std::vector<Equation> none;
    auto it = mEquations.find(dof.Id());
    return it == mEquations.end() ? none : it->second;
}

void QuasistaticSolver::Check(int status, const char* caller) const
{
    if (status == NUTO_ERR_NO_CONVERGENCE)
        throw NoConvergence(caller, nuto_last_error(mCtx));
    if (status != NUTO_OK)
        throw Exception(caller, nuto_last_error(mCtx));
}

void QuasistaticSolver::Init(const CellGroup& cells, const std::vector<DofType>& dofs, const nuto_params& p, int dim,
                             int device)
{
    if (nuto_create(&mCtx, device) != NUTO_OK)
        throw Exception("QuasistaticSolver", "could not create a device context");
    mDofs = dofs;
    for (const DofType& d : dofs)
        if (mDofInfo.Total(d) == 0)
            throw Exception("QuasistaticSolver", "You did not provide a dof numbering for DofType " + d.GetName());
    const int nCells = cells.NumCells();
    Check(nuto_set_cells(mCtx, cells.coordType, nCells, (int)(cells.coords.size() / dim), cells.coords.data(),
                         cells.conn.data(), cells.cellIds.empty() ? nullptr : cells.cellIds.data()),
          "QuasistaticSolver::SetCells");
    for (size_t s = 0; s < dofs.size(); s++)
    {
        auto it = cells.dofNumbering.find(dofs[s].Id());
        if (it == cells.dofNumbering.end())
            throw Exception("QuasistaticSolver", "the cell group has no numbering for DofType " + dofs[s].GetName());
        Check(nuto_set_dof_type(mCtx, (int)s, it->second.first, dofs[s].GetNum(), mDofInfo.Total(dofs[s]),
                                it->second.second.data()),
              "QuasistaticSolver::SetDofType");
    }
    Check(nuto_set_integrand(mCtx, &p), "QuasistaticSolver::SetIntegrand");
    int kind = 0, n = 0;
    mNip = DefaultRule(cells, dofs, kind, n);
    Check(nuto_set_rule(mCtx, kind, n), "QuasistaticSolver::SetRule");
    Check(nuto_build_pattern(mCtx, &mNnz, &mRows), "QuasistaticSolver::BuildPattern");
    Check(nuto_device_alloc(mCtx, &mX, (size_t)mRows * sizeof(double)), "QuasistaticSolver");
    std::vector<double> zero(mRows, 0.0);
    Check(nuto_copy_to_device(mCtx, mX, zero.data(), zero.size() * sizeof(double)), "QuasistaticSolver");
}

QuasistaticSolver::QuasistaticSolver(const CellGroup& cells, GradientDamage& f, DofInfo dofInfo, int device)
    : mDofInfo(std::move(dofInfo))
    , mGdm(&f)
{
    const nuto_params p = GdmParams(f);
    Init(cells, {f.disp, f.eeq}, p, 3, device);
    const size_t nk = (size_t)cells.NumCells() * mNip;
    if (f.kappas.size() != nk)
        f.kappas.assign(nk, 0.0);
    mKappaBytes = nk * sizeof(double);
    Check(nuto_device_alloc(mCtx, &mKappa, mKappaBytes), "QuasistaticSolver");
    Check(nuto_copy_to_device(mCtx, mKappa, f.kappas.data(), mKappaBytes), "QuasistaticSolver");
}

QuasistaticSolver::QuasistaticSolver(const CellGroup& cells, const MomentumBalance& f, DofInfo dofInfo, int device)
    : mDofInfo(std::move(dofInfo))
{
    const int dim = cells.coordType == NUTO_QUAD4 ? 2 : 3;
    const nuto_params p{NUTO_MOMENTUM_LINEAR_ELASTIC, dim, f.E, f.nu, f.planeState, 0, 0, 0, 0, 0, 0};
    Init(cells, {f.disp}, p, dim, device);
}

QuasistaticSolver::~QuasistaticSolver()
{
    if (mCtx)
    {
        if (mX)
            nuto_device_free(mCtx, mX);
        if (mKappa)
            nuto_device_free(mCtx, mKappa);
        nuto_destroy(mCtx);
    }
}

void QuasistaticSolver::SetConstraints(Constraints constraints)
{
    mConstraints = std::move(constraints);
    for (size_t s = 0; s < mDofs.size(); s++)
    {
        const std::vector<Equation>& eqs = mConstraints.Get(mDofs[s]);
        std::vector<int> dep, ptr{0}, dof;
        std::vector<double> coef;
        for (const Equation& e : eqs)
        {
            dep.push_back(e.dependentDof);
            for (const auto& t : e.terms)
            {
                dof.push_back(t.first);
                coef.push_back(t.second);
            }
            ptr.push_back((int)dof.size());
        }
        Check(nuto_set_constraints(mCtx, (int)s, (int)dep.size(), dep.data(), ptr.data(), dof.data(), coef.data()),
              "QuasistaticSolver::SetConstraints");
    }
}

std::vector<double> QuasistaticSolver::Rhs(double t) const
{
    std::vector<double> r;
    for (const DofType& d : mDofs)
        for (const Equation& e : mConstraints.Get(d))
            r.push_back(e.rhs(t));
    return r;
}

int QuasistaticSolver::DoStep(double newGlobalTime, std::string solverType)
{
    const std::vector<double> ro = Rhs(mGlobalTime), rn = Rhs(newGlobalTime);
    void *dro = nullptr, *drn = nullptr;
    const size_t nb = std::max<size_t>(ro.size(), 1) * sizeof(double);
    Check(nuto_device_alloc(mCtx, &dro, nb), "QuasistaticSolver::DoStep");
    Check(nuto_device_alloc(mCtx, &drn, nb), "QuasistaticSolver::DoStep");
    nuto_newton_info info{};
    int status = NUTO_OK;
    if (!ro.empty())
    {
        status = nuto_copy_to_device(mCtx, dro, ro.data(), ro.size() * sizeof(double));
        if (status == NUTO_OK)
            status = nuto_copy_to_device(mCtx, drn, rn.data(), rn.size() * sizeof(double));
    }
    if (status == NUTO_OK)
    {
        const int method = solverType == "EigenConjugateGradient" ? NUTO_SOLVER_CG : NUTO_SOLVER_BICGSTAB;
        const nuto_newton_options o{mTolerance, 6, 6, {method, mLinearTolerance, 0}};
        status = nuto_quasistatic_step(mCtx, static_cast<double*>(mX), static_cast<double*>(mKappa),
                                       static_cast<const double*>(dro), static_cast<const double*>(drn), &o, &info);
    }
    const std::string msg = status == NUTO_OK ? "" : nuto_last_error(mCtx);
    nuto_device_free(mCtx, dro);
    nuto_device_free(mCtx, drn);
    if (status == NUTO_ERR_NO_CONVERGENCE)
        throw NoConvergence("QuasistaticSolver::DoStep", msg);
    if (status != NUTO_OK)
        throw Exception("QuasistaticSolver::DoStep", msg);
    mGlobalTime = newGlobalTime;
    if (mGdm)
        Check(nuto_copy_to_host(mCtx, mGdm->kappas.data(), mKappa, mKappaBytes), "QuasistaticSolver::DoStep");
    return info.iterations;
}

DofVector<double> QuasistaticSolver::State() const
{
    std::vector<double> all(mRows);
    Check(nuto_copy_to_host(mCtx, all.data(), mX, (size_t)mRows * sizeof(double)), "QuasistaticSolver::State");
    DofVector<double> x;
    size_t off = 0;
    for (const DofType& d : mDofs)
    {
        const size_t n = mDofInfo.Total(d);
        x[d].assign(all.begin() + off, all.begin() + off + n);
        off += n;
    }
    return x;
}

// ---- MeshGmsh ----------------------------------------------------------------------------------
MeshGmsh::MeshGmsh(const std::string& fileName)
{
    const int r = nuto_gmsh_read(fileName.c_str(), &mMesh);
    if (r != NUTO_OK)
    {
        const std::string msg = mMesh ? nuto_gmsh_last_error(mMesh) : "out of memory";
        nuto_gmsh_free(mMesh);
        mMesh = nullptr;
        throw Exception("MeshGmsh::MeshGmsh", msg);
    }
    int nNodes = 0, nElements = 0, connSize = 0;
    nuto_gmsh_sizes(mMesh, &mDim, &nNodes, &nElements, &connSize);
    mCoords.resize((size_t)nNodes * mDim);
    mType.resize(nElements);
    mOffsets.resize(nElements + 1);
    mConn.resize(connSize);
    nuto_gmsh_nodes(mMesh, mCoords.data());
    nuto_gmsh_elements(mMesh, mType.data(), mOffsets.data(), mConn.data(), nullptr);
}

MeshGmsh::~MeshGmsh()
{
    nuto_gmsh_free(mMesh);
}

std::vector<int> MeshGmsh::GetPhysicalGroup(int physicalGroupId) const
{
    int n = 0;
    if (nuto_gmsh_group(mMesh, physicalGroupId, &n, nullptr) != NUTO_OK)
        throw Exception("MeshGmsh::GetPhysicalGroup", nuto_gmsh_last_error(mMesh));
    std::vector<int> elements(n);
    nuto_gmsh_group(mMesh, physicalGroupId, &n, elements.data());
    return elements;
}

std::vector<int> MeshGmsh::GetPhysicalGroup(std::string physicalName) const
{
    int id = 0;
    if (nuto_gmsh_group_id(mMesh, physicalName.c_str(), &id) != NUTO_OK)
        throw Exception("MeshGmsh::GetPhysicalGroup", nuto_gmsh_last_error(mMesh));
    return GetPhysicalGroup(id);
}

std::vector<int> MeshGmsh::ElementNodes(int element) const
{
    if (element < 0 || element >= NumElements())
        throw Exception("MeshGmsh::ElementNodes", "element index out of range");
    return std::vector<int>(mConn.begin() + mOffsets[element], mConn.begin() + mOffsets[element + 1]);
}

CellGroup MeshGmsh::Cells(int coordType, int physicalGroupId) const
{
    int n = 0;
    if (nuto_gmsh_cells(mMesh, physicalGroupId, coordType, &n, nullptr) != NUTO_OK)
        throw Exception("MeshGmsh::Cells", nuto_gmsh_last_error(mMesh));
    const int nodes[] = {4, 8, 4, 10, 20};
    CellGroup g;
    g.coordType = coordType;
    g.coords = mCoords;
    g.conn.resize((size_t)n * nodes[coordType]);
    nuto_gmsh_cells(mMesh, physicalGroupId, coordType, &n, g.conn.data());
    return g;
}

void TimeDependentProblem::AddGradientFunction(const CellGroup& group, const GradientDamage& f)
{
    mGradient.push_back({&group, &f, nullptr, nullptr});
}

void TimeDependentProblem::AddGradientFunction(const CellGroup& group, const MomentumBalance& f)
{
    mGradient.push_back({&group, nullptr, &f, nullptr});
}

void TimeDependentProblem::AddHessian0Function(const CellGroup& group, const GradientDamage& f)
{
    mHessian0.push_back({&group, &f, nullptr, nullptr});
}

void TimeDependentProblem::AddHessian0Function(const CellGroup& group, const MomentumBalance& f)
{
    mHessian0.push_back({&group, nullptr, &f, nullptr});
}

void TimeDependentProblem::AddUpdateFunction(const CellGroup& group, GradientDamage& f)
{
    mUpdate.push_back({&group, &f, nullptr, &f});
}

void TimeDependentProblem::CheckDofs(const Term& term, const std::vector<DofType>& dofs)
{
    const bool ok = term.gdm ? dofs.size() == 2 && dofs[0].Id() == term.gdm->disp.Id() &&
                                   dofs[1].Id() == term.gdm->eeq.Id()
                             : !dofs.empty() && dofs[0].Id() == term.mb->disp.Id();
    if (!ok)
        throw Exception("TimeDependentProblem",
                        "the integrand's dof types are not a leading part of the requested dof types");
}

DofVector<double> TimeDependentProblem::Gradient(const DofVector<double>& dofValues, std::vector<DofType> dofs,
                                                 double /*t*/, double /*dt*/)
{
    DofVector<double> gradient;
    for (const DofType& d : dofs)
        gradient[d].assign(mDofInfo.Total(d), 0.0);
    for (const Term& term : mGradient)
    {
        CheckDofs(term, dofs);
        const DofVector<double> r = term.gdm ? mAssembler.BuildVector(*term.group, *term.gdm, dofValues)
                                             : mAssembler.BuildVector(*term.group, *term.mb, dofValues);
        for (const DofType& d : dofs)
        {
            if (!r.Has(d))
                continue;
            const std::vector<double>& v = r[d];
            std::vector<double>& g = gradient[d];
            for (size_t i = 0; i < v.size(); i++)
                g[i] += v[i];
        }
    }
    return gradient;
}

DofMatrixSparse<double> TimeDependentProblem::Hessian0(const DofVector<double>& dofValues,
                                                       std::vector<DofType> dofs, double /*t*/, double /*dt*/)
{
    DofMatrixSparse<double> hessian0;
    for (const Term& term : mHessian0)
    {
        CheckDofs(term, dofs);
        hessian0 += term.gdm ? mAssembler.BuildMatrix(*term.group, *term.gdm, dofValues)
                             : mAssembler.BuildMatrix(*term.group, *term.mb, dofValues);
    }
    return hessian0;
}

void TimeDependentProblem::UpdateHistory(const DofVector<double>& dofValues, std::vector<DofType> dofs,
                                         double /*t*/, double /*dt*/)
{
    for (const Term& term : mUpdate)
    {
        CheckDofs(term, dofs);
        mAssembler.UpdateHistory(*term.group, *term.update, dofValues);
    }
}

}  // namespace host
}  // namespace nuto_amd
