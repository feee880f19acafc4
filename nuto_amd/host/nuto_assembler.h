// C++ host mirror of NuTo's SimpleAssembler interface over the nuto_amd C ABI (no Eigen, no Boost).
//
// Types mirror the reference's (names, argument meaning, error behaviour):
//   Exception          NuTo::Exception (nuto/base/Exception.h:9-26), thrown for every failing call
//   DofType            NuTo::DofType (nuto/mechanics/dofs/DofType.h:8-50): name, components, unique id
//   DofVector<T>       NuTo::DofVector (DofVector.h:11-149): one value vector per DofType
//   SparseMatrix<T>    the blocks' Eigen::SparseMatrix<T> (a CSR here: columns ascending per row,
//                      explicit zeros kept as setFromTriplets keeps them)
//   DofMatrixSparse<T> NuTo::DofMatrixSparse (DofMatrixSparse.h:9-10): one SparseMatrix per dof type
//                      pair, K(dofI, dofJ) (DofMatrixContainer.h:16-24); ToEigen(K, dofs) exports the
//                      blocks of any dof list monolithically (DofMatrixSparseConvertEigen.h:40-67)
//   DofInfo            NuTo::DofInfo (DofInfo.h:7-17): numIndependent / numDependent per DofType
//   SimpleAssembler    BuildVector / BuildMatrix (SimpleAssembler.h:14-21, SimpleAssembler.cpp:38-156)
//   Constraints        Constraint::Constraints / Equation (Constraints.h:47-130, Equation.h:12-66): per DofType,
//                      1 x_dep + sum c x = rhs(t)
//   TimeDependentProblem  Gradient / Hessian0 / UpdateHistory over (group, integrand) pairs
//                      (TimeDependentProblem.cpp:34-85)
//   QuasistaticSolver  QuasistaticSolver::DoStep (QuasistaticSolver.h:14-90) on the device
//   MeshGmsh           MeshGmsh (MeshGmsh.h:16-60)
// NuTo's VectorFunction / MatrixFunction are opaque std::functions; here the integrand is a typed
// descriptor (GradientDamage, MomentumBalance), and the cell group is the mesh itself (CellGroup).
// The node values that NuTo reads from its nodes (NodalValueMerger) are passed as a DofVector.
#pragma once

#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nuto_amd.h"

namespace nuto_amd
{
namespace host
{

class Exception : public std::runtime_error
{
public:
    Exception(const std::string& caller, const std::string& message) : std::runtime_error(caller + ":\n" + message)
    {
    }
};

class DofType
{
public:
    DofType(std::string name, int num);
    const std::string& GetName() const { return mName; }
    int GetNum() const { return mNum; }
    int Id() const { return mId; }

private:
    std::string mName;
    int mNum;
    int mId;
};

template <typename T>
class DofVector
{
public:
    std::vector<T>& operator[](const DofType& d) { return mValues[d.Id()]; }
    const std::vector<T>& operator[](const DofType& d) const;
    bool Has(const DofType& d) const { return mValues.count(d.Id()) != 0; }

private:
    std::map<int, std::vector<T>> mValues;
};

template <typename T>
const std::vector<T>& DofVector<T>::operator[](const DofType& d) const
{
    auto it = mValues.find(d.Id());
    if (it == mValues.end())
        throw std::out_of_range("DofVector: no values for dof type " + d.GetName());  // std::map::at in NuTo
    return it->second;
}

// One block: rows x cols, CSR with columns ascending within each row; explicit zeros are entries.
template <typename T>
struct SparseMatrix
{
    int rows = 0, cols = 0;
    std::vector<int> rowPtr{0}, col;
    std::vector<T> val;
    SparseMatrix() = default;
    SparseMatrix(int r, int c)
        : rows(r)
        , cols(c)
        , rowPtr(r + 1, 0)
    {
    }
    int Rows() const { return rows; }
    int Cols() const { return cols; }
    long long NonZeros() const { return (long long)val.size(); }
    T Coeff(int row, int column) const;  // 0 outside the pattern
    T& CoeffRef(int row, int column);    // inserts a zero entry when absent, as Eigen's coeffRef does
    SparseMatrix& operator+=(const SparseMatrix& b);  // Eigen's sparse sum: the union of both patterns
};

template <typename T>
T SparseMatrix<T>::Coeff(int row, int column) const
{
    for (int k = rowPtr[row]; k < rowPtr[row + 1]; k++)
        if (col[k] == column)
            return val[k];
    return T(0);
}

template <typename T>
T& SparseMatrix<T>::CoeffRef(int row, int column)
{
    if (row < 0 || row >= rows || column < 0 || column >= cols)
        throw std::out_of_range("SparseMatrix::CoeffRef: index outside the matrix");
    int k = rowPtr[row];
    while (k < rowPtr[row + 1] && col[k] < column)
        k++;
    if (k < rowPtr[row + 1] && col[k] == column)
        return val[k];
    col.insert(col.begin() + k, column);
    val.insert(val.begin() + k, T(0));
    for (int r = row + 1; r <= rows; r++)
        rowPtr[r]++;
    return val[k];
}

template <typename T>
SparseMatrix<T>& SparseMatrix<T>::operator+=(const SparseMatrix<T>& b)
{
    if (b.rows != rows || b.cols != cols)
        throw std::invalid_argument("SparseMatrix +=: the blocks differ in size");
    SparseMatrix<T> s(rows, cols);
    s.col.reserve(col.size() + b.col.size());
    s.val.reserve(col.size() + b.col.size());
    for (int i = 0; i < rows; i++)
    {
        int p = rowPtr[i], pe = rowPtr[i + 1], q = b.rowPtr[i], qe = b.rowPtr[i + 1];
        while (p < pe || q < qe)  // both rows have ascending columns
        {
            if (q == qe || (p < pe && col[p] < b.col[q]))
            {
                s.col.push_back(col[p]);
                s.val.push_back(val[p++]);
            }
            else if (p == pe || b.col[q] < col[p])
            {
                s.col.push_back(b.col[q]);
                s.val.push_back(b.val[q++]);
            }
            else
            {
                s.col.push_back(col[p]);
                s.val.push_back(val[p++] + b.val[q++]);
            }
        }
        s.rowPtr[i + 1] = (int)s.col.size();
    }
    *this = std::move(s);
    return *this;
}

// DofMatrixContainer<SparseMatrix<T>> (DofMatrixContainer.h:12-122): K(dofI, dofJ) creates the block
// when absent (std::map::operator[]); the const access throws std::out_of_range (std::map::at).
template <typename T>
class DofMatrixSparse
{
public:
    SparseMatrix<T>& operator()(const DofType& d0, const DofType& d1) { return mData[{d0.Id(), d1.Id()}]; }
    const SparseMatrix<T>& operator()(const DofType& d0, const DofType& d1) const
    {
        auto it = mData.find({d0.Id(), d1.Id()});
        if (it == mData.end())
            throw std::out_of_range("DofMatrixSparse: no block (" + d0.GetName() + ", " + d1.GetName() + ")");
        return it->second;
    }
    bool Has(const DofType& d0, const DofType& d1) const { return mData.count({d0.Id(), d1.Id()}) != 0; }
    // DofMatrixContainer::operator+= (DofMatrixContainer.h:27-37): blocks present in both are summed,
    // blocks only in rhs are copied
    DofMatrixSparse& operator+=(const DofMatrixSparse& rhs)
    {
        for (const auto& entry : rhs.mData)
        {
            auto it = mData.find(entry.first);
            if (it == mData.end())
                mData[entry.first] = entry.second;
            else
                it->second += entry.second;
        }
        return *this;
    }

private:
    std::map<std::pair<int, int>, SparseMatrix<T>> mData;
};

// TotalRows / TotalCols / TotalNonZeros (DofMatrixSparseConvertEigen.h:7-30)
template <typename T>
int TotalRows(const DofMatrixSparse<T>& v, const std::vector<DofType>& dofs)
{
    int total = 0;
    for (const DofType& d : dofs)
        total += v(d, d).Rows();
    return total;
}

template <typename T>
int TotalCols(const DofMatrixSparse<T>& v, const std::vector<DofType>& dofs)
{
    int total = 0;
    for (const DofType& d : dofs)
        total += v(d, d).Cols();
    return total;
}

template <typename T>
long long TotalNonZeros(const DofMatrixSparse<T>& v, const std::vector<DofType>& dofs)
{
    long long total = 0;
    for (const DofType& d0 : dofs)
        for (const DofType& d1 : dofs)
            total += v(d0, d1).NonZeros();
    return total;
}

// ToEigen (DofMatrixSparseConvertEigen.h:40-67): the blocks of `dofs` as one matrix, rows and columns
// in the order of `dofs` (row block of dofRow offset by the rows of the dof types before it, column
// block by the columns of the blocks before it in that row). One dof type: its block.
template <typename T>
SparseMatrix<T> ToEigen(const DofMatrixSparse<T>& v, std::vector<DofType> dofs)
{
    if (dofs.size() == 1)
        return v(dofs.front(), dofs.front());
    SparseMatrix<T> out(TotalRows(v, dofs), TotalCols(v, dofs));
    out.col.reserve((size_t)TotalNonZeros(v, dofs));
    out.val.reserve((size_t)TotalNonZeros(v, dofs));
    int startRow = 0;
    for (const DofType& dofRow : dofs)
    {
        const int nr = v(dofRow, dofRow).Rows();
        for (int i = 0; i < nr; i++)
        {
            int startCol = 0;
            for (const DofType& dofCol : dofs)
            {
                const SparseMatrix<T>& m = v(dofRow, dofCol);
                if (i < m.Rows())
                    for (int k = m.rowPtr[i]; k < m.rowPtr[i + 1]; k++)
                    {
                        out.col.push_back(m.col[k] + startCol);
                        out.val.push_back(m.val[k]);
                    }
                startCol += m.Cols();
            }
            out.rowPtr[startRow + i + 1] = (int)out.col.size();
        }
        startRow += nr;
    }
    return out;
}

struct DofInfo
{
    std::map<int, int> numIndependent, numDependent;
    int Total(const DofType& d) const;
};

// The cell group: coordinate elements and, per DofType, the cell dof numbering
// (Cell::DofNumbering, node-major, component-minor, Cell.h:60-63).
struct CellGroup
{
    int coordType = NUTO_HEX8;
    std::vector<double> coords;  // nodes x dim
    std::vector<int> conn;       // cells x nodes(coordType)
    std::vector<int> cellIds;    // Cell::Id per cell (empty: 0..n-1)
    std::map<int, std::pair<int, std::vector<int>>> dofNumbering;  // DofType id -> (interpolation, numbering)
    int NumCells() const;
};

// GradientDamage<3> with LinearElasticDamage, ModifiedMisesStrainNorm and DamageLawExponential
// (GradientDamage.h:36, LinearElasticDamage.h:32, ModifiedMisesStrainNorm.h:38, DamageLawExponential.h:26)
struct GradientDamage
{
    DofType disp, eeq;
    double c, E, nu, k, kappa0, beta, alpha;
    int damageApplication = 0;   // 0 FULL, 1 UNILATERAL
    std::vector<double> kappas;  // history, cells x IPs column-major (mKappas, GradientDamage.h:137)
    // NonlocalInteraction (GradientDamage.h:150-184): 0 Constant (the default), 1 Decreasing(R, eta)
    int interaction = 0;
    double interactionR = 0.005, interactionEta = 5.0;
};

// MomentumBalance<D> with LinearElastic<D> (MomentumBalance.h:16, LinearElastic.h:16)
struct MomentumBalance
{
    DofType disp;
    double E, nu;
    int planeState = 0;  // 2D: 0 PLANE_STRESS, 1 PLANE_STRAIN
};

// DynamicMomentumBalance<D> (DynamicMomentumBalance.h:10-37): MomentumBalance whose Hessian2 is the
// consistent mass N^T N rho
struct DynamicMomentumBalance : MomentumBalance
{
    double rho = 1.0;
};

// fields of SimpleAssembler::Eval (Cell::Eval with the integrand's quantities, Cell.h:80-93)
enum class IpField
{
    Strain = NUTO_IP_STRAIN,      // Nabla::Strain of the displacements (engineering Voigt)
    Stress = NUTO_IP_STRESS,      // the law's stress
    Kappa = NUTO_IP_KAPPA,        // GradientDamage::Kappa (GradientDamage.h:126-129)
    Damage = NUTO_IP_DAMAGE,      // DamageLawExponential::Damage(kappa)
    EqStrain = NUTO_IP_EQ_STRAIN  // ModifiedMisesStrainNorm::Value(strain)
};

// MeshGmsh (MeshGmsh.h:16-60, MeshGmsh.cpp:440-550): a gmsh MSH 2 file (ASCII or binary) read
// through nuto_gmsh_read; reading errors throw Exception with MeshGmsh's messages. Elements are
// indexed in file order (point elements dropped); GetPhysicalGroup returns element indices.
class MeshGmsh
{
public:
    explicit MeshGmsh(const std::string& fileName);
    ~MeshGmsh();
    MeshGmsh(const MeshGmsh&) = delete;
    MeshGmsh& operator=(const MeshGmsh&) = delete;

    // MeshGmsh::GetPhysicalGroup(int) / (std::string): unknown ids and names throw
    std::vector<int> GetPhysicalGroup(int physicalGroupId) const;
    std::vector<int> GetPhysicalGroup(std::string physicalName) const;

    int Dimension() const { return mDim; }
    int NumNodes() const { return (int)mCoords.size() / mDim; }
    int NumElements() const { return (int)mType.size(); }
    const std::vector<double>& Coordinates() const { return mCoords; }  // nodes x dim, file order
    int GmshType(int element) const { return mType[element]; }
    std::vector<int> ElementNodes(int element) const;  // 0-based, interpolation node order

    // the elements of coordType's gmsh type (quad4, hex8, tet4, tet10) of a group (< 0: all) as a
    // cell group: coordinates and connectivity, no dof numbering yet
    CellGroup Cells(int coordType, int physicalGroupId = -1) const;

private:
    nuto_gmsh* mMesh = nullptr;
    int mDim = 1;
    std::vector<double> mCoords;
    std::vector<int> mType, mOffsets, mConn;
};

// Constraint::Equation / Constraints (Equation.h:12-66, Constraints.h:47-130): for one DofType,
// 1 * dependentDof + sum coefficient * dof = rhs(t) (RhsRamp(slope) is rhs(t) = slope * t)
struct Equation
{
    int dependentDof = 0;
    std::vector<std::pair<int, double>> terms;
    std::function<double(double)> rhs = [](double) { return 0.0; };
};

class Constraints
{
public:
    void Add(const DofType& dof, Equation e) { mEquations[dof.Id()].push_back(std::move(e)); }
    const std::vector<Equation>& Get(const DofType& dof) const;

private:
    std::map<int, std::vector<Equation>> mEquations;
};

// NewtonRaphson::NoConvergence (NewtonRaphson.h:12-20)
class NoConvergence : public Exception
{
public:
    using Exception::Exception;
};

// QuasistaticSolver (QuasistaticSolver.h:14-90) for one cell group and integrand: the dof vector
// and the history stay on the device between steps; DoStep runs nuto_quasistatic_step (trial
// state, Newton-Raphson with line search, history update). The integrand's history
// (GradientDamage::kappas) is copied back after every converged step, as NuTo updates it in place.
class QuasistaticSolver
{
public:
    QuasistaticSolver(const CellGroup& cells, GradientDamage& f, DofInfo dofInfo, int device = 0);
    QuasistaticSolver(const CellGroup& cells, const MomentumBalance& f, DofInfo dofInfo, int device = 0);
    ~QuasistaticSolver();
    QuasistaticSolver(const QuasistaticSolver&) = delete;
    QuasistaticSolver& operator=(const QuasistaticSolver&) = delete;

    void SetConstraints(Constraints constraints);
    void SetGlobalTime(double globalTime) { mGlobalTime = globalTime; }
    double GlobalTime() const { return mGlobalTime; }
    // solverType "EigenConjugateGradient" selects conjugate gradients, anything else BiCGSTAB
    int DoStep(double newGlobalTime, std::string solverType = "EigenBiCGSTAB");
    DofVector<double> State() const;  // the converged dof values (mX)

    double mTolerance = 1.e-10;   // QuasistaticSolver::mTolerance
    double mLinearTolerance = 1.e-12;

private:
    void Check(int status, const char* caller) const;
    std::vector<double> Rhs(double t) const;
    void Init(const CellGroup& cells, const std::vector<DofType>& dofs, const nuto_params& p, int dim, int device);

    nuto_ctx* mCtx = nullptr;
    DofInfo mDofInfo;
    std::vector<DofType> mDofs;
    GradientDamage* mGdm = nullptr;
    Constraints mConstraints;
    double mGlobalTime = 0.0;
    int mRows = 0;
    long long mNnz = 0;
    void* mX = nullptr;      // device [u; e]
    void* mKappa = nullptr;  // device history
    size_t mKappaBytes = 0;
    int mNip = 0;            // IPs per cell of the group's rule
};

class SimpleAssembler
{
public:
    explicit SimpleAssembler(DofInfo dofInfo = DofInfo(), int device = 0);
    ~SimpleAssembler();
    SimpleAssembler(const SimpleAssembler&) = delete;
    SimpleAssembler& operator=(const SimpleAssembler&) = delete;

    void SetDofInfo(DofInfo dofInfo) { mDofInfo = std::move(dofInfo); }

    // SimpleAssembler::BuildVector / BuildMatrix for the given integrand and node values
    DofVector<double> BuildVector(const CellGroup& cells, const GradientDamage& f, const DofVector<double>& state);
    DofMatrixSparse<double> BuildMatrix(const CellGroup& cells, const GradientDamage& f,
                                        const DofVector<double>& state);
    DofVector<double> BuildVector(const CellGroup& cells, const MomentumBalance& f, const DofVector<double>& state);
    DofMatrixSparse<double> BuildMatrix(const CellGroup& cells, const MomentumBalance& f,
                                        const DofVector<double>& state);

    // SimpleAssembler::BuildDiagonallyLumpedMatrix (SimpleAssembler.cpp:67-98) with
    // DynamicMomentumBalance::Hessian2 (DynamicMomentumBalance.h:22-30): the displacement dofs only
    DofVector<double> BuildDiagonallyLumpedMatrix(const CellGroup& cells, const DynamicMomentumBalance& f);

    // Cell::Apply(GradientDamage::Update) over the group (TimeDependentProblem.cpp:78-85,
    // GradientDamage.h:122-125): f.kappas(cell, ip) = max(kappa, eeq(ip)), in place
    void UpdateHistory(const CellGroup& cells, GradientDamage& f, const DofVector<double>& state);

    // Cell::Eval (Cell.h:80-93) of one field for every cell of the group: values [cell][ip][comp]
    std::vector<double> Eval(const CellGroup& cells, const GradientDamage& f, const DofVector<double>& state,
                             IpField field);
    std::vector<double> Eval(const CellGroup& cells, const MomentumBalance& f, const DofVector<double>& state,
                             IpField field);

private:
    void Check(int status, const char* caller) const;
    void Setup(const CellGroup& cells, const std::vector<DofType>& dofs, const nuto_params& p, int dim);
    void Run(const std::vector<DofType>& dofs, const DofVector<double>& state, const std::vector<double>* kappas,
             bool wantR, bool wantK, DofVector<double>* R, DofMatrixSparse<double>* K);
    std::vector<double> RunEval(const std::vector<DofType>& dofs, const DofVector<double>& state,
                                const std::vector<double>* kappas, int nCells, int dim, IpField field);

    DofInfo mDofInfo;
    nuto_ctx* mCtx = nullptr;
    long long mNnz = 0;
    int mRows = 0;
    int mNip = 0;  // IPs per cell of the rule of the last Setup (tetrahedron or tensor-product Gauss)
};

// TimeDependentProblem (TimeDependentProblem.h:20-90, TimeDependentProblem.cpp:34-85): (cell group,
// integrand) pairs. Gradient and Hessian0 assemble every pair with SimpleAssembler and sum the
// results in the order the pairs were added (DofVector +=, DofMatrixSparse += per block);
// UpdateHistory updates every pair's history in place. Groups and integrands are held by
// reference: the caller keeps them alive, as NuTo's Bind captures the integrand by reference
// (Bind.h:14,21). The node values are the dofValues argument (NuTo merges them into its nodes
// first, NodalValueMerger::Merge). dofs: the result's dof types in order; each pair's integrand acts
// on a leading part of them ({displacements} or {displacements, nonlocal eq. strain}), otherwise
// Exception. t and dt are bound into the integrands as TimeDependentProblem's Apply does; these
// integrands do not depend on time.
class TimeDependentProblem
{
public:
    explicit TimeDependentProblem(DofInfo dofInfo = DofInfo(), int device = 0)
        : mDofInfo(dofInfo)
        , mAssembler(dofInfo, device)
    {
    }
    void SetDofInfo(DofInfo dofInfo)
    {
        mDofInfo = dofInfo;
        mAssembler.SetDofInfo(dofInfo);
    }

    void AddGradientFunction(const CellGroup& group, const GradientDamage& f);
    void AddGradientFunction(const CellGroup& group, const MomentumBalance& f);
    void AddHessian0Function(const CellGroup& group, const GradientDamage& f);
    void AddHessian0Function(const CellGroup& group, const MomentumBalance& f);
    void AddUpdateFunction(const CellGroup& group, GradientDamage& f);

    DofVector<double> Gradient(const DofVector<double>& dofValues, std::vector<DofType> dofs, double t, double dt);
    DofMatrixSparse<double> Hessian0(const DofVector<double>& dofValues, std::vector<DofType> dofs, double t,
                                     double dt);
    void UpdateHistory(const DofVector<double>& dofValues, std::vector<DofType> dofs, double t, double dt);

private:
    struct Term
    {
        const CellGroup* group;
        const GradientDamage* gdm;   // one of gdm / mb
        const MomentumBalance* mb;
        GradientDamage* update;      // history owner of an update term
    };
    static void CheckDofs(const Term& term, const std::vector<DofType>& dofs);

    DofInfo mDofInfo;
    SimpleAssembler mAssembler;
    std::vector<Term> mGradient, mHessian0, mUpdate;
};

}  // namespace host
}  // namespace nuto_amd
