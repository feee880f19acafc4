// C++ test of nuto_amd::host::SimpleAssembler (the SimpleAssembler mirror over the C ABI) against
// the oracle's C API: pattern bit-equal, R and K within 1e-12 of the contribution magnitudes, and
// the reference's error behaviour (missing dof numbering, missing DofVector entry). Needs a GPU.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../nuto_amd/host/nuto_assembler.h"

extern "C" {
void* orc_mesh_from_arrays(int coordType, int nNodes, const double* X, int nCells, const int* conn);
void orc_mesh_free(void* m);
int orc_mesh_add_dof(void* mp, int type, int ncomp);
void* orc_problem(void* mp, const nuto_params* p, int ruleKind, int ruleN, int layer0, int layer1);
void orc_problem_free(void* P);
long long orc_problem_assemble(void* Pp, const double* u, const double* e, const double* kappa, int wantR, int wantK,
                               int nthreads);
void orc_problem_R(void* Pp, double* R, double* Rb);
void orc_problem_K(void* Pp, int* rp, int* ci, double* v, double* b);
void orc_problem_lumped_mass(void* Pp, double rho, double* M, double* Mb);
int orc_problem_eval_ip(void* Pp, const double* u, const double* e, const double* kappa, int field, double* out);
void orc_problem_set_theta(void* Pp, const double* E, const double* ft, const double* gf);
}

using namespace nuto_amd::host;

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

static double Uniform(unsigned long long& s, double lo, double hi)  // SplitMix64, top 53 bits
{
    unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return lo + (hi - lo) * ((z >> 11) * (1.0 / 9007199254740992.0));
}

struct Mesh
{
    int coordType, nNodes, nCells, nDofNodes;
    std::vector<double> X;
    std::vector<int> C, ranks;
};

static Mesh MakeMesh(int coordType, int nx, int ny, int nz)
{
    Mesh m;
    m.coordType = coordType;
    nuto_mesh_structured(coordType, nx, ny, nz, nx, ny, nz, 0.0, 0, &m.nNodes, &m.nCells, nullptr, nullptr);
    const int dim = coordType == NUTO_QUAD4 ? 2 : 3, nn = coordType == NUTO_QUAD4 ? 4 : 8;
    m.X.resize((size_t)m.nNodes * dim);
    m.C.resize((size_t)m.nCells * nn);
    nuto_mesh_structured(coordType, nx, ny, nz, nx, ny, nz, 0.0, 0, &m.nNodes, &m.nCells, m.X.data(), m.C.data());
    m.ranks.resize(m.C.size());
    nuto_dof_layer(coordType, m.nCells, m.C.data(), coordType, m.ranks.data(), &m.nDofNodes);
    return m;
}

static std::vector<int> CellDofs(const std::vector<int>& ranks, int ncomp)
{
    std::vector<int> d;
    for (int r : ranks)
        for (int c = 0; c < ncomp; c++)
            d.push_back(r * ncomp + c);
    return d;
}

// R and the monolithic export of K (ToEigen over the integrand's dof types, rows [d0; d1]) against
// the oracle's assembly
static void Compare(const char* what, void* P, int nRows, const std::vector<double>& R,
                    const SparseMatrix<double>& K)
{
    std::vector<double> rR(nRows), rB(nRows);
    orc_problem_R(P, rR.data(), rB.data());
    EXPECT(K.Rows() == nRows && K.Cols() == nRows, (std::string(what) + " size").c_str());
    for (int i = 0; i < nRows; i++)
        EXPECT(std::fabs(R[i] - rR[i]) <= 1e-12 * rB[i], (std::string(what) + " R").c_str());
    std::vector<int> rp(nRows + 1), ci(K.col.size());
    std::vector<double> v(K.col.size()), b(K.col.size());
    orc_problem_K(P, rp.data(), ci.data(), v.data(), b.data());
    EXPECT(rp == K.rowPtr, (std::string(what) + " row_ptr").c_str());
    EXPECT(ci == K.col, (std::string(what) + " col").c_str());
    for (size_t k = 0; k < v.size(); k++)
        if (!(std::fabs(K.val[k] - v[k]) <= 1e-12 * b[k]))
        {
            EXPECT(false, (std::string(what) + " K").c_str());
            break;
        }
}

// The blocks of a two-field matrix: sizes, the export in the other dof order (rows and columns
// [eeq; disp]) as a permutation of [disp; eeq], and the const access to a missing block.
static void BlockChecks(const DofMatrixSparse<double>& K, const DofType& disp, const DofType& eeq)
{
    const int nu = K(disp, disp).Rows(), ne = K(eeq, eeq).Rows();
    EXPECT(K(disp, disp).Cols() == nu && K(disp, eeq).Rows() == nu && K(disp, eeq).Cols() == ne &&
               K(eeq, disp).Rows() == ne && K(eeq, disp).Cols() == nu && K(eeq, eeq).Cols() == ne,
           "block sizes");
    EXPECT(K(disp, eeq).NonZeros() == K(eeq, disp).NonZeros(), "ue and eu blocks have transposed patterns");
    const SparseMatrix<double> a = ToEigen(K, {disp, eeq}), b = ToEigen(K, {eeq, disp});
    EXPECT(a.NonZeros() == b.NonZeros() && TotalNonZeros(K, {disp, eeq}) == a.NonZeros(), "export nonzeros");
    bool same = true;
    for (int i = 0; i < a.Rows() && same; i++)
        for (int k = a.rowPtr[i]; k < a.rowPtr[i + 1]; k++)
        {
            const int i2 = i < nu ? i + ne : i - nu, j2 = a.col[k] < nu ? a.col[k] + ne : a.col[k] - nu;
            same = same && b.Coeff(i2, j2) == a.val[k];
        }
    EXPECT(same, "ToEigen in the order {eeq, disp} permutes the {disp, eeq} export");
    bool threw = false;
    try
    {
        DofType other("Temperature", 1);
        (void)K(disp, other);
    }
    catch (const std::out_of_range&)
    {
        threw = true;
    }
    EXPECT(threw, "missing DofMatrixSparse block throws std::out_of_range");
}

// The reference's container tests (test/mechanics/dofs/DofMatrixSparseConvertEigen.cpp:4-49) and the
// block sum: host code only, no device.
static void ContainerCases()
{
    DofType dof0("foo", 1), dof1("bar", 1);
    {
        DofMatrixSparse<double> m;
        m(dof0, dof0) = SparseMatrix<double>(2, 2);
        m(dof0, dof1) = SparseMatrix<double>(2, 4);
        m(dof1, dof0) = SparseMatrix<double>(4, 2);
        m(dof1, dof1) = SparseMatrix<double>(4, 4);
        double e[6][6] = {};
        m(dof0, dof0).CoeffRef(0, 1) = 1;
        m(dof0, dof0).CoeffRef(1, 0) = 10;
        e[0][1] = 1;
        e[1][0] = 10;
        m(dof0, dof1).CoeffRef(0, 3) = 4;
        e[0][3 + 2] = 4;
        m(dof1, dof0).CoeffRef(2, 1) = 5;
        e[2 + 2][1] = 5;
        m(dof1, dof1).CoeffRef(2, 3) = 6;
        e[2 + 2][3 + 2] = 6;
        const SparseMatrix<double> x = ToEigen(m, {dof0, dof1});
        double err = 0.0;
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++)
                err += (x.Coeff(i, j) - e[i][j]) * (x.Coeff(i, j) - e[i][j]);
        EXPECT(x.Rows() == 6 && x.Cols() == 6 && std::sqrt(err) < 1e-10, "DofMatrixExportEigen");
        // the other order: row/column blocks swap
        const SparseMatrix<double> y = ToEigen(m, {dof1, dof0});
        EXPECT(y.Coeff(2, 1 + 4) == 5 && y.Coeff(0 + 4, 3) == 4 && y.Coeff(2, 3) == 6 && y.Coeff(4, 5) == 1 &&
                   y.Coeff(5, 4) == 10 && y.NonZeros() == 5,
               "ToEigen in the order {dof1, dof0}");
        // DofMatrixContainer::operator+=: summed blocks take the union of the patterns
        DofMatrixSparse<double> s;
        s += m;
        s += m;
        EXPECT(s(dof0, dof1).Coeff(0, 3) == 8 && s(dof1, dof1).NonZeros() == 1, "DofMatrixSparse +=");
    }
    {
        DofMatrixSparse<double> m;
        m(dof0, dof0) = SparseMatrix<double>(2, 8);
        m(dof0, dof1) = SparseMatrix<double>(0, 0);
        m(dof1, dof0) = SparseMatrix<double>(0, 0);
        m(dof1, dof1) = SparseMatrix<double>(0, 10);
        bool ok = true;
        SparseMatrix<double> e;
        try
        {
            e = ToEigen(m, {dof0, dof1});
        }
        catch (...)
        {
            ok = false;
        }
        EXPECT(ok && e.Rows() == 2 && e.Cols() == 18, "DofMatrixZeros");
    }
}

static void GradientDamageCase()
{
    Mesh m = MakeMesh(NUTO_HEX8, 3, 3, 2);
    DofType disp("Displacements", 3), eeq("NonlocalEqStrain", 1);
    DofInfo info;
    info.numIndependent[disp.Id()] = 3 * m.nDofNodes;
    info.numIndependent[eeq.Id()] = m.nDofNodes;
    CellGroup cells;
    cells.coordType = NUTO_HEX8;
    cells.coords = m.X;
    cells.conn = m.C;
    cells.dofNumbering[disp.Id()] = {NUTO_HEX8, CellDofs(m.ranks, 3)};
    cells.dofNumbering[eeq.Id()] = {NUTO_HEX8, CellDofs(m.ranks, 1)};
    GradientDamage f{disp, eeq, 4.0, 30000.0, 0.2, 10.0, 1e-4, 100.0, 0.99, 0, {}};
    unsigned long long seed = 12345;
    DofVector<double> state;
    for (int i = 0; i < 3 * m.nDofNodes; i++)
        state[disp].push_back(Uniform(seed, -5e-3, 5e-3));
    for (int i = 0; i < m.nDofNodes; i++)
        state[eeq].push_back(Uniform(seed, 0.0, 5e-4));
    for (int i = 0; i < m.nCells * 8; i++)
        f.kappas.push_back(Uniform(seed, 0.0, 3e-4));

    // NuTo throws when the dof numbering is missing (SimpleAssembler.cpp:16-23)
    {
        SimpleAssembler none;
        bool threw = false;
        try
        {
            none.BuildVector(cells, f, state);
        }
        catch (const Exception& e)
        {
            threw = std::string(e.what()).find("dof numbering") != std::string::npos;
        }
        EXPECT(threw, "missing dof info throws Exception");
    }
    SimpleAssembler asmbl(info);
    DofVector<double> R = asmbl.BuildVector(cells, f, state);
    DofMatrixSparse<double> K = asmbl.BuildMatrix(cells, f, state);
    std::vector<double> Rall = R[disp];
    Rall.insert(Rall.end(), R[eeq].begin(), R[eeq].end());

    void* om = orc_mesh_from_arrays(NUTO_HEX8, m.nNodes, m.X.data(), m.nCells, m.C.data());
    const int Ld = orc_mesh_add_dof(om, NUTO_HEX8, 3), Le = orc_mesh_add_dof(om, NUTO_HEX8, 1);
    nuto_params p{NUTO_GRADIENT_DAMAGE, 3, 30000.0, 0.2, 0, 10.0, 1e-4, 100.0, 0.99, 4.0, 0};
    void* P = orc_problem(om, &p, 0, 2, Ld, Le);
    orc_problem_assemble(P, state[disp].data(), state[eeq].data(), f.kappas.data(), 1, 1, 1);
    Compare("gradient damage", P, 4 * m.nDofNodes, Rall, ToEigen(K, {disp, eeq}));
    BlockChecks(K, disp, eeq);
    orc_problem_free(P);
    orc_mesh_free(om);

    bool outOfRange = false;
    try
    {
        DofType other("Temperature", 1);
        (void)static_cast<const DofVector<double>&>(R)[other];
    }
    catch (const std::out_of_range&)
    {
        outOfRange = true;
    }
    EXPECT(outOfRange, "missing DofVector entry throws std::out_of_range");
}

// GradientDamage with NonlocalInteraction::Decreasing(0.005, 5) through the mirror (the device runs
// it on the generic row-owner kernel) against the oracle
static void DecreasingInteractionCase()
{
    Mesh m = MakeMesh(NUTO_HEX8, 3, 2, 2);
    DofType disp("Displacements", 3), eeq("NonlocalEqStrain", 1);
    DofInfo info;
    info.numIndependent[disp.Id()] = 3 * m.nDofNodes;
    info.numIndependent[eeq.Id()] = m.nDofNodes;
    CellGroup cells;
    cells.coordType = NUTO_HEX8;
    cells.coords = m.X;
    cells.conn = m.C;
    cells.dofNumbering[disp.Id()] = {NUTO_HEX8, CellDofs(m.ranks, 3)};
    cells.dofNumbering[eeq.Id()] = {NUTO_HEX8, CellDofs(m.ranks, 1)};
    GradientDamage f{disp, eeq, 4.0, 30000.0, 0.2, 10.0, 1e-4, 100.0, 0.99, 0, {}};
    f.interaction = 1;
    unsigned long long seed = 99;
    DofVector<double> state;
    for (int i = 0; i < 3 * m.nDofNodes; i++)
        state[disp].push_back(Uniform(seed, -5e-3, 5e-3));
    for (int i = 0; i < m.nDofNodes; i++)
        state[eeq].push_back(Uniform(seed, 0.0, 5e-4));
    for (int i = 0; i < m.nCells * 8; i++)
        f.kappas.push_back(Uniform(seed, 0.0, 3e-4));
    SimpleAssembler asmbl(info);
    DofVector<double> R = asmbl.BuildVector(cells, f, state);
    DofMatrixSparse<double> K = asmbl.BuildMatrix(cells, f, state);
    std::vector<double> Rall = R[disp];
    Rall.insert(Rall.end(), R[eeq].begin(), R[eeq].end());
    void* om = orc_mesh_from_arrays(NUTO_HEX8, m.nNodes, m.X.data(), m.nCells, m.C.data());
    const int Ld = orc_mesh_add_dof(om, NUTO_HEX8, 3), Le = orc_mesh_add_dof(om, NUTO_HEX8, 1);
    nuto_params p{NUTO_GRADIENT_DAMAGE, 3, 30000.0, 0.2, 0, 10.0, 1e-4, 100.0, 0.99, 4.0, 0, 1, 0.005, 5.0};
    void* P = orc_problem(om, &p, 0, 2, Ld, Le);
    orc_problem_assemble(P, state[disp].data(), state[eeq].data(), f.kappas.data(), 1, 1, 1);
    Compare("gradient damage, decreasing interaction", P, 4 * m.nDofNodes, Rall, ToEigen(K, {disp, eeq}));
    orc_problem_free(P);
    orc_mesh_free(om);
}

static void MomentumBalanceCase()
{
    Mesh m = MakeMesh(NUTO_QUAD4, 5, 4, 0);
    DofType disp("Displacements", 2);
    DofInfo info;
    info.numIndependent[disp.Id()] = 2 * m.nDofNodes;
    CellGroup cells;
    cells.coordType = NUTO_QUAD4;
    cells.coords = m.X;
    cells.conn = m.C;
    cells.dofNumbering[disp.Id()] = {NUTO_QUAD4, CellDofs(m.ranks, 2)};
    MomentumBalance f{disp, 30000.0, 0.2, 0};
    unsigned long long seed = 777;
    DofVector<double> state;
    for (int i = 0; i < 2 * m.nDofNodes; i++)
        state[disp].push_back(Uniform(seed, -1e-3, 1e-3));
    SimpleAssembler asmbl(info);
    DofVector<double> R = asmbl.BuildVector(cells, f, state);
    DofMatrixSparse<double> K = asmbl.BuildMatrix(cells, f, state);
    void* om = orc_mesh_from_arrays(NUTO_QUAD4, m.nNodes, m.X.data(), m.nCells, m.C.data());
    const int Ld = orc_mesh_add_dof(om, NUTO_QUAD4, 2);
    nuto_params p{NUTO_MOMENTUM_LINEAR_ELASTIC, 2, 30000.0, 0.2, 0, 0, 0, 0, 0, 0, 0};
    void* P = orc_problem(om, &p, 0, 2, Ld, -1);
    orc_problem_assemble(P, state[disp].data(), nullptr, nullptr, 1, 1, 1);
    // one dof type: hessian(d, d) is the whole matrix (test/mechanics/cell/SimpleAssembler.cpp:113)
    Compare("momentum balance", P, 2 * m.nDofNodes, R[disp], K(disp, disp));
    orc_problem_free(P);
    orc_mesh_free(om);
}

// Ten-node displacements on tetrahedra (bricks split into six tetrahedra): the mirror picks the
// order-2 tetrahedron rule for the quadratic interpolation; R, K and Eval (4 IPs per cell) against
// the oracle.
static void QuadraticTetrahedraCase()
{
    int nNodes = 0, nCells = 0;
    nuto_mesh_structured(NUTO_TET4, 2, 2, 1, 2.0, 2.0, 1.0, 0.0, 0, &nNodes, &nCells, nullptr, nullptr);
    std::vector<double> X((size_t)nNodes * 3);
    std::vector<int> C((size_t)nCells * 4);
    nuto_mesh_structured(NUTO_TET4, 2, 2, 1, 2.0, 2.0, 1.0, 0.0, 0, &nNodes, &nCells, X.data(), C.data());
    std::vector<int> ranks((size_t)nCells * 10);
    int nDofNodes = 0;
    nuto_dof_layer(NUTO_TET4, nCells, C.data(), NUTO_TET10, ranks.data(), &nDofNodes);
    DofType disp("Displacements", 3);
    DofInfo info;
    info.numIndependent[disp.Id()] = 3 * nDofNodes;
    CellGroup cells;
    cells.coordType = NUTO_TET4;
    cells.coords = X;
    cells.conn = C;
    cells.dofNumbering[disp.Id()] = {NUTO_TET10, CellDofs(ranks, 3)};
    MomentumBalance f{disp, 30000.0, 0.2, 0};
    unsigned long long seed = 4242;
    DofVector<double> state;
    for (int i = 0; i < 3 * nDofNodes; i++)
        state[disp].push_back(Uniform(seed, -1e-3, 1e-3));
    SimpleAssembler asmbl(info);
    DofVector<double> R = asmbl.BuildVector(cells, f, state);
    DofMatrixSparse<double> K = asmbl.BuildMatrix(cells, f, state);
    std::vector<double> stress = asmbl.Eval(cells, f, state, IpField::Stress);
    EXPECT(stress.size() == (size_t)nCells * 4 * 6, "Eval: 4 IPs of the order-2 rule per cell");

    void* om = orc_mesh_from_arrays(NUTO_TET4, nNodes, X.data(), nCells, C.data());
    const int Ld = orc_mesh_add_dof(om, NUTO_TET10, 3);
    nuto_params p{NUTO_MOMENTUM_LINEAR_ELASTIC, 3, 30000.0, 0.2, 0, 0, 0, 0, 0, 0, 0};
    void* P = orc_problem(om, &p, 1, 2, Ld, -1);
    orc_problem_assemble(P, state[disp].data(), nullptr, nullptr, 1, 1, 1);
    Compare("ten-node tetrahedra", P, 3 * nDofNodes, R[disp], ToEigen(K, {disp}));
    std::vector<double> ref(stress.size());
    orc_problem_eval_ip(P, state[disp].data(), nullptr, nullptr, NUTO_IP_STRESS, ref.data());
    double worst = 0.0, scale = 0.0;
    for (size_t i = 0; i < ref.size(); i++)
    {
        worst = std::max(worst, std::fabs(stress[i] - ref[i]));
        scale = std::max(scale, std::fabs(ref[i]));
    }
    EXPECT(worst <= 1e-12 * scale, "Eval stress of ten-node tetrahedra");
    orc_problem_free(P);
    orc_mesh_free(om);
}

// BuildDiagonallyLumpedMatrix with DynamicMomentumBalance and Eval of the stress, against the oracle
static void LumpedMassAndEvalCase()
{
    Mesh m = MakeMesh(NUTO_HEX8, 3, 2, 2);
    DofType disp("Displacements", 3);
    DofInfo info;
    info.numIndependent[disp.Id()] = 3 * m.nDofNodes;
    CellGroup cells;
    cells.coordType = NUTO_HEX8;
    cells.coords = m.X;
    cells.conn = m.C;
    cells.dofNumbering[disp.Id()] = {NUTO_HEX8, CellDofs(m.ranks, 3)};
    DynamicMomentumBalance f{{disp, 30000.0, 0.2, 0}, 2.4e-9};
    SimpleAssembler asmbl(info);
    DofVector<double> M = asmbl.BuildDiagonallyLumpedMatrix(cells, f);
    void* om = orc_mesh_from_arrays(NUTO_HEX8, m.nNodes, m.X.data(), m.nCells, m.C.data());
    const int Ld = orc_mesh_add_dof(om, NUTO_HEX8, 3);
    nuto_params p{NUTO_MOMENTUM_LINEAR_ELASTIC, 3, 30000.0, 0.2, 0, 0, 0, 0, 0, 0, 0};
    void* P = orc_problem(om, &p, 0, 2, Ld, -1);
    const int n = 3 * m.nDofNodes;
    std::vector<double> ref(n), bound(n);
    orc_problem_lumped_mass(P, f.rho, ref.data(), bound.data());
    for (int i = 0; i < n; i++)
        EXPECT(std::fabs(M[disp][i] - ref[i]) <= 1e-12 * bound[i], "lumped mass");
    unsigned long long seed = 4242;
    DofVector<double> state;
    for (int i = 0; i < n; i++)
        state[disp].push_back(Uniform(seed, -1e-3, 1e-3));
    std::vector<double> sig = asmbl.Eval(cells, f, state, IpField::Stress);
    std::vector<double> sref((size_t)m.nCells * 8 * 6);
    EXPECT(orc_problem_eval_ip(P, state[disp].data(), nullptr, nullptr, NUTO_IP_STRESS, sref.data()) == 6,
           "eval components");
    double scale = 0.0;
    for (double v : sref)
        scale = std::fmax(scale, std::fabs(v));
    EXPECT(sig.size() == sref.size(), "eval size");
    for (size_t i = 0; i < sref.size() && i < sig.size(); i++)
        if (!(std::fabs(sig[i] - sref[i]) <= 1e-12 * (std::fabs(sref[i]) + scale)))
        {
            EXPECT(false, "eval stress");
            break;
        }
    orc_problem_free(P);
    orc_mesh_free(om);
}

// QuasistaticSolver::DoStep on the homogeneous gradient-damage patch: every boundary node's
// displacement follows u = t A x (RhsRamp), eeq is free. At t = 1 the exact state is the linear field,
// eeq = the modified Mises norm of A everywhere and kappa = eeq (the step damages the block).
static void QuasistaticCase()
{
    Mesh m = MakeMesh(NUTO_HEX8, 3, 2, 2);
    DofType disp("Displacements", 3), eeq("NonlocalEquivalentStrains", 1);
    DofInfo info;
    info.numIndependent[disp.Id()] = 3 * m.nDofNodes;
    info.numIndependent[eeq.Id()] = m.nDofNodes;
    CellGroup cells;
    cells.coordType = NUTO_HEX8;
    cells.coords = m.X;
    cells.conn = m.C;
    cells.dofNumbering[disp.Id()] = {NUTO_HEX8, CellDofs(m.ranks, 3)};
    cells.dofNumbering[eeq.Id()] = {NUTO_HEX8, CellDofs(m.ranks, 1)};
    GradientDamage gdm{disp, eeq, 4.0, 30000.0, 0.2, 10.0, 1e-4, 100.0, 0.99, 0, {}};
    const double A[3][3] = {{2e-4, 0.5e-4, 0.0}, {0.5e-4, -0.4e-4, 0.3e-4}, {0.0, 0.3e-4, -0.6e-4}};
    std::vector<double> xs(3 * m.nDofNodes);
    for (int c = 0; c < m.nCells; c++)
        for (int a = 0; a < 8; a++)
            for (int d = 0; d < 3; d++)
                xs[3 * m.ranks[c * 8 + a] + d] = m.X[3 * m.C[c * 8 + a] + d];
    const double hi[3] = {3.0, 2.0, 2.0};
    Constraints cons;
    std::vector<double> field(3 * m.nDofNodes);
    for (int r = 0; r < m.nDofNodes; r++)
    {
        bool boundary = false;
        for (int d = 0; d < 3; d++)
            boundary = boundary || std::fabs(xs[3 * r + d]) < 1e-12 || std::fabs(xs[3 * r + d] - hi[d]) < 1e-12;
        for (int i = 0; i < 3; i++)
        {
            field[3 * r + i] = A[i][0] * xs[3 * r] + A[i][1] * xs[3 * r + 1] + A[i][2] * xs[3 * r + 2];
            if (boundary)
            {
                const double v = field[3 * r + i];
                cons.Add(disp, Equation{3 * r + i, {}, [v](double t) { return v * t; }});
            }
        }
    }
    QuasistaticSolver solver(cells, gdm, info);
    solver.mLinearTolerance = 1e-14;
    solver.SetConstraints(cons);
    const int iterations = solver.DoStep(1.0);
    DofVector<double> x = solver.State();
    // modified Mises norm of A: K1 I1 + sqrt(K1^2 I1^2 + K2 J2), J2 = dev(A) : dev(A) / 2
    const double nu = 0.2, k = 10.0, K1 = (k - 1.0) / (2.0 * k * (1.0 - 2.0 * nu)), K2 = 3.0 / (k * (1 + nu) * (1 + nu));
    const double I1 = A[0][0] + A[1][1] + A[2][2];
    double J2 = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
        {
            const double dv = A[i][j] - (i == j ? I1 / 3.0 : 0.0);
            J2 += 0.5 * dv * dv;
        }
    const double e = K1 * I1 + std::sqrt(K1 * K1 * I1 * I1 + K2 * J2);
    EXPECT(iterations >= 1 && e > 1e-4, "nonlinear damaging step");
    for (int i = 0; i < 3 * m.nDofNodes; i++)
        if (std::fabs(x[disp][i] - field[i]) > 1e-10 * 6e-4)
        {
            EXPECT(false, "displacements of the patch");
            break;
        }
    for (int r = 0; r < m.nDofNodes; r++)
        if (std::fabs(x[eeq][r] - e) > 1e-9 * e)
        {
            EXPECT(false, "nonlocal strains of the patch");
            break;
        }
    for (double kap : gdm.kappas)
        if (std::fabs(kap - e) > 1e-9 * e)
        {
            EXPECT(false, "history of the patch");
            break;
        }
    // a step that cannot converge throws NoConvergence and keeps the state
    solver.mTolerance = 1e-30;
    bool threw = false;
    try
    {
        solver.DoStep(2.0);
    }
    catch (const NoConvergence&)
    {
        threw = true;
    }
    EXPECT(threw && solver.GlobalTime() == 1.0, "NoConvergence keeps the state");
}

// TimeDependentProblem with two cell groups of one mesh (cells left / right of x = 2) and two
// gradient-damage materials: the summed R and K (union of the groups' patterns) against the oracle
// on the whole mesh with per-cell material fields; UpdateHistory against Eval(Kappa).
static void TimeDependentProblemCase()
{
    Mesh m = MakeMesh(NUTO_HEX8, 4, 3, 2);
    DofType disp("Displacements", 3), eeq("NonlocalEqStrain", 1);
    DofInfo info;
    info.numIndependent[disp.Id()] = 3 * m.nDofNodes;
    info.numIndependent[eeq.Id()] = m.nDofNodes;
    CellGroup groups[2];
    std::vector<int> cellsOf[2];
    for (int c = 0; c < m.nCells; c++)
    {
        double xc = 0.0;
        for (int k = 0; k < 8; k++)
            xc += m.X[(size_t)m.C[c * 8 + k] * 3] / 8.0;
        const int g = xc < 2.0 ? 0 : 1;
        cellsOf[g].push_back(c);
        groups[g].conn.insert(groups[g].conn.end(), m.C.begin() + c * 8, m.C.begin() + c * 8 + 8);
    }
    const double Eg[2] = {30000.0, 45000.0}, k0g[2] = {1e-4, 2e-4}, betag[2] = {100.0, 150.0};
    std::vector<GradientDamage> f;
    unsigned long long seed = 777;
    for (int g = 0; g < 2; g++)
    {
        groups[g].coordType = NUTO_HEX8;
        groups[g].coords = m.X;
        std::vector<int> ranks;
        for (int c : cellsOf[g])
            ranks.insert(ranks.end(), m.ranks.begin() + c * 8, m.ranks.begin() + c * 8 + 8);
        groups[g].dofNumbering[disp.Id()] = {NUTO_HEX8, CellDofs(ranks, 3)};
        groups[g].dofNumbering[eeq.Id()] = {NUTO_HEX8, CellDofs(ranks, 1)};
        f.push_back(GradientDamage{disp, eeq, 4.0, Eg[g], 0.2, 10.0, k0g[g], betag[g], 0.99, 0, {}});
        for (size_t i = 0; i < cellsOf[g].size() * 8; i++)
            f[g].kappas.push_back(Uniform(seed, 0.0, 3e-4));
    }
    EXPECT(!cellsOf[0].empty() && !cellsOf[1].empty(), "both groups hold cells");
    DofVector<double> state;
    for (int i = 0; i < 3 * m.nDofNodes; i++)
        state[disp].push_back(Uniform(seed, -5e-3, 5e-3));
    for (int i = 0; i < m.nDofNodes; i++)
        state[eeq].push_back(Uniform(seed, 0.0, 5e-4));

    TimeDependentProblem problem(info);
    for (int g = 0; g < 2; g++)
    {
        problem.AddGradientFunction(groups[g], f[g]);
        problem.AddHessian0Function(groups[g], f[g]);
        problem.AddUpdateFunction(groups[g], f[g]);
    }
    DofVector<double> R = problem.Gradient(state, {disp, eeq}, 0.0, 0.0);
    DofMatrixSparse<double> K = problem.Hessian0(state, {disp, eeq}, 0.0, 0.0);
    std::vector<double> Rall = R[disp];
    Rall.insert(Rall.end(), R[eeq].begin(), R[eeq].end());

    // oracle: the whole mesh, per-cell E, f_t = kappa0 E, G_f = f_t / beta, history in global cell order
    std::vector<double> E(m.nCells), ft(m.nCells), gf(m.nCells), kap((size_t)m.nCells * 8);
    for (int g = 0; g < 2; g++)
        for (size_t l = 0; l < cellsOf[g].size(); l++)
        {
            const int c = cellsOf[g][l];
            E[c] = Eg[g];
            ft[c] = k0g[g] * Eg[g];
            gf[c] = ft[c] / betag[g];
            for (int ip = 0; ip < 8; ip++)
                kap[c + (size_t)ip * m.nCells] = f[g].kappas[l + ip * cellsOf[g].size()];
        }
    void* om = orc_mesh_from_arrays(NUTO_HEX8, m.nNodes, m.X.data(), m.nCells, m.C.data());
    const int Ld = orc_mesh_add_dof(om, NUTO_HEX8, 3), Le = orc_mesh_add_dof(om, NUTO_HEX8, 1);
    nuto_params p{NUTO_GRADIENT_DAMAGE, 3, 30000.0, 0.2, 0, 10.0, 1e-4, 100.0, 0.99, 4.0, 0};
    void* P = orc_problem(om, &p, 0, 2, Ld, Le);
    orc_problem_set_theta(P, E.data(), ft.data(), gf.data());
    orc_problem_assemble(P, state[disp].data(), state[eeq].data(), kap.data(), 1, 1, 1);
    Compare("time dependent problem, two groups", P, 4 * m.nDofNodes, Rall, ToEigen(K, {disp, eeq}));
    orc_problem_free(P);
    orc_mesh_free(om);

    // UpdateHistory: kappa <- max(kappa, eeq) = GradientDamage::Kappa before the update
    std::vector<std::vector<double>> expect;
    SimpleAssembler asmbl(info);
    for (int g = 0; g < 2; g++)
        expect.push_back(asmbl.Eval(groups[g], f[g], state, IpField::Kappa));
    problem.UpdateHistory(state, {disp, eeq}, 0.0, 0.0);
    for (int g = 0; g < 2; g++)
    {
        bool same = expect[g].size() == f[g].kappas.size();
        const size_t n = cellsOf[g].size();
        for (size_t c = 0; same && c < n; c++)
            for (int ip = 0; ip < 8; ip++)  // Eval: [cell][ip]; history: cell + ip * n
            {
                const double a = expect[g][c * 8 + ip], b = f[g].kappas[c + ip * n];
                same = same && std::fabs(a - b) <= 1e-14 * std::fabs(a);
            }
        EXPECT(same, "UpdateHistory equals Kappa before the update");
    }

    bool threw = false;
    try
    {
        problem.Gradient(state, {disp}, 0.0, 0.0);  // a gradient-damage term needs {disp, eeq}
    }
    catch (const Exception&)
    {
        threw = true;
    }
    EXPECT(threw, "dof types that do not fit the integrand throw Exception");
}

int main(int argc, char** argv)
{
    // "containers": the host-only container cases (no device needed)
    const bool containersOnly = argc > 1 && std::string(argv[1]) == "containers";
    try
    {
        ContainerCases();
        if (containersOnly)
        {
            std::printf(gFailures ? "FAILED (%d)\n" : "ok\n", gFailures);
            return gFailures ? 1 : 0;
        }
        GradientDamageCase();
        DecreasingInteractionCase();
        MomentumBalanceCase();
        LumpedMassAndEvalCase();
        QuasistaticCase();
        TimeDependentProblemCase();
        QuadraticTetrahedraCase();
    }
    catch (const std::exception& e)
    {
        std::printf("FAIL exception: %s\n", e.what());
        return 1;
    }
    std::printf(gFailures ? "FAILED (%d)\n" : "ok\n", gFailures);
    return gFailures ? 1 : 0;
}
