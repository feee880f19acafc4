// CPU restatement of NuTo's element-assembly path: TEST INFRASTRUCTURE ONLY.
//
// This file is the oracle. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker or as the timed CPU
// baseline. The product (nuto_amd/) never links or calls it.
//
// It restates, without Eigen/Boost, the reference algorithm in the reference's
// operation order (citations are to files under the NuTo source tree):
//   quadrature      nuto/math/Legendre.cpp:8-66, nuto/math/Quadrature.cpp:26-39,
//                   nuto/mechanics/integrationtypes/IntegrationTypeTensorProduct.cpp:15-61,
//                   nuto/mechanics/integrationtypes/IntegrationTypeTetrahedron.cpp:3-10,3083-3107
//   shapes          nuto/mechanics/interpolation/Interpolation{QuadLinear,BrickLinear,
//                   TetrahedronLinear,TetrahedronQuadratic,BrickQuadratic}.cpp
//   Jacobian        nuto/mechanics/cell/Jacobian.h:16-48,100-103,131-146
//   B operators     nuto/mechanics/cell/DifferentialOperators.h:16-91, Matrix.h:8-14
//   laws            nuto/mechanics/constitutive/LinearElastic.h:50-121,
//                   LinearElasticDamage.h:32-99, ModifiedMisesStrainNorm.h:70-268,
//                   EngineeringStrainInvariants.h:32-71, damageLaws/DamageLawExponential.h:26-52,
//                   LocalIsotropicDamage.h:52-133
//   integrands      nuto/mechanics/integrands/MomentumBalance.h:26-46, GradientDamage.h:70-135
//   cell integral   nuto/mechanics/cell/Cell.h:105-119 (result += (f*detJ)*w, first IP assigns)
//   assembly        nuto/mechanics/cell/SimpleAssembler.cpp:38-65,100-156 (one thread, cell order,
//                   triplet list per block, setFromTriplets: duplicates summed in list order)
//   mesh/numbering  nuto/mechanics/mesh/UnitMeshFem.cpp:22-125, MeshFemDofConvert.cpp:9-218,
//                   nuto/mechanics/dofs/DofNumbering.cpp:7-58, MeshFem.cpp:104-115
//   monolithic      nuto/mechanics/dofs/DofMatrixSparseConvertEigen.h:40-67 (rows [d0;d1], cols [d0,d1])
//
// Built with -O3 -msse4 and no -mfma / fast-math, like scripts/cmake/SetCompilerFlags.cmake:3-4.
//
// Parity pins (tests/test_oracle_*.py): 17-digit Legendre roots (test/math/Legendre.cpp:10),
// Gauss exactness, partition of unity / Kronecker / FD of every shape, det J closed forms,
// Brick8N golden K_e/R_e (applications/integrationtests/mechanics/results/Brick8N*.txt),
// Tetrahedron10N golden K_e, mock-cell assembly KATs (test/mechanics/cell/SimpleAssembler.cpp:57-114),
// law closed forms and FD tangents, element-level Hessian0 vs central FD of Gradient,
// K.u == R for linear elasticity, structured closed-form nnz counts and first-touch ranks.
// Functions whose last bits depend on Eigen's product kernels are pinned only to 1e-12
// (the Eigen boundary is "parity unpinned" at the last bit, see the test docstrings).

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <list>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc
{

// ---------------------------------------------------------------------------------------------
// quadrature
// ---------------------------------------------------------------------------------------------

// k-th derivative of the Legendre polynomial P_n at x (nuto/math/Legendre.cpp:8-21).
static double Legendre(int n, double x, int k)
{
    auto fact = [](int i) -> int { return (int)std::tgamma(i + 1); };
    std::vector<double> vals(n + 1, 0.);
    if (n >= k)
        vals[k] = fact(2 * k) / (std::pow(2, k) * fact(k));
    for (int i = k + 1; i <= n; i++)
    {
        double vm2 = i >= 2 ? vals[i - 2] : 0;
        vals[i] = ((2. * i - 1.) * x * vals[i - 1] - (i - 1. + k) * vm2) / (i - k);
    }
    return vals.back();
}

// Newton iteration on P_n from the cosine first guess (nuto/math/Legendre.cpp:23-55).
static std::vector<double> LegendreRoots(int n)
{
    std::vector<double> xs(n);
    for (int i = 1; i <= n; i++)
        xs[i - 1] = std::cos(M_PI * (4. * i - 1) / (4. * n + 2));
    for (double& g : xs)
    {
        double x = g;
        while (true)
        {
            double xn = x - Legendre(n, x, 0) / Legendre(n, x, 1);
            if (std::abs(x - xn) < 1e-15)
            {
                x = xn;
                break;
            }
            x = xn;
        }
        g = x;
    }
    std::sort(xs.begin(), xs.end());
    return xs;
}

// nuto/math/Quadrature.cpp:26-39
static void Gauss1D(int n, std::vector<double>& w, std::vector<double>& x)
{
    x = LegendreRoots(n);
    w.clear();
    for (int i = 0; i < n; i++)
    {
        double dl = Legendre(n, x[i], 1);
        w.push_back(2. / (1. - x[i] * x[i]) / (dl * dl));
    }
}

struct Rule
{
    std::vector<std::array<double, 3>> xi;
    std::vector<double> w;
    int Num() const
    {
        return (int)w.size();
    }
};

// IntegrationTypeTensorProduct.cpp:15-38: IP i -> index_d = (i / n^d) % n, w = prod w_1d.
static Rule TensorGauss(int dim, int n)
{
    std::vector<double> w1, x1;
    Gauss1D(n, w1, x1);
    int num = (int)std::pow(n, dim);
    Rule r;
    r.xi.assign(num, {0., 0., 0.});
    r.w.assign(num, 1.);
    for (int i = 0; i < num; i++)
    {
        int power = 1;
        for (int d = 0; d < dim; d++)
        {
            int idx = (i / power) % n;
            r.w[i] *= w1[idx];
            r.xi[i][d] = x1[idx];
            power *= n;
        }
    }
    return r;
}

// Tabulated tetrahedron rules of order 1 and 2 (IntegrationTypeTetrahedron.cpp:3-10,3083-3107).
// The digits are the reference's own table entries (they differ from 1/6 and 1/24 in the last bits).
static Rule TetRule(int order)
{
    Rule r;
    if (order <= 1)
    {
        r.xi.push_back({0.25, 0.25, 0.25});
        r.w.push_back(0.16666666666666663);
        return r;
    }
    const double a = 0.13819660112501048, b = 0.58541019662496852, w = 0.041666666666666623;
    r.xi = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
    r.w = {w, w, w, w};
    return r;
}

// ---------------------------------------------------------------------------------------------
// shapes
// ---------------------------------------------------------------------------------------------
enum ElemType
{
    QUAD4 = 0,
    HEX8 = 1,
    TET4 = 2,
    TET10 = 3,
    HEX20 = 4
};

static int NumNodes(int t)
{
    static const int n[] = {4, 8, 4, 10, 20};
    return n[t];
}
static int Dim(int t)
{
    return t == QUAD4 ? 2 : 3;
}
static bool SameShape(int a, int b)
{
    auto fam = [](int t) { return t == QUAD4 ? 0 : (t == HEX8 || t == HEX20) ? 1 : 2; };
    return fam(a) == fam(b);
}

static std::array<double, 3> LocalCoords(int t, int i)
{
    static const double q4[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    static const double h8[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                    {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    static const double t10[10][3] = {{0, 0, 0},   {1, 0, 0},     {0, 1, 0},   {0, 0, 1},     {0.5, 0, 0},
                                      {0.5, 0.5, 0}, {0, 0.5, 0}, {0, 0, 0.5}, {0, 0.5, 0.5}, {0.5, 0, 0.5}};
    static const double h20[20][3] = {
            {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},
            {-1, 1, 1},   {0, -1, -1}, {1, 0, -1}, {0, 1, -1},  {-1, 0, -1}, {-1, -1, 0}, {1, -1, 0},
            {1, 1, 0},    {-1, 1, 0},  {0, -1, 1}, {1, 0, 1},   {0, 1, 1},   {-1, 0, 1}};
    switch (t)
    {
    case QUAD4:
        return {q4[i][0], q4[i][1], 0.};
    case HEX8:
        return {h8[i][0], h8[i][1], h8[i][2]};
    case TET4:
    case TET10:
        return {t10[i][0], t10[i][1], t10[i][2]};
    default:
        return {h20[i][0], h20[i][1], h20[i][2]};
    }
}

// N(xi): InterpolationQuadLinear.cpp:54-62, InterpolationBrickLinear.cpp:59-81,
// InterpolationTetrahedronLinear.cpp:52-60, InterpolationTetrahedronQuadratic.cpp:66-101,
// InterpolationBrickQuadratic.cpp:99-143.
static void Shape(int t, const double* x, double* N)
{
    double r = x[0], s = x[1], tt = t == QUAD4 ? 0. : x[2];
    switch (t)
    {
    case QUAD4:
        N[0] = 0.25 * (1. - r) * (1. - s);
        N[1] = 0.25 * (1. + r) * (1. - s);
        N[2] = 0.25 * (1. + r) * (1. + s);
        N[3] = 0.25 * (1. - r) * (1. + s);
        return;
    case HEX8:
    {
        double pr = 1.0 + r, ps = 1.0 + s, pt = 1.0 + tt, mr = 1.0 - r, ms = 1.0 - s, mt = 1.0 - tt;
        N[0] = 0.125 * mr * ms * mt;
        N[1] = 0.125 * pr * ms * mt;
        N[2] = 0.125 * pr * ps * mt;
        N[3] = 0.125 * mr * ps * mt;
        N[4] = 0.125 * mr * ms * pt;
        N[5] = 0.125 * pr * ms * pt;
        N[6] = 0.125 * pr * ps * pt;
        N[7] = 0.125 * mr * ps * pt;
        return;
    }
    case TET4:
        N[0] = 1. - ((r + s) + tt);
        N[1] = r;
        N[2] = s;
        N[3] = tt;
        return;
    case TET10:
        N[0] = 1. - 3. * (r + s + tt) + 2. * (r * r + s * s + tt * tt) + 4. * (r * s + r * tt + s * tt);
        N[1] = -r + 2. * r * r;
        N[2] = -s + 2. * s * s;
        N[3] = -tt + 2. * tt * tt;
        N[4] = 4. * r * (1. - r - s - tt);
        N[5] = 4. * r * s;
        N[6] = 4. * s * (1. - r - s - tt);
        N[7] = 4. * tt * (1. - r - s - tt);
        N[8] = 4. * s * tt;
        N[9] = 4. * r * tt;
        return;
    default:
    {
        double pr = 1.0 + r, ps = 1.0 + s, pt = 1.0 + tt, mr = 1.0 - r, ms = 1.0 - s, mt = 1.0 - tt;
        double dr = 1.0 - r * r, ds = 1.0 - s * s, dt = 1.0 - tt * tt;
        N[0] = 0.125 * mr * ms * mt * (-r - s - tt - 2);
        N[1] = 0.125 * pr * ms * mt * (r - s - tt - 2);
        N[2] = 0.125 * pr * ps * mt * (r + s - tt - 2);
        N[3] = 0.125 * mr * ps * mt * (-r + s - tt - 2);
        N[4] = 0.125 * mr * ms * pt * (-r - s + tt - 2);
        N[5] = 0.125 * pr * ms * pt * (r - s + tt - 2);
        N[6] = 0.125 * pr * ps * pt * (r + s + tt - 2);
        N[7] = 0.125 * mr * ps * pt * (-r + s + tt - 2);
        N[8] = 0.25 * dr * ms * mt;
        N[9] = 0.25 * pr * ds * mt;
        N[10] = 0.25 * dr * ps * mt;
        N[11] = 0.25 * mr * ds * mt;
        N[12] = 0.25 * mr * ms * dt;
        N[13] = 0.25 * pr * ms * dt;
        N[14] = 0.25 * pr * ps * dt;
        N[15] = 0.25 * mr * ps * dt;
        N[16] = 0.25 * dr * ms * pt;
        N[17] = 0.25 * pr * ds * pt;
        N[18] = 0.25 * dr * ps * pt;
        N[19] = 0.25 * mr * ds * pt;
        return;
    }
    }
}

// dN/dxi, row-major n x dim (same files, DerivativeShapeFunctions).
static void DShape(int t, const double* x, double* D)
{
    double r = x[0], s = x[1], tt = t == QUAD4 ? 0. : x[2];
    auto set = [&](int n, double a, double b, double c) {
        D[3 * n] = a;
        D[3 * n + 1] = b;
        D[3 * n + 2] = c;
    };
    switch (t)
    {
    case QUAD4:
        D[0] = -0.25 * (1. - s);
        D[1] = -0.25 * (1. - r);
        D[2] = +0.25 * (1. - s);
        D[3] = -0.25 * (1. + r);
        D[4] = +0.25 * (1. + s);
        D[5] = +0.25 * (1. + r);
        D[6] = -0.25 * (1. + s);
        D[7] = +0.25 * (1. - r);
        return;
    case HEX8:
    {
        double pr = 1.0 + r, ps = 1.0 + s, pt = 1.0 + tt, mr = 1.0 - r, ms = 1.0 - s, mt = 1.0 - tt;
        set(0, -0.125 * ms * mt, -0.125 * mr * mt, -0.125 * mr * ms);
        set(1, 0.125 * ms * mt, -0.125 * pr * mt, -0.125 * pr * ms);
        set(2, 0.125 * ps * mt, 0.125 * pr * mt, -0.125 * pr * ps);
        set(3, -0.125 * ps * mt, 0.125 * mr * mt, -0.125 * mr * ps);
        set(4, -0.125 * ms * pt, -0.125 * mr * pt, 0.125 * mr * ms);
        set(5, 0.125 * ms * pt, -0.125 * pr * pt, 0.125 * pr * ms);
        set(6, 0.125 * ps * pt, 0.125 * pr * pt, 0.125 * pr * ps);
        set(7, -0.125 * ps * pt, 0.125 * mr * pt, 0.125 * mr * ps);
        return;
    }
    case TET4:
        set(0, -1, -1, -1);
        set(1, 1, 0, 0);
        set(2, 0, 1, 0);
        set(3, 0, 0, 1);
        return;
    case TET10:
    {
        double a = -3. + 4. * (r + s + tt);
        set(0, a, a, a);
        set(1, -1. + 4. * r, 0, 0);
        set(2, 0, -1. + 4. * s, 0);
        set(3, 0, 0, -1. + 4. * tt);
        set(4, 4. - 8. * r - 4. * s - 4. * tt, -4. * r, -4. * r);
        set(5, 4. * s, 4. * r, 0.);
        set(6, -4. * s, 4. - 8. * s - 4. * r - 4. * tt, -4. * s);
        set(7, -4. * tt, -4. * tt, 4. - 8. * tt - 4. * r - 4. * s);
        set(8, 0., 4. * tt, 4. * s);
        set(9, 4. * tt, 0, 4. * r);
        return;
    }
    default:
    {
        double pr = 1.0 + r, ps = 1.0 + s, pt = 1.0 + tt, mr = 1.0 - r, ms = 1.0 - s, mt = 1.0 - tt;
        double dr = 1.0 - r * r, ds = 1.0 - s * s, dt = 1.0 - tt * tt;
        set(0, 0.125 * ms * mt * (2 * r + s + tt + 1), 0.125 * mr * mt * (r + 2 * s + tt + 1),
            0.125 * mr * ms * (r + s + 2 * tt + 1));
        set(1, 0.125 * ms * mt * (2 * r - s - tt - 1), -0.125 * pr * mt * (r - 2 * s - tt - 1),
            -0.125 * pr * ms * (r - s - 2 * tt - 1));
        set(2, 0.125 * ps * mt * (2 * r + s - tt - 1), 0.125 * pr * mt * (r + 2 * s - tt - 1),
            -0.125 * pr * ps * (r + s - 2 * tt - 1));
        set(3, 0.125 * ps * mt * (2 * r - s + tt + 1), -0.125 * mr * mt * (r - 2 * s + tt + 1),
            0.125 * mr * ps * (r - s + 2 * tt + 1));
        set(4, 0.125 * ms * pt * (2 * r + s - tt + 1), 0.125 * mr * pt * (r + 2 * s - tt + 1),
            -0.125 * mr * ms * (r + s - 2 * tt + 1));
        set(5, 0.125 * ms * pt * (2 * r - s + tt - 1), -0.125 * pr * pt * (r - 2 * s + tt - 1),
            0.125 * pr * ms * (r - s + 2 * tt - 1));
        set(6, 0.125 * ps * pt * (2 * r + s + tt - 1), 0.125 * pr * pt * (r + 2 * s + tt - 1),
            0.125 * pr * ps * (r + s + 2 * tt - 1));
        set(7, 0.125 * ps * pt * (2 * r - s - tt + 1), -0.125 * mr * pt * (r - 2 * s - tt + 1),
            -0.125 * mr * ps * (r - s - 2 * tt + 1));
        set(8, -0.5 * r * ms * mt, -0.25 * dr * mt, -0.25 * dr * ms);
        set(9, +0.25 * ds * mt, -0.5 * s * pr * mt, -0.25 * pr * ds);
        set(10, -0.5 * r * ps * mt, 0.25 * dr * mt, -0.25 * dr * ps);
        set(11, -0.25 * ds * mt, -0.5 * s * mr * mt, -0.25 * mr * ds);
        set(12, -0.25 * ms * dt, -0.25 * mr * dt, -0.5 * tt * mr * ms);
        set(13, 0.25 * ms * dt, -0.25 * pr * dt, -0.5 * tt * pr * ms);
        set(14, 0.25 * ps * dt, 0.25 * pr * dt, -0.5 * tt * pr * ps);
        set(15, -0.25 * ps * dt, 0.25 * mr * dt, -0.5 * tt * mr * ps);
        set(16, -0.5 * r * ms * pt, -0.25 * dr * pt, 0.25 * dr * ms);
        set(17, 0.25 * ds * pt, -0.5 * s * pr * pt, 0.25 * pr * ds);
        set(18, -0.5 * r * ps * pt, 0.25 * dr * pt, 0.25 * dr * ps);
        set(19, -0.25 * ds * pt, -0.5 * s * mr * pt, 0.25 * mr * ds);
        return;
    }
    }
}

// dN/dxi as a dense n x dim matrix (the dim-column view used by the Jacobian).
static void DShapeDim(int t, const double* x, double* D)
{
    int n = NumNodes(t), d = Dim(t);
    double tmp[60];
    DShape(t, x, tmp);
    if (d == 3)
    {
        std::memcpy(D, tmp, sizeof(double) * 3 * n);
        return;
    }
    for (int i = 0; i < n; i++)
    {
        D[2 * i] = tmp[2 * i];
        D[2 * i + 1] = tmp[2 * i + 1];
    }
}

// ---------------------------------------------------------------------------------------------
// dense helpers (row-major)
// ---------------------------------------------------------------------------------------------
struct Mat
{
    int r = 0, c = 0;
    std::vector<double> v;
    Mat() = default;
    Mat(int rr, int cc)
        : r(rr)
        , c(cc)
        , v((size_t)rr * cc, 0.)
    {
    }
    double& operator()(int i, int j)
    {
        return v[(size_t)i * c + j];
    }
    double operator()(int i, int j) const
    {
        return v[(size_t)i * c + j];
    }
};

static Mat Mul(const Mat& a, const Mat& b)
{
    if (a.c != b.r)
        throw std::runtime_error("Mul: shape mismatch");
    Mat o(a.r, b.c);
    for (int i = 0; i < a.r; i++)
        for (int j = 0; j < b.c; j++)
        {
            double s = 0.;
            for (int k = 0; k < a.c; k++)
                s += a(i, k) * b(k, j);
            o(i, j) = s;
        }
    return o;
}
static Mat T(const Mat& a)
{
    Mat o(a.c, a.r);
    for (int i = 0; i < a.r; i++)
        for (int j = 0; j < a.c; j++)
            o(j, i) = a(i, j);
    return o;
}
static Mat Scale(Mat a, double s)
{
    for (double& x : a.v)
        x *= s;
    return a;
}
static Mat Add(Mat a, const Mat& b)
{
    for (size_t i = 0; i < a.v.size(); i++)
        a.v[i] += b.v[i];
    return a;
}
static Mat Col(const std::vector<double>& x)
{
    Mat o((int)x.size(), 1);
    o.v = x;
    return o;
}

// Magnitude shadows: |A| and |A| * |B|. The parity scale of an output entry is the same expression
// evaluated on absolute values (the componentwise scale of the rounding-error bound of a sum of
// products); an entry that is a cancellation result is compared against the size of its terms.
static Mat Abs(Mat a)
{
    for (double& x : a.v)
        x = std::abs(x);
    return a;
}
static Mat MulAbs(const Mat& a, const Mat& b)
{
    return Mul(Abs(a), Abs(b));
}

// ---------------------------------------------------------------------------------------------
// Jacobian (Jacobian.h:16-48,100-103,131-146): J = X(dim x n) * dN(n x dim), inverse by the
// adjugate, determinant by cofactor expansion along the first row.
// ---------------------------------------------------------------------------------------------
// Planted mistakes (tests only, off by default): tests/test_parity_sensitivity.py switches one on
// and checks that the parity rule used against the device path rejects the result.
//   1 drop the c B_e^T B_e term of K_ee        2 flip the sign of K_eu
//   3 halve the engineering shear strains       4 kappa = eeq (history max dropped)
//   5 flip the sign of the c B_e^T grad(eeq) term of R_e
//   6 dN/dX with J^-T instead of J^-1 (transposed operand)
static int g_mutation = 0;

struct Jacobian
{
    int d = 0;
    double J[9] = {0}, inv[9] = {0}, det = 0;
    // magnitude shadow of inv: the adjugate's products taken in absolute value, over |det|. An
    // inverse entry that is a cancellation result is compared against the size of its terms.
    double invMag[9] = {0};
};

static Jacobian MakeJacobian(int dim, int n, const double* X /* node-major */, const double* dN /* n x dim */)
{
    Jacobian jac;
    jac.d = dim;
    for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++)
        {
            double s = 0.;
            for (int k = 0; k < n; k++)
                s += X[k * dim + i] * dN[k * dim + j];
            jac.J[i * dim + j] = s;
        }
    const double* m = jac.J;
    if (dim == 2)
    {
        double det = m[0] * m[3] - m[2] * m[1];
        double id = 1. / det;
        jac.inv[0] = m[3] * id;
        jac.inv[2] = -m[2] * id;
        jac.inv[1] = -m[1] * id;
        jac.inv[3] = m[0] * id;
        jac.det = det;
        for (int i = 0; i < 4; i++)
            jac.invMag[i] = std::abs(jac.inv[i]);
    }
    else
    {
        auto M = [&](int i, int j) { return m[i * 3 + j]; };
        auto cof = [&](int i, int j) {
            int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            return M(i1, j1) * M(i2, j2) - M(i1, j2) * M(i2, j1);
        };
        double c0[3] = {cof(0, 0), cof(1, 0), cof(2, 0)};
        double detc = c0[0] * M(0, 0) + c0[1] * M(1, 0) + c0[2] * M(2, 0);
        double id = 1. / detc;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                jac.inv[i * 3 + j] = (i == 0 ? c0[j] : cof(j, i)) * id;
        auto cofMag = [&](int i, int j) {
            int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            return std::abs(M(i1, j1) * M(i2, j2)) + std::abs(M(i1, j2) * M(i2, j1));
        };
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                jac.invMag[i * 3 + j] = cofMag(j, i) * std::abs(id);
        auto h = [&](int a, int b, int c) { return M(0, a) * (M(1, b) * M(2, c) - M(1, c) * M(2, b)); };
        jac.det = h(0, 1, 2) - h(1, 0, 2) + h(2, 0, 1);
    }
    return jac;
}

// dN/dX = dN/dxi * J^-1 (Jacobian.h:100-103)
static Mat GlobalDerivatives(const Jacobian& jac, const double* dN, int n)
{
    int d = jac.d;
    Mat o(n, d);
    for (int k = 0; k < n; k++)
        for (int j = 0; j < d; j++)
        {
            double s = 0.;
            for (int i = 0; i < d; i++)
                s += dN[k * d + i] * jac.inv[i * d + j];
            o(k, j) = s;
        }
    return o;
}

// magnitude shadow of dN/dX: |dN/dxi| * |J^-1| (shadow), so that a derivative that cancels to
// (nearly) zero, e.g. the barycentric gradient of a tetrahedron's vertex opposite a face that lies
// in a coordinate plane, is compared against the size of its terms
static Mat GlobalDerivativesMag(const Jacobian& jac, const double* dN, int n)
{
    int d = jac.d;
    Mat o(n, d);
    for (int k = 0; k < n; k++)
        for (int j = 0; j < d; j++)
        {
            double s = 0.;
            for (int i = 0; i < d; i++)
                s += std::abs(dN[k * d + i]) * jac.invMag[i * d + j];
            o(k, j) = s;
        }
    return o;
}

// Nabla::Strain (DifferentialOperators.h:24-86): rows [xx,yy,xy] in 2D, [xx,yy,zz,yz,zx,xy] in 3D
static Mat StrainB(const Mat& dNdX)
{
    int d = dNdX.c, n = dNdX.r;
    if (d == 2)
    {
        Mat B(3, 2 * n);
        for (int a = 0; a < n; a++)
        {
            B(0, 2 * a) = dNdX(a, 0);
            B(1, 2 * a + 1) = dNdX(a, 1);
            B(2, 2 * a) = dNdX(a, 1);
            B(2, 2 * a + 1) = dNdX(a, 0);
        }
        return B;
    }
    Mat B(6, 3 * n);
    for (int a = 0; a < n; a++)
    {
        double x = dNdX(a, 0), y = dNdX(a, 1), z = dNdX(a, 2);
        int c = 3 * a;
        B(0, c) = x;
        B(1, c + 1) = y;
        B(2, c + 2) = z;
        B(3, c + 1) = z;
        B(3, c + 2) = y;
        B(4, c) = z;
        B(4, c + 2) = x;
        B(5, c) = y;
        B(5, c + 1) = x;
    }
    return B;
}

// Matrix::N (Matrix.h:8-14)
static Mat NMatrix(const std::vector<double>& N, int dim)
{
    int n = (int)N.size();
    Mat o(dim, dim * n);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < dim; d++)
            o(d, i * dim + d) = N[i];
    return o;
}

// ---------------------------------------------------------------------------------------------
// constitutive laws
// ---------------------------------------------------------------------------------------------
struct Params
{
    int kind;  // 0 MomentumBalance+LinearElastic, 1 GradientDamage, 2 MomentumBalance+LocalIsotropicDamage
    int dim;
    double E, nu;
    int plane_state;  // 0 PLANE_STRESS, 1 PLANE_STRAIN (LinearElastic<2> only)
    double k, kappa0, beta, alpha, c;
    int damage_application;  // 0 FULL, 1 UNILATERAL
    int interaction;         // GradientDamage's NonlocalInteraction: 0 Constant, 1 Decreasing
    double interaction_r, interaction_eta;
};

// NonlocalInteraction::Constant::Factor/Derivative and Decreasing::Factor/Derivative
// (GradientDamage.h:153-184)
static void Interaction(const Params& p, double omega, double& g, double& dgdw)
{
    if (p.interaction == 0)
    {
        g = 1.;
        dgdw = 0.;
        return;
    }
    const double R = p.interaction_r, eta = p.interaction_eta;
    g = ((1. - R) * std::exp(-eta * omega) + R - std::exp(-eta)) / (1. - std::exp(-eta));
    dgdw = ((1. - R) * std::exp(-eta * omega)) / (1. - std::exp(-eta)) * -eta;
}

// LinearElastic<D>::CalculateC (LinearElastic.h:50-121)
static Mat ElasticC(int dim, double E, double nu, int planeState)
{
    if (dim == 2)
    {
        double C11, C12, C33;
        if (planeState == 0)
        {
            double f = E / (1.0 - (nu * nu));
            C11 = f;
            C12 = f * nu;
            C33 = f * 0.5 * (1.0 - nu);
        }
        else
        {
            double f = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
            C11 = f * (1.0 - nu);
            C12 = f * nu;
            C33 = E / (2. * (1.0 + nu));
        }
        Mat C(3, 3);
        C(0, 0) = C11;
        C(1, 0) = C12;
        C(0, 1) = C12;
        C(1, 1) = C11;
        C(2, 2) = C33;
        return C;
    }
    double f = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double C11 = f * (1.0 - nu), C12 = f * nu, C44 = E / (2. * (1.0 + nu));
    Mat C(6, 6);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            C(i, j) = i == j ? C11 : C12;
    for (int i = 3; i < 6; i++)
        C(i, i) = C44;
    return C;
}

// Laws::LinearElasticDamage<3> (LinearElasticDamage.h:32-99), 3D only.
struct ElasticDamage3
{
    double m3K, m2G;
    int app;
    double D[6] = {1, 1, 1, 0, 0, 0};
    double Pinv[6] = {1, 1, 1, 0.5, 0.5, 0.5};
    Mat Iv, PinvId;
    ElasticDamage3(double E, double nu, int application)
        : m3K(3 * E / (3 - 6 * nu))
        , m2G(2 * E / (2 + 2 * nu))
        , app(application)
        , Iv(6, 6)
        , PinvId(6, 6)
    {
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++)
                Iv(i, j) = 1. / 3. * D[i] * D[j];
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++)
                PinvId(i, j) = Pinv[i] * ((i == j ? 1. : 0.) - Iv(i, j));
    }
    double H(double eV) const
    {
        return app == 1 ? (eV < 0 ? 0 : 1) : 1;
    }
    double EV(const double* s) const
    {
        double t = 0.;
        for (int i = 0; i < 6; i++)
            t += D[i] * s[i];
        return 1. / 3. * t;
    }
    std::vector<double> Stress(const double* s, double w) const
    {
        double eV = EV(s);
        double sV = (1. - w * H(eV)) * m3K * eV;
        std::vector<double> o(6);
        for (int i = 0; i < 6; i++)
        {
            double e = s[i] - D[i] * eV;
            o[i] = D[i] * sV + (1. - w) * m2G * (Pinv[i] * e);
        }
        return o;
    }
    Mat DstressDstrain(const double* s, double w) const
    {
        double eV = EV(s);
        Mat o(6, 6);
        double a = m3K * (1. - w * H(eV)), b = (1. - w) * m2G;
        for (int i = 0; i < 36; i++)
            o.v[i] = a * Iv.v[i] + b * PinvId.v[i];
        return o;
    }
    std::vector<double> DstressDomega(const double* s) const
    {
        double eV = EV(s);
        double dsV = -m3K * eV * H(eV);
        std::vector<double> o(6);
        for (int i = 0; i < 6; i++)
        {
            double e = s[i] - D[i] * eV;
            o[i] = D[i] * dsV - m2G * (Pinv[i] * e);
        }
        return o;
    }
};

// Constitutive::ModifiedMisesStrainNorm<3> (ModifiedMisesStrainNorm.h:70-84,136-146,231-268)
struct ModMises3
{
    double K1, K2;
    ModMises3(double nu, double k)
        : K1((k - 1.) / (2. * k * (1 - 2 * nu)))
        , K2(3. / (k * (1. + nu) * (1. + nu)))
    {
    }
    static double I1(const double* v)
    {
        return v[0] + v[1] + v[2];
    }
    static double J2(const double* v)
    {
        double xx = v[0], yy = v[1], zz = v[2], yz = .5 * v[3], zx = .5 * v[4], xy = .5 * v[5];
        return 1. / 6. * ((xx - yy) * (xx - yy) + (yy - zz) * (yy - zz) + (zz - xx) * (zz - xx)) + xy * xy +
               yz * yz + zx * zx;
    }
    double Value(const double* s) const
    {
        double i1 = I1(s), j2 = J2(s);
        double A = std::sqrt(K1 * K1 * i1 * i1 + K2 * j2);
        return K1 * i1 + A;
    }
    std::vector<double> Derivative(const double* s) const
    {
        double i1 = I1(s), j2 = J2(s);
        double A = std::sqrt(K1 * K1 * i1 * i1 + K2 * j2);
        std::vector<double> d(6);
        if (A == 0)
        {
            d = {K1, K1, K1, 0, 0, 0};
            return d;
        }
        double dxx = 1. / 3. * (2 * s[0] - s[1] - s[2]);
        double dyy = 1. / 3. * (2 * s[1] - s[0] - s[2]);
        double dzz = 1. / 3. * (2 * s[2] - s[0] - s[1]);
        d[0] = K1 + 1. / (2 * A) * (2 * K1 * K1 * i1 + K2 * dxx);
        d[1] = K1 + 1. / (2 * A) * (2 * K1 * K1 * i1 + K2 * dyy);
        d[2] = K1 + 1. / (2 * A) * (2 * K1 * K1 * i1 + K2 * dzz);
        d[3] = 1. / (2 * A) * K2 * (0.5 * s[3]);
        d[4] = 1. / (2 * A) * K2 * (0.5 * s[4]);
        d[5] = 1. / (2 * A) * K2 * (0.5 * s[5]);
        return d;
    }
};

// Constitutive::DamageLawExponential (DamageLawExponential.h:26-52)
struct DamageExp
{
    double k0, beta, alpha;
    double Damage(double k) const
    {
        if (k < k0)
            return 0.;
        return 1 - k0 / k * (1 - alpha + alpha * std::exp(beta * (k0 - k)));
    }
    double Derivative(double k) const
    {
        if (k < k0)
            return 0.;
        return k0 / k * ((1 / k + beta) * alpha * std::exp(beta * (k0 - k)) + (1 - alpha) / k);
    }
};

// ---------------------------------------------------------------------------------------------
// mesh, dof layers, first-touch numbering
// ---------------------------------------------------------------------------------------------
struct DofLayer
{
    int type = 0, ncomp = 1, nnodes = 0;
    std::vector<int> conn;  // nCells x NumNodes(type): dof-node ids (creation order)
};

struct Mesh
{
    int dim = 0, coordType = 0, nCells = 0, nNodes = 0;
    std::vector<double> X;  // nNodes x dim, node-major
    std::vector<int> conn;  // nCells x NumNodes(coordType)
    std::vector<DofLayer> dofs;
};

// UnitMeshFem::CreateQuads (UnitMeshFem.cpp:22-33,67-81)
static Mesh CreateQuads(int nx, int ny)
{
    Mesh m;
    m.dim = 2;
    m.coordType = QUAD4;
    for (int iy = 0; iy < ny + 1; ++iy)
        for (int ix = 0; ix < nx + 1; ++ix)
        {
            m.X.push_back(static_cast<double>(ix) / nx);
            m.X.push_back(static_cast<double>(iy) / ny);
        }
    m.nNodes = (nx + 1) * (ny + 1);
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix)
        {
            int n0 = ix + iy * (nx + 1);
            m.conn.insert(m.conn.end(), {n0, n0 + 1, n0 + 1 + (nx + 1), n0 + (nx + 1)});
        }
    m.nCells = nx * ny;
    return m;
}

// UnitMeshFem::CreateBricks (UnitMeshFem.cpp:83-114)
static Mesh CreateBricks(int nx, int ny, int nz)
{
    Mesh m;
    m.dim = 3;
    m.coordType = HEX8;
    int ex = nx + 1, ey = ny + 1, ez = nz + 1;
    for (int iz = 0; iz < ez; ++iz)
        for (int iy = 0; iy < ey; ++iy)
            for (int ix = 0; ix < ex; ++ix)
            {
                m.X.push_back(static_cast<double>(ix) / nx);
                m.X.push_back(static_cast<double>(iy) / ny);
                m.X.push_back(static_cast<double>(iz) / nz);
            }
    m.nNodes = ex * ey * ez;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix)
            {
                auto id = [&](int a, int b, int c) { return a + b * ex + c * ex * ey; };
                m.conn.insert(m.conn.end(), {id(ix, iy, iz), id(ix + 1, iy, iz), id(ix + 1, iy + 1, iz),
                                             id(ix, iy + 1, iz), id(ix, iy, iz + 1), id(ix + 1, iy, iz + 1),
                                             id(ix + 1, iy + 1, iz + 1), id(ix, iy + 1, iz + 1)});
            }
    m.nCells = nx * ny * nz;
    return m;
}

// Interpolate(element, xi) = N(xi) * nodeValues (ElementInterface.h:31-34), padded to 3D.
static std::array<double, 3> InterpolateCoord(const Mesh& m, int cell, const std::array<double, 3>& xi)
{
    int n = NumNodes(m.coordType);
    double N[20];
    Shape(m.coordType, xi.data(), N);
    std::array<double, 3> o = {0., 0., 0.};
    for (int d = 0; d < m.dim; d++)
    {
        double s = 0.;
        for (int k = 0; k < n; k++)
            s += N[k] * m.X[(size_t)m.conn[(size_t)cell * n + k] * m.dim + d];
        o[d] = s;
    }
    return o;
}

// AddDofInterpolation (MeshFemDofConvert.cpp:164-210): walk the elements in order; a dof node is
// re-used when an existing one lies within eps = 1e-10 of the interpolated coordinate, otherwise
// it is created. The sub-box grid (300 boxes per used direction, domain grown by eps) is kept
// sparse here; the 27-neighbour search makes the match independent of box boundaries.
static int AddDofInterpolation(Mesh& m, int type, int ncomp)
{
    if (!SameShape(type, m.coordType))
        throw std::runtime_error("AddDofInterpolation: shape mismatch");
    const double eps = 1.e-10;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    int nnC = NumNodes(m.coordType);
    for (int d = 0; d < m.dim; d++)
        lo[d] = hi[d] = m.X[(size_t)m.conn[0] * m.dim + d];
    for (int c = 0; c < m.nCells; c++)
        for (int k = 0; k < nnC; k++)
            for (int d = 0; d < m.dim; d++)
            {
                double x = m.X[(size_t)m.conn[(size_t)c * nnC + k] * m.dim + d];
                lo[d] = std::min(lo[d], x);
                hi[d] = std::max(hi[d], x);
            }
    int subs[3] = {1, 1, 1};
    double start[3], delta[3];
    for (int d = 0; d < 3; d++)
    {
        start[d] = lo[d] - eps;
        if (d < m.dim)
            subs[d] = 300;
        delta[d] = ((hi[d] + eps) - start[d]) / subs[d];
    }
    std::unordered_map<long long, std::vector<int>> boxes;
    std::vector<std::array<double, 3>> pts;
    auto key = [&](int i, int j, int k) { return (long long)i + (long long)j * subs[0] + (long long)k * subs[0] * subs[1]; };
    DofLayer L;
    L.type = type;
    L.ncomp = ncomp;
    int nnD = NumNodes(type);
    L.conn.resize((size_t)m.nCells * nnD);
    for (int c = 0; c < m.nCells; c++)
        for (int a = 0; a < nnD; a++)
        {
            std::array<double, 3> x = InterpolateCoord(m, c, LocalCoords(type, a));
            int i = (int)((x[0] - start[0]) / delta[0]);
            int j = (int)((x[1] - start[1]) / delta[1]);
            int k = (int)((x[2] - start[2]) / delta[2]);
            int found = -1;
            for (int kk : {k, k - 1, k + 1})
                for (int jj : {j, j - 1, j + 1})
                    for (int ii : {i, i - 1, i + 1})
                    {
                        if (found >= 0 || ii < 0 || ii >= subs[0] || jj < 0 || jj >= subs[1] || kk < 0 ||
                            kk >= subs[2])
                            continue;
                        auto it = boxes.find(key(ii, jj, kk));
                        if (it == boxes.end())
                            continue;
                        for (int p : it->second)
                        {
                            double dx = pts[p][0] - x[0], dy = pts[p][1] - x[1], dz = pts[p][2] - x[2];
                            if (dx * dx + dy * dy + dz * dz < eps * eps)
                            {
                                found = p;
                                break;
                            }
                        }
                    }
            if (found < 0)
            {
                found = (int)pts.size();
                pts.push_back(x);
                boxes[key(i, j, k)].push_back(found);
            }
            L.conn[(size_t)c * nnD + a] = found;
        }
    L.nnodes = (int)pts.size();
    m.dofs.push_back(L);
    return (int)m.dofs.size() - 1;
}

// DofNumbering::Build (DofNumbering.cpp:7-14,32-58) over MeshFem::NodesTotal(dof)
// (MeshFem.cpp:104-115): nodes in first-appearance order, component-minor. Because
// NodesTotal walks the elements in the same order that created the nodes, the rank of a dof
// node is its creation index; this is recomputed here by the walk itself.
static std::vector<int> DofNumbers(const Mesh& m, int layer)
{
    const DofLayer& L = m.dofs[layer];
    int nn = NumNodes(L.type);
    std::vector<int> rank(L.nnodes, -1);
    int next = 0;
    for (int c = 0; c < m.nCells; c++)
        for (int a = 0; a < nn; a++)
        {
            int node = L.conn[(size_t)c * nn + a];
            if (rank[node] < 0)
                rank[node] = next++;
        }
    std::vector<int> num((size_t)L.nnodes * L.ncomp);
    for (int node = 0; node < L.nnodes; node++)
        for (int comp = 0; comp < L.ncomp; comp++)
            num[(size_t)node * L.ncomp + comp] = rank[node] * L.ncomp + comp;
    return num;
}

// ---------------------------------------------------------------------------------------------
// per-cell integration and assembly
// ---------------------------------------------------------------------------------------------
struct Problem
{
    Mesh mesh;
    Params p;
    Rule rule;
    int layer[2] = {-1, -1};  // dof layers: [disp] or [disp, eeq]
    std::vector<int> numbers[2];
    int ndof[2] = {0, 0};
    int ntypes = 1;
    // per-element material fields (E, ft, Gf), optional (GradientDamage only)
    std::vector<double> thE, thFt, thGf;
    // results
    std::vector<double> R;  // monolithic [d0; d1]
    std::vector<int> rowPtr, col;
    std::vector<double> val, bound;
    std::vector<double> Rbound;
};

struct CellResult
{
    std::vector<double> R[2];
    Mat K[2][2];
    std::vector<double> Rm[2];  // magnitude scales (see Abs / MulAbs)
    Mat Km[2][2];
};

// magnitudes of the 3D law outputs for a strain of magnitude em (em = |B| |u|)
struct LawMag
{
    double sig[6], dsw[6], eeq, m[6];
};

static LawMag LawMagnitudes(const double* em, double omega, double E, double nu, double k, int app, double A)
{
    LawMag o;
    const double m3K = 3 * E / (3 - 6 * nu), m2G = 2 * E / (2 + 2 * nu);
    const double eV = (em[0] + em[1] + em[2]) / 3.0;
    const double H = 1.0;  // bounds use the larger (damaged) volumetric factor for UNILATERAL as well
    for (int i = 0; i < 3; i++)
    {
        o.sig[i] = std::abs(1. - omega * H) * m3K * eV + std::abs(1. - omega) * m2G * (em[i] + eV);
        o.dsw[i] = m3K * eV * H + m2G * (em[i] + eV);
    }
    for (int i = 3; i < 6; i++)
    {
        o.sig[i] = std::abs(1. - omega) * m2G * 0.5 * em[i];
        o.dsw[i] = m2G * 0.5 * em[i];
    }
    (void)app;
    const double K1 = (k - 1.) / (2. * k * (1 - 2 * nu)), K2 = 3. / (k * (1. + nu) * (1. + nu));
    const double I1 = em[0] + em[1] + em[2];
    const double J2 = ((em[0] + em[1]) * (em[0] + em[1]) + (em[1] + em[2]) * (em[1] + em[2]) +
                       (em[2] + em[0]) * (em[2] + em[0])) / 6. +
                      0.25 * (em[3] * em[3] + em[4] * em[4] + em[5] * em[5]);
    o.eeq = K1 * I1 + std::sqrt(K1 * K1 * I1 * I1 + K2 * J2);
    if (A == 0)
    {
        for (int i = 0; i < 3; i++)
            o.m[i] = K1;
        for (int i = 3; i < 6; i++)
            o.m[i] = 0;
    }
    else
    {
        const double h = 1. / (2 * A);
        for (int i = 0; i < 3; i++)
            o.m[i] = K1 + h * (2 * K1 * K1 * I1 + K2 * (2 * em[i] + em[(i + 1) % 3] + em[(i + 2) % 3]) / 3.);
        for (int i = 3; i < 6; i++)
            o.m[i] = h * K2 * 0.5 * em[i];
    }
    return o;
}

static double MisesA(const double* s, double nu, double k)
{
    const double K1 = (k - 1.) / (2. * k * (1 - 2 * nu)), K2 = 3. / (k * (1. + nu) * (1. + nu));
    double i1 = ModMises3::I1(s), j2 = ModMises3::J2(s);
    return std::sqrt(K1 * K1 * i1 * i1 + K2 * j2);
}

// element-local material for cell c
struct CellLaw
{
    double E, kappa0, beta;
};

static CellLaw LawFor(const Problem& P, int c)
{
    CellLaw L{P.p.E, P.p.kappa0, P.p.beta};
    if (!P.thE.empty())
    {
        L.E = P.thE[c];
        L.kappa0 = P.thFt[c] / P.thE[c];
        L.beta = P.thFt[c] / P.thGf[c];
    }
    return L;
}

// One cell: Cell::IntegrateGeneric (Cell.h:105-119) with the integrand of P.p.kind.
// wantR / wantK select Gradient and/or Hessian0.
static void IntegrateCell(const Problem& P, int c, const double* u, const double* e, const double* kappa,
                          bool wantR, bool wantK, CellResult& out)
{
    const Mesh& m = P.mesh;
    int dim = m.dim, nnC = NumNodes(m.coordType);
    std::vector<double> Xc((size_t)nnC * dim);
    for (int k = 0; k < nnC; k++)
        for (int d = 0; d < dim; d++)
            Xc[k * dim + d] = m.X[(size_t)m.conn[(size_t)c * nnC + k] * dim + d];

    // node values per dof type (ElementFem::ExtractNodeValues, ElementFem.h:33-40)
    std::vector<double> nodeVal[2];
    const double* src[2] = {u, e};
    int nnD[2] = {0, 0};
    for (int t = 0; t < P.ntypes; t++)
    {
        const DofLayer& L = m.dofs[P.layer[t]];
        nnD[t] = NumNodes(L.type);
        for (int a = 0; a < nnD[t]; a++)
        {
            int node = L.conn[(size_t)c * nnD[t] + a];
            for (int comp = 0; comp < L.ncomp; comp++)
                nodeVal[t].push_back(src[t][P.numbers[t][(size_t)node * L.ncomp + comp]]);
        }
    }

    CellLaw law = LawFor(P, c);
    bool first = true;
    for (int ip = 0; ip < P.rule.Num(); ip++)
    {
        const double* xi = P.rule.xi[ip].data();
        double w = P.rule.w[ip];
        double dNc[60];
        DShapeDim(m.coordType, xi, dNc);
        Jacobian jac = MakeJacobian(dim, nnC, Xc.data(), dNc);
        double detJ = jac.det;
        if (g_mutation == 6)
            for (int a = 0; a < dim; a++)
                for (int b = a + 1; b < dim; b++)
                    std::swap(jac.inv[a * dim + b], jac.inv[b * dim + a]);

        // per dof type: N, dN/dX
        std::vector<double> Nv[2];
        Mat dNdX[2], dNdXm[2];
        for (int t = 0; t < P.ntypes; t++)
        {
            int ty = m.dofs[P.layer[t]].type;
            Nv[t].resize(nnD[t]);
            Shape(ty, xi, Nv[t].data());
            double dN[60];
            DShapeDim(ty, xi, dN);
            dNdX[t] = GlobalDerivatives(jac, dN, nnD[t]);
            dNdXm[t] = GlobalDerivativesMag(jac, dN, nnD[t]);
        }
        std::vector<double> Rip[2], Rmag[2];
        Mat Kip[2][2], Kmag[2][2];

        if (P.p.kind == 0 || P.p.kind == 2)
        {
            Mat B = StrainB(dNdX[0]);
            const Mat Bm = StrainB(dNdXm[0]);  // shadow of B
            Mat strainM = Mul(B, Col(nodeVal[0]));
            const double* strain = strainM.v.data();
            Mat strainMag = MulAbs(Bm, Col(nodeVal[0]));
            if (P.p.kind == 0)
            {
                // MomentumBalance::Gradient/Hessian0 (MomentumBalance.h:26-46) with LinearElastic
                Mat C = ElasticC(dim, law.E, P.p.nu, P.p.plane_state);
                if (wantR)
                {
                    Rip[0] = Mul(T(B), Mul(C, strainM)).v;
                    Rmag[0] = MulAbs(T(Bm), MulAbs(C, strainMag)).v;
                }
                if (wantK)
                {
                    Kip[0][0] = Mul(Mul(T(B), C), B);
                    Kmag[0][0] = MulAbs(MulAbs(T(Bm), C), Bm);
                }
            }
            else
            {
                // MomentumBalance<3> + LocalIsotropicDamage<3> (LocalIsotropicDamage.h:52-68,118-128)
                ElasticDamage3 led(law.E, P.p.nu, P.p.damage_application);
                ModMises3 norm(P.p.nu, P.p.k);
                DamageExp dl{law.kappa0, law.beta, P.p.alpha};
                double kh = kappa[(size_t)c + (size_t)ip * m.nCells];
                double eeq = norm.Value(strain);
                double kap = std::max(eeq, kh);
                double omega = dl.Damage(kap);
                LawMag lm = LawMagnitudes(strainMag.v.data(), omega, law.E, P.p.nu, P.p.k, P.p.damage_application,
                                          MisesA(strain, P.p.nu, P.p.k));
                if (wantR)
                {
                    Rip[0] = Mul(T(B), Col(led.Stress(strain, omega))).v;
                    Rmag[0] = MulAbs(T(Bm), Col(std::vector<double>(lm.sig, lm.sig + 6))).v;
                }
                if (wantK)
                {
                    Mat D = led.DstressDstrain(strain, omega);
                    double dw = dl.Derivative(kap);
                    std::vector<double> dk(6, 0.);
                    if (norm.Value(strain) >= kh)
                        dk = norm.Derivative(strain);
                    std::vector<double> dsw = led.DstressDomega(strain);
                    Mat tang(6, 6);
                    for (int i = 0; i < 6; i++)
                        for (int j = 0; j < 6; j++)
                            tang(i, j) = D(i, j) + (dsw[i] * dw) * dk[j];
                    Kip[0][0] = Mul(Mul(T(B), tang), B);
                    Mat tm(6, 6);
                    const bool active = norm.Value(strain) >= kh;
                    for (int i = 0; i < 6; i++)
                        for (int j = 0; j < 6; j++)
                            tm(i, j) = std::abs(D(i, j)) + (active ? lm.dsw[i] * std::abs(dw) * lm.m[j] : 0.);
                    Kmag[0][0] = MulAbs(MulAbs(T(Bm), tm), Bm);
                }
            }
        }
        else
        {
            // GradientDamage<3>::Gradient/Hessian0 (GradientDamage.h:70-135)
            ElasticDamage3 led(law.E, P.p.nu, P.p.damage_application);
            ModMises3 norm(P.p.nu, P.p.k);
            DamageExp dl{law.kappa0, law.beta, P.p.alpha};
            Mat Bd = StrainB(dNdX[0]);
            Mat Be = T(dNdX[1]);  // Nabla::Gradient
            const Mat Bdm = StrainB(dNdXm[0]), Bem = T(dNdXm[1]);  // shadows
            Mat Ne(1, nnD[1]);
            Ne.v = Nv[1];
            double eeq = Mul(Ne, Col(nodeVal[1])).v[0];
            double kh = kappa[(size_t)c + (size_t)ip * m.nCells];
            double kap = g_mutation == 4 ? eeq : std::max(kh, eeq);
            double omega = dl.Damage(kap);
            Mat eeqGrad = Mul(Be, Col(nodeVal[1]));
            Mat strainM = Mul(Bd, Col(nodeVal[0]));
            if (g_mutation == 3)
                for (int i = 3; i < 6; i++)
                    strainM.v[i] *= 0.5;
            const double* strain = strainM.v.data();
            double g, dgdw;
            Interaction(P.p, omega, g, dgdw);
            // the nonlocal length c of R_e and K_ee; mutations 7 and 8 (tests only) change it
            // consistently in both, which the finite-difference check of K cannot see
            const double cNl = g_mutation == 7 ? P.p.c * P.p.c : (g_mutation == 8 ? -P.p.c : P.p.c);
            Mat strainMag = MulAbs(Bdm, Col(nodeVal[0]));
            LawMag lm = LawMagnitudes(strainMag.v.data(), omega, law.E, P.p.nu, P.p.k, P.p.damage_application,
                                      MisesA(strain, P.p.nu, P.p.k));
            if (wantR)
            {
                Rip[0] = Mul(T(Bd), Col(led.Stress(strain, omega))).v;
                Mat r1 = Scale(T(Ne), eeq - norm.Value(strain));
                Mat r2 = Mul(T(Be), Scale(eeqGrad, (g_mutation == 5 ? -1.0 : 1.0) * cNl * g));
                Rip[1] = Add(r1, r2).v;
                Rmag[0] = MulAbs(T(Bdm), Col(std::vector<double>(lm.sig, lm.sig + 6))).v;
                Mat m1 = Scale(Abs(T(Ne)), std::abs(eeq) + lm.eeq);
                Mat m2 = MulAbs(T(Bem), Scale(MulAbs(Bem, Col(nodeVal[1])), P.p.c));
                Rmag[1] = Add(m1, m2).v;
            }
            if (wantK)
            {
                double dkde = eeq >= kh ? 1 : 0;
                double dwdk = dl.Derivative(kap);
                Kip[0][0] = Mul(Mul(T(Bd), led.DstressDstrain(strain, omega)), Bd);
                Mat dn(1, 6);
                dn.v = norm.Derivative(strain);
                Kip[1][0] = Mul(Mul(Scale(T(Ne), g_mutation == 2 ? 1. : -1.), dn), Bd);
                Mat ee = g_mutation == 1 ? Mul(T(Ne), Ne) : Add(Mul(T(Ne), Ne), Scale(Mul(T(Be), Be), cNl * g));
                Mat ee3 = Mul(Scale(Scale(Mul(T(Be), Scale(eeqGrad, cNl)), dgdw), dwdk * dkde), Ne);
                Kip[1][1] = Add(ee, ee3);
                Mat dsw = Col(led.DstressDomega(strain));
                Kip[0][1] = Mul(Mul(T(Bd), Scale(Scale(dsw, dwdk), dkde)), Ne);
                Kmag[0][0] = MulAbs(MulAbs(T(Bdm), led.DstressDstrain(strain, omega)), Bdm);
                Mat mm(1, 6);
                mm.v.assign(lm.m, lm.m + 6);
                Kmag[1][0] = MulAbs(MulAbs(T(Ne), mm), Bdm);
                Kmag[1][1] = Add(MulAbs(T(Ne), Ne), Scale(MulAbs(T(Bem), Bem), P.p.c));
                if (dgdw != 0. && dwdk * dkde != 0.)  // shadow of the g' term
                    Kmag[1][1] = Add(Kmag[1][1], Mul(Scale(MulAbs(T(Bem), MulAbs(Bem, Col(nodeVal[1]))),
                                                           P.p.c * std::abs(dgdw * dwdk)),
                                                     Abs(Ne)));
                Mat dm = Col(std::vector<double>(lm.dsw, lm.dsw + 6));
                Kmag[0][1] = MulAbs(MulAbs(T(Bdm), Scale(dm, std::abs(dwdk) * dkde)), Ne);
            }
        }
        // result += (f * detJ) * w ; first IP assigns ("uninitialized addition")
        for (int t = 0; t < P.ntypes; t++)
        {
            const double sw = std::abs(detJ) * w;
            if (wantR)
            {
                for (double& x : Rip[t])
                    x = (x * detJ) * w;
                for (double& x : Rmag[t])
                    x = x * sw;
                if (first)
                {
                    out.R[t] = Rip[t];
                    out.Rm[t] = Rmag[t];
                }
                else
                    for (size_t i = 0; i < Rip[t].size(); i++)
                    {
                        out.R[t][i] += Rip[t][i];
                        out.Rm[t][i] += Rmag[t][i];
                    }
            }
            if (wantK)
                for (int s = 0; s < P.ntypes; s++)
                {
                    for (double& x : Kip[t][s].v)
                        x = (x * detJ) * w;
                    Kmag[t][s] = Scale(Kmag[t][s], sw);
                    if (first)
                    {
                        out.K[t][s] = Kip[t][s];
                        out.Km[t][s] = Kmag[t][s];
                    }
                    else
                    {
                        out.K[t][s] = Add(out.K[t][s], Kip[t][s]);
                        out.Km[t][s] = Add(out.Km[t][s], Kmag[t][s]);
                    }
                }
        }
        first = false;
    }
}

// Cell::DofNumbering -> ElementFem::GetDofNumbering (ElementFem.h:62-75): node-major, comp-minor
static std::vector<int> CellDofs(const Problem& P, int c, int t)
{
    const DofLayer& L = P.mesh.dofs[P.layer[t]];
    int nn = NumNodes(L.type);
    std::vector<int> o;
    for (int a = 0; a < nn; a++)
        for (int comp = 0; comp < L.ncomp; comp++)
            o.push_back(P.numbers[t][(size_t)L.conn[(size_t)c * nn + a] * L.ncomp + comp]);
    return o;
}

struct Triplet
{
    int i, j;
    double v;
    double m = 0.;  // magnitude scale of the contribution
};

// setFromTriplets-equivalent: bucket by outer index in list order, collapse duplicates
// (summing in list order), then emit sorted. Produces CSR (outer = row) of the block.
static void TripletsToCSR(const std::list<Triplet>& trip, int nrows, std::vector<int>& rp, std::vector<int>& ci,
                          std::vector<double>& v, std::vector<double>* absBound)
{
    std::vector<std::vector<std::pair<int, std::pair<double, double>>>> rows(nrows);
    for (const Triplet& t : trip)
        rows[t.i].push_back({t.j, {t.v, t.m != 0. ? t.m : std::abs(t.v)}});
    rp.assign(nrows + 1, 0);
    ci.clear();
    v.clear();
    if (absBound)
        absBound->clear();
    for (int r = 0; r < nrows; r++)
    {
        auto& lst = rows[r];
        // collapse duplicates in insertion order
        std::vector<std::pair<int, std::pair<double, double>>> uniq;
        std::unordered_map<int, int> pos;
        for (auto& e : lst)
        {
            auto it = pos.find(e.first);
            if (it == pos.end())
            {
                pos[e.first] = (int)uniq.size();
                uniq.push_back(e);
            }
            else
            {
                uniq[it->second].second.first += e.second.first;
                uniq[it->second].second.second += e.second.second;
            }
        }
        std::sort(uniq.begin(), uniq.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (auto& e : uniq)
        {
            ci.push_back(e.first);
            v.push_back(e.second.first);
            if (absBound)
                absBound->push_back(e.second.second);
        }
        rp[r + 1] = (int)ci.size();
        std::vector<std::pair<int, std::pair<double, double>>>().swap(lst);
    }
}

// SimpleAssembler::BuildVector + BuildMatrix (SimpleAssembler.cpp:38-65,100-156), then the
// monolithic export (DofMatrixSparseConvertEigen.h:40-67) as CSR rows [d0;d1] x cols [d0,d1].
static void Assemble(Problem& P, const double* u, const double* e, const double* kappa, bool wantR, bool wantK,
                     int nthreads)
{
    int nt = P.ntypes;
    int off[2] = {0, P.ndof[0]};
    int ntot = P.ndof[0] + (nt > 1 ? P.ndof[1] : 0);
    if (wantR)
    {
        P.R.assign(ntot, 0.);
        P.Rbound.assign(ntot, 0.);
    }
    std::list<Triplet> trip[2][2];
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        std::vector<double> localR(wantR ? ntot : 0, 0.), localB(wantR ? ntot : 0, 0.);
        std::list<Triplet> local[2][2];
#ifdef _OPENMP
#pragma omp for nowait
#endif
        for (int c = 0; c < P.mesh.nCells; c++)
        {
            CellResult cr;
            IntegrateCell(P, c, u, e, kappa, wantR, wantK, cr);
            std::vector<int> num[2];
            for (int t = 0; t < nt; t++)
                num[t] = CellDofs(P, c, t);
            if (wantR)
                for (int t = 0; t < nt; t++)
                    for (size_t i = 0; i < num[t].size(); i++)
                    {
                        localR[off[t] + num[t][i]] += cr.R[t][i];
                        localB[off[t] + num[t][i]] += cr.Rm[t][i];
                    }
            if (wantK)
                for (int t = 0; t < nt; t++)
                    for (int s = 0; s < nt; s++)
                    {
                        const Mat& K = cr.K[t][s];
                        const Mat& Km = cr.Km[t][s];
                        for (size_t i = 0; i < num[t].size(); i++)
                            for (size_t j = 0; j < num[s].size(); j++)
                                local[t][s].push_back({num[t][i], num[s][j], K((int)i, (int)j), Km((int)i, (int)j)});
                    }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            if (wantR)
                for (int i = 0; i < ntot; i++)
                {
                    P.R[i] += localR[i];
                    P.Rbound[i] += localB[i];
                }
            for (int t = 0; t < nt; t++)
                for (int s = 0; s < nt; s++)
                    trip[t][s].splice(trip[t][s].end(), local[t][s]);
        }
    }
    if (!wantK)
        return;
    // per-block setFromTriplets, then the monolithic concatenation
    std::vector<int> brp[2][2], bci[2][2];
    std::vector<double> bv[2][2], bb[2][2];
    for (int t = 0; t < nt; t++)
        for (int s = 0; s < nt; s++)
        {
            TripletsToCSR(trip[t][s], P.ndof[t], brp[t][s], bci[t][s], bv[t][s], &bb[t][s]);
            trip[t][s].clear();
        }
    P.rowPtr.assign(ntot + 1, 0);
    P.col.clear();
    P.val.clear();
    P.bound.clear();
    for (int t = 0; t < nt; t++)
        for (int r = 0; r < P.ndof[t]; r++)
        {
            for (int s = 0; s < nt; s++)
                for (int k = brp[t][s][r]; k < brp[t][s][r + 1]; k++)
                {
                    P.col.push_back(off[s] + bci[t][s][k]);
                    P.val.push_back(bv[t][s][k]);
                    P.bound.push_back(bb[t][s][k]);
                }
            P.rowPtr[off[t] + r + 1] = (int)P.col.size();
        }
}

// ---------------------------------------------------------------------------------------------
// lumped mass and fields at the integration points
// ---------------------------------------------------------------------------------------------
// SimpleAssembler::BuildDiagonallyLumpedMatrix (SimpleAssembler.cpp:67-98) with the integrand
// DynamicMomentumBalance::Hessian2 = N^T N rho (DynamicMomentumBalance.h:22-30) on dof type 0, one
// thread in cell order: per cell the consistent mass (Cell.h:105-119), localDiagonalDof *=
// (fullSum / diagonalSum) / dof.GetNum(), added at the cell's dof numbers. Mb: the same with
// |detJ| (the magnitude scale of each entry).
static void LumpedMass(const Problem& P, double rho, std::vector<double>& M, std::vector<double>& Mb)
{
    const Mesh& m = P.mesh;
    const int dim = m.dim, nnC = NumNodes(m.coordType);
    const DofLayer& L = m.dofs[P.layer[0]];
    const int nnD = NumNodes(L.type), nc = L.ncomp, n = nc * nnD;
    M.assign(P.ndof[0], 0.0);
    Mb.assign(P.ndof[0], 0.0);
    std::vector<double> Xc((size_t)nnC * dim);
    for (int c = 0; c < m.nCells; c++)
    {
        for (int k = 0; k < nnC; k++)
            for (int d = 0; d < dim; d++)
                Xc[k * dim + d] = m.X[(size_t)m.conn[(size_t)c * nnC + k] * dim + d];
        Mat Me, Mm;
        for (int ip = 0; ip < P.rule.Num(); ip++)
        {
            const double* xi = P.rule.xi[ip].data();
            const double w = P.rule.w[ip];
            double dNc[60];
            DShapeDim(m.coordType, xi, dNc);
            Jacobian jac = MakeJacobian(dim, nnC, Xc.data(), dNc);
            std::vector<double> N(nnD);
            Shape(L.type, xi, N.data());
            Mat Nm = NMatrix(N, nc);
            Mat f = Scale(Mul(T(Nm), Nm), rho);
            Mat fm = Scale(Abs(f), std::abs(jac.det) * w);
            for (double& x : f.v)
                x = (x * jac.det) * w;
            if (ip == 0)
            {
                Me = f;
                Mm = fm;
            }
            else
            {
                Me = Add(Me, f);
                Mm = Add(Mm, fm);
            }
        }
        double diagSum = 0.0, fullSum = 0.0;
        for (int i = 0; i < n; i++)
            diagSum += Me(i, i);
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++)
                fullSum += Me(i, j);
        const double scale = (fullSum / diagSum) / nc;
        std::vector<int> dofs = CellDofs(P, c, 0);
        for (int i = 0; i < n; i++)
        {
            M[dofs[i]] += Me(i, i) * scale;
            Mb[dofs[i]] += Mm(i, i) * std::abs(scale);
        }
    }
}

// Cell::Eval (Cell.h:80-93) of one field of the problem's integrand at every (cell, IP), written
// to out[(c * nip + ip) * ncomp + comp]:
//   0 strain  Nabla::Strain (DifferentialOperators.h:24-86) applied to u
//   1 stress  LinearElastic (LinearElastic.h:76-95) or LinearElasticDamage::Stress at omega(kappa)
//             (LinearElasticDamage.h:65-72)
//   2 kappa   GradientDamage::Kappa = max(history, eeq) (GradientDamage.h:126-129); local damage:
//             EvolutionImplicit, max(history, eps_eq(strain))
//   3 damage  DamageLawExponential::Damage(kappa) (DamageLawExponential.h:39-45)
//   4 eps_eq  ModifiedMisesStrainNorm::Value(strain) (ModifiedMisesStrainNorm.h:70-84)
// Returns the number of components (0: field not defined for the integrand).
static int EvalIp(const Problem& P, const double* u, const double* e, const double* kappa, int field,
                  std::vector<double>& out)
{
    const Mesh& m = P.mesh;
    const int dim = m.dim, nnC = NumNodes(m.coordType), nip = P.rule.Num();
    const int nv = dim == 2 ? 3 : 6;
    const int ncomp = field <= 1 ? nv : 1;
    if (field > 1 && P.p.kind == 0)
        return 0;
    out.assign((size_t)m.nCells * nip * ncomp, 0.0);
    std::vector<double> Xc((size_t)nnC * dim);
    for (int c = 0; c < m.nCells; c++)
    {
        for (int k = 0; k < nnC; k++)
            for (int d = 0; d < dim; d++)
                Xc[k * dim + d] = m.X[(size_t)m.conn[(size_t)c * nnC + k] * dim + d];
        std::vector<double> nodeVal[2];
        const double* src[2] = {u, e};
        int nnD[2] = {0, 0};
        for (int t = 0; t < P.ntypes; t++)
        {
            const DofLayer& L = m.dofs[P.layer[t]];
            nnD[t] = NumNodes(L.type);
            for (int a = 0; a < nnD[t]; a++)
            {
                int node = L.conn[(size_t)c * nnD[t] + a];
                for (int comp = 0; comp < L.ncomp; comp++)
                    nodeVal[t].push_back(src[t][P.numbers[t][(size_t)node * L.ncomp + comp]]);
            }
        }
        CellLaw law = LawFor(P, c);
        for (int ip = 0; ip < nip; ip++)
        {
            const double* xi = P.rule.xi[ip].data();
            double dNc[60];
            DShapeDim(m.coordType, xi, dNc);
            Jacobian jac = MakeJacobian(dim, nnC, Xc.data(), dNc);
            const int ty = m.dofs[P.layer[0]].type;
            double dN[60];
            DShapeDim(ty, xi, dN);
            Mat B = StrainB(GlobalDerivatives(jac, dN, nnD[0]));
            Mat strainM = Mul(B, Col(nodeVal[0]));
            const double* strain = strainM.v.data();
            std::vector<double> val;
            if (field == 0)
                val = strainM.v;
            else if (P.p.kind == 0)
                val = Mul(ElasticC(dim, law.E, P.p.nu, P.p.plane_state), strainM).v;
            else
            {
                ModMises3 norm(P.p.nu, P.p.k);
                const double eq = norm.Value(strain);
                double drive = eq;
                if (P.p.kind == 1)
                {
                    std::vector<double> Ne(nnD[1]);
                    Shape(m.dofs[P.layer[1]].type, xi, Ne.data());
                    Mat Nr(1, nnD[1]);
                    Nr.v = Ne;
                    drive = Mul(Nr, Col(nodeVal[1])).v[0];
                }
                const double kh = kappa[(size_t)c + (size_t)ip * m.nCells];
                const double kap = std::max(kh, drive);
                DamageExp dl{law.kappa0, law.beta, P.p.alpha};
                const double omega = dl.Damage(kap);
                if (field == 1)
                    val = ElasticDamage3(law.E, P.p.nu, P.p.damage_application).Stress(strain, omega);
                else if (field == 2)
                    val = {kap};
                else if (field == 3)
                    val = {omega};
                else
                    val = {eq};
            }
            for (int k = 0; k < ncomp; k++)
                out[((size_t)c * nip + ip) * ncomp + k] = val[k];
        }
    }
    return ncomp;
}

// ---------------------------------------------------------------------------------------------
// constraints and the constrained solve
// ---------------------------------------------------------------------------------------------
// Constraints::BuildUnitConstraintMatrix (Constraints.cpp:89-137) for one dof type, dense
// row-major nDofs x nInd: JK numbering (GetJKNumbering, Constraints.cpp:53-86) = independent dofs
// ascending, then the dependent dofs in equation order; a 1 at (jk(i), i) for i < nInd; -coef at
// (dependent dof, reverseJK(term dof)) for every term with |coef| > 1e-18, duplicates summed.
// Returns nInd, or -1 for a dependent dof out of range.
static int UnitConstraintMatrix(int nDofs, int nEq, const int* dep, const int* termPtr, const int* termDof,
                                const double* coef, std::vector<double>& C)
{
    std::vector<char> isDep(nDofs, 0);
    for (int e = 0; e < nEq; e++)
    {
        if (dep[e] < 0 || dep[e] >= nDofs)
            return -1;
        isDep[dep[e]] = 1;
    }
    std::vector<int> jk, rev(nDofs, -1);
    for (int g = 0; g < nDofs; g++)
        if (!isDep[g])
            jk.push_back(g);
    const int nInd = (int)jk.size();
    for (int e = 0; e < nEq; e++)
        jk.push_back(dep[e]);
    for (int i = 0; i < nDofs; i++)
        rev[jk[i]] = i;
    C.assign((size_t)nDofs * nInd, 0.0);
    for (int i = 0; i < nInd; i++)
        C[(size_t)jk[i] * nInd + i] += 1.0;
    for (int e = 0; e < nEq; e++)
        for (int t = termPtr[e]; t < termPtr[e + 1]; t++)
            if (std::abs(coef[t]) > 1.e-18)
                C[(size_t)dep[e] * nInd + rev[termDof[t]]] += -coef[t];
    return nInd;
}

// Dense solve of A x = b (n x n, row-major) by Gaussian elimination with partial pivoting.
static void DenseSolve(int n, std::vector<double> A, std::vector<double> b, std::vector<double>& x)
{
    for (int k = 0; k < n; k++)
    {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (std::abs(A[(size_t)i * n + k]) > std::abs(A[(size_t)p * n + k]))
                p = i;
        if (p != k)
        {
            for (int j = 0; j < n; j++)
                std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
            std::swap(b[k], b[p]);
        }
        const double piv = A[(size_t)k * n + k];
        for (int i = k + 1; i < n; i++)
        {
            const double l = A[(size_t)i * n + k] / piv;
            if (l == 0.0)
                continue;
            for (int j = k; j < n; j++)
                A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
            b[i] -= l * b[k];
        }
    }
    x.assign(n, 0.0);
    for (int i = n - 1; i >= 0; i--)
    {
        double s = b[i];
        for (int j = i + 1; j < n; j++)
            s -= A[(size_t)i * n + j] * x[j];
        x[i] = s / A[(size_t)i * n + i];
    }
}

}  // namespace orc

// =============================================================================================
// C API (ctypes)
// =============================================================================================
using namespace orc;

extern "C" {

int orc_legendre_roots(int n, double* out)
{
    auto r = LegendreRoots(n);
    std::copy(r.begin(), r.end(), out);
    return n;
}

int orc_gauss_1d(int n, double* w, double* x)
{
    std::vector<double> ww, xx;
    Gauss1D(n, ww, xx);
    std::copy(ww.begin(), ww.end(), w);
    std::copy(xx.begin(), xx.end(), x);
    return n;
}

// rule: kind 0 = tensor Gauss (dim, n per dir), kind 1 = tetrahedron (order n)
int orc_rule(int kind, int dim, int n, double* xi /* num x 3 */, double* w)
{
    Rule r = kind == 0 ? TensorGauss(dim, n) : TetRule(n);
    for (int i = 0; i < r.Num(); i++)
    {
        if (xi)
            for (int d = 0; d < 3; d++)
                xi[i * 3 + d] = r.xi[i][d];
        if (w)
            w[i] = r.w[i];
    }
    return r.Num();
}

int orc_shape(int type, const double* xi, double* N, double* dN /* n x dim */)
{
    if (N)
        Shape(type, xi, N);
    if (dN)
        DShapeDim(type, xi, dN);
    return NumNodes(type);
}

int orc_local_coords(int type, double* out /* n x 3 */)
{
    for (int i = 0; i < NumNodes(type); i++)
    {
        auto c = LocalCoords(type, i);
        out[3 * i] = c[0];
        out[3 * i + 1] = c[1];
        out[3 * i + 2] = c[2];
    }
    return NumNodes(type);
}

void orc_jacobian(int dim, int n, const double* X, const double* dN, double* J, double* inv, double* det)
{
    Jacobian j = MakeJacobian(dim, n, X, dN);
    std::memcpy(J, j.J, sizeof(double) * dim * dim);
    std::memcpy(inv, j.inv, sizeof(double) * dim * dim);
    *det = j.det;
}

// laws (3D): out arrays sized 6 / 36
void orc_led_stress(double E, double nu, int app, const double* strain, double omega, double* sigma, double* D,
                    double* dsdw)
{
    ElasticDamage3 l(E, nu, app);
    auto s = l.Stress(strain, omega);
    std::copy(s.begin(), s.end(), sigma);
    Mat d = l.DstressDstrain(strain, omega);
    std::copy(d.v.begin(), d.v.end(), D);
    auto w = l.DstressDomega(strain);
    std::copy(w.begin(), w.end(), dsdw);
}

double orc_mises(double nu, double k, const double* strain, double* deriv)
{
    ModMises3 m(nu, k);
    if (deriv)
    {
        auto d = m.Derivative(strain);
        std::copy(d.begin(), d.end(), deriv);
    }
    return m.Value(strain);
}

// NonlocalInteraction Factor (return value) and Derivative (*deriv) at omega
double orc_interaction(int interaction, double r, double eta, double omega, double* deriv)
{
    Params p{};
    p.interaction = interaction;
    p.interaction_r = r;
    p.interaction_eta = eta;
    double g, dg;
    Interaction(p, omega, g, dg);
    if (deriv)
        *deriv = dg;
    return g;
}

double orc_damage_exp(double k0, double beta, double alpha, double kappa, double* deriv)
{
    DamageExp d{k0, beta, alpha};
    if (deriv)
        *deriv = d.Derivative(kappa);
    return d.Damage(kappa);
}

void orc_elastic_c(int dim, double E, double nu, int planeState, double* C)
{
    Mat c = ElasticC(dim, E, nu, planeState);
    std::copy(c.v.begin(), c.v.end(), C);
}

// ---- meshes ------------------------------------------------------------------------------
void* orc_mesh_quads(int nx, int ny)
{
    return new Mesh(CreateQuads(nx, ny));
}
void* orc_mesh_bricks(int nx, int ny, int nz)
{
    return new Mesh(CreateBricks(nx, ny, nz));
}
void orc_set_mutation(int m)
{
    g_mutation = m;
}

void* orc_mesh_from_arrays(int coordType, int nNodes, const double* X, int nCells, const int* conn)
{
    Mesh* m = new Mesh;
    m->coordType = coordType;
    m->dim = Dim(coordType);
    m->nNodes = nNodes;
    m->nCells = nCells;
    m->X.assign(X, X + (size_t)nNodes * m->dim);
    m->conn.assign(conn, conn + (size_t)nCells * NumNodes(coordType));
    return m;
}
void orc_mesh_free(void* m)
{
    delete (Mesh*)m;
}
// UnitMeshFem::Transform (UnitMeshFem.cpp:117-125) with f(x) = x * s
void orc_mesh_scale(void* mp, double s)
{
    for (double& x : ((Mesh*)mp)->X)
        x = x * s;
}
int orc_mesh_info(void* mp, int* out /* dim, coordType, nCells, nNodes, nLayers */)
{
    Mesh* m = (Mesh*)mp;
    out[0] = m->dim;
    out[1] = m->coordType;
    out[2] = m->nCells;
    out[3] = m->nNodes;
    out[4] = (int)m->dofs.size();
    return 0;
}
void orc_mesh_coords(void* mp, double* X)
{
    Mesh* m = (Mesh*)mp;
    std::copy(m->X.begin(), m->X.end(), X);
}
void orc_mesh_conn(void* mp, int* conn)
{
    Mesh* m = (Mesh*)mp;
    std::copy(m->conn.begin(), m->conn.end(), conn);
}
int orc_mesh_add_dof(void* mp, int type, int ncomp)
{
    try
    {
        return AddDofInterpolation(*(Mesh*)mp, type, ncomp);
    }
    catch (std::exception&)
    {
        return -1;
    }
}
int orc_dof_info(void* mp, int layer, int* out /* type, ncomp, nnodes */)
{
    Mesh* m = (Mesh*)mp;
    const DofLayer& L = m->dofs[layer];
    out[0] = L.type;
    out[1] = L.ncomp;
    out[2] = L.nnodes;
    return 0;
}
void orc_dof_conn(void* mp, int layer, int* conn)
{
    Mesh* m = (Mesh*)mp;
    std::copy(m->dofs[layer].conn.begin(), m->dofs[layer].conn.end(), conn);
}
void orc_dof_numbers(void* mp, int layer, int* num)
{
    auto v = DofNumbers(*(Mesh*)mp, layer);
    std::copy(v.begin(), v.end(), num);
}

// ---- problems ----------------------------------------------------------------------------
// rule: kind 0 tensor Gauss with n points per direction, kind 1 tetrahedron of order n
void* orc_problem(void* mp, const Params* p, int ruleKind, int ruleN, int layer0, int layer1)
{
    Problem* P = new Problem;
    P->mesh = *(Mesh*)mp;
    P->p = *p;
    P->rule = ruleKind == 0 ? TensorGauss(P->mesh.dim, ruleN) : TetRule(ruleN);
    P->layer[0] = layer0;
    P->layer[1] = layer1;
    P->ntypes = layer1 >= 0 ? 2 : 1;
    for (int t = 0; t < P->ntypes; t++)
    {
        P->numbers[t] = DofNumbers(P->mesh, P->layer[t]);
        P->ndof[t] = (int)P->numbers[t].size();
    }
    return P;
}
// Constraints::BuildUnitConstraintMatrix of one dof type, dense row-major nDofs x nInd into C
// (NULL: query nInd); returns nInd or -1.
int orc_unit_constraint_matrix(int nDofs, int nEq, const int* dep, const int* termPtr, const int* termDof,
                               const double* coef, double* C)
{
    std::vector<double> c;
    const int nInd = UnitConstraintMatrix(nDofs, nEq, dep, termPtr, termDof, coef, c);
    if (nInd >= 0 && C)
        std::copy(c.begin(), c.end(), C);
    return nInd;
}

// NuTo::Solve (trial 0) / NuTo::SolveTrialState (trial 1) (Solve.cpp:8-78) on the monolithic system
// of nSlots dof types (n[s] dofs each, rows [slot 0; slot 1]): C block diagonal of the slots' unit
// constraint matrices, Kmod = C^T K C, fmod = C^T f (trial: C^T (f + K delta), delta = rhsNew -
// rhsOld at the dependent dofs), u by a direct solve, x = C u (trial: - delta). Equations of the
// slots are concatenated: nEq[s] each, termPtr global (first entry 0). K in CSR.
// Returns the number of independent dofs, or -1.
int orc_solve_constrained(int nSlots, const int* n, const int* nEq, const int* dep, const int* termPtr,
                          const int* termDof, const double* coef, const int* kRow, const int* kCol, const double* kVal,
                          const double* f, int trial, const double* rhsOld, const double* rhsNew, double* x)
{
    int N = 0;
    for (int s = 0; s < nSlots; s++)
        N += n[s];
    std::vector<std::vector<double>> Cs(nSlots);
    std::vector<int> nInd(nSlots), eqOff(nSlots + 1, 0);
    int M = 0;
    for (int s = 0; s < nSlots; s++)
    {
        eqOff[s + 1] = eqOff[s] + nEq[s];
        // shift the slot's term pointers to start at 0
        std::vector<int> tp(nEq[s] + 1);
        for (int e = 0; e <= nEq[s]; e++)
            tp[e] = termPtr[eqOff[s] + e] - termPtr[eqOff[s]];
        nInd[s] = UnitConstraintMatrix(n[s], nEq[s], dep + eqOff[s], tp.data(), termDof + termPtr[eqOff[s]],
                                       coef + termPtr[eqOff[s]], Cs[s]);
        if (nInd[s] < 0)
            return -1;
        M += nInd[s];
    }
    std::vector<double> C((size_t)N * M, 0.0);  // monolithic, block diagonal
    for (int s = 0, ro = 0, co = 0; s < nSlots; ro += n[s], co += nInd[s], s++)
        for (int i = 0; i < n[s]; i++)
            for (int j = 0; j < nInd[s]; j++)
                C[(size_t)(ro + i) * M + co + j] = Cs[s][(size_t)i * nInd[s] + j];
    std::vector<double> delta(N, 0.0), g(f, f + N);
    if (trial)
    {
        for (int s = 0, ro = 0; s < nSlots; ro += n[s], s++)
            for (int e = 0; e < nEq[s]; e++)
            {
                const int k = eqOff[s] + e;
                delta[ro + dep[k]] = rhsNew[k] - (rhsOld ? rhsOld[k] : 0.0);
            }
        for (int r = 0; r < N; r++)
            for (int k = kRow[r]; k < kRow[r + 1]; k++)
                g[r] += kVal[k] * delta[kCol[k]];
    }
    // KC (N x M), Kmod = C^T (K C), fmod = C^T g
    std::vector<double> KC((size_t)N * M, 0.0);
    for (int r = 0; r < N; r++)
        for (int k = kRow[r]; k < kRow[r + 1]; k++)
        {
            const double v = kVal[k];
            const double* crow = &C[(size_t)kCol[k] * M];
            for (int j = 0; j < M; j++)
                KC[(size_t)r * M + j] += v * crow[j];
        }
    std::vector<double> Kmod((size_t)M * M, 0.0), fmod(M, 0.0);
    for (int r = 0; r < N; r++)
        for (int i = 0; i < M; i++)
        {
            const double c = C[(size_t)r * M + i];
            if (c == 0.0)
                continue;
            fmod[i] += c * g[r];
            for (int j = 0; j < M; j++)
                Kmod[(size_t)i * M + j] += c * KC[(size_t)r * M + j];
        }
    std::vector<double> u;
    DenseSolve(M, Kmod, fmod, u);
    for (int r = 0; r < N; r++)
    {
        double s = 0.0;
        for (int j = 0; j < M; j++)
            s += C[(size_t)r * M + j] * u[j];
        x[r] = s - (trial ? delta[r] : 0.0);
    }
    return M;
}
// lumped mass of dof type 0 (length ndof[0]) and its magnitude scale
void orc_problem_lumped_mass(void* Pp, double rho, double* M, double* Mb)
{
    std::vector<double> m, mb;
    LumpedMass(*(Problem*)Pp, rho, m, mb);
    std::copy(m.begin(), m.end(), M);
    if (Mb)
        std::copy(mb.begin(), mb.end(), Mb);
}
// field values at every (cell, IP); returns the number of components (0: not defined); with out
// NULL only the count
int orc_problem_eval_ip(void* Pp, const double* u, const double* e, const double* kappa, int field, double* out)
{
    const Problem& P = *(Problem*)Pp;
    if (!out)
        return field <= 1 ? (P.mesh.dim == 2 ? 3 : 6) : (P.p.kind == 0 ? 0 : 1);
    std::vector<double> o;
    const int nc = EvalIp(P, u, e, kappa, field, o);
    std::copy(o.begin(), o.end(), out);
    return nc;
}
void orc_problem_free(void* P)
{
    delete (Problem*)P;
}
void orc_problem_set_theta(void* Pp, const double* E, const double* ft, const double* gf)
{
    Problem* P = (Problem*)Pp;
    int n = P->mesh.nCells;
    P->thE.assign(E, E + n);
    P->thFt.assign(ft, ft + n);
    P->thGf.assign(gf, gf + n);
}
int orc_problem_ndof(void* Pp, int t)
{
    return ((Problem*)Pp)->ndof[t];
}
int orc_problem_nip(void* Pp)
{
    return ((Problem*)Pp)->rule.Num();
}
// returns nnz (0 if K not requested)
long long orc_problem_assemble(void* Pp, const double* u, const double* e, const double* kappa, int wantR, int wantK,
                               int nthreads)
{
    Problem* P = (Problem*)Pp;
    try
    {
        Assemble(*P, u, e, kappa, wantR != 0, wantK != 0, nthreads);
    }
    catch (std::exception&)
    {
        return -1;
    }
    return (long long)P->val.size();
}
void orc_problem_R(void* Pp, double* R, double* Rb)
{
    Problem* P = (Problem*)Pp;
    std::copy(P->R.begin(), P->R.end(), R);
    if (Rb)
        std::copy(P->Rbound.begin(), P->Rbound.end(), Rb);
}
void orc_problem_K(void* Pp, int* rp, int* ci, double* v, double* b)
{
    Problem* P = (Problem*)Pp;
    std::copy(P->rowPtr.begin(), P->rowPtr.end(), rp);
    std::copy(P->col.begin(), P->col.end(), ci);
    if (v)
        std::copy(P->val.begin(), P->val.end(), v);
    if (b)
        std::copy(P->bound.begin(), P->bound.end(), b);
}
// one cell's element vector / matrix (dense, monolithic local order [d0 dofs, d1 dofs])
int orc_problem_cell(void* Pp, int c, const double* u, const double* e, const double* kappa, double* Re, double* Ke)
{
    Problem* P = (Problem*)Pp;
    CellResult cr;
    IntegrateCell(*P, c, u, e, kappa, Re != nullptr, Ke != nullptr, cr);
    int n[2] = {0, 0};
    for (int t = 0; t < P->ntypes; t++)
    {
        const DofLayer& L = P->mesh.dofs[P->layer[t]];
        n[t] = NumNodes(L.type) * L.ncomp;
    }
    int tot = n[0] + n[1];
    if (Re)
    {
        int o = 0;
        for (int t = 0; t < P->ntypes; t++)
            for (int i = 0; i < n[t]; i++)
                Re[o++] = cr.R[t][i];
    }
    if (Ke)
    {
        int ro = 0;
        for (int t = 0; t < P->ntypes; t++)
        {
            int co = 0;
            for (int s = 0; s < P->ntypes; s++)
            {
                for (int i = 0; i < n[t]; i++)
                    for (int j = 0; j < n[s]; j++)
                        Ke[(size_t)(ro + i) * tot + co + j] = cr.K[t][s](i, j);
                co += n[s];
            }
            ro += n[t];
        }
    }
    return tot;
}
// cell dof numbering (global numbers per type, concatenated)
int orc_problem_cell_dofs(void* Pp, int c, int* out)
{
    Problem* P = (Problem*)Pp;
    int o = 0;
    for (int t = 0; t < P->ntypes; t++)
        for (int d : CellDofs(*P, c, t))
            out[o++] = d;
    return o;
}

// Mock-cell assembly (test/mechanics/cell/SimpleAssembler.cpp:6-114): cells given directly by
// their local vectors/matrices and numberings; same triplet/setFromTriplets path.
long long orc_mock_assemble(int nCells, int nloc, const int* numbering, const double* Re, const double* Ke, int ndof,
                            double* R, int* rp, int* ci, double* v)
{
    std::vector<double> Rv(ndof, 0.);
    std::list<Triplet> trip;
    for (int c = 0; c < nCells; c++)
    {
        const int* num = numbering + (size_t)c * nloc;
        for (int i = 0; i < nloc; i++)
            Rv[num[i]] += Re[(size_t)c * nloc + i];
        for (int i = 0; i < nloc; i++)
            for (int j = 0; j < nloc; j++)
                trip.push_back({num[i], num[j], Ke[(size_t)c * nloc * nloc + i * nloc + j]});
    }
    std::vector<int> rpv, civ;
    std::vector<double> vv;
    TripletsToCSR(trip, ndof, rpv, civ, vv, nullptr);
    std::copy(Rv.begin(), Rv.end(), R);
    if (rp)
    {
        std::copy(rpv.begin(), rpv.end(), rp);
        std::copy(civ.begin(), civ.end(), ci);
        std::copy(vv.begin(), vv.end(), v);
    }
    return (long long)vv.size();
}

}  // extern "C"
