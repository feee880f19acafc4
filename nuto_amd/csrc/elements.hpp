// Element shape functions and integration rules for the assembly kernels (host side: the
// tables are evaluated once per launch configuration and uploaded; kernels read them as
// wave-uniform data).
//
// Node orders follow NuTo's interpolations so that per-cell dof numberings can be passed
// through unchanged:
//   quad4  InterpolationQuadLinear.cpp:37-52        hex8  InterpolationBrickLinear.cpp:37-57
//   tet4   InterpolationTetrahedronLinear.cpp:36-50  tet10 InterpolationTetrahedronQuadratic.cpp:37-64
//   hex20  InterpolationBrickQuadratic.cpp:37-96
// Integration: Gauss-Legendre tensor rules with xi fastest (IntegrationTypeTensorProduct.cpp:28-38)
// and the tabulated tetrahedron rules of order 1 and 2 (IntegrationTypeTetrahedron.cpp:3-10).
#pragma once

#include <cmath>
#include <vector>

namespace nuto_amd
{

enum ElemType : int
{
    kQuad4 = 0,
    kHex8 = 1,
    kTet4 = 2,
    kTet10 = 3,
    kHex20 = 4
};

inline int NodesOf(int t)
{
    static const int n[] = {4, 8, 4, 10, 20};
    return (t >= 0 && t <= 4) ? n[t] : 0;
}
inline int DimOf(int t)
{
    return t == kQuad4 ? 2 : 3;
}
// shape family: 0 quadrilateral, 1 hexahedron, 2 tetrahedron
inline int FamilyOf(int t)
{
    return t == kQuad4 ? 0 : (t == kHex8 || t == kHex20) ? 1 : 2;
}

// natural coordinates of the element nodes
inline void NodeCoords(int t, int a, double x[3])
{
    static const signed char hexs[20][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1},
                                            {1, -1, 1},   {1, 1, 1},   {-1, 1, 1}, {0, -1, -1}, {1, 0, -1},
                                            {0, 1, -1},   {-1, 0, -1}, {-1, -1, 0}, {1, -1, 0}, {1, 1, 0},
                                            {-1, 1, 0},   {0, -1, 1},  {1, 0, 1},  {0, 1, 1},  {-1, 0, 1}};
    static const double tets[10][3] = {{0, 0, 0},     {1, 0, 0},   {0, 1, 0},   {0, 0, 1},     {0.5, 0, 0},
                                       {0.5, 0.5, 0}, {0, 0.5, 0}, {0, 0, 0.5}, {0, 0.5, 0.5}, {0.5, 0, 0.5}};
    x[0] = x[1] = x[2] = 0.0;
    if (t == kQuad4)
    {
        static const signed char q[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
        x[0] = q[a][0];
        x[1] = q[a][1];
    }
    else if (FamilyOf(t) == 1)
    {
        for (int d = 0; d < 3; d++)
            x[d] = hexs[a][d];
    }
    else
    {
        for (int d = 0; d < 3; d++)
            x[d] = tets[a][d];
    }
}

// N and dN/dxi (row-major n x 3; the third column is zero in 2D)
inline void EvalShape(int t, const double* xi, double* N, double* dN)
{
    const int n = NodesOf(t);
    double nc[3];
    if (t == kQuad4 || t == kHex8)
    {
        // products of 1D linear factors (1 + s_d xi_d) / 2
        const int dim = DimOf(t);
        for (int a = 0; a < n; a++)
        {
            NodeCoords(t, a, nc);
            double f[3], df[3];
            for (int d = 0; d < 3; d++)
            {
                if (d < dim)
                {
                    f[d] = 0.5 * (1.0 + nc[d] * xi[d]);
                    df[d] = 0.5 * nc[d];
                }
                else
                {
                    f[d] = 1.0;
                    df[d] = 0.0;
                }
            }
            N[a] = f[0] * f[1] * f[2];
            dN[3 * a + 0] = df[0] * f[1] * f[2];
            dN[3 * a + 1] = f[0] * df[1] * f[2];
            dN[3 * a + 2] = dim == 3 ? f[0] * f[1] * df[2] : 0.0;
        }
        return;
    }
    if (t == kTet4 || t == kTet10)
    {
        // barycentric L0 = 1 - r - s - t, L1 = r, L2 = s, L3 = t
        const double L[4] = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
        const double dL[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        if (t == kTet4)
        {
            for (int a = 0; a < 4; a++)
            {
                N[a] = L[a];
                for (int d = 0; d < 3; d++)
                    dN[3 * a + d] = dL[a][d];
            }
            return;
        }
        for (int a = 0; a < 4; a++)
        {
            N[a] = L[a] * (2.0 * L[a] - 1.0);
            for (int d = 0; d < 3; d++)
                dN[3 * a + d] = (4.0 * L[a] - 1.0) * dL[a][d];
        }
        // edge nodes 4..9 sit between these vertex pairs
        static const int edge[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {2, 3}, {1, 3}};
        for (int k = 0; k < 6; k++)
        {
            const int i = edge[k][0], j = edge[k][1];
            N[4 + k] = 4.0 * L[i] * L[j];
            for (int d = 0; d < 3; d++)
                dN[3 * (4 + k) + d] = 4.0 * (dL[i][d] * L[j] + L[i] * dL[j][d]);
        }
        return;
    }
    // hex20 serendipity
    for (int a = 0; a < 20; a++)
    {
        NodeCoords(t, a, nc);
        const double r = xi[0], s = xi[1], u = xi[2];
        const double pr = 1.0 + nc[0] * r, ps = 1.0 + nc[1] * s, pu = 1.0 + nc[2] * u;
        if (a < 8)
        {
            const double q = nc[0] * r + nc[1] * s + nc[2] * u - 2.0;
            N[a] = 0.125 * pr * ps * pu * q;
            dN[3 * a + 0] = 0.125 * nc[0] * ps * pu * (q + pr);
            dN[3 * a + 1] = 0.125 * nc[1] * pr * pu * (q + ps);
            dN[3 * a + 2] = 0.125 * nc[2] * pr * ps * (q + pu);
        }
        else
        {
            // the coordinate that is zero at the node carries the bubble (1 - x^2)
            double f[3], df[3];
            const double x[3] = {r, s, u};
            for (int d = 0; d < 3; d++)
            {
                if (nc[d] == 0.0)
                {
                    f[d] = 1.0 - x[d] * x[d];
                    df[d] = -2.0 * x[d];
                }
                else
                {
                    f[d] = 1.0 + nc[d] * x[d];
                    df[d] = nc[d];
                }
            }
            N[a] = 0.25 * f[0] * f[1] * f[2];
            dN[3 * a + 0] = 0.25 * df[0] * f[1] * f[2];
            dN[3 * a + 1] = 0.25 * f[0] * df[1] * f[2];
            dN[3 * a + 2] = 0.25 * f[0] * f[1] * df[2];
        }
    }
}

struct Rule
{
    std::vector<double> xi;  // num x 3
    std::vector<double> w;
    int Num() const
    {
        return (int)w.size();
    }
};

// Gauss-Legendre points/weights on [-1, 1] (IntegrationTypeTensorProduct.cpp, Quadrature.cpp:26-39)
// for n = 1..3 in closed form (the roots of P_n); these equal the Newton fixed point of the reference
// rule to the last bit for n = 1, 2 and to 1 ulp for 3. For n = 4..kMaxGauss: Newton's method on P_n
// from x0 = cos(pi (i + 3/4) / (n + 1/2)) (the textbook start, Abramowitz & Stegun 22.16.6), P_n and
// P_n' by the three-term recurrence, until the step stops changing x; w = 2 / ((1 - x^2) P_n'(x)^2).
// Points ascending, the negative half mirrored from the positive one (exact symmetry).
constexpr int kMaxGauss = 10;

inline void LegendreAt(int n, double x, double& p, double& dp)
{
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; k++)
    {
        const double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    p = p1;                                   // P_n (n >= 1)
    dp = n * (x * p1 - p0) / (x * x - 1.0);   // P_n' from P_n and P_{n-1} (x strictly inside (-1, 1))
}

inline bool Gauss1D(int n, std::vector<double>& x, std::vector<double>& w)
{
    if (n >= 4 && n <= kMaxGauss)
    {
        x.assign(n, 0.0);
        w.assign(n, 0.0);
        for (int i = 0; i < (n + 1) / 2; i++)
        {
            // the i-th largest root
            double r = std::cos(3.14159265358979323846 * (i + 0.75) / (n + 0.5));
            for (int it = 0; it < 100; it++)
            {
                double p, dp;
                LegendreAt(n, r, p, dp);
                const double rn = r - p / dp;
                if (rn == r)
                    break;
                r = rn;
            }
            if (n % 2 == 1 && i == n / 2)
                r = 0.0;  // the middle root of an odd rule
            double p, dp;
            LegendreAt(n, r, p, dp);
            const double wt = 2.0 / ((1.0 - r * r) * dp * dp);
            x[n - 1 - i] = r;
            x[i] = -r;
            w[n - 1 - i] = wt;
            w[i] = wt;
        }
        return true;
    }
    switch (n)
    {
    case 1:
        x = {0.0};
        w = {2.0};
        return true;
    case 2:
    {
        const double g = std::sqrt(1.0 / 3.0);
        x = {-g, g};
        w = {1.0, 1.0};
        return true;
    }
    case 3:
    {
        const double g = std::sqrt(0.6);
        x = {-g, 0.0, g};
        w = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
        return true;
    }
    default:
        return false;
    }
}

// kind 0: tensor Gauss with n points per direction in `dim` dimensions (xi fastest)
// kind 1: tetrahedron rule of order n (1: centroid, 2: four points); the digits are the
//         reference table entries (IntegrationTypeTetrahedron.cpp:4-10)
inline bool MakeRule(int kind, int dim, int n, Rule& r)
{
    r.xi.clear();
    r.w.clear();
    if (kind == 0)
    {
        std::vector<double> x, w;
        if (!Gauss1D(n, x, w))
            return false;
        int num = 1;
        for (int d = 0; d < dim; d++)
            num *= n;
        for (int i = 0; i < num; i++)
        {
            double p[3] = {0, 0, 0};
            double wt = 1.0;
            int rem = i;
            for (int d = 0; d < dim; d++)
            {
                const int idx = rem % n;
                rem /= n;
                p[d] = x[idx];
                wt *= w[idx];
            }
            r.xi.insert(r.xi.end(), p, p + 3);
            r.w.push_back(wt);
        }
        return true;
    }
    if (kind == 1 && n == 1)
    {
        r.xi = {0.25, 0.25, 0.25};
        r.w = {0.16666666666666663};
        return true;
    }
    if (kind == 1 && n == 2)
    {
        const double a = 0.13819660112501048, b = 0.58541019662496852, w = 0.041666666666666623;
        r.xi = {a, a, a, b, a, a, a, b, a, a, a, b};
        r.w = {w, w, w, w};
        return true;
    }
    return false;
}

}  // namespace nuto_amd
