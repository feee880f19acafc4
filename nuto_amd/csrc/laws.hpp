// Integration-point physics of the three integrands on the path, in a B-sparse closed form.
//
// The reference forms the dense products B^T * D * B per integration point
// (MomentumBalance.h:26-46, GradientDamage.h:70-120). Here the isotropic structure of D is used
// instead: with g_a = dN_a/dX,
//   B_a^T (alpha_v Iv + alpha_d Pinv Id) B_b
//     = (alpha_v - alpha_d)/3 g_a g_b^T + alpha_d/2 g_b g_a^T + alpha_d/2 (g_a . g_b) I,
// which is LinearElasticDamage::DstressDstrain (LinearElasticDamage.h:74-82) sandwiched between
// Nabla::Strain operators (DifferentialOperators.h:49-86). Vectors B_a^T v use the Voigt order
// [xx, yy, zz, yz, zx, xy]. All arithmetic is FP64.
#pragma once

#include <hip/hip_runtime.h>

namespace nuto_amd
{

struct LawParams
{
    double E, nu;
    double k, kappa0, beta, alpha, c;
    int damage_application;  // 0 FULL, 1 UNILATERAL
    int plane_state;         // 0 PLANE_STRESS, 1 PLANE_STRAIN (2D linear elasticity)
    // derived once per launch (see MakeDerived): 3K = cK E, 2G = cG E, modified Mises K1, K2
    double cK, cG, K1, K2;
    int interaction;         // gradient damage: 0 NonlocalInteraction::Constant, 1 Decreasing
    double iR, iEta;         // Decreasing(R, eta)
    // derived: c1 / c2 of the Gram form of the isotropic tangents (cell_gram.hpp), independent of E:
    // linear elasticity lambda / mu; damage with FULL application ((3K - 2G) / 3) / G
    double rLE, rLD;
};

// NonlocalInteraction::Constant / Decreasing (GradientDamage.h:150-184): the factor g(omega) of the
// nonlocal term and its derivative dg/domega
__device__ __host__ inline void NonlocalFactor(const LawParams& P, double omega, double& g, double& dg)
{
    if (P.interaction == 0)
    {
        g = 1.0;
        dg = 0.0;
        return;
    }
    const double e = exp(-P.iEta * omega), eInf = exp(-P.iEta);
    g = ((1. - P.iR) * e + P.iR - eInf) / (1. - eInf);
    dg = ((1. - P.iR) * e) / (1. - eInf) * -P.iEta;
}

inline LawParams MakeDerived(LawParams p)
{
    p.cK = 3.0 / (3.0 - 6.0 * p.nu);
    p.cG = 2.0 / (2.0 + 2.0 * p.nu);
    p.K1 = (p.k - 1.0) / (2.0 * p.k * (1.0 - 2.0 * p.nu));
    p.K2 = 3.0 / (p.k * (1.0 + p.nu) * (1.0 + p.nu));
    p.rLE = 2.0 * p.nu / (1.0 - 2.0 * p.nu);
    p.rLD = (p.cK - p.cG) / (1.5 * p.cG);
    return p;
}

// B_a^T v for a 3D strain-type 6-vector v: component i collects the rows of B_a that act on u_i
__device__ __host__ inline void BTv3(const double g[3], const double v[6], double o[3])
{
    o[0] = g[0] * v[0] + g[2] * v[4] + g[1] * v[5];
    o[1] = g[1] * v[1] + g[2] * v[3] + g[0] * v[5];
    o[2] = g[2] * v[2] + g[1] * v[3] + g[0] * v[4];
}

// DamageLawExponential (DamageLawExponential.h:39-52)
__device__ __host__ inline void DamageExp(double kappa, double k0, double beta, double alpha, double& w, double& dw)
{
    if (kappa < k0)
    {
        w = 0.0;
        dw = 0.0;
        return;
    }
    const double ex = exp(beta * (k0 - kappa));
    const double ik = 1.0 / kappa;
    const double r = k0 * ik;
    w = 1.0 - r * (1.0 - alpha + alpha * ex);
    dw = r * ((ik + beta) * alpha * ex + (1.0 - alpha) * ik);
}

// ModifiedMisesStrainNorm<3> value and derivative (ModifiedMisesStrainNorm.h:70-84,136-146,231-268)
__device__ __host__ inline double Mises3(const double e[6], double nu, double k, double d[6])
{
    const double K1 = (k - 1.0) / (2.0 * k * (1.0 - 2.0 * nu));
    const double K2 = 3.0 / (k * (1.0 + nu) * (1.0 + nu));
    const double I1 = e[0] + e[1] + e[2];
    const double a = e[0] - e[1], b = e[1] - e[2], cc = e[2] - e[0];
    const double J2 = (a * a + b * b + cc * cc) * (1.0 / 6.0) + 0.25 * (e[3] * e[3] + e[4] * e[4] + e[5] * e[5]);
    const double A = sqrt(K1 * K1 * I1 * I1 + K2 * J2);
    if (A == 0.0)
    {
        d[0] = d[1] = d[2] = K1;
        d[3] = d[4] = d[5] = 0.0;
    }
    else
    {
        const double h = 0.5 / A;
        const double t = 2.0 * K1 * K1 * I1;
        const double K23 = K2 * (1.0 / 3.0);
        d[0] = K1 + h * (t + K23 * (2.0 * e[0] - e[1] - e[2]));
        d[1] = K1 + h * (t + K23 * (2.0 * e[1] - e[0] - e[2]));
        d[2] = K1 + h * (t + K23 * (2.0 * e[2] - e[0] - e[1]));
        d[3] = h * K2 * 0.5 * e[3];
        d[4] = h * K2 * 0.5 * e[4];
        d[5] = h * K2 * 0.5 * e[5];
    }
    return K1 * I1 + A;
}

// LinearElasticDamage<3>::Stress and DstressDomega (LinearElasticDamage.h:65-91)
struct DamageStress3
{
    double alpha_v, alpha_d;  // tangent = alpha_v Iv + alpha_d Pinv Id
    double sig[6], dsdw[6];
};

// with the moduli 3K = 3E / (3 - 6 nu) and 2G = 2E / (2 + 2 nu) given
__device__ __host__ inline void ElasticDamage3K(const double e[6], double omega, double m3K, double m2G, int app,
                                                DamageStress3& o)
{
    const double eV = (e[0] + e[1] + e[2]) * (1.0 / 3.0);
    const double H = (app == 1 && eV < 0.0) ? 0.0 : 1.0;
    const double sV = (1.0 - omega * H) * m3K * eV;
    const double d2G = (1.0 - omega) * m2G;
    for (int i = 0; i < 3; i++)
    {
        const double dev = e[i] - eV;
        o.sig[i] = sV + d2G * dev;
        o.dsdw[i] = -m3K * eV * H - m2G * dev;
    }
    for (int i = 3; i < 6; i++)
    {
        o.sig[i] = d2G * 0.5 * e[i];
        o.dsdw[i] = -m2G * 0.5 * e[i];
    }
    o.alpha_v = m3K * (1.0 - omega * H);
    o.alpha_d = d2G;
}

__device__ __host__ inline void ElasticDamage3(const double e[6], double omega, double E, double nu, int app,
                                               DamageStress3& o)
{
    ElasticDamage3K(e, omega, 3.0 * E / (3.0 - 6.0 * nu), 2.0 * E / (2.0 + 2.0 * nu), app, o);
}

}  // namespace nuto_amd
