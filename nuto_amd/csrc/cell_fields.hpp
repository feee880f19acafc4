// Outputs of a cell group besides R and K: the diagonally lumped mass matrix and fields at the
// integration points.
//
// kLumpedMass: SimpleAssembler::BuildDiagonallyLumpedMatrix (SimpleAssembler.cpp:67-98) with the
//   integrand DynamicMomentumBalance::Hessian2 = N^T N rho (DynamicMomentumBalance.h:22-30) on dof
//   type slot 0. Row-owner form like the assembly kernel: one lane per dof node, its incident cells
//   in group order, so the per-node sum runs in cell order without atomics.
// kEvalIp: Cell::Eval (Cell.h:80-93) of the integrand's fields, one lane per (cell, IP).
#pragma once

#include <hip/hip_runtime.h>

#include "assemble_rows.hpp"

namespace nuto_amd
{

// fields of nuto_eval_ip (values per IP: strain and stress NV = 3 (2D) or 6 (3D), the others 1)
enum IpField : int
{
    kIpStrain = 0,    // B u, engineering Voigt (Nabla::Strain, DifferentialOperators.h:24-86)
    kIpStress = 1,    // the law's stress at that strain (and damage)
    kIpKappa = 2,     // GradientDamage::Kappa = max(history, eeq) (GradientDamage.h:126-129);
                      // local damage: max(history, eps_eq(strain)) (EvolutionImplicit)
    kIpDamage = 3,    // DamageLawExponential::Damage(kappa) (DamageLawExponential.h:39-45)
    kIpEqStrain = 4,  // ModifiedMisesStrainNorm::Value(strain) (ModifiedMisesStrainNorm.h:70-84)
};

// J = X dN/dxi of the coordinate element of `cell` at one IP, its inverse and determinant
// (Jacobian.h:16-48)
template <int DIM, int NNC>
__device__ inline double CellJacobian(const RowArgs& A, int cell, const double* dNc, double inv[DIM][DIM])
{
    double X[NNC][DIM];
#pragma unroll
    for (int k = 0; k < NNC; k++)
    {
        const int n = A.cconn[(size_t)cell * NNC + k];
        X[k][0] = A.x[n];
        X[k][1] = A.y[n];
        if constexpr (DIM == 3)
            X[k][2] = A.z[n];
    }
    double J[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int j = 0; j < DIM; j++)
        {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < NNC; k++)
                s += X[k][i] * dNc[k * DIM + j];
            J[i][j] = s;
        }
    return Inverse<DIM>(J, inv);
}

// Lumped mass of dof node p: over the incident cells in group order, the cell's consistent-mass
// diagonal entry m_p = sum_ip ((N_p N_p) rho detJ) w (first IP assigns, Cell.h:105-119), times
// (sum of all entries of the cell's block / sum of its diagonal) / ncomp. The block has ncomp
// identical copies of the nodal mass matrix, so the two sums are ncomp sum_ip ((sum_a N_a)^2 rho
// detJ) w and ncomp sum_a m_a, and their ratio needs no ncomp. Every component of p gets the
// same value.
template <int DIM, int NNC, int NND, int NIP>
__global__ void __launch_bounds__(64) kLumpedMass(const RowArgs A, double rho, double* __restrict__ M)
{
    constexpr int S = 1 + NNC * DIM + NND + NND * DIM;
    const int p = blockIdx.x * 64 + threadIdx.x;
    if (p >= A.nNodes)
        return;
    double acc = 0.0;
    for (int r = A.incPtr[p]; r < A.incPtr[p + 1]; ++r)
    {
        const int cell = A.incCell[r];
        const int la = A.incLocal[r];
        double m[NND];
        double full = 0.0;
        for (int ip = 0; ip < NIP; ip++)
        {
            const double* t = A.tab + ip * S;
            const double w = t[0];
            const double* dNc = t + 1;
            const double* Nd = dNc + NNC * DIM;
            double inv[DIM][DIM];
            const double detJ = CellJacobian<DIM, NNC>(A, cell, dNc, inv);
            double sN = 0.0;
#pragma unroll
            for (int b = 0; b < NND; b++)
            {
                const double v = (((Nd[b] * Nd[b]) * rho) * detJ) * w;
                m[b] = ip == 0 ? v : m[b] + v;
                sN += Nd[b];
            }
            const double f = (((sN * sN) * rho) * detJ) * w;
            full = ip == 0 ? f : full + f;
        }
        double diag = 0.0;
#pragma unroll
        for (int b = 0; b < NND; b++)
            diag += m[b];
        double mp = 0.0;
#pragma unroll
        for (int b = 0; b < NND; b++)
            mp = b == la ? m[b] : mp;
        acc += mp * ((full / diag) / DIM);
    }
#pragma unroll
    for (int c = 0; c < DIM; c++)
        M[(size_t)p * DIM + c] = acc;
}

// Field `field` at every (cell, IP), written to out[(cell * NIP + ip) * ncomp + comp] (cells in
// group order, IPs in rule order).
template <int KIND, int DIM, int NNC, int NND, int NIP>
__global__ void __launch_bounds__(256) kEvalIp(const RowArgs A, int field, double* __restrict__ out)
{
    constexpr int S = 1 + NNC * DIM + NND + NND * DIM;
    constexpr int NV = DIM == 2 ? 3 : 6;
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)A.nCells * NIP)
        return;
    const int cell = (int)(gid / NIP), ip = (int)(gid % NIP);
    const double* t = A.tab + ip * S;
    const double* dNc = t + 1;
    const double* Nd = dNc + NNC * DIM;
    const double* dNd = Nd + NND;
    double inv[DIM][DIM];
    CellJacobian<DIM, NNC>(A, cell, dNc, inv);
    double eps[NV];
#pragma unroll
    for (int i = 0; i < NV; i++)
        eps[i] = 0.0;
    double eN = 0.0;
    for (int b = 0; b < NND; b++)
    {
        const int q = A.dconn[(size_t)cell * NND + b];
        double g[DIM];
        Grad<DIM, NND>(dNd, b, inv, g);
        if constexpr (DIM == 2)
        {
            const double u0 = A.u[(size_t)q * 2], u1 = A.u[(size_t)q * 2 + 1];
            eps[0] += g[0] * u0;
            eps[1] += g[1] * u1;
            eps[2] += g[1] * u0 + g[0] * u1;
        }
        else
        {
            const double u0 = A.u[(size_t)q * 3], u1 = A.u[(size_t)q * 3 + 1], u2 = A.u[(size_t)q * 3 + 2];
            eps[0] += g[0] * u0;
            eps[1] += g[1] * u1;
            eps[2] += g[2] * u2;
            eps[3] += g[2] * u1 + g[1] * u2;
            eps[4] += g[2] * u0 + g[0] * u2;
            eps[5] += g[1] * u0 + g[0] * u1;
        }
        if constexpr (KIND == kGradientDamage)
            eN += Nd[b] * A.e[q];
    }
    double* o = out + (size_t)gid * (field <= kIpStress ? NV : 1);
    if (field == kIpStrain)
    {
#pragma unroll
        for (int i = 0; i < NV; i++)
            o[i] = eps[i];
        return;
    }
    const LawParams& P = A.P;
    const double E = A.thE ? A.thE[cell] : P.E;
    if constexpr (KIND == kLinearElastic)
    {
        // field == kIpStress (the caller rejects the damage fields for linear elasticity)
        if constexpr (DIM == 2)
        {
            double C11, C12, C33;
            if (P.plane_state == 0)
            {
                const double f = E / (1.0 - P.nu * P.nu);
                C11 = f;
                C12 = f * P.nu;
                C33 = f * 0.5 * (1.0 - P.nu);
            }
            else
            {
                const double f = E / ((1.0 + P.nu) * (1.0 - 2.0 * P.nu));
                C11 = f * (1.0 - P.nu);
                C12 = f * P.nu;
                C33 = E / (2.0 * (1.0 + P.nu));
            }
            o[0] = C11 * eps[0] + C12 * eps[1];
            o[1] = C12 * eps[0] + C11 * eps[1];
            o[2] = C33 * eps[2];
        }
        else
        {
            const double f = E / ((1.0 + P.nu) * (1.0 - 2.0 * P.nu));
            const double lam = f * P.nu, mu = E / (2.0 * (1.0 + P.nu));
            const double tr = eps[0] + eps[1] + eps[2];
            for (int i = 0; i < 3; i++)
                o[i] = lam * tr + 2.0 * mu * eps[i];
            for (int i = 3; i < 6; i++)
                o[i] = mu * eps[i];
        }
    }
    else if constexpr (DIM == 3)
    {
        double d[6];
        const double eq = Mises3(eps, P.nu, P.k, d);
        if (field == kIpEqStrain)
        {
            o[0] = eq;
            return;
        }
        const int cid = A.cellId ? A.cellId[cell] : cell;
        const double kh = A.kappa[(size_t)cid + (size_t)ip * A.kappaLd];
        const double drive = KIND == kGradientDamage ? eN : eq;
        const double kap = kh < drive ? drive : kh;
        if (field == kIpKappa)
        {
            o[0] = kap;
            return;
        }
        const double k0 = A.thE ? A.thFt[cell] / E : P.kappa0;
        const double beta = A.thE ? A.thFt[cell] / A.thGf[cell] : P.beta;
        double om, dom;
        DamageExp(kap, k0, beta, P.alpha, om, dom);
        if (field == kIpDamage)
        {
            o[0] = om;
            return;
        }
        DamageStress3 ds;
        ElasticDamage3(eps, om, E, P.nu, P.damage_application, ds);
#pragma unroll
        for (int i = 0; i < 6; i++)
            o[i] = ds.sig[i];
    }
}

}  // namespace nuto_amd
