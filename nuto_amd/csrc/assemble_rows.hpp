// Row-owner assembly kernel ("gather" form of SimpleAssembler::BuildVector/BuildMatrix).
//
// One lane owns one dof node p (all rows of p in the monolithic CSR and its R entries). It walks
// the cells incident to p in group order (SimpleAssembler.cpp:49 / :113 cell order), integrates
// the row block of p for each of them (Cell.h:108-117: IPs in order, weight detJ * w) and adds it
// into the CSR rows of p. A nonzero is touched only by its row owner, so the result is
// deterministic without atomics, and contributions of different cells to one nonzero are summed
// in cell order, the order setFromTriplets uses (SimpleAssembler.cpp:149-155).
//
// Generic over element types; the hex8 gradient-damage hot path has its own tiled kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "laws.hpp"

namespace nuto_amd
{

enum Kind : int
{
    kLinearElastic = 0,
    kGradientDamage = 1,
    kLocalDamage = 2
};

struct RowArgs
{
    const double* x;
    const double* y;
    const double* z;
    const int* cconn;  // n_cells x NNC coordinate node ids
    const int* dconn;  // n_cells x NND dof node ranks (slot 0; slot 1 shares them)
    const int* cellId;
    int nNodes, nCells, kappaLd, nU;
    const int* incPtr;
    const int* incCell;
    const unsigned char* incLocal;
    const unsigned short* slot;
    const int* adjPtr;
    const int* rowPtr;
    const double* u;
    const double* e;
    const double* kappa;
    const double* thE;
    const double* thFt;
    const double* thGf;
    const double* tab;  // per IP: w, dNc[NNC][DIM], N[NND], dNd[NND][DIM]
    LawParams P;
    double* R;
    double* K;
};

template <int DIM>
__device__ inline double Inverse(const double J[DIM][DIM], double inv[DIM][DIM])
{
    if constexpr (DIM == 2)
    {
        const double det = J[0][0] * J[1][1] - J[1][0] * J[0][1];
        const double id = 1.0 / det;
        inv[0][0] = J[1][1] * id;
        inv[0][1] = -J[0][1] * id;
        inv[1][0] = -J[1][0] * id;
        inv[1][1] = J[0][0] * id;
        return det;
    }
    else
    {
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        const double id = 1.0 / det;
        inv[0][0] = c00 * id;
        inv[1][0] = c01 * id;
        inv[2][0] = c02 * id;
        inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
        inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
        inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
        inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
        inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
        inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
        return det;
    }
}

template <int DIM, int NND>
__device__ inline void Grad(const double* dNd, int b, const double inv[DIM][DIM], double g[DIM])
{
#pragma unroll
    for (int j = 0; j < DIM; j++)
    {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; i++)
            s += dNd[b * DIM + i] * inv[i][j];
        g[j] = s;
    }
}

// Rows of dof node p (p < pEnd <= A.nNodes): one lane per node, incident cells in group order.
template <int KIND, int DIM, int NNC, int NND, int NIP, int BH, bool WR, bool WK>
__device__ inline void AssembleRowsBody(const RowArgs& A, int p, int pEnd)
{
    constexpr int F = KIND == kGradientDamage ? 4 : DIM;
    constexpr int S = 1 + NNC * DIM + NND + NND * DIM;
    if (p >= pEnd)
        return;
    const int beg = A.incPtr[p], end = A.incPtr[p + 1];
    const int nA = A.adjPtr[p + 1] - A.adjPtr[p];
    const int L = nA * F;
    const int ru0 = WK ? A.rowPtr[p * DIM] : 0;
    const int re0 = (WK && KIND == kGradientDamage) ? A.rowPtr[A.nU + p] : 0;
    const LawParams& P = A.P;

    double racc[F];
#pragma unroll
    for (int i = 0; i < F; i++)
        racc[i] = 0.0;

    for (int r = beg; r < end; ++r)
    {
        const int cell = A.incCell[r];
        const int la = A.incLocal[r];
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
        double U[NND][DIM];
        double EE[KIND == kGradientDamage ? NND : 1];
#pragma unroll
        for (int b = 0; b < NND; b++)
        {
            const int q = A.dconn[(size_t)cell * NND + b];
#pragma unroll
            for (int c = 0; c < DIM; c++)
                U[b][c] = A.u[(size_t)q * DIM + c];
            if constexpr (KIND == kGradientDamage)
                EE[b] = A.e[q];
        }
        double E = P.E, k0 = P.kappa0, beta = P.beta;
        if (A.thE)
        {
            E = A.thE[cell];
            k0 = A.thFt[cell] / E;
            beta = A.thFt[cell] / A.thGf[cell];
        }
        const int cid = A.cellId ? A.cellId[cell] : cell;

#pragma unroll
        for (int b0 = 0; b0 < NND; b0 += BH)
        {
            constexpr int NB = BH;
            double kacc[NB][F * F];
#pragma unroll
            for (int bb = 0; bb < NB; bb++)
#pragma unroll
                for (int i = 0; i < F * F; i++)
                    kacc[bb][i] = 0.0;

            for (int ip = 0; ip < NIP; ip++)
            {
                const double* t = A.tab + ip * S;
                const double w = t[0];
                const double* dNc = t + 1;
                const double* Nd = dNc + NNC * DIM;
                const double* dNd = Nd + NND;
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
                double inv[DIM][DIM];
                const double detJ = Inverse<DIM>(J, inv);
                const double s = detJ * w;

                // strain (Voigt 2D [xx,yy,xy], 3D [xx,yy,zz,yz,zx,xy]), nonlocal value and gradient
                constexpr int NV = DIM == 2 ? 3 : 6;
                double eps[NV];
#pragma unroll
                for (int i = 0; i < NV; i++)
                    eps[i] = 0.0;
                double eN = 0.0, ge[3] = {0.0, 0.0, 0.0};
#pragma unroll
                for (int b = 0; b < NND; b++)
                {
                    double g[DIM];
                    Grad<DIM, NND>(dNd, b, inv, g);
                    if constexpr (DIM == 2)
                    {
                        eps[0] += g[0] * U[b][0];
                        eps[1] += g[1] * U[b][1];
                        eps[2] += g[1] * U[b][0] + g[0] * U[b][1];
                    }
                    else
                    {
                        eps[0] += g[0] * U[b][0];
                        eps[1] += g[1] * U[b][1];
                        eps[2] += g[2] * U[b][2];
                        eps[3] += g[2] * U[b][1] + g[1] * U[b][2];
                        eps[4] += g[2] * U[b][0] + g[0] * U[b][2];
                        eps[5] += g[1] * U[b][0] + g[0] * U[b][1];
                    }
                    if constexpr (KIND == kGradientDamage)
                    {
                        eN += Nd[b] * EE[b];
                        ge[0] += g[0] * EE[b];
                        ge[1] += g[1] * EE[b];
                        ge[2] += g[2] * EE[b];
                    }
                }

                // row node a
                double ga[DIM];
                Grad<DIM, NND>(dNd, la, inv, ga);
                const double Na = Nd[la];

                // material response -> per-IP coefficients
                double sig[NV];
                double c1 = 0.0, c2 = 0.0, cm = 0.0;  // uu: c1 ga gb^T + c2 gb ga^T + c2 (ga.gb) I
                double t6[6] = {0, 0, 0, 0, 0, 0}, m6[6] = {0, 0, 0, 0, 0, 0};
                double rE = 0.0, cq = 0.0;  // cq: s c dg/dw dw/dk dk/de (grad eeq . g_a), the K_ee term of g'
                double C11 = 0, C12 = 0, C33 = 0;
                if constexpr (KIND == kLinearElastic)
                {
                    if constexpr (DIM == 2)
                    {
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
                        sig[0] = C11 * eps[0] + C12 * eps[1];
                        sig[1] = C12 * eps[0] + C11 * eps[1];
                        sig[2] = C33 * eps[2];
                    }
                    else
                    {
                        const double f = E / ((1.0 + P.nu) * (1.0 - 2.0 * P.nu));
                        const double lam = f * P.nu, mu = E / (2.0 * (1.0 + P.nu));
                        const double tr = eps[0] + eps[1] + eps[2];
                        for (int i = 0; i < 3; i++)
                            sig[i] = lam * tr + 2.0 * mu * eps[i];
                        for (int i = 3; i < 6; i++)
                            sig[i] = mu * eps[i];
                        c1 = s * lam;
                        c2 = s * mu;
                    }
                }
                else
                {
                    double mdev[6];
                    const double eq = Mises3(eps, P.nu, P.k, mdev);
                    const double kh = A.kappa[(size_t)cid + (size_t)ip * A.kappaLd];
                    double kap, dk;
                    if constexpr (KIND == kGradientDamage)
                    {
                        kap = kh < eN ? eN : kh;
                        dk = eN >= kh ? 1.0 : 0.0;
                    }
                    else
                    {
                        kap = eq < kh ? kh : eq;
                        dk = eq >= kh ? 1.0 : 0.0;
                    }
                    double om, dom;
                    DamageExp(kap, k0, beta, P.alpha, om, dom);
                    DamageStress3 ds;
                    ElasticDamage3(eps, om, E, P.nu, P.damage_application, ds);
                    for (int i = 0; i < 6; i++)
                        sig[i] = ds.sig[i];
                    c1 = s * (ds.alpha_v - ds.alpha_d) / 3.0;
                    c2 = s * 0.5 * ds.alpha_d;
                    const double ct = s * dom * dk;
                    // gradient damage: K_eu = -N_a (B_b^T m) s; local damage: s dom dk (B_a^T dsdw)(B_b^T m)^T
                    const double cmv = KIND == kGradientDamage ? s : 1.0;
                    for (int i = 0; i < 6; i++)
                    {
                        t6[i] = ct * ds.dsdw[i];
                        m6[i] = cmv * mdev[i];
                    }
                    cm = P.c * s;
                    if constexpr (KIND == kGradientDamage)
                    {
                        double gi, dgi;
                        NonlocalFactor(P, om, gi, dgi);
                        const double gdot = ga[0] * ge[0] + ga[1] * ge[1] + ga[2] * ge[2];
                        rE = s * (Na * (eN - eq) + P.c * gi * gdot);
                        cm *= gi;
                        cq = s * P.c * gdot * dgi * dom * dk;
                    }
                }

                if constexpr (WR)
                {
                    if (b0 == 0)
                    {
                        if constexpr (DIM == 2)
                        {
                            racc[0] += s * (ga[0] * sig[0] + ga[1] * sig[2]);
                            racc[1] += s * (ga[1] * sig[1] + ga[0] * sig[2]);
                        }
                        else
                        {
                            double bt[3];
                            BTv3(ga, sig, bt);
                            racc[0] += s * bt[0];
                            racc[1] += s * bt[1];
                            racc[2] += s * bt[2];
                        }
                        if constexpr (KIND == kGradientDamage)
                            racc[3] += rE;
                    }
                }

                if constexpr (WK)
                {
                    double va[3] = {0, 0, 0};
                    if constexpr (KIND != kLinearElastic)
                        BTv3(ga, t6, va);
#pragma unroll
                    for (int bb = 0; bb < NB; bb++)
                    {
                        const int b = b0 + bb;
                        if (b >= NND)
                            break;
                        double gb[DIM];
                        Grad<DIM, NND>(dNd, b, inv, gb);
                        double* kb = kacc[bb];
                        if constexpr (DIM == 2)
                        {
                            const double sa0 = s * ga[0], sa1 = s * ga[1];
                            kb[0] += C11 * sa0 * gb[0] + C33 * sa1 * gb[1];
                            kb[1] += C12 * sa0 * gb[1] + C33 * sa1 * gb[0];
                            kb[2] += C12 * sa1 * gb[0] + C33 * sa0 * gb[1];
                            kb[3] += C11 * sa1 * gb[1] + C33 * sa0 * gb[0];
                        }
                        else
                        {
                            const double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
                            for (int i = 0; i < 3; i++)
#pragma unroll
                                for (int j = 0; j < 3; j++)
                                    kb[i * F + j] += c1 * ga[i] * gb[j] + c2 * gb[i] * ga[j] + (i == j ? c2 * dot : 0.0);
                            if constexpr (KIND == kLocalDamage)
                            {
                                double wb[3];
                                BTv3(gb, m6, wb);
#pragma unroll
                                for (int i = 0; i < 3; i++)
#pragma unroll
                                    for (int j = 0; j < 3; j++)
                                        kb[i * F + j] += va[i] * wb[j];
                            }
                            if constexpr (KIND == kGradientDamage)
                            {
                                double wb[3];
                                BTv3(gb, m6, wb);
                                const double Nb = Nd[b];
                                kb[0 * 4 + 3] += va[0] * Nb;
                                kb[1 * 4 + 3] += va[1] * Nb;
                                kb[2 * 4 + 3] += va[2] * Nb;
                                kb[3 * 4 + 0] -= Na * wb[0];
                                kb[3 * 4 + 1] -= Na * wb[1];
                                kb[3 * 4 + 2] -= Na * wb[2];
                                kb[15] += s * Na * Nb + cm * dot + cq * Nb;
                            }
                        }
                    }
                }
            }  // ip

            if constexpr (WK)
            {
                const unsigned short* sl = A.slot + (size_t)r * NND;
#pragma unroll
                for (int bb = 0; bb < NB; bb++)
                {
                    const int b = b0 + bb;
                    if (b >= NND)
                        break;
                    const unsigned short v = sl[b];
                    const int q = v & 0x7fff;
                    const bool first = (v & 0x8000) != 0;
                    const double* kb = kacc[bb];
#pragma unroll
                    for (int i = 0; i < F; i++)
                    {
                        double* row;
                        if (KIND == kGradientDamage && i == 3)
                            row = A.K + re0;
                        else
                            row = A.K + ru0 + (size_t)i * L;
#pragma unroll
                        for (int j = 0; j < F; j++)
                        {
                            size_t col;
                            if (KIND == kGradientDamage && j == 3)
                                col = (size_t)3 * nA + q;
                            else
                                col = (size_t)(KIND == kGradientDamage ? 3 : F) * q + j;
                            if (first)
                                row[col] = kb[i * F + j];
                            else
                                row[col] += kb[i * F + j];
                        }
                    }
                }
            }
        }  // b0
    }

    if constexpr (WR)
    {
#pragma unroll
        for (int c = 0; c < DIM; c++)
            A.R[(size_t)p * DIM + c] = racc[c];
        if constexpr (KIND == kGradientDamage)
            A.R[(size_t)A.nU + p] = racc[3];
    }
}

template <int KIND, int DIM, int NNC, int NND, int NIP, int BH, bool WR, bool WK>
__global__ void __launch_bounds__(64) kAssembleRows(const RowArgs A)
{
    AssembleRowsBody<KIND, DIM, NNC, NND, NIP, BH, WR, WK>(A, blockIdx.x * blockDim.x + threadIdx.x, A.nNodes);
}

// The rows of the dof nodes nodes[0 .. n) only (the nodes a specialised kernel leaves out)
template <int KIND, int DIM, int NNC, int NND, int NIP, int BH, bool WR, bool WK>
__global__ void __launch_bounds__(64) kAssembleRowsList(const RowArgs A, const int* __restrict__ nodes, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
        AssembleRowsBody<KIND, DIM, NNC, NND, NIP, BH, WR, WK>(A, nodes[i], A.nNodes);
}

// History update (GradientDamage::Update, GradientDamage.h:122-125, or EvolutionImplicit::Update,
// LocalIsotropicDamage.h:130-133): one lane per (cell, ip).
template <int KIND, int DIM, int NNC, int NND, int NIP>
__global__ void __launch_bounds__(256) kUpdateHistory(const RowArgs A, double* kappa)
{
    constexpr int S = 1 + NNC * DIM + NND + NND * DIM;
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)A.nCells * NIP)
        return;
    const int cell = (int)(gid / NIP), ip = (int)(gid % NIP);
    const double* t = A.tab + ip * S;
    const double* dNc = t + 1;
    const double* Nd = dNc + NNC * DIM;
    const double* dNd = Nd + NND;
    const int cid = A.cellId ? A.cellId[cell] : cell;
    double& kh = kappa[(size_t)cid + (size_t)ip * A.kappaLd];
    double val;
    if constexpr (KIND == kGradientDamage)
    {
        val = 0.0;
        for (int b = 0; b < NND; b++)
            val += Nd[b] * A.e[A.dconn[(size_t)cell * NND + b]];
    }
    else
    {
        double X[NNC][DIM];
        for (int k = 0; k < NNC; k++)
        {
            const int n = A.cconn[(size_t)cell * NNC + k];
            X[k][0] = A.x[n];
            X[k][1] = A.y[n];
            if constexpr (DIM == 3)
                X[k][2] = A.z[n];
        }
        double J[DIM][DIM];
        for (int i = 0; i < DIM; i++)
            for (int j = 0; j < DIM; j++)
            {
                double s = 0.0;
                for (int k = 0; k < NNC; k++)
                    s += X[k][i] * dNc[k * DIM + j];
                J[i][j] = s;
            }
        double inv[DIM][DIM];
        Inverse<DIM>(J, inv);
        double eps[6] = {0, 0, 0, 0, 0, 0};
        for (int b = 0; b < NND; b++)
        {
            const int q = A.dconn[(size_t)cell * NND + b];
            double g[DIM];
            Grad<DIM, NND>(dNd, b, inv, g);
            const double u0 = A.u[(size_t)q * DIM], u1 = A.u[(size_t)q * DIM + 1], u2 = A.u[(size_t)q * DIM + 2];
            eps[0] += g[0] * u0;
            eps[1] += g[1] * u1;
            eps[2] += g[2] * u2;
            eps[3] += g[2] * u1 + g[1] * u2;
            eps[4] += g[2] * u0 + g[0] * u2;
            eps[5] += g[1] * u0 + g[0] * u1;
        }
        double d[6];
        val = Mises3(eps, A.P.nu, A.P.k, d);
    }
    kh = kh < val ? val : kh;
}

}  // namespace nuto_amd
