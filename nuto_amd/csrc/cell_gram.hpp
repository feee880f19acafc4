// Cell matrices of 3D momentum balance (linear elasticity, local isotropic damage) in Gram form:
// phase 1 of the two-phase assembly, writing the same staging as kCellMatrices (K_e row-major per
// cell, R_e per cell) for the row gathers.
//
// Cell::Integrate (Cell.h:105-119) of MomentumBalance<3>::Hessian0 (MomentumBalance.h:37-46) with
// the law's tangent D = c1 (1 x 1) + 2 c2 I_sym (LinearElastic<3>, LinearElasticDamage<3> at omega)
// plus, for LocalIsotropicDamage (LocalIsotropicDamage.h:59-68), the rank-one term
// (d sigma / d omega) omega' (d eps_eq / d eps)^T while loading. With g_n = dN_n/dX, the 3x3 block
// (a, b) of B^T D B is
//     c1 g_a g_b^T + c2 g_b g_a^T + c2 (g_a . g_b) I + u_a v_b^T,
// u_a = B_a^T t (t = s omega' kappa' d sigma / d omega), v_b = B_b^T (d eps_eq / d eps). Summed over
// the IPs (weight s = detJ w folded into c1, c2, t):
//     K_ab = G1 + G2^T + tr(G2) I + Σ u_a v_b^T,  G1 = Σ c1 g_a g_b^T,  G2 = Σ c2 g_a g_b^T.
// When c1 / c2 is the same at every IP (linear elasticity: lambda / mu; FULL damage application:
// (3K - 2G) / 3G, since both moduli carry (1 - omega)) one Gram suffices: G1 = r G2.
//
// Work split per cell: the IP state (Jacobian, strain, law) once per IP; g_n, u_n, v_n and the R
// terms s B_n^T sigma once per (node, IP); each lane then owns node a and the column nodes b and
// b + NND/2 (NND is even for tet4, tet10, hex8, hex20) and accumulates their two blocks over the
// IPs from LDS. The IPs are processed in chunks of ICH so that the per-(node, IP) data fit in LDS.
// Every K_e and R_e entry is summed by one lane in IP order: deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include "assemble_rows.hpp"
#include "laws.hpp"

namespace nuto_amd
{

namespace gram
{
constexpr int kIpRec = 30;  // per IP: J^-1 (9), s, c1, c2, sigma (6), t (6), d eps_eq / d eps (6)
}

// per-cell material (E, kappa0, beta; per-cell fields when given)
__device__ inline void CellMaterial(const RowArgs& A, int cell, double& E, double& k0, double& beta)
{
    E = A.P.E;
    k0 = A.P.kappa0;
    beta = A.P.beta;
    if (A.thE)
    {
        E = A.thE[cell];
        k0 = A.thFt[cell] / E;
        beta = A.thFt[cell] / A.thGf[cell];
    }
}

// Phase 1 of the cell-matrix kernels for the cells cellBase .. cellBase + CPW - 1 (clamped to c1 - 1)
// on NT threads: coordinates and displacements to LDS, J and the reference displacement gradient of
// every IP, then each IP's state record (gram::kIpRec doubles: J^-1, s = detJ w, c1, c2, sigma, t,
// d eps_eq / d eps). Ends with the records written (no barrier after them).
template <int KIND, int NNC, int NND, int NIP, int CPW, int NT>
__device__ inline void CellIpRecords(const RowArgs& A, int c1, int cellBase, double (*sX)[NNC * 3],
                                     double (*sU)[NND * 3], double (*sJG)[NIP][18],
                                     double (*sIp)[NIP][gram::kIpRec])
{
    using namespace gram;
    constexpr int S = 1 + NNC * 3 + NND + NND * 3;  // IP table row: w, dNc, N, dNd
    const int t = (int)threadIdx.x;
    const LawParams& P = A.P;
    // ---- cell data: coordinates, displacements ------------------------------------------------------
    for (int k = t; k < CPW * NNC * 3; k += NT)
    {
        const int cl = k / (NNC * 3), r = k - cl * NNC * 3;
        const int cell = cellBase + cl < c1 ? cellBase + cl : c1 - 1;
        const int n = A.cconn[(size_t)cell * NNC + r / 3];
        const int comp = r % 3;
        sX[cl][r] = comp == 0 ? A.x[n] : (comp == 1 ? A.y[n] : A.z[n]);
    }
    for (int k = t; k < CPW * NND * 3; k += NT)
    {
        const int cl = k / (NND * 3), r = k - cl * NND * 3;
        const int cell = cellBase + cl < c1 ? cellBase + cl : c1 - 1;
        const int q = A.dconn[(size_t)cell * NND + r / 3];
        sU[cl][r] = A.u[(size_t)q * 3 + r % 3];
    }


    __syncthreads();  // cell data loaded
    // ---- J = X^T dNc and the reference displacement gradient U^T dNd of every IP ----------------
    for (int k = t; k < CPW * NIP * 18; k += NT)
    {
        const int cl = k / (NIP * 18), r = k - cl * NIP * 18;
        const int ip = r / 18, e = r - 18 * ip;
        const double* dNc = A.tab + ip * S + 1;
        double v = 0.0;
        if (e < 9)
        {
            const int row = e / 3, m = e - 3 * row;
#pragma unroll
            for (int n = 0; n < NNC; n++)
                v += sX[cl][n * 3 + row] * dNc[n * 3 + m];
        }
        else
        {
            const double* dNd = dNc + NNC * 3 + NND;
            const int row = (e - 9) / 3, m = e - 9 - 3 * row;
#pragma unroll
            for (int n = 0; n < NND; n++)
                v += sU[cl][n * 3 + row] * dNd[n * 3 + m];
        }
        sJG[cl][ip][e] = v;
    }
    __syncthreads();
    // ---- IP state of every IP: J^-1, s = detJ w, strain, law (before the block accumulators live) ----
    for (int k = t; k < CPW * NIP; k += NT)
    {
        {
            const int cl = k / NIP, ip = k - cl * NIP, i = ip;
            const int cell = cellBase + cl < c1 ? cellBase + cl : c1 - 1;
            const double* jg = sJG[cl][i];
            double Jm[3][3], inv[3][3];
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int m = 0; m < 3; m++)
                    Jm[r][m] = jg[3 * r + m];
            const double detJ = Inverse<3>(Jm, inv);
            const double s = detJ * A.tab[ip * S];
            double H[3][3];
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int m = 0; m < 3; m++)
                    H[r][m] = jg[9 + 3 * r] * inv[0][m] + jg[9 + 3 * r + 1] * inv[1][m] + jg[9 + 3 * r + 2] * inv[2][m];
            const double eps[6] = {H[0][0], H[1][1], H[2][2], H[1][2] + H[2][1], H[0][2] + H[2][0], H[0][1] + H[1][0]};
            double* rec = sIp[cl][i];
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int m = 0; m < 3; m++)
                    rec[3 * r + m] = inv[r][m];
            rec[9] = s;
            double E, k0, beta;
            CellMaterial(A, cell, E, k0, beta);
            if constexpr (KIND == kLinearElastic)
            {
                const double f = E / ((1.0 + P.nu) * (1.0 - 2.0 * P.nu));
                const double lam = f * P.nu, mu = E / (2.0 * (1.0 + P.nu));
                const double tr = eps[0] + eps[1] + eps[2];
                rec[10] = s * lam;
                rec[11] = s * mu;
                for (int r = 0; r < 3; r++)
                    rec[12 + r] = lam * tr + 2.0 * mu * eps[r];
                for (int r = 3; r < 6; r++)
                    rec[12 + r] = mu * eps[r];
            }
            else
            {
                const double m3K = 3.0 * E / (3.0 - 6.0 * P.nu), m2G = 2.0 * E / (2.0 + 2.0 * P.nu);
                double mdev[6];
                const double eq = Mises3(eps, P.nu, P.k, mdev);
                const int cid = A.cellId ? A.cellId[cell] : cell;
                const double kh = A.kappa[(size_t)cid + (size_t)ip * A.kappaLd];
                const double kap = eq < kh ? kh : eq;
                const double dk = eq >= kh ? 1.0 : 0.0;
                double om, dom;
                DamageExp(kap, k0, beta, P.alpha, om, dom);
                DamageStress3 ds;
                ElasticDamage3K(eps, om, m3K, m2G, P.damage_application, ds);
                rec[10] = s * (ds.alpha_v - ds.alpha_d) * (1.0 / 3.0);
                rec[11] = s * 0.5 * ds.alpha_d;
                const double ct = s * dom * dk;
                for (int r = 0; r < 6; r++)
                {
                    rec[12 + r] = ds.sig[r];
                    rec[18 + r] = ct * ds.dsdw[r];
                    rec[24 + r] = mdev[r];
                }
            }
        }
    }
}

// CPW cells per workgroup of NT threads; lanes [0, CPW * LPC) own blocks. AP = 1: a lane owns row
// node a and column nodes p, p + NND/2 (LPC = NND^2 / 2 lanes per cell); AP = 2: row nodes a, a +
// NND/2 as well (LPC = NND^2 / 4), four blocks per lane from the same g and u, v reads (fewer LDS
// reads per block, more registers)
template <int KIND, int NNC, int NND, int NIP, int CPW, int NT, int ICH, int UNR, int AP, bool TWO>
__global__ void __launch_bounds__(NT) kCellGram(const RowArgs A, int c0, int c1, double* __restrict__ Ke,
                                                double* __restrict__ Re)
{
    using namespace gram;
    constexpr int ND = 3 * NND, HALF = NND / 2, LPC = (AP == 2 ? HALF : NND) * HALF;
    static_assert(AP == 1 || AP == 2, "one or two row nodes per lane");
    constexpr int S = 1 + NNC * 3 + NND + NND * 3;  // IP table row: w, dNc, N, dNd
    constexpr bool LD = KIND == kLocalDamage;
    static_assert(NND % 2 == 0, "column nodes come in pairs");
    static_assert(CPW * LPC <= NT, "a lane per (cell, a, pair)");

    __shared__ double sX[CPW][NNC * 3];
    __shared__ double sU[CPW][NND * 3];
    __shared__ double sJG[CPW][NIP][18];
    __shared__ double sIp[CPW][NIP][kIpRec];
    // per (node, IP): g (0..2), u (3..5), v (6..8); the odd record length (3 or 9 doubles) puts the
    // records that the lanes of a wave read at once (consecutive nodes) on distinct LDS banks
    constexpr int NR = LD ? 9 : 3;
    __shared__ double sGU[CPW][ICH][NND][NR];
    __shared__ double sRb[CPW][ICH][NND][3];
    __shared__ double sR[CPW][NND * 3];

    const int t = (int)threadIdx.x;
    const int cellBase = c0 + (int)blockIdx.x * CPW;

    for (int k = t; k < CPW * NND * 3; k += NT)
        sR[k / (NND * 3)][k % (NND * 3)] = 0.0;
    CellIpRecords<KIND, NNC, NND, NIP, CPW, NT>(A, c1, cellBase, sX, sU, sJG, sIp);

    // block lane: cell bc, row nodes a (+ HALF when AP = 2), column nodes b1 = p, b2 = p + HALF
    const bool blk = t < CPW * LPC;
    const int bc = blk ? t / LPC : 0;
    const int a = blk ? (t - bc * LPC) / HALF : 0;
    const int p = blk ? t - bc * LPC - a * HALF : 0;
    double G[AP][2][9], G1[AP][2][TWO ? 9 : 1], R1[AP][2][LD ? 9 : 1];
#pragma unroll
    for (int sa = 0; sa < AP; sa++)
#pragma unroll
        for (int q = 0; q < 2; q++)
        {
#pragma unroll
            for (int k = 0; k < 9; k++)
            {
                G[sa][q][k] = 0.0;
                if constexpr (TWO)
                    G1[sa][q][k] = 0.0;
                if constexpr (LD)
                    R1[sa][q][k] = 0.0;
            }
        }

#pragma unroll 1
    for (int ip0 = 0; ip0 < NIP; ip0 += ICH)
    {
        __syncthreads();  // IP states written / the previous chunk's LDS consumed
        // ---- per (node, IP): g_n = J^-T dN_n/dxi, u_n, v_n and s B_n^T sigma --------------------------
        for (int k = t; k < CPW * ICH * NND; k += NT)
        {
            const int cl = k / (ICH * NND), r = k - cl * ICH * NND;
            const int i = r / NND, n = r - i * NND, ip = ip0 + i;
            if (ip >= NIP)
                continue;
            const double* rec = sIp[cl][ip];
            const double* dN = A.tab + ip * S + 1 + NNC * 3 + NND + n * 3;
            double g[3];
#pragma unroll
            for (int c = 0; c < 3; c++)
                g[c] = dN[0] * rec[c] + dN[1] * rec[3 + c] + dN[2] * rec[6 + c];
            double* gd = sGU[cl][i][n];
            gd[0] = g[0];
            gd[1] = g[1];
            gd[2] = g[2];
            double bt[3];
            BTv3(g, rec + 12, bt);
            sRb[cl][i][n][0] = rec[9] * bt[0];
            sRb[cl][i][n][1] = rec[9] * bt[1];
            sRb[cl][i][n][2] = rec[9] * bt[2];
            if constexpr (LD)
            {
                double w[3];
                BTv3(g, rec + 18, w);
                gd[3] = w[0];
                gd[4] = w[1];
                gd[5] = w[2];
                BTv3(g, rec + 24, w);
                gd[6] = w[0];
                gd[7] = w[1];
                gd[8] = w[2];
            }
        }
        __syncthreads();
        // ---- R_e: one lane per (cell, n, component) adds the chunk's IPs in order ---------------------
        for (int k = t; k < CPW * NND * 3; k += NT)
        {
            const int cl = k / (NND * 3), r = k - cl * NND * 3;
            double acc = sR[cl][r];
#pragma unroll
            for (int i = 0; i < ICH; i++)
                if (ip0 + i < NIP)
                    acc += sRb[cl][i][r / 3][r % 3];
            sR[cl][r] = acc;
        }
        // ---- blocks (a, b1), (a, b2) [and (a + HALF, b1), (a + HALF, b2)] over the chunk's IPs --------
        if (blk)
        {
#pragma unroll UNR
            for (int i = 0; i < ICH; i++)
            {
                if (ip0 + i >= NIP)
                    break;
                const double* rec = sIp[bc][ip0 + i];
                const double cw1 = rec[10], cw2 = rec[11];
                double ha[AP][3], h1[AP][TWO ? 3 : 1], ua[AP][LD ? 3 : 1];
#pragma unroll
                for (int sa = 0; sa < AP; sa++)
                {
                    const double* ga = sGU[bc][i][a + sa * HALF];
#pragma unroll
                    for (int r = 0; r < 3; r++)
                    {
                        ha[sa][r] = cw2 * ga[r];
                        if constexpr (TWO)
                            h1[sa][r] = cw1 * ga[r];
                    }
                    if constexpr (LD)
                    {
                        ua[sa][0] = ga[3];
                        ua[sa][1] = ga[4];
                        ua[sa][2] = ga[5];
                    }
                }
#pragma unroll
                for (int q = 0; q < 2; q++)
                {
                    const int b = p + q * HALF;
                    const double* gb = sGU[bc][i][b];
                    const double g0 = gb[0], g1 = gb[1], g2 = gb[2];
                    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
                    if constexpr (LD)
                    {
                        v0 = gb[6];
                        v1 = gb[7];
                        v2 = gb[8];
                    }
#pragma unroll
                    for (int sa = 0; sa < AP; sa++)
                    {
#pragma unroll
                        for (int r = 0; r < 3; r++)
                        {
                            G[sa][q][3 * r + 0] = fma(ha[sa][r], g0, G[sa][q][3 * r + 0]);
                            G[sa][q][3 * r + 1] = fma(ha[sa][r], g1, G[sa][q][3 * r + 1]);
                            G[sa][q][3 * r + 2] = fma(ha[sa][r], g2, G[sa][q][3 * r + 2]);
                        }
                        if constexpr (TWO)
                        {
#pragma unroll
                            for (int r = 0; r < 3; r++)
                            {
                                G1[sa][q][3 * r + 0] = fma(h1[sa][r], g0, G1[sa][q][3 * r + 0]);
                                G1[sa][q][3 * r + 1] = fma(h1[sa][r], g1, G1[sa][q][3 * r + 1]);
                                G1[sa][q][3 * r + 2] = fma(h1[sa][r], g2, G1[sa][q][3 * r + 2]);
                            }
                        }
                        if constexpr (LD)
                        {
#pragma unroll
                            for (int r = 0; r < 3; r++)
                            {
                                R1[sa][q][3 * r + 0] = fma(ua[sa][r], v0, R1[sa][q][3 * r + 0]);
                                R1[sa][q][3 * r + 1] = fma(ua[sa][r], v1, R1[sa][q][3 * r + 1]);
                                R1[sa][q][3 * r + 2] = fma(ua[sa][r], v2, R1[sa][q][3 * r + 2]);
                            }
                        }
                    }
                }
            }
        }
    }
    __syncthreads();

    // ---- K_e blocks: G1 + G2^T + tr(G2) I (+ the rank-one sum), rows 3a + i, columns 3b + j ----------
    const int cell = cellBase + bc;
    if (blk && cell < c1)
    {
        // one Gram: G1 = r G2 with r = c1 / c2 of the law (E cancels: one value per launch)
        const double r = TWO ? 0.0 : (KIND == kLinearElastic ? A.P.rLE : A.P.rLD);
#pragma unroll
        for (int sa = 0; sa < AP; sa++)
#pragma unroll
            for (int q = 0; q < 2; q++)
            {
                const int b = p + q * HALF;
                const double tr = G[sa][q][0] + G[sa][q][4] + G[sa][q][8];
                double* dst = Ke + ((size_t)(cell - c0) * ND + 3 * (a + sa * HALF)) * ND + 3 * b;
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++)
                    {
                        double v = (TWO ? G1[sa][q][TWO ? 3 * i + j : 0] : r * G[sa][q][3 * i + j]) + G[sa][q][3 * j + i];
                        if (i == j)
                            v += tr;
                        if constexpr (LD)
                            v += R1[sa][q][3 * i + j];
                        dst[(size_t)i * ND + j] = v;
                    }
            }
    }
    for (int k = t; k < CPW * NND * 3; k += NT)
    {
        const int cl = k / (NND * 3), rr = k - cl * NND * 3;
        if (cellBase + cl < c1)
            Re[(size_t)(cellBase + cl - c0) * ND + rr] = sR[cl][rr];
    }
}

}  // namespace nuto_amd
