// Two-phase generic assembly for 3D momentum balance (linear elasticity, local isotropic damage):
// element matrices first, then a row gather into the CSR.
//
// Phase 1 (kCellMatrices): Cell::Integrate (Cell.h:105-119) of one cell per lane group. Lane j of
// the group owns column j of the cell matrix K_e (dof j = 3 b + c of dof node b) and its R_e entry.
// Before the IP loop the group computes, for all IPs at once, J = X^T dN/dxi and the reference
// displacement gradient U^T dN/dxi (one 3x3 entry per lane and pass). Per IP each lane then forms
// J^-1, its own dN_b/dX (shared through LDS), the strain from (U^T dN/dxi) J^-1, the law's stress and tangent (the same closed forms as the row-owner kernel), then
// the stress of a unit displacement of its dof, dsig_j = D B_j (D with the local-damage rank-one
// term t m^T), and accumulates column j: K_e(3 a + i, j) += (B_a^T dsig_j)_i for every node a.
// The cell's state is computed once per IP per group (the row-owner form recomputes it for every
// row node and column block). K_e (row-major) and R_e go to a staging buffer.
//
// Phase 2 (kGatherRows): one lane per dof node p adds the staged rows of p's incident cells, in
// group order, into p's CSR rows (the slot table maps a cell's column nodes to row positions) and
// sums R_e the same way: the nonzero pattern and the summation order over cells are those of the
// row-owner kernel (and of setFromTriplets), IPs are summed per cell first as Cell::Integrate does.
#pragma once

#include <hip/hip_runtime.h>

#include "assemble_rows.hpp"

namespace nuto_amd
{

template <int KIND, int NNC, int NND, int NIP>
__global__ void __launch_bounds__(64) kCellMatrices(const RowArgs A, int c0, int c1, double* __restrict__ Ke,
                                                    double* __restrict__ Re)
{
    constexpr int ND = NND * 3;     // dofs (columns) per cell
    constexpr int CPW = 64 / ND;    // cells per wave
    constexpr int S = 1 + NNC * 3 + NND + NND * 3;
    static_assert(CPW >= 1, "a cell's columns must fit one wave");
    // per node 4 doubles (x, y, z, pad): 16-byte aligned pairs for wide LDS reads
    __shared__ __align__(16) double sX[CPW][NNC * 4];
    __shared__ __align__(16) double sU[CPW][NND * 4];
    __shared__ __align__(16) double sG[CPW][NND * 4];
    // per IP: J = X^T dNc (9) and the reference displacement gradient U^T dNd (9), computed once per
    // cell by the group's lanes together instead of by every lane
    __shared__ __align__(16) double sJG[CPW][NIP * 18];
    // per IP: J^-1 (9) and detJ * w
    __shared__ double sInv[CPW][NIP * 10];

    const int lane = (int)threadIdx.x;
    const int grp = lane / ND < CPW ? lane / ND : CPW - 1;  // lanes past the last group shadow it
    const int j = lane - grp * ND < ND ? lane - grp * ND : ND - 1;
    const int cellRaw = c0 + (int)blockIdx.x * CPW + lane / ND;
    const bool act = lane / ND < CPW && cellRaw < c1;
    const int cell = act ? cellRaw : (c0 + (int)blockIdx.x * CPW + grp < c1 ? c0 + (int)blockIdx.x * CPW + grp : c1 - 1);
    const int b = j / 3, c = j - 3 * b;

    if (act)
    {
        for (int k = j; k < NNC * 3; k += ND)
        {
            const int n = A.cconn[(size_t)cell * NNC + k / 3];
            const int comp = k % 3;
            sX[grp][(k / 3) * 4 + comp] = comp == 0 ? A.x[n] : (comp == 1 ? A.y[n] : A.z[n]);
        }
        const int q = A.dconn[(size_t)cell * NND + b];
        sU[grp][b * 4 + c] = A.u[(size_t)q * 3 + c];
    }
    __syncthreads();
    if (act)
        for (int tt = j; tt < NIP * 18; tt += ND)
        {
            const int ip = tt / 18, k = tt - 18 * ip;
            const double* dNc = A.tab + ip * S + 1;
            const double* dNd = dNc + NNC * 3 + NND;
            double v = 0.0;
            if (k < 9)
            {
                const int i = k / 3, m = k - 3 * i;
#pragma unroll
                for (int n = 0; n < NNC; n++)
                    v += sX[grp][n * 4 + i] * dNc[n * 3 + m];
            }
            else
            {
                const int i = (k - 9) / 3, m = k - 9 - 3 * i;
#pragma unroll
                for (int n = 0; n < NND; n++)
                    v += sU[grp][n * 4 + i] * dNd[n * 3 + m];
            }
            sJG[grp][tt] = v;
        }
    __syncthreads();
    if (act)
        for (int ip = j; ip < NIP; ip += ND)
        {
            const double* jg = &sJG[grp][ip * 18];
            double Jm[3][3], inv[3][3];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int m = 0; m < 3; m++)
                    Jm[i][m] = jg[3 * i + m];
            const double detJ = Inverse<3>(Jm, inv);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int m = 0; m < 3; m++)
                    sInv[grp][ip * 10 + 3 * i + m] = inv[i][m];
            sInv[grp][ip * 10 + 9] = detJ * A.tab[ip * S];
        }
    __syncthreads();

    const LawParams& P = A.P;
    double E = P.E, k0 = P.kappa0, beta = P.beta;
    if (A.thE)
    {
        E = A.thE[cell];
        k0 = A.thFt[cell] / E;
        beta = A.thFt[cell] / A.thGf[cell];
    }
    const int cid = A.cellId ? A.cellId[cell] : cell;
    const double m3K = 3.0 * E / (3.0 - 6.0 * P.nu), m2G = 2.0 * E / (2.0 + 2.0 * P.nu);

    double kcol[ND];
#pragma unroll
    for (int r = 0; r < ND; r++)
        kcol[r] = 0.0;
    double rj = 0.0;

#pragma unroll 1
    for (int ip = 0; ip < NIP; ip++)
    {
        const double* dNd = A.tab + ip * S + 1 + NNC * 3 + NND;
        const double* jg = &sJG[grp][ip * 18];
        double inv[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int m = 0; m < 3; m++)
                inv[i][m] = sInv[grp][ip * 10 + 3 * i + m];
        const double s = sInv[grp][ip * 10 + 9];
        double gb[3];
        Grad<3, NND>(dNd, b, inv, gb);
        if (act && c == 0)
        {
            sG[grp][b * 4 + 0] = gb[0];
            sG[grp][b * 4 + 1] = gb[1];
            sG[grp][b * 4 + 2] = gb[2];
        }
        __syncthreads();

        // strain (Voigt [xx, yy, zz, yz, zx, xy], engineering shear) from the displacement gradient
        // H = (U^T dNd) J^-1, H_ik = du_i / dx_k
        double H[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int k = 0; k < 3; k++)
                H[i][k] = jg[9 + 3 * i] * inv[0][k] + jg[9 + 3 * i + 1] * inv[1][k] + jg[9 + 3 * i + 2] * inv[2][k];
        const double eps[6] = {H[0][0], H[1][1], H[2][2], H[1][2] + H[2][1], H[0][2] + H[2][0], H[0][1] + H[1][0]};

        // law: stress, isotropic tangent (c1 tr + 2 c2 eps, shear c2 gamma), rank-one damage term
        double sig[6], c1, c2;
        double t6[6] = {0, 0, 0, 0, 0, 0}, m6[6] = {0, 0, 0, 0, 0, 0};
        if constexpr (KIND == kLinearElastic)
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
        else
        {
            double mdev[6];
            const double eq = Mises3(eps, P.nu, P.k, mdev);
            const double kh = A.kappa[(size_t)cid + (size_t)ip * A.kappaLd];
            const double kap = eq < kh ? kh : eq;
            const double dk = eq >= kh ? 1.0 : 0.0;
            double om, dom;
            DamageExp(kap, k0, beta, P.alpha, om, dom);
            DamageStress3 ds;
            ElasticDamage3K(eps, om, m3K, m2G, P.damage_application, ds);
            for (int i = 0; i < 6; i++)
                sig[i] = ds.sig[i];
            c1 = s * (ds.alpha_v - ds.alpha_d) * (1.0 / 3.0);
            c2 = s * 0.5 * ds.alpha_d;
            const double ct = s * dom * dk;
            for (int i = 0; i < 6; i++)
            {
                t6[i] = ct * ds.dsdw[i];
                m6[i] = mdev[i];
            }
        }

        // R_e: column j of B^T sig
        {
            double bt[3];
            BTv3(gb, sig, bt);
            rj += s * (c == 0 ? bt[0] : (c == 1 ? bt[1] : bt[2]));
        }

        // dsig_j = D B_j for the unit displacement of dof j = 3 b + c
        double de[6];
        de[0] = c == 0 ? gb[0] : 0.0;
        de[1] = c == 1 ? gb[1] : 0.0;
        de[2] = c == 2 ? gb[2] : 0.0;
        de[3] = c == 1 ? gb[2] : (c == 2 ? gb[1] : 0.0);
        de[4] = c == 0 ? gb[2] : (c == 2 ? gb[0] : 0.0);
        de[5] = c == 0 ? gb[1] : (c == 1 ? gb[0] : 0.0);
        const double tr = de[0] + de[1] + de[2];
        double ds6[6];
#pragma unroll
        for (int i = 0; i < 3; i++)
            ds6[i] = c1 * tr + 2.0 * c2 * de[i];
#pragma unroll
        for (int i = 3; i < 6; i++)
            ds6[i] = c2 * de[i];
        if constexpr (KIND == kLocalDamage)
        {
            const double mb = m6[0] * de[0] + m6[1] * de[1] + m6[2] * de[2] + m6[3] * de[3] + m6[4] * de[4] +
                              m6[5] * de[5];
#pragma unroll
            for (int i = 0; i < 6; i++)
                ds6[i] += t6[i] * mb;
        }

        // column j: K_e(3 a + i, j) += (B_a^T dsig_j)_i, three FMAs into each entry
#pragma unroll
        for (int a = 0; a < NND; a++)
        {
            const double gx = sG[grp][a * 4], gy = sG[grp][a * 4 + 1], gz = sG[grp][a * 4 + 2];
            kcol[a * 3 + 0] = fma(gx, ds6[0], fma(gz, ds6[4], fma(gy, ds6[5], kcol[a * 3 + 0])));
            kcol[a * 3 + 1] = fma(gy, ds6[1], fma(gz, ds6[3], fma(gx, ds6[5], kcol[a * 3 + 1])));
            kcol[a * 3 + 2] = fma(gz, ds6[2], fma(gy, ds6[3], fma(gx, ds6[4], kcol[a * 3 + 2])));
        }
        __syncthreads();  // sG is rewritten by the next IP
    }

    if (act)
    {
        double* dst = Ke + (size_t)(cell - c0) * ND * ND + j;
#pragma unroll
        for (int r = 0; r < ND; r++)
            dst[(size_t)r * ND] = kcol[r];
        Re[(size_t)(cell - c0) * ND + j] = rj;
    }
}

template <int NND, bool WR>
__global__ void __launch_bounds__(64) kGatherRows(const RowArgs A, int p0, int p1, int s0, const double* __restrict__ Ke,
                                                  const double* __restrict__ Re)
{
    constexpr int ND = NND * 3;
    const int p = p0 + (int)(blockIdx.x * 64 + threadIdx.x);
    if (p >= p1)
        return;
    const int beg = A.incPtr[p], end = A.incPtr[p + 1];
    const int nA = A.adjPtr[p + 1] - A.adjPtr[p];
    const int L = nA * 3;
    const int ru0 = A.rowPtr[p * 3];
    double racc[3] = {0.0, 0.0, 0.0};
    for (int r = beg; r < end; ++r)
    {
        const int cell = A.incCell[r];
        const int la = A.incLocal[r];
        const unsigned short* sl = A.slot + (size_t)r * NND;
        const double* src = Ke + ((size_t)(cell - s0) * ND + la * 3) * ND;
        if constexpr (WR)
        {
            const double* rs = Re + (size_t)(cell - s0) * ND + la * 3;
            racc[0] += rs[0];
            racc[1] += rs[1];
            racc[2] += rs[2];
        }
        for (int b = 0; b < NND; b++)
        {
            const unsigned short v = sl[b];
            const int q = v & 0x7fff;
            const bool first = (v & 0x8000) != 0;
#pragma unroll
            for (int i = 0; i < 3; i++)
            {
                double* row = A.K + ru0 + (size_t)i * L + 3 * q;
                const double* s = src + i * ND + b * 3;
#pragma unroll
                for (int jj = 0; jj < 3; jj++)
                {
                    if (first)
                        row[jj] = s[jj];
                    else
                        row[jj] += s[jj];
                }
            }
        }
    }
    if constexpr (WR)
    {
        A.R[(size_t)p * 3 + 0] = racc[0];
        A.R[(size_t)p * 3 + 1] = racc[1];
        A.R[(size_t)p * 3 + 2] = racc[2];
    }
}

// Phase 2 with one wave per dof node: the node's three rows are summed in LDS (3 L doubles, L =
// 3 |adj(p)|, dynamic shared memory of 3 maxL doubles), each incident cell's three staged rows are
// read with coalesced 16-byte loads (3 ND is even and the rows start 16-byte aligned) and added at
// their slots in cell order, then the rows are written with coalesced streaming stores (the rows of
// p are contiguous in the CSR and are not read again here). The loads of CB cells are issued before
// the first of them is added, so that a wave has CB cells' rows in flight at once.
template <int NND, bool WR>
__global__ void __launch_bounds__(64) kGatherRowsWave(const RowArgs A, int p0, int p1, int s0, const double* __restrict__ Ke,
                                                      const double* __restrict__ Re)
{
    constexpr int ND = NND * 3;
    constexpr int NH = 3 * ND / 2;          // 16-byte pairs of a cell's three staged rows
    constexpr int NF = (NH + 63) / 64;      // pairs per lane and cell
    constexpr int CB = NND >= 10 ? 2 : 4;   // cells in flight
    static_assert(ND % 2 == 0, "pairs of a row");
    extern __shared__ double rows[];
    const int p = p0 + (int)blockIdx.x;
    const int lane = (int)threadIdx.x;
    if (p >= p1)
        return;
    // pair h = lane + 64 k holds entries 2h, 2h + 1 of row i (a row has an even length): column
    // nodes b and components jj of both
    int fRow[NF], fNode[NF][2], fComp[NF][2];
#pragma unroll
    for (int k = 0; k < NF; k++)
    {
        const int f = 2 * (lane + 64 * k);
        const int i = f / ND, rem = f - i * ND;
        fRow[k] = i;
#pragma unroll
        for (int w = 0; w < 2; w++)
        {
            fNode[k][w] = (rem + w) / 3;
            fComp[k][w] = rem + w - 3 * fNode[k][w];
        }
    }
    const int beg = A.incPtr[p], end = A.incPtr[p + 1];
    const int L = (A.adjPtr[p + 1] - A.adjPtr[p]) * 3;
    const int ru0 = A.rowPtr[p * 3];
    for (int f = lane; f < 3 * L; f += 64)
        rows[f] = 0.0;
    double racc = 0.0;
    __syncthreads();
    for (int r0 = beg; r0 < end; r0 += CB)
    {
        double v[CB][NF][2], rv[CB];
        int o[CB][NF][2];
#pragma unroll
        for (int c = 0; c < CB; c++)
        {
            const int r = r0 + c < end ? r0 + c : end - 1;  // past the last cell: loads of the last, unused
            const int cell = A.incCell[r];
            const int la = A.incLocal[r];
            const unsigned short* sl = A.slot + (size_t)r * NND;
            const double2* src = reinterpret_cast<const double2*>(Ke + ((size_t)(cell - s0) * ND + la * 3) * ND);
#pragma unroll
            for (int k = 0; k < NF; k++)
            {
                const int h = lane + 64 * k;
                if (h < NH)
                {
                    const double2 t = src[h];
                    v[c][k][0] = t.x;
                    v[c][k][1] = t.y;
#pragma unroll
                    for (int w = 0; w < 2; w++)
                        o[c][k][w] = fRow[k] * L + 3 * (sl[fNode[k][w]] & 0x7fff) + fComp[k][w];
                }
            }
            rv[c] = (WR && lane < 3) ? Re[(size_t)(cell - s0) * ND + la * 3 + lane] : 0.0;
        }
#pragma unroll
        for (int c = 0; c < CB; c++)
        {
            if (r0 + c < end)
            {
#pragma unroll
                for (int k = 0; k < NF; k++)
                    if (lane + 64 * k < NH)
                    {
                        rows[o[c][k][0]] += v[c][k][0];
                        rows[o[c][k][1]] += v[c][k][1];
                    }
                racc += rv[c];
                __syncthreads();  // the next cell may add at the same entries
            }
        }
    }
    double* dst = A.K + ru0;
    for (int f = lane; f < 3 * L; f += 64)
        __builtin_nontemporal_store(rows[f], dst + f);
    if (WR && lane < 3)
        A.R[(size_t)p * 3 + lane] = racc;
}

}  // namespace nuto_amd
