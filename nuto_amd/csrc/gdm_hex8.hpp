// Hex8 gradient-damage R+K kernel (the hot path): one 64-lane workgroup per dof node p.
//
// Lanes are (r, j) with r = the r-th cell incident to p (group order) and j = 0..7.
//   phase A  lane (r, ip=j): the integration-point state of cell r at IP ip (Jacobian, strain,
//            nonlocal strain, damage, stress, equivalent-strain derivative) -> LDS record
//   phase B  lane (r, b=j):  the 4x4 block (a_r, b) of cell r's element matrix for p's rows
//            (a_r = local index of p), summed over the 8 IPs in order (Cell.h:108-117)
//   reduce   slot q of p's neighbourhood = sum over r (cell order, the order setFromTriplets
//            uses, SimpleAssembler.cpp:149-155) of the blocks that land on q
//   write    p's three displacement rows (contiguous) and its eeq row, coalesced, written once.
// R is reduced the same way (cells in order, IPs in order). No atomics: the result is
// deterministic, and every nonzero and R entry is written exactly once per assembly.
//
// Block algebra (per IP, weight s = detJ * w; g = dN/dX; FULL damage application):
//   uu = lambda' S + mu' S^T + mu' tr(S) I with S = sum_ip s (1-omega) g_a g_b^T,
//        lambda' = (3K - 2G) / 3, mu' = G  (LinearElasticDamage::DstressDstrain, H = 1)
//   ue = sum_ip (B_a^T t) N_b, t = s * omega' * kappa' * dSigma/dOmega
//   eu = -sum_ip N_a B_b^T (s * d eps_eq / d eps)
//   ee = sum_ip s N_a N_b + c s g_a . g_b
// UNILATERAL damage application keeps the volumetric factor per IP, so uu is accumulated in the
// general form c1 g_a g_b^T + c2 g_b g_a^T + c2 (g_a . g_b) I.
//
// Usable when every node has at most 8 incident cells and at most 32 neighbours (structured hex
// meshes: 8 and 27) and the coordinate and dof layers share their nodes; other meshes use the
// generic row-owner kernel.
//
// Gathers are three dependent loads deep: per-node incidence tables padded to 8 cells (cell,
// local index, history column, neighbour slots), then the cell's dof node ranks and its history,
// then the node values (coordinates are stored by dof node for this).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <string>
#include <vector>

#include "assemble_rows.hpp"
#include "devbuf.hpp"
#include "laws.hpp"
#include "mesh.hpp"

namespace nuto_amd
{

namespace gdm8
{
constexpr int kMaxInc = 8;   // incident cells per node
constexpr int kMaxAdj = 32;  // neighbours per node
constexpr int kSt = 30;      // doubles per (ip, cell) record; the 8 records of one IP start 60 dwords
                             // apart (8 disjoint 4-bank groups) and are 16-byte aligned
constexpr int kNd = 66;      // doubles per cell of node data (8 nodes x 8 + 2)
constexpr int kPs = 17;      // doubles per partial block (odd: 64 lanes spread over the banks)
constexpr int kTab = 57;     // per-IP table row: w, dNc[8][3], N[8], dNd[8][3]

// record layout
constexpr int sInv = 0;   // J^-1 (9)
constexpr int sHa = 9;    // FULL: s (1-omega) g_a; UNILATERAL: c1 g_a (3)
constexpr int sVa = 12;   // B_a^T t (3)
constexpr int sMn = 15;   // -N_a s d eps_eq / d eps (6)
constexpr int sCg = 21;   // c s g_a (3)
constexpr int sSn = 24;   // s N_a (1)
constexpr int sKa = 25;   // UNILATERAL: c2 g_a (3)

constexpr double kGauss = 0.57735026918962573;  // sqrt(1/3), the 2-point Gauss abscissa
constexpr double kThird = 1.0 / 3.0;
constexpr double kSixth = 1.0 / 6.0;

// dN_n/dxi and N_n of hex8 node n (InterpolationBrickLinear.cpp:42-57 order) at 2x2x2 Gauss point
// ip (xi fastest, IntegrationTypeTensorProduct.cpp:28-38): products of the factors 1 +- g
__device__ inline void ShapeAt(int n, int ip, double d[3], double& N)
{
    const int sx = ((n + 1) >> 1) & 1, sy = (n >> 1) & 1, sz = n >> 2;
    const double f0 = (sx == (ip & 1)) ? 1.0 + kGauss : 1.0 - kGauss;
    const double f1 = (sy == ((ip >> 1) & 1)) ? 1.0 + kGauss : 1.0 - kGauss;
    const double f2 = (sz == (ip >> 2)) ? 1.0 + kGauss : 1.0 - kGauss;
    const double e = 0.125;
    d[0] = (sx ? e : -e) * f1 * f2;
    d[1] = (sy ? e : -e) * f0 * f2;
    d[2] = (sz ? e : -e) * f0 * f1;
    N = e * f0 * f1 * f2;
}

// g = J^-T d  (dN/dX = dN/dxi J^-1, Jacobian.h:100-103)
__device__ inline void GradOf(const double d[3], const double inv[3][3], double g[3])
{
#pragma unroll
    for (int c = 0; c < 3; c++)
        g[c] = d[0] * inv[0][c] + d[1] * inv[1][c] + d[2] * inv[2][c];
}
}  // namespace gdm8

// Per-node tables of the node kernel (built once per pattern, see GdmHex8Plan::Prepare).
struct GdmHex8Tables
{
    const int* nodeCell;            // [node][8]: r-th incident cell (group order), -1 past the last
    const int* nodeHist;            // [node][8]: history column Cell::Id of that cell (nullptr: the cell index)
    const unsigned char* nodeLa;    // [node][8]: local index of the node in that cell
    const unsigned char* nodeSlot;  // [node][8][8]: neighbour slot (column block) of local node j
    const double* xd;               // coordinates by dof node
    const double* yd;
    const double* zd;
    int xcdMap;                     // 1: contiguous node ranges per XCD
};

// Inputs of lane (r, j) of node p.
struct NodeFetch
{
    int cell, la, slot;
    double v[7];  // x y z u0 u1 u2 e of node j of cell r
    double th[3];
    double kh;  // history kappa of cell r at IP j
    double w;   // weight of IP j
};

__device__ inline void FetchNode(const RowArgs& A, const GdmHex8Tables& T, int p, int r, int j, NodeFetch& f)
{
    const size_t pr = (size_t)p * 8 + r;
    f.cell = T.nodeCell[pr];
    f.la = T.nodeLa[pr];
    f.slot = T.nodeSlot[pr * 8 + j];
    f.w = A.tab[j * gdm8::kTab];
    f.kh = 0.0;
    if (f.cell >= 0)
    {
        const int hist = T.nodeHist ? T.nodeHist[pr] : f.cell;
        const int q = A.dconn[(size_t)f.cell * 8 + j];
        f.kh = A.kappa[(size_t)hist + (size_t)j * A.kappaLd];
        f.v[0] = T.xd[q];
        f.v[1] = T.yd[q];
        f.v[2] = T.zd[q];
        f.v[3] = A.u[(size_t)q * 3];
        f.v[4] = A.u[(size_t)q * 3 + 1];
        f.v[5] = A.u[(size_t)q * 3 + 2];
        f.v[6] = A.e[q];
        if (A.thE)
        {
            f.th[0] = A.thE[f.cell];
            f.th[1] = A.thFt[f.cell];
            f.th[2] = A.thGf[f.cell];
        }
    }
}

// One workgroup per node. Workgroups are dispatched round-robin over the 8 XCDs; with
// xcdMap the node of block b is chosen so that each XCD processes a contiguous range of nodes
// (neighbouring nodes share cells and node data, which then stay in that XCD's L2).
template <bool WR, bool WK, bool FULLDMG>
__global__ void __launch_bounds__(64) kGdmHex8Node(const RowArgs A, const GdmHex8Tables T)
{
    using namespace gdm8;
    __shared__ __align__(16) double state[8 * 8 * kSt];  // [ip][r][kSt]; later partial blocks + R terms
    __shared__ __align__(16) double rowbuf[8 * kNd];      // node data of the incident cells, later p's rows
    __shared__ __align__(8) unsigned char slotOf[kMaxAdj][kMaxInc];

    const int lane = threadIdx.x;
    const int r = lane >> 3, j = lane & 7;
    const LawParams& P = A.P;
    int p = blockIdx.x;
    if (T.xcdMap)
    {
        const int per = (int)((gridDim.x + 7) >> 3);
        p = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
        if (p >= A.nNodes)
            return;
    }
    // p's rows: 3 displacement rows (consecutive in the CSR) and the eeq row, each 4 nA long
    const int ru0 = A.rowPtr[(size_t)p * 3];
    const int L = A.rowPtr[(size_t)p * 3 + 1] - ru0;
    const int nA = L >> 2;
    const int re0 = A.rowPtr[(size_t)A.nU + p];
    NodeFetch f;
    FetchNode(A, T, p, r, j, f);
    const bool valid = f.cell >= 0;  // r < number of cells incident to p
    const int la = f.la;

    // ---- node data of p's cells to LDS; map slots -> (r, b) -----------------------------------------
    double* nodeData = rowbuf;  // [r][n][8] (+2 per cell): x y z u0 u1 u2 e -
    for (int i = lane; i < kMaxAdj * kMaxInc; i += 64)
        (&slotOf[0][0])[i] = 0xFF;
    __syncthreads();
    if (valid)
    {
        double* nd = nodeData + r * kNd + j * 8;
#pragma unroll
        for (int k = 0; k < 7; k++)
            nd[k] = f.v[k];
        if (WK)
            slotOf[f.slot][r] = (unsigned char)j;
    }
    // material of the cell (per-cell fields: E, kappa0 = f_t / E, beta = f_t / G_f)
    double E = P.E, k0 = P.kappa0, beta = P.beta;
    if (A.thE && valid)
    {
        E = f.th[0];
        k0 = f.th[1] / E;
        beta = f.th[1] / f.th[2];
    }
    const double m3K = P.cK * E, m2G = P.cG * E;
    __syncthreads();

    // ---- phase A: lane (r, ip) ---------------------------------------------------------------------
    double Ru0 = 0.0, Ru1 = 0.0, Ru2 = 0.0, Re = 0.0;
    if (valid)
    {
        const int ip = j;
        const double w = f.w;
        const double* nd = nodeData + r * kNd;
        // J = sum_n X_n dN_n^T, Hu = sum_n u_n dN_n^T, he = sum_n e_n dN_n, eN = sum_n N_n e_n
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        double Hu[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        double he[3] = {0, 0, 0};
        double eN = 0.0;
#pragma unroll
        for (int n = 0; n < 8; n++)
        {
            double d[3], Nv;
            ShapeAt(n, ip, d, Nv);
            const double2 xy = *reinterpret_cast<const double2*>(nd + n * 8);
            const double2 zu = *reinterpret_cast<const double2*>(nd + n * 8 + 2);
            const double2 uv = *reinterpret_cast<const double2*>(nd + n * 8 + 4);
            const double en = nd[n * 8 + 6];
            const double X[3] = {xy.x, xy.y, zu.x}, U[3] = {zu.y, uv.x, uv.y};
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int k = 0; k < 3; k++)
                {
                    J[i][k] += X[i] * d[k];
                    Hu[i][k] += U[i] * d[k];
                }
            he[0] += en * d[0];
            he[1] += en * d[1];
            he[2] += en * d[2];
            eN += Nv * en;
        }
        double inv[3][3];
        const double detJ = Inverse<3>(J, inv);
        const double s = detJ * w;
        // displacement gradient Gu = Hu J^-1, strain in Voigt order [xx, yy, zz, yz, zx, xy]
        double Gu[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int c = 0; c < 3; c++)
                Gu[i][c] = Hu[i][0] * inv[0][c] + Hu[i][1] * inv[1][c] + Hu[i][2] * inv[2][c];
        const double eps[6] = {Gu[0][0], Gu[1][1], Gu[2][2], Gu[1][2] + Gu[2][1], Gu[2][0] + Gu[0][2],
                               Gu[0][1] + Gu[1][0]};
        double ge[3];
        GradOf(he, inv, ge);

        // kappa = max(history, eeq), d kappa / d eeq (GradientDamage.h:127-135)
        const double kh = f.kh;
        const double kap = kh < eN ? eN : kh;
        const double dk = eN >= kh ? 1.0 : 0.0;
        // DamageLawExponential (DamageLawExponential.h:39-52)
        double om = 0.0, dom = 0.0;
        if (!(kap < k0))
        {
            const double ik = 1.0 / kap;
            const double ex = exp(beta * (k0 - kap));
            const double rr = k0 * ik;
            om = 1.0 - rr * (1.0 - P.alpha + P.alpha * ex);
            dom = rr * ((ik + beta) * P.alpha * ex + (1.0 - P.alpha) * ik);
        }
        // LinearElasticDamage<3> stress and dStress/dOmega (LinearElasticDamage.h:65-91)
        const double eV = (eps[0] + eps[1] + eps[2]) * kThird;
        const double H = (!FULLDMG && eV < 0.0) ? 0.0 : 1.0;
        const double sV = (1.0 - om * H) * m3K * eV;
        const double d2G = (1.0 - om) * m2G;
        double sig[6], dsw[6];
#pragma unroll
        for (int i = 0; i < 3; i++)
        {
            const double dev = eps[i] - eV;
            sig[i] = sV + d2G * dev;
            dsw[i] = -m3K * eV * H - m2G * dev;
        }
#pragma unroll
        for (int i = 3; i < 6; i++)
        {
            sig[i] = 0.5 * d2G * eps[i];
            dsw[i] = -0.5 * m2G * eps[i];
        }
        // ModifiedMisesStrainNorm<3> value and derivative (ModifiedMisesStrainNorm.h:136-146,231-268)
        const double I1 = eps[0] + eps[1] + eps[2];
        const double a0 = eps[0] - eps[1], a1 = eps[1] - eps[2], a2 = eps[2] - eps[0];
        const double J2 = (a0 * a0 + a1 * a1 + a2 * a2) * kSixth +
                          0.25 * (eps[3] * eps[3] + eps[4] * eps[4] + eps[5] * eps[5]);
        const double Am = sqrt(P.K1 * P.K1 * I1 * I1 + P.K2 * J2);
        const double eq = P.K1 * I1 + Am;

        double da[3], Na, ga[3];
        ShapeAt(la, ip, da, Na);
        GradOf(da, inv, ga);
        if (WR)
        {
            double bt[3];
            BTv3(ga, sig, bt);
            Ru0 = s * bt[0];
            Ru1 = s * bt[1];
            Ru2 = s * bt[2];
            Re = s * (Na * (eN - eq) + P.c * (ga[0] * ge[0] + ga[1] * ge[1] + ga[2] * ge[2]));
        }
        if (WK)
        {
            double m[6];
            if (Am == 0.0)
            {
                m[0] = m[1] = m[2] = P.K1;
                m[3] = m[4] = m[5] = 0.0;
            }
            else
            {
                const double h = 0.5 / Am;
                const double t2 = 2.0 * P.K1 * P.K1 * I1;
                m[0] = P.K1 + h * (t2 + P.K2 * (2.0 * eps[0] - eps[1] - eps[2]) * kThird);
                m[1] = P.K1 + h * (t2 + P.K2 * (2.0 * eps[1] - eps[0] - eps[2]) * kThird);
                m[2] = P.K1 + h * (t2 + P.K2 * (2.0 * eps[2] - eps[0] - eps[1]) * kThird);
                m[3] = h * P.K2 * 0.5 * eps[3];
                m[4] = h * P.K2 * 0.5 * eps[4];
                m[5] = h * P.K2 * 0.5 * eps[5];
            }
            double* st = state + (ip * 8 + r) * kSt;
            double rec[kSt];
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++)
                    rec[sInv + 3 * a + b] = inv[a][b];
            const double ct = s * dom * dk;
            double t6[6], va[3];
#pragma unroll
            for (int i = 0; i < 6; i++)
            {
                t6[i] = ct * dsw[i];
                rec[sMn + i] = -Na * s * m[i];
            }
            BTv3(ga, t6, va);
            const double c1 = FULLDMG ? s * (1.0 - om) : s * (m3K * (1.0 - om * H) - d2G) * kThird;
#pragma unroll
            for (int i = 0; i < 3; i++)
            {
                rec[sHa + i] = c1 * ga[i];
                rec[sVa + i] = va[i];
                rec[sCg + i] = P.c * s * ga[i];
                rec[sKa + i] = FULLDMG ? 0.0 : 0.5 * s * d2G * ga[i];
            }
            rec[sSn] = s * Na;
            rec[28] = 0.0;
            rec[29] = 0.0;
#pragma unroll
            for (int k = 0; k < kSt; k += 2)
                *reinterpret_cast<double2*>(st + k) = make_double2(rec[k], rec[k + 1]);
        }
    }
    __syncthreads();

    // ---- phase B: lane (r, b) accumulates block (a_r, b) over the 8 IPs ------------------------------
    double acc[16];
#pragma unroll
    for (int i = 0; i < 16; i++)
        acc[i] = 0.0;
    if (WK && valid)
    {
        const int b = j;
#pragma unroll 1
        for (int ip = 0; ip < 8; ip++)
        {
            double db[3], Nb;
            ShapeAt(b, ip, db, Nb);
            const double* st = state + (ip * 8 + r) * kSt;
            double rec[28];
#pragma unroll
            for (int k = 0; k < 28; k += 2)
            {
                const double2 v = *reinterpret_cast<const double2*>(st + k);
                rec[k] = v.x;
                rec[k + 1] = v.y;
            }
            const double inv[3][3] = {{rec[0], rec[1], rec[2]}, {rec[3], rec[4], rec[5]}, {rec[6], rec[7], rec[8]}};
            double gb[3];
            GradOf(db, inv, gb);
            const double* ha = rec + sHa;
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int c = 0; c < 3; c++)
                    acc[i * 4 + c] += ha[i] * gb[c];
            if (!FULLDMG)
            {
                const double* ka = rec + sKa;
                const double d = ka[0] * gb[0] + ka[1] * gb[1] + ka[2] * gb[2];
#pragma unroll
                for (int i = 0; i < 3; i++)
                {
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        acc[i * 4 + c] += gb[i] * ka[c];
                    acc[i * 5] += d;
                }
            }
            acc[3] += rec[sVa] * Nb;
            acc[7] += rec[sVa + 1] * Nb;
            acc[11] += rec[sVa + 2] * Nb;
            double wb[3];
            BTv3(gb, rec + sMn, wb);
            acc[12] += wb[0];
            acc[13] += wb[1];
            acc[14] += wb[2];
            acc[15] += rec[sSn] * Nb + rec[sCg] * gb[0] + rec[sCg + 1] * gb[1] + rec[sCg + 2] * gb[2];
        }
        if (FULLDMG)
        {
            // uu = lambda' S + mu' S^T + mu' tr(S) I
            const double lam = (m3K - m2G) * kThird, mu = 0.5 * m2G;
            const double S[3][3] = {{acc[0], acc[1], acc[2]}, {acc[4], acc[5], acc[6]}, {acc[8], acc[9], acc[10]}};
            const double tr = S[0][0] + S[1][1] + S[2][2];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int c = 0; c < 3; c++)
                    acc[i * 4 + c] = lam * S[i][c] + mu * S[c][i] + (i == c ? mu * tr : 0.0);
        }
    }
    __syncthreads();

    // ---- partial blocks and R terms to LDS -----------------------------------------------------------
    double* part = state;                   // [r][b][kPs]
    double* rterm = state + 8 * 8 * kPs;    // [4][r][ip]
    double* rcell = rterm + 4 * 64;         // [4][r]: R terms summed over the IPs of cell r
    double* zero = rcell + 32;              // 8 zeros: the block of an absent (cell, slot) pair
    if (WK)
    {
#pragma unroll
        for (int i = 0; i < 16; i++)
            part[lane * kPs + i] = acc[i];
        if (lane < 8)
            zero[lane] = 0.0;
    }
    if (WR)
    {
        rterm[0 * 64 + lane] = Ru0;
        rterm[1 * 64 + lane] = Ru1;
        rterm[2 * 64 + lane] = Ru2;
        rterm[3 * 64 + lane] = Re;
    }
    __syncthreads();

    // ---- reduce: lane (q, half) sums its 8 block entries over the cells in order ---------------------
    // (absent cells add the zero block, so every lane runs the same 8 steps)
    if (WK)
    {
        const int q = lane >> 1, h = lane & 1;
        if (q < nA)
        {
            const unsigned long long row = *reinterpret_cast<const unsigned long long*>(&slotOf[q][0]);
            double sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int rr = 0; rr < 8; rr++)
            {
                const unsigned b = (unsigned)(row >> (8 * rr)) & 0xFFu;
                const double* src = b == 0xFFu ? zero : part + (rr * 8 + (int)b) * kPs + h * 8;
#pragma unroll
                for (int k = 0; k < 8; k++)
                    sum[k] += src[k];
            }
            // block entry (i, c): i = row component (0..2 displacement rows, 3 eeq row),
            // c = column component (0..2 displacement columns, 3 eeq column)
#pragma unroll
            for (int k = 0; k < 8; k++)
            {
                const int e = h * 8 + k;
                const int i = e >> 2, c = e & 3;
                const int col = c < 3 ? 3 * q + c : 3 * nA + q;
                rowbuf[i * L + col] = sum[k];
            }
        }
    }
    // R: per cell over its IPs in order (Cell.h:108-117), then over the cells in order
    if (WR && lane < 32)
    {
        const double* t = rterm + (lane >> 3) * 64 + (lane & 7) * 8;
        double s = t[0];
#pragma unroll
        for (int ip = 1; ip < 8; ip++)
            s += t[ip];
        rcell[lane] = s;
    }
    __syncthreads();

    // ---- write p's rows (displacement rows are consecutive in the CSR) -------------------------------
    if (WK)
    {
        double v[6], ve[2];
#pragma unroll
        for (int it = 0; it < 6; it++)
        {
            const int i = lane + 64 * it;
            v[it] = i < 3 * L ? rowbuf[i] : 0.0;
        }
#pragma unroll
        for (int it = 0; it < 2; it++)
        {
            const int i = lane + 64 * it;
            ve[it] = i < L ? rowbuf[3 * L + i] : 0.0;
        }
#pragma unroll
        for (int it = 0; it < 6; it++)
        {
            const int i = lane + 64 * it;
            if (i < 3 * L)
                A.K[(size_t)ru0 + i] = v[it];
        }
#pragma unroll
        for (int it = 0; it < 2; it++)
        {
            const int i = lane + 64 * it;
            if (i < L)
                A.K[(size_t)re0 + i] = ve[it];
        }
    }
    if (WR && lane < 4)
    {
        const double* t = rcell + lane * 8;
        double s = 0.0;
#pragma unroll
        for (int rr = 0; rr < 8; rr++)
            s += t[rr];
        if (lane < 3)
            A.R[(size_t)p * 3 + lane] = s;
        else
            A.R[(size_t)A.nU + p] = s;
    }
}

struct GdmHex8Plan
{
    bool ready = false;
    int nNodes = 0;
    int grid = 0;     // workgroups: one per node (rounded up to a multiple of 8 with the XCD mapping)
    int xcdMap = 1;   // contiguous node ranges per XCD (env NUTO_AMD_GDM_XCD_MAP=0 switches it off)
    DevBuf nodeCell, nodeHist, nodeLa, nodeSlot, xd, yd, zd;

    void Release()
    {
        ready = false;
        nNodes = 0;
        for (DevBuf* b : {&nodeCell, &nodeHist, &nodeLa, &nodeSlot, &xd, &yd, &zd})
            b->Release();
    }

    bool Usable(int kind, int coordType, int dofType, int nip) const
    {
        return ready && kind == kGradientDamage && coordType == kHex8 && dofType == kHex8 && nip == 8;
    }

    // Builds the per-node tables from the node graph. Returns false (plan not used) when the mesh
    // exceeds the kernel's bounds or a dof node sits on coordinate nodes with different
    // coordinates; returns a HIP error through `err` when an upload fails.
    bool Prepare(int kind, int coordType, int dofType, int nip, const NodeGraph& g, int nCells, const int* dconn,
                 const int* cconn, const double* hx, const double* hy, const double* hz, const int* cellIds,
                 hipStream_t stream, hipError_t& err)
    {
        Release();
        err = hipSuccess;
        if (kind != kGradientDamage || coordType != kHex8 || dofType != kHex8 || nip != 8)
            return false;
        const int n = g.nNodes;
        for (int p = 0; p < n; p++)
        {
            if (g.incPtr[p + 1] - g.incPtr[p] > gdm8::kMaxInc || g.adjPtr[p + 1] - g.adjPtr[p] > gdm8::kMaxAdj)
                return false;
        }
        // coordinates by dof node: every coordinate node behind one dof node must agree exactly
        std::vector<double> X(n), Y(n), Z(n);
        std::vector<unsigned char> seen(n, 0);
        for (size_t i = 0; i < (size_t)nCells * 8; i++)
        {
            const int q = dconn[i], c = cconn[i];
            if (!seen[q])
            {
                seen[q] = 1;
                X[q] = hx[c];
                Y[q] = hy[c];
                Z[q] = hz[c];
            }
            else if (X[q] != hx[c] || Y[q] != hy[c] || Z[q] != hz[c])
                return false;
        }
        std::vector<int> cell((size_t)n * 8, -1), hist(cellIds ? (size_t)n * 8 : 0, 0);
        std::vector<unsigned char> la((size_t)n * 8, 0), slot((size_t)n * 64, 0);
        for (int p = 0; p < n; p++)
        {
            for (int k = g.incPtr[p], r = 0; k < g.incPtr[p + 1]; k++, r++)
            {
                const size_t pr = (size_t)p * 8 + r;
                cell[pr] = g.incCell[k];
                la[pr] = g.incLocal[k];
                if (cellIds)
                    hist[pr] = cellIds[g.incCell[k]];
                for (int j = 0; j < 8; j++)
                    slot[pr * 8 + j] = (unsigned char)(g.slot[(size_t)k * 8 + j] & 0x7fff);
            }
        }
        if ((err = nodeCell.Upload(cell.data(), cell.size() * sizeof(int), stream)) != hipSuccess ||
            (err = nodeLa.Upload(la.data(), la.size(), stream)) != hipSuccess ||
            (err = nodeSlot.Upload(slot.data(), slot.size(), stream)) != hipSuccess ||
            (err = xd.Upload(X.data(), X.size() * sizeof(double), stream)) != hipSuccess ||
            (err = yd.Upload(Y.data(), Y.size() * sizeof(double), stream)) != hipSuccess ||
            (err = zd.Upload(Z.data(), Z.size() * sizeof(double), stream)) != hipSuccess ||
            (cellIds && (err = nodeHist.Upload(hist.data(), hist.size() * sizeof(int), stream)) != hipSuccess))
        {
            Release();
            return false;
        }
        nNodes = n;
        const char* env = std::getenv("NUTO_AMD_GDM_XCD_MAP");
        xcdMap = env ? std::atoi(env) : 1;
        grid = xcdMap ? (nNodes + 7) / 8 * 8 : nNodes;
        ready = true;
        return true;
    }

    GdmHex8Tables Tables() const
    {
        return GdmHex8Tables{nodeCell.As<int>(), nodeHist.p ? nodeHist.As<int>() : nullptr,
                             nodeLa.As<unsigned char>(), nodeSlot.As<unsigned char>(), xd.As<double>(),
                             yd.As<double>(), zd.As<double>(), xcdMap};
    }

    template <bool FULLDMG>
    void LaunchT(const RowArgs& a, bool wr, bool wk, hipStream_t stream)
    {
        const GdmHex8Tables t = Tables();
        if (wr && wk)
            hipLaunchKernelGGL((kGdmHex8Node<true, true, FULLDMG>), dim3(grid), dim3(64), 0, stream, a, t);
        else if (wr)
            hipLaunchKernelGGL((kGdmHex8Node<true, false, FULLDMG>), dim3(grid), dim3(64), 0, stream, a, t);
        else
            hipLaunchKernelGGL((kGdmHex8Node<false, true, FULLDMG>), dim3(grid), dim3(64), 0, stream, a, t);
    }

    void Launch(const RowArgs& a, bool wr, bool wk, hipStream_t stream)
    {
        if (nNodes == 0)
            return;
        if (a.P.damage_application == 0)
            LaunchT<true>(a, wr, wk, stream);
        else
            LaunchT<false>(a, wr, wk, stream);
    }
};

}  // namespace nuto_amd
