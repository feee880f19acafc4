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
// meshes: 8 and 27); other meshes use the generic row-owner kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "assemble_rows.hpp"
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

// Inputs of one node gathered from global memory (issued one node ahead of the computation).
struct NodeFetch
{
    int beg, cnt, nA, cell, la, slot;
    double v[7];  // x y z u0 u1 u2 e of node j of cell r
    double th[3];
};

__device__ inline void FetchNode(const RowArgs& A, int p, int r, int j, NodeFetch& f)
{
    f.beg = A.incPtr[p];
    f.cnt = A.incPtr[p + 1] - f.beg;
    f.nA = A.adjPtr[p + 1] - A.adjPtr[p];
    f.cell = 0;
    f.la = 0;
    f.slot = 0;
    if (r < f.cnt)
    {
        f.cell = A.incCell[f.beg + r];
        f.la = A.incLocal[f.beg + r];
        f.slot = A.slot[(size_t)(f.beg + r) * 8 + j] & 0x7fff;
        const int cn = A.cconn[(size_t)f.cell * 8 + j];
        const int q = A.dconn[(size_t)f.cell * 8 + j];
        f.v[0] = A.x[cn];
        f.v[1] = A.y[cn];
        f.v[2] = A.z[cn];
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
__global__ void __launch_bounds__(64) kGdmHex8Node(const RowArgs A, int xcdMap)
{
    using namespace gdm8;
    __shared__ __align__(16) double state[8 * 8 * kSt];  // [ip][r][kSt]; later partial blocks + R terms
    __shared__ __align__(16) double rowbuf[8 * kNd];      // node data of the incident cells, later p's rows
    __shared__ unsigned char slotOf[kMaxAdj][kMaxInc];

    const int lane = threadIdx.x;
    const int r = lane >> 3, j = lane & 7;
    const LawParams& P = A.P;
    int p = blockIdx.x;
    if (xcdMap)
    {
        const int per = (int)((gridDim.x + 7) >> 3);
        p = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
        if (p >= A.nNodes)
            return;
    }
    NodeFetch f;
    FetchNode(A, p, r, j, f);
    const int beg = f.beg;
    const int cnt = f.cnt;
    const int nA = f.nA;
    const int L = 4 * nA;
    const int cell = f.cell, la = f.la;

    // ---- node data of p's cells to LDS; map slots -> (r, b) -----------------------------------------
    double* nodeData = rowbuf;  // [r][n][8] (+2 per cell): x y z u0 u1 u2 e -
    for (int i = lane; i < kMaxAdj * kMaxInc; i += 64)
        (&slotOf[0][0])[i] = 0xFF;
    __syncthreads();
    if (r < cnt)
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
    if (A.thE && r < cnt)
    {
        E = f.th[0];
        k0 = f.th[1] / E;
        beta = f.th[1] / f.th[2];
    }
    const double m3K = P.cK * E, m2G = P.cG * E;
    __syncthreads();

    // ---- phase A: lane (r, ip) ---------------------------------------------------------------------
    double Ru0 = 0.0, Ru1 = 0.0, Ru2 = 0.0, Re = 0.0;
    if (r < cnt)
    {
        const int ip = j;
        const double w = A.tab[ip * kTab];
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
        const int cid = A.cellId ? A.cellId[cell] : cell;
        const double kh = A.kappa[(size_t)cid + (size_t)ip * A.kappaLd];
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
    if (WK && r < cnt)
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
    double* part = state;                 // [r][b][kPs]
    double* rterm = state + 8 * 8 * kPs;  // [4][64]
    if (WK)
    {
#pragma unroll
        for (int i = 0; i < 16; i++)
            part[lane * kPs + i] = acc[i];
    }
    if (WR)
    {
        rterm[0 * 64 + lane] = Ru0;
        rterm[1 * 64 + lane] = Ru1;
        rterm[2 * 64 + lane] = Ru2;
        rterm[3 * 64 + lane] = Re;
    }
    for (int i = lane; i < 16 * kMaxAdj; i += 64)
        rowbuf[i] = 0.0;  // p's four rows, 16 * kMaxAdj <= 8 * kNd
    __syncthreads();

    // ---- reduce: lane (q, half) sums its 8 block entries over the cells in order ---------------------
    if (WK)
    {
        const int q = lane >> 1, h = lane & 1;
        if (q < nA)
        {
            double sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int rr = 0; rr < cnt; rr++)
            {
                const int b = slotOf[q][rr];
                if (b != 0xFF)
                {
                    const double* src = part + (rr * 8 + b) * kPs + h * 8;
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        sum[k] += src[k];
                }
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
    __syncthreads();

    // ---- write p's rows (displacement rows are consecutive in the CSR) -------------------------------
    if (WK)
    {
        const int ru0 = A.rowPtr[p * 3];
        const int re0 = A.rowPtr[A.nU + p];
        for (int i = lane; i < 3 * L; i += 64)
            A.K[(size_t)ru0 + i] = rowbuf[i];
        for (int i = lane; i < L; i += 64)
            A.K[(size_t)re0 + i] = rowbuf[3 * L + i];
    }
    if (WR && lane < 4)
    {
        double s = 0.0;
        for (int l = 0; l < 8 * cnt; l++)
            s += rterm[lane * 64 + l];
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

    void Release()
    {
        ready = false;
    }

    bool Usable(int kind, int coordType, int dofType, int nip, bool) const
    {
        return ready && kind == kGradientDamage && coordType == kHex8 && dofType == kHex8 && nip == 8;
    }

    void Prepare(int kind, int coordType, int dofType, int nip, bool, const NodeGraph& g, int, hipStream_t)
    {
        ready = false;
        if (kind != kGradientDamage || coordType != kHex8 || dofType != kHex8 || nip != 8)
            return;
        for (int p = 0; p < g.nNodes; p++)
        {
            if (g.incPtr[p + 1] - g.incPtr[p] > gdm8::kMaxInc || g.adjPtr[p + 1] - g.adjPtr[p] > gdm8::kMaxAdj)
                return;
        }
        nNodes = g.nNodes;
        const char* env = std::getenv("NUTO_AMD_GDM_XCD_MAP");
        xcdMap = env ? std::atoi(env) : 1;
        grid = xcdMap ? (nNodes + 7) / 8 * 8 : nNodes;
        ready = true;
    }

    int grid = 0;     // workgroups: one per node, rounded up to a multiple of 8
    int xcdMap = 1;   // contiguous node ranges per XCD

    template <bool FULLDMG>
    void LaunchT(const RowArgs& a, bool wr, bool wk, hipStream_t stream)
    {
        if (wr && wk)
            hipLaunchKernelGGL((kGdmHex8Node<true, true, FULLDMG>), dim3(grid), dim3(64), 0, stream, a, xcdMap);
        else if (wr)
            hipLaunchKernelGGL((kGdmHex8Node<true, false, FULLDMG>), dim3(grid), dim3(64), 0, stream, a, xcdMap);
        else
            hipLaunchKernelGGL((kGdmHex8Node<false, true, FULLDMG>), dim3(grid), dim3(64), 0, stream, a, xcdMap);
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
