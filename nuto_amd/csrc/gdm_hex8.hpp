// Hex8 gradient-damage R+K kernel (the hot path): one 64-lane workgroup per dof node p.
//
// Lanes are (r, j) with r = the r-th cell incident to p (group order) and j = 0..7.
//   phase A  lane (r, ip=j): the integration-point state of cell r at IP ip (Jacobian, strain,
//            nonlocal strain, damage, stress, equivalent-strain derivative) -> LDS
//   phase B  lane (r, b=j):  the 4x4 block (a_r, b) of cell r's element matrix for p's rows
//            (a_r = local index of p), summed over the 8 IPs in order (Cell.h:108-117)
//   reduce   slot q of p's neighbourhood = sum over r (cell order, the order setFromTriplets
//            uses, SimpleAssembler.cpp:149-155) of the blocks that land on q
//   write    p's three displacement rows (contiguous) and its eeq row, coalesced, written once.
// R is reduced the same way (cells in order, IPs in order). No atomics: the result is
// deterministic, and every nonzero and R entry is written exactly once per assembly.
//
// Usable when every node has at most 8 incident cells and at most 32 neighbours (structured hex
// meshes: 8 and 27); other meshes use the generic row-owner kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "assemble_rows.hpp"
#include "laws.hpp"
#include "mesh.hpp"

namespace nuto_amd
{

namespace gdm8
{
constexpr int kMaxInc = 8;    // incident cells per node
constexpr int kMaxAdj = 32;   // neighbours per node
constexpr int kSt = 25;       // doubles per (ip, cell) state record (odd stride: no LDS bank conflicts)
constexpr int kTab = 57;      // per-IP table: w, dNc[8][3], N[8], dNd[8][3]

// state record layout
constexpr int sInv = 0;   // J^-1 (9)
constexpr int sGa = 9;    // g_a = dN_a/dX (3)
constexpr int sVa = 12;   // B_a^T t, t = s * dOmega/dKappa * dKappa/deeq * dSigma/dOmega (3)
constexpr int sMm = 15;   // s * d eps_eq / d eps (6)
constexpr int sC1 = 21;   // uu: c1 g_a g_b^T + c2 g_b g_a^T + c2 (g_a.g_b) I
constexpr int sC2 = 22;
constexpr int sS = 23;    // s = detJ * w
constexpr int sCm = 24;   // c * s
}  // namespace gdm8

template <bool WR, bool WK>
__global__ void __launch_bounds__(64) kGdmHex8Node(const RowArgs A)
{
    using namespace gdm8;
    __shared__ double state[8 * 8 * kSt];    // [ip][r][kSt]; later partial blocks [r][b][16] + R terms
    __shared__ double rowbuf[16 * kMaxAdj];  // node data of the incident cells, later p's four rows
    __shared__ unsigned char slotOf[kMaxAdj][kMaxInc];

    const int p = blockIdx.x;
    const int lane = threadIdx.x;
    const int r = lane >> 3, j = lane & 7;
    const int beg = A.incPtr[p];
    const int cnt = A.incPtr[p + 1] - beg;
    const int nA = A.adjPtr[p + 1] - A.adjPtr[p];
    const int L = 4 * nA;
    const LawParams& P = A.P;

    // ---- load: lane (r, n=j) brings node n of cell r into LDS; map slots -> (r, b) --------------
    double* nodeData = rowbuf;  // [r][n][7]: x y z u0 u1 u2 e
    for (int i = lane; i < kMaxAdj * kMaxInc; i += 64)
        (&slotOf[0][0])[i] = 0xFF;
    __syncthreads();
    int cell = 0, la = 0;
    if (r < cnt)
    {
        cell = A.incCell[beg + r];
        la = A.incLocal[beg + r];
        const int cn = A.cconn[(size_t)cell * 8 + j];
        const int q = A.dconn[(size_t)cell * 8 + j];
        double* nd = nodeData + (r * 8 + j) * 7;
        nd[0] = A.x[cn];
        nd[1] = A.y[cn];
        nd[2] = A.z[cn];
        nd[3] = A.u[(size_t)q * 3];
        nd[4] = A.u[(size_t)q * 3 + 1];
        nd[5] = A.u[(size_t)q * 3 + 2];
        nd[6] = A.e[q];
        if (WK)
            slotOf[A.slot[(size_t)(beg + r) * 8 + j] & 0x7fff][r] = (unsigned char)j;
    }
    __syncthreads();

    // ---- phase A: lane (r, ip) ---------------------------------------------------------------------
    double Ru0 = 0.0, Ru1 = 0.0, Ru2 = 0.0, Re = 0.0;
    if (r < cnt)
    {
        const int ip = j;
        const double* t = A.tab + ip * kTab;
        const double w = t[0];
        const double* dN = t + 1;
        const double* Nn = t + 25;
        const double* nd = nodeData + r * 8 * 7;
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int n = 0; n < 8; n++)
        {
            const double d0 = dN[3 * n], d1 = dN[3 * n + 1], d2 = dN[3 * n + 2];
#pragma unroll
            for (int i = 0; i < 3; i++)
            {
                const double xi = nd[n * 7 + i];
                J[i][0] += xi * d0;
                J[i][1] += xi * d1;
                J[i][2] += xi * d2;
            }
        }
        double inv[3][3];
        const double detJ = Inverse<3>(J, inv);
        const double s = detJ * w;
        double eps[6] = {0, 0, 0, 0, 0, 0};
        double eN = 0.0, ge0 = 0.0, ge1 = 0.0, ge2 = 0.0;
#pragma unroll
        for (int n = 0; n < 8; n++)
        {
            double g[3];
            Grad<3, 8>(dN, n, inv, g);
            const double u0 = nd[n * 7 + 3], u1 = nd[n * 7 + 4], u2 = nd[n * 7 + 5], en = nd[n * 7 + 6];
            eps[0] += g[0] * u0;
            eps[1] += g[1] * u1;
            eps[2] += g[2] * u2;
            eps[3] += g[2] * u1 + g[1] * u2;
            eps[4] += g[2] * u0 + g[0] * u2;
            eps[5] += g[1] * u0 + g[0] * u1;
            eN += Nn[n] * en;
            ge0 += g[0] * en;
            ge1 += g[1] * en;
            ge2 += g[2] * en;
        }
        double E = P.E, k0 = P.kappa0, beta = P.beta;
        if (A.thE)
        {
            E = A.thE[cell];
            k0 = A.thFt[cell] / E;
            beta = A.thFt[cell] / A.thGf[cell];
        }
        const int cid = A.cellId ? A.cellId[cell] : cell;
        const double kh = A.kappa[(size_t)cid + (size_t)ip * A.kappaLd];
        const double kap = kh < eN ? eN : kh;
        const double dk = eN >= kh ? 1.0 : 0.0;
        double om, dom;
        DamageExp(kap, k0, beta, P.alpha, om, dom);
        DamageStress3 ds;
        ElasticDamage3(eps, om, E, P.nu, P.damage_application, ds);
        double m[6];
        const double eq = Mises3(eps, P.nu, P.k, m);
        double ga[3];
        Grad<3, 8>(dN, la, inv, ga);
        const double Na = Nn[la];
        if (WR)
        {
            double bt[3];
            BTv3(ga, ds.sig, bt);
            Ru0 = s * bt[0];
            Ru1 = s * bt[1];
            Ru2 = s * bt[2];
            Re = s * (Na * (eN - eq) + P.c * (ga[0] * ge0 + ga[1] * ge1 + ga[2] * ge2));
        }
        if (WK)
        {
            double* st = state + (ip * 8 + r) * kSt;
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++)
                    st[sInv + 3 * a + b] = inv[a][b];
            st[sGa] = ga[0];
            st[sGa + 1] = ga[1];
            st[sGa + 2] = ga[2];
            double t6[6], va[3];
            const double ct = s * dom * dk;
#pragma unroll
            for (int i = 0; i < 6; i++)
            {
                t6[i] = ct * ds.dsdw[i];
                st[sMm + i] = s * m[i];
            }
            BTv3(ga, t6, va);
            st[sVa] = va[0];
            st[sVa + 1] = va[1];
            st[sVa + 2] = va[2];
            st[sC1] = s * (ds.alpha_v - ds.alpha_d) / 3.0;
            st[sC2] = s * 0.5 * ds.alpha_d;
            st[sS] = s;
            st[sCm] = P.c * s;
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
#pragma unroll 2
        for (int ip = 0; ip < 8; ip++)
        {
            const double* t = A.tab + ip * kTab;
            const double* dN = t + 1;
            const double Nb = t[25 + b], Na = t[25 + la];
            const double* st = state + (ip * 8 + r) * kSt;
            double inv[3][3];
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int c = 0; c < 3; c++)
                    inv[a][c] = st[sInv + 3 * a + c];
            double gb[3];
            Grad<3, 8>(dN, b, inv, gb);
            const double ga0 = st[sGa], ga1 = st[sGa + 1], ga2 = st[sGa + 2];
            const double c1 = st[sC1], c2 = st[sC2];
            const double dot = ga0 * gb[0] + ga1 * gb[1] + ga2 * gb[2];
            const double ga[3] = {ga0, ga1, ga2};
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int c = 0; c < 3; c++)
                    acc[i * 4 + c] += c1 * ga[i] * gb[c] + c2 * gb[i] * ga[c] + (i == c ? c2 * dot : 0.0);
            double mm[6];
#pragma unroll
            for (int i = 0; i < 6; i++)
                mm[i] = st[sMm + i];
            double wb[3];
            BTv3(gb, mm, wb);
            acc[3] += st[sVa] * Nb;
            acc[7] += st[sVa + 1] * Nb;
            acc[11] += st[sVa + 2] * Nb;
            acc[12] -= Na * wb[0];
            acc[13] -= Na * wb[1];
            acc[14] -= Na * wb[2];
            acc[15] += st[sS] * Na * Nb + st[sCm] * dot;
        }
    }
    __syncthreads();

    // ---- partial blocks and R terms to LDS -----------------------------------------------------------
    double* part = state;              // [r][b][16]
    double* rterm = state + 8 * 8 * 16;  // [4][64]
    if (WK)
    {
#pragma unroll
        for (int i = 0; i < 16; i++)
            part[lane * 16 + i] = acc[i];
    }
    if (WR)
    {
        rterm[0 * 64 + lane] = Ru0;
        rterm[1 * 64 + lane] = Ru1;
        rterm[2 * 64 + lane] = Ru2;
        rterm[3 * 64 + lane] = Re;
    }
    for (int i = lane; i < 16 * kMaxAdj; i += 64)
        rowbuf[i] = 0.0;
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
                    const double* src = part + (rr * 8 + b) * 16 + h * 8;
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

    // kind / element / rule / history-id conditions for this kernel
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
        ready = true;
    }

    void Launch(const RowArgs& a, bool wr, bool wk, hipStream_t stream)
    {
        if (nNodes == 0)
            return;
        if (wr && wk)
            hipLaunchKernelGGL((kGdmHex8Node<true, true>), dim3(nNodes), dim3(64), 0, stream, a);
        else if (wr)
            hipLaunchKernelGGL((kGdmHex8Node<true, false>), dim3(nNodes), dim3(64), 0, stream, a);
        else
            hipLaunchKernelGGL((kGdmHex8Node<false, true>), dim3(nNodes), dim3(64), 0, stream, a);
    }
};

}  // namespace nuto_amd
