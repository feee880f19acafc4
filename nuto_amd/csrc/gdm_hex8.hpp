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
// local index, history column, neighbour slots), then the cell's dof node ranks, its history and
// its geometry, then the node values. The geometry of every (cell, IP) - J^-1 and detJ * w
// (Jacobian.h:35-47, Cell.h:108-117) - depends on the coordinates only; it is computed once per
// pattern by kGdmHex8Geometry with the same arithmetic and stored (80 bytes per IP).
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
#include "morton.hpp"

namespace nuto_amd
{

namespace gdm8
{
constexpr int kMaxInc = 8;   // incident cells per node
constexpr int kMaxAdj = 32;  // neighbours per node
// doubles per (ip, cell) record: 26 for FULL damage application, 30 with the UNILATERAL terms.
// Records are 13 (15) 16-byte chunks long: an odd count, so 16 consecutive records start on 16
// distinct 4-bank groups.
template <bool FULLDMG>
constexpr int kStOf = FULLDMG ? 26 : 30;
constexpr int kGeo = 10;     // doubles per (cell, IP) geometry record: J^-1 (row-major), detJ * w
// per (cell, IP) state record of the pre-pass (kGdmHex8Ip): with e'' = [e0, e1, e2, e3/2, e4/2, e5/2]
// (tensor strain) the stress, t = s omega' kappa' d sigma / d omega and m = d eps_eq / d eps are all
// of the form alpha [1, 1, 1, 0, 0, 0] + beta e'', so B_a^T of each is alpha g_a + beta (e'' g_a)
constexpr int kSr = 18;
constexpr int rEps = 0;   // e'' (6)
constexpr int rSig = 6;   // s alpha_sigma, s beta_sigma
constexpr int rT = 8;     // alpha_t, beta_t (s included)
constexpr int rM = 10;    // s alpha_m, s beta_m
constexpr int rC1 = 12;   // FULL: s (1 - omega); UNILATERAL: s (3K (1 - omega H) - 2G (1 - omega)) / 3
constexpr int rC2 = 13;   // UNILATERAL: s G (1 - omega); FULL: 0
constexpr int rRe = 14;   // s (eeq - eps_eq)
constexpr int rGe = 15;   // s c grad eeq (3)
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
constexpr int kRed = 1160;  // doubles after phase B: partial blocks 64 x kPs, R sums 32, 8 unused, slot map 32

// value of lane (lane & ~7) + n within each group of 8 lanes (ds_swizzle bit mode:
// and-mask 0x18 keeps the group, or-mask n selects the lane)
template <int n>
__device__ inline double GroupLane(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_swizzle((int)(b & 0xffffffffLL), 0x18 | (n << 5));
    const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), 0x18 | (n << 5));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

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
// Shape data of a hex8 node at the 2x2x2 Gauss point whose corner differs from the node's in the
// axes of the set bits of t: f_k = 1 - g if bit k of t is set, else 1 + g (g = sqrt(1/3)), and
// [|dN/dxi_0|, |dN/dxi_1|, |dN/dxi_2|, N] = [f1 f2, f0 f2, f0 f1, f0 f1 f2] / 8, with the products
// taken in ShapeAt's order (so the values are bitwise ShapeAt's, up to the gradient signs).
struct ShapeRelTable
{
    double v[8][4];
};
constexpr ShapeRelTable MakeShapeRel()
{
    ShapeRelTable t{};
    for (int k = 0; k < 8; k++)
    {
        const double f0 = (k & 1) ? 1.0 - kGauss : 1.0 + kGauss;
        const double f1 = (k & 2) ? 1.0 - kGauss : 1.0 + kGauss;
        const double f2 = (k & 4) ? 1.0 - kGauss : 1.0 + kGauss;
        t.v[k][0] = 0.125 * f1 * f2;
        t.v[k][1] = 0.125 * f0 * f2;
        t.v[k][2] = 0.125 * f0 * f1;
        t.v[k][3] = 0.125 * f0 * f1 * f2;
    }
    return t;
}
__constant__ ShapeRelTable kShapeRel = MakeShapeRel();

__device__ inline void GradOf(const double d[3], const double inv[3][3], double g[3])
{
#pragma unroll
    for (int c = 0; c < 3; c++)
        g[c] = d[0] * inv[0][c] + d[1] * inv[1][c] + d[2] * inv[2][c];
}
}  // namespace gdm8

// Integration-point state of GradientDamage<3> (GradientDamage.h:70-135) from the displacement
// gradient w.r.t. xi (Hu), the nonlocal-strain gradient w.r.t. xi (he), the nonlocal strain (eN),
// J^-1 and the history kappa.
struct GdmIp
{
    double eps[6];   // engineering strain, Voigt [xx, yy, zz, yz, zx, xy]
    double ge[3];    // grad eeq
    double eN;       // eeq at the IP
    double dk;       // d kappa / d eeq
    double kap, om, dom;  // kappa, damage and its derivative
    double H;        // 1, or 0 for UNILATERAL application under volumetric compression
    double d2G;      // (1 - omega) 2G
    double sig[6], dsw[6];  // stress and d stress / d omega
    double I1, Am, eq;      // modified Mises: I1, sqrt term, value
};

template <bool FULLDMG>
__device__ inline void GdmIpState(const double Hu[3][3], const double he[3], double eN, const double inv[3][3],
                                  double kh, double m3K, double m2G, double k0, double beta, const LawParams& P,
                                  GdmIp& o)
{
    using namespace gdm8;
    // displacement gradient Gu = Hu J^-1
    double Gu[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            Gu[i][c] = Hu[i][0] * inv[0][c] + Hu[i][1] * inv[1][c] + Hu[i][2] * inv[2][c];
    o.eps[0] = Gu[0][0];
    o.eps[1] = Gu[1][1];
    o.eps[2] = Gu[2][2];
    o.eps[3] = Gu[1][2] + Gu[2][1];
    o.eps[4] = Gu[2][0] + Gu[0][2];
    o.eps[5] = Gu[0][1] + Gu[1][0];
    GradOf(he, inv, o.ge);
    o.eN = eN;
    // kappa = max(history, eeq), d kappa / d eeq (GradientDamage.h:127-135)
    o.kap = kh < eN ? eN : kh;
    o.dk = eN >= kh ? 1.0 : 0.0;
    // DamageLawExponential (DamageLawExponential.h:39-52)
    o.om = 0.0;
    o.dom = 0.0;
    if (!(o.kap < k0))
    {
        const double ik = 1.0 / o.kap;
        const double ex = exp(beta * (k0 - o.kap));
        const double rr = k0 * ik;
        o.om = 1.0 - rr * (1.0 - P.alpha + P.alpha * ex);
        o.dom = rr * ((ik + beta) * P.alpha * ex + (1.0 - P.alpha) * ik);
    }
    // LinearElasticDamage<3> stress and dStress/dOmega (LinearElasticDamage.h:65-91)
    const double* eps = o.eps;
    const double eV = (eps[0] + eps[1] + eps[2]) * kThird;
    o.H = (!FULLDMG && eV < 0.0) ? 0.0 : 1.0;
    const double sV = (1.0 - o.om * o.H) * m3K * eV;
    o.d2G = (1.0 - o.om) * m2G;
#pragma unroll
    for (int i = 0; i < 3; i++)
    {
        const double dev = eps[i] - eV;
        o.sig[i] = sV + o.d2G * dev;
        o.dsw[i] = -m3K * eV * o.H - m2G * dev;
    }
#pragma unroll
    for (int i = 3; i < 6; i++)
    {
        o.sig[i] = 0.5 * o.d2G * eps[i];
        o.dsw[i] = -0.5 * m2G * eps[i];
    }
    // ModifiedMisesStrainNorm<3> value (ModifiedMisesStrainNorm.h:136-146)
    o.I1 = eps[0] + eps[1] + eps[2];
    const double a0 = eps[0] - eps[1], a1 = eps[1] - eps[2], a2 = eps[2] - eps[0];
    const double J2 =
        (a0 * a0 + a1 * a1 + a2 * a2) * kSixth + 0.25 * (eps[3] * eps[3] + eps[4] * eps[4] + eps[5] * eps[5]);
    o.Am = sqrt(P.K1 * P.K1 * o.I1 * o.I1 + P.K2 * J2);
    o.eq = P.K1 * o.I1 + o.Am;
}

// d eps_eq / d eps of the modified Mises norm (ModifiedMisesStrainNorm.h:231-268, A == 0 branch)
__device__ inline void GdmMisesDerivative(const GdmIp& o, const LawParams& P, double m[6])
{
    using namespace gdm8;
    const double* eps = o.eps;
    if (o.Am == 0.0)
    {
        m[0] = m[1] = m[2] = P.K1;
        m[3] = m[4] = m[5] = 0.0;
    }
    else
    {
        const double h = 0.5 / o.Am;
        const double t2 = 2.0 * P.K1 * P.K1 * o.I1;
        m[0] = P.K1 + h * (t2 + P.K2 * (2.0 * eps[0] - eps[1] - eps[2]) * kThird);
        m[1] = P.K1 + h * (t2 + P.K2 * (2.0 * eps[1] - eps[0] - eps[2]) * kThird);
        m[2] = P.K1 + h * (t2 + P.K2 * (2.0 * eps[2] - eps[0] - eps[1]) * kThird);
        m[3] = h * P.K2 * 0.5 * eps[3];
        m[4] = h * P.K2 * 0.5 * eps[4];
        m[5] = h * P.K2 * 0.5 * eps[5];
    }
}

// Hu, he, eN of cell r at IP ip from the node values held by the 8 lanes of the cell
// (node n by lane (r, n)): Hu = sum_n u_n dN_n/dxi^T, he = sum_n e_n dN_n/dxi, eN = sum_n N_n e_n
__device__ inline void GatherGradients(const double v[4], int ip, double Hu[3][3], double he[3], double& eN)
{
    using namespace gdm8;
#pragma unroll
    for (int i = 0; i < 3; i++)
    {
        he[i] = 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++)
            Hu[i][k] = 0.0;
    }
    eN = 0.0;
#pragma unroll
    for (int n = 0; n < 8; n++)
    {
        double d[3], Nv;
        ShapeAt(n, ip, d, Nv);
        double U[3], en;
        switch (n)
        {
#define NUTO_GROUP_LANE(N)                                                                                 \
    case N:                                                                                                \
        U[0] = GroupLane<N>(v[0]);                                                                         \
        U[1] = GroupLane<N>(v[1]);                                                                         \
        U[2] = GroupLane<N>(v[2]);                                                                         \
        en = GroupLane<N>(v[3]);                                                                           \
        break;
            NUTO_GROUP_LANE(0)
            NUTO_GROUP_LANE(1)
            NUTO_GROUP_LANE(2)
            NUTO_GROUP_LANE(3)
            NUTO_GROUP_LANE(4)
            NUTO_GROUP_LANE(5)
            NUTO_GROUP_LANE(6)
            NUTO_GROUP_LANE(7)
#undef NUTO_GROUP_LANE
        }
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int k = 0; k < 3; k++)
                Hu[i][k] += U[i] * d[k];
        he[0] += en * d[0];
        he[1] += en * d[1];
        he[2] += en * d[2];
        eN += Nv * en;
    }
}

// Per-node tables of the node kernel (built once per pattern, see GdmHex8Plan::Prepare), stored in
// traversal order: position s holds the node order[s] (nodes sorted along a Morton curve).
struct GdmHex8Tables
{
    const int4* nodeHead;           // [pos]: node p, first CSR entry of its displacement rows, their
                                    //        length 4 nA, first CSR entry of its eeq row
    const int* nodeCell;            // [pos][8]: r-th incident cell (group order), -1 past the last
    const int* nodeHist;            // [pos][8]: history column Cell::Id of that cell (nullptr: the cell index)
    const unsigned char* nodeLa;    // [pos][8]: local index of the node in that cell
    const unsigned char* nodeSlot;  // [pos][8][8]: neighbour slot (column block) of local node j
    const double* geo;              // [cell][ip][kGeo] geometry records
    const double* ipState;          // [cell][ip][kSr] pre-pass state records (nullptr: recompute)
    const double* geoCell;          // [cell][kGeo] IP 0's geometry record (affine meshes, else nullptr)
    int xcdMap;                     // 1: contiguous node ranges per XCD
    int pBegin, pEnd;               // launch indices [pBegin, pEnd)
    const int* order;               // launch index -> position (nullptr: the index is the position)
};

// Table position of this workgroup, -1 past the end. Workgroups are dispatched round-robin over the
// 8 XCDs; with xcdMap, XCD x takes the x-th contiguous eighth of the traversal, which follows space
// (Morton order), so the nodes one XCD works on at a time share their cells' records in its L2.
// Launches over a node range (one mesh of a batch) index by node and map it through `order`.
__device__ inline int NodeOfBlock(const GdmHex8Tables& T)
{
    int i = blockIdx.x;
    if (T.xcdMap)
    {
        const int per = (int)((gridDim.x + 7) >> 3);
        i = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    }
    i += T.pBegin;
    if (i >= T.pEnd)
        return -1;
    return T.order ? T.order[i] : i;
}

// Inputs of lane (r, j) of node p.
struct NodeFetch
{
    int cell, la, slot;
    int q;  // dof node of (r, j)
    double v[4];  // u0 u1 u2 e of node j of cell r
    double g[gdm8::kGeo];  // geometry of cell r at IP j
    double sr[gdm8::kSr];  // pre-pass state of cell r at IP j
    double th[3];
    double kh;  // history kappa of cell r at IP j
};

// GEO0: the cell's geometry record (IP 0's) in every lane of the group (affine cells).
// CLAMP: lanes past the node's last incident cell (f.cell < 0) load cell 0's records instead of
// skipping the loads, so that every lane runs the same straight-line code (their results are
// never used: the slot map of the reduction leaves absent cells out).
template <bool PRE, bool GEO0 = false, bool CLAMP = false>
__device__ inline void FetchNode(const RowArgs& A, const GdmHex8Tables& T, int pos, int r, int j, NodeFetch& f)
{
    const size_t pr = (size_t)pos * 8 + r;
    f.cell = T.nodeCell[pr];
    f.la = T.nodeLa[pr];
    f.slot = T.nodeSlot[pr * 8 + j];
    f.kh = 0.0;
    f.q = 0;
    if (CLAMP || f.cell >= 0)
    {
        const int cl = CLAMP && f.cell < 0 ? 0 : f.cell;
        const double2* gr = reinterpret_cast<const double2*>(
            GEO0 ? T.geoCell + (size_t)cl * gdm8::kGeo : T.geo + ((size_t)cl * 8 + j) * gdm8::kGeo);
#pragma unroll
        for (int k = 0; k < gdm8::kGeo / 2; k++)
        {
            const double2 v = gr[k];
            f.g[2 * k] = v.x;
            f.g[2 * k + 1] = v.y;
        }
        if constexpr (PRE)
        {
            const double2* sp = reinterpret_cast<const double2*>(T.ipState + ((size_t)cl * 8 + j) * gdm8::kSr);
#pragma unroll
            for (int k = 0; k < gdm8::kSr / 2; k++)
            {
                const double2 v = sp[k];
                f.sr[2 * k] = v.x;
                f.sr[2 * k + 1] = v.y;
            }
        }
        else
        {
            const int hist = T.nodeHist ? T.nodeHist[pr] : f.cell;
            const int q = A.dconn[(size_t)f.cell * 8 + j];
            f.q = q;
            f.kh = A.kappa[(size_t)hist + (size_t)j * A.kappaLd];
            f.v[0] = A.u[(size_t)q * 3];
            f.v[1] = A.u[(size_t)q * 3 + 1];
            f.v[2] = A.u[(size_t)q * 3 + 2];
            f.v[3] = A.e[q];
        }
        if (A.thE)
        {
            f.th[0] = A.thE[cl];
            f.th[1] = A.thFt[cl];
            f.th[2] = A.thGf[cl];
        }
    }
}

// v of lane i + dpp-selected lane (CTRL: a DPP control word; out-of-row reads give 0)
template <int CTRL>
__device__ inline double DppMov(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// sum over the 8 lanes of a group, ((t0 + t1) + (t2 + t3)) + ((t4 + t5) + (t6 + t7)), valid in
// the group's lane 0 (DPP: no LDS instruction)
__device__ inline double GroupSum0(double v)
{
    v += DppMov<0xB1>(v);   // quad_perm [1, 0, 3, 2]: lane i ^ 1
    v += DppMov<0x4E>(v);   // quad_perm [2, 3, 0, 1]: lane i ^ 2
    v += DppMov<0x104>(v);  // row_shl:4: lane i + 4
    return v;
}

// geometry and pre-pass state records of (cell, ip)
__device__ inline void LoadIpRecords(const GdmHex8Tables& T, int cell, int ip, double g[gdm8::kGeo],
                                     double sr[gdm8::kSr])
{
    const double2* gr = reinterpret_cast<const double2*>(T.geo + ((size_t)cell * 8 + ip) * gdm8::kGeo);
#pragma unroll
    for (int k = 0; k < gdm8::kGeo / 2; k++)
    {
        const double2 v = gr[k];
        g[2 * k] = v.x;
        g[2 * k + 1] = v.y;
    }
    const double2* sp = reinterpret_cast<const double2*>(T.ipState + ((size_t)cell * 8 + ip) * gdm8::kSr);
#pragma unroll
    for (int k = 0; k < gdm8::kSr / 2; k++)
    {
        const double2 v = sp[k];
        sr[2 * k] = v.x;
        sr[2 * k + 1] = v.y;
    }
}

// Terms of row node a (local index la) at IP ip of a cell from its pre-pass state record: the LDS
// record for phase B (J^-1, s(1-w) g_a or c1 g_a, B_a^T t, -N_a s m, c s g_a, s N_a, c2 g_a) and the
// R terms s B_a^T sigma, s (N_a (eeq - eps_eq) + c g_a . grad eeq).
// WITHINV false: the record omits J^-1 (affine cells keep one J^-1 per cell) and the other terms
// start at 0 instead of 9.
template <bool FULLDMG, int kSt, bool WITHINV = true>
__device__ inline void PreRecord(const double* sr, const double* g, int la, int ip, const LawParams& P, double rec[kSt],
                                 double Rv[4])
{
    using namespace gdm8;
    constexpr int o = WITHINV ? 0 : -sHa;  // index shift of the terms after J^-1
    const double inv[3][3] = {{g[0], g[1], g[2]}, {g[3], g[4], g[5]}, {g[6], g[7], g[8]}};
    const double s = g[9];
    double da[3], Na, ga[3];
    ShapeAt(la, ip, da, Na);
    GradOf(da, inv, ga);
    // e'' g_a: B_a^T of the Voigt vector e''
    const double Eg[3] = {sr[rEps] * ga[0] + sr[rEps + 5] * ga[1] + sr[rEps + 4] * ga[2],
                          sr[rEps + 5] * ga[0] + sr[rEps + 1] * ga[1] + sr[rEps + 3] * ga[2],
                          sr[rEps + 4] * ga[0] + sr[rEps + 3] * ga[1] + sr[rEps + 2] * ga[2]};
#pragma unroll
    for (int i = 0; i < 3; i++)
        Rv[i] = sr[rSig] * ga[i] + sr[rSig + 1] * Eg[i];
    Rv[3] = Na * sr[rRe] + (ga[0] * sr[rGe] + ga[1] * sr[rGe + 1] + ga[2] * sr[rGe + 2]);
    if constexpr (WITHINV)
    {
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++)
                rec[sInv + 3 * a + b] = inv[a][b];
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
    {
        rec[o + sHa + i] = sr[rC1] * ga[i];
        rec[o + sVa + i] = sr[rT] * ga[i] + sr[rT + 1] * Eg[i];
        rec[o + sCg + i] = P.c * s * ga[i];
        if (!FULLDMG)
            rec[o + sKa + i] = sr[rC2] * ga[i];
    }
#pragma unroll
    for (int i = 0; i < 6; i++)
        rec[o + sMn + i] = -Na * ((i < 3 ? sr[rM] : 0.0) + sr[rM + 1] * sr[rEps + i]);
    rec[o + sSn] = s * Na;
#pragma unroll
    for (int k = o + (FULLDMG ? sSn + 1 : sKa + 3); k < kSt; k++)
        rec[k] = 0.0;  // padding
}

template <int kSt>
__device__ inline void StoreRecord(double* dst, const double rec[kSt])
{
#pragma unroll
    for (int k = 0; k < kSt; k += 2)
        *reinterpret_cast<double2*>(dst + k) = make_double2(rec[k], rec[k + 1]);
}

// Adds one IP's contribution to block (a, b): h = the record's terms after J^-1 (c1 g_a, B_a^T t,
// -N_a m, c g_a, N_a, [c2 g_a]), gb = g_b, Nb = N_b.
template <bool FULLDMG>
__device__ inline void AccumulateIp(const double* h, const double gb[3], double Nb, double acc[16])
{
    using namespace gdm8;
    const double* ha = h;
#pragma unroll
    for (int ii = 0; ii < 3; ii++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            acc[ii * 4 + c] += ha[ii] * gb[c];
    if (!FULLDMG)
    {
        const double* ka = h + (sKa - sHa);
        const double d = ka[0] * gb[0] + ka[1] * gb[1] + ka[2] * gb[2];
#pragma unroll
        for (int ii = 0; ii < 3; ii++)
        {
#pragma unroll
            for (int c = 0; c < 3; c++)
                acc[ii * 4 + c] += gb[ii] * ka[c];
            acc[ii * 5] += d;
        }
    }
    const double* va = h + (sVa - sHa);
    acc[3] += va[0] * Nb;
    acc[7] += va[1] * Nb;
    acc[11] += va[2] * Nb;
    // B_b^T (-N_a m) and the eeq-eeq entry, each term fused into the accumulator
    const double* mn = h + (sMn - sHa);
    acc[12] = fma(gb[1], mn[5], fma(gb[2], mn[4], fma(gb[0], mn[0], acc[12])));
    acc[13] = fma(gb[0], mn[5], fma(gb[2], mn[3], fma(gb[1], mn[1], acc[13])));
    acc[14] = fma(gb[0], mn[4], fma(gb[1], mn[3], fma(gb[2], mn[2], acc[14])));
    const double* cg = h + (sCg - sHa);
    acc[15] = fma(cg[2], gb[2], fma(cg[1], gb[1], fma(cg[0], gb[0], fma(h[sSn - sHa], Nb, acc[15]))));
}

// Phase B: lane (r, b) accumulates block (a_r, b) over IPs ip0 .. ip0 + n - 1 (records at
// state[(r * n + i) * kSt], i = ip - ip0), in IP order. The records of one cell are adjacent, so
// the record stores of consecutive lanes (r, i) hit consecutive records: with an odd number of
// 16-byte chunks per record (kSt / 2 odd) any 16 of them fall on distinct 4-bank groups.
template <bool FULLDMG, int kSt>
__device__ inline void PhaseB(const double* state, int r, int b, int ip0, int n, double acc[16])
{
    using namespace gdm8;
#pragma unroll 1
    for (int i = 0; i < n; i++)
    {
        const int ip = ip0 + i;
        double db[3], Nb;
        ShapeAt(b, ip, db, Nb);
        const double* st = state + (r * n + i) * kSt;
        constexpr int kRd = FULLDMG ? 26 : 28;
        double rec[kRd];
#pragma unroll
        for (int k = 0; k < kRd; k += 2)
        {
            const double2 v = *reinterpret_cast<const double2*>(st + k);
            rec[k] = v.x;
            rec[k + 1] = v.y;
        }
        const double inv[3][3] = {{rec[0], rec[1], rec[2]}, {rec[3], rec[4], rec[5]}, {rec[6], rec[7], rec[8]}};
        double gb[3];
        GradOf(db, inv, gb);
        AccumulateIp<FULLDMG>(rec + sHa, gb, Nb, acc);
    }
}

// Phase B of affine cells (MODE 2): lane (r, b) visits the IPs of cell r in the order
// ip = t ^ m(b), m(b) = b's corner (bit k set: + side along axis k). Node b's shape data at that
// IP depends on t only (kShapeRel, uniform over the wave); the signs of b's gradient are folded
// into J^-1 once, so g_b is bitwise GradOf's. The 8 lanes of a cell read 8 different records:
// 16 consecutive lanes read 16 consecutive records (distinct 4-bank groups).
template <bool FULLDMG, int kSt>
__device__ inline void PhaseBAffine(const double* state, int r, int b, const double inv[3][3], double acc[16])
{
    using namespace gdm8;
    const int sx = ((b + 1) >> 1) & 1, sy = (b >> 1) & 1, sz = b >> 2;
    const int m = sx | (sy << 1) | (sz << 2);
    double sinv[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
        sinv[0][c] = sx ? inv[0][c] : -inv[0][c];
        sinv[1][c] = sy ? inv[1][c] : -inv[1][c];
        sinv[2][c] = sz ? inv[2][c] : -inv[2][c];
    }
    const double* cellRec = state + r * 8 * kSt;
#pragma unroll 2
    for (int t = 0; t < 8; t++)
    {
        const double d0 = kShapeRel.v[t][0], d1 = kShapeRel.v[t][1], d2 = kShapeRel.v[t][2], Nb = kShapeRel.v[t][3];
        const double* st = cellRec + (t ^ m) * kSt;
        constexpr int kRd = FULLDMG ? 16 : 20;
        double rec[kRd];
#pragma unroll
        for (int k = 0; k < kRd; k += 2)
        {
            const double2 v = *reinterpret_cast<const double2*>(st + k);
            rec[k] = v.x;
            rec[k + 1] = v.y;
        }
        double gb[3];
#pragma unroll
        for (int c = 0; c < 3; c++)
            gb[c] = d0 * sinv[0][c] + d1 * sinv[1][c] + d2 * sinv[2][c];
        AccumulateIp<FULLDMG>(rec, gb, Nb, acc);
    }
}

// One workgroup per node. Workgroups are dispatched round-robin over the 8 XCDs; with
// xcdMap the node of block b is chosen so that each XCD processes a contiguous range of nodes
// (neighbouring nodes share cells and node data, which then stay in that XCD's L2).
// MODE 0: the node kernel evaluates the IP states itself (8 records per cell in LDS).
// MODE 1: the IP states come from kGdmHex8Ip; phase B runs in two passes of 4 IPs so that only 32
//         records are in LDS at once (9.3 KB per workgroup: 4 waves per SIMD at <= 128 VGPRs).
// MODE 2: as 1 for meshes of affine cells (J^-1 and detJ w equal at the 8 IPs of every cell, e.g.
//         bricks): IP 0's geometry for all IPs; records without J^-1 (18 doubles); all 8 IPs in one
//         pass (9.3 KB).
// K is written once and not read again here: streaming (non-temporal) stores keep the records,
// which are re-read, in the caches.
typedef double D2v __attribute__((ext_vector_type(2)));

// Stores n consecutive doubles from LDS to dst, the 64 lanes writing consecutive 16-byte pairs (a
// leading double when dst is not 16-byte aligned, a trailing one when the rest is odd).
__device__ inline void StoreRows(double* dst, const double* src, int n, int lane)
{
    const int head = (reinterpret_cast<unsigned long long>(dst) & 15) ? 1 : 0;
    if (head && lane == 0)
        __builtin_nontemporal_store(src[0], dst);
    const int np = (n - head) >> 1;
    for (int k = lane; k < np; k += 64)
        __builtin_nontemporal_store(D2v{src[head + 2 * k], src[head + 2 * k + 1]},
                                    reinterpret_cast<D2v*>(dst + head + 2 * k));
    if (((n - head) & 1) && lane == 0)
        __builtin_nontemporal_store(src[n - 1], dst + n - 1);
}

template <bool FULLDMG, int MODE>
struct NodeLds
{
    static constexpr int kSt = MODE == 2 ? (FULLDMG ? 18 : 22) : gdm8::kStOf<FULLDMG>;
    static constexpr int kRecIps = MODE == 1 ? 4 : 8;                         // IPs of the records in LDS at a time
    static constexpr int kRecs = 8 * kRecIps * kSt;                           // records
    static constexpr int kState = kRecs > gdm8::kRed ? kRecs : gdm8::kRed;   // doubles of LDS per workgroup
};

// The rows of node p (one 64-lane workgroup); `state` is the workgroup's LDS (NodeLds::kState doubles):
// records [i][r][kSt]; after phase B: partial blocks, R sums and the slot map.
template <bool WR, bool WK, bool FULLDMG, int MODE>
__device__ inline void NodeBody(const RowArgs& A, const GdmHex8Tables& T, int pos, double* state)
{
    using namespace gdm8;
    constexpr bool PRE = MODE >= 1, AFF = MODE == 2;
    constexpr int kSt = NodeLds<FULLDMG, MODE>::kSt;

    const int lane = threadIdx.x;
    const int r = lane >> 3, j = lane & 7;
    const LawParams& P = A.P;
    // node p at table position pos; its rows: 3 displacement rows (consecutive in the CSR) and the
    // eeq row, each 4 nA long
    const int4 head = T.nodeHead[pos];
    const int p = head.x, ru0 = head.y, L = head.z, re0 = head.w;
    const int nA = L >> 2;
    NodeFetch f;
    FetchNode<PRE, AFF, AFF>(A, T, pos, r, j, f);
    const bool valid = f.cell >= 0;  // r < number of cells incident to p
    const int la = f.la;

    // material of the cell (per-cell fields: E, kappa0 = f_t / E, beta = f_t / G_f)
    double E = P.E, k0 = P.kappa0, beta = P.beta;
    if (A.thE && valid)
    {
        E = f.th[0];
        k0 = f.th[1] / E;
        beta = f.th[1] / f.th[2];
    }
    const double m3K = P.cK * E, m2G = P.cG * E;

    double Rv[4] = {0.0, 0.0, 0.0, 0.0};
    double acc[16];
#pragma unroll
    for (int i = 0; i < 16; i++)
        acc[i] = 0.0;
    if constexpr (AFF)
    {
        // every lane, absent cells included (their records and blocks are never read; their R terms
        // are zeroed)
        double rec[kSt];
        PreRecord<FULLDMG, kSt, false>(f.sr, f.g, la, j, P, rec, Rv);
        if (!valid)
            Rv[0] = Rv[1] = Rv[2] = Rv[3] = 0.0;
        if (WK)
        {
            StoreRecord<kSt>(state + (r * 8 + j) * kSt, rec);
            __syncthreads();
            // lane (r, b) works on cell r, whose J^-1 the group holds
            const double inv[3][3] = {{f.g[0], f.g[1], f.g[2]}, {f.g[3], f.g[4], f.g[5]}, {f.g[6], f.g[7], f.g[8]}};
            PhaseBAffine<FULLDMG, kSt>(state, r, j, inv, acc);
        }
    }
    else if constexpr (PRE)
    {
        double rec[kSt];
        if (valid)
            PreRecord<FULLDMG, kSt>(f.sr, f.g, la, j, P, rec, Rv);
        if (WK)
        {
            // pass 0: IPs 0..3 (records of lanes j < 4)
            if (valid && j < 4)
                StoreRecord<kSt>(state + (r * 4 + j) * kSt, rec);
            __syncthreads();
            if (valid)
                PhaseB<FULLDMG, kSt>(state, r, j, 0, 4, acc);
            __syncthreads();
            // pass 1: IPs 4..7; lanes j >= 4 rebuild their records from the (cached) state records
            if (valid && j >= 4)
            {
                double g2[kGeo], sr2[kSr], Rd[4];
                LoadIpRecords(T, f.cell, j, g2, sr2);
                PreRecord<FULLDMG, kSt>(sr2, g2, la, j, P, rec, Rd);
                StoreRecord<kSt>(state + (r * 4 + j - 4) * kSt, rec);
            }
            __syncthreads();
            if (valid)
                PhaseB<FULLDMG, kSt>(state, r, j, 4, 4, acc);
        }
    }
    else
    {
        if (valid)
        {
            const int ip = j;
            double Hu[3][3], he[3], eN;
            GatherGradients(f.v, ip, Hu, he, eN);
            const double inv[3][3] = {{f.g[0], f.g[1], f.g[2]}, {f.g[3], f.g[4], f.g[5]}, {f.g[6], f.g[7], f.g[8]}};
            const double s = f.g[9];
            GdmIp st;
            GdmIpState<FULLDMG>(Hu, he, eN, inv, f.kh, m3K, m2G, k0, beta, P, st);
            const double* sig = st.sig;
            const double* dsw = st.dsw;
            const double* ge = st.ge;
            const double om = st.om, dom = st.dom, dk = st.dk, H = st.H, d2G = st.d2G, eq = st.eq;

            double da[3], Na, ga[3];
            ShapeAt(la, ip, da, Na);
            GradOf(da, inv, ga);
            if (WR)
            {
                double bt[3];
                BTv3(ga, sig, bt);
                Rv[0] = s * bt[0];
                Rv[1] = s * bt[1];
                Rv[2] = s * bt[2];
                Rv[3] = s * (Na * (eN - eq) + P.c * (ga[0] * ge[0] + ga[1] * ge[1] + ga[2] * ge[2]));
            }
            if (WK)
            {
                double m[6];
                GdmMisesDerivative(st, P, m);
                double* rp = state + (r * 8 + ip) * kSt;
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
                    if (!FULLDMG)
                        rec[sKa + i] = 0.5 * s * d2G * ga[i];
                }
                rec[sSn] = s * Na;
                if (FULLDMG)
                    rec[25] = 0.0;
                else
                    rec[28] = rec[29] = 0.0;
#pragma unroll
                for (int k = 0; k < kSt; k += 2)
                    *reinterpret_cast<double2*>(rp + k) = make_double2(rec[k], rec[k + 1]);
            }
        }
    }
    if constexpr (!PRE)
    {
        __syncthreads();
        if (WK && valid)
            PhaseB<FULLDMG, kSt>(state, r, j, 0, 8, acc);
    }
    if (WK && (valid || AFF) && FULLDMG)
    {
        // uu = lambda' S + mu' S^T + mu' tr(S) I
        const double lam = (m3K - m2G) * kThird, mu = 0.5 * m2G;
        const double S[3][3] = {{acc[0], acc[1], acc[2]}, {acc[4], acc[5], acc[6]}, {acc[8], acc[9], acc[10]}};
        const double tr = S[0][0] + S[1][1] + S[2][2];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int c = 0; c < 3; c++)
            {
                if (i == c)
                    acc[i * 4 + c] = lam * S[i][c] + mu * S[c][i] + mu * tr;
                else
                    acc[i * 4 + c] = lam * S[i][c] + mu * S[c][i];
            }
    }
    // R: per cell over its IPs (butterfly within the cell's 8 lanes), then over the cells in order
    if (WR)
    {
#pragma unroll
        for (int k = 0; k < 4; k++)
            Rv[k] = GroupSum0(Rv[k]);  // stored by lane j == 0
    }
    __syncthreads();

    // ---- partial blocks, R sums and the slot map to LDS (the records are dead now) ------------------
    double* part = state;               // [r][b][kPs]
    double* rcell = state + 8 * 8 * kPs;  // [4][r]: R terms summed over the IPs of cell r
    // slotOf[q][r] = local index b of neighbour q in cell r, 0xFF if q is not a node of cell r
    unsigned char(*slotOf)[kMaxInc] = reinterpret_cast<unsigned char(*)[kMaxInc]>(rcell + 40);
    if (WK)
    {
        reinterpret_cast<unsigned int*>(slotOf)[lane] = 0xFFFFFFFFu;  // 64 x 4 bytes = the whole map
#pragma unroll
        for (int i = 0; i < 16; i++)
            part[lane * kPs + i] = acc[i];
    }
    if (WR && j == 0)
    {
#pragma unroll
        for (int k = 0; k < 4; k++)
            rcell[k * 8 + r] = Rv[k];
    }
    __syncthreads();
    if (WK && valid)
        slotOf[f.slot][r] = (unsigned char)j;
    __syncthreads();

    // ---- reduce: lane (q, half) sums its 8 block entries over the cells in order and writes them ----
    // (only the cells that hold q are read: the lanes of absent (cell, q) pairs sit out that step, so
    // the LDS reads are the 64 partial blocks, not 8 per neighbour). Row layout: the three
    // displacement rows of p are consecutive in the CSR (4 nA columns each), then the eeq row.
    if (WK)
    {
        const int q = lane >> 1, h = lane & 1;
        double sum[8];
#pragma unroll
        for (int k = 0; k < 8; k++)
            sum[k] = 0.0;
        if (q < nA)
        {
            const unsigned long long row = *reinterpret_cast<const unsigned long long*>(&slotOf[q][0]);
#pragma unroll
            for (int rr = 0; rr < 8; rr++)
            {
                const unsigned b = (unsigned)(row >> (8 * rr)) & 0xFFu;
                if (b != 0xFFu)
                {
                    const double* src = part + (rr * 8 + (int)b) * kPs + h * 8;
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        sum[k] += src[k];
                }
            }
        }
        __syncthreads();  // the rows below overwrite the partial blocks
        // p's rows in CSR order in LDS: [3 L] displacement rows, then [L] the eeq row. Block entry
        // (i, c): i = row component (0..2 displacement rows, 3 eeq row), c = column component
        // (0..2 displacement columns, 3 eeq column)
        double* rows = state;
        if (q < nA)
        {
#pragma unroll
            for (int k = 0; k < 8; k++)
            {
                const int e = h * 8 + k;
                const int i = e >> 2, c = e & 3;
                rows[i * L + (c < 3 ? 3 * q + c : 3 * nA + q)] = sum[k];
            }
        }
        __syncthreads();
        StoreRows(A.K + ru0, rows, 3 * L, lane);
        StoreRows(A.K + re0, rows + 3 * L, L, lane);
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

// MODE >= 1 targets 4 waves per SIMD (<= 128 VGPRs; the LDS of modes 1 and 2 allows 16 workgroups per CU)
template <bool WR, bool WK, bool FULLDMG, int MODE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MODE >= 1 ? 4 : 1)))
kGdmHex8Node(const RowArgs A, const GdmHex8Tables T)
{
    __shared__ __align__(16) double state[NodeLds<FULLDMG, MODE>::kState];
    const int pos = NodeOfBlock(T);
    if (pos < 0)
        return;
    NodeBody<WR, WK, FULLDMG, MODE>(A, T, pos, state);
}

// Pre-pass: the state of every (cell, IP) once, for the node kernel (which would otherwise evaluate
// it in each of the cell's 8 node workgroups). Lane (c, ip) of 8-lane group c; node j of the cell is
// gathered by lane (c, j) and shared by swizzle, as in the node kernel.
// AFF: affine cells, IP 0's geometry (T.geoCell) at every IP.
// Workgroup w handles cells 8w .. 8w + 7 (lane (r, j): cell 8w + r, IP j). Their geometry and their
// records are contiguous blocks in memory; both pass through LDS so that every global load and
// store of them is 16 contiguous bytes per lane across the wave (the address and data paths, not
// the arithmetic, bound this kernel).
template <bool FULLDMG, bool AFF>
__global__ void __launch_bounds__(64) kGdmHex8Ip(const RowArgs A, const GdmHex8Tables T, double* __restrict__ out)
{
    using namespace gdm8;
    __shared__ __align__(16) double stage[64 * kSr];  // the block's geometry, then its records
    const int lane = threadIdx.x;
    const int j = lane & 7, r = lane >> 3;
    const long long cell0 = (long long)blockIdx.x * 8;
    const int nc = (int)(A.nCells - cell0 < 8 ? A.nCells - cell0 : 8);  // cells of this block, >= 1
    const bool valid = r < nc;  // uniform within each 8-lane group
    const int cell = valid ? (int)(cell0 + r) : 0;
    const LawParams& P = A.P;
    {
        constexpr int kPerCell = AFF ? kGeo / 2 : 8 * kGeo / 2;  // 16-byte chunks of geometry per cell
        const double2* src =
            reinterpret_cast<const double2*>(AFF ? T.geoCell + cell0 * kGeo : T.geo + cell0 * 8 * kGeo);
        double2* dst = reinterpret_cast<double2*>(stage);
#pragma unroll
        for (int i = 0; i < (8 * kPerCell + 63) / 64; i++)
        {
            const int c = lane + 64 * i;
            if (c < nc * kPerCell)
                dst[c] = src[c];
        }
    }
    double v[4] = {0.0, 0.0, 0.0, 0.0}, g[kGeo];
    double kh = 0.0, E = P.E, k0 = P.kappa0, beta = P.beta;
    if (valid)
    {
        const int q = A.dconn[(size_t)cell * 8 + j];
        v[0] = A.u[(size_t)q * 3];
        v[1] = A.u[(size_t)q * 3 + 1];
        v[2] = A.u[(size_t)q * 3 + 2];
        v[3] = A.e[q];
        const int hist = A.cellId ? A.cellId[cell] : cell;
        kh = A.kappa[(size_t)hist + (size_t)j * A.kappaLd];
        if (A.thE)
        {
            E = A.thE[cell];
            k0 = A.thFt[cell] / E;
            beta = A.thFt[cell] / A.thGf[cell];
        }
    }
    __syncthreads();
    {
        const double2* gs = reinterpret_cast<const double2*>(stage + (AFF ? r : r * 8 + j) * kGeo);
#pragma unroll
        for (int k = 0; k < kGeo / 2; k++)
        {
            const double2 w = gs[k];
            g[2 * k] = w.x;
            g[2 * k + 1] = w.y;
        }
    }
    double Hu[3][3], he[3], eN;
    GatherGradients(v, j, Hu, he, eN);
    __syncthreads();  // the records overwrite the geometry
    if (valid)
    {
        const double inv[3][3] = {{g[0], g[1], g[2]}, {g[3], g[4], g[5]}, {g[6], g[7], g[8]}};
        const double s = g[9];
        const double m3K = P.cK * E, m2G = P.cG * E;
        GdmIp st;
        GdmIpState<FULLDMG>(Hu, he, eN, inv, kh, m3K, m2G, k0, beta, P, st);
        const double eV = st.I1 * kThird;
        const double sV = (1.0 - st.om * st.H) * m3K * eV;
        double am = P.K1, bm = 0.0;  // m = am [1,1,1,0,0,0] + bm e'' (A == 0: m = K1 [1,1,1,0,0,0])
        if (st.Am != 0.0)
        {
            const double h = 0.5 / st.Am;
            am = P.K1 + h * (2.0 * P.K1 * P.K1 * st.I1) - h * P.K2 * st.I1 * kThird;
            bm = h * P.K2;
        }
        const double ct = s * st.dom * st.dk;
        double rec[kSr];
        rec[rEps] = st.eps[0];
        rec[rEps + 1] = st.eps[1];
        rec[rEps + 2] = st.eps[2];
        rec[rEps + 3] = 0.5 * st.eps[3];
        rec[rEps + 4] = 0.5 * st.eps[4];
        rec[rEps + 5] = 0.5 * st.eps[5];
        rec[rSig] = s * (sV - st.d2G * eV);
        rec[rSig + 1] = s * st.d2G;
        rec[rT] = -ct * (m3K * st.H - m2G) * eV;
        rec[rT + 1] = -ct * m2G;
        rec[rM] = s * am;
        rec[rM + 1] = s * bm;
        rec[rC1] = FULLDMG ? s * (1.0 - st.om) : s * (m3K * (1.0 - st.om * st.H) - st.d2G) * kThird;
        rec[rC2] = FULLDMG ? 0.0 : 0.5 * s * st.d2G;
        rec[rRe] = s * (st.eN - st.eq);
        rec[rGe] = s * P.c * st.ge[0];
        rec[rGe + 1] = s * P.c * st.ge[1];
        rec[rGe + 2] = s * P.c * st.ge[2];
        // record (r, j) at row r * 8 + j: 9 chunks (odd), consecutive lanes on distinct bank groups
        StoreRecord<kSr>(stage + lane * kSr, rec);
    }
    __syncthreads();
    {
        constexpr int kPerCell = 8 * kSr / 2;  // 16-byte chunks of records per cell
        const double2* src = reinterpret_cast<const double2*>(stage);
        double2* dst = reinterpret_cast<double2*>(out + cell0 * 8 * kSr);
#pragma unroll
        for (int i = 0; i < (8 * kPerCell + 63) / 64; i++)
        {
            const int c = lane + 64 * i;
            if (c < nc * kPerCell)
                dst[c] = src[c];
        }
    }
}

// Geometry of every (cell, IP): J = sum_n X_n dN_n/dxi^T over the coordinate nodes in local
// order (Jacobian.h:16-48), J^-1 and detJ (Jacobian.h:35-47), detJ * w (Cell.h:108-117).
__global__ void __launch_bounds__(256) kGdmHex8Geometry(const double* x, const double* y, const double* z,
                                                         const int* cconn, const double* tab, int nCells,
                                                         double* geo)
{
    using namespace gdm8;
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;  // (cell, ip)
    if (t >= (long long)nCells * 8)
        return;
    const int c = (int)(t >> 3), ip = (int)(t & 7);
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int n = 0; n < 8; n++)
    {
        double d[3], Nv;
        ShapeAt(n, ip, d, Nv);
        const int cn = cconn[(size_t)c * 8 + n];
        const double X[3] = {x[cn], y[cn], z[cn]};
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int k = 0; k < 3; k++)
                J[i][k] += X[i] * d[k];
    }
    double inv[3][3];
    const double detJ = Inverse<3>(J, inv);
    double* g = geo + (size_t)t * kGeo;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
            g[3 * a + b] = inv[a][b];
    g[9] = detJ * tab[ip * kTab];
}

// Counts the cells that are not parallelepipeds (0 means every cell is affine and the node kernel
// may keep one J^-1 per cell). For a parallelepiped J is the same at every point of the cell; the
// J^-1 computed at the 8 IPs then differ only by rounding, so a cell counts as affine when they
// agree to 1e-13 of its largest entry and detJ w to 1e-13 of IP 0's (equal weights, as in 2x2x2
// Gauss). IP 0's geometry is then used for all IPs: a relative change of order 1e-16 in the results.
__global__ void __launch_bounds__(256) kGdmHex8CountNonAffine(const double* geo, int nCells, int* count)
{
    using namespace gdm8;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= nCells)
        return;
    const double* g0 = geo + (size_t)c * 8 * kGeo;
    double scale = 0.0;
    for (int k = 0; k < 9; k++)
        scale = fmax(scale, fabs(g0[k]));
    bool same = true;
    for (int ip = 1; ip < 8; ip++)
    {
        for (int k = 0; k < 9; k++)
            same = same && (fabs(g0[ip * kGeo + k] - g0[k]) <= 1e-13 * scale);
        same = same && (fabs(g0[ip * kGeo + 9] - g0[9]) <= 1e-13 * fabs(g0[9]));
    }
    if (!same)
        atomicAdd(count, 1);
}

// Nodes sorted along a Morton (Z-order) curve of their coordinates (MortonSort): each node takes
// the coordinates of its first incident cell's coordinate node (hex8 dof nodes are the coordinate
// nodes); nodes without cells go first.
inline hipError_t MortonOrder(const NodeGraph& g, int nCells, int nCoord, const double* x, const double* y,
                              const double* z, const int* cconn, hipStream_t stream, std::vector<int>& order)
{
    std::vector<double> hx(nCoord), hy(nCoord), hz(nCoord);
    std::vector<int> hc((size_t)nCells * 8);
    hipError_t err;
    if ((err = hipMemcpyAsync(hx.data(), x, nCoord * sizeof(double), hipMemcpyDeviceToHost, stream)) != hipSuccess ||
        (err = hipMemcpyAsync(hy.data(), y, nCoord * sizeof(double), hipMemcpyDeviceToHost, stream)) != hipSuccess ||
        (err = hipMemcpyAsync(hz.data(), z, nCoord * sizeof(double), hipMemcpyDeviceToHost, stream)) != hipSuccess ||
        (err = hipMemcpyAsync(hc.data(), cconn, hc.size() * sizeof(int), hipMemcpyDeviceToHost, stream)) !=
            hipSuccess ||
        (err = hipStreamSynchronize(stream)) != hipSuccess)
        return err;
    const int n = g.nNodes;
    std::vector<double> pts((size_t)n * 3, 0.0);
    std::vector<char> valid(n, 0);
    for (int p = 0; p < n; p++)
    {
        if (g.incPtr[p + 1] == g.incPtr[p])
            continue;
        const int k = g.incPtr[p];
        const int c = hc[(size_t)g.incCell[k] * 8 + g.incLocal[k]];
        pts[(size_t)p * 3] = hx[c];
        pts[(size_t)p * 3 + 1] = hy[c];
        pts[(size_t)p * 3 + 2] = hz[c];
        valid[p] = 1;
    }
    MortonSort(n, 3, pts.data(), order, valid.data());
    return hipSuccess;
}

struct GdmHex8Plan
{
    bool ready = false;
    int nNodes = 0;
    int grid = 0;     // workgroups: one per node (rounded up to a multiple of 8 with the XCD mapping)
    int xcdMap = 1;   // contiguous node ranges per XCD (env NUTO_AMD_GDM_XCD_MAP=0 switches it off)
    DevBuf nodeHead, nodeCell, nodeHist, nodeLa, nodeSlot, geo;
    DevBuf posOf;        // node -> table position (Morton order; NUTO_AMD_GDM_ORDER=0: node order, empty)
    DevBuf ipState;      // pre-pass state records (NUTO_AMD_GDM_PREPASS=0: recompute in the node kernel)
    DevBuf geoCell;      // affine meshes: IP 0's geometry record per cell
    DevBuf extra;        // nodes beyond the kernel's bounds (more than 8 cells or 32 neighbours)
    int nExtra = 0;      // ... assembled by the generic row-owner kernel
    bool prepass = true;
    bool affine = false;  // every cell affine: MODE 2 of the node kernel
    int nCells = 0;

    void Release()
    {
        ready = false;
        nNodes = 0;
        for (DevBuf* b : {&nodeHead, &nodeCell, &nodeHist, &nodeLa, &nodeSlot, &geo, &ipState, &posOf, &geoCell, &extra})
            b->Release();
        nExtra = 0;
    }

    bool Usable(int kind, int coordType, int dofType, int nip) const
    {
        return ready && kind == kGradientDamage && coordType == kHex8 && dofType == kHex8 && nip == 8;
    }

    // Builds the per-node tables from the node graph and the geometry records from the device
    // coordinates (x, y, z by coordinate node, cconn) and the IP table. Returns false (plan not
    // used) when the mesh exceeds the kernel's bounds; a HIP failure is returned through `err`.
    bool Prepare(int kind, int coordType, int dofType, int nip, const NodeGraph& g, int nCells, const int* cellIds,
                 int nCoord, const double* x, const double* y, const double* z, const int* cconn, const double* tab,
                 const int* rowPtr, int nRows, hipStream_t stream, hipError_t& err)
    {
        Release();
        err = hipSuccess;
        if (kind != kGradientDamage || coordType != kHex8 || dofType != kHex8 || nip != 8)
            return false;
        const int n = g.nNodes;
        if (nRows != 4 * n)  // 3 displacement rows and 1 eeq row per node
            return false;
        // nodes beyond the kernel's bounds stay out of the tables (the generic kernel assembles them)
        std::vector<char> out(n, 0);
        std::vector<int> extraNodes;
        for (int p = 0; p < n; p++)
        {
            if (g.incPtr[p + 1] - g.incPtr[p] > gdm8::kMaxInc || g.adjPtr[p + 1] - g.adjPtr[p] > gdm8::kMaxAdj)
            {
                out[p] = 1;
                extraNodes.push_back(p);
            }
        }
        if (n > 0 && (int)extraNodes.size() == n)
            return false;
        // traversal order: position s holds node ord[s]
        std::vector<int> ord(n);
        bool reordered = false;
        const char* ordEnv = std::getenv("NUTO_AMD_GDM_ORDER");
        if (n > 0 && nCells > 0 && !(ordEnv && ordEnv[0] == '0'))
        {
            if ((err = MortonOrder(g, nCells, nCoord, x, y, z, cconn, stream, ord)) != hipSuccess)
                return false;
            reordered = true;
        }
        else
        {
            for (int s = 0; s < n; s++)
                ord[s] = s;
        }
        if (!extraNodes.empty())
        {
            // drop the outside nodes from the traversal; they keep position -1
            ord.erase(std::remove_if(ord.begin(), ord.end(), [&](int p) { return out[p] != 0; }), ord.end());
            reordered = true;
        }
        const int nIn = (int)ord.size();
        // CSR rows: displacement rows 3p .. 3p + 2, eeq row 3n + p
        std::vector<int> rp((size_t)4 * n + 1);
        if ((err = hipMemcpyAsync(rp.data(), rowPtr, rp.size() * sizeof(int), hipMemcpyDeviceToHost, stream)) !=
                hipSuccess ||
            (err = hipStreamSynchronize(stream)) != hipSuccess)
            return false;
        std::vector<int4> head(nIn);
        std::vector<int> cell((size_t)nIn * 8, -1), hist(cellIds ? (size_t)nIn * 8 : 0, 0), pos(reordered ? n : 0, -1);
        std::vector<unsigned char> la((size_t)nIn * 8, 0), slot((size_t)nIn * 64, 0);
        for (int s = 0; s < nIn; s++)
        {
            const int p = ord[s];
            head[s] = make_int4(p, rp[(size_t)3 * p], rp[(size_t)3 * p + 1] - rp[(size_t)3 * p], rp[(size_t)3 * n + p]);
            if (reordered)
                pos[p] = s;
            for (int k = g.incPtr[p], r = 0; k < g.incPtr[p + 1]; k++, r++)
            {
                const size_t pr = (size_t)s * 8 + r;
                cell[pr] = g.incCell[k];
                la[pr] = g.incLocal[k];
                if (cellIds)
                    hist[pr] = cellIds[g.incCell[k]];
                for (int j = 0; j < 8; j++)
                    slot[pr * 8 + j] = (unsigned char)(g.slot[(size_t)k * 8 + j] & 0x7fff);
            }
        }
        if ((err = nodeHead.Upload(head.data(), head.size() * sizeof(int4), stream)) != hipSuccess ||
            (reordered && (err = posOf.Upload(pos.data(), pos.size() * sizeof(int), stream)) != hipSuccess) ||
            (err = nodeCell.Upload(cell.data(), cell.size() * sizeof(int), stream)) != hipSuccess ||
            (err = nodeLa.Upload(la.data(), la.size(), stream)) != hipSuccess ||
            (!extraNodes.empty() &&
             (err = extra.Upload(extraNodes.data(), extraNodes.size() * sizeof(int), stream)) != hipSuccess) ||
            (err = nodeSlot.Upload(slot.data(), slot.size(), stream)) != hipSuccess ||
            (cellIds && (err = nodeHist.Upload(hist.data(), hist.size() * sizeof(int), stream)) != hipSuccess) ||
            (err = geo.Alloc((size_t)nCells * 8 * gdm8::kGeo * sizeof(double))) != hipSuccess)
        {
            Release();
            return false;
        }
        if (nCells > 0)
        {
            const long long items = (long long)nCells * 8;
            hipLaunchKernelGGL(kGdmHex8Geometry, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, stream, x, y,
                               z, cconn, tab, nCells, geo.As<double>());
            if ((err = hipGetLastError()) != hipSuccess || (err = hipStreamSynchronize(stream)) != hipSuccess)
            {
                Release();
                return false;
            }
        }
        const char* pre = std::getenv("NUTO_AMD_GDM_PREPASS");
        prepass = !(pre && pre[0] == '0');
        const char* aff = std::getenv("NUTO_AMD_GDM_AFFINE");
        affine = false;
        if (prepass && nCells > 0 && !(aff && aff[0] == '0'))
        {
            DevBuf cnt;
            int h = -1;
            if ((err = cnt.Alloc(sizeof(int))) != hipSuccess ||
                (err = hipMemsetAsync(cnt.p, 0, sizeof(int), stream)) != hipSuccess)
            {
                Release();
                return false;
            }
            hipLaunchKernelGGL(kGdmHex8CountNonAffine, dim3((unsigned)((nCells + 255) / 256)), dim3(256), 0, stream,
                               geo.As<double>(), nCells, cnt.As<int>());
            if ((err = hipGetLastError()) != hipSuccess ||
                (err = hipMemcpyAsync(&h, cnt.p, sizeof(int), hipMemcpyDeviceToHost, stream)) != hipSuccess ||
                (err = hipStreamSynchronize(stream)) != hipSuccess)
            {
                Release();
                return false;
            }
            affine = h == 0;
            if (affine && ((err = geoCell.Alloc((size_t)nCells * gdm8::kGeo * sizeof(double))) != hipSuccess ||
                           (err = hipMemcpy2DAsync(geoCell.p, gdm8::kGeo * sizeof(double), geo.p,
                                                   8 * gdm8::kGeo * sizeof(double), gdm8::kGeo * sizeof(double),
                                                   nCells, hipMemcpyDeviceToDevice, stream)) != hipSuccess ||
                           (err = hipStreamSynchronize(stream)) != hipSuccess))
            {
                Release();
                return false;
            }
        }
        if (prepass && (err = ipState.Alloc((size_t)nCells * 8 * gdm8::kSr * sizeof(double))) != hipSuccess)
        {
            Release();
            return false;
        }
        this->nCells = nCells;
        nNodes = nIn;
        nExtra = (int)extraNodes.size();
        const char* env = std::getenv("NUTO_AMD_GDM_XCD_MAP");
        xcdMap = env ? std::atoi(env) : 1;
        grid = xcdMap ? (nNodes + 7) / 8 * 8 : nNodes;
        ready = true;
        return true;
    }

    GdmHex8Tables Tables() const
    {
        return GdmHex8Tables{nodeHead.As<int4>(), nodeCell.As<int>(), nodeHist.p ? nodeHist.As<int>() : nullptr,
                             nodeLa.As<unsigned char>(), nodeSlot.As<unsigned char>(), geo.As<double>(),
                             ipState.p ? ipState.As<double>() : nullptr, geoCell.p ? geoCell.As<double>() : nullptr,
                             xcdMap, 0, nNodes, nullptr};
    }

    template <bool FULLDMG, int MODE>
    void LaunchNodes(const RowArgs& a, const GdmHex8Tables& t, bool wr, bool wk, int grid_, hipStream_t stream)
    {
        if (wr && wk)
            hipLaunchKernelGGL((kGdmHex8Node<true, true, FULLDMG, MODE>), dim3(grid_), dim3(64), 0, stream, a, t);
        else if (wr)
            hipLaunchKernelGGL((kGdmHex8Node<true, false, FULLDMG, MODE>), dim3(grid_), dim3(64), 0, stream, a, t);
        else
            hipLaunchKernelGGL((kGdmHex8Node<false, true, FULLDMG, MODE>), dim3(grid_), dim3(64), 0, stream, a, t);
    }

    // doubles of the pre-pass state records (written in full by kGdmHex8Ip before the node kernel reads them)
    size_t StateDoubles() const
    {
        return prepass ? (size_t)nCells * 8 * gdm8::kSr : 0;
    }

    template <bool FULLDMG>
    void LaunchT(const RowArgs& a, bool wr, bool wk, double* state, hipStream_t stream)
    {
        GdmHex8Tables t = Tables();
        if (prepass)
        {
            t.ipState = state;
            const dim3 ipGrid((unsigned)(((long long)nCells * 8 + 63) / 64));
            if (affine)
            {
                if (nCells > 0)
                    hipLaunchKernelGGL((kGdmHex8Ip<FULLDMG, true>), ipGrid, dim3(64), 0, stream, a, t, state);
                LaunchNodes<FULLDMG, 2>(a, t, wr, wk, grid, stream);
            }
            else
            {
                if (nCells > 0)
                    hipLaunchKernelGGL((kGdmHex8Ip<FULLDMG, false>), ipGrid, dim3(64), 0, stream, a, t, state);
                LaunchNodes<FULLDMG, 1>(a, t, wr, wk, grid, stream);
            }
        }
        else
            LaunchNodes<FULLDMG, 0>(a, t, wr, wk, grid, stream);
    }

    // state: device memory of StateDoubles() doubles for the pre-pass records (nullptr: the plan's own)
    void Launch(const RowArgs& a, bool wr, bool wk, hipStream_t stream, double* state = nullptr)
    {
        if (nNodes > 0)
        {
            if (!state && ipState.p)
                state = ipState.As<double>();
            if (a.P.damage_application == 0)
                LaunchT<true>(a, wr, wk, state, stream);
            else
                LaunchT<false>(a, wr, wk, state, stream);
        }
        LaunchExtras(a, wr, wk, stream);
    }

    // R and K rows of the nodes beyond the node kernel's bounds (the generic row-owner body)
    void LaunchExtras(const RowArgs& a, bool wr, bool wk, hipStream_t stream)
    {
        if (nExtra > 0)
        {
            const dim3 grid((unsigned)((nExtra + 63) / 64));
            const int* list = extra.As<int>();
            if (wr && wk)
                hipLaunchKernelGGL((kAssembleRowsList<kGradientDamage, 3, 8, 8, 8, 4, true, true>), grid, dim3(64), 0,
                                   stream, a, list, nExtra);
            else if (wr)
                hipLaunchKernelGGL((kAssembleRowsList<kGradientDamage, 3, 8, 8, 8, 4, true, false>), grid, dim3(64), 0,
                                   stream, a, list, nExtra);
            else
                hipLaunchKernelGGL((kAssembleRowsList<kGradientDamage, 3, 8, 8, 8, 4, false, true>), grid, dim3(64), 0,
                                   stream, a, list, nExtra);
        }
    }

    // R + K of the nodes [first, end) only (one mesh of a batch), plain block order; the nodes beyond
    // the bounds are not visited (LaunchExtras covers them for the whole plan)
    void LaunchRange(const RowArgs& a, int first, int end, hipStream_t stream)
    {
        if (end <= first)
            return;
        GdmHex8Tables t = Tables();
        t.xcdMap = 0;
        t.pBegin = first;
        t.pEnd = end;
        t.order = posOf.p ? posOf.As<int>() : nullptr;  // launch index = node
        // per-mesh launches recompute the IP state in the node kernel (no pre-pass over the mesh's cells)
        if (a.P.damage_application == 0)
            LaunchNodes<true, 0>(a, t, true, true, end - first, stream);
        else
            LaunchNodes<false, 0>(a, t, true, true, end - first, stream);
    }
};

}  // namespace nuto_amd
