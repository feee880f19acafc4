// Adjoint of the hex8 gradient-damage residual: the vector-Jacobian product of
// R(u, eeq; theta) = BuildVector(GradientDamage<3>::Gradient) (SimpleAssembler.cpp:38-65,
// GradientDamage.h:70-90) with a cotangent lambda (length n_rows, layout of R):
//   d_state = (dR / d(u, eeq))^T lambda = K^T lambda   (K = BuildMatrix(Hessian0), GradientDamage.h:92-120)
//   d_theta = (dR / d theta)^T lambda for the per-cell fields theta = (E, f_t, G_f)
//             (kappa0 = f_t / E, beta = f_t / G_f, DamageLawExponential.h:32-37)
// without forming K. The history kappa is a constant of the residual.
//
// Per IP (weight s = detJ w), with eps_l = B_u lambda_u, g_l = B_e lambda_e, l_N = N lambda_e:
//   tau    = (d sigma / d eps) eps_l - l_N (d eps_eq / d eps)       (K_uu^T, K_eu^T parts)
//   d_u,a  += s B_a^T tau
//   d_e,a  += s (N_a (omega' kappa' (d sigma / d omega . eps_l) + l_N) + c g_a . g_l)   (K_ue^T, K_ee^T)
//   d_E    += s (sigma / E . eps_l + (d sigma / d omega . eps_l) (d omega / d kappa0) (-kappa0 / E))
//   d_ft   += s (d sigma / d omega . eps_l) (d omega / d kappa0 / E + d omega / d beta / G_f)
//   d_Gf   += s (d sigma / d omega . eps_l) (d omega / d beta) (-beta / G_f)
// (d sigma / d eps is symmetric, LinearElasticDamage.h:74-82; eps_eq and R_e do not depend on theta).
//
// Same decomposition as the forward node kernel: one 64-lane workgroup per dof node p, lane (r, ip)
// recomputes the IP state of incident cell r (nothing is stored between forward and adjoint),
// p's entries of d_state are summed per cell over its IPs in order, then over the cells in
// order; d_theta of a cell is written by the workgroup of the cell's local node 0. Nodes beyond the
// node kernel's bounds get the same computation from kGdmHex8VjpList, over any number of cells.
#pragma once

#include "gdm_hex8.hpp"

namespace nuto_amd
{

// Terms of lane (r, ip = j) of a node with local index la in incident cell r: t[0..3] the node's
// d_state terms (d_u, d_e) at that IP, t[4..6] the cell's d_theta terms when OWN (the node is the
// cell's local node 0 and d_theta is requested). The 8 lanes of the cell call it together (the node
// values are shared between them).
template <bool FULLDMG>
__device__ inline void VjpTerms(const RowArgs& A, const NodeFetch& f, const double lv[4], int j, bool own,
                                double t[7])
{
    using namespace gdm8;
    const LawParams& P = A.P;
    const int la = f.la;
    // material of the cell: theta = (E, f_t, G_f); scalar parameters map to f_t = kappa0 E, G_f = f_t / beta
    double E = P.E, k0 = P.kappa0, beta = P.beta;
    double ft = k0 * E, gf = ft / beta;
    if (A.thE)
    {
        E = f.th[0];
        ft = f.th[1];
        gf = f.th[2];
        k0 = ft / E;
        beta = ft / gf;
    }
    const double m3K = P.cK * E, m2G = P.cG * E;

    const int ip = j;
    double Hu[3][3], he[3], eN;
    GatherGradients(f.v, ip, Hu, he, eN);
    double Hl[3][3], hl[3], lN;
    GatherGradients(lv, ip, Hl, hl, lN);
    const double inv[3][3] = {{f.g[0], f.g[1], f.g[2]}, {f.g[3], f.g[4], f.g[5]}, {f.g[6], f.g[7], f.g[8]}};
    const double s = f.g[9];
    GdmIp st;
    GdmIpState<FULLDMG>(Hu, he, eN, inv, f.kh, m3K, m2G, k0, beta, P, st);
    double m[6];
    GdmMisesDerivative(st, P, m);
    // eps_l = B_u lambda_u, g_l = B_e lambda_e
    double Gl[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            Gl[i][c] = Hl[i][0] * inv[0][c] + Hl[i][1] * inv[1][c] + Hl[i][2] * inv[2][c];
    const double el[6] = {Gl[0][0], Gl[1][1], Gl[2][2], Gl[1][2] + Gl[2][1], Gl[2][0] + Gl[0][2],
                          Gl[0][1] + Gl[1][0]};
    double gl[3];
    GradOf(hl, inv, gl);
    // tau = (d sigma / d eps) eps_l - l_N m
    const double elV = (el[0] + el[1] + el[2]) * kThird;
    const double sVl = (1.0 - st.om * st.H) * m3K * elV;
    double tau[6];
#pragma unroll
    for (int i = 0; i < 3; i++)
        tau[i] = sVl + st.d2G * (el[i] - elV) - lN * m[i];
#pragma unroll
    for (int i = 3; i < 6; i++)
        tau[i] = 0.5 * st.d2G * el[i] - lN * m[i];
    double dsel = 0.0, sgel = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++)
    {
        dsel += st.dsw[i] * el[i];
        sgel += st.sig[i] * el[i];
    }
    double da[3], Na, ga[3];
    ShapeAt(la, ip, da, Na);
    GradOf(da, inv, ga);
    double bt[3];
    BTv3(ga, tau, bt);
    t[0] = s * bt[0];
    t[1] = s * bt[1];
    t[2] = s * bt[2];
    t[3] = s * (Na * (st.dom * st.dk * dsel + lN) + P.c * (ga[0] * gl[0] + ga[1] * gl[1] + ga[2] * gl[2]));
    if (own)
    {
        // d omega / d kappa0 and d omega / d beta of DamageLawExponential (zero below kappa0)
        double dwk = 0.0, dwb = 0.0;
        if (!(st.kap < k0))
        {
            const double ik = 1.0 / st.kap;
            const double ex = exp(beta * (k0 - st.kap));
            dwk = -ik * (1.0 - P.alpha + P.alpha * ex) - k0 * ik * P.alpha * beta * ex;
            dwb = -k0 * ik * P.alpha * ex * (k0 - st.kap);
        }
        t[4] = s * (sgel / E + dsel * dwk * (-k0 / E));
        t[5] = s * dsel * (dwk / E + dwb / gf);
        t[6] = s * dsel * dwb * (-beta / gf);
    }
}

// The lanes' terms (t, 7 per lane) of up to 8 cells summed per cell over its IPs in order
// (Cell.h:108-117): rcell[c * 8 + r] the d_state component c of cell r; the d_theta of the cells
// with ownerOf[r] written to dtheta. red, rcell: the workgroup's LDS.
__device__ inline void VjpCellSums(const RowArgs& A, const double t[7], double* red, double* rcell, const int* cellOf,
                                   const int* ownerOf, double* dtheta)
{
    const int lane = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 7; k++)
        red[k * 64 + lane] = t[k];
    __syncthreads();
    if (lane < 32)
    {
        const double* x = red + (lane >> 3) * 64 + (lane & 7) * 8;
        double a = x[0];
#pragma unroll
        for (int ip = 1; ip < 8; ip++)
            a += x[ip];
        rcell[lane] = a;
    }
    else if (dtheta && lane < 56)
    {
        const int k = (lane - 32) >> 3, rr = lane & 7;
        if (ownerOf[rr])
        {
            const double* x = red + (4 + k) * 64 + rr * 8;
            double a = x[0];
#pragma unroll
            for (int ip = 1; ip < 8; ip++)
                a += x[ip];
            dtheta[(size_t)k * A.nCells + cellOf[rr]] = a;
        }
    }
    __syncthreads();
}

template <bool FULLDMG>
__global__ void __launch_bounds__(64) kGdmHex8Vjp(const RowArgs A, const GdmHex8Tables T, const double* lam,
                                                   double* dstate, double* dtheta)
{
    using namespace gdm8;
    __shared__ double red[7 * 64];  // [term][r][ip]: d_u (3), d_e, d_E, d_ft, d_Gf
    __shared__ double rcell[4 * 8];
    __shared__ int cellOf[8];
    __shared__ int ownerOf[8];

    const int lane = threadIdx.x;
    const int r = lane >> 3, j = lane & 7;
    const int pos = NodeOfBlock(T);
    if (pos < 0)
        return;
    const int p = T.nodeHead[pos].x;
    NodeFetch f;
    FetchNode<false>(A, T, pos, r, j, f);
    const bool valid = f.cell >= 0;
    double lv[4] = {0.0, 0.0, 0.0, 0.0};  // lambda of node j of cell r
    if (valid)
    {
        lv[0] = lam[(size_t)f.q * 3];
        lv[1] = lam[(size_t)f.q * 3 + 1];
        lv[2] = lam[(size_t)f.q * 3 + 2];
        lv[3] = lam[(size_t)A.nU + f.q];
    }
    if (j == 0)
    {
        cellOf[r] = f.cell;
        ownerOf[r] = valid && f.la == 0;
    }
    double t[7] = {0, 0, 0, 0, 0, 0, 0};
    if (valid)
        VjpTerms<FULLDMG>(A, f, lv, j, dtheta && f.la == 0, t);
    VjpCellSums(A, t, red, rcell, cellOf, ownerOf, dtheta);
    // over the cells in order
    if (lane < 4)
    {
        const double* x = rcell + lane * 8;
        double a = 0.0;
#pragma unroll
        for (int rr = 0; rr < 8; rr++)
            a += x[rr];
        if (lane < 3)
            dstate[(size_t)p * 3 + lane] = a;
        else
            dstate[(size_t)A.nU + p] = a;
    }
}

// The adjoint of the nodes beyond the node kernel's bounds (more than 8 incident cells or 32
// neighbours; listed in `nodes`): one workgroup per listed node, which visits the node's incident
// cells (the node graph's incidence, in order) 8 at a time with the lanes of kGdmHex8Vjp, the
// geometry records of the plan and the IP state recomputed from u, eeq and kappa; the cells in
// order as in kGdmHex8Vjp. It writes d_theta of the cells whose local node 0 is a listed node.
template <bool FULLDMG>
__global__ void __launch_bounds__(64) kGdmHex8VjpList(const RowArgs A, const GdmHex8Tables T, const int* nodes, int n,
                                                       const double* lam, double* dstate, double* dtheta)
{
    using namespace gdm8;
    __shared__ double red[7 * 64];
    __shared__ double rcell[4 * 8];
    __shared__ int cellOf[8];
    __shared__ int ownerOf[8];

    const int lane = threadIdx.x;
    const int r = lane >> 3, j = lane & 7;
    if ((int)blockIdx.x >= n)
        return;
    const int p = nodes[blockIdx.x];
    const int beg = A.incPtr[p], end = A.incPtr[p + 1];
    double acc = 0.0;  // lanes 0..3: d_state component over the cells in order
    for (int b0 = beg; b0 < end; b0 += 8)
    {
        const int k = b0 + r;
        const bool valid = k < end;  // uniform within the cell's 8 lanes
        NodeFetch f;
        f.cell = valid ? A.incCell[k] : -1;
        f.la = valid ? A.incLocal[k] : 0;
        double lv[4] = {0.0, 0.0, 0.0, 0.0};
        if (valid)
        {
            const double* gr = T.geo + ((size_t)f.cell * 8 + j) * kGeo;
#pragma unroll
            for (int i = 0; i < kGeo; i++)
                f.g[i] = gr[i];
            const int q = A.dconn[(size_t)f.cell * 8 + j];
            f.q = q;
            const int cid = A.cellId ? A.cellId[f.cell] : f.cell;
            f.kh = A.kappa[(size_t)cid + (size_t)j * A.kappaLd];
            f.v[0] = A.u[(size_t)q * 3];
            f.v[1] = A.u[(size_t)q * 3 + 1];
            f.v[2] = A.u[(size_t)q * 3 + 2];
            f.v[3] = A.e[q];
            if (A.thE)
            {
                f.th[0] = A.thE[f.cell];
                f.th[1] = A.thFt[f.cell];
                f.th[2] = A.thGf[f.cell];
            }
            lv[0] = lam[(size_t)q * 3];
            lv[1] = lam[(size_t)q * 3 + 1];
            lv[2] = lam[(size_t)q * 3 + 2];
            lv[3] = lam[(size_t)A.nU + q];
        }
        if (j == 0)
        {
            cellOf[r] = f.cell;
            ownerOf[r] = valid && f.la == 0;
        }
        double t[7] = {0, 0, 0, 0, 0, 0, 0};
        if (valid)
            VjpTerms<FULLDMG>(A, f, lv, j, dtheta && f.la == 0, t);
        VjpCellSums(A, t, red, rcell, cellOf, ownerOf, dtheta);
        if (lane < 4)
        {
#pragma unroll
            for (int rr = 0; rr < 8; rr++)
                acc += rcell[lane * 8 + rr];
        }
        __syncthreads();  // the next batch rewrites red, rcell and the cell tables
    }
    if (lane < 3)
        dstate[(size_t)p * 3 + lane] = acc;
    else if (lane == 3)
        dstate[(size_t)A.nU + p] = acc;
}

inline void LaunchGdmHex8Vjp(const GdmHex8Plan& plan, const RowArgs& a, const double* lam, double* dstate,
                             double* dtheta, hipStream_t stream)
{
    const GdmHex8Tables t = plan.Tables();
    const bool full = a.P.damage_application == 0;
    if (plan.nNodes > 0)
    {
        if (full)
            hipLaunchKernelGGL(kGdmHex8Vjp<true>, dim3(plan.grid), dim3(64), 0, stream, a, t, lam, dstate, dtheta);
        else
            hipLaunchKernelGGL(kGdmHex8Vjp<false>, dim3(plan.grid), dim3(64), 0, stream, a, t, lam, dstate, dtheta);
    }
    if (plan.nExtra > 0)
    {
        const int* list = plan.extra.As<int>();
        if (full)
            hipLaunchKernelGGL(kGdmHex8VjpList<true>, dim3(plan.nExtra), dim3(64), 0, stream, a, t, list, plan.nExtra,
                               lam, dstate, dtheta);
        else
            hipLaunchKernelGGL(kGdmHex8VjpList<false>, dim3(plan.nExtra), dim3(64), 0, stream, a, t, list,
                               plan.nExtra, lam, dstate, dtheta);
    }
}

}  // namespace nuto_amd
