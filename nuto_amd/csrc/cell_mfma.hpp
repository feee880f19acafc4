// Cell matrices of 3D momentum balance on the FP64 matrix cores: the algebra of kCellGram
// (K_ab = r G_ab + G_ab^T + tr(G_ab) I + R1_ab, Cell.h:105-119 over MomentumBalance<3>::Hessian0)
// with the sums over the IPs as products of LDS-staged matrices (ND = 3 x nodes, row 3n + c):
//     G  = (c2 B)^T B,  B[ip][3n + c] = g_n[c] = (dN_n/dX)_c at IP ip, c2 = s G (1 - omega)
//     R1 = U^T V,       U[ip][3n + c] = (B_n^T t)_c, V[ip][3n + c] = (B_n^T d eps_eq / d eps)_c
// The IP count is padded to NK (a multiple of 4, zero rows) and ND to NDP (a multiple of 16, zero
// columns). One workgroup of NDP / 16 waves per cell: wave w owns the 16-row tile row w of G and R1
// (NDP / 16 tiles of each). Each IP term is read from LDS once per wave instead of once per block
// lane as in kCellGram, whose LDS reads bound it.
//
// v_mfma_f64_16x16x4f64 lane layout (gfx950): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]
// and holds D[l / 16 + 4 i][l % 16] in output i (i < 4).
#pragma once

#include <hip/hip_runtime.h>

#include "cell_gram.hpp"

namespace nuto_amd
{

typedef double MfmaF64x4 __attribute__((ext_vector_type(4)));

template <int NND>
struct MfmaShape
{
    static constexpr int ND = 3 * NND;
    static constexpr int NDP = (ND + 15) / 16 * 16;  // padded dofs
    static constexpr int TILES = NDP / 16;            // tile rows = waves, tile columns per wave
    static constexpr int NT = 64 * TILES;
};

// UNILATERAL damage application (a second Gram) is not provided here: the caller uses kCellGram.
template <int KIND, int NNC, int NND, int NIP>
__global__ void __launch_bounds__(MfmaShape<NND>::NT) kCellMfma(const RowArgs A, int c0, int c1,
                                                                double* __restrict__ Ke, double* __restrict__ Re)
{
    using namespace gram;
    using Sh = MfmaShape<NND>;
    constexpr int ND = Sh::ND, NDP = Sh::NDP, TILES = Sh::TILES, NT = Sh::NT;
    constexpr int NK = (NIP + 3) / 4 * 4;
    constexpr int LDM = NDP + 2;  // row stride of B, U, V: the 4 IP rows of one read fall on other banks
    constexpr int LDG = NDP + 1;  // row stride of G
    constexpr bool LD = KIND == kLocalDamage;
    constexpr int NMAT = LD ? 3 : 1;
    constexpr int STAGE = NMAT * NK * LDM > NDP * LDG ? NMAT * NK * LDM : NDP * LDG;
    constexpr int S = 1 + NNC * 3 + NND + NND * 3;  // IP table row: w, dNc, N, dNd

    __shared__ double sX[1][NNC * 3];
    __shared__ double sU[1][NND * 3];
    __shared__ double sJG[1][NIP][18];
    __shared__ double sIp[1][NIP][kIpRec];
    __shared__ double sC2[NK];
    __shared__ double sRb[NIP][ND];
    __shared__ double sStage[STAGE];  // B, U, V for the products, then G
    double* sB = sStage;
    double* sUm = sStage + NK * LDM;
    double* sVm = sStage + 2 * NK * LDM;
    double* sG = sStage;

    const int t = (int)threadIdx.x;
    const int cell = c0 + (int)blockIdx.x;  // grid: c1 - c0 workgroups
    CellIpRecords<KIND, NNC, NND, NIP, 1, NT>(A, c1, cell, sX, sU, sJG, sIp);
    __syncthreads();

    // ---- per (node, IP): g_n, u_n, v_n into B, U, V; the R terms s B_n^T sigma ------------------------
    for (int k = t; k < NIP * NND; k += NT)
    {
        const int ip = k / NND, n = k - ip * NND;
        const double* rec = sIp[0][ip];
        const double* dN = A.tab + ip * S + 1 + NNC * 3 + NND + n * 3;
        double g[3];
#pragma unroll
        for (int c = 0; c < 3; c++)
            g[c] = dN[0] * rec[c] + dN[1] * rec[3 + c] + dN[2] * rec[6 + c];
        double* b = sB + ip * LDM + 3 * n;
        b[0] = g[0];
        b[1] = g[1];
        b[2] = g[2];
        double bt[3];
        BTv3(g, rec + 12, bt);
        sRb[ip][3 * n] = rec[9] * bt[0];
        sRb[ip][3 * n + 1] = rec[9] * bt[1];
        sRb[ip][3 * n + 2] = rec[9] * bt[2];
        if constexpr (LD)
        {
            double w[3];
            BTv3(g, rec + 18, w);
            double* u = sUm + ip * LDM + 3 * n;
            u[0] = w[0];
            u[1] = w[1];
            u[2] = w[2];
            BTv3(g, rec + 24, w);
            double* v = sVm + ip * LDM + 3 * n;
            v[0] = w[0];
            v[1] = w[1];
            v[2] = w[2];
        }
    }
    // zero padding: IP rows NIP .. NK - 1, dof columns ND .. NDP - 1 of each matrix
    constexpr int PADR = (NK - NIP) * NDP, PADC = NIP * (NDP - ND), PW = NDP > ND ? NDP - ND : 1;
    for (int k = t; k < NMAT * (PADR + PADC); k += NT)
    {
        const int m = k / (PADR + PADC), e = k - m * (PADR + PADC);
        const int row = e < PADR ? NIP + e / NDP : (e - PADR) / PW;
        const int col = e < PADR ? e % NDP : ND + (e - PADR) % PW;
        sStage[m * NK * LDM + row * LDM + col] = 0.0;
    }
    for (int k = t; k < NK; k += NT)
        sC2[k] = k < NIP ? sIp[0][k][11] : 0.0;
    __syncthreads();

    // ---- products: wave w, tile row w, all tile columns ------------------------------------------
    const int w = t >> 6, l = t & 63;
    const int lr = l & 15, lk = l >> 4;
    MfmaF64x4 accG[TILES], accR[LD ? TILES : 1];
#pragma unroll
    for (int tc = 0; tc < TILES; tc++)
    {
        accG[tc] = MfmaF64x4{0.0, 0.0, 0.0, 0.0};
        if constexpr (LD)
            accR[tc] = MfmaF64x4{0.0, 0.0, 0.0, 0.0};
    }
#pragma unroll
    for (int ks = 0; ks < NK / 4; ks++)
    {
        const int k = 4 * ks + lk;
        const double aG = sC2[k] * sB[k * LDM + 16 * w + lr];
        const double aU = LD ? sUm[k * LDM + 16 * w + lr] : 0.0;
#pragma unroll
        for (int tc = 0; tc < TILES; tc++)
        {
            accG[tc] = __builtin_amdgcn_mfma_f64_16x16x4f64(aG, sB[k * LDM + 16 * tc + lr], accG[tc], 0, 0, 0);
            if constexpr (LD)
                accR[tc] = __builtin_amdgcn_mfma_f64_16x16x4f64(aU, sVm[k * LDM + 16 * tc + lr], accR[tc], 0, 0, 0);
        }
    }
    __syncthreads();  // B, U, V consumed: the stage now holds G
#pragma unroll
    for (int tc = 0; tc < TILES; tc++)
#pragma unroll
        for (int i = 0; i < 4; i++)
            sG[(16 * w + lk + 4 * i) * LDG + 16 * tc + lr] = accG[tc][i];
    __syncthreads();

    // ---- K_e = r G + blockwise G^T + tr I (+ R1); own entries of the lane ------------------------------
    if (cell < c1)
    {
        double r;  // one Gram: c1 / c2 of the law
        if constexpr (KIND == kLinearElastic)
            r = 2.0 * A.P.nu / (1.0 - 2.0 * A.P.nu);  // lambda / mu
        else
        {
            double E, k0, beta;
            CellMaterial(A, cell, E, k0, beta);
            const double m3K = 3.0 * E / (3.0 - 6.0 * A.P.nu), m2G = 2.0 * E / (2.0 + 2.0 * A.P.nu);
            r = (m3K - m2G) / (1.5 * m2G);  // ((3K - 2G) / 3) / G
        }
        double* ke = Ke + (size_t)(cell - c0) * ND * ND;
#pragma unroll
        for (int tc = 0; tc < TILES; tc++)
#pragma unroll
            for (int i = 0; i < 4; i++)
            {
                const int row = 16 * w + lk + 4 * i, col = 16 * tc + lr;
                if (row < ND && col < ND)
                {
                    const int a = row / 3, ii = row - 3 * a, b = col / 3, jj = col - 3 * b;
                    double v = r * accG[tc][i] + sG[(3 * a + jj) * LDG + 3 * b + ii];
                    if (ii == jj)
                        v += sG[3 * a * LDG + 3 * b] + sG[(3 * a + 1) * LDG + 3 * b + 1] +
                             sG[(3 * a + 2) * LDG + 3 * b + 2];
                    if constexpr (LD)
                        v += accR[tc][i];
                    ke[(size_t)row * ND + col] = v;
                }
            }
        // R_e: the IPs in order per dof
        for (int k = t; k < ND; k += NT)
        {
            double acc = 0.0;
            for (int ip = 0; ip < NIP; ip++)
                acc += sRb[ip][k];
            Re[(size_t)(cell - c0) * ND + k] = acc;
        }
    }
}

}  // namespace nuto_amd
