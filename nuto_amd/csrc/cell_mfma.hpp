// Cell matrices of 3D momentum balance on the FP64 matrix cores: the algebra of kCellGram
// (K_ab = r G_ab + G_ab^T + tr(G_ab) I + R1_ab, Cell.h:105-119 over MomentumBalance<3>::Hessian0)
// with the sums over the IPs as products of LDS-staged matrices (ND = 3 x nodes, row 3n + c):
//     G  = (c2 B)^T B,  B[ip][3n + c] = g_n[c] = (dN_n/dX)_c at IP ip, c2 = s G (1 - omega)
//     R1 = U^T V,       U[ip][3n + c] = (B_n^T t)_c, V[ip][3n + c] = (B_n^T d eps_eq / d eps)_c
// The IPs are summed in chunks padded to a multiple of 4 (zero rows) and ND is padded to NDP (a
// multiple of 16, zero columns). One workgroup of NDP / 16 waves per cell: wave w owns the 16-row tile row w of G and R1
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
// The IPs are processed in chunks of KC (12, 8 for rules of 8 to 11 points, or NIP rounded up to 4
// for smaller rules; hex20: 240 of the 256 lanes stage a chunk's (node, IP) terms), so only one
// chunk of B, U, V is staged at a time; the IP states, the chunk's staging and the final G share
// one LDS pool (35 KB for hex20 with the block traces), which leaves room for 4 workgroups per CU.
template <int KIND, int NNC, int NND, int NIP>
__global__ void __launch_bounds__(MfmaShape<NND>::NT) __attribute__((amdgpu_waves_per_eu(4)))
kCellMfma(const RowArgs A, int c0, int c1, double* __restrict__ Ke, double* __restrict__ Re)
{
    using namespace gram;
    using Sh = MfmaShape<NND>;
    constexpr int ND = Sh::ND, NDP = Sh::NDP, TILES = Sh::TILES, NT = Sh::NT;
    constexpr int KC = NIP >= 12 ? 12 : (NIP >= 8 ? 8 : (NIP + 3) / 4 * 4);  // IPs per chunk (a multiple of 4)
    constexpr int NCH = (NIP + KC - 1) / KC;
    constexpr int LDM = NDP + 2;  // row stride of B, U, V: the 4 IP rows of one read fall on other banks
    constexpr int LDG = ND + 1;   // row stride of G (rows and columns below ND only)
    constexpr bool LD = KIND == kLocalDamage;
    constexpr int NMAT = LD ? 3 : 1;
    constexpr int S = 1 + NNC * 3 + NND + NND * 3;  // IP table row: w, dNc, N, dNd
    // pool: [J/grad records NIP x 18 | IP states NIP x kIpRec | B, U, V chunks | R terms KC x ND], then G
    constexpr int OJG = 0, OIP = OJG + NIP * 18, OST = OIP + NIP * kIpRec, ORB = OST + NMAT * KC * LDM;
    constexpr int POOLA = ORB + KC * ND, POOLB = ND * LDG + NND * NND;  // G, then the block traces
    constexpr int POOL = POOLA > POOLB ? POOLA : POOLB;

    __shared__ double sX[1][NNC * 3];
    __shared__ double sU[1][NND * 3];
    __shared__ double sC2[KC];
    __shared__ double sR[ND];
    __shared__ __align__(16) double sPool[POOL];
    auto sJG = reinterpret_cast<double(*)[NIP][18]>(sPool + OJG);
    auto sIp = reinterpret_cast<double(*)[NIP][kIpRec]>(sPool + OIP);
    double* sB = sPool + OST;
    double* sUm = sPool + OST + KC * LDM;
    double* sVm = sPool + OST + 2 * KC * LDM;
    double* sRb = sPool + ORB;  // [i][dof] of the chunk
    double* sG = sPool;         // after the last chunk
    double* sTr = sPool + ND * LDG;  // tr(G_ab) at a * NND + b, after G

    const int t = (int)threadIdx.x;
    const int cell = c0 + (int)blockIdx.x;  // grid: c1 - c0 workgroups
    for (int k = t; k < ND; k += NT)
        sR[k] = 0.0;
    CellIpRecords<KIND, NNC, NND, NIP, 1, NT>(A, c1, cell, sX, sU, sJG, sIp);
    // zero dof columns ND .. NDP - 1 of the staged matrices (never written by the chunks)
    if constexpr (NDP > ND)
    {
        for (int k = t; k < NMAT * KC * (NDP - ND); k += NT)
        {
            const int m = k / (KC * (NDP - ND)), e = k - m * (KC * (NDP - ND));
            sB[m * KC * LDM + (e / (NDP - ND)) * LDM + ND + e % (NDP - ND)] = 0.0;
        }
    }

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

#pragma unroll 1
    for (int ch = 0; ch < NCH; ch++)
    {
        const int ip0 = ch * KC;
        // rows of this chunk: KC, the last chunk's IPs rounded up to a multiple of 4
        const int rows = NIP - ip0 >= KC ? KC : (NIP - ip0 + 3) / 4 * 4;
        __syncthreads();  // IP states written / the previous chunk's staging consumed
        // ---- per (node, IP of the chunk): g_n, u_n, v_n into B, U, V; the R terms s B_n^T sigma ------
        for (int k = t; k < rows * NND; k += NT)
        {
            const int i = k / NND, n = k - i * NND, ip = ip0 + i;
            double* b = sB + i * LDM + 3 * n;
            double* u = sUm + i * LDM + 3 * n;
            double* v = sVm + i * LDM + 3 * n;
            if (ip >= NIP)
            {
                // padding rows of the last chunk
                b[0] = b[1] = b[2] = 0.0;
                if constexpr (LD)
                {
                    u[0] = u[1] = u[2] = 0.0;
                    v[0] = v[1] = v[2] = 0.0;
                }
                continue;
            }
            const double* rec = sIp[0][ip];
            const double* dN = A.tab + ip * S + 1 + NNC * 3 + NND + n * 3;
            double g[3];
#pragma unroll
            for (int c = 0; c < 3; c++)
                g[c] = dN[0] * rec[c] + dN[1] * rec[3 + c] + dN[2] * rec[6 + c];
            b[0] = g[0];
            b[1] = g[1];
            b[2] = g[2];
            double bt[3];
            BTv3(g, rec + 12, bt);
            sRb[i * ND + 3 * n] = rec[9] * bt[0];
            sRb[i * ND + 3 * n + 1] = rec[9] * bt[1];
            sRb[i * ND + 3 * n + 2] = rec[9] * bt[2];
            if constexpr (LD)
            {
                double wv[3];
                BTv3(g, rec + 18, wv);
                u[0] = wv[0];
                u[1] = wv[1];
                u[2] = wv[2];
                BTv3(g, rec + 24, wv);
                v[0] = wv[0];
                v[1] = wv[1];
                v[2] = wv[2];
            }
        }
        for (int k = t; k < KC; k += NT)
            sC2[k] = ip0 + k < NIP ? sIp[0][ip0 + k][11] : 0.0;
        __syncthreads();
        // ---- R_e: the chunk's IPs in order per dof ---------------------------------------------------
        for (int k = t; k < ND; k += NT)
        {
            double acc = sR[k];
            for (int i = 0; i < KC && ip0 + i < NIP; i++)
                acc += sRb[i * ND + k];
            sR[k] = acc;
        }
        // ---- products: wave w, tile row w, all tile columns --------------------------------------------
#pragma unroll
        for (int ks = 0; ks < KC / 4; ks++)
        {
            if (4 * ks >= rows)
                break;  // uniform over the workgroup
            const int k = 4 * ks + lk;
            const double aG = sC2[k] * sB[k * LDM + 16 * w + lr];
            const double aU = LD ? sUm[k * LDM + 16 * w + lr] : 0.0;
#pragma unroll
            for (int tc = 0; tc < TILES; tc++)
            {
                accG[tc] = __builtin_amdgcn_mfma_f64_16x16x4f64(aG, sB[k * LDM + 16 * tc + lr], accG[tc], 0, 0, 0);
                if constexpr (LD)
                    accR[tc] =
                        __builtin_amdgcn_mfma_f64_16x16x4f64(aU, sVm[k * LDM + 16 * tc + lr], accR[tc], 0, 0, 0);
            }
        }
    }
    __syncthreads();  // every chunk consumed: the pool now holds G
#pragma unroll
    for (int tc = 0; tc < TILES; tc++)
#pragma unroll
        for (int i = 0; i < 4; i++)
        {
            const int row = 16 * w + lk + 4 * i, col = 16 * tc + lr;
            if (row < ND && col < ND)
                sG[row * LDG + col] = accG[tc][i];
        }
    __syncthreads();
    for (int k = t; k < NND * NND; k += NT)
    {
        const int a = k / NND, b = k - a * NND;
        const double* g = sG + 3 * a * LDG + 3 * b;
        sTr[k] = g[0] + g[LDG + 1] + g[2 * LDG + 2];
    }
    __syncthreads();

    // ---- K_e = r G + blockwise G^T + tr I (+ R1); own entries of the lane ------------------------------
    // Lane entries: rows 16 w + lk + 4 i (node a, component ii), columns 16 tc + lr (node b, component
    // jj); G_ab^T[ii][jj] = G[3 a + jj][3 b + ii]. Rows and columns past ND are clamped for the LDS
    // reads and not stored.
    if (cell < c1)
    {
        // one Gram: c1 / c2 of the law (E cancels: one value per launch)
        const double r = KIND == kLinearElastic ? A.P.rLE : A.P.rLD;
        int rowT[4], rowTr[4], rowII[4];
        bool rowIn[4];
#pragma unroll
        for (int i = 0; i < 4; i++)
        {
            const int row = 16 * w + lk + 4 * i;
            const int rc = row < ND ? row : ND - 1;
            const int a = rc / 3, ii = rc - 3 * a;
            rowT[i] = 3 * a * LDG + ii;
            rowTr[i] = a * NND;
            rowII[i] = ii;
            rowIn[i] = row < ND;
        }
        int colT[TILES], colTr[TILES], colJJ[TILES];
        bool colIn[TILES];
#pragma unroll
        for (int tc = 0; tc < TILES; tc++)
        {
            const int col = 16 * tc + lr;
            const int cc = col < ND ? col : ND - 1;
            const int b = cc / 3, jj = cc - 3 * b;
            colT[tc] = jj * LDG + 3 * b;
            colTr[tc] = b;
            colJJ[tc] = jj;
            colIn[tc] = col < ND;
        }
        double* ke = Ke + (size_t)(cell - c0) * ND * ND + (size_t)(16 * w + lk) * ND + lr;
#pragma unroll
        for (int tc = 0; tc < TILES; tc++)
#pragma unroll
            for (int i = 0; i < 4; i++)
            {
                double v = r * accG[tc][i] + sG[rowT[i] + colT[tc]];
                const double tr = sTr[rowTr[i] + colTr[tc]];
                v = rowII[i] == colJJ[tc] ? v + tr : v;
                if constexpr (LD)
                    v += accR[tc][i];
                if (rowIn[i] && colIn[tc])
                    ke[(size_t)(4 * i) * ND + 16 * tc] = v;
            }
        for (int k = t; k < ND; k += NT)
            Re[(size_t)(cell - c0) * ND + k] = sR[k];
    }
}

}  // namespace nuto_amd
