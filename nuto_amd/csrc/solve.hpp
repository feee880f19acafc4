// Constrained linear solve on the device: ConstrainedSystemSolver::Solve / SolveTrialState
// (Solve.cpp:8-78) with the unit constraint matrix C of Constraints::BuildUnitConstraintMatrix
// (Constraints.cpp:89-137).
//
// The reduced system (C^T K C) u = C^T f is never formed: the operator is applied as three CSR
// products (C, K, C^T) with K resident in HBM, and solved with Jacobi-preconditioned conjugate
// gradients or BiCGSTAB (the iterative solvers NuTo selects by name, EigenSparseSolve.cpp:44-49,
// whose default preconditioner is the diagonal). Every reduction runs in a fixed order (per-lane
// strided sums, then fixed trees), so a solve is bit-reproducible.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "devbuf.hpp"

namespace nuto_amd
{

// y = A x for a CSR matrix, LPR lanes per row: lane l sums the entries l, l + LPR, ... of its row
// in order, then the LPR partial sums are added by an xor butterfly (a fixed order).
template <int LPR>
__global__ void __launch_bounds__(256) kCsrSpmv(int nRows, const int* __restrict__ rowPtr, const int* __restrict__ col,
                                                const double* __restrict__ val, const double* __restrict__ x,
                                                double* __restrict__ y)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long r = t / LPR;
    const int l = (int)(t % LPR);
    double s = 0.0;
    if (r < nRows)
        for (int k = rowPtr[r] + l; k < rowPtr[r + 1]; k += LPR)
            s += val[k] * x[col[k]];
#pragma unroll
    for (int m = LPR / 2; m >= 1; m >>= 1)
        s += __shfl_xor(s, m, LPR);
    if (r < nRows && l == 0)
        y[r] = s;
}

// y = K x for the assembled monolithic K: the C0 field-0 rows of dof node p (rows C0 p + i) and
// its field-1 row (nU + p, when C1 = 1) hold the same L columns, so one wave per node reads each
// column index and x entry once for all of them (values stream coalesced, 8 B per nonzero plus
// 4 / (C0 + C1) B of indices). Lane l sums the entries l, l + 64, ... of each row in order, then
// the 64 partial sums are added by an xor butterfly (a fixed order).
template <int C0, int C1>
__global__ void __launch_bounds__(256) kNodeSpmv(int nNodes, int nU, const int* __restrict__ rowPtr,
                                                 const int* __restrict__ col, const double* __restrict__ val,
                                                 const double* __restrict__ x, double* __restrict__ y)
{
    constexpr int F = C0 + C1;
    const int p = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    const int l = (int)(threadIdx.x & 63);
    double acc[F];
#pragma unroll
    for (int i = 0; i < F; i++)
        acc[i] = 0.0;
    if (p < nNodes)
    {
        const int r0 = rowPtr[C0 * p];
        const int L = rowPtr[C0 * p + 1] - r0;
        const int re = C1 ? rowPtr[nU + p] : 0;
        for (int k = l; k < L; k += 64)
        {
            const double xv = x[col[r0 + k]];
#pragma unroll
            for (int i = 0; i < C0; i++)
                acc[i] += val[r0 + i * L + k] * xv;
            if constexpr (C1 == 1)
                acc[C0] += val[re + k] * xv;
        }
    }
#pragma unroll
    for (int i = 0; i < F; i++)
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1)
            acc[i] += __shfl_xor(acc[i], m, 64);
    if (p < nNodes && l == 0)
    {
#pragma unroll
        for (int i = 0; i < C0; i++)
            y[C0 * p + i] = acc[i];
        if constexpr (C1 == 1)
            y[nU + p] = acc[C0];
    }
}

// 1 / diag(C^T K C)_j, diag_j = sum over the entries (a, c_a), (b, c_b) of column j of C of
// c_a c_b K(a, b) (column j of C = row j of C^T; K's columns are ascending in every row: binary
// search)
__global__ void __launch_bounds__(256) kDiagCtKC(int nInd, const int* __restrict__ ctRow, const int* __restrict__ ctCol,
                                                 const double* __restrict__ ctVal, const int* __restrict__ kRow,
                                                 const int* __restrict__ kCol, const double* __restrict__ kVal,
                                                 double* __restrict__ d)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= nInd)
        return;
    double s = 0.0;
    for (int p = ctRow[j]; p < ctRow[j + 1]; p++)
    {
        const int a = ctCol[p];
        for (int q = ctRow[j]; q < ctRow[j + 1]; q++)
        {
            const int b = ctCol[q];
            int lo = kRow[a], hi = kRow[a + 1];
            while (lo < hi)
            {
                const int mid = (lo + hi) >> 1;
                if (kCol[mid] < b)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            if (lo < kRow[a + 1] && kCol[lo] == b)
                s += ctVal[p] * ctVal[q] * kVal[lo];
        }
    }
    d[j] = s != 0.0 ? 1.0 / s : 1.0;  // the inverse diagonal (1 where the diagonal is zero)
}

constexpr int kDotBlocks = 1024;

// partial[b] = sum over i = b * 256 + tid + k * gridDim * 256 of a[i] b[i] (k increasing per lane,
// then a fixed tree over the block's lanes)
__global__ void __launch_bounds__(256) kDotPartial(long long n, const double* __restrict__ a,
                                                   const double* __restrict__ b, double* __restrict__ partial)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        s += a[i] * b[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1)
    {
        if ((int)threadIdx.x < m)
            sh[threadIdx.x] += sh[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partial[blockIdx.x] = sh[0];
}

// out[slot] = sum of the np partials (one block, same fixed tree)
__global__ void __launch_bounds__(256) kDotFinal(int np, const double* __restrict__ partial, double* __restrict__ out)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += 256)
        s += partial[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1)
    {
        if ((int)threadIdx.x < m)
            sh[threadIdx.x] += sh[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[0] = sh[0];
}

// y = a x + b y (x, y of length n); a zero factor leaves its operand unread (y = 0 with a = b = 0)
__global__ void __launch_bounds__(256) kAxpby(long long n, double a, const double* __restrict__ x, double b,
                                              double* __restrict__ y)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        y[i] = (a == 0.0 ? 0.0 : a * x[i]) + (b == 0.0 ? 0.0 : b * y[i]);
}

// z = invd x (the diagonal preconditioner)
__global__ void __launch_bounds__(256) kJacobi(long long n, const double* __restrict__ invd,
                                               const double* __restrict__ x, double* __restrict__ z)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        z[i] = invd[i] * x[i];
}

// BiCGSTAB's p = r + beta (p - w v) and y = invd p (the preconditioned p)
__global__ void __launch_bounds__(256) kBicgPY(long long n, const double* __restrict__ r, double beta, double w,
                                               const double* __restrict__ v, const double* __restrict__ invd,
                                               double* __restrict__ p, double* __restrict__ y)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
    {
        const double pi = r[i] + beta * (p[i] - w * v[i]);
        p[i] = pi;
        y[i] = invd[i] * pi;
    }
}

// scal[2] = num / (the dot product of kDotPartial's np partials, kDotFinal's tree): BiCGSTAB's
// alpha = rho / (r0 . v), conjugate gradients' alpha = (r . z) / (p . q)
__global__ void __launch_bounds__(256) kBicgAlpha(int np, const double* __restrict__ partial, double num,
                                                  double* __restrict__ scal)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += 256)
        s += partial[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1)
    {
        if ((int)threadIdx.x < m)
            sh[threadIdx.x] += sh[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        scal[2] = num / sh[0];
}

// Block sums of two per-lane values with kDotPartial's tree: partial[blockIdx] = the block's sum of
// s0, partial[gridDim + blockIdx] = its sum of s1
__device__ inline void BlockTree2(double s0, double s1, double* __restrict__ partial)
{
    __shared__ double sh[2][256];
    sh[0][threadIdx.x] = s0;
    sh[1][threadIdx.x] = s1;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1)
    {
        if ((int)threadIdx.x < m)
        {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + m];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + m];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        partial[blockIdx.x] = sh[0][0];
        partial[gridDim.x + blockIdx.x] = sh[1][0];
    }
}

// Two dot products in one pass, (a . b) and (c . d), in kDotPartial's element order
__global__ void __launch_bounds__(256) kDot2Partial(long long n, const double* __restrict__ a,
                                                    const double* __restrict__ b, const double* __restrict__ c,
                                                    const double* __restrict__ d, double* __restrict__ partial)
{
    double s0 = 0.0, s1 = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    {
        s0 += a[i] * b[i];
        s1 += c[i] * d[i];
    }
    BlockTree2(s0, s1, partial);
}

// out[0], out[1] = the sums of the two sets of np partials (one block of 512 threads); OMEGA: also
// out[3] = out[1] / out[0] when out[0] > 0, else 0 (BiCGSTAB's w = (t . s) / (t . t))
template <bool OMEGA>
__global__ void __launch_bounds__(512) kDot2Final(int np, const double* __restrict__ partial, double* __restrict__ out)
{
    __shared__ double sh[2][256];
    const int h = (int)threadIdx.x >> 8, t = (int)threadIdx.x & 255;
    double s = 0.0;
    for (int i = t; i < np; i += 256)
        s += partial[h * np + i];
    sh[h][t] = s;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1)
    {
        if (t < m)
            sh[h][t] += sh[h][t + m];
        __syncthreads();
    }
    if (t == 0)
        out[h] = sh[h][0];
    if constexpr (OMEGA)
    {
        __syncthreads();
        if (threadIdx.x == 0)
            out[3] = sh[0][0] > 0.0 ? sh[1][0] / sh[0][0] : 0.0;
    }
}

// BiCGSTAB's s = r - alpha v and z = invd s (the preconditioned s); alpha = scal[2]
__global__ void __launch_bounds__(256) kBicgS(long long n, const double* __restrict__ r, const double* __restrict__ scal,
                                              const double* __restrict__ v, const double* __restrict__ invd,
                                              double* __restrict__ s, double* __restrict__ z)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const double alpha = scal[2];
    if (i < n)
    {
        const double si = -alpha * v[i] + r[i];
        s[i] = si;
        z[i] = invd[i] * si;
    }
}

// BiCGSTAB's update x += alpha y, x += w z, r = s - w t, with the partial sums of (r . r) and
// (r0 . r) of the new residual (kDot2Partial's element order and layout; grid-stride over n);
// alpha = scal[2], w = scal[3]
__global__ void __launch_bounds__(256) kBicgUpdate(long long n, const double* __restrict__ scal,
                                                   const double* __restrict__ y, const double* __restrict__ z,
                                                   double* __restrict__ x, const double* __restrict__ s,
                                                   const double* __restrict__ t, double* __restrict__ r,
                                                   const double* __restrict__ r0, double* __restrict__ partial)
{
    const double alpha = scal[2], w = scal[3];
    double s0 = 0.0, s1 = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    {
        double xi = alpha * y[i] + x[i];
        xi = w * z[i] + xi;
        x[i] = xi;
        const double ri = -w * t[i] + s[i];
        r[i] = ri;
        s0 += ri * ri;
        s1 += r0[i] * ri;
    }
    BlockTree2(s0, s1, partial);
}

// Conjugate gradients' update x += alpha p, r -= alpha q, z = invd r, with the partial sums of
// (r . r) and (r . z) (kDot2Partial's element order and layout; grid-stride over n); alpha = scal[2]
__global__ void __launch_bounds__(256) kCgUpdate(long long n, const double* __restrict__ scal,
                                                 const double* __restrict__ p, const double* __restrict__ q,
                                                 double* __restrict__ x, double* __restrict__ r,
                                                 const double* __restrict__ invd, double* __restrict__ z,
                                                 double* __restrict__ partial)
{
    const double alpha = scal[2];
    double s0 = 0.0, s1 = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    {
        x[i] = alpha * p[i] + x[i];
        const double ri = -alpha * q[i] + r[i];
        r[i] = ri;
        const double zi = invd[i] * ri;
        z[i] = zi;
        s0 += ri * ri;
        s1 += ri * zi;
    }
    BlockTree2(s0, s1, partial);
}

// delta[dep[e]] = rhsNew[e] - rhsOld[e] (dependent dofs are distinct)
__global__ void __launch_bounds__(256) kScatterDelta(int nEq, const int* __restrict__ dep,
                                                     const double* __restrict__ rhsOld,
                                                     const double* __restrict__ rhsNew, double* __restrict__ delta)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < nEq)
        delta[dep[e]] = rhsNew[e] - (rhsOld ? rhsOld[e] : 0.0);
}

// Device CSR view
struct Csr
{
    int nRows = 0;
    const int* row = nullptr;
    const int* col = nullptr;
    const double* val = nullptr;
};

// node structure of an assembled K (rows per node of field 0 and 1), for the node product
struct NodeRows
{
    int nNodes = 0, nU = 0, c0 = 0, c1 = 0;
};

// Krylov solver on the operator C^T K C. Its vectors live in a workspace the caller keeps between
// solves; the stream orders every launch, and the host reads the scalars it branches on (one small
// copy per reduction or pair of reductions).
class ReducedSolver
{
public:
    // doubles of the workspace a constrained solve uses (SolveConstrained's vectors included): the
    // operator's two full-length temporaries, the reductions' partials and results, 8 Krylov vectors
    // of the reduced size, then (for the caller) 2 full-length and 3 reduced vectors
    static size_t WorkDoubles(long long nFull, int nInd)
    {
        return 4 * (size_t)nFull + 11 * (size_t)nInd + 2 * kDotBlocks + 8;
    }
    // work: WorkDoubles(nFull, nInd) doubles of device memory; host: 8 doubles of pinned memory (the
    // reductions' results). Both belong to the caller and are reused between solves.
    ReducedSolver(hipStream_t s, Csr K, Csr C, Csr Ct, long long nFull, int nInd, NodeRows nodes, double* work,
                  double* host)
        : mS(s), mK(K), mC(C), mCt(Ct), mM(nInd), mNodes(nodes), mHost(host)
    {
        t1 = work;
        t2 = t1 + nFull;
        part = t2 + nFull;
        scal = part + 2 * kDotBlocks;
        krylov = scal + 8;
        extra = krylov + 8 * (size_t)nInd;
    }
    // the caller's vectors after the solver's own: 2 of length nFull, then 3 of length nInd
    double* Extra() const
    {
        return extra;
    }

    // y (nInd) = C^T K C x (nInd)
    hipError_t Apply(const double* x, double* y)
    {
        Spmv(mC, x, t1);
        SpmvK(t1, t2);
        Spmv(mCt, t2, y);
        return hipGetLastError();
    }

    // C and C^T: one or two entries per row, one lane per row
    void Spmv(const Csr& A, const double* x, double* y)
    {
        if (A.nRows == 0)
            return;
        hipLaunchKernelGGL(kCsrSpmv<1>, dim3((unsigned)((A.nRows + 255) / 256)), dim3(256), 0, mS, A.nRows, A.row,
                           A.col, A.val, x, y);
    }
    void SpmvK(const double* x, double* y)
    {
        if (mK.nRows == 0)
            return;
        const NodeRows& nr = mNodes;
        if (nr.nNodes > 0)
        {
            const dim3 grid((unsigned)((nr.nNodes + 3) / 4)), block(256);
            if (nr.c0 == 3 && nr.c1 == 1)
            {
                hipLaunchKernelGGL((kNodeSpmv<3, 1>), grid, block, 0, mS, nr.nNodes, nr.nU, mK.row, mK.col, mK.val, x, y);
                return;
            }
            if (nr.c0 == 3 && nr.c1 == 0)
            {
                hipLaunchKernelGGL((kNodeSpmv<3, 0>), grid, block, 0, mS, nr.nNodes, nr.nU, mK.row, mK.col, mK.val, x, y);
                return;
            }
            if (nr.c0 == 2 && nr.c1 == 0)
            {
                hipLaunchKernelGGL((kNodeSpmv<2, 0>), grid, block, 0, mS, nr.nNodes, nr.nU, mK.row, mK.col, mK.val, x, y);
                return;
            }
        }
        const long long threads = (long long)mK.nRows * 16;
        hipLaunchKernelGGL(kCsrSpmv<16>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, mS, mK.nRows, mK.row,
                           mK.col, mK.val, x, y);
    }
    double Dot(long long n, const double* a, const double* b, hipError_t& err)
    {
        const int g = (int)std::min<long long>(kDotBlocks, (n + 255) / 256 > 0 ? (n + 255) / 256 : 1);
        hipLaunchKernelGGL(kDotPartial, dim3(g), dim3(256), 0, mS, n, a, b, part);
        hipLaunchKernelGGL(kDotFinal, dim3(1), dim3(256), 0, mS, g, part, scal);
        if ((err = hipGetLastError()) != hipSuccess ||
            (err = hipMemcpyAsync(mHost, scal, sizeof(double), hipMemcpyDeviceToHost, mS)) != hipSuccess ||
            (err = hipStreamSynchronize(mS)) != hipSuccess)
            return 0.0;
        return mHost[0];
    }
    // (a . b) and (c . d) with one host read: out[0], out[1]
    void Dot2(long long n, const double* a, const double* b, const double* c, const double* d, double out[2],
              hipError_t& err)
    {
        const int g = DotGrid(n);
        hipLaunchKernelGGL(kDot2Partial, dim3(g), dim3(256), 0, mS, n, a, b, c, d, part);
        Dot2Read(g, out, err);
    }
    // the two sums of the 2 x g partials in `part`
    void Dot2Read(int g, double out[2], hipError_t& err)
    {
        hipLaunchKernelGGL(kDot2Final<false>, dim3(1), dim3(512), 0, mS, g, part, scal);
        if ((err = hipGetLastError()) != hipSuccess ||
            (err = hipMemcpyAsync(mHost, scal, 2 * sizeof(double), hipMemcpyDeviceToHost, mS)) != hipSuccess ||
            (err = hipStreamSynchronize(mS)) != hipSuccess)
        {
            out[0] = out[1] = 0.0;
            return;
        }
        out[0] = mHost[0];
        out[1] = mHost[1];
    }
    static int DotGrid(long long n)
    {
        return (int)std::min<long long>(kDotBlocks, (n + 255) / 256 > 0 ? (n + 255) / 256 : 1);
    }
    void Axpby(long long n, double a, const double* x, double b, double* y)
    {
        if (n > 0)
            hipLaunchKernelGGL(kAxpby, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, mS, n, a, x, b, y);
    }
    void Precond(long long n, const double* d, const double* x, double* z)
    {
        if (n > 0)
            hipLaunchKernelGGL(kJacobi, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, mS, n, d, x, z);
    }

    // Preconditioned conjugate gradients from x = 0 on A = C^T K C: stops when |r|^2 < tol^2 |b|^2
    // or after maxIt iterations; `iters` counts the iterations before the converging one.
    hipError_t Cg(const double* b, const double* invd, double* x, double tol, int maxIt, int& iters, double& relRes)
    {
        const long long m = mM;
        hipError_t err = hipSuccess;
        double *R = krylov, *Z = krylov + m, *P = krylov + 2 * m, *Q = krylov + 3 * m;
        iters = 0;
        relRes = 0.0;
        Axpby(m, 0.0, nullptr, 0.0, x);
        const double bb = Dot(m, b, b, err);
        if (err != hipSuccess)
            return err;
        if (bb == 0.0)
            return hipSuccess;
        const double thr = std::max(tol * tol * bb, 1e-300);
        if ((err = hipMemcpyAsync(R, b, m * 8, hipMemcpyDeviceToDevice, mS)) != hipSuccess)
            return err;
        double rr = bb;
        if (rr < thr)
            return hipSuccess;
        Precond(m, invd, R, P);
        double absNew = Dot(m, R, P, err);
        int i = 0;
        while (i < maxIt && err == hipSuccess)
        {
            if ((err = Apply(P, Q)) != hipSuccess)
                break;
            // alpha = |r|_M^2 / (p . q) on the device (scal[2]); then x, r, z = invd r and the pair
            // (r . r), (r . z) in one pass, one host read
            const int g = DotGrid(m);
            hipLaunchKernelGGL(kDotPartial, dim3(g), dim3(256), 0, mS, m, P, Q, part);
            hipLaunchKernelGGL(kBicgAlpha, dim3(1), dim3(256), 0, mS, g, part, absNew, scal);
            hipLaunchKernelGGL(kCgUpdate, dim3(g), dim3(256), 0, mS, m, scal, P, Q, x, R, invd, Z, part);
            double d2[2];
            Dot2Read(g, d2, err);
            rr = d2[0];
            if (rr < thr)
                break;
            const double absOld = absNew;
            absNew = d2[1];
            Axpby(m, 1.0, Z, absNew / absOld, P);
            i++;
        }
        iters = i;
        relRes = std::sqrt(rr / bb);
        return err;
    }

    // Preconditioned BiCGSTAB from x = 0 on A = C^T K C (with a restart when rho becomes tiny)
    hipError_t Bicgstab(const double* b, const double* invd, double* x, double tol, int maxIt, int& iters,
                        double& relRes)
    {
        const long long m = mM;
        hipError_t err = hipSuccess;
        double *R = krylov, *R0 = krylov + m, *P = krylov + 2 * m, *V = krylov + 3 * m, *Y = krylov + 4 * m,
               *Z = krylov + 5 * m, *S = krylov + 6 * m, *T = krylov + 7 * m;
        iters = 0;
        relRes = 0.0;
        Axpby(m, 0.0, nullptr, 0.0, x);
        const double bb = Dot(m, b, b, err);
        if (err != hipSuccess)
            return err;
        if (bb == 0.0)
            return hipSuccess;
        if ((err = hipMemcpyAsync(R, b, m * 8, hipMemcpyDeviceToDevice, mS)) != hipSuccess ||
            (err = hipMemcpyAsync(R0, b, m * 8, hipMemcpyDeviceToDevice, mS)) != hipSuccess)
            return err;
        Axpby(m, 0.0, nullptr, 0.0, P);
        Axpby(m, 0.0, nullptr, 0.0, V);
        double r0sq = bb, rho = 1.0, alpha = 1.0, w = 1.0;
        const double tol2 = tol * tol * bb, eps2 = 2.220446049250313e-16 * 2.220446049250313e-16;
        double rr = bb, rhoNext = bb;  // (r . r) and (r0 . r) of the current residual (r = r0 = b)
        const unsigned vg = (unsigned)((m + 255) / 256);
        int i = 0, restarts = 0;
        while (rr > tol2 && i < maxIt && err == hipSuccess)
        {
            const double rhoOld = rho;
            rho = rhoNext;
            if (std::fabs(rho) < eps2 * r0sq)
            {
                // restart with the current residual r = b - A x
                if ((err = Apply(x, T)) != hipSuccess)
                    break;
                if ((err = hipMemcpyAsync(R, b, m * 8, hipMemcpyDeviceToDevice, mS)) != hipSuccess)
                    break;
                Axpby(m, -1.0, T, 1.0, R);
                if ((err = hipMemcpyAsync(R0, R, m * 8, hipMemcpyDeviceToDevice, mS)) != hipSuccess)
                    break;
                rho = r0sq = Dot(m, R, R, err);
                if (restarts++ == 0)
                    i = 0;
            }
            const double beta = (rho / rhoOld) * (alpha / w);
            hipLaunchKernelGGL(kBicgPY, dim3(vg), dim3(256), 0, mS, m, R, beta, w, V, invd, P, Y);
            if ((err = Apply(Y, V)) != hipSuccess)
                break;
            // alpha = rho / (r0 . v), w = (t . s) / (t . t) on the device; the host reads them with
            // (r . r) and the next rho at the end of the iteration
            const int g = DotGrid(m);
            hipLaunchKernelGGL(kDotPartial, dim3(g), dim3(256), 0, mS, m, R0, V, part);
            hipLaunchKernelGGL(kBicgAlpha, dim3(1), dim3(256), 0, mS, g, part, rho, scal);
            hipLaunchKernelGGL(kBicgS, dim3(vg), dim3(256), 0, mS, m, R, scal, V, invd, S, Z);
            if ((err = Apply(Z, T)) != hipSuccess)
                break;
            hipLaunchKernelGGL(kDot2Partial, dim3(g), dim3(256), 0, mS, m, T, T, T, S, part);
            hipLaunchKernelGGL(kDot2Final<true>, dim3(1), dim3(512), 0, mS, g, part, scal);
            hipLaunchKernelGGL(kBicgUpdate, dim3(g), dim3(256), 0, mS, m, scal, Y, Z, x, S, T, R, R0, part);
            hipLaunchKernelGGL(kDot2Final<false>, dim3(1), dim3(512), 0, mS, g, part, scal);
            if ((err = hipGetLastError()) != hipSuccess ||
                (err = hipMemcpyAsync(mHost, scal, 4 * sizeof(double), hipMemcpyDeviceToHost, mS)) != hipSuccess ||
                (err = hipStreamSynchronize(mS)) != hipSuccess)
                break;
            rr = mHost[0];
            rhoNext = mHost[1];
            alpha = mHost[2];
            w = mHost[3];
            ++i;
        }
        iters = i;
        relRes = std::sqrt(rr / bb);
        return err;
    }


private:
    hipStream_t mS;
    Csr mK, mC, mCt;
    int mM;
    NodeRows mNodes;
    double* mHost = nullptr;
    double *t1 = nullptr, *t2 = nullptr, *part = nullptr, *scal = nullptr, *krylov = nullptr, *extra = nullptr;
};

}  // namespace nuto_amd
