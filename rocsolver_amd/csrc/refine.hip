// refine.hip -- mixed-precision batched solve: FP32 LU on MFMA, FP64
// iterative refinement, FP64 fallback per matrix.
//
// Per matrix it follows LAPACK dsgesv (no rocSOLVER counterpart):
//   SA = fp32(A) (entries beyond the FP32 range -> ITER = -2, fallback)
//   SA = P L U in FP32 (zero pivot -> ITER = -3, fallback)
//   X  = fp64(U^-1 L^-1 P fp32(B))
//   repeat: R = B - A X in FP64; if every column has
//           ||R_c||_inf <= ||X_c||_inf * ||A||_inf * sqrt(n) * 2^-53: done (ITER = k)
//           X += fp64(U^-1 L^-1 P fp32(R))
//   after ITERMAX corrections: ITER = -(ITERMAX+1), fallback
//   fallback: X = A^-1 B with an FP64 LU of A (info = its first zero pivot).
// Every kernel of the sequence runs over the whole batch and skips the
// matrices whose state says so, so the launch sequence is fixed; the host
// only reads one counter per refinement step to stop early.
#include "internal.hpp"

#include <cfloat>
#include <cmath>

namespace rsa
{

constexpr int IR_THREADS = 256;

// state per matrix: askip != 0 -> not refining (converged or sent to the
// fallback); fskip != 0 -> no FP64 fallback needed
__global__ __launch_bounds__(IR_THREADS) void ir_init_kernel(int n, int nrhs, Batch<double> BB,
                                                             int ldb, const int* __restrict__ info32,
                                                             int* __restrict__ askip,
                                                             int* __restrict__ fskip,
                                                             int* __restrict__ iter,
                                                             int* __restrict__ info,
                                                             int* __restrict__ counters, int ncount)
{
    __shared__ int s_big;
    const int64_t b = blockIdx.x;
    const int tid = threadIdx.x;
    if(b == 0)
        for(int i = tid; i < ncount; i += IR_THREADS)
            counters[i] = 0;
    if(tid == 0)
        s_big = 0;
    __syncthreads();
    const double* Bm = BB[b];
    int big = 0;
    for(int e = tid; e < n * nrhs; e += IR_THREADS)
    {
        const double v = Bm[(e % n) + int64_t(e / n) * ldb];
        if(v > double(FLT_MAX) || v < -double(FLT_MAX))
            big = 1;
    }
    if(big)
        s_big = 1;
    __syncthreads();
    if(tid == 0)
    {
        info[b] = 0;
        if(s_big)
        {
            iter[b] = -2;
            askip[b] = 1;
            fskip[b] = 0;
        }
        else if(info32[b] != 0)
        {
            iter[b] = -3;
            askip[b] = 1;
            fskip[b] = 0;
        }
        else
        {
            iter[b] = 0;
            askip[b] = 0;
            fskip[b] = 1;
        }
    }
}

// R = B - A X (FP64, fma accumulation over the columns of A in order), then
// the dsgesv stopping test.  One workgroup per matrix, thread i owns row i;
// the first pass also measures ||A||_inf and the FP32 range of A.
template <int NR>
__global__ __launch_bounds__(IR_THREADS) void ir_residual_kernel(
    int n, int nrhs, Batch<double> AA, int lda, Batch<double> BB, int ldb, Batch<double> XX,
    int ldx, Batch<double> RR, int ldr, double* __restrict__ anorm, int it, int itermax,
    int* __restrict__ askip, int* __restrict__ fskip, int* __restrict__ iter,
    double* __restrict__ berr, int* __restrict__ counters)
{
    constexpr int NWV = IR_THREADS / 64;
    __shared__ double sX[IR_THREADS][NR];
    __shared__ double sred[NWV][2 * NR];
    __shared__ double s_an;
    __shared__ int s_ok;
    const int64_t b = blockIdx.x;
    if(askip[b] != 0)
        return;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const double* A = AA[b];
    const double* Bm = BB[b];
    const double* X = XX[b];
    double* R = RR[b];
    const bool first = (it == 0);

    if(tid == 0)
    {
        s_ok = 1;
        s_an = first ? 0.0 : anorm[b];
    }
    double an_part = 0, amax = 0;
    double worst = 0; // max over columns of ||r|| / (||A|| ||x||)
    for(int c0 = 0; c0 < nrhs; c0 += NR)
    {
        const bool measure = first && c0 == 0;
        double acc[NR], rn[NR], xn[NR];
#pragma unroll
        for(int q = 0; q < NR; ++q)
        {
            rn[q] = 0;
            xn[q] = 0;
        }
        for(int r0 = 0; r0 < n; r0 += IR_THREADS)
        {
            const int i = r0 + tid;
            double rs = 0;
#pragma unroll
            for(int q = 0; q < NR; ++q)
                acc[q] = (i < n && c0 + q < nrhs) ? Bm[i + int64_t(c0 + q) * ldb] : 0.0;
            for(int j0 = 0; j0 < n; j0 += IR_THREADS)
            {
                __syncthreads();
                const int jj = j0 + tid;
#pragma unroll
                for(int q = 0; q < NR; ++q)
                    sX[tid][q] = (jj < n && c0 + q < nrhs) ? X[jj + int64_t(c0 + q) * ldx] : 0.0;
                __syncthreads();
                const int jend = (n - j0) < IR_THREADS ? (n - j0) : IR_THREADS;
                if(i < n)
                    for(int j = 0; j < jend; ++j)
                    {
                        const double a = A[i + int64_t(j0 + j) * lda];
                        if(measure)
                        {
                            rs += fabs(a);
                            amax = fmax(amax, fabs(a));
                        }
#pragma unroll
                        for(int q = 0; q < NR; ++q)
                            acc[q] = fma(-a, sX[j][q], acc[q]);
                    }
            }
            an_part = fmax(an_part, rs);
            if(i < n)
            {
#pragma unroll
                for(int q = 0; q < NR; ++q)
                    if(c0 + q < nrhs)
                    {
                        R[i + int64_t(c0 + q) * ldr] = acc[q];
                        rn[q] = fmax(rn[q], fabs(acc[q]));
                        xn[q] = fmax(xn[q], fabs(X[i + int64_t(c0 + q) * ldx]));
                    }
            }
        }
        if(measure)
        {
            for(int off = 32; off > 0; off >>= 1)
            {
                an_part = fmax(an_part, __shfl_xor(an_part, off));
                amax = fmax(amax, __shfl_xor(amax, off));
            }
            __syncthreads();
            if(lane == 0)
            {
                sred[w][0] = an_part;
                sred[w][1] = amax;
            }
            __syncthreads();
            double an = sred[0][0], am = sred[0][1];
            for(int v = 1; v < NWV; ++v)
            {
                an = fmax(an, sred[v][0]);
                am = fmax(am, sred[v][1]);
            }
            if(am > double(FLT_MAX)) // uniform: FP32 copy of A overflowed
            {
                if(tid == 0)
                {
                    iter[b] = -2;
                    askip[b] = 1;
                    fskip[b] = 0;
                }
                return;
            }
            if(tid == 0)
            {
                s_an = an;
                anorm[b] = an;
            }
        }
        // column maxima over the workgroup
#pragma unroll
        for(int q = 0; q < NR; ++q)
            for(int off = 32; off > 0; off >>= 1)
            {
                rn[q] = fmax(rn[q], __shfl_xor(rn[q], off));
                xn[q] = fmax(xn[q], __shfl_xor(xn[q], off));
            }
        __syncthreads();
        if(lane == 0)
#pragma unroll
            for(int q = 0; q < NR; ++q)
            {
                sred[w][q] = rn[q];
                sred[w][NR + q] = xn[q];
            }
        __syncthreads();
        if(tid == 0)
        {
            const double an = s_an;
            const double cte = an * 0x1p-53 * sqrt(double(n));
            for(int q = 0; q < NR && c0 + q < nrhs; ++q)
            {
                double r = sred[0][q], x = sred[0][NR + q];
                for(int v = 1; v < NWV; ++v)
                {
                    r = fmax(r, sred[v][q]);
                    x = fmax(x, sred[v][NR + q]);
                }
                if(r > x * cte) // dsgesv.f: IF( RNRM.GT.XNRM*CTE ) not converged
                    s_ok = 0;
                const double d = an * x;
                worst = fmax(worst, d > 0 ? r / d : (r > 0 ? double(INFINITY) : 0.0));
            }
        }
    }
    __syncthreads();
    if(tid == 0)
    {
        if(s_ok)
        {
            iter[b] = it;
            askip[b] = 1;
            if(berr)
                berr[b] = worst;
        }
        else if(it >= itermax)
        {
            iter[b] = -(itermax + 1);
            askip[b] = 1;
            fskip[b] = 0;
        }
        else
            atomicAdd(&counters[it], 1);
    }
}

// after the FP64 fallback LU: no solve for systems whose FP64 LU is singular
__global__ void ir_mark_singular_kernel(const int* __restrict__ info, int* __restrict__ fskip,
                                        double* __restrict__ berr, int bc)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if(b < bc && fskip[b] == 0)
    {
        if(info[b] != 0)
            fskip[b] = 1;
        if(berr)
            berr[b] = -1.0; // refined error not measured for the fallback path
    }
}

size_t dsgesv_workspace_bytes(int n, int nrhs, int bc, int itermax)
{
    const size_t per = sizeof(double) * (size_t(n) * n + size_t(n) * nrhs + 1)
                       + sizeof(int) * (size_t(n) + 3);
    return per * size_t(bc) + sizeof(int) * (itermax + 1) + 256 * 8;
}

rocblas_status dsgesv_refine(hipStream_t stream, int n, int nrhs, Batch<double> A, int lda,
                             Batch<double> B, int ldb, Batch<double> X, int ldx, int* iter,
                             int* info, double* berr, int bc, int itermax, void* work,
                             bool early_exit)
{
    if(n <= SMALL_MAX || n > FUSED_MAX)
        return rocblas_status_not_implemented;
    // carve the workspace (all pieces 8-byte aligned)
    char* p = static_cast<char*>(work);
    auto take = [&](size_t bytes) {
        char* q = p;
        p += (bytes + 255) / 256 * 256;
        return q;
    };
    double* lu = reinterpret_cast<double*>(take(sizeof(double) * size_t(n) * n * bc));
    double* R = reinterpret_cast<double*>(take(sizeof(double) * size_t(n) * nrhs * bc));
    double* anorm = reinterpret_cast<double*>(take(sizeof(double) * bc));
    int* ipiv = reinterpret_cast<int*>(take(sizeof(int) * size_t(n) * bc));
    int* info32 = reinterpret_cast<int*>(take(sizeof(int) * bc));
    int* askip = reinterpret_cast<int*>(take(sizeof(int) * bc));
    int* fskip = reinterpret_cast<int*>(take(sizeof(int) * bc));
    int* counters = reinterpret_cast<int*>(take(sizeof(int) * (itermax + 1)));

    const int64_t slot = int64_t(n) * n; // LU slot per matrix, in doubles
    Batch<float> LU32 = strided(reinterpret_cast<float*>(lu), 2 * slot);
    Batch<double> LU64 = strided(lu, slot);
    Batch<double> Rb = strided(R, int64_t(n) * nrhs);
    rocblas_status st;
#define RSA_TRY(x)                        \
    do                                    \
    {                                     \
        st = (x);                         \
        if(st != rocblas_status_success)  \
            return st;                    \
    } while(0)

    RSA_TRY((lu_fused<double, float>(stream, n, A, lda, LU32, n, ipiv, n, info32, bc)));
    hipLaunchKernelGGL(ir_init_kernel, dim3(bc), dim3(IR_THREADS), 0, stream, n, nrhs, B, ldb,
                       info32, askip, fskip, iter, info, counters, itermax + 1);
    RSA_LAUNCH_CHECK();
    RSA_TRY((getrs_fused<float, double>(stream, false, n, nrhs, LU32, n, ipiv, n, B, ldb, X, ldx,
                                        bc, askip, false)));
    // one pinned counter per host thread, kept for the life of the process
    static thread_local int* host_count = nullptr;
    if(early_exit && !host_count && hipHostMalloc((void**)&host_count, sizeof(int)) != hipSuccess)
    {
        host_count = nullptr;
        return rocblas_status_memory_error;
    }
    for(int it = 0; it <= itermax; ++it)
    {
        if(nrhs <= 1)
            hipLaunchKernelGGL((ir_residual_kernel<1>), dim3(bc), dim3(IR_THREADS), 0, stream, n,
                               nrhs, A, lda, B, ldb, X, ldx, Rb, n, anorm, it, itermax, askip,
                               fskip, iter, berr, counters);
        else
            hipLaunchKernelGGL((ir_residual_kernel<8>), dim3(bc), dim3(IR_THREADS), 0, stream, n,
                               nrhs, A, lda, B, ldb, X, ldx, Rb, n, anorm, it, itermax, askip,
                               fskip, iter, berr, counters);
        RSA_LAUNCH_CHECK();
        if(it == itermax)
            break;
        if(early_exit)
        {
            if(hipMemcpyAsync(host_count, counters + it, sizeof(int), hipMemcpyDeviceToHost,
                              stream)
                   != hipSuccess
               || hipStreamSynchronize(stream) != hipSuccess)
                return rocblas_status_internal_error;
            if(*host_count == 0)
                break;
        }
        RSA_TRY((getrs_fused<float, double>(stream, false, n, nrhs, LU32, n, ipiv, n, Rb, n, X,
                                            ldx, bc, askip, true)));
    }
    // FP64 fallback for the matrices that did not converge
    RSA_TRY((lu_fused<double, double>(stream, n, A, lda, LU64, n, ipiv, n, info, bc, fskip)));
    hipLaunchKernelGGL(ir_mark_singular_kernel, dim3(ceil_div(bc, 256)), dim3(256), 0, stream,
                       info, fskip, berr, bc);
    RSA_LAUNCH_CHECK();
    RSA_TRY((getrs_fused<double, double>(stream, false, n, nrhs, LU64, n, ipiv, n, B, ldb, X, ldx,
                                         bc, fskip, false)));
#undef RSA_TRY
    return rocblas_status_success;
}

} // namespace rsa
