// refine.hip -- mixed-precision batched solve: FP32 LU on MFMA, FP64
// iterative refinement, FP64 fallback per matrix.
//
// Per matrix it follows LAPACK dsgesv (no rocSOLVER counterpart):
//   entries of B beyond the FP32 range -> ITER = -2, fallback
//   SA = fp32(A) (entries beyond the FP32 range -> ITER = -2, fallback)
//   SA = P L U in FP32 (zero pivot -> ITER = -3, fallback)
//   X  = fp64(U^-1 L^-1 P fp32(B))
//   repeat: R = B - A X in FP64; if every column has
//           ||R_c||_inf <= ||X_c||_inf * ||A||_inf * sqrt(n) * 2^-53: done (ITER = k)
//           fp32(R) out of range -> ITER = -2, fallback
//           X += fp64(U^-1 L^-1 P fp32(R))
//   after ITERMAX corrections: ITER = -(ITERMAX+1), fallback
//   fallback: X = A^-1 B with an FP64 LU of A (info = its first zero pivot).
//
// Launch sequence (fixed, no host synchronisation, so it can be captured in
// a graph): FP32 LU of the batch -> one refinement kernel that runs the whole
// loop of every system in one workgroup (solution in registers, A and the
// factors re-read from memory per step) -> FP64 LU and solve of the batch,
// skipping every system that converged.
//
// Half-width factorisations (no LAPACK counterpart): A is first equilibrated
// with power-of-two row and column scales R, C (exact in FP64), the FP32 LU
// of R A C is computed with FP16 or BF16 matrix-core operands in its trailing
// updates, and every correction solves (R A C) y = R r, x += C y; residuals
// and the stopping test use the original A in FP64.  FP32 overflow of A is
// not a failure there (the scaled matrix is what gets rounded), and the
// right-hand side of each FP32 solve is normalised by a power of two, so an
// FP32 out-of-range right-hand side cannot occur either.
//
// Right-hand sides are refined in groups of up to 8 columns (4 when n > 256).
// With nrhs within one group this is exactly dsgesv; with more, each group
// iterates until its own columns pass the test and ITER is the largest group
// count.
#include "internal.hpp"
#include "trsv_rows.hpp"

#include <cfloat>
#include <cmath>
#include <map>
#include <mutex>
#include <utility>

namespace rsa
{

__device__ __forceinline__ bool beyond_fp32(double v)
{
    // dlag2s.f: (A(I,J).LT.-RMAX) .OR. (A(I,J).GT.RMAX)
    return v < -double(FLT_MAX) || v > double(FLT_MAX);
}

// Power of two 2^-e with m * 2^-e in [1/2, 1); 1 for a zero or non-finite m.
// e is kept within [-1020, 1020] so the scale itself is a finite normal number
// (a subnormal m is then only brought up to 2^-1020 times its size).
__device__ __forceinline__ double pow2_scale(double m)
{
    if(!(m > 0.0) || m > DBL_MAX)
        return 1.0;
    int e;
    frexp(m, &e);
    e = e < -1020 ? -1020 : (e > 1020 ? 1020 : e);
    return ldexp(1.0, -e);
}

// Whole refinement of one system per workgroup (n <= NMAX threads, thread i
// owns row i).  fskip[b] = 1 when the system needs no FP64 fallback.
template <int NMAX, int NR>
__global__ __launch_bounds__(NMAX, NMAX <= 128 ? 3 : NMAX <= 256 ? 4 : 1) void ir_fused_kernel(
    int n, int nrhs, Batch<double> AA, int lda, Batch<float> LL, int ldl,
    const int* __restrict__ ipivA, int64_t strideP, const int* __restrict__ info32,
    Batch<double> BB, int ldb, Batch<double> XX, int ldx, int itermax, int* __restrict__ iter,
    int* __restrict__ info, int* __restrict__ fskip, double* __restrict__ berr,
    const double* __restrict__ rsA, const double* __restrict__ csA, int64_t strideS)
{
    constexpr int NW = NMAX / 64;
    // columns of A per load group in the residual (fewer at 512 rows x 8
    // columns, where the registers are short)
    constexpr int JU = 8;
    __shared__ double sX[NMAX][NR];
    __shared__ float sR[NMAX][NR];
    __shared__ float sx[2][GBS][NR];
    __shared__ float sL[GBS][GBS + 1];
    __shared__ int sipiv[NMAX];
    __shared__ int ssrc[NMAX];
    __shared__ double sred[NW][2 * NR + 4];
    __shared__ int sflag;
    __shared__ double snorm[NW];

    const int64_t b = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const double* A = AA[b];
    const MatrixBuffer<double> Ab(A, lda, n);
    const MatrixBuffer<float> LU(LL[b], ldl, n);
    const double* Bm = BB[b];
    const MatrixBuffer<double> Xbuf(XX[b], ldx, nrhs);
    const bool row = tid < n;
    const MatrixBuffer<double> Bbuf(Bm, ldb, nrhs);
    // B[t, c] (0 for rows past n or columns past nrhs), branch-free
    auto bload_at = [&](int t, int c) {
        return Bbuf.load((t < n && c < nrhs) ? Bbuf.row_offset(t) + Bbuf.col_offset(c)
                                             : MatrixBuffer<double>::OUT_OF_RANGE,
                         0);
    };
    // the same for this thread's row, the lane index re-read at each use so
    // the loads are not hoisted out of the refinement loop and held in
    // registers
    auto bload = [&](int c) { return bload_at(fresh_tid(), c); };
    const bool scaled = rsA != nullptr;
    // this row's equilibration factor and this solution component's (read
    // where used: fewer live registers)
    auto rsc = [&]() { return scaled ? rsA[b * strideS + (row ? tid : 0)] : 1.0; };
    auto csc = [&]() { return scaled ? csA[b * strideS + (row ? tid : 0)] : 1.0; };
    // the scales are powers of two: work with their exponents
    auto rexp = [&]() { return scaled ? ilogb(rsc()) : 0; };
    auto cexp = [&]() { return scaled ? ilogb(csc()) : 0; };
    constexpr double NO_EXP = -1.0e9; // exponent of a zero vector

    auto finish = [&](int code, bool ok, double err) {
        if(tid == 0)
        {
            iter[b] = code;
            info[b] = 0;
            fskip[b] = ok ? 1 : 0;
            if(berr)
                berr[b] = ok ? err : -1.0;
        }
    };
    // largest v over the workgroup, in every thread (two barriers)
    auto wg_max = [&](double v) {
        v = group_max<64>(v);
        if(lane == 0)
            snorm[w] = v;
        __syncthreads();
        double m = snorm[0];
        for(int k = 1; k < NW; ++k)
            m = fmax(m, snorm[k]);
        __syncthreads();
        return m;
    };
    // any thread's flag -> every thread (one barrier before, one after)
    auto any = [&](bool f) {
        if(tid == 0)
            sflag = 0;
        __syncthreads();
        if(f)
            sflag = 1;
        __syncthreads();
        const bool r = sflag != 0;
        __syncthreads();
        return r;
    };

    // ---- B and A range checks, FP32 LU status (dsgesv order: B, A, LU)
    bool big = false;
    if(!scaled)
    {
#pragma unroll 8
        for(int c = 0; c < nrhs; ++c)
            big = big || beyond_fp32(bload_at(tid, c));
    }
    if(any(big))
    {
        finish(-2, false, 0);
        return;
    }
    if(info32[b] != 0)
    {
        // only then is A scanned here; otherwise the first residual pass does it
        big = false;
        if(row && !scaled)
            for(int j = 0; j < n; ++j)
                if(beyond_fp32(A[tid + int64_t(j) * lda]))
                    big = true;
        finish(any(big) ? -2 : -3, false, 0);
        return;
    }

    const int pos = pivot_position(n, ipivA + b * strideP, false, sipiv, ssrc);
    (void)pos;
    const int src = ssrc[tid];
    // uniform: kept in scalar registers
    const double cte_scale = uniform_value(0x1p-53 * sqrt(double(n)));
    double anorm = 0;
    int it_all = 0;
    double worst = 0;

    for(int c0 = 0; c0 < nrhs; c0 += NR)
    {
        const bool measure_a = (c0 == 0);
        // B is re-read at every residual instead of held in registers
        auto bval = [&](int q) { return bload(c0 + q); };
        // the FP64 solution lives in LDS only (sX[row][q]); each thread updates its
        // own row, every thread reads all rows in the residual
        float d[NR];

        // correction solve of the FP32 right-hand side rv (row order) -> d
        auto solve = [&](const float (&rv)[NR]) {
#pragma unroll
            for(int q = 0; q < NR; ++q)
                sR[tid][q] = rv[q];
            __syncthreads();
#pragma unroll
            for(int q = 0; q < NR; ++q)
                d[q] = row ? sR[src][q] : 0.f;
            solve_lower_rows<float, NR, false, true>(n, LU, d, sx, sL);
            solve_upper_rows<float, NR, false, true>(n, LU, d, sx, sL);
        };

        {
            // equilibrated paths: the right-hand side R b of every FP32 solve is
            // normalised by 2^-E, E the largest exponent of its entries, so it
            // neither underflows nor overflows in FP32; the solve is linear, so
            // the correction is scaled back by 2^E.  Exponents are combined as
            // integers (ldexp), so no intermediate scale factor can overflow.
            const int ki = rexp();
            double e = NO_EXP;
            if(scaled)
            {
#pragma unroll
                for(int q = 0; q < NR; ++q)
                {
                    const double v = bval(q);
                    if(v != 0.0 && isfinite(v))
                        e = fmax(e, double(ilogb(v) + ki));
                }
                e = wg_max(e);
            }
            const int E = scaled ? int(e) : 0;
            float rv[NR];
#pragma unroll
            for(int q = 0; q < NR; ++q)
                rv[q] = float(ldexp(bval(q), ki - E));
            solve(rv);
            const int kj = cexp() + E;
#pragma unroll
            for(int q = 0; q < NR; ++q)
                sX[tid][q] = row ? ldexp(double(d[q]), kj) : 0.0;
        }

        int code = 0; // 0: running, 1: converged, <0: fallback code
        int it = 0;
        double err_c = 0;
        for(;; ++it)
        {
            // per-lane addresses recomputed in every step, not kept from entry
            const int tid = fresh_tid();
            const int lane = tid & 63, w = tid >> 6;
            const bool row = tid < n;
            // ---- R = B - A X (fma over the columns of A in order)
            __syncthreads(); // every row of sX written
            double acc[NR];
#pragma unroll
            for(int q = 0; q < NR; ++q)
                acc[q] = bval(q);
            const bool measure = measure_a && it == 0;
            double rs = 0, amax = 0;
            if(row)
            {
                const uint32_t off = Ab.row_offset(tid);
                double cur[JU], nxt[JU];
#pragma unroll
                for(int u = 0; u < JU; ++u)
                    cur[u] = u < n ? Ab.load(off, Ab.col_offset(u)) : 0.0;
                for(int j0 = 0; j0 < n; j0 += JU)
                {
#pragma unroll
                    for(int u = 0; u < JU; ++u)
                    {
                        const int j = j0 + JU + u;
                        nxt[u] = Ab.load(off, Ab.col_offset(j < n ? j : n - 1));
                    }
#pragma unroll
                    for(int u = 0; u < JU; ++u)
                    {
                        if(j0 + u < n) // uniform
                        {
                            const double a = cur[u];
                            if(measure)
                            {
                                rs += fabs(a);
                                amax = fmax(amax, fabs(a));
                            }
#pragma unroll
                            for(int q = 0; q < NR; ++q)
                                acc[q] = fma(-a, sX[j0 + u][q], acc[q]);
                        }
                    }
#pragma unroll
                    for(int u = 0; u < JU; ++u)
                        cur[u] = nxt[u];
                }
            }
            // ---- column maxima of |R| and |X| (and ||A||_inf, max |a| once)
            //      and the largest exponent of R_scaled r (normalisation of
            //      the next solve).  Per wavefront: one exchange-reduction of
            //      the 2 NR columns (lane l ends with column l / (32 / NR));
            //      across wavefronts: lane q combines column q through LDS.
            constexpr int K = 2 * NR;
            constexpr int LPC = 64 / K; // lanes per column after the exchanges
            double remax = NO_EXP;
            {
                const int ki = rexp();
                double v[K];
#pragma unroll
                for(int q = 0; q < NR; ++q)
                {
                    v[q] = row ? fabs(acc[q]) : 0.0;
                    v[NR + q] = fabs(sX[tid][q]);
                    if(row && c0 + q < nrhs && acc[q] != 0.0 && isfinite(acc[q]))
                        remax = fmax(remax, double(ilogb(acc[q]) + ki));
                }
                if constexpr(NR >= 2)
                {
                    // |R| and |X| reduced one after the other: half the
                    // registers of one 2 NR-column reduction
                    constexpr int L1 = 64 / NR;
                    double vr[NR], vx[NR];
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                        vr[q] = v[q];
                    const double cr = wave_colmax<NR>(vr);
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                        vx[q] = fabs(sX[tid][q]);
                    const double cx = wave_colmax<NR>(vx);
                    if(lane % L1 == 0)
                    {
                        sred[w][lane / L1] = cr;
                        sred[w][NR + lane / L1] = cx;
                    }
                }
                else
                {
                    const double cm = wave_colmax<K>(v);
                    if(lane % LPC == 0)
                        sred[w][lane / LPC] = cm;
                }
            }
            remax = group_max<64>(remax);
            if(measure)
            {
                rs = group_max<64>(rs);
                amax = group_max<64>(amax);
            }
            if(lane == 0)
            {
                sred[w][K] = remax;
                sred[w][K + 1] = rs;
                sred[w][K + 2] = amax;
            }
            __syncthreads();
            // lane q < NR: ||R_q||, ||X_q|| of column q; every lane: the rest
            double rnl = 0, xnl = 0;
            if(lane < NR)
                for(int v = 0; v < NW; ++v)
                {
                    rnl = fmax(rnl, sred[v][lane]);
                    xnl = fmax(xnl, sred[v][NR + lane]);
                }
            remax = sred[0][K];
            for(int v = 1; v < NW; ++v)
                remax = fmax(remax, sred[v][K]);
            if(measure)
            {
                double an = sred[0][K + 1], am = sred[0][K + 2];
                for(int v = 1; v < NW; ++v)
                {
                    an = fmax(an, sred[v][K + 1]);
                    am = fmax(am, sred[v][K + 2]);
                }
                anorm = an;
                if(!scaled && am > double(FLT_MAX)) // the FP32 copy of A overflowed
                    code = -2;
            }
            __syncthreads(); // sred is rewritten by the next step
            if(code != 0)
                break;
            const double cte = anorm * cte_scale;
            const bool qv = lane < NR && c0 + lane < nrhs;
            // dsgesv.f: IF( RNRM.GT.XNRM*CTE ) not converged
            const bool conv = __builtin_amdgcn_ballot_w64(qv && rnl > xnl * cte) == 0;
            // dlag2s.f on the next right-hand side (normalised on the scaled paths)
            const bool rbig = !scaled && __builtin_amdgcn_ballot_w64(qv && beyond_fp32(rnl)) != 0;
            if(conv)
            {
                const double dn = anorm * xnl;
                err_c = group_max<64>(
                    qv ? (dn > 0 ? rnl / dn : (rnl > 0 ? double(INFINITY) : 0.0)) : 0.0);
            }
            if(conv)
            {
                code = 1;
                break;
            }
            if(it >= itermax)
            {
                code = -(itermax + 1);
                break;
            }
            if(rbig)
            {
                code = -2;
                break;
            }
            // ---- X += fp64(U^-1 L^-1 P fp32(R))
            float rv[NR];
            const int E = scaled ? int(remax) : 0;
            const int ki2 = rexp() - E;
#pragma unroll
            for(int q = 0; q < NR; ++q)
                rv[q] = float(ldexp(acc[q], ki2));
            solve(rv);
            const int kj = cexp() + E;
#pragma unroll
            for(int q = 0; q < NR; ++q)
                if(row)
                    sX[tid][q] += ldexp(double(d[q]), kj);
        }
        if(code < 0)
        {
            finish(code, false, 0);
            return;
        }
        it_all = it > it_all ? it : it_all;
        worst = fmax(worst, err_c);
        {
            // running offset (one register, out of range for rows past n)
            uint32_t off = Xbuf.offset(row, tid, c0);
#pragma unroll
            for(int q = 0; q < NR; ++q)
            {
                if(c0 + q < nrhs) // uniform
                    Xbuf.store(sX[tid][q], off, 0);
                off += Xbuf.ld_bytes;
                opaque(off);
            }
        }
    }
    finish(it_all, true, worst);
}

// Row scales rs_i (largest |a_ij| of row i), then column scales cs_j (largest
// |rs_i a_ij| of column j), powers of two (the radix-power choice of LAPACK
// dgeequb, so the scaled entries are exact).  One workgroup per matrix.
template <int NT>
__global__ __launch_bounds__(NT) void equilibrate_kernel(int n, Batch<double> AA, int lda,
                                                         double* __restrict__ rs,
                                                         double* __restrict__ cs,
                                                         int64_t strideS)
{
    __shared__ double sr[FUSED_MAX];
    const int64_t b = blockIdx.x;
    const int tid = threadIdx.x;
    const double* A = AA[b];
    for(int i = tid; i < n; i += NT)
    {
        double m = 0;
        for(int j = 0; j < n; ++j)
            m = fmax(m, fabs(A[i + int64_t(j) * lda]));
        sr[i] = pow2_scale(m);
        rs[b * strideS + i] = sr[i];
    }
    __syncthreads();
    for(int j = tid; j < n; j += NT)
    {
        double m = 0;
        for(int i = 0; i < n; ++i)
            m = fmax(m, fabs(A[i + int64_t(j) * lda]) * sr[i]);
        cs[b * strideS + j] = pow2_scale(m);
    }
}

// after the FP64 fallback LU: no solve for systems whose FP64 LU is singular
__global__ void ir_mark_singular_kernel(const int* __restrict__ info, int* __restrict__ fskip,
                                        int bc)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if(b < bc && fskip[b] == 0 && info[b] != 0)
        fskip[b] = 1;
}

size_t dsgesv_workspace_bytes(int n, int nrhs, int bc, int itermax)
{
    (void)nrhs;
    (void)itermax;
    const size_t per = sizeof(double) * (size_t(n) * n + 2 * size_t(n) + size_t(lu_fused_inv_elems(n)))
                       + sizeof(int) * (size_t(n) + 2);
    return per * size_t(bc) + 7 * 256;
}

template <int NMAX, int NR>
static void launch_ir(hipStream_t stream, int n, int nrhs, Batch<double> A, int lda,
                      Batch<float> LU, const int* ipiv, const int* info32, Batch<double> B,
                      int ldb, Batch<double> X, int ldx, int itermax, int* iter, int* info,
                      int* fskip, double* berr, int bc, const double* rs, const double* cs)
{
    hipLaunchKernelGGL((ir_fused_kernel<NMAX, NR>), dim3(bc), dim3(NMAX), 0, stream, n, nrhs, A,
                       lda, LU, n, ipiv, int64_t(n), info32, B, ldb, X, ldx, itermax, iter, info,
                       fskip, berr, rs, cs, int64_t(n));
}

// Batches from this size on are solved in two halves on two streams.
constexpr int SPLIT_MIN = 256;

// One internal stream per caller stream (default priority, non-blocking),
// created on first use and kept for the life of the process; nullptr if it
// cannot be created (the solve then runs on the call stream alone).  Keyed
// by the caller's stream, so calls on different streams -- different
// handles, threads, or a graph being captured on one of them -- never share
// an internal stream; calls on one stream are ordered by that stream anyway.
static hipStream_t aux_stream(hipStream_t caller)
{
    static std::mutex m;
    static std::map<std::pair<int, hipStream_t>, hipStream_t> streams;
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess)
        return nullptr;
    std::lock_guard<std::mutex> lock(m);
    const auto key = std::make_pair(dev, caller);
    auto it = streams.find(key);
    if(it != streams.end())
        return it->second;
    hipStream_t s = nullptr;
    if(hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        s = nullptr;
    streams[key] = s;
    return s;
}

rocblas_status dsgesv_refine(hipStream_t stream, int factor, int n, int nrhs, Batch<double> A,
                             int lda, Batch<double> B, int ldb, Batch<double> X, int ldx,
                             int* iter, int* info, double* berr, int bc, int itermax, void* work)
{
    if(factor < 0 || factor > 2)
        return rocblas_status_invalid_value;
    if(n > FUSED_MAX)
        return rocblas_status_not_implemented;
    if(bc <= 0 || nrhs <= 0)
        return rocblas_status_success;
    // carve the workspace (pieces 256-byte aligned)
    char* p = static_cast<char*>(work);
    auto take = [&](size_t bytes) {
        char* q = p;
        p += (bytes + 255) / 256 * 256;
        return q;
    };
    double* lu = reinterpret_cast<double*>(take(sizeof(double) * size_t(n) * n * bc));
    int* ipiv = reinterpret_cast<int*>(take(sizeof(int) * size_t(n) * bc));
    int* info32 = reinterpret_cast<int*>(take(sizeof(int) * bc));
    int* fskip = reinterpret_cast<int*>(take(sizeof(int) * bc));
    double* rs = reinterpret_cast<double*>(take(sizeof(double) * size_t(n) * bc));
    double* cs = reinterpret_cast<double*>(take(sizeof(double) * size_t(n) * bc));
    // inverted diagonal blocks of L for the FP64 fallback LU (the FP32 LU
    // returns its diagonal blocks inverted)
    double* winv64
        = reinterpret_cast<double*>(take(sizeof(double) * size_t(lu_fused_inv_elems(n)) * bc));
    if(factor == 0)
        rs = cs = nullptr;

    const int64_t slot = int64_t(n) * n; // LU slot per matrix, in doubles
    Batch<float> LU32 = strided(reinterpret_cast<float*>(lu), 2 * slot);
    Batch<double> LU64 = strided(lu, slot);
    Batch<double> WI64 = strided(winv64, lu_fused_inv_elems(n));
    rocblas_status st;
#define RSA_TRY(x)                        \
    do                                    \
    {                                     \
        st = (x);                         \
        if(st != rocblas_status_success)  \
            return st;                    \
    } while(0)

    // the per-matrix stages on the systems [k0, k0 + cnt): FP32 (or
    // half-width) LU, diagonal-block inversion, refinement
    auto sub = [&](auto batch, int64_t k) {
        using BT = decltype(batch);
        return batch.ptrs ? BT{nullptr, batch.ptrs + k, 0}
                          : BT{batch.base + k * batch.stride, nullptr, batch.stride};
    };
    auto factor_range = [&](hipStream_t st_, int64_t k0, int cnt) -> rocblas_status {
        if(factor == 0)
            return lu_fused_invdiag(st_, n, sub(A, k0), lda, sub(LU32, k0), n, ipiv + k0 * n, n,
                                    info32 + k0, cnt);
        return lu_fused_lowp(st_, factor, n, sub(A, k0), lda, sub(LU32, k0), n, ipiv + k0 * n, n,
                             info32 + k0, cnt, rs + k0 * n, cs + k0 * n, int64_t(n));
    };
    auto refine_range = [&](hipStream_t st_, int64_t k0, int cnt) -> rocblas_status {
        const double* rs_ = rs ? rs + k0 * n : nullptr;
        const double* cs_ = cs ? cs + k0 * n : nullptr;
        double* berr_ = berr ? berr + k0 : nullptr;
        auto by_nr = [&](auto nmax) {
            constexpr int NMAX = decltype(nmax)::value;
            auto go = [&](auto nr) {
                constexpr int NR = decltype(nr)::value;
                launch_ir<NMAX, NR>(st_, n, nrhs, sub(A, k0), lda, sub(LU32, k0), ipiv + k0 * n,
                                    info32 + k0, sub(B, k0), ldb, sub(X, k0), ldx, itermax,
                                    iter + k0, info + k0, fskip + k0, berr_, cnt, rs_, cs_);
            };
            if(nrhs <= 1)
                go(std::integral_constant<int, 1>{});
            else if constexpr(NMAX >= 512) // groups of 4 beyond 256 rows (8 would spill)
                go(std::integral_constant<int, 4>{});
            else if(nrhs <= 4)
                go(std::integral_constant<int, 4>{});
            else
                go(std::integral_constant<int, 8>{});
        };
        if(n <= 128)
            by_nr(std::integral_constant<int, 128>{});
        else if(n <= 256)
            by_nr(std::integral_constant<int, 256>{});
        else
            by_nr(std::integral_constant<int, 512>{});
        RSA_LAUNCH_CHECK();
        return rocblas_status_success;
    };

    if(factor != 0)
    {
        if(n <= 128)
            hipLaunchKernelGGL((equilibrate_kernel<128>), dim3(bc), dim3(128), 0, stream, n, A,
                               lda, rs, cs, int64_t(n));
        else
            hipLaunchKernelGGL((equilibrate_kernel<256>), dim3(bc), dim3(256), 0, stream, n, A,
                               lda, rs, cs, int64_t(n));
        RSA_LAUNCH_CHECK();
    }
    hipStream_t aux = bc >= SPLIT_MIN ? aux_stream(stream) : nullptr;
    if(aux)
    {
        // two halves at once, the second on an internal stream: each half's
        // latency-bound LU and refinement fill the issue slots the other
        // leaves idle; event fork and join keep the call stream-ordered and
        // capturable in a graph
        const int h = bc / 2;
        hipEvent_t forked, joined;
        if(hipEventCreateWithFlags(&forked, hipEventDisableTiming) != hipSuccess)
            return rocblas_status_internal_error;
        if(hipEventCreateWithFlags(&joined, hipEventDisableTiming) != hipSuccess)
        {
            (void)hipEventDestroy(forked);
            return rocblas_status_internal_error;
        }
        bool ok = hipEventRecord(forked, stream) == hipSuccess
                  && hipStreamWaitEvent(aux, forked, 0) == hipSuccess;
        rocblas_status st1 = factor_range(stream, 0, h);
        rocblas_status st2 = ok ? factor_range(aux, h, bc - h) : rocblas_status_internal_error;
        if(st2 == rocblas_status_success)
            st2 = refine_range(aux, h, bc - h);
        // the join is recorded whatever happened on the internal stream, so
        // the call stream never waits on an unrecorded event
        ok = hipEventRecord(joined, aux) == hipSuccess;
        if(st1 == rocblas_status_success)
            st1 = refine_range(stream, 0, h);
        ok = ok && hipStreamWaitEvent(stream, joined, 0) == hipSuccess;
        (void)hipEventDestroy(forked);
        (void)hipEventDestroy(joined);
        if(st1 != rocblas_status_success)
            return st1;
        if(st2 != rocblas_status_success)
            return st2;
        if(!ok)
            return rocblas_status_internal_error;
    }
    else
    {
        RSA_TRY(factor_range(stream, 0, bc));
        RSA_TRY(refine_range(stream, 0, bc));
    }
    // FP64 fallback for the systems that did not converge
    RSA_TRY((lu_fused<double, double>(stream, n, A, lda, LU64, n, ipiv, n, info, bc, WI64,
                                      fskip)));
    hipLaunchKernelGGL(ir_mark_singular_kernel, dim3(ceil_div(bc, 256)), dim3(256), 0, stream,
                       info, fskip, bc);
    RSA_LAUNCH_CHECK();
    RSA_TRY((getrs_fused<double, double>(stream, false, n, nrhs, LU64, n, ipiv, n, B, ldb, X, ldx,
                                         bc, fskip, false)));
#undef RSA_TRY
    return rocblas_status_success;
}

} // namespace rsa
