// lu_small.hpp -- register-resident LU / solve for matrices with m, n <= 64.
//
// Layout: a "team" of TS lanes (TS = 16, 32 or 64; TS >= m) owns one matrix;
// a 64-lane wavefront carries 64/TS matrices.  Lane r of a team holds the
// original row r in NC registers (NC >= n, padded with zeros), so the whole
// matrix is read once and written once, coalesced per column.
//
// Pivoting follows the reference small kernel
// (library/src/specialized/roclapack_getf2_specialized_kernels.hpp:72-126):
//   * pivot = first row of max |a| in the column (strict comparison, ties to
//     the lowest row, NaN never displaces the candidate) -- here a DPP /
//     permlane arg-max over the team with (key, logical row) pairs;
//   * rows are swapped lazily: lanes never move data, they exchange their
//     logical row labels, and each lane writes its row to its final label;
//   * zero pivot -> info = first k+1, no swap, no scaling;
//   * the column is scaled by the reciprocal of the pivot, then a rank-1
//     update a -= l*u, one fma per element.
// The pivot row is broadcast to the team through LDS (one write by the pivot
// lane, one broadcast read per element).  Because the operation order equals
// the unblocked LAPACK-style algorithm, results are bit-identical to it.
//
// Every operation is a team-level device function (team_getrf, team_gesv,
// team_getrs) that synchronises only within its wavefront, so a team's order
// n may differ from the other teams' (the variable-size batched kernel packs
// matrices of different orders into one wave); the kernels below run one
// batch of equal orders, one wavefront per workgroup.
#pragma once

#include "internal.hpp"

#include <type_traits>

namespace rsa
{

// One team factors one matrix held in registers r[NC].  On return `label`
// is this lane's logical row, `mypiv` its ipiv entry (1-based, valid when
// label < min(m,n)) and `zero_at` the first zero pivot (1-based) or 0.
// prow: NC LDS elements for this team (16-byte aligned).
template <typename T, int NC, int TS, bool PIVOT>
__device__ __forceinline__ void team_lu(T (&r)[NC], T* prow, int m, int n, bool live,
                                        int& label, int& mypiv, int& zero_at)
{
    const int kmax = m < n ? m : n;
#pragma unroll
    for(int k = 0; k < NC; ++k)
    {
        if(k < kmax) // uniform over the team
        {
            int plab = k;
            if(PIVOT)
            {
                T key = (live && label >= k) ? pivot_key(r[k], label == k) : T(-1);
                int lab = label;
                group_argmax<TS>(key, lab);
                plab = lab;
            }
            const bool is_piv = (label == plab);
            if(is_piv)
            {
                // whole 16-byte groups from the one holding column k (an
                // aligned ds_write_b128 each; entries left of k are unused)
                constexpr int G = 16 / int(sizeof(T));
#pragma unroll
                for(int j = 0; j < NC; ++j)
                    if(j / G >= k / G)
                        prow[j] = r[j];
            }
            wave_sync();
            const T pv = prow[k];
            const bool nz = (pv != T(0));
            if(!nz && zero_at == 0)
                zero_at = k + 1;
            if(is_piv)
            {
                mypiv = plab + 1;
                label = k;
            }
            else if(label == k)
                label = plab;
            if(live && label > k)
            {
                // a zero pivot: with pivoting every candidate is zero and the
                // column stays; without, the reference npvt small kernel
                // multiplies the column by its scale of 0
                // (roclapack_getf2_specialized_kernels.hpp:185-201)
                if(nz)
                    r[k] *= T(1) / pv;
                else if(!PIVOT)
                    r[k] *= T(0);
                const T l = r[k];
#pragma unroll
                for(int j = k + 1; j < NC; ++j)
                    r[j] = fma_t(-l, prow[j], r[j]);
            }
            wave_sync();
        }
    }
}

// Element (i, j) of a team's column-major matrix, through a plain global
// pointer (any batch layout) ...
template <typename T>
struct PtrAccess
{
    T* a;
    int ld;
    __device__ __forceinline__ T load(bool live, int i, int j) const
    {
        return live ? a[i + int64_t(j) * ld] : T(0);
    }
    __device__ __forceinline__ void store(int i, int j, T v) const
    {
        a[i + int64_t(j) * ld] = v;
    }
};

// ... or through one buffer resource for the whole wavefront (strided
// batches whose matrices of one wavefront span less than 2 GiB): the team's
// matrix starts team_off bytes into the buffer, the row offset is one vector
// register and the column offset a scalar, so a column costs one load or
// store instruction and no address arithmetic.
template <typename T>
struct BufAccess
{
    MatrixBuffer<T> buf;
    uint32_t team_off;
    __device__ __forceinline__ T load(bool live, int i, int j) const
    {
        return buf.load(live ? team_off + uint32_t(i) * uint32_t(sizeof(T))
                             : MatrixBuffer<T>::OUT_OF_RANGE,
                        buf.col_offset(j));
    }
    __device__ __forceinline__ void store(int i, int j, T v) const
    {
        buf.store(v, team_off + uint32_t(i) * uint32_t(sizeof(T)), buf.col_offset(j));
    }
};

// getrf of one m x n matrix by the team (row = lane index within the team,
// valid = the team has a matrix).  prow: LDS, NC elements for this team.
// Returns the first zero pivot (1-based) or 0, in every lane of the team.
template <typename T, int NC, int TS, bool PIVOT, typename Acc>
__device__ __forceinline__ int team_getrf(int row, bool valid, int m, int n, const Acc& Am,
                                          int* ipivb, T* prow)
{
    const bool live = valid && row < m;
    T r[NC];
#pragma unroll
    for(int j = 0; j < NC; ++j)
        r[j] = (j < n) ? Am.load(live, row, j) : T(0);

    int label = row, mypiv = row + 1, zero_at = 0;
    team_lu<T, NC, TS, PIVOT>(r, prow, m, n, live, label, mypiv, zero_at);

    const int kmax = m < n ? m : n;
    if(live)
    {
#pragma unroll
        for(int j = 0; j < NC; ++j)
            if(j < n)
                Am.store(label, j, r[j]);
        if(PIVOT && label < kmax)
            ipivb[label] = mypiv;
    }
    return zero_at;
}

template <typename T, int NC, int TS, bool PIVOT>
__device__ __forceinline__ int team_getrf(int row, bool valid, int m, int n, T* Ab, int lda,
                                          int* ipivb, T* prow)
{
    return team_getrf<T, NC, TS, PIVOT>(row, valid, m, n, PtrAccess<T>{Ab, lda}, ipivb, prow);
}

// Fused gesv of one n x n system by the team: factor in registers, then
// solve for NR right-hand sides at a time without ever materialising P*B:
// lane r holds original row r of B, and its label is that row's position in
// P*B.  Forward (unit L) and backward (non-unit U, true division)
// substitution run in logical order, the solved value broadcast through LDS
// -- the column-oriented order of the reference substitution kernels
// (roclapack_trsm_specialized_kernels.hpp:166-566).  A singular system
// leaves B untouched (roclapack_gesv.hpp:162-175).  prow: NC, pb: NR LDS
// elements for this team.  Returns the first zero pivot or 0.
template <typename T, int NC, int TS, int NR>
__device__ __forceinline__ int team_gesv(int row, bool valid, int n, int nrhs, T* Ab, int lda,
                                         int* ipivb, T* Bb, int ldb, T* prow, T* pb)
{
    const bool live = valid && row < n;
    T r[NC];
#pragma unroll
    for(int j = 0; j < NC; ++j)
        r[j] = (live && j < n) ? Ab[row + int64_t(j) * lda] : T(0);

    int label = row, mypiv = row + 1, zero_at = 0;
    team_lu<T, NC, TS, true>(r, prow, n, n, live, label, mypiv, zero_at);

    if(live)
    {
#pragma unroll
        for(int j = 0; j < NC; ++j)
            if(j < n)
                Ab[label + int64_t(j) * lda] = r[j];
        ipivb[label] = mypiv;
    }

    const bool solve = live && zero_at == 0;
    for(int c0 = 0; c0 < nrhs; c0 += NR)
    {
        T x[NR];
#pragma unroll
        for(int q = 0; q < NR; ++q)
            x[q] = (solve && c0 + q < nrhs) ? Bb[row + int64_t(c0 + q) * ldb] : T(0);
        // L y = P b
#pragma unroll
        for(int k = 0; k < NC - 1; ++k)
        {
            if(k < n - 1)
            {
                if(label == k)
                {
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                        pb[q] = x[q];
                }
                wave_sync();
                if(label > k)
                {
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                        x[q] = fma_t(-r[k], pb[q], x[q]);
                }
                wave_sync();
            }
        }
        // U x = y
#pragma unroll
        for(int k = NC - 1; k >= 0; --k)
        {
            if(k < n)
            {
                if(label == k)
                {
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                    {
                        x[q] = x[q] / r[k];
                        pb[q] = x[q];
                    }
                }
                wave_sync();
                if(label < k)
                {
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                        x[q] = fma_t(-r[k], pb[q], x[q]);
                }
                wave_sync();
            }
        }
        if(solve)
        {
#pragma unroll
            for(int q = 0; q < NR; ++q)
                if(c0 + q < nrhs)
                    Bb[label + int64_t(c0 + q) * ldb] = x[q];
        }
    }
    return zero_at;
}

// getrs of one system by the team.  Lane r holds row r of the LU (no
// transpose) or column r (transpose), so every substitution step uses only
// its own registers plus one LDS broadcast.  Row interchanges are composed
// into one permutation per team (the same swaps, in the same order, as the
// reference laswp: rocauxiliary_laswp.hpp:41-62) and applied as a gather on
// load (trans = N) or a scatter on store (trans = T).  pb: NR, perm and spiv:
// TS LDS elements for this team.
template <typename T, int NC, int TS, int NR>
__device__ __forceinline__ void team_getrs(int row, bool valid, bool trans, int n, int nrhs,
                                           const T* Ab, int lda, const int* ipivb, T* Bb, int ldb,
                                           bool pivot, T* pb, int* perm, int* spiv)
{
    const bool live = valid && row < n;
    T r[NC];
    if(!trans)
    {
#pragma unroll
        for(int j = 0; j < NC; ++j)
            r[j] = (live && j < n) ? Ab[row + int64_t(j) * lda] : T(0);
    }
    else
    {
#pragma unroll
        for(int j = 0; j < NC; ++j)
            r[j] = (live && j < n) ? Ab[j + int64_t(row) * lda] : T(0);
    }

    // position of this lane's row after the interchanges (a pivot outside
    // [0, n) is ignored instead of addressed)
    int pv_in = (pivot && live) ? ipivb[row] - 1 : row;
    if(pv_in < 0 || pv_in >= n)
        pv_in = row;
    spiv[row] = pv_in;
    wave_sync();
    int pos = row;
    if(pivot)
    {
        if(!trans)
        {
            for(int k = 0; k < n; ++k)
            {
                const int p = spiv[k];
                if(pos == k)
                    pos = p;
                else if(pos == p)
                    pos = k;
            }
        }
        else
        {
            for(int k = n - 1; k >= 0; --k)
            {
                const int p = spiv[k];
                if(pos == k)
                    pos = p;
                else if(pos == p)
                    pos = k;
            }
        }
    }
    perm[pos] = row; // row that lands on position pos
    wave_sync();
    const int src = (!trans && live) ? perm[row] : row;

    for(int c0 = 0; c0 < nrhs; c0 += NR)
    {
        T x[NR];
#pragma unroll
        for(int q = 0; q < NR; ++q)
            x[q] = (live && c0 + q < nrhs) ? Bb[src + int64_t(c0 + q) * ldb] : T(0);
        if(!trans)
        {
#pragma unroll
            for(int k = 0; k < NC - 1; ++k)
            {
                if(k < n - 1)
                {
                    if(row == k)
                    {
#pragma unroll
                        for(int q = 0; q < NR; ++q)
                            pb[q] = x[q];
                    }
                    wave_sync();
                    if(row > k)
                    {
#pragma unroll
                        for(int q = 0; q < NR; ++q)
                            x[q] = fma_t(-r[k], pb[q], x[q]);
                    }
                    wave_sync();
                }
            }
#pragma unroll
            for(int k = NC - 1; k >= 0; --k)
            {
                if(k < n)
                {
                    if(row == k)
                    {
#pragma unroll
                        for(int q = 0; q < NR; ++q)
                        {
                            x[q] = x[q] / r[k];
                            pb[q] = x[q];
                        }
                    }
                    wave_sync();
                    if(row < k)
                    {
#pragma unroll
                        for(int q = 0; q < NR; ++q)
                            x[q] = fma_t(-r[k], pb[q], x[q]);
                    }
                    wave_sync();
                }
            }
            // every lane of the team has read its right-hand side (the gather
            // above) before any row is stored
            if(live)
            {
#pragma unroll
                for(int q = 0; q < NR; ++q)
                    if(c0 + q < nrhs)
                        Bb[row + int64_t(c0 + q) * ldb] = x[q];
            }
        }
        else
        {
            // U^T y = b: lane r holds column r, r[k] = U[k][r] for k <= r
#pragma unroll
            for(int k = 0; k < NC; ++k)
            {
                if(k < n)
                {
                    if(row == k)
                    {
#pragma unroll
                        for(int q = 0; q < NR; ++q)
                        {
                            x[q] = x[q] / r[k];
                            pb[q] = x[q];
                        }
                    }
                    wave_sync();
                    if(row > k)
                    {
#pragma unroll
                        for(int q = 0; q < NR; ++q)
                            x[q] = fma_t(-r[k], pb[q], x[q]);
                    }
                    wave_sync();
                }
            }
            // L^T x = y: r[k] = L[k][r] for k > r
#pragma unroll
            for(int k = NC - 1; k > 0; --k)
            {
                if(k < n)
                {
                    if(row == k)
                    {
#pragma unroll
                        for(int q = 0; q < NR; ++q)
                            pb[q] = x[q];
                    }
                    wave_sync();
                    if(row < k)
                    {
#pragma unroll
                        for(int q = 0; q < NR; ++q)
                            x[q] = fma_t(-r[k], pb[q], x[q]);
                    }
                    wave_sync();
                }
            }
            if(live)
            {
#pragma unroll
                for(int q = 0; q < NR; ++q)
                    if(c0 + q < nrhs)
                        Bb[pos + int64_t(c0 + q) * ldb] = x[q];
            }
        }
    }
}

// ------------------------------------------------------------- kernels
// One wavefront per workgroup, 64/TS matrices of the same order per wave.

// BUF: strided batch read and written through one buffer resource per
// wavefront (the host checks that the wavefront's matrices span < 2 GiB).
template <typename T, int NC, int TS, bool PIVOT, bool BUF>
__global__ __launch_bounds__(64) void getrf_small_kernel(int m, int n, Batch<T> A, int lda,
                                                         int* __restrict__ ipiv, int64_t strideP,
                                                         int* __restrict__ info, int bc)
{
    constexpr int TPW = 64 / TS;
    __shared__ __align__(16) T prow[TPW][NC];
    const int lane = threadIdx.x;
    const int team = lane / TS;
    const int row = lane % TS;
    const int64_t b0 = int64_t(blockIdx.x) * TPW;
    const int64_t b = b0 + team;
    const bool valid = b < bc;
    int zero_at;
    if constexpr(BUF)
    {
        const uint64_t span
            = (uint64_t(TPW - 1) * uint64_t(A.stride) + uint64_t(lda) * uint64_t(n)) * sizeof(T);
        const BufAccess<T> Am{MatrixBuffer<T>(A[b0], lda, span),
                              uint32_t(team) * uint32_t(A.stride) * uint32_t(sizeof(T))};
        zero_at = team_getrf<T, NC, TS, PIVOT>(row, valid, m, n, Am,
                                               ipiv + (valid ? b : 0) * strideP, prow[team]);
    }
    else
        zero_at = team_getrf<T, NC, TS, PIVOT>(row, valid, m, n, valid ? A[b] : nullptr, lda,
                                               ipiv + (valid ? b : 0) * strideP, prow[team]);
    if(valid && row == 0)
        info[b] = zero_at;
}

template <typename T, int NC, int TS, int NR>
__global__ __launch_bounds__(64) void gesv_small_kernel(int n, int nrhs, Batch<T> A, int lda,
                                                        int* __restrict__ ipiv, int64_t strideP,
                                                        Batch<T> B, int ldb,
                                                        int* __restrict__ info, int bc)
{
    constexpr int TPW = 64 / TS;
    __shared__ __align__(16) T prow[TPW][NC];
    __shared__ T pb[TPW][NR];
    const int lane = threadIdx.x;
    const int team = lane / TS;
    const int row = lane % TS;
    const int64_t b = int64_t(blockIdx.x) * TPW + team;
    const bool valid = b < bc;
    const int zero_at = team_gesv<T, NC, TS, NR>(
        row, valid, n, nrhs, valid ? A[b] : nullptr, lda, ipiv + (valid ? b : 0) * strideP,
        valid ? B[b] : nullptr, ldb, prow[team], pb[team]);
    if(valid && row == 0)
        info[b] = zero_at;
}

template <typename T, int NC, int TS, int NR>
__global__ __launch_bounds__(64) void getrs_small_kernel(bool trans, int n, int nrhs, Batch<T> A,
                                                         int lda, const int* __restrict__ ipiv,
                                                         int64_t strideP, Batch<T> B, int ldb,
                                                         int bc, bool pivot)
{
    constexpr int TPW = 64 / TS;
    __shared__ T pb[TPW][NR];
    __shared__ int perm[TPW][TS];
    __shared__ int spiv[TPW][TS];
    const int lane = threadIdx.x;
    const int team = lane / TS;
    const int row = lane % TS;
    const int64_t b = int64_t(blockIdx.x) * TPW + team;
    const bool valid = b < bc;
    team_getrs<T, NC, TS, NR>(row, valid, trans, n, nrhs, valid ? A[b] : nullptr, lda,
                              ipiv + (valid ? b : 0) * strideP, valid ? B[b] : nullptr, ldb, pivot,
                              pb[team], perm[team], spiv[team]);
}

// ------------------------------------------------------------- dispatch
// NC (padded column count) and TS (team size) are compile-time so the whole
// matrix lives in registers with static indexing.  getrf picks TS from m and
// NC from n independently; gesv/getrs are square, so (TS, NC) follow n.
// Each translation unit instantiates one kernel family for one type.

template <int V>
using ic = std::integral_constant<int, V>;

template <typename F>
void team_size_switch(int m, F&& f)
{
    if(m <= 16)
        f(ic<16>{});
    else if(m <= 32)
        f(ic<32>{});
    else
        f(ic<64>{});
}

template <typename F>
void ncols_switch(int n, F&& f)
{
    if(n <= 8)
        f(ic<8>{});
    else if(n <= 16)
        f(ic<16>{});
    else if(n <= 32)
        f(ic<32>{});
    else
        f(ic<64>{});
}

// (TS, NC) pairs for square problems of order n
template <typename F>
void square_switch(int n, F&& f)
{
    if(n <= 8)
        f(ic<16>{}, ic<8>{});
    else if(n <= 16)
        f(ic<16>{}, ic<16>{});
    else if(n <= 24)
        f(ic<32>{}, ic<24>{});
    else if(n <= 32)
        f(ic<32>{}, ic<32>{});
    else if(n <= 48)
        f(ic<64>{}, ic<48>{});
    else
        f(ic<64>{}, ic<64>{});
}

template <typename F>
void nrhs_switch(int nrhs, F&& f)
{
    if(nrhs <= 1)
        f(ic<1>{});
    else
        f(ic<8>{});
}

template <typename T>
rocblas_status getrf_small_launch(hipStream_t stream, int m, int n, Batch<T> A, int lda,
                                  int* ipiv, int64_t strideP, int* info, int bc, bool pivot)
{
    team_size_switch(m, [&](auto ts) {
        ncols_switch(n, [&](auto nc) {
            constexpr int TS = decltype(ts)::value;
            constexpr int NC = decltype(nc)::value;
            dim3 grid(ceil_div(bc, 64 / TS));
            // one buffer per wavefront when its matrices span < 2 GiB
            const bool buf = A.ptrs == nullptr && A.stride >= 0
                             && (uint64_t(64 / TS - 1) * uint64_t(A.stride)
                                 + uint64_t(lda) * uint64_t(n))
                                        * sizeof(T)
                                    < (uint64_t(1) << 31);
            if(pivot && buf)
                hipLaunchKernelGGL((getrf_small_kernel<T, NC, TS, true, true>), grid, dim3(64), 0,
                                   stream, m, n, A, lda, ipiv, strideP, info, bc);
            else if(pivot)
                hipLaunchKernelGGL((getrf_small_kernel<T, NC, TS, true, false>), grid, dim3(64),
                                   0, stream, m, n, A, lda, ipiv, strideP, info, bc);
            else if(buf)
                hipLaunchKernelGGL((getrf_small_kernel<T, NC, TS, false, true>), grid, dim3(64),
                                   0, stream, m, n, A, lda, ipiv, strideP, info, bc);
            else
                hipLaunchKernelGGL((getrf_small_kernel<T, NC, TS, false, false>), grid, dim3(64),
                                   0, stream, m, n, A, lda, ipiv, strideP, info, bc);
        });
    });
    RSA_LAUNCH_CHECK();
    return rocblas_status_success;
}

template <typename T>
rocblas_status gesv_small_launch(hipStream_t stream, int n, int nrhs, Batch<T> A, int lda,
                                 int* ipiv, int64_t strideP, Batch<T> B, int ldb, int* info,
                                 int bc)
{
    square_switch(n, [&](auto ts, auto nc) {
        nrhs_switch(nrhs, [&](auto nr) {
            constexpr int TS = decltype(ts)::value;
            constexpr int NC = decltype(nc)::value;
            constexpr int NR = decltype(nr)::value;
            dim3 grid(ceil_div(bc, 64 / TS));
            hipLaunchKernelGGL((gesv_small_kernel<T, NC, TS, NR>), grid, dim3(64), 0, stream, n,
                               nrhs, A, lda, ipiv, strideP, B, ldb, info, bc);
        });
    });
    RSA_LAUNCH_CHECK();
    return rocblas_status_success;
}

template <typename T>
rocblas_status getrs_small_launch(hipStream_t stream, bool trans, int n, int nrhs, Batch<T> A,
                                  int lda, const int* ipiv, int64_t strideP, Batch<T> B, int ldb,
                                  int bc, bool pivot)
{
    square_switch(n, [&](auto ts, auto nc) {
        nrhs_switch(nrhs, [&](auto nr) {
            constexpr int TS = decltype(ts)::value;
            constexpr int NC = decltype(nc)::value;
            constexpr int NR = decltype(nr)::value;
            dim3 grid(ceil_div(bc, 64 / TS));
            hipLaunchKernelGGL((getrs_small_kernel<T, NC, TS, NR>), grid, dim3(64), 0, stream,
                               trans, n, nrhs, A, lda, ipiv, strideP, B, ldb, bc, pivot);
        });
    });
    RSA_LAUNCH_CHECK();
    return rocblas_status_success;
}

} // namespace rsa
