// vbatched.hpp -- variable-size batched getrf / getrs / gesv (orders 0..512
// mixed in one batch) in one persistent launch.  No counterpart in the
// reference; the per-matrix algorithms are the fixed-size ones (lu_small.hpp
// for n <= 64, lu_fused.hpp + getrs_fused.hpp above), so every matrix gets
// bit-for-bit the result the fixed-size routines give it.
//
// Two launches, no host synchronisation (graph-capturable):
//   1. plan (one workgroup): validates every matrix (an invalid one gets
//      info = -1 and is skipped), then orders the batch into a work queue by
//      a counting sort on the device: first every matrix with n > 64 as its
//      own item, largest order first (longest work first), then the small
//      matrices grouped by size bucket (n <= 8, 16, 24, 32, 48, 64) in packs;
//   2. work (one persistent workgroup of 512 threads per CU slot): every
//      workgroup takes the next item from an atomic counter until the queue
//      is empty.  A large matrix is factored (and solved) by the whole
//      workgroup; a pack of small matrices is spread over the 8 wavefronts,
//      each wavefront running teams of TS lanes, one matrix per team.
// The team size of each bucket, the pack size and the number of persistent
// workgroups per CU come from a tuning table (VbTuning) chosen by the
// autotuner (rocsolver_amd/autotune.py) and installed through the C ABI.
#pragma once

#include "getrs_fused.hpp"
#include "internal.hpp"
#include "lu_fused.hpp"
#include "lu_small.hpp"

namespace rsa
{

constexpr int VB_THREADS = 512;
constexpr int VB_WAVES = VB_THREADS / 64;
constexpr int VB_NR = 4; // right-hand sides per solve pass
constexpr int VB_KIND_BIG = VB_BUCKETS;

enum VbOp : int
{
    VB_GETRF = 0,
    VB_GETRS = 1,
    VB_GESV = 2
};

__host__ __device__ constexpr int vb_bucket_order(int k)
{
    return k == 0 ? 8 : k == 1 ? 16 : k == 2 ? 24 : k == 3 ? 32 : k == 4 ? 48 : 64;
}

__device__ __forceinline__ int vb_bucket(int n)
{
    return n <= 8 ? 0 : n <= 16 ? 1 : n <= 24 ? 2 : n <= 32 ? 3 : n <= 48 ? 4 : 5;
}

struct VbItem
{
    int kind; // bucket index, or VB_KIND_BIG
    int first; // position in the order list
    int count; // matrices in the item
    int pad;
};

template <typename T>
struct VbArgs
{
    const int* n;
    const int* nrhs;
    T* const* A;
    const int* lda;
    int* const* ipiv;
    T* const* B;
    const int* ldb;
    int* info;
    bool trans;
};

// ------------------------------------------------------------------ plan
// ctrl[0] = number of items, ctrl[1] = queue head (reset to 0 here).
__global__ __launch_bounds__(1024) void vb_plan_kernel(int op, int bc, const int* __restrict__ n,
                                                       const int* __restrict__ nrhs,
                                                       const int* __restrict__ lda,
                                                       const int* __restrict__ ldb,
                                                       int* __restrict__ info, int elem_bytes,
                                                       VbTuning tun, int* __restrict__ order,
                                                       VbItem* __restrict__ items,
                                                       int* __restrict__ ctrl);

// ------------------------------------------------------------------ work
template <typename T, int OP, int NC, int TS>
__device__ __forceinline__ void vb_small_pack(const VbArgs<T>& a, const int* __restrict__ order,
                                              const VbItem& item, T (*prow)[64], T (*pb)[VB_NR],
                                              int (*perm)[64], int (*spiv)[64])
{
    constexpr int TPW = 64 / TS;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int team = lane / TS;
    const int row = lane % TS;
    // uniform trip count per wavefront; a team past the end idles
    for(int j0 = w * TPW; j0 < item.count; j0 += VB_WAVES * TPW)
    {
        const int j = j0 + team;
        const bool valid = j < item.count;
        const int b = valid ? order[item.first + j] : 0;
        const int nb = valid ? a.n[b] : 0;
        T* Ab = valid ? a.A[b] : nullptr;
        int* ipb = valid ? a.ipiv[b] : nullptr;
        const int ldab = valid ? a.lda[b] : 1;
        if(OP == VB_GETRF)
        {
            const int z = team_getrf<T, NC, TS, true>(row, valid, nb, nb, Ab, ldab, ipb,
                                                      prow[team]);
            if(valid && row == 0)
                a.info[b] = z;
        }
        else
        {
            const int nr = valid ? a.nrhs[b] : 0;
            T* Bb = valid ? a.B[b] : nullptr;
            const int ldbb = valid ? a.ldb[b] : 1;
            if(OP == VB_GESV)
            {
                const int z = team_gesv<T, NC, TS, VB_NR>(row, valid, nb, nr, Ab, ldab, ipb, Bb,
                                                          ldbb, prow[team], pb[team]);
                if(valid && row == 0)
                    a.info[b] = z;
            }
            else
                team_getrs<T, NC, TS, VB_NR>(row, valid, a.trans, nb, nr, Ab, ldab, ipb, Bb, ldbb,
                                             true, pb[team], perm[team], spiv[team]);
        }
    }
}

template <typename T, int OP>
__device__ __forceinline__ void vb_big(const VbArgs<T>& a, int b)
{
    const int nb = a.n[b];
    T* Ab = a.A[b];
    const int ldab = a.lda[b];
    int* ipb = a.ipiv[b];
    if(OP == VB_GETRS)
    {
        if(a.trans)
            getrs_fused_body<T, T, VB_THREADS, VB_NR, false, true>(nb, a.nrhs[b], Ab, ldab, ipb,
                                                                   a.B[b], a.ldb[b], a.B[b],
                                                                   a.ldb[b]);
        else
            getrs_fused_body<T, T, VB_THREADS, VB_NR, false, false>(nb, a.nrhs[b], Ab, ldab, ipb,
                                                                    a.B[b], a.ldb[b], a.B[b],
                                                                    a.ldb[b]);
        return;
    }
    const int z = lu_fused_body<T, T, VB_THREADS>(nb, Ab, ldab, Ab, ldab, ipb);
    if(threadIdx.x == 0)
        a.info[b] = z;
    if(OP == VB_GESV && z == 0) // uniform; a singular system leaves B untouched
    {
        __syncthreads(); // the factors written by every wave are visible
        getrs_fused_body<T, T, VB_THREADS, VB_NR, false, false>(nb, a.nrhs[b], Ab, ldab, ipb,
                                                                a.B[b], a.ldb[b], a.B[b],
                                                                a.ldb[b]);
    }
}

// One pack of small matrices: the bucket's team size picks the instantiation.
template <typename T, int OP>
__device__ __forceinline__ void vb_small_item(const VbArgs<T>& a, const VbTuning& tun,
                                              const int* __restrict__ order, const VbItem& item,
                                              T (*prow)[64], T (*pb)[VB_NR], int (*perm)[64],
                                              int (*spiv)[64])
{
    const int ts = tun.team[item.kind];
#define RSA_VB_PACK(NC_, TS_) vb_small_pack<T, OP, NC_, TS_>(a, order, item, prow, pb, perm, spiv)
    switch(item.kind)
    {
    case 0:
        if(ts == 16)
            RSA_VB_PACK(8, 16);
        else if(ts == 32)
            RSA_VB_PACK(8, 32);
        else
            RSA_VB_PACK(8, 64);
        break;
    case 1:
        if(ts == 16)
            RSA_VB_PACK(16, 16);
        else if(ts == 32)
            RSA_VB_PACK(16, 32);
        else
            RSA_VB_PACK(16, 64);
        break;
    case 2:
        if(ts == 32)
            RSA_VB_PACK(24, 32);
        else
            RSA_VB_PACK(24, 64);
        break;
    case 3:
        if(ts == 32)
            RSA_VB_PACK(32, 32);
        else
            RSA_VB_PACK(32, 64);
        break;
    case 4: RSA_VB_PACK(48, 64); break;
    default: RSA_VB_PACK(64, 64); break;
    }
#undef RSA_VB_PACK
}

template <typename T, int OP>
__global__ __launch_bounds__(VB_THREADS, 2) void vb_work_kernel(VbArgs<T> a, VbTuning tun,
                                                                const int* __restrict__ order,
                                                                const VbItem* __restrict__ items,
                                                                int* __restrict__ ctrl)
{
    __shared__ int s_item;
    __shared__ __align__(16) T prow[VB_WAVES][4][64];
    __shared__ T pb[VB_WAVES][4][VB_NR];
    __shared__ int perm[OP == VB_GETRS ? VB_WAVES : 1][4][64];
    __shared__ int spiv[OP == VB_GETRS ? VB_WAVES : 1][4][64];
    const int w = threadIdx.x >> 6;
    const int wp = OP == VB_GETRS ? w : 0;
    const int nitems = ctrl[0];
    // Thread 0 claims the next item at the END of an iteration, and a barrier
    // follows the claim, so the loop's back-edge is reached by every lane
    // only through that barrier.  With the claim at the top of the loop the
    // compiler may merge the divergent `threadIdx.x == 0` branch ending the
    // previous item (the info store of a large getrf) with the claim's, and
    // structure the result as an inner loop that the other lanes re-enter at
    // the barrier without waiting for lane 0's claim: they read the old item
    // and factor it again, forever.
    if(threadIdx.x == 0)
        s_item = atomicAdd(&ctrl[1], 1);
    __syncthreads();
    for(;;)
    {
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads(); // every wave has read s_item before it is rewritten
        if(it >= nitems)
            break;
        const VbItem item = items[it];
        if(item.kind == VB_KIND_BIG)
            vb_big<T, OP>(a, order[item.first]);
        else
            vb_small_item<T, OP>(a, tun, order, item, prow[w], pb[w], perm[wp], spiv[wp]);
        if(threadIdx.x == 0)
            s_item = atomicAdd(&ctrl[1], 1);
        __syncthreads();
    }
}

// One persistent launch of vb_work_kernel<T, OP>; instantiated per (T, OP)
// in vbatched_<op>_<type>.hip.
template <typename T, int OP>
void vb_launch_work(hipStream_t stream, dim3 grid, const VbArgs<T>& a, const VbTuning& tun,
                    const int* order, const VbItem* items, int* ctrl);

#define RSA_VB_INSTANTIATE(T, OP)                                                              \
    template <>                                                                                 \
    void vb_launch_work<T, OP>(hipStream_t stream, dim3 grid, const VbArgs<T>& a,               \
                               const VbTuning& tun, const int* order, const VbItem* items,       \
                               int* ctrl)                                                       \
    {                                                                                           \
        hipLaunchKernelGGL((vb_work_kernel<T, OP>), grid, dim3(VB_THREADS), 0, stream, a, tun,  \
                           order, items, ctrl);                                                 \
    }

} // namespace rsa
