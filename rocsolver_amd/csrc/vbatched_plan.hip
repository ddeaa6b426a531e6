// vbatched_plan.hip -- work queue of the variable-size batched routines
// (see vbatched.hpp), the tuning table and the workspace layout.
#include "vbatched.hpp"

#include <mutex>

namespace rsa
{

namespace
{
std::mutex g_tuning_mutex;
// defaults: the smallest team that holds the bucket's order, 2 rounds of
// teams per item, one persistent workgroup per CU
VbTuning g_tuning = {{16, 16, 32, 32, 64, 64}, {2, 2, 2, 2, 2, 2}, 1};
} // namespace

VbTuning vbatched_tuning()
{
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    return g_tuning;
}

bool vbatched_set_tuning(const int* params, int len)
{
    if(!params || len != VB_TUNING_LEN)
        return false;
    VbTuning t;
    for(int k = 0; k < VB_BUCKETS; ++k)
    {
        t.team[k] = params[k];
        t.pack[k] = params[VB_BUCKETS + k];
        const int ts = t.team[k];
        if((ts != 16 && ts != 32 && ts != 64) || ts < vb_bucket_order(k))
            return false;
        if(t.pack[k] < 1 || t.pack[k] > 64)
            return false;
    }
    t.wg_per_cu = params[2 * VB_BUCKETS];
    if(t.wg_per_cu < 1 || t.wg_per_cu > 4)
        return false;
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    g_tuning = t;
    return true;
}

int device_cu_count()
{
    static int cache[64] = {0};
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return 256;
    if(cache[dev] == 0)
    {
        int cus = 0;
        if(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess
           || cus <= 0)
            cus = 256;
        cache[dev] = cus;
    }
    return cache[dev];
}

static_assert(sizeof(VbItem) == 16, "work items are 16 bytes");

// Work queue of rocsolver_amd_dgesv_vbatched_noalloc, one copy per GPU for
// the life of the process.  Host code never uses this name as an address:
// the device address comes from hipGetSymbolAddress, per device.
static __device__ __align__(256) char scratch[vbatched_workspace_bytes(VB_NOALLOC_MAX_BATCH)];

void* vbatched_noalloc_scratch()
{
    static std::mutex m;
    static void* ptr[64] = {nullptr};
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return nullptr;
    std::lock_guard<std::mutex> lock(m);
    if(!ptr[dev])
    {
        void* p = nullptr;
        if(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return nullptr;
        ptr[dev] = p;
    }
    return ptr[dev];
}

// Matrix b: -2 invalid arguments, -1 nothing to compute, 0..5 small bucket,
// VB_KIND_BIG large.
__device__ __forceinline__ int vb_classify(int op, int b, const int* __restrict__ n,
                                           const int* __restrict__ nrhs,
                                           const int* __restrict__ lda,
                                           const int* __restrict__ ldb, int elem_bytes, int& nb)
{
    nb = n[b];
    const int la = lda[b];
    bool ok = nb >= 0 && nb <= FUSED_MAX && la >= (nb > 1 ? nb : 1)
              && uint64_t(la) * uint64_t(nb) * uint64_t(elem_bytes) < (uint64_t(1) << 31);
    int nr = 1;
    if(op != VB_GETRF)
    {
        nr = nrhs[b];
        const int lb = ldb[b];
        ok = ok && nr >= 0 && lb >= (nb > 1 ? nb : 1)
             && uint64_t(lb) * uint64_t(nr) * uint64_t(elem_bytes) < (uint64_t(1) << 31);
    }
    if(!ok)
        return -2;
    if(nb == 0 || nr == 0)
        return -1;
    return nb > SMALL_MAX ? VB_KIND_BIG : vb_bucket(nb);
}

__global__ __launch_bounds__(1024) void vb_plan_kernel(int op, int bc, const int* __restrict__ n,
                                                       const int* __restrict__ nrhs,
                                                       const int* __restrict__ lda,
                                                       const int* __restrict__ ldb,
                                                       int* __restrict__ info, int elem_bytes,
                                                       VbTuning tun, int* __restrict__ order,
                                                       VbItem* __restrict__ items,
                                                       int* __restrict__ ctrl)
{
    __shared__ int hist[FUSED_MAX + 1]; // large matrices per order
    __shared__ int next_big[FUSED_MAX + 1]; // next free position per order
    __shared__ int cnt[VB_BUCKETS];
    __shared__ int start_small[VB_BUCKETS];
    __shared__ int next_small[VB_BUCKETS];
    __shared__ int item_base[VB_BUCKETS];
    __shared__ int per_item[VB_BUCKETS];
    __shared__ int s_nbig;
    const int tid = threadIdx.x;
    for(int i = tid; i <= FUSED_MAX; i += blockDim.x)
        hist[i] = 0;
    if(tid < VB_BUCKETS)
        cnt[tid] = 0;
    __syncthreads();

    for(int b = tid; b < bc; b += blockDim.x)
    {
        int nb;
        const int c = vb_classify(op, b, n, nrhs, lda, ldb, elem_bytes, nb);
        if(c == -2 || c == -1 || op == VB_GETRS)
        {
            if(info)
                info[b] = c == -2 ? -1 : 0;
        }
        if(c == VB_KIND_BIG)
            atomicAdd(&hist[nb], 1);
        else if(c >= 0)
            atomicAdd(&cnt[c], 1);
    }
    __syncthreads();
    if(tid == 0)
    {
        // large matrices first, largest order first; then buckets 5 .. 0
        int pos = 0;
        for(int v = FUSED_MAX; v > SMALL_MAX; --v)
        {
            next_big[v] = pos;
            pos += hist[v];
        }
        s_nbig = pos;
        int nitems = pos;
        for(int k = VB_BUCKETS - 1; k >= 0; --k)
        {
            start_small[k] = pos;
            next_small[k] = pos;
            pos += cnt[k];
            per_item[k] = tun.pack[k] * VB_WAVES * (64 / tun.team[k]);
            item_base[k] = nitems;
            nitems += (cnt[k] + per_item[k] - 1) / per_item[k];
        }
        ctrl[0] = nitems;
        ctrl[1] = 0;
    }
    __syncthreads();
    for(int b = tid; b < bc; b += blockDim.x)
    {
        int nb;
        const int c = vb_classify(op, b, n, nrhs, lda, ldb, elem_bytes, nb);
        if(c == VB_KIND_BIG)
            order[atomicAdd(&next_big[nb], 1)] = b;
        else if(c >= 0)
            order[atomicAdd(&next_small[c], 1)] = b;
    }
    for(int i = tid; i < s_nbig; i += blockDim.x)
        items[i] = VbItem{VB_KIND_BIG, i, 1, 0};
    for(int k = 0; k < VB_BUCKETS; ++k)
    {
        const int per = per_item[k];
        const int nk = (cnt[k] + per - 1) / per;
        for(int i = tid; i < nk; i += blockDim.x)
        {
            const int first = start_small[k] + i * per;
            const int left = cnt[k] - i * per;
            items[item_base[k] + i] = VbItem{k, first, left < per ? left : per, 0};
        }
    }
}

template <typename T>
rocblas_status vbatched(hipStream_t stream, int op, bool trans, const int* n, const int* nrhs,
                        T* const* A, const int* lda, int* const* ipiv, T* const* B,
                        const int* ldb, int* info, int bc, void* work)
{
    if(bc <= 0)
        return rocblas_status_success;
    const VbTuning tun = vbatched_tuning();
    char* p = static_cast<char*>(work);
    int* order = reinterpret_cast<int*>(p);
    p += (sizeof(int) * size_t(bc) + 255) / 256 * 256;
    VbItem* items = reinterpret_cast<VbItem*>(p);
    p += (sizeof(VbItem) * (size_t(bc) + VB_BUCKETS) + 255) / 256 * 256;
    int* ctrl = reinterpret_cast<int*>(p);

    hipLaunchKernelGGL(vb_plan_kernel, dim3(1), dim3(1024), 0, stream, op, bc, n, nrhs, lda, ldb,
                       info, int(sizeof(T)), tun, order, items, ctrl);
    RSA_LAUNCH_CHECK();
    const VbArgs<T> a{n, nrhs, A, lda, ipiv, B, ldb, info, trans};
    const dim3 grid(device_cu_count() * tun.wg_per_cu);
    if(op == VB_GETRF)
        vb_launch_work<T, VB_GETRF>(stream, grid, a, tun, order, items, ctrl);
    else if(op == VB_GETRS)
        vb_launch_work<T, VB_GETRS>(stream, grid, a, tun, order, items, ctrl);
    else
        vb_launch_work<T, VB_GESV>(stream, grid, a, tun, order, items, ctrl);
    RSA_LAUNCH_CHECK();
    return rocblas_status_success;
}

template rocblas_status vbatched<float>(hipStream_t, int, bool, const int*, const int*,
                                        float* const*, const int*, int* const*, float* const*,
                                        const int*, int*, int, void*);
template rocblas_status vbatched<double>(hipStream_t, int, bool, const int*, const int*,
                                         double* const*, const int*, int* const*, double* const*,
                                         const int*, int*, int, void*);

} // namespace rsa
