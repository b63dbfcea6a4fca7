// fa_prim.hip -- scan and sort primitives (rocPRIM device algorithms) used to
// turn per-query counts into NeighborList offsets and to order cluster labels.
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "fa_common.h"

namespace fa {

namespace {
struct widen
{
    __device__ __host__ uint64_t operator()(uint32_t v) const { return v; }
};
} // namespace

int exclusive_scan_u32(const uint32_t* d_in, uint32_t* d_out, uint32_t n, uint64_t* total, const uint32_t* d_extra,
                       uint32_t* h_extra, int n_extra)
{
    size_t const extra_bytes = sizeof(uint32_t) * size_t(std::min(std::max(n_extra, 1), 2));
    hipStream_t s = stream();
    FA_HIP(hipMemsetAsync(d_out, 0, sizeof(uint32_t), s));
    if (n == 0)
    {
        if (total)
            *total = 0;
        if (d_extra)
        {
            FA_HIP(hipMemcpyAsync(h_extra, d_extra, extra_bytes, hipMemcpyDeviceToHost, s));
            FA_HIP(hipStreamSynchronize(s));
        }
        return FA_OK;
    }
    size_t bytes_scan = 0, bytes_red = 0;
    FA_HIP(rocprim::inclusive_scan(nullptr, bytes_scan, d_in, d_out + 1, size_t(n), rocprim::plus<uint32_t>(), s));
    auto it = rocprim::make_transform_iterator(d_in, widen());
    uint64_t* d_total = nullptr;
    FA_HIP(rocprim::reduce(nullptr, bytes_red, it, d_total, uint64_t(0), size_t(n), rocprim::plus<uint64_t>(), s));
    void* tmp = nullptr;
    size_t const need = bytes_scan + bytes_red + 512;
    if (int rc = scratch(0, need, &tmp))
        return rc;
    char* base = static_cast<char*>(tmp);
    d_total = reinterpret_cast<uint64_t*>(base);
    void* t_scan = base + 256;
    void* t_red = base + 256 + ((bytes_scan + 255) / 256) * 256;
    FA_HIP(rocprim::inclusive_scan(t_scan, bytes_scan, d_in, d_out + 1, size_t(n), rocprim::plus<uint32_t>(), s));
    if (total)
        FA_HIP(rocprim::reduce(t_red, bytes_red, it, d_total, uint64_t(0), size_t(n), rocprim::plus<uint64_t>(), s));
    if (!total && !d_extra)
        return FA_OK;
    // Both values come back through pinned host memory: the two copies are
    // queued without host staging and one synchronisation covers them (a
    // pageable copy would synchronise by itself).
    void* stage = nullptr;
    if (int rc = readback_stage(2 * sizeof(uint64_t), &stage))
        return rc;
    uint64_t* h_stage = static_cast<uint64_t*>(stage);
    if (total)
        FA_HIP(hipMemcpyAsync(h_stage, d_total, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    if (d_extra)
        FA_HIP(hipMemcpyAsync(h_stage + 1, d_extra, extra_bytes, hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (total)
        *total = h_stage[0];
    if (d_extra)
        std::memcpy(h_extra, h_stage + 1, extra_bytes);
    return FA_OK;
}

namespace {
__global__ void zero_u32_kernel(uint32_t* __restrict__ d, uint64_t n)
{
    for (uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += uint64_t(gridDim.x) * blockDim.x)
        d[t] = 0u;
}
} // namespace

int zero_u32(uint32_t* d, uint64_t n)
{
    if (n == 0)
        return FA_OK;
    uint64_t const blocks = std::min<uint64_t>((n + 255) / 256, 4096);
    FA_KLAUNCH(zero_u32_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream(), d, n);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

size_t scan_tmp_bytes(uint32_t n)
{
    size_t bytes = 0;
    if (n == 0)
        return 256;
    uint32_t* dummy = nullptr;
    if (rocprim::inclusive_scan(nullptr, bytes, dummy, dummy, size_t(n), rocprim::plus<uint32_t>(), stream()) !=
        hipSuccess)
        return 0;
    return bytes + 256;
}

int exclusive_scan_u32_async(const uint32_t* d_in, uint32_t* d_out, uint32_t n, void* tmp, size_t tmp_bytes)
{
    if (int rc = zero_u32(d_out, 1))
        return rc;
    if (n == 0)
        return FA_OK;
    size_t bytes = tmp_bytes;
    FA_HIP(rocprim::inclusive_scan(tmp, bytes, d_in, d_out + 1, size_t(n), rocprim::plus<uint32_t>(), stream()));
    return FA_OK;
}

int sort_u64(uint64_t* d_keys, uint32_t n)
{
    if (n < 2)
        return FA_OK;
    hipStream_t s = stream();
    size_t bytes = 0;
    FA_HIP(rocprim::radix_sort_keys(nullptr, bytes, d_keys, d_keys, size_t(n), 0, 64, s));
    void* tmp = nullptr;
    size_t const kb = size_t(n) * sizeof(uint64_t);
    if (int rc = scratch(1, bytes + kb + 256, &tmp))
        return rc;
    uint64_t* alt = reinterpret_cast<uint64_t*>(static_cast<char*>(tmp));
    void* t = static_cast<char*>(tmp) + ((kb + 255) / 256) * 256;
    FA_HIP(rocprim::radix_sort_keys(t, bytes, d_keys, alt, size_t(n), 0, 64, s));
    FA_HIP(hipMemcpyAsync(d_keys, alt, kb, hipMemcpyDeviceToDevice, s));
    return FA_OK;
}

} // namespace fa
