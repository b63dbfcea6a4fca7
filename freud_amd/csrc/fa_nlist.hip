// fa_nlist.hip -- device-side NeighborList operations
// (freud/locality/NeighborList.cc): construction from arrays with the
// reference's validation, filter / filter_r, sort, copy.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>

#include "fa_common.h"

namespace {

constexpr int kBlock = 256;

// NeighborList.cc:61-80: first failing bond decides the message; code packed
// with the bond index so an atomicMin finds the earliest one.
__global__ void validate_arrays_kernel(uint64_t nb, const uint32_t* __restrict__ qi, const uint32_t* __restrict__ pj,
                                       uint32_t nq, uint32_t np, unsigned long long* __restrict__ first_bad)
{
    uint64_t const e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nb)
        return;
    uint32_t const idx = qi[e];
    unsigned code = 0;
    if (e > 0 && idx < qi[e - 1])
        code = 1;
    else if (idx >= nq)
        code = 2;
    else if (pj[e] >= np)
        code = 3;
    if (code)
        atomicMin(first_bad, (unsigned long long) ((e << 2) | code));
}

__global__ void from_arrays_kernel(uint64_t nb, const uint32_t* __restrict__ qi, const uint32_t* __restrict__ pj,
                                   const float* __restrict__ vec, const float* __restrict__ w,
                                   uint32_t* __restrict__ nbr, float* __restrict__ dist, float* __restrict__ wout,
                                   float* __restrict__ vout)
{
    uint64_t const e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nb)
        return;
    float const x = vec[3 * e], y = vec[3 * e + 1], z = vec[3 * e + 2];
    reinterpret_cast<uint2*>(nbr)[e] = make_uint2(qi[e], pj[e]);
    // NeighborBond.h:40-43
    dist[e] = sqrtf((x * x + y * y) + z * z);
    wout[e] = w[e];
    vout[3 * e] = x;
    vout[3 * e + 1] = y;
    vout[3 * e + 2] = z;
}

// offsets[i] = first bond with query index >= i (lower bound over the i-sorted list)
__global__ void offsets_from_sorted_kernel(const uint32_t* __restrict__ nbr, uint64_t nb, uint32_t nq,
                                           uint32_t* __restrict__ offsets)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nq)
        return;
    uint64_t lo = 0, hi = nb;
    while (lo < hi)
    {
        uint64_t const mid = (lo + hi) >> 1;
        if (nbr[2 * mid] < i)
            lo = mid + 1;
        else
            hi = mid;
    }
    offsets[i] = uint32_t(lo);
}

__global__ void keep_r_kernel(uint64_t nb, const float* __restrict__ dist, float r_max, float r_min,
                              uint32_t* __restrict__ flag)
{
    uint64_t const e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < nb)
        flag[e] = (dist[e] >= r_min && dist[e] < r_max) ? 1u : 0u;
}

__global__ void keep_u8_kernel(uint64_t nb, const uint8_t* __restrict__ keep, uint32_t* __restrict__ flag)
{
    uint64_t const e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < nb)
        flag[e] = keep[e] ? 1u : 0u;
}

__global__ void compact_kernel(uint64_t nb, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                               const uint32_t* __restrict__ nbr, const float* __restrict__ dist,
                               const float* __restrict__ w, const float* __restrict__ vec, uint32_t* __restrict__ onbr,
                               float* __restrict__ odist, float* __restrict__ ow, float* __restrict__ ovec)
{
    uint64_t const e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nb || !flag[e])
        return;
    uint32_t const p = pos[e];
    reinterpret_cast<uint2*>(onbr)[p] = reinterpret_cast<const uint2*>(nbr)[e];
    odist[p] = dist[e];
    ow[p] = w[e];
    ovec[3 * p] = vec[3 * e];
    ovec[3 * p + 1] = vec[3 * e + 1];
    ovec[3 * p + 2] = vec[3 * e + 2];
}

__device__ inline uint32_t order_bits(float v)
{
    uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// sort keys for the two stable LSD passes of NeighborList::sort
__global__ void sort_keys_kernel(uint64_t nb, int pass, int by_distance, const uint32_t* __restrict__ nbr,
                                 const float* __restrict__ dist, const float* __restrict__ w,
                                 const uint32_t* __restrict__ perm, uint64_t* __restrict__ keys)
{
    uint64_t const t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nb)
        return;
    uint32_t const e = perm[t];
    uint2 const ij = reinterpret_cast<const uint2*>(nbr)[e];
    uint64_t k;
    if (!by_distance)
        // (i, j, w, d): first (w, d), then (i, j)
        k = pass == 0 ? (uint64_t(order_bits(w[e])) << 32) | order_bits(dist[e]) : (uint64_t(ij.x) << 32) | ij.y;
    else
        // (i, d, j, w): first (j, w), then (i, d)
        k = pass == 0 ? (uint64_t(ij.y) << 32) | order_bits(w[e]) : (uint64_t(ij.x) << 32) | order_bits(dist[e]);
    keys[t] = k;
}

__global__ void iota_kernel(uint64_t n, uint32_t* __restrict__ v)
{
    uint64_t const t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n)
        v[t] = uint32_t(t);
}

__global__ void gather_kernel(uint64_t nb, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ nbr,
                              const float* __restrict__ dist, const float* __restrict__ w,
                              const float* __restrict__ vec, uint32_t* __restrict__ onbr, float* __restrict__ odist,
                              float* __restrict__ ow, float* __restrict__ ovec)
{
    uint64_t const t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nb)
        return;
    uint32_t const e = perm[t];
    reinterpret_cast<uint2*>(onbr)[t] = reinterpret_cast<const uint2*>(nbr)[e];
    odist[t] = dist[e];
    ow[t] = w[e];
    ovec[3 * t] = vec[3 * e];
    ovec[3 * t + 1] = vec[3 * e + 1];
    ovec[3 * t + 2] = vec[3 * e + 2];
}

unsigned nblk(uint64_t n) { return unsigned(std::max<uint64_t>(1, (n + kBlock - 1) / kBlock)); }

int recompute_offsets(fa_nlist* nl)
{
    FA_KLAUNCH(offsets_from_sorted_kernel, dim3(nblk(uint64_t(nl->num_query_points) + 1)), dim3(kBlock), 0,
                       fa::stream(), nl->neighbors, nl->num_bonds, nl->num_query_points, nl->offsets);
    FA_HIP(hipGetLastError());
    return fa::nlist_finish_counts(nl);
}

// keep bonds whose flag is 1 (flags in a u32 scratch array)
int compact(fa_nlist* nl, uint32_t* flag)
{
    hipStream_t s = fa::stream();
    uint64_t const nb = nl->num_bonds;
    void* tmp = nullptr;
    if (int rc = fa::scratch(9, sizeof(uint32_t) * (nb + 1) + 256, &tmp))
        return rc;
    uint32_t* pos = static_cast<uint32_t*>(tmp);
    uint64_t total = 0;
    if (int rc = fa::exclusive_scan_u32(flag, pos, uint32_t(nb), &total))
        return rc;
    fa_nlist fresh{};
    if (int rc = fa::nlist_reserve(&fresh, total, nl->num_query_points))
        return rc;
    if (nb)
    {
        FA_KLAUNCH(compact_kernel, dim3(nblk(nb)), dim3(kBlock), 0, s, nb, flag, pos, nl->neighbors,
                           nl->distances, nl->weights, nl->vectors, fresh.neighbors, fresh.distances, fresh.weights,
                           fresh.vectors);
        FA_HIP(hipGetLastError());
    }
    fresh.num_bonds = total;
    fresh.num_points = nl->num_points;
    FA_HIP(hipStreamSynchronize(s));
    fa_nlist_free(nl);
    *nl = fresh;
    return recompute_offsets(nl);
}

} // namespace

extern "C" {

int fa_nlist_from_arrays(uint32_t num_query_points, uint32_t num_points, const uint32_t* d_qi, const uint32_t* d_pj,
                         const float* d_vectors, const float* d_weights, uint64_t num_bonds, fa_nlist* out)
{
    if (int rc = fa_init())
        return rc;
    hipStream_t s = fa::stream();
    if (num_bonds)
    {
        void* tmp = nullptr;
        if (int rc = fa::scratch(10, 64, &tmp))
            return rc;
        auto* bad = static_cast<unsigned long long*>(tmp);
        unsigned long long const none = ~0ull;
        FA_HIP(hipMemcpyAsync(bad, &none, sizeof(none), hipMemcpyHostToDevice, s));
        FA_KLAUNCH(validate_arrays_kernel, dim3(nblk(num_bonds)), dim3(kBlock), 0, s, num_bonds, d_qi, d_pj,
                           num_query_points, num_points, bad);
        FA_HIP(hipGetLastError());
        unsigned long long hb = 0;
        FA_HIP(hipMemcpyAsync(&hb, bad, sizeof(hb), hipMemcpyDeviceToHost, s));
        FA_HIP(hipStreamSynchronize(s));
        if (hb != none)
        {
            switch (hb & 3)
            {
            case 1:
                return fa::fail(FA_ERR_INVALID, "NeighborList query_point_index must be sorted.");
            case 2:
                return fa::fail(FA_ERR_INVALID,
                                "NeighborList query_point_index values must be less than num_query_points.");
            default:
                return fa::fail(FA_ERR_INVALID, "NeighborList point_index values must be less than num_points.");
            }
        }
    }
    if (int rc = fa::nlist_reserve(out, num_bonds, num_query_points))
        return rc;
    out->num_bonds = num_bonds;
    out->num_points = num_points;
    if (num_bonds)
    {
        FA_KLAUNCH(from_arrays_kernel, dim3(nblk(num_bonds)), dim3(kBlock), 0, s, num_bonds, d_qi, d_pj,
                           d_vectors, d_weights, out->neighbors, out->distances, out->weights, out->vectors);
        FA_HIP(hipGetLastError());
    }
    if (int rc = recompute_offsets(out))
        return rc;
    FA_HIP(hipStreamSynchronize(s));
    return FA_OK;
}

int fa_nlist_filter(fa_nlist* nl, const uint8_t* d_keep)
{
    if (int rc = fa_init())
        return rc;
    void* tmp = nullptr;
    if (int rc = fa::scratch(11, sizeof(uint32_t) * (nl->num_bonds + 1) + 64, &tmp))
        return rc;
    uint32_t* flag = static_cast<uint32_t*>(tmp);
    if (nl->num_bonds)
    {
        FA_KLAUNCH(keep_u8_kernel, dim3(nblk(nl->num_bonds)), dim3(kBlock), 0, fa::stream(), nl->num_bonds,
                           d_keep, flag);
        FA_HIP(hipGetLastError());
    }
    if (int rc = compact(nl, flag))
        return rc;
    FA_HIP(hipStreamSynchronize(fa::stream()));
    return FA_OK;
}

int fa_nlist_filter_r(fa_nlist* nl, float r_max, float r_min)
{
    if (int rc = fa_init())
        return rc;
    // NeighborList.cc:288-299
    if (r_max <= 0)
        return fa::fail(FA_ERR_INVALID, "NeighborList.filter_r requires r_max to be positive.");
    if (r_min < 0)
        return fa::fail(FA_ERR_INVALID, "NeighborList.filter_r requires r_min to be non-negative.");
    if (r_max <= r_min)
        return fa::fail(FA_ERR_INVALID, "NeighborList.filter_r requires that r_max must be greater than r_min.");
    void* tmp = nullptr;
    if (int rc = fa::scratch(11, sizeof(uint32_t) * (nl->num_bonds + 1) + 64, &tmp))
        return rc;
    uint32_t* flag = static_cast<uint32_t*>(tmp);
    if (nl->num_bonds)
    {
        FA_KLAUNCH(keep_r_kernel, dim3(nblk(nl->num_bonds)), dim3(kBlock), 0, fa::stream(), nl->num_bonds,
                           nl->distances, r_max, r_min, flag);
        FA_HIP(hipGetLastError());
    }
    if (int rc = compact(nl, flag))
        return rc;
    FA_HIP(hipStreamSynchronize(fa::stream()));
    return FA_OK;
}

int fa_nlist_sort(fa_nlist* nl, int by_distance)
{
    if (int rc = fa_init())
        return rc;
    uint64_t const nb = nl->num_bonds;
    if (nb < 2)
        return FA_OK;
    hipStream_t s = fa::stream();
    size_t tbytes = 0;
    FA_HIP(rocprim::radix_sort_pairs(nullptr, tbytes, (uint64_t*) nullptr, (uint64_t*) nullptr, (uint32_t*) nullptr,
                                     (uint32_t*) nullptr, size_t(nb), 0, 64, s));
    size_t const kb = sizeof(uint64_t) * nb, pb = sizeof(uint32_t) * nb;
    void* tmp = nullptr;
    if (int rc = fa::scratch(12, 2 * kb + 2 * pb + tbytes + 1024, &tmp))
        return rc;
    char* base = static_cast<char*>(tmp);
    auto align = [](size_t v) { return (v + 255) & ~size_t(255); };
    uint64_t* k0 = reinterpret_cast<uint64_t*>(base);
    uint64_t* k1 = reinterpret_cast<uint64_t*>(base + align(kb));
    uint32_t* p0 = reinterpret_cast<uint32_t*>(base + 2 * align(kb));
    uint32_t* p1 = reinterpret_cast<uint32_t*>(base + 2 * align(kb) + align(pb));
    void* t = base + 2 * align(kb) + 2 * align(pb);
    FA_KLAUNCH(iota_kernel, dim3(nblk(nb)), dim3(kBlock), 0, s, nb, p0);
    for (int pass = 0; pass < 2; ++pass)
    {
        FA_KLAUNCH(sort_keys_kernel, dim3(nblk(nb)), dim3(kBlock), 0, s, nb, pass, by_distance, nl->neighbors,
                           nl->distances, nl->weights, p0, k0);
        FA_HIP(hipGetLastError());
        FA_HIP(rocprim::radix_sort_pairs(t, tbytes, k0, k1, p0, p1, size_t(nb), 0, 64, s));
        std::swap(p0, p1);
    }
    fa_nlist fresh{};
    if (int rc = fa::nlist_reserve(&fresh, nb, nl->num_query_points))
        return rc;
    FA_KLAUNCH(gather_kernel, dim3(nblk(nb)), dim3(kBlock), 0, s, nb, p0, nl->neighbors, nl->distances,
                       nl->weights, nl->vectors, fresh.neighbors, fresh.distances, fresh.weights, fresh.vectors);
    FA_HIP(hipGetLastError());
    fresh.num_bonds = nb;
    fresh.num_points = nl->num_points;
    FA_HIP(hipStreamSynchronize(s));
    fa_nlist_free(nl);
    *nl = fresh;
    if (int rc = recompute_offsets(nl))
        return rc;
    FA_HIP(hipStreamSynchronize(s));
    return FA_OK;
}

int fa_nlist_copy(const fa_nlist* src, fa_nlist* dst)
{
    if (int rc = fa_init())
        return rc;
    if (int rc = fa::nlist_reserve(dst, src->num_bonds, src->num_query_points))
        return rc;
    hipStream_t s = fa::stream();
    uint64_t const nb = src->num_bonds;
    if (nb)
    {
        FA_HIP(hipMemcpyAsync(dst->neighbors, src->neighbors, 8 * nb, hipMemcpyDeviceToDevice, s));
        FA_HIP(hipMemcpyAsync(dst->distances, src->distances, 4 * nb, hipMemcpyDeviceToDevice, s));
        FA_HIP(hipMemcpyAsync(dst->weights, src->weights, 4 * nb, hipMemcpyDeviceToDevice, s));
        FA_HIP(hipMemcpyAsync(dst->vectors, src->vectors, 12 * nb, hipMemcpyDeviceToDevice, s));
    }
    uint32_t const nq = src->num_query_points;
    FA_HIP(hipMemcpyAsync(dst->offsets, src->offsets, 4 * (size_t(nq) + 1), hipMemcpyDeviceToDevice, s));
    if (nq)
    {
        FA_HIP(hipMemcpyAsync(dst->counts, src->counts, 4 * size_t(nq), hipMemcpyDeviceToDevice, s));
        FA_HIP(hipMemcpyAsync(dst->segments, src->segments, 4 * size_t(nq), hipMemcpyDeviceToDevice, s));
    }
    dst->num_bonds = nb;
    dst->num_points = src->num_points;
    dst->num_query_points = nq;
    FA_HIP(hipStreamSynchronize(s));
    return FA_OK;
}

} // extern "C"
