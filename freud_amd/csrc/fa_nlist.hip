// fa_nlist.hip -- device-side NeighborList operations
// (freud/locality/NeighborList.cc): construction from arrays with the
// reference's validation, filter / filter_r, sort, copy; the SANN and RAD
// NeighborList filters (freud/locality/FilterSANN.cc, FilterRAD.cc).
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
    FA_API_LOCK;
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
    FA_API_LOCK;
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
    FA_API_LOCK;
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
    FA_API_LOCK;
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
    FA_API_LOCK;
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

// ---------------------------------------------------------------- filters
// FilterSANN / FilterRAD walk each query point's segment of the unfiltered
// list sorted by (i, d, j, w) (NeighborList::sort(true)) and keep a subset of
// it; the kept bonds stay in that order, which is the (i, d, j, w) order the
// reference's final sort gives (FilterSANN.cc:99-100, FilterRAD.cc:127-128).
namespace {

// grid-stride cap of the filter kernels: bounds the atomics that collect the
// largest kept point index and the unfilled count (one per block)
constexpr unsigned kFilterBlocks = 2048;

__device__ inline uint32_t wave_max_u32(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = max(v, uint32_t(__shfl_xor(int(v), o, 64)));
    return v;
}

__device__ inline uint32_t wave_sum_u32(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += uint32_t(__shfl_xor(int(v), o, 64));
    return v;
}

// counters[0] += sum over the block of n_unf; counters[1] = max(.., max of mj).
// Every thread of the block calls it.
__device__ inline void block_counters(uint32_t n_unf, uint32_t mj, uint32_t* __restrict__ counters)
{
    __shared__ uint32_t s_sum[kBlock / 64], s_max[kBlock / 64];
    n_unf = wave_sum_u32(n_unf);
    mj = wave_max_u32(mj);
    if ((threadIdx.x & 63) == 0)
    {
        s_sum[threadIdx.x >> 6] = n_unf;
        s_max[threadIdx.x >> 6] = mj;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t a = 0, m = 0;
        for (int w = 0; w < kBlock / 64; ++w)
        {
            a += s_sum[w];
            m = max(m, s_max[w]);
        }
        if (a)
            atomicAdd(&counters[0], a);
        if (m)
            atomicMax(&counters[1], m);
    }
}

// FilterSANN.cc:52-83, one thread per query point: the first three bonds, then
// while sum / (m - 2.0) > d_m, in double as the reference's 2.0 literal makes
// it.  The kept bonds are the first kept[i] of the segment.
__global__ void __launch_bounds__(kBlock) sann_kernel(uint32_t nq, const uint32_t* __restrict__ offsets,
                                                      const float* __restrict__ dist, uint32_t* __restrict__ kept,
                                                      uint8_t* __restrict__ unfilled, uint32_t* __restrict__ counters)
{
    uint32_t n_unf = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x)
    {
        uint32_t const first = offsets[i], n = offsets[i + 1] - first;
        const float* d = dist + first;
        uint32_t m = 0;
        float sum = 0.0f;
        for (; m < 3 && m < n; ++m)
            sum += d[m];
        while (m < n && (double(sum) / (double(float(m)) - 2.0)) > double(d[m]))
        {
            sum += d[m];
            ++m;
        }
        kept[i] = m;
        // FilterSANN.cc:86-93: the list ran out before the shell closed
        uint8_t const u = m == n ? 1 : 0;
        unfilled[i] = u;
        n_unf += u;
    }
    block_counters(n_unf, 0, counters);
}

// bond flags of the SANN prefixes, and the largest kept point index + 1
__global__ void __launch_bounds__(kBlock) sann_flags_kernel(uint64_t nb, const uint32_t* __restrict__ nbr,
                                                            const uint32_t* __restrict__ offsets,
                                                            const uint32_t* __restrict__ kept,
                                                            uint32_t* __restrict__ flag,
                                                            uint32_t* __restrict__ counters)
{
    uint32_t mj = 0;
    for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < nb; e += uint64_t(gridDim.x) * blockDim.x)
    {
        uint2 const ij = reinterpret_cast<const uint2*>(nbr)[e];
        bool const k = e - offsets[ij.x] < kept[ij.x];
        flag[e] = k ? 1u : 0u;
        if (k)
            mj = max(mj, ij.y + 1);
    }
    block_counters(0, mj, counters);
}

// FilterRAD.cc:70,76: v = box.wrap(q_i - p_j) per bond (the kept bonds carry
// it, FilterRAD.cc:97-100, with weight 1), and r_j^2 for the blocking test
__global__ void __launch_bounds__(kBlock) rad_vectors_kernel(fa::DevBox b, uint64_t nb,
                                                             const uint32_t* __restrict__ nbr,
                                                             const float* __restrict__ pts,
                                                             const float* __restrict__ qpts,
                                                             const float* __restrict__ radii,
                                                             float* __restrict__ vec, float* __restrict__ w,
                                                             float* __restrict__ rsq)
{
    for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < nb; e += uint64_t(gridDim.x) * blockDim.x)
    {
        uint2 const ij = reinterpret_cast<const uint2*>(nbr)[e];
        float x, y, z;
        fa::wrap(b, qpts[3 * ij.x] - pts[3 * ij.y], qpts[3 * ij.x + 1] - pts[3 * ij.y + 1],
                 qpts[3 * ij.x + 2] - pts[3 * ij.y + 2], x, y, z);
        vec[3 * e] = x;
        vec[3 * e + 1] = y;
        vec[3 * e + 2] = z;
        w[e] = 1.0f;
        float const r = radii ? radii[ij.y] : 1.0f;
        rsq[e] = r * r;
    }
}

// FilterRAD.cc:61-118, one wave per query point.  Whether a closer k blocks j
// does not depend on which closer bonds were kept, so the lanes test 64
// candidates j at once, each against every closer k in order:
//   ((r_j^2 |v_k|^2) d_j) d_k < (r_k^2 (v_j . v_k)) |v_j|^2
// with dot = (x x + y y) + z z (VectorMath.h:270-273).  RAD-open keeps the
// unblocked j; RAD-closed keeps the segment up to its first blocked j and
// reports the query point unfilled when none is blocked (FilterRAD.cc:107-115).
__global__ void __launch_bounds__(kBlock) rad_kernel(uint32_t nq, const uint32_t* __restrict__ offsets,
                                                     const uint32_t* __restrict__ nbr, const float* __restrict__ dist,
                                                     const float* __restrict__ vec, const float* __restrict__ rsq,
                                                     int terminate, uint32_t* __restrict__ flag,
                                                     uint8_t* __restrict__ unfilled, uint32_t* __restrict__ counters)
{
    uint32_t const lane = threadIdx.x & 63;
    uint32_t const waves = gridDim.x * (kBlock / 64);
    uint32_t mj = 0, n_unf = 0;
    for (uint32_t i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); i < nq; i += waves)
    {
        uint32_t const first = offsets[i], n = offsets[i + 1] - first;
        bool closed = false;
        for (uint32_t c = 0; c < n; c += 64)
        {
            uint32_t const t = c + lane;
            bool const active = t < n;
            bool blocked = false;
            uint32_t j = 0;
            if (active)
            {
                uint64_t const e = uint64_t(first) + t;
                j = nbr[2 * e + 1];
                float const x1 = vec[3 * e], y1 = vec[3 * e + 1], z1 = vec[3 * e + 2];
                float const dj = dist[e], rj_sq = rsq[e];
                float const v1v1 = (x1 * x1 + y1 * y1) + z1 * z1;
                for (uint32_t k = 0; k < t; ++k)
                {
                    uint64_t const ek = uint64_t(first) + k;
                    float const x2 = vec[3 * ek], y2 = vec[3 * ek + 1], z2 = vec[3 * ek + 2];
                    float const v2v2 = (x2 * x2 + y2 * y2) + z2 * z2;
                    float const v1v2 = (x1 * x2 + y1 * y2) + z1 * z2;
                    float const lhs = ((rj_sq * v2v2) * dj) * dist[ek];
                    float const rhs = (rsq[ek] * v1v2) * v1v1;
                    if (lhs < rhs)
                    {
                        blocked = true;
                        break;
                    }
                }
            }
            unsigned long long const bm = __ballot(blocked);
            bool keep = active && !blocked;
            if (terminate && bm)
                keep = active && lane < uint32_t(__ffsll(bm) - 1);
            if (active)
                flag[uint64_t(first) + t] = keep ? 1u : 0u;
            if (keep)
                mj = max(mj, j + 1);
            if (terminate && bm)
            {
                // FilterRAD.cc:102-106: nothing past the first blocked neighbour
                for (uint32_t r = c + 64 + lane; r < n; r += 64)
                    flag[uint64_t(first) + r] = 0u;
                closed = true;
                break;
            }
        }
        if (lane == 0)
        {
            uint8_t const u = (terminate && !closed) ? 1 : 0;
            unfilled[i] = u;
            n_unf += u;
        }
    }
    block_counters(n_unf, mj, counters);
}

// compact the flagged bonds of the sorted copy and give the result the index
// ranges of a NeighborList built from bonds (NeighborList.cc:130-173: one past
// the largest kept query point and point indices)
int filter_finish(fa_nlist* out, uint32_t* flag, const uint32_t* counters, uint32_t* n_unfilled)
{
    hipStream_t s = fa::stream();
    uint32_t const nq_in = out->num_query_points, np_in = out->num_points;
    if (int rc = compact(out, flag))
        return rc;
    uint32_t h[2] = {0, 0};
    FA_HIP(hipMemcpyAsync(h, counters, sizeof(h), hipMemcpyDeviceToHost, s));
    uint32_t last[2] = {0, 0};
    if (out->num_bonds)
        FA_HIP(hipMemcpyAsync(last, out->neighbors + 2 * (out->num_bonds - 1), sizeof(last), hipMemcpyDeviceToHost,
                              s));
    FA_HIP(hipStreamSynchronize(s));
    *n_unfilled = h[0];
    if (out->num_bonds)
    {
        out->num_query_points = last[0] + 1;
        out->num_points = h[1];
        if (int rc = recompute_offsets(out))
            return rc;
        FA_HIP(hipStreamSynchronize(s));
    }
    else
    {
        out->num_query_points = nq_in;
        out->num_points = np_in;
    }
    return FA_OK;
}

int filter_check(const fa_nlist* nl, const fa_nlist* out, const uint8_t* d_unfilled, const uint32_t* n_unfilled)
{
    if (!nl || !out || !n_unfilled || out == nl)
        return fa::fail(FA_ERR_INVALID, "filter: nl, out (a different list) and n_unfilled are required");
    if (nl->num_query_points && !d_unfilled)
        return fa::fail(FA_ERR_INVALID, "filter: d_unfilled must hold one byte per query point");
    return FA_OK;
}

// sorted copy of nl in *out plus the filter's scratch: flags (nb + 1 u32), a
// per-query-point or per-bond float array, and two zeroed counters
int filter_setup(const fa_nlist* nl, fa_nlist* out, uint8_t* d_unfilled, uint32_t* n_unfilled, size_t extra_bytes,
                 uint32_t** flag, void** extra, uint32_t** counters)
{
    (void) d_unfilled;
    (void) n_unfilled;
    if (int rc = fa_nlist_copy(nl, out))
        return rc;
    if (int rc = fa_nlist_sort(out, 1))
        return rc;
    auto align = [](size_t v) { return (v + 255) & ~size_t(255); };
    size_t const fb = align(sizeof(uint32_t) * (out->num_bonds + 1));
    void* tmp = nullptr;
    if (int rc = fa::scratch(23, fb + align(extra_bytes) + 256, &tmp))
        return rc;
    char* base = static_cast<char*>(tmp);
    *flag = reinterpret_cast<uint32_t*>(base);
    *extra = base + fb;
    *counters = reinterpret_cast<uint32_t*>(base + fb + align(extra_bytes));
    FA_HIP(hipMemsetAsync(*counters, 0, 2 * sizeof(uint32_t), fa::stream()));
    return FA_OK;
}

unsigned filter_grid(uint64_t threads) { return unsigned(std::min<uint64_t>(nblk(threads), kFilterBlocks)); }

// ---- per-segment path.  A NeighborList is grouped by query point, so the
// (i, d, j, w) order only has to be made inside each query point's segment:
// one wave per query point ranks its segment (at most 64 P bonds) in LDS by
// (d, j, w), applies the filter to the ranked bonds, and either counts the
// kept bonds (count pass) or writes them at the offsets the scan of the
// counts gives (emit pass).  The list is never copied or globally sorted.
constexpr uint32_t kSegMax = 512;
constexpr unsigned kSegBlocks = 8192;

template<int P> constexpr int seg_waves() { return P <= 2 ? 4 : (P == 4 ? 2 : 1); }

struct SegArgs
{
    uint32_t nq;
    const uint32_t* offsets;
    const uint32_t* nbr;
    const float* dist;
    const float* w;
    const float* vec;
    const float* pts;
    const float* qpts;
    const float* radii;
    int terminate;
    uint32_t* kept;     // count pass: kept bonds per query point
    uint8_t* unfilled;  // count pass
    uint32_t* counters; // [0] largest kept i + 1, [1] largest kept j + 1, [2] unfilled query points
    uint16_t* code;     // count pass: per input bond, its rank among its query point's kept bonds, or 0xffff
    const uint32_t* base; // emit pass: first output bond of each query point
    uint32_t* onbr;
    float* odist;
    float* ow;
    float* ovec;
};

__global__ void __launch_bounds__(kBlock) seg_max_kernel(uint32_t nq, const uint32_t* __restrict__ offsets,
                                                         uint32_t* __restrict__ counters)
{
    uint32_t m = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x)
        m = max(m, offsets[i + 1] - offsets[i]);
    block_counters(0, m, counters);
}

// FilterSANN.cc:73 tests RN(sum / (m - 2.0)) > d_m in double.  With sum and
// d_m floats and k = m - 2 a positive integer below 2^29, d_m k is exact in
// double and the test equals sum > d_m k: rounding is monotonic, so a
// quotient at most d_m (a double) rounds to at most d_m; and if sum / k > d_m
// then sum > d_m k >= d_m, sum and d_m k are multiples of the float spacing
// u of d_m, and sum / k - d_m >= u / k, far above the double spacing at d_m.
__device__ inline bool sann_continues(float sum, float d, uint32_t k)
{
    return double(sum) > double(d) * double(k);
}

// FilterRAD.cc:82-86: does the closer bond k block bond j?  v = (wrap(q_i - p),
// r^2); products left to right, dot = (x x + y y) + z z (VectorMath.h:270-273)
__device__ inline bool rad_blocks(float4 vj, float dj, float vjvj, float4 vk, float dk)
{
    float const vkvk = (vk.x * vk.x + vk.y * vk.y) + vk.z * vk.z;
    float const vjvk = (vj.x * vk.x + vj.y * vk.y) + vj.z * vk.z;
    return ((vj.w * vkvk) * dj) * dk < (vk.w * vjvk) * vjvj;
}

// the (d, j, w) order of NeighborBond::less_as_distance (NeighborBond.h:97-112)
// on order-preserving bits, ties by segment position
__device__ inline bool key_less(uint4 u, uint32_t pu, uint4 t, uint32_t pt)
{
    return u.x < t.x || (u.x == t.x && (u.y < t.y || (u.y == t.y && (u.z < t.z || (u.z == t.z && pu < pt)))));
}

// bond e of query point i: its key (d, j, w bits), distance and, for RAD,
// (box.wrap(query_points[i] - points[j]), r_j^2) (FilterRAD.cc:70,76-80)
template<bool RAD>
__device__ inline void seg_load(const fa::DevBox& b, const SegArgs& a, uint32_t i, uint64_t e, uint4& key, float& d,
                                float4& v)
{
    uint2 const ij = reinterpret_cast<const uint2*>(a.nbr)[e];
    d = a.dist[e];
    key = make_uint4(order_bits(d), ij.y, order_bits(a.w[e]), 0u);
    if (RAD)
    {
        float x, y, z;
        fa::wrap(b, a.qpts[3 * i] - a.pts[3 * ij.y], a.qpts[3 * i + 1] - a.pts[3 * ij.y + 1],
                 a.qpts[3 * i + 2] - a.pts[3 * ij.y + 2], x, y, z);
        float const r = a.radii ? a.radii[ij.y] : 1.0f;
        v = make_float4(x, y, z, r * r);
    }
}

__device__ inline unsigned long long low_mask(uint32_t n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }

// Segments of at most 64 bonds, one wave per block.  Each pass ranks, filters
// and counts or writes up to two segments side by side: query points 2h and
// 2h + 1 share the 64 lanes when their segments hold at most 64 bonds
// together (segment A on lanes [0, nA), B on [nA, nA + nB)), else they take
// one pass each.
template<bool RAD>
__global__ void __launch_bounds__(64) seg_pair_kernel(fa::DevBox b, SegArgs a)
{
    __shared__ uint2 s_key[64];    // segment order: (d bits, j)
    __shared__ uint32_t s_kw[64];  // segment order: w bits
    __shared__ uint32_t s_src[64]; // ranked slot -> segment position
    __shared__ uint32_t s_j[64];
    __shared__ float s_d[64];
    __shared__ float4 s_v[RAD ? 64 : 1];  // ranked (wrap(q_i - p_j), |v|^2)
    __shared__ float s_r2[RAD ? 64 : 1];  // ranked r_j^2
    uint32_t const lane = threadIdx.x;
    uint32_t n_unf = 0, mj = 0, mi = 0;
    uint32_t const npairs = (a.nq + 1) / 2;
    for (uint32_t h = blockIdx.x; h < npairs; h += gridDim.x)
    {
        uint32_t const i0 = 2 * h;
        bool const two = i0 + 1 < a.nq;
        uint32_t const f0 = a.offsets[i0], f1 = a.offsets[i0 + 1], f2 = two ? a.offsets[i0 + 2] : f1;
        uint32_t const n0 = f1 - f0, n1 = f2 - f1;
        bool const packed = n0 + n1 <= 64;
        for (int pass = 0; pass < (packed ? 1 : 2); ++pass)
        {
            uint32_t const iA = pass ? i0 + 1 : i0;
            uint32_t const fA = pass ? f1 : f0;
            uint32_t const nA = (packed || !pass) ? n0 : n1;
            uint32_t const nB = packed ? n1 : 0;
            bool const inB = lane >= nA;
            uint32_t const base = inB ? nA : 0, own = inB ? nB : nA, t = lane - base;
            bool const act = t < own;
            uint32_t const iq = inB ? iA + 1 : iA;
            uint64_t const e = uint64_t(inB ? f1 : fA) + t;
            uint4 key = make_uint4(~0u, ~0u, ~0u, 0u);
            float d = 0.0f;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (act)
            {
                seg_load<RAD>(b, a, iq, e, key, d, v);
                s_key[lane] = make_uint2(key.x, key.y);
                s_kw[lane] = key.z;
            }
            __syncthreads();
            // rank by (d, j); a segment holding one (d, j) twice shows as two
            // bonds of one rank and is ranked again by (d, j, w, position)
            uint64_t const k64 = (uint64_t(key.x) << 32) | key.y;
            uint32_t rank = 0;
            for (uint32_t u = 0; u < own; ++u)
            {
                uint2 const ku = s_key[base + u];
                rank += ((uint64_t(ku.x) << 32) | ku.y) < k64 ? 1u : 0u;
            }
            if (act)
                s_src[base + rank] = t;
            __syncthreads();
            if (__ballot(act && s_src[base + rank] != t))
            {
                rank = 0;
                for (uint32_t u = 0; u < own; ++u)
                {
                    uint2 const ku = s_key[base + u];
                    rank += key_less(make_uint4(ku.x, ku.y, s_kw[base + u], 0u), u, key, t) ? 1u : 0u;
                }
            }
            if (act)
            {
                uint32_t const slot = base + rank;
                s_src[slot] = t;
                s_j[slot] = key.y;
                s_d[slot] = d;
                if (RAD)
                {
                    s_v[slot] = make_float4(v.x, v.y, v.z, (v.x * v.x + v.y * v.y) + v.z * v.z);
                    s_r2[slot] = v.w;
                }
            }
            __syncthreads();
            // this lane's ranked bond: position t of its segment, slot lane
            uint32_t rs = 0, rj = 0;
            float rd = 0.0f, rr2 = 0.0f;
            float4 rv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (act)
            {
                rs = s_src[lane];
                rj = s_j[lane];
                rd = s_d[lane];
                if (RAD)
                {
                    rv = s_v[lane];
                    rr2 = s_r2[lane];
                }
            }
            unsigned long long const maskA = low_mask(nA), maskB = low_mask(nA + nB) & ~maskA;
            bool keep;
            uint32_t cA, cB;
            bool uA, uB;
            if (!RAD)
            {
                // FilterSANN.cc:62-83 on the first lane of each segment
                uint32_t m = 0;
                if (act && t == 0)
                {
                    float sum = 0.0f;
                    for (; m < 3 && m < own; ++m)
                        sum += s_d[base + m];
                    while (m < own && sann_continues(sum, s_d[base + m], m - 2))
                    {
                        sum += s_d[base + m];
                        ++m;
                    }
                }
                cA = nA ? uint32_t(__shfl(int(m), 0, 64)) : 0u;
                cB = nB ? uint32_t(__shfl(int(m), int(nA), 64)) : 0u;
                keep = act && t < (inB ? cB : cA);
                uA = cA == nA;
                uB = cB == nB;
            }
            else
            {
                // FilterRAD.cc:82-86 with |v|^2 of each ranked bond kept beside it
                bool blocked = false;
                if (act)
                    for (uint32_t k = 0; k < t; ++k)
                    {
                        float4 const vk = s_v[base + k];
                        float const vjvk = (rv.x * vk.x + rv.y * vk.y) + rv.z * vk.z;
                        if (((rr2 * vk.w) * rd) * s_d[base + k] < (s_r2[base + k] * vjvk) * rv.w)
                        {
                            blocked = true;
                            break;
                        }
                    }
                unsigned long long const bm = __ballot(blocked);
                if (a.terminate)
                {
                    // RAD-closed: each segment's ranked prefix before its first blocked bond
                    cA = (bm & maskA) ? uint32_t(__ffsll(bm & maskA) - 1) : nA;
                    cB = (bm & maskB) ? uint32_t(__ffsll(bm & maskB) - 1) - nA : nB;
                    keep = act && t < (inB ? cB : cA);
                    uA = cA == nA;
                    uB = cB == nB;
                }
                else
                {
                    keep = act && !blocked;
                    unsigned long long const km = __ballot(keep);
                    cA = uint32_t(__popcll(km & maskA));
                    cB = uint32_t(__popcll(km & maskB));
                    uA = uB = false;
                }
            }
            {
                // this bond's rank among its segment's kept bonds
                unsigned long long const km = __ballot(keep) & (inB ? maskB : maskA);
                if (act)
                    a.code[uint64_t(inB ? f1 : fA) + rs] =
                        keep ? uint16_t(__popcll(km & ((1ull << lane) - 1ull))) : uint16_t(0xffff);
                if (lane == 0)
                {
                    a.kept[iA] = cA;
                    a.unfilled[iA] = uA ? 1 : 0;
                    n_unf += uA ? 1u : 0u;
                    if (cA)
                        mi = max(mi, iA + 1);
                    if (packed && two)
                    {
                        a.kept[iA + 1] = cB;
                        a.unfilled[iA + 1] = uB ? 1 : 0;
                        n_unf += uB ? 1u : 0u;
                        if (cB)
                            mi = max(mi, iA + 2);
                    }
                }
                if (keep)
                    mj = max(mj, rj + 1);
            }
            __syncthreads();
        }
    }
    {
        mj = wave_max_u32(mj);
        if (lane == 0)
        {
            if (mi)
                atomicMax(&a.counters[0], mi);
            if (mj)
                atomicMax(&a.counters[1], mj);
            if (n_unf)
                atomicAdd(&a.counters[2], n_unf);
        }
    }
}

// Segments of up to 64 P bonds (P = 2, 4, 8), one wave per query point; the
// lane holding segment position t also holds t + 64, t + 128, ...
template<int P, bool RAD>
__global__ void __launch_bounds__(64 * seg_waves<P>()) seg_filter_kernel(fa::DevBox b, SegArgs a)
{
    constexpr int W = seg_waves<P>();
    constexpr int S = 64 * P;
    __shared__ uint4 s_key[W][S];    // segment order
    __shared__ uint32_t s_src[W][S]; // ranked position -> segment position
    __shared__ uint32_t s_j[W][S];
    __shared__ float s_d[W][S];
    __shared__ float4 s_v[W][RAD ? S : 1];
    __shared__ uint32_t s_red[W][3];
    uint32_t const lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t n_unf = 0, mj = 0, mi = 0;
    for (uint32_t g = blockIdx.x * W; g < a.nq; g += gridDim.x * W)
    {
        uint32_t const i = g + wv;
        uint32_t first = 0, n = 0;
        if (i < a.nq)
        {
            first = a.offsets[i];
            n = a.offsets[i + 1] - first;
        }
        uint4 key[P];
        float dd[P];
        float4 v[P];
#pragma unroll
        for (int p = 0; p < P; ++p)
        {
            uint32_t const t = lane + 64 * p;
            key[p] = make_uint4(~0u, ~0u, ~0u, 0u);
            dd[p] = 0.0f;
            v[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (t < n)
            {
                seg_load<RAD>(b, a, i, uint64_t(first) + t, key[p], dd[p], v[p]);
                s_key[wv][t] = key[p];
            }
        }
        __syncthreads();
        uint32_t rank[P];
#pragma unroll
        for (int p = 0; p < P; ++p)
            rank[p] = 0;
        for (uint32_t u = 0; u < n; ++u)
        {
            uint4 const ku = s_key[wv][u];
#pragma unroll
            for (int p = 0; p < P; ++p)
                rank[p] += key_less(ku, u, key[p], lane + 64 * p) ? 1u : 0u;
        }
#pragma unroll
        for (int p = 0; p < P; ++p)
            if (lane + 64 * p < n)
            {
                s_src[wv][rank[p]] = lane + 64 * p;
                s_j[wv][rank[p]] = key[p].y;
                s_d[wv][rank[p]] = dd[p];
                if (RAD)
                    s_v[wv][rank[p]] = v[p];
            }
        __syncthreads();
        bool keep[P];
        uint32_t cnt = 0;
        bool unf = false;
        if (!RAD)
        {
            // FilterSANN.cc:62-83 (one lane; the sum is sequential)
            uint32_t m = 0;
            if (lane == 0)
            {
                float sum = 0.0f;
                for (; m < 3 && m < n; ++m)
                    sum += s_d[wv][m];
                while (m < n && sann_continues(sum, s_d[wv][m], m - 2))
                {
                    sum += s_d[wv][m];
                    ++m;
                }
            }
            m = uint32_t(__shfl(int(m), 0, 64));
#pragma unroll
            for (int p = 0; p < P; ++p)
                keep[p] = lane + 64 * p < m;
            cnt = m;
            unf = m == n;
        }
        else
        {
            unsigned long long bm[P];
            bool blocked[P];
#pragma unroll
            for (int p = 0; p < P; ++p)
            {
                uint32_t const sp = lane + 64 * p;
                blocked[p] = false;
                if (sp < n)
                {
                    float4 const vj = s_v[wv][sp];
                    float const dj = s_d[wv][sp];
                    float const vv = (vj.x * vj.x + vj.y * vj.y) + vj.z * vj.z;
                    for (uint32_t k = 0; k < sp; ++k)
                        if (rad_blocks(vj, dj, vv, s_v[wv][k], s_d[wv][k]))
                        {
                            blocked[p] = true;
                            break;
                        }
                }
                bm[p] = __ballot(blocked[p]);
            }
            if (a.terminate)
            {
                uint32_t fb = n;
#pragma unroll
                for (int p = P - 1; p >= 0; --p)
                    if (bm[p])
                        fb = 64 * p + uint32_t(__ffsll(bm[p]) - 1);
#pragma unroll
                for (int p = 0; p < P; ++p)
                    keep[p] = lane + 64 * p < fb;
                cnt = fb;
                unf = fb == n;
            }
            else
            {
#pragma unroll
                for (int p = 0; p < P; ++p)
                {
                    keep[p] = lane + 64 * p < n && !blocked[p];
                    cnt += uint32_t(__popcll(__ballot(keep[p])));
                }
            }
        }
        if (lane == 0 && i < a.nq)
        {
            a.kept[i] = cnt;
            a.unfilled[i] = unf ? 1 : 0;
            n_unf += unf ? 1u : 0u;
            if (cnt)
                mi = i + 1;
        }
        // each bond's rank among the segment's kept bonds
        uint32_t before = 0;
#pragma unroll
        for (int p = 0; p < P; ++p)
        {
            uint32_t const sp = lane + 64 * p;
            unsigned long long const km = __ballot(keep[p]);
            if (sp < n)
                a.code[uint64_t(first) + s_src[wv][sp]] =
                    keep[p] ? uint16_t(before + __popcll(km & ((1ull << lane) - 1ull))) : uint16_t(0xffff);
            if (keep[p])
                mj = max(mj, s_j[wv][sp] + 1);
            before += uint32_t(__popcll(km));
        }
        __syncthreads();
    }
    {
        n_unf = wave_sum_u32(n_unf);
        mj = wave_max_u32(mj);
        mi = wave_max_u32(mi);
        if (lane == 0)
        {
            s_red[wv][0] = mi;
            s_red[wv][1] = mj;
            s_red[wv][2] = n_unf;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            uint32_t r0 = 0, r1 = 0, r2 = 0;
            for (int q = 0; q < W; ++q)
            {
                r0 = max(r0, s_red[q][0]);
                r1 = max(r1, s_red[q][1]);
                r2 += s_red[q][2];
            }
            if (r0)
                atomicMax(&a.counters[0], r0);
            if (r1)
                atomicMax(&a.counters[1], r1);
            if (r2)
                atomicAdd(&a.counters[2], r2);
        }
    }
}


// Writes each kept bond of the input list at its query point's first output
// bond plus its rank among the kept bonds (the code of the count pass).
template<bool RAD>
__global__ void __launch_bounds__(kBlock) seg_emit_kernel(fa::DevBox b, SegArgs a, uint64_t nb)
{
    for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < nb; e += uint64_t(gridDim.x) * blockDim.x)
    {
        uint32_t const c = a.code[e];
        if (c == 0xffffu)
            continue;
        uint2 const ij = reinterpret_cast<const uint2*>(a.nbr)[e];
        uint64_t const o = uint64_t(a.base[ij.x]) + c;
        reinterpret_cast<uint2*>(a.onbr)[o] = ij;
        a.odist[o] = a.dist[e];
        if (RAD)
        {
            // the vector of the count pass, recomputed (FilterRAD.cc:97-100: weight 1)
            float x, y, z;
            fa::wrap(b, a.qpts[3 * ij.x] - a.pts[3 * ij.y], a.qpts[3 * ij.x + 1] - a.pts[3 * ij.y + 1],
                     a.qpts[3 * ij.x + 2] - a.pts[3 * ij.y + 2], x, y, z);
            a.ow[o] = 1.0f;
            a.ovec[3 * o] = x;
            a.ovec[3 * o + 1] = y;
            a.ovec[3 * o + 2] = z;
        }
        else
        {
            a.ow[o] = a.w[e];
            a.ovec[3 * o] = a.vec[3 * e];
            a.ovec[3 * o + 1] = a.vec[3 * e + 1];
            a.ovec[3 * o + 2] = a.vec[3 * e + 2];
        }
    }
}

template<int P, bool RAD>
int seg_launch(const fa::DevBox& b, const SegArgs& a)
{
    constexpr int W = seg_waves<P>();
    unsigned const blocks =
        unsigned(std::max<uint64_t>(1, std::min<uint64_t>((uint64_t(a.nq) + W - 1) / W, kSegBlocks)));
    FA_KLAUNCH((seg_filter_kernel<P, RAD>), dim3(blocks), dim3(64 * W), 0, fa::stream(), b, a);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

// the count pass for the longest segment maxlen
template<bool RAD>
int seg_count(uint32_t maxlen, const fa::DevBox& b, const SegArgs& a)
{
    if (maxlen <= 64)
    {
        unsigned const blocks =
            unsigned(std::max<uint64_t>(1, std::min<uint64_t>((uint64_t(a.nq) + 1) / 2, kSegBlocks)));
        // separate launch sites, so that the library's timings name the two filters apart
        if (RAD)
            FA_KLAUNCH(seg_pair_kernel<true>, dim3(blocks), dim3(64), 0, fa::stream(), b, a);
        else
            FA_KLAUNCH(seg_pair_kernel<false>, dim3(blocks), dim3(64), 0, fa::stream(), b, a);
        FA_HIP(hipGetLastError());
        return FA_OK;
    }
    if (maxlen <= 128)
        return seg_launch<2, RAD>(b, a);
    if (maxlen <= 256)
        return seg_launch<4, RAD>(b, a);
    return seg_launch<8, RAD>(b, a);
}

// The per-segment filter of nl into *out; *done = false (nothing written)
// when some segment is longer than kSegMax bonds.
int filter_segments(bool rad, const fa_nlist* nl, const fa::DevBox& b, const float* pts, const float* qpts,
                    const float* radii, int terminate, fa_nlist* out, uint8_t* d_unfilled, uint32_t* n_unfilled,
                    bool* done)
{
    *done = false;
    hipStream_t s = fa::stream();
    uint32_t const nq = nl->num_query_points;
    void* tmp = nullptr;
    uint64_t const nb = nl->num_bonds;
    size_t const kept_bytes = (sizeof(uint32_t) * (size_t(nq) + 1) + 255) & ~size_t(255);
    if (int rc = fa::scratch(24, 256 + kept_bytes + sizeof(uint16_t) * (nb + 1), &tmp))
        return rc;
    uint32_t* counters = static_cast<uint32_t*>(tmp);
    uint32_t* kept = counters + 64;
    uint16_t* code = reinterpret_cast<uint16_t*>(static_cast<char*>(tmp) + 256 + kept_bytes);
    FA_HIP(hipMemsetAsync(counters, 0, 4 * sizeof(uint32_t), s));
    uint32_t maxlen = 0;
    if (nq)
    {
        // the longest segment into counters[3] (block_counters maxes into its second word)
        FA_KLAUNCH(seg_max_kernel, dim3(std::min(nblk(nq), 256u)), dim3(kBlock), 0, s, nq, nl->offsets,
                   counters + 2);
        FA_HIP(hipGetLastError());
        if (int rc = fa::read_u32(counters + 3, &maxlen))
            return rc;
    }
    if (maxlen > kSegMax)
        return FA_OK;
    SegArgs a{};
    a.nq = nq;
    a.offsets = nl->offsets;
    a.nbr = nl->neighbors;
    a.dist = nl->distances;
    a.w = nl->weights;
    a.vec = nl->vectors;
    a.pts = pts;
    a.qpts = qpts;
    a.radii = radii;
    a.terminate = terminate;
    a.kept = kept;
    a.unfilled = d_unfilled;
    a.counters = counters;
    a.code = code;
    if (nq)
        if (int rc = rad ? seg_count<true>(maxlen, b, a) : seg_count<false>(maxlen, b, a))
            return rc;
    // the scan of the kept counts is the output's offsets array
    if (int rc = fa::nlist_reserve(out, 0, nq))
        return rc;
    uint64_t total = 0;
    uint32_t h[2] = {0, 0};
    if (int rc = fa::exclusive_scan_u32(kept, out->offsets, nq, &total, counters, h, 2))
        return rc;
    if (int rc = fa::nlist_reserve(out, total, nq))
        return rc;
    if (total)
    {
        a.base = out->offsets;
        a.onbr = out->neighbors;
        a.odist = out->distances;
        a.ow = out->weights;
        a.ovec = out->vectors;
        unsigned const blocks = unsigned(std::min<uint64_t>(nblk(nb), 8192));
        if (rad)
            FA_KLAUNCH(seg_emit_kernel<true>, dim3(blocks), dim3(kBlock), 0, s, b, a, nb);
        else
            FA_KLAUNCH(seg_emit_kernel<false>, dim3(blocks), dim3(kBlock), 0, s, b, a, nb);
        FA_HIP(hipGetLastError());
    }
    out->num_bonds = total;
    // NeighborList.cc:130-173: one past the largest kept indices
    out->num_query_points = total ? h[0] : nq;
    out->num_points = total ? h[1] : nl->num_points;
    if (int rc = fa::nlist_finish_counts(out))
        return rc;
    if (int rc = fa::read_u32(counters + 2, n_unfilled))
        return rc;
    *done = true;
    return FA_OK;
}

} // namespace

extern "C" {

int fa_filter_sann(const fa_nlist* nl, fa_nlist* out, uint8_t* d_unfilled, uint32_t* n_unfilled)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    if (int rc = filter_check(nl, out, d_unfilled, n_unfilled))
        return rc;
    bool done = false;
    if (int rc = filter_segments(false, nl, fa::DevBox{}, nullptr, nullptr, nullptr, 0, out, d_unfilled, n_unfilled,
                                 &done))
        return rc;
    if (done)
        return FA_OK;
    // segments longer than the per-segment path takes: sorted copy, then filter
    uint32_t *flag = nullptr, *counters = nullptr;
    void* extra = nullptr;
    if (int rc = filter_setup(nl, out, d_unfilled, n_unfilled, sizeof(uint32_t) * (size_t(nl->num_query_points) + 1),
                              &flag, &extra, &counters))
        return rc;
    hipStream_t s = fa::stream();
    uint32_t* kept = static_cast<uint32_t*>(extra);
    uint32_t const nq = out->num_query_points;
    uint64_t const nb = out->num_bonds;
    if (nq)
    {
        FA_KLAUNCH(sann_kernel, dim3(filter_grid(nq)), dim3(kBlock), 0, s, nq, out->offsets, out->distances, kept,
                   d_unfilled, counters);
        FA_HIP(hipGetLastError());
    }
    if (nb)
    {
        FA_KLAUNCH(sann_flags_kernel, dim3(filter_grid(nb)), dim3(kBlock), 0, s, nb, out->neighbors, out->offsets,
                   kept, flag, counters);
        FA_HIP(hipGetLastError());
    }
    return filter_finish(out, flag, counters, n_unfilled);
}

int fa_filter_rad(const fa_box* box, const float* d_points, uint32_t n_points, const float* d_query_points,
                  const float* d_points_radii, const fa_nlist* nl, int terminate_after_blocked, fa_nlist* out,
                  uint8_t* d_unfilled, uint32_t* n_unfilled)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    if (!box || !nl)
        return fa::fail(FA_ERR_INVALID, "filter: box and nl are required");
    // NeighborList::validate (NeighborList.cc:348-358) as makeDefaultNlist applies it
    if (nl->num_points != n_points)
        return fa::fail(FA_ERR_INVALID, "NeighborList found inconsistent array sizes.");
    if (nl->num_bonds && (!d_points || !d_query_points))
        return fa::fail(FA_ERR_INVALID, "filter: points and query points are required");
    if (int rc = filter_check(nl, out, d_unfilled, n_unfilled))
        return rc;
    fa::DevBox const b = fa::make_devbox(*box);
    bool done = false;
    if (int rc = filter_segments(true, nl, b, d_points, d_query_points, d_points_radii, terminate_after_blocked, out,
                                 d_unfilled, n_unfilled, &done))
        return rc;
    if (done)
        return FA_OK;
    // segments longer than the per-segment path takes: sorted copy, then filter
    uint32_t *flag = nullptr, *counters = nullptr;
    void* extra = nullptr;
    if (int rc = filter_setup(nl, out, d_unfilled, n_unfilled, sizeof(float) * (nl->num_bonds + 1), &flag, &extra,
                              &counters))
        return rc;
    hipStream_t s = fa::stream();
    float* rsq = static_cast<float*>(extra);
    uint32_t const nq = out->num_query_points;
    uint64_t const nb = out->num_bonds;
    if (nb)
    {
        FA_KLAUNCH(rad_vectors_kernel, dim3(filter_grid(nb)), dim3(kBlock), 0, s, b, nb,
                   out->neighbors, d_points, d_query_points, d_points_radii, out->vectors, out->weights, rsq);
        FA_HIP(hipGetLastError());
    }
    if (nq)
    {
        FA_KLAUNCH(rad_kernel, dim3(filter_grid(uint64_t(nq) * 64)), dim3(kBlock), 0, s, nq, out->offsets,
                   out->neighbors, out->distances, out->vectors, rsq, terminate_after_blocked, flag, d_unfilled,
                   counters);
        FA_HIP(hipGetLastError());
    }
    return filter_finish(out, flag, counters, n_unfilled);
}

} // extern "C"

