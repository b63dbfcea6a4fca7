// fa_rdf.hip -- radial distribution function (freud/density/RDF.cc).
//
// The fused path never materialises bonds: every query point walks its cell
// stencil and bins each accepted pair into a block-private uint32 histogram in
// LDS (the reference's per-thread ThreadLocalHistogram, Histogram.h:207-283);
// each block merges its histogram into the device counts once.  Integer counts
// make the result independent of the order of accumulation, so the bin counts
// are bit-identical to the reference's.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fa_common.h"

namespace {

using fa::AABB_FORMULA;
using fa::DevBox;
using fa::DevGrid;
using fa::LINKCELL_FORMULA;

constexpr int kBlock = 256;
constexpr uint32_t kMaxLdsBins = 8192;

// freud/util/Histogram.h:152-174 RegularAxis::bin; returns bins for overflow
__device__ inline uint32_t regular_bin(float value, float vmin, float vmax, float inv_width, uint32_t nbins)
{
    if (value < vmin || value >= vmax)
        return nbins;
    float const val = (value - vmin) * inv_width;
    uint32_t bin = uint32_t(val);
    if (bin == nbins)
        return bin - 1;
    return bin;
}

// Query points taken in cell order straight from the grid (query points ==
// points): consecutive lanes share stencils, so candidate reads hit in cache.
template<int FORMULA, bool LDS_HIST>
__global__ __launch_bounds__(kBlock) void rdf_self_kernel(DevBox b, DevGrid g, uint32_t n, float q_rmax_sq,
                                                          float q_rmin_sq, int exclude_ii, int R, uint32_t nbins,
                                                          float vmin, float vmax, float inv_width,
                                                          uint32_t* __restrict__ counts)
{
    extern __shared__ uint32_t hist[];
    if (LDS_HIST)
    {
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            hist[t] = 0;
        __syncthreads();
    }
    uint32_t* h = LDS_HIST ? hist : counts;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
    {
        float4 const q = g.cpos[t];
        uint32_t const qi = __float_as_uint(q.w);
        fa::for_each_candidate<FORMULA>(b, g, q.x, q.y, q.z, R, false,
                                        [&](uint32_t j, float, float, float, float r_sq, int) {
                                            if (exclude_ii && j == qi)
                                                return;
                                            if (r_sq < q_rmax_sq && r_sq >= q_rmin_sq)
                                            {
                                                uint32_t const bin = regular_bin(sqrtf(r_sq), vmin, vmax, inv_width,
                                                                                 nbins);
                                                if (bin < nbins)
                                                    atomicAdd(&h[bin], 1u);
                                            }
                                        });
    }
    if (LDS_HIST)
    {
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            if (hist[t])
                atomicAdd(&counts[t], hist[t]);
    }
}

// Distinct query points (query_points != points), in caller order.
template<int FORMULA, bool LDS_HIST>
__global__ __launch_bounds__(kBlock) void rdf_query_kernel(DevBox b, DevGrid g, const float* __restrict__ qpts,
                                                           uint32_t nq, float q_rmax_sq, float q_rmin_sq,
                                                           int exclude_ii, int R, uint32_t nbins, float vmin,
                                                           float vmax, float inv_width, uint32_t* __restrict__ counts)
{
    extern __shared__ uint32_t hist[];
    if (LDS_HIST)
    {
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            hist[t] = 0;
        __syncthreads();
    }
    uint32_t* h = LDS_HIST ? hist : counts;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x)
    {
        float const qx = qpts[3 * i], qy = qpts[3 * i + 1], qz = qpts[3 * i + 2];
        fa::for_each_candidate<FORMULA>(b, g, qx, qy, qz, R, false,
                                        [&](uint32_t j, float, float, float, float r_sq, int) {
                                            if (exclude_ii && j == i)
                                                return;
                                            if (r_sq < q_rmax_sq && r_sq >= q_rmin_sq)
                                            {
                                                uint32_t const bin = regular_bin(sqrtf(r_sq), vmin, vmax, inv_width,
                                                                                 nbins);
                                                if (bin < nbins)
                                                    atomicAdd(&h[bin], 1u);
                                            }
                                        });
    }
    if (LDS_HIST)
    {
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            if (hist[t])
                atomicAdd(&counts[t], hist[t]);
    }
}

// Histogram of an existing NeighborList's distances.
template<bool LDS_HIST>
__global__ __launch_bounds__(kBlock) void rdf_nlist_kernel(const float* __restrict__ dist, uint64_t nb, uint32_t nbins,
                                                           float vmin, float vmax, float inv_width,
                                                           uint32_t* __restrict__ counts)
{
    extern __shared__ uint32_t hist[];
    if (LDS_HIST)
    {
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            hist[t] = 0;
        __syncthreads();
    }
    uint32_t* h = LDS_HIST ? hist : counts;
    for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < nb; e += uint64_t(gridDim.x) * blockDim.x)
    {
        uint32_t const bin = regular_bin(dist[e], vmin, vmax, inv_width, nbins);
        if (bin < nbins)
            atomicAdd(&h[bin], 1u);
    }
    if (LDS_HIST)
    {
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            if (hist[t])
                atomicAdd(&counts[t], hist[t]);
    }
}

int grid_blocks(uint64_t work)
{
    uint64_t const b = (work + kBlock - 1) / kBlock;
    return int(std::max<uint64_t>(1, std::min<uint64_t>(b, 4096)));
}

} // namespace

extern "C" {

int fa_rdf_accumulate_query(fa_nq* nq, const float* d_query_points, uint32_t n_query, float q_r_max, float q_r_min,
                            int exclude_ii, uint32_t bins, float r_min, float r_max, uint32_t* d_counts)
{
    if (int rc = fa_init())
        return rc;
    // RDF.cc:27-42 argument checks happen in the constructor (host side)
    fa_query_args a{FA_MODE_BALL, 0xffffffffu, q_r_max, q_r_min, -1.0f, -1.0f, exclude_ii};
    if (int rc = fa::validate_query_args(nq, a, n_query))
        return rc;
    if (n_query == 0)
        return FA_OK;
    hipStream_t s = fa::stream();
    int const R = 1;
    int dims[3];
    fa::grid_dims_for(nq->box, nq->n, q_r_max, R, dims);
    if (int rc = fa::ensure_grid(nq, dims))
        return rc;
    DevGrid const g = nq->cur->dev();
    DevBox const& db = nq->cur->dbox;
    float const width = (r_max - r_min) / float(bins);
    float const inv_width = 1.0f / width;
    float const qmax_sq = q_r_max * q_r_max, qmin_sq = q_r_min * q_r_min;
    bool const lds = bins <= kMaxLdsBins;
    size_t const shmem = lds ? sizeof(uint32_t) * bins : 0;
    bool const self = (d_query_points == nq->d_points) && (n_query == nq->n);
    int const blocks = grid_blocks(n_query);
#define FA_RDF_LAUNCH(KERNEL, FORM, ...)                                                                      \
    do                                                                                                        \
    {                                                                                                         \
        if (lds)                                                                                              \
            FA_KLAUNCH((KERNEL<FORM, true>), dim3(blocks), dim3(kBlock), shmem, s, __VA_ARGS__);      \
        else                                                                                                  \
            FA_KLAUNCH((KERNEL<FORM, false>), dim3(blocks), dim3(kBlock), 0, s, __VA_ARGS__);         \
    } while (0)
    if (self)
    {
        if (nq->kind == FA_NQ_AABB)
            FA_RDF_LAUNCH(rdf_self_kernel, AABB_FORMULA, db, g, n_query, qmax_sq, qmin_sq, exclude_ii, R, bins, r_min,
                          r_max, inv_width, d_counts);
        else
            FA_RDF_LAUNCH(rdf_self_kernel, LINKCELL_FORMULA, db, g, n_query, qmax_sq, qmin_sq, exclude_ii, R, bins,
                          r_min, r_max, inv_width, d_counts);
    }
    else
    {
        if (nq->kind == FA_NQ_AABB)
            FA_RDF_LAUNCH(rdf_query_kernel, AABB_FORMULA, db, g, d_query_points, n_query, qmax_sq, qmin_sq,
                          exclude_ii, R, bins, r_min, r_max, inv_width, d_counts);
        else
            FA_RDF_LAUNCH(rdf_query_kernel, LINKCELL_FORMULA, db, g, d_query_points, n_query, qmax_sq, qmin_sq,
                          exclude_ii, R, bins, r_min, r_max, inv_width, d_counts);
    }
#undef FA_RDF_LAUNCH
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int fa_rdf_accumulate_nlist(const fa_nlist* nl, uint32_t bins, float r_min, float r_max, uint32_t* d_counts)
{
    if (int rc = fa_init())
        return rc;
    if (nl->num_bonds == 0)
        return FA_OK;
    float const width = (r_max - r_min) / float(bins);
    float const inv_width = 1.0f / width;
    bool const lds = bins <= kMaxLdsBins;
    int const blocks = grid_blocks(nl->num_bonds);
    if (lds)
        FA_KLAUNCH(rdf_nlist_kernel<true>, dim3(blocks), dim3(kBlock), sizeof(uint32_t) * bins, fa::stream(),
                           nl->distances, nl->num_bonds, bins, r_min, r_max, inv_width, d_counts);
    else
        FA_KLAUNCH(rdf_nlist_kernel<false>, dim3(blocks), dim3(kBlock), 0, fa::stream(), nl->distances,
                           nl->num_bonds, bins, r_min, r_max, inv_width, d_counts);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

// RDF.cc:44-63 (shell volumes) and RDF.cc:73-99 (normalisation), float order
// of the reference.
int fa_rdf_reduce(const fa_box* box, const uint32_t* counts, uint32_t bins, float r_min, float r_max,
                  uint32_t n_points, uint32_t n_query_points, uint32_t n_frames, int finite_size, float* rdf,
                  float* n_r)
{
    if (bins == 0)
        return fa::fail(FA_ERR_INVALID, "RDF requires a nonzero number of bins.");
    float const width = (r_max - r_min) / float(bins);
    std::vector<float> edges(bins + 1);
    for (uint32_t i = 0; i <= bins; ++i)
    {
        volatile float t = float(i) * width;
        edges[i] = r_min + t;
    }
    float const volume_prefactor = float((float(4.0) / float(3.0)) * M_PI);
    std::vector<float> vol(bins);
    for (uint32_t i = 0; i < bins; ++i)
    {
        float const r = edges[i], nr = edges[i + 1];
        if (box->is2D)
        {
            volatile float a = nr * nr;
            volatile float c = r * r;
            vol[i] = float(M_PI * double(float(a - c)));
        }
        else
        {
            volatile float a = nr * nr;
            volatile float a3 = a * nr;
            volatile float c = r * r;
            volatile float c3 = c * r;
            volatile float d = a3 - c3;
            vol[i] = volume_prefactor * d;
        }
    }
    float const nqp = float(n_query_points);
    float number_density = nqp / fa::box_volume(*box);
    if (finite_size)
        number_density *= float(n_query_points - 1) / float(n_query_points);
    float const npf = float(n_points);
    float const nf = float(n_frames);
    volatile float den = npf * number_density;
    volatile float den2 = den * nf;
    float prefactor = float(1.0) / den2;
    for (uint32_t i = 0; i < bins; ++i)
    {
        volatile float t = float(counts[i]) * prefactor;
        rdf[i] = t / vol[i];
    }
    volatile float den3 = nqp * float(n_frames);
    prefactor = float(1.0) / den3;
    {
        volatile float t = float(counts[0]) * prefactor;
        n_r[0] = t;
    }
    for (uint32_t i = 1; i < bins; ++i)
    {
        volatile float t = float(counts[i]) * prefactor;
        n_r[i] = n_r[i - 1] + t;
    }
    return FA_OK;
}

} // extern "C"
