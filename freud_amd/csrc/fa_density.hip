// fa_density.hip -- LocalDensity (freud/density/LocalDensity.cc) and PMFTXY
// (freud/pmft/PMFTXY.cc, freud/pmft/PMFT.h) over a NeighborList: the other
// consumers of the query path beside the RDF.
//
// LocalDensity is one pass over each query point's bond segment (the bonds of
// a query point are contiguous in a NeighborList), summing the reference's
// per-bond weight in bond order.  PMFTXY bins every bond vector, rotated into
// the query point's frame, into a 2D RegularAxis histogram
// (freud/util/Histogram.h:123-179) held in LDS per block and merged once per
// block; integer counts make the result independent of the merge order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fa_common.h"

namespace {

constexpr int kBlock = 256;
// LDS histograms up to this many bins (64 KB); larger ones count in device memory
constexpr uint32_t kMaxLdsBins2D = 16384;

// LocalDensity.cc:50-84.  Per query point, in bond order:
//   num += 1                                        if d < r_max - diameter / 2
//   num += 1 + (r_max - (d + diameter / 2)) / diameter   otherwise
// density = num / area (2D) or num / volume (3D); query points without bonds
// keep 0 (the reference's arrays start zeroed and are assigned per bond).
__global__ void local_density_kernel(const uint32_t* __restrict__ offsets, const float* __restrict__ dist,
                                     uint32_t nq, float r_max, float diameter, int is2D, float area, float volume,
                                     float* __restrict__ density, float* __restrict__ num_out)
{
    uint32_t const i = fa::xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= nq)
        return;
    uint32_t const beg = offsets[i], end = offsets[i + 1];
    float const half = diameter / 2.0f;
    float const inner = r_max - half;
    float num = 0.0f;
    for (uint32_t b = beg; b < end; ++b)
    {
        float const d = dist[b];
        if (d < inner)
            num += 1.0f;
        else
            num += 1.0f + (r_max - (d + half)) / diameter;
    }
    float dens = 0.0f;
    if (end > beg)
        dens = is2D ? num / area : num / volume;
    num_out[i] = num;
    density[i] = dens;
}

// RegularAxis::bin (Histogram.h:152-174): kOverflow outside [min, max)
constexpr uint32_t kOverflow = 0xffffffffu;
__device__ inline uint32_t axis_bin(float v, float vmin, float vmax, float inv_width, uint32_t n)
{
    if (v < vmin || v >= vmax)
        return kOverflow;
    float const val = (v - vmin) * inv_width;
    uint32_t const b = uint32_t(val);
    return b == n ? n - 1u : b;
}

struct Axis2
{
    float xmin, xmax, xinv;
    float ymin, ymax, yinv;
    uint32_t nx, ny;
};

// One bond: the vector (x, y) rotated by -theta of its query point
// (rotmat2 * vec2 = (dot(row0, v), dot(row1, v)), rows (c, -s) and (s, c)),
// then the row-major 2D bin (ManagedArray::getIndex).
__device__ inline uint32_t pmftxy_bin(const Axis2& A, float2 cs, float x, float y)
{
    float const c = cs.x, s = cs.y;
    float const rx = c * x + (-s) * y;
    float const ry = s * x + c * y;
    uint32_t const bx = axis_bin(rx, A.xmin, A.xmax, A.xinv, A.nx);
    uint32_t const by = axis_bin(ry, A.ymin, A.ymax, A.yinv, A.ny);
    if (bx == kOverflow || by == kOverflow)
        return kOverflow;
    return bx * A.ny + by;
}

template<bool LDS>
__global__ __launch_bounds__(kBlock) void pmftxy_kernel(const uint32_t* __restrict__ ij, const float* __restrict__ vec,
                                                        uint64_t nb, const float2* __restrict__ cs, Axis2 A,
                                                        uint32_t* __restrict__ counts)
{
    extern __shared__ uint32_t s_hist[];
    uint32_t const nbins = A.nx * A.ny;
    if (LDS)
    {
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            s_hist[t] = 0u;
        __syncthreads();
    }
    uint64_t const stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t b = uint64_t(fa::xcd_tile(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x; b < nb; b += stride)
    {
        uint32_t const i = ij[2 * b];
        uint32_t const bin = pmftxy_bin(A, cs[i], vec[3 * b], vec[3 * b + 1]);
        if (bin != kOverflow)
        {
            if (LDS)
                atomicAdd(&s_hist[bin], 1u);
            else
                atomicAdd(&counts[bin], 1u);
        }
    }
    if (LDS)
    {
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            if (s_hist[t])
                atomicAdd(&counts[t], s_hist[t]);
    }
}

int grid_for(uint64_t work, int cap)
{
    uint64_t const b = (work + kBlock - 1) / kBlock;
    return int(std::max<uint64_t>(1, std::min<uint64_t>(b, uint64_t(cap))));
}

} // namespace

extern "C" {

int fa_local_density(const fa_nlist* nl, float r_max, float diameter, int is2D, float* d_density,
                     float* d_num_neighbors)
{
    if (int rc = fa_init())
        return rc;
    // LocalDensity.cc:26-37
    if (r_max <= 0)
        return fa::fail(FA_ERR_INVALID, "LocalDensity requires r_max to be positive.");
    if (diameter < 0)
        return fa::fail(FA_ERR_INVALID, "LocalDensity requires diameter to be non-negative.");
    uint32_t const nq = nl->num_query_points;
    if (nq == 0)
        return FA_OK;
    // LocalDensity.cc:48-49: the area in double, rounded once; the volume in float
    float const area = float(M_PI * double(r_max) * double(r_max));
    volatile float v1 = float(4.0 / 3.0 * M_PI) * r_max;
    volatile float v2 = v1 * r_max;
    float const volume = v2 * r_max;
    unsigned const blocks = (nq + kBlock - 1) / kBlock;
    FA_KLAUNCH(local_density_kernel, dim3(blocks), dim3(kBlock), 0, fa::stream(), nl->offsets, nl->distances, nq,
               r_max, diameter, is2D, area, volume, d_density, d_num_neighbors);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int fa_pmftxy_accumulate(const fa_nlist* nl, const float* query_orientations, float x_max, float y_max,
                         uint32_t n_x, uint32_t n_y, uint32_t* d_counts)
{
    if (int rc = fa_init())
        return rc;
    // PMFTXY.cc:25-41
    if (n_x < 1)
        return fa::fail(FA_ERR_INVALID, "PMFTXY requires at least 1 bin in X.");
    if (n_y < 1)
        return fa::fail(FA_ERR_INVALID, "PMFTXY requires at least 1 bin in Y.");
    if (x_max < 0)
        return fa::fail(FA_ERR_INVALID, "PMFTXY requires that x_max must be positive.");
    if (y_max < 0)
        return fa::fail(FA_ERR_INVALID, "PMFTXY requires that y_max must be positive.");
    if (uint64_t(n_x) * n_y > 0xffffffffull)
        return fa::fail(FA_ERR_INVALID, "PMFTXY histogram too large");
    uint32_t const nq = nl->num_query_points;
    if (nl->num_bonds == 0 || nq == 0)
        return FA_OK;
    hipStream_t s = fa::stream();
    // cos(-theta), sin(-theta) per query point with the host C library's
    // float std::cos / std::sin, as rotmat2::fromAngle evaluates them
    // (VectorMath.h:912-921): those are not correctly rounded everywhere, and a
    // last-bit difference in the rotation can move a bond across a bin edge
    std::vector<float> h_cs(2 * size_t(nq));
    for (uint32_t i = 0; i < nq; ++i)
    {
        float const t = -query_orientations[i];
        h_cs[2 * i] = std::cos(t);
        h_cs[2 * i + 1] = std::sin(t);
    }
    float2* cs = nullptr;
    if (int rc = fa::scratch(20, sizeof(float2) * size_t(nq) + 256, (void**) &cs))
        return rc;
    FA_HIP(hipMemcpyAsync(cs, h_cs.data(), sizeof(float2) * size_t(nq), hipMemcpyHostToDevice, s));
    // the copy reads h_cs, which ends with this call
    FA_HIP(hipStreamSynchronize(s));
    // RegularAxis(n, -max, max) (Histogram.h:126-138)
    Axis2 A;
    A.xmin = -x_max;
    A.xmax = x_max;
    A.ymin = -y_max;
    A.ymax = y_max;
    volatile float wx = (x_max - A.xmin) / float(n_x);
    volatile float wy = (y_max - A.ymin) / float(n_y);
    A.xinv = 1.0f / wx;
    A.yinv = 1.0f / wy;
    A.nx = n_x;
    A.ny = n_y;
    uint32_t const nbins = n_x * n_y;
    uint64_t const nb = nl->num_bonds;
    if (nbins <= kMaxLdsBins2D)
        FA_KLAUNCH(pmftxy_kernel<true>, dim3(grid_for(nb, 2048)), dim3(kBlock), sizeof(uint32_t) * nbins, s,
                   nl->neighbors, nl->vectors, nb, cs, A, d_counts);
    else
        FA_KLAUNCH(pmftxy_kernel<false>, dim3(grid_for(nb, 4096)), dim3(kBlock), 0, s, nl->neighbors, nl->vectors,
                   nb, cs, A, d_counts);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int fa_pmftxy_reduce(const fa_box* box, const uint32_t* counts, uint32_t n_x, uint32_t n_y, float x_max, float y_max,
                     uint32_t n_points, uint32_t n_query_points, uint32_t n_frames, float* pcf)
{
    // PMFTXY.cc:48-52,62-66 and PMFT.h:76-87, float operation order of the reference
    volatile float dx = float(2.0) * x_max / float(n_x);
    volatile float dy = float(2.0) * y_max / float(n_y);
    volatile float jacobian = dx * dy;
    float const jacobian_factor = float(1.0) / jacobian;
    float const inv_num_dens = fa::box_volume(*box) / float(n_query_points);
    volatile float den = float(n_frames) * float(n_points);
    float const norm_factor = float(1.0) / den;
    float const prefactor = inv_num_dens * norm_factor;
    uint64_t const nbins = uint64_t(n_x) * n_y;
    for (uint64_t t = 0; t < nbins; ++t)
    {
        volatile float a = float(counts[t]) * prefactor;
        pcf[t] = a * jacobian_factor;
    }
    return FA_OK;
}

} // extern "C"
