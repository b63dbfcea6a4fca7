// fa_rdf.hip -- radial distribution function (freud/density/RDF.cc).
//
// The fused paths never materialise bonds; pairs are binned into a
// block-private uint32 histogram in LDS (the reference's per-thread
// ThreadLocalHistogram, Histogram.h:207-283) that is merged into the device
// counts once per block.  Integer counts make the result independent of the
// accumulation order, so bin counts are bit-identical to the reference's.
//
// Binning without a square root: the reference bins d = sqrtf(r_sq) with
// RegularAxis::bin (Histogram.h:152-174).  That map from the float r_sq to a
// bin is monotone, so it is fully described by nbins + 1 float thresholds
// T[k] = min{ r_sq : bin(r_sq) >= k } (T[0]: d >= r_min, T[nbins]: d >= r_max),
// found on the host by bisection over float bit patterns with the reference's
// own expression.  A kernel estimates the bin with the hardware square root and
// corrects it against T, giving exactly the reference's bin.
//
// Self RDF (query points == points, AABB formula, all points inside the box):
// a cluster-pair kernel over cells at least r_max/2 wide; each wave takes a
// cluster of up to 8 cell-ordered points and its 64 lanes test 8 x 8 pairs per
// step against 8-point chunks of the 5x5x5 stencil cells.  For two cells related
// without a periodic shift the AABB formula gives bitwise-equal r_sq for (i,j)
// and (j,i) (x - y == -(y - x) in IEEE arithmetic), so such pairs are evaluated
// once with weight 2; pairs across the box edge are evaluated in both
// directions because (q + image) rounds differently for the two directions.
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
constexpr uint32_t kMaxLdsBins = 4096;

// ------------------------------------------------------------------ thresholds
// reference bin of a float r_sq: -1 below r_min, nbins at or above r_max
int ref_bin_of_rsq(float x, float vmin, float vmax, float inv_width, uint32_t nbins)
{
    volatile float d = std::sqrt(x);
    if (d < vmin)
        return -1;
    if (d >= vmax)
        return int(nbins);
    volatile float t = d - vmin;
    volatile float v = t * inv_width;
    size_t bin = (size_t) (float(v));
    if (bin == nbins)
        bin = nbins - 1;
    return int(bin);
}

float first_rsq_with_bin_at_least(int k, float vmin, float vmax, float inv_width, uint32_t nbins)
{
    uint32_t lo = 0, hi = 0x7f800000u; // [+0, +inf]
    if (ref_bin_of_rsq(__builtin_bit_cast(float, hi), vmin, vmax, inv_width, nbins) < k)
        return __builtin_bit_cast(float, hi);
    while (lo < hi)
    {
        uint32_t const mid = lo + (hi - lo) / 2;
        if (ref_bin_of_rsq(__builtin_bit_cast(float, mid), vmin, vmax, inv_width, nbins) >= k)
            hi = mid;
        else
            lo = mid + 1;
    }
    return __builtin_bit_cast(float, lo);
}

struct Thresholds
{
    std::vector<float> host;
    float* dev = nullptr;
    uint32_t nbins = 0;
    float vmin = 0, vmax = 0;
};
Thresholds g_thr;

int thresholds(uint32_t nbins, float vmin, float vmax, float** d_out)
{
    if (g_thr.dev && g_thr.nbins == nbins && g_thr.vmin == vmin && g_thr.vmax == vmax)
    {
        *d_out = g_thr.dev;
        return FA_OK;
    }
    float const width = (vmax - vmin) / float(nbins);
    float const inv_width = 1.0f / width;
    std::vector<float> T(nbins + 1);
    for (uint32_t k = 0; k <= nbins; ++k)
        T[k] = first_rsq_with_bin_at_least(int(k), vmin, vmax, inv_width, nbins);
    if (!g_thr.dev || g_thr.nbins < nbins)
    {
        fa::device_free(g_thr.dev);
        g_thr.dev = nullptr;
        if (int rc = fa::device_alloc((void**) &g_thr.dev, sizeof(float) * (nbins + 1)))
            return rc;
    }
    FA_HIP(hipMemcpyAsync(g_thr.dev, T.data(), sizeof(float) * (nbins + 1), hipMemcpyHostToDevice, fa::stream()));
    g_thr.host = T;
    g_thr.nbins = nbins;
    g_thr.vmin = vmin;
    g_thr.vmax = vmax;
    *d_out = g_thr.dev;
    return FA_OK;
}

// The reference's bin of r_sq, given T in LDS; nbins when outside [T0, Tn).
__device__ inline uint32_t bin_of_rsq(float x, const float* __restrict__ T, uint32_t nbins, float vmin,
                                      float inv_width)
{
    if (!(x >= T[0]) || !(x < T[nbins]))
        return nbins;
    float const d = __builtin_amdgcn_sqrtf(x);
    float v = (d - vmin) * inv_width;
    int k = int(fminf(fmaxf(v, 0.0f), float(nbins - 1)));
#pragma unroll
    for (int it = 0; it < 2; ++it)
    {
        if (k > 0 && x < T[k])
            --k;
        if (k < int(nbins) - 1 && x >= T[k + 1])
            ++k;
    }
    return uint32_t(k);
}

template<bool LDS_HIST>
__device__ inline void hist_prologue(uint32_t* hist, float* sT, const float* __restrict__ T, uint32_t nbins)
{
    for (uint32_t t = threadIdx.x; t <= nbins; t += blockDim.x)
    {
        sT[t] = T[t];
        if (LDS_HIST && t < nbins)
            hist[t] = 0;
    }
    __syncthreads();
}

template<bool LDS_HIST> __device__ inline void hist_epilogue(const uint32_t* hist, uint32_t nbins, uint32_t* counts)
{
    if (!LDS_HIST)
        return;
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
        if (hist[t])
            atomicAdd(&counts[t], hist[t]);
}

// ------------------------------------------------------------------ self, cluster pairs
// Clusters: runs of up to 8 consecutive cell-ordered points of one cell.
__global__ void cluster_count_kernel(const uint32_t* __restrict__ cstart, uint32_t ncell, uint32_t* __restrict__ nclus)
{
    uint32_t const c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncell)
        nclus[c] = (cstart[c + 1] - cstart[c] + 7u) / 8u;
}

__global__ void cluster_table_kernel(const uint32_t* __restrict__ cstart, uint32_t ncell,
                                     const uint32_t* __restrict__ cl_off, uint2* __restrict__ table)
{
    uint32_t const c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell)
        return;
    uint32_t const beg = cstart[c], n = cstart[c + 1] - beg;
    uint32_t const o = cl_off[c];
    for (uint32_t k = 0; k * 8 < n; ++k)
        table[o + k] = make_uint2(c, ((beg + 8 * k) << 4) | min(8u, n - 8 * k));
}

// One wave per cluster I of cell C; lane = (a, b) = (lane / 8, lane % 8):
// query point a of I against point b of each 8-point chunk of every stencil
// cell (cells at least r_max / R wide).  Non-wrapping cell pairs are taken
// once (lexicographically positive offset, or a < b inside C) with weight 2.
template<int R, bool LDS_HIST>
__global__ __launch_bounds__(kBlock) void rdf_cluster_kernel(DevBox b, DevGrid g, const uint2* __restrict__ table,
                                                             const uint32_t* __restrict__ ncl_ptr, float q_rmax_sq,
                                                             float q_rmin_sq, const float* __restrict__ T,
                                                             uint32_t nbins, float vmin, float inv_width,
                                                             uint32_t* __restrict__ counts)
{
    extern __shared__ float smem[];
    float* sT = smem;
    uint32_t* hist = reinterpret_cast<uint32_t*>(smem + nbins + 1);
    uint32_t* h = LDS_HIST ? hist : counts;
    hist_prologue<LDS_HIST>(hist, sT, T, nbins);

    uint32_t const lane = threadIdx.x & 63u;
    uint32_t const I = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (I < *ncl_ptr)
    {
        int const dim0 = g.dim[0], dim1 = g.dim[1], dim2 = g.dim[2];
        uint2 const e = table[I];
        uint32_t const C = e.x;
        uint32_t const ia0 = e.y >> 4, na = e.y & 15u;
        int const cx0 = int(C % uint32_t(dim0));
        int const cy0 = int((C / uint32_t(dim0)) % uint32_t(dim1));
        int const cz0 = int(C / (uint32_t(dim0) * uint32_t(dim1)));
        uint32_t const a = lane >> 3, bl = lane & 7u;
        bool const a_ok = a < na;
        float4 pa = g.cpos[ia0 + (a_ok ? a : 0u)];
        if (b.is2D)
            pa.z = 0.0f;
        uint32_t const pos_a = ia0 + a;
        int const Rz = b.is2D ? 0 : R;
        for (int dz = -Rz; dz <= Rz; ++dz)
        {
            int const uz = cz0 + dz;
            if (uz < -dim2 || uz >= 2 * dim2)
                continue;
            int const sz = uz < 0 ? -1 : (uz >= dim2 ? 1 : 0);
            int const cz = uz - sz * dim2;
            for (int dy = -R; dy <= R; ++dy)
            {
                int const uy = cy0 + dy;
                if (uy < -dim1 || uy >= 2 * dim1)
                    continue;
                int const sy = uy < 0 ? -1 : (uy >= dim1 ? 1 : 0);
                int const cy = uy - sy * dim1;
                for (int dx = -R; dx <= R; ++dx)
                {
                    int const ux = cx0 + dx;
                    if (ux < -dim0 || ux >= 2 * dim0)
                        continue;
                    int const sx = ux < 0 ? -1 : (ux >= dim0 ? 1 : 0);
                    bool const wraps = (sx | sy | sz) != 0;
                    bool const self = !wraps && dx == 0 && dy == 0 && dz == 0;
                    if (!wraps && !self && !(dz > 0 || (dz == 0 && (dy > 0 || (dy == 0 && dx > 0)))))
                        continue;
                    uint32_t const weight = wraps ? 1u : 2u;
                    int const cx = ux - sx * dim0;
                    uint32_t const B = uint32_t((cz * dim1 + cy) * dim0 + cx);
                    uint32_t const b_beg = g.cstart[B];
                    uint32_t const b_end = g.cstart[B + 1];
                    int const code = (-sx + 1) * 9 + (-sy + 1) * 3 + (-sz + 1);
                    // AABB formula: r = p_j - (q_i + image)
                    float const qx = pa.x + b.img[code][0];
                    float const qy = pa.y + b.img[code][1];
                    float const qz = pa.z + b.img[code][2];
                    for (uint32_t jb = self ? ia0 : b_beg; jb < b_end; jb += 8)
                    {
                        uint32_t const pos_b = jb + bl;
                        bool valid = a_ok && pos_b < b_end;
                        if (self)
                            valid = valid && pos_a < pos_b;
                        float4 pj = g.cpos[valid ? pos_b : b_beg];
                        float const pjz = b.is2D ? 0.0f : pj.z;
                        float const rx = pj.x - qx;
                        float const ry = pj.y - qy;
                        float const rz = pjz - qz;
                        float const x = (rx * rx + ry * ry) + rz * rz;
                        if (valid && x < q_rmax_sq && x >= q_rmin_sq)
                        {
                            uint32_t const bin = bin_of_rsq(x, sT, nbins, vmin, inv_width);
                            if (bin < nbins)
                                atomicAdd(&h[bin], weight);
                        }
                    }
                }
            }
        }
    }
    hist_epilogue<LDS_HIST>(hist, nbins, counts);
}

// ------------------------------------------------------------------ generic (per query point)
// Query points in cell order straight from the grid (query points == points).
template<int FORMULA, bool LDS_HIST>
__global__ __launch_bounds__(kBlock) void rdf_self_kernel(DevBox b, DevGrid g, uint32_t n, float q_rmax_sq,
                                                          float q_rmin_sq, int exclude_ii, int R,
                                                          const float* __restrict__ T, uint32_t nbins, float vmin,
                                                          float inv_width, uint32_t* __restrict__ counts)
{
    extern __shared__ float smem[];
    float* sT = smem;
    uint32_t* hist = reinterpret_cast<uint32_t*>(smem + nbins + 1);
    uint32_t* h = LDS_HIST ? hist : counts;
    hist_prologue<LDS_HIST>(hist, sT, T, nbins);
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
                                                uint32_t const bin = bin_of_rsq(r_sq, sT, nbins, vmin, inv_width);
                                                if (bin < nbins)
                                                    atomicAdd(&h[bin], 1u);
                                            }
                                        });
    }
    hist_epilogue<LDS_HIST>(hist, nbins, counts);
}

// Distinct query points (query_points != points), in caller order.
template<int FORMULA, bool LDS_HIST>
__global__ __launch_bounds__(kBlock) void rdf_query_kernel(DevBox b, DevGrid g, const float* __restrict__ qpts,
                                                           uint32_t nq, float q_rmax_sq, float q_rmin_sq,
                                                           int exclude_ii, int R, const float* __restrict__ T,
                                                           uint32_t nbins, float vmin, float inv_width,
                                                           uint32_t* __restrict__ counts)
{
    extern __shared__ float smem[];
    float* sT = smem;
    uint32_t* hist = reinterpret_cast<uint32_t*>(smem + nbins + 1);
    uint32_t* h = LDS_HIST ? hist : counts;
    hist_prologue<LDS_HIST>(hist, sT, T, nbins);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x)
    {
        float const qx = qpts[3 * i], qy = qpts[3 * i + 1], qz = qpts[3 * i + 2];
        fa::for_each_candidate<FORMULA>(b, g, qx, qy, qz, R, false,
                                        [&](uint32_t j, float, float, float, float r_sq, int) {
                                            if (exclude_ii && j == i)
                                                return;
                                            if (r_sq < q_rmax_sq && r_sq >= q_rmin_sq)
                                            {
                                                uint32_t const bin = bin_of_rsq(r_sq, sT, nbins, vmin, inv_width);
                                                if (bin < nbins)
                                                    atomicAdd(&h[bin], 1u);
                                            }
                                        });
    }
    hist_epilogue<LDS_HIST>(hist, nbins, counts);
}

// freud/util/Histogram.h:152-174 RegularAxis::bin on a distance; nbins for overflow
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

// Histogram of an existing NeighborList's distances.
template<bool LDS_HIST>
__global__ __launch_bounds__(kBlock) void rdf_nlist_kernel(const float* __restrict__ dist, uint64_t nb, uint32_t nbins,
                                                           float vmin, float vmax, float inv_width,
                                                           uint32_t* __restrict__ counts)
{
    extern __shared__ uint32_t hist_only[];
    if (LDS_HIST)
    {
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            hist_only[t] = 0;
        __syncthreads();
    }
    uint32_t* h = LDS_HIST ? hist_only : counts;
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
            if (hist_only[t])
                atomicAdd(&counts[t], hist_only[t]);
    }
}

__global__ void add_self_pairs_kernel(uint32_t* counts, uint32_t bin, uint32_t n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        counts[bin] += n;
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
    float const width = (r_max - r_min) / float(bins);
    float const inv_width = 1.0f / width;
    float const qmax_sq = q_r_max * q_r_max, qmin_sq = q_r_min * q_r_min;
    float* T = nullptr;
    if (int rc = thresholds(bins, r_min, r_max, &T))
        return rc;
    bool const lds = bins <= kMaxLdsBins;
    size_t const shmem = sizeof(float) * (bins + 1) + (lds ? sizeof(uint32_t) * bins : 0);
    bool const self = (d_query_points == nq->d_points) && (n_query == nq->n);

    if (self && nq->kind == FA_NQ_AABB)
    {
        // cluster-pair kernel over cells at least r_max/2 wide
        int const R = 2;
        int dims[3];
        fa::grid_dims_for(nq->box, nq->n, q_r_max, R, dims);
        if (int rc = fa::ensure_grid(nq, dims))
            return rc;
        if (!nq->cur->any_outside)
        {
            DevGrid const g = nq->cur->dev();
            DevBox const& db = nq->cur->dbox;
            uint32_t const ncell = uint32_t(dims[0]) * dims[1] * dims[2];
            uint64_t const max_clusters = uint64_t(nq->n) / 8 + ncell;
            void* tmp = nullptr;
            size_t const cb = sizeof(uint32_t) * (size_t(ncell) + 1);
            if (int rc = fa::scratch(13, 2 * cb + sizeof(uint2) * max_clusters + 256, &tmp))
                return rc;
            uint32_t* nclus = static_cast<uint32_t*>(tmp);
            uint32_t* cl_off = nclus + (ncell + 1);
            uint2* table = reinterpret_cast<uint2*>(static_cast<char*>(tmp) + ((2 * cb + 15) & ~size_t(15)));
            unsigned const cblocks = (ncell + kBlock - 1) / kBlock;
            FA_KLAUNCH(cluster_count_kernel, dim3(cblocks), dim3(kBlock), 0, s, g.cstart, ncell, nclus);
            if (int rc = fa::exclusive_scan_u32(nclus, cl_off, ncell, nullptr))
                return rc;
            FA_KLAUNCH(cluster_table_kernel, dim3(cblocks), dim3(kBlock), 0, s, g.cstart, ncell, cl_off, table);
            unsigned const wblocks = unsigned((max_clusters + kBlock / 64 - 1) / (kBlock / 64));
            if (lds)
                FA_KLAUNCH((rdf_cluster_kernel<2, true>), dim3(wblocks), dim3(kBlock), shmem, s, db, g, table,
                           cl_off + ncell, qmax_sq, qmin_sq, T, bins, r_min, inv_width, d_counts);
            else
                FA_KLAUNCH((rdf_cluster_kernel<2, false>), dim3(wblocks), dim3(kBlock), shmem, s, db, g, table,
                           cl_off + ncell, qmax_sq, qmin_sq, T, bins, r_min, inv_width, d_counts);
            FA_HIP(hipGetLastError());
            if (!exclude_ii && qmin_sq <= 0.0f)
            {
                // (i, i) pairs at distance 0 (NeighborQuery.h:317-328 keeps them when exclude_ii is off)
                int const b0 = ref_bin_of_rsq(0.0f, r_min, r_max, inv_width, bins);
                if (b0 >= 0 && b0 < int(bins))
                    FA_KLAUNCH(add_self_pairs_kernel, dim3(1), dim3(64), 0, s, d_counts, uint32_t(b0), n_query);
            }
            return FA_OK;
        }
    }
    int const R = 1;
    int dims[3];
    fa::grid_dims_for(nq->box, nq->n, q_r_max, R, dims);
    if (int rc = fa::ensure_grid(nq, dims))
        return rc;
    DevGrid const g = nq->cur->dev();
    DevBox const& db = nq->cur->dbox;
    int const blocks = grid_blocks(n_query);
#define FA_RDF_LAUNCH(KERNEL, FORM, ...)                                                          \
    do                                                                                            \
    {                                                                                             \
        if (lds)                                                                                  \
            FA_KLAUNCH((KERNEL<FORM, true>), dim3(blocks), dim3(kBlock), shmem, s, __VA_ARGS__);  \
        else                                                                                      \
            FA_KLAUNCH((KERNEL<FORM, false>), dim3(blocks), dim3(kBlock), shmem, s, __VA_ARGS__); \
    } while (0)
    if (self)
    {
        if (nq->kind == FA_NQ_AABB)
            FA_RDF_LAUNCH(rdf_self_kernel, AABB_FORMULA, db, g, n_query, qmax_sq, qmin_sq, exclude_ii, R, T, bins,
                          r_min, inv_width, d_counts);
        else
            FA_RDF_LAUNCH(rdf_self_kernel, LINKCELL_FORMULA, db, g, n_query, qmax_sq, qmin_sq, exclude_ii, R, T, bins,
                          r_min, inv_width, d_counts);
    }
    else
    {
        if (nq->kind == FA_NQ_AABB)
            FA_RDF_LAUNCH(rdf_query_kernel, AABB_FORMULA, db, g, d_query_points, n_query, qmax_sq, qmin_sq,
                          exclude_ii, R, T, bins, r_min, inv_width, d_counts);
        else
            FA_RDF_LAUNCH(rdf_query_kernel, LINKCELL_FORMULA, db, g, d_query_points, n_query, qmax_sq, qmin_sq,
                          exclude_ii, R, T, bins, r_min, inv_width, d_counts);
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
