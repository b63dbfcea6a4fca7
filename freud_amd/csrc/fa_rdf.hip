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

// The reference's bin of an r_sq already known to lie in [T[0], T[nbins]).
// The hardware square root is within an ulp of the correctly rounded one, so
// the estimate is within one bin of the reference's unless bins are narrower
// than that rounding (far beyond 65536 bins): one compare with each
// neighbouring threshold then gives the exact bin.
__device__ inline uint32_t bin_in_range(float x, const float* __restrict__ T, uint32_t nbins, float vmin,
                                        float inv_width)
{
    float const d = __builtin_amdgcn_sqrtf(x);
    float const v = (d - vmin) * inv_width;
    int k = int(fminf(fmaxf(v, 0.0f), float(nbins - 1)));
    float const t0 = T[k], t1 = T[k + 1];
    k += int(x >= t1) - int(x < t0);
    if (nbins > 65536u)
    {
        while (k > 0 && x < T[k])
            --k;
        while (k < int(nbins) - 1 && x >= T[k + 1])
            ++k;
    }
    return uint32_t(k);
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
    // the estimate is off by at most one bin; the loops are exact for any k
    float const t0 = T[k], t1 = T[k + 1];
    if (x < t0)
    {
        --k;
        while (k > 0 && x < T[k])
            --k;
    }
    else if (x >= t1)
    {
        ++k;
        while (k < int(nbins) - 1 && x >= T[k + 1])
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

// ------------------------------------------------------------------ self, row clusters
// A row is the run of cells (cx = 0..dim0-1) at fixed (cy, cz); its points are
// contiguous in cell order.  Clusters are runs of up to 8 consecutive points of
// one row, so every wave of the pair kernel is full except at row ends.
// Cluster entry: (row, start, count, cx_lo | cx_hi << 16), cx_lo/cx_hi the
// cells of its first and last point.
__global__ void row_cluster_count_kernel(const uint32_t* __restrict__ cstart, uint32_t dim0, uint32_t nrows,
                                         uint32_t* __restrict__ nclus)
{
    uint32_t const r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrows)
        nclus[r] = (cstart[(r + 1) * dim0] - cstart[r * dim0] + 7u) / 8u;
}

// One thread per cluster: the row by binary search over the row offsets, the
// cells of the first and last point by binary search over the row's cstart.
__global__ void row_cluster_table_kernel(const uint32_t* __restrict__ cstart, uint32_t dim0, uint32_t nrows,
                                         const uint32_t* __restrict__ cl_off, uint32_t max_clusters,
                                         uint4* __restrict__ table)
{
    uint32_t const c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= max_clusters || c >= cl_off[nrows])
        return;
    uint32_t lo = 0, hi = nrows; // last row with cl_off[row] <= c
    while (hi - lo > 1)
    {
        uint32_t const mid = (lo + hi) >> 1;
        if (cl_off[mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    uint32_t const r = lo;
    const uint32_t* cs = cstart + size_t(r) * dim0;
    uint32_t const p = cs[0] + 8u * (c - cl_off[r]);
    uint32_t const last = min(p + 8u, cs[dim0]) - 1u;
    auto cell_of_pos = [&](uint32_t q) { // last cell with cs[cell] <= q
        uint32_t l = 0, h = dim0;
        while (h - l > 1)
        {
            uint32_t const m = (l + h) >> 1;
            if (cs[m] <= q)
                l = m;
            else
                h = m;
        }
        return l;
    };
    table[c] = make_uint4(r, p, last - p + 1u, cell_of_pos(p) | (cell_of_pos(last) << 16));
}

// One wave per cluster; lane = (a, b) = (lane / 8, lane % 8): query point a of
// the cluster against point b of each 8-point chunk of the stencil rows.  The
// stencil spans cells [cx_lo - R, cx_hi + R] x [cy - R, cy + R] x [cz - R, cz + R]
// (cells at least r_max / R wide), walked as contiguous pieces of a row that
// share one periodic shift.  A pair related without a periodic shift is found
// from both of its points with bitwise-equal r_sq (x - y == -(y - x)), so it
// is counted once, from the point earlier in cell order, with weight 2; pairs
// across the box edge are counted from both sides with weight 1, as the
// reference's (q + image) arithmetic rounds differently for the two directions.
//
// Latency hiding: lane t first resolves stencil row t (its up to 3 periodic
// pieces and their cell-range bounds), so the cstart reads of all rows are in
// flight together; the wave then walks the rows, broadcasting each row's
// bounds with readlane, and loads the next chunk before binning the current.
//
// The pair work of one cluster entry e = (frame-local row, start, count,
// cx_lo | cx_hi << 16) for one wave, binned into h (LDS).
template<int R>
__device__ inline void rdf_rows_wave(const DevBox& b, const DevGrid& g, uint4 e, float q_rmax_sq, float q_rmin_sq,
                                     const float* sT, uint32_t nbins, float vmin, float inv_width, uint32_t* h)
{
    constexpr int kRowsY = 2 * R + 1;
    uint32_t const lane = threadIdx.x & 63u;
    int const dim0 = g.dim[0], dim1 = g.dim[1], dim2 = g.dim[2];
    // pairs counted: q_rmin_sq <= r_sq < q_rmax_sq (the query) and inside the histogram
    float const lo = fmaxf(q_rmin_sq, sT[0]), hi = fminf(q_rmax_sq, sT[nbins]);
    // row trimming: a candidate row (dy, dz) is at least (|dy|-1) cell extents away
    // along y and (|dz|-1) along z; conservative margins cover the rounding of
    // the cell assignment
    float const r_lim = sqrtf(q_rmax_sq) * 1.0001f;
    float const wx = g.cw[0] * 0.9999f, wy = g.cw[1] * 0.9999f, wz = g.cw[2] * 0.9999f;
    bool const ortho = b.xy == 0.0f && b.xz == 0.0f && b.yz == 0.0f;
    uint32_t const A0 = e.y, na = e.z;
    int const cx_lo = int(e.w & 0xffffu), cx_hi = int(e.w >> 16);
    int const cy0 = int(e.x % uint32_t(dim1));
    int const cz0 = int(e.x / uint32_t(dim1));
    uint32_t const a = lane >> 3, bl = lane & 7u;
    bool const a_ok = a < na;
    uint32_t const pos_a = A0 + (a_ok ? a : 0u);
    float4 pa = g.cpos[pos_a];
    if (b.is2D)
        pa.z = 0.0f;
    int const Rz = b.is2D ? 0 : R;
    int const nrows = kRowsY * (2 * Rz + 1);
    for (int r0 = 0; r0 < nrows; r0 += 64)
    {
        // lane t: stencil row r0 + t -> pieces (beg, end, image code); code < 0: no piece
        uint32_t pb[3] = {0u, 0u, 0u}, pe[3] = {0u, 0u, 0u};
        int pc[3] = {-1, -1, -1};
        int const rt = r0 + int(lane);
        if (rt < nrows)
        {
            int const dz = rt / kRowsY - Rz, dy = rt % kRowsY - R;
            int const uz = cz0 + dz, uy = cy0 + dy;
            int const sz = fa::floor_div(uz, dim2), sy = fa::floor_div(uy, dim1);
            float const gy = float(max(0, abs(dy) - 1)) * wy, gz = float(max(0, abs(dz) - 1)) * wz;
            int xr = R;
            bool reach = true;
            if (ortho)
            {
                float const rem = r_lim * r_lim - (gy * gy + gz * gz);
                reach = rem > 0.0f;
                if (reach && wx > 0.0f)
                    xr = min(R, int(sqrtf(rem) / wx) + 1);
            }
            else
                reach = fmaxf(gy, gz) < r_lim;
            if (reach && sz >= -1 && sz <= 1 && sy >= -1 && sy <= 1)
            {
                uint32_t const rowbase = uint32_t((uz - sz * dim2) * dim1 + (uy - sy * dim1)) * uint32_t(dim0);
                // shifts beyond one box length are not images the reference visits,
                // so the clipped range has at most 3 pieces (shift -1, 0, 1)
                int ux = max(cx_lo - xr, -dim0);
                int const uxend = min(cx_hi + xr, 2 * dim0 - 1);
#pragma unroll
                for (int k = 0; k < 3; ++k)
                {
                    if (ux <= uxend)
                    {
                        int const sx = fa::floor_div(ux, dim0);
                        int const piece_end = min(uxend, (sx + 1) * dim0 - 1);
                        if (sx >= -1 && sx <= 1)
                        {
                            int const cx = ux - sx * dim0;
                            pb[k] = g.cstart[rowbase + cx];
                            pe[k] = g.cstart[rowbase + cx + (piece_end - ux) + 1];
                            // AABB formula: r = p_j - (q_i + image), image = -shift
                            pc[k] = (-sx + 1) * 9 + (-sy + 1) * 3 + (-sz + 1);
                        }
                        ux = piece_end + 1;
                    }
                }
            }
        }
        int const nr = min(64, nrows - r0);
        for (int r = 0; r < nr; ++r)
        {
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                int const code = __builtin_amdgcn_readlane(pc[k], r);
                if (code < 0)
                    continue;
                uint32_t beg = uint32_t(__builtin_amdgcn_readlane(int(pb[k]), r));
                uint32_t const end = uint32_t(__builtin_amdgcn_readlane(int(pe[k]), r));
                bool const wraps = code != 13;
                if (!wraps)
                    beg = max(beg, A0);
                if (beg >= end)
                    continue;
                uint32_t const weight = wraps ? 1u : 2u;
                float const qx = pa.x + b.img[code][0];
                float const qy = pa.y + b.img[code][1];
                float const qz = pa.z + b.img[code][2];
                float4 cur = g.cpos[min(beg + bl, end - 1u)];
                for (uint32_t jb = beg; jb < end; jb += 8)
                {
                    float4 nxt = cur;
                    if (jb + 8u < end)
                        nxt = g.cpos[min(jb + 8u + bl, end - 1u)];
                    uint32_t const pos_b = jb + bl;
                    bool const valid = a_ok && pos_b < end && (wraps || pos_b > pos_a);
                    float const pjz = b.is2D ? 0.0f : cur.z;
                    float const rx = cur.x - qx;
                    float const ry = cur.y - qy;
                    float const rz = pjz - qz;
                    float const x = (rx * rx + ry * ry) + rz * rz;
                    if (valid && x >= lo && x < hi)
                        atomicAdd(&h[bin_in_range(x, sT, nbins, vmin, inv_width)], weight);
                    cur = nxt;
                }
            }
        }
    }
}

template<int R, bool LDS_HIST>
__global__ __launch_bounds__(kBlock) void rdf_rows_kernel(DevBox b, DevGrid g, const uint4* __restrict__ table,
                                                          const uint32_t* __restrict__ ncl_ptr, float q_rmax_sq,
                                                          float q_rmin_sq, const float* __restrict__ T,
                                                          uint32_t nbins, float vmin, float inv_width,
                                                          uint32_t* __restrict__ counts)
{
    extern __shared__ float smem[];
    float* sT = smem;
    uint32_t* hist = reinterpret_cast<uint32_t*>(smem + nbins + 1);
    hist_prologue<LDS_HIST>(hist, sT, T, nbins);
    constexpr uint32_t kWaves = kBlock / 64;
    uint32_t const I = fa::xcd_tile(blockIdx.x, gridDim.x) * kWaves + (threadIdx.x >> 6);
    if (I < *ncl_ptr)
        rdf_rows_wave<R>(b, g, table[I], q_rmax_sq, q_rmin_sq, sT, nbins, vmin, inv_width,
                         LDS_HIST ? hist : counts);
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

struct RowsArgs
{
    fa_nq* nq;
    float qmax_sq, qmin_sq;
    const float* T;
    uint32_t bins;
    float r_min, inv_width, q_r_max;
};
constexpr int kRowVariants = 3; // R = 2, 3, 4

// Self RDF through the row-cluster kernel with cells at least r_max / R wide.
// Sets *handled = false (and does nothing) when the kernel does not apply:
// points outside the box, or rows too long for the cluster encoding.
int run_rows(const RowsArgs& a, int R, uint32_t* d_counts, bool* handled)
{
    *handled = false;
    fa_nq* nq = a.nq;
    hipStream_t s = fa::stream();
    int dims[3];
    fa::grid_dims_for(nq->box, nq->n, a.q_r_max, R, dims);
    if (dims[0] > 65535)
        return FA_OK;
    if (int rc = fa::ensure_grid(nq, dims))
        return rc;
    if (nq->cur->any_outside)
        return FA_OK;
    *handled = true;
    DevGrid const g = nq->cur->dev();
    DevBox const& db = nq->cur->dbox;
    uint32_t const nrows = uint32_t(dims[1]) * uint32_t(dims[2]);
    uint64_t const max_clusters = uint64_t(nq->n) / 8 + nrows;
    void* tmp = nullptr;
    size_t const rb = sizeof(uint32_t) * (size_t(nrows) + 1);
    if (int rc = fa::scratch(13, 2 * rb + sizeof(uint4) * max_clusters + 256, &tmp))
        return rc;
    uint32_t* nclus = static_cast<uint32_t*>(tmp);
    uint32_t* cl_off = nclus + (nrows + 1);
    uint4* table = reinterpret_cast<uint4*>(static_cast<char*>(tmp) + ((2 * rb + 15) & ~size_t(15)));
    unsigned const rblocks = (nrows + kBlock - 1) / kBlock;
    FA_KLAUNCH(row_cluster_count_kernel, dim3(rblocks), dim3(kBlock), 0, s, g.cstart, uint32_t(dims[0]), nrows,
               nclus);
    if (int rc = fa::exclusive_scan_u32(nclus, cl_off, nrows, nullptr))
        return rc;
    unsigned const tblocks = unsigned((max_clusters + kBlock - 1) / kBlock);
    FA_KLAUNCH(row_cluster_table_kernel, dim3(tblocks), dim3(kBlock), 0, s, g.cstart, uint32_t(dims[0]), nrows,
               cl_off, uint32_t(max_clusters), table);
    unsigned const wblocks = unsigned((max_clusters + kBlock / 64 - 1) / (kBlock / 64));
    bool const lds = a.bins <= kMaxLdsBins;
    size_t const shmem = sizeof(float) * (a.bins + 1) + (lds ? sizeof(uint32_t) * a.bins : 0);
#define FA_ROWS(RR)                                                                                                 \
    case RR:                                                                                                        \
        if (lds)                                                                                                    \
            FA_KLAUNCH((rdf_rows_kernel<RR, true>), dim3(wblocks), dim3(kBlock), shmem, s, db, g, table,            \
                       cl_off + nrows, a.qmax_sq, a.qmin_sq, a.T, a.bins, a.r_min,        \
                       a.inv_width, d_counts);                                                                      \
        else                                                                                                        \
            FA_KLAUNCH((rdf_rows_kernel<RR, false>), dim3(wblocks), dim3(kBlock), shmem, s, db, g, table,           \
                       cl_off + nrows, a.qmax_sq, a.qmin_sq, a.T, a.bins, a.r_min,        \
                       a.inv_width, d_counts);                                                                      \
        break;
    switch (R)
    {
        FA_ROWS(2)
        FA_ROWS(3)
        FA_ROWS(4)
    default:
        return fa::fail(FA_ERR_RUNTIME, "unsupported RDF cell ratio");
    }
#undef FA_ROWS
    FA_HIP(hipGetLastError());
    return FA_OK;
}

// ------------------------------------------------------------------ batches
// The frame of a batch row (frames' rows are consecutive; binary search).
__device__ inline uint32_t frame_of_row(const fa::BatchFrame* __restrict__ frames, uint32_t nf, uint32_t row)
{
    uint32_t lo = 0, hi = nf;
    while (hi - lo > 1)
    {
        uint32_t const mid = (lo + hi) >> 1;
        if (frames[mid].row_base <= row)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ void brow_count_kernel(fa::BatchGrid G, uint32_t* __restrict__ nclus)
{
    uint32_t const r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= G.total_rows)
        return;
    const fa::BatchFrame& F = G.frames[frame_of_row(G.frames, G.n_frames, r)];
    uint32_t const dim0 = uint32_t(F.dim[0]);
    const uint32_t* cs = G.cstart + F.cell_base + size_t(r - F.row_base) * dim0;
    nclus[r] = (cs[dim0] - cs[0] + 7u) / 8u;
}

// Entries (frame-local row, start, count, cx_lo | cx_hi << 16) plus the frame
// of each entry in `cl_frame`.
__global__ void brow_table_kernel(fa::BatchGrid G, const uint32_t* __restrict__ cl_off, uint32_t max_clusters,
                                  uint4* __restrict__ table, uint32_t* __restrict__ cl_frame)
{
    uint32_t const c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= max_clusters || c >= cl_off[G.total_rows])
        return;
    uint32_t lo = 0, hi = G.total_rows;
    while (hi - lo > 1)
    {
        uint32_t const mid = (lo + hi) >> 1;
        if (cl_off[mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    uint32_t const r = lo;
    uint32_t const f = frame_of_row(G.frames, G.n_frames, r);
    const fa::BatchFrame& F = G.frames[f];
    uint32_t const dim0 = uint32_t(F.dim[0]);
    const uint32_t* cs = G.cstart + F.cell_base + size_t(r - F.row_base) * dim0;
    uint32_t const p = cs[0] + 8u * (c - cl_off[r]);
    uint32_t const last = min(p + 8u, cs[dim0]) - 1u;
    auto cell_of_pos = [&](uint32_t q) {
        uint32_t l = 0, h = dim0;
        while (h - l > 1)
        {
            uint32_t const m = (l + h) >> 1;
            if (cs[m] <= q)
                l = m;
            else
                h = m;
        }
        return l;
    };
    table[c] = make_uint4(r - F.row_base, p, last - p + 1u, cell_of_pos(p) | (cell_of_pos(last) << 16));
    cl_frame[c] = f;
}

// Flush a wave's private histogram into frame f's counts and clear it.
__device__ inline void flush_wave_hist(uint32_t* wh, uint32_t nbins, uint32_t* __restrict__ counts)
{
    uint32_t const lane = threadIdx.x & 63u;
    for (uint32_t k = lane; k < nbins; k += 64)
    {
        uint32_t const v = wh[k];
        if (v)
        {
            atomicAdd(&counts[k], v);
            wh[k] = 0u;
        }
    }
}

// Persistent batched RDF: every wave takes chunks of kChunk cluster entries
// from a device work queue until the batch is exhausted, so frames of any size
// share one launch and the GPU stays full to the end.  There is one queue per
// XCD over a contiguous eighth of the clusters (workgroups are dispatched
// round-robin over the 8 XCDs, so workgroup b runs on XCD b % 8): the waves of
// an XCD work through neighbouring clusters and share its L2; a wave whose
// range is exhausted takes chunks from the other XCDs' queues.  Each wave bins
// into its own LDS histogram and flushes it to its frame's counts (bins per
// frame) when the frame changes and at exit.  Frames with points outside their
// box take the generic per-point search for the cluster's points (weight 1 per
// ordered pair).
constexpr uint32_t kChunk = 4;

template<int R>
__global__ __launch_bounds__(kBlock) void brdf_rows_kernel(fa::BatchGrid G, const uint4* __restrict__ table,
                                                           const uint32_t* __restrict__ cl_frame,
                                                           const uint32_t* __restrict__ ncl_ptr,
                                                           uint32_t* __restrict__ queue, float q_rmax_sq,
                                                           const float* __restrict__ T, uint32_t nbins, float vmin,
                                                           float inv_width, uint32_t* __restrict__ counts)
{
    extern __shared__ float smem[];
    float* sT = smem;
    uint32_t const wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    uint32_t* wh = reinterpret_cast<uint32_t*>(smem + nbins + 1) + wave * nbins;
    for (uint32_t t = threadIdx.x; t <= nbins; t += blockDim.x)
        sT[t] = T[t];
    for (uint32_t t = lane; t < nbins; t += 64)
        wh[t] = 0u;
    __syncthreads();
    uint32_t const ncl = *ncl_ptr;
    uint32_t cur = 0xffffffffu;
    uint32_t const x0 = blockIdx.x % fa::kXcds;
    for (uint32_t xi = 0; xi < fa::kXcds;)
    {
        uint32_t const x = (x0 + xi) % fa::kXcds;
        uint32_t const r_lo = uint32_t((uint64_t(ncl) * x) / fa::kXcds);
        uint32_t const r_hi = uint32_t((uint64_t(ncl) * (x + 1)) / fa::kXcds);
        uint32_t off = 0;
        if (lane == 0)
            off = atomicAdd(&queue[x], kChunk);
        off = __builtin_amdgcn_readfirstlane(off);
        uint32_t const base = r_lo + off;
        if (base >= r_hi)
        {
            ++xi; // this range is done: help the next XCD's
            continue;
        }
        uint32_t const stop = min(base + kChunk, r_hi);
        for (uint32_t I = base; I < stop; ++I)
        {
            uint32_t const f = cl_frame[I];
            if (f != cur)
            {
                if (cur != 0xffffffffu)
                    flush_wave_hist(wh, nbins, counts + size_t(cur) * nbins);
                cur = f;
            }
            const fa::BatchFrame& F = G.frames[f];
            DevGrid const g = fa::frame_grid(F, G);
            uint4 const e = table[I];
            if (!F.any_outside)
            {
                rdf_rows_wave<R>(F.box, g, e, q_rmax_sq, 0.0f, sT, nbins, vmin, inv_width, wh);
            }
            else if (lane < e.z)
            {
                float4 const q = g.cpos[e.y + lane];
                uint32_t const qi = __float_as_uint(q.w);
                fa::for_each_candidate<AABB_FORMULA>(F.box, g, q.x, q.y, q.z, R, false,
                                                     [&](uint32_t j, float, float, float, float r_sq, int) {
                                                         if (j == qi || !(r_sq < q_rmax_sq))
                                                             return;
                                                         uint32_t const bin =
                                                             bin_of_rsq(r_sq, sT, nbins, vmin, inv_width);
                                                         if (bin < nbins)
                                                             atomicAdd(&wh[bin], 1u);
                                                     });
            }
        }
    }
    if (cur != 0xffffffffu)
        flush_wave_hist(wh, nbins, counts + size_t(cur) * nbins);
}

int grid_blocks(uint64_t work)
{
    uint64_t const b = (work + kBlock - 1) / kBlock;
    return int(std::max<uint64_t>(1, std::min<uint64_t>(b, 4096)));
}

} // namespace

namespace fa {

// Batched self RDF with cells at least r_max / R wide (R in 2..4): counts has
// n_frames x bins entries.  Bins beyond 2048 are not supported here (each wave
// keeps a private LDS histogram).
int batch_rdf(const BatchGrid& G, int R, float r_max, uint32_t bins, const BatchRdfScratch& S, uint32_t* d_counts)
{
    hipStream_t s = stream();
    if (bins == 0 || bins > 2048)
        return fail(FA_ERR_INVALID, "batched RDF supports 1..2048 bins");
    const float* T = S.T;
    float const inv_width = 1.0f / ((r_max - 0.0f) / float(bins));
    unsigned const rblocks = (G.total_rows + kBlock - 1) / kBlock;
    if (G.total_rows)
        FA_KLAUNCH(brow_count_kernel, dim3(rblocks), dim3(kBlock), 0, s, G, S.nclus);
    if (int rc = exclusive_scan_u32_async(S.nclus, S.cl_off, G.total_rows, S.scan_tmp, S.scan_bytes))
        return rc;
    unsigned const tblocks = unsigned((S.max_clusters + kBlock - 1) / kBlock);
    uint32_t* cl_frame = reinterpret_cast<uint32_t*>(S.table + S.max_clusters);
    if (S.max_clusters)
        FA_KLAUNCH(brow_table_kernel, dim3(tblocks), dim3(kBlock), 0, s, G, S.cl_off, uint32_t(S.max_clusters),
                   S.table, cl_frame);
    if (int rc = zero_u32(S.queue, kXcds))
        return rc;
    // persistent grid: enough workgroups to fill every CU a few times over
    static int n_cu = 0;
    if (!n_cu)
    {
        int dev = 0;
        FA_HIP(hipGetDevice(&dev));
        FA_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    unsigned const blocks = unsigned(std::max(1, n_cu) * 8);
    size_t const shmem = sizeof(float) * (bins + 1) + sizeof(uint32_t) * bins * (kBlock / 64);
    float const qmax_sq = r_max * r_max;
#define FA_BROWS(RR)                                                                                                \
    case RR:                                                                                                        \
        FA_KLAUNCH(brdf_rows_kernel<RR>, dim3(blocks), dim3(kBlock), shmem, s, G, S.table, cl_frame,               \
                   S.cl_off + G.total_rows, S.queue, qmax_sq, T, bins, 0.0f, inv_width, d_counts);                  \
        break;
    switch (R)
    {
        FA_BROWS(2)
        FA_BROWS(3)
        FA_BROWS(4)
    default:
        return fail(FA_ERR_INVALID, "unsupported RDF cell ratio");
    }
#undef FA_BROWS
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int thresholds_for(uint32_t bins, float r_max, float* d_T)
{
    float const inv_width = 1.0f / ((r_max - 0.0f) / float(bins));
    std::vector<float> T(bins + 1);
    for (uint32_t k = 0; k <= bins; ++k)
        T[k] = first_rsq_with_bin_at_least(int(k), 0.0f, r_max, inv_width, bins);
    FA_HIP(hipMemcpy(d_T, T.data(), sizeof(float) * (bins + 1), hipMemcpyHostToDevice));
    return FA_OK;
}

} // namespace fa

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
        // row-cluster kernel; the cell size (r_max / R) is picked by the autotuner
        RowsArgs ra{nq, qmax_sq, qmin_sq, T, bins, r_min, inv_width, q_r_max};
        double const density = double(nq->n) / double(fa::box_volume(nq->box));
        int const key = fa::tune_key(density, q_r_max);
        int variant = fa::tuned_variant(fa::TUNE_RDF_SELF, key);
        bool handled = false;
        if (variant < 0 && fa::tuning_enabled())
        {
            uint32_t* trial = nullptr;
            if (int rc = fa::scratch(14, sizeof(uint32_t) * bins + 256, (void**) &trial))
                return rc;
            float best = 0.0f;
            for (int v = 0; v < kRowVariants; ++v)
            {
                float ms = 0.0f;
                bool ok = false;
                // first run builds the grid; the second is timed
                if (int rc = run_rows(ra, 2 + v, trial, &ok))
                    return rc;
                if (!ok)
                    break;
                if (int rc = fa::time_launch([&] { return run_rows(ra, 2 + v, trial, &ok); }, &ms))
                    return rc;
                if (variant < 0 || ms < best)
                {
                    best = ms;
                    variant = v;
                }
            }
            if (variant >= 0)
                fa::record_variant(fa::TUNE_RDF_SELF, key, variant);
        }
        if (int rc = run_rows(ra, 2 + std::max(variant, 0), d_counts, &handled))
            return rc;
        if (handled)
        {
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
