// fa_grid.hip -- cell grid over fractional coordinates (the GPU counterpart of
// freud's LinkCell / CellQuery cell lists, freud/locality/LinkCell.cc:316-367,
// freud/locality/CellQuery.cc): counting sort of the points into cells, stored
// as cell-ordered float4 (x, y, z, bits(j)) for coalesced stencil reads.
#include <cmath>
#include <cstring>

#include "fa_common.h"

namespace {

// arr[idx] += 1 for every valid lane of the wave, as one atomic per distinct
// idx among the first groups of equal indices (points arriving in a spatially
// coherent order share cells); returns each lane's old value plus its rank in
// its group, i.e. what a per-lane atomicAdd(&arr[idx], 1) could have returned.
__device__ inline uint32_t wave_aggregated_increment(uint32_t* arr, uint32_t idx, bool valid)
{
    uint32_t const lane = threadIdx.x & 63u;
    uint64_t const below = (uint64_t(1) << lane) - 1u;
    uint64_t pending = __builtin_amdgcn_ballot_w64(valid);
    uint32_t leader = lane, rank = 0u, cnt = 1u;
    for (int it = 0; it < 16 && pending; ++it)
    {
        int const L = __builtin_ctzll(pending);
        uint32_t const lidx = uint32_t(__builtin_amdgcn_readlane(int(idx), L));
        uint64_t const same = __builtin_amdgcn_ballot_w64(idx == lidx) & pending;
        if ((same >> lane) & 1u)
        {
            leader = uint32_t(L);
            rank = uint32_t(__builtin_popcountll(same & below));
        }
        if (lane == uint32_t(L))
            cnt = uint32_t(__builtin_popcountll(same));
        pending &= ~same;
    }
    // lanes left over (many distinct cells in the wave) count for themselves
    if ((pending >> lane) & 1u)
    {
        leader = lane;
        rank = 0u;
        cnt = 1u;
    }
    uint32_t base = 0u;
    if (valid && leader == lane)
        base = atomicAdd(&arr[idx], cnt);
    return uint32_t(__shfl(int(base), int(leader))) + rank;
}

// rank_of_point (nullable) receives each point's rank among the points of its
// cell (the value its counter increment returned), which places it in the
// scatter without a second round of atomics.
__global__ void grid_assign_kernel(fa::DevBox b, const float* __restrict__ pts, uint32_t n, int d0, int d1, int d2,
                                   uint32_t* __restrict__ cell_of_point, int* __restrict__ wrap_of_point,
                                   uint32_t* __restrict__ counts, int* __restrict__ any_outside,
                                   uint32_t* __restrict__ rank_of_point)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool const valid = i < n;
    uint32_t cell = 0u;
    if (valid)
    {
        int const dim[3] = {d0, d1, d2};
        int c[3], w[3];
        fa::cell_of(b, dim, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], c, w);
        cell = uint32_t((c[2] * d1 + c[1]) * d0 + c[0]);
        cell_of_point[i] = cell;
        if (wrap_of_point)
            wrap_of_point[i] = fa::pack_wrap(w);
        if (w[0] != 0 || w[1] != 0 || w[2] != 0)
            atomicOr(any_outside, 1);
    }
    uint32_t const rank = wave_aggregated_increment(counts, cell, valid);
    if (valid && rank_of_point)
        rank_of_point[i] = rank;
}

__global__ void grid_scatter_kernel(const float* __restrict__ pts, uint32_t n,
                                    const uint32_t* __restrict__ cell_of_point, const int* __restrict__ wrap_of_point,
                                    const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ rank_of_point,
                                    float4* __restrict__ cpos, int* __restrict__ wrapc)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t const slot = cstart[cell_of_point[i]] + rank_of_point[i];
    cpos[slot] = make_float4(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], __uint_as_float(i));
    wrapc[slot] = wrap_of_point[i];
}

// ---------------------------------------------------------------- batches
// Per tile (one frame per workgroup): the global cell of each point, its rank
// among its cell's points and the cell counts; flags the frame when a point
// lies outside its box.
__global__ __launch_bounds__(256) void bgrid_assign_kernel(fa::BatchGrid G, const uint32_t* __restrict__ tile_frame,
                                                           const float* __restrict__ pts)
{
    uint32_t const t = blockIdx.x;
    const fa::BatchFrame& F = G.frames[tile_frame[t]];
    uint32_t const local = (t - F.tile_base) * fa::kTile + threadIdx.x;
    bool const valid = local < F.n;
    uint32_t cell = 0u;
    if (valid)
    {
        uint32_t const i = F.pt_base + local;
        int c[3], w[3];
        fa::cell_of(F.box, F.dim, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], c, w);
        cell = F.cell_base + uint32_t((c[2] * F.dim[1] + c[1]) * F.dim[0] + c[0]);
        G.cell_of_point[i] = cell;
        if (w[0] != 0 || w[1] != 0 || w[2] != 0)
            atomicOr(&G.frames[tile_frame[t]].any_outside, 1);
    }
    uint32_t const rank = wave_aggregated_increment(G.fill, cell, valid);
    if (valid)
        G.rank_of_point[F.pt_base + local] = rank;
}

__global__ __launch_bounds__(256) void bgrid_scatter_kernel(fa::BatchGrid G, const uint32_t* __restrict__ tile_frame,
                                                            const float* __restrict__ pts)
{
    uint32_t const t = blockIdx.x;
    const fa::BatchFrame& F = G.frames[tile_frame[t]];
    uint32_t const local = (t - F.tile_base) * fa::kTile + threadIdx.x;
    bool const valid = local < F.n;
    if (!valid)
        return;
    uint32_t const i = F.pt_base + local;
    uint32_t const slot = G.cstart[G.cell_of_point[i]] + G.rank_of_point[i];
    float const x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    G.cpos[slot] = make_float4(x, y, z, __uint_as_float(local));
    int c[3], w[3];
    fa::cell_of(F.box, F.dim, x, y, z, c, w);
    G.wrapc[slot] = fa::pack_wrap(w);
}

__global__ void bgrid_reset_kernel(fa::BatchGrid G)
{
    uint32_t const f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < G.n_frames)
        G.frames[f].any_outside = 0;
}

// Counting-sort placement of point indices: order[slot] = i.
__global__ void order_scatter_kernel(uint32_t n, const uint32_t* __restrict__ cell_of_point,
                                     const uint32_t* __restrict__ cstart, uint32_t* __restrict__ fill,
                                     uint32_t* __restrict__ order)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t const cell = cell_of_point[i];
    order[cstart[cell] + atomicAdd(&fill[cell], 1u)] = i;
}

// How scattered the caller's order is: of kProbe evenly spaced index pairs
// (i, i + 1), the number whose points lie more than `far` apart (minimum
// image).  Points stored in a spatially coherent order (lattice order,
// simulation cell sorting) have consecutive points close together.
constexpr uint32_t kProbe = 4096;
__global__ void order_probe_kernel(fa::DevBox b, const float* __restrict__ pts, uint32_t n, float far_sq,
                                   uint32_t* __restrict__ count)
{
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kProbe)
        return;
    uint32_t const i = uint32_t((uint64_t(t) * (n - 1)) / kProbe);
    float dx, dy, dz;
    fa::wrap(b, pts[3 * i + 3] - pts[3 * i], pts[3 * i + 4] - pts[3 * i + 1], pts[3 * i + 5] - pts[3 * i + 2], dx, dy,
             dz);
    if ((dx * dx + dy * dy) + dz * dz > far_sq)
        atomicAdd(count, 1u);
}

struct OrderCache
{
    const float* pts = nullptr;
    uint32_t n = 0;
    fa_box box{};
    bool use = false; // false: the caller's order is already coherent
    uint32_t* order = nullptr;
    uint32_t cap = 0;
    uint64_t stamp = 0; // last use (0: never)
};
// A few point arrays keep their decision and order at once (a workload that
// alternates between systems would otherwise rebuild the order per call).
// The order only steers which thread handles which point, so an entry whose
// array was overwritten in place still yields correct results.
constexpr int kOrderEntries = 4;
OrderCache g_order[kOrderEntries];
uint64_t g_order_clock = 0;

} // namespace

namespace fa {

int spatial_order(const fa_box& box, const float* d_points, uint32_t n, const uint32_t** out)
{
    *out = nullptr;
    if (n < kSpatialOrderMin)
        return FA_OK;
    int victim = 0;
    for (int e = 0; e < kOrderEntries; ++e)
    {
        OrderCache& c = g_order[e];
        if (c.pts == d_points && c.n == n && std::memcmp(&c.box, &box, sizeof(box)) == 0)
        {
            c.stamp = ++g_order_clock;
            *out = c.use ? c.order : nullptr;
            return FA_OK;
        }
        if (c.stamp < g_order[victim].stamp)
            victim = e;
    }
    OrderCache& c = g_order[victim];
    c.pts = nullptr;
    hipStream_t s = stream();
    double const v = double(box_volume(box));
    double const spacing = box.is2D ? std::sqrt(v / n) : std::cbrt(v / n);
    {
        // keep the caller's order when at most a quarter of consecutive points
        // lie more than 4 mean spacings apart
        uint32_t* cnt = nullptr;
        if (int rc = scratch(18, 256, (void**) &cnt))
            return rc;
        FA_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t), s));
        float const far = float(4.0 * spacing);
        FA_KLAUNCH(order_probe_kernel, dim3(kProbe / 256), dim3(256), 0, s, make_devbox(box), d_points, n, far * far,
                   cnt);
        uint32_t h = 0;
        if (int rc = read_u32(cnt, &h))
            return rc;
        if (h <= kProbe / 4)
        {
            c.pts = d_points;
            c.n = n;
            c.box = box;
            c.use = false;
            c.stamp = ++g_order_clock;
            return FA_OK;
        }
    }
    // cells holding about four points each
    double const r = box.is2D ? std::sqrt(4.0 * v / n) : std::cbrt(4.0 * v / n);
    int dims[3];
    grid_dims_for(box, n, float(r), 1, dims);
    uint32_t const ncell = uint32_t(dims[0]) * dims[1] * dims[2];
    if (c.cap < n)
    {
        device_free(c.order);
        c.order = nullptr;
        c.cap = 0;
        c.pts = nullptr;
        if (int rc = device_alloc((void**) &c.order, sizeof(uint32_t) * n))
            return rc;
        c.cap = n;
    }
    void* tmp = nullptr;
    size_t const cb = sizeof(uint32_t) * (size_t(ncell) + 1);
    if (int rc = scratch(17, sizeof(uint32_t) * size_t(n) + 3 * cb + 256, &tmp))
        return rc;
    uint32_t* cell_of_point = static_cast<uint32_t*>(tmp);
    uint32_t* counts = cell_of_point + n;
    uint32_t* cstart = counts + ncell + 1;
    uint32_t* fill = cstart + ncell + 1;
    int* d_any = reinterpret_cast<int*>(fill + ncell);
    FA_HIP(hipMemsetAsync(counts, 0, cb, s));
    FA_HIP(hipMemsetAsync(fill, 0, sizeof(uint32_t) * ncell + sizeof(int), s));
    DevBox const db = make_devbox(box);
    unsigned const blocks = (n + 255) / 256;
    FA_KLAUNCH(grid_assign_kernel, dim3(blocks), dim3(256), 0, s, db, d_points, n, dims[0], dims[1], dims[2],
               cell_of_point, nullptr, counts, d_any, nullptr);
    if (int rc = exclusive_scan_u32(counts, cstart, ncell, nullptr))
        return rc;
    FA_KLAUNCH(order_scatter_kernel, dim3(blocks), dim3(256), 0, s, n, cell_of_point, cstart, fill, c.order);
    FA_HIP(hipGetLastError());
    c.pts = d_points;
    c.n = n;
    c.box = box;
    c.use = true;
    c.stamp = ++g_order_clock;
    *out = c.order;
    return FA_OK;
}

int batch_grid_build(const BatchGrid& G, const uint32_t* d_tile_frame, uint32_t ntiles, const float* d_pts, void* tmp,
                     size_t tmp_bytes)
{
    hipStream_t s = stream();
    FA_KLAUNCH(bgrid_reset_kernel, dim3((G.n_frames + 255) / 256), dim3(256), 0, s, G);
    if (int rc = zero_u32(G.fill, G.total_cells))
        return rc;
    if (ntiles)
        FA_KLAUNCH(bgrid_assign_kernel, dim3(ntiles), dim3(kTile), 0, s, G, d_tile_frame, d_pts);
    if (int rc = exclusive_scan_u32_async(G.fill, G.cstart, G.total_cells, tmp, tmp_bytes))
        return rc;
    if (ntiles)
        FA_KLAUNCH(bgrid_scatter_kernel, dim3(ntiles), dim3(kTile), 0, s, G, d_tile_frame, d_pts);
    FA_HIP(hipGetLastError());
    return FA_OK;
}


Grid::~Grid()
{
    device_free(cpos);
    device_free(wrapc);
    device_free(cstart);
    device_free(cell_of_point);
    device_free(outside);
}

int build_grid(Grid& g, const fa_box& box, const float* d_points, uint32_t n, const int dims[3])
{
    hipStream_t s = stream();
    uint64_t const ncell64 = uint64_t(dims[0]) * dims[1] * dims[2];
    if (ncell64 > (1ull << 30))
        return fail(FA_ERR_RUNTIME, "cell grid too large");
    uint32_t const ncell = uint32_t(ncell64);
    bool const realloc_pts = (g.n < n) || !g.cpos;
    bool const realloc_cells = !g.cstart || (uint64_t(g.dim[0]) * g.dim[1] * g.dim[2] < ncell64);
    if (realloc_pts)
    {
        device_free(g.cpos);
        device_free(g.wrapc);
        device_free(g.cell_of_point);
        g.cpos = nullptr;
        g.wrapc = nullptr;
        g.cell_of_point = nullptr;
        if (int rc = device_alloc((void**) &g.cpos, sizeof(float4) * (size_t(n) + kCposPad)))
            return rc;
        if (int rc = device_alloc((void**) &g.wrapc, sizeof(int) * size_t(n)))
            return rc;
        if (int rc = device_alloc((void**) &g.cell_of_point, sizeof(uint32_t) * size_t(n)))
            return rc;
    }
    if (realloc_cells)
    {
        device_free(g.cstart);
        g.cstart = nullptr;
        if (int rc = device_alloc((void**) &g.cstart, sizeof(uint32_t) * (size_t(ncell) + 1)))
            return rc;
    }
    g.box = box;
    g.dbox = make_devbox(box);
    g.dim[0] = dims[0];
    g.dim[1] = dims[1];
    g.dim[2] = dims[2];
    g.n = n;
    {
        float npd[3];
        nearest_plane_distance(box, npd);
        for (int k = 0; k < 3; ++k)
            g.cw[k] = (k == 2 && box.is2D) ? 0.0f : npd[k] / float(dims[k]);
    }
    // scratch: counts (ncell) + each point's rank in its cell + each point's
    // packed image offset; the outside flag lives with the grid, read by the
    // kernels (no host round trip)
    void* tmp = nullptr;
    size_t const cb = sizeof(uint32_t) * size_t(ncell);
    if (int rc = scratch(2, cb + sizeof(uint32_t) * size_t(n) + sizeof(int) * size_t(n) + 256, &tmp))
        return rc;
    uint32_t* counts = static_cast<uint32_t*>(tmp);
    uint32_t* rank_of_point = counts + ncell;
    int* wrap_of_point = reinterpret_cast<int*>(rank_of_point + n);
    if (!g.outside)
        if (int rc = device_alloc((void**) &g.outside, 256))
            return rc;
    int* d_any = g.outside;
    FA_HIP(hipMemsetAsync(counts, 0, cb, s));
    FA_HIP(hipMemsetAsync(d_any, 0, sizeof(int), s));
    unsigned const blocks = (n + 255) / 256;
    if (n)
    {
        FA_KLAUNCH(grid_assign_kernel, dim3(blocks), dim3(256), 0, s, g.dbox, d_points, n, dims[0], dims[1],
                           dims[2], g.cell_of_point, wrap_of_point, counts, d_any, rank_of_point);
        FA_HIP(hipGetLastError());
    }
    if (int rc = exclusive_scan_u32(counts, g.cstart, ncell, nullptr))
        return rc;
    if (n)
    {
        FA_KLAUNCH(grid_scatter_kernel, dim3(blocks), dim3(256), 0, s, d_points, n, g.cell_of_point, wrap_of_point,
                   g.cstart, rank_of_point, g.cpos, g.wrapc);
        FA_HIP(hipGetLastError());
    }
    return FA_OK;
}

int build_grid_into(const GridBuffers& B, const DevBox& db, const float* d_points, uint32_t n, const int dims[3])
{
    hipStream_t s = stream();
    uint32_t const ncell = uint32_t(dims[0]) * uint32_t(dims[1]) * uint32_t(dims[2]);
    if (int rc = zero_u32(B.counts, ncell))
        return rc;
    if (int rc = zero_u32(reinterpret_cast<uint32_t*>(B.outside), 1))
        return rc;
    unsigned const blocks = (n + 255) / 256;
    if (n)
    {
        FA_KLAUNCH(grid_assign_kernel, dim3(blocks), dim3(256), 0, s, db, d_points, n, dims[0], dims[1], dims[2],
                   B.cell_of_point, B.wrap_of_point, B.counts, B.outside, B.rank_of_point);
        FA_HIP(hipGetLastError());
    }
    if (int rc = exclusive_scan_u32_async(B.counts, B.cstart, ncell, B.scan_tmp, B.scan_bytes))
        return rc;
    if (n)
    {
        FA_KLAUNCH(grid_scatter_kernel, dim3(blocks), dim3(256), 0, s, d_points, n, B.cell_of_point,
                   B.wrap_of_point, B.cstart, B.rank_of_point, B.cpos, B.wrapc);
        FA_HIP(hipGetLastError());
    }
    return FA_OK;
}

} // namespace fa
