// fa_batch.hip -- batched trajectories: many frames (any box shape, any size)
// analysed together, one launch per analysis step for the whole batch.
//
// The reference analyses one frame per call (freud's compute(..., reset=False)
// loop over a trajectory, e.g. RDF.cc:101-110 per frame); a batch gives every
// frame the same results as the single-frame entry points (ball-query
// NeighborList bit for bit, RDF counts per frame bit for bit, q_l per particle)
// while replacing F x (launches + host synchronisations) with one sequence of
// about twenty launches:
//
//   grid build (per-frame cells, one scan over all cells)
//   ball count -> scan -> total -> ball fill            (per-frame NeighborLists)
//   RDF grid build -> row clusters -> persistent pair kernel
//   q_lm / q_l over the ball bonds -> per-frame order
//
// The sequence never synchronises with the host and uses no memset, so after
// the first (eager) run it is captured into a HIP graph and replayed.  The
// bond arrays are sized from the previous run; a replay whose bonds would not
// fit writes no bonds (the kernels check the total on the device) and the host
// reruns with room for them.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fa_common.h"

namespace {

using fa::BatchFrame;
using fa::BatchGrid;

__global__ void sum_counts_kernel(const uint32_t* __restrict__ counts, uint32_t n,
                                  unsigned long long* __restrict__ total)
{
    __shared__ unsigned long long s[256];
    unsigned long long acc = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        acc += counts[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1)
    {
        if (threadIdx.x < uint32_t(k))
            s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicAdd(total, s[0]);
}

__global__ void zero_u64_kernel(unsigned long long* p)
{
    *p = 0ull;
}

// Per-frame bond counts (from the batch offsets) for the host.
__global__ void frame_bonds_kernel(const BatchFrame* __restrict__ frames, uint32_t nf,
                                   const uint32_t* __restrict__ offsets, uint32_t* __restrict__ out)
{
    uint32_t const f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nf)
        out[f] = offsets[frames[f].pt_base + frames[f].n] - offsets[frames[f].pt_base];
}

// A frame's sub-list as a NeighborList of its own: rebased offsets, counts and
// segments (0 for points without bonds, NeighborList.cc:199-232).
__global__ void frame_nlist_kernel(const uint32_t* __restrict__ offsets, uint32_t base, uint32_t n,
                                   uint32_t* __restrict__ o, uint32_t* __restrict__ counts,
                                   uint32_t* __restrict__ segments)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n)
        return;
    uint32_t const b0 = offsets[base];
    o[i] = offsets[base + i] - b0;
    if (i < n)
    {
        uint32_t const c = offsets[base + i + 1] - offsets[base + i];
        counts[i] = c;
        segments[i] = c ? offsets[base + i] - b0 : 0u;
    }
}

} // namespace

struct fa_batch
{
    uint32_t n_frames = 0;
    uint32_t total = 0;  // points
    uint32_t ntiles = 0;
    std::vector<fa_box> boxes;
    std::vector<uint64_t> pt_off;
    fa_batch_plan plan{};

    float* d_points = nullptr;
    uint32_t* d_tile_frame = nullptr;

    // ball query
    int ball_R = 1;
    BatchGrid gb;
    uint32_t* d_counts = nullptr;  // total
    uint32_t* d_offsets = nullptr; // total + 1
    uint64_t capacity = 0;
    uint32_t* d_nb = nullptr;
    float* d_dist = nullptr;
    float* d_wts = nullptr;
    float* d_vec = nullptr;
    void* d_gkeys = nullptr;
    uint32_t ball_C = 0;        // keys per point the walk stages (0: count and fill)
    uint32_t* d_pos = nullptr;  // total: where each point's keys start in d_rows
    uint32_t* d_rows = nullptr; // ntiles * kTile * ball_C packed keys

    // RDF
    int rdf_R = 2;
    int rdf_F = 1; // RDF cells are F times narrower along x
    BatchGrid gr;
    fa::BatchRdfScratch rs{};
    uint32_t* d_rdf = nullptr; // n_frames x bins
    float* d_T = nullptr;

    // Steinhardt
    float* d_qlm = nullptr;
    float* d_ql = nullptr;
    float* d_partial = nullptr;
    float* d_order = nullptr;

    // status read back after every run
    unsigned long long* d_total_bonds = nullptr;
    uint32_t* d_frame_bonds = nullptr;
    unsigned long long* h_status = nullptr; // pinned: [total bonds]
    std::vector<uint32_t> frame_bonds;
    uint64_t last_total_bonds = 0;

    void* d_scan = nullptr;
    size_t scan_bytes = 0;

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool tuned = false;

    ~fa_batch();
};

namespace {

void free_grid(BatchGrid& g)
{
    fa::device_free(g.frames);
    fa::device_free(g.cpos);
    fa::device_free(g.wrapc);
    fa::device_free(g.cstart);
    fa::device_free(g.cell_of_point);
    fa::device_free(g.rank_of_point);
    fa::device_free(g.fill);
    g = BatchGrid{};
}

void drop_graph(fa_batch* b)
{
    if (b->exec)
        (void) hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// RDF launch shapes the tuner tries: (R, Fx), cells r / R wide and Fx times
// narrower along x.
constexpr int kRdfShape[][2] = {{2, 1}, {3, 1}, {4, 1}, {2, 4}, {2, 8}};
constexpr int kRdfShapes = int(sizeof(kRdfShape) / sizeof(kRdfShape[0]));

// Frame descriptions and device arrays of a grid whose cells are at least
// r / R wide in every frame, and r / (R Fx) along x where the cell count
// allows it.
int make_grid(fa_batch* b, float r, int R, BatchGrid& g, int Fx = 1)
{
    free_grid(g);
    std::vector<BatchFrame> fr(b->n_frames);
    uint64_t cells = 0, rows = 0, tiles = 0;
    for (uint32_t f = 0; f < b->n_frames; ++f)
    {
        BatchFrame& F = fr[f];
        std::memset(&F, 0, sizeof(F));
        F.box = fa::make_devbox(b->boxes[f]);
        uint32_t const n = uint32_t(b->pt_off[f + 1] - b->pt_off[f]);
        fa::grid_dims_for(b->boxes[f], n, r, R, F.dim);
        int fx = Fx;
        if (F.dim[0] * fx > 65535 ||
            double(F.dim[0]) * fx * F.dim[1] * F.dim[2] > std::max<double>(8.0 * n, 64.0))
            fx = 1;
        F.dim[0] *= fx;
        F.rx = R * fx;
        {
            float npd[3];
            fa::nearest_plane_distance(b->boxes[f], npd);
            for (int k = 0; k < 3; ++k)
                F.cw[k] = (k == 2 && b->boxes[f].is2D) ? 0.0f : npd[k] / float(F.dim[k]);
        }
        F.cell_base = uint32_t(cells);
        F.row_base = uint32_t(rows);
        F.pt_base = uint32_t(b->pt_off[f]);
        F.n = n;
        F.tile_base = uint32_t(tiles);
        cells += uint64_t(F.dim[0]) * F.dim[1] * F.dim[2];
        rows += uint64_t(F.dim[1]) * F.dim[2];
        tiles += (n + fa::kTile - 1) / fa::kTile;
        if (F.dim[0] > 65535)
            return fa::fail(FA_ERR_RUNTIME, "batch frame grid too long along x");
    }
    if (cells >= (1ull << 31) || rows >= (1ull << 31))
        return fa::fail(FA_ERR_RUNTIME, "batch cell grid too large");
    g.n_frames = b->n_frames;
    g.total_points = b->total;
    g.total_cells = uint32_t(cells);
    g.total_rows = uint32_t(rows);
    size_t const np = std::max<size_t>(b->total, 1);
    if (int rc = fa::device_alloc((void**) &g.frames, sizeof(BatchFrame) * std::max<size_t>(fr.size(), 1)))
        return rc;
    if (int rc = fa::device_alloc((void**) &g.cpos, sizeof(float4) * (np + fa::kCposPad)))
        return rc;
    if (int rc = fa::device_alloc((void**) &g.wrapc, sizeof(int) * np))
        return rc;
    if (int rc = fa::device_alloc((void**) &g.cell_of_point, sizeof(uint32_t) * np))
        return rc;
    if (int rc = fa::device_alloc((void**) &g.rank_of_point, sizeof(uint32_t) * np))
        return rc;
    if (int rc = fa::device_alloc((void**) &g.cstart, sizeof(uint32_t) * (cells + 1)))
        return rc;
    if (int rc = fa::device_alloc((void**) &g.fill, sizeof(uint32_t) * (cells + 1)))
        return rc;
    if (!fr.empty())
        FA_HIP(hipMemcpy(g.frames, fr.data(), sizeof(BatchFrame) * fr.size(), hipMemcpyHostToDevice));
    // one scan buffer serves the cell, row and point scans
    size_t const need = std::max({fa::scan_tmp_bytes(uint32_t(cells)), fa::scan_tmp_bytes(uint32_t(rows)),
                                  fa::scan_tmp_bytes(b->total)});
    if (need == 0)
        return fa::fail(FA_ERR_RUNTIME, "scan size query failed");
    if (need > b->scan_bytes)
    {
        fa::device_free(b->d_scan);
        b->d_scan = nullptr;
        b->scan_bytes = 0;
        if (int rc = fa::device_alloc(&b->d_scan, need))
            return rc;
        b->scan_bytes = need;
    }
    return FA_OK;
}

int alloc_bonds(fa_batch* b, uint64_t cap)
{
    fa::device_free(b->d_nb);
    fa::device_free(b->d_dist);
    fa::device_free(b->d_wts);
    fa::device_free(b->d_vec);
    fa::device_free(b->d_gkeys);
    b->d_nb = nullptr;
    b->d_dist = b->d_wts = b->d_vec = nullptr;
    b->d_gkeys = nullptr;
    b->capacity = 0;
    cap = std::max<uint64_t>(cap, 1024);
    if (cap > 0xffffffffull)
        return fa::fail(FA_ERR_RUNTIME, "batch holds more than 2^32 - 1 bonds");
    if (int rc = fa::device_alloc((void**) &b->d_nb, sizeof(uint32_t) * 2 * cap))
        return rc;
    if (int rc = fa::device_alloc((void**) &b->d_dist, sizeof(float) * cap))
        return rc;
    if (int rc = fa::device_alloc((void**) &b->d_wts, sizeof(float) * cap))
        return rc;
    if (int rc = fa::device_alloc((void**) &b->d_vec, sizeof(float) * 3 * cap))
        return rc;
    if (int rc = fa::device_alloc(&b->d_gkeys, sizeof(uint32_t) * cap))
        return rc;
    b->capacity = cap;
    return FA_OK;
}

int alloc_rdf_scratch(fa_batch* b)
{
    fa::device_free(b->rs.nclus);
    fa::device_free(b->rs.table);
    b->rs = fa::BatchRdfScratch{};
    uint64_t const rows = b->gr.total_rows;
    uint64_t const maxc = uint64_t(b->total) / 8 + rows + 1;
    uint32_t* u = nullptr;
    if (int rc = fa::device_alloc((void**) &u, sizeof(uint32_t) * (2 * (rows + 1) + 64)))
        return rc;
    b->rs.nclus = u;
    b->rs.cl_off = u + rows + 1;
    b->rs.queue = u + 2 * (rows + 1);
    // table entries followed by their frames
    if (int rc = fa::device_alloc((void**) &b->rs.table, sizeof(uint4) * maxc + sizeof(uint32_t) * maxc))
        return rc;
    b->rs.max_clusters = maxc;
    b->rs.T = b->d_T;
    b->rs.scan_tmp = b->d_scan;
    b->rs.scan_bytes = b->scan_bytes;
    return FA_OK;
}

bool do_ball(const fa_batch* b)
{
    return b->plan.ball_r_max > 0.0f;
}
bool do_rdf(const fa_batch* b)
{
    return b->plan.rdf_bins > 0;
}
bool do_ql(const fa_batch* b)
{
    return do_ball(b) && b->plan.l != 0xffffffffu;
}

// Every stage of a run, enqueued on the library stream.
int enqueue_ball(fa_batch* b)
{
    if (int rc = fa::batch_grid_build(b->gb, b->d_tile_frame, b->ntiles, b->d_points, b->d_scan, b->scan_bytes))
        return rc;
    if (b->ball_C)
    {
        if (int rc = fa::batch_ball_walk(b->gb, b->d_tile_frame, b->ntiles, b->plan.ball_r_max, b->ball_R, b->ball_C,
                                         b->d_counts, b->d_pos, b->d_rows))
            return rc;
    }
    else if (int rc = fa::batch_ball_count(b->gb, b->d_tile_frame, b->ntiles, b->plan.ball_r_max, b->ball_R,
                                           b->d_counts))
        return rc;
    if (int rc = fa::exclusive_scan_u32_async(b->d_counts, b->d_offsets, b->total, b->d_scan, b->scan_bytes))
        return rc;
    hipStream_t s = fa::stream();
    FA_KLAUNCH(zero_u64_kernel, dim3(1), dim3(1), 0, s, b->d_total_bonds);
    if (b->total)
        FA_KLAUNCH(sum_counts_kernel, dim3(std::min<uint32_t>((b->total + 255) / 256, 1024)), dim3(256), 0, s,
                   b->d_counts, b->total, b->d_total_bonds);
    if (b->ball_C)
    {
        if (int rc = fa::batch_ball_emit(b->gb, b->d_tile_frame, b->ntiles, b->d_points, b->plan.ball_r_max,
                                         b->ball_R, b->ball_C, b->d_pos, b->d_rows, b->d_offsets, b->d_total_bonds,
                                         b->capacity, b->d_nb, b->d_dist, b->d_wts, b->d_vec, b->d_gkeys))
            return rc;
    }
    else if (int rc = fa::batch_ball_fill(b->gb, b->d_tile_frame, b->ntiles, b->d_points, b->plan.ball_r_max,
                                          b->ball_R, b->d_offsets, b->d_total_bonds, b->capacity, b->d_nb, b->d_dist,
                                          b->d_wts, b->d_vec, b->d_gkeys))
        return rc;
    FA_KLAUNCH(frame_bonds_kernel, dim3((b->n_frames + 255) / 256), dim3(256), 0, s, b->gb.frames, b->n_frames,
               b->d_offsets, b->d_frame_bonds);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int enqueue_rdf(fa_batch* b)
{
    if (int rc = fa::batch_grid_build(b->gr, b->d_tile_frame, b->ntiles, b->d_points, b->d_scan, b->scan_bytes))
        return rc;
    if (int rc = fa::zero_u32(b->d_rdf, uint64_t(b->n_frames) * b->plan.rdf_bins))
        return rc;
    b->rs.scan_tmp = b->d_scan;
    b->rs.scan_bytes = b->scan_bytes;
    return fa::batch_rdf(b->gr, b->rdf_R, b->plan.rdf_r_max, b->plan.rdf_bins, b->rs, b->d_rdf);
}

int enqueue_ql(fa_batch* b)
{
    return fa::batch_steinhardt(b->gb, b->d_tile_frame, b->ntiles, b->d_points, b->d_nb, b->d_dist, b->d_offsets,
                                b->d_total_bonds, b->capacity, b->plan.l, b->d_qlm, b->d_ql, b->d_partial,
                                b->d_order);
}

int enqueue_all(fa_batch* b)
{
    if (do_ball(b))
        if (int rc = enqueue_ball(b))
            return rc;
    if (do_rdf(b))
        if (int rc = enqueue_rdf(b))
            return rc;
    if (do_ql(b))
        if (int rc = enqueue_ql(b))
            return rc;
    return FA_OK;
}

// Reads the run's bond total; true when the bonds fitted.
int read_status(fa_batch* b, bool* fitted)
{
    hipStream_t s = fa::stream();
    *fitted = true;
    if (!do_ball(b))
        return FA_OK;
    FA_HIP(hipMemcpyAsync(b->h_status, b->d_total_bonds, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    b->last_total_bonds = b->h_status[0];
    *fitted = b->last_total_bonds <= b->capacity;
    return FA_OK;
}

// The autotuner's choice of cell size for the ball query (R = 1, 2) and the
// RDF (the kRdfShape launch shapes), timed on this batch's own data.
int tune(fa_batch* b)
{
    double pts = 0, vol = 0;
    for (uint32_t f = 0; f < b->n_frames; ++f)
    {
        pts += double(b->pt_off[f + 1] - b->pt_off[f]);
        vol += double(fa::box_volume(b->boxes[f]));
    }
    double const density = vol > 0 ? pts / vol : 0.0;
    if (do_ball(b))
    {
        int const key = fa::tune_key(density, b->plan.ball_r_max) + fa::kBatchTuneKey;
        int v = fa::tuned_variant(fa::TUNE_BALL, key);
        if (v < 0 && fa::tuning_enabled())
        {
            float best = 0;
            for (int cand = 0; cand < 2; ++cand)
            {
                b->ball_R = 1 + cand;
                if (int rc = make_grid(b, b->plan.ball_r_max, b->ball_R, b->gb))
                    return rc;
                float ms = 0;
                auto run = [&] {
                    if (int rc = fa::batch_grid_build(b->gb, b->d_tile_frame, b->ntiles, b->d_points, b->d_scan,
                                                      b->scan_bytes))
                        return rc;
                    return fa::batch_ball_count(b->gb, b->d_tile_frame, b->ntiles, b->plan.ball_r_max, b->ball_R,
                                                b->d_counts);
                };
                if (int rc = run())
                    return rc;
                if (int rc = fa::time_launch(run, &ms))
                    return rc;
                if (v < 0 || ms < best)
                {
                    best = ms;
                    v = cand;
                }
            }
            fa::record_variant(fa::TUNE_BALL, key, v);
        }
        b->ball_R = 1 + std::max(v, 0);
        if (int rc = make_grid(b, b->plan.ball_r_max, b->ball_R, b->gb))
            return rc;
    }
    if (do_rdf(b))
    {
        int const key = fa::tune_key(density, b->plan.rdf_r_max) + fa::kBatchTuneKey;
        int v = fa::tuned_variant(fa::TUNE_RDF_SELF, key);
        if (v < 0 && fa::tuning_enabled())
        {
            float best = 0;
            for (int cand = 0; cand < kRdfShapes; ++cand)
            {
                b->rdf_R = kRdfShape[cand][0];
                b->rdf_F = kRdfShape[cand][1];
                if (int rc = make_grid(b, b->plan.rdf_r_max, b->rdf_R, b->gr, b->rdf_F))
                    return rc;
                if (int rc = alloc_rdf_scratch(b))
                    return rc;
                float ms = 0;
                if (int rc = enqueue_rdf(b))
                    return rc;
                if (int rc = fa::time_launch([&] { return enqueue_rdf(b); }, &ms))
                    return rc;
                if (v < 0 || ms < best)
                {
                    best = ms;
                    v = cand;
                }
            }
            fa::record_variant(fa::TUNE_RDF_SELF, key, v);
        }
        b->rdf_R = kRdfShape[std::max(v, 0) % kRdfShapes][0];
        b->rdf_F = kRdfShape[std::max(v, 0) % kRdfShapes][1];
        if (int rc = make_grid(b, b->plan.rdf_r_max, b->rdf_R, b->gr, b->rdf_F))
            return rc;
        if (int rc = alloc_rdf_scratch(b))
            return rc;
    }
    b->tuned = true;
    return FA_OK;
}

int capture(fa_batch* b)
{
    drop_graph(b);
    bool const prof = fa::profiling();
    if (prof)
        fa_profile_enable(0);
    hipStream_t s = fa::stream();
    FA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_all(b);
    hipGraph_t g = nullptr;
    hipError_t const e = hipStreamEndCapture(s, &g);
    if (prof)
        fa_profile_enable(1);
    if (rc)
    {
        if (g)
            (void) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess)
        return fa::fail(FA_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    FA_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    return FA_OK;
}

} // namespace

int fa::batch_rdf_shape_count()
{
    return kRdfShapes;
}

fa_batch::~fa_batch()
{
    drop_graph(this);
    free_grid(gb);
    free_grid(gr);
    fa::device_free(d_points);
    fa::device_free(d_tile_frame);
    fa::device_free(d_counts);
    fa::device_free(d_offsets);
    fa::device_free(d_nb);
    fa::device_free(d_dist);
    fa::device_free(d_wts);
    fa::device_free(d_vec);
    fa::device_free(d_gkeys);
    fa::device_free(d_pos);
    fa::device_free(d_rows);
    fa::device_free(rs.nclus);
    fa::device_free(rs.table);
    fa::device_free(d_rdf);
    fa::device_free(d_T);
    fa::device_free(d_qlm);
    fa::device_free(d_ql);
    fa::device_free(d_partial);
    fa::device_free(d_order);
    fa::device_free(d_total_bonds);
    fa::device_free(d_frame_bonds);
    fa::device_free(d_scan);
    if (h_status)
        (void) hipHostFree(h_status);
}

extern "C" {

int fa_batch_create(uint32_t n_frames, const fa_box* boxes, const uint64_t* point_offsets, const fa_batch_plan* plan,
                    fa_batch** out)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    if (!out || !plan || (n_frames && (!boxes || !point_offsets)))
        return fa::fail(FA_ERR_INVALID, "fa_batch_create: null argument");
    *out = nullptr;
    if (point_offsets && point_offsets[0] != 0)
        return fa::fail(FA_ERR_INVALID, "point_offsets must start at 0");
    for (uint32_t f = 0; f < n_frames; ++f)
        if (point_offsets[f + 1] < point_offsets[f])
            return fa::fail(FA_ERR_INVALID, "point_offsets must be non-decreasing");
    uint64_t const total = n_frames ? point_offsets[n_frames] : 0;
    if (total >= (1ull << 31))
        return fa::fail(FA_ERR_INVALID, "a batch holds fewer than 2^31 points");
    if (plan->rdf_bins > 2048)
        return fa::fail(FA_ERR_INVALID, "batched RDF supports at most 2048 bins");
    if (plan->rdf_bins && !(plan->rdf_r_max > 0.0f))
        return fa::fail(FA_ERR_INVALID, "RDF requires r_max to be positive.");
    if (plan->l != 0xffffffffu && plan->l > 20)
        return fa::fail(FA_ERR_INVALID, "freud_amd supports l <= 20");
    // the ball query and the RDF search one periodic image per pair, as the
    // reference's AABBQuery requires (NeighborQuery.h:282-290)
    for (uint32_t f = 0; f < n_frames; ++f)
    {
        float npd[3];
        fa::nearest_plane_distance(boxes[f], npd);
        float const r = std::max(plan->ball_r_max, plan->rdf_bins ? plan->rdf_r_max : 0.0f);
        if (npd[0] <= r * 2.0 || npd[1] <= r * 2.0 || (!boxes[f].is2D && npd[2] <= r * 2.0))
            return fa::fail(FA_ERR_RUNTIME, "The AABBQuery r_max is too large for this box.");
    }
    auto* b = new fa_batch();
    b->n_frames = n_frames;
    b->total = uint32_t(total);
    b->boxes.assign(boxes, boxes + n_frames);
    b->pt_off.assign(point_offsets, point_offsets + n_frames + 1);
    if (!n_frames)
        b->pt_off.assign(1, 0);
    b->plan = *plan;
    auto fail_out = [&](int rc) {
        delete b;
        return rc;
    };
    std::vector<uint32_t> tile_frame;
    for (uint32_t f = 0; f < n_frames; ++f)
    {
        uint64_t const n = point_offsets[f + 1] - point_offsets[f];
        for (uint64_t t = 0; t < (n + fa::kTile - 1) / fa::kTile; ++t)
            tile_frame.push_back(f);
    }
    b->ntiles = uint32_t(tile_frame.size());
    size_t const np = std::max<size_t>(b->total, 1);
    if (int rc = fa::device_alloc((void**) &b->d_points, sizeof(float) * 3 * np))
        return fail_out(rc);
    if (int rc = fa::device_alloc((void**) &b->d_tile_frame, sizeof(uint32_t) * std::max<size_t>(b->ntiles, 1)))
        return fail_out(rc);
    if (b->ntiles)
    {
        hipError_t const e = hipMemcpy(b->d_tile_frame, tile_frame.data(), sizeof(uint32_t) * b->ntiles,
                                       hipMemcpyHostToDevice);
        if (e != hipSuccess)
            return fail_out(fa::fail(FA_ERR_DEVICE, hipGetErrorString(e)));
    }
    if (int rc = fa::device_alloc((void**) &b->d_counts, sizeof(uint32_t) * np))
        return fail_out(rc);
    if (int rc = fa::device_alloc((void**) &b->d_offsets, sizeof(uint32_t) * (np + 1)))
        return fail_out(rc);
    if (int rc = fa::device_alloc((void**) &b->d_total_bonds, sizeof(unsigned long long)))
        return fail_out(rc);
    if (int rc = fa::device_alloc((void**) &b->d_frame_bonds, sizeof(uint32_t) * std::max<uint32_t>(n_frames, 1)))
        return fail_out(rc);
    if (hipHostMalloc((void**) &b->h_status, sizeof(unsigned long long) * 4, hipHostMallocDefault) != hipSuccess)
        return fail_out(fa::fail(FA_ERR_DEVICE, "hipHostMalloc failed"));
    if (do_rdf(b))
    {
        if (int rc = fa::device_alloc((void**) &b->d_T, sizeof(float) * fa::rdf_table_floats(plan->rdf_bins)))
            return fail_out(rc);
        if (int rc = fa::thresholds_for(plan->rdf_bins, plan->rdf_r_max, b->d_T))
            return fail_out(rc);
        if (int rc = fa::device_alloc((void**) &b->d_rdf,
                                      sizeof(uint32_t) * std::max<uint64_t>(uint64_t(n_frames) * plan->rdf_bins, 1)))
            return fail_out(rc);
    }
    if (do_ql(b))
    {
        size_t const nm = 2 * plan->l + 1;
        if (int rc = fa::device_alloc((void**) &b->d_qlm, sizeof(float) * 2 * nm * np))
            return fail_out(rc);
        if (int rc = fa::device_alloc((void**) &b->d_ql, sizeof(float) * np))
            return fail_out(rc);
        if (int rc = fa::device_alloc((void**) &b->d_partial,
                                      sizeof(float) * 2 * nm * std::max<size_t>(b->ntiles, 1)))
            return fail_out(rc);
        if (int rc = fa::device_alloc((void**) &b->d_order, sizeof(float) * std::max<uint32_t>(n_frames, 1)))
            return fail_out(rc);
    }
    // first bond capacity: ideal-gas estimate of each frame's pairs, +25 %
    if (do_ball(b))
    {
        double est = 0;
        double const r = plan->ball_r_max;
        for (uint32_t f = 0; f < n_frames; ++f)
        {
            double const n = double(point_offsets[f + 1] - point_offsets[f]);
            double const v = fa::box_volume(boxes[f]);
            double const ball = boxes[f].is2D ? M_PI * r * r : 4.0 / 3.0 * M_PI * r * r * r;
            est += n * std::min(n, n / v * ball);
        }
        if (int rc = alloc_bonds(b, uint64_t(est * 1.25) + 1024))
            return fail_out(rc);
        // keys the one-walk query stages per point: enough for the densest
        // frame (none when a frame needs more than the 32 an LDS stage holds)
        uint32_t C = 0;
        for (uint32_t f = 0; f < n_frames; ++f)
        {
            double const v = std::max(double(fa::box_volume(boxes[f])), 1e-30);
            uint32_t const c = fa::batch_ball_slots(double(point_offsets[f + 1] - point_offsets[f]) / v,
                                                    boxes[f].is2D, float(r));
            if (c == 0)
            {
                C = 0;
                break;
            }
            C = std::max(C, c);
        }
        if (C && b->ntiles)
        {
            if (int rc = fa::device_alloc((void**) &b->d_pos, sizeof(uint32_t) * np))
                return fail_out(rc);
            if (int rc = fa::device_alloc((void**) &b->d_rows, sizeof(uint32_t) * size_t(b->ntiles) * fa::kTile * C))
                return fail_out(rc);
            b->ball_C = C;
        }
    }
    *out = b;
    return FA_OK;
}

int fa_batch_destroy(fa_batch* b)
{
    FA_API_LOCK;
    delete b;
    return FA_OK;
}

int fa_batch_points(fa_batch* b, float** d_points)
{
    FA_API_LOCK;
    if (!b || !d_points)
        return fa::fail(FA_ERR_INVALID, "null argument");
    *d_points = b->d_points;
    return FA_OK;
}

int fa_batch_run(fa_batch* b, int use_graph)
{
    FA_API_LOCK;
    if (!b)
        return fa::fail(FA_ERR_INVALID, "null batch");
    if (b->total == 0)
        return FA_OK;
    if (!b->tuned)
        if (int rc = tune(b))
            return rc;
    bool fitted = true;
    if (use_graph && b->exec)
    {
        FA_HIP(hipGraphLaunch(b->exec, fa::stream()));
        if (int rc = read_status(b, &fitted))
            return rc;
        if (fitted)
            return FA_OK;
    }
    else
    {
        if (int rc = enqueue_all(b))
            return rc;
        if (int rc = read_status(b, &fitted))
            return rc;
    }
    if (!fitted)
    {
        // the bonds did not fit: grow, run eagerly again, and re-capture
        drop_graph(b);
        if (int rc = alloc_bonds(b, b->last_total_bonds + b->last_total_bonds / 4 + 1024))
            return rc;
        if (int rc = enqueue_all(b))
            return rc;
        if (int rc = read_status(b, &fitted))
            return rc;
        if (!fitted)
            return fa::fail(FA_ERR_RUNTIME, "batch bond count changed between identical runs");
    }
    if (use_graph && !b->exec)
        if (int rc = capture(b))
            return rc;
    return FA_OK;
}

int fa_batch_num_bonds(fa_batch* b, uint64_t* per_frame)
{
    FA_API_LOCK;
    if (!b || !per_frame)
        return fa::fail(FA_ERR_INVALID, "null argument");
    if (!do_ball(b))
        return fa::fail(FA_ERR_INVALID, "the batch plan has no ball query");
    std::vector<uint32_t> h(b->n_frames);
    if (b->n_frames)
        FA_HIP(hipMemcpyAsync(h.data(), b->d_frame_bonds, sizeof(uint32_t) * b->n_frames, hipMemcpyDeviceToHost,
                              fa::stream()));
    FA_HIP(hipStreamSynchronize(fa::stream()));
    for (uint32_t f = 0; f < b->n_frames; ++f)
        per_frame[f] = h[f];
    return FA_OK;
}

int fa_batch_nlist(fa_batch* b, uint32_t frame, fa_nlist* out)
{
    FA_API_LOCK;
    if (!b || !out)
        return fa::fail(FA_ERR_INVALID, "null argument");
    if (frame >= b->n_frames)
        return fa::fail(FA_ERR_RANGE, "frame index out of range");
    if (!do_ball(b))
        return fa::fail(FA_ERR_INVALID, "the batch plan has no ball query");
    hipStream_t s = fa::stream();
    uint32_t const base = uint32_t(b->pt_off[frame]);
    uint32_t const n = uint32_t(b->pt_off[frame + 1] - b->pt_off[frame]);
    uint32_t be[2];
    FA_HIP(hipMemcpyAsync(&be[0], b->d_offsets + base, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipMemcpyAsync(&be[1], b->d_offsets + base + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    uint64_t const nb = be[1] - be[0];
    if (int rc = fa::nlist_reserve(out, nb, n))
        return rc;
    out->num_bonds = nb;
    out->num_points = n;
    if (nb)
    {
        FA_HIP(hipMemcpyAsync(out->neighbors, b->d_nb + 2 * size_t(be[0]), sizeof(uint32_t) * 2 * nb,
                              hipMemcpyDeviceToDevice, s));
        FA_HIP(hipMemcpyAsync(out->distances, b->d_dist + be[0], sizeof(float) * nb, hipMemcpyDeviceToDevice, s));
        FA_HIP(hipMemcpyAsync(out->weights, b->d_wts + be[0], sizeof(float) * nb, hipMemcpyDeviceToDevice, s));
        FA_HIP(hipMemcpyAsync(out->vectors, b->d_vec + 3 * size_t(be[0]), sizeof(float) * 3 * nb,
                              hipMemcpyDeviceToDevice, s));
    }
    FA_KLAUNCH(frame_nlist_kernel, dim3((n + 1 + 255) / 256), dim3(256), 0, s, b->d_offsets, base, n, out->offsets,
               out->counts, out->segments);
    FA_HIP(hipGetLastError());
    FA_HIP(hipStreamSynchronize(s));
    return FA_OK;
}

int fa_batch_rdf_counts(fa_batch* b, uint32_t** d_counts)
{
    FA_API_LOCK;
    if (!b || !d_counts)
        return fa::fail(FA_ERR_INVALID, "null argument");
    if (!do_rdf(b))
        return fa::fail(FA_ERR_INVALID, "the batch plan has no RDF");
    *d_counts = b->d_rdf;
    return FA_OK;
}

int fa_batch_ql(fa_batch* b, float** d_ql, float** d_qlm, float* order)
{
    FA_API_LOCK;
    if (!b)
        return fa::fail(FA_ERR_INVALID, "null argument");
    if (!do_ql(b))
        return fa::fail(FA_ERR_INVALID, "the batch plan has no Steinhardt order");
    if (d_ql)
        *d_ql = b->d_ql;
    if (d_qlm)
        *d_qlm = b->d_qlm;
    if (order && b->n_frames)
    {
        FA_HIP(hipMemcpyAsync(order, b->d_order, sizeof(float) * b->n_frames, hipMemcpyDeviceToHost, fa::stream()));
        FA_HIP(hipStreamSynchronize(fa::stream()));
    }
    return FA_OK;
}

int fa_batch_config(fa_batch* b, int* ball_R, int* rdf_R, int* rdf_F, uint64_t* capacity, int* graph_ready)
{
    FA_API_LOCK;
    if (!b)
        return fa::fail(FA_ERR_INVALID, "null argument");
    if (ball_R)
        *ball_R = b->ball_R;
    if (rdf_R)
        *rdf_R = b->rdf_R;
    if (rdf_F)
        *rdf_F = b->rdf_F;
    if (capacity)
        *capacity = b->capacity;
    if (graph_ready)
        *graph_ready = b->exec != nullptr;
    return FA_OK;
}

} // extern "C"
