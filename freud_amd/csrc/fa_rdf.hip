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
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>

#include "fa_common.h"

namespace {

using fa::AABB_FORMULA;
using fa::CELLQUERY_FORMULA;
using fa::DevBox;
using fa::DevGrid;
using fa::LINKCELL_FORMULA;

constexpr int kBlock = 256;
// word 3 of a raw buffer resource on gfx9-family GPUs (32-bit data format)
constexpr int kBufferWord3 = 0x00020000;
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

// Bin table: the nbins + 1 thresholds T, then inv_delta and x_fast (floats),
// then a kLut-entry lookup table over r_sq in [0, T[nbins]).  The cell of an
// r_sq x is uint32(x * inv_delta); inv_delta is the largest float not above
// kLut / T[nbins] for which every x < T[nbins] has a cell below kLut.  A cell's
// entry is a bin k such that every x of the cell inside [T[0], T[nbins]) has
// bin k or k + 1, so one compare with T[k + 1] gives the reference's bin.
// Cells that span more bins (near r = 0, where bins are thin in r_sq) hold
// kLutNone; they all lie below x_fast (the first r_sq of the cell after the
// last such cell), and r_sq below x_fast take the square-root path.  Entries
// are found on the host by bisection over float bit patterns with the device's
// own float expression for the cell.
constexpr uint32_t kLut = 4096;
constexpr uint16_t kLutNone = 0xffffu;
constexpr uint32_t kLutMaxBins = 16382; // LDS entries hold 4 * (k + 1) < kLutNone

uint32_t lut_cell(float x, float inv_delta)
{
    volatile float p = x * inv_delta;
    float const q = p;
    return q >= float(kLut) ? kLut : uint32_t(q);
}

__host__ __device__ constexpr uint32_t table_floats(uint32_t nbins)
{
    return nbins + 3 + uint32_t((kLut * sizeof(uint16_t) + sizeof(float) - 1) / sizeof(float));
}

std::vector<float> build_table(uint32_t nbins, float vmin, float vmax)
{
    float const width = (vmax - vmin) / float(nbins);
    float const inv_width = 1.0f / width;
    std::vector<float> tab(table_floats(nbins), 0.0f);
    float* T = tab.data();
    for (uint32_t k = 0; k <= nbins; ++k)
        T[k] = first_rsq_with_bin_at_least(int(k), vmin, vmax, inv_width, nbins);
    float const tn = T[nbins];
    float inv_delta = 0.0f;
    if (tn > 0.0f && std::isfinite(tn))
    {
        volatile float idv = float(kLut) / tn;
        inv_delta = idv;
        float const last = std::nextafter(tn, 0.0f);
        while (inv_delta > 0.0f && lut_cell(last, inv_delta) >= kLut)
            inv_delta = std::nextafter(inv_delta, 0.0f);
    }
    T[nbins + 1] = inv_delta;
    T[nbins + 2] = std::numeric_limits<float>::infinity(); // x_fast
    uint16_t* lut = reinterpret_cast<uint16_t*>(T + nbins + 3);
    auto bin_of = [&](float x) { // reference bin of x in [T0, Tn)
        return uint32_t(std::upper_bound(T, T + nbins + 1, x) - T) - 1;
    };
    if (!(inv_delta > 0.0f) || nbins > kLutMaxBins)
    {
        for (uint32_t c = 0; c < kLut; ++c)
            lut[c] = kLutNone;
        return tab;
    }
    // first float x >= 0 whose cell is >= c
    auto first_in = [&](uint32_t c) {
        uint32_t lo = 0, hi = __builtin_bit_cast(uint32_t, tn);
        if (lut_cell(tn, inv_delta) < c)
            return tn;
        while (lo < hi)
        {
            uint32_t const mid = lo + (hi - lo) / 2;
            if (lut_cell(__builtin_bit_cast(float, mid), inv_delta) >= c)
                hi = mid;
            else
                lo = mid + 1;
        }
        return __builtin_bit_cast(float, lo);
    };
    float start = first_in(0);
    float x_fast = 0.0f;
    for (uint32_t c = 0; c < kLut; ++c)
    {
        float const next = (c + 1 < kLut) ? first_in(c + 1) : tn;
        float a = std::max(start, T[0]);
        float b = std::nextafter(std::min(next, tn), 0.0f); // last x of the cell below Tn
        if (!(a <= b))
            lut[c] = 0; // no r_sq of this cell is binned
        else
        {
            uint32_t const ka = bin_of(a), kb = bin_of(b);
            lut[c] = (kb - ka <= 1) ? uint16_t(ka) : kLutNone;
            if (lut[c] == kLutNone)
                x_fast = next;
        }
        start = next;
    }
    T[nbins + 2] = x_fast;
    return tab;
}

struct Thresholds
{
    float* dev = nullptr;
    size_t cap = 0;
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
    std::vector<float> const tab = build_table(nbins, vmin, vmax);
    if (!g_thr.dev || g_thr.cap < tab.size())
    {
        fa::device_free(g_thr.dev);
        g_thr.dev = nullptr;
        if (int rc = fa::device_alloc((void**) &g_thr.dev, sizeof(float) * tab.size()))
            return rc;
        g_thr.cap = tab.size();
    }
    FA_HIP(hipMemcpy(g_thr.dev, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice));
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

// LDS form of the bin table: T[0..nbins], inv_delta and x_fast, then the kLut
// cell entries as uint32 LDS byte addresses of T[k + 1] (cells below x_fast
// point at T[1]: their lookups are discarded), then the histogram.
__host__ __device__ constexpr uint32_t lds_table_floats(uint32_t nbins)
{
    return nbins + 3 + kLut;
}

typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) const uint32_t lds_cu32;
typedef __attribute__((address_space(3))) const float lds_cf32;

__device__ inline uint32_t lds_addr(const void* p)
{
    return uint32_t(uintptr_t((const __attribute__((address_space(3))) char*) (p)));
}

// Binning targets of the row kernels; add2 bins two r_sq values whose flags
// say they are counted.  LutBins: table and histogram in LDS (up to
// kMaxLdsBins bins).  For r_sq >= x_fast the entry e of the cell is the
// address of T[k + 1], so one LDS read and one compare pick bin k or k + 1 and
// the histogram slot is e (+ 4) + hist_c; both lookups are issued for every
// lane before the masked atomics, so their LDS latencies overlap.  r_sq below
// x_fast (pairs at r near 0) take the square-root path in a branch that the
// wave skips when none of its lanes needs it.  GlobalBins: thresholds and counts in device memory.
struct LutBins
{
    const float* sT;
    uint32_t nbins;
    float inv_delta, x_fast, vmin, inv_width;
    uint32_t lut, hist, hist_c;
    __device__ LutBins(const float* t, uint32_t nb, float vmn, float iw, const uint32_t* h)
        : sT(t), nbins(nb), inv_delta(t[nb + 1]), x_fast(t[nb + 2]), vmin(vmn), inv_width(iw),
          lut(lds_addr(t + nb + 3)), hist(lds_addr(h)), hist_c(lds_addr(h) - lds_addr(t) - 4u)
    {
    }
    __device__ float t0() const { return sT[0]; }
    __device__ float tn() const { return sT[nbins]; }
    __device__ __forceinline__ uint32_t fast_slot(float x) const
    {
        uint32_t const e = *(lds_cu32*) (uintptr_t) (lut + 4u * min(uint32_t(x * inv_delta), kLut - 1u));
        float const t = *(lds_cf32*) (uintptr_t) e;
        return e + (x >= t ? 4u : 0u) + hist_c;
    }
    __device__ static void atomic(uint32_t a, uint32_t w)
    {
        __atomic_fetch_add((lds_u32*) (uintptr_t) a, w, __ATOMIC_RELAXED);
    }
    // one r_sq of any value, weight 1
    __device__ void add1(float x) const
    {
        uint32_t const bin = bin_of_rsq(x, sT, nbins, vmin, inv_width);
        if (bin < nbins)
            atomic(hist + 4u * bin, 1u);
    }
    __device__ __forceinline__ void add2(float x0, bool ok0, float x1, bool ok1, uint32_t w) const
    {
        bool const f0 = ok0 && x0 >= x_fast, f1 = ok1 && x1 >= x_fast;
        // fast_slot of both values, the two table reads of each issued together
        uint32_t const e0 = *(lds_cu32*) (uintptr_t) (lut + 4u * min(uint32_t(x0 * inv_delta), kLut - 1u));
        uint32_t const e1 = *(lds_cu32*) (uintptr_t) (lut + 4u * min(uint32_t(x1 * inv_delta), kLut - 1u));
        float const t0v = *(lds_cf32*) (uintptr_t) e0;
        float const t1v = *(lds_cf32*) (uintptr_t) e1;
        uint32_t const a0 = e0 + (x0 >= t0v ? 4u : 0u) + hist_c, a1 = e1 + (x1 >= t1v ? 4u : 0u) + hist_c;
        asm volatile("" ::"v"(a0), "v"(a1)); // both lookups ahead of the masked atomics
        if (f0)
            atomic(a0, w);
        if (f1)
            atomic(a1, w);
        bool const s0 = ok0 && !f0, s1 = ok1 && !f1;
        if (s0 || s1) // a divergent branch: skipped when no lane takes it
        {
            if (s0)
                atomic(hist + 4u * bin_in_range(x0, sT, nbins, vmin, inv_width), w);
            if (s1)
                atomic(hist + 4u * bin_in_range(x1, sT, nbins, vmin, inv_width), w);
        }
    }
};

struct GlobalBins
{
    const float* T;
    uint32_t nbins;
    float vmin, inv_width;
    uint32_t* counts;
    __device__ float t0() const { return T[0]; }
    __device__ float tn() const { return T[nbins]; }
    __device__ void add1(float x) const
    {
        uint32_t const bin = bin_of_rsq(x, T, nbins, vmin, inv_width);
        if (bin < nbins)
            atomicAdd(counts + bin, 1u);
    }
    __device__ void add2(float x0, bool ok0, float x1, bool ok1, uint32_t w) const
    {
        if (ok0)
            atomicAdd(counts + bin_in_range(x0, T, nbins, vmin, inv_width), w);
        if (ok1)
            atomicAdd(counts + bin_in_range(x1, T, nbins, vmin, inv_width), w);
    }
};


// LDS copy of the bin table (lds_table_floats layout); hist_prologue also
// clears the LDS histogram that follows it.  Kernels with device-memory counts
// read the table from device memory instead.
__device__ inline void table_prologue(float* sT, const float* __restrict__ T, uint32_t nbins)
{
    uint32_t const nt = nbins + 3;
    for (uint32_t t = threadIdx.x; t < nt; t += blockDim.x)
        sT[t] = T[t];
    const uint16_t* glut = reinterpret_cast<const uint16_t*>(T + nt);
    uint32_t* slut = reinterpret_cast<uint32_t*>(sT + nt);
    uint32_t const tb = lds_addr(sT);
    for (uint32_t t = threadIdx.x; t < kLut; t += blockDim.x)
    {
        uint32_t const k = glut[t];
        slut[t] = tb + 4u * (k == kLutNone ? 1u : k + 1u);
    }
}

template<bool LDS_HIST>
__device__ inline void hist_prologue(uint32_t* hist, float* sT, const float* __restrict__ T, uint32_t nbins)
{
    if (!LDS_HIST)
        return;
    table_prologue(sT, T, nbins);
    for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
        hist[t] = 0;
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

// Exclusive prefix sum over the 64 lanes of a wave.
__device__ inline uint32_t wave_exclusive_scan(uint32_t v)
{
    uint32_t const lane = threadIdx.x & 63u;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        uint32_t const y = __shfl_up(x, d);
        if (lane >= uint32_t(d))
            x += y;
    }
    return x - v;
}

// One wave per cluster (up to 8 consecutive cell-ordered points of a row).
// The stencil spans cells [cx_lo - R, cx_hi + R] x [cy - R, cy + R] x
// [cz - R, cz + R] (cells at least r_max / R wide), walked as contiguous pieces
// of a row that share one periodic shift.  A pair related without a periodic
// shift is found from both of its points with bitwise-equal r_sq
// (x - y == -(y - x)), so it is counted once, from the point earlier in cell
// order, with weight 2; pairs across the box edge are counted from both sides
// with weight 1, as the reference's (q + image) arithmetic rounds differently
// for the two directions.
//
// Lane t first resolves stencil row t (its up to 3 periodic pieces and their
// cell-range bounds), so the cstart reads of all rows are in flight together.
// Unshifted pieces (the bulk) are concatenated in row order and taken 64
// candidates at a time, one per lane, against the cluster's query points held
// in scalar registers: every candidate is loaded once per wave and only the
// last block of a cluster is partly empty.  Shifted pieces (clusters near the
// box edge) run as 8 x 8 chunks: lane (a, b) = (lane / 8, lane % 8) pairs query
// a with point b of each 8-point chunk; the chunk loop runs on the scalar unit
// with the next two chunks in flight, and only the last chunk of a piece (which
// may run past it; the point arrays are padded for it) tests positions per
// lane.  Query lanes and candidates that do not exist carry a NaN coordinate,
// which fails every range compare.
template<int R, bool IS2D, bool HAS_LO, class Bins>
__device__ inline void rows_body(const DevBox& b, const DevGrid& g, uint32_t row, uint32_t A0, uint32_t na,
                                 int cx_lo, int cx_hi, int RX, float q_rmax_sq, float lo, float hi, const Bins& B)
{
    constexpr int kRowsY = 2 * R + 1;
    constexpr int Rz = IS2D ? 0 : R;
    constexpr int nrows = kRowsY * (2 * Rz + 1);
    uint32_t const lane = threadIdx.x & 63u;
    int const dim0 = g.dim[0], dim1 = g.dim[1], dim2 = g.dim[2];
    // row trimming: a candidate row (dy, dz) is at least (|dy|-1) cell extents away
    // along y and (|dz|-1) along z; conservative margins cover the rounding of
    // the cell assignment
    float const r_lim = sqrtf(q_rmax_sq) * 1.0001f;
    float const wx = g.cw[0] * 0.9999f, wy = g.cw[1] * 0.9999f, wz = g.cw[2] * 0.9999f;
    bool const ortho = b.xy == 0.0f && b.xz == 0.0f && b.yz == 0.0f;
    int const cy0 = int(row % uint32_t(dim1));
    int const cz0 = int(row / uint32_t(dim1));
    uint32_t const a = lane >> 3, bl = lane & 7u;
    uint32_t const pos_a = A0 + min(a, na - 1u);
    float4 pa = g.cpos[pos_a];
    if (IS2D)
        pa.z = 0.0f;
    if (a >= na)
        pa.x = __builtin_nanf("");
    // the cluster's query points, wave-uniform (NaN beyond na)
    float qsx[8], qsy[8], qsz[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
        qsx[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pa.x), 8 * k));
        qsy[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pa.y), 8 * k));
        qsz[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pa.z), 8 * k));
    }
    // The cluster's extent (orthorhombic boxes): a stencil row or cell whose
    // slab lies farther than r_max from every query point of the cluster holds
    // no pair.  Cell c along axis k covers [lo_k + c w_k, lo_k + (c + 1) w_k)
    // up to the rounding of the fractional coordinate, which the margins
    // (1e-5 of the box length) cover; unwrapped cell indices place periodic
    // images in the same frame.
    float xmn = qsx[0], xmx = qsx[0], ymn = qsy[0], ymx = qsy[0], zmn = qsz[0], zmx = qsz[0];
#pragma unroll
    for (int k = 1; k < 8; ++k)
    { // qsx is NaN beyond na (ignored by fminf/fmaxf); qsy, qsz repeat point na - 1
        xmn = fminf(xmn, qsx[k]);
        xmx = fmaxf(xmx, qsx[k]);
        ymn = fminf(ymn, qsy[k]);
        ymx = fmaxf(ymx, qsy[k]);
        zmn = fminf(zmn, qsz[k]);
        zmx = fmaxf(zmx, qsz[k]);
    }
    float const mgx = 1e-5f * b.L[0], mgy = 1e-5f * b.L[1], mgz = 1e-5f * b.L[2];
    // chunk loads through a buffer resource: wave-uniform chunk offset in an
    // SGPR, the lane's offset in a VGPR (point arrays stay below 4 GiB)
    uint32_t const voff = bl * uint32_t(sizeof(float4));
    __amdgpu_buffer_rsrc_t const rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(g.cpos), short(0), -1, kBufferWord3);
    {
        // pairs within the cluster: lane (a, b) with a < b < na, no periodic shift
        float4 const c = g.cpos[A0 + min(bl, na - 1u)];
        float const rx = c.x - pa.x;
        float const ry = c.y - pa.y;
        float x = rx * rx + ry * ry;
        if constexpr (!IS2D)
        {
            float const rz = c.z - pa.z;
            x = x + rz * rz;
        }
        bool ok = x < hi && bl > a && bl < na;
        if constexpr (HAS_LO)
            ok = ok && x >= lo;
        B.add2(x, ok, 0.0f, false, 2u);
    }
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
            float gy = float(max(0, abs(dy) - 1)) * wy, gz = float(max(0, abs(dz) - 1)) * wz;
            int xr = RX;
            bool reach = true;
            // cell range along x from the cluster's x extent (all cells otherwise)
            int ux_lo = -dim0, ux_hi = 2 * dim0 - 1;
            if (ortho)
            {
                float const ys = b.lo[1] + float(uy) * g.cw[1];
                gy = fmaxf(gy, fmaxf(ys - ymx, ymn - (ys + g.cw[1])) - mgy);
                if constexpr (!IS2D)
                {
                    float const zs = b.lo[2] + float(uz) * g.cw[2];
                    gz = fmaxf(gz, fmaxf(zs - zmx, zmn - (zs + g.cw[2])) - mgz);
                }
                float const rem = r_lim * r_lim - (gy * gy + gz * gz);
                reach = rem > 0.0f;
                if (reach && wx > 0.0f)
                {
                    float const h = sqrtf(rem);
                    xr = min(RX, int(h / wx) + 1);
                    float const ix = 1.0f / g.cw[0];
                    ux_lo = int(fmaxf(floorf((xmn - h - mgx - b.lo[0]) * ix), float(-dim0)));
                    ux_hi = int(fminf(floorf((xmx + h + mgx - b.lo[0]) * ix), float(2 * dim0 - 1)));
                }
            }
            else
                reach = fmaxf(gy, gz) < r_lim;
            if (reach && sz >= -1 && sz <= 1 && sy >= -1 && sy <= 1)
            {
                uint32_t const rowbase = uint32_t((uz - sz * dim2) * dim1 + (uy - sy * dim1)) * uint32_t(dim0);
                // shifts beyond one box length are not images the reference visits,
                // so the clipped range has at most 3 pieces (shift -1, 0, 1)
                int ux = max(max(cx_lo - xr, -dim0), ux_lo);
                int const uxend = min(min(cx_hi + xr, 2 * dim0 - 1), ux_hi);
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
        // Unshifted pieces (image code 13): one per row at most, from the point
        // after the cluster on (pairs are counted from their earlier point, weight
        // 2; pairs within the cluster were counted above).  Their candidates are
        // concatenated in row order and taken 64 at a time, one candidate per
        // lane, against the cluster's query points in turn: each candidate is
        // loaded once per wave.
        uint32_t b13 = 0u, l13 = 0u;
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            if (pc[k] == 13)
            {
                uint32_t const bb = max(pb[k], A0 + na);
                if (pe[k] > bb)
                {
                    b13 = bb;
                    l13 = pe[k] - bb;
                }
            }
        }
        uint32_t const p13 = wave_exclusive_scan(l13); // virtual start of row t's candidates
        uint32_t const e13 = p13 + l13;
        uint32_t const sh13 = b13 - p13; // position = virtual index + sh13 within row t
        uint32_t const tot = uint32_t(__builtin_amdgcn_readlane(int(e13), 63));
        uint64_t live = __builtin_amdgcn_ballot_w64(l13 > 0u); // rows not yet exhausted
        // positions of the candidates [v0, v0 + 64) (one per lane); rows that end
        // inside the block are retired
        auto locate = [&](uint32_t v0) {
            uint32_t const v = v0 + lane;
            uint32_t pos = v + uint32_t(__builtin_amdgcn_readlane(int(sh13), __builtin_ctzll(live)));
            for (uint64_t m = live & (live - 1u); m; m &= m - 1u)
            {
                int const r = __builtin_ctzll(m);
                uint32_t const sp = uint32_t(__builtin_amdgcn_readlane(int(p13), r));
                if (sp >= v0 + 64u)
                    break;
                uint32_t const p2 = v + uint32_t(__builtin_amdgcn_readlane(int(sh13), r));
                pos = v >= sp ? p2 : pos;
            }
            while (live && uint32_t(__builtin_amdgcn_readlane(int(e13), __builtin_ctzll(live))) <= v0 + 64u)
                live &= live - 1u;
            return pos;
        };
        auto fetch = [&](uint32_t pos) {
            return __builtin_amdgcn_raw_buffer_load_b128(rs, pos * uint32_t(sizeof(float4)), 0u, 0);
        };
        // the next block's candidates are loaded before the current block is binned
        auto cvn = fetch(tot ? locate(0u) : 0u);
        for (uint32_t v0 = 0; v0 < tot; v0 += 64u)
        {
            auto const cv = cvn;
            if (v0 + 64u < tot)
                cvn = fetch(locate(v0 + 64u));
            float const cx = v0 + lane < tot ? __uint_as_float(cv[0]) : __builtin_nanf("");
            float const cy = __uint_as_float(cv[1]);
            float const cz = IS2D ? 0.0f : __uint_as_float(cv[2]);
            auto rsq = [&](int k) {
                float const rx = cx - qsx[k];
                float const ry = cy - qsy[k];
                if constexpr (IS2D)
                    return rx * rx + ry * ry;
                float const rz = cz - qsz[k];
                return (rx * rx + ry * ry) + rz * rz;
            };
#pragma unroll
            for (int k = 0; k < 8; k += 2)
            {
                if (uint32_t(k) >= na)
                    break;
                float const x0 = rsq(k), x1 = rsq(k + 1); // x1 is NaN when k + 1 >= na
                bool ok0 = x0 < hi, ok1 = x1 < hi;
                if constexpr (HAS_LO)
                {
                    ok0 = ok0 && x0 >= lo;
                    ok1 = ok1 && x1 >= lo;
                }
                B.add2(x0, ok0, x1, ok1, 2u);
            }
        }
        // shifted pieces (across the box edge), weight 1, as 8 x 8 chunks
        for (int r = 0; r < nr; ++r)
        {
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                int const code = __builtin_amdgcn_readlane(pc[k], r);
                if (code < 0 || code == 13)
                    continue;
                uint32_t const beg = uint32_t(__builtin_amdgcn_readlane(int(pb[k]), r));
                uint32_t const end = uint32_t(__builtin_amdgcn_readlane(int(pe[k]), r));
                if (beg >= end)
                    continue;
                float const qx = pa.x + b.img[code][0];
                float const qy = pa.y + b.img[code][1];
                float const qz = pa.z + b.img[code][2];
                auto ld = [&](uint32_t p) {
                    auto const v = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, p * uint32_t(sizeof(float4)), 0);
                    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), 0.0f);
                };
                auto rsq = [&](const float4& c) {
                    float const rx = c.x - qx;
                    float const ry = c.y - qy;
                    if constexpr (IS2D)
                        return rx * rx + ry * ry;
                    float const rz = c.z - qz;
                    return (rx * rx + ry * ry) + rz * rz;
                };
                auto two = [&](const float4& u, const float4& v, uint32_t jb) {
                    float const x0 = rsq(u), x1 = rsq(v);
                    bool ok0 = x0 < hi, ok1 = x1 < hi;
                    if constexpr (HAS_LO)
                    {
                        ok0 = ok0 && x0 >= lo;
                        ok1 = ok1 && x1 >= lo;
                    }
                    if (jb + 16u > end)
                    {
                        ok0 = ok0 && jb + bl < end;
                        ok1 = ok1 && jb + 8u + bl < end;
                    }
                    B.add2(x0, ok0, x1, ok1, 1u);
                };
                // two register sets in turn: the loads of the next 16 points are in
                // flight while the current 16 are binned
                float4 c0 = ld(beg), c1 = ld(beg + 8u);
                for (uint32_t jb = beg;;)
                {
                    float4 const d0 = ld(jb + 16u), d1 = ld(jb + 24u);
                    two(c0, c1, jb);
                    jb += 16u;
                    if (jb >= end)
                        break;
                    c0 = ld(jb + 16u);
                    c1 = ld(jb + 24u);
                    two(d0, d1, jb);
                    jb += 16u;
                    if (jb >= end)
                        break;
                }
            }
        }
    }
}

// The pair work of one cluster entry e = (frame-local row, start, count,
// cx_lo | cx_hi << 16) for one wave (e is wave-uniform).
template<int R, class Bins>
__device__ inline void rdf_rows_wave(const DevBox& b, const DevGrid& g, uint4 e, int RX, float q_rmax_sq,
                                     float q_rmin_sq, const Bins& B)
{
    uint32_t const row = __builtin_amdgcn_readfirstlane(e.x);
    uint32_t const A0 = __builtin_amdgcn_readfirstlane(e.y);
    uint32_t const na = __builtin_amdgcn_readfirstlane(e.z);
    uint32_t const cx = __builtin_amdgcn_readfirstlane(e.w);
    int const cx_lo = int(cx & 0xffffu), cx_hi = int(cx >> 16);
    if (*g.outside)
    {
        // some point of the grid lies outside the box: lane a searches the stencil
        // of query point a through the general cell walk (AABB formula with each
        // point's own image offset) and counts each ordered pair once
        uint32_t const lane = threadIdx.x & 63u;
        if (lane < na)
        {
            float4 const q = g.cpos[A0 + lane];
            uint32_t const qi = __float_as_uint(q.w);
            fa::for_each_candidate<AABB_FORMULA>(
                b, g, q.x, q.y, q.z, R, false,
                [&](uint32_t j, float, float, float, float r_sq, int) {
                    if (j != qi && r_sq < q_rmax_sq && r_sq >= q_rmin_sq)
                        B.add1(r_sq);
                },
                RX);
        }
        return;
    }
    // pairs counted: q_rmin_sq <= r_sq < q_rmax_sq (the query) and inside the histogram
    float const lo = fmaxf(q_rmin_sq, B.t0()), hi = fminf(q_rmax_sq, B.tn());
    if (b.is2D)
    {
        if (lo > 0.0f)
            rows_body<R, true, true>(b, g, row, A0, na, cx_lo, cx_hi, RX, q_rmax_sq, lo, hi, B);
        else
            rows_body<R, true, false>(b, g, row, A0, na, cx_lo, cx_hi, RX, q_rmax_sq, lo, hi, B);
    }
    else
    {
        if (lo > 0.0f)
            rows_body<R, false, true>(b, g, row, A0, na, cx_lo, cx_hi, RX, q_rmax_sq, lo, hi, B);
        else
            rows_body<R, false, false>(b, g, row, A0, na, cx_lo, cx_hi, RX, q_rmax_sq, lo, hi, B);
    }
}

template<int R, bool LDS_HIST>
__global__ __launch_bounds__(kBlock, 8) void rdf_rows_kernel(DevBox b, DevGrid g, const uint4* __restrict__ table,
                                                          const uint32_t* __restrict__ ncl_ptr, int RX, float q_rmax_sq,
                                                          float q_rmin_sq, const float* __restrict__ T,
                                                          uint32_t nbins, float vmin, float inv_width,
                                                          uint32_t* __restrict__ counts)
{
    extern __shared__ float smem[];
    float* sT = smem;
    uint32_t* hist = reinterpret_cast<uint32_t*>(smem + lds_table_floats(nbins));
    hist_prologue<LDS_HIST>(hist, sT, T, nbins);
    constexpr uint32_t kWaves = kBlock / 64;
    uint32_t const I = fa::xcd_tile(blockIdx.x, gridDim.x) * kWaves + (threadIdx.x >> 6);
    if (I < *ncl_ptr)
    {
        if constexpr (LDS_HIST)
            rdf_rows_wave<R>(b, g, table[I], RX, q_rmax_sq, q_rmin_sq, LutBins(sT, nbins, vmin, inv_width, hist));
        else
            rdf_rows_wave<R>(b, g, table[I], RX, q_rmax_sq, q_rmin_sq,
                             GlobalBins{T, nbins, vmin, inv_width, counts});
    }
    hist_epilogue<LDS_HIST>(hist, nbins, counts);
}

// ------------------------------------------------------------------ generic (per query point)
// Query points in cell order straight from the grid (query points == points).
template<int FORMULA, bool LDS_HIST>
__global__ __launch_bounds__(kBlock) void rdf_self_kernel(DevBox b, DevGrid g, uint32_t n, float q_rmax_sq,
                                                          float q_rmin_sq, int exclude_ii, int R,
                                                          const float* __restrict__ T, uint32_t nbins, float vmin,
                                                          float inv_width, uint32_t* __restrict__ counts,
                                                          fa::CQGrid cq)
{
    extern __shared__ float smem[];
    float* sT = smem;
    uint32_t* hist = reinterpret_cast<uint32_t*>(smem + lds_table_floats(nbins));
    uint32_t* h = LDS_HIST ? hist : counts;
    const float* tab = LDS_HIST ? sT : T;
    hist_prologue<LDS_HIST>(hist, sT, T, nbins);
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
    {
        float4 const q = g.cpos[t];
        uint32_t const qi = __float_as_uint(q.w);
        fa::visit_bonds<FORMULA>(b, g, cq, q.x, q.y, q.z, R,
                                        [&](uint32_t j, float, float, float, float r_sq, int) {
                                            if (r_sq < q_rmax_sq && r_sq >= q_rmin_sq && !(exclude_ii && j == qi))
                                            {
                                                uint32_t const bin = bin_of_rsq(r_sq, tab, nbins, vmin, inv_width);
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
                                                           uint32_t* __restrict__ counts, fa::CQGrid cq)
{
    extern __shared__ float smem[];
    float* sT = smem;
    uint32_t* hist = reinterpret_cast<uint32_t*>(smem + lds_table_floats(nbins));
    uint32_t* h = LDS_HIST ? hist : counts;
    const float* tab = LDS_HIST ? sT : T;
    hist_prologue<LDS_HIST>(hist, sT, T, nbins);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x)
    {
        float const qx = qpts[3 * i], qy = qpts[3 * i + 1], qz = qpts[3 * i + 2];
        fa::visit_bonds<FORMULA>(b, g, cq, qx, qy, qz, R,
                                        [&](uint32_t j, float, float, float, float r_sq, int) {
                                            if (r_sq < q_rmax_sq && r_sq >= q_rmin_sq && !(exclude_ii && j == i))
                                            {
                                                uint32_t const bin = bin_of_rsq(r_sq, tab, nbins, vmin, inv_width);
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
// Launch variants of the self RDF: cells at least r_max / R wide across the
// rows and F times narrower along them (x), so a row's candidate range follows
// the sphere's chord more closely.
struct RowVariant
{
    int R, F;
};
constexpr RowVariant kRowVariant[] = {{2, 1}, {3, 1}, {4, 1}, {2, 2}, {2, 4}, {3, 2}, {2, 8}, {3, 4}, {2, 16}};
constexpr int kRowVariants = int(sizeof(kRowVariant) / sizeof(kRowVariant[0]));

// Grid shape of the row-cluster self RDF for variant rv: cells at least
// r_max / R wide across the rows, F times narrower along x (F falls back to 1
// when the refined grid would hold more than 8 cells per point).  Returns
// false when the kernel does not apply (rows too long for the cluster
// encoding, or too many points for 32-bit chunk offsets).
bool rows_grid_dims(const fa_box& box, uint32_t n, float q_r_max, RowVariant rv, int dims[3], int* F_out)
{
    fa::grid_dims_for(box, n, q_r_max, rv.R, dims);
    int F = rv.F;
    if (dims[0] * F > 65535 || double(dims[0]) * F * dims[1] * dims[2] > std::max<double>(8.0 * n, 64.0))
        F = 1;
    dims[0] *= F;
    *F_out = F;
    // cluster entries hold cell x in 16 bits; chunk loads use 32-bit byte offsets
    return !(dims[0] > 65535 || uint64_t(n) + fa::kCposPad > (1ull << 28));
}

// Row-cluster buffers of a grid with nrows rows over n points: cluster counts
// and offsets (nrows + 1 each), the cluster table (max_clusters uint4) and
// the scan's temporary storage.
struct RowsBuffers
{
    uint32_t* nclus;
    uint32_t* cl_off;
    uint4* table;
    void* scan_tmp; // nullptr: the library's scan scratch
    size_t scan_bytes;
};
inline uint64_t rows_max_clusters(uint32_t n, uint32_t nrows) { return uint64_t(n) / 8 + nrows; }

// The cluster table of a built grid, then the row-cluster kernel.
int launch_rows(const DevBox& db, const DevGrid& g, const int dims[3], int R, int F, uint32_t n,
                const RowsBuffers& B, float qmax_sq, float qmin_sq, const float* T, uint32_t bins, float r_min,
                float inv_width, uint32_t* d_counts)
{
    hipStream_t s = fa::stream();
    uint32_t const nrows = uint32_t(dims[1]) * uint32_t(dims[2]);
    uint64_t const max_clusters = rows_max_clusters(n, nrows);
    unsigned const rblocks = (nrows + kBlock - 1) / kBlock;
    FA_KLAUNCH(row_cluster_count_kernel, dim3(rblocks), dim3(kBlock), 0, s, g.cstart, uint32_t(dims[0]), nrows,
               B.nclus);
    if (B.scan_tmp)
    {
        if (int rc = fa::exclusive_scan_u32_async(B.nclus, B.cl_off, nrows, B.scan_tmp, B.scan_bytes))
            return rc;
    }
    else if (int rc = fa::exclusive_scan_u32(B.nclus, B.cl_off, nrows, nullptr))
        return rc;
    unsigned const tblocks = unsigned((max_clusters + kBlock - 1) / kBlock);
    FA_KLAUNCH(row_cluster_table_kernel, dim3(tblocks), dim3(kBlock), 0, s, g.cstart, uint32_t(dims[0]), nrows,
               B.cl_off, uint32_t(max_clusters), B.table);
    unsigned const wblocks = unsigned((max_clusters + kBlock / 64 - 1) / (kBlock / 64));
    bool const lds = bins <= kMaxLdsBins;
    size_t const shmem = lds ? sizeof(float) * lds_table_floats(bins) + sizeof(uint32_t) * bins : 0;
#define FA_ROWS(RR)                                                                                                 \
    case RR:                                                                                                        \
        if (lds)                                                                                                    \
            FA_KLAUNCH((rdf_rows_kernel<RR, true>), dim3(wblocks), dim3(kBlock), shmem, s, db, g, B.table,          \
                       B.cl_off + nrows, R * F, qmax_sq, qmin_sq, T, bins, r_min, inv_width, d_counts);            \
        else                                                                                                        \
            FA_KLAUNCH((rdf_rows_kernel<RR, false>), dim3(wblocks), dim3(kBlock), shmem, s, db, g, B.table,         \
                       B.cl_off + nrows, R * F, qmax_sq, qmin_sq, T, bins, r_min, inv_width, d_counts);            \
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

// Self RDF through the row-cluster kernel with cells at least r_max / R wide
// (r_max / (R F) along x).  Sets *handled = false (and does nothing) when the
// kernel does not apply: rows too long for the cluster encoding, or too many
// points (points outside the box are handled inside the kernel).
int run_rows(const RowsArgs& a, RowVariant rv, uint32_t* d_counts, bool* handled)
{
    *handled = false;
    fa_nq* nq = a.nq;
    int dims[3], F = 1;
    if (!rows_grid_dims(nq->box, nq->n, a.q_r_max, rv, dims, &F))
        return FA_OK;
    if (int rc = fa::ensure_grid(nq, dims))
        return rc;
    *handled = true;
    uint32_t const nrows = uint32_t(dims[1]) * uint32_t(dims[2]);
    uint64_t const max_clusters = rows_max_clusters(nq->n, nrows);
    void* tmp = nullptr;
    size_t const rb = sizeof(uint32_t) * (size_t(nrows) + 1);
    if (int rc = fa::scratch(13, 2 * rb + sizeof(uint4) * max_clusters + 256, &tmp))
        return rc;
    RowsBuffers B{};
    B.nclus = static_cast<uint32_t*>(tmp);
    B.cl_off = B.nclus + (nrows + 1);
    B.table = reinterpret_cast<uint4*>(static_cast<char*>(tmp) + ((2 * rb + 15) & ~size_t(15)));
    return launch_rows(nq->cur->dbox, nq->cur->dev(), dims, rv.R, F, nq->n, B, a.qmax_sq, a.qmin_sq, a.T, a.bins,
                       a.r_min, a.inv_width, d_counts);
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
constexpr uint32_t kChunk = 2;

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
    uint32_t* wh = reinterpret_cast<uint32_t*>(smem + lds_table_floats(nbins)) + wave * nbins;
    for (uint32_t t = lane; t < nbins; t += 64)
        wh[t] = 0u;
    table_prologue(sT, T, nbins);
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
            rdf_rows_wave<R>(F.box, g, table[I], F.rx, q_rmax_sq, 0.0f, LutBins(sT, nbins, vmin, inv_width, wh));
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
int rdf_row_variant_count()
{
    return kRowVariants;
}

// Batched self RDF with cells at least r_max / R wide (R in 2..4): counts has
// n_frames x bins entries.  Bins beyond 2048 are not supported here (each wave
// keeps a private LDS histogram).
int batch_rdf(const BatchGrid& G, int R, float r_max, uint32_t bins, const BatchRdfScratch& S, uint32_t* d_counts)
{
    hipStream_t s = stream();
    if (bins == 0 || bins > 2048)
        return fail(FA_ERR_INVALID, "batched RDF supports 1..2048 bins");
    if (uint64_t(G.total_points) + kCposPad > (1ull << 28))
        return fail(FA_ERR_INVALID, "batched RDF supports up to 2^28 - 32 points per batch");
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
    size_t const shmem = sizeof(float) * lds_table_floats(bins) + sizeof(uint32_t) * bins * (kBlock / 64);
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
    std::vector<float> const tab = build_table(bins, 0.0f, r_max);
    FA_HIP(hipMemcpy(d_T, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice));
    return FA_OK;
}

size_t rdf_table_floats(uint32_t bins)
{
    return table_floats(bins);
}

} // namespace fa

// ------------------------------------------------------------------ no-allocation self RDF
// fa_rdf_accumulate_self_noalloc keeps every device buffer it needs (bin
// table, cell grid, cluster table, scan storage) in the statically allocated
// array fa_noalloc_scratch: each GPU the library runs on holds its own copy,
// from module load to the end of the process.  Host code never passes the
// array's name to a kernel (in host code it names a host-side stand-in); it
// uses the device address that noalloc_setup looks up once per device.
namespace {
constexpr size_t kNoAllocBytes = size_t(64) << 20;
constexpr uint32_t kNoAllocMaxPoints = 1u << 20;
constexpr size_t kNoAllocTableBytes = size_t(32) << 10; // table_floats(kMaxLdsBins) floats
static_assert(sizeof(float) * table_floats(kMaxLdsBins) <= kNoAllocTableBytes, "bin table region too small");
constexpr int kNoAllocMaxDevices = 64;
} // namespace

// External linkage on purpose: this library is built without relocatable
// device code, and a `static` device variable then has a local symbol in the
// code object, which the HIP runtime cannot resolve by name (the lookup
// aborts).  `used` keeps the array in the code object although no kernel
// names it (the kernels receive its address as an argument).
__device__ __attribute__((used)) __align__(256) unsigned char fa_noalloc_scratch[kNoAllocBytes];

namespace {
struct NoAllocDevice
{
    char* base = nullptr; // device address of fa_noalloc_scratch on this device
    bool table_valid = false;
    uint32_t bins = 0;
    float r_min = 0.0f, r_max = 0.0f;
};
NoAllocDevice g_noalloc[kNoAllocMaxDevices];

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
} // namespace

namespace fa {
int noalloc_setup()
{
    int dev = 0;
    FA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return fail(FA_ERR_RUNTIME, "device index beyond the no-allocation scratch table");
    void* ptr = nullptr;
    // qualified: inside namespace fa the unqualified name would find other declarations first
    hipError_t const e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(::fa_noalloc_scratch));
    if (e != hipSuccess || !ptr)
        return fail(FA_ERR_DEVICE, std::string("hipGetSymbolAddress(fa_noalloc_scratch): ") + hipGetErrorString(e));
    g_noalloc[dev].base = static_cast<char*>(ptr);
    g_noalloc[dev].table_valid = false;
    return FA_OK;
}
} // namespace fa

extern "C" {

int fa_rdf_bin_table(uint32_t bins, float r_min, float r_max, float* out, uint64_t cap)
{
    FA_API_LOCK;
    if (bins == 0 || !(r_max > r_min))
        return fa::fail(FA_ERR_INVALID, "bin table needs bins > 0 and r_max > r_min");
    std::vector<float> const tab = build_table(bins, r_min, r_max);
    if (out)
    {
        if (cap < tab.size())
            return fa::fail(FA_ERR_INVALID, "output too small for the bin table");
        std::memcpy(out, tab.data(), sizeof(float) * tab.size());
    }
    return int(tab.size());
}

int fa_rdf_accumulate_query(fa_nq* nq, const float* d_query_points, uint32_t n_query, float q_r_max, float q_r_min,
                            int exclude_ii, uint32_t bins, float r_min, float r_max, uint32_t* d_counts)
{
    FA_API_LOCK;
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
    size_t const shmem = lds ? sizeof(float) * lds_table_floats(bins) + sizeof(uint32_t) * bins : 0;
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
                if (int rc = run_rows(ra, kRowVariant[v], trial, &ok))
                    return rc;
                if (!ok)
                    break;
                if (int rc = fa::time_launch([&] { return run_rows(ra, kRowVariant[v], trial, &ok); }, &ms))
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
        if (int rc = run_rows(ra, kRowVariant[std::max(variant, 0) % kRowVariants], d_counts, &handled))
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
    fa::CQGrid const cq = nq->kind == FA_NQ_CELLQUERY ? fa::cq_grid_for(nq->box, q_r_max) : fa::CQGrid{};
    if (self)
    {
        if (nq->kind == FA_NQ_AABB)
            FA_RDF_LAUNCH(rdf_self_kernel, AABB_FORMULA, db, g, n_query, qmax_sq, qmin_sq, exclude_ii, R, T, bins,
                          r_min, inv_width, d_counts, cq);
        else if (nq->kind == FA_NQ_CELLQUERY)
            FA_RDF_LAUNCH(rdf_self_kernel, CELLQUERY_FORMULA, db, g, n_query, qmax_sq, qmin_sq, exclude_ii, R, T,
                          bins, r_min, inv_width, d_counts, cq);
        else
            FA_RDF_LAUNCH(rdf_self_kernel, LINKCELL_FORMULA, db, g, n_query, qmax_sq, qmin_sq, exclude_ii, R, T, bins,
                          r_min, inv_width, d_counts, cq);
    }
    else
    {
        if (nq->kind == FA_NQ_AABB)
            FA_RDF_LAUNCH(rdf_query_kernel, AABB_FORMULA, db, g, d_query_points, n_query, qmax_sq, qmin_sq,
                          exclude_ii, R, T, bins, r_min, inv_width, d_counts, cq);
        else if (nq->kind == FA_NQ_CELLQUERY)
            FA_RDF_LAUNCH(rdf_query_kernel, CELLQUERY_FORMULA, db, g, d_query_points, n_query, qmax_sq, qmin_sq,
                          exclude_ii, R, T, bins, r_min, inv_width, d_counts, cq);
        else
            FA_RDF_LAUNCH(rdf_query_kernel, LINKCELL_FORMULA, db, g, d_query_points, n_query, qmax_sq, qmin_sq,
                          exclude_ii, R, T, bins, r_min, inv_width, d_counts, cq);
    }
#undef FA_RDF_LAUNCH
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int fa_rdf_accumulate_self_noalloc(const fa_box* box, const float* d_points, uint32_t n, float q_r_max,
                                   float q_r_min, uint32_t bins, float r_min, float r_max, uint32_t* d_counts)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    int dev = 0;
    FA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices || !g_noalloc[dev].base)
        return fa::fail(FA_ERR_RUNTIME, "the no-allocation scratch is not set up on the current device");
    NoAllocDevice& D = g_noalloc[dev];
    if (!box || !d_counts || !d_points)
        return fa::fail(FA_ERR_INVALID, "box, points and counts must not be null");
    // NeighborQuery.h:97-100 and the AABBQuery ball-query checks
    if (n == 0)
        return fa::fail(FA_ERR_INVALID, "Cannot create a NeighborQuery with 0 particles.");
    if (n > kNoAllocMaxPoints)
        return fa::fail(FA_ERR_INVALID, "fa_rdf_accumulate_self_noalloc takes at most 1048576 points");
    if (bins == 0 || bins > kMaxLdsBins || !(r_max > r_min))
        return fa::fail(FA_ERR_INVALID, "fa_rdf_accumulate_self_noalloc needs 1 <= bins <= 4096 and r_max > r_min");
    fa_nq shape{};
    shape.box = *box;
    shape.d_points = d_points;
    shape.n = n;
    shape.kind = FA_NQ_AABB;
    fa_query_args a{FA_MODE_BALL, 0xffffffffu, q_r_max, q_r_min, -1.0f, -1.0f, 1};
    if (int rc = fa::validate_query_args(&shape, a, n))
        return rc;

    // launch shape: the tuner's recorded cell shape for this problem class, else (R, F) = (2, 8)
    double const density = double(n) / double(fa::box_volume(*box));
    int variant = fa::tuned_variant(fa::TUNE_RDF_SELF, fa::tune_key(density, q_r_max));
    if (variant < 0)
        variant = 6;
    RowVariant const rv = kRowVariant[variant % kRowVariants];
    int dims[3], F = 1;
    if (!rows_grid_dims(*box, n, q_r_max, rv, dims, &F))
        return fa::fail(FA_ERR_INVALID, "cell grid rows too long for fa_rdf_accumulate_self_noalloc");
    uint64_t const ncell = uint64_t(dims[0]) * uint64_t(dims[1]) * uint64_t(dims[2]);
    uint32_t const nrows = uint32_t(dims[1]) * uint32_t(dims[2]);
    uint64_t const max_clusters = rows_max_clusters(n, nrows);

    // carve the scratch: table | outside flag | cpos | wrapc | cell_of_point |
    // rank_of_point | wrap_of_point | cstart | counts | nclus | cl_off | cluster table | scan storage
    size_t const scan_bytes = std::max(fa::scan_tmp_bytes(uint32_t(ncell)), fa::scan_tmp_bytes(nrows));
    if (scan_bytes == 0)
        return fa::fail(FA_ERR_DEVICE, "scan storage query failed");
    size_t off[13];
    size_t const sizes[13] = {kNoAllocTableBytes,
                              256,
                              sizeof(float4) * (size_t(n) + fa::kCposPad),
                              sizeof(int) * size_t(n),
                              sizeof(uint32_t) * size_t(n),
                              sizeof(uint32_t) * size_t(n),
                              sizeof(int) * size_t(n),
                              sizeof(uint32_t) * size_t(ncell + 1),
                              sizeof(uint32_t) * size_t(ncell),
                              sizeof(uint32_t) * (size_t(nrows) + 1),
                              sizeof(uint32_t) * (size_t(nrows) + 1),
                              sizeof(uint4) * size_t(max_clusters),
                              scan_bytes};
    size_t total = 0;
    for (int k = 0; k < 13; ++k)
    {
        off[k] = total;
        total += align256(sizes[k]);
    }
    if (total > kNoAllocBytes)
        return fa::fail(FA_ERR_INVALID, "fa_rdf_accumulate_self_noalloc: the cell grid for these points and r_max "
                                        "needs " + std::to_string(total) + " bytes of scratch, more than its " +
                                            std::to_string(kNoAllocBytes));
    char* const base = D.base;
    hipStream_t s = fa::stream();

    float* T = reinterpret_cast<float*>(base + off[0]);
    if (!D.table_valid || D.bins != bins || D.r_min != r_min || D.r_max != r_max)
    {
        std::vector<float> const tab = build_table(bins, r_min, r_max);
        FA_HIP(hipMemcpyAsync(T, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice, s));
        D.table_valid = true;
        D.bins = bins;
        D.r_min = r_min;
        D.r_max = r_max;
    }
    fa::GridBuffers G{};
    G.outside = reinterpret_cast<int*>(base + off[1]);
    G.cpos = reinterpret_cast<float4*>(base + off[2]);
    G.wrapc = reinterpret_cast<int*>(base + off[3]);
    G.cell_of_point = reinterpret_cast<uint32_t*>(base + off[4]);
    G.rank_of_point = reinterpret_cast<uint32_t*>(base + off[5]);
    G.wrap_of_point = reinterpret_cast<int*>(base + off[6]);
    G.cstart = reinterpret_cast<uint32_t*>(base + off[7]);
    G.counts = reinterpret_cast<uint32_t*>(base + off[8]);
    G.scan_tmp = base + off[12];
    G.scan_bytes = scan_bytes;
    DevBox const db = fa::make_devbox(*box);
    if (int rc = fa::build_grid_into(G, db, d_points, n, dims))
        return rc;
    DevGrid g{};
    g.cpos = G.cpos;
    g.wrapc = G.wrapc;
    g.cstart = G.cstart;
    g.outside = G.outside;
    float npd[3];
    fa::nearest_plane_distance(*box, npd);
    for (int k = 0; k < 3; ++k)
    {
        g.dim[k] = dims[k];
        g.cw[k] = (k == 2 && box->is2D) ? 0.0f : npd[k] / float(dims[k]);
    }
    RowsBuffers B{};
    B.nclus = reinterpret_cast<uint32_t*>(base + off[9]);
    B.cl_off = reinterpret_cast<uint32_t*>(base + off[10]);
    B.table = reinterpret_cast<uint4*>(base + off[11]);
    B.scan_tmp = base + off[12];
    B.scan_bytes = scan_bytes;
    float const width = (r_max - r_min) / float(bins);
    float const inv_width = 1.0f / width;
    return launch_rows(db, g, dims, rv.R, F, n, B, q_r_max * q_r_max, q_r_min * q_r_min, T, bins, r_min, inv_width,
                       d_counts);
}

int fa_rdf_accumulate_nlist(const fa_nlist* nl, uint32_t bins, float r_min, float r_max, uint32_t* d_counts)
{
    FA_API_LOCK;
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
    FA_API_LOCK;
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
