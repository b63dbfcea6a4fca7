// fa_query.hip -- ball and k-nearest-neighbour queries producing freud's
// NeighborList (freud/locality/NeighborQuery.h:434-481 toNeighborList,
// AABBQuery.cc:77-281, LinkCell.cc:496-679).
//
// Ball query: a count pass (one thread per query point), an exclusive scan of
// the counts into segment offsets, and a fill pass.  The fill pass stages each
// block's keys (j << 5 | image) in LDS, sorts every query point's segment by
// point index (the reference's (i, j, w, d) order, NeighborBond.h:80-95) and
// then writes the block's contiguous bond range cooperatively, one bond per
// lane, so the NeighborList stores are coalesced.
//
// kNN: one thread per query point walks Chebyshev rings of cells, keeping the
// k best (d, j) keys in registers, and stops once the k-th distance is inside
// the completely searched radius.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "fa_common.h"

namespace {

using fa::AABB_FORMULA;
using fa::CELLQUERY_FORMULA;
using fa::DevBox;
using fa::DevGrid;
using fa::LINKCELL_FORMULA;

constexpr int kBlock = 256;
constexpr int kKeyBytes = 32768; // largest LDS key staging per block
// keys per point that the one-walk ball query stages in LDS (32 KB per block)
constexpr uint32_t kWalkMaxSlots = kKeyBytes / (kBlock * sizeof(uint32_t));
extern __shared__ __align__(16) unsigned char fa_dyn_lds[];

// Dynamic LDS bytes for a fill block's keys: 1.3 x the mean bonds of kBlock
// query points, between 4 KB and kKeyBytes.
size_t fill_key_bytes(uint64_t bonds, uint64_t points, size_t key_size)
{
    double const per = points ? double(bonds) / double(points) : 0.0;
    double const want = std::ceil(per * kBlock * 1.3) * double(key_size);
    size_t const b = size_t(std::min(std::max(want, 4096.0), double(kKeyBytes)));
    return (b + 15) & ~size_t(15);
}

__device__ inline uint32_t float_order_bits(float v)
{
    uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Query point at processing position p: in SELF mode (query points are the
// grid's own points) positions run in cell order through the grid's sorted
// copy, so neighbouring lanes share stencil cells; otherwise in caller order.
template<bool SELF>
__device__ inline uint32_t query_at(const DevGrid& g, const float* __restrict__ qpts, uint32_t p, float& qx, float& qy,
                                    float& qz)
{
    if (SELF)
    {
        float4 const q = g.cpos[p];
        qx = q.x;
        qy = q.y;
        qz = q.z;
        return __float_as_uint(q.w);
    }
    qx = qpts[3 * p];
    qy = qpts[3 * p + 1];
    qz = qpts[3 * p + 2];
    return p;
}

// exclusive prefix sum over a block of kBlock threads; returns the total
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp, uint32_t& excl)
{
    uint32_t const tid = threadIdx.x;
    s_tmp[tid] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1)
    {
        uint32_t const add = (tid >= uint32_t(off)) ? s_tmp[tid - off] : 0u;
        __syncthreads();
        s_tmp[tid] += add;
        __syncthreads();
    }
    uint32_t const incl = s_tmp[tid];
    uint32_t const total = s_tmp[kBlock - 1];
    __syncthreads();
    excl = incl - v;
    return total;
}

// Count of query position p (p < pend).  In SELF mode positions index the
// grid's cell-ordered points and the point index comes from the grid.
template<int FORMULA, bool SELF>
__device__ inline void ball_count_body(const DevBox& b, const DevGrid& g, const float* __restrict__ qpts, uint32_t p,
                                       float r_max_sq, float r_min_sq, int exclude_ii, int R,
                                       uint32_t* __restrict__ counts, const fa::CQGrid& cq = fa::CQGrid{})
{
    float qx, qy, qz;
    uint32_t const i = query_at<SELF>(g, qpts, p, qx, qy, qz);
    uint32_t c = 0;
    fa::visit_bonds<FORMULA>(b, g, cq, qx, qy, qz, R,
                                    [&](uint32_t j, float, float, float, float r_sq, int) {
                                        // range first: the index is needed for in-range candidates only
                                        if (r_sq < r_max_sq && r_sq >= r_min_sq && !(exclude_ii && j == i))
                                            ++c;
                                    });
    counts[i] = c;
}

template<typename K> __device__ inline void insertion_sort(K* a, uint32_t n)
{
    for (uint32_t x = 1; x < n; ++x)
    {
        K const v = a[x];
        uint32_t y = x;
        while (y > 0 && a[y - 1] > v)
        {
            a[y] = a[y - 1];
            --y;
        }
        a[y] = v;
    }
}

// Writes the bond of query point i (at qx, qy, qz) given its key, at bond
// index bidx.
// NT: non-temporal stores, for a block's contiguous range of the
// NeighborList (full lines; the output then does not push the gathered
// positions out of L2).  Scattered segments keep ordinary stores, whose
// partial lines merge in L2.
template<int FORMULA, bool BYDIST, bool NT = false>
__device__ inline void write_bond_at(const DevBox& b, float qx, float qy, float qz, const float* __restrict__ pts,
                                     uint32_t i, uint64_t key, uint64_t bidx, uint32_t* __restrict__ nb_out,
                                     float* __restrict__ d_out, float* __restrict__ w_out, float* __restrict__ v_out)
{
    uint32_t const jc = uint32_t(key);
    uint32_t const j = jc >> 5;
    int const code = int(jc & 31u);
    float rx, ry, rz;
    fa::bond_vector<FORMULA>(b, qx, qy, qz, pts[3 * j], pts[3 * j + 1], pts[3 * j + 2], code, rx, ry, rz);
    float const r_sq = (rx * rx + ry * ry) + rz * rz;
    if (NT)
    {
        __builtin_nontemporal_store((uint64_t(j) << 32) | uint64_t(i), reinterpret_cast<uint64_t*>(nb_out) + bidx);
        __builtin_nontemporal_store(sqrtf(r_sq), d_out + bidx);
        __builtin_nontemporal_store(1.0f, w_out + bidx);
        __builtin_nontemporal_store(rx, v_out + 3 * bidx);
        __builtin_nontemporal_store(ry, v_out + 3 * bidx + 1);
        __builtin_nontemporal_store(rz, v_out + 3 * bidx + 2);
        return;
    }
    reinterpret_cast<uint2*>(nb_out)[bidx] = make_uint2(i, j);
    d_out[bidx] = sqrtf(r_sq);
    w_out[bidx] = 1.0f;
    v_out[3 * bidx] = rx;
    v_out[3 * bidx + 1] = ry;
    v_out[3 * bidx + 2] = rz;
}

template<int FORMULA, bool BYDIST>
__device__ inline void write_bond(const DevBox& b, const float* __restrict__ qpts, const float* __restrict__ pts,
                                  uint32_t i, uint64_t key, uint64_t bidx, uint32_t* __restrict__ nb_out,
                                  float* __restrict__ d_out, float* __restrict__ w_out, float* __restrict__ v_out)
{
    write_bond_at<FORMULA, BYDIST>(b, qpts[3 * i], qpts[3 * i + 1], qpts[3 * i + 2], pts, i, key, bidx, nb_out, d_out,
                                   w_out, v_out);
}

// The point lo (< nloc) of a block whose key segment [s_off[lo], s_off[lo + 1])
// holds element t < s_off[nloc].
__device__ inline uint32_t segment_of(const uint32_t* s_off, uint32_t nloc, uint32_t t)
{
    uint32_t lo = 0, hi = nloc;
    while (hi - lo > 1)
    {
        uint32_t const mid = (lo + hi) >> 1;
        if (s_off[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Block-cooperative store of a block's bonds: keys[t] for t < total belong to
// the block's query points in processing order (exclusive offsets s_off,
// point indices s_i, positions s_q); bond t of point lo goes to
// offsets[s_i[lo]] + t - s_off[lo].
template<int FORMULA, bool BYDIST, typename K, bool NT = false>
__device__ inline void store_block_bonds(const DevBox& b, const float* s_q, const float* __restrict__ pts,
                                         const K* keys, uint32_t total, uint32_t nloc, const uint32_t* s_off,
                                         const uint32_t* s_i, const uint32_t* __restrict__ offsets,
                                         uint32_t* __restrict__ nb_out, float* __restrict__ d_out,
                                         float* __restrict__ w_out, float* __restrict__ v_out)
{
    for (uint32_t t = threadIdx.x; t < total; t += kBlock)
    {
        uint32_t const lo = segment_of(s_off, nloc, t);
        uint32_t const i = s_i[lo];
        write_bond_at<FORMULA, BYDIST, NT>(b, s_q[3 * lo], s_q[3 * lo + 1], s_q[3 * lo + 2], pts, i, uint64_t(keys[t]),
                                       uint64_t(offsets[i]) + (t - s_off[lo]), nb_out, d_out, w_out, v_out);
    }
}

// One block's share of the fill pass: query positions [p0, p0 + nloc).  Point
// indices i are relative to `offsets`, `qpts` and `pts` (a frame of a batch
// passes pointers advanced to its first point); bond indices are absolute.
template<int FORMULA, bool BYDIST, bool SELF>
__device__ inline void ball_fill_body(const DevBox& b, const DevGrid& g, const float* __restrict__ qpts, uint32_t p0,
                                      uint32_t nloc, const float* __restrict__ pts, float r_max_sq, float r_min_sq,
                                      int exclude_ii, int R, const uint32_t* __restrict__ offsets,
                                      uint32_t* __restrict__ nb_out, float* __restrict__ d_out,
                                      float* __restrict__ w_out, float* __restrict__ v_out,
                                      void* __restrict__ gkeys_raw, uint32_t key_cap,
                                      const fa::CQGrid& cq = fa::CQGrid{})
{
    using K = typename std::conditional<BYDIST, uint64_t, uint32_t>::type;
    // key staging: dynamic LDS sized by the host from the bonds per point, so
    // the kernel keeps its occupancy (blocks over the capacity use gkeys_raw)
    K* s_keys = reinterpret_cast<K*>(fa_dyn_lds);
    uint32_t const CAP = key_cap;
    __shared__ uint32_t s_off[kBlock + 1];
    __shared__ uint32_t s_i[kBlock];
    __shared__ uint32_t s_tmp[kBlock];
    __shared__ float s_q[3 * kBlock];
    uint32_t const tid = threadIdx.x;
    float qx = 0.0f, qy = 0.0f, qz = 0.0f;
    uint32_t i = 0, cnt = 0;
    if (tid < nloc)
    {
        i = query_at<SELF>(g, qpts, p0 + tid, qx, qy, qz);
        cnt = offsets[i + 1] - offsets[i];
        s_i[tid] = i;
        s_q[3 * tid] = qx;
        s_q[3 * tid + 1] = qy;
        s_q[3 * tid + 2] = qz;
    }
    uint32_t excl;
    uint32_t const total = block_exclusive_scan(cnt, s_tmp, excl);
    if (tid < nloc)
        s_off[tid] = excl;
    if (tid == 0)
        s_off[nloc] = total;
    __syncthreads();
    bool const in_lds = total <= CAP;
    // overflow: keys go to the point's own segment of a global scratch array
    K* keys = in_lds ? s_keys : reinterpret_cast<K*>(gkeys_raw);
    if (tid < nloc)
    {
        uint32_t const beg = in_lds ? excl : offsets[i];
        uint32_t const end = beg + cnt;
        uint32_t k = beg;
        fa::visit_bonds<FORMULA>(b, g, cq, qx, qy, qz, R,
                                        [&](uint32_t j, float, float, float, float r_sq, int code) {
                                            if (r_sq < r_max_sq && r_sq >= r_min_sq && k < end &&
                                                !(exclude_ii && j == i))
                                            {
                                                uint32_t const jc = (j << 5) | uint32_t(code);
                                                if (BYDIST)
                                                    keys[k] = K((uint64_t(float_order_bits(sqrtf(r_sq))) << 32) | jc);
                                                else
                                                    keys[k] = K(jc);
                                                ++k;
                                            }
                                        });
        insertion_sort(keys + beg, k - beg);
        if (!in_lds)
            for (uint32_t m = 0; m < k - beg; ++m)
                write_bond<FORMULA, BYDIST>(b, qpts, pts, i, uint64_t(keys[beg + m]), uint64_t(offsets[i]) + m,
                                            nb_out, d_out, w_out, v_out);
    }
    __syncthreads();
    if (in_lds)
        store_block_bonds<FORMULA, BYDIST>(b, s_q, pts, keys, total, nloc, s_off, s_i, offsets, nb_out, d_out,
                                           w_out, v_out);
}

// ------------------------------------------------------------------ ball query: one walk
// The ball query walks the cell stencil once.  The walk runs in the order that
// keeps a wave's stencils together (cell order for self queries) and keeps
// each query point's first C keys (j << 5 | image code, unsorted) in LDS; the
// block then packs its points' keys into its own region of a scratch array
// with coalesced stores, and records each point's bond count and where its
// keys start.  The emission pass runs in index order: a block's 256 points own
// one contiguous range of the NeighborList, so its 28-byte bonds go out with
// full-line stores (written in cell order they would land as ~13-bond
// segments scattered over partly written lines).  A point with more than C
// bonds walks its stencil again inside the emission pass.
// One block's share of the walk: processing positions [p0, p0 + nloc), keys
// packed into rows[region ..] (region: the block's kBlock * C slots); counts
// and pos are indexed by point index (a frame of a batch passes them advanced
// to its first point; pos values are absolute slots of rows).
template<int FORMULA, bool SELF>
__device__ inline void ball_walk_body(const DevBox& b, const DevGrid& g, const float* __restrict__ qpts, uint32_t p0,
                                      uint32_t nloc, uint32_t region, float r_max_sq, float r_min_sq,
                                      int exclude_ii, int R, const fa::CQGrid& cq, uint32_t C,
                                      uint32_t* __restrict__ counts, uint32_t* __restrict__ pos,
                                      uint32_t* __restrict__ rows)
{
    // s_rows[m * kBlock + tid]: key m of thread tid (consecutive lanes, consecutive banks)
    uint32_t* s_rows = reinterpret_cast<uint32_t*>(fa_dyn_lds);
    __shared__ uint32_t s_off[kBlock + 1];
    __shared__ uint32_t s_tmp[kBlock];
    uint32_t const tid = threadIdx.x;
    uint32_t c = 0, i = 0;
    if (tid < nloc)
    {
        float qx, qy, qz;
        i = query_at<SELF>(g, qpts, p0 + tid, qx, qy, qz);
        fa::visit_bonds<FORMULA>(b, g, cq, qx, qy, qz, R,
                                 [&](uint32_t j, float, float, float, float r_sq, int code) {
                                     if (r_sq < r_max_sq && r_sq >= r_min_sq && !(exclude_ii && j == i))
                                     {
                                         if (c < C)
                                             s_rows[c * kBlock + tid] = (j << 5) | uint32_t(code);
                                         ++c;
                                     }
                                 });
    }
    uint32_t const kept = min(c, C);
    uint32_t excl;
    uint32_t const total = block_exclusive_scan(kept, s_tmp, excl);
    if (tid < nloc)
    {
        s_off[tid] = excl;
        counts[i] = c;
        pos[i] = region + excl;
    }
    if (tid == 0)
        s_off[nloc] = total;
    __syncthreads();
    for (uint32_t t = tid; t < total; t += kBlock)
    {
        uint32_t lo = 0, hi = nloc;
        while (hi - lo > 1)
        {
            uint32_t const mid = (lo + hi) >> 1;
            if (s_off[mid] <= t)
                lo = mid;
            else
                hi = mid;
        }
        rows[region + t] = s_rows[(t - s_off[lo]) * kBlock + lo];
    }
}

template<int FORMULA, bool SELF>
__global__ __launch_bounds__(kBlock) void ball_walk_kernel(DevBox b, DevGrid g, const float* __restrict__ qpts,
                                                           uint32_t nq, float r_max_sq, float r_min_sq,
                                                           int exclude_ii, int R, fa::CQGrid cq, uint32_t C,
                                                           uint32_t* __restrict__ counts, uint32_t* __restrict__ pos,
                                                           uint32_t* __restrict__ rows)
{
    uint32_t const tile = fa::xcd_tile(blockIdx.x, gridDim.x);
    uint32_t const p0 = tile * kBlock;
    ball_walk_body<FORMULA, SELF>(b, g, qpts, p0, min(uint32_t(kBlock), nq - p0), tile * kBlock * C, r_max_sq,
                                  r_min_sq, exclude_ii, R, cq, C, counts, pos, rows);
}

// Emission: with INDEX, query points [p0, p0 + 256) of the caller's order,
// whose bonds are one contiguous range; otherwise the walk's processing order
// (the block's keys are its own packed region, and its neighbours' positions
// stay in cache, which matters more for points stored in random order).
// Keys come from the point's row (or, past C bonds, from a second walk), are
// sorted per point (by j and image, or by distance), and the block writes its
// bonds.
// One block's share of the emission: query positions [p0, p0 + nloc) (point
// indices with INDEX, else processing positions).  Point indices are relative
// to qpts, pts, pos and offsets (a frame of a batch passes them advanced to its
// first point); bond indices and rows slots are absolute.
template<int FORMULA, bool BYDIST, bool SELF, bool INDEX>
__device__ inline void ball_emit_body(const DevBox& b, const DevGrid& g, const float* __restrict__ qpts, uint32_t p0,
                                      uint32_t nloc, const float* __restrict__ pts, float r_max_sq, float r_min_sq,
                                      int exclude_ii, int R, const fa::CQGrid& cq, uint32_t C,
                                      const uint32_t* __restrict__ pos, const uint32_t* __restrict__ rows,
                                      const uint32_t* __restrict__ offsets, uint32_t* __restrict__ nb_out,
                                      float* __restrict__ d_out, float* __restrict__ w_out,
                                      float* __restrict__ v_out, void* __restrict__ gkeys_raw, uint32_t key_cap)
{
    using K = typename std::conditional<BYDIST, uint64_t, uint32_t>::type;
    K* s_keys = reinterpret_cast<K*>(fa_dyn_lds);
    __shared__ uint32_t s_off[kBlock + 1];
    __shared__ uint32_t s_i[kBlock];
    __shared__ uint32_t s_tmp[kBlock];
    __shared__ uint32_t s_row[kBlock];
    __shared__ float s_q[3 * kBlock];
    uint32_t const tid = threadIdx.x;
    uint32_t i = p0 + tid;
    float qx = 0.0f, qy = 0.0f, qz = 0.0f;
    uint32_t cnt = 0;
    if (tid < nloc)
    {
        if (INDEX)
        {
            qx = qpts[3 * i];
            qy = qpts[3 * i + 1];
            qz = qpts[3 * i + 2];
        }
        else
            i = query_at<SELF>(g, qpts, p0 + tid, qx, qy, qz);
        cnt = offsets[i + 1] - offsets[i];
        s_i[tid] = i;
        s_q[3 * tid] = qx;
        s_q[3 * tid + 1] = qy;
        s_q[3 * tid + 2] = qz;
    }
    uint32_t excl;
    uint32_t const total = block_exclusive_scan(cnt, s_tmp, excl);
    if (tid < nloc)
        s_off[tid] = excl;
    if (tid == 0)
        s_off[nloc] = total;
    __syncthreads();
    bool const in_lds = total <= key_cap;
    K* keys = in_lds ? s_keys : reinterpret_cast<K*>(gkeys_raw);
    auto key_of = [&](uint32_t jc, float x, float y, float z) -> K {
        if (BYDIST)
        {
            uint32_t const j = jc >> 5;
            float rx, ry, rz;
            fa::bond_vector<FORMULA>(b, x, y, z, pts[3 * j], pts[3 * j + 1], pts[3 * j + 2], int(jc & 31u), rx, ry,
                                     rz);
            float const r_sq = (rx * rx + ry * ry) + rz * rz;
            return K((uint64_t(float_order_bits(sqrtf(r_sq))) << 32) | jc);
        }
        return K(jc);
    };
    bool const coop = in_lds && C > 0;
    if (coop)
    {
        // The staged key rows into LDS, element t by thread t: one load
        // instruction covers whole rows (a lane reading its own row key by key
        // lets the rows of a CU's blocks fall out of L2 between its loads).
        // Points past C bonds (no row) are walked again below.
        if (tid < nloc)
            s_row[tid] = cnt <= C ? pos[i] : ~0u;
        __syncthreads();
        for (uint32_t t = tid; t < total; t += kBlock)
        {
            uint32_t const lo = segment_of(s_off, nloc, t);
            uint32_t const r0 = s_row[lo];
            if (r0 != ~0u)
                s_keys[t] = key_of(rows[r0 + (t - s_off[lo])], s_q[3 * lo], s_q[3 * lo + 1], s_q[3 * lo + 2]);
        }
        __syncthreads();
    }
    if (tid < nloc)
    {
        uint32_t const beg = in_lds ? excl : offsets[i];
        if (cnt <= C)
        {
            if (!coop)
            {
                const uint32_t* row = rows + pos[i];
                for (uint32_t m = 0; m < cnt; ++m)
                    keys[beg + m] = key_of(row[m], qx, qy, qz);
            }
        }
        else
        {
            uint32_t k = beg;
            uint32_t const end = beg + cnt;
            fa::visit_bonds<FORMULA>(b, g, cq, qx, qy, qz, R,
                                     [&](uint32_t j, float, float, float, float r_sq, int code) {
                                         if (r_sq < r_max_sq && r_sq >= r_min_sq && k < end && !(exclude_ii && j == i))
                                         {
                                             uint32_t const jc = (j << 5) | uint32_t(code);
                                             if (BYDIST)
                                                 keys[k] = K((uint64_t(float_order_bits(sqrtf(r_sq))) << 32) | jc);
                                             else
                                                 keys[k] = K(jc);
                                             ++k;
                                         }
                                     });
        }
        insertion_sort(keys + beg, cnt);
        if (!in_lds)
            for (uint32_t m = 0; m < cnt; ++m)
                write_bond_at<FORMULA, BYDIST>(b, qx, qy, qz, pts, i, uint64_t(keys[beg + m]),
                                               uint64_t(offsets[i]) + m, nb_out, d_out, w_out, v_out);
    }
    __syncthreads();
    if (in_lds)
        store_block_bonds<FORMULA, BYDIST, K, INDEX>(b, s_q, pts, keys, total, nloc, s_off, s_i, offsets, nb_out,
                                                     d_out, w_out, v_out);
}

template<int FORMULA, bool BYDIST, bool SELF, bool INDEX>
__global__ __launch_bounds__(kBlock) void ball_emit_kernel(DevBox b, DevGrid g, const float* __restrict__ qpts,
                                                           uint32_t nq, const float* __restrict__ pts,
                                                           float r_max_sq, float r_min_sq, int exclude_ii, int R,
                                                           fa::CQGrid cq, uint32_t C, const uint32_t* __restrict__ pos,
                                                           const uint32_t* __restrict__ rows,
                                                           const uint32_t* __restrict__ offsets,
                                                           uint32_t* __restrict__ nb_out, float* __restrict__ d_out,
                                                           float* __restrict__ w_out, float* __restrict__ v_out,
                                                           void* __restrict__ gkeys_raw, uint32_t key_cap)
{
    uint32_t const p0 = fa::xcd_tile(blockIdx.x, gridDim.x) * kBlock;
    ball_emit_body<FORMULA, BYDIST, SELF, INDEX>(b, g, qpts, p0, min(uint32_t(kBlock), nq - p0), pts, r_max_sq,
                                                 r_min_sq, exclude_ii, R, cq, C, pos, rows, offsets, nb_out, d_out,
                                                 w_out, v_out, gkeys_raw, key_cap);
}

// Segments with the reference's convention (0 for query points without bonds).
__global__ void counts_segments_kernel(const uint32_t* __restrict__ offsets, uint32_t nq, uint32_t* __restrict__ counts,
                                       uint32_t* __restrict__ segments)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq)
        return;
    uint32_t const c = offsets[i + 1] - offsets[i];
    counts[i] = c;
    segments[i] = c ? offsets[i] : 0u;
}

// ------------------------------------------------------------------ kNN
template<int KMAX> struct TopK
{
    uint64_t v[KMAX];
    __device__ void init()
    {
#pragma unroll
        for (int m = 0; m < KMAX; ++m)
            v[m] = ~0ull;
    }
    // keep the KMAX smallest keys in ascending order
    __device__ void insert(uint64_t key)
    {
        if (key >= v[KMAX - 1])
            return;
#pragma unroll
        for (int m = KMAX - 1; m > 0; --m)
            v[m] = min(v[m], max(v[m - 1], key));
        v[0] = min(v[0], key);
    }
    __device__ uint64_t at(uint32_t idx) const
    {
        uint64_t r = ~0ull;
#pragma unroll
        for (int m = 0; m < KMAX; ++m)
            if (uint32_t(m) == idx)
                r = v[m];
        return r;
    }
};

// kNN selection keys order candidates by (d, j, image) with d = sqrtf(r_sq),
// the reference's distance (AABBQuery.cc:227,246 and LinkCell.cc:656 sort the
// bonds by their float distance; equal distances go to the lower index here).
// Two different r_sq can share one d, so the key holds the bits of d itself
// (monotone as an unsigned integer for d >= 0), not those of r_sq.
__device__ inline uint64_t knn_key(float r_sq, uint32_t jc)
{
    return (uint64_t(__float_as_uint(sqrtf(r_sq))) << 32) | jc;
}
__device__ inline float knn_key_distance(uint64_t key)
{
    return __uint_as_float(uint32_t(key >> 32));
}
// The square root is only taken for candidates that can enter the selection:
// once it is full, a candidate needs d <= d_last, which implies r_sq below
// d_last^2 (1 + 2^-22) for the correctly rounded sqrtf; the bound below leaves
// a margin of 4x that plus an absolute term for subnormal squares.  A bound
// that admits too much only costs a square root, never a different result.
template<int KMAX> __device__ inline float knn_rsq_cut(const TopK<KMAX>& best, float r_sq_limit)
{
    uint64_t const last = best.v[KMAX - 1];
    if (last == ~0ull)
        return r_sq_limit;
    float const d = knn_key_distance(last);
    return fminf(r_sq_limit, d * d * (1.0f + 0x1p-20f) + 1e-37f);
}

// AABB formula kNN (AABBQuery.cc:152-281).  The candidate set is every (j,
// image) with r_sq < rlim_sq (rlim = the last ball radius of the reference's
// expanding search) and sqrt(r_sq) >= r_min; the k smallest by (r_sq, j) are
// kept.  A query whose answer needs radii beyond half the smallest plane
// distance (where one point can appear through two images) is flagged for the
// exhaustive kernel below.
template<int KMAX, bool SELF>
__device__ inline void knn_ring_search(const DevBox& b, const DevGrid& g, const float* __restrict__ qpts, uint32_t p,
                                       uint32_t k, float rlim_sq, float rlim, float rmin_thr, int exclude_ii,
                                       float wmin, float half_min_plane, int rmax_ring, uint32_t* __restrict__ counts,
                                       uint64_t* __restrict__ sel, uint32_t* __restrict__ flagged,
                                       uint32_t* __restrict__ nflag, uint32_t* __restrict__ pos)
{
    float qx, qy, qz;
    uint32_t const i = query_at<SELF>(g, qpts, p, qx, qy, qz);
    if (SELF)
        pos[i] = p;
    TopK<KMAX> best;
    best.init();
    uint32_t found = 0;
    bool extended = false;
    float cut = rlim_sq; // candidates above it cannot enter the selection
    for (int R = 0; R <= rmax_ring; ++R)
    {
        fa::for_each_candidate<AABB_FORMULA>(b, g, qx, qy, qz, R, true,
                                             [&](uint32_t j, float, float, float, float r_sq, int code) {
                                                 if (!(r_sq < rlim_sq))
                                                     return;
                                                 if (exclude_ii && j == i)
                                                     return;
                                                 // sqrtf(r_sq) >= r_min, as one compare (rmin_thr
                                                 // is the smallest r_sq that passes it)
                                                 if (!(r_sq >= rmin_thr))
                                                     return;
                                                 ++found;
                                                 if (!(r_sq <= cut))
                                                     return;
                                                 best.insert(knn_key(r_sq, (j << 5) | uint32_t(code)));
                                                 cut = knn_rsq_cut(best, rlim_sq);
                                             });
        float const complete = float(R) * wmin * (1.0f - 1e-4f);
        if (found >= k)
        {
            float const kth = knn_key_distance(best.at(k - 1));
            if (kth < complete)
                break;
        }
        if (complete >= rlim)
            break;
        if (complete >= half_min_plane && rlim > half_min_plane)
        {
            extended = true;
            break;
        }
    }
    // Two images of one point are at least a plane distance apart, so only a
    // selection reaching half the smallest plane distance can hold a point
    // twice; such queries take the exhaustive, deduplicating kernel.
    uint32_t const nsel = min(found, k);
    if (!extended && rlim > half_min_plane && nsel > 0)
    {
        float const last = knn_key_distance(best.at(nsel - 1));
        if (last >= half_min_plane)
            extended = true;
    }
    if (extended)
    {
        uint32_t const slot = atomicAdd(nflag, 1u);
        flagged[slot] = p;
        counts[i] = 0;
        return;
    }
    uint32_t const n = min(found, k);
    counts[i] = n;
    for (uint32_t m = 0; m < n; ++m)
        sel[uint64_t(p) * k + m] = best.at(m);
}

// The ring search over every query position, or (LIST) over the processing
// positions list[0 .. *nlist), grid-stride, for the points the staged search
// below hands back.
template<int KMAX, bool SELF, bool LIST>
__global__ __launch_bounds__(kBlock) void knn_aabb_kernel(DevBox b, DevGrid g, const float* __restrict__ qpts,
                                                          uint32_t nq, uint32_t k, float rlim_sq, float rlim,
                                                          float rmin_thr, int exclude_ii, float wmin,
                                                          float half_min_plane, int rmax_ring,
                                                          uint32_t* __restrict__ counts, uint64_t* __restrict__ sel,
                                                          uint32_t* __restrict__ flagged, uint32_t* __restrict__ nflag,
                                                          uint32_t* __restrict__ pos, const uint32_t* __restrict__ list,
                                                          const uint32_t* __restrict__ nlist)
{
    if (LIST)
    {
        uint32_t const nl = *nlist;
        for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nl; t += gridDim.x * blockDim.x)
            knn_ring_search<KMAX, SELF>(b, g, qpts, list[t], k, rlim_sq, rlim, rmin_thr, exclude_ii, wmin,
                                        half_min_plane, rmax_ring, counts, sel, flagged, nflag, pos);
        return;
    }
    uint32_t const p = fa::xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p < nq)
        knn_ring_search<KMAX, SELF>(b, g, qpts, p, k, rlim_sq, rlim, rmin_thr, exclude_ii, wmin, half_min_plane,
                                    rmax_ring, counts, sel, flagged, nflag, pos);
}

// Staged AABB kNN: the same rings and the same answer as the ring search, with
// the selection moved out of the walk.  Inserting into a register top-k
// during the walk costs every lane of a wave the insertion whenever one lane
// inserts, which is nearly every candidate.  Here a candidate with
// d < r_stage (a radius holding ~k + sqrt(k) + 1 points at the system's
// mean density; stage_thr is the smallest r_sq whose sqrtf reaches it) is
// only recorded in the lane's LDS stage, as its cell-order slot and image
// code (4 bytes, so the stage leaves room for several blocks per CU); the k
// smallest are picked after the walk, their distances recomputed from the
// slot with the walk's own arithmetic.  A full stage is cut to its k smallest,
// whose k-th then bounds what may enter.  The rings end as in the ring search,
// once k staged distances lie below the radius the walked rings cover
// (counted as they are staged, for the first three rings).  A lane whose
// answer may hold candidates the stage did not take (fewer than k staged
// once the rings cover r_stage), or that needs a fourth ring or the
// half-box radius, is handed back (`refine`) to the ring search.
template<int KMAX, bool SELF>
__global__ __launch_bounds__(kBlock) void knn_stage_kernel(DevBox b, DevGrid g, const float* __restrict__ qpts,
                                                           uint32_t nq, uint32_t k, float rlim_sq, float rlim,
                                                           float rmin_thr, int exclude_ii, float wmin,
                                                           float half_min_plane, int rmax_ring, float stage_thr,
                                                           float r_stage, uint32_t C, uint32_t* __restrict__ counts,
                                                           uint64_t* __restrict__ sel, uint32_t* __restrict__ flagged,
                                                           uint32_t* __restrict__ nflag, uint32_t* __restrict__ refine,
                                                           uint32_t* __restrict__ nrefine, uint32_t* __restrict__ pos)
{
    // s_e[m * kBlock + tid]: staged entry m of thread tid, slot << 5 | image code
    uint32_t* s_e = reinterpret_cast<uint32_t*>(fa_dyn_lds);
    uint32_t const tid = threadIdx.x;
    uint32_t const p0 = fa::xcd_tile(blockIdx.x, gridDim.x) * blockDim.x;
    uint32_t const p = p0 + tid;
    TopK<KMAX> best; // the lane's selection, ascending
    best.init();
    if (p < nq)
    {
        float qx, qy, qz;
        uint32_t const i = query_at<SELF>(g, qpts, p, qx, qy, qz);
        if (SELF)
            pos[i] = p;
        float const qzz = b.is2D ? 0.0f : qz;
        // selection key of a staged entry: (d, j << 5 | code), r_sq as the walk computed it
        auto key_of = [&](uint32_t e, float& r_sq) -> uint64_t {
            uint32_t const code = e & 31u;
            float4 const pj = g.cpos[e >> 5];
            float const pz = b.is2D ? 0.0f : pj.z;
            float const rx = pj.x - (qx + b.img[code][0]);
            float const ry = pj.y - (qy + b.img[code][1]);
            float const rz = pz - (qzz + b.img[code][2]);
            r_sq = (rx * rx + ry * ry) + rz * rz;
            return knn_key(r_sq, (__float_as_uint(pj.w) << 5) | code);
        };
        // squared radii covered after rings 1, 2, 3 (termination counts)
        float const w1 = wmin * (1.0f - 1e-4f);
        float const c1 = w1 * w1, c2 = (2.0f * w1) * (2.0f * w1), c3 = (3.0f * w1) * (3.0f * w1);
        uint32_t found = 0, c = 0, in1 = 0, in2 = 0, in3 = 0;
        bool hand_back = false;
        float cut = INFINITY; // after a cut: candidates above it cannot be among the k smallest
        auto reduce = [&]() {
            TopK<KMAX> best;
            best.init();
            float r_sq;
            for (uint32_t m = 0; m < c; ++m)
                best.insert(key_of(s_e[m * kBlock + tid], r_sq));
            uint64_t const kth = best.at(k - 1);
            uint32_t w = 0;
            in1 = in2 = in3 = 0;
            for (uint32_t m = 0; m < c; ++m)
            {
                uint32_t const e = s_e[m * kBlock + tid];
                if (key_of(e, r_sq) <= kth)
                {
                    s_e[w * kBlock + tid] = e;
                    ++w;
                    in1 += r_sq < c1;
                    in2 += r_sq < c2;
                    in3 += r_sq < c3;
                }
            }
            c = w;
            float const d = knn_key_distance(kth);
            cut = d * d * (1.0f + 0x1p-20f) + 1e-37f;
        };
        int R = 0;
        for (; R <= rmax_ring; ++R)
        {
            fa::for_each_candidate<AABB_FORMULA>(
                b, g, qx, qy, qz, R, true,
                [&](uint32_t j, float, float, float, float r_sq, int code, const float4&, uint32_t slot) {
                    if (!(r_sq < rlim_sq))
                        return;
                    if (exclude_ii && j == i)
                        return;
                    if (!(r_sq >= rmin_thr))
                        return;
                    ++found;
                    if (!(r_sq < stage_thr && r_sq <= cut))
                        return;
                    if (c == C)
                    {
                        reduce();
                        if (!(r_sq <= cut))
                            return;
                    }
                    s_e[c * kBlock + tid] = (slot << 5) | uint32_t(code);
                    ++c;
                    in1 += r_sq < c1;
                    in2 += r_sq < c2;
                    in3 += r_sq < c3;
                });
            float const complete = float(R) * w1;
            uint32_t const in = R == 1 ? in1 : R == 2 ? in2 : R == 3 ? in3 : 0u;
            if (c >= k && in >= k)
                break;
            if (c < k && found > c && complete >= r_stage)
            {
                hand_back = true; // the answer holds candidates at or above r_stage
                break;
            }
            if (complete >= rlim)
                break;
            if ((complete >= half_min_plane && rlim > half_min_plane) || R == 3)
            {
                hand_back = true; // longer walks, and the exhaustive flag, are the ring search's
                break;
            }
        }
        uint32_t const n = min(found, k);
        if (hand_back || c < n)
            refine[atomicAdd(nrefine, 1u)] = p;
        else
        {
            // c >= n: the n smallest keys are staged
            float r_sq;
            for (uint32_t m = 0; m < c; ++m)
                best.insert(key_of(s_e[m * kBlock + tid], r_sq));
            float const last = n ? knn_key_distance(best.at(n - 1)) : 0.0f;
            // two images of one point are at least a plane distance apart (see the ring search)
            if (rlim > half_min_plane && n > 0 && last >= half_min_plane)
            {
                flagged[atomicAdd(nflag, 1u)] = p;
                counts[i] = 0;
            }
            else
                counts[i] = n;
        }
    }
    // The block's selection rows (positions p0 .. p0 + 255) are one contiguous
    // range of `sel`: staged in LDS once the stage is free, then stored with
    // coalesced writes.  Entries past a lane's count, and the rows of lanes
    // handed back or flagged, carry no result (the searches that take those
    // lanes rewrite their rows later on the stream; entries past a count are
    // never read).
    __syncthreads();
    uint64_t* s_out = reinterpret_cast<uint64_t*>(fa_dyn_lds);
#pragma unroll
    for (int m = 0; m < KMAX; ++m)
        if (uint32_t(m) < k)
            s_out[tid * k + m] = best.v[m];
    __syncthreads();
    uint32_t const nrow = min(uint32_t(kBlock), nq - p0) * k;
    for (uint32_t t = tid; t < nrow; t += kBlock)
        sel[uint64_t(p0) * k + t] = s_out[t];
}

// LinkCell formula kNN (LinkCell.cc:575-679): every point in cells within
// max_range - 1 Chebyshev shells of the reference's grid, each cell once.
template<int KMAX>
__global__ __launch_bounds__(kBlock) void knn_linkcell_kernel(DevBox b, DevGrid g, const float* __restrict__ qpts,
                                                              uint32_t nq, uint32_t k, float r_max_sq,
                                                              float r_min_sq, int exclude_ii, int R,
                                                              uint32_t* __restrict__ counts,
                                                              uint64_t* __restrict__ sel)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq)
        return;
    float const qx = qpts[3 * i], qy = qpts[3 * i + 1], qz = qpts[3 * i + 2];
    TopK<KMAX> best;
    best.init();
    uint32_t found = 0;
    float cut = r_max_sq;
    fa::for_each_candidate<LINKCELL_FORMULA>(b, g, qx, qy, qz, R, false,
                                             [&](uint32_t j, float, float, float, float r_sq, int) {
                                                 if (r_sq < r_max_sq && r_sq >= r_min_sq && !(exclude_ii && j == i))
                                                 {
                                                     ++found;
                                                     if (!(r_sq <= cut))
                                                         return;
                                                     best.insert(knn_key(r_sq, (j << 5) | 13u));
                                                     cut = knn_rsq_cut(best, r_max_sq);
                                                 }
                                             });
    uint32_t const n = min(found, k);
    counts[i] = n;
    for (uint32_t m = 0; m < n; ++m)
        sel[uint64_t(i) * k + m] = best.at(m);
}

// Exhaustive kNN for flagged queries (one block per query): every point and
// every image in {-1,0,1}^d; per point the closest image is kept (the
// reference's minimum-distance map, AABBQuery.cc:188-205); a point whose
// closest image is nearer than r_min is dropped (AABBQuery.cc:197-201, 235-244).
__global__ __launch_bounds__(kBlock) void knn_exhaustive_kernel(DevBox b, const float* __restrict__ pts, uint32_t np,
                                                                const float* __restrict__ qpts, uint32_t nq,
                                                                const float4* __restrict__ cpos,
                                                                const uint32_t* __restrict__ flagged, uint32_t k,
                                                                float rlim_sq, float r_min, int exclude_ii,
                                                                uint64_t* __restrict__ work,
                                                                uint32_t* __restrict__ counts,
                                                                uint64_t* __restrict__ sel,
                                                                uint32_t* __restrict__ pos)
{
    // flagged holds processing positions: with cpos (queries in cell order)
    // the point index is the cell-ordered entry's, else the position itself
    uint32_t const p = flagged[blockIdx.x];
    uint32_t const i = cpos ? __float_as_uint(cpos[p].w) : p;
    if (cpos && threadIdx.x == 0)
        pos[i] = p;
    uint64_t* w = work + uint64_t(blockIdx.x) * np;
    float const qx = qpts[3 * i], qy = qpts[3 * i + 1];
    float const qz = b.is2D ? 0.0f : qpts[3 * i + 2];
    int const nimg_z = b.is2D ? 0 : 1;
    for (uint32_t j = threadIdx.x; j < np; j += blockDim.x)
    {
        uint64_t bestk = ~0ull;
        if (!(exclude_ii && j == i))
        {
            float const px = pts[3 * j], py = pts[3 * j + 1];
            float const pz = b.is2D ? 0.0f : pts[3 * j + 2];
            for (int ix = -1; ix <= 1; ++ix)
                for (int iy = -1; iy <= 1; ++iy)
                    for (int iz = -nimg_z; iz <= nimg_z; ++iz)
                    {
                        int const code = (ix + 1) * 9 + (iy + 1) * 3 + (iz + 1);
                        float const rx = px - (qx + b.img[code][0]);
                        float const ry = py - (qy + b.img[code][1]);
                        float const rz = pz - (qz + b.img[code][2]);
                        float const r_sq = (rx * rx + ry * ry) + rz * rz;
                        if (r_sq < rlim_sq)
                            bestk = min(bestk, knn_key(r_sq, (j << 5) | uint32_t(code)));
                    }
            if (bestk != ~0ull && !(knn_key_distance(bestk) >= r_min))
                bestk = ~0ull;
        }
        w[j] = bestk;
    }
    __syncthreads();
    // k rounds of block-wide minimum extraction
    __shared__ uint64_t s_red[kBlock];
    __shared__ uint32_t s_cnt;
    if (threadIdx.x == 0)
        s_cnt = 0;
    __syncthreads();
    for (uint32_t m = 0; m < k; ++m)
    {
        uint64_t lm = ~0ull;
        for (uint32_t j = threadIdx.x; j < np; j += blockDim.x)
            lm = min(lm, w[j]);
        s_red[threadIdx.x] = lm;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1)
        {
            if (threadIdx.x < s)
                s_red[threadIdx.x] = min(s_red[threadIdx.x], s_red[threadIdx.x + s]);
            __syncthreads();
        }
        uint64_t const mn = s_red[0];
        __syncthreads();
        if (mn == ~0ull)
            break;
        if (threadIdx.x == 0)
        {
            sel[uint64_t(p) * k + m] = mn;
            s_cnt = m + 1;
        }
        uint32_t const jm = uint32_t(mn) >> 5;
        if (jm % blockDim.x == threadIdx.x)
            w[jm] = ~0ull;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        counts[i] = s_cnt;
}

// Writes a kNN result.  With INDEX a block takes 256 consecutive query points,
// whose bonds are one contiguous range of the NeighborList (full-line
// stores); otherwise the search's processing order, which keeps neighbour
// positions in cache for points stored in random order.  Each point's
// selected keys (point-major rows by processing position: pos[i] for
// cell-ordered self queries) are sorted into the NeighborList order ((i, j) or
// (i, d, j)), staged in LDS and stored cooperatively.
template<int FORMULA, bool BYDIST, bool SELF, bool INDEX>
__global__ __launch_bounds__(kBlock) void knn_write_kernel(DevBox b, DevGrid g, const float* __restrict__ qpts, uint32_t nq,
                                                           const float* __restrict__ pts, uint32_t k,
                                                           const uint64_t* __restrict__ sel,
                                                           const uint32_t* __restrict__ pos,
                                                           const uint32_t* __restrict__ offsets,
                                                           uint32_t* __restrict__ nb_out, float* __restrict__ d_out,
                                                           float* __restrict__ w_out, float* __restrict__ v_out,
                                                           uint64_t* __restrict__ tmp, uint32_t key_cap)
{
    using K = typename std::conditional<BYDIST, uint64_t, uint32_t>::type;
    K* s_keys = reinterpret_cast<K*>(fa_dyn_lds); // key_cap entries (dynamic LDS)
    uint32_t const CAP = key_cap;
    __shared__ uint32_t s_off[kBlock + 1];
    __shared__ uint32_t s_i[kBlock];
    __shared__ uint32_t s_tmp[kBlock];
    __shared__ uint32_t s_row[kBlock];
    __shared__ float s_q[3 * kBlock];
    uint32_t const tid = threadIdx.x;
    uint32_t const p0 = fa::xcd_tile(blockIdx.x, gridDim.x) * kBlock;
    uint32_t const nloc = min(uint32_t(kBlock), nq - p0);
    uint32_t i = p0 + tid;
    uint32_t n = 0;
    if (tid < nloc)
    {
        float qx, qy, qz;
        if (INDEX)
        {
            qx = qpts[3 * i];
            qy = qpts[3 * i + 1];
            qz = qpts[3 * i + 2];
        }
        else
            i = query_at<SELF>(g, qpts, p0 + tid, qx, qy, qz);
        n = offsets[i + 1] - offsets[i];
        s_i[tid] = i;
        s_q[3 * tid] = qx;
        s_q[3 * tid + 1] = qy;
        s_q[3 * tid + 2] = qz;
    }
    uint32_t excl;
    uint32_t const total = block_exclusive_scan(n, s_tmp, excl);
    if (tid < nloc)
        s_off[tid] = excl;
    if (tid == 0)
        s_off[nloc] = total;
    __syncthreads();
    bool const in_lds = total <= CAP;
    auto key_of = [&](uint64_t sk) -> K {
        uint32_t const jc = uint32_t(sk);
        if (BYDIST)
            return K((uint64_t(float_order_bits(knn_key_distance(sk))) << 32) | jc);
        return K(jc);
    };
    if (in_lds)
    {
        // The block's selection rows into LDS, element t by thread t: one load
        // instruction covers whole rows.  (Each lane reading its own row in k
        // steps keeps ~12 MB of rows per XCD half read, far beyond its L2, and
        // fetched them about four times.)
        if (tid < nloc)
            s_row[tid] = INDEX && SELF ? pos[i] : p0 + tid;
        __syncthreads();
        for (uint32_t t = tid; t < total; t += kBlock)
        {
            uint32_t const lo = segment_of(s_off, nloc, t);
            s_keys[t] = key_of(sel[uint64_t(s_row[lo]) * k + (t - s_off[lo])]);
        }
        __syncthreads();
        if (tid < nloc)
            insertion_sort(s_keys + excl, n);
    }
    else if (tid < nloc)
    {
        K* keys = reinterpret_cast<K*>(tmp + uint64_t(i) * k);
        const uint64_t* row = sel + uint64_t(INDEX && SELF ? pos[i] : p0 + tid) * k;
        for (uint32_t m = 0; m < n; ++m)
            keys[m] = key_of(row[m]);
        insertion_sort(keys, n);
        for (uint32_t m = 0; m < n; ++m)
            write_bond<FORMULA, BYDIST>(b, qpts, pts, i, uint64_t(keys[m]), uint64_t(offsets[i]) + m, nb_out, d_out,
                                        w_out, v_out);
    }
    __syncthreads();
    if (in_lds)
        store_block_bonds<FORMULA, BYDIST, K, INDEX>(b, s_q, pts, s_keys, total, nloc, s_off, s_i, offsets, nb_out, d_out,
                                           w_out, v_out);
}

// ------------------------------------------------------------------ all pairs
__global__ void all_pairs_kernel(DevBox b, const float* __restrict__ pts, uint32_t np, const float* __restrict__ qpts,
                                 uint32_t nq, int exclude_ii, uint32_t* __restrict__ nb_out, float* __restrict__ d_out,
                                 float* __restrict__ w_out, float* __restrict__ v_out)
{
    uint64_t const t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= uint64_t(np) * nq)
        return;
    uint32_t const i = uint32_t(t / np), j = uint32_t(t % np);
    if (exclude_ii && i == j)
        return;
    // NeighborList.cc:103-125: row start i*np - min(i, np); the vector is wrap(q_i - p_j)
    uint64_t idx = uint64_t(i) * np + j;
    if (exclude_ii)
        idx -= uint64_t(min(i, np)) + ((i < np && j > i) ? 1 : 0);
    float rx, ry, rz;
    fa::wrap(b, qpts[3 * i] - pts[3 * j], qpts[3 * i + 1] - pts[3 * j + 1], qpts[3 * i + 2] - pts[3 * j + 2], rx, ry,
             rz);
    reinterpret_cast<uint2*>(nb_out)[idx] = make_uint2(i, j);
    d_out[idx] = sqrtf((rx * rx + ry * ry) + rz * rz);
    w_out[idx] = 1.0f;
    v_out[3 * idx] = rx;
    v_out[3 * idx + 1] = ry;
    v_out[3 * idx + 2] = rz;
}

// ------------------------------------------------------------------ batches
// Ball query over every frame of a batch (query points = the frame's points,
// AABB formula, exclude_ii): one workgroup per 256-point tile of one frame,
// positions in the frame's cell order.  Counts and offsets are indexed by
// batch point index; bonds hold frame-local (i, j).
__global__ __launch_bounds__(kBlock) void bball_count_kernel(fa::BatchGrid G, const uint32_t* __restrict__ tile_frame,
                                                             uint32_t ntiles, float r_max_sq, int R,
                                                             uint32_t* __restrict__ counts)
{
    uint32_t const t = fa::xcd_tile(blockIdx.x, ntiles);
    const fa::BatchFrame& F = G.frames[tile_frame[t]];
    uint32_t const local = (t - F.tile_base) * fa::kTile + threadIdx.x;
    if (local >= F.n)
        return;
    DevGrid const g = fa::frame_grid(F, G);
    ball_count_body<AABB_FORMULA, true>(F.box, g, nullptr, F.pt_base + local, r_max_sq, 0.0f, 1, R,
                                        counts + F.pt_base);
}

__global__ __launch_bounds__(kBlock) void bball_fill_kernel(fa::BatchGrid G, const uint32_t* __restrict__ tile_frame,
                                                            uint32_t ntiles, const float* __restrict__ pts,
                                                            float r_max_sq, int R,
                                                            const uint32_t* __restrict__ offsets,
                                                            const unsigned long long* __restrict__ total_bonds,
                                                            uint64_t capacity, uint32_t* __restrict__ nb_out,
                                                            float* __restrict__ d_out, float* __restrict__ w_out,
                                                            float* __restrict__ v_out, void* __restrict__ gkeys,
                                                            uint32_t key_cap)
{
    // a run whose bonds exceed the allocation writes nothing (the host sees
    // the 64-bit total and runs again with room for it)
    if (*total_bonds > capacity)
        return;
    uint32_t const t = fa::xcd_tile(blockIdx.x, ntiles);
    const fa::BatchFrame& F = G.frames[tile_frame[t]];
    uint32_t const local0 = (t - F.tile_base) * fa::kTile;
    DevGrid const g = fa::frame_grid(F, G);
    const float* fp = pts + 3 * size_t(F.pt_base);
    ball_fill_body<AABB_FORMULA, false, true>(F.box, g, fp, F.pt_base + local0, min(fa::kTile, F.n - local0), fp,
                                              r_max_sq, 0.0f, 1, R, offsets + F.pt_base, nb_out, d_out, w_out, v_out,
                                              gkeys, key_cap);
}

// The batched one-walk ball query (tile t of the batch, its frame's grid and
// cell order; see ball_walk_body and ball_emit_body).  The emission runs in the
// walk's processing order, where each tile's keys are one packed region and
// its neighbours' positions stay in cache (the frames' points may be stored in
// any order).
__global__ __launch_bounds__(kBlock) void bball_walk_kernel(fa::BatchGrid G, const uint32_t* __restrict__ tile_frame,
                                                            uint32_t ntiles, float r_max_sq, int R, uint32_t C,
                                                            uint32_t* __restrict__ counts, uint32_t* __restrict__ pos,
                                                            uint32_t* __restrict__ rows)
{
    uint32_t const t = fa::xcd_tile(blockIdx.x, ntiles);
    const fa::BatchFrame& F = G.frames[tile_frame[t]];
    uint32_t const local0 = (t - F.tile_base) * fa::kTile;
    DevGrid const g = fa::frame_grid(F, G);
    ball_walk_body<AABB_FORMULA, true>(F.box, g, nullptr, F.pt_base + local0, min(fa::kTile, F.n - local0),
                                       t * fa::kTile * C, r_max_sq, 0.0f, 1, R, fa::CQGrid{}, C, counts + F.pt_base,
                                       pos + F.pt_base, rows);
}

__global__ __launch_bounds__(kBlock) void bball_emit_kernel(fa::BatchGrid G, const uint32_t* __restrict__ tile_frame,
                                                            uint32_t ntiles, const float* __restrict__ pts,
                                                            float r_max_sq, int R, uint32_t C,
                                                            const uint32_t* __restrict__ pos,
                                                            const uint32_t* __restrict__ rows,
                                                            const uint32_t* __restrict__ offsets,
                                                            const unsigned long long* __restrict__ total_bonds,
                                                            uint64_t capacity, uint32_t* __restrict__ nb_out,
                                                            float* __restrict__ d_out, float* __restrict__ w_out,
                                                            float* __restrict__ v_out, void* __restrict__ gkeys,
                                                            uint32_t key_cap)
{
    // a run whose bonds exceed the allocation writes nothing (the host sees
    // the 64-bit total and runs again with room for it)
    if (*total_bonds > capacity)
        return;
    uint32_t const t = fa::xcd_tile(blockIdx.x, ntiles);
    const fa::BatchFrame& F = G.frames[tile_frame[t]];
    uint32_t const local0 = (t - F.tile_base) * fa::kTile;
    DevGrid const g = fa::frame_grid(F, G);
    const float* fp = pts + 3 * size_t(F.pt_base);
    ball_emit_body<AABB_FORMULA, false, true, false>(F.box, g, fp, F.pt_base + local0, min(fa::kTile, F.n - local0),
                                                     fp, r_max_sq, 0.0f, 1, R, fa::CQGrid{}, C, pos + F.pt_base, rows,
                                                     offsets + F.pt_base, nb_out, d_out, w_out, v_out, gkeys,
                                                     key_cap);
}

} // namespace

namespace fa {

int batch_ball_count(const BatchGrid& G, const uint32_t* d_tile_frame, uint32_t ntiles, float r_max, int R,
                     uint32_t* d_counts)
{
    if (ntiles)
        FA_KLAUNCH(bball_count_kernel, dim3(ntiles), dim3(kBlock), 0, stream(), G, d_tile_frame, ntiles,
                   r_max * r_max, R, d_counts);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int batch_ball_fill(const BatchGrid& G, const uint32_t* d_tile_frame, uint32_t ntiles, const float* d_pts, float r_max,
                    int R, const uint32_t* d_offsets, const unsigned long long* d_total, uint64_t capacity,
                    uint32_t* nb, float* dist, float* wts, float* vec, void* gkeys)
{
    // the key staging is sized from the allocation (bonds per point of the batch)
    size_t const kb = fill_key_bytes(capacity, G.total_points, sizeof(uint32_t));
    if (ntiles)
        FA_KLAUNCH(bball_fill_kernel, dim3(ntiles), dim3(kBlock), kb, stream(), G, d_tile_frame, ntiles, d_pts,
                   r_max * r_max, R, d_offsets, d_total, capacity, nb, dist, wts, vec, gkeys,
                   uint32_t(kb / sizeof(uint32_t)));
    FA_HIP(hipGetLastError());
    return FA_OK;
}


uint32_t batch_ball_slots(double max_points_per_volume, bool is2D, float r_max)
{
    double const r = double(r_max);
    double const expect = is2D ? max_points_per_volume * M_PI * r * r
                               : max_points_per_volume * (4.0 / 3.0) * M_PI * r * r * r;
    double const c = std::ceil((expect * 1.15 + 4.0) / 4.0) * 4.0;
    if (c > double(kWalkMaxSlots))
        return 0;
    return uint32_t(std::max(c, 8.0));
}

int batch_ball_walk(const BatchGrid& G, const uint32_t* d_tile_frame, uint32_t ntiles, float r_max, int R, uint32_t C,
                    uint32_t* d_counts, uint32_t* d_pos, uint32_t* d_rows)
{
    if (ntiles)
        FA_KLAUNCH(bball_walk_kernel, dim3(ntiles), dim3(kBlock), size_t(C) * kBlock * sizeof(uint32_t), stream(), G,
                   d_tile_frame, ntiles, r_max * r_max, R, C, d_counts, d_pos, d_rows);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int batch_ball_emit(const BatchGrid& G, const uint32_t* d_tile_frame, uint32_t ntiles, const float* d_pts, float r_max,
                    int R, uint32_t C, const uint32_t* d_pos, const uint32_t* d_rows, const uint32_t* d_offsets,
                    const unsigned long long* d_total, uint64_t capacity, uint32_t* nb, float* dist, float* wts,
                    float* vec, void* gkeys)
{
    size_t const kb = fill_key_bytes(capacity, G.total_points, sizeof(uint32_t));
    if (ntiles)
        FA_KLAUNCH(bball_emit_kernel, dim3(ntiles), dim3(kBlock), kb, stream(), G, d_tile_frame, ntiles, d_pts,
                   r_max * r_max, R, C, d_pos, d_rows, d_offsets, d_total, capacity, nb, dist, wts, vec, gkeys,
                   uint32_t(kb / sizeof(uint32_t)));
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int nlist_reserve(fa_nlist* nl, uint64_t nb, uint32_t nq)
{
    if (nl->capacity < nb || !nl->neighbors)
    {
        device_free(nl->neighbors);
        device_free(nl->distances);
        device_free(nl->weights);
        device_free(nl->vectors);
        nl->neighbors = nullptr;
        nl->distances = nl->weights = nl->vectors = nullptr;
        nl->capacity = 0;
        uint64_t const cap = std::max<uint64_t>(nb, 1);
        if (int rc = device_alloc((void**) &nl->neighbors, sizeof(uint32_t) * 2 * cap))
            return rc;
        if (int rc = device_alloc((void**) &nl->distances, sizeof(float) * cap))
            return rc;
        if (int rc = device_alloc((void**) &nl->weights, sizeof(float) * cap))
            return rc;
        if (int rc = device_alloc((void**) &nl->vectors, sizeof(float) * 3 * cap))
            return rc;
        nl->capacity = cap;
    }
    if (!nl->counts || nl->num_query_points < nq)
    {
        device_free(nl->counts);
        device_free(nl->segments);
        device_free(nl->offsets);
        nl->counts = nl->segments = nl->offsets = nullptr;
        if (int rc = device_alloc((void**) &nl->counts, sizeof(uint32_t) * std::max<uint32_t>(nq, 1)))
            return rc;
        if (int rc = device_alloc((void**) &nl->segments, sizeof(uint32_t) * std::max<uint32_t>(nq, 1)))
            return rc;
        if (int rc = device_alloc((void**) &nl->offsets, sizeof(uint32_t) * (size_t(nq) + 1)))
            return rc;
    }
    nl->num_query_points = nq;
    return FA_OK;
}

int nlist_finish_counts(fa_nlist* nl)
{
    uint32_t const nq = nl->num_query_points;
    if (nq == 0)
        return FA_OK;
    FA_KLAUNCH(counts_segments_kernel, dim3((nq + 255) / 256), dim3(256), 0, stream(), nl->offsets, nq,
                       nl->counts, nl->segments);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

// Cell-grid shape for a stencil of half-width R covering r: every cell is at
// least r/R (+0.01 %) wide along each plane normal, and the cell count stays
// within a small multiple of the point count.
void grid_dims_for(const fa_box& box, uint32_t n, float r, int R, int dims[3])
{
    float npd[3];
    nearest_plane_distance(box, npd);
    double const w = double(r) * (1.0 + 1e-4) / R;
    for (int k = 0; k < 3; ++k)
    {
        double d = (k == 2 && box.is2D) ? 1.0 : std::floor(double(npd[k]) / w);
        if (!(d >= 1.0))
            d = 1.0;
        dims[k] = int(std::min(d, 1.0e6));
    }
    double const max_cells = std::max<double>(8.0 * n, 64.0);
    while (double(dims[0]) * dims[1] * dims[2] > max_cells)
    {
        for (int k = 0; k < 3; ++k)
            if (!(k == 2 && box.is2D))
                dims[k] = std::max(1, int(dims[k] * 0.8));
    }
}

CQGrid cq_grid_for(const fa_box& box, float r_cut)
{
    // CellQuery::setupGrid (CellQuery.h:136-160) in the reference's float order
    DevBox const db = make_devbox(box);
    float const Lx = db.L[0], Ly = db.L[1], Lz = db.L[2];
    CQGrid c;
    c.inv = 1.0f / r_cut;
    float const w_x = (Lx + Ly * std::abs(box.xy)) + Lz * std::abs(box.xz);
    float const w_y = Ly + Lz * std::abs(box.yz);
    float const w_z = Lz;
    c.n[0] = int(w_x * c.inv) + 3;
    c.n[1] = int(w_y * c.inv) + 3;
    c.n[2] = int(w_z * c.inv) + 3;
    float const box_min_x = Ly * std::min(0.0f, box.xy) + Lz * std::min(0.0f, box.xz);
    float const box_min_y = Lz * std::min(0.0f, box.yz);
    // makeAbsolute(0, 0, 0) (Box.h:212-222): lo + 0 * L, then the tilt terms
    float ax = db.lo[0] + 0.0f * Lx, ay = db.lo[1] + 0.0f * Ly, az = db.lo[2] + 0.0f * Lz;
    ax += box.xy * ay + box.xz * az;
    ay += box.yz * az;
    if (db.is2D)
        az = 0.0f;
    c.min[0] = (box_min_x - r_cut) + ax;
    c.min[1] = (box_min_y - r_cut) + ay;
    c.min[2] = -r_cut + az;
    // CellQuery.cc:76-79: r_cut / nearest plane distance per axis
    float npd[3];
    nearest_plane_distance(box, npd);
    for (int k = 0; k < 3; ++k)
        c.frc[k] = r_cut / npd[k];
    return c;
}

int ensure_grid(fa_nq* nq, const int dims[3])
{
    ++nq->tick;
    int victim = 0;
    for (int s = 0; s < fa_nq::kGrids; ++s)
    {
        if (nq->grid_valid[s] && nq->grid_dims[s][0] == dims[0] && nq->grid_dims[s][1] == dims[1]
            && nq->grid_dims[s][2] == dims[2])
        {
            nq->grid_used[s] = nq->tick;
            nq->cur = &nq->grids[s];
            return FA_OK;
        }
        if (!nq->grid_valid[s] || (nq->grid_valid[victim] && nq->grid_used[s] < nq->grid_used[victim]))
            victim = s;
    }
    nq->grid_valid[victim] = false;
    if (int rc = build_grid(nq->grids[victim], nq->box, nq->d_points, nq->n, dims))
        return rc;
    nq->grid_valid[victim] = true;
    nq->grid_used[victim] = nq->tick;
    for (int k = 0; k < 3; ++k)
        nq->grid_dims[victim][k] = dims[k];
    nq->cur = &nq->grids[victim];
    return FA_OK;
}

// NeighborQuery::validateQueryArgs / validateNearestNeighborArgs and the
// per-point iterator checks (NeighborQuery.h:195-292, 317-328; AABBQuery.h:62-69).
int validate_query_args(const fa_nq* nq, fa_query_args& a, uint32_t n_query)
{
    constexpr uint32_t kDefaultK = 0xffffffffu;
    if (a.mode == FA_MODE_NONE)
    {
        if (a.num_neighbors != kDefaultK)
            a.mode = FA_MODE_NEAREST;
        else if (a.r_max != -1.0f)
            a.mode = FA_MODE_BALL;
    }
    if (a.mode == FA_MODE_BALL)
    {
        if (a.r_max == -1.0f)
            return fail(FA_ERR_RUNTIME, "You must set r_max in the query arguments when performing ball queries.");
        if (a.num_neighbors != kDefaultK)
            return fail(FA_ERR_RUNTIME,
                        "You cannot set num_neighbors in the query arguments when performing ball queries.");
    }
    else if (a.mode == FA_MODE_NEAREST)
    {
        if (a.num_neighbors == kDefaultK)
            return fail(FA_ERR_RUNTIME, "You must set num_neighbors in the query arguments when performing number "
                                        "of neighbor queries.");
        if (a.r_max == -1.0f)
            a.r_max = std::numeric_limits<float>::infinity();
        if (nq->kind == FA_NQ_AABB)
        {
            if (a.scale == -1.0f)
                a.scale = 1.1f;
            else if (a.scale <= 1.0f)
                return fail(FA_ERR_RUNTIME, "The scale query argument must be greater than 1.");
            if (a.r_guess == -1.0f)
            {
                float const vol = box_volume(nq->box);
                float const r_guess = std::cbrt((3.0f * float(a.num_neighbors) * vol)
                                                / (4.0f * float(M_PI) * float(nq->n)));
                float npd[3];
                nearest_plane_distance(nq->box, npd);
                float mpd = std::min(npd[0], npd[1]);
                if (!nq->box.is2D)
                    mpd = std::min(mpd, npd[2]);
                a.r_guess = std::min(r_guess, mpd / 2.0f);
            }
            a.r_guess = std::min(a.r_guess, a.r_max);
        }
    }
    else
        return fail(FA_ERR_RUNTIME, "Unknown mode");
    if (nq->kind == FA_NQ_CELLQUERY)
    {
        // CellQuery::validateQueryArgs (CellQuery.h:175-204), run by query()
        // whatever the number of query points
        if (a.mode == FA_MODE_NEAREST)
            return fail(FA_ERR_RUNTIME, "CellQuery only supports ball queries (r_max), not nearest queries "
                                        "(num_neighbors). Use AABBQuery for nearest neighbor queries.");
        float npd[3];
        nearest_plane_distance(nq->box, npd);
        if (npd[0] <= a.r_max * 2.0f || npd[1] <= a.r_max * 2.0f || (!nq->box.is2D && npd[2] <= a.r_max * 2.0f))
            return fail(FA_ERR_RUNTIME, "The CellQuery r_max is too large for this box.");
        if (a.r_max <= 0)
            return fail(FA_ERR_INVALID, "r_max must be positive.");
        if (a.r_max <= a.r_min)
            return fail(FA_ERR_INVALID, "r_max must be greater than r_min.");
    }
    if (n_query == 0)
        return FA_OK;
    if (a.r_max <= 0)
        return fail(FA_ERR_INVALID, "NeighborQuery requires r_max to be positive.");
    if (a.r_max <= a.r_min)
        return fail(FA_ERR_INVALID, "NeighborQuery requires that r_max must be greater than r_min.");
    if (nq->kind == FA_NQ_AABB && a.mode == FA_MODE_BALL)
    {
        float npd[3];
        nearest_plane_distance(nq->box, npd);
        if (npd[0] <= a.r_max * 2.0 || npd[1] <= a.r_max * 2.0 || (!nq->box.is2D && npd[2] <= a.r_max * 2.0))
            return fail(FA_ERR_RUNTIME, "The AABBQuery r_max is too large for this box.");
    }
    return FA_OK;
}

} // namespace fa

namespace {

// The smallest float x >= 0 with sqrtf(x) >= r (correctly rounded square
// root, monotone in x), by bisection over float bit patterns; 0 for r <= 0.
float sqrt_threshold(float r)
{
    if (!(r > 0.0f))
        return 0.0f;
    uint32_t lo = 0, hi = 0x7f800000u;
    while (lo < hi)
    {
        uint32_t const mid = lo + (hi - lo) / 2;
        volatile float const x = __builtin_bit_cast(float, mid);
        volatile float const q = std::sqrt(float(x));
        if (float(q) >= r)
            hi = mid;
        else
            lo = mid + 1;
    }
    return __builtin_bit_cast(float, lo);
}

// Staged-search parameters for k neighbours at number density dens: the
// stage radius holds ~k + sqrt(k) + 1 points on average (regular points, for
// which the staged search is chosen, rarely have fewer than k inside it; each
// staged point beyond k costs a selection step) and the stage ~2 standard
// deviations more (at most 32 entries of 4 bytes per lane).
struct KnnStage
{
    float thr = 0.0f; // smallest r_sq whose sqrtf reaches the stage radius
    float r = 0.0f;   // the stage radius
    uint32_t C = 0;   // 0: no staged search (k > 16)
};
KnnStage knn_stage_params(bool is2D, double dens, uint32_t k)
{
    KnnStage st;
    if (k > 16 || !(dens > 0.0))
        return st;
    double const lam = k + std::sqrt(double(k)) + 1.0;
    double const r = is2D ? std::sqrt(lam / (M_PI * dens)) : std::cbrt(3.0 * lam / (4.0 * M_PI * dens));
    uint32_t C = uint32_t(std::ceil((lam + 2.0 * std::sqrt(lam)) / 4.0)) * 4;
    C = std::min<uint32_t>(32, std::max<uint32_t>(C, k + 4));
    st.r = float(r);
    st.thr = sqrt_threshold(st.r);
    st.C = C;
    return st;
}

template<int KMAX>
int launch_knn_aabb(const fa::DevBox& db, const DevGrid& g, const float* qpts, uint32_t nq, uint32_t k, float rlim,
                    float r_min, int exclude_ii, float wmin, float hmp, int rmax_ring, const KnnStage& st,
                    uint32_t* counts, uint64_t* sel, uint32_t* flagged, uint32_t* nflag, uint32_t* refine,
                    uint32_t* nrefine, const float* self_points, uint32_t self_n, uint32_t* pos)
{
    float const rlim_sq = rlim * rlim;
    float const rmin_thr = sqrt_threshold(r_min);
    bool const self = (qpts == self_points) && (nq == self_n);
    unsigned const blocks = (nq + kBlock - 1) / kBlock;
    if (st.C == 0)
    {
        FA_PROF_BEGIN("knn_aabb_kernel");
        if (self)
            hipLaunchKernelGGL((knn_aabb_kernel<KMAX, true, false>), dim3(blocks), dim3(kBlock), 0, fa::stream(), db,
                               g, qpts, nq, k, rlim_sq, rlim, rmin_thr, exclude_ii, wmin, hmp, rmax_ring, counts, sel,
                               flagged, nflag, pos, nullptr, nullptr);
        else
            hipLaunchKernelGGL((knn_aabb_kernel<KMAX, false, false>), dim3(blocks), dim3(kBlock), 0, fa::stream(),
                               db, g, qpts, nq, k, rlim_sq, rlim, rmin_thr, exclude_ii, wmin, hmp, rmax_ring, counts,
                               sel, flagged, nflag, pos, nullptr, nullptr);
        FA_PROF_END();
        FA_HIP(hipGetLastError());
        return FA_OK;
    }
    // the stage, and afterwards the block's selection rows
    size_t const lds = std::max(size_t(st.C) * kBlock * sizeof(uint32_t), size_t(kBlock) * k * sizeof(uint64_t));
    FA_PROF_BEGIN("knn_stage_kernel");
    if (self)
        hipLaunchKernelGGL((knn_stage_kernel<KMAX, true>), dim3(blocks), dim3(kBlock), lds, fa::stream(), db, g, qpts,
                           nq, k, rlim_sq, rlim, rmin_thr, exclude_ii, wmin, hmp, rmax_ring, st.thr, st.r, st.C, counts, sel,
                           flagged, nflag, refine, nrefine, pos);
    else
        hipLaunchKernelGGL((knn_stage_kernel<KMAX, false>), dim3(blocks), dim3(kBlock), lds, fa::stream(), db, g,
                           qpts, nq, k, rlim_sq, rlim, rmin_thr, exclude_ii, wmin, hmp, rmax_ring, st.thr, st.r, st.C,
                           counts, sel, flagged, nflag, refine, nrefine, pos);
    FA_PROF_END();
    FA_HIP(hipGetLastError());
    // the handed-back points: a quarter of the full grid, grid-stride over
    // the list (its length stays on the device)
    unsigned const rblocks = std::max(1u, (blocks + 3) / 4);
    FA_PROF_BEGIN("knn_refine_kernel");
    if (self)
        hipLaunchKernelGGL((knn_aabb_kernel<KMAX, true, true>), dim3(rblocks), dim3(kBlock), 0, fa::stream(), db, g,
                           qpts, nq, k, rlim_sq, rlim, rmin_thr, exclude_ii, wmin, hmp, rmax_ring, counts, sel,
                           flagged, nflag, pos, refine, nrefine);
    else
        hipLaunchKernelGGL((knn_aabb_kernel<KMAX, false, true>), dim3(rblocks), dim3(kBlock), 0, fa::stream(), db, g,
                           qpts, nq, k, rlim_sq, rlim, rmin_thr, exclude_ii, wmin, hmp, rmax_ring, counts, sel,
                           flagged, nflag, pos, refine, nrefine);
    FA_PROF_END();
    FA_HIP(hipGetLastError());
    return FA_OK;
}

template<int KMAX>
int launch_knn_lc(const fa::DevBox& db, const DevGrid& g, const float* qpts, uint32_t nq, uint32_t k, float r_max,
                  float r_min, int exclude_ii, int R, uint32_t* counts, uint64_t* sel)
{
    FA_KLAUNCH(knn_linkcell_kernel<KMAX>, dim3((nq + kBlock - 1) / kBlock), dim3(kBlock), 0, fa::stream(),
                       db, g, qpts, nq, k, r_max * r_max, r_min * r_min, exclude_ii, R, counts, sel);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

// Keys kept per point by the one-walk ball query: the bonds a point of a
// uniform system at this density has, with slack; points beyond it are walked
// again by the emission pass.  When that exceeds 32 (a 32 KB LDS stage per
// block) the walk only counts (0: every point is walked again, in processing
// order, as a count pass and a fill pass would).
uint32_t ball_row_slots(const fa_nq* nq, float r_max)
{
    double const vol = std::max(double(fa::box_volume(nq->box)), 1e-30);
    double const dens = double(nq->n) / vol;
    double const r = double(r_max);
    double const expect = nq->box.is2D ? dens * M_PI * r * r : dens * (4.0 / 3.0) * M_PI * r * r * r;
    double const c = std::ceil((expect * 1.15 + 4.0) / 4.0) * 4.0;
    if (c > double(kWalkMaxSlots))
        return 0;
    return uint32_t(std::max(c, 8.0));
}

int run_ball(fa_nq* nq, const float* qpts, uint32_t n_query, const fa_query_args& a, int by_distance, fa_nlist* out)
{
    hipStream_t s = fa::stream();
    int const R = 1;
    int dims[3];
    fa::grid_dims_for(nq->box, nq->n, a.r_max, R, dims);
    if (int rc = fa::ensure_grid(nq, dims))
        return rc;
    DevGrid const g = nq->cur->dev();
    fa::DevBox const& db = nq->cur->dbox;
    if (int rc = fa::nlist_reserve(out, out->capacity, n_query))
        return rc;
    float const r_max_sq = a.r_max * a.r_max, r_min_sq = a.r_min * a.r_min;
    unsigned const blocks = (n_query + kBlock - 1) / kBlock;
    // query points == the grid's points: the walk takes them in cell order
    bool const self = (qpts == nq->d_points) && (n_query == nq->n);
    fa::CQGrid const cq = nq->kind == FA_NQ_CELLQUERY ? fa::cq_grid_for(nq->box, a.r_max) : fa::CQGrid{};
    uint32_t const C = ball_row_slots(nq, a.r_max);
    uint32_t* rows = nullptr;
    uint32_t* pos = nullptr;
    if (n_query)
    {
        // each block owns kBlock * C slots; positions are 32-bit
        size_t const slots = size_t(blocks) * kBlock * C;
        if (slots >= (size_t(1) << 32))
            return fa::fail(FA_ERR_RUNTIME, "ball query: too many query points for the key staging");
        void* tmp = nullptr;
        if (int rc = fa::scratch(21, sizeof(uint32_t) * (slots + size_t(n_query)) + 256, &tmp))
            return rc;
        rows = static_cast<uint32_t*>(tmp);
        pos = rows + slots;
    }
    size_t const walk_lds = sizeof(uint32_t) * size_t(kBlock) * C;
    auto walk = [&](auto kern) {
        FA_PROF_BEGIN("ball_walk_kernel");
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), walk_lds, s, db, g, qpts, n_query, r_max_sq, r_min_sq,
                           a.exclude_ii, R, cq, C, out->counts, pos, rows);
        FA_PROF_END();
    };
    if (n_query == 0)
        ;
    else if (nq->kind == FA_NQ_AABB)
        self ? walk(ball_walk_kernel<AABB_FORMULA, true>) : walk(ball_walk_kernel<AABB_FORMULA, false>);
    else if (nq->kind == FA_NQ_CELLQUERY)
        self ? walk(ball_walk_kernel<CELLQUERY_FORMULA, true>) : walk(ball_walk_kernel<CELLQUERY_FORMULA, false>);
    else
        self ? walk(ball_walk_kernel<LINKCELL_FORMULA, true>) : walk(ball_walk_kernel<LINKCELL_FORMULA, false>);
    FA_HIP(hipGetLastError());
    uint64_t total = 0;
    if (int rc = fa::exclusive_scan_u32(out->counts, out->offsets, n_query, &total))
        return rc;
    if (total > 0xffffffffull)
        return fa::fail(FA_ERR_RUNTIME, "NeighborList would exceed 2^32 bonds");
    if (int rc = fa::nlist_reserve(out, total, n_query))
        return rc;
    out->num_bonds = total;
    void* gkeys = nullptr;
    if (int rc = fa::scratch(3, std::max<uint64_t>(total, 1) * (by_distance ? 8 : 4), &gkeys))
        return rc;
    size_t const kb = fill_key_bytes(total, n_query, by_distance ? 8 : 4);
    uint32_t const key_cap = uint32_t(kb / (by_distance ? 8 : 4));
    auto emit = [&](auto kern) {
        FA_PROF_BEGIN("ball_emit_kernel");
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), kb, s, db, g, qpts, n_query, nq->d_points, r_max_sq,
                           r_min_sq, a.exclude_ii, R, cq, C, pos, rows, out->offsets, out->neighbors, out->distances,
                           out->weights, out->vectors, gkeys, key_cap);
        FA_PROF_END();
    };
    // self queries of points stored in random order emit in processing order,
    // and so does a walk that kept no keys (the emission pass walks again)
    bool index_order = C > 0;
    if (self && n_query && index_order)
    {
        const uint32_t* sorder = nullptr;
        if (int rc = fa::spatial_order(nq->box, nq->d_points, nq->n, &sorder))
            return rc;
        index_order = sorder == nullptr;
    }
#define FA_BALL_EMIT(FORM)                                                                                       \
    do                                                                                                           \
    {                                                                                                            \
        if (!self)                                                                                               \
            by_distance ? emit(ball_emit_kernel<FORM, true, false, true>)                                        \
                        : emit(ball_emit_kernel<FORM, false, false, true>);                                      \
        else if (index_order)                                                                                    \
            by_distance ? emit(ball_emit_kernel<FORM, true, true, true>)                                         \
                        : emit(ball_emit_kernel<FORM, false, true, true>);                                       \
        else                                                                                                     \
            by_distance ? emit(ball_emit_kernel<FORM, true, true, false>)                                        \
                        : emit(ball_emit_kernel<FORM, false, true, false>);                                      \
    } while (0)
    if (n_query)
    {
        if (nq->kind == FA_NQ_AABB)
            FA_BALL_EMIT(AABB_FORMULA);
        else if (nq->kind == FA_NQ_CELLQUERY)
            FA_BALL_EMIT(CELLQUERY_FORMULA);
        else
            FA_BALL_EMIT(LINKCELL_FORMULA);
        FA_HIP(hipGetLastError());
    }
#undef FA_BALL_EMIT
    return fa::nlist_finish_counts(out);
}

// Radius of the last ball of the reference's expanding kNN search
// (AABBQuery.cc:174-263): r_cur starts at r_guess and is multiplied by scale
// until it reaches r_max or the largest plane distance.
float knn_radius_limit(const fa_nq* nq, const fa_query_args& a)
{
    float npd[3];
    fa::nearest_plane_distance(nq->box, npd);
    float max_pd = std::max(npd[0], npd[1]);
    if (!nq->box.is2D)
        max_pd = std::max(max_pd, npd[2]);
    float r_cur = a.r_guess;
    float last = std::min(r_cur, a.r_max);
    for (int it = 0; it < 100000; ++it)
    {
        last = std::min(r_cur, a.r_max);
        r_cur *= a.scale;
        if (r_cur >= a.r_max || r_cur >= max_pd)
            break;
    }
    return last;
}

// The exhaustive kNN kernel costs (images + k) candidate passes over all
// points per query: num_neighbors > 64 sends every query there, and so does
// an answer reaching half the smallest plane distance.  Work beyond this many
// candidate evaluations (a few seconds) is refused with an error instead of
// being started.
constexpr double kExhaustiveWorkLimit = 68719476736.0; // 2^36

int exhaustive_work_check(const fa_nq* nq, uint32_t n_exhaustive, uint32_t k)
{
    double const images = nq->box.is2D ? 9.0 : 27.0;
    double const work = double(n_exhaustive) * double(nq->n) * (images + double(k));
    if (work > kExhaustiveWorkLimit)
        return fa::fail(FA_ERR_RUNTIME,
                        "nearest-neighbour query: " + std::to_string(n_exhaustive) +
                            " query points need the exhaustive search (num_neighbors > 64, or neighbours beyond "
                            "half the box), " + std::to_string(n_exhaustive) + " x " + std::to_string(nq->n) +
                            " points is beyond its limit; use num_neighbors <= 64 or fewer points");
    return FA_OK;
}

int run_knn(fa_nq* nq, const float* qpts, uint32_t n_query, const fa_query_args& a, int by_distance, fa_nlist* out)
{
    hipStream_t s = fa::stream();
    uint32_t const k = a.num_neighbors;
    if (int rc = fa::nlist_reserve(out, out->capacity, n_query))
        return rc;
    if (n_query == 0 || k == 0)
    {
        FA_HIP(hipMemsetAsync(out->offsets, 0, sizeof(uint32_t) * (size_t(n_query) + 1), s));
        out->num_bonds = 0;
        return fa::nlist_finish_counts(out);
    }
    // scratch: selections (nq*k u64, point-major by processing position),
    // write temp (nq*k u64), the flagged and handed-back counts and lists,
    // processing position per point
    void* tmp = nullptr;
    uint64_t const selb = uint64_t(n_query) * k * sizeof(uint64_t);
    if (int rc = fa::scratch(4, 2 * selb + sizeof(uint32_t) * (3 * size_t(n_query) + 128), &tmp))
        return rc;
    uint64_t* sel = static_cast<uint64_t*>(tmp);
    uint64_t* wtmp = sel + uint64_t(n_query) * k;
    uint32_t* nflag = reinterpret_cast<uint32_t*>(wtmp + uint64_t(n_query) * k);
    uint32_t* nrefine = nflag + 1;
    uint32_t* flagged = nflag + 16;
    uint32_t* pos = flagged + n_query + 48;
    uint32_t* refine = pos + n_query + 16;
    FA_HIP(hipMemsetAsync(nflag, 0, 2 * sizeof(uint32_t), s));
    float npd[3];
    fa::nearest_plane_distance(nq->box, npd);
    float min_pd = std::min(npd[0], npd[1]);
    if (!nq->box.is2D)
        min_pd = std::min(min_pd, npd[2]);
    const fa::DevBox* dbp = nullptr;
    uint32_t h_nflag = 0;
    float rlim = 0.0f;
    // the tuner entry behind a staged search (revised after it, see below)
    int staged_analysis = 0, staged_key = 0, staged_variant = 0;
    if (nq->kind == FA_NQ_AABB)
    {
        rlim = knn_radius_limit(nq, a);
        // cells a fraction of the radius rk that holds ~2k points; the
        // autotuner picks the fraction (rk, rk/1.2, rk/1.5, rk/2): smaller
        // cells mean fewer candidates but more stencil cells per query
        float const vol = fa::box_volume(nq->box);
        double const dens = double(nq->n) / std::max(double(vol), 1e-30);
        double const rk =
            nq->box.is2D ? std::sqrt(2.0 * k / (M_PI * dens)) : std::cbrt(3.0 * 2.0 * k / (4.0 * M_PI * dens));
        static constexpr double kFrac[fa::kKnnCellSizes] = {1.0, 1.0 / 1.2, 1.0 / 1.5, 0.5};
        // the staged search (k <= 16) against the ring search: the tuner times
        // both, as the better one depends on how regular the points are
        KnnStage const stage_params = knn_stage_params(nq->box.is2D, dens, k);
        auto run_dims = [&](const int dims[3], bool staged) -> int {
            KnnStage const stage = staged ? stage_params : KnnStage{};
            if (int rc = fa::ensure_grid(nq, dims))
                return rc;
            dbp = &nq->cur->dbox;
            float wmin = std::numeric_limits<float>::infinity();
            for (int d = 0; d < (nq->box.is2D ? 2 : 3); ++d)
                wmin = std::min(wmin, npd[d] / float(dims[d]));
            int const rmax_ring = std::max(dims[0], std::max(dims[1], dims[2])) * 2 + 2;
            FA_HIP(hipMemsetAsync(nflag, 0, 2 * sizeof(uint32_t), s));
#define FA_KNN_LAUNCH(KM)                                                                                     \
    launch_knn_aabb<KM>(*dbp, nq->cur->dev(), qpts, n_query, k, rlim, a.r_min, a.exclude_ii, wmin, min_pd / 2.0f, \
                        rmax_ring, stage, out->counts, sel, flagged, nflag, refine, nrefine, nq->d_points, nq->n, pos)
            if (k <= 8)
                return FA_KNN_LAUNCH(8);
            if (k <= 12)
                return FA_KNN_LAUNCH(12);
            if (k <= 16)
                return FA_KNN_LAUNCH(16);
            if (k <= 32)
                return FA_KNN_LAUNCH(32);
            return FA_KNN_LAUNCH(64);
#undef FA_KNN_LAUNCH
        };
        auto run_variant = [&](int v) -> int {
            float const wcell = float(std::max(rk * kFrac[v % fa::kKnnCellSizes], 1e-6));
            int dims[3];
            fa::grid_dims_for(nq->box, nq->n, wcell, 1, dims);
            return run_dims(dims, v >= fa::kKnnCellSizes);
        };
        // variants the tuner tries: the staged ones only where they exist
        int const n_variants = stage_params.C ? fa::kKnnVariants : fa::kKnnCellSizes;
        // A grid that another query on these points has built serves when its
        // cells lie within the range of sizes the tuner chooses from (rk / 2 to
        // rk, with 25 % slack above): every cell size selects the same
        // neighbours, and the grid build is saved.
        int reuse = -1;
        for (int gs = 0; gs < fa_nq::kGrids && reuse < 0; ++gs)
        {
            if (!nq->grid_valid[gs])
                continue;
            bool fits = true;
            for (int d = 0; d < (nq->box.is2D ? 2 : 3); ++d)
            {
                double const w = double(npd[d]) / double(nq->grid_dims[gs][d]);
                fits = fits && w >= 0.5 * rk && w <= 1.25 * rk;
            }
            if (fits)
                reuse = gs;
        }
        int rc = FA_OK;
        int const key = fa::tune_key(dens, rk) * 128 + int(std::min<uint32_t>(k, 127u));
        if (k <= 64 && reuse >= 0)
        {
            int const dims[3] = {nq->grid_dims[reuse][0], nq->grid_dims[reuse][1], nq->grid_dims[reuse][2]};
            int mode = stage_params.C ? fa::tuned_variant(fa::TUNE_KNN_REUSE, key) : 0;
            if (mode < 0 && fa::tuning_enabled())
            {
                float best = 0.0f;
                for (int v = 0; v < 2; ++v)
                {
                    float ms = 0.0f;
                    if (int rc2 = run_dims(dims, v == 1))
                        return rc2;
                    if (int rc2 = fa::time_launch([&] { return run_dims(dims, v == 1); }, &ms))
                        return rc2;
                    if (mode < 0 || ms < best)
                    {
                        best = ms;
                        mode = v;
                    }
                }
                fa::record_variant(fa::TUNE_KNN_REUSE, key, mode);
            }
            rc = run_dims(dims, mode == 1);
            if (mode == 1 && stage_params.C)
            {
                staged_analysis = fa::TUNE_KNN_REUSE;
                staged_key = key;
                staged_variant = 0;
            }
        }
        else if (k <= 64)
        {
            int variant = fa::tuned_variant(fa::TUNE_KNN, key);
            if (variant < 0 && fa::tuning_enabled())
            {
                float best = 0.0f;
                for (int v = 0; v < n_variants; ++v)
                {
                    float ms = 0.0f;
                    if (int rc2 = run_variant(v)) // builds the grid
                        return rc2;
                    if (int rc2 = fa::time_launch([&] { return run_variant(v); }, &ms))
                        return rc2;
                    if (variant < 0 || ms < best)
                    {
                        best = ms;
                        variant = v;
                    }
                }
                fa::record_variant(fa::TUNE_KNN, key, variant);
            }
            int const v = std::max(variant, 0) % fa::kKnnVariants;
            rc = run_variant(v);
            if (v >= fa::kKnnCellSizes && stage_params.C)
            {
                staged_analysis = fa::TUNE_KNN;
                staged_key = key;
                staged_variant = v - fa::kKnnCellSizes;
            }
        }
        else
        {
            // every query takes the exhaustive path; refuse work that would
            // run for minutes (the check is repeated below for flagged queries)
            if (int rc2 = exhaustive_work_check(nq, n_query, k))
                return rc2;
            int dims[3];
            fa::grid_dims_for(nq->box, nq->n, float(std::max(rk, 1e-6)), 1, dims);
            if (int rc2 = fa::ensure_grid(nq, dims))
                return rc2;
            dbp = &nq->cur->dbox;
            std::vector<uint32_t> all(n_query);
            for (uint32_t t = 0; t < n_query; ++t)
                all[t] = t;
            FA_HIP(hipMemcpyAsync(flagged, all.data(), sizeof(uint32_t) * n_query, hipMemcpyHostToDevice, s));
            FA_HIP(hipMemcpyAsync(nflag, &n_query, sizeof(uint32_t), hipMemcpyHostToDevice, s));
            FA_HIP(hipStreamSynchronize(s));
        }
        if (rc)
            return rc;
    }
    else
    {
        // the reference's LinkCell grid and shell range (LinkCell.cc:584-585)
        int dims[3] = {nq->lc_dim[0], nq->lc_dim[1], nq->lc_dim[2]};
        if (int rc = fa::ensure_grid(nq, dims))
            return rc;
        dbp = &nq->cur->dbox;
        int const max_range = int(std::ceil(min_pd / (2 * nq->cell_width))) + 1;
        int const R = std::max(0, max_range - 1);
        int rc;
        if (k <= 8)
            rc = launch_knn_lc<8>(*dbp, nq->cur->dev(), qpts, n_query, k, a.r_max, a.r_min, a.exclude_ii, R,
                                  out->counts, sel);
        else if (k <= 16)
            rc = launch_knn_lc<16>(*dbp, nq->cur->dev(), qpts, n_query, k, a.r_max, a.r_min, a.exclude_ii, R,
                                   out->counts, sel);
        else if (k <= 32)
            rc = launch_knn_lc<32>(*dbp, nq->cur->dev(), qpts, n_query, k, a.r_max, a.r_min, a.exclude_ii, R,
                                   out->counts, sel);
        else if (k <= 64)
            rc = launch_knn_lc<64>(*dbp, nq->cur->dev(), qpts, n_query, k, a.r_max, a.r_min, a.exclude_ii, R,
                                   out->counts, sel);
        else
            return fa::fail(FA_ERR_RUNTIME, "LinkCell nearest-neighbour queries support num_neighbors <= 64");
        if (rc)
            return rc;
    }
    // The number of queries flagged for the exhaustive kernel comes back with
    // the bond total (one synchronisation when none is flagged).  Flagged
    // queries hold count 0 until the exhaustive kernel has run; the offsets
    // are then scanned again.
    uint64_t total = 0;
    uint32_t h_counts[2] = {0, 0}; // flagged, handed back
    if (int rc = fa::exclusive_scan_u32(out->counts, out->offsets, n_query, &total,
                                        nq->kind == FA_NQ_AABB ? nflag : nullptr, h_counts, 2))
        return rc;
    h_nflag = h_counts[0];
    // The staged search pays a serial ring walk for each handed-back point; on
    // points irregular enough to hand back more than 1/256 of the queries (a
    // liquid, say, against a crystal at the same density, which the tuner's key
    // does not tell apart) the ring search is faster, so a tuner choice of the
    // staged search is revised for the next call (a caller's preset is kept).
    if (staged_analysis && h_counts[1] > n_query / 256 && !fa::variant_preset(staged_analysis, staged_key))
        fa::record_variant(staged_analysis, staged_key, staged_variant);
    if (h_nflag)
    {
        if (int rc = exhaustive_work_check(nq, h_nflag, k))
            return rc;
        // bounded work space: at most ~256 MiB of per-query candidate keys at once
        uint64_t const per_q = uint64_t(nq->n) * sizeof(uint64_t);
        uint32_t const chunk = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(4096, (256ull << 20) / per_q)));
        for (uint32_t c0 = 0; c0 < h_nflag; c0 += chunk)
        {
            uint32_t const nc = std::min(chunk, h_nflag - c0);
            void* work = nullptr;
            if (int rc2 = fa::scratch(5, uint64_t(nc) * nq->n * sizeof(uint64_t) + 64, &work))
                return rc2;
            // flagged positions are in the grid's cell order when the queries
            // are the points themselves (the AABB kernel's processing order)
            bool const cell_order = (qpts == nq->d_points) && (n_query == nq->n);
            FA_KLAUNCH(knn_exhaustive_kernel, dim3(nc), dim3(kBlock), 0, s, *dbp, nq->d_points, nq->n, qpts, n_query,
                       cell_order ? nq->cur->cpos : nullptr, flagged + c0, k, rlim * rlim, a.r_min, a.exclude_ii,
                       static_cast<uint64_t*>(work), out->counts, sel, pos);
            FA_HIP(hipGetLastError());
        }
        if (int rc = fa::exclusive_scan_u32(out->counts, out->offsets, n_query, &total))
            return rc;
    }
    if (int rc = fa::nlist_reserve(out, total, n_query))
        return rc;
    out->num_bonds = total;
    unsigned const blocks = (n_query + kBlock - 1) / kBlock;
    // the AABB search takes self queries in cell order (rows at pos[i])
    bool const self = (qpts == nq->d_points) && (n_query == nq->n);
    // at most k keys per query point: the staging holds a whole block's
    size_t const ksz = by_distance ? 8 : 4;
    size_t const kb = std::min<size_t>(size_t(k) * kBlock * ksz, size_t(kKeyBytes));
    uint32_t const key_cap = uint32_t(kb / ksz);
    DevGrid const g = nq->cur->dev();
    auto launch = [&](auto kern) {
        FA_PROF_BEGIN("knn_write_kernel");
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), kb, s, *dbp, g, qpts, n_query, nq->d_points, k, sel,
                           pos, out->offsets, out->neighbors, out->distances, out->weights, out->vectors, wtmp,
                           key_cap);
        FA_PROF_END();
    };
    if (nq->kind == FA_NQ_AABB)
    {
        // self queries of points stored in random order write in processing order
        bool index_order = true;
        if (self)
        {
            const uint32_t* sorder = nullptr;
            if (int rc = fa::spatial_order(nq->box, nq->d_points, nq->n, &sorder))
                return rc;
            index_order = sorder == nullptr;
        }
        if (!self)
            by_distance ? launch(knn_write_kernel<AABB_FORMULA, true, false, true>)
                        : launch(knn_write_kernel<AABB_FORMULA, false, false, true>);
        else if (index_order)
            by_distance ? launch(knn_write_kernel<AABB_FORMULA, true, true, true>)
                        : launch(knn_write_kernel<AABB_FORMULA, false, true, true>);
        else
            by_distance ? launch(knn_write_kernel<AABB_FORMULA, true, true, false>)
                        : launch(knn_write_kernel<AABB_FORMULA, false, true, false>);
    }
    else
    {
        // the LinkCell search runs in index order and stores its selections by
        // point index
        by_distance ? launch(knn_write_kernel<LINKCELL_FORMULA, true, false, true>)
                    : launch(knn_write_kernel<LINKCELL_FORMULA, false, false, true>);
    }
    FA_HIP(hipGetLastError());
    return fa::nlist_finish_counts(out);
}

} // namespace

extern "C" {

int fa_nq_create(const fa_box* box, const float* d_points, uint32_t n, int kind, float cell_width, fa_nq** out)
{
    FA_API_LOCK;
    *out = nullptr;
    if (int rc = fa_init())
        return rc;
    // NeighborQuery.h:97-100
    if (n == 0)
        return fa::fail(FA_ERR_INVALID, "Cannot create a NeighborQuery with 0 particles.");
    if (n >= (1u << 27))
        return fa::fail(FA_ERR_INVALID, "freud_amd supports fewer than 2^27 points per NeighborQuery");
    if (kind != FA_NQ_AABB && kind != FA_NQ_LINKCELL && kind != FA_NQ_CELLQUERY)
        return fa::fail(FA_ERR_INVALID, "unknown NeighborQuery kind");
    fa_nq* nq = new fa_nq();
    nq->box = *box;
    nq->d_points = d_points;
    nq->n = n;
    nq->kind = kind;
    nq->cell_width = cell_width;
    if (kind == FA_NQ_LINKCELL)
    {
        // LinkCell.cc:222-260, 278-314
        if (cell_width == 0.0f)
        {
            unsigned const desired = std::max(n / 10u, 1u);
            nq->cell_width = std::cbrt(fa::box_volume(*box) / float(desired));
        }
        float npd[3];
        fa::nearest_plane_distance(*box, npd);
        float const w = nq->cell_width;
        nq->lc_dim[0] = std::max(1, int((unsigned) (npd[0] / w)));
        nq->lc_dim[1] = std::max(1, int((unsigned) (npd[1] / w)));
        nq->lc_dim[2] = box->is2D ? 1 : std::max(1, int((unsigned) (npd[2] / w)));
        if ((w * 2.0 > npd[0]) || (w * 2.0 > npd[1]) || (!box->is2D && w * 2.0 > npd[2]))
        {
            delete nq;
            return fa::fail(FA_ERR_RUNTIME, "Cannot generate a cell list where cell_width is larger than half the box.");
        }
    }
    *out = nq;
    return FA_OK;
}

int fa_nq_destroy(fa_nq* nq)
{
    FA_API_LOCK;
    if (nq)
    {
        (void) hipStreamSynchronize(fa::stream());
        delete nq;
    }
    return FA_OK;
}

int fa_nq_cell_width(const fa_nq* nq, float* width)
{
    FA_API_LOCK;
    *width = nq->cell_width;
    return FA_OK;
}

int fa_cellquery_grid(const fa_box* box, float r_cut, int32_t* n_xyz, float* min_pos, float* inverse_width)
{
    FA_API_LOCK;
    // CellQuery.cc:60-63
    if (!(r_cut > 0))
        return fa::fail(FA_ERR_RUNTIME, "CellQuery::buildGrid called with invalid r_cut (must be positive).");
    fa::CQGrid const c = fa::cq_grid_for(*box, r_cut);
    for (int k = 0; k < 3; ++k)
    {
        n_xyz[k] = c.n[k];
        min_pos[k] = c.min[k];
    }
    *inverse_width = c.inv;
    return FA_OK;
}

int fa_query(fa_nq* nq, const float* d_query_points, uint32_t n_query, const fa_query_args* args,
             int sort_by_distance, fa_nlist* out)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    fa_query_args a = *args;
    if (int rc = fa::validate_query_args(nq, a, n_query))
        return rc;
    out->num_points = nq->n;
    int rc = (a.mode == FA_MODE_BALL) ? run_ball(nq, d_query_points, n_query, a, sort_by_distance, out)
                                      : run_knn(nq, d_query_points, n_query, a, sort_by_distance, out);
    if (rc)
        return rc;
    FA_HIP(hipStreamSynchronize(fa::stream()));
    return FA_OK;
}

int fa_nlist_free(fa_nlist* nl)
{
    FA_API_LOCK;
    if (!nl)
        return FA_OK;
    (void) hipStreamSynchronize(fa::stream());
    fa::device_free(nl->neighbors);
    fa::device_free(nl->distances);
    fa::device_free(nl->weights);
    fa::device_free(nl->vectors);
    fa::device_free(nl->counts);
    fa::device_free(nl->segments);
    fa::device_free(nl->offsets);
    *nl = fa_nlist{};
    return FA_OK;
}

int fa_nlist_all_pairs(const fa_box* box, const float* d_points, uint32_t n_points, const float* d_query_points,
                       uint32_t n_query, int exclude_ii, fa_nlist* out)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    uint32_t const nii = exclude_ii ? std::min(n_points, n_query) : 0;
    uint64_t const nb = uint64_t(n_points) * n_query - nii;
    if (nb > 0xffffffffull)
        return fa::fail(FA_ERR_RUNTIME, "NeighborList would exceed 2^32 bonds");
    if (int rc = fa::nlist_reserve(out, nb, n_query))
        return rc;
    out->num_bonds = nb;
    out->num_points = n_points;
    fa::DevBox const db = fa::make_devbox(*box);
    uint64_t const total = uint64_t(n_points) * n_query;
    if (total)
    {
        FA_KLAUNCH(all_pairs_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, fa::stream(), db,
                           d_points, n_points, d_query_points, n_query, exclude_ii, out->neighbors, out->distances,
                           out->weights, out->vectors);
        FA_HIP(hipGetLastError());
    }
    // offsets: row i starts at i*np - min(i, np)
    std::vector<uint32_t> off(size_t(n_query) + 1);
    for (uint32_t i = 0; i <= n_query; ++i)
        off[i] = uint32_t(uint64_t(i) * n_points - (exclude_ii ? std::min(i, n_points) : 0));
    FA_HIP(hipMemcpyAsync(out->offsets, off.data(), sizeof(uint32_t) * off.size(), hipMemcpyHostToDevice,
                          fa::stream()));
    if (int rc = fa::nlist_finish_counts(out))
        return rc;
    FA_HIP(hipStreamSynchronize(fa::stream()));
    return FA_OK;
}

} // extern "C"
