// fa_common.h -- shared device/host definitions of the freud_amd HIP library.
//
// All translation units are compiled with -ffp-contract=off and
// -fhip-fp32-correctly-rounded-divide-sqrt: bond vectors, distances and
// histogram bins are reproduced bit for bit from the reference's IEEE-single
// arithmetic, which is formed without FMA and with correctly rounded divide and
// square root (freud/util/VectorMath.h, freud/box/Box.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <mutex>
#include <string>
#include <type_traits>

#include "../../include/freud_amd.h"

namespace fa {

// ------------------------------------------------------------------ errors
// Status codes returned through the C-ABI; the Python layer maps them to the
// exception types the reference raises (nanobind's default translation of
// std::invalid_argument, std::runtime_error, std::domain_error, std::out_of_range).
void set_error(int code, const std::string& msg);
int fail(int code, const std::string& msg);

// Library calls share process-wide state (scratch buffers, the spatial-order
// cache, the pinned read-back buffer, the autotuner table): every C-ABI entry
// point holds this lock while it runs, so calls from several host threads run
// one after another.  Recursive, as an entry point may call another.  GPU work
// stays ordered on the library's one stream.
std::recursive_mutex& api_mutex();
#define FA_API_LOCK std::lock_guard<std::recursive_mutex> fa_api_lock_(::fa::api_mutex())

#define FA_HIP(call)                                                                        \
    do                                                                                      \
    {                                                                                       \
        hipError_t _e = (call);                                                             \
        if (_e != hipSuccess)                                                               \
            return ::fa::fail(FA_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

hipStream_t stream();
// Pinned host memory for small read-backs on the library stream: copies into
// it are queued asynchronously, so one synchronisation covers several.  *h is
// null when `bytes` exceeds its kReadbackBytes (the caller then copies into
// pageable memory).
constexpr size_t kReadbackBytes = 64 * 1024;
int readback_stage(size_t bytes, void** h);
// One device word read back through the pinned buffer (synchronises the stream).
int read_u32(const uint32_t* d, uint32_t* h);

// Per-kernel event timing on the library stream (off unless enabled through
// fa_profile_enable); each FA_PROF_BEGIN/END pair brackets one launch.
bool profiling();
void prof_begin(const char* name);
void prof_end();
#define FA_PROF_BEGIN(name)     \
    if (::fa::profiling())      \
    ::fa::prof_begin(name)
#define FA_PROF_END()           \
    if (::fa::profiling())      \
    ::fa::prof_end()
// Kernel launch on the library stream, timed when profiling is enabled.
#define FA_KLAUNCH(kernel, ...)                    \
    do                                             \
    {                                              \
        FA_PROF_BEGIN(#kernel);                    \
        hipLaunchKernelGGL(kernel, __VA_ARGS__);   \
        FA_PROF_END();                             \
    } while (0)

// Autotuner table (fa_tune.hip).  Analyses that the tuner covers:
enum TunedAnalysis
{
    TUNE_RDF_SELF = 1, // variant = a cell shape (R, F): r_max / R wide, r_max / (R F) along x
    TUNE_BALL = 2,     // variant = cells r_max / R wide, R = 1 + variant (batched ball query)
    TUNE_KNN = 3,      // variant = cells rk, rk / 1.2, rk / 1.5, rk / 2 wide (rk: radius holding ~2k points),
                       // ring search (0-3) or staged search (4-7)
    TUNE_KNN_REUSE = 4, // variant = ring search (0) or staged search (1) on a grid another query built
};
constexpr int kKnnCellSizes = 4;
constexpr int kKnnVariants = 2 * kKnnCellSizes;
// Keys of the batched analyses are offset by this (the single-frame keys are
// small integers around 0).
constexpr int kBatchTuneKey = 1000;
bool tuning_enabled();
int tuned_variant(int analysis, int key); // -1 when not tuned yet
void record_variant(int analysis, int key, int variant);
bool variant_preset(int analysis, int key); // set by the caller (fa_tune_set), not by the tuner
// Number of launch variants that the tuner chooses among.
int rdf_row_variant_count();   // single-frame self RDF
int batch_rdf_shape_count();   // batched self RDF
// Elapsed time of the work `launch` enqueues on the library stream (synchronises).
int time_launch(const std::function<int()>& launch, float* ms);
// Problem class of a query: expected points in an r^3 volume, in half-octaves.
int tune_key(double points_per_volume, double r);

// Device scratch that persists between calls (grown on demand, never shrunk).
int scratch(size_t slot, size_t bytes, void** ptr);
int device_alloc(void** p, size_t bytes);
void device_free(void* p);

// ------------------------------------------------------------------ launch shape
// MI355X dispatches the workgroups of a grid round-robin over its 8 XCDs, each
// with its own L2.  xcd_tile maps a workgroup id to a tile so that every XCD
// works through one contiguous range of tiles: spatially sorted data then stays
// within one L2 instead of being fetched by all eight.  A bijection on
// [0, ntiles) for any ntiles.
constexpr uint32_t kXcds = 8;
__device__ inline uint32_t xcd_tile(uint32_t bid, uint32_t ntiles)
{
    uint32_t const x = bid % kXcds, k = bid / kXcds;
    uint32_t const per = ntiles / kXcds, rem = ntiles % kXcds;
    return x * per + min(x, rem) + k;
}

// ------------------------------------------------------------------ box
// Box with the 27 periodic image vectors precomputed on the host with the
// reference's float expression float(i)*a + float(j)*b + float(k)*c
// (freud/locality/NeighborQuery.h:549-556).  Image (i,j,k) in {-1,0,1}^3 is
// stored at (i+1)*9 + (j+1)*3 + (k+1).
struct DevBox
{
    float L[3];
    float lo[3];
    float xy, xz, yz;
    int is2D;
    float img[27][3];
};

DevBox make_devbox(const fa_box& b);
// freud/box/Box.h:489-497
void nearest_plane_distance(const fa_box& b, float out[3]);
float box_volume(const fa_box& b);

// freud/box/Box.h:243-255
__host__ __device__ inline void make_fractional(const DevBox& b, float vx, float vy, float vz, float& fx, float& fy,
                                                float& fz)
{
    float dx = vx - b.lo[0];
    float dy = vy - b.lo[1];
    float dz = vz - b.lo[2];
    float t1 = b.yz * b.xy;
    float t2 = b.xz - t1;
    float t3 = t2 * vz;
    float t4 = b.xy * vy;
    dx = dx - (t3 + t4);
    dy = dy - b.yz * vz;
    fx = dx / b.L[0];
    fy = dy / b.L[1];
    fz = dz / b.L[2];
    if (b.is2D)
        fz = 0.0f;
}

// freud/box/Box.h:212-222
__host__ __device__ inline void make_absolute(const DevBox& b, float fx, float fy, float fz, float& vx, float& vy,
                                              float& vz)
{
    float ax = b.lo[0] + fx * b.L[0];
    float ay = b.lo[1] + fy * b.L[1];
    float az = b.lo[2] + fz * b.L[2];
    float s = b.xy * ay;
    float u = b.xz * az;
    ax = ax + (s + u);
    ay = ay + b.yz * az;
    if (b.is2D)
        az = 0.0f;
    vx = ax;
    vy = ay;
    vz = az;
}

// freud/util/utils.h:29-32
// fmod(fmod(a, 1) + 1, 1).  For |a| < 1 (the fractional coordinates of any
// vector shorter than half the box) fmod(a, 1) == a, and t = a + 1 lies in
// [0, 2], where fmod(t, 1) == t - floor(t) exactly (Sterbenz): the same bits
// without the two library fmods.
__host__ __device__ inline float modulus_positive(float a)
{
    if (fabsf(a) < 1.0f)
    {
        float const t = a + 1.0f;
        return t - floorf(t);
    }
    return fmodf(fmodf(a, 1.0f) + 1.0f, 1.0f);
}

// freud/box/Box.h:307-329 (every dimension periodic)
__host__ __device__ inline void wrap(const DevBox& b, float vx, float vy, float vz, float& ox, float& oy, float& oz)
{
    float fx, fy, fz;
    make_fractional(b, vx, vy, vz, fx, fy, fz);
    fx = modulus_positive(fx);
    fy = modulus_positive(fy);
    fz = modulus_positive(fz);
    make_absolute(b, fx, fy, fz, ox, oy, oz);
}

// ------------------------------------------------------------------ grid
// Cell grid over fractional coordinates.  Points are stored cell-sorted as
// float4 (x, y, z, bits(j)); cells are x-fastest so the cells of one stencil
// row are contiguous in memory.  `wrapc` packs each point's integer image
// offset floor(f) (10 bits per axis) and is only read when some point lies
// outside the box.  The cell-ordered point arrays carry kCposPad entries past
// the last point, so a kernel may load whole 8-point chunks beyond the end of a
// range (the extra points are masked off) without clamping every address.
constexpr uint32_t kCposPad = 64;

struct DevGrid
{
    const float4* cpos;
    const int* wrapc;
    const uint32_t* cstart; // ncells + 1
    int dim[3];
    const int* outside; // device flag: some point lies outside the box (set by the grid build)
    float cw[3];        // cell extent along each plane normal (nearest plane distance / dim)
};

__device__ __host__ inline int floor_div(int a, int b)
{
    int q = a / b;
    return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

__device__ inline void cell_of(const DevBox& b, const int dim[3], float x, float y, float z, int c[3], int w[3])
{
    float f[3];
    make_fractional(b, x, y, z, f[0], f[1], f[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        float s = floorf(f[k] * float(dim[k]));
        s = fminf(fmaxf(s, -1.0e6f), 1.0e6f); // NaN-free clamp for far-outside points
        int C = int(s);
        int ww = floor_div(C, dim[k]);
        c[k] = C - ww * dim[k];
        w[k] = ww;
    }
}

__device__ inline int pack_wrap(const int w[3])
{
    auto cl = [](int v) { return (v < -511 ? -511 : (v > 511 ? 511 : v)) & 1023; };
    return cl(w[0]) | (cl(w[1]) << 10) | (cl(w[2]) << 20);
}
__device__ inline void unpack_wrap(int p, int w[3])
{
    w[0] = (p << 22) >> 22;
    w[1] = (p << 12) >> 22;
    w[2] = (p << 2) >> 22;
}

enum Formula
{
    AABB_FORMULA = 0,      // r = p_j - (q_i + image)   freud/locality/AABBQuery.cc:93,125
    LINKCELL_FORMULA = 1,  // r = box.wrap(p_j - q_i)   freud/locality/LinkCell.cc:522,616
    CELLQUERY_FORMULA = 2, // r = (p_j + ghost shift) - q_i   freud/locality/CellIterator.h:177-186
};

// The grid a CellQuery query builds (freud/locality/CellQuery.h:136-160
// setupGrid, CellQuery.cc:56-113 buildGrid): cubic cells of width r_cut (the
// query's r_max) over the box's bounding parallelepiped padded by one cell,
// with ghost copies of points within r_cut (in fractional terms) of a face.
// The GPU does not build it: it enumerates candidates on its own cell grid and
// keeps exactly the (point, ghost) pairs this grid would visit.
struct CQGrid
{
    float min[3]; // lowest corner of the grid
    float inv;    // 1 / r_cut
    int n[3];     // cells per axis
    float frc[3]; // r_cut / nearest plane distance (fractional ghost margin)
};

// CellQuery::cell_idx_xyz / getCellIdxSafe (CellQuery.h:53-75): false when the
// point's cell lies outside the grid.
__host__ __device__ inline bool cq_cell(const CQGrid& c, float x, float y, float z, int out[3])
{
    out[0] = int(floorf((x - c.min[0]) * c.inv));
    out[1] = int(floorf((y - c.min[1]) * c.inv));
    out[2] = int(floorf((z - c.min[2]) * c.inv));
    return out[0] >= 0 && out[1] >= 0 && out[2] >= 0 && out[0] < c.n[0] && out[1] < c.n[1] && out[2] < c.n[2];
}

// Box::getLatticeVector (freud/box/Box.h:503-518) as CellQuery.cc:80-83 takes
// them (the third is zero in 2D).
__host__ __device__ inline void cq_lattice(const fa_box& b, float lat[3][3])
{
    lat[0][0] = b.Lx;
    lat[0][1] = 0.0f;
    lat[0][2] = 0.0f;
    lat[1][0] = b.Ly * b.xy;
    lat[1][1] = b.Ly;
    lat[1][2] = 0.0f;
    if (b.is2D)
    {
        lat[2][0] = lat[2][1] = lat[2][2] = 0.0f;
    }
    else
    {
        lat[2][0] = b.Lz * b.xz;
        lat[2][1] = b.Lz * b.yz;
        lat[2][2] = b.Lz;
    }
}

// The ghost displacement for shift s (components in {-1, 0, 1}), summed in
// CellQuery.h:243-258's order: +-a, then +-b, then +-c, onto zero.
__host__ __device__ inline void cq_shift(const float lat[3][3], int sx, int sy, int sz, float& dx, float& dy, float& dz)
{
    dx = 0.0f;
    dy = 0.0f;
    dz = 0.0f;
    int const sv[3] = {sx, sy, sz};
    for (int a = 0; a < 3; ++a)
    {
        if (sv[a] == 0)
            continue;
        float const sg = sv[a] == 1 ? 1.0f : -1.0f;
        dx = dx + sg * lat[a][0];
        dy = dy + sg * lat[a][1];
        dz = dz + sg * lat[a][2];
    }
}

// Per-dimension stencil interval [lo, hi] of cell offsets.  For the AABB
// formula every offset is a distinct (cell, image) pair, so the interval is
// always [-R, R].  For the LinkCell formula a cell must be visited once, so an
// interval wider than the grid is clipped to `dim` distinct cells (the
// reference deduplicates with an unordered_set, LinkCell.cc:546-556).
template<int FORMULA> __device__ inline void stencil_interval(int R, int dim, int& lo, int& hi)
{
    if (FORMULA == LINKCELL_FORMULA && 2 * R + 1 >= dim)
    {
        lo = -((dim - 1) / 2);
        hi = dim - 1 + lo;
    }
    else
    {
        lo = -R;
        hi = R;
    }
}

// Calls a candidate visitor; visitors that take a seventh argument also get
// the candidate's cell-ordered record (x, y, z, bits(j)).
template<typename F>
__device__ inline void visit_candidate(F& f, uint32_t j, float rx, float ry, float rz, float r_sq, int code,
                                       const float4& p, uint32_t slot)
{
    if constexpr (std::is_invocable_v<F&, uint32_t, float, float, float, float, int, const float4&, uint32_t>)
        f(j, rx, ry, rz, r_sq, code, p, slot);
    else if constexpr (std::is_invocable_v<F&, uint32_t, float, float, float, float, int, const float4&>)
        f(j, rx, ry, rz, r_sq, code, p);
    else
        f(j, rx, ry, rz, r_sq, code);
}

template<int FORMULA, typename F>
__device__ inline void visit_piece(const DevBox& b, const DevGrid& g, float qx, float qy, float qzz, float qz,
                                   const int wq[3], int rowbase, int ux, int ux_end, int sy, int sz, F&& f)
{
    // cells ux..ux_end (unwrapped x indices) share one x image shift
    int const sx = floor_div(ux, g.dim[0]);
    int const cx = ux - sx * g.dim[0];
    int const cxe = cx + (ux_end - ux);
    uint32_t const beg = g.cstart[rowbase + cx];
    uint32_t const end = g.cstart[rowbase + cxe + 1];
    if (FORMULA == AABB_FORMULA)
    {
        int const nb0 = -wq[0] - sx, nb1 = -wq[1] - sy, nb2 = -wq[2] - sz;
        if (!*g.outside)
        {
            if (nb0 < -1 || nb0 > 1 || nb1 < -1 || nb1 > 1 || nb2 < -1 || nb2 > 1)
                return;
            int const code = (nb0 + 1) * 9 + (nb1 + 1) * 3 + (nb2 + 1);
            float const ix = qx + b.img[code][0];
            float const iy = qy + b.img[code][1];
            float const iz = qzz + b.img[code][2];
            if (beg >= end)
                return;
            // the next candidate is loaded before the current one is used (the
            // point arrays are padded, so cpos[end] is readable)
            float4 p = g.cpos[beg];
            for (uint32_t t = beg; t < end; ++t)
            {
                float4 const pn = g.cpos[t + 1];
                float const pz = b.is2D ? 0.0f : p.z;
                float const rx = p.x - ix;
                float const ry = p.y - iy;
                float const rz = pz - iz;
                float const r_sq = (rx * rx + ry * ry) + rz * rz;
                visit_candidate(f, __float_as_uint(p.w), rx, ry, rz, r_sq, code, p, t);
                p = pn;
            }
        }
        else
        {
            for (uint32_t t = beg; t < end; ++t)
            {
                int wj[3];
                unpack_wrap(g.wrapc[t], wj);
                int const n0 = nb0 + wj[0], n1 = nb1 + wj[1], n2 = b.is2D ? 0 : nb2 + wj[2];
                if (n0 < -1 || n0 > 1 || n1 < -1 || n1 > 1 || n2 < -1 || n2 > 1)
                    continue;
                int const code = (n0 + 1) * 9 + (n1 + 1) * 3 + (n2 + 1);
                float4 const p = g.cpos[t];
                float const pz = b.is2D ? 0.0f : p.z;
                float const rx = p.x - (qx + b.img[code][0]);
                float const ry = p.y - (qy + b.img[code][1]);
                float const rz = pz - (qzz + b.img[code][2]);
                float const r_sq = (rx * rx + ry * ry) + rz * rz;
                visit_candidate(f, __float_as_uint(p.w), rx, ry, rz, r_sq, code, p, t);
            }
        }
    }
    else
    {
        if (beg >= end)
            return;
        float4 p = g.cpos[beg];
        for (uint32_t t = beg; t < end; ++t)
        {
            float4 const pn = g.cpos[t + 1];
            float rx, ry, rz;
            wrap(b, p.x - qx, p.y - qy, p.z - qz, rx, ry, rz);
            float const r_sq = (rx * rx + ry * ry) + rz * rz;
            visit_candidate(f, __float_as_uint(p.w), rx, ry, rz, r_sq, 13, p, t);
            p = pn;
        }
    }
}

// Visit every candidate point of the stencil of half-width R around q, row by
// row (a row's cells are contiguous in memory).  For each candidate calls
//   f(j, rx, ry, rz, r_sq, code)
// with the reference's bond vector for FORMULA; `code` is the image index
// (0..26) used (AABB) or 13 (LinkCell).  `ring_only` restricts the visit to
// cells at Chebyshev offset exactly R (AABB formula only).
template<int FORMULA, typename F>
__device__ inline void for_each_candidate(const DevBox& b, const DevGrid& g, float qx, float qy, float qz, int R,
                                          bool ring_only, F&& f, int RX = -1)
{
    int cq[3], wq[3];
    cell_of(b, g.dim, qx, qy, qz, cq, wq);
    int lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
        stencil_interval<FORMULA>((k == 0 && RX >= 0) ? RX : R, g.dim[k], lo[k], hi[k]);
    if (b.is2D)
        lo[2] = hi[2] = 0;
    float const qzz = (FORMULA == AABB_FORMULA && b.is2D) ? 0.0f : qz;
    for (int dz = lo[2]; dz <= hi[2]; ++dz)
    {
        int const uz = cq[2] + dz;
        int const sz = floor_div(uz, g.dim[2]);
        int const cz = uz - sz * g.dim[2];
        for (int dy = lo[1]; dy <= hi[1]; ++dy)
        {
            int const uy = cq[1] + dy;
            int const sy = floor_div(uy, g.dim[1]);
            int const cy = uy - sy * g.dim[1];
            int const rowbase = (cz * g.dim[1] + cy) * g.dim[0];
            bool const inner = ring_only && R > 0 && (dy > -R && dy < R) && (b.is2D || (dz > -R && dz < R));
            if (inner)
            {
                visit_piece<FORMULA>(b, g, qx, qy, qzz, qz, wq, rowbase, cq[0] - R, cq[0] - R, sy, sz, f);
                visit_piece<FORMULA>(b, g, qx, qy, qzz, qz, wq, rowbase, cq[0] + R, cq[0] + R, sy, sz, f);
                continue;
            }
            int ux = cq[0] + lo[0];
            int const uxend = cq[0] + hi[0];
            while (ux <= uxend)
            {
                int const sx = floor_div(ux, g.dim[0]);
                int const piece_end = min(uxend, (sx + 1) * g.dim[0] - 1);
                visit_piece<FORMULA>(b, g, qx, qy, qzz, qz, wq, rowbase, ux, piece_end, sy, sz, f);
                ux = piece_end + 1;
            }
        }
    }
}

// Recompute the bond vector of (q, p_j) for an image code (AABB) or by wrap
// (LinkCell); identical arithmetic to for_each_candidate.
template<int FORMULA>
__device__ inline void bond_vector(const DevBox& b, float qx, float qy, float qz, float px, float py, float pz,
                                   int code, float& rx, float& ry, float& rz)
{
    if (FORMULA == CELLQUERY_FORMULA)
    {
        // ghost p_j + shift (shift = -image of the AABB image code), then minus q
        fa_box const fb{b.L[0], b.L[1], b.L[2], b.xy, b.xz, b.yz, b.is2D};
        float lat[3][3];
        cq_lattice(fb, lat);
        float dx, dy, dz;
        cq_shift(lat, 1 - code / 9, 1 - (code / 3) % 3, 1 - code % 3, dx, dy, dz);
        // a real point is taken as stored, a ghost as p_j + shift (CellQuery.cc:95-112)
        bool const real = code == 13;
        float const gx = real ? px : px + dx;
        float const gy = real ? py : py + dy;
        float const gz = real ? pz : pz + dz;
        rx = gx - qx;
        ry = gy - qy;
        rz = gz - qz;
        return;
    }
    if (FORMULA == AABB_FORMULA)
    {
        float const qzz = b.is2D ? 0.0f : qz;
        float const pzz = b.is2D ? 0.0f : pz;
        rx = px - (qx + b.img[code][0]);
        ry = py - (qy + b.img[code][1]);
        rz = pzz - (qzz + b.img[code][2]);
    }
    else
    {
        wrap(b, px - qx, py - qy, pz - qz, rx, ry, rz);
    }
}

// CellQuery's view of an AABB-enumerated candidate p (image code `code`) for
// query point q whose CellQuery cell is cq_q (CellIterator.h:94-187): the
// candidate is the point itself (code 13) or its ghost p + shift, shift = minus
// the image; the ghost must be one generateGhosts makes (CellQuery.h:221-272:
// shifts only toward faces within the fractional margin), and the point or
// ghost must lie in the grid, in the 3 x 3 x 3 cells around q's cell.  Then
// r = (ghost) - q and r_sq = (x*x + y*y) + z*z.
__device__ inline bool cq_candidate(const DevBox& b, const CQGrid& c, const int cq_q[3], float qx, float qy,
                                    float qz, const float4& p, int code, float& rx, float& ry, float& rz,
                                    float& r_sq)
{
    float gx = p.x, gy = p.y, gz = p.z;
    if (code != 13)
    {
        int const sx = 1 - code / 9, sy = 1 - (code / 3) % 3, sz = 1 - code % 3;
        float fx, fy, fz;
        make_fractional(b, p.x, p.y, p.z, fx, fy, fz);
        int const dx = int(fx <= c.frc[0]) - int(double(fx) >= 1.0 - double(c.frc[0]));
        int const dy = int(fy <= c.frc[1]) - int(double(fy) >= 1.0 - double(c.frc[1]));
        int const dz = b.is2D ? 0 : int(fz <= c.frc[2]) - int(double(fz) >= 1.0 - double(c.frc[2]));
        if ((sx != 0 && sx != dx) || (sy != 0 && sy != dy) || (sz != 0 && sz != dz))
            return false;
        fa_box const fb{b.L[0], b.L[1], b.L[2], b.xy, b.xz, b.yz, b.is2D};
        float lat[3][3];
        cq_lattice(fb, lat);
        float ddx, ddy, ddz;
        cq_shift(lat, sx, sy, sz, ddx, ddy, ddz);
        gx = p.x + ddx;
        gy = p.y + ddy;
        gz = p.z + ddz;
    }
    int cg[3];
    if (!cq_cell(c, gx, gy, gz, cg))
        return false;
    if (abs(cg[0] - cq_q[0]) > 1 || abs(cg[1] - cq_q[1]) > 1 || abs(cg[2] - cq_q[2]) > 1)
        return false;
    rx = gx - qx;
    ry = gy - qy;
    rz = gz - qz;
    r_sq = (rx * rx + ry * ry) + rz * rz;
    return true;
}

// The bonds of query point q with the arithmetic of FORMULA: f(j, rx, ry, rz,
// r_sq, code) for every candidate of the stencil (the CellQuery formula visits
// the AABB candidates and keeps those its own grid would see; a query point
// outside that grid has no bonds, CellIterator.h:52-57).
template<int FORMULA, typename F>
__device__ inline void visit_bonds(const DevBox& b, const DevGrid& g, const CQGrid& cq, float qx, float qy,
                                   float qz, int R, F&& f)
{
    if constexpr (FORMULA == CELLQUERY_FORMULA)
    {
        int cq_q[3];
        if (!cq_cell(cq, qx, qy, qz, cq_q))
            return;
        for_each_candidate<AABB_FORMULA>(
            b, g, qx, qy, qz, R, false, [&](uint32_t j, float, float, float, float, int code, const float4& p) {
                float rx, ry, rz, r_sq;
                if (cq_candidate(b, cq, cq_q, qx, qy, qz, p, code, rx, ry, rz, r_sq))
                    f(j, rx, ry, rz, r_sq, code);
            });
    }
    else
        for_each_candidate<FORMULA>(b, g, qx, qy, qz, R, false, f);
}

// ------------------------------------------------------------------ host-side objects
struct Grid
{
    fa_box box;
    DevBox dbox;
    int dim[3];
    float width; // requested minimum cell extent along every plane normal
    uint32_t n;
    float4* cpos = nullptr;
    int* wrapc = nullptr;
    uint32_t* cstart = nullptr;
    uint32_t* cell_of_point = nullptr; // scratch
    int* outside = nullptr;            // device flag: some point lies outside the box
    float cw[3] = {0.0f, 0.0f, 0.0f};
    DevGrid dev() const
    {
        DevGrid g;
        g.cpos = cpos;
        g.wrapc = wrapc;
        g.cstart = cstart;
        g.dim[0] = dim[0];
        g.dim[1] = dim[1];
        g.dim[2] = dim[2];
        g.outside = outside;
        g.cw[0] = cw[0];
        g.cw[1] = cw[1];
        g.cw[2] = cw[2];
        return g;
    }
    ~Grid();
};

// Build (or rebuild) a grid over device points with cells at least `width`
// wide along each plane normal (dims clamped to >= 1, z = 1 in 2D).
int build_grid(Grid& g, const fa_box& box, const float* d_points, uint32_t n, const int dims[3]);

// The same build into device buffers the caller provides (nothing is
// allocated; counters are cleared by kernels, not hipMemsetAsync).  Sizes:
// cpos n + kCposPad, wrapc / cell_of_point / rank_of_point / wrap_of_point n,
// cstart ncell + 1, counts ncell, outside one int, scan_tmp
// scan_tmp_bytes(ncell) bytes.
struct GridBuffers
{
    float4* cpos;
    int* wrapc;
    uint32_t* cell_of_point;
    uint32_t* cstart;
    int* outside;
    uint32_t* counts;
    uint32_t* rank_of_point;
    int* wrap_of_point;
    void* scan_tmp;
    size_t scan_bytes;
};
int build_grid_into(const GridBuffers& B, const DevBox& db, const float* d_points, uint32_t n, const int dims[3]);

// Looks up, for the current device, the address of the statically allocated
// scratch array of the no-allocation RDF entry point (called once per device
// by the library's initialisation, with that device current).
int noalloc_setup();

// A processing order of the points that follows space (cells of about four
// points, counting-sorted), for the averaged q_lm, which gathers whole q_lm
// rows of the neighbours: threads that run together then reuse each other's
// rows in cache.  (The bond passes run in index order instead, where a tile's
// bonds are one range of the NeighborList.)  It is only built when a probe
// finds the caller's order scattered (random order); a coherent caller order
// is kept (*order = nullptr, identity), as it also keeps the per-point outputs
// and bond segments streaming.  Any order gives the same per-point results, so
// the decision is cached by (pointer, n, box) and a stale one costs speed
// only.  Identity below kSpatialOrderMin points.
constexpr uint32_t kSpatialOrderMin = 65536;
int spatial_order(const fa_box& box, const float* d_points, uint32_t n, const uint32_t** order);

// Exclusive scan of n uint32 counts into n+1 offsets (offsets[n] = total);
// returns the 64-bit total through *total (host) when total != nullptr.
// d_out[0] = 0, d_out[t + 1] = sum of d_in[0..t]; *total (if given) = the 64-bit
// sum, read back with one synchronisation.  If d_extra is given, the n_extra
// (1 or 2) u32 it points to are copied to h_extra in the same synchronisation.
int exclusive_scan_u32(const uint32_t* d_in, uint32_t* d_out, uint32_t n, uint64_t* total,
                       const uint32_t* d_extra = nullptr, uint32_t* h_extra = nullptr, int n_extra = 1);
// The same scan for code that may be captured into a HIP graph: no memset, no
// host synchronisation; `tmp` holds scan_tmp_bytes(n) bytes.
size_t scan_tmp_bytes(uint32_t n);
int exclusive_scan_u32_async(const uint32_t* d_in, uint32_t* d_out, uint32_t n, void* tmp, size_t tmp_bytes);
// Sets n uint32 words to zero with a kernel (graph-safe, unlike hipMemsetAsync).
int zero_u32(uint32_t* d, uint64_t n);

// ------------------------------------------------------------------ batches
// A batch concatenates the points of many frames (any box shape, any size).
// Each frame is cut into 256-point tiles that never straddle frames, so every
// workgroup of a batched kernel works on one frame and reads that frame's
// description with uniform (scalar) loads.
constexpr uint32_t kTile = 256;

struct BatchFrame
{
    DevBox box;
    int dim[3];         // the frame's cell grid
    uint32_t cell_base; // first cell of the frame in the batch's cell arrays
    uint32_t row_base;  // first x-row (dim[1] * dim[2] rows per frame)
    uint32_t pt_base;   // first point of the frame in the batch
    uint32_t n;         // points in the frame
    uint32_t tile_base; // first tile of the frame
    int any_outside;    // set by the grid build when a point lies outside the box
    float cw[3];        // cell extent along each plane normal
    int rx;             // cells along x that cover the grid's r (R times the x refinement)
};

// One cell grid over every frame of a batch (device arrays).
struct BatchGrid
{
    BatchFrame* frames = nullptr; // n_frames
    float4* cpos = nullptr;       // cell-ordered points, (x, y, z, bits(frame-local index))
    int* wrapc = nullptr;
    uint32_t* cstart = nullptr; // total_cells + 1
    uint32_t* cell_of_point = nullptr;
    uint32_t* rank_of_point = nullptr; // each point's rank among its cell's points
    uint32_t* fill = nullptr;          // total_cells + 1 (cell counts)
    uint32_t n_frames = 0, total_points = 0, total_cells = 0, total_rows = 0;
};

__device__ inline DevGrid frame_grid(const BatchFrame& f, const BatchGrid& G)
{
    DevGrid g;
    g.cpos = G.cpos;
    g.wrapc = G.wrapc;
    g.cstart = G.cstart + f.cell_base;
    g.dim[0] = f.dim[0];
    g.dim[1] = f.dim[1];
    g.dim[2] = f.dim[2];
    g.outside = &f.any_outside;
    g.cw[0] = f.cw[0];
    g.cw[1] = f.cw[1];
    g.cw[2] = f.cw[2];
    return g;
}

// Batched launches (all enqueue on the library stream without synchronising,
// so a batch run can be captured into a HIP graph).
int batch_grid_build(const BatchGrid& G, const uint32_t* d_tile_frame, uint32_t ntiles, const float* d_pts, void* tmp,
                     size_t tmp_bytes);
int batch_ball_count(const BatchGrid& G, const uint32_t* d_tile_frame, uint32_t ntiles, float r_max, int R,
                     uint32_t* d_counts);
// Writes nothing when *d_total (the 64-bit bond total) exceeds capacity.
int batch_ball_fill(const BatchGrid& G, const uint32_t* d_tile_frame, uint32_t ntiles, const float* d_pts, float r_max,
                    int R, const uint32_t* d_offsets, const unsigned long long* d_total, uint64_t capacity,
                    uint32_t* nb, float* dist, float* wts, float* vec, void* gkeys);
// The ball query in one walk: the walk records each point's bond count
// (d_counts), its first C keys packed per tile into d_rows (ntiles * kTile * C
// slots) and where they start (d_pos); the emission (after the counts are
// scanned into d_offsets) sorts and writes them, walking again only points
// past C bonds.  C from batch_ball_slots (0: use count and fill instead).
uint32_t batch_ball_slots(double max_points_per_volume, bool is2D, float r_max);
int batch_ball_walk(const BatchGrid& G, const uint32_t* d_tile_frame, uint32_t ntiles, float r_max, int R, uint32_t C,
                    uint32_t* d_counts, uint32_t* d_pos, uint32_t* d_rows);
int batch_ball_emit(const BatchGrid& G, const uint32_t* d_tile_frame, uint32_t ntiles, const float* d_pts, float r_max,
                    int R, uint32_t C, const uint32_t* d_pos, const uint32_t* d_rows, const uint32_t* d_offsets,
                    const unsigned long long* d_total, uint64_t capacity, uint32_t* nb, float* dist, float* wts,
                    float* vec, void* gkeys);
struct BatchRdfScratch
{
    uint32_t* nclus;  // total_rows + 1
    uint32_t* cl_off; // total_rows + 1
    uint4* table;     // max_clusters
    uint32_t* queue;  // kXcds work-queue counters
    const float* T;   // bins + 1 r_sq thresholds (thresholds_for)
    void* scan_tmp;
    size_t scan_bytes;
    uint64_t max_clusters;
};
int batch_rdf(const BatchGrid& G, int R, float r_max, uint32_t bins, const BatchRdfScratch& S, uint32_t* d_counts);
// The bins + 1 r_sq thresholds of RegularAxis(bins, 0, r_max) binning of
// sqrt(r_sq), computed on the host and copied (synchronously) to d_T.
int thresholds_for(uint32_t bins, float r_max, float* d_T);
// Floats of the bin table thresholds_for writes (thresholds + lookup table).
size_t rdf_table_floats(uint32_t bins);
int batch_steinhardt(const BatchGrid& G, const uint32_t* d_tile_frame, uint32_t ntiles, const float* d_pts,
                     const uint32_t* nb, const float* dist, const uint32_t* d_offsets,
                     const unsigned long long* d_total, uint64_t capacity, uint32_t l, float* d_qlm, float* d_ql,
                     float* d_partial, float* d_order);
int sort_u64(uint64_t* d_keys, uint32_t n);

} // namespace fa

struct fa_nq
{
    fa_box box;
    const float* d_points; // caller-owned device array, N x 3
    uint32_t n;
    int kind;         // FA_NQ_AABB, FA_NQ_LINKCELL or FA_NQ_CELLQUERY
    float cell_width; // LinkCell's cell width (reference semantics)
    int lc_dim[3];    // LinkCell's grid dims (LinkCell.cc:278-314)
    // small cache of grids (one per query shape: e.g. ball query, RDF, kNN);
    // `cur` is the grid selected by the last ensure_grid
    static constexpr int kGrids = 3;
    fa::Grid grids[kGrids];
    int grid_dims[kGrids][3] = {};
    bool grid_valid[kGrids] = {};
    uint64_t grid_used[kGrids] = {};
    uint64_t tick = 0;
    fa::Grid* cur = nullptr;
};

namespace fa {
// Shared by the query, RDF and analysis translation units.
int validate_query_args(const fa_nq* nq, fa_query_args& a, uint32_t n_query);
int ensure_grid(fa_nq* nq, const int dims[3]);
void grid_dims_for(const fa_box& box, uint32_t n, float r, int R, int dims[3]);
// The grid a CellQuery ball query of radius r_cut builds (CellQuery.h:136-160)
CQGrid cq_grid_for(const fa_box& box, float r_cut);
int nlist_reserve(fa_nlist* nl, uint64_t nb, uint32_t nq);
int nlist_finish_counts(fa_nlist* nl);
} // namespace fa
