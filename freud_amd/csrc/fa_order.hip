// fa_order.hip -- Steinhardt q_l / w_l (freud/order/Steinhardt.cc), SolidLiquid
// (freud/order/SolidLiquid.cc), Cluster (freud/cluster/Cluster.cc) and Hexatic
// (freud/order/HexaticTranslational.cc) over a device NeighborList.
//
// Spherical harmonics: the reference evaluates Y_lm with the fsph submodule
// from (theta, phi) = (acos(z/d), atan2(y, x)).  Here they are evaluated
// straight from the bond direction (cos theta = z/d, e^{i phi} = (x + i y)/rho)
// with the normalised associated-Legendre recurrence
//   P(m,m)   = -sqrt((2m+1)/(2m)) sin(t) P(m-1,m-1),   P(0,0) = 1/sqrt(4 pi)
//   P(m+1,m) = sqrt(2m+3) cos(t) P(m,m)
//   P(l,m)   = a(l,m) (cos(t) P(l-1,m) - b(l,m) P(l-2,m))
// (Condon-Shortley phase included, the convention Steinhardt.cc:31-52 stores).
// Only m >= 0 is accumulated: for real bond weights the m < 0 sums follow from
// q_{l,-m} = (-1)^m conj(q_{l,m}).  No transcendental function is evaluated per
// bond.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "fa_common.h"

extern "C" double fa_wigner3j_lll(int l, int m1, int m2, int m3);

namespace {

constexpr int kLMax = 20;
constexpr int kBlock = 256;

struct YlmCoef
{
    float amm[kLMax + 1];                 // -sqrt((2m+1)/(2m)), m >= 1
    float bm1[kLMax + 1];                 // sqrt(2m+3)
    float a[kLMax + 1][kLMax + 1];        // a(l, m)
    float b[kLMax + 1][kLMax + 1];        // b(l, m)
};
// Square root by Newton's iteration from above, for the compile-time table
// (within an ulp of the double result; the host check below confirms that
// every coefficient rounds to the same float as std::sqrt).
constexpr double ce_sqrt(double x)
{
    if (!(x > 0.0))
        return 0.0;
    double g = x > 1.0 ? x : 1.0;
    for (int it = 0; it < 100; ++it)
    {
        double const n = 0.5 * (g + x / g);
        if (!(n < g))
            break;
        g = n;
    }
    return g;
}

template<class Sqrt> constexpr YlmCoef make_ylm_coef(Sqrt sq)
{
    YlmCoef h{};
    for (int m = 0; m <= kLMax; ++m)
    {
        h.amm[m] = m ? float(-sq((2.0 * m + 1.0) / (2.0 * m))) : 0.0f;
        h.bm1[m] = float(sq(2.0 * m + 3.0));
        for (int l = m + 2; l <= kLMax; ++l)
        {
            h.a[l][m] = float(sq((4.0 * l * l - 1.0) / (double(l) * l - double(m) * m)));
            h.b[l][m] = float(sq(((l - 1.0) * (l - 1.0) - double(m) * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0)));
        }
    }
    return h;
}

// The coefficients are compile-time constants: unrolled recurrences carry
// them as instruction literals instead of holding ~400 of them in registers.
constexpr YlmCoef kYlm = make_ylm_coef([](double x) { return ce_sqrt(x); });
bool g_ylm_checked = false;

int ensure_ylm_coef()
{
    if (g_ylm_checked)
        return FA_OK;
    YlmCoef const h = make_ylm_coef([](double x) { return std::sqrt(x); });
    if (std::memcmp(&h, &kYlm, sizeof(h)) != 0)
        return fa::fail(FA_ERR_RUNTIME, "Y_lm recurrence coefficients differ from the correctly rounded values");
    g_ylm_checked = true;
    return FA_OK;
}

// Bond direction in the reference's terms: delta = box.wrap(p_j - p_i)
// (Steinhardt.cc:155), cos(theta) = clamp(delta.z / d, -1, 1) with theta = 0
// for d == 0 (Steinhardt.cc:167-180), e^{i phi} with phi = atan2(delta.y, delta.x).
// (vx, vy, vz) = p_j - p_i before wrapping.
__device__ inline void bond_angles(const fa::DevBox& b, float vx, float vy, float vz, float d, float& ct, float& st,
                                   float& cp, float& sp)
{
    float dx, dy, dz;
    fa::wrap(b, vx, vy, vz, dx, dy, dz);
    if (d == 0.0f)
    {
        ct = 1.0f;
        st = 0.0f;
    }
    else
    {
        ct = fminf(fmaxf(dz / d, -1.0f), 1.0f);
        st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    }
    float const rho2 = dx * dx + dy * dy;
    if (rho2 > 0.0f)
    {
        float const inv = 1.0f / sqrtf(rho2);
        cp = dx * inv;
        sp = dy * inv;
    }
    else
    {
        cp = signbit(dx) ? -1.0f : 1.0f;
        sp = 0.0f;
    }
}

// One associated-Legendre sweep up to LMAX for one bond: for m = 0..LMAX the
// recurrence runs l = m..LMAX and calls sink(l, m, P_lm, cos m phi,
// sin m phi) for every l >= LMIN (every even l >= LMIN when EVEN); l and m are
// constants once the loops are unrolled.  Products and sums here and in the
// sinks are contracted to FMA (the q_lm results are compared within a
// tolerance; the bond geometry before the sweep is evaluated uncontracted).
// Every q_lm path shares this code, so a given (l, m) gets the same float
// operations whatever else is evaluated beside it.
template<int LMAX, int LMIN, bool EVEN, class Sink>
__device__ inline void ylm_sweep(float ct, float st, float cp, float sp, Sink&& sink)
{
#pragma clang fp contract(fast)
    float pmm = 0.28209479177387814f;
    float er = 1.0f, ei = 0.0f;
#pragma unroll
    for (int m = 0; m <= LMAX; ++m)
    {
        if (m > 0)
        {
            pmm = kYlm.amm[m] * st * pmm;
            float const nr = er * cp - ei * sp;
            float const ni = er * sp + ei * cp;
            er = nr;
            ei = ni;
        }
        float p0 = pmm, p1 = 0.0f;
#pragma unroll
        for (int l = m; l <= LMAX; ++l)
        {
            float p;
            if (l == m)
                p = pmm;
            else if (l == m + 1)
            {
                p1 = kYlm.bm1[m] * ct * pmm;
                p = p1;
            }
            else
            {
                p = kYlm.a[l][m] * (ct * p1 - kYlm.b[l][m] * p0);
                p0 = p1;
                p1 = p;
            }
            if (l >= LMIN && (!EVEN || (l & 1) == 0))
                sink(l, m, p, er, ei);
        }
    }
}

// Accumulate w * Y_L^m (m = 0..L) of one bond into acc.
template<int L> __device__ inline void accumulate_ylm(float ct, float st, float cp, float sp, float w, float2* acc)
{
    ylm_sweep<L, L, false>(ct, st, cp, sp, [&](int, int m, float p, float er, float ei) {
#pragma clang fp contract(fast)
        float const wp = w * p;
        acc[m].x += wp * er;
        if (m > 0) // e^{i 0 phi} = 1: the m = 0 imaginary sum stays +0
            acc[m].y += wp * ei;
    });
}

// Where a kernel reads a bond's neighbour index and distance.  A block's
// points are consecutive, so their bonds form one contiguous range of the
// NeighborList, which the block copies into LDS with coalesced loads; each
// thread then walks its own bonds in LDS instead of keeping a private strided
// stream through L2 (thousands of such streams per CU thrash the cache).
// Ranges longer than the staging capacity read global memory.
constexpr uint32_t kStage = 4096;
struct BondSrc
{
    const uint32_t* nbr;
    const float* dist;
    const uint32_t* sj;
    const float* sd;
    uint32_t base;
    bool staged;
    __device__ uint32_t j(uint32_t e) const { return staged ? sj[e - base] : nbr[2 * e + 1]; }
    __device__ float d(uint32_t e) const { return staged ? sd[e - base] : dist[e]; }
};

// Block-uniform: stages the bonds of points [p0, p0 + nloc) (offsets-relative)
// when they fit in cap bonds (cap == 0: never).  Every thread of the block must
// call it.
__device__ inline BondSrc stage_bonds(uint32_t p0, uint32_t nloc, const uint32_t* __restrict__ offsets,
                                      const uint32_t* __restrict__ nbr, const float* __restrict__ dist,
                                      uint32_t* s_j, float* s_d, uint32_t cap)
{
    BondSrc src{nbr, dist, s_j, s_d, 0u, false};
    if (cap == 0 || nloc == 0)
        return src;
    uint32_t const b0 = offsets[p0], b1 = offsets[p0 + nloc];
    if (b1 - b0 > cap)
        return src;
    for (uint32_t t = threadIdx.x; t < b1 - b0; t += blockDim.x)
    {
        s_j[t] = reinterpret_cast<const uint2*>(nbr)[b0 + t].y;
        s_d[t] = dist[b0 + t];
    }
    __syncthreads();
    src.base = b0;
    src.staged = true;
    return src;
}

// stage_bonds for kernels that read neighbour indices only.
__device__ inline BondSrc stage_neighbors(uint32_t p0, uint32_t nloc, const uint32_t* __restrict__ offsets,
                                          const uint32_t* __restrict__ nbr, uint32_t* s_j, uint32_t cap)
{
    BondSrc src{nbr, nullptr, s_j, nullptr, 0u, false};
    if (cap == 0 || nloc == 0)
        return src;
    uint32_t const b0 = offsets[p0], b1 = offsets[p0 + nloc];
    if (b1 - b0 > cap)
        return src;
    for (uint32_t t = threadIdx.x; t < b1 - b0; t += blockDim.x)
        s_j[t] = reinterpret_cast<const uint2*>(nbr)[b0 + t].y;
    __syncthreads();
    src.base = b0;
    src.staged = true;
    return src;
}

// Calls f(ct, st, cp, sp, w) for the bonds e0 <= e < e1 of point i, in order.
// The neighbour position, distance and weight of bond e + 1 and the neighbour
// index of bond e + 2 are loaded before bond e is evaluated, so that the
// gathers of the next bond overlap the arithmetic of this one.
template<class F>
__device__ inline void for_each_bond(const fa::DevBox& b, const float* __restrict__ pts, uint32_t i,
                                     const BondSrc& src, const float* __restrict__ wts, bool weighted, uint32_t e0,
                                     uint32_t e1, F&& f)
{
    if (e0 >= e1)
        return;
    float const xi = pts[3 * i], yi = pts[3 * i + 1], zi = pts[3 * i + 2];
    uint32_t jn = src.j(e0);
    float xj = pts[3 * jn], yj = pts[3 * jn + 1], zj = pts[3 * jn + 2];
    float d = src.d(e0);
    float w = weighted ? wts[e0] : 1.0f;
    if (e0 + 1 < e1)
        jn = src.j(e0 + 1);
    for (uint32_t e = e0; e < e1; ++e)
    {
        float const vx = xj - xi, vy = yj - yi, vz = zj - zi, de = d, we = w;
        if (e + 1 < e1)
        {
            xj = pts[3 * jn];
            yj = pts[3 * jn + 1];
            zj = pts[3 * jn + 2];
            d = src.d(e + 1);
            w = weighted ? wts[e + 1] : 1.0f;
            if (e + 2 < e1)
                jn = src.j(e + 2);
        }
        float ct, st, cp, sp;
        bond_angles(b, vx, vy, vz, de, ct, st, cp, sp);
        f(ct, st, cp, sp, we);
    }
}

// Dynamic LDS of the kernels over one NeighborList that stage bonds: cap
// neighbour indices followed by cap distances, cap sized by the host from the
// mean bond count (stage_capacity).
extern __shared__ uint32_t fa_stage_lds[];

// Steinhardt.cc:120-222: q_lm(i) = sum_b w Y_lm / sum_b w into q (freud's m
// order, interleaved re/im), returns q_l(i) = sqrt(4 pi / (2l+1) sum_m |q_lm|^2).
template<int L>
__device__ inline float qlm_point(const fa::DevBox& b, const float* __restrict__ pts, uint32_t i, const BondSrc& src,
                                  const float* __restrict__ wts, const uint32_t* __restrict__ offsets, int weighted,
                                  float* q)
{
    constexpr int NM = 2 * L + 1;
    float2 acc[L + 1];
#pragma unroll
    for (int m = 0; m <= L; ++m)
        acc[m] = make_float2(0.0f, 0.0f);
    float total_w = 0.0f;
    for_each_bond(b, pts, i, src, wts, weighted, offsets[i], offsets[i + 1],
                  [&](float ct, float st, float cp, float sp, float w) {
                      accumulate_ylm<L>(ct, st, cp, sp, w, acc);
                      total_w += w;
                  });
    float norm = 0.0f;
#pragma unroll
    for (int m = 0; m <= L; ++m)
    {
        float const re = acc[m].x / total_w, im = acc[m].y / total_w;
        q[2 * m] = re;
        q[2 * m + 1] = im;
        if (m > 0)
        {
            float const sgn = (m & 1) ? -1.0f : 1.0f;
            q[2 * (L + m)] = sgn * re;
            q[2 * (L + m) + 1] = -sgn * im;
        }
    }
#pragma unroll
    for (int t = 0; t < NM; ++t)
        norm += q[2 * t] * q[2 * t] + q[2 * t + 1] * q[2 * t + 1];
    float const nf = float(4.0 * M_PI / NM);
    return sqrtf(norm * nf);
}

// Block partial sum of q_lm(i) * invn over the block's valid points (system
// q_lm, Steinhardt.cc:213-216), written to out[0..NM).  Each wave folds its
// 64 lanes with butterfly shuffles, one barrier, then 2 NM threads add the
// block's four wave sums.  Every loop over q is unrolled so that q stays in
// registers.  Every thread of the (kBlock-thread) block must call it.
template<int L> __device__ inline void block_qlm_partial(bool valid, const float* q, float invn, float2* out)
{
    constexpr int NM = 2 * L + 1;
    constexpr int kWaves = kBlock / 64;
    __shared__ float s_w[kWaves][2 * NM];
    uint32_t const lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < 2 * NM; ++t)
    {
        float v = valid ? q[t] * invn : 0.0f;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1)
            v += __shfl_xor(v, d);
        if (lane == 0)
            s_w[wave][t] = v;
    }
    __syncthreads();
    if (threadIdx.x < 2 * NM)
    {
        float v = s_w[0][threadIdx.x];
#pragma unroll
        for (int w = 1; w < kWaves; ++w)
            v += s_w[w][threadIdx.x];
        reinterpret_cast<float*>(out)[threadIdx.x] = v;
    }
}

template<int L>
__global__ __launch_bounds__(kBlock) void qlm_kernel(fa::DevBox b, const float* __restrict__ pts, uint32_t n,
                                                     const uint32_t* __restrict__ nbr, const float* __restrict__ dist,
                                                     const float* __restrict__ wts, const uint32_t* __restrict__ offsets,
                                                     int weighted, int want_sys, float2* __restrict__ qlm,
                                                     float* __restrict__ ql, int ql_stride, int ql_col,
                                                     float2* __restrict__ sys_partial, uint32_t stage_cap)
{
    constexpr int NM = 2 * L + 1;
    uint32_t const tile0 = fa::xcd_tile(blockIdx.x, gridDim.x) * blockDim.x;
    uint32_t const i = tile0 + threadIdx.x;
    bool const valid = i < n;
    BondSrc const src = stage_bonds(tile0, min(blockDim.x, n - tile0), offsets, nbr, dist, fa_stage_lds,
                                    reinterpret_cast<float*>(fa_stage_lds + stage_cap), stage_cap);
    float q[2 * NM];
    if (valid)
    {
        float const v = qlm_point<L>(b, pts, i, src, wts, offsets, weighted, q);
        float2* out = qlm + uint64_t(i) * NM;
#pragma unroll
        for (int t = 0; t < NM; ++t)
            out[t] = make_float2(q[2 * t], q[2 * t + 1]);
        ql[uint64_t(i) * ql_stride + ql_col] = v;
    }
    if (want_sys)
        block_qlm_partial<L>(valid, q, 1.0f / float(n), sys_partial + uint64_t(blockIdx.x) * NM);
}

// Several even l <= LMAX in one bond pass (Steinhardt with a list of l):
// the bond geometry is computed once and one Legendre sweep up to LMAX feeds
// every even l; q_lm, q_l and the block partial sums are then written for the
// l in `mask`.  Accumulators: (l/2)^2 + ... for the even l (49 complex for
// LMAX = 12).
constexpr int kFuseLMax = 12;
struct MultiLOut
{
    float2* qlm[kLMax + 1];  // (n, 2l + 1) per requested l
    float2* part[kLMax + 1]; // block partial sums per requested l (nullptr: none)
    int col[kLMax + 1];      // column of l in ql (n, ql_stride)
};

__host__ __device__ constexpr int even_base(int l)
{
    return (l / 2) * (l / 2);
}

template<int L, int LMAX>
__device__ inline void emit_even_l(const float2* acc, float total_w, uint32_t mask, bool valid, uint32_t i, float invn,
                                   const MultiLOut& o, float* __restrict__ ql, int ql_stride)
{
    if constexpr (L <= LMAX)
    {
        if ((mask >> L) & 1u)
        {
            constexpr int NM = 2 * L + 1;
            float q[2 * NM];
            float norm = 0.0f;
#pragma unroll
            for (int m = 0; m <= L; ++m)
            {
                float const re = acc[even_base(L) + m].x / total_w, im = acc[even_base(L) + m].y / total_w;
                q[2 * m] = re;
                q[2 * m + 1] = im;
                if (m > 0)
                {
                    float const sgn = (m & 1) ? -1.0f : 1.0f;
                    q[2 * (L + m)] = sgn * re;
                    q[2 * (L + m) + 1] = -sgn * im;
                }
            }
#pragma unroll
            for (int t = 0; t < NM; ++t)
                norm += q[2 * t] * q[2 * t] + q[2 * t + 1] * q[2 * t + 1];
            if (valid)
            {
                float2* out = o.qlm[L] + uint64_t(i) * NM;
#pragma unroll
                for (int t = 0; t < NM; ++t)
                    out[t] = make_float2(q[2 * t], q[2 * t + 1]);
                ql[uint64_t(i) * ql_stride + o.col[L]] = sqrtf(norm * float(4.0 * M_PI / NM));
            }
            if (o.part[L])
                block_qlm_partial<L>(valid, q, invn, o.part[L] + uint64_t(blockIdx.x) * NM);
        }
        emit_even_l<L + 2, LMAX>(acc, total_w, mask, valid, i, invn, o, ql, ql_stride);
    }
}

template<int LMAX, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void qlm_even_multi_kernel(fa::DevBox b, const float* __restrict__ pts, uint32_t n,
                                                                const uint32_t* __restrict__ nbr,
                                                                const float* __restrict__ dist,
                                                                const float* __restrict__ wts,
                                                                const uint32_t* __restrict__ offsets, uint32_t mask,
                                                                MultiLOut o, float* __restrict__ ql, int ql_stride,
                                                                uint32_t stage_cap)
{
    uint32_t const tile0 = fa::xcd_tile(blockIdx.x, gridDim.x) * blockDim.x;
    uint32_t const i = tile0 + threadIdx.x;
    bool const valid = i < n;
    BondSrc const src = stage_bonds(tile0, min(blockDim.x, n - tile0), offsets, nbr, dist, fa_stage_lds,
                                    reinterpret_cast<float*>(fa_stage_lds + stage_cap), stage_cap);
    float2 acc[even_base(LMAX + 2)];
#pragma unroll
    for (int t = 0; t < even_base(LMAX + 2); ++t)
        acc[t] = make_float2(0.0f, 0.0f);
    float total_w = 0.0f;
    if (valid)
        for_each_bond(b, pts, i, src, wts, WEIGHTED, offsets[i], offsets[i + 1],
                      [&](float ct, float st, float cp, float sp, float w) {
                          ylm_sweep<LMAX, 0, true>(ct, st, cp, sp, [&](int l, int m, float p, float er, float ei) {
#pragma clang fp contract(fast)
                              // w * p == p exactly for w == 1
                              float const wp = WEIGHTED ? w * p : p;
                              acc[even_base(l) + m].x += wp * er;
                              if (m > 0) // e^{i 0 phi} = 1: the m = 0 imaginary sum stays +0
                                  acc[even_base(l) + m].y += wp * ei;
                          });
                          total_w += w;
                      });
    emit_even_l<0, LMAX>(acc, total_w, mask, valid, i, 1.0f / float(n), o, ql, ql_stride);
}

// Batched q_l: one workgroup per 256-point tile of one frame; bonds hold
// frame-local indices, offsets are batch-indexed; per-tile partial sums of
// q_lm / N_frame.
template<int L>
__global__ __launch_bounds__(kBlock) void bqlm_kernel(fa::BatchGrid G, const uint32_t* __restrict__ tile_frame,
                                                      uint32_t ntiles, const float* __restrict__ pts,
                                                      const uint32_t* __restrict__ nbr, const float* __restrict__ dist,
                                                      const uint32_t* __restrict__ offsets,
                                                      const unsigned long long* __restrict__ total_bonds,
                                                      uint64_t capacity, float2* __restrict__ qlm,
                                                      float* __restrict__ ql, float2* __restrict__ partial)
{
    constexpr int NM = 2 * L + 1;
    // bonds beyond the allocation were not written (the host reruns)
    if (*total_bonds > capacity)
        return;
    __shared__ uint32_t s_j[kStage];
    __shared__ float s_d[kStage];
    uint32_t const t = fa::xcd_tile(blockIdx.x, ntiles);
    const fa::BatchFrame& F = G.frames[tile_frame[t]];
    uint32_t const i0 = (t - F.tile_base) * fa::kTile;
    uint32_t const i = i0 + threadIdx.x;
    bool const valid = i < F.n;
    BondSrc const src = stage_bonds(i0, min(fa::kTile, F.n - i0), offsets + F.pt_base, nbr, dist, s_j, s_d, kStage);
    float q[2 * NM];
    if (valid)
    {
        float const v = qlm_point<L>(F.box, pts + 3 * size_t(F.pt_base), i, src, nullptr, offsets + F.pt_base, 0, q);
        float2* out = qlm + uint64_t(F.pt_base + i) * NM;
#pragma unroll
        for (int m = 0; m < NM; ++m)
            out[m] = make_float2(q[2 * m], q[2 * m + 1]);
        ql[F.pt_base + i] = v;
    }
    block_qlm_partial<L>(valid, q, 1.0f / float(F.n), partial + uint64_t(t) * NM);
}

// Per frame: system q_lm from its tiles' partials, then the frame's order
// sqrt(4 pi / (2l+1) sum_m |Q_lm|^2) (Steinhardt.cc:291-359 with the
// reference's float operations).
__global__ __launch_bounds__(kBlock) void bframe_order_kernel(fa::BatchGrid G, int nm, const float2* __restrict__ part,
                                                              float* __restrict__ order)
{
    __shared__ float2 s_red[kBlock];
    __shared__ float s_norm;
    const fa::BatchFrame& F = G.frames[blockIdx.x];
    uint32_t const ntl = (F.n + fa::kTile - 1) / fa::kTile;
    if (threadIdx.x == 0)
        s_norm = 0.0f;
    for (int t = 0; t < nm; ++t)
    {
        float2 acc = make_float2(0.0f, 0.0f);
        for (uint32_t bl = threadIdx.x; bl < ntl; bl += blockDim.x)
        {
            float2 const v = part[uint64_t(F.tile_base + bl) * nm + t];
            acc.x += v.x;
            acc.y += v.y;
        }
        s_red[threadIdx.x] = acc;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1)
        {
            if (threadIdx.x < s)
            {
                s_red[threadIdx.x].x += s_red[threadIdx.x + s].x;
                s_red[threadIdx.x].y += s_red[threadIdx.x + s].y;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0)
            s_norm += s_red[0].x * s_red[0].x + s_red[0].y * s_red[0].y;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        order[blockIdx.x] = sqrtf(s_norm * float(4.0 * M_PI / nm));
}

// Steinhardt.cc:224-289: qbar_lm(i) = (sum_j q_lm(j) + q_lm(i)) / (n_b + 1)
template<int L>
__global__ __launch_bounds__(kBlock) void qlm_ave_kernel(uint32_t n, const uint32_t* __restrict__ nbr,
                                                         const uint32_t* __restrict__ offsets,
                                                         const float2* __restrict__ qlm, float2* __restrict__ qlmave,
                                                         float* __restrict__ qlave, int ql_stride, int ql_col,
                                                         float2* __restrict__ sys_partial,
                                                         const uint32_t* __restrict__ order, uint32_t stage_cap)
{
    constexpr int NM = 2 * L + 1;
    uint32_t const tile0 = fa::xcd_tile(blockIdx.x, gridDim.x) * blockDim.x;
    uint32_t const t0 = tile0 + threadIdx.x;
    bool const valid = t0 < n;
    uint32_t const i = (order && valid) ? order[t0] : t0;
    // in index order a tile's neighbour indices are staged in LDS
    BondSrc const src =
        stage_neighbors(tile0, min(blockDim.x, n - tile0), offsets, nbr, fa_stage_lds, order ? 0u : stage_cap);
    float2 acc[NM];
#pragma unroll
    for (int t = 0; t < NM; ++t)
        acc[t] = make_float2(0.0f, 0.0f);
    if (valid)
    {
        uint32_t const e0 = offsets[i], e1 = offsets[i + 1];
        // the neighbour index of bond e + 1 is loaded before bond e's row is
        // gathered and added
        uint32_t jn = e0 < e1 ? src.j(e0) : 0u;
        for (uint32_t e = e0; e < e1; ++e)
        {
            const float2* row = qlm + uint64_t(jn) * NM;
            if (e + 1 < e1)
                jn = src.j(e + 1);
#pragma unroll
            for (int t = 0; t < NM; ++t)
            {
                float2 const v = row[t];
                acc[t].x += v.x;
                acc[t].y += v.y;
            }
        }
        float const cnt = float(e1 - e0 + 1);
        const float2* own = qlm + uint64_t(i) * NM;
        float norm = 0.0f;
        float2* out = qlmave + uint64_t(i) * NM;
#pragma unroll
        for (int t = 0; t < NM; ++t)
        {
            float2 const v = own[t];
            acc[t].x = (acc[t].x + v.x) / cnt;
            acc[t].y = (acc[t].y + v.y) / cnt;
            out[t] = acc[t];
            norm += acc[t].x * acc[t].x + acc[t].y * acc[t].y;
        }
        float const nf = float(4.0 * M_PI / NM);
        qlave[uint64_t(i) * ql_stride + ql_col] = sqrtf(norm * nf);
    }
    block_qlm_partial<L>(valid, reinterpret_cast<const float*>(acc), 1.0f / float(n),
                         sys_partial + uint64_t(blockIdx.x) * NM);
}

// Sum of the blocks' partial q_lm rows (nm = 2l + 1 <= 41 complex values each), one
// workgroup of kSumThreads threads per sum.  The rows are read as one flat
// array with a thread stride T that is a multiple of nm, so every thread
// always adds the same column and consecutive threads read consecutive
// values (coalesced); four loads are in flight per thread.  The T / nm
// per-thread sums of each column are then added in thread order through LDS
// (the same order on every run).
constexpr int kSumThreads = 1024;
// Workgroup b sums the nblocks rows of jobs.nm[b] complex values at
// jobs.part[b] into jobs.out[b] (the system q_lm of one l).
constexpr int kSumJobs = 16;
struct SumJobs
{
    const float2* part[kSumJobs];
    float2* out[kSumJobs];
    int nm[kSumJobs];
};

__global__ __launch_bounds__(kSumThreads) void sum_partials_kernel(SumJobs jobs, uint32_t nblocks)
{
    const float2* __restrict__ const part = jobs.part[blockIdx.x];
    float2* __restrict__ const out = jobs.out[blockIdx.x];
    uint32_t const nm = uint32_t(jobs.nm[blockIdx.x]);
    uint32_t const per = kSumThreads / nm; // threads per column
    uint32_t const T = per * nm;
    uint32_t const t = threadIdx.x;
    __shared__ float2 s_col[kSumThreads];
    float2 acc = make_float2(0.0f, 0.0f);
    if (t < T)
    {
        uint64_t const total = uint64_t(nblocks) * nm;
        for (uint64_t e = t; e < total; e += 4ull * T)
        {
            float2 v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
            {
                uint64_t const x = e + uint64_t(r) * T;
                v[r] = x < total ? part[x] : make_float2(0.0f, 0.0f);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r)
            {
                acc.x += v[r].x;
                acc.y += v[r].y;
            }
        }
    }
    s_col[t] = acc;
    __syncthreads();
    if (t < nm)
    {
        float2 r = make_float2(0.0f, 0.0f);
        for (uint32_t k = 0; k < per; ++k)
        {
            r.x += s_col[t + k * nm].x;
            r.y += s_col[t + k * nm].y;
        }
        out[t] = r;
    }
}

// Wigner3j.cc:22-56 reduction for every particle, w_l normalisation of
// Steinhardt.cc:345-355.
__device__ inline int lm_index(int l, int m) { return m < 0 ? l - m : m; }

__global__ __launch_bounds__(kBlock) void wl_kernel(uint32_t n, int l, const float2* __restrict__ src,
                                                    const float* __restrict__ w3j, const float* __restrict__ qlnorm,
                                                    int wl_normalize, float* __restrict__ wl, int stride, int col)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int const nm = 2 * l + 1;
    const float2* q = src + uint64_t(i) * nm;
    float result = 0.0f;
    unsigned counter = 0;
    for (int m1 = -l; m1 <= l; ++m1)
    {
        float2 const a = q[lm_index(l, m1)];
        int const m2lo = max(-l - m1, -l), m2hi = min(l - m1, l);
        for (int m2 = m2lo; m2 <= m2hi; ++m2)
        {
            float2 const bq = q[lm_index(l, m2)];
            float2 const c = q[lm_index(l, -m1 - m2)];
            float const w = w3j[counter++];
            // ((w * a) * b) * c, real part, in std::complex<float> order
            float const ar = w * a.x, ai = w * a.y;
            float const abr = ar * bq.x - ai * bq.y;
            float const abi = ar * bq.y + ai * bq.x;
            result += abr * c.x - abi * c.y;
        }
    }
    if (wl_normalize)
    {
        float const nf = float(4.0 * M_PI / nm);
        float const norm = sqrtf(nf) / qlnorm[uint64_t(i) * stride + col];
        result *= norm * norm * norm;
    }
    wl[uint64_t(i) * stride + col] = result;
}

// The same reduction with q_lm held in registers and the (m1, m2) loops
// unrolled at compile time (coefficients are wave-uniform loads).
template<int L>
__global__ __launch_bounds__(kBlock) void wl_kernel_t(uint32_t n, const float2* __restrict__ src,
                                                      const float* __restrict__ w3j, const float* __restrict__ qlnorm,
                                                      int wl_normalize, float* __restrict__ wl, int stride, int col)
{
    constexpr int NM = 2 * L + 1;
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float2 q[NM];
    const float2* qi = src + uint64_t(i) * NM;
#pragma unroll
    for (int t = 0; t < NM; ++t)
        q[t] = qi[t];
    float result = 0.0f;
    int counter = 0;
#pragma unroll
    for (int m1 = -L; m1 <= L; ++m1)
    {
        float2 const a = q[m1 < 0 ? L - m1 : m1];
#pragma unroll
        for (int m2 = (-L - m1 > -L ? -L - m1 : -L); m2 <= (L - m1 < L ? L - m1 : L); ++m2)
        {
            int const m3 = -m1 - m2;
            float2 const bq = q[m2 < 0 ? L - m2 : m2];
            float2 const c = q[m3 < 0 ? L - m3 : m3];
            float const w = w3j[counter++];
            float const ar = w * a.x, ai = w * a.y;
            float const abr = ar * bq.x - ai * bq.y;
            float const abi = ar * bq.y + ai * bq.x;
            result += abr * c.x - abi * c.y;
        }
    }
    if (wl_normalize)
    {
        float const nf = float(4.0 * M_PI / NM);
        float const norm = sqrtf(nf) / qlnorm[uint64_t(i) * stride + col];
        result *= norm * norm * norm;
    }
    wl[uint64_t(i) * stride + col] = result;
}

// ------------------------------------------------------------------ learned SolidLiquid features
// Rotation-invariant features of a particle for the learned SolidLiquid
// classifier, from one pass over its bonds: q_l and normalised w_l for
// l = 4, 6, 8, 10, 12 (the unweighted Steinhardt definitions,
// Steinhardt.cc:120-222 and Wigner3j.cc:17-56 with wl_normalize) and the bond
// count.  One associated-Legendre recurrence up to l = 12 serves all five l.
// The particle's q_6m and q_6 are also written, with exactly the arithmetic of
// the plain Steinhardt kernel, so that the exact SolidLiquid fallback computes
// the reference's bond dot products from them.
//
// w_l here sums each unordered triple (m1 <= m2 <= m3, m1 + m2 + m3 = 0) once,
// weighted by its number of distinct orderings: every ordering contributes the
// same product and the same 3j symbol, so this equals the full sum up to float
// rounding (about 80 terms instead of 469 at l = 12).
constexpr int kFeat = 12; // q4..q12, w4..w12, n_b, padding

__host__ __device__ constexpr int feat_base(int l)
{
    return ((l - 4) / 2) * ((l - 4) / 2) + 4 * ((l - 4) / 2);
}

template<int L> __device__ inline float2 qm_signed(const float2* q, int m)
{
    // q_{l,-m} = (-1)^m conj(q_{l,m})
    if (m >= 0)
        return q[m];
    float2 const v = q[-m];
    return (m & 1) ? make_float2(-v.x, v.y) : make_float2(v.x, -v.y);
}

template<int L> __device__ inline float wl_sorted(const float2* q, const float* __restrict__ coef)
{
    float result = 0.0f;
    int counter = 0;
#pragma unroll
    for (int m1 = -L; m1 <= 0; ++m1)
    {
#pragma unroll
        for (int m2 = (m1 > -L - m1 ? m1 : -L - m1); 2 * m2 <= -m1; ++m2)
        {
            int const m3 = -m1 - m2;
            float2 const a = qm_signed<L>(q, m1);
            float2 const bq = qm_signed<L>(q, m2);
            float2 const c = qm_signed<L>(q, m3);
            float const w = coef[counter++];
            float const abr = a.x * bq.x - a.y * bq.y;
            float const abi = a.x * bq.y + a.y * bq.x;
            result += w * (abr * c.x - abi * c.y);
        }
    }
    return result;
}

template<int L> __device__ inline void feat_l(const float2* acc, float inv_nb, const float* coef, float* qout, float* wout)
{
    float2 q[L + 1];
    float s = 0.0f;
#pragma unroll
    for (int m = 0; m <= L; ++m)
    {
        q[m] = make_float2(acc[m].x * inv_nb, acc[m].y * inv_nb);
        float const mag = q[m].x * q[m].x + q[m].y * q[m].y;
        s += m ? 2.0f * mag : mag;
    }
    *qout = sqrtf(s * float(4.0 * M_PI / (2 * L + 1)));
    float const w = wl_sorted<L>(q, coef);
    *wout = w / (s * sqrtf(s));
}

// LM: the largest l of the features (8 or 12; columns of larger l are 0).
template<int LM>
__global__ __launch_bounds__(kBlock) void sl_features_kernel(fa::DevBox b, const float* __restrict__ pts, uint32_t n,
                                                             const uint32_t* __restrict__ nbr,
                                                             const float* __restrict__ dist,
                                                             const uint32_t* __restrict__ offsets,
                                                             const float* __restrict__ coef4,
                                                             const float* __restrict__ coef6,
                                                             const float* __restrict__ coef8,
                                                             const float* __restrict__ coef10,
                                                             const float* __restrict__ coef12, float* __restrict__ feat,
                                                             float2* __restrict__ qlm6, float* __restrict__ ql6,
                                                             uint32_t stage_cap)
{
    static_assert(LM == 8 || LM == 12, "feature sets: l <= 8 or l <= 12");
    uint32_t const tile0 = fa::xcd_tile(blockIdx.x, gridDim.x) * blockDim.x;
    uint32_t const t0 = tile0 + threadIdx.x;
    BondSrc const src = stage_bonds(tile0, min(blockDim.x, n - tile0), offsets, nbr, dist, fa_stage_lds,
                                    reinterpret_cast<float*>(fa_stage_lds + stage_cap), stage_cap);
    if (t0 >= n)
        return;
    uint32_t const i = t0;
    float2 acc[feat_base(LM + 2)];
#pragma unroll
    for (int t = 0; t < feat_base(LM + 2); ++t)
        acc[t] = make_float2(0.0f, 0.0f);
    uint32_t const e0 = offsets[i], e1 = offsets[i + 1];
    for_each_bond(b, pts, i, src, nullptr, false, e0, e1, [&](float ct, float st, float cp, float sp, float) {
        ylm_sweep<LM, 4, true>(ct, st, cp, sp, [&](int l, int m, float p, float er, float ei) {
#pragma clang fp contract(fast)
            acc[feat_base(l) + m].x += p * er;
            if (m > 0)
                acc[feat_base(l) + m].y += p * ei;
        });
    });
    uint32_t const nb = e1 - e0;
    float const total_w = float(nb);
    float const inv_nb = 1.0f / total_w;
    float* f = feat + uint64_t(i) * kFeat;
    feat_l<4>(acc + feat_base(4), inv_nb, coef4, f + 0, f + 5);
    feat_l<6>(acc + feat_base(6), inv_nb, coef6, f + 1, f + 6);
    feat_l<8>(acc + feat_base(8), inv_nb, coef8, f + 2, f + 7);
    if constexpr (LM == 12)
    {
        feat_l<10>(acc + feat_base(10), inv_nb, coef10, f + 3, f + 8);
        feat_l<12>(acc + feat_base(12), inv_nb, coef12, f + 4, f + 9);
    }
    else
    {
        f[3] = f[4] = f[8] = f[9] = 0.0f;
    }
    f[10] = total_w;
    f[11] = 0.0f;
    // q_6m and q_6 as the plain Steinhardt kernel computes them
    constexpr int L6 = 6, NM = 13;
    const float2* a6 = acc + feat_base(6);
    float q[2 * NM];
    float norm = 0.0f;
#pragma unroll
    for (int m = 0; m <= L6; ++m)
    {
        float const re = a6[m].x / total_w, im = a6[m].y / total_w;
        q[2 * m] = re;
        q[2 * m + 1] = im;
        if (m > 0)
        {
            float const sgn = (m & 1) ? -1.0f : 1.0f;
            q[2 * (L6 + m)] = sgn * re;
            q[2 * (L6 + m) + 1] = -sgn * im;
        }
    }
#pragma unroll
    for (int t = 0; t < NM; ++t)
        norm += q[2 * t] * q[2 * t] + q[2 * t + 1] * q[2 * t + 1];
    float2* out = qlm6 + uint64_t(i) * NM;
#pragma unroll
    for (int t = 0; t < NM; ++t)
        out[t] = make_float2(q[2 * t], q[2 * t + 1]);
    ql6[i] = sqrtf(norm * float(4.0 * M_PI / NM));
}

// SolidLiquid.cc:51-78 (bond dot products) and :90-96 (solid-bond counts)
// Solid-like bond count of particle i; writes each bond's q_ij when ql_ij is
// not null.  NM = 2l + 1 known at compile time keeps q_lm(i) in registers and
// issues the 2l + 1 loads of a neighbour's q_lm together; NM = 0: runtime l.
// With `tile` staged, bond e's neighbour index comes from tile.sj[e - base]
// and its ql_ij goes to tile.so[e - base] (the caller stores the tile's values
// to ql_ij with coalesced writes).
struct SolidTile
{
    const uint32_t* sj = nullptr;
    float* so = nullptr;
    uint32_t base = 0;
    bool staged = false;
};

template<int NM>
__device__ inline uint32_t solid_bond_count(uint32_t i, int l, const uint32_t* __restrict__ nbr,
                                            const uint32_t* __restrict__ offsets, const float2* __restrict__ qlm,
                                            const float* __restrict__ ql, int normalize_q, float q_threshold,
                                            float* __restrict__ ql_ij, const SolidTile& tile = SolidTile())
{
    int const nm = NM ? NM : 2 * l + 1;
    float const nf = float(4.0 * M_PI / nm);
    const float2* qi = qlm + uint64_t(i) * nm;
    uint32_t const e0 = offsets[i], e1 = offsets[i + 1];
    uint32_t cnt = 0;
    float2 a[NM ? NM : 1];
    if (NM)
    {
#pragma unroll
        for (int t = 0; t < NM; ++t)
            a[t] = qi[t];
    }
    for (uint32_t e = e0; e < e1; ++e)
    {
        uint32_t const j = tile.staged ? tile.sj[e - tile.base] : nbr[2 * e + 1];
        const float2* qj = qlm + uint64_t(j) * nm;
        float re = 0.0f;
        if (NM)
        {
            float2 c[NM ? NM : 1];
#pragma unroll
            for (int t = 0; t < NM; ++t)
                c[t] = qj[t];
#pragma unroll
            for (int t = 0; t < NM; ++t)
                re += a[t].x * c[t].x + a[t].y * c[t].y;
        }
        else
        {
            for (int t = 0; t < nm; ++t)
            {
                float2 const av = qi[t], cv = qj[t];
                re += av.x * cv.x + av.y * cv.y;
            }
        }
        if (normalize_q)
            re *= nf / (ql[i] * ql[j]);
        if (tile.staged)
            tile.so[e - tile.base] = re;
        else if (ql_ij)
            ql_ij[e] = re;
        cnt += (re > q_threshold) ? 1u : 0u;
    }
    return cnt;
}

__global__ __launch_bounds__(kBlock) void solid_bonds_kernel(uint32_t n, int l, const uint32_t* __restrict__ nbr,
                                                             const uint32_t* __restrict__ offsets,
                                                             const float2* __restrict__ qlm,
                                                             const float* __restrict__ ql, int normalize_q,
                                                             float q_threshold, float* __restrict__ ql_ij,
                                                             uint32_t* __restrict__ num_conn, uint32_t stage_cap)
{
    // A tile's bonds are one range of the NeighborList: when it fits in
    // stage_cap bonds, the neighbour indices are read into LDS and the ql_ij
    // values written back from LDS, both with coalesced accesses.
    uint32_t const tile0 = fa::xcd_tile(blockIdx.x, gridDim.x) * blockDim.x;
    uint32_t const i = tile0 + threadIdx.x;
    uint32_t const nloc = min(blockDim.x, n - tile0);
    SolidTile tile;
    uint32_t b0 = 0, nb = 0;
    if (stage_cap && ql_ij)
    {
        b0 = offsets[tile0];
        nb = offsets[tile0 + nloc] - b0;
        if (nb <= stage_cap)
        {
            tile.sj = fa_stage_lds;
            tile.so = reinterpret_cast<float*>(fa_stage_lds + stage_cap);
            tile.base = b0;
            tile.staged = true;
            for (uint32_t t = threadIdx.x; t < nb; t += blockDim.x)
                fa_stage_lds[t] = reinterpret_cast<const uint2*>(nbr)[b0 + t].y;
            __syncthreads();
        }
    }
    if (i < n)
        num_conn[i] = l == 6
                          ? solid_bond_count<13>(i, l, nbr, offsets, qlm, ql, normalize_q, q_threshold, ql_ij, tile)
                          : solid_bond_count<0>(i, l, nbr, offsets, qlm, ql, normalize_q, q_threshold, ql_ij, tile);
    if (tile.staged)
    {
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < nb; t += blockDim.x)
            ql_ij[b0 + t] = tile.so[t];
    }
}

// Learned SolidLiquid labels: a particle whose classifier logit is at least
// `margin` away from 0 takes the classifier's label; the others (and NaN
// logits, e.g. particles without bonds) count their solid-like bonds exactly,
// with the arithmetic of solid_bonds_kernel.  fallback[i] marks the latter.
// Labels from the classifier: a particle whose logit lies at least `margin`
// from 0 takes its sign; the others are appended to `list` (the exact
// bond-dot pass below labels them).  They are ~1 % of the particles and
// scattered, so labelling them here would hold nearly every wave for one
// lane's serial walk over its bonds.
__global__ __launch_bounds__(kBlock) void sl_decide_kernel(uint32_t n, const float* __restrict__ logits, float margin,
                                                           uint8_t* __restrict__ solid, uint8_t* __restrict__ fallback,
                                                           uint32_t* __restrict__ n_fallback,
                                                           uint32_t* __restrict__ list)
{
    uint32_t const i = fa::xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    bool fb = false;
    if (i < n)
    {
        float const y = logits[i];
        fb = !(fabsf(y) >= margin);
        solid[i] = fb ? uint8_t(0) : uint8_t(y > 0.0f);
        fallback[i] = fb;
    }
    // list slots: per wave within the block, then one device atomic per block
    __shared__ uint32_t s_cnt, s_base;
    if (threadIdx.x == 0)
        s_cnt = 0;
    __syncthreads();
    uint64_t const ball = __ballot(fb);
    uint32_t const lane = threadIdx.x & 63u;
    uint32_t woff = 0;
    if (lane == 0 && ball)
        woff = atomicAdd(&s_cnt, uint32_t(__popcll(ball)));
    woff = __shfl(woff, 0);
    __syncthreads();
    if (threadIdx.x == 0)
        s_base = s_cnt ? atomicAdd(n_fallback, s_cnt) : 0u;
    __syncthreads();
    if (fb)
        list[s_base + woff + uint32_t(__popcll(ball & ((1ull << lane) - 1ull)))] = i;
}

// SolidLiquid.cc:51-96: the exact label of each listed particle (grid-stride
// over list[0 .. *n_list)).
__global__ __launch_bounds__(kBlock) void sl_exact_kernel(const uint32_t* __restrict__ list,
                                                          const uint32_t* __restrict__ n_list,
                                                          const uint32_t* __restrict__ nbr,
                                                          const uint32_t* __restrict__ offsets,
                                                          const float2* __restrict__ qlm6,
                                                          const float* __restrict__ ql6, int normalize_q,
                                                          float q_threshold, uint32_t solid_threshold,
                                                          uint8_t* __restrict__ solid)
{
    uint32_t const nl = *n_list;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nl; t += gridDim.x * blockDim.x)
    {
        uint32_t const i = list[t];
        solid[i] = solid_bond_count<13>(i, 6, nbr, offsets, qlm6, ql6, normalize_q, q_threshold, nullptr) >=
                   solid_threshold;
    }
}

// SolidLiquid.cc:100-110: solid-like bonds between particles with at least
// solid_threshold solid-like bonds each
__global__ void solid_keep_kernel(uint64_t nb, const uint32_t* __restrict__ nbr, const float* __restrict__ ql_ij,
                                  const uint32_t* __restrict__ num_conn, float q_threshold, uint32_t solid_threshold,
                                  uint8_t* __restrict__ keep)
{
    uint64_t const e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nb)
        return;
    uint2 const ij = reinterpret_cast<const uint2*>(nbr)[e];
    keep[e] = (ql_ij[e] > q_threshold) && num_conn[ij.x] >= solid_threshold && num_conn[ij.y] >= solid_threshold;
}

// ------------------------------------------------------------------ Cluster
// Lock-free union-find: roots always hook to the smaller index, so every root
// is the smallest member of its component (the "min id" of Cluster.cc:62-66).
__device__ inline uint32_t uf_find(uint32_t* parent, uint32_t x)
{
    uint32_t p = parent[x];
    while (p != x)
    {
        uint32_t const gp = parent[p];
        if (gp != p)
            parent[x] = gp; // path halving (benign race: any ancestor is valid)
        x = p;
        p = parent[x];
    }
    return x;
}

__global__ void uf_init_kernel(uint32_t n, uint32_t* parent, uint32_t* size)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
    {
        parent[i] = i;
        size[i] = 0;
    }
}

// Before the union pass: every point hooks to the smallest of itself and its
// (kept) neighbours.  Each link is a bond and points to a smaller index, so
// this is a valid forest whose roots are the smallest members of their trees;
// the union pass then mostly finds both ends of a bond in one tree already
// instead of merging single points one CAS at a time.
// A tile's neighbour indices and keep flags are read into LDS with coalesced
// loads when they fit in stage_cap bonds.
__global__ __launch_bounds__(kBlock) void uf_hook_min_kernel(uint32_t n, const uint32_t* __restrict__ nbr,
                                                             const uint32_t* __restrict__ offsets,
                                                             const uint8_t* __restrict__ keep,
                                                             uint32_t* __restrict__ parent, uint32_t stage_cap)
{
    uint32_t const tile0 = blockIdx.x * blockDim.x;
    uint32_t const i = tile0 + threadIdx.x;
    uint32_t const nloc = min(blockDim.x, n - tile0);
    uint32_t* const sj = fa_stage_lds;
    uint8_t* const sk = reinterpret_cast<uint8_t*>(fa_stage_lds + stage_cap);
    uint32_t const b0 = offsets[tile0], nb = offsets[tile0 + nloc] - b0;
    bool const staged = stage_cap && nb <= stage_cap;
    if (staged)
    {
        for (uint32_t t = threadIdx.x; t < nb; t += blockDim.x)
        {
            sj[t] = reinterpret_cast<const uint2*>(nbr)[b0 + t].y;
            if (keep)
                sk[t] = keep[b0 + t];
        }
        __syncthreads();
    }
    if (i < n)
    {
        uint32_t m = i;
        for (uint32_t e = offsets[i]; e < offsets[i + 1]; ++e)
        {
            bool const k = !keep || (staged ? sk[e - b0] : keep[e]);
            if (k)
                m = min(m, staged ? sj[e - b0] : nbr[2 * e + 1]);
        }
        parent[i] = m;
    }
}

__global__ void uf_union_kernel(uint64_t nb, const uint32_t* __restrict__ nbr, const uint8_t* __restrict__ keep,
                                uint32_t* parent)
{
    uint64_t const e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nb || (keep && !keep[e]))
        return;
    uint2 const ij = reinterpret_cast<const uint2*>(nbr)[e];
    uint32_t a = ij.x, c = ij.y;
    while (true)
    {
        a = uf_find(parent, a);
        c = uf_find(parent, c);
        if (a == c)
            break;
        if (a < c)
        {
            uint32_t const t = a;
            a = c;
            c = t;
        }
        // a > c: hook a under c if a is still a root
        uint32_t const old = atomicCAS(&parent[a], a, c);
        if (old == a)
            break;
        a = old;
    }
}

// One round of pointer jumping, parent <- its third ancestor (three hops), so
// that the distance to a point's parent grows at least fourfold per round.
// Parallel hooking can leave long chains (bonds (i, i+1) processed together
// hook i+1 under i for every i); ceil(log4 n) + 1 rounds bring every point
// next to its root, so the flatten below walks at most a step or two.
// Parents only move to ancestors, so the in-place updates are safe.
__global__ void uf_jump_kernel(uint32_t n, uint32_t* parent)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t const p = parent[i];
    uint32_t const p2 = parent[p];
    uint32_t const p3 = parent[p2];
    if (p3 != p)
        parent[i] = p3;
}

// Root of every point, compressing paths in place as it goes: a node's parent
// only ever moves to one of its ancestors, and roots never change, so the
// concurrent halving of shared paths is safe and later threads walk the
// shortened paths (a long hooking chain would otherwise cost O(depth) per
// point).
// Cluster sizes are counted with one atomic per distinct root in a wave: the
// members of a large cluster would otherwise all add into one counter.
__global__ void uf_flatten_kernel(uint32_t n, uint32_t* parent, uint32_t* __restrict__ root, uint32_t* size)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    bool pending = i < n;
    uint32_t r = 0;
    if (pending)
    {
        r = uf_find(parent, i);
        root[i] = r;
    }
    uint32_t const lane = threadIdx.x & 63u;
    while (__ballot(pending))
    {
        uint64_t const act = __ballot(pending);
        uint32_t const leader = uint32_t(__ffsll(static_cast<unsigned long long>(act)) - 1);
        uint32_t const lr = __shfl(r, int(leader));
        uint64_t const same = __ballot(pending && r == lr);
        if (lane == leader)
            atomicAdd(&size[lr], uint32_t(__popcll(same)));
        if (pending && r == lr)
            pending = false;
    }
}

// key = (n - size) << 32 | root sorts clusters by size descending, then by
// smallest member ascending (Cluster.cc:110-128).  Slots come from one counter
// atomic per workgroup (most points of a liquid are roots; a device-wide
// counter serialises its atomics).
__global__ __launch_bounds__(kBlock) void uf_keys_kernel(uint32_t n, const uint32_t* __restrict__ parent,
                                                         const uint32_t* __restrict__ size,
                                                         uint32_t* __restrict__ counter, uint64_t* __restrict__ keys)
{
    __shared__ uint32_t s_wave[kBlock / 64 + 1];
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    bool const is_root = i < n && parent[i] == i;
    uint64_t const roots = __ballot(is_root);
    uint32_t const lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0)
        s_wave[wave] = uint32_t(__popcll(roots));
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t total = 0;
        for (uint32_t w = 0; w < kBlock / 64; ++w)
        {
            uint32_t const c = s_wave[w];
            s_wave[w] = total;
            total += c;
        }
        s_wave[kBlock / 64] = total ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    if (is_root)
    {
        uint32_t const slot =
            s_wave[kBlock / 64] + s_wave[wave] + uint32_t(__popcll(roots & ((uint64_t(1) << lane) - 1u)));
        keys[slot] = (uint64_t(n - size[i]) << 32) | i;
    }
}

__global__ void uf_rank_kernel(uint32_t n, uint32_t ncl, const uint64_t* __restrict__ keys,
                               uint32_t* __restrict__ rank_of_root, uint32_t* __restrict__ sizes)
{
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ncl)
    {
        rank_of_root[uint32_t(keys[t])] = t;
        if (sizes)
            sizes[t] = n - uint32_t(keys[t] >> 32);
    }
}

__global__ void uf_label_kernel(uint32_t n, const uint32_t* __restrict__ parent, const uint32_t* __restrict__ rank_of_root,
                                uint32_t* __restrict__ cluster_idx)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        cluster_idx[i] = rank_of_root[parent[i]];
}

// ------------------------------------------------------------------ Hexatic
__global__ __launch_bounds__(kBlock) void hexatic_kernel(fa::DevBox b, const float* __restrict__ pts, uint32_t n,
                                                         const uint32_t* __restrict__ nbr,
                                                         const float* __restrict__ wts,
                                                         const uint32_t* __restrict__ offsets, float k, int weighted,
                                                         float2* __restrict__ psi)
{
    uint32_t const i = fa::xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float re = 0.0f, im = 0.0f, total = 0.0f;
    for (uint32_t e = offsets[i]; e < offsets[i + 1]; ++e)
    {
        uint32_t const j = nbr[2 * e + 1];
        float dx, dy, dz;
        fa::wrap(b, pts[3 * j] - pts[3 * i], pts[3 * j + 1] - pts[3 * i + 1], pts[3 * j + 2] - pts[3 * i + 2], dx, dy,
                 dz);
        float const w = weighted ? wts[e] : 1.0f;
        float const th = atan2f(dy, dx);
        float s, c;
        sincosf(k * th, &s, &c);
        re += w * c;
        im += w * s;
        total += w;
    }
    psi[i] = make_float2(re / total, im / total);
}

// ------------------------------------------------------------------ host
std::vector<std::vector<float>> g_w3j_host(kLMax + 1);
float* g_w3j_dev[kLMax + 1] = {nullptr};

int wigner_table(int l, float** d)
{
    if (l > kLMax)
        return fa::fail(FA_ERR_INVALID, "w_l is available for l <= 20 (Wigner3j.cc:59)");
    if (!g_w3j_dev[l])
    {
        std::vector<float> v;
        for (int m1 = -l; m1 <= l; ++m1)
            for (int m2 = std::max(-l - m1, -l); m2 <= std::min(l - m1, l); ++m2)
                v.push_back(float(fa_wigner3j_lll(l, m1, m2, -m1 - m2)));
        g_w3j_host[l] = v;
        if (int rc = fa::device_alloc((void**) &g_w3j_dev[l], sizeof(float) * v.size()))
            return rc;
        FA_HIP(hipMemcpy(g_w3j_dev[l], v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice));
    }
    *d = g_w3j_dev[l];
    return FA_OK;
}

// host reduction of Wigner3j.cc:22-56 in std::complex<float>
float reduce_wigner3j_host(const std::complex<float>* src, int l)
{
    const auto& w = g_w3j_host[l];
    float result = 0;
    unsigned counter = 0;
    auto idx = [l](int m) { return m < 0 ? l - m : m; };
    for (int m1 = -l; m1 <= l; ++m1)
        for (int m2 = std::max(-l - m1, -l); m2 <= std::min(l - m1, l); ++m2)
        {
            result += (w[counter] * src[idx(m1)] * src[idx(m2)] * src[idx(-m1 - m2)]).real();
            ++counter;
        }
    return result;
}

// Bond staging capacity of a kBlock-point tile: 1.3 x the mean bonds per point,
// rounded to 64, within [512, kStage].
uint32_t stage_capacity(const fa_nlist* nl, uint32_t n, uint32_t max_cap = kStage)
{
    if (n == 0)
        return 0;
    double const want = std::ceil(double(nl->num_bonds) / double(n) * kBlock * 1.3);
    uint32_t const cap = uint32_t(std::min(std::max(want, 512.0), double(max_cap)));
    return (cap + 63u) & ~63u;
}

template<int L>
void launch_qlm_l(const fa::DevBox& db, const float* pts, uint32_t n, const fa_nlist* nl, int weighted, int want_sys,
                  float2* qlm, float* ql, int stride, int col, float2* part)
{
    unsigned const blocks = (n + kBlock - 1) / kBlock;
    uint32_t const cap = stage_capacity(nl, n);
    FA_KLAUNCH(qlm_kernel<L>, dim3(blocks), dim3(kBlock), size_t(cap) * 8, fa::stream(), db, pts, n, nl->neighbors,
               nl->distances, nl->weights, nl->offsets, weighted, want_sys, qlm, ql, stride, col, part, cap);
}

template<int L>
void launch_ave_l(uint32_t n, const fa_nlist* nl, const float2* qlm, float2* qlmave, float* qlave, int stride, int col,
                  float2* part, const uint32_t* order)
{
    unsigned const blocks = (n + kBlock - 1) / kBlock;
    uint32_t const cap = order ? 0u : stage_capacity(nl, n);
    FA_KLAUNCH(qlm_ave_kernel<L>, dim3(blocks), dim3(kBlock), size_t(cap) * 4, fa::stream(), n, nl->neighbors,
               nl->offsets, qlm, qlmave, qlave, stride, col, part, order, cap);
}

template<int L>
int dispatch_l(int l, const fa::DevBox& db, const float* pts, uint32_t n, const fa_nlist* nl, int weighted,
               int want_sys, float2* qlm, float* ql, int stride, int col, float2* part)
{
    if (l == L)
    {
        launch_qlm_l<L>(db, pts, n, nl, weighted, want_sys, qlm, ql, stride, col, part);
        return 1;
    }
    if constexpr (L < kLMax)
        return dispatch_l<L + 1>(l, db, pts, n, nl, weighted, want_sys, qlm, ql, stride, col, part);
    return 0;
}

template<int LMAX>
int launch_qlm_even_multi(int lmax, const fa::DevBox& db, const float* pts, uint32_t n, const fa_nlist* nl,
                          int weighted, uint32_t mask, const MultiLOut& o, float* ql, int stride)
{
    if (lmax == LMAX)
    {
        unsigned const blocks = (n + kBlock - 1) / kBlock;
        uint32_t const cap = stage_capacity(nl, n, 2 * kStage);
        if (weighted)
            FA_KLAUNCH((qlm_even_multi_kernel<LMAX, true>), dim3(blocks), dim3(kBlock), size_t(cap) * 8, fa::stream(),
                       db, pts, n, nl->neighbors, nl->distances, nl->weights, nl->offsets, mask, o, ql, stride, cap);
        else
            FA_KLAUNCH((qlm_even_multi_kernel<LMAX, false>), dim3(blocks), dim3(kBlock), size_t(cap) * 8,
                       fa::stream(), db, pts, n, nl->neighbors, nl->distances, nl->weights, nl->offsets, mask, o, ql,
                       stride, cap);
        FA_HIP(hipGetLastError());
        return FA_OK;
    }
    if constexpr (LMAX + 2 <= kFuseLMax)
        return launch_qlm_even_multi<LMAX + 2>(lmax, db, pts, n, nl, weighted, mask, o, ql, stride);
    return fa::fail(FA_ERR_INVALID, "fused q_lm pass: l out of range");
}

int launch_qlm_even_multi(int lmax, const fa::DevBox& db, const float* pts, uint32_t n, const fa_nlist* nl,
                          int weighted, uint32_t mask, const MultiLOut& o, float* ql, int stride)
{
    return launch_qlm_even_multi<2>(lmax, db, pts, n, nl, weighted, mask, o, ql, stride);
}

template<int L>
int dispatch_ave(int l, uint32_t n, const fa_nlist* nl, const float2* qlm, float2* qlmave, float* qlave, int stride,
                 int col, float2* part, const uint32_t* order)
{
    if (l == L)
    {
        launch_ave_l<L>(n, nl, qlm, qlmave, qlave, stride, col, part, order);
        return 1;
    }
    if constexpr (L < kLMax)
        return dispatch_ave<L + 1>(l, n, nl, qlm, qlmave, qlave, stride, col, part, order);
    return 0;
}

} // namespace

namespace fa {
// Steinhardt::compute for one NeighborList; see the C-ABI below.
int steinhardt_run(const fa_box* box, const float* d_points, uint32_t n, const fa_nlist* nl, const uint32_t* ls,
                   int n_l, int flags, float* d_qlm, float* d_ql, float* d_qlave, float* d_qlmave, float* d_wl,
                   float* order)
{
    if (int rc = ensure_ylm_coef())
        return rc;
    if (nl->num_query_points != n)
        return fail(FA_ERR_RUNTIME, "NeighborList found inconsistent array sizes.");
    bool const average = flags & FA_ST_AVERAGE, wl = flags & FA_ST_WL, weighted = flags & FA_ST_WEIGHTED,
               wl_norm = flags & FA_ST_WL_NORMALIZE;
    hipStream_t s = stream();
    DevBox const db = make_devbox(*box);
    unsigned const blocks = (n + kBlock - 1) / kBlock;
    // The bond passes (q_lm) take the points in index order: a tile's bonds
    // are then one range of the NeighborList, staged in LDS with coalesced
    // loads, which beats any processing order with per-thread bond streams
    // even for points stored in random order.  The averaged q_lm gathers whole
    // neighbour rows; it takes the cached spatial order where the points are
    // not stored coherently.
    const uint32_t* sorder = nullptr;
    if (average)
        if (int rc = spatial_order(*box, d_points, n, &sorder))
            return rc;
    // Partial sums: blocks x (2l + 1) complex per l for the base and for the
    // averaged q_lm (every l live at once: the fused pass writes all of them,
    // and all system sums are reduced and copied back together), then
    // 2l + 1 complex per l for the system q_lm.
    uint32_t even_mask = 0;
    int lmax = 0;
    bool fuse = n_l >= 2;
    size_t part_total = 0, nm_total = 0;
    for (int k = 0; k < n_l; ++k)
    {
        int const l = int(ls[k]);
        if (l > kLMax)
            return fail(FA_ERR_INVALID, "freud_amd supports l <= 20");
        fuse = fuse && (l % 2 == 0) && l <= kFuseLMax && !((even_mask >> l) & 1u);
        even_mask |= 1u << l;
        lmax = std::max(lmax, l);
        part_total += size_t(blocks) * (2 * l + 1);
        nm_total += 2 * l + 1;
    }
    void* tmp = nullptr;
    if (int rc = scratch(6, sizeof(float2) * (2 * part_total + nm_total) + 256, &tmp))
        return rc;
    float2* const part_base = static_cast<float2*>(tmp);
    float2* const part_ave_base = part_base + part_total;
    float2* const sys_base = part_ave_base + part_total;
    if (fuse && n)
    {
        MultiLOut o{};
        size_t mo = 0, po = 0;
        for (int k = 0; k < n_l; ++k)
        {
            int const l = int(ls[k]);
            o.qlm[l] = reinterpret_cast<float2*>(d_qlm) + size_t(n) * mo;
            o.part[l] = average ? nullptr : part_base + po;
            o.col[l] = k;
            mo += 2 * l + 1;
            po += size_t(blocks) * (2 * l + 1);
        }
        if (int rc = launch_qlm_even_multi(lmax, db, d_points, n, nl, weighted, even_mask, o, d_ql, n_l))
            return rc;
    }
    size_t moff = 0, poff = 0;
    for (int k = 0; k < n_l; ++k)
    {
        int const l = int(ls[k]);
        int const nm = 2 * l + 1;
        float2* const part = part_base + poff;
        float2* const part_ave = part_ave_base + poff;
        poff += size_t(blocks) * nm;
        float2* qlm_l = reinterpret_cast<float2*>(d_qlm) + size_t(n) * moff;
        if (n && !fuse)
        {
            dispatch_l<0>(l, db, d_points, n, nl, weighted, !average, qlm_l, d_ql, n_l, k, part);
            FA_HIP(hipGetLastError());
        }
        float2* src = qlm_l;
        if (average && n)
        {
            float2* qlmave_l = reinterpret_cast<float2*>(d_qlmave) + size_t(n) * moff;
            dispatch_ave<0>(l, n, nl, qlm_l, qlmave_l, d_qlave, n_l, k, part_ave, sorder);
            FA_HIP(hipGetLastError());
            src = qlmave_l;
        }
        if (wl && n)
        {
            float* w3 = nullptr;
            if (int rc = wigner_table(l, &w3))
                return rc;
            const float* qn = average ? d_qlave : d_ql;
            switch (l)
            {
#define FA_WL_CASE(LL)                                                                                        \
    case LL:                                                                                                  \
        FA_KLAUNCH(wl_kernel_t<LL>, dim3(blocks), dim3(kBlock), 0, s, n, src, w3, qn, int(wl_norm), d_wl, n_l, \
                   k);                                                                                        \
        break;
                FA_WL_CASE(0)
                FA_WL_CASE(1)
                FA_WL_CASE(2)
                FA_WL_CASE(3)
                FA_WL_CASE(4)
                FA_WL_CASE(5)
                FA_WL_CASE(6)
                FA_WL_CASE(7)
                FA_WL_CASE(8)
                FA_WL_CASE(9)
                FA_WL_CASE(10)
                FA_WL_CASE(11)
                FA_WL_CASE(12)
#undef FA_WL_CASE
            default:
                FA_KLAUNCH(wl_kernel, dim3(blocks), dim3(kBlock), 0, s, n, l, src, w3, qn, int(wl_norm), d_wl, n_l,
                           k);
            }
            FA_HIP(hipGetLastError());
        }
        moff += nm;
    }
    // system order (Steinhardt.cc:291-327): every l's system q_lm reduced on
    // the device, one copy back, one synchronisation
    std::vector<std::complex<float>> hsys(nm_total, 0.0f);
    if (n)
    {
        size_t po = 0, so = 0;
        for (int k0 = 0; k0 < n_l; k0 += kSumJobs)
        {
            SumJobs jobs{};
            int const nj = std::min(kSumJobs, n_l - k0);
            for (int t = 0; t < nj; ++t)
            {
                int const nm = 2 * int(ls[k0 + t]) + 1;
                jobs.part[t] = (average ? part_ave_base : part_base) + po;
                jobs.out[t] = sys_base + so;
                jobs.nm[t] = nm;
                po += size_t(blocks) * nm;
                so += nm;
            }
            FA_KLAUNCH(sum_partials_kernel, dim3(nj), dim3(kSumThreads), 0, s, jobs, blocks);
            FA_HIP(hipGetLastError());
        }
        size_t const sys_bytes = sizeof(float2) * nm_total;
        void* stage = nullptr;
        if (int rc = fa::readback_stage(sys_bytes, &stage))
            return rc;
        FA_HIP(hipMemcpyAsync(stage ? stage : hsys.data(), sys_base, sys_bytes, hipMemcpyDeviceToHost, s));
        FA_HIP(hipStreamSynchronize(s));
        if (stage)
            std::memcpy(hsys.data(), stage, sys_bytes);
    }
    size_t so = 0;
    for (int k = 0; k < n_l; ++k)
    {
        int const l = int(ls[k]);
        int const nm = 2 * l + 1;
        const std::complex<float>* hs = hsys.data() + so;
        so += nm;
        float const nf = float(4.0 * M_PI / nm);
        float calc = 0;
        for (int t = 0; t < nm; ++t)
            calc += std::norm(hs[t]);
        float const ql_sys = std::sqrt(calc * nf);
        if (wl)
        {
            float* w3 = nullptr;
            if (int rc = wigner_table(l, &w3))
                return rc;
            float v = reduce_wigner3j_host(hs, l);
            if (wl_norm)
            {
                float const wn = std::sqrt(nf) / ql_sys;
                v *= wn * wn * wn;
            }
            order[k] = v;
        }
        else
            order[k] = ql_sys;
    }
    return FA_OK;
}
// Coefficients of wl_sorted<l>: W3j(l; m1, m2, m3) times the number of
// distinct orderings of (m1, m2, m3), in the kernel's loop order.
int sorted_wigner_table(int l, float** d)
{
    static float* dev[kLMax + 1] = {};
    if (!dev[l])
    {
        std::vector<float> v;
        for (int m1 = -l; m1 <= 0; ++m1)
            for (int m2 = std::max(m1, -l - m1); 2 * m2 <= -m1; ++m2)
            {
                int const m3 = -m1 - m2;
                int const mult = (m1 == m2 && m2 == m3) ? 1 : ((m1 == m2 || m2 == m3) ? 3 : 6);
                v.push_back(float(double(mult) * fa_wigner3j_lll(l, m1, m2, m3)));
            }
        if (int rc = device_alloc((void**) &dev[l], sizeof(float) * v.size()))
            return rc;
        FA_HIP(hipMemcpy(dev[l], v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice));
    }
    *d = dev[l];
    return FA_OK;
}

template<int L>
int launch_batch_qlm(uint32_t l, const BatchGrid& G, const uint32_t* d_tile_frame, uint32_t ntiles, const float* d_pts,
                     const uint32_t* nb, const float* dist, const uint32_t* d_offsets,
                     const unsigned long long* d_total, uint64_t capacity, float* d_qlm, float* d_ql, float* d_partial)
{
    if (uint32_t(L) == l)
    {
        FA_KLAUNCH(bqlm_kernel<L>, dim3(ntiles), dim3(kBlock), 0, stream(), G, d_tile_frame, ntiles, d_pts, nb, dist,
                   d_offsets, d_total, capacity, reinterpret_cast<float2*>(d_qlm), d_ql,
                   reinterpret_cast<float2*>(d_partial));
        FA_HIP(hipGetLastError());
        return FA_OK;
    }
    if constexpr (L < kLMax)
        return launch_batch_qlm<L + 1>(l, G, d_tile_frame, ntiles, d_pts, nb, dist, d_offsets, d_total, capacity,
                                       d_qlm, d_ql, d_partial);
    return fail(FA_ERR_INVALID, "freud_amd supports l <= 20");
}

// Batched Steinhardt q_l over the batch's ball-query bonds: d_qlm (total x
// (2l+1) complex), d_ql (total), d_partial (ntiles x (2l+1) complex scratch),
// d_order (n_frames, device).
int batch_steinhardt(const BatchGrid& G, const uint32_t* d_tile_frame, uint32_t ntiles, const float* d_pts,
                     const uint32_t* nb, const float* dist, const uint32_t* d_offsets,
                     const unsigned long long* d_total, uint64_t capacity, uint32_t l, float* d_qlm, float* d_ql,
                     float* d_partial, float* d_order)
{
    if (int rc = ensure_ylm_coef())
        return rc;
    if (ntiles == 0)
        return FA_OK;
    if (int rc = launch_batch_qlm<0>(l, G, d_tile_frame, ntiles, d_pts, nb, dist, d_offsets, d_total, capacity,
                                     d_qlm, d_ql, d_partial))
        return rc;
    FA_KLAUNCH(bframe_order_kernel, dim3(G.n_frames), dim3(kBlock), 0, stream(), G, int(2 * l + 1),
               reinterpret_cast<const float2*>(d_partial), d_order);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

} // namespace fa

extern "C" {

// (l l l; m1 m2 m3) by the Racah formula in long double (log-factorials)
double fa_wigner3j_lll(int l, int m1, int m2, int m3)
{
    if (m1 + m2 + m3 != 0 || std::abs(m1) > l || std::abs(m2) > l || std::abs(m3) > l)
        return 0.0;
    auto lf = [](int v) { return std::lgamma((long double) v + 1.0L); };
    int const j = l;
    long double const ltri = 3.0L * lf(j) - lf(3 * j + 1);
    long double const lpre = 0.5L * (ltri + lf(j + m1) + lf(j - m1) + lf(j + m2) + lf(j - m2) + lf(j + m3) + lf(j - m3));
    int const kmin = std::max(0, std::max(-m1, m2));
    int const kmax = std::min(j, std::min(j - m1, j + m2));
    long double sum = 0.0L;
    for (int k = kmin; k <= kmax; ++k)
    {
        long double const lt = lf(k) + lf(j - k) + lf(j - m1 - k) + lf(j + m2 - k) + lf(m1 + k) + lf(-m2 + k);
        long double const term = std::exp(lpre - lt);
        sum += (k & 1) ? -term : term;
    }
    return double((((-m3) % 2) != 0) ? -sum : sum);
}

int fa_steinhardt(const fa_box* box, const float* d_points, uint32_t n, const fa_nlist* nl, const uint32_t* ls,
                  int n_l, int flags, float* d_qlm, float* d_ql, float* d_qlave, float* d_qlmave, float* d_wl,
                  float* order)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    return fa::steinhardt_run(box, d_points, n, nl, ls, n_l, flags, d_qlm, d_ql, d_qlave, d_qlmave, d_wl, order);
}

int fa_sl_features(const fa_box* box, const float* d_points, uint32_t n, const fa_nlist* nl, int lmax,
                   float* d_features, float* d_qlm6, float* d_ql6)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    if (lmax != 8 && lmax != 12)
        return fa::fail(FA_ERR_INVALID, "learned SolidLiquid features take lmax 8 or 12");
    if (int rc = ensure_ylm_coef())
        return rc;
    if (nl->num_query_points != n)
        return fa::fail(FA_ERR_RUNTIME, "NeighborList found inconsistent array sizes.");
    float* c[5];
    for (int k = 0; k < 5; ++k)
        if (int rc = fa::sorted_wigner_table(4 + 2 * k, &c[k]))
            return rc;
    if (n == 0)
        return FA_OK;
    fa::DevBox const db = fa::make_devbox(*box);
    unsigned const blocks = (n + kBlock - 1) / kBlock;
    uint32_t const cap = stage_capacity(nl, n);
    if (lmax == 8)
        FA_KLAUNCH(sl_features_kernel<8>, dim3(blocks), dim3(kBlock), size_t(cap) * 8, fa::stream(), db, d_points, n,
                   nl->neighbors, nl->distances, nl->offsets, c[0], c[1], c[2], c[3], c[4], d_features,
                   reinterpret_cast<float2*>(d_qlm6), d_ql6, cap);
    else
        FA_KLAUNCH(sl_features_kernel<12>, dim3(blocks), dim3(kBlock), size_t(cap) * 8, fa::stream(), db, d_points,
                   n, nl->neighbors, nl->distances, nl->offsets, c[0], c[1], c[2], c[3], c[4], d_features,
                   reinterpret_cast<float2*>(d_qlm6), d_ql6, cap);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int fa_sl_decide(uint32_t n, const fa_nlist* nl, const float* d_logits, float margin, const float* d_qlm6,
                 const float* d_ql6, int normalize_q, float q_threshold, uint32_t solid_threshold, uint8_t* d_solid,
                 uint8_t* d_fallback, uint32_t* n_fallback)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    if (nl->num_query_points != n)
        return fa::fail(FA_ERR_RUNTIME, "NeighborList found inconsistent array sizes.");
    // the fallback count, then the list of fallback particles
    uint32_t* cnt = nullptr;
    if (int rc = fa::scratch(15, 256 + sizeof(uint32_t) * size_t(n), (void**) &cnt))
        return rc;
    uint32_t* list = cnt + 64;
    if (int rc = fa::zero_u32(cnt, 1))
        return rc;
    if (n)
    {
        unsigned const blocks = (n + kBlock - 1) / kBlock;
        FA_KLAUNCH(sl_decide_kernel, dim3(blocks), dim3(kBlock), 0, fa::stream(), n, d_logits, margin, d_solid,
                   d_fallback, cnt, list);
        FA_HIP(hipGetLastError());
        // an eighth of the full grid, grid-stride over the list (its length stays on the device)
        FA_KLAUNCH(sl_exact_kernel, dim3(std::max(1u, blocks / 8)), dim3(kBlock), 0, fa::stream(), list, cnt,
                   nl->neighbors, nl->offsets, reinterpret_cast<const float2*>(d_qlm6), d_ql6, normalize_q,
                   q_threshold, solid_threshold, d_solid);
    }
    FA_HIP(hipGetLastError());
    if (n_fallback)
        return fa::read_u32(cnt, n_fallback);
    return FA_OK;
}

int fa_cluster(uint32_t n, const fa_nlist* nl, const uint8_t* d_keep, uint32_t* d_cluster_idx,
               uint32_t* d_cluster_sizes, uint32_t* num_clusters)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    hipStream_t s = fa::stream();
    if (n == 0)
    {
        *num_clusters = 0;
        return FA_OK;
    }
    void* tmp = nullptr;
    size_t const nb4 = sizeof(uint32_t) * size_t(n);
    if (int rc = fa::scratch(7, 4 * nb4 + sizeof(uint64_t) * size_t(n) + 512, &tmp))
        return rc;
    uint32_t* parent = static_cast<uint32_t*>(tmp);
    uint32_t* size = parent + n;
    uint32_t* rank = size + n;
    uint32_t* root = rank + n;
    uint32_t* counter = root + n;
    uint64_t* keys = reinterpret_cast<uint64_t*>((reinterpret_cast<uintptr_t>(counter + 4) + 15) & ~uintptr_t(15));
    unsigned const blocks = (n + kBlock - 1) / kBlock;
    FA_KLAUNCH(uf_init_kernel, dim3(blocks), dim3(kBlock), 0, s, n, parent, size);
    if (nl->num_bonds && nl->num_query_points == n)
    {
        uint32_t const cap = stage_capacity(nl, n);
        FA_KLAUNCH(uf_hook_min_kernel, dim3(blocks), dim3(kBlock), (size_t(cap) * 5 + 7) & ~size_t(7), s, n,
                   nl->neighbors, nl->offsets, d_keep, parent, cap);
    }
    if (nl->num_bonds)
        FA_KLAUNCH(uf_union_kernel, dim3(unsigned((nl->num_bonds + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                           nl->num_bonds, nl->neighbors, d_keep, parent);
    for (uint64_t span = 1; span < n; span *= 4)
        FA_KLAUNCH(uf_jump_kernel, dim3(blocks), dim3(kBlock), 0, s, n, parent);
    FA_KLAUNCH(uf_flatten_kernel, dim3(blocks), dim3(kBlock), 0, s, n, parent, root, size);
    FA_HIP(hipMemsetAsync(counter, 0, sizeof(uint32_t), s));
    FA_KLAUNCH(uf_keys_kernel, dim3(blocks), dim3(kBlock), 0, s, n, parent, size, counter, keys);
    FA_HIP(hipGetLastError());
    uint32_t ncl = 0;
    if (int rc = fa::read_u32(counter, &ncl))
        return rc;
    if (ncl == 0)
        return fa::fail(FA_ERR_RUNTIME, "cluster labelling found no roots");
    if (int rc = fa::sort_u64(keys, ncl))
        return rc;
    FA_KLAUNCH(uf_rank_kernel, dim3((ncl + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, ncl, keys, rank,
                       d_cluster_sizes);
    FA_KLAUNCH(uf_label_kernel, dim3(blocks), dim3(kBlock), 0, s, n, root, rank, d_cluster_idx);
    FA_HIP(hipGetLastError());
    FA_HIP(hipStreamSynchronize(s));
    *num_clusters = ncl;
    return FA_OK;
}

int fa_solidliquid(const fa_box* box, const float* d_points, uint32_t n, const fa_nlist* nl, uint32_t l,
                   float q_threshold, uint32_t solid_threshold, int normalize_q, float* d_qlm, float* d_ql,
                   float* d_ql_ij, uint32_t* d_num_connections, uint32_t* d_cluster_idx, uint32_t* d_cluster_sizes,
                   uint32_t* num_clusters)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    // SolidLiquid.cc:20-29
    if (q_threshold < 0.0f)
        return fa::fail(FA_ERR_INVALID,
                        "SolidLiquid requires that the dot product cutoff q_threshold must be non-negative.");
    float order = 0.0f;
    if (int rc = fa::steinhardt_run(box, d_points, n, nl, &l, 1, 0, d_qlm, d_ql, nullptr, nullptr, nullptr, &order))
        return rc;
    hipStream_t s = fa::stream();
    unsigned const blocks = (n + kBlock - 1) / kBlock;
    if (n)
    {
        uint32_t const cap = stage_capacity(nl, n);
        FA_KLAUNCH(solid_bonds_kernel, dim3(blocks), dim3(kBlock), size_t(cap) * 8, s, n, int(l), nl->neighbors,
                   nl->offsets, reinterpret_cast<const float2*>(d_qlm), d_ql, normalize_q, q_threshold, d_ql_ij,
                   d_num_connections, cap);
    }
    FA_HIP(hipGetLastError());
    void* keep = nullptr;
    if (int rc = fa::scratch(8, std::max<uint64_t>(nl->num_bonds, 1) + 64, &keep))
        return rc;
    if (nl->num_bonds)
        FA_KLAUNCH(solid_keep_kernel, dim3(unsigned((nl->num_bonds + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           s, nl->num_bonds, nl->neighbors, d_ql_ij, d_num_connections, q_threshold, solid_threshold,
                           static_cast<uint8_t*>(keep));
    FA_HIP(hipGetLastError());
    return fa_cluster(n, nl, static_cast<uint8_t*>(keep), d_cluster_idx, d_cluster_sizes, num_clusters);
}

int fa_hexatic(const fa_box* box, const float* d_points, uint32_t n, const fa_nlist* nl, uint32_t k, int weighted,
               float* d_psi)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    // Box.h:571-577 enforce2D
    if (!box->is2D)
        return fa::fail(FA_ERR_INVALID, "A 3D box was provided to a class that only supports 2D systems.");
    if (n == 0)
        return FA_OK;
    fa::DevBox const db = fa::make_devbox(*box);
    FA_KLAUNCH(hexatic_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, fa::stream(), db, d_points, n,
               nl->neighbors, nl->weights, nl->offsets, float(k), weighted, reinterpret_cast<float2*>(d_psi));
    FA_HIP(hipGetLastError());
    return FA_OK;
}

} // extern "C"
