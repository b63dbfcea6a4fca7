// fa_density.hip -- LocalDensity (freud/density/LocalDensity.cc),
// CorrelationFunction (freud/density/CorrelationFunction.cc), BondOrder
// (freud/environment/BondOrder.cc) and the PMFT family (freud/pmft/PMFTXY.cc, PMFTXYT.cc, PMFTR12.cc, PMFTXYZ.cc, PMFT.h)
// over a NeighborList: the other consumers of the query path beside the RDF.
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

// ---------------------------------------------------------------- PMFTXYT, PMFTR12, PMFTXYZ
// Three-axis RegularAxis histogram, row-major (Histogram.h:327-351): a bond
// counts only if every coordinate falls inside its axis.
struct Axis3
{
    float min[3], max[3], inv[3];
    uint32_t n[3];
};

__device__ inline uint32_t bin3(const Axis3& A, float v0, float v1, float v2)
{
    uint32_t const b0 = axis_bin(v0, A.min[0], A.max[0], A.inv[0], A.n[0]);
    uint32_t const b1 = axis_bin(v1, A.min[1], A.max[1], A.inv[1], A.n[1]);
    uint32_t const b2 = axis_bin(v2, A.min[2], A.max[2], A.inv[2], A.n[2]);
    if (b0 == kOverflow || b1 == kOverflow || b2 == kOverflow)
        return kOverflow;
    return (b0 * A.n[1] + b1) * A.n[2] + b2;
}

// constants::TWO_PI (freud/box/Box.h:24), the float nearest 2 pi
constexpr float kTwoPi = float(2.0 * M_PI);

// util::modulusPositive (utils.h:29-32): fmod(fmod(a, b) + b, b), each fmod exact
__device__ inline float modulus_positive(float a, float b) { return fmodf(fmodf(a, b) + b, b); }

// The angle of (x, y): atan2 in double, rounded once to float (correctly
// rounded but for rare double roundings).  The reference calls the host C
// library's float atan2, whose last bit depends on the library; bins can only
// differ for bonds whose angle lies within an ulp or two of a bin edge.
__device__ inline float angle_of(float y, float x) { return float(atan2(double(y), double(x))); }

// quat rotate (VectorMath.h:810-822): (s^2 - v.v) b + 2 s (v x b) + 2 (v.b) v,
// in that float order, q = (s, x, y, z)
__device__ inline void quat_rotate(float4 q, float bx, float by, float bz, float& ox, float& oy, float& oz)
{
    float const vv = (q.y * q.y + q.z * q.z) + q.w * q.w;
    float const c1 = q.x * q.x - vv;
    float const vb = (q.y * bx + q.z * by) + q.w * bz;
    float const c2 = 2.0f * q.x;
    float const c3 = 2.0f * vb;
    float const cx = q.z * bz - q.w * by;
    float const cy = q.w * bx - q.y * bz;
    float const cz = q.y * by - q.z * bx;
    ox = (c1 * bx + c2 * cx) + c3 * q.y;
    oy = (c1 * by + c2 * cy) + c3 * q.z;
    oz = (c1 * bz + c2 * cz) + c3 * q.w;
}

enum
{
    PMFT_XYT = 0,
    PMFT_R12 = 1,
    PMFT_XYZ = 2,
};

// One pass over the bonds.  Per-point inputs: XYT takes cos / sin of minus the
// query angle (q_cs) and the points' angles (p_theta); R12 the points' and the
// query points' angles; XYZ the query quaternions and the equivalent
// orientations (every bond is binned once per equivalent orientation).
template<int MODE, bool LDS>
__global__ __launch_bounds__(kBlock) void pmft3_kernel(const uint32_t* __restrict__ ij, const float* __restrict__ vec,
                                                       const float* __restrict__ dist, uint64_t nb,
                                                       const float2* __restrict__ q_cs,
                                                       const float* __restrict__ p_theta,
                                                       const float* __restrict__ q_theta,
                                                       const float4* __restrict__ q_quat,
                                                       const float4* __restrict__ equiv, uint32_t n_equiv, Axis3 A,
                                                       uint32_t* __restrict__ counts)
{
    extern __shared__ uint32_t s_hist3[];
    uint32_t const nbins = A.n[0] * A.n[1] * A.n[2];
    if (LDS)
    {
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            s_hist3[t] = 0u;
        __syncthreads();
    }
    auto add = [&](uint32_t bin) {
        if (bin != kOverflow)
        {
            if (LDS)
                atomicAdd(&s_hist3[bin], 1u);
            else
                atomicAdd(&counts[bin], 1u);
        }
    };
    uint64_t const stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t b = uint64_t(fa::xcd_tile(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x; b < nb; b += stride)
    {
        uint32_t const i = ij[2 * b], j = ij[2 * b + 1];
        float const dx = vec[3 * b], dy = vec[3 * b + 1], dz = vec[3 * b + 2];
        if constexpr (MODE == PMFT_XYT)
        {
            // PMFTXYT.cc:84-99
            float2 const cs = q_cs[i];
            float const rx = cs.x * dx + (-cs.y) * dy;
            float const ry = cs.y * dx + cs.x * dy;
            float const d_theta = angle_of(-dy, -dx);
            float const t = modulus_positive(p_theta[j] - d_theta, kTwoPi);
            add(bin3(A, rx, ry, t));
        }
        else if constexpr (MODE == PMFT_R12)
        {
            // PMFTR12.cc:96-110
            float const d_theta1 = angle_of(dy, dx);
            float const d_theta2 = angle_of(-dy, -dx);
            float const t1 = modulus_positive(p_theta[j] - d_theta1, kTwoPi);
            float const t2 = modulus_positive(q_theta[i] - d_theta2, kTwoPi);
            add(bin3(A, dist[b], t1, t2));
        }
        else
        {
            // PMFTXYZ.cc:125-142: v = rotate(conj(q_i), delta), then rotate(equiv_k, v)
            float4 q = q_quat[i];
            q.y = -q.y;
            q.z = -q.z;
            q.w = -q.w;
            float vx, vy, vz;
            quat_rotate(q, dx, dy, dz, vx, vy, vz);
            for (uint32_t k = 0; k < n_equiv; ++k)
            {
                float wx, wy, wz;
                quat_rotate(equiv[k], vx, vy, vz, wx, wy, wz);
                add(bin3(A, wx, wy, wz));
            }
        }
    }
    if (LDS)
    {
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            if (s_hist3[t])
                atomicAdd(&counts[t], s_hist3[t]);
    }
}

// ---------------------------------------------------------------- CorrelationFunction
// CorrelationFunction.cc:84-98: per bond, the distance bin (RegularAxis(bins,
// 0, r_max)) gets one count and conj(values[j]) * query_values[i] in complex
// double.  Block-local histograms in LDS up to kMaxLdsCorrBins bins, merged
// with device atomics; the sum order over bonds is unspecified, as the
// reference's thread-local sums are.
constexpr uint32_t kMaxLdsCorrBins = 2048;

template<bool LDS>
__global__ __launch_bounds__(kBlock) void correlation_kernel(const uint32_t* __restrict__ ij,
                                                             const float* __restrict__ dist, uint64_t nb,
                                                             const double2* __restrict__ values,
                                                             const double2* __restrict__ qvalues, uint32_t nbins,
                                                             float r_max, float inv_width,
                                                             uint32_t* __restrict__ counts,
                                                             double* __restrict__ sums)
{
    extern __shared__ double s_corr[];
    double* s_sum = s_corr;                                          // 2 * nbins
    uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_corr + 2 * nbins); // nbins
    if (LDS)
    {
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
        {
            s_sum[2 * t] = 0.0;
            s_sum[2 * t + 1] = 0.0;
            s_cnt[t] = 0u;
        }
        __syncthreads();
    }
    uint64_t const stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t b = uint64_t(fa::xcd_tile(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x; b < nb; b += stride)
    {
        uint32_t const bin = axis_bin(dist[b], 0.0f, r_max, inv_width, nbins);
        if (bin == kOverflow)
            continue;
        uint32_t const i = ij[2 * b], j = ij[2 * b + 1];
        double2 const x = values[j], y = qvalues[i];
        // conj(x) * y = (x.re y.re + x.im y.im, x.re y.im - x.im y.re)
        double const re = x.x * y.x + x.y * y.y;
        double const im = x.x * y.y - x.y * y.x;
        if (LDS)
        {
            atomicAdd(&s_cnt[bin], 1u);
            atomicAdd(&s_sum[2 * bin], re);
            atomicAdd(&s_sum[2 * bin + 1], im);
        }
        else
        {
            atomicAdd(&counts[bin], 1u);
            atomicAdd(&sums[2 * bin], re);
            atomicAdd(&sums[2 * bin + 1], im);
        }
    }
    if (LDS)
    {
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            if (s_cnt[t])
            {
                atomicAdd(&counts[t], s_cnt[t]);
                atomicAdd(&sums[2 * t], s_sum[2 * t]);
                atomicAdd(&sums[2 * t + 1], s_sum[2 * t + 1]);
            }
    }
}

// ---------------------------------------------------------------- BondOrder
// BondOrder.cc:87-136: per bond, the (possibly rotated) direction's azimuth
// theta = modulusPositive(atan2(y, x), 2 pi) and polar angle
// phi = acos(z / |v|), binned into n_theta x n_phi over [0, 2 pi) x [0, pi).
// Modes (BondOrder.h): 0 bod, 1 lbod, 2 obcd, 3 oocd.  ref_q is the point's
// orientation, q the query point's.  acos, like atan2, runs in double and is
// rounded once (the reference's come from the host C library).
template<int MODE, bool LDS>
__global__ __launch_bounds__(kBlock) void bond_order_kernel(const uint32_t* __restrict__ ij,
                                                            const float* __restrict__ vec, uint64_t nb,
                                                            const float4* __restrict__ p_quat,
                                                            const float4* __restrict__ q_quat, Axis2 A,
                                                            uint32_t* __restrict__ counts)
{
    extern __shared__ uint32_t s_hist_bo[];
    uint32_t const nbins = A.nx * A.ny;
    if (LDS)
    {
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            s_hist_bo[t] = 0u;
        __syncthreads();
    }
    uint64_t const stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t b = uint64_t(fa::xcd_tile(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x; b < nb; b += stride)
    {
        uint32_t const i = ij[2 * b], j = ij[2 * b + 1];
        float vx = vec[3 * b], vy = vec[3 * b + 1], vz = vec[3 * b + 2];
        if constexpr (MODE != 0)
        {
            float4 ref_c = p_quat[j];
            ref_c.y = -ref_c.y;
            ref_c.z = -ref_c.z;
            ref_c.w = -ref_c.w;
            float4 const q = q_quat[i];
            float ox, oy, oz;
            if constexpr (MODE == 2)
            {
                quat_rotate(ref_c, vx, vy, vz, ox, oy, oz);
                quat_rotate(q, ox, oy, oz, vx, vy, vz);
            }
            else if constexpr (MODE == 1)
            {
                quat_rotate(ref_c, vx, vy, vz, ox, oy, oz);
                vx = ox;
                vy = oy;
                vz = oz;
            }
            else
            {
                float zx, zy, zz;
                quat_rotate(q, 0.0f, 0.0f, 1.0f, zx, zy, zz);
                quat_rotate(ref_c, zx, zy, zz, vx, vy, vz);
            }
        }
        float const theta = modulus_positive(angle_of(vy, vx), kTwoPi);
        float const len = sqrtf((vx * vx + vy * vy) + vz * vz);
        float const phi = float(acos(double(vz / len)));
        uint32_t const bt = axis_bin(theta, A.xmin, A.xmax, A.xinv, A.nx);
        uint32_t const bp = axis_bin(phi, A.ymin, A.ymax, A.yinv, A.ny);
        if (bt == kOverflow || bp == kOverflow)
            continue;
        uint32_t const bin = bt * A.ny + bp;
        if (LDS)
            atomicAdd(&s_hist_bo[bin], 1u);
        else
            atomicAdd(&counts[bin], 1u);
    }
    if (LDS)
    {
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < nbins; t += blockDim.x)
            if (s_hist_bo[t])
                atomicAdd(&counts[t], s_hist_bo[t]);
    }
}

// RegularAxis(n, min, max) (Histogram.h:126-138)
void set_axis(Axis3& A, int k, uint32_t n, float mn, float mx)
{
    A.min[k] = mn;
    A.max[k] = mx;
    volatile float w = (mx - mn) / float(n);
    A.inv[k] = 1.0f / w;
    A.n[k] = n;
}

int grid_for(uint64_t work, int cap);

// Uploads host arrays of the given byte sizes into one scratch region and
// returns device pointers (null where the host pointer is null).
int upload_arrays(const void* const* host, const size_t* bytes, int count, void** dev)
{
    size_t total = 0;
    for (int a = 0; a < count; ++a)
        total += (bytes[a] + 255) & ~size_t(255);
    char* base = nullptr;
    if (int rc = fa::scratch(22, total + 256, (void**) &base))
        return rc;
    hipStream_t s = fa::stream();
    size_t off = 0;
    for (int a = 0; a < count; ++a)
    {
        dev[a] = nullptr;
        if (host[a] && bytes[a])
        {
            dev[a] = base + off;
            FA_HIP(hipMemcpyAsync(dev[a], host[a], bytes[a], hipMemcpyHostToDevice, s));
        }
        off += (bytes[a] + 255) & ~size_t(255);
    }
    // the copies read caller memory that may change after this call
    FA_HIP(hipStreamSynchronize(s));
    return FA_OK;
}

template<int MODE>
int launch_pmft3(const fa_nlist* nl, const Axis3& A, const float2* q_cs, const float* p_theta, const float* q_theta,
                 const float4* q_quat, const float4* equiv, uint32_t n_equiv, uint32_t* d_counts)
{
    uint64_t const nbins = uint64_t(A.n[0]) * A.n[1] * A.n[2];
    uint64_t const nb = nl->num_bonds;
    hipStream_t s = fa::stream();
    if (nbins <= kMaxLdsBins2D)
        FA_KLAUNCH((pmft3_kernel<MODE, true>), dim3(grid_for(nb, 2048)), dim3(kBlock), sizeof(uint32_t) * nbins, s,
                   nl->neighbors, nl->vectors, nl->distances, nb, q_cs, p_theta, q_theta, q_quat, equiv, n_equiv, A,
                   d_counts);
    else
        FA_KLAUNCH((pmft3_kernel<MODE, false>), dim3(grid_for(nb, 4096)), dim3(kBlock), 0, s, nl->neighbors,
                   nl->vectors, nl->distances, nb, q_cs, p_theta, q_theta, q_quat, equiv, n_equiv, A, d_counts);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int check_bins3(uint32_t a, uint32_t b, uint32_t c)
{
    if (uint64_t(a) * b * c > 0xffffffffull)
        return fa::fail(FA_ERR_INVALID, "PMFT histogram too large");
    return FA_OK;
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
    FA_API_LOCK;
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
    FA_API_LOCK;
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

int fa_pmftxyt_accumulate(const fa_nlist* nl, const float* orientations, const float* query_orientations, float x_max,
                          float y_max, uint32_t n_x, uint32_t n_y, uint32_t n_t, uint32_t* d_counts)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    // PMFTXYT.cc:27-46
    if (n_x < 1)
        return fa::fail(FA_ERR_INVALID, "PMFTXYT requires at least 1 bin in X.");
    if (n_y < 1)
        return fa::fail(FA_ERR_INVALID, "PMFTXYT requires at least 1 bin in Y.");
    if (n_t < 1)
        return fa::fail(FA_ERR_INVALID, "PMFTXYT requires at least 1 bin in T.");
    if (x_max < 0)
        return fa::fail(FA_ERR_INVALID, "PMFTXYT requires that x_max must be positive.");
    if (y_max < 0)
        return fa::fail(FA_ERR_INVALID, "PMFTXYT requires that y_max must be positive.");
    if (int rc = check_bins3(n_x, n_y, n_t))
        return rc;
    uint32_t const nq = nl->num_query_points, np = nl->num_points;
    if (nl->num_bonds == 0 || nq == 0)
        return FA_OK;
    // cos / sin of minus each query angle with the host C library
    // (rotmat2::fromAngle, VectorMath.h:912-921)
    std::vector<float> h_cs(2 * size_t(nq));
    for (uint32_t i = 0; i < nq; ++i)
    {
        float const t = -query_orientations[i];
        h_cs[2 * i] = std::cos(t);
        h_cs[2 * i + 1] = std::sin(t);
    }
    const void* host[2] = {h_cs.data(), orientations};
    size_t const bytes[2] = {sizeof(float2) * size_t(nq), sizeof(float) * size_t(np)};
    void* dev[2];
    if (int rc = upload_arrays(host, bytes, 2, dev))
        return rc;
    Axis3 A;
    set_axis(A, 0, n_x, -x_max, x_max);
    set_axis(A, 1, n_y, -y_max, y_max);
    set_axis(A, 2, n_t, 0.0f, kTwoPi);
    return launch_pmft3<PMFT_XYT>(nl, A, static_cast<const float2*>(dev[0]), static_cast<const float*>(dev[1]),
                                  nullptr, nullptr, nullptr, 0, d_counts);
}

int fa_pmftr12_accumulate(const fa_nlist* nl, const float* orientations, const float* query_orientations, float r_max,
                          uint32_t n_r, uint32_t n_t1, uint32_t n_t2, uint32_t* d_counts)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    // PMFTR12.cc:29-46
    if (n_r < 1)
        return fa::fail(FA_ERR_INVALID, "PMFTR12 requires at least 1 bin in R.");
    if (n_t1 < 1)
        return fa::fail(FA_ERR_INVALID, "PMFTR12 requires at least 1 bin in T1.");
    if (n_t2 < 1)
        return fa::fail(FA_ERR_INVALID, "PMFTR12 requires at least 1 bin in T2.");
    if (r_max < 0)
        return fa::fail(FA_ERR_INVALID, "PMFTR12 requires that r_max must be positive.");
    if (int rc = check_bins3(n_r, n_t1, n_t2))
        return rc;
    uint32_t const nq = nl->num_query_points, np = nl->num_points;
    if (nl->num_bonds == 0 || nq == 0)
        return FA_OK;
    const void* host[2] = {orientations, query_orientations};
    size_t const bytes[2] = {sizeof(float) * size_t(np), sizeof(float) * size_t(nq)};
    void* dev[2];
    if (int rc = upload_arrays(host, bytes, 2, dev))
        return rc;
    Axis3 A;
    set_axis(A, 0, n_r, 0.0f, r_max);
    set_axis(A, 1, n_t1, 0.0f, kTwoPi);
    set_axis(A, 2, n_t2, 0.0f, kTwoPi);
    return launch_pmft3<PMFT_R12>(nl, A, nullptr, static_cast<const float*>(dev[0]),
                                  static_cast<const float*>(dev[1]), nullptr, nullptr, 0, d_counts);
}

int fa_pmftxyz_accumulate(const fa_nlist* nl, const float* query_orientations, const float* equiv_orientations,
                          uint32_t n_equiv, float x_max, float y_max, float z_max, uint32_t n_x, uint32_t n_y,
                          uint32_t n_z, uint32_t* d_counts)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    // PMFTXYZ.cc:22-48
    if (n_x < 1)
        return fa::fail(FA_ERR_INVALID, "PMFTXYZ requires at least 1 bin in X.");
    if (n_y < 1)
        return fa::fail(FA_ERR_INVALID, "PMFTXYZ requires at least 1 bin in Y.");
    if (n_z < 1)
        return fa::fail(FA_ERR_INVALID, "PMFTXYZ requires at least 1 bin in Z.");
    if (x_max < 0)
        return fa::fail(FA_ERR_INVALID, "PMFTXYZ requires that x_max must be positive.");
    if (y_max < 0)
        return fa::fail(FA_ERR_INVALID, "PMFTXYZ requires that y_max must be positive.");
    if (z_max < 0)
        return fa::fail(FA_ERR_INVALID, "PMFTXYZ requires that z_max must be positive.");
    if (int rc = check_bins3(n_x, n_y, n_z))
        return rc;
    uint32_t const nq = nl->num_query_points;
    if (nl->num_bonds == 0 || nq == 0 || n_equiv == 0)
        return FA_OK;
    const void* host[2] = {query_orientations, equiv_orientations};
    size_t const bytes[2] = {sizeof(float4) * size_t(nq), sizeof(float4) * size_t(n_equiv)};
    void* dev[2];
    if (int rc = upload_arrays(host, bytes, 2, dev))
        return rc;
    Axis3 A;
    set_axis(A, 0, n_x, -x_max, x_max);
    set_axis(A, 1, n_y, -y_max, y_max);
    set_axis(A, 2, n_z, -z_max, z_max);
    return launch_pmft3<PMFT_XYZ>(nl, A, nullptr, nullptr, nullptr, static_cast<const float4*>(dev[0]),
                                  static_cast<const float4*>(dev[1]), n_equiv, d_counts);
}

int fa_correlation_accumulate(const fa_nlist* nl, const double* values, const double* query_values, uint32_t bins,
                              float r_max, uint32_t* d_counts, double* d_sums)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    // CorrelationFunction.cc:27-35
    if (bins == 0)
        return fa::fail(FA_ERR_INVALID, "CorrelationFunction  requires a nonzero number of bins.");
    if (r_max <= 0)
        return fa::fail(FA_ERR_INVALID, "CorrelationFunction requires r_max to be positive.");
    uint32_t const nq = nl->num_query_points, np = nl->num_points;
    uint64_t const nb = nl->num_bonds;
    if (nb == 0 || nq == 0)
        return FA_OK;
    const void* host[2] = {values, query_values};
    size_t const bytes[2] = {2 * sizeof(double) * size_t(np), 2 * sizeof(double) * size_t(nq)};
    void* dev[2];
    if (int rc = upload_arrays(host, bytes, 2, dev))
        return rc;
    volatile float w = (r_max - 0.0f) / float(bins);
    float const inv_width = 1.0f / w;
    hipStream_t s = fa::stream();
    if (bins <= kMaxLdsCorrBins)
        FA_KLAUNCH(correlation_kernel<true>, dim3(grid_for(nb, 2048)), dim3(kBlock),
                   (2 * sizeof(double) + sizeof(uint32_t)) * size_t(bins), s, nl->neighbors, nl->distances, nb,
                   static_cast<const double2*>(dev[0]), static_cast<const double2*>(dev[1]), bins, r_max, inv_width,
                   d_counts, d_sums);
    else
        FA_KLAUNCH(correlation_kernel<false>, dim3(grid_for(nb, 4096)), dim3(kBlock), 0, s, nl->neighbors,
                   nl->distances, nb, static_cast<const double2*>(dev[0]), static_cast<const double2*>(dev[1]), bins,
                   r_max, inv_width, d_counts, d_sums);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int fa_bond_order_accumulate(const fa_nlist* nl, const float* orientations, const float* query_orientations, int mode,
                             uint32_t n_theta, uint32_t n_phi, uint32_t* d_counts)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    // BondOrder.cc:16-24
    if (n_theta < 2)
        return fa::fail(FA_ERR_INVALID, "BondOrder requires at least 2 bins in theta.");
    if (n_phi < 2)
        return fa::fail(FA_ERR_INVALID, "BondOrder requires at least 2 bins in phi.");
    if (mode < 0 || mode > 3)
        return fa::fail(FA_ERR_INVALID, "Unknown BondOrder mode");
    if (uint64_t(n_theta) * n_phi > 0xffffffffull)
        return fa::fail(FA_ERR_INVALID, "BondOrder histogram too large");
    uint32_t const nq = nl->num_query_points, np = nl->num_points;
    uint64_t const nb = nl->num_bonds;
    if (nb == 0 || nq == 0)
        return FA_OK;
    const void* host[2] = {orientations, query_orientations};
    size_t const bytes[2] = {sizeof(float4) * size_t(np), sizeof(float4) * size_t(nq)};
    void* dev[2];
    if (int rc = upload_arrays(host, bytes, 2, dev))
        return rc;
    // RegularAxis(n_theta, 0, TWO_PI) and RegularAxis(n_phi, 0, M_PI) (float(M_PI))
    Axis2 A;
    A.xmin = 0.0f;
    A.xmax = kTwoPi;
    volatile float wt = (kTwoPi - 0.0f) / float(n_theta);
    A.xinv = 1.0f / wt;
    A.ymin = 0.0f;
    A.ymax = float(M_PI);
    volatile float wp = (float(M_PI) - 0.0f) / float(n_phi);
    A.yinv = 1.0f / wp;
    A.nx = n_theta;
    A.ny = n_phi;
    uint32_t const nbins = n_theta * n_phi;
    auto const* pq = static_cast<const float4*>(dev[0]);
    auto const* qq = static_cast<const float4*>(dev[1]);
    hipStream_t s = fa::stream();
#define FA_BO_LAUNCH(M)                                                                                        \
    do                                                                                                         \
    {                                                                                                          \
        if (nbins <= kMaxLdsBins2D)                                                                            \
            FA_KLAUNCH((bond_order_kernel<M, true>), dim3(grid_for(nb, 2048)), dim3(kBlock),                   \
                       sizeof(uint32_t) * nbins, s, nl->neighbors, nl->vectors, nb, pq, qq, A, d_counts);      \
        else                                                                                                   \
            FA_KLAUNCH((bond_order_kernel<M, false>), dim3(grid_for(nb, 4096)), dim3(kBlock), 0, s,            \
                       nl->neighbors, nl->vectors, nb, pq, qq, A, d_counts);                                   \
    } while (0)
    if (mode == 0)
        FA_BO_LAUNCH(0);
    else if (mode == 1)
        FA_BO_LAUNCH(1);
    else if (mode == 2)
        FA_BO_LAUNCH(2);
    else
        FA_BO_LAUNCH(3);
#undef FA_BO_LAUNCH
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int fa_bond_order_reduce(const uint32_t* counts, uint32_t n_theta, uint32_t n_phi, uint32_t n_frames, float* bond_order)
{
    FA_API_LOCK;
    // BondOrder.cc:26-53 (surface areas with the host C library's cos) and :68-74
    volatile float dt = kTwoPi / float(n_theta);
    volatile float dp = M_PI / float(n_phi);
    for (uint32_t t = 0; t < n_theta; ++t)
        for (uint32_t p = 0; p < n_phi; ++p)
        {
            volatile float phi = float(p) * dp;
            volatile float phi2 = phi + dp;
            volatile float c1 = std::cos(float(phi));
            volatile float c2 = std::cos(float(phi2));
            volatile float dc = c1 - c2;
            volatile float sa = dt * dc;
            uint64_t const k = uint64_t(t) * n_phi + p;
            volatile float a = float(counts[k]) / sa;
            bond_order[k] = a / float(n_frames);
        }
    return FA_OK;
}

int fa_correlation_reduce(const uint32_t* counts, const double* sums, uint32_t bins, double* correlation)
{
    FA_API_LOCK;
    // CorrelationFunction.cc:50-60: the complex sum divided by the bin's count
    for (uint32_t t = 0; t < bins; ++t)
    {
        double re = sums[2 * t], im = sums[2 * t + 1];
        if (counts[t] != 0)
        {
            re /= double(counts[t]);
            im /= double(counts[t]);
        }
        correlation[2 * t] = re;
        correlation[2 * t + 1] = im;
    }
    return FA_OK;
}

int fa_pmft_reduce(int kind, const fa_box* box, const uint32_t* counts, const uint32_t* nbins, const float* maxes,
                   uint32_t n_points, uint32_t n_query_points, uint32_t n_frames, uint32_t n_equiv, float* pcf)
{
    FA_API_LOCK;
    // PMFT.h:76-87 (XYT, R12) and PMFTXYZ.cc:92-103, in the reference's float order
    uint64_t const total = uint64_t(nbins[0]) * nbins[1] * nbins[2];
    float const inv_num_dens = fa::box_volume(*box) / float(n_query_points);
    volatile float den = float(n_frames) * float(n_points);
    if (kind == FA_PMFT_XYZ)
        den = den * float(n_equiv);
    float const norm_factor = float(1.0) / den;
    float const prefactor = inv_num_dens * norm_factor;
    std::vector<float> jf;
    if (kind == FA_PMFT_XYT)
    {
        // PMFTXYT.cc:52-57,68-72
        volatile float dx = float(2.0) * maxes[0] / float(nbins[0]);
        volatile float dy = float(2.0) * maxes[1] / float(nbins[1]);
        volatile float dt = 1 / float(nbins[2]);
        volatile float dxy = dx * dy;
        volatile float jac = dxy * dt;
        jf.assign(1, float(1.0) / jac);
    }
    else if (kind == FA_PMFT_XYZ)
    {
        // PMFTXYZ.cc:53-70,96
        volatile float dx = float(2.0) * maxes[0] / float(nbins[0]);
        volatile float dy = float(2.0) * maxes[1] / float(nbins[1]);
        volatile float dz = float(2.0) * maxes[2] / float(nbins[2]);
        volatile float dxy = dx * dy;
        volatile float jac = dxy * dz;
        jf.assign(1, float(1.0) / jac);
    }
    else if (kind == FA_PMFT_R12)
    {
        // PMFTR12.cc:58-77: 1 / (r_center * (dr * dt1 * dt2)) per r bin;
        // centers (edge_i + edge_{i+1}) / 2 of RegularAxis(n_r, 0, r_max)
        uint32_t const n_r = nbins[0];
        volatile float w = (maxes[0] - 0.0f) / float(n_r);
        float const width = w;
        volatile float dr = maxes[0] / float(n_r);
        volatile float dt1 = kTwoPi / float(nbins[1]);
        volatile float dt2 = 1 / float(nbins[2]);
        volatile float p1 = dr * dt1;
        volatile float product = p1 * dt2;
        jf.resize(n_r);
        for (uint32_t r = 0; r < n_r; ++r)
        {
            volatile float e0 = 0.0f + float(r) * width;
            volatile float e1 = 0.0f + float(r + 1) * width;
            volatile float c = (e0 + e1) / float(2.0);
            volatile float rp = c * product;
            jf[r] = float(1.0) / rp;
        }
    }
    else
        return fa::fail(FA_ERR_INVALID, "unknown PMFT kind");
    uint64_t const per_r = uint64_t(nbins[1]) * nbins[2];
    for (uint64_t t = 0; t < total; ++t)
    {
        float const j = kind == FA_PMFT_R12 ? jf[t / per_r] : jf[0];
        volatile float a = float(counts[t]) * prefactor;
        pcf[t] = a * j;
    }
    return FA_OK;
}

int fa_pmftxy_reduce(const fa_box* box, const uint32_t* counts, uint32_t n_x, uint32_t n_y, float x_max, float y_max,
                     uint32_t n_points, uint32_t n_query_points, uint32_t n_frames, float* pcf)
{
    FA_API_LOCK;
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
