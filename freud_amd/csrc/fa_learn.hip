// fa_learn.hip -- the learned SolidLiquid classifier (no counterpart in the
// reference): an 11 -> 64 -> 64 -> 1 ReLU network over per-particle
// rotation invariants (q_l and normalised w_l for l = 4..12 even, the bond
// count) that predicts whether SolidLiquid (SolidLiquid.cc:31-114) would count
// at least solid_threshold solid-like bonds for the particle.
//
// Matrix work runs on the MI355X matrix cores: v_mfma_f32_16x16x16_bf16 with
// BF16 weights and activations and FP32 accumulation.  The layers are computed
// transposed, H^T = W X^T, so that a wave's 16 particles are the MFMA's N
// dimension: the D fragment of one 16-row block of a layer is then, lane for
// lane, the B fragment of the matching 16-deep K block of the next layer, and
// activations never leave registers.  Operand layouts (per lane l, q = l / 16,
// r = l % 16, j = 0..3):  A[r][4q + j],  B[4q + j][r],  D[4q + j][r].
//
// Training uses the same forward pass (storing pre-activations), an FP32
// backward pass with block-level gradient reductions, and Adam on FP32 master
// weights, which are re-rounded to BF16 after every step.
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "fa_common.h"

namespace {

using bf16x4 = __attribute__((ext_vector_type(4))) short;
using s16x8 = __attribute__((ext_vector_type(8))) short;
using bf16x8 = __attribute__((ext_vector_type(8))) __bf16;
using f32x4 = __attribute__((ext_vector_type(4))) float;

constexpr int kBlock = 256;
constexpr int kIn = 16;        // input width: 11 features, zero padded
constexpr int kFeatUsed = 11;
constexpr int kFeatStride = 12; // feature rows as written by fa_sl_features
constexpr int kH = 64;

// parameter vector layout (FP32 master copy)
constexpr int OFF_W1 = 0;                   // [64][16]
constexpr int OFF_B1 = OFF_W1 + kH * kIn;   // [64]
constexpr int OFF_W2 = OFF_B1 + kH;         // [64][64]
constexpr int OFF_B2 = OFF_W2 + kH * kH;    // [64]
constexpr int OFF_W3 = OFF_B2 + kH;         // [64]
constexpr int OFF_B3 = OFF_W3 + kH;         // [1]
constexpr int kParams = OFF_B3 + 1;

// round to nearest even; NaN stays NaN
__device__ __host__ inline uint16_t to_bf16(float x)
{
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u)
        return uint16_t((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__device__ inline float from_bf16(uint16_t h)
{
    return __uint_as_float(uint32_t(h) << 16);
}

// Per-lane operands of the network for one wave.  Layer 1 (K = 16) uses
// the 16x16x16 BF16 MFMA, layer 2 (K = 64) gfx950's 16x16x32 one in two
// K blocks.  Lane (r, q) = (lane % 16, lane / 16) holds K entries 8q .. 8q+7
// of a 16x16x32 block; K block c is ordered so that these are the hidden
// units 16(2c) + 4q + 0..3 and 16(2c+1) + 4q + 0..3, which are the units the
// lane holds after layer 1 (any K order gives the same sum when W2 follows it).
struct MlpRegs
{
    bf16x4 w1[4];     // W1 block mt
    s16x8 w2[4][2];   // W2 block (mt, K block c)
    float b1[4][4];   // b1[16 mt + 4q + j]
    float b2[4][4];
    float w3[4][4];
    float mu[4], isig[4]; // feature 4q + j
    float b3;
};

__device__ inline bf16x4 load4(const uint16_t* p)
{
    uint2 const v = *reinterpret_cast<const uint2*>(p);
    bf16x4 r;
    r[0] = short(v.x & 0xffffu);
    r[1] = short(v.x >> 16);
    r[2] = short(v.y & 0xffffu);
    r[3] = short(v.y >> 16);
    return r;
}

__device__ inline void load_regs(MlpRegs& R, const uint16_t* __restrict__ w1b, const uint16_t* __restrict__ w2b,
                                 const float* __restrict__ P, const float* __restrict__ norm)
{
    uint32_t const lane = threadIdx.x & 63u, q = lane >> 4, r = lane & 15u;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
    {
        R.w1[mt] = load4(w1b + (mt * 16 + r) * kIn + 4 * q);
#pragma unroll
        for (int c = 0; c < 2; ++c)
        {
            bf16x4 const lo = load4(w2b + (mt * 16 + r) * kH + (2 * c) * 16 + 4 * q);
            bf16x4 const hi = load4(w2b + (mt * 16 + r) * kH + (2 * c + 1) * 16 + 4 * q);
            R.w2[mt][c] = s16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            int const o = mt * 16 + 4 * q + j;
            R.b1[mt][j] = P[OFF_B1 + o];
            R.b2[mt][j] = P[OFF_B2 + o];
            R.w3[mt][j] = P[OFF_W3 + o];
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
        R.mu[j] = norm[4 * q + j];
        R.isig[j] = norm[kIn + 4 * q + j];
    }
    R.b3 = P[OFF_B3];
}

// Features 4q .. 4q+3 of feature row `row` (row == ~0u: padding; zeros past
// the used features).  Rows are 48 bytes, so the 4 values are one aligned
// 16-byte load.
__device__ inline float4 mlp_features(const float* __restrict__ feat, uint32_t row)
{
    uint32_t const q = (threadIdx.x & 63u) >> 4;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (row != ~0u && 4 * int(q) < kFeatUsed)
    {
        v = *reinterpret_cast<const float4*>(feat + uint64_t(row) * kFeatStride + 4 * q);
        static_assert(kFeatUsed == 11 && kFeatStride == 12, "feature 11 is the row's pad");
        if (q == 2)
            v.w = 0.0f;
    }
    return v;
}

// Forward pass of the wave's 16 particles (raw features from mlp_features,
// one particle per r).  Returns the logit (the same in the 4 lanes of a
// particle).  With z1/z2 set, stores the pre-activations of sample `slot`.
__device__ inline float mlp_wave(const MlpRegs& R, float4 v, float* z1, float* z2, uint32_t slot)
{
    uint32_t const lane = threadIdx.x & 63u, q = lane >> 4;
    bf16x4 xb;
    float const vf[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
        int const f = 4 * int(q) + j;
        float const x = f < kFeatUsed ? vf[j] : 0.0f;
        xb[j] = short(to_bf16((x - R.mu[j]) * R.isig[j]));
    }
    f32x4 const zero = {0.0f, 0.0f, 0.0f, 0.0f};
    bf16x4 hb[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
    {
        f32x4 const d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w1[mt], xb, zero, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            float const z = d[j] + R.b1[mt][j];
            if (z1)
                z1[uint64_t(slot) * kH + mt * 16 + 4 * q + j] = z;
            hb[mt][j] = short(to_bf16(fmaxf(z, 0.0f)));
        }
    }
    float partial = 0.0f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
    {
        f32x4 d = zero;
#pragma unroll
        for (int c = 0; c < 2; ++c)
        {
            s16x8 const hk = {hb[2 * c][0],     hb[2 * c][1],     hb[2 * c][2],     hb[2 * c][3],
                              hb[2 * c + 1][0], hb[2 * c + 1][1], hb[2 * c + 1][2], hb[2 * c + 1][3]};
            d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, R.w2[mt][c]),
                                                         __builtin_bit_cast(bf16x8, hk), d, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            float const z = d[j] + R.b2[mt][j];
            if (z2)
                z2[uint64_t(slot) * kH + mt * 16 + 4 * q + j] = z;
            partial += R.w3[mt][j] * fmaxf(z, 0.0f);
        }
    }
    float s = partial + __shfl_xor(partial, 16);
    s = s + __shfl_xor(s, 32);
    return s + R.b3;
}

// Inference: persistent waves over groups of 16 particles.
__global__ __launch_bounds__(kBlock) void mlp_infer_kernel(const uint16_t* __restrict__ w1b,
                                                           const uint16_t* __restrict__ w2b,
                                                           const float* __restrict__ P, const float* __restrict__ norm,
                                                           const float* __restrict__ feat, uint32_t n,
                                                           float* __restrict__ logits)
{
    MlpRegs R;
    load_regs(R, w1b, w2b, P, norm);
    uint32_t const lane = threadIdx.x & 63u;
    uint32_t const waves = gridDim.x * (kBlock / 64);
    uint32_t const ngroups = (n + 15) / 16;
    uint32_t g = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    auto row_of = [&](uint32_t gg) {
        uint32_t const p = gg * 16 + (lane & 15u);
        return gg < ngroups && p < n ? p : ~0u;
    };
    // the next group's features are loaded before the current group's layers run
    float4 next = mlp_features(feat, row_of(g));
    for (; g < ngroups; g += waves)
    {
        float4 const cur = next;
        next = mlp_features(feat, row_of(g + waves));
        uint32_t const p = g * 16 + (lane & 15u);
        float const y = mlp_wave(R, cur, nullptr, nullptr, 0);
        if (lane < 16 && p < n)
            logits[p] = y;
    }
}

// Training forward: sample b of the batch is feature row idx[b].
__global__ __launch_bounds__(kBlock) void mlp_train_fwd_kernel(const uint16_t* __restrict__ w1b,
                                                               const uint16_t* __restrict__ w2b,
                                                               const float* __restrict__ P,
                                                               const float* __restrict__ norm,
                                                               const float* __restrict__ feat,
                                                               const uint32_t* __restrict__ idx, uint32_t nb,
                                                               float* __restrict__ z1, float* __restrict__ z2,
                                                               float* __restrict__ y)
{
    MlpRegs R;
    load_regs(R, w1b, w2b, P, norm);
    uint32_t const lane = threadIdx.x & 63u;
    uint32_t const g = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    uint32_t const b = g * 16 + (lane & 15u);
    if (g * 16 >= nb)
        return;
    bool const ok = b < nb;
    float const v = mlp_wave(R, mlp_features(feat, ok ? idx[b] : ~0u), ok ? z1 : nullptr, ok ? z2 : nullptr, b);
    if (lane < 16 && ok)
        y[b] = v;
}

// Backward (FP32): 64 samples per block, 4 threads per sample (16 units
// each).  Loss = mean over the batch of BCE-with-logits(y, t).  Gradients
// are accumulated into G (atomically across blocks), the loss into *loss.
constexpr int kBS = 64;

__global__ __launch_bounds__(kBlock) void mlp_train_bwd_kernel(const uint16_t* __restrict__ w2b,
                                                               const float* __restrict__ P,
                                                               const float* __restrict__ norm,
                                                               const float* __restrict__ feat,
                                                               const uint8_t* __restrict__ labels,
                                                               const uint32_t* __restrict__ idx, uint32_t nb,
                                                               const float* __restrict__ z1,
                                                               const float* __restrict__ z2,
                                                               const float* __restrict__ y, float* __restrict__ G,
                                                               float* __restrict__ loss)
{
    __shared__ float s_dz2[kBS][kH + 1];
    __shared__ float s_h1[kBS][kH + 1];
    __shared__ float s_dz1[kBS][kH + 1];
    __shared__ float s_x[kBS][kIn];
    __shared__ float s_dy[kBS];
    __shared__ float s_dw3[kH];
    uint32_t const t = threadIdx.x, sl = t >> 2, qq = t & 3u;
    if (t < kH)
        s_dw3[t] = 0.0f;
    __syncthreads();
    uint32_t const b = blockIdx.x * kBS + sl;
    bool const ok = b < nb;
    float const inv_b = 1.0f / float(nb);
    float dy = 0.0f;
    if (ok)
    {
        float const yy = y[b];
        float const tt = labels[idx[b]] ? 1.0f : 0.0f;
        float const sg = 1.0f / (1.0f + expf(-yy));
        dy = (sg - tt) * inv_b;
        if (qq == 0)
        {
            // softplus(y) - t y, computed stably
            float const l = fmaxf(yy, 0.0f) - tt * yy + log1pf(expf(-fabsf(yy)));
            atomicAdd(loss, l * inv_b);
        }
    }
    if (qq == 0)
        s_dy[sl] = dy;
    for (int u = 0; u < 16; ++u)
    {
        int const o = int(qq) * 16 + u;
        float const a2 = ok ? z2[uint64_t(b) * kH + o] : 0.0f;
        float const a1 = ok ? z1[uint64_t(b) * kH + o] : 0.0f;
        s_dz2[sl][o] = a2 > 0.0f ? dy * P[OFF_W3 + o] : 0.0f;
        atomicAdd(&s_dw3[o], dy * fmaxf(a2, 0.0f));
        s_h1[sl][o] = from_bf16(to_bf16(fmaxf(a1, 0.0f)));
    }
    for (int u = 0; u < 4; ++u)
    {
        int const f = int(qq) * 4 + u;
        float v = 0.0f;
        if (ok && f < kFeatUsed)
            v = feat[uint64_t(idx[b]) * kFeatStride + f];
        s_x[sl][f] = ok ? from_bf16(to_bf16((v - norm[f]) * norm[kIn + f])) : 0.0f;
    }
    __syncthreads();
    // dh1 = W2b^T dz2, dz1 = dh1 * [z1 > 0]
    for (int u = 0; u < 16; ++u)
    {
        int const k = int(qq) * 16 + u;
        float acc = 0.0f;
        for (int o = 0; o < kH; ++o)
            acc += from_bf16(w2b[o * kH + k]) * s_dz2[sl][o];
        float const a1 = ok ? z1[uint64_t(b) * kH + k] : 0.0f;
        s_dz1[sl][k] = a1 > 0.0f ? acc : 0.0f;
    }
    __syncthreads();
    // weight gradients, summed over the block's samples
    for (int e = t; e < kH * kH; e += kBlock)
    {
        int const o = e / kH, k = e % kH;
        float acc = 0.0f;
        for (int s = 0; s < kBS; ++s)
            acc += s_dz2[s][o] * s_h1[s][k];
        atomicAdd(&G[OFF_W2 + e], acc);
    }
    for (int e = t; e < kH * kIn; e += kBlock)
    {
        int const o = e / kIn, f = e % kIn;
        float acc = 0.0f;
        for (int s = 0; s < kBS; ++s)
            acc += s_dz1[s][o] * s_x[s][f];
        atomicAdd(&G[OFF_W1 + e], acc);
    }
    if (t < kH)
    {
        float g1 = 0.0f, g2 = 0.0f;
        for (int s = 0; s < kBS; ++s)
        {
            g1 += s_dz1[s][t];
            g2 += s_dz2[s][t];
        }
        atomicAdd(&G[OFF_B1 + t], g1);
        atomicAdd(&G[OFF_B2 + t], g2);
        atomicAdd(&G[OFF_W3 + t], s_dw3[t]);
    }
    if (t == 0)
    {
        float g = 0.0f;
        for (int s = 0; s < kBS; ++s)
            g += s_dy[s];
        atomicAdd(&G[OFF_B3], g);
    }
}

// Adam step on the FP32 parameters (then zero the gradient) and the BF16
// copies of W1 and W2.
__global__ void mlp_adam_kernel(float* __restrict__ P, float* __restrict__ G, float* __restrict__ M,
                                float* __restrict__ V, float lr, float b1, float b2, float eps, float c1, float c2,
                                uint16_t* __restrict__ w1b, uint16_t* __restrict__ w2b)
{
    int const k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kParams)
        return;
    float const g = G[k];
    float const m = b1 * M[k] + (1.0f - b1) * g;
    float const v = b2 * V[k] + (1.0f - b2) * g * g;
    M[k] = m;
    V[k] = v;
    float const p = P[k] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P[k] = p;
    G[k] = 0.0f;
    if (k < OFF_B1)
        w1b[k - OFF_W1] = to_bf16(p);
    else if (k >= OFF_W2 && k < OFF_B2)
        w2b[k - OFF_W2] = to_bf16(p);
}

__global__ void mlp_pack_kernel(const float* __restrict__ P, uint16_t* __restrict__ w1b, uint16_t* __restrict__ w2b)
{
    int const k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < kH * kIn)
        w1b[k] = to_bf16(P[OFF_W1 + k]);
    if (k < kH * kH)
        w2b[k] = to_bf16(P[OFF_W2 + k]);
}

} // namespace

struct fa_mlp
{
    float* P = nullptr;     // kParams
    float* G = nullptr;     // kParams
    float* M = nullptr;     // Adam first moments
    float* V = nullptr;     // Adam second moments
    float* norm = nullptr;  // mu[16], 1/sigma[16]
    uint16_t* w1b = nullptr;
    uint16_t* w2b = nullptr;
    float* loss = nullptr;  // 1
    float* z1 = nullptr;    // batch x 64
    float* z2 = nullptr;
    float* y = nullptr;
    uint32_t* idx = nullptr;
    uint32_t cap = 0;       // batch capacity
    uint64_t steps = 0;
    ~fa_mlp()
    {
        for (void* p : {(void*) P, (void*) G, (void*) M, (void*) V, (void*) norm, (void*) w1b, (void*) w2b,
                        (void*) loss, (void*) z1, (void*) z2, (void*) y, (void*) idx})
            fa::device_free(p);
    }
};

namespace {

int ensure_batch(fa_mlp* m, uint32_t batch)
{
    if (batch <= m->cap)
        return FA_OK;
    fa::device_free(m->z1);
    fa::device_free(m->z2);
    fa::device_free(m->y);
    fa::device_free(m->idx);
    m->z1 = m->z2 = m->y = nullptr;
    m->idx = nullptr;
    m->cap = 0;
    size_t const b = batch + 16;
    if (int rc = fa::device_alloc((void**) &m->z1, sizeof(float) * kH * b))
        return rc;
    if (int rc = fa::device_alloc((void**) &m->z2, sizeof(float) * kH * b))
        return rc;
    if (int rc = fa::device_alloc((void**) &m->y, sizeof(float) * b))
        return rc;
    if (int rc = fa::device_alloc((void**) &m->idx, sizeof(uint32_t) * b))
        return rc;
    m->cap = batch;
    return FA_OK;
}

// Forward + backward of one batch (gradients accumulate into G, loss into
// m->loss).
int grad_batch(fa_mlp* m, const float* d_feat, const uint8_t* d_labels, const uint32_t* d_idx, uint32_t nb)
{
    hipStream_t s = fa::stream();
    uint32_t const groups = (nb + 15) / 16;
    unsigned const fblocks = (groups + kBlock / 64 - 1) / (kBlock / 64);
    FA_KLAUNCH(mlp_train_fwd_kernel, dim3(fblocks), dim3(kBlock), 0, s, m->w1b, m->w2b, m->P, m->norm, d_feat,
               d_idx, nb, m->z1, m->z2, m->y);
    unsigned const bblocks = (nb + kBS - 1) / kBS;
    FA_KLAUNCH(mlp_train_bwd_kernel, dim3(bblocks), dim3(kBlock), 0, s, m->w2b, m->P, m->norm, d_feat, d_labels,
               d_idx, nb, m->z1, m->z2, m->y, m->G, m->loss);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

} // namespace

extern "C" {

int fa_mlp_create(uint64_t seed, fa_mlp** out)
{
    FA_API_LOCK;
    if (int rc = fa_init())
        return rc;
    auto* m = new fa_mlp();
    auto bad = [&](int rc) {
        delete m;
        return rc;
    };
    for (float** p : {&m->P, &m->G, &m->M, &m->V})
        if (int rc = fa::device_alloc((void**) p, sizeof(float) * kParams))
            return bad(rc);
    if (int rc = fa::device_alloc((void**) &m->norm, sizeof(float) * 2 * kIn))
        return bad(rc);
    if (int rc = fa::device_alloc((void**) &m->w1b, sizeof(uint16_t) * kH * kIn))
        return bad(rc);
    if (int rc = fa::device_alloc((void**) &m->w2b, sizeof(uint16_t) * kH * kH))
        return bad(rc);
    if (int rc = fa::device_alloc((void**) &m->loss, sizeof(float) * 4))
        return bad(rc);
    // He-uniform initial weights, zero biases; identity normalisation
    std::mt19937_64 rng(seed);
    std::vector<float> h(kParams, 0.0f);
    auto fill = [&](int off, int fan_out, int fan_in, int used_in) {
        float const a = std::sqrt(6.0f / float(used_in));
        std::uniform_real_distribution<float> u(-a, a);
        for (int o = 0; o < fan_out; ++o)
            for (int i = 0; i < fan_in; ++i)
                h[off + o * fan_in + i] = i < used_in ? u(rng) : 0.0f;
    };
    fill(OFF_W1, kH, kIn, kFeatUsed);
    fill(OFF_W2, kH, kH, kH);
    fill(OFF_W3, 1, kH, kH);
    std::vector<float> nrm(2 * kIn, 0.0f);
    for (int f = 0; f < kFeatUsed; ++f)
        nrm[kIn + f] = 1.0f;
    hipStream_t s = fa::stream();
    auto hip_ok = [&](hipError_t e) { return e == hipSuccess; };
    if (!hip_ok(hipMemcpy(m->P, h.data(), sizeof(float) * kParams, hipMemcpyHostToDevice)) ||
        !hip_ok(hipMemcpy(m->norm, nrm.data(), sizeof(float) * 2 * kIn, hipMemcpyHostToDevice)))
        return bad(fa::fail(FA_ERR_DEVICE, "parameter upload failed"));
    for (float* p : {m->G, m->M, m->V})
        if (int rc = fa::zero_u32(reinterpret_cast<uint32_t*>(p), kParams))
            return bad(rc);
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((kH * kH + 255) / 256), dim3(256), 0, s, m->P, m->w1b, m->w2b);
    if (hipStreamSynchronize(s) != hipSuccess)
        return bad(fa::fail(FA_ERR_DEVICE, "initialisation failed"));
    *out = m;
    return FA_OK;
}

int fa_mlp_destroy(fa_mlp* m)
{
    FA_API_LOCK;
    delete m;
    return FA_OK;
}

int fa_mlp_num_params(void)
{
    FA_API_LOCK;
    return kParams;
}

int fa_mlp_set_params(fa_mlp* m, const float* params)
{
    FA_API_LOCK;
    if (!m || !params)
        return fa::fail(FA_ERR_INVALID, "null argument");
    FA_HIP(hipMemcpyAsync(m->P, params, sizeof(float) * kParams, hipMemcpyHostToDevice, fa::stream()));
    FA_KLAUNCH(mlp_pack_kernel, dim3((kH * kH + 255) / 256), dim3(256), 0, fa::stream(), m->P, m->w1b, m->w2b);
    FA_HIP(hipStreamSynchronize(fa::stream()));
    return FA_OK;
}

int fa_mlp_get_params(fa_mlp* m, float* params)
{
    FA_API_LOCK;
    if (!m || !params)
        return fa::fail(FA_ERR_INVALID, "null argument");
    FA_HIP(hipMemcpyAsync(params, m->P, sizeof(float) * kParams, hipMemcpyDeviceToHost, fa::stream()));
    FA_HIP(hipStreamSynchronize(fa::stream()));
    return FA_OK;
}

int fa_mlp_set_norm(fa_mlp* m, const float* mu, const float* sigma)
{
    FA_API_LOCK;
    if (!m || !mu || !sigma)
        return fa::fail(FA_ERR_INVALID, "null argument");
    std::vector<float> nrm(2 * kIn, 0.0f);
    for (int f = 0; f < kFeatUsed; ++f)
    {
        if (!(sigma[f] > 0.0f))
            return fa::fail(FA_ERR_INVALID, "feature scales must be positive");
        nrm[f] = mu[f];
        nrm[kIn + f] = 1.0f / sigma[f];
    }
    FA_HIP(hipMemcpyAsync(m->norm, nrm.data(), sizeof(float) * 2 * kIn, hipMemcpyHostToDevice, fa::stream()));
    FA_HIP(hipStreamSynchronize(fa::stream()));
    return FA_OK;
}

int fa_mlp_forward(fa_mlp* m, const float* d_features, uint32_t n, float* d_logits)
{
    FA_API_LOCK;
    if (!m)
        return fa::fail(FA_ERR_INVALID, "null argument");
    if (n == 0)
        return FA_OK;
    static int n_cu = 0;
    if (!n_cu)
    {
        int dev = 0;
        FA_HIP(hipGetDevice(&dev));
        FA_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    uint32_t const groups = (n + 15) / 16;
    unsigned const blocks = std::min<unsigned>((groups + 3) / 4, unsigned(std::max(1, n_cu)) * 8);
    FA_KLAUNCH(mlp_infer_kernel, dim3(blocks), dim3(kBlock), 0, fa::stream(), m->w1b, m->w2b, m->P, m->norm,
               d_features, n, d_logits);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

int fa_mlp_gradient(fa_mlp* m, const float* d_features, const uint8_t* d_labels, const uint32_t* d_idx,
                    uint32_t batch, float* grads, float* loss)
{
    FA_API_LOCK;
    if (!m || !grads)
        return fa::fail(FA_ERR_INVALID, "null argument");
    if (batch == 0)
        return fa::fail(FA_ERR_INVALID, "empty batch");
    if (int rc = ensure_batch(m, batch))
        return rc;
    if (int rc = fa::zero_u32(reinterpret_cast<uint32_t*>(m->G), kParams))
        return rc;
    if (int rc = fa::zero_u32(reinterpret_cast<uint32_t*>(m->loss), 1))
        return rc;
    if (int rc = grad_batch(m, d_features, d_labels, d_idx, batch))
        return rc;
    FA_HIP(hipMemcpyAsync(grads, m->G, sizeof(float) * kParams, hipMemcpyDeviceToHost, fa::stream()));
    if (loss)
        FA_HIP(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, fa::stream()));
    FA_HIP(hipStreamSynchronize(fa::stream()));
    return fa::zero_u32(reinterpret_cast<uint32_t*>(m->G), kParams);
}

int fa_mlp_train(fa_mlp* m, const float* d_features, const uint8_t* d_labels, uint32_t n, uint32_t epochs,
                 uint32_t batch, float lr, uint64_t seed, float* epoch_loss)
{
    FA_API_LOCK;
    if (!m)
        return fa::fail(FA_ERR_INVALID, "null argument");
    if (n == 0 || batch == 0)
        return fa::fail(FA_ERR_INVALID, "empty training set or batch");
    batch = std::min(batch, n);
    if (int rc = ensure_batch(m, batch))
        return rc;
    hipStream_t s = fa::stream();
    std::mt19937_64 rng(seed);
    std::vector<uint32_t> perm(n);
    uint32_t* d_perm = nullptr;
    if (int rc = fa::scratch(16, sizeof(uint32_t) * size_t(n) + 256, (void**) &d_perm))
        return rc;
    float const b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t ep = 0; ep < epochs; ++ep)
    {
        for (uint32_t k = 0; k < n; ++k)
            perm[k] = k;
        std::shuffle(perm.begin(), perm.end(), rng);
        FA_HIP(hipMemcpyAsync(d_perm, perm.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
        if (int rc = fa::zero_u32(reinterpret_cast<uint32_t*>(m->loss), 1))
            return rc;
        uint32_t nbatches = 0;
        for (uint32_t off = 0; off + batch <= n; off += batch, ++nbatches)
        {
            if (int rc = grad_batch(m, d_features, d_labels, d_perm + off, batch))
                return rc;
            ++m->steps;
            float const c1 = 1.0f - std::pow(b1, float(m->steps));
            float const c2 = 1.0f - std::pow(b2, float(m->steps));
            FA_KLAUNCH(mlp_adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, s, m->P, m->G, m->M, m->V, lr,
                       b1, b2, eps, c1, c2, m->w1b, m->w2b);
        }
        FA_HIP(hipGetLastError());
        if (epoch_loss)
        {
            float l = 0.0f;
            FA_HIP(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, s));
            FA_HIP(hipStreamSynchronize(s));
            epoch_loss[ep] = nbatches ? l / float(nbatches) : 0.0f;
        }
    }
    FA_HIP(hipStreamSynchronize(s));
    return FA_OK;
}

} // extern "C"
