// fa_runtime.hip -- device context, memory, errors, box helpers and the
// box-transform kernel of the freud_amd library.
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "fa_common.h"

namespace fa {

std::recursive_mutex& api_mutex()
{
    static std::recursive_mutex m;
    return m;
}

namespace {
thread_local std::string g_err;
std::mutex g_mutex;
hipStream_t g_stream = nullptr;
// pinned host words for small device-to-host read-backs (allocated once)
uint64_t* g_stage = nullptr;
bool g_init = false;
std::vector<std::pair<void*, size_t>> g_scratch;
hipEvent_t g_t0 = nullptr, g_t1 = nullptr;

// profiling state
bool g_prof = false;
struct Pending
{
    std::string name;
    hipEvent_t a, b;
};
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_event_pool;
struct Stat
{
    std::string name;
    uint64_t launches;
    double total_ms;
};
std::vector<Stat> g_stats;

hipEvent_t take_event()
{
    if (!g_event_pool.empty())
    {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void) hipEventCreate(&e);
    return e;
}

void collect()
{
    if (g_pending.empty())
        return;
    (void) hipStreamSynchronize(g_stream);
    for (auto& p : g_pending)
    {
        float ms = 0.0f;
        (void) hipEventElapsedTime(&ms, p.a, p.b);
        bool found = false;
        for (auto& s : g_stats)
            if (s.name == p.name)
            {
                s.launches += 1;
                s.total_ms += ms;
                found = true;
                break;
            }
        if (!found)
            g_stats.push_back({p.name, 1, double(ms)});
        g_event_pool.push_back(p.a);
        g_event_pool.push_back(p.b);
    }
    g_pending.clear();
}
} // namespace

bool profiling() { return g_prof; }

void prof_begin(const char* name)
{
    Pending p{name, take_event(), take_event()};
    (void) hipEventRecord(p.a, g_stream);
    g_pending.push_back(p);
}

void prof_end()
{
    if (!g_pending.empty())
        (void) hipEventRecord(g_pending.back().b, g_stream);
    if (g_pending.size() > 4096)
        collect();
}

void set_error(int code, const std::string& msg)
{
    (void) code;
    g_err = msg;
}

int fail(int code, const std::string& msg)
{
    set_error(code, msg);
    return code;
}

static int ensure_init()
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_init)
        return FA_OK;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
        return fail(FA_ERR_DEVICE, "no HIP device available");
    FA_HIP(hipSetDevice(0));
    FA_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    FA_HIP(hipEventCreate(&g_t0));
    FA_HIP(hipEventCreate(&g_t1));
    FA_HIP(hipHostMalloc(reinterpret_cast<void**>(&g_stage), kReadbackBytes, hipHostMallocDefault));
    if (int rc = noalloc_setup())
        return rc;
    g_init = true;
    return FA_OK;
}

hipStream_t stream() { return g_stream; }

int read_u32(const uint32_t* d, uint32_t* h)
{
    void* stage = nullptr;
    if (int rc = readback_stage(sizeof(uint32_t), &stage))
        return rc;
    FA_HIP(hipMemcpyAsync(stage, d, sizeof(uint32_t), hipMemcpyDeviceToHost, g_stream));
    FA_HIP(hipStreamSynchronize(g_stream));
    std::memcpy(h, stage, sizeof(uint32_t));
    return FA_OK;
}

int readback_stage(size_t bytes, void** h)
{
    if (!g_stage)
        return fail(FA_ERR_DEVICE, "library not initialised (no pinned read-back buffer)");
    *h = bytes <= kReadbackBytes ? g_stage : nullptr;
    return FA_OK;
}

// Caching device allocator.  hipFree synchronises the whole device and
// hipMalloc is slow, and a frame-by-frame analysis (a new NeighborQuery with its
// cell grids per frame) would pay both every frame.  Freed blocks go to a
// size-keyed cache instead and are handed out again.  All library work runs in
// order on the library stream, so a block reused by a later call is only
// touched after the earlier call's kernels finished.  Sizes are rounded up
// (256 B granules up to 1 MiB, 2 MiB granules above) and a cached block is
// reused for requests between half and all of its size.  At most
// kCacheLimit bytes stay cached; an allocation failure first returns the
// cache to HIP and retries.
namespace {
std::multimap<size_t, void*> g_cache;
std::unordered_map<void*, size_t> g_live;
size_t g_cached_bytes = 0;
constexpr size_t kCacheLimit = size_t(16) << 30;

size_t round_size(size_t b)
{
    if (b <= (size_t(1) << 20))
        return (b + 255) & ~size_t(255);
    return (b + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
}

void release_cache()
{
    if (g_cache.empty())
        return;
    (void) hipStreamSynchronize(g_stream);
    for (auto& kv : g_cache)
        (void) hipFree(kv.second);
    g_cache.clear();
    g_cached_bytes = 0;
}
} // namespace

int device_alloc(void** p, size_t bytes)
{
    *p = nullptr;
    size_t const want = round_size(bytes == 0 ? 16 : bytes);
    auto it = g_cache.lower_bound(want);
    if (it != g_cache.end() && it->first <= 2 * want)
    {
        *p = it->second;
        g_live[*p] = it->first;
        g_cached_bytes -= it->first;
        g_cache.erase(it);
        return FA_OK;
    }
    hipError_t e = hipMalloc(p, want);
    if (e != hipSuccess)
    {
        (void) hipGetLastError();
        release_cache();
        e = hipMalloc(p, want);
    }
    if (e != hipSuccess)
    {
        *p = nullptr;
        return fail(FA_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    g_live[*p] = want;
    return FA_OK;
}

void device_free(void* p)
{
    if (!p)
        return;
    auto it = g_live.find(p);
    if (it == g_live.end())
    {
        (void) hipFree(p);
        return;
    }
    size_t const sz = it->second;
    g_live.erase(it);
    if (g_cached_bytes + sz > kCacheLimit)
    {
        (void) hipStreamSynchronize(g_stream);
        (void) hipFree(p);
        return;
    }
    g_cache.emplace(sz, p);
    g_cached_bytes += sz;
}

int scratch(size_t slot, size_t bytes, void** ptr)
{
    if (g_scratch.size() <= slot)
        g_scratch.resize(slot + 1, {nullptr, 0});
    auto& s = g_scratch[slot];
    if (s.second < bytes)
    {
        if (s.first)
        {
            FA_HIP(hipStreamSynchronize(g_stream));
            FA_HIP(hipFree(s.first));
        }
        s.first = nullptr;
        s.second = 0;
        size_t const want = bytes + bytes / 4 + 256;
        FA_HIP(hipMalloc(&s.first, want));
        s.second = want;
    }
    *ptr = s.first;
    return FA_OK;
}

// Box.h:97-115 (setL) and NeighborQuery.h:496-564 (image vectors)
DevBox make_devbox(const fa_box& b)
{
    DevBox d;
    d.is2D = b.is2D ? 1 : 0;
    d.L[0] = b.Lx;
    d.L[1] = b.Ly;
    d.L[2] = d.is2D ? 0.0f : b.Lz;
    for (int k = 0; k < 3; ++k)
        d.lo[k] = -(d.L[k] / 2.0f);
    d.xy = b.xy;
    d.xz = b.xz;
    d.yz = b.yz;
    // lattice vectors, Box.h:503-518
    float const a[3] = {d.L[0], 0.0f, 0.0f};
    float const bb[3] = {d.L[1] * b.xy, d.L[1], 0.0f};
    float c[3] = {0.0f, 0.0f, 0.0f};
    if (!d.is2D)
    {
        c[0] = d.L[2] * b.xz;
        c[1] = d.L[2] * b.yz;
        c[2] = d.L[2];
    }
    for (int i = -1; i <= 1; ++i)
        for (int j = -1; j <= 1; ++j)
            for (int k = -1; k <= 1; ++k)
            {
                int const code = (i + 1) * 9 + (j + 1) * 3 + (k + 1);
                float const fi = float(i), fj = float(j), fk = float(k);
                for (int q = 0; q < 3; ++q)
                {
                    volatile float ia = fi * a[q];
                    volatile float jb = fj * bb[q];
                    volatile float kc = fk * c[q];
                    volatile float s = ia + jb;
                    d.img[code][q] = s + kc;
                }
            }
    return d;
}

void nearest_plane_distance(const fa_box& b, float out[3])
{
    float const t = b.xy * b.yz - b.xz;
    volatile float s1 = 1.0f + b.xy * b.xy;
    volatile float s2 = s1 + t * t;
    out[0] = b.Lx / std::sqrt(float(s2));
    volatile float s3 = 1.0f + b.yz * b.yz;
    out[1] = b.Ly / std::sqrt(float(s3));
    out[2] = b.is2D ? 0.0f : b.Lz;
}

float box_volume(const fa_box& b)
{
    if (b.is2D)
        return b.Lx * b.Ly;
    volatile float xy = b.Lx * b.Ly;
    return xy * b.Lz;
}

} // namespace fa

// ------------------------------------------------------------------ kernels
namespace {
__global__ void box_transform_kernel(fa::DevBox b, int op, const float* in, float* out, uint64_t n)
{
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float x = in[3 * i], y = in[3 * i + 1], z = in[3 * i + 2], ox, oy, oz;
    if (op == 0)
        fa::make_fractional(b, x, y, z, ox, oy, oz);
    else if (op == 1)
        fa::make_absolute(b, x, y, z, ox, oy, oz);
    else
        fa::wrap(b, x, y, z, ox, oy, oz);
    out[3 * i] = ox;
    out[3 * i + 1] = oy;
    out[3 * i + 2] = oz;
}
} // namespace

// ------------------------------------------------------------------ C-ABI
extern "C" {

const char* fa_last_error(void) { return fa::g_err.c_str(); }

int fa_init(void)
{
    FA_API_LOCK;
    return fa::ensure_init();
}

int fa_device_name(char* buf, int len)
{
    FA_API_LOCK;
    if (int rc = fa::ensure_init())
        return rc;
    hipDeviceProp_t prop;
    FA_HIP(hipGetDeviceProperties(&prop, 0));
    std::snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return FA_OK;
}

int fa_malloc(void** d_ptr, size_t bytes)
{
    FA_API_LOCK;
    if (int rc = fa::ensure_init())
        return rc;
    return fa::device_alloc(d_ptr, bytes);
}

int fa_free(void* d_ptr)
{
    FA_API_LOCK;
    fa::device_free(d_ptr);
    return FA_OK;
}

int fa_memcpy_h2d(void* d_dst, const void* src, size_t bytes)
{
    FA_API_LOCK;
    if (int rc = fa::ensure_init())
        return rc;
    if (bytes == 0)
        return FA_OK;
    FA_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, fa::stream()));
    FA_HIP(hipStreamSynchronize(fa::stream()));
    return FA_OK;
}

int fa_memcpy_d2h(void* dst, const void* d_src, size_t bytes)
{
    FA_API_LOCK;
    if (int rc = fa::ensure_init())
        return rc;
    if (bytes == 0)
        return FA_OK;
    FA_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, fa::stream()));
    FA_HIP(hipStreamSynchronize(fa::stream()));
    return FA_OK;
}

int fa_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes)
{
    FA_API_LOCK;
    if (int rc = fa::ensure_init())
        return rc;
    if (bytes == 0)
        return FA_OK;
    FA_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, fa::stream()));
    return FA_OK;
}

int fa_memset(void* d_dst, int value, size_t bytes)
{
    FA_API_LOCK;
    if (int rc = fa::ensure_init())
        return rc;
    if (bytes == 0)
        return FA_OK;
    FA_HIP(hipMemsetAsync(d_dst, value, bytes, fa::stream()));
    return FA_OK;
}

int fa_synchronize(void)
{
    FA_API_LOCK;
    if (int rc = fa::ensure_init())
        return rc;
    FA_HIP(hipStreamSynchronize(fa::stream()));
    return FA_OK;
}

int fa_timer_start(void)
{
    FA_API_LOCK;
    if (int rc = fa::ensure_init())
        return rc;
    FA_HIP(hipEventRecord(fa::g_t0, fa::stream()));
    return FA_OK;
}

int fa_timer_stop(float* ms)
{
    FA_API_LOCK;
    FA_HIP(hipEventRecord(fa::g_t1, fa::stream()));
    FA_HIP(hipEventSynchronize(fa::g_t1));
    FA_HIP(hipEventElapsedTime(ms, fa::g_t0, fa::g_t1));
    return FA_OK;
}

int fa_profile_enable(int on)
{
    FA_API_LOCK;
    if (int rc = fa::ensure_init())
        return rc;
    fa::collect();
    fa::g_prof = on != 0;
    return FA_OK;
}

int fa_profile_reset(void)
{
    FA_API_LOCK;
    fa::collect();
    fa::g_stats.clear();
    return FA_OK;
}

int fa_profile_count(void)
{
    FA_API_LOCK;
    fa::collect();
    return int(fa::g_stats.size());
}

int fa_profile_get(int idx, char* name, int len, uint64_t* launches, double* total_ms)
{
    FA_API_LOCK;
    fa::collect();
    if (idx < 0 || idx >= int(fa::g_stats.size()))
        return fa::fail(FA_ERR_RANGE, "profile index out of range");
    std::snprintf(name, len, "%s", fa::g_stats[idx].name.c_str());
    *launches = fa::g_stats[idx].launches;
    *total_ms = fa::g_stats[idx].total_ms;
    return FA_OK;
}

int fa_box_transform(const fa_box* box, int op, const float* d_in, float* d_out, uint64_t n)
{
    FA_API_LOCK;
    if (int rc = fa::ensure_init())
        return rc;
    if (n == 0)
        return FA_OK;
    fa::DevBox b = fa::make_devbox(*box);
    unsigned const blocks = unsigned((n + 255) / 256);
    hipLaunchKernelGGL(box_transform_kernel, dim3(blocks), dim3(256), 0, fa::stream(), b, op, d_in, d_out, n);
    FA_HIP(hipGetLastError());
    return FA_OK;
}

} // extern "C"
