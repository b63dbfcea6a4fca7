// fa_tune.hip -- the autotuner's table of launch variants.
//
// Analyses with several correct launch shapes (cell size relative to r_max,
// waves per workgroup) register the variants they can run.  The first call of
// an analysis on an unseen problem class (a key built from the expected number
// of points per cell) times every variant once on the real input, on the
// library stream, and records the fastest; later calls of that class use the
// recorded variant.  Every variant produces bit-identical results, so tuning
// changes speed only.  The table can be read and preset through the C-ABI
// (fa_tune_*), which is how a caller persists tuning across processes.
#include <map>
#include <set>
#include <utility>

#include "fa_common.h"

namespace fa {
namespace {

std::map<std::pair<int, int>, int> g_table;
std::set<std::pair<int, int>> g_preset; // entries set through fa_tune_set (not revised by the library)
bool g_tune_on = true;
hipEvent_t g_e0 = nullptr, g_e1 = nullptr;

} // namespace

bool tuning_enabled()
{
    return g_tune_on;
}

int tuned_variant(int analysis, int key)
{
    auto it = g_table.find({analysis, key});
    return it == g_table.end() ? -1 : it->second;
}

void record_variant(int analysis, int key, int variant)
{
    g_table[{analysis, key}] = variant;
    g_preset.erase({analysis, key});
}

bool variant_preset(int analysis, int key)
{
    return g_preset.count({analysis, key}) != 0;
}

int time_launch(const std::function<int()>& launch, float* ms)
{
    if (!g_e0)
    {
        FA_HIP(hipEventCreate(&g_e0));
        FA_HIP(hipEventCreate(&g_e1));
    }
    FA_HIP(hipEventRecord(g_e0, stream()));
    if (int rc = launch())
        return rc;
    FA_HIP(hipEventRecord(g_e1, stream()));
    FA_HIP(hipEventSynchronize(g_e1));
    FA_HIP(hipEventElapsedTime(ms, g_e0, g_e1));
    return FA_OK;
}

int tune_key(double points_per_volume, double r)
{
    // expected points in an r^3 volume, in half-octaves
    double const v = points_per_volume * r * r * r;
    return v > 0 ? int(std::floor(2.0 * std::log2(v) + 0.5)) : -1000;
}

} // namespace fa

extern "C" {

int fa_tune_enable(int on)
{
    FA_API_LOCK;
    fa::g_tune_on = on != 0;
    return FA_OK;
}

int fa_tune_count(void)
{
    FA_API_LOCK;
    return int(fa::g_table.size());
}

int fa_tune_get(int idx, int* analysis, int* key, int* variant)
{
    FA_API_LOCK;
    if (idx < 0 || idx >= int(fa::g_table.size()))
        return fa::fail(FA_ERR_RANGE, "tuning entry index out of range");
    auto it = fa::g_table.begin();
    std::advance(it, idx);
    *analysis = it->first.first;
    *key = it->first.second;
    *variant = it->second;
    return FA_OK;
}

int fa_tune_set(int analysis, int key, int variant)
{
    FA_API_LOCK;
    if (variant < 0)
    {
        fa::g_table.erase({analysis, key});
        fa::g_preset.erase({analysis, key});
    }
    else
    {
        fa::record_variant(analysis, key, variant);
        fa::g_preset.insert({analysis, key});
    }
    return FA_OK;
}

int fa_tune_reset(void)
{
    FA_API_LOCK;
    fa::g_table.clear();
    fa::g_preset.clear();
    return FA_OK;
}

int fa_tune_variants(int analysis, int key)
{
    FA_API_LOCK;
    switch (analysis)
    {
    case fa::TUNE_RDF_SELF:
        return key >= fa::kBatchTuneKey / 2 ? fa::batch_rdf_shape_count() : fa::rdf_row_variant_count();
    case fa::TUNE_BALL:
        return 2;
    case fa::TUNE_KNN:
        return fa::kKnnVariants;
    case fa::TUNE_KNN_REUSE:
        return 2;
    default:
        return fa::fail(FA_ERR_INVALID, "unknown tuned analysis");
    }
}

} // extern "C"
