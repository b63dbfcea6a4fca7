// freud_oracle.cpp -- CPU restatement of freud 3.5.0's pair-analysis path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU
// implementation in freud_amd/ is compared against.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product path never does.
//
// Every function restates the reference arithmetic op-for-op in IEEE single
// precision (built with -O2 -ffp-contract=off, no -march, so no FMA is formed,
// matching a Release x86-64 build of the reference).  Citations are to files of
// glotzerlab/freud 3.5.0.
//
// Pinning status (what checks each function; the checks live in tests/):
//   box wrap/fractional/absolute ... closed forms + known values from
//                                    tests/test_box_box.py
//   ball query (AABB, LinkCell) .... numpy brute force with the same per-formula
//                                    float32 arithmetic; hand-built known answers
//   ball query (CellQuery) ......... the reference's AABB == CellQuery bond-set
//                                    tests; numpy brute force of the ghost
//                                    arithmetic (p_j + shift) - q_i; closed-form
//                                    grid sizes; the reference's error cases
//   kNN ............................ brute-force k-smallest (also with r_min);
//                                    known answers incl. the reference's r_min
//                                    case; ties are broken by point index
//                                    (reference order is an unstable sort: tie
//                                    order is unpinned)
//   all_pairs ...................... the reference's all_pairs test (counts,
//                                    index order, distances) and hand-built
//                                    vectors wrap(q_i - p_j)
//   RDF counts / g(r) / N(r) ....... golden tests/integration/files/lj/rdf.npz;
//                                    np.histogram_bin_edges; n_r closed form
//   Y_lm / q_lm .................... scipy.special.sph_harm_y (closed form)
//   q_l, w_l, averages ............. GC validation files (atol 2e-5); perfect FCC
//                                    constants; Wigner-3j vs sympy exact values
//   weighted Steinhardt ............ the reference's re-weighted FCC test, plus
//                                    invariants tying it to the unweighted
//                                    branch (uniform weight, weight 0 = removed
//                                    bond, weight 2 = doubled bond, dominant
//                                    weight -> q_l = 1)
//   SolidLiquid, Cluster, Hexatic .. FCC known answer; scipy connected components;
//                                    closed forms
//   LocalDensity ................... the reference's known answer
//                                    (test_density_local_density.py:80-98), its
//                                    statistical bounds, and a float64 brute force
//   PMFTXY counts / pcf ............ the reference's known answers
//                                    (test_pmft.py: two-particle bins, nearest-
//                                    neighbour bins, orientation tests) and a
//                                    numpy brute force of the rotated bins
//   system `order` ................. parity unpinned bitwise (the reference's value
//                                    depends on TBB's thread partition); pinned to
//                                    tolerance by the FCC constant.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

// ---------------------------------------------------------------- threading
template<typename F> void parallel_for(size_t n, int nthreads, F&& fn)
{
    if (nthreads <= 1 || n < 64)
    {
        fn(size_t(0), n, 0);
        return;
    }
    std::vector<std::thread> pool;
    size_t const chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t)
    {
        size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e)
            break;
        pool.emplace_back([&, b, e, t]() { fn(b, e, t); });
    }
    for (auto& th : pool)
        th.join();
}

struct v3
{
    float x, y, z;
};
inline v3 add(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline v3 sub(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// dot order of freud/util/VectorMath.h:270-273
inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// ---------------------------------------------------------------- box (Box.h)
struct Box
{
    float L[3], lo[3], xy, xz, yz;
    bool is2D;
};

Box make_box(const float* p, int is2D)
{
    // Box.h:97-115 setL; m_hi = m_L / 2, m_lo = -m_hi
    Box b;
    b.is2D = is2D != 0;
    b.L[0] = p[0];
    b.L[1] = p[1];
    b.L[2] = b.is2D ? 0.0f : p[2];
    for (int k = 0; k < 3; ++k)
        b.lo[k] = -(b.L[k] / 2.0f);
    b.xy = p[3];
    b.xz = p[4];
    b.yz = p[5];
    return b;
}

// Box.h:243-255
inline v3 make_fractional(const Box& b, v3 v)
{
    v3 d = {v.x - b.lo[0], v.y - b.lo[1], v.z - b.lo[2]};
    d.x -= (b.xz - b.yz * b.xy) * v.z + b.xy * v.y;
    d.y -= b.yz * v.z;
    d.x = d.x / b.L[0];
    d.y = d.y / b.L[1];
    d.z = d.z / b.L[2];
    if (b.is2D)
        d.z = 0.0f;
    return d;
}

// Box.h:212-222
inline v3 make_absolute(const Box& b, v3 f)
{
    v3 v = {b.lo[0] + f.x * b.L[0], b.lo[1] + f.y * b.L[1], b.lo[2] + f.z * b.L[2]};
    v.x += b.xy * v.y + b.xz * v.z;
    v.y += b.yz * v.z;
    if (b.is2D)
        v.z = 0.0f;
    return v;
}

// utils.h:29-32
inline float modulus_positive(float a, float b) { return std::fmod(std::fmod(a, b) + b, b); }

// Box.h:307-329 (all three dimensions periodic: queries reject other boxes)
inline v3 wrap(const Box& b, v3 v)
{
    v3 f = make_fractional(b, v);
    f.x = modulus_positive(f.x, 1.0f);
    f.y = modulus_positive(f.y, 1.0f);
    f.z = modulus_positive(f.z, 1.0f);
    return make_absolute(b, f);
}

// Box.h:489-497
inline v3 nearest_plane_distance(const Box& b)
{
    v3 d;
    float const t = b.xy * b.yz - b.xz;
    d.x = b.L[0] / std::sqrt(1.0f + b.xy * b.xy + t * t);
    d.y = b.L[1] / std::sqrt(1.0f + b.yz * b.yz);
    d.z = b.L[2];
    return d;
}

float box_volume(const Box& b)
{
    // Box.h:198-205
    if (b.is2D)
        return b.L[0] * b.L[1];
    return b.L[0] * b.L[1] * b.L[2];
}

// NeighborQuery.h:496-564 (periodic in every dimension; 2D drops z images)
std::vector<v3> image_vectors(const Box& b)
{
    // lattice vectors Box.h:503-518
    v3 a = {b.L[0], 0.0f, 0.0f};
    v3 bb = {b.L[1] * b.xy, b.L[1], 0.0f};
    v3 c = {0.0f, 0.0f, 0.0f};
    if (!b.is2D)
        c = {b.L[2] * b.xz, b.L[2] * b.yz, b.L[2]};
    unsigned n_images = b.is2D ? 9 : 27;
    std::vector<v3> img(n_images);
    img[0] = {0.0f, 0.0f, 0.0f};
    unsigned n = 1;
    for (int i = -1; i <= 1 && n < n_images; ++i)
        for (int j = -1; j <= 1 && n < n_images; ++j)
            for (int k = -1; k <= 1 && n < n_images; ++k)
            {
                if (i == 0 && j == 0 && k == 0)
                    continue;
                if (k != 0 && b.is2D)
                    continue;
                float fi = float(i), fj = float(j), fk = float(k);
                v3 ia = {fi * a.x, fi * a.y, fi * a.z};
                v3 jb = {fj * bb.x, fj * bb.y, fj * bb.z};
                v3 kc = {fk * c.x, fk * c.y, fk * c.z};
                img[n++] = add(add(ia, jb), kc);
            }
    return img;
}

// ---------------------------------------------------------------- bonds
struct Bond
{
    uint32_t i, j;
    float d, w;
    v3 v;
};

// NeighborBond.h:80-95 less_as_tuple and :97-112 less_as_distance
inline bool less_tuple(const Bond& a, const Bond& b)
{
    if (a.i != b.i)
        return a.i < b.i;
    if (a.j != b.j)
        return a.j < b.j;
    if (a.w != b.w)
        return a.w < b.w;
    return a.d < b.d;
}
inline bool less_distance(const Bond& a, const Bond& b)
{
    if (a.i != b.i)
        return a.i < b.i;
    if (a.d != b.d)
        return a.d < b.d;
    if (a.j != b.j)
        return a.j < b.j;
    return a.w < b.w;
}

// ---------------------------------------------------------------- AABB tree
// A bounding-volume hierarchy built the way AABBTree.h:383-571 describes:
// recursive split at the centre of the longest axis, leaves of <= 16 points,
// nodes laid out in build (pre-)order with a `skip` count so that the query is
// a stackless walk (AABBQuery.cc:97-143).
struct Node
{
    float lo[3], hi[3];
    uint32_t skip;
    uint32_t first, count; // particles [first, first+count) of `order` (leaf)
    bool leaf;
};

struct AABBTree
{
    std::vector<Node> nodes;
    std::vector<uint32_t> order;

    void build(const std::vector<v3>& pts)
    {
        uint32_t n = pts.size();
        order.resize(n);
        std::iota(order.begin(), order.end(), 0u);
        nodes.clear();
        nodes.reserve(2 * (n / 8 + 1));
        build_node(pts, 0, n);
    }

    uint32_t build_node(const std::vector<v3>& pts, uint32_t start, uint32_t len)
    {
        Node nd;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t t = start; t < start + len; ++t)
        {
            v3 p = pts[order[t]];
            float c[3] = {p.x, p.y, p.z};
            for (int k = 0; k < 3; ++k)
            {
                lo[k] = std::min(lo[k], c[k]);
                hi[k] = std::max(hi[k], c[k]);
            }
        }
        std::memcpy(nd.lo, lo, sizeof lo);
        std::memcpy(nd.hi, hi, sizeof hi);
        nd.skip = 0;
        uint32_t my = nodes.size();
        if (len <= 16)
        {
            nd.leaf = true;
            nd.first = start;
            nd.count = len;
            nodes.push_back(nd);
            return 1;
        }
        nd.leaf = false;
        nd.first = nd.count = 0;
        nodes.push_back(nd);
        float ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
        int axis = (ext[0] > ext[1] && ext[0] > ext[2]) ? 0 : (ext[1] > ext[2] ? 1 : 2);
        float mid = (lo[axis] + hi[axis]) * 0.5f;
        uint32_t right = len;
        if (len != 2)
        {
            for (uint32_t t = 0; t < right; ++t)
            {
                v3 p = pts[order[start + t]];
                float c = axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
                if (!(c < mid))
                {
                    std::swap(order[start + t], order[start + right - 1]);
                    --right;
                    --t;
                }
            }
        }
        if (right == len)
            right = len - 1;
        if (right == 0)
            right = 1;
        uint32_t sl = build_node(pts, start, right);
        uint32_t sr = build_node(pts, start + right, len - right);
        nodes[my].skip = sl + sr;
        return sl + sr + 1;
    }

    // AABB.h:306-326: closest point of the box to the sphere centre
    static inline bool overlap(const Node& nd, v3 c, float r)
    {
        float px = std::min(std::max(c.x, nd.lo[0]), nd.hi[0]) - c.x;
        float py = std::min(std::max(c.y, nd.lo[1]), nd.hi[1]) - c.y;
        float pz = std::min(std::max(c.z, nd.lo[2]), nd.hi[2]) - c.z;
        return (px * px + py * py) + pz * pz < r * r;
    }
};

struct QueryParams
{
    int mode; // 1 ball, 2 nearest
    uint32_t k;
    float r_max, r_min, r_guess, scale;
    int exclude_ii;
};

// AABBQuery.cc:77-150: one point, one radius, all images; calls emit(j, r_ij, r_sq)
template<typename Emit>
void aabb_ball_point(const AABBTree& tree, const std::vector<v3>& pts, const std::vector<v3>& images,
                     v3 q, uint32_t qi, float r_max, float r_min, bool exclude_ii, bool is2D, Emit&& emit)
{
    float const r_max_sq = r_max * r_max;
    float const r_min_sq = r_min * r_min;
    if (is2D)
        q.z = 0.0f;
    for (const v3& img : images)
    {
        v3 const qimg = add(q, img);
        uint32_t node = 0, nn = tree.nodes.size();
        while (node < nn)
        {
            const Node& nd = tree.nodes[node];
            if (AABBTree::overlap(nd, qimg, r_max))
            {
                if (nd.leaf)
                {
                    for (uint32_t t = nd.first; t < nd.first + nd.count; ++t)
                    {
                        uint32_t j = tree.order[t];
                        if (exclude_ii && j == qi)
                            continue;
                        v3 pj = pts[j];
                        if (is2D)
                            pj.z = 0.0f;
                        v3 r = sub(pj, qimg);
                        float r_sq = dot(r, r);
                        if (r_sq < r_max_sq && r_sq >= r_min_sq)
                            emit(j, r, r_sq);
                    }
                }
            }
            else
            {
                node += nd.skip;
            }
            ++node;
        }
    }
}

// ---------------------------------------------------------------- LinkCell
struct LinkCell
{
    Box box;
    float w;
    uint32_t dim[3];
    std::vector<uint32_t> head, next;

    // LinkCell.cc:356-367
    void cell_coord(v3 p, int c[3]) const
    {
        v3 a = make_fractional(box, p);
        float f[3] = {a.x, a.y, a.z};
        for (int k = 0; k < 3; ++k)
        {
            unsigned int u = (unsigned int) std::floor(f[k] * float(dim[k]));
            c[k] = u % dim[k];
        }
    }
    uint32_t index(int x, int y, int z) const { return (uint32_t(z) * dim[1] + uint32_t(y)) * dim[0] + x; }
};

// LinkCell.cc:222-260, 278-314, 316-336
int build_linkcell(LinkCell& lc, const Box& b, const std::vector<v3>& pts, float cell_width)
{
    lc.box = b;
    uint32_t n = pts.size();
    if (cell_width == 0.0f)
    {
        unsigned desired = std::max(n / 10u, 1u);
        cell_width = std::cbrt(box_volume(b) / float(desired));
    }
    lc.w = cell_width;
    v3 L = nearest_plane_distance(b);
    lc.dim[0] = (unsigned) (L.x / cell_width);
    lc.dim[1] = (unsigned) (L.y / cell_width);
    lc.dim[2] = b.is2D ? 1u : (unsigned) (L.z / cell_width);
    for (int k = 0; k < 3; ++k)
        if (lc.dim[k] == 0)
            lc.dim[k] = 1;
    if ((cell_width * 2.0 > L.x) || (cell_width * 2.0 > L.y) || (!b.is2D && cell_width * 2.0 > L.z))
        return -3; // runtime_error: cell_width larger than half the box
    if (b.is2D)
        lc.dim[2] = 1;
    uint32_t nc = lc.dim[0] * lc.dim[1] * lc.dim[2];
    lc.head.assign(nc, 0xffffffffu);
    lc.next.assign(n, 0xffffffffu);
    for (uint32_t i = n - 1; i != 0xffffffffu; --i)
    {
        int c[3];
        lc.cell_coord(pts[i], c);
        uint32_t cell = lc.index(c[0], c[1], c[2]);
        lc.next[i] = lc.head[cell];
        lc.head[cell] = i;
    }
    return 0;
}

// Cells of Chebyshev shell `range` around `pc` (LinkCell.cc:58-213 enumerates the
// same set; the order only affects emission order, which is sorted later).
void shell_cells(const LinkCell& lc, const int pc[3], int range, std::vector<uint32_t>& out)
{
    out.clear();
    int zr = lc.box.is2D ? 0 : range;
    for (int dz = -zr; dz <= zr; ++dz)
        for (int dy = -range; dy <= range; ++dy)
            for (int dx = -range; dx <= range; ++dx)
            {
                int m = std::max(std::abs(dx), std::max(std::abs(dy), std::abs(dz)));
                if (m != range)
                    continue;
                int c[3] = {pc[0] + dx, pc[1] + dy, pc[2] + dz};
                for (int k = 0; k < 3; ++k)
                {
                    int d = int(lc.dim[k]);
                    c[k] %= d;
                    if (c[k] < 0)
                        c[k] += d;
                }
                out.push_back(lc.index(c[0], c[1], c[2]));
            }
}

// LinkCell.cc:496-573 (ball) -- emits in shell order
template<typename Emit>
void linkcell_ball_point(const LinkCell& lc, const std::vector<v3>& pts, v3 q, uint32_t qi, float r_max,
                         float r_min, bool exclude_ii, Emit&& emit)
{
    float const r_max_sq = r_max * r_max;
    float const r_min_sq = r_min * r_min;
    int const extra = (r_max == lc.w) ? 0 : 1; // LinkCell.h:321-328
    int pc[3];
    lc.cell_coord(q, pc);
    std::unordered_set<uint32_t> searched;
    std::vector<uint32_t> cells;
    for (int range = 0;; ++range)
    {
        if (range > 0 && float(range - extra) * lc.w > r_max)
            break;
        shell_cells(lc, pc, range, cells);
        for (uint32_t cell : cells)
        {
            if (!searched.insert(cell).second)
                continue;
            for (uint32_t j = lc.head[cell]; j != 0xffffffffu; j = lc.next[j])
            {
                if (exclude_ii && j == qi)
                    continue;
                v3 r = wrap(lc.box, sub(pts[j], q));
                float r_sq = dot(r, r);
                if (r_sq < r_max_sq && r_sq >= r_min_sq)
                    emit(j, r, r_sq);
            }
        }
        // every cell of the grid has been seen: nothing further can be found
        if (searched.size() == lc.head.size())
            break;
    }
}

// ---------------------------------------------------------------- CellQuery
// A cubic grid of width r_cut over the box's bounding parallelepiped plus one
// cell of padding, holding every point and the ghost copies generateGhosts
// makes for points within r_cut (fractionally) of a face.
struct CellQueryGrid
{
    float inv;
    v3 min;
    int n[3];
    std::vector<uint32_t> start; // cell c holds entries [start[c], start[c + 1])
    std::vector<v3> pos;         // point or ghost position
    std::vector<uint32_t> id;    // index of the real point
    std::vector<uint32_t> real;  // real points per cell

    // CellQuery.h:53-75 (cell_idx_xyz / getCellIdxSafe)
    bool cell_of(v3 p, int c[3]) const
    {
        c[0] = int(std::floor((p.x - min.x) * inv));
        c[1] = int(std::floor((p.y - min.y) * inv));
        c[2] = int(std::floor((p.z - min.z) * inv));
        return c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[0] < n[0] && c[1] < n[1] && c[2] < n[2];
    }
};

// CellQuery.h:136-160 (setupGrid), CellQuery.cc:56-113 (buildGrid) and
// CellQuery.h:221-298 (generateGhosts: faces, then edges, then the corner;
// each shift sums +-a, +-b, +-c in that order onto zero).  Entries of a cell
// are kept real points first, then ghosts (placeParticlesInSortedOrder).
void build_cellquery(CellQueryGrid& g, const Box& b, const std::vector<v3>& pts, float r_cut)
{
    g.inv = 1.0f / r_cut;
    float const Lx = b.L[0], Ly = b.L[1], Lz = b.L[2];
    float const w_x = Lx + (Ly * std::abs(b.xy)) + (Lz * std::abs(b.xz));
    float const w_y = Ly + (Lz * std::abs(b.yz));
    float const w_z = Lz;
    g.n[0] = int(w_x * g.inv) + 3;
    g.n[1] = int(w_y * g.inv) + 3;
    g.n[2] = int(w_z * g.inv) + 3;
    float const box_min_x = (Ly * std::min(0.0f, b.xy)) + (Lz * std::min(0.0f, b.xz));
    float const box_min_y = Lz * std::min(0.0f, b.yz);
    v3 const a0 = make_absolute(b, {0.0f, 0.0f, 0.0f});
    g.min = add({box_min_x - r_cut, box_min_y - r_cut, -r_cut}, a0);
    v3 const npd = nearest_plane_distance(b);
    v3 const frc = {r_cut / npd.x, r_cut / npd.y, r_cut / npd.z};
    v3 const la = {Lx, 0.0f, 0.0f};
    v3 const lb = {Ly * b.xy, Ly, 0.0f};
    v3 const lc = b.is2D ? v3{0.0f, 0.0f, 0.0f} : v3{Lz * b.xz, Lz * b.yz, Lz};
    size_t const ncell = size_t(g.n[0]) * g.n[1] * g.n[2];
    std::vector<uint32_t> real_count(ncell, 0), count(ncell, 0);
    struct Entry
    {
        uint32_t cell;
        v3 p;
        uint32_t id;
        bool ghost;
    };
    std::vector<Entry> entries;
    entries.reserve(pts.size());
    auto cell_index = [&](const int c[3]) { return uint32_t((c[2] * g.n[1] + c[1]) * g.n[0] + c[0]); };
    for (size_t i = 0; i < pts.size(); ++i)
    {
        v3 const p = pts[i];
        v3 const f = make_fractional(b, p);
        int const dx = int(f.x <= frc.x) - int(double(f.x) >= 1.0 - double(frc.x));
        int const dy = int(f.y <= frc.y) - int(double(f.y) >= 1.0 - double(frc.y));
        int dz = int(f.z <= frc.z) - int(double(f.z) >= 1.0 - double(frc.z));
        if (b.is2D)
            dz = 0;
        std::vector<v3> shifts;
        auto site = [&](int si, int sj, int sk) {
            v3 sh = {0.0f, 0.0f, 0.0f};
            if (si != 0)
                sh = add(sh, si == 1 ? la : v3{-la.x, -la.y, -la.z});
            if (sj != 0)
                sh = add(sh, sj == 1 ? lb : v3{-lb.x, -lb.y, -lb.z});
            if (sk != 0)
                sh = add(sh, sk == 1 ? lc : v3{-lc.x, -lc.y, -lc.z});
            shifts.push_back(sh);
        };
        if (dx != 0 || dy != 0 || dz != 0)
        {
            if (dx != 0)
                site(dx, 0, 0);
            if (dy != 0)
                site(0, dy, 0);
            if (dz != 0)
                site(0, 0, dz);
            if (dx != 0 && dy != 0)
                site(dx, dy, 0);
            if (dx != 0 && dz != 0)
                site(dx, 0, dz);
            if (dy != 0 && dz != 0)
                site(0, dy, dz);
            if (dx != 0 && dy != 0 && dz != 0)
                site(dx, dy, dz);
        }
        int c[3];
        for (const v3& sh : shifts)
        {
            v3 const gp = add(p, sh);
            if (g.cell_of(gp, c))
            {
                entries.push_back({cell_index(c), gp, uint32_t(i), true});
                ++count[cell_index(c)];
            }
        }
        if (g.cell_of(p, c))
        {
            entries.push_back({cell_index(c), p, uint32_t(i), false});
            ++count[cell_index(c)];
            ++real_count[cell_index(c)];
        }
    }
    g.start.assign(ncell + 1, 0);
    for (size_t c = 0; c < ncell; ++c)
        g.start[c + 1] = g.start[c] + count[c];
    g.real = real_count;
    std::vector<uint32_t> real_fill(ncell, 0), ghost_fill(ncell, 0);
    g.pos.resize(entries.size());
    g.id.resize(entries.size());
    for (const Entry& e : entries)
    {
        uint32_t const at = e.ghost ? g.start[e.cell] + real_count[e.cell] + ghost_fill[e.cell]++
                                    : g.start[e.cell] + real_fill[e.cell]++;
        g.pos[at] = e.p;
        g.id[at] = e.id;
    }
}

// CellIterator.h:94-187 (CellQueryBallIterator): the 3 x 3 x 3 cells around the
// query point's cell (none if that cell is outside the grid), r = entry - q.
template<typename Emit>
void cellquery_ball_point(const CellQueryGrid& g, v3 q, uint32_t qi, float r_max, float r_min, bool exclude_ii,
                          Emit&& emit)
{
    float const r_max_sq = r_max * r_max;
    float const r_min_sq = r_min * r_min;
    int c[3];
    if (!g.cell_of(q, c))
        return;
    for (int dz = -1; dz <= 1; ++dz)
    {
        int const z = c[2] + dz;
        if (z < 0 || z >= g.n[2])
            continue;
        for (int dy = -1; dy <= 1; ++dy)
        {
            int const y = c[1] + dy;
            if (y < 0 || y >= g.n[1])
                continue;
            int const x0 = std::max(c[0] - 1, 0), x1 = std::min(c[0] + 1, g.n[0] - 1);
            uint32_t const row = uint32_t((z * g.n[1] + y) * g.n[0]);
            for (uint32_t k = g.start[row + x0]; k < g.start[row + x1 + 1]; ++k)
            {
                uint32_t const j = g.id[k];
                if (exclude_ii && j == qi)
                    continue;
                v3 const r = sub(g.pos[k], q);
                float const r_sq = dot(r, r);
                if (r_sq < r_max_sq && r_sq >= r_min_sq)
                    emit(j, r, r_sq);
            }
        }
    }
}

thread_local char g_err[256];

// CellQuery::validateQueryArgs (CellQuery.h:175-204) after the base class's
// mode inference and ball / nearest checks (NeighborQuery.h:195-228).
int validate_cellquery(const Box& b, QueryParams* qp)
{
    if (qp->mode == 0)
    {
        if (qp->k != 0xffffffffu)
            qp->mode = 2;
        else if (qp->r_max != -1.0f)
            qp->mode = 1;
    }
    if (qp->mode == 1)
    {
        if (qp->r_max == -1.0f)
        {
            std::snprintf(g_err, sizeof g_err, "You must set r_max in the query arguments when performing ball queries.");
            return -2;
        }
        if (qp->k != 0xffffffffu)
        {
            std::snprintf(g_err, sizeof g_err, "You cannot set num_neighbors in the query arguments when performing ball queries.");
            return -2;
        }
    }
    else if (qp->mode == 2)
    {
        if (qp->k == 0xffffffffu)
        {
            std::snprintf(g_err, sizeof g_err, "You must set num_neighbors in the query arguments when performing number of neighbor queries.");
            return -2;
        }
        std::snprintf(g_err, sizeof g_err,
                      "CellQuery only supports ball queries (r_max), not nearest queries (num_neighbors). Use "
                      "AABBQuery for nearest neighbor queries.");
        return -2;
    }
    else
    {
        std::snprintf(g_err, sizeof g_err, "Unknown mode");
        return -2;
    }
    v3 const npd = nearest_plane_distance(b);
    if (npd.x <= qp->r_max * 2.0f || npd.y <= qp->r_max * 2.0f || (!b.is2D && npd.z <= qp->r_max * 2.0f))
    {
        std::snprintf(g_err, sizeof g_err, "The CellQuery r_max is too large for this box.");
        return -2;
    }
    if (qp->r_max <= 0)
    {
        std::snprintf(g_err, sizeof g_err, "r_max must be positive.");
        return -1;
    }
    if (qp->r_max <= qp->r_min)
    {
        std::snprintf(g_err, sizeof g_err, "r_max must be greater than r_min.");
        return -1;
    }
    return 0;
}

inline bool knn_less(const Bond& a, const Bond& b)
{
    // distance first; ties broken by point index (reference tie order is unpinned)
    if (a.d != b.d)
        return a.d < b.d;
    return a.j < b.j;
}

struct NlistOut
{
    uint32_t* ij;
    float* dist;
    float* w;
    float* vec;
    uint64_t n;
};

void fill_nlist(std::vector<Bond>& bonds, int sort_by_distance, NlistOut* out)
{
    // NeighborQuery.h:457-478
    if (sort_by_distance)
        std::sort(bonds.begin(), bonds.end(), less_distance);
    else
        std::sort(bonds.begin(), bonds.end(), less_tuple);
    uint64_t n = bonds.size();
    out->n = n;
    out->ij = (uint32_t*) std::malloc(sizeof(uint32_t) * 2 * (n ? n : 1));
    out->dist = (float*) std::malloc(sizeof(float) * (n ? n : 1));
    out->w = (float*) std::malloc(sizeof(float) * (n ? n : 1));
    out->vec = (float*) std::malloc(sizeof(float) * 3 * (n ? n : 1));
    for (uint64_t b = 0; b < n; ++b)
    {
        out->ij[2 * b] = bonds[b].i;
        out->ij[2 * b + 1] = bonds[b].j;
        out->dist[b] = bonds[b].d;
        out->w[b] = bonds[b].w;
        out->vec[3 * b] = bonds[b].v.x;
        out->vec[3 * b + 1] = bonds[b].v.y;
        out->vec[3 * b + 2] = bonds[b].v.z;
    }
}

std::vector<v3> to_v3(const float* p, uint32_t n)
{
    std::vector<v3> v(n);
    for (uint32_t i = 0; i < n; ++i)
        v[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    return v;
}

} // namespace

// ============================================================== C interface
extern "C" {

const char* orc_last_error() { return g_err; }

void orc_nlist_free(NlistOut* o)
{
    std::free(o->ij);
    std::free(o->dist);
    std::free(o->w);
    std::free(o->vec);
    o->ij = nullptr;
    o->dist = o->w = o->vec = nullptr;
    o->n = 0;
}

void orc_box_op(const float* boxp, int is2D, int op, const float* v, uint64_t n, float* out)
{
    // op 0 make_fractional, 1 make_absolute, 2 wrap
    Box b = make_box(boxp, is2D);
    for (uint64_t i = 0; i < n; ++i)
    {
        v3 a = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
        v3 r = op == 0 ? make_fractional(b, a) : (op == 1 ? make_absolute(b, a) : wrap(b, a));
        out[3 * i] = r.x;
        out[3 * i + 1] = r.y;
        out[3 * i + 2] = r.z;
    }
}

void orc_nearest_plane_distance(const float* boxp, int is2D, float* out)
{
    Box b = make_box(boxp, is2D);
    v3 d = nearest_plane_distance(b);
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

// NeighborQuery.h:195-292: mode inference, defaults and validation.
// returns 0 ok; -1 invalid_argument (ValueError); -2 runtime_error
int orc_validate_query(const float* boxp, int is2D, uint32_t n_points, QueryParams* qp)
{
    Box b = make_box(boxp, is2D);
    if (qp->mode == 0)
    {
        if (qp->k != 0xffffffffu)
            qp->mode = 2;
        else if (qp->r_max != -1.0f)
            qp->mode = 1;
    }
    if (qp->mode == 1)
    {
        if (qp->r_max == -1.0f)
        {
            std::snprintf(g_err, sizeof g_err, "You must set r_max in the query arguments when performing ball queries.");
            return -2;
        }
        if (qp->k != 0xffffffffu)
        {
            std::snprintf(g_err, sizeof g_err, "You cannot set num_neighbors in the query arguments when performing ball queries.");
            return -2;
        }
    }
    else if (qp->mode == 2)
    {
        if (qp->k == 0xffffffffu)
        {
            std::snprintf(g_err, sizeof g_err, "You must set num_neighbors in the query arguments when performing number of neighbor queries.");
            return -2;
        }
        if (qp->r_max == -1.0f)
            qp->r_max = INFINITY;
        if (qp->scale == -1.0f)
            qp->scale = 1.1f;
        else if (qp->scale <= 1.0f)
        {
            std::snprintf(g_err, sizeof g_err, "The scale query argument must be greater than 1.");
            return -2;
        }
        if (qp->r_guess == -1.0f)
        {
            float r_guess = std::cbrt((3.0f * float(qp->k) * box_volume(b)) / (4.0f * float(M_PI) * float(n_points)));
            v3 npd = nearest_plane_distance(b);
            float mpd = std::min(npd.x, npd.y);
            if (!b.is2D)
                mpd = std::min(mpd, npd.z);
            qp->r_guess = std::min(r_guess, mpd / 2.0f);
        }
        qp->r_guess = std::min(qp->r_guess, qp->r_max);
    }
    else
    {
        std::snprintf(g_err, sizeof g_err, "Unknown mode");
        return -2;
    }
    if (qp->r_max <= 0)
    {
        std::snprintf(g_err, sizeof g_err, "NeighborQuery requires r_max to be positive.");
        return -1;
    }
    if (qp->r_max <= qp->r_min)
    {
        std::snprintf(g_err, sizeof g_err, "NeighborQuery requires that r_max must be greater than r_min.");
        return -1;
    }
    return 0;
}

// Full query -> sorted NeighborList.  nq_type 0 = AABBQuery, 1 = LinkCell,
// 2 = CellQuery (ball queries only).
int orc_query(const float* boxp, int is2D, const float* pts_, uint32_t np, const float* qpts_, uint32_t nq,
              int nq_type, float cell_width, QueryParams qp, int sort_by_distance, int nthreads,
              NlistOut* out)
{
    out->ij = nullptr;
    out->dist = out->w = out->vec = nullptr;
    out->n = 0;
    Box b = make_box(boxp, is2D);
    int rc = nq_type == 2 ? validate_cellquery(b, &qp) : orc_validate_query(boxp, is2D, np, &qp);
    if (rc)
        return rc;
    std::vector<v3> pts = to_v3(pts_, np), qpts = to_v3(qpts_, nq);
    v3 npd = nearest_plane_distance(b);
    std::vector<std::vector<Bond>> per_thread(std::max(nthreads, 1));

    if (nq_type == 2)
    {
        CellQueryGrid grid;
        build_cellquery(grid, b, pts, qp.r_max);
        parallel_for(nq, nthreads, [&](size_t beg, size_t end, int t) {
            auto& local = per_thread[t];
            for (size_t i = beg; i < end; ++i)
                cellquery_ball_point(grid, qpts[i], uint32_t(i), qp.r_max, qp.r_min, qp.exclude_ii,
                                     [&](uint32_t j, v3 r, float r_sq) {
                                         local.push_back({uint32_t(i), j, std::sqrt(r_sq), 1.0f, r});
                                     });
        });
    }
    else if (nq_type == 0)
    {
        if (qp.mode == 1)
        {
            // NeighborQuery.h:504-509
            if (npd.x <= qp.r_max * 2.0 || npd.y <= qp.r_max * 2.0 || (!b.is2D && npd.z <= qp.r_max * 2.0))
            {
                std::snprintf(g_err, sizeof g_err, "The AABBQuery r_max is too large for this box.");
                return -2;
            }
        }
        std::vector<v3> tpts = pts;
        if (b.is2D)
            for (auto& p : tpts)
                p.z = 0.0f;
        AABBTree tree;
        tree.build(tpts);
        std::vector<v3> images = image_vectors(b);
        float min_pd = std::min(npd.x, npd.y), max_pd = std::max(npd.x, npd.y);
        if (!b.is2D)
        {
            min_pd = std::min(min_pd, npd.z);
            max_pd = std::max(max_pd, npd.z);
        }
        parallel_for(nq, nthreads, [&](size_t beg, size_t end, int t) {
            auto& local = per_thread[t];
            std::vector<Bond> cur;
            std::map<uint32_t, Bond> min_map;
            std::unordered_set<uint32_t> below;
            for (size_t i = beg; i < end; ++i)
            {
                v3 q = qpts[i];
                if (qp.mode == 1)
                {
                    aabb_ball_point(tree, pts, images, q, i, qp.r_max, qp.r_min, qp.exclude_ii, b.is2D,
                                    [&](uint32_t j, v3 r, float r_sq) {
                                        local.push_back({uint32_t(i), j, std::sqrt(r_sq), 1.0f, r});
                                    });
                    continue;
                }
                // AABBQuery.cc:152-281 expanding ball search
                float r_cur = qp.r_guess;
                bool extended = false;
                while (true)
                {
                    cur.clear();
                    min_map.clear();
                    below.clear();
                    aabb_ball_point(tree, pts, images, q, i, std::min(r_cur, qp.r_max), 0.0f, qp.exclude_ii,
                                    b.is2D, [&](uint32_t j, v3 r, float r_sq) {
                                        Bond nb = {uint32_t(i), j, std::sqrt(r_sq), 1.0f, r};
                                        if (extended)
                                        {
                                            auto it = min_map.find(j);
                                            if (it == min_map.end() || it->second.d > nb.d)
                                            {
                                                min_map[j] = nb;
                                                if (nb.d < qp.r_min)
                                                    below.insert(j);
                                            }
                                        }
                                        else if (nb.d >= qp.r_min)
                                            cur.push_back(nb);
                                    });
                    r_cur *= qp.scale;
                    if (cur.size() >= qp.k)
                    {
                        std::sort(cur.begin(), cur.end(), knn_less);
                        break;
                    }
                    if (r_cur >= qp.r_max || r_cur >= max_pd || (min_map.size() - below.size()) >= qp.k)
                    {
                        for (auto& kv : min_map)
                            if (kv.second.d >= qp.r_min)
                                cur.push_back(kv.second);
                        std::sort(cur.begin(), cur.end(), knn_less);
                        break;
                    }
                    if (r_cur > min_pd / 2)
                        extended = true;
                }
                for (size_t m = 0; m < qp.k && m < cur.size(); ++m)
                {
                    if (cur[m].d > qp.r_max)
                        break;
                    local.push_back(cur[m]);
                }
            }
        });
    }
    else
    {
        LinkCell lc;
        int rc2 = build_linkcell(lc, b, pts, cell_width);
        if (rc2)
        {
            std::snprintf(g_err, sizeof g_err, "Cannot generate a cell list where cell_width is larger than half the box.");
            return -2;
        }
        float min_pd = std::min(npd.x, npd.y);
        if (!b.is2D)
            min_pd = std::min(min_pd, npd.z);
        parallel_for(nq, nthreads, [&](size_t beg, size_t end, int t) {
            auto& local = per_thread[t];
            std::vector<Bond> cur;
            for (size_t i = beg; i < end; ++i)
            {
                v3 q = qpts[i];
                if (qp.mode == 1)
                {
                    linkcell_ball_point(lc, pts, q, i, qp.r_max, qp.r_min, qp.exclude_ii,
                                        [&](uint32_t j, v3 r, float r_sq) {
                                            local.push_back({uint32_t(i), j, std::sqrt(r_sq), 1.0f, r});
                                        });
                    continue;
                }
                // LinkCell.cc:575-679: every cell within max_range shells
                unsigned max_range = unsigned(std::ceil(min_pd / (2 * lc.w))) + 1;
                float const r_max_sq = qp.r_max * qp.r_max, r_min_sq = qp.r_min * qp.r_min;
                cur.clear();
                int pc[3];
                lc.cell_coord(q, pc);
                std::unordered_set<uint32_t> searched;
                std::vector<uint32_t> cells;
                for (int range = 0; range < int(max_range); ++range)
                {
                    shell_cells(lc, pc, range, cells);
                    for (uint32_t cell : cells)
                    {
                        if (!searched.insert(cell).second)
                            continue;
                        for (uint32_t j = lc.head[cell]; j != 0xffffffffu; j = lc.next[j])
                        {
                            if (qp.exclude_ii && j == i)
                                continue;
                            v3 r = wrap(b, sub(pts[j], q));
                            float r_sq = dot(r, r);
                            if (r_sq < r_max_sq && r_sq >= r_min_sq)
                                cur.push_back({uint32_t(i), j, std::sqrt(r_sq), 1.0f, r});
                        }
                    }
                }
                std::sort(cur.begin(), cur.end(), knn_less);
                for (size_t m = 0; m < qp.k && m < cur.size(); ++m)
                {
                    if (cur[m].d > qp.r_max)
                        break;
                    local.push_back(cur[m]);
                }
            }
        });
    }
    std::vector<Bond> all;
    for (auto& v : per_thread)
        all.insert(all.end(), v.begin(), v.end());
    fill_nlist(all, sort_by_distance, out);
    return 0;
}

// CellQuery grid for radius r_cut over a point set (CellQuery.h:136-160,
// CellQuery.cc:56-113): dims, lower corner, inverse width; when counts is not
// null, total and real occupancy per cell (ncell = nx*ny*nz entries, index
// (z * ny + y) * nx + x).  Returns the number of entries, points plus ghosts.
uint32_t orc_cellquery_grid(const float* boxp, int is2D, const float* pts_, uint32_t np, float r_cut, int* n,
                            float* mn, float* inv, uint32_t* counts, uint32_t* counts_real)
{
    Box b = make_box(boxp, is2D);
    CellQueryGrid g;
    build_cellquery(g, b, to_v3(pts_, np), r_cut);
    for (int k = 0; k < 3; ++k)
        n[k] = g.n[k];
    mn[0] = g.min.x;
    mn[1] = g.min.y;
    mn[2] = g.min.z;
    *inv = g.inv;
    size_t const ncell = size_t(g.n[0]) * g.n[1] * g.n[2];
    if (counts)
        for (size_t c = 0; c < ncell; ++c)
        {
            counts[c] = g.start[c + 1] - g.start[c];
            counts_real[c] = g.real[c];
        }
    return uint32_t(g.pos.size());
}

// NeighborList.cc:84-128 all_pairs (note the reference's vector is wrap(q_i - p_j))
int orc_all_pairs(const float* boxp, int is2D, const float* pts_, uint32_t np, const float* qpts_, uint32_t nq,
                  int exclude_ii, NlistOut* out)
{
    Box b = make_box(boxp, is2D);
    std::vector<v3> pts = to_v3(pts_, np), qpts = to_v3(qpts_, nq);
    std::vector<Bond> all;
    for (uint32_t i = 0; i < nq; ++i)
        for (uint32_t j = 0; j < np; ++j)
        {
            if (exclude_ii && i == j)
                continue;
            v3 dr = wrap(b, sub(qpts[i], pts[j]));
            all.push_back({i, j, std::sqrt(dot(dr, dr)), 1.0f, dr});
        }
    // already in (i, j) order; fill without re-sorting
    out->n = all.size();
    uint64_t n = all.size();
    out->ij = (uint32_t*) std::malloc(sizeof(uint32_t) * 2 * (n ? n : 1));
    out->dist = (float*) std::malloc(sizeof(float) * (n ? n : 1));
    out->w = (float*) std::malloc(sizeof(float) * (n ? n : 1));
    out->vec = (float*) std::malloc(sizeof(float) * 3 * (n ? n : 1));
    for (uint64_t k = 0; k < n; ++k)
    {
        out->ij[2 * k] = all[k].i;
        out->ij[2 * k + 1] = all[k].j;
        out->dist[k] = all[k].d;
        out->w[k] = all[k].w;
        out->vec[3 * k] = all[k].v.x;
        out->vec[3 * k + 1] = all[k].v.y;
        out->vec[3 * k + 2] = all[k].v.z;
    }
    return 0;
}

// ---------------------------------------------------------------- RDF
// Histogram.h:123-174 RegularAxis.  Returns the bin or -1 for overflow.
static inline long regular_bin(float value, float mn, float mx, float inv_width, size_t nbins)
{
    if (value < mn || value >= mx)
        return -1;
    float const val = (value - mn) * inv_width;
    size_t bin = (size_t) (val);
    if (bin == nbins)
        return long(bin - 1);
    return long(bin);
}

void orc_regular_axis(uint32_t bins, float r_min, float r_max, float* edges, float* centers, float* width,
                      float* inv_width)
{
    float const w = (r_max - r_min) / float(bins);
    *width = w;
    *inv_width = 1.0f / w;
    for (uint32_t i = 0; i <= bins; ++i)
        edges[i] = r_min + float(i) * w;
    for (uint32_t i = 0; i < bins; ++i)
        centers[i] = (edges[i] + edges[i + 1]) / 2.0f;
}

// Bin bond distances into counts (counts are added to, not cleared): RDF.cc:101-110
void orc_rdf_bin(const float* dist, uint64_t nb, uint32_t bins, float r_min, float r_max, uint32_t* counts)
{
    float const w = (r_max - r_min) / float(bins);
    float const inv = 1.0f / w;
    for (uint64_t b = 0; b < nb; ++b)
    {
        long k = regular_bin(dist[b], r_min, r_max, inv, bins);
        if (k >= 0)
            counts[k] += 1;
    }
}

// Fused query + binning without materialising bonds (the reference's default RDF path:
// loopOverNeighbors over per-point AABB/LinkCell iterators with thread-local histograms)
int orc_rdf_accumulate(const float* boxp, int is2D, const float* pts_, uint32_t np, const float* qpts_, uint32_t nq,
                       int nq_type, float cell_width, float q_r_max, float q_r_min, int exclude_ii, uint32_t bins,
                       float r_min, float r_max, int nthreads, uint32_t* counts)
{
    Box b = make_box(boxp, is2D);
    QueryParams qp = {1, 0xffffffffu, q_r_max, q_r_min, -1.0f, -1.0f, exclude_ii};
    int rc = nq_type == 2 ? validate_cellquery(b, &qp) : orc_validate_query(boxp, is2D, np, &qp);
    if (rc)
        return rc;
    v3 npd = nearest_plane_distance(b);
    float const w = (r_max - r_min) / float(bins);
    float const inv = 1.0f / w;
    std::vector<v3> pts = to_v3(pts_, np), qpts = to_v3(qpts_, nq);
    int nt = std::max(nthreads, 1);
    std::vector<std::vector<uint32_t>> local(nt, std::vector<uint32_t>(bins, 0));
    if (nq_type == 2)
    {
        CellQueryGrid grid;
        build_cellquery(grid, b, pts, qp.r_max);
        parallel_for(nq, nthreads, [&](size_t beg, size_t end, int t) {
            auto& h = local[t];
            for (size_t i = beg; i < end; ++i)
                cellquery_ball_point(grid, qpts[i], uint32_t(i), qp.r_max, qp.r_min, qp.exclude_ii,
                                     [&](uint32_t, v3, float r_sq) {
                                         long k = regular_bin(std::sqrt(r_sq), r_min, r_max, inv, bins);
                                         if (k >= 0)
                                             h[k] += 1;
                                     });
        });
    }
    else if (nq_type == 0)
    {
        if (npd.x <= qp.r_max * 2.0 || npd.y <= qp.r_max * 2.0 || (!b.is2D && npd.z <= qp.r_max * 2.0))
        {
            std::snprintf(g_err, sizeof g_err, "The AABBQuery r_max is too large for this box.");
            return -2;
        }
        std::vector<v3> tpts = pts;
        if (b.is2D)
            for (auto& p : tpts)
                p.z = 0.0f;
        AABBTree tree;
        tree.build(tpts);
        std::vector<v3> images = image_vectors(b);
        parallel_for(nq, nthreads, [&](size_t beg, size_t end, int t) {
            auto& h = local[t];
            for (size_t i = beg; i < end; ++i)
                aabb_ball_point(tree, pts, images, qpts[i], i, qp.r_max, qp.r_min, qp.exclude_ii, b.is2D,
                                [&](uint32_t, v3, float r_sq) {
                                    long k = regular_bin(std::sqrt(r_sq), r_min, r_max, inv, bins);
                                    if (k >= 0)
                                        h[k] += 1;
                                });
        });
    }
    else
    {
        LinkCell lc;
        if (build_linkcell(lc, b, pts, cell_width))
        {
            std::snprintf(g_err, sizeof g_err, "Cannot generate a cell list where cell_width is larger than half the box.");
            return -2;
        }
        parallel_for(nq, nthreads, [&](size_t beg, size_t end, int t) {
            auto& h = local[t];
            for (size_t i = beg; i < end; ++i)
                linkcell_ball_point(lc, pts, qpts[i], i, qp.r_max, qp.r_min, qp.exclude_ii,
                                    [&](uint32_t, v3, float r_sq) {
                                        long k = regular_bin(std::sqrt(r_sq), r_min, r_max, inv, bins);
                                        if (k >= 0)
                                            h[k] += 1;
                                    });
        });
    }
    for (auto& h : local)
        for (uint32_t k = 0; k < bins; ++k)
            counts[k] += h[k];
    return 0;
}

// RDF.cc:25-64 (shell volumes) and :73-99 (reduce)
void orc_rdf_reduce(const float* boxp, int is2D, const uint32_t* counts, uint32_t bins, float r_min, float r_max,
                    uint32_t n_points, uint32_t n_query_points, uint32_t n_frames, int finite_size, float* rdf,
                    float* n_r)
{
    Box b = make_box(boxp, is2D);
    float const w = (r_max - r_min) / float(bins);
    std::vector<float> edges(bins + 1);
    for (uint32_t i = 0; i <= bins; ++i)
        edges[i] = r_min + float(i) * w;
    float const volume_prefactor = (float(4.0) / float(3.0)) * M_PI;
    std::vector<float> vol(bins);
    for (uint32_t i = 0; i < bins; ++i)
    {
        float const r = edges[i], nr = edges[i + 1];
        if (b.is2D)
            vol[i] = M_PI * (nr * nr - r * r);
        else
            vol[i] = volume_prefactor * (nr * nr * nr - r * r * r);
    }
    float const nqp = float(n_query_points);
    float number_density = nqp / box_volume(b);
    if (finite_size)
        number_density *= float(n_query_points - 1) / float(n_query_points);
    float const npf = float(n_points);
    float const nf = float(n_frames);
    float prefactor = float(1.0) / (npf * number_density * nf);
    for (uint32_t i = 0; i < bins; ++i)
        rdf[i] = float(counts[i]) * prefactor / vol[i];
    prefactor = float(1.0) / (nqp * float(n_frames));
    n_r[0] = float(counts[0]) * prefactor;
    for (uint32_t i = 1; i < bins; ++i)
        n_r[i] = n_r[i - 1] + float(counts[i]) * prefactor;
}

} // extern "C"

namespace {
// ---------------------------------------------------------------- Y_lm
// Orthonormal spherical harmonics with the Condon-Shortley phase, the convention of
// scipy.special.sph_harm_y and of the values Steinhardt.cc:31-52 stores.  The
// reference evaluates them with the absent fsph submodule; this is an independent
// evaluation by the standard normalised associated-Legendre recurrence:
//   P(m,m)   = -sqrt((2m+1)/(2m)) sin(t) P(m-1,m-1),  P(0,0) = 1/sqrt(4 pi)
//   P(m+1,m) = sqrt(2m+3) cos(t) P(m,m)
//   P(l,m)   = a(l,m) (cos(t) P(l-1,m) - b(l,m) P(l-2,m))
// with a = sqrt((4l^2-1)/(l^2-m^2)), b = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1)).
// Output layout per l: m = 0..l, then -1..-l (Steinhardt.cc:44-49 / order.py:563).
template<typename T>
void ylm_all(int lmax, T ct, T st, T cphi, T sphi, std::vector<T>& P, std::vector<std::complex<T>>& eimphi)
{
    // P indexed [l*(l+1)/2 + m]
    P.assign((lmax + 1) * (lmax + 2) / 2, T(0));
    auto idx = [](int l, int m) { return l * (l + 1) / 2 + m; };
    P[0] = T(0.28209479177387814347403972578038629292202531466449942842204286);
    for (int m = 1; m <= lmax; ++m)
        P[idx(m, m)] = -std::sqrt(T(2 * m + 1) / T(2 * m)) * st * P[idx(m - 1, m - 1)];
    for (int m = 0; m < lmax; ++m)
        P[idx(m + 1, m)] = std::sqrt(T(2 * m + 3)) * ct * P[idx(m, m)];
    for (int m = 0; m <= lmax; ++m)
        for (int l = m + 2; l <= lmax; ++l)
        {
            T a = std::sqrt(T(4 * l * l - 1) / T(l * l - m * m));
            T b = std::sqrt(T((l - 1) * (l - 1) - m * m) / T(4 * (l - 1) * (l - 1) - 1));
            P[idx(l, m)] = a * (ct * P[idx(l - 1, m)] - b * P[idx(l - 2, m)]);
        }
    eimphi.assign(lmax + 1, std::complex<T>(1, 0));
    std::complex<T> e1(cphi, sphi);
    for (int m = 1; m <= lmax; ++m)
        eimphi[m] = eimphi[m - 1] * e1;
}

template<typename T> void ylm_l(int l, const std::vector<T>& P, const std::vector<std::complex<T>>& e, std::complex<T>* Y)
{
    int const base = l * (l + 1) / 2;
    for (int m = 0; m <= l; ++m)
        Y[m] = P[base + m] * e[m];
    for (int m = 1; m <= l; ++m)
        Y[l + m] = ((m & 1) ? T(-1) : T(1)) * std::conj(Y[m]);
}
} // namespace

extern "C" {

// Evaluate Y_l^m for a list of (theta, phi) (test hook against scipy)
void orc_ylm(int l, const double* theta, const double* phi, uint64_t n, int f64, double* out_re, double* out_im)
{
    for (uint64_t i = 0; i < n; ++i)
    {
        if (f64)
        {
            std::vector<double> P;
            std::vector<std::complex<double>> e;
            ylm_all<double>(l, std::cos(theta[i]), std::sin(theta[i]), std::cos(phi[i]), std::sin(phi[i]), P, e);
            std::vector<std::complex<double>> Y(2 * l + 1);
            ylm_l<double>(l, P, e, Y.data());
            for (int k = 0; k < 2 * l + 1; ++k)
            {
                out_re[i * (2 * l + 1) + k] = Y[k].real();
                out_im[i * (2 * l + 1) + k] = Y[k].imag();
            }
        }
        else
        {
            float t = float(theta[i]), p = float(phi[i]);
            std::vector<float> P;
            std::vector<std::complex<float>> e;
            ylm_all<float>(l, std::cos(t), std::sin(t), std::cos(p), std::sin(p), P, e);
            std::vector<std::complex<float>> Y(2 * l + 1);
            ylm_l<float>(l, P, e, Y.data());
            for (int k = 0; k < 2 * l + 1; ++k)
            {
                out_re[i * (2 * l + 1) + k] = Y[k].real();
                out_im[i * (2 * l + 1) + k] = Y[k].imag();
            }
        }
    }
}

// ---------------------------------------------------------------- Wigner 3j
// (l l l; m1 m2 m3) by the Racah formula, evaluated with long-double log-factorials.
// The reference stores a sympy-generated table of the same quantities
// (Wigner3j.cc:59-5833) in the loop order of Wigner3j.cc:41-53; tests check these
// values against sympy's exact ones.
static long double lfact(int n) { return std::lgamma((long double) n + 1.0L); }

double wigner3j_lll(int l, int m1, int m2, int m3)
{
    if (m1 + m2 + m3 != 0 || std::abs(m1) > l || std::abs(m2) > l || std::abs(m3) > l)
        return 0.0;
    int const j1 = l, j2 = l, j3 = l;
    // triangle coefficient
    long double const ltri = lfact(j1 + j2 - j3) + lfact(j1 - j2 + j3) + lfact(-j1 + j2 + j3) - lfact(j1 + j2 + j3 + 1);
    long double const lpre = 0.5L * (ltri + lfact(j1 + m1) + lfact(j1 - m1) + lfact(j2 + m2) + lfact(j2 - m2)
                                     + lfact(j3 + m3) + lfact(j3 - m3));
    int kmin = std::max(0, std::max(j2 - j3 - m1, j1 - j3 + m2));
    int kmax = std::min(j1 + j2 - j3, std::min(j1 - m1, j2 + m2));
    long double sum = 0.0L;
    for (int k = kmin; k <= kmax; ++k)
    {
        long double lt = lfact(k) + lfact(j1 + j2 - j3 - k) + lfact(j1 - m1 - k) + lfact(j2 + m2 - k)
            + lfact(j3 - j2 + m1 + k) + lfact(j3 - j1 - m2 + k);
        long double term = std::exp(lpre - lt);
        sum += (k & 1) ? -term : term;
    }
    int const phase = j1 - j2 - m3;
    return double(((phase % 2) != 0) ? -sum : sum);
}

// Coefficients in the order of Wigner3j.cc:41-53
uint32_t orc_wigner3j_table(int l, double* out)
{
    uint32_t c = 0;
    for (int m1 = -l; m1 <= l; ++m1)
        for (int m2 = std::max(-l - m1, -l); m2 <= std::min(l - m1, l); ++m2)
        {
            if (out)
                out[c] = wigner3j_lll(l, m1, m2, -m1 - m2);
            ++c;
        }
    return c;
}

} // extern "C"

namespace {
inline int lm_index(int l, int m) { return m < 0 ? l - m : m; }

// Wigner3j.cc:22-56
template<typename T> T reduce_wigner3j(const std::complex<T>* src, int l, const std::vector<double>& w3j)
{
    T result = 0;
    unsigned counter = 0;
    for (int m1 = -l; m1 <= l; ++m1)
    {
        int const i1 = lm_index(l, m1);
        for (int m2 = std::max(-l - m1, -l); m2 <= std::min(l - m1, l); ++m2)
        {
            int const i2 = lm_index(l, m2);
            int const i3 = lm_index(l, -m1 - m2);
            result += (T(w3j[counter]) * src[i1] * src[i2] * src[i3]).real();
            ++counter;
        }
    }
    return result;
}

template<typename T>
int steinhardt_impl(const Box& b, const std::vector<v3>& pts, uint32_t np, const uint32_t* ij, const float* dist,
                    const float* wts, uint64_t nb, const uint32_t* ls, int nl, int average, int wl, int weighted,
                    int wl_normalize, int nthreads, T* qlm_out, T* ql_out, T* qlave_out, T* wl_out, T* order_out,
                    T* qlmave_out)
{
    int lmax = 0;
    std::vector<int> nms(nl);
    size_t total_m = 0;
    std::vector<size_t> moff(nl);
    for (int k = 0; k < nl; ++k)
    {
        lmax = std::max(lmax, int(ls[k]));
        nms[k] = 2 * ls[k] + 1;
        moff[k] = total_m;
        total_m += nms[k];
    }
    // segments of the (i-sorted) neighbor list: NeighborList.cc:199-232
    std::vector<uint64_t> seg(np + 1, 0);
    for (uint64_t e = 0; e < nb; ++e)
        seg[ij[2 * e] + 1]++;
    for (uint32_t i = 0; i < np; ++i)
        seg[i + 1] += seg[i];
    // qlm per particle: [i][total_m] complex
    std::vector<std::complex<T>> qlm(size_t(np) * total_m, std::complex<T>(0, 0));
    std::vector<T> ql(size_t(np) * nl, T(0));
    // Steinhardt.cc:120-222
    parallel_for(np, nthreads, [&](size_t beg, size_t end, int) {
        std::vector<T> P;
        std::vector<std::complex<T>> e;
        std::vector<std::complex<T>> Y(2 * lmax + 1);
        for (size_t i = beg; i < end; ++i)
        {
            T total_weight = 0;
            v3 const ref = pts[i];
            for (uint64_t bnd = seg[i]; bnd < seg[i + 1]; ++bnd)
            {
                uint32_t j = ij[2 * bnd + 1];
                v3 const delta = wrap(b, sub(pts[j], ref));
                T const weight = weighted ? T(wts[bnd]) : T(1);
                T const dx = T(delta.x), dy = T(delta.y), dz = T(delta.z), d = T(dist[bnd]);
                T const phi = std::atan2(dy, dx);
                T cz = dz / d;
                cz = std::max(T(-1), std::min(cz, T(1)));
                T theta = std::acos(cz);
                if (d == T(0))
                    theta = 0;
                ylm_all<T>(lmax, std::cos(theta), std::sin(theta), std::cos(phi), std::sin(phi), P, e);
                for (int k = 0; k < nl; ++k)
                {
                    ylm_l<T>(int(ls[k]), P, e, Y.data());
                    std::complex<T>* q = &qlm[i * total_m + moff[k]];
                    for (int m = 0; m < nms[k]; ++m)
                        q[m] += weight * Y[m];
                }
                total_weight += weight;
            }
            for (int k = 0; k < nl; ++k)
            {
                T const nf = T(4.0 * M_PI / nms[k]);
                std::complex<T>* q = &qlm[i * total_m + moff[k]];
                T acc = 0;
                for (int m = 0; m < nms[k]; ++m)
                {
                    q[m] /= total_weight;
                    acc += std::norm(q[m]);
                }
                acc *= nf;
                ql[i * nl + k] = std::sqrt(acc);
            }
        }
    });
    // system qlm accumulates qlm/Np (or qlmAve/Np): Steinhardt.cc:211-216, 276
    std::vector<std::complex<T>> qlmave;
    std::vector<T> qlave;
    std::vector<std::complex<T>> sys(total_m, std::complex<T>(0, 0));
    if (average)
    {
        // Steinhardt.cc:224-289
        qlmave.assign(size_t(np) * total_m, std::complex<T>(0, 0));
        qlave.assign(size_t(np) * nl, T(0));
        parallel_for(np, nthreads, [&](size_t beg, size_t end, int) {
            for (size_t i = beg; i < end; ++i)
            {
                unsigned count = 1;
                for (uint64_t bnd = seg[i]; bnd < seg[i + 1]; ++bnd)
                {
                    uint32_t j = ij[2 * bnd + 1];
                    for (size_t m = 0; m < total_m; ++m)
                        qlmave[i * total_m + m] += qlm[size_t(j) * total_m + m];
                    ++count;
                }
                for (int k = 0; k < nl; ++k)
                {
                    T const nf = T(4.0 * M_PI / nms[k]);
                    T acc = 0;
                    for (int m = 0; m < nms[k]; ++m)
                    {
                        size_t const idx = i * total_m + moff[k] + m;
                        qlmave[idx] += qlm[idx];
                        qlmave[idx] /= T(count);
                        acc += std::norm(qlmave[idx]);
                    }
                    acc *= nf;
                    qlave[i * nl + k] = std::sqrt(acc);
                }
            }
        });
    }
    const auto& src = average ? qlmave : qlm;
    for (size_t i = 0; i < np; ++i)
        for (size_t m = 0; m < total_m; ++m)
            sys[m] += src[i * total_m + m] / T(np);
    // w_l: Steinhardt.cc:329-359
    std::vector<std::vector<double>> w3j(nl);
    for (int k = 0; k < nl; ++k)
    {
        w3j[k].resize(orc_wigner3j_table(ls[k], nullptr));
        orc_wigner3j_table(ls[k], w3j[k].data());
    }
    const auto& norm_src = average ? qlave : ql;
    if (wl)
    {
        for (size_t i = 0; i < np; ++i)
            for (int k = 0; k < nl; ++k)
            {
                T const nf = T(4.0 * M_PI / nms[k]);
                T v = reduce_wigner3j<T>(&src[i * total_m + moff[k]], ls[k], w3j[k]);
                if (wl_normalize)
                {
                    T const norm = std::sqrt(nf) / norm_src[i * nl + k];
                    v *= norm * norm * norm;
                }
                wl_out[i * nl + k] = v;
            }
    }
    // system normalisation: Steinhardt.cc:291-327
    for (int k = 0; k < nl; ++k)
    {
        T const nf = T(4.0 * M_PI / nms[k]);
        T calc = 0;
        for (int m = 0; m < nms[k]; ++m)
            calc += std::norm(sys[moff[k] + m]);
        T const ql_sys = std::sqrt(calc * nf);
        if (wl)
        {
            T v = reduce_wigner3j<T>(&sys[moff[k]], ls[k], w3j[k]);
            if (wl_normalize)
            {
                T const wn = std::sqrt(nf) / ql_sys;
                v *= wn * wn * wn;
            }
            order_out[k] = v;
        }
        else
            order_out[k] = ql_sys;
    }
    // outputs: qlm per l contiguous [l][i][m] as interleaved re/im
    for (int k = 0; k < nl; ++k)
    {
        size_t const base = size_t(np) * moff[k] * 2;
        for (size_t i = 0; i < np; ++i)
            for (int m = 0; m < nms[k]; ++m)
            {
                qlm_out[base + (i * nms[k] + m) * 2] = qlm[i * total_m + moff[k] + m].real();
                qlm_out[base + (i * nms[k] + m) * 2 + 1] = qlm[i * total_m + moff[k] + m].imag();
                if (average && qlmave_out)
                {
                    qlmave_out[base + (i * nms[k] + m) * 2] = qlmave[i * total_m + moff[k] + m].real();
                    qlmave_out[base + (i * nms[k] + m) * 2 + 1] = qlmave[i * total_m + moff[k] + m].imag();
                }
            }
    }
    for (size_t t = 0; t < size_t(np) * nl; ++t)
    {
        ql_out[t] = ql[t];
        if (average)
            qlave_out[t] = qlave[t];
    }
    return 0;
}

// union-find with path halving; union by smaller root (deterministic)
struct DSU
{
    std::vector<uint32_t> p;
    explicit DSU(uint32_t n) : p(n) { std::iota(p.begin(), p.end(), 0u); }
    uint32_t find(uint32_t x)
    {
        while (p[x] != x)
        {
            p[x] = p[p[x]];
            x = p[x];
        }
        return x;
    }
    void unite(uint32_t a, uint32_t b)
    {
        a = find(a);
        b = find(b);
        if (a == b)
            return;
        if (a < b)
            std::swap(a, b);
        p[a] = b;
    }
};

// Cluster.cc:21-128: components over bonds, relabelled by (size desc, min id asc)
uint32_t cluster_impl(uint32_t np, const uint32_t* ij, uint64_t nb, const uint8_t* keep, uint32_t* cluster_idx)
{
    DSU dj(np);
    for (uint64_t e = 0; e < nb; ++e)
        if (!keep || keep[e])
            dj.unite(ij[2 * e], ij[2 * e + 1]);
    std::vector<size_t> label(np, np), count(np, 0), min_id(np, np);
    uint32_t ncl = 0;
    for (uint32_t i = 0; i < np; ++i)
    {
        uint32_t s = dj.find(i);
        if (label[s] == np)
        {
            label[s] = ncl;
            min_id[ncl] = i;
            ++ncl;
        }
        count[label[s]]++;
    }
    std::vector<size_t> idx(ncl);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](size_t a, size_t c) {
        if (count[a] != count[c])
            return count[a] > count[c];
        return min_id[a] < min_id[c];
    });
    std::vector<size_t> inv(ncl);
    for (size_t t = 0; t < ncl; ++t)
        inv[idx[t]] = t;
    for (uint32_t i = 0; i < np; ++i)
        cluster_idx[i] = uint32_t(inv[label[dj.find(i)]]);
    return ncl;
}
} // namespace

extern "C" {

int orc_steinhardt(const float* boxp, int is2D, const float* pts_, uint32_t np, const uint32_t* ij, const float* dist,
                   const float* w, uint64_t nb, const uint32_t* ls, int nl, int average, int wl, int weighted,
                   int wl_normalize, int f64, int nthreads, void* qlm_out, void* ql_out, void* qlave_out, void* wl_out,
                   void* order_out, void* qlmave_out)
{
    Box b = make_box(boxp, is2D);
    std::vector<v3> pts = to_v3(pts_, np);
    if (f64)
        return steinhardt_impl<double>(b, pts, np, ij, dist, w, nb, ls, nl, average, wl, weighted, wl_normalize,
                                       nthreads, (double*) qlm_out, (double*) ql_out, (double*) qlave_out,
                                       (double*) wl_out, (double*) order_out, (double*) qlmave_out);
    return steinhardt_impl<float>(b, pts, np, ij, dist, w, nb, ls, nl, average, wl, weighted, wl_normalize, nthreads,
                                  (float*) qlm_out, (float*) ql_out, (float*) qlave_out, (float*) wl_out,
                                  (float*) order_out, (float*) qlmave_out);
}

uint32_t orc_cluster(uint32_t np, const uint32_t* ij, uint64_t nb, uint32_t* cluster_idx)
{
    return cluster_impl(np, ij, nb, nullptr, cluster_idx);
}

// SolidLiquid.cc:31-114.  qlm is the Steinhardt(l) q_lm (np x (2l+1), interleaved
// complex64) and ql its q_l.  Outputs: ql_ij (nb), num_connections (np),
// cluster_idx (np); returns the number of clusters.
uint32_t orc_solidliquid(uint32_t np, const uint32_t* ij, uint64_t nb, int l, const float* qlm, const float* ql,
                         float q_threshold, uint32_t solid_threshold, int normalize_q, float* ql_ij,
                         uint32_t* num_connections, uint32_t* cluster_idx)
{
    int const nm = 2 * l + 1;
    float const nf = float(4.0 * M_PI / nm);
    const std::complex<float>* q = reinterpret_cast<const std::complex<float>*>(qlm);
    for (uint64_t e = 0; e < nb; ++e)
    {
        uint32_t i = ij[2 * e], j = ij[2 * e + 1];
        std::complex<float> acc = 0;
        for (int k = 0; k < nm; ++k)
            acc += q[size_t(i) * nm + k] * std::conj(q[size_t(j) * nm + k]);
        if (normalize_q)
            acc *= nf / (ql[i] * ql[j]);
        ql_ij[e] = acc.real();
    }
    std::fill(num_connections, num_connections + np, 0u);
    for (uint64_t e = 0; e < nb; ++e)
        if (ql_ij[e] > q_threshold)
            num_connections[ij[2 * e]]++;
    std::vector<uint8_t> keep(nb, 0);
    for (uint64_t e = 0; e < nb; ++e)
        keep[e] = (ql_ij[e] > q_threshold) && num_connections[ij[2 * e]] >= solid_threshold
            && num_connections[ij[2 * e + 1]] >= solid_threshold;
    return cluster_impl(np, ij, nb, keep.data(), cluster_idx);
}

// HexaticTranslational.cc:22-74 (psi interleaved complex64)
void orc_hexatic(const float* boxp, int is2D, const float* pts_, uint32_t np, const uint32_t* ij, const float* w,
                 uint64_t nb, uint32_t k, int weighted, float* psi)
{
    Box b = make_box(boxp, is2D);
    std::vector<v3> pts = to_v3(pts_, np);
    std::vector<uint64_t> seg(np + 1, 0);
    for (uint64_t e = 0; e < nb; ++e)
        seg[ij[2 * e] + 1]++;
    for (uint32_t i = 0; i < np; ++i)
        seg[i + 1] += seg[i];
    for (uint32_t i = 0; i < np; ++i)
    {
        std::complex<float> acc(0, 0);
        float total = 0;
        for (uint64_t e = seg[i]; e < seg[i + 1]; ++e)
        {
            v3 const d = wrap(b, sub(pts[ij[2 * e + 1]], pts[i]));
            float const wt = weighted ? w[e] : 1.0f;
            float const th = std::atan2(d.y, d.x);
            acc += wt * std::exp(std::complex<float>(0, float(k) * th));
            total += wt;
        }
        acc /= std::complex<float>(total);
        psi[2 * i] = acc.real();
        psi[2 * i + 1] = acc.imag();
    }
}

// ---------------------------------------------------------------- LocalDensity, PMFTXY
// LocalDensity.cc:39-84 over an i-sorted bond list (bond order within a point).
void orc_local_density(uint32_t nq, const uint32_t* ij, const float* dist, uint64_t nb, float r_max, float diameter,
                       int is2D, float* density, float* num_neighbors)
{
    float const area = M_PI * r_max * r_max;
    float const volume = static_cast<float>(4.0 / 3.0 * M_PI) * r_max * r_max * r_max;
    std::fill(density, density + nq, 0.0f);
    std::fill(num_neighbors, num_neighbors + nq, 0.0f);
    uint64_t e = 0;
    for (uint32_t i = 0; i < nq; ++i)
    {
        float num = 0;
        for (; e < nb && ij[2 * e] == i; ++e)
        {
            float const d = dist[e];
            if (d < (r_max - diameter / float(2.0)))
                num += float(1.0);
            else
                num += float(1.0) + (r_max - (d + diameter / float(2.0))) / diameter;
            num_neighbors[i] = num;
            density[i] = is2D ? num / area : num / volume;
        }
    }
}

// PMFTXY.cc:25-88 (accumulate): each bond vector rotated by rotmat2::fromAngle(-theta_i)
// (VectorMath.h:905-921, std::cos / std::sin of a float) and binned on two
// RegularAxis (Histogram.h:123-179), row-major index bx * n_y + by.
void orc_pmftxy_accumulate(const uint32_t* ij, const float* vec, uint64_t nb, const float* theta, float x_max,
                           float y_max, uint32_t n_x, uint32_t n_y, uint32_t* counts)
{
    float const wx = (x_max - (-x_max)) / static_cast<float>(n_x);
    float const wy = (y_max - (-y_max)) / static_cast<float>(n_y);
    float const ix = static_cast<float>(1.0) / wx;
    float const iy = static_cast<float>(1.0) / wy;
    for (uint64_t e = 0; e < nb; ++e)
    {
        float const t = -theta[ij[2 * e]];
        float const c = std::cos(t), sn = std::sin(t);
        float const r0x = c, r0y = -sn, r1x = sn, r1y = c;
        float const x = vec[3 * e], y = vec[3 * e + 1];
        float const rx = r0x * x + r0y * y;
        float const ry = r1x * x + r1y * y;
        long const bx = regular_bin(rx, -x_max, x_max, ix, n_x);
        long const by = regular_bin(ry, -y_max, y_max, iy, n_y);
        if (bx < 0 || by < 0)
            continue;
        counts[size_t(bx) * n_y + size_t(by)] += 1;
    }
}

// PMFTXY.cc:62-66 with PMFT.h:76-87
void orc_pmftxy_reduce(const float* boxp, int is2D, const uint32_t* counts, uint32_t n_x, uint32_t n_y, float x_max,
                       float y_max, uint32_t n_points, uint32_t n_query_points, uint32_t n_frames, float* pcf)
{
    Box b = make_box(boxp, is2D);
    float const dx = float(2.0) * x_max / float(n_x);
    float const dy = float(2.0) * y_max / float(n_y);
    float const jacobian = dx * dy;
    float const jacobian_factor = (float) 1.0 / jacobian;
    float const inv_num_dens = box_volume(b) / static_cast<float>(n_query_points);
    float const norm_factor = float(1.0) / (static_cast<float>(n_frames) * static_cast<float>(n_points));
    float const prefactor = inv_num_dens * norm_factor;
    for (uint64_t t = 0; t < uint64_t(n_x) * n_y; ++t)
        pcf[t] = static_cast<float>(counts[t]) * prefactor * jacobian_factor;
}

static float modulus_positive_f(float a, float b) { return std::fmod(std::fmod(a, b) + b, b); }

struct quatf
{
    float s;
    v3 v;
};
// VectorMath.h:810-822
static v3 quat_rotate(const quatf& a, v3 b)
{
    float const c1 = a.s * a.s - dot(a.v, a.v);
    v3 const cr = {a.v.y * b.z - a.v.z * b.y, a.v.z * b.x - a.v.x * b.z, a.v.x * b.y - a.v.y * b.x};
    float const c2 = 2 * a.s;
    float const c3 = 2 * dot(a.v, b);
    v3 const t1 = {c1 * b.x, c1 * b.y, c1 * b.z};
    v3 const t2 = {c2 * cr.x, c2 * cr.y, c2 * cr.z};
    v3 const t3 = {c3 * a.v.x, c3 * a.v.y, c3 * a.v.z};
    return add(add(t1, t2), t3);
}

// BondOrder.cc:87-136 (modes 0 bod, 1 lbod, 2 obcd, 3 oocd) and reduce
// (BondOrder.cc:26-53, 68-74).  Quaternions (s, x, y, z).
void orc_bond_order_accumulate(const uint32_t* ij, const float* vec, uint64_t nb, const float* p_quat,
                               const float* q_quat, int mode, uint32_t n_theta, uint32_t n_phi, uint32_t* counts)
{
    float const two_pi = float(2.0 * M_PI);
    float const pi_f = float(M_PI);
    float const it = static_cast<float>(1.0) / ((two_pi - 0.0f) / static_cast<float>(n_theta));
    float const ip = static_cast<float>(1.0) / ((pi_f - 0.0f) / static_cast<float>(n_phi));
    for (uint64_t e = 0; e < nb; ++e)
    {
        uint32_t const i = ij[2 * e], j = ij[2 * e + 1];
        quatf const ref_q = {p_quat[4 * j], {p_quat[4 * j + 1], p_quat[4 * j + 2], p_quat[4 * j + 3]}};
        quatf const ref_c = {ref_q.s, {-ref_q.v.x, -ref_q.v.y, -ref_q.v.z}};
        quatf const q = {q_quat[4 * i], {q_quat[4 * i + 1], q_quat[4 * i + 2], q_quat[4 * i + 3]}};
        v3 v = {vec[3 * e], vec[3 * e + 1], vec[3 * e + 2]};
        if (mode == 2)
        {
            v = quat_rotate(ref_c, v);
            v = quat_rotate(q, v);
        }
        else if (mode == 1)
            v = quat_rotate(ref_c, v);
        else if (mode == 3)
        {
            v3 z = {0.0f, 0.0f, 1.0f};
            z = quat_rotate(q, z);
            v = quat_rotate(ref_c, z);
        }
        float theta = std::atan2(v.y, v.x);
        theta = modulus_positive_f(theta, two_pi);
        float const phi = std::acos(v.z / std::sqrt(dot(v, v)));
        long const bt = regular_bin(theta, 0.0f, two_pi, it, n_theta);
        long const bp = regular_bin(phi, 0.0f, pi_f, ip, n_phi);
        if (bt < 0 || bp < 0)
            continue;
        counts[size_t(bt) * n_phi + size_t(bp)] += 1;
    }
}

void orc_bond_order_reduce(const uint32_t* counts, uint32_t n_theta, uint32_t n_phi, uint32_t n_frames, float* bo)
{
    float const dt = float(2.0 * M_PI) / float(n_theta);
    float const dp = M_PI / float(n_phi);
    for (uint32_t t = 0; t < n_theta; ++t)
        for (uint32_t p = 0; p < n_phi; ++p)
        {
            float const phi = (float) p * dp;
            float const sa = dt * (std::cos(phi) - std::cos(phi + dp));
            size_t const k = size_t(t) * n_phi + p;
            bo[k] = float(counts[k]) / sa / static_cast<float>(n_frames);
        }
}

// CorrelationFunction.cc:84-98 in bond order: counts and the complex double
// sums of conj(values[j]) * query_values[i] per distance bin.
void orc_correlation_accumulate(const uint32_t* ij, const float* dist, uint64_t nb, const double* values,
                                const double* qvalues, uint32_t bins, float r_max, uint32_t* counts, double* sums)
{
    float const w = (r_max - 0.0f) / static_cast<float>(bins);
    float const inv = static_cast<float>(1.0) / w;
    for (uint64_t e = 0; e < nb; ++e)
    {
        long const b = regular_bin(dist[e], 0.0f, r_max, inv, bins);
        if (b < 0)
            continue;
        std::complex<double> const x(values[2 * ij[2 * e + 1]], values[2 * ij[2 * e + 1] + 1]);
        std::complex<double> const y(qvalues[2 * ij[2 * e]], qvalues[2 * ij[2 * e] + 1]);
        std::complex<double> const p = std::conj(x) * y;
        counts[b] += 1;
        sums[2 * b] += p.real();
        sums[2 * b + 1] += p.imag();
    }
}

// PMFTXYT.cc:76-100, PMFTR12.cc:90-111, PMFTXYZ.cc:106-145 (kind 0, 1, 2).
// n, lo, hi: the three RegularAxis(n, lo, hi).  XYT and R12 take angles
// (p_theta per point, q_theta per query point); XYZ takes query quaternions
// (s, x, y, z) and n_equiv equivalent orientations.
void orc_pmft3_accumulate(int kind, const uint32_t* ij, const float* vec, const float* dist, uint64_t nb,
                          const float* p_theta, const float* q_theta, const float* q_quat, const float* equiv,
                          uint32_t n_equiv, const uint32_t* n, const float* lo, const float* hi, uint32_t* counts)
{
    float const two_pi = float(2.0 * M_PI); // constants::TWO_PI
    float inv[3];
    for (int k = 0; k < 3; ++k)
    {
        float const w = (hi[k] - lo[k]) / static_cast<float>(n[k]);
        inv[k] = static_cast<float>(1.0) / w;
    }
    auto bin = [&](float a, float b, float c) {
        long const b0 = regular_bin(a, lo[0], hi[0], inv[0], n[0]);
        long const b1 = regular_bin(b, lo[1], hi[1], inv[1], n[1]);
        long const b2 = regular_bin(c, lo[2], hi[2], inv[2], n[2]);
        if (b0 < 0 || b1 < 0 || b2 < 0)
            return;
        counts[(size_t(b0) * n[1] + size_t(b1)) * n[2] + size_t(b2)] += 1;
    };
    for (uint64_t e = 0; e < nb; ++e)
    {
        uint32_t const i = ij[2 * e], j = ij[2 * e + 1];
        v3 const d = {vec[3 * e], vec[3 * e + 1], vec[3 * e + 2]};
        if (kind == 0)
        {
            // rotmat2::fromAngle(-theta_i) * (x, y), then the point's relative angle
            float const t = -q_theta[i];
            float const c = std::cos(t), sn = std::sin(t);
            float const rx = c * d.x + (-sn) * d.y;
            float const ry = sn * d.x + c * d.y;
            float const d_theta = std::atan2(-d.y, -d.x);
            float const tt = modulus_positive_f(p_theta[j] - d_theta, two_pi);
            bin(rx, ry, tt);
        }
        else if (kind == 1)
        {
            float const d_theta1 = std::atan2(d.y, d.x);
            float const d_theta2 = std::atan2(-d.y, -d.x);
            float const t1 = modulus_positive_f(p_theta[j] - d_theta1, two_pi);
            float const t2 = modulus_positive_f(q_theta[i] - d_theta2, two_pi);
            bin(dist[e], t1, t2);
        }
        else
        {
            quatf const q = {q_quat[4 * i], {-q_quat[4 * i + 1], -q_quat[4 * i + 2], -q_quat[4 * i + 3]}};
            for (uint32_t k = 0; k < n_equiv; ++k)
            {
                v3 v = quat_rotate(q, d);
                quatf const eq = {equiv[4 * k], {equiv[4 * k + 1], equiv[4 * k + 2], equiv[4 * k + 3]}};
                v = quat_rotate(eq, v);
                bin(v.x, v.y, v.z);
            }
        }
    }
}

// PMFT::reduce (PMFT.h:76-87) with the Jacobians of PMFTXYT.cc:52-57 and
// PMFTR12.cc:58-77; PMFTXYZ::reduce (PMFTXYZ.cc:92-103).  maxes as the library.
void orc_pmft3_reduce(int kind, const float* boxp, int is2D, const uint32_t* counts, const uint32_t* n,
                      const float* maxes, uint32_t n_points, uint32_t n_query_points, uint32_t n_frames,
                      uint32_t n_equiv, float* pcf)
{
    Box b = make_box(boxp, is2D);
    float const two_pi = float(2.0 * M_PI);
    size_t const total = size_t(n[0]) * n[1] * n[2];
    std::vector<float> jf(kind == 1 ? n[0] : 1);
    float inv_num_dens = box_volume(b) / static_cast<float>(n_query_points);
    float norm_factor;
    if (kind == 2)
        norm_factor = (float) 1.0 / ((float) n_frames * (float) n_points * (float) n_equiv);
    else
        norm_factor = float(1.0) / (static_cast<float>(n_frames) * static_cast<float>(n_points));
    float const prefactor = inv_num_dens * norm_factor;
    if (kind == 0)
    {
        float const dx = float(2.0) * maxes[0] / float(n[0]);
        float const dy = float(2.0) * maxes[1] / float(n[1]);
        float const dt = 1 / float(n[2]);
        float const jacobian = dx * dy * dt;
        jf[0] = (float) 1.0 / jacobian;
    }
    else if (kind == 2)
    {
        float const dx = float(2.0) * maxes[0] / float(n[0]);
        float const dy = float(2.0) * maxes[1] / float(n[1]);
        float const dz = float(2.0) * maxes[2] / float(n[2]);
        float const jacobian = dx * dy * dz;
        jf[0] = (float) 1.0 / jacobian;
    }
    else
    {
        float const width = (maxes[0] - 0.0f) / static_cast<float>(n[0]);
        float const dr = maxes[0] / float(n[0]);
        float const dt1 = two_pi / float(n[1]);
        float const dt2 = 1 / float(n[2]);
        float const product = dr * dt1 * dt2;
        for (uint32_t r = 0; r < n[0]; ++r)
        {
            float const e0 = 0.0f + static_cast<float>(r) * width;
            float const e1 = 0.0f + static_cast<float>(r + 1) * width;
            float const c = (e0 + e1) / static_cast<float>(2.0);
            jf[r] = (float) 1.0 / (c * product);
        }
    }
    size_t const per_r = size_t(n[1]) * n[2];
    for (size_t t = 0; t < total; ++t)
        pcf[t] = static_cast<float>(counts[t]) * prefactor * (kind == 1 ? jf[t / per_r] : jf[0]);
}

} // extern "C"

// ---------------------------------------------------------------- NeighborList filters
extern "C" {

// FilterSANN.cc:22-104 over a NeighborList sorted by (query point, distance,
// point) (NeighborList::sort(true), NeighborBond.h:97-112): keep[e] marks the
// SANN bonds (the first 3 of each query point, then while the distance sum
// over (m - 2), evaluated in double as the reference's 2.0 literal makes it,
// exceeds the next distance); unfilled[i] = 1 where the list runs out first.
void orc_filter_sann(uint32_t nq, const uint32_t* ij, const float* dist, uint64_t nb, uint8_t* keep,
                     uint8_t* unfilled)
{
    uint64_t e = 0;
    for (uint32_t i = 0; i < nq; ++i)
    {
        uint64_t const first = e;
        while (e < nb && ij[2 * e] == i)
            ++e;
        uint32_t const n = uint32_t(e - first);
        uint32_t m = 0;
        float sum = 0.0f;
        for (; m < 3 && m < n; ++m)
        {
            sum += dist[first + m];
            keep[first + m] = 1;
        }
        while (m < n && (double(sum) / (double(float(m)) - 2.0)) > double(dist[first + m]))
        {
            sum += dist[first + m];
            keep[first + m] = 1;
            ++m;
        }
        for (uint32_t t = m; t < n; ++t)
            keep[first + t] = 0;
        unfilled[i] = m == n ? 1 : 0;
    }
}

// FilterRAD.cc:22-131 over the same sorted list: neighbour j of query point i
// is blocked by a closer k when
//   r_j^2 |v2|^2 d_j d_k < r_k^2 (v1 . v2) |v1|^2,  v1 = wrap(q_i - p_j), v2 = wrap(q_i - p_k)
// (products left to right as written there); kept bonds carry v1 (vec) and
// weight 1.  terminate: RAD-closed (stop at the first blocked neighbour);
// unfilled[i] = 1 when RAD-closed never met a blocked neighbour.  radii may
// be null (all 1).
void orc_filter_rad(const float* boxp, int is2D, const float* pts_, const float* qpts_, const float* radii,
                    uint32_t nq, const uint32_t* ij, const float* dist, uint64_t nb, int terminate, uint8_t* keep,
                    float* vec, uint8_t* unfilled)
{
    Box const b = make_box(boxp, is2D);
    const v3* pts = reinterpret_cast<const v3*>(pts_);
    const v3* qpts = reinterpret_cast<const v3*>(qpts_);
    auto rad = [&](uint32_t j) { return radii ? radii[j] : 1.0f; };
    uint64_t e = 0;
    for (uint32_t i = 0; i < nq; ++i)
    {
        uint64_t const first = e;
        while (e < nb && ij[2 * e] == i)
            ++e;
        uint32_t const n = uint32_t(e - first);
        for (uint32_t t = 0; t < n; ++t)
            keep[first + t] = 0;
        bool good = true;
        for (uint32_t j = 0; j < n; ++j)
        {
            uint32_t const pj = ij[2 * (first + j) + 1];
            v3 const v1 = wrap(b, sub(qpts[i], pts[pj]));
            good = true;
            for (uint32_t k = 0; k < j; ++k)
            {
                uint32_t const pk = ij[2 * (first + k) + 1];
                v3 const v2 = wrap(b, sub(qpts[i], pts[pk]));
                float const rj = rad(pj), rk = rad(pk);
                float const rj_sq = rj * rj, rk_sq = rk * rk;
                if ((rj_sq * dot(v2, v2) * dist[first + j] * dist[first + k]) < (rk_sq * dot(v1, v2) * dot(v1, v1)))
                {
                    good = false;
                    break;
                }
            }
            if (good)
            {
                keep[first + j] = 1;
                vec[3 * (first + j)] = v1.x;
                vec[3 * (first + j) + 1] = v1.y;
                vec[3 * (first + j) + 2] = v1.z;
            }
            else if (terminate)
                break;
        }
        unfilled[i] = (good && terminate) ? 1 : 0;
    }
}

} // extern "C"

