// Host-side grid and input generation for the sweep (libssw_grid.so).
//
// The sweep consumes a grid of cells with face lists (the reference's
// grid::Cell, src/sweep/grid/cell.rs:92-97). The reference obtains it from an
// incremental Delaunay triangulation (src/voronoi/, 6.3 kLoC) or from a
// cartesian constructor (src/sweep/grid/cartesian.rs). This file provides:
//
//  * ssw_grid_cartesian: the cartesian constructor's semantics on one rank
//    (src/sweep/grid/cartesian.rs:231-395): 6 faces per cell in the order
//    -x,+x,-y,+y,-z,+z, normal = normalize(neighbour_pos - pos), area = h^2,
//    volume = h^3, size = h, out-of-box neighbours periodic or boundary.
//  * ssw_voronoi_build: a periodic 3-D Voronoi tessellation built cell by cell
//    by clipping a cube with the bisector planes of the nearest generators
//    (and their periodic images) in order of distance, stopping once no farther
//    generator can cut the cell. Face quantities follow the reference's grid
//    contract: normal = normalize(p_neighbour - p_cell) with glam's
//    v * (1/|v|) (src/voronoi/cell.rs:63-71), area = triangle fan from the
//    first vertex (src/voronoi/primitives/polygon3d.rs:10-16), volume = sum of
//    pyramid volumes (src/voronoi/cell.rs:153-157,190-197), size =
//    cbrt(3V/(4 pi)) (src/voronoi/cell.rs:186-188). Faces to a periodic image
//    are kind PERIODIC, like the reference's LocalPeriodic; with a non-periodic
//    sweep they act as boundary faces (src/voronoi/constructor/mod.rs:125-135).
//    Mirror faces are made exactly consistent: a face is kept only if both
//    cells found it, and both copies carry the same area.
//  * generators for jittered-lattice and clustered points, lognormal densities,
//    and the nearest-cell source assignment of set_source_terms_system
//    (src/source_systems.rs:48-80: Euclidean, non-periodic, rates summed).
//
// All random numbers come from SplitMix64 streams seeded by the caller.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline double length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return a * (1.0 / length(a)); }  // glam DVec3::normalize

// SplitMix64
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double normal() {  // Box-Muller, one value per call
        double u1 = uniform();
        double u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};
// Independent stream per index: seed mixed with the index.
inline Rng stream(uint64_t seed, uint64_t idx) {
    Rng r(seed ^ (0xD1B54A32D192ED03ull * (idx + 1)));
    r.next();
    return r;
}

// ---- convex polyhedron clipped by half-spaces --------------------------------
struct Face {
    std::vector<V3> pts;  // ordered polygon
    int64_t label;        // neighbour index, or -1 for an initial box face
    int32_t shift[3];     // periodic image shift of the neighbour
    V3 plane_n;           // plane normal (unnormalised) for initial faces
};

struct Poly {
    std::vector<Face> faces;

    void init_box(V3 lo, V3 hi) {
        faces.clear();
        V3 c[8];
        for (int i = 0; i < 8; i++) c[i] = {(i & 1) ? hi.x : lo.x, (i & 2) ? hi.y : lo.y, (i & 4) ? hi.z : lo.z};
        // counter-clockwise seen from outside
        const int q[6][4] = {{0, 4, 6, 2}, {1, 3, 7, 5}, {0, 1, 5, 4}, {2, 6, 7, 3}, {0, 2, 3, 1}, {4, 5, 7, 6}};
        const V3 nn[6] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
        for (int f = 0; f < 6; f++) {
            Face fc;
            for (int k = 0; k < 4; k++) fc.pts.push_back(c[q[f][k]]);
            fc.label = -1 - f;
            fc.shift[0] = fc.shift[1] = fc.shift[2] = 0;
            fc.plane_n = nn[f];
            faces.push_back(fc);
        }
    }

    double max_dist2(V3 p) const {
        double m = 0.0;
        for (const Face& f : faces)
            for (const V3& v : f.pts) {
                V3 d = v - p;
                m = std::max(m, dot(d, d));
            }
        return m;
    }

    // Keep {y : dot(y - m, n) <= 0}. Returns true if the polyhedron changed.
    bool clip(V3 m, V3 n, int64_t label, const int32_t* shift, double tol) {
        // quick test
        bool any_out = false;
        for (const Face& f : faces) {
            for (const V3& v : f.pts)
                if (dot(v - m, n) > tol) {
                    any_out = true;
                    break;
                }
            if (any_out) break;
        }
        if (!any_out) return false;
        std::vector<V3> cut;
        std::vector<Face> out;
        out.reserve(faces.size() + 1);
        for (Face& f : faces) {
            std::vector<V3> np;
            size_t k = f.pts.size();
            for (size_t i = 0; i < k; i++) {
                V3 a = f.pts[i], b = f.pts[(i + 1) % k];
                double sa = dot(a - m, n), sb = dot(b - m, n);
                bool ina = sa <= tol, inb = sb <= tol;
                if (ina) {
                    np.push_back(a);
                    if (sa >= -tol) cut.push_back(a);
                }
                if ((sa < -tol && sb > tol) || (sa > tol && sb < -tol)) {
                    double t = sa / (sa - sb);
                    V3 ip = a + (b - a) * t;
                    np.push_back(ip);
                    cut.push_back(ip);
                }
            }
            if (np.size() >= 3) {
                f.pts.swap(np);
                out.push_back(std::move(f));
            }
        }
        // dedupe cut points and order them around their centroid
        std::vector<V3> u;
        for (const V3& p : cut) {
            bool dup = false;
            for (const V3& q : u) {
                V3 d = p - q;
                if (dot(d, d) <= tol * tol * 1e-2) {
                    dup = true;
                    break;
                }
            }
            if (!dup) u.push_back(p);
        }
        if (u.size() >= 3) {
            V3 cen{0, 0, 0};
            for (const V3& p : u) cen = cen + p;
            cen = cen * (1.0 / (double)u.size());
            V3 nn = normalize(n);
            V3 a = (std::fabs(nn.x) < 0.9) ? V3{1, 0, 0} : V3{0, 1, 0};
            V3 e1 = normalize(cross(nn, a));
            V3 e2 = cross(nn, e1);
            std::vector<std::pair<double, V3>> ang;
            for (const V3& p : u) {
                V3 d = p - cen;
                ang.push_back({std::atan2(dot(d, e2), dot(d, e1)), p});
            }
            std::sort(ang.begin(), ang.end(), [](const auto& l, const auto& r) { return l.first < r.first; });
            Face nf;
            for (auto& pr : ang) nf.pts.push_back(pr.second);
            nf.label = label;
            nf.shift[0] = shift[0];
            nf.shift[1] = shift[1];
            nf.shift[2] = shift[2];
            nf.plane_n = n;
            out.push_back(std::move(nf));
        }
        faces.swap(out);
        return true;
    }
};

// Polygon3d::area: triangle fan from points[0] over periodic windows
// (src/voronoi/primitives/polygon3d.rs:10-16).
double polygon_area(const std::vector<V3>& pts) {
    V3 r = pts[0];
    double s = 0.0;
    size_t k = pts.size();
    for (size_t i = 0; i < k; i++) {
        V3 p1 = pts[i], p2 = pts[(i + 1) % k];
        s += 0.5 * length(cross(r - p1, r - p2));
    }
    return s;
}

struct BuiltFace {
    int64_t nb;
    int32_t shift[3];
    double area;
    V3 normal;
    double height;  // |n . (c - poly[0])| for the pyramid volume
};

struct Grid {
    int64_t N = 0;
    std::vector<int64_t> off;
    std::vector<int32_t> nbr;
    std::vector<uint8_t> kind;
    std::vector<double> area, normal, volume, size;
    int64_t dropped_faces = 0;
    int64_t open_cells = 0;
    std::string error;
};

thread_local std::string g_err;
int g_progress = 0;  // print construction progress to stderr (long builds)

struct Cand {
    double d2;
    int64_t j;
    int32_t s[3];
};

// k-d tree over generator positions (and their periodic images) for the
// k-nearest candidate queries of the cell construction.
struct KdTree {
    struct Item {
        V3 p;
        int64_t j;
        int32_t s[3];
    };
    struct Node {
        double lo[3], hi[3];
        int64_t b, e;
        int32_t l, r;  // children, -1 for a leaf
    };
    std::vector<Item> items;
    std::vector<Node> nodes;

    static double axis(const V3& p, int k) { return k == 0 ? p.x : (k == 1 ? p.y : p.z); }

    int build_rec(int64_t b, int64_t e) {
        Node nd;
        for (int k = 0; k < 3; k++) {
            nd.lo[k] = HUGE_VAL;
            nd.hi[k] = -HUGE_VAL;
        }
        for (int64_t i = b; i < e; i++)
            for (int k = 0; k < 3; k++) {
                nd.lo[k] = std::min(nd.lo[k], axis(items[i].p, k));
                nd.hi[k] = std::max(nd.hi[k], axis(items[i].p, k));
            }
        nd.b = b;
        nd.e = e;
        nd.l = nd.r = -1;
        const int id = (int)nodes.size();
        nodes.push_back(nd);
        if (e - b > 8) {
            int dim = 0;
            for (int k = 1; k < 3; k++)
                if (nd.hi[k] - nd.lo[k] > nd.hi[dim] - nd.lo[dim]) dim = k;
            const int64_t mid = (b + e) / 2;
            std::nth_element(items.begin() + b, items.begin() + mid, items.begin() + e,
                             [dim](const Item& x, const Item& y) { return axis(x.p, dim) < axis(y.p, dim); });
            const int l = build_rec(b, mid);
            const int r = build_rec(mid, e);
            nodes[id].l = l;
            nodes[id].r = r;
        }
        return id;
    }

    void build() {
        nodes.clear();
        nodes.reserve((size_t)(items.size() / 4 + 16));
        if (!items.empty()) build_rec(0, (int64_t)items.size());
    }

    static double box_d2(const Node& n, const V3& q) {
        double d2 = 0.0;
        for (int k = 0; k < 3; k++) {
            const double v = axis(q, k);
            const double d = v < n.lo[k] ? n.lo[k] - v : (v > n.hi[k] ? v - n.hi[k] : 0.0);
            d2 += d * d;
        }
        return d2;
    }

    // The k nearest items to q (excluding item `self` with zero shift), as
    // candidates relative to `pa`, sorted by (d2, index, shift).
    void knn(const V3& q, const V3& pa, int64_t self, size_t k, std::vector<Cand>& out,
             std::vector<std::pair<int, double>>& stack) const {
        out.clear();
        if (nodes.empty()) return;
        auto worse = [](const Cand& a, const Cand& b) { return a.d2 < b.d2; };  // max-heap on d2
        stack.clear();
        stack.push_back({0, box_d2(nodes[0], q)});
        while (!stack.empty()) {
            auto [ni, nd2] = stack.back();
            stack.pop_back();
            if (out.size() == k && nd2 > out.front().d2) continue;
            const Node& n = nodes[ni];
            if (n.l < 0) {
                for (int64_t i = n.b; i < n.e; i++) {
                    const Item& it = items[i];
                    if (it.j == self && it.s[0] == 0 && it.s[1] == 0 && it.s[2] == 0) continue;
                    const V3 d = it.p - pa;
                    const double d2 = dot(d, d);
                    if (out.size() < k) {
                        out.push_back({d2, it.j, {it.s[0], it.s[1], it.s[2]}});
                        std::push_heap(out.begin(), out.end(), worse);
                    } else if (d2 < out.front().d2) {
                        std::pop_heap(out.begin(), out.end(), worse);
                        out.back() = {d2, it.j, {it.s[0], it.s[1], it.s[2]}};
                        std::push_heap(out.begin(), out.end(), worse);
                    }
                }
            } else {
                const double dl = box_d2(nodes[n.l], q), dr = box_d2(nodes[n.r], q);
                if (dl <= dr) {
                    stack.push_back({n.r, dr});
                    stack.push_back({n.l, dl});
                } else {
                    stack.push_back({n.l, dl});
                    stack.push_back({n.r, dr});
                }
            }
        }
        std::sort(out.begin(), out.end(), [](const Cand& a, const Cand& b) {
            if (a.d2 != b.d2) return a.d2 < b.d2;
            if (a.j != b.j) return a.j < b.j;
            for (int k = 0; k < 3; k++)
                if (a.s[k] != b.s[k]) return a.s[k] < b.s[k];
            return false;
        });
    }
};

}  // namespace

extern "C" {

const char* ssw_grid_last_error(void) { return g_err.c_str(); }

// ---- points and densities ------------------------------------------------------

// n^3 points ((i + 1/2 + u) h) with u ~ U(-jitter, jitter) per axis, wrapped into
// the box; index = (iz * n + iy) * n + ix.
void ssw_points_jittered_lattice(int64_t n, double box, double jitter, uint64_t seed, double* out) {
    double h = box / (double)n;
#pragma omp parallel for schedule(static)
    for (int64_t iz = 0; iz < n; iz++)
        for (int64_t iy = 0; iy < n; iy++)
            for (int64_t ix = 0; ix < n; ix++) {
                int64_t idx = (iz * n + iy) * n + ix;
                Rng r = stream(seed, (uint64_t)idx);
                int64_t ii[3] = {ix, iy, iz};
                for (int k = 0; k < 3; k++) {
                    double u = (2.0 * r.uniform() - 1.0) * jitter;
                    double v = ((double)ii[k] + 0.5 + u) * h;
                    if (v < 0.0) v += box;
                    if (v >= box) v -= box;
                    out[3 * idx + k] = v;
                }
            }
}

// n points uniform in [0, box)^3.
void ssw_points_uniform(int64_t n, double box, uint64_t seed, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        Rng r = stream(seed, (uint64_t)i);
        for (int k = 0; k < 3; k++) out[3 * i + k] = r.uniform() * box;
    }
}

// rho = rho_mean * exp(sigma g - sigma^2/2), g ~ N(0,1) i.i.d. per cell.
void ssw_lognormal_density(int64_t n, double rho_mean, double sigma_ln, uint64_t seed, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        Rng r = stream(seed, (uint64_t)i);
        double g = r.normal();
        out[i] = rho_mean * std::exp(sigma_ln * g - 0.5 * sigma_ln * sigma_ln);
    }
}

// Nearest cell of each source (Euclidean, non-periodic), rates summed
// (src/source_systems.rs:48-80). Ties go to the lower cell index.
int ssw_assign_sources(int64_t n, const double* pts, int64_t nsrc, const double* src_pos, const double* src_rate,
                       double* source_out, int64_t* cell_out) {
    for (int64_t i = 0; i < n; i++) source_out[i] = 0.0;
    std::vector<int64_t> best(nsrc, -1);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t s = 0; s < nsrc; s++) {
        double bd = INFINITY;
        int64_t bi = -1;
        const double* q = src_pos + 3 * s;
        for (int64_t i = 0; i < n; i++) {
            double dx = pts[3 * i] - q[0], dy = pts[3 * i + 1] - q[1], dz = pts[3 * i + 2] - q[2];
            double d2 = (dx * dx + dy * dy) + dz * dz;
            if (d2 < bd) {
                bd = d2;
                bi = i;
            }
        }
        best[s] = bi;
    }
    for (int64_t s = 0; s < nsrc; s++) {
        if (best[s] < 0) return -1;
        source_out[best[s]] += src_rate[s];
        if (cell_out) cell_out[s] = best[s];
    }
    return 0;
}

// ---- cartesian grid (src/sweep/grid/cartesian.rs:231-395, one rank) ------------
// n^3 cells of side h = box/n, cell index = (ix * n + iy) * n + iz as the
// constructor enumerates positions (x outermost, :179-190). Arrays sized N+1,
// 6N, 6N, 6N, 18N, N, N.
void ssw_grid_cartesian(int64_t n, double box, int periodic, int64_t* off, int32_t* nbr, uint8_t* kind,
                        double* area, double* normal, double* volume, double* size) {
    double h = box / (double)n;  // NumCellsSpec::CellSize
    int64_t N = n * n * n;
    for (int64_t c = 0; c <= N; c++) off[c] = 6 * c;
    const int dirs[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int64_t ix = 0; ix < n; ix++)
        for (int64_t iy = 0; iy < n; iy++)
            for (int64_t iz = 0; iz < n; iz++) {
                int64_t c = (ix * n + iy) * n + iz;
                // to_pos: side * i / n
                V3 pos{box * (double)ix / (double)n, box * (double)iy / (double)n, box * (double)iz / (double)n};
                for (int f = 0; f < 6; f++) {
                    int64_t j[3] = {ix + dirs[f][0], iy + dirs[f][1], iz + dirs[f][2]};
                    V3 np{box * (double)j[0] / (double)n, box * (double)j[1] / (double)n,
                          box * (double)j[2] / (double)n};
                    V3 nn = normalize(np - pos);
                    bool inside = true;
                    for (int k = 0; k < 3; k++) {
                        if (j[k] < 0) {
                            j[k] += n;
                            inside = false;
                        } else if (j[k] >= n) {
                            j[k] -= n;
                            inside = false;
                        }
                    }
                    int64_t fi = 6 * c + f;
                    nbr[fi] = (int32_t)((j[0] * n + j[1]) * n + j[2]);
                    kind[fi] = inside ? 0 : (periodic ? 1 : 2);
                    area[fi] = h * h;  // cell_size.powi::<2>()
                    normal[3 * fi] = nn.x;
                    normal[3 * fi + 1] = nn.y;
                    normal[3 * fi + 2] = nn.z;
                }
                volume[c] = h * h * h;
                size[c] = h;
            }
}

// ---- periodic Voronoi -----------------------------------------------------------
// mode 0: periodic box [0,box)^3 (faces to images are PERIODIC);
// mode 1: free space clipped to [lo, hi] (walls become BOUNDARY faces).
void* ssw_voronoi_build(int64_t n, const double* pts, const double* box, int mode, const double* lo_hi,
                        int num_threads) {
    Grid* g = new Grid();
    g->N = n;
    double boxv[3] = {box[0], box[1], box[2]};
    if (mode == 1) {
        // shift points into a box covering [lo, hi]
        boxv[0] = lo_hi[3] - lo_hi[0];
        boxv[1] = lo_hi[4] - lo_hi[1];
        boxv[2] = lo_hi[5] - lo_hi[2];
    }
    std::vector<double> local(3 * n);
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) local[3 * i + k] = (mode == 1) ? pts[3 * i + k] - lo_hi[k] : pts[3 * i + k];
    const double mean_sp = std::cbrt(boxv[0] * boxv[1] * boxv[2] / (double)n);
    // Candidate generators: the points and, in the periodic box, their images
    // within `margin` of the box. Cells whose security sphere leaves the imaged
    // region are built again with a tree of wider margin.
    auto make_tree = [&](double margin, KdTree& kd) {
        kd.items.clear();
        kd.items.reserve((size_t)n * (mode == 0 ? 2 : 1));
        for (int64_t j = 0; j < n; j++) {
            const V3 q{local[3 * j], local[3 * j + 1], local[3 * j + 2]};
            for (int sz = -1; sz <= 1; sz++)
                for (int sy = -1; sy <= 1; sy++)
                    for (int sx = -1; sx <= 1; sx++) {
                        if (mode == 1 && (sx || sy || sz)) continue;
                        const V3 im{q.x + sx * boxv[0], q.y + sy * boxv[1], q.z + sz * boxv[2]};
                        if (im.x < -margin || im.x > boxv[0] + margin || im.y < -margin ||
                            im.y > boxv[1] + margin || im.z < -margin || im.z > boxv[2] + margin)
                            continue;
                        kd.items.push_back({im, j, {sx, sy, sz}});
                    }
        }
        kd.build();
    };
    std::vector<std::vector<BuiltFace>> cells(n);
    std::vector<double> vols(n, 0.0);
    std::atomic<int64_t> open_cells{0};
    std::atomic<int64_t> done_cells{0};
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#endif
    const double bmax = std::max(boxv[0], std::max(boxv[1], boxv[2]));
    double margin = (mode == 0) ? std::min(0.5, std::max(0.1, 8.0 * mean_sp / bmax)) * bmax : 0.0;
    std::vector<int64_t> todo(n);
    for (int64_t i = 0; i < n; i++) todo[i] = i;
    std::vector<double> need(n, 0.0);  // margin a cell needs beyond the current one
    for (;;) {
        KdTree kd;
        make_tree(margin, kd);
#pragma omp parallel
        {
            Poly poly;
            std::vector<Cand> cands;
            std::vector<std::pair<int, double>> stack;
#pragma omp for schedule(dynamic, 256)
            for (size_t ti = 0; ti < todo.size(); ti++) {
                const int64_t i = todo[ti];
                if (g_progress) {
                    const int64_t k = ++done_cells;
                    const int64_t step = std::max<int64_t>(1, (int64_t)todo.size() / 10);
                    if (k % step == 0) {
                        fprintf(stderr, "voronoi: %lld of %lld cells\n", (long long)k, (long long)todo.size());
                        fflush(stderr);
                    }
                }
                // work relative to the generator for accuracy
                V3 pa{local[3 * i], local[3 * i + 1], local[3 * i + 2]};
                V3 p{0.0, 0.0, 0.0};
                double tol = 1e-12 * mean_sp;
                if (mode == 0) poly.init_box(V3{-bmax, -bmax, -bmax}, V3{bmax, bmax, bmax});
                else poly.init_box(V3{0, 0, 0} - pa, V3{boxv[0], boxv[1], boxv[2]} - pa);
                // Clip with the k nearest candidates in order of distance; the
                // cell is complete once no candidate beyond the k-th can cut it
                // (distance > 2 x the farthest vertex). A larger k repeats the
                // smaller query's candidates first (same total order); those are
                // skipped, so every plane is applied once.
                int done_r = -1;
                const size_t total = kd.items.size() - 1;
                bool have_last = false;
                Cand last{0.0, 0, {0, 0, 0}};
                auto after_last = [&](const Cand& c) {
                    if (!have_last) return true;
                    if (c.d2 != last.d2) return c.d2 > last.d2;
                    if (c.j != last.j) return c.j > last.j;
                    for (int k = 0; k < 3; k++)
                        if (c.s[k] != last.s[k]) return c.s[k] > last.s[k];
                    return false;
                };
                for (size_t k = std::min<size_t>(64, total);; k = std::min(total, k * 4)) {
                    kd.knn(pa, pa, i, k, cands, stack);
                    double rmax2 = poly.max_dist2(p);
                    bool stop = false;
                    for (const Cand& c : cands) {
                        if (!after_last(c)) continue;
                        last = c;
                        have_last = true;
                        if (c.d2 > 4.0 * rmax2 * (1.0 + 1e-9)) {
                            stop = true;
                            break;
                        }
                        V3 q{local[3 * c.j] + c.s[0] * boxv[0] - pa.x, local[3 * c.j + 1] + c.s[1] * boxv[1] - pa.y,
                             local[3 * c.j + 2] + c.s[2] * boxv[2] - pa.z};
                        V3 m = q * 0.5;
                        if (poly.clip(m, q, c.j, c.s, tol)) rmax2 = poly.max_dist2(p);
                    }
                    const bool all = cands.size() >= total;
                    if (stop || all || 4.0 * rmax2 * (1.0 + 1e-9) < cands.back().d2) {
                        done_r = 1;
                        break;
                    }
                }
                need[i] = 0.0;
                if (mode == 0 && done_r > 0) {
                    // the security sphere must lie inside the imaged region
                    const double r = 2.0 * std::sqrt(poly.max_dist2(p)) * (1.0 + 1e-9);
                    double over = 0.0;
                    over = std::max(over, r - pa.x);
                    over = std::max(over, r - pa.y);
                    over = std::max(over, r - pa.z);
                    over = std::max(over, pa.x + r - boxv[0]);
                    over = std::max(over, pa.y + r - boxv[1]);
                    over = std::max(over, pa.z + r - boxv[2]);
                    if (over > margin) {
                        if (margin < bmax) {
                            need[i] = over;
                            continue;  // built again with a wider margin
                        }
                        done_r = -1;
                    }
                }
                if (done_r < 0 && mode == 0) open_cells++;
                // collect faces
                std::vector<BuiltFace>& out = cells[i];
                double vol = 0.0;
                for (const Face& f : poly.faces) {
                    double a = polygon_area(f.pts);
                    BuiltFace bf;
                    bf.nb = f.label;
                    bf.shift[0] = f.shift[0];
                    bf.shift[1] = f.shift[1];
                    bf.shift[2] = f.shift[2];
                    bf.area = a;
                    if (f.label >= 0) {
                        V3 q{local[3 * f.label] + f.shift[0] * boxv[0], local[3 * f.label + 1] + f.shift[1] * boxv[1],
                             local[3 * f.label + 2] + f.shift[2] * boxv[2]};
                        bf.normal = normalize(q - pa);
                    } else {
                        bf.normal = normalize(f.plane_n);
                    }
                    // pyramid_volume (src/voronoi/cell.rs:153-157)
                    bf.height = std::fabs(dot(bf.normal, p - f.pts[0]));
                    vol += 1.0 / 3.0 * a * bf.height;
                    bool merged = false;
                    for (BuiltFace& e : out) {
                        if (e.nb == bf.nb && bf.nb >= 0 && e.shift[0] == bf.shift[0] && e.shift[1] == bf.shift[1] &&
                            e.shift[2] == bf.shift[2]) {
                            if (bf.area > e.area) e = bf;
                            merged = true;
                        }
                    }
                    if (!merged) out.push_back(bf);
                }
                vols[i] = vol;
            }
        }
        std::vector<int64_t> again;
        double widest = 0.0;
        for (int64_t i : todo)
            if (need[i] > 0.0) {
                again.push_back(i);
                widest = std::max(widest, need[i]);
            }
        if (again.empty()) break;
        todo.swap(again);
        margin = std::min(bmax, widest * 1.05);
    }
    g->open_cells = open_cells.load();
    // Mirror consistency: keep neighbour faces whose twin exists, share the area.
    auto find_twin = [&](int64_t i, const BuiltFace& bf) -> const BuiltFace* {
        const int32_t ms[3] = {-bf.shift[0], -bf.shift[1], -bf.shift[2]};
        for (const BuiltFace& t : cells[bf.nb])
            if (t.nb == i && t.shift[0] == ms[0] && t.shift[1] == ms[1] && t.shift[2] == ms[2]) return &t;
        return nullptr;
    };
    g->off.assign(n + 1, 0);
    std::vector<std::vector<BuiltFace>> kept(n);
    int64_t dropped = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : dropped)
    for (int64_t i = 0; i < n; i++) {
        for (const BuiltFace& bf : cells[i]) {
            if (bf.nb < 0) {
                if (mode == 1) kept[i].push_back(bf);
                continue;
            }
            const BuiltFace* t = find_twin(i, bf);
            if (!t || !(bf.area > 0.0) || !(t->area > 0.0)) {
                dropped++;
                continue;
            }
            BuiltFace k = bf;
            // the smaller index's polygon area is used for both copies
            bool mine = (i < bf.nb) || (i == bf.nb && (bf.shift[0] + bf.shift[1] * 3 + bf.shift[2] * 9) > 0);
            k.area = mine ? bf.area : t->area;
            kept[i].push_back(k);
        }
        // deterministic order: neighbour index, then shift
        std::sort(kept[i].begin(), kept[i].end(), [](const BuiltFace& a, const BuiltFace& b) {
            if (a.nb != b.nb) return a.nb < b.nb;
            for (int k = 0; k < 3; k++)
                if (a.shift[k] != b.shift[k]) return a.shift[k] < b.shift[k];
            return false;
        });
    }
    for (int64_t i = 0; i < n; i++) g->off[i + 1] = g->off[i] + (int64_t)kept[i].size();
    g->dropped_faces = dropped;
    int64_t F = g->off[n];
    g->nbr.resize(F);
    g->kind.resize(F);
    g->area.resize(F);
    g->normal.resize(3 * F);
    g->volume.resize(n);
    g->size.resize(n);
    for (int64_t i = 0; i < n; i++) {
        int64_t f0 = g->off[i];
        for (size_t k = 0; k < kept[i].size(); k++) {
            const BuiltFace& bf = kept[i][k];
            int64_t f = f0 + (int64_t)k;
            bool img = bf.shift[0] != 0 || bf.shift[1] != 0 || bf.shift[2] != 0;
            g->nbr[f] = bf.nb >= 0 ? (int32_t)bf.nb : 0;
            g->kind[f] = bf.nb < 0 ? 2 : (img ? 1 : 0);
            g->area[f] = bf.area;
            g->normal[3 * f] = bf.normal.x;
            g->normal[3 * f + 1] = bf.normal.y;
            g->normal[3 * f + 2] = bf.normal.z;
        }
        g->volume[i] = vols[i];
        g->size[i] = std::cbrt(3.0 * vols[i] / (4.0 * M_PI));
    }
    return g;
}

void ssw_grid_set_progress(int enabled) { g_progress = enabled; }

int64_t ssw_voronoi_num_faces(void* h) { return static_cast<Grid*>(h)->off.back(); }
int64_t ssw_voronoi_dropped_faces(void* h) { return static_cast<Grid*>(h)->dropped_faces; }
int64_t ssw_voronoi_open_cells(void* h) { return static_cast<Grid*>(h)->open_cells; }

void ssw_voronoi_get(void* h, int64_t* off, int32_t* nbr, uint8_t* kind, double* area, double* normal, double* volume,
                     double* size) {
    Grid* g = static_cast<Grid*>(h);
    int64_t F = g->off.back();
    std::memcpy(off, g->off.data(), (g->N + 1) * sizeof(int64_t));
    std::memcpy(nbr, g->nbr.data(), F * sizeof(int32_t));
    std::memcpy(kind, g->kind.data(), F);
    std::memcpy(area, g->area.data(), F * sizeof(double));
    std::memcpy(normal, g->normal.data(), 3 * F * sizeof(double));
    std::memcpy(volume, g->volume.data(), g->N * sizeof(double));
    std::memcpy(size, g->size.data(), g->N * sizeof(double));
}

void ssw_voronoi_free(void* h) { delete static_cast<Grid*>(h); }

}  // extern "C"
