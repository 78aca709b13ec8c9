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

struct BucketGrid {
    int nb[3];
    double L[3];
    double bs[3];
    std::vector<int64_t> start;  // nb^3 + 1
    std::vector<int64_t> items;
    void build(const double* pts, int64_t n, const double* box, double target_per_bucket) {
        double vol = box[0] * box[1] * box[2];
        double bsz = std::cbrt(vol * target_per_bucket / (double)n);
        for (int k = 0; k < 3; k++) {
            L[k] = box[k];
            nb[k] = std::max(1, (int)std::floor(box[k] / bsz));
            if (nb[k] > 2048) nb[k] = 2048;
            bs[k] = box[k] / nb[k];
        }
        int64_t nt = (int64_t)nb[0] * nb[1] * nb[2];
        std::vector<int64_t> cnt(nt + 1, 0);
        std::vector<int64_t> b(n);
        for (int64_t i = 0; i < n; i++) {
            b[i] = bucket_of(pts + 3 * i);
            cnt[b[i] + 1]++;
        }
        for (int64_t i = 0; i < nt; i++) cnt[i + 1] += cnt[i];
        start = cnt;
        items.assign(n, 0);
        std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
        for (int64_t i = 0; i < n; i++) items[fill[b[i]]++] = i;
    }
    int coord(double v, int k) const {
        int c = (int)std::floor(v / bs[k]);
        if (c < 0) c = 0;
        if (c >= nb[k]) c = nb[k] - 1;
        return c;
    }
    int64_t bucket_of(const double* p) const {
        return ((int64_t)coord(p[2], 2) * nb[1] + coord(p[1], 1)) * nb[0] + coord(p[0], 0);
    }
};

struct Cand {
    double d2;
    int64_t j;
    int32_t s[3];
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
    BucketGrid bg;
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
    bg.build(local.data(), n, boxv, 4.0);
    double mean_sp = std::cbrt(boxv[0] * boxv[1] * boxv[2] / (double)n);
    std::vector<std::vector<BuiltFace>> cells(n);
    std::vector<double> vols(n, 0.0);
    std::atomic<int64_t> open_cells{0};
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#endif
#pragma omp parallel
    {
        Poly poly;
        std::vector<Cand> cands;
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; i++) {
            // work relative to the generator for accuracy
            V3 pa{local[3 * i], local[3 * i + 1], local[3 * i + 2]};
            V3 p{0.0, 0.0, 0.0};
            double tol = 1e-12 * mean_sp;
            if (mode == 0) {
                double hb = std::max(boxv[0], std::max(boxv[1], boxv[2]));
                poly.init_box(V3{-hb, -hb, -hb}, V3{hb, hb, hb});
            } else {
                poly.init_box(V3{0, 0, 0} - pa, V3{boxv[0], boxv[1], boxv[2]} - pa);
            }
            int bc[3] = {bg.coord(pa.x, 0), bg.coord(pa.y, 1), bg.coord(pa.z, 2)};
            int maxr = std::max(bg.nb[0], std::max(bg.nb[1], bg.nb[2])) + 1;
            int done_r = -1;
            double applied2 = -1.0;  // candidates with d2 <= applied2 were clipped already
            for (int r = 1; r <= maxr; r++) {
                cands.clear();
                for (int oz = -r; oz <= r; oz++)
                    for (int oy = -r; oy <= r; oy++)
                        for (int ox = -r; ox <= r; ox++) {
                            int o[3] = {ox, oy, oz};
                            int bi[3];
                            int32_t sh[3];
                            bool skip = false;
                            for (int k = 0; k < 3; k++) {
                                int v = bc[k] + o[k];
                                int m = bg.nb[k];
                                int s = (int)std::floor((double)v / m);
                                bi[k] = v - s * m;
                                sh[k] = s;
                                if (mode == 1 && s != 0) skip = true;
                            }
                            if (skip) continue;
                            int64_t b = ((int64_t)bi[2] * bg.nb[1] + bi[1]) * bg.nb[0] + bi[0];
                            for (int64_t t = bg.start[b]; t < bg.start[b + 1]; t++) {
                                int64_t j = bg.items[t];
                                if (j == i && sh[0] == 0 && sh[1] == 0 && sh[2] == 0) continue;
                                V3 q{local[3 * j] + sh[0] * boxv[0], local[3 * j + 1] + sh[1] * boxv[1],
                                     local[3 * j + 2] + sh[2] * boxv[2]};
                                V3 d = q - pa;
                                double d2 = dot(d, d);
                                if (d2 <= applied2) continue;
                                cands.push_back({d2, j, {sh[0], sh[1], sh[2]}});
                            }
                        }
                std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) {
                    if (a.d2 != b.d2) return a.d2 < b.d2;
                    if (a.j != b.j) return a.j < b.j;
                    for (int k = 0; k < 3; k++)
                        if (a.s[k] != b.s[k]) return a.s[k] < b.s[k];
                    return false;
                });
                // points within radius r*min(bs) are complete
                double complete = r * std::min(bg.bs[0], std::min(bg.bs[1], bg.bs[2]));
                double complete2 = complete * complete;
                double rmax2 = poly.max_dist2(p);
                bool stop = false;
                for (const Cand& c : cands) {
                    if (c.d2 > complete2) break;
                    if (c.d2 > 4.0 * rmax2 * (1.0 + 1e-9)) {
                        stop = true;
                        break;
                    }
                    V3 q{local[3 * c.j] + c.s[0] * boxv[0] - pa.x, local[3 * c.j + 1] + c.s[1] * boxv[1] - pa.y,
                         local[3 * c.j + 2] + c.s[2] * boxv[2] - pa.z};
                    V3 m = q * 0.5;
                    if (poly.clip(m, q, c.j, c.s, tol)) rmax2 = poly.max_dist2(p);
                }
                applied2 = complete2;
                if (stop || 4.0 * rmax2 <= complete2) {
                    done_r = r;
                    break;
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
    g->open_cells = open_cells.load();
    // Mirror consistency: keep neighbour faces whose twin exists, share the area.
    auto key = [](int64_t a, int64_t b, const int32_t* s) {
        uint64_t h = (uint64_t)a * 0x9E3779B97F4A7C15ull ^ ((uint64_t)b + 0x632BE59BD9B4E019ull) * 0xC2B2AE3D27D4EB4Full;
        h ^= ((uint64_t)(s[0] + 8) << 0) | ((uint64_t)(s[1] + 8) << 8) | ((uint64_t)(s[2] + 8) << 16);
        return h;
    };
    // index of faces by (cell, nb, shift)
    std::unordered_map<uint64_t, std::vector<std::pair<int64_t, int>>> idx;
    idx.reserve((size_t)n * 16);
    for (int64_t i = 0; i < n; i++)
        for (int f = 0; f < (int)cells[i].size(); f++) {
            const BuiltFace& bf = cells[i][f];
            if (bf.nb < 0) continue;
            idx[key(i, bf.nb, bf.shift)].push_back({i, f});
        }
    auto find_twin = [&](int64_t i, const BuiltFace& bf) -> const BuiltFace* {
        int32_t ms[3] = {-bf.shift[0], -bf.shift[1], -bf.shift[2]};
        auto it = idx.find(key(bf.nb, i, ms));
        if (it == idx.end()) return nullptr;
        for (auto& pr : it->second) {
            const BuiltFace& t = cells[pr.first][pr.second];
            if (pr.first == bf.nb && t.nb == i && t.shift[0] == ms[0] && t.shift[1] == ms[1] && t.shift[2] == ms[2])
                return &t;
        }
        return nullptr;
    };
    g->off.assign(n + 1, 0);
    std::vector<std::vector<BuiltFace>> kept(n);
    int64_t dropped = 0;
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
        g->off[i + 1] = g->off[i] + (int64_t)kept[i].size();
    }
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
