// libsubsweep_amd: MI355X (gfx950) sweep radiative transfer + hydrogen chemistry.
//
// Device layout (one handle = one resident problem):
//   faces      CSR over cells: off[N+1] (u32), geo[F] = {nx, ny, nz, area} (f64x4),
//              nk[F] = neighbour | kind << 30 (u32); for the Nper cells with a
//              periodic face also a compact list of those faces
//   cell state rho, x, T, source/D, attenuation, timestep, previous rate,
//              change timescale (f64[N]); timestep level (u8[N])
//   fluxes     direction-major [d * N + c]:
//              X = outgoing rate / downwind effective area (FP64 or scaled FP32)
//              I = incoming rate at solve time (FP64 or scaled FP32)
//              P[d * Nper + slot] periodic-source snapshot of the periodic cells
//   schedule   the (cell, direction) tasks of the full dependency graph in Kahn
//              order, level-major, within a level sorted by (direction, cell);
//              full_start[k] is where level set k begins. Two replay forms:
//              geometric  task records {cell | face_count << 25, first_face,
//                         direction, periodic slot} (16 B per task)
//              compiled   per task a header {cell | direction << 25, periodic
//                         slot, overflow start, upwind count}, the effective area
//                         A, and its upwind local faces as (neighbour, weight)
//                         pairs in kPlanes planes: entry q of the task at
//                         position p of a level set starting at b with width W
//                         is planes[kPlanes * b + q * W + (p - b)] in FP64 and
//                         planes[kPlanes * b + ((q / 2) * W + (p - b)) * 2 + q % 2]
//                         in FP32, plane pairs of 16 bytes (coalesced across
//                         tasks, addressable without reading the header);
//                         entries q >= kPlanes (rare) sit in an overflow list,
//                         indexed per (direction, cell).
//              A sweep at timestep level l solves the cells with level >= l.
//              Restricted to those cells, the full graph's level sets are still
//              a valid order (an active task's active upwind neighbours are in
//              earlier level sets), so such a sweep replays the full schedule
//              through a per-level list of the active positions (idx[l]),
//              made by a stream compaction when the active set changes.
//
// The transport is the reference's upwind sweep (src/sweep/mod.rs:412-513)
// restated in pull form. A cell that is solved distributes the change of its
// outgoing rate over its downwind faces in proportion to area * (n . d)
// (:453-461), so the incoming rate of a cell equals, up to rounding,
//   I[c,d] = sum_{upwind local faces f} area_f (-n_f . d) * O[nb,d] / A[nb,d]
// with A[nb,d] = sum of area (n . d) over nb's downwind faces. Storing
// X = O / A lets every task gather its upwind neighbours' X instead of
// scattering with atomics. The upwind sum is evaluated in face order, one term
// after the other, and A by a fixed 16-way tree, in every kernel, so the
// geometric and the compiled replay give bitwise the same result, whatever
// order tasks run in. The level sets (Kahn layering of init_counts' counts,
// src/sweep/mod.rs:346-386) are built on the device once per grid by counter
// decrements with block-aggregated appends; the level-set parity hook builds
// the layering of a restricted active set the same way.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <mutex>
#include <shared_mutex>
#include <thread>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/subsweep_amd.h"
#include "chemistry.h"
#include "direction_bins.h"
#include "level_math.h"

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kKindLocal = 0;
constexpr uint32_t kKindPeriodic = 1;
constexpr uint32_t kKindBoundary = 2;
constexpr uint32_t kNbrMask = 0x3FFFFFFFu;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kLanes = 16;                 // lanes per task in the build / geometric kernels
constexpr int kTasksPerBlock = kBlock / kLanes;
constexpr uint32_t kCellBits = 25;         // task record: cell index bits
constexpr uint32_t kCellMask = (1u << kCellBits) - 1u;
constexpr int kMaxFaces = 127;             // faces per cell
constexpr int kQueueCap = kTasksPerBlock * kMaxFaces;
constexpr int kPlanes = 12;                // upwind entries per task held in planes
static_assert(kPlanes % 2 == 0, "FP32 plane entries are stored in pairs");
constexpr int kChunk = 4096;               // compaction chunk (tasks, within one level set)
constexpr int kRunBlock = 1024;            // workgroup of a run of narrow level sets (tasks per level set)
constexpr int kGroupT = 64;                // grouped replay: root tasks per workgroup (default; 64 or 128)
constexpr int kGroupLevels = 4;            // grouped replay: level sets per group (default)
constexpr int kConeCap = 3072;             // grouped replay build: cone members / pending entries per workgroup
constexpr int kSortCap = 4096;             // grouped replay build: entries of one level set sorted in LDS
constexpr int kMaxGroup = 8;               // grouped replay: level sets per group
constexpr uint32_t kNarrowGroup = 2048;    // grouped replay: level sets this narrow go kMaxGroup to a group
constexpr uint32_t kXcds = 8;              // MI355X: accelerator dies, each with its own L2

// Staging of ssw_download_noalloc, which allocates nothing: a file-scope
// device array that every device the library runs on holds until the process
// ends. One output array of at most kNoAllocBytes / 8 cells passes through it
// at a time. Host code never uses the name `scratch` as an address (in host
// code it is a stand-in, not the device array): it uses the per-device address
// from hipGetSymbolAddress (noalloc_scratch below).
constexpr size_t kNoAllocBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

thread_local std::string g_err;
// Handles alive in the process.
std::atomic<int> g_live_handles{0};
// A stream capture is invalidated by the device-synchronising calls
// (allocation, free, blocking copies, stream creation and destruction) that
// other threads make while it runs. Those calls hold this lock shared; a
// capture holds it exclusively, so that handles driven from several threads
// (the request loop) capture their replays too.
std::shared_mutex g_capture_mu;
std::atomic<uint64_t> g_next_uid{0};

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess)                                                                   \
            return fail(_e == hipErrorOutOfMemory ? SSW_ERR_OOM : SSW_ERR_HIP,                  \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                     \
    } while (0)

#define TRY(expr)                    \
    do {                             \
        int _r = (expr);             \
        if (_r != SSW_OK) return _r; \
    } while (0)

// ---------------------------------------------------------------------------
// device helpers

__device__ __forceinline__ double fdot(const double4& g, double dx, double dy, double dz) {
    // glam DVec3::dot: (x*x') + (y*y') + (z*z'), no contraction
    return (g.x * dx + g.y * dy) + g.z * dz;
}

// Upwind entry of the compiled schedule.
template <typename R>
struct Entry;
template <>
struct __align__(8) Entry<float> {
    uint32_t nb;
    float w;
};
template <>
struct __align__(16) Entry<double> {
    uint32_t nb;
    uint32_t pad;
    double w;
};

// The plane entries of a level set of width W, from the level set's first
// entry (kPlanes * W entries of storage; p is a task's offset in the level
// set). FP32 entries (8 bytes) are stored in plane pairs, entries 2h and
// 2h+1 of a task side by side, so that a task reads its kPlanes entries with
// kPlanes / 2 loads of 16 bytes (half the load instructions, and half the
// cache-line lookups where a workgroup's tasks sit at scattered positions).
// FP64 entries are split: the upwind positions four to 16 bytes (entry q of
// task p at uint32 index ((q / 4) * W + p) * 4 + q % 4), then the weights two
// to 16 bytes (at double index 6 * W + ((q / 2) * W + p) * 2 + q % 2), so that
// a task reads them with kPlanes / 4 + kPlanes / 2 loads of 16 bytes and no
// padding; the last quarter of the level set's storage is unused.
template <typename R>
__host__ __device__ __forceinline__ size_t plane_off(uint32_t q, uint32_t W, uint32_t p) {
    return ((size_t)(q >> 1) * W + p) * 2 + (q & 1u);  // FP32
}
__host__ __device__ __forceinline__ size_t plane_nb_off(uint32_t q, uint32_t W, uint32_t p) {
    return ((size_t)(q >> 2) * W + p) * 4 + (q & 3u);  // FP64 positions, uint32 index
}
__host__ __device__ __forceinline__ size_t plane_w_off(uint32_t q, uint32_t W, uint32_t p) {
    return (size_t)6 * W + ((size_t)(q >> 1) * W + p) * 2 + (q & 1u);  // FP64 weights, double index
}
static_assert(kPlanes % 4 == 0, "FP64 upwind positions are stored four to 16 bytes");

struct Geo {
    int64_t N;
    int D;
    int L;
    const uint32_t* off;
    const double4* geo;
    const uint32_t* nk;
    const double* dirs;      // D x 3
    const uint8_t* level;
    const uint32_t* pslot;   // N: periodic slot or kNone
    const uint32_t* pcells;  // Nper
    const uint32_t* pfoff;   // Nper + 1: periodic faces of each periodic cell
    const uint32_t* pfidx;   // face indices
    int64_t Nper;
    uint4* flat;             // Kahn layering: task records in level order (D*N)
    uint32_t* lstart;        // Kahn layering: level starts
    uint32_t* ctr;           // Kahn layering: 3 rotating append counters
    uint32_t* cnt;           // D x N missing-upwind counters
    uint32_t* need;          // L: the active set of a sweep level changed
};

// Per-cell {transport source / D in the storage scale, attenuation} as the
// transport reads them: in its storage precision (the FP32 transport rounds
// both to float before use, so storing them rounded changes no result).
template <typename R>
struct SAOf {
    using T = double2;
};
template <>
struct SAOf<float> {
    using T = float2;
};
template <typename R>
using SA = typename SAOf<R>::T;

template <typename R>
struct Flux {
    R* X;                    // D x N, at xidx(d, c)
    const uint32_t* xpos;    // compiled replay: X of (c, d) at the task's schedule position xpos[d*N + c]
    R* I;                    // D x N
    R* P;                    // D x Nper
    // After a rotation of the directions (ssw_rotate_directions), the rates
    // that the remap of the state left in place, as offsets to the pull sums:
    // the incoming rate of (c, d) is offI[d*N + c] + sum of the upwind X, the
    // periodic source offP[d*Nper + slot] + its periodic sum. NULL before any
    // rotation (offsets 0).
    R* offI;                 // D x N
    R* offP;                 // D x Nper
    const SA<R>* sa;         // N: {source / D in the storage scale, attenuation} (one load per task)
    double thr;              // threshold in the storage scale
    double inv_lref;         // storage scale of rates
    double inv_aref;         // storage scale of areas
};

// Where X of (cell c, direction d) is stored: direction-major d*N + c, or in
// the compiled replay the task's position in the schedule (level-major), so
// that a level set's gathers read what the previous level sets wrote densely.
template <typename R>
__device__ __forceinline__ size_t xidx(const Flux<R>& fl, int64_t N, uint32_t d, uint32_t c) {
    const size_t o = (size_t)d * N + c;
    return fl.xpos ? (size_t)fl.xpos[o] : o;
}

// Descriptor ranges of a grouped launch, one per build pass (the group's
// last level set's roots first): workgroups [off[r], off[r+1]) solve pass r's
// descriptors. n = 0: no ranges (the workgroup order is the descriptor order).
struct PassTab {
    uint32_t off[kMaxGroup + 1];
    int n;
};

// The descriptor of workgroup b: the hardware deals workgroups to the dies
// round-robin (b mod kXcds). Inside the pass range [o, o + c) that holds b,
// the workgroups of one die are those with the same x = (b - o) mod kXcds;
// they are given the contiguous run of the range's descriptors that starts
// at o + x*(c / kXcds) + min(x, c mod kXcds), in order. Neighbouring cones,
// which read the same schedule and X lines, then share one die's L2, and
// every die still gets an eighth of every pass.
__device__ __forceinline__ uint32_t pass_item(const PassTab& pt, uint32_t b) {
    if (pt.n == 0) return b;
    int r = 0;
    while (r + 1 < pt.n && b >= pt.off[r + 1]) r++;
    const uint32_t o = pt.off[r], c = pt.off[r + 1] - o, i = b - o;
    const uint32_t x = i % kXcds, q = c / kXcds, rem = c % kXcds;
    return o + x * q + min(x, rem) + i / kXcds;
}

// Block b of nb equal blocks in die order: the hardware deals workgroups to
// the dies round-robin (b mod kXcds); die x takes the contiguous run of
// blocks [x*q + min(x, r), ...) (q = nb / kXcds, r = nb mod kXcds), so that
// neighbouring blocks, whose gathers share cache lines, share its L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t x = b % kXcds, q = nb / kXcds, r = nb % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

__device__ __forceinline__ uint4 task_record(uint32_t c, uint32_t f0, uint32_t f1, uint32_t d, uint32_t ps) {
    return make_uint4(c | ((f1 - f0) << kCellBits), f0, d, ps);
}

// Appends `rec` for every lane with `pred` at dst[*counter ...], one atomic per
// wave; lanes outside the current control flow take no part.
__device__ __forceinline__ void wave_append(bool pred, uint4 rec, uint32_t* counter, uint4* dst) {
    unsigned long long m = __ballot(pred);
    if (m == 0ull) return;
    unsigned long long act = __ballot(1);
    int leader = __ffsll((long long)act) - 1;
    int lane = __lane_id();
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (pred) dst[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = rec;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}

// Per-cell inputs of a task's outgoing rate, loaded as soon as the task's cell
// is known so that they arrive with the upwind gathers.
template <typename R>
struct CellIn {
    double src;
    double att;
    R per;
    R off;  // incoming-rate offset after a rotation (0 before)
};

template <typename R>
__device__ __forceinline__ CellIn<R> cell_inputs(const Geo& g, const Flux<R>& fl, uint32_t c, uint32_t d,
                                                 uint32_t ps) {
    CellIn<R> ci;
    const SA<R> sa = fl.sa[c];
    ci.src = sa.x;
    ci.att = sa.y;
    // the periodic load uses slot 0 for cells without one (always a valid index)
    const R p = fl.P ? fl.P[(size_t)d * g.Nper + (ps != kNone ? ps : 0u)] : (R)0;
    ci.per = (ps != kNone) ? p : (R)0;
    ci.off = fl.offI ? fl.offI[(size_t)d * g.N + c] : (R)0;
    return ci;
}

// Outgoing rate of a solved task (Site::get_rate + HydrogenOnly::get_outgoing_rate,
// src/sweep/site.rs:49-56, src/chemistry/hydrogen_only/mod.rs:72-88), stored as X = O / A.
// X goes to ox, I to oi; returns the stored X.
template <typename R>
__device__ __forceinline__ R finish_task(const Flux<R>& fl, size_t ox, size_t oi, R inc, const CellIn<R>& ci,
                                         R A) {
    if (fl.offI) {
        // after a rotation the incoming rate is the remapped rate plus the
        // corrections since, offset + upwind sum; its persistent clamp
        // (Photons::make_positive) sets the offset to -sum, so that the rate
        // is 0 now and grows by the later corrections only
        R t = ci.off + inc;
        if (t < (R)0) {
            fl.offI[oi] = -inc;
            t = 0;
        }
        inc = t;
    } else if (inc < (R)0) {
        inc = 0;  // Photons::make_positive
    }
    R in = (inc + (R)ci.src) + ci.per;
    // the threshold in the storage scale is compared in double: in FP32 it is
    // threshold / lref (e.g. 1e-5 / 2^160), far below the float range
    R out = ((double)in < fl.thr) ? (R)0 : in * (R)ci.att;
    const R xv = (A > (R)0) ? out / A : (R)0;
    fl.X[ox] = xv;
    fl.I[oi] = inc;
    return xv;
}

// The face loop of one task, kLanes lanes per task (one face each), shared by
// the Kahn build and the geometric replay. Lane `sub` handles faces sub,
// sub + kLanes, ...; the loop runs to the longest face list of the wave so
// that every lane takes part in the shuffles. The upwind sum is accumulated
// in face order on every lane of the group; A is the xor-butterfly over the
// lanes' partial sums.
//   kTransport  gather the upwind X and sum
//   kBuild      decrement the downwind neighbours' counters and queue the
//               tasks that become ready into the block queue `sq`
template <typename R, bool kTransport, bool kBuild>
__device__ __forceinline__ void task_faces(const Geo& g, const Flux<R>& fl, bool valid, uint4 r, uint32_t sub,
                                           uint32_t nfmax, double& A_out, R& inc_out, uint4* sq, uint32_t* sq_n) {
    const uint32_t nf = valid ? (r.x >> kCellBits) : 0u;
    const uint32_t d = r.z;
    const double dx = g.dirs[3 * d], dy = g.dirs[3 * d + 1], dz = g.dirs[3 * d + 2];
    const size_t dN = (size_t)d * g.N;
    const int lane = __lane_id();
    const int gbase = lane & ~(kLanes - 1);
    double A = 0.0;
    R inc = 0;
    for (uint32_t base = 0; base < nfmax; base += kLanes) {
        const uint32_t j = base + sub;
        const bool fa = j < nf;
        const uint32_t f = r.y + j;
        const double4 a = fa ? g.geo[f] : make_double4(0.0, 0.0, 0.0, 0.0);
        const uint32_t nk = fa ? g.nk[f] : (kKindBoundary << 30);
        const uint32_t nb = nk & kNbrMask;
        const bool local = (nk >> 30) == kKindLocal;
        const double dd = fdot(a, dx, dy, dz);
        const bool up = fa && local && dd < 0.0;
        const bool down = fa && dd > 0.0;
        // nb is a valid cell index on every lane (0 for lanes without a face
        // and for boundary faces), so the gathers and the neighbour loads are
        // unconditional and issue together; the products are selected after.
        const R w = up ? (R)(a.w * fl.inv_aref * -dd) : (R)0;
        const R xv = kTransport ? fl.X[dN + nb] : (R)0;
        const R v = up ? w * xv : (R)0;
        bool ready = false;
        uint4 rec = make_uint4(0, 0, 0, 0);
        if (kBuild) {
            const uint32_t f0n = g.off[nb], f1n = g.off[nb + 1], psn = g.pslot[nb];
            if (down && local) ready = atomicSub(&g.cnt[dN + nb], 1u) == 1u;
            rec = task_record(nb, f0n, f1n, d, psn);
        }
        if (down) A += a.w * fl.inv_aref * dd;
        if (kTransport) {
#pragma unroll
            for (int l = 0; l < kLanes; l++) inc += __shfl(v, gbase + l);
        }
        if (kBuild) {
            const unsigned long long m = __ballot(ready);
            if (m) {
                uint32_t b = 0;
                if (lane == 0) b = atomicAdd(sq_n, (uint32_t)__popcll(m));
                b = __shfl(b, 0);
                if (ready) sq[b + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = rec;
            }
        }
    }
#pragma unroll
    for (int m = kLanes / 2; m > 0; m >>= 1) A += __shfl_xor(A, m, kLanes);
    A_out = A;
    inc_out = inc;
}

// ---------------------------------------------------------------------------
// grid checks and statistics (once per grid)

// Every local face c->nb needs a mirror face nb->c with exactly the negated
// normal, which makes the upwind/downwind classification of the pair agree.
__global__ void k_check_mirror(Geo g, unsigned long long* bad) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    for (uint32_t f = g.off[c]; f < g.off[c + 1]; f++) {
        uint32_t nk = g.nk[f];
        if ((nk >> 30) != kKindLocal) continue;
        uint32_t nb = nk & kNbrMask;
        if ((int64_t)nb >= g.N || (int64_t)nb == c) {
            atomicAdd(bad, 1ull);
            continue;
        }
        double4 a = g.geo[f];
        bool found = false;
        for (uint32_t h = g.off[nb]; h < g.off[nb + 1]; h++) {
            uint32_t hk = g.nk[h];
            if ((hk >> 30) != kKindLocal || (int64_t)(hk & kNbrMask) != c) continue;
            double4 b = g.geo[h];
            if (b.x == -a.x && b.y == -a.y && b.z == -a.z) found = true;
        }
        if (!found) atomicAdd(bad, 1ull);
    }
}

// Thread t handles (cell t / D, direction t % D): the lanes of a wave share one
// or two cells, so the face loads are broadcasts.
__device__ __forceinline__ bool cell_dir(const Geo& g, uint32_t t, uint32_t& c, uint32_t& d) {
    c = t / (uint32_t)g.D;
    d = t - c * (uint32_t)g.D;
    return (int64_t)c < g.N;
}

// Per (cell, direction): downwind faces by kind (the byte model) and the
// number of upwind local faces beyond kPlanes (overflow entries).
__global__ void k_face_stats(Geo g, unsigned long long* out4, uint32_t* ovf_count) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c, d;
    unsigned long long n[4] = {0, 0, 0, 0};
    if (cell_dir(g, t, c, d)) {
        const double dx = g.dirs[3 * d], dy = g.dirs[3 * d + 1], dz = g.dirs[3 * d + 2];
        uint32_t nup = 0;
        for (uint32_t f = g.off[c]; f < g.off[c + 1]; f++) {
            const uint32_t kind = g.nk[f] >> 30;
            const double dd = fdot(g.geo[f], dx, dy, dz);
            if (dd > 0.0) n[kind]++;
            else if (dd < 0.0 && kind == kKindLocal) nup++;
        }
        const uint32_t extra = nup > (uint32_t)kPlanes ? nup - kPlanes : 0u;
        if (ovf_count) ovf_count[(size_t)d * g.N + c] = extra;
        n[3] = extra;
    }
    for (int k = 0; k < 4; k++) {
        unsigned long long v = n[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
        if (__lane_id() == 0 && v) atomicAdd(&out4[k], v);
    }
}

// Start of the overflow entries (upwind entries q >= kPlanes, in face order)
// of every (cell, direction) in the overflow list, ovf_start[d*N + c]
// (assigned by an atomic cursor); the schedule writes the entries.
__global__ void k_overflow(Geo g, const uint32_t* ovf_count, uint32_t* ovf_start, uint32_t* cursor) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c, d;
    if (!cell_dir(g, t, c, d)) return;
    const size_t o = (size_t)d * g.N + c;
    const uint32_t extra = ovf_count[o];
    ovf_start[o] = extra ? atomicAdd(cursor, extra) : 0u;
}

// ---------------------------------------------------------------------------
// Kahn layering (once per grid for the full graph; on demand for a restricted
// active set, for the parity hook)

// Resets the append counters and the first level start.
__global__ void k_kahn_reset(Geo g) {
    if (threadIdx.x < 3) g.ctr[threadIdx.x] = 0;
    if (threadIdx.x == 0) g.lstart[0] = 0;
}

// init_counts (src/sweep/mod.rs:346-386) + get_initial_tasks (:388-398) for the
// active set {level >= lvl}: counts the active local upwind neighbours of every
// active (cell, direction) and queues the tasks with count 0 as level set 0
// (unless only_counts). Thread t handles (cell t / D, direction t % D).
__global__ void k_count(Geo g, int lvl, int only_counts) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c, d;
    const bool in = cell_dir(g, t, c, d);
    const bool active = in && g.level[c] >= lvl;
    uint32_t n = 0, f0 = 0, f1 = 0;
    if (active) {
        const double dx = g.dirs[3 * d], dy = g.dirs[3 * d + 1], dz = g.dirs[3 * d + 2];
        f0 = g.off[c];
        f1 = g.off[c + 1];
        for (uint32_t f = f0; f < f1; f++) {
            const uint32_t nk = g.nk[f];
            if ((nk >> 30) != kKindLocal) continue;
            if (!(fdot(g.geo[f], dx, dy, dz) < 0.0)) continue;
            if (g.level[nk & kNbrMask] >= lvl) n++;
        }
    }
    if (in) g.cnt[(size_t)d * g.N + c] = active ? n : kNone;
    if (only_counts) return;
    const bool ready = active && n == 0;
    wave_append(ready, ready ? task_record(c, f0, f1, d, g.pslot[c]) : make_uint4(0, 0, 0, 0), &g.ctr[0], g.flat);
}

// Level set k+1 of the layering from level set k: kLanes lanes per task
// decrement the counters of the local downwind neighbours (an inactive
// neighbour's counter starts at 0xFFFFFFFF and never reaches 0); the tasks that
// become ready are gathered per block in LDS and appended with one atomic per
// block (handle_local_neighbour, src/sweep/mod.rs:487-503).
__global__ void __launch_bounds__(kBlock) k_kahn(Geo g, int k) {
    __shared__ uint4 sq[kQueueCap];
    __shared__ uint32_t sq_n, sq_base, sq_cnt;
    uint32_t* ls = g.lstart;
    uint32_t* ctr = g.ctr;
    const uint32_t beg = ls[k];
    const uint32_t end = beg + ctr[k % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ls[k + 1] = end;
        ctr[(k + 2) % 3] = 0;
    }
    if (beg >= end) return;
    uint32_t* next_ctr = ctr + (k + 1) % 3;
    const uint32_t sub = threadIdx.x % kLanes;
    const uint32_t stride = gridDim.x * kTasksPerBlock;
    Flux<float> fl{};
    if (threadIdx.x == 0) sq_n = 0;
    __syncthreads();
    for (uint32_t i0 = beg + blockIdx.x * kTasksPerBlock; i0 < end; i0 += stride) {
        const uint32_t i = i0 + threadIdx.x / kLanes;
        const bool valid = i < end;
        const uint4 r = valid ? g.flat[i] : make_uint4(0, 0, 0, 0);
        const uint32_t nfmax = wave_max(valid ? (r.x >> kCellBits) : 0u);
        double A;
        float inc;
        task_faces<float, false, true>(g, fl, valid, r, sub, nfmax, A, inc, sq, &sq_n);
        // flush the block queue into level set k+1
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t n = sq_n;
            sq_cnt = n;
            sq_base = n ? atomicAdd(next_ctr, n) : 0u;
            sq_n = 0;
        }
        __syncthreads();
        const uint32_t n = sq_cnt, b = end + sq_base;
        for (uint32_t t = threadIdx.x; t < n; t += kBlock) g.flat[b + t] = sq[t];
        __syncthreads();
    }
}

// Kahn depth of every task of a layering: entry i of the level-major array is
// in level set k with lstart[k] <= i < lstart[k+1].
__global__ void k_scatter_depth(Geo g, int kmax, uint32_t* depth) {
    const uint32_t total = g.lstart[kmax];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        int a = 0, b = kmax - 1;  // last k with lstart[k] <= i
        while (a < b) {
            int m = (a + b + 1) >> 1;
            if (g.lstart[m] <= i) a = m;
            else b = m - 1;
        }
        const uint4 r = g.flat[i];
        depth[(size_t)r.z * g.N + (r.x & kCellMask)] = (uint32_t)a;
    }
}

__global__ void k_iota(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// Schedule position of every task: xpos[sorted[p]] = p.
__global__ void k_invert(const uint32_t* sorted, uint32_t* xpos, uint32_t n) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) xpos[sorted[p]] = p;
}

// The full schedule from the tasks sorted by (depth, direction, cell): task t =
// d*N + c at position p of level set k (start b, width W). Compiled: header,
// A, plane entries and overflow entries, each entry naming the upwind task's
// schedule position (one thread per task, faces in order; A as the same
// 16-lane butterfly the geometric kernels form). Geometric: the task record.
template <typename R>
__global__ void k_schedule(Geo g, double inv_aref, const uint32_t* sorted, const uint32_t* fstart, int kmax,
                           int compiled, const uint32_t* xpos, const uint32_t* ovf_start, uint4* hdr, R* Av,
                           Entry<R>* planes, Entry<R>* ovf, uint4* recs) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if ((int64_t)p >= (int64_t)g.D * g.N) return;
    int a = 0, bb = kmax - 1;  // level set of position p
    while (a < bb) {
        int m = (a + bb + 1) >> 1;
        if (fstart[m] <= p) a = m;
        else bb = m - 1;
    }
    const uint32_t b = fstart[a], W = fstart[a + 1] - b;
    const uint32_t t = sorted[p];
    const uint32_t d = t / (uint32_t)g.N, c = t - d * (uint32_t)g.N;
    const uint32_t f0 = g.off[c], f1 = g.off[c + 1], ps = g.pslot[c];
    if (!compiled) {
        recs[p] = task_record(c, f0, f1, d, ps);
        return;
    }
    const double dx = g.dirs[3 * d], dy = g.dirs[3 * d + 1], dz = g.dirs[3 * d + 2];
    double part[kLanes];
#pragma unroll
    for (int l = 0; l < kLanes; l++) part[l] = 0.0;
    uint32_t q = 0;
    const uint32_t os = ovf_start[(size_t)d * g.N + c];
    for (uint32_t f = f0; f < f1; f++) {
        const double4 fa = g.geo[f];
        const uint32_t nk = g.nk[f];
        const double dd = fdot(fa, dx, dy, dz);
        if (dd > 0.0) {
            const uint32_t l = (f - f0) % kLanes;  // the lane that holds face f in the 16-lane kernels
            const double term = fa.w * inv_aref * dd;
#pragma unroll
            for (int u = 0; u < kLanes; u++)
                if ((uint32_t)u == l) part[u] += term;
        } else if (dd < 0.0 && (nk >> 30) == kKindLocal) {
            Entry<R> e;
            e.nb = xpos[(size_t)d * g.N + (nk & kNbrMask)];
            if constexpr (sizeof(R) == 8) e.pad = 0;
            e.w = (R)(fa.w * inv_aref * -dd);
            if (q < (uint32_t)kPlanes) {
                Entry<R>* base = planes + (size_t)kPlanes * b;
                if constexpr (sizeof(R) == 4) {
                    base[plane_off<R>(q, W, p - b)] = e;
                } else {
                    reinterpret_cast<uint32_t*>(base)[plane_nb_off(q, W, p - b)] = e.nb;
                    reinterpret_cast<double*>(base)[plane_w_off(q, W, p - b)] = e.w;
                }
            } else {
                ovf[os + q - kPlanes] = e;
            }
            q++;
        }
    }
    // the xor butterfly of the 16-lane kernels, as seen from lane 0
#pragma unroll
    for (int m = kLanes / 2; m > 0; m >>= 1) {
        double nxt[kLanes];
#pragma unroll
        for (int l = 0; l < kLanes; l++) nxt[l] = part[l] + part[l ^ m];
#pragma unroll
        for (int l = 0; l < kLanes; l++) part[l] = nxt[l];
    }
    hdr[p] = make_uint4(c | (d << kCellBits), ps, q > (uint32_t)kPlanes ? os : 0u, q);
    Av[p] = (R)part[0];
}

// ---------------------------------------------------------------------------
// compaction of the full schedule to an active set (per sweep level);
// chunks[j] = {level set, first position, end position, 0}, never across level sets

__device__ __forceinline__ uint32_t task_cell(const uint4* tasks, uint32_t p) { return tasks[p].x & kCellMask; }

__global__ void __launch_bounds__(kBlock) k_compact_count(Geo g, const uint4* tasks, const uint4* chunks,
                                                          uint32_t nchunks, int lvl, uint32_t* counts) {
    const uint32_t j = blockIdx.x;
    if (j >= nchunks) return;
    const uint4 ch = chunks[j];
    __shared__ uint32_t total;
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    uint32_t n = 0;
    for (uint32_t p = ch.y + threadIdx.x; p < ch.z; p += kBlock) n += g.level[task_cell(tasks, p)] >= lvl ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) n += __shfl_down(n, o);
    if (__lane_id() == 0 && n) atomicAdd(&total, n);
    __syncthreads();
    if (threadIdx.x == 0) counts[j] = total;
}

// Writes the active positions of each chunk, in order, relative to the start
// of their level set: idx[offset[j] + rank] = p - fstart[level].
__global__ void __launch_bounds__(kBlock) k_compact_write(Geo g, const uint4* tasks, const uint4* chunks,
                                                          uint32_t nchunks, const uint32_t* fstart, int lvl,
                                                          const uint32_t* offset, uint32_t* idx) {
    const uint32_t j = blockIdx.x;
    if (j >= nchunks) return;
    const uint4 ch = chunks[j];
    const uint32_t b = fstart[ch.x];
    __shared__ uint32_t wave_tot[kBlock / 64];
    uint32_t base = offset[j];
    const int lane = __lane_id(), wv = threadIdx.x / 64;
    for (uint32_t p0 = ch.y; p0 < ch.z; p0 += kBlock) {
        const uint32_t p = p0 + threadIdx.x;
        const bool act = p < ch.z && g.level[task_cell(tasks, p)] >= lvl;
        const unsigned long long m = __ballot(act);
        if (lane == 0) wave_tot[wv] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, all = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            if (w < wv) before += wave_tot[w];
            all += wave_tot[w];
        }
        if (act) idx[base + before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = p - b;
        base += all;
        __syncthreads();
    }
}

// Compacted level starts: cstart[k] = offset of the level set's first chunk.
__global__ void k_compact_starts(const uint32_t* first_chunk, const uint32_t* offset, int kmax, uint32_t total,
                                 uint32_t* cstart) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < kmax) cstart[k] = offset[first_chunk[k]];
    if (k == kmax) cstart[k] = total;
}

// ---------------------------------------------------------------------------
// replay kernels

// Geometric replay of one level set: `count` tasks, kLanes lanes per task; with
// an index list, task i is record recs[idx[i]].
template <typename R>
__global__ void __launch_bounds__(kBlock) k_sweep_rec(Geo g, Flux<R> fl, const uint4* recs, const uint32_t* idx,
                                                      uint32_t count) {
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kTasksPerBlock + threadIdx.x / kLanes;
    const uint32_t sub = threadIdx.x % kLanes;
    const bool valid = i < count;
    const uint32_t p = valid ? (idx ? idx[i] : i) : 0u;
    const uint4 r = valid ? recs[p] : make_uint4(0, 0, 0, 0);
    const uint32_t ps = r.w;
    const uint32_t c = r.x & kCellMask, d = r.z;
    const CellIn<R> ci = cell_inputs<R>(g, fl, c, d, ps);
    const uint32_t nfmax = wave_max(valid ? (r.x >> kCellBits) : 0u);
    double A;
    R inc;
    task_faces<R, true, false>(g, fl, valid, r, sub, nfmax, A, inc, nullptr, nullptr);
    if (valid && sub == 0) finish_task<R>(fl, (size_t)d * g.N + c, (size_t)d * g.N + c, inc, ci, (R)A);
}

// One task of the compiled replay in three steps, so that the kernels that
// solve several level sets in one workgroup can load the next level set's
// schedule data while the current level set's gathers are in flight:
//   load_task   the schedule data of the task at position p of a level set
//               of width W whose header, A and plane entries start at hdr,
//               Av and planes (independent loads, no dependence on X);
//   issue_task  the upwind X gathers (at the upwind tasks' schedule
//               positions; entries past the upwind count gather the task's
//               own X, a valid index, and are not summed) and the cell inputs;
//   finish_task_c  the upwind sum in entry order, the overflow entries, and
//               the stores (X at the task's schedule position `self`).
template <typename R>
struct TaskS {
    Entry<R> e[kPlanes];
    uint4 h;
    R A;
};

// Overflow entries of a task read together with its upwind gathers (the
// rest, rarer, one after the other).
constexpr int kEarlyOvf = 2;

template <typename R>
struct TaskG {
    R x[kPlanes];
    Entry<R> o[kEarlyOvf];  // first overflow entries (tasks with more than kPlanes upwind faces)
    CellIn<R> ci;
};

template <typename R>
__device__ __forceinline__ void load_task(TaskS<R>& t, const uint4* hdr, const R* Av, const Entry<R>* planes,
                                          uint32_t W, uint32_t p) {
    if constexpr (sizeof(R) == 4) {  // plane pairs: one 16-byte load per two entries
        const uint4* pp = reinterpret_cast<const uint4*>(planes);
#pragma unroll
        for (int h = 0; h < kPlanes / 2; h++) {
            const uint4 v = pp[(size_t)h * W + p];
            t.e[2 * h].nb = v.x;
            t.e[2 * h].w = __uint_as_float(v.y);
            t.e[2 * h + 1].nb = v.z;
            t.e[2 * h + 1].w = __uint_as_float(v.w);
        }
    } else {  // positions four to a 16-byte load, weights two
        const uint4* pn = reinterpret_cast<const uint4*>(planes);
        const double2* pw = reinterpret_cast<const double2*>(reinterpret_cast<const double*>(planes) + 6 * (size_t)W);
#pragma unroll
        for (int h = 0; h < kPlanes / 4; h++) {
            const uint4 v = pn[(size_t)h * W + p];
            t.e[4 * h].nb = v.x;
            t.e[4 * h + 1].nb = v.y;
            t.e[4 * h + 2].nb = v.z;
            t.e[4 * h + 3].nb = v.w;
        }
#pragma unroll
        for (int h = 0; h < kPlanes / 2; h++) {
            const double2 v = pw[(size_t)h * W + p];
            t.e[2 * h].w = v.x;
            t.e[2 * h + 1].w = v.y;
        }
    }
    t.h = hdr[p];
    t.A = Av[p];
}

template <typename R>
__device__ __forceinline__ void issue_task(TaskG<R>& G, const Geo& g, const Flux<R>& fl, const TaskS<R>& t,
                                           const Entry<R>* ovf, uint32_t self) {
    const uint32_t c = t.h.x & kCellMask, d = t.h.x >> kCellBits, ps = t.h.y, nup = t.h.w;
    if (nup > (uint32_t)kPlanes) {  // indexes clamped to the task's own overflow entries
        const uint32_t last = nup - kPlanes - 1u;
#pragma unroll
        for (int k = 0; k < kEarlyOvf; k++) G.o[k] = ovf[t.h.z + min((uint32_t)k, last)];
    }
    G.ci = cell_inputs<R>(g, fl, c, d, ps);
#pragma unroll
    for (int q = 0; q < kPlanes; q++) G.x[q] = fl.X[(uint32_t)q < nup ? t.e[q].nb : self];
}

template <typename R>
__device__ __forceinline__ void finish_task_c(const Geo& g, const Flux<R>& fl, const TaskS<R>& t, const TaskG<R>& G,
                                              const Entry<R>* ovf, uint32_t self) {
    const uint32_t c = t.h.x & kCellMask, d = t.h.x >> kCellBits, nup = t.h.w;
    R inc = 0;
#pragma unroll
    for (int q = 0; q < kPlanes; q++) inc += ((uint32_t)q < nup) ? t.e[q].w * G.x[q] : (R)0;
    if (nup > (uint32_t)kPlanes) {  // the overflow entries, in entry order
        R xo[kEarlyOvf];
#pragma unroll
        for (int k = 0; k < kEarlyOvf; k++) xo[k] = fl.X[G.o[k].nb];
#pragma unroll
        for (int k = 0; k < kEarlyOvf; k++)
            if ((uint32_t)(kPlanes + k) < nup) inc += G.o[k].w * xo[k];
        for (uint32_t q = kPlanes + kEarlyOvf; q < nup; q++) {
            const Entry<R> o = ovf[t.h.z + q - kPlanes];
            inc += o.w * fl.X[o.nb];
        }
    }
    finish_task<R>(fl, self, (size_t)d * g.N + c, inc, G.ci, t.A);
}

// The three steps back to back: the task at position p of the level set that
// starts at schedule position b and has width W.
template <typename R>
__device__ __forceinline__ void comp_task(const Geo& g, const Flux<R>& fl, const uint4* hdr, const R* Av,
                                          const Entry<R>* planes, const Entry<R>* ovf, uint32_t b, uint32_t W,
                                          uint32_t p) {
    TaskS<R> t;
    TaskG<R> G;
    load_task<R>(t, hdr, Av, planes, W, p);
    issue_task<R>(G, g, fl, t, ovf, b + p);
    finish_task_c<R>(g, fl, t, G, ovf, b + p);
}

// Compiled replay of one level set of width W: one lane per task. With an
// index list, lane i solves the task at position idx[i] of the level set.
template <typename R, int kRecBlock>
__global__ void __launch_bounds__(kRecBlock) k_sweep_comp(Geo g, Flux<R> fl, const uint4* hdr, const R* Av,
                                                          const Entry<R>* planes, const Entry<R>* ovf, uint32_t b,
                                                          uint32_t W, const uint32_t* idx, uint32_t count) {
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kRecBlock + threadIdx.x;
    if (i >= count) return;
    comp_task<R>(g, fl, hdr, Av, planes, ovf, b, W, idx ? idx[i] : i);
}

// Lanes of the workgroup that solves a run of narrow level sets: at most this
// many active tasks per level set (FP64 tasks hold twice the registers).
template <typename R>
constexpr int run_block() {
    return sizeof(R) == 4 ? kRunBlock : kRunBlock / 2;
}

// Compiled replay of a run of consecutive narrow level sets (at most
// run_block<R>() active tasks each) by one workgroup, level set after level
// set, with a workgroup barrier in between: every wave of the workgroup runs
// on one CU, whose L1 is coherent within the workgroup, so the X stored for
// one level set is seen by the next after the barrier. run[j] = {first
// position b, width W, first index list entry, active count} of the run's
// j-th level set; without an index list entry i is position i.
// Software pipelined: while level set j's gathers are in flight, level set
// j+1's schedule data, level set j+2's index entries and run[j+3] are loaded,
// none of which level set j waits for. Every load is unconditional (lanes
// without a task, and level sets past the run's end, read a valid clamped
// entry) so that the waits stay exact; only the stores are predicated.
template <typename R>
__global__ void __launch_bounds__(run_block<R>()) k_sweep_comp_run(Geo g, Flux<R> fl, const uint4* __restrict__ hdr,
                                                                   const R* __restrict__ Av,
                                                                   const Entry<R>* __restrict__ planes,
                                                                   const Entry<R>* __restrict__ ovf,
                                                                   const uint32_t* __restrict__ idx,
                                                                   const uint4* __restrict__ run, int levels) {
    const uint32_t i = threadIdx.x;
    const int last = levels - 1;
    // position of lane i in level set r (a valid position for every lane)
    auto pos = [&](const uint4& r) -> uint32_t {
        const uint32_t k = min(i, r.w - 1u);
        return idx ? idx[r.z + k] : k;
    };
    uint4 r = run[0];
    uint32_t p = pos(r);
    TaskS<R> cur;
    load_task<R>(cur, hdr + r.x, Av + r.x, planes + (size_t)kPlanes * r.x, r.y, p);
    uint4 rn = run[min(1, last)];
    uint32_t pn = pos(rn);
    uint4 rnn = run[min(2, last)];
    for (int j = 0; j < levels; j++) {
        TaskG<R> G;
        issue_task<R>(G, g, fl, cur, ovf, r.x + p);
        TaskS<R> nxt;
        load_task<R>(nxt, hdr + rn.x, Av + rn.x, planes + (size_t)kPlanes * rn.x, rn.y, pn);
        const uint32_t pnn = pos(rnn);
        const uint4 rnnn = run[min(j + 3, last)];
        if (i < r.w) finish_task_c<R>(g, fl, cur, G, ovf, r.x + p);
        __syncthreads();
        r = rn;
        p = pn;
        cur = nxt;
        rn = rnn;
        pn = pnn;
        rnn = rnnn;
    }
}

// Direction sweeps (compiled replay): the tasks of one direction depend only
// on tasks of the same direction, so one workgroup per direction solves its
// direction's whole sweep, level set after level set with a workgroup barrier
// in between, and no dependency crosses workgroups or launches. The schedule
// is sorted by (depth, direction, cell), so direction d's tasks of level set k
// are one range of it: run[d * kmax + k] = {first position b of level set k,
// its width W (bit 31: wait for all memory operations at the end of this
// level set), offset of d's tasks in it, their count}; d has levels[d]
// non-empty level sets, the first ones.
//
// In this mode X is stored in each direction's Kahn order: the task at offset
// j of d's tasks in level set k has X index d*N + q, q = (d's tasks in level
// sets before k) + j, and the upwind entries of the schedule name these
// indices. The workgroup keeps the last kR values it solved in an LDS ring
// (slot q mod kR), written for the last kR tasks of each level set, so that
// an upwind value with q >= (end of the current level set) - kR is read from
// LDS and the barrier between level sets waits for LDS operations only.
// Older values are read from memory: the host flags a level set for a full
// wait whenever the level sets since the last flagged one would hold more
// than kR tasks, so every value older than the ring was stored before a
// completed wait and barrier.
//
// In a restricted sweep (lvl > 0; the host passes 0 when every cell is
// active) every task still passes through: a task whose cell is below level
// lvl is not solved and puts its stored X into the ring. Software pipelined
// two level sets deep: while level set j is solved, level set j+1's cell
// inputs, first overflow entries (and level and stored X) and level set
// j+2's schedule data are loaded. Same arithmetic per task as the other
// replays.
constexpr int kDirBlock = 512;  // lanes of a direction sweep's workgroup
constexpr uint32_t kDirWait = 0x80000000u;      // run entry: a full wait at the end of this level set
constexpr uint32_t kDirWidth = 0x7FFFFFFFu;     // run entry: the level set's width

// overflow entries loaded ahead per task (FP32: tasks of up to 18 upwind
// faces in LDS only; FP64 keeps two, within its registers)
template <typename R>
constexpr int dir_ovf() {
    return sizeof(R) == 4 ? 6 : 2;
}

template <typename R>
constexpr uint32_t dir_ring() {
    return sizeof(R) == 4 ? 32768u : 16384u;  // 128 KiB of LDS
}

template <typename R>
__global__ void __launch_bounds__(kDirBlock) k_sweep_dir(Geo g, Flux<R> fl, const uint4* __restrict__ hdr,
                                                         const R* __restrict__ Av, const Entry<R>* __restrict__ planes,
                                                         const Entry<R>* __restrict__ ovf,
                                                         const uint4* __restrict__ runs, int kmax,
                                                         const uint32_t* __restrict__ levels, int lvl) {
    constexpr uint32_t kR = dir_ring<R>();
    constexpr int kDirOvf = dir_ovf<R>();
    __shared__ R ring[kR];
    const uint32_t i = threadIdx.x;
    const uint32_t d = blockIdx.x;
    const uint4* run = runs + (size_t)d * kmax;
    const int nl = (int)levels[d];
    if (nl == 0) return;
    const int last = nl - 1;
    const uint32_t dN = d * (uint32_t)g.N;
    // position in its level set of lane i's task in pass q0 (a valid one for every lane)
    auto pos = [&](const uint4& r, uint32_t q0) -> uint32_t { return r.z + min(q0 + i, r.w - 1u); };
    auto ld = [&](TaskS<R>& t, const uint4& r, uint32_t p) {
        load_task<R>(t, hdr + r.x, Av + r.x, planes + (size_t)kPlanes * r.x, r.y & kDirWidth, p);
    };
    // cell inputs, and in a restricted sweep the cell's level and the stored
    // X (compared and used only when the task is solved, a level set later)
    struct Pre {
        CellIn<R> ci;
        R xold;
        uint32_t lev;
        Entry<R> o[kDirOvf];  // the first overflow entries (index clamped to the task's own)
    };
    auto pre = [&](const TaskS<R>& t, uint32_t self) -> Pre {
        Pre q;
        const uint32_t c = t.h.x & kCellMask;
        q.ci = cell_inputs<R>(g, fl, c, d, t.h.y);
        if (t.h.w > (uint32_t)kPlanes) {
            const uint32_t lastq = t.h.w - kPlanes - 1u;
#pragma unroll
            for (int k = 0; k < kDirOvf; k++) q.o[k] = ovf[t.h.z + min((uint32_t)k, lastq)];
        } else {
#pragma unroll
            for (int k = 0; k < kDirOvf; k++) {
                q.o[k].nb = self;
                q.o[k].w = (R)0;
            }
        }
        if (lvl > 0) {
            q.lev = g.level[c];
            q.xold = fl.X[self];
        } else {
            q.lev = 0;
            q.xold = (R)0;
        }
        return q;
    };
    // an upwind value: from the ring when it holds it, else from memory
    auto up = [&](uint32_t nb, uint32_t hi) -> R {
        const uint32_t qu = nb - dN;
        return (qu + kR >= hi) ? ring[qu & (kR - 1u)] : fl.X[nb];
    };
    // the task's upwind sum in entry order, its stores, and its ring slot.
    // Waves whose tasks find every upwind value in the ring (and have at most
    // kDirOvf overflow entries, loaded a level set ahead) read LDS only, so
    // that no wait on memory loads (which would also wait for the loads of
    // the next level sets) enters the chain; other waves take the slow path.
    auto solve = [&](const TaskS<R>& t, const Pre& pr, uint32_t self, uint32_t hi, bool lane) {
        const uint32_t nup = t.h.w;
        bool near = nup <= (uint32_t)(kPlanes + kDirOvf);
#pragma unroll
        for (int q = 0; q < kPlanes; q++)
            near = near && ((uint32_t)q >= nup || (t.e[q].nb - dN) + kR >= hi);
#pragma unroll
        for (int k = 0; k < kDirOvf; k++) near = near && (pr.o[k].nb - dN) + kR >= hi;
        R x[kPlanes];
        R inc = 0;
        if (__ballot(!near) == 0ull) {
#pragma unroll
            for (int q = 0; q < kPlanes; q++) x[q] = ring[(((uint32_t)q < nup ? t.e[q].nb : self) - dN) & (kR - 1u)];
            R xo[kDirOvf];
#pragma unroll
            for (int k = 0; k < kDirOvf; k++) xo[k] = ring[(pr.o[k].nb - dN) & (kR - 1u)];
#pragma unroll
            for (int q = 0; q < kPlanes; q++) inc += ((uint32_t)q < nup) ? t.e[q].w * x[q] : (R)0;
#pragma unroll
            for (int k = 0; k < kDirOvf; k++)  // overflow entries, in entry order
                if ((uint32_t)(kPlanes + k) < nup) inc += pr.o[k].w * xo[k];
        } else {
#pragma unroll
            for (int q = 0; q < kPlanes; q++) x[q] = up((uint32_t)q < nup ? t.e[q].nb : self, hi);
#pragma unroll
            for (int q = 0; q < kPlanes; q++) inc += ((uint32_t)q < nup) ? t.e[q].w * x[q] : (R)0;
            for (uint32_t q = kPlanes; q < nup; q++) {  // overflow entries, in entry order
                const Entry<R> o = ovf[t.h.z + q - kPlanes];
                inc += o.w * up(o.nb, hi);
            }
        }
        if (lane) {
            const bool act = pr.lev >= (uint32_t)lvl;
            const R xv = act ? finish_task<R>(fl, self, (size_t)d * g.N + (t.h.x & kCellMask), inc, pr.ci, t.A)
                             : pr.xold;
            const uint32_t qs = self - dN;
            if (qs + kR >= hi) ring[qs & (kR - 1u)] = xv;
        }
    };
    // level sets j (r0), j+1 (r1), j+2 (r2) and their first q (past the last
    // level set the entries repeat the last one)
    uint4 r0 = run[0], r1 = run[min(1, last)], r2 = run[min(2, last)];
    uint32_t qs0 = 0;
    uint32_t qs1 = last >= 1 ? r0.w : 0u;
    uint32_t qs2 = last >= 2 ? qs1 + r1.w : qs1;
    uint32_t p0 = pos(r0, 0), p1 = pos(r1, 0);
    TaskS<R> t0, t1;
    ld(t0, r0, p0);
    ld(t1, r1, p1);
    Pre c0 = pre(t0, dN + qs0 + (p0 - r0.z));
    for (int j = 0; j < nl; j++) {
        const Pre c1 = pre(t1, dN + qs1 + (p1 - r1.z));
        const uint32_t p2 = pos(r2, 0);
        TaskS<R> t2;
        ld(t2, r2, p2);
        const uint4 r3 = run[min(j + 3, last)];
        const uint32_t hi = qs0 + r0.w;
        solve(t0, c0, dN + qs0 + (p0 - r0.z), hi, i < r0.w);
        for (uint32_t q0 = kDirBlock; q0 < r0.w; q0 += kDirBlock) {  // wide level sets: further passes
            const uint32_t pq = pos(r0, q0);
            const uint32_t self = dN + qs0 + (pq - r0.z);
            TaskS<R> t;
            ld(t, r0, pq);
            solve(t, pre(t, self), self, hi, q0 + i < r0.w);
        }
        if (r0.y & kDirWait) {  // every store so far complete before the barrier
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        } else {  // LDS operations only
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        }
        r0 = r1;
        p0 = p1;
        t0 = t1;
        c0 = c1;
        qs0 = qs1;
        r1 = r2;
        p1 = p2;
        t1 = t2;
        qs1 = qs2;
        qs2 = (j + 3 <= last) ? qs2 + r2.w : qs2;
        r2 = r3;
    }
}

// The lowest timestep level of the cells (direction sweeps: a restricted
// sweep whose active set is every cell runs as a full one).
__global__ void k_min_level(const uint8_t* __restrict__ level, int64_t n, uint32_t* out) {
    uint32_t m = 255u;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
        m = min(m, (uint32_t)level[c]);
    for (int o = 32; o > 0; o >>= 1) m = min(m, (uint32_t)__shfl_down((int)m, o));
    if (__lane_id() == 0) atomicMin(out, m);
}

// X index of every task in the direction sweeps' order: the task at position
// p of level set k (sorted[p] = d*N + c) goes to d*N + qstart + (p - b_k - off)
// with oq[k*D + d] = {off, qstart} of d's tasks in k.
__global__ void k_invert_dir(const uint32_t* __restrict__ sorted, const uint32_t* __restrict__ fstart, int kmax,
                             const uint2* __restrict__ oq, uint32_t N, uint32_t D, uint32_t* __restrict__ xpos,
                             uint32_t n) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int a = 0, bb = kmax - 1;  // level set of position p
    while (a < bb) {
        const int m = (a + bb + 1) >> 1;
        if (fstart[m] <= p) a = m;
        else bb = m - 1;
    }
    const uint32_t t = sorted[p];
    const uint32_t d = t / N;
    const uint2 v = oq[(size_t)a * D + d];
    xpos[t] = d * N + v.y + (p - fstart[a] - v.x);
}

// Tasks per (level set, direction): cnt[depth[t] * D + t / N] for every task t = d*N + c.
__global__ void k_dir_counts(const uint32_t* __restrict__ depth, uint32_t N, uint32_t D, uint32_t n,
                             uint32_t* __restrict__ cnt) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) atomicAdd(&cnt[(size_t)depth[t] * D + t / N], 1u);
}

// ---------------------------------------------------------------------------
// grouped replay (full sweeps): m consecutive level sets per launch. Each
// workgroup owns a block of root tasks (kGroupT consecutive tasks of one level
// set of the group) and solves the roots' upwind cone inside the group, level
// set by level set with a workgroup barrier in between, so that every
// dependency inside the launch stays inside the workgroup (L1 coherent within
// the workgroup). Cones of different workgroups may overlap: such a task is
// solved by each of them from the same inputs with the same arithmetic and
// stores the same values. The roots of level set j are the tasks of j in no
// cone of a higher level set of the group.

// Sorts a[0, n2) ascending (n2 a power of two) in LDS, all lanes taking part.
__device__ void lds_sort(uint32_t* a, uint32_t n2) {
    for (uint32_t k = 2; k <= n2; k <<= 1)
        for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
            for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) {
                const uint32_t ixj = i ^ jj;
                if (ixj > i) {
                    const uint32_t x = a[i], y = a[ixj];
                    if ((x > y) == ((i & k) == 0)) {
                        a[i] = y;
                        a[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
}

// Appends the upwind schedule positions in [lo, hi) of the task at position
// P (level set starting at b, width W) to cand.
template <typename R>
__device__ void push_upwind(const uint4* hdr, const Entry<R>* planes, const Entry<R>* ovf, uint32_t P, uint32_t b,
                            uint32_t W, uint32_t lo, uint32_t hi, uint32_t* cand, uint32_t* ncand, uint32_t* over) {
    const uint4 h = hdr[P];
    for (uint32_t q = 0; q < h.w; q++) {
        const Entry<R>* base = planes + (size_t)kPlanes * b;
        uint32_t v;
        if (q >= (uint32_t)kPlanes) v = ovf[h.z + q - kPlanes].nb;
        else if constexpr (sizeof(R) == 4) v = base[plane_off<R>(q, W, P - b)].nb;
        else v = reinterpret_cast<const uint32_t*>(base)[plane_nb_off(q, W, P - b)];
        if (v < lo || v >= hi) continue;
        const uint32_t k = atomicAdd(ncand, 1u);
        if (k < (uint32_t)kConeCap) cand[k] = v;
        else *over = 1u;
    }
}

// Workgroup descriptor of the grouped replay, two uint4: {list offset, level
// sets, counts of level sets 0|1, 2|3} and {counts 4|5, 6|7, 0, 0} (16 bits
// per count, level sets counted from the group's first).
__device__ __forceinline__ uint32_t desc_count(const uint4& d0, const uint4& d1, int ph) {
    const uint32_t w = ph < 2 ? d0.z : ph < 4 ? d0.w : ph < 6 ? d1.x : d1.y;
    return (ph & 1) ? (w >> 16) : (w & 0xFFFFu);
}

// One workgroup per block of at most T roots of one level set j of a group
// starting at level set k0, meta[block] = {k0, j, first root, roots}: its cone
// per level set (sorted, unique), written with the roots to the task list,
// level set k0 first, and its descriptor. Members and roots are marked done
// in todo. Roots are roots[first + r] or, without a root list, positions
// first + r. Pending upwind entries are kept per level set: those of level
// set i are taken out when i is processed.
template <typename R, int T>
__global__ void __launch_bounds__(T) k_cone(const uint4* hdr, const Entry<R>* planes, const Entry<R>* ovf,
                                            const uint32_t* fstart, const uint4* meta, const uint32_t* roots,
                                            uint8_t* todo, uint32_t* list, uint32_t list_cap, uint32_t* cursor,
                                            uint4* desc, uint32_t* fail) {
    __shared__ uint32_t candA[kConeCap], candB[kConeCap], tmp[kSortCap], memb[kConeCap];
    __shared__ uint32_t ncand, nkeep, ntmp, nmemb, over, base;
    __shared__ uint32_t cnt[kMaxGroup], psum[T];
    uint32_t* cand = candA;
    uint32_t* keep = candB;
    const uint32_t tid = threadIdx.x;
    const uint4 mb = meta[blockIdx.x];
    const int k0 = (int)mb.x, j = (int)mb.y;
    const uint32_t n = min((uint32_t)T, mb.w);
    if (tid == 0) {
        ncand = 0;
        nmemb = 0;
        over = 0;
    }
    if (tid < (uint32_t)kMaxGroup) cnt[tid] = 0;
    __syncthreads();
    const uint32_t lo = fstart[k0];
    uint32_t root = 0;
    if (tid < n) {
        root = roots ? roots[mb.z + tid] : mb.z + tid;
        push_upwind<R>(hdr, planes, ovf, root, fstart[j], fstart[j + 1] - fstart[j], lo, fstart[j], cand, &ncand,
                       &over);
    }
    __syncthreads();
    for (int i = j - 1; i >= k0; i--) {
        const uint32_t bi = fstart[i], ei = fstart[i + 1];
        if (tid == 0) {
            ntmp = 0;
            nkeep = 0;
        }
        __syncthreads();
        // pending entries of level set i to tmp, of lower level sets to keep
        const uint32_t nc = min(ncand, (uint32_t)kConeCap);
        for (uint32_t c = tid; c < nc; c += T) {
            const uint32_t v = cand[c];
            if (v >= bi) tmp[atomicAdd(&ntmp, 1u)] = v;  // v < ei: entries are below the level sets done
            else keep[atomicAdd(&nkeep, 1u)] = v;
        }
        __syncthreads();
        {
            uint32_t* x = cand;
            cand = keep;
            keep = x;
        }
        if (tid == 0) ncand = nkeep;
        const uint32_t nt = ntmp;
        uint32_t n2 = 1;
        while (n2 < nt) n2 <<= 1;
        for (uint32_t c = nt + tid; c < n2; c += T) tmp[c] = kNone;
        __syncthreads();
        lds_sort(tmp, n2);
        // unique, in order: each lane compacts a contiguous chunk
        const uint32_t per = (nt + T - 1) / T;
        const uint32_t c0 = min(nt, tid * per), c1 = min(nt, c0 + per);
        uint32_t u = 0;
        for (uint32_t c = c0; c < c1; c++) u += (c == 0 || tmp[c] != tmp[c - 1]) ? 1u : 0u;
        psum[tid] = u;
        __syncthreads();
        if (tid == 0) {
            uint32_t s = 0;
            for (int l = 0; l < T; l++) {
                const uint32_t x = psum[l];
                psum[l] = s;
                s += x;
            }
            cnt[i - k0] = s;
            if (nmemb + s > (uint32_t)kConeCap) over = 1u;
        }
        __syncthreads();
        if (over) break;
        const uint32_t m0 = nmemb;
        uint32_t w = m0 + psum[tid];
        for (uint32_t c = c0; c < c1; c++)
            if (c == 0 || tmp[c] != tmp[c - 1]) memb[w++] = tmp[c];
        __syncthreads();
        if (tid == 0) nmemb = m0 + cnt[i - k0];
        __syncthreads();
        const uint32_t m1 = nmemb;
        if (i > k0)
            for (uint32_t c = m0 + tid; c < m1; c += T)
                push_upwind<R>(hdr, planes, ovf, memb[c], bi, ei - bi, lo, bi, cand, &ncand, &over);
        __syncthreads();
    }
    if (over) {
        if (tid == 0) atomicOr(fail, 1u);
        return;
    }
    const uint32_t total = nmemb + n;
    if (tid == 0) {
        base = atomicAdd(cursor, total);
        if (base + total > list_cap) {
            atomicOr(fail, 2u);
            base = kNone;
        }
    }
    __syncthreads();
    if (base == kNone) return;
    // members were appended from level set j-1 down to k0; the list holds them
    // from k0 up, then the roots
    uint32_t src = nmemb, dst = 0;
    for (int i = k0; i < j; i++) {
        const uint32_t ci = cnt[i - k0];
        src -= ci;
        for (uint32_t c = tid; c < ci; c += T) {
            const uint32_t v = memb[src + c];
            list[base + dst + c] = v;
            todo[v] = 0;
        }
        dst += ci;
    }
    if (tid < n) {
        list[base + dst + tid] = root;
        todo[root] = 0;
    }
    if (tid == 0) {
        uint32_t c[kMaxGroup];
        for (int k = 0; k < kMaxGroup; k++) c[k] = 0;
        for (int i = k0; i < j; i++) c[i - k0] = cnt[i - k0];
        c[j - k0] = n;
        desc[2 * blockIdx.x] = make_uint4(base, (uint32_t)(j - k0 + 1), c[0] | (c[1] << 16), c[2] | (c[3] << 16));
        desc[2 * blockIdx.x + 1] = make_uint4(c[4] | (c[5] << 16), c[6] | (c[7] << 16), 0u, 0u);
    }
}

// Cone build: flags[p] = todo[p] for the positions of the level sets marked
// in lvlflag (binary search of p's level set in fstart), else 0.
__global__ void k_cone_flags(const uint32_t* fstart, int kmax, const uint8_t* lvlflag, const uint8_t* todo,
                             uint8_t* flags, uint32_t n) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int a = 0, b = kmax - 1;
    while (a < b) {
        const int m = (a + b + 1) >> 1;
        if (fstart[m] <= p) a = m;
        else b = m - 1;
    }
    flags[p] = lvlflag[a] ? todo[p] : (uint8_t)0;
}

// Cone build: for level set lv[i], the range [lo, hi) of the sorted root list
// (nsel entries) holding its positions.
__global__ void k_cone_bounds(const uint32_t* roots, const uint32_t* nsel, const uint32_t* fstart, const int* lv,
                              int nlv, uint32_t* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlv) return;
    const uint32_t n = *nsel;
    for (int e = 0; e < 2; e++) {
        const uint32_t key = fstart[lv[i] + e];
        uint32_t a = 0, z = n;
        while (a < z) {
            const uint32_t m = (a + z) >> 1;
            if (roots[m] < key) a = m + 1;
            else z = m;
        }
        out[2 * i + e] = a;
    }
}

// Descriptors into group order: out[i] = in[perm[i]] (two uint4 each).
__global__ void k_permute_desc(const uint4* in, const uint32_t* perm, uint4* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[2 * i] = in[2 * perm[i]];
    out[2 * i + 1] = in[2 * perm[i] + 1];
}

// Grouped replay of the level sets k0 .. k0+phases-1 with E lanes per
// workgroup: workgroup w solves the tasks of its descriptor, level set by
// level set, barrier in between. Each lane's first task of every level set is
// read from the list up front, and the next level set's schedule data is
// loaded while the current level set's gathers are in flight; loads are
// unconditional (lanes without a task read the level set's first position)
// so that the waits stay exact. Tasks beyond the first E of a level set are
// solved without this overlap.
// PH > 0: the group has exactly PH level sets (known at compile time, so no
// list entry is read and no schedule data is prefetched for a level set past
// the group's last); PH = 0: phases level sets, at most kMaxGroup.
template <typename R, int E, int PH>
__device__ __forceinline__ void group_body(const Geo& g, const Flux<R>& fl, const uint4* __restrict__ hdr,
                                           const R* __restrict__ Av, const Entry<R>* __restrict__ planes,
                                           const Entry<R>* __restrict__ ovf, const uint32_t* __restrict__ fstart,
                                           int k0, int phases, const uint4* __restrict__ desc,
                                           const uint32_t* __restrict__ list, uint32_t list_len, const PassTab& pt,
                                           uint32_t blk) {
    constexpr int MaxPh = PH > 0 ? PH : kMaxGroup;
    if constexpr (PH > 0) phases = PH;
    const uint32_t w = pass_item(pt, blk);
    const uint4 d0 = desc[2 * w], d1 = desc[2 * w + 1];
    const uint32_t tid = threadIdx.x;
    uint32_t n[MaxPh], off[MaxPh], bb[MaxPh], WW[MaxPh], P[MaxPh];
#pragma unroll
    for (int ph = 0; ph < MaxPh; ph++) {
        n[ph] = desc_count(d0, d1, ph);
        off[ph] = ph == 0 ? d0.x : off[ph - 1] + n[ph - 1];
        const int k = k0 + min(ph, phases - 1);
        bb[ph] = fstart[k];
        WW[ph] = fstart[k + 1] - bb[ph];
        const uint32_t v = list[min(off[ph] + tid, list_len - 1u)];
        P[ph] = (ph < phases && tid < n[ph]) ? v : bb[ph];
    }
    TaskS<R> cur;
    load_task<R>(cur, hdr + bb[0], Av + bb[0], planes + (size_t)kPlanes * bb[0], WW[0], P[0] - bb[0]);
#pragma unroll
    for (int ph = 0; ph < MaxPh; ph++) {
        if (ph < phases) {
            TaskG<R> G;
            issue_task<R>(G, g, fl, cur, ovf, P[ph]);
            TaskS<R> nxt;
            const int q = ph + 1 < MaxPh ? ph + 1 : ph;
            if (PH == 0 || ph + 1 < MaxPh)
                load_task<R>(nxt, hdr + bb[q], Av + bb[q], planes + (size_t)kPlanes * bb[q], WW[q], P[q] - bb[q]);
            if (tid < n[ph]) finish_task_c<R>(g, fl, cur, G, ovf, P[ph]);
            for (uint32_t i = tid + E; i < n[ph]; i += E)
                comp_task<R>(g, fl, hdr + bb[ph], Av + bb[ph], planes + (size_t)kPlanes * bb[ph], ovf, bb[ph],
                             WW[ph], list[off[ph] + i] - bb[ph]);
            __syncthreads();
            if (PH == 0 || ph + 1 < MaxPh) cur = nxt;
        }
    }
}

template <typename R, int E, int PH>
__global__ void __launch_bounds__(E) k_sweep_group(Geo g, Flux<R> fl, const uint4* __restrict__ hdr,
                                                   const R* __restrict__ Av, const Entry<R>* __restrict__ planes,
                                                   const Entry<R>* __restrict__ ovf,
                                                   const uint32_t* __restrict__ fstart, int k0, int phases,
                                                   const uint4* __restrict__ desc, const uint32_t* __restrict__ list,
                                                   uint32_t list_len, PassTab pt) {
    group_body<R, E, PH>(g, fl, hdr, Av, planes, ovf, fstart, k0, phases, desc, list, list_len, pt, blockIdx.x);
}

// One problem's grouped launch inside a packed launch: everything the grouped
// kernel reads of its handle, and its workgroups [start, start + blocks) of
// the packed grid (start a multiple of kXcds, so that the die-aware order of
// its descriptors holds as in a launch of its own).
template <typename R>
struct GroupJob {
    int64_t N, Nper;
    R* X;
    R* I;
    R* P;
    R* offI;
    const SA<R>* sa;
    double thr;
    const uint4* hdr;
    const R* A;
    const Entry<R>* planes;
    const Entry<R>* ovf;
    const uint32_t* fstart;
    const uint4* desc;
    const uint32_t* list;
    uint32_t list_len;
    int k0, phases;
    uint32_t start, blocks;
    PassTab pt;
};

constexpr int kMaxPack = 8;  // problems per packed launch (kernel arguments of ~1.5 KB)

template <typename R>
struct PackArgs {
    GroupJob<R> job[kMaxPack];
    int n;
};

// Grouped launches of up to kMaxPack problems in one launch: workgroup b runs
// the job whose range holds it, as k_sweep_group would run it alone.
template <typename R, int E, int PH>
__global__ void __launch_bounds__(E) k_sweep_group_pack(const PackArgs<R> a) {
    int j = 0;
    for (int k = 1; k < a.n; k++)
        if (blockIdx.x >= a.job[k].start) j = k;
    const GroupJob<R>& J = a.job[j];
    const uint32_t b = blockIdx.x - J.start;
    if (b >= J.blocks) return;  // padding up to the next job's die-aligned start
    Geo g{};
    g.N = J.N;
    g.Nper = J.Nper;
    Flux<R> fl{};
    fl.X = J.X;
    fl.I = J.I;
    fl.P = J.P;
    fl.offI = J.offI;
    fl.sa = J.sa;
    fl.thr = J.thr;
    group_body<R, E, PH>(g, fl, J.hdr, J.A, J.planes, J.ovf, J.fstart, J.k0, J.phases, J.desc, J.list, J.list_len,
                         J.pt, b);
}

// The grouped replay of a restricted sweep (active set {level >= lvl}):
// workgroup w filters descriptor w's task list to the active tasks, in order,
// into the same place of list_out, and, if any is left, writes its filtered
// descriptor to desc_tmp[w] and kept[w] = 1 (k_place_desc then packs the kept
// descriptors of each group in order), so that a launch starts only
// workgroups with active tasks; the active tasks and the kept descriptors of
// each group are summed into gcount and gblocks (gmap: group of each
// descriptor). The cones of the full graph contain the active cones, so every
// dependency of an active task inside a launch stays inside its workgroup;
// the workgroups of a launch are independent, so their order is free.
__global__ void __launch_bounds__(kBlock) k_group_compact(Geo g, const uint4* __restrict__ hdr, int lvl,
                                                          const uint4* __restrict__ desc_in,
                                                          const uint32_t* __restrict__ list_in,
                                                          const uint32_t* __restrict__ gmap, uint32_t ndesc,
                                                          uint4* desc_tmp, uint32_t* kept, uint32_t* list_out,
                                                          uint32_t* gcount, uint32_t* gblocks) {
    const uint32_t w = blockIdx.x;
    if (w >= ndesc) return;
    __shared__ uint32_t wave_tot[kBlock / 64];
    const uint4 d0 = desc_in[2 * w], d1 = desc_in[2 * w + 1];
    uint32_t m[kMaxGroup];
    const int lane = __lane_id(), wv = threadIdx.x / 64;
    uint32_t src = d0.x, dst = d0.x;
    for (int ph = 0; ph < kMaxGroup; ph++) {
        const uint32_t n = desc_count(d0, d1, ph);
        const uint32_t base = dst;
        for (uint32_t p0 = 0; p0 < n; p0 += kBlock) {
            const uint32_t i = p0 + threadIdx.x;
            uint32_t P = 0;
            bool act = false;
            if (i < n) {
                P = list_in[src + i];
                act = g.level[hdr[P].x & kCellMask] >= lvl;
            }
            const unsigned long long bal = __ballot(act);
            if (lane == 0) wave_tot[wv] = (uint32_t)__popcll(bal);
            __syncthreads();
            uint32_t before = 0, all = 0;
            for (int k = 0; k < kBlock / 64; k++) {
                if (k < wv) before += wave_tot[k];
                all += wave_tot[k];
            }
            if (act) list_out[dst + before + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = P;
            dst += all;
            __syncthreads();
        }
        m[ph] = dst - base;
        src += n;
    }
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int k = 0; k < kMaxGroup; k++) total += m[k];
        kept[w] = total ? 1u : 0u;
        if (total) {
            const uint32_t gr = gmap[w];
            atomicAdd(&gcount[gr], total);
            atomicAdd(&gblocks[gr], 1u);
            desc_tmp[2 * w] = make_uint4(d0.x, d0.y, m[0] | (m[1] << 16), m[2] | (m[3] << 16));
            desc_tmp[2 * w + 1] = make_uint4(m[4] | (m[5] << 16), m[6] | (m[7] << 16), 0u, 0u);
        }
    }
}

// Packs the kept filtered descriptors: descriptor w of group g (whose
// descriptors start at gbase[g]) goes to gbase[g] + (pos[w] - pos[gbase[g]]),
// pos the exclusive prefix sum of kept, so each group's kept descriptors keep
// their order (and with it the pass ranges of the die-aware launch order).
__global__ void k_place_desc(const uint4* __restrict__ desc_tmp, const uint32_t* __restrict__ kept,
                             const uint32_t* __restrict__ pos, const uint32_t* __restrict__ gmap,
                             const uint32_t* __restrict__ gbase, uint4* desc_out, uint32_t ndesc) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= ndesc || !kept[w]) return;
    const uint32_t b = gbase[gmap[w]];
    const uint32_t o = b + (pos[w] - pos[b]);
    desc_out[2 * o] = desc_tmp[2 * w];
    desc_out[2 * o + 1] = desc_tmp[2 * w + 1];
}

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx, uint32_t* out,
                             uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = src[idx[i]];
}

// ---------------------------------------------------------------------------
// per-sweep kernels

// Periodic source of every periodic cell from the current X, after each
// sweep: the chemistry of the sweep reads it (the reference's periodic
// corrections have all arrived when update_chemistry runs), and the next
// sweep's transport reads it as the value at its start (X does not change in
// between). Thread t handles (periodic cell t / D, direction t % D).
template <typename R>
__global__ void k_periodic(Geo g, Flux<R> fl) {
    if (fl.P == nullptr) return;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = t / (uint32_t)g.D;
    const uint32_t d = t - s * (uint32_t)g.D;
    if ((int64_t)s >= g.Nper) return;
    const double dx = g.dirs[3 * d], dy = g.dirs[3 * d + 1], dz = g.dirs[3 * d + 2];
    R acc = fl.offP ? fl.offP[(size_t)d * g.Nper + s] : (R)0;
    for (uint32_t e = g.pfoff[s]; e < g.pfoff[s + 1]; e++) {
        const uint32_t f = g.pfidx[e];
        const uint32_t nk = g.nk[f];
        const double4 a = g.geo[f];
        const double dd = fdot(a, dx, dy, dz);
        if (dd < 0.0) acc += (R)(a.w * fl.inv_aref * -dd) * fl.X[xidx(fl, g.N, d, nk & kNbrMask)];
    }
    fl.P[(size_t)d * g.Nper + s] = acc;
}

struct ChemArgs {
    double* rho;
    double* x;
    double* T;
    double* src;             // source per direction bin (SI)
    void* sa;                // SA<R>[N], .y: attenuation of the next sweep
    double* species_ts;
    double* prev_rate;
    double* change_ts;
    const double* vol;
    const double* size;
    double max_ts;
    double thr;
    double safety;
    double scale_factor;
    int prevent_cooling;
    double lref;             // storage scale of I (rate = I * lref)
    unsigned long long* failures;
    unsigned long long* active;
    unsigned long long* limiting;  // [3] TimescaleCounter: temperature, ionisation fraction, photon rate
};

// Adds the number of lanes with `pred` to *counter with one atomic per wave
// (lanes outside the current control flow take no part).
__device__ __forceinline__ void wave_count(bool pred, unsigned long long* counter) {
    const unsigned long long m = __ballot(pred);
    if (m && __lane_id() == __ffsll((long long)__ballot(1)) - 1) atomicAdd(counter, (unsigned long long)__popcll(m));
}

// update_chemistry (src/sweep/mod.rs:549-574) for the active cells, then the
// attenuation of the updated cell for the next sweep.
template <typename R>
__global__ void k_chemistry(Geo g, Flux<R> fl, ChemArgs ca, int lvl) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) g.need[lvl] = 0;  // the active-set list of this sweep level is current
    if (c >= g.N) return;
    int lev = g.level[c];
    {
        unsigned long long m = __ballot(lev >= lvl);
        if (__lane_id() == __ffsll((long long)__ballot(1)) - 1 && m)
            atomicAdd(ca.active, (unsigned long long)__popcll(m));
    }
    if (lev < lvl) return;
    const double dt = ca.max_ts * ssw::powi(0.5, lev);
    const uint32_t ps = g.pslot[c];
    const bool has_p = ps != kNone && fl.P;
    const double src = ca.src[c];
    double rate = 0.0;
    for (int d = 0; d < g.D; d++) {
        const double inc = (double)fl.I[(size_t)d * g.N + c] * ca.lref;
        // current periodic source (k_periodic ran after the sweep)
        const double per = has_p ? (double)fl.P[(size_t)d * g.Nper + ps] * ca.lref : 0.0;
        rate += (inc + src) + per;  // sum over d of Site::get_rate, in direction order
    }
    double rel;
    if (fabs(rate) < fabs(ca.thr)) {
        rel = 0.0;
    } else {  // PhotonRate::relative_change_to(..).abs() (src/chemistry/mod.rs:73-77)
        rel = fabs(fmin(fabs(fabs(rate - ca.prev_rate[c]) / rate), ssw::kInvEps));
    }
    ca.prev_rate[c] = rate;
    const double rate_ts = dt / rel;
    ssw::Cell cell;
    cell.x = ca.x[c];
    cell.T = ca.T[c];
    cell.rho = ca.rho[c];
    cell.vol = ca.vol[c];
    cell.len = ca.size[c];
    cell.rate = rate;
    cell.a = ca.scale_factor;
    ssw::Floor fl0{ca.prevent_cooling != 0, cell.T, cell.x};
    int proc = 0;
    bool failed = false;
    double used = ssw::perform_timestep(cell, fl0, dt, ca.safety, proc, failed);
    if (failed) atomicAdd(ca.failures, 1ull);
    ca.x[c] = cell.x;
    ca.T[c] = cell.T;
    ca.species_ts[c] = used;
    // Timescale::min (src/chemistry/timescale.rs:32-38): a tie goes to the chemistry
    const bool rate_limits = rate_ts < used;
    const double change = rate_limits ? rate_ts : used;
    ca.change_ts[c] = change;
    // TimescaleCounter::count (src/chemistry/timescale.rs:80-87)
    const int process = rate_limits ? 2 : proc;
    const bool counted = change < ca.max_ts;
    wave_count(counted && process == 0, &ca.limiting[0]);
    wave_count(counted && process == 1, &ca.limiting[1]);
    wave_count(counted && process == 2, &ca.limiting[2]);
    ((SA<R>*)ca.sa)[c].y = (R)ssw::non_absorbed_fraction(cell.rho, cell.x, cell.len);
}

// The optional chemistry outputs (src/sweep/chemistry_output.rs:25-67) of
// every cell: Sweep::get_solver (src/sweep/mod.rs:612-632) with the rate
// summed over directions in direction order as in update_chemistry, the
// current x and T, and no floor; output `which` of ssw::chemistry_output.
template <typename R>
__global__ void k_chem_outputs(Geo g, Flux<R> fl, ChemArgs ca, int which, double* out) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const uint32_t ps = g.pslot[c];
    const bool has_p = ps != kNone && fl.P;
    const double src = ca.src[c];
    double rate = 0.0;
    for (int d = 0; d < g.D; d++) {
        const double inc = (double)fl.I[(size_t)d * g.N + c] * ca.lref;
        const double per = has_p ? (double)fl.P[(size_t)d * g.Nper + ps] * ca.lref : 0.0;
        rate += (inc + src) + per;  // Site::get_rate (src/sweep/site.rs:53-56)
    }
    ssw::Cell cell;
    cell.x = ca.x[c];
    cell.T = ca.T[c];
    cell.rho = ca.rho[c];
    cell.vol = ca.vol[c];
    cell.len = ca.size[c];
    cell.rate = rate;
    cell.a = ca.scale_factor;
    out[c] = ssw::chemistry_output(cell, which);
}

// update_timestep_levels (src/sweep/mod.rs:576-589) for all cells. Marks the
// sweep levels whose active set changed.
__global__ void k_levels(Geo g, const double* change_ts, uint8_t* level, double max_ts, double safety,
                         int lowest) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    double desired = safety * change_ts[c];
    int lev = ssw::level_from_ratio(max_ts / desired, g.L);
    if (lev < lowest) lev = lowest;
    int old = level[c];
    level[c] = (uint8_t)lev;
    for (int l = 0; l < g.L; l++)
        if ((old >= l) != (lev >= l)) g.need[l] = 1;
}

// Pull evaluation of the current incoming / outgoing / periodic rates of every
// (cell, direction) in SI units (parity hooks and the photon-rate output).
template <typename R>
__global__ void k_fluxes(Geo g, Flux<R> fl, double lref, double* I_out, double* O_out, double* P_out,
                         double* photon_rate, const uint32_t* rate_idx) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    double total = 0.0;
    for (int d = 0; d < g.D; d++) {
        double dx = g.dirs[3 * d], dy = g.dirs[3 * d + 1], dz = g.dirs[3 * d + 2];
        double A = 0.0;
        R inc = 0, per = 0;
        for (uint32_t f = g.off[c]; f < g.off[c + 1]; f++) {
            double4 a = g.geo[f];
            uint32_t nk = g.nk[f];
            double dd = fdot(a, dx, dy, dz);
            if (dd > 0.0) {
                A += a.w * dd;
            } else if (dd < 0.0 && (nk >> 30) != kKindBoundary) {
                R v = (R)(a.w * fl.inv_aref * -dd) * fl.X[xidx(fl, g.N, d, nk & kNbrMask)];
                if ((nk >> 30) == kKindLocal) inc += v;
                else per += v;
            }
        }
        if (fl.offI) inc = fl.offI[(size_t)d * g.N + c] + inc;
        if (fl.offP && g.pslot[c] != kNone) per = fl.offP[(size_t)d * g.Nper + g.pslot[c]] + per;
        double inc_si = (double)inc * lref;
        total += inc_si;
        size_t o = (size_t)d * g.N + c;
        if (I_out) I_out[o] = inc_si;
        if (P_out) P_out[o] = (double)per * lref;
        if (O_out) O_out[o] = (double)fl.X[xidx(fl, g.N, (uint32_t)d, (uint32_t)c)] * (A * fl.inv_aref) * lref;
    }
    // rate_idx: the caller's index of each internal cell (NULL: internal order)
    if (photon_rate) photon_rate[rate_idx ? rate_idx[c] : c] = total;
}

// rotate_directions_system (src/sweep/direction/mod.rs:158-175), state part,
// one thread per (cell c, direction d) after the schedule of the rotated
// directions is built: remap (:190-205) leaves every rate of direction d equal
// to the old rate of the last direction times kept[d] (1 or 0), for incoming,
// outgoing and periodic rates alike. Old rates (FP64, SI, [d*N + c]) come
// from k_fluxes before the rotation. Writes X = O / A with the rotated
// downwind area, I, the periodic source, and the remapped I and periodic
// source into the offsets (k_rotate_offsets subtracts the pull sums).
template <typename R>
__global__ void k_rotate_state(Geo g, Flux<R> fl, const double* __restrict__ kept, const double* __restrict__ Iold,
                               const double* __restrict__ Oold, const double* __restrict__ Pold) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)g.D * g.N) return;
    const uint32_t d = (uint32_t)(t / g.N);
    const uint32_t c = (uint32_t)(t - (int64_t)d * g.N);
    const size_t last = (size_t)(g.D - 1) * g.N + c;
    const double k = kept[d];
    const double inc = Iold[last] * k, out = Oold[last] * k, per = Pold[last] * k;
    const double dx = g.dirs[3 * d], dy = g.dirs[3 * d + 1], dz = g.dirs[3 * d + 2];
    double A = 0.0;
    for (uint32_t f = g.off[c]; f < g.off[c + 1]; f++) {
        const double4 a = g.geo[f];
        const double dd = fdot(a, dx, dy, dz);
        if (dd > 0.0) A += a.w * fl.inv_aref * dd;
    }
    const R Ar = (R)A;
    const R o = (R)(out * fl.inv_lref);
    fl.X[xidx(fl, g.N, d, c)] = (Ar > (R)0) ? o / Ar : (R)0;
    const size_t oi = (size_t)d * g.N + c;
    fl.I[oi] = (R)(inc * fl.inv_lref);
    fl.offI[oi] = (R)(inc * fl.inv_lref);
    const uint32_t ps = g.pslot[c];
    if (ps != kNone && fl.P) {
        fl.P[(size_t)d * g.Nper + ps] = (R)(per * fl.inv_lref);
        fl.offP[(size_t)d * g.Nper + ps] = (R)(per * fl.inv_lref);
    }
}

// The offsets of the rotated state: remapped rate minus the pull sum of the
// rotated X over the upwind faces (local for I, periodic for the periodic
// source), so that offset + sum gives the remapped rate back and grows by the
// corrections of later solves, as the reference's accumulated rates do.
template <typename R>
__global__ void k_rotate_offsets(Geo g, Flux<R> fl) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)g.D * g.N) return;
    const uint32_t d = (uint32_t)(t / g.N);
    const uint32_t c = (uint32_t)(t - (int64_t)d * g.N);
    const double dx = g.dirs[3 * d], dy = g.dirs[3 * d + 1], dz = g.dirs[3 * d + 2];
    R inc = 0, per = 0;
    for (uint32_t f = g.off[c]; f < g.off[c + 1]; f++) {
        const double4 a = g.geo[f];
        const uint32_t nk = g.nk[f];
        const double dd = fdot(a, dx, dy, dz);
        if (dd < 0.0 && (nk >> 30) != kKindBoundary) {
            const R v = (R)(a.w * fl.inv_aref * -dd) * fl.X[xidx(fl, g.N, d, nk & kNbrMask)];
            if ((nk >> 30) == kKindLocal) inc += v;
            else per += v;
        }
    }
    const size_t oi = (size_t)d * g.N + c;
    fl.offI[oi] = fl.offI[oi] - inc;
    const uint32_t ps = g.pslot[c];
    if (ps != kNone && fl.P) fl.offP[(size_t)d * g.Nper + ps] = fl.offP[(size_t)d * g.Nper + ps] - per;
}

// dst[idx[i]] = src[i]: a per-cell array from the internal into the caller's order.
template <typename T>
__global__ void k_to_caller(const T* __restrict__ src, const uint32_t* __restrict__ idx, T* __restrict__ dst,
                            int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[idx[i]] = src[i];
}

// Direction-major copy of X from the schedule-position order.
template <typename R>
__global__ void k_gather_x(const R* X, const uint32_t* xpos, R* out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = X[xpos[i]];
}

// X from the direction-major order into the schedule-position order.
template <typename R>
__global__ void k_scatter_x(R* X, const uint32_t* xpos, const R* in, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) X[xpos[i]] = in[i];
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename R>
__global__ void k_attenuation(const double* rho, const double* x, const double* size, SA<R>* sa, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sa[i].y = (R)ssw::non_absorbed_fraction(rho[i], x[i], size[i]);
}

// ---------------------------------------------------------------------------
// host side

// TimestepState (src/sweep/timestep_state.rs:4-80)
struct TimestepState {
    double max_ts = 0;
    int L = 1;
    int lowest = 0;
    bool first_done = false;
    void init(double m, int n) {
        max_ts = m;
        L = n;
        lowest = n - 1;
        first_done = false;
    }
    std::vector<int> sweep_order() const {  // :22-27, 70-75
        int num = L - lowest;
        std::vector<int> out;
        for (uint32_t i = 0; i < (1u << (num - 1)); i++) {
            int fb = i ? __builtin_ctz(i) : num - 1;
            out.push_back(lowest + (num - 1 - fb));
        }
        return out;
    }
    void advance() {  // :37-48
        if (first_done && lowest > 0) lowest -= 1;
        if (!first_done) first_done = true;
    }
    double current_max_timestep() const { return max_ts * ssw::powi(0.5, lowest); }
};

unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

}  // namespace

// Shared with the library's other translation units (placement.hip): the
// thread's error message and the lock that device allocations hold shared
// while a stream capture holds it exclusively.
namespace ssw_detail {
void set_error(const std::string& msg) { g_err = msg; }
std::shared_mutex& capture_mutex() { return g_capture_mu; }
}  // namespace ssw_detail

struct ssw_sweep {
    ssw_sweep_parameters p{};
    int D = 0;
    std::vector<double> dirs;
    int device = 0;
    hipStream_t stream = nullptr;
    bool fp32 = false;
    int mode_request = SSW_TRANSPORT_MODE_AUTO;
    int rec_block = 64;                             // compiled replay block size
    bool graphs_on = true;                          // capture replays as graphs
    bool narrow_runs = true;                        // runs of narrow level sets in one workgroup
    bool dir_sweeps = false;                        // full sweeps: one workgroup per direction
    uint4* d_dirrun = nullptr;                      // direction sweeps' run table (D x kmax)
    uint32_t* d_dirlev = nullptr;                   // direction sweeps: level sets per direction
    std::vector<uint2> dir_oq;                      // direction sweeps: {offset, first q} per (level set, direction)
    int group_request = kGroupLevels;               // level sets per grouped launch (1: off)
    int group_block = kGroupT;                      // grouped replay: root tasks per workgroup
    // grouped replay of full sweeps (compiled): per group of level sets its
    // first level set, level-set count and workgroup descriptors
    struct Group {
        int k0, levels;
        uint32_t desc0, ndesc;
        PassTab passes;  // descriptor ranges of the build passes, relative to desc0
    };
    std::vector<Group> groups;
    uint4* d_gdesc = nullptr;
    uint32_t* d_glist = nullptr;
    uint32_t* d_gmap = nullptr;                     // group of each descriptor
    uint32_t ndesc = 0;
    // restricted sweeps (level >= 1): filtered descriptors and lists, active
    // tasks per group
    std::vector<uint4*> d_gdesc_l;
    std::vector<uint32_t*> d_glist_l;
    uint32_t* d_gcount = nullptr;
    uint32_t* d_gbase = nullptr;                    // first descriptor of each group
    uint32_t* d_gblocks = nullptr;                  // restricted sweep: workgroups with active tasks per group
    // global descriptor index of every pass boundary of every group (group
    // gi's npass + 1 boundaries from pbound_start[gi]); per restricted sweep
    // level the pass ranges of each group's kept descriptors
    uint32_t* d_pbound = nullptr;
    std::vector<uint32_t> pbound_start;
    // restricted sweeps' filter scratch, held from the cone build on: filtered
    // descriptors in place, kept flags, their prefix sum, the pass bounds'
    // packed positions and the scan's temporary storage
    uint4* d_gf_desc = nullptr;
    uint32_t *d_gf_kept = nullptr, *d_gf_pos = nullptr, *d_gf_pbval = nullptr;
    uint8_t* d_gf_scan = nullptr;
    size_t gf_scan_bytes = 0;
    std::vector<std::vector<PassTab>> gpass_l;
    std::vector<std::vector<uint32_t>> gcount, gblocks;
    int64_t group_tasks = 0;                        // tasks solved per full sweep by the grouped replay
    uint32_t glist_len = 0;                         // entries of the grouped replay's task list
    bool group_fallback = false;                    // groups requested but a cone did not fit
    bool compiled = false;
    // grid
    int64_t N = 0, F = 0, Nper = 0;
    bool have_grid = false, have_state = false;
    // internal cell order: internal index perm[c] for caller's cell c (Morton
    // order of the positions given with ssw_set_cell_positions), inv = perm^-1
    std::vector<double> pos;
    std::vector<uint32_t> perm, inv;
    bool permuted = false;
    uint32_t* d_inv = nullptr;                      // inv on the device (permuted grids only)
    uint32_t *d_off = nullptr, *d_nk = nullptr, *d_pslot = nullptr, *d_pcells = nullptr;
    std::vector<uint32_t> h_pcells;                 // internal cell of each periodic slot
    uint32_t *d_pfoff = nullptr, *d_pfidx = nullptr;
    double4* d_geo = nullptr;
    double *d_dirs = nullptr, *d_vol = nullptr, *d_size = nullptr;
    double aref = 1.0;
    double dw[3] = {0, 0, 0};
    int64_t ovf_total = 0;
    // Kahn layering scratch (full graph at upload, restricted sets for the
    // hooks): flat and cnt are held only while a layering or count runs
    uint4* d_flat = nullptr;
    uint32_t *d_lstart = nullptr, *d_ctr = nullptr, *d_cnt = nullptr, *d_need = nullptr;
    int kcap = 0;
    // full schedule
    int kmax = 0;
    std::vector<uint32_t> full_start;               // kmax + 1 level starts
    uint32_t* d_fstart = nullptr;
    uint4* d_recs = nullptr;                        // geometric
    uint4* d_hdr = nullptr;                         // compiled
    void* d_A = nullptr;
    void* d_planes = nullptr;
    uint32_t* d_ovf_start = nullptr;
    void* d_ovf = nullptr;
    uint32_t* d_xpos = nullptr;                     // compiled: schedule position of each (d, c)
    // compaction to active sets
    uint32_t nchunks = 0;
    uint4* d_chunks = nullptr;
    uint32_t* d_first_chunk = nullptr;              // kmax: first chunk of each level set
    uint32_t *d_ccount = nullptr, *d_coff = nullptr, *d_cstart = nullptr;
    void* d_scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    std::vector<uint32_t*> d_idx;                   // per sweep level >= 1: active positions
    std::vector<std::vector<uint32_t>> cstart;      // per sweep level: level starts in idx
    std::vector<int> list_ok;                       // per sweep level: idx / cstart current
    // per sweep level: the grouped replay's filter kept every task (the
    // active set is all cells), so the level replays the full sweep's launches
    std::vector<int> list_full;
    std::vector<int> replays;                       // per sweep level: replays since the list was made
    std::vector<hipGraphExec_t> graphs;             // per sweep level: captured replay
    // replay plan per sweep level: one launch per level set, or one launch per
    // run of consecutive narrow level sets (compiled replay)
    struct ReplayItem {
        int k;          // level set (single launch) or first level set of the run / group
        int levels;     // 0: single launch; else the run's non-empty level sets (or the group's level sets)
        uint32_t tab;   // the run's first entry in the run table (the group's index)
        int group;      // 1: a group of the grouped replay
    };
    std::vector<std::vector<ReplayItem>> plan;
    std::vector<std::vector<uint4>> runtab;         // host copies of the run tables
    std::vector<uint4*> d_runtab;
    // state
    double *d_rho = nullptr, *d_x = nullptr, *d_T = nullptr, *d_src = nullptr;
    void* d_sa = nullptr;                           // SA<R>: {transport source / D, attenuation}
    double *d_sts = nullptr, *d_prev = nullptr, *d_change = nullptr;
    unsigned long long* d_active = nullptr;
    unsigned long long* d_limit = nullptr;          // [3] limiting processes of the last step
    uint8_t* d_level = nullptr;
    void *d_X = nullptr, *d_I = nullptr, *d_P = nullptr;
    void *d_offI = nullptr, *d_offP = nullptr;      // rate offsets after a rotation (Flux::offI / offP)
    bool rotated = false;                           // directions rotated since ssw_create
    unsigned long long* d_fail = nullptr;
    double lref = 1.0;
    TimestepState ts;
    int64_t bytes = 0, peak_bytes = 0;
    std::unordered_map<void*, size_t> alloc;
    int64_t sweeps_last = 0;
    int64_t builds_last = 0;
    int64_t capture_failures = 0;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    hipEvent_t sync_ev = nullptr;                   // orders the handle's stream against a packed step
    uint64_t uid = 0;                               // process-unique handle id (packed-step graphs)
    uint64_t epoch = 0;                             // bumped whenever a replay plan or its graph changes
    double t_transport_total = 0, t_chem_total = 0, t_build_total = 0;
    int64_t n_transport = 0;
};

namespace {

template <typename T>
void dfree(ssw_sweep* s, T*& p) {
    if (p) {
        std::shared_lock<std::shared_mutex> lock(g_capture_mu);
        auto it = s->alloc.find((void*)p);
        if (it != s->alloc.end()) {
            s->bytes -= (int64_t)it->second;
            s->alloc.erase(it);
        }
        (void)hipFree((void*)p);
    }
    p = nullptr;
}

template <typename T>
int dalloc(ssw_sweep* s, T** p, size_t count) {
    dfree(s, *p);
    if (count == 0) count = 1;
    hipError_t e;
    {
        std::shared_lock<std::shared_mutex> lock(g_capture_mu);
        e = hipMalloc((void**)p, count * sizeof(T));
    }
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(SSW_ERR_OOM, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                     " bytes failed: " + hipGetErrorString(e));
    }
    s->bytes += (int64_t)(count * sizeof(T));
    s->peak_bytes = std::max(s->peak_bytes, s->bytes);
    s->alloc[(void*)*p] = count * sizeof(T);
    return SSW_OK;
}

// Runs its function when the scope ends, also on an early error return: the
// transient device buffers of a call are freed on every path.
struct Defer {
    std::function<void()> f;
    ~Defer() { f(); }
};

size_t real_size(const ssw_sweep* s) { return s->fp32 ? 4 : 8; }
size_t entry_size(const ssw_sweep* s) { return s->fp32 ? sizeof(Entry<float>) : sizeof(Entry<double>); }
size_t sa_size(const ssw_sweep* s) { return s->fp32 ? sizeof(float2) : sizeof(double2); }

Geo make_geo(ssw_sweep* s) {
    Geo g;
    g.N = s->N;
    g.D = s->D;
    g.L = s->p.num_timestep_levels;
    g.off = s->d_off;
    g.geo = s->d_geo;
    g.nk = s->d_nk;
    g.dirs = s->d_dirs;
    g.level = s->d_level;
    g.pslot = s->d_pslot;
    g.pcells = s->d_pcells;
    g.pfoff = s->d_pfoff;
    g.pfidx = s->d_pfidx;
    g.Nper = s->Nper;
    g.flat = s->d_flat;
    g.lstart = s->d_lstart;
    g.ctr = s->d_ctr;
    g.cnt = s->d_cnt;
    g.need = s->d_need;
    return g;
}

template <typename R>
Flux<R> make_flux(ssw_sweep* s) {
    Flux<R> f;
    f.X = (R*)s->d_X;
    f.xpos = s->compiled ? s->d_xpos : nullptr;
    f.I = (R*)s->d_I;
    f.P = (R*)s->d_P;
    f.offI = (R*)s->d_offI;
    f.offP = (R*)s->d_offP;
    f.sa = (const SA<R>*)s->d_sa;
    f.inv_lref = 1.0 / s->lref;
    f.inv_aref = 1.0 / s->aref;
    f.thr = s->p.significant_rate_threshold * f.inv_lref;
    return f;
}

// Morton (Z-order) code of a position quantised to 21 bits per axis.
uint64_t morton3(const double* p, const double* lo, const double* inv_ext) {
    uint64_t code = 0;
    uint64_t q[3];
    for (int k = 0; k < 3; k++) {
        double u = (p[k] - lo[k]) * inv_ext[k];
        u = std::min(std::max(u, 0.0), 1.0);
        q[k] = (uint64_t)(u * 2097151.0);
    }
    for (int b = 0; b < 21; b++)
        for (int k = 0; k < 3; k++) code |= ((q[k] >> b) & 1ull) << (3 * b + k);
    return code;
}

// The internal cell order for the current grid: Morton order of the given
// positions, or the caller's order.
void make_order(ssw_sweep* s, int64_t n) {
    s->permuted = (int64_t)s->pos.size() == 3 * n;
    s->perm.resize(n);
    s->inv.resize(n);
    if (!s->permuted) {
        for (int64_t c = 0; c < n; c++) s->perm[c] = s->inv[c] = (uint32_t)c;
        return;
    }
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL}, ie[3];
    for (int64_t c = 0; c < n; c++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], s->pos[3 * c + k]);
            hi[k] = std::max(hi[k], s->pos[3 * c + k]);
        }
    for (int k = 0; k < 3; k++) ie[k] = hi[k] > lo[k] ? 1.0 / (hi[k] - lo[k]) : 0.0;
    std::vector<std::pair<uint64_t, uint32_t>> key(n);
    for (int64_t c = 0; c < n; c++) key[c] = {morton3(&s->pos[3 * c], lo, ie), (uint32_t)c};
    std::sort(key.begin(), key.end());
    for (int64_t i = 0; i < n; i++) {
        s->inv[i] = key[i].second;
        s->perm[key[i].second] = (uint32_t)i;
    }
}

// Per-cell host array in the caller's order -> internal order.
template <typename T>
const T* to_internal(const ssw_sweep* s, const T* in, std::vector<T>& tmp) {
    if (!s->permuted) return in;
    tmp.resize(s->N);
    for (int64_t i = 0; i < s->N; i++) tmp[i] = in[s->inv[i]];
    return tmp.data();
}

// `blocks` consecutive per-cell arrays (e.g. one per direction) in internal
// order -> the caller's order, in place.
template <typename T>
void to_caller(const ssw_sweep* s, T* a, int64_t blocks) {
    if (!s->permuted) return;
    std::vector<T> tmp(s->N);
    for (int64_t b = 0; b < blocks; b++) {
        T* x = a + b * s->N;
        for (int64_t i = 0; i < s->N; i++) tmp[s->inv[i]] = x[i];
        std::copy(tmp.begin(), tmp.end(), x);
    }
}

int set_need_all(ssw_sweep* s, uint32_t v) {
    k_fill_u32<<<1, kBlock, 0, s->stream>>>(s->d_need, v, s->p.num_timestep_levels);
    HIP_TRY(hipGetLastError());
    return SSW_OK;
}

// The per-cell {source / D, attenuation} pairs to and from the device, in
// the transport's storage precision there (synchronous).
int upload_sa(ssw_sweep* s, const std::vector<double2>& sa) {
    const size_t n = sa.size();
    if (s->fp32) {
        std::vector<float2> f(n);
        for (size_t c = 0; c < n; c++) f[c] = make_float2((float)sa[c].x, (float)sa[c].y);
        HIP_TRY(hipMemcpyAsync(s->d_sa, f.data(), n * sizeof(float2), hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    } else {
        HIP_TRY(hipMemcpyAsync(s->d_sa, sa.data(), n * sizeof(double2), hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    return SSW_OK;
}

int download_sa(ssw_sweep* s, std::vector<double2>& sa) {
    const size_t n = sa.size();
    if (s->fp32) {
        std::vector<float2> f(n);
        HIP_TRY(hipMemcpyAsync(f.data(), s->d_sa, n * sizeof(float2), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        for (size_t c = 0; c < n; c++) sa[c] = make_double2((double)f[c].x, (double)f[c].y);
    } else {
        HIP_TRY(hipMemcpyAsync(sa.data(), s->d_sa, n * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    return SSW_OK;
}

void drop_list(ssw_sweep* s, int lvl) {
    s->epoch++;
    if (lvl >= (int)s->list_ok.size()) return;
    s->list_ok[lvl] = 0;
    s->list_full[lvl] = 0;
    s->replays[lvl] = 0;
    if (s->graphs[lvl]) {
        (void)hipGraphExecDestroy(s->graphs[lvl]);
        s->graphs[lvl] = nullptr;
    }
}

void drop_all_lists(ssw_sweep* s) {
    for (int l = 0; l < (int)s->list_ok.size(); l++) drop_list(s, l);
}

// The layering scratch: counters (cnt) and, with `flat`, the level-major task
// records of a layering.
int ensure_layering_scratch(ssw_sweep* s, bool flat = true) {
    const size_t DN = (size_t)s->D * s->N;
    if (!s->d_cnt) TRY(dalloc(s, &s->d_cnt, DN));
    if (flat && !s->d_flat) TRY(dalloc(s, &s->d_flat, DN));
    return SSW_OK;
}

void release_layering_scratch(ssw_sweep* s) {
    dfree(s, s->d_flat);
    dfree(s, s->d_cnt);
}

// Kahn layering of the active set {level >= lvl} into the scratch records:
// counting, then one launch per level set until the queue is empty. The
// launches go out in batches of kKahnBatch level sets between host reads of
// the level starts (a launch past the last level set finds it empty and only
// repeats its start); the grid of a batch is sized from the last width read.
// `expected` active tasks must all be placed, else the graph has a cycle.
constexpr int kKahnBatch = 32;

int kahn_layering(ssw_sweep* s, int lvl, int64_t expected, int* depth_out, std::vector<uint32_t>* starts) {
    Geo g = make_geo(s);
    k_kahn_reset<<<1, 64, 0, s->stream>>>(g);
    k_count<<<grid_for((int64_t)s->N * s->D), kBlock, 0, s->stream>>>(g, lvl, 0);
    HIP_TRY(hipGetLastError());
    if (starts) starts->assign(1, 0u);
    uint32_t c3[3];
    HIP_TRY(hipMemcpyAsync(c3, s->d_ctr, sizeof(c3), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (c3[0] == 0) {
        if (expected != 0)
            return fail(SSW_ERR_CYCLE, "dependency cycle: " + std::to_string(expected) +
                                           " cell-direction pairs never become ready");
        k_kahn<<<1, kBlock, 0, s->stream>>>(g, 0);  // no active task: one empty level set
        HIP_TRY(hipGetLastError());
        if (starts) starts->push_back(0u);
        *depth_out = 1;
        return SSW_OK;
    }
    uint32_t hint = c3[0], prev = 0;
    std::vector<uint32_t> part(kKahnBatch);
    for (int k = 0;;) {
        const int kend = std::min(k + kKahnBatch, s->kcap - 1);
        if (k >= kend) return fail(SSW_ERR_CYCLE, "level-set depth exceeds its bound");
        const unsigned blocks =
            (unsigned)std::max<int64_t>(1, std::min<int64_t>(16384, (2 * (int64_t)hint + kTasksPerBlock - 1) /
                                                                        kTasksPerBlock));
        for (int kk = k; kk < kend; kk++) k_kahn<<<blocks, kBlock, 0, s->stream>>>(g, kk);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(part.data(), s->d_lstart + k + 1, (size_t)(kend - k) * 4, hipMemcpyDeviceToHost,
                               s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        for (int kk = k; kk < kend; kk++) {
            const uint32_t cur = part[kk - k];
            if (cur == prev) {  // level set kk is empty: the layering ended with kk level sets
                if ((int64_t)prev != expected)
                    return fail(SSW_ERR_CYCLE, "dependency cycle: " + std::to_string(expected - (int64_t)prev) +
                                                   " cell-direction pairs never become ready");
                *depth_out = std::max(1, kk);
                return SSW_OK;
            }
            if (starts) starts->push_back(cur);
            hint = cur - prev;
            prev = cur;
        }
        k = kend;
    }
}

// Device bytes of the grouped replay for DN tasks, n cells and L timestep
// levels with T roots per workgroup: the full graph's task list (bound: DN +
// DN/4 + 2^20 entries), descriptors and group map, and per restricted sweep
// level a filtered list and descriptors (replacing the active-set lists).
int64_t group_bytes(int64_t DN, int64_t n, int64_t L, int64_t T) {
    const int64_t kmx = std::min<int64_t>(n + 2, 1 << 22);
    const int64_t dcap = DN / T + kmx + 1;
    const int64_t list = (DN + DN / 4 + (1 << 20)) * 4;
    // + the transient scratch of a restricted sweep's filter (descriptors in
    // place, kept flags, their prefix sum)
    return list + dcap * 36 + kmx * 4 + (L - 1) * (list + dcap * 32) + (L > 1 ? dcap * 40 + (1 << 20) : 0);
}

// The replay plan of sweep level lvl from its level starts (the full
// schedule's for level 0, the active-set list's otherwise): in the compiled
// replay, each maximal run of at least two consecutive non-empty level sets of
// at most run_block<R>() active tasks (empty level sets inside a run are skipped)
// becomes one single-workgroup launch over a run table entry per level set;
// every other non-empty level set is one launch.
int plan_replay(ssw_sweep* s, int lvl) {
    s->epoch++;  // graphs of packed sweeps captured from the old plan no longer match
    const std::vector<uint32_t>& fs = s->full_start;
    const std::vector<uint32_t>& cs = lvl == 0 ? fs : s->cstart[lvl];
    std::vector<ssw_sweep::ReplayItem>& plan = s->plan[lvl];
    std::vector<uint4>& tab = s->runtab[lvl];
    plan.clear();
    tab.clear();
    if (s->d_dirrun) {  // direction sweeps: one launch, restricted sweeps included
        plan.push_back({0, s->kmax, 0u, 2});
        dfree(s, s->d_runtab[lvl]);
        return SSW_OK;
    }
    if (!s->groups.empty()) {  // the grouped replay (restricted sweeps: groups with active tasks)
        for (size_t i = 0; i < s->groups.size(); i++)
            if (s->groups[i].ndesc && (lvl == 0 || s->gcount[lvl][i] > 0))
                plan.push_back({s->groups[i].k0, s->groups[i].levels, (uint32_t)i, 1});
        dfree(s, s->d_runtab[lvl]);
        return SSW_OK;
    }
    const bool runs = s->compiled && s->narrow_runs;
    const uint32_t rb = s->fp32 ? (uint32_t)run_block<float>() : (uint32_t)run_block<double>();
    int k = 0;
    while (k < s->kmax) {
        const uint32_t count = cs[k + 1] - cs[k];
        if (count == 0) {
            k++;
            continue;
        }
        if (!runs || count > rb) {
            plan.push_back({k, 0, 0u, 0});
            k++;
            continue;
        }
        const uint32_t t0 = (uint32_t)tab.size();
        int k1 = k;
        for (; k1 < s->kmax && cs[k1 + 1] - cs[k1] <= rb; k1++) {
            const uint32_t c = cs[k1 + 1] - cs[k1];
            if (c) tab.push_back(make_uint4(fs[k1], fs[k1 + 1] - fs[k1], cs[k1], c));
        }
        const int n = (int)(tab.size() - t0);
        if (n >= 2) {
            plan.push_back({k, n, t0, 0});
        } else {
            tab.resize(t0);
            plan.push_back({k, 0, 0u, 0});
        }
        k = k1;
    }
    if (!tab.empty()) {
        TRY(dalloc(s, &s->d_runtab[lvl], tab.size()));
        HIP_TRY(hipMemcpyAsync(s->d_runtab[lvl], tab.data(), tab.size() * sizeof(uint4), hipMemcpyHostToDevice,
                               s->stream));
    } else {
        dfree(s, s->d_runtab[lvl]);
    }
    return SSW_OK;
}

// The grouped replay of full sweeps (compiled replay, group_request >= 2):
// groups of m consecutive level sets; in each group, from its last level set
// down to its first, the level set's tasks in no cone yet become roots, in
// blocks of kGroupT, whose cones are built (k_cone). If a cone exceeds the
// LDS capacity or the task list its bound, full sweeps keep the per-level
// replay (no groups).
template <typename R>
int build_groups_t(ssw_sweep* s) {
    const int m = s->group_request;
    const size_t DN = (size_t)s->D * s->N;
    const std::vector<uint32_t>& fs = s->full_start;
    const int T = s->group_block;
    const size_t desc_cap = DN / T + (size_t)s->kmax + 1;  // at most ceil(width / T) per level set
    const size_t list_cap = DN + DN / 4 + (1u << 20);
    // groups of m level sets, of kMaxGroup where that many consecutive level
    // sets hold at most kNarrowGroup tasks each (the narrow ends of the
    // graph, whose cones stay small); per pass r the level set
    // k0 + levels - 1 - r of each group that has it (pass 0: the last level
    // sets, all tasks roots)
    s->groups.clear();
    int mmax = 0;
    for (int k0 = 0; k0 < s->kmax;) {
        int mk = m;
        if (m < kMaxGroup && k0 + kMaxGroup <= s->kmax) {
            bool narrow = true;
            for (int k = k0; narrow && k < k0 + kMaxGroup; k++) narrow = fs[k + 1] - fs[k] <= kNarrowGroup;
            if (narrow) mk = kMaxGroup;
        }
        const int lv = std::min(s->kmax, k0 + mk) - k0;
        s->groups.push_back({k0, lv, 0u, 0u, PassTab{}});
        mmax = std::max(mmax, lv);
        k0 += mk;
    }
    const size_t ng = s->groups.size();
    size_t pass_roots = 1;
    for (int r = 1; r < mmax; r++) {
        size_t w = 0;
        for (const ssw_sweep::Group& gr : s->groups)
            if (gr.levels > r) w += fs[gr.k0 + gr.levels - r] - fs[gr.k0 + gr.levels - 1 - r];
        pass_roots = std::max(pass_roots, w);
    }
    uint8_t *todo = nullptr, *flags = nullptr, *lvlflag = nullptr, *tmp = nullptr;
    uint32_t *roots = nullptr, *ctl = nullptr, *bounds = nullptr, *perm = nullptr;  // ctl: cursor, fail, selected
    int* lv = nullptr;
    uint4 *meta = nullptr, *desc_rm = nullptr;
    Defer cleanup{[&] {
        for (uint8_t** p : {&todo, &flags, &lvlflag, &tmp}) dfree(s, *p);
        for (uint32_t** p : {&roots, &ctl, &bounds, &perm}) dfree(s, *p);
        dfree(s, lv);
        dfree(s, meta);
        dfree(s, desc_rm);
    }};
    TRY(dalloc(s, &todo, DN));
    TRY(dalloc(s, &ctl, 3));
    TRY(dalloc(s, &s->d_glist, list_cap));
    TRY(dalloc(s, &desc_rm, 2 * desc_cap));
    HIP_TRY(hipMemsetAsync(todo, 1, DN, s->stream));
    HIP_TRY(hipMemsetAsync(ctl, 0, 12, s->stream));
    size_t tb = 0;
    if (mmax > 1) {
        TRY(dalloc(s, &flags, DN));
        TRY(dalloc(s, &lvlflag, (size_t)s->kmax));
        TRY(dalloc(s, &roots, pass_roots));
        TRY(dalloc(s, &lv, ng));
        TRY(dalloc(s, &bounds, 2 * ng));
        HIP_TRY(hipcub::DeviceSelect::Flagged(nullptr, tb, hipcub::CountingInputIterator<uint32_t>(0u), flags, roots,
                                              ctl + 2, (int)DN, s->stream));
        TRY(dalloc(s, &tmp, tb));
    }
    const uint4* hdr = s->d_hdr;
    const Entry<R>* planes = (const Entry<R>*)s->d_planes;
    const Entry<R>* ovf = (const Entry<R>*)s->d_ovf;
    // blocks of group g in pass r: [rstart[r][g], rstart[r][g] + rcount[r][g]) of the pass-major descriptors
    std::vector<std::vector<uint32_t>> rstart(mmax, std::vector<uint32_t>(ng, 0u)), rcount(rstart);
    uint32_t ndesc = 0;
    std::vector<uint4> hmeta;
    std::vector<uint8_t> hflag(s->kmax);
    std::vector<int> hlv;
    std::vector<uint32_t> hb;
    for (int r = 0; r < mmax; r++) {
        hmeta.clear();
        hlv.clear();
        const uint32_t* rl = nullptr;
        if (r > 0) {  // roots: the tasks of the pass's level sets in no cone yet, in position order
            std::fill(hflag.begin(), hflag.end(), 0);
            for (const ssw_sweep::Group& gr : s->groups)
                if (gr.levels > r) {
                    hflag[gr.k0 + gr.levels - 1 - r] = 1;
                    hlv.push_back(gr.k0 + gr.levels - 1 - r);
                }
            if (hlv.empty()) break;
            HIP_TRY(hipMemcpyAsync(lvlflag, hflag.data(), hflag.size(), hipMemcpyHostToDevice, s->stream));
            HIP_TRY(hipMemcpyAsync(lv, hlv.data(), hlv.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
            k_cone_flags<<<grid_for((int64_t)DN), kBlock, 0, s->stream>>>(s->d_fstart, s->kmax, lvlflag, todo, flags,
                                                                          (uint32_t)DN);
            size_t tbr = tb;
            HIP_TRY(hipcub::DeviceSelect::Flagged(tmp, tbr, hipcub::CountingInputIterator<uint32_t>(0u), flags, roots,
                                                  ctl + 2, (int)DN, s->stream));
            k_cone_bounds<<<grid_for((int64_t)hlv.size()), kBlock, 0, s->stream>>>(roots, ctl + 2, s->d_fstart, lv,
                                                                                    (int)hlv.size(), bounds);
            HIP_TRY(hipGetLastError());
            hb.resize(2 * hlv.size());
            HIP_TRY(hipMemcpyAsync(hb.data(), bounds, hb.size() * 4, hipMemcpyDeviceToHost, s->stream));
            HIP_TRY(hipStreamSynchronize(s->stream));
            rl = roots;
        }
        size_t li = 0;
        for (size_t gi = 0; gi < ng; gi++) {
            const ssw_sweep::Group& gr = s->groups[gi];
            if (gr.levels <= r) continue;
            const int j = gr.k0 + gr.levels - 1 - r;
            uint32_t first, n;
            if (r == 0) {
                first = fs[j];
                n = fs[j + 1] - fs[j];
            } else {
                first = hb[2 * li];
                n = hb[2 * li + 1] - hb[2 * li];
                li++;
            }
            rstart[r][gi] = ndesc + (uint32_t)hmeta.size();
            for (uint32_t b = 0; b < n; b += (uint32_t)T)
                hmeta.push_back(make_uint4((uint32_t)gr.k0, (uint32_t)j, first + b, std::min<uint32_t>(T, n - b)));
            rcount[r][gi] = ndesc + (uint32_t)hmeta.size() - rstart[r][gi];
        }
        if (hmeta.empty()) continue;
        if ((size_t)ndesc + hmeta.size() > desc_cap) return fail(SSW_ERR_HIP, "grouped replay: descriptor bound exceeded");
        TRY(dalloc(s, &meta, hmeta.size()));
        HIP_TRY(hipMemcpyAsync(meta, hmeta.data(), hmeta.size() * sizeof(uint4), hipMemcpyHostToDevice, s->stream));
        const unsigned nb = (unsigned)hmeta.size();
        if (T == 32)
            k_cone<R, 32><<<nb, 32, 0, s->stream>>>(hdr, planes, ovf, s->d_fstart, meta, rl, todo, s->d_glist,
                                                    (uint32_t)list_cap, ctl, desc_rm + 2 * (size_t)ndesc, ctl + 1);
        else if (T == 64)
            k_cone<R, 64><<<nb, 64, 0, s->stream>>>(hdr, planes, ovf, s->d_fstart, meta, rl, todo, s->d_glist,
                                                    (uint32_t)list_cap, ctl, desc_rm + 2 * (size_t)ndesc, ctl + 1);
        else
            k_cone<R, 128><<<nb, 128, 0, s->stream>>>(hdr, planes, ovf, s->d_fstart, meta, rl, todo, s->d_glist,
                                                      (uint32_t)list_cap, ctl, desc_rm + 2 * (size_t)ndesc, ctl + 1);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s->stream));  // meta is reallocated by the next pass
        ndesc += nb;
    }
    uint32_t hctl[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(hctl, ctl, 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s, meta);
    dfree(s, tmp);
    dfree(s, bounds);
    dfree(s, lv);
    dfree(s, roots);
    dfree(s, lvlflag);
    dfree(s, flags);
    dfree(s, ctl);
    dfree(s, todo);
    if (hctl[1] != 0) {  // a cone or the list did not fit: per-level replay
        s->groups.clear();
        dfree(s, desc_rm);
        dfree(s, s->d_glist);
        s->group_fallback = true;
        return SSW_OK;
    }
    // descriptors in group order (a group's workgroups contiguous)
    std::vector<uint32_t> hperm;
    hperm.reserve(ndesc);
    for (size_t gi = 0; gi < ng; gi++) {
        ssw_sweep::Group& gr = s->groups[gi];
        gr.desc0 = (uint32_t)hperm.size();
        gr.passes = PassTab{};
        for (int r = 0; r < mmax; r++) {
            if (rcount[r][gi] == 0) continue;
            gr.passes.off[gr.passes.n++] = (uint32_t)hperm.size() - gr.desc0;
            for (uint32_t k = 0; k < rcount[r][gi]; k++) hperm.push_back(rstart[r][gi] + k);
        }
        gr.ndesc = (uint32_t)hperm.size() - gr.desc0;
        gr.passes.off[gr.passes.n] = gr.ndesc;
    }
    TRY(dalloc(s, &s->d_gdesc, 2 * std::max<size_t>(1, ndesc)));
    TRY(dalloc(s, &perm, std::max<size_t>(1, ndesc)));
    if (ndesc) {
        HIP_TRY(hipMemcpyAsync(perm, hperm.data(), hperm.size() * 4, hipMemcpyHostToDevice, s->stream));
        k_permute_desc<<<grid_for(ndesc), kBlock, 0, s->stream>>>(desc_rm, perm, s->d_gdesc, ndesc);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s, perm);
    dfree(s, desc_rm);
    s->group_tasks = hctl[0];
    s->glist_len = std::max<uint32_t>(1u, hctl[0]);
    s->ndesc = ndesc;
    std::vector<uint32_t> gmap(std::max<uint32_t>(1u, ndesc), 0u);
    for (size_t i = 0; i < s->groups.size(); i++)
        for (uint32_t k = 0; k < s->groups[i].ndesc; k++) gmap[s->groups[i].desc0 + k] = (uint32_t)i;
    TRY(dalloc(s, &s->d_gmap, gmap.size()));
    HIP_TRY(hipMemcpyAsync(s->d_gmap, gmap.data(), gmap.size() * 4, hipMemcpyHostToDevice, s->stream));
    TRY(dalloc(s, &s->d_gcount, std::max<size_t>(1, s->groups.size())));
    TRY(dalloc(s, &s->d_gblocks, std::max<size_t>(1, s->groups.size())));
    std::vector<uint32_t> gbase(std::max<size_t>(1, s->groups.size()), 0u);
    for (size_t i = 0; i < s->groups.size(); i++) gbase[i] = s->groups[i].desc0;
    TRY(dalloc(s, &s->d_gbase, gbase.size()));
    HIP_TRY(hipMemcpyAsync(s->d_gbase, gbase.data(), gbase.size() * 4, hipMemcpyHostToDevice, s->stream));
    std::vector<uint32_t> pb;
    s->pbound_start.assign(s->groups.size(), 0u);
    for (size_t i = 0; i < s->groups.size(); i++) {
        const ssw_sweep::Group& gr = s->groups[i];
        s->pbound_start[i] = (uint32_t)pb.size();
        for (int r = 0; r <= gr.passes.n; r++) pb.push_back(gr.desc0 + gr.passes.off[r]);
    }
    if (pb.empty()) pb.push_back(0u);
    TRY(dalloc(s, &s->d_pbound, pb.size()));
    HIP_TRY(hipMemcpyAsync(s->d_pbound, pb.data(), pb.size() * 4, hipMemcpyHostToDevice, s->stream));
    if (s->p.num_timestep_levels > 1 && ndesc) {  // restricted sweeps filter the cones
        s->gf_scan_bytes = 0;
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, s->gf_scan_bytes, s->d_gf_kept, s->d_gf_pos, (int)ndesc + 1,
                                                 s->stream));
        TRY(dalloc(s, &s->d_gf_desc, 2 * (size_t)ndesc));
        TRY(dalloc(s, &s->d_gf_kept, (size_t)ndesc + 1));
        TRY(dalloc(s, &s->d_gf_pos, (size_t)ndesc + 1));
        TRY(dalloc(s, &s->d_gf_pbval, pb.size()));
        TRY(dalloc(s, &s->d_gf_scan, std::max<size_t>(s->gf_scan_bytes, 1)));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SSW_OK;
}

int build_groups(ssw_sweep* s) {
    s->groups.clear();
    s->group_tasks = 0;
    s->group_fallback = false;
    if (!s->compiled || s->group_request < 2 || s->d_dirrun) return SSW_OK;  // direction sweeps need no cones
    return s->fp32 ? build_groups_t<float>(s) : build_groups_t<double>(s);
}

// The run table of the direction sweeps from every task's depth: per
// direction d and level set k, {first position of k, width of k, offset of
// d's tasks in k, their count}, and d's number of level sets (its tasks of
// depth k exist for every k below its deepest, since a task of depth k has an
// upwind task of depth k - 1 in its own direction).
int build_dir_runs(ssw_sweep* s, const uint32_t* d_depth) {
    const size_t DN = (size_t)s->D * s->N;
    const size_t KD = (size_t)s->kmax * s->D;
    uint32_t* d_cnt = nullptr;
    Defer cleanup{[&] { dfree(s, d_cnt); }};
    TRY(dalloc(s, &d_cnt, KD));
    k_fill_u32<<<grid_for((int64_t)KD), kBlock, 0, s->stream>>>(d_cnt, 0u, (int64_t)KD);
    k_dir_counts<<<grid_for((int64_t)DN), kBlock, 0, s->stream>>>(d_depth, (uint32_t)s->N, (uint32_t)s->D,
                                                                  (uint32_t)DN, d_cnt);
    HIP_TRY(hipGetLastError());
    std::vector<uint32_t> cnt(KD);
    HIP_TRY(hipMemcpyAsync(cnt.data(), d_cnt, KD * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    const std::vector<uint32_t>& fs = s->full_start;
    std::vector<uint4> run(KD, make_uint4(0, 0, 0, 0));
    std::vector<uint32_t> lev(s->D, 0), qs(s->D, 0);
    s->dir_oq.assign(KD, make_uint2(0, 0));
    for (int k = 0; k < s->kmax; k++) {
        uint32_t off = 0;
        for (int d = 0; d < s->D; d++) {
            const uint32_t c = cnt[(size_t)k * s->D + d];
            run[(size_t)d * s->kmax + k] = make_uint4(fs[k], fs[k + 1] - fs[k], off, c);
            s->dir_oq[(size_t)k * s->D + d] = make_uint2(off, qs[d]);
            if (c) {
                if (lev[d] != (uint32_t)k) return fail(SSW_ERR_HIP, "direction sweeps: a direction skips a level set");
                lev[d] = (uint32_t)k + 1;
            }
            off += c;
            qs[d] += c;
        }
        if (off != fs[k + 1] - fs[k]) return fail(SSW_ERR_HIP, "direction sweeps: level set counts differ");
        if (off > kDirWidth) return fail(SSW_ERR_INVALID, "direction sweeps: a level set is too wide");
    }
    // full waits (kDirWait): at the end of level set k - 1 whenever the
    // level sets since the last wait, k included, would hold more tasks than
    // the LDS ring, so that every upwind value older than the ring was stored
    // before a completed wait and barrier
    const uint64_t ring = s->fp32 ? dir_ring<float>() : dir_ring<double>();
    for (int d = 0; d < s->D; d++) {
        uint64_t acc = 0;
        for (uint32_t k = 0; k < lev[d]; k++) {
            const uint64_t c = run[(size_t)d * s->kmax + k].w;
            if (k > 0 && acc + c > ring) {
                run[(size_t)d * s->kmax + k - 1].y |= kDirWait;
                acc = 0;
            }
            acc += c;
        }
    }
    TRY(dalloc(s, &s->d_dirrun, KD));
    TRY(dalloc(s, &s->d_dirlev, (size_t)s->D));
    HIP_TRY(hipMemcpyAsync(s->d_dirrun, run.data(), KD * sizeof(uint4), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_dirlev, lev.data(), (size_t)s->D * 4, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SSW_OK;
}

// The full schedule, once per grid: Kahn layering of the whole graph, each
// task's depth, a stable sort by depth of the tasks in (direction, cell)
// order, then the compiled or geometric per-task data in that order, and the
// chunk table of the compaction.
int build_schedule(ssw_sweep* s) {
    const size_t DN = (size_t)s->D * s->N;
    s->kcap = (int)std::min<int64_t>(s->N + 2, 1 << 22);
    TRY(dalloc(s, &s->d_lstart, (size_t)s->kcap + 2));
    TRY(ensure_layering_scratch(s));
    int depth = 0;
    TRY(kahn_layering(s, 0, (int64_t)DN, &depth, &s->full_start));
    s->kmax = depth;
    dfree(s, s->d_cnt);
    // sort the tasks by depth, stable over t = d*N + c (double-buffered: four
    // DN-word arrays and a small temporary)
    uint32_t *d_depth = nullptr, *d_keys = nullptr, *d_vals = nullptr, *d_vals2 = nullptr, *d_sorted = nullptr;
    uint8_t* d_tmp = nullptr;
    Defer cleanup{[&] {
        for (uint32_t** p : {&d_depth, &d_keys, &d_vals, &d_vals2, &d_sorted}) dfree(s, *p);
        dfree(s, d_tmp);
    }};
    TRY(dalloc(s, &d_depth, DN));
    Geo g = make_geo(s);
    k_scatter_depth<<<grid_for(std::min<int64_t>((int64_t)DN, 1 << 22)), kBlock, 0, s->stream>>>(g, s->kmax,
                                                                                                  d_depth);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->compiled && s->dir_sweeps) TRY(build_dir_runs(s, d_depth));
    dfree(s, s->d_flat);
    TRY(dalloc(s, &d_keys, DN));
    TRY(dalloc(s, &d_vals, DN));
    TRY(dalloc(s, &d_vals2, DN));
    k_iota<<<grid_for((int64_t)DN), kBlock, 0, s->stream>>>(d_vals, (uint32_t)DN);
    HIP_TRY(hipGetLastError());
    int bits = 1;
    while ((1ll << bits) <= s->kmax) bits++;
    hipcub::DoubleBuffer<uint32_t> kb(d_depth, d_keys), vb(d_vals, d_vals2);
    size_t tmp_bytes = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, kb, vb, (int)DN, 0, bits, s->stream));
    TRY(dalloc(s, &d_tmp, tmp_bytes));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp_bytes, kb, vb, (int)DN, 0, bits, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    // the sorted order moves to d_sorted: the buffer it is stays owned by
    // exactly one pointer, so that it is freed once
    d_sorted = vb.Current();
    if (d_sorted == d_vals) d_vals = nullptr;
    else d_vals2 = nullptr;
    dfree(s, d_tmp);
    dfree(s, d_keys);
    dfree(s, d_depth);
    dfree(s, d_vals);
    dfree(s, d_vals2);
    TRY(dalloc(s, &s->d_fstart, s->full_start.size()));
    HIP_TRY(hipMemcpyAsync(s->d_fstart, s->full_start.data(), s->full_start.size() * 4, hipMemcpyHostToDevice,
                           s->stream));
    const double inv_aref = 1.0 / s->aref;
    if (s->compiled) {
        TRY(dalloc(s, &s->d_xpos, DN));
        if (s->d_dirrun) {  // direction sweeps: X in each direction's Kahn order
            uint2* d_oq = nullptr;
            Defer free_oq{[&] { dfree(s, d_oq); }};
            TRY(dalloc(s, &d_oq, s->dir_oq.size()));
            HIP_TRY(hipMemcpyAsync(d_oq, s->dir_oq.data(), s->dir_oq.size() * sizeof(uint2), hipMemcpyHostToDevice,
                                   s->stream));
            k_invert_dir<<<grid_for((int64_t)DN), kBlock, 0, s->stream>>>(d_sorted, s->d_fstart, s->kmax, d_oq,
                                                                         (uint32_t)s->N, (uint32_t)s->D, s->d_xpos,
                                                                         (uint32_t)DN);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(s->stream));
            s->dir_oq.clear();
            s->dir_oq.shrink_to_fit();
        } else {
            k_invert<<<grid_for((int64_t)DN), kBlock, 0, s->stream>>>(d_sorted, s->d_xpos, (uint32_t)DN);
        }
        TRY(dalloc(s, &s->d_hdr, DN));
        TRY(dalloc(s, (uint8_t**)&s->d_A, DN * real_size(s)));
        TRY(dalloc(s, (uint8_t**)&s->d_planes, (size_t)kPlanes * DN * entry_size(s)));
        if (s->fp32)
            k_schedule<float><<<grid_for((int64_t)DN), kBlock, 0, s->stream>>>(
                g, inv_aref, d_sorted, s->d_fstart, s->kmax, 1, s->d_xpos, s->d_ovf_start, s->d_hdr,
                (float*)s->d_A, (Entry<float>*)s->d_planes, (Entry<float>*)s->d_ovf, nullptr);
        else
            k_schedule<double><<<grid_for((int64_t)DN), kBlock, 0, s->stream>>>(
                g, inv_aref, d_sorted, s->d_fstart, s->kmax, 1, s->d_xpos, s->d_ovf_start, s->d_hdr,
                (double*)s->d_A, (Entry<double>*)s->d_planes, (Entry<double>*)s->d_ovf, nullptr);
    } else {
        TRY(dalloc(s, &s->d_recs, DN));
        k_schedule<double><<<grid_for((int64_t)DN), kBlock, 0, s->stream>>>(
            g, inv_aref, d_sorted, s->d_fstart, s->kmax, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
            s->d_recs);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s, d_sorted);
    dfree(s, s->d_ovf_start);  // copied into the headers
    // chunk table: chunks of at most kChunk tasks inside one level set
    std::vector<uint4> chunks;
    std::vector<uint32_t> first(s->kmax);
    for (int k = 0; k < s->kmax; k++) {
        first[k] = (uint32_t)chunks.size();
        const uint32_t b = s->full_start[k], e = s->full_start[k + 1];
        if (b == e) chunks.push_back(make_uint4((uint32_t)k, b, e, 0u));
        for (uint32_t p = b; p < e; p += kChunk)
            chunks.push_back(make_uint4((uint32_t)k, p, std::min(e, p + (uint32_t)kChunk), 0u));
    }
    s->nchunks = (uint32_t)chunks.size();
    TRY(dalloc(s, &s->d_chunks, chunks.size()));
    TRY(dalloc(s, &s->d_first_chunk, first.size()));
    TRY(dalloc(s, &s->d_ccount, chunks.size()));
    TRY(dalloc(s, &s->d_coff, chunks.size()));
    TRY(dalloc(s, &s->d_cstart, (size_t)s->kmax + 1));
    HIP_TRY(hipMemcpyAsync(s->d_chunks, chunks.data(), chunks.size() * sizeof(uint4), hipMemcpyHostToDevice,
                           s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_first_chunk, first.data(), first.size() * 4, hipMemcpyHostToDevice, s->stream));
    s->scan_tmp_bytes = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, s->scan_tmp_bytes, s->d_ccount, s->d_coff, (int)s->nchunks,
                                             s->stream));
    TRY(dalloc(s, (uint8_t**)&s->d_scan_tmp, s->scan_tmp_bytes));
    HIP_TRY(hipStreamSynchronize(s->stream));
    const int L = s->p.num_timestep_levels;
    for (uint32_t*& p : s->d_idx) dfree(s, p);
    s->d_idx.assign(L, nullptr);
    s->cstart.assign(L, std::vector<uint32_t>());
    s->list_ok.assign(L, 0);
    s->list_full.assign(L, 0);
    s->replays.assign(L, 0);
    for (hipGraphExec_t ge : s->graphs)
        if (ge) (void)hipGraphExecDestroy(ge);
    s->graphs.assign(L, nullptr);
    for (uint4*& p : s->d_runtab) dfree(s, p);
    s->d_runtab.assign(L, nullptr);
    for (uint4*& p : s->d_gdesc_l) dfree(s, p);
    for (uint32_t*& p : s->d_glist_l) dfree(s, p);
    s->d_gdesc_l.assign(L, nullptr);
    s->d_glist_l.assign(L, nullptr);
    s->gcount.assign(L, std::vector<uint32_t>());
    s->gblocks.assign(L, std::vector<uint32_t>());
    s->runtab.assign(L, std::vector<uint4>());
    s->plan.assign(L, std::vector<ssw_sweep::ReplayItem>());
    TRY(build_groups(s));
    TRY(plan_replay(s, 0));
    return SSW_OK;
}

const uint4* schedule_tasks(ssw_sweep* s) { return s->compiled ? s->d_hdr : s->d_recs; }

// The active-set list of sweep level lvl >= 1: the positions of the full
// schedule whose cell has level >= lvl, per level set, and its level starts
// on the host. Sweep level 0 covers every task and needs no list.
int make_group_list(ssw_sweep* s, int lvl) {
    if (!s->d_gdesc_l[lvl]) TRY(dalloc(s, &s->d_gdesc_l[lvl], 2 * (size_t)std::max<uint32_t>(1u, s->ndesc)));
    if (!s->d_glist_l[lvl]) TRY(dalloc(s, &s->d_glist_l[lvl], s->glist_len));
    const size_t ng = std::max<size_t>(1, s->groups.size());
    HIP_TRY(hipMemsetAsync(s->d_gcount, 0, ng * 4, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_gblocks, 0, ng * 4, s->stream));
    Geo g = make_geo(s);
    const uint32_t nd = s->ndesc;
    const size_t npb = std::max<size_t>(1, s->pbound_start.empty() ? 1 : (size_t)s->pbound_start.back() +
                                                                          s->groups.back().passes.n + 1);
    std::vector<uint32_t> pbv(npb, 0u);
    if (nd) {
        // filter, then pack each group's kept descriptors in order (scratch
        // held since the cone build)
        uint4* tmp_desc = s->d_gf_desc;
        uint32_t *kept = s->d_gf_kept, *pos = s->d_gf_pos, *pbval = s->d_gf_pbval;
        uint8_t* scan_tmp = s->d_gf_scan;
        size_t tb = s->gf_scan_bytes;
        if (!tmp_desc || !kept || !pos || !pbval || !scan_tmp)
            return fail(SSW_ERR_INVALID, "grouped replay: restricted-sweep scratch missing");
        k_fill_u32<<<1, 1, 0, s->stream>>>(kept + nd, 0u, 1);
        k_group_compact<<<nd, kBlock, 0, s->stream>>>(g, s->d_hdr, lvl, s->d_gdesc, s->d_glist, s->d_gmap, nd,
                                                       tmp_desc, kept, s->d_glist_l[lvl], s->d_gcount,
                                                       s->d_gblocks);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(scan_tmp, tb, kept, pos, (int)nd + 1, s->stream));
        k_place_desc<<<grid_for(nd), kBlock, 0, s->stream>>>(tmp_desc, kept, pos, s->d_gmap, s->d_gbase,
                                                             s->d_gdesc_l[lvl], nd);
        k_gather_u32<<<grid_for((int64_t)npb), kBlock, 0, s->stream>>>(pos, s->d_pbound, pbval, (uint32_t)npb);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(pbv.data(), pbval, npb * 4, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    std::vector<uint32_t>& gc = s->gcount[lvl];
    gc.assign(ng, 0u);
    HIP_TRY(hipMemcpyAsync(gc.data(), s->d_gcount, ng * 4, hipMemcpyDeviceToHost, s->stream));
    std::vector<uint32_t>& gb = s->gblocks[lvl];
    gb.assign(ng, 0u);
    HIP_TRY(hipMemcpyAsync(gb.data(), s->d_gblocks, ng * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    // pass ranges of each group's kept descriptors
    if ((int)s->gpass_l.size() <= lvl) s->gpass_l.resize(lvl + 1);
    s->gpass_l[lvl].assign(s->groups.size(), PassTab{});
    for (size_t i = 0; i < s->groups.size(); i++) {
        const ssw_sweep::Group& gr = s->groups[i];
        PassTab& pt = s->gpass_l[lvl][i];
        pt.n = gr.passes.n;
        const uint32_t b = pbv[s->pbound_start[i]];
        for (int r = 0; r <= gr.passes.n; r++) pt.off[r] = pbv[s->pbound_start[i] + r] - b;
    }
    int64_t kept = 0;
    for (uint32_t v : gc) kept += v;
    // every list entry kept: the same tasks in the same per-workgroup order as
    // the full sweep, whose launches (in their die-aware order) are then used
    s->list_full[lvl] = (kept == s->group_tasks) ? 1 : 0;
    s->list_ok[lvl] = 1;
    s->replays[lvl] = 0;
    s->epoch++;
    if (s->graphs[lvl]) {
        (void)hipGraphExecDestroy(s->graphs[lvl]);
        s->graphs[lvl] = nullptr;
    }
    TRY(plan_replay(s, lvl));
    return SSW_OK;
}

int make_list(ssw_sweep* s, int lvl) {
    if (s->d_dirrun) {  // direction sweeps pass every task and test its level: no list
        uint32_t* d_min = nullptr;
        Defer free_min{[&] { dfree(s, d_min); }};
        TRY(dalloc(s, &d_min, 1));
        k_fill_u32<<<1, 1, 0, s->stream>>>(d_min, 255u, 1);
        k_min_level<<<std::min<unsigned>(grid_for(s->N), 1024u), kBlock, 0, s->stream>>>(s->d_level, s->N, d_min);
        HIP_TRY(hipGetLastError());
        uint32_t mn = 0;
        HIP_TRY(hipMemcpyAsync(&mn, d_min, 4, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        s->list_full[lvl] = mn >= (uint32_t)lvl ? 1 : 0;  // every cell active: a full sweep
        s->list_ok[lvl] = 1;
        s->replays[lvl] = 0;
        s->epoch++;
        if (s->graphs[lvl]) {
            (void)hipGraphExecDestroy(s->graphs[lvl]);
            s->graphs[lvl] = nullptr;
        }
        return plan_replay(s, lvl);
    }
    if (!s->groups.empty()) return make_group_list(s, lvl);
    const size_t DN = (size_t)s->D * s->N;
    if (!s->d_idx[lvl]) TRY(dalloc(s, &s->d_idx[lvl], DN));
    Geo g = make_geo(s);
    const uint4* tasks = schedule_tasks(s);
    k_compact_count<<<s->nchunks, kBlock, 0, s->stream>>>(g, tasks, s->d_chunks, s->nchunks, lvl, s->d_ccount);
    HIP_TRY(hipGetLastError());
    size_t tb = s->scan_tmp_bytes;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(s->d_scan_tmp, tb, s->d_ccount, s->d_coff, (int)s->nchunks, s->stream));
    k_compact_write<<<s->nchunks, kBlock, 0, s->stream>>>(g, tasks, s->d_chunks, s->nchunks, s->d_fstart, lvl,
                                                          s->d_coff, s->d_idx[lvl]);
    HIP_TRY(hipGetLastError());
    // total = offset + count of the last chunk
    uint32_t last[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(&last[0], s->d_coff + s->nchunks - 1, 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(&last[1], s->d_ccount + s->nchunks - 1, 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    k_compact_starts<<<grid_for((int64_t)s->kmax + 1), kBlock, 0, s->stream>>>(s->d_first_chunk, s->d_coff, s->kmax,
                                                                               last[0] + last[1], s->d_cstart);
    HIP_TRY(hipGetLastError());
    std::vector<uint32_t>& cs = s->cstart[lvl];
    cs.resize(s->kmax + 1);
    HIP_TRY(hipMemcpyAsync(cs.data(), s->d_cstart, (s->kmax + 1) * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    TRY(plan_replay(s, lvl));
    s->list_ok[lvl] = 1;
    s->replays[lvl] = 0;
    s->epoch++;
    if (s->graphs[lvl]) {
        (void)hipGraphExecDestroy(s->graphs[lvl]);
        s->graphs[lvl] = nullptr;
    }
    return SSW_OK;
}

// Launches one item of sweep level lvl's replay plan on stream st: a group
// of level sets (its cones one workgroup each), a run of narrow level sets
// (one workgroup), or one level set.
template <typename R>
void launch_item_t(ssw_sweep* s, const Geo& g, const Flux<R>& fl, int lvl, const ssw_sweep::ReplayItem& it,
                   hipStream_t st) {
    const std::vector<uint32_t>& fs = s->full_start;
    const bool all = lvl == 0 || s->list_full[lvl];
    const uint32_t* idx = all ? nullptr : s->d_idx[lvl];
    const Entry<R>* ovf = (const Entry<R>*)s->d_ovf;
    if (it.group == 2) {  // direction sweeps: one workgroup per direction
        k_sweep_dir<R><<<s->D, kDirBlock, 0, st>>>(g, fl, s->d_hdr, (const R*)s->d_A, (const Entry<R>*)s->d_planes,
                                                    ovf, s->d_dirrun, s->kmax, s->d_dirlev, all ? 0 : lvl);
        return;
    }
    if (it.group) {
        const ssw_sweep::Group& gr = s->groups[it.tab];
        const uint4* dsc = (all ? s->d_gdesc : s->d_gdesc_l[lvl]) + 2 * (size_t)gr.desc0;
        const uint32_t* gl = all ? s->d_glist : s->d_glist_l[lvl];
        const uint32_t nblk = all ? gr.ndesc : s->gblocks[lvl][it.tab];  // restricted: workgroups with active tasks
        const Entry<R>* pl = (const Entry<R>*)s->d_planes;
        // full sweeps: die-aware order inside each pass; restricted sweeps
        // launch their filtered descriptors in order
        const PassTab pt = all ? gr.passes : s->gpass_l[lvl][it.tab];
        // groups of exactly 4 level sets (all but the narrow and last ones)
        // run the kernel that knows its level-set count at compile time
        using GroupKernel = void (*)(Geo, Flux<R>, const uint4*, const R*, const Entry<R>*, const Entry<R>*,
                                     const uint32_t*, int, int, const uint4*, const uint32_t*, uint32_t, PassTab);
        const bool four = gr.levels == 4;
        GroupKernel kern;
        int lanes;
        if (s->group_block == 32) {
            lanes = 64;
            kern = four ? k_sweep_group<R, 64, 4> : k_sweep_group<R, 64, 0>;
        } else if (s->group_block == 64) {
            lanes = 128;
            kern = four ? k_sweep_group<R, 128, 4> : k_sweep_group<R, 128, 0>;
        } else {
            lanes = 256;
            kern = four ? k_sweep_group<R, 256, 4> : k_sweep_group<R, 256, 0>;
        }
        hipLaunchKernelGGL(kern, dim3(nblk), dim3(lanes), 0, st, g, fl, s->d_hdr, (const R*)s->d_A, pl, ovf,
                           s->d_fstart, gr.k0, gr.levels, dsc, gl, s->glist_len, pt);
        return;
    }
    if (it.levels > 0) {  // a run of narrow level sets, one workgroup
        k_sweep_comp_run<R><<<1, run_block<R>(), 0, st>>>(g, fl, s->d_hdr, (const R*)s->d_A,
                                                          (const Entry<R>*)s->d_planes, ovf, idx,
                                                          s->d_runtab[lvl] + it.tab, it.levels);
        return;
    }
    const std::vector<uint32_t>& cs = all ? fs : s->cstart[lvl];
    const int k = it.k;
    const uint32_t count = cs[k + 1] - cs[k];
    const uint32_t b = fs[k], W = fs[k + 1] - fs[k];
    const uint32_t* ik = idx ? idx + cs[k] : nullptr;
    if (s->compiled) {
        const uint4* hdr = s->d_hdr + b;
        const R* A = (const R*)s->d_A + b;
        const Entry<R>* planes = (const Entry<R>*)s->d_planes + (size_t)kPlanes * b;
        switch (s->rec_block) {
            case 128:
                k_sweep_comp<R, 128><<<(count + 127) / 128, 128, 0, st>>>(g, fl, hdr, A, planes, ovf, b, W, ik, count);
                break;
            case 256:
                k_sweep_comp<R, 256><<<(count + 255) / 256, 256, 0, st>>>(g, fl, hdr, A, planes, ovf, b, W, ik, count);
                break;
            default:
                k_sweep_comp<R, 64><<<(count + 63) / 64, 64, 0, st>>>(g, fl, hdr, A, planes, ovf, b, W, ik, count);
        }
    } else {
        k_sweep_rec<R><<<(count + kTasksPerBlock - 1) / kTasksPerBlock, kBlock, 0, st>>>(g, fl, s->d_recs + b, ik,
                                                                                        count);
    }
}

// Launches the replay of sweep level lvl's level sets on the handle's stream,
// one exact-size launch per plan item; with `ev` each launch is bracketed by
// events.
template <typename R>
void launch_replay_t(ssw_sweep* s, const Geo& g, int lvl, std::vector<std::pair<hipEvent_t, hipEvent_t>>* ev,
                     const std::function<hipEvent_t()>* mark) {
    const Flux<R> fl = make_flux<R>(s);
    for (const ssw_sweep::ReplayItem& it : s->plan[lvl]) {
        hipEvent_t a = ev ? (*mark)() : nullptr;
        launch_item_t<R>(s, g, fl, lvl, it, s->stream);
        if (ev) ev->push_back({a, (*mark)()});
    }
}

void launch_replay(ssw_sweep* s, const Geo& g, int lvl, std::vector<std::pair<hipEvent_t, hipEvent_t>>* ev,
                   const std::function<hipEvent_t()>* mark = nullptr) {
    if (s->fp32) launch_replay_t<float>(s, g, lvl, ev, mark);
    else launch_replay_t<double>(s, g, lvl, ev, mark);
}

// Captures the replay of sweep level lvl as one graph (nothing is executed). A
// failed capture leaves no graph; the level sets are then launched one by one
// (counted in ssw_info).
void capture(ssw_sweep* s, const Geo& g, int lvl) {
    hipGraph_t graph = nullptr;
    std::unique_lock<std::shared_mutex> lock(g_capture_mu);
    if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
        (void)hipGetLastError();
        s->capture_failures++;
        return;
    }
    launch_replay(s, g, lvl, nullptr);
    hipError_t ce = hipStreamEndCapture(s->stream, &graph);
    lock.unlock();
    if (ce != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        s->capture_failures++;
        return;
    }
    hipGraphExec_t ge = nullptr;
    hipError_t ie = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        (void)hipGetLastError();
        s->capture_failures++;
        return;
    }
    s->graphs[lvl] = ge;
}

ChemArgs chem_args(ssw_sweep* s) {
    ChemArgs ca;
    ca.rho = s->d_rho;
    ca.x = s->d_x;
    ca.T = s->d_T;
    ca.src = s->d_src;
    ca.sa = s->d_sa;
    ca.species_ts = s->d_sts;
    ca.prev_rate = s->d_prev;
    ca.change_ts = s->d_change;
    ca.vol = s->d_vol;
    ca.size = s->d_size;
    ca.max_ts = s->p.max_timestep;
    ca.thr = s->p.significant_rate_threshold;
    ca.safety = s->p.chemistry_timestep_safety_factor;
    ca.scale_factor = s->p.scale_factor;
    ca.prevent_cooling = s->p.prevent_cooling;
    ca.lref = s->lref;
    ca.failures = s->d_fail;
    ca.active = s->d_active;
    ca.limiting = s->d_limit;
    return ca;
}

template <typename R>
int run_step(ssw_sweep* s, double* elapsed) {
    Geo g = make_geo(s);
    Flux<R> fl = make_flux<R>(s);
    const ChemArgs ca = chem_args(s);
    const std::vector<int> order = s->ts.sweep_order();
    const int L = s->p.num_timestep_levels;
    k_fill_u32<<<1, 2, 0, s->stream>>>((uint32_t*)s->d_active, 0u, 2);
    k_fill_u32<<<1, 6, 0, s->stream>>>((uint32_t*)s->d_limit, 0u, 6);
    const unsigned gper = grid_for(std::max<int64_t>(s->Nper, 1) * s->D);
    const unsigned gcells = grid_for(s->N);
    s->sweeps_last = 0;
    s->builds_last = 0;
    s->t_transport_total = s->t_chem_total = s->t_build_total = 0;
    s->n_transport = 0;
    size_t evn = 0;
    std::function<hipEvent_t()> mark = [&](void) -> hipEvent_t {
        if (evn >= s->ev.size()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            s->ev.push_back(e);
        }
        hipEvent_t e = s->ev[evn++];
        (void)hipEventRecord(e, s->stream);
        return e;
    };
    std::vector<std::pair<hipEvent_t, hipEvent_t>> tr, ch, bu;
    // sweep levels whose active set changed (device flags set by the level
    // update of the previous step)
    std::vector<uint32_t> need(L, 1u);
    if (L > 1) {
        HIP_TRY(hipMemcpyAsync(need.data(), s->d_need, L * 4, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    for (int l = 1; l < L; l++)
        if (need[l]) drop_list(s, l);
    for (int lvl : order) {
        if (lvl > 0 && !s->list_ok[lvl]) {
            hipEvent_t b0 = s->timing ? mark() : nullptr;
            TRY(make_list(s, lvl));
            if (s->timing) bu.push_back({b0, mark()});
            s->builds_last++;
        }
        // timing: one event pair around the sweep's whole replay (per-launch
        // events would add their own packets between the launches); the
        // average launch duration is the total over the replay's launches
        hipEvent_t t0 = s->timing ? mark() : nullptr;
        if (s->graphs_on && !s->graphs[lvl] && s->replays[lvl] >= 1)
            capture(s, g, lvl);
        if (s->graphs[lvl]) HIP_TRY(hipGraphLaunch(s->graphs[lvl], s->stream));
        else launch_replay(s, g, lvl, nullptr);
        if (s->timing) {
            tr.push_back({t0, mark()});
            s->n_transport += (int64_t)s->plan[lvl].size();
        }
        HIP_TRY(hipGetLastError());
        s->replays[lvl]++;
        hipEvent_t a = s->timing ? mark() : nullptr;
        k_periodic<R><<<gper, kBlock, 0, s->stream>>>(g, fl);
        k_chemistry<R><<<gcells, kBlock, 0, s->stream>>>(g, fl, ca, lvl);
        if (s->timing) ch.push_back({a, mark()});
        s->sweeps_last++;
    }
    HIP_TRY(hipGetLastError());
    double el = s->ts.current_max_timestep();
    s->ts.advance();
    k_levels<<<gcells, kBlock, 0, s->stream>>>(g, s->d_change, s->d_level, s->p.max_timestep,
                                               s->p.timestep_safety_factor, s->ts.lowest);
    HIP_TRY(hipGetLastError());
    if (s->timing) {
        HIP_TRY(hipStreamSynchronize(s->stream));
        for (auto& pr : tr) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, pr.first, pr.second);
            s->t_transport_total += ms;
        }
        for (auto& pr : ch) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, pr.first, pr.second);
            s->t_chem_total += ms;
        }
        for (auto& pr : bu) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, pr.first, pr.second);
            s->t_build_total += ms;
        }
    }
    if (elapsed) *elapsed = el;
    return SSW_OK;
}

// ---------------------------------------------------------------------------
// Packed steps: one step of each of several handles, their grouped launches
// packed into shared launches (k_sweep_group_pack).
//
// The handles' sweeps are aligned by index: sweep i of every handle whose
// step has more than i sweeps runs together. Inside sweep i, the t-th items of
// the handles' replay plans are launched together: grouped items of handles
// with the same transport precision and group block go into one packed launch
// (up to kMaxPack per launch), other items are launched alone. Each handle's
// items still run in its plan's order (everything is on one stream), and a
// packed workgroup computes exactly what it computes in a launch of its own,
// so every handle's results are bitwise those of stepping it alone.

struct BatchSweep {
    ssw_sweep* s;
    std::vector<int> order;
    Geo g;
};

template <typename R>
GroupJob<R> group_job(ssw_sweep* s, int lvl, const ssw_sweep::ReplayItem& it, uint32_t* nblk) {
    const bool all = lvl == 0 || s->list_full[lvl];
    const ssw_sweep::Group& gr = s->groups[it.tab];
    const Flux<R> fl = make_flux<R>(s);
    const Geo g = make_geo(s);
    GroupJob<R> J;
    J.N = g.N;
    J.Nper = g.Nper;
    J.X = fl.X;
    J.I = fl.I;
    J.P = fl.P;
    J.offI = fl.offI;
    J.sa = fl.sa;
    J.thr = fl.thr;
    J.hdr = s->d_hdr;
    J.A = (const R*)s->d_A;
    J.planes = (const Entry<R>*)s->d_planes;
    J.ovf = (const Entry<R>*)s->d_ovf;
    J.fstart = s->d_fstart;
    J.desc = (all ? s->d_gdesc : s->d_gdesc_l[lvl]) + 2 * (size_t)gr.desc0;
    J.list = all ? s->d_glist : s->d_glist_l[lvl];
    J.list_len = s->glist_len;
    J.k0 = gr.k0;
    J.phases = gr.levels;
    J.start = 0;
    *nblk = all ? gr.ndesc : s->gblocks[lvl][it.tab];
    J.blocks = *nblk;
    J.pt = all ? gr.passes : s->gpass_l[lvl][it.tab];
    return J;
}

// A packed launch under construction: jobs of one precision and group block.
template <typename R>
struct Pack {
    PackArgs<R> a{};  // value-initialised: no jobs yet
    uint32_t grid = 0;
};

template <typename R>
void flush_pack(Pack<R>& p, int roots, hipStream_t st) {
    if (p.a.n <= 0) return;
    bool four = true;
    for (int k = 0; k < p.a.n; k++) four = four && p.a.job[k].phases == 4;
    if (roots == 32) {
        if (four) k_sweep_group_pack<R, 64, 4><<<p.grid, 64, 0, st>>>(p.a);
        else k_sweep_group_pack<R, 64, 0><<<p.grid, 64, 0, st>>>(p.a);
    } else if (roots == 64) {
        if (four) k_sweep_group_pack<R, 128, 4><<<p.grid, 128, 0, st>>>(p.a);
        else k_sweep_group_pack<R, 128, 0><<<p.grid, 128, 0, st>>>(p.a);
    } else {
        if (four) k_sweep_group_pack<R, 256, 4><<<p.grid, 256, 0, st>>>(p.a);
        else k_sweep_group_pack<R, 256, 0><<<p.grid, 256, 0, st>>>(p.a);
    }
    p.a.n = 0;
    p.grid = 0;
}

template <typename R>
void add_job(Pack<R>& p, const GroupJob<R>& J, int roots, hipStream_t st) {
    if (p.a.n >= kMaxPack) flush_pack<R>(p, roots, st);
    GroupJob<R> j = J;
    j.start = p.grid;
    p.grid += (J.blocks + kXcds - 1) / kXcds * kXcds;
    p.a.job[p.a.n++] = j;
}

// Packs keyed by (precision, group block).
struct Packs {
    Pack<float> f[3];
    Pack<double> d[3];
    static int slot(int roots) { return roots == 32 ? 0 : roots == 64 ? 1 : 2; }
    void flush(hipStream_t st) {
        const int roots[3] = {32, 64, 128};
        for (int k = 0; k < 3; k++) {
            flush_pack<float>(f[k], roots[k], st);
            flush_pack<double>(d[k], roots[k], st);
        }
    }
};

template <typename R>
void packed_item(ssw_sweep* s, const Geo& g, int lvl, const ssw_sweep::ReplayItem& it, Pack<R>& pk, hipStream_t st) {
    if (it.group == 1) {
        uint32_t nblk = 0;
        const GroupJob<R> J = group_job<R>(s, lvl, it, &nblk);
        if (nblk) add_job<R>(pk, J, s->group_block, st);
        return;
    }
    launch_item_t<R>(s, g, make_flux<R>(s), lvl, it, st);
}

template <typename R>
void batch_chemistry(ssw_sweep* s, const Geo& g, int lvl, hipStream_t st) {
    const Flux<R> fl = make_flux<R>(s);
    k_periodic<R><<<grid_for(std::max<int64_t>(s->Nper, 1) * s->D), kBlock, 0, st>>>(g, fl);
    k_chemistry<R><<<grid_for(s->N), kBlock, 0, st>>>(g, fl, chem_args(s), lvl);
}

// Graphs of packed sweeps, keyed by (uid, epoch, sweep level) of every
// handle in the sweep: a key seen once is captured the second time it comes
// and replayed from then on, as a handle's own replays are. The epoch of a
// handle changes whenever one of its replay plans is rebuilt, so a graph is
// only replayed with the plans (and buffers) it was captured from; graphs of
// destroyed handles are never matched again (uids are not reused) and leave
// the cache by age.
struct PackedGraph {
    std::vector<uint64_t> key;
    int device;
    hipGraphExec_t ge;  // null: seen once, not captured yet
    uint64_t last_use;
};
std::mutex g_pack_mu;
std::vector<PackedGraph> g_pack_graphs;
uint64_t g_pack_clock = 0;
int64_t g_pack_captured = 0, g_pack_replayed = 0;
constexpr size_t kPackGraphs = 256;

// Under g_pack_mu: the entry of key on device dev, or nullptr.
PackedGraph* find_packed(const std::vector<uint64_t>& key, int dev) {
    for (PackedGraph& e : g_pack_graphs)
        if (e.device == dev && e.key == key) {
            e.last_use = ++g_pack_clock;
            return &e;
        }
    return nullptr;
}

// Under g_pack_mu: adds an entry, evicting the least recently used one when full.
void add_packed(const std::vector<uint64_t>& key, int dev, hipGraphExec_t ge) {
    if (g_pack_graphs.size() >= kPackGraphs) {
        size_t old = 0;
        for (size_t k = 1; k < g_pack_graphs.size(); k++)
            if (g_pack_graphs[k].last_use < g_pack_graphs[old].last_use) old = k;
        if (g_pack_graphs[old].ge) (void)hipGraphExecDestroy(g_pack_graphs[old].ge);
        g_pack_graphs.erase(g_pack_graphs.begin() + (std::ptrdiff_t)old);
    }
    g_pack_graphs.push_back(PackedGraph{key, dev, ge, ++g_pack_clock});
}

// Captures what `enqueue` launches on st as one graph (nothing runs). Returns
// null if the capture failed; the caller then launches the work directly.
hipGraphExec_t capture_packed(hipStream_t st, const std::function<void()>& enqueue) {
    hipGraph_t graph = nullptr;
    std::unique_lock<std::shared_mutex> lock(g_capture_mu);
    if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    enqueue();
    hipError_t ce = hipStreamEndCapture(st, &graph);
    lock.unlock();
    if (ce != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return nullptr;
    }
    hipGraphExec_t ge = nullptr;
    hipError_t ie = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return ge;
}

// One packed step of the prepared handles bs[0..n) on the stream of the
// first (the others' streams wait for it before and it waits for theirs).
int step_pack(BatchSweep* bs, int n, double* elapsed) {
    const hipStream_t st = bs[0].s->stream;
    for (int h = 1; h < n; h++) {
        ssw_sweep* s = bs[h].s;
        if (!s->sync_ev) HIP_TRY(hipEventCreateWithFlags(&s->sync_ev, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(s->sync_ev, s->stream));
        HIP_TRY(hipStreamWaitEvent(st, s->sync_ev, 0));
    }
    for (int h = 0; h < n; h++) {
        k_fill_u32<<<1, 2, 0, st>>>((uint32_t*)bs[h].s->d_active, 0u, 2);
        k_fill_u32<<<1, 6, 0, st>>>((uint32_t*)bs[h].s->d_limit, 0u, 6);
    }
    size_t sweeps = 0;
    for (int h = 0; h < n; h++) sweeps = std::max(sweeps, bs[h].order.size());
    bool graphs = true;
    for (int h = 0; h < n; h++) graphs = graphs && bs[h].s->graphs_on;
    const int dev = bs[0].s->device;
    for (size_t i = 0; i < sweeps; i++) {
        // sweep i of every handle: the t-th plan items together, then each
        // handle's periodic sources and chemistry
        const std::function<void()> enqueue = [&]() {
            Packs pk;
            size_t items = 0;
            for (int h = 0; h < n; h++)
                if (i < bs[h].order.size()) items = std::max(items, bs[h].s->plan[bs[h].order[i]].size());
            for (size_t t = 0; t < items; t++) {
                for (int h = 0; h < n; h++) {
                    BatchSweep& b = bs[h];
                    if (i >= b.order.size()) continue;
                    const int lvl = b.order[i];
                    if (t >= b.s->plan[lvl].size()) continue;
                    const ssw_sweep::ReplayItem& it = b.s->plan[lvl][t];
                    const int k = Packs::slot(b.s->group_block);
                    if (b.s->fp32) packed_item<float>(b.s, b.g, lvl, it, pk.f[k], st);
                    else packed_item<double>(b.s, b.g, lvl, it, pk.d[k], st);
                }
                pk.flush(st);
            }
            for (int h = 0; h < n; h++) {
                BatchSweep& b = bs[h];
                if (i >= b.order.size()) continue;
                if (b.s->fp32) batch_chemistry<float>(b.s, b.g, b.order[i], st);
                else batch_chemistry<double>(b.s, b.g, b.order[i], st);
            }
        };
        bool done = false;
        if (graphs) {
            std::vector<uint64_t> key;
            for (int h = 0; h < n; h++)
                if (i < bs[h].order.size())
                    key.insert(key.end(), {bs[h].s->uid, bs[h].s->epoch, (uint64_t)bs[h].order[i]});
            std::lock_guard<std::mutex> lk(g_pack_mu);
            PackedGraph* e = find_packed(key, dev);
            if (e && !e->ge) {  // the second time: capture
                e->ge = capture_packed(st, enqueue);
                if (e->ge) g_pack_captured++;
            }
            if (e && e->ge) {
                HIP_TRY(hipGraphLaunch(e->ge, st));
                g_pack_replayed++;
                done = true;
            } else if (!e) {
                add_packed(key, dev, nullptr);
            }
        }
        if (!done) enqueue();
        for (int h = 0; h < n; h++)
            if (i < bs[h].order.size()) bs[h].s->sweeps_last++;
        HIP_TRY(hipGetLastError());
    }
    for (int h = 0; h < n; h++) {
        ssw_sweep* s = bs[h].s;
        const double el = s->ts.current_max_timestep();
        s->ts.advance();
        k_levels<<<grid_for(s->N), kBlock, 0, st>>>(bs[h].g, s->d_change, s->d_level, s->p.max_timestep,
                                                   s->p.timestep_safety_factor, s->ts.lowest);
        if (elapsed) elapsed[h] = el;
    }
    HIP_TRY(hipGetLastError());
    // the handles' own streams continue after the packed step
    ssw_sweep* s0 = bs[0].s;
    if (!s0->sync_ev) HIP_TRY(hipEventCreateWithFlags(&s0->sync_ev, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(s0->sync_ev, st));
    for (int h = 1; h < n; h++) HIP_TRY(hipStreamWaitEvent(bs[h].s->stream, s0->sync_ev, 0));
    return SSW_OK;
}

// Packs of about kPackSize handles, on up to kPackStreams streams: the device
// runs the launches of a few streams at once (one hardware queue per stream,
// dealt to streams in turn as they are created), so a few packs side by side
// beat one pack of all of them. Handle k goes to pack uid(k) mod npack
// (creation turn; evened out where the turns are uneven), so that the packs'
// streams sit on different queues and every pack holds a mix of problem
// sizes; within a pack the longest plan leads. Measured on 16 small problems
// (8000-27000 cells, 16-84 directions): 3.46 ms per step against 4.28 ms for
// packs of similar sizes and 5.34 ms for one stream per problem.
constexpr int kPackSize = 4;
constexpr int kPackStreams = 4;

int step_batch(ssw_sweep* const* hs, int n, double* elapsed) {
    std::vector<BatchSweep> bs((size_t)n);
    // per handle, on its own stream: the sweep levels whose active set
    // changed, and their active-set lists (the lists depend only on the
    // timestep levels, which change at the end of a step)
    for (int h = 0; h < n; h++) {
        ssw_sweep* s = hs[h];
        BatchSweep& b = bs[h];
        b.s = s;
        b.order = s->ts.sweep_order();
        b.g = make_geo(s);
        const int L = s->p.num_timestep_levels;
        s->sweeps_last = 0;
        s->builds_last = 0;
        s->t_transport_total = s->t_chem_total = s->t_build_total = 0;
        s->n_transport = 0;
        std::vector<uint32_t> need(L, 1u);
        if (L > 1) {
            HIP_TRY(hipMemcpyAsync(need.data(), s->d_need, L * 4, hipMemcpyDeviceToHost, s->stream));
            HIP_TRY(hipStreamSynchronize(s->stream));
        }
        for (int l = 1; l < L; l++)
            if (need[l]) drop_list(s, l);
        for (int lvl : b.order)
            if (lvl > 0 && !s->list_ok[lvl]) {
                TRY(make_list(s, lvl));
                s->builds_last++;
            }
    }
    // handles by their full sweep's plan length (a proxy of their per-step
    // chain), longest first, dealt to the packs by creation turn
    std::vector<int> order((size_t)n);
    for (int h = 0; h < n; h++) order[h] = h;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return bs[x].s->plan[0].size() > bs[y].s->plan[0].size(); });
    const int npack = std::min(kPackStreams, (n + kPackSize - 1) / kPackSize);
    std::vector<std::vector<int>> packs((size_t)npack);
    for (int r = 0; r < n; r++) packs[bs[order[r]].s->uid % (uint64_t)npack].push_back(order[r]);
    for (;;) {  // even out: no pack more than one above another
        auto big = std::max_element(packs.begin(), packs.end(),
                                    [](const std::vector<int>& x, const std::vector<int>& y) { return x.size() < y.size(); });
        auto small = std::min_element(packs.begin(), packs.end(),
                                      [](const std::vector<int>& x, const std::vector<int>& y) { return x.size() < y.size(); });
        if (big->size() <= small->size() + 1) break;
        small->push_back(big->back());
        big->pop_back();
    }
    // the lead (whose stream a pack runs on) of each pack: its longest member,
    // unless an evened-out pack needs another member to sit at a place in the
    // turn of the hardware queues (creation order modulo the queue count) that
    // no earlier pack's lead holds
    uint32_t used = 0;
    for (size_t k = 0; k < packs.size(); k++) {
        std::vector<int>& m = packs[k];
        for (size_t j = 0; j < m.size(); j++) {
            const uint32_t q = 1u << (bs[m[j]].s->uid % (uint64_t)kPackStreams);
            if (!(used & q)) {
                std::swap(m[0], m[j]);
                break;
            }
        }
        used |= 1u << (bs[m[0]].s->uid % (uint64_t)kPackStreams);
    }
    // each pack from a thread of its own (the first on the caller's): issuing
    // the launches of a step takes host time comparable to running them
    const int np = (int)packs.size();
    std::vector<int> rc((size_t)np, SSW_OK);
    std::vector<std::string> err((size_t)np);
    auto run = [&](int k) {
        std::vector<BatchSweep> sub;
        for (int h : packs[k]) sub.push_back(bs[h]);
        std::vector<double> el(sub.size(), 0.0);
        rc[k] = step_pack(sub.data(), (int)sub.size(), el.data());
        if (rc[k] != SSW_OK) err[k] = g_err;
        if (elapsed)
            for (size_t j = 0; j < sub.size(); j++) elapsed[packs[k][j]] = el[j];
    };
    std::vector<std::thread> th;
    int spawned = 1;
    try {
        for (; spawned < np; spawned++) th.emplace_back(run, spawned);
    } catch (const std::exception&) {  // no thread to spare: the rest on this one
    }
    run(0);
    for (int k = spawned; k < np; k++) run(k);
    for (std::thread& t : th) t.join();
    for (int k = 0; k < np; k++)
        if (rc[k] != SSW_OK) return fail(rc[k], err[k]);
    return SSW_OK;
}

void free_state(ssw_sweep* s) {
    dfree(s, s->d_rho);
    dfree(s, s->d_x);
    dfree(s, s->d_T);
    dfree(s, s->d_src);
    dfree(s, s->d_sa);
    dfree(s, s->d_sts);
    dfree(s, s->d_prev);
    dfree(s, s->d_change);
    dfree(s, s->d_X);
    dfree(s, s->d_I);
    dfree(s, s->d_P);
    dfree(s, s->d_offI);
    dfree(s, s->d_offP);
}

void free_grid(ssw_sweep* s) {
    drop_all_lists(s);
    for (uint32_t*& p : s->d_idx) dfree(s, p);
    for (uint4*& p : s->d_runtab) dfree(s, p);
    dfree(s, s->d_gdesc);
    dfree(s, s->d_glist);
    dfree(s, s->d_gmap);
    dfree(s, s->d_gcount);
    dfree(s, s->d_gbase);
    dfree(s, s->d_gblocks);
    dfree(s, s->d_pbound);
    dfree(s, s->d_gf_desc);
    dfree(s, s->d_gf_kept);
    dfree(s, s->d_gf_pos);
    dfree(s, s->d_gf_pbval);
    dfree(s, s->d_gf_scan);
    s->gf_scan_bytes = 0;
    for (uint4*& p : s->d_gdesc_l) dfree(s, p);
    for (uint32_t*& p : s->d_glist_l) dfree(s, p);
    s->groups.clear();
    dfree(s, s->d_off);
    dfree(s, s->d_inv);
    dfree(s, s->d_geo);
    dfree(s, s->d_nk);
    dfree(s, s->d_vol);
    dfree(s, s->d_size);
    dfree(s, s->d_pslot);
    dfree(s, s->d_pcells);
    dfree(s, s->d_pfoff);
    dfree(s, s->d_pfidx);
    dfree(s, s->d_flat);
    dfree(s, s->d_lstart);
    dfree(s, s->d_ctr);
    dfree(s, s->d_cnt);
    dfree(s, s->d_level);
    dfree(s, s->d_fstart);
    dfree(s, s->d_dirrun);
    dfree(s, s->d_dirlev);
    dfree(s, s->d_recs);
    dfree(s, s->d_hdr);
    dfree(s, s->d_A);
    dfree(s, s->d_planes);
    dfree(s, s->d_ovf_start);
    dfree(s, s->d_ovf);
    dfree(s, s->d_xpos);
    dfree(s, s->d_chunks);
    dfree(s, s->d_first_chunk);
    dfree(s, s->d_ccount);
    dfree(s, s->d_coff);
    dfree(s, s->d_cstart);
    dfree(s, s->d_scan_tmp);
}

// Overflow entries of the compiled schedule (once per grid).
int setup_overflow(ssw_sweep* s, uint32_t* d_ovf_count) {
    const size_t DN = (size_t)s->D * s->N;
    TRY(dalloc(s, &s->d_ovf_start, DN));
    TRY(dalloc(s, (uint8_t**)&s->d_ovf, (size_t)std::max<int64_t>(1, s->ovf_total) * entry_size(s)));
    uint32_t* d_cursor = nullptr;
    Defer cleanup{[&] { dfree(s, d_cursor); }};
    TRY(dalloc(s, &d_cursor, 1));
    HIP_TRY(hipMemsetAsync(d_cursor, 0, 4, s->stream));
    Geo g = make_geo(s);
    k_overflow<<<grid_for((int64_t)DN), kBlock, 0, s->stream>>>(g, d_ovf_count, s->d_ovf_start, d_cursor);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s, d_cursor);
    return SSW_OK;
}

// The device address of `scratch` on device `dev`, looked up once per device
// with that device current (each device holds its own copy of the array at
// its own address) and kept for the life of the process.
std::mutex g_scratch_mu;
std::array<void*, kMaxDevices> g_scratch_ptr{};

int noalloc_scratch(int dev, uint8_t** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(SSW_ERR_INVALID, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = -1;
        HIP_TRY(hipGetDevice(&cur));
        if (cur != dev) HIP_TRY(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) HIP_TRY(hipSetDevice(cur));
        if (e != hipSuccess || !p)
            return fail(SSW_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = (uint8_t*)g_scratch_ptr[dev];
    return SSW_OK;
}

// One per-cell output of ssw_download_noalloc: into the caller's order through
// the staging buffer when the grid is permuted, else copied directly.
template <typename T>
int noalloc_out(ssw_sweep* s, const T* src, T* dst, uint8_t* stage) {
    const int64_t n = s->N;
    if (s->permuted) {
        k_to_caller<T><<<grid_for(n), kBlock, 0, s->stream>>>(src, s->d_inv, (T*)stage, n);
        HIP_TRY(hipGetLastError());
        src = (const T*)stage;
    }
    // the next output reuses the staging buffer: stream order puts its kernel
    // after this copy
    HIP_TRY(hipMemcpyAsync(dst, src, (size_t)n * sizeof(T), hipMemcpyDeviceToHost, s->stream));
    return SSW_OK;
}

// The per-direction dependency graph of the uploaded faces under the current
// directions: mirror check (first build), downwind statistics, overflow
// counts, the transport mode (choose_mode) and the schedule. d_cnt is held.
int build_direction_graph(ssw_sweep* s, bool choose_mode) {
    const int64_t n = s->N;
    const int L = s->p.num_timestep_levels;
    const size_t DN = (size_t)s->D * n;
    unsigned long long* d_stats = nullptr;
    Defer cleanup{[&] { dfree(s, d_stats); }};
    TRY(dalloc(s, &d_stats, 5));
    HIP_TRY(hipMemsetAsync(d_stats, 0, 5 * 8, s->stream));
    Geo g = make_geo(s);
    if (choose_mode) k_check_mirror<<<grid_for(n), kBlock, 0, s->stream>>>(g, d_stats);
    k_face_stats<<<grid_for((int64_t)DN), kBlock, 0, s->stream>>>(g, d_stats + 1, s->d_cnt);
    HIP_TRY(hipGetLastError());
    unsigned long long hb[5] = {0, 0, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(hb, d_stats, sizeof(hb), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s, d_stats);
    if (hb[0] != 0)
        return fail(SSW_ERR_GRID, std::to_string(hb[0]) +
                                      " local faces have no mirror face with the exactly negated normal");
    for (int k = 0; k < 3; k++) s->dw[k] = (double)hb[k + 1] / ((double)n * s->D);
    s->ovf_total = (int64_t)hb[4];
    if (!choose_mode) {
        if (s->compiled) TRY(setup_overflow(s, s->d_cnt));
        return build_schedule(s);
    }
    // transport mode: compiled when, at its peak (schedule arrays, overflow
    // starts and sorted order while the schedule is written; then the
    // active-set lists and the transport state), it needs at most 85% of the
    // free device memory (automatic), else geometric
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const size_t rs = real_size(s);
    const size_t comp_bytes = DN * (16 + rs + kPlanes * entry_size(s) + 4) + (size_t)s->ovf_total * entry_size(s);
    const bool grouped = s->group_request >= 2;
    const size_t lists = grouped ? (size_t)group_bytes((int64_t)DN, n, L, s->group_block) + DN * 5
                                 : DN * 4 * (size_t)(L - 1);  // grouped: + build scratch
    const size_t comp_need = comp_bytes + std::max<size_t>(DN * 8, lists + DN * 2 * rs);
    if (s->mode_request == SSW_TRANSPORT_MODE_COMPILED) s->compiled = true;
    else if (s->mode_request == SSW_TRANSPORT_MODE_GEOMETRIC) s->compiled = false;
    else s->compiled = (double)comp_need <= 0.85 * (double)free_b;
    if (s->compiled) TRY(setup_overflow(s, s->d_cnt));
    return build_schedule(s);
}

// The row-major matrix times every direction (multiply_by_matrix,
// src/sweep/direction/mod.rs:150-155), and kept[i] = kernel[i][D-1] of remap
// (:177-205): 1 if the last rotated direction is the last of the rotated
// directions most aligned with old direction i (Iterator::max_by_key keeps
// the last maximum of 1 + old_i . new_j), else 0.
void rotate_table(const double* m, std::vector<double>& dirs, std::vector<double>& kept) {
    const int D = (int)dirs.size() / 3;
    const std::vector<double> old = dirs;
    for (int d = 0; d < D; d++) {
        const double x = old[3 * d], y = old[3 * d + 1], z = old[3 * d + 2];
        dirs[3 * d] = x * m[0] + y * m[1] + z * m[2];
        dirs[3 * d + 1] = x * m[3] + y * m[4] + z * m[5];
        dirs[3 * d + 2] = x * m[6] + y * m[7] + z * m[8];
    }
    kept.assign(D, 0.0);
    for (int i = 0; i < D; i++) {
        int best = 0;
        double best_v = 0.0;
        for (int j = 0; j < D; j++) {
            const double v =
                1.0 + ((old[3 * i] * dirs[3 * j] + old[3 * i + 1] * dirs[3 * j + 1]) + old[3 * i + 2] * dirs[3 * j + 2]);
            if (j == 0 || v >= best_v) {
                best = j;
                best_v = v;
            }
        }
        kept[i] = best == D - 1 ? 1.0 : 0.0;
    }
}

template <typename R>
int rotate_state_t(ssw_sweep* s, const double* d_kept, const double* dI, const double* dO, const double* dP) {
    const int64_t DN = (int64_t)s->D * s->N;
    const size_t rs = sizeof(R);
    if (!s->d_offI) TRY(dalloc(s, (uint8_t**)&s->d_offI, (size_t)DN * rs));
    if (!s->d_offP) TRY(dalloc(s, (uint8_t**)&s->d_offP, (size_t)s->D * std::max<int64_t>(1, s->Nper) * rs));
    const Geo g = make_geo(s);
    const Flux<R> fl = make_flux<R>(s);
    k_rotate_state<R><<<grid_for(DN), kBlock, 0, s->stream>>>(g, fl, d_kept, dI, dO, dP);
    k_rotate_offsets<R><<<grid_for(DN), kBlock, 0, s->stream>>>(g, fl);
    HIP_TRY(hipGetLastError());
    return SSW_OK;
}

}  // namespace

extern "C" {

const char* ssw_last_error(void) { return g_err.c_str(); }

void ssw_default_parameters(ssw_sweep_parameters* p) {
    std::memset(p, 0, sizeof(*p));
    p->num_directions = 84;
    p->explicit_directions = nullptr;
    p->num_timestep_levels = 1;
    p->periodic = 0;
    p->max_timestep = 1e6 * 3.15576e7;
    p->significant_rate_threshold = 0.0;
    p->timestep_safety_factor = 0.01 * 10.0;  // percent(10.0)
    p->chemistry_timestep_safety_factor = 0.01 * 10.0;
    p->prevent_cooling = 1;
    p->transport_precision = SSW_TRANSPORT_FP64;
    p->scale_factor = 1.0;
}

int ssw_create(const ssw_sweep_parameters* params, ssw_sweep** out) {
    if (!params || !out) return fail(SSW_ERR_INVALID, "null argument");
    *out = nullptr;
    const ssw_sweep_parameters& p = *params;
    if (p.num_timestep_levels < 1 || p.num_timestep_levels >= 32)
        return fail(SSW_ERR_INVALID, "num_timestep_levels must be in [1, 31]");
    if (p.transport_precision != SSW_TRANSPORT_FP64 && p.transport_precision != SSW_TRANSPORT_FP32)
        return fail(SSW_ERR_INVALID, "transport_precision must be 32 or 64");
    if (!(p.max_timestep > 0.0)) return fail(SSW_ERR_INVALID, "max_timestep must be positive");
    std::vector<double> dirs;
    if (p.explicit_directions) {
        if (p.num_directions < 1 || p.num_directions > 127)
            return fail(SSW_ERR_INVALID, "explicit direction count must be in [1, 127]");
        for (int d = 0; d < p.num_directions; d++) {
            const double* v = p.explicit_directions + 3 * d;
            double len = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
            if (!(len > 0.0)) return fail(SSW_ERR_INVALID, "zero explicit direction");
            double r = 1.0 / len;  // glam normalize: v * (1/|v|)
            dirs.push_back(v[0] * r);
            dirs.push_back(v[1] * r);
            dirs.push_back(v[2] * r);
        }
    } else {
        const double(*bins)[3] = nullptr;
        static const double one[1][3] = {{1.0, 0.0, 0.0}};
        switch (p.num_directions) {  // Directions::from_num (src/sweep/direction/mod.rs:58-75)
            case 1: bins = one; break;
            case 16: bins = SSW_DIRECTION_BINS_16; break;
            case 21: bins = SSW_DIRECTION_BINS_21; break;
            case 32: bins = SSW_DIRECTION_BINS_32; break;
            case 64: bins = SSW_DIRECTION_BINS_64; break;
            case 84: bins = SSW_DIRECTION_BINS_84; break;
            default: return fail(SSW_ERR_INVALID, "num_directions must be 1, 16, 21, 32, 64 or 84");
        }
        for (int d = 0; d < p.num_directions; d++)
            for (int k = 0; k < 3; k++) dirs.push_back(bins[d][k]);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(SSW_ERR_NO_DEVICE, "no HIP device");
    ssw_sweep* s = new ssw_sweep();
    s->uid = g_next_uid.fetch_add(1) + 1;
    g_live_handles++;
    s->p = p;
    s->p.explicit_directions = nullptr;
    s->D = p.num_directions;
    s->dirs = dirs;
    s->fp32 = p.transport_precision == SSW_TRANSPORT_FP32;
    (void)hipGetDevice(&s->device);
    hipError_t se;
    {
        std::shared_lock<std::shared_mutex> lock(g_capture_mu);
        se = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    }
    if (se != hipSuccess) {
        delete s;
        g_live_handles--;
        return fail(SSW_ERR_HIP, "hipStreamCreate failed");
    }
    int r = dalloc(s, &s->d_dirs, dirs.size());
    if (r == SSW_OK) r = dalloc(s, &s->d_need, (size_t)p.num_timestep_levels);
    if (r == SSW_OK) r = dalloc(s, &s->d_fail, 1);
    if (r == SSW_OK) r = dalloc(s, &s->d_active, 1);
    if (r == SSW_OK) r = dalloc(s, &s->d_limit, 3);
    if (r != SSW_OK) {
        ssw_destroy(s);
        return r;
    }
    if (hipMemcpyAsync(s->d_dirs, dirs.data(), dirs.size() * 8, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemsetAsync(s->d_fail, 0, 8, s->stream) != hipSuccess ||
        hipMemsetAsync(s->d_active, 0, 8, s->stream) != hipSuccess ||
        hipMemsetAsync(s->d_limit, 0, 24, s->stream) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess) {
        ssw_destroy(s);
        return fail(SSW_ERR_HIP, "initialisation of the handle's device memory failed");
    }
    s->ts.init(p.max_timestep, p.num_timestep_levels);
    *out = s;
    return SSW_OK;
}

void ssw_destroy(ssw_sweep* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_state(s);
    free_grid(s);
    dfree(s, s->d_dirs);
    dfree(s, s->d_need);
    dfree(s, s->d_fail);
    dfree(s, s->d_active);
    dfree(s, s->d_limit);
    for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
    if (s->sync_ev) (void)hipEventDestroy(s->sync_ev);
    if (s->stream) {
        std::shared_lock<std::shared_mutex> lock(g_capture_mu);
        (void)hipStreamDestroy(s->stream);
    }
    delete s;
    g_live_handles--;
}

int ssw_set_transport_mode(ssw_sweep* s, int32_t mode) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (mode != SSW_TRANSPORT_MODE_AUTO && mode != SSW_TRANSPORT_MODE_GEOMETRIC && mode != SSW_TRANSPORT_MODE_COMPILED)
        return fail(SSW_ERR_INVALID, "transport mode " + std::to_string(mode) +
                                         " is not available (0 automatic, 3 geometric, 4 compiled)");
    if (s->have_grid) return fail(SSW_ERR_INVALID, "the transport mode must be set before ssw_upload_grid");
    s->mode_request = mode;
    return SSW_OK;
}

int ssw_set_replay_block(ssw_sweep* s, int32_t block) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (block != 0 && block != 64 && block != 128 && block != 256)
        return fail(SSW_ERR_INVALID, "replay block size must be 0, 64, 128 or 256");
    s->rec_block = block ? block : 64;
    s->epoch++;
    for (hipGraphExec_t& ge : s->graphs)  // graphs hold the old launch shape
        if (ge) {
            (void)hipGraphExecDestroy(ge);
            ge = nullptr;
        }
    std::fill(s->replays.begin(), s->replays.end(), 0);
    return SSW_OK;
}

int ssw_set_cell_positions(ssw_sweep* s, int64_t n, const double* positions) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (s->have_grid) return fail(SSW_ERR_INVALID, "cell positions must be set before ssw_upload_grid");
    if (!positions || n <= 0) {
        s->pos.clear();
        return SSW_OK;
    }
    s->pos.assign(positions, positions + 3 * n);
    return SSW_OK;
}

int ssw_set_group_levels(ssw_sweep* s, int32_t levels) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (levels < 1 || levels > kMaxGroup) return fail(SSW_ERR_INVALID, "group levels must be 1 to 8");
    if (s->have_grid) return fail(SSW_ERR_INVALID, "ssw_set_group_levels must be called before ssw_upload_grid");
    s->group_request = levels;
    return SSW_OK;
}

int ssw_set_direction_sweeps(ssw_sweep* s, int32_t enabled) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (s->have_grid) return fail(SSW_ERR_INVALID, "ssw_set_direction_sweeps must be called before ssw_upload_grid");
    s->dir_sweeps = enabled != 0;
    return SSW_OK;
}

int ssw_direction_table(int32_t num_directions, double* out) {
    if (!out) return fail(SSW_ERR_INVALID, "null argument");
    const double(*bins)[3] = nullptr;
    static const double one[1][3] = {{1.0, 0.0, 0.0}};
    switch (num_directions) {  // Directions::from_num (src/sweep/direction/mod.rs:58-75)
        case 1: bins = one; break;
        case 16: bins = SSW_DIRECTION_BINS_16; break;
        case 21: bins = SSW_DIRECTION_BINS_21; break;
        case 32: bins = SSW_DIRECTION_BINS_32; break;
        case 64: bins = SSW_DIRECTION_BINS_64; break;
        case 84: bins = SSW_DIRECTION_BINS_84; break;
        default: return fail(SSW_ERR_INVALID, "num_directions must be 1, 16, 21, 32, 64 or 84");
    }
    for (int d = 0; d < num_directions; d++)
        for (int k = 0; k < 3; k++) out[3 * d + k] = bins[d][k];
    return SSW_OK;
}

int ssw_set_group_block(ssw_sweep* s, int32_t roots) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (roots != 0 && roots != 32 && roots != 64 && roots != 128)
        return fail(SSW_ERR_INVALID, "group block must be 0, 32, 64 or 128");
    if (s->have_grid) return fail(SSW_ERR_INVALID, "ssw_set_group_block must be called before ssw_upload_grid");
    s->group_block = roots ? roots : kGroupT;
    return SSW_OK;
}

int ssw_set_narrow_runs(ssw_sweep* s, int32_t enabled) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (s->have_grid) return fail(SSW_ERR_INVALID, "ssw_set_narrow_runs must be called before ssw_upload_grid");
    s->narrow_runs = enabled != 0;
    return SSW_OK;
}

int ssw_set_graph_replay(ssw_sweep* s, int32_t enabled) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    s->graphs_on = enabled != 0;
    s->epoch++;
    if (!s->graphs_on)
        for (hipGraphExec_t& ge : s->graphs)
            if (ge) {
                (void)hipGraphExecDestroy(ge);
                ge = nullptr;
            }
    return SSW_OK;
}

int ssw_upload_grid(ssw_sweep* s, int64_t n, const int64_t* off, const int32_t* nbr, const uint8_t* kind,
                    const double* area, const double* normal, const double* volume, const double* size) {
    if (!s || !off || !nbr || !kind || !area || !normal || !volume || !size)
        return fail(SSW_ERR_INVALID, "null argument");
    if (n <= 0 || n > (int64_t)kCellMask) return fail(SSW_ERR_INVALID, "cell count must be in [1, 2^25 - 1]");
    if ((int64_t)s->D * n >= (int64_t)0x7FFFFFFFll)
        return fail(SSW_ERR_INVALID, "cells x directions must be below 2^31");
    if (off[0] != 0) return fail(SSW_ERR_INVALID, "face_offsets[0] must be 0");
    for (int64_t c = 0; c < n; c++) {
        if (off[c + 1] < off[c]) return fail(SSW_ERR_INVALID, "face_offsets must be non-decreasing");
        if (off[c + 1] - off[c] > kMaxFaces) return fail(SSW_ERR_INVALID, "a cell has more than 127 faces");
    }
    const int64_t F = off[n];
    if (F >= (int64_t)0xFFFFFFFFll) return fail(SSW_ERR_INVALID, "too many faces");
    free_state(s);
    free_grid(s);
    s->have_grid = s->have_state = false;
    s->N = n;
    s->F = F;
    make_order(s, n);
    // pack faces in the internal cell order, each cell's faces in the caller's
    // order; a non-periodic sweep treats periodic faces as boundaries
    std::vector<double4> geo(F);
    std::vector<uint32_t> nk(F);
    std::vector<uint32_t> off32(n + 1);
    std::vector<uint32_t> pslot(n, kNone), pcells, pfoff(1, 0), pfidx;
    double asum = 0.0;
    int64_t nfaces = 0;
    for (int64_t f = 0; f < F; f++) asum += area[f];  // caller's order: independent of the internal order
    nfaces = F;
    off32[0] = 0;
    for (int64_t i = 0; i < n; i++) {
        const int64_t c = s->inv[i];
        const int64_t fi0 = off32[i];
        off32[i + 1] = (uint32_t)(fi0 + (off[c + 1] - off[c]));
        bool has_p = false;
        for (int64_t f = off[c]; f < off[c + 1]; f++) {
            const int64_t fi = fi0 + (f - off[c]);
            uint32_t k = kind[f];
            if (k > 2) return fail(SSW_ERR_INVALID, "face_kind must be 0, 1 or 2");
            if (k == kKindPeriodic && !s->p.periodic) k = kKindBoundary;
            int64_t nb = (k == kKindBoundary) ? 0 : nbr[f];
            if (k != kKindBoundary && (nb < 0 || nb >= n)) return fail(SSW_ERR_INVALID, "face_neighbour out of range");
            if (k != kKindBoundary) nb = s->perm[nb];
            geo[fi] = make_double4(normal[3 * f], normal[3 * f + 1], normal[3 * f + 2], area[f]);
            nk[fi] = (uint32_t)nb | (k << 30);
            if (k == kKindPeriodic) has_p = true;
        }
        if (has_p) {
            pslot[i] = (uint32_t)pcells.size();
            pcells.push_back((uint32_t)i);
            for (int64_t fi = fi0; fi < (int64_t)off32[i + 1]; fi++)
                if ((nk[fi] >> 30) == kKindPeriodic) pfidx.push_back((uint32_t)fi);
            pfoff.push_back((uint32_t)pfidx.size());
        }
    }
    // Areas are carried in units of a power of two near the mean face area, so
    // that X = O / A keeps the range of O (SI areas reach 1e42 m^2); scaling
    // by a power of two is exact.
    s->aref = (nfaces > 0 && asum > 0.0) ? std::ldexp(1.0, (int)std::lround(std::log2(asum / (double)nfaces))) : 1.0;
    s->Nper = (int64_t)pcells.size();
    s->h_pcells = pcells;
    const int L = s->p.num_timestep_levels;
    const size_t DN = (size_t)s->D * n;
    TRY(dalloc(s, &s->d_off, n + 1));
    TRY(dalloc(s, &s->d_geo, F));
    TRY(dalloc(s, &s->d_nk, F));
    TRY(dalloc(s, &s->d_vol, n));
    TRY(dalloc(s, &s->d_size, n));
    TRY(dalloc(s, &s->d_pslot, n));
    TRY(dalloc(s, &s->d_pcells, std::max<int64_t>(1, s->Nper)));
    TRY(dalloc(s, &s->d_pfoff, pfoff.size()));
    TRY(dalloc(s, &s->d_pfidx, std::max<size_t>(1, pfidx.size())));
    TRY(dalloc(s, &s->d_level, n));
    TRY(dalloc(s, &s->d_ctr, 3));
    TRY(dalloc(s, &s->d_cnt, DN));
    if (s->permuted) {
        TRY(dalloc(s, &s->d_inv, n));
        HIP_TRY(hipMemcpyAsync(s->d_inv, s->inv.data(), n * 4, hipMemcpyHostToDevice, s->stream));
    }
    HIP_TRY(hipMemcpyAsync(s->d_off, off32.data(), (n + 1) * 4, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_geo, geo.data(), F * sizeof(double4), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_nk, nk.data(), F * 4, hipMemcpyHostToDevice, s->stream));
    std::vector<double> vol_i, size_i;
    HIP_TRY(hipMemcpyAsync(s->d_vol, to_internal(s, volume, vol_i), n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_size, to_internal(s, size, size_i), n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_pslot, pslot.data(), n * 4, hipMemcpyHostToDevice, s->stream));
    if (s->Nper) {
        HIP_TRY(hipMemcpyAsync(s->d_pcells, pcells.data(), s->Nper * 4, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->d_pfidx, pfidx.data(), pfidx.size() * 4, hipMemcpyHostToDevice, s->stream));
    }
    HIP_TRY(hipMemcpyAsync(s->d_pfoff, pfoff.data(), pfoff.size() * 4, hipMemcpyHostToDevice, s->stream));
    std::vector<uint8_t> lev(n, (uint8_t)(L - 1));
    HIP_TRY(hipMemcpyAsync(s->d_level, lev.data(), n, hipMemcpyHostToDevice, s->stream));
    TRY(build_direction_graph(s, true));
    s->have_grid = true;
    return SSW_OK;
}

int ssw_upload_state(ssw_sweep* s, const double* density, const double* xhii, const double* temp,
                     const double* source) {
    if (!s || !density || !xhii || !temp || !source) return fail(SSW_ERR_INVALID, "null argument");
    if (!s->have_grid) return fail(SSW_ERR_INVALID, "ssw_upload_grid must be called first");
    drop_all_lists(s);
    free_state(s);
    const int64_t n = s->N;
    const int L = s->p.num_timestep_levels;
    std::vector<double> rho_i, x_i, T_i, src_i;
    density = to_internal(s, density, rho_i);
    xhii = to_internal(s, xhii, x_i);
    temp = to_internal(s, temp, T_i);
    source = to_internal(s, source, src_i);
    std::vector<double> sod(n);
    double lmax = 0.0;
    for (int64_t c = 0; c < n; c++) {
        sod[c] = source[c] / (double)s->D;  // Site::source_per_direction_bin
        lmax = std::max(lmax, std::fabs(sod[c]));
    }
    // FP32 transport carries rates in units of a power of two near the largest
    // source per direction bin (photon rates reach 1e52 s^-1, above FP32 range).
    s->lref = (s->fp32 && lmax > 0.0) ? std::ldexp(1.0, (int)std::lround(std::log2(lmax))) : 1.0;
    const size_t es = real_size(s);
    TRY(dalloc(s, &s->d_rho, n));
    TRY(dalloc(s, &s->d_x, n));
    TRY(dalloc(s, &s->d_T, n));
    TRY(dalloc(s, &s->d_src, n));
    TRY(dalloc(s, (uint8_t**)&s->d_sa, n * sa_size(s)));
    TRY(dalloc(s, &s->d_sts, n));
    TRY(dalloc(s, &s->d_prev, n));
    TRY(dalloc(s, &s->d_change, n));
    TRY(dalloc(s, (uint8_t**)&s->d_X, (size_t)s->D * n * es));
    TRY(dalloc(s, (uint8_t**)&s->d_I, (size_t)s->D * n * es));
    if (s->Nper) TRY(dalloc(s, (uint8_t**)&s->d_P, (size_t)s->D * s->Nper * es));
    HIP_TRY(hipMemcpyAsync(s->d_rho, density, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_x, xhii, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_T, temp, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_src, sod.data(), n * 8, hipMemcpyHostToDevice, s->stream));
    // the transport reads source/D in its storage scale (FP32: divided by the
    // power of two lref, exact); the attenuation follows below
    std::vector<double2> sa(n);
    for (int64_t c = 0; c < n; c++) sa[c] = make_double2(s->fp32 ? sod[c] / s->lref : sod[c], 0.0);
    TRY(upload_sa(s, sa));
    HIP_TRY(hipMemsetAsync(s->d_sts, 0, n * 8, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_prev, 0, n * 8, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_change, 0, n * 8, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_X, 0, (size_t)s->D * n * es, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_I, 0, (size_t)s->D * n * es, s->stream));
    if (s->Nper) HIP_TRY(hipMemsetAsync(s->d_P, 0, (size_t)s->D * s->Nper * es, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_fail, 0, 8, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_active, 0, 8, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_limit, 0, 24, s->stream));
    std::vector<uint8_t> lev(n, (uint8_t)(L - 1));  // src/sweep/mod.rs:206
    HIP_TRY(hipMemcpyAsync(s->d_level, lev.data(), n, hipMemcpyHostToDevice, s->stream));
    if (s->fp32)
        k_attenuation<float><<<grid_for(n), kBlock, 0, s->stream>>>(s->d_rho, s->d_x, s->d_size, (float2*)s->d_sa, n);
    else
        k_attenuation<double><<<grid_for(n), kBlock, 0, s->stream>>>(s->d_rho, s->d_x, s->d_size, (double2*)s->d_sa,
                                                                     n);
    HIP_TRY(hipGetLastError());
    TRY(set_need_all(s, 1u));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->ts.init(s->p.max_timestep, L);
    s->have_state = true;
    return SSW_OK;
}

int ssw_step(ssw_sweep* s, double* elapsed) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (!s->have_state) return fail(SSW_ERR_INVALID, "ssw_upload_state must be called first");
    return s->fp32 ? run_step<float>(s, elapsed) : run_step<double>(s, elapsed);
}

int ssw_packed_graph_stats(int64_t* captured, int64_t* replayed) {
    std::lock_guard<std::mutex> lk(g_pack_mu);
    if (captured) *captured = g_pack_captured;
    if (replayed) *replayed = g_pack_replayed;
    return SSW_OK;
}

int ssw_step_batch(ssw_sweep* const* handles, int32_t n, double* elapsed) {
    if (!handles || n < 1) return fail(SSW_ERR_INVALID, "no handles");
    for (int h = 0; h < n; h++) {
        if (!handles[h]) return fail(SSW_ERR_INVALID, "null handle");
        if (!handles[h]->have_state) return fail(SSW_ERR_INVALID, "ssw_upload_state must be called first");
        if (handles[h]->device != handles[0]->device)
            return fail(SSW_ERR_INVALID, "the handles of a packed step must be on one device");
        for (int k = 0; k < h; k++)
            if (handles[k] == handles[h]) return fail(SSW_ERR_INVALID, "a handle appears twice in a packed step");
    }
    return step_batch(handles, n, elapsed);
}

int ssw_rotate_directions(ssw_sweep* s, const double* matrix) {
    if (!s || !matrix) return fail(SSW_ERR_INVALID, "null argument");
    if (!s->have_grid) return fail(SSW_ERR_INVALID, "ssw_upload_grid must be called first");
    for (int k = 0; k < 9; k++)
        if (!std::isfinite(matrix[k])) return fail(SSW_ERR_INVALID, "rotation matrix must be finite");
    const int64_t n = s->N;
    const size_t DN = (size_t)s->D * n;
    // the rates before the rotation, as the reference's sites hold them
    double *dI = nullptr, *dO = nullptr, *dP = nullptr, *d_kept = nullptr;
    Defer cleanup{[&] {
        for (double** p : {&dI, &dO, &dP, &d_kept}) dfree(s, *p);
    }};
    if (s->have_state) {
        TRY(dalloc(s, &dI, DN));
        TRY(dalloc(s, &dO, DN));
        TRY(dalloc(s, &dP, DN));
        const Geo g = make_geo(s);
        if (s->fp32)
            k_fluxes<float><<<grid_for(n), kBlock, 0, s->stream>>>(g, make_flux<float>(s), s->lref, dI, dO, dP, nullptr,
                                                                  nullptr);
        else
            k_fluxes<double><<<grid_for(n), kBlock, 0, s->stream>>>(g, make_flux<double>(s), s->lref, dI, dO, dP,
                                                                   nullptr, nullptr);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    std::vector<double> kept;
    rotate_table(matrix, s->dirs, kept);
    HIP_TRY(hipMemcpyAsync(s->d_dirs, s->dirs.data(), s->dirs.size() * 8, hipMemcpyHostToDevice, s->stream));
    // the dependency graph and schedule of the rotated directions
    drop_all_lists(s);
    TRY(dalloc(s, &s->d_cnt, DN));
    TRY(build_direction_graph(s, false));
    s->rotated = true;
    if (s->have_state) {
        TRY(dalloc(s, &d_kept, kept.size()));
        HIP_TRY(hipMemcpyAsync(d_kept, kept.data(), kept.size() * 8, hipMemcpyHostToDevice, s->stream));
        TRY(s->fp32 ? rotate_state_t<float>(s, d_kept, dI, dO, dP) : rotate_state_t<double>(s, d_kept, dI, dO, dP));
        TRY(set_need_all(s, 1u));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SSW_OK;
}

int ssw_rotation_matrix(const double* axis, double angle, double* out) {
    if (!axis || !out) return fail(SSW_ERR_INVALID, "null argument");
    // get_rotation_matrix (src/sweep/direction/mod.rs:112-135)
    const double x = axis[0], y = axis[1], z = axis[2];
    const double c = std::cos(angle), sn = std::sin(angle);
    const double m[9] = {c + x * x * (1.0 - c),     x * y * (1.0 - c) - z * sn, x * z * (1.0 - c) + y * sn,
                         y * x * (1.0 - c) + z * sn, c + y * y * (1.0 - c),     y * z * (1.0 - c) - x * sn,
                         z * x * (1.0 - c) - y * sn, z * y * (1.0 - c) + x * sn, c + z * z * (1.0 - c)};
    std::memcpy(out, m, sizeof(m));
    return SSW_OK;
}

int ssw_random_rotation_matrix(double phi, double u, double psi, double* out) {
    if (!out) return fail(SSW_ERR_INVALID, "null argument");
    // get_random_rotation_matrix (src/sweep/direction/mod.rs:137-148) for its
    // three draws: phi in [0, 2 pi), u in [0, 1), psi in [0, 2 pi)
    const double theta = std::acos(2.0 * u - 1.0);
    const double axis[3] = {std::cos(phi) * std::sin(theta), std::sin(phi) * std::sin(theta), std::cos(theta)};
    return ssw_rotation_matrix(axis, psi, out);
}

int ssw_synchronize(ssw_sweep* s) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SSW_OK;
}

int ssw_download(ssw_sweep* s, double* x, double* T, double* sts, double* photon, uint8_t* level, double* change) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (!s->have_state) return fail(SSW_ERR_INVALID, "ssw_upload_state must be called first");
    const int64_t n = s->N;
    if (x) HIP_TRY(hipMemcpyAsync(x, s->d_x, n * 8, hipMemcpyDeviceToHost, s->stream));
    if (T) HIP_TRY(hipMemcpyAsync(T, s->d_T, n * 8, hipMemcpyDeviceToHost, s->stream));
    if (sts) HIP_TRY(hipMemcpyAsync(sts, s->d_sts, n * 8, hipMemcpyDeviceToHost, s->stream));
    if (change) HIP_TRY(hipMemcpyAsync(change, s->d_change, n * 8, hipMemcpyDeviceToHost, s->stream));
    if (level) HIP_TRY(hipMemcpyAsync(level, s->d_level, n, hipMemcpyDeviceToHost, s->stream));
    if (photon) {
        double* d_pr = nullptr;
        Defer cleanup{[&] { dfree(s, d_pr); }};
        TRY(dalloc(s, &d_pr, n));
        Geo g = make_geo(s);
        if (s->fp32)
            k_fluxes<float><<<grid_for(n), kBlock, 0, s->stream>>>(g, make_flux<float>(s), s->lref, nullptr, nullptr,
                                                                  nullptr, d_pr, nullptr);
        else
            k_fluxes<double><<<grid_for(n), kBlock, 0, s->stream>>>(g, make_flux<double>(s), s->lref, nullptr,
                                                                   nullptr, nullptr, d_pr, nullptr);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(photon, d_pr, n * 8, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        dfree(s, d_pr);
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (x) to_caller(s, x, 1);
    if (T) to_caller(s, T, 1);
    if (sts) to_caller(s, sts, 1);
    if (change) to_caller(s, change, 1);
    if (level) to_caller(s, level, 1);
    if (photon) to_caller(s, photon, 1);
    return SSW_OK;
}

int ssw_download_noalloc(ssw_sweep* s, double* x, double* T, double* sts, double* photon, uint8_t* level,
                         double* change) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (!s->have_state) return fail(SSW_ERR_INVALID, "ssw_upload_state must be called first");
    const int64_t n = s->N;
    if ((size_t)n * 8 > kNoAllocBytes)
        return fail(SSW_ERR_INVALID, "ssw_download_noalloc: " + std::to_string(n) + " cells exceed its limit of " +
                                         std::to_string(kNoAllocBytes / 8) + " cells (use ssw_download)");
    uint8_t* stage = nullptr;
    TRY(noalloc_scratch(s->device, &stage));
    if (x) TRY(noalloc_out(s, s->d_x, x, stage));
    if (T) TRY(noalloc_out(s, s->d_T, T, stage));
    if (sts) TRY(noalloc_out(s, s->d_sts, sts, stage));
    if (change) TRY(noalloc_out(s, s->d_change, change, stage));
    if (level) TRY(noalloc_out(s, s->d_level, level, stage));
    if (photon) {
        // the photon rate of every cell lands at its caller index in the staging buffer
        const Geo g = make_geo(s);
        double* pr = (double*)stage;
        const uint32_t* ridx = s->permuted ? s->d_inv : nullptr;
        if (s->fp32)
            k_fluxes<float><<<grid_for(n), kBlock, 0, s->stream>>>(g, make_flux<float>(s), s->lref, nullptr, nullptr,
                                                                  nullptr, pr, ridx);
        else
            k_fluxes<double><<<grid_for(n), kBlock, 0, s->stream>>>(g, make_flux<double>(s), s->lref, nullptr,
                                                                   nullptr, nullptr, pr, ridx);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(photon, pr, (size_t)n * 8, hipMemcpyDeviceToHost, s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SSW_OK;
}

int ssw_download_chemistry_outputs(ssw_sweep* s, double* photoionization, double* heating, double* recombination,
                                   double* collisional_ionization) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (!s->have_state) return fail(SSW_ERR_INVALID, "ssw_upload_state must be called first");
    const int64_t n = s->N;
    double* outs[4] = {photoionization, heating, recombination, collisional_ionization};
    double* d_out = nullptr;
    Defer cleanup{[&] { dfree(s, d_out); }};
    TRY(dalloc(s, &d_out, n));
    const Geo g = make_geo(s);
    const ChemArgs ca = chem_args(s);
    for (int k = 0; k < 4; k++) {
        if (!outs[k]) continue;
        if (s->fp32)
            k_chem_outputs<float><<<grid_for(n), kBlock, 0, s->stream>>>(g, make_flux<float>(s), ca, k, d_out);
        else
            k_chem_outputs<double><<<grid_for(n), kBlock, 0, s->stream>>>(g, make_flux<double>(s), ca, k, d_out);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(outs[k], d_out, n * 8, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        to_caller(s, outs[k], 1);
    }
    dfree(s, d_out);
    return SSW_OK;
}

int ssw_download_counts(ssw_sweep* s, int32_t lvl, uint32_t* counts) {
    if (!s || !counts) return fail(SSW_ERR_INVALID, "null argument");
    if (!s->have_grid) return fail(SSW_ERR_INVALID, "ssw_upload_grid must be called first");
    if (lvl < 0 || lvl >= s->p.num_timestep_levels) return fail(SSW_ERR_INVALID, "sweep_level out of range");
    Defer cleanup{[&] { release_layering_scratch(s); }};
    TRY(ensure_layering_scratch(s, false));
    Geo g = make_geo(s);
    k_count<<<grid_for((int64_t)s->N * s->D), kBlock, 0, s->stream>>>(g, lvl, 1);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(counts, s->d_cnt, (size_t)s->D * s->N * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    release_layering_scratch(s);
    to_caller(s, counts, s->D);
    return SSW_OK;
}

int ssw_download_levelsets(ssw_sweep* s, int32_t lvl, uint32_t* depth) {
    if (!s || !depth) return fail(SSW_ERR_INVALID, "null argument");
    if (!s->have_grid) return fail(SSW_ERR_INVALID, "ssw_upload_grid must be called first");
    if (lvl < 0 || lvl >= s->p.num_timestep_levels) return fail(SSW_ERR_INVALID, "sweep_level out of range");
    // Kahn layering of the restricted graph in the scratch records (the replay
    // schedule is not touched)
    std::vector<uint8_t> lev(s->N);
    HIP_TRY(hipMemcpyAsync(lev.data(), s->d_level, s->N, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    int64_t active = 0;
    for (uint8_t v : lev) active += v >= lvl;
    int k = 0;
    uint32_t* d_depth = nullptr;
    Defer cleanup{[&] {
        dfree(s, d_depth);
        release_layering_scratch(s);
    }};
    TRY(ensure_layering_scratch(s));
    TRY(kahn_layering(s, lvl, active * s->D, &k, nullptr));
    const int64_t DN = (int64_t)s->D * s->N;
    TRY(dalloc(s, &d_depth, (size_t)DN));
    k_fill_u32<<<grid_for(DN), kBlock, 0, s->stream>>>(d_depth, kNone, DN);
    Geo g = make_geo(s);
    k_scatter_depth<<<grid_for(std::min<int64_t>(DN, 1 << 22)), kBlock, 0, s->stream>>>(g, k, d_depth);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(depth, d_depth, (size_t)DN * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s, d_depth);
    release_layering_scratch(s);
    to_caller(s, depth, s->D);
    return SSW_OK;
}

int ssw_download_fluxes(ssw_sweep* s, double* incoming, double* outgoing, double* periodic) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (!s->have_state) return fail(SSW_ERR_INVALID, "ssw_upload_state must be called first");
    const size_t m = (size_t)s->D * s->N;
    double *dI = nullptr, *dO = nullptr, *dP = nullptr;
    Defer cleanup{[&] {
        for (double** p : {&dI, &dO, &dP}) dfree(s, *p);
    }};
    TRY(dalloc(s, &dI, m));
    TRY(dalloc(s, &dO, m));
    TRY(dalloc(s, &dP, m));
    Geo g = make_geo(s);
    if (s->fp32)
        k_fluxes<float><<<grid_for(s->N), kBlock, 0, s->stream>>>(g, make_flux<float>(s), s->lref, dI, dO, dP, nullptr,
                                                                  nullptr);
    else
        k_fluxes<double><<<grid_for(s->N), kBlock, 0, s->stream>>>(g, make_flux<double>(s), s->lref, dI, dO, dP,
                                                                  nullptr, nullptr);
    HIP_TRY(hipGetLastError());
    if (incoming) HIP_TRY(hipMemcpyAsync(incoming, dI, m * 8, hipMemcpyDeviceToHost, s->stream));
    if (outgoing) HIP_TRY(hipMemcpyAsync(outgoing, dO, m * 8, hipMemcpyDeviceToHost, s->stream));
    if (periodic) HIP_TRY(hipMemcpyAsync(periodic, dP, m * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s, dI);
    dfree(s, dO);
    dfree(s, dP);
    if (incoming) to_caller(s, incoming, s->D);
    if (outgoing) to_caller(s, outgoing, s->D);
    if (periodic) to_caller(s, periodic, s->D);
    return SSW_OK;
}

int ssw_download_transport(ssw_sweep* s, void* X, void* I) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (!s->have_state) return fail(SSW_ERR_INVALID, "ssw_upload_state must be called first");
    const int64_t DN = (int64_t)s->D * s->N;
    const size_t m = (size_t)DN * real_size(s);
    uint8_t* dtmp = nullptr;
    Defer cleanup{[&] { dfree(s, dtmp); }};
    if (X && s->compiled) {  // X is stored in schedule-position order
        TRY(dalloc(s, &dtmp, m));
        if (s->fp32)
            k_gather_x<float><<<grid_for(DN), kBlock, 0, s->stream>>>((const float*)s->d_X, s->d_xpos,
                                                                      (float*)dtmp, DN);
        else
            k_gather_x<double><<<grid_for(DN), kBlock, 0, s->stream>>>((const double*)s->d_X, s->d_xpos,
                                                                       (double*)dtmp, DN);
        HIP_TRY(hipGetLastError());
    }
    if (X) HIP_TRY(hipMemcpyAsync(X, dtmp ? (void*)dtmp : s->d_X, m, hipMemcpyDeviceToHost, s->stream));
    if (I) HIP_TRY(hipMemcpyAsync(I, s->d_I, m, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s, dtmp);
    for (void* a : {X, I}) {
        if (!a) continue;
        if (s->fp32) to_caller(s, (float*)a, s->D);
        else to_caller(s, (double*)a, s->D);
    }
    return SSW_OK;
}

// ---------------------------------------------------------------------------
// checkpoint of the dynamic state (no counterpart in the reference, which
// restarts from snapshots of the ECS components): a header, then per cell in
// the caller's order density, x, T, source per direction, attenuation,
// species timestep, previous rate, change timescale and level, then X and I
// direction-major in the caller's cell order and the periodic sources per
// direction for the periodic cells in increasing caller index, all in the
// storage precision and scale of the transport.

}  // extern "C"

namespace {

struct CkptHeader {
    uint64_t magic;
    uint32_t version, precision;
    int64_t N, Nper;
    int32_t D, L;
    double lref, max_ts;
    int32_t lowest, first_done;
};
constexpr uint64_t kCkptMagic = 0x31544b5057535353ull;  // "SSSWPKT1"

int64_t ckpt_bytes(const ssw_sweep* s) {
    const int64_t n = s->N, DN = (int64_t)s->D * n, rs = (int64_t)real_size(s);
    return (int64_t)sizeof(CkptHeader) + n * 8 * 8 + ((n + 7) / 8) * 8 + 2 * DN * rs + (int64_t)s->D * s->Nper * rs;
}

// Periodic slots ordered by the caller index of their cell.
std::vector<uint32_t> slots_by_caller(const ssw_sweep* s) {
    std::vector<std::pair<uint32_t, uint32_t>> v(s->h_pcells.size());
    for (size_t k = 0; k < v.size(); k++) {
        const uint32_t ci = s->h_pcells[k];
        v[k] = {s->permuted ? s->inv[ci] : ci, (uint32_t)k};
    }
    std::sort(v.begin(), v.end());
    std::vector<uint32_t> out(v.size());
    for (size_t k = 0; k < v.size(); k++) out[k] = v[k].second;
    return out;
}

template <typename R>
int save_t(ssw_sweep* s, uint8_t* buf) {
    const int64_t n = s->N, DN = (int64_t)s->D * n;
    CkptHeader h{kCkptMagic, 1u, (uint32_t)(s->fp32 ? 32 : 64), n, s->Nper, s->D, s->p.num_timestep_levels,
                 s->lref, s->p.max_timestep, s->ts.lowest, s->ts.first_done ? 1 : 0};
    std::memcpy(buf, &h, sizeof(h));
    uint8_t* p = buf + sizeof(h);
    std::vector<double2> sa(n);
    double* cell = reinterpret_cast<double*>(p);
    HIP_TRY(hipMemcpyAsync(cell + 0 * n, s->d_rho, n * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(cell + 1 * n, s->d_x, n * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(cell + 2 * n, s->d_T, n * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(cell + 3 * n, s->d_src, n * 8, hipMemcpyDeviceToHost, s->stream));
    TRY(download_sa(s, sa));
    HIP_TRY(hipMemcpyAsync(cell + 5 * n, s->d_sts, n * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(cell + 6 * n, s->d_prev, n * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(cell + 7 * n, s->d_change, n * 8, hipMemcpyDeviceToHost, s->stream));
    uint8_t* lev = p + n * 8 * 8;
    HIP_TRY(hipMemcpyAsync(lev, s->d_level, n, hipMemcpyDeviceToHost, s->stream));
    R* X = reinterpret_cast<R*>(lev + ((n + 7) / 8) * 8);
    R* I = X + DN;
    R* P = I + DN;
    R* tmp = nullptr;
    Defer cleanup{[&] { dfree(s, tmp); }};
    if (s->compiled) {
        TRY(dalloc(s, &tmp, (size_t)DN));
        k_gather_x<R><<<grid_for(DN), kBlock, 0, s->stream>>>((const R*)s->d_X, s->d_xpos, tmp, DN);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(X, tmp ? (void*)tmp : s->d_X, DN * sizeof(R), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(I, s->d_I, DN * sizeof(R), hipMemcpyDeviceToHost, s->stream));
    std::vector<R> Pd((size_t)s->D * s->Nper);
    if (s->Nper)
        HIP_TRY(hipMemcpyAsync(Pd.data(), s->d_P, Pd.size() * sizeof(R), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s, tmp);
    for (int64_t c = 0; c < n; c++) cell[4 * n + c] = sa[c].y;
    for (int k = 0; k < 8; k++) to_caller(s, cell + k * n, 1);
    to_caller(s, lev, 1);
    to_caller(s, X, s->D);
    to_caller(s, I, s->D);
    const std::vector<uint32_t> order = slots_by_caller(s);
    for (int d = 0; d < s->D; d++)
        for (size_t j = 0; j < order.size(); j++) P[(size_t)d * s->Nper + j] = Pd[(size_t)d * s->Nper + order[j]];
    return SSW_OK;
}

template <typename R>
int load_t(ssw_sweep* s, const uint8_t* buf) {
    const int64_t n = s->N, DN = (int64_t)s->D * n;
    CkptHeader h;
    std::memcpy(&h, buf, sizeof(h));
    const uint8_t* p = buf + sizeof(h);
    std::vector<double> cell((size_t)n * 8);
    std::memcpy(cell.data(), p, (size_t)n * 8 * 8);
    std::vector<uint8_t> lev(p + n * 8 * 8, p + n * 8 * 9);
    const R* Xc = reinterpret_cast<const R*>(p + n * 8 * 8 + ((n + 7) / 8) * 8);
    std::vector<R> X(Xc, Xc + DN), I(Xc + DN, Xc + 2 * DN);
    const R* Pc = Xc + 2 * DN;
    std::vector<double> tmpc;
    for (int k = 0; k < 8; k++) {
        const double* in = to_internal(s, cell.data() + k * n, tmpc);
        if (in != cell.data() + k * n) std::copy(in, in + n, cell.begin() + k * n);
    }
    {
        std::vector<uint8_t> tl;
        const uint8_t* in = to_internal(s, lev.data(), tl);
        if (in != lev.data()) lev.assign(in, in + n);
    }
    for (int d = 0; d < s->D; d++) {
        std::vector<R> tr;
        const R* inx = to_internal(s, X.data() + (size_t)d * n, tr);
        if (inx != X.data() + (size_t)d * n) std::copy(inx, inx + n, X.begin() + (size_t)d * n);
        const R* ini = to_internal(s, I.data() + (size_t)d * n, tr);
        if (ini != I.data() + (size_t)d * n) std::copy(ini, ini + n, I.begin() + (size_t)d * n);
    }
    const std::vector<uint32_t> order = slots_by_caller(s);
    std::vector<R> Pd((size_t)s->D * s->Nper);
    for (int d = 0; d < s->D; d++)
        for (size_t j = 0; j < order.size(); j++) Pd[(size_t)d * s->Nper + order[j]] = Pc[(size_t)d * s->Nper + j];
    s->lref = h.lref;
    std::vector<double2> sa(n);
    for (int64_t c = 0; c < n; c++)
        sa[c] = make_double2(s->fp32 ? cell[3 * n + c] / s->lref : cell[3 * n + c], cell[4 * n + c]);
    HIP_TRY(hipMemcpyAsync(s->d_rho, cell.data() + 0 * n, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_x, cell.data() + 1 * n, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_T, cell.data() + 2 * n, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_src, cell.data() + 3 * n, n * 8, hipMemcpyHostToDevice, s->stream));
    TRY(upload_sa(s, sa));
    HIP_TRY(hipMemcpyAsync(s->d_sts, cell.data() + 5 * n, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_prev, cell.data() + 6 * n, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_change, cell.data() + 7 * n, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_level, lev.data(), n, hipMemcpyHostToDevice, s->stream));
    R* tmp = nullptr;
    Defer cleanup{[&] { dfree(s, tmp); }};
    if (s->compiled) {
        TRY(dalloc(s, &tmp, (size_t)DN));
        HIP_TRY(hipMemcpyAsync(tmp, X.data(), DN * sizeof(R), hipMemcpyHostToDevice, s->stream));
        k_scatter_x<R><<<grid_for(DN), kBlock, 0, s->stream>>>((R*)s->d_X, s->d_xpos, tmp, DN);
        HIP_TRY(hipGetLastError());
    } else {
        HIP_TRY(hipMemcpyAsync(s->d_X, X.data(), DN * sizeof(R), hipMemcpyHostToDevice, s->stream));
    }
    HIP_TRY(hipMemcpyAsync(s->d_I, I.data(), DN * sizeof(R), hipMemcpyHostToDevice, s->stream));
    if (s->Nper)
        HIP_TRY(hipMemcpyAsync(s->d_P, Pd.data(), Pd.size() * sizeof(R), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s, tmp);
    s->ts.lowest = h.lowest;
    s->ts.first_done = h.first_done != 0;
    drop_all_lists(s);
    TRY(set_need_all(s, 1u));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SSW_OK;
}

}  // namespace

extern "C" {

int64_t ssw_checkpoint_bytes(ssw_sweep* s) {
    if (!s || !s->have_state) return 0;
    return ckpt_bytes(s);
}

int ssw_save_checkpoint(ssw_sweep* s, void* buf, int64_t bytes) {
    if (!s || !buf) return fail(SSW_ERR_INVALID, "null argument");
    if (!s->have_state) return fail(SSW_ERR_INVALID, "ssw_upload_state must be called first");
    if (bytes < ckpt_bytes(s)) return fail(SSW_ERR_INVALID, "checkpoint buffer too small");
    if (s->rotated) return fail(SSW_ERR_INVALID, "checkpoints of a handle with rotated directions are not supported");
    return s->fp32 ? save_t<float>(s, (uint8_t*)buf) : save_t<double>(s, (uint8_t*)buf);
}

int ssw_load_checkpoint(ssw_sweep* s, const void* buf, int64_t bytes) {
    if (!s || !buf) return fail(SSW_ERR_INVALID, "null argument");
    if (!s->have_state) return fail(SSW_ERR_INVALID, "ssw_upload_state must be called first");
    if (bytes < (int64_t)sizeof(CkptHeader)) return fail(SSW_ERR_INVALID, "checkpoint too small");
    if (s->rotated) return fail(SSW_ERR_INVALID, "checkpoints of a handle with rotated directions are not supported");
    CkptHeader h;
    std::memcpy(&h, buf, sizeof(h));
    if (h.magic != kCkptMagic || h.version != 1u) return fail(SSW_ERR_INVALID, "not a checkpoint of this library");
    if (h.N != s->N || h.D != s->D || h.L != s->p.num_timestep_levels || h.Nper != s->Nper ||
        h.precision != (uint32_t)(s->fp32 ? 32 : 64) || h.max_ts != s->p.max_timestep)
        return fail(SSW_ERR_INVALID, "checkpoint of a different problem (cells, directions, levels, periodic "
                                     "cells, precision or maximum timestep differ)");
    if (bytes < ckpt_bytes(s)) return fail(SSW_ERR_INVALID, "checkpoint truncated");
    return s->fp32 ? load_t<float>(s, (const uint8_t*)buf) : load_t<double>(s, (const uint8_t*)buf);
}

extern "C++" {
namespace {
// Workgroups of the grouped replay kernel one CU holds at once (0 if unknown).
template <typename R>
int group_blocks_per_cu(int roots) {
    int n = 0;
    hipError_t e;
    if (roots == 32) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_sweep_group<R, 64, 4>, 64, 0);
    else if (roots == 64) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_sweep_group<R, 128, 4>, 128, 0);
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_sweep_group<R, 256, 4>, 256, 0);
    return e == hipSuccess ? n : 0;
}
}  // namespace
}

int ssw_get_info(ssw_sweep* s, ssw_info* info) {
    if (!s || !info) return fail(SSW_ERR_INVALID, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->num_cells = s->N;
    info->num_faces = s->F;
    info->num_directions = s->D;
    info->max_depth = s->kmax;
    info->sweeps_last_step = s->sweeps_last;
    info->downwind_local = s->dw[0];
    info->downwind_periodic = s->dw[1];
    info->downwind_boundary = s->dw[2];
    info->device_bytes = s->bytes;
    info->peak_device_bytes = s->peak_bytes;
    info->transport_mode = s->have_grid ? (s->compiled ? SSW_TRANSPORT_MODE_COMPILED : SSW_TRANSPORT_MODE_GEOMETRIC)
                                        : s->mode_request;
    info->list_builds_last_step = (int32_t)s->builds_last;
    info->overflow_entries = s->ovf_total;
    info->graph_capture_failures = s->capture_failures;
    info->replay_block = s->rec_block;
    info->graph_replay = s->graphs_on ? 1 : 0;
    info->narrow_runs = (s->compiled && s->narrow_runs) ? 1 : 0;
    info->direction_sweeps = s->d_dirrun ? 1 : 0;
    info->group_levels = s->groups.empty() ? 1 : s->group_request;
    info->group_fallback = s->group_fallback ? 1 : 0;
    info->group_tasks = s->group_tasks;
    info->group_block = s->group_block;
    if (!s->groups.empty()) {
        info->group_launches = (int64_t)s->groups.size();
        for (const ssw_sweep::Group& gr : s->groups) info->group_workgroups += gr.ndesc;
        info->group_blocks_per_cu =
            s->fp32 ? group_blocks_per_cu<float>(s->group_block) : group_blocks_per_cu<double>(s->group_block);
    }
    {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device) == hipSuccess)
            info->compute_units = cus;
    }
    info->locality_order = s->permuted ? 1 : 0;
    if (s->have_state) {
        unsigned long long f = 0, a = 0, lim[3] = {0, 0, 0};
        HIP_TRY(hipMemcpyAsync(&f, s->d_fail, 8, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipMemcpyAsync(&a, s->d_active, 8, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipMemcpyAsync(lim, s->d_limit, 24, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        info->chemistry_failures = (int64_t)f;
        info->updates_last_step = (int64_t)a * s->D;
        info->limiting_temperature = (int64_t)lim[0];
        info->limiting_ionization_fraction = (int64_t)lim[1];
        info->limiting_photon_rate = (int64_t)lim[2];
    }
    return SSW_OK;
}

int32_t ssw_timestep_level(int32_t num_levels, double max_timestep, double desired_timestep) {
    if (num_levels < 1) return 0;
    return ssw::level_from_ratio(max_timestep / desired_timestep, num_levels);
}

double ssw_chemistry_timestep(double* state, double density, double volume, double length, double rate,
                              double scale_factor, int32_t prevent_cooling, double timestep, double safety,
                              int32_t* process, int32_t* failed) {
    ssw::Cell c;
    c.x = state[0];
    c.T = state[1];
    c.rho = density;
    c.vol = volume;
    c.len = length;
    c.rate = rate;
    c.a = scale_factor;
    ssw::Floor f{prevent_cooling != 0, c.T, c.x};
    int proc = 0;
    bool fl = false;
    double ts = ssw::perform_timestep(c, f, timestep, safety, proc, fl);
    state[0] = c.x;
    state[1] = c.T;
    if (process) *process = proc;
    if (failed) *failed = fl ? 1 : 0;
    return ts;
}

int64_t ssw_estimate_device_bytes(const ssw_sweep_parameters* p, int64_t n, int64_t F, int64_t num_periodic_cells,
                                  int32_t mode) {
    if (!p || n <= 0) return 0;
    const int64_t D = p->num_directions, L = p->num_timestep_levels;
    const int64_t rs = p->transport_precision == SSW_TRANSPORT_FP32 ? 4 : 8;
    const int64_t es = rs == 4 ? 8 : 16;
    const int64_t DN = D * n;
    const bool comp = mode == SSW_TRANSPORT_MODE_COMPILED;
    const int64_t ovf = comp ? DN * es / 16 : 0;   // overflow entries (bound used for the estimate)
    // held from the grid upload on: grid arrays (with the caller-order map of
    // a permuted grid), level starts, chunk table
    const int64_t grid = (n + 1) * 4 + F * 32 + F * 4 + n * 16 + n * 4 + num_periodic_cells * 8 + F * 4 + n +
                         (n + 4) * 4 + n * 4 + DN / kChunk * 40 + (1 << 20);
    const int64_t sched = comp ? DN * (16 + rs + kPlanes * es + 4) + ovf : DN * 16;  // compiled: + positions
    // compiled replay with the default grouped replay (4 level sets, 64 roots):
    // group lists instead of active-set lists, and the cone build's scratch
    const int64_t grp = comp ? group_bytes(DN, n, L, kGroupT) : 0;
    const int64_t lists = comp ? grp : (L - 1) * DN * 4;
    // transient peak of the schedule build: counters + layering records (+
    // overflow starts), the four sort arrays, or the schedule being written
    // with the sorted order and overflow starts
    const int64_t ostart = comp ? DN * 4 + ovf : 0;
    int64_t build = std::max(DN * 20 + ostart, std::max(DN * 16 + ostart, sched + DN * 4 + ostart));
    if (comp) build = std::max(build, sched + grp + DN * 5);
    // afterwards: schedule, active-set (or group) lists, state, the photon-rate download
    const int64_t run = sched + lists + n * 8 * 9 + 2 * DN * rs + D * num_periodic_cells * rs + n * 8;
    const int64_t b = grid + std::max(build, run);
    return b + b / 20 + (1 << 20);
}

int ssw_set_timing(ssw_sweep* s, int32_t enabled) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    s->timing = enabled != 0;
    return SSW_OK;
}

int ssw_get_timing(ssw_sweep* s, double* avg, double* total, int64_t* launches, double* chem, double* build) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (avg) *avg = s->n_transport ? s->t_transport_total / (double)s->n_transport : 0.0;
    if (total) *total = s->t_transport_total;
    if (launches) *launches = s->n_transport;
    if (chem) *chem = s->t_chem_total;
    if (build) *build = s->t_build_total;
    return SSW_OK;
}

}  // extern "C"
