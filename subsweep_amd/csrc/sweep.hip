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
//   schedule   per sweep level l (the active set {level >= l}): the (cell,
//              direction) tasks in Kahn order, level-major; lstart[l][k] is where
//              level set k begins. Two replay forms:
//              geometric  task records {cell | face_count << 25, first_face,
//                         direction, periodic slot} (16 B per task and sweep level)
//              compiled   per task a header {cell | direction << 25, periodic
//                         slot, overflow start, upwind count}, the effective area
//                         A, and its upwind local faces as (neighbour, weight)
//                         pairs in kPlanes planes: entry q of the task at
//                         position i of a level set starting at b with width W
//                         is planes[kPlanes * b + q * W + (i - b)] (coalesced
//                         across tasks, addressable without reading the header);
//                         entries q >= kPlanes (rare) sit in a shared overflow
//                         list, indexed per (direction, cell).
//
// The transport is the reference's upwind sweep (src/sweep/mod.rs:412-513)
// restated in pull form. A cell that is solved distributes the change of its
// outgoing rate over its downwind faces in proportion to area * (n . d)
// (:453-461), so the incoming rate of a cell equals, up to rounding,
//   I[c,d] = sum_{upwind local faces f} area_f (-n_f . d) * O[nb,d] / A[nb,d]
// with A[nb,d] = sum of area (n . d) over nb's downwind faces. Storing
// X = O / A lets every task gather its upwind neighbours' X instead of
// scattering with atomics. The upwind sum is evaluated in face order, one term
// after the other, in every kernel, so the build, the geometric replay and the
// compiled replay give bitwise the same result, whatever order tasks run in.
// Tasks are run level set by level set: level k holds the (cell, direction)
// pairs whose active upwind neighbours are all in levels < k (Kahn layering of
// init_counts' counts, src/sweep/mod.rs:346-386). The level sets of an active
// set are built on the device by the first sweep that needs them (counter
// decrements, block-aggregated appends) and replayed by later sweeps, as one
// captured HIP graph of exact-size launches, until a timestep-level change
// alters the active set.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
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
constexpr int kRecBlock = 64;              // compiled replay: one lane per task
constexpr int kDefaultGroups = 1;          // direction groups run side by side

thread_local std::string g_err;
// Handles alive in the process. Replay graphs are captured only while a single
// handle exists: a stream capture is invalidated by the device-synchronising
// calls (allocation, free, blocking copies) that other handles' threads make.
std::atomic<int> g_live_handles{0};

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

struct Geo {
    int64_t N;
    int D;
    int L;
    int G;                   // direction groups (schedules run side by side)
    int Kcap;                // level starts per (slot, group) = Kcap + 2
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
    const uint8_t* dgroup;   // D: direction group of each direction
    const uint32_t* gbase;   // G: first task position of a group within a slot (N x its first direction)
    uint4* flat;             // task records: per slot (geometric) or one slot of scratch (compiled)
    size_t flat_slot_stride; // D * N (geometric) or 0 (compiled)
    uint32_t* lstart;        // L x G x (Kcap + 2), positions relative to the group's base
    uint32_t* ctr;           // L x G x 3 rotating append counters
    uint32_t* cnt;           // D x N missing-upwind counters
    uint32_t* need;          // L: schedule must be (re)built
};

template <typename R>
struct Flux {
    R* X;                    // D x N
    R* I;                    // D x N
    R* P;                    // D x Nper
    const double* src_over_d;
    const double* att;
    double thr;              // threshold in the storage scale
    double inv_lref;         // storage scale of rates
    double inv_aref;         // storage scale of areas
};

// Compiled schedule (null pointers in geometric mode).
template <typename R>
struct Comp {
    uint4* hdr;              // L x D*N
    R* A;                    // L x D*N
    Entry<R>* planes;        // L x kPlanes*D*N
    const uint32_t* ovf_start;  // D x N: first overflow entry of a task with > kPlanes upwind entries
    const Entry<R>* ovf;
};

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
};

template <typename R>
__device__ __forceinline__ CellIn<R> cell_inputs(const Geo& g, const Flux<R>& fl, uint32_t c, uint32_t d,
                                                 uint32_t ps) {
    CellIn<R> ci;
    ci.src = fl.src_over_d[c];
    ci.att = fl.att[c];
    // the periodic load uses slot 0 for cells without one (always a valid index)
    const R p = fl.P ? fl.P[(size_t)d * g.Nper + (ps != kNone ? ps : 0u)] : (R)0;
    ci.per = (ps != kNone) ? p : (R)0;
    return ci;
}

// Outgoing rate of a solved task (Site::get_rate + HydrogenOnly::get_outgoing_rate,
// src/sweep/site.rs:49-56, src/chemistry/hydrogen_only/mod.rs:72-88), stored as X = O / A.
template <typename R>
__device__ __forceinline__ void finish_task(const Flux<R>& fl, size_t o, R inc, const CellIn<R>& ci, R A) {
    if (inc < (R)0) inc = 0;  // Photons::make_positive
    R in = (inc + (R)ci.src) + ci.per;
    R out = (in < (R)fl.thr) ? (R)0 : in * (R)ci.att;
    fl.X[o] = (A > (R)0) ? out / A : (R)0;
    fl.I[o] = inc;
}

// The face loop of one task, kLanes lanes per task, shared by the build and
// the geometric replay so that both evaluate identical arithmetic. Lane `sub`
// handles faces sub, sub + kLanes, ...; the loop runs to the longest face list
// of the wave so that every lane takes part in the shuffles. The upwind sum is
// accumulated in face order on every lane of the group; A is a fixed-shape
// tree over the lanes' partial sums.
//   kTransport  gather the upwind X and sum
//   kBuild      decrement the downwind neighbours' counters and queue the
//               tasks that become ready into the block queue `sq`
//   kCompile    write the task's upwind entries into the planes
template <typename R, bool kTransport, bool kBuild, bool kCompile>
__device__ __forceinline__ void task_faces(const Geo& g, const Flux<R>& fl, bool valid, uint4 r, uint32_t sub,
                                           uint32_t nfmax, double& A_out, R& inc_out, uint32_t& nup_out,
                                           uint4* sq, uint32_t* sq_n, Entry<R>* planes, uint32_t W,
                                           uint32_t pos) {
    const uint32_t c = r.x & kCellMask;
    const uint32_t nf = valid ? (r.x >> kCellBits) : 0u;
    const uint32_t d = r.z;
    const double dx = g.dirs[3 * d], dy = g.dirs[3 * d + 1], dz = g.dirs[3 * d + 2];
    const size_t dN = (size_t)d * g.N;
    const int lane = __lane_id();
    const int gbase = lane & ~(kLanes - 1);
    double A = 0.0;
    R inc = 0;
    uint32_t nup = 0;
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
        if (kCompile) {
            const unsigned long long m = __ballot(up);
            const uint32_t seg = (uint32_t)(m >> gbase) & 0xFFFFu;
            const uint32_t q = nup + (uint32_t)__popc(seg & ((1u << sub) - 1u));
            if (up && q < (uint32_t)kPlanes) {
                Entry<R> e;
                e.nb = nb;
                if constexpr (sizeof(R) == 8) e.pad = 0;
                e.w = w;
                planes[(size_t)q * W + pos] = e;
            }
            nup += (uint32_t)__popc(seg);
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
    nup_out = nup;
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

// Overflow entries (upwind entries q >= kPlanes) of every (cell, direction), in
// face order, at ovf_start[d*N + c] (assigned here by an atomic cursor).
template <typename R>
__global__ void k_overflow(Geo g, double inv_aref, const uint32_t* ovf_count, uint32_t* ovf_start,
                           Entry<R>* ovf, uint32_t* cursor) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c, d;
    if (!cell_dir(g, t, c, d)) return;
    const size_t o = (size_t)d * g.N + c;
    const uint32_t extra = ovf_count[o];
    if (extra == 0) {
        ovf_start[o] = 0;
        return;
    }
    const uint32_t s = atomicAdd(cursor, extra);
    ovf_start[o] = s;
    const double dx = g.dirs[3 * d], dy = g.dirs[3 * d + 1], dz = g.dirs[3 * d + 2];
    uint32_t q = 0;
    for (uint32_t f = g.off[c]; f < g.off[c + 1]; f++) {
        const uint32_t nk = g.nk[f];
        const double4 a = g.geo[f];
        const double dd = fdot(a, dx, dy, dz);
        if (!((nk >> 30) == kKindLocal && dd < 0.0)) continue;
        if (q >= (uint32_t)kPlanes) {
            Entry<R> e;
            e.nb = nk & kNbrMask;
            if constexpr (sizeof(R) == 8) e.pad = 0;
            e.w = (R)(a.w * inv_aref * -dd);
            ovf[s + q - kPlanes] = e;
        }
        q++;
    }
}

// ---------------------------------------------------------------------------
// per-sweep kernels

// Resets the append counters of slot `lvl` when its schedule is rebuilt and
// takes the periodic-source snapshot of the active periodic cells (the
// transport reads the periodic source as it was when the sweep started).
// Thread t handles (periodic cell t / D, direction t % D).
template <typename R>
__global__ void k_begin(Geo g, Flux<R> fl, int lvl) {
    if (blockIdx.x == 0 && threadIdx.x < 3 * g.G && g.need[lvl] != 0) {
        g.ctr[lvl * 3 * g.G + threadIdx.x] = 0;
        if (threadIdx.x < g.G) g.lstart[((size_t)lvl * g.G + threadIdx.x) * (g.Kcap + 2)] = 0;
    }
    if (fl.P == nullptr) return;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = t / (uint32_t)g.D;
    const uint32_t d = t - s * (uint32_t)g.D;
    if ((int64_t)s >= g.Nper) return;
    const uint32_t c = g.pcells[s];
    if (g.level[c] < lvl) return;
    const double dx = g.dirs[3 * d], dy = g.dirs[3 * d + 1], dz = g.dirs[3 * d + 2];
    R acc = 0;
    for (uint32_t e = g.pfoff[s]; e < g.pfoff[s + 1]; e++) {
        const uint32_t f = g.pfidx[e];
        const uint32_t nk = g.nk[f];
        const double4 a = g.geo[f];
        const double dd = fdot(a, dx, dy, dz);
        if (dd < 0.0) acc += (R)(a.w * fl.inv_aref * -dd) * fl.X[(size_t)d * g.N + (nk & kNbrMask)];
    }
    fl.P[(size_t)d * g.Nper + s] = acc;
}

// init_counts (src/sweep/mod.rs:346-386) + get_initial_tasks (:388-398):
// counts the active local upwind neighbours of every active (cell, direction)
// and queues the tasks with count 0 as level set 0. Thread t handles
// (cell t / D, direction t % D).
__global__ void k_count(Geo g, int lvl, int only_counts) {
    const bool build = only_counts || g.need[lvl] != 0;
    if (!build) return;
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
    uint4* flat = g.flat + (size_t)lvl * g.flat_slot_stride;
    const bool ready = active && n == 0;
    const uint32_t grp = in ? g.dgroup[d] : 0u;
    const uint4 rec = ready ? task_record(c, f0, f1, d, g.pslot[c]) : make_uint4(0, 0, 0, 0);
    for (int q = 0; q < g.G; q++)  // one aggregated append per group present in the wave
        wave_append(ready && grp == (uint32_t)q, rec, &g.ctr[((size_t)lvl * g.G + q) * 3 + 0], flat + g.gbase[q]);
}

// Builds level set k+1 of slot `lvl` while solving level set k (kTransport) or
// only building (schedules for the parity hook and the depth bound). kLanes
// lanes per task; ready tasks are gathered per block in LDS and appended with
// one atomic per block. With kCompile also writes the compiled schedule.
template <typename R, bool kTransport, bool kCompile>
__global__ void __launch_bounds__(kBlock) k_sweep_build(Geo g, Flux<R> fl, Comp<R> cp, int lvl, int k, int grp) {
    __shared__ uint4 sq[kQueueCap];
    __shared__ uint32_t sq_n, sq_base, sq_cnt;
    const size_t unit = (size_t)lvl * g.G + grp;
    uint32_t* ls = g.lstart + unit * (g.Kcap + 2);
    uint32_t* ctr = g.ctr + unit * 3;
    const uint32_t beg = ls[k];
    const uint32_t end = beg + ctr[k % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ls[k + 1] = end;
        ctr[(k + 2) % 3] = 0;
    }
    if (beg >= end) return;
    const uint32_t W = end - beg;
    const uint32_t gb = g.gbase[grp];
    uint4* flat = g.flat + (size_t)lvl * g.flat_slot_stride + gb;
    uint32_t* next_ctr = ctr + (k + 1) % 3;
    const size_t DN = (size_t)g.D * g.N;
    Entry<R>* planes = kCompile ? cp.planes + (size_t)lvl * kPlanes * DN + (size_t)kPlanes * (gb + beg) : nullptr;
    const uint32_t sub = threadIdx.x % kLanes;
    const uint32_t stride = gridDim.x * kTasksPerBlock;
    if (threadIdx.x == 0) sq_n = 0;
    __syncthreads();
    for (uint32_t i0 = beg + blockIdx.x * kTasksPerBlock; i0 < end; i0 += stride) {
        const uint32_t i = i0 + threadIdx.x / kLanes;
        const bool valid = i < end;
        const uint4 r = valid ? flat[i] : make_uint4(0, 0, 0, 0);
        const uint32_t nfmax = wave_max(valid ? (r.x >> kCellBits) : 0u);
        const uint32_t c = r.x & kCellMask, d = r.z, ps = r.w;
        CellIn<R> ci{0.0, 0.0, (R)0};
        if (kTransport) ci = cell_inputs<R>(g, fl, c, d, ps);
        double A;
        R inc;
        uint32_t nup;
        task_faces<R, kTransport, true, kCompile>(g, fl, valid, r, sub, nfmax, A, inc, nup, sq, &sq_n, planes, W,
                                                  i - beg);
        if (valid && sub == 0) {
            const size_t o = (size_t)d * g.N + c;
            if (kTransport) finish_task<R>(fl, o, inc, ci, (R)A);
            if (kCompile) {
                const size_t so = (size_t)lvl * DN + gb + i;
                cp.hdr[so] = make_uint4(c | (d << kCellBits), ps, nup > (uint32_t)kPlanes ? cp.ovf_start[o] : 0u, nup);
                cp.A[so] = (R)A;
            }
        }
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
        for (uint32_t t = threadIdx.x; t < n; t += kBlock) flat[b + t] = sq[t];
        __syncthreads();
    }
}

// Geometric replay of one level set: `count` task records, kLanes lanes per task.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_sweep_rec(Geo g, Flux<R> fl, const uint4* recs, uint32_t count) {
    const uint32_t i = blockIdx.x * kTasksPerBlock + threadIdx.x / kLanes;
    const uint32_t sub = threadIdx.x % kLanes;
    const bool valid = i < count;
    const uint4 r = valid ? recs[i] : make_uint4(0, 0, 0, 0);
    const uint32_t ps = r.w;
    const uint32_t c = r.x & kCellMask, d = r.z;
    const CellIn<R> ci = cell_inputs<R>(g, fl, c, d, ps);
    const uint32_t nfmax = wave_max(valid ? (r.x >> kCellBits) : 0u);
    double A;
    R inc;
    uint32_t nup;
    task_faces<R, true, false, false>(g, fl, valid, r, sub, nfmax, A, inc, nup, nullptr, nullptr, nullptr, 0, 0);
    if (valid && sub == 0) finish_task<R>(fl, (size_t)d * g.N + c, inc, ci, (R)A);
}

// Compiled replay of one level set of width W: one lane per task. The header,
// A and the plane entries of a task are independent loads; the upwind X
// gathers follow. The sum runs over the upwind entries in face order.
template <typename R>
__global__ void __launch_bounds__(kRecBlock) k_sweep_comp(Geo g, Flux<R> fl, const uint4* hdr, const R* Av,
                                                          const Entry<R>* planes, const Entry<R>* ovf, uint32_t W) {
    const uint32_t i = blockIdx.x * kRecBlock + threadIdx.x;
    if (i >= W) return;
    // first memory step: the plane entries, the header and A (independent)
    Entry<R> e[kPlanes];
#pragma unroll
    for (int q = 0; q < kPlanes; q++) e[q] = planes[(size_t)q * W + i];
    const uint4 h = hdr[i];
    const R A = Av[i];
    const uint32_t c = h.x & kCellMask, d = h.x >> kCellBits, ps = h.y, nup = h.w;
    const size_t dN = (size_t)d * g.N;
    // second memory step: the upwind X (entries past the upwind count gather
    // the cell's own X, a valid index, and are not summed), the cell inputs
    const CellIn<R> ci = cell_inputs<R>(g, fl, c, d, ps);
    R x[kPlanes];
#pragma unroll
    for (int q = 0; q < kPlanes; q++) x[q] = fl.X[dN + ((uint32_t)q < nup ? e[q].nb : c)];
    R inc = 0;
#pragma unroll
    for (int q = 0; q < kPlanes; q++) inc += ((uint32_t)q < nup) ? e[q].w * x[q] : (R)0;
    for (uint32_t q = kPlanes; q < nup; q++) {
        const Entry<R> o = ovf[h.z + q - kPlanes];
        inc += o.w * fl.X[dN + o.nb];
    }
    finish_task<R>(fl, dN + c, inc, ci, A);
}

struct ChemArgs {
    double* rho;
    double* x;
    double* T;
    double* src;             // source per direction bin (SI)
    double* att;
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
};

// update_chemistry (src/sweep/mod.rs:549-574) for the active cells, then the
// attenuation of the updated cell for the next sweep.
template <typename R>
__global__ void k_chemistry(Geo g, Flux<R> fl, ChemArgs ca, int lvl) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) g.need[lvl] = 0;  // the schedule of this active set is complete
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
    double rate = 0.0;
    for (int d = 0; d < g.D; d++) {
        double inc = (double)fl.I[(size_t)d * g.N + c] * ca.lref;
        double per = 0.0;
        if (has_p) {  // current periodic source (the sweep has finished)
            double dx = g.dirs[3 * d], dy = g.dirs[3 * d + 1], dz = g.dirs[3 * d + 2];
            R acc = 0;
            for (uint32_t e = g.pfoff[ps]; e < g.pfoff[ps + 1]; e++) {
                const uint32_t f = g.pfidx[e];
                uint32_t nk = g.nk[f];
                double4 a = g.geo[f];
                double dd = fdot(a, dx, dy, dz);
                if (dd < 0.0) acc += (R)(a.w * fl.inv_aref * -dd) * fl.X[(size_t)d * g.N + (nk & kNbrMask)];
            }
            per = (double)acc * ca.lref;
        }
        rate += (inc + ca.src[c]) + per;  // sum over d of Site::get_rate, in direction order
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
    ca.change_ts[c] = (rate_ts < used) ? rate_ts : used;  // Timescale::min
    ca.att[c] = ssw::non_absorbed_fraction(cell.rho, cell.x, cell.len);
}

// update_timestep_levels (src/sweep/mod.rs:576-589) for all cells. Marks the
// schedules whose active set changed.
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
                         double* photon_rate) {
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
                R v = (R)(a.w * fl.inv_aref * -dd) * fl.X[(size_t)d * g.N + (nk & kNbrMask)];
                if ((nk >> 30) == kKindLocal) inc += v;
                else per += v;
            }
        }
        double inc_si = (double)inc * lref;
        total += inc_si;
        size_t o = (size_t)d * g.N + c;
        if (I_out) I_out[o] = inc_si;
        if (P_out) P_out[o] = (double)per * lref;
        if (O_out) O_out[o] = (double)fl.X[o] * (A * fl.inv_aref) * lref;
    }
    if (photon_rate) photon_rate[c] = total;
}

// Kahn depth of every task of a built schedule: entry i of the level-major
// array is in level set k with lstart[k] <= i < lstart[k+1]. `compiled`
// selects the header layout of the entries.
__global__ void k_scatter_depth(Geo g, const uint4* tasks, const uint32_t* ls, int kmax, int compiled,
                                uint32_t* depth) {
    const uint32_t total = ls[kmax];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        int a = 0, b = kmax - 1;  // last k with ls[k] <= i
        while (a < b) {
            int m = (a + b + 1) >> 1;
            if (ls[m] <= i) a = m;
            else b = m - 1;
        }
        const uint4 r = tasks[i];
        const uint32_t c = r.x & kCellMask;
        const uint32_t d = compiled ? (r.x >> kCellBits) : r.z;
        depth[(size_t)d * g.N + c] = (uint32_t)a;
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_attenuation(const double* rho, const double* x, const double* size, double* att, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) att[i] = ssw::non_absorbed_fraction(rho[i], x[i], size[i]);
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

// Host threads that submit the launches of direction groups 1..G-1 while the
// calling thread submits group 0 (a kernel launch costs microseconds of host
// time, so one thread cannot keep several streams fed).
struct SubmitPool {
    std::vector<std::thread> threads;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::function<void(int)> job;
    uint64_t generation = 0;
    int pending = 0;
    bool quit = false;

    void start(int workers) {
        for (int i = 0; i < workers; i++)
            threads.emplace_back([this, i] {
                uint64_t seen = 0;
                for (;;) {
                    std::function<void(int)> j;
                    {
                        std::unique_lock<std::mutex> lk(mu);
                        cv.wait(lk, [&] { return quit || generation != seen; });
                        if (quit) return;
                        seen = generation;
                        j = job;
                    }
                    j(i + 1);
                    {
                        std::lock_guard<std::mutex> lk(mu);
                        pending--;
                    }
                    cv_done.notify_one();
                }
            });
    }
    // Runs f(0) here and f(1..workers) on the pool; returns when all are done.
    void run(const std::function<void(int)>& f) {
        {
            std::lock_guard<std::mutex> lk(mu);
            job = f;
            pending = (int)threads.size();
            generation++;
        }
        cv.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(mu);
        cv_done.wait(lk, [&] { return pending == 0; });
    }
    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu);
            quit = true;
        }
        cv.notify_all();
        for (std::thread& th : threads) th.join();
        threads.clear();
        quit = false;
    }
};

}  // namespace

struct ssw_sweep {
    ssw_sweep_parameters p{};
    int D = 0;
    std::vector<double> dirs;
    int device = 0;
    hipStream_t stream = nullptr;
    bool fp32 = false;
    int mode_request = SSW_TRANSPORT_MODE_AUTO;
    bool compiled = false;
    // grid
    int64_t N = 0, F = 0, Nper = 0;
    bool have_grid = false, have_state = false;
    uint32_t *d_off = nullptr, *d_nk = nullptr, *d_pslot = nullptr, *d_pcells = nullptr;
    uint32_t *d_pfoff = nullptr, *d_pfidx = nullptr;
    double4* d_geo = nullptr;
    double *d_dirs = nullptr, *d_vol = nullptr, *d_size = nullptr;
    double aref = 1.0;
    int kmax = 0;                                   // deepest group
    double dw[3] = {0, 0, 0};
    int64_t ovf_total = 0;
    // direction groups: each group's level sets run on its own stream
    int G = 1;
    int G_request = 0;                              // 0: library default
    std::vector<uint32_t> gbase;                    // first task position of each group in a slot
    std::vector<int> kmax_g;                        // depth bound of each group
    uint8_t* d_dgroup = nullptr;
    uint32_t* d_gbase = nullptr;
    std::vector<hipStream_t> gstreams;
    std::vector<hipEvent_t> gjoin;
    hipEvent_t fork = nullptr;
    SubmitPool pool;                                // G - 1 submitting threads
    std::vector<std::vector<uint32_t>> full_start; // per group: level starts of the full graph
    // schedule
    uint4* d_flat = nullptr;
    uint32_t *d_lstart = nullptr, *d_ctr = nullptr, *d_cnt = nullptr, *d_need = nullptr;
    uint4* d_hdr = nullptr;
    void* d_A = nullptr;
    void* d_planes = nullptr;
    uint32_t* d_ovf_start = nullptr;
    void* d_ovf = nullptr;
    int64_t compiled_bytes = 0;
    int kcap = 0;
    std::vector<int> replay_ok;                     // per slot: host copy of the level starts valid
    std::vector<int> replays;                       // per slot: replays since its last build
    std::vector<std::vector<uint32_t>> level_start; // per (slot, group): level starts of its last build
    std::vector<hipGraphExec_t> graphs;             // per (slot, group): replay of its level sets
    // state
    double *d_rho = nullptr, *d_x = nullptr, *d_T = nullptr, *d_src = nullptr, *d_att = nullptr;
    double *d_sts = nullptr, *d_prev = nullptr, *d_change = nullptr, *d_src_t = nullptr;
    unsigned long long* d_active = nullptr;
    uint8_t* d_level = nullptr;
    void *d_X = nullptr, *d_I = nullptr, *d_P = nullptr;
    unsigned long long* d_fail = nullptr;
    double lref = 1.0;
    TimestepState ts;
    int64_t bytes = 0;
    std::unordered_map<void*, size_t> alloc;
    int64_t sweeps_last = 0;
    int64_t builds_last = 0;
    int64_t capture_failures = 0;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    double t_transport_total = 0, t_chem_total = 0, t_build_total = 0;
    int64_t n_transport = 0;
};

namespace {

template <typename T>
void dfree(ssw_sweep* s, T*& p) {
    if (p) {
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
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(SSW_ERR_OOM, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                     " bytes failed: " + hipGetErrorString(e));
    }
    s->bytes += (int64_t)(count * sizeof(T));
    s->alloc[(void*)*p] = count * sizeof(T);
    return SSW_OK;
}

size_t real_size(const ssw_sweep* s) { return s->fp32 ? 4 : 8; }
size_t entry_size(const ssw_sweep* s) { return s->fp32 ? sizeof(Entry<float>) : sizeof(Entry<double>); }

Geo make_geo(ssw_sweep* s) {
    Geo g;
    g.N = s->N;
    g.D = s->D;
    g.L = s->p.num_timestep_levels;
    g.G = s->G;
    g.Kcap = s->kcap;
    g.dgroup = s->d_dgroup;
    g.gbase = s->d_gbase;
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
    g.flat_slot_stride = s->compiled ? 0 : (size_t)s->D * s->N;
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
    f.I = (R*)s->d_I;
    f.P = (R*)s->d_P;
    f.src_over_d = s->fp32 ? s->d_src_t : s->d_src;
    f.att = s->d_att;
    f.inv_lref = 1.0 / s->lref;
    f.inv_aref = 1.0 / s->aref;
    f.thr = s->p.significant_rate_threshold * f.inv_lref;
    return f;
}

template <typename R>
Comp<R> make_comp(ssw_sweep* s) {
    Comp<R> c;
    c.hdr = s->d_hdr;
    c.A = (R*)s->d_A;
    c.planes = (Entry<R>*)s->d_planes;
    c.ovf_start = s->d_ovf_start;
    c.ovf = (const Entry<R>*)s->d_ovf;
    return c;
}

int set_need_all(ssw_sweep* s, uint32_t v) {
    k_fill_u32<<<1, kBlock, 0, s->stream>>>(s->d_need, v, s->p.num_timestep_levels);
    HIP_TRY(hipGetLastError());
    return SSW_OK;
}

void invalidate_replays(ssw_sweep* s) {
    for (hipGraphExec_t& ge : s->graphs)
        if (ge) {
            (void)hipGraphExecDestroy(ge);
            ge = nullptr;
        }
    std::fill(s->replay_ok.begin(), s->replay_ok.end(), 0);
}

hipStream_t gstream(ssw_sweep* s, int grp) { return s->G > 1 ? s->gstreams[grp] : s->stream; }

// Blocks for building level set k of (slot, group): from the widths of its
// last build, else of the full graph, with headroom; the kernel's grid-stride
// loop covers wider level sets.
unsigned build_blocks(ssw_sweep* s, int lvl, int grp, int k) {
    const std::vector<uint32_t>& last = s->level_start[(size_t)lvl * s->G + grp];
    const std::vector<uint32_t>& st = last.size() > (size_t)k + 1 ? last : s->full_start[grp];
    int64_t w = 1024;
    if (st.size() > (size_t)k + 1) w = (int64_t)(st[k + 1] - st[k]);
    w = w + w / 4 + 64;
    return (unsigned)std::min<int64_t>(16384, (w + kTasksPerBlock - 1) / kTasksPerBlock);
}

template <typename R>
void launch_build(ssw_sweep* s, const Geo& g, int lvl, int grp, int k, bool transport, unsigned blocks,
                  hipStream_t st) {
    Flux<R> fl = make_flux<R>(s);
    Comp<R> cp = make_comp<R>(s);
    if (transport) {
        if (s->compiled) k_sweep_build<R, true, true><<<blocks, kBlock, 0, st>>>(g, fl, cp, lvl, k, grp);
        else k_sweep_build<R, true, false><<<blocks, kBlock, 0, st>>>(g, fl, cp, lvl, k, grp);
    } else {
        if (s->compiled) k_sweep_build<R, false, true><<<blocks, kBlock, 0, st>>>(g, fl, cp, lvl, k, grp);
        else k_sweep_build<R, false, false><<<blocks, kBlock, 0, st>>>(g, fl, cp, lvl, k, grp);
    }
}

void launch_build_any(ssw_sweep* s, const Geo& g, int lvl, int grp, int k, bool transport, unsigned blocks,
                      hipStream_t st) {
    if (s->fp32) launch_build<float>(s, g, lvl, grp, k, transport, blocks, st);
    else launch_build<double>(s, g, lvl, grp, k, transport, blocks, st);
}

// Builds the schedule of the full graph (all cells active) once per grid, with
// no compiled output: the depth bound of every group and its level widths; a
// graph that never empties its queue has a cycle. The level starts are sized
// by a guess and the build restarts with more room if a group is deeper.
int build_full_schedule(ssw_sweep* s) {
    const int L = s->p.num_timestep_levels, G = s->G;
    const bool was_compiled = s->compiled;
    for (int kcap = 1024;; kcap *= 4) {
        s->kcap = kcap;
        TRY(dalloc(s, &s->d_lstart, (size_t)L * G * (kcap + 2)));
        TRY(set_need_all(s, 1u));
        s->compiled = false;  // records only
        Geo g = make_geo(s);
        g.flat_slot_stride = 0;
        Flux<double> fl{};
        Comp<double> cp{};
        k_begin<double><<<1, kBlock, 0, s->stream>>>(g, fl, 0);
        k_count<<<grid_for((int64_t)s->N * s->D), kBlock, 0, s->stream>>>(g, 0, 0);
        HIP_TRY(hipGetLastError());
        s->full_start.assign(G, std::vector<uint32_t>(1, 0u));
        std::vector<int64_t> total(G, 0);
        std::vector<int> depth(G, -1);
        std::vector<uint32_t> c3(3 * G);
        bool grow = false;
        int rc = SSW_OK;
        for (int k = 0;; k++) {
            HIP_TRY(hipMemcpyAsync(c3.data(), s->d_ctr, 3 * G * 4, hipMemcpyDeviceToHost, s->stream));
            HIP_TRY(hipStreamSynchronize(s->stream));
            bool any = false;
            for (int q = 0; q < G; q++) {
                if (depth[q] >= 0) continue;
                const uint32_t width = c3[3 * q + k % 3];
                if (width == 0) {
                    depth[q] = k;
                    continue;
                }
                any = true;
                total[q] += width;
                s->full_start[q].push_back((uint32_t)total[q]);
            }
            if (!any) break;
            if (k + 2 > kcap) {
                grow = true;
                break;
            }
            if (k > s->N + 1) {
                rc = fail(SSW_ERR_CYCLE, "dependency graph deeper than the number of cells");
                break;
            }
            for (int q = 0; q < G; q++) {
                if (depth[q] >= 0) continue;
                const int64_t width = (int64_t)(s->full_start[q][k + 1] - s->full_start[q][k]);
                const unsigned blocks = (unsigned)std::min<int64_t>(16384, (width + kTasksPerBlock - 1) / kTasksPerBlock);
                k_sweep_build<double, false, false><<<blocks, kBlock, 0, s->stream>>>(g, fl, cp, 0, k, q);
            }
            HIP_TRY(hipGetLastError());
        }
        s->compiled = was_compiled;
        if (rc != SSW_OK) return rc;
        if (grow) continue;
        int64_t all = 0;
        for (int q = 0; q < G; q++) all += total[q];
        if (all != s->N * (int64_t)s->D)
            return fail(SSW_ERR_CYCLE, "dependency cycle: " + std::to_string(s->N * (int64_t)s->D - all) +
                                           " cell-direction pairs never become ready");
        s->kmax_g.assign(G, 0);
        s->kmax = 1;
        for (int q = 0; q < G; q++) {
            s->kmax_g[q] = std::max(1, depth[q]);
            s->kmax = std::max(s->kmax, s->kmax_g[q]);
        }
        break;
    }
    s->kcap = s->kmax;
    TRY(dalloc(s, &s->d_lstart, (size_t)L * G * (s->kcap + 2)));
    s->replay_ok.assign(L, 0);
    s->replays.assign(L, 0);
    s->level_start.assign((size_t)L * G, std::vector<uint32_t>());
    for (hipGraphExec_t ge : s->graphs)
        if (ge) (void)hipGraphExecDestroy(ge);
    s->graphs.assign((size_t)L * G, nullptr);
    return SSW_OK;
}

// Launches the replay of (slot, group)'s level sets on stream `st`, one
// exact-size launch per non-empty level set; with `ev` each launch is
// bracketed by events.
template <typename R>
void launch_replay_t(ssw_sweep* s, const Geo& g, int lvl, int grp, hipStream_t st,
                     std::vector<std::pair<hipEvent_t, hipEvent_t>>* ev,
                     const std::function<hipEvent_t(hipStream_t)>* mark) {
    Flux<R> fl = make_flux<R>(s);
    Comp<R> cp = make_comp<R>(s);
    const std::vector<uint32_t>& ls = s->level_start[(size_t)lvl * s->G + grp];
    const size_t DN = (size_t)s->D * s->N;
    const uint32_t gb = s->gbase[grp];
    for (int k = 0; k < s->kmax_g[grp]; k++) {
        const uint32_t beg = gb + ls[k], count = ls[k + 1] - ls[k];
        if (count == 0) continue;
        hipEvent_t a = ev ? (*mark)(st) : nullptr;
        if (s->compiled) {
            const size_t so = (size_t)lvl * DN + beg;
            k_sweep_comp<R><<<(count + kRecBlock - 1) / kRecBlock, kRecBlock, 0, st>>>(
                g, fl, cp.hdr + so, cp.A + so, cp.planes + (size_t)lvl * kPlanes * DN + (size_t)kPlanes * beg, cp.ovf,
                count);
        } else {
            const uint4* recs = s->d_flat + (size_t)lvl * DN + beg;
            k_sweep_rec<R><<<(count + kTasksPerBlock - 1) / kTasksPerBlock, kBlock, 0, st>>>(g, fl, recs, count);
        }
        if (ev) ev->push_back({a, (*mark)(st)});
    }
}

void launch_replay(ssw_sweep* s, const Geo& g, int lvl, int grp, hipStream_t st,
                   std::vector<std::pair<hipEvent_t, hipEvent_t>>* ev,
                   const std::function<hipEvent_t(hipStream_t)>* mark = nullptr) {
    if (s->fp32) launch_replay_t<float>(s, g, lvl, grp, st, ev, mark);
    else launch_replay_t<double>(s, g, lvl, grp, st, ev, mark);
}

// After a schedule is built: copies its level starts to the host. Its replay
// graphs are captured on its second replay (a schedule that is rebuilt after
// one use never pays for a capture).
int build_replay_data(ssw_sweep* s, const Geo& g, int lvl) {
    (void)g;
    const int G = s->G;
    std::vector<uint32_t> all((size_t)G * (s->kcap + 2));
    HIP_TRY(hipMemcpyAsync(all.data(), s->d_lstart + (size_t)lvl * G * (s->kcap + 2), all.size() * 4,
                           hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int q = 0; q < G; q++) {
        std::vector<uint32_t>& ls = s->level_start[(size_t)lvl * G + q];
        ls.assign(all.begin() + (size_t)q * (s->kcap + 2), all.begin() + (size_t)q * (s->kcap + 2) + s->kmax_g[q] + 1);
        hipGraphExec_t& slot = s->graphs[(size_t)lvl * G + q];
        if (slot) {
            (void)hipGraphExecDestroy(slot);
            slot = nullptr;
        }
    }
    s->replay_ok[lvl] = 1;
    s->replays[lvl] = 0;
    return SSW_OK;
}

// Captures the replay of (slot, group)'s level sets as one graph on the
// group's stream (nothing is executed). A failed capture leaves no graph and
// the level sets are launched one by one (counted in ssw_info).
void capture_group(ssw_sweep* s, const Geo& g, int lvl, int q, int64_t* failures) {
    const hipStream_t st = gstream(s, q);
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess) {
        (void)hipGetLastError();
        (*failures)++;
        return;
    }
    launch_replay(s, g, lvl, q, st, nullptr);
    hipError_t ce = hipStreamEndCapture(st, &graph);
    if (ce != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        (*failures)++;
        return;
    }
    hipGraphExec_t ge = nullptr;
    hipError_t ie = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        (void)hipGetLastError();
        (*failures)++;
        return;
    }
    s->graphs[(size_t)lvl * s->G + q] = ge;
}

void drop_replay(ssw_sweep* s, int lvl) {
    s->replay_ok[lvl] = 0;
    s->replays[lvl] = 0;
    for (int q = 0; q < s->G; q++) {
        hipGraphExec_t& slot = s->graphs[(size_t)lvl * s->G + q];
        if (slot) {
            (void)hipGraphExecDestroy(slot);
            slot = nullptr;
        }
    }
}

template <typename R>
int run_step(ssw_sweep* s, double* elapsed) {
    Geo g = make_geo(s);
    Flux<R> fl = make_flux<R>(s);
    ChemArgs ca;
    ca.rho = s->d_rho;
    ca.x = s->d_x;
    ca.T = s->d_T;
    ca.src = s->d_src;
    ca.att = s->d_att;
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
    const std::vector<int> order = s->ts.sweep_order();
    const int L = s->p.num_timestep_levels;
    k_fill_u32<<<1, 2, 0, s->stream>>>((uint32_t*)s->d_active, 0u, 2);
    const unsigned gper = grid_for(std::max<int64_t>(s->Nper, 1) * s->D);
    const unsigned gcount = grid_for((int64_t)s->N * s->D);
    const unsigned gcells = grid_for(s->N);
    s->sweeps_last = 0;
    s->builds_last = 0;
    s->t_transport_total = s->t_chem_total = s->t_build_total = 0;
    s->n_transport = 0;
    size_t evn = 0;
    std::function<hipEvent_t(hipStream_t)> mark = [&](hipStream_t st) -> hipEvent_t {
        if (evn >= s->ev.size()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            s->ev.push_back(e);
        }
        hipEvent_t e = s->ev[evn++];
        (void)hipEventRecord(e, st);
        return e;
    };
    std::vector<std::pair<hipEvent_t, hipEvent_t>> tr, ch, bu;
    // Which schedules are (re)built in this step (device flags set by the level
    // update of the previous step).
    std::vector<uint32_t> need(L, 1u);
    HIP_TRY(hipMemcpyAsync(need.data(), s->d_need, L * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int l = 0; l < L; l++)
        if (need[l] && s->replay_ok[l]) drop_replay(s, l);
    const int G = s->G;
    for (int lvl : order) {
        k_begin<R><<<gper, kBlock, 0, s->stream>>>(g, fl, lvl);
        if (!need[lvl] && !s->replay_ok[lvl]) TRY(build_replay_data(s, g, lvl));  // built by a parity hook
        const bool rebuild = need[lvl] != 0;
        hipEvent_t b0 = (s->timing && rebuild) ? mark(s->stream) : nullptr;
        if (rebuild) k_count<<<gcount, kBlock, 0, s->stream>>>(g, lvl, 0);
        if (G > 1) HIP_TRY(hipEventRecord(s->fork, s->stream));
        const bool capture = !rebuild && !s->timing && s->replays[lvl] >= 1 && g_live_handles.load() == 1;
        std::vector<int64_t> cap_fail(G, 0);
        std::vector<hipError_t> herr(G, hipSuccess);
        // group q's launches on its stream: replay or build-while-solving
        auto submit = [&](int q) {
            const hipStream_t st = gstream(s, q);
            hipError_t e = hipSuccess;
            if (G > 1) e = hipStreamWaitEvent(st, s->fork, 0);
            if (!rebuild) {
                hipGraphExec_t& ge = s->graphs[(size_t)lvl * G + q];
                if (s->timing) {
                    launch_replay(s, g, lvl, q, st, &tr, &mark);
                } else {
                    if (!ge && capture) capture_group(s, g, lvl, q, &cap_fail[q]);
                    if (ge) e = e != hipSuccess ? e : hipGraphLaunch(ge, st);
                    else launch_replay(s, g, lvl, q, st, nullptr);
                }
            } else {
                for (int k = 0; k < s->kmax_g[q]; k++)
                    launch_build<R>(s, g, lvl, q, k, true, build_blocks(s, lvl, q, k), st);
            }
            if (e == hipSuccess) e = hipGetLastError();  // launch errors are per host thread
            if (G > 1 && e == hipSuccess) e = hipEventRecord(s->gjoin[q], st);
            herr[q] = e;
        };
        if (G > 1 && !s->timing && !s->pool.threads.empty()) {
            s->pool.run(submit);
        } else {
            for (int q = 0; q < G; q++) submit(q);
        }
        for (int q = 0; q < G; q++) {
            s->capture_failures += cap_fail[q];
            HIP_TRY(herr[q]);
            if (G > 1) HIP_TRY(hipStreamWaitEvent(s->stream, s->gjoin[q], 0));
        }
        HIP_TRY(hipGetLastError());
        if (rebuild) {
            if (s->timing) bu.push_back({b0, mark(s->stream)});
            TRY(build_replay_data(s, g, lvl));
            need[lvl] = 0;
            s->builds_last++;
        } else {
            s->replays[lvl]++;
        }
        hipEvent_t a = s->timing ? mark(s->stream) : nullptr;
        k_chemistry<R><<<gcells, kBlock, 0, s->stream>>>(g, fl, ca, lvl);
        if (s->timing) ch.push_back({a, mark(s->stream)});
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
            s->n_transport++;
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

void free_state(ssw_sweep* s) {
    dfree(s, s->d_rho);
    dfree(s, s->d_x);
    dfree(s, s->d_T);
    dfree(s, s->d_src);
    dfree(s, s->d_att);
    dfree(s, s->d_sts);
    dfree(s, s->d_prev);
    dfree(s, s->d_change);
    dfree(s, s->d_src_t);
    dfree(s, s->d_X);
    dfree(s, s->d_I);
    dfree(s, s->d_P);
}

void free_grid(ssw_sweep* s) {
    invalidate_replays(s);
    dfree(s, s->d_off);
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
    dfree(s, s->d_hdr);
    dfree(s, s->d_A);
    dfree(s, s->d_planes);
    dfree(s, s->d_ovf_start);
    dfree(s, s->d_ovf);
    dfree(s, s->d_dgroup);
    dfree(s, s->d_gbase);
}

// Allocates the compiled schedule and the overflow entries (once per grid).
int setup_compiled(ssw_sweep* s, uint32_t* d_ovf_count) {
    const size_t DN = (size_t)s->D * s->N;
    const int L = s->p.num_timestep_levels;
    TRY(dalloc(s, &s->d_hdr, (size_t)L * DN));
    TRY(dalloc(s, (uint8_t**)&s->d_A, (size_t)L * DN * real_size(s)));
    TRY(dalloc(s, (uint8_t**)&s->d_planes, (size_t)L * kPlanes * DN * entry_size(s)));
    TRY(dalloc(s, &s->d_ovf_start, DN));
    TRY(dalloc(s, (uint8_t**)&s->d_ovf, (size_t)std::max<int64_t>(1, s->ovf_total) * entry_size(s)));
    uint32_t* d_cursor = nullptr;
    TRY(dalloc(s, &d_cursor, 1));
    HIP_TRY(hipMemsetAsync(d_cursor, 0, 4, s->stream));
    Geo g = make_geo(s);
    const double inv_aref = 1.0 / s->aref;
    const unsigned gt = grid_for((int64_t)DN);
    if (s->fp32)
        k_overflow<float><<<gt, kBlock, 0, s->stream>>>(g, inv_aref, d_ovf_count, s->d_ovf_start,
                                                        (Entry<float>*)s->d_ovf, d_cursor);
    else
        k_overflow<double><<<gt, kBlock, 0, s->stream>>>(g, inv_aref, d_ovf_count, s->d_ovf_start,
                                                         (Entry<double>*)s->d_ovf, d_cursor);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s, d_cursor);
    s->compiled_bytes = (int64_t)((size_t)L * DN * (16 + real_size(s) + kPlanes * entry_size(s)) + DN * 4 +
                                  (size_t)std::max<int64_t>(1, s->ovf_total) * entry_size(s));
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
    g_live_handles++;
    s->p = p;
    s->p.explicit_directions = nullptr;
    s->D = p.num_directions;
    s->dirs = dirs;
    s->fp32 = p.transport_precision == SSW_TRANSPORT_FP32;
    (void)hipGetDevice(&s->device);
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        g_live_handles--;
        return fail(SSW_ERR_HIP, "hipStreamCreate failed");
    }
    int r = dalloc(s, &s->d_dirs, dirs.size());
    if (r == SSW_OK) r = dalloc(s, &s->d_need, (size_t)p.num_timestep_levels);
    if (r == SSW_OK) r = dalloc(s, &s->d_fail, 1);
    if (r == SSW_OK) r = dalloc(s, &s->d_active, 1);
    if (r != SSW_OK) {
        ssw_destroy(s);
        return r;
    }
    if (hipMemcpy(s->d_dirs, dirs.data(), dirs.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(s->d_fail, 0, 8) != hipSuccess || hipMemset(s->d_active, 0, 8) != hipSuccess) {
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
    s->pool.stop();
    free_state(s);
    free_grid(s);
    dfree(s, s->d_dirs);
    dfree(s, s->d_need);
    dfree(s, s->d_fail);
    dfree(s, s->d_active);
    for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : s->gjoin) (void)hipEventDestroy(e);
    if (s->fork) (void)hipEventDestroy(s->fork);
    for (hipStream_t st : s->gstreams) (void)hipStreamDestroy(st);
    if (s->stream) (void)hipStreamDestroy(s->stream);
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

int ssw_upload_grid(ssw_sweep* s, int64_t n, const int64_t* off, const int32_t* nbr, const uint8_t* kind,
                    const double* area, const double* normal, const double* volume, const double* size) {
    if (!s || !off || !nbr || !kind || !area || !normal || !volume || !size)
        return fail(SSW_ERR_INVALID, "null argument");
    if (n <= 0 || n > (int64_t)kCellMask) return fail(SSW_ERR_INVALID, "cell count must be in [1, 2^25 - 1]");
    if ((int64_t)s->D * n >= (int64_t)0xFFFFFFFFll) return fail(SSW_ERR_INVALID, "cells x directions exceeds 2^32");
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
    // pack faces; a non-periodic sweep treats periodic faces as boundaries
    std::vector<double4> geo(F);
    std::vector<uint32_t> nk(F);
    std::vector<uint32_t> off32(n + 1);
    std::vector<uint32_t> pslot(n, kNone), pcells, pfoff(1, 0), pfidx;
    double asum = 0.0;
    int64_t nfaces = 0;
    for (int64_t c = 0; c <= n; c++) off32[c] = (uint32_t)off[c];
    for (int64_t c = 0; c < n; c++) {
        bool has_p = false;
        for (int64_t f = off[c]; f < off[c + 1]; f++) {
            uint32_t k = kind[f];
            if (k > 2) return fail(SSW_ERR_INVALID, "face_kind must be 0, 1 or 2");
            if (k == kKindPeriodic && !s->p.periodic) k = kKindBoundary;
            int64_t nb = (k == kKindBoundary) ? 0 : nbr[f];
            if (k != kKindBoundary && (nb < 0 || nb >= n)) return fail(SSW_ERR_INVALID, "face_neighbour out of range");
            geo[f] = make_double4(normal[3 * f], normal[3 * f + 1], normal[3 * f + 2], area[f]);
            nk[f] = (uint32_t)nb | (k << 30);
            if (k == kKindPeriodic) has_p = true;
            asum += area[f];
            nfaces++;
        }
        if (has_p) {
            pslot[c] = (uint32_t)pcells.size();
            pcells.push_back((uint32_t)c);
            for (int64_t f = off[c]; f < off[c + 1]; f++)
                if ((nk[f] >> 30) == kKindPeriodic) pfidx.push_back((uint32_t)f);
            pfoff.push_back((uint32_t)pfidx.size());
        }
    }
    // Areas are carried in units of a power of two near the mean face area, so
    // that X = O / A keeps the range of O (SI areas reach 1e42 m^2); scaling
    // by a power of two is exact.
    s->aref = (nfaces > 0 && asum > 0.0) ? std::ldexp(1.0, (int)std::lround(std::log2(asum / (double)nfaces))) : 1.0;
    s->Nper = (int64_t)pcells.size();
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
    // direction groups: contiguous blocks of directions; each group's level
    // sets run on a stream of their own
    s->G = s->G_request > 0 ? std::min(s->G_request, s->D) : std::min(kDefaultGroups, s->D);
    std::vector<uint8_t> dgroup(s->D);
    s->gbase.assign(s->G, 0u);
    for (int q = 0; q < s->G; q++) {
        const int d0 = q * s->D / s->G, d1 = (q + 1) * s->D / s->G;
        s->gbase[q] = (uint32_t)((int64_t)d0 * n);
        for (int d = d0; d < d1; d++) dgroup[d] = (uint8_t)q;
    }
    while ((int)s->gstreams.size() < s->G) {
        hipStream_t st = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        s->gstreams.push_back(st);
        hipEvent_t e = nullptr;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        s->gjoin.push_back(e);
    }
    if (!s->fork) HIP_TRY(hipEventCreateWithFlags(&s->fork, hipEventDisableTiming));
    if ((int)s->pool.threads.size() != s->G - 1) {
        s->pool.stop();
        if (s->G > 1) s->pool.start(s->G - 1);
    }
    TRY(dalloc(s, &s->d_dgroup, (size_t)s->D));
    TRY(dalloc(s, &s->d_gbase, (size_t)s->G));
    HIP_TRY(hipMemcpyAsync(s->d_dgroup, dgroup.data(), s->D, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_gbase, s->gbase.data(), s->G * 4, hipMemcpyHostToDevice, s->stream));
    TRY(dalloc(s, &s->d_ctr, (size_t)L * s->G * 3));
    TRY(dalloc(s, &s->d_cnt, DN));
    HIP_TRY(hipMemcpyAsync(s->d_off, off32.data(), (n + 1) * 4, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_geo, geo.data(), F * sizeof(double4), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_nk, nk.data(), F * 4, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_vol, volume, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_size, size, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_pslot, pslot.data(), n * 4, hipMemcpyHostToDevice, s->stream));
    if (s->Nper) {
        HIP_TRY(hipMemcpyAsync(s->d_pcells, pcells.data(), s->Nper * 4, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->d_pfidx, pfidx.data(), pfidx.size() * 4, hipMemcpyHostToDevice, s->stream));
    }
    HIP_TRY(hipMemcpyAsync(s->d_pfoff, pfoff.data(), pfoff.size() * 4, hipMemcpyHostToDevice, s->stream));
    std::vector<uint8_t> lev(n, (uint8_t)(L - 1));
    HIP_TRY(hipMemcpyAsync(s->d_level, lev.data(), n, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_ctr, 0, (size_t)L * s->G * 3 * 4, s->stream));
    // mirror check, downwind statistics and overflow counts
    unsigned long long* d_stats = nullptr;
    TRY(dalloc(s, &d_stats, 5));
    HIP_TRY(hipMemsetAsync(d_stats, 0, 5 * 8, s->stream));
    Geo g = make_geo(s);
    k_check_mirror<<<grid_for(n), kBlock, 0, s->stream>>>(g, d_stats);
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
    // transport mode: compiled when its arrays fit (automatic: within half of
    // the free device memory), else geometric
    const size_t comp_bytes = (size_t)L * DN * (16 + real_size(s) + kPlanes * entry_size(s)) + DN * (16 + 4) +
                              (size_t)s->ovf_total * entry_size(s);
    const size_t geom_bytes = (size_t)L * DN * 16;
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    if (s->mode_request == SSW_TRANSPORT_MODE_COMPILED) s->compiled = true;
    else if (s->mode_request == SSW_TRANSPORT_MODE_GEOMETRIC) s->compiled = false;
    else s->compiled = comp_bytes <= free_b / 2;
    if (!s->compiled && geom_bytes > free_b)
        return fail(SSW_ERR_OOM, "task records need " + std::to_string(geom_bytes) + " bytes, " +
                                     std::to_string(free_b) + " free");
    TRY(dalloc(s, &s->d_flat, s->compiled ? DN : (size_t)L * DN));
    if (s->compiled) TRY(setup_compiled(s, s->d_cnt));
    TRY(build_full_schedule(s));
    s->have_grid = true;
    return SSW_OK;
}

int ssw_upload_state(ssw_sweep* s, const double* density, const double* xhii, const double* temp,
                     const double* source) {
    if (!s || !density || !xhii || !temp || !source) return fail(SSW_ERR_INVALID, "null argument");
    if (!s->have_grid) return fail(SSW_ERR_INVALID, "ssw_upload_grid must be called first");
    invalidate_replays(s);
    free_state(s);
    const int64_t n = s->N;
    const int L = s->p.num_timestep_levels;
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
    TRY(dalloc(s, &s->d_att, n));
    TRY(dalloc(s, &s->d_sts, n));
    TRY(dalloc(s, &s->d_prev, n));
    TRY(dalloc(s, &s->d_change, n));
    if (s->fp32) TRY(dalloc(s, &s->d_src_t, n));
    TRY(dalloc(s, (uint8_t**)&s->d_X, (size_t)s->D * n * es));
    TRY(dalloc(s, (uint8_t**)&s->d_I, (size_t)s->D * n * es));
    if (s->Nper) TRY(dalloc(s, (uint8_t**)&s->d_P, (size_t)s->D * s->Nper * es));
    HIP_TRY(hipMemcpyAsync(s->d_rho, density, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_x, xhii, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_T, temp, n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_src, sod.data(), n * 8, hipMemcpyHostToDevice, s->stream));
    std::vector<double> sod_scaled;
    if (s->fp32) {  // the FP32 transport reads source/D in its storage scale
        sod_scaled.resize(n);
        for (int64_t c = 0; c < n; c++) sod_scaled[c] = sod[c] / s->lref;
        HIP_TRY(hipMemcpyAsync(s->d_src_t, sod_scaled.data(), n * 8, hipMemcpyHostToDevice, s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemsetAsync(s->d_sts, 0, n * 8, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_prev, 0, n * 8, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_change, 0, n * 8, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_X, 0, (size_t)s->D * n * es, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_I, 0, (size_t)s->D * n * es, s->stream));
    if (s->Nper) HIP_TRY(hipMemsetAsync(s->d_P, 0, (size_t)s->D * s->Nper * es, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_fail, 0, 8, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_active, 0, 8, s->stream));
    std::vector<uint8_t> lev(n, (uint8_t)(L - 1));  // src/sweep/mod.rs:206
    HIP_TRY(hipMemcpyAsync(s->d_level, lev.data(), n, hipMemcpyHostToDevice, s->stream));
    k_attenuation<<<grid_for(n), kBlock, 0, s->stream>>>(s->d_rho, s->d_x, s->d_size, s->d_att, n);
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
        TRY(dalloc(s, &d_pr, n));
        Geo g = make_geo(s);
        if (s->fp32)
            k_fluxes<float><<<grid_for(n), kBlock, 0, s->stream>>>(g, make_flux<float>(s), s->lref, nullptr, nullptr,
                                                                  nullptr, d_pr);
        else
            k_fluxes<double><<<grid_for(n), kBlock, 0, s->stream>>>(g, make_flux<double>(s), s->lref, nullptr,
                                                                   nullptr, nullptr, d_pr);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(photon, d_pr, n * 8, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        dfree(s, d_pr);
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SSW_OK;
}

int ssw_download_counts(ssw_sweep* s, int32_t lvl, uint32_t* counts) {
    if (!s || !counts) return fail(SSW_ERR_INVALID, "null argument");
    if (!s->have_grid) return fail(SSW_ERR_INVALID, "ssw_upload_grid must be called first");
    if (lvl < 0 || lvl >= s->p.num_timestep_levels) return fail(SSW_ERR_INVALID, "sweep_level out of range");
    Geo g = make_geo(s);
    k_count<<<grid_for((int64_t)s->N * s->D), kBlock, 0, s->stream>>>(g, lvl, 1);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(counts, s->d_cnt, (size_t)s->D * s->N * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SSW_OK;
}

int ssw_download_levelsets(ssw_sweep* s, int32_t lvl, uint32_t* depth) {
    if (!s || !depth) return fail(SSW_ERR_INVALID, "null argument");
    if (!s->have_grid) return fail(SSW_ERR_INVALID, "ssw_upload_grid must be called first");
    if (lvl < 0 || lvl >= s->p.num_timestep_levels) return fail(SSW_ERR_INVALID, "sweep_level out of range");
    Geo g = make_geo(s);
    uint32_t need = 0;
    HIP_TRY(hipMemcpyAsync(&need, s->d_need + lvl, 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (need) {  // build the schedule of this active set without transport
        if (s->fp32) {
            Flux<float> fl{};
            k_begin<float><<<1, kBlock, 0, s->stream>>>(g, fl, lvl);
        } else {
            Flux<double> fl{};
            k_begin<double><<<1, kBlock, 0, s->stream>>>(g, fl, lvl);
        }
        k_count<<<grid_for((int64_t)s->N * s->D), kBlock, 0, s->stream>>>(g, lvl, 0);
        for (int q = 0; q < s->G; q++)
            for (int k = 0; k < s->kmax_g[q]; k++)
                launch_build_any(s, g, lvl, q, k, false, build_blocks(s, lvl, q, k), s->stream);
        k_fill_u32<<<1, 1, 0, s->stream>>>(s->d_need + lvl, 0u, 1);
        HIP_TRY(hipGetLastError());
        if ((int)s->replay_ok.size() > lvl) drop_replay(s, lvl);
    }
    uint32_t* d_depth = nullptr;
    const int64_t DN = (int64_t)s->D * s->N;
    TRY(dalloc(s, &d_depth, (size_t)DN));
    k_fill_u32<<<grid_for(DN), kBlock, 0, s->stream>>>(d_depth, kNone, DN);
    for (int q = 0; q < s->G; q++) {
        const size_t base = (size_t)lvl * DN + s->gbase[q];
        const uint4* tasks = s->compiled ? s->d_hdr + base : s->d_flat + base;
        k_scatter_depth<<<grid_for(std::min<int64_t>(DN, 1 << 22)), kBlock, 0, s->stream>>>(
            g, tasks, s->d_lstart + ((size_t)lvl * s->G + q) * (s->kcap + 2), s->kmax_g[q], s->compiled ? 1 : 0,
            d_depth);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(depth, d_depth, (size_t)DN * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s, d_depth);
    return SSW_OK;
}

int ssw_download_fluxes(ssw_sweep* s, double* incoming, double* outgoing, double* periodic) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (!s->have_state) return fail(SSW_ERR_INVALID, "ssw_upload_state must be called first");
    const size_t m = (size_t)s->D * s->N;
    double *dI = nullptr, *dO = nullptr, *dP = nullptr;
    TRY(dalloc(s, &dI, m));
    TRY(dalloc(s, &dO, m));
    TRY(dalloc(s, &dP, m));
    Geo g = make_geo(s);
    if (s->fp32)
        k_fluxes<float><<<grid_for(s->N), kBlock, 0, s->stream>>>(g, make_flux<float>(s), s->lref, dI, dO, dP, nullptr);
    else
        k_fluxes<double><<<grid_for(s->N), kBlock, 0, s->stream>>>(g, make_flux<double>(s), s->lref, dI, dO, dP,
                                                                  nullptr);
    HIP_TRY(hipGetLastError());
    if (incoming) HIP_TRY(hipMemcpyAsync(incoming, dI, m * 8, hipMemcpyDeviceToHost, s->stream));
    if (outgoing) HIP_TRY(hipMemcpyAsync(outgoing, dO, m * 8, hipMemcpyDeviceToHost, s->stream));
    if (periodic) HIP_TRY(hipMemcpyAsync(periodic, dP, m * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    dfree(s, dI);
    dfree(s, dO);
    dfree(s, dP);
    return SSW_OK;
}

int ssw_download_transport(ssw_sweep* s, void* X, void* I) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (!s->have_state) return fail(SSW_ERR_INVALID, "ssw_upload_state must be called first");
    const size_t m = (size_t)s->D * s->N * real_size(s);
    if (X) HIP_TRY(hipMemcpyAsync(X, s->d_X, m, hipMemcpyDeviceToHost, s->stream));
    if (I) HIP_TRY(hipMemcpyAsync(I, s->d_I, m, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SSW_OK;
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
    info->transport_mode = s->have_grid ? (s->compiled ? SSW_TRANSPORT_MODE_COMPILED : SSW_TRANSPORT_MODE_GEOMETRIC)
                                        : s->mode_request;
    info->schedule_builds_last_step = s->builds_last;
    info->overflow_entries = s->ovf_total;
    info->graph_capture_failures = s->capture_failures;
    info->direction_groups = s->have_grid ? s->G : s->G_request;
    if (s->have_state) {
        unsigned long long f = 0, a = 0;
        HIP_TRY(hipMemcpyAsync(&f, s->d_fail, 8, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipMemcpyAsync(&a, s->d_active, 8, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        info->chemistry_failures = (int64_t)f;
        info->updates_last_step = (int64_t)a * s->D;
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
    int64_t b = 0;
    b += (n + 1) * 4 + F * 32 + F * 4 + n * 16 + n * 4 + num_periodic_cells * 8 + F * 4 + n + L * 12 + DN * 4;
    b += L * (n + 2) * 4;                          // level starts (depth < number of cells)
    b += (mode == SSW_TRANSPORT_MODE_COMPILED) ? DN * 16 : L * DN * 16;
    if (mode == SSW_TRANSPORT_MODE_COMPILED) b += L * DN * (16 + rs + kPlanes * es) + DN * 4 + DN * es / 16;
    b += n * 8 * 9 + 2 * DN * rs + D * num_periodic_cells * rs;
    b += n * 8;                                    // photon-rate download
    return b + b / 20 + (1 << 20);
}

int ssw_set_direction_groups(ssw_sweep* s, int32_t groups) {
    if (!s) return fail(SSW_ERR_INVALID, "null handle");
    if (groups < 0 || groups > 16) return fail(SSW_ERR_INVALID, "direction groups must be in [0, 16]");
    if (s->have_grid) return fail(SSW_ERR_INVALID, "direction groups must be set before ssw_upload_grid");
    s->G_request = groups;
    return SSW_OK;
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
