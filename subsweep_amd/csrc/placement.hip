// Source placement on the device: the nearest cell of each point source
// (set_source_terms_system, src/source_systems.rs:48-80: kiddo's
// nearest_one over the cells' positions with squared Euclidean distance,
// non-periodic; rates of the sources that fall in one cell add up).
//
// The search is brute force, which the device does at its read bandwidth:
// workgroup (t, c) compares kSrcTile sources with chunk c of the cells, each
// lane over a strided subset, keeping the first minimum (lowest cell index)
// of the squared distance (dx*dx + dy*dy) + dz*dz, evaluated as written
// (no contraction), so the result is that of the host loop over cells in
// index order, ties going to the lower cell index; a second kernel takes the
// minimum over the chunks in chunk order. The rates are then added on the
// host, in source order, as the host placement adds them.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

#include "../../include/subsweep_amd.h"

namespace ssw_detail {
void set_error(const std::string& msg);
std::shared_mutex& capture_mutex();
}  // namespace ssw_detail

namespace {

constexpr int kSrcTile = 8;        // sources per workgroup
constexpr int kPlaceBlock = 256;   // lanes per workgroup
constexpr int64_t kChunk = 1 << 18;  // cells per workgroup column

// (d2, i) comes before (e2, j): i is a cell (not -1) and j is none, or the
// distance is smaller, or equal with the lower index.
__device__ __forceinline__ bool before(double d2, int64_t i, double e2, int64_t j) {
    if (i < 0) return false;
    if (j < 0) return true;
    return d2 < e2 || (d2 == e2 && i < j);
}

__global__ void __launch_bounds__(kPlaceBlock) k_nearest_chunk(const double* __restrict__ pts, int64_t n,
                                                               const double* __restrict__ spos, int64_t nsrc,
                                                               int64_t nchunks, double* __restrict__ part_d,
                                                               int64_t* __restrict__ part_i) {
    const int64_t s0 = (int64_t)blockIdx.x * kSrcTile;
    const int64_t chunk = blockIdx.y;
    const int64_t c0 = chunk * kChunk;
    const int64_t c1 = c0 + kChunk < n ? c0 + kChunk : n;
    double qx[kSrcTile], qy[kSrcTile], qz[kSrcTile], bd[kSrcTile];
    int64_t bi[kSrcTile];
#pragma unroll
    for (int k = 0; k < kSrcTile; k++) {
        const int64_t s = s0 + k < nsrc ? s0 + k : nsrc - 1;  // a valid source for every slot
        qx[k] = spos[3 * s];
        qy[k] = spos[3 * s + 1];
        qz[k] = spos[3 * s + 2];
        bd[k] = INFINITY;
        bi[k] = -1;
    }
    for (int64_t i = c0 + threadIdx.x; i < c1; i += kPlaceBlock) {
        const double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
#pragma unroll
        for (int k = 0; k < kSrcTile; k++) {
            const double dx = px - qx[k], dy = py - qy[k], dz = pz - qz[k];
            const double d2 = (dx * dx + dy * dy) + dz * dz;
            if (d2 < bd[k]) {  // strictly smaller: the lowest index of a tie stays
                bd[k] = d2;
                bi[k] = i;
            }
        }
    }
    __shared__ double sd[kPlaceBlock];
    __shared__ int64_t si[kPlaceBlock];
    for (int k = 0; k < kSrcTile; k++) {
        sd[threadIdx.x] = bd[k];
        si[threadIdx.x] = bi[k];
        __syncthreads();
        for (int w = kPlaceBlock / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) {
                const double e2 = sd[threadIdx.x + w];
                const int64_t j = si[threadIdx.x + w];
                if (before(e2, j, sd[threadIdx.x], si[threadIdx.x])) {
                    sd[threadIdx.x] = e2;
                    si[threadIdx.x] = j;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0 && s0 + k < nsrc) {
            part_d[(s0 + k) * nchunks + chunk] = sd[0];
            part_i[(s0 + k) * nchunks + chunk] = si[0];
        }
        __syncthreads();
    }
}

// Per source, the first minimum over the chunks (chunks hold ascending cell
// indices, so the strict comparison keeps the lowest index of a tie).
__global__ void k_nearest_final(const double* __restrict__ part_d, const int64_t* __restrict__ part_i,
                                int64_t nsrc, int64_t nchunks, int64_t* __restrict__ best) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsrc) return;
    double bd = INFINITY;
    int64_t bi = -1;
    for (int64_t c = 0; c < nchunks; c++) {
        const double d2 = part_d[s * nchunks + c];
        const int64_t i = part_i[s * nchunks + c];
        if (i >= 0 && (bi < 0 || d2 < bd)) {
            bd = d2;
            bi = i;
        }
    }
    best[s] = bi;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) {
            std::shared_lock<std::shared_mutex> lock(ssw_detail::capture_mutex());
            (void)hipFree(p);
        }
    }
};

int place_fail(int code, const std::string& msg) {
    ssw_detail::set_error(msg);
    return code;
}

}  // namespace

extern "C" {

int ssw_assign_sources_device(int64_t n, const double* pts, int64_t nsrc, const double* src_pos,
                              const double* src_rate, double* source_out, int64_t* cell_out) {
    if (n <= 0 || nsrc < 0 || !pts || !source_out || (nsrc > 0 && (!src_pos || !src_rate)))
        return place_fail(SSW_ERR_INVALID, "invalid arguments");
    for (int64_t i = 0; i < n; i++) source_out[i] = 0.0;
    if (nsrc == 0) return SSW_OK;
    const int64_t nchunks = (n + kChunk - 1) / kChunk;
    const int64_t tiles = (nsrc + kSrcTile - 1) / kSrcTile;
    if (tiles > 0x7FFFFFFF || nchunks > 65535) return place_fail(SSW_ERR_INVALID, "too many sources or cells");
    DevBuf d_pts, d_spos, d_pd, d_pi, d_best;
    auto alloc = [&](DevBuf& b, size_t bytes) {
        std::shared_lock<std::shared_mutex> lock(ssw_detail::capture_mutex());
        return hipMalloc(&b.p, bytes) == hipSuccess;
    };
    if (!alloc(d_pts, (size_t)n * 24) || !alloc(d_spos, (size_t)nsrc * 24) ||
        !alloc(d_pd, (size_t)nsrc * nchunks * 8) || !alloc(d_pi, (size_t)nsrc * nchunks * 8) ||
        !alloc(d_best, (size_t)nsrc * 8)) {
        (void)hipGetLastError();
        return place_fail(SSW_ERR_OOM, "device memory for the source placement");
    }
    if (hipMemcpy(d_pts.p, pts, (size_t)n * 24, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_spos.p, src_pos, (size_t)nsrc * 24, hipMemcpyHostToDevice) != hipSuccess)
        return place_fail(SSW_ERR_HIP, "copy of the positions");
    k_nearest_chunk<<<dim3((unsigned)tiles, (unsigned)nchunks), kPlaceBlock>>>(
        (const double*)d_pts.p, n, (const double*)d_spos.p, nsrc, nchunks, (double*)d_pd.p, (int64_t*)d_pi.p);
    k_nearest_final<<<(unsigned)((nsrc + 255) / 256), 256>>>((const double*)d_pd.p, (const int64_t*)d_pi.p, nsrc,
                                                            nchunks, (int64_t*)d_best.p);
    if (hipGetLastError() != hipSuccess) return place_fail(SSW_ERR_HIP, "source placement launch");
    std::vector<int64_t> best((size_t)nsrc);
    if (hipMemcpy(best.data(), d_best.p, (size_t)nsrc * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return place_fail(SSW_ERR_HIP, "copy of the nearest cells");
    for (int64_t s = 0; s < nsrc; s++) {
        if (best[s] < 0) return place_fail(SSW_ERR_INVALID, "a source has no nearest cell (non-finite position)");
        source_out[best[s]] += src_rate[s];
        if (cell_out) cell_out[s] = best[s];
    }
    return SSW_OK;
}

}  // extern "C"
