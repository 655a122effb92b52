// tpm_chainmesh.hip -- chain mesh on the device.
//
// The reference bins particles into chain_dim^3 cells with a recursive
// in-place partition on the CPU (chainmesh.cpp:168-213, particles_sort
// particles.cpp:414-445) and keeps {pbegin, pend} per cell in a hash map
// (chainmesh.cpp:30-36).  Here: one key per particle (cell id in the
// partition's memory order, z slowest, times 64 plus a 4x4x4 sub-cell Morton
// code), a stable radix sort of (key, index), a gather of every particle
// array, and a dense cell table cell_begin[C^3 + 1] found from the sorted
// keys.  Cell ranges equal the reference partition's; the in-cell order is
// deterministic (sub-cell, then previous order) so CPU and GPU agree bit for
// bit.  Tiles of <= ctx->tile consecutive same-cell particles are the PP
// work items.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "tpm_internal.hpp"

namespace {

__global__ void key_kernel(int64_t n, const uint32_t* __restrict__ x, const uint32_t* __restrict__ y,
                           const uint32_t* __restrict__ z, int C, const uint32_t* __restrict__ bnd,
                           uint32_t* __restrict__ key, uint32_t* __restrict__ idx) {
    __shared__ uint32_t b[TPM_MAX_CHAIN];
    for (int k = threadIdx.x; k < C; k += blockDim.x) b[k] = bnd[k];
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int sx, sy, sz;
        const int cx = tpm_chain_coord(x[i], C, b, &sx);
        const int cy = tpm_chain_coord(y[i], C, b, &sy);
        const int cz = tpm_chain_coord(z[i], C, b, &sz);
        const uint32_t cid = ((uint32_t)cz * C + cy) * C + cx;
        const uint32_t m = ((sz >> 1) << 5) | ((sy >> 1) << 4) | ((sx >> 1) << 3) | ((sz & 1) << 2) |
                           ((sy & 1) << 1) | (sx & 1);
        key[i] = cid * 64u + m;
        idx[i] = (uint32_t)i;
    }
}

template <class T>
__global__ void gather_kernel(int64_t n, const uint32_t* __restrict__ perm, const T* __restrict__ src,
                              T* __restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[perm[i]];
}

// cell_begin[c] = first sorted position with cell id >= c.  Every entry is
// written by exactly one thread: the particle that opens the run of cells
// (cprev, c], and the last particle fills the tail.
__global__ void cell_bounds_kernel(int64_t n, const uint32_t* __restrict__ key, int64_t ncell,
                                   int32_t* __restrict__ cell_begin) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = key[i] >> 6;
        const int64_t cprev = i > 0 ? (int64_t)(key[i - 1] >> 6) : -1;
        for (int64_t cc = cprev + 1; cc <= c; cc++) cell_begin[cc] = (int32_t)i;
        if (i == n - 1)
            for (int64_t cc = c + 1; cc <= ncell; cc++) cell_begin[cc] = (int32_t)n;
    }
}

__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void tiles_per_cell_kernel(int64_t ncell, int ts, const int32_t* __restrict__ cell_begin,
                                      int32_t* __restrict__ nt) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
        const int32_t cnt = cell_begin[c + 1] - cell_begin[c];
        nt[c] = (cnt + ts - 1) / ts;
    }
}

__global__ void tiles_write_kernel(int64_t ncell, int ts, const int32_t* __restrict__ cell_begin,
                                   const int32_t* __restrict__ tile_end, int4* __restrict__ tiles,
                                   int32_t* __restrict__ ntiles) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = cell_begin[c], e = cell_begin[c + 1];
        const int32_t nt = (e - b + ts - 1) / ts;
        const int32_t t0 = tile_end[c] - nt;
        for (int32_t t = 0; t < nt; t++) {
            const int32_t tb = b + t * ts;
            const int32_t te = min(tb + ts, e);
            tiles[t0 + t] = make_int4((int)c, tb, te, 0);
        }
        if (c == ncell - 1) *ntiles = tile_end[c];
    }
}

}  // namespace

tpm_status tpm_chainmesh_build(tpm_ctx* ctx) {
    if (!ctx->params_set) return tpm_fail(ctx, TPM_ERR_STATE, "tpm_chainmesh_create: parameters not set");
    const int C = ctx->C;
    const int64_t n = ctx->n;
    const int64_t ncell = (int64_t)C * C * C;
    hipStream_t s = ctx->stream;

    // Split planes b[k] = fixed32(k * (1.0 / C)), computed exactly as the
    // reference does in double (chainmesh.cpp:198-199).
    if (ctx->chain_C != C) {
        std::vector<uint32_t> hb(C);
        const double ninv = 1.0 / C;
        for (int k = 0; k < C; k++) hb[k] = (uint32_t)(4294967296.0 * ((double)k * ninv));
        if (ctx->bounds) TPM_HIP(ctx, hipFree(ctx->bounds));
        TPM_HIP(ctx, hipMalloc(&ctx->bounds, sizeof(uint32_t) * C));
        TPM_HIP(ctx, hipMemcpy(ctx->bounds, hb.data(), sizeof(uint32_t) * C, hipMemcpyHostToDevice));
    }
    if (ncell + 1 > ctx->cells_cap) {
        if (ctx->cell_begin) TPM_HIP(ctx, hipFree(ctx->cell_begin));
        if (ctx->tile_end) TPM_HIP(ctx, hipFree(ctx->tile_end));
        if (ctx->tile_cnt) TPM_HIP(ctx, hipFree(ctx->tile_cnt));
        TPM_HIP(ctx, hipMalloc(&ctx->cell_begin, sizeof(int32_t) * (ncell + 1)));
        TPM_HIP(ctx, hipMalloc(&ctx->tile_end, sizeof(int32_t) * (ncell + 1)));
        TPM_HIP(ctx, hipMalloc(&ctx->tile_cnt, sizeof(int32_t) * (ncell + 1)));
        ctx->cells_cap = ncell + 1;
    }
    const int ts = ctx->tile;
    const int64_t tiles_max = n / ts + ncell + 1;
    if (tiles_max > ctx->tiles_cap) {
        if (ctx->tiles) TPM_HIP(ctx, hipFree(ctx->tiles));
        TPM_HIP(ctx, hipMalloc(&ctx->tiles, sizeof(int4) * tiles_max));
        ctx->tiles_cap = tiles_max;
    }
    if (!ctx->ntiles) TPM_HIP(ctx, hipMalloc(&ctx->ntiles, sizeof(int32_t) * 4));

    if (n == 0) {
        fill_i32_kernel<<<grid_for(ncell + 1, 256), 256, 0, s>>>(ctx->cell_begin, ncell + 1, 0);
        fill_i32_kernel<<<1, 64, 0, s>>>(ctx->ntiles, 1, 0);
        TPM_HIP(ctx, hipGetLastError());
        ctx->chain_valid = true;
        ctx->chain_C = C;
        ctx->chain_tile = ts;
        return TPM_OK;
    }

    key_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, ctx->x, ctx->y, ctx->z, C, ctx->bounds, ctx->keys, ctx->idx);
    TPM_HIP(ctx, hipGetLastError());

    unsigned end_bit = 1;
    while ((uint64_t(1) << end_bit) < (uint64_t)ncell * 64) end_bit++;
    size_t need = 0;
    TPM_HIP(ctx, rocprim::radix_sort_pairs(nullptr, need, ctx->keys, ctx->keys2, ctx->idx, ctx->perm, (size_t)n, 0,
                                           end_bit, s));
    if (need > ctx->sort_tmp_bytes) {
        if (ctx->sort_tmp) TPM_HIP(ctx, hipFree(ctx->sort_tmp));
        TPM_HIP(ctx, hipMalloc(&ctx->sort_tmp, need));
        ctx->sort_tmp_bytes = need;
    }
    TPM_HIP(ctx, rocprim::radix_sort_pairs(ctx->sort_tmp, need, ctx->keys, ctx->keys2, ctx->idx, ctx->perm,
                                           (size_t)n, 0, end_bit, s));

    const unsigned g = grid_for(n, 256);
    gather_kernel<uint32_t><<<g, 256, 0, s>>>(n, ctx->perm, ctx->x, ctx->x2);
    gather_kernel<uint32_t><<<g, 256, 0, s>>>(n, ctx->perm, ctx->y, ctx->y2);
    gather_kernel<uint32_t><<<g, 256, 0, s>>>(n, ctx->perm, ctx->z, ctx->z2);
    gather_kernel<float><<<g, 256, 0, s>>>(n, ctx->perm, ctx->vx, ctx->vx2);
    gather_kernel<float><<<g, 256, 0, s>>>(n, ctx->perm, ctx->vy, ctx->vy2);
    gather_kernel<float><<<g, 256, 0, s>>>(n, ctx->perm, ctx->vz, ctx->vz2);
    gather_kernel<int8_t><<<g, 256, 0, s>>>(n, ctx->perm, ctx->rung, ctx->rung2);
    TPM_HIP(ctx, hipGetLastError());
    std::swap(ctx->x, ctx->x2);
    std::swap(ctx->y, ctx->y2);
    std::swap(ctx->z, ctx->z2);
    std::swap(ctx->vx, ctx->vx2);
    std::swap(ctx->vy, ctx->vy2);
    std::swap(ctx->vz, ctx->vz2);
    std::swap(ctx->rung, ctx->rung2);

    cell_bounds_kernel<<<g, 256, 0, s>>>(n, ctx->keys2, ncell, ctx->cell_begin);
    // tiles per cell, then their inclusive scan.
    tiles_per_cell_kernel<<<grid_for(ncell, 256), 256, 0, s>>>(ncell, ts, ctx->cell_begin, ctx->tile_cnt);
    TPM_HIP(ctx, hipGetLastError());
    size_t need2 = 0;
    TPM_HIP(ctx, rocprim::inclusive_scan(nullptr, need2, ctx->tile_cnt, ctx->tile_end, (size_t)ncell,
                                         rocprim::plus<int32_t>(), s));
    if (need2 > ctx->scan_tmp_bytes) {
        if (ctx->scan_tmp) TPM_HIP(ctx, hipFree(ctx->scan_tmp));
        TPM_HIP(ctx, hipMalloc(&ctx->scan_tmp, need2));
        ctx->scan_tmp_bytes = need2;
    }
    TPM_HIP(ctx, rocprim::inclusive_scan(ctx->scan_tmp, need2, ctx->tile_cnt, ctx->tile_end, (size_t)ncell,
                                         rocprim::plus<int32_t>(), s));
    tiles_write_kernel<<<grid_for(ncell, 256), 256, 0, s>>>(ncell, ts, ctx->cell_begin, ctx->tile_end, ctx->tiles,
                                                            ctx->ntiles);
    TPM_HIP(ctx, hipGetLastError());
    ctx->chain_valid = true;
    ctx->chain_C = C;
    ctx->chain_tile = ts;
    return TPM_OK;
}
