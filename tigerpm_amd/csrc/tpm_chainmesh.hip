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

// Onesweep radix sort with 9-bit digits: the 512^3 keys (26 bits) take three
// passes instead of four (6.0 vs 6.5 ms for the whole chain mesh, same
// permutation: the sort is stable, so the digit width cannot change it).
using ChainSortConfig = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<512, 12>, rocprim::kernel_config<512, 12>, 9,
                                        rocprim::block_radix_rank_algorithm::match>>;

__global__ void key_kernel(int64_t n, const uint32_t* __restrict__ x, const uint32_t* __restrict__ y,
                           const uint32_t* __restrict__ z, int C, const uint32_t* __restrict__ bnd,
                           uint32_t* __restrict__ key, uint32_t* __restrict__ idx) {
    __shared__ uint32_t b[TPM_MAX_CHAIN];
    for (int k = threadIdx.x; k < C; k += blockDim.x) b[k] = bnd[k];
    __syncthreads();
    // 4 particles per thread and pass, their loads issued together (a
    // capped grid: the split-plane table is loaded once per workgroup)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i0 < n; i0 += 4 * stride) {
        uint32_t px[4], py[4], pz[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t i = min(i0 + u * stride, n - 1);
            px[u] = x[i];
            py[u] = y[i];
            pz[u] = z[i];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t i = i0 + u * stride;
            if (i >= n) break;
            int sx, sy, sz;
            const int cx = tpm_chain_coord(px[u], C, b, &sx);
            const int cy = tpm_chain_coord(py[u], C, b, &sy);
            const int cz = tpm_chain_coord(pz[u], C, b, &sz);
            const uint32_t cid = ((uint32_t)cz * C + cy) * C + cx;
            const uint32_t m = ((sz >> 1) << 5) | ((sy >> 1) << 4) | ((sx >> 1) << 3) | ((sz & 1) << 2) |
                               ((sy & 1) << 1) | (sx & 1);
            key[i] = cid * 64u + m;
            idx[i] = (uint32_t)i;
        }
    }
}

// The seven particle arrays in sorted order, in one pass (perm read once).
struct GatherArgs {
    const uint32_t* perm;
    const uint32_t *x, *y, *z;
    const float *vx, *vy, *vz;
    const int8_t* rung;
    uint32_t *x2, *y2, *z2;
    float *vx2, *vy2, *vz2;
    int8_t* rung2;
};

__global__ void gather_kernel(int64_t n, GatherArgs a) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t j = a.perm[i];
        a.x2[i] = a.x[j];
        a.y2[i] = a.y[j];
        a.z2[i] = a.z[j];
        a.vx2[i] = a.vx[j];
        a.vy2[i] = a.vy[j];
        a.vz2[i] = a.vz[j];
        a.rung2[i] = a.rung[j];
    }
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

// Tile order: cells are visited in z-major blocks of TILE_BLOCK^3 cells
// (blocks z-major, cells z-major inside a block), so the tiles the PP kernel
// runs close together in time share their 27-cell neighbourhoods: a block's
// sources (10^3 cells at 512^3, ~2 MB) and its stretch of the potential mesh
// fit one XCD's L2.  rank(c) is this order's position of cell c, a bijection
// onto [0, C^3) (edge blocks are smaller).  Each tile is computed
// independently, so the order changes no result.
#define TILE_BLOCK 8
__device__ __forceinline__ int32_t tile_rank(int64_t c, int C) {
    const int x = (int)(c % C), y = (int)((c / C) % C), z = (int)(c / ((int64_t)C * C));
    const int bx = x / TILE_BLOCK, by = y / TILE_BLOCK, bz = z / TILE_BLOCK;
    const int sy = min(TILE_BLOCK, C - by * TILE_BLOCK), sz = min(TILE_BLOCK, C - bz * TILE_BLOCK);
    const int sx = min(TILE_BLOCK, C - bx * TILE_BLOCK);
    const int64_t before = (int64_t)bz * TILE_BLOCK * C * C + (int64_t)sz * by * TILE_BLOCK * C +
                           (int64_t)sz * sy * bx * TILE_BLOCK;
    const int local = ((z - bz * TILE_BLOCK) * sy + (y - by * TILE_BLOCK)) * sx + (x - bx * TILE_BLOCK);
    return (int32_t)(before + local);
}

__global__ void tiles_per_cell_kernel(int64_t ncell, int C, int ts, const int32_t* __restrict__ cell_begin,
                                      int32_t* __restrict__ nt) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
        const int32_t cnt = cell_begin[c + 1] - cell_begin[c];
        nt[tile_rank(c, C)] = (cnt + ts - 1) / ts;
    }
}
__global__ void tiles_write_kernel(int64_t ncell, int C, int ts, const int32_t* __restrict__ cell_begin,
                                   const int32_t* __restrict__ tile_end, int4* __restrict__ tiles,
                                   int32_t* __restrict__ ntiles) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = cell_begin[c], e = cell_begin[c + 1];
        const int32_t nt = (e - b + ts - 1) / ts;
        const int32_t r = tile_rank(c, C);
        const int32_t t0 = tile_end[r] - nt;
        for (int32_t t = 0; t < nt; t++) {
            const int32_t tb = b + t * ts;
            const int32_t te = min(tb + ts, e);
            tiles[t0 + t] = make_int4((int)c, tb, te, 0);
        }
        if (r == ncell - 1) *ntiles = tile_end[r];
    }
}

// ---------------------------------------------------------------------------
// The order of the reference partition (chainmesh.cpp:168-213 with the
// two-pointer particles_sort, particles.cpp:414-445), found level by level.
//
// The recursion splits a box of chain cells on its last axis that is more
// than one cell wide (z until the box is one z plane, then y, then x) at
// mid = begin + span / 2.  particles_sort partitions the box's particle range
// [s, e) in place: lo walks up, and every particle lo meets that belongs to
// the right is swapped with the next particle that belongs to the left,
// searched downwards from the top.  The sweep has a closed form: with m =
// s + (particles of the left box), the k-th right particle of [s, m)
// (counted upwards) trades places with the k-th left particle of [m, e)
// (counted downwards); nothing else moves.  Every box of the recursion is a
// contiguous run of cell ids (z-major order), so s, m and e are entries of
// the finished cell table.  One level moves every range of that depth at
// once: a flag per position, one exclusive scan of
// (right-in-left << 32 | left-in-right), the positions of the misplaced
// particles by scan index, and the exchange.
// ---------------------------------------------------------------------------
#define REF_LEVELS 10  // bisection levels per axis: ceil(log2(TPM_MAX_CHAIN)) <= 9

// Box of the recursion at level j for the particle of cell c: its cell-id
// range [c0, c1) split at cm; false when the particle's cell is final.
__device__ __forceinline__ bool ref_box(uint32_t c, int C, int j, const int4* __restrict__ tab, int64_t* c0,
                                        int64_t* cm, int64_t* c1, bool* right) {
    const int ix = (int)(c % (uint32_t)C), iy = (int)((c / (uint32_t)C) % (uint32_t)C),
              iz = (int)(c / ((uint32_t)C * (uint32_t)C));
    const int dz = tab[iz * REF_LEVELS].w, dy = tab[iy * REF_LEVELS].w, dx = tab[ix * REF_LEVELS].w;
    int4 t;
    int64_t base, stride;
    int coord;
    if (j < dz) {
        t = tab[iz * REF_LEVELS + j];
        base = 0;
        stride = (int64_t)C * C;
        coord = iz;
    } else if (j < dz + dy) {
        t = tab[iy * REF_LEVELS + (j - dz)];
        base = (int64_t)iz * C * C;
        stride = C;
        coord = iy;
    } else if (j < dz + dy + dx) {
        t = tab[ix * REF_LEVELS + (j - dz - dy)];
        base = ((int64_t)iz * C + iy) * C;
        stride = 1;
        coord = ix;
    } else {
        return false;
    }
    *c0 = base + t.x * stride;
    *cm = base + t.y * stride;
    *c1 = base + t.z * stride;
    *right = coord >= t.y;
    return true;
}

constexpr unsigned long long REF_LO = 0xffffffffull;

// cid[h] = cell of the caller's particle h, idx[h] = h.
__global__ void ref_init_kernel(int64_t n, const uint32_t* __restrict__ key, const uint32_t* __restrict__ h2d,
                                uint32_t* __restrict__ cid, uint32_t* __restrict__ idx) {
    for (int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; h < n; h += (int64_t)gridDim.x * blockDim.x) {
        cid[h] = key[h2d ? h2d[h] : (uint32_t)h] >> 6;
        idx[h] = (uint32_t)h;
    }
}

__global__ void ref_flag_kernel(int64_t n, int C, int j, const int4* __restrict__ tab,
                                const int32_t* __restrict__ cb, const uint32_t* __restrict__ cid,
                                unsigned long long* __restrict__ flag) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p <= n; p += (int64_t)gridDim.x * blockDim.x) {
        unsigned long long f = 0ull;
        int64_t c0, cm, c1;
        bool right;
        if (p < n && ref_box(cid[p], C, j, tab, &c0, &cm, &c1, &right)) {
            const int64_t m = cb[cm];
            if (p < m && right) f = 1ull << 32;
            if (p >= m && !right) f = 1ull;
        }
        flag[p] = f;
    }
}

__global__ void ref_scatter_kernel(int64_t n, const unsigned long long* __restrict__ flag,
                                   const unsigned long long* __restrict__ scan, uint32_t* __restrict__ posa,
                                   uint32_t* __restrict__ posb) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long f = flag[p];
        if (f >> 32) posa[scan[p] >> 32] = (uint32_t)p;
        if (f & REF_LO) posb[scan[p] & REF_LO] = (uint32_t)p;
    }
}

__global__ void ref_move_kernel(int64_t n, int C, int j, const int4* __restrict__ tab, const int32_t* __restrict__ cb,
                                const unsigned long long* __restrict__ flag,
                                const unsigned long long* __restrict__ scan, const uint32_t* __restrict__ posa,
                                const uint32_t* __restrict__ posb, const uint32_t* __restrict__ cid,
                                const uint32_t* __restrict__ idx, uint32_t* __restrict__ cid2,
                                uint32_t* __restrict__ idx2) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t c = cid[p];
        const unsigned long long f = flag[p];
        int64_t dst = p;
        if (f) {
            int64_t c0, cm, c1;
            bool right;
            ref_box(c, C, j, tab, &c0, &cm, &c1, &right);
            const unsigned long long ss = scan[cb[c0]], se = scan[cb[c1]];
            if (f >> 32) {
                // k-th right particle of [s, m) <-> k-th left particle of [m, e) from the top
                const unsigned long long k = (scan[p] >> 32) - (ss >> 32);
                dst = posb[(se & REF_LO) - 1ull - k];
            } else {
                const unsigned long long k = (se & REF_LO) - 1ull - (scan[p] & REF_LO);
                dst = posa[(ss >> 32) + k];
            }
        }
        cid2[dst] = c;
        idx2[dst] = idx[p];
    }
}

// inv[perm[d]] = d: device position after the sort of the particle at
// device position perm[d] before it.
__global__ void invert_kernel(int64_t n, const uint32_t* __restrict__ perm, uint32_t* __restrict__ inv) {
    for (int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; d < n; d += (int64_t)gridDim.x * blockDim.x)
        inv[perm[d]] = (uint32_t)d;
}

// h2d_new[r] = device position of the particle the caller now sees at r.
__global__ void ref_compose_kernel(int64_t n, const uint32_t* __restrict__ ref, const uint32_t* __restrict__ h2d_old,
                                   const uint32_t* __restrict__ inv, uint32_t* __restrict__ h2d_new) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t h = ref[r];
        h2d_new[r] = inv[h2d_old ? h2d_old[h] : h];
    }
}

}  // namespace

// Per axis coordinate i and bisection level k (rows of REF_LEVELS): the
// interval [lo, hi) of the reference partition's boxes that holds i and its
// split point mid = lo + (hi - lo) / 2 (chainmesh.cpp:189-201); w = the
// number of levels until i's interval is one cell wide.
static void ref_axis_table(int C, std::vector<int4>& tab) {
    tab.assign((size_t)C * REF_LEVELS, make_int4(0, 0, 0, 0));
    for (int i = 0; i < C; i++) {
        int lo = 0, hi = C, k = 0;
        while (hi - lo > 1) {
            const int mid = lo + (hi - lo) / 2;
            tab[(size_t)i * REF_LEVELS + k] = make_int4(lo, mid, hi, 0);
            if (i < mid)
                hi = mid;
            else
                lo = mid;
            k++;
        }
        for (int q = 0; q < REF_LEVELS; q++) tab[(size_t)i * REF_LEVELS + q].w = k;
    }
}

// The reference partition of the caller's order, after tpm_chainmesh_build
// has sorted and gathered (ctx->keys: cell keys in the device order before
// the sort, ctx->perm: the sort's permutation, ctx->cell_begin: the table).
static tpm_status ref_order(tpm_ctx* ctx) {
    const int C = ctx->C;
    const int64_t n = ctx->n;
    hipStream_t s = ctx->stream;
    if (n > ctx->ref_cap || !ctx->h2d) {
        uint32_t** u32s[] = {&ctx->h2d,      &ctx->h2d2,     &ctx->ref_perm, &ctx->ref_cid, &ctx->ref_cid2,
                             &ctx->ref_idx2, &ctx->ref_posa, &ctx->ref_posb, &ctx->ref_inv};
        for (uint32_t** b : u32s) {
            if (*b) TPM_HIP(ctx, hipFree(*b));
            *b = nullptr;
            TPM_HIP(ctx, hipMalloc(b, sizeof(uint32_t) * std::max<int64_t>(n, 1)));
        }
        unsigned long long** u64s[] = {&ctx->ref_flag, &ctx->ref_scan};
        for (unsigned long long** b : u64s) {
            if (*b) TPM_HIP(ctx, hipFree(*b));
            *b = nullptr;
            TPM_HIP(ctx, hipMalloc(b, sizeof(unsigned long long) * (n + 1)));
        }
        ctx->ref_cap = std::max<int64_t>(n, 1);
        ctx->h2d_identity = true;  // the old map went with its buffer; only reached after an upload
    }
    if (ctx->ref_tab_C != C) {
        std::vector<int4> tab;
        ref_axis_table(C, tab);
        if (ctx->ref_tab) TPM_HIP(ctx, hipFree(ctx->ref_tab));
        ctx->ref_tab = nullptr;
        TPM_HIP(ctx, hipMalloc(&ctx->ref_tab, sizeof(int4) * tab.size()));
        TPM_HIP(ctx, hipMemcpy(ctx->ref_tab, tab.data(), sizeof(int4) * tab.size(), hipMemcpyHostToDevice));
        ctx->ref_tab_C = C;
    }
    int depth = 0;
    while ((1 << depth) < C) depth++;
    const uint32_t* old_h2d = ctx->h2d_identity ? nullptr : ctx->h2d;
    const unsigned g = grid_for(n, 256), g1 = grid_for(n + 1, 256);
    uint32_t *cid = ctx->ref_cid, *cid2 = ctx->ref_cid2, *idx = ctx->ref_perm, *idx2 = ctx->ref_idx2;
    ref_init_kernel<<<g, 256, 0, s>>>(n, ctx->keys, old_h2d, cid, idx);
    TPM_HIP(ctx, hipGetLastError());
    size_t need = 0;
    TPM_HIP(ctx, rocprim::exclusive_scan(nullptr, need, ctx->ref_flag, ctx->ref_scan, 0ull, (size_t)(n + 1),
                                         rocprim::plus<unsigned long long>(), s));
    if (need > ctx->ref_scan_tmp_bytes) {
        if (ctx->ref_scan_tmp) TPM_HIP(ctx, hipFree(ctx->ref_scan_tmp));
        ctx->ref_scan_tmp = nullptr;
        TPM_HIP(ctx, hipMalloc(&ctx->ref_scan_tmp, need));
        ctx->ref_scan_tmp_bytes = need;
    }
    for (int j = 0; j < 3 * depth; j++) {
        ref_flag_kernel<<<g1, 256, 0, s>>>(n, C, j, ctx->ref_tab, ctx->cell_begin, cid, ctx->ref_flag);
        TPM_HIP(ctx, hipGetLastError());
        TPM_HIP(ctx, rocprim::exclusive_scan(ctx->ref_scan_tmp, need, ctx->ref_flag, ctx->ref_scan, 0ull,
                                             (size_t)(n + 1), rocprim::plus<unsigned long long>(), s));
        ref_scatter_kernel<<<g, 256, 0, s>>>(n, ctx->ref_flag, ctx->ref_scan, ctx->ref_posa, ctx->ref_posb);
        ref_move_kernel<<<g, 256, 0, s>>>(n, C, j, ctx->ref_tab, ctx->cell_begin, ctx->ref_flag, ctx->ref_scan,
                                          ctx->ref_posa, ctx->ref_posb, cid, idx, cid2, idx2);
        TPM_HIP(ctx, hipGetLastError());
        std::swap(cid, cid2);
        std::swap(idx, idx2);
    }
    // idx now holds the reference permutation; keep it as ref_perm
    ctx->ref_perm = idx;
    ctx->ref_idx2 = idx2;
    ctx->ref_cid = cid;
    ctx->ref_cid2 = cid2;
    invert_kernel<<<g, 256, 0, s>>>(n, ctx->perm, ctx->ref_inv);
    ref_compose_kernel<<<g, 256, 0, s>>>(n, ctx->ref_perm, old_h2d, ctx->ref_inv, ctx->h2d2);
    TPM_HIP(ctx, hipGetLastError());
    std::swap(ctx->h2d, ctx->h2d2);
    ctx->h2d_identity = false;
    ctx->ref_valid = true;
    return TPM_OK;
}

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
        ctx->chain_epoch++;
        ctx->ref_valid = ctx->host_order == TPM_ORDER_REFERENCE;
        return TPM_OK;
    }

    key_kernel<<<std::min(grid_for(n, 256 * 4), 8192u), 256, 0, s>>>(n, ctx->x, ctx->y, ctx->z, C, ctx->bounds,
                                                                        ctx->keys, ctx->idx);
    TPM_HIP(ctx, hipGetLastError());

    unsigned end_bit = 1;
    while ((uint64_t(1) << end_bit) < (uint64_t)ncell * 64) end_bit++;
    size_t need = 0;
    TPM_HIP(ctx, rocprim::radix_sort_pairs<ChainSortConfig>(nullptr, need, ctx->keys, ctx->keys2, ctx->idx, ctx->perm, (size_t)n, 0,
                                           end_bit, s));
    if (need > ctx->sort_tmp_bytes) {
        if (ctx->sort_tmp) TPM_HIP(ctx, hipFree(ctx->sort_tmp));
        TPM_HIP(ctx, hipMalloc(&ctx->sort_tmp, need));
        ctx->sort_tmp_bytes = need;
    }
    TPM_HIP(ctx, rocprim::radix_sort_pairs<ChainSortConfig>(ctx->sort_tmp, need, ctx->keys, ctx->keys2, ctx->idx, ctx->perm,
                                           (size_t)n, 0, end_bit, s));

    const unsigned g = grid_for(n, 256);
    gather_kernel<<<g, 256, 0, s>>>(n, GatherArgs{ctx->perm, ctx->x, ctx->y, ctx->z, ctx->vx, ctx->vy, ctx->vz,
                                                  ctx->rung, ctx->x2, ctx->y2, ctx->z2, ctx->vx2, ctx->vy2,
                                                  ctx->vz2, ctx->rung2});
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
    tiles_per_cell_kernel<<<grid_for(ncell, 256), 256, 0, s>>>(ncell, C, ts, ctx->cell_begin, ctx->tile_cnt);
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
    tiles_write_kernel<<<grid_for(ncell, 256), 256, 0, s>>>(ncell, C, ts, ctx->cell_begin, ctx->tile_end, ctx->tiles,
                                                            ctx->ntiles);
    TPM_HIP(ctx, hipGetLastError());
    ctx->chain_valid = true;
    ctx->chain_C = C;
    ctx->chain_tile = ts;
    ctx->chain_epoch++;
    ctx->ref_valid = false;
    if (ctx->host_order == TPM_ORDER_REFERENCE) TPM_TRY(ref_order(ctx));
    return TPM_OK;
}

// ---------------------------------------------------------------------------
// Batched chain mesh over every box of a tpm_batch.
//
// The same order as tpm_chainmesh_build (cells of the box z-major, in a cell
// by sub-cell Morton code then previous position), found without a radix
// sort so that the step can be captured in a HIP graph (no memset nodes):
//   key + count: per particle its global cell (box cell offset + cell id)
//                and an atomic slot in that cell's count; particles of
//                departed boxes go to a trash cell after the last one;
//   scan:        exclusive scan of the counts -> cell_begin;
//   scatter:     particle -> cell_begin[cell] + slot (order within a cell
//                still arbitrary);
//   cell order:  inside each cell, the rank of every particle by (Morton
//                code, previous position) -> its final place;
//   gather:      every particle array into the second buffer set.
// The counts and ranks are exact integers, so the result is deterministic.
// ---------------------------------------------------------------------------
namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_ITEMS;

// cnt[c] += 1 for every active lane, one atomic per distinct c in the wave;
// returns the lane's slot (count before it).  Which lane of a cell gets which
// slot depends on timing; the in-cell ranking later fixes the order.
__device__ __forceinline__ uint32_t wave_count_slot(int32_t* cnt, uint32_t c) {
    const int lane = (int)(threadIdx.x & 63);
    unsigned long long todo = __ballot(1);
    uint32_t slot = 0;
    while (todo) {
        const int leader = __builtin_ctzll(todo);
        const uint32_t lc = (uint32_t)__shfl((int)c, leader);
        const unsigned long long same = __ballot(c == lc) & todo;
        int base = 0;
        if (lane == leader) base = atomicAdd(&cnt[lc], __popcll(same));
        base = __shfl(base, leader);
        if (c == lc) slot = (uint32_t)base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        todo &= ~same;
    }
    return slot;
}

__global__ __launch_bounds__(256) void bkey_kernel(int64_t n, int nbox, const tpm_box_dev* __restrict__ boxes,
                                                   int32_t trash, const uint32_t* __restrict__ x,
                                                   const uint32_t* __restrict__ y, const uint32_t* __restrict__ z,
                                                   uint32_t* __restrict__ key, uint32_t* __restrict__ slot,
                                                   int32_t* __restrict__ cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int b = tpm_upper_index(nbox, i, [&](int k) { return boxes[k].p0; });
        uint32_t gc = (uint32_t)trash, m = 0;
        if (b >= 0 && boxes[b].live && i < boxes[b].p0 + boxes[b].n) {
            const int C = boxes[b].C;
            const double ninv = boxes[b].ninv;
            int sx, sy, sz;
            const int cx = tpm_chain_coord_f(x[i], C, ninv, &sx);
            const int cy = tpm_chain_coord_f(y[i], C, ninv, &sy);
            const int cz = tpm_chain_coord_f(z[i], C, ninv, &sz);
            gc = (uint32_t)boxes[b].cell0 + ((uint32_t)cz * C + cy) * C + cx;
            m = ((sz >> 1) << 5) | ((sy >> 1) << 4) | ((sx >> 1) << 3) | ((sz & 1) << 2) | ((sy & 1) << 1) | (sx & 1);
        }
        key[i] = gc * 64u + m;
        slot[i] = wave_count_slot(cnt, gc);
    }
}

__global__ void bfill_i32_kernel(int32_t* __restrict__ p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t* sh, int32_t* total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < SCAN_THREADS; o <<= 1) {
        const int32_t a = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    const int32_t incl = sh[t];
    *total = sh[SCAN_THREADS - 1];
    __syncthreads();
    return incl - v;
}

// Phase 1: the sum of every SCAN_CHUNK-item chunk.
__global__ __launch_bounds__(SCAN_THREADS) void scan_partials_kernel(const int32_t* __restrict__ in, int64_t n,
                                                                     int32_t* __restrict__ part) {
    __shared__ int32_t sh[SCAN_THREADS];
    const int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK + (int64_t)threadIdx.x * SCAN_ITEMS;
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) s += base + k < n ? in[base + k] : 0;
    int32_t tot;
    block_exclusive_scan(s, sh, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// Phase 2 (one block): exclusive scan of the chunk sums in place; out[n] =
// the grand total.
__global__ __launch_bounds__(SCAN_THREADS) void scan_part_kernel(int32_t* __restrict__ part, int64_t nchunk,
                                                                 int32_t* __restrict__ out_total) {
    __shared__ int32_t sh[SCAN_THREADS];
    const int64_t per = (nchunk + SCAN_THREADS - 1) / SCAN_THREADS;
    const int64_t b0 = (int64_t)threadIdx.x * per;
    int32_t s = 0;
    for (int64_t k = b0; k < b0 + per && k < nchunk; k++) s += part[k];
    int32_t tot;
    int32_t run = block_exclusive_scan(s, sh, &tot);
    for (int64_t k = b0; k < b0 + per && k < nchunk; k++) {
        const int32_t v = part[k];
        part[k] = run;
        run += v;
    }
    if (threadIdx.x == 0) *out_total = tot;
}

// Phase 3: chunk-local exclusive scan plus the chunk offset.
__global__ __launch_bounds__(SCAN_THREADS) void scan_apply_kernel(const int32_t* __restrict__ in, int64_t n,
                                                                  const int32_t* __restrict__ part,
                                                                  int32_t* __restrict__ out) {
    __shared__ int32_t sh[SCAN_THREADS];
    const int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK + (int64_t)threadIdx.x * SCAN_ITEMS;
    int32_t v[SCAN_ITEMS];
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    int32_t tot;
    int32_t run = block_exclusive_scan(s, sh, &tot) + part[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

__global__ void bscatter_kernel(int64_t n, const uint32_t* __restrict__ key, const uint32_t* __restrict__ slot,
                                const int32_t* __restrict__ cell_begin, uint32_t* __restrict__ tmp) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        tmp[cell_begin[key[i] >> 6] + slot[i]] = (uint32_t)i;
}

// One wave per cell (grid-stride over the cells): the rank of each particle
// of the cell by (Morton code, previous position); every (code, position)
// pair is distinct, so the ranks are a permutation of the cell's slots.
// Cells of up to BCO_CAP particles: the rank is the number of particles with
// a smaller code (a histogram of the 64 codes and its scan) plus those of the
// same code with a smaller position (each code's particles are gathered in
// LDS and counted, a handful per code).  Larger cells (clustered boxes):
// every particle is compared with every other.
constexpr int BCO_CAP = 1024;
__global__ __launch_bounds__(256) void bcell_order_kernel(int32_t ncell, const int32_t* __restrict__ cell_begin,
                                                          const uint32_t* __restrict__ tmp,
                                                          const uint32_t* __restrict__ key,
                                                          uint32_t* __restrict__ perm) {
    __shared__ uint32_t s_fill[4][64], s_start[4][64], s_item[4][BCO_CAP];
    const int lane = threadIdx.x & 63;
    const int wl = (int)(threadIdx.x >> 6);
    uint32_t* const fill = s_fill[wl];
    uint32_t* const first = s_start[wl];
    uint32_t* const item = s_item[wl];
    // the LDS words of one wave are written and read by its own lanes
    const auto wsync = [] {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    };
    const int wave = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
    const int nwave = (int)(((int64_t)gridDim.x * blockDim.x) >> 6);
    for (int c = wave; c < ncell; c += nwave) {
        const int b = cell_begin[c];
        const int n = cell_begin[c + 1] - b;
        if (n == 1 && lane == 0) perm[b] = tmp[b];
        if (n <= 1) continue;
        if (n <= BCO_CAP) {
            fill[lane] = 0u;
            wsync();
            for (int g = lane; g < n; g += 64) atomicAdd(&fill[key[tmp[b + g]] & 63u], 1u);
            wsync();
            // exclusive scan of the code counts, lane l holding code l
            const uint32_t h = fill[lane];
            uint32_t x = h;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o);
                if (lane >= o) x += y;
            }
            wsync();
            first[lane] = x - h;
            fill[lane] = x - h;
            wsync();
            // each code's particles into its range, in any order
            for (int g = lane; g < n; g += 64) {
                const uint32_t me = tmp[b + g];
                item[atomicAdd(&fill[key[me] & 63u], 1u)] = me;
            }
            wsync();
            // fill[code] is now the end of the code's range
            for (int g = lane; g < n; g += 64) {
                const uint32_t me = tmp[b + g];
                const uint32_t code = key[me] & 63u;
                const int s0 = (int)first[code], s1 = (int)fill[code];
                int rank = s0;
                for (int k = s0; k < s1; k++) rank += item[k] < me ? 1 : 0;
                perm[b + rank] = me;
            }
            wsync();
            continue;
        }
        for (int g = 0; g < n; g += 64) {
            const bool valid = g + lane < n;
            const uint32_t me = valid ? tmp[b + g + lane] : 0u;
            const uint64_t mine = valid ? (((uint64_t)(key[me] & 63u) << 32) | me) : ~0ull;
            int rank = 0;
            for (int h = 0; h < n; h += 64) {
                const uint32_t other = h + lane < n ? tmp[b + h + lane] : 0u;
                const uint64_t oc = h + lane < n ? (((uint64_t)(key[other] & 63u) << 32) | other) : ~0ull;
                const unsigned olo = (unsigned)oc, ohi = (unsigned)(oc >> 32);
                const int cntq = min(64, n - h);
                for (int q = 0; q < cntq; q++) {
                    const uint64_t cq = ((uint64_t)(unsigned)__builtin_amdgcn_readlane((int)ohi, q) << 32) |
                                        (unsigned)__builtin_amdgcn_readlane((int)olo, q);
                    rank += cq < mine ? 1 : 0;
                }
            }
            if (valid) perm[b + rank] = me;
        }
    }
}


__global__ void btiles_count_kernel(int32_t ncell, int ts, const int32_t* __restrict__ cell_begin,
                                    int32_t* __restrict__ nt) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x)
        nt[c] = (cell_begin[c + 1] - cell_begin[c] + ts - 1) / ts;
}

// Tiles (global cell, begin, end, box) in cell order.
__global__ void btiles_write_kernel(int32_t ncell, int ts, int nbox, const tpm_box_dev* __restrict__ boxes,
                                    const int32_t* __restrict__ cell_begin, const int32_t* __restrict__ tile_off,
                                    int4* __restrict__ tiles) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
        const int box = tpm_upper_index(nbox, c, [&](int k) { return (int64_t)boxes[k].cell0; });
        const int32_t b = cell_begin[c], e = cell_begin[c + 1];
        const int32_t t0 = tile_off[c];
        for (int32_t t = 0; b + t * ts < e; t++) {
            const int32_t tb = b + t * ts;
            tiles[t0 + t] = make_int4((int)c, tb, min(tb + ts, e), box);
        }
    }
}

}  // namespace

// Exclusive scan of in[0, n) into out[0, n] (out[n] = total) with the
// three kernels above; part holds ceil(n / 2048) ints.
tpm_status tpm_scan_i32(std::string& err, hipStream_t s, const int32_t* in, int64_t n, int32_t* part,
                        int32_t* out) {
    static_assert(SCAN_CHUNK == 2048, "part size documented as n / 2048");
    const int64_t nchunk = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    if (nchunk < 1 || nchunk > INT32_MAX) {
        err = "scan: bad size";
        return TPM_ERR_INVALID;
    }
    scan_partials_kernel<<<(unsigned)nchunk, SCAN_THREADS, 0, s>>>(in, n, part);
    scan_part_kernel<<<1, SCAN_THREADS, 0, s>>>(part, nchunk, out + n);
    scan_apply_kernel<<<(unsigned)nchunk, SCAN_THREADS, 0, s>>>(in, n, part, out);
    TPM_HIPE(err, hipGetLastError());
    return TPM_OK;
}

tpm_status tpm_batch_chainmesh(tpm_batch* bt, const tpm_batch_layout& L) {
    hipStream_t s = bt->stream;
    const int32_t ncell = L.ncell;
    const int64_t n = L.n_arena;
    const int ts = bt->tune.tile_sinks;
    // counts of the live cells and of the trash cell ncell
    bfill_i32_kernel<<<grid_for(ncell + 1, 256), 256, 0, s>>>(bt->cnt, ncell + 1, 0);
    if (n > 0)
        bkey_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, L.nbox, bt->d_boxes, ncell, bt->x, bt->y, bt->z, bt->key,
                                                     bt->slot, bt->cnt);
    TPM_HIP(bt, hipGetLastError());
    TPM_TRY(tpm_scan_i32(bt->err, s, bt->cnt, ncell + 1, bt->scan_part, bt->cell_begin));
    if (n > 0) {
        bscatter_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, bt->key, bt->slot, bt->cell_begin, bt->tmp);
        const int64_t waves = std::min<int64_t>(std::max<int32_t>(ncell, 1), 65536);
        bcell_order_kernel<<<(unsigned)((waves + 3) / 4), 256, 0, s>>>(ncell, bt->cell_begin, bt->tmp, bt->key,
                                                                        bt->perm);
        TPM_HIP(bt, hipGetLastError());
    }
    if (L.n_live > 0) {
        GatherArgs g{bt->perm, bt->x,  bt->y,  bt->z,  bt->vx,  bt->vy,  bt->vz, bt->rung,
                     bt->x2,   bt->y2, bt->z2, bt->vx2, bt->vy2, bt->vz2, bt->rung2};
        gather_kernel<<<grid_for(L.n_live, 256), 256, 0, s>>>(L.n_live, g);
        TPM_HIP(bt, hipGetLastError());
    }
    std::swap(bt->x, bt->x2);
    std::swap(bt->y, bt->y2);
    std::swap(bt->z, bt->z2);
    std::swap(bt->vx, bt->vx2);
    std::swap(bt->vy, bt->vy2);
    std::swap(bt->vz, bt->vz2);
    std::swap(bt->rung, bt->rung2);
    // tiles of <= ts sinks per cell
    if (ncell > 0) {
        btiles_count_kernel<<<grid_for(ncell, 256), 256, 0, s>>>(ncell, ts, bt->cell_begin, bt->tile_cnt);
        TPM_HIP(bt, hipGetLastError());
        TPM_TRY(tpm_scan_i32(bt->err, s, bt->tile_cnt, ncell, bt->scan_part, bt->tile_off));
        btiles_write_kernel<<<grid_for(ncell, 256), 256, 0, s>>>(ncell, ts, L.nbox, bt->d_boxes, bt->cell_begin,
                                                                 bt->tile_off, bt->tiles);
    } else {
        bfill_i32_kernel<<<1, 64, 0, s>>>(bt->tile_off, 1, 0);
    }
    TPM_HIP(bt, hipGetLastError());
    return TPM_OK;
}
