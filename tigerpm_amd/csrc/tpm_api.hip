// tpm_api.hip -- the C ABI (include/tigerpm_hip.h).
#include <cmath>
#include <cstring>
#include <mutex>

#include "tpm_internal.hpp"

tpm_status tpm_fft_release(tpm_ctx* ctx);

// dst[r] = src[map[r]]: device arrays in the caller's order.
__global__ static void gather_u32_kernel(int64_t n, const uint32_t* __restrict__ map, const uint32_t* __restrict__ src,
                                         uint32_t* __restrict__ dst) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        dst[r] = src[map[r]];
}
__global__ static void gather_u8_kernel(int64_t n, const uint32_t* __restrict__ map, const uint8_t* __restrict__ src,
                                        uint8_t* __restrict__ dst) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        dst[r] = src[map[r]];
}

static std::once_flag g_fft_once;
static bool g_fft_ok = false;

extern "C" tpm_status tpm_ctx_create(int device, tpm_ctx** out) {
    if (!out) return TPM_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return TPM_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return TPM_ERR_HIP;
    std::call_once(g_fft_once, []() { g_fft_ok = rocfft_setup() == rocfft_status_success; });
    if (!g_fft_ok) return TPM_ERR_FFT;
    std::string err;
    if (tpm_deposit_lds_setup(err) != TPM_OK) return TPM_ERR_HIP;
    tpm_ctx* ctx = new tpm_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return TPM_ERR_HIP;
    }
    *out = ctx;
    return TPM_OK;
}

extern "C" tpm_status tpm_ctx_destroy(tpm_ctx* ctx) {
    if (!ctx) return TPM_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    tpm_fft_release(ctx);
    void* bufs[] = {ctx->x,     ctx->y,       ctx->z,        ctx->vx,        ctx->vy,       ctx->vz,    ctx->rung,
                    ctx->gx,    ctx->gy,      ctx->gz,       ctx->pot,       ctx->x2,       ctx->y2,    ctx->z2,
                    ctx->vx2,   ctx->vy2,     ctx->vz2,      ctx->rung2,     ctx->keys,     ctx->keys2, ctx->idx,
                    ctx->perm,  ctx->sort_tmp, ctx->bounds,  ctx->cell_begin, ctx->tile_cnt, ctx->tile_end,
                    ctx->tiles, ctx->ntiles,  ctx->scan_tmp, ctx->rho_q,     ctx->rho,      ctx->rho_hat, ctx->rho_hat1,
                    ctx->phi,   ctx->green,   ctx->counters, ctx->dscratch, ctx->tree_pos, ctx->tree_tmp,
                    ctx->tree_cnt, ctx->node_begin, ctx->bucket_begin, ctx->tree_part, ctx->tree_flags, ctx->nodes,
                    ctx->buckets, ctx->h2d, ctx->h2d2, ctx->ref_perm, ctx->ref_cid, ctx->ref_cid2, ctx->ref_idx2,
                    ctx->ref_posa, ctx->ref_posb, ctx->ref_inv, ctx->ref_flag, ctx->ref_scan, ctx->ref_scan_tmp,
                    ctx->ref_tab, ctx->moment_buf, ctx->box_tiles, ctx->box_count, ctx->box_buckets, ctx->tree_moved,
                    ctx->gsink, ctx->gsrc, ctx->gmesh, ctx->mesh_force, ctx->green_t};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (hipEvent_t e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->prof_pool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TPM_OK;
}

extern "C" const char* tpm_last_error(const tpm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" tpm_status tpm_sync(tpm_ctx* ctx) {
    if (!ctx) return TPM_ERR_INVALID;
    TPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return TPM_OK;
}

// Parameter checks mirror the reference's own (options.cpp:81-85 aborts on
// an inconsistent parts/chain ratio) plus what the single-device kernels
// need: the 27-cell stencil must cover rcut (cell >= rcut) and hold 27
// distinct cells (C >= 3).
tpm_status tpm_check_params(std::string& err, tpm_params* p) {
    tpm_params& q = *p;
    auto bad = [&](const char* m) {
        err = m;
        return TPM_ERR_INVALID;
    };
    if (q.mesh_dim < 4 || q.mesh_dim > 2048) return bad("mesh_dim out of range [4, 2048]");
    if (q.order < 2 || q.order > 4) return bad("order must be 2 (CIC), 3 (TSC) or 4");
    if (q.long_type != TPM_LONG_PM && q.long_type != TPM_LONG_PME)
        return bad("long_type must be TPM_LONG_PM or TPM_LONG_PME");
    if (q.interlace != 0 && q.interlace != 1) return bad("interlace must be 0 or 1");
    if (!(q.rs > 0.0) || !(q.rcut > 0.0) || !(q.hsoft > 0.0) || !(q.rcut < 0.5))
        return bad("rs, rcut, hsoft must be positive and rcut < 1/2");
    if (!(q.eta > 0.0) || !std::isfinite(q.GM) || !(q.n_total >= 0.0))
        return bad("eta must be positive, GM finite, n_total >= 0");
    if (q.chain_dim == 0) q.chain_dim = (int)std::floor(1.0 / q.rcut);
    if (q.chain_dim < 3 || q.chain_dim > TPM_MAX_CHAIN) return bad("chain_dim must be in [3, 406]");
    if ((double)q.chain_dim * q.rcut > 1.0)
        return bad("chain cells smaller than rcut: the 27-cell stencil would miss pairs");
    return TPM_OK;
}

extern "C" tpm_status tpm_set_params(tpm_ctx* ctx, const tpm_params* p) {
    if (!ctx || !p) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    tpm_params q = *p;
    TPM_TRY(tpm_check_params(ctx->err, &q));
    if (ctx->params_set && (q.mesh_dim != ctx->p.mesh_dim || q.order != ctx->p.order || q.rs != ctx->p.rs ||
                            q.long_type != ctx->p.long_type || q.interlace != ctx->p.interlace))
        ctx->pm_valid = false;
    ctx->p = q;
    ctx->C = q.chain_dim;
    ctx->M = q.mesh_dim;
    ctx->params_set = true;
    return TPM_OK;
}

extern "C" tpm_status tpm_get_params(const tpm_ctx* ctx, tpm_params* p) {
    if (!ctx || !p || !ctx->params_set) return TPM_ERR_STATE;
    *p = ctx->p;
    return TPM_OK;
}

extern "C" tpm_status tpm_set_tuning(tpm_ctx* ctx, const tpm_tuning* t) {
    if (!ctx || !t) return TPM_ERR_INVALID;
    if (t->tile_sinks != 8 && t->tile_sinks != 16 && t->tile_sinks != 32 && t->tile_sinks != 64)
        return tpm_fail(ctx, TPM_ERR_INVALID, "tile_sinks must be 8, 16, 32 or 64");
    if (t->deposit_block < -1 || t->deposit_block > 8)
        return tpm_fail(ctx, TPM_ERR_INVALID, "deposit_block must be in [-1, 8]");
    if (t->pp_waves != 0 && t->pp_waves != 1 && t->pp_waves != 4)
        return tpm_fail(ctx, TPM_ERR_INVALID, "pp_waves must be 0, 1 or 4");
    if (t->pp_lists < -1 || t->pp_lists > 1) return tpm_fail(ctx, TPM_ERR_INVALID, "pp_lists must be -1, 0 or 1");
    if (t->pp_newton < -1 || t->pp_newton > 1) return tpm_fail(ctx, TPM_ERR_INVALID, "pp_newton must be -1, 0 or 1");
    if (t->fft < 0 || t->fft > 1) return tpm_fail(ctx, TPM_ERR_INVALID, "fft must be 0 or 1");
    ctx->tile = t->tile_sinks;
    ctx->deposit_block = t->deposit_block;
    ctx->pp_waves = t->pp_waves == 0 ? TPM_PP_WAVES_DEFAULT : t->pp_waves;
    ctx->pp_lists = t->pp_lists < 0 ? TPM_PP_LISTS_DEFAULT : t->pp_lists;
    ctx->pp_newton = t->pp_newton < 0 ? TPM_PP_NEWTON_DEFAULT : t->pp_newton;
    if (ctx->fft_rocfft != t->fft) {
        ctx->fft_rocfft = t->fft;
        ctx->pm_valid = false;
    }
    return TPM_OK;
}

extern "C" tpm_status tpm_get_tuning(const tpm_ctx* ctx, tpm_tuning* t) {
    if (!ctx || !t) return TPM_ERR_INVALID;
    t->tile_sinks = ctx->tile;
    t->deposit_block = ctx->deposit_block;
    t->pp_waves = ctx->pp_waves;
    t->pp_lists = ctx->pp_lists;
    t->pp_newton = ctx->pp_newton;
    t->fft = ctx->fft_rocfft;
    return TPM_OK;
}

extern "C" tpm_status tpm_set_host_order(tpm_ctx* ctx, int order) {
    if (!ctx) return TPM_ERR_INVALID;
    if (order != TPM_ORDER_DEVICE && order != TPM_ORDER_REFERENCE)
        return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_set_host_order: TPM_ORDER_DEVICE or TPM_ORDER_REFERENCE");
    (void)hipSetDevice(ctx->device);
    TPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // from now on the caller's order is the device's current one (trees
    // start from the caller's order: build them again)
    ctx->host_order = order;
    ctx->h2d_identity = true;
    ctx->ref_valid = false;
    ctx->tree_epoch = -1;
    return TPM_OK;
}

tpm_status tpm_ensure_particles(tpm_ctx* ctx, int64_t n) {
    if (n > (int64_t)INT32_MAX - 1024) return tpm_fail(ctx, TPM_ERR_INVALID, "particle count exceeds 2^31");
    if (n <= ctx->cap && ctx->x) return TPM_OK;
    const int64_t c = std::max<int64_t>(n, 1);
    void** u32s[] = {(void**)&ctx->x, (void**)&ctx->y, (void**)&ctx->z, (void**)&ctx->x2, (void**)&ctx->y2,
                     (void**)&ctx->z2, (void**)&ctx->keys, (void**)&ctx->keys2, (void**)&ctx->idx, (void**)&ctx->perm};
    void** f32s[] = {(void**)&ctx->vx, (void**)&ctx->vy, (void**)&ctx->vz, (void**)&ctx->vx2, (void**)&ctx->vy2,
                     (void**)&ctx->vz2, (void**)&ctx->gx, (void**)&ctx->gy, (void**)&ctx->gz, (void**)&ctx->pot};
    void** i8s[] = {(void**)&ctx->rung, (void**)&ctx->rung2};
    for (void** b : u32s) {
        if (*b) TPM_HIP(ctx, hipFree(*b));
        *b = nullptr;
        TPM_HIP(ctx, hipMalloc(b, sizeof(uint32_t) * c));
    }
    for (void** b : f32s) {
        if (*b) TPM_HIP(ctx, hipFree(*b));
        *b = nullptr;
        TPM_HIP(ctx, hipMalloc(b, sizeof(float) * c));
    }
    for (void** b : i8s) {
        if (*b) TPM_HIP(ctx, hipFree(*b));
        *b = nullptr;
        TPM_HIP(ctx, hipMalloc(b, c));
    }
    ctx->cap = c;
    return TPM_OK;
}

extern "C" tpm_status tpm_particles_upload(tpm_ctx* ctx, int64_t n, const uint32_t* x, const uint32_t* y,
                                           const uint32_t* z, const float* vx, const float* vy, const float* vz,
                                           const int8_t* rung) {
    if (!ctx || n < 0 || (n > 0 && (!x || !y || !z))) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    TPM_TRY(tpm_ensure_particles(ctx, n));
    hipStream_t s = ctx->stream;
    if (n > 0) {
        TPM_HIP(ctx, hipMemcpyAsync(ctx->x, x, 4 * n, hipMemcpyHostToDevice, s));
        TPM_HIP(ctx, hipMemcpyAsync(ctx->y, y, 4 * n, hipMemcpyHostToDevice, s));
        TPM_HIP(ctx, hipMemcpyAsync(ctx->z, z, 4 * n, hipMemcpyHostToDevice, s));
        const float* vs[3] = {vx, vy, vz};
        float* dv[3] = {ctx->vx, ctx->vy, ctx->vz};
        for (int d = 0; d < 3; d++) {
            if (vs[d])
                TPM_HIP(ctx, hipMemcpyAsync(dv[d], vs[d], 4 * n, hipMemcpyHostToDevice, s));
            else
                TPM_HIP(ctx, hipMemsetAsync(dv[d], 0, 4 * n, s));
        }
        if (rung)
            TPM_HIP(ctx, hipMemcpyAsync(ctx->rung, rung, n, hipMemcpyHostToDevice, s));
        else
            TPM_HIP(ctx, hipMemsetAsync(ctx->rung, 0, n, s));
        float* gs[4] = {ctx->gx, ctx->gy, ctx->gz, ctx->pot};
        for (float* g : gs) TPM_HIP(ctx, hipMemsetAsync(g, 0, 4 * n, s));
    }
    TPM_HIP(ctx, hipStreamSynchronize(s));
    ctx->n = n;
    ctx->chain_valid = false;
    ctx->pm_valid = false;
    ctx->h2d_identity = true;
    ctx->ref_valid = false;
    return TPM_OK;
}

extern "C" tpm_status tpm_particles_download(tpm_ctx* ctx, uint32_t* x, uint32_t* y, uint32_t* z, float* vx,
                                             float* vy, float* vz, int8_t* rung, float* gx, float* gy, float* gz,
                                             float* pot) {
    if (!ctx) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    const int64_t n = ctx->n;
    hipStream_t s = ctx->stream;
    if (n > 0) {
        struct {
            void* dst;
            const void* src;
            size_t bytes;
        } cp[] = {{x, ctx->x, 4}, {y, ctx->y, 4}, {z, ctx->z, 4},   {vx, ctx->vx, 4},   {vy, ctx->vy, 4},   {vz, ctx->vz, 4},
                  {rung, ctx->rung, 1}, {gx, ctx->gx, 4}, {gy, ctx->gy, 4}, {gz, ctx->gz, 4}, {pot, ctx->pot, 4}};
        // The caller's order differs from the device's: gather each array
        // through h2d into a staging buffer first (stream order keeps the
        // staging buffer's copy ahead of the next gather).
        const bool mapped = ctx->host_order == TPM_ORDER_REFERENCE && !ctx->h2d_identity;
        for (auto& c : cp) {
            if (!c.dst) continue;
            const void* src = c.src;
            if (mapped) {
                if (c.bytes == 4)
                    gather_u32_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, ctx->h2d, (const uint32_t*)c.src,
                                                                       ctx->ref_cid2);
                else
                    gather_u8_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, ctx->h2d, (const uint8_t*)c.src,
                                                                      (uint8_t*)ctx->ref_cid2);
                TPM_HIP(ctx, hipGetLastError());
                src = ctx->ref_cid2;
            }
            TPM_HIP(ctx, hipMemcpyAsync(c.dst, src, c.bytes * n, hipMemcpyDeviceToHost, s));
        }
    }
    TPM_HIP(ctx, hipStreamSynchronize(s));
    return TPM_OK;
}

extern "C" int64_t tpm_particles_size(const tpm_ctx* ctx) { return ctx ? ctx->n : -1; }

extern "C" tpm_status tpm_chainmesh_create(tpm_ctx* ctx) {
    if (!ctx) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    return tpm_chainmesh_build(ctx);
}

extern "C" tpm_status tpm_chainmesh_cells(tpm_ctx* ctx, int64_t* cell_begin) {
    if (!ctx || !cell_begin) return TPM_ERR_INVALID;
    if (!ctx->chain_valid) return tpm_fail(ctx, TPM_ERR_STATE, "chain mesh not built");
    (void)hipSetDevice(ctx->device);
    const int64_t nc = (int64_t)ctx->C * ctx->C * ctx->C + 1;
    std::vector<int32_t> h(nc);
    TPM_HIP(ctx, hipMemcpyAsync(h.data(), ctx->cell_begin, sizeof(int32_t) * nc, hipMemcpyDeviceToHost, ctx->stream));
    TPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t c = 0; c < nc; c++) cell_begin[c] = h[c];
    return TPM_OK;
}

extern "C" tpm_status tpm_chainmesh_get(tpm_ctx* ctx, int ix, int iy, int iz, int64_t* pbegin, int64_t* pend) {
    if (!ctx || !pbegin || !pend) return TPM_ERR_INVALID;
    if (!ctx->chain_valid) return tpm_fail(ctx, TPM_ERR_STATE, "chain mesh not built");
    (void)hipSetDevice(ctx->device);
    const int C = ctx->C;
    ix = ((ix % C) + C) % C;
    iy = ((iy % C) + C) % C;
    iz = ((iz % C) + C) % C;
    const int64_t cid = ((int64_t)iz * C + iy) * C + ix;
    int32_t h[2];
    TPM_HIP(ctx, hipMemcpyAsync(h, ctx->cell_begin + cid, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    TPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *pbegin = h[0];
    *pend = h[1];
    return TPM_OK;
}

extern "C" tpm_status tpm_chainmesh_perm(tpm_ctx* ctx, int64_t* perm) {
    if (!ctx || !perm) return TPM_ERR_INVALID;
    if (!ctx->chain_valid) return tpm_fail(ctx, TPM_ERR_STATE, "chain mesh not built");
    (void)hipSetDevice(ctx->device);
    const bool ref = ctx->host_order == TPM_ORDER_REFERENCE;
    if (ref && !ctx->ref_valid)
        return tpm_fail(ctx, TPM_ERR_STATE, "tpm_chainmesh_perm: the chain mesh was built before TPM_ORDER_REFERENCE");
    std::vector<uint32_t> h(ctx->n);
    if (ctx->n > 0) {
        TPM_HIP(ctx, hipMemcpyAsync(h.data(), ref ? ctx->ref_perm : ctx->perm, sizeof(uint32_t) * ctx->n,
                                    hipMemcpyDeviceToHost, ctx->stream));
        TPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    for (int64_t i = 0; i < ctx->n; i++) perm[i] = h[i];
    return TPM_OK;
}

extern "C" tpm_status tpm_gravity_long_compute(tpm_ctx* ctx) {
    if (!ctx) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    return tpm_pm_solve(ctx);
}

static tpm_status copy_mesh(tpm_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx->pm_valid) return tpm_fail(ctx, TPM_ERR_STATE, "PM mesh not computed");
    TPM_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return TPM_OK;
}

extern "C" tpm_status tpm_gravity_long_get_phi(tpm_ctx* ctx, float* phi) {
    if (!ctx || !phi) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    const int64_t M = ctx->M;
    return copy_mesh(ctx, phi, ctx->phi, sizeof(float) * M * M * M);
}

extern "C" tpm_status tpm_gravity_long_get_density(tpm_ctx* ctx, uint64_t* grid) {
    if (!ctx || !grid) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    const int64_t M = ctx->M;
    return copy_mesh(ctx, grid, ctx->rho_q, sizeof(uint64_t) * M * M * M);
}

extern "C" tpm_status tpm_gravity_long_get_green(tpm_ctx* ctx, float* table) {
    if (!ctx || !table) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    const int64_t M = ctx->M;
    return copy_mesh(ctx, table, ctx->green, sizeof(float) * M * M * (M / 2 + 1));
}

// One 3-D transform through the context's mesh buffers: real data in rho,
// spectra in rho_hat (fft.cpp:51-84).
static tpm_status fft3d_host(tpm_ctx* ctx, int M, bool inverse, const float* in, float* out) {
    if (!ctx || !in || !out || M < 4 || M > 2048) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    ctx->pm_valid = false;
    TPM_TRY(tpm_ensure_mesh(ctx, M));
    const int64_t M3 = (int64_t)M * M * M, MC = (int64_t)M * M * (M / 2 + 1);
    hipStream_t s = ctx->stream;
    if (!inverse) {
        TPM_HIP(ctx, hipMemcpyAsync(ctx->rho, in, sizeof(float) * M3, hipMemcpyHostToDevice, s));
        TPM_TRY(tpm_fft_forward(ctx, ctx->rho, ctx->rho_hat));
        TPM_HIP(ctx, hipMemcpyAsync(out, ctx->rho_hat, sizeof(float2) * MC, hipMemcpyDeviceToHost, s));
    } else {
        TPM_HIP(ctx, hipMemcpyAsync(ctx->rho_hat, in, sizeof(float2) * MC, hipMemcpyHostToDevice, s));
        TPM_TRY(tpm_fft_inverse(ctx, ctx->rho_hat, ctx->rho));
        TPM_HIP(ctx, hipMemcpyAsync(out, ctx->rho, sizeof(float) * M3, hipMemcpyDeviceToHost, s));
    }
    TPM_HIP(ctx, hipStreamSynchronize(s));
    return TPM_OK;
}

extern "C" tpm_status tpm_fft3d_r2c(tpm_ctx* ctx, int M, const float* in, float* out) {
    return fft3d_host(ctx, M, false, in, out);
}

extern "C" tpm_status tpm_fft3d_c2r(tpm_ctx* ctx, int M, const float* in, float* out) {
    return fft3d_host(ctx, M, true, in, out);
}

// ---------------------------------------------------------------- box kicks
namespace {

struct BoxCost {
    int64_t bytes = 0;   // working set of one launch over the box
    int64_t tiles = 0;   // sink tiles of the box's cells
};

// The working set of a kick of the sinks in box (the bytes its launch
// touches; the terms of mem_requirements, kick_pme.cu:110-124 and
// kick_treepm.cu:140-157, with this build's layout), from the host copy of
// the cell table and, on the tree path, the node / bucket offsets.
BoxCost box_cost(const tpm_ctx* ctx, const tpm_cell_box& b, const std::vector<int32_t>& cb,
                 const std::vector<int32_t>* nodes, const std::vector<int32_t>* buckets) {
    const int C = ctx->C, M = ctx->M, ts = ctx->tile;
    BoxCost r;
    int64_t sinks = 0, sources = 0, pad_cells = 0, tree_nodes = 0, tree_buckets = 0;
    auto cid = [C](int x, int y, int z) { return ((int64_t)z * C + y) * C + x; };
    for (int z = b.lo[2]; z < b.hi[2]; z++)
        for (int y = b.lo[1]; y < b.hi[1]; y++)
            for (int x = b.lo[0]; x < b.hi[0]; x++) {
                const int64_t c = cid(x, y, z);
                const int64_t n = cb[c + 1] - cb[c];
                sinks += n;
                r.tiles += (n + ts - 1) / ts;
                if (buckets) tree_buckets += (*buckets)[c + 1] - (*buckets)[c];
            }
    // the box padded by one cell, periodic, each cell once
    std::vector<int> ax[3];
    for (int d = 0; d < 3; d++) {
        const int span = b.hi[d] - b.lo[d];
        if (span + 2 >= C) {
            for (int i = 0; i < C; i++) ax[d].push_back(i);
        } else {
            for (int i = b.lo[d] - 1; i <= b.hi[d]; i++) ax[d].push_back((i + C) % C);
        }
    }
    for (int z : ax[2])
        for (int y : ax[1])
            for (int x : ax[0]) {
                const int64_t c = cid(x, y, z);
                sources += cb[c + 1] - cb[c];
                pad_cells++;
                if (nodes) tree_nodes += (*nodes)[c + 1] - (*nodes)[c];
            }
    // potential mesh nodes under the padded box: the 6-point stencil reaches
    // 3 nodes beyond a sink's cell (PHI_BW, tigerpm.hpp:41)
    int64_t phivol = 1;
    for (int d = 0; d < 3; d++) {
        const int64_t w = ((int64_t)(b.hi[d] - b.lo[d]) * M + C - 1) / C + 1 + 6;
        phivol *= std::min<int64_t>(w, M);
    }
    r.bytes = 12 * sources + 41 * sinks + 16 * r.tiles + 4 * phivol + 8 * pad_cells + 48 * tree_nodes +
              (int64_t)sizeof(tpm_tree_bucket) * tree_buckets;
    return r;
}

struct BoxKick {
    tpm_ctx* ctx;
    int min_rung;
    double scale, t0;
    int first_call;
    unsigned flags;
    tpm_stats* acc;  // accumulated statistics, or nullptr
    std::vector<int32_t> cb, nodes, buckets;
};

tpm_status box_kick_rec(BoxKick& K, const tpm_cell_box& b) {
    tpm_ctx* ctx = K.ctx;
    const bool tree = (K.flags & TPM_KICK_TREE) != 0;
    const BoxCost cost = box_cost(ctx, b, K.cb, tree ? &K.nodes : nullptr, tree ? &K.buckets : nullptr);
    if (ctx->budget > 0 && cost.bytes > ctx->budget) {
        const int vol = (b.hi[0] - b.lo[0]) * (b.hi[1] - b.lo[1]) * (b.hi[2] - b.lo[2]);
        if (vol <= 1)
            return tpm_fail(ctx, TPM_ERR_BUDGET,
                            "kick: one chain cell needs " + std::to_string(cost.bytes) + " bytes, budget " +
                                std::to_string(ctx->budget));
        // range::split (range.hpp:118-133): the first longest axis, at its middle
        int dim = 0, span = 0;
        for (int d = 0; d < 3; d++)
            if (b.hi[d] - b.lo[d] > span) {
                span = b.hi[d] - b.lo[d];
                dim = d;
            }
        const int mid = (b.hi[dim] + b.lo[dim]) / 2;
        tpm_cell_box l = b, r = b;
        l.hi[dim] = r.lo[dim] = mid;
        TPM_TRY(box_kick_rec(K, l));
        return box_kick_rec(K, r);
    }
    tpm_stats st{};
    tpm_stats* sp = K.acc ? &st : nullptr;
    if (tree)
        TPM_TRY(tpm_tree_kick(ctx, K.min_rung, K.scale, K.t0, K.first_call, K.flags, sp, &b));
    else
        TPM_TRY(tpm_pp_kick(ctx, K.min_rung, K.scale, K.t0, K.first_call, K.flags, sp, &b, cost.tiles));
    if (K.acc) {
        K.acc->candidate_pairs += st.candidate_pairs;
        K.acc->executed_pairs += st.executed_pairs;
        K.acc->pc_interactions += st.pc_interactions;
        K.acc->ms_pp += st.ms_pp;
        K.acc->n_tiles += st.n_tiles;
        K.acc->max_rung = std::max(K.acc->max_rung, st.max_rung);
    }
    return TPM_OK;
}

}  // namespace

tpm_status tpm_kick_any(tpm_ctx* ctx, const tpm_cell_box* box, int min_rung, double scale, double t0, int first_call,
                        unsigned flags, tpm_stats* stats) {
    const bool tree = (flags & TPM_KICK_TREE) != 0;
    if (!box && ctx->budget <= 0) {
        // the whole box in one launch
        if (tree) return tpm_tree_kick(ctx, min_rung, scale, t0, first_call, flags, stats);
        return tpm_pp_kick(ctx, min_rung, scale, t0, first_call, flags, stats);
    }
    if (!ctx->chain_valid || ctx->chain_C != ctx->C || (!tree && ctx->chain_tile != ctx->tile))
        return tpm_fail(ctx, TPM_ERR_STATE, "tpm_kick: chain mesh not built (call tpm_chainmesh_create)");
    const int C = ctx->C;
    tpm_cell_box b{{0, 0, 0}, {C, C, C}};
    if (box) {
        b = *box;
        for (int d = 0; d < 3; d++)
            if (b.lo[d] < 0 || b.hi[d] > C || b.lo[d] >= b.hi[d])
                return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_kick_box: need 0 <= lo < hi <= chain_dim on every axis");
    }
    if (tree) TPM_TRY(tpm_tree_ensure(ctx));
    BoxKick K{ctx, min_rung, scale, t0, first_call, flags, stats, {}, {}, {}};
    const int64_t ncell = (int64_t)C * C * C;
    hipStream_t s = ctx->stream;
    if (ctx->h_cells_epoch != ctx->chain_epoch) {
        ctx->h_cells.resize((size_t)ncell + 1);
        TPM_HIP(ctx, hipMemcpyAsync(ctx->h_cells.data(), ctx->cell_begin, sizeof(int32_t) * (ncell + 1),
                                    hipMemcpyDeviceToHost, s));
        TPM_HIP(ctx, hipStreamSynchronize(s));
        ctx->h_cells_epoch = ctx->chain_epoch;
    }
    K.cb = ctx->h_cells;
    if (tree) {
        K.nodes.resize((size_t)ncell + 1);
        K.buckets.resize((size_t)ncell + 1);
        TPM_HIP(ctx, hipMemcpyAsync(K.nodes.data(), ctx->node_begin, sizeof(int32_t) * (ncell + 1),
                                    hipMemcpyDeviceToHost, s));
        TPM_HIP(ctx, hipMemcpyAsync(K.buckets.data(), ctx->bucket_begin, sizeof(int32_t) * (ncell + 1),
                                    hipMemcpyDeviceToHost, s));
        TPM_HIP(ctx, hipStreamSynchronize(s));
    }
    return box_kick_rec(K, b);
}

extern "C" tpm_status tpm_set_budget(tpm_ctx* ctx, int64_t bytes) {
    if (!ctx || bytes < 0) return TPM_ERR_INVALID;
    ctx->budget = bytes;
    return TPM_OK;
}

static tpm_status check_kick_flags(tpm_ctx* ctx, unsigned flags) {
    if (flags & ~(TPM_SKIP_PP | TPM_SKIP_PM | TPM_SKIP_KICK | TPM_KICK_PME | TPM_KICK_TREE))
        return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_kick: unknown flags");
    if ((flags & TPM_KICK_PME) && (flags & TPM_KICK_TREE))
        return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_kick: TPM_KICK_PME and TPM_KICK_TREE exclude each other");
    return TPM_OK;
}

extern "C" tpm_status tpm_kick(tpm_ctx* ctx, int min_rung, double scale, double t0, int first_call, unsigned flags,
                               tpm_stats* stats) {
    if (!ctx) return TPM_ERR_INVALID;
    TPM_TRY(check_kick_flags(ctx, flags));
    (void)hipSetDevice(ctx->device);
    if (stats) std::memset(stats, 0, sizeof(*stats));
    TPM_TRY(tpm_kick_any(ctx, nullptr, min_rung, scale, t0, first_call, flags, stats));
    if (stats) {
        stats->ms_total = stats->ms_pp;
        TPM_TRY(tpm_ctx_moments(ctx, &stats->moments));
    }
    return TPM_OK;
}

extern "C" tpm_status tpm_kick_box(tpm_ctx* ctx, const int32_t lo[3], const int32_t hi[3], int min_rung, double scale,
                                   double t0, int first_call, unsigned flags, tpm_stats* stats) {
    if (!ctx || !lo || !hi) return TPM_ERR_INVALID;
    TPM_TRY(check_kick_flags(ctx, flags));
    (void)hipSetDevice(ctx->device);
    if (stats) std::memset(stats, 0, sizeof(*stats));
    const tpm_cell_box b{{lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}};
    TPM_TRY(tpm_kick_any(ctx, &b, min_rung, scale, t0, first_call, flags, stats));
    if (stats) {
        stats->ms_total = stats->ms_pp;
        TPM_TRY(tpm_ctx_moments(ctx, &stats->moments));
    }
    return TPM_OK;
}

extern "C" tpm_status tpm_moments_get(tpm_ctx* ctx, tpm_moments* out) {
    if (!ctx || !out) return TPM_ERR_INVALID;
    if (!ctx->params_set) return tpm_fail(ctx, TPM_ERR_STATE, "tpm_moments_get: parameters not set (GM)");
    (void)hipSetDevice(ctx->device);
    return tpm_ctx_moments(ctx, out);
}

extern "C" tpm_status tpm_kick_treepm(tpm_ctx* ctx, int min_rung, double scale, double t0, int first_call,
                                      tpm_stats* stats) {
    if (!ctx) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (!ctx->ev[0]) {
        for (int e = 0; e < 8; e++) TPM_HIP(ctx, hipEventCreate(&ctx->ev[e]));
    }
    hipStream_t s = ctx->stream;
    std::array<hipEvent_t, 4> ev = {ctx->ev[0], ctx->ev[1], ctx->ev[4], ctx->ev[5]};
    if (ctx->prof_on) {
        for (auto& e : ev) {
            e = tpm_prof_event(ctx);
            if (!e) return tpm_fail(ctx, TPM_ERR_HIP, "hipEventCreate failed");
        }
        ctx->prof_step.push_back(ev);
    }
    TPM_HIP(ctx, hipEventRecord(ev[0], s));
    TPM_TRY(tpm_chainmesh_build(ctx));
    TPM_HIP(ctx, hipEventRecord(ev[1], s));
    TPM_TRY(tpm_pm_solve(ctx));
    TPM_HIP(ctx, hipEventRecord(ev[2], s));
    TPM_TRY(tpm_kick_any(ctx, nullptr, min_rung, scale, t0, first_call, 0u, stats));
    TPM_HIP(ctx, hipEventRecord(ev[3], s));
    if (stats) {
        TPM_HIP(ctx, hipEventSynchronize(ev[3]));
        float a = 0.f, b = 0.f, c = 0.f;
        TPM_HIP(ctx, hipEventElapsedTime(&a, ev[0], ev[1]));
        TPM_HIP(ctx, hipEventElapsedTime(&b, ev[1], ev[2]));
        TPM_HIP(ctx, hipEventElapsedTime(&c, ev[0], ev[3]));
        stats->ms_chainmesh = a;
        stats->ms_pm = b;
        stats->ms_total = c;
        TPM_TRY(tpm_ctx_moments(ctx, &stats->moments));
    }
    return TPM_OK;
}

extern "C" tpm_status tpm_profile(tpm_ctx* ctx, int enable) {
    if (!ctx) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    if (enable) {
        TPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->prof_used = 0;
        ctx->prof_pp.clear();
        ctx->prof_step.clear();
    }
    ctx->prof_on = enable != 0;
    return TPM_OK;
}

extern "C" tpm_status tpm_profile_read(tpm_ctx* ctx, double* ms, int64_t* pp_launches, int64_t* steps) {
    if (!ctx || !ms) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    TPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < 4; k++) ms[k] = 0.0;
    for (auto& pr : ctx->prof_pp) {
        float t = 0.f;
        TPM_HIP(ctx, hipEventElapsedTime(&t, pr.first, pr.second));
        ms[2] += t;
    }
    for (auto& ev : ctx->prof_step) {
        float a = 0.f, b = 0.f, c = 0.f;
        TPM_HIP(ctx, hipEventElapsedTime(&a, ev[0], ev[1]));
        TPM_HIP(ctx, hipEventElapsedTime(&b, ev[1], ev[2]));
        TPM_HIP(ctx, hipEventElapsedTime(&c, ev[0], ev[3]));
        ms[0] += a;
        ms[1] += b;
        ms[3] += c;
    }
    if (pp_launches) *pp_launches = (int64_t)ctx->prof_pp.size();
    if (steps) *steps = (int64_t)ctx->prof_step.size();
    return TPM_OK;
}

extern "C" tpm_status tpm_ewald(tpm_ctx* ctx, int64_t nsamples, const int64_t* samples, double alpha, double* phi,
                                double* gx, double* gy, double* gz) {
    if (!ctx || nsamples < 0 || (nsamples > 0 && (!samples || !phi || !gx || !gy || !gz))) return TPM_ERR_INVALID;
    if (!ctx->params_set) return tpm_fail(ctx, TPM_ERR_STATE, "tpm_ewald: parameters not set (hsoft)");
    (void)hipSetDevice(ctx->device);
    return tpm_ewald_run(ctx, nsamples, samples, nullptr, alpha, phi, gx, gy, gz);
}

extern "C" tpm_status tpm_ewald_at(tpm_ctx* ctx, int64_t nsamples, const uint32_t* sx, const uint32_t* sy,
                                   const uint32_t* sz, double alpha, double* phi, double* gx, double* gy, double* gz) {
    if (!ctx || nsamples < 0 || (nsamples > 0 && (!sx || !sy || !sz || !phi || !gx || !gy || !gz)))
        return TPM_ERR_INVALID;
    if (!ctx->params_set) return tpm_fail(ctx, TPM_ERR_STATE, "tpm_ewald_at: parameters not set (hsoft)");
    (void)hipSetDevice(ctx->device);
    const uint32_t* pos[3] = {sx, sy, sz};
    return tpm_ewald_run(ctx, nsamples, nullptr, pos, alpha, phi, gx, gy, gz);
}

extern "C" tpm_status tpm_ewald_at_noalloc(tpm_ctx* ctx, int64_t nsamples, const uint32_t* d_sx, const uint32_t* d_sy,
                                           const uint32_t* d_sz, double alpha, double* d_phi, double* d_gx,
                                           double* d_gy, double* d_gz) {
    if (!ctx || nsamples < 0 || (nsamples > 0 && (!d_sx || !d_sy || !d_sz || !d_phi || !d_gx || !d_gy || !d_gz)))
        return TPM_ERR_INVALID;
    if (!ctx->params_set) return tpm_fail(ctx, TPM_ERR_STATE, "tpm_ewald_at_noalloc: parameters not set (hsoft)");
    TPM_HIP(ctx, hipSetDevice(ctx->device));
    return tpm_ewald_noalloc_run(ctx, nsamples, d_sx, d_sy, d_sz, alpha, d_phi, d_gx, d_gy, d_gz);
}

extern "C" tpm_status tpm_ic_zeldovich(tpm_ctx* ctx, int n_side, uint64_t seed, double rms_disp,
                                       double spectral_index) {
    if (!ctx) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    return tpm_ic_zeldovich_run(ctx, n_side, seed, rms_disp, spectral_index);
}

extern "C" tpm_status tpm_ic_uniform(tpm_ctx* ctx, int64_t n, uint64_t seed) {
    if (!ctx || n < 0) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    return tpm_ic_uniform_run(ctx, n, seed);
}
