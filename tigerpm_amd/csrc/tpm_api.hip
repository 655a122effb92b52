// tpm_api.hip -- the C ABI (include/tigerpm_hip.h).
#include <cmath>
#include <cstring>
#include <mutex>

#include "tpm_internal.hpp"

tpm_status tpm_fft_release(tpm_ctx* ctx);

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
                    ctx->buckets};
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
    ctx->tile = t->tile_sinks;
    ctx->deposit_block = t->deposit_block;
    return TPM_OK;
}

extern "C" tpm_status tpm_get_tuning(const tpm_ctx* ctx, tpm_tuning* t) {
    if (!ctx || !t) return TPM_ERR_INVALID;
    t->tile_sinks = ctx->tile;
    t->deposit_block = ctx->deposit_block;
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
        for (auto& c : cp)
            if (c.dst) TPM_HIP(ctx, hipMemcpyAsync(c.dst, c.src, c.bytes * n, hipMemcpyDeviceToHost, s));
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
    std::vector<uint32_t> h(ctx->n);
    if (ctx->n > 0) {
        TPM_HIP(ctx, hipMemcpyAsync(h.data(), ctx->perm, sizeof(uint32_t) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
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

extern "C" tpm_status tpm_kick(tpm_ctx* ctx, int min_rung, double scale, double t0, int first_call, unsigned flags,
                               tpm_stats* stats) {
    if (!ctx) return TPM_ERR_INVALID;
    if (flags & ~(TPM_SKIP_PP | TPM_SKIP_PM | TPM_SKIP_KICK | TPM_KICK_PME | TPM_KICK_TREE))
        return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_kick: unknown flags");
    if ((flags & TPM_KICK_PME) && (flags & TPM_KICK_TREE))
        return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_kick: TPM_KICK_PME and TPM_KICK_TREE exclude each other");
    (void)hipSetDevice(ctx->device);
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (flags & TPM_KICK_TREE)
        TPM_TRY(tpm_tree_kick(ctx, min_rung, scale, t0, first_call, flags, stats));
    else
        TPM_TRY(tpm_pp_kick(ctx, min_rung, scale, t0, first_call, flags, stats));
    if (stats) stats->ms_total = stats->ms_pp;
    return TPM_OK;
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
    TPM_TRY(tpm_pp_kick(ctx, min_rung, scale, t0, first_call, 0u, stats));
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
    return tpm_ewald_run(ctx, nsamples, samples, alpha, phi, gx, gy, gz);
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
