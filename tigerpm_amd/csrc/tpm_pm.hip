// tpm_pm.hip -- long-range PM solve on one device.
//
// Reference: gravity_long_compute (gravity_long.cpp:36-48): compute_source
// (cloud4 deposit in float under per-plane spinlocks, :142-199), distributed
// FFTW r2c (fft.cpp:51-67), apply_laplacian (:201-245), inverse FFT
// normalised by 1/N^3 (fft.cpp:468, 512).  Here the deposit accumulates
// integer Q32 weights with 64-bit atomics (order-independent, so the grid is
// bit-exact and deterministic), the 3-D transforms run in rocFFT, and the
// deposit factor 4 pi M, the PME filter, the assignment deconvolution and the
// 1/M^3 normalisation live in one FP32 Green's table built in FP64.
#include "tpm_internal.hpp"

namespace {

__global__ void fill_u64_kernel(unsigned long long* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

__device__ __forceinline__ int wrapi(int i, int M) {
    i += (i < 0) ? M : 0;
    i -= (i >= M) ? M : 0;
    return i;
}

// One thread per particle: 1-D integer weights per axis, 3-D products in
// Q32 with the truncation deficit added to the centre node, 64-bit atomics.
template <int ORDER>
__global__ __launch_bounds__(256) void deposit_kernel(int64_t n, const uint32_t* __restrict__ x,
                                                      const uint32_t* __restrict__ y,
                                                      const uint32_t* __restrict__ z, int M,
                                                      unsigned long long* __restrict__ grid) {
    constexpr int NP = ORDER;
    const int64_t M2 = (int64_t)M * M;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t wx[4], wy[4], wz[4];
        int bx, by, bz, q;
        tpm_weights_1d(ORDER, x[i], M, &bx, wx, &q);
        tpm_weights_1d(ORDER, y[i], M, &by, wy, &q);
        tpm_weights_1d(ORDER, z[i], M, &bz, wz, &q);
        const int shift = 3 * q - 32;
        uint64_t w3[NP * NP * NP];
        uint64_t sum = 0;
#pragma unroll
        for (int a = 0; a < NP; a++)
#pragma unroll
            for (int b = 0; b < NP; b++)
#pragma unroll
                for (int c = 0; c < NP; c++) {
                    const uint64_t v = (wx[a] * wy[b] * wz[c]) >> shift;
                    w3[(a * NP + b) * NP + c] = v;
                    sum += v;
                }
        constexpr int centre = (ORDER == 2) ? 0 : ((1 * NP + 1) * NP + 1);
        w3[centre] += (1ull << 32) - sum;
#pragma unroll
        for (int a = 0; a < NP; a++) {
            const int64_t ix = wrapi(bx + a, M);
#pragma unroll
            for (int b = 0; b < NP; b++) {
                const int64_t iy = wrapi(by + b, M);
#pragma unroll
                for (int c = 0; c < NP; c++) {
                    const int64_t iz = wrapi(bz + c, M);
                    atomicAdd(&grid[ix * M2 + iy * M + iz], (unsigned long long)w3[(a * NP + b) * NP + c]);
                }
            }
        }
    }
}

// Deposit of a block of B^3 chain cells per workgroup (particles already in
// chain-mesh order): the block's mesh footprint lives in LDS, weights are
// added there with 64-bit LDS atomics and each nonzero node is flushed with
// one global atomic, instead of ORDER^3 global atomics per particle.  The
// integers added are the same, so the grid is the same bit for bit.
struct BlockedDepositArgs {
    const uint32_t* __restrict__ x;
    const uint32_t* __restrict__ y;
    const uint32_t* __restrict__ z;
    const int32_t* __restrict__ cell_begin;
    const uint32_t* __restrict__ bounds;  // C split planes
    unsigned long long* __restrict__ grid;
    int C, M, B, nb;                      // nb = ceil(C / B) blocks per axis
    int wmax;                             // LDS footprint capacity per axis
};

// Mesh nodes [lo, lo + w) touched by particles of chain cells [c0, c1].
__device__ __forceinline__ void block_nodes(const uint32_t* b, int C, int M, int c0, int c1, int* lo, int* w) {
    const uint64_t xlo = b[c0];
    const uint64_t xhi = (c1 + 1 < C ? (uint64_t)b[c1 + 1] : (1ull << 32)) - 1;
    const int l = (int)((xlo * (uint64_t)M) >> 32) - 1;
    const int h = (int)((xhi * (uint64_t)M) >> 32) + 2;
    *lo = l;
    *w = h - l + 1;
}

template <int ORDER>
__global__ __launch_bounds__(256) void deposit_blocked_kernel(BlockedDepositArgs a) {
    extern __shared__ unsigned long long sg[];
    constexpr int NP = ORDER;
    const int C = a.C, M = a.M, B = a.B, nb = a.nb;
    const int bid = blockIdx.x;
    const int bx = bid % nb, by = (bid / nb) % nb, bz = bid / (nb * nb);
    const int c0x = bx * B, c0y = by * B, c0z = bz * B;
    const int c1x = min(c0x + B, C) - 1, c1y = min(c0y + B, C) - 1, c1z = min(c0z + B, C) - 1;
    int lox, loy, loz, wx, wy, wz;
    block_nodes(a.bounds, C, M, c0x, c1x, &lox, &wx);
    block_nodes(a.bounds, C, M, c0y, c1y, &loy, &wy);
    block_nodes(a.bounds, C, M, c0z, c1z, &loz, &wz);
    const int wtot = wx * wy * wz;
    for (int l = threadIdx.x; l < wtot; l += blockDim.x) sg[l] = 0ull;
    __syncthreads();
    for (int cz = c0z; cz <= c1z; cz++) {
        for (int cy = c0y; cy <= c1y; cy++) {
            // cells c0x..c1x of one (cy, cz) row are contiguous in memory
            const int rb = a.cell_begin[(cz * C + cy) * C + c0x];
            const int re = a.cell_begin[(cz * C + cy) * C + c1x + 1];
            for (int i = rb + threadIdx.x; i < re; i += blockDim.x) {
                uint64_t wxv[4], wyv[4], wzv[4];
                int bxn, byn, bzn, q;
                tpm_weights_1d(ORDER, a.x[i], M, &bxn, wxv, &q);
                tpm_weights_1d(ORDER, a.y[i], M, &byn, wyv, &q);
                tpm_weights_1d(ORDER, a.z[i], M, &bzn, wzv, &q);
                const int shift = 3 * q - 32;
                uint64_t w3[NP * NP * NP];
                uint64_t sum = 0;
#pragma unroll
                for (int p = 0; p < NP; p++)
#pragma unroll
                    for (int r = 0; r < NP; r++)
#pragma unroll
                        for (int s = 0; s < NP; s++) {
                            const uint64_t v = (wxv[p] * wyv[r] * wzv[s]) >> shift;
                            w3[(p * NP + r) * NP + s] = v;
                            sum += v;
                        }
                constexpr int centre = (ORDER == 2) ? 0 : ((1 * NP + 1) * NP + 1);
                w3[centre] += (1ull << 32) - sum;
                const int ox = bxn - lox, oy = byn - loy, oz = bzn - loz;
#pragma unroll
                for (int p = 0; p < NP; p++)
#pragma unroll
                    for (int r = 0; r < NP; r++)
#pragma unroll
                        for (int s = 0; s < NP; s++)
                            atomicAdd(&sg[((ox + p) * wy + (oy + r)) * wz + (oz + s)],
                                      (unsigned long long)w3[(p * NP + r) * NP + s]);
            }
        }
    }
    __syncthreads();
    const int64_t M2 = (int64_t)M * M;
    for (int l = threadIdx.x; l < wtot; l += blockDim.x) {
        const unsigned long long v = sg[l];
        if (v == 0ull) continue;
        int ix = lox + l / (wy * wz), iy = loy + (l / wz) % wy, iz = loz + l % wz;
        ix = wrapi(ix, M);
        iy = wrapi(iy, M);
        iz = wrapi(iz, M);
        atomicAdd(&a.grid[ix * M2 + (int64_t)iy * M + iz], v);
    }
}

// Unit masses per node: float(double(sum) * 2^-32).
__global__ void q32_to_float_kernel(int64_t n, const unsigned long long* __restrict__ q, float* __restrict__ rho) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        rho[i] = (float)((double)q[i] * 0x1p-32);
}

__device__ __forceinline__ double sincd(double x) { return x != 0.0 ? sin(x) / x : 1.0; }

// G(k) = -4 pi M / M^3 / k^2 exp(-k^2 (rs M)^2) prod sinc(k_i / 2)^-p, in FP64
// (apply_laplacian, gravity_long.cpp:213-242), k in radians per cell.
__global__ void green_kernel(int M, double rs, int order, float* __restrict__ table) {
    const int Mh = M / 2 + 1;
    const int64_t total = (int64_t)M * M * Mh;
    const double c0 = 2.0 * M_PI / M;
    const double rs2 = (rs * M) * (rs * M);
    const double scale = 4.0 * M_PI * M / ((double)M * M * M);
    for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < total; l += (int64_t)gridDim.x * blockDim.x) {
        const int64_t iz = l % Mh;
        const int64_t iy = (l / Mh) % M;
        const int64_t ix = l / ((int64_t)Mh * M);
        const double kx = c0 * (double)(2 * ix < M ? ix : ix - M);
        const double ky = c0 * (double)(2 * iy < M ? iy : iy - M);
        const double kz = c0 * (double)(2 * iz < M ? iz : iz - M);
        const double k2 = kx * kx + ky * ky + kz * kz;
        double g = 0.0;
        if (k2 > 0.0) {
            const double s = sincd(0.5 * kx) * sincd(0.5 * ky) * sincd(0.5 * kz);
            g = -scale / k2 * exp(-k2 * rs2) * pow(s, -(double)order);
        }
        table[l] = (float)g;
    }
}

__global__ void green_apply_kernel(int64_t n, float2* __restrict__ c, const float* __restrict__ g) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gg = g[i];
        float2 v = c[i];
        v.x *= gg;
        v.y *= gg;
        c[i] = v;
    }
}

}  // namespace

tpm_status tpm_fill_u64(tpm_ctx* ctx, unsigned long long* p, int64_t n) {
    fill_u64_kernel<<<grid_for(n, 256), 256, 0, ctx->stream>>>(p, n);
    TPM_HIP(ctx, hipGetLastError());
    return TPM_OK;
}

static tpm_status fft_destroy(tpm_ctx* ctx) {
    tpm_fft& f = ctx->fft;
    if (f.fwd) rocfft_plan_destroy(f.fwd);
    if (f.inv) rocfft_plan_destroy(f.inv);
    if (f.info_fwd) rocfft_execution_info_destroy(f.info_fwd);
    if (f.info_inv) rocfft_execution_info_destroy(f.info_inv);
    if (f.work) TPM_HIP(ctx, hipFree(f.work));
    f = tpm_fft{};
    return TPM_OK;
}

tpm_status tpm_fft_plan(tpm_ctx* ctx, int M) {
    tpm_fft& f = ctx->fft;
    if (f.M == M && f.fwd) return TPM_OK;
    TPM_TRY(fft_destroy(ctx));
    const size_t len[3] = {(size_t)M, (size_t)M, (size_t)M};  // fastest first; cubic
    if (rocfft_plan_create(&f.fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                           rocfft_precision_single, 3, len, 1, nullptr) != rocfft_status_success)
        return tpm_fail(ctx, TPM_ERR_FFT, "rocfft_plan_create (forward) failed");
    if (rocfft_plan_create(&f.inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                           rocfft_precision_single, 3, len, 1, nullptr) != rocfft_status_success)
        return tpm_fail(ctx, TPM_ERR_FFT, "rocfft_plan_create (inverse) failed");
    size_t w1 = 0, w2 = 0;
    rocfft_plan_get_work_buffer_size(f.fwd, &w1);
    rocfft_plan_get_work_buffer_size(f.inv, &w2);
    f.work_bytes = std::max(w1, w2);
    if (f.work_bytes) TPM_HIP(ctx, hipMalloc(&f.work, f.work_bytes));
    rocfft_execution_info_create(&f.info_fwd);
    rocfft_execution_info_create(&f.info_inv);
    rocfft_execution_info_set_stream(f.info_fwd, ctx->stream);
    rocfft_execution_info_set_stream(f.info_inv, ctx->stream);
    if (f.work_bytes) {
        rocfft_execution_info_set_work_buffer(f.info_fwd, f.work, f.work_bytes);
        rocfft_execution_info_set_work_buffer(f.info_inv, f.work, f.work_bytes);
    }
    f.M = M;
    return TPM_OK;
}

tpm_status tpm_fft_forward(tpm_ctx* ctx, float* in, float2* out) {
    void* ip[1] = {in};
    void* op[1] = {out};
    if (rocfft_execute(ctx->fft.fwd, ip, op, ctx->fft.info_fwd) != rocfft_status_success)
        return tpm_fail(ctx, TPM_ERR_FFT, "rocfft_execute (forward) failed");
    return TPM_OK;
}

tpm_status tpm_fft_inverse(tpm_ctx* ctx, float2* in, float* out) {
    void* ip[1] = {in};
    void* op[1] = {out};
    if (rocfft_execute(ctx->fft.inv, ip, op, ctx->fft.info_inv) != rocfft_status_success)
        return tpm_fail(ctx, TPM_ERR_FFT, "rocfft_execute (inverse) failed");
    return TPM_OK;
}

tpm_status tpm_ensure_mesh(tpm_ctx* ctx, int M) {
    const int64_t M3 = (int64_t)M * M * M;
    const int64_t MC = (int64_t)M * M * (M / 2 + 1);
    if (M3 > ctx->mesh_cap) {
        if (ctx->rho_q) TPM_HIP(ctx, hipFree(ctx->rho_q));
        if (ctx->rho) TPM_HIP(ctx, hipFree(ctx->rho));
        if (ctx->rho_hat) TPM_HIP(ctx, hipFree(ctx->rho_hat));
        if (ctx->phi) TPM_HIP(ctx, hipFree(ctx->phi));
        if (ctx->green) TPM_HIP(ctx, hipFree(ctx->green));
        TPM_HIP(ctx, hipMalloc(&ctx->rho_q, sizeof(unsigned long long) * M3));
        TPM_HIP(ctx, hipMalloc(&ctx->rho, sizeof(float) * M3));
        TPM_HIP(ctx, hipMalloc(&ctx->rho_hat, sizeof(float2) * MC));
        TPM_HIP(ctx, hipMalloc(&ctx->phi, sizeof(float) * M3));
        TPM_HIP(ctx, hipMalloc(&ctx->green, sizeof(float) * MC));
        ctx->mesh_cap = M3;
        ctx->green_M = 0;
    }
    return tpm_fft_plan(ctx, M);
}

tpm_status tpm_pm_solve(tpm_ctx* ctx) {
    if (!ctx->params_set) return tpm_fail(ctx, TPM_ERR_STATE, "tpm_gravity_long_compute: parameters not set");
    const int M = ctx->M;
    const int64_t M3 = (int64_t)M * M * M;
    const int64_t MC = (int64_t)M * M * (M / 2 + 1);
    hipStream_t s = ctx->stream;
    TPM_TRY(tpm_ensure_mesh(ctx, M));
    // GRAVITY_LONG_PM has no Gaussian filter: the table with rs = 0.
    const double rs_eff = ctx->p.long_type == TPM_LONG_PM ? 0.0 : ctx->p.rs;
    if (ctx->green_M != M || ctx->green_rs != rs_eff || ctx->green_order != ctx->p.order) {
        green_kernel<<<grid_for(MC, 256), 256, 0, s>>>(M, rs_eff, ctx->p.order, ctx->green);
        TPM_HIP(ctx, hipGetLastError());
        ctx->green_M = M;
        ctx->green_rs = rs_eff;
        ctx->green_order = ctx->p.order;
    }
    TPM_TRY(tpm_fill_u64(ctx, ctx->rho_q, M3));
    // Blocked deposit needs particles in chain-mesh order; pick the largest
    // block edge (cells) whose LDS footprint fits 64 KiB.
    int B = 0, wmax = 0;
    if (ctx->n > 0 && ctx->chain_valid && ctx->chain_C == ctx->C && ctx->deposit_block >= 0) {
        const int C = ctx->C;
        std::vector<uint64_t> b(C + 1);
        const double ninv = 1.0 / C;
        for (int k = 0; k < C; k++) b[k] = (uint32_t)(4294967296.0 * ((double)k * ninv));
        b[C] = 1ull << 32;
        const int want = ctx->deposit_block > 0 ? ctx->deposit_block : 3;
        for (int cand = want; cand >= 1; cand--) {
            int wm = 0;
            for (int c0 = 0; c0 < C; c0 += cand) {
                const int c1 = std::min(c0 + cand, C) - 1;
                const int lo = (int)((b[c0] * (uint64_t)M) >> 32) - 1;
                const int hi = (int)(((b[c1 + 1] - 1) * (uint64_t)M) >> 32) + 2;
                wm = std::max(wm, hi - lo + 1);
            }
            if ((int64_t)wm * wm * wm * 8 <= 65536) {
                B = cand;
                wmax = wm;
                break;
            }
        }
    }
    if (B > 0) {
        BlockedDepositArgs da;
        da.x = ctx->x;
        da.y = ctx->y;
        da.z = ctx->z;
        da.cell_begin = ctx->cell_begin;
        da.bounds = ctx->bounds;
        da.grid = ctx->rho_q;
        da.C = ctx->C;
        da.M = M;
        da.B = B;
        da.nb = (ctx->C + B - 1) / B;
        da.wmax = wmax;
        const unsigned nblk = (unsigned)((int64_t)da.nb * da.nb * da.nb);
        const size_t lds = sizeof(unsigned long long) * (size_t)wmax * wmax * wmax;
        switch (ctx->p.order) {
            case 2: deposit_blocked_kernel<2><<<nblk, 256, lds, s>>>(da); break;
            case 3: deposit_blocked_kernel<3><<<nblk, 256, lds, s>>>(da); break;
            default: deposit_blocked_kernel<4><<<nblk, 256, lds, s>>>(da); break;
        }
        TPM_HIP(ctx, hipGetLastError());
    } else if (ctx->n > 0) {
        const unsigned g = grid_for(ctx->n, 256);
        switch (ctx->p.order) {
            case 2: deposit_kernel<2><<<g, 256, 0, s>>>(ctx->n, ctx->x, ctx->y, ctx->z, M, ctx->rho_q); break;
            case 3: deposit_kernel<3><<<g, 256, 0, s>>>(ctx->n, ctx->x, ctx->y, ctx->z, M, ctx->rho_q); break;
            default: deposit_kernel<4><<<g, 256, 0, s>>>(ctx->n, ctx->x, ctx->y, ctx->z, M, ctx->rho_q); break;
        }
        TPM_HIP(ctx, hipGetLastError());
    }
    q32_to_float_kernel<<<grid_for(M3, 256), 256, 0, s>>>(M3, ctx->rho_q, ctx->rho);
    TPM_HIP(ctx, hipGetLastError());
    TPM_TRY(tpm_fft_forward(ctx, ctx->rho, ctx->rho_hat));
    green_apply_kernel<<<grid_for(MC, 256), 256, 0, s>>>(MC, ctx->rho_hat, ctx->green);
    TPM_HIP(ctx, hipGetLastError());
    TPM_TRY(tpm_fft_inverse(ctx, ctx->rho_hat, ctx->phi));
    ctx->pm_valid = true;
    return TPM_OK;
}

tpm_status tpm_fft_release(tpm_ctx* ctx) { return fft_destroy(ctx); }
