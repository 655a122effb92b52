// tpm_ic.hip -- synthetic initial conditions on the device.
//
// Reference: initialize() (initialize.cpp:91-221) draws a Gaussian random
// field from a tabulated P(k), builds the Zel'dovich displacement per axis
// through inverse FFTs and displaces a particle lattice.  The benchmark
// workloads use a power law instead:
//   P(k) ~ k^n exp(-(k / k_Nyq)^2),
// white noise from a counter-based hash (splitmix64 + Box-Muller, so a seed
// fixes the field on any device), displacement psi_hat = i k / k^2 delta_hat,
// scaled so that the rms 3-D displacement equals the requested number of
// lattice cells.  Positions are (i + 1/2 + psi) / n_side, wrapped into
// [0, 1) and stored as fixed32.
#include <algorithm>

#include "tpm_internal.hpp"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void noise_kernel(int64_t n, uint64_t seed, float* __restrict__ out) {
    const uint64_t sk = mix64(seed);
    for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t h1 = mix64(sk ^ mix64(2 * (uint64_t)l));
        const uint64_t h2 = mix64(sk ^ mix64(2 * (uint64_t)l + 1));
        const double u1 = (double)((h1 >> 11) + 1) * 0x1p-53;
        const double u2 = (double)(h2 >> 11) * 0x1p-53;
        out[l] = (float)(sqrt(-2.0 * log(u1)) * cospi(2.0 * u2));
    }
}

// One displacement component: psi_hat_d = i k_d / k^2 sqrt(P(k)) noise_hat.
// Modes on a Nyquist plane (index M/2 for even M) are dropped.
__global__ void displace_kernel(int M, int dim, double spectral_index, const float2* __restrict__ in,
                                float2* __restrict__ out) {
    const int Mh = M / 2 + 1;
    const int64_t total = (int64_t)M * M * Mh;
    const double c0 = 2.0 * M_PI / M;
    for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < total; l += (int64_t)gridDim.x * blockDim.x) {
        const int iz = (int)(l % Mh);
        const int iy = (int)((l / Mh) % M);
        const int ix = (int)(l / ((int64_t)Mh * M));
        const int ii[3] = {ix, iy, iz};
        double k[3];
        bool nyq = false;
        for (int d = 0; d < 3; d++) {
            k[d] = c0 * (double)(2 * ii[d] < M ? ii[d] : ii[d] - M);
            nyq |= ((M % 2) == 0 && 2 * ii[d] == M);
        }
        const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        float2 v = make_float2(0.f, 0.f);
        if (k2 > 0.0 && !nyq) {
            const double amp = pow(k2, 0.25 * spectral_index) * exp(-0.5 * k2 / (M_PI * M_PI));
            const double f = amp * k[dim] / k2;
            const float2 a = in[l];
            v.x = (float)(-f * a.y);
            v.y = (float)(f * a.x);
        }
        out[l] = v;
    }
}

__global__ void sumsq_kernel(int64_t n, const float* __restrict__ a, const float* __restrict__ b,
                             const float* __restrict__ c, double* __restrict__ out) {
    double acc = 0.0;
    for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x)
        acc += (double)a[l] * a[l] + (double)b[l] * b[l] + (double)c[l] * c[l];
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Sum of the per-block partials in a fixed order (deterministic).
__global__ void sum_parts_kernel(int nparts, double* __restrict__ parts) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < nparts; k++) s += parts[k];
        parts[nparts] = s;
    }
}

__device__ __forceinline__ uint32_t to_fixed(double pos) {
    pos -= floor(pos);
    double r = pos * 4294967296.0;
    if (r >= 4294967295.0) r = 4294967295.0;
    return (uint32_t)r;
}

__global__ void lattice_kernel(int n_side, double scale, const float* __restrict__ px, const float* __restrict__ py,
                               const float* __restrict__ pz, uint32_t* __restrict__ x, uint32_t* __restrict__ y,
                               uint32_t* __restrict__ z, float* __restrict__ vx, float* __restrict__ vy,
                               float* __restrict__ vz, int8_t* __restrict__ rung) {
    const int64_t n = (int64_t)n_side * n_side * n_side;
    const double inv = 1.0 / n_side;
    for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = l % n_side;
        const int64_t j = (l / n_side) % n_side;
        const int64_t i = l / ((int64_t)n_side * n_side);
        x[l] = to_fixed(((double)i + 0.5 + scale * px[l]) * inv);
        y[l] = to_fixed(((double)j + 0.5 + scale * py[l]) * inv);
        z[l] = to_fixed(((double)k + 0.5 + scale * pz[l]) * inv);
        vx[l] = vy[l] = vz[l] = 0.f;
        rung[l] = 0;
    }
}

__global__ void uniform_kernel(int64_t n, uint64_t seed, uint32_t* __restrict__ x, uint32_t* __restrict__ y,
                               uint32_t* __restrict__ z, float* __restrict__ vx, float* __restrict__ vy,
                               float* __restrict__ vz, int8_t* __restrict__ rung) {
    const uint64_t sk = mix64(seed);
    for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t h1 = mix64(sk ^ mix64(2 * (uint64_t)l));
        const uint64_t h2 = mix64(sk ^ mix64(2 * (uint64_t)l + 1));
        x[l] = (uint32_t)h1;
        y[l] = (uint32_t)(h1 >> 32);
        z[l] = (uint32_t)h2;
        vx[l] = vy[l] = vz[l] = 0.f;
        rung[l] = 0;
    }
}

}  // namespace

tpm_status tpm_ic_uniform_run(tpm_ctx* ctx, int64_t n, uint64_t seed) {
    TPM_TRY(tpm_ensure_particles(ctx, n));
    ctx->n = n;
    uniform_kernel<<<grid_for(n, 256), 256, 0, ctx->stream>>>(n, seed, ctx->x, ctx->y, ctx->z, ctx->vx, ctx->vy,
                                                              ctx->vz, ctx->rung);
    TPM_HIP(ctx, hipGetLastError());
    TPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->chain_valid = false;
    ctx->h2d_identity = true;
    ctx->ref_valid = false;
    ctx->pm_valid = false;
    return TPM_OK;
}

tpm_status tpm_ic_zeldovich_run(tpm_ctx* ctx, int n_side, uint64_t seed, double rms_disp, double spectral_index) {
    if (n_side < 2 || n_side > 2048) return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_ic_zeldovich: bad n_side");
    const int M = n_side;
    const int64_t n = (int64_t)M * M * M;
    const int64_t MC = (int64_t)M * M * (M / 2 + 1);
    hipStream_t s = ctx->stream;
    TPM_TRY(tpm_ensure_particles(ctx, n));
    TPM_TRY(tpm_ensure_mesh(ctx, M));  // FFT plans and mesh buffers of size M
    // Scratch: noise (rho), its transform (rho_hat), one component spectrum
    // (green), and three displacement grids (phi, gx, gy of the particle
    // store, sized n = M^3).
    if (!ctx->rho_hat1 || ctx->rho_hat1_cap < MC) {
        if (ctx->rho_hat1) TPM_HIP(ctx, hipFree(ctx->rho_hat1));
        ctx->rho_hat1 = nullptr;
        TPM_HIP(ctx, hipMalloc(&ctx->rho_hat1, sizeof(float2) * MC));
        ctx->rho_hat1_cap = MC;
    }
    float2* spec = ctx->rho_hat1;
    constexpr unsigned NPART = 2048;  // sum-of-squares blocks
    if (!ctx->dscratch) TPM_HIP(ctx, hipMalloc(&ctx->dscratch, sizeof(double) * (NPART + 8)));
    noise_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, seed, ctx->rho);
    TPM_HIP(ctx, hipGetLastError());
    TPM_TRY(tpm_fft_forward(ctx, ctx->rho, ctx->rho_hat));
    float* comp[3] = {ctx->phi, ctx->gx, ctx->gy};
    for (int d = 0; d < 3; d++) {
        displace_kernel<<<grid_for(MC, 256), 256, 0, s>>>(M, d, spectral_index, ctx->rho_hat, spec);
        TPM_HIP(ctx, hipGetLastError());
        TPM_TRY(tpm_fft_inverse(ctx, spec, comp[d]));
    }
    double* parts = ctx->dscratch;
    const unsigned nb = std::min<unsigned>(grid_for(n, 256), NPART);
    sumsq_kernel<<<nb, 256, 0, s>>>(n, comp[0], comp[1], comp[2], parts);
    sum_parts_kernel<<<1, 64, 0, s>>>((int)nb, parts);
    double* d_sum = parts + nb;
    TPM_HIP(ctx, hipGetLastError());
    double sum = 0.0;
    TPM_HIP(ctx, hipMemcpyAsync(&sum, d_sum, sizeof(double), hipMemcpyDeviceToHost, s));
    TPM_HIP(ctx, hipStreamSynchronize(s));
    const double rms = sqrt(sum / (double)n);
    const double scale = rms > 0.0 ? rms_disp / rms : 0.0;
    lattice_kernel<<<grid_for(n, 256), 256, 0, s>>>(M, scale, comp[0], comp[1], comp[2], ctx->x, ctx->y, ctx->z,
                                                    ctx->vx, ctx->vy, ctx->vz, ctx->rung);
    TPM_HIP(ctx, hipGetLastError());
    TPM_HIP(ctx, hipStreamSynchronize(s));
    ctx->n = n;
    ctx->chain_valid = false;
    ctx->h2d_identity = true;
    ctx->ref_valid = false;
    ctx->pm_valid = false;  // phi was used as scratch
    return TPM_OK;
}
