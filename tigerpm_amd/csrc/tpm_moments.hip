// tpm_moments.hip -- FP64 global momentum and energy reductions.
//
// No reference counterpart: the reference prints timings and pair counts
// after a kick (kick_treepm.cu:1084-1087) but no conserved quantities.  Per
// particle the kick epilogue (kick_treepm.cu:716-754) leaves v, g and phi;
// these kernels reduce, over every resident particle of a box, with masses
// m = GM (the kick's g and phi already carry GM):
//   momentum  m sum v,   force  m sum g,
//   kinetic   m/2 sum |v|^2,   potential  m/2 sum phi.
// FP64 throughout, in a fixed order: the particles of a box are cut into
// TPM_MOMENT_BLOCKS contiguous chunks (the cut depends on n only), each
// chunk is summed by one workgroup (thread-strided, then a fixed tree in
// LDS), and one workgroup sums the chunk partials in chunk order.  The
// result is bit-identical across runs for the same particle order.
#include "tpm_internal.hpp"

namespace {

constexpr int MT = 256;  // threads per workgroup

__device__ __forceinline__ void block_sum8(double v[8], double* sh) {
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 8; k++) sh[k * MT + t] = v[k];
    __syncthreads();
    for (int o = MT / 2; o > 0; o >>= 1) {
        if (t < o) {
#pragma unroll
            for (int k = 0; k < 8; k++) sh[k * MT + t] += sh[k * MT + t + o];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = sh[k * MT];
}

// part[8 b + k]: chunk b's sums of vx, vy, vz, gx, gy, gz, |v|^2, phi.
__global__ __launch_bounds__(MT) void moments_partial_kernel(int64_t n, int nchunk, const float* __restrict__ vx,
                                                             const float* __restrict__ vy,
                                                             const float* __restrict__ vz,
                                                             const float* __restrict__ gx,
                                                             const float* __restrict__ gy,
                                                             const float* __restrict__ gz,
                                                             const float* __restrict__ pot,
                                                             double* __restrict__ part) {
    __shared__ double sh[8 * MT];
    const int b = blockIdx.x;
    if (b >= nchunk) return;
    const int64_t per = (n + nchunk - 1) / nchunk;
    const int64_t i0 = (int64_t)b * per, i1 = min(n, i0 + per);
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = i0 + threadIdx.x; i < i1; i += MT) {
        const double ux = vx[i], uy = vy[i], uz = vz[i];
        v[0] += ux;
        v[1] += uy;
        v[2] += uz;
        v[3] += (double)gx[i];
        v[4] += (double)gy[i];
        v[5] += (double)gz[i];
        v[6] += ux * ux + uy * uy + uz * uz;
        v[7] += (double)pot[i];
    }
    block_sum8(v, sh);
    if (threadIdx.x < 8) part[8 * b + threadIdx.x] = v[threadIdx.x];
}

// out: momentum[3], force[3], kinetic, potential (tpm_moments layout).
__global__ __launch_bounds__(MT) void moments_final_kernel(int nchunk, const double* __restrict__ part, double m,
                                                           double* __restrict__ out) {
    __shared__ double sh[8 * MT];
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nchunk; b += MT) {
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] += part[8 * b + k];
    }
    block_sum8(v, sh);
    if (threadIdx.x < 8) out[threadIdx.x] = threadIdx.x < 6 ? m * v[threadIdx.x] : 0.5 * m * v[threadIdx.x];
}

}  // namespace

int tpm_moment_chunks(int64_t n) {
    const int64_t c = (n + 4095) / 4096;
    return (int)std::max<int64_t>(1, std::min<int64_t>(TPM_MOMENT_BLOCKS, c));
}

tpm_status tpm_moments_enqueue(std::string& err, hipStream_t s, int64_t n, const float* vx, const float* vy,
                               const float* vz, const float* gx, const float* gy, const float* gz, const float* pot,
                               double m, double* part, double* out) {
    const int nchunk = tpm_moment_chunks(n);
    if (n > 0)
        moments_partial_kernel<<<nchunk, MT, 0, s>>>(n, nchunk, vx, vy, vz, gx, gy, gz, pot, part);
    moments_final_kernel<<<1, MT, 0, s>>>(n > 0 ? nchunk : 0, part, m, out);
    TPM_HIPE(err, hipGetLastError());
    return TPM_OK;
}

tpm_status tpm_ctx_moments(tpm_ctx* ctx, tpm_moments* out) {
    if (!ctx->moment_buf)
        TPM_HIP(ctx, hipMalloc(&ctx->moment_buf, sizeof(double) * (8 * TPM_MOMENT_BLOCKS + 8)));
    double* res = ctx->moment_buf + 8 * TPM_MOMENT_BLOCKS;
    TPM_TRY(tpm_moments_enqueue(ctx->err, ctx->stream, ctx->n, ctx->vx, ctx->vy, ctx->vz, ctx->gx, ctx->gy, ctx->gz,
                                ctx->pot, ctx->p.GM, ctx->moment_buf, res));
    double h[8];
    TPM_HIP(ctx, hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    TPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int d = 0; d < 3; d++) {
        out->momentum[d] = h[d];
        out->force[d] = h[3 + d];
    }
    out->kinetic = h[6];
    out->potential = h[7];
    return TPM_OK;
}
