// tpm_probe.hip -- the issue-rate ceiling of the direct kernel's pair body.
//
// pair_rate_kernel runs pp_pair (the device function pp_kick_kernel uses)
// with pp_kick_kernel's 16-sink tile layout (lane l: sink l % 16, source
// stream l / 16), its flush loop (two pairs per trip, the next source read
// one step ahead from LDS) and its 8-wave launch bound, over 256 sources
// held in LDS, with nothing else in the loop: no staging, culling or
// epilogue.  Sources fill a cube of side 2 rcut around the sinks, so about
// half of the pairs are inside rcut and the softening branch is rare, as in
// the kernel.  Pairs per second of this loop bound what the kernel's pair
// loop can reach at the same occupancy.
#include "tpm_pp_device.hpp"

namespace {

constexpr int PROBE_SRC = 256;
constexpr int PROBE_TS = 16;
constexpr int PROBE_SPLIT = 64 / PROBE_TS;

__global__ __launch_bounds__(64, 8) void pair_rate_kernel(tpm_box_dev B, int iters, float* __restrict__ out) {
    __shared__ int4 sbuf[PROBE_SRC + PROBE_SPLIT];
    const int lane = (int)lane_id();
    const int slot = lane % PROBE_TS, stream = lane / PROBE_TS;
    const float rc = sqrtf(B.rcut2s);
    for (int k = lane; k < PROBE_SRC + PROBE_SPLIT; k += 64) {
        // a hash of (block, k): offsets uniform in [-rcut, rcut)^3
        uint32_t h = (uint32_t)(blockIdx.x * (PROBE_SRC + PROBE_SPLIT) + k) * 2654435761u;
        int c[3];
        for (int d = 0; d < 3; d++) {
            h ^= h >> 15;
            h *= 2246822519u;
            h ^= h >> 13;
            c[d] = (int)((float)(h >> 8) * 0x1p-24f * 2.f * rc - rc);
        }
        sbuf[k] = make_int4(c[0], c[1], c[2], k);
    }
    __syncthreads();
    const int ox = slot * 4099 - 32768, oy = 20011 - slot * 3001, oz = slot * 1777;
    float gx = 0.f, gy = 0.f, gz = 0.f, ph = 0.f;
    for (int it = 0; it < iters; it++) {
        int4 s = sbuf[stream];
#pragma unroll 2
        for (int qq = 0; qq < PROBE_SRC; qq += PROBE_SPLIT) {
            const int4 sn = sbuf[qq + PROBE_SPLIT + stream];
            pp_pair<false>(B, ox, oy, oz, s, gx, gy, gz, ph);
            s = sn;
        }
    }
    out[blockIdx.x * 64 + lane] = gx + gy + gz + ph;
}

}  // namespace

extern "C" tpm_status tpm_probe_pair_rate(tpm_ctx* ctx, int waves_per_simd, double* pairs_per_second) {
    if (!ctx || !pairs_per_second) return TPM_ERR_INVALID;
    if (!ctx->params_set) return tpm_fail(ctx, TPM_ERR_STATE, "tpm_probe_pair_rate: parameters not set");
    if (waves_per_simd < 1 || waves_per_simd > 8)
        return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_probe_pair_rate: waves_per_simd must be in [1, 8]");
    (void)hipSetDevice(ctx->device);
    hipDeviceProp_t prop;
    TPM_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
    const int blocks = prop.multiProcessorCount * 4 * waves_per_simd;  // one wave64 per block, 4 SIMDs per CU
    const int iters = 64;
    tpm_box_dev B;
    tpm_box_fill(&B, ctx->p, 0, 1, 1.0, 1.0);
    float* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TPM_HIP(ctx, hipMalloc(&out, sizeof(float) * (size_t)blocks * 64));
    tpm_status st = TPM_OK;
    float ms = 0.f;
    const int reps = 5;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        st = tpm_fail(ctx, TPM_ERR_HIP, "tpm_probe_pair_rate: hipEventCreate failed");
    } else {
        pair_rate_kernel<<<blocks, 64, 0, ctx->stream>>>(B, iters, out);  // warm-up
        (void)hipEventRecord(e0, ctx->stream);
        for (int r = 0; r < reps; r++) pair_rate_kernel<<<blocks, 64, 0, ctx->stream>>>(B, iters, out);
        (void)hipEventRecord(e1, ctx->stream);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.f)
            st = tpm_fail(ctx, TPM_ERR_HIP, "tpm_probe_pair_rate: launch or timing failed");
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (st != TPM_OK) return st;
    const double pairs = (double)blocks * iters * (PROBE_SRC / PROBE_SPLIT) * 64.0 * reps;
    *pairs_per_second = pairs / (ms * 1e-3);
    return TPM_OK;
}
