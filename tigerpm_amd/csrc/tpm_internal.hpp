// tpm_internal.hpp -- context layout and device helpers shared by the
// libtigerpm_hip kernels.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <array>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "tigerpm_hip.h"

#define TPM_MAX_CHAIN 406    // C^3 * 64 must fit the 32-bit sort key

struct tpm_fft {
    int M = 0;
    rocfft_plan fwd = nullptr;
    rocfft_plan inv = nullptr;
    rocfft_execution_info info_fwd = nullptr;
    rocfft_execution_info info_inv = nullptr;
    void* work = nullptr;
    size_t work_bytes = 0;
};

struct tpm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    tpm_params p{};
    bool params_set = false;
    int C = 0;  // chain_dim
    int M = 0;  // mesh_dim

    // Resident particles (SoA, reference particles.hpp:17-25) and the
    // second set of buffers the chain-mesh gather writes into.
    int64_t n = 0;
    int64_t cap = 0;
    uint32_t *x = nullptr, *y = nullptr, *z = nullptr;
    float *vx = nullptr, *vy = nullptr, *vz = nullptr;
    int8_t* rung = nullptr;
    float *gx = nullptr, *gy = nullptr, *gz = nullptr, *pot = nullptr;
    uint32_t *x2 = nullptr, *y2 = nullptr, *z2 = nullptr;
    float *vx2 = nullptr, *vy2 = nullptr, *vz2 = nullptr;
    int8_t* rung2 = nullptr;

    // Chain mesh.
    uint32_t *keys = nullptr, *keys2 = nullptr, *idx = nullptr, *perm = nullptr;
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    uint32_t* bounds = nullptr;  // C split planes (raw values)
    int32_t* cell_begin = nullptr;
    int64_t cells_cap = 0;
    int32_t* tile_cnt = nullptr;  // tiles per cell
    int32_t* tile_end = nullptr;  // inclusive scan of tile_cnt
    int4* tiles = nullptr;        // (cell id, begin, end, 0)
    int64_t tiles_cap = 0;
    int32_t* ntiles = nullptr;    // device scalar
    void* scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    bool chain_valid = false;
    int chain_C = 0;
    int tile = 16;        // sinks per PP tile: 16, 32 or 64 (tuning)
    int chain_tile = 0;   // tile size the current tile list was built with
    int deposit_block = 0;  // chain cells per deposit block edge; 0 auto, -1 per-particle atomics

    // PM mesh.
    int64_t mesh_cap = 0;  // M^3 capacity
    unsigned long long* rho_q = nullptr;  // Q32 fixed-point density
    float* rho = nullptr;                 // FFT real input
    float2* rho_hat = nullptr;            // M*M*(M/2+1)
    float2* rho_hat1 = nullptr;           // transform of the half-cell-shifted deposit
    int64_t rho_hat1_cap = 0;
    float* phi = nullptr;                 // potential mesh
    float* green = nullptr;
    int green_M = 0, green_order = 0;
    double green_rs = 0.0;
    tpm_fft fft;
    bool pm_valid = false;

    // Diagnostics.
    unsigned long long* counters = nullptr;  // [0] executed pairs, [1] max rung
    double* dscratch = nullptr;
    hipEvent_t ev[8] = {};

    // Launch timing log (tpm_profile).
    bool prof_on = false;
    std::vector<hipEvent_t> prof_pool;       // all events ever created
    size_t prof_used = 0;                    // events handed out since enable
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_pp;     // PP kernel
    std::vector<std::array<hipEvent_t, 4>> prof_step;           // start, chain, PM, end
};

// Next event from the context's pool (created on first use).
static inline hipEvent_t tpm_prof_event(tpm_ctx* ctx) {
    if (ctx->prof_used == ctx->prof_pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->prof_pool.push_back(e);
    }
    return ctx->prof_pool[ctx->prof_used++];
}

#define TPM_HIP(ctx, call)                                                              \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);             \
            return TPM_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)

#define TPM_TRY(call)                     \
    do {                                          \
        tpm_status s_ = (call);                   \
        if (s_ != TPM_OK) return s_;              \
    } while (0)

static inline tpm_status tpm_fail(tpm_ctx* ctx, tpm_status s, const std::string& msg) {
    ctx->err = msg;
    return s;
}

template <class T>
static inline tpm_status tpm_grow(tpm_ctx* ctx, T** ptr, int64_t count, int64_t* cap_field_or_null,
                                  int64_t old_cap) {
    (void)cap_field_or_null;
    if (*ptr != nullptr && count <= old_cap) return TPM_OK;
    if (*ptr) {
        TPM_HIP(ctx, hipFree(*ptr));
        *ptr = nullptr;
    }
    if (count <= 0) count = 1;
    TPM_HIP(ctx, hipMalloc(reinterpret_cast<void**>(ptr), sizeof(T) * (size_t)count));
    return TPM_OK;
}

static inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 1048576) g = 1048576;
    return (unsigned)g;
}

// ---------------------------------------------------------------------------
// Device helpers (integer parts bit-identical to the CPU restatement).
// ---------------------------------------------------------------------------

// Chain cell coordinate: b[k] <= raw < b[k+1] for the reference's split
// planes b[k] = fixed32(k * (1.0 / C)) (chainmesh.cpp:198-199,
// particles.cpp:417), plus the 2-bit sub-cell used for the in-cell order.
__device__ __forceinline__ int tpm_chain_coord(uint32_t raw, int C, const uint32_t* b, int* sub) {
    const uint64_t P = (uint64_t)raw * (uint64_t)C;
    int k = (int)(P >> 32);
    int s = (int)((uint32_t)P >> 30);
    if (k + 1 < C && raw >= b[k + 1]) {
        k++;
        s = 0;
    } else if (raw < b[k]) {
        k--;
        s = 3;
    }
    *sub = s;
    return k;
}

// One-axis integer mass-assignment weights (see the CPU restatement for the
// derivation): CIC/TSC in Q21, cloud4 in Q16.  Returns the stencil width.
// shift = 2^31 deposits onto the grid displaced by half a cell (interlacing).
__device__ __forceinline__ int tpm_weights_1d(int order, uint32_t raw, int M, uint64_t shift, int* base,
                                              uint64_t w[4], int* qbits) {
    const uint64_t P = (uint64_t)raw * (uint64_t)M + shift;
    if (order == 2) {
        const uint64_t t = (uint32_t)P >> 11;
        *base = (int)(P >> 32);
        w[0] = (1ull << 21) - t;
        w[1] = t;
        *qbits = 21;
        return 2;
    } else if (order == 3) {
        const uint64_t j0 = (P + 0x80000000ull) >> 32;
        const int64_t d = (int64_t)(P - (j0 << 32));
        const int64_t ds = d >> 11;
        const int64_t hp = 1ll << 20;
        const uint64_t a = (uint64_t)(hp - ds);
        const uint64_t c = (uint64_t)(hp + ds);
        const uint64_t wm = (a * a) >> 22;
        const uint64_t wp = (c * c) >> 22;
        *base = (int)j0 - 1;
        w[0] = wm;
        w[1] = (1ull << 21) - wm - wp;
        w[2] = wp;
        *qbits = 21;
        return 3;
    } else {
        const int64_t t = (int64_t)((uint32_t)P >> 16);
        const int64_t u = 65536 - t;
        const int64_t one = 1ll << 16;
        w[0] = (uint64_t)((u * u * u / 6) >> 32);
        w[3] = (uint64_t)((t * t * t / 6) >> 32);
        const int64_t num = 4 * one * one * one - 6 * one * t * t + 3 * t * t * t;
        w[1] = (uint64_t)((num / 6) >> 32);
        w[2] = (uint64_t)one - w[0] - w[1] - w[3];
        *base = (int)(P >> 32) - 1;
        *qbits = 16;
        return 4;
    }
}

// Internal entry points implemented in the .hip files.
tpm_status tpm_chainmesh_build(tpm_ctx* ctx);
tpm_status tpm_pm_solve(tpm_ctx* ctx);
tpm_status tpm_pp_kick(tpm_ctx* ctx, int min_rung, double scale, double t0, int first_call, unsigned flags,
                       tpm_stats* stats);
tpm_status tpm_fft_plan(tpm_ctx* ctx, int M);
tpm_status tpm_fft_forward(tpm_ctx* ctx, float* in, float2* out);
tpm_status tpm_fft_inverse(tpm_ctx* ctx, float2* in, float* out);
tpm_status tpm_ensure_particles(tpm_ctx* ctx, int64_t n);
tpm_status tpm_ensure_mesh(tpm_ctx* ctx, int M);
tpm_status tpm_ewald_run(tpm_ctx* ctx, int64_t ns, const int64_t* samples, double alpha, double* phi, double* gx,
                         double* gy, double* gz);
tpm_status tpm_ic_zeldovich_run(tpm_ctx* ctx, int n_side, uint64_t seed, double rms_disp, double spectral_index);
tpm_status tpm_ic_uniform_run(tpm_ctx* ctx, int64_t n, uint64_t seed);
tpm_status tpm_fill_u64(tpm_ctx* ctx, unsigned long long* p, int64_t n);
