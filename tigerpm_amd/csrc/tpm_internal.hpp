// tpm_internal.hpp -- context layout and device helpers shared by the
// libtigerpm_hip kernels.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

// The direct kernel's per-kick counters are spread over slots on separate
// 128-byte lines (slot = workgroup index mod TPM_COUNTER_SLOTS), so the
// millions of per-wave atomics of a launch do not all meet on one address.
constexpr int TPM_COUNTER_SLOTS = 64;
constexpr int TPM_COUNTER_STRIDE = 16;
constexpr int TPM_COUNTER_WORDS = TPM_COUNTER_SLOTS * TPM_COUNTER_STRIDE;

// The counters of all slots (h: TPM_COUNTER_WORDS words copied from the
// device) -> out[0] executed pairs, [1] max rung, [2] candidate pairs,
// [3] multipole interactions.
inline void tpm_reduce_counters(const unsigned long long* h, unsigned long long out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0;
    for (int k = 0; k < TPM_COUNTER_SLOTS; k++) {
        const unsigned long long* c = h + k * TPM_COUNTER_STRIDE;
        out[0] += c[0];
        out[1] = c[1] > out[1] ? c[1] : out[1];
        out[2] += c[2];
        out[3] += c[3];
    }
}
#include <rocfft/rocfft.h>

#include <array>
#include <iterator>
#include <map>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "tigerpm_hip.h"
#include "tpm_alloc.hpp"

#define TPM_MAX_CHAIN 406    // C^3 * 64 must fit the 32-bit sort key

// Plans of one cubic size (single precision, out of place), shared by every
// context and batch on the device: made once per (device, M, backend) and
// kept until the process ends.  Backend: rocFFT, or this library's LDS-tiled
// transform (tpm_fft.hip) for power-of-two M in [16, 1024], whose plan is its
// twiddle table.
struct tpm_fft_plans {
    rocfft_plan fwd = nullptr;  // R2C
    rocfft_plan inv = nullptr;  // C2R
    float2* tw = nullptr;       // LDS-tiled: e^{-2 pi i q / M}, q < M
    size_t work_bytes = 0;
};
// Returns nullptr (and sets *err) when the size cannot be planned.
const tpm_fft_plans* tpm_fft_plans_get(int device, int M, bool lds, std::string* err);

// What one caller needs to execute cached plans on its stream: execution
// infos bound to the stream and a work buffer large enough for every plan it
// has used.  The LDS-tiled transform's inverse overwrites its input, as
// rocFFT's real inverse may.
struct tpm_fft {
    int M = 0;
    const tpm_fft_plans* plans = nullptr;
    rocfft_execution_info info = nullptr;
    hipStream_t stream = nullptr;
    void* work = nullptr;
    size_t work_bytes = 0;
};
bool tpm_lds_fft_size(int M);
tpm_status tpm_lds_fft_twiddles(std::string& err, int M, float2** tf);
// in: M^3 floats, or with q32 the int64 Q32 density grid; green: the PM's
// Green's table applied to the spectrum as it is stored (nullptr: none).
tpm_status tpm_lds_fft_r2c(std::string& err, hipStream_t st, int M, const void* in, bool q32, float2* out,
                           float2* work, const float2* tf, const float* green);
tpm_status tpm_lds_fft_c2r(std::string& err, hipStream_t st, int M, float2* in, float* out, float2* work,
                           const float2* tf);
// The Green's table regrouped into the fused x pass's workgroup tiles: with
// gt null only *count (floats) is set, else gt[0 .. *count) is written.
tpm_status tpm_lds_green_tiles(std::string& err, hipStream_t st, int M, const float* green, float* gt,
                               int64_t* count);
// The PM's Poisson solve on the LDS-tiled transform: Q32 density q ->
// potential phi (Green's factors green_tiles, from tpm_lds_green_tiles,
// applied between the transforms); spec and work are M * M * (M/2 + 1)
// complex scratch.
tpm_status tpm_lds_pm_solve(std::string& err, hipStream_t st, int M, const unsigned long long* q, float* phi,
                            float2* spec, float2* work, const float2* tf, const float* green_tiles);

// Per-box constants of the fused PP + interpolation + kick kernel, and the
// box's place in the batched arenas.  One entry per box: the single-box path
// passes it by value, the batched path keeps a table in device memory.
struct tpm_box_dev {
    // layout
    int64_t p0, n;       // particle range in the arena (before this step's sort)
    int64_t q0;          // first particle after the sort
    int64_t mesh0;       // offset of the box's real meshes (M^3 nodes)
    int64_t k0;          // offset of the box's spectra (M^2 (M/2+1) modes)
    int64_t green0;      // offset of the box's Green's table
    int32_t cell0;       // first global chain cell of the box
    int32_t C, M;        // chain_dim, mesh_dim
    int32_t live;        // 0: particles leave the arena at the next sort
    double ninv;         // 1.0 / C (split planes fixed32(k * ninv))
    int32_t dep0;        // first deposit work item of the box
    int32_t B, nb;       // deposit block edge (chain cells) and blocks per axis
    int32_t min_rung, first_call, do_phi, cull;
    // pair constants in 2^-32 box units
    float rcut2s, cull2s, inv2rs_s, tsp, h2s, h2inv_s, hinv_s, h3inv_s;
    float e2s;  // -log2(e) / (4 rs^2) in 2^-64 units: exp(-x^2) = exp2(r2 * e2s)
    float p_inv2rs_s, tsp_inv2rs_s;  // 0.3275911 / (2 rs) and 2 / sqrt(pi) / (2 rs), 2^-32 units
    // inside the softening length: 1 / (2 rs)^2 in 2^-64 units (x^2 = r2 *
    // inv4rs2_s) and 4 / (3 sqrt(pi)) / (2 rs)^3 in 2^-96 units
    float inv4rs2_s, k3_s;
    // kick epilogue
    float GM, phi0, eta, t0, scale, hsoft;
    double drift;  // t0 of the drift step (batched path)
};

// Fills the pair, epilogue and layout constants of a box from its
// parameters (compute_pp_interaction constants, kick_treepm.cu:833-853).
void tpm_box_fill(tpm_box_dev* b, const tpm_params& p, int min_rung, int first_call, double t0, double scale);

// PP tiles per workgroup when the tuning asks for the default.
#define TPM_PP_WAVES_DEFAULT 1
// Per-sink source lists in the PP kernel (16-sink tiles) unless the tuning
// says otherwise, for single boxes and batches alike.
#define TPM_PP_LISTS_DEFAULT 0
// Whole-box kicks with every particle active evaluate each pair once
// (tpm_pp_sym.hip) unless the tuning says otherwise.
#define TPM_PP_NEWTON_DEFAULT 0

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
    int pp_waves = TPM_PP_WAVES_DEFAULT;  // PP tiles per workgroup (1 or 4)
    int pp_lists = TPM_PP_LISTS_DEFAULT;  // per-sink source lists (16-sink tiles)
    int pp_newton = TPM_PP_NEWTON_DEFAULT;  // whole-box kicks: each pair evaluated once
    int fft_rocfft = 0;                     // 1: rocFFT for every mesh size
    // one-sided pairs: sink sums [n], 13 slots of source shares [13 n],
    // mesh force and potential [n]
    void* gsink = nullptr;
    void* gsrc = nullptr;
    void* gmesh = nullptr;
    int64_t sym_cap = 0;
    void* mesh_force = nullptr;  // float4 per particle: the mesh term of whole-box kicks
    int64_t mesh_force_cap = 0;
    int mesh_force_w = 0;        // nodes per axis around one chain cell
    int64_t chain_epoch = 0;  // counts chain-mesh builds (trees remember the one they follow)

    // The order the caller sees (tpm_set_host_order).  With
    // TPM_ORDER_REFERENCE each chain-mesh build also finds the order the
    // reference partition leaves the caller's previous order in: h2d[r] is
    // the device position of the caller's particle r (h2d_identity: the two
    // orders agree, as after an upload), ref_perm[r] the caller's position of
    // that particle before the build.
    int host_order = TPM_ORDER_DEVICE;
    bool h2d_identity = true;
    bool ref_valid = false;        // ref_perm belongs to the current chain mesh
    int64_t ref_cap = 0;
    uint32_t *h2d = nullptr, *h2d2 = nullptr, *ref_perm = nullptr;
    uint32_t *ref_cid = nullptr, *ref_cid2 = nullptr, *ref_idx2 = nullptr;
    uint32_t *ref_posa = nullptr, *ref_posb = nullptr, *ref_inv = nullptr;
    unsigned long long *ref_flag = nullptr, *ref_scan = nullptr;
    void* ref_scan_tmp = nullptr;
    size_t ref_scan_tmp_bytes = 0;
    int4* ref_tab = nullptr;       // per axis coordinate and level: (lo, mid, hi, depth)
    int ref_tab_C = 0;

    // Per-cell trees (tpm_tree.hip).
    int4* tree_pos = nullptr;        // tree order: x, y, z, chain-mesh index
    int4* tree_tmp = nullptr;        // partition ping-pong buffer
    int32_t* tree_moved = nullptr;   // partition scratch of cells built in device memory
    int64_t tree_pos_cap = 0;
    int32_t* tree_cnt = nullptr;     // per cell: node counts [0, ncell), bucket counts [ncell, 2 ncell)
    int32_t* node_begin = nullptr;   // ncell + 1 exclusive offsets
    int32_t* bucket_begin = nullptr; // ncell + 1
    int32_t* tree_part = nullptr;    // scan partials
    int32_t* tree_flags = nullptr;   // [0] depth exceeded, [1] walk stack overflow
    int64_t tree_cells_cap = 0;
    tpm_tree_node* nodes = nullptr;
    int64_t nodes_cap = 0;
    tpm_tree_bucket* buckets = nullptr;
    int64_t buckets_cap = 0;
    int64_t n_nodes = 0, n_buckets = 0;
    int64_t tree_epoch = -1;         // chain_epoch the trees were built from

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
    float* green_t = nullptr;             // green regrouped for the fused x pass
    int64_t green_t_cap = 0;
    bool green_t_valid = false;
    tpm_fft fft;
    bool pm_valid = false;

    // Diagnostics.
    // TPM_COUNTER_SLOTS slots of TPM_COUNTER_STRIDE words: [0] executed
    // pairs, [1] max rung, [2] candidate pairs, [3] multipole interactions
    // (tpm_reduce_counters)
    unsigned long long* counters = nullptr;
    double* dscratch = nullptr;
    double* moment_buf = nullptr;  // TPM_MOMENT_BLOCKS x 8 partials + 8 results
    hipEvent_t ev[8] = {};

    // Kicks restricted to a box of chain cells (tpm_kick_box) and the
    // device budget that splits them (tpm_set_budget).
    int64_t budget = 0;              // bytes of one kick's working set; 0 = no limit
    int4* box_tiles = nullptr;       // the tiles of the box being kicked
    int64_t box_tiles_cap = 0;
    int32_t* box_count = nullptr;    // device scalar: tiles in box_tiles
    int32_t* box_buckets = nullptr;  // tree walk: the buckets of the box
    int64_t box_buckets_cap = 0;
    std::vector<int32_t> h_cells;    // host copy of cell_begin ...
    int64_t h_cells_epoch = -1;      // ... of this chain-mesh build

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

#define TPM_HIPE(errstr, call)                                                          \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            (errstr) = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return TPM_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)
#define TPM_HIP(ctx, call) TPM_HIPE((ctx)->err, call)

#define TPM_TRY(call)                     \
    do {                                          \
        tpm_status s_ = (call);                   \
        if (s_ != TPM_OK) return s_;              \
    } while (0)

template <class T>
static inline tpm_status tpm_fail(T* ctx, tpm_status s, const std::string& msg) {
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
// Exact products on the 24-bit multipliers (v_mul_u32_u24 and
// v_mul_hi_u32_u24 issue at full rate; 32-bit and 64-bit integer multiplies
// at a quarter of it).  a, b < 2^24.
__device__ __forceinline__ uint64_t tpm_mul24(uint32_t a, uint32_t b) {
    return (uint64_t)(a & 0xFFFFFFu) * (uint64_t)(b & 0xFFFFFFu);
}
// x < 2^48, c < 2^24 and x c < 2^64: x c exactly.
__device__ __forceinline__ uint64_t tpm_mul48x24(uint64_t x, uint32_t c) {
    return tpm_mul24((uint32_t)x, c) + (tpm_mul24((uint32_t)(x >> 24), c) << 24);
}

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
        const uint64_t wm = tpm_mul24((uint32_t)a, (uint32_t)a) >> 22;  // a, c <= 2^21
        const uint64_t wp = tpm_mul24((uint32_t)c, (uint32_t)c) >> 22;
        *base = (int)j0 - 1;
        w[0] = wm;
        w[1] = (1ull << 21) - wm - wp;
        w[2] = wp;
        *qbits = 21;
        return 3;
    } else {
        // floor(floor(a / 6) / 2^32) = floor(floor(a / 2^32) / 6) for a >= 0,
        // so each cubic is shifted first and divided as a 32-bit integer
        const uint32_t t = (uint32_t)P >> 16;
        const uint32_t u = 65536u - t;
        const uint64_t one3 = 1ull << 48;
        const uint64_t t2 = tpm_mul24(t, t), t3 = tpm_mul48x24(t2, t);
        const uint64_t u3 = tpm_mul48x24(tpm_mul24(u, u), u);
        w[0] = (uint32_t)(u3 >> 32) / 6u;
        w[3] = (uint32_t)(t3 >> 32) / 6u;
        // 4 - 6 t^2 + 3 t^3 (Q48) > 0 for t in [0, 1]
        const uint64_t num = 4 * one3 - 6 * (t2 << 16) + 3 * t3;
        w[1] = (uint32_t)(num >> 32) / 6u;
        w[2] = (1ull << 16) - w[0] - w[1] - w[3];
        *base = (int)(P >> 32) - 1;
        *qbits = 16;
        return 4;
    }
}

// Split plane k of a chain mesh of C cells: fixed32(k * (1.0 / C)), the
// reference's chainmesh.cpp:198-199 in double, identical on host and device.
__host__ __device__ __forceinline__ uint32_t tpm_split_plane(int k, double ninv) {
    return (uint32_t)(4294967296.0 * ((double)k * ninv));
}

// tpm_chain_coord with the split planes computed instead of looked up.
__device__ __forceinline__ int tpm_chain_coord_f(uint32_t raw, int C, double ninv, int* sub) {
    const uint64_t P = (uint64_t)raw * (uint64_t)C;
    int k = (int)(P >> 32);
    int s = (int)((uint32_t)P >> 30);
    if (k + 1 < C && raw >= tpm_split_plane(k + 1, ninv)) {
        k++;
        s = 0;
    } else if (raw < tpm_split_plane(k, ninv)) {
        k--;
        s = 3;
    }
    *sub = s;
    return k;
}

// Index of the last entry with start(e) <= v in a table sorted by start
// (ties: the last one, which is the only one of them whose range is not
// empty).  Returns -1 when v is below every start.
template <class F>
__host__ __device__ __forceinline__ int tpm_upper_index(int n, int64_t v, F start) {
    int lo = 0, hi = n;  // first entry with start > v is in [lo, hi]
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (start(mid) <= v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo - 1;
}

// ---------------------------------------------------------------------------
// Batch of independent boxes resident on one device (tpm_batch_* API).
// ---------------------------------------------------------------------------

// Spectrum region of one live box, the batch keeps them sorted by k0.
struct tpm_kregion {
    int64_t k0, nk;
    int32_t M, pad_;
};

struct tpm_batch_box {
    int id = -1;
    tpm_box_spec spec{};
    int64_t n = 0;
    int64_t p0 = 0;  // particle range [p0, p0 + n) in the arena
    int steps = 0;
    bool live = true;
    int C = 0, M = 0;
    int64_t mesh0 = -1, k0 = -1;  // arena regions: M^3 real nodes, M^2 (M/2+1) modes
    int B = 0, nb = 0, wmax = 0;  // blocked deposit
};

struct tpm_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    tpm_batch_limits lim{};
    int64_t pcap = 0, mcap = 0, kcap = 0, ccap = 0, tcap = 0;

    // particle arena (SoA) and the buffers the chain-mesh gather writes into
    uint32_t *x = nullptr, *y = nullptr, *z = nullptr, *x2 = nullptr, *y2 = nullptr, *z2 = nullptr;
    float *vx = nullptr, *vy = nullptr, *vz = nullptr, *vx2 = nullptr, *vy2 = nullptr, *vz2 = nullptr;
    int8_t *rung = nullptr, *rung2 = nullptr;
    float *gx = nullptr, *gy = nullptr, *gz = nullptr, *pot = nullptr;
    float4* mesh_force = nullptr;  // mesh force and potential per particle (separate pass)
    uint32_t *key = nullptr, *slot = nullptr, *tmp = nullptr, *perm = nullptr;
    // chain cells and tiles of all live boxes
    int32_t *cnt = nullptr, *cell_begin = nullptr, *tile_cnt = nullptr, *tile_off = nullptr, *scan_part = nullptr;
    int4* tiles = nullptr;
    // mesh arenas: real nodes share mesh0, spectra and Green's tables share k0
    unsigned long long* rho_q = nullptr;
    float *rho = nullptr, *phi = nullptr, *green = nullptr;
    float2 *rho_hat = nullptr, *rho_hat1 = nullptr;
    tpm_range_alloc mesh_alloc, k_alloc;

    std::vector<tpm_batch_box> boxes;  // arena order
    int next_id = 0;
    int64_t n_arena = 0;  // particles in the arena, departed boxes included until the next sort
    tpm_box_dev* d_boxes = nullptr;   // device table (arena order)
    tpm_box_dev* h_boxes = nullptr;   // pinned staging of the table
    tpm_kregion* d_kreg = nullptr;    // live boxes' spectrum regions by k0
    tpm_kregion* h_kreg = nullptr;
    hipEvent_t table_ev = nullptr;    // the last table upload
    int64_t table_cap = 0;

    tpm_fft fft;
    // per-box FFTs fan out over side streams (fork/join with events; a
    // captured step records them as parallel branches)
    static constexpr int NSIDE = 4;
    hipStream_t side[NSIDE] = {};
    tpm_fft side_fft[NSIDE];
    hipEvent_t ev_fork = nullptr, ev_join[NSIDE] = {};
    float* rho1 = nullptr;  // float image of the shifted deposit (interlace)
    tpm_ctx* scratch = nullptr;  // initial conditions are generated here
    // counter slots as in tpm_ctx (tpm_reduce_counters); words 8..15 of
    // slot 0: the persistent PP launch's tile queues
    unsigned long long* counters = nullptr;
    double* moment_buf = nullptr;  // tpm_batch_moments: partials + results
    tpm_batch_tuning tune{16, 0, 0, 0, TPM_PP_LISTS_DEFAULT};
    int persist_grid = 0;  // workgroups of the persistent PP launch

    // captured steps, one per buffer parity (which of the two particle
    // buffer sets holds the state when the step starts)
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    uint64_t graph_version[2] = {0, 0};
    uint64_t layout_version = 1;  // changes with the set of boxes and the tuning
    int parity = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_stage[3] = {nullptr, nullptr, nullptr};  // after chain mesh, PM, PP
    bool stage_events = false;                             // record them (eager steps with stats)
    hipEvent_t copy_ev = nullptr;  // the last copy into the arena from the scratch context
};

// Layout of one step, fixed by the host before the kernels are enqueued.
struct tpm_batch_layout {
    int nbox = 0;          // entries of the device table (departed boxes included)
    int64_t n_arena = 0;   // particles the key kernel reads
    int64_t n_live = 0;    // particles after the sort
    int32_t ncell = 0;     // cells of live boxes (the trash cell is ncell)
    int32_t ndep = 0;      // deposit work items
    int64_t tiles_max = 0;
    int64_t mesh_hw = 0, k_hw = 0;
    int nkreg = 0;
    int order = 3, interlace = 0, lds = 0;
    int mfw = 0;           // potential nodes per axis around a chain cell (0: no separate mesh-force pass)
};
// Potential nodes per axis around one chain cell of C (the separate
// mesh-force pass's LDS cube); 0 when the cube exceeds 32 KiB.
int tpm_mesh_force_width(int C, int M);

// Internal entry points implemented in the .hip files.
tpm_status tpm_check_params(std::string& err, tpm_params* p);
tpm_status tpm_chainmesh_build(tpm_ctx* ctx);
// Exclusive scan of in[0, n) into out[0, n] (out[n] = total); part holds
// n / 2048 + 1 ints.
tpm_status tpm_scan_i32(std::string& err, hipStream_t s, const int32_t* in, int64_t n, int32_t* part, int32_t* out);
struct tpm_cell_box;
// Adds W_sr (candidate pairs of the 27-cell stencil) of the context's chain
// mesh to *dst (device); box: only the sinks of these cells.
tpm_status tpm_candidate_pairs(tpm_ctx* ctx, unsigned long long* dst, const tpm_cell_box* box = nullptr);
// Trees of the current chain mesh (no-op when they are current).
tpm_status tpm_tree_ensure(tpm_ctx* ctx);
// A box of chain cells [lo, hi) per axis (x, y, z), 0 <= lo < hi <= C.
struct tpm_cell_box {
    int lo[3], hi[3];
};
// Tree-walk short-range sum + mesh interpolation + kick (TPM_KICK_TREE).
// box: only the sink buckets of these cells (nullptr: all).
tpm_status tpm_tree_kick(tpm_ctx* ctx, int min_rung, double scale, double t0, int first_call, unsigned flags,
                         tpm_stats* stats, const tpm_cell_box* box = nullptr);
tpm_status tpm_pm_solve(tpm_ctx* ctx);
// The direct kernel over every tile, or (box != nullptr) over the box_tiles
// tiles of the box's cells (ntiles_box of them).
bool tpm_pp_sym_eligible(const tpm_ctx* ctx, int min_rung, unsigned flags, const tpm_cell_box* box);
tpm_status tpm_pp_kick_sym(tpm_ctx* ctx, double scale, double t0, int first_call, unsigned flags, tpm_stats* stats);
tpm_status tpm_pp_kick(tpm_ctx* ctx, int min_rung, double scale, double t0, int first_call, unsigned flags,
                       tpm_stats* stats, const tpm_cell_box* box = nullptr, int64_t ntiles_box = 0);
// Kick of the sinks in box (nullptr: all) on the direct or tree path,
// split recursively while its working set exceeds the context's budget.
tpm_status tpm_kick_any(tpm_ctx* ctx, const tpm_cell_box* box, int min_rung, double scale, double t0,
                        int first_call, unsigned flags, tpm_stats* stats);
tpm_status tpm_fft_plan(tpm_ctx* ctx, int M);
tpm_status tpm_fft_use(std::string& err, tpm_fft& f, int device, hipStream_t s, int M, bool lds = true);
tpm_status tpm_fft_r2c(std::string& err, tpm_fft& f, float* in, float2* out);
tpm_status tpm_fft_c2r(std::string& err, tpm_fft& f, float2* in, float* out);
void tpm_fft_free(tpm_fft& f);
tpm_status tpm_fft_forward(tpm_ctx* ctx, float* in, float2* out);
tpm_status tpm_fft_inverse(tpm_ctx* ctx, float2* in, float* out);
tpm_status tpm_ensure_particles(tpm_ctx* ctx, int64_t n);
tpm_status tpm_ensure_mesh(tpm_ctx* ctx, int M);
// Samples are particle indices (samples != nullptr) or fixed32 positions
// (pos[0..2]).
tpm_status tpm_ewald_run(tpm_ctx* ctx, int64_t ns, const int64_t* samples, const uint32_t* const pos[3],
                         double alpha, double* phi, double* gx, double* gy, double* gz);
tpm_status tpm_ewald_noalloc_run(tpm_ctx* ctx, int64_t ns, const uint32_t* spx, const uint32_t* spy,
                                 const uint32_t* spz, double alpha, double* phi, double* gx, double* gy, double* gz);
tpm_status tpm_ic_zeldovich_run(tpm_ctx* ctx, int n_side, uint64_t seed, double rms_disp, double spectral_index);
tpm_status tpm_ic_uniform_run(tpm_ctx* ctx, int64_t n, uint64_t seed);
tpm_status tpm_fill_u64(tpm_ctx* ctx, unsigned long long* p, int64_t n);

// FP64 momentum / energy reductions (tpm_moments.hip).  part holds
// 8 * TPM_MOMENT_BLOCKS doubles, out 8 (tpm_moments layout, on the device).
constexpr int TPM_MOMENT_BLOCKS = 1024;
int tpm_moment_chunks(int64_t n);
tpm_status tpm_moments_enqueue(std::string& err, hipStream_t s, int64_t n, const float* vx, const float* vy,
                               const float* vz, const float* gx, const float* gy, const float* gz, const float* pot,
                               double m, double* part, double* out);
tpm_status tpm_ctx_moments(tpm_ctx* ctx, tpm_moments* out);

// Batched stages (tpm_chainmesh.hip, tpm_pm.hip, tpm_pp.hip).
tpm_status tpm_batch_chainmesh(tpm_batch* bt, const tpm_batch_layout& L);
tpm_status tpm_batch_pm(tpm_batch* bt, const tpm_batch_layout& L);
tpm_status tpm_batch_pp(tpm_batch* bt, const tpm_batch_layout& L, unsigned flags);
tpm_status tpm_batch_drift(tpm_batch* bt, const tpm_batch_layout& L);
tpm_status tpm_batch_candidates(tpm_batch* bt, const tpm_batch_layout& L);
tpm_status tpm_batch_fill_u64(tpm_batch* bt, unsigned long long* p, int64_t n);
int tpm_pp_persistent_grid(int device, int ts, bool lists);
tpm_status tpm_green_table(std::string& err, hipStream_t s, int M, double rs, int order, float* table);
// Deposit blocking of one box: block edge B (chain cells) and the widest
// footprint in mesh nodes, so that w^3 int64 nodes fit 64 KiB of LDS.
// LDS of one workgroup on gfx950 (160 KiB), the cap of a requested
// deposit block's footprint.
#define TPM_DEPOSIT_LDS_MAX (160 * 1024)
#define TPM_MAX_DEPOSIT_BLOCK 8       // chain cells per deposit block edge
#define TPM_DEPOSIT_STATIC_LDS 1024   // the block's row table, beside its mesh nodes
#ifndef TPM_DEPOSIT_THREADS
#define TPM_DEPOSIT_THREADS 512       // per deposit block (workgroup)
#endif
#define TPM_DEPOSIT_LDS_AUTO (80 * 1024)  // automatic block edge: 2 workgroups per CU
tpm_status tpm_deposit_lds_setup(std::string& err);  // allow the deposit kernels their LDS
void tpm_deposit_blocking(int C, int M, int interlace, int want, int* B, int* wmax);
