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
// Optionally (interlace) a second deposit onto the grid displaced by half a
// cell is combined in k space, cancelling the odd aliases of the assignment.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

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

// 3-D integer weights of one particle: 1-D weights per axis, products in
// Q32 with the truncation deficit added to the centre node, so that every
// particle deposits exactly 2^32.
template <int ORDER>
__device__ __forceinline__ void weights_3d(uint32_t x, uint32_t y, uint32_t z, int M, uint64_t shift, int* bx,
                                           int* by, int* bz, uint64_t w3[ORDER * ORDER * ORDER]) {
    constexpr int NP = ORDER;
    uint64_t wx[4], wy[4], wz[4];
    int q;
    tpm_weights_1d(ORDER, x, M, shift, bx, wx, &q);
    tpm_weights_1d(ORDER, y, M, shift, by, wy, &q);
    tpm_weights_1d(ORDER, z, M, shift, bz, wz, &q);
    const int s = 3 * q - 32;
    uint64_t sum = 0;
    // 1-D weights are at most 2^q <= 2^21, so the products are exact on the
    // 24-bit multipliers
#pragma unroll
    for (int a = 0; a < NP; a++)
#pragma unroll
        for (int b = 0; b < NP; b++) {
            const uint64_t xy = tpm_mul24((uint32_t)wx[a], (uint32_t)wy[b]);
#pragma unroll
            for (int c = 0; c < NP; c++) {
                const uint64_t v = tpm_mul48x24(xy, (uint32_t)wz[c]) >> s;
                w3[(a * NP + b) * NP + c] = v;
                sum += v;
            }
        }
    constexpr int centre = (ORDER == 2) ? 0 : ((1 * NP + 1) * NP + 1);
    w3[centre] += (1ull << 32) - sum;
}

// One thread per particle, ORDER^3 global 64-bit atomics (any particle order).
template <int ORDER>
__global__ __launch_bounds__(256) void deposit_kernel(int64_t n, const uint32_t* __restrict__ x,
                                                      const uint32_t* __restrict__ y,
                                                      const uint32_t* __restrict__ z, int M, uint64_t shift,
                                                      unsigned long long* __restrict__ grid) {
    constexpr int NP = ORDER;
    const int64_t M2 = (int64_t)M * M;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int bx, by, bz;
        uint64_t w3[NP * NP * NP];
        weights_3d<ORDER>(x[i], y[i], z[i], M, shift, &bx, &by, &bz, w3);
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
    uint64_t shift;
    int C, M, B, nb;                      // nb = ceil(C / B) blocks per axis
};

// Mesh nodes [lo, lo + w) touched by particles of chain cells [c0, c1].
__host__ __device__ __forceinline__ void block_nodes(uint64_t xlo, uint64_t xhi_incl, int M, uint64_t shift, int* lo,
                                                     int* w) {
    const int l = (int)((xlo * (uint64_t)M + shift) >> 32) - 1;
    const int h = (int)((xhi_incl * (uint64_t)M + shift) >> 32) + 2;
    *lo = l;
    *w = h - l + 1;
}

// The deposit of block bid (of nb^3) of one box; plane(k) gives split plane
// k < C of the box's chain mesh, cell0 its first global cell.
template <int ORDER, class Plane>
__device__ __forceinline__ void deposit_block(const uint32_t* __restrict__ x, const uint32_t* __restrict__ y,
                                              const uint32_t* __restrict__ z, const int32_t* __restrict__ cell_begin,
                                              int cell0, unsigned long long* __restrict__ grid, uint64_t shift, int C,
                                              int M, int B, int nb, int bid, Plane plane,
                                              unsigned long long* __restrict__ sg) {
    constexpr int NP = ORDER;
    const int bx = bid % nb, by = (bid / nb) % nb, bz = bid / (nb * nb);
    const int c0[3] = {bx * B, by * B, bz * B};
    int lo[3], w[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int c1 = min(c0[d] + B, C) - 1;
        const uint64_t xhi = (c1 + 1 < C ? (uint64_t)plane(c1 + 1) : (1ull << 32)) - 1;
        block_nodes(plane(c0[d]), xhi, M, shift, &lo[d], &w[d]);
    }
    const int c1x = min(c0[0] + B, C) - 1, c1y = min(c0[1] + B, C) - 1, c1z = min(c0[2] + B, C) - 1;
    const int wy = w[1], wz = w[2];
    const int wtot = w[0] * wy * wz;
    // the block's particles: the cells c0x..c1x of each (cy, cz) row are
    // contiguous in memory; the rows are numbered into one index space so
    // that every thread has work until the block's last particle
    __shared__ int row_first[TPM_MAX_DEPOSIT_BLOCK * TPM_MAX_DEPOSIT_BLOCK];
    __shared__ int row_pre[TPM_MAX_DEPOSIT_BLOCK * TPM_MAX_DEPOSIT_BLOCK + 1];
    const int ny = c1y - c0[1] + 1, nrows = ny * (c1z - c0[2] + 1);
    for (int l = threadIdx.x; l < wtot; l += blockDim.x) sg[l] = 0ull;
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int r = 0; r < nrows; r++) {
            const int cz = c0[2] + r / ny, cy = c0[1] + r % ny;
            const int rb = cell_begin[cell0 + (cz * C + cy) * C + c0[0]];
            const int re = cell_begin[cell0 + (cz * C + cy) * C + c1x + 1];
            row_first[r] = rb;
            row_pre[r] = acc;
            acc += re - rb;
        }
        row_pre[nrows] = acc;
    }
    __syncthreads();
    const int total = row_pre[nrows];
    int row = 0;
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
        while (k >= row_pre[row + 1]) row++;
        const int i = row_first[row] + (k - row_pre[row]);
        int bxn, byn, bzn;
        uint64_t w3[NP * NP * NP];
        weights_3d<ORDER>(x[i], y[i], z[i], M, shift, &bxn, &byn, &bzn, w3);
        const int ox = bxn - lo[0], oy = byn - lo[1], oz = bzn - lo[2];
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int r = 0; r < NP; r++)
#pragma unroll
                for (int q = 0; q < NP; q++)
                    atomicAdd(&sg[((ox + p) * wy + (oy + r)) * wz + (oz + q)],
                              (unsigned long long)w3[(p * NP + r) * NP + q]);
    }
    __syncthreads();
    const int64_t M2 = (int64_t)M * M;
    for (int l = threadIdx.x; l < wtot; l += blockDim.x) {
        const unsigned long long v = sg[l];
        if (v == 0ull) continue;
        const int ix = wrapi(lo[0] + l / (wy * wz), M);
        const int iy = wrapi(lo[1] + (l / wz) % wy, M);
        const int iz = wrapi(lo[2] + l % wz, M);
        atomicAdd(&grid[ix * M2 + (int64_t)iy * M + iz], v);
    }
}

template <int ORDER>
__global__ __launch_bounds__(TPM_DEPOSIT_THREADS) void deposit_blocked_kernel(BlockedDepositArgs a) {
    extern __shared__ unsigned long long sg[];
    const uint32_t* bounds = a.bounds;
    deposit_block<ORDER>(a.x, a.y, a.z, a.cell_begin, 0, a.grid, a.shift, a.C, a.M, a.B, a.nb, (int)blockIdx.x,
                         [&](int k) { return bounds[k]; }, sg);
}

// Batched: work item -> (box, block) through the boxes' first items dep0.
template <int ORDER>
__global__ __launch_bounds__(TPM_DEPOSIT_THREADS) void bdeposit_kernel(const uint32_t* __restrict__ x, const uint32_t* __restrict__ y,
                                                       const uint32_t* __restrict__ z,
                                                       const int32_t* __restrict__ cell_begin, int nbox,
                                                       const tpm_box_dev* __restrict__ boxes,
                                                       unsigned long long* __restrict__ grid, uint64_t shift) {
    extern __shared__ unsigned long long sg[];
    const int item = (int)blockIdx.x;
    const int b = tpm_upper_index(nbox, item, [&](int k) { return (int64_t)boxes[k].dep0; });
    const tpm_box_dev& bx = boxes[b];
    const double ninv = bx.ninv;
    deposit_block<ORDER>(x, y, z, cell_begin, bx.cell0, grid + bx.mesh0, shift, bx.C, bx.M, bx.B, bx.nb,
                         item - bx.dep0, [&](int k) { return tpm_split_plane(k, ninv); }, sg);
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

// Interlaced combination fused with the Green's multiply:
// c0 <- G (c0 + exp(i (kx + ky + kz) / 2) c1) / 2.
__global__ void interlace_green_kernel(int M, float2* __restrict__ c0, const float2* __restrict__ c1,
                                       const float* __restrict__ g) {
    const int Mh = M / 2 + 1;
    const int64_t total = (int64_t)M * M * Mh;
    const float cph = (float)(M_PI / M);  // half of 2 pi / M
    for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < total; l += (int64_t)gridDim.x * blockDim.x) {
        const int iz = (int)(l % Mh);
        const int iy = (int)((l / Mh) % M);
        const int ix = (int)(l / ((int64_t)Mh * M));
        const int ksum = (2 * ix < M ? ix : ix - M) + (2 * iy < M ? iy : iy - M) + (2 * iz < M ? iz : iz - M);
        float sn, cs;
        sincosf(cph * (float)ksum, &sn, &cs);
        const float2 a = c0[l], b = c1[l];
        const float gg = 0.5f * g[l];
        float2 v;
        v.x = (a.x + cs * b.x - sn * b.y) * gg;
        v.y = (a.y + sn * b.x + cs * b.y) * gg;
        c0[l] = v;
    }
}

// Batched Green's multiply over the spectrum arena [0, n): each mode finds
// its box's region; modes outside every region are left alone.  With c1 the
// interlaced combination of interlace_green_kernel.
__global__ void bgreen_apply_kernel(int64_t n, int nreg, const tpm_kregion* __restrict__ reg,
                                    float2* __restrict__ c0, const float2* __restrict__ c1,
                                    const float* __restrict__ g) {
    for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
        const int r = tpm_upper_index(nreg, l, [&](int k) { return reg[k].k0; });
        if (r < 0 || l >= reg[r].k0 + reg[r].nk) continue;
        const float gg = g[l];
        float2 a = c0[l];
        if (c1) {
            const int M = reg[r].M;
            const int Mh = M / 2 + 1;
            const int64_t q = l - reg[r].k0;
            const int iz = (int)(q % Mh);
            const int iy = (int)((q / Mh) % M);
            const int ix = (int)(q / ((int64_t)Mh * M));
            const int ksum = (2 * ix < M ? ix : ix - M) + (2 * iy < M ? iy : iy - M) + (2 * iz < M ? iz : iz - M);
            float sn, cs;
            sincosf((float)(M_PI / M) * (float)ksum, &sn, &cs);
            const float2 b = c1[l];
            const float h = 0.5f * gg;
            a = make_float2((a.x + cs * b.x - sn * b.y) * h, (a.y + sn * b.x + cs * b.y) * h);
        } else {
            a.x *= gg;
            a.y *= gg;
        }
        c0[l] = a;
    }
}

}  // namespace

tpm_status tpm_green_table(std::string& err, hipStream_t s, int M, double rs, int order, float* table) {
    const int64_t MC = (int64_t)M * M * (M / 2 + 1);
    green_kernel<<<grid_for(MC, 256), 256, 0, s>>>(M, rs, order, table);
    TPM_HIPE(err, hipGetLastError());
    return TPM_OK;
}

tpm_status tpm_batch_fill_u64(tpm_batch* bt, unsigned long long* p, int64_t n) {
    fill_u64_kernel<<<grid_for(n, 256), 256, 0, bt->stream>>>(p, n);
    TPM_HIP(bt, hipGetLastError());
    return TPM_OK;
}

// Dynamic LDS above 64 KiB must be allowed per kernel: the deposit
// kernels may take up to TPM_DEPOSIT_LDS_MAX (once per process and device;
// every context does it when it is made, before any graph capture).
tpm_status tpm_deposit_lds_setup(std::string& err) {
    static std::mutex mu;
    static bool done[64] = {};
    int dev = 0;
    TPM_HIPE(err, hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if (dev >= 0 && dev < 64 && done[dev]) return TPM_OK;
    const void* k[6] = {(const void*)deposit_blocked_kernel<2>, (const void*)deposit_blocked_kernel<3>,
                        (const void*)deposit_blocked_kernel<4>, (const void*)bdeposit_kernel<2>,
                        (const void*)bdeposit_kernel<3>,        (const void*)bdeposit_kernel<4>};
    for (const void* f : k)
        TPM_HIPE(err, hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          TPM_DEPOSIT_LDS_MAX - TPM_DEPOSIT_STATIC_LDS));
    if (dev >= 0 && dev < 64) done[dev] = true;
    return TPM_OK;
}

// Deposit (grid shift) of every live box into its region of rho_q, then the
// float image of the used part of the arena into out.
static tpm_status batch_deposit(tpm_batch* bt, const tpm_batch_layout& L, uint64_t shift, float* out) {
    hipStream_t s = bt->stream;
    TPM_TRY(tpm_batch_fill_u64(bt, bt->rho_q, L.mesh_hw));
    if (L.ndep > 0) {
        const size_t lds = (size_t)L.lds;
        switch (L.order) {
            case 2:
                bdeposit_kernel<2><<<(unsigned)L.ndep, TPM_DEPOSIT_THREADS, lds, s>>>(bt->x, bt->y, bt->z, bt->cell_begin, L.nbox,
                                                                       bt->d_boxes, bt->rho_q, shift);
                break;
            case 3:
                bdeposit_kernel<3><<<(unsigned)L.ndep, TPM_DEPOSIT_THREADS, lds, s>>>(bt->x, bt->y, bt->z, bt->cell_begin, L.nbox,
                                                                       bt->d_boxes, bt->rho_q, shift);
                break;
            default:
                bdeposit_kernel<4><<<(unsigned)L.ndep, TPM_DEPOSIT_THREADS, lds, s>>>(bt->x, bt->y, bt->z, bt->cell_begin, L.nbox,
                                                                       bt->d_boxes, bt->rho_q, shift);
                break;
        }
        TPM_HIP(bt, hipGetLastError());
    }
    q32_to_float_kernel<<<grid_for(L.mesh_hw, 256), 256, 0, s>>>(L.mesh_hw, bt->rho_q, out);
    TPM_HIP(bt, hipGetLastError());
    return TPM_OK;
}

// Live boxes, largest mesh first (side stream k takes boxes k, k + NSIDE, ...).
static std::vector<const tpm_batch_box*> boxes_by_size(const tpm_batch* bt) {
    std::vector<const tpm_batch_box*> v;
    for (const tpm_batch_box& b : bt->boxes)
        if (b.live) v.push_back(&b);
    std::stable_sort(v.begin(), v.end(), [](const tpm_batch_box* a, const tpm_batch_box* b) { return a->M > b->M; });
    return v;
}

// Forks the side streams off the main stream (they wait for its work so far).
static tpm_status fork_sides(tpm_batch* bt) {
    TPM_HIP(bt, hipEventRecord(bt->ev_fork, bt->stream));
    for (int k = 0; k < tpm_batch::NSIDE; k++) TPM_HIP(bt, hipStreamWaitEvent(bt->side[k], bt->ev_fork, 0));
    return TPM_OK;
}

// The main stream waits for everything the side streams were given.
static tpm_status join_sides(tpm_batch* bt) {
    for (int k = 0; k < tpm_batch::NSIDE; k++) {
        TPM_HIP(bt, hipEventRecord(bt->ev_join[k], bt->side[k]));
        TPM_HIP(bt, hipStreamWaitEvent(bt->stream, bt->ev_join[k], 0));
    }
    return TPM_OK;
}

// One transform per live box on the side streams: R2C from real to
// spectrum (forward) or C2R from spectrum to real.
static tpm_status side_ffts(tpm_batch* bt, const std::vector<const tpm_batch_box*>& v, bool forward, float* real,
                            float2* spec) {
    for (size_t i = 0; i < v.size(); i++) {
        const tpm_batch_box& b = *v[i];
        const int k = (int)(i % tpm_batch::NSIDE);
        tpm_fft& f = bt->side_fft[k];
        TPM_TRY(tpm_fft_use(bt->err, f, bt->device, bt->side[k], b.M));
        if (forward)
            TPM_TRY(tpm_fft_r2c(bt->err, f, real + b.mesh0, spec + b.k0));
        else
            TPM_TRY(tpm_fft_c2r(bt->err, f, spec + b.k0, real + b.mesh0));
    }
    return TPM_OK;
}

tpm_status tpm_batch_pm(tpm_batch* bt, const tpm_batch_layout& L) {
    if (L.mesh_hw == 0) return TPM_OK;
    const auto v = boxes_by_size(bt);
    TPM_TRY(batch_deposit(bt, L, 0, bt->rho));
    TPM_TRY(fork_sides(bt));
    TPM_TRY(side_ffts(bt, v, true, bt->rho, bt->rho_hat));
    if (L.interlace) {
        // the shifted deposit runs on the main stream meanwhile
        TPM_TRY(batch_deposit(bt, L, 1ull << 31, bt->rho1));
        TPM_TRY(fork_sides(bt));
        TPM_TRY(side_ffts(bt, v, true, bt->rho1, bt->rho_hat1));
    }
    TPM_TRY(join_sides(bt));
    bgreen_apply_kernel<<<grid_for(L.k_hw, 256), 256, 0, bt->stream>>>(
        L.k_hw, L.nkreg, bt->d_kreg, bt->rho_hat, L.interlace ? bt->rho_hat1 : nullptr, bt->green);
    TPM_HIP(bt, hipGetLastError());
    TPM_TRY(fork_sides(bt));
    TPM_TRY(side_ffts(bt, v, false, bt->phi, bt->rho_hat));
    TPM_TRY(join_sides(bt));
    return TPM_OK;
}

tpm_status tpm_fill_u64(tpm_ctx* ctx, unsigned long long* p, int64_t n) {
    fill_u64_kernel<<<grid_for(n, 256), 256, 0, ctx->stream>>>(p, n);
    TPM_HIP(ctx, hipGetLastError());
    return TPM_OK;
}

// ---------------------------------------------------------------------------
// rocFFT plans: one pair per (device, M), shared by every caller.
// ---------------------------------------------------------------------------
namespace {
std::mutex g_plan_mu;
std::map<std::tuple<int, int, bool>, tpm_fft_plans*> g_plans;
}  // namespace

const tpm_fft_plans* tpm_fft_plans_get(int device, int M, bool lds, std::string* err) {
    lds = lds && tpm_lds_fft_size(M);
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find({device, M, lds});
    if (it != g_plans.end()) return it->second;
    auto* pl = new tpm_fft_plans();
    if (lds) {
        std::string e;
        if (tpm_lds_fft_twiddles(e, M, &pl->tw) != TPM_OK) {
            delete pl;
            if (err) *err = e;
            return nullptr;
        }
        pl->work_bytes = sizeof(float2) * (size_t)M * M * (M / 2 + 1);
        g_plans[{device, M, lds}] = pl;
        return pl;
    }
    const size_t len[3] = {(size_t)M, (size_t)M, (size_t)M};  // fastest first; cubic
    if (rocfft_plan_create(&pl->fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                           rocfft_precision_single, 3, len, 1, nullptr) != rocfft_status_success ||
        rocfft_plan_create(&pl->inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                           rocfft_precision_single, 3, len, 1, nullptr) != rocfft_status_success) {
        if (pl->fwd) rocfft_plan_destroy(pl->fwd);
        delete pl;
        if (err) *err = "rocfft_plan_create failed for M = " + std::to_string(M);
        return nullptr;
    }
    size_t w1 = 0, w2 = 0;
    rocfft_plan_get_work_buffer_size(pl->fwd, &w1);
    rocfft_plan_get_work_buffer_size(pl->inv, &w2);
    pl->work_bytes = std::max(w1, w2);
    g_plans[{device, M, lds}] = pl;
    return pl;
}

tpm_status tpm_fft_use(std::string& err, tpm_fft& f, int device, hipStream_t s, int M, bool lds) {
    lds = lds && tpm_lds_fft_size(M);
    if (f.M == M && f.plans && (f.plans->tw != nullptr) == lds) return TPM_OK;
    const tpm_fft_plans* pl = tpm_fft_plans_get(device, M, lds, &err);
    if (!pl) return TPM_ERR_FFT;
    if (!f.info) {
        if (rocfft_execution_info_create(&f.info) != rocfft_status_success ||
            rocfft_execution_info_set_stream(f.info, s) != rocfft_status_success) {
            err = "rocfft_execution_info_create failed";
            return TPM_ERR_FFT;
        }
    }
    f.stream = s;
    if (pl->work_bytes > f.work_bytes) {
        if (f.work) TPM_HIPE(err, hipFree(f.work));
        f.work = nullptr;
        TPM_HIPE(err, hipMalloc(&f.work, pl->work_bytes));
        f.work_bytes = pl->work_bytes;
        if (rocfft_execution_info_set_work_buffer(f.info, f.work, f.work_bytes) != rocfft_status_success) {
            err = "rocfft_execution_info_set_work_buffer failed";
            return TPM_ERR_FFT;
        }
    }
    f.M = M;
    f.plans = pl;
    return TPM_OK;
}

tpm_status tpm_fft_r2c(std::string& err, tpm_fft& f, float* in, float2* out) {
    if (f.plans->tw) return tpm_lds_fft_r2c(err, f.stream, f.M, in, false, out, (float2*)f.work, f.plans->tw, nullptr);
    void* ip[1] = {in};
    void* op[1] = {out};
    if (rocfft_execute(f.plans->fwd, ip, op, f.info) != rocfft_status_success) {
        err = "rocfft_execute (forward) failed";
        return TPM_ERR_FFT;
    }
    return TPM_OK;
}

tpm_status tpm_fft_c2r(std::string& err, tpm_fft& f, float2* in, float* out) {
    if (f.plans->tw) return tpm_lds_fft_c2r(err, f.stream, f.M, in, out, (float2*)f.work, f.plans->tw);
    void* ip[1] = {in};
    void* op[1] = {out};
    if (rocfft_execute(f.plans->inv, ip, op, f.info) != rocfft_status_success) {
        err = "rocfft_execute (inverse) failed";
        return TPM_ERR_FFT;
    }
    return TPM_OK;
}

void tpm_fft_free(tpm_fft& f) {
    if (f.info) rocfft_execution_info_destroy(f.info);
    if (f.work) (void)hipFree(f.work);
    f = tpm_fft{};
}

tpm_status tpm_fft_plan(tpm_ctx* ctx, int M) {
    return tpm_fft_use(ctx->err, ctx->fft, ctx->device, ctx->stream, M, !ctx->fft_rocfft);
}

tpm_status tpm_fft_forward(tpm_ctx* ctx, float* in, float2* out) { return tpm_fft_r2c(ctx->err, ctx->fft, in, out); }

tpm_status tpm_fft_inverse(tpm_ctx* ctx, float2* in, float* out) { return tpm_fft_c2r(ctx->err, ctx->fft, in, out); }

tpm_status tpm_ensure_mesh(tpm_ctx* ctx, int M) {
    const int64_t M3 = (int64_t)M * M * M;
    const int64_t MC = (int64_t)M * M * (M / 2 + 1);
    if (M3 > ctx->mesh_cap) {
        void** bufs[] = {(void**)&ctx->rho_q, (void**)&ctx->rho, (void**)&ctx->rho_hat, (void**)&ctx->rho_hat1,
                         (void**)&ctx->phi, (void**)&ctx->green};
        for (void** b : bufs) {
            if (*b) TPM_HIP(ctx, hipFree(*b));
            *b = nullptr;
        }
        TPM_HIP(ctx, hipMalloc(&ctx->rho_q, sizeof(unsigned long long) * M3));
        TPM_HIP(ctx, hipMalloc(&ctx->rho, sizeof(float) * M3));
        TPM_HIP(ctx, hipMalloc(&ctx->rho_hat, sizeof(float2) * MC));
        TPM_HIP(ctx, hipMalloc(&ctx->phi, sizeof(float) * M3));
        TPM_HIP(ctx, hipMalloc(&ctx->green, sizeof(float) * MC));
        ctx->mesh_cap = M3;
        ctx->green_M = 0;
        ctx->green_t_valid = false;
    }
    return tpm_fft_plan(ctx, M);
}

void tpm_deposit_blocking(int C, int M, int interlace, int want, int* Bout, int* wout) {
    // the largest block edge (cells, <= want or 3) whose footprint of int64
    // nodes fits the LDS of one workgroup: 80 KiB when chosen automatically
    // (want = 0: two workgroups per CU), up to gfx950's 160 KiB when the
    // caller asks for a block edge (fewer, larger blocks: less halo traffic).
    // At 512^3 (5.6 mesh cells per chain cell) the automatic edge is 3
    // cells: 2.5 ms against 3.0 ms for 2-cell blocks and 2.9-3.2 ms for 4.
    const int64_t lds_cap = (want > 0 ? TPM_DEPOSIT_LDS_MAX : TPM_DEPOSIT_LDS_AUTO) - TPM_DEPOSIT_STATIC_LDS;
    const double ninv = 1.0 / C;
    std::vector<uint64_t> b(C + 1);
    for (int k = 0; k < C; k++) b[k] = tpm_split_plane(k, ninv);
    b[C] = 1ull << 32;
    *Bout = 0;
    *wout = 0;
    for (int cand = want > 0 ? want : 3; cand >= 1; cand--) {
        int wm = 0;
        for (int sh = 0; sh <= (interlace ? 1 : 0); sh++)
            for (int c0 = 0; c0 < C; c0 += cand) {
                const int c1 = std::min(c0 + cand, C) - 1;
                int lo, w;
                block_nodes(b[c0], b[c1 + 1] - 1, M, sh ? (1ull << 31) : 0ull, &lo, &w);
                wm = std::max(wm, w);
            }
        if ((int64_t)wm * wm * wm * 8 <= lds_cap) {
            *Bout = cand;
            *wout = wm;
            return;
        }
    }
}

// Integer deposit (grid displaced by shift / 2^32 cells) into rho_q and,
// with to_float, its float image into rho.
static tpm_status deposit_density(tpm_ctx* ctx, uint64_t shift, bool to_float = true) {
    const int M = ctx->M;
    const int64_t M3 = (int64_t)M * M * M;
    hipStream_t s = ctx->stream;
    TPM_TRY(tpm_fill_u64(ctx, ctx->rho_q, M3));
    // Blocked deposit needs particles in chain-mesh order.
    int B = 0, wmax = 0;
    if (ctx->n > 0 && ctx->chain_valid && ctx->chain_C == ctx->C && ctx->deposit_block >= 0)
        tpm_deposit_blocking(ctx->C, M, ctx->p.interlace, ctx->deposit_block, &B, &wmax);
    if (B > 0) {
        BlockedDepositArgs da;
        da.x = ctx->x;
        da.y = ctx->y;
        da.z = ctx->z;
        da.cell_begin = ctx->cell_begin;
        da.bounds = ctx->bounds;
        da.grid = ctx->rho_q;
        da.shift = shift;
        da.C = ctx->C;
        da.M = M;
        da.B = B;
        da.nb = (ctx->C + B - 1) / B;
        const unsigned nblk = (unsigned)((int64_t)da.nb * da.nb * da.nb);
        const size_t lds = sizeof(unsigned long long) * (size_t)wmax * wmax * wmax;
        switch (ctx->p.order) {
            case 2: deposit_blocked_kernel<2><<<nblk, TPM_DEPOSIT_THREADS, lds, s>>>(da); break;
            case 3: deposit_blocked_kernel<3><<<nblk, TPM_DEPOSIT_THREADS, lds, s>>>(da); break;
            default: deposit_blocked_kernel<4><<<nblk, TPM_DEPOSIT_THREADS, lds, s>>>(da); break;
        }
        TPM_HIP(ctx, hipGetLastError());
    } else if (ctx->n > 0) {
        const unsigned g = grid_for(ctx->n, 256);
        switch (ctx->p.order) {
            case 2: deposit_kernel<2><<<g, 256, 0, s>>>(ctx->n, ctx->x, ctx->y, ctx->z, M, shift, ctx->rho_q); break;
            case 3: deposit_kernel<3><<<g, 256, 0, s>>>(ctx->n, ctx->x, ctx->y, ctx->z, M, shift, ctx->rho_q); break;
            default: deposit_kernel<4><<<g, 256, 0, s>>>(ctx->n, ctx->x, ctx->y, ctx->z, M, shift, ctx->rho_q); break;
        }
        TPM_HIP(ctx, hipGetLastError());
    }
    if (to_float) {
        q32_to_float_kernel<<<grid_for(M3, 256), 256, 0, s>>>(M3, ctx->rho_q, ctx->rho);
        TPM_HIP(ctx, hipGetLastError());
    }
    return TPM_OK;
}

tpm_status tpm_pm_solve(tpm_ctx* ctx) {
    if (!ctx->params_set) return tpm_fail(ctx, TPM_ERR_STATE, "tpm_gravity_long_compute: parameters not set");
    const int M = ctx->M;
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
        ctx->green_t_valid = false;
    }
    if (ctx->fft.plans->tw && !ctx->p.interlace) {
        // LDS-tiled transform: the Q32 conversion inside its first pass, the
        // Green's multiply between the forward and inverse x transforms of
        // one pass (same arithmetic as the separate kernels), its factors
        // read from a copy of the table grouped by that pass's workgroups
        if (!ctx->green_t_valid) {
            int64_t n = 0;
            TPM_TRY(tpm_lds_green_tiles(ctx->err, s, M, nullptr, nullptr, &n));
            if (n > ctx->green_t_cap) {
                if (ctx->green_t) TPM_HIP(ctx, hipFree(ctx->green_t));
                ctx->green_t = nullptr;
                ctx->green_t_cap = 0;
                TPM_HIP(ctx, hipMalloc(&ctx->green_t, sizeof(float) * n));
                ctx->green_t_cap = n;
            }
            TPM_TRY(tpm_lds_green_tiles(ctx->err, s, M, ctx->green, ctx->green_t, &n));
            ctx->green_t_valid = true;
        }
        TPM_TRY(deposit_density(ctx, 0, false));
        TPM_TRY(tpm_lds_pm_solve(ctx->err, s, M, ctx->rho_q, ctx->phi, ctx->rho_hat, (float2*)ctx->fft.work,
                                 ctx->fft.plans->tw, ctx->green_t));
        ctx->pm_valid = true;
        return TPM_OK;
    }
    TPM_TRY(deposit_density(ctx, 0));
    TPM_TRY(tpm_fft_forward(ctx, ctx->rho, ctx->rho_hat));
    if (ctx->p.interlace) {
        if (!ctx->rho_hat1 || ctx->rho_hat1_cap < MC) {
            if (ctx->rho_hat1) TPM_HIP(ctx, hipFree(ctx->rho_hat1));
            TPM_HIP(ctx, hipMalloc(&ctx->rho_hat1, sizeof(float2) * MC));
            ctx->rho_hat1_cap = MC;
        }
        TPM_TRY(deposit_density(ctx, 1ull << 31));
        TPM_TRY(tpm_fft_forward(ctx, ctx->rho, ctx->rho_hat1));
        interlace_green_kernel<<<grid_for(MC, 256), 256, 0, s>>>(M, ctx->rho_hat, ctx->rho_hat1, ctx->green);
    } else {
        green_apply_kernel<<<grid_for(MC, 256), 256, 0, s>>>(MC, ctx->rho_hat, ctx->green);
    }
    TPM_HIP(ctx, hipGetLastError());
    TPM_TRY(tpm_fft_inverse(ctx, ctx->rho_hat, ctx->phi));
    ctx->pm_valid = true;
    return TPM_OK;
}

tpm_status tpm_fft_release(tpm_ctx* ctx) {
    tpm_fft_free(ctx->fft);
    return TPM_OK;
}
