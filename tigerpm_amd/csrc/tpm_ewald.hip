// tpm_ewald.hip -- FP64 Ewald sum at sampled particles (the accuracy reference).
//
// Reference: gravity_ewald_kernel (gravity_short.cu:93-201), one block per
// sample over all sources, FP32 per-pair Ewald sums with alpha = 2 (real
// images |n| <= 4, k-space |h|^2 <= 8) and FP64 block reduction.  Here the
// sum is classic Ewald summation organised for many sources: with
// alpha >= 9.6 the real-space part needs the minimum image only
// (rc = 4.8/alpha <= 1/2, erfc(4.8) ~ 6e-12) and the reciprocal part goes
// through the structure factor S(h) = sum_j exp(2 pi i h.x_j) over the half
// space 0 < |h| <= 5.2 alpha/pi.  All arithmetic is FP64; source phases are
// formed exactly in 32-bit fixed point (h.x mod 1) before sincospi.
// Unit masses.  At a sample particle, phi excludes the particle's own
// singular term but includes its image/background term (the convention of
// the CPU restatement); at an arbitrary position every particle counts,
// softened inside hsoft.
#include <mutex>

#include "tpm_internal.hpp"

namespace {

#define EW_CHUNK 512

// S(h): one thread per h vector, all threads of a block stream the same
// source chunk through LDS; partial sums are added with FP64 atomics.
__global__ __launch_bounds__(256) void structure_factor_kernel(int64_t n, const uint32_t* __restrict__ x,
                                                               const uint32_t* __restrict__ y,
                                                               const uint32_t* __restrict__ z, int nh,
                                                               const int4* __restrict__ hl, int64_t per_block,
                                                               double* __restrict__ Sre, double* __restrict__ Sim) {
    __shared__ uint32_t sx[EW_CHUNK], sy[EW_CHUNK], sz[EW_CHUNK];
    const int h = blockIdx.y * blockDim.x + threadIdx.x;
    int4 hv = make_int4(0, 0, 0, 0);
    if (h < nh) hv = hl[h];
    const int64_t j0 = (int64_t)blockIdx.x * per_block;
    const int64_t j1 = min(n, j0 + per_block);
    double re = 0.0, im = 0.0;
    for (int64_t jb = j0; jb < j1; jb += EW_CHUNK) {
        const int cnt = (int)min((int64_t)EW_CHUNK, j1 - jb);
        __syncthreads();
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
            sx[q] = x[jb + q];
            sy[q] = y[jb + q];
            sz[q] = z[jb + q];
        }
        __syncthreads();
        if (h < nh) {
            for (int q = 0; q < cnt; q++) {
                const uint32_t ph = (uint32_t)hv.x * sx[q] + (uint32_t)hv.y * sy[q] + (uint32_t)hv.z * sz[q];
                double s, c;
                sincospi((double)ph * 0x1p-31, &s, &c);
                re += c;
                im += s;
            }
        }
    }
    if (h < nh) {
        atomicAdd(&Sre[h], re);
        atomicAdd(&Sim[h], im);
    }
}

// Reciprocal-space potential and force at each sample.
__global__ void kspace_kernel(int64_t ns, const uint32_t* __restrict__ spx, const uint32_t* __restrict__ spy,
                              const uint32_t* __restrict__ spz, int nh,
                              const int4* __restrict__ hl, const double* __restrict__ coef,
                              const double* __restrict__ Sre, const double* __restrict__ Sim,
                              double* __restrict__ phi, double* __restrict__ gx, double* __restrict__ gy,
                              double* __restrict__ gz) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const uint32_t xi = spx[s], yi = spy[s], zi = spz[s];
    double P = 0.0, G0 = 0.0, G1 = 0.0, G2 = 0.0;
    for (int k = 0; k < nh; k++) {
        const int4 hv = hl[k];
        const uint32_t ph = (uint32_t)hv.x * xi + (uint32_t)hv.y * yi + (uint32_t)hv.z * zi;
        double es, ec;
        sincospi((double)ph * 0x1p-31, &es, &ec);
        const double re = ec * Sre[k] + es * Sim[k];
        const double im = es * Sre[k] - ec * Sim[k];
        const double c = coef[k];
        P -= 2.0 / M_PI * c * re;
        G0 -= 4.0 * c * hv.x * im;
        G1 -= 4.0 * c * hv.y * im;
        G2 -= 4.0 * c * hv.z * im;
    }
    phi[s] += P;
    gx[s] += G0;
    gy[s] += G1;
    gz[s] += G2;
}

// Real space on the minimum image: 64 samples per block (one per lane) over
// a chunk of sources staged in LDS.
// excl[s] (when given) is the particle sitting at sample s, whose own pair
// is left out.
__global__ __launch_bounds__(64) void realspace_kernel(int64_t n, const uint32_t* __restrict__ x,
                                                       const uint32_t* __restrict__ y,
                                                       const uint32_t* __restrict__ z, int64_t ns,
                                                       const uint32_t* __restrict__ spx,
                                                       const uint32_t* __restrict__ spy,
                                                       const uint32_t* __restrict__ spz,
                                                       const int64_t* __restrict__ excl, int64_t per_block,
                                                       double alpha, double hsoft, double* __restrict__ phi,
                                                       double* __restrict__ gx, double* __restrict__ gy,
                                                       double* __restrict__ gz) {
    __shared__ uint32_t sx[EW_CHUNK], sy[EW_CHUNK], sz[EW_CHUNK];
    const int64_t s = (int64_t)blockIdx.y * 64 + threadIdx.x;
    const bool ok = s < ns;
    int64_t i = -1;
    uint32_t xi = 0, yi = 0, zi = 0;
    if (ok) {
        i = excl ? excl[s] : -1;
        xi = spx[s];
        yi = spy[s];
        zi = spz[s];
    }
    const double a = alpha;
    const double rc = 4.8 / a;
    const double rc2 = rc * rc;
    const double h2 = hsoft * hsoft;
    const double tsp = 2.0 / sqrt(M_PI);
    double P = 0.0, G0 = 0.0, G1 = 0.0, G2 = 0.0;
    const int64_t j0 = (int64_t)blockIdx.x * per_block;
    const int64_t j1 = min(n, j0 + per_block);
    for (int64_t jb = j0; jb < j1; jb += EW_CHUNK) {
        const int cnt = (int)min((int64_t)EW_CHUNK, j1 - jb);
        __syncthreads();
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
            sx[q] = x[jb + q];
            sy[q] = y[jb + q];
            sz[q] = z[jb + q];
        }
        __syncthreads();
        if (!ok) continue;
        for (int q = 0; q < cnt; q++) {
            const double dx = (double)(int)(xi - sx[q]) * 0x1p-32;
            const double dy = (double)(int)(yi - sy[q]) * 0x1p-32;
            const double dz = (double)(int)(zi - sz[q]) * 0x1p-32;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 >= rc2 || jb + q == i) continue;
            const double r = sqrt(r2);
            double pp, dr;
            if (r2 < h2) {
                const double q2 = r2 / h2;
                const double d1 = ((15.0 / 8.0 * q2 - 21.0 / 4.0) * q2 + 35.0 / 8.0) / (h2 * hsoft);
                const double d0 = (((-5.0 / 16.0 * q2 + 21.0 / 16.0) * q2 - 35.0 / 16.0) * q2 + 35.0 / 16.0) / hsoft;
                G0 -= d1 * dx;
                G1 -= d1 * dy;
                G2 -= d1 * dz;
                if (r2 > 0.0) {
                    const double e = erf(a * r);
                    pp = e / r - d0;
                    dr = (tsp * a * exp(-a * a * r2) * r - e) / (r2 * r);
                } else {
                    pp = tsp * a - d0;
                    dr = 0.0;
                }
            } else {
                const double ec = erfc(a * r);
                pp = -ec / r;
                dr = (ec / r2 + tsp * a * exp(-a * a * r2) / r) / r;
            }
            P += pp;
            G0 -= dr * dx;
            G1 -= dr * dy;
            G2 -= dr * dz;
        }
    }
    if (ok) {
        atomicAdd(&phi[s], P);
        atomicAdd(&gx[s], G0);
        atomicAdd(&gy[s], G1);
        atomicAdd(&gz[s], G2);
    }
}

__global__ void ewald_init_kernel(int64_t ns, double c, double* phi, double* gx, double* gy, double* gz) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s < ns) {
        phi[s] = c;
        gx[s] = gy[s] = gz[s] = 0.0;
    }
}

// Positions of the sample particles (their device positions through map,
// when the caller's order differs from the device's).
__global__ void sample_pos_kernel(int64_t ns, const uint32_t* __restrict__ map, int64_t* __restrict__ samp,
                                  const uint32_t* __restrict__ x, const uint32_t* __restrict__ y,
                                  const uint32_t* __restrict__ z, uint32_t* __restrict__ spx,
                                  uint32_t* __restrict__ spy, uint32_t* __restrict__ spz) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= ns) return;
    int64_t i = samp[s];
    if (map) i = map[i];
    samp[s] = i;
    spx[s] = x[i];
    spy[s] = y[i];
    spz[s] = z[i];
}

__global__ void fill_f64_kernel(double* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// ---- Work space of tpm_ewald_at_noalloc ----------------------------------
// The wave vectors, their coefficients and S(h) for up to EW_WS_MAX_H
// vectors live in one fixed device array (one copy per GPU, held until the
// process ends) instead of per-call allocations.
#define EW_WS_MAX_H 16384
#define EW_WS_HL 0                                       // int4[EW_WS_MAX_H]
#define EW_WS_COEF (EW_WS_HL + 16 * EW_WS_MAX_H)         // double[EW_WS_MAX_H]
#define EW_WS_S (EW_WS_COEF + 8 * EW_WS_MAX_H)           // double[2 * EW_WS_MAX_H]
#define EW_WS_BYTES (EW_WS_S + 16 * EW_WS_MAX_H)         // 640 KiB
static __device__ __align__(256) unsigned char scratch[EW_WS_BYTES];

// Number of wave vectors 0 < |h| <= hmax in the half space the host loop of
// tpm_ewald_run enumerates.
int ew_count_h(int hmax) {
    int nh = 0;
    for (int hx = 0; hx <= hmax; hx++)
        for (int hy = -hmax; hy <= hmax; hy++)
            for (int hz = -hmax; hz <= hmax; hz++) {
                const int h2 = hx * hx + hy * hy + hz * hz;
                if (h2 == 0 || h2 > hmax * hmax) continue;
                if (hx == 0 && (hy < 0 || (hy == 0 && hz < 0))) continue;
                nh++;
            }
    return nh;
}

// The same vectors in the same order as that host loop, made on the device
// by one workgroup: candidates of the (hmax+1) x (2 hmax+1)^2 box in loop
// order, 256 at a time, compacted with a block-wide scan of ballots.
__global__ __launch_bounds__(256) void ew_hlist_kernel(int hmax, double alpha, int4* __restrict__ hl,
                                                        double* __restrict__ coef) {
    __shared__ int wcount[4];
    const int side = 2 * hmax + 1;
    const int64_t total = (int64_t)(hmax + 1) * side * side;
    const int wave = (int)threadIdx.x >> 6;
    const int lane = (int)threadIdx.x & 63;
    int base = 0;
    for (int64_t c0 = 0; c0 < total; c0 += 256) {
        const int64_t c = c0 + threadIdx.x;
        bool keep = false;
        int hx = 0, hy = 0, hz = 0, h2 = 0;
        if (c < total) {
            hx = (int)(c / ((int64_t)side * side));
            hy = (int)((c / side) % side) - hmax;
            hz = (int)(c % side) - hmax;
            h2 = hx * hx + hy * hy + hz * hz;
            keep = h2 != 0 && h2 <= hmax * hmax && !(hx == 0 && (hy < 0 || (hy == 0 && hz < 0)));
        }
        const unsigned long long m = __ballot(keep);
        if (lane == 0) wcount[wave] = __popcll(m);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; w++) off += wcount[w];
        const int below = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        if (keep) {
            hl[off + below] = make_int4(hx, hy, hz, 0);
            coef[off + below] = exp(-M_PI * M_PI * h2 / (alpha * alpha)) / h2;
        }
        base += wcount[0] + wcount[1] + wcount[2] + wcount[3];
        __syncthreads();
    }
}

}  // namespace

// Device address of this GPU's work space, looked up once per device (with
// that device current) and kept; nullptr (and the message) on failure.
#define EW_MAX_DEVICES 64
static void* ew_scratch_ptr(tpm_ctx* ctx) {
    static std::mutex mu;
    static void* ptr[EW_MAX_DEVICES] = {};
    if (ctx->device < 0 || ctx->device >= EW_MAX_DEVICES) {
        ctx->err = "tpm_ewald_at_noalloc: device index out of range";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[ctx->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            ctx->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        ptr[ctx->device] = p;
    }
    return ptr[ctx->device];
}

tpm_status tpm_ewald_noalloc_run(tpm_ctx* ctx, int64_t ns, const uint32_t* spx, const uint32_t* spy,
                                 const uint32_t* spz, double alpha, double* phi, double* gx, double* gy, double* gz) {
    if (alpha <= 0.0) alpha = 9.6;
    if (alpha < 9.6)
        return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_ewald_at_noalloc: alpha must be >= 9.6 (minimum image only)");
    const int hmax = (int)floor(5.2 * alpha / M_PI);
    const int nh = ew_count_h(hmax);
    if (nh > EW_WS_MAX_H)
        return tpm_fail(ctx, TPM_ERR_INVALID,
                        "tpm_ewald_at_noalloc: alpha = " + std::to_string(alpha) + " needs " + std::to_string(nh) +
                            " wave vectors, more than the work space's " + std::to_string(EW_WS_MAX_H) +
                            " (alpha below 12.08)");
    if (ns > (int64_t)TPM_EWALD_NOALLOC_MAX_SAMPLES)
        return tpm_fail(ctx, TPM_ERR_INVALID,
                        "tpm_ewald_at_noalloc: more than " + std::to_string(TPM_EWALD_NOALLOC_MAX_SAMPLES) + " samples");
    if (ns <= 0) return TPM_OK;
    unsigned char* ws = static_cast<unsigned char*>(ew_scratch_ptr(ctx));
    if (!ws) return TPM_ERR_HIP;
    int4* hl = reinterpret_cast<int4*>(ws + EW_WS_HL);
    double* coef = reinterpret_cast<double*>(ws + EW_WS_COEF);
    double* S = reinterpret_cast<double*>(ws + EW_WS_S);
    hipStream_t st = ctx->stream;
    ew_hlist_kernel<<<1, 256, 0, st>>>(hmax, alpha, hl, coef);
    // S(h) accumulates by atomics: cleared on every call (by a kernel, so the
    // call can be captured in a graph)
    fill_f64_kernel<<<grid_for(2 * nh, 256), 256, 0, st>>>(S, 2 * nh);
    ewald_init_kernel<<<grid_for(ns, 256), 256, 0, st>>>(ns, (double)ctx->n * M_PI / (alpha * alpha), phi, gx, gy,
                                                          gz);
    TPM_HIP(ctx, hipGetLastError());
    const int64_t n = ctx->n;
    if (n > 0) {
        int64_t sf_chunks = std::max<int64_t>(1, std::min<int64_t>(4096, (n + 32767) / 32768));
        const int64_t sf_per = (n + sf_chunks - 1) / sf_chunks;
        sf_chunks = (n + sf_per - 1) / sf_per;
        dim3 g1((unsigned)sf_chunks, (unsigned)((nh + 255) / 256));
        structure_factor_kernel<<<g1, 256, 0, st>>>(n, ctx->x, ctx->y, ctx->z, nh, hl, sf_per, S, S + nh);
        TPM_HIP(ctx, hipGetLastError());
    }
    kspace_kernel<<<grid_for(ns, 64), 64, 0, st>>>(ns, spx, spy, spz, nh, hl, coef, S, S + nh, phi, gx, gy, gz);
    TPM_HIP(ctx, hipGetLastError());
    if (n > 0) {
        const int64_t sblocks = (ns + 63) / 64;
        int64_t rs_chunks =
            std::max<int64_t>(1, std::min<int64_t>((n + 4095) / 4096, (8192 + sblocks - 1) / sblocks));
        const int64_t rs_per = (n + rs_chunks - 1) / rs_chunks;
        rs_chunks = (n + rs_per - 1) / rs_per;
        dim3 g2((unsigned)rs_chunks, (unsigned)sblocks);
        realspace_kernel<<<g2, 64, 0, st>>>(n, ctx->x, ctx->y, ctx->z, ns, spx, spy, spz, nullptr, rs_per, alpha,
                                            ctx->p.hsoft, phi, gx, gy, gz);
        TPM_HIP(ctx, hipGetLastError());
    }
    return TPM_OK;
}

tpm_status tpm_ewald_run(tpm_ctx* ctx, int64_t ns, const int64_t* samples, const uint32_t* const pos[3],
                         double alpha, double* phi, double* gx, double* gy, double* gz) {
    if (ns <= 0) return TPM_OK;
    if (samples)
        for (int64_t s = 0; s < ns; s++)
            if (samples[s] < 0 || samples[s] >= ctx->n)
                return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_ewald: bad sample index");
    if (alpha <= 0.0) alpha = 9.6;
    if (alpha < 9.6) return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_ewald: alpha must be >= 9.6 (minimum image only)");
    const int hmax = (int)floor(5.2 * alpha / M_PI);
    std::vector<int4> hl;
    std::vector<double> coef;
    for (int hx = 0; hx <= hmax; hx++)
        for (int hy = -hmax; hy <= hmax; hy++)
            for (int hz = -hmax; hz <= hmax; hz++) {
                const int h2 = hx * hx + hy * hy + hz * hz;
                if (h2 == 0 || h2 > hmax * hmax) continue;
                if (hx == 0 && (hy < 0 || (hy == 0 && hz < 0))) continue;
                hl.push_back(make_int4(hx, hy, hz, 0));
                coef.push_back(exp(-M_PI * M_PI * h2 / (alpha * alpha)) / h2);
            }
    const int nh = (int)hl.size();
    hipStream_t st = ctx->stream;
    int4* d_hl = nullptr;
    double *d_coef = nullptr, *d_S = nullptr, *d_out = nullptr;
    int64_t* d_samp = nullptr;
    uint32_t* d_pos = nullptr;
    auto cleanup = [&]() {
        if (d_hl) (void)hipFree(d_hl);
        if (d_coef) (void)hipFree(d_coef);
        if (d_S) (void)hipFree(d_S);
        if (d_out) (void)hipFree(d_out);
        if (d_samp) (void)hipFree(d_samp);
        if (d_pos) (void)hipFree(d_pos);
    };
#define EW_HIP(call)                                                   \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) {                                        \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_); \
            cleanup();                                                 \
            return TPM_ERR_HIP;                                        \
        }                                                              \
    } while (0)
    EW_HIP(hipMalloc(&d_hl, sizeof(int4) * nh));
    EW_HIP(hipMalloc(&d_coef, sizeof(double) * nh));
    EW_HIP(hipMalloc(&d_S, sizeof(double) * 2 * nh));
    EW_HIP(hipMalloc(&d_out, sizeof(double) * 4 * ns));
    EW_HIP(hipMalloc(&d_pos, sizeof(uint32_t) * 3 * ns));
    EW_HIP(hipMemcpyAsync(d_hl, hl.data(), sizeof(int4) * nh, hipMemcpyHostToDevice, st));
    EW_HIP(hipMemcpyAsync(d_coef, coef.data(), sizeof(double) * nh, hipMemcpyHostToDevice, st));
    uint32_t *spx = d_pos, *spy = d_pos + ns, *spz = d_pos + 2 * ns;
    if (samples) {
        EW_HIP(hipMalloc(&d_samp, sizeof(int64_t) * ns));
        EW_HIP(hipMemcpyAsync(d_samp, samples, sizeof(int64_t) * ns, hipMemcpyHostToDevice, st));
        const uint32_t* map =
            ctx->host_order == TPM_ORDER_REFERENCE && !ctx->h2d_identity ? ctx->h2d : nullptr;
        sample_pos_kernel<<<grid_for(ns, 256), 256, 0, st>>>(ns, map, d_samp, ctx->x, ctx->y, ctx->z, spx, spy, spz);
    } else {
        for (int d = 0; d < 3; d++)
            EW_HIP(hipMemcpyAsync(d_pos + d * ns, pos[d], sizeof(uint32_t) * ns, hipMemcpyHostToDevice, st));
    }
    fill_f64_kernel<<<grid_for(2 * nh, 256), 256, 0, st>>>(d_S, 2 * nh);
    double* o_phi = d_out;
    double* o_gx = d_out + ns;
    double* o_gy = d_out + 2 * ns;
    double* o_gz = d_out + 3 * ns;
    // background of the n sources; for a sample particle, + 2 alpha / sqrt(pi)
    // restores its image term (its own pair is left out of the real space)
    const double tsp = 2.0 / sqrt(M_PI);
    ewald_init_kernel<<<grid_for(ns, 256), 256, 0, st>>>(
        ns, (double)ctx->n * M_PI / (alpha * alpha) + (samples ? tsp * alpha : 0.0), o_phi, o_gx, o_gy, o_gz);
    EW_HIP(hipGetLastError());
    // S(h): source chunks of per_block particles x h blocks of 256.
    const int64_t n = ctx->n;
    int64_t sf_chunks = std::max<int64_t>(1, std::min<int64_t>(4096, (n + 32767) / 32768));
    const int64_t sf_per = (n + sf_chunks - 1) / sf_chunks;
    sf_chunks = (n + sf_per - 1) / sf_per;
    dim3 g1((unsigned)sf_chunks, (unsigned)((nh + 255) / 256));
    structure_factor_kernel<<<g1, 256, 0, st>>>(n, ctx->x, ctx->y, ctx->z, nh, d_hl, sf_per, d_S, d_S + nh);
    EW_HIP(hipGetLastError());
    kspace_kernel<<<grid_for(ns, 64), 64, 0, st>>>(ns, spx, spy, spz, nh, d_hl, d_coef, d_S, d_S + nh, o_phi, o_gx,
                                                   o_gy, o_gz);
    EW_HIP(hipGetLastError());
    const int64_t sblocks = (ns + 63) / 64;
    int64_t rs_chunks = std::max<int64_t>(1, std::min<int64_t>((n + 4095) / 4096, (8192 + sblocks - 1) / sblocks));
    const int64_t rs_per = (n + rs_chunks - 1) / rs_chunks;
    rs_chunks = (n + rs_per - 1) / rs_per;
    dim3 g2((unsigned)rs_chunks, (unsigned)sblocks);
    realspace_kernel<<<g2, 64, 0, st>>>(n, ctx->x, ctx->y, ctx->z, ns, spx, spy, spz, d_samp, rs_per, alpha,
                                        ctx->p.hsoft, o_phi, o_gx, o_gy, o_gz);
    EW_HIP(hipGetLastError());
    EW_HIP(hipMemcpyAsync(phi, o_phi, sizeof(double) * ns, hipMemcpyDeviceToHost, st));
    EW_HIP(hipMemcpyAsync(gx, o_gx, sizeof(double) * ns, hipMemcpyDeviceToHost, st));
    EW_HIP(hipMemcpyAsync(gy, o_gy, sizeof(double) * ns, hipMemcpyDeviceToHost, st));
    EW_HIP(hipMemcpyAsync(gz, o_gz, sizeof(double) * ns, hipMemcpyDeviceToHost, st));
    EW_HIP(hipStreamSynchronize(st));
#undef EW_HIP
    cleanup();
    return TPM_OK;
}
