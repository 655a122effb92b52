/*
 * tpm_oracle.c -- CPU restatement of the tigerpm TreePM kick path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product (tigerpm_amd/csrc) and the timed CPU baseline; the product never
 * links, loads or calls it.  Every function cites the reference file:line
 * whose semantics it restates (paths relative to the reference checkout).
 *
 * Build: oracle/Makefile (gcc -O2 -std=c11 -ffp-contract=off -fopenmp).
 * FP contraction is off so that every FP32 expression is evaluated exactly
 * as written; integer stages (fixed point, chain mesh, deposit) are exact.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#include <omp.h>

#define ORC_API __attribute__((visibility("default")))

typedef struct {
    int mesh_dim;   /* M: PM mesh cells per side (reference four_dim)            */
    int chain_dim;  /* C: chain-mesh cells per side (reference chain_dim)        */
    int order;      /* mass assignment: 2 CIC, 3 TSC, 4 cloud4 (cubic B-spline)  */
    int do_phi;     /* accumulate the potential (reference do_phi)               */
    int interlace;  /* second deposit on the half-cell-shifted grid (not in the reference) */
    int pme;        /* 1: kick_pme's pair law (no cutoff, plain softening, no phi0) */
    double rs;      /* split scale, box units                                    */
    double rcut;    /* short-range cutoff, box units                             */
    double hsoft;   /* Plummer-polynomial softening length, box units            */
    double GM;      /* multiplies g and phi (reference GM)                       */
    double eta;     /* time-step accuracy parameter                              */
    double n_total; /* number of unit masses in the box (mean density)          */
} orc_params;

/* ------------------------------------------------------------------------ */
/* a1: fixed-point positions                                                */
/* ------------------------------------------------------------------------ */

/* fixed.hpp:63-75: raw = T(float(2^32) * x), truncating conversion. */
ORC_API uint32_t orc_fixed_from_double(double x) {
    return (uint32_t)(4294967296.0 * x);
}

/* fixed.hpp:115-118: float(i) * 2^-32. */
ORC_API float orc_fixed_to_float(uint32_t raw) {
    return (float)raw * 0x1p-32f;
}

/* fixed.hpp:126-129. */
ORC_API double orc_fixed_to_double(uint32_t raw) {
    return (double)raw * 0x1p-32;
}

/* gravity_short.hpp:20-22: periodic minimum image through int32 wrap. */
ORC_API float orc_distance(uint32_t a, uint32_t b) {
    return (float)(int32_t)(a - b) * 0x1p-32f;
}

/* ------------------------------------------------------------------------ */
/* a3: chain mesh                                                           */
/* ------------------------------------------------------------------------ */

/* Split planes of the recursive chain-mesh partition.  chainmesh.cpp:198-199
 * takes mid_x = mid_box * (1.0 / N) in double and particles.cpp:417 turns it
 * into fixed32 (truncation), so plane k sits at raw value b[k].  b[C] = 2^32. */
ORC_API void orc_chain_bounds(int C, uint64_t* b) {
    const double ninv = 1.0 / C;
    for (int k = 0; k < C; k++) {
        b[k] = (uint64_t)(uint32_t)(4294967296.0 * ((double)k * ninv));
    }
    b[C] = 1ull << 32;
}

/* Chain cell coordinate along one axis: the k with b[k] <= raw < b[k+1],
 * which is where the reference's partition (particles.cpp:414-445, left iff
 * x < xmid) leaves the particle.  Also returns the 2-bit sub-cell used to
 * order particles inside a cell (the device sort's deterministic in-cell order). */
static inline int chain_coord(uint32_t raw, int C, const uint64_t* b, int* sub) {
    const uint64_t P = (uint64_t)raw * (uint64_t)C;
    int k = (int)(P >> 32);
    int s = (int)((uint32_t)P >> 30);
    if (k + 1 < C && (uint64_t)raw >= b[k + 1]) {
        k++;
        s = 0;
    } else if ((uint64_t)raw < b[k]) {
        k--;
        s = 3;
    }
    *sub = s;
    return k;
}

ORC_API int orc_chain_coord(uint32_t raw, int C) {
    uint64_t* b = malloc(sizeof(uint64_t) * (C + 1));
    orc_chain_bounds(C, b);
    int s;
    int k = chain_coord(raw, C, b, &s);
    free(b);
    return k;
}

/* Sort key: cell id in the reference's memory order (z slowest: the
 * partition splits the last axis with span > 1 first, chainmesh.cpp:189-197)
 * times 64, plus the Morton code of the 4x4x4 sub-cell. */
ORC_API void orc_chain_keys(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z,
                            int C, uint32_t* key) {
    uint64_t* b = malloc(sizeof(uint64_t) * (C + 1));
    orc_chain_bounds(C, b);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        int sx, sy, sz;
        const int cx = chain_coord(x[i], C, b, &sx);
        const int cy = chain_coord(y[i], C, b, &sy);
        const int cz = chain_coord(z[i], C, b, &sz);
        const uint32_t cid = ((uint32_t)cz * C + cy) * C + cx;
        const uint32_t m = ((sz >> 1) << 5) | ((sy >> 1) << 4) | ((sx >> 1) << 3) | ((sz & 1) << 2) |
                           ((sy & 1) << 1) | (sx & 1);
        key[i] = cid * 64u + m;
    }
    free(b);
}

/* Stable counting sort of particles by key.  perm[new] = old;
 * cell_begin[cid] (C^3 + 1 entries) are the reference's chaincell pbegin
 * values (chainmesh.hpp:8-11), cell_begin[cid+1] its pend. */
ORC_API void orc_chainmesh_sort(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z,
                                int C, int64_t* perm, int64_t* cell_begin) {
    uint32_t* key = malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    orc_chain_keys(n, x, y, z, C, key);
    const int64_t ncell = (int64_t)C * C * C;
    const int64_t nkey = ncell * 64;
    int64_t* cnt = calloc((size_t)nkey + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) cnt[key[i] + 1]++;
    for (int64_t k = 0; k < nkey; k++) cnt[k + 1] += cnt[k];
    for (int64_t c = 0; c <= ncell; c++) cell_begin[c] = cnt[c * 64];
    for (int64_t i = 0; i < n; i++) perm[cnt[key[i]]++] = i;
    free(cnt);
    free(key);
}

/* Restatement of the reference's recursive in-place partition
 * (chainmesh.cpp:168-213 with particles_sort, particles.cpp:414-445) over
 * an index array.  Used only to pin that the sorted order above yields the
 * reference's cell ranges and per-cell particle sets. */
typedef struct {
    const uint32_t* X[3];
    int64_t* idx;
    int C;
    const uint64_t* b;
    int64_t* cell_begin;
    int64_t* cell_end;
} part_ctx;

static int64_t ref_particles_sort(part_ctx* pc, int64_t begin, int64_t end, uint32_t xmid, int xdim) {
    int64_t lo = begin;
    int64_t hi = end;
    const uint32_t* xd = pc->X[xdim];
    while (lo < hi) {
        if (xd[pc->idx[lo]] >= xmid) {
            while (lo != hi) {
                hi--;
                if (xd[pc->idx[hi]] < xmid) {
                    int64_t t = pc->idx[hi];
                    pc->idx[hi] = pc->idx[lo];
                    pc->idx[lo] = t;
                    break;
                }
            }
        }
        lo++;
    }
    return hi;
}

static void ref_chain_sort(part_ctx* pc, const int bb[3], const int be[3], int64_t pbegin, int64_t pend) {
    const int vol = (be[0] - bb[0]) * (be[1] - bb[1]) * (be[2] - bb[2]);
    if (vol == 1) {
        const int C = pc->C;
        const int64_t cid = ((int64_t)bb[2] * C + bb[1]) * C + bb[0];
        pc->cell_begin[cid] = pbegin;
        pc->cell_end[cid] = pend;
        return;
    }
    int long_dim = 0, long_span = -1;
    for (int d = 0; d < 3; d++) {
        const int span = be[d] - bb[d];
        if (span > 1) {
            long_span = span;
            long_dim = d;
        }
    }
    const int mid_box = bb[long_dim] + long_span / 2;
    const uint32_t xmid = (uint32_t)pc->b[mid_box];
    int lb[3], le[3], rb[3], re[3];
    for (int d = 0; d < 3; d++) {
        lb[d] = rb[d] = bb[d];
        le[d] = re[d] = be[d];
    }
    le[long_dim] = rb[long_dim] = mid_box;
    const int64_t pmid = ref_particles_sort(pc, pbegin, pend, xmid, long_dim);
    ref_chain_sort(pc, lb, le, pbegin, pmid);
    ref_chain_sort(pc, rb, re, pmid, pend);
}

ORC_API void orc_chainmesh_partition_ref(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z,
                                         int C, int64_t* idx, int64_t* cell_begin, int64_t* cell_end) {
    uint64_t* b = malloc(sizeof(uint64_t) * (C + 1));
    orc_chain_bounds(C, b);
    part_ctx pc = {{x, y, z}, idx, C, b, cell_begin, cell_end};
    for (int64_t i = 0; i < n; i++) idx[i] = i;
    const int bb[3] = {0, 0, 0};
    const int be[3] = {C, C, C};
    ref_chain_sort(&pc, bb, be, 0, n);
    free(b);
}

/* ------------------------------------------------------------------------ */
/* a4: mass assignment with integer weights                                 */
/* ------------------------------------------------------------------------ */

/* One-axis weights of a particle at raw on an M-node periodic mesh (nodes at
 * j/M, the grid of gravity_long.cpp:166-194).  Integer arithmetic only:
 *   P = raw*M (32 fractional bits).
 *   CIC: nodes I, I+1; weights (1-t, t), t in Q21.
 *   TSC: nearest node j0 = round(P); d = P - j0 in [-1/2, 1/2);
 *        w(+-1) = (1/2 +- d)^2 / 2, w(0) = 1 - w(-1) - w(+1), Q21.
 *   cloud4 (gravity_long.cpp:127-140, cubic B-spline): nodes I-1..I+2,
 *        Q16 so that the cubic fits 64 bits.
 * Returns the first node index (may be -1 or M; wrapped by the caller), the
 * number of weights and the weight precision in bits. */
static inline int weights_1d(int order, uint32_t raw, int M, uint64_t shift, int64_t* base, uint64_t w[4],
                             int* qbits) {
    /* shift = 2^31 deposits onto the grid displaced by half a cell (node j
     * at (j - 1/2) / M), the second grid of an interlaced deposit. */
    const uint64_t P = (uint64_t)raw * (uint64_t)M + shift;
    if (order == 2) {
        const uint64_t t = (uint32_t)P >> 11;
        *base = (int64_t)(P >> 32);
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
        *base = (int64_t)j0 - 1;
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
        *base = (int64_t)(P >> 32) - 1;
        *qbits = 16;
        return 4;
    }
}

ORC_API int orc_weights_1d(int order, uint32_t raw, int M, int64_t* base, uint64_t* w, int* qbits) {
    return weights_1d(order, raw, M, 0, base, w, qbits);
}

/* 3-D weights in Q32 (2^32 = one unit mass).  Products are truncated to Q32
 * and the truncation deficit is added to the centre node, so every particle
 * deposits exactly 2^32 and the grid total is exactly n * 2^32. */
static inline int weights_3d(int order, uint32_t x, uint32_t y, uint32_t z, int M, uint64_t grid_shift,
                             int64_t base[3], uint64_t w3[64]) {
    uint64_t wx[4], wy[4], wz[4];
    int q;
    const int np = weights_1d(order, x, M, grid_shift, &base[0], wx, &q);
    weights_1d(order, y, M, grid_shift, &base[1], wy, &q);
    weights_1d(order, z, M, grid_shift, &base[2], wz, &q);
    const int shift = 3 * q - 32;
    uint64_t sum = 0;
    for (int a = 0; a < np; a++)
        for (int b2 = 0; b2 < np; b2++)
            for (int c = 0; c < np; c++) {
                const uint64_t v = (wx[a] * wy[b2] * wz[c]) >> shift;
                w3[(a * np + b2) * np + c] = v;
                sum += v;
            }
    const int centre = (order == 2) ? 0 : ((1 * np + 1) * np + 1);
    w3[centre] += (1ull << 32) - sum;
    return np;
}

/* compute_source (gravity_long.cpp:142-199) with integer accumulation.
 * grid is [x][y][z], z fastest (range::index, range.hpp:149-165). */
ORC_API void orc_deposit_shifted(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int M,
                                 int order, uint64_t shift, uint64_t* grid) {
    const int64_t M3 = (int64_t)M * M * M;
    memset(grid, 0, sizeof(uint64_t) * (size_t)M3);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        int64_t base[3];
        uint64_t w3[64];
        const int np = weights_3d(order, x[i], y[i], z[i], M, shift, base, w3);
        for (int a = 0; a < np; a++) {
            const int64_t ix = (base[0] + a + M) % M;
            for (int b2 = 0; b2 < np; b2++) {
                const int64_t iy = (base[1] + b2 + M) % M;
                for (int c = 0; c < np; c++) {
                    const int64_t iz = (base[2] + c + M) % M;
                    __atomic_fetch_add(&grid[(ix * M + iy) * M + iz], w3[(a * np + b2) * np + c],
                                       __ATOMIC_RELAXED);
                }
            }
        }
    }
}

ORC_API void orc_deposit(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int M, int order,
                         uint64_t* grid) {
    orc_deposit_shifted(n, x, y, z, M, order, 0, grid);
}

/* Mass per node in unit masses: float(double(sum) * 2^-32) (both roundings
 * are IEEE round-to-nearest, so CPU and GPU agree bit for bit). */
ORC_API void orc_density_float(int64_t M3, const uint64_t* grid, float* rho) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < M3; i++) rho[i] = (float)((double)grid[i] * 0x1p-32);
}

/* ------------------------------------------------------------------------ */
/* a5: FFT (double precision, any length: radix-2 or Bluestein)             */
/* ------------------------------------------------------------------------ */

typedef double complex cplx;

static void fft_pow2(cplx* a, int n, int sign) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            cplx t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const double ang = sign * 2.0 * M_PI / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < len / 2; k++) {
                const cplx w = cos(ang * k) + I * sin(ang * k);
                const cplx u = a[i + k];
                const cplx v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
        }
    }
}

/* Unnormalized DFT X[k] = sum_j a[j] exp(sign * 2 pi i j k / n). */
static void fft_any(cplx* a, int n, int sign, cplx* scratch) {
    if ((n & (n - 1)) == 0) {
        fft_pow2(a, n, sign);
        return;
    }
    int m = 1;
    while (m < 2 * n - 1) m <<= 1;
    cplx* A = scratch;
    cplx* B = scratch + m;
    for (int j = 0; j < m; j++) A[j] = B[j] = 0;
    for (int j = 0; j < n; j++) {
        const double t = M_PI * (double)(((int64_t)j * j) % (2 * n)) / n;
        const cplx wj = cos(t) + I * sign * sin(t);
        A[j] = a[j] * wj;
        B[j] = conj(wj);
        if (j) B[m - j] = conj(wj);
    }
    fft_pow2(A, m, -1);
    fft_pow2(B, m, -1);
    for (int j = 0; j < m; j++) A[j] *= B[j];
    fft_pow2(A, m, +1);
    for (int k = 0; k < n; k++) {
        const double t = M_PI * (double)(((int64_t)k * k) % (2 * n)) / n;
        const cplx wk = cos(t) + I * sign * sin(t);
        a[k] = A[k] / m * wk;
    }
}

static int scratch_len(int n) {
    int m = 1;
    while (m < 2 * n - 1) m <<= 1;
    return 2 * m + n;
}

/* In-place 3-D transform of a [x][y][z] complex array. */
static void fft3d(cplx* a, int M, int sign) {
    const int64_t M2 = (int64_t)M * M;
#pragma omp parallel
    {
        cplx* line = malloc(sizeof(cplx) * (size_t)scratch_len(M));
        cplx* scr = line + M;
#pragma omp for schedule(static)
        for (int64_t l = 0; l < M2; l++) fft_any(a + l * M, M, sign, scr);
#pragma omp for schedule(static)
        for (int64_t l = 0; l < M2; l++) {
            const int64_t ix = l / M, iz = l % M;
            for (int j = 0; j < M; j++) line[j] = a[(ix * M + j) * M + iz];
            fft_any(line, M, sign, scr);
            for (int j = 0; j < M; j++) a[(ix * M + j) * M + iz] = line[j];
        }
#pragma omp for schedule(static)
        for (int64_t l = 0; l < M2; l++) {
            const int64_t iy = l / M, iz = l % M;
            for (int j = 0; j < M; j++) line[j] = a[((int64_t)j * M + iy) * M + iz];
            fft_any(line, M, sign, scr);
            for (int j = 0; j < M; j++) a[((int64_t)j * M + iy) * M + iz] = line[j];
        }
        free(line);
    }
}

/* Forward real-to-complex transform, output [x][y][0..M/2] (the layout of
 * the reference's complex box after fft3d_execute, fft.cpp:51-67), as
 * interleaved (re, im) doubles. */
ORC_API void orc_fft3d_r2c(int M, const double* in, double* out) {
    const int64_t M3 = (int64_t)M * M * M;
    const int Mh = M / 2 + 1;
    cplx* a = malloc(sizeof(cplx) * (size_t)M3);
    for (int64_t i = 0; i < M3; i++) a[i] = in[i];
    fft3d(a, M, -1);
    for (int64_t ix = 0; ix < M; ix++)
        for (int64_t iy = 0; iy < M; iy++)
            for (int iz = 0; iz < Mh; iz++) {
                const cplx v = a[(ix * M + iy) * M + iz];
                const int64_t o = (ix * M + iy) * Mh + iz;
                out[2 * o] = creal(v);
                out[2 * o + 1] = cimag(v);
            }
    free(a);
}

/* Unnormalized complex-to-real inverse (fft.cpp:69-84 without the 1/N,
 * which the device code folds into the Green's table). */
ORC_API void orc_fft3d_c2r(int M, const double* in, double* out) {
    const int64_t M3 = (int64_t)M * M * M;
    const int Mh = M / 2 + 1;
    cplx* a = malloc(sizeof(cplx) * (size_t)M3);
    for (int64_t ix = 0; ix < M; ix++)
        for (int64_t iy = 0; iy < M; iy++)
            for (int64_t iz = 0; iz < M; iz++) {
                cplx v;
                if (iz < Mh) {
                    const int64_t o = (ix * M + iy) * Mh + iz;
                    v = in[2 * o] + I * in[2 * o + 1];
                } else {
                    const int64_t jx = (M - ix) % M, jy = (M - iy) % M, jz = M - iz;
                    const int64_t o = (jx * M + jy) * Mh + jz;
                    v = in[2 * o] - I * in[2 * o + 1];
                }
                a[(ix * M + iy) * M + iz] = v;
            }
    fft3d(a, M, +1);
    for (int64_t i = 0; i < M3; i++) out[i] = creal(a[i]);
    free(a);
}

/* ------------------------------------------------------------------------ */
/* a6: Green's function table                                               */
/* ------------------------------------------------------------------------ */

static double sinc(double x) { return x != 0.0 ? sin(x) / x : 1.0; }

/* apply_laplacian (gravity_long.cpp:201-245) in FP64, folded with the
 * deposit factor 4 pi M (gravity_long.cpp:185) and the inverse-FFT 1/M^3
 * (fft.cpp:468, 512):
 *   G(k) = -4 pi M / M^3 / k^2 * exp(-k^2 (rs M)^2) * prod_i sinc(k_i/2)^-p,
 *   k_i = 2 pi (i < M/2 ? i : i - M) / M, G(0) = 0, p = assignment order. */
ORC_API void orc_green_table(int M, double rs, int order, float* table) {
    const int Mh = M / 2 + 1;
    const double c0 = 2.0 * M_PI / M;
    const double rs2 = (rs * M) * (rs * M);
    const double scale = 4.0 * M_PI * M / ((double)M * M * M);
#pragma omp parallel for schedule(static)
    for (int64_t ix = 0; ix < M; ix++) {
        for (int64_t iy = 0; iy < M; iy++) {
            for (int64_t iz = 0; iz < Mh; iz++) {
                const double kx = c0 * (double)(2 * ix < M ? ix : ix - M);
                const double ky = c0 * (double)(2 * iy < M ? iy : iy - M);
                const double kz = c0 * (double)(2 * iz < M ? iz : iz - M);
                const double k2 = kx * kx + ky * ky + kz * kz;
                double g = 0.0;
                if (k2 > 0.0) {
                    const double s = sinc(0.5 * kx) * sinc(0.5 * ky) * sinc(0.5 * kz);
                    g = -scale / k2 * exp(-k2 * rs2) * pow(s, -(double)order);
                }
                table[(ix * M + iy) * Mh + iz] = (float)g;
            }
        }
    }
}

/* Potential mesh: phi = C2R(table * R2C(rho)), in double, rounded to float. */
ORC_API void orc_pm_phi(int M, const float* rho, const float* table, float* phi) {
    const int64_t M3 = (int64_t)M * M * M;
    const int64_t MC = (int64_t)M * M * (M / 2 + 1);
    double* r = malloc(sizeof(double) * (size_t)M3);
    double* c = malloc(sizeof(double) * 2 * (size_t)MC);
    for (int64_t i = 0; i < M3; i++) r[i] = rho[i];
    orc_fft3d_r2c(M, r, c);
    for (int64_t i = 0; i < MC; i++) {
        c[2 * i] *= table[i];
        c[2 * i + 1] *= table[i];
    }
    orc_fft3d_c2r(M, c, r);
    for (int64_t i = 0; i < M3; i++) phi[i] = (float)r[i];
    free(r);
    free(c);
}

/* Interlaced potential: rho1 is the deposit onto the grid displaced by half
 * a cell along every axis (node j at (j - 1/2)/M).  A mode aliased by an
 * integer vector n carries (-1)^(nx+ny+nz) on that grid relative to the
 * unshifted one, so rho_hat = (rho0_hat + exp(i (kx+ky+kz)/2) rho1_hat) / 2
 * cancels the aliases with odd nx+ny+nz. */
ORC_API void orc_pm_phi_interlaced(int M, const float* rho0, const float* rho1, const float* table, float* phi) {
    const int64_t M3 = (int64_t)M * M * M;
    const int Mh = M / 2 + 1;
    const int64_t MC = (int64_t)M * M * Mh;
    double* r = malloc(sizeof(double) * (size_t)M3);
    double* c0 = malloc(sizeof(double) * 2 * (size_t)MC);
    double* c1 = malloc(sizeof(double) * 2 * (size_t)MC);
    for (int64_t i = 0; i < M3; i++) r[i] = rho0[i];
    orc_fft3d_r2c(M, r, c0);
    for (int64_t i = 0; i < M3; i++) r[i] = rho1[i];
    orc_fft3d_r2c(M, r, c1);
    const double c = 2.0 * M_PI / M;
    for (int64_t ix = 0; ix < M; ix++)
        for (int64_t iy = 0; iy < M; iy++)
            for (int64_t iz = 0; iz < Mh; iz++) {
                const int64_t l = (ix * M + iy) * Mh + iz;
                const double th = 0.5 * c *
                                  (double)((2 * ix < M ? ix : ix - M) + (2 * iy < M ? iy : iy - M) +
                                           (2 * iz < M ? iz : iz - M));
                const double cs = cos(th), sn = sin(th);
                const double re = 0.5 * (c0[2 * l] + cs * c1[2 * l] - sn * c1[2 * l + 1]);
                const double im = 0.5 * (c0[2 * l + 1] + sn * c1[2 * l] + cs * c1[2 * l + 1]);
                c0[2 * l] = re * table[l];
                c0[2 * l + 1] = im * table[l];
            }
    orc_fft3d_c2r(M, c0, r);
    for (int64_t i = 0; i < M3; i++) phi[i] = (float)r[i];
    free(r);
    free(c0);
    free(c1);
}

/* ------------------------------------------------------------------------ */
/* a8: 6-point interpolation (kick_treepm.cu:505-573)                        */
/* ------------------------------------------------------------------------ */

/* Quintic 6-point weights at offsets -2..+3 and their derivatives, written
 * exactly as kick_treepm.cu:527-549. */
ORC_API void orc_interp6_weights(float x, float* w, float* dw) {
    float x1 = x;
    float x2 = x * x1;
    float x3 = x1 * x2;
    float x4 = x2 * x2;
    float x5 = x3 * x2;
    w[0] = (1.f / 12.f) * x1 - (1.f / 24.f) * x2 - (3.f / 8.f) * x3 + (13.f / 24.f) * x4 - (5.f / 24.f) * x5;
    w[1] = -(2.f / 3.f) * x1 + (2.f / 3.f) * x2 + (13.f / 8.f) * x3 - (8.f / 3.f) * x4 + (25.f / 24.f) * x5;
    w[2] = 1.0f - (5.f / 4.f) * x2 - (35.f / 12.f) * x3 + (21.f / 4.f) * x4 - (25.f / 12.f) * x5;
    w[3] = (2.f / 3.f) * x1 + (2.f / 3.f) * x2 + (11.f / 4.f) * x3 - (31.f / 6.f) * x4 + (25.f / 12.f) * x5;
    w[4] = -(1.f / 12.f) * x1 - (1.f / 24.f) * x2 - (11.f / 8.f) * x3 + (61.f / 24.f) * x4 - (25.f / 24.f) * x5;
    w[5] = (7.f / 24.f) * x3 - (0.5f) * x4 + (5.f / 24.f) * x5;
    x5 = 5.0f * x4;
    x4 = 4.0f * x3;
    x3 = 3.0f * x2;
    x2 = 2.0f * x1;
    x1 = 1.0f;
    dw[0] = (1.f / 12.f) * x1 - (1.f / 24.f) * x2 - (3.f / 8.f) * x3 + (13.f / 24.f) * x4 - (5.f / 24.f) * x5;
    dw[1] = -(2.f / 3.f) * x1 + (2.f / 3.f) * x2 + (13.f / 8.f) * x3 - (8.f / 3.f) * x4 + (25.f / 24.f) * x5;
    dw[2] = -(5.f / 4.f) * x2 - (35.f / 12.f) * x3 + (21.f / 4.f) * x4 - (25.f / 12.f) * x5;
    dw[3] = (2.f / 3.f) * x1 + (2.f / 3.f) * x2 + (11.f / 4.f) * x3 - (31.f / 6.f) * x4 + (25.f / 12.f) * x5;
    dw[4] = -(1.f / 12.f) * x1 - (1.f / 24.f) * x2 - (11.f / 8.f) * x3 + (61.f / 24.f) * x4 - (25.f / 24.f) * x5;
    dw[5] = (7.f / 24.f) * x3 - (0.5f) * x4 + (5.f / 24.f) * x5;
}

/* Mesh interpolation of one sink: g -= M * sum dw w w phi, pot += sum w w w
 * phi (kick_treepm.cu:550-572), on the periodic mesh.  The cell index and
 * fraction come from the exact product raw * M. */
static void interp6_one(int M, const float* phi, uint32_t x, uint32_t y, uint32_t z, float* g, float* pot) {
    const uint32_t r[3] = {x, y, z};
    int64_t I0[3];
    float w[3][6], dw[3][6];
    for (int d = 0; d < 3; d++) {
        const uint64_t P = (uint64_t)r[d] * (uint64_t)M;
        I0[d] = (int64_t)(P >> 32) - 2;
        const float X = (float)((uint32_t)P >> 8) * 0x1p-24f;
        orc_interp6_weights(X, w[d], dw[d]);
    }
    const float Mf = (float)M;
    float gx = 0.f, gy = 0.f, gz = 0.f, p = 0.f;
    for (int a = 0; a < 6; a++) {
        const int64_t jx = (I0[0] + a + 2 * M) % M;
        for (int b = 0; b < 6; b++) {
            const int64_t jy = (I0[1] + b + 2 * M) % M;
            for (int c = 0; c < 6; c++) {
                const int64_t jz = (I0[2] + c + 2 * M) % M;
                const float p0 = phi[(jx * M + jy) * M + jz];
                const float p1 = p0 * Mf;
                gx -= dw[0][a] * w[1][b] * w[2][c] * p1;
                gy -= w[0][a] * dw[1][b] * w[2][c] * p1;
                gz -= w[0][a] * w[1][b] * dw[2][c] * p1;
                p += w[0][a] * w[1][b] * w[2][c] * p0;
            }
        }
    }
    g[0] = gx;
    g[1] = gy;
    g[2] = gz;
    *pot = p;
}

ORC_API void orc_interp6(int M, const float* phi, int64_t n, const uint32_t* x, const uint32_t* y,
                         const uint32_t* z, float* gx, float* gy, float* gz, float* pot) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        float g[3], p;
        interp6_one(M, phi, x[i], y[i], z[i], g, &p);
        gx[i] = g[0];
        gy[i] = g[1];
        gz[i] = g[2];
        pot[i] = p;
    }
}

/* ------------------------------------------------------------------------ */
/* a9: short-range pair interaction                                          */
/* ------------------------------------------------------------------------ */

/* erfcexp (util.hpp:36-59): Abramowitz & Stegun 7.1.26. */
static inline float erfcexp(float x, float* e) {
    const float p = 0.3275911f;
    const float a1 = 0.254829592f;
    const float a2 = -0.284496736f;
    const float a3 = 1.421413741f;
    const float a4 = -1.453152027f;
    const float a5 = 1.061405429f;
    const float t1 = 1.f / fmaf(p, x, 1.f);
    const float t2 = t1 * t1;
    const float t3 = t2 * t1;
    const float t4 = t2 * t2;
    const float t5 = t2 * t3;
    *e = expf(-x * x);
    return fmaf(a1, t1, fmaf(a2, t2, fmaf(a3, t3, fmaf(a4, t4, a5 * t5)))) * *e;
}

typedef struct {
    float rcut2, inv2rs, twooversqrtpi, h2, hinv, h2inv, h3inv;
    int do_phi, pme;
} pp_consts;

static pp_consts make_pp_consts(const orc_params* p) {
    pp_consts c;
    const float rcut = (float)p->rcut;
    const float rs = (float)p->rs;
    const float h = (float)p->hsoft;
    c.rcut2 = rcut * rcut;
    c.inv2rs = 1.0f / rs / 2.0f;
    c.twooversqrtpi = 2.0f / sqrtf((float)M_PI);
    c.h2 = h * h;
    c.hinv = 1.f / h;
    c.h2inv = c.hinv * c.hinv;
    c.h3inv = c.hinv * c.h2inv;
    c.do_phi = p->do_phi;
    c.pme = p->pme;
    return c;
}

/* compute_pp_interaction (kick_treepm.cu:242-283), FP32, as written. */
static inline void pp_fp32(const pp_consts* c, float dx, float dy, float dz, float* g, float* phi) {
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    float rinv, rinv3;
    if (r2 < c->rcut2) {
        float exp0;
        rinv = r2 > 0.0f ? 1.0f / sqrtf(r2) : 0.0f;
        const float r = r2 * rinv;
        const float r0 = r * c->inv2rs;
        const float erfc0 = erfcexp(r0, &exp0);
        if (r2 > c->h2) {
            rinv3 = (erfc0 + c->twooversqrtpi * r0 * exp0) * rinv * rinv * rinv;
            rinv *= erfc0;
        } else {
            const float q2 = r2 * c->h2inv;
            float d1 = +15.0f / 8.0f;
            d1 = fmaf(d1, q2, -21.0f / 4.0f);
            d1 = fmaf(d1, q2, +35.0f / 8.0f);
            d1 *= c->h3inv;
            rinv3 = ((erfc0 - 1.0f) + c->twooversqrtpi * r0 * exp0) * rinv * rinv * rinv + d1;
            if (c->do_phi) {
                float d0 = -5.0f / 16.0f;
                d0 = fmaf(d0, q2, 21.0f / 16.0f);
                d0 = fmaf(d0, q2, -35.0f / 16.0f);
                d0 = fmaf(d0, q2, 35.0f / 16.0f);
                d0 *= c->hinv;
                rinv = (erfc0 - 1.0f) * rinv + d0;
            }
        }
        g[0] -= dx * rinv3;
        g[1] -= dy * rinv3;
        g[2] -= dz * rinv3;
        *phi -= rinv;
    }
}

/* kick_pme_kernel's pair law (kick_pme.cu:280-310), FP32, as written: every
 * pair of the 27 cells (no cutoff test); outside h the erfc-split force,
 * inside h the plain Plummer polynomials without the (erfc - 1) terms. */
static inline void pp_fp32_pme(const pp_consts* c, float dx, float dy, float dz, float* g, float* phi) {
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    float rinv, rinv3;
    if (r2 > c->h2) {
        const float r = sqrtf(r2);
        rinv = 1.0f / sqrtf(r2);
        const float r0 = r * c->inv2rs;
        float exp0;
        const float erfc0 = erfcexp(r0, &exp0);
        rinv3 = (erfc0 + c->twooversqrtpi * r0 * exp0) * rinv * rinv * rinv;
        rinv *= erfc0;
    } else {
        const float q = sqrtf(r2) * c->hinv;
        const float q2 = q * q;
        rinv3 = +15.0f / 8.0f;
        rinv3 = fmaf(rinv3, q2, -21.0f / 4.0f);
        rinv3 = fmaf(rinv3, q2, +35.0f / 8.0f);
        rinv3 *= c->h3inv;
        rinv = -5.0f / 16.0f;
        rinv = fmaf(rinv, q2, 21.0f / 16.0f);
        rinv = fmaf(rinv, q2, -35.0f / 16.0f);
        rinv = fmaf(rinv, q2, 35.0f / 16.0f);
        rinv *= c->hinv;
    }
    g[0] -= dx * rinv3;
    g[1] -= dy * rinv3;
    g[2] -= dz * rinv3;
    *phi -= rinv;
}

/* The same interaction in FP64 with libm erfc (the truth the FP32 formula
 * approximates).  Inside the softening the (erfc - 1) terms are evaluated as
 * the series -x^3 G(x^2) and -x E(x^2) to avoid cancellation. */
static inline void pp_fp64(const orc_params* p, double dx, double dy, double dz, double* g, double* phi) {
    const double rcut2 = p->rcut * p->rcut;
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < rcut2)) return;
    const double rs = p->rs, h = p->hsoft;
    const double tsp = 2.0 / sqrt(M_PI);
    const double r = sqrt(r2);
    const double x = r / (2.0 * rs);
    double f, ph;
    if (r2 > h * h) {
        const double e = exp(-x * x);
        const double ec = erfc(x);
        f = (ec + tsp * x * e) / (r2 * r);
        ph = ec / r;
    } else {
        const double u = x * x;
        /* G(u) = (4/sqrt(pi)) sum (-1)^k u^k / ((2k+3) k!),
         * E(u) = (2/sqrt(pi)) sum (-1)^k u^k / ((2k+1) k!). */
        double Gs = 0.0, Es = 0.0, term = 1.0;
        for (int k = 0; k < 40; k++) {
            Gs += term / (2 * k + 3);
            Es += term / (2 * k + 1);
            term *= -u / (k + 1);
        }
        Gs *= 2.0 * tsp;
        Es *= tsp;
        const double q2 = r2 / (h * h);
        const double d1 = ((15.0 / 8.0 * q2 - 21.0 / 4.0) * q2 + 35.0 / 8.0) / (h * h * h);
        const double d0 = (((-5.0 / 16.0 * q2 + 21.0 / 16.0) * q2 - 35.0 / 16.0) * q2 + 35.0 / 16.0) / h;
        f = d1 - Gs / (8.0 * rs * rs * rs);
        ph = d0 - Es / (2.0 * rs);
    }
    g[0] -= dx * f;
    g[1] -= dy * f;
    g[2] -= dz * f;
    *phi -= ph;
}

/* One pair through either formula (tests). out = {gx, gy, gz, phi}. */
ORC_API void orc_pp_pair(const orc_params* p, double dx, double dy, double dz, int fp64, double* out) {
    if (fp64) {
        double g[3] = {0, 0, 0}, ph = 0;
        pp_fp64(p, dx, dy, dz, g, &ph);
        out[0] = g[0];
        out[1] = g[1];
        out[2] = g[2];
        out[3] = ph;
    } else {
        pp_consts c = make_pp_consts(p);
        float g[3] = {0, 0, 0}, ph = 0;
        if (c.pme)
            pp_fp32_pme(&c, (float)dx, (float)dy, (float)dz, g, &ph);
        else
            pp_fp32(&c, (float)dx, (float)dy, (float)dz, g, &ph);
        out[0] = g[0];
        out[1] = g[1];
        out[2] = g[2];
        out[3] = ph;
    }
}

/* Short-range sum over the 27 neighbour chain cells (the candidate set of
 * kick_pme.cu:265-318, with the cutoff test of kick_treepm.cu:254) for the
 * sinks listed in sinks[] (indices into the cell-sorted arrays).  Returns
 * unit-mass sums (no GM, no phi0).  Also counts candidate pairs. */
ORC_API double orc_pp_sum(const orc_params* p, int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z,
                          const int64_t* cell_begin, int64_t nsinks, const int64_t* sinks, int fp64, float* gx,
                          float* gy, float* gz, float* pot) {
    const int C = p->chain_dim;
    uint64_t* b = malloc(sizeof(uint64_t) * (C + 1));
    orc_chain_bounds(C, b);
    pp_consts c = make_pp_consts(p);
    double cand = 0.0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : cand)
    for (int64_t s = 0; s < nsinks; s++) {
        const int64_t i = sinks[s];
        int sub;
        const int cx = chain_coord(x[i], C, b, &sub);
        const int cy = chain_coord(y[i], C, b, &sub);
        const int cz = chain_coord(z[i], C, b, &sub);
        float g[3] = {0, 0, 0}, ph = 0;
        double G[3] = {0, 0, 0}, PH = 0;
        for (int oz = -1; oz <= 1; oz++)
            for (int oy = -1; oy <= 1; oy++)
                for (int ox = -1; ox <= 1; ox++) {
                    const int jx = (cx + ox + C) % C, jy = (cy + oy + C) % C, jz = (cz + oz + C) % C;
                    const int64_t cid = ((int64_t)jz * C + jy) * C + jx;
                    const int64_t sb = cell_begin[cid], se = cell_begin[cid + 1];
                    cand += (double)(se - sb);
                    for (int64_t j = sb; j < se; j++) {
                        if (fp64) {
                            const double dx = (double)(int32_t)(x[i] - x[j]) * 0x1p-32;
                            const double dy = (double)(int32_t)(y[i] - y[j]) * 0x1p-32;
                            const double dz = (double)(int32_t)(z[i] - z[j]) * 0x1p-32;
                            pp_fp64(p, dx, dy, dz, G, &PH);
                        } else {
                            if (c.pme)
                                pp_fp32_pme(&c, orc_distance(x[i], x[j]), orc_distance(y[i], y[j]),
                                            orc_distance(z[i], z[j]), g, &ph);
                            else
                                pp_fp32(&c, orc_distance(x[i], x[j]), orc_distance(y[i], y[j]),
                                        orc_distance(z[i], z[j]), g, &ph);
                        }
                    }
                }
        if (fp64) {
            gx[s] = (float)G[0];
            gy[s] = (float)G[1];
            gz[s] = (float)G[2];
            pot[s] = (float)PH;
        } else {
            gx[s] = g[0];
            gy[s] = g[1];
            gz[s] = g[2];
            pot[s] = ph;
        }
    }
    free(b);
    return cand;
}

/* Exact candidate-pair count W_sr = sum_c n_c * sum_{c' in N27(c)} n_c'. */
ORC_API double orc_candidate_pairs(int C, const int64_t* cell_begin) {
    double w = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : w)
    for (int64_t cz = 0; cz < C; cz++)
        for (int64_t cy = 0; cy < C; cy++)
            for (int64_t cx = 0; cx < C; cx++) {
                const int64_t cid = (cz * C + cy) * C + cx;
                const double nc = (double)(cell_begin[cid + 1] - cell_begin[cid]);
                double s = 0.0;
                for (int oz = -1; oz <= 1; oz++)
                    for (int oy = -1; oy <= 1; oy++)
                        for (int ox = -1; ox <= 1; ox++) {
                            const int64_t j = (((cz + oz + C) % C) * C + (cy + oy + C) % C) * C + (cx + ox + C) % C;
                            s += (double)(cell_begin[j + 1] - cell_begin[j]);
                        }
                w += nc * s;
            }
    return w;
}

/* ------------------------------------------------------------------------ */
/* Per-cell trees, tree walk and multipole (PC) interactions of kick_treepm  */
/* (tree.cpp:7-233, kernels.hpp:13-152, kick_treepm.cu:226-240, 586-713).   */
/* ------------------------------------------------------------------------ */

#define ORC_TREE_LEAF 32    /* min(SOURCE_BUCKET_SIZE, SINK_BUCKET_SIZE), tigerpm.hpp:32-33 */
#define ORC_TREE_BUCKET 128 /* SINK_BUCKET_SIZE, tigerpm.hpp:32 */
#define ORC_TREE_DEPTH 50   /* MAX_DEPTH, tree.cpp:5 */
#define ORC_TREE_THETA 0.4f /* params.theta, kick_treepm.cu:829 */

/* tree_node (tree.hpp:17-27) with the children as indices inside the cell's
 * own tree (preorder: a node precedes its subtrees, child 0 first). */
typedef struct {
    uint32_t x[3];      /* expansion centre, fixed32 */
    float radius;       /* bounds the node's particles, plus hsoft */
    int32_t child[2];   /* -1 for a leaf */
    int32_t begin, end; /* particle range, tree order */
    float m[4];         /* particle count and -(sum of offsets from the centre) */
} orc_tree_node;

/* sink_bucket (tree.hpp:29-36). */
typedef struct {
    int32_t begin, end;
    uint32_t x[3];
    float radius;
} orc_tree_bucket;

/* fixed32 from double (fixed.hpp:63-75): 2^32 x truncated; a centre that
 * falls just outside [0, 1) wraps, which distance() treats periodically. */
static inline uint32_t tree_fixed(double x) { return (uint32_t)(int64_t)(4294967296.0 * x); }

/* sqr(a, b, c) (util.hpp:32-34) is fmaf(a, a, fmaf(b, b, c * c)); fmaf takes
 * float, so the double instantiation that tree.cpp uses rounds to FP32. */
static inline double sqr3_d(double a, double b, double c) {
    return (double)fmaf((float)a, (float)a, fmaf((float)b, (float)b, (float)(c * c)));
}
static inline float sqr3_f(float a, float b, float c) { return fmaf(a, a, fmaf(b, b, c * c)); }
static inline double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static inline double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */

typedef struct {
    orc_tree_node* nodes;
    int64_t nn, ncap;
    orc_tree_bucket* buckets;
    int64_t nb, bcap;
    uint32_t* pos[3]; /* tree-order positions (whole arrays, permuted per cell) */
    int32_t* idx;     /* tree order -> input order */
    uint32_t* tpos[3];
    int32_t* tidx; /* partition scratch, one cell long */
    double hsoft;
    int err;
} tree_cell;

/* particles_sort (particles.cpp:414-445), called by tree.cpp's sort()
 * (tree.cpp:101-107) on the longest side's midpoint: lo walks up, and a
 * particle with x >= xmid that lo meets is swapped with the next particle
 * below the top with x < xmid.  Returns the first position of the right
 * side.  The swap sequence fixes the order of the particles in a leaf. */
static int tree_partition(tree_cell* t, int begin, int end, uint32_t thr, int dim) {
    int lo = begin, hi = end;
    const uint32_t* xd = t->pos[dim];
    while (lo < hi) {
        if (xd[lo] >= thr) {
            while (lo != hi) {
                hi--;
                if (xd[hi] < thr) {
                    for (int d = 0; d < 3; d++) {
                        const uint32_t v = t->pos[d][hi];
                        t->pos[d][hi] = t->pos[d][lo];
                        t->pos[d][lo] = v;
                    }
                    const int32_t k = t->idx[hi];
                    t->idx[hi] = t->idx[lo];
                    t->idx[lo] = k;
                    break;
                }
            }
        }
        lo++;
    }
    return hi;
}

/* Centre of the bounding box of [begin, end) (the box centre when empty)
 * and the largest squared distance of a particle from it, in double
 * (tree.cpp:25-59, 88-99 and 173-206). */
static double tree_bbox(const tree_cell* t, const double bb[3], const double be[3], int begin, int end, double x[3]) {
    double xmin[3] = {1.0, 1.0, 1.0}, xmax[3] = {0.0, 0.0, 0.0};
    for (int i = begin; i < end; i++)
        for (int d = 0; d < 3; d++) {
            const double v = orc_fixed_to_double(t->pos[d][i]);
            xmax[d] = dmax(xmax[d], v);
            xmin[d] = dmin(xmin[d], v);
        }
    for (int d = 0; d < 3; d++) x[d] = (end > begin) ? (xmax[d] + xmin[d]) * 0.5 : (bb[d] + be[d]) * 0.5;
    double r2max = 0.0;
    for (int i = begin; i < end; i++) {
        double r2 = 0.0;
        for (int d = 0; d < 3; d++) {
            const double e = x[d] - orc_fixed_to_double(t->pos[d][i]);
            r2 += e * e;
        }
        r2max = dmax(r2max, r2);
    }
    return r2max;
}

/* sort() of tree.cpp:7-214: builds the node of [begin, end) in the box
 * [bb, be), recursing into the two halves of its longest side. */
static int tree_sort(tree_cell* t, const double bb[3], const double be[3], int begin, int end, int depth, int sunk) {
    if (depth > ORC_TREE_DEPTH) {
        t->err = 1; /* tree.cpp:16-19 aborts */
        return -1;
    }
    if (t->nn == t->ncap) {
        t->ncap = t->ncap ? 2 * t->ncap : 64;
        t->nodes = realloc(t->nodes, sizeof(orc_tree_node) * (size_t)t->ncap);
    }
    const int index = (int)t->nn++;
    orc_tree_node node;
    node.begin = begin;
    node.end = end;
    if (end - begin <= ORC_TREE_LEAF) {
        double x[3];
        const double r2max = tree_bbox(t, bb, be, begin, end, x);
        float m[4] = {0.f, 0.f, 0.f, 0.f};
        for (int i = begin; i < end; i++) {
            /* P2M_kernel (kernels.hpp:121-131) of the FP32 offset */
            float dx[3];
            for (int d = 0; d < 3; d++) dx[d] = (float)(orc_fixed_to_double(t->pos[d][i]) - x[d]);
            m[0] += 1.0f;
            m[1] += -dx[0];
            m[2] += -dx[1];
            m[3] += -dx[2];
        }
        for (int j = 0; j < 4; j++) node.m[j] = m[j];
        for (int d = 0; d < 3; d++) node.x[d] = tree_fixed(x[d]);
        node.radius = (float)(sqrt(r2max) + t->hsoft);
        node.child[0] = node.child[1] = -1;
    } else {
        /* range::split (range.hpp:118-133): the first strictly longest side */
        int dim = 0;
        double span = 0.0;
        for (int d = 0; d < 3; d++)
            if (be[d] - bb[d] > span) {
                span = be[d] - bb[d];
                dim = d;
            }
        const double mid = (be[dim] + bb[dim]) / 2.0;
        const int m = tree_partition(t, begin, end, tree_fixed(mid), dim);
        const int this_sunk = end - begin <= ORC_TREE_BUCKET;
        double lb[3], le[3], rb[3], re[3];
        for (int d = 0; d < 3; d++) {
            lb[d] = rb[d] = bb[d];
            le[d] = re[d] = be[d];
        }
        le[dim] = rb[dim] = mid;
        const int i0 = tree_sort(t, lb, le, begin, m, depth + 1, this_sunk);
        const int i1 = tree_sort(t, rb, re, m, end, depth + 1, this_sunk);
        if (t->err) return -1;
        const orc_tree_node* c0 = &t->nodes[i0];
        const orc_tree_node* c1 = &t->nodes[i1];
        const float mass0 = c0->m[0], mass1 = c1->m[0];
        double n[3], x[3];
        for (int d = 0; d < 3; d++) n[d] = orc_fixed_to_double(c1->x[d]) - orc_fixed_to_double(c0->x[d]);
        const double norminv = 1.0 / sqrt(sqr3_d(n[0], n[1], n[2]));
        for (int d = 0; d < 3; d++) n[d] *= norminv;
        for (int d = 0; d < 3; d++) {
            const double x0 = orc_fixed_to_double(c0->x[d]), x1 = orc_fixed_to_double(c1->x[d]);
            if (mass1 == 0.0f)
                x[d] = x0;
            else if (mass0 == 0.0f)
                x[d] = x1;
            else
                x[d] = (x0 + x1 + n[d] * (double)(c1->radius - c0->radius)) * 0.5;
        }
        /* M2M_kernel (kernels.hpp:134-152) of both children, summed */
        float mm[2][4];
        const orc_tree_node* cs[2] = {c0, c1};
        for (int k = 0; k < 2; k++) {
            float X[3];
            for (int d = 0; d < 3; d++) X[d] = -(float)(orc_fixed_to_double(cs[k]->x[d]) - x[d]);
            const float* Ma = cs[k]->m;
            mm[k][0] = Ma[0];
            mm[k][1] = fmaf(X[0], Ma[0], Ma[1]);
            mm[k][3] = fmaf(X[2], Ma[0], Ma[3]);
            mm[k][2] = fmaf(X[1], Ma[0], Ma[2]);
        }
        for (int j = 0; j < 4; j++) node.m[j] = mm[0][j] + mm[1][j];
        for (int d = 0; d < 3; d++) node.x[d] = tree_fixed(x[d]);
        double r20 = 0.0, r21 = 0.0;
        for (int d = 0; d < 3; d++) {
            if (mass0 != 0.0f) {
                const double e = orc_fixed_to_double(c0->x[d]) - x[d];
                r20 += e * e;
            }
            if (mass1 != 0.0f) {
                const double e = orc_fixed_to_double(c1->x[d]) - x[d];
                r21 += e * e;
            }
        }
        double r2_max = 0.0;
        for (int corner = 0; corner < 8; corner++) {
            const double cx = (corner & 4) ? be[0] : bb[0];
            const double cy = (corner & 2) ? be[1] : bb[1];
            const double cz = (corner & 1) ? be[2] : bb[2];
            r2_max = dmax(r2_max, sqr3_d(cx - x[0], cy - x[1], cz - x[2]));
        }
        node.radius = (float)dmin(dmax(sqrt(r20) + (double)c0->radius, sqrt(r21) + (double)c1->radius),
                                  sqrt(r2_max) + t->hsoft);
        node.child[0] = i0;
        node.child[1] = i1;
    }
    if (end - begin <= ORC_TREE_BUCKET && !sunk) {
        if (t->nb == t->bcap) {
            t->bcap = t->bcap ? 2 * t->bcap : 16;
            t->buckets = realloc(t->buckets, sizeof(orc_tree_bucket) * (size_t)t->bcap);
        }
        orc_tree_bucket* bk = &t->buckets[t->nb++];
        double x[3];
        const double r2max = tree_bbox(t, bb, be, begin, end, x);
        for (int d = 0; d < 3; d++) bk->x[d] = tree_fixed(x[d]);
        bk->radius = (float)(sqrt(r2max) + t->hsoft);
        bk->begin = begin;
        bk->end = end;
    }
    t->nodes[index] = node;
    return index;
}

typedef struct {
    int64_t ncell;
    int64_t* node_begin;   /* ncell + 1 */
    int64_t* bucket_begin; /* ncell + 1 */
    orc_tree_node* nodes;
    orc_tree_bucket* buckets;
    uint32_t* pos[3];
    int32_t* idx;
    int err;
} orc_forest;

static void forest_free(orc_forest* f) {
    free(f->node_begin);
    free(f->bucket_begin);
    free(f->nodes);
    free(f->buckets);
    for (int d = 0; d < 3; d++) free(f->pos[d]);
    free(f->idx);
}

/* tree_create (tree.cpp:216-245) for every chain cell, as the loop of
 * kick_treepm_begin (kick_treepm.cpp:26-44) does: the root box of cell i is
 * [i / C, (i + 1) / C) per axis in double. */
static void forest_build(const orc_params* p, int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z,
                         const int64_t* cell_begin, orc_forest* f) {
    const int C = p->chain_dim;
    const int64_t ncell = (int64_t)C * C * C;
    const double ninv = 1.0 / (double)C;
    f->ncell = ncell;
    f->err = 0;
    const uint32_t* src[3] = {x, y, z};
    for (int d = 0; d < 3; d++) {
        f->pos[d] = malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
        memcpy(f->pos[d], src[d], sizeof(uint32_t) * (size_t)n);
    }
    f->idx = malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) f->idx[i] = (int32_t)i;
    orc_tree_node** cn = calloc((size_t)ncell, sizeof(orc_tree_node*));
    orc_tree_bucket** cbk = calloc((size_t)ncell, sizeof(orc_tree_bucket*));
    f->node_begin = calloc((size_t)ncell + 1, sizeof(int64_t));
    f->bucket_begin = calloc((size_t)ncell + 1, sizeof(int64_t));
    int64_t maxcell = 1;
    for (int64_t c = 0; c < ncell; c++)
        if (cell_begin[c + 1] - cell_begin[c] > maxcell) maxcell = cell_begin[c + 1] - cell_begin[c];
    int err = 0;
#pragma omp parallel reduction(| : err)
    {
        tree_cell t;
        memset(&t, 0, sizeof(t));
        for (int d = 0; d < 3; d++) {
            t.pos[d] = f->pos[d];
            t.tpos[d] = malloc(sizeof(uint32_t) * (size_t)maxcell);
        }
        t.idx = f->idx;
        t.tidx = malloc(sizeof(int32_t) * (size_t)maxcell);
        t.hsoft = p->hsoft;
#pragma omp for schedule(dynamic, 16)
        for (int64_t c = 0; c < ncell; c++) {
            const int ix = (int)(c % C), iy = (int)((c / C) % C), iz = (int)(c / ((int64_t)C * C));
            const int ic[3] = {ix, iy, iz};
            double bb[3], be[3];
            for (int d = 0; d < 3; d++) {
                bb[d] = ic[d] * ninv;
                be[d] = (ic[d] + 1) * ninv;
            }
            t.nn = t.nb = 0;
            t.err = 0;
            tree_sort(&t, bb, be, (int)cell_begin[c], (int)cell_begin[c + 1], 0, 0);
            err |= t.err;
            cn[c] = malloc(sizeof(orc_tree_node) * (size_t)(t.nn > 0 ? t.nn : 1));
            memcpy(cn[c], t.nodes, sizeof(orc_tree_node) * (size_t)t.nn);
            cbk[c] = malloc(sizeof(orc_tree_bucket) * (size_t)(t.nb > 0 ? t.nb : 1));
            memcpy(cbk[c], t.buckets, sizeof(orc_tree_bucket) * (size_t)t.nb);
            f->node_begin[c + 1] = t.nn;
            f->bucket_begin[c + 1] = t.nb;
        }
        for (int d = 0; d < 3; d++) free(t.tpos[d]);
        free(t.tidx);
        free(t.nodes);
        free(t.buckets);
    }
    f->err = err;
    for (int64_t c = 0; c < ncell; c++) {
        f->node_begin[c + 1] += f->node_begin[c];
        f->bucket_begin[c + 1] += f->bucket_begin[c];
    }
    f->nodes = malloc(sizeof(orc_tree_node) * (size_t)(f->node_begin[ncell] + 1));
    f->buckets = malloc(sizeof(orc_tree_bucket) * (size_t)(f->bucket_begin[ncell] + 1));
    for (int64_t c = 0; c < ncell; c++) {
        memcpy(f->nodes + f->node_begin[c], cn[c], sizeof(orc_tree_node) * (size_t)(f->node_begin[c + 1] - f->node_begin[c]));
        memcpy(f->buckets + f->bucket_begin[c], cbk[c],
               sizeof(orc_tree_bucket) * (size_t)(f->bucket_begin[c + 1] - f->bucket_begin[c]));
        free(cn[c]);
        free(cbk[c]);
    }
    free(cn);
    free(cbk);
}

/* The trees of every chain cell.  order[i] = input index of tree position i;
 * tx/ty/tz = positions in tree order; node_begin/bucket_begin get C^3 + 1
 * offsets.  Node and bucket arrays are filled when their capacity suffices.
 * Returns 0, 1 (depth limit exceeded) or 2 (capacity too small). */
ORC_API int orc_tree_build(const orc_params* p, int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z,
                           const int64_t* cell_begin, int32_t* order, uint32_t* tx, uint32_t* ty, uint32_t* tz,
                           int64_t* node_begin, int64_t* bucket_begin, orc_tree_node* nodes, int64_t node_cap,
                           orc_tree_bucket* buckets, int64_t bucket_cap) {
    orc_forest f;
    forest_build(p, n, x, y, z, cell_begin, &f);
    memcpy(order, f.idx, sizeof(int32_t) * (size_t)n);
    memcpy(tx, f.pos[0], sizeof(uint32_t) * (size_t)n);
    memcpy(ty, f.pos[1], sizeof(uint32_t) * (size_t)n);
    memcpy(tz, f.pos[2], sizeof(uint32_t) * (size_t)n);
    memcpy(node_begin, f.node_begin, sizeof(int64_t) * (size_t)(f.ncell + 1));
    memcpy(bucket_begin, f.bucket_begin, sizeof(int64_t) * (size_t)(f.ncell + 1));
    int rc = f.err ? 1 : 0;
    if (!rc) {
        if (f.node_begin[f.ncell] <= node_cap && f.bucket_begin[f.ncell] <= bucket_cap) {
            memcpy(nodes, f.nodes, sizeof(orc_tree_node) * (size_t)f.node_begin[f.ncell]);
            memcpy(buckets, f.buckets, sizeof(orc_tree_bucket) * (size_t)f.bucket_begin[f.ncell]);
        } else {
            rc = 2;
        }
    }
    forest_free(&f);
    return rc;
}

/* compute_pc_interaction (kick_treepm.cu:226-240): greens_function
 * (kernels.hpp:13-66) and the 4-component M2L_kernel (kernels.hpp:69-88),
 * as written.  d = sink - expansion centre. */
static inline void pc_fp32(const pp_consts* c, float dx, float dy, float dz, const float* M, float* g, float* phi) {
    float D[10] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const float r2 = sqr3_f(dx, dy, dz);
    const float rinv = 1.0f / sqrtf(r2);
    const float r = r2 * rinv;
    const float r0 = r * c->inv2rs;
    float exp0;
    const float erfc0 = erfcexp(r0, &exp0);
    const float c0 = -2.f * r * c->inv2rs;
    const float d0 = -erfc0 * rinv;
    float e0 = c->twooversqrtpi * exp0 * rinv * c->inv2rs;
    const float d1 = fmaf(-1.f * d0, rinv, e0);
    e0 *= c0;
    const float d2 = fmaf(-3.f * d1, rinv, e0);
    const float D00 = d0 * 1.f, D10 = d1 * 1.f, D11 = d1 * rinv, D20 = d2 * 1.f;
    const float x100 = dx * rinv, x010 = dy * rinv, x001 = dz * rinv;
    const float x002 = x001 * x001, x011 = x010 * x001, x020 = x010 * x010;
    const float x101 = x100 * x001, x110 = x100 * x010, x200 = x100 * x100;
    float x21 = x002;
    x21 += x020;
    x21 += x200;
    x21 *= D11;
    D[0] = fmaf(1.f, D00, D[0]);
    D[7] += x21;
    D[3] = fmaf(x001, D10, D[3]);
    D[1] = fmaf(x100, D10, D[1]);
    D[9] = fmaf(x002, D20, D[9]);
    D[6] = fmaf(x101, D20, D[6]);
    D[9] += x21;
    D[5] = fmaf(x110, D20, D[5]);
    D[2] = fmaf(x010, D10, D[2]);
    D[4] = fmaf(x200, D20, D[4]);
    D[8] = fmaf(x011, D20, D[8]);
    D[4] += x21;
    D[7] = fmaf(x020, D20, D[7]);
    float L[4] = {0.f, 0.f, 0.f, 0.f};
    if (c->do_phi) {
        L[0] = fmaf(M[0], D[0], L[0]);
        L[0] = fmaf(M[3], D[3], L[0]);
        L[0] = fmaf(M[2], D[2], L[0]);
        L[0] = fmaf(M[1], D[1], L[0]);
    }
    L[1] = fmaf(M[0], D[1], L[1]);
    L[2] = fmaf(M[2], D[7], L[2]);
    L[1] = fmaf(M[3], D[6], L[1]);
    L[2] = fmaf(M[1], D[5], L[2]);
    L[1] = fmaf(M[2], D[5], L[1]);
    L[3] = fmaf(M[0], D[3], L[3]);
    L[1] = fmaf(M[1], D[4], L[1]);
    L[3] = fmaf(M[3], D[9], L[3]);
    L[2] = fmaf(M[0], D[2], L[2]);
    L[3] = fmaf(M[2], D[8], L[3]);
    L[2] = fmaf(M[3], D[8], L[2]);
    L[3] = fmaf(M[1], D[6], L[3]);
    g[0] -= L[1];
    g[1] -= L[2];
    g[2] -= L[3];
    *phi += L[0];
}

/* One PC interaction (tests).  M = {mass, -dipole}; out = {gx, gy, gz, phi}. */
ORC_API void orc_pc_pair(const orc_params* p, float dx, float dy, float dz, const float* M, float* out) {
    pp_consts c = make_pp_consts(p);
    float g[3] = {0.f, 0.f, 0.f}, ph = 0.f;
    pc_fp32(&c, dx, dy, dz, M, g, &ph);
    out[0] = g[0];
    out[1] = g[1];
    out[2] = g[2];
    out[3] = ph;
}

typedef struct {
    int32_t *check, *next, *tmp, *multi, *part;
    int64_t cap;
} walk_lists;

static void lists_reserve(walk_lists* w, int64_t need) {
    if (need <= w->cap) return;
    w->cap = need;
    int32_t** all[5] = {&w->check, &w->next, &w->tmp, &w->multi, &w->part};
    for (int k = 0; k < 5; k++) *all[k] = realloc(*all[k], sizeof(int32_t) * (size_t)need);
}

/* Short-range sum of kick_treepm_kernel (kick_treepm.cu:586-713) for every
 * active sink (rung >= min_rung): per sink bucket, for each of the 27
 * neighbour trees (x outermost, kick_treepm.cu:930-940), a breadth-first walk
 * that drops nodes beyond rcut of the bucket (cutoff), takes nodes far from
 * the bucket (R > (r_src + r_snk) / theta) as multipoles, opens the others,
 * and sends each near leaf to PC when every active sink is beyond
 * r_src / theta of it and to PP otherwise; then PC (compute_pc_interaction)
 * over the multipole list and PP (compute_pp_interaction) over the leaves'
 * particles.  Unit-mass sums in input order (inactive sinks get 0).
 * stats = {PC interactions, PP pairs, multipole-list nodes, PP-list leaves}.
 * pc_exact != 0 (a check, not the reference): nodes of the multipole list
 * are summed particle by particle with the pair law instead.  Returns 0, or
 * 1 when a tree exceeds the depth limit. */
ORC_API int orc_tree_sum(const orc_params* p, int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z,
                         const int64_t* cell_begin, const int8_t* rung, int min_rung, int pc_exact, float* gx,
                         float* gy, float* gz, float* pot, double* stats) {
    orc_forest f;
    forest_build(p, n, x, y, z, cell_begin, &f);
    for (int64_t i = 0; i < n; i++) gx[i] = gy[i] = gz[i] = pot[i] = 0.f;
    stats[0] = stats[1] = stats[2] = stats[3] = 0.0;
    if (f.err) {
        forest_free(&f);
        return 1;
    }
    const int C = p->chain_dim;
    const pp_consts c = make_pp_consts(p);
    const float rcut = (float)p->rcut;
    const float theta2inv = 1.0f / (ORC_TREE_THETA * ORC_TREE_THETA);
    int64_t maxnodes = 1;
    for (int64_t q = 0; q < f.ncell; q++)
        if (f.node_begin[q + 1] - f.node_begin[q] > maxnodes) maxnodes = f.node_begin[q + 1] - f.node_begin[q];
    double npc = 0, npp = 0, nmulti = 0, npart = 0;
    const uint32_t* const* P = (const uint32_t* const*)f.pos;
#pragma omp parallel reduction(+ : npc, npp, nmulti, npart)
    {
        walk_lists w;
        memset(&w, 0, sizeof(w));
        lists_reserve(&w, 2 * maxnodes + 2);
        int32_t act[ORC_TREE_BUCKET];
#pragma omp for schedule(dynamic, 4)
        for (int64_t cell = 0; cell < f.ncell; cell++) {
            const int cx = (int)(cell % C), cy = (int)((cell / C) % C), cz = (int)(cell / ((int64_t)C * C));
            for (int64_t bi = f.bucket_begin[cell]; bi < f.bucket_begin[cell + 1]; bi++) {
                const orc_tree_bucket* bk = &f.buckets[bi];
                int na = 0;
                for (int i = bk->begin; i < bk->end; i++)
                    if (rung[f.idx[i]] >= min_rung) act[na++] = i;
                if (na == 0) continue;
                float g[ORC_TREE_BUCKET][3], ph[ORC_TREE_BUCKET];
                for (int s = 0; s < na; s++) g[s][0] = g[s][1] = g[s][2] = ph[s] = 0.f;
                for (int ox = -1; ox <= 1; ox++)
                    for (int oy = -1; oy <= 1; oy++)
                        for (int oz = -1; oz <= 1; oz++) {
                            const int jx = (cx + ox + C) % C, jy = (cy + oy + C) % C, jz = (cz + oz + C) % C;
                            const int64_t tc = ((int64_t)jz * C + jy) * C + jx;
                            const orc_tree_node* tr = f.nodes + f.node_begin[tc];
                            int ncheck = 1, nmul = 0, npar = 0;
                            w.check[0] = 0;
                            while (ncheck) {
                                int nnext = 0, ntmp = 0;
                                for (int ci = 0; ci < ncheck; ci++) {
                                    const orc_tree_node* nd = &tr[w.check[ci]];
                                    const float dx = orc_distance(bk->x[0], nd->x[0]);
                                    const float dy = orc_distance(bk->x[1], nd->x[1]);
                                    const float dz = orc_distance(bk->x[2], nd->x[2]);
                                    const float R2 = sqr3_f(dx, dy, dz);
                                    const float R = sqrtf(R2);
                                    const int cutoff = R - nd->radius - bk->radius > rcut;
                                    const float rr = nd->radius + bk->radius;
                                    const int far = R2 > rr * rr * theta2inv;
                                    const int leaf = nd->end - nd->begin <= ORC_TREE_LEAF;
                                    if (cutoff) continue;
                                    if (far)
                                        w.multi[nmul++] = w.check[ci];
                                    else if (leaf)
                                        w.tmp[ntmp++] = w.check[ci];
                                    else {
                                        w.next[nnext++] = nd->child[0];
                                        w.next[nnext++] = nd->child[1];
                                    }
                                }
                                int32_t* sw = w.check;
                                w.check = w.next;
                                w.next = sw;
                                ncheck = nnext;
                                for (int k = 0; k < ntmp; k++) {
                                    const orc_tree_node* nd = &tr[w.tmp[k]];
                                    const float lim = nd->radius * nd->radius * theta2inv;
                                    int near = 0;
                                    for (int s = 0; s < na && !near; s++) {
                                        const int i = act[s];
                                        const float dx = orc_distance(P[0][i], nd->x[0]);
                                        const float dy = orc_distance(P[1][i], nd->x[1]);
                                        const float dz = orc_distance(P[2][i], nd->x[2]);
                                        near = !(sqr3_f(dx, dy, dz) > lim);
                                    }
                                    if (near)
                                        w.part[npar++] = w.tmp[k];
                                    else
                                        w.multi[nmul++] = w.tmp[k];
                                }
                            }
                            nmulti += nmul;
                            npart += npar;
                            for (int s = 0; s < na; s++) {
                                const int i = act[s];
                                for (int k = 0; k < nmul; k++) {
                                    const orc_tree_node* nd = &tr[w.multi[k]];
                                    if (pc_exact) {
                                        for (int j = nd->begin; j < nd->end; j++)
                                            pp_fp32(&c, orc_distance(P[0][i], P[0][j]), orc_distance(P[1][i], P[1][j]),
                                                    orc_distance(P[2][i], P[2][j]), g[s], &ph[s]);
                                    } else {
                                        pc_fp32(&c, orc_distance(P[0][i], nd->x[0]), orc_distance(P[1][i], nd->x[1]),
                                                orc_distance(P[2][i], nd->x[2]), nd->m, g[s], &ph[s]);
                                    }
                                }
                                for (int k = 0; k < npar; k++) {
                                    const orc_tree_node* nd = &tr[w.part[k]];
                                    for (int j = nd->begin; j < nd->end; j++)
                                        pp_fp32(&c, orc_distance(P[0][i], P[0][j]), orc_distance(P[1][i], P[1][j]),
                                                orc_distance(P[2][i], P[2][j]), g[s], &ph[s]);
                                    npp += nd->end - nd->begin;
                                }
                                npc += nmul;
                            }
                        }
                for (int s = 0; s < na; s++) {
                    const int32_t o = f.idx[act[s]];
                    gx[o] = g[s][0];
                    gy[o] = g[s][1];
                    gz[o] = g[s][2];
                    pot[o] = ph[s];
                }
            }
        }
        free(w.check);
        free(w.next);
        free(w.tmp);
        free(w.multi);
        free(w.part);
    }
    stats[0] = npc;
    stats[1] = npp;
    stats[2] = nmulti;
    stats[3] = npart;
    forest_free(&f);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Drift step of the batched serving loop (no reference counterpart: the      */
/* reference only kicks).  x <- x + v t0 in fixed32: the displacement v t0 in */
/* box units, formed in double, is rounded to the nearest raw unit (llrint,   */
/* ties to even) and added modulo 2^32, i.e. periodically.                    */
/* ------------------------------------------------------------------------ */
ORC_API void orc_drift(int64_t n, uint32_t* x, uint32_t* y, uint32_t* z, const float* vx, const float* vy,
                       const float* vz, double t0) {
    const double f = t0 * 4294967296.0;
    for (int64_t i = 0; i < n; i++) {
        x[i] += (uint32_t)(uint64_t)llrint((double)vx[i] * f);
        y[i] += (uint32_t)(uint64_t)llrint((double)vy[i] * f);
        z[i] += (uint32_t)(uint64_t)llrint((double)vz[i] * f);
    }
}

/* ------------------------------------------------------------------------ */
/* a12: kick epilogue (kick_treepm.cu:716-754; rung_dt fmmpm.cu:18-21)        */
/* ------------------------------------------------------------------------ */

/* Input g/pot are the unit-mass totals (PM + PP).  Applies phi0 and GM,
 * stores the FORCE_TEST outputs and updates velocity and rung.  The rung is
 * clamped to [0, 31] (the reference asserts it).  rung_dt[r] = 2^-r. */
ORC_API void orc_kick_epilogue(const orc_params* p, int64_t n, float* gx, float* gy, float* gz, float* pot,
                               float* vx, float* vy, float* vz, int8_t* rung, int min_rung, double scale, double t0,
                               int first_call) {
    const float GM = (float)p->GM;
    /* kick_treepm's phi0 (kick_treepm.cu:839); kick_pme starts phi at 0 */
    const float phi0 = p->pme ? 0.f : (float)(p->n_total * 4.0 * M_PI * p->rs * p->rs + (35.0 / 16.0) / p->hsoft);
    const float eta = (float)p->eta;
    const float hsoft = (float)p->hsoft;
    const float t0f = (float)t0;
    const float scalef = (float)scale;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        if (rung[i] < min_rung) continue;
        const float g0 = gx[i] * GM, g1 = gy[i] * GM, g2v = gz[i] * GM;
        const float ph = (pot[i] + phi0) * GM;
        gx[i] = g0;
        gy[i] = g1;
        gz[i] = g2v;
        pot[i] = ph;
        int r = rung[i];
        float dt = 0.5f * ldexpf(1.0f, -r) * t0f;
        if (!first_call) {
            vx[i] = fmaf(g0, dt, vx[i]);
            vy[i] = fmaf(g1, dt, vy[i]);
            vz[i] = fmaf(g2v, dt, vz[i]);
        }
        const float g2 = fmaf(g0, g0, fmaf(g1, g1, g2v * g2v));
        const float factor = eta * sqrtf(scalef * hsoft);
        dt = fminf(factor * (1.0f / sqrtf(sqrtf(g2))), t0f);
        float nr = fmaxf(ceilf(log2f(t0f) - log2f(dt)), (float)(r - 1));
        if (nr < 0.f) nr = 0.f;
        if (nr > 31.f) nr = 31.f;
        r = (int)nr;
        rung[i] = (int8_t)r;
        dt = 0.5f * ldexpf(1.0f, -r) * t0f;
        vx[i] = fmaf(g0, dt, vx[i]);
        vy[i] = fmaf(g1, dt, vy[i]);
        vz[i] = fmaf(g2v, dt, vz[i]);
    }
}

/* ------------------------------------------------------------------------ */
/* Whole kick on the CPU (kick_treepm_test, test.cpp:237-290, with the     */
/* direct 27-cell PP of kick_pme.cu in place of the tree walk).              */
/* ------------------------------------------------------------------------ */

/* Sorts the particle arrays in place into chain-mesh order, runs the PM
 * solve, PP + interpolation and the kick epilogue.  Outputs are in sorted
 * order.  cell_begin receives C^3 + 1 entries.  times[0..4] = seconds in
 * chain mesh, deposit, FFT+Green, PP+interp, epilogue.  fp64_pp: 0 the FP32
 * direct 27-cell sum, 1 the same in FP64, 2 kick_treepm's tree walk with
 * multipoles (times[3] = -1 when a tree exceeds the depth limit). */
ORC_API void orc_kick_full(const orc_params* p, int64_t n, uint32_t* x, uint32_t* y, uint32_t* z, float* vx,
                           float* vy, float* vz, int8_t* rung, float* gx, float* gy, float* gz, float* pot,
                           int64_t* cell_begin, int min_rung, double scale, double t0, int first_call,
                           int fp64_pp, double* times) {
    const int M = p->mesh_dim;
    const int C = p->chain_dim;
    const int64_t M3 = (int64_t)M * M * M;
    double t = omp_get_wtime();
    int64_t* perm = malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    orc_chainmesh_sort(n, x, y, z, C, perm, cell_begin);
    {
        uint32_t* tu = malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
        float* tf = (float*)tu;
        int8_t* tr = malloc((size_t)(n > 0 ? n : 1));
        uint32_t* arrs[3] = {x, y, z};
        for (int a = 0; a < 3; a++) {
            for (int64_t i = 0; i < n; i++) tu[i] = arrs[a][perm[i]];
            memcpy(arrs[a], tu, sizeof(uint32_t) * (size_t)n);
        }
        float* farrs[3] = {vx, vy, vz};
        for (int a = 0; a < 3; a++) {
            for (int64_t i = 0; i < n; i++) tf[i] = farrs[a][perm[i]];
            memcpy(farrs[a], tf, sizeof(float) * (size_t)n);
        }
        for (int64_t i = 0; i < n; i++) tr[i] = rung[perm[i]];
        memcpy(rung, tr, (size_t)n);
        free(tu);
        free(tr);
    }
    free(perm);
    times[0] = omp_get_wtime() - t;
    t = omp_get_wtime();
    uint64_t* grid = malloc(sizeof(uint64_t) * (size_t)M3);
    orc_deposit(n, x, y, z, M, p->order, grid);
    float* rho = malloc(sizeof(float) * (size_t)M3);
    orc_density_float(M3, grid, rho);
    float* rho1 = NULL;
    if (p->interlace) {
        orc_deposit_shifted(n, x, y, z, M, p->order, 1ull << 31, grid);
        rho1 = malloc(sizeof(float) * (size_t)M3);
        orc_density_float(M3, grid, rho1);
    }
    free(grid);
    times[1] = omp_get_wtime() - t;
    t = omp_get_wtime();
    float* table = malloc(sizeof(float) * (size_t)M * M * (M / 2 + 1));
    orc_green_table(M, p->rs, p->order, table);
    float* phi = malloc(sizeof(float) * (size_t)M3);
    if (p->interlace)
        orc_pm_phi_interlaced(M, rho, rho1, table, phi);
    else
        orc_pm_phi(M, rho, table, phi);
    free(table);
    free(rho);
    free(rho1);
    times[2] = omp_get_wtime() - t;
    t = omp_get_wtime();
    int tree_err = 0;
    if (fp64_pp == 2) {
        double st[4];
        tree_err = orc_tree_sum(p, n, x, y, z, cell_begin, rung, min_rung, 0, gx, gy, gz, pot, st);
    } else {
        int64_t* sinks = malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
        for (int64_t i = 0; i < n; i++) sinks[i] = i;
        orc_pp_sum(p, n, x, y, z, cell_begin, n, sinks, fp64_pp, gx, gy, gz, pot);
        free(sinks);
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        float g[3], pp;
        interp6_one(M, phi, x[i], y[i], z[i], g, &pp);
        gx[i] += g[0];
        gy[i] += g[1];
        gz[i] += g[2];
        pot[i] += pp;
    }
    free(phi);
    times[3] = tree_err ? -1.0 : omp_get_wtime() - t;
    t = omp_get_wtime();
    orc_kick_epilogue(p, n, gx, gy, gz, pot, vx, vy, vz, rung, min_rung, scale, t0, first_call);
    times[4] = omp_get_wtime() - t;
}

/* ------------------------------------------------------------------------ */
/* a14: FP64 Ewald sum (the accuracy reference)                              */
/* ------------------------------------------------------------------------ */

/* Periodic potential of a unit mass in the unit box with a neutralizing
 * background, and its gradient, by direct Ewald sums (the decomposition of
 * gravity_short.cu:117-164 in FP64):
 *   psi(d) = -sum_n erfc(a|d+n|)/|d+n|
 *            - (1/pi) sum_{h!=0} exp(-pi^2 h^2/a^2)/h^2 cos(2 pi h.d) + pi/a^2.
 * With self != 0 the n = 0 real-space term is replaced by its regular part
 * erf(a r)/r (the limit psi(d) + 1/|d|). */
static void ewald_psi(double dx, double dy, double dz, double a, int nimg, int hmax, int self, double* psi,
                      double* grad) {
    const double tsp = 2.0 / sqrt(M_PI);
    double ps = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    for (int ix = -nimg; ix <= nimg; ix++)
        for (int iy = -nimg; iy <= nimg; iy++)
            for (int iz = -nimg; iz <= nimg; iz++) {
                const double X = dx + ix, Y = dy + iy, Z = dz + iz;
                const double r2 = X * X + Y * Y + Z * Z;
                const double r = sqrt(r2);
                if (ix == 0 && iy == 0 && iz == 0 && self) {
                    if (r2 == 0.0) {
                        ps += tsp * a;
                    } else {
                        const double e = erf(a * r);
                        ps += e / r;
                        const double dr = (tsp * a * exp(-a * a * r2) * r - e) / r2;
                        gx += dr * X / r;
                        gy += dr * Y / r;
                        gz += dr * Z / r;
                    }
                    continue;
                }
                const double ec = erfc(a * r);
                ps -= ec / r;
                const double dr = ec / r2 + tsp * a * exp(-a * a * r2) / r;
                gx += dr * X / r;
                gy += dr * Y / r;
                gz += dr * Z / r;
            }
    for (int hx = -hmax; hx <= hmax; hx++)
        for (int hy = -hmax; hy <= hmax; hy++)
            for (int hz = -hmax; hz <= hmax; hz++) {
                const int h2 = hx * hx + hy * hy + hz * hz;
                if (h2 == 0 || h2 > hmax * hmax) continue;
                const double c = exp(-M_PI * M_PI * h2 / (a * a)) / h2 / M_PI;
                const double om = 2.0 * M_PI * (hx * dx + hy * dy + hz * dz);
                ps -= c * cos(om);
                const double s = sin(om) * 2.0 * M_PI * c;
                gx += s * hx;
                gy += s * hy;
                gz += s * hz;
            }
    ps += M_PI / (a * a);
    *psi = ps;
    grad[0] = gx;
    grad[1] = gy;
    grad[2] = gz;
}

ORC_API void orc_ewald_psi(double dx, double dy, double dz, double alpha, int nimg, int hmax, int self,
                           double* out) {
    double g[3];
    ewald_psi(dx, dy, dz, alpha, nimg, hmax, self, &out[0], g);
    out[1] = g[0];
    out[2] = g[1];
    out[3] = g[2];
}

/* Softened Plummer-polynomial potential and force factor (kick_treepm.cu:
 * 263-276, gravity_short.cu:165-182): phi = -d0/h, g = -d1/h^3 * d. */
static inline void soft_poly(double r2, double h, double* d0, double* d1) {
    const double q2 = r2 / (h * h);
    *d1 = ((15.0 / 8.0 * q2 - 21.0 / 4.0) * q2 + 35.0 / 8.0) / (h * h * h);
    *d0 = (((-5.0 / 16.0 * q2 + 21.0 / 16.0) * q2 - 35.0 / 16.0) * q2 + 35.0 / 16.0) / h;
}

/* Brute-force Ewald sum for small N (pair by pair with ewald_psi, alpha 2,
 * images |n| <= 4, |h| <= 4): for each sample particle s, phi = sum_{j != s}
 * psi_soft(x_s - x_j) + xi, g = -grad phi, unit masses.  psi_soft is psi with
 * the Newtonian -1/r replaced by the softened polynomial for r < hsoft;
 * xi = lim (psi + 1/r) is the particle's own image + background term. */
ORC_API void orc_ewald_direct(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int64_t ns,
                              const int64_t* samples, double hsoft, double* phi, double* gx, double* gy,
                              double* gz) {
    const double a = 2.0;
    double xi, g0[3];
    ewald_psi(0, 0, 0, a, 4, 4, 1, &xi, g0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t s = 0; s < ns; s++) {
        const int64_t i = samples[s];
        double P = xi, G[3] = {0, 0, 0};
        for (int64_t j = 0; j < n; j++) {
            if (j == i) continue;
            const double dx = (double)(int32_t)(x[i] - x[j]) * 0x1p-32;
            const double dy = (double)(int32_t)(y[i] - y[j]) * 0x1p-32;
            const double dz = (double)(int32_t)(z[i] - z[j]) * 0x1p-32;
            const double r2 = dx * dx + dy * dy + dz * dz;
            double ps, gr[3];
            if (r2 < hsoft * hsoft) {
                ewald_psi(dx, dy, dz, a, 4, 4, 1, &ps, gr);
                double d0, d1;
                soft_poly(r2, hsoft, &d0, &d1);
                P += ps - d0;
                G[0] -= gr[0] + d1 * dx;
                G[1] -= gr[1] + d1 * dy;
                G[2] -= gr[2] + d1 * dz;
            } else {
                ewald_psi(dx, dy, dz, a, 4, 4, 0, &ps, gr);
                P += ps;
                G[0] -= gr[0];
                G[1] -= gr[1];
                G[2] -= gr[2];
            }
        }
        phi[s] = P;
        gx[s] = G[0];
        gy[s] = G[1];
        gz[s] = G[2];
    }
}

/* Structure-factor Ewald sum (the algorithm the GPU's tpm_ewald runs): real
 * space on the minimum image only (rc = 4.8/alpha < 1/2, erfc(4.8) ~ 6e-12),
 * reciprocal space through S(h) = sum_j exp(2 pi i h.x_j) over the half
 * space 0 < |h| <= 5.2 alpha/pi.  Same outputs and conventions as
 * orc_ewald_direct. */
static void ewald_sf_core(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int64_t ns,
                          const int64_t* samples, const uint32_t* sx, const uint32_t* sy, const uint32_t* sz,
                          double hsoft, double alpha, double* phi, double* gx, double* gy, double* gz) {
    const double a = alpha;
    const double rc = 4.8 / a;
    const double rc2 = rc * rc;
    const double tsp = 2.0 / sqrt(M_PI);
    const int hmax = (int)floor(5.2 * a / M_PI);
    /* half-space h list */
    int nh = 0;
    int* hl = malloc(sizeof(int) * 3 * (size_t)(2 * hmax + 1) * (2 * hmax + 1) * (2 * hmax + 1));
    for (int hx = 0; hx <= hmax; hx++)
        for (int hy = -hmax; hy <= hmax; hy++)
            for (int hz = -hmax; hz <= hmax; hz++) {
                const int h2 = hx * hx + hy * hy + hz * hz;
                if (h2 == 0 || h2 > hmax * hmax) continue;
                if (hx == 0 && (hy < 0 || (hy == 0 && hz < 0))) continue;
                hl[3 * nh] = hx;
                hl[3 * nh + 1] = hy;
                hl[3 * nh + 2] = hz;
                nh++;
            }
    double* Sre = calloc((size_t)nh, sizeof(double));
    double* Sim = calloc((size_t)nh, sizeof(double));
#pragma omp parallel
    {
        double* lre = calloc((size_t)nh, sizeof(double));
        double* lim = calloc((size_t)nh, sizeof(double));
#pragma omp for schedule(static)
        for (int64_t j = 0; j < n; j++) {
            const double px = (double)x[j] * 0x1p-32, py = (double)y[j] * 0x1p-32, pz = (double)z[j] * 0x1p-32;
            for (int k = 0; k < nh; k++) {
                const double om = 2.0 * M_PI * (hl[3 * k] * px + hl[3 * k + 1] * py + hl[3 * k + 2] * pz);
                lre[k] += cos(om);
                lim[k] += sin(om);
            }
        }
#pragma omp critical
        for (int k = 0; k < nh; k++) {
            Sre[k] += lre[k];
            Sim[k] += lim[k];
        }
        free(lre);
        free(lim);
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t s = 0; s < ns; s++) {
        /* a sample particle: its own pair is left out and tsp a restores the
         * image + background term; a position: every particle counts */
        const int64_t i = samples ? samples[s] : -1;
        const uint32_t qx = samples ? x[i] : sx[s], qy = samples ? y[i] : sy[s], qz = samples ? z[i] : sz[s];
        double P = (double)n * M_PI / (a * a) + (samples ? tsp * a : 0.0);
        double G[3] = {0, 0, 0};
        for (int64_t j = 0; j < n; j++) {
            if (j == i) continue;
            const double dx = (double)(int32_t)(qx - x[j]) * 0x1p-32;
            const double dy = (double)(int32_t)(qy - y[j]) * 0x1p-32;
            const double dz = (double)(int32_t)(qz - z[j]) * 0x1p-32;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 >= rc2) continue;
            const double r = sqrt(r2);
            if (r2 < hsoft * hsoft) {
                double d0, d1;
                soft_poly(r2, hsoft, &d0, &d1);
                P -= d0;
                G[0] -= d1 * dx;
                G[1] -= d1 * dy;
                G[2] -= d1 * dz;
                if (r2 > 0.0) {
                    const double e = erf(a * r);
                    P += e / r;
                    const double dr = (tsp * a * exp(-a * a * r2) * r - e) / r2;
                    G[0] -= dr * dx / r;
                    G[1] -= dr * dy / r;
                    G[2] -= dr * dz / r;
                } else {
                    P += tsp * a;
                }
            } else {
                const double ec = erfc(a * r);
                P -= ec / r;
                const double dr = ec / r2 + tsp * a * exp(-a * a * r2) / r;
                G[0] -= dr * dx / r;
                G[1] -= dr * dy / r;
                G[2] -= dr * dz / r;
            }
        }
        const double px = (double)qx * 0x1p-32, py = (double)qy * 0x1p-32, pz = (double)qz * 0x1p-32;
        for (int k = 0; k < nh; k++) {
            const int hx = hl[3 * k], hy = hl[3 * k + 1], hz = hl[3 * k + 2];
            const int h2 = hx * hx + hy * hy + hz * hz;
            const double c = exp(-M_PI * M_PI * h2 / (a * a)) / h2;
            const double om = 2.0 * M_PI * (hx * px + hy * py + hz * pz);
            const double er = cos(om), ei = sin(om);
            const double re = er * Sre[k] + ei * Sim[k];
            const double im = ei * Sre[k] - er * Sim[k];
            P -= 2.0 / M_PI * c * re;
            G[0] -= 4.0 * c * hx * im;
            G[1] -= 4.0 * c * hy * im;
            G[2] -= 4.0 * c * hz * im;
        }
        phi[s] = P;
        gx[s] = G[0];
        gy[s] = G[1];
        gz[s] = G[2];
    }
    free(hl);
    free(Sre);
    free(Sim);
}

ORC_API void orc_ewald_sf(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int64_t ns,
                          const int64_t* samples, double hsoft, double alpha, double* phi, double* gx, double* gy,
                          double* gz) {
    ewald_sf_core(n, x, y, z, ns, samples, NULL, NULL, NULL, hsoft, alpha, phi, gx, gy, gz);
}

/* The same sum at arbitrary positions (gravity_short_ewald_call_kernel's
 * sinks, gravity_short.cu:10-91): the potential and force of all n unit
 * masses at each position, each softened inside hsoft; a particle sitting
 * exactly at the position adds -35/16 h^-1 to phi and nothing to g. */
ORC_API void orc_ewald_sf_at(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int64_t ns,
                             const uint32_t* sx, const uint32_t* sy, const uint32_t* sz, double hsoft, double alpha,
                             double* phi, double* gx, double* gy, double* gz) {
    ewald_sf_core(n, x, y, z, ns, NULL, sx, sy, sz, hsoft, alpha, phi, gx, gy, gz);
}

/* Pair-by-pair Ewald sum at arbitrary positions (alpha 2, images |n| <= 4,
 * |h| <= 4), the independent check of orc_ewald_sf_at: sum over every
 * particle of psi_soft(q - x_j), where inside hsoft the Newtonian -1/r of
 * psi is replaced by the softened polynomial (at r = 0: psi's regular
 * part minus 35/16 h^-1). */
ORC_API void orc_ewald_direct_at(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int64_t ns,
                                 const uint32_t* sx, const uint32_t* sy, const uint32_t* sz, double hsoft,
                                 double* phi, double* gx, double* gy, double* gz) {
    const double a = 2.0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t s = 0; s < ns; s++) {
        double P = 0.0, G[3] = {0, 0, 0};
        for (int64_t j = 0; j < n; j++) {
            const double dx = (double)(int32_t)(sx[s] - x[j]) * 0x1p-32;
            const double dy = (double)(int32_t)(sy[s] - y[j]) * 0x1p-32;
            const double dz = (double)(int32_t)(sz[s] - z[j]) * 0x1p-32;
            const double r2 = dx * dx + dy * dy + dz * dz;
            double ps, gr[3];
            if (r2 < hsoft * hsoft) {
                ewald_psi(dx, dy, dz, a, 4, 4, 1, &ps, gr);
                double d0, d1;
                soft_poly(r2, hsoft, &d0, &d1);
                P += ps - d0;
                G[0] -= gr[0] + d1 * dx;
                G[1] -= gr[1] + d1 * dy;
                G[2] -= gr[2] + d1 * dz;
            } else {
                ewald_psi(dx, dy, dz, a, 4, 4, 0, &ps, gr);
                P += ps;
                G[0] -= gr[0];
                G[1] -= gr[1];
                G[2] -= gr[2];
            }
        }
        phi[s] = P;
        gx[s] = G[0];
        gy[s] = G[1];
        gz[s] = G[2];
    }
}

ORC_API int orc_num_threads(void) { return omp_get_max_threads(); }
