// tpm_fft.hip -- 3-D real FFTs of power-of-two cubic meshes (16 .. 1024),
// LDS-tiled, in the layouts of fft3d_execute / fft3d_inv_execute
// (fft.cpp:51-84): R2C unnormalised, C2R unnormalised (the PM folds 1/M^3
// into its Green's function, fft.cpp:468, 512).
//
//   real  r[ix][iy][iz]           M^3 floats, iz fastest
//   spec  S[kx][ky][kz]           M * M * H complex, H = M/2 + 1, kz fastest
//   work  A[ix][kz][iy]           M * H * M complex, iy fastest
//   work  B[kz][ky][ix]           H * M * M complex, ix fastest
//
// Forward:  r --Z--> A --Y--> B --X--> S       (A in the output buffer)
// Inverse:  S --X'-> B --Y'-> A --Z'-> r       (A in the input buffer)
//
// Every pass reads and writes each element once: a workgroup loads T whole
// rows of its input (each row contiguous in memory), transforms them in LDS
// and writes T-element runs of its output, so the transposes that put the
// next pass's rows in contiguous memory are part of the pass.  Z and Z' make
// a real row of M points from a complex transform of M/2 points
// (z_n = x_2n + i x_2n+1) with the usual split of even and odd parts.
//
// The 1-D transforms are Stockham autosort FFTs (radix 8, 4 and 2 stages;
// each stage reads its butterflies' inputs from LDS, twiddles, applies the
// radix-R DFT and writes the outputs at their autosorted places), the
// twiddles read from a table e^{-2 pi i q / M} built in FP64.  The LDS tile
// holds T transforms with element (t, idx) at idx*T + (t ^ (idx % T)): rows
// of T consecutive t are whole bank rows, and the xor spreads a single
// transform's consecutive elements over the banks when rows are loaded.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tpm_internal.hpp"

namespace {

constexpr int FFT_THREADS = 256;

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 conjc(float2 a) { return make_float2(a.x, -a.y); }
// a * (DIR i)
template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a) {
    return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// In-place DFTs of radix 2, 4, 8 with kernel e^{DIR 2 pi i r k / R},
// natural order in and out.
template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<2, DIR> {
    static __device__ __forceinline__ void run(float2* v) {
        const float2 a = v[0];
        v[0] = cadd(a, v[1]);
        v[1] = csub(a, v[1]);
    }
};

template <int DIR>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
    const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_i<DIR>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a1 = cadd(t1, t3);
    a2 = csub(t0, t2);
    a3 = csub(t1, t3);
}

template <int DIR>
struct Dft<4, DIR> {
    static __device__ __forceinline__ void run(float2* v) { dft4<DIR>(v[0], v[1], v[2], v[3]); }
};

template <int DIR>
struct Dft<8, DIR> {
    static __device__ __forceinline__ void run(float2* v) {
        // even and odd halves by radix 4, then y_k = E_k + w^k O_k and
        // y_k+4 = E_k - w^k O_k with w = e^{DIR 2 pi i / 8}
        float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4<DIR>(e0, e1, e2, e3);
        dft4<DIR>(o0, o1, o2, o3);
        const float h = 0.70710678118654752f;
        const float d = (float)DIR;
        o1 = make_float2(h * (o1.x - d * o1.y), h * (o1.y + d * o1.x));    // (1 + DIR i) / sqrt 2
        o2 = mul_i<DIR>(o2);                                                // DIR i
        o3 = make_float2(h * (-o3.x - d * o3.y), h * (-o3.y + d * o3.x));  // (-1 + DIR i) / sqrt 2
        v[0] = cadd(e0, o0);
        v[4] = csub(e0, o0);
        v[1] = cadd(e1, o1);
        v[5] = csub(e1, o1);
        v[2] = cadd(e2, o2);
        v[6] = csub(e2, o2);
        v[3] = cadd(e3, o3);
        v[7] = csub(e3, o3);
    }
};

template <int T>
__device__ __forceinline__ int lds_at(int t, int idx) {
    return idx * T + (t ^ (idx & (T - 1)));
}

// f(e) for e = threadIdx.x + 256 u < COUNT, fully unrolled, so that a
// phase's global loads are all in flight before the first wait.
template <int COUNT, class F>
__device__ __forceinline__ void for_each_elem(F f) {
    constexpr int NU = (COUNT + FFT_THREADS - 1) / FFT_THREADS;
#pragma unroll
    for (int u = 0; u < NU; u++) {
        const int e = (int)threadIdx.x + FFT_THREADS * u;
        if (COUNT % FFT_THREADS == 0 || e < COUNT) f(e);
    }
}

// Loads for e < COUNT into registers, then stores: every load of the phase
// is issued before any result is used.
template <int COUNT, class L, class S>
__device__ __forceinline__ void move_elems(L load, S store) {
    constexpr int NU = (COUNT + FFT_THREADS - 1) / FFT_THREADS;
    float2 v[NU];
#pragma unroll
    for (int u = 0; u < NU; u++) {
        const int e = (int)threadIdx.x + FFT_THREADS * u;
        if (COUNT % FFT_THREADS == 0 || e < COUNT) v[u] = load(e);
    }
#pragma unroll
    for (int u = 0; u < NU; u++) {
        const int e = (int)threadIdx.x + FFT_THREADS * u;
        if (COUNT % FFT_THREADS == 0 || e < COUNT) store(e, v[u]);
    }
}

// One Stockham stage of radix R over T transforms of length L whose first
// Ns-point sub-transforms are done: butterfly j (k = j mod Ns) takes
// x[j + r L/R] times e^{DIR 2 pi i r k / (Ns R)} and writes the R-point DFT
// to y[(j - k) R + k + r Ns].  Twiddles come from tf[q] = e^{-2 pi i q / M},
// M = L * SCALE.
template <int L, int T, int R, int Ns, int DIR, int SCALE>
__device__ __forceinline__ void fft_stage(float2* s, const float2* __restrict__ tf) {
    constexpr int NB = T * L / R;
    constexpr int NV = (NB + FFT_THREADS - 1) / FFT_THREADS;
    const int tid = (int)threadIdx.x;
    float2 v[NV][R];
#pragma unroll
    for (int u = 0; u < NV; u++) {
        const int b = tid + FFT_THREADS * u;
        if (NB % FFT_THREADS == 0 || b < NB) {
            const int t = b % T, j = b / T;
            const int k = j & (Ns - 1);
#pragma unroll
            for (int r = 0; r < R; r++) v[u][r] = s[lds_at<T>(t, j + r * (L / R))];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    const float2 w = tf[r * k * (L / (Ns * R)) * SCALE];
                    v[u][r] = cmul(v[u][r], DIR < 0 ? w : conjc(w));
                }
            }
            Dft<R, DIR>::run(v[u]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NV; u++) {
        const int b = tid + FFT_THREADS * u;
        if (NB % FFT_THREADS == 0 || b < NB) {
            const int t = b % T, j = b / T;
            const int k = j & (Ns - 1);
            const int i0 = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) s[lds_at<T>(t, i0 + r * Ns)] = v[u][r];
        }
    }
    __syncthreads();
}

// All stages: radix 8 while at least 8 points remain, except that 16 points
// go as 4 x 4 and 4 points as one radix-4 stage, 2 as radix 2.
template <int L, int T, int DIR, int SCALE, int Ns = 1>
__device__ __forceinline__ void fft_rows(float2* s, const float2* __restrict__ tf) {
    if constexpr (Ns < L) {
        constexpr int rem = ilog2c(L / Ns);
        constexpr int R = rem == 1 ? 2 : (rem == 2 || rem == 4) ? 4 : 8;
        fft_stage<L, T, R, Ns, DIR, SCALE>(s, tf);
        fft_rows<L, T, DIR, SCALE, Ns * R>(s, tf);
    }
}

// Z: real rows r[ix][iy0 + t][*] -> A[ix][kz][iy0 + t].  Grid M * (M / T).
// With Q32 the rows are the integer density grid (unit masses times 2^32,
// 16-byte aligned), converted as float(double(q) 2^-32) on the way in.
template <int M, int T, bool Q32>
__global__ __launch_bounds__(FFT_THREADS) void fft_z_r2c_kernel(const void* __restrict__ in, float2* __restrict__ A,
                                                               const float2* __restrict__ tf) {
    constexpr int N = M / 2, H = N + 1;
    __shared__ float2 s[T * N];
    const int iy0 = (int)(blockIdx.x % (M / T)) * T;
    const int64_t ix = blockIdx.x / (M / T);
    const int64_t row0 = (ix * M + iy0) * M;
    const float* src = static_cast<const float*>(in) + row0;
    const ulonglong2* qsrc = reinterpret_cast<const ulonglong2*>(static_cast<const unsigned long long*>(in) + row0);
    const bool aligned = (reinterpret_cast<uintptr_t>(src) & 7) == 0;
    move_elems<T * N>(
        [&](int e) {
            if (Q32) {
                const ulonglong2 q = qsrc[e];
                return make_float2((float)((double)q.x * 0x1p-32), (float)((double)q.y * 0x1p-32));
            }
            if (aligned) return reinterpret_cast<const float2*>(src)[e];
            return make_float2(src[2 * e], src[2 * e + 1]);
        },
        [&](int e, float2 v) { s[lds_at<T>(e / N, e % N)] = v; });
    __syncthreads();
    fft_rows<N, T, -1, 2>(s, tf);
    // X_k = (Z_k + conj Z_N-k) / 2 - i W^k (Z_k - conj Z_N-k) / 2, W = e^{-2 pi i / M}
    float2* dst = A + ix * H * M + iy0;
    for_each_elem<T * H>([&](int e) {
        const int k = e / T, t = e % T;
        const float2 a = s[lds_at<T>(t, k & (N - 1))];
        const float2 b = conjc(s[lds_at<T>(t, (N - k) & (N - 1))]);
        const float2 ev = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
        const float2 od = mul_i<-1>(make_float2(0.5f * (a.x - b.x), 0.5f * (a.y - b.y)));
        dst[(int64_t)k * M + t] = cadd(ev, cmul(tf[k & (M - 1)], od));
    });
}

// Z': A[ix][kz][iy0 + t] -> real rows r[ix][iy0 + t][*].  Grid M * (M / T).
template <int M, int T>
__global__ __launch_bounds__(FFT_THREADS) void fft_z_c2r_kernel(const float2* __restrict__ A, float* __restrict__ out,
                                                               const float2* __restrict__ tf) {
    constexpr int N = M / 2, H = N + 1;
    constexpr int NV = (T * N + FFT_THREADS - 1) / FFT_THREADS;
    __shared__ float2 s[T * H];
    const int tid = (int)threadIdx.x;
    const int iy0 = (int)(blockIdx.x % (M / T)) * T;
    const int64_t ix = blockIdx.x / (M / T);
    const float2* src = A + ix * H * M + iy0;
    move_elems<T * H>([&](int e) { return src[(int64_t)(e / T) * M + e % T]; },
                      [&](int e, float2 v) { s[lds_at<T>(e % T, e / T)] = v; });
    __syncthreads();
    // Z'_k = (X_k + conj X_N-k) + i (X_k - conj X_N-k) conj(W^k), k < N;
    // the inverse N-point transform of Z' is x_2n + i x_2n+1 times M
    float2 z[NV];
#pragma unroll
    for (int u = 0; u < NV; u++) {
        const int e = tid + FFT_THREADS * u;
        if ((T * N) % FFT_THREADS == 0 || e < T * N) {
            const int k = e / T, t = e % T;
            const float2 a = s[lds_at<T>(t, k)];
            const float2 b = conjc(s[lds_at<T>(t, N - k)]);
            z[u] = cadd(cadd(a, b), mul_i<1>(cmul(csub(a, b), conjc(tf[k]))));
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NV; u++) {
        const int e = tid + FFT_THREADS * u;
        if ((T * N) % FFT_THREADS == 0 || e < T * N) s[lds_at<T>(e % T, e / T)] = z[u];
    }
    __syncthreads();
    fft_rows<N, T, 1, 2>(s, tf);
    float* dst = out + (ix * M + iy0) * M;
    const bool aligned = (reinterpret_cast<uintptr_t>(dst) & 7) == 0;
    for_each_elem<T * N>([&](int e) {
        const float2 v = s[lds_at<T>(e / N, e % N)];
        if (aligned) {
            reinterpret_cast<float2*>(dst)[e] = v;
        } else {
            dst[2 * e] = v.x;
            dst[2 * e + 1] = v.y;
        }
    });
}

// Y (FWD): A[ix0 + t][kz][*] -> B[kz][ky][ix0 + t];  Y' (!FWD): the reverse.
// Grid H * (M / T).
template <int M, int T, bool FWD>
__global__ __launch_bounds__(FFT_THREADS) void fft_y_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                                           const float2* __restrict__ tf) {
    constexpr int H = M / 2 + 1;
    __shared__ float2 s[T * M];
    const int ix0 = (int)(blockIdx.x % (M / T)) * T;
    const int64_t kz = blockIdx.x / (M / T);
    if (FWD)
        move_elems<T * M>([&](int e) { return in[((int64_t)(ix0 + e / M) * H + kz) * M + e % M]; },
                          [&](int e, float2 v) { s[lds_at<T>(e / M, e % M)] = v; });
    else
        move_elems<T * M>([&](int e) { return in[(kz * M + e / T) * M + ix0 + e % T]; },
                          [&](int e, float2 v) { s[lds_at<T>(e % T, e / T)] = v; });
    __syncthreads();
    fft_rows<M, T, FWD ? -1 : 1, 1>(s, tf);
    if (FWD)
        for_each_elem<T * M>([&](int e) { out[(kz * M + e / T) * M + ix0 + e % T] = s[lds_at<T>(e % T, e / T)]; });
    else
        for_each_elem<T * M>(
            [&](int e) { out[((int64_t)(ix0 + e / M) * H + kz) * M + e % M] = s[lds_at<T>(e / M, e % M)]; });
}

// X (FWD): B[kz0 + t][ky][*] -> S[kx][ky][kz0 + t];  X' (!FWD): the reverse.
// Grid M * ceil(H / T); rows past kz = M/2 are zeros and not stored.  With
// a Green's table g (FWD only, layout of S) each stored mode is g times the
// transform, as a separate multiply pass would make it.
template <int M, int T, bool FWD>
__global__ __launch_bounds__(FFT_THREADS) void fft_x_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                                           const float2* __restrict__ tf,
                                                           const float* __restrict__ g = nullptr) {
    constexpr int H = M / 2 + 1;
    constexpr int NKT = (H + T - 1) / T;
    __shared__ float2 s[T * M];
    const int kz0 = (int)(blockIdx.x % NKT) * T;
    const int64_t ky = blockIdx.x / NKT;
    const int tn = min(T, H - kz0);
    const float2 zero = make_float2(0.f, 0.f);
    // the Green's factors of this thread's stored modes, loaded before any
    // store (a load issued after stores waits for them too)
    constexpr int NU = T * M / FFT_THREADS;
    static_assert(T * M % FFT_THREADS == 0, "whole tiles per thread");
    float gv[FWD ? NU : 1];
    if (FWD && g) {
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int e = (int)threadIdx.x + FFT_THREADS * u;
            gv[u] = e % T < tn ? g[((int64_t)(e / T) * M + ky) * H + kz0 + e % T] : 0.f;
        }
    }
    if (FWD)
        move_elems<T * M>(
            [&](int e) { return e / M < tn ? in[((int64_t)(kz0 + e / M) * M + ky) * M + e % M] : zero; },
            [&](int e, float2 v) { s[lds_at<T>(e / M, e % M)] = v; });
    else
        move_elems<T * M>(
            [&](int e) { return e % T < tn ? in[((int64_t)(e / T) * M + ky) * H + kz0 + e % T] : zero; },
            [&](int e, float2 v) { s[lds_at<T>(e % T, e / T)] = v; });
    __syncthreads();
    fft_rows<M, T, FWD ? -1 : 1, 1>(s, tf);
    if (FWD) {
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int e = (int)threadIdx.x + FFT_THREADS * u;
            if (e % T < tn) {
                float2 v = s[lds_at<T>(e % T, e / T)];
                if (g) {
                    v.x *= gv[u];
                    v.y *= gv[u];
                }
                out[((int64_t)(e / T) * M + ky) * H + kz0 + e % T] = v;
            }
        }
    }
    else
        for_each_elem<T * M>([&](int e) {
            if (e / M < tn) out[((int64_t)(kz0 + e / M) * M + ky) * M + e % M] = s[lds_at<T>(e / M, e % M)];
        });
}

// X, Green's multiply and X' in one pass, in place on B (the PM's Poisson
// solve): the rows B[kz0 + t][ky][*] go forward along x, each mode (kx, ky,
// kz0 + t) is multiplied by its Green's factor, and the rows go back.  The
// same float operations as X, a Green's pass and X', without storing S.
// gt is the table regrouped by green_tiles_kernel: workgroup w reads the
// T * M factors gt[w * T * M ...] in one contiguous run.
// Grid M * ceil(H / T).
template <int M, int T>
__global__ __launch_bounds__(FFT_THREADS) void fft_xgx_kernel(float2* __restrict__ b, const float2* __restrict__ tf,
                                                             const float* __restrict__ gt) {
    constexpr int H = M / 2 + 1;
    constexpr int NKT = (H + T - 1) / T;
    constexpr int NU = T * M / FFT_THREADS;
    static_assert(T * M % FFT_THREADS == 0, "whole tiles per thread");
    __shared__ float2 s[T * M];
    const int kz0 = (int)(blockIdx.x % NKT) * T;
    const int64_t ky = blockIdx.x / NKT;
    const int tn = min(T, H - kz0);
    float gv[NU];
#pragma unroll
    for (int u = 0; u < NU; u++) gv[u] = gt[(int64_t)blockIdx.x * (T * M) + threadIdx.x + FFT_THREADS * u];
    move_elems<T * M>(
        [&](int e) {
            return e / M < tn ? b[((int64_t)(kz0 + e / M) * M + ky) * M + e % M] : make_float2(0.f, 0.f);
        },
        [&](int e, float2 v) { s[lds_at<T>(e / M, e % M)] = v; });
    __syncthreads();
    fft_rows<M, T, -1, 1>(s, tf);
#pragma unroll
    for (int u = 0; u < NU; u++) {
        const int e = (int)threadIdx.x + FFT_THREADS * u;
        float2& v = s[lds_at<T>(e % T, e / T)];
        v.x *= gv[u];
        v.y *= gv[u];
    }
    __syncthreads();
    fft_rows<M, T, 1, 1>(s, tf);
    for_each_elem<T * M>([&](int e) {
        if (e / M < tn) b[((int64_t)(kz0 + e / M) * M + ky) * M + e % M] = s[lds_at<T>(e / M, e % M)];
    });
}

#ifndef TPM_FFT_TILE
#define TPM_FFT_TILE 16
#endif
template <int M>
constexpr int tile_of() {
    return M * TPM_FFT_TILE <= 8192 ? TPM_FFT_TILE : 8192 / M;  // T * M complex <= 64 KiB of LDS
}

// Rows per workgroup of the fused x pass.  Its loads and stores are whole
// rows of B whatever T is, so T only sets the LDS per workgroup: with 4 rows
// (16 KiB at M = 512) eight workgroups share a CU, against two at 16 rows
// (64 KiB), and the pass takes 0.29 instead of 0.69 ms at M = 512 (8 rows:
// 0.30 ms).
#ifndef TPM_FFT_XGX_TILE
#define TPM_FFT_XGX_TILE 4
#endif
template <int M>
constexpr int xgx_tile_of() {
    constexpr int t = M * TPM_FFT_XGX_TILE <= 8192 ? TPM_FFT_XGX_TILE : 8192 / M;
    return t * M < FFT_THREADS ? FFT_THREADS / M : t;
}

// The Green's table (layout of S: [kx][ky][kz], kz < H) regrouped for
// fft_xgx_kernel: the factors of its workgroup w = ky * NKT + kz tile, in the
// order e = kx * T + t its threads use, zeros past kz = M/2.  The values are
// copied, not recomputed.
template <int M>
__global__ void green_tiles_kernel(const float* __restrict__ g, float* __restrict__ gt) {
    constexpr int T = xgx_tile_of<M>(), H = M / 2 + 1, NKT = (H + T - 1) / T;
    const int64_t total = (int64_t)M * NKT * T * M;
    for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < total; l += (int64_t)gridDim.x * blockDim.x) {
        const int e = (int)(l % (T * M));
        const int64_t w = l / (T * M);
        const int kz = (int)(w % NKT) * T + e % T;
        const int64_t ky = w / NKT;
        gt[l] = kz < H ? g[((int64_t)(e / T) * M + ky) * H + kz] : 0.f;
    }
}

template <int M>
int64_t green_tiles(hipStream_t st, const float* g, float* gt) {
    constexpr int T = xgx_tile_of<M>(), H = M / 2 + 1;
    const int64_t n = (int64_t)M * ((H + T - 1) / T) * T * M;
    if (gt) green_tiles_kernel<M><<<(unsigned)std::min<int64_t>((n + 255) / 256, 65536), 256, 0, st>>>(g, gt);
    return n;
}

template <int M>
void launch_r2c(hipStream_t st, const void* in, bool q32, float2* out, float2* work, const float2* tf,
                const float* green) {
    constexpr int T = tile_of<M>(), H = M / 2 + 1;
    if (q32)
        fft_z_r2c_kernel<M, T, true><<<M * (M / T), FFT_THREADS, 0, st>>>(in, out, tf);
    else
        fft_z_r2c_kernel<M, T, false><<<M * (M / T), FFT_THREADS, 0, st>>>(in, out, tf);
    fft_y_kernel<M, T, true><<<H * (M / T), FFT_THREADS, 0, st>>>(out, work, tf);
    fft_x_kernel<M, T, true><<<M * ((H + T - 1) / T), FFT_THREADS, 0, st>>>(work, out, tf, green);
}

template <int M>
void launch_c2r(hipStream_t st, float2* in, float* out, float2* work, const float2* tf) {
    constexpr int T = tile_of<M>(), H = M / 2 + 1;
    fft_x_kernel<M, T, false><<<M * ((H + T - 1) / T), FFT_THREADS, 0, st>>>(in, work, tf);
    fft_y_kernel<M, T, false><<<H * (M / T), FFT_THREADS, 0, st>>>(work, in, tf);
    fft_z_c2r_kernel<M, T><<<M * (M / T), FFT_THREADS, 0, st>>>(in, out, tf);
}

// The PM's Poisson solve: Q32 density -> phi through spec (A layout) and
// work (B layout), both M * M * (M/2 + 1) complex.
template <int M>
void launch_pm(hipStream_t st, const unsigned long long* q, float* phi, float2* spec, float2* work,
               const float2* tf, const float* green) {
    constexpr int T = tile_of<M>(), H = M / 2 + 1;
    fft_z_r2c_kernel<M, T, true><<<M * (M / T), FFT_THREADS, 0, st>>>(q, spec, tf);
    fft_y_kernel<M, T, true><<<H * (M / T), FFT_THREADS, 0, st>>>(spec, work, tf);
    constexpr int TX = xgx_tile_of<M>();
    fft_xgx_kernel<M, TX><<<M * ((H + TX - 1) / TX), FFT_THREADS, 0, st>>>(work, tf, green);
    fft_y_kernel<M, T, false><<<H * (M / T), FFT_THREADS, 0, st>>>(work, spec, tf);
    fft_z_c2r_kernel<M, T><<<M * (M / T), FFT_THREADS, 0, st>>>(spec, phi, tf);
}

}  // namespace

bool tpm_lds_fft_size(int M) { return M >= 16 && M <= 1024 && (M & (M - 1)) == 0; }

tpm_status tpm_lds_fft_twiddles(std::string& err, int M, float2** tf) {
    std::vector<float2> h(M);
    for (int q = 0; q < M; q++) {
        const double a = -2.0 * M_PI * (double)q / (double)M;
        h[q] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    TPM_HIPE(err, hipMalloc(tf, sizeof(float2) * M));
    TPM_HIPE(err, hipMemcpy(*tf, h.data(), sizeof(float2) * M, hipMemcpyHostToDevice));
    return TPM_OK;
}

tpm_status tpm_lds_fft_r2c(std::string& err, hipStream_t st, int M, const void* in, bool q32, float2* out,
                           float2* work, const float2* tf, const float* green) {
    switch (M) {
        case 16: launch_r2c<16>(st, in, q32, out, work, tf, green); break;
        case 32: launch_r2c<32>(st, in, q32, out, work, tf, green); break;
        case 64: launch_r2c<64>(st, in, q32, out, work, tf, green); break;
        case 128: launch_r2c<128>(st, in, q32, out, work, tf, green); break;
        case 256: launch_r2c<256>(st, in, q32, out, work, tf, green); break;
        case 512: launch_r2c<512>(st, in, q32, out, work, tf, green); break;
        case 1024: launch_r2c<1024>(st, in, q32, out, work, tf, green); break;
        default: err = "LDS FFT: unsupported size " + std::to_string(M); return TPM_ERR_FFT;
    }
    TPM_HIPE(err, hipGetLastError());
    return TPM_OK;
}

tpm_status tpm_lds_fft_c2r(std::string& err, hipStream_t st, int M, float2* in, float* out, float2* work,
                           const float2* tf) {
    switch (M) {
        case 16: launch_c2r<16>(st, in, out, work, tf); break;
        case 32: launch_c2r<32>(st, in, out, work, tf); break;
        case 64: launch_c2r<64>(st, in, out, work, tf); break;
        case 128: launch_c2r<128>(st, in, out, work, tf); break;
        case 256: launch_c2r<256>(st, in, out, work, tf); break;
        case 512: launch_c2r<512>(st, in, out, work, tf); break;
        case 1024: launch_c2r<1024>(st, in, out, work, tf); break;
        default: err = "LDS FFT: unsupported size " + std::to_string(M); return TPM_ERR_FFT;
    }
    TPM_HIPE(err, hipGetLastError());
    return TPM_OK;
}

tpm_status tpm_lds_green_tiles(std::string& err, hipStream_t st, int M, const float* green, float* gt,
                               int64_t* count) {
    int64_t n = 0;
    switch (M) {
        case 16: n = green_tiles<16>(st, green, gt); break;
        case 32: n = green_tiles<32>(st, green, gt); break;
        case 64: n = green_tiles<64>(st, green, gt); break;
        case 128: n = green_tiles<128>(st, green, gt); break;
        case 256: n = green_tiles<256>(st, green, gt); break;
        case 512: n = green_tiles<512>(st, green, gt); break;
        case 1024: n = green_tiles<1024>(st, green, gt); break;
        default: err = "LDS FFT: unsupported size " + std::to_string(M); return TPM_ERR_FFT;
    }
    if (count) *count = n;
    TPM_HIPE(err, hipGetLastError());
    return TPM_OK;
}

tpm_status tpm_lds_pm_solve(std::string& err, hipStream_t st, int M, const unsigned long long* q, float* phi,
                            float2* spec, float2* work, const float2* tf, const float* green) {
    switch (M) {
        case 16: launch_pm<16>(st, q, phi, spec, work, tf, green); break;
        case 32: launch_pm<32>(st, q, phi, spec, work, tf, green); break;
        case 64: launch_pm<64>(st, q, phi, spec, work, tf, green); break;
        case 128: launch_pm<128>(st, q, phi, spec, work, tf, green); break;
        case 256: launch_pm<256>(st, q, phi, spec, work, tf, green); break;
        case 512: launch_pm<512>(st, q, phi, spec, work, tf, green); break;
        case 1024: launch_pm<1024>(st, q, phi, spec, work, tf, green); break;
        default: err = "LDS FFT: unsupported size " + std::to_string(M); return TPM_ERR_FFT;
    }
    TPM_HIPE(err, hipGetLastError());
    return TPM_OK;
}
