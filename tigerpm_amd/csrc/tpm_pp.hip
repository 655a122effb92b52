// tpm_pp.hip -- fused short-range PP + mesh interpolation + kick.
//
// Reference: kick_treepm_kernel (kick_treepm.cu:460-759) and kick_pme_kernel
// (kick_pme.cu:128-357).  Work item here = one tile of <= TS consecutive
// sinks of one chain cell, processed by one wave64: lane l holds sink
// l % TS and walks source stream l / TS (64 / TS streams over one shared
// source list), the streams are summed at the end.
//
// Source staging: the wave walks the 27 neighbour cells (the candidate set
// of kick_pme.cu:265-318) 64 particles at a time; each lane tests one source
// against the integer bounding box of the tile's active sinks and survivors
// are compacted (ballot + mbcnt) into an LDS buffer as exact int32 offsets
// from the tile's first particle plus the source index.  A source whose box
// distance exceeds rcut has r >= rcut to every sink, so dropping it changes
// no result; smaller tiles have tighter boxes and drop more.  The inner loop
// reads one staged source per stream with a ds_read_b128 (prefetched one
// step ahead) and evaluates compute_pp_interaction (kick_treepm.cu:242-283).
//
// Pair distances are int32(sink - source) converted to float, which is the
// reference's distance() (gravity_short.hpp:20-22) in units of 2^-32; every
// constant is pre-scaled by the matching power of two so the FP32 arithmetic
// is the reference's up to exact power-of-two factors.  The self pair, which
// the reference evaluates inside the softening branch (it contributes
// -35/16 h^-1 to phi and nothing to g), is masked out of the loop and added
// once per sink, so the softening branch is taken only by distinct pairs
// closer than h.
#include "tpm_internal.hpp"

#include <climits>

#define SBUF 256  // staged sources per flush (4 KiB LDS)

namespace {

struct PPArgs {
    const uint32_t* __restrict__ x;
    const uint32_t* __restrict__ y;
    const uint32_t* __restrict__ z;
    float* __restrict__ vx;
    float* __restrict__ vy;
    float* __restrict__ vz;
    int8_t* __restrict__ rung;
    float* __restrict__ gx;
    float* __restrict__ gy;
    float* __restrict__ gz;
    float* __restrict__ pot;
    const int32_t* __restrict__ cell_begin;
    const int4* __restrict__ tiles;
    const int32_t* __restrict__ ntiles;
    const float* __restrict__ phi;
    unsigned long long* __restrict__ counters;
    int C, M;
    int min_rung, first_call, do_phi, cull;
    unsigned flags;
    // pair constants in 2^-32 box units
    float rcut2s, cull2s, inv2rs_s, tsp, h2s, h2inv_s, hinv_s, h3inv_s;
    float e2s;  // -log2(e) / (4 rs^2), 2^-64 units: exp(-x^2) = exp2(r2 * e2s)
    // epilogue
    float GM, phi0, eta, t0, scale, hsoft;
};

__device__ __forceinline__ unsigned lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ unsigned count_below(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

// compute_pp_interaction (kick_treepm.cu:242-283) with erfcexp
// (util.hpp:36-59, Abramowitz & Stegun 7.1.26), in 2^-32 units.  self is
// the sink's own particle index: that pair is excluded here.
__device__ __forceinline__ void pp_pair(const PPArgs& a, int ox, int oy, int oz, int self, int4 s, float& gx,
                                        float& gy, float& gz, float& ph) {
    const float dx = (float)(int)((unsigned)ox - (unsigned)s.x);
    const float dy = (float)(int)((unsigned)oy - (unsigned)s.y);
    const float dz = (float)(int)((unsigned)oz - (unsigned)s.z);
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    // Distinct particles differ by >= 1 raw unit, so r2 >= 1 unless this is
    // the self pair; the clamp keeps rinv^3 finite for it.
    const float rinv = __builtin_amdgcn_rsqf(fmaxf(r2, 1.0f));
    const float r = r2 * rinv;
    const float r0 = r * a.inv2rs_s;
    const float t1 = __builtin_amdgcn_rcpf(fmaf(0.3275911f, r0, 1.f));
    // The cutoff test (kick_treepm.cu:254) and the self exclusion zero the
    // Gaussian factor, which zeroes both erfc and the force factor below.
    const bool other = s.w != self;
    const bool in = r2 < a.rcut2s && other;
    const float e = in ? __builtin_amdgcn_exp2f(r2 * a.e2s) : 0.f;
    const float poly =
        t1 * fmaf(t1, fmaf(t1, fmaf(t1, fmaf(t1, 1.061405429f, -1.453152027f), 1.421413741f), -0.284496736f),
                  0.254829592f);
    const float erfc0 = poly * e;
    const float tre = a.tsp * r0 * e;
    float rinv3 = fmaf(a.tsp, r0, poly) * e * (rinv * rinv) * rinv;
    float rinv1 = rinv * erfc0;
    if (r2 <= a.h2s && other) {
        // Distinct particles closer than h: Plummer polynomial plus the
        // (erfc - 1) correction (kick_treepm.cu:263-276).
        const float ri = r2 > 0.f ? rinv : 0.f;
        const float q2 = r2 * a.h2inv_s;
        float d1 = 15.0f / 8.0f;
        d1 = fmaf(d1, q2, -21.0f / 4.0f);
        d1 = fmaf(d1, q2, 35.0f / 8.0f);
        d1 *= a.h3inv_s;
        rinv3 = ((erfc0 - 1.0f) + tre) * ri * ri * ri + d1;
        float d0 = -5.0f / 16.0f;
        d0 = fmaf(d0, q2, 21.0f / 16.0f);
        d0 = fmaf(d0, q2, -35.0f / 16.0f);
        d0 = fmaf(d0, q2, 35.0f / 16.0f);
        d0 *= a.hinv_s;
        rinv1 = (erfc0 - 1.0f) * ri + d0;
    }
    gx = fmaf(-dx, rinv3, gx);
    gy = fmaf(-dy, rinv3, gy);
    gz = fmaf(-dz, rinv3, gz);
    ph -= rinv1;
}

// 6-point quintic weights (kick_treepm.cu:527-549) in Horner form.
__device__ __forceinline__ void interp6_weights(float x, float w[6], float dw[6]) {
    w[0] = x * ((1.f / 12.f) + x * (-(1.f / 24.f) + x * (-(3.f / 8.f) + x * ((13.f / 24.f) + x * -(5.f / 24.f)))));
    w[1] = x * (-(2.f / 3.f) + x * ((2.f / 3.f) + x * ((13.f / 8.f) + x * (-(8.f / 3.f) + x * (25.f / 24.f)))));
    w[2] = 1.0f + x * x * (-(5.f / 4.f) + x * (-(35.f / 12.f) + x * ((21.f / 4.f) + x * -(25.f / 12.f))));
    w[3] = x * ((2.f / 3.f) + x * ((2.f / 3.f) + x * ((11.f / 4.f) + x * (-(31.f / 6.f) + x * (25.f / 12.f)))));
    w[4] = x * (-(1.f / 12.f) + x * (-(1.f / 24.f) + x * (-(11.f / 8.f) + x * ((61.f / 24.f) + x * -(25.f / 24.f)))));
    w[5] = x * x * x * ((7.f / 24.f) + x * (-0.5f + x * (5.f / 24.f)));
    dw[0] = (1.f / 12.f) + x * (-(2.f / 24.f) + x * (-(9.f / 8.f) + x * ((52.f / 24.f) + x * -(25.f / 24.f))));
    dw[1] = -(2.f / 3.f) + x * ((4.f / 3.f) + x * ((39.f / 8.f) + x * (-(32.f / 3.f) + x * (125.f / 24.f))));
    dw[2] = x * (-(10.f / 4.f) + x * (-(105.f / 12.f) + x * ((84.f / 4.f) + x * -(125.f / 12.f))));
    dw[3] = (2.f / 3.f) + x * ((4.f / 3.f) + x * ((33.f / 4.f) + x * (-(124.f / 6.f) + x * (125.f / 12.f))));
    dw[4] = -(1.f / 12.f) + x * (-(2.f / 24.f) + x * (-(33.f / 8.f) + x * ((244.f / 24.f) + x * -(125.f / 24.f))));
    dw[5] = x * x * ((21.f / 24.f) + x * (-2.0f + x * (25.f / 24.f)));
}

template <int TS>
__global__ __launch_bounds__(64) void pp_kick_kernel(PPArgs a) {
    constexpr int SPLIT = 64 / TS;  // source streams per wave
    __shared__ int4 sbuf[SBUF];
    // XCD-aware tile order: the hardware deals workgroups round-robin over
    // the 8 XCDs, so give each XCD one contiguous (spatially coherent) run of
    // tiles and let its L2 keep the shared neighbour cells.
    const int nt = *a.ntiles;
    const int per_xcd = (nt + 7) >> 3;
    const int k = blockIdx.x >> 3;
    if (k >= per_xcd) return;
    const int t = (int)(blockIdx.x & 7) * per_xcd + k;
    if (t >= nt) return;

    const int4 tile = a.tiles[t];
    const int lane = (int)lane_id();
    const int slot = lane % TS;
    const int stream = lane / TS;
    const int i = tile.y + slot;
    const bool valid = i < tile.z;
    const uint32_t rx = a.x[tile.y], ry = a.y[tile.y], rz = a.z[tile.y];
    uint32_t px = rx, py = ry, pz = rz;
    int r = 0;
    if (valid) {
        px = a.x[i];
        py = a.y[i];
        pz = a.z[i];
        r = a.rung[i];
    }
    const bool active = valid && r >= a.min_rung;
    const unsigned long long amask = __ballot(active);
    if (amask == 0ull) return;
    const int ox = (int)(px - rx), oy = (int)(py - ry), oz = (int)(pz - rz);

    float gx = 0.f, gy = 0.f, gz = 0.f, ph = 0.f;
    if (!(a.flags & TPM_SKIP_PP)) {
        // Tile bounding box as floats.  float(int32) rounds by <= 64 raw
        // units, so box distances below carry <= 192 units of error, inside
        // the 512-unit margin of cull2s: the test never drops a source with
        // r < rcut to any sink.
        const float lox = (float)wave_min(active ? ox : INT_MAX), hix = (float)wave_max(active ? ox : INT_MIN);
        const float loy = (float)wave_min(active ? oy : INT_MAX), hiy = (float)wave_max(active ? oy : INT_MIN);
        const float loz = (float)wave_min(active ? oz : INT_MAX), hiz = (float)wave_max(active ? oz : INT_MIN);
        const int C = a.C;
        const int cid = tile.x;
        const int cx = cid % C, cy = (cid / C) % C, cz = cid / (C * C);
        int nbuf = 0;
        unsigned long long staged = 0ull;
        // Far dummy that pads the list to a multiple of SPLIT: r >= rcut.
        const int4 far = make_int4(0x40000000, 0x40000000, 0x40000000, -1);
        auto flush = [&]() {
            const int padded = (nbuf + SPLIT - 1) / SPLIT * SPLIT;
            if (lane < padded - nbuf) sbuf[nbuf + lane] = far;
            __syncthreads();
            int4 s = sbuf[stream];
#pragma unroll 2
            for (int q = 0; q < padded; q += SPLIT) {
                const int4 sn = sbuf[min(q + SPLIT, padded - SPLIT) + stream];
                pp_pair(a, ox, oy, oz, i, s, gx, gy, gz, ph);
                s = sn;
            }
            staged += (unsigned long long)nbuf;
            nbuf = 0;
            __syncthreads();
        };
        for (int nb = 0; nb < 27; nb++) {
            int jx = cx + (nb % 3) - 1, jy = cy + (nb / 3) % 3 - 1, jz = cz + nb / 9 - 1;
            jx += (jx < 0) ? C : 0;
            jx -= (jx >= C) ? C : 0;
            jy += (jy < 0) ? C : 0;
            jy -= (jy >= C) ? C : 0;
            jz += (jz < 0) ? C : 0;
            jz -= (jz >= C) ? C : 0;
            const int c2 = (jz * C + jy) * C + jx;
            const int sb = a.cell_begin[c2], se = a.cell_begin[c2 + 1];
            for (int j0 = sb; j0 < se; j0 += 64) {
                const int j = j0 + lane;
                bool keep = false;
                int4 sv = make_int4(0, 0, 0, j);
                if (j < se) {
                    sv.x = (int)(a.x[j] - rx);
                    sv.y = (int)(a.y[j] - ry);
                    sv.z = (int)(a.z[j] - rz);
                    const float fx = (float)sv.x, fy = (float)sv.y, fz = (float)sv.z;
                    const float ex = fmaxf(fmaxf(lox - fx, fx - hix), 0.f);
                    const float ey = fmaxf(fmaxf(loy - fy, fy - hiy), 0.f);
                    const float ez = fmaxf(fmaxf(loz - fz, fz - hiz), 0.f);
                    keep = !a.cull || fmaf(ex, ex, fmaf(ey, ey, ez * ez)) < a.cull2s;
                }
                const unsigned long long m = __ballot(keep);
                if (keep) sbuf[nbuf + (int)count_below(m)] = sv;
                nbuf += __popcll(m);
                if (nbuf > SBUF - 2 * 64) flush();
            }
        }
        if (nbuf > 0) flush();
        // sum the source streams
#pragma unroll
        for (int o = TS; o < 64; o <<= 1) {
            gx += __shfl_xor(gx, o);
            gy += __shfl_xor(gy, o);
            gz += __shfl_xor(gz, o);
            ph += __shfl_xor(ph, o);
        }
        if (lane == 0 && a.counters)
            atomicAdd(&a.counters[0], staged * (unsigned long long)__popcll(amask & ((TS == 64) ? ~0ull : ((1ull << TS) - 1))));
        // the self pair (kick_treepm.cu:263-276 at r = 0): -35/16 h^-1
        ph -= (35.0f / 16.0f) * a.hinv_s;
        // back to box units: g in 2^-64, phi in 2^-32
        gx *= 0x1p64f;
        gy *= 0x1p64f;
        gz *= 0x1p64f;
        ph *= 0x1p32f;
    }
    if (!active || stream != 0) return;

    // Mesh force and potential (kick_treepm.cu:505-573) on the periodic mesh.
    float pgx = 0.f, pgy = 0.f, pgz = 0.f, pph = 0.f;
    if (a.phi && !(a.flags & TPM_SKIP_PM)) {
        const int M = a.M;
        const uint32_t pr[3] = {px, py, pz};
        int I0[3];
        float w[3][6], dw[3][6];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const uint64_t P = (uint64_t)pr[d] * (uint64_t)M;
            I0[d] = (int)(P >> 32) - 2 + M;
            const float X = (float)((uint32_t)P >> 8) * 0x1p-24f;
            interp6_weights(X, w[d], dw[d]);
        }
        const float Mf = (float)M;
        for (int aa = 0; aa < 6; aa++) {
            int jx = I0[0] + aa;
            jx -= (jx >= M) ? M : 0;
            jx -= (jx >= M) ? M : 0;
            for (int bb = 0; bb < 6; bb++) {
                int jy = I0[1] + bb;
                jy -= (jy >= M) ? M : 0;
                jy -= (jy >= M) ? M : 0;
                const float* row = a.phi + ((int64_t)jx * M + jy) * M;
                const float wxy = w[0][aa] * w[1][bb];
                const float dxy = dw[0][aa] * w[1][bb];
                const float xdy = w[0][aa] * dw[1][bb];
#pragma unroll
                for (int cc = 0; cc < 6; cc++) {
                    int jz = I0[2] + cc;
                    jz -= (jz >= M) ? M : 0;
                    jz -= (jz >= M) ? M : 0;
                    const float p0 = row[jz];
                    const float p1 = p0 * Mf;
                    pgx = fmaf(-dxy * w[2][cc], p1, pgx);
                    pgy = fmaf(-xdy * w[2][cc], p1, pgy);
                    pgz = fmaf(-wxy * dw[2][cc], p1, pgz);
                    pph = fmaf(wxy * w[2][cc], p0, pph);
                }
            }
        }
    }

    // Kick epilogue (kick_treepm.cu:716-754); rung_dt[r] = 2^-r.
    const float GM = a.GM;
    const float g0 = (pgx + gx) * GM;
    const float g1 = (pgy + gy) * GM;
    const float g2 = (pgz + gz) * GM;
    const float pt = a.do_phi ? (a.phi0 + pph + ph) * GM : 0.f;
    a.gx[i] = g0;
    a.gy[i] = g1;
    a.gz[i] = g2;
    a.pot[i] = pt;
    if (a.flags & TPM_SKIP_KICK) return;
    float vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
    float dt = 0.5f * ldexpf(1.0f, -r) * a.t0;
    if (!a.first_call) {
        vx = fmaf(g0, dt, vx);
        vy = fmaf(g1, dt, vy);
        vz = fmaf(g2, dt, vz);
    }
    const float gg = fmaf(g0, g0, fmaf(g1, g1, g2 * g2));
    const float factor = a.eta * sqrtf(a.scale * a.hsoft);
    dt = fminf(factor * (1.0f / sqrtf(sqrtf(gg))), a.t0);
    float nr = fmaxf(ceilf(log2f(a.t0) - log2f(dt)), (float)(r - 1));
    nr = fminf(fmaxf(nr, 0.f), 31.f);
    r = (int)nr;
    dt = 0.5f * ldexpf(1.0f, -r) * a.t0;
    vx = fmaf(g0, dt, vx);
    vy = fmaf(g1, dt, vy);
    vz = fmaf(g2, dt, vz);
    a.vx[i] = vx;
    a.vy[i] = vy;
    a.vz[i] = vz;
    a.rung[i] = (int8_t)r;
    if (a.counters) atomicMax(&a.counters[1], (unsigned long long)r);
}

// W_sr = sum_c n_c * sum_{27 neighbours} n_c' (exact, 64-bit integer).
__global__ void candidate_pairs_kernel(int C, const int32_t* __restrict__ cell_begin,
                                       unsigned long long* __restrict__ out) {
    const int64_t ncell = (int64_t)C * C * C;
    unsigned long long acc = 0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
        const int cx = (int)(c % C), cy = (int)((c / C) % C), cz = (int)(c / ((int64_t)C * C));
        const int64_t nc = cell_begin[c + 1] - cell_begin[c];
        int64_t s = 0;
        for (int nb = 0; nb < 27; nb++) {
            const int jx = (cx + nb % 3 - 1 + C) % C, jy = (cy + (nb / 3) % 3 - 1 + C) % C,
                      jz = (cz + nb / 9 - 1 + C) % C;
            const int64_t c2 = ((int64_t)jz * C + jy) * C + jx;
            s += cell_begin[c2 + 1] - cell_begin[c2];
        }
        acc += (unsigned long long)(nc * s);
    }
    atomicAdd(out, acc);
}

}  // namespace

tpm_status tpm_pp_kick(tpm_ctx* ctx, int min_rung, double scale, double t0, int first_call, unsigned flags,
                       tpm_stats* stats) {
    if (!ctx->chain_valid || ctx->chain_C != ctx->C || ctx->chain_tile != ctx->tile)
        return tpm_fail(ctx, TPM_ERR_STATE, "tpm_kick: chain mesh not built (call tpm_chainmesh_create)");
    const bool use_pm = !(flags & TPM_SKIP_PM);
    if (use_pm && (!ctx->pm_valid || ctx->green_M != ctx->M))
        return tpm_fail(ctx, TPM_ERR_STATE, "tpm_kick: PM mesh not computed (call tpm_gravity_long_compute)");
    if (t0 <= 0.0) return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_kick: t0 must be positive");
    hipStream_t s = ctx->stream;
    const tpm_params& p = ctx->p;
    if (!ctx->counters) TPM_HIP(ctx, hipMalloc(&ctx->counters, sizeof(unsigned long long) * 4));
    TPM_TRY(tpm_fill_u64(ctx, ctx->counters, 4));

    PPArgs a;
    a.x = ctx->x;
    a.y = ctx->y;
    a.z = ctx->z;
    a.vx = ctx->vx;
    a.vy = ctx->vy;
    a.vz = ctx->vz;
    a.rung = ctx->rung;
    a.gx = ctx->gx;
    a.gy = ctx->gy;
    a.gz = ctx->gz;
    a.pot = ctx->pot;
    a.cell_begin = ctx->cell_begin;
    a.tiles = ctx->tiles;
    a.ntiles = ctx->ntiles;
    a.phi = use_pm ? ctx->phi : nullptr;
    a.counters = ctx->counters;
    a.C = ctx->C;
    a.M = ctx->M;
    a.min_rung = min_rung;
    a.first_call = first_call ? 1 : 0;
    a.do_phi = p.do_phi;
    a.flags = flags;
    // Constants of compute_pp_interaction (kick_treepm.cu:833-853), scaled
    // to 2^-32 units: lengths x 2^32, squares x 2^64.
    const float rcut = (float)p.rcut, rs = (float)p.rs, h = (float)p.hsoft;
    const float S = 0x1p32f;
    a.rcut2s = (rcut * rcut) * (S * S);
    // Culling keeps a margin of 512 raw units over rcut (float rounding of
    // the box distance stays below ~200); off for C < 5, where a source's
    // offset from the tile origin can wrap to the far image.
    const double cull = (double)p.rcut * 4294967296.0 + 512.0;
    a.e2s = (float)(-1.4426950408889634 / (4.0 * p.rs * p.rs) / (4294967296.0 * 4294967296.0));
    a.cull2s = (float)(cull * cull);
    a.cull = ctx->C >= 5 ? 1 : 0;
    a.inv2rs_s = (1.0f / rs / 2.0f) / S;
    a.tsp = 2.0f / sqrtf((float)M_PI);
    a.h2s = (h * h) * (S * S);
    const float hinv = 1.f / h;
    a.hinv_s = hinv / S;
    a.h2inv_s = (hinv * hinv) / (S * S);
    a.h3inv_s = ((hinv * hinv * hinv) / (S * S)) / S;
    a.GM = (float)p.GM;
    a.phi0 = (float)(p.n_total * 4.0 * M_PI * p.rs * p.rs + (35.0 / 16.0) / p.hsoft);
    a.eta = (float)p.eta;
    a.t0 = (float)t0;
    a.scale = (float)scale;
    a.hsoft = (float)p.hsoft;

    const int64_t ncell = (int64_t)ctx->C * ctx->C * ctx->C;
    const int64_t tiles_max = ctx->n / ctx->tile + ncell + 1;
    const int64_t blocks = (tiles_max + 7) / 8 * 8;
    if (blocks > INT_MAX) return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_kick: too many tiles");
    if (!ctx->ev[0]) {
        for (int e = 0; e < 8; e++) TPM_HIP(ctx, hipEventCreate(&ctx->ev[e]));
    }
    hipEvent_t e0 = ctx->ev[2], e1 = ctx->ev[3];
    if (ctx->prof_on) {
        e0 = tpm_prof_event(ctx);
        e1 = tpm_prof_event(ctx);
        if (!e0 || !e1) return tpm_fail(ctx, TPM_ERR_HIP, "hipEventCreate failed");
        ctx->prof_pp.emplace_back(e0, e1);
    }
    TPM_HIP(ctx, hipEventRecord(e0, s));
    if (ctx->n > 0) {
        switch (ctx->tile) {
            case 16: pp_kick_kernel<16><<<(unsigned)blocks, 64, 0, s>>>(a); break;
            case 64: pp_kick_kernel<64><<<(unsigned)blocks, 64, 0, s>>>(a); break;
            default: pp_kick_kernel<32><<<(unsigned)blocks, 64, 0, s>>>(a); break;
        }
        TPM_HIP(ctx, hipGetLastError());
    }
    TPM_HIP(ctx, hipEventRecord(e1, s));
    if (stats) {
        candidate_pairs_kernel<<<grid_for(ncell, 256), 256, 0, s>>>(ctx->C, ctx->cell_begin, ctx->counters + 2);
        TPM_HIP(ctx, hipGetLastError());
        unsigned long long hc[4];
        int32_t nt = 0;
        TPM_HIP(ctx, hipMemcpyAsync(hc, ctx->counters, sizeof(hc), hipMemcpyDeviceToHost, s));
        TPM_HIP(ctx, hipMemcpyAsync(&nt, ctx->ntiles, sizeof(nt), hipMemcpyDeviceToHost, s));
        TPM_HIP(ctx, hipStreamSynchronize(s));
        float ms = 0.f;
        TPM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
        stats->executed_pairs = (double)hc[0];
        stats->max_rung = (int)hc[1];
        stats->candidate_pairs = (double)hc[2];
        stats->ms_pp = ms;
        stats->n_tiles = nt;
    }
    return TPM_OK;
}
