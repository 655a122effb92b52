// tpm_pp_device.hpp -- device code shared by the fused PP kernel
// (tpm_pp.hip) and the tree-walk kernel (tpm_tree.hip): wave primitives,
// the pair law, 6-point mesh interpolation and the kick epilogue.
//
// Every fused multiply-add here is written out (fmaf) and the compiler may
// fuse nothing else, so each includer produces the same bits.
#pragma once

#include "tpm_internal.hpp"

#pragma clang fp contract(off)

namespace {

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
// (util.hpp:36-59, Abramowitz & Stegun 7.1.26), in 2^-32 units.  The sink's
// own particle is one of its sources, as in the reference: at r = 0 the
// softening branch gives no force and exactly -35/16 h^-1 to phi.
// PME = true: kick_pme_kernel's law instead (kick_pme.cu:280-310): no cutoff,
// and inside h the plain softened polynomials without the (erfc - 1) terms.
// pp_law gives the pair's float offset (sink - source) and its force and
// potential factors; pp_pair adds them to the sink's sums.
template <bool PME>
__device__ __forceinline__ void pp_law(const tpm_box_dev& a, int ox, int oy, int oz, int4 s, float& dx, float& dy,
                                       float& dz, float& rinv3, float& rinv1) {
    dx = (float)(int)((unsigned)ox - (unsigned)s.x);
    dy = (float)(int)((unsigned)oy - (unsigned)s.y);
    dz = (float)(int)((unsigned)oz - (unsigned)s.z);
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    // Distinct particles differ by >= 1 raw unit, so r2 = 0 only for the
    // sink itself: there rinv = inf and every value below that depends on it
    // is replaced in the softening branch.
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    // x = r / (2 rs); t = 1 / (1 + p x) with p / (2 rs) folded into one constant
    const float t1 = __builtin_amdgcn_rcpf(fmaf(a.p_inv2rs_s, r, 1.f));
    // exp(-x^2); the cutoff test (kick_treepm.cu:254) zeroes it, which
    // zeroes both erfc and the force factor below.
    const float e = (PME || r2 < a.rcut2s) ? __builtin_amdgcn_exp2f(r2 * a.e2s) : 0.f;
    const float poly =
        t1 * fmaf(t1, fmaf(t1, fmaf(t1, fmaf(t1, 1.061405429f, -1.453152027f), 1.421413741f), -0.284496736f),
                  0.254829592f);
    const float er = e * rinv;
    // erfc(x) / r and (erfc(x) + 2 x / sqrt(pi) exp(-x^2)) / r^3, the latter
    // as (erfc(x) / r + (2 / sqrt(pi)) / (2 rs) exp(-x^2)) / r^2
    rinv1 = poly * er;
    rinv3 = fmaf(a.tsp_inv2rs_s, e, rinv1) * (rinv * rinv);
    if (r2 <= a.h2s) {
        // Closer than h: Plummer polynomial plus the (erfc - 1) correction
        // (kick_treepm.cu:263-276).  At r = 0 (the sink itself) only d1
        // (times dx = 0) and d0 remain.
        const float q2 = r2 * a.h2inv_s;
        float d1 = 15.0f / 8.0f;
        d1 = fmaf(d1, q2, -21.0f / 4.0f);
        d1 = fmaf(d1, q2, 35.0f / 8.0f);
        d1 *= a.h3inv_s;
        float d0 = -5.0f / 16.0f;
        d0 = fmaf(d0, q2, 21.0f / 16.0f);
        d0 = fmaf(d0, q2, -35.0f / 16.0f);
        d0 = fmaf(d0, q2, 35.0f / 16.0f);
        d0 *= a.hinv_s;
        if (PME) {
            rinv3 = d1;
            rinv1 = d0;
        } else {
            const bool pos = r2 > 0.f;
            const float ri = pos ? rinv : 0.f;
            const float em1 = pos ? fmaf(poly, e, -1.0f) : 0.f;
            const float tre = pos ? a.tsp_inv2rs_s * r * e : 0.f;
            // (erfc(x) - 1) / r and (erfc(x) - 1 + 2 x e^{-x^2} / sqrt(pi)) / r^3
            // are differences of O(1) terms that cancel to O(x) and O(x^3):
            // at r << h the formula as written keeps no digits in FP32 (1 %
            // off at r = 0.013 h).  For x^2 = u < 0.04 their Taylor series are
            // exact to 3e-8:  -(2 / sqrt(pi)) / (2 rs) (1 - u/3 + u^2/10 - u^3/42)
            // and -(4 / (3 sqrt(pi))) / (2 rs)^3 (1 - 3u/5 + 3u^2/14 - u^3/18).
            // At r = 0 (the sink itself) both terms stay zero, as written.
            const float u = r2 * a.inv4rs2_s;
            const float s1 = fmaf(fmaf(fmaf(-1.0f / 42.0f, u, 0.1f), u, -1.0f / 3.0f), u, 1.0f);
            const float s3 = fmaf(fmaf(fmaf(-1.0f / 18.0f, u, 3.0f / 14.0f), u, -0.6f), u, 1.0f);
            const bool ser = pos && u < 0.04f;
            rinv3 = ser ? fmaf(-a.k3_s, s3, d1) : fmaf((em1 + tre) * ri * ri, ri, d1);
            rinv1 = ser ? fmaf(-a.tsp_inv2rs_s, s1, d0) : fmaf(em1, ri, d0);
        }
    }
}

template <bool PME>
__device__ __forceinline__ void pp_pair(const tpm_box_dev& a, int ox, int oy, int oz, int4 s, float& gx, float& gy,
                                        float& gz, float& ph) {
    float dx, dy, dz, rinv3, rinv1;
    pp_law<PME>(a, ox, oy, oz, s, dx, dy, dz, rinv3, rinv1);
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

// Weight and derivative k of the same 6-point set (k a compile-time index
// when the caller's loop is unrolled).
__device__ __forceinline__ void interp6_weight(float x, int k, float* w, float* dw) {
    switch (k) {
        case 0:
            *w = x * ((1.f / 12.f) + x * (-(1.f / 24.f) + x * (-(3.f / 8.f) + x * ((13.f / 24.f) + x * -(5.f / 24.f)))));
            *dw = (1.f / 12.f) + x * (-(2.f / 24.f) + x * (-(9.f / 8.f) + x * ((52.f / 24.f) + x * -(25.f / 24.f))));
            break;
        case 1:
            *w = x * (-(2.f / 3.f) + x * ((2.f / 3.f) + x * ((13.f / 8.f) + x * (-(8.f / 3.f) + x * (25.f / 24.f)))));
            *dw = -(2.f / 3.f) + x * ((4.f / 3.f) + x * ((39.f / 8.f) + x * (-(32.f / 3.f) + x * (125.f / 24.f))));
            break;
        case 2:
            *w = 1.0f + x * x * (-(5.f / 4.f) + x * (-(35.f / 12.f) + x * ((21.f / 4.f) + x * -(25.f / 12.f))));
            *dw = x * (-(10.f / 4.f) + x * (-(105.f / 12.f) + x * ((84.f / 4.f) + x * -(125.f / 12.f))));
            break;
        case 3:
            *w = x * ((2.f / 3.f) + x * ((2.f / 3.f) + x * ((11.f / 4.f) + x * (-(31.f / 6.f) + x * (25.f / 12.f)))));
            *dw = (2.f / 3.f) + x * ((4.f / 3.f) + x * ((33.f / 4.f) + x * (-(124.f / 6.f) + x * (125.f / 12.f))));
            break;
        case 4:
            *w = x * (-(1.f / 12.f) + x * (-(1.f / 24.f) + x * (-(11.f / 8.f) + x * ((61.f / 24.f) + x * -(25.f / 24.f)))));
            *dw = -(1.f / 12.f) + x * (-(2.f / 24.f) + x * (-(33.f / 8.f) + x * ((244.f / 24.f) + x * -(125.f / 24.f))));
            break;
        default:
            *w = x * x * x * ((7.f / 24.f) + x * (-0.5f + x * (5.f / 24.f)));
            *dw = x * x * ((21.f / 24.f) + x * (-2.0f + x * (25.f / 24.f)));
            break;
    }
}

// Mesh force and potential of one sink (kick_treepm.cu:505-573) on the
// periodic mesh phi (M^3), unit-mass: g = -M sum dw w w phi, p = sum w w w phi.
// Separable: along z (w, dw), then y, then x, so only the y and z weights
// are live; the x weights are made where used.  This part adds the x planes
// aa0, aa0 + astep, ... (< 6) to the raw sums sx, sy, sz (gradient, before
// the factor -M) and sp, so several lanes can share one sink's stencil.
__device__ __forceinline__ void mesh_interp6_part(const float* __restrict__ phi, int M, uint32_t px, uint32_t py,
                                                  uint32_t pz, int aa0, int astep, float& sx, float& sy, float& sz,
                                                  float& sp) {
    const uint32_t pr[3] = {px, py, pz};
    int I0[3];
    float X[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const uint64_t P = (uint64_t)pr[d] * (uint64_t)M;
        I0[d] = (int)(P >> 32) - 2 + M;
        X[d] = (float)((uint32_t)P >> 8) * 0x1p-24f;
    }
    float wy[6], dwy[6], wz[6], dwz[6];
    interp6_weights(X[1], wy, dwy);
    interp6_weights(X[2], wz, dwz);
    int jz[6];
#pragma unroll
    for (int cc = 0; cc < 6; cc++) {
        int j = I0[2] + cc;
        j -= (j >= M) ? M : 0;
        j -= (j >= M) ? M : 0;
        jz[cc] = j;
    }
#pragma unroll 1
    for (int aa = aa0; aa < 6; aa += astep) {
        int jx = I0[0] + aa;
        jx -= (jx >= M) ? M : 0;
        jx -= (jx >= M) ? M : 0;
        float u = 0.f, v = 0.f, t = 0.f;  // sum_y w s, sum_y dw s, sum_y w tz
#pragma unroll 1
        for (int bb = 0; bb < 6; bb++) {
            int jy = I0[1] + bb;
            jy -= (jy >= M) ? M : 0;
            jy -= (jy >= M) ? M : 0;
            const float* row = phi + ((int64_t)jx * M + jy) * M;
            float s0 = 0.f, s1 = 0.f;  // sum_z w phi, sum_z dw phi
#pragma unroll
            for (int cc = 0; cc < 6; cc++) {
                const float p0 = row[jz[cc]];
                s0 = fmaf(wz[cc], p0, s0);
                s1 = fmaf(dwz[cc], p0, s1);
            }
            u = fmaf(wy[bb], s0, u);
            v = fmaf(dwy[bb], s0, v);
            t = fmaf(wy[bb], s1, t);
        }
        // x weight of this plane, made here: hoisted out of the loop (the
        // plane index is per lane) the six candidates would spill.
        float xa = X[0];
        asm volatile("" : "+v"(xa));
        float wx, dwx;
        interp6_weight(xa, aa, &wx, &dwx);
        sx = fmaf(dwx, u, sx);
        sy = fmaf(wx, v, sy);
        sz = fmaf(wx, t, sz);
        sp = fmaf(wx, u, sp);
    }
}

// The arithmetic of mesh_interp6_part (the same products and sums in the
// same order, so the same bits) with every stencil value of a lane's planes
// fetched before the first product: load(aa, bb, cc) gives phi at x plane aa,
// y row bb, z column cc of the sink's stencil.  NPL = planes per lane, at
// least ceil((6 - aa0) / astep).  X: the sink's in-cell offsets (as made by
// mesh_interp6_part).
// wy, dwy, wz, dwz: interp6_weights of X[1] and X[2].
template <int NPL, class Load>
__device__ __forceinline__ void interp6_planes(const float X[3], const float wy[6], const float dwy[6],
                                               const float wz[6], const float dwz[6], int aa0, int astep, Load load,
                                               float& sx, float& sy, float& sz, float& sp) {
    float p[NPL][6][6];
#pragma unroll
    for (int l = 0; l < NPL; l++) {
        const int aa = min(aa0 + l * astep, 5);
#pragma unroll
        for (int bb = 0; bb < 6; bb++)
#pragma unroll
            for (int cc = 0; cc < 6; cc++) p[l][bb][cc] = load(aa, bb, cc);
    }
#pragma unroll
    for (int l = 0; l < NPL; l++) {
        const int aa = aa0 + l * astep;
        if (aa >= 6) break;
        float u = 0.f, v = 0.f, t = 0.f;
#pragma unroll
        for (int bb = 0; bb < 6; bb++) {
            float s0 = 0.f, s1 = 0.f;
#pragma unroll
            for (int cc = 0; cc < 6; cc++) {
                s0 = fmaf(wz[cc], p[l][bb][cc], s0);
                s1 = fmaf(dwz[cc], p[l][bb][cc], s1);
            }
            u = fmaf(wy[bb], s0, u);
            v = fmaf(dwy[bb], s0, v);
            t = fmaf(wy[bb], s1, t);
        }
        float wx, dwx;
        interp6_weight(X[0], aa, &wx, &dwx);
        sx = fmaf(dwx, u, sx);
        sy = fmaf(wx, v, sy);
        sz = fmaf(wx, t, sz);
        sp = fmaf(wx, u, sp);
    }
}

// The stencil's first node (mesh index - 2) and in-cell offset per axis.
__device__ __forceinline__ void interp6_origin(uint32_t px, uint32_t py, uint32_t pz, int M, int I0[3], float X[3]) {
    const uint32_t pr[3] = {px, py, pz};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const uint64_t P = (uint64_t)pr[d] * (uint64_t)M;
        I0[d] = (int)(P >> 32) - 2;
        X[d] = (float)((uint32_t)P >> 8) * 0x1p-24f;
    }
}

// The whole stencil in one lane.
__device__ __forceinline__ void mesh_interp6(const float* __restrict__ phi, int M, uint32_t px, uint32_t py,
                                             uint32_t pz, float& pgx, float& pgy, float& pgz, float& pph) {
    float sx = 0.f, sy = 0.f, sz = 0.f, sp = 0.f;
    mesh_interp6_part(phi, M, px, py, pz, 0, 1, sx, sy, sz, sp);
    const float Mf = (float)M;
    pgx = -sx * Mf;
    pgy = -sy * Mf;
    pgz = -sz * Mf;
    pph = sp;
}

// Where the kick epilogue writes (particle arrays of a context or batch).
struct tpm_kick_out {
    float* __restrict__ gx;
    float* __restrict__ gy;
    float* __restrict__ gz;
    float* __restrict__ pot;
    float* __restrict__ vx;
    float* __restrict__ vy;
    float* __restrict__ vz;
    int8_t* __restrict__ rung;
    unsigned long long* __restrict__ counters;  // [1]: max rung
};

// Kick epilogue of one active sink i with rung r (kick_treepm.cu:716-754;
// rung_dt[r] = 2^-r): total unit-mass force = mesh (pg*) + short range
// (g*), scaled by GM; potential with phi0; velocity and rung update unless
// TPM_SKIP_KICK.
// Returns the new rung (-1 without the kick); the caller reduces it over
// the wave and updates the maximum rung once per wave.
__device__ __forceinline__ int kick_epilogue(const tpm_box_dev& B, unsigned flags, const tpm_kick_out& o, int i,
                                             int r, float pgx, float pgy, float pgz, float pph, float gx, float gy,
                                             float gz, float ph) {
    const float GM = B.GM;
    const float g0 = (pgx + gx) * GM;
    const float g1 = (pgy + gy) * GM;
    const float g2 = (pgz + gz) * GM;
    o.gx[i] = g0;
    o.gy[i] = g1;
    o.gz[i] = g2;
    o.pot[i] = B.do_phi ? (B.phi0 + pph + ph) * GM : 0.f;
    if (!(flags & TPM_SKIP_KICK)) {
        const float t0 = B.t0;
        float vx = o.vx[i], vy = o.vy[i], vz = o.vz[i];
        float dt = 0.5f * ldexpf(1.0f, -r) * t0;
        if (!B.first_call) {
            vx = fmaf(g0, dt, vx);
            vy = fmaf(g1, dt, vy);
            vz = fmaf(g2, dt, vz);
        }
        const float gg = fmaf(g0, g0, fmaf(g1, g1, g2 * g2));
        const float factor = B.eta * sqrtf(B.scale * B.hsoft);
        dt = fminf(factor * (1.0f / sqrtf(sqrtf(gg))), t0);
        float nr = fmaxf(ceilf(log2f(t0) - log2f(dt)), (float)(r - 1));
        nr = fminf(fmaxf(nr, 0.f), 31.f);
        r = (int)nr;
        dt = 0.5f * ldexpf(1.0f, -r) * t0;
        vx = fmaf(g0, dt, vx);
        vy = fmaf(g1, dt, vy);
        vz = fmaf(g2, dt, vz);
        o.vx[i] = vx;
        o.vy[i] = vy;
        o.vz[i] = vz;
        o.rung[i] = (int8_t)r;
        return r;
    }
    return -1;
}

// The maximum rung of a wave's kicked sinks into counters[1]: one atomic per
// wave (all lanes must call it; lanes without a kick pass -1).
__device__ __forceinline__ void max_rung_update(const tpm_kick_out& o, int r) {
    const int m = wave_max(r);
    if (lane_id() == 0 && m >= 0 && o.counters) atomicMax(&o.counters[1], (unsigned long long)m);
}

}  // namespace
