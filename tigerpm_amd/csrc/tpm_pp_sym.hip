// tpm_pp_sym.hip -- the short-range kick with each pair evaluated once.
//
// The reference evaluates every pair twice, once from each end (its sink
// loop over the 27 cells, kick_treepm.cu:358-458, pair law :242-283).  The
// two evaluations are exact negatives of each other: the int32 offset of j
// from i is minus that of i from j, its conversion to float is symmetric and
// r^2 is the same, so the force factor and the potential term are the same
// bits.  Here each pair is evaluated once and both ends take their share:
//
//  * cell c's wave pairs the sinks of c with the sources of its 13 forward
//    neighbour cells c + d (d lexicographically positive, z first) and with
//    the sources b > a of c itself, so every unordered pair of neighbouring
//    cells is met exactly once (a particle's own term once, at the sink);
//  * the sinks of c and their sums sit in LDS; the sources come 32 at a
//    time into registers (4 per lane, a group of 8 lanes holds all 32), and
//    the 8 groups walk the sinks within the cull distance of that source
//    tile's box, as the two-sided kernel walks culled sources;
//  * a sink's share goes to its LDS sum; a source's share is summed in
//    registers over the sinks, then over the groups, and stored once into
//    slot k(d) of the source (for a source in c itself: into its LDS sum).
//    Slot k(d) of a particle of cell c' is written by the wave of cell
//    c' - d alone (with zeros when that cell is empty), so no two waves
//    touch the same word, no sum is formed by atomics and nothing is read
//    back;
//  * the same wave interpolates the mesh force of its particles (the 6^3
//    stencil split over 4 lanes per particle), so that its memory latency
//    hides behind the pair work of other waves;
//  * pp_sym_epilogue_kernel adds, per particle and in a fixed order, its
//    sink sum and its 13 slots, then the mesh force and the kick, as the
//    two-sided kernel's epilogue does.
//
// Results are deterministic and equal the two-sided kernel's up to FP32
// summation order.  Every particle must be active (min_rung 0): a pair is met
// at one end only, so an inactive end could not be skipped.  Used for whole-
// box kicks with chain_dim >= 8 (see the far sources below).
#include <vector>

#include "tpm_pp_device.hpp"

#pragma clang fp contract(off)

#define SYM_SELF 13   // the 14th cell: c itself (its pairs need no slot)
#define SYM_SLOTS 13  // slots of source shares per particle
#define SYM_SMAX 256  // sinks of a cell per chunk (LDS: 10 KiB per wave)
#define SYM_WAVES 4   // waves per SIMD (the LDS allows 16 per CU)

namespace {

// The 13 forward neighbour offsets (x, y, z): z = +1, then z = 0 and
// y = +1, then (1, 0, 0).  Slot k of a particle holds what the cell at
// minus offset k contributed to it.
__device__ __forceinline__ void sym_offset(int k, int& dx, int& dy, int& dz) {
    if (k < 9) {
        dx = k % 3 - 1;
        dy = k / 3 - 1;
        dz = 1;
    } else if (k < 12) {
        dx = k - 10;
        dy = 1;
        dz = 0;
    } else {
        dx = 1;
        dy = 0;
        dz = 0;
    }
}

__device__ __forceinline__ int sym_wrap(int v, int C) {
    v += (v < 0) ? C : 0;
    v -= (v >= C) ? C : 0;
    return v;
}

struct SymArgs {
    const uint32_t* __restrict__ x;
    const uint32_t* __restrict__ y;
    const uint32_t* __restrict__ z;
    const int32_t* __restrict__ cell_begin;
    const float* __restrict__ phi;  // nullptr: no mesh term
    float4* gsink;  // [n] sink sums
    float4* gsrc;   // [13 n] source shares, slot k of particle j at k n + j
    float4* gmesh;  // [n] mesh force and potential (with phi)
    int64_t n;
    int ncell;
    unsigned long long* __restrict__ counters;
    tpm_box_dev box0;
};

// v as read from another lane by DPP control CTRL (quad permutations,
// row rotations); every source lane is in range for these controls
template <int CTRL>
__device__ __forceinline__ float dpp_quad(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ int dpp_int(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

// The 4-lane quad sum (every lane of the quad gets it; the same bits in
// each, as float addition commutes).
__device__ __forceinline__ float quad_sum(float v) {
    v += dpp_quad<0xB1>(v);        // lanes [1,0,3,2]
    return v + dpp_quad<0x4E>(v);  // lanes [2,3,0,1]
}

// Reduce-scatter of (x, y, z, w) over a 4-lane quad: lane q gets the quad
// sum of component ((q & 1) << 1) | (q >> 1), so lanes 0..3 hold the sums
// of x, z, y and w.
__device__ __forceinline__ float quad_reduce_scatter(float x, float y, float z, float w, bool odd, bool hi) {
    float k0 = odd ? z : x, k1 = odd ? w : y;
    const float s0 = odd ? x : z, s1 = odd ? y : w;
    k0 += dpp_quad<0xB1>(s0);
    k1 += dpp_quad<0xB1>(s1);
    const float m = hi ? k1 : k0;
    const float t = hi ? k0 : k1;
    return m + dpp_quad<0x4E>(t);
}

// The 8-lane form: lanes l and 7 - l of each half row hold the same
// component after quad_reduce_scatter when the upper quad numbers its lanes
// backwards (qq = 3 - (l & 3)); one more step with the mirrored lane sums
// the 8 lanes.
__device__ __forceinline__ float oct_reduce_scatter(float x, float y, float z, float w, bool odd, bool hi) {
    const float v = quad_reduce_scatter(x, y, z, w, odd, hi);
    return v + dpp_quad<0x141>(v);  // row_half_mirror
}

// One wave per chain cell c; the cells of one XCD form one contiguous run.
// The sinks of c sit in LDS (SMAX at a time, a chunk) with their sums; the
// sources of each of the 14 cells come 32 at a time (a source tile) into
// registers: lane q of an 8-lane group holds sources 4q .. 4q + 3, and the 8
// groups (streams) walk the sinks of the chunk that lie within the cull
// distance of the tile's box.  A sink's share is summed over the group and
// added to its LDS sum (one lane per component, no two lanes on one word,
// in program order); a source's share stays in registers until the tile
// ends, then is summed over the streams and written once to its slot.
// Pairs inside c are met once: sink a with source b only for a < b (a
// tile's own pairs in masked rounds), and the sources' shares go to the
// sinks' sums.  With more than SMAX particles in c the chunks run in turn:
// slots are then added to (this wave wrote them), and shares of sinks of a
// later chunk go through gsink, cleared first.  SMAX staged sinks per tile
// fit the list.
template <int SMAX, int WAVES>
__global__ __launch_bounds__(64, WAVES) void pp_sym_kernel(SymArgs a) {
    static_assert(SMAX % 32 == 0 && 2 * (SMAX + 8) <= 128 * 16, "chunk size");
    __shared__ uint4 spos[SMAX];    // sink positions of the chunk (raw x, y, z)
    __shared__ float4 sacc[SMAX];   // sink sums of the chunk
    __shared__ float4 sred[128];    // a tile's source sums per 16-lane row; before
                                    // them the staged sinks (chunk-local, +8 read ahead)
    uint16_t* const slist = reinterpret_cast<uint16_t*>(sred);
    const int ncell = a.ncell;
    const int per_xcd = (ncell + 7) >> 3;
    const int xcd = (int)(blockIdx.x & 7);
    const int kk = (int)(blockIdx.x >> 3);
    if (kk >= per_xcd) return;
    const int c = xcd * per_xcd + kk;
    if (c >= ncell) return;
    const tpm_box_dev& B = a.box0;
    const int C = B.C;
    const int cx = c % C, cy = (c / C) % C, cz = c / (C * C);
    const int lane = (int)lane_id();
    const int q = lane & 7;        // this lane's sources: 4q .. 4q + 3 of the tile
    const int stream = lane >> 3;  // sink stream 0 .. 7
    const int qq = (lane & 4) ? 3 - (lane & 3) : (lane & 3);
    const bool qodd = (qq & 1) != 0, qhi = (qq & 2) != 0;
    const int qcomp = ((qq & 1) << 1) | (qq >> 1);  // component of oct_reduce_scatter
    const bool qlow = (lane & 4) == 0;              // the lane of a mirrored pair that adds
    const int64_t n = a.n;
    float* const saccf = reinterpret_cast<float*>(sacc);
    // the cell of slot k: c + offset k, or c itself (wave-uniform)
    auto slot_cell = [&](int k) {
        if (k == SYM_SELF) return c;
        int dx, dy, dz;
        sym_offset(k, dx, dy, dz);
        return ((sym_wrap(cz + dz, C) * C) + sym_wrap(cy + dy, C)) * C + sym_wrap(cx + dx, C);
    };
    const int cb = a.cell_begin[c], ce = a.cell_begin[c + 1];
    if (cb == ce) {
        // No sinks, so no source tiles either: the slots this cell owns in
        // its 13 forward neighbours still get their (zero) shares.
        for (int k = 0; k < SYM_SELF; k++) {
            const int ck = slot_cell(k);
            const int b2 = a.cell_begin[ck], e2 = a.cell_begin[ck + 1];
            for (int j = b2 + lane; j < e2; j += 64) a.gsrc[k * n + j] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        return;
    }
    const bool chunked = ce - cb > SMAX;
    if (chunked) {
        for (int j = cb + lane; j < ce; j += 64) a.gsink[j] = make_float4(0.f, 0.f, 0.f, 0.f);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    const float cull2s = B.cull2s;
    unsigned long long executed = 0ull;
    for (int s0 = cb; s0 < ce; s0 += SMAX) {
        const int sn = min(SMAX, ce - s0);
        __syncthreads();
        for (int l = lane; l < sn; l += 64) {
            spos[l] = make_uint4(a.x[s0 + l], a.y[s0 + l], a.z[s0 + l], 0u);
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (chunked) {
                // what earlier chunks added (this wave's stores, read in L2)
                const float* p = reinterpret_cast<const float*>(a.gsink + s0 + l);
                v.x = __hip_atomic_load(p + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                v.y = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                v.z = __hip_atomic_load(p + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                v.w = __hip_atomic_load(p + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            sacc[l] = v;
        }
        __syncthreads();
        // The source tiles in order: the 13 forward cells, then c itself
        // from the chunk on (its sources before the chunk come before every
        // sink of the chunk).  Each tile's positions are loaded one tile
        // ahead, so their latency hides behind the current tile's pairs.
        auto next_tile = [&](int& ks, int& t0, int& je) {
            t0 += 32;
            while (t0 >= je) {
                if (++ks == 14) return false;
                const int c2 = slot_cell(ks);
                t0 = ks == SYM_SELF ? s0 : __builtin_amdgcn_readfirstlane(a.cell_begin[c2]);
                je = __builtin_amdgcn_readfirstlane(a.cell_begin[c2 + 1]);
            }
            return true;
        };
        // lane q's 4 sources of the tile at t0 (past its end: its last)
        auto load_tile = [&](int t0, int je, uint32_t* px, uint32_t* py, uint32_t* pz) {
            const uint32_t* bx = a.x + t0;
            const uint32_t* by = a.y + t0;
            const uint32_t* bz = a.z + t0;
            const int last = min(32, je - t0) - 1;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const unsigned j = (unsigned)min(4 * q + k, last);
                px[k] = bx[j];
                py[k] = by[j];
                pz[k] = bz[j];
            }
        };
        int nks = -1, nt0 = -32, nje = 0;
        bool have = next_tile(nks, nt0, nje);
        uint32_t nx[4], ny[4], nz[4];
        if (have) load_tile(nt0, nje, nx, ny, nz);
        while (have) {
            const int ks = nks, t0 = nt0, je = nje;
            uint32_t px[4], py[4], pz[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                px[k] = nx[k];
                py[k] = ny[k];
                pz[k] = nz[k];
            }
            have = next_tile(nks, nt0, nje);
            if (have) load_tile(nt0, nje, nx, ny, nz);
            const int tn = min(32, je - t0);
            // the tile's origin: its first source (lane 0, k = 0)
            const uint32_t rx = __builtin_amdgcn_readfirstlane(px[0]);
            const uint32_t ry = __builtin_amdgcn_readfirstlane(py[0]);
            const uint32_t rz = __builtin_amdgcn_readfirstlane(pz[0]);
            // the lane's 4 sources (past the tile's end: half the box away
            // on every axis; every sink lies within 2/C of the tile's
            // origin, so with C >= 8 such a source is at least 1/4 of the
            // box, beyond rcut, from each and adds exactly zero to both
            // ends).  The box takes the loaded positions, whose repeats of
            // the last source change nothing.
            int4 sv[4];
            int lox = INT_MAX, hix = INT_MIN, loy = INT_MAX, hiy = INT_MIN, loz = INT_MAX, hiz = INT_MIN;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t far = (4 * q + k < tn) ? 0u : 0x80000000u;
                sv[k] = make_int4((int)(px[k] + far), (int)(py[k] + far), (int)(pz[k] + far), 0);
                const int ox = (int)(px[k] - rx), oy = (int)(py[k] - ry), oz = (int)(pz[k] - rz);
                lox = min(lox, ox);
                hix = max(hix, ox);
                loy = min(loy, oy);
                hiy = max(hiy, oy);
                loz = min(loz, oz);
                hiz = max(hiz, oz);
            }
            // tile bounding box over lanes 0..7 (float(int32) rounds by <= 64
            // raw units: box distances carry <= 192 units of error, inside
            // cull2s's margin)
            auto grp_min = [](int v) {
                v = min(v, dpp_int<0xB1>(v));
                v = min(v, dpp_int<0x4E>(v));
                return min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 4));
            };
            auto grp_max = [](int v) {
                v = max(v, dpp_int<0xB1>(v));
                v = max(v, dpp_int<0x4E>(v));
                return max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 4));
            };
            const float flox = (float)grp_min(lox), fhix = (float)grp_max(hix);
            const float floy = (float)grp_min(loy), fhiy = (float)grp_max(hiy);
            const float floz = (float)grp_min(loz), fhiz = (float)grp_max(hiz);
            float ax[4], ay[4], az[4], ap[4];
#pragma unroll
            for (int k = 0; k < 4; k++) ax[k] = ay[k] = az[k] = ap[k] = 0.f;
            // The sinks: in c itself those before the tile, plus the tile's
            // own pairs when it lies in the chunk; elsewhere all.
            const bool diag = ks == SYM_SELF && t0 < s0 + sn;
            const int lim = diag ? t0 - s0 : sn;
            if (diag) {
                // stream s takes sinks t0 + s, s + 8, ...; sink i takes the
                // pair when source j >= i (j == i: its own potential term),
                // the source when j > i
#pragma unroll 1
                for (int si = stream; si < 32; si += 8) {
                    if (si < tn) {
                        const int li = t0 - s0 + si;
                        const uint4 p = spos[li];
                        float sx = 0.f, sy = 0.f, sz = 0.f, sp = 0.f;
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            float dx, dy, dz, r3, r1;
                            pp_law<false>(B, (int)p.x, (int)p.y, (int)p.z, sv[k], dx, dy, dz, r3, r1);
                            const int sk = 4 * q + k;
                            const float a3 = sk >= si ? r3 : 0.f, a1 = sk >= si ? r1 : 0.f;
                            const float b3 = sk > si ? r3 : 0.f, b1 = sk > si ? r1 : 0.f;
                            sx = fmaf(-dx, a3, sx);
                            sy = fmaf(-dy, a3, sy);
                            sz = fmaf(-dz, a3, sz);
                            sp -= a1;
                            ax[k] = fmaf(dx, b3, ax[k]);
                            ay[k] = fmaf(dy, b3, ay[k]);
                            az[k] = fmaf(dz, b3, az[k]);
                            ap[k] -= b1;
                        }
                        const float v = oct_reduce_scatter(sx, sy, sz, sp, qodd, qhi);
                        if (qlow) atomicAdd(saccf + 4 * li + qcomp, v);
                    }
                }
                executed += (unsigned long long)tn * (unsigned long long)tn;
            }
            // stage the sinks within the cull distance of the tile's box
            int nbuf = 0;
            for (int l0 = 0; l0 < lim; l0 += 64) {
                const int l = l0 + lane;
                bool keep = false;
                if (l < lim) {
                    const uint4 p = spos[l];
                    const float fx = (float)(int)(p.x - rx), fy = (float)(int)(p.y - ry), fz = (float)(int)(p.z - rz);
                    const float ex = fmaxf(fmaxf(flox - fx, fx - fhix), 0.f);
                    const float ey = fmaxf(fmaxf(floy - fy, fy - fhiy), 0.f);
                    const float ez = fmaxf(fmaxf(floz - fz, fz - fhiz), 0.f);
                    keep = fmaf(ex, ex, fmaf(ey, ey, ez * ez)) < cull2s;
                }
                const unsigned long long m = __ballot(keep);
                if (keep) slist[nbuf + (int)count_below(m)] = (uint16_t)l;
                nbuf += __popcll(m);
            }
            if (nbuf > 0) {
                __syncthreads();
                // stream s takes staged sinks s, s + 8, ...
                int e = stream;
                int li = min((int)slist[e], SMAX - 1);
                uint4 p = spos[li];
#pragma unroll 1
                for (int e0 = 0; e0 < nbuf; e0 += 8) {
                    // the next round's sink, read ahead (past nbuf: any
                    // index, clamped into the chunk, unused)
                    const int lin = min((int)slist[e + 8], SMAX - 1);
                    const uint4 pn = spos[lin];
                    if (e < nbuf) {
                        float sx = 0.f, sy = 0.f, sz = 0.f, sp = 0.f;
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            float dx, dy, dz, r3, r1;
                            pp_law<false>(B, (int)p.x, (int)p.y, (int)p.z, sv[k], dx, dy, dz, r3, r1);
                            sx = fmaf(-dx, r3, sx);
                            sy = fmaf(-dy, r3, sy);
                            sz = fmaf(-dz, r3, sz);
                            sp -= r1;
                            ax[k] = fmaf(dx, r3, ax[k]);
                            ay[k] = fmaf(dy, r3, ay[k]);
                            az[k] = fmaf(dz, r3, az[k]);
                            ap[k] -= r1;
                            // the sums are formed here, so this source's pair
                            // values die before the next one's are made
                            asm volatile("" : "+v"(ax[k]), "+v"(ay[k]), "+v"(az[k]), "+v"(ap[k]), "+v"(sx), "+v"(sy),
                                         "+v"(sz), "+v"(sp));
                        }
                        const float v = oct_reduce_scatter(sx, sy, sz, sp, qodd, qhi);
                        if (qlow) atomicAdd(saccf + 4 * li + qcomp, v);
                    }
                    li = lin;
                    p = pn;
                    e += 8;
                }
                executed += (unsigned long long)nbuf * (unsigned long long)tn;
            }
            __syncthreads();
            // The sources' shares, summed over the 8 streams: the 2 streams
            // of a 16-lane row by a row rotation, then the 4 rows through
            // sred, where lane l adds value l and l + 64 (source v / 4,
            // component v % 4).
#pragma unroll
            for (int k = 0; k < 4; k++) {
                ax[k] += dpp_quad<0x128>(ax[k]);  // row_ror:8
                ay[k] += dpp_quad<0x128>(ay[k]);
                az[k] += dpp_quad<0x128>(az[k]);
                ap[k] += dpp_quad<0x128>(ap[k]);
            }
            if ((lane & 15) < 8) {
#pragma unroll
                for (int k = 0; k < 4; k++) sred[32 * (lane >> 4) + 4 * q + k] = make_float4(ax[k], ay[k], az[k], ap[k]);
            }
            __syncthreads();
            const float* part = reinterpret_cast<const float*>(sred);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int vi = lane + 64 * h;
                const float v = ((part[vi] + part[128 + vi]) + part[256 + vi]) + part[384 + vi];
                const int s = vi >> 2, u = vi & 3;
                const int j = t0 + s;
                if (s < tn) {
                    if (ks != SYM_SELF) {
                        float* o = reinterpret_cast<float*>(a.gsrc + ks * n + j) + u;
                        *o = s0 == cb ? v : __hip_atomic_load(o, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + v;
                    } else if (j < s0 + sn) {
                        atomicAdd(saccf + 4 * (j - s0) + u, v);
                    } else {
                        float* o = reinterpret_cast<float*>(a.gsink + j) + u;
                        *o = __hip_atomic_load(o, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + v;
                    }
                }
            }
            __syncthreads();
        }
        // the chunk's sink sums out
        for (int l = lane; l < sn; l += 64) a.gsink[s0 + l] = sacc[l];
        if (chunked) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    // mesh force and potential of c's particles, 16 at a time: lane =
    // particle + 16 part, the four parts split the six x planes of the stencil
    if (a.phi) {
        for (int t0 = cb; t0 < ce; t0 += 16) {
            const int tn = min(16, ce - t0);
            const int sk = lane & 15, part = lane >> 4;
            const int i = t0 + min(sk, tn - 1);
            float sx = 0.f, sy = 0.f, sz = 0.f, sp = 0.f;
            mesh_interp6_part(a.phi + B.mesh0, B.M, a.x[i], a.y[i], a.z[i], part, 4, sx, sy, sz, sp);
#pragma unroll
            for (int o = 16; o < 64; o <<= 1) {
                sx += __shfl_xor(sx, o);
                sy += __shfl_xor(sy, o);
                sz += __shfl_xor(sz, o);
                sp += __shfl_xor(sp, o);
            }
            const float Mf = (float)B.M;
            if (part == 0 && sk < tn) a.gmesh[t0 + sk] = make_float4(-sx * Mf, -sy * Mf, -sz * Mf, sp);
        }
    }
    if (lane == 0 && a.counters)
        atomicAdd(&a.counters[TPM_COUNTER_STRIDE * (blockIdx.x % TPM_COUNTER_SLOTS)], executed);
}

struct SymEpiArgs {
    const float4* __restrict__ gsink;
    const float4* __restrict__ gsrc;
    const float4* __restrict__ gmesh;  // nullptr: no mesh term
    int64_t n;
    tpm_kick_out o;
    unsigned flags;
    tpm_box_dev box0;
};

// Per particle: the sink sum plus the 13 slots in a fixed order, the mesh
// force, and the kick epilogue.
__global__ __launch_bounds__(256) void pp_sym_epilogue_kernel(SymEpiArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < a.n;
    const int64_t ii = valid ? i : 0;
    const tpm_box_dev& B = a.box0;
    float4 v = a.gsink[ii];
#pragma unroll
    for (int s = 0; s < SYM_SLOTS; s++) {
        const float4 w = a.gsrc[s * a.n + ii];
        v.x += w.x;
        v.y += w.y;
        v.z += w.z;
        v.w += w.w;
    }
    // back to box units: g in 2^-64, phi in 2^-32
    const float gx = v.x * 0x1p64f, gy = v.y * 0x1p64f, gz = v.z * 0x1p64f, ph = v.w * 0x1p32f;
    const float4 pm = a.gmesh ? a.gmesh[ii] : make_float4(0.f, 0.f, 0.f, 0.f);
    // the maximum rung goes to this workgroup's counter slot: millions of
    // waves on one address would queue in one L2 channel
    tpm_kick_out o = a.o;
    if (o.counters) o.counters += TPM_COUNTER_STRIDE * (blockIdx.x % TPM_COUNTER_SLOTS);
    int nr = -1;
    if (valid) nr = kick_epilogue(B, a.flags, o, (int)i, o.rung[i], pm.x, pm.y, pm.z, pm.w, gx, gy, gz, ph);
    max_rung_update(o, nr);
}

}  // namespace

bool tpm_pp_sym_eligible(const tpm_ctx* ctx, int min_rung, unsigned flags, const tpm_cell_box* box) {
    return ctx->pp_newton && !box && min_rung == 0 && !(flags & (TPM_KICK_PME | TPM_SKIP_PP)) && ctx->C >= 8 &&
           ctx->n > 0 && ctx->n <= INT_MAX;
}

// The one-sided-pair kick of the whole box (every particle active).
tpm_status tpm_pp_kick_sym(tpm_ctx* ctx, double scale, double t0, int first_call, unsigned flags, tpm_stats* stats) {
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->n;
    if (n > ctx->sym_cap) {
        void** bufs[3] = {&ctx->gsink, &ctx->gsrc, &ctx->gmesh};
        for (void** b : bufs) {
            if (*b) TPM_HIP(ctx, hipFree(*b));
            *b = nullptr;
        }
        ctx->sym_cap = 0;
        TPM_HIP(ctx, hipMalloc(&ctx->gsink, sizeof(float4) * n));
        TPM_HIP(ctx, hipMalloc(&ctx->gsrc, sizeof(float4) * SYM_SLOTS * n));
        TPM_HIP(ctx, hipMalloc(&ctx->gmesh, sizeof(float4) * n));
        ctx->sym_cap = n;
    }
    if (!ctx->counters) TPM_HIP(ctx, hipMalloc(&ctx->counters, sizeof(unsigned long long) * TPM_COUNTER_WORDS));
    TPM_TRY(tpm_fill_u64(ctx, ctx->counters, TPM_COUNTER_WORDS));
    const bool use_pm = !(flags & TPM_SKIP_PM);
    tpm_box_dev box;
    tpm_box_fill(&box, ctx->p, 0, first_call, t0, scale);
    box.cell0 = 0;
    box.mesh0 = 0;
    SymArgs a{};
    a.x = ctx->x;
    a.y = ctx->y;
    a.z = ctx->z;
    a.cell_begin = ctx->cell_begin;
    a.phi = use_pm ? ctx->phi : nullptr;
    a.gsink = reinterpret_cast<float4*>(ctx->gsink);
    a.gsrc = reinterpret_cast<float4*>(ctx->gsrc);
    a.gmesh = reinterpret_cast<float4*>(ctx->gmesh);
    a.n = n;
    a.ncell = ctx->C * ctx->C * ctx->C;
    a.counters = ctx->counters;
    a.box0 = box;
    SymEpiArgs b{};
    b.gsink = a.gsink;
    b.gsrc = a.gsrc;
    b.gmesh = use_pm ? a.gmesh : nullptr;
    b.n = n;
    b.o = tpm_kick_out{ctx->gx, ctx->gy, ctx->gz, ctx->pot, ctx->vx, ctx->vy, ctx->vz, ctx->rung, ctx->counters};
    b.flags = flags & (TPM_SKIP_PP | TPM_SKIP_PM | TPM_SKIP_KICK);
    b.box0 = box;
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
    const int64_t per_xcd = ((int64_t)a.ncell + 7) / 8;
    TPM_HIP(ctx, hipEventRecord(e0, s));
    pp_sym_kernel<SYM_SMAX, SYM_WAVES><<<(unsigned)(8 * per_xcd), 64, 0, s>>>(a);
    TPM_HIP(ctx, hipGetLastError());
    pp_sym_epilogue_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(b);
    TPM_HIP(ctx, hipGetLastError());
    TPM_HIP(ctx, hipEventRecord(e1, s));
    if (stats) {
        TPM_TRY(tpm_candidate_pairs(ctx, ctx->counters + 2, nullptr));
        std::vector<unsigned long long> hw(TPM_COUNTER_WORDS);
        TPM_HIP(ctx, hipMemcpyAsync(hw.data(), ctx->counters, sizeof(unsigned long long) * TPM_COUNTER_WORDS,
                                    hipMemcpyDeviceToHost, s));
        int32_t nt = 0;
        TPM_HIP(ctx, hipMemcpyAsync(&nt, ctx->ntiles, sizeof(nt), hipMemcpyDeviceToHost, s));
        TPM_HIP(ctx, hipStreamSynchronize(s));
        float ms = 0.f;
        TPM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
        unsigned long long hc[4];
        tpm_reduce_counters(hw.data(), hc);
        stats->executed_pairs = (double)hc[0];
        stats->max_rung = (int)hc[1];
        stats->candidate_pairs = (double)hc[2];
        stats->ms_pp = ms;
        stats->n_tiles = nt;
    }
    return TPM_OK;
}
