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
// is the reference's up to exact power-of-two factors, except that inside
// the softening length the two (erfc - 1) corrections are evaluated by their
// series in x^2 (pp_law), where the formula as written cancels to nothing in
// FP32.  The self pair is
// evaluated like any other source, as in the reference: inside the
// softening branch it contributes -35/16 h^-1 to phi and nothing to g.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "tpm_pp_device.hpp"

#include <climits>

#define SBUF 256  // staged sources per flush (4 KiB LDS)
// Per-sink source lists (LISTS, 16-sink tiles): a flush of <= LCAP staged
// sources (staging flushes above LFLUSH), the far dummy at LFAR.
#define LCAP 144
#define LFLUSH (LCAP - 64)
#define LSBUF 160
#define LFAR (LSBUF - 1)

// Every fused multiply-add in this file is written out (fmaf) and the
// compiler may fuse nothing else, so the results come out the same bits in
// every instantiation of the kernel (single box, batch, persistent).
#pragma clang fp contract(off)

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
    const float4* __restrict__ mesh;  // mesh force and potential per sink, made before the launch
                                      // by mesh_force_kernel (nullptr: no mesh term)
    unsigned long long* __restrict__ counters;
    int* __restrict__ queue;        // 8 per-XCD tile queues (persistent launch)
    const tpm_box_dev* __restrict__ boxes;  // batched: per-box table, tile.w = box
    tpm_box_dev box0;                        // single box
    unsigned flags;
};

// The fused kernel.  Work item = one tile: TS sinks of one chain cell
// against the culled sources of the 27 neighbour cells, then mesh
// interpolation and the kick epilogue.
//   PERSIST = false: one workgroup per tile.  The hardware deals workgroups
//     round-robin over the 8 XCDs, so each XCD gets one contiguous
//     (spatially coherent) run of tiles and its L2 keeps the shared
//     neighbour cells.
//   PERSIST = true: resident workgroups take tiles from 8 queues, one
//     contiguous run per XCD; a workgroup drains its own XCD's queue first
//     and then helps the others.  Every queue ends, so every workgroup exits.
// BATCH = true: tile.w names the box whose constants apply (device table).
// PME = true: kick_pme's pair law (the caller turns culling and phi0 off).
// The box record of a tile.  A batch's table is read word by word through
// the constant address space, so the record arrives by scalar loads in SGPRs
// (the table is not written during the launch); read as ordinary global
// memory, its fields would take vector registers.
template <bool BATCH>
__device__ __forceinline__ tpm_box_dev tile_box(const PPArgs& a, int w) {
    if constexpr (BATCH) {
        static_assert(sizeof(tpm_box_dev) % 4 == 0, "box record of whole words");
        constexpr int NW = (int)(sizeof(tpm_box_dev) / 4);
        using cword = const __attribute__((address_space(4))) int;
        cword* src = (cword*)(a.boxes + __builtin_amdgcn_readfirstlane(w));
        int words[NW];
#pragma unroll
        for (int k = 0; k < NW; k++) words[k] = src[k];
        tpm_box_dev b;
        __builtin_memcpy(&b, words, sizeof b);
        return b;
    } else {
        return a.box0;
    }
}

// f(integral_constant<int, k>) for k = 0 .. N-1, k a constant expression
// in each call.
template <int K, int N>
struct static_for_t {
    template <class F>
    static __device__ __forceinline__ void run(F& f) {
        f(std::integral_constant<int, K>{});
        static_for_t<K + 1, N>::run(f);
    }
};
template <int N>
struct static_for_t<N, N> {
    template <class F>
    static __device__ __forceinline__ void run(F&) {}
};
template <int N, class F>
__device__ __forceinline__ void static_for(F& f) {
    static_for_t<0, N>::run(f);
}

// WPG = tiles per workgroup, one wave each (1 or 4).  The waves of a
// workgroup run on one CU and take consecutive tiles of their XCD's run,
// mostly of one chain cell, so their source loads meet in that CU's L1;
// they share nothing else and never wait for each other.
// LISTS = true (16-sink tiles, culling on): each flush of staged sources is
// sorted into one list per sink of the sources within the cull distance of
// that sink (not of the tile's box), and each sink's 4 source streams walk
// its own list; a flush takes as many steps as its longest list.
template <int TS, bool BATCH, bool PERSIST, bool PME = false, int WPG = 1, bool LISTS = false>
// Occupancy: 8 waves per SIMD (64 VGPRs, no spills) for one box and for a
// batch (whose box record tile_box keeps in SGPRs); 7 with per-sink lists
// (72 VGPRs: at 64 the list build spills).
__global__ __launch_bounds__(64 * WPG, PERSIST ? 1 : (LISTS ? 7 : 8)) void pp_kick_kernel(PPArgs a) {
    static_assert(WPG == 1 || (WPG == 4 && !PERSIST), "several tiles per workgroup: one-shot launch only");
    static_assert(!LISTS || (TS == 16 && !PME), "per-sink lists: 16-sink tiles, kick_treepm's law");
    constexpr int SPLIT = 64 / TS;  // source streams per wave
    constexpr int NSB = LISTS ? LSBUF : SBUF;
    constexpr int FLUSH_AT = LISTS ? LFLUSH : SBUF - 2 * 64;
    __shared__ int4 sbufs[WPG][NSB];
    __shared__ uint8_t lsts[WPG][LISTS ? 16 * LCAP + 16 : 4];
    __shared__ float4 sinkfs[WPG][LISTS ? 16 : 1];
    const int wslot = WPG == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    int4* const sbuf = sbufs[wslot];
    uint8_t* const lst = lsts[wslot];
    float4* const sinkf = sinkfs[wslot];
    // the staged list is written and read by the lanes of one wave
    auto lds_sync = [&]() {
        if constexpr (WPG == 1) {
            __syncthreads();
        } else {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
        }
    };
    const int nt = *a.ntiles;
    const int per_xcd = (nt + 7) >> 3;
    const int xcd = (int)(blockIdx.x & 7);
    int q = xcd, qi = 0;  // persistent: current queue and queues visited
    int t;
    if constexpr (PERSIST) {
        t = INT_MAX;
    } else {
        const int k = (int)(blockIdx.x >> 3) * WPG + (WPG == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6));
        if (k >= per_xcd) return;
        t = xcd * per_xcd + k;
        if (t >= nt) return;
    }
    for (;;) {
        if constexpr (PERSIST) {
            // next tile: own queue first, then the others
            for (;;) {
                int k = 0;
                if (lane_id() == 0) k = atomicAdd(&a.queue[q], 1);
                k = __builtin_amdgcn_readfirstlane(__shfl(k, 0));
                t = q * per_xcd + k;
                if (t < min((q + 1) * per_xcd, nt)) break;
                if (++qi == 8) return;
                q = (xcd + qi) & 7;
            }
        }
        const int4 tile = a.tiles[t];
        const tpm_box_dev B = tile_box<BATCH>(a, tile.w);
        unsigned long long* const cnt =
            a.counters ? a.counters + TPM_COUNTER_STRIDE * (blockIdx.x % TPM_COUNTER_SLOTS) : nullptr;
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
        const bool active = valid && r >= B.min_rung;
        const unsigned long long amask = __ballot(active);
        if (amask == 0ull) {
            if constexpr (PERSIST)
                continue;
            else
                return;
        }
        {
            const int ox = (int)(px - rx), oy = (int)(py - ry), oz = (int)(pz - rz);
            float gx = 0.f, gy = 0.f, gz = 0.f, ph = 0.f;
            if (!(a.flags & TPM_SKIP_PP)) {
                // Tile bounding box as floats.  float(int32) rounds by <= 64
                // raw units, so box distances below carry <= 192 units of
                // error, inside the 512-unit margin of cull2s: the test never
                // drops a source with r < rcut to any sink.
                const float lox = (float)wave_min(active ? ox : INT_MAX), hix = (float)wave_max(active ? ox : INT_MIN);
                const float loy = (float)wave_min(active ? oy : INT_MAX), hiy = (float)wave_max(active ? oy : INT_MIN);
                const float loz = (float)wave_min(active ? oz : INT_MAX), hiz = (float)wave_max(active ? oz : INT_MIN);
                const int C = B.C;
                const int cid = tile.x - B.cell0;
                const int cx = cid % C, cy = (cid / C) % C, cz = cid / (C * C);
                const bool cull = B.cull != 0;
                const float cull2s = B.cull2s;
                int nbuf = 0;
                unsigned long long executed = 0ull;
                const int nact = __popcll(amask & ((TS == 64) ? ~0ull : ((1ull << TS) - 1)));
                // Far dummy of the per-sink lists: offset -2^31 on every
                // axis, i.e. at least 1/2 - 1/C of the box from every sink of
                // the tile (a sink lies within one cell of the tile's first
                // particle), so r > rcut whenever culling is on (C >= 5).
                const int4 far = make_int4(INT_MIN, INT_MIN, INT_MIN, -1);
                // Per-sink lists: the tile's sinks as floats in LDS, read by
                // every lane alike (an inactive sink sits far from every
                // source: empty list).
                if constexpr (LISTS) {
                    if (lane < 16)
                        sinkf[lane] = make_float4(active ? (float)ox : 3.0e9f, active ? (float)oy : 3.0e9f,
                                                  active ? (float)oz : 3.0e9f, 0.f);
                    if (lane == 0) sbuf[LFAR] = far;
                }
                // Sources in range of no sink cost nothing but the test: a
                // source joins sink k's list when its float distance to the
                // sink is inside the cull distance (float(int32) offsets are
                // off by <= 8 raw units here, far inside the 512-unit margin),
                // so no pair with r < rcut is dropped and the pair law's own
                // cutoff decides as before.
                auto flush_lists = [&]() {
                    lds_sync();
                    int cntv = 0;  // lane k < 16: entries of sink k's list
                    for (int r0 = 0; r0 < nbuf; r0 += 64) {
                        const int e = r0 + lane;
                        // lanes past the staged sources test the far dummy,
                        // which is in range of no sink
                        const int4 sv = sbuf[e < nbuf ? e : LFAR];
                        const float fx = (float)sv.x, fy = (float)sv.y, fz = (float)sv.z;
                        // sink k's position is read from LDS in its own step
                        // (48 positions kept live would cost occupancy); its
                        // count lives in lane k of cntv
                        auto sink_test = [&](auto kc) {
                            constexpr int k = decltype(kc)::value;
                            const float4 sk = sinkf[k];
                            const float dx = fx - sk.x, dy = fy - sk.y, dz = fz - sk.z;
                            const bool in = fmaf(dx, dx, fmaf(dy, dy, dz * dz)) < cull2s;
                            const unsigned long long m = __ballot(in);
                            const int base = __builtin_amdgcn_readlane(cntv, k);
                            if (in) lst[k * LCAP + base + (int)count_below(m)] = (uint8_t)e;
                            const int nb = base + __popcll(m);
                            asm volatile("v_writelane_b32 %0, %1, %2" : "+v"(cntv) : "s"(nb), "i"(k));
                        };
                        static_for<16>(sink_test);
                    }
                    lds_sync();
                    const int mycnt = __shfl(cntv, slot);
                    const int mx = wave_max(cntv);
                    int tot = cntv;
#pragma unroll
                    for (int o = 1; o < 16; o <<= 1) tot += __shfl_xor(tot, o);
                    tot = __builtin_amdgcn_readfirstlane(tot);
                    // stream s of sink k takes entries s, s + 4, ... of list k;
                    // the index read one step ahead stays inside the lists'
                    // slack (4 * ceil(LCAP / 4) + 3 < LCAP + 16)
                    const uint8_t* my = lst + slot * LCAP + stream;
                    const int steps = (mx + 3) >> 2;
                    int pidx = stream;
                    int4 sc = sbuf[pidx < mycnt ? (int)my[0] : LFAR];
#pragma unroll 2
                    for (int i = 0; i < steps; i++) {
                        const int pn = pidx + 4;
                        const int4 sn = sbuf[pn < mycnt ? (int)my[4 * i + 4] : LFAR];
                        pp_pair<PME>(B, ox, oy, oz, sc, gx, gy, gz, ph);
                        sc = sn;
                        pidx = pn;
                    }
                    executed += (unsigned long long)tot;
                    nbuf = 0;
                    lds_sync();
                };
                auto flush = [&]() {
                    if constexpr (LISTS) {
                        if (cull) {
                            flush_lists();
                            return;
                        }
                    }
                    // whole rounds of SPLIT sources, then the streams that
                    // still have one (no padding source: with C < 8 a
                    // stand-in placed anywhere can come within rcut of some
                    // sink, and kick_pme's law has no cutoff at all)
                    const int full = nbuf / SPLIT * SPLIT;
                    lds_sync();
                    int4 s = sbuf[stream];
                    // the read one round ahead stays below full + 2 SPLIT <=
                    // FLUSH_AT + 64 + SPLIT < NSB (and below LFAR with lists);
                    // what it reads past nbuf is not used
#pragma unroll 2
                    for (int qq = 0; qq < full; qq += SPLIT) {
                        const int4 sn = sbuf[qq + SPLIT + stream];
                        pp_pair<PME>(B, ox, oy, oz, s, gx, gy, gz, ph);
                        s = sn;
                    }
                    if (full + stream < nbuf) pp_pair<PME>(B, ox, oy, oz, s, gx, gy, gz, ph);
                    executed += (unsigned long long)nbuf * (unsigned long long)nact;
                    nbuf = 0;
                    lds_sync();
                };
                for (int nb = 0; nb < 27; nb++) {
                    int jx = cx + (nb % 3) - 1, jy = cy + (nb / 3) % 3 - 1, jz = cz + nb / 9 - 1;
                    jx += (jx < 0) ? C : 0;
                    jx -= (jx >= C) ? C : 0;
                    jy += (jy < 0) ? C : 0;
                    jy -= (jy >= C) ? C : 0;
                    jz += (jz < 0) ? C : 0;
                    jz -= (jz >= C) ? C : 0;
                    const int c2 = B.cell0 + (jz * C + jy) * C + jx;
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
                            keep = !cull || fmaf(ex, ex, fmaf(ey, ey, ez * ez)) < cull2s;
                        }
                        const unsigned long long m = __ballot(keep);
                        if (keep) sbuf[nbuf + (int)count_below(m)] = sv;
                        nbuf += __popcll(m);
                        if (nbuf > FLUSH_AT) flush();
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
                if (lane == 0 && cnt) atomicAdd(&cnt[0], executed);
                // back to box units: g in 2^-64, phi in 2^-32
                gx *= 0x1p64f;
                gy *= 0x1p64f;
                gz *= 0x1p64f;
                ph *= 0x1p32f;
            }
            {
                // Then the kick, by each active sink's stream-0 lane, with the
                // mesh force and potential made before the launch.
                const tpm_kick_out o{a.gx, a.gy, a.gz, a.pot, a.vx, a.vy, a.vz, a.rung, cnt};
                int nr = -1;
                if (active && stream == 0) {
                    float pgx = 0.f, pgy = 0.f, pgz = 0.f, pph = 0.f;
                    if (a.mesh) {
                        const float4 m = a.mesh[i];
                        pgx = m.x;
                        pgy = m.y;
                        pgz = m.z;
                        pph = m.w;
                    }
                    nr = kick_epilogue(B, a.flags, o, i, r, pgx, pgy, pgz, pph, gx, gy, gz, ph);
                }
                max_rung_update(o, nr);
            }
        }
        if constexpr (!PERSIST) return;
    }
}

// The mesh force and potential of every active sink of one box, as the PP
// kernel's epilogue makes them, one workgroup per chain cell: the potential
// around the cell (W^3 nodes, periodic) is read once into LDS, then each
// sink's 6 x planes go in the epilogue's SPLIT = 64 / TS parts with the
// arithmetic of mesh_interp6_part on the LDS copy, the parts summed in the
// order of the epilogue's xor shuffles -- the same bits as the epilogue.  Inside
// the PP kernel the stencil's 216 scattered loads per sink took 11 ms of a
// 204 ms launch at 512^3; here each potential value is fetched once per
// neighbouring cell.  Cell k spans raw [bounds[k], bounds[k + 1]).
// One chain cell's sinks [cb, ce); r0: the cell's first stencil node per
// axis (mesh index of its lower split plane - 2).  The W^3 nodes from r0 on
// are read into LDS first; GLOBAL: the stencil is read from phi directly (a
// cell's neighbourhood larger than the LDS cube).
template <int TS, bool GLOBAL>
__device__ __forceinline__ void mesh_force_cell(float* sphi, const uint32_t* __restrict__ x,
                                                const uint32_t* __restrict__ y, const uint32_t* __restrict__ z,
                                                const int8_t* __restrict__ rung, int min_rung, int cb, int ce,
                                                const int r0[3], const float* __restrict__ phi, int M, int W,
                                                float4* __restrict__ out) {
    constexpr int SPLIT = 64 / TS;
    const int W2 = W * W;
    if (!GLOBAL && W <= M) {
        // node l = (a W + b) W + c is r0 + (a, b, c), periodic (r0 >= -2, so
        // every coordinate lies in [-2, 2M)); a thread's next node is
        // blockDim.x further on, its (a, b, c) advanced by carries instead of
        // divisions
        const int d = (int)blockDim.x, dc = d % W, db = (d / W) % W, da = d / W2;
        int l = (int)threadIdx.x, a = l / W2, b = (l / W) % W, c = l % W;
        for (; l < W2 * W; l += d) {
            int gx = r0[0] + a, gy = r0[1] + b, gz = r0[2] + c;
            gx = gx < 0 ? gx + M : (gx >= M ? gx - M : gx);
            gy = gy < 0 ? gy + M : (gy >= M ? gy - M : gy);
            gz = gz < 0 ? gz + M : (gz >= M ? gz - M : gz);
            sphi[l] = phi[((int64_t)gx * M + gy) * M + gz];
            c += dc;
            if (c >= W) {
                c -= W;
                b++;
            }
            b += db;
            if (b >= W) {
                b -= W;
                a++;
            }
            a += da;
        }
    } else {
        // W may exceed M on tiny meshes
        for (int l = (int)threadIdx.x; l < (GLOBAL ? 0 : W2 * W); l += (int)blockDim.x) {
            const int gx = (r0[0] + l / W2 + M) % M, gy = (r0[1] + (l / W) % W + M) % M,
                      gz = (r0[2] + l % W + M) % M;
            sphi[l] = phi[((int64_t)gx * M + gy) * M + gz];
        }
    }
    __syncthreads();
    const float Mf = (float)M;
    // one lane per sink: its SPLIT parts in turn, summed as the xor shuffles
    // of a wave that splits each sink over 64 / TS lanes would sum them
    // (pairs of neighbouring parts, then pairs of pairs, ...)
    for (int i = cb + (int)threadIdx.x; i < ce; i += (int)blockDim.x) {
        if (rung[i] < min_rung) continue;
        int I0[3];
        float X[3];
        interp6_origin(x[i], y[i], z[i], M, I0, X);
        const float* base = sphi + ((I0[0] - r0[0]) * W + (I0[1] - r0[1])) * W + (I0[2] - r0[2]);
        int jx[6], jy[6], jz[6];
        if (GLOBAL) {
#pragma unroll
            for (int k = 0; k < 6; k++) {
                jx[k] = (I0[0] + k + M) % M;
                jy[k] = (I0[1] + k + M) % M;
                jz[k] = (I0[2] + k + M) % M;
            }
        }
        const auto load = [&](int aa, int bb, int cc) {
            if constexpr (GLOBAL)
                return phi[((int64_t)jx[aa] * M + jy[bb]) * M + jz[cc]];
            else
                return base[(aa * W + bb) * W + cc];
        };
        float wy[6], dwy[6], wz[6], dwz[6];
        interp6_weights(X[1], wy, dwy);
        interp6_weights(X[2], wz, dwz);
        float4 part[SPLIT];
#pragma unroll
        for (int k = 0; k < SPLIT; k++) {
            float sx = 0.f, sy = 0.f, sz = 0.f, sp = 0.f;
            interp6_planes<(6 + SPLIT - 1) / SPLIT>(X, wy, dwy, wz, dwz, k, SPLIT, load, sx, sy, sz, sp);
            part[k] = make_float4(sx, sy, sz, sp);
        }
#pragma unroll
        for (int h = 1; h < SPLIT; h <<= 1)
#pragma unroll
            for (int k = 0; k < SPLIT; k += 2 * h) {
                part[k].x += part[k + h].x;
                part[k].y += part[k + h].y;
                part[k].z += part[k + h].z;
                part[k].w += part[k + h].w;
            }
        out[i] = make_float4(-part[0].x * Mf, -part[0].y * Mf, -part[0].z * Mf, part[0].w);
    }
}

// One box: cell c = (cz C + cy) C + cx spans raw [bounds[k], bounds[k + 1]);
// only the cells in [lo, hi) per axis.
template <int TS, bool GLOBAL>
__global__ __launch_bounds__(256) void mesh_force_kernel(const uint32_t* __restrict__ x,
                                                        const uint32_t* __restrict__ y,
                                                        const uint32_t* __restrict__ z,
                                                        const int8_t* __restrict__ rung, int min_rung,
                                                        const int32_t* __restrict__ cell_begin,
                                                        const uint32_t* __restrict__ bounds, int C,
                                                        const float* __restrict__ phi, int M, int W,
                                                        int3 lo, int3 hi, float4* __restrict__ out) {
    extern __shared__ float sphi[];  // W^3, z fastest
    // Each of the 8 XCDs (workgroup index mod 8) takes one contiguous run of
    // cells in an order with z fastest, the potential's fastest axis, so
    // that the workgroups running together on one XCD read neighbouring
    // cubes through its L2 (in cell-id order, x fastest, the cubes of
    // consecutive workgroups lie a megabyte apart at 512^3 and every cube
    // came from HBM: 20.7 GB per launch).
    const int nc = C * C * C, per = (nc + 7) >> 3;
    const int rk = (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
    if (rk >= nc) return;
    const int k3[3] = {rk / (C * C), (rk / C) % C, rk % C};
    const int c = (k3[2] * C + k3[1]) * C + k3[0];
    if (k3[0] < lo.x || k3[0] >= hi.x || k3[1] < lo.y || k3[1] >= hi.y || k3[2] < lo.z || k3[2] >= hi.z) return;
    const int cb = cell_begin[c], ce = cell_begin[c + 1];
    if (cb == ce) return;
    int r0[3];
#pragma unroll
    for (int d = 0; d < 3; d++) r0[d] = (int)(((uint64_t)bounds[k3[d]] * (uint64_t)M) >> 32) - 2;
    mesh_force_cell<TS, GLOBAL>(sphi, x, y, z, rung, min_rung, cb, ce, r0, phi, M, W, out);
}

// Batched: global cell c belongs to the last box whose cell0 <= c; its split
// planes are fixed32(k / C) and its mesh starts at mesh0.
template <int TS, bool GLOBAL>
__global__ __launch_bounds__(256) void bmesh_force_kernel(const uint32_t* __restrict__ x,
                                                         const uint32_t* __restrict__ y,
                                                         const uint32_t* __restrict__ z,
                                                         const int8_t* __restrict__ rung,
                                                         const int32_t* __restrict__ cell_begin, int nbox,
                                                         const tpm_box_dev* __restrict__ boxes,
                                                         const float* __restrict__ phi, int W, int ncell,
                                                         float4* __restrict__ out) {
    extern __shared__ float sphi[];
    // per-XCD runs of the global cell space, each box's cells with z
    // fastest, as in mesh_force_kernel
    const int per = (ncell + 7) >> 3;
    const int rk = (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
    if (rk >= ncell) return;
    const int b = tpm_upper_index(nbox, rk, [&](int k) { return (int64_t)boxes[k].cell0; });
    const tpm_box_dev& B = boxes[b];
    const int C = B.C, M = B.M, rl = rk - B.cell0;
    int k3[3];
    if (rl < C * C * C) {
        k3[0] = rl / (C * C);
        k3[1] = (rl / C) % C;
        k3[2] = rl % C;
    } else {
        k3[0] = rl % C;
        k3[1] = (rl / C) % C;
        k3[2] = rl / (C * C);
    }
    const int c = B.cell0 + (k3[2] * C + k3[1]) * C + k3[0];
    const int cb = cell_begin[c], ce = cell_begin[c + 1];
    if (cb == ce) return;
    int r0[3];
#pragma unroll
    for (int d = 0; d < 3; d++)
        r0[d] = (int)(((uint64_t)tpm_split_plane(k3[d], B.ninv) * (uint64_t)M) >> 32) - 2;
    mesh_force_cell<TS, GLOBAL>(sphi, x, y, z, rung, B.min_rung, cb, ce, r0, phi + B.mesh0, M, W, out);
}

// W_sr = sum_c n_c * sum_{27 neighbours} n_c' (exact, 64-bit integer) over
// the global cells [0, ncell); with a box table, cell c belongs to the last
// box whose cell0 <= c.
// A single box's sinks may be restricted to the cells [lo, hi).
__global__ void candidate_pairs_kernel(int64_t ncell, int nbox, const tpm_box_dev* __restrict__ boxes, int C0,
                                       const int32_t* __restrict__ cell_begin, unsigned long long* __restrict__ out,
                                       int3 lo = make_int3(0, 0, 0), int3 hi = make_int3(INT_MAX, INT_MAX, INT_MAX)) {
    unsigned long long acc = 0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
        int C = C0;
        int64_t base = 0;
        if (boxes) {
            const int b = tpm_upper_index(nbox, c, [&](int k) { return (int64_t)boxes[k].cell0; });
            C = boxes[b].C;
            base = boxes[b].cell0;
        }
        const int64_t l = c - base;
        const int cx = (int)(l % C), cy = (int)((l / C) % C), cz = (int)(l / ((int64_t)C * C));
        if (cx < lo.x || cx >= hi.x || cy < lo.y || cy >= hi.y || cz < lo.z || cz >= hi.z) continue;
        const int64_t nc = cell_begin[c + 1] - cell_begin[c];
        int64_t s = 0;
        for (int nb = 0; nb < 27; nb++) {
            const int jx = (cx + nb % 3 - 1 + C) % C, jy = (cy + (nb / 3) % 3 - 1 + C) % C,
                      jz = (cz + nb / 9 - 1 + C) % C;
            const int64_t c2 = base + ((int64_t)jz * C + jy) * C + jx;
            s += cell_begin[c2 + 1] - cell_begin[c2];
        }
        acc += (unsigned long long)(nc * s);
    }
    atomicAdd(out, acc);
}

// The tiles whose cell lies in [lo, hi), appended to out (one atomic per
// wave).  The order of out does not matter: every tile is computed on its
// own.
__global__ void box_tiles_kernel(const int4* __restrict__ tiles, const int32_t* __restrict__ ntiles, int C, int3 lo,
                                 int3 hi, int4* __restrict__ out, int32_t* __restrict__ count) {
    const int nt = *ntiles;
    const int lane = (int)lane_id();
    for (int64_t t0 = (int64_t)blockIdx.x * blockDim.x; t0 < nt; t0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = t0 + threadIdx.x;
        int4 tile = make_int4(0, 0, 0, 0);
        bool in = false;
        if (t < nt) {
            tile = tiles[t];
            const int cx = tile.x % C, cy = (tile.x / C) % C, cz = tile.x / (C * C);
            in = cx >= lo.x && cx < hi.x && cy >= lo.y && cy < hi.y && cz >= lo.z && cz < hi.z;
        }
        const unsigned long long m = __ballot(in);
        int base = 0;
        if (lane == 0 && m) base = atomicAdd(count, __popcll(m));
        base = __shfl(base, 0);
        if (in) out[base + (int)count_below(m)] = tile;
    }
}

__global__ void set_i32_kernel(int32_t* p, int32_t v) { *p = v; }

// Drift of every live particle by v * t0 of its box (fixed32, periodic):
// raw += round(v t0 2^32) mod 2^32.
__global__ void drift_kernel(int64_t n, int nbox, const tpm_box_dev* __restrict__ boxes, uint32_t* __restrict__ x,
                             uint32_t* __restrict__ y, uint32_t* __restrict__ z, const float* __restrict__ vx,
                             const float* __restrict__ vy, const float* __restrict__ vz) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int b = tpm_upper_index(nbox, i, [&](int k) { return boxes[k].q0; });
        const double f = boxes[b].drift * 4294967296.0;
        x[i] += (uint32_t)(uint64_t)llrint((double)vx[i] * f);
        y[i] += (uint32_t)(uint64_t)llrint((double)vy[i] * f);
        z[i] += (uint32_t)(uint64_t)llrint((double)vz[i] * f);
    }
}

template <bool BATCH>
// grid = workgroups (a multiple of 8); wpg = tiles per workgroup (1 or 4;
// the persistent and kick_pme launches take one).
static void pp_launch(int ts, bool persistent, unsigned grid, hipStream_t s, const PPArgs& a, bool pme = false,
                      int wpg = 1, bool lists = false) {
    if (lists && ts == 16 && !pme) {
        if (persistent)
            pp_kick_kernel<16, BATCH, true, false, 1, true><<<grid, 64, 0, s>>>(a);
        else if (wpg == 4)
            pp_kick_kernel<16, BATCH, false, false, 4, true><<<grid, 256, 0, s>>>(a);
        else
            pp_kick_kernel<16, BATCH, false, false, 1, true><<<grid, 64, 0, s>>>(a);
        return;
    }
    if (pme) {
        switch (ts) {
            case 8: pp_kick_kernel<8, BATCH, false, true><<<grid, 64, 0, s>>>(a); break;
            case 16: pp_kick_kernel<16, BATCH, false, true><<<grid, 64, 0, s>>>(a); break;
            case 64: pp_kick_kernel<64, BATCH, false, true><<<grid, 64, 0, s>>>(a); break;
            default: pp_kick_kernel<32, BATCH, false, true><<<grid, 64, 0, s>>>(a); break;
        }
    } else if (persistent) {
        switch (ts) {
            case 8: pp_kick_kernel<8, BATCH, true><<<grid, 64, 0, s>>>(a); break;
            case 16: pp_kick_kernel<16, BATCH, true><<<grid, 64, 0, s>>>(a); break;
            case 64: pp_kick_kernel<64, BATCH, true><<<grid, 64, 0, s>>>(a); break;
            default: pp_kick_kernel<32, BATCH, true><<<grid, 64, 0, s>>>(a); break;
        }
    } else if (wpg == 4) {
        switch (ts) {
            case 8: pp_kick_kernel<8, BATCH, false, false, 4><<<grid, 256, 0, s>>>(a); break;
            case 16: pp_kick_kernel<16, BATCH, false, false, 4><<<grid, 256, 0, s>>>(a); break;
            case 64: pp_kick_kernel<64, BATCH, false, false, 4><<<grid, 256, 0, s>>>(a); break;
            default: pp_kick_kernel<32, BATCH, false, false, 4><<<grid, 256, 0, s>>>(a); break;
        }
    } else {
        switch (ts) {
            case 8: pp_kick_kernel<8, BATCH, false><<<grid, 64, 0, s>>>(a); break;
            case 16: pp_kick_kernel<16, BATCH, false><<<grid, 64, 0, s>>>(a); break;
            case 64: pp_kick_kernel<64, BATCH, false><<<grid, 64, 0, s>>>(a); break;
            default: pp_kick_kernel<32, BATCH, false><<<grid, 64, 0, s>>>(a); break;
        }
    }
}

}  // namespace

void tpm_box_fill(tpm_box_dev* b, const tpm_params& p, int min_rung, int first_call, double t0, double scale) {
    b->C = p.chain_dim;
    b->M = p.mesh_dim;
    b->ninv = 1.0 / p.chain_dim;
    b->min_rung = min_rung;
    b->first_call = first_call ? 1 : 0;
    b->do_phi = p.do_phi;
    // Constants of compute_pp_interaction (kick_treepm.cu:833-853), scaled
    // to 2^-32 units: lengths x 2^32, squares x 2^64.
    const float rcut = (float)p.rcut, rs = (float)p.rs, h = (float)p.hsoft;
    const float S = 0x1p32f;
    b->rcut2s = (rcut * rcut) * (S * S);
    // Culling keeps a margin of 512 raw units over rcut (float rounding of
    // the box distance stays below ~200); off for C < 5, where a source's
    // offset from the tile origin can wrap to the far image.
    const double cull = (double)p.rcut * 4294967296.0 + 512.0;
    b->e2s = (float)(-1.4426950408889634 / (4.0 * p.rs * p.rs) / (4294967296.0 * 4294967296.0));
    b->cull2s = (float)(cull * cull);
    b->cull = p.chain_dim >= 5 ? 1 : 0;
    b->inv2rs_s = (1.0f / rs / 2.0f) / S;
    b->tsp = 2.0f / sqrtf((float)M_PI);
    // A&S 7.1.26 p and 2/sqrt(pi), each times 1 / (2 rs) in 2^-32 units
    b->p_inv2rs_s = (float)(0.3275911 / (2.0 * p.rs) / 4294967296.0);
    b->tsp_inv2rs_s = (float)(2.0 / std::sqrt(M_PI) / (2.0 * p.rs) / 4294967296.0);
    b->inv4rs2_s = (float)(1.0 / (4.0 * p.rs * p.rs) / (4294967296.0 * 4294967296.0));
    b->k3_s = (float)(4.0 / (3.0 * std::sqrt(M_PI)) / std::pow(2.0 * p.rs * 4294967296.0, 3));
    b->h2s = (h * h) * (S * S);
    const float hinv = 1.f / h;
    b->hinv_s = hinv / S;
    b->h2inv_s = (hinv * hinv) / (S * S);
    b->h3inv_s = ((hinv * hinv * hinv) / (S * S)) / S;
    b->GM = (float)p.GM;
    b->phi0 = (float)(p.n_total * 4.0 * M_PI * p.rs * p.rs + (35.0 / 16.0) / p.hsoft);
    b->eta = (float)p.eta;
    b->t0 = (float)t0;
    b->scale = (float)scale;
    b->hsoft = (float)p.hsoft;
    b->drift = t0;
}

// Nodes per axis of the potential around one chain cell: the sinks of cell k
// (raw [b_k, b_k+1)) take stencil nodes from mesh index (b_k M >> 32) - 2 to
// ((b_k+1 - 1) M >> 32) + 3.  0 when the largest cube exceeds 32 KiB.
int tpm_mesh_force_width(int C, int M) {
    const double ninv = 1.0 / C;
    int w = 0;
    for (int k = 0; k < C; k++) {
        const uint64_t lo = tpm_split_plane(k, ninv);
        const uint64_t hi = (k + 1 < C ? tpm_split_plane(k + 1, ninv) : (1ull << 32)) - 1;
        const int a = (int)((lo * (uint64_t)M) >> 32) - 2, b = (int)((hi * (uint64_t)M) >> 32) + 3;
        w = std::max(w, b - a + 1);
    }
    return (int64_t)w * w * w * 4 <= 32768 ? w : 0;
}

tpm_status tpm_pp_kick(tpm_ctx* ctx, int min_rung, double scale, double t0, int first_call, unsigned flags,
                       tpm_stats* stats, const tpm_cell_box* box, int64_t ntiles_box) {
    if (!ctx->chain_valid || ctx->chain_C != ctx->C || ctx->chain_tile != ctx->tile)
        return tpm_fail(ctx, TPM_ERR_STATE, "tpm_kick: chain mesh not built (call tpm_chainmesh_create)");
    const bool use_pm = !(flags & TPM_SKIP_PM);
    if (use_pm && (!ctx->pm_valid || ctx->green_M != ctx->M))
        return tpm_fail(ctx, TPM_ERR_STATE, "tpm_kick: PM mesh not computed (call tpm_gravity_long_compute)");
    if (t0 <= 0.0) return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_kick: t0 must be positive");
    if (tpm_pp_sym_eligible(ctx, min_rung, flags, box)) return tpm_pp_kick_sym(ctx, scale, t0, first_call, flags, stats);
    hipStream_t s = ctx->stream;
    if (!ctx->counters) TPM_HIP(ctx, hipMalloc(&ctx->counters, sizeof(unsigned long long) * TPM_COUNTER_WORDS));
    TPM_TRY(tpm_fill_u64(ctx, ctx->counters, TPM_COUNTER_WORDS));

    PPArgs a{};
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
    float4* gmesh = nullptr;
    a.counters = ctx->counters;
    a.queue = nullptr;
    a.boxes = nullptr;
    a.flags = flags;
    const int64_t ncell = (int64_t)ctx->C * ctx->C * ctx->C;
    int64_t tiles_max = ctx->n / ctx->tile + ncell + 1;
    // the mesh force and potential of the active sinks, made before the
    // pair kernel (LDS cube per chain cell when it fits 32 KiB)
    ctx->mesh_force_w = tpm_mesh_force_width(ctx->C, ctx->M);
    if (use_pm && ctx->n > 0) {
        const int64_t n = ctx->n;
        if (n > ctx->mesh_force_cap) {
            if (ctx->mesh_force) TPM_HIP(ctx, hipFree(ctx->mesh_force));
            ctx->mesh_force = nullptr;
            ctx->mesh_force_cap = 0;
            TPM_HIP(ctx, hipMalloc(&ctx->mesh_force, sizeof(float4) * n));
            ctx->mesh_force_cap = n;
        }
        gmesh = static_cast<float4*>(ctx->mesh_force);
    }
    int3 blo = make_int3(0, 0, 0), bhi = make_int3(INT_MAX, INT_MAX, INT_MAX);
    if (box) {
        // the box's tiles into box_tiles (ntiles_box of them, counted by the caller)
        blo = make_int3(box->lo[0], box->lo[1], box->lo[2]);
        bhi = make_int3(box->hi[0], box->hi[1], box->hi[2]);
        // room for every tile of the context, so the append cannot overrun
        if (tiles_max > ctx->box_tiles_cap) {
            if (ctx->box_tiles) TPM_HIP(ctx, hipFree(ctx->box_tiles));
            ctx->box_tiles = nullptr;
            TPM_HIP(ctx, hipMalloc(&ctx->box_tiles, sizeof(int4) * tiles_max));
            ctx->box_tiles_cap = tiles_max;
        }
        if (ntiles_box > tiles_max) return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_kick_box: tile count");
        if (!ctx->box_count) TPM_HIP(ctx, hipMalloc(&ctx->box_count, sizeof(int32_t)));
        set_i32_kernel<<<1, 1, 0, s>>>(ctx->box_count, 0);
        box_tiles_kernel<<<grid_for(tiles_max, 256), 256, 0, s>>>(ctx->tiles, ctx->ntiles, ctx->C, blo, bhi,
                                                                  ctx->box_tiles, ctx->box_count);
        TPM_HIP(ctx, hipGetLastError());
        a.tiles = ctx->box_tiles;
        a.ntiles = ctx->box_count;
        tiles_max = ntiles_box;
    }
    tpm_box_fill(&a.box0, ctx->p, min_rung, first_call, t0, scale);
    a.box0.cell0 = 0;
    a.box0.mesh0 = 0;
    const bool pme = (flags & TPM_KICK_PME) != 0;
    if (pme) {
        // kick_pme: every pair of the 27 cells counts, and phi starts at 0
        a.box0.cull = 0;
        a.box0.phi0 = 0.f;
    }
    a.flags = flags & (TPM_SKIP_PP | TPM_SKIP_PM | TPM_SKIP_KICK);
    a.mesh = gmesh;

    // tiles per workgroup (kick_pme's launch: one)
    const int wpg = pme ? 1 : ctx->pp_waves;
    const int64_t per_xcd = (tiles_max + 7) / 8;
    const int64_t blocks = 8 * ((per_xcd + wpg - 1) / wpg);
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
    if (gmesh) {
        const int W = ctx->mesh_force_w;
        const size_t lds = sizeof(float) * (size_t)W * W * W;
        const unsigned nc = (unsigned)(8 * ((ncell + 7) / 8));  // 8 XCD runs
#define TPM_MF(TSV)                                                                                               \
    if (W > 0)                                                                                                  \
        mesh_force_kernel<TSV, false><<<nc, 256, lds, s>>>(ctx->x, ctx->y, ctx->z, ctx->rung, min_rung,         \
                                                           ctx->cell_begin, ctx->bounds, ctx->C, ctx->phi,      \
                                                           ctx->M, W, blo, bhi, gmesh);                         \
    else                                                                                                        \
        mesh_force_kernel<TSV, true><<<nc, 256, 0, s>>>(ctx->x, ctx->y, ctx->z, ctx->rung, min_rung,            \
                                                        ctx->cell_begin, ctx->bounds, ctx->C, ctx->phi, ctx->M, \
                                                        0, blo, bhi, gmesh)
        switch (ctx->tile) {
            case 8: TPM_MF(8); break;
            case 16: TPM_MF(16); break;
            case 32: TPM_MF(32); break;
            default: TPM_MF(64); break;
        }
#undef TPM_MF
        TPM_HIP(ctx, hipGetLastError());
    }
    if (ctx->n > 0 && blocks > 0) {
        pp_launch<false>(ctx->tile, false, (unsigned)blocks, s, a, pme, wpg, ctx->pp_lists != 0);
        TPM_HIP(ctx, hipGetLastError());
    }
    TPM_HIP(ctx, hipEventRecord(e1, s));
    if (stats) {
        candidate_pairs_kernel<<<grid_for(ncell, 256), 256, 0, s>>>(ncell, 0, nullptr, ctx->C, ctx->cell_begin,
                                                                    ctx->counters + 2, blo, bhi);
        TPM_HIP(ctx, hipGetLastError());
        std::vector<unsigned long long> hw(TPM_COUNTER_WORDS);
        int32_t nt = 0;
        TPM_HIP(ctx, hipMemcpyAsync(hw.data(), ctx->counters, sizeof(unsigned long long) * TPM_COUNTER_WORDS,
                                    hipMemcpyDeviceToHost, s));
        TPM_HIP(ctx, hipMemcpyAsync(&nt, a.ntiles, sizeof(nt), hipMemcpyDeviceToHost, s));
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

tpm_status tpm_candidate_pairs(tpm_ctx* ctx, unsigned long long* dst, const tpm_cell_box* box) {
    const int64_t ncell = (int64_t)ctx->C * ctx->C * ctx->C;
    int3 lo = make_int3(0, 0, 0), hi = make_int3(INT_MAX, INT_MAX, INT_MAX);
    if (box) {
        lo = make_int3(box->lo[0], box->lo[1], box->lo[2]);
        hi = make_int3(box->hi[0], box->hi[1], box->hi[2]);
    }
    candidate_pairs_kernel<<<grid_for(ncell, 256), 256, 0, ctx->stream>>>(ncell, 0, nullptr, ctx->C, ctx->cell_begin,
                                                                          dst, lo, hi);
    TPM_HIP(ctx, hipGetLastError());
    return TPM_OK;
}

// The fused kernel over the tiles of every live box of a batch.
tpm_status tpm_batch_pp(tpm_batch* bt, const tpm_batch_layout& L, unsigned flags) {
    if (L.n_live == 0 || L.ncell == 0) return TPM_OK;
    PPArgs a{};
    a.x = bt->x;
    a.y = bt->y;
    a.z = bt->z;
    a.vx = bt->vx;
    a.vy = bt->vy;
    a.vz = bt->vz;
    a.rung = bt->rung;
    a.gx = bt->gx;
    a.gy = bt->gy;
    a.gz = bt->gz;
    a.pot = bt->pot;
    a.cell_begin = bt->cell_begin;
    a.tiles = bt->tiles;
    a.ntiles = bt->tile_off + L.ncell;
    a.mesh = nullptr;
    a.counters = bt->counters;
    a.queue = reinterpret_cast<int*>(bt->counters + 8);
    a.boxes = bt->d_boxes;
    a.flags = flags & (TPM_SKIP_PP | TPM_SKIP_PM | TPM_SKIP_KICK);
    if (!(flags & TPM_SKIP_PM)) {
        // the mesh force of every box's active sinks, before the pair kernel
        const size_t lds = sizeof(float) * (size_t)L.mfw * L.mfw * L.mfw;
        const unsigned nc = (unsigned)(8 * ((L.ncell + 7) / 8));  // 8 XCD runs
#define TPM_BMF(TSV)                                                                                              \
    if (L.mfw > 0)                                                                                              \
        bmesh_force_kernel<TSV, false><<<nc, 256, lds, bt->stream>>>(bt->x, bt->y, bt->z, bt->rung,             \
                                                                     bt->cell_begin, L.nbox, bt->d_boxes,       \
                                                                     bt->phi, L.mfw, (int)L.ncell,              \
                                                                     bt->mesh_force);                           \
    else                                                                                                        \
        bmesh_force_kernel<TSV, true><<<nc, 256, 0, bt->stream>>>(bt->x, bt->y, bt->z, bt->rung,                \
                                                                  bt->cell_begin, L.nbox, bt->d_boxes, bt->phi, \
                                                                  0, (int)L.ncell, bt->mesh_force)
        switch (bt->tune.tile_sinks) {
            case 8: TPM_BMF(8); break;
            case 16: TPM_BMF(16); break;
            case 32: TPM_BMF(32); break;
            default: TPM_BMF(64); break;
        }
#undef TPM_BMF
        TPM_HIP(bt, hipGetLastError());
        a.mesh = bt->mesh_force;
    }
    const bool persistent = bt->tune.sched == 1;
    int64_t grid = persistent ? bt->persist_grid : (L.tiles_max + 7) / 8 * 8;
    if (grid < 8) grid = 8;
    if (grid > INT_MAX) return tpm_fail(bt, TPM_ERR_INVALID, "tpm_batch_step: too many tiles");
    pp_launch<true>(bt->tune.tile_sinks, persistent, (unsigned)grid, bt->stream, a, false, 1, bt->tune.pp_lists != 0);
    TPM_HIP(bt, hipGetLastError());
    return TPM_OK;
}

tpm_status tpm_batch_candidates(tpm_batch* bt, const tpm_batch_layout& L) {
    if (L.ncell == 0) return TPM_OK;
    candidate_pairs_kernel<<<grid_for(L.ncell, 256), 256, 0, bt->stream>>>(L.ncell, L.nbox, bt->d_boxes, 0,
                                                                           bt->cell_begin, bt->counters + 2);
    TPM_HIP(bt, hipGetLastError());
    return TPM_OK;
}

tpm_status tpm_batch_drift(tpm_batch* bt, const tpm_batch_layout& L) {
    if (L.n_live == 0) return TPM_OK;
    drift_kernel<<<grid_for(L.n_live, 256), 256, 0, bt->stream>>>(L.n_live, L.nbox, bt->d_boxes, bt->x, bt->y, bt->z,
                                                                  bt->vx, bt->vy, bt->vz);
    TPM_HIP(bt, hipGetLastError());
    return TPM_OK;
}

// Workgroups of one persistent PP launch: resident workgroups per CU x CUs,
// a multiple of the 8 XCDs.
int tpm_pp_persistent_grid(int device, int ts, bool lists) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 0;
    int per_cu = 0;
    hipError_t e;
    if (lists && ts == 16) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pp_kick_kernel<16, true, true, false, 1, true>, 64, 0);
        if (e != hipSuccess || per_cu < 1) per_cu = 8;
        const int g = per_cu * prop.multiProcessorCount;
        return (g + 7) / 8 * 8;
    }
    switch (ts) {
        case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pp_kick_kernel<8, true, true>, 64, 0); break;
        case 16: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pp_kick_kernel<16, true, true>, 64, 0); break;
        case 64: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pp_kick_kernel<64, true, true>, 64, 0); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pp_kick_kernel<32, true, true>, 64, 0); break;
    }
    if (e != hipSuccess || per_cu < 1) per_cu = 8;
    const int g = per_cu * prop.multiProcessorCount;
    return (g + 7) / 8 * 8;
}
