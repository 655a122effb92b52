// tpm_tree.hip -- per-cell trees and kick_treepm's tree walk with multipole
// (PC) interactions (tree.cpp:7-245, kernels.hpp:13-152,
// kick_treepm.cu:226-240, 460-759), on the device.
//
// Build, one wave64 per chain cell:
//   tree_partition_kernel: the recursive midpoint split of tree.cpp's sort()
//     as a depth-first walk over an LDS stack.  Each split is a stable
//     wave-wide partition (ballot + mbcnt) from one buffer into the other, so
//     the leaves end in tree order as int4 records (x, y, z, chain-mesh
//     index); the kernel counts the cell's nodes and sink buckets.  A cell of
//     <= TREE_LDS_CAP particles is split in LDS (16-bit indices), a larger
//     one in device memory.
//   Exclusive scans of the counts give each cell's node and bucket base.
//   tree_nodes_kernel: the same walk over the partitioned positions (in LDS
//     up to the same size; a split point is the ballot count of particles
//     left of the plane), nodes numbered in preorder and written after their
//     children: a leaf's centre, radius and moments from its particles
//     (bounding box and farthest distance by wave reductions, exact in FP64;
//     moment sums in particle order), an internal node from its two children
//     as tree.cpp:101-171 does, and the sink buckets (tree.cpp:173-211).
//   The arithmetic is the reference's, operation for operation, so the trees
//   are bit-identical to the CPU restatement's.
//
// Walk, one wave64 per sink bucket of <= 128 sinks: lane l holds sinks
// (l & 31) + 32 k for k < K = ceil(sinks / 32) and takes every other
// multipole node and staged particle (source stream l >> 5), so a bucket of
// 89 sinks keeps 93 % of the lanes' sink slots busy.  The roots of the 27
// neighbour trees share one LDS stack; nodes are taken 64 at a time, each
// lane tests one (cutoff: R - r_src - r_snk > rcut; far: R^2 >
// (r_src + r_snk)^2 / theta^2); far nodes go to the multipole list, near
// internal nodes push their children, near leaves go to the particle list
// unless every active sink is beyond r_src / theta of them
// (kick_treepm.cu:630-680).  The multipole list is applied with
// compute_pc_interaction (greens_function + M2L, kernels.hpp:13-88); the
// leaves' particles are staged in LDS and, for each block of 32 sinks, those
// within reach of the block are applied with the pair law of the direct
// kernel (the others would add exact zeros).  The decisions do not depend on
// the visiting order, so they are the reference's breadth-first walk's.
// Then the mesh force and the kick epilogue, as in the direct kernel.
#include "tpm_pp_device.hpp"

#include <vector>

#define TREE_LEAF 32     // min(SOURCE_BUCKET_SIZE, SINK_BUCKET_SIZE), tigerpm.hpp:32-33
#define TREE_BUCKET 128  // SINK_BUCKET_SIZE
#define TREE_DEPTH 50    // MAX_DEPTH, tree.cpp:5
#define TREE_STACK 64    // partition frames: a depth-first walk holds <= depth + 2
#define NODE_STACK (2 * TREE_DEPTH + 4)  // numbering frames: the path and its pending right siblings
#define TREE_LDS_CAP 512  // cells up to this many particles are built in LDS
#define WALK_STACK 512   // walk: pending nodes of the 27 neighbour trees (>= 192 + depth suffices)
#define WALK_LIST 128    // multipole / leaf list entries between flushes
#define WALK_SBUF 256    // staged PP sources

namespace {

// fixed32 from double (fixed.hpp:63-75): 2^32 x truncated; a centre just
// outside [0, 1) wraps, which distance() treats periodically.
__device__ __forceinline__ uint32_t tree_fixed(double x) { return (uint32_t)(int64_t)(4294967296.0 * x); }
__device__ __forceinline__ double raw_to_double(uint32_t r) { return (double)r * 0x1p-32; }
// sqr(a, b, c) of util.hpp:32-34 is fmaf(a, a, fmaf(b, b, c * c)); the
// double instantiation tree.cpp uses rounds through FP32.
__device__ __forceinline__ double sqr3_d(double a, double b, double c) {
    return (double)fmaf((float)a, (float)a, fmaf((float)b, (float)b, (float)(c * c)));
}
__device__ __forceinline__ float sqr3_f(float a, float b, float c) { return fmaf(a, a, fmaf(b, b, c * c)); }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ uint32_t comp(int4 v, int d) { return (uint32_t)(d == 0 ? v.x : (d == 1 ? v.y : v.z)); }
// distance() (gravity_short.hpp:20-22) in box units.
__device__ __forceinline__ float dist(uint32_t a, uint32_t b) { return (float)(int)(a - b) * 0x1p-32f; }

__device__ __forceinline__ uint32_t wave_umin(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}
__device__ __forceinline__ uint32_t wave_umax(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}
__device__ __forceinline__ double wave_dmax(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = dmax(v, __shfl_xor(v, o));
    return v;
}

// range::split (range.hpp:118-133): the first strictly longest side.
__device__ __forceinline__ int split_dim(const double* b, const double* e) {
    int dim = 0;
    double span = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++)
        if (e[d] - b[d] > span) {
            span = e[d] - b[d];
            dim = d;
        }
    return dim;
}

struct PartFrame {
    double b[3], e[3];
    int begin, end, depth, sunk;
};

// Where a cell's records live while it is split.  A split at depth d reads
// buffer d & 1 and writes buffer (d + 1) & 1; a leaf leaves its particles in
// `pos` in tree order.
//   PartLds: cells of <= TREE_LDS_CAP particles.  Positions are read once
//     into LDS and the splits move 16-bit local indices; each leaf writes its
//     records (x, y, z, chain-mesh index) to `pos` once.
//   PartGlobal: larger cells; int4 records move between pos and tmp.
//   Both keep, for each split, the positions of the particles the
//   two-pointer sweep moves (misplaced right ones below the split point in
//   `ma`, misplaced left ones above it in `mb`), and map[] gives the device
//   index of the cell's k-th particle in the order the trees start from
//   (nullptr: the device order).
struct PartLds {
    const uint32_t *xs, *ys, *zs;
    uint16_t *ia, *ib;
    uint16_t *ma, *mb;
    int b0;
    int4* pos;
    const uint32_t* map;
    using rec = int;
    __device__ rec load(int d, int i) const { return ((d & 1) ? ib : ia)[i - b0]; }
    __device__ uint32_t coord(rec k, int dim) const { return dim == 0 ? xs[k] : (dim == 1 ? ys[k] : zs[k]); }
    __device__ void store(int d, int i, rec k) const { ((d & 1) ? ib : ia)[i - b0] = (uint16_t)k; }
    __device__ void set_a(int k, int i) const { ma[k] = (uint16_t)(i - b0); }
    __device__ void set_b(int k, int i) const { mb[k] = (uint16_t)(i - b0); }
    __device__ int get_a(int k) const { return b0 + ma[k]; }
    __device__ int get_b(int k) const { return b0 + mb[k]; }
    __device__ void leaf(int d, int begin, int end) const {
        const uint16_t* buf = (d & 1) ? ib : ia;
        for (int i = begin + (int)lane_id(); i < end; i += 64) {
            const int k = buf[i - b0];
            pos[i] = make_int4((int)xs[k], (int)ys[k], (int)zs[k], map ? (int)map[b0 + k] : b0 + k);
        }
    }
};
struct PartGlobal {
    int4 *pos, *tmp;
    int32_t* moved;  // this cell's [b0, e0) of a device buffer: a in the low half, b in the high half
    int half;        // (e0 - b0) / 2
    __device__ void set_a(int k, int i) const { moved[k] = i; }
    __device__ void set_b(int k, int i) const { moved[half + k] = i; }
    __device__ int get_a(int k) const { return moved[k]; }
    __device__ int get_b(int k) const { return moved[half + k]; }
    using rec = int4;
    // (a pointer chosen by d & 1 would be looked up in scratch memory)
    __device__ rec load(int d, int i) const { return (d & 1) ? tmp[i] : pos[i]; }
    __device__ uint32_t coord(rec v, int dim) const { return comp(v, dim); }
    __device__ void store(int d, int i, rec v) const {
        if (d & 1)
            tmp[i] = v;
        else
            pos[i] = v;
    }
    __device__ void leaf(int d, int begin, int end) const {
        if (d & 1)
            for (int i = begin + (int)lane_id(); i < end; i += 64) pos[i] = tmp[i];
    }
};

// tree.cpp's sort() for one cell as a depth-first walk; returns the node
// and sink-bucket counts, or false past the depth limit.
template <class S>
__device__ __forceinline__ bool partition_cell(const S s, PartFrame* st, int c, int C, int b0, int e0, int& nodes, int& buckets) {
    const int lane = (int)lane_id();
    if (lane == 0) {
        // tree_create (tree.cpp:216-245): the cell's box [i / C, (i + 1) / C)
        const double ninv = 1.0 / (double)C;
        const int ic[3] = {c % C, (c / C) % C, c / (C * C)};
        PartFrame f;
        for (int d = 0; d < 3; d++) {
            f.b[d] = ic[d] * ninv;
            f.e[d] = (ic[d] + 1) * ninv;
        }
        f.begin = b0;
        f.end = e0;
        f.depth = 0;
        f.sunk = 0;
        st[0] = f;
    }
    __syncthreads();
    int sp = 1;
    nodes = buckets = 0;
    while (sp > 0) {
        --sp;
        PartFrame f;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            f.b[d] = st[sp].b[d];
            f.e[d] = st[sp].e[d];
        }
        f.begin = st[sp].begin;
        f.end = st[sp].end;
        f.depth = st[sp].depth;
        f.sunk = st[sp].sunk;
        const int dim = split_dim(f.b, f.e);
        const double mid = (st[sp].e[dim] + st[sp].b[dim]) / 2.0;  // from LDS: f.b[dim] would go to scratch
        __syncthreads();
        if (f.depth > TREE_DEPTH) return false;  // tree.cpp:16-19 aborts
        nodes++;
        const int n = f.end - f.begin;
        if (n <= TREE_BUCKET && !f.sunk) buckets++;
        if (n <= TREE_LEAF) {
            s.leaf(f.depth, f.begin, f.end);
            continue;
        }
        const uint32_t thr = tree_fixed(mid);
        // particles_sort (particles.cpp:414-445): x < xmid goes left.  Its
        // two-pointer sweep in closed form: with m = begin + (particles
        // left of the plane), the k-th right particle of [begin, m) counted
        // upwards trades places with the k-th left particle of [m, end)
        // counted downwards; the others stay.
        int nl = 0;
        for (int i0 = f.begin; i0 < f.end; i0 += 64) {
            const int i = i0 + lane;
            const bool left = i < f.end && s.coord(s.load(f.depth, i), dim) < thr;
            nl += __popcll(__ballot(left));
        }
        const int m = f.begin + nl;
        int na = 0, nb = 0;
        for (int i0 = f.begin; i0 < f.end; i0 += 64) {
            const int i = i0 + lane;
            const bool valid = i < f.end;
            const bool left = valid && s.coord(s.load(f.depth, i), dim) < thr;
            const bool a_ = valid && i < m && !left, b_ = valid && i >= m && left;
            const unsigned long long ma = __ballot(a_), mb = __ballot(b_);
            if (a_) s.set_a(na + (int)count_below(ma), i);
            if (b_) s.set_b(nb + (int)count_below(mb), i);
            na += __popcll(ma);
            nb += __popcll(mb);
        }
        __syncthreads();
        int ka = 0, kb = 0;
        for (int i0 = f.begin; i0 < f.end; i0 += 64) {
            const int i = i0 + lane;
            const bool valid = i < f.end;
            typename S::rec v{};
            if (valid) v = s.load(f.depth, i);
            const bool left = valid && s.coord(v, dim) < thr;
            const bool a_ = valid && i < m && !left, b_ = valid && i >= m && left;
            const unsigned long long ma = __ballot(a_), mb = __ballot(b_);
            int dst = i;
            // a of upward rank k <-> b of downward rank k, i.e. upward rank nb - 1 - k
            if (a_) dst = s.get_b(nb - 1 - (ka + (int)count_below(ma)));
            if (b_) dst = s.get_a(nb - 1 - (kb + (int)count_below(mb)));
            if (valid) s.store(f.depth + 1, dst, v);
            ka += __popcll(ma);
            kb += __popcll(mb);
        }
        __syncthreads();
        if (sp + 2 > TREE_STACK) return false;  // unreachable below the depth limit
        if (lane == 0) {
            PartFrame& R = st[sp];
            PartFrame& L = st[sp + 1];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                L.b[d] = R.b[d] = f.b[d];
                L.e[d] = R.e[d] = f.e[d];
            }
            L.e[dim] = mid;  // LDS, indexed at run time
            R.b[dim] = mid;
            L.begin = f.begin;
            R.end = f.end;
            L.end = R.begin = f.begin + nl;
            L.depth = R.depth = f.depth + 1;
            L.sunk = R.sunk = n <= TREE_BUCKET;
        }
        sp += 2;
        __syncthreads();
    }
    return true;
}

// map (or nullptr): map[i] = device index of the particle the trees'
// starting order holds at position i of its cell (the caller's order when it
// differs from the device's), so the two-pointer sweeps start where the
// reference's tree_create starts.  moved: n ints of scratch for the
// partition of cells larger than TREE_LDS_CAP.
__global__ __launch_bounds__(64) void tree_partition_kernel(int C, const int32_t* __restrict__ cell_begin,
                                                            const uint32_t* __restrict__ x,
                                                            const uint32_t* __restrict__ y,
                                                            const uint32_t* __restrict__ z,
                                                            const uint32_t* __restrict__ map, int4* __restrict__ pos,
                                                            int4* __restrict__ tmp, int32_t* __restrict__ moved,
                                                            int32_t* __restrict__ cnt, int32_t* __restrict__ flags) {
    __shared__ PartFrame st[TREE_STACK];
    __shared__ uint32_t xs[TREE_LDS_CAP], ys[TREE_LDS_CAP], zs[TREE_LDS_CAP];
    __shared__ uint16_t ia[TREE_LDS_CAP], ib[TREE_LDS_CAP];
    __shared__ uint16_t ma[TREE_LDS_CAP / 2], mb[TREE_LDS_CAP / 2];
    const int c = blockIdx.x;
    const int ncell = C * C * C;
    const int lane = (int)lane_id();
    const int b0 = cell_begin[c], e0 = cell_begin[c + 1];
    int nodes = 0, buckets = 0;
    bool ok;
    if (e0 - b0 <= TREE_LDS_CAP) {
        for (int k = lane; k < e0 - b0; k += 64) {
            const int dk = map ? (int)map[b0 + k] : b0 + k;
            xs[k] = x[dk];
            ys[k] = y[dk];
            zs[k] = z[dk];
            ia[k] = (uint16_t)k;
        }
        ok = partition_cell(PartLds{xs, ys, zs, ia, ib, ma, mb, b0, pos, map}, st, c, C, b0, e0, nodes, buckets);
    } else {
        for (int i = b0 + lane; i < e0; i += 64) {
            const int di = map ? (int)map[i] : i;
            pos[i] = make_int4((int)x[di], (int)y[di], (int)z[di], di);
        }
        ok = partition_cell(PartGlobal{pos, tmp, moved + b0, (e0 - b0) / 2}, st, c, C, b0, e0, nodes, buckets);
    }
    if (lane == 0) {
        cnt[c] = nodes;
        cnt[ncell + c] = buckets;
        if (!ok) atomicOr(&flags[0], 1);
    }
}

// Read access to a cell's partitioned positions, pos[i] for absolute i:
// from LDS for cells of <= TREE_LDS_CAP particles, else from the records.
struct PosLds {
    const uint32_t *xs, *ys, *zs;
    int b0;
    __device__ int4 operator[](int i) const {
        const int k = i - b0;
        return make_int4((int)xs[k], (int)ys[k], (int)zs[k], 0);
    }
};
struct PosGlobal {
    const int4* __restrict__ p;
    __device__ int4 operator[](int i) const { return p[i]; }
};

// Centre of the bounding box of [begin, end) (the box centre when empty) and
// the largest squared distance of a particle from it (tree.cpp:25-59,
// 88-99); every lane returns the same values.
template <class P>
__device__ double tree_bbox(const P& pos, int begin, int end, const double* bb, const double* be, double* xc) {
    // min and max in fixed point: raw -> double is exact and monotonic, so
    // these are the reference's FP64 extremes
    const int lane = (int)lane_id();
    uint32_t mn[3] = {~0u, ~0u, ~0u}, mx[3] = {0u, 0u, 0u};
    for (int i = begin + lane; i < end; i += 64) {
        const int4 v = pos[i];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            mx[d] = max(mx[d], comp(v, d));
            mn[d] = min(mn[d], comp(v, d));
        }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double lo = raw_to_double(wave_umin(mn[d])), hi = raw_to_double(wave_umax(mx[d]));
        xc[d] = (end > begin) ? (hi + lo) * 0.5 : (bb[d] + be[d]) * 0.5;
    }
    double r2max = 0.0;
    for (int i = begin + lane; i < end; i += 64) {
        const int4 v = pos[i];
        double r2 = 0.0;
        for (int d = 0; d < 3; d++) {
            const double e = xc[d] - raw_to_double(comp(v, d));
            r2 += e * e;
        }
        r2max = dmax(r2max, r2);
    }
    return wave_dmax(r2max);
}

// A leaf of <= 32 particles (tree.cpp:24-100): lane k holds particle k; the
// moments {N, -sum dx} are summed in particle order.
template <class P>
__device__ tpm_tree_node tree_leaf(const P& pos, int begin, int end, const double* bb, const double* be,
                                   double hsoft) {
    const int lane = (int)lane_id();
    const int n = end - begin;
    double xc[3];
    const double r2max = tree_bbox(pos, begin, end, bb, be, xc);
    float dx0 = 0.f, dx1 = 0.f, dx2 = 0.f;
    if (lane < n) {
        const int4 v = pos[begin + lane];
        dx0 = (float)(raw_to_double(comp(v, 0)) - xc[0]);
        dx1 = (float)(raw_to_double(comp(v, 1)) - xc[1]);
        dx2 = (float)(raw_to_double(comp(v, 2)) - xc[2]);
    }
    float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f;
    for (int k = 0; k < n; k++) {  // k is wave-uniform: readlane, no LDS round trip
        m0 += 1.0f;
        m1 += -__int_as_float(__builtin_amdgcn_readlane(__float_as_int(dx0), k));
        m2 += -__int_as_float(__builtin_amdgcn_readlane(__float_as_int(dx1), k));
        m3 += -__int_as_float(__builtin_amdgcn_readlane(__float_as_int(dx2), k));
    }
    tpm_tree_node nd;
    for (int d = 0; d < 3; d++) nd.x[d] = tree_fixed(xc[d]);
    nd.radius = (float)(sqrt(r2max) + hsoft);
    nd.child[0] = nd.child[1] = -1;
    nd.begin = begin;
    nd.end = end;
    nd.m[0] = m0;
    nd.m[1] = m1;
    nd.m[2] = m2;
    nd.m[3] = m3;
    return nd;
}

// An internal node from its children (tree.cpp:101-171).
__device__ tpm_tree_node tree_internal(const tpm_tree_node& c0, const tpm_tree_node& c1, const double* bb,
                                       const double* be, double hsoft) {
    const float mass0 = c0.m[0], mass1 = c1.m[0];
    double n[3], xc[3];
    for (int d = 0; d < 3; d++) n[d] = raw_to_double(c1.x[d]) - raw_to_double(c0.x[d]);
    const double norminv = 1.0 / sqrt(sqr3_d(n[0], n[1], n[2]));
    for (int d = 0; d < 3; d++) n[d] *= norminv;
    for (int d = 0; d < 3; d++) {
        const double x0 = raw_to_double(c0.x[d]), x1 = raw_to_double(c1.x[d]);
        if (mass1 == 0.0f)
            xc[d] = x0;
        else if (mass0 == 0.0f)
            xc[d] = x1;
        else
            xc[d] = (x0 + x1 + n[d] * (double)(c1.radius - c0.radius)) * 0.5;
    }
    // M2M_kernel (kernels.hpp:134-152) of each child about the new centre
    tpm_tree_node nd;
    float mm[2][4];
    const tpm_tree_node* cs[2] = {&c0, &c1};
    for (int k = 0; k < 2; k++) {
        float X[3];
        for (int d = 0; d < 3; d++) X[d] = -(float)(raw_to_double(cs[k]->x[d]) - xc[d]);
        const float* Ma = cs[k]->m;
        mm[k][0] = Ma[0];
        mm[k][1] = fmaf(X[0], Ma[0], Ma[1]);
        mm[k][3] = fmaf(X[2], Ma[0], Ma[3]);
        mm[k][2] = fmaf(X[1], Ma[0], Ma[2]);
    }
    for (int j = 0; j < 4; j++) nd.m[j] = mm[0][j] + mm[1][j];
    for (int d = 0; d < 3; d++) nd.x[d] = tree_fixed(xc[d]);
    double r20 = 0.0, r21 = 0.0;
    for (int d = 0; d < 3; d++) {
        if (mass0 != 0.0f) {
            const double e = raw_to_double(c0.x[d]) - xc[d];
            r20 += e * e;
        }
        if (mass1 != 0.0f) {
            const double e = raw_to_double(c1.x[d]) - xc[d];
            r21 += e * e;
        }
    }
    double r2_max = 0.0;
    for (int corner = 0; corner < 8; corner++) {
        const double cx = (corner & 4) ? be[0] : bb[0];
        const double cy = (corner & 2) ? be[1] : bb[1];
        const double cz = (corner & 1) ? be[2] : bb[2];
        r2_max = dmax(r2_max, sqr3_d(cx - xc[0], cy - xc[1], cz - xc[2]));
    }
    nd.radius = (float)dmin(dmax(sqrt(r20) + (double)c0.radius, sqrt(r21) + (double)c1.radius), sqrt(r2_max) + hsoft);
    nd.begin = c0.begin;
    nd.end = c1.end;
    return nd;
}

struct NodeFrame {
    double b[3], e[3];
    int begin, end;
    int depth;
    int sunk;   // inside a sink bucket already
    int state;  // 0: not entered; 1: children pending
    int which;  // 0: child 0 of its parent, 1: child 1
    int index;  // preorder index in the cell's tree
    int c1;     // child 1's index
};

// The numbering walk over one cell.  The stack holds the path from the root
// and each ancestor's pending right child (<= 2 depth + 3 frames); a
// finished node is handed to its parent through kids[parent depth][which],
// which only the parent's two children write while the parent is on the
// path.  Returns false when the walk disagrees with the partition's counts.
template <class P>
__device__ __forceinline__ bool number_cell(const P pos, NodeFrame* st, tpm_tree_node (*kids)[2], int c, int C, int b0, int e0,
                            tpm_tree_node* __restrict__ out, tpm_tree_bucket* __restrict__ bout, int nmax, int bmax,
                            double hsoft) {
    const int lane = (int)lane_id();
    if (lane == 0) {
        const double ninv = 1.0 / (double)C;
        const int ic[3] = {c % C, (c / C) % C, c / (C * C)};
        NodeFrame f;
        for (int d = 0; d < 3; d++) {
            f.b[d] = ic[d] * ninv;
            f.e[d] = (ic[d] + 1) * ninv;
        }
        f.begin = b0;
        f.end = e0;
        f.depth = 0;
        f.sunk = 0;
        f.state = 0;
        f.which = 0;
        f.index = 0;
        f.c1 = -1;
        st[0] = f;
    }
    __syncthreads();
    int sp = 1, next = 0, nb = 0;
    while (sp > 0) {
        NodeFrame& f = st[sp - 1];
        const int begin = f.begin, end = f.end, n = end - begin, depth = f.depth;
        double bb[3], be[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            bb[d] = f.b[d];
            be[d] = f.e[d];
        }
        tpm_tree_node done;
        int index;
        if (f.state == 0) {
            index = next++;
            if (index >= nmax || sp + 2 > NODE_STACK || depth > TREE_DEPTH ||
                (n <= TREE_BUCKET && !f.sunk && nb >= bmax))
                return false;
            if (n <= TREE_BUCKET && !f.sunk) {
                double xc[3];
                const double r2max = tree_bbox(pos, begin, end, bb, be, xc);
                if (lane == 0) {
                    tpm_tree_bucket bk;
                    bk.begin = begin;
                    bk.end = end;
                    for (int d = 0; d < 3; d++) bk.x[d] = tree_fixed(xc[d]);
                    bk.radius = (float)(sqrt(r2max) + hsoft);
                    bk.cell = c;
                    bk.pad_ = 0;
                    bout[nb] = bk;
                }
                nb++;
            }
            if (n > TREE_LEAF) {
                const int dim = split_dim(bb, be);
                const double mid = (f.e[dim] + f.b[dim]) / 2.0;  // f is in LDS
                const uint32_t thr = tree_fixed(mid);
                // first particle of the right half: the records are
                // partitioned, so it is begin + the number left of thr
                int lo = begin;
                for (int i0 = begin; i0 < end; i0 += 64) {
                    const int i = i0 + lane;
                    lo += __popcll(__ballot(i < end && comp(pos[i], dim) < thr));
                }
                __syncthreads();
                if (lane == 0) {
                    f.state = 1;
                    f.index = index;
                    NodeFrame& R = st[sp];
                    NodeFrame& L = st[sp + 1];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        L.b[d] = R.b[d] = bb[d];
                        L.e[d] = R.e[d] = be[d];
                    }
                    L.e[dim] = mid;  // LDS, indexed at run time
                    R.b[dim] = mid;
                    L.begin = begin;
                    L.end = R.begin = lo;
                    R.end = end;
                    L.depth = R.depth = depth + 1;
                    L.sunk = R.sunk = n <= TREE_BUCKET;
                    L.state = R.state = 0;
                    L.which = 0;
                    R.which = 1;
                    L.index = R.index = 0;
                    L.c1 = R.c1 = -1;
                }
                sp += 2;
                __syncthreads();
                continue;
            }
            done = tree_leaf(pos, begin, end, bb, be, hsoft);
        } else {
            index = f.index;
            done = tree_internal(kids[depth][0], kids[depth][1], bb, be, hsoft);
            done.child[0] = index + 1;
            done.child[1] = f.c1;
        }
        // finished: store it, hand it to the parent, pop
        const int which = f.which;
        __syncthreads();
        if (lane == 0) {
            out[index] = done;
            if (sp > 1) {
                kids[depth - 1][which] = done;
                if (which == 1) st[sp - 2].c1 = index;  // the parent is just below its last child
            }
        }
        sp--;
        __syncthreads();
    }
    return true;
}

__global__ __launch_bounds__(64) void tree_nodes_kernel(int C, const int32_t* __restrict__ cell_begin,
                                                        const int4* __restrict__ pos,
                                                        const int32_t* __restrict__ node_begin,
                                                        const int32_t* __restrict__ bucket_begin,
                                                        tpm_tree_node* __restrict__ nodes,
                                                        tpm_tree_bucket* __restrict__ buckets, double hsoft,
                                                        int32_t* __restrict__ flags) {
    __shared__ NodeFrame st[NODE_STACK];
    __shared__ tpm_tree_node kids[TREE_DEPTH + 1][2];
    __shared__ uint32_t xs[TREE_LDS_CAP], ys[TREE_LDS_CAP], zs[TREE_LDS_CAP];
    const int c = blockIdx.x;
    const int lane = (int)lane_id();
    const int b0 = cell_begin[c], e0 = cell_begin[c + 1];
    tpm_tree_node* out = nodes + node_begin[c];
    tpm_tree_bucket* bout = buckets + bucket_begin[c];
    // the partition kernel counted these; a walk that disagrees stops
    const int nmax = node_begin[c + 1] - node_begin[c], bmax = bucket_begin[c + 1] - bucket_begin[c];
    bool ok;
    if (e0 - b0 <= TREE_LDS_CAP) {
        for (int k = lane; k < e0 - b0; k += 64) {
            const int4 v = pos[b0 + k];
            xs[k] = (uint32_t)v.x;
            ys[k] = (uint32_t)v.y;
            zs[k] = (uint32_t)v.z;
        }
        ok = number_cell(PosLds{xs, ys, zs, b0}, st, kids, c, C, b0, e0, out, bout, nmax, bmax, hsoft);
    } else {
        ok = number_cell(PosGlobal{pos}, st, kids, c, C, b0, e0, out, bout, nmax, bmax, hsoft);
    }
    if (!ok && lane == 0) atomicOr(&flags[2], 1);
}

// ---------------------------------------------------------------------------
// The walk.
// ---------------------------------------------------------------------------

struct WalkArgs {
    const int4* __restrict__ pos;
    const tpm_tree_node* __restrict__ nodes;
    const int32_t* __restrict__ node_begin;
    const tpm_tree_bucket* __restrict__ buckets;
    const float* __restrict__ phi;
    int32_t* __restrict__ flags_dev;  // [1]: walk stack overflow
    tpm_kick_out o;
    tpm_box_dev B;
    int nbuckets, per_xcd;
    const int32_t* __restrict__ blist;   // buckets to walk (nullptr: all)
    int n;                               // particles
    float rcut, theta2inv, inv2rs, tsp;  // box units
    float e2;                            // -log2(e) inv2rs^2: exp(-r0^2) = exp2(r^2 e2)
    float thetainv;                      // 1 / theta
    unsigned flags;
};

// compute_pc_interaction (kick_treepm.cu:226-240): greens_function
// (kernels.hpp:13-66) and the 4-component M2L_kernel (kernels.hpp:69-88),
// in box units; d = sink - expansion centre.
__device__ __forceinline__ void pc_interaction(const WalkArgs& a, float dx, float dy, float dz, const float* M,
                                               float& gx, float& gy, float& gz, float& ph) {
    const float r2 = sqr3_f(dx, dy, dz);
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    const float r0 = r * a.inv2rs;
    // erfcexp (util.hpp:36-59), with the hardware reciprocal and exp2 as in
    // the pair law
    const float t1 = __builtin_amdgcn_rcpf(fmaf(0.3275911f, r0, 1.f));
    const float exp0 = __builtin_amdgcn_exp2f(r2 * a.e2);
    const float t2 = t1 * t1, t3 = t2 * t1, t4 = t2 * t2, t5 = t2 * t3;
    const float erfc0 = fmaf(0.254829592f, t1,
                             fmaf(-0.284496736f, t2, fmaf(1.421413741f, t3, fmaf(-1.453152027f, t4, 1.061405429f * t5)))) *
                        exp0;
    const float c0 = -2.f * r * a.inv2rs;
    const float d0 = -erfc0 * rinv;
    float e0 = a.tsp * exp0 * rinv * a.inv2rs;
    const float d1 = fmaf(-1.f * d0, rinv, e0);
    e0 *= c0;
    const float d2 = fmaf(-3.f * d1, rinv, e0);
    const float D11 = d1 * rinv;
    const float x100 = dx * rinv, x010 = dy * rinv, x001 = dz * rinv;
    const float x002 = x001 * x001, x011 = x010 * x001, x020 = x010 * x010;
    const float x101 = x100 * x001, x110 = x100 * x010, x200 = x100 * x100;
    float x21 = x002;
    x21 += x020;
    x21 += x200;
    x21 *= D11;
    const float D0 = d0, D1 = x100 * d1, D2 = x010 * d1, D3 = x001 * d1;
    const float D4 = fmaf(x200, d2, 0.f) + x21;
    const float D5 = x110 * d2;
    const float D6 = x101 * d2;
    const float D7 = fmaf(x020, d2, x21);
    const float D8 = x011 * d2;
    const float D9 = fmaf(x002, d2, 0.f) + x21;
    float L0 = 0.f, L1 = 0.f, L2 = 0.f, L3 = 0.f;
    if (a.B.do_phi) {
        L0 = fmaf(M[0], D0, L0);
        L0 = fmaf(M[3], D3, L0);
        L0 = fmaf(M[2], D2, L0);
        L0 = fmaf(M[1], D1, L0);
    }
    L1 = fmaf(M[0], D1, L1);
    L2 = fmaf(M[2], D7, L2);
    L1 = fmaf(M[3], D6, L1);
    L2 = fmaf(M[1], D5, L2);
    L1 = fmaf(M[2], D5, L1);
    L3 = fmaf(M[0], D3, L3);
    L1 = fmaf(M[1], D4, L1);
    L3 = fmaf(M[3], D9, L3);
    L2 = fmaf(M[0], D2, L2);
    L3 = fmaf(M[2], D8, L3);
    L2 = fmaf(M[3], D8, L2);
    L3 = fmaf(M[1], D6, L3);
    gx -= L1;
    gy -= L2;
    gz -= L3;
    ph += L0;
}

struct WalkShared {
    int stack[WALK_STACK];  // (neighbour << 27) | node index within that tree, as unsigned bits
    int mlist[WALK_LIST];   // global node indices
    int plist[WALK_LIST];
    int4 sbuf[WALK_SBUF];   // staged sources of the pair list's leaves
    int4 cbuf[WALK_SBUF + 4];  // those within reach of one block (+ far padding)
    float box[4][6];        // each block's active sinks: bounding box of the offsets
    int nb_base[27], nb_cnt[27];  // the neighbour trees' first node and node count
};

// One sink bucket of <= 32 K sinks.  Lane l holds sinks (l & 31) + 32 k,
// k < K, and works on source stream l >> 5 (every other multipole node and
// staged particle); the two streams' sums meet at the end.
// HALF: the last block holds the <= 16 remaining sinks, lane l sink
// (l & 15) of it with source stream l >> 4 (four streams, half the pair
// iterations of a 32-sink block).
template <int K, bool HALF>
__device__ __forceinline__ void walk_bucket(const WalkArgs& a, const tpm_tree_bucket& bk, WalkShared& sh) {
    const tpm_box_dev& B = a.B;
    const int lane = (int)lane_id();
    const int slot = lane & 31, stream = lane >> 5, s4 = lane >> 4;
    const int ns = bk.end - bk.begin;
    // per sink slot: offsets from the bucket centre (exact int32, as the
    // direct kernel's; the position is the centre plus the offset) and the
    // chain-mesh index; active slots as a bit mask
    int ox[K], oy[K], oz[K], w[K];
    unsigned act = 0u;
    float pg[K][4], cg[K][4];  // PP sums (2^-64 / 2^-32 units), PC sums (box units)
    int nact = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const bool hb = HALF && k == K - 1;  // this block's sinks are held four times
        const int j = hb ? 32 * k + (lane & 15) : slot + 32 * k;
        ox[k] = oy[k] = oz[k] = 0;
        w[k] = -1;
        bool on = false;
        if (j < ns) {
            const int4 v = a.pos[bk.begin + j];
            ox[k] = (int)((uint32_t)v.x - bk.x[0]);
            oy[k] = (int)((uint32_t)v.y - bk.x[1]);
            oz[k] = (int)((uint32_t)v.z - bk.x[2]);
            w[k] = v.w;
            on = a.o.rung[v.w] >= B.min_rung;
        }
        act |= on ? 1u << k : 0u;
        nact += __popcll(__ballot(on && (hb ? s4 == 0 : stream == 0)));
#pragma unroll
        for (int c = 0; c < 4; c++) pg[k][c] = cg[k][c] = 0.f;
    }
    // distance() from slot k's sink to a point: the offset plus the centre's
    auto sdist = [&](int o, uint32_t centre, uint32_t to) { return dist((uint32_t)o + centre, to); };
    if (nact == 0) return;
    // Bounding box of each block's active sinks (offsets as floats: <= 64
    // raw units of rounding, inside the 512-unit margin of cull2s, so the
    // cull below never drops a source within rcut of an active sink).
    unsigned bact = 0u;  // blocks with an active sink
#pragma unroll
    for (int k = 0; k < K; k++) {
        const bool on = (act >> k) & 1u;
        bact |= __ballot(on) ? 1u << k : 0u;
        const float lo0 = (float)wave_min(on ? ox[k] : INT_MAX), hi0 = (float)wave_max(on ? ox[k] : INT_MIN);
        const float lo1 = (float)wave_min(on ? oy[k] : INT_MAX), hi1 = (float)wave_max(on ? oy[k] : INT_MIN);
        const float lo2 = (float)wave_min(on ? oz[k] : INT_MAX), hi2 = (float)wave_max(on ? oz[k] : INT_MIN);
        if (lane == 0) {
            sh.box[k][0] = lo0;
            sh.box[k][1] = hi0;
            sh.box[k][2] = lo1;
            sh.box[k][3] = hi1;
            sh.box[k][4] = lo2;
            sh.box[k][5] = hi2;
        }
    }
    unsigned long long npairs = 0, npc = 0;
    bool overflow = false;
    if (!(a.flags & TPM_SKIP_PP)) {
        const int C = B.C;
        const int cx = bk.cell % C, cy = (bk.cell / C) % C, cz = bk.cell / (C * C);
        const float rcut = a.rcut, t2 = a.theta2inv;
        const float sr = bk.radius;
        // Placeholder of lanes past the staged sources (never kept).
        const int4 far = make_int4(INT_MIN, INT_MIN, INT_MIN, 0);
        // the 27 neighbour trees (x outermost, z innermost, kick_treepm.cu:930-940)
        if (lane < 27) {
            int jx = cx + lane / 9 - 1, jy = cy + (lane / 3) % 3 - 1, jz = cz + lane % 3 - 1;
            jx += (jx < 0) ? C : 0;
            jx -= (jx >= C) ? C : 0;
            jy += (jy < 0) ? C : 0;
            jy -= (jy >= C) ? C : 0;
            jz += (jz < 0) ? C : 0;
            jz -= (jz >= C) ? C : 0;
            const int tcell = (jz * C + jy) * C + jx;
            sh.nb_base[lane] = a.node_begin[tcell];
            sh.nb_cnt[lane] = a.node_begin[tcell + 1] - a.node_begin[tcell];
        }
        __syncthreads();
        // all 27 trees are walked together from one stack, so a round takes
        // up to 64 pending nodes of any of them
        const bool root = lane < 27 && sh.nb_cnt[lane] > 0;
        const unsigned long long mr = __ballot(root);
        if (root) sh.stack[(int)count_below(mr)] = (int)((unsigned)lane << 27);
        int tn = lane < 27 ? sh.nb_cnt[lane] : 0;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) tn += __shfl_xor(tn, o);
        int sp = __popcll(mr), visited = 0, nm = 0, np = 0;
        __syncthreads();
        // applies and empties the multipole list (each stream every other node)
        auto flush_pc = [&]() {
            // the next node's centre and moments load while this one is applied
            uint32_t nx = 0, ny = 0, nz = 0;
            float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f;
            if (stream < nm) {
                const tpm_tree_node& nd = a.nodes[sh.mlist[stream]];
                nx = nd.x[0], ny = nd.x[1], nz = nd.x[2];
                m0 = nd.m[0], m1 = nd.m[1], m2 = nd.m[2], m3 = nd.m[3];
            }
            for (int q = stream; q < nm; q += 2) {
                const uint32_t cx0 = nx, cy0 = ny, cz0 = nz;
                const float M[4] = {m0, m1, m2, m3};
                if (q + 2 < nm) {
                    const tpm_tree_node& nn = a.nodes[sh.mlist[q + 2]];
                    nx = nn.x[0], ny = nn.x[1], nz = nn.x[2];
                    m0 = nn.m[0], m1 = nn.m[1], m2 = nn.m[2], m3 = nn.m[3];
                }
#pragma unroll
                for (int k = 0; k < K; k++) {
                    // the half block's sinks count once per stream pair: lanes
                    // with (l & 16) == 0 apply the node
                    if (HALF && k == K - 1 && (lane & 16)) continue;
                    pc_interaction(a, sdist(ox[k], bk.x[0], cx0), sdist(oy[k], bk.x[1], cy0),
                                   sdist(oz[k], bk.x[2], cz0), M, cg[k][0], cg[k][1], cg[k][2], cg[k][3]);
                }
            }
            npc += (unsigned long long)nm * nact;
            nm = 0;
        };
        // stages the leaf list's particles and applies them pair by pair
        auto flush_pp = [&]() {
            int q = 0;
            while (q < np) {
                int nbuf = 0;
                while (q < np) {
                    const tpm_tree_node& nd = a.nodes[__builtin_amdgcn_readfirstlane(sh.plist[q])];
                    const int cnt = min(max(nd.end - nd.begin, 0), TREE_LEAF);
                    if (nbuf + cnt > WALK_SBUF - 3) break;
                    if (lane < cnt && nd.begin + lane < a.n) {
                        const int4 v = a.pos[nd.begin + lane];
                        sh.sbuf[nbuf + lane] = make_int4((int)((uint32_t)v.x - bk.x[0]),
                                                         (int)((uint32_t)v.y - bk.x[1]),
                                                         (int)((uint32_t)v.z - bk.x[2]), 0);
                    }
                    nbuf += cnt;
                    q++;
                }
                __syncthreads();
                // Block by block: the staged sources within reach of the
                // block's sinks (box distance < rcut + margin; a dropped pair
                // has r >= rcut and adds exactly zero), then the pairs.  The
                // count is the walk's: every staged source times every active
                // sink, as the reference evaluates them.
#pragma unroll
                for (int k = 0; k < K; k++) {
                    if (!((bact >> k) & 1u)) continue;
                    const float lx = sh.box[k][0], hx = sh.box[k][1], ly = sh.box[k][2];
                    const float hy = sh.box[k][3], lz = sh.box[k][4], hz = sh.box[k][5];
                    int m = 0;
                    for (int i0 = 0; i0 < nbuf; i0 += 64) {
                        const int i = i0 + lane;
                        bool keep = false;
                        int4 sv = far;
                        if (i < nbuf) {
                            sv = sh.sbuf[i];
                            const float fx = (float)sv.x, fy = (float)sv.y, fz = (float)sv.z;
                            const float ex = fmaxf(fmaxf(lx - fx, fx - hx), 0.f);
                            const float ey = fmaxf(fmaxf(ly - fy, fy - hy), 0.f);
                            const float ez = fmaxf(fmaxf(lz - fz, fz - hz), 0.f);
                            keep = !B.cull || fmaf(ex, ex, fmaf(ey, ey, ez * ez)) < B.cull2s;
                        }
                        const unsigned long long mk = __ballot(keep);
                        if (keep) sh.cbuf[m + (int)count_below(mk)] = sv;
                        m += __popcll(mk);
                    }
                    const bool hb = HALF && k == K - 1;
                    const int S = hb ? 4 : 2, st = hb ? s4 : stream;  // source streams
                    // whole rounds of S sources, then the streams that still
                    // have one (no padding source: with a small chain mesh
                    // a stand-in placed anywhere can come within rcut)
                    const int full = m / S * S;
                    __syncthreads();
                    // one source ahead (the read past the end stays below
                    // m + 4 <= WALK_SBUF + 4 and is not used)
                    int4 sv = sh.cbuf[st];
                    for (int j = st; j < full; j += S) {
                        const int4 sn = sh.cbuf[j + S];
                        pp_pair<false>(B, ox[k], oy[k], oz[k], sv, pg[k][0], pg[k][1], pg[k][2], pg[k][3]);
                        sv = sn;
                    }
                    if (full + st < m)
                        pp_pair<false>(B, ox[k], oy[k], oz[k], sv, pg[k][0], pg[k][1], pg[k][2], pg[k][3]);
                    __syncthreads();
                }
                npairs += (unsigned long long)nbuf * nact;
            }
            np = 0;
        };
        while (sp > 0) {
            // 64 nodes a round while the stack has room for their children;
            // near the top one at a time, which grows the stack by at most
            // the remaining tree depth (<= 50)
            const int take = (sp + 128 <= WALK_STACK) ? (sp < 64 ? sp : 64) : 1;
            visited += take;
            if (visited > tn) {  // each node is met at most once
                overflow = true;
                break;
            }
            const unsigned e = lane < take ? (unsigned)sh.stack[sp - 1 - lane] : 0u;
            const int nbr = (int)(e >> 27), li = (int)(e & 0x7ffffffu);
            const bool valid = lane < take && nbr < 27 && li < sh.nb_cnt[nbr < 27 ? nbr : 0];
            const int idx = valid ? sh.nb_base[nbr] + li : 0;  // global node index
            __syncthreads();
            sp -= take;
            bool is_multi = false, is_open = false, is_leaf = false;
            tpm_tree_node nd;
            if (valid) {
                nd = a.nodes[idx];
                const float dx = dist(bk.x[0], nd.x[0]), dy = dist(bk.x[1], nd.x[1]), dz = dist(bk.x[2], nd.x[2]);
                const float R2 = sqr3_f(dx, dy, dz);
                const float R = sqrtf(R2);
                const bool cutoff = R - nd.radius - sr > rcut;
                const float rr = nd.radius + sr;
                const bool far_node = R2 > rr * rr * t2;
                const bool leaf = nd.end - nd.begin <= TREE_LEAF;
                is_multi = !cutoff && far_node;
                is_open = !cutoff && !far_node && !leaf;
                is_leaf = !cutoff && !far_node && leaf;
            }
            // A near leaf is decided here when every active block's box lies
            // clearly within r_src / theta of it (some active sink is near:
            // pairs) or clearly beyond (no sink is: multipole); margins of
            // 512 raw units and 1e-5 cover the float offsets' rounding and the
            // exact test's FP32 distances, so the outcome is the per-sink
            // test's.  The rest are tested sink by sink below.
            bool leaf_near = false, leaf_far = false;
            if (is_leaf && B.cull) {
                const float qx = (float)(int)(nd.x[0] - bk.x[0]), qy = (float)(int)(nd.x[1] - bk.x[1]);
                const float qz = (float)(int)(nd.x[2] - bk.x[2]);
                const float R = nd.radius * a.thetainv * 0x1p32f;
                const float rin = R * (1.f - 1e-5f) - 512.f, rout = R * (1.f + 1e-5f) + 512.f;
                bool anear = false, afar = true;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    if (!((bact >> k) & 1u)) continue;
                    const float lx = sh.box[k][0], hx = sh.box[k][1], ly = sh.box[k][2];
                    const float hy = sh.box[k][3], lz = sh.box[k][4], hz = sh.box[k][5];
                    const float ex = fmaxf(fmaxf(lx - qx, qx - hx), 0.f), ey = fmaxf(fmaxf(ly - qy, qy - hy), 0.f);
                    const float ez = fmaxf(fmaxf(lz - qz, qz - hz), 0.f);
                    const float fx = fmaxf(qx - lx, hx - qx), fy = fmaxf(qy - ly, hy - qy);
                    const float fz = fmaxf(qz - lz, hz - qz);
                    anear = anear || (rin > 0.f && fmaf(fx, fx, fmaf(fy, fy, fz * fz)) < rin * rin);
                    afar = afar && fmaf(ex, ex, fmaf(ey, ey, ez * ez)) > rout * rout;
                }
                leaf_near = anear;
                leaf_far = !anear && afar;
            }
            // near internal nodes: children onto the stack
            const unsigned long long mo = __ballot(is_open);
            const int nopen = __popcll(mo);
            if (sp + 2 * nopen > WALK_STACK) {  // unreachable below the depth limit
                overflow = true;
                break;
            }
            if (is_open) {
                const int at = sp + 2 * (int)count_below(mo);
                sh.stack[at] = (int)(((unsigned)nbr << 27) | (unsigned)nd.child[1]);
                sh.stack[at + 1] = (int)(((unsigned)nbr << 27) | (unsigned)nd.child[0]);
            }
            sp += 2 * nopen;
            // far nodes and clearly distant leaves: the multipole list
            const unsigned long long mm = __ballot(is_multi || leaf_far);
            if (nm + __popcll(mm) > WALK_LIST) {
                __syncthreads();
                flush_pc();
            }
            if (is_multi || leaf_far) sh.mlist[nm + (int)count_below(mm)] = idx;
            nm += __popcll(mm);
            // clearly near leaves: the pair list
            const unsigned long long mq = __ballot(leaf_near);
            if (np + __popcll(mq) > WALK_LIST) {
                __syncthreads();
                flush_pp();
            }
            if (leaf_near) sh.plist[np + (int)count_below(mq)] = idx;
            np += __popcll(mq);
            // the other near leaves: PC when every active sink is beyond r_src / theta
            unsigned long long ml = __ballot(is_leaf && !leaf_near && !leaf_far);
            while (ml) {
                const int j = __ffsll((long long)ml) - 1;
                ml &= ml - 1;
                const int lidx = __shfl(idx, j);
                const uint32_t lx = (uint32_t)__shfl((int)nd.x[0], j), ly = (uint32_t)__shfl((int)nd.x[1], j),
                               lz = (uint32_t)__shfl((int)nd.x[2], j);
                const float lr = __shfl(nd.radius, j);
                const float lim = lr * lr * t2;
                bool nearby = false;
#pragma unroll
                for (int k = 0; k < K; k++)
                    nearby |= ((act >> k) & 1u) &&
                              !(sqr3_f(sdist(ox[k], bk.x[0], lx), sdist(oy[k], bk.x[1], ly),
                                       sdist(oz[k], bk.x[2], lz)) > lim);
                if (__ballot(nearby) != 0ull) {
                    if (np == WALK_LIST) {
                        __syncthreads();
                        flush_pp();
                    }
                    if (lane == 0) sh.plist[np] = lidx;
                    np++;
                } else {
                    if (nm == WALK_LIST) {
                        __syncthreads();
                        flush_pc();
                    }
                    if (lane == 0) sh.mlist[nm] = lidx;
                    nm++;
                }
            }
            __syncthreads();
        }
        __syncthreads();
        if (!overflow) {
            flush_pc();
            flush_pp();
        }
    }
    if (overflow) {
        if (lane == 0) atomicOr(&a.flags_dev[1], 1);
        return;
    }
    // this workgroup's counter slot (tpm_reduce_counters)
    tpm_kick_out o = a.o;
    if (o.counters) o.counters += TPM_COUNTER_STRIDE * (blockIdx.x % TPM_COUNTER_SLOTS);
    if (lane == 0 && o.counters) {
        atomicAdd(&o.counters[0], npairs);
        atomicAdd(&o.counters[3], npc);
    }
    // the two streams' sums: short range in box units (pg), then the mesh
    // force, each sink's six x planes split between the streams (cg)
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (HALF && k == K - 1) {  // four streams
                pg[k][c] += __shfl_xor(pg[k][c], 16);
                cg[k][c] += __shfl_xor(cg[k][c], 16);
            }
            pg[k][c] += __shfl_xor(pg[k][c], 32);
            cg[k][c] += __shfl_xor(cg[k][c], 32);
            pg[k][c] = pg[k][c] * (c < 3 ? 0x1p64f : 0x1p32f) + cg[k][c];
            cg[k][c] = 0.f;
        }
    const bool use_mesh = a.phi && !(a.flags & TPM_SKIP_PM);
    if (use_mesh) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const bool hb = HALF && k == K - 1;
            if ((act >> k) & 1u)
                mesh_interp6_part(a.phi, B.M, (uint32_t)ox[k] + bk.x[0], (uint32_t)oy[k] + bk.x[1],
                                  (uint32_t)oz[k] + bk.x[2], hb ? s4 : stream, hb ? 4 : 2, cg[k][0], cg[k][1],
                                  cg[k][2], cg[k][3]);
#pragma unroll
            for (int c = 0; c < 4; c++) {
                if (hb) cg[k][c] += __shfl_xor(cg[k][c], 16);
                cg[k][c] += __shfl_xor(cg[k][c], 32);
            }
        }
    }
    const float Mf = (float)B.M;
    int rmax = -1;  // the new rungs' maximum (the kicks are done by stream 0)
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (stream != 0 || !((act >> k) & 1u)) continue;
        if (HALF && k == K - 1 && lane >= 16) continue;  // the half block's copies
        const float pgx = use_mesh ? -cg[k][0] * Mf : 0.f, pgy = use_mesh ? -cg[k][1] * Mf : 0.f;
        const float pgz = use_mesh ? -cg[k][2] * Mf : 0.f, pph = use_mesh ? cg[k][3] : 0.f;
        rmax = max(rmax, kick_epilogue(B, a.flags, a.o, w[k], a.o.rung[w[k]], pgx, pgy, pgz, pph, pg[k][0],
                                       pg[k][1], pg[k][2], pg[k][3]));
    }
    max_rung_update(o, rmax);
}

__global__ __launch_bounds__(64) void tree_walk_kernel(WalkArgs a) {
    __shared__ WalkShared sh;
    // one contiguous run of buckets per XCD (workgroups are dealt round-robin)
    const int wg = blockIdx.x;
    const int k = wg >> 3;
    if (k >= a.per_xcd) return;
    const int item = (wg & 7) * a.per_xcd + k;
    if (item >= a.nbuckets) return;
    const int bi = a.blist ? a.blist[item] : item;
    const tpm_tree_bucket bk = a.buckets[bi];
    const int ns = bk.end - bk.begin;
    const int K = (ns + 31) >> 5;          // sinks per lane
    const bool half = ns - 32 * (K - 1) <= 16;  // the last block fits 16 lanes
    switch (2 * K + (half ? 1 : 0)) {
        case 2: walk_bucket<1, false>(a, bk, sh); break;
        case 3: walk_bucket<1, true>(a, bk, sh); break;
        case 4: walk_bucket<2, false>(a, bk, sh); break;
        case 5: walk_bucket<2, true>(a, bk, sh); break;
        case 6: walk_bucket<3, false>(a, bk, sh); break;
        case 7: walk_bucket<3, true>(a, bk, sh); break;
        case 8: walk_bucket<4, false>(a, bk, sh); break;
        case 9: walk_bucket<4, true>(a, bk, sh); break;
        default: break;  // an empty bucket (buckets hold <= 128 sinks)
    }
}

// order[i] = index of tree position i's particle: its device index, or
// through d2h its place in the caller's order.
__global__ void tree_order_kernel(int64_t n, const int4* __restrict__ pos, const int32_t* __restrict__ d2h,
                                  int32_t* __restrict__ order) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        order[i] = d2h ? d2h[pos[i].w] : pos[i].w;
}

__global__ void tree_invert_kernel(int64_t n, const uint32_t* __restrict__ h2d, int32_t* __restrict__ d2h) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        d2h[h2d[r]] = (int32_t)r;
}

// The caller's order becomes the trees' order: h2d[r] = device index of
// tree position r.
__global__ void tree_host_order_kernel(int64_t n, const int4* __restrict__ pos, uint32_t* __restrict__ h2d) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        h2d[r] = (uint32_t)pos[r].w;
}

__global__ void tree_fill_i32_kernel(int32_t* p, int n, int32_t v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <class T>
tpm_status grow(tpm_ctx* ctx, T** p, int64_t* cap, int64_t need) {
    if (need <= *cap && *p) return TPM_OK;
    if (*p) TPM_HIP(ctx, hipFree(*p));
    *p = nullptr;
    *cap = 0;
    TPM_HIP(ctx, hipMalloc((void**)p, sizeof(T) * (size_t)(need > 0 ? need : 1)));
    *cap = need;
    return TPM_OK;
}

}  // namespace

tpm_status tpm_tree_ensure(tpm_ctx* ctx) {
    if (!ctx->chain_valid || ctx->chain_C != ctx->C)
        return tpm_fail(ctx, TPM_ERR_STATE, "tree: chain mesh not built (call tpm_chainmesh_create)");
    if (ctx->tree_epoch == ctx->chain_epoch) return TPM_OK;
    hipStream_t s = ctx->stream;
    const int C = ctx->C;
    const int64_t ncell = (int64_t)C * C * C;
    const int64_t n = ctx->n;
    if (n > ctx->tree_pos_cap || !ctx->tree_pos || !ctx->tree_tmp || !ctx->tree_moved) {
        int64_t c1 = 0, c2 = 0, c3 = 0;
        if (ctx->tree_pos) TPM_HIP(ctx, hipFree(ctx->tree_pos));
        if (ctx->tree_tmp) TPM_HIP(ctx, hipFree(ctx->tree_tmp));
        if (ctx->tree_moved) TPM_HIP(ctx, hipFree(ctx->tree_moved));
        ctx->tree_pos = ctx->tree_tmp = nullptr;
        ctx->tree_moved = nullptr;
        ctx->tree_pos_cap = 0;
        TPM_TRY(grow(ctx, &ctx->tree_pos, &c1, n));
        TPM_TRY(grow(ctx, &ctx->tree_tmp, &c2, n));
        TPM_TRY(grow(ctx, &ctx->tree_moved, &c3, n));
        ctx->tree_pos_cap = n;
    }
    if (ncell > ctx->tree_cells_cap || !ctx->tree_cnt) {
        int64_t c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0;
        if (ctx->tree_cnt) TPM_HIP(ctx, hipFree(ctx->tree_cnt));
        if (ctx->node_begin) TPM_HIP(ctx, hipFree(ctx->node_begin));
        if (ctx->bucket_begin) TPM_HIP(ctx, hipFree(ctx->bucket_begin));
        if (ctx->tree_part) TPM_HIP(ctx, hipFree(ctx->tree_part));
        if (ctx->tree_flags) TPM_HIP(ctx, hipFree(ctx->tree_flags));
        ctx->tree_cnt = ctx->node_begin = ctx->bucket_begin = ctx->tree_part = ctx->tree_flags = nullptr;
        TPM_TRY(grow(ctx, &ctx->tree_cnt, &c1, 2 * ncell));
        TPM_TRY(grow(ctx, &ctx->node_begin, &c2, ncell + 1));
        TPM_TRY(grow(ctx, &ctx->bucket_begin, &c3, ncell + 1));
        TPM_TRY(grow(ctx, &ctx->tree_part, &c4, ncell / 2048 + 2));
        TPM_TRY(grow(ctx, &ctx->tree_flags, &c5, 4));
        ctx->tree_cells_cap = ncell;
    }
    tree_fill_i32_kernel<<<1, 64, 0, s>>>(ctx->tree_flags, 4, 0);
    // the trees start from the caller's order (tree_create partitions the
    // caller's arrays in place, tree.cpp:101-107)
    const uint32_t* map = ctx->host_order == TPM_ORDER_REFERENCE && !ctx->h2d_identity ? ctx->h2d : nullptr;
    tree_partition_kernel<<<(unsigned)ncell, 64, 0, s>>>(C, ctx->cell_begin, ctx->x, ctx->y, ctx->z, map,
                                                          ctx->tree_pos, ctx->tree_tmp, ctx->tree_moved, ctx->tree_cnt,
                                                          ctx->tree_flags);
    TPM_HIP(ctx, hipGetLastError());
    TPM_TRY(tpm_scan_i32(ctx->err, s, ctx->tree_cnt, ncell, ctx->tree_part, ctx->node_begin));
    TPM_TRY(tpm_scan_i32(ctx->err, s, ctx->tree_cnt + ncell, ncell, ctx->tree_part, ctx->bucket_begin));
    int32_t tot[3] = {0, 0, 0};
    TPM_HIP(ctx, hipMemcpyAsync(&tot[0], ctx->node_begin + ncell, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    TPM_HIP(ctx, hipMemcpyAsync(&tot[1], ctx->bucket_begin + ncell, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    TPM_HIP(ctx, hipMemcpyAsync(&tot[2], ctx->tree_flags, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    TPM_HIP(ctx, hipStreamSynchronize(s));
    if (tot[2])
        return tpm_fail(ctx, TPM_ERR_INVALID,
                        "tree: depth limit (50) exceeded -- more than 32 particles too close to be split");
    TPM_TRY(grow(ctx, &ctx->nodes, &ctx->nodes_cap, tot[0]));
    TPM_TRY(grow(ctx, &ctx->buckets, &ctx->buckets_cap, tot[1]));
    tree_nodes_kernel<<<(unsigned)ncell, 64, 0, s>>>(C, ctx->cell_begin, ctx->tree_pos, ctx->node_begin,
                                                      ctx->bucket_begin, ctx->nodes, ctx->buckets, ctx->p.hsoft,
                                                      ctx->tree_flags);
    TPM_HIP(ctx, hipGetLastError());
    TPM_HIP(ctx, hipMemcpyAsync(&tot[2], ctx->tree_flags + 2, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    TPM_HIP(ctx, hipStreamSynchronize(s));
    if (tot[2]) return tpm_fail(ctx, TPM_ERR_HIP, "tree: node numbering disagrees with the partition counts");
    ctx->n_nodes = tot[0];
    ctx->n_buckets = tot[1];
    ctx->tree_epoch = ctx->chain_epoch;
    return TPM_OK;
}

tpm_status tpm_tree_kick(tpm_ctx* ctx, int min_rung, double scale, double t0, int first_call, unsigned flags,
                         tpm_stats* stats, const tpm_cell_box* box) {
    const bool use_pm = !(flags & TPM_SKIP_PM);
    if (!ctx->chain_valid || ctx->chain_C != ctx->C)
        return tpm_fail(ctx, TPM_ERR_STATE, "tpm_kick: chain mesh not built (call tpm_chainmesh_create)");
    if (use_pm && (!ctx->pm_valid || ctx->green_M != ctx->M))
        return tpm_fail(ctx, TPM_ERR_STATE, "tpm_kick: PM mesh not computed (call tpm_gravity_long_compute)");
    if (t0 <= 0.0) return tpm_fail(ctx, TPM_ERR_INVALID, "tpm_kick: t0 must be positive");
    TPM_TRY(tpm_tree_ensure(ctx));
    hipStream_t s = ctx->stream;
    if (!ctx->counters) TPM_HIP(ctx, hipMalloc(&ctx->counters, sizeof(unsigned long long) * TPM_COUNTER_WORDS));
    TPM_TRY(tpm_fill_u64(ctx, ctx->counters, TPM_COUNTER_WORDS));
    tree_fill_i32_kernel<<<1, 64, 0, s>>>(ctx->tree_flags, 4, 0);
    WalkArgs a{};
    a.pos = ctx->tree_pos;
    a.nodes = ctx->nodes;
    a.node_begin = ctx->node_begin;
    a.buckets = ctx->buckets;
    a.phi = use_pm ? ctx->phi : nullptr;
    a.flags_dev = ctx->tree_flags;
    a.o = tpm_kick_out{ctx->gx, ctx->gy, ctx->gz, ctx->pot, ctx->vx, ctx->vy, ctx->vz, ctx->rung, ctx->counters};
    tpm_box_fill(&a.B, ctx->p, min_rung, first_call, t0, scale);
    a.B.cell0 = 0;
    a.B.mesh0 = 0;
    a.nbuckets = (int)ctx->n_buckets;
    a.blist = nullptr;
    if (box) {
        // the buckets of the box's cells, from the bucket offsets per cell
        const int C = ctx->C;
        const int64_t ncell = (int64_t)C * C * C;
        std::vector<int32_t> bb((size_t)ncell + 1), list;
        TPM_HIP(ctx, hipMemcpyAsync(bb.data(), ctx->bucket_begin, sizeof(int32_t) * (ncell + 1),
                                    hipMemcpyDeviceToHost, s));
        TPM_HIP(ctx, hipStreamSynchronize(s));
        for (int z = box->lo[2]; z < box->hi[2]; z++)
            for (int y = box->lo[1]; y < box->hi[1]; y++)
                for (int x = box->lo[0]; x < box->hi[0]; x++) {
                    const int64_t c = ((int64_t)z * C + y) * C + x;
                    for (int32_t b = bb[c]; b < bb[c + 1]; b++) list.push_back(b);
                }
        if ((int64_t)list.size() > ctx->box_buckets_cap || !ctx->box_buckets) {
            if (ctx->box_buckets) TPM_HIP(ctx, hipFree(ctx->box_buckets));
            ctx->box_buckets = nullptr;
            TPM_HIP(ctx, hipMalloc(&ctx->box_buckets, sizeof(int32_t) * std::max<size_t>(list.size(), 1)));
            ctx->box_buckets_cap = (int64_t)std::max<size_t>(list.size(), 1);
        }
        if (!list.empty())
            TPM_HIP(ctx, hipMemcpyAsync(ctx->box_buckets, list.data(), sizeof(int32_t) * list.size(),
                                        hipMemcpyHostToDevice, s));
        // the host list must outlive the copy
        TPM_HIP(ctx, hipStreamSynchronize(s));
        a.blist = ctx->box_buckets;
        a.nbuckets = (int)list.size();
    }
    a.n = (int)ctx->n;
    a.per_xcd = (a.nbuckets + 7) / 8;
    // treepm_params (kick_treepm.cu:826-853): rcut, theta = 0.4, 1 / (2 rs)
    a.rcut = (float)ctx->p.rcut;
    const float theta = 0.4f;
    a.theta2inv = 1.0f / (theta * theta);
    a.thetainv = 1.0f / theta;
    a.inv2rs = 1.0f / (float)ctx->p.rs / 2.0f;
    a.e2 = (float)(-1.4426950408889634 / (4.0 * ctx->p.rs * ctx->p.rs));
    a.tsp = 2.0f / sqrtf((float)M_PI);
    a.flags = flags & (TPM_SKIP_PP | TPM_SKIP_PM | TPM_SKIP_KICK);
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
    if (ctx->n > 0 && a.per_xcd > 0) {
        tree_walk_kernel<<<(unsigned)(8 * (int64_t)a.per_xcd), 64, 0, s>>>(a);
        TPM_HIP(ctx, hipGetLastError());
    }
    TPM_HIP(ctx, hipEventRecord(e1, s));
    int32_t fl[2] = {0, 0};
    TPM_HIP(ctx, hipMemcpyAsync(fl, ctx->tree_flags, sizeof(fl), hipMemcpyDeviceToHost, s));
    std::vector<unsigned long long> hw(TPM_COUNTER_WORDS, 0ull);
    if (stats) {
        TPM_TRY(tpm_candidate_pairs(ctx, ctx->counters + 2, box));
        TPM_HIP(ctx, hipMemcpyAsync(hw.data(), ctx->counters, sizeof(unsigned long long) * TPM_COUNTER_WORDS,
                                    hipMemcpyDeviceToHost, s));
    }
    TPM_HIP(ctx, hipStreamSynchronize(s));
    unsigned long long hc[4];
    tpm_reduce_counters(hw.data(), hc);
    if (fl[1])
        return tpm_fail(ctx, TPM_ERR_HIP, "tpm_kick: tree walk stopped (stack exhausted or a node met twice)");
    if (stats) {
        float ms = 0.f;
        TPM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
        stats->executed_pairs = (double)hc[0];
        stats->max_rung = (int)hc[1];
        stats->candidate_pairs = (double)hc[2];
        stats->pc_interactions = (double)hc[3];
        stats->ms_pp = ms;
        stats->n_tiles = a.nbuckets;
    }
    return TPM_OK;
}

extern "C" tpm_status tpm_tree_build(tpm_ctx* ctx) {
    if (!ctx) return TPM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    return tpm_tree_ensure(ctx);
}

extern "C" tpm_status tpm_tree_info(tpm_ctx* ctx, int64_t* n_nodes, int64_t* n_buckets) {
    if (!ctx) return TPM_ERR_INVALID;
    if (ctx->tree_epoch != ctx->chain_epoch || !ctx->chain_valid)
        return tpm_fail(ctx, TPM_ERR_STATE, "tpm_tree_info: trees not built for the current chain mesh");
    if (n_nodes) *n_nodes = ctx->n_nodes;
    if (n_buckets) *n_buckets = ctx->n_buckets;
    return TPM_OK;
}

extern "C" tpm_status tpm_tree_download(tpm_ctx* ctx, int64_t* node_begin, int64_t* bucket_begin,
                                        tpm_tree_node* nodes, tpm_tree_bucket* buckets, int32_t* order) {
    if (!ctx) return TPM_ERR_INVALID;
    if (ctx->tree_epoch != ctx->chain_epoch || !ctx->chain_valid)
        return tpm_fail(ctx, TPM_ERR_STATE, "tpm_tree_download: trees not built for the current chain mesh");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const int64_t ncell = (int64_t)ctx->C * ctx->C * ctx->C;
    std::vector<int32_t> nb((size_t)ncell + 1), bb((size_t)ncell + 1);
    TPM_HIP(ctx, hipMemcpyAsync(nb.data(), ctx->node_begin, sizeof(int32_t) * (ncell + 1), hipMemcpyDeviceToHost, s));
    TPM_HIP(ctx,
            hipMemcpyAsync(bb.data(), ctx->bucket_begin, sizeof(int32_t) * (ncell + 1), hipMemcpyDeviceToHost, s));
    if (nodes && ctx->n_nodes)
        TPM_HIP(ctx, hipMemcpyAsync(nodes, ctx->nodes, sizeof(tpm_tree_node) * ctx->n_nodes, hipMemcpyDeviceToHost, s));
    if (buckets && ctx->n_buckets)
        TPM_HIP(ctx, hipMemcpyAsync(buckets, ctx->buckets, sizeof(tpm_tree_bucket) * ctx->n_buckets,
                                    hipMemcpyDeviceToHost, s));
    int32_t *dorder = nullptr, *d2h = nullptr;
    if (order && ctx->n) {
        const bool mapped = ctx->host_order == TPM_ORDER_REFERENCE && !ctx->h2d_identity;
        TPM_HIP(ctx, hipMallocAsync((void**)&dorder, sizeof(int32_t) * ctx->n, s));
        if (mapped) {
            TPM_HIP(ctx, hipMallocAsync((void**)&d2h, sizeof(int32_t) * ctx->n, s));
            tree_invert_kernel<<<grid_for(ctx->n, 256), 256, 0, s>>>(ctx->n, ctx->h2d, d2h);
        }
        tree_order_kernel<<<grid_for(ctx->n, 256), 256, 0, s>>>(ctx->n, ctx->tree_pos, d2h, dorder);
        TPM_HIP(ctx, hipGetLastError());
        TPM_HIP(ctx, hipMemcpyAsync(order, dorder, sizeof(int32_t) * ctx->n, hipMemcpyDeviceToHost, s));
        TPM_HIP(ctx, hipFreeAsync(dorder, s));
        if (d2h) TPM_HIP(ctx, hipFreeAsync(d2h, s));
    }
    TPM_HIP(ctx, hipStreamSynchronize(s));
    for (int64_t c = 0; c <= ncell; c++) {
        if (node_begin) node_begin[c] = nb[c];
        if (bucket_begin) bucket_begin[c] = bb[c];
    }
    return TPM_OK;
}

extern "C" tpm_status tpm_tree_sort_particles(tpm_ctx* ctx) {
    if (!ctx) return TPM_ERR_INVALID;
    if (ctx->host_order != TPM_ORDER_REFERENCE)
        return tpm_fail(ctx, TPM_ERR_STATE, "tpm_tree_sort_particles: needs TPM_ORDER_REFERENCE");
    (void)hipSetDevice(ctx->device);
    TPM_TRY(tpm_tree_ensure(ctx));
    const int64_t n = ctx->n;
    if (n > 0) {
        if (!ctx->h2d || ctx->ref_cap < n)
            return tpm_fail(ctx, TPM_ERR_STATE, "tpm_tree_sort_particles: build the chain mesh first");
        tree_host_order_kernel<<<grid_for(n, 256), 256, 0, ctx->stream>>>(n, ctx->tree_pos, ctx->h2d);
        TPM_HIP(ctx, hipGetLastError());
    }
    ctx->h2d_identity = false;
    ctx->ref_valid = false;
    return TPM_OK;
}
