// tpm_batch.hip -- many independent periodic boxes resident on one device and
// kicked together (the tpm_batch_* C ABI).
//
// No reference counterpart: kick_treepm kicks one box per call and allocates,
// copies and frees per call (kick_treepm.cu:815-829, 1078-1082).  Here the
// boxes live in fixed arenas sized once at creation (the HBM budget):
//   particles  one SoA arena in box order, compacted by every step's
//              chain-mesh sort (particles of departed boxes sort into a trash
//              cell and drop off the end; new boxes are appended);
//   cells      the chain cells of all live boxes in one global cell space;
//   meshes     real-space regions (M^3) and spectrum regions (M^2 (M/2+1))
//              handed out by first-fit allocators for a box's lifetime.
// A step runs every stage once over all boxes, driven by a device table of
// per-box constants (tpm_box_dev): chain mesh, deposit, Green's multiply,
// the fused PP + interpolation + kick kernel over every (box, tile) item, and
// optionally the drift; only the FFTs are per box (one rocFFT plan pair per
// mesh size).  The kernels of a box see exactly the data and the order they
// would see alone, so every box's result equals its tpm_kick_treepm result
// bit for bit (same tile size).  With tuning.graph the step is captured once
// per (set of boxes, tuning, buffer parity) and replayed.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "tpm_internal.hpp"

tpm_status tpm_ensure_particles(tpm_ctx* ctx, int64_t n);

namespace {

constexpr int64_t MESH_ALIGN = 64;  // nodes (256 B of floats)
constexpr int64_t K_ALIGN = 32;     // modes (256 B of float2)

int64_t kmodes(int M) { return (int64_t)M * M * (M / 2 + 1); }
int64_t cells_of(int C) { return (int64_t)C * C * C; }
int64_t tiles_of(int64_t n, int C) { return n / 8 + cells_of(C) + 1; }  // any tile size

struct Caps {
    int64_t p, m, k, c, t, nbox;
};

Caps caps_of(const tpm_batch_limits& l) {
    Caps c;
    c.p = l.max_particles;
    c.m = l.max_mesh_nodes;
    c.nbox = l.max_boxes;
    // M^2 (M/2 + 1) <= 3/4 M^3 for M >= 4, plus alignment slack per box
    c.k = l.max_mesh_nodes / 4 * 3 + K_ALIGN * (l.max_boxes + 1);
    c.m += MESH_ALIGN * (l.max_boxes + 1);
    c.c = l.max_cells > 0 ? l.max_cells : l.max_mesh_nodes / 64 + 4096;
    c.t = c.p / 8 + c.c + l.max_boxes + 64;
    return c;
}

int64_t arena_bytes(const tpm_batch_limits& l) {
    const Caps c = caps_of(l);
    // positions, velocities (x2), rungs, g/phi, sort, mesh force
    const int64_t per_p = 12 * 2 + 12 * 2 + 2 + 16 + 16 + 16;
    const int64_t per_m = 8 + 4 + 4 + (l.interlace ? 4 : 0);  // Q32 density, float densities, phi
    const int64_t per_k = 8 + 4 + (l.interlace ? 8 : 0);  // spectrum, Green's, shifted spectrum
    const int64_t per_c = 4 * 4;                          // counts, cell_begin, tiles per cell, tile offsets
    return c.p * per_p + c.m * per_m + c.k * per_k + (c.c + 2) * per_c + c.t * 16 +
           ((c.c + 2) / 2048 + 2) * 4 + c.nbox * (int64_t)(sizeof(tpm_box_dev) + sizeof(tpm_kregion)) +
           (int64_t)sizeof(unsigned long long) * TPM_COUNTER_WORDS + 128;
}

template <class T>
tpm_status alloc(tpm_batch* bt, T** p, int64_t count) {
    TPM_HIP(bt, hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
    return TPM_OK;
}

tpm_batch_box* find_box(tpm_batch* bt, int id) {
    for (auto& b : bt->boxes)
        if (b.id == id && b.live) return &b;
    return nullptr;
}

// Resources the live boxes use now.
void usage(const tpm_batch* bt, int64_t* n_live, int64_t* cells, int64_t* tiles, int* nlive) {
    *n_live = *cells = *tiles = 0;
    *nlive = 0;
    for (const auto& b : bt->boxes) {
        if (!b.live) continue;
        *n_live += b.n;
        *cells += cells_of(b.C);
        *tiles += tiles_of(b.n, b.C);
        (*nlive)++;
    }
}

// The device table for the next step and the step's sizes.
tpm_status build_layout(tpm_batch* bt, tpm_batch_layout* L) {
    TPM_HIP(bt, hipEventSynchronize(bt->table_ev));  // previous upload done with the staging copy
    const int ts = bt->tune.tile_sinks;
    int64_t cell = 0, dep = 0, q = 0, tiles = 0;
    int lds = 0, nk = 0, mfw = 0;
    const int nb = (int)bt->boxes.size();
    for (int i = 0; i < nb; i++) {
        const tpm_batch_box& b = bt->boxes[i];
        tpm_box_dev& d = bt->h_boxes[i];
        std::memset(&d, 0, sizeof(d));
        tpm_box_fill(&d, b.spec.params, b.spec.min_rung, b.steps == 0, b.spec.t0, b.spec.scale);
        d.p0 = b.p0;
        d.n = b.n;
        d.q0 = q;
        d.live = b.live ? 1 : 0;
        d.cell0 = (int32_t)cell;
        d.dep0 = (int32_t)dep;
        if (b.live) {
            d.mesh0 = b.mesh0;
            d.k0 = b.k0;
            d.green0 = b.k0;
            d.B = b.B;
            d.nb = b.nb;
            cell += cells_of(b.C);
            dep += (int64_t)b.nb * b.nb * b.nb;
            q += b.n;
            tiles += b.n / ts + cells_of(b.C) + 1;
            lds = std::max(lds, b.wmax * b.wmax * b.wmax * 8);
            const int w = tpm_mesh_force_width(b.C, b.M);
            mfw = (w == 0 || mfw < 0) ? -1 : std::max(mfw, w);
            bt->h_kreg[nk++] = tpm_kregion{b.k0, kmodes(b.M), b.M, 0};
        }
    }
    std::sort(bt->h_kreg, bt->h_kreg + nk, [](const tpm_kregion& a, const tpm_kregion& b) { return a.k0 < b.k0; });
    if (cell > INT32_MAX / 64 - 1 || dep > INT32_MAX) return tpm_fail(bt, TPM_ERR_INVALID, "batch too large");
    L->nbox = nb;
    L->n_arena = bt->n_arena;
    L->n_live = q;
    L->ncell = (int32_t)cell;
    L->ndep = (int32_t)dep;
    L->tiles_max = tiles;
    L->mesh_hw = bt->mesh_alloc.high_water();
    L->k_hw = bt->k_alloc.high_water();
    L->nkreg = nk;
    L->order = bt->lim.order;
    L->interlace = bt->lim.interlace;
    L->lds = lds;
    L->mfw = mfw > 0 ? mfw : 0;
    if (nb > 0)
        TPM_HIP(bt, hipMemcpyAsync(bt->d_boxes, bt->h_boxes, sizeof(tpm_box_dev) * nb, hipMemcpyHostToDevice,
                                   bt->stream));
    if (nk > 0)
        TPM_HIP(bt, hipMemcpyAsync(bt->d_kreg, bt->h_kreg, sizeof(tpm_kregion) * nk, hipMemcpyHostToDevice,
                                   bt->stream));
    TPM_HIP(bt, hipEventRecord(bt->table_ev, bt->stream));
    return TPM_OK;
}

// The kernels of one step, in stream order (capturable: kernels and rocFFT
// executions only).
tpm_status enqueue_step(tpm_batch* bt, const tpm_batch_layout& L, unsigned flags) {
    auto mark = [&](int k) -> tpm_status {
        if (bt->stage_events) TPM_HIP(bt, hipEventRecord(bt->ev_stage[k], bt->stream));
        return TPM_OK;
    };
    TPM_TRY(tpm_batch_fill_u64(bt, bt->counters, TPM_COUNTER_WORDS));
    TPM_TRY(tpm_batch_chainmesh(bt, L));
    TPM_TRY(mark(0));
    if (!(flags & TPM_SKIP_PM)) TPM_TRY(tpm_batch_pm(bt, L));
    TPM_TRY(mark(1));
    TPM_TRY(tpm_batch_pp(bt, L, flags));
    TPM_TRY(mark(2));
    if (flags & TPM_BATCH_DRIFT) TPM_TRY(tpm_batch_drift(bt, L));
    return TPM_OK;
}

void swap_sets(tpm_batch* bt) {
    std::swap(bt->x, bt->x2);
    std::swap(bt->y, bt->y2);
    std::swap(bt->z, bt->z2);
    std::swap(bt->vx, bt->vx2);
    std::swap(bt->vy, bt->vy2);
    std::swap(bt->vz, bt->vz2);
    std::swap(bt->rung, bt->rung2);
}

void drop_graphs(tpm_batch* bt) {
    for (int k = 0; k < 2; k++) {
        if (bt->graph_exec[k]) (void)hipGraphExecDestroy(bt->graph_exec[k]);
        if (bt->graph[k]) (void)hipGraphDestroy(bt->graph[k]);
        bt->graph_exec[k] = nullptr;
        bt->graph[k] = nullptr;
        bt->graph_version[k] = 0;
    }
}

// Host bookkeeping after a step (or a compaction) was enqueued: departed
// boxes are gone from the arena, live boxes start where the sort put them.
void after_sort(tpm_batch* bt, bool advanced) {
    std::vector<tpm_batch_box> keep;
    keep.reserve(bt->boxes.size());
    int64_t q = 0;
    bool dropped = false;
    for (auto& b : bt->boxes) {
        if (!b.live) {
            dropped = true;
            continue;
        }
        b.p0 = q;
        q += b.n;
        if (advanced) b.steps++;
        keep.push_back(b);
    }
    bt->boxes.swap(keep);
    bt->n_arena = q;
    bt->parity ^= 1;
    if (dropped) bt->layout_version++;
}

tpm_status run_step(tpm_batch* bt, unsigned flags, tpm_batch_stats* st) {
    (void)hipSetDevice(bt->device);
    tpm_batch_layout L;
    TPM_TRY(build_layout(bt, &L));
    TPM_HIP(bt, hipEventRecord(bt->ev0, bt->stream));
    const int par = bt->parity;
    const uint64_t key = bt->layout_version * 64 + (uint64_t)flags;
    if (bt->tune.graph && L.n_live > 0) {
        if (bt->graph_version[par] != key) {
            if (bt->graph_exec[par]) TPM_HIP(bt, hipGraphExecDestroy(bt->graph_exec[par]));
            if (bt->graph[par]) TPM_HIP(bt, hipGraphDestroy(bt->graph[par]));
            bt->graph_exec[par] = nullptr;
            bt->graph[par] = nullptr;
            bt->graph_version[par] = 0;
            TPM_HIP(bt, hipStreamBeginCapture(bt->stream, hipStreamCaptureModeRelaxed));
            const tpm_status s = enqueue_step(bt, L, flags);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(bt->stream, &g);
            if (s != TPM_OK) {
                if (g) (void)hipGraphDestroy(g);
                return s;
            }
            TPM_HIPE(bt->err, e);
            bt->graph[par] = g;
            TPM_HIP(bt, hipGraphInstantiate(&bt->graph_exec[par], g, nullptr, nullptr, 0));
            bt->graph_version[par] = key;
            swap_sets(bt);  // enqueue_step swapped the sets: undo, the replay below does it again
        }
        TPM_HIP(bt, hipGraphLaunch(bt->graph_exec[par], bt->stream));
        swap_sets(bt);
    } else {
        bt->stage_events = st != nullptr;
        const tpm_status s = enqueue_step(bt, L, flags);
        bt->stage_events = false;
        TPM_TRY(s);
    }
    TPM_HIP(bt, hipEventRecord(bt->ev1, bt->stream));
    const bool staged = st != nullptr && !(bt->tune.graph && L.n_live > 0);
    after_sort(bt, !(flags & TPM_SKIP_KICK));
    if (st) {
        TPM_TRY(tpm_batch_candidates(bt, L));
        std::vector<unsigned long long> hw(TPM_COUNTER_WORDS);
        TPM_HIP(bt, hipMemcpyAsync(hw.data(), bt->counters, sizeof(unsigned long long) * TPM_COUNTER_WORDS,
                                   hipMemcpyDeviceToHost, bt->stream));
        TPM_HIP(bt, hipStreamSynchronize(bt->stream));
        unsigned long long hc[4];
        tpm_reduce_counters(hw.data(), hc);
        float ms = 0.f;
        TPM_HIP(bt, hipEventElapsedTime(&ms, bt->ev0, bt->ev1));
        st->executed_pairs = (double)hc[0];
        st->max_rung = (int32_t)hc[1];
        st->candidate_pairs = (double)hc[2];
        st->ms_step = ms;
        st->ms_chainmesh = st->ms_pm = st->ms_pp = -1.0;
        if (staged) {
            float a = 0.f, b = 0.f, c = 0.f;
            TPM_HIP(bt, hipEventElapsedTime(&a, bt->ev0, bt->ev_stage[0]));
            TPM_HIP(bt, hipEventElapsedTime(&b, bt->ev_stage[0], bt->ev_stage[1]));
            TPM_HIP(bt, hipEventElapsedTime(&c, bt->ev_stage[1], bt->ev_stage[2]));
            st->ms_chainmesh = a;
            st->ms_pm = b;
            st->ms_pp = c;
        }
        st->n_particles = L.n_live;
        int nl = 0;
        for (const auto& b : bt->boxes) nl += b.live ? 1 : 0;
        st->n_boxes = nl;
    }
    return TPM_OK;
}

// Sorts the arena without kicking (drops departed boxes' particles).
tpm_status compact(tpm_batch* bt) {
    tpm_batch_layout L;
    TPM_TRY(build_layout(bt, &L));
    TPM_TRY(tpm_batch_chainmesh(bt, L));
    after_sort(bt, false);
    return TPM_OK;
}

// Common part of tpm_batch_add*: admission, arena regions, Green's table,
// then the particle copy (kind: H2D or D2D).
tpm_status add_box(tpm_batch* bt, const tpm_box_spec* spec_in, int64_t n, const uint32_t* x, const uint32_t* y,
                   const uint32_t* z, const float* vx, const float* vy, const float* vz, const int8_t* rung,
                   hipMemcpyKind kind, int* id) {
    if (!spec_in || !id || n < 1 || !x || !y || !z) return tpm_fail(bt, TPM_ERR_INVALID, "tpm_batch_add: bad arguments");
    (void)hipSetDevice(bt->device);
    tpm_box_spec spec = *spec_in;
    TPM_TRY(tpm_check_params(bt->err, &spec.params));
    const tpm_params& p = spec.params;
    if (p.order != bt->lim.order || p.interlace != bt->lim.interlace)
        return tpm_fail(bt, TPM_ERR_INVALID, "tpm_batch_add: order and interlace must match the batch limits");
    if (!(spec.t0 > 0.0) || !(spec.scale > 0.0))
        return tpm_fail(bt, TPM_ERR_INVALID, "tpm_batch_add: t0 and scale must be positive");
    const int C = p.chain_dim, M = p.mesh_dim;
    int B = 0, wmax = 0;
    tpm_deposit_blocking(C, M, p.interlace, bt->tune.deposit_block, &B, &wmax);
    if (B < 1) return tpm_fail(bt, TPM_ERR_INVALID, "tpm_batch_add: a chain cell's mesh footprint exceeds 64 KiB of LDS");
    int64_t n_live, cells, tiles;
    int nlive;
    usage(bt, &n_live, &cells, &tiles, &nlive);
    if (n_live + n > bt->pcap || cells + cells_of(C) > bt->ccap || tiles + tiles_of(n, C) > bt->tcap ||
        (int64_t)bt->boxes.size() + 1 > bt->lim.max_boxes)
        return tpm_fail(bt, TPM_ERR_BUDGET, "tpm_batch_add: particle, cell or box capacity exhausted");
    const int64_t m0 = bt->mesh_alloc.alloc((int64_t)M * M * M, MESH_ALIGN);
    if (m0 < 0) return tpm_fail(bt, TPM_ERR_BUDGET, "tpm_batch_add: mesh arena exhausted");
    const int64_t k0 = bt->k_alloc.alloc(kmodes(M), K_ALIGN);
    if (k0 < 0) {
        bt->mesh_alloc.release(m0, (int64_t)M * M * M);
        return tpm_fail(bt, TPM_ERR_BUDGET, "tpm_batch_add: spectrum arena exhausted");
    }
    auto undo = [&](tpm_status s) {
        bt->mesh_alloc.release(m0, (int64_t)M * M * M);
        bt->k_alloc.release(k0, kmodes(M));
        return s;
    };
    if (bt->n_arena + n > bt->pcap) {
        const tpm_status s = compact(bt);
        if (s != TPM_OK) return undo(s);
    }
    // FFT plans of this size (and a work buffer large enough) before any
    // step is captured; the Green's table of the box.
    tpm_status s = tpm_fft_use(bt->err, bt->fft, bt->device, bt->stream, M);
    for (int k = 0; k < tpm_batch::NSIDE && s == TPM_OK; k++)
        s = tpm_fft_use(bt->err, bt->side_fft[k], bt->device, bt->side[k], M);
    if (s != TPM_OK) return undo(s);
    const double rs_eff = p.long_type == TPM_LONG_PM ? 0.0 : p.rs;
    s = tpm_green_table(bt->err, bt->stream, M, rs_eff, p.order, bt->green + k0);
    if (s != TPM_OK) return undo(s);
    const int64_t o = bt->n_arena;
    hipStream_t st = bt->stream;
#define TPM_BT(call)                                          \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) {                               \
            bt->err = std::string(#call) + ": " + hipGetErrorString(e_); \
            return undo(TPM_ERR_HIP);                         \
        }                                                     \
    } while (0)
    TPM_BT(hipMemcpyAsync(bt->x + o, x, 4 * n, kind, st));
    TPM_BT(hipMemcpyAsync(bt->y + o, y, 4 * n, kind, st));
    TPM_BT(hipMemcpyAsync(bt->z + o, z, 4 * n, kind, st));
    const float* vs[3] = {vx, vy, vz};
    float* dv[3] = {bt->vx + o, bt->vy + o, bt->vz + o};
    for (int d = 0; d < 3; d++) {
        if (vs[d])
            TPM_BT(hipMemcpyAsync(dv[d], vs[d], 4 * n, kind, st));
        else
            TPM_BT(hipMemsetAsync(dv[d], 0, 4 * n, st));
    }
    if (rung)
        TPM_BT(hipMemcpyAsync(bt->rung + o, rung, n, kind, st));
    else
        TPM_BT(hipMemsetAsync(bt->rung + o, 0, n, st));
    if (kind == hipMemcpyHostToDevice) TPM_BT(hipStreamSynchronize(st));  // pageable sources
    TPM_BT(hipEventRecord(bt->copy_ev, st));
#undef TPM_BT
    tpm_batch_box b;
    b.id = bt->next_id++;
    b.spec = spec;
    b.n = n;
    b.p0 = o;
    b.C = C;
    b.M = M;
    b.mesh0 = m0;
    b.k0 = k0;
    b.B = B;
    b.nb = (C + B - 1) / B;
    b.wmax = wmax;
    bt->boxes.push_back(b);
    bt->n_arena += n;
    bt->layout_version++;
    *id = b.id;
    return TPM_OK;
}

}  // namespace

extern "C" tpm_status tpm_batch_arena_bytes(const tpm_batch_limits* l, int64_t* bytes) {
    if (!l || !bytes) return TPM_ERR_INVALID;
    *bytes = arena_bytes(*l);
    return TPM_OK;
}

extern "C" tpm_status tpm_batch_create(int device, const tpm_batch_limits* l, tpm_batch** out) {
    if (!l || !out) return TPM_ERR_INVALID;
    *out = nullptr;
    if (l->max_particles < 1 || l->max_particles > (int64_t)INT32_MAX - 1024 || l->max_mesh_nodes < 64 ||
        l->max_boxes < 1 || l->max_boxes > 65536 || l->order < 2 || l->order > 4 ||
        (l->interlace != 0 && l->interlace != 1) || l->max_cells < 0)
        return TPM_ERR_INVALID;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return TPM_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return TPM_ERR_HIP;
    // rocFFT is set up by the first context
    tpm_ctx* probe = nullptr;
    tpm_status s = tpm_ctx_create(device, &probe);
    if (s != TPM_OK) return s;
    tpm_batch* bt = new tpm_batch();
    bt->scratch = probe;
    bt->device = device;
    bt->lim = *l;
    const Caps c = caps_of(*l);
    bt->pcap = c.p;
    bt->mcap = c.m;
    bt->kcap = c.k;
    bt->ccap = c.c;
    bt->tcap = c.t;
    if (c.c + 1 > INT32_MAX / 64) {
        tpm_batch_destroy(bt);
        return TPM_ERR_INVALID;
    }
    auto fail = [&](tpm_status st) {
        tpm_batch_destroy(bt);
        return st;
    };
    if (hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking) != hipSuccess) return fail(TPM_ERR_HIP);
    const int64_t P = c.p;
    if (alloc(bt, &bt->x, P) || alloc(bt, &bt->y, P) || alloc(bt, &bt->z, P) || alloc(bt, &bt->x2, P) ||
        alloc(bt, &bt->y2, P) || alloc(bt, &bt->z2, P) || alloc(bt, &bt->vx, P) || alloc(bt, &bt->vy, P) ||
        alloc(bt, &bt->vz, P) || alloc(bt, &bt->vx2, P) || alloc(bt, &bt->vy2, P) || alloc(bt, &bt->vz2, P) ||
        alloc(bt, &bt->rung, P) || alloc(bt, &bt->rung2, P) || alloc(bt, &bt->gx, P) || alloc(bt, &bt->gy, P) ||
        alloc(bt, &bt->gz, P) || alloc(bt, &bt->pot, P) || alloc(bt, &bt->key, P) || alloc(bt, &bt->slot, P) ||
        alloc(bt, &bt->tmp, P) || alloc(bt, &bt->perm, P) || alloc(bt, &bt->cnt, c.c + 2) ||
        alloc(bt, &bt->cell_begin, c.c + 2) || alloc(bt, &bt->tile_cnt, c.c + 2) ||
        alloc(bt, &bt->tile_off, c.c + 2) || alloc(bt, &bt->scan_part, (c.c + 2) / 2048 + 2) ||
        alloc(bt, &bt->tiles, c.t) || alloc(bt, &bt->rho_q, c.m) || alloc(bt, &bt->rho, c.m) ||
        alloc(bt, &bt->phi, c.m) || alloc(bt, &bt->rho_hat, c.k) || alloc(bt, &bt->green, c.k) ||
        (l->interlace && alloc(bt, &bt->rho_hat1, c.k)) || alloc(bt, &bt->d_boxes, l->max_boxes) ||
        alloc(bt, &bt->d_kreg, l->max_boxes) || alloc(bt, &bt->counters, TPM_COUNTER_WORDS) ||
        alloc(bt, &bt->moment_buf, 8 * TPM_MOMENT_BLOCKS + 8) || alloc(bt, &bt->mesh_force, P))
        return fail(TPM_ERR_BUDGET);
    if (l->interlace && alloc(bt, &bt->rho1, c.m)) return fail(TPM_ERR_BUDGET);
    for (int k = 0; k < tpm_batch::NSIDE; k++)
        if (hipStreamCreateWithFlags(&bt->side[k], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&bt->ev_join[k], hipEventDisableTiming) != hipSuccess)
            return fail(TPM_ERR_HIP);
    if (hipEventCreateWithFlags(&bt->ev_fork, hipEventDisableTiming) != hipSuccess) return fail(TPM_ERR_HIP);
    if (hipHostMalloc(reinterpret_cast<void**>(&bt->h_boxes), sizeof(tpm_box_dev) * l->max_boxes, 0) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&bt->h_kreg), sizeof(tpm_kregion) * l->max_boxes, 0) != hipSuccess)
        return fail(TPM_ERR_HIP);
    if (hipEventCreate(&bt->ev0) != hipSuccess || hipEventCreate(&bt->ev1) != hipSuccess ||
        hipEventCreate(&bt->ev_stage[0]) != hipSuccess || hipEventCreate(&bt->ev_stage[1]) != hipSuccess ||
        hipEventCreate(&bt->ev_stage[2]) != hipSuccess ||
        hipEventCreateWithFlags(&bt->table_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&bt->copy_ev, hipEventDisableTiming) != hipSuccess)
        return fail(TPM_ERR_HIP);
    if (hipEventRecord(bt->table_ev, bt->stream) != hipSuccess || hipEventRecord(bt->copy_ev, bt->stream) != hipSuccess)
        return fail(TPM_ERR_HIP);
    bt->mesh_alloc.reset(c.m);
    bt->k_alloc.reset(c.k);
    bt->persist_grid = tpm_pp_persistent_grid(device, bt->tune.tile_sinks, bt->tune.pp_lists != 0);
    *out = bt;
    return TPM_OK;
}

extern "C" tpm_status tpm_batch_destroy(tpm_batch* bt) {
    if (!bt) return TPM_OK;
    (void)hipSetDevice(bt->device);
    if (bt->stream) (void)hipStreamSynchronize(bt->stream);
    for (int k = 0; k < tpm_batch::NSIDE; k++)
        if (bt->side[k]) (void)hipStreamSynchronize(bt->side[k]);
    drop_graphs(bt);
    tpm_fft_free(bt->fft);
    for (int k = 0; k < tpm_batch::NSIDE; k++) {
        tpm_fft_free(bt->side_fft[k]);
        if (bt->side[k]) (void)hipStreamDestroy(bt->side[k]);
        if (bt->ev_join[k]) (void)hipEventDestroy(bt->ev_join[k]);
    }
    if (bt->ev_fork) (void)hipEventDestroy(bt->ev_fork);
    if (bt->rho1) (void)hipFree(bt->rho1);
    void* bufs[] = {bt->x,     bt->y,        bt->z,        bt->x2,       bt->y2,     bt->z2,      bt->vx,
                    bt->vy,    bt->vz,       bt->vx2,      bt->vy2,      bt->vz2,    bt->rung,    bt->rung2,
                    bt->gx,    bt->gy,       bt->gz,       bt->pot,      bt->key,    bt->slot,    bt->tmp,
                    bt->perm,  bt->cnt,      bt->cell_begin, bt->tile_cnt, bt->tile_off, bt->scan_part, bt->tiles,
                    bt->rho_q, bt->rho,      bt->phi,      bt->rho_hat,  bt->rho_hat1, bt->green,  bt->d_boxes,
                    bt->d_kreg, bt->counters, bt->moment_buf, bt->mesh_force};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (bt->h_boxes) (void)hipHostFree(bt->h_boxes);
    if (bt->h_kreg) (void)hipHostFree(bt->h_kreg);
    hipEvent_t evs[] = {bt->ev0, bt->ev1, bt->table_ev, bt->copy_ev, bt->ev_stage[0], bt->ev_stage[1],
                        bt->ev_stage[2]};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    if (bt->stream) (void)hipStreamDestroy(bt->stream);
    if (bt->scratch) tpm_ctx_destroy(bt->scratch);
    delete bt;
    return TPM_OK;
}

extern "C" const char* tpm_batch_last_error(const tpm_batch* bt) { return bt ? bt->err.c_str() : "null batch"; }

extern "C" tpm_status tpm_batch_fits(tpm_batch* bt, const tpm_box_spec* spec, int64_t n, int* fits) {
    if (!bt || !spec || !fits) return TPM_ERR_INVALID;
    tpm_params p = spec->params;
    TPM_TRY(tpm_check_params(bt->err, &p));
    int64_t n_live, cells, tiles;
    int nlive;
    usage(bt, &n_live, &cells, &tiles, &nlive);
    const int M = p.mesh_dim;
    bool ok = n_live + n <= bt->pcap && cells + cells_of(p.chain_dim) <= bt->ccap &&
              tiles + tiles_of(n, p.chain_dim) <= bt->tcap && (int64_t)bt->boxes.size() + 1 <= bt->lim.max_boxes;
    if (ok) {
        // trial allocation on copies of the allocators
        tpm_range_alloc m = bt->mesh_alloc, k = bt->k_alloc;
        ok = m.alloc((int64_t)M * M * M, MESH_ALIGN) >= 0 && k.alloc(kmodes(M), K_ALIGN) >= 0;
    }
    *fits = ok ? 1 : 0;
    return TPM_OK;
}

extern "C" tpm_status tpm_batch_add(tpm_batch* bt, const tpm_box_spec* spec, int64_t n, const uint32_t* x,
                                    const uint32_t* y, const uint32_t* z, const float* vx, const float* vy,
                                    const float* vz, const int8_t* rung, int* id) {
    if (!bt) return TPM_ERR_INVALID;
    return add_box(bt, spec, n, x, y, z, vx, vy, vz, rung, hipMemcpyHostToDevice, id);
}

extern "C" tpm_status tpm_batch_add_zeldovich(tpm_batch* bt, const tpm_box_spec* spec, int n_side, uint64_t seed,
                                              double rms_disp, double spectral_index, int* id) {
    if (!bt || !spec) return TPM_ERR_INVALID;
    if (n_side < 2 || n_side > 1290) return tpm_fail(bt, TPM_ERR_INVALID, "tpm_batch_add_zeldovich: bad n_side");
    (void)hipSetDevice(bt->device);
    int fits = 0;
    TPM_TRY(tpm_batch_fits(bt, spec, (int64_t)n_side * n_side * n_side, &fits));
    if (!fits) return tpm_fail(bt, TPM_ERR_BUDGET, "tpm_batch_add_zeldovich: the box does not fit the arenas");
    tpm_ctx* sc = bt->scratch;
    // the scratch buffers are free once the previous box's copy has run
    TPM_HIP(bt, hipStreamWaitEvent(sc->stream, bt->copy_ev, 0));
    tpm_status s = tpm_ic_zeldovich_run(sc, n_side, seed, rms_disp, spectral_index);
    if (s != TPM_OK) return tpm_fail(bt, s, sc->err);
    return add_box(bt, spec, sc->n, sc->x, sc->y, sc->z, sc->vx, sc->vy, sc->vz, sc->rung, hipMemcpyDeviceToDevice,
                   id);
}

extern "C" tpm_status tpm_batch_prepare(tpm_batch* bt, const tpm_box_spec* spec, int n_side) {
    if (!bt || !spec || n_side < 0) return TPM_ERR_INVALID;
    (void)hipSetDevice(bt->device);
    tpm_params p = spec->params;
    TPM_TRY(tpm_check_params(bt->err, &p));
    TPM_TRY(tpm_fft_use(bt->err, bt->fft, bt->device, bt->stream, p.mesh_dim));
    for (int k = 0; k < tpm_batch::NSIDE; k++)
        TPM_TRY(tpm_fft_use(bt->err, bt->side_fft[k], bt->device, bt->side[k], p.mesh_dim));
    if (n_side > 0) {
        tpm_ctx* sc = bt->scratch;
        TPM_HIP(bt, hipStreamWaitEvent(sc->stream, bt->copy_ev, 0));
        tpm_status s = tpm_ensure_particles(sc, (int64_t)n_side * n_side * n_side);
        if (s == TPM_OK) s = tpm_ensure_mesh(sc, n_side);
        if (s != TPM_OK) return tpm_fail(bt, s, sc->err);
    }
    return TPM_OK;
}

extern "C" tpm_status tpm_batch_remove(tpm_batch* bt, int id) {
    if (!bt) return TPM_ERR_INVALID;
    tpm_batch_box* b = find_box(bt, id);
    if (!b) return tpm_fail(bt, TPM_ERR_INVALID, "tpm_batch_remove: no such box");
    // regions are reused by later boxes; their writes are stream-ordered
    // after every step that read them
    bt->mesh_alloc.release(b->mesh0, (int64_t)b->M * b->M * b->M);
    bt->k_alloc.release(b->k0, kmodes(b->M));
    b->live = false;
    bt->layout_version++;
    return TPM_OK;
}

extern "C" tpm_status tpm_batch_step(tpm_batch* bt, unsigned flags, tpm_batch_stats* st) {
    if (!bt) return TPM_ERR_INVALID;
    if (flags & ~(TPM_SKIP_PP | TPM_SKIP_PM | TPM_SKIP_KICK | TPM_BATCH_DRIFT))
        return tpm_fail(bt, TPM_ERR_INVALID, "tpm_batch_step: unknown flags");
    return run_step(bt, flags, st);
}

extern "C" tpm_status tpm_batch_sync(tpm_batch* bt) {
    if (!bt) return TPM_ERR_INVALID;
    TPM_HIP(bt, hipStreamSynchronize(bt->stream));
    return TPM_OK;
}

extern "C" tpm_status tpm_batch_box_info(tpm_batch* bt, int id, int64_t* n, int* steps) {
    if (!bt) return TPM_ERR_INVALID;
    tpm_batch_box* b = find_box(bt, id);
    if (!b) return tpm_fail(bt, TPM_ERR_INVALID, "tpm_batch_box_info: no such box");
    if (n) *n = b->n;
    if (steps) *steps = b->steps;
    return TPM_OK;
}

extern "C" tpm_status tpm_batch_download(tpm_batch* bt, int id, uint32_t* x, uint32_t* y, uint32_t* z, float* vx,
                                         float* vy, float* vz, int8_t* rung, float* gx, float* gy, float* gz,
                                         float* pot) {
    if (!bt) return TPM_ERR_INVALID;
    tpm_batch_box* b = find_box(bt, id);
    if (!b) return tpm_fail(bt, TPM_ERR_INVALID, "tpm_batch_download: no such box");
    (void)hipSetDevice(bt->device);
    const int64_t o = b->p0, n = b->n;
    hipStream_t s = bt->stream;
    struct {
        void* dst;
        const void* src;
        size_t bytes;
    } cp[] = {{x, bt->x + o, 4 * (size_t)n},    {y, bt->y + o, 4 * (size_t)n},   {z, bt->z + o, 4 * (size_t)n},
              {vx, bt->vx + o, 4 * (size_t)n},  {vy, bt->vy + o, 4 * (size_t)n}, {vz, bt->vz + o, 4 * (size_t)n},
              {rung, bt->rung + o, (size_t)n},  {gx, bt->gx + o, 4 * (size_t)n}, {gy, bt->gy + o, 4 * (size_t)n},
              {gz, bt->gz + o, 4 * (size_t)n},  {pot, bt->pot + o, 4 * (size_t)n}};
    for (auto& c : cp)
        if (c.dst) TPM_HIP(bt, hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, s));
    TPM_HIP(bt, hipStreamSynchronize(s));
    return TPM_OK;
}

extern "C" tpm_status tpm_batch_moments(tpm_batch* bt, int id, tpm_moments* out) {
    if (!bt || !out) return TPM_ERR_INVALID;
    tpm_batch_box* b = find_box(bt, id);
    if (!b) return tpm_fail(bt, TPM_ERR_INVALID, "tpm_batch_moments: no such box");
    (void)hipSetDevice(bt->device);
    const int64_t o = b->p0, n = b->n;
    double* res = bt->moment_buf + 8 * TPM_MOMENT_BLOCKS;
    TPM_TRY(tpm_moments_enqueue(bt->err, bt->stream, n, bt->vx + o, bt->vy + o, bt->vz + o, bt->gx + o, bt->gy + o,
                                bt->gz + o, bt->pot + o, b->spec.params.GM, bt->moment_buf, res));
    double h[8];
    TPM_HIP(bt, hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, bt->stream));
    TPM_HIP(bt, hipStreamSynchronize(bt->stream));
    for (int d = 0; d < 3; d++) {
        out->momentum[d] = h[d];
        out->force[d] = h[3 + d];
    }
    out->kinetic = h[6];
    out->potential = h[7];
    return TPM_OK;
}

extern "C" tpm_status tpm_batch_to_ctx(tpm_batch* bt, int id, tpm_ctx* ctx) {
    if (!bt || !ctx) return TPM_ERR_INVALID;
    tpm_batch_box* b = find_box(bt, id);
    if (!b) return tpm_fail(bt, TPM_ERR_INVALID, "tpm_batch_to_ctx: no such box");
    if (ctx->device != bt->device) return tpm_fail(bt, TPM_ERR_INVALID, "tpm_batch_to_ctx: different device");
    (void)hipSetDevice(bt->device);
    tpm_status s = tpm_set_params(ctx, &b->spec.params);
    if (s == TPM_OK) s = tpm_ensure_particles(ctx, b->n);
    if (s != TPM_OK) return tpm_fail(bt, s, ctx->err);
    const int64_t o = b->p0, n = b->n;
    hipStream_t st = bt->stream;
    TPM_HIP(bt, hipStreamSynchronize(ctx->stream));
    TPM_HIP(bt, hipMemcpyAsync(ctx->x, bt->x + o, 4 * n, hipMemcpyDeviceToDevice, st));
    TPM_HIP(bt, hipMemcpyAsync(ctx->y, bt->y + o, 4 * n, hipMemcpyDeviceToDevice, st));
    TPM_HIP(bt, hipMemcpyAsync(ctx->z, bt->z + o, 4 * n, hipMemcpyDeviceToDevice, st));
    TPM_HIP(bt, hipMemcpyAsync(ctx->vx, bt->vx + o, 4 * n, hipMemcpyDeviceToDevice, st));
    TPM_HIP(bt, hipMemcpyAsync(ctx->vy, bt->vy + o, 4 * n, hipMemcpyDeviceToDevice, st));
    TPM_HIP(bt, hipMemcpyAsync(ctx->vz, bt->vz + o, 4 * n, hipMemcpyDeviceToDevice, st));
    TPM_HIP(bt, hipMemcpyAsync(ctx->rung, bt->rung + o, n, hipMemcpyDeviceToDevice, st));
    TPM_HIP(bt, hipMemcpyAsync(ctx->gx, bt->gx + o, 4 * n, hipMemcpyDeviceToDevice, st));
    TPM_HIP(bt, hipMemcpyAsync(ctx->gy, bt->gy + o, 4 * n, hipMemcpyDeviceToDevice, st));
    TPM_HIP(bt, hipMemcpyAsync(ctx->gz, bt->gz + o, 4 * n, hipMemcpyDeviceToDevice, st));
    TPM_HIP(bt, hipMemcpyAsync(ctx->pot, bt->pot + o, 4 * n, hipMemcpyDeviceToDevice, st));
    TPM_HIP(bt, hipStreamSynchronize(st));
    ctx->n = n;
    ctx->chain_valid = false;
    ctx->h2d_identity = true;
    ctx->ref_valid = false;
    ctx->pm_valid = false;
    return TPM_OK;
}

extern "C" tpm_status tpm_batch_set_tuning(tpm_batch* bt, const tpm_batch_tuning* t) {
    if (!bt || !t) return TPM_ERR_INVALID;
    if (t->tile_sinks != 8 && t->tile_sinks != 16 && t->tile_sinks != 32 && t->tile_sinks != 64)
        return tpm_fail(bt, TPM_ERR_INVALID, "tile_sinks must be 8, 16, 32 or 64");
    if (t->deposit_block < 0 || t->deposit_block > 8)
        return tpm_fail(bt, TPM_ERR_INVALID, "deposit_block must be in [0, 8]");
    if (t->sched != 0 && t->sched != 1) return tpm_fail(bt, TPM_ERR_INVALID, "sched must be 0 or 1");
    if (t->graph != 0 && t->graph != 1) return tpm_fail(bt, TPM_ERR_INVALID, "graph must be 0 or 1");
    if (t->pp_lists < -1 || t->pp_lists > 1) return tpm_fail(bt, TPM_ERR_INVALID, "pp_lists must be -1, 0 or 1");
    (void)hipSetDevice(bt->device);
    if (t->deposit_block != bt->tune.deposit_block) {
        // re-block every resident box; keep the old blocking if one does not fit
        std::vector<std::pair<int, int>> nb(bt->boxes.size());
        for (size_t i = 0; i < bt->boxes.size(); i++) {
            const auto& b = bt->boxes[i];
            tpm_deposit_blocking(b.C, b.M, bt->lim.interlace, t->deposit_block, &nb[i].first, &nb[i].second);
            if (nb[i].first < 1) return tpm_fail(bt, TPM_ERR_INVALID, "deposit_block does not fit a resident box");
        }
        for (size_t i = 0; i < bt->boxes.size(); i++) {
            auto& b = bt->boxes[i];
            b.B = nb[i].first;
            b.wmax = nb[i].second;
            b.nb = (b.C + b.B - 1) / b.B;
        }
    }
    const int lists = t->pp_lists < 0 ? TPM_PP_LISTS_DEFAULT : t->pp_lists;
    if (t->tile_sinks != bt->tune.tile_sinks || lists != bt->tune.pp_lists)
        bt->persist_grid = tpm_pp_persistent_grid(bt->device, t->tile_sinks, lists != 0);
    bt->tune = *t;
    bt->tune.pp_lists = lists;
    drop_graphs(bt);
    bt->layout_version++;
    return TPM_OK;
}

extern "C" tpm_status tpm_batch_get_tuning(const tpm_batch* bt, tpm_batch_tuning* t) {
    if (!bt || !t) return TPM_ERR_INVALID;
    *t = bt->tune;
    return TPM_OK;
}

// Autotuner.  The timed steps skip the kick epilogue and the drift, so the
// boxes do not advance (their chain-mesh order is a function of their
// positions and stays the same); the candidates run in the same process on
// the same boxes, so their times compare directly.
extern "C" tpm_status tpm_batch_autotune(tpm_batch* bt, int reps, tpm_batch_tuning* best, double* ms_best) {
    if (!bt || reps < 1) return TPM_ERR_INVALID;
    const tpm_batch_tuning keep = bt->tune;
    auto time_it = [&](const tpm_batch_tuning& t, double* ms) -> tpm_status {
        TPM_TRY(tpm_batch_set_tuning(bt, &t));
        TPM_TRY(run_step(bt, TPM_SKIP_KICK, nullptr));  // warm-up (and capture)
        double sum = 0.0;
        for (int r = 0; r < reps; r++) {
            tpm_batch_stats st{};
            TPM_TRY(run_step(bt, TPM_SKIP_KICK, &st));
            sum += st.ms_step;
        }
        *ms = sum / reps;
        return TPM_OK;
    };
    tpm_batch_tuning bestt = keep;
    double bms = 1e300;
    // (tile size, per-sink lists): lists only with 16-sink tiles
    const int tiles[5][2] = {{8, 0}, {16, 0}, {16, 1}, {32, 0}, {64, 0}};
    for (int ti = 0; ti < 5; ti++)
        for (int sc = 0; sc <= 1; sc++) {
            tpm_batch_tuning t = keep;
            t.tile_sinks = tiles[ti][0];
            t.pp_lists = tiles[ti][1];
            t.sched = sc;
            double ms = 0.0;
            TPM_TRY(time_it(t, &ms));
            if (ms < bms) {
                bms = ms;
                bestt = t;
            }
        }
    for (int db = 1; db <= 3; db++) {
        tpm_batch_tuning t = bestt;
        t.deposit_block = db;
        bool ok = true;
        for (const auto& b : bt->boxes) {
            int B, w;
            tpm_deposit_blocking(b.C, b.M, bt->lim.interlace, db, &B, &w);
            ok = ok && B >= 1;
        }
        if (!ok || db == bestt.deposit_block) continue;
        double ms = 0.0;
        TPM_TRY(time_it(t, &ms));
        if (ms < bms) {
            bms = ms;
            bestt = t;
        }
    }
    TPM_TRY(tpm_batch_set_tuning(bt, &bestt));
    if (best) *best = bestt;
    if (ms_best) *ms_best = bms;
    return TPM_OK;
}
