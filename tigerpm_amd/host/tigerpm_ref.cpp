// tigerpm_ref.cpp -- the reference's kick-path entry points over the C ABI.
//
// State mirrors the reference: module-global particle arrays (particles.cpp:
// 6-12), options (options.cpp:21), the chain-mesh cell table (chainmesh.cpp:
// 33) and the potential mesh (gravity_long.cpp:8-9).  The particles live on
// the device between calls; each entry point moves exactly the arrays the
// reference's version reads or writes.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>

#include "tigerpm_hip.h"
#include "tigerpm_ref.hpp"

std::array<std::vector<fixed32>, NDIM> particles_X;
std::array<std::vector<float>, NDIM> particles_U;
std::vector<char> particles_R;
std::vector<float> particles_P;
std::array<std::vector<float>, NDIM> particles_G;

namespace {

options g_opts;
bool g_opts_set = false;
tpm_ctx* g_ctx = nullptr;
std::vector<int64_t> g_cells;  // C^3 + 1 offsets of the last chain mesh
std::vector<float> g_phi;      // host copy of the potential mesh
bool g_phi_valid = false;
int g_long_type = GRAVITY_LONG_PME;
bool g_pm_computed = false;
kick_report g_report{};
// Trees of the current chain mesh on the host (tree_create): the device's
// forest, downloaded once per chain-mesh build.
struct host_forest {
    bool valid = false;
    std::vector<int64_t> node_begin, bucket_begin;
    std::vector<tpm_tree_node> nodes;
    std::vector<tpm_tree_bucket> buckets;
    std::vector<int32_t> order;     // tree position -> host index when built
    std::vector<char> cell_sorted;  // the host arrays of this cell are in tree order
};
host_forest g_forest;
// tree_create reordered host cells: the device's view of the host order
// must follow (tpm_tree_sort_particles) before the next download
bool g_tree_pending = false;
bool g_chain_valid = false;  // the device's chain mesh is of the current host particles

void check(tpm_status s, const char* what) {
    if (s != TPM_OK) {
        std::string msg = std::string(what) + " failed: " + (g_ctx ? tpm_last_error(g_ctx) : "no context");
        std::fprintf(stderr, "%s\n", msg.c_str());
        throw std::runtime_error(msg);
    }
}

tpm_ctx* ctx() {
    if (!g_ctx) {
        check(tpm_ctx_create(0, &g_ctx), "tpm_ctx_create");
        // the host arrays follow the reference partition's order: the device
        // keeps its own order and maps on download
        check(tpm_set_host_order(g_ctx, TPM_ORDER_REFERENCE), "tpm_set_host_order");
    }
    return g_ctx;
}

void derive(options& o) {
    o.parts_o_chain = o.parts_o_four * o.four_o_chain;
    o.four_dim = o.four_dim_override > 0 ? o.four_dim_override : o.parts_dim / o.parts_o_four;
    o.rs = (o.rs_cells > 0.0 ? o.rs_cells : 1.0) / o.four_dim;
    if (o.rcut_rs > 0.0) {
        o.rcut = o.rcut_rs * o.rs;
        o.chain_dim = (int)std::floor(1.0 / o.rcut);
    } else {
        if (o.parts_dim % o.parts_o_chain != 0) {
            std::fprintf(stderr, "Parts dim must be a multiple of %i\n", o.parts_o_chain);
            throw std::runtime_error("parts_dim not a multiple of parts_o_chain");
        }
        o.chain_dim = o.parts_dim / o.parts_o_chain;
        o.rcut = 1.0 / o.chain_dim;
    }
    const double n = std::pow((double)o.parts_dim, 3.0);
    o.GM = o.unit_mass ? 1.0 : 1.0 / n;
    o.hsoft = 2.0 / 25.0 / o.parts_dim;
    o.eta = 0.2 / std::sqrt(2.0);
}

tpm_params device_params(int long_type) {
    const options& o = get_options();
    tpm_params p{};
    p.mesh_dim = o.four_dim;
    p.chain_dim = o.chain_dim;
    p.order = o.order;
    p.do_phi = 1;
    p.long_type = long_type;
    p.interlace = o.interlace;
    p.rs = o.rs;
    p.rcut = o.rcut;
    p.hsoft = o.hsoft;
    p.GM = o.GM;
    p.eta = o.eta;
    p.n_total = (double)particles_size();
    return p;
}

void push_params(int long_type) {
    tpm_params p = device_params(long_type);
    check(tpm_set_params(ctx(), &p), "tpm_set_params");
    tpm_tuning t{get_options().tile_sinks, 0};
    check(tpm_set_tuning(ctx(), &t), "tpm_set_tuning");
}

void upload() {
    const int64_t n = particles_size();
    static_assert(sizeof(fixed32) == sizeof(uint32_t), "layout");
    check(tpm_particles_upload(ctx(), n, reinterpret_cast<const uint32_t*>(particles_X[0].data()),
                               reinterpret_cast<const uint32_t*>(particles_X[1].data()),
                               reinterpret_cast<const uint32_t*>(particles_X[2].data()), particles_U[0].data(),
                               particles_U[1].data(), particles_U[2].data(),
                               reinterpret_cast<const int8_t*>(particles_R.data())),
          "tpm_particles_upload");
    g_pm_computed = false;
    g_tree_pending = false;
    g_forest.valid = false;
    g_chain_valid = false;
}

// Positions, velocities and rungs back from the device (after a sort or IC).
void download_state(bool forces) {
    if (g_tree_pending) {
        // every cell into tree order, as tree_create leaves the ones it met
        check(tpm_tree_sort_particles(ctx()), "tpm_tree_sort_particles");
        g_tree_pending = false;
    }
    check(tpm_particles_download(ctx(), reinterpret_cast<uint32_t*>(particles_X[0].data()),
                                 reinterpret_cast<uint32_t*>(particles_X[1].data()),
                                 reinterpret_cast<uint32_t*>(particles_X[2].data()), particles_U[0].data(),
                                 particles_U[1].data(), particles_U[2].data(),
                                 reinterpret_cast<int8_t*>(particles_R.data()),
                                 forces ? particles_G[0].data() : nullptr, forces ? particles_G[1].data() : nullptr,
                                 forces ? particles_G[2].data() : nullptr, forces ? particles_P.data() : nullptr),
          "tpm_particles_download");
}

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// ------------------------------------------------------------------ options
bool process_options(int argc, char* argv[]) {
    options o;
    for (int a = 1; a < argc; a++) {
        std::string s = argv[a];
        if (s == "--help") {
            std::printf("options: --parts_dim --parts_o_four --four_o_chain --test --four_dim --rs_cells --rcut_rs "
                        "--order --mass=unit|1/N --ic=uniform|zeldovich --rms_disp --seed --tile_sinks --interlace "
                        "--short=direct|tree\n");
            return false;
        }
        const size_t eq = s.find('=');
        if (s.rfind("--", 0) != 0 || eq == std::string::npos) {
            std::fprintf(stderr, "unrecognised option %s\n", s.c_str());
            return false;
        }
        const std::string k = s.substr(2, eq - 2), v = s.substr(eq + 1);
        if (k == "parts_dim") o.parts_dim = std::stoi(v);
        else if (k == "parts_o_four") o.parts_o_four = std::stoi(v);
        else if (k == "four_o_chain") o.four_o_chain = std::stoi(v);
        else if (k == "test") o.test = v;
        else if (k == "box_size") o.box_size = std::stod(v);
        else if (k == "four_dim") o.four_dim_override = std::stoi(v);
        else if (k == "rs_cells") o.rs_cells = std::stod(v);
        else if (k == "rcut_rs") o.rcut_rs = std::stod(v);
        else if (k == "order") o.order = std::stoi(v);
        else if (k == "mass") o.unit_mass = (v == "unit");
        else if (k == "ic") o.ic = v;
        else if (k == "rms_disp") o.rms_disp = std::stod(v);
        else if (k == "seed") o.seed = std::stoull(v);
        else if (k == "tile_sinks") o.tile_sinks = std::stoi(v);
        else if (k == "interlace") o.interlace = std::stoi(v);
        else if (k == "short" && (v == "direct" || v == "tree")) o.short_range = v;
        else {
            std::fprintf(stderr, "unrecognised option %s\n", s.c_str());
            return false;
        }
    }
    set_options(o);
    const options& d = get_options();
    std::printf("Simulation Options\nparts_dim = %d\nfour_dim = %d\nchain_dim = %d\nrs = %e\nrcut = %e\nhsoft = %e\n"
                "order = %d\nshort = %s\ntest = %s\n",
                d.parts_dim, d.four_dim, d.chain_dim, d.rs, d.rcut, d.hsoft, d.order, d.short_range.c_str(),
                d.test.c_str());
    return true;
}

const options& get_options() {
    if (!g_opts_set) {
        derive(g_opts);
        g_opts_set = true;
    }
    return g_opts;
}

void set_options(const options& opts) {
    g_opts = opts;
    derive(g_opts);
    g_opts_set = true;
}

// ---------------------------------------------------------------- particles
int particles_size() { return (int)particles_X[0].size(); }
int particles_local_size() { return particles_size(); }

void particles_resize(size_t n) {
    for (int d = 0; d < NDIM; d++) {
        particles_X[d].resize(n);
        particles_U[d].resize(n);
        particles_G[d].resize(n);
    }
    particles_R.resize(n);
    particles_P.resize(n);
}

void particles_random_init() {
    const options& o = get_options();
    const int64_t n = (int64_t)o.parts_dim * o.parts_dim * o.parts_dim;
    particles_resize(n);
    check(tpm_ic_uniform(ctx(), n, o.seed), "tpm_ic_uniform");
    download_state(false);
    g_pm_computed = false;
    g_chain_valid = false;
    g_forest.valid = false;
}

void particles_domain_sort() {}

particle particles_get_particle(int i) {
    particle p;
    for (int d = 0; d < NDIM; d++) {
        p.x[d] = particles_X[d][i];
        p.v[d] = particles_U[d][i];
        p.g[d] = particles_G[d][i];
    }
    p.r = particles_R[i];
    p.p = particles_P[i];
    return p;
}

std::vector<particle> particles_sample(int nsamples) {
    std::mt19937_64 rng(get_options().seed + 1000);
    std::vector<particle> out;
    for (int s = 0; s < nsamples; s++) out.push_back(particles_get_particle((int)(rng() % particles_size())));
    return out;
}

std::array<int, NDIM> particles_mesh_loc(int index) {
    const double N = get_options().chain_dim;
    std::array<int, NDIM> i;
    for (int d = 0; d < NDIM; d++) i[d] = (int)(particles_X[d][index].to_double() * N);
    return i;
}

void initialize() {
    const options& o = get_options();
    const int64_t n = (int64_t)o.parts_dim * o.parts_dim * o.parts_dim;
    particles_resize(n);
    check(tpm_ic_zeldovich(ctx(), o.parts_dim, o.seed, o.rms_disp, -2.0), "tpm_ic_zeldovich");
    download_state(false);
    g_pm_computed = false;
    g_chain_valid = false;
    g_forest.valid = false;
}

// ---------------------------------------------------------------- chainmesh
void chainmesh_create() {
    push_params(g_long_type);
    upload();
    check(tpm_chainmesh_create(ctx()), "tpm_chainmesh_create");
    g_forest.valid = false;
    g_chain_valid = true;
    download_state(false);  // the reference sorts the host arrays in place
    const int C = get_options().chain_dim;
    g_cells.assign((size_t)C * C * C + 1, 0);
    check(tpm_chainmesh_cells(ctx(), g_cells.data()), "tpm_chainmesh_cells");
}

void chainmesh_exchange_begin() {}
void chainmesh_exchange_end() {}

range<int> chainmesh_interior_box() {
    range<int> r;
    const int C = get_options().chain_dim;
    r.begin = {0, 0, 0};
    r.end = {C, C, C};
    return r;
}

chaincell chainmesh_get(const std::array<int, NDIM>& i) {
    const int C = get_options().chain_dim;
    if (g_cells.empty()) throw std::runtime_error("chainmesh_get before chainmesh_create");
    const int ix = ((i[0] % C) + C) % C, iy = ((i[1] % C) + C) % C, iz = ((i[2] % C) + C) % C;
    const int64_t cid = ((int64_t)iz * C + iy) * C + ix;
    return chaincell{(int)g_cells[cid], (int)g_cells[cid + 1]};
}

// ------------------------------------------------------------- gravity_long
void gravity_long_compute(gravity_long_type type) {
    g_long_type = type;
    push_params(type);
    upload();
    check(tpm_gravity_long_compute(ctx()), "tpm_gravity_long_compute");
    g_pm_computed = true;
    g_phi_valid = false;
}

static const std::vector<float>& host_phi() {
    if (!g_phi_valid) {
        const int64_t M = get_options().four_dim;
        g_phi.resize(M * M * M);
        if (!g_pm_computed) {
            push_params(g_long_type);
            check(tpm_gravity_long_compute(ctx()), "tpm_gravity_long_compute");
            g_pm_computed = true;
        }
        check(tpm_gravity_long_get_phi(ctx(), g_phi.data()), "tpm_gravity_long_get_phi");
        g_phi_valid = true;
    }
    return g_phi;
}

std::vector<float> gravity_long_get_phi(const range<int>& box) {
    const auto& phi = host_phi();
    const int M = get_options().four_dim;
    std::vector<float> out(box.volume());
    std::array<int, NDIM> i;
    for (i[0] = box.begin[0]; i[0] < box.end[0]; i[0]++)
        for (i[1] = box.begin[1]; i[1] < box.end[1]; i[1]++)
            for (i[2] = box.begin[2]; i[2] < box.end[2]; i[2]++) {
                const int64_t a = ((i[0] % M) + M) % M, b = ((i[1] % M) + M) % M, c = ((i[2] % M) + M) % M;
                out[box.index(i)] = phi[(a * M + b) * M + c];
            }
    return out;
}

// The reference keeps phi on the chain box times four_o_chain padded by
// PHI_BW = 3 (gravity_long.cpp:150-157); here the periodic mesh itself.
range<int> gravity_long_get_phi_box() {
    range<int> r;
    const int M = get_options().four_dim;
    r.begin = {-3, -3, -3};
    r.end = {M + 3, M + 3, M + 3};
    return r;
}

// gravity_long_force_at (gravity_long.cpp:52-125): 4-point cubic (Keys
// a = -1/2) interpolation of phi and its analytic derivative, in double.
std::pair<float, std::array<float, NDIM>> gravity_long_force_at(const std::array<double, NDIM>& pos) {
    const auto& phi = host_phi();
    const int M = get_options().four_dim;
    const double N = M;
    std::array<int, NDIM> I;
    std::array<double, NDIM> X;
    double w[NDIM][4], dw[NDIM][4];
    for (int d = 0; d < NDIM; d++) {
        X[d] = pos[d] * N;
        I[d] = (int)X[d];
        X[d] -= (double)I[d];
        I[d] -= 1;
        const double x1 = X[d], x2 = x1 * x1, x3 = x1 * x2;
        w[d][0] = -0.5 * x1 + x2 - 0.5 * x3;
        w[d][1] = 1.0 - 2.5 * x2 + 1.5 * x3;
        w[d][2] = 0.5 * x1 + 2.0 * x2 - 1.5 * x3;
        w[d][3] = -0.5 * x2 + 0.5 * x3;
        dw[d][0] = -0.5 + 2.0 * x1 - 1.5 * x2;
        dw[d][1] = -5.0 * x1 + 4.5 * x2;
        dw[d][2] = 0.5 + 4.0 * x1 - 4.5 * x2;
        dw[d][3] = -x1 + 1.5 * x2;
    }
    double g[NDIM] = {0, 0, 0}, p = 0.0;
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++)
            for (int c = 0; c < 4; c++) {
                const int64_t ia = ((I[0] + a) % M + M) % M, ib = ((I[1] + b) % M + M) % M,
                              ic = ((I[2] + c) % M + M) % M;
                const double ph = phi[(ia * M + ib) * M + ic];
                g[0] += dw[0][a] * w[1][b] * w[2][c] * ph * N;
                g[1] += w[0][a] * dw[1][b] * w[2][c] * ph * N;
                g[2] += w[0][a] * w[1][b] * dw[2][c] * ph * N;
                p += w[0][a] * w[1][b] * w[2][c] * ph;
            }
    return {(float)p, {(float)-g[0], (float)-g[1], (float)-g[2]}};
}

// ------------------------------------------------------------------- kicks
// The device holds the current chain mesh and potential; rebuilds them when
// the particles were uploaded since (tpm_kick reports TPM_ERR_STATE).
static void device_ready() {
    push_params(g_long_type);
    if (!g_chain_valid) chainmesh_create();
    if (!g_pm_computed) {
        check(tpm_gravity_long_compute(ctx()), "tpm_gravity_long_compute");
        g_pm_computed = true;
    }
}

static void record(const tpm_stats& st) {
    g_report.ms_pp = st.ms_pp;
    g_report.candidate_pairs = st.candidate_pairs;
    g_report.executed_pairs = st.executed_pairs;
    g_report.pc_interactions = st.pc_interactions;
}

// The kick of the sinks in box (nullptr: all cells) with the given flags.
static void device_kick(const range<int>* box, int min_rung, double scale, double t0, bool first_call,
                        unsigned flags) {
    device_ready();
    tpm_stats st{};
    int32_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (box)
        for (int d = 0; d < NDIM; d++) {
            lo[d] = box->begin[d];
            hi[d] = box->end[d];
        }
    auto run = [&]() {
        return box ? tpm_kick_box(ctx(), lo, hi, min_rung, scale, t0, first_call ? 1 : 0, flags, &st)
                   : tpm_kick(ctx(), min_rung, scale, t0, first_call ? 1 : 0, flags, &st);
    };
    tpm_status s = run();
    if (s == TPM_ERR_STATE) {  // particles changed on the host since the sort
        chainmesh_create();
        check(tpm_gravity_long_compute(ctx()), "tpm_gravity_long_compute");
        g_pm_computed = true;
        s = run();
    }
    check(s, box ? "tpm_kick_box" : "tpm_kick");
    record(st);
    download_state(true);
}

static unsigned short_range_flags() { return get_options().short_range == "tree" ? TPM_KICK_TREE : 0u; }

// The device's forest on the host, once per chain-mesh build.
static host_forest& forest() {
    if (g_forest.valid) return g_forest;
    device_ready();
    check(tpm_tree_build(ctx()), "tpm_tree_build");
    int64_t nn = 0, nb = 0;
    check(tpm_tree_info(ctx(), &nn, &nb), "tpm_tree_info");
    const int C = get_options().chain_dim;
    const size_t ncell = (size_t)C * C * C;
    host_forest& f = g_forest;
    f.node_begin.assign(ncell + 1, 0);
    f.bucket_begin.assign(ncell + 1, 0);
    f.nodes.resize((size_t)nn);
    f.buckets.resize((size_t)nb);
    f.order.resize((size_t)particles_size());
    check(tpm_tree_download(ctx(), f.node_begin.data(), f.bucket_begin.data(), f.nodes.data(), f.buckets.data(),
                            f.order.data()),
          "tpm_tree_download");
    f.cell_sorted.assign(ncell, 0);
    f.valid = true;
    return f;
}

static int64_t wrap_cell(const std::array<int, NDIM>& i) {
    const int C = get_options().chain_dim;
    const int ix = ((i[0] % C) + C) % C, iy = ((i[1] % C) + C) % C, iz = ((i[2] % C) + C) % C;
    return ((int64_t)iz * C + iy) * C + ix;
}

// tree_create (tree.cpp:216-245)
std::pair<tree, std::vector<sink_bucket>> tree_create(const std::array<int, NDIM>& cell_index, chaincell cell) {
    host_forest& f = forest();
    const int64_t c = wrap_cell(cell_index);
    const chaincell mine = chainmesh_get(cell_index);
    if (mine.pbegin != cell.pbegin || mine.pend != cell.pend)
        throw std::runtime_error("tree_create: the cell's range is not the current chain mesh's");
    std::pair<tree, std::vector<sink_bucket>> rc;
    for (int64_t k = f.node_begin[c]; k < f.node_begin[c + 1]; k++) {
        const tpm_tree_node& d = f.nodes[(size_t)k];
        tree_node t;
        for (int q = 0; q < NDIM; q++) t.x[q] = fixed32::from_raw(d.x[q]);
        t.radius = d.radius;
        t.children = {d.child[0], d.child[1]};
        t.src_begin = t.snk_begin = d.begin;
        t.src_end = t.snk_end = d.end;
        for (int q = 0; q < MULTIPOLE_SIZE; q++) t.multi[q] = d.m[q];
        rc.first.nodes.push_back(t);
    }
    for (int64_t k = f.bucket_begin[c]; k < f.bucket_begin[c + 1]; k++) {
        const tpm_tree_bucket& d = f.buckets[(size_t)k];
        sink_bucket b;
        b.src_begin = b.snk_begin = d.begin;
        b.src_end = b.snk_end = d.end;
        b.radius = d.radius;
        for (int q = 0; q < NDIM; q++) b.x[q] = fixed32::from_raw(d.x[q]);
        rc.second.push_back(b);
    }
    if (!f.cell_sorted[(size_t)c]) {
        // the cell's particles into tree order: what particles_sort leaves
        // in X, U and R (particles.cpp:426-434)
        const int b = cell.pbegin, n = cell.pend - cell.pbegin;
        std::vector<fixed32> x(n);
        std::vector<float> u(n);
        std::vector<char> r(n);
        for (int d = 0; d < NDIM; d++) {
            for (int k = 0; k < n; k++) x[k] = particles_X[d][f.order[(size_t)(b + k)]];
            std::copy(x.begin(), x.end(), particles_X[d].begin() + b);
            for (int k = 0; k < n; k++) u[k] = particles_U[d][f.order[(size_t)(b + k)]];
            std::copy(u.begin(), u.end(), particles_U[d].begin() + b);
        }
        for (int k = 0; k < n; k++) r[k] = particles_R[f.order[(size_t)(b + k)]];
        std::copy(r.begin(), r.end(), particles_R.begin() + b);
        f.cell_sorted[(size_t)c] = 1;
        g_tree_pending = true;
    }
    return rc;
}

// kick_treepm_begin (kick_treepm.cpp:10-54 + kick_treepm.cu:761-1089): the
// trees of every cell -- on the device, the whole forest at once, with the
// particles left in tree order as tree_create leaves them -- then the kick
// of the interior box.
void kick_treepm_begin(int min_rung, double scale, double t0, bool first_call) {
    device_ready();
    check(tpm_tree_build(ctx()), "tpm_tree_build");
    g_tree_pending = true;
    device_kick(nullptr, min_rung, scale, t0, first_call, short_range_flags());
}

// kick_treepm (kick_treepm.hpp:14)
void kick_treepm(std::vector<tree> trees, std::vector<std::vector<sink_bucket>> buckets, range<int> box, int min_rung,
                 double scale, double t0, bool first_call) {
    const range<int> bigbox = box.pad(1);
    if ((int)trees.size() != bigbox.volume() || (int)buckets.size() != box.volume())
        throw std::runtime_error("kick_treepm: need the trees of box.pad(1) and the sink buckets of box");
    host_forest& f = forest();
    std::array<int, NDIM> i;
    for (i[0] = bigbox.begin[0]; i[0] < bigbox.end[0]; i[0]++)
        for (i[1] = bigbox.begin[1]; i[1] < bigbox.end[1]; i[1]++)
            for (i[2] = bigbox.begin[2]; i[2] < bigbox.end[2]; i[2]++) {
                const int64_t c = wrap_cell(i);
                if ((int64_t)trees[bigbox.index(i)].size() != f.node_begin[c + 1] - f.node_begin[c] ||
                    (box.contains(i) &&
                     (int64_t)buckets[box.index(i)].size() != f.bucket_begin[c + 1] - f.bucket_begin[c]))
                    throw std::runtime_error("kick_treepm: the trees are not tree_create's of the current chain mesh");
            }
    device_kick(&box, min_rung, scale, t0, first_call, short_range_flags());
}

void kick_treepm_end() {}  // no ghost particles to drop on one device

// kick_pme (kick_pme.hpp:15, kick_pme.cu:390-621): kick_pme's pair law (no
// cutoff, plain softening inside h, no phi0) over the sinks of box.
void kick_pme(range<int> box, int min_rung, double scale, double t0, bool first_call) {
    device_kick(&box, min_rung, scale, t0, first_call, TPM_KICK_PME);
}
// kick_pme_begin (kick_pme.cpp): kick_pme of the interior box.
void kick_pme_begin(int min_rung, double scale, double t0, bool first_call) {
    device_kick(nullptr, min_rung, scale, t0, first_call, TPM_KICK_PME);
}
void kick_pme_end() {}

void set_device_budget(int64_t bytes) { check(tpm_set_budget(ctx(), bytes), "tpm_set_budget"); }

// kick_pm (kick_pm.cpp:9-39): PM force at every particle on the CPU.
void kick_pm() {
    const int n = particles_size();
    host_phi();
    for (int i = 0; i < n; i++) {
        const std::array<double, NDIM> pos = {particles_X[0][i].to_double(), particles_X[1][i].to_double(),
                                              particles_X[2][i].to_double()};
        const auto gf = gravity_long_force_at(pos);
        for (int d = 0; d < NDIM; d++) particles_G[d][i] = gf.second[d];
        particles_P[i] = gf.first;
    }
}

kick_report last_kick_report() { return g_report; }

// ------------------------------------------------------------------- ewald
static std::vector<int64_t> sample_indices(int nsamples) {
    std::mt19937_64 rng(get_options().seed + 2000);
    const int n = particles_size();
    std::vector<int64_t> s;
    std::vector<char> taken(n, 0);
    while ((int)s.size() < std::min(nsamples, n)) {
        const int64_t i = (int64_t)(rng() % n);
        if (!taken[i]) {
            taken[i] = 1;
            s.push_back(i);
        }
    }
    return s;
}

// gravity_short_ewald_call_kernel (gravity_short.cu:10-91): the periodic
// potential and force of every resident particle at each sink position
// (gravity_ewald_kernel, :93-201: every source counts, softened inside
// hsoft, so a sink on a particle gets its -35/16 h^-1).  FP64 Ewald sum
// (tpm_ewald_at); unit masses, as the reference's kernel.
std::pair<std::vector<double>, std::array<std::vector<double>, NDIM>> gravity_short_ewald_call_kernel(
    const std::vector<fixed32>& sinkx, const std::vector<fixed32>& sinky, const std::vector<fixed32>& sinkz) {
    const size_t ns = sinkx.size();
    if (sinky.size() != ns || sinkz.size() != ns)
        throw std::runtime_error("gravity_short_ewald_call_kernel: sink arrays of different lengths");
    push_params(g_long_type);
    upload();
    std::pair<std::vector<double>, std::array<std::vector<double>, NDIM>> rc;
    rc.first.resize(ns);
    for (int d = 0; d < NDIM; d++) rc.second[d].resize(ns);
    static_assert(sizeof(fixed32) == sizeof(uint32_t), "layout");
    check(tpm_ewald_at(ctx(), (int64_t)ns, reinterpret_cast<const uint32_t*>(sinkx.data()),
                       reinterpret_cast<const uint32_t*>(sinky.data()), reinterpret_cast<const uint32_t*>(sinkz.data()),
                       0.0, rc.first.data(), rc.second[0].data(), rc.second[1].data(), rc.second[2].data()),
          "tpm_ewald_at");
    return rc;
}

ewald_errors gravity_short_ewald_errors(int nsamples) {
    const options& o = get_options();
    push_params(g_long_type);
    upload();
    const auto s = sample_indices(nsamples);
    const size_t ns = s.size();
    std::vector<double> ph(ns), gx(ns), gy(ns), gz(ns);
    check(tpm_ewald(ctx(), (int64_t)ns, s.data(), 0.0, ph.data(), gx.data(), gy.data(), gz.data()), "tpm_ewald");
    // unit-mass Ewald sums times GM; the kick's potential carries the mesh
    // self term -GM / (sqrt(pi) rs) that the Ewald potential does not.
    const double GM = o.GM;
    const double self = g_long_type == GRAVITY_LONG_PME ? GM / (std::sqrt(M_PI) * o.rs) : 0.0;
    double l2f = 0, n2f = 0, lmax = 0, l2v = 0, l2p = 0, n2p = 0;
    for (size_t k = 0; k < ns; k++) {
        const int64_t i = s[k];
        const double e[3] = {gx[k] * GM, gy[k] * GM, gz[k] * GM};
        const double g[3] = {particles_G[0][i], particles_G[1][i], particles_G[2][i]};
        const double g1 = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        const double g2 = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        l2f += (g1 - g2) * (g1 - g2);
        lmax += std::fabs(g1 - g2);
        n2f += g2 * g2;
        for (int d = 0; d < 3; d++) l2v += (g[d] - e[d]) * (g[d] - e[d]);
        const double p1 = particles_P[i] + self, p2 = ph[k] * GM;
        l2p += (p1 - p2) * (p1 - p2);
        n2p += p2 * p2;
    }
    ewald_errors r;
    r.force_rms_magnitude = std::sqrt(l2f / n2f);
    r.force_max = lmax / std::sqrt(n2f);
    r.force_rms_vector = std::sqrt(l2v / n2f);
    r.potential_rms = std::sqrt(l2p / n2p);
    r.nsamples = (int)ns;
    return r;
}

// gravity_short_ewald_compare (gravity_short.cpp:35-82): prints the
// reference's accuracy report.
void gravity_short_ewald_compare(int nsamples) {
    const ewald_errors e = gravity_short_ewald_errors(nsamples);
    std::printf("Force RMS Error     = %e\n", e.force_rms_magnitude);
    std::printf("Force Max Error     = %e\n", e.force_max);
    std::printf("Force RMS Vector    = %e\n", e.force_rms_vector);
    std::printf("Potential RMS Error = %e\n", e.potential_rms);
    std::fflush(stdout);
}

// -------------------------------------------------------------------- tests
// The kick scenarios of test.cpp:120-346 (timings printed like the reference).
static void timed(const char* label, void (*f)()) {
    std::printf("%s\n", label);
    const double t = now();
    f();
    std::printf("%e s\n", now() - t);
    std::fflush(stdout);
}

void run_test(std::string test) {
    std::printf("Testing\n");
    const options& o = get_options();
    auto init = []() {
        if (get_options().ic == "zeldovich")
            initialize();
        else
            particles_random_init();
    };
    if (test == "treepm" || test == "kick_pme") {
        // kick_treepm_test / kick_pme_test (test.cpp:237-290, :182-235); the
        // kick_pme scenario starts from random particles
        const bool pme = test == "kick_pme";
        if (pme)
            timed("INIT", particles_random_init);
        else
            timed("INIT", +[]() {
                if (get_options().ic == "zeldovich")
                    initialize();
                else
                    particles_random_init();
            });
        timed("DOMAIN SORT", particles_domain_sort);
        timed("FOURIER", +[]() { gravity_long_compute(GRAVITY_LONG_PME); });
        timed("SORT", chainmesh_create);
        timed("BOUNDARIES", +[]() {
            chainmesh_exchange_begin();
            chainmesh_exchange_end();
        });
        if (pme)
            timed("KICK", +[]() {
                kick_pme_begin(0, 1.0, 1.0, true);
                kick_pme_end();
            });
        else
            timed("KICK", +[]() {
                kick_treepm_begin(0, 1.0, 1.0, true);
                kick_treepm_end();
            });
        const kick_report r = last_kick_report();
        std::printf("PP kernel %e ms, candidate pairs %e, evaluated pairs %e, multipole interactions %e\n", r.ms_pp,
                    r.candidate_pairs, r.executed_pairs, r.pc_interactions);
        std::printf("COMPARE\n");
        gravity_short_ewald_compare(100);
    } else if (test == "kick_pm") {
        init();
        timed("FOURIER", +[]() { gravity_long_compute(GRAVITY_LONG_PM); });
        timed("KICK", kick_pm);
        gravity_short_ewald_compare(100);
    } else if (test == "chainmesh" || test == "sort") {
        init();
        timed("SORT", chainmesh_create);
    } else if (test == "gravity_long") {
        init();
        timed("FOURIER", +[]() { gravity_long_compute(GRAVITY_LONG_PME); });
    } else {
        std::printf("%s is an unknown test.\n", test.c_str());
        (void)o;
        throw std::runtime_error("unknown test");
    }
}
