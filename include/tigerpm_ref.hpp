// tigerpm_ref.hpp -- the reference's C++ entry points for the TreePM kick
// path, implemented over the C ABI of libtigerpm_hip.so (tigerpm_hip.h).
//
// Names, argument meanings and the module-global particle store follow the
// reference headers (options.hpp, particles.hpp, chainmesh.hpp,
// gravity_long.hpp, gravity_short.hpp, kick_treepm.hpp, kick_pme.hpp,
// kick_pm.hpp, initialize.hpp, range.hpp, fixed.hpp), so code written
// against them links here unchanged on one GPU.  Multi-locality pieces
// (HPX fan-out, halo exchange, domain sort) reduce to no-ops on one device.
// Errors from the device library throw std::runtime_error carrying
// tpm_last_error(); the reference prints and aborts on the same conditions.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#define NDIM 3
#define XDIM 0
#define YDIM 1
#define ZDIM 2
#define MAX_RUNG 32
#define SELF_PHI float(-35.0 / 16.0)

// ---- fixed.hpp:23-235 ------------------------------------------------------
// 32-bit fixed point in [0, 1): raw = uint32(2^32 * x) (truncating).
class fixed32 {
    uint32_t i_ = 0;

public:
    fixed32() = default;
    fixed32(double x) : i_((uint32_t)(4294967296.0 * x)) {}
    static fixed32 from_raw(uint32_t r) {
        fixed32 f;
        f.i_ = r;
        return f;
    }
    uint32_t raw() const { return i_; }
    float to_float() const { return (float)i_ * 0x1p-32f; }
    double to_double() const { return (double)i_ * 0x1p-32; }
    bool operator<(fixed32 o) const { return i_ < o.i_; }
    bool operator==(fixed32 o) const { return i_ == o.i_; }
};
static_assert(sizeof(fixed32) == 4, "fixed32 must be the raw 32-bit word");

// gravity_short.hpp:20-22: periodic minimum image.
inline float distance(fixed32 a, fixed32 b) { return (float)(int32_t)(a.raw() - b.raw()) * 0x1p-32f; }

// ---- range.hpp:31-203 (the parts the kick path uses) ------------------------
template <class T>
struct range {
    std::array<T, NDIM> begin{};
    std::array<T, NDIM> end{};
    T volume() const {
        T v = 1;
        for (int d = 0; d < NDIM; d++) v *= (end[d] > begin[d]) ? end[d] - begin[d] : 0;
        return v;
    }
    // row-major, z fastest (range.hpp:149-165)
    T index(const std::array<T, NDIM>& i) const {
        return (end[2] - begin[2]) * ((end[1] - begin[1]) * (i[0] - begin[0]) + (i[1] - begin[1])) + (i[2] - begin[2]);
    }
    bool contains(const std::array<T, NDIM>& i) const {
        for (int d = 0; d < NDIM; d++)
            if (i[d] < begin[d] || i[d] >= end[d]) return false;
        return true;
    }
    range pad(T dx = 1) const {
        range r = *this;
        for (int d = 0; d < NDIM; d++) {
            r.begin[d] -= dx;
            r.end[d] += dx;
        }
        return r;
    }
};

// ---- options.hpp:5-49, options.cpp:37-118 ----------------------------------
struct options {
    int four_dim = 0;
    int chain_dim = 0;
    int parts_dim = 130;
    int parts_o_four = 3;
    int parts_o_chain = 15;
    int four_o_chain = 5;
    double eta = 0.0;
    double GM = 1.0;
    double hsoft = 0.0;
    double box_size = 1.0;
    double rs = 0.0;
    std::string config_file;
    std::string test;
    // Workload extensions (defaults reproduce the reference):
    double rs_cells = 0.0;   // split scale in mesh cells; 0 = reference (1 / four_dim)
    double rcut_rs = 0.0;    // cutoff in units of rs; 0 = reference (1 / chain_dim)
    double rcut = 0.0;       // derived
    int four_dim_override = 0;  // mesh cells per side; 0 = parts_dim / parts_o_four
    int order = 4;           // mass assignment: 4 cloud4 (reference), 3 TSC, 2 CIC
    bool unit_mass = true;   // true: unit masses, GM = 1 (reference); false: m = 1/N
    std::string ic = "uniform";  // particles_random_init / initialize(): uniform | zeldovich
    double rms_disp = 0.5;   // Zel'dovich rms displacement in lattice cells
    uint64_t seed = 42;
    int tile_sinks = 16;     // PP work-item size (tpm_tuning)
    int interlace = 0;       // 1: interlaced (two-grid) deposit
    std::string short_range = "direct";  // kick_treepm's short range: direct 27-cell sum | tree (theta = 0.4 walk)
};
// Parses --parts_dim --parts_o_four --four_o_chain --test plus the extensions
// (--four_dim --rs_cells --rcut_rs --order --mass=unit|1/N --ic --rms_disp
// --seed --tile_sinks --interlace --short=direct|tree); derives the dependent
// fields (options.cpp:81-96).
bool process_options(int argc, char* argv[]);
const options& get_options();
void set_options(const options& opts);  // derives and installs

// ---- particles.hpp:17-122 ----------------------------------------------------
struct particle {
    std::array<fixed32, NDIM> x;
    std::array<float, NDIM> v;
    char r;
    std::array<float, NDIM> g;
    float p;
};
extern std::array<std::vector<fixed32>, NDIM> particles_X;
extern std::array<std::vector<float>, NDIM> particles_U;
extern std::vector<char> particles_R;
extern std::vector<float> particles_P;
extern std::array<std::vector<float>, NDIM> particles_G;

int particles_size();
int particles_local_size();
void particles_resize(size_t new_size);
void particles_random_init();  // parts_dim^3 uniform positions (seeded)
void particles_domain_sort();  // one device: nothing to migrate
particle particles_get_particle(int index);
std::vector<particle> particles_sample(int nsamples);
std::array<int, NDIM> particles_mesh_loc(int index);

// ---- initialize.hpp ------------------------------------------------------------
void initialize();  // Zel'dovich lattice (options::ic / rms_disp / seed)

// ---- chainmesh.hpp:8-17 ------------------------------------------------------
struct chaincell {
    int pbegin;
    int pend;
};
void chainmesh_create();
void chainmesh_exchange_begin();  // one device: periodic neighbours are local
void chainmesh_exchange_end();
range<int> chainmesh_interior_box();
chaincell chainmesh_get(const std::array<int, NDIM>& cell);

// ---- kernels.hpp:6-10, tree.hpp:17-110 ------------------------------------------
#define NCHILD 2
#define MULTIPOLE_SIZE 4
using multipole = std::array<float, MULTIPOLE_SIZE>;  // {N, -sum dx, -sum dy, -sum dz}

struct tree_node {
    std::array<fixed32, NDIM> x;  // expansion centre
    float radius;
    std::array<int, NCHILD> children;  // within the cell's tree; -1 for a leaf
    int src_begin;
    int src_end;
    int snk_begin;
    int snk_end;
    multipole multi;
};

struct sink_bucket {
    int src_begin;
    int src_end;
    int snk_begin;
    int snk_end;
    float radius;
    std::array<fixed32, NDIM> x;
};

// One chain cell's tree, nodes in preorder (the reference's tree class
// holds the same nodes behind a raw pointer with a device copy; here the
// device keeps its own copy).
class tree {
public:
    std::vector<tree_node> nodes;
    size_t size() const { return nodes.size(); }
    void resize(int new_size) { nodes.resize(new_size); }
    int allocate() {
        nodes.emplace_back();
        return (int)nodes.size() - 1;
    }
    fixed32 get_x(int dim, int i) const { return nodes[i].x[dim]; }
    int get_pbegin(int i) const { return nodes[i].src_begin; }
    int get_pend(int i) const { return nodes[i].src_end; }
    int get_snk_begin(int i) const { return nodes[i].snk_begin; }
    int get_snk_end(int i) const { return nodes[i].snk_end; }
    multipole get_multipole(int i) const { return nodes[i].multi; }
    float get_mass(int i) const { return nodes[i].multi[0]; }
    bool is_leaf(int i) const { return nodes[i].children[0] == -1; }
    float get_radius(int i) const { return nodes[i].radius; }
    std::array<int, NCHILD> get_children(int i) const { return nodes[i].children; }
    void set(tree_node node, int i) { nodes[i] = node; }
    void adjust_src_indexes(int dif) {
        for (auto& n : nodes) {
            n.src_begin += dif;
            n.src_end += dif;
        }
    }
    void adjust_snk_indexes(int dif) {
        for (auto& n : nodes) {
            n.snk_begin += dif;
            n.snk_end += dif;
        }
    }
};

// ---- gravity_long.hpp:17-25 --------------------------------------------------
using gravity_long_type = int;
#define GRAVITY_LONG_PM 0
#define GRAVITY_LONG_PME 1
void gravity_long_compute(gravity_long_type type = GRAVITY_LONG_PM);
std::vector<float> gravity_long_get_phi(const range<int>& box);  // periodic
range<int> gravity_long_get_phi_box();
std::pair<float, std::array<float, NDIM>> gravity_long_force_at(const std::array<double, NDIM>& pos);

// tree.hpp:108: the tree and sink buckets of one chain cell, its particles
// reordered in place (particles_X/U/R, [cell.pbegin, cell.pend)) into tree
// order, as the reference's two-pointer partitions leave them.  The trees
// are built on the device from the current chain mesh (tpm_tree_build);
// node and bucket ranges index the global particle arrays.
std::pair<tree, std::vector<sink_bucket>> tree_create(const std::array<int, NDIM>& cell_index, chaincell cell);

// ---- kick_treepm.hpp / kick_pme.hpp / kick_pm.hpp ----------------------------
// kick_treepm_begin: the trees of every cell (particles into tree order, as
// tree_create for each cell of the interior box) and the kick of the
// interior box.  The short-range sum is options::short_range: "direct" (the
// erfc-split pair force over the 27 neighbour cells, the default) or "tree"
// (the theta = 0.4 walk of the trees with monopole + dipole multipoles, the
// reference's kick_treepm kernel).
void kick_treepm_begin(int min_rung, double scale, double t0, bool first_call);
// kick_treepm.hpp:14: the kick of the sinks in box (chain cells), with the
// trees of box.pad(1) and the sink buckets of box as tree_create made them
// (their sizes are checked against the device's copies, which the walk
// uses).  Over the device budget (set_device_budget) the box is split in
// halves and kicked one half after the other (kick_treepm.cu:815-823).
void kick_treepm(std::vector<tree> trees, std::vector<std::vector<sink_bucket>> buckets, range<int> box, int min_rung,
                 double scale, double t0, bool first_call);
void kick_treepm_end();
// kick_pme.hpp:15: kick_pme's pair law over the sinks of box (split over
// the device budget as kick_pme.cu:422-429 does).
void kick_pme(range<int> box, int min_rung, double scale, double t0, bool first_call);
void kick_pme_begin(int min_rung, double scale, double t0, bool first_call);
void kick_pme_end();
void kick_pm();
// Device memory a kick's working set may use (bytes, 0 = no limit): the
// reference compares with 85 % of the free device memory
// (kick_treepm.cu:817); here the budget is explicit (tpm_set_budget).
void set_device_budget(int64_t bytes);

// ---- gravity_short.hpp:25-28 -------------------------------------------------
std::pair<std::vector<double>, std::array<std::vector<double>, NDIM>> gravity_short_ewald_call_kernel(
    const std::vector<fixed32>& sinkx, const std::vector<fixed32>& sinky, const std::vector<fixed32>& sinkz);
void gravity_short_ewald_compare(int nsamples);
// The numbers gravity_short_ewald_compare prints (gravity_short.cpp:53-79)
// plus the vector rms, for callers that want to check them.
struct ewald_errors {
    double force_rms_magnitude;  // reference metric: sqrt(sum (|g|-|gE|)^2 / sum |gE|^2)
    double force_max;            // reference "max": sum ||g|-|gE|| / sqrt(sum |gE|^2)
    double force_rms_vector;     // sqrt(sum |g-gE|^2 / sum |gE|^2)
    double potential_rms;
    int nsamples;
};
ewald_errors gravity_short_ewald_errors(int nsamples);

// ---- test.hpp -----------------------------------------------------------------
void run_test(std::string test);
// Last kick's device statistics (milliseconds and pair counts).
struct kick_report {
    double ms_chainmesh, ms_pm, ms_pp;
    double candidate_pairs, executed_pairs;
    double pc_interactions;  // --short=tree: sink x multipole interactions
};
kick_report last_kick_report();
