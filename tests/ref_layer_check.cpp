// ref_layer_check.cpp -- drives the reference-named C++ layer
// (libtigerpm_ref.so, include/tigerpm_ref.hpp) the way reference code would
// call it and writes what it leaves in the global arrays to raw files, for
// tests/test_gpu_driver.py to check against the CPU oracle.
//
//   ref_layer_check <mode> <outdir> <parts_dim>
//   force_at: random particles, gravity_long_compute(PME), kick_pm; the
//             potential mesh; gravity_long_force_at at mesh nodes and at
//             random points.
//   kick_pme: random particles with velocities and rungs, FOURIER, SORT,
//             kick_pme over a sub-box, then the whole interior box split by a
//             small device budget.
//   trees:    Zel'dovich particles, FOURIER, SORT, tree_create for every
//             cell of the padded interior box, kick_treepm(trees, buckets,
//             box) with the tree walk.
#include <cstdio>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "tigerpm_ref.hpp"

static std::string g_dir;

template <class T>
static void dump(const std::string& name, const std::vector<T>& v) {
    FILE* f = std::fopen((g_dir + "/" + name).c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + name);
    if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

static std::vector<uint32_t> raw(const std::vector<fixed32>& a) {
    std::vector<uint32_t> r(a.size());
    for (size_t i = 0; i < a.size(); i++) r[i] = a[i].raw();
    return r;
}

static void dump_particles(const std::string& tag) {
    for (int d = 0; d < NDIM; d++) {
        dump(tag + "_x" + std::to_string(d), raw(particles_X[d]));
        dump(tag + "_u" + std::to_string(d), particles_U[d]);
        dump(tag + "_g" + std::to_string(d), particles_G[d]);
    }
    dump(tag + "_p", particles_P);
    dump(tag + "_r", particles_R);
}

static void setup(int parts_dim, const char* ic, uint64_t seed) {
    options o;
    o.parts_dim = parts_dim;
    o.four_dim_override = parts_dim;
    o.rs_cells = 1.25;
    o.rcut_rs = 4.5;
    o.order = 3;
    o.unit_mass = false;
    o.ic = ic;
    o.rms_disp = 1.0;
    o.seed = seed;
    set_options(o);
}

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: ref_layer_check force_at|kick_pme|trees <outdir> <parts_dim>\n");
        return 2;
    }
    const std::string mode = argv[1];
    g_dir = argv[2];
    const int parts_dim = std::stoi(argv[3]);
    if (mode == "force_at") {
        setup(parts_dim, "uniform", 11);
        particles_random_init();
        gravity_long_compute(GRAVITY_LONG_PME);
        kick_pm();
        dump_particles("pm");
        const int M = get_options().four_dim;
        range<int> box;
        box.begin = {0, 0, 0};
        box.end = {M, M, M};
        dump("phi", gravity_long_get_phi(box));
        // nodes: i / M is exact for M a power of two
        std::vector<double> q, out;
        std::mt19937_64 rng(5);
        for (int k = 0; k < 200; k++) {
            const int i = (int)(rng() % M), j = (int)(rng() % M), l = (int)(rng() % M);
            q.insert(q.end(), {(double)i / M, (double)j / M, (double)l / M});
        }
        std::uniform_real_distribution<double> U(0.0, 1.0);
        for (int k = 0; k < 2000; k++) q.insert(q.end(), {U(rng), U(rng), U(rng)});
        for (size_t k = 0; k < q.size(); k += 3) {
            const auto r = gravity_long_force_at({q[k], q[k + 1], q[k + 2]});
            out.insert(out.end(), {(double)r.first, (double)r.second[0], (double)r.second[1], (double)r.second[2]});
        }
        dump("query", q);
        dump("force_at", out);
    } else if (mode == "kick_pme") {
        setup(parts_dim, "uniform", 12);
        particles_random_init();
        std::mt19937_64 rng(6);
        std::normal_distribution<float> N(0.f, 1e-3f);
        for (int i = 0; i < particles_size(); i++) {
            for (int d = 0; d < NDIM; d++) particles_U[d][i] = N(rng);
            particles_R[i] = (char)(rng() % 3);
        }
        gravity_long_compute(GRAVITY_LONG_PME);
        chainmesh_create();
        dump_particles("sorted");
        const int C = get_options().chain_dim;
        range<int> sub;
        sub.begin = {1, 0, 2};
        sub.end = {C - 1, C, 3};
        kick_pme(sub, 1, 1.0, 1e-2, false);
        dump_particles("sub");
        set_device_budget(200000);
        kick_pme(chainmesh_interior_box(), 1, 1.0, 1e-2, false);
        set_device_budget(0);
        dump_particles("split");
    } else if (mode == "trees") {
        setup(parts_dim, "zeldovich", 13);
        {
            options o = get_options();
            o.short_range = "tree";
            set_options(o);
        }
        initialize();
        gravity_long_compute(GRAVITY_LONG_PME);
        chainmesh_create();
        dump_particles("sorted");
        const range<int> box = chainmesh_interior_box();
        const range<int> bigbox = box.pad(1);
        std::vector<tree> trees(bigbox.volume());
        std::vector<std::vector<sink_bucket>> buckets(box.volume());
        std::vector<int32_t> counts;
        std::array<int, NDIM> i;
        for (i[0] = bigbox.begin[0]; i[0] < bigbox.end[0]; i[0]++)
            for (i[1] = bigbox.begin[1]; i[1] < bigbox.end[1]; i[1]++)
                for (i[2] = bigbox.begin[2]; i[2] < bigbox.end[2]; i[2]++) {
                    auto rc = tree_create(i, chainmesh_get(i));
                    if (box.contains(i)) {
                        counts.push_back((int32_t)rc.first.size());
                        counts.push_back((int32_t)rc.second.size());
                        buckets[box.index(i)] = rc.second;
                    }
                    trees[bigbox.index(i)] = std::move(rc.first);
                }
        dump("tree_counts", counts);  // interior cells in box.index order: nodes, buckets
        dump_particles("treeorder");  // X, U, R permuted by tree_create; G, P as before
        kick_treepm(trees, buckets, box, 0, 1.0, 1.0, true);
        dump_particles("kicked");
    } else {
        std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 2;
    }
    std::printf("ok %s\n", mode.c_str());
    return 0;
}
