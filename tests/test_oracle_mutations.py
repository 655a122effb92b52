"""Plants plausible mistakes in a copy of the CPU oracle and checks that the
oracle tests (test_oracle.py, test_oracle_tree.py) catch every one (the pins are not retyped formulas).  CPU only:
each mutant is compiled with gcc into a temporary directory and loaded
through TPM_ORACLE_LIB in a child pytest process."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "tpm_oracle.c")

MUTANTS = {
    "tsc_weight_sign": ("const uint64_t a = (uint64_t)(hp - ds);", "const uint64_t a = (uint64_t)(hp + ds);"),
    "deposit_deficit_dropped": ("w3[centre] += (1ull << 32) - sum;", "w3[centre] += 0;"),
    "green_deconvolution_dropped": ("* pow(s, -(double)order);", ";"),
    "green_sign": ("g = -scale / k2", "g = scale / k2"),
    "soft_erfc_minus_one_dropped": ("rinv3 = ((erfc0 - 1.0f) + c->twooversqrtpi", "rinv3 = ((erfc0) + c->twooversqrtpi"),
    "force_gaussian_term_dropped": ("rinv3 = (erfc0 + c->twooversqrtpi * r0 * exp0) * rinv * rinv * rinv;",
                                    "rinv3 = (erfc0) * rinv * rinv * rinv;"),
    "neighbour_cell_skipped": ("for (int ox = -1; ox <= 1; ox++) {\n                    const int jx",
                               "for (int ox = -1; ox <= 0; ox++) {\n                    const int jx"),
    "interp_gradient_sign": ("gx -= dw[0][a] * w[1][b] * w[2][c] * p1;", "gx += dw[0][a] * w[1][b] * w[2][c] * p1;"),
    "cell_order_transposed": ("const uint32_t cid = ((uint32_t)cz * C + cy) * C + cx;",
                              "const uint32_t cid = ((uint32_t)cx * C + cy) * C + cz;"),
    "ewald_kspace_force_sign": ("G[0] -= 4.0 * c * hx * im;", "G[0] += 4.0 * c * hx * im;"),
    "ewald_self_constant_dropped": ("double P = (double)n * M_PI / (a * a) + (samples ? tsp * a : 0.0);",
                                    "double P = (double)n * M_PI / (a * a);"),
    "ewald_at_self_constant_kept": ("double P = (double)n * M_PI / (a * a) + (samples ? tsp * a : 0.0);",
                                    "double P = (double)n * M_PI / (a * a) + tsp * a;"),
    "tree_partition_idx_not_swapped": ("                    t->idx[hi] = t->idx[lo];\n                    t->idx[lo] = k;",
                                       "                    (void)k;"),
    "partition_plane_goes_left": ("if (xd[pc->idx[lo]] >= xmid) {", "if (xd[pc->idx[lo]] > xmid) {"),
    "partition_stable_instead_of_swaps": ("                    int64_t t = pc->idx[hi];\n"
                                          "                    pc->idx[hi] = pc->idx[lo];\n"
                                          "                    pc->idx[lo] = t;",
                                          "                    int64_t t = pc->idx[hi];\n"
                                          "                    for (int64_t q = hi; q > lo; q--) pc->idx[q] = pc->idx[q - 1];\n"
                                          "                    pc->idx[lo] = t;"),
    "rung_rounding": ("float nr = fmaxf(ceilf(log2f(t0f)", "float nr = fmaxf(floorf(log2f(t0f)"),
    "pme_soft_variable": ("        const float q2 = q * q;\n        rinv3 = +15.0f / 8.0f;",
                          "        const float q2 = q;\n        rinv3 = +15.0f / 8.0f;"),
    "drift_truncated": ("x[i] += (uint32_t)(uint64_t)llrint((double)vx[i] * f);",
                        "x[i] += (uint32_t)(uint64_t)(int64_t)((double)vx[i] * f);"),
    "tree_leaf_test_with_sink_radius": ("const float lim = nd->radius * nd->radius * theta2inv;",
                                        "const float lim = (nd->radius + bk->radius) * (nd->radius + bk->radius) * "
                                        "theta2inv;"),
    "tree_dipole_sign": ("            m[1] += -dx[0];", "            m[1] += dx[0];"),
    "tree_cutoff_radius_sign": ("const int cutoff = R - nd->radius - bk->radius > rcut;",
                                "const int cutoff = R + nd->radius - bk->radius > rcut;"),
    "pc_m2l_index": ("    L[2] = fmaf(M[2], D[7], L[2]);", "    L[2] = fmaf(M[2], D[8], L[2]);"),
    "pc_d1_sign": ("const float d1 = fmaf(-1.f * d0, rinv, e0);", "const float d1 = fmaf(d0, rinv, e0);"),
    "tree_leaf_radius_without_hsoft": ("        node.radius = (float)(sqrt(r2max) + t->hsoft);\n        node.child[0]",
                                       "        node.radius = (float)(sqrt(r2max));\n        node.child[0]"),
    "tree_split_last_longest_side": ("            if (be[d] - bb[d] > span) {", "            if (be[d] - bb[d] >= span) {"),
}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_is_caught(name, tmp_path):
    src = open(SRC).read()
    old, new = MUTANTS[name]
    assert old in src, "mutation anchor not found: %s" % name
    c = tmp_path / "m.c"
    c.write_text(src.replace(old, new, 1))
    so = tmp_path / "lib.so"
    subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fPIC", "-fopenmp", "-shared", "-o", str(so),
                    str(c), "-lm"], check=True)
    env = dict(os.environ, TPM_ORACLE_LIB=str(so), OMP_NUM_THREADS="4")
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_oracle.py"),
                        os.path.join(ROOT, "tests", "test_oracle_tree.py"), "-q", "-x",
                        "-m", "not gpu", "-p", "no:cacheprovider"], env=env, capture_output=True, text=True,
                       timeout=600, cwd=ROOT)
    assert r.returncode != 0, "mutant %s survived the oracle tests" % name
