"""The reference's --test= scenarios (test.cpp:120-346) run through the C++
layer with the reference's entry-point names (libtigerpm_ref.so), on the GPU."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tigerpm_amd", "bin", "tigerpm")


def run(*args):
    r = subprocess.run([DRIVER] + list(args), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = dict(re.findall(r"^(Force RMS Error|Force RMS Vector|Force Max Error)\s*= (\S+)", r.stdout, re.M))
    return r.stdout, {k: float(v) for k, v in vals.items()}


def test_treepm_scenario_baseline_parameters():
    # 64^3 uniform particles, 64^3 mesh, rs = 1.25 cells, rcut = 4.5 rs, TSC, m = 1/N
    out, e = run("--test=treepm", "--parts_dim=64", "--four_dim=64", "--rs_cells=1.25", "--rcut_rs=4.5",
                 "--order=3", "--mass=1/N")
    assert "KICK" in out and "COMPARE" in out
    assert e["Force RMS Vector"] < 1e-3 and e["Force RMS Error"] < 1e-3


def test_treepm_scenario_reference_defaults():
    # reference defaults (options.cpp:81-96): rs = 1 mesh cell, rcut = 1/chain_dim, cloud4, unit masses
    out, e = run("--test=treepm", "--parts_dim=90")
    assert e["Force RMS Error"] < 1e-2


def test_treepm_scenario_tree_walk():
    # --short=tree: the reference's own short-range algorithm (per-cell trees,
    # theta = 0.4 monopole + dipole, PP for near leaves) at its defaults
    out, e = run("--test=treepm", "--parts_dim=90", "--short=tree")
    m = re.search(r"multipole interactions (\S+)", out)
    assert m and float(m.group(1)) > 0
    _, d = run("--test=treepm", "--parts_dim=90")
    # the multipoles add their approximation error to the direct sum's
    assert d["Force RMS Error"] < e["Force RMS Error"] < 5e-2


def test_kick_pm_scenario():
    out, e = run("--test=kick_pm", "--parts_dim=60")
    assert "Force RMS Error" in out


def test_kick_pme_scenario():
    # kick_pme_test (test.cpp:182-235): random particles, kick_pme_begin
    out, e = run("--test=kick_pme", "--parts_dim=60")
    assert "KICK" in out and "COMPARE" in out
    assert 0.0 < e["Force RMS Error"] < 0.5
