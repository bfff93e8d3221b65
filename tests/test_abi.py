"""The C-ABI library without a GPU: it loads, exports exactly what
include/pdipc.h declares, and its host precompute (the part of initialization
that needs no device) matches the reference fixtures and the oracle's
independent construction."""

import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import pdipc
import scenes

from conftest import ROOT, gpu_available


def _header_functions():
    text = open(os.path.join(ROOT, "include", "pdipc.h")).read()
    return set(re.findall(r"^\s*(?:int|void|const char\*)\s+(pdipc_\w+)\s*\(", text, re.M))


def test_library_exports_every_declared_symbol():
    L = pdipc.lib()
    declared = _header_functions()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), name
    assert declared == set(pdipc.SIGNATURES), declared ^ set(pdipc.SIGNATURES)


def test_param_struct_layout_matches_header():
    text = open(os.path.join(ROOT, "include", "pdipc.h")).read()
    body = text[text.index("typedef struct pdipc_params"):text.index("} pdipc_params;")]
    names = re.findall(r"^\s*(?:int|double)\s+(\w+);", body, re.M)
    assert names == [n for n, _ in pdipc.Params._fields_]
    body = text[text.index("typedef struct pdipc_counters"):text.index("} pdipc_counters;")]
    names = re.findall(r"(?:int|double|long long)\s+([\w, ]+);", body)
    flat = [n.strip() for group in names for n in group.split(",")]
    assert flat == [n for n, _ in pdipc.Counters._fields_]


def test_default_params_are_the_reference_constants():
    p = pdipc.default_params()
    assert (p.lg_iters, p.cheb_iters, p.max_outer, p.ccd_iters) == (10, 15, 30, 10)
    assert p.residual_tol == 1e-5 and abs(p.omega - 2 / 3) < 1e-16 and p.rho == 0.9992
    assert (p.sigma_lo, p.sigma_hi, p.ccd_filter, p.toi_scale) == (0.999, 1.001, 5e-3, 0.98)
    assert (p.ccd_step, p.ccd_gap, p.static_mass, p.dhat, p.kappa) == (0.9, 0.1, 1e4, 1e-3, 1e4)
    for k, v in scenes.DEFAULT_PARAMS.items():
        if k != "kappa":
            assert getattr(p, k) == v, k


def test_host_setup_reproduces_reference_fixtures(golden):
    for name in ("cloth1", "cloth2"):
        s = scenes.make_scene()
        scenes.add_cloth(s, golden[name + "_v"].astype(np.float64), golden[name + "_f"], strain=1e6, bend=25, pos=1e12,
                         density=2500, patch_dim=50, fixed=golden[name + "_fixed"])
        sim = pdipc.Simulator.from_scene(s, init=False)
        sim.host_setup()
        assert np.array_equal(sim.model_dat(0, 0), golden[name + "_boundary"])
        assert np.array_equal(sim.model_dat(0, 1), golden[name + "_patch"])


@pytest.mark.parametrize("which", ["cloth", "cloth_tets"])
def test_host_operator_sizes_match_oracle(which):
    if which == "cloth":
        s = scenes.cloth_over_sphere(16, 2, ground=True, seed=1, patch_dim=16, sphere_patch_dim=16)
    else:
        s = scenes.config2(cells=3, cloth_n=8)
    sim = pdipc.Simulator.from_scene(s, init=False)
    sim.host_setup()
    o = oracle.Oracle(s)
    c, oc = sim.counters(), o.counters()
    assert c["nnz_r1"] == oc["nnz_r1"]
    assert c["nnz_r2"] == oc["nnz_r2"]
    assert c["n_2paths"] == oc["n_2paths"]
    assert c["patches"] == o.patches
    assert c["boundary_faces"] == o.bfaces and c["boundary_edges"] == o.bedges
    assert c["sim_points"] == o.sim and c["total_points"] == o.total


def test_bad_mesh_is_an_error_not_an_abort():
    sim = pdipc.Simulator()
    v = np.zeros((3, 3))
    sim.add_cloth(v, np.array([[0, 1, 5]], np.int32), 1e4, 1e-2, 1e12, 0.2, 4)
    with pytest.raises(pdipc.PdipcError, match="out of range"):
        sim.host_setup()
    sim2 = pdipc.Simulator()
    with pytest.raises(pdipc.PdipcError):
        sim2.host_setup()  # no model


def test_calls_before_init_fail_cleanly():
    sim = pdipc.Simulator()
    with pytest.raises(pdipc.PdipcError, match="pdipc_init"):
        sim.run(1)


@pytest.mark.skipif(gpu_available(), reason="checks the no-device path")
def test_init_without_device_reports_nodevice():
    s = scenes.cloth_over_sphere(4, 0, seed=0, patch_dim=4, sphere_patch_dim=4)
    sim = pdipc.Simulator.from_scene(s, init=False)
    with pytest.raises(pdipc.PdipcError, match="-6"):
        sim.initialization()
