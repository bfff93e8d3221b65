"""Mutation tests of the closed forms: the oracle rebuilt with one planted
misreading of the reference at a time (ORC_MUTANT in pdipc_oracle.cpp) must
fail the closed-form tests that target it, while the unmodified oracle passes
them.  Each build runs the closed-form tests in a separate process."""

import os
import subprocess
import sys

import pytest

from conftest import ROOT

ORACLE = os.path.join(ROOT, "oracle")

# planted mistake -> the closed-form tests (a pytest -k expression) that must catch it
MUTANTS = {
    1: ("collision split ratios m_q/(m_p+m_q) and m_p/(m_p+m_q) swapped", "barrier_projection_closed_form and not rejects"),
    2: ("triangle targets weighted by barycentric coordinates", "barrier_projection_closed_form and not rejects"),
    3: ("ACCD step factor 0.8 instead of 0.9", "accd_head_on or accd_geometric"),
    4: ("ACCD gap test after the increment", "accd_head_on or accd_geometric"),
    5: ("spring ratio |d|/l0 instead of l0/|d|", "uniform_scaling"),
    6: ("bending rescale |a|/H0 instead of H0/|a|", "uniform_scaling"),
    7: ("static collision mass 1 instead of the static_mass parameter", "barrier_projection_closed_form and not rejects"),
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("oracle_mutants"))  # outside the repository
    subprocess.run(["make", "-s", "-j8", "-C", ORACLE, "mutants", f"MUTDIR={out}"], check=True)
    return {k: os.path.join(out, f"liboracle_mut{k}.so") for k in MUTANTS}


def _run_closed_forms(lib, kexpr):
    env = dict(os.environ)
    if lib:
        env["ORACLE_LIB"] = lib
    cmd = [sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
           os.path.join(ROOT, "tests", "test_oracle_closed_forms.py"), "-k", kexpr]
    return subprocess.run(cmd, cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=600)


def test_unmodified_oracle_passes_closed_forms():
    r = _run_closed_forms(None, "closed_form or geometric or scaling")
    assert r.returncode == 0, r.stdout.decode()[-2000:]


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_planted_mistake_fails_closed_forms(mutant_libs, k):
    what, kexpr = MUTANTS[k]
    r = _run_closed_forms(mutant_libs[k], kexpr)
    out = r.stdout.decode()
    assert r.returncode == 1 and " failed" in out, f"mutant {k} ({what}) was not caught:\n{out[-2000:]}"
