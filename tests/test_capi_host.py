"""CPU tests of libpf.so: every entry point declared in include/pf.h is
exported, and the host-side helpers (ran_2, k_space, gamma, initial conditions)
agree with the oracle and the reference fixtures.  No GPU calls."""
import json
import os
import re

import numpy as np
import pytest

import oracle
from oracle import ics

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "pf.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pf_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported(pf):
    L = pf.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_python_binding_covers_header(pf):
    assert set(declared_symbols()) <= set(pf.EXPORTED)


def test_ran2_matches_reference_fixture(pf):
    fx = json.load(open(os.path.join(GOLD, "ran2_ref.json")))
    for seed, d in fx.items():
        r = pf.Ran2(int(seed))
        assert [r() for _ in range(len(d["values"]))] == d["values"]


def test_k_space_bit_exact(pf):
    f = np.load(os.path.join(GOLD, "kspace_ref.npz"))
    nx, ny, nz = int(f["nx"]), int(f["ny"]), int(f["nz"])
    kf, ksq, k4 = pf.k_space(nx, ny, nz)
    for mine, ref in ((kf, f["kf"]), (ksq, f["kf_sq"]), (k4, f["kf_4"])):
        assert np.array_equal(np.ascontiguousarray(mine).view(np.float64).ravel(), ref)


def test_gamma(pf):
    assert np.array_equal(pf.gamma(7, 1.5, 0.25), ics.gamma(7, 1.5, 0.25))


def test_initial_conditions_match_oracle(pf):
    a = pf.ic_random(8, 12, 16, 3, -5)
    b = oracle.ic_random(8, 12, 16, 3, -5)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    a = pf.ic_voronoi(16, 12, 8, 5, 9, -3)
    b = oracle.ic_voronoi(16, 12, 8, 5, 9, -3)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_errors_are_reported_not_crashing(pf):
    with pytest.raises(pf.PfError) as e:
        pf.Simulation(pf.make_params(), 10, 8, 8, 3)  # 10 is not a supported length
    assert e.value.code == pf.PF_EINVAL
    with pytest.raises(pf.PfError):
        pf.k_space(1, 8, 8)


def test_missing_library_fails_loudly(pf, monkeypatch):
    """Without libpf.so there is no CPU path: every entry point raises."""
    monkeypatch.setattr(pf, "_lib", None)
    monkeypatch.setattr(pf, "LIB_PATH", os.path.join(ROOT, "no_such_dir", "libpf.so"))
    with pytest.raises(ImportError):
        pf.lib()
    with pytest.raises(ImportError):
        pf.Simulation(pf.make_params(), 16, 16, 16, 3, "fp32_mixed")
    with pytest.raises(ImportError):
        pf.k_space(8, 8, 8)
