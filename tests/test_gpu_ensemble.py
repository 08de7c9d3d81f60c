"""Ensemble directories (simulation.inp with nsims) through the C ABI
(piafs_ensemble_create / piafs_ensemble_solve) and the drop-in executable,
against the files the reference writes for the same directory
(tests/golden/ensembles/, made by make_golden.py --ensembles with the
reference binary built from source):

* the same file names: op_<i>[_NNNNN].bin, op_species_z<k>_<i>_.bin,
  errors_<i>.dat, conservation_<i>.dat, function_counts_<i>.dat
  (OutputSolution.cpp:42-62, ChemistryWriteSpecies.c:44-50,
  SimulationWriteErrors.c:39-61);
* every solution and species file to 1e-12 (Euler) / 1e-11 (NS with
  chemistry) of each variable's largest magnitude, the grids exactly;
* conservation errors to 1e-11, function counts exactly, the exact-solution
  errors of errors_<i>.dat (none: -1) exactly;
* each member's final solution equal, bit for bit, to the same deck run alone;
* the executable writes the same files, byte for byte, as the C ABI run."""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
import piafs_amd
from piafs_amd import decks

pytestmark = pytest.mark.gpu

META = json.load(open(os.path.join(GOLDEN, "ensembles", "ensembles.json")))
TOL = {"ens_sod_crweno_comp_3": 1e-12, "ens_ns2d_chem_2": 1e-11}


def _members(name):
    m = META[name]
    out = []
    for fn, kw in m["members"]:
        dk = getattr(decks, fn)(**kw)
        dk.solver.update(m["solver"])
        out.append(dk)
    return out


def _rel(got, ref):
    scale = np.abs(ref).max(axis=0)
    err = np.abs(got - ref).max(axis=0)
    return np.where(scale > 0, err / np.where(scale > 0, scale, 1.0), err)


def _compare_dir(name, d):
    ref_dir = os.path.join(GOLDEN, "ensembles", name)
    files = META[name]["files"]
    made = sorted(f for f in os.listdir(d) if f.endswith((".bin", ".dat")))
    assert made == sorted(files)
    for fn in files:
        got, ref = os.path.join(d, fn), os.path.join(ref_dir, fn)
        if fn.endswith(".bin"):
            g, r = decks.read_binary_op(got), decks.read_binary_op(ref)
            assert g[:3] == r[:3], fn
            np.testing.assert_array_equal(g[3], r[3])
            assert np.all(_rel(g[4], r[4]) <= TOL[name]), (fn, _rel(g[4], r[4]))
        elif fn.startswith("conservation"):
            g, r = np.loadtxt(got), np.loadtxt(ref)
            np.testing.assert_allclose(g, r, rtol=0, atol=1e-11)
        elif fn.startswith("function_counts"):
            assert open(got).read().split() == open(ref).read().split()
        elif fn.startswith("errors"):
            g, r = np.loadtxt(got), np.loadtxt(ref)
            np.testing.assert_array_equal(g[:-2], r[:-2])  # all but the two wall times


@pytest.mark.parametrize("name", sorted(META))
def test_ensemble_directory_matches_reference(name):
    members = _members(name)
    with tempfile.TemporaryDirectory() as d:
        decks.write_ensemble(d, members)
        with piafs_amd.Batch.ensemble(d) as b:
            assert b.slots == list(range(len(members)))
            for k, dk in enumerate(members):
                inf = b.info(k)
                assert [inf.dim[i] for i in range(inf.ndims)] == dk.dims
            b.solve(d)
            finals = [b.solution(k) for k in range(len(members))]
        _compare_dir(name, d)
    # every member as it runs alone
    for k, dk in enumerate(members):
        with tempfile.TemporaryDirectory() as d:
            dk.write(d)
            with piafs_amd.Case(d) as c:
                c.step(int(dk.solver["n_iter"]))
                assert np.array_equal(c.solution(), finals[k]), k


def test_ensemble_executable_writes_the_same_files():
    name = "ens_sod_crweno_comp_3"
    members = _members(name)
    exe = os.path.join(ROOT, "piafs_amd", "bin", "piafs_amd")
    with tempfile.TemporaryDirectory() as d1, tempfile.TemporaryDirectory() as d2:
        decks.write_ensemble(d1, members)
        decks.write_ensemble(d2, members)
        with piafs_amd.Batch.ensemble(d1) as b:
            b.solve(d1)
        r = subprocess.run([exe], cwd=d2, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        assert "-- Ensemble Simulation --" in r.stdout
        for fn in META[name]["files"]:
            if fn.startswith("errors"):
                continue  # wall times
            assert open(os.path.join(d1, fn), "rb").read() == open(os.path.join(d2, fn), "rb").read(), fn
        _compare_dir(name, d2)


def test_ensemble_refuses_a_changed_member_set():
    members = _members("ens_sod_crweno_comp_3")
    with tempfile.TemporaryDirectory() as d:
        decks.write_ensemble(d, members)
        with piafs_amd.Batch.ensemble(d) as b:
            b.remove(1)
            with pytest.raises(piafs_amd.PiafsError):
                b.solve(d)
