"""Files a run writes, against the reference's own output on the same decks.

tests/golden/outputs/<case>/ holds every file the reference PIAFS wrote for a
small deck (solution files in binary, text and Tecplot layouts with the
indexed names of op_overwrite "no", species files, a restart run,
errors.dat against exact.inp, conservation.dat with and without the
conservation check, function_counts.dat).  The driver piafs_amd runs the same
decks on the GPU; the test checks the set of file names, every header and
index column exactly, and the numbers to the parity tolerance (the runtimes
in errors.dat are the only fields not compared)."""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from piafs_amd import decks

pytestmark = pytest.mark.gpu

OUT = os.path.join(GOLDEN, "outputs")
META = json.load(open(os.path.join(OUT, "outputs.json")))
CLI = os.path.join(ROOT, "piafs_amd", "bin", "piafs_amd")
RTOL = 1e-11   # FP64 parity bar for the state, relative to each column's scale


def _run_cli(name, d):
    m = META[name]
    dk = getattr(decks, m["builder"])(**m["kwargs"])
    dk.solver.update(m["solver"])
    dk.write(d)
    if m["exact"]:
        shutil.copy(os.path.join(d, "initial.inp"), os.path.join(d, "exact.inp"))
    before = set(os.listdir(d))
    r = subprocess.run([CLI, d], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return sorted(set(os.listdir(d)) - before), r.stdout


def _close_cols(ours, ref, what):
    """columns of floats: each column to RTOL of its largest magnitude"""
    assert ours.shape == ref.shape, what
    scale = np.maximum(np.abs(ref).max(axis=0), 1e-300)
    err = np.abs(ours - ref).max(axis=0) / scale
    assert np.all(err <= RTOL), (what, err)


def _compare_text(a, b, fn):
    la, lb = open(a).read().splitlines(), open(b).read().splitlines()
    assert len(la) == len(lb), fn
    hdr = 2 if lb and lb[0].startswith("VARIABLES") else 0
    assert la[:hdr] == lb[:hdr], fn
    ra = [r.split() for r in la[hdr:]]
    rb = [r.split() for r in lb[hdr:]]
    nint = sum(1 for t in rb[0] if "E" not in t)
    assert [r[:nint] for r in ra] == [r[:nint] for r in rb], fn          # the index columns
    fa = np.array([[float(t) for t in r[nint:]] for r in ra])
    fb = np.array([[float(t) for t in r[nint:]] for r in rb])
    _close_cols(fa, fb, fn)


def _compare_binary(a, b, fn):
    nda, nva, da, xa, ua = decks.read_binary_op(a)
    ndb, nvb, db, xb, ub = decks.read_binary_op(b)
    assert (nda, nva, list(da)) == (ndb, nvb, list(db)), fn
    np.testing.assert_array_equal(xa, xb)
    _close_cols(ua, ub, fn)


def _compare_run_file(a, b, fn, ndims):
    ta, tb = open(a).read().split(), open(b).read().split()
    assert len(ta) == len(tb), fn
    if fn == "function_counts.dat":
        assert ta == tb
        return
    head = 2 * ndims + 1                       # sizes, process counts, dt
    assert ta[:head] == tb[:head], fn
    va = np.array([float(t) for t in ta[head:]])
    vb = np.array([float(t) for t in tb[head:]])
    if fn == "errors.dat":
        va, vb = va[:3], vb[:3]                # L1, L2, Linf (then two runtimes)
        np.testing.assert_allclose(va, vb, rtol=1e-8, err_msg=fn)
    else:                                      # conservation errors: -1 unchecked, else round-off
        np.testing.assert_allclose(va, vb, rtol=0, atol=1e-12, err_msg=fn)


@pytest.mark.parametrize("name", sorted(META))
def test_run_writes_the_reference_files(name):
    m = META[name]
    with tempfile.TemporaryDirectory() as d:
        made, log = _run_cli(name, d)
        assert made == sorted(m["files"]), (made, m["files"])
        ndims = 2 if m["builder"] != "sod_1d" else 1
        for fn in made:
            a, b = os.path.join(d, fn), os.path.join(OUT, name, fn)
            if fn in ("errors.dat", "conservation.dat", "function_counts.dat"):
                _compare_run_file(a, b, fn, ndims)
            elif fn.endswith(".bin"):
                _compare_binary(a, b, fn)
            else:
                _compare_text(a, b, fn)
    if m["exact"]:
        assert "Computed errors for domain 0:" in log
    if m["solver"].get("conservation_check") == "yes":
        assert "Conservation Errors:" in log
