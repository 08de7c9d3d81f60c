"""Serving through the C ABI (piafs_server_*): requests that arrive while
others run are admitted into the resident ensembles of their class, and each
result is exactly the result of the same deck run alone (bitwise; the same
device code runs on the same data) and matches the oracle to FP64 relative
L2 <= 1e-11."""
import math
import os
import tempfile
import time

import numpy as np
import pytest

import _oracle as orc
import piafs_amd
from piafs_amd import decks
from test_gpu_parity import TOL_STEPS, _rel

pytestmark = pytest.mark.gpu


def _mix():
    return [decks.ns2d_photochem(nx=32, ny=24, n_iter=12),
            decks.sod_1d(N=200, n_iter=60, t_final=0.2 * 60 / 176),
            decks.ns2d_photochem(nx=20, ny=40, n_iter=7, scheme="weno5", weights="z"),
            decks.sod_1d(N=96, n_iter=30, t_final=0.2 * 30 / 84, interp="components", scheme="crweno5"),
            decks.ns2d_photochem(nx=48, ny=16, n_iter=20, weights="mapped", Re=300.0)]


def test_served_results_equal_solo_runs():
    with tempfile.TemporaryDirectory() as tmp:
        dirs = []
        for k, dk in enumerate(_mix()):
            dirs.append(os.path.join(tmp, f"r{k}"))
            dk.write(dirs[-1])
        with piafs_amd.Server() as srv:
            ids = []
            for k, d in enumerate(dirs):
                ids.append(srv.submit(d))
                time.sleep(0.002 * k)   # staggered arrivals join running ensembles
            srv.wait_all()
            got = []
            for i in ids:
                st, lat = srv.poll(i)
                assert st == srv.DONE and lat >= 0.0
                got.append(srv.result(i))
            assert srv.stats()["running"] == 0
        for k, d in enumerate(dirs):
            with piafs_amd.Case(d) as c:
                c.step(c.info.n_iter)
                alone = c.solution()
            assert np.array_equal(got[k], alone), k
            o = orc.OCase(d)
            o.step(o.n_iter if hasattr(o, "n_iter") else c_iters(d))
            assert np.all(_rel(got[k], o.solution()) <= TOL_STEPS), k


def c_iters(d):
    for line in open(os.path.join(d, "solver.inp")):
        t = line.split()
        if t and t[0] == "n_iter":
            return int(t[1])
    raise AssertionError("n_iter missing")


def test_budget_admission_serialises_when_tight():
    """With a budget that holds one request at a time the requests still all
    complete, one after another, with the same results."""
    mix = _mix()[:3]
    with tempfile.TemporaryDirectory() as tmp:
        dirs = []
        for k, dk in enumerate(mix):
            dirs.append(os.path.join(tmp, f"r{k}"))
            dk.write(dirs[-1])
        need = max(piafs_amd.estimate_bytes(d) for d in dirs)
        with piafs_amd.Server(hbm_budget=need) as srv:
            ids = [srv.submit(d) for d in dirs]
            srv.wait_all()
            res = [srv.result(i) for i in ids]
        for d, r in zip(dirs, res):
            with piafs_amd.Case(d) as c:
                c.step(c.info.n_iter)
                assert np.array_equal(r, c.solution())


def test_oversized_request_is_refused():
    with tempfile.TemporaryDirectory() as tmp:
        decks.sod_1d(N=64, n_iter=2).write(tmp)
        with piafs_amd.Server(hbm_budget=1024) as srv:
            with pytest.raises(piafs_amd.PiafsError, match="budget"):
                srv.submit(tmp)


def test_both_lanes_busy_with_many_admissions():
    """Both lanes (1D and 2D) admit and release many requests while the other
    lane is stepping: plan rebuilds and graph captures of one lane interleave
    with the other lane's rounds.  Every result equals its solo run."""
    mix = []
    for k in range(12):
        if k % 2:
            n = 48 + 16 * k
            mix.append(decks.sod_1d(N=n, n_iter=10 + 3 * k, t_final=0.2 * (10 + 3 * k) / math.ceil(0.88 * (n - 1))))
        else:
            mix.append(decks.ns2d_photochem(nx=16 + 4 * k, ny=20 + 2 * k, n_iter=4 + k,
                                            scheme="crweno5" if k % 4 == 0 else "weno5"))
    with tempfile.TemporaryDirectory() as tmp:
        dirs = []
        for k, dk in enumerate(mix):
            dirs.append(os.path.join(tmp, f"q{k}"))
            dk.write(dirs[-1])
        with piafs_amd.Server() as srv:
            ids = []
            for d in dirs:
                ids.append(srv.submit(d))
                time.sleep(0.001)
            srv.wait_all()
            got = []
            for i in ids:
                st, _ = srv.poll(i)
                assert st == srv.DONE, srv._err()
                got.append(srv.result(i))
        for k, d in enumerate(dirs):
            with piafs_amd.Case(d) as c:
                c.step(c.info.n_iter)
                assert np.array_equal(got[k], c.solution()), k
