"""Ensembles and continuous batching through the C ABI (piafs_batch_*).

A batch advances all member cases with one launch per kernel class and RK
stage; each block reads its own case's descriptor.  The bar: every member's
state equals what the same case gives alone (bitwise -- the same device code
runs on the same data), and the reference fixtures to FP64 relative L2
<= 1e-11, whatever else is in the batch and whenever it was admitted."""
import tempfile
import os

import numpy as np
import pytest

import _oracle as orc
from conftest import build_deck
import piafs_amd
from piafs_amd import decks
from test_gpu_parity import FIXTURES, TOL_STEPS, _rel

pytestmark = pytest.mark.gpu


def test_mixed_fixture_batch_matches_reference(ref_runs):
    """All reference fixtures (1D Euler and 2D NS, WENO/CRWENO, RK4 and SSPRK3,
    different grids) in one batch; each is removed once it has taken its own
    n_iter steps (the order of the shipped runs)."""
    data, meta = ref_runs
    with tempfile.TemporaryDirectory() as tmp:
        dirs, want = [], []
        for name in FIXTURES:
            dk = build_deck(meta[name])
            d = os.path.join(tmp, name)
            dk.write(d)
            dirs.append(d)
            want.append(int(dk.solver["n_iter"]))
        with piafs_amd.Batch(dirs) as b:
            done, live = 0, set(range(len(dirs)))
            for k in sorted(live, key=lambda k: want[k]):
                if want[k] > done:
                    b.step(want[k] - done)
                    done = want[k]
                assert b.steps_done(k) == want[k]
                got = b.solution(k)
                err = _rel(got, data[FIXTURES[k]])
                assert np.all(err <= TOL_STEPS), (FIXTURES[k], err)
                b.remove(k)


def _ns2d_members():
    return [decks.ns2d_photochem(nx=24, ny=20, n_iter=8),
            decks.ns2d_photochem(nx=17, ny=31, n_iter=8, scheme="weno5", weights="z"),
            decks.ns2d_photochem(nx=40, ny=16, n_iter=8, weights="mapped", Re=300.0),
            decks.sod_1d(N=96, n_iter=8, t_final=0.02)]  # CFL about 0.3


def test_graph_replay_equals_eager_launches():
    with tempfile.TemporaryDirectory() as tmp:
        dirs = []
        for k, dk in enumerate(_ns2d_members()):
            dirs.append(os.path.join(tmp, f"m{k}"))
            dk.write(dirs[-1])
        with piafs_amd.Batch(dirs) as g, piafs_amd.Batch(dirs) as e:
            e.use_graph(False)
            for n in (1, 3, 4):   # replays after the first capture too
                g.step(n)
                e.step(n)
            g.sync()
            e.sync()
            for k in range(len(dirs)):
                assert np.array_equal(g.solution(k), e.solution(k)), k


def test_admission_and_release_are_mix_invariant():
    """Members admitted and released between steps give exactly what they give
    alone, and match the oracle at their own step counts."""
    mem = _ns2d_members()
    with tempfile.TemporaryDirectory() as tmp:
        dirs = []
        for k, dk in enumerate(mem):
            dirs.append(os.path.join(tmp, f"m{k}"))
            dk.write(dirs[-1])
        with piafs_amd.Batch(dirs[:2]) as b:
            b.step(3)
            s2 = b.add(dirs[2])
            b.step(2)
            b.remove(0)
            s3 = b.add(dirs[3])
            assert s3 == 0  # the released slot is reused
            b.step(4)
            res = {1: (b.solution(1), 9), 2: (b.solution(s2), 6), 3: (b.solution(s3), 4)}
            assert b.steps_done(1) == 9 and b.steps_done(s2) == 6 and b.steps_done(s3) == 4
        for k, (got, n) in res.items():
            with piafs_amd.Case(dirs[k]) as c:
                c.step(n)
                alone = c.solution()
            assert np.array_equal(got, alone), k
            o = orc.OCase(dirs[k])
            o.step(n)
            err = _rel(got, o.solution())
            assert np.all(err <= TOL_STEPS), (k, err)


def test_batch_of_one_equals_case():
    dk = decks.ns2d_photochem(nx=32, ny=24, n_iter=5)
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        with piafs_amd.Batch([d]) as b:
            b.step(5)
            got = b.solution(0)
            inf = b.info(0)
        with piafs_amd.Case(d) as c:
            c.step(5)
            alone = c.solution()
    assert np.array_equal(got, alone)
    assert inf.dim[0] == 32 and inf.dim[1] == 24 and inf.nvars == 11


def test_batch_create_reports_bad_member():
    with tempfile.TemporaryDirectory() as tmp:
        good = os.path.join(tmp, "good")
        decks.sod_1d(N=32, n_iter=1).write(good)
        with pytest.raises(piafs_amd.PiafsError, match="case 1"):
            piafs_amd.Batch([good, os.path.join(tmp, "missing")])


def test_autotune_keeps_results_bitwise():
    """The segment length the autotuner picks changes only the speed: the line
    solves are exact for every team shape."""
    with tempfile.TemporaryDirectory() as tmp:
        dirs = []
        for k, dk in enumerate(_ns2d_members()):
            dirs.append(os.path.join(tmp, f"m{k}"))
            dk.write(dirs[-1])
        with piafs_amd.Batch(dirs) as a, piafs_amd.Batch(dirs) as b:
            rows, ms = b.autotune()
            assert rows in (4, 8, 12, 16) and ms > 0.0
            a.step(6)
            b.step(6)
            for k in range(len(dirs)):
                assert np.array_equal(a.solution(k), b.solution(k)), k
        with piafs_amd.Case(dirs[0]) as c:
            rows, ms = c.autotune()
            c.step(8)
            tuned = c.solution()
        with piafs_amd.Case(dirs[0]) as c:
            c.step(8)
            assert np.array_equal(tuned, c.solution())


def test_batch_with_species_fluxes_in_the_reconstruction():
    """A member whose lines take 64-lane teams in both sweeps (600 x 560, NS2D
    with chemistry, paired Roe), where the reconstruction writes the species'
    fluxes itself, batched with members that do not (a 48 x 40 case of 16-lane
    teams and a WENO5 case): every member equals its solo run bit for bit."""
    periodic = [("periodic", 0, 1, 0, 0, -1e99, 1e99), ("periodic", 0, -1, 0, 0, -1e99, 1e99),
                ("periodic", 1, 1, -1e99, 1e99, 0, 0), ("periodic", 1, -1, -1e99, 1e99, 0, 0)]
    mem = [decks.ns2d_open_box(nx=600, ny=560, n_iter=3, dt=5e-4, flow=(1.0, 0.3, 0.2, None), scheme="crweno5",
                               include_chemistry=True, Re=1000.0, boundary=periodic),
           decks.ns2d_photochem(nx=48, ny=40, n_iter=3),
           decks.ns2d_open_box(nx=64, ny=600, n_iter=3, dt=5e-4, flow=(1.0, 0.3, 0.2, None), scheme="weno5",
                               include_chemistry=True, Re=1000.0, boundary=periodic)]
    with tempfile.TemporaryDirectory() as tmp:
        dirs = []
        for k, dk in enumerate(mem):
            dirs.append(os.path.join(tmp, f"m{k}"))
            dk.write(dirs[-1])
        with piafs_amd.Batch(dirs) as b:
            b.step(3)
            got = [b.solution(k) for k in range(len(dirs))]
        for k, d in enumerate(dirs):
            with piafs_amd.Case(d) as c:
                c.step(3)
                assert np.array_equal(got[k], c.solution()), k
