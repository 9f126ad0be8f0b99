"""bct.dat reading and interpolation (crimsonflowsolver_amd.bct) on the
reference's basic case (tests/golden/basic_bct.npz).

The pin is the reference's own output: restart.5.1 of its restart test, the
state after 5 steps of 0.001 on the same mesh and bct.dat, holds at the 26
prescribed inlet nodes exactly the velocities BCint gives at t = 0.005.  The
other checks follow from linear interpolation and the period.
"""
import os

import numpy as np
import pytest

from crimsonflowsolver_amd import bct

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def case():
    g = np.load(os.path.join(HERE, "golden", "basic_case.npz"))
    b = np.load(os.path.join(HERE, "golden", "basic_bct.npz"))
    return g, b


def _write(path, xs, us, t):
    with open(path, "w") as f:
        f.write(f"{len(xs)} {len(t)}\n")
        for x, u in zip(xs, us):
            f.write(f"{x[0]:14.8f} {x[1]:14.8f} {x[2]:14.8f} {len(t)}\n")
            for k in range(len(t)):
                f.write(f"{u[k, 0]:14.8f} {u[k, 1]:14.8f} {u[k, 2]:14.8f} {t[k]:14.8f}\n")


@pytest.fixture(scope="module")
def records(case, tmp_path_factory):
    _, b = case
    p = tmp_path_factory.mktemp("bct") / "bct.dat"
    _write(p, b["x"], b["u"], b["t"])
    return bct.read_bct(p)


def test_read_round_trip(case, records):
    _, b = case
    assert len(records) == 26
    for k, (x, tab) in enumerate(records):
        assert np.array_equal(x, b["x"][k])
        assert np.array_equal(tab[:, :3], b["u"][k]) and np.array_equal(tab[:, 3], b["t"])


def test_attaches_every_record_to_a_prescribed_inlet_node(case, records):
    g, b = case
    bt = bct.BCt(records, g["x"], g["iBC"])
    assert np.array_equal(bt.nodes, b["u5_nodes"])
    assert np.all(((g["iBC"][bt.nodes] >> 3) & 7) == 7)
    d = g["x"][bt.nodes] - b["x"]
    assert np.all(np.sum(d * d, axis=1) < bct.EPSD)


def test_velocities_at_step_5_equal_the_reference_restart(case, records):
    g, b = case
    bt = bct.BCt(records, g["x"], g["iBC"])
    BC = np.full((g["x"].shape[0], 7), 9.0)
    bt.apply(int(b["step5"]) * 0.001, BC)
    assert np.array_equal(BC[b["u5_nodes"], 2:5], b["u5"])
    rest = np.setdiff1d(np.arange(BC.shape[0]), bt.nodes)
    assert np.all(BC[rest, 2:5] == 0.0)  # no mesh motion: the other nodes' velocity columns are cleared
    assert np.all(BC[:, [0, 1, 5, 6]] == 9.0)  # other columns untouched
    # a neighbouring time gives something else
    BC4 = bt.apply(0.004, np.zeros_like(BC))
    assert np.abs(BC4[b["u5_nodes"], 2:5] - b["u5"]).max() > 0.1


def test_linear_between_samples_and_periodic(case, records):
    g, b = case
    bt = bct.BCt(records, g["x"], g["iBC"])
    u, t = b["u"], b["t"]
    n = np.zeros((g["x"].shape[0], 7))
    for j, wr in ((37, 0.25), (400, 0.5), (1099, 0.9)):
        tt = t[j] + wr * (t[j + 1] - t[j])
        got = bt.apply(tt, n.copy())[bt.nodes, 2:5]
        want = (1.0 - wr) * u[:, j] + wr * u[:, j + 1]
        assert np.allclose(got, want, rtol=0, atol=1e-9 * np.abs(u).max()), j
        T = t[-1]
        again = bt.apply(tt + 3 * T, n.copy())[bt.nodes, 2:5]
        assert np.allclose(again, want, rtol=0, atol=1e-9 * np.abs(u).max()), j


def test_time_scale_and_unmatched_records(case, records):
    g, b = case
    scaled = bct.BCt(records, g["x"], g["iBC"], time_scale=2.0)
    plain = bct.BCt(records, g["x"], g["iBC"])
    n = np.zeros((g["x"].shape[0], 7))
    assert np.array_equal(scaled.apply(0.010, n.copy()), plain.apply(0.005, n.copy()))
    far = [(x + 100.0, tab) for x, tab in records[:3]] + records[3:]
    assert bct.BCt(far, g["x"], g["iBC"]).nodes.size == 23
    # a node whose velocity is not fully prescribed is never attached
    ibc = g["iBC"].copy()
    ibc[b["u5_nodes"][0]] = 8 | 16
    assert b["u5_nodes"][0] not in bct.BCt(records, g["x"], ibc).nodes


def test_truncated_file_is_refused(tmp_path):
    p = tmp_path / "bct.dat"
    p.write_text("1 3\n0 0 0 3\n1 2 3 0.0\n1 2 3 0.1\n")
    with pytest.raises(ValueError):
        bct.read_bct(p)
