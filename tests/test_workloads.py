"""Synthetic workloads: the counter-based SplitMix64 stream reproduces the
published SplitMix64 sequence (seed 0: e220a8397b1dcdaf, 6e789e6aa1b965f4,
06c45d188009454f), and every generator satisfies its stated invariants."""
import math

import numpy as np

from briscola_amd import workloads as W


def test_splitmix64_known_answers():
    G = W.GOLDEN
    with np.errstate(over="ignore"):
        got = [int(W._mix64(np.uint64(i) * G)) for i in (1, 2, 3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_streams_are_counter_based():
    a = W.uniform_pm1(2, 7, 1, 100)
    b = W.uniform_pm1(2, 7, 1, 40, start=60)
    assert np.array_equal(a[60:], b)
    assert a.min() >= -1.0 and a.max() < 1.0
    assert not np.array_equal(a, W.uniform_pm1(2, 8, 1, 100))


def test_uniform_neumann_consistent():
    p = W.uniform_neumann(16, seed=1)
    assert p.b.shape == (18, 18, 18)
    assert abs(p.b.sum()) < 1e-15
    assert np.all(p.b[0] == 0) and np.all(p.b[:, :, -1] == 0)


def test_channel_mesh_and_rhs():
    p = W.channel(32, seed=0)
    vx, vy, vz = p.vertices
    assert np.isclose(vx[-1], 4 * math.pi) and np.isclose(vz[-1], 2 * math.pi)
    assert vy[0] == -1.0 and vy[-1] == 1.0
    dy = np.diff(vy)
    assert dy[0] < dy[len(dy) // 2] / 5          # stretched towards the walls
    assert p.bc == (W.PERIODIC, W.PERIODIC, W.NEUMANN, W.NEUMANN, W.PERIODIC, W.PERIODIC)
    assert abs(p.b.sum()) < 1e-14 * np.abs(p.b).sum()
    assert np.array_equal(p.b, W.channel(32, seed=0).b)


def test_spheres_coefficients_and_rhs():
    p = W.spheres(24, seed=3, nspheres=6)
    lam = p.lam[0][1:-1, 1:-1, 1:-1]
    assert lam.min() >= 1.0 / 1000 - 1e-15 and lam.max() <= 1.0 + 1e-15
    assert lam.max() > 0.5                      # some faces inside spheres
    assert abs(p.b.sum()) < 1e-14 * np.abs(p.b).sum()
