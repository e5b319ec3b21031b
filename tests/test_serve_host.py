"""Host side of the serving loop: the arrival trace and template snapping."""

import math

import numpy as np


def test_trace_is_poisson_loguniform():
    from kcfd_amd import serve as S
    tr = S.make_trace(rate=200.0, duration=60.0, seed=5000)
    t = np.array([x[0] for x in tr])
    n = np.array([x[2] for x in tr])
    assert abs(len(tr) - 12000) < 4 * math.sqrt(12000)
    assert np.all(np.diff(t) > 0) and t[-1] < 60.0
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 200) < 0.05 / 200 * 4
    assert abs(np.std(gaps) / gaps.mean() - 1.0) < 0.05          # exponential: CV = 1
    assert n.min() >= 1000 and n.max() <= 1_000_000
    assert abs(np.log(n).mean() - 0.5 * (math.log(1e3) + math.log(1e6))) < 0.1
    frac = [sum(x[1] == k for x in tr) / len(tr) for k in S.KINDS]
    assert abs(frac[0] - 0.5) < 0.03 and abs(frac[1] - 0.3) < 0.03 and abs(frac[2] - 0.2) < 0.03
    assert S.make_trace(rate=200.0, duration=1.0, seed=5000) == [x for x in tr if x[0] < 1.0]


def test_template_sizes_and_snapping():
    from kcfd_amd import serve as S
    sz = S.template_sizes(per_kind=7)
    assert sz[0] == 1000 and sz[-1] == 1_000_000 and len(sz) == 7
    assert all(b > a for a, b in zip(sz, sz[1:]))
    assert S.nearest_template(sz, 1000) == 0 and S.nearest_template(sz, 10 ** 6) == 6
    assert S.nearest_template(sz, 31623) == 3


def test_template_cases_share_a_mesh():
    from kcfd_amd import serve as S
    for kind in S.KINDS:
        p, m = S.build_template_cases(kind, 1500, 11)
        assert p.mesh is m.mesh
        assert p.recipe[0][0] == "pressure"
