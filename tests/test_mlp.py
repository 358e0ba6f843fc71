"""The result-size predictor on the CPU: the oracle's backpropagation is
pinned by central finite differences (and a planted mistake shows the
check has teeth); features and sizing helpers on a small table."""
from __future__ import annotations

import numpy as np

import virginian_amd as va
from oracle import features as FE
from oracle import mlp as M
from tests.helpers import CONFIG1_SQL, QUERY_SUITE
from virginian_amd import predict as PR
from virginian_amd import workload as W


def _fd(p, x, t, nin, nh, eps=1e-6):
    g = np.zeros_like(p)
    for i in range(len(p)):
        a, b = p.copy(), p.copy()
        a[i] += eps
        b[i] -= eps
        g[i] = (M.loss(a, x, t, nin, nh) - M.loss(b, x, t, nin, nh)) / (2 * eps)
    return g


def _case(seed, nin=5, nh=4, B=7):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.7, M.layout(nin, nh)["n"])
    x = rng.normal(0, 1, (B, nin))
    t = rng.uniform(0, 1, B)
    return p, x, t, nin, nh


def test_oracle_gradient_matches_finite_differences():
    for seed in range(4):
        p, x, t, nin, nh = _case(seed)
        _, g = M.gradient(p, x, t, nin, nh)
        fd = _fd(p, x, t, nin, nh)
        assert np.allclose(g, fd, rtol=1e-5, atol=1e-8), (seed, np.abs(g - fd).max())


def test_oracle_absolute_error_gradient_matches_finite_differences():
    """Mean absolute error: away from y = t the loss is smooth and the
    gradient is backpropagated sign(y - t)."""
    for seed in range(4):
        p, x, t, nin, nh = _case(seed + 20)
        y, _ = M.forward(p, x, nin, nh)
        assert np.abs(y - t).min() > 1e-4
        _, g = M.gradient(p, x, t, nin, nh, 0.0, True)
        fd = np.zeros_like(p)
        for i in range(len(p)):
            a, b = p.copy(), p.copy()
            a[i] += 1e-6
            b[i] -= 1e-6
            fd[i] = (M.loss(a, x, t, nin, nh, 0.0, True) - M.loss(b, x, t, nin, nh, 0.0, True)) / 2e-6
        assert np.allclose(g, fd, rtol=1e-5, atol=1e-8), (seed, np.abs(g - fd).max())
        # and it is not the squared error's gradient
        _, g2 = M.gradient(p, x, t, nin, nh)
        assert not np.allclose(g, g2, rtol=1e-3)


def test_oracle_pinball_gradient_matches_finite_differences():
    """Pinball loss of quantile tau: away from y = t the gradient is
    backpropagated 1 - tau (above) / -tau (below); tau = 1/2 is half the
    absolute error; and at zero weights (y = 1/2) only b3 moves, by the
    closed form mean(dl) / 4."""
    for seed, tau in zip(range(4), (0.1, 0.5, 0.9, 0.97)):
        p, x, t, nin, nh = _case(seed + 40)
        y, _ = M.forward(p, x, nin, nh)
        assert np.abs(y - t).min() > 1e-4
        _, g = M.gradient(p, x, t, nin, nh, 0.0, tau=tau)
        fd = np.zeros_like(p)
        for i in range(len(p)):
            a, b = p.copy(), p.copy()
            a[i] += 1e-6
            b[i] -= 1e-6
            fd[i] = (M.loss(a, x, t, nin, nh, tau=tau) - M.loss(b, x, t, nin, nh, tau=tau)) / 2e-6
        assert np.allclose(g, fd, rtol=1e-5, atol=1e-8), (seed, tau, np.abs(g - fd).max())
    p, x, t, nin, nh = _case(45)
    assert np.isclose(M.loss(p, x, t, nin, nh, tau=0.5), 0.5 * M.loss(p, x, t, nin, nh, absolute=True))
    nin, nh = 3, 2
    p = np.zeros(M.layout(nin, nh)["n"])
    x = np.ones((4, nin))
    t = np.array([0.0, 1.0, 0.25, 0.75])
    loss, g = M.gradient(p, x, t, nin, nh, tau=0.8)
    assert np.isclose(loss, np.mean([0.2 * 0.5, 0.8 * 0.5, 0.2 * 0.25, 0.8 * 0.25]))
    L = M.layout(nin, nh)
    assert np.isclose(g[L["b3"][0]], np.mean([0.2, -0.8, 0.2, -0.8]) * 0.25)
    assert np.count_nonzero(g) == 1


def test_finite_difference_check_catches_a_dropped_relu_mask():
    p, x, t, nin, nh = _case(11)
    w = M.unpack(p, nin, nh)
    y, (h1, h2) = M.forward(p, x, nin, nh)
    d3 = (2.0 * (y - t) / len(t) * y * (1 - y))[:, None]
    d2_wrong = d3 @ w["W3"].T                  # the ReLU mask left out
    gW2_wrong = h1.T @ d2_wrong
    L = M.layout(nin, nh)
    fd = _fd(p, x, t, nin, nh)[L["W2"][0]:L["W2"][1]]
    assert not np.allclose(gW2_wrong.ravel(), fd, rtol=1e-3, atol=1e-6)


def test_oracle_closed_form_at_zero_weights():
    nin, nh = 3, 2
    p = np.zeros(M.layout(nin, nh)["n"])
    x = np.ones((4, nin))
    t = np.array([0.0, 1.0, 0.25, 0.5])
    y, _ = M.forward(p, x, nin, nh)
    assert np.allclose(y, 0.5)                 # sigmoid(0)
    loss, g = M.gradient(p, x, t, nin, nh)
    assert np.isclose(loss, np.mean((0.5 - t) ** 2))
    L = M.layout(nin, nh)
    # only b3 has a gradient: mean 2 (y - t) y (1 - y)
    assert np.isclose(g[L["b3"][0]], np.mean(2 * (0.5 - t) * 0.25))
    assert np.count_nonzero(g) == 1


def test_features_and_estimates(db):
    db.gen_table("test", 300_000, 1 << 16, seed=3)
    tid = db.table_id("test")
    st = PR.TableStats(db, tid)
    assert len(st.rows) == 5 and st.rows.sum() == 300_000
    # uniformi < 50: half the rows on every tablet
    est = FE.passing(st, 0, va.OP_Lt, 50.0)
    assert np.all(np.abs(est - 0.5) < 0.02)
    # uniformf > 90 (uniform on [0, 100))
    est = FE.passing(st, 3, va.OP_Gt, 90.0)
    assert np.all(np.abs(est - 0.1) < 0.02)
    vm = db.compile(CONFIG1_SQL)
    try:
        ops = va.program(vm)
    finally:
        db.free_vm(vm)
    f = PR.program_features(ops, st)
    assert f.shape == (5, PR.NFEATURES) and f.dtype == np.float32
    assert np.all(np.isfinite(f))
    # two compares, both on a bare column against a constant
    assert f[0, 4] == 2 / 4 and f[0, 5] == 0
    # uniformi < 50 and normali5 > 0: 0.5 * P(N(0,5) rounds above 0) = 0.23
    assert np.all(np.abs(f[:, 3] - 0.23) < 0.03), f[:, 3]
    # under independence the AND keeps the product of its parts, per tablet
    both = FE.passing(st, 0, va.OP_Lt, 50.0) * FE.passing(st, 1, va.OP_Gt, 0.0)
    assert np.allclose(f[:, 3], both, rtol=1e-6, atol=1e-7), (f[:, 3], both)


def test_features_match_the_restatement(db):
    """vg_program_features (C) against oracle/features.py on seeded random
    programs over a table of several tablets, and on the published queries."""
    db.gen_table("test", 200_000, 1 << 15, seed=4)
    tid = db.table_id("test")
    st = PR.TableStats(db, tid)
    rng = np.random.default_rng(17)
    progs = [W.random_program(rng, max_ops=40, selectivity=float(rng.uniform(0.0001, 0.8)) if k % 3 else None,
                              table=tid) for k in range(300)]
    # constant first: the compare is read mirrored (50 < uniformi is uniformi > 50)
    mirrored = ["select id from test where 50 < uniformi", "select id from test where 50.0 > uniformf",
                "select id from test where 10 <= normali20 and 3 = uniformi",
                "select id from test where -2.5 >= normalf5 or 7 != uniformi"]
    for sql in QUERY_SUITE + [CONFIG1_SQL] + mirrored:
        vm = db.compile(sql)
        try:
            progs.append(va.program(vm))
        finally:
            db.free_vm(vm)
    for k, ops in enumerate(progs):
        got = PR.program_features(ops, st)
        want = FE.program_features(ops, st)
        assert got.shape == want.shape == (len(st.rows), PR.NFEATURES)
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-6, err_msg=f"program {k}")


def test_capacity_and_calibration():
    rows = np.array([1000.0, 1000.0])
    assert PR.capacity(np.array([0.1, 0.2]), rows, 0.0) == 300 + 64
    assert PR.capacity(np.array([0.9, 0.9]), rows, 0.5) == 2000 + 64
    preds = [np.array([0.1]), np.array([0.1])]
    m = PR.calibrate(preds, [100.0, 300.0], [np.array([1000.0])] * 2, 0.0)
    assert PR.capacity(preds[1], np.array([1000.0]), m) >= 300


def test_oracle_skip_gradient_matches_finite_differences():
    p, x, t, nin, nh = _case(21)
    _, g = M.gradient(p, x, t, nin, nh, 2.5)
    eps = 1e-6
    fd = np.zeros_like(p)
    for i in range(len(p)):
        a, b = p.copy(), p.copy()
        a[i] += eps
        b[i] -= eps
        fd[i] = (M.loss(a, x, t, nin, nh, 2.5) - M.loss(b, x, t, nin, nh, 2.5)) / (2 * eps)
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-8)
