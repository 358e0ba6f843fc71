"""The result-size predictor on the GPU: the HIP forward and backward
kernels against the CPU oracle (tolerance for f32 sums of up to a few
hundred terms: 1e-4 relative), training on the executor's own per-tablet
counts, and what the trained network gains over sizing without it."""
from __future__ import annotations

import numpy as np
import pytest

import virginian_amd as va
from oracle import mlp as M
from oracle import oracle
from tests.helpers import assert_same
from virginian_amd import predict as PR
from virginian_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("loss", ["squared", "absolute", "quantile"])
def test_forward_and_gradient_match_oracle(gpu_db, loss):
    absolute = loss == "absolute"
    tau = 0.85 if loss == "quantile" else None
    for nin, nh, B, seed, skip in [(7, 5, 9, 1, 0.0), (PR.NFEATURES, 32, 300, 2, PR.SKIP), (40, 32, 1, 3, 0.0),
                                   (12, 32, 64, 4, 2.5)]:
        net = PR.MLP(nin, nh, seed=seed, skip=skip, loss=loss, tau=tau or 0.9)
        rng = np.random.default_rng(seed)
        # biases away from zero so that ReLUs are on and off
        p = net.params().astype(np.float64)
        p += rng.normal(0, 0.05, len(p))
        net.set_params(p.astype(np.float32))
        p = net.params().astype(np.float64)
        x = rng.normal(0, 1, (B, nin)).astype(np.float32)
        t = rng.uniform(0, 1, B).astype(np.float32)
        y = net.forward(x)
        y_ref, _ = M.forward(p, x.astype(np.float64), nin, nh, skip)
        assert np.allclose(y, y_ref, rtol=1e-4, atol=1e-6), (nin, np.abs(y - y_ref).max())
        # the host forward pass (same parameters, copied when set) agrees too
        yh = net.forward_host(x)
        assert np.allclose(yh, y_ref, rtol=1e-4, atol=1e-6), (nin, np.abs(yh - y_ref).max())
        assert np.allclose(yh, y, rtol=1e-5, atol=1e-6), (nin, np.abs(yh - y).max())
        lv, g = net.gradient(x, t)
        loss_ref, g_ref = M.gradient(p, x.astype(np.float64), t.astype(np.float64), nin, nh, skip, absolute,
                                     tau=tau)
        assert np.isclose(lv, loss_ref, rtol=1e-4)
        scale = np.abs(g_ref).max()
        assert np.allclose(g, g_ref, rtol=1e-4, atol=1e-5 * scale), (nin, np.abs(g - g_ref).max(), scale)
        net.close()


def test_training_fits_a_known_function(gpu_db):
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (2000, 6)).astype(np.float32)
    t = (1 / (1 + np.exp(-(2 * x[:, 0] - x[:, 1] * x[:, 2])))).astype(np.float32)
    net = PR.MLP(6, 32, seed=3, skip=0.0, loss="squared")
    before = M.loss(net.params().astype(np.float64), x, t, 6, 32)
    after = net.train(x, t, epochs=40, batch=128, lr=3e-3)
    assert after < before / 10 and after < 2e-3, (before, after)
    # the loss the GPU reports is the oracle's loss of the trained weights
    assert np.isclose(after, M.loss(net.params().astype(np.float64), x, t, 6, 32), rtol=1e-3)
    # the host copy of the parameters follows training
    assert np.allclose(net.forward_host(x), net.forward(x), rtol=1e-5, atol=1e-6)
    net.close()


@pytest.fixture
def learned(gpu_db):
    """A trained predictor over three tables and held-out programs."""
    rng = np.random.default_rng(17)
    sizes = [200_000, 700_000, 1_500_000]
    stats = {}
    for k, rows in enumerate(sizes):
        gpu_db.gen_table(f"t{k}", rows, 1 << 17, seed=40 + k)
        tid = gpu_db.table_id(f"t{k}")
        stats[tid] = PR.TableStats(gpu_db, tid)
    tids = list(stats)

    def progs(n):
        out = []
        for _ in range(n):
            sel = float(rng.uniform(0.0001, 0.8)) if rng.random() < 0.8 else None
            out.append(W.random_program(rng, max_ops=40, selectivity=sel, table=int(rng.choice(tids))))
        return out

    train, val, held = progs(400), progs(100), progs(150)
    x, t = PR.collect(gpu_db, train, stats)
    xv, tv = PR.collect(gpu_db, val, stats)
    net = PR.MLP(seed=7)
    fit = net.fit(x, t, xv, tv)
    print("fit", fit, flush=True)
    qnet = PR.MLP(seed=7, loss="quantile", tau=0.9)
    print("quantile fit", qnet.fit(x, t, xv, tv), flush=True)
    yield gpu_db, net, stats, train, held, (x, t), qnet
    net.close()
    qnet.close()


def test_predictor_gains_over_sizing_without_it(learned):
    """At the same overflow rate on held-out programs, buffers sized by
    the network reserve less than buffers sized from the mean selectivity
    of the training runs (the method without it), and far less than the
    table (what a single pass needs with no estimate at all)."""
    db, net, stats, train, held, (x, t), qnet = learned
    hx, ht = PR.collect(db, held, stats)
    mae_net = float(np.mean(np.abs(net.forward(hx) - ht)))
    mae_const = float(np.mean(np.abs(np.mean(t) - ht)))
    print(f"held-out MAE per tablet: network {mae_net:.4f}, mean selectivity {mae_const:.4f}")
    mae_est = float(np.mean(np.abs(hx[:, 3] - ht)))
    print(f"  (the independence estimate alone: {mae_est:.4f})")
    assert mae_net < 0.5 * mae_const
    # the network gains over the estimate it corrects, on programs it never saw
    assert mae_net < mae_est, (mae_net, mae_est)
    pred_net, pred_const, true_rows, rows = [], [], [], []
    for ops in held:
        st = stats[PR.table_of(ops)]
        f = PR.program_features(ops, st)
        pred_net.append(net.forward(f))
        pred_const.append(np.full(len(st.rows), float(np.mean(t))))
        rows.append(st.rows)
    k = 0
    for ops, r in zip(held, rows):
        true_rows.append(float((ht[k:k + len(r)] * r).sum()))
        k += len(r)
    m_net = PR.calibrate(pred_net, true_rows, rows, 0.05)
    m_const = PR.calibrate(pred_const, true_rows, rows, 0.05)
    res_net = sum(PR.capacity(p, r, m_net) for p, r in zip(pred_net, rows))
    res_const = sum(PR.capacity(p, r, m_const) for p, r in zip(pred_const, rows))
    res_all = sum(r.sum() for r in rows)
    print(f"reserved rows at <=5% overflow: network {res_net} (margin {m_net:.3f}), "
          f"mean selectivity {res_const} (margin {m_const:.3f}), whole tables {res_all:.0f}, "
          f"selected {sum(true_rows):.0f}")
    assert res_net < 0.7 * res_const


def test_quantile_sizing_reserves_less_than_the_estimate(learned):
    """What the network gains over the method it corrects: trained for the
    0.9-quantile of each tablet's selectivity, it sizes buffers by how
    uncertain each program is, so at the same overflow rate on held-out
    programs (margins calibrated for each method alike) it reserves fewer
    rows than the independence estimate of the histograms."""
    db, net, stats, train, held, (x, t), qnet = learned
    hx, ht = PR.collect(db, held, stats)
    pq, pe, true_rows, rows, k = [], [], [], [], 0
    yq = qnet.forward(hx)
    for ops in held:
        st = stats[PR.table_of(ops)]
        n = len(st.rows)
        pq.append(yq[k:k + n])
        pe.append(hx[k:k + n, 3])
        true_rows.append(float((ht[k:k + n] * st.rows).sum()))
        rows.append(st.rows)
        k += n
    for target in (0.02, 0.05):
        m_q = PR.calibrate(pq, true_rows, rows, target)
        m_e = PR.calibrate(pe, true_rows, rows, target)
        res_q = sum(PR.capacity(p, r, m_q) for p, r in zip(pq, rows))
        res_e = sum(PR.capacity(p, r, m_e) for p, r in zip(pe, rows))
        over_q = sum(PR.capacity(p, r, m_q) < n for p, n, r in zip(pq, true_rows, rows))
        over_e = sum(PR.capacity(p, r, m_e) < n for p, n, r in zip(pe, true_rows, rows))
        print(f"<= {target:.0%} overflow: quantile network {res_q} rows (margin {m_q:.3f}, {over_q} over), "
              f"estimate {res_e} rows (margin {m_e:.3f}, {over_e} over), selected {sum(true_rows):.0f}")
        assert over_q <= target * len(held) and over_e <= target * len(held)
        assert res_q < res_e, (target, res_q, res_e)


def test_single_pass_with_predicted_sizes(learned):
    """Single pass into network-sized buffers; a query that did not fit is
    rerun in two passes; every result equals the oracle's."""
    db, net, stats, train, held, _, _ = learned
    qs, want = [], []
    for ops in held[:40]:
        st = stats[PR.table_of(ops)]
        cap = PR.capacity(net.forward(PR.program_features(ops, st)), st.rows, 0.02)
        qs.append(db.prepare(ops, capacity=cap, flags=va.VG_QUERY_LIGHT))
        want.append(oracle.select(db, ops)[0])
    b = va.PersistentBatch(qs, tiles_per_item=4, single=1)
    try:
        rows, _, flags = b.results()
    finally:
        b.close()
    print(f"{sum(1 for f in flags if f)} of {len(qs)} did not fit", flush=True)
    for i, (q, w, f) in enumerate(zip(qs, want, flags)):
        n = len(next(iter(w.values())))
        assert rows[i] == n
        if f:
            q2 = db.prepare(held[i], capacity=n, flags=va.VG_QUERY_LIGHT)
            q2.run()
            assert_same(q2.fetch(), w, f"query {i} rerun")
            q2.close()
        else:
            assert_same(q.fetch(), w, f"query {i} single pass")
    for q in qs:
        q.close()
