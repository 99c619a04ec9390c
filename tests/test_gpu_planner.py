"""The planner's estimator kernels against the CPU oracle (oracle/mlp_ref.py), and the planned
batched path against the oracle's query results: a plan may only reorder and skip, never change a
result."""
import numpy as np
import pytest

from helpers import ExprGen, small_spec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(engine):
    from oracle.oracle import OracleTable

    ds, ts = small_spec(seed=31, n_rows=6000, L=500, chunk=4096)
    ot = OracleTable(ts)
    et = engine.table(ts)
    yield ds, ts, ot, et
    et.close()
    ot.close()


def test_forward_and_gradients_match_oracle(engine):
    """GPU forward and backward (BF16 parameters, FP32 accumulation) against the float64 oracle on
    the same BF16 parameter values.  Tolerance: the kernels accumulate in FP32, so the difference
    is FP32 rounding of sums of at most 300 products, relative 1e-4 of the largest magnitude."""
    from oracle import mlp_ref

    m = engine.mlp(n_in=32, n_hidden=128, n_hidden_layers=3, seed=5)
    try:
        dims = [32, 128, 128, 128, 1]
        p = m.parameters()
        assert len(p) == m.n_params == sum(o * i + o for o, i in mlp_ref.layer_shapes(dims))
        # the parameters the kernels use are BF16 values
        assert np.array_equal(p, mlp_ref.bf16_round(p))
        rng = np.random.default_rng(1)
        x = rng.uniform(0, 1, (300, 32)).astype(np.float32)
        y = rng.uniform(0, 1, 300).astype(np.float32)
        out_ref, _ = mlp_ref.forward(p, dims, x)
        out = m.forward(x)
        assert np.allclose(out, out_ref, rtol=1e-4, atol=1e-4 * np.abs(out_ref).max())
        loss, g = m.gradients(x, y)
        loss_ref, g_ref = mlp_ref.loss_and_grads(p, dims, x, y)
        assert abs(loss - loss_ref) <= 1e-4 * loss_ref
        assert np.allclose(g, g_ref, rtol=1e-3, atol=1e-4 * np.abs(g_ref).max())
    finally:
        m.free()


def test_set_parameters_round_trip_and_training(engine):
    from oracle import mlp_ref

    m = engine.mlp(n_in=32, n_hidden=64, n_hidden_layers=2, seed=2)
    try:
        rng = np.random.default_rng(3)
        p = rng.normal(0, 0.1, m.n_params).astype(np.float32)
        m.set_parameters(p)
        assert np.array_equal(m.parameters(), mlp_ref.bf16_round(p))
        # a learnable target: y = mean of the first four features
        x = rng.uniform(0, 1, (8192, 32)).astype(np.float32)
        y = x[:, :4].mean(axis=1).astype(np.float32)
        before = float(np.mean((m.forward(x) - y) ** 2))
        last = m.train(x, y, epochs=20, batch=512, lr=3e-3, seed=4)
        after = float(np.mean((m.forward(x) - y) ** 2))
        assert after < before / 20 and after < 2e-3, (before, after, last)
    finally:
        m.free()


def test_plan_samples_targets_are_exact(small):
    """Every node's target is the normalised exact cardinality of its own sub-filter."""
    from silo_amd.engine import plan_estimates, plan_samples

    ds, ts, ot, et = small
    gen = ExprGen(ds, seed=12)
    filters = [gen.expr(3) for _ in range(20)]
    f, y, tree = plan_samples(et, filters)
    assert f.shape[1] == 32 and len(y) == len(tree) > len(filters)
    est, ind, ex, tr = plan_estimates(et, filters)
    assert np.array_equal(tr, tree)
    assert np.allclose(y, np.log1p(ex.astype(np.float64)) / np.log1p(ts.n_rows), atol=1e-6)
    # the last node of each tree is its root: the filter's own cardinality
    for q, e in enumerate(filters):
        root = np.flatnonzero(tree == q)[-1]
        assert int(ex[root]) == ot.count(e), e


@pytest.mark.parametrize("mode", ["independence", "learned"])
def test_planned_batches_are_unchanged(small, engine, mode):
    ds, ts, ot, et = small
    gen = ExprGen(ds, seed=40)
    filters = [gen.expr(4) for _ in range(48)]
    queries = []
    for i, e in enumerate(filters):
        if i % 3 == 0:
            queries.append({"filter": e, "action": "count"})
        elif i % 3 == 1:
            queries.append({"filter": e, "action": "mutations", "column": "main", "min_proportion": 0.05})
        else:
            queries.append({"filter": e, "action": "aggregate", "dims": ["country"]})
    want = []
    for q in queries:
        if q["action"] == "count":
            want.append(ot.count(q["filter"]))
        elif q["action"] == "mutations":
            want.append(ot.mutations(q["filter"], "main", 0.05))
        else:
            want.append(ot.aggregate(q["filter"], ["country"]))
    mlp = engine.mlp(seed=9) if mode == "learned" else None
    try:
        engine.set_planner(mode, mlp)
        assert et.batch(queries) == want
        engine.set_planner("reference")
        assert et.batch(queries) == want
    finally:
        engine.set_planner("reference")
        if mlp is not None:
            mlp.free()


def test_trained_estimator_beats_independence(engine):
    """What the trained network gains over the method without it: trained on the engine's exact node
    cardinalities of seeded trees, it estimates held-out trees' node cardinalities with a lower p95
    q-error than the independence estimate (product / complement / Poisson-binomial rules) that the
    planner uses without it."""
    from silo_amd import datagen, planner_train, table

    # the config-1 generator (lineage tree over the full genome) at 200,000 rows: leaves are
    # correlated through the lineages, which the independence rules ignore
    ds = datagen.generate_config(1, n_rows=200000)
    et = engine.table(table.spec_from_dataset(ds))
    try:
        x, y = planner_train.samples(et, planner_train.trees(ds, 3000, seed=3))
        mlp = engine.mlp(n_in=32, n_hidden=128, n_hidden_layers=3, seed=7)
        mlp.train(x, y, epochs=20, batch=4096, lr=1e-3, seed=8)
        ev = planner_train.evaluate(et, mlp, planner_train.trees(ds, 500, seed=4))
        learned, indep = ev["learned"], ev["independence"]
        assert learned["nodes_p95"] < indep["nodes_p95"], ev
        assert learned["roots_p95"] < indep["roots_p95"], ev
    finally:
        et.close()
