"""Pins the CPU oracle of the learned warm start (oracle/mlp_oracle.py)."""
import numpy as np

from oracle import mlp_oracle as mo


def _params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    return (scale * rng.standard_normal(mo.layout()[3])).astype(np.float32)


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    p = _params()
    feat = rng.standard_normal((7, 20))
    q = rng.standard_normal((7, 6))
    t = rng.standard_normal(7)
    loss, g = mo.loss_grad(p, feat, q, t, 0.37, round_weights=False, bf16_ops=False)
    w_off, b_off, _, n_all = mo.layout()
    idx = [w_off[0] + 5, w_off[1] + 1000, w_off[2] + 77, w_off[3] + 100, b_off[0] + 3, b_off[3] + 2,
           b_off[2] + 50, w_off[3] + 767]
    p64 = p.astype(np.float64)
    for i in idx:
        h = 1e-6
        pp, pm = p64.copy(), p64.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = mo.loss_grad(pp, feat, q, t, 0.37, round_weights=False, bf16_ops=False)
        lm, _ = mo.loss_grad(pm, feat, q, t, 0.37, round_weights=False, bf16_ops=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38], np.float32)
    r = mo.bf16_round(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                 # halfway, 1.0 has even mantissa
    assert r[2] == 1.0 + 4 * 2 ** -8   # halfway, rounds up to even
    assert r[3] == 1.0
    assert r[4] == -2.5
    assert np.isfinite(r[5])


def test_features_on_a_hand_built_body():
    # one body rotated 90 degrees about z: body-frame values undo the rotation
    th = np.pi / 2
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    rng = np.random.default_rng(3)
    local = rng.standard_normal((20, 3))
    nloc = local / np.linalg.norm(local, axis=1, keepdims=True)
    c = np.array([1.0, 2.0, 3.0])
    y = local @ R.T + c
    nrm = nloc @ R.T
    vw = np.array([[1.0, 0.0, 0.0, 0.0, 0.0, 2.0]])
    f, q = mo.features(y, nrm, np.arange(20.0), np.zeros(20, np.int64), np.array([0, 20]), R.reshape(1, 9),
                       c[None], vw)
    assert np.allclose(f[:, 0:3], local)
    assert np.allclose(f[:, 3:6], nloc)
    assert np.allclose(q[0], [0.0, -1.0, 0.0, 0.0, 0.0, 2.0])   # R^T v, R^T w
    # neighbours: max over the 16 nearest of (r_k - r_j)
    j = 4
    d = np.sum((local - local[j]) ** 2, axis=1)
    d[j] = np.inf
    nb = np.argsort(d)[:16]
    assert np.allclose(f[j, 13:16], (local[nb] - local[j]).max(axis=0))
    assert f[j, 19] == max(nb)


def test_adam_first_step_moves_by_lr_times_sign():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 0.0])
    p2, m, v = mo.adam(p, g, np.zeros(3), np.zeros(3), 1, 1e-3)
    assert np.allclose(p2, p - 1e-3 * np.sign(g), atol=1e-9)


def test_bf16_operands_leave_bf16_exact_values_alone():
    """With sparse weights of +-0.5, features and velocities in {-1, 0, 1}
    and targets equal to the outputs except at one point (a residual of 1
    there), every product operand -- the activations forward, the one-point
    back-propagated gradients backward -- is a short binary fraction, hence a
    BF16 value, so rounding the operands (bf16_ops) must change nothing:
    rounding happens at the operands of the products only, never to their
    results (the weight gradients themselves need not be BF16 values)."""
    rng = np.random.default_rng(5)
    n_all = mo.layout()[3]
    p = (rng.integers(-1, 2, n_all) * 0.5).astype(np.float32)
    p[rng.random(n_all) < 0.95] = 0.0         # sparse: sums stay short binary fractions
    feat = rng.integers(-1, 2, (9, 20)).astype(np.float64)
    q = rng.integers(-1, 2, (9, 6)).astype(np.float64)
    o1, A1 = mo.forward(p, feat, bf16_ops=True)
    o0, A0 = mo.forward(p, feat, bf16_ops=False)
    exact = lambda a: np.array_equal(mo.bf16_round(np.asarray(a, np.float32)), a)
    assert all(exact(a) for a in A0)          # the premise: every operand is a BF16 value
    assert np.array_equal(o1, o0)
    t = np.sum(o0 * q, axis=1)
    t[3] -= 1.0
    l1, g1 = mo.loss_grad(p, feat, q, t, 0.25, bf16_ops=True)
    l0, g0 = mo.loss_grad(p, feat, q, t, 0.25, bf16_ops=False)
    assert l1 == l0 == 0.25 and np.array_equal(g1, g0)
    assert np.count_nonzero(g0) > 100


def test_bf16_operands_move_results_by_about_the_bf16_unit():
    """On generic data the rounded-operand network differs from the exact
    one by about BF16's unit roundoff (2^-9 relative per operand), far more
    than FP32 rounding and far less than O(1)."""
    rng = np.random.default_rng(6)
    p = _params(seed=2, scale=0.15)
    feat = rng.standard_normal((50, 20))
    o1, _ = mo.forward(p, feat, bf16_ops=True)
    o0, _ = mo.forward(p, feat, bf16_ops=False)
    rel = np.linalg.norm(o1 - o0) / np.linalg.norm(o0)
    assert 1e-4 < rel < 2e-2, rel
