"""Pins the CPU fp64 statement of the scheduler network (CPU only): its
hand-derived gradient against central finite differences, and its bf16
rounding against exhaustive float32 cases."""
import numpy as np

from oracle import mlp


def _random_problem(seed, n=40):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.4, mlp.PARAMS)
    x = rng.random((n, mlp.F))
    rule = rng.integers(0, 70, n)  # ids past 63 wrap onto the table
    y = (rng.random(n) < 0.4).astype(np.float64)
    return p, x, rule, y


def test_gradient_matches_finite_differences():
    p, x, rule, y = _random_problem(1)
    loss, g, _ = mlp.loss_and_grad(p, x, rule, y)
    rng = np.random.default_rng(2)
    # every block of the layout, plus random parameters
    idx = [mlp.OFF_B3, mlp.OFF_W3 + 5, mlp.OFF_B2 + 7, mlp.OFF_W2 + 123, mlp.OFF_B1 + 3, mlp.OFF_W1 + 70,
           mlp.OFF_W1 + 15 * 64 + 2, mlp.OFF_E + (rule[0] % 64) * 8 + 4]
    idx += list(rng.integers(0, mlp.PARAMS, 40))
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss_and_grad(pp, x, rule, y)[0] - mlp.loss_and_grad(pm, x, rule, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_embedding_gradient_only_on_used_rows():
    p, x, rule, y = _random_problem(3)
    rule = np.full_like(rule, 5)
    _, g, _ = mlp.loss_and_grad(p, x, rule, y)
    gE = g[mlp.OFF_E:mlp.OFF_W1].reshape(64, 8)
    assert np.count_nonzero(np.abs(gE).sum(1)) <= 1 and np.abs(gE[5]).sum() > 0


def test_loss_closed_forms():
    # all-zero network: logit 0 for every sample, loss = log 2, dL/db3 = mean(0.5 - y)
    p = np.zeros(mlp.PARAMS)
    x = np.random.default_rng(0).random((10, 8))
    y = np.array([1, 0, 1, 1, 0, 0, 0, 1, 1, 1], float)
    loss, g, z = mlp.loss_and_grad(p, x, np.zeros(10, int), y)
    assert np.allclose(z, 0) and abs(loss - np.log(2)) < 1e-12
    assert abs(g[mlp.OFF_B3] - np.mean(0.5 - y)) < 1e-12


def test_bf16_rounding():
    vals = np.array([1.0, -2.5, 3.1415926, 1e-3, 65504.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 0.0], np.float32)
    b = mlp.f32_to_bf16(vals)
    back = mlp.bf16_to_f64(b)
    # exact for representable values, nearest-even otherwise
    assert back[0] == 1.0 and back[1] == -2.5 and back[7] == 0.0
    assert back[5] == 1.0  # tie 1 + 2^-8 rounds to even (1.0)
    assert back[6] == 1.0 + 2 ** -7  # 1 + 3*2^-9 is nearer to 1 + 2^-7 than to 1
    rel = np.abs(back[vals != 0] - vals[vals != 0]) / np.abs(vals[vals != 0])
    assert (rel <= 2 ** -8).all()
