"""CPU fp64 statement of the learned tile scheduler's network.

TEST INFRASTRUCTURE ONLY: the tests use it as the checker of the engine's HIP
forward / backward kernels.  Nothing in the engine calls it.

The network has no counterpart in the reference: it is this engine's learned
scheduler.  Its layout follows the C ABI, CFPQ_SCHED_PARAMS = 5825 parameters:

  E[64][8]   rule embedding (row = rule id mod 64)
  W1[16][64], b1[64]   input = 8 tile features ++ E[rule]
  W2[64][64], b2[64]   ReLU after each hidden layer
  W3[64], b3           one logit

The loss is the mean binary cross-entropy with logits.  Gradients are derived
by hand here and checked against central finite differences in the tests.
"""
import numpy as np

F, E_DIM, IN, H, RULES = 8, 8, 16, 64, 64
OFF_E = 0
OFF_W1 = OFF_E + RULES * E_DIM
OFF_B1 = OFF_W1 + IN * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
PARAMS = OFF_B3 + 1


def bf16_to_f64(w: np.ndarray) -> np.ndarray:
    return (np.asarray(w, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def f32_to_bf16(x) -> np.ndarray:
    """Round to nearest even (finite inputs)."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = u + 0x7FFF + ((u >> 16) & 1)
    return (u >> 16).astype(np.uint16)


def unpack(p: np.ndarray):
    return (p[OFF_E:OFF_W1].reshape(RULES, E_DIM), p[OFF_W1:OFF_B1].reshape(IN, H), p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p: np.ndarray, x: np.ndarray, rule: np.ndarray):
    E, W1, b1, W2, b2, W3, b3 = unpack(p)
    inp = np.concatenate([np.asarray(x, np.float64), E[np.asarray(rule) % RULES]], axis=1)
    a1 = inp @ W1 + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ W3 + b3
    return z, (inp, a1, h1, a2, h2)


def loss_and_grad(p: np.ndarray, x, rule, y):
    """Mean BCE-with-logits and its gradient (fp64) at parameters p (fp64)."""
    y = np.asarray(y, np.float64)
    n = y.shape[0]
    z, (inp, a1, h1, a2, h2) = forward(p, x, rule)
    loss = np.mean(np.maximum(z, 0) - y * z + np.log1p(np.exp(-np.abs(z))))
    E, W1, b1, W2, b2, W3, b3 = unpack(p)
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / n
    g = np.zeros(PARAMS)
    g[OFF_B3] = dz.sum()
    g[OFF_W3:OFF_B3] = h2.T @ dz
    dh2 = np.outer(dz, W3) * (a2 > 0)
    g[OFF_B2:OFF_W3] = dh2.sum(0)
    g[OFF_W2:OFF_B2] = (h1.T @ dh2).ravel()
    dh1 = (dh2 @ W2.T) * (a1 > 0)
    g[OFF_B1:OFF_W2] = dh1.sum(0)
    g[OFF_W1:OFF_B1] = (inp.T @ dh1).ravel()
    din = dh1 @ W1.T
    gE = np.zeros((RULES, E_DIM))
    np.add.at(gE, np.asarray(rule) % RULES, din[:, F:])
    g[OFF_E:OFF_W1] = gE.ravel()
    return loss, g, z
