"""CPU pins of the oracle's optimizer restatement (Adam with all tcnn keys,
ExponentialDecay schedule, EMA) against closed forms computed independently
in float64 here, and of the training-loop restatement against
torch.optim.Adam on a plain network."""
import numpy as np
import pytest

from oracle import model as M


def _adam_closed_form(w0, grads, n_matrix, lr, b1, b2, eps, l2=0.0, nm_factor=1.0, rel=0.0, ab=0.0, clip=0.0,
                      adabound=False):
    """Per-parameter Adam recursion in float64 (textbook Adam with tcnn's rules:
    zero gradients skip non-matrix parameters, L2 only on matrix parameters)."""
    w = np.array(w0, np.float64)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    t = np.zeros(w.size, np.int64)
    for step, g_all in enumerate(grads, start=1):
        for i in range(w.size):
            g = float(g_all[i])
            matrix = i < n_matrix
            if not matrix and g == 0.0:
                continue
            if matrix:
                g += l2 * w[i]
            m[i] = b1 * m[i] + (1 - b1) * g
            v[i] = b2 * v[i] + (1 - b2) * g * g
            t[i] += 1
            a = lr * (1.0 if matrix else nm_factor)
            a *= np.sqrt(1 - b2 ** t[i]) / (1 - b1 ** t[i])
            eff = a / (np.sqrt(v[i]) + eps)
            if adabound:
                lo = 0.1 - 0.1 / ((1 - b2) * step + 1)
                hi = 0.1 + 0.1 / ((1 - b2) * step)
                eff = min(max(eff, lo), hi)
            nw = (1 - rel * a) * w[i] - np.copysign(ab * a, w[i]) - eff * m[i]
            if clip:
                nw = min(max(nw, -clip), clip)
            w[i] = nw
    return w


@pytest.mark.parametrize("extra", [
    {},
    {"l2_reg": 1e-2},
    {"non_matrix_learning_rate_factor": 0.25},
    {"relative_decay": 0.1, "absolute_decay": 1e-3},
    {"clipping_magnitude": 0.3},
    {"adabound": True},
])
def test_adam_three_steps_closed_form(extra):
    rng = np.random.default_rng(0)
    n, n_matrix = 40, 24
    w0 = rng.uniform(-0.5, 0.5, n).astype(np.float32)
    grads = [rng.normal(size=n).astype(np.float32) for _ in range(3)]
    grads[1][30] = 0.0  # a grid entry untouched in step 2: skipped, its step count lags
    grads[2][5] = 0.0   # a matrix entry with zero gradient: still updated (moments decay)
    lr, b1, b2, eps = 1e-2, 0.9, 0.99, 1e-8
    w = w0.copy()
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    steps = np.zeros(n, np.uint32)
    l2 = extra.get("l2_reg", 0.0)
    oracle_extra = {k: val for k, val in extra.items() if k != "l2_reg"}
    for s, g in enumerate(grads, start=1):
        M.adam_step(w, (g * 128.0).astype(np.float32), m, v, steps, n_matrix, 128.0, lr, b1, b2, eps, l2,
                    opt_step=s, **oracle_extra)
    ref = _adam_closed_form(w0, grads, n_matrix, lr, b1, b2, eps, l2=l2,
                            nm_factor=extra.get("non_matrix_learning_rate_factor", 1.0),
                            rel=extra.get("relative_decay", 0.0), ab=extra.get("absolute_decay", 0.0),
                            clip=extra.get("clipping_magnitude", 0.0), adabound=extra.get("adabound", False))
    assert steps[30] == 2 and steps[5] == 3
    np.testing.assert_allclose(w, ref, rtol=2e-5, atol=2e-6)


def test_adam_flags_freeze_parameter_groups():
    rng = np.random.default_rng(1)
    n, n_matrix = 20, 8
    w0 = rng.uniform(-1, 1, n).astype(np.float32)
    g = rng.normal(size=n).astype(np.float32)
    for flag, frozen in (("optimize_matrix_params", slice(0, n_matrix)), ("optimize_non_matrix_params",
                                                                          slice(n_matrix, n))):
        w = w0.copy()
        M.adam_step(w, g, np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.uint32), n_matrix, 1.0,
                    1e-2, 0.9, 0.99, 1e-8, 0.0, **{flag: False})
        assert np.array_equal(w[frozen], w0[frozen])
        other = np.ones(n, bool)
        other[frozen] = False
        assert (w[other] != w0[other]).all()


def test_exponential_decay_schedule_closed_form():
    base, start, interval, end = 0.33, 8, 4, 30
    f = M.exp_decay_factors(50, base, start, interval, end)
    for s in range(50):
        k = 0 if s < start else (min(s, end) - start) // interval + 1
        assert abs(f[s] - base ** k) <= 1e-6 * base ** k, s


def test_ema_closed_form():
    rng = np.random.default_rng(2)
    ws = [rng.normal(size=16).astype(np.float32) for _ in range(6)]
    d = 0.95
    ema = np.zeros(16, np.float32)
    for t, w in enumerate(ws, start=1):
        M.ema_step(w, ema, d, t)
        # debiased exponential average of the weights seen so far
        num = sum((1 - d) * d ** (t - k) * ws[k - 1].astype(np.float64) for k in range(1, t + 1))
        np.testing.assert_allclose(ema, num / (1 - d ** t), rtol=1e-5, atol=1e-6)
    const = np.full(16, 0.7, np.float32)
    e = np.zeros(16, np.float32)
    for t in range(1, 10):
        M.ema_step(const, e, d, t)
    np.testing.assert_allclose(e, const, rtol=1e-6)


def test_train_loop_matches_torch_adam():
    """The oracle's training loop on a network without grid parameters equals
    torch.optim.Adam driving the FP64 oracle gradients (the tcnn Adam without
    L2 / decay is textbook Adam with eps outside the sqrt)."""
    torch = pytest.importorskip("torch")
    enc = {"otype": "Identity"}
    net = {"otype": "FullyFusedMLP", "n_neurons": 16, "n_hidden_layers": 2, "activation": "ReLU",
           "output_activation": "None"}
    L = M.Layout(3, 1, enc, net)
    p0 = M.xavier_uniform(L, 3)
    rng = np.random.default_rng(4)
    batches = []
    for _ in range(5):
        x = rng.uniform(0, 1, (64, 3)).astype(np.float32)
        batches.append((x, (x.sum(axis=1, keepdims=True) * 0.5).astype(np.float32)))
    opt = {"otype": "Adam", "learning_rate": 1e-2, "beta1": 0.9, "beta2": 0.99, "epsilon": 1e-8, "l2_reg": 0.0}
    losses, w, _ = M.train(L, p0, batches, opt, "L2", loss_scale=1.0)
    wt = torch.tensor(p0.astype(np.float64), requires_grad=True)
    topt = torch.optim.Adam([wt], lr=1e-2, betas=(0.9, 0.99), eps=1e-8)
    for s, (x, t) in enumerate(batches):
        cur = wt.detach().numpy().astype(np.float32)
        out = M.forward(L, cur, x)
        _, g_out = M.loss("L2", out[:, :1], t, 1, 1.0)
        d = np.zeros((64, L.net.n_output_padded))
        d[:, :1] = g_out
        g, _ = M.backward(L, cur, x, d)
        wt.grad = torch.from_numpy(g.astype(np.float32).astype(np.float64))
        topt.step()
    np.testing.assert_allclose(w, wt.detach().numpy(), rtol=1e-4, atol=1e-6)
    assert losses[-1] < losses[0]
