"""FP64 restatement of the control-variate MLP (test infrastructure only).

Network: z = 2 (x - lo) / (hi - lo) - 1 per coordinate; h1 = tanh(W1 z + b1),
h2 = tanh(W2 h1 + b2), h3 = tanh(W3 h2 + b3), o = W4 h3 + b4 with o = (u, gx, gy):
2 -> 64 -> 64 -> 64 -> 3, 8707 parameters, packed as
[W1 (64x2), b1, W2 (64x64), b2, W3 (64x64), b3, W4 (3x64), b4], W row-major [out][in].

Loss over a batch of n samples with targets u_k:
    L = (1/n) sum_k (o0_k - u_k)^2 + lam ((o1_k - gx_k)^2 + (o2_k - gy_k)^2)
where (gx_k, gy_k) = d o0 / d(x, y) at x_k computed by forward-mode through the
network and held fixed (no gradient flows through them): the gradient heads learn
the derivative of the u head, so one forward pass gives the control-variate field.

Only tests/ import this module, as the checker of the HIP training kernels.
Pinned by tests/test_mlp.py: the analytic backward equals central finite
differences of the loss; adam_run's first step moves every parameter with a
nonzero gradient by lr0 against the gradient's sign (the bias-corrected first
step), and a constant gradient moves it by lr at every step.
"""
import numpy as np

H = 64
SIZES = [(H, 2), (H,), (H, H), (H,), (H, H), (H,), (3, H), (3,)]
N_PARAMS = sum(int(np.prod(s)) for s in SIZES)


def unpack(theta):
    out, o = [], 0
    for s in SIZES:
        n = int(np.prod(s))
        out.append(np.asarray(theta[o:o + n], dtype=np.float64).reshape(s))
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(p, dtype=np.float64).ravel() for p in parts])


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (the host initialisation of the library)."""
    rng = np.random.default_rng(seed)
    parts = []
    for s in SIZES:
        if len(s) == 2:
            lim = np.sqrt(6.0 / (s[0] + s[1]))
            parts.append(rng.uniform(-lim, lim, size=s))
        else:
            parts.append(np.zeros(s))
    return pack(parts)


def forward(theta, x, y, domain):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(theta)
    lo = np.array(domain[:2])
    span = np.array(domain[2:]) - lo
    z = np.stack([2 * (x - lo[0]) / span[0] - 1, 2 * (y - lo[1]) / span[1] - 1])  # (2, n)
    a1 = W1 @ z + b1[:, None]
    h1 = np.tanh(a1)
    h2 = np.tanh(W2 @ h1 + b2[:, None])
    h3 = np.tanh(W3 @ h2 + b3[:, None])
    o = W4 @ h3 + b4[:, None]
    # forward-mode d o0 / d x, d o0 / d y
    g = []
    for d in range(2):
        t1 = (1 - h1 ** 2) * (W1[:, d:d + 1] * (2.0 / span[d]))
        t2 = (1 - h2 ** 2) * (W2 @ t1)
        t3 = (1 - h3 ** 2) * (W3 @ t2)
        g.append((W4[0] @ t3))
    return dict(z=z, h1=h1, h2=h2, h3=h3, o=o, g=np.stack(g))


def loss_grad(theta, x, y, u, lam, domain):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(theta)
    f = forward(theta, x, y, domain)
    n = x.size
    o, g = f["o"], f["g"]
    r0 = o[0] - u
    r1 = o[1] - g[0]
    r2 = o[2] - g[1]
    loss = np.mean(r0 ** 2 + lam * (r1 ** 2 + r2 ** 2))
    dO = np.stack([2 * r0, 2 * lam * r1, 2 * lam * r2]) / n  # (3, n)
    dW4 = dO @ f["h3"].T
    db4 = dO.sum(1)
    d3 = (W4.T @ dO) * (1 - f["h3"] ** 2)
    dW3 = d3 @ f["h2"].T
    db3 = d3.sum(1)
    d2 = (W3.T @ d3) * (1 - f["h2"] ** 2)
    dW2 = d2 @ f["h1"].T
    db2 = d2.sum(1)
    d1 = (W2.T @ d2) * (1 - f["h1"] ** 2)
    dW1 = d1 @ f["z"].T
    db1 = d1.sum(1)
    return loss, pack([dW1, db1, dW2, db2, dW3, db3, dW4, db4])


def loss_only(theta, x, y, u, lam, domain, g_fixed):
    """Loss with the gradient targets held at g_fixed (for finite differences)."""
    f = forward(theta, x, y, domain)
    o = f["o"]
    return np.mean((o[0] - u) ** 2 + lam * ((o[1] - g_fixed[0]) ** 2 + (o[2] - g_fixed[1]) ** 2))


def adam(theta, grad, steps, lr0, lr1, beta2, beta1=0.9, eps=1e-8):
    """`steps` Adam steps in FP64 on the gradient function grad(theta) (Kingma & Ba, "Adam: a
    method for stochastic optimization", ICLR 2015, Algorithm 1: bias-corrected first and second
    moments from zero), with the learning rate decayed geometrically from lr0 at the first step
    to lr1 at the last (lr1 <= 0: constant lr0).  Returns the parameters after the last step."""
    th = np.array(theta, dtype=np.float64)
    m = np.zeros_like(th)
    v = np.zeros_like(th)
    for k in range(steps):
        g = grad(th)
        t = k + 1
        lr = lr0 * (lr1 / lr0) ** (k / (steps - 1)) if (lr1 > 0 and steps > 1) else lr0
        m = beta1 * m + (1 - beta1) * g
        v = beta2 * v + (1 - beta2) * g * g
        th = th - lr * (m / (1 - beta1 ** t)) / (np.sqrt(v / (1 - beta2 ** t)) + eps)
    return th


def adam_run(theta, x, y, u, lam, domain, steps, lr0, lr1, beta2):
    """adam() on the full-batch training loss of the network."""
    return adam(theta, lambda th: loss_grad(th, x, y, u, lam, domain)[1], steps, lr0, lr1, beta2)
