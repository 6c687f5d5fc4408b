"""The control-variate MLP: oracle pins (CPU) and the HIP kernels against the
oracle (GPU).

Pins of oracle/mlp_oracle.py, independent of its backward code:
- the analytic parameter gradient equals central finite differences of the
  loss (gradient targets held fixed, as the loss defines them);
- the forward-mode d o0 / d(x, y) equals central finite differences of o0.
GPU checks, all through the C ABI:
- FP32 forward and loss gradient against the FP64 oracle at the same
  parameters (relative 1e-4: FP32 arithmetic, sequential sums);
- BF16-MFMA inference against a numpy model of the same arithmetic (weights
  and hidden activations rounded to bf16, FP32 accumulation): 3e-3 absolute,
  far below what a wrong lane layout or operand order would give;
- the table equals inference at the grid points;
- training reduces the loss on a smooth target by orders of magnitude.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo
from spac_amd.problems import franke_u

DOM = (0.0, 0.0, 1.0, 1.0)


def _samples(n, seed=0):
    rng = np.random.default_rng(seed)
    x, y = rng.random(n), rng.random(n)
    return x, y, franke_u(x, y) + 0.01 * rng.standard_normal(n)


def test_oracle_parameter_count():
    assert mo.N_PARAMS == 8707
    th = mo.init_params(1)
    parts = mo.unpack(th)
    assert [p.shape for p in parts] == [tuple(s) for s in mo.SIZES]
    assert np.array_equal(mo.pack(parts), th)


def test_oracle_gradient_matches_finite_differences():
    x, y, u = _samples(40, 1)
    th = mo.init_params(2)
    th += 0.05 * np.random.default_rng(3).standard_normal(th.size)  # nonzero biases too
    lam = 0.3
    loss, g = mo.loss_grad(th, x, y, u, lam, DOM)
    gfix = mo.forward(th, x, y, DOM)["g"]
    rng = np.random.default_rng(4)
    idx = np.concatenate([rng.choice(th.size, 60, replace=False), [0, 128, 192, 4288, 8514, 8706]])
    eps = 1e-6
    for k in idx:
        tp, tm = th.copy(), th.copy()
        tp[k] += eps
        tm[k] -= eps
        fd = (mo.loss_only(tp, x, y, u, lam, DOM, gfix) - mo.loss_only(tm, x, y, u, lam, DOM, gfix)) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-9, (k, fd, g[k])
    assert abs(loss - mo.loss_only(th, x, y, u, lam, DOM, gfix)) < 1e-15


def test_oracle_forward_mode_gradient():
    th = mo.init_params(5)
    x, y, _ = _samples(25, 6)
    f = mo.forward(th, x, y, DOM)
    eps = 1e-6
    fx = (mo.forward(th, x + eps, y, DOM)["o"][0] - mo.forward(th, x - eps, y, DOM)["o"][0]) / (2 * eps)
    fy = (mo.forward(th, x, y + eps, DOM)["o"][0] - mo.forward(th, x, y - eps, DOM)["o"][0]) / (2 * eps)
    assert np.abs(f["g"][0] - fx).max() < 1e-8
    assert np.abs(f["g"][1] - fy).max() < 1e-8


# ------------------------------------------------------------------ GPU --
def _bf16(a):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16 << 16
    return b.astype(np.uint32).view(np.float32)


def _mfma_model(th, x, y):
    """The inference arithmetic: layer 1 and the heads in FP32, hidden layers with
    bf16 weights and bf16 inputs, FP32 accumulation."""
    W1, b1, W2, b2, W3, b3, W4, b4 = [p.astype(np.float32) for p in mo.unpack(th.astype(np.float64))]
    z = np.stack([2 * (x - DOM[0]) / (DOM[2] - DOM[0]) - 1, 2 * (y - DOM[1]) / (DOM[3] - DOM[1]) - 1]).astype(np.float32)
    h1 = np.tanh(W1 @ z + b1[:, None])
    h2 = np.tanh(_bf16(W2).astype(np.float64) @ _bf16(h1).astype(np.float64) + b2[:, None])
    h3 = np.tanh(_bf16(W3).astype(np.float64) @ _bf16(h2).astype(np.float64) + b3[:, None])
    return (W4.astype(np.float64) @ h3 + b4[:, None]).T


@pytest.fixture(scope="module")
def mlp(gpu):
    from spac_amd.mlp import MLP

    with MLP(DOM, seed=7) as m:
        th = mo.init_params(8)
        th += 0.05 * np.random.default_rng(9).standard_normal(th.size)
        m.params = th.astype(np.float32)
        yield m, m.params.astype(np.float64)


def test_oracle_adam_constant_gradient():
    """With a constant gradient the bias-corrected moments equal g and g^2 at every step, so
    each step moves a parameter by lr_k g / (|g| + eps): a closed form for the schedule."""
    g = np.array([1.0, -2.0, 1e-3, 0.5, -7.0])
    steps, lr0, lr1 = 6, 1e-2, 1e-3
    th = mo.adam(np.zeros(g.size), lambda t: g, steps, lr0, lr1, 0.95)
    lrs = [lr0 * (lr1 / lr0) ** (k / (steps - 1)) for k in range(steps)]
    assert np.allclose(th, -sum(lrs) * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)


def test_oracle_adam_first_step_on_network_loss():
    """Adam's first step moves every parameter by lr0 g / (|g| + eps) (m_hat = g, v_hat = g^2)."""
    th0 = mo.init_params(5)
    x, y, u = _samples(300, 21)
    _, g = mo.loss_grad(th0, x, y, u, 0.03, DOM)
    th1 = mo.adam_run(th0, x, y, u, 0.03, DOM, 1, 1e-3, 0.0, 0.95)
    assert np.allclose(th1 - th0, -1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-10, atol=1e-18)


@pytest.mark.gpu
def test_params_round_trip(mlp):
    m, th = mlp
    assert np.array_equal(m.params.astype(np.float64), th)


@pytest.mark.gpu
def test_gpu_forward_matches_oracle(mlp):
    m, th = mlp
    x, y, _ = _samples(1000, 10)
    out = m.forward_fp32(x, y)
    ref = mo.forward(th, x, y, DOM)["o"]
    assert np.abs(out - ref).max() < 1e-5 * max(1.0, np.abs(ref).max())


@pytest.mark.gpu
def test_gpu_loss_gradient_matches_oracle(mlp):
    m, th = mlp
    x, y, u = _samples(1000, 11)
    for lam in (0.0, 0.3):
        loss, g = m.loss_grad(x, y, u, lam)
        ref_loss, ref_g = mo.loss_grad(th, x, y, u, lam, DOM)
        assert abs(loss - ref_loss) < 1e-5 * ref_loss + 1e-9
        err = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
        print(f"lam {lam}: loss {loss:.6e} vs {ref_loss:.6e}; grad rel err {err:.2e}", flush=True)
        assert err < 1e-4
        assert np.abs(g - ref_g).max() < 1e-4 * np.abs(ref_g).max()


@pytest.mark.gpu
def test_mfma_inference_matches_bf16_model(mlp):
    m, th = mlp
    x, y, _ = _samples(4099, 12)  # not a multiple of 32: the last wave is partial
    out = m.infer(x, y)
    model = _mfma_model(th, x, y)
    err = np.abs(out[:, :3] - model).max()
    ref = mo.forward(th, x, y, DOM)["o"].T
    err32 = np.abs(out[:, :3] - ref).max()
    print(f"mfma vs bf16 model {err:.2e}; vs fp64 network {err32:.2e}", flush=True)
    assert np.all(out[:, 3] == 0.0)
    assert err < 3e-3
    assert err32 < 5e-2


@pytest.mark.gpu
def test_table_equals_inference_on_grid(mlp):
    m, th = mlp
    t = 64
    m.build_table(t)
    tab = m.read_table()
    xs = DOM[0] + np.arange(t) * (DOM[2] - DOM[0]) / (t - 1)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    ref = m.infer(X.ravel(), Y.ravel()).reshape(t, t, 4)
    assert np.abs(tab - ref).max() < 1e-6


@pytest.mark.gpu
def test_training_reduces_loss(gpu):
    from spac_amd.mlp import MLP

    x, y, _ = _samples(4096, 13)
    u = franke_u(x, y)
    with MLP(DOM, seed=1) as m:
        l0, _ = m.loss_grad(x, y, u, 0.03)
        rep = m.train(x, y, u, steps=3000, lr=3e-3, lr_final=1e-5, grad_weight=0.03)
        print(f"loss {l0:.3e} -> {rep['final_loss']:.3e} in {rep['train_ms']:.1f} ms", flush=True)
        assert rep["final_loss"] < 0.01 * l0
        xv, yv, _ = _samples(2000, 14)
        pred = m.infer(xv, yv)[:, 0]
        assert np.sqrt(np.mean((pred - franke_u(xv, yv)) ** 2)) < 0.03


@pytest.mark.gpu
def test_eager_training_equals_graph_training(gpu):
    from spac_amd.mlp import MLP

    x, y, _ = _samples(1000, 15)
    u = franke_u(x, y)
    out = []
    for eager in (False, True):
        with MLP(DOM, seed=3) as m:
            m.train(x, y, u, steps=120, lr=3e-3, lr_final=1e-4, grad_weight=0.03, eager=eager)
            out.append(m.params)
    assert np.array_equal(out[0], out[1])


@pytest.mark.gpu
def test_gpu_adam_matches_oracle(gpu):
    """110 GPU training steps (two replays of a captured 50-step graph plus a 10-step graph, and
    eagerly) against FP64 Adam on the oracle's gradients from the same start: the step counter
    (bias correction), its parity across graph replays and the learning-rate schedule.  FP32
    gradients leave differences ~2e-8 (p99 ~1e-7) against ~7e-3 of movement; a step counter
    stuck at 1 moves parameters by ~8e-3."""
    from spac_amd.mlp import MLP

    x, y, _ = _samples(512, 17)
    u = franke_u(x, y)
    steps, lr0, lr1 = 110, 1e-3, 3e-4
    for eager in (False, True):
        with MLP(DOM, seed=4) as m:
            th0 = m.params.astype(np.float64)
            m.train(x, y, u, steps=steps, lr=lr0, lr_final=lr1, grad_weight=0.03, eager=eager, beta2=0.95)
            th = m.params.astype(np.float64)
        ref = mo.adam_run(th0, x, y, u, 0.03, DOM, steps, lr0, lr1, 0.95)
        d = np.abs(th - ref)
        print(f"eager={eager}: movement median {np.median(np.abs(ref - th0)):.2e}; "
              f"|gpu - oracle| median {np.median(d):.2e} p99 {np.percentile(d, 99):.2e}", flush=True)
        assert np.median(d) < 1e-6 and np.percentile(d, 99) < 1e-5


@pytest.mark.gpu
def test_training_beta2_option(gpu):
    """beta2 of Adam is selectable (0 = 0.999); out-of-range values are rejected."""
    from spac_amd import abi
    from spac_amd.mlp import MLP

    x, y, _ = _samples(1024, 16)
    u = franke_u(x, y)
    with MLP(DOM, seed=2) as m:
        l0, _ = m.loss_grad(x, y, u, 0.03)
        rep = m.train(x, y, u, steps=400, lr=2e-2, lr_final=1e-4, grad_weight=0.03, beta2=0.95)
        assert rep["final_loss"] < 0.05 * l0
        with pytest.raises(abi.SpacError):
            m.train(x, y, u, steps=10, lr=1e-3, lr_final=1e-4, grad_weight=0.03, beta2=1.5)


def test_first_layer_scale_rejected_out_of_range():
    """A negative or non-finite first-layer scale is an argument error, checked before the device."""
    import ctypes as C
    from spac_amd import abi
    from spac_amd.mlp import MLP  # noqa: F401  (binds the MLP entry points)

    lib = abi.load()
    for bad in (-1.0, float("nan"), 1e6):
        d = abi.MlpDesc()
        d.domain[:] = DOM
        d.seed = 0
        d.first_layer_scale = bad
        h = C.c_void_p()
        assert lib.spac_mlp_create(C.byref(d), C.byref(h)) == abi.SPAC_E_INVALID_ARG
        assert "first_layer_scale" in abi.last_error(lib)


@pytest.mark.gpu
def test_first_layer_scale_initialisation(gpu):
    """first_layer_scale s: W1 and b1 uniform in (-s, s), the deeper layers Glorot with zero
    biases; s = 0: Glorot W1 and zero b1 throughout."""
    from spac_amd.mlp import MLP

    W1, b1, W2, b2, W3, b3, W4, b4 = range(8)
    with MLP(DOM, seed=3, first_layer_scale=2.0) as m:
        p = mo.unpack(m.params)
    assert np.abs(p[W1]).max() < 2.0 and np.abs(p[W1]).max() > 1.5
    assert np.abs(p[b1]).max() < 2.0 and np.abs(p[b1]).max() > 1.5
    assert np.abs(p[W2]).max() <= np.sqrt(6.0 / 128) and not p[b2].any() and not p[b4].any()
    with MLP(DOM, seed=3) as m:
        q = mo.unpack(m.params)
    assert np.abs(q[W1]).max() <= np.sqrt(6.0 / 66) and not q[b1].any()
