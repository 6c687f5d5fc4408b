"""Pins for the shared FP64 primitives (spac_amd/csrc/spac_math.h) on the CPU.

- exp/log/sinpi/cospi/atan2 against libm/numpy on sweeps (ulp bounds);
- Philox4x32-10 against the published Random123 known-answer vectors;
- Box–Muller normals: moments and a Kolmogorov–Smirnov test.
"""
import json
from pathlib import Path

import numpy as np
import pytest
from scipy import stats as sst

import oracle_lib as o

GOLDEN = Path(__file__).resolve().parent / "golden"


def ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    sp = np.spacing(np.abs(b))
    sp[sp == 0] = np.spacing(0.0)
    return np.abs(a - b) / sp


def test_exp_ulp():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-745, 709, 100000), rng.uniform(-60, 1, 100000), rng.uniform(-1e-3, 1e-3, 5000)])
    assert ulps(o.math_eval(0, x), np.exp(x)).max() <= 1.0
    sp = o.math_eval(0, np.array([0.0, 710.0, -746.0, -745.0]))
    assert sp[0] == 1.0 and np.isinf(sp[1]) and sp[2] == 0.0 and sp[3] > 0.0


def test_log_ulp():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(1e-300, 1, 100000), 10 ** rng.uniform(-307, 300, 50000),
                        np.array([5e-324, 1e-310, 1.0, 2.0, 0.5])])
    assert ulps(o.math_eval(1, x), np.log(x)).max() <= 2.0
    assert o.math_eval(1, np.array([1.0]))[0] == 0.0


def test_exp_finite_and_log_normal_equal_the_full_functions():
    """spac_exp_finite (no NaN / overflow selects, one-step 2^k scaling) gives spac_exp's bits for
    every x < 709.78, subnormal results and underflow to zero included; spac_log_normal gives
    spac_log's bits for every normal x > 0."""
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-760, 709.78, 400000), rng.uniform(-746, -700, 100000),
                        np.array([0.0, -0.0, -708.4, -745.1332191019412, -745.14, -746.0, -800.0, 709.7])])
    assert np.array_equal(o.math_eval(9, x).view(np.uint64), o.math_eval(0, x).view(np.uint64))
    assert np.isinf(o.math_eval(9, np.array([709.8, 800.0, 1e300]))).all()
    u = np.concatenate([rng.uniform(2.3e-308, 1.0, 200000), 10 ** rng.uniform(-307, 300, 100000),
                        2.0 ** -np.arange(0, 1022, dtype=np.float64), np.array([1.0, 2.0, 0.5])])
    assert np.array_equal(o.math_eval(10, u).view(np.uint64), o.math_eval(1, u).view(np.uint64))


def _sincospi_ref(x):
    pi = np.longdouble("3.14159265358979323846264338327950288")
    xl = x.astype(np.longdouble)
    q = np.rint(xl * 2)
    r = xl - q / 2
    s, c = np.sin(pi * r), np.cos(pi * r)
    qm = q % 4
    sin = np.select([qm == 0, qm == 1, qm == 2], [s, c, -s], -c)
    cos = np.select([qm == 0, qm == 1, qm == 2], [c, -s, -c], s)
    return sin.astype(np.float64), cos.astype(np.float64)


def test_sincospi_ulp():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-4, 4, 100000), rng.uniform(0, 2, 100000), np.arange(-16, 16) * 0.25])
    rs, rc = _sincospi_ref(x)
    assert ulps(o.math_eval(2, x), rs).max() <= 2.0
    assert ulps(o.math_eval(3, x), rc).max() <= 2.0
    # exact values at quarter points
    q = np.array([0.0, 0.5, 1.0, 1.5])
    assert np.array_equal(o.math_eval(2, q), [0.0, 1.0, 0.0, -1.0])
    assert np.array_equal(o.math_eval(3, q), [1.0, 0.0, -1.0, 0.0])
    # odd / even symmetry: sinpi(-a) = -sinpi(a), cospi(-a) = cospi(a)
    assert np.array_equal(o.math_eval(2, -x), -o.math_eval(2, x))
    assert np.array_equal(o.math_eval(3, -x), o.math_eval(3, x))


def test_atan2_ulp():
    rng = np.random.default_rng(4)
    y = rng.uniform(-3, 3, 200000)
    x = rng.uniform(-3, 3, 200000)
    assert ulps(o.math_eval(4, y, x), np.arctan2(y, x)).max() <= 4.0
    ys = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0])
    xs = np.array([1.0, 1.0, -1.0, -1.0, 0.0, 0.0])
    got = o.math_eval(4, ys, xs)
    ref = np.arctan2(ys, xs)
    assert np.all(ulps(got, ref) <= 1.0) and np.array_equal(np.signbit(got), np.signbit(ref))


def test_philox_known_answers():
    kat = json.loads((GOLDEN / "philox4x32_10_kat.json").read_text())
    assert len(kat["vectors"]) >= 3
    for v in kat["vectors"]:
        out = o.philox(v["rounds"], [int(c, 16) for c in v["ctr"]], [int(k, 16) for k in v["key"]])
        assert [f"{w:08x}" for w in out] == v["out"]


def test_box_muller_normals():
    # 4 streams x 8 paths x 4000 steps = 256k normals
    z = np.concatenate([o.normals(0, s, p, 4000) for s in range(4) for p in range(8)])
    n = z.size
    assert abs(z.mean()) < 4.0 / np.sqrt(n)
    assert abs(z.var() - 1.0) < 4.0 * np.sqrt(2.0 / n)
    assert abs(sst.skew(z)) < 4.0 * np.sqrt(6.0 / n)
    assert sst.kstest(z, "norm").pvalue > 1e-3
    # distinct streams / paths / seeds give distinct sequences
    a, b, c = o.normals(0, 0, 0, 16), o.normals(0, 1, 0, 16), o.normals(1, 0, 0, 16)
    assert not np.array_equal(a, b) and not np.array_equal(a, c)


@pytest.mark.parametrize("seed", [0, 123456789, 2**40 + 7])
def test_normals_come_from_one_philox_block_per_step_pair(seed):
    """Steps 2k and 2k+1 of a path take words (0, 1) and (2, 3) of the Philox block with counter
    (k, path, seed_hi, 0) and key (seed_lo, stream), through Box-Muller on 32-bit uniforms; numpy's
    log/cos/sin against the library's own (a few ulp apart)."""
    stream, path, nsteps = 5, 9, 12
    z = o.normals(seed, stream, path, nsteps).reshape(-1, 2)
    for step in range(nsteps):
        w = o.philox(10, [step // 2, path, seed >> 32, 0], [seed & 0xFFFFFFFF, stream])
        a, b = (w[0], w[1]) if step % 2 == 0 else (w[2], w[3])
        u1, u2 = 1.0 - a * 2.0 ** -32, b * 2.0 ** -32
        rad = np.sqrt(-2.0 * np.log(u1))
        want = np.array([rad * np.cos(2 * np.pi * u2), rad * np.sin(2 * np.pi * u2)])
        assert np.allclose(z[step], want, rtol=1e-13, atol=1e-13), step


@pytest.mark.parametrize("seed", [0, 123456789, 2**40 + 7])
def test_normals_of_consecutive_steps_independent(seed):
    z = o.normals(seed, 5, 9, 20000).reshape(-1, 2)
    for c in (0, 1):  # step 2k against step 2k+1, which share a Philox block
        r = np.corrcoef(z[0::2, c], z[1::2, c])[0, 1]
        assert abs(r) < 4.0 / np.sqrt(z.shape[0] // 2)


@pytest.mark.parametrize("seed", [0, 123456789, 2**40 + 7])
def test_normals_pairs_independent(seed):
    z = o.normals(seed, 5, 9, 20000).reshape(-1, 2)
    r = np.corrcoef(z[:, 0], z[:, 1])[0, 1]
    assert abs(r) < 4.0 / np.sqrt(z.shape[0])


def test_gobet_test_in_squares_equals_square_root_form():
    """Dentro's stop test d < k |N^T sigma| compared in squares, for both signs of k, against
    the square-root form wherever the two sides are not within rounding of each other."""
    rng = np.random.default_rng(11)
    n = 200000
    k = rng.uniform(-0.05, 0.05, n)
    d = rng.uniform(-0.1, 0.1, n)
    d[:2000] = 0.0
    k[2000:4000] = 0.0
    rhs = k * np.sqrt(1.7)
    want = d < rhs
    got = o.math_eval(11, d, k) == 1.0
    far = np.abs(d - rhs) > 1e-12 * (np.abs(d) + np.abs(rhs)) + 1e-300
    far |= (d == 0.0) | (k == 0.0)
    assert far.mean() > 0.99
    assert np.array_equal(got[far], want[far])
