"""The CPU oracle of the learned EARSM closure, pinned: the network's gradient by central
finite differences, BF16 rounding and SiLU by closed forms, and the numpy restatement of
EARSM1 against the C oracle's BIJ (itself bit for bit the reference Fortran's)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from oracle.oracle_py import OracleBlock
from parity import BIJ, EARSM_CASES, make_earsm_case


def _net(seed=3, n=37):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed).astype(np.float64)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(0, 0.3, M.NH)
    p[M.OFF_B2:M.OFF_W3] = rng.normal(0, 0.3, M.NH)
    p[M.OFF_B3:] = rng.normal(0, 0.3, M.NOUT)
    norm = np.concatenate([rng.normal(0, 1, 5), rng.uniform(0.5, 2, 5), rng.normal(0, 1, 3),
                           rng.uniform(0.5, 2, 3), -np.ones(5), np.ones(5)]).astype(np.float32)
    x = rng.normal(0, 2, (n, 5)).astype(np.float32)
    y = (M.SIGN * np.exp(rng.normal(0, 2, (n, 3)))).astype(np.float32)
    return p, norm, x, y


def test_gradient_matches_central_differences():
    p, norm, x, y = _net()
    loss, g = M.loss_grad(p, norm, x, y, emulate=False)
    rng = np.random.default_rng(7)
    picks = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 8, replace=False),
                            rng.choice(np.arange(M.OFF_B1, M.OFF_W2), 6, replace=False),
                            rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 12, replace=False),
                            rng.choice(np.arange(M.OFF_B2, M.OFF_W3), 6, replace=False),
                            rng.choice(np.arange(M.OFF_W3, M.OFF_B3), 8, replace=False),
                            np.arange(M.OFF_B3, M.NP)])
    h = 1e-6
    for j in picks:
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        fd = (M.loss_grad(pp, norm, x, y, emulate=False)[0]
              - M.loss_grad(pm, norm, x, y, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(g[j])), (j, fd, g[j])


def test_a_dropped_term_is_caught():
    """The finite-difference check fails for a gradient missing SiLU's z s (1 - s) term."""
    p, norm, x, y = _net(seed=5, n=11)
    real = M.dsilu
    try:
        M.dsilu = lambda z: 1.0 / (1.0 + np.exp(-z))
        _, g = M.loss_grad(p, norm, x, y, emulate=False)
    finally:
        M.dsilu = real
    j = M.OFF_W1 + 3
    h = 1e-6
    pp, pm = p.copy(), p.copy()
    pp[j] += h
    pm[j] -= h
    fd = (M.loss_grad(pp, norm, x, y, emulate=False)[0]
          - M.loss_grad(pm, norm, x, y, emulate=False)[0]) / (2 * h)
    assert abs(fd - g[j]) > 1e-4 * max(1.0, abs(fd))


def test_silu_closed_forms():
    z = np.array([0.0, 40.0, -40.0, 1.0])
    np.testing.assert_allclose(M.silu(z), [0.0, 40.0, -40.0 * np.exp(-40.0) / (1 + np.exp(-40.0)),
                                           1.0 / (1.0 + np.exp(-1.0))], rtol=1e-15, atol=1e-300)
    np.testing.assert_allclose(M.dsilu(np.array([0.0, 40.0])), [0.5, 1.0], rtol=1e-15)
    zz = np.linspace(-6, 6, 101)
    h = 1e-6
    np.testing.assert_allclose(M.dsilu(zz), (M.silu(zz + h) - M.silu(zz - h)) / (2 * h), rtol=1e-7,
                               atol=1e-9)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 1.0 + 2.0 ** -7 + 2.0 ** -8],
                 dtype=np.float32)
    # 7 stored mantissa bits; ties go to the even mantissa (1 + 2^-8 down, 1 + 3 2^-8 up)
    assert M.bf16(v).tolist() == [1.0, 1.0, 1.0 + 2.0 ** -7, -2.5, 1.0 + 2.0 ** -6]


def test_emulated_forward_is_close_to_exact():
    p, norm, x, y = _net(seed=9, n=64)
    a = M.predict(p, norm, x, emulate=True)
    b = M.predict(p, norm, x, emulate=False)
    assert np.abs(a - b).max() <= 0.05 * np.abs(b).max()


def test_fit_norm_box():
    rng = np.random.default_rng(4)
    x = rng.normal(3, 5, (500, 5))
    y = M.SIGN * np.exp(rng.normal(-1, 0.1, (500, 3)))
    nrm = M.fit_norm(x, y)
    xn = (M.feature(x) - nrm[0:5]) / nrm[5:10]
    yn = (M.target(y) - nrm[10:13]) / nrm[13:16]
    np.testing.assert_allclose(yn.mean(0), 0, atol=1e-5)
    np.testing.assert_allclose(xn.mean(0), 0, atol=1e-6)
    np.testing.assert_allclose(xn.std(0), 1, rtol=1e-6)
    np.testing.assert_allclose(nrm[16:21], xn.min(0), rtol=1e-6)
    np.testing.assert_allclose(nrm[21:26], xn.max(0), rtol=1e-6)


def _earsm_block(case):
    name, dims, pover, ckw = case
    prm, blk = make_earsm_case(dims, pover, ckw)
    o = OracleBlock(prm, blk)
    o.save_old()
    o.state()
    o.turb()
    S = np.stack([blk.interior(f).ravel() for f in ("S11", "S12", "S13", "S22", "S23", "S33")], 1)
    W = np.stack([blk.interior(f).ravel() for f in ("W12", "W13", "W23")], 1)
    T = blk.interior("TTS").ravel()
    B = np.stack([blk.interior(f).ravel() for f in BIJ], 1)
    return S, W, T, B


@pytest.mark.parametrize("case", EARSM_CASES, ids=[c[0] for c in EARSM_CASES])
def test_earsm_restatement_matches_oracle(case):
    """The numpy EARSM1 gives the C oracle's BIJ (the libm pow/acos/cos of numpy may round
    differently in the last bit)."""
    S, W, T, B = _earsm_block(case)
    inv, beta, terms = M.earsm_parts(S, W, T)
    got = M.bij(beta, terms)
    scale = np.abs(B).max(0)
    assert np.all(np.abs(got - B) <= 1e-12 * scale + 1e-300)
    # both branches of the cubic are exercised by these cases together
    assert np.isfinite(inv).all()


def test_learned_outputs_determine_the_betas():
    """N = -beta6/beta9 and beta3 = 2 IV beta9^2/beta6 hold for the exact coefficients, and
    the betas formed from N alone are EARSM1's, so the learned outputs carry all of b_ij."""
    S, W, T, B = _earsm_block(EARSM_CASES[0])
    inv, beta, terms = M.earsm_parts(S, W, T)
    np.testing.assert_allclose(beta[:, 0], 2.0 * inv[:, 3] * beta[:, 3] ** 2 / beta[:, 2],
                               rtol=1e-12, atol=1e-300)
    rec = M.betas_from_learned(beta[:, 1:], inv)
    np.testing.assert_allclose(rec, beta, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(M.bij(rec, terms), B, rtol=1e-10, atol=1e-12 * np.abs(B).max())
