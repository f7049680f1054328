"""CPU oracle of the learned EARSM closure (TEST INFRASTRUCTURE ONLY).

Two parts:

* `earsm_parts`: a numpy restatement of EARSM1's invariants, cubic solve, N correction and
  betas (earsm.f:72-205) and of its b_ij assembly (earsm.f:209-219).  It is pinned against
  the C oracle's BIJ (itself pinned bit for bit against the reference Fortran) in
  tests/test_mlp_oracle.py.
* the surrogate network 5 -> 64 -> 64 -> 3 with SiLU: forward, loss and gradient in float64,
  optionally emulating the device kernels' BF16 rounding of MFMA operands (weights,
  normalised inputs, activations and back-propagated signals; FP32 accumulation and biases).
  The gradient is pinned by central finite differences (tests/test_mlp_oracle.py).

The network's inputs are the five strain/rotation invariants of the T-scaled tensors
(II_S = SII, II_W = WII, III_S = tr S^3, IV = SWWIV, V = SSWWV), seen through asinh; its
outputs are log(sign * beta) of the coefficients (beta4, beta6, beta9), whose signs are fixed
(beta4, beta6 < 0 < beta9).  The closure takes N = -beta6/beta9 from them (the exact model
has beta6/beta9 = -N, earsm.f:202-205) and forms the four betas from N with EARSM1's own
formulas, so b_ij follows from the same tensor terms as in EARSM1.
"""
import numpy as np

NIN, NH, NOUT = 5, 64, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NOUT * NH
NP = OFF_B3 + NOUT                 # 4739 parameters
# normalisation block: input mean (5), input scale (5), output mean (3), output scale (3),
# training hull of the normalised inputs: low (5), high (5)
NNORM = 26


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NH, NIN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(NOUT, NH), p[OFF_B3:NP])


def init_params(seed=0):
    """Xavier-uniform weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, dtype=np.float32)

    def xav(fo, fi):
        a = np.sqrt(6.0 / (fi + fo))
        return rng.uniform(-a, a, size=(fo, fi)).astype(np.float32)
    p[OFF_W1:OFF_B1] = xav(NH, NIN).ravel()
    p[OFF_W2:OFF_B2] = xav(NH, NH).ravel()
    p[OFF_W3:OFF_B3] = xav(NOUT, NH).ravel()
    return p


def bf16(a):
    """Round float32 values to bfloat16 (round to nearest even), returned as float32."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


SIGN = np.array([-1.0, -1.0, 1.0])


def feature(x):
    """The network's view of the invariants: asinh of each (float32, as on the device)."""
    return np.arcsinh(np.asarray(x, dtype=np.float32))


def target(y):
    """log(sign * beta) of (beta4, beta6, beta9), whose signs are fixed (float32)."""
    y = np.asarray(y, dtype=np.float32)
    return np.log(np.maximum(SIGN.astype(np.float32) * y, np.float32(1e-37)))


def fit_norm(x, y):
    """Means and scales of the transformed inputs and outputs, and the box of the
    normalised inputs."""
    x = feature(x).astype(np.float64)
    y = target(y).astype(np.float64)
    xm, xs = x.mean(0), x.std(0)
    ym, ys = y.mean(0), y.std(0)
    xs = np.where(xs > 0, xs, 1.0)
    ys = np.where(ys > 0, ys, 1.0)
    xn = (x - xm) / xs
    return np.concatenate([xm, xs, ym, ys, xn.min(0), xn.max(0)]).astype(np.float32)


def _q(a, emulate):
    return bf16(a).astype(np.float64) if emulate else np.asarray(a, dtype=np.float64)


def forward(p, norm, x, emulate=True):
    """Normalised outputs yn (n, 3) and the intermediate values of the pass."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    norm = np.asarray(norm, dtype=np.float64)
    xm, xs = norm[0:5], norm[5:10]
    xn = (feature(x) - xm.astype(np.float32)) / xs.astype(np.float32)
    xq = _q(xn, emulate)
    W1q, W2q, W3q = _q(W1, emulate), _q(W2, emulate), _q(W3, emulate)
    z1 = xq @ W1q.T + b1
    h1 = silu(z1)
    h1q = _q(h1, emulate)
    z2 = h1q @ W2q.T + b2
    h2 = silu(z2)
    h2q = _q(h2, emulate)
    yn = h2q @ W3q.T + b3
    return yn, dict(xq=xq, z1=z1, h1q=h1q, z2=z2, h2q=h2q, W1q=W1q, W2q=W2q, W3q=W3q)


def predict(p, norm, x, emulate=True):
    """Physical outputs (beta4, beta6, beta9)."""
    norm = np.asarray(norm, dtype=np.float64)
    yn, _ = forward(p, norm, x, emulate)
    return SIGN * np.exp(yn * norm[13:16] + norm[10:13])


def loss_grad(p, norm, x, y, emulate=True):
    """Mean over samples of the squared error of the normalised outputs, summed over the
    three outputs, and its gradient with respect to the parameters."""
    norm = np.asarray(norm, dtype=np.float64)
    n = len(x)
    yn, c = forward(p, norm, x, emulate)
    tn = (target(y) - norm[10:13].astype(np.float32)) / norm[13:16].astype(np.float32)
    r = yn - tn
    loss = float((r * r).sum() / n)
    dy = 2.0 * r / n
    dyq = _q(dy, emulate)
    dh2 = dyq @ c["W3q"]
    dz2 = dh2 * dsilu(c["z2"])
    dz2q = _q(dz2, emulate)
    dh1 = dz2q @ c["W2q"]
    dz1 = dh1 * dsilu(c["z1"])
    dz1q = _q(dz1, emulate)
    g = np.zeros(NP)
    g[OFF_W1:OFF_B1] = (dz1q.T @ c["xq"]).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W2:OFF_B2] = (dz2q.T @ c["h1q"]).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W3:OFF_B3] = (dyq.T @ c["h2q"]).ravel()
    g[OFF_B3:NP] = dy.sum(0)
    return loss, g


# ---------------------------------------------------------------------------------------
# EARSM1 restated (earsm.f:72-219), vectorised over cells
# ---------------------------------------------------------------------------------------
def earsm_parts(S, W, T):
    """S: (n, 6) = S11 S12 S13 S22 S23 S33; W: (n, 3) = W12 W13 W23; T: (n,) the time scale.
    Returns the five invariants (n, 5), the exact betas (n, 4) = beta3 beta4 beta6 beta9 and
    the tensor terms (n, 4, 5): for beta3, beta4, beta6, beta9 the components 11, 22, 12, 13,
    23."""
    S = np.asarray(S, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    p3 = 1.0 / 3.0
    p6, tt = 0.5 * p3, 2.0 * p3
    c1 = 1.8
    c1p = 9.0 * (c1 - 1.0) / 4.0
    c1p3, c1psq = c1p * p3, c1p * c1p
    s11, s12, s13, s22, s23, s33 = (T * S[:, q] for q in range(6))
    w12, w13, w23 = (T * W[:, q] for q in range(3))
    s11sq, s22sq, s33sq = s11 * s11, s22 * s22, s33 * s33
    s12sq, s13sq, s23sq = s12 * s12, s13 * s13, s23 * s23
    w12sq, w13sq, w23sq = w12 * w12, w13 * w13, w23 * w23
    sii = s11sq + s22sq + s33sq + 2.0 * (s12sq + s13sq + s23sq)
    wii = -(2.0 * (w12sq + w13sq + w23sq))
    wiip3 = p3 * wii
    swwiv = (-(s11 * (w12sq + w13sq)) - s22 * (w12sq + w23sq) - s33 * (w13sq + w23sq)
             + 2.0 * (-(s12 * w13 * w23) + s13 * w12 * w23 - s23 * w12 * w13))
    swwivtt = tt * swwiv
    sswwv = (2.0 * (-((s12 * s13 + s22 * s23 + s23 * s33) * w12 * w13)
                    + (s11 * s13 + s12 * s23 + s13 * s33) * w12 * w23
                    - (s11 * s12 + s12 * s22 + s13 * s23) * w13 * w23)
             - (s11sq + s12sq + s13sq) * (w12sq + w13sq)
             - (s12sq + s22sq + s23sq) * (w12sq + w23sq)
             - (s13sq + s23sq + s33sq) * (w13sq + w23sq))
    iiis = (s11 * s11sq + s22 * s22sq + s33 * s33sq
            + 3.0 * (s11 * (s12sq + s13sq) + s22 * (s12sq + s23sq) + s33 * (s13sq + s23sq))
            + 6.0 * s12 * s13 * s23)
    t3 = [-w12sq - w13sq - wiip3, -w12sq - w23sq - wiip3, -(w13 * w23), w12 * w23, -(w12 * w13)]
    t4 = [-(2.0 * (s12 * w12 + s13 * w13)), 2.0 * (s12 * w12 - s23 * w23),
          (s11 - s22) * w12 - s23 * w13 - s13 * w23,
          -(s23 * w12) + (s11 - s33) * w13 + s12 * w23,
          s13 * w12 + s12 * w13 + (s22 - s33) * w23]
    t6 = [-(2.0 * ((s12 * w13 - s13 * w12) * w23 + s11 * (w12sq + w13sq))) - swwivtt - wii * s11,
          -(2.0 * ((s23 * w12 + s12 * w23) * w13 + s22 * (w12sq + w23sq))) - swwivtt - wii * s22,
          -(s12 * (2.0 * w12sq + w13sq + w23sq)) - (s13 * w13 - s23 * w23) * w12
          - (s11 + s22) * w13 * w23 - wii * s12,
          -(s13 * (w12sq + 2.0 * w13sq + w23sq)) - (s12 * w12 + s23 * w23) * w13
          + (s11 + s33) * w12 * w23 - wii * s13,
          -(s23 * (w12sq + w13sq + 2.0 * w23sq)) + (s12 * w12 - s13 * w13) * w23
          - (s22 + s33) * w12 * w13 - wii * s23]
    t9 = [-(2.0 * ((s12 * w12 + s13 * w13 - s23 * w23) * w12sq
                   + (s12 * w12 + s13 * w13 + s23 * w23) * w13sq
                   + (s22 - s33) * w12 * w13 * w23)),
          -(2.0 * ((-(s12 * w12) - s13 * w13 + s23 * w23) * w12sq
                   + (-(s12 * w12) + s13 * w13 + s23 * w23) * w23sq
                   + (-s11 + s33) * w12 * w13 * w23)),
          ((s11 - s22) * w12 - 2.0 * (s13 * w23 + s23 * w13)) * w12sq
          + ((s11 - s33) * w12 - 2.0 * s13 * w23) * w13sq
          + ((s33 - s22) * w12 - 2.0 * s23 * w13) * w23sq,
          ((s11 - s22) * w13 + 2.0 * s12 * w23) * w12sq
          + ((s11 - s33) * w13 + 2.0 * (s12 * w23 - s23 * w12)) * w13sq
          + ((s22 - s33) * w13 - 2.0 * s23 * w12) * w23sq,
          ((s22 - s11) * w23 + 2.0 * s12 * w13) * w12sq
          + ((s11 - s33) * w23 + 2.0 * s13 * w12) * w13sq
          + ((s22 - s33) * w23 + 2.0 * (s12 * w13 + s13 * w12)) * w23sq]
    # N_c from the cubic (earsm.f:170-184)
    p1 = (c1psq / 27.0 + 0.45 * sii - tt * wii) * c1p
    cb = c1psq * (p3 * p3) + 0.9 * sii + tt * wii
    p2 = p1 * p1 - cb * cb * cb
    fnc = np.empty_like(p1)
    pos = p2 >= 0.0
    with np.errstate(invalid="ignore"):
        sqp2 = np.sqrt(np.where(pos, p2, 0.0))
        pm = p1 - sqp2
        pmp = np.abs(pm) ** p3
        fnc_pos = c1p3 + (p1 + sqp2) ** p3 + np.copysign(pmp, pm)
        pmn = p1 * p1 - p2
        sqpm = np.sqrt(np.where(pos, 1.0, pmn))
        alpo = np.clip(p1 / sqpm, -1.0, 1.0)
        fnc_neg = c1p3 + 2.0 * np.where(pos, 1.0, pmn) ** p6 * np.cos(p3 * np.arccos(alpo))
    fnc = np.where(pos, fnc_pos, fnc_neg)
    # N correction and betas (earsm.f:187-205)
    dd = (20.0 * (fnc * fnc * fnc * fnc) * (fnc - 0.5 * c1p)
          - wii * (10.0 * fnc + 15.0 * c1p) * (fnc * fnc) + 10.0 * c1p * (wii * wii))
    fn = fnc + 162.0 * (swwiv * swwiv + (sswwv - 0.5 * sii * wii) * (fnc * fnc)) / dd
    fnsq = fn * fn
    q = 5.0 * p6 * (fnsq - 2.0 * wii) * (2.0 * fnsq - wii)
    pq = 1.0 / q
    beta = np.stack([-(pq * 12.0 * swwiv / fn), -(pq * 2.0 * (fnsq - 2.0 * wii)),
                     -(pq * 6.0 * fn), pq * 6.0], axis=1)
    inv = np.stack([sii, wii, iiis, swwiv, sswwv], axis=1)
    terms = np.stack([np.stack(t, axis=1) for t in (t3, t4, t6, t9)], axis=1)
    return inv, beta, terms


def bij(beta, terms):
    """b_ij (n, 6) = B11 B12 B13 B22 B23 B33 from the four betas and the tensor terms
    (earsm.f:209-219)."""
    b = 0.5 * (beta[:, 0:1] * terms[:, 0] + beta[:, 1:2] * terms[:, 1]
               + beta[:, 2:3] * terms[:, 2] + beta[:, 3:4] * terms[:, 3])
    b11, b22, b12, b13, b23 = (b[:, q] for q in range(5))
    return np.stack([b11, b12, b13, b22, b23, -b11 - b22], axis=1)


def betas_from_n(fn, wii, swwiv):
    """(beta3, beta4, beta6, beta9) from N (earsm.f:196-205)."""
    p6 = 0.5 * (1.0 / 3.0)
    fnsq = fn * fn
    q = 5.0 * p6 * (fnsq - 2.0 * wii) * (2.0 * fnsq - wii)
    pq = 1.0 / q
    return np.stack([-(pq * 12.0 * swwiv / fn), -(pq * 2.0 * (fnsq - 2.0 * wii)),
                     -(pq * 6.0 * fn), pq * 6.0], axis=1)


def betas_from_learned(y, inv):
    """The closure's (beta3, beta4, beta6, beta9): N = -beta6/beta9 of the learned
    coefficients y = (beta4, beta6, beta9), then EARSM1's formulas; inv are the invariants
    (SII, WII, III_S, IV, V)."""
    y = np.asarray(y, dtype=np.float64)
    inv = np.asarray(inv, dtype=np.float64)
    return betas_from_n(-y[:, 1] / y[:, 2], inv[:, 1], inv[:, 3])
