"""Host helpers of the learned EARSM closure (the kernels are in csrc/ffmlp.hip)."""
import numpy as np

from .ffhip import FF_MLP_NPARAM

NIN, NH, NOUT = 5, 64, 3


def init_params(seed=0):
    """Seeded Xavier-uniform weights and zero biases in the parameter layout of
    include/ffhip.h: W1[64][5], b1[64], W2[64][64], b2[64], W3[3][64], b3[3] (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fo, fi in ((NH, NIN), (NH, NH), (NOUT, NH)):
        a = np.sqrt(6.0 / (fi + fo))
        parts += [rng.uniform(-a, a, size=(fo, fi)).ravel(), np.zeros(fo)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == FF_MLP_NPARAM
    return p
