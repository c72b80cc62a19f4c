"""The FP32 (1 - b, b) pair of the BGK combinations (split1 in
mpdugks_amd/csrc/fused.h): from a = fl(1 - b), b' = fl(1 - a) and
a' = fl(1 - b').  The FP32 kernels write a x + b x_eq + c x_src as
a' x + (b' x_eq) + c x_src and rely on a' + b' == 1 exactly (the unit
coefficient of x + b (x_eq - x)), which Sterbenz' lemma gives: one of the two
subtractions has an operand in [1/2, 2] and is exact.  Checked here in IEEE
single precision (numpy float32 rounds every operation to nearest even) over
random and edge values of a, with the size of the move b' - (1 - a) that the
pair costs: at most half an ulp of 1 (2^-25), i.e. a relative change of the
relaxation rate below 2^-25 / b.
"""
import numpy as np


def _split(a):
    a = np.asarray(a, dtype=np.float32)
    b = np.float32(1) - a
    omb = np.float32(1) - b
    return omb, b


def test_pair_sums_to_one_exactly():
    rng = np.random.default_rng(5)
    a = np.concatenate([
        rng.uniform(0, 1, 1_000_000).astype(np.float32),
        np.float32(2.0) ** -np.arange(1, 40, dtype=np.float32),            # tiny a (b near 1)
        np.float32(1) - np.float32(2.0) ** -np.arange(2, 25, dtype=np.float32),  # a near 1 (b tiny)
        np.nextafter(np.float32(0.5), np.float32([0, 1]), dtype=np.float32),
        np.float32([0.5, 0.25, 0.75, 1.0, 0.0]),
    ]).astype(np.float32)
    omb, b = _split(a)
    assert omb.dtype == np.float32 and b.dtype == np.float32
    total = omb.astype(np.float64) + b.astype(np.float64)   # exact in FP64
    assert np.all(total == 1.0)
    # the pair moves 1 - a by at most half an ulp of 1
    move = np.abs(b.astype(np.float64) - (1.0 - a.astype(np.float64)))
    assert move.max() <= 2.0 ** -25
    # where a >= 1/2, b is 1 - a exactly and a is kept
    hi = a >= 0.5
    assert np.all(omb[hi] == a[hi])


def test_plain_rounding_misses_one():
    """The pair is needed: 1 - b and b rounded separately do not sum to 1 for
    most b < 1/2 (the bias the FP32 kernels would otherwise add every step)."""
    rng = np.random.default_rng(6)
    bd = rng.uniform(0.01, 0.49, 100_000)
    a = (1.0 - bd).astype(np.float32)
    b = bd.astype(np.float32)
    total = a.astype(np.float64) + b.astype(np.float64)
    assert np.mean(total != 1.0) > 0.5
