"""CPU tests pinning the MLP restatement (oracle/mlp_ref.py): the exact
(float64, unrounded) gradient against central finite differences, and the
BF16 rounding against values fixed by the format."""
import numpy as np

from oracle import mlp_ref


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7 + 2 ** -8, -2.5,
                  3.0e38], dtype=np.float32)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-9 rounds up to 1 + 2^-7;
    # 1 + 2^-7 + 2^-8 is a tie between 1 + 2^-7 (odd) and 1 + 2^-6 -> 1 + 2^-6
    expect = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 2.9910792e38],
                      dtype=np.float32)
    np.testing.assert_array_equal(mlp_ref.bf16(x)[:5], expect[:5])
    assert abs(mlp_ref.bf16(x)[5] / x[5] - 1) < 2 ** -8


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = rng.normal(0, 0.3, mlp_ref.NP)
    lo, hi = mlp_ref.HULL_LO, mlp_ref.HULL_HI
    X = lo + (hi - lo) * rng.random((37, 4))
    T = rng.normal(0, 1, 37)
    loss, g = mlp_ref.loss_grad(P, X, T, rounded=False, dtype=np.float64)
    idx = np.r_[rng.choice(256, 8, replace=False), 256 + rng.choice(64, 4, replace=False),
                320 + rng.choice(4096, 8, replace=False), 4416 + rng.choice(64, 4, replace=False),
                4480 + rng.choice(4096, 8, replace=False), 8576 + rng.choice(64, 4, replace=False),
                8640 + rng.choice(64, 4, replace=False), [8704]]
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        lp, _ = mlp_ref.loss_grad(Pp, X, T, rounded=False, dtype=np.float64)
        lm, _ = mlp_ref.loss_grad(Pm, X, T, rounded=False, dtype=np.float64)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_rounded_forward_is_close_to_exact():
    """The BF16 rounding points perturb the output at the 1e-2 level only."""
    rng = np.random.default_rng(4)
    P = rng.normal(0, 0.2, mlp_ref.NP).astype(np.float32)
    lo, hi = mlp_ref.HULL_LO, mlp_ref.HULL_HI
    X = lo + (hi - lo) * rng.random((256, 4))
    y64 = mlp_ref.forward(P.astype(np.float64), X, rounded=False, dtype=np.float64)
    y16 = mlp_ref.forward(P, X, rounded=True)
    assert np.max(np.abs(y16 - y64)) < 3e-2 * max(1.0, np.max(np.abs(y64)))
