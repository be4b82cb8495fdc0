"""Shared helpers of the parity tests (tolerance rule, FNV, closed forms)."""
import numpy as np

# Regrets and strategy sums after a launch must satisfy
#   |x - x_hat| <= REL_TOL * max(|x_hat|, sum|delta|)
# where x_hat = pre-launch value + FP64 sum of the launch's contributions and
# sum|delta| is the FP64 L1 mass of those contributions (the denominator
# matters because regret deltas cancel: +u/q - u against -u).
REL_TOL = 1e-5


def within_tol(x, x_hat, l1):
    x = np.asarray(x, np.float64)
    x_hat = np.asarray(x_hat, np.float64)
    bound = REL_TOL * np.maximum(np.abs(x_hat), np.asarray(l1, np.float64))
    return np.abs(x - x_hat) <= bound


def fnv1a(data: bytes) -> int:
    h = 2166136261
    for b in data:
        h = ((h ^ b) * 16777619) & 0xFFFFFFFF
    return h


def expected_post(R, S, dR, dS, cfr_plus=True):
    r = R.astype(np.float64) + dR
    if cfr_plus:
        r = np.maximum(r, 0.0)
    return r, S.astype(np.float64) + dS


def check_launch_sparse(R0, S0, R1, S1, idx, dR, dS, aR, aS, cfr_plus=True):
    """A launch's tables against its sparse FP64 replay: every touched entry
    (idx) within tolerance of pre-launch value + contributions (regrets
    clamped at 0 under CFR+), every other entry unchanged bit for bit.
    Overwrites the touched entries of R1 and S1 (full-size tables are not
    copied)."""
    eR = R0[idx].astype(np.float64) + dR
    if cfr_plus:
        eR = np.maximum(eR, 0.0)
    eS = S0[idx].astype(np.float64) + dS
    assert within_tol(R1[idx], eR, aR).all()
    assert within_tol(S1[idx], eS, aS).all()
    R1[idx] = R0[idx]
    S1[idx] = S0[idx]
    assert np.array_equal(R1.view(np.uint32), R0.view(np.uint32))
    assert np.array_equal(S1.view(np.uint32), S0.view(np.uint32))
