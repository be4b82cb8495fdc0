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
