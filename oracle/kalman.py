"""Exact log-evidence of the linear-Gaussian model by the Kalman filter.

Test infrastructure (an oracle): for x_t = A x_{t-1} + q z, y_t = x_t + r v,
x_0 ~ N(0, s0^2 I), the marginal likelihood log p(y_0..y_{T-1}) is closed
form.  A bootstrap particle filter's log-evidence estimate (the running sum
of log mean weights, example/include/move_smc.hpp:64-71) converges to it.
"""
import numpy as np


def kalman_logz(A, y, q, r, s0):
    """Returns the running log p(y_0..y_t) for every t (FP64)."""
    A = np.asarray(A, np.float64)
    y = np.asarray(y, np.float64)
    d = A.shape[0]
    Q = q * q * np.eye(d)
    R = r * r * np.eye(d)
    m = np.zeros(d)
    P = s0 * s0 * np.eye(d)
    out = np.zeros(len(y))
    lz = 0.0
    for t in range(len(y)):
        if t > 0:
            m = A @ m
            P = A @ P @ A.T + Q
        S = P + R
        e = y[t] - m
        L = np.linalg.cholesky(S)
        sol = np.linalg.solve(L, e)
        lz += -0.5 * (sol @ sol) - np.log(np.diag(L)).sum() - 0.5 * d * np.log(2 * np.pi)
        out[t] = lz
        K = np.linalg.solve(S, P).T  # P S^-1 (S, P symmetric)
        m = m + K @ e
        P = P - K @ P
    return out
