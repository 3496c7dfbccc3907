"""The Kalman log-evidence oracle (oracle/kalman.py) pinned by the closed form
it must equal: y_0..y_{T-1} of the linear-Gaussian model are jointly Gaussian
with mean 0 and covariance Cov(y_s, y_t) = A^(t-s) Var(x_s) (+ r^2 I on the
diagonal), Var(x_0) = s0^2 I, Var(x_t) = A Var(x_{t-1}) A^T + q^2 I.  The
joint density evaluated directly must match the filter's running sum."""
import numpy as np
import pytest

import vsmc_amd as V
from kalman import kalman_logz


def _joint_logpdf(A, y, q, r, s0):
    T, d = y.shape
    V_ = [s0 * s0 * np.eye(d)]
    for t in range(1, T):
        V_.append(A @ V_[-1] @ A.T + q * q * np.eye(d))
    S = np.zeros((T * d, T * d))
    for s in range(T):
        for t in range(s, T):
            C = np.linalg.matrix_power(A, t - s) @ V_[s]
            S[t * d:(t + 1) * d, s * d:(s + 1) * d] = C
            S[s * d:(s + 1) * d, t * d:(t + 1) * d] = C.T
        S[s * d:(s + 1) * d, s * d:(s + 1) * d] += r * r * np.eye(d)
    v = y.reshape(-1)
    sign, logdet = np.linalg.slogdet(S)
    assert sign > 0
    return -0.5 * (v @ np.linalg.solve(S, v)) - 0.5 * logdet - 0.5 * T * d * np.log(2 * np.pi)


@pytest.mark.parametrize("d,T", [(1, 1), (2, 5), (3, 7), (10, 4)])
def test_kalman_equals_joint_gaussian(d, T):
    A = V.lg_matrix(d, seed=d)
    y = V.lg_simulate(A, T, seed=T, q=0.3, r=0.2, s0=0.5)
    lz = kalman_logz(A, y, 0.3, 0.2, 0.5)
    for t in range(T):
        assert abs(lz[t] - _joint_logpdf(A, y[:t + 1], 0.3, 0.2, 0.5)) < 1e-9 * (1 + abs(lz[t]))


def test_kalman_scalar_closed_form():
    """d = 1, T = 1: y_0 ~ N(0, s0^2 + r^2)."""
    y = np.array([[0.7]])
    lz = kalman_logz(np.array([[0.5]]), y, 0.1, 0.2, 0.3)
    v = 0.3 ** 2 + 0.2 ** 2
    assert abs(lz[0] - (-0.5 * 0.49 / v - 0.5 * np.log(2 * np.pi * v))) < 1e-14


def test_lg_matrix_is_stable():
    A = V.lg_matrix(10, seed=4)
    assert abs(np.max(np.abs(np.linalg.eigvals(A))) - 0.9) < 1e-12
