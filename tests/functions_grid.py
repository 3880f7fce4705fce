"""Argument grids of the function-level comparisons (exact and tabulated
DCS, dilogarithm, DEL polar angles) shared by the CPU and GPU tests."""
import numpy as np

PROCESSES = {0: "bremsstrahlung", 1: "pair production", 2: "photonuclear",
             3: "ionisation"}


def dcs_grid():
    """(K, q) over 1e-2..1e9 GeV and q / K over 1e-7..1 (log spaced)."""
    K = np.logspace(-2, 9, 45)
    nu = np.logspace(-7, 0, 36)
    kk, vv = np.meshgrid(K, nu, indexing="ij")
    return kk.ravel().copy(), (kk * vv).ravel().copy()


def dilog_grid():
    x = np.concatenate([np.linspace(-30.0, 30.0, 40001),
                        -np.logspace(-12, 1, 2001), np.logspace(-12, 1, 2001),
                        [-2.0, -1.0, -0.5, 0.0, 0.5, 1.0, 2.0, 10.0]])
    return x, np.zeros_like(x)


def polar_grid():
    """ki over 1e-1..1e6 GeV, kf = ki (1 - nu), nu over 1e-3..0.999."""
    ki = np.logspace(-1, 6, 40)
    nu = np.concatenate([np.logspace(-3, -0.01, 24), [0.5, 0.9, 0.999]])
    kk, vv = np.meshgrid(ki, nu, indexing="ij")
    return kk.ravel().copy(), (kk * (1.0 - vv)).ravel().copy()
