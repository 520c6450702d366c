"""Error metrics of the interaction path.

normwise: per output component c, max_cells |a_c - b_c| / max_cells |b_c|; the
1e-12 gate uses this because components such as centre-of-mass dipole terms
cancel towards zero in single cells.  elementwise: max |a - b| / max(|b|, floor).
"""
import numpy as np


def normwise(a, b):
    a = np.asarray(a).reshape(-1, np.asarray(a).shape[-1])
    b = np.asarray(b).reshape(-1, np.asarray(b).shape[-1])
    worst = 0.0
    for c in range(b.shape[1]):
        scale = np.abs(b[:, c]).max()
        d = np.abs(a[:, c] - b[:, c]).max()
        if scale == 0.0:
            if d != 0.0:
                return np.inf
            continue
        worst = max(worst, d / scale)
    return worst


def elementwise(a, b, floor=1e-300):
    return float((np.abs(a - b) / np.maximum(np.abs(b), floor)).max())


def force_torque(L, Lc, mass, x):
    """Total force and torque on the given cells: g = -(L[1..3] + L_c),
    F_i = m_i g_i, tau = sum x_i x F_i (G = 1)."""
    g = -(L[..., 1:4] + Lc)
    F = mass[..., None] * g
    Ft = F.reshape(-1, 3).sum(axis=0)
    tau = np.cross(x.reshape(-1, 3), F.reshape(-1, 3)).sum(axis=0)
    scale_f = np.linalg.norm(F.reshape(-1, 3), axis=1).sum()
    scale_t = np.linalg.norm(np.cross(x.reshape(-1, 3), F.reshape(-1, 3)), axis=1).sum()
    return Ft, tau, scale_f, scale_t


def force_torque_error(L, Lc, Lr, Lcr, mass, x):
    F, t, sf, st = force_torque(L, Lc, mass, x)
    Fr, tr, _, _ = force_torque(Lr, Lcr, mass, x)
    return float(np.linalg.norm(F - Fr) / sf), float(np.linalg.norm(t - tr) / st)
