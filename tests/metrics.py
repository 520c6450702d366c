"""Error metrics of the interaction path.

normwise: per output component c, max_cells |a_c - b_c| / max_cells |b_c|; the
1e-12 gate uses this because components such as centre-of-mass dipole terms
cancel towards zero in single cells.  elementwise: max |a - b| / max(|b|, floor).

Total force and torque: F = sum_i m_i g_i with g_i = -(L_i[1..3] + L_c,i) and
tau = sum_i x_i x F_i (G = 1, the leaf assembly of src/grid_fmm.cpp:1613-1629).
The gate is the error of each total relative to the reference's own total,
|F - F_ref| / |F_ref| and |tau - tau_ref| / |tau_ref|.  The same error
relative to the sum of per-cell magnitudes (sum_i |F_i|, sum_i |x_i x F_i|) is
kept as a secondary figure: it is smaller by that sum over the total, which is
25-66 on the single-sub-grid cases.
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
    """Total force and torque on the given cells, and the sums of the per-cell
    magnitudes: (F, tau, sum |F_i|, sum |x_i x F_i|)."""
    g = -(L[..., 1:4] + Lc)
    F = (mass[..., None] * g).reshape(-1, 3)
    xf = x.reshape(-1, 3)
    tq = np.cross(xf, F)
    return F.sum(axis=0), tq.sum(axis=0), np.linalg.norm(F, axis=1).sum(), np.linalg.norm(tq, axis=1).sum()


def force_torque_errors(L, Lc, Lr, Lcr, mass, x):
    """Errors of the total force and torque against the reference totals:
    'force' = |F - F_ref| / |F_ref|, 'torque' = |tau - tau_ref| / |tau_ref|
    (the gate), and the same differences over the sums of per-cell magnitudes
    ('force_per_cell_sum', 'torque_per_cell_sum'), with those sums over the
    totals ('force_cancellation', 'torque_cancellation')."""
    F, t, _, _ = force_torque(L, Lc, mass, x)
    Fr, tr, sf, st = force_torque(Lr, Lcr, mass, x)
    nF, nt = np.linalg.norm(Fr), np.linalg.norm(tr)
    dF, dt = np.linalg.norm(F - Fr), np.linalg.norm(t - tr)
    return {
        "force": float(dF / nF) if nF > 0 else (0.0 if dF == 0 else np.inf),
        "torque": float(dt / nt) if nt > 0 else (0.0 if dt == 0 else np.inf),
        "force_per_cell_sum": float(dF / sf) if sf > 0 else 0.0,
        "torque_per_cell_sum": float(dt / st) if st > 0 else 0.0,
        "force_cancellation": float(sf / nF) if nF > 0 else np.inf,
        "torque_cancellation": float(st / nt) if nt > 0 else np.inf,
    }


def force_torque_error(L, Lc, Lr, Lcr, mass, x):
    """(|F - F_ref| / |F_ref|, |tau - tau_ref| / |tau_ref|): the gate."""
    e = force_torque_errors(L, Lc, Lr, Lcr, mass, x)
    return e["force"], e["torque"]
