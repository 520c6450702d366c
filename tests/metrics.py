"""Error metrics of the interaction path.

normwise: per output component c, max_cells |a_c - b_c| / max_cells |b_c|; the
1e-12 gate uses this because components such as centre-of-mass dipole terms
cancel towards zero in single cells.  elementwise: max |a - b| / max(|b|, floor).

Total force and torque: F = sum_i m_i g_i with g_i = -(L_i[1..3] + L_c,i) and
tau = sum_i x_i x F_i (G = 1, the leaf assembly of src/grid_fmm.cpp:1613-1629).
The gate is the error of each total relative to the reference's own total,
|F - F_ref| / |F_ref| and |tau - tau_ref| / |tau_ref|.

Rounding floor.  Where the totals cancel (Newton's third law makes the total
force of an isolated sub-grid vanish up to the truncation error: on the root
sub-grid |F_ref| is 1e-4 of sum_i |F_i|), two FP64 evaluations that differ only
in summation order cannot agree on the total to better than the rounding of
the per-cell values it is summed from.  The gate therefore accepts an error up
to max(tol * |F_ref|, 4 eps sum_i |F_i|), and reports both.  Where the totals
do not cancel that far (sum_i |F_i| / |F_ref| < tol / 4 eps, about 1100 for
tol = 1e-12) the floor does not enter.  The error relative to sum_i |F_i| is
kept as a secondary figure.
"""
import numpy as np

EPS = np.finfo(np.float64).eps
FLOOR_ULPS = 4.0


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


def _ratio(d, n):
    return float(d / n) if n > 0 else (0.0 if d == 0 else np.inf)


def force_torque_errors(L, Lc, Lr, Lcr, mass, x):
    """Errors of the total force and torque against the reference totals:
    'force' = |F - F_ref| / |F_ref|, 'torque' = |tau - tau_ref| / |tau_ref|;
    'force_floor' / 'torque_floor' = the rounding floor 4 eps sum_i |F_i| in
    the same units; the same differences over the sums of per-cell magnitudes
    ('force_per_cell_sum', 'torque_per_cell_sum'), and those sums over the
    totals ('force_cancellation', 'torque_cancellation')."""
    F, t, _, _ = force_torque(L, Lc, mass, x)
    Fr, tr, sf, st = force_torque(Lr, Lcr, mass, x)
    nF, nt = np.linalg.norm(Fr), np.linalg.norm(tr)
    dF, dt = np.linalg.norm(F - Fr), np.linalg.norm(t - tr)
    return {
        "force": _ratio(dF, nF),
        "torque": _ratio(dt, nt),
        "force_floor": _ratio(FLOOR_ULPS * EPS * sf, nF),
        "torque_floor": _ratio(FLOOR_ULPS * EPS * st, nt),
        "force_per_cell_sum": _ratio(dF, sf),
        "torque_per_cell_sum": _ratio(dt, st),
        "force_cancellation": _ratio(sf, nF),
        "torque_cancellation": _ratio(st, nt),
    }


def force_torque_gate(L, Lc, Lr, Lcr, mass, x, tol):
    """(passed, errors): total force and torque each within
    max(tol, rounding floor) of the reference totals, relative to them."""
    e = force_torque_errors(L, Lc, Lr, Lcr, mass, x)
    ok = e["force"] <= max(tol, e["force_floor"]) and e["torque"] <= max(tol, e["torque_floor"])
    return ok, e


def describe(e):
    return (f"force {e['force']:.3e} (floor {e['force_floor']:.1e}) "
            f"torque {e['torque']:.3e} (floor {e['torque_floor']:.1e}); "
            f"per-cell sums {e['force_per_cell_sum']:.3e} {e['torque_per_cell_sum']:.3e}")
