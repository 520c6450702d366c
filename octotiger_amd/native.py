"""ctypes binding of liboctofmm.so, the gfx950 C-ABI library (include/octofmm.h).

The library is the only compute path: when it is missing, or no HIP device is
present, every call raises -- there is no CPU fallback.
"""
import ctypes
import os
import weakref

import numpy as np

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboctofmm.so")

EXPORTED = [
    "octofmm_abi_version", "octofmm_status_string", "octofmm_last_error", "octofmm_create",
    "octofmm_destroy", "octofmm_compute_multipole_interactions",
    "octofmm_compute_multipole_interactions_noalloc",
    "octofmm_compute_monopole_interactions", "octofmm_level_plan_create",
    "octofmm_level_plan_execute", "octofmm_level_plan_stats", "octofmm_level_plan_time_stages",
    "octofmm_level_plan_destroy", "octofmm_tree_plan_create", "octofmm_tree_plan_execute",
    "octofmm_tree_plan_stats", "octofmm_tree_plan_time_phases", "octofmm_tree_plan_destroy",
    "octofmm_measure_fma_rate",
]

_lib = None


class OctofmmError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"octofmm status {status}: {msg}")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (python -m octotiger_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.octofmm_abi_version.restype = ctypes.c_int32
        L.octofmm_status_string.restype = ctypes.c_char_p
        L.octofmm_status_string.argtypes = [ctypes.c_int]
        L.octofmm_last_error.restype = ctypes.c_char_p
        L.octofmm_last_error.argtypes = [vp]
        L.octofmm_create.argtypes = [ctypes.POINTER(_abi.Config), ctypes.POINTER(vp)]
        L.octofmm_destroy.argtypes = [vp]
        L.octofmm_compute_multipole_interactions.argtypes = [
            vp, ctypes.POINTER(_abi.Subgrid), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.octofmm_compute_multipole_interactions_noalloc.argtypes = [
            vp, ctypes.POINTER(_abi.Subgrid), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.octofmm_compute_monopole_interactions.argtypes = [
            vp, ctypes.POINTER(_abi.Subgrid), ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(ctypes.c_double)]
        L.octofmm_level_plan_create.argtypes = [vp, ctypes.POINTER(_abi.Level), ctypes.POINTER(vp)]
        L.octofmm_level_plan_execute.argtypes = [vp, vp]
        L.octofmm_level_plan_stats.argtypes = [vp, ctypes.POINTER(_abi.PlanStats)]
        L.octofmm_level_plan_time_stages.argtypes = [vp, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
        L.octofmm_level_plan_destroy.argtypes = [vp]
        L.octofmm_tree_plan_create.argtypes = [vp, ctypes.POINTER(_abi.Tree), ctypes.POINTER(vp)]
        L.octofmm_tree_plan_execute.argtypes = [vp, vp]
        L.octofmm_tree_plan_stats.argtypes = [vp, ctypes.POINTER(_abi.TreeStats)]
        L.octofmm_tree_plan_time_phases.argtypes = [vp, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
        L.octofmm_tree_plan_destroy.argtypes = [vp]
        L.octofmm_measure_fma_rate.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
        for fn in EXPORTED:
            getattr(L, fn).restype = getattr(L, fn).restype or ctypes.c_int
        L.octofmm_status_string.restype = ctypes.c_char_p
        L.octofmm_last_error.restype = ctypes.c_char_p
        L.octofmm_abi_version.restype = ctypes.c_int32
        _lib = L
    return _lib


class Context:
    """One GPU, one opening angle; stencils are built at creation."""

    def __init__(self, theta=0.5, device=0, check_finite=False):
        self._h = ctypes.c_void_p()
        cfg = _abi.Config(theta=theta, device=device, flags=_abi.CHECK_FINITE if check_finite else 0)
        st = lib().octofmm_create(ctypes.byref(cfg), ctypes.byref(self._h))
        if st != _abi.OK:
            raise OctofmmError(st, lib().octofmm_status_string(st).decode())
        self.theta = theta
        self.device = device
        self._plans = weakref.WeakSet()   # plans hold the context: closed before it

    def _check(self, st):
        if st != _abi.OK:
            raise OctofmmError(st, lib().octofmm_last_error(self._h).decode())

    def close(self):
        """Destroys the context, after every plan made from it that is still
        open (a plan must not outlive its context)."""
        if self._h:
            for p in list(self._plans):
                p.close()
            lib().octofmm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def multipole_interactions(self, subgrid, L, L_c):
        """L [512,20] accumulated, L_c [512,3] overwritten (RHO)."""
        self._check(lib().octofmm_compute_multipole_interactions(
            self._h, ctypes.byref(subgrid), _abi.dptr(L), _abi.dptr(L_c)))

    def multipole_interactions_noalloc(self, subgrid, L, L_c):
        """As multipole_interactions, with the library's static device
        workspace (no device allocation)."""
        self._check(lib().octofmm_compute_multipole_interactions_noalloc(
            self._h, ctypes.byref(subgrid), _abi.dptr(L), _abi.dptr(L_c)))

    def monopole_interactions(self, subgrid, contains_multipole_neighbor, L, L_c):
        self._check(lib().octofmm_compute_monopole_interactions(
            self._h, ctypes.byref(subgrid), 1 if contains_multipole_neighbor else 0, _abi.dptr(L),
            _abi.dptr(L_c) if L_c is not None else None))

    def level_plan(self, level, d_M, d_com, d_mon, d_L, d_L_c, type_=_abi.RHO,
                   precision=_abi.PRECISION_FP64, mixed_far_d2=4, flags=_abi.PLAN_USE_GRAPH,
                   targets=None, partial=None):
        """Plan over a workloads.Level whose state lives in device buffers
        (addresses as ints, e.g. torch tensor data_ptr()).  partial: halo
        slots handed over partially, a list of dicts with 'slot', 'cells'
        (int32 cell indices in list order) and the device addresses 'M' and
        'com' (non-leaf slot) or 'mon' (leaf slot) of the compact arrays; the
        plan rewrites those slots' rows of d_M / d_com / d_mon at every
        execute (octofmm_partial_slot)."""
        return LevelPlan(self, level, d_M, d_com, d_mon, d_L, d_L_c, type_, precision, mixed_far_d2,
                         flags, targets, partial)

    def measure_fma_rate(self, precision=_abi.PRECISION_FP64):
        """Dense vector FMA rate of the device in TFLOP/s (FP64, or unpacked
        FP32 with PRECISION_MIXED), from one timed launch of 8 independent
        FMA chains per lane over the whole chip."""
        out = ctypes.c_double()
        self._check(lib().octofmm_measure_fma_rate(self._h, precision, ctypes.byref(out)))
        return float(out.value)

    def tree_plan(self, tree, dev, **kw):
        """Plan of the whole gravity solve over a tree.Tree (see TreePlan)."""
        return TreePlan(self, tree, dev, **kw)


class TreePlan:
    """A whole-tree gravity solve (upward M2M, interactions of every level,
    downward L2L and the leaf force assembly) over device arrays, one set per
    level: dev[l] = dict of int addresses M, com, mon, L, L_c, g."""

    def __init__(self, ctx, tree, dev, type_=_abi.RHO, precision=_abi.PRECISION_FP64, mixed_far_d2=4,
                 flags=_abi.PLAN_USE_GRAPH, G=1.0, correct_am=True):
        self.ctx = ctx
        self._keep = []
        levels = (_abi.TreeLevel * len(tree.levels))()
        for i, lv in enumerate(tree.levels):
            arrs = [np.ascontiguousarray(lv.neighbors, dtype=np.int32), np.ascontiguousarray(lv.kind, dtype=np.int32),
                    np.ascontiguousarray(lv.parent, dtype=np.int32), np.ascontiguousarray(lv.octant, dtype=np.int32),
                    np.ascontiguousarray(lv.xbase, dtype=np.float64)]
            self._keep.append(arrs)
            nb, kd, pa, oc, xb = arrs
            t = levels[i]
            t.nnodes = lv.nnodes
            t.neighbors, t.kind, t.parent, t.octant = _abi.iptr(nb), _abi.iptr(kd), _abi.iptr(pa), _abi.iptr(oc)
            t.xbase = _abi.dptr(xb)
            t.dx = float(lv.dx)
            d = dev[i]
            t.d_M, t.d_com, t.d_mon, t.d_L, t.d_L_c = d["M"], d["com"], d["mon"], d["L"], d["L_c"]
            t.d_g = d.get("g")
        self._levels = levels
        tr = _abi.Tree()
        tr.nlevels = len(tree.levels)
        tr.levels = ctypes.cast(levels, ctypes.POINTER(_abi.TreeLevel))
        tr.type, tr.precision, tr.mixed_far_d2, tr.flags = type_, precision, mixed_far_d2, flags
        tr.G = G
        tr.correct_am = 1 if correct_am else 0
        self._h = ctypes.c_void_p()
        ctx._check(lib().octofmm_tree_plan_create(ctx._h, ctypes.byref(tr), ctypes.byref(self._h)))
        ctx._plans.add(self)

    def execute(self, stream=0):
        self.ctx._check(lib().octofmm_tree_plan_execute(self._h, ctypes.c_void_p(stream)))

    def stats(self):
        s = _abi.TreeStats()
        self.ctx._check(lib().octofmm_tree_plan_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _abi.TreeStats._fields_}

    def time_phases(self, stream=0, iters=5):
        ms = (ctypes.c_double * 3)()
        self.ctx._check(lib().octofmm_tree_plan_time_phases(self._h, ctypes.c_void_p(stream), iters, ms))
        return list(ms)

    def close(self):
        if self._h:
            lib().octofmm_tree_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LevelPlan:
    def __init__(self, ctx, level, d_M, d_com, d_mon, d_L, d_L_c, type_, precision, mixed_far_d2, flags,
                 targets, partial=None):
        self.ctx = ctx
        self._keep = [np.ascontiguousarray(level.neighbors, dtype=np.int32),
                      np.ascontiguousarray(level.kind, dtype=np.int32),
                      np.ascontiguousarray(level.xbase, dtype=np.float64),
                      np.ascontiguousarray(level.targets if targets is None else targets, dtype=np.int32)]
        nb, kd, xb, tg = self._keep
        lv = _abi.Level()
        lv.nslots = int(kd.shape[0])
        lv.neighbors = _abi.iptr(nb)
        lv.kind = _abi.iptr(kd)
        lv.xbase = _abi.dptr(xb)
        lv.dx = float(level.dx)
        lv.ntargets = int(tg.size)
        lv.targets = _abi.iptr(tg)
        lv.type = type_
        lv.precision = precision
        lv.mixed_far_d2 = mixed_far_d2
        lv.flags = flags
        lv.d_M, lv.d_com, lv.d_mon, lv.d_L, lv.d_L_c = d_M, d_com, d_mon, d_L, d_L_c
        if partial:
            ps = (_abi.PartialSlot * len(partial))()
            for e, d in zip(ps, partial):
                cells = np.ascontiguousarray(d["cells"], dtype=np.int32)
                self._keep.append(cells)
                e.slot, e.count, e.cells = int(d["slot"]), int(cells.size), _abi.iptr(cells)
                e.d_M, e.d_com, e.d_mon = d.get("M"), d.get("com"), d.get("mon")
            self._keep.append(ps)
            lv.npartial, lv.partial = len(partial), ps
        self._h = ctypes.c_void_p()
        ctx._check(lib().octofmm_level_plan_create(ctx._h, ctypes.byref(lv), ctypes.byref(self._h)))
        ctx._plans.add(self)

    def execute(self, stream=0):
        self.ctx._check(lib().octofmm_level_plan_execute(self._h, ctypes.c_void_p(stream)))

    def stats(self):
        s = _abi.PlanStats()
        self.ctx._check(lib().octofmm_level_plan_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _abi.PlanStats._fields_}

    def time_stages(self, stream=0, iters=10):
        ms = (ctypes.c_double * 3)()
        self.ctx._check(lib().octofmm_level_plan_time_stages(self._h, ctypes.c_void_p(stream), iters, ms))
        return list(ms)

    def close(self):
        if self._h:
            lib().octofmm_level_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
