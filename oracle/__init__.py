"""Python binding of the CPU FP64 oracle (oracle/liboctofmm_oracle.so) and of
the reference-template harness (oracle/_ref/libocto_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline.
The product (octotiger_amd) never imports this module.
"""
import ctypes
import os

import numpy as np

from octotiger_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboctofmm_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libocto_ref.so")
REF_SIMD_PATHS = {8: os.path.join(HERE, "_ref", "libocto_ref_simd8.so"),
                  4: os.path.join(HERE, "_ref", "libocto_ref_simd4.so")}

_lib = None
_ref = None
_ref_simd = None

_dp = ctypes.POINTER(ctypes.c_double)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_ip = ctypes.POINTER(ctypes.c_int32)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_build_stencils.argtypes = [ctypes.c_double, _u8, _u8, _u8, _dp]
        L.oracle_multipole_stencil.argtypes = [ctypes.c_double, _ip, _u8, _ip]
        L.oracle_d_factors.argtypes = [_dp, _dp]
        L.oracle_pair_m2l.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int32]
        L.oracle_pair_p2m.argtypes = [_dp, _dp, _dp, _dp, _dp, ctypes.c_int32]
        L.oracle_multipole_interactions.argtypes = [ctypes.c_double, ctypes.POINTER(_abi.Subgrid), _dp, _dp]
        L.oracle_monopole_interactions.argtypes = [ctypes.c_double, ctypes.POINTER(_abi.Subgrid),
                                                   ctypes.c_int32, _dp, _dp]
        L.oracle_level_interactions.argtypes = [
            ctypes.c_double, ctypes.c_int32, _ip, _ip, _dp, ctypes.c_double, ctypes.c_int32, _ip,
            ctypes.c_int32, _dp, _dp, _dp, _dp, _dp, ctypes.c_int32]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference harness missing: {REF_PATH}")
        R = ctypes.CDLL(REF_PATH)
        R.ref_build_stencils.argtypes = [ctypes.c_double, _u8, _u8, _u8, _dp]
        R.ref_d_factors.argtypes = [_dp, _dp]
        R.ref_pair_m2l.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int32]
        R.ref_pair_p2m.argtypes = [_dp, _dp, _dp, _dp, _dp, ctypes.c_int32]
        R.ref_multipole_interactions.argtypes = [ctypes.c_double, ctypes.POINTER(_abi.Subgrid), _dp, _dp]
        R.ref_monopole_interactions.argtypes = [ctypes.c_double, ctypes.POINTER(_abi.Subgrid),
                                                ctypes.c_int32, _dp, _dp]
        R.ref_level_interactions.argtypes = [
            ctypes.c_double, ctypes.c_int32, _ip, _ip, _dp, ctypes.c_double, ctypes.c_int32, _ip,
            ctypes.c_int32, _dp, _dp, _dp, _dp, _dp, ctypes.c_int32]
        _ref = R
    return _ref


def host_simd_width():
    """Vector width (doubles) of the host: 8 with AVX-512F, 4 with AVX2 and
    FMA, else 0 (read from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            flags = next((ln.split(":", 1)[1].split() for ln in f if ln.startswith("flags")), [])
    except OSError:
        return 0
    if {"avx512f", "avx512dq", "avx512vl", "avx512bw"} <= set(flags):
        return 8
    if {"avx2", "fma"} <= set(flags):
        return 4
    return 0


def ref_simd_available():
    w = host_simd_width()
    return w in REF_SIMD_PATHS and os.path.exists(REF_SIMD_PATHS[w])


def ref_simd_lib():
    """The timed CPU baseline build of the reference templates for this host's
    vector width (oracle/_ref/libocto_ref_simd{8,4}.so)."""
    global _ref_simd
    if _ref_simd is None:
        w = host_simd_width()
        if w not in REF_SIMD_PATHS or not os.path.exists(REF_SIMD_PATHS[w]):
            raise RuntimeError(f"no vector build of the reference harness for this host (width {w})")
        R = ctypes.CDLL(REF_SIMD_PATHS[w])
        R.ref_simd_width.restype = ctypes.c_int32
        R.ref_level_interactions_simd.argtypes = [
            ctypes.c_double, ctypes.c_int32, _ip, _ip, _dp, ctypes.c_double, ctypes.c_int32, _ip,
            ctypes.c_int32, _dp, _dp, _dp, _dp, _dp, ctypes.c_int32]
        _ref_simd = R
    return _ref_simd


def level_interactions_simd(level, theta=0.5, type_=_abi.RHO, targets=None, nthreads=None):
    """The reference's interaction templates walked as the Kokkos host kernel
    walks them (z-rows of 8 cells as vectors of the host's width), -O3,
    OpenMP over the target sub-grids: the timed CPU baseline.  Non-leaf
    targets only.  Returns (L, L_c) shaped like level.M / level.com."""
    targets = level.targets if targets is None else np.ascontiguousarray(targets, dtype=np.int32)
    L = np.zeros((level.nslots, 512, 20))
    L_c = np.zeros((level.nslots, 512, 3))
    nthreads = nthreads or min(8, os.cpu_count() or 1)
    st = ref_simd_lib().ref_level_interactions_simd(
        theta, level.nslots, _abi.iptr(level.neighbors), _abi.iptr(level.kind), _abi.dptr(level.xbase),
        level.dx, int(targets.size), _abi.iptr(targets), type_, _abi.dptr(level.M),
        _abi.dptr(level.com), _abi.dptr(level.mon), _abi.dptr(L), _abi.dptr(L_c), nthreads)
    if st != 0:
        raise RuntimeError(f"reference SIMD level_interactions failed ({st})")
    return L, L_c


def _u8p(a):
    return a.ctypes.data_as(_u8)


def stencils(theta=0.5, use_ref=False):
    """(multipole mask, phase indicator, p2p mask, four constants) over the
    11^3 stencil box."""
    mm = np.zeros(1331, np.uint8)
    mi = np.zeros(1331, np.uint8)
    pm = np.zeros(1331, np.uint8)
    four = np.zeros((1331, 4))
    f = ref_lib().ref_build_stencils if use_ref else lib().oracle_build_stencils
    st = f(theta, _u8p(mm), _u8p(mi), _u8p(pm), _abi.dptr(four))
    if st != 0:
        raise ValueError(f"stencil construction failed for theta={theta} (status {st})")
    return mm, mi, pm, four


def multipole_stencil(theta=0.5):
    el = np.zeros((1331, 3), np.int32)
    ph = np.zeros(1331, np.uint8)
    n = ctypes.c_int32(0)
    st = lib().oracle_multipole_stencil(theta, _abi.iptr(el), _u8p(ph), ctypes.byref(n))
    if st != 0:
        raise ValueError(f"stencil construction failed for theta={theta}")
    return el[: n.value].copy(), ph[: n.value].astype(bool)


def d_factors(dX, use_ref=False):
    out = np.zeros(25)
    dX = np.ascontiguousarray(dX, dtype=np.float64)
    (ref_lib().ref_d_factors if use_ref else lib().oracle_d_factors)(_abi.dptr(dX), _abi.dptr(out))
    return out


def pair_m2l(X, Y, mp, mc, L, corr, rho=True, use_ref=False):
    """Accumulates one M2L pair into L[20], corr[3] (in place)."""
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (X, Y, mp, mc)]
    f = ref_lib().ref_pair_m2l if use_ref else lib().oracle_pair_m2l
    f(*[_abi.dptr(a) for a in args], _abi.dptr(L), _abi.dptr(corr), 1 if rho else 0)


def pair_p2m(X, Y, mp, L, corr, rho=True, use_ref=False):
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (X, Y, mp)]
    f = ref_lib().ref_pair_p2m if use_ref else lib().oracle_pair_p2m
    f(*[_abi.dptr(a) for a in args], _abi.dptr(L), _abi.dptr(corr), 1 if rho else 0)


def multipole_interactions(subgrid, theta=0.5, L=None, L_c=None, use_ref=False):
    """Runs the reference-semantics multipole kernel on one sub-grid
    descriptor; returns (L [512,20], L_c [512,3])."""
    L = np.zeros((512, 20)) if L is None else L
    L_c = np.zeros((512, 3)) if L_c is None else L_c
    f = ref_lib().ref_multipole_interactions if use_ref else lib().oracle_multipole_interactions
    st = f(theta, ctypes.byref(subgrid), _abi.dptr(L), _abi.dptr(L_c))
    if st != 0:
        raise RuntimeError(f"oracle multipole_interactions failed ({st})")
    return L, L_c


def monopole_interactions(subgrid, contains_multipole_neighbor, theta=0.5, L=None, L_c=None, use_ref=False):
    L = np.zeros((512, 20)) if L is None else L
    L_c = np.zeros((512, 3)) if L_c is None else L_c
    f = ref_lib().ref_monopole_interactions if use_ref else lib().oracle_monopole_interactions
    st = f(theta, ctypes.byref(subgrid), 1 if contains_multipole_neighbor else 0, _abi.dptr(L), _abi.dptr(L_c))
    if st != 0:
        raise RuntimeError(f"oracle monopole_interactions failed ({st})")
    return L, L_c


def level_interactions(level, theta=0.5, type_=_abi.RHO, targets=None, nthreads=None, use_ref=False):
    """Oracle (or, with use_ref, the reference-template harness) over every
    target slot of a workloads.Level; returns (L, L_c) arrays shaped like
    level.M / level.com (zeros outside targets)."""
    targets = level.targets if targets is None else np.ascontiguousarray(targets, dtype=np.int32)
    L = np.zeros((level.nslots, 512, 20))
    L_c = np.zeros((level.nslots, 512, 3))
    nthreads = nthreads or min(8, os.cpu_count() or 1)
    f = ref_lib().ref_level_interactions if use_ref else lib().oracle_level_interactions
    st = f(
        theta, level.nslots, _abi.iptr(level.neighbors), _abi.iptr(level.kind), _abi.dptr(level.xbase),
        level.dx, int(targets.size), _abi.iptr(targets), type_, _abi.dptr(level.M),
        _abi.dptr(level.com), _abi.dptr(level.mon), _abi.dptr(L), _abi.dptr(L_c), nthreads)
    if st != 0:
        raise RuntimeError(f"oracle level_interactions failed ({st})")
    return L, L_c


# ---------------------------------------------------------------------------
# Tree passes and the whole adaptive-tree solve
# ---------------------------------------------------------------------------
def _tree_lib():
    L = lib()
    if not getattr(L, "_tree_ready", False):
        L.oracle_multipole_shift.argtypes = [_dp, _dp, _dp]
        L.oracle_local_shift.argtypes = [_dp, _dp, _dp]
        L.oracle_m2m_level.argtypes = [ctypes.c_int32, _ip, _ip, _ip, _dp, ctypes.c_double, ctypes.c_int32,
                                       _dp, _dp, _dp, _dp, _dp]
        L.oracle_l2l_level.argtypes = [ctypes.c_int32, _ip, _ip, _ip, ctypes.c_int32, _dp, _dp, _dp, _dp,
                                       _dp, _dp, ctypes.c_double, ctypes.c_int32, _dp]
        L._tree_ready = True
    return L


def multipole_shift(A, X):
    """taylor::operator>>= : the multipole A moved by X (old - new centre)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    B = np.zeros(20)
    _tree_lib().oracle_multipole_shift(_abi.dptr(A), _abi.dptr(X), _abi.dptr(B))
    return B


def local_shift(A, X):
    """taylor::operator<<= : the local expansion A re-expanded at +X."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    B = np.zeros(20)
    _tree_lib().oracle_local_shift(_abi.dptr(A), _abi.dptr(X), _abi.dptr(B))
    return B


class _LevelView:
    """What level_interactions reads, for one tree level."""

    def __init__(self, lv, M, com, mon):
        self.nslots = lv.nnodes
        self.neighbors = np.ascontiguousarray(lv.neighbors, dtype=np.int32)
        self.kind = np.ascontiguousarray(lv.kind, dtype=np.int32)
        self.xbase = np.ascontiguousarray(lv.xbase, dtype=np.float64)
        self.dx = lv.dx
        self.targets = np.arange(lv.nnodes, dtype=np.int32)
        self.M, self.com, self.mon = M, com, mon


def tree_solve(tree, theta=0.5, type_=_abi.RHO, G=1.0, correct_am=True, nthreads=None, use_ref=False, coms=None):
    """The full gravity solve over an octree.Tree, as node_server::compute_fmm
    runs it level by level (src/node_server.cpp:589-798): M2M up the tree,
    interactions on every level (the root with the root stencil), then L2L
    down the tree and the leaf force assembly.  Returns per-level dicts with
    M, com, L, L_c and g ([n, 512, 4] = phi, gx, gy, gz; leaves only).
    DRHODT reuses the centres of mass of an RHO solve: pass its per-level
    com arrays as coms."""
    T = _tree_lib()
    rho = type_ == _abi.RHO
    out = []
    for lv in tree.levels:
        n = lv.nnodes
        out.append({
            "M": np.zeros((n, 512, 20)), "com": np.zeros((n, 512, 3)),
            "mon": np.ascontiguousarray(lv.mon, dtype=np.float64),
            "L": np.zeros((n, 512, 20)), "L_c": np.zeros((n, 512, 3)), "g": np.zeros((n, 512, 4)),
        })
    if coms is not None:
        for o, c in zip(out, coms):
            o["com"][...] = c
    elif not rho:
        raise ValueError("a DRHODT solve needs the centres of mass of an RHO solve (coms=)")
    ints = [tuple(np.ascontiguousarray(a, dtype=np.int32) for a in (lv.kind, lv.parent, lv.octant))
            for lv in tree.levels]
    # upward pass, finest level first
    for li in range(len(tree.levels) - 1, -1, -1):
        lv, o = tree.levels[li], out[li]
        kind, parent, octant = ints[li]
        up = out[li - 1] if li > 0 else None
        st = T.oracle_m2m_level(lv.nnodes, _abi.iptr(kind), _abi.iptr(parent), _abi.iptr(octant),
                                _abi.dptr(np.ascontiguousarray(lv.xbase)), lv.dx, type_, _abi.dptr(o["M"]),
                                _abi.dptr(o["com"]), _abi.dptr(o["mon"]),
                                _abi.dptr(up["M"]) if up else None, _abi.dptr(up["com"]) if up else None)
        if st != 0:
            raise RuntimeError(f"oracle_m2m_level failed on level {li} ({st})")
    # interactions on every level
    for li, lv in enumerate(tree.levels):
        o = out[li]
        if li == 0 and lv.kind[0] == _abi.SLOT_NONLEAF:
            sg = _abi.Subgrid()
            sg.M = _abi.dptr(o["M"][0])
            sg.com = _abi.dptr(o["com"][0])
            sg.mon = _abi.dptr(o["mon"][0])
            sg.dx = lv.dx
            for k in range(3):
                sg.xbase[k] = lv.xbase[0, k]
            sg.type = type_
            sg.use_root_stencil = 1
            multipole_interactions(sg, theta, L=o["L"][0], L_c=o["L_c"][0], use_ref=use_ref)
        else:
            Lv, Lcv = level_interactions(_LevelView(lv, o["M"], o["com"], o["mon"]), theta, type_,
                                         nthreads=nthreads, use_ref=use_ref)
            o["L"] += Lv
            if rho:
                o["L_c"][...] = Lcv
    # downward pass, root first
    for li, lv in enumerate(tree.levels):
        o = out[li]
        kind, parent, octant = ints[li]
        up = out[li - 1] if li > 0 else None
        st = T.oracle_l2l_level(lv.nnodes, _abi.iptr(kind), _abi.iptr(parent), _abi.iptr(octant), type_,
                                _abi.dptr(o["com"]), _abi.dptr(up["L"]) if up else None,
                                _abi.dptr(up["L_c"]) if up else None, _abi.dptr(up["com"]) if up else None,
                                _abi.dptr(o["L"]), _abi.dptr(o["L_c"]), G, 1 if correct_am else 0,
                                _abi.dptr(o["g"]))
        if st != 0:
            raise RuntimeError(f"oracle_l2l_level failed on level {li} ({st})")
    return out
