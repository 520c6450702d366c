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

_lib = None
_ref = None

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
