"""CPU oracle binding (TEST INFRASTRUCTURE ONLY).

Loads oracle/_build/liboracle.so, the C restatement of the reference's
count semantics (oracle.c).  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this package; the product library
never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
CLI_PATH = os.path.join(_HERE, "_build", "pm_oracle")
MAXK = 7


class OrPattern(C.Structure):
    _fields_ = [("k", C.c_int), ("adj", (C.c_int * MAXK) * MAXK), ("vlabel", C.c_int * MAXK)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build the oracle (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.or_count.restype = C.c_int
        L.or_count.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.POINTER(OrPattern), C.c_int, C.c_uint64, C.c_uint64, C.c_int,
                               C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.or_plan_info.restype = C.c_int
        L.or_plan_info.argtypes = [C.POINTER(OrPattern), C.c_int, C.POINTER(C.c_longlong),
                                   C.POINTER(C.c_int), C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_int)]
        _lib = L
    return _lib


def parse_g(text: str, labeled: bool = True) -> OrPattern:
    """Parse .g text the way src/pattern.h:67-100 does."""
    p = OrPattern()
    k = 0
    state = 0
    for line in text.splitlines():
        if state == 0 and line.startswith("#"):
            continue
        if state <= 1 and line.startswith("v"):
            _, _, lab = line.split()[:3]
            p.vlabel[k] = int(lab) if labeled else 1
            k += 1
            state = 1
            continue
        state = 2
        f = line.split()
        if len(f) < 4:
            continue
        a, b, lab = int(f[1]), int(f[2]), int(f[3])
        p.adj[a][b] = lab
        p.adj[b][a] = lab
    p.k = k
    return p


def plan_info(p: OrPattern, labeled: bool):
    aut = C.c_longlong()
    deg = (C.c_int * MAXK)()
    A, B, X = (C.c_uint32 * MAXK)(), (C.c_uint32 * MAXK)(), (C.c_uint32 * MAXK)()
    rc = C.c_int()
    if lib().or_plan_info(C.byref(p), 1 if labeled else 0, C.byref(aut), deg, A, B, X,
                          C.byref(rc)):
        raise ValueError("pattern violates the matcher's preconditions")
    k = p.k
    return dict(aut=aut.value, deg=list(deg)[:k], A=list(A)[:k], B=list(B)[:k], X=list(X)[:k],
                rc_order=rc.value)


def count(rowptr, colidx, labels, p: OrPattern, labeled: bool = True, edge_lo: int = 0,
          edge_hi: int = 0, threads: int = 0, per_edge: bool = False, induced: bool = False):
    """Total count (x|Aut| when unlabeled), job count, optional per-edge counts.
    induced=True counts vertex-induced matches (pattern non-edges map to
    non-adjacent vertices)."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.uint64)
    colidx = np.ascontiguousarray(colidx, dtype=np.uint32)
    lab = None
    if labeled:
        lab = np.ascontiguousarray(labels, dtype=np.uint32)
    n = rowptr.size - 1
    m = colidx.size
    pe = np.zeros(m, dtype=np.uint64) if per_edge else None
    out, jobs = C.c_uint64(), C.c_uint64()
    rc = lib().or_count(n, m, rowptr.ctypes.data, colidx.ctypes.data if m else None,
                        None if lab is None else lab.ctypes.data, C.byref(p),
                        (1 if labeled else 0) | (2 if induced else 0), edge_lo, edge_hi, threads,
                        None if pe is None else pe.ctypes.data, C.byref(out), C.byref(jobs))
    if rc:
        raise RuntimeError(f"oracle failed ({rc})")
    return int(out.value), int(jobs.value), pe
