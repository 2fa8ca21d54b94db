"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use this module, to compare against; the product never imports it.
The library is a C++ restatement of the reference example tree code
(example/tree.cpp, example/calculate.cpp) built by ``oracle/Makefile``.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

_VARIANTS = {(0, 4): "liboracle_so0_pf.so", (0, 1): "liboracle_so0_p.so", (0, 3): "liboracle_so0_f.so",
             (1, 3): "liboracle_so1_f.so", (2, 4): "liboracle_so2_pf.so"}
_libs = {}

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int64_p = ctypes.POINTER(ctypes.c_int64)


def build(jobs=8):
    subprocess.run(["make", "-C", HERE, f"-j{jobs}"], check=True,
                   stdout=subprocess.DEVNULL)


def lib(source_order=0, outputsize=4):
    key = (source_order, outputsize)
    if key in _libs:
        return _libs[key]
    if key not in _VARIANTS:
        raise ValueError(f"no oracle variant for source_order={source_order}, outputsize={outputsize}")
    _libs[key] = _load(os.path.join(BUILD, _VARIANTS[key]))
    return _libs[key]


def mutant_lib(k):
    """The charge oracle (potential + field) built with planted tree/list
    mistake k (ORACLE_MUTANT in fmm_oracle.cpp), for tests of the tests."""
    key = ("mutant", k)
    if key not in _libs:
        _libs[key] = _load(os.path.join(BUILD, f"mutant_{k}.so"))
    return _libs[key]


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
    L = ctypes.CDLL(path)
    L.oracle_build_tree.restype = ctypes.c_void_p
    L.oracle_build_tree.argtypes = [c_double_p, c_double_p, ctypes.c_size_t, ctypes.c_size_t,
                                    ctypes.c_size_t, ctypes.c_double]
    for nm in ("oracle_ncells", "oracle_n_m2l", "oracle_n_p2p"):
        getattr(L, nm).restype = ctypes.c_size_t
        getattr(L, nm).argtypes = [ctypes.c_void_p]
    L.oracle_free_tree.argtypes = [ctypes.c_void_p]
    L.oracle_export_cells.argtypes = [ctypes.c_void_p, c_double_p, c_int64_p]
    L.oracle_export_leaf_of.argtypes = [ctypes.c_void_p, c_int64_p]
    L.oracle_export_lists.argtypes = [ctypes.c_void_p, c_int64_p, c_int64_p]
    L.oracle_export_ML.argtypes = [ctypes.c_void_p, c_double_p, c_double_p]
    L.oracle_compute_field_fmm.argtypes = [ctypes.c_void_p, c_double_p]
    L.oracle_compute_field_fmm_phases.argtypes = [ctypes.c_void_p, c_double_p, ctypes.c_int]
    L.oracle_compute_field_bh.argtypes = [ctypes.c_void_p, c_double_p]
    L.oracle_direct_targets.argtypes = [c_double_p, c_double_p, ctypes.c_size_t, c_int64_p,
                                        ctypes.c_size_t, c_double_p]
    for nm in ("oracle_op_P2M",):
        getattr(L, nm).argtypes = [ctypes.c_double] * 4 + [c_double_p, ctypes.c_int]
    for nm in ("oracle_op_M2M", "oracle_op_M2L", "oracle_op_L2L", "oracle_op_L2P", "oracle_op_M2P"):
        getattr(L, nm).argtypes = [ctypes.c_double] * 3 + [c_double_p, c_double_p, ctypes.c_int]
    L.oracle_op_P2P.argtypes = [ctypes.c_double] * 3 + [c_double_p, c_double_p]
    L.oracle_set_runtime_schedule.argtypes = [ctypes.c_int]
    return L


def _p(a):
    return a.ctypes.data_as(c_double_p)


def _pi(a):
    return a.ctypes.data_as(c_int64_p)


class OracleTree:
    """build_tree + Tree::compute_field_* of the example, on the CPU."""

    def __init__(self, pos, S, ncrit, order, theta, source_order=0, outputsize=4, library=None):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        self.n = self.pos.shape[0]
        self.S = np.ascontiguousarray(S, dtype=np.float64).reshape(self.n, -1)
        self.outputsize = outputsize
        self.lib = library if library is not None else lib(source_order, outputsize)
        self.h = self.lib.oracle_build_tree(_p(self.pos), _p(self.S), self.n, ncrit, order, float(theta))
        self.ncrit, self.order, self.theta = ncrit, order, theta

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle_free_tree(self.h)
            self.h = None

    @property
    def ncells(self):
        return self.lib.oracle_ncells(self.h)

    def cells(self):
        nc = self.ncells
        geo = np.zeros((nc, 5))
        meta = np.zeros((nc, 12), dtype=np.int64)
        self.lib.oracle_export_cells(self.h, _p(geo), _pi(meta))
        return geo, meta

    def leaf_of(self):
        out = np.full(self.n, -1, dtype=np.int64)
        self.lib.oracle_export_leaf_of(self.h, _pi(out))
        return out

    def lists(self):
        m2l = np.zeros((self.lib.oracle_n_m2l(self.h), 2), dtype=np.int64)
        p2p = np.zeros((self.lib.oracle_n_p2p(self.h), 2), dtype=np.int64)
        self.lib.oracle_export_lists(self.h, _pi(m2l), _pi(p2p))
        return m2l, p2p

    def compute_field_fmm(self, mask=63):
        F = np.zeros((self.n, self.outputsize))
        self.lib.oracle_compute_field_fmm_phases(self.h, _p(F), int(mask))
        return F

    def compute_field_bh(self):
        F = np.zeros((self.n, self.outputsize))
        self.lib.oracle_compute_field_bh(self.h, _p(F))
        return F

    def ML(self, msize, lsize):
        M = np.zeros((self.ncells, msize))
        L = np.zeros((self.ncells, lsize))
        self.lib.oracle_export_ML(self.h, _p(M), _p(L))
        return M, L


def direct_targets(pos, S, targets, source_order=0, outputsize=4):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    S = np.ascontiguousarray(S, dtype=np.float64).reshape(pos.shape[0], -1)
    t = np.ascontiguousarray(targets, dtype=np.int64)
    F = np.zeros((t.shape[0], outputsize))
    lib(source_order, outputsize).oracle_direct_targets(_p(pos), _p(S), pos.shape[0], _pi(t), t.shape[0], _p(F))
    return F


def canonical_paths(meta):
    """Octant path of every cell ('' for the root), from exported meta."""
    nc = meta.shape[0]
    paths = [None] * nc
    paths[0] = ""
    # parents precede children in creation order
    for c in range(nc):
        for o in range(8):
            ch = meta[c, 4 + o]
            if ch >= 0:
                paths[ch] = paths[c] + str(o)
    return paths
