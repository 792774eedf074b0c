"""ctypes front end of the CPU FP64 oracle (oracle/sph_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB overrides the library path (another build of the same sources,
# e.g. one with a planted mistake that the pins must catch)
_LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "_build", "libsph_oracle.so")
_lib = None

PH_NEIGHBOURS, PH_STRAIN, PH_NORMALS, PH_CONSTITUTIVE, PH_TIMESTEP, PH_MOMENTUM, \
    PH_BOUNDCOND, PH_VELOCITY, PH_MOVE = range(9)
NPHASES = 9

F_X, F_V, F_A, F_RHO, F_P, F_C, F_MASS, F_H, F_TRD, F_MINDIST, F_NORMAL, F_NNORMAL, \
    F_N, F_KAPPA, F_CRITTS, F_RHOOLD, F_ACTIVE, F_VABS, F_Q, F_SIGMA_DIAG, F_SIGMA_OFF, \
    F_ROD_DIAG, F_ROD_OFF = range(23)
_VEC = {F_X, F_V, F_A, F_NORMAL, F_NNORMAL, F_SIGMA_DIAG, F_SIGMA_OFF, F_ROD_DIAG, F_ROD_OFF}

GLOBAL_NAMES = ["dt", "dtold", "ptime", "critts", "timestep", "hmax",
                "cmm_min0", "cmm_min1", "cmm_min2", "cmm_max0", "cmm_max1", "cmm_max2", "ngp",
                "rbR0", "rbR1", "rbR2", "rbV0", "rbV1", "rbV2", "rbW0", "rbW1", "rbW2",
                "rbA0", "rbA1", "rbA2", "rbAL0", "rbAL1", "rbAL2", "done"]


def build() -> str:
    """Compiles the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_last_error.argtypes = [ctypes.c_void_p]
        L.orc_phase.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_int)]
        L.orc_np.argtypes = [ctypes.c_void_p]
        L.orc_ngp.argtypes = [ctypes.c_void_p]
        L.orc_globals.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.orc_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.orc_get_neighbours.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int]
        L.orc_kernel_w.restype = ctypes.c_double
        L.orc_kernel_w.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.orc_kernel_gradw.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.c_double]
        L.orc_kernel_dwdr.restype = ctypes.c_double
        L.orc_kernel_dwdr.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


class Oracle:
    """One scene advanced by the CPU restatement of MSolution::Solution."""

    def __init__(self, deck):
        L = lib()
        from sph07_amd.deck import CDeck
        assert L.orc_deck_size() == ctypes.sizeof(CDeck), "deck layout mismatch"
        c, self._keep = deck.to_c()
        err = ctypes.create_string_buffer(256)
        self._h = L.orc_create(ctypes.byref(c), err, 256)
        if not self._h:
            raise OracleError(err.value.decode())
        self.maxnbr = deck.maxnbr
        self.ndim = deck.ndim

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    def _check(self, rc):
        if rc != 0:
            raise OracleError(lib().orc_last_error(self._h).decode())

    @property
    def np(self) -> int:
        return lib().orc_np(self._h)

    def phase(self, ph: int):
        self._check(lib().orc_phase(self._h, ph))

    def run(self, nsteps: int) -> int:
        steps = ctypes.c_int(0)
        done = ctypes.c_int(0)
        self._check(lib().orc_run(self._h, nsteps, ctypes.byref(steps), ctypes.byref(done)))
        return steps.value

    def globals(self) -> dict:
        out = np.zeros(len(GLOBAL_NAMES))
        lib().orc_globals(self._h, out.ctypes.data, len(out))
        return dict(zip(GLOBAL_NAMES, out.tolist()))

    def field(self, f: int) -> np.ndarray:
        n = self.np
        out = np.zeros((n, 3) if f in _VEC else n)
        self._check(lib().orc_get_field(self._h, f, out.ctypes.data))
        return out

    def set_field(self, f: int, data):
        """Overwrites a field between phases (velocities only)."""
        L = lib()
        L.orc_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        arr = np.ascontiguousarray(data, np.float64).reshape(self.np, 3)
        self._check(L.orc_set_field(self._h, f, arr.ctypes.data))

    def linked_list(self) -> dict:
        """Grid and linked lists of the last SetupLinkedList + GhostSetup."""
        L = lib()
        L.orc_get_linked_list.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_longlong] + \
            [ctypes.c_void_p] * 3 + [ctypes.c_int]
        gmin, gsize, glim = np.zeros(3), np.zeros(3), np.zeros(3, np.int32)
        L.orc_get_linked_list(self._h, gmin.ctypes.data, gsize.ctypes.data, glim.ctypes.data,
                              None, 0, None, None, None, 0)
        ngp = L.orc_ngp(self._h)
        ncell = int(np.prod(glim.astype(np.int64)))
        grid = np.zeros(ncell, np.int32)
        real_next = np.zeros(self.np, np.int32)
        gnext = np.zeros(max(ngp, 1), np.int32)
        gpar = np.zeros(max(ngp, 1), np.int32)
        rc = L.orc_get_linked_list(self._h, None, None, None, grid.ctypes.data, ncell,
                                   real_next.ctypes.data, gnext.ctypes.data, gpar.ctypes.data,
                                   max(ngp, 1))
        assert rc == 0
        return {"gridmin": gmin, "gridsize": gsize, "gridlim": glim,
                "llgrid": grid.reshape(tuple(int(g) for g in glim)), "real_next": real_next,
                "ghost_next": gnext[:ngp], "ghost_par": gpar[:ngp]}

    def neighbours(self):
        n = self.np
        nn = np.zeros(n, np.int32)
        canon = np.zeros((n, self.maxnbr), np.int64)
        lib().orc_get_neighbours(self._h, nn.ctypes.data, canon.ctypes.data, self.maxnbr)
        return nn, canon


def kernel_w(xi, xj, ndim, havg) -> float:
    xi = np.ascontiguousarray(xi, np.float64)
    xj = np.ascontiguousarray(xj, np.float64)
    return lib().orc_kernel_w(xi.ctypes.data, xj.ctypes.data, ndim, havg)


def kernel_gradw(xi, xj, ndim, havg) -> np.ndarray:
    xi = np.ascontiguousarray(xi, np.float64)
    xj = np.ascontiguousarray(xj, np.float64)
    out = np.zeros(3)
    lib().orc_kernel_gradw(out.ctypes.data, xi.ctypes.data, xj.ctypes.data, ndim, havg)
    return out


def kernel_dwdr(xi, xj, ndim, havg) -> float:
    """dWdr of MKernelBSpline::GradW (static): the radial derivative dW/dr
    (0 at r = 0)."""
    xi = np.ascontiguousarray(xi, np.float64)
    xj = np.ascontiguousarray(xj, np.float64)
    return lib().orc_kernel_dwdr(xi.ctypes.data, xj.ctypes.data, ndim, havg)


def threads() -> int:
    """OpenMP threads of the oracle's per-particle loops (OMP_NUM_THREADS)."""
    return int(lib().orc_threads())
