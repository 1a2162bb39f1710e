"""ctypes wrapper of the C/OpenMP restatement (oracle/c/jxf_cpu.c).

TEST INFRASTRUCTURE ONLY: used by tests and by the CPU-baseline leg of bench.py.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "c", "libjxf_cpu.so")

VAR = {"PRIMITIVE": 0, "CONSERVATIVE": 1, "CHAR-PRIMITIVE": 2}
STENCIL = {"WENO5-JS": 0, "WENO5-Z": 1}
RIEMANN = {"HLLC": 0, "RUSANOV": 1}
SS = {"ARITHMETIC": 0, "RUSANOV": 1, "DAVIS": 2, "EINFELDT": 3, "TORO": 4}
BC = {"PERIODIC": 0, "ZEROGRADIENT": 1, "SYMMETRY": 2}
FACES = ("east", "west", "north", "south", "top", "bottom")


class CGrid(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("nh", ctypes.c_int), ("dx", ctypes.c_double * 3)]


class CScheme(ctypes.Structure):
    _fields_ = [("var", ctypes.c_int), ("stencil", ctypes.c_int), ("riemann", ctypes.c_int), ("ss", ctypes.c_int),
                ("gamma", ctypes.c_double), ("eps", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from oracle.c import build as b  # noqa
        _lib = ctypes.CDLL(LIB)
        P = ctypes.POINTER(ctypes.c_double)
        _lib.jxf_cpu_rk3_step.argtypes = [P, P, P, P, P, ctypes.POINTER(CGrid), ctypes.POINTER(CScheme),
                                          ctypes.POINTER(ctypes.c_int), ctypes.c_double]
        _lib.jxf_cpu_compute_dt.argtypes = [P, ctypes.POINTER(CGrid), ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double]
        _lib.jxf_cpu_compute_dt.restype = ctypes.c_double
        _lib.jxf_cpu_compute_rhs.argtypes = [P, P, ctypes.POINTER(CGrid), ctypes.POINTER(CScheme), P, P]
        _lib.jxf_cpu_max_threads.restype = ctypes.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class CPUStepper:
    """Holds scratch for repeated steps of one domain."""

    def __init__(self, dom, s):
        self.g = CGrid()
        for i in range(3):
            self.g.n[i] = dom.cells[i]
            self.g.dx[i] = dom.dx[i]
        self.g.nh = dom.nh
        self.sc = CScheme(VAR[s.variable], STENCIL[s.stencil], RIEMANN[s.riemann], SS[s.signal_speed or "EINFELDT"],
                          s.gamma, s.eps_stencil)
        self.bc = (ctypes.c_int * 6)(*[BC[s.bcs.get(f, "PERIODIC")] for f in FACES])
        self.s = s
        shape = (5,) + dom.shape
        ncell = int(np.prod(dom.cells))
        nf = max(int(np.prod([c + (i == a) for i, c in enumerate(dom.cells)])) for a in dom.active)
        self.Un = np.empty(shape)
        self.rhs = np.empty(5 * ncell)
        self.flux = np.empty(5 * nf)

    def step(self, W, U, dt):
        lib().jxf_cpu_rk3_step(_p(W), _p(U), _p(self.Un), _p(self.rhs), _p(self.flux), ctypes.byref(self.g),
                               ctypes.byref(self.sc), self.bc, float(dt))

    def dt(self, W, eps):
        return lib().jxf_cpu_compute_dt(_p(W), ctypes.byref(self.g), self.s.gamma, self.s.cfl, eps)


def max_threads():
    return lib().jxf_cpu_max_threads()
