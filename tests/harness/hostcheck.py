"""ctypes wrapper of libhostcheck.so (the CPU harness of the kernel arithmetic; test-only)."""

import ctypes

import numpy as np

_P = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(_P)


_GRAD = [ctypes.c_int, ctypes.c_int] + [_P] * 8 + [ctypes.c_int, _P, ctypes.c_double, ctypes.c_double, _P]


class Harness:
    def __init__(self, path):
        L = ctypes.CDLL(path)
        self.full = hasattr(L, "hc_run")          # the rsq-model library exports the device variant only
        if self.full:
            L.hc_run.argtypes = [ctypes.c_int, ctypes.c_int] + [_P] * 7 + [ctypes.c_double, ctypes.c_int] + [_P] * 6 + [ctypes.c_long]
            L.hc_run.restype = ctypes.c_long
            L.hc_oneel.argtypes = [ctypes.c_int, ctypes.c_int] + [_P] * 14
            L.hc_grad_two.argtypes = _GRAD
            L.hc_grad_two_variant.argtypes = [ctypes.c_int] + _GRAD
            L.hc_grad_two_dev_float.argtypes = _GRAD
        L.hc_grad_two_dev_double.argtypes = _GRAD
        self.L = L

    @staticmethod
    def _arrays(b):
        return [np.ascontiguousarray(a) for a in (b.shell_atom, b.shell_l, b.shell_prim, b.shell_block,
                                                   b.prim_exp, b.prim_coef, np.asarray(b.positions, float))]

    def jk(self, b, D, thr=1e-12):
        """(J, K, Q_shell, Q_block, quartets) through the kernel code (D: (nset, n, n))."""
        at, l, pr, blk, ex, co, pos = self._arrays(b)
        D = np.ascontiguousarray(D, dtype=float)
        J, K = np.zeros_like(D), np.zeros_like(D)
        npair = b.nshell * (b.nshell + 1) // 2
        nbp = b.nblock * (b.nblock + 1) // 2
        qs, qb = np.zeros(npair), np.zeros(nbp)
        cap = npair * (npair + 1) // 2
        qu = np.zeros((cap, 2), np.int64)
        n = self.L.hc_run(b.nshell, b.natm, _p(at), _p(l), _p(pr), _p(blk), _p(ex), _p(co), _p(pos), float(thr),
                          D.shape[0], _p(D), _p(J), _p(K), _p(qs), _p(qb), _p(qu), cap)
        return J, K, qs, qb, qu[:n]

    def one_electron(self, b, D, W):
        at, l, pr, blk, ex, co, pos = self._arrays(b)
        n = b.nao
        S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        M = np.zeros((3, n, n))
        g = np.zeros((b.natm, 3))
        Z = b.nuclear_charges.astype(float)
        self.L.hc_oneel(b.nshell, b.natm, _p(at), _p(l), _p(pr), _p(ex), _p(co), _p(pos), _p(Z),
                        _p(np.ascontiguousarray(D, float)), _p(np.ascontiguousarray(W, float)),
                        _p(S), _p(T), _p(V), _p(M), _p(g))
        return S, T, V, M, g

    def grad_two(self, b, Dc, Dx, kfac, thr=1e-14):
        at, l, pr, blk, ex, co, pos = self._arrays(b)
        Dx = np.ascontiguousarray(np.asarray(Dx, float).reshape(-1, b.nao, b.nao))
        g = np.zeros((b.natm, 3))
        self.L.hc_grad_two(b.nshell, b.natm, _p(at), _p(l), _p(pr), _p(blk), _p(ex), _p(co), _p(pos),
                           _p(np.ascontiguousarray(Dc, float)), Dx.shape[0], _p(Dx), float(kfac), float(thr), _p(g))
        return g

    def grad_two_variant(self, variant, b, Dc, Dx, kfac, thr=1e-14):
        """variant 1: the device kernel's work items in FP64; 2: the same in FP32."""
        at, l, pr, blk, ex, co, pos = self._arrays(b)
        Dx = np.ascontiguousarray(np.asarray(Dx, float).reshape(-1, b.nao, b.nao))
        g = np.zeros((b.natm, 3))
        fn = {1: self.L.hc_grad_two_dev_double, 2: getattr(self.L, "hc_grad_two_dev_float", None)}[int(variant)]
        fn(b.nshell, b.natm, _p(at), _p(l), _p(pr), _p(blk), _p(ex), _p(co), _p(pos),
           _p(np.ascontiguousarray(Dc, float)), Dx.shape[0], _p(Dx), float(kfac), float(thr), _p(g))
        return g
