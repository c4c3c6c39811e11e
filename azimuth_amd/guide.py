"""Python handle on the learned path guide (include/azimuth_guide.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native as N

LAYERS = [(128, 64), (64, 64), (64, 64), (64, 64), (64, 40)]


def pack_params(Ws, bs):
    """Per layer W [in][out] row-major then b [out] -> one float32 vector."""
    return np.concatenate([np.concatenate([np.asarray(W, np.float32).ravel(), np.asarray(b, np.float32)])
                           for W, b in zip(Ws, bs)])


def unpack_params(p):
    Ws, bs, o = [], [], 0
    for i, n in LAYERS:
        Ws.append(p[o:o + i * n].reshape(i, n))
        o += i * n
        bs.append(p[o:o + n])
        o += n
    return Ws, bs


class Guide:
    def __init__(self, device=0, seed=0):
        self.lib = N.load()
        self.h = C.c_void_p()
        N.check(self.lib.az_guide_create(C.byref(self.h), int(device), int(seed)), "az_guide_create")
        self.count = self.lib.az_guide_param_count()

    def _check(self, st, what):
        if st != N.AZ_OK:
            raise N.AzError(f"{what} failed with {N.STATUS_NAMES.get(st, st)}: "
                            f"{self.lib.az_guide_last_error(self.h).decode(errors='replace')}")

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._check(self.lib.az_guide_set_params(self.h, p.ctypes.data, p.size), "az_guide_set_params")

    def params(self):
        p = np.zeros(self.count, np.float32)
        self._check(self.lib.az_guide_get_params(self.h, p.ctypes.data, self.count), "az_guide_get_params")
        return p

    def grad(self):
        g = np.zeros(self.count, np.float32)
        self._check(self.lib.az_guide_get_grad(self.h, g.ctypes.data, self.count), "az_guide_get_grad")
        return g

    def zero_grad(self):
        self._check(self.lib.az_guide_zero_grad(self.h), "az_guide_zero_grad")

    def infer(self, p01, nrm, rough):
        n = len(p01)
        q = np.zeros((n, 8), np.float32)
        q[:, 0:3], q[:, 3:6], q[:, 6] = p01, nrm, rough
        out = np.zeros((n, 40), np.float32)
        self._check(self.lib.az_guide_infer(self.h, q.ctypes.data, n, out.ctypes.data), "az_guide_infer")
        return out.reshape(n, 8, 5)

    def infer_packed(self, p01, nrm, rough, fill=0):
        """The renderer's packed lobes: (n, 8, 2) uint32 words per query;
        queries of negative roughness keep the fill value."""
        n = len(p01)
        q = np.zeros((n, 8), np.float32)
        q[:, 0:3], q[:, 3:6], q[:, 6] = p01, nrm, rough
        out = np.full((n, 16), fill, np.uint32)
        self._check(self.lib.az_guide_infer_packed(self.h, q.ctypes.data, n, out.ctypes.data), "az_guide_infer_packed")
        return out.reshape(n, 8, 2)

    def train(self, p01, nrm, rough, dirs, wgt, update=True):
        n = len(p01)
        s = np.zeros((n, 12), np.float32)
        s[:, 0:3], s[:, 3:6], s[:, 6], s[:, 7:10], s[:, 10] = p01, nrm, rough, dirs, wgt
        loss = C.c_float()
        self._check(self.lib.az_guide_train(self.h, s.ctypes.data, n, 1 if update else 0, C.byref(loss)),
                    "az_guide_train")
        return loss.value

    def mfma_selftest(self, A, B):
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        D = np.zeros((16, 16), np.float32)
        self._check(self.lib.az_guide_mfma_selftest(self.h, A.ctypes.data, B.ctypes.data, D.ctypes.data),
                    "az_guide_mfma_selftest")
        return D

    def close(self):
        if self.h:
            self.lib.az_guide_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
