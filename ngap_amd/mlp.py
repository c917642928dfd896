"""Python mirror of the frontier-size predictor's C ABI (ngap_mlp_*): a
264 -> 64 -> 64 -> 1 ReLU MLP whose forward, backward and Adam kernels are
the HIP kernels in csrc/mlp.hip."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

IN, HID = 264, 64


def split_params(p):
    """Flat parameters -> (W1[264,64], b1, W2[64,64], b2, W3[64], b3)."""
    p = np.asarray(p)
    o = 0
    out = []
    for shape in ((IN, HID), (HID,), (HID, HID), (HID,), (HID,), (1,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    return out


class MLP:
    def __init__(self, device=0, seed=0):
        self._h = C.c_void_p()
        check(_lib.lib().ngap_mlp_create(int(device), int(seed), C.byref(self._h)))
        self.num_params = _lib.lib().ngap_mlp_num_params()

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().ngap_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def params(self):
        out = np.zeros(self.num_params, np.float32)
        check(_lib.lib().ngap_mlp_get_params(self._h, out.ctypes.data))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        check(_lib.lib().ngap_mlp_set_params(self._h, p.ctypes.data))

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(x.shape[0], np.float32)
        check(_lib.lib().ngap_mlp_predict(self._h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def predict_pairs(self, hist, stats, pairs):
        hist = np.ascontiguousarray(hist, np.float32)
        stats = np.ascontiguousarray(stats, np.float32)
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        y = np.zeros(pairs.shape[0], np.float32)
        check(_lib.lib().ngap_mlp_predict_pairs(self._h, hist.ctypes.data, hist.shape[0],
                                                stats.ctypes.data, stats.shape[0],
                                                pairs.ctypes.data, pairs.shape[0], y.ctypes.data))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.num_params, np.float32)
        loss = C.c_float()
        check(_lib.lib().ngap_mlp_gradient(self._h, x.ctypes.data, t.ctypes.data, x.shape[0],
                                           g.ctypes.data, C.byref(loss)))
        return g, loss.value

    def train(self, x, t, epochs=4, batch=4096, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float()
        check(_lib.lib().ngap_mlp_train(self._h, x.ctypes.data, t.ctypes.data, x.shape[0],
                                        int(epochs), int(batch), float(lr), int(seed),
                                        C.byref(loss)))
        return loss.value
