"""Learned cost model of the autotuner (include/symcost.h): an FP32 MLP, features -> 64 -> 64
-> 1, predicting log(launch time).  Training and inference run as HIP kernels on the GPU."""
import ctypes

import numpy as np

from . import _native

NUM_FEATURES = 24
HIDDEN = 64


class CostModel:
    def __init__(self, handle):
        self._h = handle

    @classmethod
    def create(cls, num_features=NUM_FEATURES, hidden=HIDDEN, seed=0):
        h = ctypes.c_void_p()
        _native.check(_native.lib().symcost_create(ctypes.byref(h), num_features, hidden, seed), "symcost_create")
        return cls(h)

    @classmethod
    def load(cls, path):
        h = ctypes.c_void_p()
        _native.check(_native.lib().symcost_load(ctypes.byref(h), str(path).encode()), "symcost_load")
        return cls(h)

    def save(self, path):
        _native.check(_native.lib().symcost_save(self._h, str(path).encode()), "symcost_save")

    def train(self, X, y, epochs=200, batch=256, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = ctypes.c_float()
        _native.check(_native.lib().symcost_train(self._h, X.ctypes.data, y.ctypes.data, len(y), epochs, batch,
                                                   lr, seed, ctypes.byref(loss)), "symcost_train")
        return loss.value

    def step(self, X, y, lr):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        _native.check(_native.lib().symcost_step(self._h, X.ctypes.data, y.ctypes.data, len(y), lr), "symcost_step")

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.empty(len(X), np.float32)
        _native.check(_native.lib().symcost_predict(self._h, X.ctypes.data, len(X), out.ctypes.data),
                      "symcost_predict")
        return out

    def params(self):
        n = ctypes.c_size_t()
        _native.check(_native.lib().symcost_num_params(self._h, ctypes.byref(n)), "symcost_num_params")
        out = np.empty(n.value, np.float32)
        _native.check(_native.lib().symcost_get_params(self._h, out.ctypes.data, out.size), "symcost_get_params")
        return out

    def gradient(self):
        g = np.empty(self.params().size, np.float32)
        _native.check(_native.lib().symcost_get_gradient(self._h, g.ctypes.data, g.size), "symcost_get_gradient")
        return g

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _native.check(_native.lib().symcost_set_params(self._h, p.ctypes.data, p.size), "symcost_set_params")

    def scaling(self, num_features=NUM_FEATURES):
        mean = np.empty(num_features, np.float32)
        dev = np.empty(num_features, np.float32)
        y = np.empty(2, np.float32)
        _native.check(_native.lib().symcost_get_scaling(self._h, mean.ctypes.data, dev.ctypes.data, y.ctypes.data),
                      "symcost_get_scaling")
        return mean, dev, y

    def set_scaling(self, mean, dev, y_mean_dev):
        mean = np.ascontiguousarray(mean, np.float32)
        dev = np.ascontiguousarray(dev, np.float32)
        y = np.ascontiguousarray(y_mean_dev, np.float32)
        _native.check(_native.lib().symcost_set_scaling(self._h, mean.ctypes.data, dev.ctypes.data, y.ctypes.data),
                      "symcost_set_scaling")

    def close(self):
        if self._h:
            _native.lib().symcost_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def unpack(params, num_features=NUM_FEATURES, hidden=HIDDEN):
    """Parameter vector -> (W1, b1, W2, b2, w3, b3) as numpy arrays (layout of symcost.h)."""
    F, H = num_features, hidden
    o = 0
    W1 = params[o:o + H * F].reshape(H, F); o += H * F
    b1 = params[o:o + H]; o += H
    W2 = params[o:o + H * H].reshape(H, H); o += H * H
    b2 = params[o:o + H]; o += H
    w3 = params[o:o + H]; o += H
    b3 = params[o]
    return W1, b1, W2, b2, w3, b3
