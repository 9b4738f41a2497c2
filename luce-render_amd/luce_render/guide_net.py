"""The learned guide on its own (the C ABI's lr_guide_net_*): a small MLP on
the matrix cores that maps a surface point and its normal to a mixture of
four von Mises-Fisher lobes over incident directions, trained by Adam on
-weight * log q(direction). The renderer trains and uses it internally when
guiding is 2; this module exposes it for tests and standalone use.

    net = GuideNet(device, lo, hi)
    net.step(samples, rate=5e-3)     # samples: (n, 10) position, normal, direction, weight
    mix = net.mixtures(points)       # points: (n, 6) -> (n, 20)
"""
import ctypes

import numpy as np

from . import _lib


class GuideNet:
    def __init__(self, device, lo, hi, seed=1):
        self.device = device
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().lr_guide_net_create(device.handle, _lib.ptr(lo), _lib.ptr(hi), seed, ctypes.byref(h)))
        self.handle = h
        self.count = int(_lib.lib().lr_guide_net_param_count())

    def params(self):
        out = np.empty(self.count, np.float32)
        _lib.check(_lib.lib().lr_guide_net_get_params(self.handle, _lib.ptr(out), self.count))
        return out

    def set_params(self, params):
        params = np.ascontiguousarray(params, np.float32)
        _lib.check(_lib.lib().lr_guide_net_set_params(self.handle, _lib.ptr(params), params.size))

    def gradient(self, samples):
        samples = np.ascontiguousarray(samples, np.float32)
        grad = np.empty(self.count, np.float32)
        loss = ctypes.c_float()
        _lib.check(_lib.lib().lr_guide_net_gradient(self.handle, _lib.ptr(samples), len(samples), _lib.ptr(grad),
                                                    self.count, ctypes.byref(loss)))
        return grad, loss.value

    def step(self, samples, rate=5e-3):
        samples = np.ascontiguousarray(samples, np.float32)
        loss = ctypes.c_float()
        _lib.check(_lib.lib().lr_guide_net_step(self.handle, _lib.ptr(samples), len(samples), rate,
                                                ctypes.byref(loss)))
        return loss.value

    def mixtures(self, points):
        points = np.ascontiguousarray(points, np.float32)
        out = np.empty((len(points), 20), np.float32)
        _lib.check(_lib.lib().lr_guide_net_mixtures(self.handle, _lib.ptr(points), len(points), _lib.ptr(out)))
        return out

    def destroy(self):
        if self.handle:
            _lib.lib().lr_guide_net_destroy(self.handle)
            self.handle = None
