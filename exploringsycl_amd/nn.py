"""Python side of the learned preconditioner's network (libteahip.so)."""

import ctypes

import numpy as np

from . import native
from .native import check

# (out, in) channels of the six 3x3 layers, in parameter order; each layer's
# weights [out][in][3][3] are followed by its biases [out].
LAYER_SHAPES = [(32, 3)] + [(32, 32)] * 4 + [(1, 32)]

_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)


_vp, _i, _f, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_size_t
# Argument types of the tea_nn_* functions of include/tea_hip.h.
SIGNATURES = {
    "tea_nn_num_params": [],
    "tea_nn_create": [_i, _i, _i, _i, ctypes.POINTER(_vp)],
    "tea_nn_destroy": [_vp],
    "tea_nn_set_params": [_vp, _F, _sz],
    "tea_nn_get_params": [_vp, _F, _sz],
    "tea_nn_forward": [_vp, _i, _F, _F],
    "tea_nn_loss_grad": [_vp, _i, _F, _F, _F, _F],
    "tea_nn_train_step": [_vp, _i, _F, _F, _f, _F],
    "tea_nn_adam_step": [_vp, _F, _f],
    "tea_nn_get_adam_state": [_vp, _F, _F, _sz, ctypes.POINTER(ctypes.c_longlong)],
    "tea_nn_load_dataset": [_vp, _i, _F, _F],
    "tea_nn_train_epoch": [_vp, _I, _i, _i, _f, _F],
}


def _bind(lib):
    if getattr(lib, "_nn_bound", False):
        return lib
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib._nn_bound = True
    return lib


def _fp(a):
    return a.ctypes.data_as(_F)


class Net:
    """The network on the GPU for batches of [batch][height][width] inputs."""

    def __init__(self, max_batch, height, width, device=0):
        self.lib = _bind(native.load())
        self.max_batch, self.height, self.width = max_batch, height, width
        h = ctypes.c_void_p()
        check(self.lib.tea_nn_create(device, max_batch, height, width, ctypes.byref(h)),
              "tea_nn_create")
        self.handle = h
        self.num_params = self.lib.tea_nn_num_params()

    def close(self):
        if self.handle:
            check(self.lib.tea_nn_destroy(self.handle), "tea_nn_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(self.lib.tea_nn_set_params(self.handle, _fp(p), p.size), "set_params")

    def get_params(self):
        p = np.empty(self.num_params, dtype=np.float32)
        check(self.lib.tea_nn_get_params(self.handle, _fp(p), p.size), "get_params")
        return p

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty(x.shape[:3], dtype=np.float32)
        check(self.lib.tea_nn_forward(self.handle, x.shape[0], _fp(x), _fp(out)), "forward")
        return out

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = ctypes.c_float()
        g = np.empty(self.num_params, dtype=np.float32)
        check(self.lib.tea_nn_loss_grad(self.handle, x.shape[0], _fp(x), _fp(t),
                                        ctypes.byref(loss), _fp(g)), "loss_grad")
        return loss.value, g

    def train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = ctypes.c_float()
        check(self.lib.tea_nn_train_step(self.handle, x.shape[0], _fp(x), _fp(t), lr,
                                         ctypes.byref(loss)), "train_step")
        return loss.value

    def adam_step(self, grad, lr):
        """One Adam step on the master parameters with the given gradient."""
        grad = np.ascontiguousarray(grad, dtype=np.float32)
        check(self.lib.tea_nn_adam_step(self.handle, _fp(grad), lr), "adam_step")

    def adam_state(self):
        """(m, v, step): the Adam moments and the number of steps taken."""
        m = np.empty(self.num_params, dtype=np.float32)
        v = np.empty(self.num_params, dtype=np.float32)
        t = ctypes.c_longlong()
        check(self.lib.tea_nn_get_adam_state(self.handle, _fp(m), _fp(v), m.size,
                                             ctypes.byref(t)), "get_adam_state")
        return m, v, t.value

    def load_dataset(self, inputs, targets):
        inputs = np.ascontiguousarray(inputs, dtype=np.float32)
        targets = np.ascontiguousarray(targets, dtype=np.float32)
        check(self.lib.tea_nn_load_dataset(self.handle, inputs.shape[0], _fp(inputs),
                                           _fp(targets)), "load_dataset")

    def train_epoch(self, order, batch, lr):
        order = np.ascontiguousarray(order, dtype=np.int32)
        loss = ctypes.c_float()
        check(self.lib.tea_nn_train_epoch(self.handle,
                                          order.ctypes.data_as(_I), order.size, batch, lr,
                                          ctypes.byref(loss)), "train_epoch")
        return loss.value
