"""Learned pair-cost model of the batch scheduler (csrc/costmodel.hip).

A 4 -> 64 -> 64 -> 1 ReLU MLP with BF16 weights and FP32 accumulation predicts
each pair's WFA score from cheap features (log lengths, log length difference,
fraction of 16 sampled pattern 12-mers found near their expected text
position).  Every computation -- features, forward, backward, Adam -- runs in
the library's HIP kernels; this module only moves host arrays.

The prediction feeds ``DeviceBatch.set_schedule(SCHEDULE_LEARNED, score)``: it
decides each pair's starting workspace tier and its place in the queue, never a
result.  The labels come from the aligner itself (score-only runs over a
generated training set).
"""
import ctypes

import numpy as np

from . import (BiwfaError, SCHEDULE_LEARNED, WavefrontAligner, _ptr, generate_dataset, lib)
from .configs import CONFIGS, PENALTIES

N_FEATURES = 4


def _L():
    L = lib()
    if not getattr(L, "_cm_ready", False):
        L.bw_mlp_new.restype = ctypes.c_void_p
        L.bw_mlp_new.argtypes = [ctypes.c_int, ctypes.c_uint64]
        L.bw_mlp_free.argtypes = [ctypes.c_void_p]
        L.bw_mlp_last_error.restype = ctypes.c_char_p
        L.bw_mlp_get_params.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.bw_mlp_set_params.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.bw_mlp_predict.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.bw_mlp_gradient.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_void_p]
        L.bw_mlp_train.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64,
                                   ctypes.POINTER(ctypes.c_float)]
        L.bw_batch_features.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L._cm_ready = True
    return L


def _err(what, st):
    raise BiwfaError("%s failed (%d): %s" % (what, st, _L().bw_mlp_last_error().decode()))


def score_to_target(score):
    """Regression target of a WFA score: log2(1 + score) / 16."""
    return (np.log2(1.0 + np.maximum(np.asarray(score, np.float64), 0.0)) / 16.0).astype(np.float32)


def target_to_score(y):
    return np.exp2(16.0 * np.asarray(y, np.float64)) - 1.0


def batch_features(packed, device=0):
    """X[n, 4] of a PackedPairs batch, computed on the device."""
    n = len(packed)
    X = np.zeros((n, N_FEATURES), np.float32)
    if n == 0:
        return X
    b = packed._struct()
    st = _L().bw_batch_features(int(device), ctypes.byref(b), _ptr(X))
    if st != 0:
        _err("bw_batch_features", st)
    return X


class CostModel:
    """The device MLP (bw_mlp_t)."""

    def __init__(self, device=0, seed=0):
        self._L = _L()
        self.n_params = self._L.bw_mlp_num_params()
        self._h = self._L.bw_mlp_new(int(device), int(seed))
        if not self._h:
            raise BiwfaError("bw_mlp_new: " + self._L.bw_mlp_last_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.bw_mlp_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        """(FP32 master parameters, their BF16-rounded copy used by the kernels)."""
        m = np.zeros(self.n_params, np.float32)
        w = np.zeros(self.n_params, np.float32)
        st = self._L.bw_mlp_get_params(self._h, _ptr(m), _ptr(w))
        if st != 0:
            _err("bw_mlp_get_params", st)
        return m, w

    def set_params(self, master):
        m = np.ascontiguousarray(master, np.float32)
        if m.shape != (self.n_params,):
            raise BiwfaError("expected %d parameters" % self.n_params)
        st = self._L.bw_mlp_set_params(self._h, _ptr(m))
        if st != 0:
            _err("bw_mlp_set_params", st)

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        y = np.zeros(len(X), np.float32)
        if len(X):
            st = self._L.bw_mlp_predict(self._h, _ptr(X), len(X), _ptr(y))
            if st != 0:
                _err("bw_mlp_predict", st)
        return y

    def predict_score(self, X):
        return target_to_score(self.predict(X))

    def gradient(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(self.n_params, np.float32)
        st = self._L.bw_mlp_gradient(self._h, _ptr(X), _ptr(y), len(X), _ptr(g))
        if st != 0:
            _err("bw_mlp_gradient", st)
        return g

    def train(self, X, y, epochs=30, batch=256, lr=3e-3, seed=0):
        """Mini-batch Adam on the device; returns the full-set MSE afterwards."""
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = ctypes.c_float()
        st = self._L.bw_mlp_train(self._h, _ptr(X), _ptr(y), len(X), int(epochs), int(batch), float(lr),
                                  int(seed), ctypes.byref(loss))
        if st != 0:
            _err("bw_mlp_train", st)
        return float(loss.value)


def aligner_labels(packed, device=0, penalties=PENALTIES):
    """Exact WFA scores of a batch from the aligner itself (score-only BiWFA)."""
    with WavefrontAligner(penalties=penalties, score_only=True, device=device) as a:
        res = a.align_batch(packed)
    if np.any(res.status != 0):
        raise BiwfaError("score-only labelling failed for %d pairs" % int(np.sum(res.status != 0)))
    return res.wf_score.astype(np.int64)


def train_cost_model(cfg=None, num_pairs=None, device=0, seed=0, epochs=30, batch=256, lr=3e-3):
    """Generates the training set (default: the 50,000-pair mixed-length set), labels it
    with the aligner, computes features on the device and trains a fresh model.
    Returns (model, info)."""
    cfg = cfg or CONFIGS["C4_TRAIN"]
    ds = generate_dataset(cfg, num_pairs=num_pairs)
    packed = ds.pack()
    score = aligner_labels(packed, device=device)
    X = batch_features(packed, device=device)
    y = score_to_target(score)
    model = CostModel(device=device, seed=seed)
    loss0 = float(np.mean((model.predict(X) - y) ** 2))
    loss = model.train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return model, {"pairs": len(ds), "initial_mse": loss0, "final_mse": loss,
                   "X": X, "y": y, "score": score}


def learned_schedule(dbatch, model, packed, device=0):
    """Sets the learned schedule of a DeviceBatch from the model's predictions."""
    pred = model.predict_score(batch_features(packed, device=device)).astype(np.float32)
    dbatch.set_schedule(SCHEDULE_LEARNED, pred)
    return pred
