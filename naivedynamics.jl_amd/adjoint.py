"""Differentiable simulation: hand-written adjoint velocity Verlet on the GPU
(nd_adj_* in libnaivehip.so).

The forward model is the simulate! step (Simulator.jl:198-225) with textbook
shifted-force Lennard-Jones and Coulomb forces, unit masses and no walls;
the loss is the squared distance between a smoothed radial distribution of
the final frame and a target.  backward() returns dL/dx0, dL/dv0, dL/dq,
dL/deps and dL/dsigma from a reverse sweep that recomputes each checkpoint
segment.  This is the role a ChainRulesCore rrule plays for the Julia
entry point: forward returns the loss, backward the pullback of a unit
cotangent.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import PREC_FP32, PREC_FP64, AdjParams, check, load

PRECISIONS = {"fp32": PREC_FP32, "fp64": PREC_FP64}


class Adjoint:
    def __init__(self, n, nsteps, dt=0.001, eps=1.0, sigma=1.0, rc=2.5, k_coul=1.0, rc_coul=2.5,
                 skin=0.3, precision="fp64", checkpoint_every=50, rebuild_every=10, nbins=50,
                 rdf_rmax=2.5, rdf_h=0.1, volume=1.0, ctx=None):
        self.ctx = ctx if ctx is not None else _lib.default_context()
        p = AdjParams()
        p.n, p.nsteps, p.checkpoint_every, p.rebuild_every, p.nbins = (
            int(n), int(nsteps), int(checkpoint_every), int(rebuild_every), int(nbins))
        p.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        p.eps, p.sigma, p.rc, p.k_coul, p.rc_coul = eps, sigma, rc, k_coul, rc_coul
        p.skin, p.dt, p.rdf_rmax, p.rdf_h, p.volume = skin, dt, rdf_rmax, rdf_h, volume
        self.params = p
        self.n = int(n)
        self.handle = None
        h = ctypes.c_void_p()
        check(load().nd_adj_create(self.ctx.handle, ctypes.byref(p), ctypes.byref(h)), self.ctx,
              "nd_adj_create")
        self.handle = h
        self.ctx.register(self)

    def close(self):
        if self.handle:
            if self.ctx.handle:
                load().nd_adj_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_potential(self, eps, sigma):
        check(load().nd_adj_set_potential(self.handle, float(eps), float(sigma)), self.ctx,
              "nd_adj_set_potential")

    def forward(self, x0, v0, q=None, target=None):
        """Returns (loss, rdf); loss is 0 without a target."""
        x0 = np.ascontiguousarray(x0, np.float64).reshape(self.n, 3)
        v0 = np.ascontiguousarray(v0, np.float64).reshape(self.n, 3)
        qq = None if q is None else np.ascontiguousarray(q, np.float64).reshape(self.n)
        tg = None if target is None else np.ascontiguousarray(target, np.float64)
        g = np.zeros(self.params.nbins)
        L = ctypes.c_double(0)
        P = _lib.ptr
        check(load().nd_adj_forward(self.handle, P(x0), P(v0), P(qq), P(tg), P(g), ctypes.byref(L)),
              self.ctx, "nd_adj_forward")
        return L.value, g

    def final_state(self):
        x = np.empty((self.n, 3))
        v = np.empty((self.n, 3))
        check(load().nd_adj_final_state(self.handle, _lib.ptr(x), _lib.ptr(v)), self.ctx,
              "nd_adj_final_state")
        return x, v

    def backward(self):
        """Gradients of the last forward run's loss."""
        dx = np.empty((self.n, 3))
        dv = np.empty((self.n, 3))
        dq = np.empty(self.n)
        dp = np.empty(2)
        P = _lib.ptr
        check(load().nd_adj_backward(self.handle, P(dx), P(dv), P(dq), P(dp)), self.ctx,
              "nd_adj_backward")
        return dict(dx0=dx, dv0=dv, dq=dq, deps=float(dp[0]), dsigma=float(dp[1]))


def rrule_loss(adj, x0, v0, q, target):
    """(loss, pullback): pullback(dL) gives the cotangents of (x0, v0, q, eps, sigma)."""
    loss, _ = adj.forward(x0, v0, q, target)
    g = adj.backward()

    def pullback(dl=1.0):
        return (dl * g["dx0"], dl * g["dv0"], dl * g["dq"], dl * g["deps"], dl * g["dsigma"])

    return loss, pullback


__all__ = ["Adjoint", "rrule_loss"]
