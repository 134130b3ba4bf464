"""Tangent-linear (forward-mode differentiated) form of the solid path and of
the MUSCL(+THINC/BVD) + HLLC fluid path, and the probe-trace calibration loss of the
differentiated plate impact with a torch.autograd.Function over it.

The reference has no differentiated form.  Its time loop (main.cpp:121-191)
is what is differentiated: every value of a tangent run is the FP64 form's,
bit for bit, and each direction carries d(state)/d(parameter) for one
material parameter of the Us-up solid EOS (Gamma0, s) or the perfect-plastic
return (yield stress), through initialisation, the 5-wave fluxes, the update,
plasticUpdate, cleanUpV, cleanUpAlpha and the CFL time step.

With six parameters and one scalar loss, forward mode costs one tangent pass
per parameter group and needs no stored trajectory (no checkpoints): the
derivative state rides along with the state.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import (TJ_TP_GAMMA0, TJ_TP_PREF, TJ_TP_S, TJ_TP_YIELD, GpuLevel, TenjinError, configs)

# "Gamma0" is the Gruneisen coefficient adiabaticIndex - 1 of any material
# (d/dGamma0 = d/d adiabaticIndex); "pref" the stiffened-gas reference pressure
PARAM_CODE = {"Gamma0": TJ_TP_GAMMA0, "gamma": TJ_TP_GAMMA0, "s": TJ_TP_S, "Y": TJ_TP_YIELD, "pref": TJ_TP_PREF}


class TangentRun:
    """A level (solid, or fluid with MUSCL(+THINC) + HLLC) with its tangent-linear
    form.  directions: (material, parameter) pairs, parameter one of
    "Gamma0" (alias "gamma"), "s", "Y", "pref" (or TJ_TP_* codes).
    chunk: directions per pass group (1, 2, 3 or 6; 0 = all in one group up to
    6); every group re-runs the trajectory."""

    def __init__(self, pb, directions, chunk: int = 1, device: int = 0):
        self.pb = pb
        self.g = GpuLevel(pb, device)
        self.lib = self.g.lib
        self.dirs = [(int(m), PARAM_CODE.get(p, p)) for m, p in directions]
        n = len(self.dirs)
        mats = (ctypes.c_int * n)(*[m for m, _ in self.dirs])
        pars = (ctypes.c_int * n)(*[p for _, p in self.dirs])
        self.tg = ctypes.c_void_p()
        self.g._chk(self.lib.tj_tangent_create(self.g.ctx, n, mats, pars, chunk, ctypes.byref(self.tg)),
                    "tj_tangent_create")
        torch = self.g.torch
        self.P0 = torch.from_numpy(np.ascontiguousarray(pb.prims, dtype=np.float64)).to(self.g.S.device)
        self.vP0 = self.g._view(self.P0)
        torch.cuda.synchronize()

    def run(self, nsteps: int, probe=None, final_time: float = float("inf")):
        """nsteps loop bodies from the initial primitives.  probe: (comp, i,
        j, k).  Returns (trace [nsteps, 1 + ndir], steps run): the probe's
        value after each step, then its derivative along each direction."""
        comp, i, j, k = probe if probe is not None else (0, 0, 0, 0)
        trace = np.zeros((nsteps, 1 + len(self.dirs)))
        done = ctypes.c_int()
        tp = trace.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self.g._chk(self.lib.tj_tangent_run(self.tg, ctypes.byref(self.vP0), self.pb.shift_pressure,
                                            ctypes.byref(self.g.vS), ctypes.byref(self.g.vT), final_time, nsteps,
                                            comp, i, j, k, tp, ctypes.byref(done)), "tj_tangent_run")
        return trace, done.value

    def state(self) -> np.ndarray:
        """Values of the final state."""
        return self.g.get_state()

    def state_derivative(self, d: int) -> np.ndarray:
        """d(final state)/d(parameter of direction d) (direction of the last group)."""
        out = self.g.torch.empty_like(self.g.S)
        self.g.torch.cuda.synchronize()  # the allocation is torch's; the copy runs on the library's stream
        v = self.g._view(out)
        self.g._chk(self.lib.tj_tangent_state(self.tg, d, ctypes.byref(v)), "tj_tangent_state")
        return out.cpu().numpy()

    def status(self, clear: int = 0) -> int:
        return self.g.status(clear)

    def close(self):
        if self.tg:
            self.lib.tj_tangent_destroy(self.tg)
            self.tg = ctypes.c_void_p()
        self.g.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


ALL_DIRECTIONS = [(m, p) for m in (0, 1) for p in configs.PLATE_PARAMS]


def forward_trace(pb, nsteps: int, probe, device: int = 0) -> np.ndarray:
    """Probe values after each of nsteps FP64 loop bodies (tj_advance_steps one
    step at a time)."""
    g = GpuLevel(pb, device)
    g.init_from_prims(pb.prims, pb.shift_pressure)
    g.begin_run(float("inf"))
    comp, i, j, k = probe
    out = np.empty(nsteps)
    for q in range(nsteps):
        g.advance(1)
        g.sync()
        out[q] = float(g.S[comp, k, j, i].item())
    g.close()
    return out


def probe_loss(trace_values: np.ndarray, target: np.ndarray) -> float:
    """Loss = sum_n (u_n - u*_n)^2."""
    r = np.asarray(trace_values) - np.asarray(target)
    return float(np.dot(r, r))


def loss_and_grad(theta, target: np.ndarray, n: int = 512, steps: int = 200, chunk: int = 1, device: int = 0,
                  seed: int = 4):
    """Loss of the differentiated plate impact with parameters theta
    ([Gamma0, s, Y] of copper then aluminium) against a target trace, and its
    gradient dL/dtheta = sum_n 2 (u_n - u*_n) du_n/dtheta."""
    pb = configs.plate_impact_differentiated(n=n, seed=seed, theta=theta, steps=steps)
    tr = TangentRun(pb, ALL_DIRECTIONS, chunk=chunk, device=device)
    try:
        trace, done = tr.run(steps, configs.plate_probe(pb))
        if done != steps:
            raise TenjinError(f"tangent run stopped after {done} of {steps} steps (status {tr.status()})")
    finally:
        tr.close()
    r = trace[:, 0] - np.asarray(target)
    return float(np.dot(r, r)), 2.0 * (r @ trace[:, 1:])


def _autograd_function():
    import torch

    class ProbeLoss(torch.autograd.Function):
        """torch.autograd over the tangent-linear form: forward evaluates the
        loss and its gradient in one tangent run; backward scales the gradient
        by the incoming one."""

        @staticmethod
        def forward(ctx, theta, target, n, steps, chunk):
            loss, grad = loss_and_grad(theta.detach().cpu().double().numpy().reshape(-1), target, n, steps, chunk)
            # gradient with theta's device, dtype and shape; the loss on theta's device
            ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype, device=theta.device).view(theta.shape))
            return torch.as_tensor(loss, dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, g):
            (grad,) = ctx.saved_tensors
            return g.to(grad.device) * grad, None, None, None, None

    return ProbeLoss


def libtorch_module():
    """The libtorch extension (tenjin_amd/tenjin_torch.so, built by
    tenjin_amd.build.build_torch): ProbeLoss as a C++ torch::autograd::Function."""
    import importlib

    import torch  # noqa: F401  (the extension links libtorch)

    return importlib.import_module("tenjin_amd.tenjin_torch")


def probe_loss_libtorch(theta, target: np.ndarray, n: int = 512, steps: int = 200, chunk: int = 1, seed: int = 4):
    """probe_loss_torch through the C++ torch::autograd::Function of the
    libtorch extension: the level, its initial primitives and the probe come
    from the differentiated plate impact; theta (a torch tensor of 6 values)
    sets its materials inside the Function, so loss.backward() fills
    theta.grad."""
    import ctypes as ct

    import torch

    from . import level_desc

    pb = configs.plate_impact_differentiated(n=n, seed=seed, steps=steps)
    d = level_desc(pb)
    desc = torch.frombuffer(bytearray(ct.string_at(ct.addressof(d), ct.sizeof(d))), dtype=torch.uint8)
    prims = torch.from_numpy(np.ascontiguousarray(pb.prims, dtype=np.float64)).to("cuda")
    tgt = torch.as_tensor(np.asarray(target, dtype=np.float64))
    probe = list(configs.plate_probe(pb))
    return libtorch_module().probe_loss(theta, prims, tgt, desc, steps, probe, chunk, pb.shift_pressure)


def probe_loss_torch(theta, target: np.ndarray, n: int = 512, steps: int = 200, chunk: int = 1):
    """The calibration loss as a differentiable function of a torch tensor
    theta (6 values); loss.backward() fills theta.grad."""
    return _autograd_function().apply(theta, target, n, steps, chunk)
