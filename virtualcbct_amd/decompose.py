"""Dual-energy elemental decomposition by gradient descent on the differentiable projector.

Several problems (phantoms of different slice counts) are optimised together by Adam on the loss
    sum over problems, spectra and rays of (p_s - t_s)^2,
p_s the polychromatic log projection of a problem's six FP32 density maps under spectrum s.

Two paths compute the same iterations:

* `decompose` (the default, fused): per problem ONE traversal of its rays gives the log projections under all
  spectra (the material line integrals do not depend on the spectrum) and the summed per-ray coefficients, and
  ONE scatter gives the gradient (vcbct_projector_multi_loss_grad); the Adam update runs on the device
  (vcbct_adam_step, step count on the device).  After one eager iteration the whole iteration of every problem
  is captured in a hipGraph and replayed.  Checkpoints (maps, Adam moments, step, loss history) are written as
  .npz files and a run can be resumed from one.
* `decompose_autograd` (the per-problem reference path): torch autograd over `poly_logproj`, one log projection
  and one backward pass per (problem, spectrum), torch.optim.Adam.

There is no reference counterpart; the reference's scatter U-Net (Unet_scatter.py) consumes such projections.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import _lib
from ._lib import stream_arg as _stream
from .autograd import poly_logproj


def decompose_autograd(problems, iters=200, lr=0.01, init=None, nonneg=True, callback=None):
    """problems: list of dicts {"projectors": [Projector per spectrum, same geometry], "targets": [p per
    spectrum, FP32 device tensors], "shape": (6, Z, N, N), "xfactor": float}.  Returns the list of fitted
    maps (FP32 [mat][z][y][x] device tensors) and the loss history."""
    maps = []
    for i, pr in enumerate(problems):
        m = (init[i].clone() if init is not None else
             torch.zeros(pr["shape"], dtype=torch.float32, device=pr["targets"][0].device))
        maps.append(m.requires_grad_(True))
    opt = torch.optim.Adam(maps, lr=lr)
    hist = []
    for it in range(iters):
        opt.zero_grad(set_to_none=True)
        total = None
        for m, pr in zip(maps, problems):
            for P, t in zip(pr["projectors"], pr["targets"]):
                p = poly_logproj(m, P, pr.get("xfactor", 1.0))
                loss = ((p.double() - t.double()) ** 2).sum()
                loss.backward()
                total = loss.detach() if total is None else total + loss.detach()
        opt.step()
        if nonneg:
            with torch.no_grad():
                for m in maps:
                    m.clamp_(min=0.0)
        hist.append(total)
        if callback:
            callback(it, total)
    hist = [float(h) for h in hist]  # one synchronisation at the end
    return [m.detach() for m in maps], hist


class FusedDecomposition:
    """The fused, graph-captured decomposition state: per problem its projector (all spectra set with
    set_spectra), stacked targets, maps, gradient, Adam moments and step count, all on the device."""

    def __init__(self, problems, lr=0.01, init=None, nonneg=True, betas=(0.9, 0.999), eps=1e-8, graph=True):
        self.lr, self.betas, self.eps = float(lr), betas, float(eps)
        self.lower = 0.0 if nonneg else -np.inf
        self.graph_on = bool(graph)
        self.items = []
        for i, pr in enumerate(problems):
            P = pr["projectors"][0]
            spectra = [Q.spectrum for Q in pr["projectors"]]
            P.set_spectra(spectra)
            T = torch.stack([t.to(torch.float32) for t in pr["targets"]]).contiguous()
            dev = T.device
            x = (init[i].clone() if init is not None else torch.zeros(pr["shape"], dtype=torch.float32, device=dev))
            x = x.contiguous()
            self.items.append({
                "P": P, "T": T, "x": x, "g": torch.empty_like(x), "m": torch.zeros_like(x), "v": torch.zeros_like(x),
                "t": torch.zeros(1, dtype=torch.int32, device=dev), "loss": torch.zeros(1, dtype=torch.float64,
                                                                                         device=dev),
                "xf": float(pr.get("xfactor", 1.0))})
        self.graph = None
        self.hist = []   # per iteration: FP64 device tensor of the summed loss (read back at the end)
        self.iteration = 0

    @property
    def channels(self):
        return [it["P"].multi_channels for it in self.items]

    def _step_eager(self):
        L = _lib.lib()
        for it in self.items:
            P = it["P"]
            P.set_volume(it["x"])
            P.multi_loss_grad(it["T"], it["xf"], grad=it["g"], loss=it["loss"])
            _lib.check(L.vcbct_adam_step(it["x"].numel(), _vp(it["x"]), _vp(it["g"]), _vp(it["m"]), _vp(it["v"]),
                                         _vp(it["t"]), self.lr, float(self.betas[0]), float(self.betas[1]),
                                         self.eps, float(self.lower), _stream()))

    def step(self):
        """One Adam iteration of every problem (replayed from the captured graph after the first)."""
        if self.graph_on and self.graph is None and self.iteration >= 1:
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._step_eager()
            self.graph = g
            # capturing records the iteration without running it
            self.graph.replay()
        elif self.graph is not None:
            self.graph.replay()
        else:
            self._step_eager()
        total = self.items[0]["loss"].clone()
        for it in self.items[1:]:
            total += it["loss"]
        self.hist.append(total)
        self.iteration += 1

    def losses(self):
        return [float(h.item()) for h in self.hist]

    def maps(self):
        return [it["x"] for it in self.items]

    # checkpoint / resume -------------------------------------------------------------------------------------
    def save(self, path):
        """Writes maps, Adam moments, step counts and the loss history to `path` (.npz, no pickled objects)."""
        arrs = {"iteration": np.array(self.iteration), "hist": np.array(self.losses(), np.float64)}
        for i, it in enumerate(self.items):
            for k in ("x", "m", "v", "t"):
                arrs[f"{k}{i}"] = it[k].cpu().numpy()
        tmp = path + ".tmp.npz"
        np.savez(tmp, **arrs)
        os.replace(tmp, path)

    def load(self, path):
        """Restores a checkpoint written by save() into this state (same problems, same shapes)."""
        d = np.load(path, allow_pickle=False)
        for i, it in enumerate(self.items):
            for k in ("x", "m", "v", "t"):
                a = d[f"{k}{i}"]
                if tuple(a.shape) != tuple(it[k].shape):
                    raise ValueError(f"checkpoint {path}: {k}{i} has shape {a.shape}, expected {tuple(it[k].shape)}")
                it[k].copy_(torch.from_numpy(a))
        self.iteration = int(d["iteration"])
        dev = self.items[0]["x"].device
        self.hist = [torch.tensor([h], dtype=torch.float64, device=dev) for h in d["hist"]]


def _vp(t):
    import ctypes as C
    return C.c_void_p(t.data_ptr())


def decompose(problems, iters=200, lr=0.01, init=None, nonneg=True, callback=None, graph=True, checkpoint=None,
              checkpoint_every=0, resume=False):
    """The fused decomposition (module docstring).  problems as decompose_autograd takes them.  checkpoint: a
    .npz path written every `checkpoint_every` iterations and at the end; resume=True continues from it when it
    exists (the iterations already done count towards `iters`).  Returns (fitted maps, loss history)."""
    st = FusedDecomposition(problems, lr=lr, init=init, nonneg=nonneg, graph=graph)
    if resume and checkpoint and os.path.exists(checkpoint):
        st.load(checkpoint)
    while st.iteration < iters:
        st.step()
        if callback:
            callback(st.iteration - 1, st.hist[-1])
        if checkpoint and checkpoint_every and st.iteration % checkpoint_every == 0:
            st.save(checkpoint)
    if checkpoint:
        st.save(checkpoint)
    return [m.detach() for m in st.maps()], st.losses()
