"""``torch.autograd.Function`` over the batched adjoint (``pf_batch_lml_grad``).

For testing the adjoint from PyTorch: the GMM parameters are tensors shared by
every sequence of a fixed batch problem (particle counts, lengths, detections,
uniforms, initial components); the forward runs the batch on the GPU and
returns the summed log-marginal-likelihood, the backward returns the adjoint's
gradients (resampling and component draws held fixed, as in
``pf_batch_lml_grad``).
"""
from __future__ import annotations

import numpy as np
import torch

from .batch import RbpfBatch
from .devmem import DeviceArray


class BatchProblem:
    """A fixed batch: one model structure (pca_proj, pca_mean, sel, R, joint
    kinds) shared by S sequences with their own detections, uniforms and
    initial components."""

    def __init__(self, Ns, Ts, pca_proj, pca_mean, sel, R, clutter, count, cand, uniforms, bins,
                 checkpoint_every=50, **cfg):
        self.Ns, self.Ts = list(Ns), list(Ts)
        self.pca_proj, self.pca_mean, self.sel, self.R = pca_proj, pca_mean, sel, R
        self.clutter = clutter
        self.bins = bins
        self.checkpoint_every = checkpoint_every
        self.cfg = cfg
        self.d_cnt = DeviceArray(np.ascontiguousarray(count, np.int32))
        self.d_cand = DeviceArray(np.ascontiguousarray(cand, np.float32))
        self.d_uni = DeviceArray(np.ascontiguousarray(uniforms, np.uint32))
        self.d_lml = DeviceArray.empty(int(np.sum(Ts)), np.float64)
        self.batch = None

    def evaluate(self, means, logits, rates, chol):
        """LML summed over sequences and frames, and its gradient (FP64 numpy)."""
        K, d = means.shape
        if self.batch is None:
            self.batch = RbpfBatch(self.Ns, self.Ts, K=K, m=len(self.sel), **self.cfg)
        b = self.batch
        w = np.exp(logits - np.logaddexp.reduce(logits))
        covs = np.einsum("kij,klj->kil", chol, chol).reshape(K * d, d)
        for s in range(b.S):
            b.load_model(s, means, covs, w, 1.0 - rates, self.pca_proj, self.pca_mean, self.sel, self.R)
        b.set_joint_clutter(self.clutter)
        for s in range(b.S):
            b.reset(s, self.bins[s])
        g = b.lml_grad(self.d_cnt.ptr, self.d_cand.ptr, self.d_uni.ptr, self.d_lml.ptr, self.checkpoint_every)
        lml = self.d_lml.get()
        return float(lml.sum()), {k: v.sum(axis=0) for k, v in g.items()}


class LmlFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means, logits, rates, chol, problem: BatchProblem):
        inputs = (means, logits, rates, chol)
        f, g = problem.evaluate(*(t.detach().cpu().double().numpy() for t in inputs))
        # each gradient on its input's device and in its dtype; the value on
        # the device of the means
        ctx.grads = tuple(torch.from_numpy(np.ascontiguousarray(g[k])).to(device=t.device, dtype=t.dtype)
                          for k, t in zip(("means", "logits", "rates", "chol"), inputs))
        return torch.tensor(f, dtype=torch.float64, device=means.device)

    @staticmethod
    def backward(ctx, gout):
        return tuple(gout.to(device=g.device, dtype=g.dtype) * g for g in ctx.grads) + (None,)


def lml(means, logits, rates, chol, problem: BatchProblem):
    return LmlFunction.apply(means, logits, rates, chol, problem)
