"""Phase-only hologram optimisation (no reference counterpart).

Projector phases theta (amplitude fixed at 1) on an n_proj^2 plane are
propagated to an n_img^2 image plane; the loss is the MSE between the
normalised intensity I / mean(I) and the target T / mean(T) (mean
normalisation keeps the loss differentiable everywhere, unlike a max, and of
order one, so Adam's epsilon does not swamp the gradients).
Adam (lr 0.05) updates theta; gradients come from the hand-written adjoint.

    python -m holo_amd.hologram --steps 200       (via __graft_entry__.load_package)
"""
from __future__ import annotations

import sys
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import FP32_SPLIT
from .torch_ops import Propagator
from .workloads import hologram as hologram_workload


def intensity_loss(z: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """MSE between I / mean(I) and `target` (already divided by its mean)."""
    I = z[:, 0] ** 2 + z[:, 1] ** 2
    return torch.mean((I / I.mean() - target) ** 2)


@dataclass
class HologramResult:
    losses: list
    ms_per_step: float
    theta: torch.Tensor
    step_ms: dict = None         # per-step HIP-event times (median, std, ...), step_events
    forward_rel_l2: float = None  # error report of the final forward field, report_error


def optimise(steps: int = 200, lr: float = 0.05, seed: int = 0, n_proj: int = 1024,
             n_img: int = 512, precision: int = FP32_SPLIT, device: int = 0,
             warmup: int = 1, log_every: int = 0, z: float = 0.35,
             step_events: bool = False, report_error: bool = False) -> HologramResult:
    """Adam on theta.  step_events: every step is bracketed by HIP events on
    torch's current stream (the stream the library launches on) and the result
    carries their median / std; report_error: the FP32-split forward field of
    the final theta is evaluated once more with its error report."""
    wl = hologram_workload(seed, n_proj, n_img, z)
    st = wl.stages[0]
    dev = torch.device("cuda", device)
    u = torch.from_numpy(st.u).to(dev)
    v = torch.from_numpy(st.v).to(dev)
    prop = Propagator(u, v, st.sign, st.lam, precision)
    target = torch.from_numpy(wl.extra["target"] / wl.extra["target"].mean()).to(dev)
    amp = torch.ones(u.shape[0], dtype=torch.float64, device=dev)
    theta = torch.from_numpy(st.phase.copy()).to(dev).requires_grad_(True)
    opt = torch.optim.Adam([theta], lr=lr)
    losses = []

    def step():
        opt.zero_grad(set_to_none=True)
        loss = intensity_loss(prop(amp, theta), target)
        loss.backward()
        opt.step()
        return loss

    for _ in range(warmup):
        step()
    torch.cuda.synchronize(dev)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps if step_events else 0)]
    t0 = time.perf_counter()
    for i in range(steps):
        if step_events:
            ev[i][0].record()
        loss = step()
        if step_events:
            ev[i][1].record()
        if log_every and i % log_every == 0:
            losses.append(float(loss.detach()))
            print(f"step {i}: loss {losses[-1]:.4e}", file=sys.stderr, flush=True)
    torch.cuda.synchronize(dev)
    dt = time.perf_counter() - t0
    if not log_every:
        losses.append(float(loss.detach()))
    res = HologramResult(losses, 1e3 * dt / max(steps, 1), theta.detach())
    if step_events and steps:
        ms = np.array([a.elapsed_time(b) for a, b in ev])
        res.step_ms = {"median_ms": float(np.median(ms)), "std_ms": float(np.std(ms)),
                       "mean_ms": float(np.mean(ms)), "min_ms": float(np.min(ms)),
                       "repeats": int(steps)}
    if report_error:
        with torch.no_grad():
            prop.plan_u.transform(amp, theta.detach(), prop.v, prop.sign, prop.lam,
                                  prop.precision, report_error=True)
            res.forward_rel_l2 = float(prop.ctx.last_rel_l2.item())
    return res


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--lr", type=float, default=0.05)
    a = ap.parse_args()
    r = optimise(a.steps, a.lr, log_every=10)
    print(f"{r.ms_per_step:.1f} ms/step, final loss {r.losses[-1]:.4e}")
    np.save("hologram_theta.npy", r.theta.cpu().numpy())
