"""Differentiable rollouts: the gradient harness over the step library's adjoint.

The library records a rollout (cpgpuWorldTapeBegin) and runs its hand-written
reverse kernels (cpgpuWorldBackward) to return dL/d of the initial state and
of the parameters.  This module maps those onto the quantities a user
differentiates -- initial velocities, masses (the moment follows the mass for
each body's geometry, i = k m) and friction coefficients -- and wraps one
rollout as a torch.autograd.Function.  Torch is used only here, for autograd
plumbing; the rollout and its adjoint run on the device through the C ABI.
"""
import numpy as np

from .cabi import PackedWorld
from .gpu import GPUWorld

DYNAMIC = 0
SPRING = 1   # CPGPU_JOINT_DAMPED_SPRING


def with_params(pw, v0=None, m=None, u=None, joint=None):
    """A copy of the packed description with initial velocities v0 [n, 2],
    masses m [n] (dynamic bodies; others ignored), friction coefficients
    u [n_shapes] and constraint parameters joint [n_constraints, 8]
    (cons_params) replaced."""
    a = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in pw.a.items()}
    if v0 is not None:
        a["body_v"] = np.ascontiguousarray(v0, dtype=np.float64).reshape(-1, 2)
    if m is not None:
        # only changed masses are rewritten, so an unchanged mass keeps its exact moment
        m = np.asarray(m, dtype=np.float64)
        ch = (a["body_type"] == DYNAMIC) & (m != a["body_m"])
        k = a["body_i"][ch] / a["body_m"][ch]
        a["body_m"][ch] = m[ch]
        a["body_m_inv"][ch] = 1.0 / m[ch]
        a["body_i"][ch] = k * m[ch]
        a["body_i_inv"][ch] = 1.0 / (k * m[ch])
    if u is not None:
        a["shape_u"] = np.ascontiguousarray(u, dtype=np.float64)
    if joint is not None:
        a["cons_params"] = np.ascontiguousarray(joint, dtype=np.float64).reshape(-1, 8)
    return PackedWorld(a)


def rollout(pw, steps, dt, checkpoint_every=10):
    """Create a world from pw, record `steps` steps.  Returns the world (its
    final state is the rollout's)."""
    g = GPUWorld(pw)
    g.tape_begin(checkpoint_every)
    g.step(dt, steps)
    return g


def param_grads(pw, adj):
    """dL/d(v0, m, u) from the library's adjoint output `adj` for the
    description pw the rollout started from."""
    dyn = pw.a["body_type"] == DYNAMIC
    m, i = pw.a["body_m"], pw.a["body_i"]
    dm = np.zeros(len(m))
    # 1/m and 1/(k m): d/dm = -1/m^2 and -k/(k m)^2 = -(i/m)/i^2
    dm[dyn] = -adj["m_inv"][dyn] / (m[dyn] * m[dyn]) - adj["i_inv"][dyn] * (i[dyn] / m[dyn]) / (i[dyn] * i[dyn])
    out = {"v0": adj["v"], "w0": adj["w"], "p0": adj["p"], "a0": adj["a"], "m": dm, "u": adj["u"],
           "joint": adj["joint"]}
    # damped springs: stiffness and damping (cons_params columns 5 and 6)
    spring = pw.a["cons_type"] == SPRING if pw.n_constraints else np.zeros(0, bool)
    out["spring_k"] = adj["joint"][spring, 5] if pw.n_constraints else np.zeros(0)
    out["spring_c"] = adj["joint"][spring, 6] if pw.n_constraints else np.zeros(0)
    return out


def position_loss(p, targets, dyn):
    """L = sum over dynamic bodies of |p - target|^2 and dL/dp."""
    d = (p - targets) * dyn[:, None]
    return float(np.sum(d * d)), 2.0 * d


def loss_and_grads(pw, targets, steps, dt, checkpoint_every=10):
    """The position loss after a rollout of pw and its gradients."""
    dyn = (pw.a["body_type"] == DYNAMIC).astype(np.float64)
    g = rollout(pw, steps, dt, checkpoint_every)
    L, gp = position_loss(g.bodies().p, targets, dyn)
    grads = param_grads(pw, g.backward(g_p=gp))
    g.close()
    return L, grads


def _torch_function():
    import torch

    class Rollout(torch.autograd.Function):
        """p_T = rollout(v0, m, u): final body positions after `steps` steps of
        the packed spaces with these initial velocities, masses and friction
        coefficients (float64 tensors)."""

        @staticmethod
        def forward(ctx, v0, m, u, pw, steps, dt, checkpoint_every):
            pwp = with_params(pw, v0.detach().cpu().numpy(), m.detach().cpu().numpy(), u.detach().cpu().numpy())
            world = rollout(pwp, steps, dt, checkpoint_every)
            ctx.world, ctx.pw = world, pwp
            return torch.from_numpy(world.bodies().p.copy())

        @staticmethod
        def backward(ctx, grad_p):
            adj = ctx.world.backward(g_p=grad_p.detach().cpu().numpy())
            ctx.world.close()
            g = param_grads(ctx.pw, adj)
            return (torch.from_numpy(g["v0"]), torch.from_numpy(g["m"]), torch.from_numpy(g["u"]),
                    None, None, None, None)

    return Rollout


def torch_rollout(v0, m, u, pw, steps, dt, checkpoint_every=10):
    """Differentiable final positions (torch.autograd) of a rollout of pw."""
    return _torch_function().apply(v0, m, u, pw, steps, dt, checkpoint_every)
