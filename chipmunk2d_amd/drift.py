"""Drift of the FP32 step against the FP64 step.

A CPGPU_FP32 world runs the same kernels as a CPGPU_FP64 world, compiled with
float arithmetic.  Its trajectory leaves the double one from the first step,
by float rounding in the broadphase bounds, the contact geometry and the
impulse sums.  In a resting or slow scene the gap stays near float precision.
In a chaotic scene (a pile of thousands of colliding shapes) it grows until
the two runs are different samples of the same physics.  The report
therefore gives the per-body state difference at several step counts,
together with quantities that a correct FP32 run must keep close: kinetic
energy (relative to the FP64 run's while the scene moves; a resting scene's
kinetic energy is round-off in both runs), the last step's contact count and
the mean height of the dynamic bodies.

The FP64 world is bit-for-bit the double-precision oracle (the device step
and the oracle perform the same IEEE operations in the same order), so drift
against it is drift against the oracle.
"""
import numpy as np

from .cabi import FP32, FP64
from .gpu import GPUWorld

DYNAMIC = 0


def _kinetic(pw, b, dyn):
    m, i = pw.a["body_m"][dyn], pw.a["body_i"][dyn]
    v, w = b.v[dyn], b.w[dyn]
    return float(0.5 * np.sum(m * np.sum(v * v, axis=1)) + 0.5 * np.sum(i * w * w))


def compare(pw, b64, b32, s64, s32):
    """One drift record from the two worlds' bodies and stats."""
    dyn = pw.a["body_type"] == DYNAMIC
    dp = np.linalg.norm(b32.p[dyn] - b64.p[dyn], axis=1) if dyn.any() else np.zeros(1)
    da = np.abs(b32.a[dyn] - b64.a[dyn]) if dyn.any() else np.zeros(1)
    dv = np.linalg.norm(b32.v[dyn] - b64.v[dyn], axis=1) if dyn.any() else np.zeros(1)
    ke64, ke32 = _kinetic(pw, b64, dyn), _kinetic(pw, b32, dyn)
    scale = float(np.max(np.abs(b64.p[dyn]))) if dyn.any() else 1.0
    return {
        "max_dp": float(np.max(dp)),
        "rms_dp": float(np.sqrt(np.mean(dp * dp))),
        "max_dp_rel": float(np.max(dp)) / max(scale, 1.0),
        "max_da": float(np.max(da)),
        "max_dv": float(np.max(dv)),
        "kinetic_fp64": ke64,
        "kinetic_fp32": ke32,
        "kinetic_rel": abs(ke32 - ke64) / max(ke64, 1e-300),
        "mean_y_fp64": float(np.mean(b64.p[dyn, 1])) if dyn.any() else 0.0,
        "mean_y_fp32": float(np.mean(b32.p[dyn, 1])) if dyn.any() else 0.0,
        "contacts_fp64": int(s64["contacts"]),
        "contacts_fp32": int(s32["contacts"]),
        "finite": bool(np.all(np.isfinite(b32.p)) and np.all(np.isfinite(b32.v))),
    }


def drift_report(pw, dt, checkpoints=(1, 10, 100, 300), w64=None, w32=None):
    """Step an FP64 and an FP32 world of the packed description `pw` side by
    side and return one drift record per checkpoint step count.  Worlds may
    be passed in (already created from pw and not yet stepped)."""
    own = w64 is None
    w64 = w64 or GPUWorld(pw, FP64)
    w32 = w32 or GPUWorld(pw, FP32)
    out, done = [], 0
    for c in sorted(checkpoints):
        if c > done:
            w64.step(dt, c - done)
            w32.step(dt, c - done)
            done = c
        rec = compare(pw, w64.bodies(), w32.bodies(), w64.stats(), w32.stats())
        rec["step"] = c
        out.append(rec)
    if own:
        w64.close()
        w32.close()
    return out
