"""TEST INFRASTRUCTURE ONLY -- FP64 reference gradient for the adjoint.

The reference has no gradients; this is the implicit-function-theorem
gradient of an image loss written out in numpy FP64, used to check
libmafm's adjoint kernels.  It is itself pinned against central finite
differences of the oracle's scan (tests/test_adjoint_cpu.py).

Pair force on the tip from site i (LJ + Coulomb, interactions.cpp:15-51):
    F_i = f(r) d,  d = p - a_i,  f = 12 A r^-14 - 6 B r^-8 + qq r^-3
    dF_i/dp = f I + (f'/r) d d^T,  f'/r = -168 A r^-16 + 48 B r^-10 - 3 qq r^-5
Tip-dummy VDW + Coulomb and the lateral spring (interactions.cpp:252-266)
add to the Jacobian J of the total tip force.  With g = dL/dFz:
    J w = grad_p Fz_ts,   v = g (e_z - w),
    dL/dA_i = sum v . 12 r^-14 d,  dL/dB_i = sum v . (-6 r^-8 d),
    dL/dqq_i = sum v . r^-3 d,     dL/da_i = -sum (f v + f'/r (d.v) d).
"""
import numpy as np


def _radial(r2, A, B, qq):
    i2 = 1.0 / r2
    i8 = i2 ** 4
    i14 = i8 * i2 ** 3
    ir3 = i2 * np.sqrt(i2)
    f = 12 * A * i14 - 6 * B * i8 + qq * ir3
    fp = -168 * A * i14 * i2 + 48 * B * i8 * i2 - 3 * qq * ir3 * i2
    return f, fp, i2, i8, i14, ir3


def tip_dummy_jacobian(td, tip, dummy):
    """td: dict(vdw, p0, p1, p2, qq, has_coulomb, k, r0); returns (P,3,3)."""
    d = tip - dummy
    r2 = (d ** 2).sum(1)
    r = np.sqrt(r2)
    if td["vdw"] == 0:
        f, fp, *_ = _radial(r2, td["p1"], td["p0"], 0.0)
    else:
        a, c = td["p1"], 2 * td["p0"] * td["p1"]
        x = np.exp(-a * (r - td["p2"]))
        f = c * (x * x - x) / r
        fp = c * ((-2 * a * x * x + a * x) / r - (x * x - x) / r2) / r
    if td.get("has_coulomb"):
        f = f + td["qq"] / (r2 * r)
        fp = fp - 3 * td["qq"] / (r2 * r) / r2
    J = f[:, None, None] * np.eye(3)[None] + fp[:, None, None] * d[:, :, None] * d[:, None, :]
    rl = np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2)
    ok = rl > 1e-10
    a_ = np.where(ok, -2 * td["k"] * (1 - td["r0"] / np.where(ok, rl, 1)), 0.0)
    b_ = np.where(ok, -2 * td["k"] * td["r0"] / np.where(ok, rl, 1) ** 3, 0.0)
    J[:, 0, 0] += a_ + b_ * d[:, 0] ** 2
    J[:, 1, 1] += a_ + b_ * d[:, 1] ** 2
    J[:, 0, 1] += b_ * d[:, 0] * d[:, 1]
    J[:, 1, 0] += b_ * d[:, 0] * d[:, 1]
    return J


def ift_gradients(A, B, qq, pos, td, tip, dummy, g):
    """Site arrays A=es12, B=es6, qq (N,), pos (N,3); positions (P,3); g (P,)."""
    d = tip[:, None, :] - pos[None, :, :]                     # (P, N, 3)
    r2 = (d ** 2).sum(-1)
    f, fp, i2, i8, i14, ir3 = _radial(r2, A[None], B[None], qq[None])
    Jts = (f.sum(1)[:, None, None] * np.eye(3)[None]
           + np.einsum("pn,pni,pnj->pij", fp, d, d))
    J = Jts + tip_dummy_jacobian(td, tip, dummy)
    u = Jts[:, 2, :]
    w = np.linalg.solve(J, u[:, :, None])[:, :, 0]
    v = -g[:, None] * w
    v[:, 2] += g
    vd = np.einsum("pni,pi->pn", d, v)
    d_B = (-6 * i8 * vd).sum(0)
    d_A = (12 * i14 * vd).sum(0)
    d_qq = (ir3 * vd).sum(0)
    d_pos = -(f[:, :, None] * v[:, None, :] + (fp * vd)[:, :, None] * d).sum(0)
    return d_B, d_A, d_qq, d_pos
