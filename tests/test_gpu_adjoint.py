"""GPU adjoint (mafm_adjoint) against the FP64 reference gradient
(oracle/adjoint_ref.py, itself pinned against finite differences on the CPU)
and, end to end through torch.autograd, against central finite differences of
GPU scans.  Bars: FP64 adjoint vs FP64 reference 1e-9 relative; FP32 pair
arithmetic vs FP64 1e-4 relative (in norm); autograd vs finite differences
1e-3 relative."""
import numpy as np
import pytest
import torch

from oracle.adjoint_ref import ift_gradients

pytestmark = pytest.mark.gpu


def _model(gpu_ctx, tmp_path, **over):
    from mechafm_amd import workloads as W
    from mechafm_amd.adjoint import ScanModel
    opts = dict(mini=True, minterm="f", ftol="1e-9", maxsteps="200000", area="0.2 0.2", zhigh="9.0", zlow="8.05")
    opts.update(over)
    wl = W.make_c5(str(tmp_path), **opts)
    return ScanModel(wl.input_path, gpu_ctx, precision="fast64")


def _td(model):
    d = model.desc.tip_dummy
    return {"vdw": d.vdw, "p0": d.p0, "p1": d.p1, "p2": d.p2, "qq": d.qq, "has_coulomb": bool(d.has_coulomb),
            "k": d.k, "r0": d.r0}


def test_adjoint_matches_fp64_reference(gpu_ctx, tmp_path):
    m = _model(gpu_ctx, tmp_path)
    es6, es12, qq = (t.detach().numpy() for t in m.site_constants(m.eps, m.sig, m.q))
    pos = m.pos.numpy()
    fz, tip, dummy = m.forward_np(es6, es12, qq, pos)
    g = np.random.default_rng(2).normal(size=len(fz))
    want = ift_gradients(es12, es6, qq, pos, _td(m), tip, dummy, g)
    for prec, tol in (("fast64", 1e-9), ("fast32", 1e-4)):
        m.adjoint_precision = prec
        got = m.adjoint_np(es6, es12, qq, pos, tip, dummy, g)
        for a, b in zip(got, want):
            assert np.linalg.norm(a - b) <= tol * np.linalg.norm(b), prec


# "cell": vdw_pbc with a 9 A cell (the reference allows it only without
# Coulomb, parse.cpp:390-399): the sites carry the +-1 in-plane images and
# every scan wraps the dummy into the cell (System::setDummyXY), so the
# relaxed tips sit at the wrapped pixels (pixel (0, 0) at (9, 9) here), not
# at the pixel coordinates the records report
CELL = dict(coulomb="off", vdw_pbc="on", cell_a="9.0 0.0 0.0", cell_b="0.0 9.0 0.0", cell_c="0.0 0.0 40.0")


@pytest.mark.parametrize("cell", [False, True], ids=["open", "cell"])
def test_autograd_matches_finite_differences(gpu_ctx, tmp_path, cell):
    m = _model(gpu_ctx, tmp_path, **(CELL if cell else {}))
    if cell:
        rec = m.forward_np(*(t.detach().numpy() for t in m.site_constants(m.eps, m.sig, m.q)), m.pos.numpy())[2]
        assert np.abs(rec[:, 0] - m.last_records["x"]).max() > 1.0, "no pixel wrapped: the case tests nothing"
    rng = np.random.default_rng(4)
    with torch.no_grad():
        target = m.image(m.eps * torch.from_numpy(1 + 0.05 * rng.choice([-1, 1], len(m.eps))),
                         m.sig * torch.from_numpy(1 + 0.05 * rng.choice([-1, 1], len(m.sig))),
                         m.q * torch.from_numpy(1 + 0.05 * rng.choice([-1, 1], len(m.q))))
    eps = m.eps.clone().requires_grad_(True)
    sig = m.sig.clone().requires_grad_(True)
    q = m.q.clone().requires_grad_(True)
    pos = m.pos.clone().requires_grad_(True)
    loss = ((m.image(eps, sig, q, pos) - target) ** 2).sum()
    loss.backward()

    def L(**kw):
        with torch.no_grad():
            args = dict(eps=m.eps, sig=m.sig, q=m.q, pos=m.pos)
            args.update(kw)
            return float(((m.image(**args) - target) ** 2).sum())

    checks = []
    for k in range(len(m.eps)):
        h = 1e-4 * float(m.eps[k])
        e_p, e_m = m.eps.clone(), m.eps.clone()
        e_p[k] += h
        e_m[k] -= h
        checks.append((f"eps[{m.type_names[k]}]", float(eps.grad[k]), (L(eps=e_p) - L(eps=e_m)) / (2 * h)))
        h = 1e-4 * float(m.sig[k])
        s_p, s_m = m.sig.clone(), m.sig.clone()
        s_p[k] += h
        s_m[k] -= h
        checks.append((f"sig[{m.type_names[k]}]", float(sig.grad[k]), (L(sig=s_p) - L(sig=s_m)) / (2 * h)))
    i = int(torch.argmax(q.grad.abs()))
    h = 1e-4 * abs(float(m.q[i]))
    q_p, q_m = m.q.clone(), m.q.clone()
    q_p[i] += h
    q_m[i] -= h
    checks.append((f"q[{i}]", float(q.grad[i]), (L(q=q_p) - L(q=q_m)) / (2 * h)))
    j = int(torch.argmax(pos.grad.abs().max(1).values))
    for c in range(3):
        p_p, p_m = m.pos.clone(), m.pos.clone()
        p_p[j, c] += 1e-5
        p_m[j, c] -= 1e-5
        checks.append((f"pos[{j},{c}]", float(pos.grad[j, c]), (L(pos=p_p) - L(pos=p_m)) / 2e-5))
    for name, ad, fd in checks:
        print(f"{name}: adjoint {ad:.6e} FD {fd:.6e}")
        assert ad == pytest.approx(fd, rel=1e-3, abs=1e-9 * float(loss.detach())), name
