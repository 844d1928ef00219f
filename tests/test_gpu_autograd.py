"""The torch.autograd wrapper of the tracer step (nemo_hip.autograd) on the
GPU, against the oracle's adjoint:

* J = 1/2 sum wgt (SST - target)^2 after six steps: the gradient torch
  computes through TracerStep (each backward the library's adjoint with the
  forward step's limiter factors frozen) equals the gradient of the oracle's
  adjoint chain (the same factors, tl_step / ad_step of the oracle) to
  1e-12, for FCT orders 2/2 and 4/4 on a closed and a doubly periodic basin;
* the same window with checkpoints every two steps (segments recomputed in
  the backward pass) gives the same gradient bit for bit;
* the gradient matches central differences of the frozen-limiter model
  along random directions (1e-8).
"""
import numpy as np
import pytest

from nemo_hip import configs
from test_gpu_parity import Env, dev, sync
from test_oracle_adjoint import KT, ad_step, coef

torch = pytest.importorskip("torch")
ag = pytest.importorskip("nemo_hip.autograd")


def oracle_tl_step(o, d, F, c, tb, tn, kn):
    s3 = (d.jpi, d.jpj, d.jpk)
    ta = np.zeros(s3, order="F")
    o.fct_tl(F["p2dt"], F["zun"], F["zvn"], F["zwn"], tb, tn, *c, ta, kn=kn)
    o.tra_ldf_lap(KT, 1, "TRA", F["ahtu"], F["ahtv"], tb, ta, 1, 1)
    o.tra_zdf_imp(KT, 1, "TRA", F["p2dt"], tb, ta, 1, F["avt"], F["avt"])
    B, N, A = (np.zeros(s3 + (2,), order="F") for _ in range(3))
    B[..., 0], N[..., 0], A[..., 0] = tb, tn, ta
    o.tra_nxt(KT, B, N, A)
    return np.asfortranarray(B[..., 0]), np.asfortranarray(N[..., 0])


def oracle_ad_step(o, d, F, c, lb, ln, kn):
    s3 = (d.jpi, d.jpj, d.jpk)
    lb, ln = lb.copy(order="F"), ln.copy(order="F")
    la = np.zeros(s3, order="F")
    o.tra_nxt_ad(lb, ln, la)
    o.zdf_imp_ad(F["p2dt"], F["avt"], la, lb)
    o.ldf_lap_ad(F["ahtu"], F["ahtv"], la, lb)
    o.fct_ad(F["zun"], F["zvn"], F["zwn"], *c, la, lb, ln, kn=kn)
    return lb, ln


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["basin", "cyc7"])
@pytest.mark.parametrize("kn", [(2, 2), (4, 4)], ids=["fct22", "fct44"])
def test_autograd_gradient_equals_oracle_adjoint(hip, case, kn):
    nsteps = 6
    with torch.cuda.stream(torch.cuda.ExternalStream(hip.stream())):
        e = Env(hip, case)
        d, S, o = e.d, e.S, e.o
        s3 = e.s3
        z = [np.zeros(s3, order="F") for _ in range(3)]
        o.tra_adv_trp(S["un"], S["vn"], S["wn"], *z)
        F = {"zun": z[0], "zvn": z[1], "zwn": z[2], "ahtu": S["ahtu"], "ahtv": S["ahtv"], "avt": S["avt"],
             "p2dt": 2.0 * d.opts["rdt"]}
        model = ag.TracerModel(hip, dev(z[0]), dev(z[1]), dev(z[2]), dev(S["ahtu"]), dev(S["ahtv"]),
                               dev(S["avt"]), F["p2dt"], kn=kn, kt=KT)
        T0 = np.asfortranarray(S["tsn"][..., 0])
        o.lbc_lnk(T0, "T")
        rng = np.random.default_rng(71)
        t, *_ = configs.masks(d)
        wgt = np.zeros(s3)
        wgt[1:-1, 1:-1, 0] = t[1:-1, 1:-1, 0]
        # the oracle: forward with each state's own factors, then the adjoint chain
        b, n, cs = T0.copy(order="F"), T0.copy(order="F"), []
        for _ in range(nsteps):
            c = coef(o, d, F, b, n, kn)
            cs.append(c)
            b, n = oracle_tl_step(o, d, F, c, b, n, kn)
        target = np.asfortranarray(n + 0.1 * rng.standard_normal(s3))
        lb, ln = np.zeros(s3, order="F"), np.asfortranarray(wgt * (n - target))
        for c in reversed(cs):
            lb, ln = oracle_ad_step(o, d, F, c, lb, ln, kn)
        g_oracle = lb + ln
        # torch through the wrapper, with and without checkpoints
        wgt_d, target_d = dev(wgt), dev(target)
        grads = []
        for every in (None, 2):
            T = dev(T0).requires_grad_(True)
            nb, nn = ag.run(model, T, T, nsteps, every=every)
            J = 0.5 * torch.sum(wgt_d * (nn - target_d) ** 2)
            J.backward()
            sync(hip)
            grads.append(T.grad.detach().clone())
        assert torch.equal(grads[0], grads[1]), "checkpointed gradient differs"
        g = grads[0].cpu().numpy().reshape(s3, order="F")
        scale = np.abs(g_oracle).max()
        assert scale > 0
        err = np.abs(g - g_oracle).max() / scale
        assert err <= 1e-12, f"{case} {kn}: gradient against the oracle adjoint {err:.2e}"
        # central differences of the frozen-limiter model
        cs_d = [[dev(a) for a in c] for c in cs]

        def Jfrozen(Tv):
            b, n = Tv.clone(), Tv.clone()
            for c in cs_d:
                b, n = model.step(c, b, n)
            return 0.5 * float(torch.sum(wgt_d * (n - target_d) ** 2))

        for q in range(3):
            dirn = dev(rng.standard_normal(s3))
            eps = 1e-3
            fd = (Jfrozen(dev(T0) + eps * dirn) - Jfrozen(dev(T0) - eps * dirn)) / (2 * eps)
            an = float(torch.dot(grads[0], dirn))
            assert abs(fd - an) <= 1e-8 * max(abs(an), 1e-30), (q, fd, an)
