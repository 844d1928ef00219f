"""The tangent linear and adjoint of the tracer step on the GPU (one tracer,
FP64; FCT of orders 2 and 4 with frozen limiter factors, tra_ldf_lap,
tra_zdf_imp, leapfrog tra_nxt):

* the limiter factors equal the oracle's bit for bit; the frozen FCT with
  the state's own factors equals the library's tra_adv_fct bit for bit;
* every tangent-linear and adjoint entry point matches the oracle
  (<= 1e-12 max relative error);
* the composed GPU step passes the dot-product test, and the gradient of
  J = 1/2 sum (SST - target)^2 after six steps matches central finite
  differences of the frozen model (<= 1e-8).
"""
import numpy as np
import pytest

from nemo_hip import configs
from test_gpu_parity import CASE_IDS, Env, check, dev, host, sync
from test_oracle_adjoint import KN, KT, coef, interior

torch = pytest.importorskip("torch")


def prep(e):
    s3 = e.s3
    z = [np.zeros(s3, order="F") for _ in range(3)]
    e.o.tra_adv_trp(e.S["un"], e.S["vn"], e.S["wn"], *z)
    F = {"zun": z[0], "zvn": z[1], "zwn": z[2], "ahtu": e.S["ahtu"], "ahtv": e.S["ahtv"], "avt": e.S["avt"],
         "p2dt": 2.0 * e.d.opts["rdt"]}
    G = {k: (dev(v) if isinstance(v, np.ndarray) else v) for k, v in F.items()}
    return F, G


class GpuStep:
    """the frozen-limiter tracer step and its adjoint through the library"""

    def __init__(self, hip, e, G, kn=(2, 2)):
        self.h, self.e, self.G, self.kn = hip, e, G, kn
        self.n3 = int(np.prod(e.s3))

    def coef(self, tb, tn):
        c = [torch.ones(self.n3, dtype=torch.float64, device="cuda") for _ in range(3)]
        G = self.G
        self.h.tra_adv_fct_coef(G["p2dt"], G["zun"], G["zvn"], G["zwn"], tb, tn, *c, kn=self.kn)
        return c

    def tl(self, c, tb, tn):
        h, G, n3 = self.h, self.G, self.n3
        ta = torch.zeros(n3, dtype=torch.float64, device="cuda")
        h.tra_adv_fct_tl(G["p2dt"], G["zun"], G["zvn"], G["zwn"], tb, tn, *c, ta, kn=self.kn)
        h.tra_ldf_lap(KT, 1, "TRA", G["ahtu"], G["ahtv"], tb, ta, 1, 1)
        h.tra_zdf_imp(KT, 1, "TRA", G["p2dt"], tb, ta, 1, G["avt"], G["avt"])
        B, N, A = (torch.zeros(2 * n3, dtype=torch.float64, device="cuda") for _ in range(3))
        B[:n3], N[:n3], A[:n3] = tb, tn, ta
        h.tra_nxt(KT, B, N, A)
        return B[:n3].clone(), N[:n3].clone()

    def ad(self, c, lb, ln):
        h, G = self.h, self.G
        lb, ln = lb.clone(), ln.clone()
        la = torch.zeros(self.n3, dtype=torch.float64, device="cuda")
        h.tra_nxt_ad(lb, ln, la)
        h.tra_zdf_imp_ad(G["p2dt"], G["avt"], la, lb)
        h.tra_ldf_lap_ad(G["ahtu"], G["ahtv"], la, lb)
        h.tra_adv_fct_ad(G["zun"], G["zvn"], G["zwn"], *c, la, lb, ln, kn=self.kn)
        return lb, ln


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
def test_fct_coef_and_tl(hip, case):
    with torch.cuda.stream(torch.cuda.ExternalStream(hip.stream())):   # torch work in order with the library's
        e = Env(hip, case)
        F, G = prep(e)
        tb, tn = np.asfortranarray(e.S["tsb"][..., 0]), np.asfortranarray(e.S["tsn"][..., 0])
        rng = np.random.default_rng(61)
        xb, xn = (np.asfortranarray(rng.standard_normal(e.s3)) for _ in range(2))
        for kn in KN:
            c_o = coef(e.o, e.d, F, tb, tn, kn=kn)
            st = GpuStep(hip, e, G, kn)
            c_g = st.coef(dev(tb), dev(tn))
            sync(hip)
            for a, b, n in zip(c_g, c_o, "xyz"):
                assert np.array_equal(host(a, e.s3), b), f"c{n} {case} {kn}"
            # with the state's own factors: tra_adv_fct itself
            ta1 = dev(np.zeros(e.s3))
            hip.tra_adv_fct_tl(G["p2dt"], G["zun"], G["zvn"], G["zwn"], dev(tb), dev(tn), *c_g, ta1, kn=kn)
            ta2 = dev(np.zeros(e.s3))
            hip.tra_adv_fct(KT, 1, "TRA", G["p2dt"], G["zun"], G["zvn"], G["zwn"], dev(tb), dev(tn), ta2, 1, *kn)
            sync(hip)
            assert torch.equal(ta1, ta2), f"frozen FCT with own factors {case} {kn}"
            # perturbations against the oracle
            want = np.zeros(e.s3, order="F")
            e.o.fct_tl(F["p2dt"], F["zun"], F["zvn"], F["zwn"], xb, xn, *c_o, want, kn=kn)
            got = dev(np.zeros(e.s3))
            hip.tra_adv_fct_tl(G["p2dt"], G["zun"], G["zvn"], G["zwn"], dev(xb), dev(xn), *c_g, got, kn=kn)
            sync(hip)
            check(host(got, e.s3), want, what=f"fct_tl {case} {kn}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
def test_adjoint_routines(hip, case):
    with torch.cuda.stream(torch.cuda.ExternalStream(hip.stream())):   # torch work in order with the library's
        e = Env(hip, case)
        F, G = prep(e)
        tb, tn = np.asfortranarray(e.S["tsb"][..., 0]), np.asfortranarray(e.S["tsn"][..., 0])
        c_o = coef(e.o, e.d, F, tb, tn)
        c_g = [dev(a) for a in c_o]
        rng = np.random.default_rng(62)
        r = lambda: np.asfortranarray(rng.standard_normal(e.s3))
        mu, b0, n0 = interior(e.d, r()), r(), r()
        # FCT, every order
        for kn in KN:
            wb, wn = b0.copy(order="F"), n0.copy(order="F")
            e.o.fct_ad(F["zun"], F["zvn"], F["zwn"], *c_o, mu, wb, wn, kn=kn)
            gb, gn = dev(b0), dev(n0)
            hip.tra_adv_fct_ad(G["zun"], G["zvn"], G["zwn"], *c_g, dev(mu), gb, gn, kn=kn)
            sync(hip)
            check(host(gb, e.s3), wb, what=f"fct_ad tb {case} {kn}")
            check(host(gn, e.s3), wn, what=f"fct_ad tn {case} {kn}")
        # laplacian
        wb = b0.copy(order="F")
        e.o.ldf_lap_ad(F["ahtu"], F["ahtv"], mu, wb)
        gb = dev(b0)
        hip.tra_ldf_lap_ad(G["ahtu"], G["ahtv"], dev(mu), gb)
        sync(hip)
        check(host(gb, e.s3), wb, what=f"ldf_lap_ad {case}")
        # implicit vertical diffusion
        la0 = r()
        wa, wb = la0.copy(order="F"), b0.copy(order="F")
        e.o.zdf_imp_ad(F["p2dt"], F["avt"], wa, wb)
        ga, gb = dev(la0), dev(b0)
        hip.tra_zdf_imp_ad(G["p2dt"], G["avt"], ga, gb)
        sync(hip)
        check(host(ga, e.s3), wa, what=f"zdf_imp_ad ta {case}")
        check(host(gb, e.s3), wb, what=f"zdf_imp_ad tb {case}")
        # leapfrog filter
        wb, wn, wa = b0.copy(order="F"), n0.copy(order="F"), np.zeros(e.s3, order="F")
        e.o.tra_nxt_ad(wb, wn, wa)
        gb, gn, ga = dev(b0), dev(n0), dev(r())
        hip.tra_nxt_ad(gb, gn, ga)
        sync(hip)
        check(host(gb, e.s3), wb, what=f"tra_nxt_ad tb {case}")
        check(host(gn, e.s3), wn, what=f"tra_nxt_ad tn {case}")
        check(host(ga, e.s3), wa, what=f"tra_nxt_ad ta {case}")
        # the transposed lbc
        w = b0.copy(order="F")
        e.o.lbc_adj(w, e.d.jpk, "T")
        g = dev(b0)
        hip.lbc_adj(g, e.d.jpk)
        sync(hip)
        check(host(g, e.s3), w, what=f"lbc_adj {case}")


@pytest.mark.gpu
@pytest.mark.parametrize("kn", [(2, 2), (4, 4)], ids=["fct22", "fct44"])
@pytest.mark.parametrize("case", ["basin", "cyc7"])
def test_step_dot_product_and_gradient(hip, case, kn):
    with torch.cuda.stream(torch.cuda.ExternalStream(hip.stream())):   # torch work in order with the library's
        e = Env(hip, case)
        F, G = prep(e)
        st = GpuStep(hip, e, G, kn)
        rng = np.random.default_rng(63)
        T0 = np.asfortranarray(e.S["tsn"][..., 0])
        e.o.lbc_lnk(T0, "T")
        cs, b, n = [], dev(T0), dev(T0)
        for _ in range(6):
            c = st.coef(b, n)
            cs.append(c)
            b, n = st.tl(c, b, n)
        # dot product of one step
        xb, xn, yb, yn = (dev(rng.standard_normal(e.s3)) for _ in range(4))
        b1, n1 = st.tl(cs[0], xb, xn)
        lb, ln = st.ad(cs[0], yb, yn)
        sync(hip)
        lhs = float(torch.dot(b1, yb) + torch.dot(n1, yn))
        rhs = float(torch.dot(xb, lb) + torch.dot(xn, ln))
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs)), (lhs, rhs)
        # gradient of J after six steps against central differences
        t, *_ = configs.masks(e.d)
        wgt = np.zeros(e.s3)
        wgt[1:-1, 1:-1, 0] = t[1:-1, 1:-1, 0]
        wgt = dev(wgt)
        target = n.clone() + 0.1 * dev(rng.standard_normal(e.s3))

        def J(T):
            b, n = T.clone(), T.clone()
            for c in cs:
                b, n = st.tl(c, b, n)
            return 0.5 * float(torch.sum(wgt * (n - target) ** 2)), n

        _, nN = J(dev(T0))
        lb, ln = torch.zeros_like(nN), wgt * (nN - target)
        for c in reversed(cs):
            lb, ln = st.ad(c, lb, ln)
        grad = lb + ln
        hip.lbc_adj(grad, e.d.jpk)
        for q in range(6):
            dirn = np.asfortranarray(interior(e.d, rng.standard_normal(e.s3)))
            eps = 1e-3
            tp, tm = T0 + eps * dirn, T0 - eps * dirn
            e.o.lbc_lnk(tp, "T")
            e.o.lbc_lnk(tm, "T")
            fd = (J(dev(tp))[0] - J(dev(tm))[0]) / (2 * eps)
            ad = float(torch.dot(grad, dev(dirn)))
            assert abs(fd - ad) <= 1e-8 * abs(fd), (q, fd, ad)
