"""torch.autograd wrapper of the tracer step and its hand-written adjoint.

The tracer step is the one the adjoint entry points of the C-ABI transpose
(include/nemo_hip.h, "Tangent linear and adjoint of the tracer step"): for
one tracer in FP64, tra_adv_fct (orders kn) with the limiter factors of
nonosc (TRA/traadv_fct.F90:404-426), tra_ldf_lap, tra_zdf_imp and the
leapfrog tra_nxt, the velocities and diffusivities fixed.  Forward computes
the factors of the state (nemo_hip_tra_adv_fct_coef) and applies them, which
is the nonlinear step bit for bit; backward applies the adjoint entry points
with those factors frozen (the transpose of the step linearised with the
limiter held).  Nothing here computes: every operation is a library call on
resident device tensors, in order on the library stream.

run() steps a window and, with `every`, keeps only the state at every
`every`-th step: the backward pass recomputes each segment from its
checkpoint (torch.utils.checkpoint), so memory holds one segment's states.
"""
import torch
import torch.utils.checkpoint as tcp

from . import NEMO_HIP_PTR_DEVICE


class TracerModel:
    """The fixed fields of the tracer step: transports (zun, zvn, zwn of
    tra_adv), lateral diffusivities, vertical diffusivity, time step, FCT
    orders and the time-step index (kt > nit000: a leapfrog step)."""

    def __init__(self, h, zun, zvn, zwn, ahtu, ahtv, avt, p2dt, kn=(4, 4), kt=5):
        if h.mode != NEMO_HIP_PTR_DEVICE:
            raise ValueError("the autograd wrapper works on resident fields (device-pointer mode)")
        self.h, self.zun, self.zvn, self.zwn = h, zun, zvn, zwn
        self.ahtu, self.ahtv, self.avt = ahtu, ahtv, avt
        self.p2dt, self.kn, self.kt = float(p2dt), tuple(kn), kt
        self.n3 = zun.numel()
        self.stream = torch.cuda.ExternalStream(h.stream())

    def _z(self, n=1):
        return torch.zeros(n * self.n3, dtype=torch.float64, device="cuda")

    def factors(self, tb, tn):
        c = [self._z() for _ in range(3)]
        self.h.tra_adv_fct_coef(self.p2dt, self.zun, self.zvn, self.zwn, tb, tn, *c, kn=self.kn)
        return c

    def step(self, c, tb, tn):
        """one step with the limiter factors c: (tb, tn) -> (tb', tn')"""
        h, n3 = self.h, self.n3
        ta = self._z()
        h.tra_adv_fct_tl(self.p2dt, self.zun, self.zvn, self.zwn, tb, tn, *c, ta, kn=self.kn)
        h.tra_ldf_lap(self.kt, 1, "TRA", self.ahtu, self.ahtv, tb, ta, 1, 1)
        h.tra_zdf_imp(self.kt, 1, "TRA", self.p2dt, tb, ta, 1, self.avt, self.avt)
        B, N, A = self._z(2), self._z(2), self._z(2)
        B[:n3].copy_(tb)
        N[:n3].copy_(tn)
        A[:n3].copy_(ta)
        h.tra_nxt(self.kt, B, N, A)
        return B[:n3].clone(), N[:n3].clone()

    def step_ad(self, c, lb, ln):
        """the adjoint of step with the factors c frozen: (lb', ln') -> (lb, ln)"""
        h = self.h
        lb, ln = lb.contiguous().clone(), ln.contiguous().clone()
        la = self._z()
        h.tra_nxt_ad(lb, ln, la)
        h.tra_zdf_imp_ad(self.p2dt, self.avt, la, lb)
        h.tra_ldf_lap_ad(self.ahtu, self.ahtv, la, lb)
        h.tra_adv_fct_ad(self.zun, self.zvn, self.zwn, *c, la, lb, ln, kn=self.kn)
        return lb, ln


class TracerStep(torch.autograd.Function):
    """(tb, tn) -> (tb', tn') of one tracer step; backward: the adjoint with
    the forward step's limiter factors frozen."""

    @staticmethod
    def forward(ctx, tb, tn, model):
        cur = torch.cuda.current_stream()
        model.stream.wait_stream(cur)          # the inputs are ready
        with torch.cuda.stream(model.stream):
            c = model.factors(tb, tn)
            nb, nn = model.step(c, tb, tn)
        cur.wait_stream(model.stream)          # the outputs are ready
        for t in (nb, nn):
            t.record_stream(cur)
        ctx.model, ctx.c = model, c
        return nb, nn

    @staticmethod
    def backward(ctx, gb, gn):
        model = ctx.model
        cur = torch.cuda.current_stream()
        model.stream.wait_stream(cur)
        with torch.cuda.stream(model.stream):
            z = torch.zeros(model.n3, dtype=torch.float64, device="cuda")
            lb, ln = model.step_ad(ctx.c, z if gb is None else gb, z if gn is None else gn)
        cur.wait_stream(model.stream)
        for t in (lb, ln):
            t.record_stream(cur)
        return lb, ln, None


def _segment(model, steps):
    def f(tb, tn):
        for _ in range(steps):
            tb, tn = TracerStep.apply(tb, tn, model)
        return tb, tn
    return f


def run(model, tb, tn, steps, every=None):
    """`steps` tracer steps from (tb, tn), differentiable; with `every`, the
    states between checkpoints are recomputed in the backward pass"""
    if not every:
        return _segment(model, steps)(tb, tn)
    done = 0
    while done < steps:
        k = min(every, steps - done)
        tb, tn = tcp.checkpoint(_segment(model, k), tb, tn, use_reentrant=False)
        done += k
    return tb, tn
