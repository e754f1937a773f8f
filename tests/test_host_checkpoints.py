"""The particle checkpoints of g3d_autograd.BunchMarch, on the CPU: the device pushes are replaced by a toy
nonlinear step with a hand-written reverse (x' = x + ds sin(p) F, p' = p + ds cos(x) F, and z, rem and the alive
codes moving with the step), so that the segment bookkeeping (which state, z, rem and alive codes enter each reverse
step, how the field adjoint accumulates) is checked against torch's own autograd of the same step for every
checkpoint spacing.  No GPU."""
import torch

import g3d_autograd as ga


def toy_step(geo, F, state, z, rem, qm, alive):
    ds = geo.kw["s1"] - geo.kw["s0"]
    x, y, zeta, px, py, pz = state
    f = F.sum()
    out = [x + ds * torch.sin(px) * f * alive, y + ds * qm * py * rem, zeta * (1 + 0.1 * z), px + ds * torch.cos(x) * f,
           py * (1 + rem), pz + ds * x * y * f]
    return out, z + ds, rem * 0.5 + 0.1, alive


def toy_raw(geo, F, state, z, rem, qm, alive):
    with torch.no_grad():
        return toy_step(geo, F, state, z, rem, qm, alive)


def toy_vjp(geo, F, state, z, rem, qm, alive, bs, gF):
    Fv = F.detach().clone().requires_grad_()
    st = [t.detach().clone().requires_grad_() for t in state]
    with torch.enable_grad():
        out, _, _, _ = toy_step(geo, Fv, st, z, rem, qm, alive)
        g = torch.autograd.grad(out, [Fv] + st, bs)
    gF += g[0]
    return list(g[1:])


def test_checkpoint_bookkeeping(monkeypatch):
    monkeypatch.setattr(ga, "bunch_push_raw", toy_raw)
    monkeypatch.setattr(ga, "bunch_push_vjp_raw", toy_vjp)
    monkeypatch.setattr(ga, "_f64", lambda t, what: t.contiguous())      # CPU tensors for the toy
    S, n = 7, 5
    geos = [ga.PushGeometry(s0=0.1 * t, s1=0.1 * (t + 1)) for t in range(S)]
    gen = torch.Generator().manual_seed(3)
    F0 = torch.randn(4, dtype=torch.float64, generator=gen)
    st0 = [torch.randn(n, dtype=torch.float64, generator=gen) for _ in range(6)]
    z, rem, qm = torch.zeros(n, dtype=torch.float64), torch.rand(n, dtype=torch.float64, generator=gen), -torch.ones(
        n, dtype=torch.float64)
    alive = torch.tensor([1, 1, 0, 1, 1], dtype=torch.int32)
    cot = [torch.randn(n, dtype=torch.float64, generator=gen) for _ in range(6)]

    def leaves():
        return [F0.clone().requires_grad_()] + [t.clone().requires_grad_() for t in st0]

    ins = leaves()                                  # reference: torch's autograd through the toy steps
    st, zz, rr = ins[1:], z, rem
    for geo in geos:
        st, zz, rr, _ = toy_step(geo, ins[0], st, zz, rr, qm, alive)
    ref = torch.autograd.grad(sum((a * c).sum() for a, c in zip(st, cot)), ins)
    for every in (1, 2, 3, S, S + 1):
        ins = leaves()
        out = ga.BunchMarch.apply(geos, every, ins[0], *ins[1:], z, rem, qm, alive)
        for a, b in zip(out, st):
            assert torch.equal(a, b.detach())
        g = torch.autograd.grad(sum((a * c).sum() for a, c in zip(out, cot)), ins)
        for a, b in zip(g, ref):
            assert torch.allclose(a, b, rtol=1e-13, atol=1e-14), every


def test_second_backward_is_refused(monkeypatch):
    """the checkpoints are released by the first reverse sweep: a second backward through the same graph
    (retain_graph=True) raises a clear error instead of failing inside the sweep"""
    import pytest
    monkeypatch.setattr(ga, "bunch_push_raw", toy_raw)
    monkeypatch.setattr(ga, "bunch_push_vjp_raw", toy_vjp)
    monkeypatch.setattr(ga, "_f64", lambda t, what: t.contiguous())
    n = 3
    geos = [ga.PushGeometry(s0=0.1 * t, s1=0.1 * (t + 1)) for t in range(4)]
    F = torch.ones(4, dtype=torch.float64, requires_grad=True)
    st = [torch.full((n,), 0.5, dtype=torch.float64, requires_grad=True) for _ in range(6)]
    z, rem, qm = torch.zeros(n, dtype=torch.float64), torch.ones(n, dtype=torch.float64), -torch.ones(
        n, dtype=torch.float64)
    alive = torch.ones(n, dtype=torch.int32)
    out = ga.BunchMarch.apply(geos, 2, F, *st, z, rem, qm, alive)
    loss = sum(o.sum() for o in out)
    loss.backward(retain_graph=True)
    with pytest.raises(RuntimeError, match="second backward"):
        loss.backward()
