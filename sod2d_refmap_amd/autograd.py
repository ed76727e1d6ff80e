"""torch.autograd binding of the assembled stage right-hand side
(sod_full_rhs) and of its hand-written adjoint (sod_full_rhs_vjp), and an RK4
time integration built on it whose gradient is taken with checkpointing.

Both directions run in libsod2d_hip.so; torch only holds the device arrays
and records the graph.  Nodal fields are torch tensors of shape (npoin,) or
(npoin, 3); element-node fields (nelem, 64) follow the reference's
element-fastest layout.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .device import to_device

TD = torch.float64
RK_A = (0.0, 0.5, 0.5, 1.0)                  # time_integ.f90:207-210
RK_B = (1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0)


def _fcol(x: torch.Tensor) -> torch.Tensor:
    """(n, c) tensor -> flat column-major storage (the reference layout)"""
    return x.t().contiguous().reshape(-1) if x.dim() == 2 else x.contiguous()


def _unfcol(flat: torch.Tensor, n: int, c: int) -> torch.Tensor:
    return flat.reshape(c, n).t()


@dataclass
class RHSContext:
    """Static data of one mesh: device copies of the reference arrays and the
    gas constants the stage RHS needs."""
    nelem: int
    npoin: int
    npoin_w: int
    connec: torch.Tensor
    lpoin_w: torch.Tensor
    He: torch.Tensor
    gpvol: torch.Tensor
    Ml: torch.Tensor
    mu_factor: torch.Tensor
    mu_sgs: torch.Tensor
    Rgas: float
    gamma: float
    Cp: float
    Prt: float
    working: torch.Tensor   # bool (npoin,): lpoin_w nodes

    @staticmethod
    def from_mesh(mesh, He, gpvol, Ml, mu_factor, gas, mu_sgs=None) -> "RHSContext":
        dev = lambda a: to_device(np.asarray(a))
        w = torch.zeros(mesh.npoin, dtype=torch.bool, device="cuda")
        w[torch.as_tensor(np.asarray(mesh.lpoin_w) - 1, device="cuda").long()] = True
        sgs = np.zeros((mesh.nelem, 64), order="F") if mu_sgs is None else mu_sgs
        return RHSContext(mesh.nelem, mesh.npoin, mesh.npoin_w, dev(mesh.connec), dev(mesh.lpoin_w),
                          He if torch.is_tensor(He) else dev(He), gpvol if torch.is_tensor(gpvol) else dev(gpvol),
                          Ml if torch.is_tensor(Ml) else dev(Ml), dev(mu_factor), dev(sgs), gas.Rgas,
                          gas.gamma_gas, gas.Cp, gas.Prt, w)

    def static_ptrs(self):
        return [self.connec.data_ptr(), self.lpoin_w.data_ptr(), self.He.data_ptr(), self.gpvol.data_ptr(),
                self.Ml.data_ptr()]


class FullRHS(torch.autograd.Function):
    """(rho, q, E, mu_e) -> (Rmass, Rmom, Rener, Reta) of one RK stage
    (time_integ.f90:266-373, assembled, not lumped-solved); the backward is
    the hand-written adjoint kernel."""

    @staticmethod
    def forward(ctx, rho, q, E, mu_e, rc: RHSContext):
        n, ne = rc.npoin, rc.nelem
        r, qf, e = rho.contiguous(), _fcol(q), E.contiguous()
        me = _fcol(mu_e)
        out = [torch.empty(n, dtype=TD, device=rho.device) for _ in range(4)]
        rq = torch.empty(3 * n, dtype=TD, device=rho.device)
        _lib.call("sod_full_rhs", ne, n, rc.npoin_w, *rc.static_ptrs(), r.data_ptr(), qf.data_ptr(), e.data_ptr(),
                  rc.mu_factor.data_ptr(), me.data_ptr(), rc.mu_sgs.data_ptr(), rc.Rgas, rc.gamma, rc.Cp, rc.Prt,
                  out[0].data_ptr(), rq.data_ptr(), out[2].data_ptr(), out[3].data_ptr())
        ctx.save_for_backward(r, qf, e, me)
        ctx.rc = rc
        return out[0], _unfcol(rq, n, 3), out[2], out[3]

    @staticmethod
    def backward(ctx, gm, gq, ge, gt):
        r, qf, e, me = ctx.saved_tensors
        rc = ctx.rc
        n, ne = rc.npoin, rc.nelem
        z = lambda k: torch.zeros(k, dtype=TD, device=r.device)
        lm = gm.contiguous() if gm is not None else z(n)
        lq = _fcol(gq) if gq is not None else z(3 * n)
        le = ge.contiguous() if ge is not None else z(n)
        lt = gt.contiguous() if gt is not None else z(n)
        g_r, g_q, g_e, g_me = z(n), z(3 * n), z(n), z(64 * ne)
        _lib.call("sod_full_rhs_vjp", ne, n, rc.npoin_w, *rc.static_ptrs(), r.data_ptr(), qf.data_ptr(),
                  e.data_ptr(), rc.mu_factor.data_ptr(), me.data_ptr(), rc.mu_sgs.data_ptr(), rc.Rgas, rc.gamma,
                  rc.Cp, rc.Prt, lm.data_ptr(), lq.data_ptr(), le.data_ptr(), lt.data_ptr(), g_r.data_ptr(),
                  g_q.data_ptr(), g_e.data_ptr(), g_me.data_ptr())
        return g_r, _unfcol(g_q, n, 3), g_e, _unfcol(g_me, ne, 64), None


def full_rhs(rho, q, E, mu_e, rc: RHSContext):
    return FullRHS.apply(rho, q, E, mu_e, rc)


def rk4_step(rc: RHSContext, rho, q, E, mu_e, dt: float):
    """One classical RK4 step of rk_4_main (time_integ.f90:234-426) with the
    entropy viscosity mu_e held fixed within the step, as the reference does.
    Residuals are lumped-solved on the working nodes (mod_solver.f90:23-58);
    other nodes keep their values.  Returns the new (rho, q, E) and the
    entropy residual Reta = -sum_i b_i Ml^-1 R_eta,i (time_integ.f90:467-471)."""
    w = rc.working
    Ml = torch.where(w, rc.Ml, torch.ones_like(rc.Ml))
    zero = torch.zeros_like(rc.Ml)
    y = (rho, q, E)
    acc = None
    Rprev = None
    for s in range(4):
        if s == 0:
            ys = y
        else:
            f = dt * RK_A[s]
            ys = tuple(yy - f * rr for yy, rr in zip(y, Rprev))
        Rm, Rq, Re, Rt = full_rhs(ys[0], ys[1], ys[2], mu_e, rc)
        R = (torch.where(w, Rm / Ml, zero), torch.where(w[:, None], Rq / Ml[:, None], zero[:, None]),
             torch.where(w, Re / Ml, zero), torch.where(w, Rt / Ml, zero))
        acc = tuple(RK_B[s] * rr for rr in R) if acc is None else tuple(a + RK_B[s] * rr for a, rr in zip(acc, R))
        Rprev = R[:3]
    new = tuple(yy - dt * a for yy, a in zip(y, acc[:3]))
    return new[0], new[1], new[2], -acc[3]


@dataclass
class EVContext:
    """Static data of the entropy viscosity (smart_visc_spectral)."""
    rc: RHSContext
    helem: torch.Tensor
    helem_l: torch.Tensor
    tables: dict
    coord: torch.Tensor

    @staticmethod
    def create(rc: RHSContext, mesh, helem, helem_l, tables) -> "EVContext":
        dev = lambda a: a if torch.is_tensor(a) else to_device(np.asarray(a))
        tab = {k: dev(getattr(tables, k)) for k in ("Ngp", "dNgp", "wgp", "invAtoIJK", "gmshAtoI", "gmshAtoJ",
                                                     "gmshAtoK")}
        return EVContext(rc, dev(helem), dev(helem_l), tab, dev(mesh.coord))


class SmartVisc(torch.autograd.Function):
    """(rho, u, csound, Tem, eta, Reta) -> mu_e(nelem, 64) of
    smart_visc_spectral (mod_entropy_viscosity.f90:14-122); the backward is
    the hand-written adjoint kernel (limiters as subgradients)."""

    @staticmethod
    def forward(ctx, rho, u, csound, Tem, eta, Reta, ec: EVContext):
        rc = ec.rc
        n, ne = rc.npoin, rc.nelem
        T = ec.tables
        args = [rho.contiguous(), _fcol(u), csound.contiguous(), Tem.contiguous(), eta.contiguous(),
                Reta.contiguous()]
        mu_e = torch.empty(64 * ne, dtype=TD, device=rho.device)
        mue_l = torch.empty(64 * ne, dtype=TD, device=rho.device)
        r, uf, cs, tm, et, rt = args
        _lib.call("sod_smart_visc_spectral", ne, n, rc.npoin_w, rc.connec.data_ptr(), rc.lpoin_w.data_ptr(),
                  rt.data_ptr(), T["Ngp"].data_ptr(), ec.coord.data_ptr(), T["dNgp"].data_ptr(), rc.gpvol.data_ptr(),
                  T["wgp"].data_ptr(), rc.gamma, r.data_ptr(), uf.data_ptr(), cs.data_ptr(), tm.data_ptr(),
                  et.data_ptr(), ec.helem_l.data_ptr(), ec.helem.data_ptr(), rc.Ml.data_ptr(), mu_e.data_ptr(),
                  T["invAtoIJK"].data_ptr(), T["gmshAtoI"].data_ptr(), T["gmshAtoJ"].data_ptr(),
                  T["gmshAtoK"].data_ptr(), mue_l.data_ptr())
        ctx.save_for_backward(*args)
        ctx.ec = ec
        return _unfcol(mu_e, ne, 64)

    @staticmethod
    def backward(ctx, g_mu):
        r, uf, cs, tm, et, rt = ctx.saved_tensors
        ec = ctx.ec
        rc = ec.rc
        n, ne = rc.npoin, rc.nelem
        z = lambda k: torch.zeros(k, dtype=TD, device=r.device)
        lam = _fcol(g_mu)
        gr, gu, gc, gT, ge, gR = z(n), z(3 * n), z(n), z(n), z(n), z(n)
        _lib.call("sod_smart_visc_spectral_vjp", ne, n, rc.npoin_w, rc.connec.data_ptr(), rc.lpoin_w.data_ptr(),
                  rt.data_ptr(), rc.gpvol.data_ptr(), rc.gamma, r.data_ptr(), uf.data_ptr(), cs.data_ptr(),
                  tm.data_ptr(), et.data_ptr(), ec.helem_l.data_ptr(), ec.helem.data_ptr(), lam.data_ptr(),
                  gr.data_ptr(), gu.data_ptr(), gc.data_ptr(), gT.data_ptr(), ge.data_ptr(), gR.data_ptr())
        return gr, _unfcol(gu, n, 3), gc, gT, ge, gR, None


def final_eos(rc: RHSContext, rho, q, E):
    """Velocity, sound speed, temperature and entropy after the last stage
    (time_integ.f90:445-463)."""
    g = rc.gamma
    u = q / rho[:, None]
    e = E / rho - 0.5 * (u * u).sum(-1)
    p = rho * (g - 1.0) * e
    c = torch.sqrt(g * p / rho)
    T = p / (rho * rc.Rgas)
    eta = (rho / (g - 1.0)) * torch.log(p / rho ** g)
    return u, c, T, eta


def rk4_step_ev(ec: EVContext, rho, q, E, mu_e, dt: float):
    """rk_4_main as a differentiable map (rho, q, E, mu_e) -> (rho, q, E,
    mu_e) of the next step: the RK4 stages with mu_e frozen, the final EOS,
    and the entropy viscosity of the new state (time_integ.f90:492-493)."""
    r, qq, e, Reta = rk4_step(ec.rc, rho, q, E, mu_e, dt)
    u, c, T, eta = final_eos(ec.rc, r, qq, e)
    return r, qq, e, SmartVisc.apply(r, u, c, T, eta, Reta, ec)


def kinetic_energy(rc: RHSContext, rho, q, rho0: float = 1.0):
    """volAvg_EK (mod_analysis.f90:10-44) of a conserved state: with the
    collocated GLL quadrature the element sum is sum_n Ml(n) rho|u|^2/2, over
    rho0 (2 * 3.14159)^3 (the reference's approximate pi)."""
    ke = 0.5 * (q * q).sum(-1) / rho
    return (rc.Ml * ke).sum() / (rho0 * (2.0 * 3.14159) ** 3)


def run_checkpointed(ec: EVContext, rho, q, E, mu_e, dts, every: int = 20, objective=None):
    """Integrates len(dts) steps of rk4_step_ev from (rho, q, E, mu_e) and
    returns objective(final rho, q, E) (default: the kinetic energy) as a
    tensor whose backward recomputes each segment of `every` steps from its
    stored starting state, so that the graph of only one segment is alive at
    a time (RK4 checkpointing)."""
    rc = ec.rc
    obj = objective or (lambda r, qq, e: kinetic_energy(rc, r, qq))
    return _Checkpointed.apply(rho, q, E, mu_e, ec, list(dts), int(every), obj)


class _Checkpointed(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, q, E, mu_e, ec, dts, every, obj):
        states = []
        y = (rho.detach(), q.detach(), E.detach(), mu_e.detach())
        with torch.no_grad():
            for i, dt in enumerate(dts):
                if i % every == 0:
                    states.append(tuple(t.clone() for t in y))
                y = rk4_step_ev(ec, *y, dt)
        ctx.ec, ctx.dts, ctx.every, ctx.states = ec, dts, every, states
        with torch.enable_grad():
            yf = tuple(t.detach().requires_grad_(True) for t in y[:3])
            J = obj(*yf)
            lam = torch.autograd.grad(J, yf, allow_unused=True)
        ctx.lam = tuple(torch.zeros_like(t) if g is None else g for t, g in zip(yf, lam))
        return J.detach()

    @staticmethod
    def backward(ctx, gJ):
        lam = tuple(gJ * l for l in ctx.lam) + (None,)
        for sidx in reversed(range(len(ctx.states))):
            start = sidx * ctx.every
            stop = min(start + ctx.every, len(ctx.dts))
            with torch.enable_grad():
                y0 = tuple(t.detach().requires_grad_(True) for t in ctx.states[sidx])
                y = y0
                for i in range(start, stop):
                    y = rk4_step_ev(ctx.ec, *y, ctx.dts[i])
                outs = [t for t, g in zip(y, lam) if g is not None]
                grads = [g for g in lam if g is not None]
                lam = torch.autograd.grad(outs, y0, grad_outputs=grads, allow_unused=True)
                lam = tuple(torch.zeros_like(t) if g is None else g for t, g in zip(y0, lam))
        return lam[0], lam[1], lam[2], lam[3], None, None, None, None
