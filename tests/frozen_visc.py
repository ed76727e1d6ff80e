"""The entropy viscosity (smart_visc_spectral, mod_entropy_viscosity.f90:14-122)
in FP64 torch with its branch choices recorded at a base state and replayed
(test infrastructure for the adjoint's finite-difference checks).

smart_visc_spectral is piecewise smooth: every max/min over nodes, the
saturation min(Ve, beta_e), the limiters max(x, ce), the |Reta| and the
normalisation max |eta - mean| pick a branch.  The hand-written adjoint
differentiates the branch taken at the point where it is evaluated (first
maximiser, strict comparisons).  A central difference whose +-h states take
other branches measures a different map, so the checks evaluate the +-h
states along the branches recorded on the unperturbed trajectory: a smooth
map that equals smart_visc_spectral at the base point and whose derivative
there is what the adjoint returns.

Branches (per element unless noted): the node of max |u| (iv), of max
csound (ic) and of max beta (ib); tanh(M^15 pi) > ce; ceM > the Jacobian
limiter; per element node the sign of Reta and Ve < beta_e; globally the
working node of max |eta - mean| and the sign of eta - mean there.
"""
from __future__ import annotations

import numpy as np
import torch

from torch_rhs import filter_matrix

TD = torch.float64


class FrozenSmartVisc:
    def __init__(self, mesh, gpvol, wgp, helem, helem_l, invAtoIJK, gamma, Rgas, ce=0.1, cmax=0.5, cglob=1.0,
                 normalise=True, device="cuda"):
        dev = lambda a, dt=TD: torch.as_tensor(np.asarray(a), dtype=dt, device=device)
        self.c0 = dev(np.asarray(mesh.connec, dtype=np.int64) - 1, torch.long)          # (ne, 64) gmsh order
        self.lw = dev(np.asarray(mesh.lpoin_w, dtype=np.int64) - 1, torch.long)
        ne = self.c0.shape[0]
        J = dev(np.asarray(gpvol).reshape(64, ne).T) / dev(wgp)[None, :]
        rj = torch.clamp(J, max=1000000.0).min(dim=1).values / torch.clamp(J, min=0.0).max(dim=1).values
        self.ced_geo = torch.clamp(1.0 - rj * rj, min=ce)      # geometry only
        self.he = dev(helem)
        self.hl2 = dev(helem_l) ** 2
        self.F = filter_matrix(invAtoIJK).to(device)
        self.gamma, self.Rgas, self.ce, self.cmax, self.cglob = gamma, Rgas, ce, cmax, cglob
        self.normalise = normalise
        self.ar = torch.arange(ne, device=device)

    def __call__(self, rho, u, csound, Tem, eta, Reta, branches=None):
        """mu_e (ne, 64) and the branches used (recorded from this state when
        `branches` is None)."""
        c0, ar = self.c0, self.ar
        rec = branches is None
        b = {} if rec else branches
        if self.normalise:
            ew = eta[self.lw]
            dev = ew - ew.sum() / self.lw.shape[0]
            if rec:
                b["norm_idx"] = int(torch.argmax(torch.abs(dev)))
                b["norm_sign"] = 1.0 if float(dev[b["norm_idx"]]) >= 0.0 else -1.0
            norm = b["norm_sign"] * dev[b["norm_idx"]]
        else:
            norm = torch.tensor(1.0, dtype=TD, device=rho.device)
        ug = u[c0]                                     # (ne, 64, 3)
        s2 = (ug * ug).sum(-1)
        cs = csound[c0]
        rg, Tg = rho[c0], Tem[c0]
        if rec:
            vmag = torch.sqrt(s2)
            b["iv"] = torch.max(vmag, dim=1).indices
            b["ic"] = torch.max(cs, dim=1).indices
            beta = (rg * self.he[:, None]) * (self.cmax / 3.0) * (vmag + torch.sqrt(self.Rgas * self.gamma * Tg))
            b["ib"] = torch.max(beta, dim=1).indices
        iv, ic, ib = b["iv"], b["ic"], b["ib"]
        maxV = torch.sqrt(s2[ar, iv])
        maxC = cs[ar, ic]
        M = maxV / maxC
        th = torch.tanh(M ** 15 * np.pi)
        if rec:
            b["tanh"] = th > self.ce
        ceM = torch.where(b["tanh"], th, torch.full_like(th, self.ce))
        if rec:
            b["ceM"] = ceM > self.ced_geo
        ced = torch.where(b["ceM"], ceM, self.ced_geo)
        betae = (rg[ar, ib] * self.he) * (self.cmax / 3.0) * (
            torch.sqrt(s2[ar, ib]) + torch.sqrt(self.Rgas * self.gamma * Tg[ar, ib]))
        Rg = Reta[c0]
        if rec:
            b["reta_sign"] = torch.sign(Rg).to(torch.int8)
        R1 = rg * (b["reta_sign"].to(TD) * Rg) / norm
        Ve = ced[:, None] * R1 * self.hl2
        if rec:
            b["via_ve"] = Ve < betae[:, None]
        mue_l = self.cglob * torch.where(b["via_ve"], Ve, betae[:, None].expand_as(Ve))
        return mue_l @ self.F.T, b


def trajectory(ec, fz, y0, dts, branches=None):
    """The differentiable step's map (RK4 stages through the library's
    sod_full_rhs with mu_e frozen within the step, final EOS, entropy
    viscosity of the new state) with the entropy viscosity evaluated by `fz`:
    recording its branches (branches None) or replaying them.  Returns the
    final (rho, q, E, mu_e) and the branches of every step."""
    from sod2d_refmap_amd.autograd import final_eos, rk4_step
    rc = ec.rc
    out = []
    y = y0
    with torch.no_grad():
        for i, dt in enumerate(dts):
            r, q, E, Reta = rk4_step(rc, y[0], y[1], y[2], y[3], dt)
            u, c, T, eta = final_eos(rc, r, q, E)
            mu, b = fz(r, u, c, T, eta, Reta, None if branches is None else branches[i])
            out.append(b)
            y = (r, q, E, mu)
    return y, out
