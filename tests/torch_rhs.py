"""FP64 pure-torch transcription of the reference element routines on the
path (test infrastructure: the reference for the adjoint, by autograd).

  full_convec_ijk            elem_convec.f90:211-460
  generic_scalar_convec_ijk  elem_convec.f90:462-570
  full_diffusion_ijk         elem_diffu.f90:12-198
  stage EOS + Sutherland     time_integ.f90:266-316, mod_fluid_viscosity.f90:21-41

Element-node data are held on the element lattice X[e, i, j, k] in the
sod2d 1-D indices (gmsh node a = invAtoIJK(i,j,k)); derivatives are
contractions of the 4x4 matrix D[m, ii] = l'_ii(xi_m) along one lattice axis,
mapped with He as in the reference (products first, then the derivative).
Checked against the oracle in tests/test_torch_rhs.py.
"""
from __future__ import annotations

import numpy as np
import torch

TD = torch.float64


class Geometry:
    """Static element data of a mesh on the lattice, as torch tensors."""

    def __init__(self, connec, He, gpvol, tables, device="cpu"):
        inv = np.asarray(tables.invAtoIJK) - 1          # (4,4,4) [i,j,k] -> gmsh node
        self.lat = torch.as_tensor(inv.reshape(-1), device=device)   # i-major flattening of [i,j,k]
        ne = connec.shape[0]
        self.nelem = ne
        c0 = torch.as_tensor(np.asarray(connec, dtype=np.int64) - 1, device=device)   # (ne, 64) by gmsh node
        self.conn = c0[:, self.lat].reshape(ne, 4, 4, 4)                                # lattice order
        He_t = torch.as_tensor(np.asarray(He), dtype=TD, device=device)                  # (3,3,64,ne)
        He_e = He_t.permute(3, 2, 0, 1)[:, self.lat]                                     # (ne, 64, 3, 3)
        self.He = He_e.reshape(ne, 4, 4, 4, 3, 3)
        gv = torch.as_tensor(np.asarray(gpvol).reshape(64, ne), dtype=TD, device=device).T[:, self.lat]
        self.gv = gv.reshape(ne, 4, 4, 4)
        dl = np.asarray(tables.dlxigp_ip)               # (64, 3, 4): D[coord_d(g)][ii]
        # D[m, ii] from the xi-derivative rows of the nodes with i = m
        Dm = np.zeros((4, 4))
        I = np.asarray(tables.gmshAtoI) - 1
        for g in range(64):
            Dm[I[g]] = dl[g, 0]
        self.D = torch.as_tensor(Dm, dtype=TD, device=device)

    def gather(self, f):
        """nodal (npoin, ...) -> lattice (ne, 4,4,4, ...)"""
        return f[self.conn]

    def scatter(self, r, npoin):
        """lattice (ne,4,4,4, ...) -> nodal sum (npoin, ...)"""
        out = torch.zeros((npoin,) + tuple(r.shape[4:]), dtype=r.dtype, device=r.device)
        return out.index_add(0, self.conn.reshape(-1), r.reshape((-1,) + tuple(r.shape[4:])))

    def dxi(self, X):
        """xi-derivatives of a lattice field X[e,i,j,k,...] -> [..., 3] (xi, eta, zeta)"""
        D = self.D
        a = torch.einsum("mi,eijk...->emjk...", D, X)
        b = torch.einsum("mj,eijk...->eimk...", D, X)
        c = torch.einsum("mk,eijk...->eijm...", D, X)
        return torch.stack((a, b, c), dim=-1)

    def grad(self, X):
        """physical gradient of a scalar lattice field: sum_j He(i,j) dX/dxi_j"""
        return torch.einsum("eijkab,eijkb->eijka", self.He, self.dxi(X))

    def grad_v(self, V):
        """gradient of a vector lattice field V[...,c]: G[..., c, i] = d V_c / d x_i"""
        return torch.einsum("eijkab,eijkcb->eijkca", self.He, self.dxi(V))

    def div(self, V):
        g = self.grad_v(V)
        return g[..., 0, 0] + g[..., 1, 1] + g[..., 2, 2]

    def weak_div(self, F):
        """sum over the 3x4 line points q through each test node of
        w_q J_q dl_own/dxi(q) He(:,d,q) . F(q)  (elem_diffu.f90:142-194)"""
        G = torch.einsum("eijkad,eijk...a->eijk...d", self.He, F)
        G = self.gv.reshape(self.gv.shape + (1,) * (G.dim() - 4)) * G
        D = self.D
        a = torch.einsum("mi,emjk...->eijk...", D, G[..., 0])
        b = torch.einsum("mj,eimk...->eijk...", D, G[..., 1])
        c = torch.einsum("mk,eijm...->eijk...", D, G[..., 2])
        return a + b + c


def full_convec(geo, u, q, rho, pr, E):
    """Rmass, Rmom, Rener of full_convec_ijk (nodal inputs, nodal outputs)."""
    npoin = rho.shape[0]
    ul, ql, rl, pl, El = geo.gather(u), geo.gather(q), geo.gather(rho), geo.gather(pr), geo.gather(E)
    kl = 0.5 * (ul * ul).sum(-1)
    fl = ql[..., :, None] * ul[..., None, :]          # f_ij = q_i u_j
    fuul = ul[..., :, None] * ul[..., None, :]
    fel = (rl * El)[..., None] * ul
    fuel = El[..., None] * ul
    fkl = (rl * kl)[..., None] * ul
    fukl = kl[..., None] * ul
    REl, Rkl = rl * El, rl * kl
    gRho, gP, gE, gRE, gk, gRk = (geo.grad(x) for x in (rl, pl, El, REl, kl, Rkl))
    gU, gQ = geo.grad_v(ul), geo.grad_v(ql)
    divU = gU[..., 0, 0] + gU[..., 1, 1] + gU[..., 2, 2]
    divQ = gQ[..., 0, 0] + gQ[..., 1, 1] + gQ[..., 2, 2]
    divFe, divFue, divFk, divFuk = (geo.div(x) for x in (fel, fuel, fkl, fukl))
    gF = torch.einsum("eijkab,eijkcdb->eijkcda", geo.He, geo.dxi(fl))     # d f_cd / dx_a
    gFuu = torch.einsum("eijkab,eijkcdb->eijkcda", geo.He, geo.dxi(fuul))
    divF = torch.einsum("eijkcdd->eijkc", gF)
    divFuu = torch.einsum("eijkcdd->eijkc", gFuu)
    Rm = 0.5 * (divQ + rl * divU) + 0.5 * (ul * gRho).sum(-1)
    Re = (0.25 * (divFe + El * divQ + REl * divU + rl * divFue) +
          0.25 * (divFk + kl * divQ + Rkl * divU + rl * divFuk) +
          (0.25 * (ql * gE + ul * gRE + fuel * gRho) + 0.25 * (ql * gk + ul * gRk + fukl * gRho)).sum(-1))
    Rq = (0.25 * (divF + ul * divQ[..., None] + ql * divU[..., None] + rl[..., None] * divFuu) + gP +
          0.25 * (torch.einsum("eijkcd,eijkd->eijkc", fuul, gRho) + torch.einsum("eijkd,eijkcd->eijkc", ql, gU) +
                  torch.einsum("eijkd,eijkcd->eijkc", ul, gQ)))
    gv = geo.gv
    return (geo.scatter(gv * Rm, npoin), geo.scatter(gv[..., None] * Rq, npoin), geo.scatter(gv * Re, npoin))


def scalar_convec(geo, q, eta, u):
    npoin = eta.shape[0]
    ul, fl, el = geo.gather(u), geo.gather(q), geo.gather(eta)
    R = 0.5 * (geo.div(fl) + el * geo.div(ul)) + 0.5 * (ul * geo.grad(el)).sum(-1)
    return geo.scatter(geo.gv * R, npoin)


def full_diffusion(geo, Cp, Pr, rho_n, rho, u, Tem, mu_fluid, mu_e_lat, mu_sgs_lat, c_rho=1.0, c_ener=1.0):
    """mu_e_lat / mu_sgs_lat are element-node fields on the lattice (ne,4,4,4)."""
    npoin = rho.shape[0]
    rl, rnl, Tl, mfl, ul = (geo.gather(x) for x in (rho, rho_n, Tem, mu_fluid, u))
    mu_f = mfl + rnl * mu_sgs_lat
    mu_t = mu_f + mu_e_lat
    gR, gT, gU = geo.grad(rl), geo.grad(Tl), geo.grad_v(ul)
    divU = gU[..., 0, 0] + gU[..., 1, 1] + gU[..., 2, 2]
    eye = torch.eye(3, dtype=TD, device=rho.device)
    tau = mu_t[..., None, None] * (gU + gU.transpose(-1, -2)) - (mu_f * (2.0 / 3.0) * divU)[..., None, None] * eye
    tauU = torch.einsum("eijkcd,eijkd->eijkc", tau, ul)
    nu_e = c_rho * mu_e_lat / rnl
    kappa = mfl * Cp / Pr + c_ener * mu_e_lat * Cp / Pr + Cp * rnl * mu_sgs_lat / 0.9
    dR = geo.weak_div(gR)
    dM = geo.weak_div(tau)             # row c: sum_a w . tau[c, a]
    dU = geo.weak_div(tauU)
    dT = geo.weak_div(gT)
    Rm = nu_e * dR
    Re = dU + kappa * dT
    return geo.scatter(Rm, npoin), geo.scatter(dM, npoin), geo.scatter(Re, npoin)


def stage_eos(rho, q, E, mu_factor, gamma, Rgas):
    u = q / rho[:, None]
    e = E / rho - 0.5 * (u * u).sum(-1)
    p = rho * (gamma - 1.0) * e
    h = (gamma / (gamma - 1.0)) * p / rho
    T = p / (rho * Rgas)
    eta = (rho / (gamma - 1.0)) * torch.log(p / rho ** gamma)
    mu = mu_factor * 0.000001458 * T ** 1.5 / (T + 110.4)
    return u, p, h, T, eta, mu


def full_rhs(geo, rho, q, E, mu_factor, mu_e_lat, mu_sgs_lat, gamma, Rgas, Cp, Prt, c_rho=1.0, c_ener=1.0):
    """The assembled RHS of one RK stage (what sod_full_rhs returns)."""
    u, p, h, T, eta, mu = stage_eos(rho, q, E, mu_factor, gamma, Rgas)
    cm, cq, ce = full_convec(geo, u, q, rho, p, h)
    dm, dq, de = full_diffusion(geo, Cp, Prt, rho, rho, u, T, mu, mu_e_lat, mu_sgs_lat, c_rho, c_ener)
    rt = scalar_convec(geo, u * eta[:, None], eta, u)
    return cm + dm, cq + dq, ce + de, rt


def lattice_from_elemnode(geo, a):
    """(nelem, 64) element-node array in gmsh node order -> lattice"""
    t = torch.as_tensor(np.asarray(a), dtype=TD)
    return t[:, geo.lat.cpu()].reshape(geo.nelem, 4, 4, 4).to(geo.D.device)


def filter_matrix(invAtoIJK):
    """27-point (1/4,1/2,1/4)^3 mirror filter of smart_visc_spectral
    (mod_entropy_viscosity.f90:105-121, convertIJK of mod_filters.f90:42-51)
    as a 64x64 matrix over gmsh node order: mu_e = F @ mue_l."""
    inv = np.asarray(invAtoIJK) - 1
    conv = [3, 1, 3, 4, 2, 3]
    w = [0.25, 0.5, 0.25]
    F = np.zeros((64, 64))
    for ii in range(1, 5):
        for jj in range(1, 5):
            for kk in range(1, 5):
                out = inv[conv[ii] - 1, conv[jj] - 1, conv[kk] - 1]
                for l in (-1, 0, 1):
                    for m in (-1, 0, 1):
                        for n in (-1, 0, 1):
                            a = inv[conv[ii + l] - 1, conv[jj + m] - 1, conv[kk + n] - 1]
                            F[out, a] += w[l + 1] * w[m + 1] * w[n + 1]
    return torch.as_tensor(F, dtype=TD)


def smart_visc(connec, lpoin_w, Reta, gpvol, wgp, gamma, rho, u, csound, Tem, eta, helem_l, helem, invAtoIJK,
               Rgas, ce=0.1, cmax=0.5, cglob=1.0, normalise=True):
    """smart_visc_spectral (mod_entropy_viscosity.f90:14-122) on torch tensors;
    max/min are taken with first-index subgradients (torch.max over a dim)."""
    c0 = torch.as_tensor(np.asarray(connec, dtype=np.int64) - 1)
    lw = torch.as_tensor(np.asarray(lpoin_w, dtype=np.int64) - 1)
    ne = c0.shape[0]
    if normalise:
        ew = eta[lw]
        mean = ew.sum() / lw.shape[0]
        norm, _ = torch.max(torch.abs(ew - mean), dim=0)
    else:
        norm = torch.tensor(1.0, dtype=TD)
    gv = torch.as_tensor(np.asarray(gpvol).reshape(64, ne).T, dtype=TD)
    J = gv / torch.as_tensor(np.asarray(wgp), dtype=TD)[None, :]
    minJe = torch.clamp(J, max=1000000.0).min(dim=1).values
    maxJe = torch.clamp(J, min=0.0).max(dim=1).values
    ug = u[c0]
    s2 = (ug * ug).sum(-1)
    vmag = torch.where(s2 > 0, torch.sqrt(torch.where(s2 > 0, s2, torch.ones_like(s2))), torch.zeros_like(s2))
    maxV, _ = torch.max(vmag, dim=1)
    maxC, _ = torch.max(csound[c0], dim=1)
    M = maxV / maxC
    th = torch.tanh(M ** 15 * np.pi)
    ceM = torch.where(th > ce, th, torch.full_like(th, ce))
    rj = minJe / maxJe
    ced_geo = torch.clamp(1.0 - rj * rj, min=ce)
    ced = torch.where(ceM > ced_geo, ceM, ced_geo)
    aux3 = torch.sqrt(Rgas * gamma * Tem[c0])
    he = torch.as_tensor(np.asarray(helem), dtype=TD)
    beta = (rho[c0] * he[:, None]) * (cmax / 3.0) * (vmag + aux3)
    betae, _ = torch.max(beta, dim=1)
    R1 = rho[c0] * torch.abs(Reta[c0]) / norm
    hl = torch.as_tensor(np.asarray(helem_l), dtype=TD)
    Ve = ced[:, None] * R1 * (hl * hl)
    mue_l = cglob * torch.where(Ve < betae[:, None], Ve, betae[:, None].expand_as(Ve))
    F = filter_matrix(invAtoIJK)
    return mue_l @ F.T, mue_l
