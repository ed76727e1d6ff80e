"""Flow cases: the compressible Taylor-Green vortex of the reference and the
synthetic perturbations used by the randomised and ensemble workloads.

TGV constants and initial condition follow TGVSolver.f90:87-103 and
TGVSolver.f90:119-185.  The reference reads node coordinates into real(4)
variables and evaluates the velocity in single precision
(TGVSolver.f90:121-123, 137-139); `ref_fp32=True` reproduces that.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mesh import Mesh, hex64_tables


@dataclass
class GasParams:
    Cp: float = 1004.0
    Prt: float = 0.71
    gamma_gas: float = 1.4
    Mach: float = 0.1
    Re: float = 1600.0
    rho0: float = 1.0
    V0: float = 1.0
    L: float = 1.0

    # TGVSolver.f90:97-103
    @property
    def mul(self) -> float:
        return (self.rho0 * 1.0 * 1.0) / self.Re

    @property
    def Rgas(self) -> float:
        return self.Cp * (self.gamma_gas - 1.0) / self.gamma_gas

    @property
    def to(self) -> float:
        return 1.0 * 1.0 / (self.gamma_gas * self.Rgas * self.Mach * self.Mach)

    @property
    def po(self) -> float:
        return self.rho0 * self.Rgas * self.to

    @property
    def mu_factor(self) -> float:
        mur = 0.000001458 * (self.to ** 1.50) / (self.to + 110.40)
        return self.mul / mur


def sutherland(Tem, mu_factor):
    """mod_fluid_viscosity.f90:35-39 (host-side, initial viscosity only)."""
    return mu_factor * 0.000001458 * (Tem ** 1.5) / (Tem + 110.4)


def fourier_velocity_perturbation(mesh: Mesh, seed: int, amp: float = 1e-2, kmax: float = 8.0):
    """Real part of sum_{0<|k|<=kmax} a_k exp(i k.x) for each velocity component,
    a_k = amp * (N(0,1) + i N(0,1)), evaluated at the unperturbed lattice
    positions (exactly periodic).  Returns (npoin,3)."""
    rng = np.random.default_rng(seed)
    km = int(np.floor(kmax))
    ks = np.arange(-km, km + 1)
    K1, K2, K3 = np.meshgrid(ks, ks, ks, indexing="ij")
    mask = (K1 ** 2 + K2 ** 2 + K3 ** 2 <= kmax ** 2) & ~((K1 == 0) & (K2 == 0) & (K3 == 0))
    nx, ny, nz = mesh.n
    Lx, Ly, Lz = mesh.L
    lat = mesh.node_lattice
    r = hex64_tables().gll_roots
    # unperturbed coordinate of each lattice index along each axis
    axes = []
    for d, (nel, Ld) in enumerate(((nx, Lx), (ny, Ly), (nz, Lz))):
        idx = np.arange(3 * nel + 1)
        h = Ld / nel
        e = np.minimum(idx // 3, nel - 1)
        xs = e * h + (1.0 + r[idx - 3 * e]) * (0.5 * h)
        axes.append(np.exp(1j * np.outer(ks, 2 * np.pi * xs / Ld)))   # (2km+1, M)
    out = np.zeros((mesh.npoin, 3))
    for c in range(3):
        a = (rng.standard_normal(K1.shape) + 1j * rng.standard_normal(K1.shape)) * amp
        a = np.where(mask, a, 0.0)
        # separable evaluation on the lattice: u[I,J,K] = sum a[k1,k2,k3] X[k1,I] Y[k2,J] Z[k3,K]
        t1 = np.einsum("abc,aI->bcI", a, axes[0])
        t2 = np.einsum("bcI,bJ->cIJ", t1, axes[1])
        field = np.einsum("cIJ,cK->IJK", t2, axes[2]).real
        out[:, c] = field[lat[:, 0], lat[:, 1], lat[:, 2]]
    return out


def fourier_velocity_perturbation_fast(mesh: Mesh, seed: int, amp: float = 1e-2, kmax: float = 8.0):
    """fourier_velocity_perturbation (the same modes, amplitudes and random
    draws) with its three separable sums as matrix products: equal to it up to
    rounding, and seconds instead of minutes on 96^3 boxes.  Returns (npoin,3)."""
    rng = np.random.default_rng(seed)
    km = int(np.floor(kmax))
    ks = np.arange(-km, km + 1)
    K1, K2, K3 = np.meshgrid(ks, ks, ks, indexing="ij")
    mask = (K1 ** 2 + K2 ** 2 + K3 ** 2 <= kmax ** 2) & ~((K1 == 0) & (K2 == 0) & (K3 == 0))
    r = hex64_tables().gll_roots
    axes = []
    for nel, Ld in zip(mesh.n, mesh.L):
        idx = np.arange(3 * nel + 1)
        h = Ld / nel
        e = np.minimum(idx // 3, nel - 1)
        xs = e * h + (1.0 + r[idx - 3 * e]) * (0.5 * h)
        axes.append(np.exp(1j * np.outer(ks, 2 * np.pi * xs / Ld)))
    nk, lat = len(ks), mesh.node_lattice
    M = [ax.shape[1] for ax in axes]
    out = np.zeros((mesh.npoin, 3))
    for c in range(3):
        a = (rng.standard_normal(K1.shape) + 1j * rng.standard_normal(K1.shape)) * amp
        a = np.where(mask, a, 0.0)
        t1 = a.reshape(nk, nk * nk).T @ axes[0]                                  # (k2 k3, I)
        t2 = t1.reshape(nk, nk * M[0]).T @ axes[1]                               # (k3 I, J)
        f = (t2.reshape(nk, M[0] * M[1]).T @ axes[2]).real.reshape(M[0], M[1], M[2])
        out[:, c] = f[lat[:, 0], lat[:, 1], lat[:, 2]]
    return out


def placeholder_state(mesh: Mesh, gas: GasParams):
    """A state of the shapes tgv_initial_state returns, for constructing a
    solver whose level-2 arrays tgv_state_on_device then fills."""
    npn, ne = mesh.npoin, mesh.nelem
    one, zero = np.ones(npn), np.zeros(npn)
    z3, ze = np.zeros((npn, 3), order="F"), np.zeros((ne, 64), order="F")
    return dict(rho=one, u=z3, q=z3, pr=one, E=one, Tem=one, e_int=one, eta=zero, csound=one, machno=zero,
                mu_fluid=zero, mu_factor=np.full(npn, gas.mu_factor), mu_e=ze, mu_sgs=ze, mue_l=ze)


def tgv_state_on_device(sol, mesh: Mesh, gas: GasParams, modes_seed: int | None = None, amp: float = 1e-2):
    """tgv_initial_state (ref_fp32, plus fourier_velocity_perturbation_fast
    when modes_seed is given) evaluated on the device from a PeriodicSolver's
    coordinates and written into its level-2 arrays, csound and mu_fluid;
    the host version takes minutes at 96^3."""
    import torch
    npn = mesh.npoin
    xyz = sol.coord.view(3, npn)
    x, y, z = (xyz[d].float() for d in range(3))          # coordinates read as real(4) (TGVSolver.f90:121-123)
    V0, L = float(np.float32(gas.V0)), float(np.float32(gas.L))
    u = torch.zeros(3, npn, dtype=torch.float64, device=xyz.device)
    u[0] = (V0 * torch.sin(x / L) * torch.cos(y / L) * torch.cos(z / L)).double()
    u[1] = (-V0 * torch.cos(x / L) * torch.sin(y / L) * torch.cos(z / L)).double()
    xd, yd, zd = x.double(), y.double(), z.double()
    Lf = float(gas.L)
    pr = gas.po + ((gas.V0 * gas.V0) / 16.0) * (torch.cos(2.0 * xd / Lf) + torch.cos(2.0 * yd / Lf)) * (
        torch.cos(2.0 * zd / Lf) + 2.0)
    if modes_seed is not None:
        pert = fourier_velocity_perturbation_fast(mesh, modes_seed, amp)
        u += torch.from_numpy(np.ascontiguousarray(pert.T)).to(xyz.device)
    rho = pr / gas.Rgas / gas.to
    gm = gas.gamma_gas
    e_int = pr / (rho * (gm - 1.0))
    Tem = pr / (rho * gas.Rgas)
    E = rho * (0.5 * (u * u).sum(0) + e_int)
    cs = torch.sqrt(gm * pr / rho)
    eta = (rho / (gm - 1.0)) * torch.log(pr / rho ** gm)
    mu = sol.mu_factor * 0.000001458 * Tem ** 1.5 / (Tem + 110.4)
    sol.level("rho").copy_(rho)
    sol.level("u").copy_(u.reshape(-1))
    sol.level("q").copy_((rho[None, :] * u).reshape(-1))
    sol.level("pr").copy_(pr)
    sol.level("E").copy_(E)
    sol.level("Tem").copy_(Tem)
    sol.level("e_int").copy_(e_int)
    sol.level("eta").copy_(eta)
    sol.csound.copy_(cs)
    sol.mu_fluid.copy_(mu)
    torch.cuda.synchronize()


def tgv_initial_state(mesh: Mesh, gas: GasParams | None = None, ref_fp32: bool = True,
                      perturb_seed: int | None = None, perturb_amp: float = 1e-2):
    """Level-2 state of TGVSolver_evalInitialConditions (TGVSolver.f90:119-185)
    plus the initial Sutherland viscosity (CFDSolverBase.f90:1059-1079).

    Returns a dict of Fortran-ordered arrays: rho, u (npoin,3), q, pr, E, Tem,
    e_int, eta, csound, machno, mu_fluid, mu_factor (npoin), mu_e, mu_sgs,
    mue_l (nelem,64)."""
    g = gas or GasParams()
    xyz = mesh.coord
    if ref_fp32:
        x, y, z = (xyz[:, d].astype(np.float32) for d in range(3))
        V0, L = np.float32(g.V0), np.float32(g.L)
        u0 = (V0 * np.sin(x / L) * np.cos(y / L) * np.cos(z / L)).astype(np.float64)
        u1 = (-V0 * np.cos(x / L) * np.sin(y / L) * np.cos(z / L)).astype(np.float64)
        xd, yd, zd = (v.astype(np.float64) for v in (x, y, z))
    else:
        xd, yd, zd = xyz[:, 0], xyz[:, 1], xyz[:, 2]
        u0 = g.V0 * np.sin(xd / g.L) * np.cos(yd / g.L) * np.cos(zd / g.L)
        u1 = -g.V0 * np.cos(xd / g.L) * np.sin(yd / g.L) * np.cos(zd / g.L)
    u2 = np.zeros_like(u0)
    Lf = float(g.L)
    pr = g.po + ((1.0 * g.V0 * g.V0) / 16.0) * (np.cos(2.0 * xd / Lf) + np.cos(2.0 * yd / Lf)) * (
        np.cos(2.0 * zd / Lf) + 2.0)
    rho = pr / g.Rgas / g.to
    u = np.stack([u0, u1, u2], axis=1)
    if perturb_seed is not None:
        u = u + fourier_velocity_perturbation(mesh, perturb_seed, perturb_amp)
    gm = g.gamma_gas
    e_int = pr / (rho * (gm - 1.0))
    Tem = pr / (rho * g.Rgas)
    E = rho * (0.5 * np.sum(u * u, axis=1) + e_int)
    q = rho[:, None] * u
    csound = np.sqrt(gm * pr / rho)
    eta = (rho / (gm - 1.0)) * np.log(pr / (rho ** gm))
    machno = np.sum(u * u, axis=1) / csound
    mu_factor = np.full(mesh.npoin, g.mu_factor)
    mu_fluid = sutherland(Tem, mu_factor)
    F = np.asfortranarray
    return dict(
        rho=F(rho), u=F(u), q=F(q), pr=F(pr), E=F(E), Tem=F(Tem), e_int=F(e_int), eta=F(eta),
        csound=F(csound), machno=F(machno), mu_fluid=F(mu_fluid), mu_factor=F(mu_factor),
        mu_e=np.zeros((mesh.nelem, 64), order="F"), mu_sgs=np.zeros((mesh.nelem, 64), order="F"),
        mue_l=np.zeros((mesh.nelem, 64), order="F"))
