"""Discrete adjoint of the RK3 trajectory on one periodic box, behind a
``torch.autograd.Function``.

``rollout(U0, field, dts, checkpoint=...)`` advances the conserved state
``U0`` (a float64 CUDA tensor of shape (5, sz, sy, sx), the ghosted layout)
by ``len(dts)`` RK3 steps with the given fixed time steps and returns the
final state; ``backward`` runs the hand-written discrete adjoint of each step
(``DeviceField.adjoint_step``) in reverse order, recomputing the states it
needs from checkpoints:

- ``"all"``: every state kept;
- ``"uniform:K"``: every K-th state kept; each segment is recomputed once,
  its K states and (keep_stages, the default) their RK3 stage states kept
  during its reverse sweep, so that the adjoint steps recompute nothing;
- ``"binomial:S"``: S snapshots placed by the binomial (revolve) schedule,
  for a fixed memory budget.

The time steps are inputs, not functions of the state: the gradient is that
of the map U0 -> U_N at the given dt sequence (for a CFL-controlled run,
record the dt sequence of the forward run and pass it).  Gradients reach
interior cells only; ghost entries of the input are ignored by the forward
(periodic ghosts are refilled) and get zero gradient.

``time_integral(U0, field, dts, density, checkpoint=...)`` is the scalar
J = sum_n dt_n density(U_{n+1}) over the same trajectory (the right end
point of every step), e.g. the time-integrated viscous dissipation with
``density = dissipation_rate``.  Its backward adds the source
dt_{n-1} d density / dU_n to the adjoint after each reverse step; the
density's own gradient comes from torch autograd on the state tensor.
The forward may run the mixed FP32/FP64 method (the field's precision); the
adjoint is then the FP64 linearisation about that trajectory.
"""
from __future__ import annotations

import math
import time

import torch

from .field import NG, DeviceField

CONSERVED = ("rho", "rhou", "rhov", "rhow", "E")


def _ptrs(t: torch.Tensor):
    return [t[n].data_ptr() for n in range(5)]


def _check_state(U: torch.Tensor, field: DeviceField):
    if U.dtype != torch.float64 or not U.is_cuda or not U.is_contiguous():
        raise ValueError("states are contiguous float64 CUDA tensors")
    if tuple(U.shape) != (5,) + field.shape:
        raise ValueError(f"state shape {tuple(U.shape)} != {(5,) + field.shape}")


def parse_checkpoint(policy: str, nsteps: int):
    """('all', n) | ('uniform', K) | ('binomial', S) from a policy string."""
    if policy == "all":
        return "all", nsteps
    kind, _, arg = policy.partition(":")
    if kind in ("uniform", "binomial") and arg.isdigit() and int(arg) >= (1 if kind == "uniform" else 0):
        return kind, int(arg)
    raise ValueError(f"checkpoint policy {policy!r}: 'all', 'uniform:K' (K >= 1) or 'binomial:S' (S >= 0)")


def binomial_beta(s: int, t: int) -> int:
    """Largest number of steps reversible with s snapshots and t forward
    sweeps: C(s + t, s) (Griewank's binomial checkpointing bound)."""
    return math.comb(s + t, s)


class _Sweep:
    """State handling of one reverse sweep on a DeviceField.  source(U, i),
    when given, is added to the adjoint after the reverse of step i (the
    objective's gradient at the state U = U_i, for i >= 1)."""

    def __init__(self, field: DeviceField, dts, stats: dict, source=None, keep_stages: bool = True):
        self.field = field
        self.dts = [float(d) for d in dts]
        self.stats = stats
        self.source = source
        self.keep_stages = keep_stages

    def load(self, U: torch.Tensor):
        self.field.set_state(_ptrs(U))

    def save(self) -> torch.Tensor:
        U = torch.empty((5,) + self.field.shape, dtype=torch.float64, device="cuda")
        self.field.get_state(_ptrs(U))
        return U

    def advance(self, i: int, m: int):
        """Steps i..i+m-1 from the loaded state."""
        for j in range(i, i + m):
            self.field.step(self.dts[j])
        self.stats["recomputed_steps"] += m

    def advance_stages(self, i: int):
        """Step i from the loaded state, keeping its RK3 stage states."""
        U1 = torch.empty((5,) + self.field.shape, dtype=torch.float64, device="cuda")
        U2 = torch.empty_like(U1)
        self.field.step_stages(self.dts[i], _ptrs(U1), _ptrs(U2))
        self.stats["recomputed_steps"] += 1
        return U1, U2

    def adjoint(self, i: int, U: torch.Tensor, adj: torch.Tensor, stages=None):
        self.load(U)
        if stages is None:
            self.field.adjoint_step(self.dts[i], _ptrs(adj))
        else:
            self.field.adjoint_step_stages(self.dts[i], _ptrs(adj), _ptrs(stages[0]), _ptrs(stages[1]))
        self.stats["adjoint_steps"] += 1
        if self.source is not None and i >= 1:
            adj += self.source(U, i)

    def reverse_binomial(self, i: int, n: int, s: int, Ui: torch.Tensor, adj: torch.Tensor):
        """Reverse steps i..i+n-1 from the state Ui at step i with s free
        snapshots (recursive binomial schedule)."""
        if n == 0:
            return
        if n == 1:
            self.adjoint(i, Ui, adj)
            return
        if s == 0:
            for j in range(n - 1, -1, -1):
                self.load(Ui)
                self.advance(i, j)
                Uj = self.save() if j else Ui
                self.adjoint(i + j, Uj, adj)
            return
        t = 1
        while binomial_beta(s, t) < n:
            t += 1
        m = max(n - binomial_beta(s - 1, t), 1)
        self.load(Ui)
        self.advance(i, m)
        Um = self.save()
        self.stats["peak_snapshots"] = max(self.stats["peak_snapshots"], self.stats["live"] + 1)
        self.stats["live"] += 1
        self.reverse_binomial(i + m, n - m, s - 1, Um, adj)
        del Um
        self.stats["live"] -= 1
        self.reverse_binomial(i, m, s, Ui, adj)


def _reverse(sweep: _Sweep, kind: str, arg: int, saved: dict, n: int, adj: torch.Tensor):
    """The reverse sweep over steps n-1..0 under a checkpoint policy; adj
    holds the adjoint of U_n on entry and of U_0 on return."""
    if kind == "all":
        for i in range(n - 1, -1, -1):
            sweep.adjoint(i, saved[i], adj)
    elif kind == "uniform":
        starts = list(range(0, n, arg))
        for s0 in reversed(starts):
            s1 = min(s0 + arg, n)
            states, stages = [saved[s0]], []
            sweep.load(saved[s0])
            if sweep.keep_stages:
                # every step of the segment once, its two stage states kept:
                # the adjoint steps then recompute nothing
                for i in range(s0, s1):
                    stages.append(sweep.advance_stages(i))
                    if i < s1 - 1:
                        states.append(sweep.save())
            else:
                for i in range(s0, s1 - 1):
                    sweep.advance(i, 1)
                    states.append(sweep.save())
            for i in range(s1 - 1, s0 - 1, -1):
                sweep.adjoint(i, states[i - s0], adj, stages.pop() if stages else None)
    else:
        sweep.reverse_binomial(0, n, arg, saved[0], adj)


def _interior_only(g: torch.Tensor) -> torch.Tensor:
    adj = g.detach().contiguous().clone()
    ghost = torch.ones_like(adj, dtype=torch.bool)
    ghost[:, NG:-NG, NG:-NG, NG:-NG] = False
    adj[ghost] = 0.0
    return adj


class RK3Rollout(torch.autograd.Function):
    """U0 -> U_N over fixed time steps, with the discrete adjoint as backward."""

    @staticmethod
    def forward(ctx, U0, field: DeviceField, dts, checkpoint: str, stats: dict | None, keep_stages: bool):
        _check_state(U0, field)
        n = len(dts)
        kind, arg = parse_checkpoint(checkpoint, n)
        field.set_stream(torch.cuda.current_stream().cuda_stream)
        sweep = _Sweep(field, dts, stats if stats is not None else {}, keep_stages=keep_stages)
        for k in ("recomputed_steps", "adjoint_steps", "peak_snapshots", "live"):
            sweep.stats.setdefault(k, 0)
        t0 = time.perf_counter()
        saved = {0: U0.detach().clone()}
        sweep.load(saved[0])
        for i in range(n):
            field.step(sweep.dts[i])
            if kind == "all" or (kind == "uniform" and (i + 1) % arg == 0 and i + 1 < n):
                saved[i + 1] = sweep.save()
        out = sweep.save()
        torch.cuda.synchronize()
        sweep.stats["forward_s"] = time.perf_counter() - t0
        sweep.stats["stored_states"] = len(saved)
        ctx.sweep, ctx.kind, ctx.arg, ctx.saved, ctx.n = sweep, kind, arg, saved, n
        return out

    @staticmethod
    def backward(ctx, gU):
        sweep, kind, arg, saved, n = ctx.sweep, ctx.kind, ctx.arg, ctx.saved, ctx.n
        sweep.field.set_stream(torch.cuda.current_stream().cuda_stream)
        t0 = time.perf_counter()
        adj = _interior_only(gU)  # only interior cells carry gradient
        _reverse(sweep, kind, arg, saved, n, adj)
        torch.cuda.synchronize()
        sweep.stats["backward_s"] = time.perf_counter() - t0
        return adj, None, None, None, None, None


class RK3TimeIntegral(torch.autograd.Function):
    """J = sum_n dt_n density(U_{n+1}) over fixed time steps, with the
    discrete adjoint (plus the density's gradient at every state) as
    backward."""

    @staticmethod
    def forward(ctx, U0, field: DeviceField, dts, density, checkpoint: str, stats: dict | None, keep_stages: bool,
                density_grad):
        _check_state(U0, field)
        n = len(dts)
        kind, arg = parse_checkpoint(checkpoint, n)
        field.set_stream(torch.cuda.current_stream().cuda_stream)
        stats = stats if stats is not None else {}

        def source(U, i):
            if density_grad is not None:
                g = density_grad(U)
            else:
                with torch.enable_grad():
                    X = U.detach().requires_grad_(True)
                    (g,) = torch.autograd.grad(density(X), X)
            return (sweep.dts[i - 1] * gJ_scale[0]) * g

        gJ_scale = [1.0]
        sweep = _Sweep(field, dts, stats, source, keep_stages)
        for k in ("recomputed_steps", "adjoint_steps", "peak_snapshots", "live"):
            sweep.stats.setdefault(k, 0)
        t0 = time.perf_counter()
        saved = {0: U0.detach().clone()}
        sweep.load(saved[0])
        J = torch.zeros((), dtype=torch.float64, device=U0.device)
        for i in range(n):
            field.step(sweep.dts[i])
            Ui = sweep.save()
            with torch.no_grad():
                J += sweep.dts[i] * density(Ui)
            if kind == "all" or (kind == "uniform" and (i + 1) % arg == 0 and i + 1 < n):
                saved[i + 1] = Ui
            del Ui
        last = sweep.save()
        torch.cuda.synchronize()
        sweep.stats["forward_s"] = time.perf_counter() - t0
        sweep.stats["stored_states"] = len(saved)
        ctx.sweep, ctx.kind, ctx.arg, ctx.saved, ctx.n, ctx.last = sweep, kind, arg, saved, n, last
        ctx.scale = gJ_scale
        return J

    @staticmethod
    def backward(ctx, gJ):
        sweep, n = ctx.sweep, ctx.n
        sweep.field.set_stream(torch.cuda.current_stream().cuda_stream)
        t0 = time.perf_counter()
        ctx.scale[0] = float(gJ)
        # adjoint of U_N: the last step's source
        adj = torch.zeros_like(ctx.last)
        adj += sweep.source(ctx.last, n)
        _reverse(sweep, ctx.kind, ctx.arg, ctx.saved, n, adj)
        torch.cuda.synchronize()
        sweep.stats["backward_s"] = time.perf_counter() - t0
        return adj, None, None, None, None, None, None, None


def time_integral(U0: torch.Tensor, field: DeviceField, dts, density, checkpoint: str = "uniform:10",
                  stats: dict | None = None, keep_stages: bool = True, density_grad=None):
    """Differentiable J = sum_n dt_n density(U_{n+1}) over the rollout (see
    the module docstring).  density_grad(U), when given, is the density's
    gradient (e.g. dissipation_rate_grad, which bounds the memory on large
    boxes); otherwise torch autograd of density."""
    return RK3TimeIntegral.apply(U0, field, list(dts), density, checkpoint, stats, keep_stages, density_grad)


def time_integral_value(field: DeviceField, U0: torch.Tensor, dts, density) -> float:
    """J = sum_n dt_n density(U_{n+1}) without checkpoints (forward only)."""
    with torch.no_grad():
        field.set_stream(torch.cuda.current_stream().cuda_stream)
        field.set_state(_ptrs(U0.contiguous()))
        U = torch.empty_like(U0)
        J = 0.0
        for dt in dts:
            field.step(float(dt))
            field.get_state(_ptrs(U))
            J += float(dt) * float(density(U))
        return J


def rollout(U0: torch.Tensor, field: DeviceField, dts, checkpoint: str = "uniform:10", stats: dict | None = None,
            keep_stages: bool = True):
    """Differentiable N-step RK3 rollout (see the module docstring)."""
    return RK3Rollout.apply(U0, field, list(dts), checkpoint, stats, keep_stages)


def interior(U: torch.Tensor) -> torch.Tensor:
    return U[:, NG:-NG, NG:-NG, NG:-NG]


def mean_kinetic_energy(U: torch.Tensor) -> torch.Tensor:
    """J = mean over interior cells of 0.5 rho |u|^2 = 0.5 |rho u|^2 / rho."""
    I = interior(U)
    return (0.5 * (I[1] * I[1] + I[2] * I[2] + I[3] * I[3]) / I[0]).mean()


def d6(f: torch.Tensor, axis: int, h: float) -> torch.Tensor:
    """6th-order central first derivative on the periodic interior along a
    tensor axis (compute_gradients' periodic stencil, ViscousFlux.cpp:27-29)."""
    r = lambda k: torch.roll(f, -k, axis)  # noqa: E731  r(k)[i] = f[i + k]
    return (r(3) - 9.0 * r(2) + 45.0 * r(1) - 45.0 * r(-1) + 9.0 * r(-2) - r(-3)) / (60.0 * h)


def _dissipation_field(I: torch.Tensor, params, grid, zh: int = 0) -> torch.Tensor:
    """tau_ij du_i/dx_j per cell from interior conserved fields I (5, z, y, x).
    zh = 0: I is the whole periodic interior.  zh = 3: I is a z-slab with
    three halo planes on each side; the result covers the slab's own planes
    (z derivatives from the halo instead of the periodic image)."""
    rho = I[0]
    vel = [I[1] / rho, I[2] / rho, I[3] / rho]
    h = (grid.dx, grid.dy, grid.dz)

    def own(f):
        return f[zh:f.shape[0] - zh] if zh else f

    def dz(f):
        if not zh:
            return d6(f, 0, h[2])
        n = f.shape[0] - 2 * zh
        s = lambda k: f[zh + k:zh + k + n]  # noqa: E731  s(k)[i] = f[i + k]
        return (s(3) - 9.0 * s(2) + 45.0 * s(1) - 45.0 * s(-1) + 9.0 * s(-2) - s(-3)) / (60.0 * h[2])

    g = [[d6(own(vel[i]), 2, h[0]), d6(own(vel[i]), 1, h[1]), dz(vel[i])] for i in range(3)]
    r, v = own(rho), [own(c) for c in vel]
    p = (params.gamma - 1.0) * (own(I[4]) - 0.5 * r * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
    T = p / (r * params.Rgas)
    mu = params.mu * T ** 1.5 * (1.0 + params.S_ref) / (T + params.S_ref)
    div = g[0][0] + g[1][1] + g[2][2]
    eps = -2.0 / 3.0 * mu * div * div
    for i in range(3):
        for j in range(3):
            eps = eps + mu * (g[i][j] + g[j][i]) * g[i][j]
    return eps


def dissipation_rate(U: torch.Tensor, params, grid) -> torch.Tensor:
    """Volume mean over the interior of the viscous dissipation tau_ij
    du_i/dx_j, with the solver's ingredients: u = rho u / rho,
    p = (gamma - 1)(E - rho |u|^2 / 2), T = p / (rho R), Sutherland mu(T)
    (field_structures.h:96-106), tau_ij = mu (du_i/dx_j + du_j/dx_i) -
    2/3 mu div(u) delta_ij (ViscousFlux.cpp:151-208) and 6th-order central
    gradients on the periodic box."""
    return _dissipation_field(interior(U), params, grid).mean()


def dissipation_rate_grad(U: torch.Tensor, params, grid, slab: int = 32) -> torch.Tensor:
    """d dissipation_rate / dU (ghosted shape, zero ghosts), formed slab by
    slab along z (each slab with its three periodic halo planes), so that the
    autograd intermediates of a large box need only a slab's memory."""
    I = interior(U)
    nz = I.shape[1]
    ncell = I[0].numel()
    G = torch.zeros_like(U)
    GI = interior(G)
    for z0 in range(0, nz, slab):
        z1 = min(z0 + slab, nz)
        idx = torch.arange(z0 - 3, z1 + 3, device=U.device) % nz
        with torch.enable_grad():
            X = I.index_select(1, idx).detach().requires_grad_(True)
            (g,) = torch.autograd.grad(_dissipation_field(X, params, grid, zh=3).sum(), X)
        GI.index_add_(1, idx, g)
    return G / ncell


def kinetic_energy_density(U: torch.Tensor) -> torch.Tensor:
    """0.5 rho |u|^2 per interior cell."""
    I = interior(U)
    return 0.5 * (I[1] * I[1] + I[2] * I[2] + I[3] * I[3]) / I[0]


def central_difference(field: DeviceField, U0: torch.Tensor, d: torch.Tensor, dts, eps: float = 1e-6,
                       checkpoint: str = "binomial:0") -> torch.Tensor:
    """(J(U0 + eps d) - J(U0 - eps d)) / (2 eps) for J = mean_kinetic_energy.

    The difference is taken cell by cell before the mean: J itself is a sum
    of O(1) terms whose rounding (~1e-16 J log N) would swamp the O(eps |dJ|)
    difference on large boxes, while the cell-wise differences are small and
    sum accurately."""
    with torch.no_grad():
        Up = rollout((U0 + eps * d).contiguous(), field, dts, checkpoint)
        Um = rollout((U0 - eps * d).contiguous(), field, dts, checkpoint)
        return (kinetic_energy_density(Up) - kinetic_energy_density(Um)).mean() / (2 * eps)


def state_tensor(field: DeviceField) -> torch.Tensor:
    """The field's conserved state as a new (5, sz, sy, sx) CUDA tensor."""
    U = torch.empty((5,) + field.shape, dtype=torch.float64, device="cuda")
    field.set_stream(torch.cuda.current_stream().cuda_stream)
    field.get_state(_ptrs(U))
    return U
