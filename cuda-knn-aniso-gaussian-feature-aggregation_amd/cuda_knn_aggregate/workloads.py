"""Synthetic, seeded inputs for the benchmark and test workloads.

All tensors are generated in float32 on the CPU with ``torch.Generator`` (so every
consumer sees identical bits) and moved afterwards.  Shapes and distributions:

* ``readme``     B=2, Q=1000, P=500, K=12, F=32 (the reference README's example,
                 README.md:26-45, with rotation matrices instead of ``torch.rand``);
                 q, p, f ~ U[0,1); sigma ~ U[0.05, 0.3]; R from normalised N(0,1)
                 quaternions.
* ``bench_small`` B=2, Q=12288, P=4096, K=16, F=64, same distributions (the shape of
                 the reference's published timing table, README.md:47-55).
* ``clustered``  B=1, Q=2^20, P=2^16, K=16, F=64: points from 64 isotropic Gaussian
                 clusters (centres U[0,1)^3, std 0.03), queries = point samples +
                 N(0, 0.01) jitter, sigma log-uniform [0.005, 0.05], random rotations.
* ``ragged``     64 clouds, P_i log-uniform [1024, 32768], Q_i log-uniform
                 [4096, 131072], K=16, F=32, readme distributions, offset-packed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import torch


def rand_rot(n: int, gen: torch.Generator, dtype=torch.float32) -> torch.Tensor:
    """Rotation matrices [n,3,3] from normalised N(0,1) quaternions (w, x, y, z)."""
    u = torch.randn(n, 4, generator=gen, dtype=torch.float64)
    u = u / u.norm(dim=1, keepdim=True)
    w, x, y, z = u.unbind(1)
    R = torch.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
        2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
        2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y),
    ], dim=1).reshape(n, 3, 3)
    return R.to(dtype)


@dataclass
class Batch:
    q: torch.Tensor
    p: torch.Tensor
    f: torch.Tensor
    sigma: torch.Tensor
    R: torch.Tensor
    K: int
    v: Optional[torch.Tensor] = None      # [F] Eikonal direction
    g: Optional[torch.Tensor] = None      # [B,Q,F] upstream gradient

    def to(self, device, dtype=None):
        def mv(t):
            if t is None:
                return None
            t = t.to(device)
            return t.to(dtype) if (dtype is not None and t.is_floating_point()) else t
        return Batch(mv(self.q), mv(self.p), mv(self.f), mv(self.sigma), mv(self.R), self.K,
                     mv(self.v), mv(self.g))

    @property
    def shape(self):
        return dict(B=self.q.shape[0], Q=self.q.shape[1], P=self.p.shape[1], F=self.f.shape[2], K=self.K)


def uniform_batch(B, Q, P, K, F, seed=0, sigma_range=(0.05, 0.3), with_grad=True) -> Batch:
    """q, p, f ~ U[0,1); sigma ~ U[lo, hi); random rotations; v, g ~ N(0,1)."""
    gen = torch.Generator().manual_seed(seed)
    q = torch.rand(B, Q, 3, generator=gen)
    p = torch.rand(B, P, 3, generator=gen)
    f = torch.rand(B, P, F, generator=gen)
    lo, hi = sigma_range
    sigma = lo + (hi - lo) * torch.rand(B, P, 3, generator=gen)
    R = rand_rot(B * P, gen).reshape(B, P, 3, 3)
    v = torch.randn(F, generator=gen)
    g = torch.randn(B, Q, F, generator=gen) if with_grad else None
    return Batch(q, p, f, sigma, R, K, v, g)


def readme(seed=0) -> Batch:
    return uniform_batch(2, 1000, 500, 12, 32, seed)


def bench_small(seed=0) -> Batch:
    return uniform_batch(2, 12288, 4096, 16, 64, seed)


def clustered_batch(Q=1 << 20, P=1 << 16, K=16, F=64, seed=0, centres=64, spread=0.03,
                    jitter=0.01, sigma_lo=0.005, sigma_hi=0.05, with_grad=False) -> Batch:
    """One clustered cloud (B=1): the 1M-query / 64K-point workload by default."""
    gen = torch.Generator().manual_seed(seed)
    c = torch.rand(centres, 3, generator=gen)
    p = c[torch.arange(P) % centres] + spread * torch.randn(P, 3, generator=gen)
    pick = torch.randint(P, (Q,), generator=gen)
    q = p[pick] + jitter * torch.randn(Q, 3, generator=gen)
    f = torch.rand(P, F, generator=gen)
    lo, hi = math.log(sigma_lo), math.log(sigma_hi)
    sigma = torch.exp(lo + (hi - lo) * torch.rand(P, 3, generator=gen))
    R = rand_rot(P, gen)
    v = torch.randn(F, generator=gen)
    g = torch.randn(1, Q, F, generator=gen) if with_grad else None
    return Batch(q[None], p[None], f[None], sigma[None], R[None], K, v, g)


def clustered(seed=0) -> Batch:
    return clustered_batch(seed=seed)


@dataclass
class Ragged:
    q: torch.Tensor        # [sum Q_i, 3]
    p: torch.Tensor        # [sum P_i, 3]
    f: torch.Tensor        # [sum P_i, F]
    sigma: torch.Tensor    # [sum P_i, 3]
    R: torch.Tensor        # [sum P_i, 3, 3]
    q_off: torch.Tensor    # [B+1] int64 (CPU)
    p_off: torch.Tensor    # [B+1] int64 (CPU)
    K: int
    v: Optional[torch.Tensor] = None

    def to(self, device, dtype=None):
        def mv(t):
            t = t.to(device)
            return t.to(dtype) if (dtype is not None and t.is_floating_point()) else t
        return Ragged(mv(self.q), mv(self.p), mv(self.f), mv(self.sigma), mv(self.R),
                      self.q_off, self.p_off, self.K, None if self.v is None else mv(self.v))

    def cloud(self, b: int) -> Batch:
        qs, qe = int(self.q_off[b]), int(self.q_off[b + 1])
        ps, pe = int(self.p_off[b]), int(self.p_off[b + 1])
        return Batch(self.q[None, qs:qe], self.p[None, ps:pe], self.f[None, ps:pe],
                     self.sigma[None, ps:pe], self.R[None, ps:pe], self.K, self.v)

    @property
    def B(self) -> int:
        return self.q_off.numel() - 1


def ragged_batch(B=64, p_range=(1024, 32768), q_range=(4096, 131072), K=16, F=32, seed=0) -> Ragged:
    """Offset-packed clouds with log-uniform sizes and readme distributions."""
    gen = torch.Generator().manual_seed(seed)

    def logu(lo, hi):
        return torch.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * torch.rand(B, generator=gen, dtype=torch.float64)).round().long()

    Ps = logu(*p_range).clamp(min=K)
    Qs = logu(*q_range)
    p_off = torch.zeros(B + 1, dtype=torch.int64)
    q_off = torch.zeros(B + 1, dtype=torch.int64)
    p_off[1:] = torch.cumsum(Ps, 0)
    q_off[1:] = torch.cumsum(Qs, 0)
    TP, TQ = int(p_off[-1]), int(q_off[-1])
    q = torch.rand(TQ, 3, generator=gen)
    p = torch.rand(TP, 3, generator=gen)
    f = torch.rand(TP, F, generator=gen)
    sigma = 0.05 + 0.25 * torch.rand(TP, 3, generator=gen)
    R = rand_rot(TP, gen)
    v = torch.randn(F, generator=gen)
    return Ragged(q, p, f, sigma, R, q_off, p_off, K, v)


def ragged_small(seed=0) -> Ragged:
    """A small ragged batch for tests: 6 clouds of a few hundred rows."""
    return ragged_batch(B=6, p_range=(40, 400), q_range=(30, 700), K=8, F=32, seed=seed)
