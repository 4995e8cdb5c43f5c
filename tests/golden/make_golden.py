"""Writes the golden fixtures in this folder from the CPU oracle (oracle/plain.py).

Run from the repository root:  python tests/golden/make_golden.py

Each fixture is an .npz (no pickles) with the float32 inputs, the upstream gradients
``g`` (first order) and ``h`` (second order), and the oracle's outputs: ``idxs`` and
``dists`` (float32, (key, index) order), ``f_out``/``w_out`` and every first- and
second-order gradient in float64.  They are regression anchors for this build's own
contract (the reference ships no test vectors).
"""
from __future__ import annotations

import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "cuda-knn-aniso-gaussian-feature-aggregation_amd"))

from cuda_knn_aggregate import workloads  # noqa: E402
from oracle import plain  # noqa: E402


def cases():
    yield "uniform_small", workloads.uniform_batch(2, 64, 48, 8, 8, seed=101, sigma_range=(1.0, 30.0))
    b = workloads.uniform_batch(1, 40, 30, 12, 5, seed=102, sigma_range=(2.0, 20.0))
    for t in (b.p, b.sigma, b.R):       # exact duplicates: tie order by lowest index
        t[0, 15:30] = t[0, 0:15]
    yield "duplicates", b
    g = torch.Generator().manual_seed(103)  # the reference README's inputs: R, sigma ~ rand
    yield "readme_rand_R", workloads.Batch(torch.rand(1, 50, 3, generator=g), torch.rand(1, 20, 3, generator=g),
                                           torch.rand(1, 20, 4, generator=g), torch.rand(1, 20, 3, generator=g),
                                           torch.rand(1, 20, 3, 3, generator=g), 6)


def make(name, b, key="aniso"):
    gen = torch.Generator().manual_seed(7)
    B, Q, F = b.q.shape[0], b.q.shape[1], b.f.shape[2]
    g = torch.randn(B, Q, F, generator=gen)
    h = torch.randn(B, Q, 3, generator=gen)
    dists, idxs = plain.knn_sort(b.q, b.p, b.sigma, b.R, b.K, key=key)
    ins = [t.double() for t in (b.q, b.p, b.f, b.sigma, b.R)]
    f_out, w_out, _ = plain.forward_given_idxs(*ins, idxs)
    dq, dp, df, ds, dR = plain.backward_closed(g.double(), *ins, idxs)
    dg2, dq2, dp2, df2, ds2, dR2 = plain.backward2_closed(h.double(), g.double(), *ins, idxs)
    arrays = dict(q=b.q, p=b.p, f=b.f, sigma=b.sigma, R=b.R, K=torch.tensor(b.K), g=g, h=h,
                  idxs=idxs, dists=dists, f_out=f_out, w_out=w_out, dq=dq, dp=dp, df=df, dsigma=ds, dR=dR,
                  dgrad_f_out2=dg2, dq2=dq2, dp2=dp2, df2=df2, dsigma2=ds2, dR2=dR2)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **{k: v.numpy() for k, v in arrays.items()})


if __name__ == "__main__":
    for name, b in cases():
        make(name, b)
        print("wrote", name)
