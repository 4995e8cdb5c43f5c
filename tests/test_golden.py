"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the
oracle): the oracle still reproduces them (CPU), and the HIP library matches them (GPU:
idxs and dists bit-exact, values and gradients within 1e-5 of the FP64 fixtures)."""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest
import torch

from oracle import knn_ref, plain

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(HERE, "*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: torch.from_numpy(z[k]) for k in z.files}


def test_fixtures_present():
    assert len(FIXTURES) >= 3


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p))
def test_oracle_reproduces_fixture(path):
    z = load(path)
    K = int(z["K"])
    d1, i1 = plain.knn_sort(z["q"], z["p"], z["sigma"], z["R"], K)
    d2, i2 = knn_ref.knn_ref(z["q"], z["p"], z["sigma"], z["R"], K)
    assert torch.equal(i1, z["idxs"]) and torch.equal(i2, z["idxs"])
    assert torch.equal(d1, z["dists"]) and torch.equal(d2, z["dists"])
    ins = [z[k].double() for k in ("q", "p", "f", "sigma", "R")]
    f_out, w_out, _ = plain.forward_given_idxs(*ins, z["idxs"])
    assert plain.rel_err(f_out, z["f_out"]) < 1e-13 and plain.rel_err(w_out, z["w_out"]) < 1e-13
    got = plain.backward_autograd(z["g"].double(), *ins, z["idxs"])
    for n, x in zip(("dq", "dp", "df", "dsigma", "dR"), got):
        assert plain.rel_err(x, z[n]) < 1e-12, n
    got2 = plain.backward2_autograd(z["h"].double(), z["g"].double(), *ins, z["idxs"])
    for n, x in zip(("dgrad_f_out2", "dq2", "dp2", "df2", "dsigma2", "dR2"), got2):
        assert plain.rel_err(x, z[n]) < 1e-12, n


@pytest.mark.gpu
@pytest.mark.parametrize("search", ["brute", "tiles", "wave"])
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p))
def test_gpu_matches_fixture(cuda, path, search):
    from cuda_knn_aggregate import _C

    z = load(path)
    K = int(z["K"])
    q, p, f, s, R, g, h = (z[k].to(cuda) for k in ("q", "p", "f", "sigma", "R", "g", "h"))
    f_out, w_out, dists, idxs = _C.knn_aggregate_aniso_forward(q, p, f, s, R, K, tune={"search": search})
    assert torch.equal(idxs.cpu(), z["idxs"]) and torch.equal(dists.cpu(), z["dists"])
    assert plain.rel_err(f_out.cpu(), z["f_out"]) < 1e-5 and plain.rel_err(w_out.cpu(), z["w_out"]) < 1e-5
    got = _C.knn_aggregate_aniso_backward(g, w_out, q, p, f, s, R, w_out, idxs)
    for n, x in zip(("dq", "dp", "df", "dsigma", "dR"), got):
        assert plain.rel_err(x.cpu(), z[n]) < 1e-5, n
    got2 = _C.knn_aggregate_aniso_backward_2nd(h, g, w_out, q, p, f, s, R, w_out, idxs)
    for n, x in zip(("dgrad_f_out2", "dq2", "dp2", "df2", "dsigma2", "dR2"), got2):
        assert plain.rel_err(x.cpu(), z[n]) < 1e-5, n
