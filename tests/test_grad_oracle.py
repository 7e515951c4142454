"""Pins of the FP64 gradient reference (oracle/diffsims_grad.py), CPU only.

The torch restatement must reproduce the NumPy oracle's patterns, and its
autograd gradients must equal central finite differences of the NumPy oracle
(every discrete choice recomputed at each shifted point).
"""

import math

import numpy as np
import pytest
import torch

from oracle import diffsims_grad as G
from oracle import diffsims_oracle as O


def mono_problem(n_atoms=5, seed=7, model="lorentzian", precession=0.0, **kw):
    rng = np.random.default_rng(seed)
    ph = O.PhaseSpec((4.3, 5.1, 4.7), (90.0, 104.0, 90.0),
                     list(rng.choice(["O", "Mg", "Si", "Fe"], n_atoms)),
                     rng.uniform(0, 1, (n_atoms, 3)), np.ones(n_atoms),
                     rng.uniform(0.3, 1.0, n_atoms))
    p = O.SimParams(kv=200, reciprocal_radius=1.3, max_excitation_error=0.03,
                    shape_model=model, precession_angle=precession, with_direct_beam=False,
                    minimum_intensity=1e-20)
    return G.Problem(ph, p, (56, 60), 1.3, 0.05, clip=1e-3, **kw)


def _torch_inputs(pr, euler):
    e = torch.tensor(euler, dtype=torch.float64, requires_grad=True)
    abc = torch.tensor(pr.ph.abc, dtype=torch.float64, requires_grad=True)
    ang = torch.tensor(pr.ph.angles, dtype=torch.float64, requires_grad=True)
    xyz = torch.tensor(pr.ph.xyz, dtype=torch.float64, requires_grad=True)
    B = torch.tensor(pr.ph.B, dtype=torch.float64, requires_grad=True)
    return e, abc, ang, xyz, B


def test_bunge_and_basis_match_numpy_oracle():
    eu = O.random_euler(16, 3)
    M = G.bunge(torch.tensor(eu)).numpy()
    assert np.allclose(M, O.bunge_matrix(eu), atol=1e-14)
    for abc, ang in [((4.05,) * 3, (90.0,) * 3), ((4.3, 5.1, 4.7), (90.0, 104.0, 90.0)),
                     ((5.0, 6.0, 7.0), (80.0, 95.0, 110.0)), ((3.0, 3.0, 5.0), (90.0, 90.0, 120.0))]:
        rec = G.reciprocal_basis(torch.tensor(abc, dtype=torch.float64),
                                 torch.tensor(ang, dtype=torch.float64)).numpy()
        assert np.allclose(rec, O.lattice_basis(*abc, *ang)[1], rtol=0, atol=1e-14)


@pytest.mark.parametrize("normalize", [True, False])
def test_torch_patterns_equal_numpy_oracle(normalize):
    pr = mono_problem(normalize=normalize)
    eu = O.random_euler(4, 11)
    M = O.bunge_matrix(eu)
    pats, frozen, tab = G.numpy_patterns(pr, M)
    assert sum(len(f.idx) for f in frozen) > 10
    e, abc, ang, xyz, B = _torch_inputs(pr, eu)
    P = G.torch_patterns(pr, G.bunge(e).transpose(-1, -2), G.reciprocal_basis(abc, ang), xyz, B,
                         frozen, tab.hkl)
    assert np.allclose(P.detach().numpy(), pats, rtol=0, atol=1e-12 * max(1.0, pats.max()))


def _fd_check(pr, eu, target, which, h=1e-4, tol=2e-5):
    """Central differences of the NumPy oracle's loss against torch autograd."""
    M = O.bunge_matrix(eu)
    _, frozen, tab = G.numpy_patterns(pr, M)
    e, abc, ang, xyz, B = _torch_inputs(pr, eu)
    L = G.torch_loss(pr, G.bunge(e).transpose(-1, -2), G.reciprocal_basis(abc, ang), xyz, B,
                     frozen, tab.hkl, target)
    L.backward()
    grads = {"euler": e.grad.numpy(), "abc": abc.grad.numpy(), "angles": ang.grad.numpy(),
             "xyz": xyz.grad.numpy(), "B": B.grad.numpy()}
    base = {"euler": eu, "abc": np.array(pr.ph.abc, float), "angles": np.array(pr.ph.angles, float),
            "xyz": np.array(pr.ph.xyz, float), "B": np.array(pr.ph.B, float)}
    scale = max(np.abs(g).max() for g in grads.values())
    def central(name, index, hh):
        vals, smooth = [], True
        for sgn in (1, -1):
            v = {k: np.array(x, dtype=np.float64, copy=True) for k, x in base.items()}
            v[name][index] += sgn * hh
            kw = dict(xyz=v["xyz"], B=v["B"], abc=tuple(v["abc"]), angles=tuple(v["angles"]))
            P, fz, _ = G.numpy_patterns(pr, O.bunge_matrix(v["euler"]), **kw)
            smooth &= all(np.array_equal(a.idx, b.idx) and np.array_equal(a.boxes, b.boxes)
                          and a.argmax == b.argmax for a, b in zip(frozen, fz))
            vals.append(float(np.sum((P - target) ** 2)))
        return smooth, (vals[0] - vals[1]) / (2 * hh)

    for name, index in which:
        # Central differences at h = 1e-4 and h/4 combined by Richardson
        # extrapolation (the O(h^2) term cancels: the narrow shape models are
        # strongly curved).  The loss is only piecewise smooth: when a clip box,
        # the survivor set or the pixel of a pattern's maximum changes inside
        # [x - h, x + h], h shrinks by 4 until it does not.
        hh = h
        for _ in range(6):
            s1, f1 = central(name, index, hh)
            s2, f2 = central(name, index, hh / 4) if s1 else (False, 0.0)
            if s1 and s2:
                break
            hh /= 4
        assert s1 and s2, (name, index)
        fd = (16 * f2 - f1) / 15
        ad = grads[name][index]
        assert abs(fd - ad) <= tol * scale + 1e-6 * abs(ad), (name, index, fd, ad, scale)


CHECKS = [("euler", (0, 0)), ("euler", (1, 1)), ("euler", (2, 2)), ("abc", 0), ("abc", 2),
          ("angles", 1), ("xyz", (0, 0)), ("xyz", (2, 1)), ("xyz", (4, 2)), ("B", 1), ("B", 3)]


@pytest.mark.parametrize("model,prec", [("lorentzian", 0.0), ("linear", 0.0), ("sin2c", 0.0),
                                        ("atanc", 0.0), ("lorentzian", 1.0)])
def test_autograd_matches_fd_of_numpy_oracle(model, prec):
    pr = mono_problem(model=model, precession=prec)
    eu = O.random_euler(3, 5)
    rng = np.random.default_rng(1)
    target = rng.uniform(0, 0.3, (3,) + pr.shape)
    _fd_check(pr, eu, target, CHECKS)


def test_fd_check_catches_a_planted_mistake():
    """A sign error in the restated excitation error must fail the FD pin."""
    pr = mono_problem()
    eu = O.random_euler(3, 5)
    target = np.random.default_rng(1).uniform(0, 0.3, (3,) + pr.shape)
    orig = G.shape_factor

    def wrong(s, r, p):
        return orig(-s + 0.004, r, p)

    G.shape_factor = wrong
    try:
        with pytest.raises(AssertionError):
            _fd_check(pr, eu, target, [("euler", (0, 0)), ("abc", 0), ("xyz", (0, 0))])
    finally:
        G.shape_factor = orig
