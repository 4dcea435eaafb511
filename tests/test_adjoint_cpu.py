"""CPU checks of the adjoint's loss and shape parametrisation (electricbubbles_amd/adjoint.py):
the deformation of ellipses against its closed form (L - B)/(L + B), invariance under
translation, rotation and reparametrisation, and the initial shapes against their formula."""
import math

import numpy as np
import torch

from electricbubbles_amd import adjoint


def ellipse(a, b, N, c=0j, rot=0.0, shift=0.0):
    th = 2 * math.pi * (np.arange(N) + shift) / N
    z = (a * np.cos(th) + 1j * b * np.sin(th)) * np.exp(1j * rot) + c
    return torch.tensor(z)[None, :]


def test_deformation_of_ellipses_is_the_closed_form():
    for a, b in ((1.2, 0.8), (1.0, 1.0 / 1.3), (2.0, 0.5)):
        D = float(adjoint.deformation(None, ellipse(a, b, 256)))
        assert abs(D - (a - b) / (a + b)) < 1e-13, (a, b, D)


def test_deformation_is_invariant():
    z0 = ellipse(1.3, 0.9, 200)
    D0 = float(adjoint.deformation(None, z0))
    for c, rot, shift in ((3.0 - 2.0j, 0.0, 0.0), (0j, 0.7, 0.0), (1.0 + 1.0j, 2.1, 0.37)):
        D = float(adjoint.deformation(None, ellipse(1.3, 0.9, 200, c, rot, shift)))
        assert abs(D - D0) < 1e-13


def test_deformation_of_a_perturbed_circle_to_first_order():
    # r = 1 + eps cos 2 th: L = 1 + eps, B = 1 - eps to first order, D = eps + O(eps^2)
    for eps in (1e-3, 1e-4):
        z = adjoint.initial_shapes(torch.tensor([0j]), torch.tensor([[0.0, eps, 0.0, 0.0, 0.0, 0.0]], dtype=torch.float64), 128)
        D = float(adjoint.deformation(None, z))
        assert abs(D - eps) < 3 * eps * eps


def test_initial_shapes_formula():
    rng = np.random.default_rng(0)
    modes = rng.uniform(0, 0.02, (3, 8))
    centres = np.array([0j, 4 + 1j, -2 - 3j])
    N, decay = 64, 2.0
    z = adjoint.initial_shapes(torch.tensor(centres), torch.tensor(modes), N, decay).numpy()
    th = 2 * np.pi * np.arange(N) / N
    m = np.arange(1, 5)
    r = 1 + (modes[:, :4] * m ** -decay) @ np.cos(np.outer(m, th)) + (modes[:, 4:] * m ** -decay) @ np.sin(np.outer(m, th))
    ref = centres[:, None] + r * np.exp(1j * th)[None, :]
    assert np.abs(z - ref).max() < 1e-14
