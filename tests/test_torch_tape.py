"""The float64 torch restatement (tests/torch_step.py) reproduces the FP64
oracle on a periodic 2D drop, so its autograd gradients are gradients of the
oracle's map; the GPU adjoint is checked against them in test_adjoint_gpu.py."""
import numpy as np
import torch

from clip_amd import Config
from tests import cases
from tests.oracle_lib import Oracle
from tests.torch_step import TorchStep, interior_phi0, wrapped_grid


def test_torch_restatement_matches_oracle():
    cfg = Config.from_string(cases.drop2d(n=24, radius=6))
    phi0 = interior_phi0(cfg)
    rng = np.random.default_rng(0)
    phi0 = phi0 + 1e-3 * rng.standard_normal(phi0.shape)
    o = Oracle(cfg).init_from_phi(wrapped_grid(phi0)).step(15)
    ts = TorchStep(cfg)
    F, G, U, Phi = ts.run(torch.tensor(phi0), torch.tensor(cfg.sigma, dtype=torch.float64),
                          torch.tensor(cfg.mobility, dtype=torch.float64), 15)
    phi_o = o.interior("phi")[..., 0, 0]
    u_o = o.interior("u")[..., 0, :]
    assert np.abs(Phi.numpy() - phi_o).max() <= 1e-13
    assert np.abs(U.numpy().transpose(1, 2, 0) - u_o).max() <= 1e-15
