"""The conditioning that decides the mixed form of the Hholtz 512^2 batch (hholtz.rs:29-197,
fdma_tensor.rs:73-235): the fast diagonalisation runs through the eigenvectors Q of the
preconditioned x operator, and FP32 transforms perturb them by dQ.  Refinement against the
FP64 residual of the tensor equation contracts the error by about ||dQ|| ||Q^-1|| per step,
so it converges when that figure is far below 1.  Measured here on the oracle's set-up
(eigenvectors from its own eigensolver): cond(Q) = 12, 21, 37, 63 at n = 64 .. 512 and a
contraction of ~6e-7 at n = 512, i.e. one refinement step takes an FP32 solve to FP64 level.
"""
import numpy as np
import pytest


@pytest.mark.parametrize("n, cond_max", [(64, 20.0), (512, 100.0)])
def test_c2_eigenvectors_well_conditioned(oracle, n, cond_max):
    s = oracle.Solver("hholtz", oracle.CD, n, oracle.CD, n, [1e-3, 1e-3])
    lam, fwd, bwd = s.tensor_setup()
    assert lam.min() > 0  # Hholtz: 1 + c lam with lam > 0, no singular mode
    kq = np.linalg.cond(bwd)
    assert kq < cond_max, kq
    b32 = bwd.astype(np.float32).astype(float)
    contraction = np.linalg.norm(bwd - b32, 2) * np.linalg.norm(np.linalg.inv(bwd), 2)
    assert contraction < 1e-6, contraction
