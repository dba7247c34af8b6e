"""Device ensemble averages of effective charges and Raman tensors
(tdscha_weighted_tn_product behind tdscha_amd.perturbations, Modules/Perturbations.py:195-375)
against the oracle's restatement."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("N,nat", [(1, 1), (777, 7), (130, 22)])
def test_averages_match_oracle(N, nat):
    from tdscha_amd import perturbations as pt
    print("case: N=%d nat=%d" % (N, nat), flush=True)
    rng = np.random.default_rng(N + nat)
    n3 = 3 * nat
    rho = np.exp(0.3 * rng.standard_normal(N))
    u = rng.standard_normal((N, nat, 3)) * 0.1
    Z = rng.standard_normal((N, nat, 3, 3))
    R = rng.standard_normal((N, 3, 3, n3))
    M = rng.standard_normal((n3, n3))
    ups = M @ M.T
    ub = u.reshape(N, -1) * pt.A_TO_BOHR
    assert rel(pt.get_d1M_dR_av(rho, Z, device=0), orc.d1M_dR_av(rho, Z)) < 1e-13
    assert rel(pt.get_d1alpha_dR_av(rho, R, device=0), orc.d1alpha_dR_av(rho, R)) < 1e-13
    assert rel(pt.get_d2M_dR_av(u, rho, Z, ups, device=0), orc.d2M_dR_av(ub, rho, Z, ups)) < 1e-12
    assert rel(pt.get_d2alpha_dR_av(u, rho, R, ups, device=0),
               orc.d2alpha_dR_av(ub, rho, R, ups)) < 1e-12


def test_stein_identity_on_device():
    import test_perturbations as tp
    from tdscha_amd import perturbations as pt
    print("case: Stein identity, nat=4 N=64", flush=True)
    rng = np.random.default_rng(9)
    nat, N = 4, 64
    n3 = 3 * nat
    u, ups = tp.stein_ensemble(rng, n3, N)
    A = rng.standard_normal((nat, 3, 3, n3))
    Z = rng.standard_normal((nat, 3, 3))[None] + np.einsum("afcd,id->iafc", A, u)
    d2M = pt.get_d2M_dR_av(u, np.ones(N), Z, ups, device=0, a_to_bohr=1.0)
    expect = np.einsum("afcd->dacf", A).reshape(n3, n3, 3)
    assert rel(d2M, expect) < 1e-10


@pytest.mark.parametrize("K,M,N", [(1, 1, 1), (5, 70, 3), (1000, 64, 64), (333, 129, 200)])
def test_weighted_tn_product_shapes(K, M, N):
    import torch
    from tdscha_amd import _lib as L
    print("case: K=%d M=%d N=%d" % (K, M, N), flush=True)
    rng = np.random.default_rng(K * 7 + M + N)
    dev = torch.device("cuda", 0)
    A = rng.standard_normal((K, M + 3))          # leading dimension M + 3
    B = rng.standard_normal((K, N))
    w = rng.standard_normal(K)
    At, Bt, wt = (torch.as_tensor(x).to(dev) for x in (A, B, w))
    C = torch.full((M, N + 2), 7.0, dtype=torch.float64, device=dev)
    L.check(L.lib().tdscha_weighted_tn_product(
        K, M, N, At.data_ptr(), M + 3, wt.data_ptr(), Bt.data_ptr(), N, 0.5, C.data_ptr(), N + 2,
        torch.cuda.current_stream(dev).cuda_stream), "tn")
    torch.cuda.synchronize()
    got = C.cpu().numpy()
    expect = 0.5 * A[:, :M].T @ (w[:, None] * B)
    assert rel(got[:, :N], expect) < 1e-13
    assert np.all(got[:, N:] == 7.0)
