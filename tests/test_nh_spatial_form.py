"""The spatial form the Neo-Hookean element kernel evaluates, pinned against
the material form GetFEM's GWFL derivative gives (src/getfem_nonlinear_
elasticity.cc:612-696 for the Bonet law, :2319-2320 for the tangent
expression): for the node pair (i, j) and one point,

    g_i,c D[(a,c),(b,n)] g_j,n
      = lambda h_i (x) h_j + (mu - lambda ln J) h_j (x) h_i + mu (g_i . g_j) I
        (+ pen (F g_i) . (F g_j) I where det F <= 0)

with D[(a,c),(b,n)] = delta_ab S_nc + F_ak A_kcmn F_bm, A = dS/dE and h = F^-T g.
The material side is written here from the law itself (numpy, four nested
indices), not from the kernel."""
import numpy as np
import pytest


def material_block(F, gi, gj, lam, mu):
    C = F.T @ F
    Ci = np.linalg.inv(C)
    ln3 = np.log(np.linalg.det(C))
    S = mu * (np.eye(3) - Ci) + 0.5 * lam * ln3 * Ci
    if np.linalg.det(F) <= 0.0:
        S = S + 1e200 * C
    # dCi_kc/dC_mn (symmetrised) and A = 2 dS/dC (the penalty term is not differentiated)
    dCi = -0.5 * (np.einsum("km,nc->kcmn", Ci, Ci) + np.einsum("kn,mc->kcmn", Ci, Ci))
    A = 2.0 * (-mu * dCi + 0.5 * lam * (np.einsum("mn,kc->kcmn", Ci, Ci) + ln3 * dCi))
    D = np.einsum("ab,nc->acbn", np.eye(3), S) + np.einsum("ak,kcmn,bm->acbn", F, A, F)
    return np.einsum("c,acbn,n->ab", gi, D, gj)


def spatial_block(F, gi, gj, lam, mu):
    Finv = np.linalg.inv(F)
    hi, hj = Finv.T @ gi, Finv.T @ gj
    lnJ = 0.5 * np.log(np.linalg.det(F.T @ F))
    pen = 1e200 if np.linalg.det(F) <= 0.0 else 0.0
    return (lam * np.outer(hi, hj) + (mu - lam * lnJ) * np.outer(hj, hi)
            + (mu * gi @ gj + pen * (F @ gi) @ (F @ gj)) * np.eye(3))


@pytest.mark.parametrize("seed", range(6))
def test_spatial_form_equals_material_form(seed):
    rng = np.random.default_rng(seed)
    lam, mu = 10.0 * rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
    F = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
    if np.linalg.det(F) <= 0.0:
        F[:, 0] *= -1.0
    gi, gj = rng.standard_normal(3), rng.standard_normal(3)
    Km, Ks = material_block(F, gi, gj, lam, mu), spatial_block(F, gi, gj, lam, mu)
    assert np.abs(Km - Ks).max() <= 1e-12 * np.abs(Km).max()
    # the block of (j, i) is the transpose: K is symmetric
    assert np.abs(spatial_block(F, gj, gi, lam, mu) - Ks.T).max() <= 1e-13 * np.abs(Ks).max()


def test_inverted_point_with_the_penalty():
    rng = np.random.default_rng(11)
    F = np.diag([-1.1, 0.9, 1.05]) + 0.05 * rng.standard_normal((3, 3))
    assert np.linalg.det(F) < 0.0
    gi, gj = rng.standard_normal(3), rng.standard_normal(3)
    Km, Ks = material_block(F, gi, gj, 10.0, 1.0), spatial_block(F, gi, gj, 10.0, 1.0)
    assert np.abs(Km - Ks).max() <= 1e-12 * np.abs(Km).max()
