"""The algebra behind the bf16x2 element-batched form, on the CPU in FP64:

* the symmetric root S = sqrt(2 mu) I + c 1 1^T (normal block, c =
  (sqrt(2 mu + 3 lambda) - sqrt(2 mu)) / 3) and sqrt(mu) (shears) squares to
  the isotropic Voigt matrix D;
* sum_q sign(w_q) F_q^T F_q with F_q = sqrt(|w_q| |J|) S B_q equals
  sum_q w_q |J| B_q^T D B_q, including rules with a negative weight;
* the Voigt rows the kernel builds per dof (xx, yy, zz, yz, xz, xy with
  engineering shears) give the isotropic elasticity pair block
  lambda g_i[a] g_j[b] + mu (g_i[b] g_j[a] + delta_ab g_i.g_j).
No GPU: these pin the formulas the kernel evaluates (kernels_eb.hip,
k_elast_bdb_mfma), so that a wrong sign, index or factor shows here."""
import numpy as np
import pytest


def voigt_d(lam, mu):
    d = np.zeros((6, 6))
    d[:3, :3] = lam
    d[[0, 1, 2], [0, 1, 2]] = lam + 2 * mu
    d[[3, 4, 5], [3, 4, 5]] = mu
    return d


def sym_root(lam, mu):
    s2m = np.sqrt(2 * mu)
    c = (np.sqrt(2 * mu + 3 * lam) - s2m) / 3
    s = np.zeros((6, 6))
    s[:3, :3] = c
    s[[0, 1, 2], [0, 1, 2]] = s2m + c
    s[[3, 4, 5], [3, 4, 5]] = np.sqrt(mu)
    return s


def voigt_b(g):
    """6 x 3n strain-displacement matrix of gradients g (n x 3), dof (i, a) =
    column 3 i + a, as the kernel's fragment rows (bdb_factor)."""
    n = g.shape[0]
    b = np.zeros((6, 3 * n))
    for i in range(n):
        gx, gy, gz = g[i]
        for ax in range(3):
            col = 3 * i + ax
            b[0, col] = gx if ax == 0 else 0.0
            b[1, col] = gy if ax == 1 else 0.0
            b[2, col] = gz if ax == 2 else 0.0
            b[3, col] = gz if ax == 1 else (gy if ax == 2 else 0.0)  # yz
            b[4, col] = gz if ax == 0 else (gx if ax == 2 else 0.0)  # xz
            b[5, col] = gy if ax == 0 else (gx if ax == 1 else 0.0)  # xy
    return b


@pytest.mark.parametrize("lam,mu", [(1.0, 0.5), (10.0, 1.0), (0.0, 2.0), (-0.5, 1.0), (3.7, 0.01)])
def test_symmetric_root_squares_to_d(lam, mu):
    s = sym_root(lam, mu)
    assert np.allclose(s, s.T)
    assert np.allclose(s @ s, voigt_d(lam, mu), rtol=1e-13, atol=1e-13 * (lam + mu))


@pytest.mark.parametrize("lam,mu", [(1.0, 0.5), (1e-6, 1.0), (-0.5, 1.0), (3.7, 0.01)])
def test_cancellation_free_c_in_fp32(lam, mu):
    """The kernel evaluates c as lambda / (sqrt(2 mu + 3 lambda) + sqrt(2 mu))
    in FP32: the same number as the difference form, and S^2 = D to FP32
    accuracy of the material scale even for lambda << mu."""
    f = np.float32
    c32 = f(lam) / (np.sqrt(f(2) * f(mu) + f(3) * f(lam)) + np.sqrt(f(2) * f(mu)))
    c64 = (np.sqrt(2 * mu + 3 * lam) - np.sqrt(2 * mu)) / 3
    assert abs(float(c32) - c64) <= 4e-7 * abs(c64) + 1e-30
    s = sym_root(lam, mu)
    s[:3, :3] += float(c32) - c64  # the FP32 c in place of the exact one
    assert np.abs(s @ s - voigt_d(lam, mu)).max() <= 2e-6 * (abs(lam) + mu)


def test_voigt_rows_give_the_isotropic_pair_block():
    rng = np.random.default_rng(3)
    g = rng.normal(size=(10, 3))
    lam, mu = 1.7, 0.6
    k = voigt_b(g).T @ voigt_d(lam, mu) @ voigt_b(g)
    for i in range(10):
        for j in range(10):
            for a in range(3):
                for b in range(3):
                    ref = lam * g[i, a] * g[j, b] + mu * (g[i, b] * g[j, a] + (a == b) * g[i] @ g[j])
                    assert abs(k[3 * i + a, 3 * j + b] - ref) <= 1e-13 * (1 + abs(ref))


@pytest.mark.parametrize("weights", [[0.2, 0.3, 0.1, 0.4], [-2 / 15, 0.075, 0.075, 0.075, 0.075]])
def test_signed_factor_sum_equals_btdb(weights):
    """Per-point factors with the sign of w_q on the C side (one XOR in the
    kernel) reproduce the weighted sum, negative centroid weight included
    (IM_TETRAHEDRON(3))."""
    rng = np.random.default_rng(11)
    n, det_j = 20, 0.037
    lam, mu = 2.3, 0.8
    s = sym_root(lam, mu)
    e_ref = np.zeros((3 * n, 3 * n))
    e = np.zeros((3 * n, 3 * n))
    for w in weights:
        bq = voigt_b(rng.normal(size=(n, 3)))
        e_ref += w * det_j * bq.T @ voigt_d(lam, mu) @ bq
        f = np.sqrt(abs(w) * det_j) * s @ bq
        e += np.sign(w) * f.T @ f
    assert np.allclose(e, e_ref, rtol=1e-12, atol=1e-12 * np.abs(e_ref).max())
    assert np.allclose(e, e.T)  # only the upper tiles are computed; the lower ones are mirrored


def test_fragment_layout_reassembles_the_product():
    """MFMA 16x16x32 bf16 lane layout as the kernel uses it: A[i][k] in lane
    i + 16 (k / 8), element k % 8; C[k][j] in lane j + 16 (k / 8); D[i][j] in
    lane j + 16 (i / 4), element i % 4.  With k = 8 (point) + Voigt row (rows 6
    and 7 zero), the per-lane fragments of F rebuild F^T F tile by tile."""
    rng = np.random.default_rng(5)
    n, npts = 20, 4
    nl = 3 * n
    nt = (nl + 15) // 16
    fq = [np.vstack([sym_root(1.0, 0.5) @ voigt_b(rng.normal(size=(n, 3))), np.zeros((2, nl))])
          for _ in range(npts)]  # 8 x nl per point
    fk = np.vstack(fq)  # 32 x nl, k = 8 q + s
    fk = np.hstack([fk, np.zeros((32, 16 * nt - nl))])
    for a in range(nt):
        for b in range(a, nt):
            acc = np.zeros((64, 4))
            for lane in range(64):
                for x in range(4):
                    i, j = 4 * (lane // 16) + x, lane % 16
                    # D[i][j] = sum_k A[i][k] C[k][j]; A[i][k] is lane i + 16 (k / 8)
                    # of the row tile's fragment, C[k][j] lane j + 16 (k / 8)
                    acc[lane, x] = sum(fk[k, 16 * a + i] * fk[k, 16 * b + j] for k in range(32))
            ref = (fk.T @ fk)[16 * a:16 * a + 16, 16 * b:16 * b + 16]
            for lane in range(64):
                for x in range(4):
                    assert abs(acc[lane, x] - ref[4 * (lane // 16) + x, lane % 16]) <= 1e-12
