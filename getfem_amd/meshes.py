"""Structured meshes and seeded synthetic inputs.

``regular_unit_mesh`` builds the same point and convex numbering as GetFEM's
``regular_unit_mesh`` (src/getfem_regular_meshes.cc:237-285): cubes visited
with x fastest, each cube's corners in reference-cube product order, each cube
cut into the six tetrahedra of the Kuhn table
(src/bgeot_convex_ref_simplexified.cc:34-39), and the final point numbering by
first encounter over the convexes' vertex lists (the copy through
add_convex_by_points at :266-280).  Vectorised with exact lattice keys.

Seeds: base S = 20261202; streams S+1 jitter, S+2 lambda, S+3 mu, S+4 rho,
S+5 u, S+6 sine modes (numpy PCG64).
"""
import numpy as np

SEED = 20261202

# Kuhn subdivision of the unit cube, corners numbered dx + 2 dy + 4 dz
KUHN6 = np.array([[3, 7, 0, 1], [7, 0, 5, 4], [7, 0, 1, 5],
                  [3, 7, 0, 2], [6, 7, 0, 4], [6, 7, 0, 2]], dtype=np.int64)
CORNERS = np.array([[k & 1, (k >> 1) & 1, (k >> 2) & 1] for k in range(8)], dtype=np.int64)


def _lattice_coords(I, n):
    """Coordinate of lattice index I along one axis as the reference stores it:
    the first cube containing the point (lowest index) adds its corner offset
    to its origin, (I-1)*h + h (getfem_regular_meshes.cc:55-60)."""
    h = 1.0 / n
    I = np.asarray(I, dtype=np.int64)
    out = np.where(I > 0, (np.maximum(I - 1, 0) * h) + h, 0.0)
    return out


def regular_unit_mesh(n, geotrans="GT_PK(3,1)"):
    """Returns (pts (n_pts,3) float64, cvx (n_cv,nv) int32)."""
    iz, iy, ix = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    cube = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)  # x fastest
    if geotrans == "GT_PK(3,1)":
        loc = CORNERS[KUHN6]                      # (6, 4, 3)
        lat = cube[:, None, None, :] + loc[None]  # (ncube, 6, 4, 3)
        lat = lat.reshape(-1, 4, 3)
    elif geotrans == "GT_QK(3,1)":
        lat = cube[:, None, :] + CORNERS[None]    # (ncube, 8, 3)
    else:
        raise ValueError("unsupported geometric transformation " + geotrans)
    nv = lat.shape[1]
    m = n + 1
    keys = (lat[..., 0] + m * lat[..., 1] + m * m * lat[..., 2]).ravel()
    uniq, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    cvx = rank[inv].reshape(-1, nv).astype(np.int32)
    ukeys = uniq[order]
    I, J, K = ukeys % m, (ukeys // m) % m, ukeys // (m * m)
    pts = np.stack([_lattice_coords(I, n), _lattice_coords(J, n), _lattice_coords(K, n)], axis=1)
    return pts, cvx


def jitter(pts, n, seed=SEED + 1, amplitude=0.1):
    """Interior nodes move by amplitude*h*U(-1,1)^3; a node is on the border
    when any coordinate is within 1e-10 of 0 or 1 (the rule of
    getfem_regular_meshes.cc:221-226)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    d = rng.uniform(-1.0, 1.0, size=pts.shape)
    border = np.any((np.abs(pts) < 1e-10) | (np.abs(pts - 1.0) < 1e-10), axis=1)
    out = pts.copy()
    out[~border] += amplitude * (1.0 / n) * d[~border]
    return out


def lognormal_lame(n_cv, seed_lambda=SEED + 2, seed_mu=SEED + 3, sigma=0.5):
    """Per-element lambda, mu = exp(N(0, sigma^2))."""
    lam = np.random.Generator(np.random.PCG64(seed_lambda)).lognormal(0.0, sigma, n_cv)
    mu = np.random.Generator(np.random.PCG64(seed_mu)).lognormal(0.0, sigma, n_cv)
    return lam, mu


def simp_density(n_cv, seed=SEED + 4):
    return np.random.Generator(np.random.PCG64(seed)).uniform(0.1, 1.0, n_cv)


def random_state(ndof, seed=SEED + 5):
    return np.random.Generator(np.random.PCG64(seed)).standard_normal(ndof)


def sine_modes(seed=SEED + 6, amplitude=0.05):
    """8 modes k_m in 2*pi*{1,2}^3, phases U(0, 2pi), coefficients N(0,1);
    returns a function x (N,3) -> u (N,3) scaled so max|u| <= amplitude on the
    evaluated points (rescaled by the caller with ``scale_to``)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ks = 2.0 * np.pi * np.array([[a, b, c] for a in (1, 2) for b in (1, 2) for c in (1, 2)], float)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=8)
    coef = rng.standard_normal((8, 3))

    def field(x, scale_to=amplitude):
        s = np.sin(x @ ks.T + phi[None, :])       # (N, 8)
        u = s @ coef                              # (N, 3)
        m = np.max(np.abs(u))
        return u * (scale_to / m if m > 0 else 1.0)

    return field


def c1_inputs(n=20):
    """P1 Poisson on the jittered n^3 x 6 tetrahedral unit cube."""
    pts, cvx = regular_unit_mesh(n)
    return jitter(pts, n), cvx


def c2_inputs(n=64):
    """Vector P2 isotropic elasticity with lognormal per-element lambda, mu."""
    pts, cvx = regular_unit_mesh(n)
    lam, mu = lognormal_lame(cvx.shape[0])
    return jitter(pts, n), cvx, lam, mu
