"""Pins of the oracle's wall metric derivatives and its edge/corner ghost
rules (no GPU).

Quadratic node maps.  On x(xi, eta, zeta) = M (xi, eta, zeta) + sum_a v_a
xi_a^2 every difference the metric formulas take is exact at the face
centre: central differences and the second-order one-sided wall derivative
.5 (-3 x0 + 4 x1 - x2) (fsrc/metrica.f03:89-124, 207-240, 320-355; mul_met,
fsrc/multigrid.f03:2537-2596, 2695-2754, 2852-2912) of a quadratic, and the
four-node averages across a face of a map that is quadratic in each index
separately.  So each face metric equals det(J) (J^-1 J^-T)_ab / pot with J
the exact Jacobian of the map (in the level's node strides) at the face
centre, and giac equals det(J) at the cell centre.  A first-order one-sided
wall derivative, a derivative taken at a shifted face or a transposed
Jacobian all fail (the test checks that the first-order value differs).

Edge and corner ghosts (fsrc/multigrid.f03:1358-1742).  Each is a copy of
one other value.  Two properties pin the rules:
  * a field that satisfies the boundary conditions of every direction
    (periodic with the box period along periodic directions, mirror-symmetric
    about both walls along wall directions) is left unchanged by every rule,
    so each source is a periodic image or the mirror neighbour across a
    wall, never another row;
  * a field constant along a set S of directions (generic otherwise) is left
    unchanged at a ghost exactly when the copy directions of that ghost lie
    in S, which pins the direction of every copy.  The copy directions are
    the reference's, listed per ghost class below with their lines.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

# ----------------------------------------------------------------- metrics

M_Q = np.array([[0.30, 0.05, 0.02], [0.04, 0.20, -0.03], [0.01, -0.07, 0.25]])
V_Q = np.array([[0.004, 0.003, -0.002],    # v_xi  (column: the xi^2 term)
                [-0.002, 0.006, 0.001],    # v_eta
                [0.003, -0.001, 0.005]]).T  # v_zeta: V_Q[:, a] multiplies xi_a^2


def quad_grid(nx, ny, nz, M, V, ip, jp, kp):
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    xi = np.stack([i, j, k], axis=-1).astype(float)
    X = xi @ np.asarray(M).T + (xi ** 2) @ np.asarray(V).T
    return O.Grid(nx, ny, nz, ip, jp, kp, *(np.ascontiguousarray(X[..., c]) for c in range(3)))


def jac(M, V, xi):
    """Exact Jacobian dX/dxi (columns) at the fine-index point xi."""
    return np.asarray(M) + 2.0 * np.asarray(V) * np.asarray(xi)[None, :]


def closed_form_metrics(M, V, shape, st):
    """Face metrics of the level with node stride st, ghosted layout, and the
    face vectors / giac of level 1.  Face centres: xi-face i of cell (j, k)
    at fine indices (i st, (j - 1/2) st, (k - 1/2) st), etc."""
    nx, ny, nz = shape
    g = [np.zeros((nz + 2, ny + 2, nx + 2)) for _ in range(9)]
    v = [np.zeros((nz + 2, ny + 2, nx + 2)) for _ in range(9)]
    giac = np.zeros((nz + 2, ny + 2, nx + 2))
    for a in range(3):
        rng = [range(1, nx + 1), range(1, ny + 1), range(1, nz + 1)]
        rng[a] = range(0, (nx, ny, nz)[a] + 1)
        for k in rng[2]:
            for j in rng[1]:
                for i in rng[0]:
                    c = [i - 0.5, j - 0.5, k - 0.5]
                    c[a] = (i, j, k)[a]
                    J = st * jac(M, V, [st * t for t in c])
                    det = np.linalg.det(J)
                    Ji = np.linalg.inv(J)
                    G = det * (Ji @ Ji.T) / st
                    for b in range(3):
                        g[3 * a + b][k, j, i] = G[a, b]
                        v[3 * a + b][k, j, i] = det * Ji[a, b]
    for k in range(1, nz + 1):
        for j in range(1, ny + 1):
            for i in range(1, nx + 1):
                giac[k, j, i] = np.linalg.det(st * jac(M, V, [st * (i - .5), st * (j - .5), st * (k - .5)]))
    return g, v, giac


FACES = [lambda nx, ny, nz: np.s_[1:nz + 1, 1:ny + 1, 0:nx + 1],
         lambda nx, ny, nz: np.s_[1:nz + 1, 0:ny + 1, 1:nx + 1],
         lambda nx, ny, nz: np.s_[0:nz + 1, 1:ny + 1, 1:nx + 1]]


@pytest.mark.parametrize("flags,V", [
    ((1, 1, 1), V_Q),                                   # walls everywhere, quadratic in every index
    ((0, 1, 0), np.outer([0.004, 0.006, -0.003], [0, 1, 0])),  # periodic xi, zeta: quadratic in eta only
])
def test_metrics_exact_on_quadratic_node_maps_all_levels(flags, V):
    M = M_Q if flags == (1, 1, 1) else np.array([[0.3, 0.05, 0.0], [0.0, 0.2, 0.0], [0.0, -0.07, 0.25]])
    g = quad_grid(16, 8, 16, M, V, *flags)
    for lev in range(3):
        st = 1 << lev
        d = g.dims(lev)
        shape = (d.nx, d.ny, d.nz)
        gm, csv, giac = O.metrics(g, lev, want_csv=(lev == 0), want_giac=(lev == 0))
        ref_g, ref_v, ref_j = closed_form_metrics(M, V, shape, st)
        for q in range(9):
            s = FACES[q // 3](*shape)
            scale = np.abs(ref_g[q][s]).max() + np.abs(ref_g[3 * (q // 3) + q // 3][s]).max()
            err = np.abs(gm[q][s] - ref_g[q][s]).max() / scale
            assert err < 1e-12, (lev, q, err)
        if lev == 0:
            for q in range(9):
                s = FACES[q // 3](*shape)
                assert np.allclose(csv[q][s], ref_v[q][s], rtol=1e-12, atol=1e-14), q
            assert np.allclose(giac[1:-1, 1:-1, 1:-1], ref_j[1:-1, 1:-1, 1:-1], rtol=1e-12)


def test_quadratic_pin_rejects_a_first_order_wall_derivative():
    """The pin above has teeth: at the wall faces of the quadratic map the
    first-order one-sided derivative x1 - x0 differs from the exact one by
    v_a st^2, which moves every wall metric by far more than the tolerance."""
    M, V = M_Q, V_Q
    for st in (1, 2, 4):
        for a in range(3):
            J2 = st * jac(M, V, [0.0 if b == a else 3.5 * st for b in range(3)])   # exact at the wall
            J1 = J2.copy()
            J1[:, a] += V[:, a] * st * st                                         # x1 - x0
            G2 = np.linalg.det(J2) * np.linalg.inv(J2) @ np.linalg.inv(J2).T
            G1 = np.linalg.det(J1) * np.linalg.inv(J1) @ np.linalg.inv(J1).T
            assert np.abs(G1[a] - G2[a]).max() / np.abs(G2[a, a]) > 1e-4, (st, a)


# ----------------------------------------------------------------- edge rules

def _profile(n, periodic, phase):
    """Values at ghosted indices 0..n+1 that satisfy the direction's boundary
    condition exactly: the ghosts are the periodic images (period n) or the
    mirror neighbours across the walls of generic interior values."""
    t = np.arange(n + 2, dtype=float)
    f = np.cos(2.1 * t / n + phase) + 0.3 * np.sin(5.3 * t / n + 2 * phase)
    if periodic:
        f[0], f[n + 1] = f[n], f[1]
    else:
        f[0], f[n + 1] = f[1], f[n]
    return f


def _bc_field(nx, ny, nz, ip, kp, rng):
    fx = _profile(nx, ip == 0, rng.uniform(0, 1))
    fy = _profile(ny, False, rng.uniform(0, 1))
    fz = _profile(nz, kp == 0, rng.uniform(0, 1))
    Z, Y, X = np.meshgrid(fz, fy, fx, indexing="ij")
    # a generic smooth function of the three profiles
    return np.ascontiguousarray(np.sin(1.3 * X + 0.7) * np.cos(0.9 * Y - 0.4) + X * Y * Z + 0.5 * Z ** 2 - Y)


LAYOUTS = [(0, 0), (1, 0), (0, 1), (1, 1)]


@pytest.mark.parametrize("ip,kp", LAYOUTS)
def test_edge_rules_keep_a_field_that_satisfies_every_boundary_condition(ip, kp):
    nx, ny, nz = 6, 5, 7
    rng = np.random.default_rng(3)
    g = O.Grid(nx, ny, nz, ip, 1, kp)
    f = _bc_field(nx, ny, nz, ip, kp, rng)
    p = f.copy()
    O.edge_fill(g, 0, p)
    assert np.array_equal(p, f)
    # a field that breaks the conditions is changed at the edges only
    h = rng.standard_normal(f.shape)
    q = h.copy()
    O.edge_fill(g, 0, q)
    changed = q != h
    kk, jj, ii = np.nonzero(changed)
    outside = [(i in (0, nx + 1)) + (j in (0, ny + 1)) + (k in (0, nz + 1)) for k, j, i in zip(kk, jj, ii)]
    assert changed.any() and min(outside) >= 2


def _ghost_classes(nx, ny, nz):
    """Non-face ghost positions grouped as the reference's loops group them:
    'xline' (i interior, j and k ghost), 'vframe' (i ghost, k ghost, any j),
    'hframe' (i ghost, j ghost, k interior)."""
    out = {"xline": [], "vframe": [], "hframe": []}
    for k in range(nz + 2):
        for j in range(ny + 2):
            for i in range(nx + 2):
                gi, gj, gk = i in (0, nx + 1), j in (0, ny + 1), k in (0, nz + 1)
                if not gi and gj and gk:
                    out["xline"].append((k, j, i))
                elif gi and gk:
                    out["vframe"].append((k, j, i))
                elif gi and gj and not gk:
                    out["hframe"].append((k, j, i))
    return out


# Copy directions of each ghost class ('x', 'y', 'z'), from the reference:
#   kp = 0: x-lines copy the periodic image in z (1373-1440); the vertical
#   frame copies the image in x and z with ip = 0 (1453-1525), in z with
#   ip = 1 (1530-1598); the horizontal frame copies along x with ip = 0
#   (1604-1611), along y (the wall neighbour) with ip = 1 (1614-1621).
#   kp = 1: x-lines copy the wall neighbour in y (1640-1650); the vertical
#   frame copies along x (image with ip = 0, 1665-1677; wall neighbour with
#   ip = 1, 1691-1703); the horizontal frame copies along x (1711-1728).
COPY_DIRS = {
    (0, 0): {"xline": "z", "vframe": "xz", "hframe": "x"},
    (1, 0): {"xline": "z", "vframe": "z", "hframe": "y"},
    (0, 1): {"xline": "y", "vframe": "x", "hframe": "x"},
    (1, 1): {"xline": "y", "vframe": "x", "hframe": "x"},
}


@pytest.mark.parametrize("ip,kp", LAYOUTS)
def test_edge_rules_copy_along_the_reference_directions(ip, kp):
    nx, ny, nz = 6, 5, 7
    g = O.Grid(nx, ny, nz, ip, 1, kp)
    rng = np.random.default_rng(11)
    classes = _ghost_classes(nx, ny, nz)
    axes = {"x": 2, "y": 1, "z": 0}
    for r in (1, 2):
        for S in itertools.combinations("xyz", r):
            # a generic field constant along the directions in S
            shape = [nz + 2, ny + 2, nx + 2]
            small = [1 if d in S else shape[axes[d]] for d in "zyx"]
            f = np.ascontiguousarray(np.broadcast_to(rng.standard_normal(small), shape))
            p = f.copy()
            O.edge_fill(g, 0, p)
            for cls, pos in classes.items():
                if cls == "vframe":
                    # vertical-frame positions whose j is a ghost row take the
                    # x-line values first: skip those corners here (the
                    # previous test covers them)
                    pos = [q for q in pos if 0 < q[1] < ny + 1]
                want = set(COPY_DIRS[(ip, kp)][cls]) <= set(S)
                kept = [p[q] == f[q] for q in pos]
                if want:
                    assert all(kept), (S, cls)
                else:
                    assert not any(kept), (S, cls)
