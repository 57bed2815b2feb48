"""Electrostatic potential of the domain: P1 finite elements on the reference's mesh.

Restates kmc_dn.init_V / update_V (kmc_dopant_networks.py:706-821), which solve
Laplace's equation with FEniCS:

* mesh: RectangleMesh((0,0), (xdim,ydim), int(xdim//res), int(ydim//res)) with
  the default "right" diagonal (:766-770); at res = 0.01 that is 99 x 99 cells;
* Dirichlet data on every boundary vertex from the first matching clause of the
  ternary expression built at :733-762 -- an electrode on the x = 0 / x = xdim
  sides covers the vertices with x[0] == e_x and |x[1] - e_y| <= xdim/10, the
  others x[1] == e_y and |x[0] - e_x| <= xdim/10; static electrodes follow the
  electrodes; every other boundary vertex gets mu;
* P1 stiffness assembled per triangle, interior solve with the boundary values
  moved to the right-hand side (f = 0, :784-787);
* V(x, y) = P1 interpolation in the cell that contains (x, y).

1D (IntervalMesh, point electrodes x[0] == e_x) is supported the same way.
3D has no reference potential (the reference's 3D branch reads an undefined
array, :852-858); :func:`potential_3d_slab` is this engine's own choice: the
2D solution of the (x, y) footprint, constant through the slab thickness.

Because the problem is linear, :class:`LaplaceBasis` solves once per electrode
(unit voltage, others 0, mu = 0) plus once for mu, and forms any operating
point by superposition -- the generalisation of the IV shortcut at
kmc_dopant_networks_utils.py:655-676.
"""
from __future__ import annotations

import functools

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


class Mesh2D:
    """Vertices and 'right'-diagonal triangles of a uniform rectangle mesh."""

    def __init__(self, xdim: float, ydim: float, nx: int, ny: int):
        self.xdim, self.ydim, self.nx, self.ny = float(xdim), float(ydim), int(nx), int(ny)
        hx = (self.xdim - 0.0) / self.nx
        hy = (self.ydim - 0.0) / self.ny
        self.hx, self.hy = hx, hy
        ix = np.arange(self.nx + 1)
        iy = np.arange(self.ny + 1)
        self.xs = 0.0 + ix * hx
        self.ys = 0.0 + iy * hy
        X, Y = np.meshgrid(self.xs, self.ys)          # vertex v = iy*(nx+1) + ix
        self.x = X.ravel()
        self.y = Y.ravel()
        n1 = self.nx + 1
        v0 = (iy[:-1, None] * n1 + ix[None, :-1]).ravel()
        v1, v2, v3 = v0 + 1, v0 + n1, v0 + n1 + 1
        # diagonal v0-v3: lower triangle (v0, v1, v3), upper triangle (v0, v2, v3)
        self.cells = np.concatenate([np.stack([v0, v1, v3], 1), np.stack([v0, v2, v3], 1)])
        self.nv = self.x.size
        bx = (ix == 0) | (ix == self.nx)
        by = (iy == 0) | (iy == self.ny)
        B = by[:, None] | bx[None, :]
        self.boundary = np.flatnonzero(B.ravel())
        self.interior = np.flatnonzero(~B.ravel())

    def stiffness(self) -> sp.csr_matrix:
        x, y, c = self.x, self.y, self.cells
        px, py = x[c], y[c]                                   # (T, 3)
        # gradients of the barycentric basis: grad phi_k = (b_k, c_k) / (2A)
        b = np.stack([py[:, 1] - py[:, 2], py[:, 2] - py[:, 0], py[:, 0] - py[:, 1]], 1)
        g = np.stack([px[:, 2] - px[:, 1], px[:, 0] - px[:, 2], px[:, 1] - px[:, 0]], 1)
        area2 = (px[:, 1] - px[:, 0]) * (py[:, 2] - py[:, 0]) - \
                (px[:, 2] - px[:, 0]) * (py[:, 1] - py[:, 0])
        ke = (b[:, :, None] * b[:, None, :] + g[:, :, None] * g[:, None, :]) / \
             (2.0 * np.abs(area2))[:, None, None]
        rows = np.repeat(c, 3, axis=1).ravel()
        cols = np.tile(c, (1, 3)).ravel()
        return sp.csr_matrix((ke.ravel(), (rows, cols)), shape=(self.nv, self.nv))

    def evaluate(self, values: np.ndarray, px, py) -> np.ndarray:
        """P1 interpolation of vertex `values` at points (px, py)."""
        px = np.asarray(px, dtype=np.float64)
        py = np.asarray(py, dtype=np.float64)
        i = np.clip(np.floor(px / self.hx).astype(np.int64), 0, self.nx - 1)
        j = np.clip(np.floor(py / self.hy).astype(np.int64), 0, self.ny - 1)
        xi = (px - self.xs[i]) / self.hx
        eta = (py - self.ys[j]) / self.hy
        n1 = self.nx + 1
        v0 = j * n1 + i
        V0, V1, V2, V3 = values[v0], values[v0 + 1], values[v0 + n1], values[v0 + n1 + 1]
        lower = (1 - xi) * V0 + (xi - eta) * V1 + eta * V3
        upper = (1 - eta) * V0 + (eta - xi) * V2 + xi * V3
        return np.where(xi >= eta, lower, upper)


def boundary_values_2d(x, y, electrodes, static_electrodes, mu, xdim) -> np.ndarray:
    """First-match ternary of kmc_dopant_networks.py:739-762 at points (x, y)."""
    surplus = xdim / 10
    out = np.full(np.shape(x), np.nan)
    for el in list(np.asarray(electrodes).reshape(-1, 4)) + \
            list(np.asarray(static_electrodes).reshape(-1, 4)):
        ex, ey, ev = el[0], el[1], el[3]
        if ex == 0 or ex == xdim:
            hit = (x == ex) & (y >= ey - surplus) & (y <= ey + surplus)
        else:
            hit = (x >= ex - surplus) & (x <= ex + surplus) & (y == ey)
        take = hit & np.isnan(out)
        out[take] = ev
    out[np.isnan(out)] = mu
    return out


class Laplace2D:
    """Reusable factorisation of the interior P1 system of one mesh."""

    def __init__(self, xdim: float, ydim: float, res: float):
        nx, ny = int(xdim // res), int(ydim // res)
        self.mesh = Mesh2D(xdim, ydim, nx, ny)
        K = self.mesh.stiffness()
        I, B = self.mesh.interior, self.mesh.boundary
        self.K_IB = K[I][:, B]
        self.lu = spla.splu(K[I][:, I].tocsc())

    def solve(self, boundary: np.ndarray) -> np.ndarray:
        m = self.mesh
        V = np.zeros(m.nv)
        V[m.boundary] = boundary
        V[m.interior] = self.lu.solve(-(self.K_IB @ boundary))
        return V

    def solve_electrodes(self, electrodes, static_electrodes=None, mu=0.0) -> np.ndarray:
        m = self.mesh
        se = np.zeros((0, 4)) if static_electrodes is None else static_electrodes
        bv = boundary_values_2d(m.x[m.boundary], m.y[m.boundary], electrodes, se, mu, m.xdim)
        return self.solve(bv)


@functools.lru_cache(maxsize=8)
def laplace2d(xdim: float, ydim: float, res: float) -> Laplace2D:
    """Shared factorisation per mesh (read-only after construction)."""
    return Laplace2D(float(xdim), float(ydim), float(res))


class Interval1D:
    """IntervalMesh(int(xdim//res), 0, xdim) with P1 (exact piecewise-linear solve)."""

    def __init__(self, xdim: float, res: float):
        self.nx = int(xdim // res)
        self.xdim = float(xdim)
        self.h = (self.xdim - 0.0) / self.nx
        self.xs = 0.0 + np.arange(self.nx + 1) * self.h

    def solve_electrodes(self, electrodes, static_electrodes=None, mu=0.0) -> np.ndarray:
        se = np.zeros((0, 4)) if static_electrodes is None else static_electrodes
        vals = np.full(self.nx + 1, np.nan)
        for k in (0, self.nx):
            x = self.xs[k]
            for el in list(np.asarray(electrodes).reshape(-1, 4)) + list(np.asarray(se).reshape(-1, 4)):
                if x == el[0]:
                    vals[k] = el[3]
                    break
            else:
                vals[k] = mu
        # Laplace in 1D with Dirichlet ends: linear between the end values
        return vals[0] + (vals[-1] - vals[0]) * (self.xs - self.xs[0]) / (self.xs[-1] - self.xs[0])

    def evaluate(self, values, px) -> np.ndarray:
        return np.interp(np.asarray(px, dtype=np.float64), self.xs, values)


def eV_constant(acceptors, electrodes, xdim, ydim, zdim=0.0, res=None, mu=0.0,
                static_electrodes=None) -> np.ndarray:
    """V at the acceptor sites (the reference's eV_constant, :868-891)."""
    acc = np.asarray(acceptors, dtype=np.float64).reshape(-1, 3)
    if ydim == 0 and zdim == 0:
        r = xdim / 100 if res is None else res
        m = Interval1D(xdim, r)
        return m.evaluate(m.solve_electrodes(electrodes, static_electrodes, mu), acc[:, 0])
    if zdim == 0:
        r = min(xdim, ydim) / 100 if res is None else res
        lap = laplace2d(xdim, ydim, r)
        V = lap.solve_electrodes(electrodes, static_electrodes, mu)
        return lap.mesh.evaluate(V, acc[:, 0], acc[:, 1])
    return potential_3d_slab(acc, electrodes, xdim, ydim, res, mu, static_electrodes)


def potential_3d_slab(acceptors, electrodes, xdim, ydim, res=None, mu=0.0,
                      static_electrodes=None) -> np.ndarray:
    """Engine-defined 3D potential: 2D solve of the footprint, constant in z."""
    acc = np.asarray(acceptors, dtype=np.float64).reshape(-1, 3)
    r = min(xdim, ydim) / 100 if res is None else res
    lap = laplace2d(xdim, ydim, r)
    V = lap.solve_electrodes(electrodes, static_electrodes, mu)
    return lap.mesh.evaluate(V, acc[:, 0], acc[:, 1])


class LaplaceBasis:
    """Unit-voltage basis of one electrode layout: V = mu*phi_mu + sum_e V_e*phi_e at the sites."""

    def __init__(self, acceptors, electrode_xy, xdim=1.0, ydim=1.0, res=None,
                 static_electrodes=None):
        acc = np.asarray(acceptors, dtype=np.float64).reshape(-1, 3)
        el = np.asarray(electrode_xy, dtype=np.float64).reshape(-1, 4).copy()
        r = min(xdim, ydim) / 100 if res is None else res
        self.lap = laplace2d(xdim, ydim, r)
        P = el.shape[0]
        self.static = np.zeros((0, 4)) if static_electrodes is None else static_electrodes
        cols = []
        for e in range(P + 1):
            v = np.zeros(P)
            if e < P:
                v[e] = 1.0
            el[:, 3] = v
            st = self.static.copy()
            st[:, 3] = 0.0
            V = self.lap.solve_electrodes(el, st, mu=1.0 if e == P else 0.0)
            cols.append(self.lap.mesh.evaluate(V, acc[:, 0], acc[:, 1]))
        self.phi = np.stack(cols, 1)       # (N, P+1): electrodes then mu
        self.P = P

    def eV(self, V, mu=0.0) -> np.ndarray:
        """eV_constant for electrode voltages V (..., P) -> (..., N)."""
        V = np.asarray(V, dtype=np.float64)
        return V @ self.phi[:, :self.P].T + mu * self.phi[:, self.P]
