"""NumPy restatement of this project's to-tolerance pressure solver.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference multigrid (pressure_solvers/multigrid_solver.py:55-241) cannot
reach ``||div|| <= 1e-6``: its smoother is Jacobi with omega = 1 (no damping of
the checkerboard mode), every level uses the fine-grid spacing, and its
residual test includes the outlet row that carries no equation.  This module
restates the project's own textbook solver for the *same* discrete problem:

  the consistent outlet-Dirichlet operator ``A``: 5-point Laplacian
  ``ax (p[i+1] + p[i-1] - 2p) + ay (p[j+1] + p[j-1] - 2p)`` with
  ``ax = 1/dx^2, ay = 1/dy^2``, edge (Neumann) ghosts on the inlet row and
  the two wall columns, and ``p[nx-1, :] = 0`` (outlet Dirichlet,
  multigrid_solver.py:75-77).  Equations live on rows ``0 .. nx-2``.

Hierarchy (level 0 = fine grid):
  * level 0: rows 0..nx-2 unknown, row nx-1 holds the Dirichlet value 0.
  * level l >= 1: n_l = nx/2^l rows, all unknown, coefficients ax/4^l,
    ay/4^l.  The fine Dirichlet point (centre of row nx-1) lies a distance
    d_l = (1/2 - 2^-(l+1)) H_l beyond the last coarse centre, so the outlet
    ghost is the linear extrapolation through 0 at that point:
    ghost = gamma_l * last, gamma_l = 1 - H_l/d_l (-3 at l=1, -> -1).
    Keeping every level's Dirichlet point where the fine grid has it gives a
    grid-independent V(2,2) factor of ~0.085 (a face-Dirichlet coarse
    outlet gives ~0.167).
  * coarsening stops when n*m <= COARSE_CELLS (65536) and n <=
    MAX_COARSE_ROWS (1024), or when a side stops being even (or after
    exactly ``levels`` levels when that is given); the
    coarsest level is solved exactly in FP64: the walls are Neumann, so the
    orthonormal DCT-II along y diagonalises the y operator (eigenvalues
    -4 ay sin^2(pi k / 2m)), leaving one tridiagonal system along x per
    cosine mode (Thomas here; parallel cyclic reduction on the GPU; the same
    linear map).  The exact coarse solve is the "coarse-grid solve" of the
    V-cycle; a 512x128 coarsest level replaces five further levels.
Smoother: red-black Gauss-Seidel (red = (i+j) even first), exact row solve
(clamped Neumann neighbours drop out of the row and the diagonal),
over-relaxed by ``OMEGA`` = 1.2 (red-black SOR smoothing:
p <- p + omega (p_GS - p); omega = 1 is plain Gauss-Seidel).
Restriction: 2x2 average (cell-centred full weighting; level-0 outlet row
contributes 0).  Prolongation: cell-centred bilinear (3/4, 1/4 per axis),
constant extrapolation at Neumann sides, 0 at the level-0 Dirichlet row and
the same gamma_l ghost beyond a coarse outlet.
V(2,2) cycles in FP32 inside an FP64 defect-correction loop whose residual
and norm are FP64.
"""
from __future__ import annotations

import numpy as np

COARSE_CELLS = 65536
MAX_COARSE_ROWS = 1024
OMEGA = 1.2  # over-relaxation of the red-black smoother


def hierarchy(nx, ny, levels=0):
    """Level shapes ``[(n0, m0), (n1, m1), ...]`` down to the coarsest.

    ``levels`` = 0: coarsen while n*m > COARSE_CELLS or n > MAX_COARSE_ROWS,
    and both sides are even; ``levels`` = L > 0: exactly L levels
    (ValueError if the sides run out)."""
    shapes = [(nx, ny)]
    n, m = nx, ny

    def more():
        return len(shapes) < levels if levels > 0 else (n * m > COARSE_CELLS or
                                                         n > MAX_COARSE_ROWS)
    while more() and n % 2 == 0 and m % 2 == 0 and n >= 4 and m >= 2:
        n //= 2
        m //= 2
        shapes.append((n, m))
    if levels > 0 and len(shapes) != levels:
        raise ValueError(f"Grid dimensions ({nx}, {ny}) not suitable for multigrid: "
                         f"{levels} levels requested, {len(shapes)} possible")
    return shapes


def outlet_gamma(level):
    """Outlet ghost factor of coarse level ``level`` >= 1."""
    return 1.0 - 1.0 / (0.5 - 2.0 ** (-level - 1))


def _coefs(level, ax, ay):
    s = 4.0 ** level
    return ax / s, ay / s


def apply_A(p, level, ax, ay):
    """Operator of ``level`` applied to ``p`` (float64 or float32 follows p).

    Level 0: returns rows 0..n-2, row n-1 set to 0 (p[n-1] is read as the
    Dirichlet value, whatever it holds, so callers keep it at 0).
    """
    T = p.dtype.type
    axl, ayl = _coefs(level, ax, ay)
    axl, ayl = T(axl), T(ayl)
    n, m = p.shape
    out = np.zeros_like(p)
    W = np.concatenate([p[:1], p[:-1]], axis=0)  # edge ghost at i=0
    if level == 0:
        E = np.concatenate([p[1:], p[-1:]], axis=0)
    else:
        E = np.concatenate([p[1:], T(outlet_gamma(level)) * p[-1:]], axis=0)
    S = np.concatenate([p[:, :1], p[:, :-1]], axis=1)
    N = np.concatenate([p[:, 1:], p[:, -1:]], axis=1)
    out[:] = axl * ((E - p) + (W - p)) + ayl * ((N - p) + (S - p))
    if level == 0:
        out[-1, :] = 0.0
    return out


def _diag_and_offsum(p, level, axl, ayl):
    """Diagonal and off-diagonal neighbour sum of the exact row solve."""
    T = p.dtype.type
    n, m = p.shape
    i = np.arange(n)[:, None]
    j = np.arange(m)[None, :]
    hasW = (i > 0)
    hasS = (j > 0)
    hasN = (j < m - 1)
    zero = T(0)
    W = np.where(hasW, np.concatenate([p[:1], p[:-1]], axis=0), zero)
    S = np.where(hasS, np.concatenate([p[:, :1], p[:, :-1]], axis=1), zero)
    Nn = np.where(hasN, np.concatenate([p[:, 1:], p[:, -1:]], axis=1), zero)
    if level == 0:
        E = np.concatenate([p[1:], np.zeros_like(p[:1])], axis=0)
        dE = np.full((n, 1), axl, dtype=p.dtype)
    else:
        E = np.concatenate([p[1:], np.zeros_like(p[:1])], axis=0)
        dE = np.where(i == n - 1, T(1.0 - outlet_gamma(level)) * axl, axl).astype(p.dtype)
    diag = dE + np.where(hasW, axl, zero) + np.where(hasS, ayl, zero) + np.where(hasN, ayl, zero)
    off = axl * (E + W) + ayl * (Nn + S)
    return diag, off


def rbgs(p, b, level, ax, ay, sweeps, omega=OMEGA):
    """``sweeps`` red-black Gauss-Seidel sweeps, red ((i+j) even) first,
    each update over-relaxed: p <- p + omega (p_GS - p)."""
    T = p.dtype.type
    axl, ayl = _coefs(level, ax, ay)
    axl, ayl = T(axl), T(ayl)
    n, m = p.shape
    ii, jj = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    color = (ii + jj) & 1
    active = np.ones((n, m), dtype=bool)
    if level == 0:
        active[-1, :] = False
    p = p.copy()
    for _ in range(sweeps):
        for c in (0, 1):
            diag, off = _diag_and_offsum(p, level, axl, ayl)
            upd = (off - b) / diag
            sel = (color == c) & active
            if omega == 1.0:
                p[sel] = upd[sel]
            else:
                p[sel] = p[sel] + T(omega) * (upd[sel] - p[sel])
    return p


def restrict(r, level):
    """2x2 average of the fine residual; level-0 outlet row counts as 0."""
    r = r.copy()
    if level == 0:
        r[-1, :] = 0.0
    return 0.25 * ((r[0::2, 0::2] + r[1::2, 0::2]) + (r[0::2, 1::2] + r[1::2, 1::2]))


def prolong(e, fine_level, fine_shape):
    """Cell-centred bilinear prolongation of the coarse correction ``e`` onto
    the ``fine_level`` grid (3/4 own, 1/4 neighbour per axis)."""
    T = e.dtype.type
    nc, mc = e.shape
    nf, mf = fine_shape
    i = np.arange(nf)
    j = np.arange(mf)
    I = i // 2
    J = j // 2
    In = np.where(i % 2 == 0, I - 1, I + 1)
    Jn = np.where(j % 2 == 0, J - 1, J + 1)
    # x neighbour value: clamp at inlet (Neumann); beyond a coarse outlet the
    # ghost is gamma * own (linear through the Dirichlet point)
    ex_sign = np.ones(nf, dtype=e.dtype)
    ex_sign[In >= nc] = outlet_gamma(fine_level + 1)
    In_c = np.clip(In, 0, nc - 1)
    Jn_c = np.clip(Jn, 0, mc - 1)
    own = e[I][:, J]
    xn = e[In_c][:, J] * ex_sign[:, None]
    yn = e[I][:, Jn_c]
    dn = e[In_c][:, Jn_c] * ex_sign[:, None]
    out = T(0.5625) * own + T(0.1875) * (xn + yn) + T(0.0625) * dn
    if fine_level == 0:
        out[-1, :] = 0.0
    return out


def coarse_tables(level, shape, ax, ay):
    """FP64 tables of the exact coarse solve: orthonormal DCT-II matrix
    C[k, j] and the per-mode tridiagonal (off-diagonal ``off``, diagonal
    ``diag[k, i]``) on the level's unknown rows."""
    n, m = shape
    rows = n - 1 if level == 0 else n
    axl, ayl = _coefs(level, ax, ay)
    k = np.arange(m)[:, None]
    j = np.arange(m)[None, :]
    C = np.sqrt(np.where(k == 0, 1.0, 2.0) / m) * np.cos(np.pi * k * (j + 0.5) / m)
    mu = -4.0 * ayl * np.sin(np.pi * np.arange(m) / (2.0 * m)) ** 2
    i = np.arange(rows)
    dW = np.where(i > 0, axl, 0.0)
    dE = np.full(rows, axl)
    if level > 0:
        dE[-1] = axl * (1.0 - outlet_gamma(level))
    diag = -(dW + dE)[None, :] + mu[:, None]
    return C, axl, diag


def coarse_solve(b, level, shape, tables):
    """Exact FP64 solve of ``A_level x = b`` on the unknown rows (the level-0
    Dirichlet row of the result is 0): DCT along y, Thomas along x per mode,
    inverse DCT."""
    C, off, diag = tables
    n, m = shape
    rows = diag.shape[1]
    W = np.asarray(b[:rows], dtype=np.float64) @ C.T      # W[i, k]
    dd = diag.T.copy()                                    # [i, k]
    X = W.copy()
    for i in range(1, rows):                              # forward elimination
        f = off / dd[i - 1]
        dd[i] -= f * off
        X[i] -= f * X[i - 1]
    X[rows - 1] /= dd[rows - 1]
    for i in range(rows - 2, -1, -1):                     # back substitution
        X[i] = (X[i] - off * X[i + 1]) / dd[i]
    out = np.zeros((n, m))
    out[:rows] = X @ C
    return out


def dense_operator(level, shape, ax, ay):
    """Dense FP64 matrix of the level operator on its unknowns."""
    n, m = shape
    rows = n - 1 if level == 0 else n
    N = rows * m
    A = np.zeros((N, N))
    for k in range(N):
        p = np.zeros((n, m))
        p.reshape(-1)[k] = 1.0
        A[:, k] = apply_A(p, level, ax, ay)[:rows].reshape(-1)
    return A


class TextbookMG:
    """FP32 V(2,2) multigrid + FP64 defect correction for one grid shape."""

    def __init__(self, nx, ny, dx, dy, pre=2, post=2, dtype=np.float32, levels=0, omega=OMEGA):
        self.nx, self.ny = nx, ny
        self.omega = omega
        self.ax = 1.0 / (dx * dx)
        self.ay = 1.0 / (dy * dy)
        self.dx, self.dy = dx, dy
        self.shapes = hierarchy(nx, ny, levels)
        self.pre, self.post = pre, post
        self.dtype = dtype
        L = len(self.shapes) - 1
        nL = self.shapes[L][0]
        rows = nL - 1 if L == 0 else nL
        if rows < 1 or rows > MAX_COARSE_ROWS:
            raise ValueError(f"Grid dimensions ({nx}, {ny}) not suitable for multigrid: "
                             f"coarsest level has {rows} rows")
        self.coarse = coarse_tables(L, self.shapes[L], self.ax, self.ay)

    def vcycle(self, b, level=0, p=None):
        """One V-cycle on ``A_level e = b`` (zero initial guess if p is None)."""
        T = self.dtype
        shape = self.shapes[level]
        L = len(self.shapes) - 1
        if level == L:
            return coarse_solve(b, level, shape, self.coarse).astype(T)
        if p is None:
            p = np.zeros(shape, dtype=T)
        p = rbgs(p, b, level, self.ax, self.ay, self.pre, self.omega)
        r = b - apply_A(p, level, self.ax, self.ay)
        rc = restrict(r, level)
        ec = self.vcycle(rc, level + 1)
        p = p + prolong(ec, level, shape)
        p = rbgs(p, b, level, self.ax, self.ay, self.post, self.omega)
        return p

    def residual64(self, p64, b64):
        r = b64 - apply_A(p64, 0, self.ax, self.ay)
        r[-1, :] = 0.0
        return r

    def norm(self, r64):
        """||r|| = sqrt(sum over equation rows of r^2 dx dy)."""
        return float(np.sqrt(np.sum(r64[:-1] ** 2) * self.dx * self.dy))

    def solve(self, rhs, tol=1e-6, p0=None, cycles_per_outer=2, max_outer=30):
        """Defect correction: r = b - A p (FP64) -> FP32 V-cycles -> p += e.

        Returns ``(p64, stats)``; ``b[nx-1,:]`` is ignored (Dirichlet row).
        """
        b64 = np.asarray(rhs, dtype=np.float64).copy()
        b64[-1, :] = 0.0
        p64 = np.zeros_like(b64) if p0 is None else np.asarray(p0, dtype=np.float64).copy()
        p64[-1, :] = 0.0
        norms = []
        cycles = 0
        for _ in range(max_outer + 1):
            r64 = self.residual64(p64, b64)
            nr = self.norm(r64)
            norms.append(nr)
            if nr <= tol or len(norms) > max_outer:
                break
            r32 = r64.astype(self.dtype)
            e = None
            # a single-level hierarchy is one exact solve per outer iteration
            for _c in range(1 if len(self.shapes) == 1 else cycles_per_outer):
                e = self.vcycle(r32, 0, e)
                cycles += 1
            p64 = p64 + e.astype(np.float64)
        return p64, {"norms": norms, "cycles": cycles, "outer": len(norms) - 1}


def consistent_operator_sparse(nx, ny, dx, dy):
    """scipy.sparse matrix of the consistent outlet-Dirichlet operator on
    rows 0..nx-2 (the converged-pressure oracle)."""
    import scipy.sparse as sp
    ax = 1.0 / (dx * dx)
    ay = 1.0 / (dy * dy)
    rows = nx - 1
    N = rows * ny
    idx = np.arange(N).reshape(rows, ny)
    data, ri, ci = [], [], []

    def add(r, c, v):
        ri.append(r)
        ci.append(c)
        data.append(v)

    for i in range(rows):
        for j in range(ny):
            k = idx[i, j]
            d = 0.0
            # east: neighbour row i+1 (Dirichlet 0 if i+1 == nx-1)
            d -= ax
            if i + 1 < rows:
                add(k, idx[i + 1, j], ax)
            if i > 0:
                d -= ax
                add(k, idx[i - 1, j], ax)
            if j > 0:
                d -= ay
                add(k, idx[i, j - 1], ay)
            if j < ny - 1:
                d -= ay
                add(k, idx[i, j + 1], ay)
            add(k, k, d)
    return sp.csr_matrix((data, (ri, ci)), shape=(N, N))


def direct_solve(rhs, dx, dy):
    """FP64 sparse direct solve of the consistent operator; p[nx-1,:] = 0."""
    import scipy.sparse.linalg as spla
    nx, ny = rhs.shape
    A = consistent_operator_sparse(nx, ny, dx, dy)
    x = spla.spsolve(A.tocsc(), np.asarray(rhs[:-1], dtype=np.float64).reshape(-1))
    p = np.zeros((nx, ny))
    p[:-1] = x.reshape(nx - 1, ny)
    return p
