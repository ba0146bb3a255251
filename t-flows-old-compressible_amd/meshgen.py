"""Synthetic finite-volume systems in the T-FlowS matrix layout.

The T-FlowS solvers act on a CSR matrix whose pattern is built once by
``TopolM`` (Process/TopolM.f90:35-94) and whose values are filled by the
equation assemblers.  The layout, which every function here reproduces, is:

* ``Acol(N+1)`` -- the row pointer (the reference names it "col"),
  1-based, ``Acol(1) = 1`` and ``Acol(N+1) = nnz + 1``
  (TopolM.f90:53-58);
* ``Arow(NONZ)`` -- the column indices, 1-based, sorted ascending inside
  each row (TopolM.f90:61-74, ISORT), with ``NONZ = nnz + 1``: one
  trailing slot that is allocated and left zero (TopolM.f90:49-51);
* ``Aval(NONZ)`` -- FP64 values, trailing slot zero;
* ``Adia(N)`` -- 1-based position of the diagonal entry of each row
  (TopolM.f90:75-77).

The pattern is structurally symmetric because every interior face inserts
both (c1,c2) and (c2,c1) (TopolM.f90:65-66).

Matrix recipes follow the reference assemblers:

* pressure correction: off-diagonal ``-A12`` and diagonal ``+A12`` per
  interior face with ``A12 = SMDPN = |S|^2/(S.d)`` for unit ``V/Asave``
  (Process/CalcPS.f90:77-94); a pinned cell gets one face-sized Dirichlet
  term on its diagonal the way OUTFLOW/PRESSURE faces do
  (CalcPS.f90:146-147, 159-160);
* momentum / scalar transport: first-order upwind ``A12 = -min(F,0)``,
  ``A21 = max(F,0)``, diffusion ``D = Gamma*|S|/d``, "m1" placement
  (Process/NewUVW.f90:356-367) and the time term ``V/dt``
  (NewUVW.f90:447-449).

Wall stretching follows the hyperbolic-tangent law of the mesh generator,
case 3 (both walls), Generate/Linija.f90:114-117; natural cell numbering
(i fastest, then j, then k) follows Generate/Calc1.f90:398-402.

Everything is vectorised numpy, so a 256^3 box (16.8M rows) builds in well
under a minute.  The generator is deterministic for a given seed.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

GENERATOR_VERSION = 1

# slot order of the 7-point stencil inside a row before any renumbering;
# the offsets are ascending, so rows come out sorted in natural order.
_SLOTS = ("-z", "-y", "-x", "self", "+x", "+y", "+z")


# --------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------
def tanh_nodes(n: int, length: float, ratio: float, origin: float = 0.0) -> np.ndarray:
    """Node coordinates clustered at both ends (Generate/Linija.f90:114-117).

    ``y(i) = y0 + 0.5*(1 + tanh(xi*atanh(pr))/pr)*L`` with
    ``xi = -1 + 2i/n``.  ``pr`` in (0,1) is found by bisection so that the
    largest ratio of adjacent cell sizes equals ``ratio``; ``ratio <= 1``
    gives uniform spacing.
    """
    i = np.arange(n + 1, dtype=np.float64)
    xi = -1.0 + 2.0 * i / n

    def nodes(pr: float) -> np.ndarray:
        return origin + 0.5 * (1.0 + np.tanh(xi * np.arctanh(pr)) / pr) * length

    def max_ratio(pr: float) -> float:
        d = np.diff(nodes(pr))
        return float(np.max(np.maximum(d[1:] / d[:-1], d[:-1] / d[1:])))

    if ratio <= 1.0 or n < 3:
        return origin + length * i / n
    lo, hi = 1e-9, 1.0 - 1e-15
    if max_ratio(hi) < ratio:
        return nodes(hi)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if max_ratio(mid) < ratio:
            lo = mid
        else:
            hi = mid
    return nodes(lo)


def geometric_wall_nodes(n: int, length: float, ratio: float, max_aspect: float,
                         ref_spacing: float) -> np.ndarray:
    """Nodes clustered at both walls with a geometric growth ``ratio``.

    The first cell is ``ref_spacing / max_aspect`` thick (so that a wall cell
    reaches the requested aspect ratio against a transverse spacing of
    ``ref_spacing``); cells grow by ``ratio`` towards the centre until they
    reach uniform spacing that fills the rest of the span.
    """
    h0 = ref_spacing / max_aspect
    half = n // 2
    sizes = h0 * ratio ** np.arange(half)
    # cap the growth so both halves fit exactly into the length
    cap_lo, cap_hi = h0, length
    for _ in range(200):
        cap = 0.5 * (cap_lo + cap_hi)
        s = np.minimum(sizes, cap)
        tot = 2.0 * s.sum() + (cap if n % 2 else 0.0)
        if tot > length:
            cap_hi = cap
        else:
            cap_lo = cap
    s = np.minimum(sizes, cap_lo)
    mid = [cap_lo] if n % 2 else []
    cells = np.concatenate([s, mid, s[::-1]])
    cells *= length / cells.sum()
    return np.concatenate([[0.0], np.cumsum(cells)])


@dataclass
class BoxMesh:
    """Tensor-product hex box, cells numbered naturally (i fastest)."""
    xn: np.ndarray
    yn: np.ndarray
    zn: np.ndarray

    @property
    def shape(self):
        return (len(self.xn) - 1, len(self.yn) - 1, len(self.zn) - 1)

    @property
    def n_cells(self) -> int:
        nx, ny, nz = self.shape
        return nx * ny * nz

    def centres(self):
        return tuple(0.5 * (a[1:] + a[:-1]) for a in (self.xn, self.yn, self.zn))

    def spacings(self):
        return tuple(np.diff(a) for a in (self.xn, self.yn, self.zn))


def channel_mesh(nx: int, ny: int, nz: int, ratio: float = 1.1) -> BoxMesh:
    """2*pi x 2 x pi channel, tanh wall stretching in y."""
    return BoxMesh(np.linspace(0.0, 2.0 * np.pi, nx + 1),
                   tanh_nodes(ny, 2.0, ratio),
                   np.linspace(0.0, np.pi, nz + 1))


def unit_cube_mesh(n: int) -> BoxMesh:
    g = np.linspace(0.0, 1.0, n + 1)
    return BoxMesh(g, g.copy(), g.copy())


# --------------------------------------------------------------------------
# the T-FlowS CSR container
# --------------------------------------------------------------------------
@dataclass
class TFlowSMatrix:
    """A CSR matrix in the T-FlowS layout (1-based, NONZ = nnz + 1)."""
    n: int
    nonz: int
    acol: np.ndarray          # int32 (N+1)  row pointer, 1-based
    arow: np.ndarray          # int32 (NONZ) column index, 1-based, last slot 0
    adia: np.ndarray          # int32 (N)    diagonal position, 1-based
    aval: np.ndarray          # f64  (NONZ)  values, last slot 0
    mirror: np.ndarray | None = None   # int64 (nnz) 0-based position of (j,i) for (i,j)
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return self.nonz - 1

    def row_of_entry(self) -> np.ndarray:
        """0-based row index of every stored entry (length nnz)."""
        return np.repeat(np.arange(self.n, dtype=np.int64),
                         np.diff(self.acol.astype(np.int64)))

    def to_scipy(self):
        import scipy.sparse as sp
        nnz = self.nnz
        return sp.csr_matrix((self.aval[:nnz], self.arow[:nnz].astype(np.int64) - 1,
                              self.acol.astype(np.int64) - 1), shape=(self.n, self.n))

    def with_values(self, aval: np.ndarray) -> "TFlowSMatrix":
        return TFlowSMatrix(self.n, self.nonz, self.acol, self.arow, self.adia,
                            np.ascontiguousarray(aval, dtype=np.float64),
                            self.mirror, dict(self.meta))

    def transpose_values(self) -> np.ndarray:
        """Values of A^T on the same (structurally symmetric) pattern."""
        out = np.zeros_like(self.aval)
        out[: self.nnz] = self.aval[: self.nnz][self.mirror]
        return out

    def digest(self) -> str:
        h = hashlib.sha256()
        for a in (self.acol, self.arow, self.adia, self.aval):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()


def vector_digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _assemble(shape, off_lo, off_hi, diag, perm=None) -> TFlowSMatrix:
    """Build the T-FlowS CSR from per-direction coefficients.

    ``off_lo[d][c]`` is the entry A(c, c - stride_d) and ``off_hi[d][c]`` is
    A(c, c + stride_d) for d = x, y, z, indexed by natural 0-based cell.
    Entries pointing outside the box are dropped.  With ``perm`` the cell
    with natural index c becomes row ``perm[c]`` and every row's columns are
    re-sorted ascending, as TopolM's ISORT does for any numbering.
    """
    nx, ny, nz = shape
    n = nx * ny * nz
    c = np.arange(n, dtype=np.int64)
    i = c % nx
    j = (c // nx) % ny
    k = c // (nx * ny)
    sx, sy, sz = 1, nx, nx * ny
    cols = np.stack([c - sz, c - sy, c - sx, c, c + sx, c + sy, c + sz], axis=1)
    valid = np.stack([k > 0, j > 0, i > 0, np.ones(n, bool),
                      i < nx - 1, j < ny - 1, k < nz - 1], axis=1)
    vals = np.stack([off_lo[2], off_lo[1], off_lo[0], diag,
                     off_hi[0], off_hi[1], off_hi[2]], axis=1).astype(np.float64)
    cols = np.where(valid, cols, 0)

    if perm is not None:
        perm = np.asarray(perm, dtype=np.int64)
        newcols = np.where(valid, perm[cols], np.iinfo(np.int64).max)
        order = np.argsort(newcols, axis=1, kind="stable")
        newcols = np.take_along_axis(newcols, order, axis=1)
        vals = np.take_along_axis(vals, order, axis=1)
        valid = np.take_along_axis(valid, order, axis=1)
        inv = np.empty(n, dtype=np.int64)
        inv[perm] = c
        cols = np.where(valid, newcols, 0)[inv]
        vals = vals[inv]
        valid = valid[inv]

    counts = valid.sum(axis=1)
    nnz = int(counts.sum())
    acol = np.empty(n + 1, dtype=np.int32)
    acol[0] = 1
    np.cumsum(counts, out=acol[1:])
    acol[1:] += 1
    arow = np.zeros(nnz + 1, dtype=np.int32)
    aval = np.zeros(nnz + 1, dtype=np.float64)
    arow[:nnz] = (cols[valid] + 1).astype(np.int32)
    aval[:nnz] = vals[valid]
    rows = np.repeat(np.arange(n, dtype=np.int64), counts)
    is_diag = (arow[:nnz].astype(np.int64) - 1) == rows
    adia = np.empty(n, dtype=np.int32)
    adia[rows[is_diag]] = (np.nonzero(is_diag)[0] + 1).astype(np.int32)
    m = TFlowSMatrix(n, nnz + 1, acol, arow, adia, aval)
    m.mirror = mirror_positions(m)
    return m


def mirror_positions(m: TFlowSMatrix) -> np.ndarray:
    """0-based position of entry (j,i) for every stored entry (i,j).

    This is the role of SidAij(1,s)/SidAij(2,s) in the reference
    (TopolM.f90:81-94), extended to the diagonal (which maps to itself).
    Requires a structurally symmetric pattern.
    """
    nnz = m.nnz
    rows = m.row_of_entry()
    cols = m.arow[:nnz].astype(np.int64) - 1
    key = rows * m.n + cols           # sorted ascending already (CSR order)
    tkey = cols * m.n + rows
    pos = np.searchsorted(key, tkey)
    if not np.array_equal(key[np.minimum(pos, nnz - 1)], tkey):
        raise ValueError("pattern is not structurally symmetric")
    return pos.astype(np.int64)


def random_permutation(n: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).permutation(n).astype(np.int64)


# --------------------------------------------------------------------------
# face geometry of a tensor-product box
# --------------------------------------------------------------------------
def _face_coeffs(mesh: BoxMesh):
    """|S|/d for interior faces in each direction, per low-side cell.

    Returns three arrays of shape (n_cells,) holding the coefficient of the
    face between cell c and c + stride_d (zero where that face does not
    exist), plus the per-cell face areas used for boundary terms.
    """
    nx, ny, nz = mesh.shape
    dx, dy, dz = mesh.spacings()
    xc, yc, zc = mesh.centres()
    DX, DY, DZ = np.meshgrid(dx, dy, dz, indexing="ij")
    # transpose to (k, j, i) order then flatten so that i runs fastest
    def flat(a):
        return np.ascontiguousarray(a.transpose(2, 1, 0)).ravel()

    ax = np.zeros((nx, ny, nz))
    ay = np.zeros((nx, ny, nz))
    az = np.zeros((nx, ny, nz))
    if nx > 1:
        ax[:-1] = (DY * DZ)[:-1] / np.diff(xc)[:, None, None]
    if ny > 1:
        ay[:, :-1] = (DX * DZ)[:, :-1] / np.diff(yc)[None, :, None]
    if nz > 1:
        az[:, :, :-1] = (DX * DY)[:, :, :-1] / np.diff(zc)[None, None, :]
    vol = DX * DY * DZ
    return (flat(ax), flat(ay), flat(az)), flat(DY * DZ), flat(DX), flat(vol)


def _shift_down(a: np.ndarray, stride: int) -> np.ndarray:
    """out[c] = a[c - stride] (zero where c < stride)."""
    out = np.zeros_like(a)
    out[stride:] = a[:-stride] if stride < len(a) else 0
    return out


# --------------------------------------------------------------------------
# system recipes
# --------------------------------------------------------------------------
@dataclass
class LinearSystem:
    matrix: TFlowSMatrix
    b: np.ndarray
    x0: np.ndarray
    tol: float
    prec: int
    niter: int
    meta: dict = field(default_factory=dict)


def pressure_matrix(mesh: BoxMesh, pin_cell: int = 1, perm=None) -> TFlowSMatrix:
    """Pressure-correction matrix (Process/CalcPS.f90:77-94, 146-147).

    For an orthogonal hex face SMDPN = |S|^2/(S.d) = A_f/d_f.  Each interior
    face contributes ``-A12`` to both off-diagonal slots and ``+A12`` to both
    diagonals; ``pin_cell`` (1-based natural index, 0 to disable) gets one
    boundary-face term A/(dx/2) on its diagonal to remove the null space.
    """
    nx, ny, nz = mesh.shape
    (ax, ay, az), area_x, dxc, _ = _face_coeffs(mesh)
    strides = (1, nx, nx * ny)
    hi = [ax, ay, az]
    lo = [_shift_down(a, s) for a, s in zip(hi, strides)]
    diag = np.zeros(mesh.n_cells)
    for a in hi:
        diag += a
    for a in lo:
        diag += a
    if pin_cell:
        c = pin_cell - 1
        diag[c] += area_x[c] / (0.5 * dxc[c])
    m = _assemble(mesh.shape, [-a for a in lo], [-a for a in hi], diag, perm)
    m.meta.update(kind="pressure", shape=mesh.shape, pin_cell=pin_cell)
    return m


def pressure_matrix_torch(mesh: BoxMesh, pin_cell: int = 1, device="cuda"):
    """pressure_matrix (natural order) evaluated with torch on ``device``, for
    sizes whose host assembly takes minutes (256^3).  The same element-wise
    formulas in the same order, so the arrays are the host recipe's bit for
    bit.  Returns (n, nonz, acol, arow, adia, aval): int32 / float64 tensors
    in the T-FlowS layout (1-based, NONZ = nnz + 1, last slot 0)."""
    import torch
    f64 = torch.float64
    nx, ny, nz = mesh.shape
    n = nx * ny * nz

    def t(a):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=f64, device=device)

    dx, dy, dz = (t(a) for a in mesh.spacings())
    xc, yc, zc = (t(a) for a in mesh.centres())
    DX, DY, DZ = torch.meshgrid(dx, dy, dz, indexing="ij")

    def flat(a):
        return a.permute(2, 1, 0).contiguous().reshape(-1)

    ax = torch.zeros((nx, ny, nz), dtype=f64, device=device)
    ay = torch.zeros_like(ax)
    az = torch.zeros_like(ax)
    if nx > 1:
        ax[:-1] = (DY * DZ)[:-1] / torch.diff(xc)[:, None, None]
    if ny > 1:
        ay[:, :-1] = (DX * DZ)[:, :-1] / torch.diff(yc)[None, :, None]
    if nz > 1:
        az[:, :, :-1] = (DX * DY)[:, :, :-1] / torch.diff(zc)[None, None, :]
    hi = [flat(ax), flat(ay), flat(az)]
    strides = (1, nx, nx * ny)
    lo = []
    for a, st in zip(hi, strides):
        b = torch.zeros_like(a)
        if st < n:
            b[st:] = a[:-st]
        lo.append(b)
    diag = torch.zeros(n, dtype=f64, device=device)
    for a in hi:
        diag += a
    for a in lo:
        diag += a
    if pin_cell:
        c = pin_cell - 1
        area_x, dxc = flat(DY * DZ), flat(DX)
        diag[c] += area_x[c] / (0.5 * dxc[c])
    del ax, ay, az, DX, DY, DZ
    c = torch.arange(n, dtype=torch.int64, device=device)
    i, j, k = c % nx, (c // nx) % ny, c // (nx * ny)
    sx, sy, sz = strides
    cols = torch.stack([c - sz, c - sy, c - sx, c, c + sx, c + sy, c + sz], dim=1)
    valid = torch.stack([k > 0, j > 0, i > 0, torch.ones_like(i, dtype=torch.bool), i < nx - 1, j < ny - 1,
                         k < nz - 1], dim=1)
    vals = torch.stack([-lo[2], -lo[1], -lo[0], diag, -hi[0], -hi[1], -hi[2]], dim=1)
    del lo, hi, diag, i, j, k
    counts = valid.sum(dim=1)
    nnz = int(counts.sum())
    acol = torch.empty(n + 1, dtype=torch.int32, device=device)
    acol[0] = 1
    acol[1:] = (torch.cumsum(counts, 0) + 1).to(torch.int32)
    arow = torch.zeros(nnz + 1, dtype=torch.int32, device=device)
    arow[:nnz] = (cols[valid] + 1).to(torch.int32)
    del cols
    aval = torch.zeros(nnz + 1, dtype=f64, device=device)
    aval[:nnz] = vals[valid]
    del vals
    adia = (acol[:-1].to(torch.int64) + valid[:, :3].sum(dim=1)).to(torch.int32)
    return n, nnz + 1, acol, arow, adia, aval


def wall_pressure_mesh(n: int | tuple, ratio: float = 1.05, max_aspect: float = 50.0) -> BoxMesh:
    """The mesh of wall_pressure_system."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    dx = 2.0 * np.pi / nx
    return BoxMesh(np.linspace(0.0, 2.0 * np.pi, nx + 1),
                   geometric_wall_nodes(ny, 2.0, ratio, max_aspect, dx),
                   np.linspace(0.0, np.pi, nz + 1))


def divergence_free_fluxes(mesh: BoxMesh, seed: int, amplitude: float = 1.0, modes: int = 4):
    """Face volume fluxes F = curl(psi) of a seeded edge vector potential.

    psi vanishes on boundary edges, so every boundary-face flux is zero and
    the discrete divergence of every cell is exactly zero (up to rounding).
    Returns (Fx, Fy, Fz) indexed by the low-side cell of each interior face
    (natural numbering), scaled so that max |F|/area = ``amplitude``.
    """
    rng = np.random.default_rng(seed)
    nx, ny, nz = mesh.shape
    xn, yn, zn = mesh.xn, mesh.yn, mesh.zn
    xc, yc, zc = mesh.centres()
    L = (xn[-1] - xn[0], yn[-1] - yn[0], zn[-1] - zn[0])

    def smooth_field(X, Y, Z):
        f = np.zeros(np.broadcast(X, Y, Z).shape)
        for _ in range(modes):
            kx, ky, kz = rng.integers(1, 4, size=3)
            ph = rng.uniform(0, 2 * np.pi, size=3)
            amp = rng.uniform(-1, 1)
            f += amp * (np.sin(np.pi * kx * X / L[0] + ph[0]) *
                        np.sin(np.pi * ky * Y / L[1] + ph[1]) *
                        np.sin(np.pi * kz * Z / L[2] + ph[2]))
        # window that vanishes on the box boundary
        w = (np.sin(np.pi * X / L[0]) * np.sin(np.pi * Y / L[1]) * np.sin(np.pi * Z / L[2]))
        return f * w

    # psi_x lives on x-edges: (xc[i], yn[j], zn[k]); psi_y on (xn, yc, zn); psi_z on (xn, yn, zc)
    psx = smooth_field(*np.meshgrid(xc, yn, zn, indexing="ij"))   # (nx, ny+1, nz+1)
    psy = smooth_field(*np.meshgrid(xn, yc, zn, indexing="ij"))   # (nx+1, ny, nz+1)
    psz = smooth_field(*np.meshgrid(xn, yn, zc, indexing="ij"))   # (nx+1, ny+1, nz)
    # circulation along edges (psi * edge length) -> fluxes through faces
    dx, dy, dz = mesh.spacings()
    cx = psx * dx[:, None, None]
    cy = psy * dy[None, :, None]
    cz = psz * dz[None, None, :]
    # flux through x-face at node plane xn[i] (shape (nx+1, ny, nz))
    Fx = (cz[:, 1:, :] - cz[:, :-1, :]) - (cy[:, :, 1:] - cy[:, :, :-1])
    Fy = (cx[:, :, 1:] - cx[:, :, :-1]) - (cz[1:, :, :] - cz[:-1, :, :])
    Fz = (cy[1:, :, :] - cy[:-1, :, :]) - (cx[:, 1:, :] - cx[:, :-1, :])
    # velocity scale: |F| / face area
    DY, DZ = np.meshgrid(dy, dz, indexing="ij")
    DX2, DZ2 = np.meshgrid(dx, dz, indexing="ij")
    DX3, DY3 = np.meshgrid(dx, dy, indexing="ij")
    umax = max(np.abs(Fx / DY[None]).max(), np.abs(Fy / DX2[:, None, :]).max(),
               np.abs(Fz / DX3[:, :, None]).max(), 1e-300)
    s = amplitude / umax
    Fx, Fy, Fz = Fx * s, Fy * s, Fz * s

    def flat(a):
        return np.ascontiguousarray(a.transpose(2, 1, 0)).ravel()

    # interior faces indexed by low-side cell: x-face between i and i+1 sits at xn[i+1]
    fx = np.zeros((nx, ny, nz)); fx[:-1] = Fx[1:-1]
    fy = np.zeros((nx, ny, nz)); fy[:, :-1] = Fy[:, 1:-1]
    fz = np.zeros((nx, ny, nz)); fz[:, :, :-1] = Fz[:, :, 1:-1]
    return flat(fx), flat(fy), flat(fz), amplitude


def transport_matrix(mesh: BoxMesh, gamma: float, flux_seed: int, cfl: float = 0.5,
                     perm=None, amplitude: float = 1.0):
    """Upwind convection-diffusion matrix with a time term.

    Follows the "m1" placement of Process/NewUVW.f90:356-367 for every
    interior face s = (c1, c2), c1 the low-side cell and F the flux c1->c2:
      A12 = -min(F,0), A21 = max(F,0), D12 = D21 = Gamma*|S|/d
      diag(c1) += A21 + D21;   A(c2,c1) -= A21 + D21
      diag(c2) += A12 + D12;   A(c1,c2) -= A12 + D12
    then diag += V/dt (NewUVW.f90:447-449) with dt = cfl*h_min/max|u|.
    """
    nx, ny, nz = mesh.shape
    (ax, ay, az), _, _, vol = _face_coeffs(mesh)
    fx, fy, fz, umax = divergence_free_fluxes(mesh, flux_seed, amplitude)
    strides = (1, nx, nx * ny)
    hi_vals, lo_vals = [], []
    diag = np.zeros(mesh.n_cells)
    for a, F, s in zip((ax, ay, az), (fx, fy, fz), strides):
        exists = a > 0
        F = np.where(exists, F, 0.0)
        D = gamma * a
        A12 = -np.minimum(F, 0.0)
        A21 = np.maximum(F, 0.0)
        diag += A21 + D                       # diag(c1)
        diag += _shift_down(A12 + D, s)       # diag(c2)
        hi_vals.append(-(A12 + D))            # A(c1, c2) stored in row c1
        lo_vals.append(_shift_down(-(A21 + D), s))  # A(c2, c1) stored in row c2
    hmin = min(float(np.min(d)) for d in mesh.spacings())
    dt = cfl * hmin / umax
    diag += vol / dt
    m = _assemble(mesh.shape, lo_vals, hi_vals, diag, perm)
    m.meta.update(kind="transport", shape=mesh.shape, gamma=gamma, dt=dt)
    return m


def seeded_rhs(n: int, seed: int, subtract_mean: bool = True, perm=None) -> np.ndarray:
    b = np.random.default_rng(seed).uniform(-1.0, 1.0, size=n)
    if subtract_mean:
        b = b - b.mean()
    if perm is not None:
        out = np.empty_like(b)
        out[np.asarray(perm)] = b
        b = out
    return np.ascontiguousarray(b)


def pressure_system(n: int | tuple, seed: int = 1, ratio: float = 1.1, prec: int = 1,
                    niter: int = 2000, rel_tol: float = 1e-8, perm_seed: int | None = None
                    ) -> LinearSystem:
    """Channel pressure-correction system (32^3 / 128^3 / ... recipe).

    2*pi x 2 x pi box, tanh wall stretching in y with adjacent-cell ratio
    ``ratio``, cell 1 pinned, b ~ U(-1,1) from ``seed`` with its mean
    subtracted, x0 = 0 and the absolute tolerance ``rel_tol*||b||_2``.
    """
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    mesh = channel_mesh(nx, ny, nz, ratio)
    perm = random_permutation(mesh.n_cells, perm_seed) if perm_seed is not None else None
    m = pressure_matrix(mesh, pin_cell=1, perm=perm)
    b = seeded_rhs(m.n, seed, True, perm)
    tol = rel_tol * float(np.sqrt(np.dot(b, b)))
    return LinearSystem(m, b, np.zeros(m.n), tol, prec, niter,
                        dict(recipe="pressure", shape=(nx, ny, nz), seed=seed, ratio=ratio,
                             rel_tol=rel_tol, perm_seed=perm_seed,
                             generator_version=GENERATOR_VERSION))


def rescaled_pressure_values(m: TFlowSMatrix, seed: int, spread: tuple = (0.5, 2.0)) -> TFlowSMatrix:
    """The same pattern with the values a later outer iteration would give
    (the face coefficients A12 change with the interpolated velocities,
    Process/CalcPS.f90:77-94): every face pair (i, j) / (j, i) scaled by one
    seeded factor in ``spread``, each diagonal reset to minus its row's new
    off-diagonal sum plus the row's original remainder (the pin), so the
    matrix stays a symmetric M-matrix."""
    rng = np.random.default_rng(seed)
    rows = m.row_of_entry()
    cols = m.arow[:m.nnz].astype(np.int64) - 1
    off = rows != cols
    lo = np.flatnonzero(off & (cols < rows))
    mirror = m.mirror if m.mirror is not None else mirror_positions(m)
    scale = np.ones(m.nnz)
    f = rng.uniform(spread[0], spread[1], lo.size)
    scale[lo] = f
    scale[mirror[lo]] = f
    a = m.aval.copy()
    a[:m.nnz] *= scale
    dia = m.adia.astype(np.int64) - 1
    pin = m.aval[dia] + np.bincount(rows[off], weights=m.aval[:m.nnz][off], minlength=m.n)
    a[dia] = pin - np.bincount(rows[off], weights=a[:m.nnz][off], minlength=m.n)
    return m.with_values(a)


def wall_pressure_system(n: int | tuple, seed: int = 5, ratio: float = 1.05, max_aspect: float = 50.0,
                         prec: int = 2, niter: int = 20000, rel_tol: float = 1e-8) -> LinearSystem:
    """Channel pressure-correction system with strong wall clustering.

    2*pi x 2 x pi box; y-nodes grow geometrically by ``ratio`` from each wall
    (geometric_wall_nodes) with a first cell of dx / ``max_aspect``, so the
    wall cells reach an aspect ratio of about ``max_aspect``:1 against the
    streamwise spacing dx = 2*pi/nx.  Cell 1 pinned, b ~ U(-1,1) with its
    mean removed, DIC preconditioning by default.
    """
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    dx = 2.0 * np.pi / nx
    mesh = wall_pressure_mesh((nx, ny, nz), ratio, max_aspect)
    m = pressure_matrix(mesh, pin_cell=1)
    b = seeded_rhs(m.n, seed, True)
    tol = rel_tol * float(np.sqrt(np.dot(b, b)))
    dy = np.diff(mesh.yn)
    return LinearSystem(m, b, np.zeros(m.n), tol, prec, niter,
                        dict(recipe="wall_pressure", shape=(nx, ny, nz), seed=seed, ratio=ratio,
                             max_aspect=float(dx / dy.min()), rel_tol=rel_tol,
                             generator_version=GENERATOR_VERSION))


def momentum_system(n: int | tuple, seed: int = 2, gamma: float = 1e-3, prec: int = 1,
                    niter: int = 2000, rel_tol: float = 1e-8, renumber: bool = True,
                    flux_seed: int | None = None) -> LinearSystem:
    """Nonsymmetric momentum-type system on a unit cube.

    Upwind convection from a seeded divergence-free velocity of amplitude 1,
    diffusion ``gamma``, time term for CFL 0.5, random cell renumbering from
    ``seed``; b ~ U(-1,1).
    """
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    mesh = BoxMesh(np.linspace(0, 1, nx + 1), np.linspace(0, 1, ny + 1), np.linspace(0, 1, nz + 1))
    perm = random_permutation(mesh.n_cells, seed) if renumber else None
    m = transport_matrix(mesh, gamma, flux_seed if flux_seed is not None else 1000 + seed,
                         perm=perm)
    b = seeded_rhs(m.n, seed, False, perm)
    tol = rel_tol * float(np.sqrt(np.dot(b, b)))
    return LinearSystem(m, b, np.zeros(m.n), tol, prec, niter,
                        dict(recipe="momentum", shape=(nx, ny, nz), seed=seed, gamma=gamma,
                             rel_tol=rel_tol, renumber=renumber,
                             generator_version=GENERATOR_VERSION))


def transport_batch(n: int = 96, count: int = 5, seed: int = 3, prec: int = 1, niter: int = 2000,
                    rel_tol: float = 1e-8, renumber: bool = True) -> list[LinearSystem]:
    """Systems sharing one pattern: the momentum components and scalars of
    one mesh, as NewUVW.f90:506 (u, v, w) and CalcZmix.f90:394 (Zmix, Y)
    solve them one call each.

    One n^3 unit-cube mesh and one random cell renumbering (seed 100*seed);
    system k (seed 100*seed + k) has its own divergence-free convecting
    velocity, its own diffusion coefficient Gamma_k = 10**U(-3.5, -2.5) and
    its own right-hand side b ~ U(-1,1); CFL 0.5 time term as in
    transport_matrix.
    """
    mesh = BoxMesh(np.linspace(0, 1, n + 1), np.linspace(0, 1, n + 1), np.linspace(0, 1, n + 1))
    base = 100 * seed
    perm = random_permutation(mesh.n_cells, base) if renumber else None
    names = ("u", "v", "w", "zmix", "y")
    out = []
    for k in range(count):
        sk = base + k
        gamma = float(10.0 ** np.random.default_rng(sk).uniform(-3.5, -2.5))
        m = transport_matrix(mesh, gamma, sk, perm=perm)
        b = seeded_rhs(m.n, sk, False, perm)
        tol = rel_tol * float(np.sqrt(np.dot(b, b)))
        out.append(LinearSystem(m, b, np.zeros(m.n), tol, prec, niter,
                                dict(recipe="transport_batch", name=names[k % len(names)], shape=(n, n, n),
                                     seed=sk, gamma=gamma, rel_tol=rel_tol, renumber=renumber,
                                     generator_version=GENERATOR_VERSION)))
    return out


def uneven_pressure_batch(count: int = 64, seed: int = 3, n_min: float = 1e4, n_max: float = 5e5,
                          prec: int = 1, niter: int = 5000, rel_tol: float = 1e-8) -> list[LinearSystem]:
    """Independent pressure-correction systems of uneven sizes.

    Cell counts N_k are log-uniform on [n_min, n_max] and the box of system
    k has seeded aspect ratios a ~ U(1,4)^3 (rng seed ``seed``): its
    dimensions are round(a_i * (N_k / prod a)^(1/3)), at least 2.  Unit
    cubic cells, cell 1 pinned (pressure_matrix), b ~ U(-1,1) from seed
    100*seed + k with its mean removed, tol = rel_tol * ||b||.
    """
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(n_min), np.log(n_max), size=count))
    aspect = rng.uniform(1.0, 4.0, size=(count, 3))
    out = []
    for k in range(count):
        a = aspect[k]
        s = (sizes[k] / float(np.prod(a))) ** (1.0 / 3.0)
        nx, ny, nz = (max(2, int(round(v * s))) for v in a)
        mesh = BoxMesh(np.arange(nx + 1, dtype=np.float64), np.arange(ny + 1, dtype=np.float64),
                       np.arange(nz + 1, dtype=np.float64))
        m = pressure_matrix(mesh, pin_cell=1)
        sk = 100 * seed + k
        b = seeded_rhs(m.n, sk, True)
        tol = rel_tol * float(np.sqrt(np.dot(b, b)))
        out.append(LinearSystem(m, b, np.zeros(m.n), tol, prec, niter,
                                dict(recipe="uneven_pressure_batch", shape=(nx, ny, nz), seed=sk,
                                     rel_tol=rel_tol, generator_version=GENERATOR_VERSION)))
    return out


def from_scipy(A, meta=None) -> TFlowSMatrix:
    """Wrap a square scipy sparse matrix (structurally symmetric pattern with a
    stored diagonal) in the T-FlowS layout, columns sorted per row."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    nnz = A.nnz
    acol = (A.indptr + 1).astype(np.int32)
    arow = np.zeros(nnz + 1, np.int32)
    arow[:nnz] = A.indices + 1
    aval = np.zeros(nnz + 1)
    aval[:nnz] = A.data
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    is_d = A.indices == rows
    if is_d.sum() != n:
        raise ValueError("every row needs a stored diagonal entry")
    adia = np.empty(n, np.int32)
    adia[rows[is_d]] = np.nonzero(is_d)[0] + 1
    m = TFlowSMatrix(n, nnz + 1, acol, arow, adia, aval, meta=dict(meta or {}))
    try:
        m.mirror = mirror_positions(m)
    except ValueError:
        m.mirror = None
    return m


# --------------------------------------------------------------------------
# face-based assembly (TopolM) and perturbed geometry
# --------------------------------------------------------------------------
@dataclass
class Topology:
    """Result of :func:`topolm`: the CSR pattern plus the positions of each
    face's two off-diagonal entries (SidAij) and of every diagonal."""
    n: int
    nonz: int
    acol: np.ndarray      # int32 (N+1), 1-based
    arow: np.ndarray      # int32 (NONZ), 1-based, last slot 0
    adia: np.ndarray      # int32 (N), 1-based
    sidaij: np.ndarray    # int64 (2, NS) 0-based positions of A(c1,c2), A(c2,c1)


def topolm(n: int, c1: np.ndarray, c2: np.ndarray) -> Topology:
    """Pattern of the system matrix from the face list (Process/TopolM.f90).

    ``c1``, ``c2`` are 1-based cells of the interior faces (SideC(1:2,s)
    with c2 > 0).  Each row holds the diagonal plus one entry per face of
    the cell, columns sorted ascending (ISORT, TopolM.f90:72-74); Adia is the
    diagonal position (:75-77) and SidAij(1,s) / SidAij(2,s) the positions of
    A(c1,c2) / A(c2,c1) (:81-94).  NONZERO = nnz + 1 (:49).
    """
    c1 = np.asarray(c1, np.int64) - 1
    c2 = np.asarray(c2, np.int64) - 1
    ns = len(c1)
    rows = np.concatenate([np.arange(n), c1, c2])
    cols = np.concatenate([np.arange(n), c2, c1])
    order = np.lexsort((cols, rows))
    nnz = len(rows)
    pos = np.empty(nnz, np.int64)
    pos[order] = np.arange(nnz)
    counts = np.bincount(rows, minlength=n)
    acol = np.empty(n + 1, np.int32)
    acol[0] = 1
    acol[1:] = 1 + np.cumsum(counts)
    arow = np.zeros(nnz + 1, np.int32)
    arow[:nnz] = cols[order] + 1
    adia = (pos[:n] + 1).astype(np.int32)
    sid = np.stack([pos[n:n + ns], pos[n + ns:]])
    return Topology(n, nnz + 1, acol, arow, adia, sid)


def box_faces(shape):
    """Interior faces (c1 < c2, 1-based, natural numbering) of a box of cells,
    x-faces, then y-faces, then z-faces."""
    nx, ny, nz = shape
    c = np.arange(nx * ny * nz).reshape(nz, ny, nx)
    f = [(c[:, :, :-1].ravel(), c[:, :, 1:].ravel()),
         (c[:, :-1, :].ravel(), c[:, 1:, :].ravel()),
         (c[:-1, :, :].ravel(), c[1:, :, :].ravel())]
    return (np.concatenate([a for a, _ in f]) + 1, np.concatenate([b for _, b in f]) + 1,
            np.concatenate([np.full(len(a), d) for d, (a, _) in enumerate(f)]))


@dataclass
class FaceGeometry:
    shape: tuple
    c1: np.ndarray            # 1-based low-side cell of each interior face
    c2: np.ndarray            # 1-based high-side cell
    coef: np.ndarray          # a_f = |S|^2 / (S.d), FP64
    centres: np.ndarray       # (N, 3) cell centres
    pin_coef: float           # boundary term added to the pinned cell


def perturbed_box_geometry(n: int, ratio: float = 1.05, perturb: float = 0.2, seed: int = 4) -> FaceGeometry:
    """Hex box with tanh wall stretching in y and every interior node moved by
    up to ``perturb`` times the local spacing in each direction (seeded).
    Face area vectors come from the quad diagonals, S = (n3-n1) x (n4-n2)/2,
    cell centres are node averages, and the pressure coefficient of a face is
    a_f = |S|^2/(S.d) with d the centre-to-centre vector -- in FP64, the
    form of Process/Calc3.f90:24, 46 and CalcPS.f90:77-78."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, 1.0, n + 1)
    ys = tanh_nodes(n, 1.0, ratio)
    zs = np.linspace(0.0, 1.0, n + 1)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")          # (n+1)^3, index [i, j, k]
    P = np.stack([X, Y, Z], axis=-1)
    h = np.stack(np.meshgrid(np.gradient(xs), np.gradient(ys), np.gradient(zs), indexing="ij"), axis=-1)
    disp = rng.uniform(-perturb, perturb, P.shape) * h
    disp[[0, -1], :, :, 0] = 0.0
    disp[:, [0, -1], :, 1] = 0.0
    disp[:, :, [0, -1], 2] = 0.0
    P = P + disp
    # cell centres (node average), cell index = i + n*(j + n*k)
    C = 0.125 * (P[:-1, :-1, :-1] + P[1:, :-1, :-1] + P[:-1, 1:, :-1] + P[1:, 1:, :-1] +
                 P[:-1, :-1, 1:] + P[1:, :-1, 1:] + P[:-1, 1:, 1:] + P[1:, 1:, 1:])
    cen = C.transpose(2, 1, 0, 3).reshape(-1, 3)

    def quad_area(a, b, c, d):    # nodes in cyclic order
        return 0.5 * np.cross(c - a, d - b)

    coef, c1s, c2s = [], [], []
    cid = np.arange(n ** 3).reshape(n, n, n, order="F")        # [i, j, k] -> i + n j + n^2 k
    # x-faces at node plane i+1 between cells (i, j, k) and (i+1, j, k)
    S = quad_area(P[1:-1, :-1, :-1], P[1:-1, 1:, :-1], P[1:-1, 1:, 1:], P[1:-1, :-1, 1:])
    d = C[1:] - C[:-1]
    coef.append((np.sum(S * S, -1) / np.sum(S * d, -1)).ravel(order="F"))
    c1s.append(cid[:-1].ravel(order="F")); c2s.append(cid[1:].ravel(order="F"))
    # y-faces
    S = quad_area(P[:-1, 1:-1, :-1], P[:-1, 1:-1, 1:], P[1:, 1:-1, 1:], P[1:, 1:-1, :-1])
    d = C[:, 1:] - C[:, :-1]
    coef.append((np.sum(S * S, -1) / np.sum(S * d, -1)).ravel(order="F"))
    c1s.append(cid[:, :-1].ravel(order="F")); c2s.append(cid[:, 1:].ravel(order="F"))
    # z-faces
    S = quad_area(P[:-1, :-1, 1:-1], P[1:, :-1, 1:-1], P[1:, 1:, 1:-1], P[:-1, 1:, 1:-1])
    d = C[:, :, 1:] - C[:, :, :-1]
    coef.append((np.sum(S * S, -1) / np.sum(S * d, -1)).ravel(order="F"))
    c1s.append(cid[:, :, :-1].ravel(order="F")); c2s.append(cid[:, :, 1:].ravel(order="F"))
    a = np.concatenate(coef)
    if np.any(a <= 0):
        raise ValueError("perturbation too strong: a face has S.d <= 0")
    # pin: one boundary face of cell 1 (its x=0 face), A/(distance to face)
    S0 = quad_area(P[0, 0, 0], P[0, 1, 0], P[0, 1, 1], P[0, 0, 1])
    d0 = C[0, 0, 0] - 0.25 * (P[0, 0, 0] + P[0, 1, 0] + P[0, 1, 1] + P[0, 0, 1])
    pin = float(np.dot(S0, S0) / abs(np.dot(S0, d0)))
    return FaceGeometry((n, n, n), np.concatenate(c1s) + 1, np.concatenate(c2s) + 1, a, cen, pin)


def renumber_faces(geo: FaceGeometry, seed: int) -> tuple[FaceGeometry, np.ndarray]:
    """Random cell renumbering: natural cell c becomes perm[c]."""
    n = len(geo.centres)
    perm = random_permutation(n, seed)
    cen = np.empty_like(geo.centres)
    cen[perm] = geo.centres
    return (FaceGeometry(geo.shape, perm[geo.c1 - 1] + 1, perm[geo.c2 - 1] + 1, geo.coef, cen, geo.pin_coef),
            perm)


def assemble_face_values(topo: Topology, c1, c2, coef, pin_cell: int, pin_coef: float) -> np.ndarray:
    """Pressure-correction values on a pattern from face coefficients:
    A(c1,c2) = A(c2,c1) = -a_f and a_f added to both diagonals
    (CalcPS.f90:77-94), plus pin_coef on the diagonal of pin_cell."""
    aval = np.zeros(topo.nonz)
    np.add.at(aval, topo.sidaij[0], -coef)
    np.add.at(aval, topo.sidaij[1], -coef)
    np.add.at(aval, topo.adia[np.asarray(c1) - 1] - 1, coef)
    np.add.at(aval, topo.adia[np.asarray(c2) - 1] - 1, coef)
    if pin_cell:
        aval[topo.adia[pin_cell - 1] - 1] += pin_coef
    return aval


def fourier_target(centres: np.ndarray, seed: int, modes: int = 4) -> np.ndarray:
    """Smooth target field: a sum of `modes` seeded Fourier modes."""
    rng = np.random.default_rng(seed)
    t = np.zeros(len(centres))
    for _ in range(modes):
        k = rng.integers(1, 4, size=3)
        ph = rng.uniform(0, 2 * np.pi, size=3)
        amp = rng.uniform(-1, 1)
        t += amp * np.prod(np.cos(np.pi * k * centres + ph), axis=1)
    return t


# --------------------------------------------------------------------------
# inputs of the pressure-correction assembly (Process/CalcPS.f90:51-202)
# --------------------------------------------------------------------------
BC_CODES = {"inflow": 1, "wall": 2, "outflow": 3, "symmetry": 4, "convect": 5, "wallfl": 6, "buffer": 11,
            "pressure": 12}   # Library/allp_mod.f90:16-24

CHANNEL_BC = {"xmin": "inflow", "xmax": "outflow", "ymin": "wall", "ymax": "wall", "zmin": "symmetry",
              "zmax": "symmetry"}


@dataclass
class CalcPSCase:
    """Everything the pressure-correction assembly reads, in the reference's
    layout: ``sidec`` is SideC(0:2,NS) as a (3, NS) int32 array (row 0
    unused), ``sidaij`` SidAij(1:2,NS) as (2, NS) int32, 1-based, 0 on
    boundary faces; the TopolM pattern (acol, arow, adia); per face Sx..Dz
    and f; and the cell arrays declared (-NbC:NC) as arrays of NbC+NC+1
    entries with x(c) at index c + NbC (boundary cells c = -1..-NbC, one
    per boundary face, numbered in face order)."""
    shape: tuple
    nc: int
    nbc: int
    ns: int
    nonz: int
    nmat: int
    dt: float
    sidec: np.ndarray
    sidaij: np.ndarray
    acol: np.ndarray
    arow: np.ndarray
    adia: np.ndarray
    typebc: np.ndarray
    sx: np.ndarray
    sy: np.ndarray
    sz: np.ndarray
    dx: np.ndarray
    dy: np.ndarray
    dz: np.ndarray
    f: np.ndarray
    volume: np.ndarray
    asave: np.ndarray
    rho_n: np.ndarray
    rho_o: np.ndarray
    rho_oo: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    p: np.ndarray
    px: np.ndarray
    py: np.ndarray
    pz: np.ndarray
    xc: np.ndarray | None = None      # (NbC+NC+1, 3) positions of the cells (face centres for boundary cells)

    def matrix(self, aval: np.ndarray) -> TFlowSMatrix:
        """The assembled system matrix with values ``aval`` (NONZ)."""
        return TFlowSMatrix(self.nc, self.nonz, self.acol, self.arow, self.adia,
                            np.ascontiguousarray(aval, np.float64), None, {"kind": "calcps", "shape": self.shape})

    def fields(self) -> dict:
        return {k: getattr(self, k) for k in ("asave", "rho_n", "rho_o", "rho_oo", "u", "v", "w", "p", "px",
                                              "py", "pz")}


def calcps_case(shape, seed: int = 7, ratio: float = 1.05, perturb: float = 0.15, bc: dict | None = None,
                face_seed: int | None = None, nmat: int = 1, dt: float = 1e-2) -> CalcPSCase:
    """A channel box (2 pi x 2 x pi, tanh wall stretching in y, interior
    nodes moved by up to ``perturb`` of the local spacing) with its faces,
    boundary faces on the six sides with the types of ``bc`` (keys xmin ..
    zmax, values from BC_CODES), and smooth seeded fields.

    Faces: S is the quad-diagonal area vector pointing from c1 to c2, d the
    centre-to-centre vector (face centre for a boundary cell), f the weight
    of c1 in the face interpolation (1 on boundary faces); volumes from the
    divergence theorem over the same faces.  Interior faces come first
    (x, y, z), then the boundary faces side by side, unless ``face_seed``
    shuffles all of them.  BUFFER cells get a volume and Asave like an
    interior cell's; other boundary cells have zero volume and Asave.
    """
    nx, ny, nz = shape
    bc = dict(CHANNEL_BC, **(bc or {}))
    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, 2.0 * np.pi, nx + 1)
    ys = tanh_nodes(ny, 2.0, ratio)
    zs = np.linspace(0.0, np.pi, nz + 1)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    P = np.stack([X, Y, Z], axis=-1)
    h = np.stack(np.meshgrid(np.gradient(xs), np.gradient(ys), np.gradient(zs), indexing="ij"), axis=-1)
    disp = rng.uniform(-perturb, perturb, P.shape) * h
    disp[[0, -1], :, :, 0] = 0.0
    disp[:, [0, -1], :, 1] = 0.0
    disp[:, :, [0, -1], 2] = 0.0
    P = P + disp
    C = 0.125 * (P[:-1, :-1, :-1] + P[1:, :-1, :-1] + P[:-1, 1:, :-1] + P[1:, 1:, :-1] +
                 P[:-1, :-1, 1:] + P[1:, :-1, 1:] + P[:-1, 1:, 1:] + P[1:, 1:, 1:])
    cid = np.arange(nx * ny * nz).reshape(nx, ny, nz, order="F")     # [i, j, k] -> i + nx (j + ny k)
    cen = C.reshape(-1, 3, order="F")

    def quad(a, b, c, d):
        return 0.5 * np.cross(c - a, d - b), 0.25 * (a + b + c + d)

    def xplane(i):   # +x area vectors of node plane i, (ny, nz, 3)
        return quad(P[i, :-1, :-1], P[i, 1:, :-1], P[i, 1:, 1:], P[i, :-1, 1:])

    def yplane(j):
        return quad(P[:-1, j, :-1], P[:-1, j, 1:], P[1:, j, 1:], P[1:, j, :-1])

    def zplane(k):
        return quad(P[:-1, :-1, k], P[1:, :-1, k], P[1:, 1:, k], P[:-1, 1:, k])

    S_l, xf_l, c1_l, c2_l = [], [], [], []
    # interior faces
    S, xf = quad(P[1:-1, :-1, :-1], P[1:-1, 1:, :-1], P[1:-1, 1:, 1:], P[1:-1, :-1, 1:])
    S_l.append(S.reshape(-1, 3, order="F")); xf_l.append(xf.reshape(-1, 3, order="F"))
    c1_l.append(cid[:-1].ravel(order="F")); c2_l.append(cid[1:].ravel(order="F"))
    S, xf = quad(P[:-1, 1:-1, :-1], P[:-1, 1:-1, 1:], P[1:, 1:-1, 1:], P[1:, 1:-1, :-1])
    S_l.append(S.reshape(-1, 3, order="F")); xf_l.append(xf.reshape(-1, 3, order="F"))
    c1_l.append(cid[:, :-1].ravel(order="F")); c2_l.append(cid[:, 1:].ravel(order="F"))
    S, xf = quad(P[:-1, :-1, 1:-1], P[1:, :-1, 1:-1], P[1:, 1:, 1:-1], P[:-1, 1:, 1:-1])
    S_l.append(S.reshape(-1, 3, order="F")); xf_l.append(xf.reshape(-1, 3, order="F"))
    c1_l.append(cid[:, :, :-1].ravel(order="F")); c2_l.append(cid[:, :, 1:].ravel(order="F"))
    n_int = sum(len(a) for a in c1_l)
    # boundary faces, outward from the adjacent cell
    sides = []
    for name, (S, xf), cells, sign in (
            ("xmin", xplane(0), cid[0], -1.0), ("xmax", xplane(nx), cid[-1], 1.0),
            ("ymin", yplane(0), cid[:, 0], -1.0), ("ymax", yplane(ny), cid[:, -1], 1.0),
            ("zmin", zplane(0), cid[:, :, 0], -1.0), ("zmax", zplane(nz), cid[:, :, -1], 1.0)):
        S_l.append(sign * S.reshape(-1, 3, order="F")); xf_l.append(xf.reshape(-1, 3, order="F"))
        c1_l.append(cells.ravel(order="F")); c2_l.append(np.full(cells.size, -1))
        sides.append(np.full(cells.size, BC_CODES[bc[name]]))
    S = np.concatenate(S_l); xf = np.concatenate(xf_l)
    c1 = np.concatenate(c1_l); c2 = np.concatenate(c2_l)
    bcode = np.concatenate([np.zeros(n_int, np.int64)] + sides)
    ns = len(c1)
    order = np.random.default_rng(face_seed).permutation(ns) if face_seed is not None else np.arange(ns)
    S, xf, c1, c2, bcode = S[order], xf[order], c1[order], c2[order], bcode[order]
    bnd = c2 < 0
    nbc = int(bnd.sum())
    nc = nx * ny * nz
    # boundary cells numbered -1 .. -NbC in face order
    c2 = c2.copy()
    c2[bnd] = -np.arange(1, nbc + 1)
    c1f = c1 + 1                                   # 1-based
    c2f = np.where(bnd, c2, c2 + 1)
    # geometry
    xc1 = cen[c1]
    xc2 = np.where(bnd[:, None], xf, cen[np.where(bnd, 0, c2)])
    d = xc2 - xc1
    a1 = np.linalg.norm(xf - xc1, axis=1)
    a2 = np.linalg.norm(xc2 - xf, axis=1)
    f = np.where(bnd, 1.0, a2 / np.where(a1 + a2 > 0, a1 + a2, 1.0))
    third = np.einsum("ij,ij->i", xf, S) / 3.0
    vol_cells = np.bincount(c1, weights=third, minlength=nc) - np.bincount(c2[~bnd], weights=third[~bnd],
                                                                            minlength=nc)
    # pattern of the interior faces (TopolM), in face order
    interior = ~bnd
    topo = topolm(nc, c1f[interior], c2f[interior])
    sidaij = np.zeros((2, ns), np.int32)
    sidaij[:, interior] = (topo.sidaij + 1).astype(np.int32)
    sidec = np.zeros((3, ns), np.int32)
    sidec[1] = c1f
    sidec[2] = c2f
    # cell arrays (-NbC:NC)
    ntot = nbc + nc + 1
    at = lambda c: c + nbc                         # noqa: E731  index of cell c
    typebc = np.zeros(ntot, np.int32)
    typebc[at(c2[bnd])] = bcode[bnd]
    volume = np.zeros(ntot)
    volume[at(np.arange(1, nc + 1))] = vol_cells
    is_buf = bnd & (bcode == BC_CODES["buffer"])
    volume[at(c2[is_buf])] = vol_cells[c1[is_buf]] * rng.uniform(0.9, 1.1, int(is_buf.sum()))
    bpos = np.zeros((nbc, 3))                      # positions of the cells (-NbC:NC)
    bpos[-c2[bnd] - 1] = xf[bnd]
    allc = np.concatenate([bpos[::-1], np.zeros((1, 3)), cen])

    nonbuf_bc = np.zeros(ntot, bool)
    nonbuf_bc[at(c2[bnd & ~is_buf])] = True
    volume[nonbuf_bc] = 0.0
    asave, rho_n, rho_o, rho_oo, u, v, w, p, px, py, pz = _smooth_fields(rng, allc, nbc, volume, nonbuf_bc, dt)
    return CalcPSCase(tuple(shape), nc, nbc, ns, topo.nonz, nmat, dt, sidec, sidaij, topo.acol, topo.arow,
                      topo.adia, typebc, np.ascontiguousarray(S[:, 0]), np.ascontiguousarray(S[:, 1]),
                      np.ascontiguousarray(S[:, 2]), np.ascontiguousarray(d[:, 0]), np.ascontiguousarray(d[:, 1]),
                      np.ascontiguousarray(d[:, 2]), f, volume, asave, rho_n, rho_o, rho_oo, u, v, w, p, px, py, pz,
                      allc)


def _smooth_fields(rng, allc, nbc, volume, nonbuf_bc, dt):
    """Smooth seeded cell fields (-NbC:NC) at the positions ``allc``: Asave
    (zero on non-BUFFER boundary cells), rho n/o/oo, u, v, w, p, px, py,
    pz; cell 0 zero."""
    ntot = len(allc)

    def smooth(mean, amp, k):
        g = np.full(ntot, float(mean))
        for _ in range(3):
            kv = rng.uniform(0.5, 2.0, 3) * k
            ph = rng.uniform(0, 2 * np.pi)
            g += amp * np.sin(allc @ kv + ph) / 3.0
        return g

    rho_n = smooth(1.0, 0.05, 1.0)
    rho_o = rho_n * (1.0 + 1e-3 * rng.standard_normal(ntot))
    rho_oo = rho_n * (1.0 + 1e-3 * rng.standard_normal(ntot))
    u = smooth(1.0, 0.3, 1.0)
    v = smooth(0.0, 0.2, 1.3)
    w = smooth(0.0, 0.2, 0.7)
    p = smooth(0.0, 0.5, 1.1)
    px, py, pz = smooth(0.0, 0.4, 1.0), smooth(0.0, 0.4, 1.2), smooth(0.0, 0.4, 0.9)
    asave = volume / dt * (1.5 + 0.2 * rng.uniform(0, 1, ntot))
    asave[nonbuf_bc] = 0.0
    asave[nbc] = 0.0                              # cell 0 is not a cell
    for a in (rho_n, rho_o, rho_oo, u, v, w, p, px, py, pz):
        a[nbc] = 0.0
    return asave, rho_n, rho_o, rho_oo, u, v, w, p, px, py, pz


def case_from_mesh(mesh, typbou, seed: int = 7, nmat: int = 1, dt: float = 1e-2) -> CalcPSCase:
    """A CalcPSCase on a mesh read from T-FlowS files (krylov.Mesh): its
    faces, pattern (TopolM), centre distances (Calc3), volumes and cell
    positions, TypeBC from bcmark and the boundary types ``typbou`` of the
    .b file, and smooth seeded fields as calcps_case makes them."""
    rng = np.random.default_rng(seed)
    nbc = mesh.nbc
    typebc = mesh.typebc(typbou)
    ntot = nbc + mesh.nc + 1
    nonbuf_bc = np.zeros(ntot, bool)
    nonbuf_bc[:nbc] = typebc[:nbc] != BC_CODES["buffer"]
    volume = mesh.volume.copy()
    volume[nonbuf_bc] = 0.0
    allc = np.stack([mesh.xc, mesh.yc, mesh.zc], axis=1)
    fields = _smooth_fields(rng, allc, nbc, volume, nonbuf_bc, dt)
    return CalcPSCase(("mesh", mesh.nc), mesh.nc, nbc, mesh.ns, mesh.nonz, nmat, dt,
                      mesh.sidec.astype(np.int32), mesh.sidaij.astype(np.int32), mesh.acol, mesh.arow, mesh.adia,
                      typebc, mesh.sx.copy(), mesh.sy.copy(), mesh.sz.copy(), mesh.dx.copy(), mesh.dy.copy(),
                      mesh.dz.copy(), mesh.f.copy(), volume, *fields, allc)


# --------------------------------------------------------------------------
# Momentum assembly inputs (Process/NewUVW.f90)
# --------------------------------------------------------------------------
FLAGS = {"AB": 1, "CN": 2, "FI": 3, "LIN": 4, "PAR": 5, "YES": 10, "NO": 11}   # Process/IniPar.f90:22-32

UVW_GRADS = ("ux", "uy", "uz", "vx", "vy", "vz", "wx", "wy", "wz")

# newuvw's arguments for var = 1, 2, 3 as Process/Processor.f90:314-340
# passes them: the component, its seven gradients (dUidi, dUidj, dUidk,
# dUjdi, dUkdi, dUjdj, dUkdk), the order of the S and D components, Pi
UVW_ARGS = {
    1: ("u", ("ux", "uy", "uz", "vx", "wx", "vy", "wz"), (0, 1, 2), "px"),
    2: ("v", ("vy", "vx", "vz", "uy", "wy", "ux", "wz"), (1, 0, 2), "py"),
    3: ("w", ("wz", "wx", "wy", "uz", "vz", "ux", "vy"), (2, 0, 1), "pz"),
}


def uvw_fields(case: CalcPSCase, seed: int = 5, nmaterials: int = 1, linear: np.ndarray | None = None,
               zero_flux_types=("wall", "wallfl", "symmetry")) -> dict:
    """What the momentum assembly reads beyond a CalcPSCase: Scoef and fF
    (as Process/Calc3.f90:23-70 forms them from the face geometry), the
    material of every cell (1 .. nmaterials, slabs in x; a boundary cell
    takes its face's cell's), a face flux Flux_u of the case's velocity
    (zero on the faces of ``zero_flux_types``), the nine velocity gradients
    ux .. wz, the old levels u_o, u_oo, .. and sources src_u, .. (cell
    arrays (-NbC:NC)).  With ``linear`` (3x3 G, rows u, v, w) the velocity
    n-level is the linear field G x (keys u, v, w) and the gradients are the
    constants G[i, j]; otherwise they are smooth seeded fields."""
    rng = np.random.default_rng(seed)
    nbc, nc, ns = case.nbc, case.nc, case.ns
    ntot = nbc + nc + 1
    sx, sy, sz, dx, dy, dz = case.sx, case.sy, case.sz, case.dx, case.dy, case.dz
    scoef = ((sx * sx + sy * sy) + sz * sz) / ((dx * sx + dy * sy) + dz * sz)
    c1 = case.sidec[1].astype(np.int64)
    c2 = case.sidec[2].astype(np.int64)
    bnd = c2 < 0
    ff = case.f.copy()
    tb = case.typebc[nbc + c2]
    ff[bnd & (tb != BC_CODES["buffer"])] = 1.0
    xc = case.xc
    material = np.ones(ntot, np.int32)
    if nmaterials > 1:
        edges = np.linspace(xc[nbc + 1:, 0].min(), xc[nbc + 1:, 0].max() + 1e-12, nmaterials + 1)
        material[nbc + 1:] = np.clip(np.searchsorted(edges, xc[nbc + 1:, 0], side="right"), 1, nmaterials)
        material[nbc + c2[bnd]] = material[nbc + c1[bnd]]
    out = {"scoef": scoef, "ff": ff, "material": material, "nmaterials": nmaterials}
    if linear is not None:
        G = np.asarray(linear, float)
        vel = {k: xc @ G[i] for i, k in enumerate("uvw")}
        for i, a in enumerate("uvw"):
            for j, b in enumerate("xyz"):
                out[a + b] = np.full(ntot, G[i, j])
        out.update(vel)
    else:
        vel = {"u": case.u, "v": case.v, "w": case.w}
        for k in UVW_GRADS:
            out[k] = 0.3 * rng.standard_normal(ntot) + rng.uniform(-1, 1)
    f = case.f
    ue = [f * vel[k][nbc + c1] + (1.0 - f) * vel[k][nbc + c2] for k in "uvw"]
    flux = (ue[0] * sx + ue[1] * sy) + ue[2] * sz
    for t in zero_flux_types:
        flux[bnd & (tb == BC_CODES[t])] = 0.0
    out["flux_u"] = flux
    for k in "uvw":
        out[k + "_o"] = vel[k] * (1.0 + 0.02 * rng.standard_normal(ntot))
        out[k + "_oo"] = vel[k] * (1.0 + 0.02 * rng.standard_normal(ntot))
        out["src_" + k] = 0.1 * rng.standard_normal(ntot)
    return out


def zmix_fields(case: CalcPSCase, seed: int = 6, nmaterials: int = 1, linear: np.ndarray | None = None,
                zero_flux_types=("wall", "wallfl", "symmetry")) -> dict:
    """The inputs of the mixture-fraction assembly (Process/CalcZmix.f90):
    those of uvw_fields (Scoef, fF, material, Flux_u) and the scalar: z
    (n-level, in [0, 1]), z_o, z_oo, its gradients zx, zy, zz and a source
    src_z, cell arrays (-NbC:NC).  With ``linear`` (3 numbers g) z = 0.5 +
    g . x and the gradients are the constants g."""
    out = uvw_fields(case, seed=seed, nmaterials=nmaterials, zero_flux_types=zero_flux_types)
    rng = np.random.default_rng(seed + 1000)
    ntot = case.nbc + case.nc + 1
    xc = case.xc
    if linear is not None:
        g = np.asarray(linear, float)
        out["z"] = 0.5 + xc @ g
        for k, gk in zip(("zx", "zy", "zz"), g):
            out[k] = np.full(ntot, gk)
    else:
        ph = rng.uniform(0, 2 * np.pi, 3)
        out["z"] = 0.5 + 0.4 * np.sin(xc[:, 0] + ph[0]) * np.cos(0.7 * xc[:, 1] + ph[1])
        for k in ("zx", "zy", "zz"):
            out[k] = 0.2 * rng.standard_normal(ntot)
    out["z_o"] = np.clip(out["z"] + 0.01 * rng.standard_normal(ntot), 0.0, 1.0)
    out["z_oo"] = np.clip(out["z"] + 0.01 * rng.standard_normal(ntot), 0.0, 1.0)
    out["src_z"] = 0.05 * rng.standard_normal(ntot)
    return out


def zmix_args(case: CalcPSCase, zf: dict) -> dict:
    """The cell and face arrays of the mixture-fraction assembly."""
    return {"rho_n": case.rho_n, "rho_o": case.rho_o, "rho_oo": case.rho_oo, "z_n": zf["z"], "z_o": zf["z_o"],
            "z_oo": zf["z_oo"], "source": zf["src_z"], "g": [zf["zx"], zf["zy"], zf["zz"]],
            "flux_u": zf["flux_u"], "flux": zf.get("flux")}


def uvw_params(nmaterials: int = 1, **kw) -> dict:
    """Flags and constants of one newuvw call (defaults: fully implicit
    convection, diffusion and cross diffusion, two-level inertia, central
    convection without the upwind part, URF 0.8)."""
    p = {"var": 1, "convec": FLAGS["FI"], "diffus": FLAGS["FI"], "cross": FLAGS["FI"], "inert": FLAGS["LIN"],
         "visc": 1e-2, "dt": 1e-2, "urf": 0.8, "blend": [FLAGS["NO"]] * nmaterials,
         "blend_tem": [FLAGS["NO"]] * nmaterials, "urfc": [1.0] * nmaterials, "visc_rho": 0}
    p.update(kw)
    return p


def uvw_args(case: CalcPSCase, uf: dict, var: int) -> dict:
    """The cell and face arrays of newuvw(var, ...) in its argument order:
    rho_n/o/oo, ui_n/o/oo, source, pi, g (the seven gradients), S and D
    (the three components each, in the order Process/Processor.f90:314-340
    passes them), flux_u."""
    comp, grads, order, pname = UVW_ARGS[var]
    S = (case.sx, case.sy, case.sz)
    D = (case.dx, case.dy, case.dz)
    return {"rho_n": case.rho_n, "rho_o": case.rho_o, "rho_oo": case.rho_oo,
            "ui_n": uf.get(comp, getattr(case, comp)), "ui_o": uf[comp + "_o"], "ui_oo": uf[comp + "_oo"],
            "source": uf["src_" + comp], "pi": getattr(case, pname), "g": [uf[k] for k in grads],
            "S": [S[q] for q in order], "D": [D[q] for q in order], "flux_u": uf["flux_u"],
            "flux": uf.get("flux")}
