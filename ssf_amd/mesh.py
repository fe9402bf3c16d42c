"""Seeded synthetic finite-element problems for the SIFEL linear-system path.

Node and element numbering follow the reference's structured generators:
hexahedra PREP/SEQMESHGEN/gensifhex.cpp:18-96 (nodes z fastest, then y, then x;
coordinates accumulated step by step; element nodes (+,+,+), (-,+,+), (-,-,+),
(+,-,+) then the same at the lower z, matching GEFEL/basefun.cpp:2299-2309) and
quadrilaterals PREP/SEQMESHGEN/gensifquad.cpp:25-78.

Equation numbers follow gtopology::gencodnum (GEFEL/gtopology.cpp:1619-1680):
per-node DOF flags 0 = fixed, 1 = free, k > 1 = coupled group, <0 = prescribed
value index; free DOFs are numbered 1..n in node order x DOF order, a coupled
group sharing the number of its first DOF. The element code-number table matches
gtopology::give_code_numbers (GEFEL/gtopology.cpp:597-626).

Everything here is host-side input construction (numpy); the arithmetic of the
path itself runs in the HIP library.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

KIND_HEX8_ELAST = 0
KIND_HEX8_HEAT = 1
KIND_QUAD4_PSTRESS = 2

KIND_NAMES = {KIND_HEX8_ELAST: "hex8_elast", KIND_HEX8_HEAT: "hex8_heat",
              KIND_QUAD4_PSTRESS: "quad4_pstress"}
NNE = {KIND_HEX8_ELAST: 8, KIND_HEX8_HEAT: 8, KIND_QUAD4_PSTRESS: 4}
NDOFN = {KIND_HEX8_ELAST: 3, KIND_HEX8_HEAT: 1, KIND_QUAD4_PSTRESS: 2}


@dataclass
class Problem:
    """One linear FE problem: mesh, material, code numbers and right-hand side(s)."""
    kind: int
    coords: np.ndarray          # (nn, 3) float64 (z = 0 for quad4)
    conn: np.ndarray            # (ne, nne) int64, 0-based node ids
    dof: np.ndarray             # (nn, ndofn) int64 code numbers (>0 eq., 0 fixed, <0 presc.)
    n: int                      # number of unknowns
    e: np.ndarray               # (ne,) Young's modulus / conductivity
    nu: np.ndarray              # (ne,) Poisson ratio (zeros for heat)
    thick: np.ndarray           # (ne,) thickness (quad4; ones otherwise)
    pv: np.ndarray              # prescribed values, indexed by -cn-1
    rhs: np.ndarray             # (nrhs, n) loads without the Dirichlet lift
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def ne(self) -> int:
        return int(self.conn.shape[0])

    @property
    def nne(self) -> int:
        return NNE[self.kind]

    @property
    def ndofn(self) -> int:
        return NDOFN[self.kind]

    @property
    def ndofe(self) -> int:
        return self.nne * self.ndofn

    def elem_codes(self) -> np.ndarray:
        """(ne, ndofe) int64 element code numbers (give_code_numbers)."""
        return self.dof[self.conn].reshape(self.ne, self.ndofe).copy()

    def xyz_flat(self) -> np.ndarray:
        if self.kind == KIND_QUAD4_PSTRESS:
            return np.ascontiguousarray(self.coords[:, :2]).reshape(-1)
        return np.ascontiguousarray(self.coords).reshape(-1)


def _axis(n: int, length: float) -> np.ndarray:
    """Coordinates accumulated as gensifhex.cpp:28-50 does (xx += lx)."""
    h = length / n
    out = np.empty(n + 1)
    v = 0.0
    for i in range(n + 1):
        out[i] = v
        v += h
    return out


def hex_grid(nex: int, ney: int, nez: int, lx: float, ly: float, lz: float):
    xs, ys, zs = _axis(nex, lx), _axis(ney, ly), _axis(nez, lz)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    ii, jj, kk = np.meshgrid(np.arange(nex), np.arange(ney), np.arange(nez), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    s1, s2 = (ney + 1) * (nez + 1), nez + 1
    lb = ii * s1 + jj * s2 + kk
    rb = (ii + 1) * s1 + jj * s2 + kk
    lt = ii * s1 + (jj + 1) * s2 + kk
    rt = (ii + 1) * s1 + (jj + 1) * s2 + kk
    conn = np.stack([rt + 1, lt + 1, lb + 1, rb + 1, rt, lt, lb, rb], axis=1).astype(np.int64)
    return coords, conn


def quad_grid(nex: int, ney: int, lx: float, ly: float):
    xs, ys = _axis(nex, lx), _axis(ney, ly)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    ii, jj = np.meshgrid(np.arange(nex), np.arange(ney), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    nl = ii * (ney + 1) + jj
    nr = (ii + 1) * (ney + 1) + jj
    conn = np.stack([nl, nr, nr + 1, nl + 1], axis=1).astype(np.int64)
    return coords, conn


def codenum(flags: np.ndarray, ordering=None) -> tuple[np.ndarray, int]:
    """gencodnum (GEFEL/gtopology.cpp:1619-1665) for node flags (nn, ndofn): 0 fixed,
    1 free, k > 1 coupled group k, < 0 prescribed. Nodes are visited in `ordering`
    (default: node order), DOFs of a node in order; a free DOF takes the next
    number, a coupled group one number at its first DOF; 0 and < 0 stay."""
    flags = np.asarray(flags, dtype=np.int64)
    shape = flags.shape
    f2 = flags.reshape(shape[0], -1)
    order = np.arange(shape[0]) if ordering is None else np.asarray(ordering, dtype=np.int64)
    visit = f2[order].reshape(-1)
    pos = np.nonzero(visit > 0)[0]
    # a token per unknown: each free DOF its own (negative), a group its flag
    key = np.where(visit[pos] == 1, -(pos + 1), visit[pos])
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    rank = np.empty(uniq.size, dtype=np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(1, uniq.size + 1)
    out = visit.copy()
    out[pos] = rank[inv]
    dof = np.empty_like(f2)
    dof[order] = out.reshape(f2.shape)
    return dof.reshape(shape), int(uniq.size)


def lognormal(rng: np.random.Generator, size: int, mean: float, sigma_log: float) -> np.ndarray:
    """Lognormal field with the given arithmetic mean: exp(mu + s Z), mu = ln(mean) - s^2/2."""
    mu = np.log(mean) - 0.5 * sigma_log * sigma_log
    return np.exp(mu + sigma_log * rng.standard_normal(size))


# ---------------------------------------------------------------------------
# the five workloads (geometry choices recorded in the design notes)
# ---------------------------------------------------------------------------

def cantilever(nex=80, ney=20, nez=20, length=4.0, width=1.0, height=1.0,
               e=210e9, nu=0.3, traction=1e6) -> Problem:
    """Linear elastic 3D cantilever: x=0 face clamped, uniform traction -z on x=L."""
    coords, conn = hex_grid(nex, ney, nez, length, width, height)
    nn = coords.shape[0]
    flags = np.ones((nn, 3), dtype=np.int64)
    flags[np.isclose(coords[:, 0], 0.0)] = 0
    dof, n = codenum(flags)
    ne = conn.shape[0]
    # consistent nodal forces of a uniform traction on the x=L face:
    # t * hy * hz * (1/4 per adjacent face element) -> {1, 1/2, 1/4}
    hy, hz = width / ney, height / nez
    f = np.zeros(n)
    face = np.nonzero(np.isclose(coords[:, 0], coords[:, 0].max()))[0]
    jy = np.rint(coords[face, 1] / hy).astype(int)
    kz = np.rint(coords[face, 2] / hz).astype(int)
    wy = np.where((jy == 0) | (jy == ney), 0.5, 1.0)
    wz = np.where((kz == 0) | (kz == nez), 0.5, 1.0)
    fz = -traction * hy * hz * wy * wz
    eq = dof[face, 2]
    np.add.at(f, eq[eq > 0] - 1, fz[eq > 0])
    return Problem(KIND_HEX8_ELAST, coords, conn, dof, n, np.full(ne, e), np.full(ne, nu),
                   np.ones(ne), np.zeros(0), f[None, :], name=f"cantilever_{nex}x{ney}x{nez}")


def heat_cube(nel=128, seed=2, k_mean=1.0, sigma_log=0.5, t_top=100.0) -> Problem:
    """3D steady heat conduction on the unit cube: T=0 on z=0, T=t_top on z=1."""
    coords, conn = hex_grid(nel, nel, nel, 1.0, 1.0, 1.0)
    nn = coords.shape[0]
    rng = np.random.default_rng(seed)
    kfield = lognormal(rng, conn.shape[0], k_mean, sigma_log)
    kz = np.rint(coords[:, 2] * nel).astype(int)
    flags = np.ones((nn, 1), dtype=np.int64)
    flags[kz == 0, 0] = -1          # prescribed value pv[0] = 0
    flags[kz == nel, 0] = -2        # prescribed value pv[1] = t_top
    dof, n = codenum(flags)
    ne = conn.shape[0]
    return Problem(KIND_HEX8_HEAT, coords, conn, dof, n, kfield, np.zeros(ne), np.ones(ne),
                   np.array([0.0, t_top]), np.zeros((1, n)), name=f"heat_{nel}^3",
                   meta={"lift": True})


def elastic_column(nex=64, ney=64, nez=256, seed=3, e_mean=30e9, sigma_log=0.3,
                   nu_lo=0.2, nu_hi=0.3, lx=1.0, ly=1.0, lz=4.0) -> Problem:
    """3D elasticity column: base z=0 clamped, random nodal body force N(0,1) kN/m^3."""
    coords, conn = hex_grid(nex, ney, nez, lx, ly, lz)
    nn = coords.shape[0]
    ne = conn.shape[0]
    rng = np.random.default_rng(seed)
    e = lognormal(rng, ne, e_mean, sigma_log)
    nu = rng.uniform(nu_lo, nu_hi, ne)
    flags = np.ones((nn, 3), dtype=np.int64)
    flags[np.isclose(coords[:, 2], 0.0)] = 0
    dof, n = codenum(flags)
    hx, hy, hz = lx / nex, ly / ney, lz / nez
    ijk = np.rint(coords / np.array([hx, hy, hz])).astype(int)
    trib = np.ones(nn) * hx * hy * hz
    for ax, nax in enumerate((nex, ney, nez)):
        trib *= np.where((ijk[:, ax] == 0) | (ijk[:, ax] == nax), 0.5, 1.0)
    b = rng.standard_normal((nn, 3)) * 1e3 * trib[:, None]
    f = np.zeros(n)
    m = dof > 0
    f[dof[m] - 1] = b[m]
    return Problem(KIND_HEX8_ELAST, coords, conn, dof, n, e, nu, np.ones(ne), np.zeros(0),
                   f[None, :], name=f"elastic_{nex}x{ney}x{nez}")


def plate(nex: int, ney: int, seed: int, nrhs=8, e_mean=30e9, sigma_log=0.3, nu=0.25,
          lx=None, ly=1.0) -> Problem:
    """2D plane-stress quad4 plate, left edge (x=0) clamped, nrhs random nodal loads."""
    lx = float(nex) / ney * ly if lx is None else lx
    coords, conn = quad_grid(nex, ney, lx, ly)
    nn = coords.shape[0]
    ne = conn.shape[0]
    rng = np.random.default_rng(seed)
    e = lognormal(rng, ne, e_mean, sigma_log)
    flags = np.ones((nn, 2), dtype=np.int64)
    flags[np.isclose(coords[:, 0], 0.0)] = 0
    dof, n = codenum(flags)
    loads = np.zeros((nrhs, n))
    m = dof > 0
    for r in range(nrhs):
        g = rng.standard_normal((nn, 2))
        loads[r, dof[m] - 1] = g[m]
    return Problem(KIND_QUAD4_PSTRESS, coords, conn, dof, n, e, np.full(ne, nu), np.ones(ne),
                   np.zeros(0), loads, name=f"plate_{nex}x{ney}")


def plate_batch(count=512, seed=4, dof_lo=1e3, dof_hi=6e4, nrhs=8):
    """Batch of independent plates: DOF log-uniform in [dof_lo, dof_hi], aspect 1..4."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        target = float(np.exp(rng.uniform(np.log(dof_lo), np.log(dof_hi))))
        aspect = float(rng.uniform(1.0, 4.0))
        nodes = target / 2.0
        ney = max(2, int(round(np.sqrt(nodes / aspect))) - 1)
        nex = max(2, int(round(nodes / (ney + 1))) - 1)
        out.append(plate(nex, ney, seed=int(rng.integers(1 << 31)), nrhs=nrhs))
    return out


def served_request(rng: np.random.Generator, dof_lo=5e3, dof_hi=2e6) -> Problem:
    """One serving request: 50/50 hex8 elasticity / heat, DOF log-uniform."""
    target = float(np.exp(rng.uniform(np.log(dof_lo), np.log(dof_hi))))
    seed = int(rng.integers(1 << 31))
    if rng.uniform() < 0.5:
        nel = max(2, int(round((target / 3.0) ** (1.0 / 3.0) / 4 ** (1.0 / 3.0))))
        return elastic_column(nel, nel, 4 * nel, seed=seed)
    nel = max(2, int(round(target ** (1.0 / 3.0))) - 1)
    return heat_cube(nel, seed=seed)
