"""Seeded synthetic inputs for the continuum-LDOS workloads.

Every workload is a tight-binding model (a list of hoppings in Wannier90
``_hr.dat`` form), a set of Wannier-function tiles sampled at tip height, an
impurity potential and the grid/energy parameters.  The generator also writes
the three input files the calcqpi driver reads: the ``_hr.dat`` model
(src/tightbinding.cc:70-149), the idl orbital file (float32 tiles of side
(2w+1)*os, src/calcqpi.cc:483-518) and the ``key=value;`` input file
(include/parser.h:29-54).

Conventions fixed by the reference and restated here:
  * tile sample (i, j) sits at lattice coordinates (i/os - w - 0.5, j/os - w - 0.5)
    (src/wannierfunctions.cc:441); hexagonal lattices sample at the Cartesian
    point u*a1 + v*a2 of those lattice coordinates;
  * tiles are rounded to float32 because the idl orbital file stores float32
    (src/idl.cc:50-74), so file round trips are exact;
  * hoppings are grouped by R and written with degeneracy 1 (the reader counts
    a new R whenever (x,y,z) changes, src/tightbinding.cc:121-127).

Builder choices (the workload text leaves them open) are stated on each
generator: the on-site term of the one-orbital model is written as -mu; the
one-orbital tile is exp(-r^2/(2 sigma^2)); angular tiles are normalised to
max |w| = 1; the 5-orbital d-set is (xz, yz, xy, x^2-y^2, 3z^2-r^2).
"""
from __future__ import annotations

import math
import dataclasses
from dataclasses import dataclass, field

import numpy as np

# must match cqpi_hopping in include/cqpi_amd.h (and or_hop in the oracle)
HOP_DTYPE = np.dtype(
    [("rx", "<f8"), ("ry", "<f8"), ("rz", "<f8"), ("o1", "<i4"), ("o2", "<i4"),
     ("tre", "<f8"), ("tim", "<f8"), ("anti", "<i4"), ("reserved", "<i4")],
    align=True,
)
assert HOP_DTYPE.itemsize == 56


@dataclass
class Workload:
    """One continuum-LDOS job: model + tiles + impurity + grid and energies."""

    name: str
    bands: int
    hops: np.ndarray            # HOP_DTYPE, 0-based orbitals, degeneracy already applied
    K: int                      # k-grid side (kpoints)
    n: int                      # lattice cells per side (lattice)
    window: int                 # neighbour window w
    oversamp: int               # sub-pixels per cell side
    energies: np.ndarray        # omega_e (eV)
    eta: float
    V: np.ndarray               # b x b complex impurity potential
    tiles: np.ndarray           # maxband x T x T float64 (float32-exact)
    spin: bool = False
    efermi: float = 0.0
    scmodel: bool = False
    lattice: str = "square"
    meta: dict = field(default_factory=dict)

    @property
    def maxband(self) -> int:
        return self.bands >> 1 if self.scmodel else self.bands

    @property
    def W(self) -> int:
        return self.oversamp * self.n

    @property
    def T(self) -> int:
        return (2 * self.window + 1) * self.oversamp

    @property
    def points(self) -> int:
        return self.W * self.W * len(self.energies)

    def V_is_diagonal(self) -> bool:
        off = self.V - np.diag(np.diag(self.V))
        return bool(np.all(off == 0))


# ----------------------------------------------------------------------------
# hoppings
# ----------------------------------------------------------------------------

def hops_from_blocks(blocks: dict, bands: int, scmodel: bool = False) -> np.ndarray:
    """blocks: {(rx, ry): b x b complex t(R)} -> HOP_DTYPE array (row-major, one
    entry per (R, o1, o2), zero entries kept so every R has b*b lines)."""
    scmax = bands >> 1 if scmodel else bands
    out = []
    for (rx, ry), t in blocks.items():
        for o1 in range(bands):
            for o2 in range(bands):
                v = complex(t[o1, o2])
                anti = int(o1 >= scmax and o2 >= scmax)
                out.append((float(rx), float(ry), 0.0, o1, o2, v.real, v.imag, anti, 0))
    return np.array(out, dtype=HOP_DTYPE)


def hk_numpy(hops: np.ndarray, bands: int, efermi: float, kx, ky, scmodel: bool = False):
    """H(k) for arrays of fractional k (vectorised numpy, used by tests and
    the gradient harness): src/tightbinding.cc:210-219, src/tightbinding.h:128-135."""
    kx = np.asarray(kx, dtype=np.float64)
    ky = np.asarray(ky, dtype=np.float64)
    shape = np.broadcast(kx, ky).shape
    H = np.zeros(shape + (bands, bands), dtype=np.complex128)
    scmax = bands >> 1 if scmodel else bands
    for i in range(bands):
        H[..., i, i] = -efermi if i < scmax else efermi
    for h in hops:
        arg = 2.0 * np.pi * (h["rx"] * kx + h["ry"] * ky)
        if h["anti"]:
            arg = -arg
        H[..., h["o1"], h["o2"]] += complex(h["tre"], h["tim"]) * np.exp(1j * arg)
    return H


def kgrid(K: int) -> np.ndarray:
    """Fractional k of index i: i/K, or (i-K)/K for i > K/2 (src/qpi.cc:465-466)."""
    i = np.arange(K, dtype=np.float64)
    return np.where(i > (K >> 1), (i - K) / K, i / K)


# ----------------------------------------------------------------------------
# tiles
# ----------------------------------------------------------------------------

def sample_grid(window: int, oversamp: int, lattice: str = "square"):
    """Cartesian (x, y) of tile sample (i, j): lattice coordinates
    (i/os - w - 0.5, j/os - w - 0.5) (src/wannierfunctions.cc:441)."""
    T = (2 * window + 1) * oversamp
    u = np.arange(T) / oversamp - window - 0.5
    U, Vv = np.meshgrid(u, u, indexing="ij")
    if lattice == "hex":
        return U + 0.5 * Vv, Vv * (math.sqrt(3.0) / 2.0)
    return U, Vv


ANGULAR = {
    "s": lambda x, y, z: np.ones_like(x),
    "xz": lambda x, y, z: x * z,
    "yz": lambda x, y, z: y * z,
    "xy": lambda x, y, z: x * y,
    "x2y2": lambda x, y, z: 0.5 * (x * x - y * y),
    "z2": lambda x, y, z: (3 * z * z - (x * x + y * y + z * z)) / (2.0 * math.sqrt(3.0)),
}


def gaussian_tiles(names, sigma, z, window, oversamp, lattice="square", normalise=True):
    x, y = sample_grid(window, oversamp, lattice)
    zz = np.full_like(x, z)
    g = np.exp(-(x * x + y * y + zz * zz) / (2.0 * sigma * sigma))
    tiles = []
    for nm in names:
        t = ANGULAR[nm](x, y, zz) * g
        if normalise:
            m = np.max(np.abs(t))
            if m > 0:
                t = t / m
        tiles.append(t)
    return np.array(tiles, dtype=np.float32).astype(np.float64)


# ----------------------------------------------------------------------------
# neighbour shells
# ----------------------------------------------------------------------------

def square_shells(nshell: int):
    """R vectors of the square lattice up to the nshell-th neighbour shell."""
    cand = {}
    for rx in range(-4, 5):
        for ry in range(-4, 5):
            cand.setdefault(rx * rx + ry * ry, []).append((rx, ry))
    d2 = sorted(cand)
    out = []
    for d in d2[: nshell + 1]:
        out += cand[d]
    return out


def hex_shells(nshell: int):
    """R = m a1 + n a2 (|a1|=|a2|=1, 60 degrees) up to the nshell-th shell,
    |R|^2 = m^2 + m n + n^2."""
    cand = {}
    for m in range(-5, 6):
        for n in range(-5, 6):
            cand.setdefault(m * m + m * n + n * n, []).append((m, n))
    d2 = sorted(cand)
    out = []
    for d in d2[: nshell + 1]:
        out += cand[d]
    return out


def hermitian_blocks(rs, bands, rng, draw, onsite):
    """t(R) drawn for one R of each {R, -R} pair, t(-R) = t(R)^H; t(0) = onsite."""
    blocks = {}
    for R in rs:
        if R == (0, 0):
            continue
        if R in blocks:
            continue
        t = draw((bands, bands))
        blocks[R] = t
        blocks[(-R[0], -R[1])] = t.conj().T
    blocks[(0, 0)] = onsite
    # deterministic order: sorted by R
    return {k: blocks[k] for k in sorted(blocks)}


def dorbital_L():
    """Orbital angular momentum L_x, L_y, L_z in the real d basis
    (xz, yz, xy, x^2-y^2, 3z^2-r^2)."""
    l = 2
    ms = np.arange(-l, l + 1)
    Lz = np.diag(ms).astype(np.complex128)
    Lp = np.zeros((5, 5), dtype=np.complex128)
    for i, m in enumerate(ms[:-1]):
        Lp[i + 1, i] = math.sqrt(l * (l + 1) - m * (m + 1))
    Lm = Lp.conj().T
    Lx = 0.5 * (Lp + Lm)
    Ly = -0.5j * (Lp - Lm)
    s2 = 1.0 / math.sqrt(2.0)
    # rows: real orbitals in terms of |m> (m=-2..2), Condon-Shortley phases
    U = np.zeros((5, 5), dtype=np.complex128)
    idx = {m: i for i, m in enumerate(ms)}
    U[0, idx[-1]], U[0, idx[1]] = s2, -s2            # xz
    U[1, idx[-1]], U[1, idx[1]] = 1j * s2, 1j * s2   # yz
    U[2, idx[-2]], U[2, idx[2]] = 1j * s2, -1j * s2  # xy
    U[3, idx[-2]], U[3, idx[2]] = s2, s2             # x^2-y^2
    U[4, idx[0]] = 1.0                               # 3z^2-r^2
    return [U.conj() @ L @ U.T for L in (Lx, Ly, Lz)]


# ----------------------------------------------------------------------------
# workloads
# ----------------------------------------------------------------------------

def square_1orb(K=256, n=64, E=11) -> Workload:
    """Single-orbital square lattice: t=-1, t'=0.3, on-site -mu with mu=0.5,
    9 R-vectors; V=1.0 on orbital 0; Gaussian tile sigma=0.35 at z=0.5,
    os=8, w=1; k-grid 256, 11 energies in [-0.5, 0.5], eta=0.01, map 64x64."""
    t, tp, mu = -1.0, 0.3, 0.5
    blocks = {}
    for R in square_shells(2):
        d2 = R[0] ** 2 + R[1] ** 2
        blocks[R] = np.array([[{0: -mu, 1: t, 2: tp}[d2]]], dtype=np.complex128)
    blocks = {k: blocks[k] for k in sorted(blocks)}
    w, os_ = 1, 8
    return Workload(
        name="square 1-orbital t=-1 t'=0.3 mu=0.5", bands=1,
        hops=hops_from_blocks(blocks, 1), K=K, n=n, window=w, oversamp=os_,
        energies=np.linspace(-0.5, 0.5, E), eta=0.01,
        V=np.array([[1.0 + 0j]]),
        tiles=gaussian_tiles(["s"], 0.35, 0.5, w, os_, normalise=False),
        meta={"mu": mu, "t": t, "tp": tp},
    )


def square_3orb(K=512, n=128, E=51, seed=1002, elo=-0.1, ehi=0.1) -> Workload:
    """3-orbital square lattice, Hermitian real hoppings up to the 3rd
    neighbour (13 R) drawn N(0,0.3), on-site diagonal U(-0.2,0.2) (on-site
    off-diagonal N(0,0.3), Hermitised); V = 0.5 on every orbital; tiles
    Gaussian(sigma=0.4) x {xz, yz, xy} at z=0.6, os=4, w=2; k-grid 512,
    51 energies in [-0.1, 0.1], eta=0.002, map 128x128."""
    rng = np.random.default_rng(seed)
    b = 3
    onsite = rng.normal(0, 0.3, (b, b))
    onsite = 0.5 * (onsite + onsite.T)
    np.fill_diagonal(onsite, rng.uniform(-0.2, 0.2, b))
    blocks = hermitian_blocks(square_shells(3), b, rng,
                              lambda s: rng.normal(0, 0.3, s).astype(np.complex128),
                              onsite.astype(np.complex128))
    w, os_ = 2, 4
    return Workload(
        name="square 3-orbital seed %d" % seed, bands=b,
        hops=hops_from_blocks(blocks, b), K=K, n=n, window=w, oversamp=os_,
        energies=np.linspace(elo, ehi, E), eta=0.002,
        V=np.diag([0.5] * b).astype(np.complex128),
        tiles=gaussian_tiles(["xz", "yz", "xy"], 0.4, 0.6, w, os_),
        meta={"seed": seed},
    )


def hex_5d_spin(K=1024, n=256, E=101, seed=1003, lam=0.05) -> Workload:
    """Hexagonal lattice, 5 d orbitals x spin = 10 bands, complex hoppings
    N(0,0.2)+iN(0,0.02) up to the 4th shell (31 R), Hermitised, spin-diagonal,
    plus on-site spin-orbit lambda L.S with lambda=0.05; full 10x10 Hermitian
    V with entries N(0,0.3); tiles Gaussian(sigma=0.4) x real d harmonics at
    z=0.5, os=6, w=3; k-grid 1024, 101 energies in [-0.2,0.2], eta=0.001,
    map 256x256; spin=true (same-spin contraction blocks)."""
    rng = np.random.default_rng(seed)
    d = 5

    def draw(s):
        return rng.normal(0, 0.2, s) + 1j * rng.normal(0, 0.02, s)

    a = draw((d, d))
    h0 = 0.5 * (a + a.conj().T)
    hb = hermitian_blocks(hex_shells(4), d, rng, draw, h0)
    L = dorbital_L()
    sig = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]]),
           np.array([[1, 0], [0, -1]], complex)]
    soc = lam * sum(np.kron(0.5 * s, Lm) for s, Lm in zip(sig, L))
    blocks = {}
    for R, t in hb.items():
        full = np.kron(np.eye(2), t)
        if R == (0, 0):
            full = full + soc
        blocks[R] = full
    A = rng.normal(0, 0.3, (2 * d, 2 * d)) + 1j * rng.normal(0, 0.3, (2 * d, 2 * d))
    V = 0.5 * (A + A.conj().T)
    w, os_ = 3, 6
    tiles5 = gaussian_tiles(["xz", "yz", "xy", "x2y2", "z2"], 0.4, 0.5, w, os_, lattice="hex")
    return Workload(
        name="hexagonal 5d x spin seed %d" % seed, bands=2 * d,
        hops=hops_from_blocks(blocks, 2 * d), K=K, n=n, window=w, oversamp=os_,
        energies=np.linspace(-0.2, 0.2, E), eta=0.001, V=V,
        tiles=np.concatenate([tiles5, tiles5]), spin=True, lattice="hex",
        meta={"seed": seed, "lambda": lam},
    )


ORB_SETS = {1: ["s"], 3: ["xz", "yz", "xy"], 5: ["xz", "yz", "xy", "x2y2", "z2"],
            6: ["s", "xz", "yz", "xy", "x2y2", "z2"], 7: ["s", "xz", "yz", "xy", "x2y2", "z2", "s"]}


def layered(b, K, n, E, seed, tz=0.12, **kw):
    """square_norb's model as one layer (its terms at rz = 0) of a stack along
    z: coupling blocks to the next layer at R = (0, 0, -1) and (1, 0, -1)
    (random, scale tz) and their conjugates at rz = +1.  Returns the Workload
    with the stacked hoppings."""
    wl = square_norb(b, K=K, n=n, E=E, seed=seed, **kw)
    rng = np.random.default_rng(seed + 1000)
    extra = []
    for rx in (0.0, 1.0):
        M = tz * (rng.standard_normal((b, b)) + 1j * rng.standard_normal((b, b))) / (2.0 if rx else 1.0)
        # <o1, R| H |o2, 0> = M[o1, o2] at R = (rx, 0, -1); the hermitian
        # partner at (-rx, 0, +1) is M^H (one block per R, as an _hr file groups them)
        for o1 in range(b):
            for o2 in range(b):
                extra.append((rx, 0.0, -1.0, o1, o2, M[o1, o2].real, M[o1, o2].imag, 0, 0))
        for o1 in range(b):
            for o2 in range(b):
                extra.append((-rx, 0.0, 1.0, o1, o2, M[o2, o1].real, -M[o2, o1].imag, 0, 0))
    hops = np.concatenate([np.asarray(wl.hops, HOP_DTYPE), np.array(extra, HOP_DTYPE)])
    return dataclasses.replace(wl, name="layered %d-orbital" % b, hops=hops)


def square_norb(b, K, n, E, seed, elo=-0.3, ehi=0.3, w=2, os_=4, eta=0.002, V=None):
    """Model and tiles generated as in config 2 for b in {1,3,5} (6 and 7 for
    the lane-group eigensolve tests)."""
    rng = np.random.default_rng(seed)
    onsite = rng.normal(0, 0.3, (b, b))
    onsite = 0.5 * (onsite + onsite.T)
    np.fill_diagonal(onsite, rng.uniform(-0.2, 0.2, b))
    blocks = hermitian_blocks(square_shells(3), b, rng,
                              lambda s: rng.normal(0, 0.3, s).astype(np.complex128),
                              onsite.astype(np.complex128))
    if V is None:
        A = rng.normal(0, 1, (b, b)) + 1j * rng.normal(0, 1, (b, b))
        Hm = 0.5 * (A + A.conj().T)
        Hm /= np.linalg.norm(Hm)
        V = rng.uniform(0.1, 2.0) * Hm
    return Workload(
        name="square %d-orbital seed %d" % (b, seed), bands=b,
        hops=hops_from_blocks(blocks, b), K=K, n=n, window=w, oversamp=os_,
        energies=np.linspace(elo, ehi, E), eta=eta, V=np.asarray(V, np.complex128),
        tiles=gaussian_tiles(orb_set(b), 0.4, 0.6, w, os_),
        meta={"seed": seed},
    )


def orb_set(b):
    """Angular factors of b tiles: the fixed sets for 1, 3, 5, 6, 7 orbitals,
    otherwise s, xz, yz, xy, x2y2, z2 repeated."""
    if b in ORB_SETS:
        return ORB_SETS[b]
    base = ["s", "xz", "yz", "xy", "x2y2", "z2"]
    return [base[i % len(base)] for i in range(b)]


def bdg_square(nb, K, n, E, seed, delta=0.08, w=1, os_=2, eta=0.01, efermi=0.05, elo=-0.06, ehi=0.06):
    """A superconducting (scmodel) square-lattice model for the Josephson map:
    the normal-state blocks t(R) of square_norb, the hole block -t(R)^T with
    the 'anti' phase (so H_hh(k) = -H(-k)^*), constant on-site s-wave pairing
    delta; V = diag(V_n, -V_n^*); tiles for the nb electron orbitals."""
    base = square_norb(nb, K, n, E, seed, elo=elo, ehi=ehi, w=w, os_=os_, eta=eta)
    rng = np.random.default_rng(seed + 500)
    onsite = rng.normal(0, 0.3, (nb, nb))
    onsite = 0.5 * (onsite + onsite.T)
    np.fill_diagonal(onsite, rng.uniform(-0.2, 0.2, nb))
    normal = hermitian_blocks(square_shells(2), nb, rng,
                              lambda sh: rng.normal(0, 0.3, sh).astype(np.complex128),
                              onsite.astype(np.complex128))
    b = 2 * nb
    blocks = {}
    for R, t in normal.items():
        m = np.zeros((b, b), np.complex128)
        m[:nb, :nb] = t
        m[nb:, nb:] = -t.T
        if R == (0, 0):
            m[:nb, nb:] = delta * np.eye(nb)
            m[nb:, :nb] = np.conj(delta) * np.eye(nb)
        blocks[R] = m
    V = np.zeros((b, b), np.complex128)
    V[:nb, :nb] = base.V
    V[nb:, nb:] = -np.conj(base.V)
    return Workload(
        name="BdG square %d-orbital seed %d" % (nb, seed), bands=b,
        hops=hops_from_blocks(blocks, b, scmodel=True), K=K, n=n, window=w, oversamp=os_,
        energies=base.energies, eta=eta, V=V, tiles=base.tiles, efermi=efermi, scmodel=True,
        meta={"seed": seed, "delta": delta},
    )


def batch_uneven(count=128, seed=1004):
    """Batch of independent instances with uneven sizes: orbitals in {1,3,5},
    k-grid in {128,256,512}, 16-64 energies in [-0.3,0.3], V = U(0.1,2.0) x a
    random Hermitian unit-norm matrix, map in {32,64,128} cells."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        b = int(rng.choice([1, 3, 5]))
        K = int(rng.choice([128, 256, 512]))
        n = int(rng.choice([32, 64, 128]))
        E = int(rng.integers(16, 65))
        out.append(square_norb(b, K, n, E, seed=seed * 1000 + i))
    return out


def square_3orb_grad(K=512, n=64, E=21, seed=1005):
    """Config-2 shapes with 21 energies and a 64x64-cell map (gradient workload)."""
    return square_3orb(K=K, n=n, E=E, seed=seed)


def perturbed(wl: Workload, seed=2005, dt=0.02, dmu=0.01, dV=0.1) -> Workload:
    """Target-map parameters: hoppings + N(0,dt) (kept Hermitian), mu + dmu,
    V + dV on the diagonal."""
    rng = np.random.default_rng(seed)
    hops = wl.hops.copy()
    # perturb one of each (R,o1,o2)/(−R,o2,o1) pair and mirror it
    key = {(h["rx"], h["ry"], h["o1"], h["o2"]): i for i, h in enumerate(hops)}
    done = set()
    for i, h in enumerate(hops):
        if i in done:
            continue
        j = key[(-h["rx"] + 0.0, -h["ry"] + 0.0, h["o2"], h["o1"])]
        d = rng.normal(0, dt)
        hops[i]["tre"] += d
        if j != i:
            hops[j]["tre"] += d
        done.update((i, j))
    out = Workload(**{**wl.__dict__})
    out.hops = hops
    out.efermi = wl.efermi + dmu
    out.V = wl.V + dV * np.eye(wl.bands)
    return out


WORKLOADS = {
    "square_1orb": square_1orb,
    "square_3orb": square_3orb,
    "hex_5d_spin": hex_5d_spin,
    "square_3orb_grad": square_3orb_grad,
}


def workload(name: str, **kw) -> Workload:
    return WORKLOADS[name](**kw)


# ----------------------------------------------------------------------------
# file writers (calcqpi input formats)
# ----------------------------------------------------------------------------

def write_hr(path, wl: Workload):
    """Wannier90 _hr.dat: name, num_wann, nrpts, degeneracies (15 per line),
    then 'x y z u v Re Im' grouped by R (1-based orbitals)."""
    hops = wl.hops
    Rs = []
    for h in hops:
        R = (int(h["rx"]), int(h["ry"]), int(h["rz"]))
        if not Rs or Rs[-1] != R:
            Rs.append(R)
    with open(path, "w") as f:
        f.write(wl.name + "\n")
        f.write("%12d\n" % wl.bands)
        f.write("%12d\n" % len(Rs))
        for i in range(0, len(Rs), 15):
            f.write("".join("%5d" % 1 for _ in Rs[i:i + 15]) + "\n")
        for h in hops:
            f.write("%5d%5d%5d%5d%5d %23.16e %23.16e\n" % (
                int(h["rx"]), int(h["ry"]), int(h["rz"]), int(h["o1"]) + 1, int(h["o2"]) + 1,
                h["tre"], h["tim"]))


def write_idl(path, data_xyz, name="", date="", hdr=(1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0)):
    """idl file (src/idl.cc:78-95): text header then float32 data, x fastest,
    then y, then layer.  data_xyz has shape (sx, sy, sz)."""
    sx, sy, sz = data_xyz.shape
    with open(path, "wb") as f:
        head = "%s\n%s\n%d\n%d\n%d\n" % (name, date, sx, sy, sz)
        head += "".join("%g\n" % v for v in hdr)
        f.write(head.encode())
        f.write(np.ascontiguousarray(np.transpose(data_xyz, (2, 1, 0)), dtype="<f4").tobytes())


def read_idl(path):
    """Returns (data[sx, sy, sz] float64, header dict)."""
    with open(path, "rb") as f:
        raw = f.read()
    lines = []
    pos = 0
    for _ in range(12):
        e = raw.index(b"\n", pos)
        lines.append(raw[pos:e].decode(errors="replace"))
        pos = e + 1
    sx, sy, sz = int(lines[2]), int(lines[3]), int(lines[4])
    vals = [float(v) for v in lines[5:12]]
    data = np.frombuffer(raw, dtype="<f4", count=sx * sy * sz, offset=pos)
    data = data.reshape(sz, sy, sx).transpose(2, 1, 0).astype(np.float64)
    return data, {"name": lines[0], "date": lines[1], "scan_range": vals[0:2],
                  "offset": vals[2:4], "bias": vals[4:6], "setpoint": vals[6]}


def write_inputs(dirpath, wl: Workload, qpifile="qpi.idl", extra=None, output="wannier"):
    """Writes model.dat, orbitals.idl and calcqpi.in into dirpath; returns the
    input-file path.  The impurity must be diagonal (the input format only
    carries per-orbital scattering, src/calcqpi.cc:254-272)."""
    import os
    os.makedirs(dirpath, exist_ok=True)
    write_hr(os.path.join(dirpath, "model.dat"), wl)
    tiles = wl.tiles
    if wl.spin and tiles.shape[0] == wl.maxband and np.array_equal(
            tiles[: wl.maxband // 2], tiles[wl.maxband // 2:]):
        tiles = tiles[: wl.maxband // 2]
    write_idl(os.path.join(dirpath, "orbitals.idl"), np.transpose(tiles, (1, 2, 0)),
              name="tiles")
    if not wl.V_is_diagonal():
        raise ValueError("calcqpi input files carry a diagonal impurity only")
    scat = ",".join("%.17g" % v.real for v in np.diag(wl.V))
    E = len(wl.energies)
    lines = [
        'tbfile="model.dat";', 'qpifile="%s";' % qpifile, "output=%s;" % output,
        "lattice=%d;" % wl.n, "kpoints=%d;" % wl.K, "oversamp=%d;" % wl.oversamp,
        "window=%d;" % wl.window, "layers=%d;" % E,
        "energies=(%.17g,%.17g);" % (wl.energies[0], wl.energies[-1]),
        "eta=%.17g;" % wl.eta, "fermi=%.17g;" % wl.efermi,
        "scattering=(%s);" % scat, 'idlorbitalfile="orbitals.idl";',
        "spin=%s;" % ("true" if wl.spin else "false"),
    ]
    if wl.scmodel:
        lines.append("scmodel=true;")
    if extra:
        lines += ["%s=%s;" % (k, v) for k, v in extra.items()]
    p = os.path.join(dirpath, "calcqpi.in")
    with open(p, "w") as f:
        f.write("# continuum-LDOS job: %s\n" % wl.name)
        f.write("\n".join(lines) + "\n")
    return p
