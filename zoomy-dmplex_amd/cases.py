"""Synthetic benchmark / parity cases (meshes, bathymetry, initial conditions).

Every random draw comes from a SplitMix64 stream with a per-case seed, so the host arrays
fed to the CPU oracle and to the GPU are identical and reproducible:

  C1  seed 1     SME level 0, 60 x 60 right-triangle grid (7 200 cells), Gaussian bump,
                 dam break h = 1 (x < 0.4) / 0.5 plus a 1e-3 Fourier perturbation, walls,
                 MUSCL + SSP-RK2, CFL 0.45, 200 steps
  C2  seed 2     SME level 2 + implicit Newtonian-slip source (nu = 1e-3, slip 0.1),
                 1000 x 1000 grid (2 000 000 cells), 16 random Gaussian bumps,
                 h = 1 - b + 0.1 * circular dam (r = 0.2), MUSCL + SSP-RK2, 500 steps
  C4  seed 4000+i ensemble member i (SME level 1, jittered grid, 8 bumps, random dam)
  C5  seed 5     calibration: SME level 1 + Manning friction, smooth random Manning field
                 n in [0.01, 0.05] and bed, dam break, h observed at 32 gauge cells

The initial condition is evaluated at the cell centroids, as
VirtualSolver::ProjectModelInitialCondition does (VirtualSolver.hpp:259-281).
"""
import math

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        u = (self.next() >> 11) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def randint(self, lo, hi):
        """integer in [lo, hi]"""
        return lo + min(int(self.uniform() * (hi - lo + 1)), hi - lo)


def gaussian_bumps(xy, rng, n, amp_hi, w_lo=0.02, w_hi=0.1):
    b = np.zeros(xy.shape[0])
    for _ in range(n):
        a = rng.uniform(0.0, amp_hi)
        w = rng.uniform(w_lo, w_hi)
        cx, cy = rng.uniform(), rng.uniform()
        r2 = (xy[:, 0] - cx) ** 2 + (xy[:, 1] - cy) ** 2
        b += a * np.exp(-r2 / (w * w))
    return b


def sme_state(n_q, b, h):
    X = np.zeros((b.shape[0], n_q))
    X[:, 0] = b
    X[:, 1] = h
    return X


def c1_state(mesh, n_q=4, seed=1):
    """C1: dam break over a Gaussian bump with a seeded Fourier perturbation."""
    rng = SplitMix64(seed)
    xy = mesh.cell_centroid[:, :2]
    x, y = xy[:, 0], xy[:, 1]
    b = 0.2 * np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.01)
    pert = np.zeros_like(x)
    for _ in range(8):
        a = rng.uniform(-1.0, 1.0)
        m = rng.randint(1, 4)
        k = rng.randint(1, 4)
        ph = rng.uniform(0.0, 2.0 * math.pi)
        pert += a * np.sin(2.0 * math.pi * (m * x + k * y) + ph)
    h = np.where(x < 0.4, 1.0, 0.5) + 1e-3 * pert / 8.0
    return sme_state(n_q, b, h)


def c2_state(mesh, n_q=8, seed=2):
    """C2: 16 random bumps, still water at depth 1 - b plus a circular dam of height 0.1."""
    rng = SplitMix64(seed)
    xy = mesh.cell_centroid[:, :2]
    b = gaussian_bumps(xy, rng, 16, 0.15)
    r = np.hypot(xy[:, 0] - 0.5, xy[:, 1] - 0.5)
    h = 1.0 - b + 0.1 * (r < 0.2)
    return sme_state(n_q, b, h)


def c3_state(mesh, seed=3, amp=0.2, x0=0.3, n_bumps=6, bump_hi=0.1, g=9.81):
    """C3 (VAM): solitary wave of amplitude `amp` on depth 1 over a seeded bump field
    (heights U(0, bump_hi)); state [b, h, q_x_0, q_x_1, q_y_0, q_y_1, r_0, r_1]."""
    rng = SplitMix64(seed)
    xy = mesh.cell_centroid[:, :2]
    b = gaussian_bumps(xy, rng, n_bumps, bump_hi)
    kappa = math.sqrt(3.0 * amp / (4.0 * (1.0 + amp)))
    eta = amp / np.cosh(kappa * (xy[:, 0] - x0) / 0.05) ** 2
    h = np.maximum(1.0 - b + eta, 1e-3)
    c = math.sqrt(g * (1.0 + amp))
    X = np.zeros((xy.shape[0], 8))
    X[:, 0] = b
    X[:, 1] = h
    X[:, 2] = c * eta
    return X


def ensemble_member(i, n_q=6, base_seed=4000, cell_lo=1.0e4, cell_hi=2.0e5):
    """C4 member i: (grid n, jitter seed, state builder)."""
    rng = SplitMix64(base_seed + i)
    ncell = math.exp(rng.uniform(math.log(cell_lo), math.log(cell_hi)))
    n = max(2, int(round(math.sqrt(ncell / 2.0))))
    mesh_seed = rng.next()
    bump_seed = rng.next()
    x0 = rng.uniform(0.3, 0.7)
    ratio = rng.uniform(0.3, 0.8)

    def state(mesh):
        r2 = SplitMix64(bump_seed)
        xy = mesh.cell_centroid[:, :2]
        b = gaussian_bumps(xy, r2, 8, 0.2)
        h = np.where(xy[:, 0] < x0, 1.0, ratio) - b
        h = np.maximum(h, 0.0)
        return sme_state(n_q, b, h)

    return n, mesh_seed, state


def lake_at_rest(mesh, n_q, seed=7, eta=1.0, amp=0.3):
    rng = SplitMix64(seed)
    xy = mesh.cell_centroid[:, :2]
    b = gaussian_bumps(xy, rng, 6, amp)
    return sme_state(n_q, b, eta - b)


def smooth_field(xy, rng, modes=6):
    """Smooth random field in [-1, 1]: normalised sum of random Fourier modes."""
    f = np.zeros(xy.shape[0])
    for _ in range(modes):
        a = rng.uniform(-1.0, 1.0)
        kx, ky = rng.uniform(0.5, 3.0), rng.uniform(0.5, 3.0)
        ph = rng.uniform(0.0, 2.0 * math.pi)
        f += a * np.sin(2.0 * math.pi * (kx * xy[:, 0] + ky * xy[:, 1]) + ph)
    m = np.max(np.abs(f))
    return f / m if m > 0 else f


def c5_fields(mesh, seed):
    """(bed, Manning n) smooth random fields: b in [0, 0.1] + a gentle slope, n in [0.01, 0.05]."""
    rng = SplitMix64(seed)
    xy = mesh.cell_centroid[:, :2]
    b = 0.05 * (1.0 + smooth_field(xy, rng)) + 0.05 * xy[:, 0]
    n = 0.03 + 0.02 * smooth_field(xy, rng)
    return b, n


def c5_case(mesh, seed=5, ngauge=32, x0=0.5, eta_hi=1.0, eta_lo=0.6):
    """C5: truth run state/aux, first-guess state/aux and the gauge cells.
    Dam break eta = eta_hi (x < x0) / eta_lo over the bed; the first guess has the bed and
    Manning fields of another seed."""
    rng = SplitMix64(seed + 1000)
    xy = mesh.cell_centroid[:, :2]

    def build(fseed):
        b, n = c5_fields(mesh, fseed)
        X = np.zeros((mesh.ncells, 6))
        X[:, 0] = b
        X[:, 1] = np.where(xy[:, 0] < x0, eta_hi, eta_lo) - b
        A = np.zeros((mesh.ncells, 2))
        A[:, 1] = n
        return X, A

    Xt, At = build(seed)
    Xg, Ag = build(seed + 1)
    gauges = []
    while len(gauges) < ngauge:
        g = rng.randint(0, mesh.ncells - 1)
        if g not in gauges:
            gauges.append(g)
    return Xt, At, Xg, Ag, np.array(gauges, dtype=np.int32)

