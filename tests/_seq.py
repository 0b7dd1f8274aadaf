"""Deterministic pair-state and particle-state sequences that sweep a contact
through formation, sliding, rolling, twisting and separation. Used to drive the
reference functor headers (golden vectors) and the oracle through identical
inputs."""
import numpy as np

from _oracle import R_PART


def pair_sequence(seed, nsteps=200, period=40, amp=0.03, v_sigma=2.0, w_sigma=1.0e8, offset=1e-6):
    """Rows of 18 doubles: xi, xj, vi, vj, wi, wj."""
    rng = np.random.default_rng(seed)
    base = rng.normal(0.0, offset, 3)
    th0, ph0 = rng.uniform(0, np.pi), rng.uniform(0, 2 * np.pi)
    out = np.zeros((nsteps, 18))
    for k in range(nsteps):
        th, ph = th0 + 0.03 * k, ph0 + 0.05 * k
        u = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        d = 2.0 * R_PART * (1.0 + amp * np.sin(2 * np.pi * k / period))
        xi = base
        xj = base + d * u
        out[k, 0:3] = xi
        out[k, 3:6] = xj
        out[k, 6:12] = rng.normal(0.0, v_sigma, 6)
        out[k, 12:18] = rng.normal(0.0, w_sigma, 6)
    return out


def bonded_sequence(seed, nsteps=200, amp=0.05, **kw):
    """Starts exactly touching (so alt_sinter bonds the pair), then stretches
    and compresses through +-amp of the contact distance."""
    s = pair_sequence(seed, nsteps=nsteps, amp=amp, **kw)
    xi, xj = s[0, 0:3], s[0, 3:6]
    u = (xj - xi) / np.linalg.norm(xj - xi)
    s[0, 3:6] = xi + 2.0 * R_PART * u
    return s


def far_sequence(seed, nsteps=200, lo=1.8, hi=4.5):
    """Centre distances from lo*r to hi*r: crosses h0 saturation, the 3r
    cutoff and the free range of the Hamaker functors."""
    rng = np.random.default_rng(seed)
    out = np.zeros((nsteps, 18))
    base = rng.normal(0.0, 1e-6, 3)
    for k in range(nsteps):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        d = R_PART * (lo + (hi - lo) * k / max(nsteps - 1, 1))
        out[k, 0:3] = base
        out[k, 3:6] = base + d * u
        out[k, 6:18] = rng.normal(0.0, 1.0, 12)
    return out


def facet_sequence(seed, nsteps=200, verts=None):
    """One particle moving above/into a triangle. Rows: x, v, w."""
    rng = np.random.default_rng(seed)
    out = np.zeros((nsteps, 9))
    for k in range(nsteps):
        # sweep over the plane, including points outside the triangle
        px, py = rng.uniform(-2.0, 2.0, 2) * 3 * R_PART
        h = R_PART * (1.0 + 0.04 * np.sin(2 * np.pi * k / 37.0)) if k % 3 else R_PART * rng.uniform(0.9, 1.4)
        out[k, 0:3] = [px, py, h]
        out[k, 3:6] = rng.normal(0.0, 1.0, 3)
        out[k, 6:9] = rng.normal(0.0, 1.0e8, 3)
    return out
