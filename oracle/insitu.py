"""NumPy restatement of the reference's in-situ reduced diagnostics, FP64.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Each function returns (per_slice, sums) for one slice the way the reference fills
m_insitu_rdata[islice + i*nslices] and adds into m_insitu_sum_rdata:
  fields  Fields::InSituComputeDiags          src/fields/Fields.cpp:1289-1346
  plasma  PlasmaParticleContainer::InSituComputeDiags
                                              src/particles/plasma/PlasmaParticleContainer.cpp:442-527
  beam    BeamParticleContainer::InSituComputeDiags
                                              src/particles/beam/BeamParticleContainer.cpp:475-594
Units are normalised (c = 1). Sums use numpy's pairwise summation; the device sums in
another order, so comparisons allow rounding-level differences.
"""
from __future__ import annotations

import numpy as np

CLIGHT = 1.0


def fields(ExmBy, EypBx, Ez, Bx, By, Bz, jz_beam, dx, dy, dz):
    """Fields.cpp:1323-1343: ten sums over the valid cells of one slice times dx*dy*dz.
    Arguments are the valid-cell planes (ny, nx). Returns (values[10], values[10]): the
    per-slice entry and the amount added to the integrated sums are the same number."""
    q = [
        (ExmBy + By * CLIGHT) ** 2,  # [Ex^2]
        (EypBx - Bx * CLIGHT) ** 2,  # [Ey^2]
        Ez ** 2,                     # [Ez^2]
        Bx ** 2,                     # [Bx^2]
        By ** 2,                     # [By^2]
        Bz ** 2,                     # [Bz^2]
        ExmBy ** 2,                  # [ExmBy^2]
        EypBx ** 2,                  # [EypBx^2]
        jz_beam,                     # [jz_beam]
        Ez * jz_beam,                # [Ez*jz_beam]
    ]
    dxdydz = dx * dy * dz
    v = np.array([float(np.sum(a)) * dxdydz for a in q])
    return v, v.copy()


def _select(valid, x, y, radius):
    return (np.asarray(valid) != 0) & ~(x * x + y * y > radius * radius)


def plasma(x, y, w, ux, uy, psi, valid, radius=np.inf):
    """PlasmaParticleContainer.cpp:471-526. Returns (per_slice[14], sums[14], Np): the
    per-slice moments 1..12 are divided by sum(w); sum(w) and [(ga-1)*(1-vz)] are not."""
    m = _select(valid, x, y, radius)
    x, y, w, ux, uy, psi = (np.asarray(a, dtype=np.float64)[m] for a in (x, y, w, ux, uy, psi))
    ux = ux / CLIGHT
    uy = uy / CLIGHT
    gamma = (1.0 + ux * ux + uy * uy + psi * psi) / (2.0 * psi)
    uz = gamma - psi
    wq = w * gamma / psi
    energy = w * (gamma - 1.0)
    terms = [wq, wq * x, wq * x * x, wq * y, wq * y * y, wq * ux, wq * ux * ux, wq * uy, wq * uy * uy,
             wq * uz, wq * uz * uz, wq * gamma, wq * gamma * gamma, energy]
    s = np.array([float(np.sum(t)) for t in terms])
    inv = 0.0 if s[0] <= 0.0 else 1.0 / s[0]
    per = s.copy()
    for i in range(1, 13):
        per[i] = s[i] * inv
    return per, s, int(m.sum())


def beam(x, y, z, ux, uy, uz, w, valid, radius=np.inf):
    """BeamParticleContainer.cpp:491-553. Returns (per_slice[22], sums[22], Np): every
    per-slice moment but sum(w) is divided by sum(w)."""
    m = _select(valid, x, y, radius)
    x, y, z, ux, uy, uz, w = (np.asarray(a, dtype=np.float64)[m] for a in (x, y, z, ux, uy, uz, w))
    ux, uy, uz = ux / CLIGHT, uy / CLIGHT, uz / CLIGHT
    with np.errstate(divide="ignore"):
        uz_inv = np.where(uz == 0.0, 0.0, 1.0 / np.where(uz == 0.0, 1.0, uz))
    gamma = np.sqrt(1.0 + ux * ux + uy * uy + uz * uz)
    terms = [w, w * x, w * x * x, w * y, w * y * y, w * z, w * z * z, w * ux, w * ux * ux, w * uy, w * uy * uy,
             w * uz, w * uz * uz, w * x * ux, w * y * uy, w * z * uz, w * x * uy, w * y * ux, w * ux * uz_inv,
             w * uy * uz_inv, w * gamma, w * gamma * gamma]
    s = np.array([float(np.sum(t)) for t in terms])
    inv = 0.0 if s[0] <= 0.0 else 1.0 / s[0]
    per = s * inv
    per[0] = s[0]
    return per, s, int(m.sum())
