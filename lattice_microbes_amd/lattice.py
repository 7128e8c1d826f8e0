"""Host lattices in the solver's device layout.

``ByteLattice`` mirrors ``lm::rdme::ByteLattice`` as its Python module exposes
it (src/rdme/ByteLattice.cpp, src/swig/lm.i:875-896): particles of up to P
(a multiple of 4) per site, one byte each, stored as P/4 planes of 32-bit
words over the sites (slot 4w+b of site i is byte b of word w of plane w;
i = x + X*y + X*Y*z), and one site-type byte per site.  ``IntLattice``
mirrors ``lm::rdme::IntLattice`` (src/rdme/IntLattice.cpp, lm.i:898-916):
one 32-bit particle per word, P planes (slot w of site i is word w of plane
w).  The view accessors return numpy arrays over that storage, shaped as the
reference's: ``getParticleLatticeView()`` (W, Z, Y, X, particles per word:
4 bytes or 1 word) and ``getSiteLatticeView()`` (Z, Y, X).
``row_major``/``sites_row_major`` give the (x, y, z, p) file layout.
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np

from . import _lib


class Lattice:
    """Planar particle storage shared by ByteLattice and IntLattice: W planes
    of words, ``_per_word`` particles of ``_dtype`` in each."""

    _dtype = np.uint8
    _per_word = 4

    def __init__(self, xSize: int, ySize: int, zSize: int, spacing: float, particlesPerSite: int):
        if particlesPerSite <= 0 or particlesPerSite % self._per_word:
            raise _lib.InvalidArgException(_lib.ERR_ARG, f"particlesPerSite must be a positive multiple of "
                                                         f"{self._per_word}")
        self.size = (int(xSize), int(ySize), int(zSize))
        self.spacing = float(spacing)
        self.P = int(particlesPerSite)
        X, Y, Z = self.size
        self._words = np.zeros((self.P // self._per_word, Z, Y, X), dtype=np.uint32)
        self._sites = np.zeros((Z, Y, X), dtype=np.uint8)

    # ---------------------------------------------------------------- geometry
    def getXSize(self) -> int:
        return self.size[0]

    def getYSize(self) -> int:
        return self.size[1]

    def getZSize(self) -> int:
        return self.size[2]

    def getSize(self) -> Tuple[int, int, int]:
        return self.size

    def getNumberSites(self) -> int:
        X, Y, Z = self.size
        return X * Y * Z

    def getMaxOccupancy(self) -> int:
        return self.P

    def getSpacing(self) -> float:
        return self.spacing

    # ---------------------------------------------------------------- views
    def getParticleLatticeView(self) -> np.ndarray:
        W, Z, Y, X = self._words.shape
        return self._words.view(self._dtype).reshape(W, Z, Y, X, self._per_word)

    def getSiteLatticeView(self) -> np.ndarray:
        return self._sites

    # ---------------------------------------------------------------- sites and particles
    def _check_site(self, x: int, y: int, z: int) -> None:
        X, Y, Z = self.size
        if not (0 <= x < X and 0 <= y < Y and 0 <= z < Z):
            raise _lib.InvalidArgException(_lib.ERR_ARG, f"InvalidSiteException: ({x}, {y}, {z})")

    def _slots(self, x: int, y: int, z: int) -> np.ndarray:
        self._check_site(x, y, z)
        return self.getParticleLatticeView()[:, z, y, x, :].reshape(-1)

    def getSiteType(self, x: int, y: int, z: int) -> int:
        self._check_site(x, y, z)
        return int(self._sites[z, y, x])

    def setSiteType(self, x: int, y: int, z: int, siteType: int) -> None:
        self._check_site(x, y, z)
        self._sites[z, y, x] = siteType

    def getOccupancy(self, x: int, y: int, z: int) -> int:
        return int(np.count_nonzero(self._slots(x, y, z)))

    def getParticle(self, x: int, y: int, z: int, particleIndex: int) -> int:
        s = self._slots(x, y, z)
        if not 0 <= particleIndex < s.size or s[particleIndex] == 0:
            raise _lib.InvalidArgException(_lib.ERR_ARG, "InvalidParticleException: no particle at that index")
        return int(s[particleIndex])

    def addParticle(self, x: int, y: int, z: int, particle: int) -> None:
        """Into the first empty slot (ByteLattice/IntLattice::addParticle)."""
        if not 0 < particle <= np.iinfo(self._dtype).max:
            raise _lib.InvalidArgException(_lib.ERR_ARG, "InvalidParticleException: particle type out of range")
        s = self._slots(x, y, z)
        k = int(np.count_nonzero(s))
        if k >= s.size:
            raise _lib.InvalidArgException(_lib.ERR_ARG, "InvalidParticleException: the site is full")
        self.getParticleLatticeView()[k // self._per_word, z, y, x, k % self._per_word] = particle

    def removeParticles(self, x: int, y: int, z: int) -> None:
        self._check_site(x, y, z)
        self.getParticleLatticeView()[:, z, y, x, :] = 0

    def removeAllParticles(self) -> None:
        self._words[:] = 0

    def getParticleCounts(self) -> Dict[int, int]:
        v = self.getParticleLatticeView().reshape(-1)
        types, counts = np.unique(v[v != 0], return_counts=True)
        return {int(p): int(c) for p, c in zip(types, counts)}

    def findParticles(self, minParticleType: int, maxParticleType: int) -> List[Tuple[int, int, int, int, int]]:
        """(x, y, z, slot, type) of every particle with type in the range."""
        v = self.getParticleLatticeView()
        w, z, y, x, b = np.nonzero((v >= minParticleType) & (v <= maxParticleType) & (v != 0))
        pw = self._per_word
        return [(int(xi), int(yi), int(zi), int(pw * wi + bi), int(v[wi, zi, yi, xi, bi]))
                for wi, zi, yi, xi, bi in zip(w, z, y, x, b)]

    # ---------------------------------------------------------------- file layout
    def row_major(self) -> np.ndarray:
        """Particles as row-major (X, Y, Z, P)."""
        v = self.getParticleLatticeView()
        W, Z, Y, X, pw = v.shape
        return np.ascontiguousarray(np.transpose(v, (3, 2, 1, 0, 4)).reshape(X, Y, Z, pw * W))

    def sites_row_major(self) -> np.ndarray:
        return np.ascontiguousarray(np.transpose(self._sites, (2, 1, 0)))

    def setFromRowMajorData(self, data, sites=None) -> None:
        """Particles from row-major (X, Y, Z, P), each site's particles moved
        to its first slots in order (Lattice::setFromRowMajorByteData,
        IntLattice::setFromRowMajorData); optionally the site types from
        row-major (X, Y, Z)."""
        X, Y, Z = self.size
        a = np.asarray(data)
        if a.size and (a.min() < 0 or a.max() > np.iinfo(self._dtype).max):
            raise _lib.InvalidArgException(_lib.ERR_ARG, "InvalidParticleException: particle type out of range")
        lat = a.astype(self._dtype, copy=False).reshape(X, Y, Z, self.P)
        zyx = np.transpose(lat, (2, 1, 0, 3))
        order = np.argsort(zyx == 0, axis=-1, kind="stable")
        comp = np.take_along_axis(zyx, order, axis=-1)
        pw = self._per_word
        self.getParticleLatticeView()[:] = np.transpose(comp.reshape(Z, Y, X, self.P // pw, pw), (3, 0, 1, 2, 4))
        if sites is not None:
            self._sites[:] = np.transpose(np.asarray(sites, dtype=np.uint8).reshape(X, Y, Z), (2, 1, 0))


class ByteLattice(Lattice):
    """lm::rdme::ByteLattice: one byte per particle, four per word."""

    _dtype = np.uint8
    _per_word = 4

    def setFromRowMajorByteData(self, data, sites=None) -> None:
        self.setFromRowMajorData(data, sites)


class IntLattice(Lattice):
    """lm::rdme::IntLattice: one 32-bit particle per word."""

    _dtype = np.uint32
    _per_word = 1
