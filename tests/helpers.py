"""Model builders for hand-constructed test cases."""
import numpy as np

from lattice_microbes_amd.models import ModelSpec

TAU = 1e-6
SPACING = 1e-8


def q_to_D(q):
    """Diffusion coefficient whose transition probability is q (T = D tau / lambda^2)."""
    return np.asarray(q, dtype=np.float64) * SPACING ** 2 / TAU


def spec(shape, P=8, nS=1, nST=1, DF=None, reactions=None, particles=(), sites=None, boundary="periodic"):
    """particles: iterable of (x, y, z, species) placed in order (slot order);
    reactions: list of (type, K, {species: stoich}, [dependencies]) in LM units."""
    X, Y, Z = shape
    lattice = np.zeros((X, Y, Z, P), dtype=np.uint8)
    fill = np.zeros((X, Y, Z), dtype=np.int64)
    for (x, y, z, sp) in particles:
        lattice[x, y, z, fill[x, y, z]] = sp
        fill[x, y, z] += 1
    reactions = reactions or []
    R = len(reactions)
    types = np.array([r[0] for r in reactions], dtype=np.uint32)
    K = np.array([r[1] for r in reactions], dtype=np.float64)
    S = np.zeros((nS, R), dtype=np.int32)
    D = np.zeros((nS, R), dtype=np.uint32)
    for j, (_, _, sto, deps) in enumerate(reactions):
        for s, v in sto.items():
            S[s, j] = v
        for s in deps:
            D[s, j] = 1
    if DF is None:
        DF = np.zeros((nST, nST, nS))
    RL = np.ones((R, nST), dtype=np.uint32)
    if sites is None:
        sites = np.zeros(shape, dtype=np.uint8)
    return ModelSpec("handmade", tuple(shape), SPACING, TAU, P, nS, nST, types, K, S, D,
                     np.asarray(DF, dtype=np.float64), RL, lattice, np.asarray(sites, dtype=np.uint8), boundary)


def site_contents(native, shape, P, x, y, z):
    """Particles of one site from a (P/4, Z, Y, X) native lattice, slot order."""
    words = native[:, z, y, x]
    out = []
    for w in words:
        for b in range(4):
            out.append(int((int(w) >> (8 * b)) & 0xFF))
    return out[:P]
