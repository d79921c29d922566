"""Deterministic density-matrix guess for the synthetic workloads.

The reference obtains P from its own SCF (Hueckel guess, SCF iterations);
that machinery is outside the XC path, so the workloads use a documented
stand-in: an atomic-occupation diagonal (every atom's electrons placed on its
own AOs: 1s core 2.0 on the first s function of heavy atoms, the valence s
and p electrons split 0.6 / 0.4 between the inner and outer contracted
functions, hydrogen 0.6 / 0.4, 0.01 on each d function taken from the outer
valence s, so that the occupations sum to the electron count), plus a
seeded symmetric off-diagonal part u_ij * eps * exp(-r_AB^2 / 4) coupling
AOs i, j on atoms A, B (u ~ U(-1, 1)).  With eps = 0 and normalised AOs the
integral of rho equals sum_i P_ii, which the tests use.  Returned as the
dense total density P (the engine is called with P_alpha = P / 2, as
DfXCFunctional::buildXcMatrix does, DfXCFunctional.cpp:134-139).
"""
import numpy as np

VALENCE = {"H": 1, "C": 4, "N": 5, "O": 6}


def atomic_occupations(mol, info):
    occ = np.zeros(info.n_aos)
    for a, s in enumerate(mol.symbols):
        idx = np.where(info.ao_atom == a)[0]
        types = info.ao_type[idx]
        s_idx = idx[types == 0]
        p_idx = [idx[types == t] for t in (1, 2, 3)]
        d_idx = idx[types >= 4]
        if s == "H":
            occ[s_idx[0]] += 0.6
            occ[s_idx[1]] += 0.4
            continue
        occ[s_idx[0]] += 2.0
        occ[s_idx[1]] += 1.2
        occ[s_idx[2]] += 0.8 - 0.01 * len(d_idx)
        npe = VALENCE[s] - 2
        for comp in p_idx:
            occ[comp[0]] += npe / 3.0 * 0.6
            occ[comp[1]] += npe / 3.0 * 0.4
        occ[d_idx] += 0.01
    return occ


def density_guess(mol, info, seed=0, eps=0.05):
    occ = atomic_occupations(mol, info)
    P = np.diag(occ)
    if eps:
        rng = np.random.default_rng(seed)
        n = info.n_aos
        U = rng.uniform(-1.0, 1.0, size=(n, n))
        U = np.tril(U, -1)
        U = U + U.T
        c = info.ao_center
        r2 = ((c[:, None, :] - c[None, :, :]) ** 2).sum(-1)
        P = P + eps * U * np.exp(-r2 / 4.0)
    return P


def pack_lower(M):
    """Dense symmetric -> RLHD packed lower triangle (row r, col c <= r)."""
    r, c = np.tril_indices(M.shape[0])
    return np.ascontiguousarray(M[r, c])


def unpack_lower(v, n):
    M = np.zeros((n, n))
    r, c = np.tril_indices(n)
    M[r, c] = v
    M[c, r] = v
    return M
