"""The augmented velocity block of the mixed solve, checked on the CPU with the
oracle's assembled K, G and F (16x16, Ra 1e6, viscosity contrast 1e3, the
benchmark's seeded temperature).

The device builds K_al = K + gamma G W G^T (W = 1 / BAB of the plain K,
BAB_e = sum_i G_ie^2 / K_ii) in K's ELL slots and returns p + gamma W G^T u.
These tests pin the facts it rests on: G W G^T has no entry outside K's
pattern; the augmented saddle system's velocity is the plain one's and its
pressure plus gamma W G^T u is the plain one's; and the Schur complement's
spectrum relative to BAB narrows (what cuts the Uzawa count).  A dropped
gamma, a W taken as BAB instead of 1 / BAB, or a shift with the wrong sign
fails one of them.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

GAMMA = 3.0


@pytest.fixture(scope="module")
def system():
    import bench
    from mantle_convection_lab_amd import Input, Mesh
    from tests.oracle_lib import VW, Oracle, csr_from_slots, ctrl_from_control, mesh_from_host_level
    keys = bench.path_keys("mixed")
    keys.update(level=5)
    inp = Input.from_keys(keys)
    mesh = Mesh(inp)
    T0 = mesh.initial_temperature(-1)
    lv = mesh.level(-1)
    o = Oracle(mesh_from_host_level(lv), ctrl_from_control(inp.control(), schur_divide=1))
    A = o.assemble(o.viscosity(T0), o.buoyancy(T0))
    n, nel = lv["neq"], lv["nel"]
    ptr, col, val = csr_from_slots(A["Kncol"], A["Kcol"], A["Kval"], VW)
    K = sp.csr_matrix((val, col, ptr), shape=(n, n))
    ptr, col, val = csr_from_slots(A["Gncol"], A["Gcol"], A["Gval"], VW)
    G = sp.csr_matrix((val, col, ptr), shape=(n, nel))
    bab = np.asarray(G.multiply(G).T @ (1.0 / K.diagonal())).ravel()
    W = 1.0 / bab
    Kal = (K + GAMMA * (G @ sp.diags(W) @ G.T)).tocsr()
    return K, G, A["F"].copy(), W, Kal


def _saddle(K, G, F):
    """Exact solve of K u + G p = F, G^T u = 0 with mean(p) = 0 (the closed
    box leaves the constant pressure free): a bordered dense system."""
    n, m = G.shape
    M = np.zeros((n + m + 1, n + m + 1))
    M[:n, :n] = K.toarray()
    M[:n, n:n + m] = G.toarray()
    M[n:n + m, :n] = G.toarray().T
    M[n + m, n:n + m] = 1.0
    M[n:n + m, n + m] = 1.0
    rhs = np.zeros(n + m + 1)
    rhs[:n] = F
    x = np.linalg.solve(M, rhs)
    return x[:n], x[n:n + m]


def test_augmented_term_stays_in_the_pattern_of_K(system):
    K, G, F, W, Kal = system
    pattern = set(zip(*K.nonzero()))
    extra = [ij for ij in zip(*Kal.nonzero()) if ij not in pattern]
    assert not extra
    # and it is not empty: the term really enters K's slots
    assert abs(Kal - K).max() > 1e-3 * abs(K).max()


def test_shifted_pressure_solves_the_plain_system(system):
    K, G, F, W, Kal = system
    u, p = _saddle(K, G, F)
    ua, pa = _saddle(Kal, G, F)
    assert np.linalg.norm(ua - u) <= 1e-10 * np.linalg.norm(u)
    shifted = pa + GAMMA * W * (G.T @ ua)
    shifted -= shifted.mean()
    assert np.linalg.norm(shifted - p) <= 1e-9 * np.linalg.norm(p)
    # momentum residual of the plain system with the shifted pressure, for a
    # velocity that is not exactly divergence-free (as after a stopped
    # iteration): zero up to rounding
    v = ua + 1e-6 * np.random.default_rng(1).standard_normal(u.size)
    res_al = F - Kal @ v - G @ pa
    res_plain = F - K @ v - G @ (pa + GAMMA * W * (G.T @ v))
    assert np.linalg.norm(res_plain - res_al) <= 1e-12 * np.linalg.norm(F)


def test_schur_spectrum_narrows(system):
    K, G, F, W, Kal = system
    bab = 1.0 / W

    def spread(Kx):
        S = G.toarray().T @ np.linalg.solve(Kx.toarray(), G.toarray())
        ev = sla.eigh(S, np.diag(bab), eigvals_only=True)
        ev = ev[ev > 1e-10 * ev.max()]      # the constant pressure is S's null vector
        return ev.max() / ev.min()

    plain, aug = spread(K), spread(Kal)
    # S_al^-1 = S^-1 + gamma W: relative to BAB the eigenvalues s become
    # s / (1 + gamma s), so the spread shrinks by about (1 + gamma s_max)
    assert aug < plain / 2
