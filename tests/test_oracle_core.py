"""CPU pins of the oracle (no GPU): every oracle function is checked against a
known-answer vector, a closed form, brute force or an invariant that a
plausible mistake (sign, shift direction, transposed operand, dropped term)
would break.  Mirrors the reference's own tests where they exist
(Tests/FermionMatrix/test_fermionMatrixHubbardExp.cpp, Tests/Action/
test_force.cpp, Tests/LinAlg/Solver/test_CG.cpp, Tests/Lattice/test_honeycomb.cpp).
"""

import math

import numpy as np
import pytest
import torch

from oracle import correlator, fermion as F, hmc, lattice, rng, solvers


# ------------------------------------------------------------------ RNG
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for c, k, want in kat:
        got = rng.philox4x32_10(np.array([c], np.uint32), np.array([k], np.uint32))[0]
        assert tuple(int(v) for v in got) == want


def test_normal_moments():
    z = rng.normal(7, 3, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    u = rng.uniform(7, 4, 100000)
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.01


# -------------------------------------------------------------- lattices
def honeycomb_bloch(L0, L1):
    """Closed-form spectrum of the periodic honeycomb: +-|1 + e^{ik.a0} + e^{ik.a1}|
    with k.a_i = 2 pi m_i / L_i."""
    ev = []
    for m0 in range(L0):
        for m1 in range(L1):
            f = abs(1 + np.exp(2j * np.pi * m0 / L0) + np.exp(2j * np.pi * m1 / L1))
            ev += [f, -f]
    return np.sort(ev)


@pytest.mark.parametrize("L", [(3, 3), (4, 3), (6, 6), (5, 7)])
def test_honeycomb_closed_form(L):
    lat = lattice.honeycomb(*L)
    K = lat.hops.real.numpy()
    assert K.shape == (2 * L[0] * L[1],) * 2
    assert np.allclose(K, K.T)
    assert (K.sum(1) == 3).all()                              # valence 3 (test_honeycomb.cpp)
    ev = np.linalg.eigvalsh(K)
    assert np.allclose(ev, honeycomb_bloch(*L), atol=1e-12)
    zero = int((np.abs(ev) < 1e-10).sum())
    assert zero == (4 if L[0] % 3 == 0 and L[1] % 3 == 0 else 0) or not (L[0] % 3 == 0 and L[1] % 3 == 0)


def test_ring_matches_reference_corner_order():
    K = lattice.ring(5).hops.real.numpy()
    assert (K.sum(1) == 2).all() and K[0, 4] == 1 and K[4, 0] == 1
    assert np.allclose(np.linalg.eigvalsh(K), np.sort(2 * np.cos(2 * np.pi * np.arange(5) / 5)))


@pytest.mark.parametrize("n,cells", [(6, 4), (6, 7), (3, 2)])
def test_armchair_nanotube(n, cells):
    lat = lattice.armchair_nanotube(n, cells)
    K = lat.hops.real.numpy()
    N = K.shape[0]
    assert N == 4 * n * cells
    deg = K.sum(1)
    assert set(np.unique(deg)) <= {2.0, 3.0}
    assert (deg == 2).sum() == 4 * n                            # 2n edge sites at each open end
    ev = np.linalg.eigvalsh(K)
    assert np.allclose(np.sort(ev), np.sort(-ev), atol=1e-10)  # bipartite: symmetric spectrum
    assert np.abs(ev).max() <= 3 + 1e-12


def test_exp_hopping_matrix():
    lat = lattice.honeycomb(3, 3)
    for a in (0.25, -0.25, 0.75):
        E = lat.exp_hopping_matrix(a)
        w, U = np.linalg.eigh(lat.hops.real.numpy())
        Eref = (U * np.exp(a * w)) @ U.T
        assert np.abs(E.numpy() - Eref).max() < 1e-13
    I = lat.exp_hopping_matrix(0.3) @ lat.exp_hopping_matrix(-0.3)
    assert torch.allclose(I, torch.eye(18, dtype=torch.complex128), atol=1e-14)


# ---------------------------------------------------------- fermion matrix
def _setup(lat, Nt, beta, seed=3, std=0.7, mu=0.0, species=F.PARTICLE, complex_phi=False):
    phi = torch.as_tensor(rng.real_field(seed, 0, (Nt, lat.sites()), std))
    if complex_phi:
        phi = phi + 0.3j * torch.as_tensor(rng.normal(seed, 1, Nt * lat.sites()).reshape(Nt, -1))
    h = F.HubbardExp(lat, Nt, beta, mu, species)
    h.populate(phi)
    return h, phi


@pytest.mark.parametrize("species", [F.PARTICLE, F.HOLE])
@pytest.mark.parametrize("lat,Nt", [(lattice.ring(4), 4), (lattice.ring(6), 8), (lattice.honeycomb(2, 2), 6)])
def test_M_against_index_formula(lat, Nt, species):
    """brute force: dense matrix from the index formula of hubbardExp.hpp:22-31"""
    h, phi = _setup(lat, Nt, 2.0, mu=0.1, species=species, complex_phi=True)
    D = h.M_dense()
    D2 = F.index_formula_dense(lat, Nt, 2.0, phi, mu=0.1, species=species)
    assert (D - D2).abs().max() < 1e-14
    N = Nt * lat.sites()
    psi = torch.as_tensor(rng.complex_normal(9, 0, (3, Nt, lat.sites())))
    Dm = D.reshape(N, N)
    md = h.Mdagger(psi).reshape(3, N)
    assert (md - (Dm.conj().T @ psi.reshape(3, N).T).T).abs().max() < 1e-13
    assert (h.MMdagger(psi) - h.M(h.Mdagger(psi))).abs().max() < 1e-13
    assert (h.MdaggerM(psi) - h.Mdagger(h.M(psi))).abs().max() < 1e-13
    A = (Dm @ Dm.conj().T)
    assert (A - A.conj().T).abs().max() < 1e-13                 # hermiticity


def test_batched_equals_single():
    lat = lattice.honeycomb(2, 2)
    h, _ = _setup(lat, 6, 2.0)
    psi = torch.as_tensor(rng.complex_normal(11, 0, (5, 6, 8)))
    for op in (h.M, h.Mdagger, h.MMdagger, h.MdaggerM):
        b = op(psi)
        for k in range(5):
            assert (b[k] - op(psi[k])).abs().max() < 1e-14


def test_free_matsubara_spectrum():
    """phi = 0: eigenvalues of M are 1 - e^{s delta eps_k} e^{i w_n}, w_n = (2n+1) pi / Nt."""
    lat = lattice.ring(5)
    Nt, beta = 6, 2.0
    h = F.HubbardExp(lat, Nt, beta)
    h.populate(torch.zeros(Nt, 5))
    ev = torch.linalg.eigvals(h.M_dense().reshape(30, 30)).numpy()
    eps = np.linalg.eigvalsh(lat.hops.real.numpy())
    w = (2 * np.arange(Nt) + 1) * np.pi / Nt
    ref = (1 - np.exp(beta / Nt * eps)[:, None] * np.exp(1j * w)[None, :]).ravel()
    for e in ref:
        assert np.min(np.abs(ev - e)) < 1e-12


@pytest.mark.parametrize("species", [F.PARTICLE, F.HOLE])
def test_logdet_pins(species):
    lat = lattice.honeycomb(2, 3)
    Nt, beta = 8, 3.0
    h, phi = _setup(lat, Nt, beta, species=species, complex_phi=True)
    ld = h.logDetM()
    sign, la = torch.linalg.slogdet(h.M_dense().reshape(Nt * 12, Nt * 12))
    ref = complex(float(la), math.atan2(float(sign.imag), float(sign.real)))
    assert abs(ld - ref) < 1e-10
    for nc in (1, 2, 3, 8):
        assert abs(h.logdet_stable(nc) - ref) < 1e-10
    # time-shift and 2 pi invariance (test_fermionMatrixHubbardExp.cpp:475-545)
    h2 = F.HubbardExp(lat, Nt, beta, species=species)
    h2.populate(torch.roll(phi, 3, 0))
    assert abs(h2.logDetM() - ld) < 1e-10
    h2.populate(phi + 2 * math.pi * torch.randint(0, 5, phi.shape).to(torch.float64))
    d = h2.logDetM() - ld
    assert abs(d.real) < 1e-10 and abs(math.remainder(d.imag, 2 * math.pi)) < 1e-10


def test_logdet_closed_forms_and_large_beta():
    lat = lattice.honeycomb(3, 3)
    eps = np.linalg.eigvalsh(lat.hops.real.numpy())
    for Nt, beta in ((16, 4.0), (32, 40.0)):
        h = F.HubbardExp(lat, Nt, beta)
        h.populate(torch.zeros(Nt, 18))
        exact = float(np.sum(np.logaddexp(0.0, beta * eps)))   # logdet(1 + e^{beta K})
        assert abs(h.logdet_stable() - exact) < 1e-11 * max(1, exact)
        if beta < 10:
            assert abs(h.logDetM() - exact) < 1e-9
        # uniform time slices: logdet(1 + e^{i sum phi_t} e^{beta K})
        th = torch.as_tensor(rng.uniform(5, 0, Nt))
        phi = th[:, None].repeat(1, 18).to(torch.complex128)
        h.populate(phi)
        Phi = float(th.sum())
        exact_u = complex(np.sum(np.log(1 + np.exp(1j * Phi) * np.exp(beta * eps))))
        got = h.logdet_stable()
        assert abs(got.real - exact_u.real) < 1e-9 * max(1, abs(exact_u.real))
        assert abs(math.remainder(got.imag - exact_u.imag, 2 * math.pi)) < 1e-8


def test_grad_logdet_relation():
    """gradLogDetM = i - s d(logdet)/d(phi) per species (the reference's
    convention; checked by central differences)."""
    lat = lattice.ring(4)
    Nt, beta = 6, 2.0
    for species, s in ((F.PARTICLE, 1), (F.HOLE, -1)):
        h, phi = _setup(lat, Nt, beta, species=species)
        g = h.gradLogDetM()
        for (t, x) in ((0, 0), (2, 3), (5, 1)):
            e = 1e-6
            pp, pm = phi.clone(), phi.clone()
            pp[t, x] += e
            pm[t, x] -= e
            hp = F.HubbardExp(lat, Nt, beta, species=species); hp.populate(pp)
            hm = F.HubbardExp(lat, Nt, beta, species=species); hm.populate(pm)
            fd = (hp.logDetM() - hm.logDetM()) / (2 * e)
            assert abs(g[t, x] - (1j - s * fd)) < 1e-7


@pytest.mark.parametrize("species", [F.PARTICLE, F.HOLE])
def test_hubbard_diag_pins(species):
    """HubbardDiag (hubbardDiag.tpp): M against the brute-force index formula,
    M^dagger = (M_dense)^H, the reference's term-by-term MMdagger/MdaggerM
    against M(Mdagger .)/Mdagger(M .), logDetM against the LU of the dense
    matrix (mu = 0), logdet_stable against logDetM, and gradLogDetM =
    i - s d(logDetM)/d(phi) by central differences (the HubbardExp relation)."""
    lat = lattice.ring(4)
    Nt, beta, s = 6, 2.0, (1 if species == F.PARTICLE else -1)
    phi = torch.as_tensor(rng.real_field(3, 0, (Nt, 4), 0.8))
    h = F.HubbardDiag(lat, Nt, beta, species=species)
    h.populate(phi)
    Md = h.M_dense().reshape(Nt * 4, Nt * 4)
    Mi = F.diag_index_formula_dense(lat, Nt, beta, phi, species=species).reshape(Nt * 4, Nt * 4)
    assert (Md - Mi).abs().max() < 1e-14
    psi = torch.as_tensor(rng.complex_normal(8, 0, (2, Nt, 4)))
    ref = (Md.conj().T @ psi.reshape(2, -1).T).T.reshape(2, Nt, 4)
    assert (h.Mdagger(psi) - ref).abs().max() < 1e-13
    assert (h.MMdagger(psi) - h.M(h.Mdagger(psi))).abs().max() < 1e-13
    assert (h.MdaggerM(psi) - h.Mdagger(h.M(psi))).abs().max() < 1e-13
    sign, logabs = torch.linalg.slogdet(Md)
    ld = h.logDetM()
    assert abs(ld.real - float(logabs)) < 1e-12
    dang = ld.imag - math.atan2(float(sign.imag), float(sign.real))
    assert abs(dang - 2 * math.pi * round(dang / (2 * math.pi))) < 1e-12
    st = h.logdet_stable(3)
    assert abs(st.real - ld.real) < 1e-12
    assert abs(math.remainder(st.imag - ld.imag, 2 * math.pi)) < 1e-12
    g = h.gradLogDetM()
    for (t, x) in ((0, 0), (2, 3), (5, 1)):
        e = 1e-6
        pp, pm = phi.clone(), phi.clone()
        pp[t, x] += e
        pm[t, x] -= e
        hp = F.HubbardDiag(lat, Nt, beta, species=species); hp.populate(pp)
        hm = F.HubbardDiag(lat, Nt, beta, species=species); hm.populate(pm)
        fd = (hp.logDetM() - hm.logDetM()) / (2 * e)
        assert abs(g[t, x] - (1j - s * fd)) < 1e-7


def test_action_force_finite_differences():
    """Tests/Action/test_force.cpp: Nt=8, ring(2), beta=2, U=3, real phi."""
    lat = lattice.ring(2)
    A = hmc.HubbardAction(lat, 8, 2.0, 3.0)
    phi = torch.as_tensor(rng.real_field(21, 0, (8, 2), 1.0)) * A.Utilde
    g = A.grad(phi)
    assert g.imag.abs().max() < 1e-12                          # force is real for real phi
    for t in range(8):
        for x in range(2):
            pp, pm = phi.clone(), phi.clone()
            pp[t, x] += 1e-6
            pm[t, x] -= 1e-6
            fd = (A.eval(pp) - A.eval(pm)) / 2e-6
            assert abs(fd - g[t, x]) < 1e-6


def test_hmc_generate_states():
    """HMC::generate restated: with saveFrequency 1 the saved states are the
    trajectory-by-trajectory chain (pinned against Delta H elsewhere); with
    saveFrequency 2 every other one; LastState after Nconf * save - 1
    trajectories is that chain's state; a chain resumed from a saved state
    and its markov time continues it exactly."""
    lat = lattice.ring(4)
    A = hmc.HubbardAction(lat, 6, 2.0, 2.0)
    phi0 = torch.as_tensor(rng.real_field(9, 0, (6, 4), 0.8))
    chain = hmc.hmc_chain(A, phi0, 6, 1.0, 4, 77)
    MC = hmc.hmc_generate(A, phi0, 7, 1.0, 4, 77)
    assert [m["markov_time"] for m in MC] == list(range(7))
    for k in range(6):
        assert torch.equal(MC[k + 1]["phi"], chain[k]["phi"]) and MC[k + 1]["accepted"] == chain[k]["accepted"]
    MC2 = hmc.hmc_generate(A, phi0, 4, 1.0, 4, 77, save_frequency=2)
    for n in range(4):
        assert torch.equal(MC2[n]["phi"], MC[2 * n]["phi"]) and MC2[n]["markov_time"] == 2 * n
    last = hmc.hmc_generate(A, phi0, 3, 1.0, 4, 77, save_frequency=2, all_states=False)[0]
    assert torch.equal(last["phi"], MC[5]["phi"]) and last["markov_time"] == 5
    res = hmc.hmc_generate(A, MC[3]["phi"], 4, 1.0, 4, 77, markov_time=3)
    for n in range(4):
        assert torch.equal(res[n]["phi"], MC[3 + n]["phi"])


def test_leapfrog_reversible_and_energy():
    lat = lattice.ring(3)
    A = hmc.HubbardAction(lat, 6, 2.0, 2.0)
    phi = torch.as_tensor(rng.real_field(2, 0, (6, 3), 0.5))
    p = hmc.momentum_draw(4, 0, (6, 3))
    q1, p1 = hmc.leapfrog_real_force(A, phi, p, 1.0, 10)
    q0, p0 = hmc.leapfrog_real_force(A, q1, -p1, 1.0, 10)
    assert (q0 - phi).abs().max() < 1e-10 and (p0 + p).abs().max() < 1e-10
    dH = []
    for nmd in (10, 20):
        q1, p1 = hmc.leapfrog_real_force(A, phi, p, 1.0, nmd)
        H0 = A.eval(phi) + 0.5 * complex((p * p).sum())
        H1 = A.eval(q1) + 0.5 * complex((p1 * p1).sum())
        dH.append(abs((H1 - H0).real))
    assert dH[1] < dH[0] / 3                                   # O(eps^2)


# ------------------------------------------------------------------ solvers
def test_cg_random_hpd():
    g = torch.Generator().manual_seed(0)
    n = 40
    X = torch.randn(n, n, dtype=torch.complex128, generator=g)
    A = X @ X.conj().T + n * torch.eye(n, dtype=torch.complex128)
    b = torch.randn(n, dtype=torch.complex128, generator=g)
    for solver in (solvers.cg, solvers.bicgstab):
        x, info = solver(lambda v: A @ v, b, eps=1e-12)
        assert info["converged"]
        assert (x - torch.linalg.solve(A, b)).abs().max() < 1e-10


def test_cg_fermion_matrix_and_not_converged():
    lat = lattice.honeycomb(3, 3)
    h, _ = _setup(lat, 16, 4.0, std=math.sqrt(0.5))
    b = torch.as_tensor(rng.complex_normal(5, 0, (2, 16, 18)))
    x, info = solvers.cg(h.MMdagger, b, eps=1e-9)
    assert info["converged"]
    assert ((h.MMdagger(x) - b).norm() / b.norm()) < 1e-9
    x1, info1 = solvers.cg(h.MMdagger, b, eps=1e-12, maxIter=3)
    assert not info1["converged"] and info1["iterations"] == 3


def test_cg_preconditioned_pins():
    """CGpreconditioned.tpp:14-118.  P = I: the inner CG from x0 = r stops at
    once with z = r, so the outer recurrence must be plain CG iterate for
    iterate; P = A: z = A^-1 r, one or two outer steps; a random HPD P: the
    direct solve; a fermion-matrix system preconditioned by the free
    (phi = 0) matrix, as in the reference's example: its FP64 true residual."""
    g = torch.Generator().manual_seed(1)
    n = 40
    X = torch.randn(n, n, dtype=torch.complex128, generator=g)
    A = X @ X.conj().T + n * torch.eye(n, dtype=torch.complex128)
    Y = torch.randn(n, n, dtype=torch.complex128, generator=g)
    P = Y @ Y.conj().T + n * torch.eye(n, dtype=torch.complex128)
    b = torch.randn(n, dtype=torch.complex128, generator=g)
    mat = lambda M: (lambda v: M @ v)
    x_cg, i_cg = solvers.cg(mat(A), b, eps=1e-12)
    x_pi, i_pi = solvers.cg_preconditioned(mat(A), lambda v: v, b, eps=1e-12)
    assert i_pi["iterations"] == i_cg["iterations"] and i_pi["inner_iterations"] == 0
    assert torch.equal(x_pi, x_cg)
    x_pa, i_pa = solvers.cg_preconditioned(mat(A), mat(A), b, eps=1e-12)
    assert i_pa["converged"] and i_pa["iterations"] <= 2
    x_pp, i_pp = solvers.cg_preconditioned(mat(A), mat(P), b, eps=1e-12)
    assert i_pp["converged"]
    assert (x_pp - torch.linalg.solve(A, b)).abs().max() < 1e-10
    lat = lattice.honeycomb(3, 3)
    h, _ = _setup(lat, 16, 4.0, std=math.sqrt(0.5))
    h0 = F.HubbardExp(lat, 16, 4.0)
    h0.populate(torch.zeros(16, 18))
    bf = torch.as_tensor(rng.complex_normal(6, 0, (2, 16, 18)))
    x, info = solvers.cg_preconditioned(h.MMdagger, h0.MMdagger, bf, eps=1e-10)
    assert info["converged"] and info["inner_iterations"] > 0
    assert ((h.MMdagger(x) - bf).norm() / bf.norm()) < 1e-9


def honeycomb_walls(L, ks):
    """Band-projected plane waves w[k, s, x] = exp(i k.r_x) / L on sublattice
    s, for momenta ks = (m C + n D) / L of the L x L honeycomb."""
    lat = lattice.honeycomb(L, L)
    pos = torch.as_tensor(lat.sites_xyz[:, :2], dtype=torch.float64)
    C = torch.tensor([4 * math.pi / 3, 0.0], dtype=torch.float64)
    D = torch.tensor([-2 * math.pi / 3, 2 * math.pi / math.sqrt(3.0)], dtype=torch.float64)
    mom = torch.stack([(m * C + n * D) / L for (m, n) in ks])
    Nx = lat.sites()
    w = torch.zeros((len(ks), 2, Nx), dtype=torch.complex128)
    for s in range(2):
        sel = torch.arange(Nx) % 2 == s
        ph = torch.exp(1j * (mom @ pos[sel].T))
        w[:, s, sel] = ph / L
    return lat, w, mom


def test_wall_correlators_linearity():
    """measureK against the point-source correlator contracted with the
    walls; the two-body pools against a dense inverse of M (both species,
    both the MM^dagger and the M^dagger M pools); the quadruple count of a
    full 3 x 3 momentum grid (momentum addition is a group: 9^3 of 9^4)."""
    L, Nt, beta = 3, 8, 2.0
    lat, w, mom = honeycomb_walls(L, [(0, 0), (1, 0), (2, 0)])
    Nx = lat.sites()
    phi = torch.as_tensor(rng.real_field(41, 0, (Nt, Nx), 0.5))
    h = F.HubbardExp(lat, Nt, beta)
    h.populate(phi)
    wall = w[1]                                              # two walls (bands) of momentum 1
    corrK, block = correlator.two_point_wall(h, Nt, 2, wall)
    corr, _ = correlator.two_point(h, Nt, Nx, 2)
    lin = torch.einsum("sx,tyx->sty", wall, corr)
    assert (corrK - lin).abs().max() < 1e-9 * lin.abs().max()
    assert (block - torch.einsum("ax,btx->tab", torch.conj(wall), lin)).abs().max() < 1e-9 * block.abs().max()
    # pools of tsrc = 3 against inv(M_dense)
    tsrc = 3

    def make(species):
        o = F.HubbardExp(lat, Nt, beta, species=species)
        o.populate(phi)
        return o

    pools = correlator.two_body_pools(make, phi, w, tsrc)
    for name, species in (("P", F.PARTICLE), ("H", F.HOLE)):
        Mi = torch.linalg.inv(make(species).M_dense().reshape(Nt * Nx, Nt * Nx)).reshape(Nt, Nx, Nt, Nx)
        Mdi = torch.linalg.inv(make(species).M_dense().reshape(Nt * Nx, Nt * Nx).conj().T).reshape(Nt, Nx, Nt, Nx)
        ts = (torch.arange(Nt) + tsrc) % Nt
        sg = torch.where(torch.arange(Nt) + tsrc >= Nt, -1.0, 1.0).to(torch.complex128)
        G = Mi[ts][:, :, tsrc, :] * sg[:, None, None]           # [t, x, y] = sign M^-1[(t+tsrc, x), (tsrc, y)]
        Gd = Mdi[ts][:, :, tsrc, :] * sg[:, None, None]
        C = torch.einsum("aix,txy,bjy->abtij", torch.conj(w), G, w)
        Dg = torch.conj(torch.einsum("aix,txy,bjy->abtij", w, Gd, torch.conj(w)))
        assert (pools["C" + name] - C).abs().max() < 1e-9 * C.abs().max()
        assert (pools["D" + name] - Dg).abs().max() < 1e-9 * Dg.abs().max()
    _, _, mom9 = honeycomb_walls(3, [(m, n) for m in range(3) for n in range(3)])
    assert len(correlator.two_body_quads(mom9)) == 9 ** 3


def test_correlator_free_propagator():
    lat = lattice.ring(4)
    Nt, beta = 8, 2.0
    h = F.HubbardExp(lat, Nt, beta)
    h.populate(torch.zeros(Nt, 4))
    corr, _ = correlator.two_point(h, Nt, 4, 2)
    Minv = torch.linalg.inv(h.M_dense().reshape(Nt * 4, Nt * 4)).reshape(Nt, 4, Nt, 4)
    # free propagator depends on t - tsrc with the antiperiodic sign
    ref = torch.zeros(Nt, 4, 4, dtype=torch.complex128)
    for tsrc in (0, 4):
        for t in range(Nt):
            tt = (t + tsrc) % Nt
            sgn = -1.0 if t + tsrc >= Nt else 1.0
            ref[t] += sgn * Minv[tt, :, tsrc, :]
    ref /= 2
    assert (corr - ref).abs().max() < 1e-10


def test_two_body_pauli():
    """Antisymmetry of the two-body channels.  The I = 1, S = 1 channels pair
    two fermions that the exchange term antisymmetrises (exchange sign -1),
    so for any field their correlator vanishes where the two sink labels
    coincide (y = z and band k = l) or the two source labels coincide (w = x
    and band i = j); the I0S1 / I1S0 channels (exchange sign +1) do not.
    A wrong exchange sign or a transposed exchange index fails this."""
    L, Nt, beta = 3, 8, 2.0
    lat, w, mom = honeycomb_walls(L, [(0, 0), (1, 0), (2, 0)])
    Nx = lat.sites()
    phi = torch.as_tensor(rng.real_field(43, 0, (Nt, Nx), 0.6))

    def make(species):
        o = F.HubbardExp(lat, Nt, beta, species=species)
        o.populate(phi)
        return o

    res = correlator.two_body(make, phi, w, mom, 2)
    b = w.shape[1]
    anti = [name for name, _, terms in correlator.TWO_BODY_TABLE if all(s == -1 for (_, _, e, s) in terms if e)]
    assert len(anti) == 8 and all(n.startswith("I1S1") for n in anti)
    scale = max(float(c.abs().max()) for ch in res.values() for c in ch.values())
    n_sink = n_src = 0
    for name, ch in res.items():
        for (qw, qx, qz, qy), c in ch.items():
            # c[t, i * b + j, l * b + k]
            if qz == qy:
                n_sink += 1
                sink = torch.stack([c[:, :, k * b + k] for k in range(b)])
                if name in anti:
                    assert float(sink.abs().max()) < 1e-12 * scale, (name, (qw, qx, qz, qy))
                else:
                    assert float(sink.abs().max()) > 1e-6 * scale, (name, (qw, qx, qz, qy))
            if qw == qx:
                n_src += 1
                src = torch.stack([c[:, i * b + i, :] for i in range(b)])
                if name in anti:
                    assert float(src.abs().max()) < 1e-12 * scale, (name, (qw, qx, qz, qy))
    assert n_sink > 0 and n_src > 0
    # away from the coincident labels no channel vanishes
    for name in anti:
        assert max(float(c.abs().max()) for c in res[name].values()) > 1e-6 * scale, name


def test_two_body_particle_hole_pairs():
    """Prefactor pin of the twelve two-body channels.  At phi = 0 and mu = 0
    on the bipartite honeycomb the free theory is symmetric under the
    particle-hole transformation c_x -> (-1)^x c_x^dagger, which exchanges the
    particle and hole pools and multiplies each band label by its sublattice
    sign.  So each Iz = +1 channel equals its Iz = -1 partner times
    s_i s_j s_l s_k (s = +1 on sublattice A, -1 on B), and the two Iz = 0
    channels of each multiplet (Sz = +1 and -1) are equal.  These six pairs
    cover all twelve channels, so a wrong value or sign of any one alpha
    fails (every channel is nonzero here).  A factor common to both members
    of a pair stays unpinned."""
    L, Nt, beta = 3, 8, 2.0
    lat, w, mom = honeycomb_walls(L, [(0, 0), (1, 0), (2, 0)])
    Nx = lat.sites()
    phi = torch.zeros((Nt, Nx), dtype=torch.float64)

    def make(species):
        o = F.HubbardExp(lat, Nt, beta, species=species)
        o.populate(phi)
        return o

    res = correlator.two_body(make, phi, w, mom, 2)
    b = w.shape[1]
    s = torch.tensor([1.0, -1.0], dtype=torch.float64)
    sub = torch.einsum("i,j,l,k->ijlk", s, s, s, s).reshape(b * b, b * b).to(torch.complex128)
    pairs = [("I1S1Iz1Sz1", "I1S1Izn1Sz1", sub), ("I1S1Iz1Sz0", "I1S1Izn1Sz0", sub),
             ("I1S1Iz1Szn1", "I1S1Izn1Szn1", sub), ("I1S0Iz1Sz0", "I1S0Izn1Sz0", sub),
             ("I1S1Iz0Sz1", "I1S1Iz0Szn1", None), ("I0S1Iz0Sz1", "I0S1Iz0Szn1", None)]
    covered = sorted(n for p in pairs for n in p[:2])
    assert covered == sorted(n for n, _, _ in correlator.TWO_BODY_TABLE)
    scale = max(float(c.abs().max()) for ch in res.values() for c in ch.values())
    for a, bname, sign in pairs:
        assert max(float(c.abs().max()) for c in res[a].values()) > 0.1 * scale, a
        for q, c in res[a].items():
            other = res[bname][q] if sign is None else res[bname][q] * sign
            assert float((c - other).abs().max()) < 1e-12 * scale, (a, bname, q)
    # the relation is not an identity of the contraction: the Iz = +1 and
    # Iz = -1 channels differ without the sublattice sign
    d = max(float((res["I1S1Iz1Sz0"][q] - res["I1S1Izn1Sz0"][q]).abs().max()) for q in res["I1S1Iz1Sz0"])
    assert d > 0.1 * scale
