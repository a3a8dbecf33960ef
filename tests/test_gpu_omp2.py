"""CD-OMP2 on the GPU: the amplitude/OPDM entry point against the oracle's
restatement of cholesky_omp2.cc:94-326, and the drop-in CholeskyOMP2 loop
against the oracle's loop (cholesky_omp2.cc:427-513)."""
import os

import numpy as np
import pytest

from conftest import ROOT, check_close, to_dev, to_np
from mshqc_amd import synthetic as syn
from oracle import omp2_ref

pytestmark = pytest.mark.gpu
BD = os.path.join(ROOT, "mshqc_amd", "data", "basis")


def _spin(rng, Q, o, v):
    B = syn.gen_B(Q, o, v, rng) if o * v else np.zeros((Q, o, v))
    eo, ev = syn.gen_eps(o, v, rng)
    return B, eo, ev


@pytest.mark.parametrize("Q,oa,va,ob,vb,zblock,seed", [
    (23, 5, 11, 4, 12, None, 1),
    (31, 6, 20, 6, 20, None, 2),
    (17, 3, 9, 0, 12, None, 3),     # no beta electrons
    (19, 1, 7, 2, 6, None, 4),      # one alpha electron: no alpha pairs
    (40, 7, 30, 5, 32, 13440, 5),   # blocks of 2 occupied indices, spin-swapped pass for G_oo^b
    (24, 20, 9, 18, 11, None, 6),   # G_oo in 2 x 2 MFMA tiles
    (20, 40, 8, 38, 10, None, 7),   # 3 x 3 tiles
    (16, 60, 6, 52, 14, None, 8),   # 4 x 4 tiles
    (12, 70, 5, 66, 9, None, 9),    # o > 64: the generic batch-reduced GEMM
])
def test_omp2_opdm_matches_oracle(ctx, Q, oa, va, ob, vb, zblock, seed):
    rng = np.random.default_rng(seed)
    Ba, eoa, eva = _spin(rng, Q, oa, va)
    Bb, eob, evb = _spin(rng, Q, ob, vb)
    taa, tbb, tab = omp2_ref.t2_blocks(Ba, Bb, eoa, eva, eob, evb)
    z = np.zeros
    E_ref = (omp2_ref.energy(taa, z(tbb.shape), z(tab.shape), eoa, eva, eob, evb),
             omp2_ref.energy(z(taa.shape), tbb, z(tab.shape), eoa, eva, eob, evb),
             omp2_ref.energy(z(taa.shape), z(tbb.shape), tab, eoa, eva, eob, evb))
    G_ref = omp2_ref.opdm(taa, tbb, tab)
    print(f"omp2_opdm Q={Q} oa={oa} va={va} ob={ob} vb={vb} zblock={zblock}", flush=True)
    if zblock:
        ctx.set_zblock(zblock)
    try:
        E, G = ctx.omp2_opdm(*to_dev(Ba, Bb, eoa, eva, eob, evb))
    finally:
        ctx.set_zblock(0)
    for got, want in zip(E, E_ref):
        assert abs(got - want) <= 1e-12 * max(abs(want), 1e-300)
    for name, got, want in zip(("Goo_a", "Gvv_a", "Goo_b", "Gvv_b"), G, G_ref):
        check_close(to_np(got), want, 1e-12, name)


def test_omp2_opdm_identical_spins(ctx):
    """Identical alpha and beta inputs: the two spins' OPDM blocks agree and
    the energy equals the RMP2 energy of the same B (E_aa + E_bb + E_ab)."""
    import oracle

    rng = np.random.default_rng(6)
    B, eo, ev = _spin(rng, 29, 6, 25)
    E, G = ctx.omp2_opdm(*to_dev(B, B, eo, ev, eo, ev))
    e_rmp2 = oracle.rmp2_energy(B, eo, ev)
    assert abs(sum(E) - e_rmp2) < 1e-12 * abs(e_rmp2)
    assert E[0] == E[1]
    # the opposite-spin terms reach the two spins through different GEMM orders
    check_close(to_np(G[2]), to_np(G[0]), 1e-13, "Goo_b vs Goo_a")
    check_close(to_np(G[3]), to_np(G[1]), 1e-13, "Gvv_b vs Gvv_a")


@pytest.mark.parametrize("Q,o,v,seed", [(23, 5, 11, 1), (31, 1, 20, 2), (17, 6, 1, 3), (40, 12, 130, 4),
                                          (1, 4, 6, 5)])
def test_omp2_grad2p_matches_oracle(ctx, Q, o, v, seed):
    """Two-particle orbital-gradient term (batch-reduced GEMMs over Q; Q = 1
    takes the unsplit path) against the oracle's einsum form, which the CPU
    tests pin by finite differences of the MP2 energy."""
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((Q, o, v))
    Loo = rng.standard_normal((Q, o, o))
    Lvv = rng.standard_normal((Q, v, v))
    Loo, Lvv = Loo + Loo.transpose(0, 2, 1), Lvv + Lvv.transpose(0, 2, 1)
    print(f"omp2_grad2p Q={Q} o={o} v={v}", flush=True)
    w2 = ctx.omp2_grad2p(*to_dev(G, Loo, Lvv))
    check_close(to_np(w2), omp2_ref.two_particle_grad(G, Loo, Lvv), 1e-12, "w2")


@pytest.fixture(scope="module")
def m(ctx):
    from mshqc_amd import native_module

    return native_module()


def _setup(m, atoms, basis, mult):
    mol = m.Molecule(0, mult)
    for Z, x, y, z in atoms:
        mol.add_atom(Z, x, y, z)
    b = m.BasisSet(basis, mol, BD)
    eng = m.IntegralEngine(mol, b)
    chol = m.CholeskyERI(1e-6)
    chol.decompose(eng.compute_eri())
    n = b.n_basis_functions()
    L = np.array(chol.get_L_vectors()).reshape(-1, n, n)
    H = eng.compute_kinetic() + eng.compute_nuclear()
    return mol, b, eng, chol, L, H


def _compare(m, mol, b, eng, chol, L, H, scf, cfg):
    print(f"CD-OMP2 n={b.n_basis_functions()} na={scf.n_occ_alpha} nb={scf.n_occ_beta} "
          f"two_particle={cfg.two_particle_gradient}", flush=True)
    res = m.CholeskyOMP2(mol, b, eng, scf, cfg, chol).compute()
    ref = omp2_ref.cd_omp2(H, L, scf.C_alpha, scf.C_beta, scf.P_alpha, scf.P_beta, scf.orbital_energies_alpha,
                           scf.orbital_energies_beta, scf.n_occ_alpha, scf.n_occ_beta, scf.energy_total,
                           max_iterations=cfg.max_iterations, energy_threshold=cfg.energy_threshold,
                           gradient_threshold=cfg.gradient_threshold, two_particle=cfg.two_particle_gradient)
    print(f"  E {res.energy_total!r} oracle {ref['energy_total']!r} iters {res.iterations}/{ref['iterations']}",
          flush=True)
    assert res.iterations == ref["iterations"] and res.converged == ref["converged"]
    assert abs(res.energy_total - ref["energy_total"]) < 1e-10
    assert abs(res.energy_mp2_corr - ref["e_mp2"]) < 1e-10
    assert abs(res.energy_mp2_os - ref["e_os"]) < 1e-10 and abs(res.energy_mp2_ss - ref["e_ss"]) < 1e-10
    assert np.abs(res.C_alpha - ref["C_a"]).max() < 1e-8 and np.abs(res.C_beta - ref["C_b"]).max() < 1e-8
    assert res.energy_scf == scf.energy_total
    assert res.n_virt_alpha == b.n_basis_functions() - scf.n_occ_alpha
    return res


def test_cd_omp2_from_rhf(m):
    """H4 chain, cc-pVDZ, from the drop-in CD-RHF result."""
    mol, b, eng, chol, L, H = _setup(m, [(1, 0.0, 0.0, 1.4 * k) for k in range(4)], "cc-pvdz", 1)
    rcfg = m.CholeskyRHFConfig()
    rcfg.print_level = 0
    scf = m.CholeskyRHF(mol, b, eng, rcfg, chol).compute()
    cfg = m.CholeskyOMP2Config()
    cfg.print_level = 0
    res = _compare(m, mol, b, eng, chol, L, H, scf, cfg)
    assert res.converged and res.iterations > 3
    # constructor 1 decomposes with the same threshold: same loop
    res1 = m.CholeskyOMP2(mol, b, eng, scf, cfg).compute()
    assert res1.iterations == res.iterations and abs(res1.energy_total - res.energy_total) < 1e-10


def test_cd_omp2_from_uhf(m):
    """H3 chain doublet (2 alpha, 1 beta), cc-pVDZ, from the drop-in CD-UHF
    result; the beta orbitals rotate on their own."""
    mol, b, eng, chol, L, H = _setup(m, [(1, 0.0, 0.0, 1.4 * k) for k in range(3)], "cc-pvdz", 2)
    ucfg = m.CholeskyUHFConfig()
    ucfg.print_level = 0
    ucfg.energy_threshold = 1e-11
    ucfg.max_iterations = 300
    uhf = m.CholeskyUHF(mol, b, eng, 2, 1, ucfg)
    uhf.set_cholesky_vectors(chol.get_L_vectors())
    scf = uhf.compute()
    cfg = m.CholeskyOMP2Config()
    cfg.print_level = 0
    cfg.energy_threshold = 1e-9
    cfg.gradient_threshold = 1e-7
    res = _compare(m, mol, b, eng, chol, L, H, scf, cfg)
    assert res.converged


@pytest.mark.parametrize("mult", [1, 2])
def test_cd_omp2_two_particle(m, mult):
    """The loop with the two-particle gradient term (not in the reference):
    H4 from CD-RHF and H3 doublet from CD-UHF, cc-pVDZ, against the oracle's
    loop with the same term."""
    nat = 4 if mult == 1 else 3
    mol, b, eng, chol, L, H = _setup(m, [(1, 0.0, 0.0, 1.4 * k) for k in range(nat)], "cc-pvdz", mult)
    if mult == 1:
        rcfg = m.CholeskyRHFConfig()
        rcfg.print_level = 0
        scf = m.CholeskyRHF(mol, b, eng, rcfg, chol).compute()
    else:
        ucfg = m.CholeskyUHFConfig()
        ucfg.print_level = 0
        ucfg.energy_threshold = 1e-11
        ucfg.max_iterations = 300
        uhf = m.CholeskyUHF(mol, b, eng, 2, 1, ucfg)
        uhf.set_cholesky_vectors(chol.get_L_vectors())
        scf = uhf.compute()
    cfg = m.CholeskyOMP2Config()
    cfg.print_level = 0
    cfg.two_particle_gradient = True
    if mult == 2:
        cfg.energy_threshold = 1e-9
        cfg.gradient_threshold = 1e-7
    res = _compare(m, mol, b, eng, chol, L, H, scf, cfg)
    assert res.converged


def test_device_orbital_step_matches_oracle(ctx):
    """mshqc_hip_omp2_fock / _grad_norm / _rotate (cholesky_omp2.cc:331-425 on
    the device) against numpy: P + C G C^T, the generalised Fock in the MO
    basis from the oracle's J/K, the gradient norm, and the oracle's step
    C exp(-kappa) (omp2_ref.rotation_step, scipy's expm), with and without a
    two-particle term."""
    import scipy.linalg as sl
    import torch

    import oracle

    rng = np.random.default_rng(97)
    n, Q, na, nb = 24, 40, 5, 3
    L = syn.gen_L(Q, n, rng)
    A = rng.standard_normal((n, n))
    H = 0.5 * (A + A.T) - 3.0 * np.eye(n)
    _, U = np.linalg.eigh(H)
    C = [U @ sl.expm(0.1 * (lambda X: X - X.T)(rng.standard_normal((n, n)))) for _ in range(2)]
    no = (na, nb)
    P = [C[s][:, :no[s]] @ C[s][:, :no[s]].T + 0.01 * np.eye(n) for s in range(2)]
    G = []
    for s in range(2):
        X = rng.standard_normal((no[s], no[s]))
        Y = rng.standard_normal((n - no[s], n - no[s]))
        G += [-0.01 * (X @ X.T), 0.01 * (Y @ Y.T)]
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    Co = [dev(C[s][:, :no[s]]) for s in range(2)]
    Cv = [dev(C[s][:, no[s]:]) for s in range(2)]
    Fd = ctx.omp2_fock(dev(L), dev(H), Co[0], Cv[0], Co[1], Cv[1], dev(P[0]), dev(P[1]), *[dev(g) for g in G])
    Pt = []
    for s in range(2):
        Gf = np.zeros((n, n))
        Gf[:no[s], :no[s]], Gf[no[s]:, no[s]:] = G[2 * s], G[2 * s + 1]
        Pt.append(P[s] + C[s] @ Gf @ C[s].T)
    J, K = oracle.jk(L, np.stack(Pt))
    for s in range(2):
        F_ref = C[s].T @ (H + J - K[s]) @ C[s]
        check_close(to_np(Fd[s]), F_ref, 1e-12, f"F spin {s}")
        w2 = 0.05 * rng.standard_normal((n - no[s], no[s])) if s == 1 else None
        w = 2.0 * F_ref[no[s]:, :no[s]] + (0.0 if w2 is None else w2)
        g2 = ctx.omp2_grad_norm(Fd[s], no[s], None if w2 is None else dev(w2))
        assert abs(g2 - np.sum(w * w)) <= 1e-12 * np.sum(w * w)
        Cn = C[s] @ sl.expm(-omp2_ref.rotation_step(w, F_ref, no[s]))
        Pd = dev(P[s])
        ctx.omp2_rotate(Fd[s], Co[s], Cv[s], Pd, None if w2 is None else dev(w2))
        check_close(to_np(Co[s]), Cn[:, :no[s]], 1e-12, f"C_occ spin {s}")
        check_close(to_np(Cv[s]), Cn[:, no[s]:], 1e-12, f"C_vir spin {s}")
        check_close(to_np(Pd), Cn[:, :no[s]] @ Cn[:, :no[s]].T, 1e-12, f"P spin {s}")
