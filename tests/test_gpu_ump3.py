"""CD-UMP3 on the GPU against the oracle's restatement of
cholesky_ump3.cc:160-637 (itself pinned against the reference's compiled
CholeskyUMP3), through the C ABI and through the drop-in classes."""
import os

import numpy as np
import pytest

import oracle
from conftest import ROOT, to_dev
from mshqc_amd import synthetic as syn
from oracle import mp3_ref

pytestmark = pytest.mark.gpu
BD = os.path.join(ROOT, "mshqc_amd", "data", "basis")


def blocks(L, C, o):
    Co, Cv = C[:, :o].copy(), C[:, o:].copy()
    return oracle.transform(L, Co, Cv), oracle.transform(L, Co, Co), oracle.transform(L, Cv, Cv)


def check_components(got, want, rel=1e-12):
    for g, w, name in zip(got, want, ("aa", "bb", "ab")):
        assert abs(g - w) <= rel * abs(w) + 1e-16, f"{name}: {g!r} vs {w!r}"


# (Q, n, oa, ob, zblock): alpha and beta occupations differ, so va != vb;
# ob = 0 leaves only the alpha component; a small zblock builds (ac|bd) in
# blocks of two a.
@pytest.mark.parametrize("Q,n,oa,ob,zblock,seed", [(23, 20, 4, 3, None, 1), (31, 33, 6, 4, None, 2),
                                                   (40, 46, 7, 5, 2, 3), (17, 12, 2, 0, None, 4),
                                                   (19, 15, 1, 1, None, 5)])
def test_ump3_matches_oracle(ctx, Q, n, oa, ob, zblock, seed):
    rng = np.random.default_rng(seed)
    L = syn.gen_L(Q, n, rng)
    Ca = syn.gen_Cocc(n, n, rng)
    Cb = syn.gen_Cocc(n, n, rng)
    va, vb = n - oa, n - ob
    eoa, eva = syn.gen_eps(oa, va, rng)
    eob, evb = syn.gen_eps(ob, vb, rng) if ob else (np.zeros(0), np.sort(rng.uniform(0.1, 2.0, vb)))
    Ba, Looa, Lvva = blocks(L, Ca, oa)
    Bb, Loob, Lvvb = blocks(L, Cb, ob)
    want = mp3_ref.ump3(Ba, Bb, Looa, Loob, Lvva, Lvvb, eoa, eva, eob, evb)
    print(f"ump3 Q={Q} oa={oa} va={va} ob={ob} vb={vb} zblock={zblock}", flush=True)
    if zblock:
        ctx.set_zblock(zblock * va * vb * vb)
    try:
        got = ctx.ump3(*to_dev(Ba, Bb, Looa, Loob, Lvva, Lvvb, eoa, eva, eob, evb))
    finally:
        ctx.set_zblock(0)
    if ob == 0:
        assert got[1] == 0.0 and got[2] == 0.0
    check_components(got, want)


def test_ump3_identical_spins(ctx):
    """Equal alpha and beta orbitals: the two same-spin components follow the
    reference's different (alpha / beta) formulas, so only the oracle fixes
    them; the opposite-spin term is the symmetric one."""
    rng = np.random.default_rng(11)
    Q, n, o = 21, 16, 3
    L = syn.gen_L(Q, n, rng)
    C = syn.gen_Cocc(n, n, rng)
    eo, ev = syn.gen_eps(o, n - o, rng)
    B, Loo, Lvv = blocks(L, C, o)
    want = mp3_ref.ump3(B, B, Loo, Loo, Lvv, Lvv, eo, ev, eo, ev)
    print(f"ump3 identical spins Q={Q} o={o} v={n - o}", flush=True)
    got = ctx.ump3(*to_dev(B, B, Loo, Loo, Lvv, Lvv, eo, ev, eo, ev))
    check_components(got, want)


@pytest.fixture(scope="module")
def m(ctx):
    from mshqc_amd import native_module

    return native_module()


def test_cd_ump3_dropin(m):
    """H3 doublet, cc-pVDZ: CD-UHF -> CD-UMP2 -> CD-UMP3 through the classes
    (from the UMP2 solver, and decomposing its own vectors from the UHF
    result), against the oracle on the same vectors and orbitals."""
    mol = m.Molecule(0, 2)
    for k in range(3):
        mol.add_atom(1, 0.0, 0.1 * k, 1.5 * k)
    b = m.BasisSet("cc-pvdz", mol, BD)
    eng = m.IntegralEngine(mol, b)
    chol = m.CholeskyERI(1e-6)
    chol.decompose(eng.compute_eri())
    n = b.n_basis_functions()
    cfg = m.CholeskyUHFConfig()
    cfg.print_level = 0
    solver = m.CholeskyUHF(mol, b, eng, 2, 1, cfg)
    solver.set_cholesky_vectors(chol.get_L_vectors())
    uhf = solver.compute()
    mc = m.CholeskyUMP2Config()
    mc.print_level = 0
    ump2 = m.CholeskyUMP2(uhf, b, eng, mc, chol)
    r2 = ump2.compute()
    c3 = m.CholeskyUMP3Config()
    c3.print_level = 0
    r3 = m.CholeskyUMP3(ump2, c3).compute()
    na, nb = uhf.n_occ_alpha, uhf.n_occ_beta
    Lu = np.array(ump2.get_cholesky().get_L_vectors()).reshape(-1, n, n)
    ea, eb = uhf.orbital_energies_alpha, uhf.orbital_energies_beta
    Ba, Looa, Lvva = blocks(Lu, np.asarray(uhf.C_alpha), na)
    Bb, Loob, Lvvb = blocks(Lu, np.asarray(uhf.C_beta), nb)
    want = mp3_ref.ump3(Ba, Bb, Looa, Loob, Lvva, Lvvb, ea[:na], ea[na:], eb[:nb], eb[nb:])
    print(f"H3 CD-UMP3: {(r3.e_mp3_ss_aa, r3.e_mp3_ss_bb, r3.e_mp3_os)!r} oracle {want!r}", flush=True)
    check_components((r3.e_mp3_ss_aa, r3.e_mp3_ss_bb, r3.e_mp3_os), want)
    assert abs(r3.e_mp3_total - sum(want)) <= 1e-12 * abs(sum(want))
    assert r3.n_cholesky_vectors == Lu.shape[0] and r2.e_corr_total != 0.0
    # own decomposition from the UHF result, at the same threshold
    r3b = m.CholeskyUMP3(uhf, b, eng, c3).compute()
    assert abs(r3b.e_mp3_total - r3.e_mp3_total) <= 1e-12 * abs(r3.e_mp3_total)
