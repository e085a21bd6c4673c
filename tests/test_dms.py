"""The DMS reader (openmm_sdm_plugin_amd/dms.py) against hand-computed values.

A synthetic DMS file is written with sqlite3 (the reference ships no reader tests); its
expected values follow desmonddmsfile75.py's conversions written out by hand.  The
reference's example files, when present, are checked against aggregates computed
straight from their tables with SQL (an independent path from the reader's)."""
import dataclasses
import math
import os
import sqlite3

import numpy as np
import pytest

from openmm_sdm_plugin_amd.dms import ANG, KCAL, DesmondDMSFile, to_sdm_system

REF_EXAMPLES = "/root/reference/example"


def _make_dms(path, box=True, vdw_rule="arithmetic", extra=None, pair_excluded=True):
    con = sqlite3.connect(path)
    c = con.cursor()
    c.execute("CREATE TABLE particle (id INTEGER PRIMARY KEY, name TEXT, anum INTEGER, "
              "resname TEXT, resid INTEGER, chain TEXT, x FLOAT, y FLOAT, z FLOAT, vx FLOAT, "
              "vy FLOAT, vz FLOAT, mass FLOAT, charge FLOAT, nbtype INTEGER)")
    atoms = [  # water (O, H, H), then a 4-atom chain C-C-C-C
        (0, "O", 8, "WAT", 1, "A", 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 15.999, -0.8, 0),
        (1, "H1", 1, "WAT", 1, "A", 0.9572, 0.0, 0.0, 0.0, 1.0, 0.0, 1.008, 0.4, 1),
        (2, "H2", 1, "WAT", 1, "A", -0.24, 0.927, 0.0, 0.0, 0.0, 1.0, 1.008, 0.4, 1),
        (3, "C1", 6, "LIG", 2, "B", 3.0, 0.0, 0.0, 0.0, 0.0, 0.0, 12.0, 0.1, 2),
        (4, "C2", 6, "LIG", 2, "B", 4.5, 0.0, 0.0, 0.0, 0.0, 0.0, 12.0, -0.1, 2),
        (5, "C3", 6, "LIG", 2, "B", 5.0, 1.4, 0.0, 0.0, 0.0, 0.0, 12.0, 0.2, 2),
        (6, "C4", 6, "LIG", 2, "B", 6.4, 1.5, 0.5, 0.0, 0.0, 0.0, 12.0, -0.2, 2),
    ]
    c.executemany("INSERT INTO particle VALUES (?,?,?,?,?,?,?,?,?,?,?,?,?,?,?)", atoms)
    c.execute("CREATE TABLE nonbonded_param (id INTEGER PRIMARY KEY, sigma FLOAT, "
              "epsilon FLOAT)")
    c.executemany("INSERT INTO nonbonded_param VALUES (?,?,?)",
                  [(0, 3.15, 0.152), (1, 0.0, 0.0), (2, 3.4, 0.086)])
    c.execute("CREATE TABLE nonbonded_info (vdw_funct TEXT, vdw_rule TEXT)")
    c.execute("INSERT INTO nonbonded_info VALUES ('vdw_12_6', ?)", (vdw_rule,))
    c.execute("CREATE TABLE bond (p0 INTEGER, p1 INTEGER)")
    c.executemany("INSERT INTO bond VALUES (?,?)", [(0, 1), (0, 2), (3, 4), (4, 5), (5, 6)])
    c.execute("CREATE TABLE stretch_harm_param (id INTEGER PRIMARY KEY, r0 FLOAT, fc FLOAT, "
              "constrained INTEGER)")
    c.executemany("INSERT INTO stretch_harm_param VALUES (?,?,?,?)",
                  [(0, 0.9572, 450.0, 1), (1, 1.5, 300.0, 0)])
    c.execute("CREATE TABLE stretch_harm_term (p0 INTEGER, p1 INTEGER, param INTEGER)")
    c.executemany("INSERT INTO stretch_harm_term VALUES (?,?,?)",
                  [(0, 1, 0), (0, 2, 0), (3, 4, 1), (4, 5, 1), (5, 6, 1)])
    c.execute("CREATE TABLE angle_harm_param (id INTEGER PRIMARY KEY, theta0 FLOAT, fc FLOAT, "
              "constrained INTEGER)")
    c.executemany("INSERT INTO angle_harm_param VALUES (?,?,?,?)", [(0, 109.5, 50.0, 0)])
    c.execute("CREATE TABLE angle_harm_term (p0 INTEGER, p1 INTEGER, p2 INTEGER, "
              "param INTEGER)")
    c.executemany("INSERT INTO angle_harm_term VALUES (?,?,?,?)", [(3, 4, 5, 0)])
    c.execute("CREATE TABLE constraint_hoh_param (id INTEGER PRIMARY KEY, r1 FLOAT, r2 FLOAT, "
              "theta FLOAT)")
    c.execute("INSERT INTO constraint_hoh_param VALUES (0, 0.9572, 0.9572, 104.52)")
    c.execute("CREATE TABLE constraint_hoh_term (p0 INTEGER, p1 INTEGER, p2 INTEGER, "
              "param INTEGER)")
    c.execute("INSERT INTO constraint_hoh_term VALUES (0, 1, 2, 0)")
    c.execute("CREATE TABLE dihedral_trig_param (id INTEGER PRIMARY KEY, phi0 FLOAT, fc0 FLOAT, "
              "fc1 FLOAT, fc2 FLOAT, fc3 FLOAT, fc4 FLOAT, fc5 FLOAT, fc6 FLOAT)")
    c.execute("INSERT INTO dihedral_trig_param VALUES "
              "(0, 180.0, 0.0, 1.5, 0.0, 0.2, 0.0, 0.0, 0.0)")
    c.execute("CREATE TABLE dihedral_trig_term (p0 INTEGER, p1 INTEGER, p2 INTEGER, p3 INTEGER, "
              "param INTEGER)")
    c.execute("INSERT INTO dihedral_trig_term VALUES (3, 4, 5, 6, 0)")
    c.execute("CREATE TABLE exclusion (p0 INTEGER, p1 INTEGER)")
    ex = [(0, 1), (0, 2), (1, 2), (3, 4), (4, 5), (5, 6), (3, 5), (4, 6)]
    if pair_excluded:
        ex.append((3, 6))
    c.executemany("INSERT INTO exclusion VALUES (?,?)", ex)
    c.execute("CREATE TABLE pair_12_6_es_param (id INTEGER PRIMARY KEY, aij FLOAT, bij FLOAT, "
              "qij FLOAT)")
    c.execute("INSERT INTO pair_12_6_es_param VALUES (0, 2.0e5, 300.0, -0.02)")
    c.execute("CREATE TABLE pair_12_6_es_term (p0 INTEGER, p1 INTEGER, param INTEGER)")
    c.execute("INSERT INTO pair_12_6_es_term VALUES (3, 6, 0)")
    c.execute("CREATE TABLE posre_harmflatbottom_param (id INTEGER PRIMARY KEY, fc FLOAT, "
              "tol FLOAT)")
    c.execute("INSERT INTO posre_harmflatbottom_param VALUES (0, 2.0, 1.5)")
    c.execute("CREATE TABLE posre_harmflatbottom_term (p0 INTEGER, x0 FLOAT, y0 FLOAT, "
              "z0 FLOAT, param INTEGER)")
    c.execute("INSERT INTO posre_harmflatbottom_term VALUES (4, 4.4, 0.1, -0.2, 0)")
    if box:
        c.execute("CREATE TABLE global_cell (id INTEGER PRIMARY KEY, x FLOAT, y FLOAT, "
                  "z FLOAT)")
        c.executemany("INSERT INTO global_cell VALUES (?,?,?,?)",
                      [(0, 30.0, 0, 0), (1, 0, 31.0, 0), (2, 0, 0, 32.0)])
    for stmt in extra or ():
        c.execute(stmt)
    con.commit()
    con.close()
    return path


def test_synthetic_dms_terms(tmp_path):
    d = DesmondDMSFile(_make_dms(str(tmp_path / "a.dms")))
    assert d.topology.getNumAtoms() == 7
    assert np.allclose(d.topology.getUnitCellDimensions(), [3.0, 3.1, 3.2])
    assert np.allclose(d.getPositions()[1], [0.09572, 0.0, 0.0])
    assert np.allclose(d.getVelocities()[2], [0.0, 0.0, 0.1])  # A/ps -> nm/ps
    ff = d.createSystem(nonbondedMethod="PME", nonbondedCutoff=0.9)
    # unconstrained bonds: r0 in nm, k doubled and converted: 2*300 kcal/mol/A^2
    assert ff.bonds.tolist() == [[3, 4], [4, 5], [5, 6]]
    assert np.allclose(ff.bond_r0, 0.15) and np.allclose(ff.bond_k, 600 * 4.184 * 100)
    assert np.allclose(ff.angle_t0, 109.5 * math.pi / 180)
    assert np.allclose(ff.angle_k, 100 * 4.184)
    # two O-H constraints from the constrained stretches, the H-H one from constraint_hoh
    cons = {tuple(sorted(p)): r for p, r in zip(ff.constraints.tolist(), ff.constraint_r)}
    assert cons[(0, 1)] == pytest.approx(0.09572)
    hh = math.sqrt(2 * 0.09572 ** 2 * (1 - math.cos(104.52 * math.pi / 180)))
    assert cons[(1, 2)] == pytest.approx(hh, rel=1e-12)
    # one periodic term per nonzero fc_n: n = 1 and n = 3, phase 180 degrees
    assert ff.torsion_n.tolist() == [1, 3]
    assert np.allclose(ff.torsion_phase, math.pi)
    assert np.allclose(ff.torsion_k, [1.5 * 4.184, 0.2 * 4.184])
    # 1-4 exception: sigma = (a/b)^(1/6), epsilon = b^2/(4a) from a, b in kJ/mol nm^12, nm^6
    a, b = 2.0e5 * KCAL * ANG ** 12, 300.0 * KCAL * ANG ** 6
    assert ff.exception_pairs.tolist() == [[3, 6]]
    assert ff.exception_sigma[0] == pytest.approx((a / b) ** (1 / 6), rel=1e-12)
    assert ff.exception_epsilon[0] == pytest.approx(b * b / (4 * a), rel=1e-12)
    assert ff.exception_qq[0] == -0.02
    # flat-bottom positional restraint: hk = fc/2 in kJ/mol/nm^2, tol in nm
    assert ff.posre_fb_atoms.tolist() == [4]
    assert ff.posre_fb_hk[0] == pytest.approx(0.5 * 2.0 * 4.184 * 100)
    assert ff.posre_fb_tol[0] == pytest.approx(0.15)
    assert np.allclose(ff.posre_fb_x0[0], [0.44, 0.01, -0.02])
    # nonbonded: sigma A -> nm, epsilon kcal -> kJ; arithmetic rule
    assert np.allclose(ff.sigma[[0, 3]], [0.315, 0.34]) and not ff.opls
    assert ff.epsilon[0] == pytest.approx(0.152 * 4.184)
    assert ff.exclusions.shape == (9, 2)
    assert ff.nonbonded_method == "PME" and ff.cutoff == 0.9


VSITE_SQL = [
    # a charge site on the water bisector (lc3), a site between C1 and C2 (lc2), and an
    # out-of-plane site of C2-C3-C4 (out3); each a massless particle appended to the file
    "INSERT INTO particle VALUES (7, 'M', 0, 'WAT', 1, 'A', 0.1, 0.1, 0.0, 0, 0, 0, 0.0, "
    "-0.1, 1)",
    "INSERT INTO particle VALUES (8, 'X', 0, 'LIG', 2, 'B', 3.7, 0, 0, 0, 0, 0, 0.0, 0.0, 1)",
    "INSERT INTO particle VALUES (9, 'Y', 0, 'LIG', 2, 'B', 5.0, 0, 1, 0, 0, 0, 0.0, 0.0, 1)",
    "CREATE TABLE virtual_lc3_param (id INTEGER PRIMARY KEY, c1 FLOAT, c2 FLOAT)",
    "INSERT INTO virtual_lc3_param VALUES (0, 0.12, 0.13)",
    "CREATE TABLE virtual_lc3_term (p0 INTEGER, p1 INTEGER, p2 INTEGER, p3 INTEGER, "
    "param INTEGER)",
    "INSERT INTO virtual_lc3_term VALUES (7, 0, 1, 2, 0)",
    "CREATE TABLE virtual_lc2_param (id INTEGER PRIMARY KEY, c1 FLOAT)",
    "INSERT INTO virtual_lc2_param VALUES (0, 0.4)",
    "CREATE TABLE virtual_lc2_term (p0 INTEGER, p1 INTEGER, p2 INTEGER, param INTEGER)",
    "INSERT INTO virtual_lc2_term VALUES (8, 3, 4, 0)",
    "CREATE TABLE virtual_out3_param (id INTEGER PRIMARY KEY, c1 FLOAT, c2 FLOAT, c3 FLOAT)",
    "INSERT INTO virtual_out3_param VALUES (0, 0.3, 0.2, 0.5)",
    "CREATE TABLE virtual_out3_term (p0 INTEGER, p1 INTEGER, p2 INTEGER, p3 INTEGER, "
    "param INTEGER)",
    "INSERT INTO virtual_out3_term VALUES (9, 4, 5, 6, 0)",
]


def test_virtual_sites(tmp_path):
    """The reference's reader makes TwoParticleAverageSite(p1, p2, 1-c1, c1),
    ThreeParticleAverageSite(p1, p2, p3, 1-c1-c2, c1, c2) and OutOfPlaneSite(p1, p2, p3,
    c1, c2, c3) (desmonddmsfile75.py:877, :892, :907), lc2 tables first."""
    path = _make_dms(str(tmp_path / "v.dms"), extra=VSITE_SQL)
    ff = DesmondDMSFile(path).createSystem(nonbondedMethod="CutoffPeriodic")
    assert ff.n == 10 and ff.mass[7] == 0.0
    assert ff.vsite_kind.tolist() == [0, 1, 2]
    assert ff.vsite_atoms[0, :3].tolist() == [8, 3, 4]
    assert ff.vsite_atoms[1].tolist() == [7, 0, 1, 2]
    assert ff.vsite_atoms[2].tolist() == [9, 4, 5, 6]
    assert np.allclose(ff.vsite_weights[0, :2], [0.6, 0.4])
    assert np.allclose(ff.vsite_weights[1], [0.75, 0.12, 0.13])
    assert np.allclose(ff.vsite_weights[2], [0.3, 0.2, 0.5])
    # offsets of a second file
    ff2 = DesmondDMSFile([path, path]).createSystem(nonbondedMethod="CutoffPeriodic")
    assert ff2.vsite_atoms[ff2.vsite_kind == 1].tolist() == [[7, 0, 1, 2], [17, 10, 11, 12]]
    s = to_sdm_system(ff, DesmondDMSFile(path).getPositions())
    assert s.virtual_sites is not None and s.virtual_sites.kind.tolist() == [0, 1, 2]
    with pytest.raises(NotImplementedError):
        DesmondDMSFile(_make_dms(str(tmp_path / "f3.dms"), extra=[
            "CREATE TABLE virtual_fdat3_term (p0 INTEGER)"])).createSystem()
    with pytest.raises(NotImplementedError):
        DesmondDMSFile(_make_dms(str(tmp_path / "o.dms"), extra=[
            "CREATE TABLE virtual_lc4_term (p0 INTEGER)",
            "INSERT INTO virtual_lc4_term VALUES (0)"])).createSystem()


def test_opls_torsions_and_rule(tmp_path):
    """With OPLS=True the reference writes torsions as f cos(n phi - phi0) (a
    CustomTorsionForce, desmonddmsfile75.py:652-656, :679-680): here two periodic terms per
    fc, whose energy is that sum exactly; the engine gets the geometric sigma rule."""
    from oracle import terms as OT
    path = _make_dms(str(tmp_path / "o.dms"), vdw_rule="geometric")
    d = DesmondDMSFile(path)
    ff = d.createSystem(nonbondedMethod="CutoffPeriodic", OPLS=True)
    assert ff.opls and ff.torsion_n.tolist() == [1, 1, 3, 3]
    x = d.getPositions()
    only_torsions = dataclasses.replace(
        ff, impropers=np.zeros((0, 4), np.int32), improper_t0=np.zeros(0),
        improper_k=np.zeros(0), posre_fb_atoms=np.zeros(0, np.int32),
        posre_fb_x0=np.zeros((0, 3)), posre_fb_hk=np.zeros(0), posre_fb_tol=np.zeros(0),
        posre_h_atoms=np.zeros(0, np.int32), posre_h_x0=np.zeros((0, 3)),
        posre_h_hk=np.zeros((0, 3)), exception_pairs=np.zeros((0, 2), np.int32),
        exception_qq=np.zeros(0), exception_sigma=np.zeros(0), exception_epsilon=np.zeros(0))
    e1, _ = OT.energy(x, only_torsions, ff.box)
    phi = OT.dihedral(*[x[i] for i in (3, 4, 5, 6)])
    want = 1.5 * KCAL * math.cos(phi - math.pi) + 0.2 * KCAL * math.cos(3 * phi - math.pi)
    assert e1 == pytest.approx(want, rel=1e-12, abs=1e-12)
    s = to_sdm_system(ff, x)
    assert s.lj_rule == 1
    assert to_sdm_system(d.createSystem(nonbondedMethod="CutoffPeriodic"), x).lj_rule == 0


def test_multi_file_offsets(tmp_path):
    f1 = _make_dms(str(tmp_path / "a.dms"))
    f2 = _make_dms(str(tmp_path / "b.dms"), box=False)
    ff = DesmondDMSFile([f1, f2]).createSystem(nonbondedMethod="CutoffPeriodic")
    assert ff.n == 14
    assert ff.bonds.tolist()[3:] == [[10, 11], [11, 12], [12, 13]]
    assert ff.exception_pairs.tolist() == [[3, 6], [10, 13]]
    assert ff.exclusions.shape == (18, 2) and ff.exclusions.max() == 13


def test_geometric_rule_and_rejections(tmp_path):
    # a "geometric" file is read with Lorentz-Berthelot unless OPLS=True is asked for, as
    # the reference's createSystem builds its forces from its OPLS argument alone
    path = _make_dms(str(tmp_path / "g.dms"), vdw_rule="geometric")
    assert not DesmondDMSFile(path).createSystem().opls
    assert DesmondDMSFile(path).createSystem(OPLS=True).opls
    with pytest.raises(NotImplementedError):
        DesmondDMSFile(_make_dms(str(tmp_path / "u.dms"), vdw_rule="harmonic")).createSystem()
    with pytest.raises(NotImplementedError):
        DesmondDMSFile(_make_dms(str(tmp_path / "p.dms"), pair_excluded=False)).createSystem()
    with pytest.raises(NotImplementedError):
        DesmondDMSFile(_make_dms(str(tmp_path / "f.dms"), extra=[
            "CREATE TABLE stretch_fbhw_term (p0 INTEGER)"])).createSystem()
    with pytest.raises(ValueError):
        DesmondDMSFile(_make_dms(str(tmp_path / "n.dms"), box=False)).createSystem(
            nonbondedMethod="PME")
    with pytest.raises(IOError):
        DesmondDMSFile(str(tmp_path / "missing.dms"))


def test_write_back_positions(tmp_path):
    path = _make_dms(str(tmp_path / "w.dms"))
    d = DesmondDMSFile(path)
    x = d.getPositions() + 0.05
    d.setPositions(x)
    d.close()
    assert np.allclose(DesmondDMSFile(path).getPositions(), x)


def _sql(path, q):
    con = sqlite3.connect("file:%s?mode=ro" % path, uri=True)
    try:
        return con.execute(q).fetchone()
    finally:
        con.close()


@pytest.mark.parametrize("name", ["temoa-g1-g4.dms", "oa-g6-g3-align-restr_0_displaced.dms"])
def test_reference_examples(name):
    path = os.path.join(REF_EXAMPLES, name)
    if not os.path.exists(path):
        pytest.skip("the reference's example files are not present here")
    d = DesmondDMSFile(path)
    ff = d.createSystem(nonbondedMethod="NoCutoff")
    n, qsum, msum = _sql(path, "SELECT COUNT(*), SUM(charge), SUM(mass) FROM particle")
    assert ff.n == n and d.topology.getNumAtoms() == n
    assert ff.charge.sum() == pytest.approx(qsum, abs=1e-9)
    assert ff.mass.sum() == pytest.approx(msum, rel=1e-12)
    nb, kb = _sql(path, "SELECT COUNT(*), SUM(fc) FROM stretch_harm_term JOIN "
                        "stretch_harm_param ON stretch_harm_term.param = stretch_harm_param.id "
                        "WHERE constrained = 0")
    assert ff.bonds.shape[0] == nb
    assert ff.bond_k.sum() == pytest.approx(2 * (kb or 0.0) * 418.4, rel=1e-12)
    na, = _sql(path, "SELECT COUNT(*) FROM angle_harm_term JOIN angle_harm_param "
                     "ON angle_harm_term.param = angle_harm_param.id WHERE constrained = 0")
    assert ff.angles.shape[0] == na
    per_order = " + ".join(
        "(SELECT COUNT(*) FROM dihedral_trig_term JOIN dihedral_trig_param ON "
        "dihedral_trig_term.param = dihedral_trig_param.id WHERE fc%d != 0)" % k
        for k in range(7))
    nt, = _sql(path, "SELECT " + per_order)
    assert ff.torsions.shape[0] == nt
    ne, = _sql(path, "SELECT COUNT(*) FROM exclusion")
    npair, = _sql(path, "SELECT COUNT(*) FROM pair_12_6_es_term")
    assert ff.exclusions.shape[0] == ne and ff.exception_pairs.shape[0] == npair
    nfb, = _sql(path, "SELECT COUNT(*) FROM posre_harmflatbottom_term")
    assert ff.posre_fb_atoms.shape[0] == nfb
    if name.startswith("temoa"):
        # 6,734 rigid waters (3 constraints each) and the constrained X-H stretches
        nw, = _sql(path, "SELECT COUNT(*) FROM constraint_hoh_term")
        nc, = _sql(path, "SELECT COUNT(*) FROM stretch_harm_term JOIN stretch_harm_param ON "
                         "stretch_harm_term.param = stretch_harm_param.id WHERE constrained = 1")
        assert nw == 6734 and ff.constraints.shape[0] == nc + nw
        assert np.allclose(ff.box, [5.64589811, 6.32925716, 5.79653227])
        assert not ff.opls
    d.close()


def test_fixture_matches_reader():
    """tests/golden/temoa_g1_g4.npz (what the GPU tests run) is the reader's output."""
    path = os.path.join(REF_EXAMPLES, "temoa-g1-g4.dms")
    if not os.path.exists(path):
        pytest.skip("the reference's example files are not present here")
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "temoa_g1_g4.npz"))
    d = DesmondDMSFile(path)
    ff = d.createSystem(nonbondedMethod="PME", nonbondedCutoff=1.0)
    assert np.array_equal(z["positions"], d.getPositions())
    assert np.array_equal(z["velocities"], d.getVelocities())
    for k in ("mass", "charge", "sigma", "epsilon", "exclusions", "constraints", "constraint_r",
              "torsions", "torsion_k", "exception_pairs", "exception_sigma", "posre_fb_x0"):
        assert np.array_equal(z[k], getattr(ff, k)), k
    d.close()
