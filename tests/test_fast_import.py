"""FAST input/output import (winds.fast: core/input_import.m,
core/output_import.m, core/input_mod.m).

No FAST files ship with the reference; tests/fastfiles.py writes file sets
on the lines the reference reads.  Checks:
- every field input_import.m reads comes from its line (turbine, blade
  table, airfoil tables and their header values, platform types 0-3);
- output_import.m: the output file chosen by YCMode (and the _Sfunc.out
  fallback), the time column rebuilt from DT (an inconsistent DT is an
  error), channels interpolated at (t0:1/f:t1), out-of-range times clamped,
  the zero-frequency quirk;
- input_mod.m: trailing-node radii, twist sign, half chord at the tip,
  cosine respacing (ends at hub and tip, nearest airfoil), its cubic
  interpolation against scipy's PCHIP, the azimuth unwrapped and a spike
  bridged;
- a FAST file set of the NREL 5-MW rotor turning rigidly gives the same
  hub motion, azimuth and blade kinematics as the modeldata/NRELrotor.m case,
  and its tables march in the oracle.
"""
import math

import numpy as np
import pytest

import fastfiles as FF
from winds import airfoils as af
from winds import cases, fast


@pytest.fixture
def nrel_set(tmp_path):
    return FF.nrel_file_set(tmp_path / "nrel")


def test_read_fst_fields(nrel_set):
    airfoils, blade, turbine, platform, wind = fast.read_fst(nrel_set)
    assert (blade.TipRad, blade.HubRad) == (63.0, 1.5)
    assert turbine.NumBl == 3 and isinstance(turbine.NumBl, int)
    assert (turbine.OverHang, turbine.TowerHt, turbine.Twr2Shft, turbine.ShftTilt) == \
        (5.0, 90.0, 0.0, 5.0)
    assert np.array_equal(turbine.PreCone, [-2.5, -2.5, -2.5])
    assert wind.filename.endswith("Wind/steady.wnd")
    assert platform.Type == "onshore" and platform.PtfmRef == 0.0
    assert np.allclose(blade.RNodes, FF.NREL_NODES) and np.allclose(blade.Chord, FF.NREL_CHORD)
    assert np.allclose(blade.AeroTwst, FF.NREL_TWIST) and np.allclose(blade.DRNodes, FF.NREL_DR)
    assert np.array_equal(blade.NFoil, FF.NREL_FOIL)
    assert airfoils.Names == list(af.NREL5MW_NAMES)
    for got, ref in zip(airfoils.profiles, af.nrel5mw_polars()):
        assert np.allclose(got.AoA, ref.AoA, atol=5e-5) and np.allclose(got.Cl, ref.Cl, atol=5e-7)
        assert np.allclose(got.Cd, ref.Cd, atol=5e-7) and np.allclose(got.Cm, ref.Cm, atol=5e-7)
        assert got.header["StallAoA"] == 15.0 and got.header["Cdmin"] == 0.006


def test_platform_files(tmp_path):
    d = tmp_path / "p"
    d.mkdir()
    FF.write_airfoil(d / "a.dat", [-10.0, 0.0, 10.0], [-1.0, 0.0, 1.0], [0.01] * 3, [0.0] * 3)
    FF.write_aerodyn(d, ["a.dat"], nodes=(10.0, 20.0), twist=(1.0, 0.0), dr=(10.0, 10.0),
                     chord=(2.0, 1.0), nfoil=(1, 1))
    FF.write_platform(d)
    fst = FF.write_fst(d, ptfm_model=2)
    _, _, _, pl, _ = fast.read_fst(fst)
    assert pl.Type == "fixedoffshore"
    assert (pl.TwrDraft, pl.PtfmCM, pl.PtfmRef, pl.PtfmDraft, pl.PtfmDiam) == \
        (-10.0, -89.9, 0.5, 20.0, 6.0)
    fst = FF.write_fst(d, ptfm_model=3)
    _, _, _, pl, _ = fast.read_fst(fst)
    assert pl.Type == "Spar"
    assert (pl.PtfmDraft, pl.PtfmDiam) == (120.0, 9.4)
    fst = FF.write_fst(d, ptfm_model=1)
    _, _, _, pl, _ = fast.read_fst(fst)
    assert pl.Type == "onshore" and pl.filename == "No platform model used."


def test_read_out_interpolates_and_clamps(nrel_set, capsys):
    fo = fast.read_out(nrel_set, (-1.0, 2.0, 8.0))
    assert "reset to 0 seconds" in capsys.readouterr().out
    assert np.allclose(fo.Time, np.arange(17) / 8.0, rtol=0, atol=1e-15)
    omega = 7 * 11.4 / 63
    # rows every 0.05 s; 1/8 s falls between rows: linear interpolation
    az = np.mod(omega * 180 / math.pi * np.arange(201) * 0.05, 360)
    ref = np.interp(fo.Time, np.arange(201) * 0.05, az)
    assert np.allclose(fo.Azimuth, ref, rtol=1e-14, atol=1e-12)
    assert np.all(fo.WindVxi == 11.4) and fo.BldPitch3.shape == fo.Time.shape
    fo = fast.read_out(nrel_set, (9.0, 12.0, 10.0))
    assert "reset to 10 seconds" in capsys.readouterr().out
    assert fo.Time[-1] == pytest.approx(10.0) and fo.Time.size == 11


def test_read_out_zero_frequency_quirk(tmp_path):
    """output_import.m:53 takes the mean step of the first row (here Time,
    Azimuth, RotSpeed at t = 0: (0, 0, 12.1), mean step 6.05), not of the
    time column: the times are 0 and 6.05 s."""
    d = tmp_path / "q"
    fst = FF.nrel_file_set(d)
    t = np.arange(201) * 0.05
    FF.write_out(d / "Test.out", {"Time": t, "Azimuth": 10 * t, "RotSpeed": 12.1 + t})
    fo = fast.read_out(fst, (0.0, 10.0, 0.0))
    assert np.allclose(fo.Time, [0.0, 6.05]) and np.allclose(fo.RotSpeed, [12.1, 18.15])
    # the reference's own example row (every channel 0 but the wind and the
    # rotor speed, equal first and last values): a zero step, no times
    FF.write_out(d / "Test.out", {"Time": t, "RotSpeed": 12.1 + t, "NacYaw": 0 * t})
    assert fast.read_out(fst, (0.0, 10.0, 0.0)).Time.size == 0


def test_output_file_choice_and_dt(tmp_path):
    fst = FF.nrel_file_set(tmp_path / "s", ycmode=2)
    fo = fast.read_out(fst, (0.0, 1.0, 10.0))
    assert fo.filename.endswith("Test_SFunc.out")
    # YCMode 0 but only a _Sfunc.out present: the fallback name
    d = tmp_path / "f"
    fst = FF.nrel_file_set(d)
    (d / "Test.out").rename(d / "Test_Sfunc.out")
    fo = fast.read_out(fst, (0.0, 1.0, 10.0))
    assert fo.filename.endswith("Test_Sfunc.out")
    # a DT that does not match the rows is an error when the time is rebuilt
    d = tmp_path / "dt"
    fst = FF.nrel_file_set(d)
    text = fst.read_text().splitlines()
    text[10] = "0.0125      DT"
    fst.write_text("\n".join(text) + "\n")
    with pytest.raises(ValueError, match="steps of DT"):
        fast.read_out(fst, (0.0, 1.0, 10.0))


def test_input_mod_blade_nodes():
    from types import SimpleNamespace as NS
    bl = NS(TipRad=10.0, HubRad=1.0, RNodes=np.array([2.0, 4.0, 7.0]),
            AeroTwst=np.array([5.0, 3.0, 1.0]), Chord=np.array([1.0, 0.8, 0.5]),
            NFoil=np.array([1.0, 2.0, 2.0]), DRNodes=np.zeros(3))
    fo = NS(Azimuth=np.arange(5.0) * 10, BldPitch1=np.zeros(5), BldPitch2=np.ones(5))
    fast.input_mod(bl, fo, NS(NumBl=2))
    # extended nodes 1, 2, 4, 7, 10: trailing radii 1, (2+4)/2, (4+7)/2, 10
    assert np.allclose(bl.RTrail, [1.0, 3.0, 5.5, 10.0])
    # DRNodes(2:end-1) of the five extended nodes: two widths for three
    # stations, as input_mod.m:42 leaves them
    assert np.array_equal(bl.RNodes, [2.0, 4.0, 7.0]) and np.array_equal(bl.DRNodes, [2.0, 3.0])
    assert np.array_equal(bl.AeroTwst, [-5.0, -3.0, -1.0])
    assert bl.NFoil.dtype == np.int32 and np.array_equal(bl.NFoil, [1, 2, 2])
    ext_r = [1.0, 2.0, 4.0, 7.0, 10.0]
    assert np.allclose(bl.ChordTrail, fast.pchip(ext_r, [1.0, 1.0, 0.8, 0.5, 0.25], bl.RTrail))
    assert bl.ChordTrail[-1] == pytest.approx(0.25) and bl.AeroTwstTrail[0] == pytest.approx(-5.0)
    assert fo.BldPitch.shape == (5, 2) and np.all(fo.BldPitch[:, 1] == 1)


def test_input_mod_cosine_respacing():
    from types import SimpleNamespace as NS
    bl = NS(TipRad=63.0, HubRad=1.5, RNodes=np.array(FF.NREL_NODES),
            AeroTwst=np.array(FF.NREL_TWIST), Chord=np.array(FF.NREL_CHORD),
            NFoil=np.array(FF.NREL_FOIL, dtype=float), DRNodes=np.array(FF.NREL_DR))
    fo = NS(Azimuth=np.zeros(3), BldPitch1=np.zeros(3), BldPitch2=np.zeros(3),
            BldPitch3=np.zeros(3))
    fast.input_mod(bl, fo, NS(NumBl=3), n=25)
    assert bl.RNodes.size == 23 and bl.RTrail.size == 24
    assert bl.RTrail[0] == 1.5 and bl.RTrail[-1] == 63.0
    assert np.all(np.diff(bl.RNodes) > 0)
    # cosine clustering: spacing shrinks towards the tip
    assert np.diff(bl.RNodes)[-1] < 0.2 * np.diff(bl.RNodes)[0]
    assert set(np.unique(bl.NFoil)) <= set(range(1, 9))


def test_pchip_against_scipy():
    from scipy.interpolate import PchipInterpolator
    rng = np.random.default_rng(4)
    x = np.sort(rng.uniform(0, 10, 12))
    y = np.column_stack([np.cumsum(rng.uniform(0, 1, 12)), rng.normal(0, 1, 12)])
    xi = np.linspace(-0.5, 10.5, 301)
    got = fast.pchip(x, y, xi)
    ref = PchipInterpolator(x, y, extrapolate=True)(xi)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)
    lin = fast.pchip(x, 2 * x + 1, xi)                 # linear data is reproduced
    assert np.allclose(lin, 2 * xi + 1, rtol=1e-13)


def test_azimuth_unwrap_and_spike():
    from types import SimpleNamespace as NS
    # 10-degree steps (input_mod.m:49 takes any step above 25 degrees for a
    # jump), wrapping at 360, with a spike at index 5
    az = np.array([340.0, 350.0, 0.0, 10.0, 20.0, 195.0, 40.0, 50.0])
    fo = NS(Azimuth=az, BldPitch1=np.zeros(8))
    bl = NS(TipRad=2.0, HubRad=0.0, RNodes=np.array([0.5, 1.0, 1.5]), AeroTwst=np.zeros(3),
            Chord=np.ones(3), NFoil=np.ones(3), DRNodes=np.zeros(3))
    fast.input_mod(bl, fo, NS(NumBl=1))
    # the 360 -> 0 wrap is unwrapped; the spike at 250 and the point before
    # it are bridged by the cubic through the good points (linear here)
    assert np.allclose(fo.Azimuth, 340.0 + 10.0 * np.arange(8), rtol=0, atol=1e-12)


def test_fast_case_matches_the_nrel_rotor_case(nrel_set):
    fc = fast.fast_case(nrel_set, (0.0, 2.0, 10.0))
    nc = cases.nrel_rotor(t=(0.0, 2.0, 10.0))
    assert np.allclose(fc.fastout.Azimuth, nc.fastout.Azimuth, rtol=0, atol=1e-9)
    tf = cases.tables(fc)
    tn = cases.tables(nc)
    assert tf.n_times == tn.n_times == 21 and tf.nb == 3 and tf.nst == 18
    assert np.allclose(tf.hub, tn.hub, rtol=0, atol=1e-12)
    assert np.allclose(tf.wind, tn.wind) and np.allclose(tf.rot_speed, tn.rot_speed)
    # with the FAST blade on the NRELrotor case, every table agrees: the
    # imported turbine, platform, output channels and wind drive the same
    # kinematics, velocities and BEM start
    nc.blade = fc.blade
    tn = cases.tables(nc)
    for k in ("quarter", "trail", "bound", "bxn", "byn", "bzn"):
        assert np.allclose(getattr(tf, k), getattr(tn, k), rtol=0, atol=1e-9), k
    assert np.allclose(tf.vblade, tn.vblade, rtol=0, atol=1e-9)
    assert np.allclose(tf.cl_init, tn.cl_init, rtol=1e-9)
    assert sorted(set(tf.foil.tolist())) == list(range(1, 9))


def test_fast_case_marches_in_the_oracle(nrel_set):
    from oracle.march import OracleMarch
    tb = cases.tables(fast.fast_case(nrel_set, (0.0, 1.0, 10.0)))
    om = OracleMarch(tb, 8, d=1e-4, co=math.inf, semantics="mex", integ="pcc")
    for _ in range(4):
        om.step()
    h = np.array(om.history)
    assert np.all(np.isfinite(h)) and np.all(h[:, 1] > 0) and np.all(h[:, 3] > 0)
