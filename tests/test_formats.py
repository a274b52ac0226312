"""The on-disk formats (thrid_amd/csrc/pgk_io.hip) against the reference's own
files: its state file read and written back byte for byte, and its out_ene.dat
energy row re-formatted from the parsed values."""
import gzip
import os

import numpy as np
import pytest

from thrid_amd import formats

GOLD = os.path.join(os.path.dirname(__file__), "golden", "formats")


@pytest.fixture(scope="module")
def sta_path(tmp_path_factory):
    p = tmp_path_factory.mktemp("sta") / "sta_peachgk-2.0.dat"
    with gzip.open(os.path.join(GOLD, "sta_peachgk-2.0.dat.gz"), "rb") as f:
        p.write_bytes(f.read())
    return p


def test_state_file_round_trip_is_byte_identical(sta_path, tmp_path):
    d = formats.read_sta(sta_path)
    assert (d["npoly"], d["nwater"], d["nmatom"], d["natom"], d["mchain"]) == (200, 0, 3584, 6364, 3)
    assert d["sections"] == (2780, 0, 3584)
    assert d["nstep"] == 1
    np.testing.assert_array_equal(d["box"], [0.38595e-8, 0.38583e-8, 0.1308e-7])
    assert d["types"][0] == "CH" and d["types"][-1] == "PT"
    # the first atom line, parsed
    np.testing.assert_array_equal(d["pos"][0], [0.1269835373332408e-08, 0.3363856188387809e-09,
                                                0.5452248512588356e-08])
    assert d["andersen"][1] == 0.8443527726351980e+09
    out = tmp_path / "out.dat"
    formats.write_sta(out, d)
    assert out.read_bytes() == sta_path.read_bytes()


def test_state_file_rejects_inconsistent_counts(sta_path, tmp_path):
    from thrid_amd.pgk import PgkError
    text = sta_path.read_text().replace("#natom      6364", "#natom      6365")
    bad = tmp_path / "bad.dat"
    bad.write_text(text)
    with pytest.raises(PgkError):
        formats.read_sta(bad)


def test_out_ene_row_matches_the_reference_file():
    lines = open(os.path.join(GOLD, "out_ene.dat")).read().splitlines()
    f = lines[1].split()
    step, vals = int(f[0]), [float(x) for x in f[1:]]
    assert formats.out_ene_row(step, vals) == lines[1] + "\n"


@pytest.mark.parametrize("v,w", [(0.0, "  0.0000000000000000E+00"), (-0.0, " -0.0000000000000000E+00"),
                                 (1.0, "  0.1000000000000000E+01"), (-2.5e-300, " -0.2500000000000000-299"),
                                 (1e100, "  0.1000000000000000+101")])
def test_fortran_e_edge_cases(v, w, tmp_path):
    # through the state writer's box field (E24.16)
    d = dict(npoly=0, nwater=0, nmatom=0, natom=0, mchain=0, sections=(0, 0, 0),
             box=[v, 1.0, 1.0], nstep=0, pos=np.zeros((0, 3)), vel=np.zeros((0, 3)), types=[],
             nhc=np.zeros((0, 2)), andersen=np.zeros(2), aniso=np.zeros(12))
    p = tmp_path / "e.dat"
    formats.write_sta(p, d)
    line = [l for l in p.read_text().splitlines() if l.startswith("#xcel")][0]
    assert line == "#xcel" + w


def test_ini_words_of_the_reference_script():
    """peachgk.ini as rdscript reads it: keys, repeated keys, Fortran reals
    with d exponents, logicals, trailing comments, END."""
    from thrid_amd.pgk import PgkError
    ini = formats.Ini(os.path.join(GOLD, "peachgk.ini"))
    assert ini.float("xcel") == 38.595000000e-10
    assert ini.float("dt_long_cal") == 1.0e-15
    assert ini.int("nstep_short") == 5
    assert ini.str("mts_vdw") == "long" and ini.str("mts_bond") == "short"
    assert ini.str("ouenename") == "out_ene.dat" and ini.bool("ouenename", 1) is True
    assert ini.bool("ouforname", 1) is False
    assert ini.bool("ifstarec") is False and ini.bool("ifcreatecor") is True
    assert ini.count("npolymoletyp") == 2
    assert [ini.int("npolymoletyp", k, occ=1) for k in range(3)] == [2, 190, 14]
    assert ini.str("npolymoletyp", 4, occ=1) == "C4H10_top.dat"
    assert ini.float("r_expand") == 1.1111111     # the "! expansion ratio ..." comment follows
    assert ini.count("nmatomtyp") == 6 and ini.str("nmatomtyp", 4, occ=1) == "localheat" and ini.float("nmatomtyp", 5, occ=1) == 146.0
    assert ini.count("no_such_key") == 0 and ini.nvalues("no_such_key") == -1
    with pytest.raises(PgkError):
        ini.float("mts_vdw")
    with pytest.raises(PgkError):
        ini.int("npolytyp", 3)
    ini.close()


def test_out_pre_header_and_row_match_the_reference_file():
    lines = open(os.path.join(GOLD, "out_pre.dat")).read().splitlines()
    assert formats.out_pre_header() == lines[0] + "\n"
    f = lines[1].split()
    assert len(f) == 82                      # nstep + nprefield = 81 (output_data.F90:700)
    assert formats.out_pre_row(int(f[0]), [float(x) for x in f[1:]]) == lines[1] + "\n"
    hi = formats.out_pre_row(int(f[0]), [float(x) for x in f[1:]], high=True)
    assert len(hi) == 10 + 81 * 24 + 1


@pytest.mark.parametrize("all_dir", [True, False])
def test_out_htf_layout(all_dir):
    """outhtf's layout (output_hf.F90:164-167, 485-686): 36 fields per
    region; for each field the regions in order, x y z of each with all_dir;
    header names name_<c>_<region> left-justified in A23; values from
    calheatf's [region, term, component] times jref, classes this library
    does not evaluate as zero."""
    from thrid_amd import abi
    nreg = 3
    nc = 3 if all_dir else 1
    head = formats.out_htf_header(nreg, all_dir)
    names = head.split()
    assert names[0] == "nstep" and len(names) == 1 + 36 * nreg * nc
    assert len(head) == 9 + 1 + 24 * 36 * nreg * nc + 1
    comps = "xyz" if all_dir else "z"
    assert names[1:1 + nreg * nc] == [f"hftot_atm_{c}_{r}" for r in range(1, nreg + 1)
                                      for c in comps]
    assert names[-1] == f"hfvirilj_ar2pt_z_{nreg}"
    assert head.index("hfkin_atm_") == 9 + 1 + 24 * nreg * nc + 1   # left-justified
    hf = np.zeros((nreg, abi.NHFJ, 3))
    for r in range(nreg):
        for t in range(abi.NHFJ):
            for c in range(3):
                hf[r, t, c] = (r + 1) * 1000 + t * 10 + c + 1
    row = formats.out_htf_row(7, hf, jref=0.5, all_dir=all_dir)
    vals = row.split()
    assert int(vals[0]) == 7 and len(vals) == len(names)
    assert len(row) == 10 + 24 * 36 * nreg * nc + 1
    got = dict(zip(names[1:], (float(v) for v in vals[1:])))
    for r in range(nreg):
        for ci, c in enumerate("xyz"):
            if c not in comps:
                continue
            k = f"_{c}_{r + 1}"
            assert got["hftot_atm" + k] == 0.5 * hf[r, abi.HFJ_TOT, ci]
            assert got["hfkin_atm" + k] == 0.5 * hf[r, abi.HFJ_KIN, ci]
            assert got["hfpotlj_atm" + k] == 0.5 * hf[r, abi.HFJ_POT + 7, ci]
            assert got["hfvirielc_atm" + k] == 0.5 * hf[r, abi.HFJ_VIRI + 6, ci]
            assert got["hfpot14_atm" + k] == 0.5 * hf[r, abi.HFJ_POT + 3, ci]
            assert got["hfpotmor_atm" + k] == 0.0 and got["hfvirilj_ar12" + k] == 0.0
