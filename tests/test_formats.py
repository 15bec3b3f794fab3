"""File formats of the reference (include/ifos2d_formats.h), host side:
the quasi-JSON parameter file (the reference's own par/in_and_out/
IFOS2D_FW.json, committed as tests/golden/IFOS2D_FW.json), binary model
files and SU seismograms, checked against numpy readings of the bytes."""
import os
import tempfile

import numpy as np
import pytest

import ifos_amd
from ifos_amd import IfosError, formats
from ifos_amd import formats as F
import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_reference_parameter_file():
    objs = formats.read_json(os.path.join(GOLDEN, "IFOS2D_FW.json"))
    assert objs["NX"] == "500" and objs["NY"] == "100" and objs["DH"] == "0.2"
    assert objs["FDORDER"] == "2" and objs["FW"] == "20" and objs["npower"] == "4.0"
    # two objects on one line, names and values cut at the blanks
    assert objs["REFRECX"] == "0.0" and objs["REFRECY"] == "0.0"
    assert objs["XREC1"] == "6.0" and objs["YREC1"] == "0.2"
    # comment lines are not objects
    assert not any("comment" in v.lower() for v in objs.values())
    c = formats.config_from_json(os.path.join(GOLDEN, "IFOS2D_FW.json"), ifos_amd.make_config())
    assert (c.NX, c.NY, c.FDORDER, c.FW, c.FREE_SURF, c.BOUNDARY) == (500, 100, 2, 20, 1, 0)
    assert c.DH == np.float32(0.2) and c.DT == np.float32(5e-5)
    # NT = iround(TIME/DT) in FP32 (src/IFOS2D.c:257)
    assert c.NT == int(np.floor(np.float32(0.5) / np.float32(5e-5) + 0.5)) == 10000
    assert abs(c.VPPML - 600.0) < 1e-3 and abs(c.k_max_PML - 8.0) < 1e-6


def test_parameter_file_errors(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{\n"NX" : "10",\n"NY, DH" : "5",\n}\n')
    with pytest.raises(ifos_amd.IfosError):
        formats.read_json(str(p))
    p.write_text('{\n"NX" : "1.5",\n"NY" : "10",\n"DH" : "1",\n"DT" : "1e-3",\n"TIME" : "0.1",\n"FDORDER" : "4"\n}\n')
    with pytest.raises(ifos_amd.IfosError):
        formats.config_from_json(str(p))
    p.write_text('{\n"NX" : "10",\n"NY" : "10",\n"DH" : "1",\n"DT" : "1e-3",\n"FDORDER" : "4"\n}\n')
    with pytest.raises(ifos_amd.IfosError):   # TIME missing
        formats.config_from_json(str(p))


def test_model_file_layout(tmp_path):
    rng = np.random.default_rng(0)
    m = rng.standard_normal((7, 11)).astype(np.float32)   # [NY][NX]
    p = str(tmp_path / "m.vp")
    formats.write_model(p, m)
    raw = np.fromfile(p, np.float32)
    # x outer, y inner (src/writemod.c): the file is the transpose, row-major
    np.testing.assert_array_equal(raw, m.T.ravel())
    np.testing.assert_array_equal(formats.read_model(p, 11, 7), m)
    with pytest.raises(ifos_amd.IfosError):
        formats.read_model(p, 12, 7)


def test_su_roundtrip_and_headers(tmp_path):
    rng = np.random.default_rng(1)
    ntr, ns, DH, DT = 5, 300, 2.0, 2.5e-4
    data = rng.standard_normal((ntr, ns)).astype(np.float32)
    rec = np.array([[10 + 3 * k, 4] for k in range(ntr)], np.int32)
    xs, ys = 52.0, 8.0
    p = str(tmp_path / "s.su")
    formats.write_su(p, data, rec, DH, DT, xs, ys, ndt=2)
    raw = np.fromfile(p, np.uint8).reshape(ntr, 240 + 4 * ns)
    hdr = raw[:, :240]
    i32 = lambda off: hdr[:, off:off + 4].copy().view(np.int32).ravel()
    i16 = lambda off: hdr[:, off:off + 2].copy().view(np.int16).ravel()
    u16 = lambda off: hdr[:, off:off + 2].copy().view(np.uint16).ravel()
    assert list(i32(0)) == [1, 2, 3, 4, 5] and list(i32(20)) == [1, 2, 3, 4, 5]     # tracl, cdp
    assert list(i16(28)) == [1] * ntr                                                 # trid
    xr = rec[:, 0] * np.float32(DH)
    yr = rec[:, 1] * np.float32(DH)
    off = np.floor(np.sqrt((xs - xr) ** 2 + (ys - yr) ** 2) * 1000.0 + 0.5)
    np.testing.assert_array_equal(i32(36), off.astype(np.int32))                     # offset (mm)
    np.testing.assert_array_equal(i32(72), np.full(ntr, 52000, np.int32))            # sx
    np.testing.assert_array_equal(i32(80), (xr * 1000).astype(np.int32))             # gx
    assert list(i16(68)) == [-1000] * ntr and list(i16(70)) == [-1000] * ntr         # scalel, scalco
    assert list(u16(114)) == [ns] * ntr and list(u16(116)) == [500] * ntr            # ns, dt (us)
    np.testing.assert_array_equal(raw[:, 240:].copy().view(np.float32), data)
    np.testing.assert_array_equal(formats.read_su(p, ns), data)
    with pytest.raises(ifos_amd.IfosError):
        formats.read_su(p, ns + 1)


def test_trkill_file(tmp_path):
    t = np.array([[0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 0]], np.int32)
    p = tmp_path / "kill.dat"
    p.write_text("\n".join(" ".join(str(v) for v in row) for row in t) + "\n")
    np.testing.assert_array_equal(F.read_trkill(str(p), 4, 3), t)
    # a file that ends inside the last value, or is short, is rejected as in calc_res
    p.write_text("\n".join(" ".join(str(v) for v in row) for row in t))
    with pytest.raises(IfosError):
        F.read_trkill(str(p), 4, 3)
    p.write_text("0 1 0\n1 1 0\n")
    with pytest.raises(IfosError):
        F.read_trkill(str(p), 4, 3)


@pytest.mark.skipif(not O.ref_available(), reason="reference tree not built into oracle/_ref")
def test_trkill_offsets_match_reference():
    import ctypes as C
    rng = np.random.default_rng(3)
    rec = np.stack([np.arange(5, 200, 7), np.full(28, 4)], 1).astype(np.int32)
    src = np.stack([rng.uniform(20, 400, 6), rng.uniform(4, 30, 6)], 1).astype(np.float32)
    DH, lo, hi = 2.0, 50.0, 250.0
    got = F.trkill_offsets(rec, src, DH, lo, hi)
    want = np.zeros((len(rec), len(src)), np.int32)
    O.ref().ref_trkill_offsets(len(rec), rec.ctypes.data_as(C.POINTER(C.c_int)), len(src),
                               src.ctypes.data_as(C.POINTER(C.c_float)), C.c_float(DH), C.c_float(lo), C.c_float(hi),
                               want.ctypes.data_as(C.POINTER(C.c_int)))
    np.testing.assert_array_equal(got, want)
    assert 0 < got.sum() < got.size
    # closed form: offsets by numpy in float64 away from the band edges
    off = np.hypot(src[None, :, 0] - rec[:, None, 0] * DH, src[None, :, 1] - rec[:, None, 1] * DH)
    far = (np.abs(off - lo) > 1e-3) & (np.abs(off - hi) > 1e-3)
    np.testing.assert_array_equal(got[far], ((off > lo) & (off < hi)).astype(np.int32)[far])
    # added to an existing table (TRKILL_OFFSET 2)
    base = np.zeros_like(got)
    base[0, :] = 1
    both = F.trkill_offsets(rec, src, DH, lo, hi, table=base)
    np.testing.assert_array_equal(both, np.maximum(base, got))


# ------------------------------------------- source / receiver files, inseis --
needs_ref_files = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")


@pytest.fixture
def short_tmp():
    """A directory with a short path: the reference keeps file names in
    char[74] buffers (STRING_SIZE, src/fd.h:22) and aborts the process when
    the truncated name does not open, and pytest's tmp_path under xdist is
    longer than that."""
    import pathlib
    with tempfile.TemporaryDirectory(prefix="f", dir="/tmp") as d:
        yield pathlib.Path(d)


def write_source_file(path, rows):
    with open(path, "w") as f:
        f.write(f"{len(rows)}\n")
        for r in rows:
            f.write(" ".join(f"{v:g}" for v in r) + "\n")


SOURCE_ROWS = [(100.0, 0.0, 20.0, 0.0, 30.0, 1.0),                      # 6 columns: azimuth 0, type default
               (210.5, 3.0, 9.75, 0.01, 25.0, 2.0e3, 45.0),               # 7 columns: type default
               (400.0, 0.0, 40.0, 0.0, 35.0, 1.0, 30.0, 4.0),             # rotated force
               (3.0, 0.0, 3.0, 0.002, 15.0, 1.5, 0.0, 2.0),
               (-40.0, 0.0, 10.0, 0.0, 20.0, 1.0, 0.0, 1.0)]               # left of the grid


@needs_ref_files
@pytest.mark.parametrize("stype", [1, 3])
def test_source_file_matches_reference(short_tmp, stype):
    p = str(short_tmp / "src.dat")
    write_source_file(p, SOURCE_ROWS)
    mine, fc = F.read_sources(p, source_type=stype)
    ref, fc_ref = O.ref_sources(p, source_type=stype)
    np.testing.assert_array_equal(mine, ref)
    assert fc == fc_ref == 35.0
    assert mine[0, 6] == 0.0 and mine[0, 7] == stype and mine[1, 7] == stype and mine[2, 7] == 4.0
    # x, y from columns 1 and 3 (column 2 is z)
    assert mine[1, 0] == 210.5 and mine[1, 1] == 9.75


@needs_ref_files
@pytest.mark.parametrize("DH", [1.0, 2.0, 3.3])
def test_sources_to_grid_matches_splitsrc(tmp_path, DH):
    p = str(tmp_path / "src.dat")
    write_source_file(p, SOURCE_ROWS)
    rows, _ = F.read_sources(p)
    for NX, NY in ((200, 30), (50, 12)):
        np.testing.assert_array_equal(F.sources_to_grid(rows, DH, NX, NY), O.ref_splitsrc(rows, DH, NX, NY))
    g = F.sources_to_grid(rows, 1.0, 500, 100)
    assert g[0, 0] == 100 and g[0, 1] == 20


def test_source_file_missing_columns(tmp_path):
    p = str(tmp_path / "src.dat")
    write_source_file(p, [(1.0, 0.0, 2.0, 0.0, 10.0)])
    with pytest.raises(ifos_amd.IfosError):
        F.read_sources(p)


@needs_ref_files
@pytest.mark.parametrize("DH,ref", [(1.0, (0.0, 0.0)), (2.5, (3.0, -1.0))])
def test_receiver_file_matches_reference(short_tmp, DH, ref):
    p = str(short_tmp / "rec.dat")
    pts = [(10.0, 5.0), (12.4, 5.0), (12.6, 5.2), (12.0, 5.0), (30.25, 7.75), (10.0, 5.0), (0.1, 5.0)]
    with open(p, "w") as f:
        for x, y in pts:
            f.write(f"{x} {y}\n")
    mine = F.read_receivers(p, DH, refrec=ref)
    np.testing.assert_array_equal(mine, O.ref_receivers(1, p, DH, refrec=ref))
    assert len(mine) < len(pts)  # duplicates on one grid point dropped


@needs_ref_files
@pytest.mark.parametrize("line,ngeoph", [((10.0, 6.0, 210.0, 6.0), 2), ((10.0, 4.0, 90.0, 30.0), 3),
                                          ((20.0, 5.0, 26.0, 95.0), 4)])
def test_receiver_line_matches_reference(line, ngeoph):
    mine = F.receiver_line(1.0, *line, ngeoph)
    np.testing.assert_array_equal(mine, O.ref_receivers(0, None, 1.0, line=line, ngeoph=ngeoph))
    assert len(mine) > 3


@needs_ref_files
def test_observed_su_matches_inseis():
    rng = np.random.default_rng(4)
    ntr, ns = 9, 57
    data = rng.standard_normal((ntr, ns)).astype(np.float32)
    rec = np.stack([np.arange(ntr) + 5, np.full(ntr, 3)], 1).astype(np.int32)
    # the reference's DATA_DIR buffer holds 74 bytes: a short prefix
    with tempfile.TemporaryDirectory(prefix="su", dir="/tmp") as d:
        prefix = os.path.join(d, "o")
        F.write_su(prefix + "_vx.su.shot2", data, rec, 2.0, 1e-3, 10.0, 4.0)
        mine = F.read_su_observed(prefix + "_vx.su.shot2", ntr, ns)
        ref = O.ref_inseis(prefix, 2, ntr, ns)
    np.testing.assert_array_equal(mine, ref)
    assert np.all(mine[:, 0] == 0) and np.array_equal(mine[:, 1:], data[:, 1:])


def test_matrix_dump_is_model_layout(tmp_path):
    m = np.arange(12, dtype=np.float32).reshape(3, 4)
    F.dump_matrix(str(tmp_path / "grad"), m)
    raw = np.fromfile(str(tmp_path / "grad.bin"), np.float32)
    np.testing.assert_array_equal(raw, m.T.ravel())  # x outer, y inner
    np.testing.assert_array_equal(F.read_model(str(tmp_path / "grad.bin"), 4, 3), m)
