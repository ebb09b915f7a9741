"""The Scheme path on the CPU: XML loading, packing, load errors, the .dat
reader and the FITS writer (librt_scheme_hip.so; no device is used).

- A scheme written in the reference's XML format (scheme.cpp:272-857) loads
  into exactly the packed table, spline pool and beam description that the
  Python packer builds from the same description (bit for bit: angles
  through the deg2rad round trip, tables through TabulatedFunction's file
  reader and the natural-spline recurrence).
- Planted faults in a scheme file raise the exceptions the reference raises
  for them, in the reference's order of checks (codes: 100 + SchemeError,
  200 + BeamError, 300 + SurfaceError, 400 + LensError, 500 + GratingError,
  600 + TabulatedFunctionError, 700 + a bare GratingError enumerator).
- load_beam_data reads Beam::Save's layout (good rays only, or all).
- rt_write_fits_image writes a primary HDU a FITS reader accepts: 80-column
  cards in 2880-byte records, BITPIX -64, NAXIS1 = xdim, big-endian data.
"""
import ctypes as C
import dataclasses
import os
import re
import struct

import numpy as np
import pytest

import branch_schemes as B
from rt_amd import scheme_xml as X
from rt_amd import lib
from rt_amd import scheme as S
from rt_amd import schemes
from test_ref_pipeline import echelle_fixed_order, mixed


def scheme_path_view(sc):
    """the scheme as the Scheme path packs it: fixed orders, no CCD frame or
    detector functions (the reference's scheme file has none)"""
    surfs = [dataclasses.replace(s, blaze_k=0.0, order_range=(-1000000, 1000000), ccd=None, qe=1.0, spectrum=1.0,
                                 order=s.order if s.type == S.RT_GRATING else 0,
                                 grating_type=s.grating_type if s.type == S.RT_GRATING else S.RT_REFLECTIVE)
             for s in sc.surfaces]
    return S.Scheme(surfs, sc.beam, name=sc.name)


SCHEMES = {
    "czerny_turner": lambda: schemes.czerny_turner(n_per_lambda=1000, nlambda=5),
    "echelle_fixed_order": lambda: echelle_fixed_order(n_per_lambda=100, nlambda=16)[0],
    "mixed": lambda: mixed(n=2000)[0],
    "branches": lambda: B.branches(2000),
}


@pytest.mark.parametrize("case", sorted(SCHEMES))
def test_xml_loads_into_the_python_packing(case, tmp_path):
    sc = SCHEMES[case]()
    path = X.write_scheme(sc, str(tmp_path))
    rc, table, pool, beam, n_rays = lib.scheme_pack_file(path)
    assert rc == 0, rc
    want = scheme_path_view(sc)
    assert table.shape == want.table.shape
    diff = np.flatnonzero(table != want.table)
    assert diff.size == 0, [(i // S.SURF_STRIDE, i % S.SURF_STRIDE, table[i], want.table[i]) for i in diff[:8]]
    assert np.array_equal(pool, want.pool)
    assert np.array_equal(beam, sc.beam.pack())
    assert n_rays == sc.beam.n_rays


def test_xml_reader_text_rules(tmp_path):
    """entities and character references are decoded, comments, the
    declaration and CDATA are handled, values keep their inner spaces and are
    trimmed of spaces and tabs only (scheme.cpp:27-41), names are upper-cased
    except file names"""
    sc = schemes.czerny_turner(n_per_lambda=10, nlambda=5)
    path = X.write_scheme(sc, str(tmp_path))
    text = open(path).read()
    text = text.replace("<name>", "<!-- a comment <surface> --><name>")
    text = text.replace("<class>CONIC</class>", "<class> \tco&#78;ic </class>", 1)
    text = text.replace("<type>MIRROR</type>", "<type><![CDATA[mirror]]></type>", 1)
    open(path, "w").write(text)
    rc, table, _, _, _ = lib.scheme_pack_file(path)
    assert rc == 0
    assert np.array_equal(table, scheme_path_view(sc).table)


def planted(tmp_path, mutate, sc=None, **kw):
    sc = sc or schemes.czerny_turner(n_per_lambda=10, nlambda=5)
    path = X.write_scheme(sc, str(tmp_path), **kw)
    if mutate:
        text = mutate(open(path).read())
        open(path, "w").write(text)
    return lib.scheme_pack_file(path)[0]


def lens_scheme():
    return mixed(n=100)[0]


def grating_scheme():
    return schemes.czerny_turner(n_per_lambda=10, nlambda=5)


FileOpenFailure, ParseFailure, BadAttrValue, NanAttrValue = 0, 1, 3, 4
NoSchemeDesc, NoBeamDesc, NoSurfaceDesc, NoRangeFile, BadRange = 6, 7, 8, 10, 11


@pytest.mark.parametrize("name,mutate,code", [
    ("unbalanced", lambda t: t.replace("</general>", ""), 100 + ParseFailure),
    ("no general", lambda t: re.sub(r"<general>.*?</general>", "", t, flags=re.S), 100 + NoSchemeDesc),
    ("no result_file", lambda t: re.sub(r"<result_file>.*?</result_file>", "", t), 100 + ParseFailure),
    ("empty name", lambda t: re.sub(r"<name>.*?</name>", "<name> </name>", t), 100 + BadAttrValue),
    ("no beam", lambda t: re.sub(r"<beam>.*?</beam>", "", t, flags=re.S), 100 + NoBeamDesc),
    ("no surfaces", lambda t: re.sub(r"<surface>.*?</surface>", "", t, flags=re.S), 100 + NoSurfaceDesc),
    ("bad class", lambda t: t.replace("<class>CONIC</class>", "<class>SPHERE</class>", 1), 100 + BadAttrValue),
    ("bad shape", lambda t: t.replace("<shape>CIRCLE</shape>", "<shape>OVAL</shape>", 1), 100 + BadAttrValue),
    ("no distance", lambda t: re.sub(r"<distance>.*?</distance>", "", t, count=1), 100 + ParseFailure),
    ("element value", lambda t: t.replace("<angles>", "<angles><x>1</x>", 1), 100 + BadAttrValue),
    ("two angles", lambda t: re.sub(r"<angles>[^<]*</angles>", "<angles>1 2</angles>", t, count=1),
     300 + 5),                                                        # InvalidSurfaceAngles
    ("two distances", lambda t: re.sub(r"<distance>[^<]*</distance>", "<distance>0 1</distance>", t, count=1),
     300 + 4),                                                        # InvalidSurfaceDistance
    ("short conic", lambda t: re.sub(r"<params>[^<]*</params>", "<params>0.001 0 2 1e-9</params>", t, count=1),
     300 + 1),                                                        # InvalidSurfaceParam
    ("one param", lambda t: re.sub(r"<params>[^<]*</params>", "<params>0.001</params>", t, count=1), 300 + 1),
    ("no center", lambda t: re.sub(r"<center>[^<]*</center>", "<center>1</center>", t, count=1), 300 + 3),
    ("reversed range", lambda t: re.sub(r"<range>[^<]*</range>", "<range>0.7 0.4</range>", t), 100 + BadRange),
    ("range file missing", lambda t: re.sub(r"<range>[^<]*</range>", "<range>no_such_file.txt</range>", t),
     100 + NoRangeFile),
    ("beam params", lambda t: re.sub(r"(<beam>.*?)<params>[^<]*</params>", r"\1<params>10 0</params>", t,
                                     flags=re.S), 200 + 1),           # Beam InvalidParams
    ("beam type", lambda t: t.replace("<type>PARALLEL</type>", "<type>LASER</type>"), 200 + 1),
    ("grating order", lambda t: t.replace("<order>1</order>", "<order>3 1</order>"), 500 + 2),   # BadOrder
    ("grating type", lambda t: t.replace("<grating_type>REFLECTIVE</grating_type>",
                                         "<grating_type>HOLOGRAPHIC</grating_type>"), 700 + 3),  # bare BadType
])
def test_planted_faults_raise_the_reference_error(name, mutate, code, tmp_path):
    assert planted(tmp_path, mutate) == code, name


def test_missing_scheme_file(tmp_path):
    assert lib.scheme_pack_file(str(tmp_path / "none.xml"))[0] == 100 + FileOpenFailure


def test_lens_index_faults(tmp_path):
    sc = lens_scheme()
    # two constants: TabulatedFunction's BadValue gets through (surface.cpp:400-408)
    plate = S.Scheme([S.Surface(S.RT_PLANE, S.RT_LENS, S.RT_CIRCLE, size=(30.0, 0.0), distance=(0, 0, 10.0),
                                n1=1.0, n2=1.5),
                      S.Surface(S.RT_PLANE, S.RT_DETECTOR, S.RT_CIRCLE, size=(30.0, 0.0), distance=(0, 0, 10.0))],
                     sc.beam, name="plate")
    assert planted(tmp_path / "a", lambda t: t.replace("<n2>1.5</n2>", "<n2>-1.5</n2>", 1), sc=plate) == 600 + 1
    # a file constructor turns any table fault into BadRefIndexFile (surface.cpp:410-470)
    bad = lambda t: re.sub(r"<n2>[^<]*n2_1\.txt</n2>", "<n2>missing_glass.txt</n2>", t, count=1)
    assert planted(tmp_path / "b", bad, sc=sc) == 400 + 1


def test_range_file_with_orders(tmp_path):
    """a range file gives one range per line and, per line, one order for each
    grating (scheme.cpp:215-264); run's first range sets the gratings' orders"""
    sc = grating_scheme()
    path = X.write_scheme(sc, str(tmp_path), ranges=[(0.40, 0.55, 2), (0.55, 0.70, 1)])
    rc, table, _, beam, _ = lib.scheme_pack_file(path)
    assert rc == 0
    t = table.reshape(sc.n_surf, S.SURF_STRIDE)
    g = [i for i, s in enumerate(sc.surfaces) if s.type == S.RT_GRATING][0]
    assert t[g, S.S_ORDER] == 2
    assert tuple(beam[S.B_RANGE:S.B_RANGE + 2]) == (0.40, 0.55)
    # a line with too few numbers is BadRange
    path = X.write_scheme(sc, str(tmp_path / "bad"), ranges=[(0.40, 0.55)])
    assert lib.scheme_pack_file(path)[0] == 100 + BadRange


def test_load_beam_data_reads_the_save_layout(tmp_path):
    from oracle import oracle as O
    n, ng = 9, 5
    arr = np.arange(7 * n, dtype=np.float64).reshape(7, n) * 0.5 + 0.25
    path = str(tmp_path / "b.dat")
    open(path, "wb").write(O.dat_bytes(n, ng, 3, arr))
    L = lib.scheme_lib()
    L.load_beam_data = getattr(L, "_Z14load_beam_dataPKcPmS1_S1_PPdS3_S3_S3_S3_S3_S3_b")
    L.load_beam_data.argtypes = ([C.c_char_p] + [C.POINTER(C.c_size_t)] * 3 +
                                 [C.POINTER(C.POINTER(C.c_double))] * 7 + [C.c_bool])
    libc = C.CDLL(None)
    for all_rays in (False, True):
        tot, good, nl = C.c_size_t(), C.c_size_t(), C.c_size_t()
        ptrs = [C.POINTER(C.c_double)() for _ in range(7)]
        args = [C.byref(p) for p in ptrs]
        args[2] = None                                  # Z skipped
        rc = L.load_beam_data(path.encode(), C.byref(tot), C.byref(good), C.byref(nl), *args, all_rays)
        assert rc == 0 and (tot.value, good.value, nl.value) == (n, ng, 3)
        m = n if all_rays else ng
        for k in range(7):
            if k == 2:
                continue
            got = np.ctypeslib.as_array(ptrs[k], shape=(m,)).copy()
            assert np.array_equal(got, arr[k][:m])
            libc.free(ptrs[k])
    assert L.load_beam_data(str(tmp_path / "none.dat").encode(), C.byref(tot), C.byref(good), C.byref(nl),
                            *[None] * 7, False) == 1


def read_fits(path):
    """a minimal FITS primary-HDU reader (FITS Standard 4.0, sections 3-4)"""
    data = open(path, "rb").read()
    assert len(data) % 2880 == 0
    cards, pos = {}, 0
    while True:
        card = data[pos:pos + 80].decode("ascii")
        pos += 80
        key = card[:8].strip()
        if key == "END":
            break
        assert card[8:10] == "= "
        v = card[10:].split(" / ")[0].strip()
        cards[key] = v
    hdr_end = (pos + 2879) // 2880 * 2880
    assert set(data[pos:hdr_end]) <= {ord(" ")}
    nx, ny = int(cards["NAXIS1"]), int(cards["NAXIS2"])
    img = np.frombuffer(data[hdr_end:hdr_end + 8 * nx * ny], dtype=">f8").reshape(ny, nx)
    assert set(data[hdr_end + 8 * nx * ny:]) <= {0}
    return cards, img


def test_fits_writer_round_trip(tmp_path):
    img = np.random.default_rng(3).standard_normal((37, 53)) * 1e6
    img[0, 0], img[1, 2] = 0.0, -0.0
    path = str(tmp_path / "x.fits")
    long_name = "/a/very/long/path/" + "d" * 80 + "/list.txt"
    assert lib.write_fits_image(path, img, [("RT_LIST", long_name), ("QE_FILE", "it's.qe")]) == 0
    cards, got = read_fits(path)
    assert cards["SIMPLE"] == "T" and cards["BITPIX"] == "-64" and cards["NAXIS"] == "2"
    assert cards["NAXIS1"] == "53" and cards["NAXIS2"] == "37"
    assert cards["QE_FILE"] == "'it''s.qe'"
    assert cards["RT_LIST"].startswith("'/a/very/long/path/") and len(cards["RT_LIST"]) == 70
    assert np.array_equal(got, img) and np.signbit(got[1, 2])
    raw = open(path, "rb").read()
    assert raw[:80] == ("SIMPLE  = %20s / file does conform to FITS standard" % "T").ljust(80).encode()
