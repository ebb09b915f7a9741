"""The C-ABI library (rt_amd/librt_hip.so) without a GPU: it loads, exports
every entry point include/rt_hip.h declares, agrees with the header's
layout constants, validates schemes on the host and generates/compiles the
per-scheme kernels (hiprtc runs without a device)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from rt_amd import lib, schemes
from rt_amd import scheme as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rt_hip.h")


def declared_functions(path):
    text = open(path).read()
    return sorted(set(re.findall(r"^int (rt_hip_\w+)\(", text, re.M)))


def test_library_exports_every_declared_entry_point():
    names = declared_functions(HEADER)
    assert "rt_hip_trace" in names and len(names) >= 8
    out = subprocess.run(["nm", "-D", "--defined-only", lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    L = lib.lib()
    for n in names:
        assert hasattr(L, n)


def _enum_values(text):
    vals = {}
    for name, val in re.findall(r"#define (RT_[A-Z0-9_]+) (\d+)", text):
        vals[name] = int(val)
    for name, val in re.findall(r"\b(RT_[A-Z0-9_]+)\s*=\s*([A-Z0-9_+* ]+?)\s*(?:,|}|/\*|\n)", text):
        vals[name] = int(eval(val, {}, dict(vals)))
    return vals


def test_python_layout_matches_header():
    v = _enum_values(open(HEADER).read())
    pairs = {"RT_SURF_STRIDE": S.SURF_STRIDE, "RT_BEAM_STRIDE": S.BEAM_STRIDE, "RT_STATS_LEN": S.STATS_LEN,
             "RT_S_C": S.S_C, "RT_S_A": S.S_A, "RT_S_CR": S.S_CR, "RT_S_DIST": S.S_DIST, "RT_S_ANG": S.S_ANG,
             "RT_S_SIN": S.S_SIN, "RT_S_COS": S.S_COS, "RT_S_GRCONST": S.S_GRCONST, "RT_S_N2VAL": S.S_N2VAL,
             "RT_S_CCD_XDIM": S.S_CCD_XDIM, "RT_S_SPVAL": S.S_SPVAL, "RT_B_ANGPAR": S.B_ANGPAR,
             "RT_B_COMPOSITE": S.B_COMPOSITE, "RT_ST_SURF": S.ST_SURF, "RT_MAX_SURF": S.RT_MAX_SURF,
             "RT_TRACE_FLUX": S.RT_TRACE_FLUX, "RT_DETECTOR": S.RT_DETECTOR, "RT_TRANSPARENT": S.RT_TRANSPARENT}
    for k, py in pairs.items():
        assert v[k] == py, k


def test_packed_angles_and_trig():
    sch = schemes.echelle(1, 4)
    t = sch.table.reshape(sch.n_surf, S.SURF_STRIDE)
    for s, surf in enumerate(sch.surfaces):
        for j in range(3):
            rad = surf.angles[j] * S.DEG2RAD
            rad = (rad / S.DEG2RAD) * S.DEG2RAD      # surface.cpp:201,227 + beam.cpp:609-611
            assert t[s, S.S_ANG + j] == rad
            assert t[s, S.S_SIN + j] == np.sin(rad) and t[s, S.S_COS + j] == np.cos(rad)


def test_validate_scheme():
    sch = schemes.echelle(1, 4)
    assert lib.validate(sch) == 0
    bad = schemes.echelle(1, 4)
    bad.table = bad.table.copy()
    bad.table[S.S_NPOLY] = 11
    assert lib.validate(bad) == 4 and "aspheric" in lib.last_error()
    bad.table[S.S_NPOLY] = 0
    bad.table[S.S_CLASS] = 7
    assert lib.validate(bad) == 4
    lens = 5 * S.SURF_STRIDE
    bad = schemes.echelle(1, 4)
    bad.table = bad.table.copy()
    bad.table[lens + S.S_N2VAL] = len(bad.pool) + 3          # spline reference past the pool
    assert lib.validate(bad) == 4 and "refractive index" in lib.last_error()
    L = lib.lib()
    assert L.rt_hip_validate_scheme(None, 0, None, 0) == 4


@pytest.mark.parametrize("make,mode", [(lambda: schemes.echelle(1, 4), S.RT_BEAM_GAUSS),
                                       (lambda: schemes.czerny_turner(1, 5), -1)])
def test_generated_kernel_compiles(make, mode):
    rc, size, text = lib.jit_check(make(), mode)
    assert rc == 0, text[-2000:]
    assert size > 10000
    assert text.count("rt::surface_step(") == make().n_surf


def test_generated_table_kernel_compiles():
    """The trace generated for a grid-wavelength beam reading a per-wavelength
    table (0x400 = RT_JIT_LTAB) compiles.  A single scheme's trace reads the
    derived table: mu for each lens and grating, the grating term for each
    grating, the flux weight for the CCD detector -- on the echelle, 2 + 2
    columns for its two gratings, 6 for its six lens surfaces and 1 for the
    detector, each loaded just before the surface that uses it.  The batched
    trace (0x100 = RT_JIT_BATCH) reads one column per spline-valued function
    that is not reused from the previous surface: the three glasses (n2 of
    each first lens surface) and the detector QE."""
    sch = schemes.echelle(1, 4)
    rc, size, text = lib.jit_check(sch, S.RT_BEAM_GAUSS | 0x400)
    assert rc == 0, text[-2000:]
    assert "double lv[11];" in text
    assert text.count("rt::load_lval(lv, c, ") == 11
    assert text.count(", lv, ") == 9
    assert text.count(", 0, true>(") == sch.n_surf
    rc, size, text = lib.jit_check(sch, S.RT_BEAM_GAUSS | 0x400 | 0x100)
    assert rc == 0, text[-2000:]
    assert "double lv[4];" in text
    assert text.count("rt::load_lval(lv, c, ") == 4
    assert text.count(", lv, ") == 4
    assert "true>(" not in text
    # without the flag: no table
    rc, size, text = lib.jit_check(sch, S.RT_BEAM_GAUSS)
    assert rc == 0 and "lv" not in text


def test_generated_adjoint_kernel_compiles():
    """The adjoint kernel generated for the echelle with every slot of every
    surface asked for (0x200 = RT_JIT_ADJOINT) compiles: one forward and one
    backward step per surface, straight-line."""
    sch = schemes.echelle(1, 4)
    rc, size, text = lib.jit_check(sch, S.RT_BEAM_GAUSS | 0x200)
    assert rc == 0, text[-2000:]
    assert size > 10000
    assert "rt_adjoint_jit" in text
    # the forward records the Newton iterates of the four iterating
    # intersections (tape offsets 7 rows per surface, 8 more where it iterates)
    assert text.count("rt::adj::fwd_step<false, true>(") == sch.n_surf
    assert text.count("rt::adj::bwd_step<true>(") == sch.n_surf
    import re
    offsets = [int(k) for k in re.findall(r"r, alive, c, t, (\d+)", text)]
    assert offsets == [0, 7, 14, 21, 28, 35, 50, 57, 72, 87, 102, 109]
    # reading a per-wavelength table (0x400).  One scheme: the derived
    # table's mu and grating columns (2 + 2 for the gratings, 6 for the lens
    # surfaces; not the detector's flux weight: no flux in the adjoint), read
    # by the forward steps and again by the backward steps
    rc, size, text = lib.jit_check(sch, S.RT_BEAM_GAUSS | 0x200 | 0x400)
    assert rc == 0, text[-2000:]
    assert text.count("rt::adj::load_lval_adj(lv, c, ") == 10
    assert text.count("rt::adj::load_lval_adj(lb, c, ") == 10
    assert text.count("rt::adj::fwd_step<true, true>(") == sch.n_surf
    assert text.count("rt::adj::bwd_step_d<true>(") == sch.n_surf
    # a batch (0x100): the three glass columns in the forward (not the
    # detector QE) and n1/n2 of the six lens surfaces in the backward, where a
    # back surface's n1 is its front surface's n2 column
    rc, size, text = lib.jit_check(sch, S.RT_BEAM_GAUSS | 0x200 | 0x400 | 0x100)
    assert rc == 0, text[-2000:]
    assert text.count("rt::adj::load_lval_adj(lv, c, ") == 3
    assert text.count("rt::adj::load_lval_adj(lb, c, ") == 6


ENGINE_HEADER = os.path.join(ROOT, "include", "rt", "rt_engine_func.h")
ENGINE_LIB = os.path.join(ROOT, "rt_amd", "librt_engine_hip.so")
COMMON_LIB = os.path.join(ROOT, "rt_amd", "librt_common_hip.so")

# The 19 entry points of common-lib/rt_engine_func.h:25-104 as the reference's
# callers link them (C++ linkage: these demangled signatures are the ABI).
PLUGIN_SIGNATURES = [
    "rt_engine_init()",
    "rt_engine_info(std::__cxx11::list<std::__cxx11::basic_string<char, std::char_traits<char>, "
    "std::allocator<char> >, std::allocator<std::__cxx11::basic_string<char, std::char_traits<char>, "
    "std::allocator<char> > > >&)",
    "beam_random(Beam::BeamShape, double*, double*, unsigned long, double*, double*, double*)",
    "beam_conic_cosins(unsigned long, double, double*, double*, double*, double*, double*)",
    "beam_gauss_cosins(unsigned long, double, double*, double*, double*)",
    "uniform_range_beam(unsigned long, double*, double*)",
    "random_range_beam(unsigned long, double*, double*)",
    "beam_transform(unsigned long, double*, double*, double*, double*, double*, double*, double*, double*)",
    "surface_intersection(Surface::SurfaceClass, double*, unsigned long, double*, double*, double*, double*, "
    "double*, double*, unsigned long*, char*)",
    "surface_reflection(Surface::SurfaceClass, double*, unsigned long, double*, double*, double*, double*, "
    "double*, double*)",
    "surface_refraction(Surface::SurfaceClass, double*, unsigned long, double*, double*, double*, double*, "
    "double*, double*, double*, TabulatedFunction&, TabulatedFunction&, unsigned long*, char*)",
    "surface_diffration(Surface::SurfaceClass, Grating::GratingType, double*, long, double, unsigned long, "
    "double*, double*, double*, double*, double*, double*, double*, TabulatedFunction&, TabulatedFunction&, "
    "unsigned long*, char*)",
    "surface_QE(unsigned long, double*, double*, TabulatedFunction&)",
    "grating_energy_distr(unsigned long, double*, long, double, double, double, double, double, double, double*)",
    "rt_engine_free_vector(double*)",
    "rt_engine_alocate_vector(unsigned long)",
    "spline_init(unsigned long, double*, double*, double*)",
    "ccd_coordinates(unsigned long, double*, double*, double, double, double, double, unsigned int*, "
    "unsigned int*)",
    "compute_ccd_image(unsigned long, unsigned int*, unsigned int*, double*, TabulatedFunction&, "
    "TabulatedFunction&, unsigned long, unsigned long, double, double, double*)",
]


def _demangled_exports(path):
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", path], capture_output=True, text=True).stdout
    return {line.split(" T ", 1)[1] for line in out.splitlines() if " T " in line}


def test_plugin_library_exports_the_19_entry_points():
    assert len(PLUGIN_SIGNATURES) == 19
    header = open(ENGINE_HEADER).read()
    declared = re.findall(r"^(?:RT_engine_error|void|real_t \*)\s*(\w+)\(", header, re.M)
    assert sorted(declared) == sorted(s.split("(")[0] for s in PLUGIN_SIGNATURES)
    exported = _demangled_exports(ENGINE_LIB)
    missing = [s for s in PLUGIN_SIGNATURES if s not in exported]
    assert not missing, f"not exported with the reference signature: {missing}"
    # the engine must not define TabulatedFunction: its callers bring their own
    assert not any(s.startswith("TabulatedFunction::") for s in exported)
    common = _demangled_exports(COMMON_LIB)
    assert "TabulatedFunction::operator[](double const&)" in common


def test_plugin_library_loads_without_a_device():
    # the engine takes TabulatedFunction::operator[] from its caller: here librt_common_hip.so
    C.CDLL(COMMON_LIB)
    L = C.CDLL(ENGINE_LIB)
    sym = "_Z11spline_initmPdS_S_"                  # spline_init(size_t, real_t*, real_t*, real_t*)
    f = getattr(L, sym)
    f.restype = C.c_int
    f.argtypes = [C.c_size_t] + [C.POINTER(C.c_double)] * 3
    x = np.linspace(0.0, 3.0, 7)
    y = np.sin(x)
    y2 = np.zeros(7)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    assert f(7, dp(x), dp(y), dp(y2)) == 0
    assert np.array_equal(y2, S.natural_spline_y2(x, y))


def test_stage_intersect_accepts_scheme_records():
    """A record taken from a packed scheme refers to the scheme's spline pool
    (refractive indices, QE); the intersection stage checks only its geometry."""
    L = lib.lib()
    sch = schemes.echelle(1, 4)
    t = np.asarray(sch.table).reshape(sch.n_surf, S.SURF_STRIDE)
    dummy = np.zeros(8)
    flag = np.ones(8, dtype=np.int8)
    for s in range(sch.n_surf):
        rec = np.ascontiguousarray(t[s])
        rc = L.rt_hip_stage_intersect(rec.ctypes.data_as(C.POINTER(C.c_double)), 0,
                                      *[dummy.ctypes.data] * 6, flag.ctypes.data, None, None, None)
        assert rc == 0, (s, lib.last_error())
    bad = np.ascontiguousarray(t[6]).copy()
    bad[S.S_NPOLY] = 11
    rc = L.rt_hip_stage_intersect(bad.ctypes.data_as(C.POINTER(C.c_double)), 0, *[dummy.ctypes.data] * 6,
                                  flag.ctypes.data, None, None, None)
    assert rc == 4 and "aspheric" in lib.last_error()


def test_batched_kernel_compiles_and_batches_are_validated():
    rc, size, text = lib.jit_check(schemes.echelle(1, 4), S.RT_BEAM_GAUSS | 0x100)
    assert rc == 0, text[-2000:]
    assert "trace_batch_body" in text
    L = lib.lib()
    a = schemes.echelle(1, 4)
    b = schemes.czerny_turner(1, 5)
    b.table = np.resize(b.table, len(a.table))           # same size, other structure
    tables = np.ascontiguousarray(np.stack([a.table, b.table]))
    counts = np.array([10, 10], dtype=np.uint64)
    h = C.c_void_p()
    rc = L.rt_hip_batch_create(tables.ctypes.data_as(C.POINTER(C.c_double)), 2, a.n_surf,
                               np.ascontiguousarray(a.pool).ctypes.data_as(C.POINTER(C.c_double)), len(a.pool),
                               counts.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(h))
    assert rc == 4 and "structure" in lib.last_error()


SCHEME_LIB = os.path.join(ROOT, "rt_amd", "librt_scheme_hip.so")

# The public interface of common-lib/scheme.h:21-66 and the reader of
# beam.h:18-19, as callers such as ray-tracing/main.cpp, ccd/ccd_frame.cpp and
# efficiency/losses_comp.cpp link them (C++ linkage).
SCHEME_SIGNATURES = [
    "Scheme::Scheme()",
    "Scheme::Scheme(std::__cxx11::basic_string<char, std::char_traits<char>, std::allocator<char> > const&)",
    "Scheme::~Scheme()",
    "Scheme::load_from_file(std::__cxx11::basic_string<char, std::char_traits<char>, std::allocator<char> > "
    "const&, bool)",
    "Scheme::destroy()",
    "Scheme::run(std::__cxx11::list<std::__cxx11::basic_string<char, std::char_traits<char>, std::allocator<char> >, "
    "std::allocator<std::__cxx11::basic_string<char, std::char_traits<char>, std::allocator<char> > > >&, "
    "std::ostream&, unsigned long)",
    "Scheme::GetNumberOfSurfaces() const",
    "Scheme::ComputeQE(unsigned long, char*)",
    "Scheme::bad_scheme::bad_scheme(Scheme::SchemeError)",
    "Scheme::bad_scheme::scheme_error() const",
    "Beam::bad_beam::bad_beam(Beam::BeamError)",
    "Surface::bad_surface::bad_surface(Surface::SurfaceError)",
    "load_beam_data(char const*, unsigned long*, unsigned long*, unsigned long*, double**, double**, double**, "
    "double**, double**, double**, double**, bool)",
]


def test_scheme_library_exports_the_scheme_api():
    exported = _demangled_exports(SCHEME_LIB)
    missing = [s for s in SCHEME_SIGNATURES if s not in exported]
    assert not missing, f"not exported with the reference signature: {missing}"
    for header in ("scheme.h", "ccd_frame.h"):
        text = open(os.path.join(ROOT, "include", "rt", header)).read()
        names = re.findall(r"^int (rt_\w+)\(", text, re.M)
        assert names
        for n in names:
            assert n in exported, n
    # the reference's enumerator values (scheme.h:26-28, beam.h:35, surface.h:29-32)
    text = open(os.path.join(ROOT, "include", "rt", "scheme.h")).read()
    assert "FileOpenFailure, ParseFailure, NoClassAttr, BadAttrValue, NanAttrValue, MemAllocFailure" in text
