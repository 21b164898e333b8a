"""C ABI checks that need no GPU: every function declared in include/rslab_amd.h
is exported by librslab_amd.so and bound by the Python mirror, the settings
defaults equal SolverSettings::default (multifrontal_ldlt.rs:583-614), and the
host-side validation reports the reference's error kinds."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import rslab_amd as R
from rslab_amd import matgen
from rslab_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rslab_amd.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(rsl_[a-z_0-9]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("rsl_analyze", "rsl_factor", "rsl_refactor", "rsl_solve", "rsl_solve_device",
                 "rsl_factor_download", "rsl_refine", "rsl_free_factor", "rsl_free_symbolic"):
        assert must in names


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    # and the ctypes mirror binds each of them
    assert set(declared_functions()) <= set(_lib.SIGNATURES)
    lib = C.CDLL(_lib.LIB_PATH)
    for n in declared_functions():
        assert getattr(lib, n) is not None


def test_settings_defaults_match_reference():
    s = _lib.Settings()
    _lib.lib.rsl_settings_default(C.byref(s))
    assert (s.ordering, s.nemin, s.relax, s.relax_width, s.relax_rows) == (0, 16, 1, 256, 64)
    assert (s.panel_nb, s.pivot, s.scaling, s.precision, s.nd_seed) == (64, 0, 1, 0, 1)
    py = R.SolverSettings().to_c()
    for f, _ in _lib.Settings._fields_:
        assert getattr(py, f) == getattr(s, f), f


def lower(n, cols, rows):
    return R.CscMatrix(n, np.asarray(cols), np.asarray(rows), np.ones(len(rows)))


@pytest.mark.parametrize("case,err", [
    ("upper", R.InvalidInput),          # entry above the diagonal
    ("unsorted", R.InvalidInput),       # rows not strictly increasing
    ("oob", R.InvalidInput),            # row index >= n
    ("colptr", R.InvalidInput),         # decreasing column pointers
    ("panel_nb", R.InvalidInput),       # panel width beyond the device limit
])
def test_analyze_rejects_invalid_input(case, err):
    st = R.SolverSettings()
    if case == "upper":
        a = lower(3, [0, 1, 3, 4], [0, 0, 1, 2])
    elif case == "unsorted":
        a = lower(3, [0, 2, 3, 4], [1, 0, 1, 2])
    elif case == "oob":
        a = lower(3, [0, 1, 2, 3], [0, 1, 5])
    elif case == "colptr":
        a = lower(3, [0, 2, 1, 3], [0, 1, 2])
    else:
        a = lower(3, [0, 1, 2, 3], [0, 1, 2])
        st = st.with_panel_nb(256)
    with pytest.raises(err):
        R.LdltSymbolic.analyze_with(a, st)


def test_factor_rejects_pattern_mismatch_before_device_work():
    a = R.CscMatrix.from_triplets(3, [0, 1, 2], [0, 1, 2], [1.0, 2.0, 3.0])
    sym = R.LdltSymbolic.analyze(a)
    b = R.CscMatrix.from_triplets(4, [0, 1, 2, 3], [0, 1, 2, 3], [1.0, 2.0, 3.0, 4.0])
    with pytest.raises(R.InvalidInput):
        sym.factor(b)


def test_factor_method_is_validated_and_fixed_by_the_plan():
    """FactorMethod (multifrontal_ldlt.rs:257-295): left- and right-looking
    plan on the host; Multifrontal and unknown values are InvalidInput; a
    factor asking for the other schedule than its symbolic's is InvalidInput
    before any device work."""
    a = matgen.grid_laplacian((6, 6), shift=-0.3)
    st = R.SolverSettings()
    assert st.method == R.FactorMethod.LeftLooking
    with pytest.raises(R.InvalidInput):
        R.LdltSymbolic.analyze_with(a, st.with_method(R.FactorMethod.Multifrontal))
    s = st.to_c()
    s.method = 7
    h = C.c_void_p()
    cp = np.ascontiguousarray(a.col_ptr, dtype=np.int64)
    ri = np.ascontiguousarray(a.row_idx, dtype=np.int64)
    assert _lib.lib.rsl_analyze(a.n, cp.ctypes.data_as(C.POINTER(C.c_int64)),
                                ri.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(s), C.byref(h)) == 2
    assert not h.value
    rl = st.with_method(R.FactorMethod.RightLooking)
    sym = R.LdltSymbolic.analyze_with(a, rl)
    with pytest.raises(R.InvalidInput, match="method"):
        sym.factor(a, st)
    # a replan switches the schedule without repeating the ordering
    sym_ll = sym.replan(st)
    i_rl, i_ll = sym.info(), sym_ll.info()
    for k in ("n", "nsuper", "panel_entries", "levels", "nnzl_padded", "factor_flops"):
        assert i_rl[k] == i_ll[k], k


def test_from_triplets_sums_duplicates_and_rejects_upper():
    # CscMatrix::from_triplets (csc.rs:41-106)
    a = R.CscMatrix.from_triplets(3, [0, 2, 2, 1], [0, 0, 0, 1], [1.0, 2.0, 3.0, 4.0])
    assert a.col_ptr.tolist() == [0, 2, 3, 3] and a.row_idx.tolist() == [0, 2, 1]
    assert a.values.tolist() == [1.0, 5.0, 4.0]
    with pytest.raises(R.InvalidInput):
        R.CscMatrix.from_triplets(3, [0], [1], [1.0])


def test_rng_matches_reference_xorshift():
    """Rng (src/matgen/mod.rs:63-96): state = (seed ^ 0x9E3779B97F4A7C15) | 1,
    xorshift64* (12, 25, 27, * 0x2545F4914F6CDD1D), uniform = (x >> 11) / 2^53."""
    st = (5 ^ 0x9E3779B97F4A7C15) | 1
    want = []
    M = (1 << 64) - 1
    for _ in range(4):
        st ^= st >> 12
        st ^= (st << 25) & M
        st ^= st >> 27
        v = (st * 0x2545F4914F6CDD1D) & M
        want.append(-2.0 + 5.0 * ((v >> 11) / float(1 << 53)))
    got = matgen.rng_uniform(5, -2.0, 3.0, 4)
    assert np.array_equal(got, np.array(want))


@pytest.mark.parametrize("case", ["oob", "negative", "colptr", "colptr0"])
def test_order_rejects_malformed_pattern(case):
    """rsl_order validates the pattern before the orderings index their work
    arrays by it (an out-of-range row, a negative row, decreasing or offset
    column pointers are InvalidInput, not out-of-bounds writes)."""
    cp, ri = [0, 2, 3, 4], [0, 1, 1, 2]
    if case == "oob":
        ri = [0, 7, 1, 2]
    elif case == "negative":
        ri = [0, -1, 1, 2]
    elif case == "colptr":
        cp = [0, 3, 2, 4]
    else:
        cp = [1, 2, 3, 4]
    for order in (R.amd_order, R.nd_order):
        with pytest.raises(R.InvalidInput):
            order(3, np.array(cp), np.array(ri))
