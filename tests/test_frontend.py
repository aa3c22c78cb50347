"""Front end (csrc/frontend.cpp) against DepthTree fixtures written by the
reference's own pareas-ast front end (tests/golden/make_golden.py), and the
generator's node accounting."""
import glob
import os

import numpy as np
import pytest

import oracle
from ptilopsis_amd import parse_source, workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "trees", "*.npz")))
FIELDS = ("node_types", "result_types", "parents", "depth", "child_idx", "node_data")


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_front_end_matches_reference_fixture(path):
    z = np.load(path)
    t = parse_source(z["source"].tobytes())
    for f in FIELDS:
        assert np.array_equal(getattr(t, f), z[f]), f
    assert t.max_depth == int(z["max_depth"][0])


def test_fixture_count():
    assert len(FIXTURES) >= 12


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_generator_node_count_is_exact(seed):
    src, n = W.program_source(seed, 10)
    assert parse_source(src).n == n
    src, n = W.program_source(seed, 3, base=W.STRESS)
    assert parse_source(src).n == n


def test_live_reference_front_end(tmp_path):
    if not os.path.exists(oracle.REF_DUMP):
        pytest.skip("reference front end not built here")
    src, _ = W.program_source(77, 20)
    p = tmp_path / "x.ptl"
    p.write_text(src)
    ref = oracle.ref_front_end(str(p), str(tmp_path / "d.bin"))
    t = parse_source(src)
    for f in FIELDS:
        assert np.array_equal(getattr(t, f), getattr(ref, f)), f


@pytest.mark.parametrize("src,msg", [
    ("function f() : int { x = 1; }", "undeclared"),
    ("function f() : int { x : int = 1.5; }", "Mismatched"),
    ("function f() : int { g(); }", "undeclared function"),
    ("function f() : int { x : int; x : int; }", "Redeclaration"),
    ("function f() : int { 1 = 2; }", "Mismatched"),
])
def test_front_end_errors(src, msg):
    with pytest.raises(ValueError, match=msg):
        parse_source(src)


def test_float_keyword_is_a_literal_like_the_reference():
    # token.hpp: float literals and the `float` keyword share one token type
    a = parse_source("function f() : float { return float; }")
    b = parse_source("function f() : float { return 0.0; }")
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f))
