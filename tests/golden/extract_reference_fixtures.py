"""Write ``reference_fixtures.json``: the numeric golden vectors the crate's own tests hold.

Run in the development container, where the reference sources are mounted read-only:

    python tests/golden/extract_reference_fixtures.py /root/reference

Only numbers are taken (inputs, expected outputs, seeds, shape and parameter values);
no source text is stored.  What is extracted, and from where:

* ``dgemm_{tn,nt,tt}``: the 3x4x5 CBLAS-captured products at (alpha, beta) = (1, 0) and
  (2.5, -1.5) -- src/blas.rs:1688-1997 (compared at 1e-13 relative, src/blas.rs:1669-1676);
* ``dgemm_tn_exact``: the small-integer case that must match bit for bit -- src/blas.rs:2008-2022;
* ``larnv``: LAPACK ``libtmglib`` stream values for seed [1988,1989,1990,1991], n = 200
  (first twelve, last two, seed after) -- src/lapack/families.rs:1182-1258;
* ``dlatms``: the twelve ``DLATMS`` matrices with the seed after each call --
  src/lapack/families.rs:2013-2470 (compared at 1e-13 relative, seeds exactly);
* ``zgemm_{tn,nt,tt}``: the 2x3x4 CBLAS-captured complex products at alpha = 2 - i,
  beta = 0.5 + 0.25i -- src/blas.rs:2029-2195 (compared at 1e-13 relative to max(|want|, 1),
  src/blas.rs:2386-2389); complex numbers are stored as [re, im] pairs;
* ``zlapack``: the complex LU/Cholesky test matrices -- ``spd3`` and ``general3``
  (src/lapack/complex.rs:422-451), the ``zgesv`` known solution (:483-494), the exactly
  singular 2x2 (:500-508), the non-positive-definite 2x2 (:577-581) and the 3x2 ``zgetrf_rect``
  panel (:598-608).
"""
from __future__ import annotations

import json
import os
import re
import sys

NUM = r"[-+]?(?:\d+\.?\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?)"


def numbers(text: str) -> list[float]:
    # strip line comments before scanning for numbers
    text = re.sub(r"//[^\n]*", "", text)
    return [float(x) for x in re.findall(NUM, text)]


def array_after(src: str, start: int, name: str) -> tuple[list[float], int]:
    """Numbers of the Rust slice literal ``let <name>: ... = [ ... ];`` after ``start``."""
    m = re.compile(r"let\s+" + re.escape(name) + r"\s*:[^=]*=\s*&?\[").search(src, start)
    if m is None:
        raise KeyError(name)
    depth, i = 1, m.end()
    while depth:
        depth += {"[": 1, "]": -1}.get(src[i], 0)
        i += 1
    return numbers(src[m.end() : i - 1]), i


def fn_body(src: str, fn_name: str) -> tuple[int, int]:
    s = src.index("fn " + fn_name + "(")
    depth, i = 0, src.index("{", s)
    start = i
    while True:
        c = src[i]
        if c == "{":
            depth += 1
        elif c == "}":
            depth -= 1
            if depth == 0:
                return start, i
        i += 1


def gemm_fixture(blas: str, fn_name: str) -> dict:
    s, e = fn_body(blas, fn_name)
    body = blas[s:e]
    out = {}
    pos = 0
    for nm in ("a", "b", "c0", "want", "want_ab"):
        vals, pos = array_after(body, pos, nm)
        out[nm] = vals
    mnk = re.search(r"let \(m, n, k\) = \((\d+)usize, (\d+)usize, (\d+)usize\)", body)
    out["m"], out["n"], out["k"] = (int(x) for x in mnk.groups())
    out["alpha_beta"] = [[1.0, 0.0], [2.5, -1.5]]
    return out


def exact_fixture(blas: str) -> dict:
    s, e = fn_body(blas, "dgemm_tn_matches_cblas_exactly_on_integers")
    body = blas[s:e]
    out = {}
    pos = 0
    for nm in ("a", "b", "want"):
        vals, pos = array_after(body, pos, nm)
        out[nm] = vals
    mnk = re.search(r"let \(m, n, k\) = \((\d+)usize, (\d+)usize, (\d+)usize\)", body)
    out["m"], out["n"], out["k"] = (int(x) for x in mnk.groups())
    return out


def larnv_fixture(fam: str) -> dict:
    s, e = fn_body(fam, "larnv_matches_the_reference_stream")
    body = fam[s:e]
    cases = []
    for m in re.finditer(r"Dist::(\w+),\s*\[(.*?)\],\s*\[(\d+),\s*(\d+),\s*(\d+),\s*(\d+)\]", body, re.S):
        dist = m.group(1)
        vals = numbers(m.group(2))
        assert len(vals) == 14, (dist, len(vals))
        cases.append({"dist": dist, "first12": vals[:12], "last2": vals[12:],
                      "seed_after": [int(m.group(k)) for k in range(3, 7)]})
    return {"seed": [1988, 1989, 1990, 1991], "n": 200, "cases": cases}


def dlatms_fixture(fam: str) -> dict:
    s, e = fn_body(fam, "dlatms_matches_the_reference_matrices")
    body = fam[s:e]
    arrays = {}
    for m in re.finditer(r"let\s+([ad]\d+)\s*:\s*\[f64;\s*(\d+)\]\s*=\s*\[", body):
        vals, _ = array_after(body, m.start(), m.group(1))
        assert len(vals) == int(m.group(2)), (m.group(1), len(vals))
        arrays[m.group(1)] = vals
    case_re = re.compile(
        r"Dlatms\s*\{\s*dist:\s*Dist::(\w+),\s*sym:\s*DlatmsSym::(\w+),\s*mode:\s*(-?\d+),\s*"
        r"cond:\s*(" + NUM + r"),\s*dmax:\s*(" + NUM + r"),\s*pack:\s*DlatmsPack::(\w+),?\s*\},\s*"
        r"(\d+),\s*(\d+),\s*\[(\d+),\s*(\d+),\s*(\d+),\s*(\d+)\],\s*&(d\d+),\s*&(a\d+)",
        re.S,
    )
    cases = []
    for m in case_re.finditer(body):
        g = m.groups()
        cases.append({
            "dist": g[0], "sym": g[1], "mode": int(g[2]), "cond": float(g[3]), "dmax": float(g[4]),
            "pack": g[5], "m": int(g[6]), "n": int(g[7]), "seed_after": [int(x) for x in g[8:12]],
            "d": arrays[g[12]], "a": arrays[g[13]],
        })
    assert len(cases) == 12, len(cases)
    return {"seed": [1988, 1989, 1990, 1991], "cases": cases}


CPX = re.compile(r"cpx!\(\s*(" + NUM + r")\s*,\s*(" + NUM + r")\s*\)")


def cpx_list(text: str) -> list[list[float]]:
    text = re.sub(r"//[^\n]*", "", text)
    return [[float(a), float(b)] for a, b in CPX.findall(text)]


def bracket_after(src: str, start: int, pattern: str) -> tuple[str, int]:
    """Text inside the first ``[...]`` after the regex ``pattern`` (searched from ``start``)."""
    m = re.compile(pattern).search(src, start)
    if m is None:
        raise KeyError(pattern)
    i = src.index("[", m.end() - 1)
    depth, j = 1, i + 1
    while depth:
        depth += {"[": 1, "]": -1}.get(src[j], 0)
        j += 1
    return src[i + 1 : j - 1], j


def zgemm_fixture(blas: str, fn_name: str) -> dict:
    s, e = fn_body(blas, fn_name)
    body = blas[s:e]
    out = {}
    pos = 0
    for nm in ("a", "b", "c0", "want"):
        txt, pos = bracket_after(body, pos, r"let\s+" + nm + r"\s*:[^=]*=\s*&\[")
        out[nm] = cpx_list(txt)
    mnk = re.search(r"let \(m, n, k\) = \((\d+)usize, (\d+)usize, (\d+)usize\)", body)
    out["m"], out["n"], out["k"] = (int(x) for x in mnk.groups())
    out["alpha"] = cpx_list(re.search(r"let alpha = (cpx!\([^)]*\))", body).group(1))[0]
    out["beta"] = cpx_list(re.search(r"let beta = (cpx!\([^)]*\))", body).group(1))[0]
    return out


def zlapack_fixture(cx: str) -> dict:
    out = {}
    for nm in ("spd3", "general3"):
        s, e = fn_body(cx, nm)
        out[nm] = cpx_list(cx[s:e])
        assert len(out[nm]) == 9, nm
    s, e = fn_body(cx, "zgesv_solves_a_system")
    txt, _ = bracket_after(cx[s:e], 0, r"let x_true = \[")
    out["zgesv_x_true"] = cpx_list(txt)
    s, e = fn_body(cx, "zgetrf_detects_a_singular_matrix")
    txt, _ = bracket_after(cx[s:e], 0, r"let mut a = vec!\[")
    out["singular2"] = cpx_list(txt)
    s, e = fn_body(cx, "zpotrf_rejects_a_non_positive_definite_matrix")
    txt, _ = bracket_after(cx[s:e], 0, r"let mut a = vec!\[")
    out["not_pd2"] = cpx_list(txt)
    s, e = fn_body(cx, "zgetrf_rect_reconstructs_a_tall_panel")
    txt, _ = bracket_after(cx[s:e], 0, r"let a0 = vec!\[")
    out["rect3x2"] = cpx_list(txt)
    shp = re.search(r"let \(nrows, ncols\) = \((\d+), (\d+)\)", cx[s:e])
    out["rect_shape"] = [int(shp.group(1)), int(shp.group(2))]
    for key, ln in (("zgesv_x_true", 3), ("singular2", 4), ("not_pd2", 4), ("rect3x2", 6)):
        assert len(out[key]) == ln, (key, len(out[key]))
    return out


def main(root: str) -> None:
    blas = open(os.path.join(root, "src", "blas.rs")).read()
    fam = open(os.path.join(root, "src", "lapack", "families.rs")).read()
    cx = open(os.path.join(root, "src", "lapack", "complex.rs")).read()
    out = {
        "provenance": "numeric test vectors of cpmech/algebra_linear; see the docstring of "
                      "tests/golden/extract_reference_fixtures.py for file:line origins",
        "dgemm_tn": gemm_fixture(blas, "dgemm_tn_matches_cblas"),
        "dgemm_nt": gemm_fixture(blas, "dgemm_nt_matches_cblas"),
        "dgemm_tt": gemm_fixture(blas, "dgemm_tt_matches_cblas"),
        "dgemm_tn_exact": exact_fixture(blas),
        "larnv": larnv_fixture(fam),
        "dlatms": dlatms_fixture(fam),
        "zgemm_tn": zgemm_fixture(blas, "zgemm_tn_matches_cblas"),
        "zgemm_nt": zgemm_fixture(blas, "zgemm_nt_matches_cblas"),
        "zgemm_tt": zgemm_fixture(blas, "zgemm_tt_matches_cblas"),
        "zlapack": zlapack_fixture(cx),
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
