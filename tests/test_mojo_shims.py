"""The Mojo drop-in shims (mojo/src/level3, mojo/src/level2) against the
C-ABI they call and the reference signatures they keep.

No Mojo toolchain is available here, so the shims are checked as text:
* every external_call names a symbol declared in include/mojoblas_hip.h
  (and exported by the built library) with as many arguments as the C
  prototype has parameters;
* each shim's public function takes the reference's parameters, in order
  (the lists below restate src/level3/gemm.mojo:56-78, trsm.mojo:20-38,
  trmm.mojo:20-38, syrk.mojo:20-37, syr2k.mojo:20-37, symm.mojo:20-41,
  src/level2/gemv.mojo:23-43, trsv.mojo:20-35), and covers float32 and
  float64 through the s / d entries;
* the level3 package exports the reference's names (src/level3/__init__.mojo:29-44).
"""
import os
import re

import pytest

from mojoblas_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOJO = os.path.join(ROOT, "mojo", "src")

REFERENCE_PARAMS = {
    "level3/gemm.mojo": ("gemm", ["trans_a", "trans_b", "m", "n", "k", "alpha", "a", "lda", "b", "ldb", "beta", "c",
                                  "ldc"]),
    "level3/trsm.mojo": ("trsm", ["side", "uplo", "trans_a", "diag", "m", "n", "alpha", "a", "lda", "b", "ldb"]),
    "level3/trmm.mojo": ("trmm", ["side", "uplo", "trans_a", "diag", "m", "n", "alpha", "a", "lda", "b", "ldb"]),
    "level3/syrk.mojo": ("syrk", ["uplo", "trans", "n", "k", "alpha", "a", "lda", "beta", "c", "ldc"]),
    "level3/syr2k.mojo": ("syr2k", ["uplo", "trans", "n", "k", "alpha", "a", "lda", "b", "ldb", "beta", "c", "ldc"]),
    "level3/symm.mojo": ("symm", ["side", "uplo", "m", "n", "alpha", "a", "lda", "b", "ldb", "beta", "c", "ldc"]),
    "level2/gemv.mojo": ("gemv", ["trans", "m", "n", "alpha", "a", "lda", "x", "incx", "beta", "y", "incy"]),
    "level2/trsv.mojo": ("trsv", ["uplo", "trans", "diag", "n", "a", "lda", "x", "incx"]),
}
REFERENCE_EXPORTS = ["gemm", "gemm_v3", "gemm_v6", "gemm_v7", "gemm_v8", "gemm_v9", "gemm_v10", "gemm_v11",
                     "gemm_v12", "symm", "syrk", "syr2k", "trmm", "trsm"]


def _split_top(s):
    """Split an argument list at top-level commas."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([":
            depth += 1
        elif ch in ")]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return [a.strip() for a in out if a.strip()]


def _balanced(text, start):
    """Contents of the parenthesis opening at text[start]."""
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "(":
            depth += 1
        elif text[i] == ")":
            depth -= 1
            if depth == 0:
                return text[start + 1:i]
    raise AssertionError("unbalanced parenthesis")


def header_prototypes():
    text = open(os.path.join(ROOT, "include", "mojoblas_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(?:int|const char\*)\s+(mbh_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.S):
        params = [p for p in _split_top(m.group(2)) if p != "void"]
        protos[m.group(1)] = len(params)
    return protos


def external_calls(text):
    calls = []
    for m in re.finditer(r'external_call\["(mbh_\w+)",\s*[^\]]+\]\(', text):
        calls.append((m.group(1), len(_split_top(_balanced(text, m.end() - 1)))))
    return calls


SHIM_FILES = sorted(REFERENCE_PARAMS) + ["level3/_hip.mojo"]


@pytest.mark.parametrize("rel", SHIM_FILES)
def test_external_calls_match_the_c_abi(rel):
    protos = header_prototypes()
    L = _lib.lib()
    calls = external_calls(open(os.path.join(MOJO, rel)).read())
    assert calls, rel
    for name, nargs in calls:
        assert name in protos, (rel, name)
        assert nargs == protos[name], (rel, name, nargs, protos[name])
        assert hasattr(L, name)


@pytest.mark.parametrize("rel", sorted(REFERENCE_PARAMS))
def test_shim_keeps_the_reference_signature(rel):
    fname, params = REFERENCE_PARAMS[rel]
    text = open(os.path.join(MOJO, rel)).read()
    m = re.search(r"\ndef %s\[(.*?)\]\((.*?)\):" % fname, text, flags=re.S)
    assert m, (rel, fname)
    got = [a.split(":")[0].strip() for a in _split_top(m.group(2))]
    assert got == params
    comp = _split_top(m.group(1))
    assert comp[-1].startswith("dtype: DType") and "//" in comp
    # float32 and float64 both reach the library, through the s and d entries
    entries = [n for n, _ in external_calls(text)]
    assert any(e.startswith("mbh_s") for e in entries) and any(e.startswith("mbh_d") for e in entries)


def test_level3_package_exports_the_reference_names():
    text = open(os.path.join(MOJO, "level3", "__init__.mojo")).read()
    names = set(re.findall(r"\b(gemm(?:_v\d+)?|symm|syrk|syr2k|trmm|trsm)\b", text))
    assert set(REFERENCE_EXPORTS) <= names
    gemm_text = open(os.path.join(MOJO, "level3", "gemm.mojo")).read()
    for v in REFERENCE_EXPORTS[1:9]:
        assert re.search(r"\bcomptime %s = gemm\b" % v, gemm_text), v
