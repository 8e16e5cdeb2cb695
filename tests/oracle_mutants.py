"""Builds of the oracle library with one planted mistake, for the tests that
show a check has teeth: the check must pass on the real library and fail
(raise AssertionError) on every build listed for it.

A mutant is (file, old text, new text[, anchor]): the first copy of `old`
in that source file of oracle/ (the first one after the text `anchor`, when
given, e.g. the function's signature) is replaced by `new`, and the three
sources are compiled with gcc into a private directory.  Mutations are
chosen so that the broken build still reads and writes only inside the
buffers the checks pass (sign, scale, index and triangle mistakes on square
or generously padded operands), so a mutant fails by value, never by
faulting.
"""
import os
import subprocess

import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
SOURCES = ("oracle.c", "baseline.c", "oracle.h")


def build(dest, file, old, new, anchor=None):
    """Compile oracle.c + baseline.c with one substitution; returns the .so path."""
    os.makedirs(dest, exist_ok=True)
    for name in SOURCES:
        with open(os.path.join(ORACLE_DIR, name)) as f:
            text = f.read()
        if name == file:
            start = 0
            if anchor is not None:
                start = text.find(anchor)
                assert start >= 0, "anchor %r not found in %s" % (anchor, file)
            at = text.find(old, start)
            assert at >= 0, "mutation target %r not found in %s after %r" % (old, file, anchor)
            text = text[:at] + new + text[at + len(old):]
        with open(os.path.join(dest, name), "w") as f:
            f.write(text)
    out = os.path.join(dest, "liboracle_mutant.so")
    subprocess.check_call(["gcc", "-O1", "-fPIC", "-std=c11", "-fopenmp", "-shared", "-o", out,
                           os.path.join(dest, "oracle.c"), os.path.join(dest, "baseline.c"), "-lm"],
                          cwd=dest)
    return out


def assert_rejects(tmp_path, check, mutants):
    """check() passes on the real library and fails on every mutant."""
    check()
    for i, (file, old, new, *anchor) in enumerate(mutants):
        path = build(str(tmp_path / ("m%d" % i)), file, old, new, anchor[0] if anchor else None)
        with oracle.swapped(path):
            with pytest.raises(AssertionError):
                check()
