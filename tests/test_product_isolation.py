"""The product never reaches the checker.

The library, its Python package and bench.py (outside its CPU-baseline leg)
must not import, load or run anything under oracle/ or tests/.  The oracle is
test infrastructure: a product path that routed through it would make every
parity test compare the oracle with itself.
"""
import ast
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "curvedspacevertexmodel_amd")

# module names and path fragments that only the checker owns
CHECKER_MODULES = {"svmtest", "mlp_oracle", "voronoi_oracle", "oracle", "tests", "conftest"}
CHECKER_FRAGMENTS = ("oracle", "tests", "svmtest", "liboracle", "mlp_oracle", "voronoi_oracle", "golden")


def _package_py():
    for d, _, files in os.walk(PKG):
        for f in files:
            if f.endswith(".py"):
                yield os.path.join(d, f)


def _offences(tree, allowed_functions=()):
    """Imports of checker modules and string constants that name a checker
    path component, outside the functions whose names start with one of
    allowed_functions."""
    bad = []

    def visit(node, allowed):
        if isinstance(node, (ast.FunctionDef, ast.AsyncFunctionDef)) and node.name.startswith(allowed_functions):
            allowed = True
        if not allowed:
            if isinstance(node, ast.Import):
                for a in node.names:
                    if a.name.split(".")[0] in CHECKER_MODULES:
                        bad.append((node.lineno, f"import {a.name}"))
            elif isinstance(node, ast.ImportFrom):
                if (node.module or "").split(".")[0] in CHECKER_MODULES:
                    bad.append((node.lineno, f"from {node.module} import"))
            elif isinstance(node, ast.Constant) and isinstance(node.value, str):
                # a path component, not prose: no spaces, names a checker file or folder
                v = node.value
                if " " not in v and any(re.search(rf"(^|[/\\.]){re.escape(f)}($|[/\\.])", v)
                                        for f in CHECKER_FRAGMENTS):
                    bad.append((node.lineno, f"path {v!r}"))
        for ch in ast.iter_child_nodes(node):
            visit(ch, allowed)

    visit(tree, False)
    return bad


def test_package_sources_do_not_reach_the_checker():
    bad = []
    for p in _package_py():
        tree = ast.parse(open(p).read(), p)
        bad += [(os.path.relpath(p, ROOT), ln, what) for ln, what in _offences(tree)]
    assert not bad, f"product modules reach oracle/ or tests/: {bad}"


def test_bench_uses_the_checker_only_in_its_cpu_baselines():
    p = os.path.join(ROOT, "bench.py")
    bad = _offences(ast.parse(open(p).read(), p), allowed_functions=("cpu_baseline",))
    assert not bad, f"bench.py reaches oracle/ or tests/ outside its cpu_baseline functions: {bad}"


def test_native_sources_do_not_include_the_checker():
    bad = []
    for d in (os.path.join(PKG, "csrc"), os.path.join(ROOT, "include")):
        for f in sorted(os.listdir(d)):
            text = open(os.path.join(d, f)).read()
            for ln, line in enumerate(text.splitlines(), 1):
                if re.match(r"\s*#\s*include", line) and re.search(r"oracle|tests/", line):
                    bad.append((f, ln, line.strip()))
                if re.search(r"dlopen\s*\(", line) and "oracle" in line:
                    bad.append((f, ln, line.strip()))
    assert not bad, bad


def test_importing_the_package_loads_no_checker_module():
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import curvedspacevertexmodel_amd, curvedspacevertexmodel_amd.warmstart, "
            "curvedspacevertexmodel_amd.trajectory, curvedspacevertexmodel_amd.voronoi\n"
            "bad = [m for m in sys.modules if m.split('.')[0] in %r]\n"
            "assert not bad, bad\n") % (ROOT, sorted(CHECKER_MODULES))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd="/")
    assert r.returncode == 0, r.stderr


def test_warm_start_initialiser_matches_the_oracle_statement():
    """The package's He initialiser draws the same weights as the oracle's
    statement of the network (same generator, same order, structural zeros)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mlp_oracle
    from curvedspacevertexmodel_amd import warmstart
    for seed in (1, 7):
        np.testing.assert_array_equal(warmstart.init_parameters(seed), mlp_oracle.init_theta(seed))
    th = warmstart.init_parameters(3)
    assert not th[~mlp_oracle.structural_mask()].any()


def test_the_guard_catches_a_planted_import():
    tree = ast.parse("import os\nimport sys\nsys.path.insert(0, 'oracle')\nfrom mlp_oracle import init_theta\n")
    found = _offences(tree)
    assert any("mlp_oracle" in w for _, w in found) and any("'oracle'" in w for _, w in found)
    assert not _offences(ast.parse("'''uses the oracle as a checker'''\nx = 'warm start oracle test'\n"))
