"""The product does not run the oracle.

oracle/ is test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg may import or load it.  These tests fail if any module of the
package (Python or native source) imports or loads anything under oracle/, or
if bench.py touches the oracle outside its cpu_baseline function.
"""
from __future__ import annotations

import ast
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "water-sim-cuda_amd"


def _oracle_refs(tree: ast.AST):
    """(line, what) of every import of `oracle` and every string that names
    the oracle's files."""
    out = []
    for node in ast.walk(tree):
        if isinstance(node, ast.Import):
            for a in node.names:
                if a.name.split(".")[0] == "oracle":
                    out.append((node.lineno, f"import {a.name}"))
        elif isinstance(node, ast.ImportFrom):
            if (node.module or "").split(".")[0] == "oracle":
                out.append((node.lineno, f"from {node.module} import ..."))
        elif isinstance(node, ast.Constant) and isinstance(node.value, str):
            v = node.value
            if "liboracle" in v or "oracle/" in v or v == "oracle" or v.startswith("oracle."):
                out.append((node.lineno, repr(v)))
        elif isinstance(node, ast.Call) and getattr(node.func, "attr", None) == "import_module":
            for a in node.args:
                if isinstance(a, ast.Constant) and str(a.value).split(".")[0] == "oracle":
                    out.append((node.lineno, f"import_module({a.value!r})"))
    return out


def test_package_python_does_not_reference_oracle():
    bad = []
    for f in sorted(PKG.rglob("*.py")):
        for line, what in _oracle_refs(ast.parse(f.read_text(), str(f))):
            bad.append(f"{f.relative_to(ROOT)}:{line}: {what}")
    assert not bad, "product code references the oracle:\n" + "\n".join(bad)


def test_package_native_does_not_include_oracle():
    bad = []
    for f in sorted(list(PKG.rglob("*.hip")) + list(PKG.rglob("*.cpp")) + list(PKG.rglob("*.hpp"))
                    + list((ROOT / "include").glob("*.h*"))):
        for i, line in enumerate(f.read_text().splitlines(), 1):
            s = line.strip()
            if s.startswith("#include") and "oracle" in s:
                bad.append(f"{f.relative_to(ROOT)}:{i}: {s}")
            if "liboracle" in s or "dlopen" in s:
                bad.append(f"{f.relative_to(ROOT)}:{i}: {s}")
    assert not bad, "native product code includes or loads the oracle:\n" + "\n".join(bad)


def test_build_links_no_oracle():
    text = (PKG / "build.py").read_text()
    assert "oracle" not in text


def test_bench_uses_oracle_only_in_cpu_baseline():
    tree = ast.parse((ROOT / "bench.py").read_text())
    allowed = set()
    for node in ast.walk(tree):
        if isinstance(node, ast.FunctionDef) and node.name == "cpu_baseline":
            allowed |= {id(n) for n in ast.walk(node)}
    assert allowed, "bench.py has no cpu_baseline function"
    bad, inside = [], 0
    for node in ast.walk(tree):
        if not isinstance(node, (ast.Import, ast.ImportFrom, ast.Constant, ast.Call)):
            continue
        refs = _oracle_refs(ast.Module(body=[ast.Expr(node)] if not isinstance(node, (ast.Import, ast.ImportFrom))
                                       else [node], type_ignores=[]))
        if not refs:
            continue
        if id(node) in allowed:
            inside += 1
        else:
            bad += [f"bench.py:{line}: {what}" for line, what in refs]
    assert not bad, "bench.py touches the oracle outside cpu_baseline:\n" + "\n".join(bad)
    assert inside >= 1, "cpu_baseline no longer uses the oracle"


def test_importing_package_loads_no_oracle():
    """Importing every package module (without a GPU) pulls in no oracle module."""
    code = (
        "import importlib, sys\n"
        f"sys.path.insert(0, {str(ROOT)!r})\n"
        "for m in ['water-sim-cuda_amd', 'water-sim-cuda_amd.learned', 'water-sim-cuda_amd.autotune',\n"
        "          'water-sim-cuda_amd.scenes', 'water-sim-cuda_amd.train_corrector']:\n"
        "    importlib.import_module(m)\n"
        "bad = [m for m in sys.modules if m == 'oracle' or m.startswith('oracle.')]\n"
        "assert not bad, bad\n"
        "print('ok')\n"
    )
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd="/")
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
