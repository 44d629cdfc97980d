"""The C and C++ bindings shown in the integration guide compile against include/.

Every fenced ```c block is compiled with gcc and every ```cpp block with g++
(-fsyntax-only), so the documented replacement of the reference's
Simulation/Engine calls (Simulation.cpp:24-51 against engine.h:14-43) cannot
drift from include/mpm.h and include/mpm_engine.hpp.
"""
from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
DOC = ROOT / "INTEGRATION.md"


def blocks(lang: str) -> list[str]:
    text = DOC.read_text()
    return re.findall(r"^```" + lang + r"\n(.*?)^```$", text, flags=re.S | re.M)


def test_snippets_present():
    assert len(blocks("c")) >= 4
    assert len(blocks("cpp")) >= 1


@pytest.mark.parametrize("lang,compiler,flags", [
    ("c", "gcc", ["-x", "c", "-std=c11"]),
    ("cpp", "g++", ["-x", "c++", "-std=c++17"]),
])
def test_snippets_compile(lang, compiler, flags, tmp_path):
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} not installed")
    for i, code in enumerate(blocks(lang)):
        src = tmp_path / f"snippet{i}.{lang}"
        src.write_text(code)
        r = subprocess.run([compiler, *flags, "-fsyntax-only", "-Wall", "-Werror", "-Wno-unused-variable",
                            "-Wno-unused-but-set-variable", f"-I{ROOT / 'include'}", str(src)],
                           capture_output=True, text=True)
        assert r.returncode == 0, f"integration guide {lang} block {i} does not compile:\n{code}\n{r.stderr}"


def test_parse_file_signature_matches_doc():
    """The documented config call passes the 7 arguments mpm.h declares."""
    header = (ROOT / "include" / "mpm.h").read_text()
    decl = re.search(r"int mpm_config_parse_file\((.*?)\);", header, flags=re.S).group(1)
    nargs = decl.count(",") + 1
    for code in blocks("c") + blocks("cpp"):
        for call in re.findall(r"mpm_config_parse_file\((.*?)\)\s*(?:!=|\|\||\))", code, flags=re.S):
            assert call.count(",") + 1 == nargs, call
