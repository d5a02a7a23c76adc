"""Static check of the Python sources: every name a function reads must be a
local, a closure variable, a module global or a builtin (catches NameErrors in
GPU-only code paths on the CPU)."""

import builtins
import symtable
from pathlib import Path

from conftest import ROOT

SOURCES = sorted((ROOT / "cusmic_amd").glob("*.py")) + sorted((ROOT / "oracle").glob("*.py")) + [
    ROOT / "bench.py", ROOT / "__graft_entry__.py"]


def undefined_names(path):
    text = Path(path).read_text()
    top = symtable.symtable(text, str(path), "exec")
    module_names = set(top.get_identifiers())
    bad = []

    def walk(table):
        for child in table.get_children():
            for sym in child.get_symbols():
                name = sym.get_name()
                if sym.is_global() and sym.is_referenced() and name not in module_names \
                        and not hasattr(builtins, name):
                    bad.append(f"{path.name}:{child.get_name()}:{name}")
            walk(child)
    walk(top)
    return bad


def test_no_undefined_names():
    problems = [p for src in SOURCES for p in undefined_names(src)]
    assert not problems, problems
