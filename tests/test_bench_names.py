"""bench.py refers to no undefined name: every name a function reads from the
module scope is bound at module level or is a builtin (a static check, so a
renamed or moved variable fails here on the CPU, not in the GPU run)."""
import builtins
import os
import symtable

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _module_bound(top):
    return {s.get_name() for s in top.get_symbols() if s.is_assigned() or s.is_imported()}


def test_bench_functions_have_no_undefined_globals():
    src = open(os.path.join(ROOT, "bench.py")).read()
    top = symtable.symtable(src, "bench.py", "exec")
    bound = _module_bound(top)
    missing = []

    def walk(table):
        for child in table.get_children():
            if child.get_type() == "function":
                for sym in child.get_symbols():
                    name = sym.get_name()
                    if sym.is_global() and sym.is_referenced() and name not in bound and not hasattr(builtins, name):
                        missing.append(f"{child.get_name()}: {name}")
            walk(child)

    walk(top)
    assert not missing, missing


def test_check_catches_an_undefined_name():
    src = "def f():\n    return undefined_thing\n"
    top = symtable.symtable(src, "x.py", "exec")
    f = top.get_children()[0]
    sym = f.lookup("undefined_thing")
    assert sym.is_global() and sym.get_name() not in _module_bound(top)
