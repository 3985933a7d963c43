"""Static checks of bench.py that need no GPU: every name a function reads is
bound in that function, an enclosing function, the module, or builtins (the
GPU-only paths are otherwise exercised only at benchmark time)."""
import ast
import builtins
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bound_in(node):
    names = set()
    for n in ast.walk(node):
        if isinstance(n, ast.Name) and isinstance(n.ctx, (ast.Store, ast.Del)):
            names.add(n.id)
        elif isinstance(n, ast.arg):
            names.add(n.arg)
        elif isinstance(n, (ast.Import, ast.ImportFrom)):
            names.update((a.asname or a.name).split(".")[0] for a in n.names)
        elif isinstance(n, (ast.FunctionDef, ast.ClassDef)):
            names.add(n.name)
        elif isinstance(n, ast.ExceptHandler) and n.name:
            names.add(n.name)
    return names


def test_bench_functions_read_only_bound_names():
    tree = ast.parse(open(os.path.join(ROOT, "bench.py")).read())
    module = set(dir(builtins))
    for n in tree.body:
        if isinstance(n, (ast.FunctionDef, ast.ClassDef)):
            module.add(n.name)
        elif isinstance(n, (ast.Assign, ast.AnnAssign, ast.Import, ast.ImportFrom)):
            module |= _bound_in(n)
    bad = []

    def visit(fn, enclosing):
        scope = enclosing | _bound_in(fn)
        for n in ast.walk(fn):
            if isinstance(n, ast.Name) and isinstance(n.ctx, ast.Load) and n.id not in scope and n.id not in module:
                bad.append(f"{fn.name}:{n.lineno} {n.id}")

    for n in tree.body:
        if isinstance(n, ast.FunctionDef):
            visit(n, set())
    assert not bad, bad
