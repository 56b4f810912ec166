"""bench.py runs only on a GPU, so a misspelt or removed name in its report
code would surface only there.  This CPU check parses it and requires every
name it loads to be bound somewhere in the file (assignment, import,
function, argument, comprehension target) or to be a builtin."""
import ast
import builtins
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _unbound_names(path):
    tree = ast.parse(open(path).read(), filename=path)
    bound = set(dir(builtins)) | {"__file__", "__name__"}
    for node in ast.walk(tree):
        if isinstance(node, (ast.FunctionDef, ast.AsyncFunctionDef, ast.ClassDef)):
            bound.add(node.name)
        elif isinstance(node, ast.Name) and isinstance(node.ctx, (ast.Store, ast.Del)):
            bound.add(node.id)
        elif isinstance(node, (ast.Import, ast.ImportFrom)):
            for a in node.names:
                bound.add((a.asname or a.name).split(".")[0])
        elif isinstance(node, ast.arg):
            bound.add(node.arg)
        elif isinstance(node, ast.ExceptHandler) and node.name:
            bound.add(node.name)
    used = {n.id for n in ast.walk(tree) if isinstance(n, ast.Name) and isinstance(n.ctx, ast.Load)}
    return sorted(used - bound)


def test_bench_names_are_bound():
    assert _unbound_names(os.path.join(ROOT, "bench.py")) == []


def test_the_check_sees_an_unbound_name(tmp_path):
    p = tmp_path / "b.py"
    p.write_text("def main():\n    x = 1\n    return x + kfn\n")
    assert _unbound_names(str(p)) == ["kfn"]
