"""The reference's sample scripts against the dnlPython surface (no GPU).

Each script under Samples/ of the reference tree is read as text and parsed
(it is not run).  Variables bound to a ``dnl.<Class>(...)`` call are typed
by that class; every ``dnl.<name>`` chain, every method called and every
attribute read or assigned on such a variable (and on chains like
``model.parallel``) must exist on the corresponding dnlPython object.  An
attribute that a script sets on a class that does not have it would be
silently ignored by Python, so assignments are checked against a constructed
instance.  Scripts that are not valid Python (one of the reference's has an unclosed
parenthesis, Samples/Element/Shear/Elastic/Shear-Elastic.py:111) are
listed, not checked.  Skipped where the reference tree is absent (it is not
shipped with this repository).
"""
import ast
import os

import pytest

import dnlPython as dnl

SAMPLES = "/root/reference/Samples"

pytestmark = pytest.mark.skipif(not os.path.isdir(SAMPLES), reason="reference tree absent")


def _scripts():
    out = []
    for root, _, files in os.walk(SAMPLES):
        for f in sorted(files):
            if f.endswith(".py"):
                path = os.path.join(root, f)
                with open(path, encoding="utf-8", errors="replace") as fh:
                    text = fh.read()
                if "dnlPython" in text:
                    out.append((path, text))
    return sorted(out)


def _instance(cls):
    for args in (("sample",), ()):
        try:
            return cls(*args)
        except TypeError:
            continue
    return None


class _Checker(ast.NodeVisitor):
    def __init__(self, alias):
        self.alias = alias
        self.types = {}      # variable -> dnlPython object (an instance or the class)
        self.missing = set()
        self.checked = 0

    def _resolve(self, node):
        """The dnlPython object an expression stands for, or None."""
        if isinstance(node, ast.Name):
            if node.id == self.alias:
                return dnl
            return self.types.get(node.id)
        if isinstance(node, ast.Attribute):
            base = self._resolve(node.value)
            if base is None:
                return None
            self.checked += 1
            if not hasattr(base, node.attr):
                self.missing.add(f"{getattr(base, '__name__', type(base).__name__)}.{node.attr}")
                return None
            return getattr(base, node.attr)
        if isinstance(node, ast.Call):
            f = self._resolve(node.func)
            if isinstance(f, type) and f.__module__.startswith("dnlPython"):
                return _instance(f) or f
            return None
        return None

    def visit_Assign(self, node):
        self.visit(node.value)
        obj = self._resolve(node.value) if isinstance(node.value, ast.Call) else None
        for t in node.targets:
            if isinstance(t, ast.Name) and obj is not None:
                self.types[t.id] = obj
            elif isinstance(t, ast.Attribute):
                base = self._resolve(t.value)
                if base is not None:
                    self.checked += 1
                    if not hasattr(base, t.attr):
                        self.missing.add(f"{type(base).__name__}.{t.attr} (assigned)")

    def visit_Call(self, node):
        self._resolve(node.func)
        self.generic_visit(node)

    def visit_Attribute(self, node):
        self._resolve(node)
        self.generic_visit(node)


def _alias(tree):
    for n in ast.walk(tree):
        if isinstance(n, ast.Import):
            for a in n.names:
                if a.name == "dnlPython":
                    return a.asname or a.name
    return None


def test_every_sample_script_uses_only_names_dnlpython_has():
    scripts = _scripts()
    assert len(scripts) >= 10
    missing, checked, broken = {}, 0, []
    for path, text in scripts:
        try:
            tree = ast.parse(text, filename=path)
        except SyntaxError as e:   # the reference's own interpreter rejects it too
            broken.append(f"{os.path.relpath(path, SAMPLES)}:{e.lineno}")
            continue
        alias = _alias(tree)
        if alias is None:
            continue
        c = _Checker(alias)
        c.visit(tree)
        checked += c.checked
        if c.missing:
            missing[os.path.relpath(path, SAMPLES)] = sorted(c.missing)
    print(f"{len(scripts)} scripts, {checked} dnlPython references checked; not valid Python: {broken}")
    assert len(broken) <= 2
    assert not missing, missing


def test_the_checker_flags_unknown_names():
    snippet = ("import dnlPython as dnl\n"
               "model = dnl.DynELA('m')\n"
               "model.createNode(1, 0.0, 0.0, 0.0)\n"
               "model.noSuchMethod(3)\n"
               "mat = dnl.Material('s')\n"
               "mat.youngModulus = 1.0\n"
               "mat.noSuchAttribute = 2.0\n"
               "x = dnl.NoSuchClass\n")
    c = _Checker("dnl")
    c.visit(ast.parse(snippet))
    assert c.missing == {"DynELA.noSuchMethod", "Material.noSuchAttribute (assigned)", "dnlPython.NoSuchClass"}
