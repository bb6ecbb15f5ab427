"""A small Mako-compatible template engine for kernel and transformation code.

reikna writes every kernel and transformation as a Mako template rendered by
grunnur (``Template``, ``DefTemplate``, ``Snippet``, ``Module``; e.g.
reikna/core/transformation.py:87-90, reikna/transformations.py:26-29).  Mako is
not available here, so this module implements the subset those templates use:

* ``${expr}`` substitutions (any Python expression, converted with ``str``);
* ``% if/elif/else/endif``, ``% for/endfor`` control lines (``%%`` escapes);
* ``<% python block %>`` and ``## comment`` lines;
* ``<%def name="f(a, b)"> ... </%def>`` definitions, callable from other
  templates and retrievable with ``Template.get_def``;
* a backslash at the end of a line joins it with the next one.

A ``Snippet`` is a def rendered inline when called; a ``Module`` is rendered
once per kernel with a unique prefix and its text is emitted before the code
that references it (``str(module)`` gives the prefix), which is how reikna's
load/store macros and grunnur's ``functions.mul`` are composed.
"""

from __future__ import annotations

import itertools
import re
import textwrap
import threading
from collections.abc import Callable, Mapping, Sequence
from typing import Any

_DEF_OPEN = re.compile(r'<%def\s+name\s*=\s*"([^"]+)"\s*>')
_DEF_CLOSE = "</%def>"


class TemplateError(Exception):
    pass


def _split_expr(text: str, start: int) -> int:
    """Index just past the ``}`` that closes the ``${`` at ``start``."""
    depth = 0
    i = start + 2
    quote = None
    while i < len(text):
        ch = text[i]
        if quote:
            if ch == "\\":
                i += 2
                continue
            if ch == quote:
                quote = None
        elif ch in "'\"":
            quote = ch
        elif ch == "{":
            depth += 1
        elif ch == "}":
            if depth == 0:
                return i + 1
            depth -= 1
        i += 1
    raise TemplateError("unterminated ${...}")


def _compile_body(body: str, name: str) -> str:
    """Python source of a function body that renders `body` into the list `__out`."""
    body = re.sub(r"\\\n", "", body)   # line continuation
    lines: list[str] = []
    indent = 1
    pad = lambda: "    " * indent  # noqa: E731

    def emit_text(text: str) -> None:
        pos = 0
        while True:
            k = text.find("${", pos)
            if k < 0:
                break
            if k > pos:
                lines.append(pad() + f"__out.append({text[pos:k]!r})")
            end = _split_expr(text, k)
            expr = text[k + 2 : end - 1].strip()
            lines.append(pad() + f"__out.append(__str({expr}))")
            pos = end
        if pos < len(text):
            lines.append(pad() + f"__out.append({text[pos:]!r})")

    i = 0
    text_buf: list[str] = []

    def flush() -> None:
        if text_buf:
            emit_text("".join(text_buf))
            text_buf.clear()

    src_lines = body.splitlines(keepends=True)
    while i < len(src_lines):
        line = src_lines[i]
        stripped = line.lstrip()
        if stripped.startswith("##"):
            i += 1
            continue
        if stripped.startswith("%") and not stripped.startswith("%%"):
            flush()
            stmt = stripped[1:].strip()
            word = stmt.split()[0].rstrip(":") if stmt else ""
            if word in ("endif", "endfor", "endwhile"):
                indent -= 1
                lines.append(pad() + "pass")
            elif word in ("elif", "else"):
                indent -= 1
                lines.append(pad() + stmt)
                indent += 1
            else:
                if not stmt.endswith(":"):
                    raise TemplateError(f"control line must end with ':' in {name}: {stmt}")
                lines.append(pad() + stmt)
                indent += 1
            i += 1
            continue
        if stripped.startswith("%%"):
            text_buf.append(line.replace("%%", "%", 1))
            i += 1
            continue
        # python blocks may start mid-line and span lines
        k = line.find("<%")
        if k >= 0 and not line.startswith("<%def", k) and not line.startswith("</%", k):
            rest = "".join(src_lines[i:])
            start = rest.find("<%")
            end = rest.find("%>", start)
            if end < 0:
                raise TemplateError(f"unterminated <% block in {name}")
            text_buf.append(rest[:start])
            flush()
            code = textwrap.dedent(rest[start + 2 : end].strip("\n"))
            for cl in code.splitlines():
                if cl.strip():
                    lines.append(pad() + cl)
            remaining = rest[end + 2 :]
            src_lines = src_lines[:i] + remaining.splitlines(keepends=True)
            # drop the newline right after %> (as Mako does for a block on its own line)
            if src_lines[i:] and src_lines[i].strip() == "":
                src_lines.pop(i)
            continue
        text_buf.append(line)
        i += 1
    flush()
    if indent != 1:
        raise TemplateError(f"unbalanced control lines in {name}")
    return "\n".join(lines) if lines else "    pass"


def _strify(x: Any) -> str:
    return str(x)


class DefTemplate:
    """A renderable template function with named positional parameters."""

    def __init__(self, name: str, parameters: Sequence[str], body: str, owner: Template | None = None):
        self.name = name
        self.parameters = list(parameters)
        self.body = body
        self.owner = owner
        params = ", ".join(["__ctx", *self.parameters])
        src = f"def __render({params}):\n    __out = []\n    __str = __ctx['__str']\n"
        src += "    globals().update(__ctx)\n"
        src += _compile_body(body, name) + "\n    return ''.join(__out)\n"
        self._source = src
        self._code = compile(src, f"<template {name}>", "exec")

    @classmethod
    def from_string(cls, name: str, parameters: Sequence[str], text: str) -> DefTemplate:
        return cls(name, parameters, text)

    def render(self, *args: Any, **globals_: Any) -> str:
        ctx: dict[str, Any] = {}
        if self.owner is not None:
            ctx.update(self.owner.defs_namespace())
        ctx.update(globals_)
        ctx["__str"] = _strify
        ns: dict[str, Any] = {}
        ns.update(ctx)
        exec(self._code, ns)
        if len(args) != len(self.parameters):
            raise TemplateError(
                f"def {self.name} takes {len(self.parameters)} arguments ({self.parameters}), got {len(args)}"
            )
        return ns["__render"](ctx, *args)

    def __repr__(self) -> str:
        return f"DefTemplate({self.name}({', '.join(self.parameters)}))"


class Template:
    """A template holding zero or more <%def>s (grunnur's Template)."""

    def __init__(self, text: str, render_globals: Mapping[str, Any] | None = None):
        self.text = text
        self.render_globals = dict(render_globals or {})
        self.defs: dict[str, DefTemplate] = {}
        body_parts = []
        pos = 0
        while True:
            m = _DEF_OPEN.search(text, pos)
            if not m:
                body_parts.append(text[pos:])
                break
            body_parts.append(text[pos : m.start()])
            # find the matching close, allowing nested defs
            depth, k = 1, m.end()
            while depth:
                nxt_open = _DEF_OPEN.search(text, k)
                nxt_close = text.find(_DEF_CLOSE, k)
                if nxt_close < 0:
                    raise TemplateError("unterminated <%def>")
                if nxt_open and nxt_open.start() < nxt_close:
                    depth += 1
                    k = nxt_open.end()
                else:
                    depth -= 1
                    k = nxt_close + len(_DEF_CLOSE)
            sig = m.group(1)
            fname, _, params = sig.partition("(")
            params = [p.strip() for p in params.rstrip(")").split(",") if p.strip()]
            inner = text[m.end() : k - len(_DEF_CLOSE)]
            self.defs[fname.strip()] = DefTemplate(fname.strip(), params, inner, owner=self)
            pos = k
        self.main = DefTemplate("__main__", [], "".join(body_parts), owner=self)

    @classmethod
    def from_string(cls, text: str, render_globals: Mapping[str, Any] | None = None) -> Template:
        return cls(text, render_globals)

    def get_def(self, name: str) -> DefTemplate:
        return self.defs[name]

    def defs_namespace(self) -> dict[str, Any]:
        ns: dict[str, Any] = dict(self.render_globals)
        for name, d in self.defs.items():
            ns[name] = _BoundDef(d, self.render_globals)
        return ns

    def render(self, **kwds: Any) -> str:
        return self.main.render(**kwds)


class _BoundDef:
    """A def callable from another template: ``${other_def(a, b)}``."""

    def __init__(self, d: DefTemplate, render_globals: Mapping[str, Any]):
        self._d = d
        self._g = render_globals

    def __call__(self, *args: Any) -> str:
        return self._d.render(*args, **self._g)


# ---------------------------------------------------------------------------
# snippets and modules
# ---------------------------------------------------------------------------
_render_state = threading.local()
_prefix_counter = itertools.count()


class Snippet:
    """A def rendered inline where it is called (grunnur's Snippet)."""

    def __init__(self, template_def: DefTemplate, render_globals: Mapping[str, Any] | None = None):
        self.template_def = template_def
        self.render_globals = dict(render_globals or {})

    @classmethod
    def from_callable(cls, func: Callable[..., str], parameters: Sequence[str] | None = None) -> Snippet:
        s = cls.__new__(cls)
        s.template_def = None
        s.render_globals = {}
        s._func = func
        return s

    @classmethod
    def from_string(
        cls, parameters: Sequence[str], code: str, render_globals: Mapping[str, Any] | None = None
    ) -> Snippet:
        return cls(DefTemplate.from_string("_snippet", parameters, code), render_globals)

    def __call__(self, *args: Any) -> str:
        if self.template_def is None:
            return self._func(*args)
        return self.template_def.render(*args, **self.render_globals)


class Module:
    """
    Code rendered once per kernel with a unique ``prefix`` (its first template
    parameter) and emitted before its users; ``str(module)`` is the prefix.
    """

    def __init__(self, template_def: DefTemplate | Callable[..., str], render_globals: Mapping[str, Any] | None = None,
                 args: Sequence[Any] = ()):
        self.template_def = template_def
        self.render_globals = dict(render_globals or {})
        self.args = list(args)
        self._id = next(_prefix_counter)

    @classmethod
    def from_string(cls, code: str, render_globals: Mapping[str, Any] | None = None) -> Module:
        return cls(DefTemplate.from_string("_module", ["prefix"], code), render_globals)

    @classmethod
    def from_callable(cls, func: Callable[..., str], name: str = "_module") -> Module:
        return cls(func)

    def _render(self, prefix: str) -> str:
        if isinstance(self.template_def, DefTemplate):
            return self.template_def.render(prefix, *self.args, **self.render_globals)
        return self.template_def(prefix, *self.args)

    def __str__(self) -> str:
        state = getattr(_render_state, "stack", None)
        if not state:
            raise TemplateError("a Module was referenced outside of a kernel render")
        reg = state[-1]
        key = id(self)
        if key in reg["prefixes"]:
            return reg["prefixes"][key]
        prefix = f"_mod{len(reg['prefixes'])}_{self._id}_"
        reg["prefixes"][key] = prefix
        reg["keep"].append(self)
        text = self._render(prefix)
        reg["sources"].append(text)
        return prefix

    def __format__(self, spec: str) -> str:
        return str(self)


def render_with_modules(render: Callable[[], str]) -> str:
    """Run ``render`` collecting every Module it references; return modules + main text."""
    stack = getattr(_render_state, "stack", None)
    if stack is None:
        stack = _render_state.stack = []
    reg = {"prefixes": {}, "sources": [], "keep": []}
    stack.append(reg)
    try:
        main = render()
    finally:
        stack.pop()
    return "\n".join(reg["sources"]) + "\n" + main
