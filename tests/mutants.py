"""Planted-mistake builds of the CPU oracle (tests only, CPU only).

A pin of an oracle function is only worth something if a plausible mistake in
that function fails it.  ``build_mutant`` copies oracle/fvm_oracle.c into a
temporary directory, applies text replacements inside one named function (each
must match exactly once there, so a stale replacement fails loudly instead of
silently testing the unmodified oracle), and compiles it with the oracle's own
flags.  The tests then run their pins against the mutant and require a failure.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "fvm_oracle.c")
CFLAGS = ["-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-fvisibility=hidden",
          "-Wno-unknown-pragmas", "-Wl,-Bsymbolic", "-DSREAL=double"]

_TMP = tempfile.mkdtemp(prefix="oracle_mutants_")
_built = {}


def _function_span(src: str, name: str) -> tuple[int, int]:
    """[start, end) of the definition of ``name`` (from its signature to the
    next top-level definition)."""
    for sig in (f"API void {name}(", f"API double {name}(", f"static inline real {name}(",
                f"static void {name}(", f"static inline real {name} ("):
        start = src.find(sig)
        if start >= 0:
            break
    else:
        raise KeyError(f"{name} not found in the oracle")
    end = src.find("\nAPI ", start + 1)
    nxt = src.find("\nstatic ", start + 1)
    ends = [e for e in (end, nxt) if e >= 0]
    return start, (min(ends) if ends else len(src))


def build_mutant(tag: str, function: str, replacements: list[tuple[str, str]]) -> str:
    """Path of a library built from the oracle with ``replacements`` applied in
    ``function``."""
    if tag in _built:
        return _built[tag]
    src = open(SRC).read()
    a, b = _function_span(src, function)
    body = src[a:b]
    for old, new in replacements:
        n = body.count(old)
        if n != 1:
            raise AssertionError(f"mutant {tag}: {old!r} occurs {n} times in {function}")
        body = body.replace(old, new)
    path_c = os.path.join(_TMP, f"{tag}.c")
    path_so = os.path.join(_TMP, f"lib_{tag}.so")
    with open(path_c, "w") as f:
        f.write(src[:a] + body + src[b:])
    subprocess.run(["gcc", *CFLAGS, "-o", path_so, path_c, "-lm"], check=True)
    _built[tag] = path_so
    return path_so


def fails(check, *args) -> bool:
    """True when ``check(*args)`` raises AssertionError (the pin caught the mutant)."""
    try:
        check(*args)
    except AssertionError:
        return True
    return False
