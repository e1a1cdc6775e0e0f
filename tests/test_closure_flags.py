"""CPU checks of the closure's attach flags: the Python gate options map to
the flag bits include/ldg.h defines (the library call is replaced by a
recorder; tests/test_gpu_closure.py checks the library's own validation)."""
import re
import os

import pytest

import ldg_amd
from ldg_amd import closure as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_flags():
    text = open(os.path.join(ROOT, "include", "ldg.h")).read()
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"#define (LDG_CLOSURE_\w+) (\d+)", text)}


def test_python_flags_match_header():
    f = header_flags()
    assert f == {"LDG_CLOSURE_DISSIPATIVE": ldg_amd.LDG_CLOSURE_DISSIPATIVE,
                 "LDG_CLOSURE_ENERGIZING": ldg_amd.LDG_CLOSURE_ENERGIZING,
                 "LDG_CLOSURE_PER_STEP": ldg_amd.LDG_CLOSURE_PER_STEP}
    assert len({v for v in f.values()}) == 3 and all(v & (v - 1) == 0 for v in f.values())


class _Lib:
    def __init__(self):
        self.calls = []

    def ldg_set_closure(self, h, c, bound, flags):
        self.calls.append((bound, flags))
        return ldg_amd.LDG_OK


class _Solver:
    _h = None


@pytest.mark.parametrize("gate,per_step,expect", [
    (False, False, 0),
    (True, False, ldg_amd.LDG_CLOSURE_DISSIPATIVE),
    ("energizing", False, ldg_amd.LDG_CLOSURE_ENERGIZING),
    ("energizing", True, ldg_amd.LDG_CLOSURE_ENERGIZING | ldg_amd.LDG_CLOSURE_PER_STEP),
    (False, True, ldg_amd.LDG_CLOSURE_PER_STEP),
])
def test_attach_passes_the_flags(monkeypatch, gate, per_step, expect):
    lib = _Lib()
    monkeypatch.setattr(C, "load_library", lambda: lib)
    C.attach(_Solver(), None, 1.0, dissipative=gate, per_step=per_step)
    assert lib.calls == [(1.0, expect)]
