"""CPU tests of the boundary: the C-ABI library loads and exports every symbol include/hgx.h
declares, the header and the Python mirror agree, and host-side validation works without
touching a device."""
import ctypes as C
import os
import re

import pytest

import hypergraphrewritingengine_amd as hgx
from hypergraphrewritingengine_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "hgx.h")).read()
    return sorted(set(re.findall(r"\b(hgx_[a-z_]+)\s*\(", text)))


def test_header_and_python_mirror_agree():
    assert sorted(hgx.EXPORTED_SYMBOLS) == _declared_symbols()


def test_library_exports_every_declared_symbol():
    lib = hgx.load_library()
    for name in _declared_symbols():
        assert hasattr(lib, name), name


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors against the header itself: sizes and field offsets as the C compiler
    lays them out."""
    import subprocess
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    structs = {"hgx_run_desc": abi.RunDesc, "hgx_sampling": abi.Sampling, "hgx_counts": abi.Counts,
               "hgx_event": abi.Event, "hgx_pair": abi.Pair, "hgx_unfolding": abi.Unfolding}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hgx.h"', 'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", inc, str(src), "-o", str(exe)])
    got = dict(l.split() for l in subprocess.check_output([str(exe)], text=True).splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def test_footprint_estimate_is_monotone():
    lib = hgx.load_library()
    from hypergraphrewritingengine_amd import workloads as W
    keep = []
    small = abi.make_run_desc(W.batched_run(1, cap=1024), keep)
    big = abi.make_run_desc(W.batched_run(1, cap=4096), keep)
    assert 0 < lib.hgx_request_footprint(C.byref(small)) < lib.hgx_request_footprint(C.byref(big))


def test_context_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(hgx.HGXError):
        hgx.Context(0)


def test_rule_encoding_roundtrip():
    r = hgx.make_rule(0).lhs([0, 1]).lhs([0, 2]).rhs([0, 2]).rhs([0, 3]).rhs([1, 3]).rhs([2, 3]).build()
    prog = hgx.encode_rules([r])
    assert prog == bytes([2, 4, 2, 0, 1, 2, 0, 2, 2, 0, 2, 2, 0, 3, 2, 1, 3, 2, 2, 3])
    assert r.new_var_mask == 0b1000
    with pytest.raises(ValueError):
        hgx.rule_from_lists([[0, 40]], [[0]])
