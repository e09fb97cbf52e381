"""The C ABI: struct layouts of include/pga.h against the ctypes mirrors,
and every function declared there exported by libpga_amd.so (loaded, not
called: no GPU is needed)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from operatorgraph_amd import abi, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pga.h")

PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "pga.h"
#define S(t) printf(#t " %zu\n", sizeof(t));
#define O(t, f) printf(#t "." #f " %zu\n", offsetof(t, f));
int main(void) {
  S(pga_dispatch_parameter) S(pga_dispatch_successor) S(pga_dispatch_entry) S(pga_procedure)
  S(pga_axiom) S(pga_instance) S(pga_config) S(pga_stats) S(pga_tune_result)
  O(pga_dispatch_entry, num_parameters) O(pga_dispatch_entry, parameters)
  O(pga_dispatch_entry, num_successors) O(pga_dispatch_entry, rule_tag_id)
  O(pga_dispatch_entry, successors) O(pga_dispatch_entry, num_edge_indices)
  O(pga_dispatch_entry, edge_indices) O(pga_dispatch_entry, sub_graph_index)
  O(pga_axiom, size) O(pga_axiom, seed) O(pga_axiom, custom) O(pga_axiom, entry_index)
  O(pga_axiom, shape_type) O(pga_instance, custom) O(pga_instance, axiom) O(pga_instance, path)
  O(pga_instance, meta) O(pga_config, stream) O(pga_tune_result, seconds)
  O(pga_tune_result, max_level_children) S(pga_vertex) O(pga_vertex, normal) O(pga_vertex, uv)
  return 0;
}
"""

MIRRORS = {
    "pga_dispatch_parameter": abi.DispatchParameter, "pga_dispatch_successor": abi.DispatchSuccessor,
    "pga_dispatch_entry": abi.DispatchEntry, "pga_procedure": abi.Procedure, "pga_axiom": abi.Axiom,
    "pga_instance": abi.Instance, "pga_config": abi.Config, "pga_stats": abi.Stats,
    "pga_tune_result": abi.TuneResult, "pga_vertex": abi.Vertex,
}


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    for line in out.strip().splitlines():
        name, value = line.split()
        value = int(value)
        if "." in name:
            t, f = name.split(".")
            assert getattr(MIRRORS[t], f).offset == value, name
        else:
            assert C.sizeof(MIRRORS[name]) == value, name
    assert C.sizeof(abi.DispatchEntry) == 1420   # PGA::DispatchTableEntry


def declared_functions():
    text = open(HEADER).read()
    return re.findall(r"PGA_API\s+[\w\s\*]+?\b(pga_\w+)\s*\(", text)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(engine.LIB_PATH):
        from operatorgraph_amd import build
        build.build()
    lib = C.CDLL(engine.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.pga_abi_version() == 1


def test_engine_refuses_without_gpu():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(engine.PGAError):
        engine.Engine()


@pytest.mark.parametrize("name", ["facade", "menger_interpreter", "stochastic", "facade_params"])
def test_specialised_kernel_compiles_without_a_device(name):
    """The program written as constant entry types compiles at run time
    (pga_specialize_check: validation, source, hipRTC for gfx950) on a
    machine without a GPU."""
    from operatorgraph_amd import engine, grammars
    st, log, _ = engine.specialize_check(getattr(grammars, name)(), 512)
    assert st == 0, log


def test_specialised_kernel_check_reports_a_bad_program():
    from operatorgraph_amd import abi, engine, grammars
    from operatorgraph_amd.program import DispatchTable, Entry, Program, Scalar
    bad = Program(grammars.INTERPRETER_PROCEDURES, DispatchTable([
        Entry(grammars.op(abi.SUBDIV, abi.BOX), [Scalar(0), Scalar(1)], [(7, 0)])]))   # no entry 7
    st, log, _ = engine.specialize_check(bad, 512)
    assert st != 0 and log
