"""GPU parity of generated partition code: each case of
tests/codegen_cases.py (a grammar's operator graph under a partition,
written as C++ by the library, compiled against the C++ front end and
lowered) is derived by the HIP engine and checked against the oracle on
the same lowered program: bit-exact terminal multisets, transforms within
1e-5 of the FP64 twin."""
import os
import sys

import pytest

from operatorgraph_amd import abi
from operatorgraph_amd.compiler import compile_grammar
from tests import codegen_cases as CC
from tests.helpers import assert_same_multiset, canon_order, rel_err
from tests.test_codegen import case_axioms

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", CC.cases(), ids=lambda c: c[0])
def test_partition_code_on_the_gpu(case):
    from oracle import oracle as O
    from operatorgraph_amd import engine
    name, text, shape, spec, kind = case
    print(f"partition code {name} ({spec})", flush=True)
    sys.stdout.flush()
    if not os.path.exists(CC.SO):
        pytest.fail("tests/cpp/_build/libgen_partitions.so is not built (run __graft_entry__.build())")
    prog = CC.LoweredProgram(name)
    ax = case_axioms(compile_grammar(text, name), kind)
    try:
        r = O.derive(prog, ax, twins=O.TWIN_F64)
    except O.OracleError:
        pytest.skip("the oracle stops on these sizes (covered by the CPU test)")
    eng = engine.Engine(technique=abi.MEGAKERNEL)
    try:
        eng.load(prog).set_axioms(ax)
        for t in range(3):
            eng.bind(t, max(r.count(t), 1) + 64)
        eng.execute()
        for t in range(3):
            got, ref = eng.instances(t), r.instances(t)
            assert_same_multiset(got, ref, f"{name} terminal {t}")
            if len(ref):
                f64 = r.instances_f64(t)[:, :12]
                assert rel_err(got[canon_order(got)]["model"], f64[canon_order(ref)]) < 1e-5
    finally:
        eng.close()
