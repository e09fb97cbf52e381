"""GPU parity on the reference's compiled test-runner scenes (the tables
src/test_runner/*.cuh generate, lowered through include/pga/core/Core.h):
both techniques and the kernel specialised to each table derive the
oracle's instances bit for bit (canonical sort; transforms within 1e-5 of
its FP64 twin), at the axiom sizes each scene's AxiomTraits give. The
polygon scenes (triangular right prisms) run the same way."""
import sys

import pytest

from operatorgraph_amd import abi
from tests.helpers import SceneProgram, scene_axioms
from tests.test_engine_gpu import TECHNIQUES, check_against_oracle
from tests.test_scenes import POLYGON_SCENES, SUPPORTED
from tests.test_spec_gpu import spec_check

pytestmark = pytest.mark.gpu

# lots per scene: the 3-branch 8-level tree is 22963 instances a lot
LOTS = {"3DTree_8_3": 8, "MengerSponge": 4}


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def E():
    from operatorgraph_amd import engine
    return engine


def _terminals(name, ax):
    from oracle import oracle as O
    r = O.derive(SceneProgram(name), ax)
    used = [t for t in range(abi.MAX_TERMINALS) if r.count(t)]
    return max(used) + 1, max(r.count(t) for t in used) + 1024


@pytest.mark.parametrize("technique", TECHNIQUES)
@pytest.mark.parametrize("name", SUPPORTED)
def test_scene_interpreter(E, name, technique):
    ax = scene_axioms(name, LOTS.get(name, 32), menger_attribute=3)
    n_terms, cap = _terminals(name, ax)
    _note(f"scene {name}: {len(ax)} lots, technique {technique}")
    check_against_oracle(E, SceneProgram(name), ax, n_terms, technique, cap=cap)


@pytest.mark.parametrize("name", SUPPORTED)
def test_scene_specialised(name):
    ax = scene_axioms(name, LOTS.get(name, 32), menger_attribute=3)
    n_terms, cap = _terminals(name, ax)
    _note(f"scene {name}: {len(ax)} lots, specialised kernel")
    spec_check(SceneProgram(name), ax, n_terms, cap=cap)


@pytest.mark.parametrize("technique", TECHNIQUES)
@pytest.mark.parametrize("name", POLYGON_SCENES)
def test_polygon_scene_interpreter(E, name, technique):
    ax = scene_axioms(name, 32)
    n_terms, cap = _terminals(name, ax)
    _note(f"scene {name}: {len(ax)} lots, technique {technique}")
    check_against_oracle(E, SceneProgram(name), ax, n_terms, technique, cap=cap)


@pytest.mark.parametrize("name", POLYGON_SCENES)
def test_polygon_scene_specialised(name):
    ax = scene_axioms(name, 32)
    n_terms, cap = _terminals(name, ax)
    _note(f"scene {name}: {len(ax)} lots, specialised kernel")
    spec_check(SceneProgram(name), ax, n_terms, cap=cap)
