"""Writes tests/golden/scene_tables.json: the partitions of the reference's
compiled test-runner scenes (src/test_runner/*.cuh, the analyzer's code
generator output: a Proc<> procedure list and a DispatchTable initialiser,
CodeGenerator.cpp:1691-1722) lowered by this build's C++ front end
(include/pga/core/Core.h) into the C ABI's procedure and entry arrays.

Run in the build container, where the reference is at /root/reference:

    python tests/golden/make_scene_tables.py [/root/reference [out.json]]

For each scene the script reads the header, keeps its procedure list and
dispatch table (the partition; the scene controllers around it are the
renderer's), drops the trailing commas the generator leaves in template
argument lists (a CUDA front end accepts them, C++17 does not), compiles the
block against Core.h with g++ and runs it. Only what it prints is stored:
the lowered arrays, the axiom traits' entry index and size, and the SHA-256
of the header the tables came from - no text of the reference. A scene the
front end refuses (polygon shapes) is stored with the reason."""
import base64
import ctypes as C
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile
import zlib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SCENES = ["MengerSponge", "SimpleHouse", "Commercial", "MC_Skyscrapers", "MC_Spaceship", "3DTree_4_3", "3DTree_8_3",
          "3DTree_12_2", "Balcony", "SuburbanHouse"]

PROGRAM = r"""
#include <pga/core/Core.h>

#include <cstdio>
#include <vector>

using namespace PGA;
using namespace PGA::Operators;
using namespace PGA::Shapes;
using namespace PGA::Parameters;

namespace Scene {
%s
}

template <class T> static void dump(const char* path, const std::vector<T>& v) {
    FILE* f = std::fopen(path, "wb");
    if (!f) throw std::runtime_error("cannot write output");
    std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

int main(int argc, char** argv) {
    try {
        const auto procs = Lowering::lower<Scene::ProcedureList>();
        auto entries = Scene::dispatchTable.toDispatchTableEntriesPtr();
        const auto prog = Lowering::expand(procs, reinterpret_cast<const pga_dispatch_entry*>(entries.get()),
                                           Scene::dispatchTable.entries.size());
        dump(argv[1], prog.procedures);
        dump(argv[2], prog.entries);
        std::printf("%%zu %%zu\n", prog.procedures.size(), prog.entries.size());
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%%s\n", e.what());
        return 1;
    }
}
"""


def partition_block(text):
    """The procedure list (with its Pn procedure structs) and the dispatch
    table of a scene header, trailing commas of template lists removed."""
    lines = text.splitlines()
    start = next(i for i, l in enumerate(lines) if re.match(r"\s*struct (P\d+|ProcedureList)\b", l))
    dt = next(i for i, l in enumerate(lines) if "DispatchTable dispatchTable" in l)
    end = next(i for i in range(dt, len(lines)) if lines[i].strip() == "};")
    block = "\n".join(lines[start:end + 1])
    return re.sub(r",(\s*)>", r"\1>", block)


def axiom_traits(text):
    """Entry index and size of the scene's AxiomTraits when both are
    constants (None otherwise)."""
    size = re.search(r"getSize\([^)]*\)\s*\{\s*return math::float3\(([-0-9.f]+),\s*([-0-9.f]+),\s*([-0-9.f]+)\)",
                     text)
    entry = re.search(r"getEntryIndex\(\)\s*\{\s*return (-?\d+);", text)
    return {"entry": int(entry.group(1)) if entry else 0,
            "size": [float(g.rstrip("f")) for g in size.groups()] if size else None}


def lower(name, text, work):
    src = os.path.join(work, name + ".cpp")
    exe = os.path.join(work, name)
    with open(src, "w") as f:
        f.write(PROGRAM % partition_block(text))
    cc = subprocess.run(["g++", "-std=c++17", "-O1", "-w", "-I", os.path.join(ROOT, "include"), src, "-o", exe],
                        capture_output=True, text=True)
    if cc.returncode != 0:
        err = [l for l in cc.stderr.splitlines() if "error" in l]
        return None, "does not compile against Core.h: " + (err[0].split("error: ", 1)[-1] if err else "?")
    procs, entries = exe + ".procs", exe + ".entries"
    run = subprocess.run([exe, procs, entries], capture_output=True, text=True)
    if run.returncode != 0:
        return None, run.stderr.strip()
    return (open(procs, "rb").read(), open(entries, "rb").read()), None


def main(ref="/root/reference", out_path=os.path.join(HERE, "scene_tables.json")):
    from operatorgraph_amd import abi
    out = {}
    with tempfile.TemporaryDirectory() as work:
        for name in SCENES:
            rel = f"src/test_runner/{name}.cuh"
            raw = open(os.path.join(ref, rel), "rb").read()
            text = raw.decode("utf-8", errors="replace")
            rec = {"header": rel, "sha256": hashlib.sha256(raw).hexdigest(), "axiom": axiom_traits(text)}
            arrays, why = lower(name, text, work)
            if arrays is None:
                rec["refused"] = why
            else:
                p, e = arrays
                rec["procedures"] = base64.b64encode(zlib.compress(p, 9)).decode()
                rec["entries"] = base64.b64encode(zlib.compress(e, 9)).decode()
                rec["num_procedures"] = len(p) // C.sizeof(abi.Procedure)
                rec["num_entries"] = len(e) // C.sizeof(abi.DispatchEntry)
            out[name] = rec
            print(name, "refused: " + why if arrays is None else f"{rec['num_procedures']} procedures, "
                  f"{rec['num_entries']} entries", flush=True)
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main(*sys.argv[1:3])
