"""Writes tests/golden/grammar_tables.json: the dispatch tables this
library's grammar front end compiles from the reference's interpreter
grammars (src/interpreter/grammars/*/grammar.pga), so GPU tests can derive
them where the reference is absent.

Each record holds the compiled entries (zlib + base64 of the
pga_dispatch_entry array), the axioms, the SHA-256 of the grammar text it
was compiled from and the instance counts the reference's configuration
file states for it (src/interpreter/grammars/*/configuration.xml,
maxNumElements per InstancedTriangleMesh index). The grammar text itself is
not stored. Grammars the front end refuses are listed with the message.

Run from the repository root after building the library:
    python tests/golden/make_grammar_tables.py
"""
import base64
import glob
import hashlib
import json
import os
import re
import sys
import zlib

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

GRAMMARS = "/root/reference/src/interpreter/grammars"
OUT = os.path.join(ROOT, "tests", "golden", "grammar_tables.json")


def config_counts(path):
    """index -> maxNumElements of each InstancedTriangleMesh element."""
    if not os.path.exists(path):
        return {}
    text = open(path, encoding="utf-8").read()
    out = {}
    for m in re.finditer(r"<InstancedTriangleMesh\b([^>]*)>", text):
        attrs = dict(re.findall(r'(\w+)="([^"]*)"', m.group(1)))
        if "index" in attrs and "maxNumElements" in attrs:
            out[int(attrs["index"])] = int(attrs["maxNumElements"])
    return out


def main():
    from operatorgraph_amd import compiler
    records = {}
    for path in sorted(glob.glob(os.path.join(GRAMMARS, "*", "grammar.pga"))):
        name = os.path.basename(os.path.dirname(path))
        text = open(path, encoding="utf-8").read()
        rec = {"sha256": hashlib.sha256(text.encode("utf-8")).hexdigest(),
               "config_max_elements": config_counts(os.path.join(os.path.dirname(path), "configuration.xml"))}
        try:
            g = compiler.compile_grammar(text, name)
            rec["entries"] = base64.b64encode(zlib.compress(g.raw, 9)).decode()
            rec["num_entries"] = len(g.program.table.entries)
            rec["axioms"] = g.axioms
        except compiler.GrammarError as e:
            rec["refused"] = str(e)
        records[name] = rec
    with open(OUT, "w", encoding="utf-8") as f:
        json.dump(records, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {OUT}: {len(records)} grammars")


if __name__ == "__main__":
    main()
