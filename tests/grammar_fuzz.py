"""Seeded random grammars in the grammar language, for parity tests.

A grammar is a chain of rules R0 .. Rk. Each rule names only later rules or
terminals, so every derivation is finite. Shape types are tracked so that
every operator has a procedure in the interpreter's list: CompSplit on
boxes, Extrude on quads. Parameters stay where the operators are defined:
* Subdivide mixes relative factors with absolute ones far below any size the
  axioms reach;
* Repeat uses relative extents;
* a quad is split only along X and Y, where it has an extent;
* scales are relative or moderate absolutes.

Operands exercise every form: numbers, variables, axes, repeat modes,
Rand, shape attributes and nested Exp.
"""
import numpy as np

SINGLE_OPS = ("Translate", "Rotate", "Scale", "SwapSize")
AXES = ("X", "Y", "Z")


class _Gen:
    def __init__(self, seed, n_rules, n_terms):
        self.rng = np.random.default_rng(seed)
        self.n_rules = n_rules
        self.n_terms = n_terms
        self.shape = {}          # rule index -> "Box" / "Quad"
        self.lines = []

    def num(self, lo, hi, nd=3):
        return f"{self.rng.uniform(lo, hi):.{nd}f}"

    def axis(self):
        return AXES[self.rng.integers(3)]

    def operand(self, lo, hi):
        """A scalar operand in about [lo, hi]."""
        k = self.rng.integers(6)
        if k == 0:
            return f"Rand({self.num(lo, (lo + hi) / 2)}, {self.num((lo + hi) / 2, hi)})"
        if k == 1:
            return f"Exp(ShapeSize({self.axis()}) * {self.num(lo / 4, hi / 4)})"
        if k == 2:
            return f"Exp(Exp({self.num(lo / 2, hi / 2)} + {self.num(lo / 2, hi / 2)}) * 1)"
        if k == 3:
            return "vhalf" if lo <= 0.5 <= hi else self.num(lo, hi)
        return self.num(lo, hi)

    def target(self, i, shape):
        """A successor of rule i with input `shape`: a later rule (taking that
        shape) or a terminal."""
        later = [j for j in range(i + 1, self.n_rules) if self.shape.get(j, shape) == shape]
        if later and self.rng.random() < 0.7:
            j = int(self.rng.choice(later))
            self.shape[j] = shape
            return f"R{j}"
        if self.rng.random() < 0.1:
            return "Discard()"
        # one declaration names a box and a quad terminal of each index
        return f"T{'B' if shape == 'Box' else 'Q'}{self.rng.integers(self.n_terms)}"

    def chain(self, shape):
        ops = []
        for _ in range(self.rng.integers(0, 3)):
            # quads keep their zero height in Z: no SwapSize
            op = SINGLE_OPS[self.rng.integers(len(SINGLE_OPS) - (shape == "Quad"))]
            if op == "Translate":
                ops.append(f"Translate({self.operand(-1, 1)}, {self.operand(-1, 1)}, {self.num(-1, 1)})")
            elif op == "Rotate":
                ops.append(f"Rotate({self.num(-90, 90, 2)}, {self.num(-90, 90, 2)}, {self.operand(-45, 45)})")
            elif op == "Scale":
                sc = [self.num(-1.2, -0.6) if self.rng.random() < 0.6 else self.num(0.5, 3) for _ in range(3)]
                ops.append(f"Scale({sc[0]}, {sc[1]}, {sc[2]})")
            else:
                p = self.rng.permutation(3)
                ops.append(f"SwapSize({AXES[p[0]]}, {AXES[p[1]]}, {AXES[p[2]]})")
        return ops

    def body(self, i, shape):
        ops = self.chain(shape)
        kinds = ["SubDiv", "Repeat", "Replicate", "IfSizeLess", "If", "RandomRule", "leaf"]
        kinds.append("CompSplit" if shape == "Box" else "Extrude")
        k = kinds[self.rng.integers(len(kinds))]
        if k == "SubDiv":
            n = int(self.rng.integers(2, 5))
            fs = []
            for _ in range(n):
                fs.append(self.num(-2, -0.2, 2) if self.rng.random() < 0.8 else self.num(0.01, 0.05, 3))
            succ = " | ".join(f"{f}: {self.target(i, shape)}" for f in fs)
            ax = self.axis() if shape == "Box" else AXES[self.rng.integers(2)]   # a quad has no Z extent
            tail = f"SubDiv({ax}) {{ {succ} }}"
        elif k == "Repeat":
            mode = ("ANCHOR_TO_START", "ANCHOR_TO_END", "ADJUST_TO_FILL")[self.rng.integers(3)]
            ax = self.axis() if shape == "Box" else AXES[self.rng.integers(2)]
            tail = f"Repeat({ax}, {self.num(-0.45, -0.2, 2)}, {mode}) {self.target(i, shape)}"
        elif k == "Replicate":
            n = int(self.rng.integers(2, 4))
            tail = "Replicate() { " + " | ".join(self.target(i, shape) for _ in range(n)) + " }"
        elif k == "IfSizeLess":
            tail = (f"IfSizeLess({self.axis()}, {self.num(0.2, 2)}) "
                    f"{{ {self.target(i, shape)} | {self.target(i, shape)} }}")
        elif k == "If":
            cond = f"Exp(ShapeSize({self.axis()}) > {self.num(0.2, 2)})"
            if self.rng.random() < 0.5:
                cond = f"Exp({cond} && Exp(ShapeSeed() < {self.num(0.2, 0.8)}))"
            tail = f"If({cond}) {{ {self.target(i, shape)} | {self.target(i, shape)} }}"
        elif k == "RandomRule":
            n = int(self.rng.integers(2, 4))
            ws = [self.num(0.1, 1.0, 2) for _ in range(n)]
            chance = "ShapeSeed()" if self.rng.random() < 0.5 else "ShapeCustomAttribute()"
            tail = f"RandomRule({chance}) {{ " + " | ".join(f"{w}: {self.target(i, shape)}" for w in ws) + " }"
        elif k == "CompSplit":
            tail = ("CompSplit() { " + " | ".join(self.target(i, "Quad") for _ in range(3)) + " }")
        elif k == "Extrude":
            tail = f"Extrude({self.num(0.2, 2)}) {self.target(i, 'Box')}"
        else:
            tail = self.target(i, shape)
            if tail.startswith("R") and not ops:
                ops.append(f"Translate(0, {self.num(-1, 1)}, 0)")   # no bare aliases
        return " ".join(ops + [tail])

    def grammar(self, axiom_shape):
        self.shape[0] = axiom_shape
        out = ["// generated", "var vhalf = 0.5;", f"axiom {axiom_shape} R0;"]
        for t in range(self.n_terms):
            out.append(f"terminal TB{t} TQ{t} (1, {t});")
        bodies = {}
        for i in range(self.n_rules):
            if i not in self.shape:
                continue     # unreachable rule: no text
            bodies[i] = self.body(i, self.shape[i])
        for i in sorted(bodies):
            out.append(f"R{i} = {bodies[i]};")
        return "\n".join(out) + "\n"


def random_grammar(seed, n_rules=8, n_terms=3):
    """(grammar text, axiom shape) for `seed`."""
    rng = np.random.default_rng(seed + 1_000_003)
    shape = "Box" if rng.random() < 0.7 else "Quad"
    return _Gen(seed, n_rules, n_terms).grammar(shape), shape


def lots(g, n, seed, quad=False):
    """n axioms for compiled grammar g: sizes 1..6 per axis (no height for a
    quad), seeds and positions from `seed`."""
    rng = np.random.default_rng(seed)
    sizes = rng.uniform(1.0, 6.0, size=(n, 3)).astype(np.float32)
    if quad:
        sizes[:, 2] = 0.0
    pos = np.stack([np.arange(n) * 20.0, np.zeros(n), np.zeros(n)], 1)
    return g.axiom_records(rng.random(n).astype(np.float32), sizes=sizes, positions=pos)
