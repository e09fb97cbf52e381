// pga_compiler.cpp - the PGA grammar language (*.pga) to a dispatch table for
// the interpreter's procedure list. Host code only.
//
// Language (src/compiler/Parser.cpp:305-467): `var` declarations, then
// `axiom` declarations, then `terminal` declarations, then production rules
// `Name = successor ;` (also `->`, `-->`, `=>`, `==>`, `>>`). A successor is
// an operator with its operands and its successors (one chained successor, a
// `{ a | b | ... }` block, or a `{ f: a | g: b }` block of parameterised
// successors), `Discard()` / `Generate()`, or a symbol (a rule or a
// terminal). Operands: numbers and variables, axes, repeat modes,
// Rand(a, b), shape attributes, (x, y) vertices and Exp(operand op operand).
// Comments are C++ comments.
//
// Table generation follows DispatchTableGenerator::fromRuleSet
// (src/compiler/DispatchTableGenerator.cpp:555-609) with its entry builders
// (:144-548), so entry numbers are the ones that function gives:
//   * axioms are queued in declaration order; a symbol's entry is allocated
//     when it is first named as a successor, or when it is taken from the
//     queue if nothing named it before;
//   * an inline operator's entry is allocated when its parent is visited,
//     with its siblings, and filled right after the parent;
//   * a symbol keeps the shape type it was first reached with;
//   * a symbol that is neither a rule nor a declared terminal becomes a
//     Generate with generation function 0 and terminal index 0
//     (visitTerminalSymbol, :216-264), as in the reference.
// Differences, all on inputs the reference does not turn into a working
// table: a symbol reached with two shape types is an error (the reference
// keeps the first silently); a rule whose body is a bare symbol (`A = B;`)
// shares B's entry (the reference leaves A's entry empty); a rule defined
// twice, an unknown variable and a stochastic rule (`A = 30%: B else: C;`,
// whose RandomRule entry the reference builds without its chance parameter)
// are errors; the dynamic-polygon operators (SetAs*) are refused because the
// interpreter's default procedure list has no procedures for the shapes
// they produce.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pga.h"

namespace {

// ---- the interpreter's procedure list (src/interpreter/PGAFacade.cu:24-62) ----
const pga_procedure kInterpreterProcedures[] = {
    // Quad
    {PGA_OP_DISCARD, PGA_SHAPE_QUAD, 0, 0}, {PGA_OP_EXTRUDE, PGA_SHAPE_QUAD, 0, 0},
    {PGA_OP_GENERATE, PGA_SHAPE_QUAD, 0, 0}, {PGA_OP_GENERATE, PGA_SHAPE_QUAD, 1, 0},
    {PGA_OP_IF, PGA_SHAPE_QUAD, 0, 0}, {PGA_OP_IFSIZELESS, PGA_SHAPE_QUAD, 0, 0},
    {PGA_OP_STOCHASTIC, PGA_SHAPE_QUAD, 0, 0}, {PGA_OP_REPEAT, PGA_SHAPE_QUAD, 0, 0},
    {PGA_OP_REPLICATE, PGA_SHAPE_QUAD, 0, 0}, {PGA_OP_ROTATE, PGA_SHAPE_QUAD, 0, 0},
    {PGA_OP_SCALE, PGA_SHAPE_QUAD, 0, 0}, {PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON, PGA_SHAPE_QUAD, 0, 0},
    {PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, PGA_SHAPE_QUAD, 0, 0},
    {PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON, PGA_SHAPE_QUAD, 0, 0},
    {PGA_OP_SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM, PGA_SHAPE_QUAD, 0, 0}, {PGA_OP_SUBDIV, PGA_SHAPE_QUAD, 0, 0},
    {PGA_OP_SWAPSIZE, PGA_SHAPE_QUAD, 0, 0}, {PGA_OP_TRANSLATE, PGA_SHAPE_QUAD, 0, 0},
    // Box
    {PGA_OP_COMPSPLIT, PGA_SHAPE_BOX, 0, 0}, {PGA_OP_DISCARD, PGA_SHAPE_BOX, 0, 0},
    {PGA_OP_GENERATE, PGA_SHAPE_BOX, 0, 0}, {PGA_OP_GENERATE, PGA_SHAPE_BOX, 1, 0},
    {PGA_OP_IF, PGA_SHAPE_BOX, 0, 0}, {PGA_OP_IFSIZELESS, PGA_SHAPE_BOX, 0, 0},
    {PGA_OP_STOCHASTIC, PGA_SHAPE_BOX, 0, 0}, {PGA_OP_REPEAT, PGA_SHAPE_BOX, 0, 0},
    {PGA_OP_REPLICATE, PGA_SHAPE_BOX, 0, 0}, {PGA_OP_ROTATE, PGA_SHAPE_BOX, 0, 0},
    {PGA_OP_SCALE, PGA_SHAPE_BOX, 0, 0}, {PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON, PGA_SHAPE_BOX, 0, 0},
    {PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, PGA_SHAPE_BOX, 0, 0},
    {PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON, PGA_SHAPE_BOX, 0, 0},
    {PGA_OP_SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM, PGA_SHAPE_BOX, 0, 0}, {PGA_OP_SUBDIV, PGA_SHAPE_BOX, 0, 0},
    {PGA_OP_SWAPSIZE, PGA_SHAPE_BOX, 0, 0}, {PGA_OP_TRANSLATE, PGA_SHAPE_BOX, 0, 0},
    // the dynamic shapes of the list built with INTERPRETER_SUPPORT_ALL_SHAPES
    // (PGAFacade.cu:65-130, 395-470), which follow Box there too
#define PGA_PLANAR_PROCS(S)                                                                                       \
    {PGA_OP_DISCARD, S, 0, 0}, {PGA_OP_EXTRUDE, S, 0, 0}, {PGA_OP_GENERATE, S, 0, 0}, {PGA_OP_GENERATE, S, 1, 0},   \
        {PGA_OP_IF, S, 0, 0}, {PGA_OP_IFSIZELESS, S, 0, 0}, {PGA_OP_STOCHASTIC, S, 0, 0},                          \
        {PGA_OP_REPEAT, S, 0, 0}, {PGA_OP_REPLICATE, S, 0, 0}, {PGA_OP_ROTATE, S, 0, 0}, {PGA_OP_SCALE, S, 0, 0}, \
        {PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON, S, 0, 0}, {PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, S, 0, 0},      \
        {PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON, S, 0, 0}, {PGA_OP_SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM, S, 0, 0},    \
        {PGA_OP_SUBDIV, S, 0, 0}, {PGA_OP_SWAPSIZE, S, 0, 0}, {PGA_OP_TRANSLATE, S, 0, 0}
#define PGA_PRISM_PROCS(S)                                                                                        \
    {PGA_OP_COMPSPLIT, S, 0, 0}, {PGA_OP_DISCARD, S, 0, 0}, {PGA_OP_GENERATE, S, 0, 0}, {PGA_OP_GENERATE, S, 1, 0}, \
        {PGA_OP_IF, S, 0, 0}, {PGA_OP_IFSIZELESS, S, 0, 0}, {PGA_OP_STOCHASTIC, S, 0, 0},                          \
        {PGA_OP_REPEAT, S, 0, 0}, {PGA_OP_REPLICATE, S, 0, 0}, {PGA_OP_ROTATE, S, 0, 0}, {PGA_OP_SCALE, S, 0, 0}, \
        {PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON, S, 0, 0}, {PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, S, 0, 0},      \
        {PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON, S, 0, 0}, {PGA_OP_SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM, S, 0, 0},    \
        {PGA_OP_SUBDIV, S, 0, 0}, {PGA_OP_SWAPSIZE, S, 0, 0}, {PGA_OP_TRANSLATE, S, 0, 0}
    PGA_PLANAR_PROCS(PGA_SHAPE_DYNAMIC_CONVEX_POLYGON),
    PGA_PLANAR_PROCS(PGA_SHAPE_DYNAMIC_POLYGON),
    PGA_PRISM_PROCS(PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM),
    PGA_PRISM_PROCS(PGA_SHAPE_DYNAMIC_RIGHT_PRISM),
#undef PGA_PLANAR_PROCS
#undef PGA_PRISM_PROCS
};
const uint32_t kNumInterpreterProcedures = sizeof(kInterpreterProcedures) / sizeof(kInterpreterProcedures[0]);

// Operator::nextShapeType (Operator.cpp:20-130) for successor k: a
// ComponentSplit of a Box gives Quads, of a dynamic prism its polygon
// (the caps, successors 0 and 1) and Quads (the sides); Extrude gives a
// planar shape's extruded type; SetAs its target
int next_shape_of(int op, int shape, size_t k) {
    const bool prism = shape == PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM || shape == PGA_SHAPE_DYNAMIC_RIGHT_PRISM;
    if (op == PGA_OP_COMPSPLIT && shape == PGA_SHAPE_BOX) return PGA_SHAPE_QUAD;
    if (op == PGA_OP_COMPSPLIT && prism)
        return k == 2 ? PGA_SHAPE_QUAD
             : shape == PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM ? PGA_SHAPE_DYNAMIC_CONVEX_POLYGON : PGA_SHAPE_DYNAMIC_POLYGON;
    if (op == PGA_OP_EXTRUDE && shape == PGA_SHAPE_QUAD) return PGA_SHAPE_BOX;
    if (op == PGA_OP_EXTRUDE && shape == PGA_SHAPE_DYNAMIC_CONVEX_POLYGON) return PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM;
    if (op == PGA_OP_EXTRUDE && shape == PGA_SHAPE_DYNAMIC_POLYGON) return PGA_SHAPE_DYNAMIC_RIGHT_PRISM;
    if (op == PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON) return PGA_SHAPE_DYNAMIC_CONVEX_POLYGON;
    if (op == PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM) return PGA_SHAPE_DYNAMIC_CONVEX_RIGHT_PRISM;
    if (op == PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON) return PGA_SHAPE_DYNAMIC_POLYGON;
    if (op == PGA_OP_SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM) return PGA_SHAPE_DYNAMIC_RIGHT_PRISM;
    return shape;
}

struct CompileError : std::runtime_error {
    explicit CompileError(const std::string& m) : std::runtime_error(m) {}
};

std::string at_line(int line) { return line > 0 ? "line " + std::to_string(line) + ": " : std::string(); }

const char* operator_name(int op) {
    switch (op) {
    case PGA_OP_TRANSLATE: return "Translate";
    case PGA_OP_ROTATE: return "Rotate";
    case PGA_OP_SCALE: return "Scale";
    case PGA_OP_EXTRUDE: return "Extrude";
    case PGA_OP_COMPSPLIT: return "CompSplit";
    case PGA_OP_SUBDIV: return "SubDiv";
    case PGA_OP_REPEAT: return "Repeat";
    case PGA_OP_DISCARD: return "Discard";
    case PGA_OP_IF: return "If";
    case PGA_OP_IFSIZELESS: return "IfSizeLess";
    case PGA_OP_IFCOLLIDES: return "IfCollides";
    case PGA_OP_GENERATE: return "Generate";
    case PGA_OP_STOCHASTIC: return "RandomRule";
    case PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON: return "SetAsDynamicConvexPolygon";
    case PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM: return "SetAsDynamicConvexRightPrism";
    case PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON: return "SetAsDynamicConcavePolygon";
    case PGA_OP_SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM: return "SetAsDynamicConcaveRightPrism";
    case PGA_OP_COLLIDER: return "Collider";
    case PGA_OP_SWAPSIZE: return "SwapSize";
    case PGA_OP_REPLICATE: return "Replicate";
    default: return "?";
    }
}

// PGA::Compiler::ShapeType order (include/pga/compiler/ShapeType.h)
const char* const kShapeNames[] = {"Triangle", "Quad", "Pentagon", "Hexagon", "Heptagon", "Octagon",
                                   "DynamicConvexPolygon", "DynamicPolygon", "Prism3", "Box", "Prism5",
                                   "Prism6", "Prism7", "Prism8", "DynamicConvexRightPrism",
                                   "DynamicRightPrism", "Sphere"};
const int kNumShapes = (int)(sizeof(kShapeNames) / sizeof(kShapeNames[0]));

std::string shape_name(int s) { return s >= 0 && s < kNumShapes ? kShapeNames[s] : std::to_string(s); }

// ---- syntax tree ----
enum OperandKind { OD_NUMBER, OD_AXIS, OD_MODE, OD_RAND, OD_ATTR, OD_VEC2, OD_EXP };
enum AttrKind { AT_POS, AT_SIZE, AT_ROTATION, AT_NORMAL, AT_SEED, AT_CUSTOM };

struct Operand {
    OperandKind kind = OD_NUMBER;
    double a = 0.0, b = 0.0;          // number, axis, mode: a; Rand and (x, y): a, b
    int attr = 0, axis = 0, component = 0;
    int op = 0;                       // Exp: the operation
    std::shared_ptr<Operand> left, right;
};

struct Successor;
typedef std::shared_ptr<Successor> SuccessorP;

struct Successor {
    bool is_symbol = false;
    std::string symbol;               // is_symbol
    int op = 0;                       // pga_operator_type
    std::vector<Operand> params;
    std::vector<double> factors;      // parameterised successors: one per successor
    std::vector<SuccessorP> next;
    int line = 0;
};

struct Axiom {
    std::string symbol;
    int shape = 0;
};

// ---- parser: recursive descent over the source text ----
class Parser {
public:
    explicit Parser(const std::string& s) : src_(s) {}

    std::vector<Axiom> axioms;
    std::map<std::string, std::vector<double>> terminals;   // name -> (generation function, index, attributes)
    std::map<std::string, SuccessorP> rules;

    void parse() {
        while (keyword("var")) parse_var();
        if (!peek_word("axiom")) fail("expecting an axiom declaration");
        while (keyword("axiom")) parse_axiom();
        while (keyword("terminal")) parse_terminal();
        if (at_end()) fail("expecting a production rule");
        while (!at_end()) parse_rule();
    }

private:
    const std::string& src_;
    size_t p_ = 0;
    std::map<std::string, double> vars_;

    int line() const {
        return 1 + (int)std::count(src_.begin(), src_.begin() + (long)std::min(p_, src_.size()), '\n');
    }
    [[noreturn]] void fail(const std::string& what) const { throw CompileError(at_line(line()) + what); }
    bool at_end() { skip(); return p_ >= src_.size(); }
    void skip() {
        for (;;) {
            while (p_ < src_.size() && std::isspace((unsigned char)src_[p_])) ++p_;
            if (src_.compare(p_, 2, "//") == 0) {
                while (p_ < src_.size() && src_[p_] != '\n') ++p_;
            } else if (src_.compare(p_, 2, "/*") == 0) {
                const size_t e = src_.find("*/", p_ + 2);
                if (e == std::string::npos) fail("unterminated comment");
                p_ = e + 2;
            } else {
                return;
            }
        }
    }
    static bool ident_char(char c) { return std::isalnum((unsigned char)c) || c == '_'; }
    bool lit(const char* s) {
        skip();
        const size_t n = std::strlen(s);
        if (src_.compare(p_, n, s) == 0) { p_ += n; return true; }
        return false;
    }
    void expect(const char* s) { if (!lit(s)) fail(std::string("expecting '") + s + "'"); }
    bool keyword(const char* s) {
        skip();
        const size_t n = std::strlen(s);
        if (src_.compare(p_, n, s) == 0 && (p_ + n >= src_.size() || !ident_char(src_[p_ + n]))) {
            p_ += n;
            return true;
        }
        return false;
    }
    bool peek_word(const char* s) {
        const size_t save = p_;
        const bool k = keyword(s);
        p_ = save;
        return k;
    }
    bool try_ident(std::string& out) {
        skip();
        const size_t b = p_;
        while (p_ < src_.size() && ident_char(src_[p_])) ++p_;
        if (p_ == b) return false;
        out = src_.substr(b, p_ - b);
        return true;
    }
    std::string ident() {
        std::string s;
        if (!try_ident(s)) fail("expecting a name");
        return s;
    }
    bool try_number(double& v) {
        skip();
        if (p_ >= src_.size()) return false;
        const char c = src_[p_];
        if (!(std::isdigit((unsigned char)c) || c == '-' || c == '+' || c == '.')) return false;
        const char* b = src_.c_str() + p_;
        char* e = nullptr;
        v = std::strtod(b, &e);
        if (e == b) return false;
        p_ += (size_t)(e - b);
        return true;
    }
    double number() {
        double v;
        if (!try_number(v)) fail("expecting a number");
        return v;
    }
    double variable(const std::string& name) const {
        auto it = vars_.find(name);
        if (it == vars_.end()) fail("unknown variable '" + name + "'");
        return it->second;
    }
    double number_or_var() {
        double v;
        if (try_number(v)) return v;
        return variable(ident());
    }

    void parse_var() {
        const std::string name = ident();
        expect("=");
        const double v = number_or_var();
        expect(";");
        vars_[name] = v;
    }

    void parse_axiom() {
        Axiom a;
        const std::string shape = ident();
        a.shape = -1;
        for (int i = 0; i < kNumShapes; ++i)
            if (shape == kShapeNames[i]) a.shape = i;
        if (a.shape < 0) fail("unknown shape type '" + shape + "'");
        if (lit("{")) {               // polygon vertices: read and not used by Box/Quad axioms
            while (lit("(")) {
                number_or_var();
                expect(",");
                number_or_var();
                expect(")");
                while (lit(",")) {}
            }
            expect("}");
        }
        a.symbol = ident();
        expect(";");
        axioms.push_back(a);
    }

    void parse_terminal() {
        std::vector<std::string> names;
        names.push_back(ident());
        while (!lit("(")) names.push_back(ident());
        std::vector<double> attrs;
        attrs.push_back(number());
        while (lit(",")) attrs.push_back(number());
        expect(")");
        expect(";");
        for (const auto& n : names) terminals[n] = attrs;
    }

    void parse_rule() {
        const std::string name = ident();
        if (!(lit("-->") || lit("->") || lit("==>") || lit("=>") || lit(">>") || lit("="))) fail("expecting '='");
        SuccessorP s = successor();
        skip();
        if (!lit(";")) fail("expecting ';' (a rule with several stochastic successors is not supported)");
        if (rules.count(name)) fail("rule '" + name + "' is defined twice");
        rules[name] = s;
    }

    static int operator_of(const std::string& s) {
        static const int ops[] = {PGA_OP_TRANSLATE, PGA_OP_ROTATE, PGA_OP_SCALE, PGA_OP_EXTRUDE, PGA_OP_COMPSPLIT,
                                  PGA_OP_SUBDIV, PGA_OP_REPEAT, PGA_OP_IF, PGA_OP_IFSIZELESS, PGA_OP_IFCOLLIDES,
                                  PGA_OP_STOCHASTIC, PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON,
                                  PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON,
                                  PGA_OP_SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM, PGA_OP_COLLIDER, PGA_OP_SWAPSIZE,
                                  PGA_OP_REPLICATE};
        for (int o : ops)
            if (s == operator_name(o)) return o;
        return -1;
    }

    // "p%:" or "else:" in front of a successor (a stochastic rule)
    bool probability_prefix() {
        const size_t save = p_;
        double v;
        if (keyword("else") && lit(":")) return true;
        p_ = save;
        if (try_number(v) && lit("%") && lit(":")) return true;
        p_ = save;
        return false;
    }

    // a factor before ':' in a block of parameterised successors
    bool factor(double& v) {
        const size_t save = p_;
        std::string name;
        if (try_number(v) && lit(":")) return true;
        p_ = save;
        if (try_ident(name) && lit(":")) {
            v = variable(name);
            return true;
        }
        p_ = save;
        return false;
    }

    SuccessorP successor() {
        if (probability_prefix()) fail("stochastic rules (p%: ... else: ...) are not supported");
        skip();
        auto s = std::make_shared<Successor>();
        s->line = line();
        const std::string name = ident();
        if ((name == "Discard" || name == "Generate") && lit("(")) {
            expect(")");
            s->op = name == "Discard" ? PGA_OP_DISCARD : PGA_OP_GENERATE;
            return s;
        }
        const int op = operator_of(name);
        if (op >= 0 && lit("(")) {
            s->op = op;
            if (!lit(")")) {
                s->params.push_back(operand());
                while (lit(",")) s->params.push_back(operand());
                expect(")");
            }
            if (lit("{")) {
                double f;
                const size_t first = p_;
                const bool parameterised = factor(f);
                p_ = first;
                do {
                    if (parameterised) {
                        if (!factor(f)) fail("expecting 'factor:' before the successor");
                        s->factors.push_back(f);
                    }
                    s->next.push_back(successor());
                } while (lit("|"));
                expect("}");
            } else {
                s->next.push_back(successor());
            }
            return s;
        }
        s->is_symbol = true;
        s->symbol = name;
        return s;
    }

    int axis_name() {
        if (keyword("X")) return 0;
        if (keyword("Y")) return 1;
        if (keyword("Z")) return 2;
        fail("expecting an axis");
    }

    int operation() {
        // longest first; OperationType values (include/pga/core/OperationType.h)
        static const std::pair<const char*, int> ops[] = {
            {"==", PGA_OPT_EQ}, {"!=", PGA_OPT_NEQ}, {"<=", PGA_OPT_LEQ}, {">=", PGA_OPT_GEQ}, {"&&", PGA_OPT_AND},
            {"||", PGA_OPT_OR}, {"+", PGA_OPT_ADD}, {"-", PGA_OPT_SUB}, {"*", PGA_OPT_MULTI}, {"/", PGA_OPT_DIV},
            {"<", PGA_OPT_LT}, {">", PGA_OPT_GT}};
        for (const auto& o : ops)
            if (lit(o.first)) return o.second;
        fail("expecting an operation");
    }

    Operand operand() {
        Operand o;
        double v;
        if (keyword("X")) { o.kind = OD_AXIS; o.a = 0; return o; }
        if (keyword("Y")) { o.kind = OD_AXIS; o.a = 1; return o; }
        if (keyword("Z")) { o.kind = OD_AXIS; o.a = 2; return o; }
        if (keyword("ANCHOR_TO_START")) { o.kind = OD_MODE; o.a = PGA_ANCHOR_TO_START; return o; }
        if (keyword("ANCHOR_TO_END")) { o.kind = OD_MODE; o.a = PGA_ANCHOR_TO_END; return o; }
        if (keyword("ADJUST_TO_FILL")) { o.kind = OD_MODE; o.a = PGA_ADJUST_TO_FILL; return o; }
        if (keyword("Rand")) {
            expect("(");
            o.kind = OD_RAND;
            o.a = number_or_var();
            expect(",");
            o.b = number_or_var();
            expect(")");
            return o;
        }
        static const std::pair<const char*, int> attrs[] = {{"ShapePos", AT_POS}, {"ShapeSize", AT_SIZE},
                                                            {"ShapeRotation", AT_ROTATION}, {"ShapeNormal", AT_NORMAL},
                                                            {"ShapeSeed", AT_SEED},
                                                            {"ShapeCustomAttribute", AT_CUSTOM}};
        for (const auto& at : attrs) {
            if (keyword(at.first)) {
                o.kind = OD_ATTR;
                o.attr = at.second;
                expect("(");
                if (!lit(")")) {
                    o.axis = axis_name();
                    while (lit(",")) o.component = axis_name();
                    expect(")");
                }
                return o;
            }
        }
        if (keyword("Exp")) {
            expect("(");
            o.kind = OD_EXP;
            o.left = std::make_shared<Operand>(operand());
            o.op = operation();
            o.right = std::make_shared<Operand>(operand());
            expect(")");
            return o;
        }
        if (lit("(")) {
            o.kind = OD_VEC2;
            o.a = number_or_var();
            expect(",");
            o.b = number_or_var();
            expect(")");
            return o;
        }
        o.kind = OD_NUMBER;
        o.a = try_number(v) ? v : variable(ident());
        return o;
    }
};

// ---- parameters (src/compiler/Parameters.cpp) ----
// PackUtils::packOperand (include/pga/core/PackUtils.h:18-23): the operand
// type in the top two bits of the float's bit pattern shifted right by two.
float pack_operand(int type, float x) {
    uint32_t r;
    std::memcpy(&r, &x, 4);
    r = (r >> 2) | ((uint32_t)(type & 3) << 30);
    float o;
    std::memcpy(&o, &r, 4);
    return o;
}
// PackUtils::packFloat2 (PackUtils.h:25-28)
float pack_float2(float x, float y) { return std::floor(x * 4095.0f) * 4096.0f + std::floor(y * 4095.0f); }

// ShapeAttr::getShapeAttrIndex and ShapeAttr::encode (Parameters.cpp:312-378)
int attr_index(const Operand& o, bool in_expression, int line) {
    switch (o.attr) {
    case AT_POS: return (o.axis + 1) * 4 - 1;
    case AT_SIZE: return o.axis + 12;
    case AT_SEED: return 15;
    case AT_CUSTOM: return 16;
    case AT_ROTATION: if (in_expression) return o.component * 4 + o.axis; break;
    case AT_NORMAL: if (in_expression) return o.axis * 4 + 2; break;
    default: break;
    }
    throw CompileError(at_line(line) + "ShapeRotation and ShapeNormal are only valid inside Exp(...)");
}

// Exp::encode (Parameters.cpp:470-475): postfix, left operand, right operand, operation
void encode(const Operand& o, std::vector<float>& b, int line) {
    switch (o.kind) {
    case OD_NUMBER: case OD_AXIS: case OD_MODE: b.push_back(pack_operand(PGA_ORT_SCALAR, (float)o.a)); break;
    case OD_RAND: b.push_back(pack_operand(PGA_ORT_RAND, pack_float2((float)o.a, (float)o.b))); break;
    case OD_ATTR: b.push_back(pack_operand(PGA_ORT_SHAPE_ATTR, (float)attr_index(o, true, line))); break;
    case OD_VEC2:   // Vec2::encode: the vector's length
        b.push_back(pack_operand(PGA_ORT_SCALAR, std::sqrt((float)o.a * (float)o.a + (float)o.b * (float)o.b)));
        break;
    case OD_EXP:
        encode(*o.left, b, line);
        encode(*o.right, b, line);
        b.push_back(pack_operand(PGA_ORT_OP, (float)o.op));
        break;
    }
}

pga_dispatch_parameter blank_param(int type) {
    pga_dispatch_parameter p;
    std::memset(&p, 0, sizeof(p));
    p.type = type;
    return p;
}

// Parameter::toDispatchTableParameter for each operand kind
pga_dispatch_parameter to_param(const Operand& o, int line) {
    pga_dispatch_parameter p = blank_param(PGA_PT_SCALAR);
    switch (o.kind) {
    case OD_NUMBER: case OD_AXIS: case OD_MODE:
        p.value[0] = (float)o.a;
        break;
    case OD_RAND:               // Rand(min, max, seed = 0)
        p.type = PGA_PT_RAND;
        p.value[0] = (float)o.a;
        p.value[1] = (float)o.b;
        break;
    case OD_ATTR:
        p.type = PGA_PT_SHAPE_ATTR;
        p.value[0] = (float)attr_index(o, false, line);
        break;
    case OD_VEC2:
        p.type = PGA_PT_VEC2;
        p.value[0] = (float)o.a;
        p.value[1] = (float)o.b;
        break;
    case OD_EXP: {              // value[0] = encoded length, then the postfix code
        std::vector<float> b;
        encode(o, b, line);
        if (b.size() + 1 > PGA_MAX_NUM_PARAMETER_VALUES) throw CompileError(at_line(line) + "expression too long");
        p.type = PGA_PT_EXP;
        p.value[0] = (float)b.size();
        for (size_t i = 0; i < b.size(); ++i) p.value[i + 1] = b[i];
        break;
    }
    }
    return p;
}

pga_dispatch_parameter scalar3(float a, float b, float c) {
    pga_dispatch_parameter p = blank_param(PGA_PT_SCALAR);
    p.value[0] = a;
    p.value[1] = b;
    p.value[2] = c;
    return p;
}

pga_dispatch_parameter scalar(float a) {
    pga_dispatch_parameter p = blank_param(PGA_PT_SCALAR);
    p.value[0] = a;
    return p;
}

// ---- table generation ----
class Generator {
public:
    explicit Generator(const Parser& parser) : g_(parser) {}

    std::vector<pga_dispatch_entry> entries;

    void run(std::vector<pga_grammar_axiom>& axioms_out) {
        std::deque<std::string> queue;
        std::set<std::string> queued;
        for (const Axiom& a : g_.axioms) {
            const std::string s = resolve(a.symbol);
            if (!check_shape(s, a.shape, 0)) shape_.emplace(s, a.shape);
            if (g_.rules.count(s) && queued.insert(s).second) queue.push_back(s);
        }
        queue_ = &queue;
        queued_ = &queued;
        while (!queue.empty()) {
            const std::string s = queue.front();
            queue.pop_front();
            int idx;
            auto it = entry_.find(s);
            if (it == entry_.end()) {
                idx = new_entry();
                entry_[s] = idx;
            } else {
                idx = it->second;
            }
            const int shape = shape_.at(s);
            auto r = g_.rules.find(s);
            if (r == g_.rules.end()) terminal_entry(idx, s, shape);
            else visit(idx, shape, r->second, s);
        }
        for (const Axiom& a : g_.axioms) {
            pga_grammar_axiom x;
            auto it = entry_.find(resolve(a.symbol));
            x.entry_index = it == entry_.end() ? -1 : it->second;   // no rule: skipped, PGAFacade.cu:708
            x.shape_type = a.shape;
            axioms_out.push_back(x);
        }
    }

private:
    const Parser& g_;
    std::map<std::string, int> entry_;
    std::map<std::string, int> shape_;
    std::deque<std::string>* queue_ = nullptr;
    std::set<std::string>* queued_ = nullptr;

    static pga_dispatch_entry blank() {
        pga_dispatch_entry e;
        std::memset(&e, 0, sizeof(e));
        for (int k = 0; k < PGA_MAX_NUM_SUCCESSORS; ++k) {
            e.successors[k].entry_index = -1;
            e.successors[k].phase_index = -1;
            e.edge_indices[k] = -1;
        }
        e.sub_graph_index = -1;
        return e;
    }
    int new_entry() {
        entries.push_back(blank());
        return (int)entries.size() - 1;
    }
    std::string resolve(const std::string& name) const {
        std::string s = name;
        for (size_t guard = 0; guard <= g_.rules.size(); ++guard) {
            auto it = g_.rules.find(s);
            if (it == g_.rules.end() || !it->second->is_symbol) return s;
            s = it->second->symbol;
        }
        throw CompileError("the rules aliasing '" + name + "' form a cycle");
    }
    // true when the symbol already has a shape type; an error if it differs
    bool check_shape(const std::string& s, int shape, int line) const {
        auto it = shape_.find(s);
        if (it == shape_.end()) return false;
        if (it->second != shape)
            throw CompileError(at_line(line) + "'" + s + "' expects shapes of type " +
                               shape_name(it->second) + " but also gets shapes of type " + shape_name(shape));
        return true;
    }
    // the successor entry of a symbol: DispatchTableGenerator.cpp:110-125
    int symbol_successor(const std::string& name, int shape, int line) {
        const std::string s = resolve(name);
        if (!check_shape(s, shape, line)) shape_.emplace(s, shape);
        auto it = entry_.find(s);
        if (it != entry_.end()) return it->second;
        const int e = new_entry();
        entry_[s] = e;
        if (queued_->insert(s).second) queue_->push_back(s);
        return e;
    }
    static int next_shape(int op, int shape, size_t k = 0) { return next_shape_of(op, shape, k); }
    int procedure(int op, int shape, int gen, int line) const {
        for (uint32_t i = 0; i < kNumInterpreterProcedures; ++i) {
            const pga_procedure& p = kInterpreterProcedures[i];
            if (p.op_type == op && p.shape_type == shape && p.gen_func_idx == gen) return (int)i;
        }
        throw CompileError(at_line(line) + "unsupported procedure (" + operator_name(op) + " on " + shape_name(shape) +
                           ", generation function " + std::to_string(gen) + ")");
    }
    // visitTerminalSymbol / buildGenerateEntry (DispatchTableGenerator.cpp:216-285)
    void generate_entry(pga_dispatch_entry& e, const std::vector<double>* attrs, int shape, int line) {
        int gen = 0;
        std::vector<double> rest;
        if (attrs && !attrs->empty()) {
            gen = (int)(*attrs)[0];
            rest.assign(attrs->begin() + 1, attrs->end());
        }
        if (rest.empty()) rest.push_back(0.0);
        if (rest.size() > 6) throw CompileError(at_line(line) + "too many terminal attributes");
        e.operator_code = procedure(PGA_OP_GENERATE, shape, gen, line);
        e.num_parameters = (uint32_t)rest.size();
        for (size_t k = 0; k < rest.size(); ++k) e.parameters[k] = scalar((float)rest[k]);
    }
    const std::vector<double>* terminal_of(const std::string& name) const {
        auto it = g_.terminals.find(name);
        return it == g_.terminals.end() ? nullptr : &it->second;
    }
    void terminal_entry(int idx, const std::string& name, int shape) {
        pga_dispatch_entry e = blank();
        generate_entry(e, terminal_of(name), shape, 0);
        entries[idx] = e;
    }
    void visit(int idx, int shape, const SuccessorP& s, const std::string& rule) {
        const int op = s->op;
        const int ln = s->line;
        pga_dispatch_entry e = blank();
        if (op == PGA_OP_GENERATE) {          // Generate(): the enclosing rule's terminal declaration
            generate_entry(e, terminal_of(rule), shape, ln);
            entries[idx] = e;
            return;
        }
        if (op == PGA_OP_IFCOLLIDES || op == PGA_OP_COLLIDER)
            throw CompileError(at_line(ln) + operator_name(op) + " is not supported (no interpreter procedure)");
        e.operator_code = procedure(op, shape, 0, ln);
        std::vector<int> succ;
        for (size_t k = 0; k < s->next.size(); ++k) {
            const SuccessorP& n = s->next[k];
            succ.push_back(n->is_symbol ? symbol_successor(n->symbol, next_shape(op, shape, k), n->line) : new_entry());
        }
        auto need = [&](size_t np, size_t nsucc) {   // Operator::check (ParserOperator.cpp:216-231)
            if (s->params.size() < np)
                throw CompileError(at_line(ln) + operator_name(op) + " requires at least " + std::to_string(np) +
                                   " parameters");
            if (nsucc && s->next.size() != nsucc)
                throw CompileError(at_line(ln) + operator_name(op) + " requires " + std::to_string(nsucc) +
                                   " successors");
        };
        auto need_factors = [&]() {
            if (s->factors.size() != s->next.size() || s->next.empty())
                throw CompileError(at_line(ln) + operator_name(op) + " requires a factor before each successor");
        };
        auto no_factors = [&]() {
            if (!s->factors.empty())
                throw CompileError(at_line(ln) + operator_name(op) + " cannot have successor parameters");
        };
        std::vector<pga_dispatch_parameter> params;
        switch (op) {
        case PGA_OP_TRANSLATE: case PGA_OP_ROTATE: case PGA_OP_SCALE: case PGA_OP_SWAPSIZE: case PGA_OP_REPEAT:
            need(3, 1);
            no_factors();
            for (int k = 0; k < 3; ++k) params.push_back(to_param(s->params[k], ln));
            break;
        case PGA_OP_EXTRUDE:      // the axis is implicit: Z first (ParserOperator.cpp:160-162)
            need(1, 1);
            no_factors();
            params.push_back(scalar(2.0f));
            params.push_back(to_param(s->params[0], ln));
            break;
        case PGA_OP_COMPSPLIT:
            need(0, 3);
            no_factors();
            break;
        case PGA_OP_SUBDIV: {     // buildSubdivideEntry (:355-383): running sums of the factors
            need(1, 0);
            need_factors();
            float rel = 0.0f, abs = 0.0f;
            std::vector<pga_dispatch_parameter> f;
            for (double x : s->factors) {
                const float v = (float)x;
                f.push_back(scalar3(v, abs, rel));
                if (v > 0) abs += v;
                else rel -= v;
            }
            params.push_back(scalar3(s->params[0].kind == OD_ATTR ? (float)attr_index(s->params[0], false, ln)
                                                                  : (float)s->params[0].a,
                                     abs, rel));
            params.insert(params.end(), f.begin(), f.end());
            break;
        }
        case PGA_OP_DISCARD:
            break;
        case PGA_OP_IF:
            need(1, 2);
            no_factors();
            params.push_back(to_param(s->params[0], ln));
            break;
        case PGA_OP_IFSIZELESS:
            need(2, 2);
            no_factors();
            params.push_back(to_param(s->params[0], ln));
            params.push_back(to_param(s->params[1], ln));
            break;
        case PGA_OP_STOCHASTIC:   // buildRandomRuleEntry (:533-548): the chance, then the factors
            need(1, 0);
            need_factors();
            params.push_back(to_param(s->params[0], ln));
            for (double x : s->factors) params.push_back(scalar((float)x));
            break;
        case PGA_OP_REPLICATE:
            if (!s->params.empty()) throw CompileError(at_line(ln) + "Replicate accepts no parameters");
            no_factors();
            break;
        case PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON: case PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM:
        case PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON: case PGA_OP_SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM: {
            // buildSetAs*Entry (DispatchTableGenerator.cpp:442-517): each
            // vertex (its first two values) as a Vec2, ordered counter-
            // clockwise from the bottom-left one (GeometryUtils.h:54-71)
            need(3, 1);
            no_factors();
            std::vector<std::pair<float, float>> v;
            for (const Operand& o : s->params) {
                const pga_dispatch_parameter q = to_param(o, ln);
                v.push_back({q.value[0], q.value[1]});
            }
            for (const auto& p : order_ccw(v)) {
                pga_dispatch_parameter q{};
                q.type = PGA_PT_VEC2;
                q.value[0] = p.first;
                q.value[1] = p.second;
                params.push_back(q);
            }
            break;
        }
        default:
            throw CompileError(at_line(ln) + operator_name(op) + " is not supported");
        }
        if (params.size() > PGA_MAX_NUM_PARAMETERS) throw CompileError(at_line(ln) + "too many parameters");
        if (succ.size() > PGA_MAX_NUM_SUCCESSORS) throw CompileError(at_line(ln) + "too many successors");
        e.num_parameters = (uint32_t)params.size();
        for (size_t k = 0; k < params.size(); ++k) e.parameters[k] = params[k];
        e.num_successors = (uint32_t)succ.size();
        for (size_t k = 0; k < succ.size(); ++k) {
            e.successors[k].entry_index = succ[k];
            e.successors[k].phase_index = 0;
        }
        entries[idx] = e;
        for (size_t i = 0; i < s->next.size(); ++i)
            if (!s->next[i]->is_symbol) visit(succ[i], next_shape(op, shape, i), s->next[i], rule);
    }
    // GeometryUtils::orderVertices_CCW: the anchor is the vertex of least y
    // (then least x); the others follow by the angle acos(d.x / |d|) of
    // d = v - anchor, ascending (FP32, as OrderByAngle computes it; equal
    // angles keep their order)
    static std::vector<std::pair<float, float>> order_ccw(const std::vector<std::pair<float, float>>& in) {
        size_t a = 0;
        for (size_t j = 1; j < in.size(); ++j)
            if (in[j].second < in[a].second || (in[j].second == in[a].second && in[j].first < in[a].first)) a = j;
        const float ax = in[a].first, ay = in[a].second;
        auto angle = [&](const std::pair<float, float>& v) {
            const float dx = v.first - ax, dy = v.second - ay;
            const float l = std::sqrt(dx * dx + dy * dy);
            return std::acos(dx / l);
        };
        std::vector<std::pair<float, float>> out;
        for (size_t j = 0; j < in.size(); ++j)
            if (j != a) out.push_back(in[j]);
        std::stable_sort(out.begin(), out.end(), [&](const auto& p, const auto& q) { return angle(p) < angle(q); });
        out.insert(out.begin(), in[a]);
        return out;
    }
};

// ---- operator graph and partition code ----
// The graph of the operators reachable from the first axiom
// (Context::Context, src/compiler/Context.cpp:11-253, in the library's own
// numbering). It has one vertex per rule body and per inline operator
// occurrence, one Generate vertex per terminal symbol (or per inline
// Generate()), and one per Discard(). Its edges are the successor links,
// numbered in depth-first pre-order. An edge that closes a cycle is marked:
// a partition must cut it.
struct GVertex {
    const Successor* node = nullptr;   // the operator (nullptr: a terminal symbol's Generate)
    std::string terminal;              // terminal symbol (Generate vertices)
    std::string rule;                  // enclosing rule (an inline Generate() takes its declaration)
    int shape = 0;
    std::vector<int> out;              // edges, in successor order
};
struct GEdge {
    int from = -1, to = -1;
    bool cycle = false;
    bool phase = false;      // enters an IfCollides that starts a later phase: a partition must cut it
};

class GraphBuilder {
public:
    explicit GraphBuilder(const Parser& parser) : g_(parser) {}

    std::vector<GVertex> V;
    std::vector<GEdge> E;
    std::vector<int> phase;   // per vertex: the phase its operator runs in (-1: none assigned)
    int nphases = 1;

    void run() {
        if (g_.axioms.empty()) throw CompileError("no axiom");
        const Axiom& a = g_.axioms[0];
        symbol_vertex(a.symbol, a.shape, 0);
        compute_phases();
    }

private:
    const Parser& g_;
    std::map<std::string, int> named_;   // rule or terminal symbol -> vertex
    std::vector<char> on_stack_;

    int op_of(int v) const { return V[v].node ? V[v].node->op : PGA_OP_GENERATE; }
    unsigned tag_of(int v) const {
        const Successor* s = V[v].node;
        if (s->params.empty() || s->params[0].kind != OD_NUMBER)
            throw CompileError(at_line(s->line) + operator_name(s->op) + " takes one number, its collider tag");
        return (unsigned)s->params[0].a;
    }

    // Graph::computePhases (src/compiler/Graph.cpp:149-272): a walk from the
    // axiom's operator, each operator in its predecessor's phase, except an
    // IfCollides(tag), which waits until every Collider(tag) has been reached
    // and then runs one phase after the latest of them; the edge into it
    // must be cut (a dispatch into that phase). The walk order is the
    // reference's: a deque whose front is taken, the children of an operator
    // pushed to the front in successor order, an IfCollides put back on a
    // waiting list (once) while colliders are missing, the waiting list's
    // front resumed when the deque runs dry. An IfCollides still waiting
    // for its colliders after that one retry keeps no phase in the
    // reference (which reports it only in verbose builds); here it is an
    // error, since its table would test against colliders of its own phase.
    void compute_phases() {
        phase.assign(V.size(), -1);
        std::map<unsigned, std::vector<int>> colliders;
        for (size_t v = 0; v < V.size(); ++v)
            if (op_of((int)v) == PGA_OP_COLLIDER) colliders[tag_of((int)v)].push_back((int)v);
        if (op_of(0) == PGA_OP_IFCOLLIDES) throw CompileError("the first operator can not be an IfCollides");
        struct Item { int src, dst, edge; };
        std::deque<Item> stack, waiting;
        std::set<int> visited;
        std::map<int, unsigned> pops;
        std::vector<int> unresolved;
        auto resume = [&]() {
            if (!stack.empty() || waiting.empty()) return;
            const Item w = waiting.front();
            waiting.pop_front();
            ++pops[w.dst];
            stack.push_front(w);
        };
        stack.push_back({-1, 0, -1});
        while (!stack.empty()) {
            const Item cur = stack.front();
            stack.pop_front();
            const int pred = cur.src >= 0 ? phase[cur.src] : 0;
            if (op_of(cur.dst) == PGA_OP_IFCOLLIDES) {
                auto it = colliders.find(tag_of(cur.dst));
                if (it != colliders.end()) {
                    int latest = -1;
                    bool missing = false;
                    for (int c : it->second) {
                        latest = std::max(latest, phase[c]);
                        missing |= visited.count(c) == 0;
                    }
                    if (missing) {
                        if (pops[cur.dst] < 1) waiting.push_back(cur);
                        else unresolved.push_back(cur.dst);
                        resume();
                        continue;
                    }
                    phase[cur.dst] = latest + 1;
                    if (cur.edge >= 0) E[cur.edge].phase = true;
                }
            } else {
                phase[cur.dst] = pred;
            }
            visited.insert(cur.dst);
            for (int e : V[cur.dst].out)
                if (!visited.count(E[e].to)) stack.push_front({cur.dst, E[e].to, e});
            resume();
        }
        for (int v : unresolved)
            if (phase[v] < 0)
                throw CompileError(at_line(V[v].node->line) + "IfCollides(" + std::to_string(tag_of(v)) +
                                   ") is reached before its colliders in the phase walk: list the branch that "
                                   "stores them first");
        for (int p : phase) nphases = std::max(nphases, p + 1);
        if (nphases > (int)PGA_MAX_PHASES) throw CompileError("more than 16 phases");
    }

    std::string resolve(const std::string& name) const {
        std::string s = name;
        for (size_t guard = 0; guard <= g_.rules.size(); ++guard) {
            auto it = g_.rules.find(s);
            if (it == g_.rules.end() || !it->second->is_symbol) return s;
            s = it->second->symbol;
        }
        throw CompileError("the rules aliasing '" + name + "' form a cycle");
    }
    int add_vertex(const Successor* node, const std::string& terminal, const std::string& rule, int shape) {
        GVertex v;
        v.node = node;
        v.terminal = terminal;
        v.rule = rule;
        v.shape = shape;
        V.push_back(v);
        on_stack_.push_back(0);
        return (int)V.size() - 1;
    }
    int symbol_vertex(const std::string& name, int shape, int line) {
        const std::string s = resolve(name);
        auto it = named_.find(s);
        if (it != named_.end()) {
            if (V[it->second].shape != shape)
                throw CompileError(at_line(line) + "'" + s + "' expects shapes of type " +
                                   shape_name(V[it->second].shape) + " but also gets shapes of type " +
                                   shape_name(shape));
            return it->second;
        }
        auto r = g_.rules.find(s);
        if (r == g_.rules.end()) {
            const int v = add_vertex(nullptr, s, s, shape);
            named_[s] = v;
            return v;
        }
        const int v = add_vertex(r->second.get(), "", s, shape);
        named_[s] = v;
        expand(v);
        return v;
    }
    void expand(int v) {
        on_stack_[v] = 1;
        const Successor* s = V[v].node;
        const std::string rule = V[v].rule;
        const int shape = V[v].shape;
        for (size_t i = 0; i < s->next.size(); ++i) {
            const SuccessorP& n = s->next[i];
            const int ns = next_shape_of(s->op, shape, i);
            const int e = (int)E.size();
            E.push_back(GEdge());
            E[e].from = v;
            V[v].out.push_back(e);
            int w;
            if (n->is_symbol) {
                w = symbol_vertex(n->symbol, ns, n->line);
            } else {
                w = add_vertex(n.get(), "", rule, ns);
                if (n->op != PGA_OP_DISCARD && n->op != PGA_OP_GENERATE) expand(w);
            }
            E[e].to = w;
            E[e].cycle = on_stack_[w] != 0;
        }
        on_stack_[v] = 0;
    }
};

// Partition code in the form CodeGenerator::fromPartition writes
// (src/compiler/CodeGenerator.cpp:1691-1722; e.g. src/test_runner/*.cuh): one
// `struct P<k> : Proc<Shape, operator tree, 1>` per subgraph, the procedure
// list and a DispatchTable with one entry per procedure. A subgraph is a
// vertex reached by a cut edge (or the axiom's) and every vertex below it
// through uncut edges, written inline. A cut edge becomes
// FSCall<procedure, 0, -1, entry>, a call to that subgraph's own entry.
// Every operand is a static template parameter: Scalar<V, N> with
// V / 10^N equal in FP32 to the grammar's value, AxisParam, RepeatModeParam,
// Rand, the shape attributes, Vec2 and Exp<A, B, Op>.
class CodeGen {
public:
    CodeGen(const Parser& p, const GraphBuilder& gb, const std::string& cuts)
        : g_(p), V_(gb.V), E_(gb.E), phase_(gb.phase), nphases_(gb.nphases), cut_(gb.E.size(), 0) {
        if (!cuts.empty() && cuts.size() != E_.size())
            throw CompileError("the partition names " + std::to_string(cuts.size()) + " edges; the graph has " +
                               std::to_string(E_.size()));
        for (size_t e = 0; e < E_.size(); ++e) {
            const bool c = cuts.empty() ? true : cuts[e] == '1';
            if (!cuts.empty() && cuts[e] != '0' && cuts[e] != '1')
                throw CompileError("a partition is a string of '0' (keep) and '1' (cut), one per edge");
            if (E_[e].cycle && !c)
                throw CompileError("edge " + std::to_string(e) + " closes a cycle and must be cut");
            if (E_[e].phase && !c)
                throw CompileError("edge " + std::to_string(e) + " enters an IfCollides of a later phase and must be cut");
            cut_[e] = c;
        }
        proc_of_.assign(V_.size(), -1);
        roots_.push_back(0);
        proc_of_[0] = 0;
        for (size_t e = 0; e < E_.size(); ++e) {
            const int w = E_[e].to;
            if (cut_[e] && proc_of_[w] < 0) {
                proc_of_[w] = (int)roots_.size();
                roots_.push_back(w);
            }
        }
    }

    std::string run(const std::string& name) {
        std::string o;
        o += "// Partition code: " + std::to_string(roots_.size()) + " procedures over " +
             std::to_string(V_.size()) + " operators and " + std::to_string(E_.size()) + " edges; partition ";
        for (size_t e = 0; e < E_.size(); ++e) o += cut_[e] ? '1' : '0';
        o += "\nnamespace " + name + " {\n\n";
        for (size_t k = 0; k < roots_.size(); ++k)
            o += "struct P" + std::to_string(k) + " : public Proc<" + shape_type(V_[roots_[k]].shape) + ", " +
                 tree(roots_[k]) + ", 1> {};\n";
        o += "\nstruct ProcedureList : T::List<";
        for (size_t k = 0; k < roots_.size(); ++k) o += (k ? ", P" : "P") + std::to_string(k);
        o += "> {};\n\nDispatchTable dispatchTable = {\n";
        for (size_t k = 0; k < roots_.size(); ++k)
            o += "    /* entries[" + std::to_string(k) + "]= */ {" + std::to_string(k) + ", {}, {}},\n";
        o += "};\n\nconst unsigned int AxiomEntry = 0;\nconst unsigned int AxiomShape = " +
             std::to_string(V_[0].shape) + ";\nconst unsigned int NumPhases = " + std::to_string(nphases_) +
             ";\n\n} // namespace " + name + "\n";
        return o;
    }

private:
    const Parser& g_;
    const std::vector<GVertex>& V_;
    const std::vector<GEdge>& E_;
    const std::vector<int>& phase_;
    int nphases_;
    std::vector<char> cut_;
    std::vector<int> proc_of_;
    std::vector<int> roots_;

    static std::string shape_type(int s) {
        if (s == PGA_SHAPE_BOX) return "Box";
        if (s == PGA_SHAPE_QUAD) return "Quad";
        throw CompileError("only Box and Quad procedures are generated");
    }
    static const char* axis_name(int a) { return a == 0 ? "X" : (a == 1 ? "Y" : "Z"); }
    // Scalar<V, N>: the smallest N with V / 10^N (FP32, as the template
    // evaluates it) equal to the FP32 value
    static std::string fixed(double x, std::string* nspart = nullptr) {
        const float want = (float)x;
        for (int n = 0; n <= 9; ++n) {
            const double scaled = std::nearbyint((double)want * std::pow(10.0, n));
            if (std::fabs(scaled) > 2147483647.0) break;
            const int v = (int)scaled;
            float p = 1.0f;
            for (int k = 0; k < n; ++k) p *= 10.0f;
            if (v / p == want) {
                if (nspart) *nspart = std::to_string(n);
                return std::to_string(v);
            }
        }
        throw CompileError("the value " + std::to_string(x) + " has no Scalar<V, N> form");
    }
    static std::string scalar(double x) {
        std::string n;
        const std::string v = fixed(x, &n);
        return "Scalar<" + v + ", " + n + ">";
    }
    static std::string operation(int op) {
        static const char* names[] = {"Add", "Sub", "Div", "Multi", "Eq", "Neq", "Lt", "Gt", "Leq", "Geq", "And", "Or"};
        if (op < 0 || op > 11) throw CompileError("unknown operation");
        return names[op];
    }
    static std::string operand(const Operand& o, bool in_exp, int line) {
        switch (o.kind) {
        case OD_NUMBER: return scalar(o.a);
        case OD_AXIS: return std::string("AxisParam<") + axis_name((int)o.a) + ">";
        case OD_MODE: {
            static const char* m[] = {"ANCHOR_TO_START", "ANCHOR_TO_END", "ADJUST_TO_FILL"};
            return std::string("RepeatModeParam<") + m[(int)o.a] + ">";
        }
        case OD_RAND: {
            std::string na, nb;
            const std::string a = fixed(o.a, &na), b = fixed(o.b, &nb);
            return "Rand<" + a + ", " + b + ", " + na + ", " + nb + ">";
        }
        case OD_VEC2: {
            std::string nx, ny;
            const std::string x = fixed(o.a, &nx), y = fixed(o.b, &ny);
            return "Vec2<" + x + ", " + y + ", " + nx + ", " + ny + ">";
        }
        case OD_ATTR:
            switch (o.attr) {
            case AT_POS: return std::string("ShapePos<") + axis_name(o.axis) + ">";
            case AT_SIZE: return std::string("ShapeSize<") + axis_name(o.axis) + ">";
            case AT_SEED: return "ShapeSeed";
            case AT_CUSTOM: return "ShapeCustomAttribute";
            case AT_ROTATION:
                if (in_exp) return std::string("ShapeRotation<") + axis_name(o.axis) + ", " + axis_name(o.component) + ">";
                break;
            case AT_NORMAL:
                if (in_exp) return std::string("ShapeNormal<") + axis_name(o.axis) + ">";
                break;
            default: break;
            }
            throw CompileError(at_line(line) + "ShapeRotation and ShapeNormal are only valid inside Exp(...)");
        case OD_EXP:
            return "Exp<" + operand(*o.left, true, line) + ", " + operand(*o.right, true, line) + ", " +
                   operation(o.op) + ">";
        }
        throw CompileError("unknown operand");
    }
    const std::vector<double>* terminal_of(const std::string& name) const {
        auto it = g_.terminals.find(name);
        return it == g_.terminals.end() ? nullptr : &it->second;
    }
    std::string generate(const std::vector<double>* attrs, int line) const {
        int gen = 0;
        std::vector<double> rest;
        if (attrs && !attrs->empty()) {
            gen = (int)(*attrs)[0];
            rest.assign(attrs->begin() + 1, attrs->end());
        }
        if (rest.empty()) rest.push_back(0.0);
        if (rest.size() > 6) throw CompileError(at_line(line) + "too many terminal attributes");
        std::string o = "Generate<false, " + std::to_string(gen);
        for (double x : rest) o += ", " + scalar(x);
        return o + ">";
    }
    std::string succ(int v, size_t i) {
        const int e = V_[v].out[i];
        const int w = E_[e].to;
        if (cut_[e]) {   // FSCall<procedure, its phase, -1, entry> (CodeGenerator.cpp:1161)
            const std::string k = std::to_string(proc_of_[w]);
            return "FSCall<" + k + ", " + std::to_string(std::max(phase_[w], 0)) + ", -1, " + k + ">";
        }
        return tree(w);
    }
    std::string tree(int v) {
        const GVertex& x = V_[v];
        if (!x.node) return generate(terminal_of(x.terminal), 0);
        const Successor* s = x.node;
        const int ln = s->line;
        auto P = [&](size_t k) {
            if (k >= s->params.size())
                throw CompileError(at_line(ln) + operator_name(s->op) + " requires at least " + std::to_string(k + 1) +
                                   " parameters");
            return operand(s->params[k], false, ln);
        };
        auto need_succ = [&](size_t n) {
            if (s->next.size() != n)
                throw CompileError(at_line(ln) + operator_name(s->op) + " requires " + std::to_string(n) +
                                   " successors");
        };
        switch (s->op) {
        case PGA_OP_DISCARD: return "Discard";
        case PGA_OP_GENERATE: return generate(terminal_of(x.rule), ln);
        case PGA_OP_TRANSLATE: case PGA_OP_ROTATE: case PGA_OP_SCALE: case PGA_OP_SWAPSIZE: {
            need_succ(1);
            static const std::map<int, const char*> names = {{PGA_OP_TRANSLATE, "Translate"}, {PGA_OP_ROTATE, "Rotate"},
                                                             {PGA_OP_SCALE, "Scale"}, {PGA_OP_SWAPSIZE, "SwapSize"}};
            return std::string(names.at(s->op)) + "<" + P(0) + ", " + P(1) + ", " + P(2) + ", " + succ(v, 0) + ">";
        }
        case PGA_OP_EXTRUDE:
            need_succ(1);
            return "Extrude<AxisParam<Z>, " + P(0) + ", " + succ(v, 0) + ">";
        case PGA_OP_COMPSPLIT:
            need_succ(3);
            return "ComponentSplit<false, " + succ(v, 0) + ", " + succ(v, 1) + ", " + succ(v, 2) + ">";
        case PGA_OP_SUBDIV: {
            if (s->factors.size() != s->next.size() || s->next.empty())
                throw CompileError(at_line(ln) + "SubDiv requires a factor before each successor");
            std::string o = "Subdivide<" + P(0);
            for (size_t i = 0; i < s->next.size(); ++i) o += ", T::Pair<" + scalar(s->factors[i]) + ", " + succ(v, i) + ">";
            return o + ">";
        }
        case PGA_OP_REPEAT:
            need_succ(1);
            return "Repeat<false, " + P(0) + ", " + P(1) + ", " + P(2) + ", " + succ(v, 0) + ">";
        case PGA_OP_IF:
            need_succ(2);
            return "If<" + P(0) + ", " + succ(v, 0) + ", " + succ(v, 1) + ">";
        case PGA_OP_IFSIZELESS:
            need_succ(2);
            return "IfSizeLess<" + P(0) + ", " + P(1) + ", " + succ(v, 0) + ", " + succ(v, 1) + ">";
        case PGA_OP_IFCOLLIDES:
            need_succ(2);
            return "IfCollides<" + P(0) + ", " + succ(v, 0) + ", " + succ(v, 1) + ">";
        case PGA_OP_COLLIDER:
            need_succ(1);
            return "Collider<" + P(0) + ", " + succ(v, 0) + ">";
        case PGA_OP_STOCHASTIC: {
            if (s->factors.size() != s->next.size() || s->next.empty())
                throw CompileError(at_line(ln) + "RandomRule requires a factor before each successor");
            std::string o = "RandomRule<" + P(0);
            for (size_t i = 0; i < s->next.size(); ++i) o += ", T::Pair<" + scalar(s->factors[i]) + ", " + succ(v, i) + ">";
            return o + ">";
        }
        case PGA_OP_REPLICATE: {
            std::string o = "Replicate<";
            for (size_t i = 0; i < s->next.size(); ++i) o += (i ? ", " : "") + succ(v, i);
            return o + ">";
        }
        default:
            throw CompileError(at_line(ln) + operator_name(s->op) + " is not supported");
        }
    }
};

} // namespace

struct pga_grammar {
    std::vector<pga_dispatch_entry> entries;
    std::vector<pga_grammar_axiom> axioms;
    std::string code;                 // partition code (pga_generate_partition)
    std::vector<uint8_t> cycle;       // per operator-graph edge: 1 = closes a cycle, 2 = enters a later
                                      // phase's IfCollides (either: a partition must cut it)
    std::string error;
};

extern "C" {

PGA_API pga_status pga_interpreter_procedures(pga_procedure* out, uint32_t capacity, uint32_t* count) {
    if (!count) return PGA_ERR_INVALID_ARG;
    *count = kNumInterpreterProcedures;
    if (!out) return PGA_OK;
    if (capacity < kNumInterpreterProcedures) return PGA_ERR_INVALID_ARG;
    std::memcpy(out, kInterpreterProcedures, sizeof(kInterpreterProcedures));
    return PGA_OK;
}

PGA_API pga_status pga_compile_grammar(const char* source, pga_grammar** out) {
    if (!out) return PGA_ERR_INVALID_ARG;
    *out = nullptr;
    pga_grammar* g = new (std::nothrow) pga_grammar();
    if (!g) return PGA_ERR_INVALID_ARG;
    *out = g;
    if (!source) {
        g->error = "null source";
        return PGA_ERR_INVALID_ARG;
    }
    try {
        const std::string src(source);
        Parser p(src);
        p.parse();
        Generator gen(p);
        gen.run(g->axioms);
        g->entries.swap(gen.entries);
    } catch (const std::exception& e) {
        g->error = e.what();
        g->entries.clear();
        g->axioms.clear();
        return PGA_ERR_BAD_PROGRAM;
    }
    return PGA_OK;
}

PGA_API const char* pga_grammar_error(const pga_grammar* g) { return g ? g->error.c_str() : "null grammar"; }

PGA_API pga_status pga_grammar_entries(const pga_grammar* g, const pga_dispatch_entry** entries, uint32_t* count) {
    if (!g || !entries || !count) return PGA_ERR_INVALID_ARG;
    *entries = g->entries.data();
    *count = (uint32_t)g->entries.size();
    return PGA_OK;
}

PGA_API pga_status pga_grammar_axioms(const pga_grammar* g, const pga_grammar_axiom** axioms, uint32_t* count) {
    if (!g || !axioms || !count) return PGA_ERR_INVALID_ARG;
    *axioms = g->axioms.data();
    *count = (uint32_t)g->axioms.size();
    return PGA_OK;
}

PGA_API void pga_grammar_destroy(pga_grammar* g) { delete g; }

PGA_API pga_status pga_generate_partition(const char* source, const char* partition, const char* name,
                                          pga_grammar** out) {
    if (!out) return PGA_ERR_INVALID_ARG;
    *out = nullptr;
    pga_grammar* g = new (std::nothrow) pga_grammar();
    if (!g) return PGA_ERR_INVALID_ARG;
    *out = g;
    if (!source || !name || !*name) {
        g->error = "null source or name";
        return PGA_ERR_INVALID_ARG;
    }
    for (const char* c = name; *c; ++c)
        if (!(std::isalnum((unsigned char)*c) || *c == '_') || std::isdigit((unsigned char)name[0])) {
            g->error = "the name must be a C++ identifier";
            return PGA_ERR_INVALID_ARG;
        }
    try {
        const std::string src(source);
        Parser p(src);
        p.parse();
        GraphBuilder gb(p);
        gb.run();
        for (const GEdge& e : gb.E) g->cycle.push_back((e.cycle ? 1u : 0u) | (e.phase ? 2u : 0u));
        CodeGen cg(p, gb, partition ? std::string(partition) : std::string());
        g->code = cg.run(name);
    } catch (const std::exception& e) {
        g->error = e.what();
        g->code.clear();
        return PGA_ERR_BAD_PROGRAM;
    }
    return PGA_OK;
}

PGA_API pga_status pga_grammar_code(const pga_grammar* g, const char** code, uint64_t* length) {
    if (!g || !code || !length) return PGA_ERR_INVALID_ARG;
    *code = g->code.c_str();
    *length = g->code.size();
    return PGA_OK;
}

PGA_API pga_status pga_grammar_edges(const pga_grammar* g, uint32_t* num_edges, const uint8_t** cycle_flags) {
    if (!g || !num_edges || !cycle_flags) return PGA_ERR_INVALID_ARG;
    *num_edges = (uint32_t)g->cycle.size();
    *cycle_flags = g->cycle.data();
    return PGA_OK;
}

} // extern "C"
