/*
 * pga/facade/Interpreter.h - the DSL interpreter's entry points
 * (src/interpreter/PGAFacade.h:17-21, PGAFacade.cu:354-923) on the C ABI:
 * getProcedureList(), initializePGA(dispatch table), executePGA(),
 * releasePGA(), destroyPGA(), and setAxioms for the interpreter's axioms
 * (Device::setAxioms, src/interpreter/GlobalVariables.cuh:47-59; each axiom
 * is a Box or Quad at the origin of size 1 whose seed and custom attribute
 * are Seeds[index % 32], PGAFacade.cu:684-729).
 *
 * The procedure list is the interpreter's for Box and Quad, in its order
 * (PGAFacade.cu:354-392), so dispatch tables compiled against the
 * interpreter's getProcedureList() index it unchanged. Header-only: the
 * evaluator lives in a function-local static shared by all translation
 * units.
 */
#pragma once

#include <memory>
#include <vector>

#include "../core/Core.h"

namespace PGA {
namespace Compiler {

// SingleOperatorProcedure / ProcedureList (ProcedureList.h:18-69)
struct SingleOperatorProcedure {
    int opType;
    int shapeType;
    int genFuncIdx;
};

struct ProcedureList {
    std::vector<SingleOperatorProcedure> procedures;
    int indexOf(int opType, int shapeType, int genFuncIdx = 0) const {
        for (size_t i = 0; i < procedures.size(); ++i)
            if (procedures[i].opType == opType && procedures[i].shapeType == shapeType &&
                procedures[i].genFuncIdx == genFuncIdx)
                return (int)i;
        return -1;
    }
    size_t size() const { return procedures.size(); }
};

} // namespace Compiler

namespace Interpreter {

using namespace PGA::Operators;
using namespace PGA::Parameters;
using PGA::Shapes::Box;
using PGA::Shapes::Quad;
typedef PGA::Shapes::DynamicPolygon<PGA::Constants::MaxNumSides, true> DCPoly;
typedef PGA::Shapes::DynamicPolygon<PGA::Constants::MaxNumSides, false> DPoly;
typedef PGA::Shapes::DynamicRightPrism<PGA::Constants::MaxNumSides, true> DCRPrism;
typedef PGA::Shapes::DynamicRightPrism<PGA::Constants::MaxNumSides, false> DRPrism;

// the procedures every dynamic shape shares, after its first one or two
// (PGAFacade.cu:397-470)
#define PGA_INTERPRETER_COMMON(S)                                                                               \
    Proc<S, Generate<false, 0, DynParam<0>>>, Proc<S, Generate<false, 1, DynParam<0>>>,                         \
        Proc<S, If<DynParam<0>, DCall<0>, DCall<1>>>,                                                           \
        Proc<S, IfSizeLess<DynParam<0>, DynParam<1>, DCall<0>, DCall<1>>>, Proc<S, RandomRule<DynParams>>,     \
        Proc<S, Repeat<false, DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>, Proc<S, Replicate<DynParams>>, \
        Proc<S, Rotate<DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>,                                        \
        Proc<S, Scale<DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>,                                         \
        Proc<S, SetAsDynamicConvexPolygon<DynParams>>, Proc<S, SetAsDynamicConvexRightPrism<DynParams>>,        \
        Proc<S, SetAsDynamicConcavePolygon<DynParams>>, Proc<S, SetAsDynamicConcaveRightPrism<DynParams>>,      \
        Proc<S, Subdivide<DynParams>>, Proc<S, SwapSize<DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>,      \
        Proc<S, Translate<DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>

// Quad: Discard .. Scale, four SetAs, Subdivide, SwapSize, Translate;
// Box: ComponentSplit first, no Extrude (PGAFacade.cu:354-392); then the
// dynamic polygons (Extrude) and prisms (ComponentSplit)
struct ProcList
    : T::List<
          Proc<Quad, Discard>, Proc<Quad, Extrude<DynParam<0>, DynParam<1>, DCall<0>>>,
          Proc<Quad, Generate<false, 0, DynParam<0>>>, Proc<Quad, Generate<false, 1, DynParam<0>>>,
          Proc<Quad, If<DynParam<0>, DCall<0>, DCall<1>>>,
          Proc<Quad, IfSizeLess<DynParam<0>, DynParam<1>, DCall<0>, DCall<1>>>, Proc<Quad, RandomRule<DynParams>>,
          Proc<Quad, Repeat<false, DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>, Proc<Quad, Replicate<DynParams>>,
          Proc<Quad, Rotate<DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>,
          Proc<Quad, Scale<DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>,
          Proc<Quad, SetAsDynamicConvexPolygon<DynParams>>, Proc<Quad, SetAsDynamicConvexRightPrism<DynParams>>,
          Proc<Quad, SetAsDynamicConcavePolygon<DynParams>>, Proc<Quad, SetAsDynamicConcaveRightPrism<DynParams>>,
          Proc<Quad, Subdivide<DynParams>>, Proc<Quad, SwapSize<DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>,
          Proc<Quad, Translate<DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>,
          Proc<Box, ComponentSplit<false, DCall<0>, DCall<1>, DCall<2>>>, Proc<Box, Discard>,
          Proc<Box, Generate<true, 0, DynParam<0>>, 16>, Proc<Box, Generate<false, 1, DynParam<0>>>,
          Proc<Box, If<DynParam<0>, DCall<0>, DCall<1>>>,
          Proc<Box, IfSizeLess<DynParam<0>, DynParam<1>, DCall<0>, DCall<1>>>, Proc<Box, RandomRule<DynParams>>,
          Proc<Box, Repeat<false, DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>, Proc<Box, Replicate<DynParams>>,
          Proc<Box, Rotate<DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>,
          Proc<Box, Scale<DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>,
          Proc<Box, SetAsDynamicConvexPolygon<DynParams>>, Proc<Box, SetAsDynamicConvexRightPrism<DynParams>>,
          Proc<Box, SetAsDynamicConcavePolygon<DynParams>>, Proc<Box, SetAsDynamicConcaveRightPrism<DynParams>>,
          Proc<Box, Subdivide<DynParams>>, Proc<Box, SwapSize<DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>,
          Proc<Box, Translate<DynParam<0>, DynParam<1>, DynParam<2>, DCall<0>>>,
          // the dynamic shapes of the list built with INTERPRETER_SUPPORT_ALL_SHAPES
          Proc<DCPoly, Discard>, Proc<DCPoly, Extrude<DynParam<0>, DynParam<1>, DCall<0>>>, PGA_INTERPRETER_COMMON(DCPoly),
          Proc<DPoly, Discard>, Proc<DPoly, Extrude<DynParam<0>, DynParam<1>, DCall<0>>>, PGA_INTERPRETER_COMMON(DPoly),
          Proc<DCRPrism, ComponentSplit<false, DCall<0>, DCall<1>, DCall<2>>>, Proc<DCRPrism, Discard>,
          PGA_INTERPRETER_COMMON(DCRPrism),
          Proc<DRPrism, ComponentSplit<false, DCall<0>, DCall<1>, DCall<2>>>, Proc<DRPrism, Discard>,
          PGA_INTERPRETER_COMMON(DRPrism)> {};
#undef PGA_INTERPRETER_COMMON

struct Configuration {
    static const PGA::GPU::Technique Technique = PGA::GPU::MEGAKERNEL;
    static const unsigned int QueueSize = 1024 * 32;   // PGAFacade.cu:883-893
    static const unsigned int MaxSharedMemory = 0;
};

// The interpreter's axiom record (src/interpreter/Axiom.h:8-15), Box and
// Quad only.
struct Axiom {
    PGA::Compiler::ShapeType shapeType;
    int entryIndex;
};

struct AxiomGenerator : PGA::ListedAxioms {};

typedef PGA::SinglePhaseEvaluator<ProcList, AxiomGenerator, PGA::Rendering::GenFuncFilter, false, 0, 0, Configuration>
    Evaluator;
typedef std::unique_ptr<Evaluator, PGA::ReleaseCallback> EvaluatorPtr;

inline EvaluatorPtr& evaluator() {
    static EvaluatorPtr e;
    return e;
}
inline std::vector<uint64_t>& capacities() {
    static std::vector<uint64_t> c;
    return c;
}

} // namespace Interpreter
} // namespace PGA

// the interpreter's axioms: a Box or Quad at the origin, size 1, seeded
// from the seed table by index
inline void setAxioms(const PGA::Interpreter::Axiom* axioms, unsigned int n) {
    std::vector<pga_axiom> a(n);
    for (unsigned int i = 0; i < n; ++i) {
        std::memset(&a[i], 0, sizeof(pga_axiom));
        a[i].model[0] = a[i].model[5] = a[i].model[10] = 1.0f;
        a[i].size[0] = a[i].size[1] = a[i].size[2] = 1.0f;
        a[i].seed = a[i].custom = PGA::GlobalVars::getSeed((int)i);
        a[i].shape_type = axioms[i].shapeType;
        a[i].entry_index = axioms[i].entryIndex;
    }
    PGA::Interpreter::AxiomGenerator::set(a.data(), n);
}

// capacity of the instance buffer of terminal type `terminal`
inline void setMaxNumInstances(unsigned int terminal, uint64_t capacity) {
    auto& c = PGA::Interpreter::capacities();
    if (c.size() <= terminal) c.resize(terminal + 1, 0);
    c[terminal] = capacity;
}

inline PGA::Compiler::ProcedureList getProcedureList() {
    PGA::Compiler::ProcedureList l;
    for (const auto& t : PGA::Lowering::lower<PGA::Interpreter::ProcList>()) {
        const auto& n = t.nodes[0];
        l.procedures.push_back({n.op_type, n.shape_type, n.gen_func});
    }
    return l;
}

inline void initializePGA(const std::unique_ptr<PGA::DispatchTable>& dispatchTable) {
    auto& e = PGA::Interpreter::evaluator();
    if (!e) e = PGA::Interpreter::EvaluatorPtr(new PGA::Interpreter::Evaluator());
    const auto& c = PGA::Interpreter::capacities();
    for (unsigned int t = 0; t < c.size(); ++t)
        if (c[t]) e->setMaxNumInstances(t, c[t]);
    e->initialize(dispatchTable->toDispatchTableEntriesPtr().get(), dispatchTable->entries.size());
}

inline double executePGA() { return PGA::Interpreter::evaluator()->execute(); }

// as the reference: gives the evaluator up without releasing it
// (PGAFacade.cu:913-916); destroyPGA releases it
inline void releasePGA() { (void)PGA::Interpreter::evaluator().release(); }

inline void destroyPGA() { PGA::Interpreter::evaluator().reset(); }

namespace PGA {
namespace Interpreter {

// a library-written entry back in the DispatchTable form
inline PGA::DispatchTable::Entry toEntry(const pga_dispatch_entry& e) {
    PGA::DispatchTable::Entry x;
    x.operatorCode = (size_t)e.operator_code;
    for (uint32_t k = 0; k < e.num_parameters && k < PGA_MAX_NUM_PARAMETERS; ++k) {
        const pga_dispatch_parameter& p = e.parameters[k];
        size_t n = PGA_MAX_NUM_PARAMETER_VALUES;
        while (n > 0 && p.value[n - 1] == 0.0f && !std::signbit(p.value[n - 1])) --n;
        x.parameters.push_back(
            PGA::DispatchTable::Parameter((PGA::ParameterType)p.type, std::vector<double>(p.value, p.value + n)));
    }
    for (uint32_t k = 0; k < e.num_successors && k < PGA_MAX_NUM_SUCCESSORS; ++k)
        x.successors.push_back({e.successors[k].entry_index, e.successors[k].phase_index});
    for (uint32_t k = 0; k < e.num_edge_indices && k < PGA_MAX_NUM_SUCCESSORS; ++k)
        x.edgeIndices.push_back(e.edge_indices[k]);
    x.subGraphIndex = e.sub_graph_index;
    return x;
}

} // namespace Interpreter
} // namespace PGA

// InterpreterApp::parse (src/interpreter/InterpreterApp.cpp:175-231) without
// the rendering side: the grammar text compiled by the library's front end
// (pga_compile_grammar) into a dispatch table over getProcedureList(), and
// the declared axioms (an axiom whose symbol has no rule gets entry -1 and
// is skipped). Throws std::runtime_error carrying the compiler's message,
// which names the line, where the reference prints its parser log and throws
// "parsing error".
inline std::unique_ptr<PGA::DispatchTable> compileGrammar(const std::string& source,
                                                          std::vector<PGA::Interpreter::Axiom>& axioms) {
    pga_grammar* g = nullptr;
    const pga_status st = pga_compile_grammar(source.c_str(), &g);
    std::unique_ptr<pga_grammar, void (*)(pga_grammar*)> guard(g, pga_grammar_destroy);
    if (st != PGA_OK) throw std::runtime_error(std::string("parsing error: ") + pga_grammar_error(g));
    const pga_dispatch_entry* entries = nullptr;
    const pga_grammar_axiom* ax = nullptr;
    uint32_t n = 0, m = 0;
    if (pga_grammar_entries(g, &entries, &n) != PGA_OK || pga_grammar_axioms(g, &ax, &m) != PGA_OK)
        throw std::runtime_error("parsing error: no compiled grammar");
    std::unique_ptr<PGA::DispatchTable> table(new PGA::DispatchTable());
    for (uint32_t i = 0; i < n; ++i) table->entries.push_back(PGA::Interpreter::toEntry(entries[i]));
    axioms.clear();
    for (uint32_t i = 0; i < m; ++i) axioms.push_back({(PGA::Compiler::ShapeType)ax[i].shape_type, ax[i].entry_index});
    return table;
}
