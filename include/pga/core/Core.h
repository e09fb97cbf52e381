/*
 * pga/core/Core.h - the reference's C++ grammar idiom over the C ABI (pga.h).
 *
 * Header-only C++17, no HIP types. A grammar written for the reference -
 * a procedure list `struct ProcedureList : T::List<Proc<Shape, Operator>...>`
 * (Proc.cuh:20-101, TStdLib.h:262-350), a `DispatchTable` initialiser
 * (DispatchTable.h:16-104), an axiom generator (Grid.cuh:113-284) and
 * `SinglePhaseEvaluator<...>` (GPUEvaluator.cuh:475-536) - compiles against
 * these names and arities and runs on the GPU engine:
 *
 *   1. every Proc's operator tree is lowered, from its template arguments,
 *      to a procedure template: one node per operator, each parameter either
 *      a constant (Scalar, AxisParam, Rand, ShapeAttr, Exp, Vec2, ...) or a
 *      reference to the dispatching entry's parameter (DynParam<k>, the whole
 *      entry for the <DynParams> forms), each successor an inline node or a
 *      call (DCall, PSCall, FSCall);
 *   2. Evaluator::initialize expands the templates over the dispatch table:
 *      table entry i becomes expanded entry i (its procedure's root
 *      operator), inline operators and calls that name a procedure other
 *      than their target entry's become further entries; calls resolve as
 *      the reference's SymbolDecorator does (SymbolDecorator.cuh:16-123):
 *      DCall<s> and PSCall<op, s> go to the current entry's successor s,
 *      FSCall<op, ...> to its fixed entry or, without one, stay on the
 *      current entry;
 *   3. the expanded table - one operator per entry, the form the engine
 *      runs - goes to pga_initialize with the procedure list of the
 *      distinct (operator, shape, generation function, static form) kinds.
 *      A call's successor keeps the phase its dispatch names, an inline
 *      operator's runs in the phase of the shape reaching it
 *      (PGA_PHASE_INHERIT), so `MultiPhaseEvaluator<...>` (with IfCollides
 *      and Collider, GPUEvaluator.cuh:395-456) runs the same tables.
 *
 * Dynamic polygons and right prisms derive (SetAs*, a prism's
 * ComponentSplit, a polygon's Extrude), except through generation function
 * 0 and the collision operators; the static shapes (Triangle .. Octagon,
 * Prism3 .. Prism8, Sphere) lower but are refused by pga_initialize
 * (PGA_ERR_UNSUPPORTED), as the engine does not derive them.
 */
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../pga.h"

#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif

namespace math {

struct float2 {
    float x, y;
    constexpr float2(float x_ = 0.0f, float y_ = 0.0f) : x(x_), y(y_) {}
};

struct float3 {
    float x, y, z;
    constexpr float3(float x_ = 0.0f, float y_ = 0.0f, float z_ = 0.0f) : x(x_), y(y_), z(z_) {}
};

template <class A, class B>
inline auto pow(A a, B b) -> decltype(std::pow(a, b)) { return std::pow(a, b); }

} // namespace math

namespace PGA {

// ---- enumerations (Axis.h, RepeatMode.h, ParameterType.h, OperationType.h,
//      OperandType.h, GPUTechnique.h, compiler/ShapeType.h, compiler/OperatorType.h)
enum Axis { X = 0, Y, Z };
enum RepeatMode { ANCHOR_TO_START = 0, ANCHOR_TO_END, ADJUST_TO_FILL };
enum ParameterType { PT_SCALAR = 0, PT_RAND, PT_SHAPE_ATTR, PT_EXP, PT_VEC2, PT_VEC4 };
enum OperationType { OPT_ADD = 0, OPT_SUB, OPT_DIV, OPT_MULTI, OPT_EQ, OPT_NEQ, OPT_LT, OPT_GT, OPT_LEQ, OPT_GEQ,
                     OPT_AND, OPT_OR };
enum OperandType { ORT_SCALAR = 0, ORT_RAND, ORT_SHAPE_ATTR, ORT_OP };

namespace GPU {
enum Technique { KERNELS, DYN_PAR, MEGAKERNEL, MEGAKERNEL_LOCAL_QUEUES };
}

namespace Compiler {
enum ShapeType { TRIANGLE, QUAD, PENTAGON, HEXAGON, HEPTAGON, OCTAGON, DYNAMIC_CONVEX_POLYGON, DYNAMIC_POLYGON,
                 PRISM3, BOX, PRISM5, PRISM6, PRISM7, PRISM8, DYNAMIC_CONVEX_RIGHT_PRISM, DYNAMIC_RIGHT_PRISM, SPHERE };
}

namespace Constants {
const unsigned int MaxNumParameters = PGA_MAX_NUM_PARAMETERS;
const unsigned int MaxNumSides = 16;
const unsigned int MaxNumExpLevels = 4;
const unsigned int MaxNumParameterValues = PGA_MAX_NUM_PARAMETER_VALUES;
const unsigned int MaxNumSuccessors = PGA_MAX_NUM_SUCCESSORS;
const unsigned int MaxNumAxioms = PGA_MAX_NUM_AXIOM_SEEDS;
} // namespace Constants

// ---- compile-time lists (TStdLib.h:138, 262) ----
namespace T {
template <class... ItemsT> struct List {};
template <class FactorT, class NextT> struct Pair {};
template <int BaseT, unsigned int ExpT> struct Power {
    static const long long Result = BaseT * Power<BaseT, ExpT - 1>::Result;
};
template <int BaseT> struct Power<BaseT, 0> {
    static const long long Result = 1;
};
} // namespace T

// ---- dispatch table (DispatchTableEntry.h:8-36, DispatchTable.h:16-104) ----
struct DispatchTableEntry {
    struct Parameter {
        float value[Constants::MaxNumParameterValues];
        int type;
    };
    struct Successor {
        int entryIndex;
        int phaseIndex;
    };
    unsigned int operatorCode;
    unsigned int numParameters;
    Parameter parameters[Constants::MaxNumParameters];
    unsigned int numSuccessors;
    unsigned char ruleTagId;
    Successor successors[Constants::MaxNumSuccessors];
    unsigned int numEdgeIndices;
    int edgeIndices[Constants::MaxNumSuccessors];
    int subGraphIndex;
};
static_assert(sizeof(DispatchTableEntry) == sizeof(pga_dispatch_entry), "entry layout");
static_assert(offsetof(DispatchTableEntry, parameters) == offsetof(pga_dispatch_entry, parameters), "entry layout");
static_assert(offsetof(DispatchTableEntry, successors) == offsetof(pga_dispatch_entry, successors), "entry layout");
static_assert(offsetof(DispatchTableEntry, edgeIndices) == offsetof(pga_dispatch_entry, edge_indices), "entry layout");
static_assert(offsetof(DispatchTableEntry, subGraphIndex) == offsetof(pga_dispatch_entry, sub_graph_index),
              "entry layout");

struct DispatchTable {
    struct Parameter {
        ParameterType type = PT_SCALAR;
        std::vector<double> values;
        Parameter() = default;
        Parameter(ParameterType t, const std::vector<double>& v) : type(t), values(v) {}
        Parameter(ParameterType t, std::initializer_list<double> v) : type(t), values(v) {}
        void fill(DispatchTableEntry::Parameter& p) const {
            if (values.size() > Constants::MaxNumParameterValues)
                throw std::runtime_error("values.size() > PGA::Constants::MaxNumParameterValues");
            p.type = type;
            for (unsigned i = 0; i < Constants::MaxNumParameterValues; ++i)
                p.value[i] = i < values.size() ? static_cast<float>(values[i]) : 0.0f;
        }
    };
    struct Successor {
        int entryIndex = -1;
        int phaseIndex = -1;
        Successor() = default;
        Successor(int e, int p) : entryIndex(e), phaseIndex(p) {}
    };
    struct Entry {
        size_t operatorCode = 0;
        std::vector<Parameter> parameters;
        std::vector<Successor> successors;
        std::vector<int> edgeIndices;
        int subGraphIndex = -1;
        Entry() = default;
        Entry(size_t op, std::initializer_list<Parameter> params, std::initializer_list<Successor> succ,
              std::initializer_list<int> edges = {}, int subGraph = -1)
            : operatorCode(op), parameters(params), successors(succ), edgeIndices(edges), subGraphIndex(subGraph) {}
        void fill(DispatchTableEntry& e) const {
            std::memset(&e, 0, sizeof(e));
            if (parameters.size() > Constants::MaxNumParameters)
                throw std::runtime_error("parameters.size() > PGA::Constants::MaxNumParameters");
            if (successors.size() > Constants::MaxNumSuccessors || edgeIndices.size() > Constants::MaxNumSuccessors)
                throw std::runtime_error("successors.size() > PGA::Constants::MaxNumSuccessors");
            e.operatorCode = static_cast<unsigned int>(operatorCode);
            e.numParameters = static_cast<unsigned int>(parameters.size());
            for (size_t i = 0; i < parameters.size(); ++i) parameters[i].fill(e.parameters[i]);
            e.numSuccessors = static_cast<unsigned int>(successors.size());
            for (unsigned i = 0; i < Constants::MaxNumSuccessors; ++i) {
                e.successors[i].entryIndex = i < successors.size() ? successors[i].entryIndex : -1;
                e.successors[i].phaseIndex = i < successors.size() ? successors[i].phaseIndex : -1;
                e.edgeIndices[i] = i < edgeIndices.size() ? edgeIndices[i] : -1;
            }
            e.numEdgeIndices = static_cast<unsigned int>(edgeIndices.size());
            e.subGraphIndex = subGraphIndex;
        }
    };

    std::vector<Entry> entries;
    DispatchTable() = default;
    DispatchTable(std::initializer_list<Entry> e) : entries(e) {}

    std::unique_ptr<DispatchTableEntry[]> toDispatchTableEntriesPtr() const {
        std::unique_ptr<DispatchTableEntry[]> p(new DispatchTableEntry[entries.empty() ? 1 : entries.size()]);
        for (size_t i = 0; i < entries.size(); ++i) entries[i].fill(p[i]);
        return p;
    }
};

// ---- shapes (the shape type each procedure's queue holds) ----
namespace Shapes {
template <int ShapeTypeT>
struct ShapeBase {
    static const int ShapeType = ShapeTypeT;
    math::float3 size{1.0f, 1.0f, 1.0f};
    math::float3 position{0.0f, 0.0f, 0.0f};
    float seed = 0.0f, customAttribute = 0.0f;
    math::float3 getSize() const { return size; }
    void setSize(const math::float3& s) { size = s; }
    math::float3 getPosition() const { return position; }
    void setPosition(const math::float3& p) { position = p; }
    float getSeed() const { return seed; }
    void setSeed(float s) { seed = s; }
    float getCustomAttribute() const { return customAttribute; }
    void setCustomAttribute(float c) { customAttribute = c; }
};
struct Triangle : ShapeBase<Compiler::TRIANGLE> {};
struct Quad : ShapeBase<Compiler::QUAD> {};
struct Pentagon : ShapeBase<Compiler::PENTAGON> {};
struct Hexagon : ShapeBase<Compiler::HEXAGON> {};
struct Heptagon : ShapeBase<Compiler::HEPTAGON> {};
struct Octagon : ShapeBase<Compiler::OCTAGON> {};
struct Prism3 : ShapeBase<Compiler::PRISM3> {};
struct Box : ShapeBase<Compiler::BOX> {};
struct Prism5 : ShapeBase<Compiler::PRISM5> {};
struct Prism6 : ShapeBase<Compiler::PRISM6> {};
struct Prism7 : ShapeBase<Compiler::PRISM7> {};
struct Prism8 : ShapeBase<Compiler::PRISM8> {};
struct Sphere : ShapeBase<Compiler::SPHERE> {};
template <unsigned int MaxNumSidesT, bool ConvexT>
struct DynamicPolygon : ShapeBase<ConvexT ? Compiler::DYNAMIC_CONVEX_POLYGON : Compiler::DYNAMIC_POLYGON> {};
template <unsigned int MaxNumSidesT, bool ConvexT>
struct DynamicRightPrism
    : ShapeBase<ConvexT ? Compiler::DYNAMIC_CONVEX_RIGHT_PRISM : Compiler::DYNAMIC_RIGHT_PRISM> {};
} // namespace Shapes

// ---- parameters (Parameters.cuh:28-800) ----
namespace Parameters {
template <int ValueT, unsigned int NegativeOrderOfMagnitudeT = 0>
struct Scalar {
    static float getValue() { return ValueT / (float)T::Power<10, NegativeOrderOfMagnitudeT>::Result; }
};
template <int XT, int YT, unsigned int XNegativeOrderOfMagnitudeT = 3, unsigned int YNegativeOrderOfMagnitudeT = 3>
struct Vec2 {
    static float getX() { return XT / (float)T::Power<10, XNegativeOrderOfMagnitudeT>::Result; }
    static float getY() { return YT / (float)T::Power<10, YNegativeOrderOfMagnitudeT>::Result; }
};
template <PGA::Axis AxisT> struct AxisParam : Scalar<AxisT, 0> {};
template <PGA::RepeatMode RepeatModeT> struct RepeatModeParam : Scalar<RepeatModeT, 0> {};
template <unsigned int IndexT> struct ShapeAttr {
    static const unsigned int Index = IndexT;
};
template <PGA::Axis AxisT> struct ShapePos : ShapeAttr<(AxisT + 1) * 4 - 1> {};
template <PGA::Axis AxisT, PGA::Axis ComponentT> struct ShapeRotation : ShapeAttr<(ComponentT * 4) + AxisT> {};
template <PGA::Axis ComponentT> struct ShapeNormal : ShapeRotation<PGA::Z, ComponentT> {};
template <PGA::Axis AxisT> struct ShapeSize : ShapeAttr<(AxisT + 12)> {};
struct ShapeSeed : ShapeAttr<15> {};
struct ShapeCustomAttribute : ShapeAttr<16> {};
template <int MinT, int MaxT, unsigned int MinNegativeOrderOfMagnitudeT = 0, unsigned int MaxNegativeOrderOfMagnitudeT = 0>
struct Rand {
    static float getMin() { return MinT / (float)T::Power<10, MinNegativeOrderOfMagnitudeT>::Result; }
    static float getMax() { return MaxT / (float)T::Power<10, MaxNegativeOrderOfMagnitudeT>::Result; }
};
template <class A, class B, OperationType OpT> struct ExpOperator {
    static const OperationType Op = OpT;
};
template <class A, class B> struct Add : ExpOperator<A, B, OPT_ADD> {};
template <class A, class B> struct Sub : ExpOperator<A, B, OPT_SUB> {};
template <class A, class B> struct Multi : ExpOperator<A, B, OPT_MULTI> {};
template <class A, class B> struct Div : ExpOperator<A, B, OPT_DIV> {};
template <class A, class B> struct Eq : ExpOperator<A, B, OPT_EQ> {};
template <class A, class B> struct Neq : ExpOperator<A, B, OPT_NEQ> {};
template <class A, class B> struct Lt : ExpOperator<A, B, OPT_LT> {};
template <class A, class B> struct Gt : ExpOperator<A, B, OPT_GT> {};
template <class A, class B> struct Leq : ExpOperator<A, B, OPT_LEQ> {};
template <class A, class B> struct Geq : ExpOperator<A, B, OPT_GEQ> {};
template <class A, class B> struct And : ExpOperator<A, B, OPT_AND> {};
template <class A, class B> struct Or : ExpOperator<A, B, OPT_OR> {};
template <class OperandAT, class OperandBT, template <class, class> class OpT> struct Exp {};
template <unsigned int ParamIdxT> struct DynParam {};
struct DynParams {};
} // namespace Parameters

// ---- operators and calls (Operators/*.cuh, Call.cuh:23-376) ----
namespace Operators {
template <class AxisT, class... ParamsT> struct Subdivide {};
template <bool ParallelT, class AxisT, class ExtentT, class RepeatModeT, class NextT> struct Repeat {};
template <bool ParallelT, class TopT, class BottomT, class SidesT> struct ComponentSplit {};
template <class ExpressionT, class IfTrueT, class IfFalseT> struct If {};
template <class AxisT, class SizeT, class IfTrueT, class IfFalseT> struct IfSizeLess {};
template <class ChanceT, class... ParamsT> struct RandomRule {};
struct Discard {};
template <class XT, class YT, class ZT, class NextT> struct Translate {};
template <class XT, class YT, class ZT, class NextT> struct Rotate {};
template <class XT, class YT, class ZT, class NextT> struct Scale {};
template <class XT, class YT, class ZT, class NextT> struct SwapSize {};
template <class AxisT, class ExtentT, class NextT> struct Extrude {};
template <class... ParamsT> struct Replicate {};
// context sensitivity (IfCollides.cuh:22-76, Collider.cuh:22-69)
template <class ColliderTagT, class NextIfTrueT, class NextIfFalseT> struct IfCollides {};
template <class ColliderTagT, class NextT> struct Collider {};
template <bool ParallelT, unsigned int GenFuncIdxT = 0, class TerminalIdxT = Parameters::Scalar<0>,
          class Attr1T = Parameters::Scalar<0>, class Attr2T = Parameters::Scalar<0>,
          class Attr3T = Parameters::Scalar<0>, class Attr4T = Parameters::Scalar<0>,
          class Attr5T = Parameters::Scalar<0>>
struct Generate {};
template <class NextT, class... VerticesT> struct SetAsDynamicConvexPolygon {};
template <class NextT, class... VerticesT> struct SetAsDynamicConvexRightPrism {};
template <class NextT, class... VerticesT> struct SetAsDynamicConcavePolygon {};
template <class NextT, class... VerticesT> struct SetAsDynamicConcaveRightPrism {};
template <unsigned int OperationCodeT, unsigned int SuccessorOffsetT, int EdgeIndexT = -1> struct PSCall {};
template <unsigned int OperationCodeT, int PhaseIndexT, int EdgeIndexT = -1, int EntryIndexT = -1> struct FSCall {};
template <unsigned int SuccessorOffsetT, int EdgeIndexT = -1> struct DCall {};
} // namespace Operators

// ---- the procedure idiom (Proc.cuh:20-27) ----
template <class ShapeT, class OperatorT, unsigned int NumThreadsT = 1>
struct Proc {
    typedef ShapeT Shape;
    typedef OperatorT Operator;
    static const unsigned int NumThreads = NumThreadsT;
};

namespace Rendering {
struct GenFuncFilter {};
}

struct ReleaseCallback {
    template <class TT> void operator()(TT* obj) {
        if (obj) obj->release();
    }
};

// ---- lowering: operator trees -> procedure templates ----
namespace Lowering {

// PackUtils.h:18-28: operand tag in bits 31:30, the float's bits shifted
// right by two; float pairs quantised to 12 bits each.
inline float pack_operand(int type, float x) {
    uint32_t r;
    std::memcpy(&r, &x, 4);
    r = (r >> 2) | ((uint32_t)(type & 3) << 30);
    float o;
    std::memcpy(&o, &r, 4);
    return o;
}
inline float pack_float2(float x, float y) {
    return std::floor(x * 4095.0f) * 4096.0f + std::floor(y * 4095.0f);
}

struct ParamRef {
    int dyn = -1;                       // >= 0: the dispatching entry's parameter dyn
    pga_dispatch_parameter value{};     // the constant otherwise
};

enum LinkKind { LINK_NODE = 0, LINK_DCALL, LINK_PSCALL, LINK_FSCALL };
struct Link {
    int kind = LINK_NODE;
    int a = -1;   // NODE: node index; DCALL: successor; PSCALL/FSCALL: procedure
    int b = -1;   // PSCALL: successor; FSCALL: fixed entry (-1: the current one)
    int phase = PGA_PHASE_INHERIT;   // FSCALL: its PhaseIndexT (the others take it from the entry)
};

struct Node {
    int op_type = 0, shape_type = 0, gen_func = 0;
    uint32_t flags = 0;
    bool whole_entry = false;           // <DynParams> form: the entry's parameters and successors
    std::vector<ParamRef> params;
    std::vector<Link> succ;
};

struct Template {
    int shape_type = 0;
    std::vector<Node> nodes;            // node 0: the procedure's root operator
};

inline ParamRef constant(int type, std::initializer_list<float> v) {
    ParamRef p;
    p.value.type = type;
    int i = 0;
    for (float x : v) p.value.value[i++] = x;
    return p;
}

// ---- parameters ----
template <class P, class = void> struct Param;   // a parameter type -> ParamRef

template <class P> struct IsShapeAttr {
    template <unsigned int I> static std::true_type test(const Parameters::ShapeAttr<I>*);
    static std::false_type test(...);
    static const bool value = decltype(test(static_cast<const P*>(nullptr)))::value;
};
template <class P> struct IsScalar {
    template <int V, unsigned N> static std::true_type test(const Parameters::Scalar<V, N>*);
    static std::false_type test(...);
    static const bool value = decltype(test(static_cast<const P*>(nullptr)))::value;
};

// expression operands, postfix (Parameters.cuh:249-258, 459-480)
template <class P, class = void> struct Operand;
template <class P> struct Operand<P, std::enable_if_t<IsScalar<P>::value>> {
    static void encode(std::vector<float>& b) { b.push_back(pack_operand(ORT_SCALAR, P::getValue())); }
};
template <class P> struct Operand<P, std::enable_if_t<IsShapeAttr<P>::value>> {
    static void encode(std::vector<float>& b) { b.push_back(pack_operand(ORT_SHAPE_ATTR, (float)P::Index)); }
};
template <int A, int B, unsigned NA, unsigned NB> struct Operand<Parameters::Rand<A, B, NA, NB>> {
    typedef Parameters::Rand<A, B, NA, NB> R;
    static void encode(std::vector<float>& b) {
        b.push_back(pack_operand(ORT_RAND, pack_float2(R::getMin(), R::getMax())));
    }
};
template <class A, class B, template <class, class> class OpT> struct Operand<Parameters::Exp<A, B, OpT>> {
    static void encode(std::vector<float>& b) {
        Operand<A>::encode(b);
        Operand<B>::encode(b);
        b.push_back(pack_operand(ORT_OP, (float)OpT<A, B>::Op));
    }
};

template <class P> struct Param<P, std::enable_if_t<IsScalar<P>::value>> {
    static ParamRef get() { return constant(PT_SCALAR, {P::getValue()}); }
};
template <class P> struct Param<P, std::enable_if_t<IsShapeAttr<P>::value>> {
    static ParamRef get() { return constant(PT_SHAPE_ATTR, {(float)P::Index}); }
};
template <int A, int B, unsigned NA, unsigned NB> struct Param<Parameters::Rand<A, B, NA, NB>> {
    typedef Parameters::Rand<A, B, NA, NB> R;
    static ParamRef get() { return constant(PT_RAND, {R::getMin(), R::getMax()}); }
};
template <int A, int B, unsigned NA, unsigned NB> struct Param<Parameters::Vec2<A, B, NA, NB>> {
    typedef Parameters::Vec2<A, B, NA, NB> V;
    static ParamRef get() { return constant(PT_VEC2, {V::getX(), V::getY()}); }
};
template <class A, class B, template <class, class> class OpT> struct Param<Parameters::Exp<A, B, OpT>> {
    static ParamRef get() {
        std::vector<float> b;
        Operand<Parameters::Exp<A, B, OpT>>::encode(b);
        if (b.size() + 1 > Constants::MaxNumParameterValues)
            throw std::runtime_error("expression longer than PGA::Constants::MaxNumParameterValues");
        ParamRef p;
        p.value.type = PT_EXP;
        p.value.value[0] = (float)b.size();
        for (size_t i = 0; i < b.size(); ++i) p.value.value[i + 1] = b[i];
        return p;
    }
};
template <unsigned int K> struct Param<Parameters::DynParam<K>> {
    static ParamRef get() {
        ParamRef p;
        p.dyn = (int)K;
        return p;
    }
};

// ---- operators ----
template <class Op, class = void> struct Lower;   // emits an operator node, returns its index

template <class Op> struct IsCall : std::false_type {};
template <unsigned O, unsigned S, int E> struct IsCall<Operators::PSCall<O, S, E>> : std::true_type {};
template <unsigned O, int P, int E, int N> struct IsCall<Operators::FSCall<O, P, E, N>> : std::true_type {};
template <unsigned S, int E> struct IsCall<Operators::DCall<S, E>> : std::true_type {};

template <class Op> struct CallLink;
template <unsigned O, unsigned S, int E> struct CallLink<Operators::PSCall<O, S, E>> {
    static Link get() { return Link{LINK_PSCALL, (int)O, (int)S}; }
};
template <unsigned O, int P, int E, int N> struct CallLink<Operators::FSCall<O, P, E, N>> {
    static Link get() { return Link{LINK_FSCALL, (int)O, N, P}; }
};
template <unsigned S, int E> struct CallLink<Operators::DCall<S, E>> {
    static Link get() { return Link{LINK_DCALL, (int)S, -1}; }
};

// a successor: a call, or an operator emitted inline on the given shape
template <class N> Link link(Template& t, int shape) {
    if constexpr (IsCall<N>::value) {
        return CallLink<N>::get();
    } else {
        return Link{LINK_NODE, Lower<N>::emit(t, shape), -1};
    }
}

inline int add_node(Template& t, int op, int shape, std::vector<ParamRef> params, int gen = 0, uint32_t flags = 0) {
    Node n;
    n.op_type = op;
    n.shape_type = shape;
    n.gen_func = gen;
    n.flags = flags;
    n.params = std::move(params);
    t.nodes.push_back(std::move(n));
    return (int)t.nodes.size() - 1;
}

inline int add_whole_entry(Template& t, int op, int shape) {
    const int i = add_node(t, op, shape, {});
    t.nodes[i].whole_entry = true;
    return i;
}

template <> struct Lower<Operators::Discard> {
    static int emit(Template& t, int shape) { return add_node(t, PGA_OP_DISCARD, shape, {}); }
};

template <class X, class Y, class Z, class N> struct Lower<Operators::Translate<X, Y, Z, N>> {
    static int emit(Template& t, int shape) {
        const int i = add_node(t, PGA_OP_TRANSLATE, shape, {Param<X>::get(), Param<Y>::get(), Param<Z>::get()});
        const Link l = link<N>(t, shape);
        t.nodes[i].succ.push_back(l);
        return i;
    }
};
template <class X, class Y, class Z, class N> struct Lower<Operators::Rotate<X, Y, Z, N>> {
    static int emit(Template& t, int shape) {
        const int i = add_node(t, PGA_OP_ROTATE, shape, {Param<X>::get(), Param<Y>::get(), Param<Z>::get()});
        const Link l = link<N>(t, shape);
        t.nodes[i].succ.push_back(l);
        return i;
    }
};
template <class X, class Y, class Z, class N> struct Lower<Operators::Scale<X, Y, Z, N>> {
    static int emit(Template& t, int shape) {
        const int i = add_node(t, PGA_OP_SCALE, shape, {Param<X>::get(), Param<Y>::get(), Param<Z>::get()});
        const Link l = link<N>(t, shape);
        t.nodes[i].succ.push_back(l);
        return i;
    }
};
template <class X, class Y, class Z, class N> struct Lower<Operators::SwapSize<X, Y, Z, N>> {
    static int emit(Template& t, int shape) {
        const int i = add_node(t, PGA_OP_SWAPSIZE, shape, {Param<X>::get(), Param<Y>::get(), Param<Z>::get()});
        const Link l = link<N>(t, shape);
        t.nodes[i].succ.push_back(l);
        return i;
    }
};
// Extrude turns a planar shape into its right prism: a Quad into a Box
template <class A, class E, class N> struct Lower<Operators::Extrude<A, E, N>> {
    static int emit(Template& t, int shape) {
        const int i = add_node(t, PGA_OP_EXTRUDE, shape, {Param<A>::get(), Param<E>::get()});
        // Operator::nextShapeType: a planar shape's extruded type
        const int to = shape == Compiler::QUAD ? (int)Compiler::BOX
                     : shape == Compiler::DYNAMIC_CONVEX_POLYGON ? (int)Compiler::DYNAMIC_CONVEX_RIGHT_PRISM
                     : shape == Compiler::DYNAMIC_POLYGON ? (int)Compiler::DYNAMIC_RIGHT_PRISM : -1;
        const Link l = link<N>(t, to);
        t.nodes[i].succ.push_back(l);
        return i;
    }
};
// Repeat<true> traps in the reference (Repeat.cuh:35-61): its procedure is
// marked parallel and the engine refuses it on load (PGA_ERR_UNSUPPORTED)
template <bool P, class A, class E, class M, class N> struct Lower<Operators::Repeat<P, A, E, M, N>> {
    static int emit(Template& t, int shape) {
        const int i = add_node(t, PGA_OP_REPEAT, shape, {Param<A>::get(), Param<E>::get(), Param<M>::get()}, 0,
                               P ? PGA_PROC_PARALLEL : 0u);
        const Link l = link<N>(t, shape);
        t.nodes[i].succ.push_back(l);
        return i;
    }
};
// a Box's faces are Quads (ComponentSplit.cuh:77-175); the <true> form
// splits through ParallelFaceSplitter (ComponentSplit.cuh:179-230), whose
// frames and cap seeds differ
template <bool P, class Top, class Bottom, class Sides> struct Lower<Operators::ComponentSplit<P, Top, Bottom, Sides>> {
    static int emit(Template& t, int shape) {
        // a Box's faces are Quads; a dynamic prism's caps are its polygon,
        // its sides Quads (GetCapFaceType, GetSideFaceType)
        const int cap = shape == Compiler::BOX ? (int)Compiler::QUAD
                      : shape == Compiler::DYNAMIC_CONVEX_RIGHT_PRISM ? (int)Compiler::DYNAMIC_CONVEX_POLYGON
                      : shape == Compiler::DYNAMIC_RIGHT_PRISM ? (int)Compiler::DYNAMIC_POLYGON : -1;
        const int side = (shape == Compiler::BOX || shape == Compiler::DYNAMIC_CONVEX_RIGHT_PRISM ||
                          shape == Compiler::DYNAMIC_RIGHT_PRISM) ? (int)Compiler::QUAD : -1;
        const int i = add_node(t, PGA_OP_COMPSPLIT, shape, {}, 0, P ? PGA_PROC_PARALLEL : 0u);
        const Link a = link<Top>(t, cap), b = link<Bottom>(t, cap), c = link<Sides>(t, side);
        t.nodes[i].succ = {a, b, c};
        return i;
    }
};
template <class E, class TT, class F> struct Lower<Operators::If<E, TT, F>> {
    static int emit(Template& t, int shape) {
        const int i = add_node(t, PGA_OP_IF, shape, {Param<E>::get()});
        const Link a = link<TT>(t, shape), b = link<F>(t, shape);
        t.nodes[i].succ = {a, b};
        return i;
    }
};
template <class A, class S, class TT, class F> struct Lower<Operators::IfSizeLess<A, S, TT, F>> {
    static int emit(Template& t, int shape) {
        const int i = add_node(t, PGA_OP_IFSIZELESS, shape, {Param<A>::get(), Param<S>::get()});
        const Link a = link<TT>(t, shape), b = link<F>(t, shape);
        t.nodes[i].succ = {a, b};
        return i;
    }
};
template <class Tag, class TT, class F> struct Lower<Operators::IfCollides<Tag, TT, F>> {
    static int emit(Template& t, int shape) {
        const int i = add_node(t, PGA_OP_IFCOLLIDES, shape, {Param<Tag>::get()});
        const Link a = link<TT>(t, shape), b = link<F>(t, shape);
        t.nodes[i].succ = {a, b};
        return i;
    }
};
template <class Tag, class N> struct Lower<Operators::Collider<Tag, N>> {
    static int emit(Template& t, int shape) {
        const int i = add_node(t, PGA_OP_COLLIDER, shape, {Param<Tag>::get()});
        const Link l = link<N>(t, shape);
        t.nodes[i].succ.push_back(l);
        return i;
    }
};

template <class PairT> struct PairParts;
template <class F, class N> struct PairParts<T::Pair<F, N>> {
    typedef F Factor;
    typedef N Next;
};

// static forms: Subdivide<Axis, Pair<Factor, Next>...> (Subdivide.cuh:22-168),
// RandomRule<Chance, Pair<Probability, Next>...> (RandomRule.cuh:20-111)
template <class A, class... Pairs> struct Lower<Operators::Subdivide<A, Pairs...>,
                                                std::enable_if_t<!std::is_same<A, Parameters::DynParams>::value>> {
    static int emit(Template& t, int shape) {
        const int i = add_node(t, PGA_OP_SUBDIV, shape,
                               {Param<A>::get(), Param<typename PairParts<Pairs>::Factor>::get()...},
                               0, PGA_PROC_STATIC);
        std::vector<Link> s = {link<typename PairParts<Pairs>::Next>(t, shape)...};
        t.nodes[i].succ = s;
        return i;
    }
};
template <> struct Lower<Operators::Subdivide<Parameters::DynParams>> {
    static int emit(Template& t, int shape) { return add_whole_entry(t, PGA_OP_SUBDIV, shape); }
};
template <class C, class... Pairs> struct Lower<Operators::RandomRule<C, Pairs...>,
                                                std::enable_if_t<!std::is_same<C, Parameters::DynParams>::value>> {
    static int emit(Template& t, int shape) {
        const int i = add_node(t, PGA_OP_STOCHASTIC, shape,
                               {Param<C>::get(), Param<typename PairParts<Pairs>::Factor>::get()...},
                               0, PGA_PROC_STATIC);
        std::vector<Link> s = {link<typename PairParts<Pairs>::Next>(t, shape)...};
        t.nodes[i].succ = s;
        return i;
    }
};
template <> struct Lower<Operators::RandomRule<Parameters::DynParams>> {
    static int emit(Template& t, int shape) { return add_whole_entry(t, PGA_OP_STOCHASTIC, shape); }
};
template <class... Ns> struct Lower<Operators::Replicate<Ns...>,
                                    std::enable_if_t<!std::is_same<std::tuple<Ns...>,
                                                                   std::tuple<Parameters::DynParams>>::value>> {
    static int emit(Template& t, int shape) {
        const int i = add_node(t, PGA_OP_REPLICATE, shape, {}, 0, PGA_PROC_STATIC);
        std::vector<Link> s = {link<Ns>(t, shape)...};
        t.nodes[i].succ = s;
        return i;
    }
};
template <> struct Lower<Operators::Replicate<Parameters::DynParams>> {
    static int emit(Template& t, int shape) { return add_whole_entry(t, PGA_OP_REPLICATE, shape); }
};
template <bool P, unsigned G, class Term, class A1, class A2, class A3, class A4, class A5>
struct Lower<Operators::Generate<P, G, Term, A1, A2, A3, A4, A5>> {
    static int emit(Template& t, int shape) {
        return add_node(t, PGA_OP_GENERATE, shape, {Param<Term>::get()}, (int)G);
    }
};

// SetAs* (SetAs.cuh:24-170): the shape becomes a dynamic polygon or prism
// whose vertices are the operator's Vec2 (or scalar) parameters; the
// <DynParams> form takes them, and its successor, from the dispatching entry.
template <int OpT> constexpr int set_as_shape() {
    return OpT == PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON ? (int)Compiler::DYNAMIC_CONVEX_POLYGON
         : OpT == PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM ? (int)Compiler::DYNAMIC_CONVEX_RIGHT_PRISM
         : OpT == PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON ? (int)Compiler::DYNAMIC_POLYGON
                                                        : (int)Compiler::DYNAMIC_RIGHT_PRISM;
}
template <int OpT, class N, class... V> struct LowerSetAs {
    static int emit(Template& t, int shape) {
        if constexpr (std::is_same<N, Parameters::DynParams>::value) {
            static_assert(sizeof...(V) == 0, "PGA::Parameters::DynParams can only be used as a single parameter");
            return add_whole_entry(t, OpT, shape);
        } else {
            const int i = add_node(t, OpT, shape, {Param<V>::get()...});
            const Link l = link<N>(t, set_as_shape<OpT>());
            t.nodes[i].succ.push_back(l);
            return i;
        }
    }
};
template <class N, class... V> struct Lower<Operators::SetAsDynamicConvexPolygon<N, V...>>
    : LowerSetAs<PGA_OP_SET_AS_DYNAMIC_CONVEX_POLYGON, N, V...> {};
template <class N, class... V> struct Lower<Operators::SetAsDynamicConvexRightPrism<N, V...>>
    : LowerSetAs<PGA_OP_SET_AS_DYNAMIC_CONVEX_RIGHT_PRISM, N, V...> {};
template <class N, class... V> struct Lower<Operators::SetAsDynamicConcavePolygon<N, V...>>
    : LowerSetAs<PGA_OP_SET_AS_DYNAMIC_CONCAVE_POLYGON, N, V...> {};
template <class N, class... V> struct Lower<Operators::SetAsDynamicConcaveRightPrism<N, V...>>
    : LowerSetAs<PGA_OP_SET_AS_DYNAMIC_CONCAVE_RIGHT_PRISM, N, V...> {};

// a procedure of the list (a Proc or a struct deriving from one)
template <class S, class Op, unsigned N> Template lower_proc(const Proc<S, Op, N>*) {
    Template t;
    t.shape_type = S::ShapeType;
    Lower<Op>::emit(t, S::ShapeType);
    return t;
}
template <class... Ps> std::vector<Template> lower_list(const T::List<Ps...>*) {
    return {lower_proc(static_cast<const Ps*>(nullptr))...};
}
template <class ProcListT> std::vector<Template> lower() {
    return lower_list(static_cast<const ProcListT*>(nullptr));
}

// ---- expansion over a dispatch table ----
struct Program {
    std::vector<pga_procedure> procedures;
    std::vector<pga_dispatch_entry> entries;
    std::map<int, int> static_roots;   // procedure -> expanded entry of an axiom dispatched to it directly
};

// Expands the procedure templates over the table. Entry i of the table is
// expanded entry i; `static_procs` are procedures axioms are dispatched to
// with no entry (DefaultAxiomTraits<Shape, OperatorCode>, Grid.cuh:113-140).
inline Program expand(const std::vector<Template>& procs, const pga_dispatch_entry* table, size_t n,
                      const std::vector<int>& static_procs = {}) {
    Program out;
    std::map<std::tuple<int, int, int>, int> index;   // (procedure, entry, node) -> expanded entry
    std::vector<std::tuple<int, int, int>> work;
    std::map<std::tuple<int, int, uint32_t, int>, int> kinds;
    auto fail = [](const std::string& m) { throw std::runtime_error("PGA shim: " + m); };
    auto proc_of_entry = [&](int e) {
        if (e < 0 || (size_t)e >= n) fail("entry index " + std::to_string(e) + " out of range");
        const int p = (int)table[e].operator_code;
        if (p < 0 || (size_t)p >= procs.size()) fail("entry " + std::to_string(e) + ": procedure out of range");
        return p;
    };
    auto slot = [&](int p, int e, int node) {
        auto key = std::make_tuple(p, e, node);
        auto it = index.find(key);
        if (it != index.end()) return it->second;
        const int id = (int)work.size();
        index[key] = id;
        work.push_back(key);
        return id;
    };
    // procedure codes follow the list: procedure i's root operator is kind i
    // (unless an earlier procedure has the same kind)
    for (const Template& t : procs) {
        const Node& r = t.nodes[0];
        const auto kind = std::make_tuple(r.op_type, r.shape_type, r.flags, r.gen_func);
        if (kinds.emplace(kind, (int)out.procedures.size()).second)
            out.procedures.push_back(pga_procedure{r.op_type, r.shape_type, r.gen_func, r.flags});
    }
    work.reserve(n);
    for (size_t i = 0; i < n; ++i) {       // roots first: expanded entry i = table entry i
        const int p = proc_of_entry((int)i);
        index[std::make_tuple(p, (int)i, 0)] = (int)i;
        work.push_back(std::make_tuple(p, (int)i, 0));
    }
    for (int p : static_procs) {
        if (p < 0 || (size_t)p >= procs.size()) fail("static axiom procedure out of range");
        out.static_roots[p] = slot(p, -1, 0);
    }
    auto call_target = [&](int p, int e) {   // procedure p on entry e
        if (e >= 0 && proc_of_entry(e) == p) return e;
        return slot(p, e, 0);
    };
    for (size_t w = 0; w < work.size(); ++w) {
        const int p = std::get<0>(work[w]), e = std::get<1>(work[w]), ni = std::get<2>(work[w]);
        const Template& t = procs[p];
        if (ni < 0 || (size_t)ni >= t.nodes.size()) fail("bad node");
        const Node& nd = t.nodes[ni];
        const pga_dispatch_entry* cur = e >= 0 ? &table[e] : nullptr;
        pga_dispatch_entry x;
        std::memset(&x, 0, sizeof(x));
        for (int k = 0; k < PGA_MAX_NUM_SUCCESSORS; ++k) {
            x.successors[k].entry_index = -1;
            x.successors[k].phase_index = -1;
            x.edge_indices[k] = -1;
        }
        x.sub_graph_index = -1;
        auto succ_entry = [&](int s) {
            if (!cur) fail("a call reads the successors of an axiom dispatched with no entry");
            if (s < 0 || (uint32_t)s >= cur->num_successors)
                fail("entry " + std::to_string(e) + ": successor " + std::to_string(s) + " out of range");
            return (int)cur->successors[s].entry_index;
        };
        std::vector<pga_dispatch_parameter> params;
        std::vector<int> succ;
        // the phase of each successor: a call dispatches into the phase its
        // entry's successor names (DCall, PSCall, Call.cuh:69-70, 88-89) or
        // its PhaseIndexT (FSCall, :184, 202); an operator chained inline
        // runs in the current one (PGA_PHASE_INHERIT)
        std::vector<int> phase;
        auto succ_phase = [&](int s) { return (int)cur->successors[s].phase_index; };
        if (nd.whole_entry) {
            if (!cur) fail("a <DynParams> operator needs a dispatching entry");
            for (uint32_t k = 0; k < cur->num_parameters; ++k) params.push_back(cur->parameters[k]);
            for (uint32_t k = 0; k < cur->num_successors; ++k) {
                const int te = (int)cur->successors[k].entry_index;
                succ.push_back(te < 0 ? -1 : call_target(proc_of_entry(te), te));
                phase.push_back(succ_phase((int)k));
            }
        } else {
            for (const ParamRef& r : nd.params) {
                if (r.dyn < 0) {
                    params.push_back(r.value);
                } else {
                    if (!cur) fail("DynParam in a procedure an axiom enters with no entry");
                    if ((uint32_t)r.dyn >= cur->num_parameters)
                        fail("entry " + std::to_string(e) + ": parameter " + std::to_string(r.dyn) + " out of range");
                    params.push_back(cur->parameters[r.dyn]);
                }
            }
            for (const Link& l : nd.succ) {
                switch (l.kind) {
                case LINK_NODE:
                    succ.push_back(slot(p, e, l.a));
                    phase.push_back(PGA_PHASE_INHERIT);
                    break;
                case LINK_DCALL: {
                    const int te = succ_entry(l.a);
                    succ.push_back(te < 0 ? -1 : call_target(proc_of_entry(te), te));
                    phase.push_back(succ_phase(l.a));
                    break;
                }
                case LINK_PSCALL: {
                    const int te = succ_entry(l.b);
                    if ((size_t)l.a >= procs.size()) fail("PSCall procedure out of range");
                    succ.push_back(call_target(l.a, te));
                    phase.push_back(succ_phase(l.b));
                    break;
                }
                default: {   // FSCall: a fixed entry, or the current one
                    if ((size_t)l.a >= procs.size()) fail("FSCall procedure out of range");
                    succ.push_back(call_target(l.a, l.b >= 0 ? l.b : e));
                    phase.push_back(l.phase);
                    break;
                }
                }
            }
        }
        if (params.size() > PGA_MAX_NUM_PARAMETERS) fail("too many parameters");
        if (succ.size() > PGA_MAX_NUM_SUCCESSORS) fail("too many successors");
        const auto kind = std::make_tuple(nd.op_type, nd.shape_type, nd.flags, nd.gen_func);
        auto it = kinds.find(kind);
        if (it == kinds.end()) {
            it = kinds.emplace(kind, (int)out.procedures.size()).first;
            out.procedures.push_back(pga_procedure{nd.op_type, nd.shape_type, nd.gen_func, nd.flags});
        }
        x.operator_code = it->second;
        x.num_parameters = (uint32_t)params.size();
        for (size_t k = 0; k < params.size(); ++k) x.parameters[k] = params[k];
        x.num_successors = (uint32_t)succ.size();
        for (size_t k = 0; k < succ.size(); ++k) {
            x.successors[k].entry_index = succ[k];
            x.successors[k].phase_index = succ[k] < 0 ? -1 : phase[k];
        }
        if (out.entries.size() <= w) out.entries.resize(w + 1);
        out.entries[w] = x;
    }
    return out;
}

} // namespace Lowering

// ---- axiom generators (Grid.cuh:113-284) ----
namespace GlobalVars {
// the seed table axioms read (Seeds[id % 32], GlobalVariables.cuh:28-32,91-98)
inline std::vector<float>& seeds() {
    static std::vector<float> s(Constants::MaxNumAxioms, 0.0f);
    return s;
}
inline float getSeed(int id) { return seeds()[(unsigned)id % Constants::MaxNumAxioms]; }
} // namespace GlobalVars

namespace Host {
// copies (n % 33) seeds, as the reference does (GlobalVariables.cuh:28-49)
inline void setSeeds(const float* s, unsigned int n) {
    const unsigned int m = n % (Constants::MaxNumAxioms + 1);
    for (unsigned int i = 0; i < m; ++i) GlobalVars::seeds()[i] = s[i];
}
} // namespace Host
namespace Device {
inline void setSeeds(const float* s, unsigned int n) { Host::setSeeds(s, n); }
} // namespace Device

namespace Grid {

template <class ShapeT, int OperatorCodeT = -1>
struct DefaultAxiomTraits {
    static ShapeT getShape() { return ShapeT(); }
    static const bool StaticDispatch = (OperatorCodeT > -1);
    static const int OperatorCode = OperatorCodeT;
    static int getEntryIndex() { return -1; }
    static bool isEnabled(int) { return true; }
    static math::float3 getSize(const ShapeT&) { return math::float3(1, 1, 1); }
    static math::float3 getPosition(int row, int column, const ShapeT& shape) {
        const math::float3 size = shape.getSize();
        return math::float3(column * (size.x * 1.5f), size.y * 0.5f, row * (size.z * 1.5f));
    }
};

namespace GlobalVars {
inline unsigned int& rows() { static unsigned int r = 1; return r; }
inline unsigned int& columns() { static unsigned int c = 1; return c; }
inline unsigned int& numElements() { static unsigned int n = 1; return n; }
inline void setNumElements(unsigned int n) { numElements() = n; }
inline unsigned int getNumElements() { return numElements(); }
inline void setDimensions(unsigned int r, unsigned int c) { rows() = r; columns() = c; }
inline unsigned int getRows() { return rows(); }
inline unsigned int getColumns() { return columns(); }
} // namespace GlobalVars

// One axiom record from the traits of starting axiom `axiomId`: the grid
// call order of the reference is kept (generate receives (column, row) in
// its (row, column) slots, Grid.cuh:222-231).
template <template <unsigned int> class TraitsT, unsigned int I, unsigned int N>
struct AxiomAt {
    static bool make(int id, unsigned int axiomId, int row, int column, pga_axiom& a, int& static_op) {
        if (axiomId != I) return AxiomAt<TraitsT, I + 1, N>::make(id, axiomId, row, column, a, static_op);
        typedef TraitsT<I> Tr;
        if (!Tr::isEnabled((int)axiomId)) return false;
        auto shape = Tr::getShape();
        const float seed = PGA::GlobalVars::getSeed(id);
        shape.setSeed(seed);
        shape.setCustomAttribute(seed);
        shape.setSize(Tr::getSize(shape));
        shape.setPosition(Tr::getPosition(row, column, shape));
        std::memset(&a, 0, sizeof(a));
        const math::float3 s = shape.getSize(), p = shape.getPosition();
        a.model[0] = 1.0f; a.model[5] = 1.0f; a.model[10] = 1.0f;
        a.model[3] = p.x; a.model[7] = p.y; a.model[11] = p.z;
        a.size[0] = s.x; a.size[1] = s.y; a.size[2] = s.z;
        a.seed = shape.getSeed();
        a.custom = shape.getCustomAttribute();
        a.shape_type = decltype(shape)::ShapeType;
        a.entry_index = Tr::StaticDispatch ? -1 : Tr::getEntryIndex();
        static_op = Tr::StaticDispatch ? Tr::OperatorCode : -1;
        return true;
    }
    static void static_ops(std::vector<int>& ops) {
        if (TraitsT<I>::StaticDispatch) ops.push_back(TraitsT<I>::OperatorCode);
        AxiomAt<TraitsT, I + 1, N>::static_ops(ops);
    }
};
template <template <unsigned int> class TraitsT, unsigned int N>
struct AxiomAt<TraitsT, N, N> {
    static bool make(int, unsigned int, int, int, pga_axiom&, int&) { return false; }
    static void static_ops(std::vector<int>&) {}
};

template <template <unsigned int> class StartingAxiomTraitsT, unsigned int NumStartingAxioms, unsigned int Rows,
          unsigned int Columns, int _NumElements = -1>
struct Static {
    static const unsigned int MaxNumElements = Rows * Columns;
    static const unsigned int NumElements = (_NumElements > 0) ? (unsigned int)_NumElements : MaxNumElements;
    static unsigned int getNumAxioms() { return NumStartingAxioms * NumElements; }
    static bool axiom(int id, pga_axiom& a, int& static_op) {
        const int axiomId = id % (int)NumStartingAxioms, groupId = id / (int)NumStartingAxioms;
        const int column = groupId % (int)Columns, row = groupId / (int)Rows;
        return AxiomAt<StartingAxiomTraitsT, 0, NumStartingAxioms>::make(id, (unsigned)axiomId, column, row, a,
                                                                          static_op);
    }
    static std::vector<int> staticOperatorCodes() {
        std::vector<int> v;
        AxiomAt<StartingAxiomTraitsT, 0, NumStartingAxioms>::static_ops(v);
        return v;
    }
};

template <template <unsigned int> class StartingAxiomTraitsT, unsigned int NumStartingAxioms>
struct Dynamic {
    static unsigned int getNumAxioms() { return NumStartingAxioms * GlobalVars::getNumElements(); }
    static bool axiom(int id, pga_axiom& a, int& static_op) {
        const int axiomId = id % (int)NumStartingAxioms, groupId = id / (int)NumStartingAxioms;
        const int column = groupId % (int)GlobalVars::getColumns(), row = groupId / (int)GlobalVars::getRows();
        return AxiomAt<StartingAxiomTraitsT, 0, NumStartingAxioms>::make(id, (unsigned)axiomId, column, row, a,
                                                                          static_op);
    }
    static std::vector<int> staticOperatorCodes() {
        std::vector<int> v;
        AxiomAt<StartingAxiomTraitsT, 0, NumStartingAxioms>::static_ops(v);
        return v;
    }
};

} // namespace Grid

// An axiom generator over caller-given records (the interpreter's
// Device::setAxioms, src/interpreter/GlobalVariables.cuh:47-59).
struct ListedAxioms {
    static std::vector<pga_axiom>& list() {
        static std::vector<pga_axiom> a;
        return a;
    }
    static void set(const pga_axiom* a, unsigned int n) { list().assign(a, a + n); }
    static unsigned int getNumAxioms() { return (unsigned int)list().size(); }
    static bool axiom(int id, pga_axiom& a, int& static_op) {
        a = list()[(size_t)id];
        static_op = -1;
        return true;
    }
    static std::vector<int> staticOperatorCodes() { return {}; }
};

// ---- evaluator (GPUEvaluator.cuh:475-536) ----
class Error : public std::runtime_error {
public:
    pga_status status;
    Error(pga_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

template <class ProcListT, class AxiomGeneratorT, class GenFuncFilterT, bool InstrumentedT, unsigned int NumSubgraphsT,
          unsigned int NumEdgesT, class ConfigurationT>
class SinglePhaseEvaluator {
public:
    SinglePhaseEvaluator() = default;
    SinglePhaseEvaluator(const SinglePhaseEvaluator&) = delete;
    SinglePhaseEvaluator& operator=(const SinglePhaseEvaluator&) = delete;
    virtual ~SinglePhaseEvaluator() { destroy(); }

    // capacity of terminal type `terminal` (the instanced-mesh buffer's
    // maxNumInstances, RenderingGlobalVariables.cuh:26-59)
    void setMaxNumInstances(unsigned int terminal, uint64_t capacity) {
        if (terminal >= PGA_MAX_TERMINALS) throw Error(PGA_ERR_INVALID_ARG, "terminal index out of range");
        if (caps_.size() <= terminal) caps_.resize(terminal + 1, 0);
        caps_[terminal] = capacity;
    }

    virtual void initialize(DispatchTableEntry* dispatchTable, size_t numEntries) {
        destroy();
        pga_config cfg{};
        cfg.technique = ConfigurationT::Technique;
        cfg.queue_capacity = ConfigurationT::QueueSize;
        cfg.device = -1;
        check(pga_create(&cfg, &ctx_));
        const std::vector<Lowering::Template> procs = Lowering::lower<ProcListT>();
        program_ = Lowering::expand(procs, reinterpret_cast<const pga_dispatch_entry*>(dispatchTable), numEntries,
                                    AxiomGeneratorT::staticOperatorCodes());
        check(pga_initialize(ctx_, program_.procedures.data(), (uint32_t)program_.procedures.size(),
                             program_.entries.data(), (uint32_t)program_.entries.size()));
        check(pga_set_seeds(ctx_, GlobalVars::seeds().data(), Constants::MaxNumAxioms));
        for (unsigned int t = 0; t < caps_.size(); ++t) {
            if (!caps_[t]) continue;
            void* p = nullptr;
            check(pga_device_alloc(caps_[t] * sizeof(pga_instance), &p));
            buffers_.push_back(p);
            check(pga_bind_instances(ctx_, t, static_cast<pga_instance*>(p), caps_[t]));
        }
    }

    virtual double execute() { return execute(AxiomGeneratorT::getNumAxioms()); }

    virtual double execute(unsigned int numAxioms) {
        if (!ctx_) throw Error(PGA_ERR_NOT_INITIALIZED, "PGA shim: execute before initialize");
        std::vector<pga_axiom> axioms;
        axioms.reserve(numAxioms);
        for (unsigned int id = 0; id < numAxioms; ++id) {
            pga_axiom a;
            int op = -1;
            if (!AxiomGeneratorT::axiom((int)id, a, op)) continue;
            if (op >= 0) a.entry_index = program_.static_roots.at(op);
            axioms.push_back(a);
        }
        check(pga_set_axioms(ctx_, axioms.data(), (uint32_t)axioms.size()));
        double seconds = 0.0;
        check(pga_execute(ctx_, (uint32_t)axioms.size(), &seconds));
        return seconds;
    }

    virtual void release() { destroy(); }

    uint64_t getNumInstances(unsigned int terminal) {
        uint64_t c[PGA_MAX_TERMINALS] = {0};
        check(pga_get_counts(ctx_, c, PGA_MAX_TERMINALS));
        return terminal < PGA_MAX_TERMINALS ? c[terminal] : 0;
    }
    // the instances of terminal type `terminal`, copied to the host
    std::vector<pga_instance> getInstances(unsigned int terminal) {
        std::vector<pga_instance> out;
        if (terminal >= caps_.size() || !caps_[terminal]) return out;
        const uint64_t n = std::min<uint64_t>(getNumInstances(terminal), caps_[terminal]);
        out.resize(n);
        size_t b = 0;
        for (unsigned int t = 0; t < terminal; ++t) b += caps_[t] ? 1 : 0;
        if (n) check(pga_copy_to_host(out.data(), buffers_[b], n * sizeof(pga_instance)));
        return out;
    }
    // generation function 0: the triangle soup of terminal type `terminal`
    // (shapes of `shapeType`, Box or Quad), copied to the host in the
    // reference's TriangleMeshData layout: 24 vertices and 36 indices per
    // box, 4 and 6 per quad, indices counted from baseVertex
    // (ShapeGenerator<Box|Quad, false>::run, ShapeGenerator.cuh:376-511, 670-708)
    void getTriangleMesh(unsigned int terminal, int shapeType, std::vector<pga_vertex>& vertices,
                         std::vector<uint32_t>& indices, uint32_t baseVertex = 0) {
        vertices.clear();
        indices.clear();
        if (terminal >= caps_.size() || !caps_[terminal]) return;
        const uint64_t n = std::min<uint64_t>(getNumInstances(terminal), caps_[terminal]);
        if (!n) return;
        const uint64_t V = shapeType == PGA_SHAPE_QUAD ? 4u : 24u;
        size_t b = 0;
        for (unsigned int t = 0; t < terminal; ++t) b += caps_[t] ? 1 : 0;
        void* dv = nullptr;
        void* di = nullptr;
        pga_status st = pga_device_alloc(n * V * sizeof(pga_vertex), &dv);
        if (st == PGA_OK) st = pga_device_alloc(n * (V + V / 2) * sizeof(uint32_t), &di);
        // the null stream: execute() has finished, and the copies below wait for it
        if (st == PGA_OK)
            st = pga_tessellate(static_cast<const pga_instance*>(buffers_[b]), n, shapeType,
                                static_cast<pga_vertex*>(dv), static_cast<uint32_t*>(di), baseVertex, nullptr);
        if (st == PGA_OK) {
            vertices.resize(n * V);
            indices.resize(n * (V + V / 2));
            st = pga_copy_to_host(vertices.data(), dv, vertices.size() * sizeof(pga_vertex));
            if (st == PGA_OK) st = pga_copy_to_host(indices.data(), di, indices.size() * sizeof(uint32_t));
        }
        pga_device_free(dv);
        pga_device_free(di);
        check(st);
    }
    const Lowering::Program& program() const { return program_; }
    pga_context* context() { return ctx_; }

private:
    void check(pga_status s) {
        if (s != PGA_OK) throw Error(s, std::string("PGA: ") + pga_status_string(s) + ": " + (ctx_ ? pga_last_error(ctx_) : ""));
    }
    void destroy() {
        if (ctx_) pga_destroy(ctx_);
        ctx_ = nullptr;
        for (void* p : buffers_) pga_device_free(p);
        buffers_.clear();
    }
    pga_context* ctx_ = nullptr;
    std::vector<uint64_t> caps_;
    std::vector<void*> buffers_;
    Lowering::Program program_;
};

// MultiPhaseEvaluator (GPUEvaluator.cuh:395-456): the phases come from the
// dispatch table's successors, and execute() runs them all, each to the end
// before the next starts, with the colliders stored so far gathered into a
// BVH between them. PhaseTraitsT, which in the reference decides which
// procedures get a queue in which phase, is accepted and not consulted:
// every procedure runs in every phase. initialize() throws when the table
// names a phase at or beyond NumPhasesT.
template <class ProcListT, class AxiomGeneratorT, class GenFuncFilterT, unsigned int NumPhasesT,
          template <class, int> class PhaseTraitsT, bool InstrumentedT, unsigned int NumSubgraphsT,
          unsigned int NumEdgesT, class ConfigurationT>
class MultiPhaseEvaluator : public SinglePhaseEvaluator<ProcListT, AxiomGeneratorT, GenFuncFilterT, InstrumentedT,
                                                        NumSubgraphsT, NumEdgesT, ConfigurationT> {
    typedef SinglePhaseEvaluator<ProcListT, AxiomGeneratorT, GenFuncFilterT, InstrumentedT, NumSubgraphsT, NumEdgesT,
                                 ConfigurationT> Base;

public:
    void initialize(DispatchTableEntry* dispatchTable, size_t numEntries) override {
        static_assert(NumPhasesT >= 1 && NumPhasesT <= PGA_MAX_PHASES, "NumPhasesT out of range");
        Base::initialize(dispatchTable, numEntries);
        uint32_t colliders = 0, phases = 0;
        pga_num_colliders(this->context(), &colliders, &phases);
        if (phases > NumPhasesT)
            throw Error(PGA_ERR_BAD_PROGRAM, "PGA: the dispatch table names phase " + std::to_string(phases - 1) +
                                             " of a " + std::to_string(NumPhasesT) + "-phase evaluator");
    }
    // capacity of the collider store (Collider stores one shape per call)
    void setMaxNumColliders(uint32_t capacity) {
        const pga_status st = pga_set_collider_capacity(this->context(), capacity);
        if (st != PGA_OK) throw Error(st, std::string("PGA: ") + pga_last_error(this->context()));
    }
    // colliders the last execute stored
    uint32_t getNumColliders() {
        uint32_t colliders = 0, phases = 0;
        pga_num_colliders(this->context(), &colliders, &phases);
        return colliders;
    }
};

} // namespace PGA
