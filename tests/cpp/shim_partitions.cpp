// Grammars in the reference's partition codegen form (a procedure list of
// Proc<Shape, operator tree> plus a DispatchTable initialiser, as
// CodeGenerator::fromPartition writes them, CodeGenerator.cpp:1691-1722)
// compiled against pga/core/Core.h and run through the C ABI.
//
// Built as a program and as a shared library exporting shim_main:
//   shim_partitions lower <scene> <dir>        lowered program -> <dir>/<scene>.procs/.entries
//   shim_partitions run <scene> <dir>          derive (GPU); <dir>/<scene>.axioms in (listed
//                                              scenes), <dir>/<scene>.t<k> instances out
//   shim_partitions interp <dir>               interpreter facade over <dir>/interp.table
//                                              (pga_dispatch_entry records) and
//                                              <dir>/interp.axioms (shape, entry pairs)
#include <pga/core/Core.h>
#include <pga/core/GPUTechnique.h>
#include <pga/core/Grid.cuh>
#include <pga/facade/Interpreter.h>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

using namespace PGA;
using namespace PGA::Operators;
using namespace PGA::Shapes;
using namespace PGA::Parameters;

namespace Menger {   // four procedures, nine entries; axioms enter at entry 2

struct ProcedureList : T::List<
    /* procedure[0]= */ Proc<Box, Subdivide<DynParam<0>, T::Pair<DynParam<1>, DCall<0>>, T::Pair<DynParam<2>, DCall<1>>,
                                            T::Pair<DynParam<3>, DCall<2>>>, 1>,
    /* procedure[1]= */ Proc<Box, Discard, 1>,
    /* procedure[2]= */ Proc<Box, IfSizeLess<DynParam<0>, DynParam<1>, DCall<0>, DCall<1>>, 1>,
    /* procedure[3]= */ Proc<Box, Generate<false, 1, DynParam<0>>, 1>> {};

DispatchTable dispatchTable = {
    /* entries[0]= */ {2, {{PT_SCALAR, {0}}, {PT_SCALAR, {1}}}, {{1, 0}, {2, 0}}},
    /* entries[1]= */ {3, {{PT_SCALAR, {0}}}, {}},
    /* entries[2]= */ {0, {{PT_SCALAR, {0}}, {PT_SCALAR, {-1}}, {PT_SCALAR, {-1}}, {PT_SCALAR, {-1}}}, {{3, 0}, {7, 0}, {3, 0}}},
    /* entries[3]= */ {0, {{PT_SCALAR, {1}}, {PT_SCALAR, {-1}}, {PT_SCALAR, {-1}}, {PT_SCALAR, {-1}}}, {{4, 0}, {5, 0}, {4, 0}}},
    /* entries[4]= */ {0, {{PT_SCALAR, {2}}, {PT_SCALAR, {-1}}, {PT_SCALAR, {-1}}, {PT_SCALAR, {-1}}}, {{0, 0}, {0, 0}, {0, 0}}},
    /* entries[5]= */ {0, {{PT_SCALAR, {2}}, {PT_SCALAR, {-1}}, {PT_SCALAR, {-1}}, {PT_SCALAR, {-1}}}, {{0, 0}, {6, 0}, {0, 0}}},
    /* entries[6]= */ {1, {}, {}},
    /* entries[7]= */ {0, {{PT_SCALAR, {1}}, {PT_SCALAR, {-1}}, {PT_SCALAR, {-1}}, {PT_SCALAR, {-1}}}, {{5, 0}, {8, 0}, {5, 0}}},
    /* entries[8]= */ {1, {}, {}},
};

template <unsigned int AxiomId>
struct AxiomTraits : Grid::DefaultAxiomTraits<Box> {
    static int getEntryIndex() { return 2; }
    static math::float3 getSize(const Box&) {
        const float size = static_cast<float>(math::pow(3, 2));
        return math::float3(size, size, size);
    }
};
typedef Grid::Static<AxiomTraits, 1, 1, 1> AxiomGenerator;
const unsigned int Terminals = 1;
const uint64_t Capacity[] = {8000};

} // namespace Menger

namespace Facade {   // the batched facade grammar as chained static operators

struct ProcedureList : T::List<
    /* procedure[0]= */ Proc<Box, ComponentSplit<false, Discard, Discard, PSCall<1, 0>>, 1>,
    /* procedure[1]= */ Proc<Quad, Repeat<false, AxisParam<Y>, Scalar<3000, 3>, RepeatModeParam<ADJUST_TO_FILL>,
        Repeat<false, AxisParam<X>, Scalar<2500, 3>, RepeatModeParam<ADJUST_TO_FILL>,
            IfSizeLess<AxisParam<X>, Scalar<1600, 3>, PSCall<2, 0>,
                Subdivide<AxisParam<X>, T::Pair<Scalar<-1000, 3>, PSCall<2, 0>>, T::Pair<Scalar<1200, 3>, PSCall<2, 1>>,
                          T::Pair<Scalar<-1000, 3>, PSCall<2, 0>>>>>>, 1>,
    /* procedure[2]= */ Proc<Quad, Generate<false, 1, DynParam<0>>, 1>> {};

DispatchTable dispatchTable = {
    /* entries[0]= */ {0, {}, {{1, 0}}},
    /* entries[1]= */ {1, {}, {{2, 0}, {3, 0}}},
    /* entries[2]= */ {2, {{PT_SCALAR, {0}}}, {}},
    /* entries[3]= */ {2, {{PT_SCALAR, {1}}}, {}},
};
const unsigned int Terminals = 2;

} // namespace Facade

namespace Mix {   // every call form, expression and static parameters, Extrude, Replicate

struct ProcedureList : T::List<
    /* procedure[0]= */ Proc<Box, Translate<Scalar<0>, Exp<ShapeSize<Y>, Scalar<500, 3>, Multi>, Scalar<0>,
        Rotate<Scalar<0>, Scalar<30000, 3>, Scalar<0>,
            ComponentSplit<false, FSCall<1, 0, -1, 2>, DCall<0>,
                Replicate<PSCall<2, 1>, Scale<Scalar<-500, 3>, Scalar<-1000, 3>, Scalar<-1000, 3>, FSCall<3, 0>>>>>>, 1>,
    /* procedure[1]= */ Proc<Quad, Extrude<AxisParam<Z>, Scalar<2000, 3>, Generate<false, 1, Scalar<0>>>, 1>,
    /* procedure[2]= */ Proc<Quad, If<Exp<ShapeSize<X>, Scalar<2500, 3>, Gt>, Generate<false, 1, DynParam<0>>, Discard>, 1>,
    /* procedure[3]= */ Proc<Quad, Generate<false, 1, DynParam<0>>, 1>> {};

DispatchTable dispatchTable = {
    /* entries[0]= */ {0, {{PT_SCALAR, {2}}}, {{1, 0}, {3, 0}}},
    /* entries[1]= */ {2, {{PT_SCALAR, {1}}}, {}},
    /* entries[2]= */ {1, {}, {}},
    /* entries[3]= */ {2, {{PT_SCALAR, {3}}}, {}},
};
const unsigned int Terminals = 4;

} // namespace Mix

namespace Yard {   // two phases: the building stored as a collider in phase 0, props that ask in phase 1

// the lot replicates into its building (the entry's successor 0, phase 0)
// and its yard (a fixed entry in phase 1); the chains inside each procedure
// stay in the phase their shape was dispatched to
struct ProcedureList : T::List<
    /* procedure[0]= */ Proc<Box, Replicate<DCall<0>, FSCall<2, 1, -1, 2>>, 1>,
    /* procedure[1]= */ Proc<Box, Scale<Scalar<-500, 3>, Scalar<-1000, 3>, Scalar<-500, 3>,
        Collider<Scalar<1>, Generate<false, 1, Scalar<0>>>>, 1>,
    /* procedure[2]= */ Proc<Box, Repeat<false, AxisParam<X>, Scalar<1000, 3>, RepeatModeParam<ADJUST_TO_FILL>,
        Repeat<false, AxisParam<Z>, Scalar<1000, 3>, RepeatModeParam<ADJUST_TO_FILL>,
            Scale<Scalar<-500, 3>, Scalar<-300, 3>, Scalar<-500, 3>,
                IfCollides<Scalar<1>, PSCall<3, 0>, PSCall<3, 1>>>>>, 1>,
    /* procedure[3]= */ Proc<Box, Generate<false, 1, DynParam<0>>, 1>> {};

DispatchTable dispatchTable = {
    /* entries[0]= */ {0, {}, {{1, 0}}},
    /* entries[1]= */ {1, {}, {}},
    /* entries[2]= */ {2, {}, {{3, 1}, {4, 1}}},
    /* entries[3]= */ {3, {{PT_SCALAR, {2}}}, {}},
    /* entries[4]= */ {3, {{PT_SCALAR, {1}}}, {}},
};
const unsigned int Terminals = 3;

// every procedure has a queue in every phase
template <class ProcedureT, int PhaseIndexT> struct AllPhases {
    static const bool Enabled = true;
};

} // namespace Yard

struct Configuration {
    static const PGA::GPU::Technique Technique = PGA::GPU::MEGAKERNEL;
    static const unsigned int QueueSize = 1u << 20;
    static const unsigned int MaxSharedMemory = 0;
};

template <class T>
static bool write_all(const std::string& path, const std::vector<T>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    return (bool)f;
}

template <class T>
static std::vector<T> read_all(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot read " + path);
    const size_t bytes = (size_t)f.tellg();
    std::vector<T> v(bytes / sizeof(T));
    f.seekg(0);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    return v;
}

template <class ProcListT, class AxiomGenT,
          class Evaluator = SinglePhaseEvaluator<ProcListT, AxiomGenT, Rendering::GenFuncFilter, false, 0, 0,
                                                 Configuration>>
static int run_scene(const char* mode, const std::string& scene, const std::string& dir, DispatchTable& table,
                     unsigned int terminals, const std::vector<uint64_t>& caps) {
    const std::string base = dir + "/" + scene;
    if (std::string(mode) == "lower") {
        const auto procs = Lowering::lower<ProcListT>();
        auto entries = table.toDispatchTableEntriesPtr();
        const auto prog = Lowering::expand(procs, reinterpret_cast<const pga_dispatch_entry*>(entries.get()),
                                           table.entries.size(), AxiomGenT::staticOperatorCodes());
        write_all(base + ".procs", prog.procedures);
        write_all(base + ".entries", prog.entries);
        std::vector<pga_axiom> axioms;
        for (unsigned int id = 0; id < AxiomGenT::getNumAxioms(); ++id) {
            pga_axiom a;
            int op = -1;
            if (AxiomGenT::axiom((int)id, a, op)) axioms.push_back(a);
        }
        write_all(base + ".gen_axioms", axioms);
        std::printf("%s: %zu procedures, %zu entries, %zu axioms\n", scene.c_str(), prog.procedures.size(),
                    prog.entries.size(), axioms.size());
        return 0;
    }
    std::unique_ptr<Evaluator, ReleaseCallback> ev(new Evaluator());
    for (unsigned int t = 0; t < terminals; ++t) ev->setMaxNumInstances(t, caps[t]);
    ev->initialize(table.toDispatchTableEntriesPtr().get(), table.entries.size());
    const double seconds = ev->execute();
    for (unsigned int t = 0; t < terminals; ++t) {
        const auto inst = ev->getInstances(t);
        write_all(base + ".t" + std::to_string(t), inst);
        std::printf("%s: terminal %u: %zu instances\n", scene.c_str(), t, inst.size());
    }
    std::printf("%s: %.6f s\n", scene.c_str(), seconds);
    return 0;
}

// The test program's entry point, also exported from the shared-library
// build (SHIM_LIBRARY) so that a process already using the GPU can run it
// in-process: cooperative launches of two processes do not share a device.
extern "C" __attribute__((visibility("default"))) int shim_main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s lower|run <scene> <dir> | interp|grammar|grammar-run <dir>\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    try {
        if (mode == "interp") {
            const std::string dir = argv[2];
            const auto raw = read_all<pga_dispatch_entry>(dir + "/interp.table");
            const auto ax = read_all<int32_t>(dir + "/interp.axioms");
            const auto caps = read_all<uint64_t>(dir + "/interp.caps");
            std::unique_ptr<DispatchTable> table(new DispatchTable());
            for (const auto& e : raw) table->entries.push_back(PGA::Interpreter::toEntry(e));
            std::vector<PGA::Interpreter::Axiom> axioms;
            for (size_t i = 0; i + 1 < ax.size(); i += 2)
                axioms.push_back({(PGA::Compiler::ShapeType)ax[i], ax[i + 1]});
            const auto procs = getProcedureList();
            std::printf("interpreter procedure list: %zu procedures, Box Subdivide at %d\n", procs.size(),
                        procs.indexOf(PGA_OP_SUBDIV, PGA::Compiler::BOX));
            setAxioms(axioms.data(), (unsigned int)axioms.size());
            for (unsigned int t = 0; t < caps.size(); ++t) setMaxNumInstances(t, caps[t]);
            initializePGA(table);
            const double seconds = executePGA();
            auto& ev = PGA::Interpreter::evaluator();
            for (unsigned int t = 0; t < caps.size(); ++t) {
                const auto inst = ev->getInstances(t);
                write_all(dir + "/interp.t" + std::to_string(t), inst);
                std::printf("interp: terminal %u: %zu instances\n", t, inst.size());
            }
            std::printf("interp: %.6f s\n", seconds);
            destroyPGA();
            return 0;
        }
        if (mode == "grammar" || mode == "grammar-run") {
            // the interpreter's flow: grammar text -> compileGrammar -> (run)
            const std::string dir = argv[2];
            const auto text = read_all<char>(dir + "/grammar.pga");
            std::vector<PGA::Interpreter::Axiom> axioms;
            const auto table = compileGrammar(std::string(text.begin(), text.end()), axioms);
            const auto entries = table->toDispatchTableEntriesPtr();
            write_all(dir + "/grammar.table",
                      std::vector<pga_dispatch_entry>(reinterpret_cast<const pga_dispatch_entry*>(entries.get()),
                                                      reinterpret_cast<const pga_dispatch_entry*>(entries.get()) +
                                                          table->entries.size()));
            std::printf("grammar: %zu entries, %zu axioms, first at entry %d\n", table->entries.size(), axioms.size(),
                        axioms.empty() ? -1 : axioms[0].entryIndex);
            if (mode == "grammar") return 0;
            const auto caps = read_all<uint64_t>(dir + "/grammar.caps");
            setAxioms(axioms.data(), (unsigned int)axioms.size());
            for (unsigned int t = 0; t < caps.size(); ++t) setMaxNumInstances(t, caps[t]);
            initializePGA(table);
            const double seconds = executePGA();
            auto& ev = PGA::Interpreter::evaluator();
            for (unsigned int t = 0; t < caps.size(); ++t) {
                const auto inst = ev->getInstances(t);
                write_all(dir + "/grammar.t" + std::to_string(t), inst);
                std::printf("grammar: terminal %u: %zu instances\n", t, inst.size());
            }
            // generation function 0: one shape type per terminal in grammar.mesh
            std::ifstream has_mesh(dir + "/grammar.mesh");
            if (has_mesh.good()) {
                const auto shapes = read_all<int32_t>(dir + "/grammar.mesh");
                for (unsigned int t = 0; t < shapes.size(); ++t) {
                    std::vector<pga_vertex> v;
                    std::vector<uint32_t> idx;
                    ev->getTriangleMesh(t, shapes[t], v, idx);
                    write_all(dir + "/grammar.t" + std::to_string(t) + ".vtx", v);
                    write_all(dir + "/grammar.t" + std::to_string(t) + ".idx", idx);
                    std::printf("grammar: terminal %u: %zu vertices, %zu indices\n", t, v.size(), idx.size());
                }
            }
            std::printf("grammar: %.6f s\n", seconds);
            destroyPGA();
            return 0;
        }
        if (argc < 4) return 2;
        const std::string scene = argv[2], dir = argv[3];
        if (scene == "menger")
            return run_scene<Menger::ProcedureList, Menger::AxiomGenerator>(
                mode.c_str(), scene, dir, Menger::dispatchTable, Menger::Terminals, {Menger::Capacity[0]});
        // listed axioms: written by the caller as pga_axiom records
        if (mode == "run") {
            const auto a = read_all<pga_axiom>(dir + "/" + scene + ".axioms");
            ListedAxioms::set(a.data(), (unsigned int)a.size());
        }
        const auto caps = mode == "run" ? read_all<uint64_t>(dir + "/" + scene + ".caps") : std::vector<uint64_t>(8, 1);
        if (scene == "facade")
            return run_scene<Facade::ProcedureList, ListedAxioms>(mode.c_str(), scene, dir, Facade::dispatchTable,
                                                                  Facade::Terminals, caps);
        if (scene == "mix")
            return run_scene<Mix::ProcedureList, ListedAxioms>(mode.c_str(), scene, dir, Mix::dispatchTable,
                                                               Mix::Terminals, caps);
        if (scene == "yard")
            return run_scene<Yard::ProcedureList, ListedAxioms,
                             MultiPhaseEvaluator<Yard::ProcedureList, ListedAxioms, Rendering::GenFuncFilter, 2,
                                                 Yard::AllPhases, false, 0, 0, Configuration>>(
                mode.c_str(), scene, dir, Yard::dispatchTable, Yard::Terminals, caps);
        std::fprintf(stderr, "unknown scene %s\n", scene.c_str());
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}

#ifndef SHIM_LIBRARY
int main(int argc, char** argv) { return shim_main(argc, argv); }
#endif
