// pga_spec_host.h - host side of the kernels specialised to one program
// (pga_spec.cpp): the generated source and its run-time compilation.
#pragma once

#include <string>
#include <vector>

#include "pga_kernels.h"

// C++ source of the persistent derivation kernel `pga_spec_mega` for the
// program `entries` (the device image's entries), workgroups of block_size
// threads.
std::string pga_spec_source(const std::vector<pga::DevEntry>& entries, uint32_t count_geom, uint32_t use_seed,
                            uint32_t block_size);
// Compiles a source from pga_spec_source for gfx950 (hipRTC) into a code
// object; false with the compiler log on failure. Repeated sources come from
// a per-process cache (*cached = true).
bool pga_spec_compile(const std::string& source, std::vector<char>& code, std::string& log, bool* cached);
