/* the C++ front end of the pga C ABI (pga/core/Core.h): PGA::Rendering::GenFuncFilter */
#pragma once
#include "../core/Core.h"
