/* the C++ front end of the pga C ABI (pga/core/Core.h) */
#pragma once
#include "Core.h"
