// meep.hpp — the reference's header name (src/meep.hpp of NanoComp/meep) for programs that
// `#include <meep.hpp>`: it provides the C++ API of the MI355X engine, meep_amd.hpp (namespace
// meep), so such a program compiles unchanged and links -lmeep_amd.
#pragma once

#include "meep_amd.hpp"
