// TEST INFRASTRUCTURE -- for the CPU emulation build only: the ROCm host
// API header plus the kernel emulation of ../../emu.hpp.
#pragma once
#include_next <hip/hip_runtime_api.h>
#include "../../emu.hpp"
