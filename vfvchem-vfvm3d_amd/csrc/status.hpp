// Status codes of the C ABI, shared by the host sources.
#pragma once
#include "../../include/vfvm.h"
