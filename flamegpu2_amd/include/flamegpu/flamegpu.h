// Umbrella header of the FLAME GPU 2 compatible API (include/flamegpu/flamegpu.h).
#pragma once

#include "flamegpu/detail/core.h"
#include "flamegpu/model.h"
#include "flamegpu/runtime/device.h"
#include "flamegpu/simulation.h"
#include "flamegpu/batch.h"
