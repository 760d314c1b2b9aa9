// reference include path io/RestoreGRO.hpp of the mrmd action API (io::restoreGRO,
// RestoreGRO.cpp:24-147), served by the gfx950 library
#pragma once

#include "mrmd_facade/core.hpp"
