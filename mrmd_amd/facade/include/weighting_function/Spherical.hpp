// reference include path weighting_function/Spherical.hpp of the mrmd action API, served by the gfx950 library
#pragma once

#include "mrmd_facade/core.hpp"
