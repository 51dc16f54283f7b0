// include/velographx/types.hpp of the reference, served by the gfx950 facade.
#pragma once

#include "velographx_amd/types.hpp"
