// include/velographx/storage/consolidation.hpp of the reference, served by the gfx950 facade.
#pragma once

#include "velographx_amd/consolidation.hpp"
