// include/velographx/incremental/connected_components.hpp of the reference, served by the gfx950 facade.
#pragma once

#include "velographx_amd/components.hpp"
