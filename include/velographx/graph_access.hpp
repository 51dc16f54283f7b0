// include/velographx/graph_access.hpp of the reference, served by the gfx950 facade.
#pragma once

#include "velographx_amd/graph_access.hpp"
