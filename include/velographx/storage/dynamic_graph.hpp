// include/velographx/storage/dynamic_graph.hpp of the reference, served by the gfx950 facade.
#pragma once

#include "velographx_amd/dynamic_graph.hpp"
