// include/velographx/csr_graph.hpp of the reference, served by the gfx950 facade.
#pragma once

#include "velographx_amd/csr_graph.hpp"
