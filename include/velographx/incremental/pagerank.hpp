// include/velographx/incremental/pagerank.hpp of the reference, served by the gfx950 facade.
#pragma once

#include "velographx_amd/pagerank.hpp"
