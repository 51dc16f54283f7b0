// include/velographx/incremental/kcore.hpp of the reference, served by the gfx950 facade.
#pragma once

#include "velographx_amd/kcore.hpp"
