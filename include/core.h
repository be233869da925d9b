// Umbrella header (same contents as the reference's src/include/core.h).
#ifndef GRIDINF_INCLUDE_CORE_H
#define GRIDINF_INCLUDE_CORE_H

#include "utils.h"
#include "matrix.h"
#include "grid.h"

#endif
