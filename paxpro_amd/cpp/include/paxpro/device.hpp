// Process-wide GPU context of the drop-in classes and the status check that
// turns a C-ABI error code into the std::runtime_error the reference throws.
#pragma once

#include <stdexcept>
#include <string>

#include "paxgpu.h"

namespace paxpro {

// Context on device $PAXPRO_DEVICE (default 0), created on first use.
pxg_ctx context();

inline void check(int rc)
{
  if (rc != PXG_OK) throw std::runtime_error(std::string("paxgpu: ") + pxg_last_error());
}

// Precision of the fields of solvers created after the call (PXG_PREC_FP64 by
// default; PXG_PREC_MIXED selects complex64 fields with FP64 phase arguments).
void setDefaultPrecision(int precision);
int defaultPrecision();

}  // namespace paxpro
