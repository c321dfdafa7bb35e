// Error codes of the drop-in classes.  The enumerators and their order are
// the reference's (cpp/Errors.hpp:4-21) because callers catch them by value;
// the C ABI reports mcerror e as status 1 + e.
#ifndef VMC_HOST_ERRORS_HPP
#define VMC_HOST_ERRORS_HPP

#include "valomc.h"

enum mcerror {
  NO_LIGHTSOURCE,
  SIZE_MISMATCH_G,
  SIZE_MISMATCH_MUA,
  SIZE_MISMATCH_MUS,
  SIZE_MISMATCH_N,
  SIZE_MISMATCH_BCN,
  MISSING_BCN,
  SIZE_MISMATCH_BCTYPE,
  SIZE_MISMATCH_LIGHT_DIRECTION,
  SIZE_MISMATCH_LIGHT_DIRECTION_TYPE,
  INCONSISTENT_H,
  INCONSISTENT_BH,
  INCONSISTENT_MESH,
  INCONSISTENT_MESH_DUPLICATE_NEIGHBORS,
  NO_GAUSSIAN_SIGMA,
  MISSING_BOUNDARY
};

// Errors outside the reference's list (undefined light direction, device
// failures) are thrown as this type.
struct vmc_device_error {
  int status;
  const char *what() const { return vmc_status_string(status); }
};

static inline const char *errorstring(mcerror e) { return vmc_status_string(1 + (int)e); }

// Throw the reference exception for a C ABI status.
static inline void vmc_throw_status(int status) {
  if (status >= 1 && status <= 16) throw (mcerror)(status - 1);
  if (status != VMC_OK) throw vmc_device_error{status};
}

#endif
