// oracle_enum.cpp - exact enumeration for small systems.
//
// TEST INFRASTRUCTURE ONLY (see ldo_oracle.h).
#include "ldo_oracle.h"

extern "C" int oracle_enumerate(void*, int32_t, int32_t, int32_t, int32_t, double*, double*) {
    return LDO_E_NOTIMPL;
}
