// Host-side handle behind the opaque mm_rate_plan* of the C ABI.
#pragma once

#include "mm_plan.h"

struct mm_rate_plan {
    MMRatePlanDev host;  // host copy (metadata + tables)
    MMRatePlanDev* dev;  // device copy
    int device;
    int block;           // threads per block of the map / point kernels (0: default)
    struct MMSpecModule* spec;       // plan-specialised trajectory kernels (rate_spec.hip), or null
    struct MMSpecModule* spec_grad;  // plan-specialised forward-sensitivity kernels, or null
    struct MMSpecModule* spec_rf;    // RF kernels with this plan as the first phase, or null
    const mm_rate_plan* spec_rf_peer;  // the second phase plan spec_rf was built with
    unsigned long long spec_rf_serial;  // and its serial number
    unsigned long long serial;          // distinct for every plan the process builds
};

// (n_ground, n_excited) pairs with compiled rate-equation kernels
#define MM_SIZE_LIST(X) \
    X(12, 4)            \
    X(36, 4)            \
    X(2, 1)             \
    X(3, 1)             \
    X(4, 2)

static inline unsigned grid_for(long n, int block) { return (unsigned)((n + block - 1) / block); }
