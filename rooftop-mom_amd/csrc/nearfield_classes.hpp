// Near-field classification on the device (nearfield_classes.hip).
#pragma once

#include "rooftop.hpp"

namespace rt {

struct NearFieldClasses {
    int classes = 0;                        // distinct 5 x 5 windows (sparse-approximate-inverse rows)
    int stencils = 0;                       // distinct stencil tuples
    DeviceArray<int> cls;                   // per rooftop: its class, -1 where no rooftop exists
    DeviceArray<int> scls;                  // per rooftop: its stencil class, -1 where no rooftop exists
    DeviceArray<int> rep_dir;               // per class: the direction of its rooftops
    DeviceArray<unsigned long long> rep_bits;   // per class: its window, bit d 25 + (dx + 2) + 5 (dy + 2)
    DeviceArray<int> tuples;                // per stencil class: class, direction, the 18 neighbours' classes
};

// Classes and stencil classes of every rooftop of p, on stream s.
NearFieldClasses classify_near_field(const Problem &p, hipStream_t s);

}  // namespace rt
