// Host representation of a gas absorption table (ocrt_gas, include/ocrt_amd.h)
// and its device packing.
#pragma once

#include <vector>

#include "common.h"

struct ocrt_gas {
    int n_lev = 0;
    std::vector<double> z;          // AFGL levels, ascending from the ground (km)
    std::vector<double> ngas[6];    // number density per level (cm^-3)
    double col_def[6] = {0, 0, 0, 0, 0, 0};
    double col_eff[6] = {0, 0, 0, 0, 0, 0};
    int overridden[6] = {0, 0, 0, 0, 0, 0};
    struct Table {
        int avail = 0, n_wl = 0, n_layers = 0;
        std::vector<double> wl, sig;   // [n_wl], [n_layers][n_wl]
    } xs[6];
};

namespace ocrt {

// Append one table to a device pool image; returns its descriptor.
DevGas pack_gas(const ocrt_gas &g, std::vector<double> &pool);

}  // namespace ocrt
