// state.hpp — per-device library state and the launch tuning table.
#pragma once
#include <stdint.h>

namespace clmsoil {

// Lanes per column (G) and workgroup size of the two fused kernels. Defaults are
// the tuned values for MI355X; clmsoil_set_tuning() replaces them.
struct Tuning {
    int richards_g = 4;
    int richards_block = 256;
    int heat_g = 4;
    int heat_block = 256;
    int richards_persistent = 1;   // resident grid with per-group column refill
};
extern Tuning g_tuning;

// Allocated once per device on first use: a 256-byte status block (word 0 device
// error bits, word 1 pivot counter, words 8..15 per-XCD work counters) and the
// heat step's pivot list.
struct DeviceState {
    unsigned *err = nullptr;
    unsigned *pivot_count = nullptr;
    int64_t *pivot_list = nullptr;
    int64_t pivot_cap = 0;
    double *dense_ws = nullptr;    // dense systems of the heat pivoting pass
};

int current_device_state(DeviceState **out);

}  // namespace clmsoil
