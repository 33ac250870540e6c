// state.hpp — per-device library state and the launch tuning table.
#pragma once
#include <stdint.h>

#include "../../include/clmsoil.h"

namespace clmsoil {

// Lanes per column (G) and workgroup size of the two fused kernels. Defaults are
// the tuned values for MI355X; clmsoil_set_tuning() replaces them.
struct Tuning {
    int richards_g = 4;
    int richards_block = 256;
    int heat_g = 4;
    int heat_block = 256;
    int richards_persistent = 1;   // resident grid with per-group column refill
    int richards_blocks_per_cu = 0;  // persistent grid size; 0 = occupancy query
    int zd_block = 128;              // Zeng-Decker 2009 step, one lane per column
    int ghf_block = 256;             // ground heat flux
};
extern Tuning g_tuning;

// Allocated once per device on first use: a 256-byte status block (word 0 device
// error bits, words 8..15 per-XCD work counters of the persistent Richards launch).
struct DeviceState {
    unsigned *err = nullptr;
};

int current_device_state(DeviceState **out);
int validate_levels(const clmsoil_levels *lev);

}  // namespace clmsoil
