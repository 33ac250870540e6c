// state.hpp — per-device library state and the launch tuning table.
#pragma once
#include <stdint.h>

#include <atomic>

#include "../../include/clmsoil.h"

namespace clmsoil {

// Lanes per column (G) and workgroup size of the two fused kernels. Defaults are
// the tuned values for MI355X; clmsoil_set_tuning() replaces them. Atomic: a
// launch reads them once; any value gives the same results.
struct Tuning {
    std::atomic<int> richards_g{4};
    std::atomic<int> richards_block{256};
    std::atomic<int> heat_g{4};
    std::atomic<int> heat_block{256};
    std::atomic<int> richards_persistent{1};   // resident grid with per-group column refill
    std::atomic<int> richards_blocks_per_cu{0};  // persistent grid size; 0 = occupancy query
    // persistent grid: refill once this many 16ths of a wavefront's groups are
    // idle, plain and with the fused first-substep network (which evaluates the
    // whole wavefront at every refill). Measured on the 1M-column production
    // state at G = 4 (ms per step): 4: 2.95, 6: 2.87, 7: 2.81, 8: 2.81, 9: 2.81,
    // 10: 2.83, 12: 2.94; learned on config 0: 7 fastest (4: +11 %, 10: +1 %,
    // 16: +9 %).
    std::atomic<int> richards_refill{7};
    std::atomic<int> richards_refill_learned{7};
    // per-layer fields of the Richards column in the LDS cache (0, 5 or 7; the
    // cache exists at G = 4 with 256-thread blocks)
    std::atomic<int> richards_lds_fields{7};
    std::atomic<int> zd_block{128};              // Zeng-Decker 2009 step, one lane per column
    std::atomic<int> ghf_block{256};             // ground heat flux
};
extern Tuning g_tuning;

// Set up once per device by clmsoil_init(): the 256-byte status block whose
// word 0 holds the device's error bits, and the device address of the static
// scratch of clmsoil_soilwm_moisture_form_solve_static.
struct DeviceState {
    unsigned *err = nullptr;
    unsigned *scratch = nullptr;
};

// Bytes of a Richards call's workspace: word 0 the call's error bits, words
// 8..15 the per-XCD work counters of the resident-grid launch.
constexpr int64_t RICHARDS_WORKSPACE_BYTES = 256;

int current_device_state(DeviceState **out);
int validate_levels(const clmsoil_levels *lev);

}  // namespace clmsoil
