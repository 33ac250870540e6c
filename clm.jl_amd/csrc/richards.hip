// richards.hip — fused adaptive-substep moisture-form Richards step (gfx950).
//
// Restates the algorithm of _soilwm_moisture_form_kernel!
// (src/biogeophys/soil_water_movement.jl:1250-1489) with a group of G lanes
// (G = 1, 2 or 4) owning one column. Layer j (1-based) of the column lives in lane
// (j-1) mod G, register slot (j-1) div G, so a wavefront load of one slot reads
// 64/G consecutive columns at G levels — whole 128-byte lines of the
// column-major [c, j] arrays. Everything per layer stays in VGPRs for the whole
// step: there is no per-column scratch in memory (the reference keeps 22
// nc x nlevsoi scratch matrices, :1555-1561).
//
// Per substep iteration:
//   * hydraulic properties (Clapp-Hornberger hk/smp and derivatives, :1281-1310)
//     and interface fluxes (:1334-1375) depend only on the water state, so they
//     are recomputed only after an accepted substep; a rejected substep reuses
//     them (bit-identical to recomputing, the state is unchanged);
//   * the tridiagonal RHS/LHS (:1381-1397) and the Thomas solve (:1421-1438) run
//     as a serial sweep over the group: the owner lane of row j computes it and a
//     DPP quad broadcast hands the carry (bet, dwat, cmx) to the next owner;
//   * error estimate, accept/reject/double (:1441-1483) are group-uniform.
// Arithmetic follows the reference's operation order (no reassociation, no
// contraction) and Julia's max/min semantics; divisions stay divisions.
//
// Scheduling. Columns need from 1 to ~40 substep iterations. A plain launch
// (one column per group for its lifetime) runs every wavefront for as many
// iterations as its slowest column. The persistent launch instead keeps a fixed
// grid resident and lets groups fetch their next columns from a work counter
// once enough of them are done (wavefront-aggregated atomics, one counter per
// XCD over a contiguous column range so neighbouring columns share an L2; idle
// groups move on to other XCDs' ranges at the end).
#include <atomic>

#include <hip/hip_runtime.h>
// The log and exp tables of clm_math.h are read from a per-block copy in LDS:
// every property update gathers 5 table entries per pow at lane-dependent
// indices, and LDS answers those gathers far sooner than the cache hierarchy.
// 8 KB per block; the values (and so the results) are the same.
static __shared__ double clm_lds_log_tab[256][3];
static __shared__ double clm_lds_exp_tab[128][2];
#define CLM_LOG_TAB(i, k) clm_lds_log_tab[i][k]
#define CLM_EXP_TAB(j, k) clm_lds_exp_tab[j][k]

#include "rcolumn.hpp"

namespace clmsoil {

// Copy the math tables to the block's LDS (all threads, before any returns).
__device__ __forceinline__ void fill_math_tables() {
    for (int t = threadIdx.x; t < 256 * 3; t += blockDim.x)
        (&clm_lds_log_tab[0][0])[t] = (&clm_log_tab[0][0])[t];
    for (int t = threadIdx.x; t < 128 * 2; t += blockDim.x)
        (&clm_lds_exp_tab[0][0])[t] = (&clm_exp_tab[0][0])[t];
    __syncthreads();
}

// LDS cache stride of a launch: the block size for G = 4 at 256 threads per block
// (35 doubles per lane, 70 KB per block beside the 8 KB of math tables), none
// otherwise (it would not fit at G = 1 or 2, and other block sizes keep the
// global-memory path).
template <int G>
constexpr int richards_cache_stride(int block) {
    return (G == 4 && block == 256) ? 256 : 0;
}
template <int G, bool EXP, int LCS, int LCF = 7>
static size_t richards_lds_bytes() {
    return (size_t)RColumn<G, EXP, false, LCS, LCF>::LC_FIELDS *
           RColumn<G, EXP, false, LCS, LCF>::S * LCS * sizeof(double);
}

// One column per group for the group's lifetime. LEARN: the first substep comes
// from the network, evaluated by the whole wavefront right after the loads.
template <int G, bool EXP, int LCS, bool LEARN>
__global__ __launch_bounds__(256, CLMSOIL_RICHARDS_WAVES) void richards_mf_kernel(RichardsParams P) {
    fill_math_tables();
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = gtid / G;
    const int k = (int)(threadIdx.x % G);
    const bool live = c < P.nc && (!P.mask || P.mask[c]);
    RColumn<G, EXP, false, LCS> col;
    if constexpr (LCS > 0) {
        extern __shared__ double richards_lds[];
        col.lc = richards_lds + threadIdx.x;
    }
    if constexpr (LEARN) {
        const bool ok = live && col.load(P, c, k);
        col.learned_first_substep(P, k, ok);   // all lanes of the wavefront
        if (!ok) return;
    } else {
        if (!live) return;
        if (!col.load(P, c, k)) return;
    }
    while (!col.iterate(P, k)) {
    }
    col.store(P, k);
}

// Persistent: a resident grid whose groups refill from per-XCD column ranges.
// Every wavefront leaves the loop once all eight ranges are exhausted.
template <int G, bool EXP, int LCS, bool LEARN, int LCF = 7, bool GRID = false>
__global__ __launch_bounds__(256, CLMSOIL_RICHARDS_WAVES) void richards_mf_persistent(RichardsParams P) {
    fill_math_tables();
    const int k = (int)(threadIdx.x % G);
    const int lane = (int)__lane_id();
    const int xcd = (int)(blockIdx.x % NUM_XCD);
    const int64_t chunk = (P.nc + NUM_XCD - 1) / NUM_XCD;
    unsigned *ctrs = P.work_counters;
    RColumn<G, EXP, false, LCS, LCF, GRID> col;
    if constexpr (LCS > 0) {
        extern __shared__ double richards_lds[];
        col.lc = richards_lds + threadIdx.x;
    }
    // refill threshold: P.refill_sixteenths / 16 of the wavefront's groups idle
    int refill_min = (64 / G) * P.refill_sixteenths / 16;
    refill_min = refill_min < 1 ? 1 : refill_min;
    bool active = false;   // group holds a column
    bool pending = false;  // group's finished column not yet written back
    bool done = false;     // no work left for this group
    int steal = 0;         // XCD ranges exhausted so far (wavefront-uniform)
    for (;;) {
        // ---- refill groups that have no column ----
        const bool want = !active && !done;
        const unsigned long long need = __ballot(want && k == 0);
        const unsigned long long busy = __ballot(active && k == 0);
        // A refill runs RColumn::load (and with LEARN the network, for the whole
        // wavefront) for the refilling groups only, with the rest of the
        // wavefront masked off; waiting until refill_min groups are idle (or none
        // is busy) shares that pass among them.
        if (need && (__popcll(need) >= refill_min || !busy)) {
            // finished columns are written back in the same masked pass
            if (pending) {
                col.store(P, k);
                pending = false;
            }
            const int leader = __ffsll((long long)need) - 1;
            const int cnt = __popcll(need);
            const int x = (xcd + steal) % NUM_XCD;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(&ctrs[x], (unsigned)cnt);
            base = (unsigned)__builtin_amdgcn_readlane((int)base, leader);
            const int64_t lo = (int64_t)x * chunk;
            const int64_t hi = lo + chunk < P.nc ? lo + chunk : P.nc;
            int64_t cidx = -1;
            if (want) {
                // rank of this group among the refilling groups of the wavefront
                const int my_k0 = lane - k;
                const unsigned long long below = need & ((1ull << my_k0) - 1ull);
                const int64_t idx = lo + (int64_t)base + __popcll(below);
                if (idx < hi) cidx = idx;
            }
            if (cidx >= 0) {
                const bool on = !P.mask || P.mask[cidx];
                active = on ? col.load(P, cidx, k) : false;
            }
            // first substeps of the columns just loaded: the whole wavefront
            if constexpr (LEARN) col.learned_first_substep(P, k, cidx >= 0 && active);
            // a wavefront that got fewer columns than it asked for moves on to the
            // next XCD range; after the last range its idle groups are done
            if (__ballot(want && k == 0 && cidx < 0)) {
                ++steal;
                if (steal >= NUM_XCD && want && cidx < 0) done = true;
            }
        }
        if (!__ballot(active)) {
            if (!__ballot(!done)) break;
            continue;
        }
        if (active) {
            if (col.iterate(P, k)) {
                pending = true;
                active = false;
            }
        }
    }
    if (pending) col.store(P, k);   // every group has refilled once more before it is done
}

template <int G, bool EXP, int LCS, bool LEARN = false>
static hipError_t launch_g(const RichardsParams &P, hipStream_t stream, int block) {
    const int64_t threads = P.nc * G;
    const int64_t grid = (threads + block - 1) / block;
    if (richards_lds_bytes<G, EXP, LCS>() > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)richards_mf_kernel<G, EXP, LCS, LEARN>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)richards_lds_bytes<G, EXP, LCS>());
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((richards_mf_kernel<G, EXP, LCS, LEARN>), dim3((unsigned)grid), dim3(block),
                       (richards_lds_bytes<G, EXP, LCS>()), stream, P);
    return hipGetLastError();
}

// Resident blocks per CU of richards_mf_persistent at a block size, from the
// occupancy calculator, cached per (device, kernel, block) after the first query.
template <int G, bool EXP, int LCS, bool LEARN, int LCF = 7, bool GRID = false>
static int resident_blocks(int dev, int block) {
    static std::atomic<int> cache[16][3];
    const int bi = block == 64 ? 0 : (block == 128 ? 1 : 2);
    if (dev < 0 || dev >= 16) dev = 0;
    int nb = cache[dev][bi].load(std::memory_order_relaxed);
    if (nb > 0) return nb;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, richards_mf_persistent<G, EXP, LCS, LEARN, LCF, GRID>,
                                                     block, richards_lds_bytes<G, EXP, LCS, LCF>()) !=
        hipSuccess)
        return -1;
    if (nb < 1) nb = 1;
    cache[dev][bi].store(nb, std::memory_order_relaxed);
    return nb;
}

template <int G, bool EXP, int LCS, bool LEARN = false, int LCF = 7, bool GRID = false>
static hipError_t launch_p(const RichardsParams &P, hipStream_t stream, int block,
                           int blocks_per_cu) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return hipErrorUnknown;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return hipErrorUnknown;
    if (richards_lds_bytes<G, EXP, LCS, LCF>() > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)richards_mf_persistent<G, EXP, LCS, LEARN, LCF, GRID>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)richards_lds_bytes<G, EXP, LCS, LCF>());
        if (e != hipSuccess) return e;
    }
    // blocks_per_cu <= 0: as many as are resident; larger than that, the blocks past
    // the resident ones would start only once the resident ones exit (all work
    // taken), so the grid is clamped to the resident count
    const int resident = resident_blocks<G, EXP, LCS, LEARN, LCF, GRID>(dev, block);
    int nb = blocks_per_cu;
    if (nb <= 0 || (resident > 0 && nb > resident)) nb = resident;
    if (nb <= 0) return hipErrorUnknown;
    // Any grid size is correct: groups only take work from the counters and never
    // wait on each other, so blocks that start late simply find less work.
    int64_t grid = (int64_t)nb * cus;
    grid = (grid + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    const int64_t need = (P.nc * G + block - 1) / block;
    if (grid > need) grid = (need + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    hipLaunchKernelGGL((richards_mf_persistent<G, EXP, LCS, LEARN, LCF, GRID>), dim3((unsigned)grid), dim3(block),
                       (richards_lds_bytes<G, EXP, LCS, LCF>()), stream, P);
    return hipGetLastError();
}

template <bool EXP>
static hipError_t launch_exp(const RichardsParams &P, int G, int block, bool persistent,
                             int blocks_per_cu, int lds_fields, hipStream_t stream) {
    // the LDS column cache needs 256-thread blocks at G = 4; lds_fields 0 keeps
    // every field in registers or memory
    const bool cached = richards_cache_stride<4>(block) > 0 && lds_fields > 0;
    if (persistent) {
        switch (G) {
            case 1: return launch_p<1, EXP, 0>(P, stream, block, blocks_per_cu);
            case 2: return launch_p<2, EXP, 0>(P, stream, block, blocks_per_cu);
            default:
                if (!cached) return launch_p<4, EXP, 0>(P, stream, block, blocks_per_cu);
                if (lds_fields == 5) return launch_p<4, EXP, 256, false, 5>(P, stream, block, blocks_per_cu);
                // the CLM5 grid (12 snow, 20 soil levels) with its level counts compiled in
                return (P.nlevsoi == MAX_NLEVSOI && P.nlevsno == MAX_NLEVSNO)
                           ? launch_p<4, EXP, 256, false, 7, true>(P, stream, block, blocks_per_cu)
                           : launch_p<4, EXP, 256>(P, stream, block, blocks_per_cu);
        }
    }
    switch (G) {
        case 1: return launch_g<1, EXP, 0>(P, stream, block);
        case 2: return launch_g<2, EXP, 0>(P, stream, block);
        default:
            return cached ? launch_g<4, EXP, 256>(P, stream, block)
                          : launch_g<4, EXP, 0>(P, stream, block);
    }
}

hipError_t launch_richards(const RichardsParams &P, int G, int block, bool persistent,
                           int blocks_per_cu, int lds_fields, hipStream_t stream) {
    if (P.nc == 0) return hipSuccess;
    return P.expensive ? launch_exp<true>(P, G, block, persistent, blocks_per_cu, lds_fields, stream)
                       : launch_exp<false>(P, G, block, persistent, blocks_per_cu, lds_fields, stream);
}

hipError_t launch_richards_learned(const RichardsParams &P, bool persistent, int blocks_per_cu,
                                   hipStream_t stream) {
    if (P.nc == 0) return hipSuccess;
    if (!P.mlp || !P.mlp_norm) return hipErrorInvalidValue;
    if (persistent)
        return P.expensive ? launch_p<4, true, 256, true>(P, stream, 256, blocks_per_cu)
                           : launch_p<4, false, 256, true>(P, stream, 256, blocks_per_cu);
    return P.expensive ? launch_g<4, true, 256, true>(P, stream, 256)
                       : launch_g<4, false, 256, true>(P, stream, 256);
}

}  // namespace clmsoil
