// mpf_state.h -- device state layout and launch arguments shared by the HIP
// kernels and the C-ABI host code.
//
// Slot layout (one buffer of the double-buffered state), cells ordered
//   c = (k*ny + m)*nx + l          (x fastest, z slowest; 2D: k = 0)
//   liq [c]                  liquid phi (phase 0), dense
//   cnt [c]                  number of stored solid slots (uint8)
//   sid [s*ncell + c]        grain id of slot s (uint16), ascending in s
//   sval[s*ncell + c]        phi of that grain (non-zero)
// A solid grain is stored at a cell only while its phi there is non-zero;
// the reference's dense phi[N][nx][ny][nz] (src/gpu_kernels_3d.py) holds the
// same information with zeros.  The slot capacity is min(KMAX, N-1), so it
// cannot overflow: after a step only active phases (at most KMAX) are
// non-zero.  Reads touch only slots [0, cnt), so the traffic is proportional
// to the number of grains present, not to the capacity.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "mpf_cell.h"

namespace mpf {

constexpr int KCAP = 64;        // largest KMAX supported
constexpr uint32_t NO_ID = 0xFFFFu;
constexpr int NO_HALT = 0x7fffffff; // halt_step value while no step has overflowed
// Work lists of a step: cells for the merge pass, cells with more active
// grains than the register path holds, and solid-liquid interface cells in
// NFB buckets by their number of active solids (WL_FRONT + n - 1), so that a
// wavefront of the interface pass evaluates the same number of solid-liquid
// pairs on every lane.
// The classification pass fills WL_FRONT.., the merge pass WL_LATE.., so
// that the interface cells of the first can be updated while the merge pass
// runs.
constexpr int NFB = 4;
enum { WL_MERGE = 0, WL_GENERAL = 1, WL_FRONT = 2, WL_LATE = WL_FRONT + NFB };
constexpr int NWL = WL_LATE + NFB;
// Counters of a step: one per work list, then the tile list of k_light.
constexpr int WC_TILES = NWL;
constexpr int NWC = NWL + 1;

// Tiles of the classification pass: LTX x LTY cells in x/y, LZC z-planes.
constexpr int LTX = 64, LTY = 8, LZC = 16;

// Occupancy mask: one uint16 per tile (index (tz*nty + ty)*ntx + tx); bit p
// is set when plane tz*LZC + p of the tile may hold a stored grain or liquid
// other than 1.  A clear bit guarantees cnt == 0 and liquid exactly 1 on
// every cell of that tile plane.  Each state buffer has its own mask; states
// written by the host get all bits set.
inline size_t occ_words(int nx, int ny, int nz, int dim)
{
    size_t ntx = (size_t)(nx + LTX - 1) / LTX, nty = (size_t)(ny + LTY - 1) / LTY;
    size_t nzc = dim == 3 ? (size_t)(nz + LZC - 1) / LZC : 1;
    return ntx * nty * nzc;
}

template <typename Real>
struct SlotView {
    Real *liq;
    uint8_t *cnt;
    uint16_t *sid;
    Real *sval;
    size_t stride; // slot s of cell c is at s*stride + c (the buffer's cell count)
};

struct Stats {
    int max_active;            // largest active count seen (status[1])
    int kmax_overflow;         // 1 if any cell exceeded KMAX (status[0])
    unsigned long long overflow_cells; // cells that exceeded KMAX (summed over steps)
    unsigned long long slow_cells;     // cells sent to the general path
    int halt_step;             // first step with a KMAX overflow (halt mode), else NO_HALT
    int pad;
    unsigned long long merge_cells;    // cells queued for the merge pass (summed over steps)
    unsigned long long front_cells;    // solid-liquid interface cells (summed over steps)
    unsigned long long light_cells;    // cells of the tile planes the classification pass computed
};

// One instance of a batch: its grid and where its cells, classification
// tiles, occupancy words and temperature planes start in the batch's arrays.
// A batch is a set of independent 3D grids stepped together: every kernel
// launch covers all instances.
struct Inst {
    int nx, ny, nz;
    int pad;
    uint32_t cell_off;
    uint32_t tile_off;
    uint32_t occ_off;
    uint32_t t_off;
};
constexpr int MAX_INST = 4096;

template <typename Real>
struct StepArgs {
    Consts<Real> K;
    SlotView<Real> in, out;
    const uint16_t *occ_in; // occupancy mask of `in`
    uint16_t *occ_out;      // occupancy mask of `out`, written by the classification pass
    const double *T_in;
    double *T_out;
    int temp_dense;  // 1: temperature per cell; 0: one value per z-plane
    int temp_f32;    // 1: the temperature is a float32 quantity
    double cooling;
    // Work lists of one step (cell indices).  A step runs five kernels: the
    // tile pass lists the tiles with planes that can change; the
    // classification pass over their cells finishes cells whose stencil holds
    // only liquid or only one grain and queues the rest (WL_MERGE); the merge
    // pass finishes cells that need no anisotropy terms and queues
    // solid-liquid interface cells (WL_LATE) and cells with more than the
    // register capacity of active grains (WL_GENERAL).
    int *wl_base;    // NWL lists of ncell entries each, list i at wl_base + i*ncell
    int *wl_count;   // [NWC][2] counters, ping-pong on the step parity
    int *tiles;      // work items of k_light (k_tiles -> k_light), n_tiles * LZC / 4 entries
    uint16_t *item_busy; // planes to compute of each item's tile
    int light_q;     // planes per work item: 16, 8 or 4
    int light_grid;  // blocks of the persistent classification pass (0: the default, 1024)
    size_t n_temp;   // temperature entries of all instances (planes, or cells when temp_dense)
    int step;        // step index (for the ping-pong counters and halt)
    int halt_mode;   // stop at the first KMAX overflow
    int tab_class;   // 1: the tables have the class form (K.cls holds the values)
    const Inst *inst; // batch instances (n_inst > 0 only for the batched kernels)
    int n_inst;
    size_t tab_stride; // table elements per instance (batched)
    int n_tiles;       // classification tiles over all instances (batched)
    Stats *stats;
    int ntx, nty;    // tiles in x and y
};

} // namespace mpf
