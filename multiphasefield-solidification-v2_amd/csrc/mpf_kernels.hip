// mpf_kernels.hip -- HIP kernels of the multi-phase-field step for gfx950.
//
// This translation unit is compiled twice: with MPF_REAL=double and
// -ffp-contract=off (the FP64 solver, bit-identical to the CPU oracle) and
// with MPF_REAL=float (the FP32 solver, contraction allowed).  Each build
// exports plain launcher functions named with the MPF_SFX suffix (_f64/_f32).
//
// Kernels
//   one time step (src/gpu_kernels_3d.py:462-486, :489-527, :530-680,
//   :683-830, :833-838) runs over the slot layout:
//   k_tiles       the work items of k_light: plane ranges of the tiles that
//                 hold a plane that can change
//   k_light       the listed planes: classification, temperature, the cells
//                 whose stencil holds only liquid or only one grain, and the
//                 routing of simple interface cells to their buckets
//   k_merge       queued cells: APT merge; finishes cells without a
//                 solid-liquid pair (or with anisotropy off)
//   k_front1      interface cells with one solid, anisotropy and torque terms
//   k_front_multi interface cells with 2-4 solids and those the merge pass
//                 found, each bucket compiled for its solid count
//   k_step_general cells with more active grains than the register path
//                 holds (one lane per solid up to 8, scratch arrays beyond)
//   k_interface   the last three as one launch
//   k_apt_slots   checked APT of the slot state into the reference's dense
//                 mf/nf layout (parity export)
//   k_pack/k_unpack dense reference layout <-> slots
//   k_init_columns slot state from a grain map and per-column z profiles
//   k_dense_*     the reference's three kernels on its own dense layout
//                 (drop-in entry points)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpf_gather.h"
#include "mpf_state.h"
#include "mpf_streams.h"

#ifndef MPF_REAL
#error "MPF_REAL must be float or double"
#endif
// MPF_PART selects the part of the file one compilation emits, so that the
// parts build in parallel: 0 the 3D step, 1 the batched 3D step, 2 the 2D
// step, 3 every other kernel and the step dispatch.  Unset: everything.
#ifndef MPF_PART
#define MPF_PART -1
#endif
#define MPF_HAS(p) (MPF_PART < 0 || MPF_PART == (p))
#define MPF_CAT2(a, b) a##b
#define MPF_CAT(a, b) MPF_CAT2(a, b)
#define MPF_NAME(n) MPF_CAT(n, MPF_SFX)

using Real = MPF_REAL;

// Everything below lives in a precision-specific namespace so that the two
// builds of this file can be linked into one library.
namespace mpf {
namespace MPF_CAT(impl, MPF_SFX) {

constexpr int FAST_A = 4;   // active solids held in registers on the fast path
constexpr int BX = 64;      // tile width in x (one wavefront per row)
constexpr int BY = 4;       // tile rows in y

// Raise the running maximum of the active count.  It is monotone, so a lane
// issues the atomic only when it would raise it.
__device__ __forceinline__ void note_active(Stats *st, int ntot, int kmax, int step, int halt_mode)
{
    if (ntot > *((volatile int *)&st->max_active)) atomicMax(&st->max_active, ntot);
    if (ntot > kmax) {
        atomicMax(&st->kmax_overflow, 1);
        atomicAdd(&st->overflow_cells, 1ull);
        if (halt_mode) atomicMin(&st->halt_step, step);
    }
}

// x is +0.0 (all bits clear; -0.0 is not)
__device__ __forceinline__ bool is_plus_zero(float x) { return __float_as_uint(x) == 0u; }
__device__ __forceinline__ bool is_plus_zero(double x) { return __double_as_longlong(x) == 0ll; }

__device__ __forceinline__ int *wl_counter(const StepArgs<Real> &a, int list, int step)
{
    return a.wl_count + list * 2 + (step & 1);
}

__device__ __forceinline__ int *wl_list(const StepArgs<Real> &a, int list)
{
    return a.wl_base + (size_t)list * a.in.stride;
}

// ------------------------------------------------------------ batches
// In a batch (B = 1) every instance is a grid of its own inside the shared
// buffers: its cells start at cell_off, its tables at instance * tab_stride,
// its temperature planes at t_off, its occupancy words at occ_off.  Work
// lists hold batch-wide cell indices.
__device__ __forceinline__ SlotView<Real> shift_view(SlotView<Real> v, size_t off)
{
    v.liq += off;
    v.cnt += off;
    v.sid += off;
    v.sval += off;
    return v;
}

// The arguments as seen by one instance.
__device__ __forceinline__ void instance_args(StepArgs<Real> &a, const Inst &d, int ii)
{
    a.K.nx = d.nx;
    a.K.ny = d.ny;
    a.K.nz = d.nz;
    a.K.tab = a.K.tab + (size_t)ii * a.tab_stride;
    a.in = shift_view(a.in, d.cell_off);
    a.out = shift_view(a.out, d.cell_off);
    a.T_in += d.t_off;
    a.T_out += d.t_off;
    a.occ_in += d.occ_off;
    a.occ_out += d.occ_off;
}

// Cell offsets of the instances in LDS for the per-cell instance search.
template <int B>
__device__ __forceinline__ void load_cell_offsets(const StepArgs<Real> &a, uint32_t *s_coff)
{
    if constexpr (B != 0) {
        for (int i = threadIdx.x; i < a.n_inst; i += blockDim.x) s_coff[i] = a.inst[i].cell_off;
        __syncthreads();
    }
}

// A queued cell resolved to its grid: constants, views, temperature and the
// cell index and coordinates within the grid.
struct Where {
    Consts<Real> K;
    SlotView<Real> in, out;
    const double *T;
    size_t c;
    int l, m, k;
};

template <int B>
__device__ __forceinline__ Where locate(const StepArgs<Real> &a, const uint32_t *s_coff, size_t cg)
{
    Where w;
    w.K = a.K;
    w.in = a.in;
    w.out = a.out;
    w.T = a.T_out;
    w.c = cg;
    if constexpr (B != 0) {
        int lo = 0, hi = a.n_inst - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if ((size_t)s_coff[mid] <= cg)
                lo = mid;
            else
                hi = mid - 1;
        }
        const Inst d = a.inst[lo];
        w.K.nx = d.nx;
        w.K.ny = d.ny;
        w.K.nz = d.nz;
        w.K.tab = a.K.tab + (size_t)lo * a.tab_stride;
        w.in = shift_view(a.in, d.cell_off);
        w.out = shift_view(a.out, d.cell_off);
        w.T = a.T_out + d.t_off;
        w.c = cg - d.cell_off;
    }
    w.l = (int)(w.c % w.K.nx);
    w.m = (int)((w.c / w.K.nx) % w.K.ny);
    w.k = (int)(w.c / ((size_t)w.K.nx * w.K.ny));
    return w;
}

// Batch-wide tile t: the arguments as its instance sees them, the first cell
// of that instance (batch-wide list entries) and the tile's index inside it.
template <int B>
__device__ __forceinline__ int enter_tile(StepArgs<Real> &a, int t, size_t &cell_base)
{
    cell_base = 0;
    if constexpr (B != 0) {
        int lo = 0, hi = a.n_inst - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if ((int)a.inst[mid].tile_off <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        const Inst d = a.inst[lo];
        instance_args(a, d, lo);
        a.ntx = (d.nx + LTX - 1) / LTX;
        a.nty = (d.ny + LTY - 1) / LTY;
        cell_base = d.cell_off;
        return t - (int)d.tile_off;
    }
    return t;
}

// Planes of tile (tx, ty, tz) that the classification pass computes (bit p:
// plane tz*LZC + p), and the occupancy of its own column over planes
// k0-1 .. k0+LZC in *colmask (bit kk-k0+1).
//
// An occupancy bit that is clear means that every cell of the tile plane
// holds liquid exactly 1 and no grain.  A cell whose whole stencil is clear
// becomes liquid 1 again (its liquid phase is active, 1 + dt*0 = 1, 1/1 = 1);
// when the output buffer's bit for the plane is clear as well, that buffer
// already holds this result, so the plane is neither read nor written.
template <int DIM>
__device__ __forceinline__ uint32_t tile_busy(const StepArgs<Real> &a, int tx, int ty, int tz, uint32_t &colmask)
{
    const Consts<Real> &K = a.K;
    const int nzc = (DIM == 3) ? (K.nz + LZC - 1) / LZC : 1;
    const int k0 = tz * LZC;
    const int np = (DIM == 3) ? min(LZC, K.nz - k0) : 1;
    const int l0 = tx * LTX, m0 = ty * LTY;
    const int w = min(LTX, K.nx - l0), h = min(LTY, K.ny - m0);
    auto tile_of = [&](int ttx, int tty, int ttz) { return ((size_t)ttz * a.nty + tty) * a.ntx + ttx; };
    uint32_t cm = (uint32_t)a.occ_in[tile_of(tx, ty, tz)] << 1;
    if (DIM == 3 && tz > 0) cm |= ((uint32_t)a.occ_in[tile_of(tx, ty, tz - 1)] >> (LZC - 1)) & 1u;
    if (DIM == 3 && tz + 1 < nzc) cm |= ((uint32_t)a.occ_in[tile_of(tx, ty, tz + 1)] & 1u) << (LZC + 1);
    colmask = cm;
    // tiles holding the x and y halo (x periodic; y periodic in 3D, mirror in 2D)
    const int xm = bc_pm(l0, K.nx) / LTX, xp = bc_pp(l0 + w - 1, K.nx) / LTX;
    const int ym = (DIM == 3 ? bc_pm(m0, K.ny) : bc_mm(m0, K.ny)) / LTY;
    const int yp = (DIM == 3 ? bc_pp(m0 + h - 1, K.ny) : bc_mp(m0 + h - 1, K.ny)) / LTY;
    const uint32_t nbr = (uint32_t)a.occ_in[tile_of(xm, ty, tz)] | (uint32_t)a.occ_in[tile_of(xp, ty, tz)] |
                         (uint32_t)a.occ_in[tile_of(tx, ym, tz)] | (uint32_t)a.occ_in[tile_of(tx, yp, tz)];
    return (cm | (cm >> 1) | (cm >> 2) | nbr | (uint32_t)a.occ_out[tile_of(tx, ty, tz)]) & ((1u << np) - 1u);
}

// Pass 0: one lane per tile lists the work items of k_light, counts their
// cells, and records that the liquid phase was active on the planes left as
// they are.  An item is one range of light_q planes (16, 8 or 4: small grids
// take shorter ranges, so that more blocks share the few tiles with work)
// holding at least one plane to compute, listed as tile * (16 / light_q) +
// range with the tile's plane mask.
template <int DIM, int B>
__global__ __launch_bounds__(256) void k_tiles(StepArgs<Real> a0)
{
    if (a0.halt_mode && a0.stats->halt_step < a0.step) return;
    const int t = blockIdx.x * 256 + threadIdx.x;
    uint32_t busy = 0;
    bool left = false;
    unsigned long long cells = 0;
    if (t < a0.n_tiles) {
        StepArgs<Real> a = a0;
        size_t cb;
        const int lb = enter_tile<B>(a, t, cb);
        const int tx = lb % a.ntx, ty = (lb / a.ntx) % a.nty, tz = lb / (a.ntx * a.nty);
        uint32_t colmask;
        busy = tile_busy<DIM>(a, tx, ty, tz, colmask);
        const int np = (DIM == 3) ? min(LZC, a.K.nz - tz * LZC) : 1;
        left = busy != ((1u << np) - 1u);
        cells = (unsigned long long)__popc(busy) * (unsigned)(min(LTX, a.K.nx - tx * LTX) * min(LTY, a.K.ny - ty * LTY));
    }
    const int lane = threadIdx.x & 63;
    const int q = a0.light_q, nsub = LZC / q;
    for (int r = 0; r < nsub; ++r) {
        const bool has = ((busy >> (r * q)) & ((1u << q) - 1u)) != 0u;
        const unsigned long long ball = __ballot(has);
        if (ball == 0ull) continue;
        const int leader = __ffsll((long long)ball) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(wl_counter(a0, WC_TILES, a0.step), __popcll(ball));
        base = __shfl(base, leader, 64);
        if (has) {
            const int e = base + __popcll(ball & ((1ull << lane) - 1ull));
            a0.tiles[e] = t * nsub + r;
            a0.item_busy[e] = (uint16_t)busy;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cells += __shfl_xor(cells, o, 64);
    if (lane == 0 && cells) atomicAdd(&a0.stats->light_cells, cells);
    if (__ballot(left) != 0ull && lane == 0 && *((volatile int *)&a0.stats->max_active) < 1)
        atomicMax(&a0.stats->max_active, 1);
}

// Pass 1 over the listed work items: classification and the cells whose
// stencil holds only liquid (no stored grain at any of the 7 points) or one
// grain with zero liquid everywhere.  Interface cells whose stencil stores at
// most one grain per point go to their interface bucket; everything else is
// queued for k_merge.  The same launch advances the temperature (per plane
// or per cell).
//
// A block takes one work item (a range of planes of an LTX x LTY column tile)
// at a time and marches up its z-planes, skipping the planes that cannot
// change.  The cell's own column is carried in registers across planes
// (z-1, z, z+1); the x/y neighbours of plane z come from an LDS tile with a
// one-cell halo.  The queued cells of an item are collected in LDS and
// appended to each list with a single atomic.  The blocks of one XCD take a
// contiguous part of the item list (listed in occupancy order, so
// neighbouring tiles mostly share an L2).
//
// The loads of the next plane are issued before the current plane is
// processed.  Every lane issues the same loads on every plane (lanes outside
// the grid read a clamped cell, lanes without a halo role read their own
// cell, and planes whose occupancy bit is clear read the first 64 cells, a
// line that stays cached, and substitute the pure-liquid cell), so the number
// of memory operations in flight is static and the compiler waits only for
// the plane it needs.  The march is unrolled by four so that the column ring
// (z-1, z, z+1, z+2) and the halo double buffer are fixed registers.  Slot 0
// is read together with the count and dropped when the count is zero.
// Offsets are cell indices of type Off (32-bit when every byte offset of the
// arrays fits in 32 bits).
struct ColCell {
    uint8_t cnt;
    Real liq;
    uint16_t sid;   // slot-0 grain (NO_ID when empty)
    Real sv;        // slot-0 value
};

// base[i] with a 32-bit byte offset when Off is 32-bit (lets the compiler use
// a scalar base with a 32-bit vector offset)
template <typename T, typename Off>
__device__ __forceinline__ T &at(T *base, Off i)
{
    if constexpr (sizeof(Off) == 4)
        return *(T *)((char *)base + (uint32_t)(i * (uint32_t)sizeof(T)));
    else
        return base[i];
}

template <typename Off>
__device__ __forceinline__ ColCell load_cell(const SlotView<Real> &S, Off c)
{
    ColCell r;
    r.cnt = at(S.cnt, c);
    r.liq = at(S.liq, c);
    r.sid = at(S.sid, c);
    r.sv = at(S.sval, c);
    return r;
}

// Cell of a tile plane whose occupancy bit is clear: liquid exactly 1, no
// grain.  occ: the bit; the loaded values are kept when it is set.
__device__ __forceinline__ void pristine_unless(ColCell &r, bool occ)
{
    r.cnt = occ ? r.cnt : (uint8_t)0;
    r.liq = occ ? r.liq : Real(1.0);
}

__device__ __forceinline__ void finish_cell(ColCell &r)
{
    const bool e = r.cnt == 0;
    r.sid = e ? (uint16_t)NO_ID : r.sid;
    r.sv = e ? Real(0.0) : r.sv;
}
// FP64: held to 128 VGPRs (four waves per SIMD, two blocks per CU; a few
// spilled): 0.51 -> 0.33 ms at the mid-run state of 512x512x256.  The FP32
// build needs 108 anyway.
template <int DIM, typename Off, int B>
__global__ __launch_bounds__(LTX * LTY) __attribute__((amdgpu_waves_per_eu(sizeof(Real) == 8 ? 4 : 1))) void k_light(StepArgs<Real> a0)
{
    if (a0.halt_mode && a0.stats->halt_step < a0.step) return;
    if (blockIdx.x == 0 && threadIdx.x < NWC) a0.wl_count[threadIdx.x * 2 + ((a0.step + 1) & 1)] = 0;
    constexpr int NP = (DIM == 3) ? 7 : 5;
    constexpr int ZC = (DIM == 3) ? LZC : 1;
    // plane tiles, double-buffered so one barrier per plane suffices
    // key = count | slot-0 grain << 8
    __shared__ uint32_t s_key[2][LTY + 2][LTX + 2];
    __shared__ Real s_liq[2][LTY + 2][LTX + 2];
    __shared__ Real s_sv[2][LTY + 2][LTX + 2];
    // queued cells of a tile as (z-plane, row, column): the merge list grows
    // from the start, the interface list (entries tagged with their bucket
    // above bit 13) from the end; a cell enters at most one of them
    constexpr int QCAP = LTX * LTY * ZC;
    __shared__ uint16_t s_q[QCAP];
    __shared__ int s_qn, s_qbase, s_any, s_fn;
    __shared__ int s_bn[NFB], s_fbase[NFB], s_wc[NFB * LTX * LTY / 64];
    __shared__ uint32_t s_occ;
    const bool need_terms = a0.K.aniso || DIM == 2;

    // this block's share of the tile list: XCD x = blockIdx % 8 takes the
    // x-th eighth, its blocks stride through it
    const int nlist = *wl_counter(a0, WC_TILES, a0.step);
    int i0 = blockIdx.x, i1 = nlist, di = gridDim.x;
    if ((gridDim.x & 7) == 0) {
        const int x = blockIdx.x & 7;
        i0 = (int)((long long)nlist * x / 8) + (blockIdx.x >> 3);
        i1 = (int)((long long)nlist * (x + 1) / 8);
        di = gridDim.x >> 3;
    }
    const int nsub = LZC / a0.light_q;
    for (int li = i0; li < i1; li += di) {
        StepArgs<Real> a = a0;
        size_t cell_base;
        const int item = a0.tiles[li];
        const int lb = enter_tile<B>(a, item / nsub, cell_base);
        const int sub = item % nsub;
        const Consts<Real> &K = a.K;
        const int tx = lb % a.ntx, ty = (lb / a.ntx) % a.nty, tz = lb / (a.ntx * a.nty);
        const int lx = threadIdx.x % LTX, ly = threadIdx.x / LTX;
        const int l0 = tx * LTX, m0 = ty * LTY;
        const int w = min(LTX, K.nx - l0), h = min(LTY, K.ny - m0);
        const int l = l0 + lx, m = m0 + ly;
        const bool valid = (lx < w) && (ly < h);
        const int k0 = tz * ZC;
        // this item's planes [ks, ke) of the tile's [k0, k0 + ZC)
        const int ks = (DIM == 3) ? k0 + sub * a0.light_q : 0;
        const int ke = (DIM == 3) ? min(ks + a0.light_q, K.nz) : 1;
        const Off plane = (Off)K.nx * (Off)K.ny;
        if (threadIdx.x == 0) {
            s_qn = 0;
            s_fn = 0;
            s_any = 0;
            s_occ = 0;
        }
        if (threadIdx.x < NFB) s_bn[threadIdx.x] = 0;
        auto tile_of = [&](int ttx, int tty, int ttz) { return ((size_t)ttz * a.nty + tty) * a.ntx + ttx; };
        // own column's occupancy over planes k0-1 .. k0+LZC (bit kk-k0+1)
        const int nzc = (DIM == 3) ? (K.nz + LZC - 1) / LZC : 1;
        uint32_t colmask = (uint32_t)a.occ_in[tile_of(tx, ty, tz)] << 1;
        if (DIM == 3 && tz > 0) colmask |= ((uint32_t)a.occ_in[tile_of(tx, ty, tz - 1)] >> (LZC - 1)) & 1u;
        if (DIM == 3 && tz + 1 < nzc) colmask |= ((uint32_t)a.occ_in[tile_of(tx, ty, tz + 1)] & 1u) << (LZC + 1);
        // planes to compute (tile_busy, evaluated by k_tiles before this
        // pass starts rewriting the output occupancy)
        const uint32_t busy = a0.item_busy[li];

        // the one halo cell this thread loads each plane (x periodic; y
        // periodic in 3D, mirror in 2D): wave 0 the row below the tile, wave
        // 1 the row above, lanes 0-7 / 8-15 of wave 2 the cells left / right
        // of the tile's rows; hrow, hcol: its place in the LDS tile (hrow < 0:
        // none).  All lanes load (the others their own cell), so every wave
        // issues the same loads.
        int hrow = -1, hcol = 0, hm = 0, hl = 0;
        if (ly < 2 && lx < w) {
            hrow = (ly == 0) ? 0 : h + 1;
            hcol = lx + 1;
            hl = l;
            if (DIM == 3)
                hm = (ly == 0) ? bc_pm(m0, K.ny) : bc_pp(m0 + h - 1, K.ny);
            else
                hm = (ly == 0) ? bc_mm(m0, K.ny) : bc_mp(m0 + h - 1, K.ny);
        } else if (ly == 2 && lx < 16 && (lx & 7) < h) {
            hrow = (lx & 7) + 1;
            hcol = (lx < 8) ? 0 : w + 1;
            hm = m0 + (lx & 7);
            hl = (lx < 8) ? bc_pm(l0, K.nx) : bc_pp(l0 + w - 1, K.nx);
        }
        const uint32_t occh = hrow >= 0 ? a.occ_in[tile_of(hl / LTX, hm / LTY, tz)] : 0u;
        // in-plane cell offsets: own (clamped into the grid) and halo (lanes
        // without a halo role use their own cell)
        const Off own_xy = (Off)min(m, K.ny - 1) * (Off)K.nx + (Off)min(l, K.nx - 1);
        const Off h_xy = hrow >= 0 ? (Off)hm * (Off)K.nx + (Off)hl : own_xy;
        const Off dum =
            (Off)(threadIdx.x & 63) < plane * (Off)(DIM == 3 ? K.nz : 1) ? (Off)(threadIdx.x & 63) : (Off)0;

        auto own_cell = [&](int kk) {
            const bool occ = (colmask >> (kk - k0 + 1)) & 1u;
            ColCell r = load_cell<Off>(a.in, occ ? own_xy + (Off)kk * plane : dum);
            pristine_unless(r, occ);
            return r;
        };
        auto halo = [&](int kk) {
            const bool occ = (occh >> (kk - k0)) & 1u;
            ColCell r = load_cell<Off>(a.in, occ ? h_xy + (Off)kk * plane : dum);
            pristine_unless(r, occ);
            return r;
        };

        // ring[r & 3] holds plane ks + r - 1 of the own column; hv[r & 1]
        // the halo cell of plane ks + r
        ColCell ring[4], hv[2];
        ring[0] = (DIM == 3) ? own_cell(bc_mm(ks, K.nz)) : ColCell{};
        ring[1] = own_cell(ks);
        ring[2] = (DIM == 3) ? own_cell(bc_mp(ks, K.nz)) : ColCell{};
        ring[3] = ColCell{};
        hv[0] = halo(ks);
        hv[1] = ColCell{};
        finish_cell(ring[0]);
        finish_cell(ring[1]);
        uint32_t occ_bits = 0;
        int nb = 0; // planes computed so far: their parity selects the LDS buffer

        for (int kb = ks; kb < ke; kb += 4) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = kb + j;
                if (k >= ke) break;
                ColCell &cp = ring[j & 3], &cc = ring[(j + 1) & 3], &cn = ring[(j + 2) & 3];
                ColCell &hx = hv[j & 1];
                // loads of the next plane (the last plane reloads itself; unused)
                if (DIM == 3) {
                    const int kn = min(k + 1, ke - 1);
                    ring[(j + 3) & 3] = own_cell(bc_mp(kn, K.nz));
                    hv[(j + 1) & 1] = halo(kn);
                }
                finish_cell(cn);
                finish_cell(hx);
                if (!((busy >> (k - k0)) & 1u)) {
                    // plane left as it is (block-uniform branch); its liquid
                    // phase was active
                    if (threadIdx.x == 0) s_any = 1;
                    continue;
                }
                const int b = nb & 1;
                ++nb;
                if (valid) {
                    s_key[b][ly + 1][lx + 1] = (uint32_t)cc.cnt | ((uint32_t)cc.sid << 8);
                    s_liq[b][ly + 1][lx + 1] = cc.liq;
                    s_sv[b][ly + 1][lx + 1] = cc.sv;
                }
                if (hrow >= 0) {
                    s_key[b][hrow][hcol] = (uint32_t)hx.cnt | ((uint32_t)hx.sid << 8);
                    s_liq[b][hrow][hcol] = hx.liq;
                    s_sv[b][hrow][hcol] = hx.sv;
                }
                __syncthreads();
                bool queue = false, dirty = false, no_liq = false;
                int fb = -1; // interface bucket
                if (valid) {
                    const Off c = own_xy + (Off)k * plane;
                    // stencil: c, x+, x-, y+, y-, z+, z-
                    uint32_t pkey[NP];
                    Real pl[NP];
                    Real psv[NP];
                    const int yy = ly + 1, xx = lx + 1;
                    pkey[P_C] = (uint32_t)cc.cnt | ((uint32_t)cc.sid << 8); pl[P_C] = cc.liq; psv[P_C] = cc.sv;
                    pkey[P_XP] = s_key[b][yy][xx + 1]; pl[P_XP] = s_liq[b][yy][xx + 1]; psv[P_XP] = s_sv[b][yy][xx + 1];
                    pkey[P_XM] = s_key[b][yy][xx - 1]; pl[P_XM] = s_liq[b][yy][xx - 1]; psv[P_XM] = s_sv[b][yy][xx - 1];
                    pkey[P_YP] = s_key[b][yy + 1][xx]; pl[P_YP] = s_liq[b][yy + 1][xx]; psv[P_YP] = s_sv[b][yy + 1][xx];
                    pkey[P_YM] = s_key[b][yy - 1][xx]; pl[P_YM] = s_liq[b][yy - 1][xx]; psv[P_YM] = s_sv[b][yy - 1][xx];
                    if constexpr (DIM == 3) {
                        pkey[P_ZP] = (uint32_t)cn.cnt | ((uint32_t)cn.sid << 8); pl[P_ZP] = cn.liq; psv[P_ZP] = cn.sv;
                        pkey[P_ZM] = (uint32_t)cp.cnt | ((uint32_t)cp.sid << 8); pl[P_ZM] = cp.liq; psv[P_ZM] = cp.sv;
                    }
                    // one grain everywhere: every key equals the centre's, with count 1
                    int sum = 0;
                    bool single = (pkey[P_C] & 0xffu) == 1u;
#pragma unroll
                    for (int p = 0; p < NP; ++p) {
                        sum += (int)(pkey[p] & 0xffu);
                        single = single && (pkey[p] == pkey[P_C]) && (pl[p] == Real(0.0));
                    }
                    dirty = sum > 0;
                    if (sum == 0) {
                        // only liquid in the stencil: active set {L} or empty
                        bool act = apt_active<NP>(pl);
                        Real ln, sn;
                        single_phase_update(pl[P_C], act, true, K.dt, ln, sn);
                        at(a.out.liq, c) = ln;
                        at(a.out.cnt, c) = 0;
                        dirty = ln != Real(1.0);
                        if (act) s_any = 1;
                    } else if (single) {
                        // one grain, no liquid: active set {g} or empty
                        bool act = apt_active<NP>(psv);
                        Real ln, sn;
                        single_phase_update(psv[P_C], act, false, K.dt, ln, sn);
                        at(a.out.liq, c) = ln;
                        if (sn != Real(0.0)) {
                            at(a.out.sid, c) = (uint16_t)cc.sid;
                            at(a.out.sval, c) = sn;
                            at(a.out.cnt, c) = (uint8_t)1;
                        } else {
                            at(a.out.cnt, c) = (uint8_t)0;
                        }
                        if (act) s_any = 1;
                    } else {
                        // A solid-liquid interface cell whose stencil stores at
                        // most one grain per point, each with a positive value,
                        // has exactly the distinct stored grains as active
                        // solids; within the register capacity and KMAX it goes
                        // straight to the interface pass, in the bucket of its
                        // solid count (the merge pass would route it there).
                        int nd = 0;
                        bool direct = need_terms && apt_active<NP>(pl);
#pragma unroll
                        for (int p = 0; p < NP; ++p) {
                            const uint32_t cp = pkey[p] & 0xffu;
                            direct = direct && (cp == 0u || (cp == 1u && psv[p] > Real(0.0)));
                            bool first = cp == 1u;
#pragma unroll
                            for (int q = 0; q < p; ++q)
                                first = first && !((pkey[q] & 0xffu) == 1u && pkey[q] == pkey[p]);
                            nd += first ? 1 : 0;
                        }
                        if (direct && nd <= NFB && nd + 1 <= K.kmax)
                            fb = nd - 1;
                        else
                            queue = true;
                        // liquid +0 (bitwise) at every stencil point: the
                        // merge pass need not read it
                        no_liq = true;
#pragma unroll
                        for (int p = 0; p < NP; ++p) no_liq = no_liq && is_plus_zero(pl[p]);
                    }
                }
                // output occupancy: an output cell can only hold a grain where
                // its stencil held one; a liquid-only stencil may leave liquid
                // below 1
                if (__ballot(dirty) != 0ull) occ_bits |= 1u << (k - k0);
                // append queued cells to the tile's LDS lists
                const int lane = threadIdx.x & 63;
                const unsigned long long lt = (1ull << lane) - 1ull;
                const uint16_t qe = (uint16_t)(((k - k0) * LTY + ly) * LTX + lx);
                unsigned long long ball = __ballot(queue);
                if (ball != 0ull) {
                    int leader = __ffsll((long long)ball) - 1;
                    int base = 0;
                    if (lane == leader) base = atomicAdd(&s_qn, __popcll(ball));
                    base = __shfl(base, leader, 64);
                    if (queue) s_q[base + __popcll(ball & lt)] = (uint16_t)(qe | (no_liq ? 0x8000u : 0u));
                }
                ball = __ballot(fb >= 0);
                if (ball != 0ull) {
                    int leader = __ffsll((long long)ball) - 1;
                    int base = 0;
                    if (lane == leader) base = atomicAdd(&s_fn, __popcll(ball));
                    base = __shfl(base, leader, 64);
                    if (fb >= 0) s_q[QCAP - 1 - (base + __popcll(ball & lt))] = (uint16_t)(qe | (fb << 13));
#pragma unroll
                    for (int q = 0; q < NFB; ++q) {
                        const unsigned long long bq = __ballot(fb == q);
                        if (lane == 0 && bq != 0ull) atomicAdd(&s_bn[q], __popcll(bq));
                    }
                }
            }
        }
        if ((threadIdx.x & 63) == 0 && occ_bits) atomicOr(&s_occ, occ_bits);
        __syncthreads();
        __syncthreads();
        const int nq = s_qn, nf = s_fn;
        if (threadIdx.x == 0) {
            if (s_any && *((volatile int *)&a.stats->max_active) < 1) atomicMax(&a.stats->max_active, 1);
            s_qbase = nq > 0 ? atomicAdd(wl_counter(a, WL_MERGE, a.step), nq) : 0;
            if (nsub == 1) {
                a.occ_out[tile_of(tx, ty, tz)] = (uint16_t)s_occ;
            } else {
                // other items write the other plane ranges of this tile's
                // word: replace only this range's bits, in the aligned
                // 32-bit word that holds the 16-bit entry
                uint16_t *o = a.occ_out + tile_of(tx, ty, tz);
                const int sh = (int)(((uintptr_t)o >> 1) & 1u) * 16;
                unsigned int *wd = (unsigned int *)((uintptr_t)o & ~(uintptr_t)3);
                const unsigned int rng = (((1u << a0.light_q) - 1u) << (sub * a0.light_q)) << sh;
                atomicAnd(wd, ~rng);
                atomicOr(wd, (s_occ << sh) & rng);
            }
        }
        if (threadIdx.x >= 64 && threadIdx.x < 64 + NFB) {
            const int q = threadIdx.x - 64;
            s_fbase[q] = s_bn[q] > 0 ? atomicAdd(wl_counter(a, WL_FRONT + q, a.step), s_bn[q]) : 0;
        }
        __syncthreads();
        auto cell_of = [&](int q) {
            q &= (1 << 13) - 1;
            const int qx = q % LTX, qy = (q / LTX) % LTY, qz = q / (LTX * LTY);
            return (int)(cell_base + cidx(l0 + qx, m0 + qy, k0 + qz, K.nx, K.ny));
        };
        int *dst = wl_list(a, WL_MERGE) + s_qbase;
        // bit 31 of a merge entry: no liquid in the cell's stencil
        for (int i = threadIdx.x; i < nq; i += LTX * LTY)
            dst[i] = (int)((uint32_t)cell_of(s_q[i]) | ((uint32_t)(s_q[i] >> 15) << 31));
        // interface cells, compacted into their buckets: per chunk of the
        // list, each wave counts its entries per bucket, and an entry's place
        // is the bucket's base, the entries of the chunks before and of the
        // waves before, and its rank in its wave (block-uniform loop)
        constexpr int NW = LTX * LTY / 64;
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        int run[NFB];
#pragma unroll
        for (int q = 0; q < NFB; ++q) run[q] = s_fbase[q];
        for (int i0 = 0; i0 < nf; i0 += LTX * LTY) {
            const int i = i0 + threadIdx.x;
            const int e = i < nf ? (int)s_q[QCAP - 1 - i] : 0;
            const int mq = i < nf ? (e >> 13) : -1;
            unsigned long long mb = 0ull;
#pragma unroll
            for (int q = 0; q < NFB; ++q) {
                const unsigned long long bq = __ballot(mq == q);
                if (mq == q) mb = bq;
                if (lane == 0) s_wc[q * NW + wave] = __popcll(bq);
            }
            __syncthreads();
            int off = 0;
#pragma unroll
            for (int q = 0; q < NFB; ++q) {
                int o = run[q], tot = 0;
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    const int v = s_wc[q * NW + w];
                    o += w < wave ? v : 0;
                    tot += v;
                }
                if (mq == q) off = o;
                run[q] += tot;
            }
            if (mq >= 0) wl_list(a, WL_FRONT + mq)[off + __popcll(mb & ((1ull << lane) - 1ull))] = cell_of(e);
            __syncthreads();
        }
        // the next tile resets the LDS counters and lists
        __syncthreads();
    }

    // temperature of the next state: one value per z-plane or per cell, the
    // arrays of all instances in one range
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < a0.n_temp; i += nth) {
        const double T = a0.T_in[i];
        a0.T_out[i] = a0.temp_f32 ? (double)((float)T - (float)a0.cooling) : T - a0.cooling;
    }
}

// Pass 2 over the queued cells: full APT merge.  Cells that need the
// solid-liquid anisotropy terms go to the interface lists of this pass
// (WL_LATE, taken by k_front_multi), cells with more active grains than the
// register capacity to k_step_general; the rest are finished here.
// Four waves per SIMD (at most 128 VGPRs, a few spilled): the pass waits on
// dependent loads of the slot lists, and the extra wave per SIMD measured
// 3.5% faster than the three the unbounded build allows.
// FP64: two waves per SIMD (202 VGPRs, none spilled); bounded to four, as
// in FP32, it spilled 82 VGPRs and took 1.05 instead of 0.70 ms at the
// mid-run state of 512x512x256.
template <int DIM, int TC, int B>
__global__ __launch_bounds__(256, sizeof(Real) == 8 ? 2 : 4) void k_merge(StepArgs<Real> a)
{
    if (a.halt_mode && a.stats->halt_step < a.step) return;
    __shared__ uint32_t s_coff[B ? MAX_INST : 1];
    load_cell_offsets<B>(a, s_coff);
    const Consts<Real> &K = a.K;
    const int n = *wl_counter(a, WL_MERGE, a.step);
    const size_t ncell = a.in.stride; // slot stride
    const bool need_terms = K.aniso || DIM == 2;
    if (blockIdx.x == 0 && threadIdx.x == 0 && n > 0) atomicAdd(&a.stats->merge_cells, (unsigned long long)n);
    // Same-address device atomics serialise (about 11 ns each on MI355X), so
    // the queued lanes of the whole block are appended with one atomic per
    // list and chunk.  s_off is double-buffered on the chunk parity: a wave
    // can only overwrite a buffer after every wave passed the barrier of the
    // chunk in between, by which time that buffer has been read.
    constexpr int NQ = 1 + NFB; // pushed lists: WL_GENERAL, then the late interface buckets
    __shared__ int s_off[2][NQ][4];
    __shared__ int s_base[2][NQ];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int par = 0;
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x, par ^= 1) {
        int i = base + threadIdx.x;
        bool valid = i < n;
        const int e = valid ? wl_list(a, WL_MERGE)[i] : 0;
        const size_t c = (size_t)(e & 0x7fffffff);
        const bool no_liq = e < 0;
        const Where W = locate<B>(a, s_coff, c);
        Cell<Real, DIM, FAST_A> C;
        int ntot = 0;
        bool fits = true, front = false;
        if (valid) {
            gather<Real, DIM, FAST_A>(W.K, W.in, ncell, W.l, W.m, W.k, C, &ntot, &fits, no_liq);
            front = fits && need_terms && C.liq && C.nsol > 0;
            // interface cells are counted by the interface pass
            if (!front) note_active(a.stats, ntot, K.kmax, a.step, a.halt_mode);
            if (fits && !front) {
                size_t ti = a.temp_dense ? W.c : (size_t)W.k;
                SlotLiq<Real> L{W.in.liq, W.K.nx, W.K.ny};
                // grain-boundary cells (no active liquid, 2-4 solids) take an
                // update compiled for their solid count
                if (!C.liq && C.nsol == 2) {
                    Cell<Real, DIM, 2> D = narrow_cell<2>(C);
                    cell_update<Real, DIM, 2, false, TC, true, 1, false>(W.K, L, W.l, W.m, W.k, W.T[ti], D);
                    write_cell<Real, DIM, 2>(W.out, ncell, W.c, D);
                } else if (!C.liq && C.nsol == 3) {
                    Cell<Real, DIM, 3> D = narrow_cell<3>(C);
                    cell_update<Real, DIM, 3, false, TC, true, 1, false>(W.K, L, W.l, W.m, W.k, W.T[ti], D);
                    write_cell<Real, DIM, 3>(W.out, ncell, W.c, D);
                } else {
                    cell_update<Real, DIM, FAST_A, false, TC>(W.K, L, W.l, W.m, W.k, W.T[ti], C);
                    write_cell<Real, DIM, FAST_A>(W.out, ncell, W.c, C);
                }
            }
        }
        // q = 0: the many-phase list; q = 1..NFB: interface cells with q solids
        const int myq = !valid ? -1 : (!fits ? 0 : (front ? min(C.nsol, NFB) : -1));
        unsigned long long bq = 0ull;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const unsigned long long b = __ballot(myq == q);
            if (q == myq) bq = b;
            if (lane == 0) s_off[par][q][wave] = __popcll(b);
        }
        __syncthreads();
        if (threadIdx.x < NQ) {
            int *o = s_off[par][threadIdx.x];
            int tot = 0;
            for (int w = 0; w < 4; ++w) {
                int v = o[w];
                o[w] = tot;
                tot += v;
            }
            const int list = threadIdx.x == 0 ? WL_GENERAL : WL_LATE + threadIdx.x - 1;
            s_base[par][threadIdx.x] = tot ? atomicAdd(wl_counter(a, list, a.step), tot) : 0;
        }
        __syncthreads();
        if (myq >= 0) {
            const int list = myq == 0 ? WL_GENERAL : WL_LATE + myq - 1;
            const unsigned long long below = (1ull << lane) - 1ull;
            wl_list(a, list)[s_base[par][myq] + s_off[par][myq][wave] + __popcll(bq & below)] = (int)c;
        }
    }
}

// FP64 interface kernels: held to two waves per SIMD (unbounded they take
// all 256 VGPRs and run one): 0.545 -> 0.409 ms at the mid-run state of
// 512x512x256; three or four waves spill far more.
constexpr int FRONT_WPE = sizeof(Real) == 8 ? 2 : 1;
// Pass 3: solid-liquid interface cells by bucket: the cells of a bucket of
// NS solids have an active liquid and exactly NS active solids (both the
// classification and the merge pass bucket by that count), so their update is
// compiled for that set with no presence tests and only the registers it
// needs.  k_front1 takes the classification pass's one-solid bucket (nearly
// all interface cells); k_front_multi, concurrently, the classification
// pass's buckets of 2-4 solids and every bucket of the merge pass.
static_assert(NFB == FAST_A, "interface buckets cover the register path's solid counts");
template <int DIM, int TC, int B, int NS, int G>
__device__ __forceinline__ void front_bucket(const StepArgs<Real> &a, const uint32_t *s_coff, int wl, int blk,
                                             int nblk)
{
    const Consts<Real> &K = a.K;
    const int n = *wl_counter(a, wl, a.step);
    const int *list = wl_list(a, wl);
    const size_t ncell = a.in.stride; // slot stride
    const bool lead = (threadIdx.x & (G - 1)) == 0;
    if (blk == 0 && threadIdx.x == 0 && n > 0) atomicAdd(&a.stats->front_cells, (unsigned long long)n);
    const int per = blockDim.x / G;
    for (int i = blk * per + (int)threadIdx.x / G; i < n; i += nblk * per) {
        const size_t c = (size_t)list[i];
        const Where W = locate<B>(a, s_coff, c);
        Cell<Real, DIM, NS> C;
        int ntot;
        bool fits;
        gather<Real, DIM, NS>(W.K, W.in, ncell, W.l, W.m, W.k, C, &ntot, &fits);
        if (lead) note_active(a.stats, ntot, K.kmax, a.step, a.halt_mode);
        size_t ti = a.temp_dense ? W.c : (size_t)W.k;
        SlotLiq<Real> L{W.in.liq, W.K.nx, W.K.ny};
        cell_update<Real, DIM, NS, true, TC, true, G>(W.K, L, W.l, W.m, W.k, W.T[ti], C);
        if (lead) write_cell<Real, DIM, NS>(W.out, ncell, W.c, C);
    }
}

template <int DIM, int TC, int B>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(FRONT_WPE))) void k_front1(StepArgs<Real> a)
{
    if (a.halt_mode && a.stats->halt_step < a.step) return;
    __shared__ uint32_t s_coff[B ? MAX_INST : 1];
    load_cell_offsets<B>(a, s_coff);
    front_bucket<DIM, TC, B, 1, 1>(a, s_coff, WL_FRONT, blockIdx.x, gridDim.x);
}

// The grid (at least 8 blocks) is split over the seven lists in proportion
// to cells x solids, with at least one block for a non-empty list; every
// block computes the same split from the counts.  (Lane groups per cell
// measured slower here: there are enough cells to keep every lane busy with
// a cell of its own.)
template <int DIM, int TC, int B>
__device__ __forceinline__ void front_multi(const StepArgs<Real> &a, const uint32_t *s_coff, int blk, int g)
{
    constexpr int NL = 2 * NFB - 1;
    // lists and their solid counts: the classification pass's buckets 2-4,
    // then the merge pass's buckets 1-4
    constexpr int wls[NL] = {WL_FRONT + 1, WL_FRONT + 2, WL_FRONT + 3, WL_LATE, WL_LATE + 1, WL_LATE + 2,
                             WL_LATE + 3};
    constexpr int nss[NL] = {2, 3, 4, 1, 2, 3, 4};
    long long w[NL], wt = 0;
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        w[q] = (long long)nss[q] * *wl_counter(a, wls[q], a.step);
        wt += w[q];
    }
    if (wt == 0) return;
    int gb[NL], sum = 0;
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        gb[q] = w[q] > 0 ? max(1, (int)(g * w[q] / wt)) : 0;
        sum += gb[q];
    }
    // give the rest to (or take the excess from) the largest shares; at most
    // NL - 1 lists were raised to one block, and g >= 8 > NL - 1
    while (sum != g) {
        int big = 0;
#pragma unroll
        for (int q = 1; q < NL; ++q)
            if (gb[q] > gb[big]) big = q;
        const int d = sum < g ? 1 : -1;
        gb[big] += d;
        sum += d;
    }
    int b = blk, q = 0;
    while (q < NL - 1 && b >= gb[q]) b -= gb[q++];
    switch (q) {
    case 0: front_bucket<DIM, TC, B, 2, 1>(a, s_coff, wls[0], b, gb[0]); break;
    case 1: front_bucket<DIM, TC, B, 3, 1>(a, s_coff, wls[1], b, gb[1]); break;
    case 2: front_bucket<DIM, TC, B, 4, 1>(a, s_coff, wls[2], b, gb[2]); break;
    case 3: front_bucket<DIM, TC, B, 1, 1>(a, s_coff, wls[3], b, gb[3]); break;
    case 4: front_bucket<DIM, TC, B, 2, 1>(a, s_coff, wls[4], b, gb[4]); break;
    case 5: front_bucket<DIM, TC, B, 3, 1>(a, s_coff, wls[5], b, gb[5]); break;
    default: front_bucket<DIM, TC, B, 4, 1>(a, s_coff, wls[6], b, gb[6]); break;
    }
}

template <int DIM, int TC, int B>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(FRONT_WPE))) void k_front_multi(StepArgs<Real> a)
{
    if (a.halt_mode && a.stats->halt_step < a.step) return;
    __shared__ uint32_t s_coff[B ? MAX_INST : 1];
    load_cell_offsets<B>(a, s_coff);
    front_multi<DIM, TC, B>(a, s_coff, blockIdx.x, gridDim.x);
}

// Pass 4: cells with more active grains than the merge pass holds.  Cells
// with up to MID_A active grains take unrolled register arrays, MID_A lanes
// per cell; beyond that one lane per cell keeps the arrays in scratch memory
// with room for KMAX phases and the loops run over the occupied entries only.
constexpr int MID_A = 8;
template <int DIM, int TC, int B>
__device__ __forceinline__ void step_general(const StepArgs<Real> &a, const uint32_t *s_coff, int blk, int nblk)
{
    const int n = *wl_counter(a, WL_GENERAL, a.step);
    const size_t ncell = a.in.stride; // slot stride
    if (blk == 0 && threadIdx.x == 0 && n > 0) atomicAdd(&a.stats->slow_cells, (unsigned long long)n);
    // MID_A lanes per cell on the register path (one per solid); lane 0 of
    // the group writes, and alone takes the scratch path
    constexpr int G = MID_A;
    const bool lead = (threadIdx.x & (G - 1)) == 0;
    const int per = blockDim.x / G;
    for (int i = blk * per + (int)threadIdx.x / G; i < n; i += nblk * per) {
        const size_t c = (size_t)wl_list(a, WL_GENERAL)[i];
        const Where W = locate<B>(a, s_coff, c);
        size_t ti = a.temp_dense ? W.c : (size_t)W.k;
        double Tn = W.T[ti];
        SlotLiq<Real> L{W.in.liq, W.K.nx, W.K.ny};
        int ntot;
        bool fits;
        {
            Cell<Real, DIM, MID_A> C;
            gather<Real, DIM, MID_A>(W.K, W.in, ncell, W.l, W.m, W.k, C, &ntot, &fits);
            if (fits) {
                cell_update<Real, DIM, MID_A, true, TC, false, G>(W.K, L, W.l, W.m, W.k, Tn, C);
                if (lead) write_cell<Real, DIM, MID_A>(W.out, ncell, W.c, C);
                continue;
            }
        }
        if (!lead) continue;
        Cell<Real, DIM, KCAP> C;
        gather<Real, DIM, KCAP>(W.K, W.in, ncell, W.l, W.m, W.k, C, &ntot, &fits);
        cell_update<Real, DIM, KCAP, true, TC>(W.K, L, W.l, W.m, W.k, Tn, C);
        write_cell<Real, DIM, KCAP>(W.out, ncell, W.c, C);
    }
}

template <int DIM, int TC, int B>
__global__ __launch_bounds__(64) void k_step_general(StepArgs<Real> a)
{
    if (a.halt_mode && a.stats->halt_step < a.step) return;
    __shared__ uint32_t s_coff[B ? MAX_INST : 1];
    load_cell_offsets<B>(a, s_coff);
    step_general<DIM, TC, B>(a, s_coff, blockIdx.x, gridDim.x);
}

// The three interface and many-phase passes as one launch: blocks
// [0, gg) take the many-phase cells, [gg, gg + gm) the multi-solid interface
// lists, the rest the one-solid interface cells.  The long serial cells get
// the first blocks, which start first; one launch needs no cross-stream
// events, whose fork and join cost about 15 us each on this GPU.
template <int DIM, int TC, int B>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(FRONT_WPE))) void k_interface(StepArgs<Real> a, int gg, int gm)
{
    if (a.halt_mode && a.stats->halt_step < a.step) return;
    __shared__ uint32_t s_coff[B ? MAX_INST : 1];
    load_cell_offsets<B>(a, s_coff);
    const int b = blockIdx.x;
    if (b < gg)
        step_general<DIM, TC, B>(a, s_coff, b, gg);
    else if (b < gg + gm)
        front_multi<DIM, TC, B>(a, s_coff, b - gg, gm);
    else
        front_bucket<DIM, TC, B, 1, 1>(a, s_coff, WL_FRONT, b - gg - gm, gridDim.x - gg - gm);
}

// Checked APT of the slot state in the reference's dense mf/nf layout
// (kernel_update_nfmf_3d_checked, src/gpu_kernels_3d.py:489-527).
#if MPF_HAS(3)
template <int DIM>
__global__ void k_apt_slots(Consts<Real> K, SlotView<Real> S, int32_t *mf, int32_t *nf, int32_t *status,
                            int depth)
{
    constexpr int NP = (DIM == 3) ? 7 : 5;
    const int nz = (DIM == 3) ? K.nz : 1;
    const size_t ncell = (size_t)K.nx * K.ny * nz;
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int l = (int)(c % K.nx);
    int m = (int)((c / K.nx) % K.ny);
    int k = (int)(c / ((size_t)K.nx * K.ny));
    size_t r = (DIM == 3) ? ridx(l, m, k, K.ny, K.nz) : ((size_t)l * K.ny + m);
    Merge<Real, DIM> M;
    stencil_cells<Real, DIM>(K, l, m, k, M.pc);
    Real lv[NP];
    for (int p = 0; p < NP; ++p) lv[p] = S.liq[M.pc[p]];
    int n = 0;
    if (apt_active<NP>(lv)) {
        n = 1;
        if (n <= K.kmax && n <= depth) mf[(size_t)(n - 1) * ncell + r] = 0;
    }
    M.init(S, S.stride);
    Real vals[NP];
    uint32_t g;
    while ((g = M.next(S, S.stride, vals)) != NO_ID) {
        if (!apt_active<NP>(vals)) continue;
        n += 1;
        if (n <= K.kmax && n <= depth) mf[(size_t)(n - 1) * ncell + r] = (int32_t)g;
    }
    if (n > K.kmax) {
        atomicMax(&status[0], 1);
        nf[r] = K.kmax;
    } else {
        nf[r] = n;
    }
    atomicMax(&status[1], n);
}

// Dense reference layout -> slots.  Cells holding more non-zero solids than
// the slot capacity raise err[0].
template <typename In>
__global__ void k_pack(const In *phi, int N, int nx, int ny, int nz, SlotView<Real> S, int scap, int *err)
{
    const size_t ncell = (size_t)nx * ny * nz;
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int l = (int)(c % nx);
    int m = (int)((c / nx) % ny);
    int k = (int)(c / ((size_t)nx * ny));
    size_t r = ridx(l, m, k, ny, nz);
    S.liq[c] = (Real)phi[r];
    int w = 0;
    for (int g = 1; g < N; ++g) {
        Real v = (Real)phi[(size_t)g * ncell + r];
        if (v != Real(0.0)) {
            if (w < scap) {
                S.sid[(size_t)w * S.stride + c] = (uint16_t)g;
                S.sval[(size_t)w * S.stride + c] = v;
            } else {
                atomicMax(err, 1);
            }
            w += 1;
        }
    }
    S.cnt[c] = (uint8_t)(w < scap ? w : scap);
}

// Slots -> dense reference layout (out must be zero-filled).
template <typename Out>
__global__ void k_unpack(SlotView<Real> S, int nx, int ny, int nz, Out *phi)
{
    const size_t ncell = (size_t)nx * ny * nz;
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int l = (int)(c % nx);
    int m = (int)((c / nx) % ny);
    int k = (int)(c / ((size_t)nx * ny));
    size_t r = ridx(l, m, k, ny, nz);
    phi[r] = (Out)S.liq[c];
    int n = S.cnt[c];
    for (int s = 0; s < n; ++s) {
        size_t o = (size_t)s * S.stride + c;
        phi[(size_t)S.sid[o] * ncell + r] = (Out)S.sval[o];
    }
}

// Columnar initial condition: cell (l,m,k) holds grain gmap[l,m] with value
// solid[prof[l,m]][k] and liquid liquid[prof[l,m]][k] (the seed layers of
// src/seed_modes_3d.py:38-57 and :117-166).
__global__ void k_init_columns(const int32_t *gmap, const int32_t *prof, const double *solid,
                               const double *liquid, int nx, int ny, int nz, SlotView<Real> S)
{
    const size_t ncell = (size_t)nx * ny * nz;
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int l = (int)(c % nx);
    int m = (int)((c / nx) % ny);
    int k = (int)(c / ((size_t)nx * ny));
    size_t col = (size_t)l * ny + m;
    int pi = prof[col];
    Real sv = (Real)solid[(size_t)pi * nz + k];
    S.liq[c] = (Real)liquid[(size_t)pi * nz + k];
    int g = gmap[col];
    if (sv != Real(0.0) && g > 0) {
        S.sid[c] = (uint16_t)g;
        S.sval[c] = sv;
        S.cnt[c] = 1;
    } else {
        S.cnt[c] = 0;
    }
}

// Temperature between layouts: dense reference [l][m][k] <-> slot order.
__global__ void k_temp_to_cells(const double *Tref, int nx, int ny, int nz, double *Tc)
{
    const size_t ncell = (size_t)nx * ny * nz;
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int l = (int)(c % nx);
    int m = (int)((c / nx) % ny);
    int k = (int)(c / ((size_t)nx * ny));
    Tc[c] = Tref[ridx(l, m, k, ny, nz)];
}

__global__ void k_temp_from_cells(const double *Tc, int nx, int ny, int nz, int dense, double *Tref)
{
    const size_t ncell = (size_t)nx * ny * nz;
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int l = (int)(c % nx);
    int m = (int)((c / nx) % ny);
    int k = (int)(c / ((size_t)nx * ny));
    Tref[ridx(l, m, k, ny, nz)] = dense ? Tc[c] : Tc[k];
}

// ------------------------------------------------- dense drop-in kernels
// kernel_update_nfmf_3d / _checked (src/gpu_kernels_3d.py:462-527) and
// kernel_update_nfmf (src/gpu_kernels.py:337-356) on phi[N][nx][ny](nz).
// Threads run along k (the contiguous index) for coalescing.
template <int DIM>
__global__ void k_dense_apt(const Real *phi, int32_t *mf, int32_t *nf, int32_t *status, int N, int nx,
                            int ny, int nz, int depth, int kmax, int checked)
{
    const int NZ = (DIM == 3) ? nz : 1;
    const size_t ncell = (size_t)nx * ny * NZ;
    size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ncell) return;
    int k = (int)(r % NZ);
    int m = (int)((r / NZ) % ny);
    int l = (int)(r / ((size_t)NZ * ny));
    int lp = bc_pp(l, nx), lm = bc_pm(l, nx);
    int mp, mm, kp = k, km = k;
    if (DIM == 3) {
        mp = bc_pp(m, ny); mm = bc_pm(m, ny);
        kp = bc_mp(k, nz); km = bc_mm(k, nz);
    } else {
        mp = bc_mp(m, ny); mm = bc_mm(m, ny);
    }
    size_t o_xp = ((size_t)lp * ny + m) * NZ + k, o_xm = ((size_t)lm * ny + m) * NZ + k;
    size_t o_yp = ((size_t)l * ny + mp) * NZ + k, o_ym = ((size_t)l * ny + mm) * NZ + k;
    size_t o_zp = ((size_t)l * ny + m) * NZ + kp, o_zm = ((size_t)l * ny + m) * NZ + km;
    int n = 0;
    for (int i = 0; i < N; ++i) {
        const Real *q = phi + (size_t)i * ncell;
        Real v = q[r];
        bool act = v > Real(0.0);
        if (!act && v == Real(0.0)) {
            act = (q[o_xp] > Real(0.0)) || (q[o_xm] > Real(0.0)) || (q[o_yp] > Real(0.0)) ||
                  (q[o_ym] > Real(0.0));
            if (DIM == 3) act = act || (q[o_zp] > Real(0.0)) || (q[o_zm] > Real(0.0));
        }
        if (act) {
            n += 1;
            if ((!checked || n <= kmax) && n <= depth) mf[(size_t)(n - 1) * ncell + r] = i;
        }
    }
    if (checked && n > kmax) {
        atomicMax(&status[0], 1);
        nf[r] = kmax;
    } else {
        nf[r] = n;
    }
    if (checked) atomicMax(&status[1], n);
}

// kernel_update_phasefield_active_3d(_switchable) (src/gpu_kernels_3d.py:530-830)
// and kernel_update_phasefield_active (src/gpu_kernels.py:361-515).
template <int DIM, int A>
__device__ __forceinline__ void dense_cell(const Consts<Real> &K, const Real *phi, Real *phi_new,
                                           const int32_t *mf, int n_act, size_t r, int l, int m, int k,
                                           double Tcell, size_t ncell)
{
    constexpr int NP = (DIM == 3) ? 7 : 5;
    constexpr bool SMALL = (A <= 8);
    constexpr int UNR_A = SMALL ? A : 1;
    const int NZ = (DIM == 3) ? K.nz : 1;
    int lp = bc_pp(l, K.nx), lm = bc_pm(l, K.nx);
    int mp, mm, kp = k, km = k;
    if (DIM == 3) {
        mp = bc_pp(m, K.ny); mm = bc_pm(m, K.ny);
        kp = bc_mp(k, K.nz); km = bc_mm(k, K.nz);
    } else {
        mp = bc_mp(m, K.ny); mm = bc_mm(m, K.ny);
    }
    size_t off[7];
    off[P_C] = r;
    off[P_XP] = ((size_t)lp * K.ny + m) * NZ + k;
    off[P_XM] = ((size_t)lm * K.ny + m) * NZ + k;
    off[P_YP] = ((size_t)l * K.ny + mp) * NZ + k;
    off[P_YM] = ((size_t)l * K.ny + mm) * NZ + k;
    off[P_ZP] = ((size_t)l * K.ny + m) * NZ + kp;
    off[P_ZM] = ((size_t)l * K.ny + m) * NZ + km;
    Cell<Real, DIM, A> C;
    C.gid[0] = 0;
    C.liq = false;
    C.nsol = 0;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        C.v[0][p] = phi[off[p]];
        C.dom1[p] = (K.N > 1) ? phi[ncell + off[p]] : Real(0.0);
    }
    // mf is ascending with the liquid first (APT order)
    for (int t = 0; t < n_act; ++t) {
        int g = mf[(size_t)t * ncell + r];
        if (g == 0) {
            C.liq = true;
            continue;
        }
        int u = C.nsol + 1;
        if constexpr (SMALL) {
#pragma unroll
            for (int w = 1; w <= A; ++w) {
                if (w == u) {
                    C.gid[w] = g;
#pragma unroll
                    for (int p = 0; p < NP; ++p) C.v[w][p] = phi[(size_t)g * ncell + off[p]];
                }
            }
        } else {
            C.gid[u] = g;
            for (int p = 0; p < NP; ++p) C.v[u][p] = phi[(size_t)g * ncell + off[p]];
        }
        C.nsol = u;
    }
    DenseLiq<Real> L{phi, K.ny, NZ};
    cell_update<Real, DIM, A>(K, L, l, m, k, Tcell, C);
    for (int g = 0; g < K.N; ++g) phi_new[(size_t)g * ncell + r] = Real(0.0);
    phi_new[r] = C.out[0];
#pragma unroll UNR_A
    for (int u = 1; u <= A; ++u)
        if (u <= C.nsol) phi_new[(size_t)C.gid[u] * ncell + r] = C.out[u];
}

template <int DIM>
__global__ __launch_bounds__(128) void k_dense_update(Consts<Real> K, const Real *phi, Real *phi_new,
                                                      const void *temp, int temp_f64, const int32_t *mf,
                                                      const int32_t *nf, int depth)
{
    const int NZ = (DIM == 3) ? K.nz : 1;
    const size_t ncell = (size_t)K.nx * K.ny * NZ;
    size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ncell) return;
    int k = (int)(r % NZ);
    int m = (int)((r / NZ) % K.ny);
    int l = (int)(r / ((size_t)NZ * K.ny));
    double T = temp_f64 ? ((const double *)temp)[r] : (double)((const float *)temp)[r];
    // nf may exceed the ids written to mf for the unchecked APT (undefined in
    // the reference); only the stored ids are used.
    int n_act = nf[r];
    n_act = n_act < depth ? n_act : depth;
    n_act = n_act < KCAP + 1 ? n_act : KCAP + 1;
    if (n_act <= FAST_A)
        dense_cell<DIM, FAST_A>(K, phi, phi_new, mf, n_act, r, l, m, k, T, ncell);
    else
        dense_cell<DIM, KCAP>(K, phi, phi_new, mf, n_act, r, l, m, k, T, ncell);
}

// kernel_update_temp_3d (src/gpu_kernels_3d.py:833-838)
__global__ void k_dense_temp(void *temp, int temp_f64, double cooling, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (temp_f64)
        ((double *)temp)[i] -= cooling;
    else
        ((float *)temp)[i] = ((float *)temp)[i] - (float)cooling;
}

// FP64 observables of a run of instances (tests/verification_3d.py:409-416,
// :612-654), one launch.  pre[0..n] are the instances' first plane blocks and
// pre[n+1..2n+1] their first column blocks (counted after all plane blocks),
// pre[2n+2..3n+2] their first columns in front[].
//
// A plane block (one z-plane of one instance): per-grain sums with LDS
// atomics, added to gvol[q][g]; scalar partials by a deterministic block
// reduction into part[block][6]: max|sum-1|, min, max, NaN count, sum of
// (1 - phi_L), stored grain slots.  A column block: front[l][m] of 256
// columns, the topmost k with 1 - phi_L > 0.5 (-1 when none).
MPF_DEV int obs_instance(const long long *pre, int n, long long b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_observe(SlotView<Real> S, const Inst *inst, const long long *pre, int n,
                                                 int N, double *gvol, double *part, int32_t *front)
{
    const long long b = blockIdx.x, planes = pre[n];
    if (b >= planes) {
        const long long cb = b - planes;
        const int q = obs_instance(pre + n + 1, n, cb);
        const Inst g = inst[q];
        const long long col = (cb - pre[n + 1 + q]) * 256 + threadIdx.x;
        if (col >= (long long)g.nx * g.ny) return;
        const int l = (int)(col % g.nx), m = (int)(col / g.nx);
        const Real *liq = S.liq + g.cell_off;
        int f = -1;
        for (int k = g.nz - 1; k >= 0; --k) {
            if (1.0 - (double)liq[cidx(l, m, k, g.nx, g.ny)] > 0.5) {
                f = k;
                break;
            }
        }
        front[pre[2 * n + 2 + q] + (long long)l * g.ny + m] = f;
        return;
    }
    extern __shared__ double sh[];
    double *vol = sh;                  // [N]
    double *red = sh + N;              // [256*6]
    const int q = obs_instance(pre, n, b);
    const Inst g = inst[q];
    const int nxy = g.nx * g.ny;
    const size_t c0 = g.cell_off + (size_t)(b - pre[q]) * nxy;
    for (int i = threadIdx.x; i < N; i += blockDim.x) vol[i] = 0.0;
    __syncthreads();
    double e_max = 0.0, p_min = 1e300, p_max = -1e300, nan_c = 0.0, solid = 0.0, slots = 0.0;
    bool any_absent = false;
    for (int i = threadIdx.x; i < nxy; i += blockDim.x) {
        const size_t c = c0 + i;
        double lq = (double)S.liq[c];
        double sum = lq;
        int ns = S.cnt[c];
        slots += (double)ns;
        if (lq != lq) nan_c += 1.0;
        p_min = lq < p_min ? lq : p_min;
        p_max = lq > p_max ? lq : p_max;
        atomicAdd(&vol[0], lq);
        for (int s = 0; s < ns; ++s) {
            size_t o = (size_t)s * S.stride + c;
            double v = (double)S.sval[o];
            sum += v;
            if (v != v) nan_c += 1.0;
            p_min = v < p_min ? v : p_min;
            p_max = v > p_max ? v : p_max;
            atomicAdd(&vol[S.sid[o]], v);
        }
        if (ns + 1 < N) any_absent = true;
        double e = fabs(sum - 1.0);
        e_max = e > e_max ? e : e_max;
        solid += 1.0 - lq;
    }
    if (any_absent) {
        p_min = 0.0 < p_min ? 0.0 : p_min;
        p_max = 0.0 > p_max ? 0.0 : p_max;
    }
    double *r = red + threadIdx.x * 6;
    r[0] = e_max; r[1] = p_min; r[2] = p_max; r[3] = nan_c; r[4] = solid; r[5] = slots;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            double *o = red + (threadIdx.x + h) * 6;
            r[0] = o[0] > r[0] ? o[0] : r[0];
            r[1] = o[1] < r[1] ? o[1] : r[1];
            r[2] = o[2] > r[2] ? o[2] : r[2];
            r[3] += o[3];
            r[4] += o[4];
            r[5] += o[5];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int i = 0; i < 6; ++i) part[(size_t)b * 6 + i] = red[i];
    double *gv = gvol + (size_t)q * N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) atomicAdd(&gv[i], vol[i]);
}

// ------------------------------------------------------------ ray probes
// Directional radius probes of the orientation scan
// (run_orientation_scan_3d.py:236-331): march from an origin along a unit
// direction in steps of step_cells, sampling phase g by trilinear
// interpolation, and return the distance (cells) of the first downward
// crossing of the threshold, linearly interpolated; NaN when the origin is
// below the threshold or outside, or no crossing is met before the grid
// boundary.  Contraction is off so that every sum and product rounds as in
// the reference's double-precision Python.

// Value of phase g at cell c: the liquid array for g = 0, else the slot that
// holds g (0 when the cell stores no slot for g).
__device__ __forceinline__ double phase_value(const SlotView<Real> &S, size_t c, int g)
{
    if (g == 0) return (double)S.liq[c];
    const int n = S.cnt[c];
    for (int s = 0; s < n; ++s) {
        const size_t o = (size_t)s * S.stride + c;
        if ((int)S.sid[o] == g) return (double)S.sval[o];
    }
    return 0.0;
}

// trilinear_sample (run_orientation_scan_3d.py:236-267)
__device__ double probe_sample(const SlotView<Real> &S, int nx, int ny, int nz, int g, double x, double y,
                               double z)
{
#pragma clang fp contract(off)
    if (x < 0.0 || y < 0.0 || z < 0.0 || x > (double)(nx - 1) || y > (double)(ny - 1) || z > (double)(nz - 1))
        return __builtin_nan("");
    const int x0 = (int)floor(x), y0 = (int)floor(y), z0 = (int)floor(z);
    const int x1 = min(x0 + 1, nx - 1), y1 = min(y0 + 1, ny - 1), z1 = min(z0 + 1, nz - 1);
    const double tx = x - (double)x0, ty = y - (double)y0, tz = z - (double)z0;
    auto at = [&](int l, int m, int k) {
        return phase_value(S, (size_t)l + (size_t)nx * ((size_t)m + (size_t)ny * (size_t)k), g);
    };
    const double c000 = at(x0, y0, z0), c100 = at(x1, y0, z0), c010 = at(x0, y1, z0), c110 = at(x1, y1, z0);
    const double c001 = at(x0, y0, z1), c101 = at(x1, y0, z1), c011 = at(x0, y1, z1), c111 = at(x1, y1, z1);
    const double c00 = c000 * (1.0 - tx) + c100 * tx;
    const double c01 = c001 * (1.0 - tx) + c101 * tx;
    const double c10 = c010 * (1.0 - tx) + c110 * tx;
    const double c11 = c011 * (1.0 - tx) + c111 * tx;
    const double c0 = c00 * (1.0 - ty) + c10 * ty;
    const double c1 = c01 * (1.0 - ty) + c11 * ty;
    return c0 * (1.0 - tz) + c1 * tz;
}

// One ray per thread: max_distance_to_boundary and directional_radius_one_side
// (run_orientation_scan_3d.py:270-331).
__global__ void k_probe(SlotView<Real> S, const Inst *inst, int n_ray, const int32_t *rinst,
                        const int32_t *rphase, const double *origin, const double *dir, double thr,
                        double step, double *out)
{
#pragma clang fp contract(off)
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_ray) return;
    const Inst d = inst[rinst[r]];
    SlotView<Real> V = S;
    V.liq += d.cell_off;
    V.cnt += d.cell_off;
    V.sid += d.cell_off;
    V.sval += d.cell_off;
    const int g = rphase[r];
    const double c[3] = {origin[3 * r], origin[3 * r + 1], origin[3 * r + 2]};
    const double u[3] = {dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]};
    const double up[3] = {(double)d.nx - 1.0, (double)d.ny - 1.0, (double)d.nz - 1.0};
    double max_s = 0.0;
    bool any = false;
    for (int a = 0; a < 3; ++a) {
        if (fabs(u[a]) < 1.0e-12) continue;
        const double lim = u[a] > 0.0 ? (up[a] - c[a]) / u[a] : (0.0 - c[a]) / u[a];
        if (lim > 0.0) {
            max_s = (!any || lim < max_s) ? lim : max_s;
            any = true;
        }
    }
    double res = __builtin_nan("");
    if (max_s > 0.0) {
        const double v0 = probe_sample(V, d.nx, d.ny, d.nz, g, c[0], c[1], c[2]);
        if (isfinite(v0) && !(v0 < thr)) {
            double prev_s = 0.0, prev_v = v0;
            long n_steps = (long)ceil(max_s / step) + 1;
            n_steps = n_steps > 2 ? n_steps : 2;
            for (long i = 1; i <= n_steps; ++i) {
                const double si = (double)i * step;
                const double s = si < max_s ? si : max_s;
                const double v = probe_sample(V, d.nx, d.ny, d.nz, g, c[0] + u[0] * s, c[1] + u[1] * s,
                                              c[2] + u[2] * s);
                if (!isfinite(v)) break;
                if (prev_v >= thr && v < thr) {
                    const double dv = prev_v - v;
                    if (fabs(dv) < 1.0e-12) {
                        res = s;
                        break;
                    }
                    double frac = (prev_v - thr) / dv;
                    frac = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
                    res = prev_s + frac * (s - prev_s);
                    break;
                }
                prev_s = s;
                prev_v = v;
            }
        }
    }
    out[r] = res;
}

// ------------------------------------------------------------ phase maps
// Dominant phase of every cell of one plane: np.argmax over the phases with
// the lowest id on ties (plot_utils_3d.save_phase_map_slice_3d,
// src/plot_utils_3d.py:36-45).  axis 0: the xy plane k = index, out[l][m];
// axis 1: xz, m = index, out[l][k]; axis 2: yz, l = index, out[m][k].
__global__ void k_phase_map(SlotView<Real> S, int nx, int ny, int nz, int axis, int index, int32_t *out)
{
    const int n0 = axis == 2 ? ny : nx;
    const int n1 = axis == 0 ? ny : nz;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n0 * n1) return;
    const int a = p / n1, b = p % n1;
    const int l = axis == 2 ? index : a;
    const int m = axis == 0 ? b : (axis == 1 ? index : a);
    const int k = axis == 0 ? index : b;
    const size_t c = (size_t)l + (size_t)nx * ((size_t)m + (size_t)ny * (size_t)k);
    Real best = S.liq[c];
    int g = 0;
    const int n = S.cnt[c];
    for (int s = 0; s < n; ++s) {
        const size_t o = (size_t)s * S.stride + c;
        const Real v = S.sval[o];
        const int id = S.sid[o];
        if (v > best || (v == best && id < g)) {
            best = v;
            g = id;
        }
    }
    out[p] = g;
}

#endif // MPF_HAS(3)

} // namespace impl_*
} // namespace mpf

// ------------------------------------------------------------ launchers
using namespace mpf;
using namespace mpf::MPF_CAT(impl, MPF_SFX);

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// One step.  Fused (the default): after the merge pass one launch of
// k_interface takes the many-phase cells and all interface cells.  With side
// streams, after the merge pass the many-phase kernel
// and the interface cells of several solids run concurrently with the
// one-solid interface kernel (every one of these kernels only reads the input
// state and writes cells no other writes); the solver stream waits for both
// before the step ends.  (Running k_front1 concurrently with the merge pass
// instead measured slower: both are throughput-bound there.)  tev
// (optional): 8 timing events: [0] before k_tiles, [1] after k_light, [2]
// after k_merge, [3]/[4] around k_front1 and [5]/[6] around k_step_general
// ([7] unused).
template <int DIM, int TC, int B>
static void launch_step_kernels(const StepArgs<Real> &a, unsigned grid, const int *grids, hipStream_t st,
                                const StepStreams &S, hipEvent_t *tev)
{
    const int gm = max(grids[1], 8);
    if (tev) hipEventRecord(tev[0], st);
    hipLaunchKernelGGL((k_tiles<DIM, B>), dim3(nblk((size_t)a.n_tiles, 256)), dim3(256), 0, st, a);
    // 32-bit offsets when every byte offset of the light pass fits (4- or
    // 8-byte values per cell are the widest elements it indexes)
    const bool off32 = a.in.stride * 8 <= 0xffffffffull;
    if (off32)
        hipLaunchKernelGGL((k_light<DIM, uint32_t, B>), dim3(grid), dim3(LTX * LTY), 0, st, a);
    else
        hipLaunchKernelGGL((k_light<DIM, size_t, B>), dim3(grid), dim3(LTX * LTY), 0, st, a);
    if (tev) hipEventRecord(tev[1], st);
    hipLaunchKernelGGL((k_merge<DIM, TC, B>), dim3(grids[0]), dim3(256), 0, st, a);
    if (tev) hipEventRecord(tev[2], st);
    if (S.side) {
        hipEventRecord(S.ev_merged, st);
        hipStreamWaitEvent(S.side, S.ev_merged, 0);
        hipStreamWaitEvent(S.side2, S.ev_merged, 0);
        if (tev) hipEventRecord(tev[5], S.side);
        hipLaunchKernelGGL((k_step_general<DIM, TC, B>), dim3(grids[2]), dim3(64), 0, S.side, a);
        if (tev) hipEventRecord(tev[6], S.side);
        hipEventRecord(S.ev_general, S.side);
        hipLaunchKernelGGL((k_front_multi<DIM, TC, B>), dim3(gm), dim3(128), 0, S.side2, a);
        hipEventRecord(S.ev_front, S.side2);
        if (tev) hipEventRecord(tev[3], st);
        hipLaunchKernelGGL((k_front1<DIM, TC, B>), dim3(grids[1]), dim3(128), 0, st, a);
        if (tev) hipEventRecord(tev[4], st);
        hipStreamWaitEvent(st, S.ev_general, 0);
        hipStreamWaitEvent(st, S.ev_front, 0);
    } else if (S.fused) {
        // general blocks in units of 128 threads
        const int gg = max(grids[2] / 2, 1);
        if (tev) hipEventRecord(tev[3], st);
        hipLaunchKernelGGL((k_interface<DIM, TC, B>), dim3(gg + gm + grids[1]), dim3(128), 0, st, a, gg, gm);
        if (tev) hipEventRecord(tev[4], st);
        if (tev) hipEventRecord(tev[5], st);
        if (tev) hipEventRecord(tev[6], st);
    } else {
        if (tev) hipEventRecord(tev[3], st);
        hipLaunchKernelGGL((k_front1<DIM, TC, B>), dim3(grids[1]), dim3(128), 0, st, a);
        if (tev) hipEventRecord(tev[4], st);
        hipLaunchKernelGGL((k_front_multi<DIM, TC, B>), dim3(gm), dim3(128), 0, st, a);
        if (tev) hipEventRecord(tev[5], st);
        hipLaunchKernelGGL((k_step_general<DIM, TC, B>), dim3(grids[2]), dim3(64), 0, st, a);
        if (tev) hipEventRecord(tev[6], st);
    }
}

extern "C" {

// The step of one kind of grid: 3D, a batch of 3D grids, 2D.
hipError_t MPF_NAME(mpf_step_3d)(const StepArgs<Real> &a, unsigned grid, const int *grids, hipStream_t st,
                                 const StepStreams *ss, hipEvent_t *tev);
hipError_t MPF_NAME(mpf_step_batch)(const StepArgs<Real> &a, unsigned grid, const int *grids, hipStream_t st,
                                    const StepStreams *ss, hipEvent_t *tev);
hipError_t MPF_NAME(mpf_step_2d)(const StepArgs<Real> &a, unsigned grid, const int *grids, hipStream_t st,
                                 const StepStreams *ss, hipEvent_t *tev);

#if MPF_HAS(0)
hipError_t MPF_NAME(mpf_step_3d)(const StepArgs<Real> &a, unsigned grid, const int *grids, hipStream_t st,
                                 const StepStreams *ss, hipEvent_t *tev)
{
    if (a.tab_class)
        launch_step_kernels<3, 1, 0>(a, grid, grids, st, *ss, tev);
    else
        launch_step_kernels<3, 0, 0>(a, grid, grids, st, *ss, tev);
    return hipGetLastError();
}
#endif

#if MPF_HAS(1)
hipError_t MPF_NAME(mpf_step_batch)(const StepArgs<Real> &a, unsigned grid, const int *grids, hipStream_t st,
                                    const StepStreams *ss, hipEvent_t *tev)
{
    if (a.tab_class)
        launch_step_kernels<3, 1, 1>(a, grid, grids, st, *ss, tev);
    else
        launch_step_kernels<3, 0, 1>(a, grid, grids, st, *ss, tev);
    return hipGetLastError();
}
#endif

#if MPF_HAS(2)
hipError_t MPF_NAME(mpf_step_2d)(const StepArgs<Real> &a, unsigned grid, const int *grids, hipStream_t st,
                                 const StepStreams *ss, hipEvent_t *tev)
{
    if (a.tab_class)
        launch_step_kernels<2, 1, 0>(a, grid, grids, st, *ss, tev);
    else
        launch_step_kernels<2, 0, 0>(a, grid, grids, st, *ss, tev);
    return hipGetLastError();
}
#endif

#if MPF_HAS(3)
hipError_t MPF_NAME(mpf_launch_step)(const StepArgs<Real> *args, int dim, const int *grids, hipStream_t st,
                                     const StepStreams *ss, hipEvent_t *tev)
{
    StepArgs<Real> a = *args;
    // persistent classification grid (by default 1024 blocks, four per CU of
    // the 256), at most one per work item
    const unsigned grid = (unsigned)min(a.n_tiles * (LZC / a.light_q), a.light_grid > 0 ? a.light_grid : 1024);
    if (a.n_inst > 0) // a batch of 3D grids: one launch per kernel covers every instance
        return MPF_NAME(mpf_step_batch)(a, grid, grids, st, ss, tev);
    a.ntx = (a.K.nx + LTX - 1) / LTX;
    a.nty = (a.K.ny + LTY - 1) / LTY;
    if (dim == 3) return MPF_NAME(mpf_step_3d)(a, grid, grids, st, ss, tev);
    return MPF_NAME(mpf_step_2d)(a, grid, grids, st, ss, tev);
}

hipError_t MPF_NAME(mpf_launch_apt_slots)(const Consts<Real> *K, SlotView<Real> S, int dim, int32_t *mf,
                                          int32_t *nf, int32_t *status, int depth, hipStream_t st)
{
    size_t ncell = (size_t)K->nx * K->ny * (dim == 3 ? K->nz : 1);
    if (dim == 3)
        hipLaunchKernelGGL(k_apt_slots<3>, dim3(nblk(ncell, 256)), dim3(256), 0, st, *K, S, mf, nf, status,
                           depth);
    else
        hipLaunchKernelGGL(k_apt_slots<2>, dim3(nblk(ncell, 256)), dim3(256), 0, st, *K, S, mf, nf, status,
                           depth);
    return hipGetLastError();
}

hipError_t MPF_NAME(mpf_launch_pack)(const void *phi, int phi_f64, int N, int nx, int ny, int nz,
                                     SlotView<Real> S, int scap, int *err, hipStream_t st)
{
    size_t ncell = (size_t)nx * ny * nz;
    if (phi_f64)
        hipLaunchKernelGGL(k_pack<double>, dim3(nblk(ncell, 256)), dim3(256), 0, st, (const double *)phi, N,
                           nx, ny, nz, S, scap, err);
    else
        hipLaunchKernelGGL(k_pack<float>, dim3(nblk(ncell, 256)), dim3(256), 0, st, (const float *)phi, N,
                           nx, ny, nz, S, scap, err);
    return hipGetLastError();
}

hipError_t MPF_NAME(mpf_launch_unpack)(SlotView<Real> S, int nx, int ny, int nz, void *phi, int phi_f64,
                                       hipStream_t st)
{
    size_t ncell = (size_t)nx * ny * nz;
    if (phi_f64)
        hipLaunchKernelGGL(k_unpack<double>, dim3(nblk(ncell, 256)), dim3(256), 0, st, S, nx, ny, nz,
                           (double *)phi);
    else
        hipLaunchKernelGGL(k_unpack<float>, dim3(nblk(ncell, 256)), dim3(256), 0, st, S, nx, ny, nz,
                           (float *)phi);
    return hipGetLastError();
}

hipError_t MPF_NAME(mpf_launch_init_columns)(const int32_t *gmap, const int32_t *prof, const double *solid,
                                             const double *liquid, int nx, int ny, int nz,
                                             SlotView<Real> S, hipStream_t st)
{
    size_t ncell = (size_t)nx * ny * nz;
    hipLaunchKernelGGL(k_init_columns, dim3(nblk(ncell, 256)), dim3(256), 0, st, gmap, prof, solid, liquid,
                       nx, ny, nz, S);
    return hipGetLastError();
}

hipError_t MPF_NAME(mpf_launch_temp_layout)(const double *src, double *dst, int nx, int ny, int nz,
                                            int to_cells, int dense, hipStream_t st)
{
    size_t ncell = (size_t)nx * ny * nz;
    if (to_cells)
        hipLaunchKernelGGL(k_temp_to_cells, dim3(nblk(ncell, 256)), dim3(256), 0, st, src, nx, ny, nz, dst);
    else
        hipLaunchKernelGGL(k_temp_from_cells, dim3(nblk(ncell, 256)), dim3(256), 0, st, src, nx, ny, nz,
                           dense, dst);
    return hipGetLastError();
}

hipError_t MPF_NAME(mpf_launch_dense_apt)(const Real *phi, int32_t *mf, int32_t *nf, int32_t *status, int N,
                                          int nx, int ny, int nz, int dim, int depth, int kmax, int checked,
                                          hipStream_t st)
{
    size_t ncell = (size_t)nx * ny * (dim == 3 ? nz : 1);
    if (dim == 3)
        hipLaunchKernelGGL(k_dense_apt<3>, dim3(nblk(ncell, 128)), dim3(128), 0, st, phi, mf, nf, status, N,
                           nx, ny, nz, depth, kmax, checked);
    else
        hipLaunchKernelGGL(k_dense_apt<2>, dim3(nblk(ncell, 128)), dim3(128), 0, st, phi, mf, nf, status, N,
                           nx, ny, nz, depth, kmax, checked);
    return hipGetLastError();
}

hipError_t MPF_NAME(mpf_launch_dense_update)(const Consts<Real> *K, int dim, const Real *phi, Real *phi_new,
                                             const void *temp, int temp_f64, const int32_t *mf,
                                             const int32_t *nf, int depth, hipStream_t st)
{
    size_t ncell = (size_t)K->nx * K->ny * (dim == 3 ? K->nz : 1);
    if (dim == 3)
        hipLaunchKernelGGL(k_dense_update<3>, dim3(nblk(ncell, 128)), dim3(128), 0, st, *K, phi, phi_new,
                           temp, temp_f64, mf, nf, depth);
    else
        hipLaunchKernelGGL(k_dense_update<2>, dim3(nblk(ncell, 128)), dim3(128), 0, st, *K, phi, phi_new,
                           temp, temp_f64, mf, nf, depth);
    return hipGetLastError();
}

hipError_t MPF_NAME(mpf_launch_dense_temp)(void *temp, int temp_f64, double cooling, size_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_dense_temp, dim3(nblk(n, 256)), dim3(256), 0, st, temp, temp_f64, cooling, n);
    return hipGetLastError();
}

hipError_t MPF_NAME(mpf_launch_probe)(SlotView<Real> S, const Inst *inst, int n_ray, const int32_t *rinst,
                                      const int32_t *rphase, const double *origin, const double *dir, double thr,
                                      double step, double *out, hipStream_t st)
{
    if (n_ray <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_probe, dim3(nblk((size_t)n_ray, 64)), dim3(64), 0, st, S, inst, n_ray, rinst, rphase,
                       origin, dir, thr, step, out);
    return hipGetLastError();
}

hipError_t MPF_NAME(mpf_launch_phase_map)(SlotView<Real> S, int nx, int ny, int nz, int axis, int index,
                                          int32_t *out, hipStream_t st)
{
    const int n0 = axis == 2 ? ny : nx, n1 = axis == 0 ? ny : nz;
    hipLaunchKernelGGL(k_phase_map, dim3(nblk((size_t)n0 * n1, 256)), dim3(256), 0, st, S, nx, ny, nz, axis, index,
                       out);
    return hipGetLastError();
}

hipError_t MPF_NAME(mpf_launch_observe)(SlotView<Real> S, const Inst *inst, const long long *pre, int n,
                                        long long blocks, int N, double *gvol, double *part, int32_t *front,
                                        hipStream_t st)
{
    size_t shm = ((size_t)N + 256 * 6) * sizeof(double);
    hipLaunchKernelGGL(k_observe, dim3((unsigned)blocks), dim3(256), shm, st, S, inst, pre, n, N, gvol, part, front);
    return hipGetLastError();
}

#endif // MPF_HAS(3)

} // extern "C"
