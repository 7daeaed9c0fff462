// mpf_api.cpp -- host side of the C ABI declared in include/mpf.h.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>
#include <vector>

#include "mpf.h"
#include "mpf_math.h"
#include "mpf_consts.h"
#include "mpf_state.h"
#include "mpf_streams.h"

using namespace mpf;

// ------------------------------------------------ launchers of both builds
#define MPF_DECLARE_LAUNCHERS(R, SFX)                                                                    \
    extern "C" hipError_t mpf_launch_step##SFX(const StepArgs<R> *, int, const int *, hipStream_t,           \
                                               const StepStreams *, hipEvent_t *);                          \
    extern "C" hipError_t mpf_launch_apt_slots##SFX(const Consts<R> *, SlotView<R>, int, int32_t *,      \
                                                    int32_t *, int32_t *, int, hipStream_t);             \
    extern "C" hipError_t mpf_launch_pack##SFX(const void *, int, int, int, int, int, SlotView<R>, int,  \
                                               int *, hipStream_t);                                      \
    extern "C" hipError_t mpf_launch_unpack##SFX(SlotView<R>, int, int, int, void *, int, hipStream_t);  \
    extern "C" hipError_t mpf_launch_init_columns##SFX(const int32_t *, const int32_t *, const double *, \
                                                       const double *, int, int, int, SlotView<R>,      \
                                                       hipStream_t);                                     \
    extern "C" hipError_t mpf_launch_temp_layout##SFX(const double *, double *, int, int, int, int, int, \
                                                      hipStream_t);                                      \
    extern "C" hipError_t mpf_launch_dense_apt##SFX(const R *, int32_t *, int32_t *, int32_t *, int,     \
                                                    int, int, int, int, int, int, int, hipStream_t);     \
    extern "C" hipError_t mpf_launch_dense_update##SFX(const Consts<R> *, int, const R *, R *,           \
                                                       const void *, int, const int32_t *,               \
                                                       const int32_t *, int, hipStream_t);               \
    extern "C" hipError_t mpf_launch_dense_temp##SFX(void *, int, double, size_t, hipStream_t);          \
    extern "C" hipError_t mpf_launch_observe##SFX(SlotView<R>, const Inst *, const long long *, int,     \
                                                  long long, int, double *, double *, int32_t *,          \
                                                  hipStream_t);                                           \
    extern "C" hipError_t mpf_launch_probe##SFX(SlotView<R>, const Inst *, int, const int32_t *,         \
                                                const int32_t *, const double *, const double *, double,  \
                                                double, double *, hipStream_t);                           \
    extern "C" hipError_t mpf_launch_phase_map##SFX(SlotView<R>, int, int, int, int, int, int32_t *,      \
                                                    hipStream_t);

MPF_DECLARE_LAUNCHERS(double, _f64)
MPF_DECLARE_LAUNCHERS(float, _f32)

// ------------------------------------------------------------------ errors
static thread_local char g_err[512] = "";

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(MPF_EHIP, "%s: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

extern "C" const char *mpf_last_error(void) { return g_err; }
extern "C" int mpf_abi_version(void) { return MPF_ABI_VERSION; }

// ------------------------------------------------------------------ handle
struct Buffers {
    void *liq = nullptr;
    uint8_t *cnt = nullptr;
    uint16_t *sid = nullptr;
    void *sval = nullptr;
    uint16_t *occ = nullptr; // occupancy mask (occ_words entries)
};

struct mpf_handle {
    mpf_config cfg;
    int dim = 3;
    size_t ncell = 0; // cells over all instances
    size_t nocc = 0;  // occupancy mask words per buffer
    int scap = 1;
    size_t rsz = 8; // sizeof(Real)
    int batch = 0;                  // created by mpf_batch_create
    std::vector<Inst> inst;         // one entry per grid; a single grid is instance 0
    Inst *inst_d = nullptr;         // device copy
    int n_tiles = 0;                // classification tiles over all instances
    size_t tab_stride = 0;          // table elements per instance (batches)
    std::vector<int> tab_set;       // per instance: tables uploaded
    std::vector<int> tab_cls;       // per instance: tables have the class form
    std::vector<int> tab_pairs;     // per instance: n111 rows in opposite pairs
    std::vector<double> cls_vals;   // per instance: 6 class values
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;            // many-phase kernel, concurrent with the interface kernel
    hipStream_t side2 = nullptr;           // interface cells of several solids, concurrent as well
    hipEvent_t ev_merged = nullptr, ev_general = nullptr, ev_front = nullptr;
    int kernel_timing = 0;                 // per-kernel events (mpf_set_kernel_timing)
    std::vector<hipEvent_t> tev;           // 8 per step of the current call
    double kern_ms[4] = {0, 0, 0, 0};      // k_tiles + k_light, k_merge, k_front1, k_step_general
    int kern_steps = 0;
    Buffers b[2];
    double *T[2] = {nullptr, nullptr};
    size_t nT = 0;
    void *tab = nullptr; // per instance: wij, aij, mij, n111 contiguous
    int *wl = nullptr;        // NWL work lists of ncell entries (mpf_state.h)
    int *wl_count = nullptr;  // [NWC][2]
    int *tiles = nullptr;     // work items of the classification pass (n_tiles * LZC / 4 entries)
    uint16_t *item_busy = nullptr; // their tiles' plane masks
    int light_q_req = 0;      // planes per classification work item (0: by grid size)
    int light_grid_req = 0;   // blocks of the classification pass (0: 1024)
    Stats *stats = nullptr;
    int cur = 0, tcur = 0;
    int step_counter = 0;
    int steps_done = 0;
    int tables_set = 0;
    int tab_class = 0;    // every instance's tables have the same class form
    int n111_pairs = 0;   // every instance's n111 rows come in opposite pairs
    double cls[3][2] = {}; // its values, exact in the solver precision
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    size_t bytes = 0;
    // launch shape: blocks of the merge, interface and many-phase kernels
    // (0: from the grid size) and the side stream for the many-phase kernel
    int grid_req[3] = {0, 0, 0};
    int use_side = 2;     // 0: one stream in sequence, 1: side streams, 2: one interface launch
    // captured steps, one per (state buffer, temperature buffer, step parity)
    int use_graphs = 0;
    hipGraphExec_t gexec[8] = {};
    // observables' buffer: results and instance offsets, kept between calls
    void *obs = nullptr;
    size_t obs_cap = 0;
};

static void drop_graphs(mpf_handle *h)
{
    for (int i = 0; i < 8; ++i) {
        if (h->gexec[i]) hipGraphExecDestroy(h->gexec[i]);
        h->gexec[i] = nullptr;
    }
}

static size_t inst_cells(const Inst &d) { return (size_t)d.nx * d.ny * d.nz; }

template <typename R>
static SlotView<R> view(mpf_handle *h, int i)
{
    SlotView<R> v;
    v.liq = (R *)h->b[i].liq;
    v.cnt = h->b[i].cnt;
    v.sid = h->b[i].sid;
    v.sval = (R *)h->b[i].sval;
    v.stride = h->ncell;
    return v;
}

// The buffer seen by one instance: its first cell at index 0, slot stride
// of the whole buffer.
template <typename R>
static SlotView<R> iview(mpf_handle *h, int i, int ii)
{
    SlotView<R> v = view<R>(h, i);
    const size_t o = h->inst[ii].cell_off;
    v.liq += o;
    v.cnt += o;
    v.sid += o;
    v.sval += o;
    return v;
}

static int dev_alloc(mpf_handle *h, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 1;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(MPF_ENOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    h->bytes += bytes;
    return 0;
}

extern "C" void mpf_destroy(mpf_handle *h)
{
    if (!h) return;
    if (h->stream) hipStreamSynchronize(h->stream);
    for (int i = 0; i < 2; ++i) {
        hipFree(h->b[i].liq);
        hipFree(h->b[i].cnt);
        hipFree(h->b[i].sid);
        hipFree(h->b[i].sval);
        hipFree(h->b[i].occ);
        hipFree(h->T[i]);
    }
    hipFree(h->tab);
    hipFree(h->wl);
    hipFree(h->wl_count);
    hipFree(h->tiles);
    hipFree(h->item_busy);
    hipFree(h->stats);
    hipFree(h->inst_d);
    hipFree(h->obs);
    drop_graphs(h);
    if (h->side) hipStreamSynchronize(h->side);
    if (h->side2) hipStreamSynchronize(h->side2);
    for (hipEvent_t e : h->tev) hipEventDestroy(e);
    for (hipEvent_t e : {h->ev_merged, h->ev_general, h->ev_front})
        if (e) hipEventDestroy(e);
    if (h->side) hipStreamDestroy(h->side);
    if (h->side2) hipStreamDestroy(h->side2);
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

static int reset_stats_dev(mpf_handle *h)
{
    Stats s;
    memset(&s, 0, sizeof(s));
    s.halt_step = NO_HALT;
    HIP_TRY(hipMemcpyAsync(h->stats, &s, sizeof(s), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemsetAsync(h->wl_count, 0, 2 * NWC * sizeof(int), h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return 0;
}

// Mode of the passes after the merge pass (mpf_set_launch): side streams on
// large grids (measured equal to one launch in FP32 and faster in FP64 at
// 512x512x256), one launch on small ones (the cross-stream fork and join
// cost more than the passes there: 128^3 steps 125 -> 82 us).
static int default_mode(const mpf_handle *h) { return h->n_tiles >= 4096 ? 1 : 2; }

// Planes per classification work item: as set (mpf_set_tiling), else
// shorter ranges on grids with few tiles, so that more blocks share the
// tiles with work.  2D grids have one plane.
static int light_planes(const mpf_handle *h)
{
    if (h->cfg.dim == 2) return LZC;
    if (h->light_q_req) return h->light_q_req;
    return h->n_tiles >= 4096 ? LZC : (h->n_tiles >= 1024 ? LZC / 2 : LZC / 4);
}

// Common creation of a single grid (batch = 0, one instance with the
// configuration's grid) and of a batch of 3D grids.
static int create_handle(const mpf_config &c, const std::vector<Inst> &grids, int batch, mpf_handle **out)
{
    if (c.dim != 2 && c.dim != 3) return fail(MPF_EINVAL, "dim must be 2 or 3 (got %d)", c.dim);
    if (c.n_phase < 1 || c.n_phase > 65535) return fail(MPF_EINVAL, "n_phase must be in [1, 65535]");
    if (c.kmax < 1 || c.kmax > KCAP) return fail(MPF_EINVAL, "kmax must be in [1, %d]", KCAP);
    if (c.precision != MPF_FP32 && c.precision != MPF_FP64) return fail(MPF_EINVAL, "bad precision");
    if (c.dim == 2 && c.n_phase < 2) return fail(MPF_EINVAL, "the 2D update needs n_phase >= 2");
    if (grids.empty()) return fail(MPF_EINVAL, "no grid");
    if (batch) {
        if (c.dim != 3) return fail(MPF_EINVAL, "a batch holds 3D grids");
        if ((int)grids.size() > MAX_INST) return fail(MPF_EINVAL, "a batch holds at most %d grids", MAX_INST);
        if (c.temp_dense) return fail(MPF_EINVAL, "a batch keeps the temperature per z-plane (temp_dense = 0)");
        if (c.halt_on_kmax) return fail(MPF_EINVAL, "halt_on_kmax is not available for a batch");
    }
    size_t ncell = 0, ntile = 0, nplane = 0;
    std::vector<Inst> inst = grids;
    for (size_t i = 0; i < inst.size(); ++i) {
        Inst &d = inst[i];
        if (d.nx < 1 || d.ny < 1 || d.nz < 1) return fail(MPF_EINVAL, "grid %zu: sizes must be positive", i);
        if (c.dim == 2 && d.nz != 1) return fail(MPF_EINVAL, "2D grids need nz == 1");
        d.pad = 0;
        d.cell_off = (uint32_t)ncell;
        d.tile_off = (uint32_t)ntile;
        d.occ_off = (uint32_t)ntile;
        d.t_off = (uint32_t)nplane;
        ncell += inst_cells(d);
        ntile += occ_words(d.nx, d.ny, d.nz, c.dim);
        nplane += (size_t)d.nz;
        if (ncell > (size_t)0x7fffffff) return fail(MPF_EINVAL, "more than 2^31-1 cells in total");
    }
    mpf_handle *h = new (std::nothrow) mpf_handle();
    if (!h) return fail(MPF_ENOMEM, "host allocation failed");
    h->cfg = c;
    h->cfg.nx = inst[0].nx;
    h->cfg.ny = inst[0].ny;
    h->cfg.nz = inst[0].nz;
    h->dim = c.dim;
    h->ncell = ncell;
    h->batch = batch;
    h->inst = inst;
    h->n_tiles = (int)ntile;
    h->use_side = default_mode(h);
    h->nocc = ntile;
    h->rsz = c.precision == MPF_FP64 ? 8 : 4;
    const size_t nn = (size_t)c.n_phase * c.n_phase;
    h->tab_stride = 3 * nn + (size_t)c.n_phase * 24;
    h->tab_set.assign(inst.size(), 0);
    h->tab_cls.assign(inst.size(), 0);
    h->tab_pairs.assign(inst.size(), 0);
    h->cls_vals.assign(inst.size() * 6, 0.0);
    int cap = c.kmax < c.n_phase - 1 ? c.kmax : c.n_phase - 1;
    if (c.slot_capacity > 0) {
        // a step may store up to min(KMAX, N-1) grains per cell
        if (c.slot_capacity < cap) {
            delete h;
            return fail(MPF_EINVAL, "slot_capacity must be at least min(kmax, n_phase-1) = %d", cap);
        }
        cap = c.slot_capacity;
    }
    if (cap > 255) {
        delete h;
        return fail(MPF_EINVAL, "slot_capacity must be at most 255");
    }
    h->scap = cap < 1 ? 1 : cap;
    int rc = 0;
#define TRY_RC(x)                       \
    do {                                \
        if ((rc = (x)) != 0) {          \
            mpf_destroy(h);             \
            return rc;                  \
        }                               \
    } while (0)
    hipError_t e = hipStreamCreate(&h->stream);
    if (e != hipSuccess) {
        delete h;
        return fail(MPF_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    // at least two slot arrays: the classification pass reads slot 1 of
    // every cell it visits (and uses it only where the count is 2)
    const size_t nslot = h->scap < 2 ? 2 : (size_t)h->scap;
    for (int i = 0; i < 2; ++i) {
        TRY_RC(dev_alloc(h, &h->b[i].liq, ncell * h->rsz));
        TRY_RC(dev_alloc(h, (void **)&h->b[i].cnt, ncell));
        TRY_RC(dev_alloc(h, (void **)&h->b[i].sid, ncell * nslot * sizeof(uint16_t)));
        TRY_RC(dev_alloc(h, &h->b[i].sval, ncell * nslot * h->rsz));
        TRY_RC(dev_alloc(h, (void **)&h->b[i].occ, h->nocc * sizeof(uint16_t)));
    }
    h->nT = c.temp_dense ? ncell : nplane;
    for (int i = 0; i < 2; ++i) TRY_RC(dev_alloc(h, (void **)&h->T[i], h->nT * sizeof(double)));
    TRY_RC(dev_alloc(h, &h->tab, inst.size() * h->tab_stride * h->rsz));
    TRY_RC(dev_alloc(h, (void **)&h->wl, (size_t)NWL * ncell * sizeof(int)));
    TRY_RC(dev_alloc(h, (void **)&h->wl_count, 2 * NWC * sizeof(int)));
    TRY_RC(dev_alloc(h, (void **)&h->tiles, (size_t)h->n_tiles * (LZC / 4) * sizeof(int)));
    TRY_RC(dev_alloc(h, (void **)&h->item_busy, (size_t)h->n_tiles * (LZC / 4) * sizeof(uint16_t)));
    TRY_RC(dev_alloc(h, (void **)&h->stats, sizeof(Stats)));
    {
        TRY_RC(dev_alloc(h, (void **)&h->inst_d, inst.size() * sizeof(Inst)));
        e = hipMemcpy(h->inst_d, inst.data(), inst.size() * sizeof(Inst), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            mpf_destroy(h);
            return fail(MPF_EHIP, "hipMemcpy: %s", hipGetErrorString(e));
        }
    }
    // empty state: all liquid, no grains
    for (int i = 0; i < 2; ++i) {
        e = hipMemsetAsync(h->b[i].cnt, 0, ncell, h->stream);
        if (e == hipSuccess) e = hipMemsetAsync(h->b[i].liq, 0, ncell * h->rsz, h->stream);
        if (e == hipSuccess) e = hipMemsetAsync(h->b[i].occ, 0xff, h->nocc * sizeof(uint16_t), h->stream);
        if (e == hipSuccess) e = hipMemsetAsync(h->T[i], 0, h->nT * sizeof(double), h->stream);
        if (e != hipSuccess) {
            mpf_destroy(h);
            return fail(MPF_EHIP, "hipMemsetAsync: %s", hipGetErrorString(e));
        }
    }
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
        hipStreamCreate(&h->side) != hipSuccess || hipStreamCreate(&h->side2) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_merged, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_general, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_front, hipEventDisableTiming) != hipSuccess) {
        mpf_destroy(h);
        return fail(MPF_EHIP, "hipEventCreate failed");
    }
    TRY_RC(reset_stats_dev(h));
#undef TRY_RC
    *out = h;
    return 0;
}

extern "C" int mpf_create(const mpf_config *cfg, mpf_handle **out)
{
    if (!cfg || !out) return fail(MPF_EINVAL, "mpf_create: null argument");
    *out = nullptr;
    Inst d{};
    d.nx = cfg->nx;
    d.ny = cfg->ny;
    d.nz = cfg->nz;
    return create_handle(*cfg, std::vector<Inst>(1, d), 0, out);
}

extern "C" int mpf_batch_create(const mpf_config *cfg, int n_inst, const int32_t *dims, mpf_handle **out)
{
    if (!cfg || !dims || !out || n_inst < 1) return fail(MPF_EINVAL, "mpf_batch_create: bad argument");
    *out = nullptr;
    std::vector<Inst> g((size_t)n_inst);
    for (int i = 0; i < n_inst; ++i) {
        g[i] = Inst{};
        g[i].nx = dims[3 * i];
        g[i].ny = dims[3 * i + 1];
        g[i].nz = dims[3 * i + 2];
    }
    return create_handle(*cfg, g, 1, out);
}

extern "C" int mpf_batch_size(mpf_handle *h) { return h ? (int)h->inst.size() : 0; }

extern "C" size_t mpf_device_bytes(mpf_handle *h) { return h ? h->bytes : 0; }

static int check_inst(mpf_handle *h, int ii, const char *fn)
{
    if (!h) return fail(MPF_EINVAL, "%s: null handle", fn);
    if (ii < 0 || ii >= (int)h->inst.size()) return fail(MPF_EINVAL, "%s: no instance %d", fn, ii);
    return 0;
}

// Tables of one instance, converted to the solver precision; the class form
// is checked per instance and used only when every instance has it with the
// same values.
static int set_tables_inst(mpf_handle *h, int ii, const double *wij, const double *aij, const double *mij,
                           const double *n111)
{
    const int N = h->cfg.n_phase;
    const size_t nn = (size_t)N * N, tot = h->tab_stride;
    std::vector<double> all(tot);
    memcpy(all.data(), wij, nn * sizeof(double));
    memcpy(all.data() + nn, aij, nn * sizeof(double));
    memcpy(all.data() + 2 * nn, mij, nn * sizeof(double));
    memcpy(all.data() + 3 * nn, n111, (size_t)N * 24 * sizeof(double));
    char *dst = (char *)h->tab + (size_t)ii * tot * h->rsz;
    int cls_ok = 0, pairs = 0;
    double cv[6];
    if (h->rsz == 8) {
        HIP_TRY(hipMemcpy(dst, all.data(), tot * sizeof(double), hipMemcpyHostToDevice));
        double cls[3][2];
        cls_ok = class_tables<double>(all.data(), N, cls);
        pairs = n111_pairs<double>(all.data() + 3 * nn, N);
        for (int t = 0; t < 6; ++t) cv[t] = cls[t / 2][t % 2];
    } else {
        std::vector<float> f(tot);
        for (size_t i = 0; i < tot; ++i) f[i] = (float)all[i];
        HIP_TRY(hipMemcpy(dst, f.data(), tot * sizeof(float), hipMemcpyHostToDevice));
        float cls[3][2];
        cls_ok = class_tables<float>(f.data(), N, cls);
        pairs = n111_pairs<float>(f.data() + 3 * nn, N);
        for (int t = 0; t < 6; ++t) cv[t] = cls[t / 2][t % 2];
    }
    drop_graphs(h); // the class form and values are part of a captured step
    h->tab_set[ii] = 1;
    h->tab_cls[ii] = cls_ok;
    h->tab_pairs[ii] = pairs;
    for (int t = 0; t < 6; ++t) h->cls_vals[(size_t)ii * 6 + t] = cv[t];
    int all_set = 1, all_cls = 1, all_pairs = 1;
    for (size_t i = 0; i < h->inst.size(); ++i) {
        all_set &= h->tab_set[i];
        all_pairs &= h->tab_set[i] && h->tab_pairs[i];
        all_cls &= h->tab_set[i] && h->tab_cls[i] &&
                   memcmp(&h->cls_vals[i * 6], &h->cls_vals[0], 6 * sizeof(double)) == 0;
    }
    h->tables_set = all_set;
    h->tab_class = all_cls;
    h->n111_pairs = all_pairs;
    for (int t = 0; t < 6; ++t) h->cls[t / 2][t % 2] = h->cls_vals[t];
    return 0;
}

extern "C" int mpf_set_tables(mpf_handle *h, const double *wij, const double *aij, const double *mij,
                              const double *n111)
{
    if (!h || !wij || !aij || !mij || !n111) return fail(MPF_EINVAL, "mpf_set_tables: null argument");
    for (int ii = 0; ii < (int)h->inst.size(); ++ii) {
        int rc = set_tables_inst(h, ii, wij, aij, mij, n111);
        if (rc) return rc;
    }
    return 0;
}

extern "C" int mpf_batch_set_tables(mpf_handle *h, int ii, const double *wij, const double *aij,
                                    const double *mij, const double *n111)
{
    if (!wij || !aij || !mij || !n111) return fail(MPF_EINVAL, "mpf_batch_set_tables: null argument");
    if (int rc = check_inst(h, ii, "mpf_batch_set_tables")) return rc;
    return set_tables_inst(h, ii, wij, aij, mij, n111);
}

// phi: host memory, or device memory with on_device (no staging copy).
static int upload_phi_inst(mpf_handle *h, int ii, const void *phi, int dtype, bool on_device = false)
{
    if (dtype != MPF_FP32 && dtype != MPF_FP64) return fail(MPF_EINVAL, "bad dtype");
    const Inst &g = h->inst[ii];
    size_t esz = dtype == MPF_FP64 ? 8 : 4;
    size_t bytes = inst_cells(g) * (size_t)h->cfg.n_phase * esz;
    void *d = nullptr;
    if (!on_device) HIP_TRY(hipMalloc(&d, bytes));
    int *err = nullptr;
    hipError_t e = hipMalloc((void **)&err, sizeof(int));
    if (on_device)
        d = const_cast<void *>(phi);
    else if (e == hipSuccess)
        e = hipMemcpy(d, phi, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(err, 0, sizeof(int));
    // the new state may hold grains anywhere
    if (e == hipSuccess) e = hipMemsetAsync(h->b[h->cur].occ, 0xff, h->nocc * sizeof(uint16_t), h->stream);
    if (e == hipSuccess) {
        const mpf_config &c = h->cfg;
        if (h->rsz == 8)
            e = mpf_launch_pack_f64(d, dtype == MPF_FP64, c.n_phase, g.nx, g.ny, g.nz, iview<double>(h, h->cur, ii),
                                    h->scap, err, h->stream);
        else
            e = mpf_launch_pack_f32(d, dtype == MPF_FP64, c.n_phase, g.nx, g.ny, g.nz, iview<float>(h, h->cur, ii),
                                    h->scap, err, h->stream);
    }
    int herr = 0;
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost);
    if (!on_device) hipFree(d);
    hipFree(err);
    if (e != hipSuccess) return fail(MPF_EHIP, "upload_phi: %s", hipGetErrorString(e));
    if (herr) return fail(MPF_ECAPACITY, "a cell holds more than %d non-zero solid phases", h->scap);
    return 0;
}

extern "C" int mpf_upload_phi(mpf_handle *h, const void *phi, int dtype)
{
    if (!h || !phi) return fail(MPF_EINVAL, "mpf_upload_phi: null argument");
    if (h->batch) return fail(MPF_EINVAL, "mpf_upload_phi: use mpf_batch_upload_phi for a batch");
    return upload_phi_inst(h, 0, phi, dtype);
}

extern "C" int mpf_batch_upload_phi(mpf_handle *h, int ii, const void *phi, int dtype)
{
    if (!phi) return fail(MPF_EINVAL, "mpf_batch_upload_phi: null argument");
    if (int rc = check_inst(h, ii, "mpf_batch_upload_phi")) return rc;
    return upload_phi_inst(h, ii, phi, dtype);
}

extern "C" int mpf_upload_phi_device(mpf_handle *h, const void *phi, int dtype)
{
    if (!h || !phi) return fail(MPF_EINVAL, "mpf_upload_phi_device: null argument");
    if (h->batch) return fail(MPF_EINVAL, "mpf_upload_phi_device: single grids only");
    return upload_phi_inst(h, 0, phi, dtype, true);
}

static int init_columns_inst(mpf_handle *h, int ii, const int32_t *grain_map, const int32_t *profile, int nprof,
                             const double *solid, const double *liquid)
{
    if (!grain_map || !profile || !solid || !liquid || nprof < 1)
        return fail(MPF_EINVAL, "init_columns: bad argument");
    const mpf_config &c = h->cfg;
    const Inst &gd = h->inst[ii];
    size_t ncol = (size_t)gd.nx * gd.ny;
    for (size_t i = 0; i < ncol; ++i) {
        if (profile[i] < 0 || profile[i] >= nprof) return fail(MPF_EINVAL, "profile index out of range");
        if (grain_map[i] < 0 || grain_map[i] >= c.n_phase) return fail(MPF_EINVAL, "grain id out of range");
    }
    void *d = nullptr;
    size_t bytes = ncol * 2 * sizeof(int32_t) + (size_t)nprof * gd.nz * 2 * sizeof(double);
    HIP_TRY(hipMalloc(&d, bytes));
    int32_t *dg = (int32_t *)d, *dp = dg + ncol;
    double *ds = (double *)(dp + ncol), *dl = ds + (size_t)nprof * gd.nz;
    hipError_t e = hipMemcpy(dg, grain_map, ncol * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dp, profile, ncol * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ds, solid, (size_t)nprof * gd.nz * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dl, liquid, (size_t)nprof * gd.nz * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(h->b[h->cur].occ, 0xff, h->nocc * sizeof(uint16_t), h->stream);
    if (e == hipSuccess) {
        if (h->rsz == 8)
            e = mpf_launch_init_columns_f64(dg, dp, ds, dl, gd.nx, gd.ny, gd.nz, iview<double>(h, h->cur, ii),
                                            h->stream);
        else
            e = mpf_launch_init_columns_f32(dg, dp, ds, dl, gd.nx, gd.ny, gd.nz, iview<float>(h, h->cur, ii),
                                            h->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hipFree(d);
    if (e != hipSuccess) return fail(MPF_EHIP, "init_columns: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int mpf_init_columns(mpf_handle *h, const int32_t *grain_map, const int32_t *profile, int nprof,
                                const double *solid, const double *liquid)
{
    if (!h) return fail(MPF_EINVAL, "mpf_init_columns: null handle");
    if (h->batch) return fail(MPF_EINVAL, "mpf_init_columns: use mpf_batch_init_columns for a batch");
    return init_columns_inst(h, 0, grain_map, profile, nprof, solid, liquid);
}

extern "C" int mpf_batch_init_columns(mpf_handle *h, int ii, const int32_t *grain_map, const int32_t *profile,
                                      int nprof, const double *solid, const double *liquid)
{
    if (int rc = check_inst(h, ii, "mpf_batch_init_columns")) return rc;
    return init_columns_inst(h, ii, grain_map, profile, nprof, solid, liquid);
}

static int upload_temp_inst(mpf_handle *h, int ii, const double *temp)
{
    const mpf_config &c = h->cfg;
    const Inst &g = h->inst[ii];
    if (!c.temp_dense) {
        HIP_TRY(hipMemcpy(h->T[h->tcur] + g.t_off, temp, (size_t)g.nz * sizeof(double), hipMemcpyHostToDevice));
        return 0;
    }
    double *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, h->ncell * sizeof(double)));
    hipError_t e = hipMemcpy(d, temp, h->ncell * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = mpf_launch_temp_layout_f64(d, h->T[h->tcur], c.nx, c.ny, c.nz, 1, 1, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hipFree(d);
    if (e != hipSuccess) return fail(MPF_EHIP, "upload_temp: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int mpf_upload_temp(mpf_handle *h, const double *temp)
{
    if (!h || !temp) return fail(MPF_EINVAL, "mpf_upload_temp: null argument");
    if (h->batch) return fail(MPF_EINVAL, "mpf_upload_temp: use mpf_batch_upload_temp for a batch");
    return upload_temp_inst(h, 0, temp);
}

extern "C" int mpf_batch_upload_temp(mpf_handle *h, int ii, const double *temp)
{
    if (!temp) return fail(MPF_EINVAL, "mpf_batch_upload_temp: null argument");
    if (int rc = check_inst(h, ii, "mpf_batch_upload_temp")) return rc;
    return upload_temp_inst(h, ii, temp);
}

static int download_temp_inst(mpf_handle *h, int ii, double *temp)
{
    const mpf_config &c = h->cfg;
    const Inst &g = h->inst[ii];
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (!c.temp_dense) {
        HIP_TRY(hipMemcpy(temp, h->T[h->tcur] + g.t_off, (size_t)g.nz * sizeof(double), hipMemcpyDeviceToHost));
        return 0;
    }
    double *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, h->ncell * sizeof(double)));
    hipError_t e = mpf_launch_temp_layout_f64(h->T[h->tcur], d, c.nx, c.ny, c.nz, 0, 1, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(temp, d, h->ncell * sizeof(double), hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(MPF_EHIP, "download_temp: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int mpf_download_temp(mpf_handle *h, double *temp)
{
    if (!h || !temp) return fail(MPF_EINVAL, "mpf_download_temp: null argument");
    if (h->batch) return fail(MPF_EINVAL, "mpf_download_temp: use mpf_batch_download_temp for a batch");
    return download_temp_inst(h, 0, temp);
}

extern "C" int mpf_batch_download_temp(mpf_handle *h, int ii, double *temp)
{
    if (!temp) return fail(MPF_EINVAL, "mpf_batch_download_temp: null argument");
    if (int rc = check_inst(h, ii, "mpf_batch_download_temp")) return rc;
    return download_temp_inst(h, ii, temp);
}

template <typename R>
static hipError_t launch_one_step(mpf_handle *h, const int *grids, hipEvent_t *tev)
{
    const mpf_config &c = h->cfg;
    StepArgs<R> a;
    memset(&a, 0, sizeof(a));
    a.K = make_consts<R>(c.s, c.nx, c.ny, c.nz, c.n_phase, c.kmax, c.enable_anisotropy,
                         c.dim == 2 ? 1 : c.enable_torque, h->tab);
    if (c.dim == 2) a.K.aniso = 1;
    a.in = view<R>(h, h->cur);
    a.out = view<R>(h, h->cur ^ 1);
    a.occ_in = h->b[h->cur].occ;
    a.occ_out = h->b[h->cur ^ 1].occ;
    a.T_in = h->T[h->tcur];
    a.T_out = h->T[h->tcur ^ 1];
    a.temp_dense = c.temp_dense;
    a.temp_f32 = c.temp_f32;
    a.cooling = c.cooling_rate;
    a.wl_base = h->wl;
    a.wl_count = h->wl_count;
    a.tiles = h->tiles;
    a.item_busy = h->item_busy;
    a.n_tiles = h->n_tiles;
    a.light_q = light_planes(h);
    a.light_grid = h->light_grid_req;
    a.n_temp = h->nT;
    a.step = h->step_counter;
    a.halt_mode = c.halt_on_kmax;
    a.tab_class = h->tab_class;
    a.K.n111_pairs = h->n111_pairs;
    for (int t = 0; t < 3; ++t) {
        a.K.cls[t][0] = (R)h->cls[t][0];
        a.K.cls[t][1] = (R)h->cls[t][1];
    }
    a.stats = h->stats;
    if (h->batch) {
        a.inst = h->inst_d;
        a.n_inst = (int)h->inst.size();
        a.tab_stride = h->tab_stride;
    }
    StepStreams ss{h->use_side == 2 ? 1 : 0, h->use_side == 1 ? h->side : nullptr, h->side2, h->ev_merged,
                   h->ev_general, h->ev_front};
    if (sizeof(R) == 8)
        return mpf_launch_step_f64((const StepArgs<double> *)&a, c.dim, grids, h->stream, &ss, tev);
    return mpf_launch_step_f32((const StepArgs<float> *)&a, c.dim, grids, h->stream, &ss, tev);
}

// Launch shape of the work-list kernels: the requested block counts, or by
// default enough waves to fill the GPU and fewer for small grids.
static void step_grids(mpf_handle *h, int grids[3])
{
    auto blocks = [&](size_t threads_per_block, int cap) {
        size_t b = (h->ncell + threads_per_block - 1) / threads_per_block;
        return (int)(b < (size_t)cap ? (b ? b : 1) : cap);
    };
    grids[0] = h->grid_req[0] > 0 ? h->grid_req[0] : blocks(256, 8192);
    grids[1] = h->grid_req[1] > 0 ? h->grid_req[1] : blocks(128, 2048);
    grids[2] = h->grid_req[2] > 0 ? h->grid_req[2] : blocks(64, 256);
}

static hipError_t launch_step(mpf_handle *h, const int *grids, hipEvent_t *tev)
{
    return (h->rsz == 8) ? launch_one_step<double>(h, grids, tev) : launch_one_step<float>(h, grids, tev);
}

// One step replayed from a graph captured on first use for this state
// parity.  The captured step covers the side-stream fork and join.
static hipError_t graph_step(mpf_handle *h, const int *grids)
{
    const int key = (h->cur & 1) | ((h->tcur & 1) << 1) | ((h->step_counter & 1) << 2);
    if (!h->gexec[key]) {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        hipError_t el = launch_step(h, grids, nullptr);
        e = hipStreamEndCapture(h->stream, &g);
        if (el != hipSuccess) e = el;
        if (e == hipSuccess) e = hipGraphInstantiate(&h->gexec[key], g, nullptr, nullptr, 0);
        if (g) hipGraphDestroy(g);
        if (e != hipSuccess) {
            h->gexec[key] = nullptr;
            return e;
        }
    }
    return hipGraphLaunch(h->gexec[key], h->stream);
}

extern "C" int mpf_step(mpf_handle *h, int nsteps)
{
    if (!h) return fail(MPF_EINVAL, "mpf_step: null handle");
    if (nsteps < 0) return fail(MPF_EINVAL, "mpf_step: negative step count");
    if (!h->tables_set) return fail(MPF_EINVAL, "mpf_step: tables not set (mpf_set_tables)");
    if (nsteps == 0) return 0;
    const mpf_config &c = h->cfg;
    // a batch starts with clear worklist counters and no halt
    HIP_TRY(hipStreamSynchronize(h->stream));
    int m1 = NO_HALT;
    HIP_TRY(hipMemcpy(&h->stats->halt_step, &m1, sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(h->wl_count, 0, 2 * NWC * sizeof(int)));
    const int s0 = h->step_counter, cur0 = h->cur, tcur0 = h->tcur;
    int grids[3];
    step_grids(h, grids);
    const bool graphs = h->use_graphs && !h->kernel_timing && !c.halt_on_kmax;
    if (h->kernel_timing) {
        const size_t need = (size_t)nsteps * 8;
        while (h->tev.size() < need) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            h->tev.push_back(e);
        }
    }
    HIP_TRY(hipEventRecord(h->ev0, h->stream));
    for (int i = 0; i < nsteps; ++i) {
        hipEvent_t *tev = h->kernel_timing ? h->tev.data() + (size_t)i * 8 : nullptr;
        hipError_t e = graphs ? graph_step(h, grids) : launch_step(h, grids, tev);
        if (e != hipSuccess) return fail(MPF_EHIP, "step launch: %s", hipGetErrorString(e));
        h->cur ^= 1;
        h->tcur ^= 1;
        h->step_counter += 1;
    }
    HIP_TRY(hipEventRecord(h->ev1, h->stream));
    HIP_TRY(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    if (h->kernel_timing) {
        HIP_TRY(hipStreamSynchronize(h->side));
        HIP_TRY(hipStreamSynchronize(h->side2));
        if (h->use_side != 1) HIP_TRY(hipStreamSynchronize(h->stream));
        for (int i = 0; i < nsteps; ++i) {
            // classification (with the tile pass), merge, one-solid
            // interface, many-phase (events of launch_step_kernels)
            hipEvent_t *t = h->tev.data() + (size_t)i * 8;
            float k[4];
            HIP_TRY(hipEventElapsedTime(&k[0], t[0], t[1]));
            HIP_TRY(hipEventElapsedTime(&k[1], t[1], t[2]));
            HIP_TRY(hipEventElapsedTime(&k[2], t[3], t[4]));
            HIP_TRY(hipEventElapsedTime(&k[3], t[5], t[6]));
            for (int j = 0; j < 4; ++j) h->kern_ms[j] += k[j];
        }
        h->kern_steps += nsteps;
    }
    int done = nsteps;
    if (c.halt_on_kmax) {
        int hs = -1;
        HIP_TRY(hipMemcpy(&hs, &h->stats->halt_step, sizeof(int), hipMemcpyDeviceToHost));
        if (hs != NO_HALT) {
            done = hs - s0;
            h->cur = cur0 ^ (done & 1);
            h->tcur = tcur0 ^ ((done + 1) & 1);
            h->step_counter = hs + 1;
        }
    }
    h->steps_done += done;
    return done;
}

// phi: host memory, or device memory with on_device (unpacked in place).
static int download_phi_inst(mpf_handle *h, int ii, void *phi, int dtype, bool on_device = false)
{
    if (dtype != MPF_FP32 && dtype != MPF_FP64) return fail(MPF_EINVAL, "bad dtype");
    const mpf_config &c = h->cfg;
    const Inst &g = h->inst[ii];
    size_t esz = dtype == MPF_FP64 ? 8 : 4;
    size_t bytes = inst_cells(g) * (size_t)c.n_phase * esz;
    void *d = phi;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (!on_device) HIP_TRY(hipMalloc(&d, bytes));
    hipError_t e = hipMemsetAsync(d, 0, bytes, h->stream);
    if (e == hipSuccess) {
        if (h->rsz == 8)
            e = mpf_launch_unpack_f64(iview<double>(h, h->cur, ii), g.nx, g.ny, g.nz, d, dtype == MPF_FP64, h->stream);
        else
            e = mpf_launch_unpack_f32(iview<float>(h, h->cur, ii), g.nx, g.ny, g.nz, d, dtype == MPF_FP64, h->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && !on_device) e = hipMemcpy(phi, d, bytes, hipMemcpyDeviceToHost);
    if (!on_device) hipFree(d);
    if (e != hipSuccess) return fail(MPF_EHIP, "download_phi: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int mpf_download_phi(mpf_handle *h, void *phi, int dtype)
{
    if (!h || !phi) return fail(MPF_EINVAL, "mpf_download_phi: null argument");
    if (h->batch) return fail(MPF_EINVAL, "mpf_download_phi: use mpf_batch_download_phi for a batch");
    return download_phi_inst(h, 0, phi, dtype);
}

extern "C" int mpf_batch_download_phi(mpf_handle *h, int ii, void *phi, int dtype)
{
    if (!phi) return fail(MPF_EINVAL, "mpf_batch_download_phi: null argument");
    if (int rc = check_inst(h, ii, "mpf_batch_download_phi")) return rc;
    return download_phi_inst(h, ii, phi, dtype);
}

extern "C" int mpf_download_phi_device(mpf_handle *h, void *phi, int dtype)
{
    if (!h || !phi) return fail(MPF_EINVAL, "mpf_download_phi_device: null argument");
    if (h->batch) return fail(MPF_EINVAL, "mpf_download_phi_device: use mpf_batch_download_phi_device");
    return download_phi_inst(h, 0, phi, dtype, true);
}

extern "C" int mpf_batch_download_phi_device(mpf_handle *h, int ii, void *phi, int dtype)
{
    if (!phi) return fail(MPF_EINVAL, "mpf_batch_download_phi_device: null argument");
    if (int rc = check_inst(h, ii, "mpf_batch_download_phi_device")) return rc;
    return download_phi_inst(h, ii, phi, dtype, true);
}

static int download_apt_inst(mpf_handle *h, int ii, int32_t *mf, int depth, int32_t *nf, int32_t *status)
{
    if (!mf || !nf || !status || depth < 1) return fail(MPF_EINVAL, "download_apt: bad argument");
    const mpf_config &c = h->cfg;
    const Inst &g = h->inst[ii];
    const size_t nc = inst_cells(g);
    size_t bm = (size_t)depth * nc * sizeof(int32_t), bn = nc * sizeof(int32_t);
    void *d = nullptr;
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMalloc(&d, bm + bn + 2 * sizeof(int32_t)));
    int32_t *dmf = (int32_t *)d, *dnf = dmf + (size_t)depth * nc, *dst = dnf + nc;
    hipError_t e = hipMemsetAsync(d, 0, bm + bn + 2 * sizeof(int32_t), h->stream);
    if (e == hipSuccess) {
        if (h->rsz == 8) {
            Consts<double> K = make_consts<double>(c.s, g.nx, g.ny, g.nz, c.n_phase, c.kmax, 1, 1, h->tab);
            e = mpf_launch_apt_slots_f64(&K, iview<double>(h, h->cur, ii), c.dim, dmf, dnf, dst, depth, h->stream);
        } else {
            Consts<float> K = make_consts<float>(c.s, g.nx, g.ny, g.nz, c.n_phase, c.kmax, 1, 1, h->tab);
            e = mpf_launch_apt_slots_f32(&K, iview<float>(h, h->cur, ii), c.dim, dmf, dnf, dst, depth, h->stream);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(mf, dmf, bm, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(nf, dnf, bn, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(status, dst, 2 * sizeof(int32_t), hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(MPF_EHIP, "download_apt: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int mpf_download_apt(mpf_handle *h, int32_t *mf, int depth, int32_t *nf, int32_t *status)
{
    if (!h) return fail(MPF_EINVAL, "mpf_download_apt: null handle");
    if (h->batch) return fail(MPF_EINVAL, "mpf_download_apt: use mpf_batch_download_apt for a batch");
    return download_apt_inst(h, 0, mf, depth, nf, status);
}

extern "C" int mpf_batch_download_apt(mpf_handle *h, int ii, int32_t *mf, int depth, int32_t *nf, int32_t *status)
{
    if (int rc = check_inst(h, ii, "mpf_batch_download_apt")) return rc;
    return download_apt_inst(h, ii, mf, depth, nf, status);
}

extern "C" int mpf_get_stats(mpf_handle *h, mpf_stats *st)
{
    if (!h || !st) return fail(MPF_EINVAL, "mpf_get_stats: null argument");
    Stats s;
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(&s, h->stats, sizeof(s), hipMemcpyDeviceToHost));
    st->max_active = s.max_active;
    st->kmax_overflow = s.kmax_overflow;
    st->overflow_cells = s.overflow_cells;
    st->general_cells = s.slow_cells;
    st->halt_step = s.halt_step == NO_HALT ? -1 : s.halt_step;
    st->steps_done = h->steps_done;
    st->last_step_ms = h->last_ms;
    st->merge_cells = s.merge_cells;
    st->front_cells = s.front_cells;
    st->light_cells = s.light_cells;
    st->n111_pairs = h->n111_pairs;
    st->class_tables = h->tab_class;
    for (int j = 0; j < 4; ++j) st->kernel_ms[j] = h->kern_ms[j];
    st->kernel_steps = h->kern_steps;
    return 0;
}

extern "C" int mpf_set_kernel_timing(mpf_handle *h, int on)
{
    if (!h) return fail(MPF_EINVAL, "mpf_set_kernel_timing: null handle");
    h->kernel_timing = on ? 1 : 0;
    return 0;
}

extern "C" int mpf_set_launch(mpf_handle *h, int merge_blocks, int front_blocks, int general_blocks,
                              int side_stream)
{
    if (!h) return fail(MPF_EINVAL, "mpf_set_launch: null handle");
    if (merge_blocks < 0 || front_blocks < 0 || general_blocks < 0)
        return fail(MPF_EINVAL, "mpf_set_launch: negative block count");
    h->grid_req[0] = merge_blocks;
    h->grid_req[1] = front_blocks;
    h->grid_req[2] = general_blocks;
    h->use_side = side_stream < 0 ? 0 : (side_stream > 2 ? 2 : side_stream);
    drop_graphs(h);
    return 0;
}

extern "C" int mpf_set_tiling(mpf_handle *h, int planes_per_item, int light_blocks)
{
    if (!h) return fail(MPF_EINVAL, "mpf_set_tiling: null handle");
    if (planes_per_item != 0 && planes_per_item != 4 && planes_per_item != 8 && planes_per_item != LZC)
        return fail(MPF_EINVAL, "mpf_set_tiling: planes_per_item must be 0, 4, 8 or %d (got %d)", LZC,
                    planes_per_item);
    if (light_blocks < 0) return fail(MPF_EINVAL, "mpf_set_tiling: negative block count");
    h->light_q_req = planes_per_item;
    h->light_grid_req = light_blocks;
    drop_graphs(h);
    return 0;
}

extern "C" int mpf_get_tiling(mpf_handle *h, int *out)
{
    if (!h || !out) return fail(MPF_EINVAL, "mpf_get_tiling: null argument");
    out[0] = light_planes(h);
    out[1] = h->light_grid_req > 0 ? h->light_grid_req : 1024;
    return 0;
}

extern "C" int mpf_get_launch(mpf_handle *h, int *out)
{
    if (!h || !out) return fail(MPF_EINVAL, "mpf_get_launch: null argument");
    int g[3];
    step_grids(h, g);
    out[0] = g[0];
    out[1] = g[1];
    out[2] = g[2];
    out[3] = h->use_side;
    return 0;
}

extern "C" int mpf_set_graphs(mpf_handle *h, int on)
{
    if (!h) return fail(MPF_EINVAL, "mpf_set_graphs: null handle");
    if (on && h->cfg.halt_on_kmax) return fail(MPF_EINVAL, "mpf_set_graphs: not with halt_on_kmax");
    h->use_graphs = on ? 1 : 0;
    if (!on) drop_graphs(h);
    return 0;
}

// Copies of the evolving state (both slot buffers hold live data only in
// `cur`, but the step reads the occupancy of `cur` and writes the other).
struct Snapshot {
    void *mem = nullptr;
};

static int snapshot(mpf_handle *h, Snapshot &sn, bool save)
{
    const size_t bl = h->ncell * h->rsz, bc = h->ncell, bs = h->ncell * h->scap * 2,
                 bv = h->ncell * h->scap * h->rsz, bo = h->nocc * 2, bt = h->nT * 8;
    const size_t tot = bl + bc + bs + bv + bo + bt;
    if (save) HIP_TRY(hipMalloc(&sn.mem, tot));
    char *p = (char *)sn.mem;
    const Buffers &b = h->b[h->cur];
    void *src[6] = {b.liq, b.cnt, b.sid, b.sval, b.occ, h->T[h->tcur]};
    const size_t sz[6] = {bl, bc, bs, bv, bo, bt};
    for (int i = 0; i < 6; ++i) {
        if (save)
            HIP_TRY(hipMemcpyAsync(p, src[i], sz[i], hipMemcpyDeviceToDevice, h->stream));
        else
            HIP_TRY(hipMemcpyAsync(src[i], p, sz[i], hipMemcpyDeviceToDevice, h->stream));
        p += sz[i];
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    return 0;
}

// Try launch shapes on the current state (trial_steps steps each, after one
// untimed step; about 15 shapes by coordinate descent) and keep the fastest;
// the state, step counter and statistics are restored afterwards.  out (optional): the chosen shape as
// mpf_get_launch reports it, and out[4] = its time per step in microseconds.
extern "C" int mpf_autotune(mpf_handle *h, int trial_steps, int *out)
{
    if (!h) return fail(MPF_EINVAL, "mpf_autotune: null handle");
    if (!h->tables_set) return fail(MPF_EINVAL, "mpf_autotune: tables not set");
    if (trial_steps < 1) trial_steps = 4;
    Snapshot sn;
    Stats st0;
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(&st0, h->stats, sizeof(st0), hipMemcpyDeviceToHost));
    if (int rc = snapshot(h, sn, true)) return rc;
    const int s0 = h->step_counter, cur0 = h->cur, tcur0 = h->tcur, done0 = h->steps_done;
    const int timing0 = h->kernel_timing, graphs0 = h->use_graphs;
    h->kernel_timing = 0;
    h->use_graphs = 0;
    // Coordinate descent from the default shape: the mode of the passes
    // after the merge pass, the planes per classification work item (3D) and
    // the classification grid, then the block counts of the merge, interface
    // and many-phase kernels, each over its candidates with the others held
    // at their best so far.  Shape: merge, interface and many-phase blocks,
    // mode, planes per item, classification blocks.
    constexpr int NS = 6;
    int cur[NS];
    {
        int g[3];
        const int r0[3] = {h->grid_req[0], h->grid_req[1], h->grid_req[2]};
        const int q0 = h->light_q_req;
        h->grid_req[0] = h->grid_req[1] = h->grid_req[2] = 0;
        h->light_q_req = 0;
        step_grids(h, g);
        cur[4] = light_planes(h);
        h->grid_req[0] = r0[0], h->grid_req[1] = r0[1], h->grid_req[2] = r0[2];
        h->light_q_req = q0;
        cur[0] = g[0], cur[1] = g[1], cur[2] = g[2], cur[3] = default_mode(h), cur[5] = 1024;
    }
    const int q_req0 = h->light_q_req, lg_req0 = h->light_grid_req;
    static const int cand_m[] = {512, 1024, 2048, 4096, 8192}, cand_g[] = {64, 128, 256, 512};
    static const int cand_q[] = {4, 8, LZC}, cand_l[] = {512, 1024, 2048};
    const int *cands[NS] = {cand_m, cand_m, cand_g, nullptr, cand_q, cand_l};
    const int ncand[NS] = {5, 5, 4, 3, h->cfg.dim == 3 ? 3 : 0, 3};
    double best = 1e300;
    int pick[NS];
    for (int j = 0; j < NS; ++j) pick[j] = cur[j];
    int rc = 0;
    // the three fastest shapes seen, for the confirmation round
    constexpr int NTOP = 3;
    int top[NTOP][NS];
    double topt[NTOP] = {1e300, 1e300, 1e300};
    auto time_shape = [&](const int *shape, int steps, double &t) -> int {
        h->grid_req[0] = shape[0];
        h->grid_req[1] = shape[1];
        h->grid_req[2] = shape[2];
        h->use_side = shape[3];
        h->light_q_req = shape[4];
        h->light_grid_req = shape[5];
        h->cur = cur0;
        h->tcur = tcur0;
        h->step_counter = s0;
        if (int r = snapshot(h, sn, false)) return r;
        if (int r = mpf_step(h, 1); r < 0) return r;
        if (int r = mpf_step(h, steps); r < 0) return r;
        t = h->last_ms / steps;
        return 0;
    };
    auto trial = [&](const int *shape) -> int {
        double t;
        if (int r = time_shape(shape, trial_steps, t)) return r;
        if (t < best) {
            best = t;
            for (int j = 0; j < NS; ++j) pick[j] = shape[j];
        }
        int q = NTOP;
        while (q > 0 && t < topt[q - 1]) --q;
        if (q < NTOP) {
            for (int k = NTOP - 1; k > q; --k) {
                topt[k] = topt[k - 1];
                for (int j = 0; j < NS; ++j) top[k][j] = top[k - 1][j];
            }
            topt[q] = t;
            for (int j = 0; j < NS; ++j) top[q][j] = shape[j];
        }
        return 0;
    };
    rc = trial(cur);
    const int order[NS] = {3, 4, 5, 0, 1, 2};
    for (int oi = 0; oi < NS && !rc; ++oi) {
        const int dmn = order[oi];
        for (int v = 0; v < ncand[dmn] && !rc; ++v) {
            int shape[NS];
            for (int j = 0; j < NS; ++j) shape[j] = pick[j];
            const int val = dmn == 3 ? 2 - v : cands[dmn][v];
            if (val == pick[dmn]) continue;
            shape[dmn] = val;
            rc = trial(shape);
        }
    }
    // Confirmation: single short trials are noisy at the 1-2% level that
    // separates the leading shapes, so the three fastest are timed again
    // over three times as many steps and the fastest of those is kept.
    if (!rc && topt[0] < 1e300) {
        double bestc = 1e300;
        for (int q = 0; q < NTOP && !rc; ++q) {
            if (!(topt[q] < 1e300)) continue;
            double t;
            rc = time_shape(top[q], 3 * trial_steps, t);
            if (!rc && t < bestc) {
                bestc = t;
                for (int j = 0; j < NS; ++j) pick[j] = top[q][j];
            }
        }
        if (!rc) best = bestc;
    }
    h->cur = cur0;
    h->tcur = tcur0;
    h->step_counter = s0;
    h->steps_done = done0;
    int rc2 = snapshot(h, sn, false);
    hipFree(sn.mem);
    h->kernel_timing = timing0;
    h->use_graphs = graphs0;
    h->light_q_req = q_req0;
    h->light_grid_req = lg_req0;
    if (rc < 0) return rc;
    if (rc2) return rc2;
    h->light_q_req = h->cfg.dim == 3 ? pick[4] : 0;
    h->light_grid_req = pick[5];
    h->grid_req[0] = pick[0];
    h->grid_req[1] = pick[1];
    h->grid_req[2] = pick[2];
    h->use_side = pick[3];
    drop_graphs(h);
    HIP_TRY(hipMemcpy(h->stats, &st0, sizeof(st0), hipMemcpyHostToDevice));
    if (out) {
        mpf_get_launch(h, out);
        out[4] = (int)(best * 1000.0 + 0.5);
    }
    return 0;
}

extern "C" int mpf_reset_stats(mpf_handle *h)
{
    if (!h) return fail(MPF_EINVAL, "mpf_reset_stats: null handle");
    for (int j = 0; j < 4; ++j) h->kern_ms[j] = 0.0;
    h->kern_steps = 0;
    return reset_stats_dev(h);
}

// Observables of instances [i0, i0 + n): per instance N phase sums, the
// topmost solid cell of every column (fronts one instance after the other)
// and 7 scalars.  One launch over all the instances' planes and columns, in
// a device buffer the handle keeps between calls (grown when a call needs
// more), one copy back.
static int observe_many(mpf_handle *h, int i0, int n, double *grain_volume, int32_t *front, double *out)
{
    if (!grain_volume || !out) return fail(MPF_EINVAL, "observables: null argument");
    const mpf_config &c = h->cfg;
    const int N = c.n_phase;
    if (((size_t)N + 256 * 6) * sizeof(double) > 64 * 1024)
        return fail(MPF_EINVAL, "observables support at most %d phases", 64 * 1024 / 8 - 1536);
    // pre: first plane block, first column block and first column per instance
    std::vector<long long> pre(3 * ((size_t)n + 1));
    long long planes = 0, colblk = 0, cols = 0;
    for (int q = 0; q < n; ++q) {
        const Inst &g = h->inst[i0 + q];
        pre[q] = planes;
        pre[n + 1 + q] = colblk;
        pre[2 * (size_t)n + 2 + q] = cols;
        planes += g.nz;
        colblk += ((long long)g.nx * g.ny + 255) / 256;
        cols += (long long)g.nx * g.ny;
    }
    pre[n] = planes;
    pre[2 * (size_t)n + 1] = colblk;
    pre[3 * (size_t)n + 2] = cols;
    if (planes + colblk > 0x7fffffffLL) return fail(MPF_EINVAL, "observables: too many planes and columns");
    const size_t gv_bytes = (size_t)n * N * 8, part_bytes = (size_t)planes * 6 * 8, fr_bytes = (size_t)cols * 4;
    const size_t res_bytes = (gv_bytes + part_bytes + fr_bytes + 7) & ~(size_t)7;
    const size_t bytes = res_bytes + pre.size() * sizeof(long long);
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (bytes > h->obs_cap) {
        hipFree(h->obs);
        h->obs = nullptr;
        h->obs_cap = 0;
        HIP_TRY(hipMalloc(&h->obs, bytes));
        h->obs_cap = bytes;
    }
    char *base = (char *)h->obs;
    double *gv = (double *)base, *part = (double *)(base + gv_bytes);
    int32_t *fr = (int32_t *)(base + gv_bytes + part_bytes);
    long long *pre_d = (long long *)(base + res_bytes);
    hipError_t e = hipMemcpyAsync(pre_d, pre.data(), pre.size() * sizeof(long long), hipMemcpyHostToDevice,
                                  h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(gv, 0, gv_bytes, h->stream);
    if (e == hipSuccess) {
        if (h->rsz == 8)
            e = mpf_launch_observe_f64(view<double>(h, h->cur), h->inst_d + i0, pre_d, n, planes + colblk, N, gv,
                                       part, fr, h->stream);
        else
            e = mpf_launch_observe_f32(view<float>(h, h->cur), h->inst_d + i0, pre_d, n, planes + colblk, N, gv,
                                       part, fr, h->stream);
    }
    std::vector<char> res(gv_bytes + part_bytes + fr_bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(res.data(), base, res.size(), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail(MPF_EHIP, "observables: %s", hipGetErrorString(e));
    memcpy(grain_volume, res.data(), gv_bytes);
    if (front) memcpy(front, res.data() + gv_bytes + part_bytes, fr_bytes);
    const double *hp = (const double *)(res.data() + gv_bytes);
    for (int q = 0; q < n; ++q) {
        const Inst &g = h->inst[i0 + q];
        double emax = 0.0, pmin = 1e300, pmax = -1e300, nanc = 0.0, solid = 0.0, slots = 0.0;
        for (int k = 0; k < g.nz; ++k) {
            const double *p = hp + (pre[q] + k) * 6;
            emax = p[0] > emax ? p[0] : emax;
            pmin = p[1] < pmin ? p[1] : pmin;
            pmax = p[2] > pmax ? p[2] : pmax;
            nanc += p[3];
            solid += p[4];
            slots += p[5];
        }
        double *o = out + (size_t)q * 7;
        o[0] = emax;
        o[1] = pmin;
        o[2] = pmax;
        o[3] = nanc;
        o[4] = solid;
        o[5] = (double)inst_cells(g);
        o[6] = slots;
    }
    return 0;
}

static int observables_inst(mpf_handle *h, int ii, double *grain_volume, int32_t *front, double *out)
{
    if (!front) return fail(MPF_EINVAL, "observables: null argument");
    return observe_many(h, ii, 1, grain_volume, front, out);
}

extern "C" int mpf_observables(mpf_handle *h, double *grain_volume, int32_t *front, double *out)
{
    if (!h) return fail(MPF_EINVAL, "mpf_observables: null handle");
    if (h->batch) return fail(MPF_EINVAL, "mpf_observables: use mpf_batch_observables for a batch");
    return observables_inst(h, 0, grain_volume, front, out);
}

extern "C" hipError_t mpf_launch_compare(const void *const *pa, int fa, size_t sa, const void *const *pb, int fb,
                                         size_t sb, size_t ncell, double tol, unsigned long long *out,
                                         hipStream_t st);

extern "C" int mpf_compare_phi(mpf_handle *a, mpf_handle *b, double tol, double *out)
{
    if (!a || !b || !out) return fail(MPF_EINVAL, "mpf_compare_phi: null argument");
    if (a->batch || b->batch) return fail(MPF_EINVAL, "mpf_compare_phi: single grids only");
    const mpf_config &ca = a->cfg, &cb = b->cfg;
    if (ca.nx != cb.nx || ca.ny != cb.ny || ca.nz != cb.nz || ca.dim != cb.dim || ca.n_phase != cb.n_phase)
        return fail(MPF_EINVAL, "mpf_compare_phi: the two handles hold different grids or phase counts");
    const Buffers &A = a->b[a->cur], &B = b->b[b->cur];
    const void *pa[4] = {A.liq, A.cnt, A.sid, A.sval};
    const void *pb[4] = {B.liq, B.cnt, B.sid, B.sval};
    unsigned long long *d = nullptr;
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipStreamSynchronize(a->stream));
    HIP_TRY(hipMalloc((void **)&d, 4 * sizeof(unsigned long long)));
    unsigned long long r[4] = {0, 0, 0, 0};
    hipError_t e = hipMemsetAsync(d, 0, sizeof(r), a->stream);
    if (e == hipSuccess)
        e = mpf_launch_compare(pa, a->rsz == 8, a->ncell, pb, b->rsz == 8, b->ncell, a->ncell, tol, d, a->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(a->stream);
    if (e == hipSuccess) e = hipMemcpy(r, d, sizeof(r), hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(MPF_EHIP, "mpf_compare_phi: %s", hipGetErrorString(e));
    double mx, sm;
    memcpy(&mx, &r[0], 8);
    memcpy(&sm, &r[1], 8);
    out[0] = mx;
    out[1] = sm;
    out[2] = (double)r[2];
    out[3] = (double)r[3];
    return 0;
}

extern "C" int mpf_batch_observables(mpf_handle *h, int ii, double *grain_volume, int32_t *front, double *out)
{
    if (int rc = check_inst(h, ii, "mpf_batch_observables")) return rc;
    return observables_inst(h, ii, grain_volume, front, out);
}

extern "C" int mpf_batch_observables_all(mpf_handle *h, double *grain_volume, int32_t *front, double *out)
{
    if (!h) return fail(MPF_EINVAL, "mpf_batch_observables_all: null handle");
    return observe_many(h, 0, (int)h->inst.size(), grain_volume, front, out);
}

static int phase_map_inst(mpf_handle *h, int ii, int axis, int index, int32_t *out)
{
    if (!out) return fail(MPF_EINVAL, "phase_map: null argument");
    if (h->cfg.dim != 3) return fail(MPF_EINVAL, "phase_map: 3D grids only");
    const Inst &g = h->inst[ii];
    const int lim = axis == 0 ? g.nz : (axis == 1 ? g.ny : g.nx);
    if (axis < 0 || axis > 2) return fail(MPF_EINVAL, "phase_map: axis must be 0 (xy), 1 (xz) or 2 (yz)");
    if (index < 0 || index >= lim) return fail(MPF_EINVAL, "phase_map: index %d outside [0, %d)", index, lim);
    const size_t n = (size_t)(axis == 2 ? g.ny : g.nx) * (size_t)(axis == 0 ? g.ny : g.nz);
    void *d = nullptr;
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMalloc(&d, n * 4));
    hipError_t e = h->rsz == 8
                       ? mpf_launch_phase_map_f64(iview<double>(h, h->cur, ii), g.nx, g.ny, g.nz, axis, index,
                                                  (int32_t *)d, h->stream)
                       : mpf_launch_phase_map_f32(iview<float>(h, h->cur, ii), g.nx, g.ny, g.nz, axis, index,
                                                  (int32_t *)d, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d, n * 4, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(MPF_EHIP, "phase_map: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int mpf_phase_map(mpf_handle *h, int axis, int index, int32_t *out)
{
    if (!h) return fail(MPF_EINVAL, "mpf_phase_map: null handle");
    if (h->batch) return fail(MPF_EINVAL, "mpf_phase_map: use mpf_batch_phase_map for a batch");
    return phase_map_inst(h, 0, axis, index, out);
}

extern "C" int mpf_batch_phase_map(mpf_handle *h, int ii, int axis, int index, int32_t *out)
{
    if (int rc = check_inst(h, ii, "mpf_batch_phase_map")) return rc;
    return phase_map_inst(h, ii, axis, index, out);
}

extern "C" int mpf_probe_rays(mpf_handle *h, int n_ray, const int32_t *inst, const int32_t *phase,
                              const double *origin, const double *dir, double threshold, double step_cells,
                              double *out)
{
    if (!h) return fail(MPF_EINVAL, "mpf_probe_rays: null handle");
    if (n_ray < 0) return fail(MPF_EINVAL, "mpf_probe_rays: negative ray count");
    if (n_ray == 0) return 0;
    if (!inst || !phase || !origin || !dir || !out) return fail(MPF_EINVAL, "mpf_probe_rays: null argument");
    if (h->cfg.dim != 3) return fail(MPF_EINVAL, "mpf_probe_rays: 3D grids only");
    if (!(step_cells > 0.0)) return fail(MPF_EINVAL, "mpf_probe_rays: step_cells must be positive");
    for (int r = 0; r < n_ray; ++r) {
        if (inst[r] < 0 || inst[r] >= (int)h->inst.size())
            return fail(MPF_EINVAL, "mpf_probe_rays: ray %d names grid %d of %d", r, inst[r], (int)h->inst.size());
        if (phase[r] < 0 || phase[r] >= h->cfg.n_phase)
            return fail(MPF_EINVAL, "mpf_probe_rays: ray %d names phase %d of %d", r, phase[r], h->cfg.n_phase);
    }
    const size_t n = (size_t)n_ray;
    const size_t bytes = n * 8 * (3 + 3 + 1) + n * 4 * 2;
    void *d = nullptr;
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMalloc(&d, bytes));
    double *d_or = (double *)d, *d_dir = d_or + 3 * n, *d_out = d_dir + 3 * n;
    int32_t *d_in = (int32_t *)(d_out + n), *d_ph = d_in + n;
    hipError_t e = hipMemcpy(d_or, origin, 3 * n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_dir, dir, 3 * n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_in, inst, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_ph, phase, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        if (h->rsz == 8)
            e = mpf_launch_probe_f64(view<double>(h, h->cur), h->inst_d, n_ray, d_in, d_ph, d_or, d_dir, threshold,
                                     step_cells, d_out, h->stream);
        else
            e = mpf_launch_probe_f32(view<float>(h, h->cur), h->inst_d, n_ray, d_in, d_ph, d_or, d_dir, threshold,
                                     step_cells, d_out, h->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d_out, n * 8, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(MPF_EHIP, "mpf_probe_rays: %s", hipGetErrorString(e));
    return 0;
}

// ------------------------------------------------------ drop-in kernels
static hipStream_t as_stream(void *s) { return (hipStream_t)s; }

extern "C" int mpf_dense_apt(const void *phi, int dtype, int32_t *mf, int32_t *nf, int32_t *status, int N,
                             int nx, int ny, int nz, int dim, int depth, int kmax, int checked, void *stream)
{
    if (!phi || !mf || !nf || (checked && !status)) return fail(MPF_EINVAL, "mpf_dense_apt: null argument");
    if (dim != 2 && dim != 3) return fail(MPF_EINVAL, "dim must be 2 or 3");
    if (depth < 1 || kmax < 1) return fail(MPF_EINVAL, "depth and kmax must be positive");
    hipError_t e;
    if (dtype == MPF_FP64)
        e = mpf_launch_dense_apt_f64((const double *)phi, mf, nf, status, N, nx, ny, nz, dim, depth, kmax, checked,
                                     as_stream(stream));
    else if (dtype == MPF_FP32)
        e = mpf_launch_dense_apt_f32((const float *)phi, mf, nf, status, N, nx, ny, nz, dim, depth, kmax, checked,
                                     as_stream(stream));
    else
        return fail(MPF_EINVAL, "bad dtype");
    if (e != hipSuccess) return fail(MPF_EHIP, "mpf_dense_apt: %s", hipGetErrorString(e));
    return 0;
}

// The update on the dense layout with the four tables gathered into `tab`
// (tab_bytes available), the contiguous block the device code indexes.
// sync: wait for the kernel before returning.
static int dense_update_into(const char *who, const void *phi, void *phi_new, int dtype, const void *temp,
                             int temp_dtype, const int32_t *mf, const int32_t *nf, const void *wij,
                             const void *aij, const void *mij, const void *n111, const mpf_scalars *s, int N,
                             int nx, int ny, int nz, int dim, int depth, int kmax, int aniso, int torque,
                             void *stream, void *tab, bool sync)
{
    const size_t esz = dtype == MPF_FP64 ? 8 : 4, nn = (size_t)N * N;
    hipStream_t st = as_stream(stream);
    hipError_t e = hipMemcpyAsync(tab, wij, nn * esz, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync((char *)tab + nn * esz, aij, nn * esz, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync((char *)tab + 2 * nn * esz, mij, nn * esz, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync((char *)tab + 3 * nn * esz, n111, (size_t)N * 24 * esz, hipMemcpyDeviceToDevice, st);
    int t64 = temp_dtype == MPF_FP64;
    int nzz = dim == 3 ? nz : 1;
    if (e == hipSuccess) {
        if (dtype == MPF_FP64) {
            Consts<double> K = make_consts<double>(*s, nx, ny, nzz, N, kmax, aniso, dim == 2 ? 1 : torque, tab);
            if (dim == 2) K.aniso = 1;
            e = mpf_launch_dense_update_f64(&K, dim, (const double *)phi, (double *)phi_new, temp, t64, mf, nf,
                                            depth, st);
        } else {
            Consts<float> K = make_consts<float>(*s, nx, ny, nzz, N, kmax, aniso, dim == 2 ? 1 : torque, tab);
            if (dim == 2) K.aniso = 1;
            e = mpf_launch_dense_update_f32(&K, dim, (const float *)phi, (float *)phi_new, temp, t64, mf, nf,
                                            depth, st);
        }
    }
    if (e == hipSuccess && sync) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(MPF_EHIP, "%s: %s", who, hipGetErrorString(e));
    return 0;
}

static int dense_update_args(const char *who, const void *phi, const void *phi_new, const void *temp,
                             const int32_t *mf, const int32_t *nf, const void *wij, const void *aij,
                             const void *mij, const void *n111, const mpf_scalars *s, int dim, int depth, int kmax)
{
    if (!phi || !phi_new || !temp || !mf || !nf || !wij || !aij || !mij || !n111 || !s)
        return fail(MPF_EINVAL, "%s: null argument", who);
    if (dim != 2 && dim != 3) return fail(MPF_EINVAL, "dim must be 2 or 3");
    if (kmax < 1 || kmax > KCAP) return fail(MPF_EINVAL, "kmax must be in [1, %d]", KCAP);
    if (depth < 1) return fail(MPF_EINVAL, "depth must be positive");
    return 0;
}

extern "C" int mpf_dense_update(const void *phi, void *phi_new, int dtype, const void *temp, int temp_dtype,
                                const int32_t *mf, const int32_t *nf, const void *wij, const void *aij,
                                const void *mij, const void *n111, const mpf_scalars *s, int N, int nx, int ny,
                                int nz, int dim, int depth, int kmax, int aniso, int torque, void *stream)
{
    if (int rc = dense_update_args("mpf_dense_update", phi, phi_new, temp, mf, nf, wij, aij, mij, n111, s, dim,
                                   depth, kmax))
        return rc;
    size_t esz = dtype == MPF_FP64 ? 8 : 4, nn = (size_t)N * N;
    size_t tb = (3 * nn + (size_t)N * 24) * esz;
    // the gathered tables live in a block allocated and freed in the
    // caller's stream order: no device-wide serialisation and no wait, so
    // the call is as asynchronous as the reference's kernel launch
    hipStream_t st = as_stream(stream);
    void *tab = nullptr;
    hipError_t e = hipMallocAsync(&tab, tb, st);
    if (e != hipSuccess) return fail(MPF_ENOMEM, "mpf_dense_update: hipMallocAsync(%zu): %s", tb, hipGetErrorString(e));
    int rc = dense_update_into("mpf_dense_update", phi, phi_new, dtype, temp, temp_dtype, mf, nf, wij, aij, mij,
                               n111, s, N, nx, ny, nz, dim, depth, kmax, aniso, torque, stream, tab, false);
    e = hipFreeAsync(tab, st);
    if (rc) return rc;
    if (e != hipSuccess) return fail(MPF_EHIP, "mpf_dense_update: hipFreeAsync: %s", hipGetErrorString(e));
    return 0;
}

// ---- the update without allocation ------------------------------------------
// The gathered tables of mpf_dense_update_noalloc live in this file-scope
// device array, one copy per GPU for the life of the process.  Host code
// never passes its name: the address comes from hipGetSymbolAddress, taken
// once per device (with that device current) on the first call there.
constexpr int NOALLOC_MAX_PHASES = 256;
constexpr size_t NOALLOC_SCRATCH_BYTES =
    (3 * (size_t)NOALLOC_MAX_PHASES * NOALLOC_MAX_PHASES + (size_t)NOALLOC_MAX_PHASES * 24) * 8;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

static int scratch_on_current_device(void **out)
{
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(MPF_EINVAL, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MPF_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" int mpf_dense_update_noalloc(const void *phi, void *phi_new, int dtype, const void *temp, int temp_dtype,
                                        const int32_t *mf, const int32_t *nf, const void *wij, const void *aij,
                                        const void *mij, const void *n111, const mpf_scalars *s, int N, int nx,
                                        int ny, int nz, int dim, int depth, int kmax, int aniso, int torque,
                                        void *stream)
{
    if (int rc = dense_update_args("mpf_dense_update_noalloc", phi, phi_new, temp, mf, nf, wij, aij, mij, n111, s,
                                   dim, depth, kmax))
        return rc;
    if (N < 1) return fail(MPF_EINVAL, "mpf_dense_update_noalloc: n_phase must be positive");
    if (dtype != MPF_FP32 && dtype != MPF_FP64) return fail(MPF_EINVAL, "mpf_dense_update_noalloc: bad dtype");
    if (N > NOALLOC_MAX_PHASES)
        return fail(MPF_EINVAL, "mpf_dense_update_noalloc: n_phase %d exceeds the %d phases its tables hold", N,
                    NOALLOC_MAX_PHASES);
    void *tab = nullptr;
    if (int rc = scratch_on_current_device(&tab)) return rc;
    return dense_update_into("mpf_dense_update_noalloc", phi, phi_new, dtype, temp, temp_dtype, mf, nf, wij, aij,
                             mij, n111, s, N, nx, ny, nz, dim, depth, kmax, aniso, torque, stream, tab, false);
}

extern "C" int mpf_dense_temp(void *temp, int temp_dtype, double cooling, size_t n, void *stream)
{
    if (!temp) return fail(MPF_EINVAL, "mpf_dense_temp: null argument");
    hipError_t e = mpf_launch_dense_temp_f64(temp, temp_dtype == MPF_FP64, cooling, n, as_stream(stream));
    if (e != hipSuccess) return fail(MPF_EHIP, "mpf_dense_temp: %s", hipGetErrorString(e));
    return 0;
}

// ------------------------------------------------------------ memory
extern "C" int mpf_malloc(void **ptr, size_t bytes)
{
    if (!ptr) return fail(MPF_EINVAL, "mpf_malloc: null argument");
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(MPF_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return 0;
}
extern "C" int mpf_free(void *ptr)
{
    HIP_TRY(hipFree(ptr));
    return 0;
}
extern "C" int mpf_memcpy_htod(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}
extern "C" int mpf_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}
extern "C" int mpf_memcpy_dtod(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return 0;
}
extern "C" int mpf_memset(void *ptr, int value, size_t bytes)
{
    HIP_TRY(hipMemset(ptr, value, bytes));
    return 0;
}
extern "C" int mpf_device_synchronize(void)
{
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}
