// Host build of the solver's per-cell code over the slot layout (test
// infrastructure).  Runs one time step the way the step kernels do -- the
// classification pass (single-phase stencils), the 4-grain merge/front path,
// the 8-grain register path and the scratch path with room for KMAX phases --
// sequentially on the CPU, so that the device code's arithmetic and APT can be
// compared with the oracle without a GPU.
//
// mode 0: the kernels' dispatch (single-phase stencils, 4-grain register
// path, scratch path); 1: every cell through an 8-grain register path
// (scratch beyond 8); 2: every cell through the scratch path.
#include <cstdint>
#include <cstring>

#include "mpf_consts.h"
#include "mpf_gather.h"

using namespace mpf;

template <typename R, int DIM, int A, int TC>
static bool run_cell(const Consts<R> &K, const SlotView<R> &in, const SlotView<R> &out, size_t ncell, int l, int m,
                     int k, double T)
{
    Cell<R, DIM, A> C;
    int ntot;
    bool fits;
    gather<R, DIM, A>(K, in, ncell, l, m, k, C, &ntot, &fits);
    if (!fits) return false;
    SlotLiq<R> L{in.liq, K.nx, K.ny};
    cell_update<R, DIM, A, true, TC>(K, L, l, m, k, T, C);
    write_cell<R, DIM, A>(out, ncell, cidx(l, m, k, K.nx, K.ny), C);
    return true;
}

template <typename R, int DIM, int TC>
static int step(const Consts<R> &K, SlotView<R> in, SlotView<R> out, const double *T, int mode, int *paths,
                int8_t *path_of)
{
    constexpr int NP = DIM == 3 ? 7 : 5;
    const int nz = DIM == 3 ? K.nz : 1;
    const size_t ncell = (size_t)K.nx * K.ny * nz;
    for (int k = 0; k < nz; ++k)
        for (int m = 0; m < K.ny; ++m)
            for (int l = 0; l < K.nx; ++l) {
                size_t c = cidx(l, m, k, K.nx, K.ny);
                if (mode == 0) {
                    size_t pc[7];
                    stencil_cells<R, DIM>(K, l, m, k, pc);
                    int sum = 0;
                    bool single = true;
                    R pl[NP], psv[NP];
                    uint32_t psid[NP];
                    for (int p = 0; p < NP; ++p) {
                        int cn = in.cnt[pc[p]];
                        pl[p] = in.liq[pc[p]];
                        psid[p] = cn ? in.sid[pc[p]] : NO_ID;
                        psv[p] = cn ? in.sval[pc[p]] : R(0);
                        sum += cn;
                        single = single && cn == 1 && pl[p] == R(0);
                    }
                    for (int p = 0; p < NP; ++p) single = single && psid[p] == psid[0];
                    if (sum == 0 || single) {
                        bool liquid = sum == 0;
                        R *vals = liquid ? pl : psv;
                        bool act = apt_active<NP, R>(vals);
                        R ln, sn;
                        single_phase_update(vals[0], act, liquid, K.dt, ln, sn);
                        out.liq[c] = ln;
                        out.cnt[c] = 0;
                        if (!liquid && sn != R(0)) {
                            out.sid[c] = (uint16_t)psid[0];
                            out.sval[c] = sn;
                            out.cnt[c] = 1;
                        }
                        paths[0]++;
                        if (path_of) path_of[c] = 0;
                        continue;
                    }
                    if (run_cell<R, DIM, 4, TC>(K, in, out, ncell, l, m, k, T[c])) {
                        paths[1]++;
                        if (path_of) path_of[c] = 1;
                        continue;
                    }
                }
                if (mode == 1 && run_cell<R, DIM, 8, TC>(K, in, out, ncell, l, m, k, T[c])) {
                    paths[2]++;
                    if (path_of) path_of[c] = 2;
                    continue;
                }
                run_cell<R, DIM, KCAP, TC>(K, in, out, ncell, l, m, k, T[c]);
                paths[3]++;
                if (path_of) path_of[c] = 3;
            }
    return 0;
}

template <typename R>
static int step_any(Consts<R> &K, int dim, int use_class, SlotView<R> in, SlotView<R> out, const double *T, int mode,
                    int *paths, int8_t *path_of)
{
    R cls[3][2];
    const bool tc = use_class && class_tables<R>(K.tab, K.N, cls);
    if (tc)
        for (int t = 0; t < 3; ++t) K.cls[t][0] = cls[t][0], K.cls[t][1] = cls[t][1];
    if (dim == 3)
        return tc ? step<R, 3, 1>(K, in, out, T, mode, paths, path_of) : step<R, 3, 0>(K, in, out, T, mode, paths, path_of);
    return tc ? step<R, 2, 1>(K, in, out, T, mode, paths, path_of) : step<R, 2, 0>(K, in, out, T, mode, paths, path_of);
}

// use_class bit 0: select the class-form table kernels when the tables have
// that form (as the solver does); returns -1 if asked for but the form is
// absent.  Bit 1: scan four facet rows when the n111 rows come in opposite
// pairs (as the solver does); returns -2 if asked for but they do not.
extern "C" int emu_step(int precision, int dim, const mpf_scalars *s, int nx, int ny, int nz, int N, int kmax,
                        int aniso, int torque, const void *tables, const void *liq_in, const uint8_t *cnt_in,
                        const uint16_t *sid_in, const void *sval_in, void *liq_out, uint8_t *cnt_out,
                        uint16_t *sid_out, void *sval_out, const double *T, int mode, int *paths, int8_t *path_of,
                        int use_class)
{
    const size_t ncell = (size_t)nx * ny * (dim == 3 ? nz : 1);
    const bool want_pairs = (use_class & 2) != 0;
    use_class &= 1;
    if (precision == MPF_FP64) {
        Consts<double> K = make_consts<double>(*s, nx, ny, nz, N, kmax, dim == 2 ? 1 : aniso, dim == 2 ? 1 : torque,
                                               tables);
        if (want_pairs && !n111_pairs<double>(K.n111(), N)) return -2;
        K.n111_pairs = want_pairs ? 1 : 0;
        double cls[3][2];
        if (use_class && !class_tables<double>(K.tab, N, cls)) return -1;
        SlotView<double> in{(double *)liq_in, (uint8_t *)cnt_in, (uint16_t *)sid_in, (double *)sval_in, ncell};
        SlotView<double> out{(double *)liq_out, cnt_out, sid_out, (double *)sval_out, ncell};
        return step_any<double>(K, dim, use_class, in, out, T, mode, paths, path_of);
    }
    Consts<float> K = make_consts<float>(*s, nx, ny, nz, N, kmax, dim == 2 ? 1 : aniso, dim == 2 ? 1 : torque,
                                         tables);
    if (want_pairs && !n111_pairs<float>(K.n111(), N)) return -2;
    K.n111_pairs = want_pairs ? 1 : 0;
    float cls[3][2];
    if (use_class && !class_tables<float>(K.tab, N, cls)) return -1;
    SlotView<float> in{(float *)liq_in, (uint8_t *)cnt_in, (uint16_t *)sid_in, (float *)sval_in, ncell};
    SlotView<float> out{(float *)liq_out, cnt_out, sid_out, (float *)sval_out, ncell};
    return step_any<float>(K, dim, use_class, in, out, T, mode, paths, path_of);
}
