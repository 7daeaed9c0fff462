// mpf_gather.h -- slot-layout access of one cell: stencil addressing, the
// k-way merge that forms the APT active set (src/gpu_kernels_3d.py:462-527)
// from the sorted grain slots of the cell and its face neighbours, and the
// write-back of the new state.  Shared by the step kernels and the host
// build used by the CPU emulation tests.
#pragma once

#include "mpf_cell.h"
#include "mpf_state.h"

namespace mpf {

// ------------------------------------------------------------------ helpers
MPF_DEV size_t cidx(int l, int m, int k, int nx, int ny)
{
    return ((size_t)k * ny + m) * nx + l;
}

// Reference dense index phi[g][l][m][k] (3D) / phi[g][l][m] (2D).
MPF_DEV size_t ridx(int l, int m, int k, int ny, int nz)
{
    return ((size_t)l * ny + m) * nz + k;
}

// Stencil point cell indices in the slot layout: c, x+, x-, y+, y-, (z+, z-).
template <typename Real, int DIM>
MPF_DEV void stencil_cells(const Consts<Real> &K, int l, int m, int k, size_t *pc)
{
    int lp = bc_pp(l, K.nx), lm = bc_pm(l, K.nx);
    if constexpr (DIM == 3) {
        int mp = bc_pp(m, K.ny), mm = bc_pm(m, K.ny);
        int kp = bc_mp(k, K.nz), km = bc_mm(k, K.nz);
        pc[P_C] = cidx(l, m, k, K.nx, K.ny);
        pc[P_XP] = cidx(lp, m, k, K.nx, K.ny);
        pc[P_XM] = cidx(lm, m, k, K.nx, K.ny);
        pc[P_YP] = cidx(l, mp, k, K.nx, K.ny);
        pc[P_YM] = cidx(l, mm, k, K.nx, K.ny);
        pc[P_ZP] = cidx(l, m, kp, K.nx, K.ny);
        pc[P_ZM] = cidx(l, m, km, K.nx, K.ny);
    } else {
        int mp = bc_mp(m, K.ny), mm = bc_mm(m, K.ny);
        pc[P_C] = cidx(l, m, 0, K.nx, K.ny);
        pc[P_XP] = cidx(lp, m, 0, K.nx, K.ny);
        pc[P_XM] = cidx(lm, m, 0, K.nx, K.ny);
        pc[P_YP] = cidx(l, mp, 0, K.nx, K.ny);
        pc[P_YM] = cidx(l, mm, 0, K.nx, K.ny);
    }
}

template <typename Real>
struct SlotLiq {
    const Real *liq;
    int nx, ny;
    MPF_DEV Real operator()(int l, int m, int k) const
    {
        return liq[cidx(l, m, k, nx, ny)];
    }
};

template <typename Real>
struct DenseLiq {
    const Real *phi; // phase 0 of the dense field
    int ny, nz;
    MPF_DEV Real operator()(int l, int m, int k) const
    {
        return phi[ridx(l, m, k, ny, nz)];
    }
};

// k-way merge of the sorted slot lists of the stencil points.  Yields the
// union of grain ids present at any of the points in ascending order with the
// value at every point (0 where absent).
//
// Each list keeps its head and the slot after it in registers: the counts
// are read in the round trip of the caller's liquid values, both slots in one
// round trip after the counts, and a refill (lists of 3 or more slots) is
// needed only two merge steps after it is issued, so the merge does not wait
// on memory once per step.  (Reading slots 0 and 1 without waiting for the
// counts, masked afterwards, measured 12% slower on the merge pass: more
// registers live across the loads and more lines fetched.)
template <typename Real, int DIM>
struct Merge {
    static constexpr int NP = (DIM == 3) ? 7 : 5;
    size_t pc[NP];
    int cnt[NP], pos[NP];   // pos: slot index of the head
    uint32_t hid[NP], nid[NP];
    Real hv[NP], nv[NP];

    MPF_DEV void init(const SlotView<Real> &S, size_t ncell)
    {
#pragma unroll
        for (int p = 0; p < NP; ++p) cnt[p] = S.cnt[pc[p]];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            pos[p] = 0;
            hid[p] = nid[p] = NO_ID;
            hv[p] = nv[p] = Real(0.0);
            if (cnt[p] > 0) {
                hid[p] = S.sid[pc[p]];
                hv[p] = S.sval[pc[p]];
            }
            if (cnt[p] > 1) {
                nid[p] = S.sid[ncell + pc[p]];
                nv[p] = S.sval[ncell + pc[p]];
            }
        }
    }

    // Next candidate; returns NO_ID when all lists are exhausted.
    MPF_DEV uint32_t next(const SlotView<Real> &S, size_t ncell, Real *vals)
    {
        uint32_t g = hid[0];
#pragma unroll
        for (int p = 1; p < NP; ++p) g = hid[p] < g ? hid[p] : g;
        if (g == NO_ID) return NO_ID;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            if (hid[p] == g) {
                vals[p] = hv[p];
                pos[p] += 1;
                hid[p] = nid[p];
                hv[p] = nv[p];
                nid[p] = NO_ID;
                if (pos[p] + 1 < cnt[p]) {
                    size_t o = (size_t)(pos[p] + 1) * ncell + pc[p];
                    nid[p] = S.sid[o];
                    nv[p] = S.sval[o];
                }
            } else {
                vals[p] = Real(0.0);
            }
        }
        return g;
    }
};

// APT rule, src/gpu_kernels_3d.py:478-482: phi > 0 at the cell, or phi == 0
// there and > 0 at a face neighbour.
// Evaluated without branches (bitwise on the comparisons), so that no load
// issued after it waits behind its operands.
template <int NP, typename Real>
MPF_DEV bool apt_active(const Real *vals)
{
    bool any = false;
#pragma unroll
    for (int p = 1; p < NP; ++p) any = any | (vals[p] > Real(0.0));
    return (vals[0] > Real(0.0)) | ((vals[0] == Real(0.0)) & any);
}

// Gather the active set of one cell (APT with the checked KMAX clamp).
// Returns the total number of active phases (before the clamp) in *ntot and
// whether the kept solids fit in A (*fits).  no_liquid: the caller knows the
// liquid is +0 at all stencil points (the classification pass saw it), so its
// values are not read.
template <typename Real, int DIM, int A>
MPF_DEV void gather(const Consts<Real> &K, const SlotView<Real> &S, size_t ncell,
                                       int l, int m, int k, Cell<Real, DIM, A> &C, int *ntot, bool *fits,
                                       bool no_liquid = false)
{
    constexpr int NP = Cell<Real, DIM, A>::NP;
    constexpr bool SMALL = (A <= 8);
    Merge<Real, DIM> M;
    stencil_cells<Real, DIM>(K, l, m, k, M.pc);
    // liquid (dense) and the counts, issued before either is used
    Real lv[NP];
    if (no_liquid) {
#pragma unroll
        for (int p = 0; p < NP; ++p) lv[p] = Real(0.0);
    } else {
#pragma unroll
        for (int p = 0; p < NP; ++p) lv[p] = S.liq[M.pc[p]];
    }
    M.init(S, ncell);
    C.liq = apt_active<NP>(lv);
    C.gid[0] = 0;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        C.v[0][p] = lv[p];
        C.dom1[p] = Real(0.0);
    }
    int keep = K.kmax - (C.liq ? 1 : 0); // solids kept under the KMAX clamp
    int nact = 0;                        // active solids found
    C.nsol = 0;
    Real vals[NP];
    if constexpr (SMALL) {
        // unrolled over output positions so that the arrays stay in registers
#pragma unroll
        for (int u = 1; u <= A; ++u) {
            if (nact < keep) {
                uint32_t g;
                bool act = false;
                while ((g = M.next(S, ncell, vals)) != NO_ID) {
                    if constexpr (DIM == 2) {
                        if (g == 1) {
#pragma unroll
                            for (int p = 0; p < NP; ++p) C.dom1[p] = vals[p];
                        }
                    }
                    if (apt_active<NP>(vals)) {
                        act = true;
                        break;
                    }
                }
                if (act) {
                    nact += 1;
                    C.nsol = u;
                    C.gid[u] = (int)g;
#pragma unroll
                    for (int p = 0; p < NP; ++p) C.v[u][p] = vals[p];
                }
            }
        }
    }
    // count (and, for the general path, store) the remaining active solids
    uint32_t g;
    while ((g = M.next(S, ncell, vals)) != NO_ID) {
        if constexpr (DIM == 2) {
            if (g == 1) {
#pragma unroll
                for (int p = 0; p < NP; ++p) C.dom1[p] = vals[p];
            }
        }
        if (!apt_active<NP>(vals)) continue;
        nact += 1;
        if constexpr (!SMALL) {
            if (nact <= keep && nact <= A) {
                C.nsol = nact;
                C.gid[nact] = (int)g;
                for (int p = 0; p < NP; ++p) C.v[nact][p] = vals[p];
            }
        }
    }
    *ntot = nact + (C.liq ? 1 : 0);
    int kept = nact < keep ? nact : keep;
    *fits = (kept <= A);
}

template <typename Real, int DIM, int A>
MPF_DEV void write_cell(const SlotView<Real> &O, size_t ncell, size_t c,
                                           const Cell<Real, DIM, A> &C)
{
    constexpr bool SMALL = (A <= 8);
    constexpr int UNR_A = SMALL ? A : 1;
    O.liq[c] = C.out[0];
    int w = 0;
#pragma unroll UNR_A
    for (int u = 1; u <= A; ++u) {
        if (u <= C.nsol && C.out[u] != Real(0.0)) {
            size_t o = (size_t)w * ncell + c;
            O.sid[o] = (uint16_t)C.gid[u];
            O.sval[o] = C.out[u];
            w += 1;
        }
    }
    O.cnt[c] = (uint8_t)w;
}

// Update of a cell whose active set has at most one phase (n_act = 0 or 1):
// dphi = 0, so phi_new = clip(phi + dt*0) and the renormalisation
// (src/gpu_kernels_3d.py:661-680) either scales it by 1/phi or, below 1e-20,
// sets the liquid to 1.
template <typename Real>
MPF_DEV void single_phase_update(Real v, bool act, bool is_liquid, Real dt, Real &liq_new,
                                                    Real &sol_new)
{
    liq_new = Real(0.0);
    sol_new = Real(0.0);
    Real s = Real(0.0), out = Real(0.0);
    if (act) {
        out = v + dt * Real(0.0);
        if (out < Real(0.0)) out = Real(0.0);
        if (out > Real(1.0)) out = Real(1.0);
        s += out;
    }
    if (s > Real(1e-20)) {
        Real invs = Real(1.0) / s;
        out *= invs;
        if (is_liquid)
            liq_new = out;
        else
            sol_new = out;
    } else {
        liq_new = Real(1.0);
        if (act && !is_liquid) sol_new = out;
    }
}


} // namespace mpf
