// mpf_cell.h -- per-cell multi-phase-field update for gfx950.
//
// The physics of one cell, independent of how the state is stored: the caller
// gathers the active set (liquid flag + ascending solid ids) and the values of
// every active phase at the cell and its face neighbours, plus a functor that
// reads the liquid field at any cell (the anisotropy terms need radius 2), and
// gets back the new values.  The slot-layout step kernel and the dense
// drop-in kernels both use it.
//
// Reference: src/gpu_kernels_3d.py:24-441 (helpers), :530-830 (update kernels)
// and src/gpu_kernels.py:25-515 (2D).  In FP64 every expression is evaluated in
// the order the reference writes it (built with -ffp-contract=off), so the
// result is bit-identical to the CPU oracle.  Values the reference recomputes
// several times per cell (liquid gradients at the 7 points, the facet search,
// Laplacians inside the O(n^3) loop) are computed once here; because they are
// deterministic functions of the same inputs this changes no bits.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MPF_DEV __device__ __forceinline__
#else
// Host build of the per-cell code (CPU emulation tests).
#include <cmath>
#include <cstddef>
#define MPF_DEV inline
#endif

#include "mpf_math.h"

namespace mpf {

// Stencil point order used everywhere: centre, x+, x-, y+, y-, z+, z-.
enum { P_C = 0, P_XP = 1, P_XM = 2, P_YP = 3, P_YM = 4, P_ZP = 5, P_ZM = 6 };

template <typename Real>
struct Consts {
    int nx, ny, nz, N, kmax;
    int aniso, torque;
    Real dx, dt;
    Real two_dx;           // 2.0*dx
    Real inv_dx2;          // 1.0/(dx*dx)
    Real q_inv_dx2;        // 0.25*inv_dx2
    Real half_over_dx;     // 0.5/dx
    Real eps0_sl, eps0_sq; // eps0, eps0*eps0
    Real w0_sl;
    Real tan_a0, delta_a, mu_a, mu_delta, p2; // mu_a*delta_a, p_round*p_round
    Real g2_floor, ksi, theta_c;
    double Sf, T_melt;
    // wij, aij, mij [N][N] and n111 [N][8][3], contiguous in that order.  One
    // pointer instead of four keeps the kernels' scalar registers free.
    const Real *tab;
    uint32_t nn; // N*N (N <= 65535)
    MPF_DEV const Real *n111() const { return tab + 3 * (size_t)nn; }
    // Values of the class-form tables (used by kernels built with TC = 1):
    // cls[table][0] for solid-liquid pairs, cls[table][1] for grain-grain
    // pairs; table 0 = wij, 1 = aij, 2 = mij.
    Real cls[3][2];
    // 1 when every grain's n111 rows come in opposite pairs, row 7 - t equal
    // to row t with every sign bit flipped (the rotated (+-1,+-1,+-1)/sqrt(3)
    // of compute_rotated_n111 in its order); checked bit for bit by the host.
    int n111_pairs;
};

// Interaction-table lookup.  TC = 1: the tables have the class form that
// build_interaction_matrices gives them (src/seed_modes.py:333-380), with +0
// on the diagonal, one value for every solid-liquid pair and one for every
// grain-grain pair.  The host checks that form bit for bit before selecting
// these kernels, so the value is selected instead of loaded and the result
// is identical.
enum { TAB_W = 0, TAB_A = 1, TAB_M = 2 };

// A kernel-uniform value moved into a vector register.  The unrolled pair
// loop selects among the class values many times; kept in scalar registers,
// which the kernel has run out of, they would be reloaded from spill lanes at
// every use.
MPF_DEV float uniform_in_vgpr(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    float r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
#else
    return x;
#endif
}

MPF_DEV double uniform_in_vgpr(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r;
    asm volatile("v_mov_b64 %0, %1" : "=v"(r) : "s"(x));
    return r;
#else
    return x;
#endif
}

// Lane of the wavefront and a value read from another lane (lane groups of
// cell_update).  The host build runs one cell per call and never groups
// lanes, so there they are lane 0 and the identity.
MPF_DEV int lane_id()
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (int)__lane_id();
#else
    return 0;
#endif
}

template <typename T>
MPF_DEV T lane_read(T v, int src)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __shfl(v, src, 64);
#else
    (void)src;
    return v;
#endif
}

template <typename Real>
struct ClassVals {
    Real v[3][2];
};

template <int TC, typename Real>
MPF_DEV ClassVals<Real> class_vals(const Consts<Real> &K)
{
    ClassVals<Real> c;
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) c.v[t][u] = TC ? uniform_in_vgpr(K.cls[t][u]) : Real(0.0);
    return c;
}

template <int TC, int T, typename Real>
MPF_DEV Real tab_at(const Consts<Real> &K, const ClassVals<Real> &cv, int i, int j)
{
    if constexpr (TC != 0) {
        (void)K;
        if (i == j) return Real(0.0);
        return (i == 0 || j == 0) ? cv.v[T][0] : cv.v[T][1];
    } else {
        (void)cv;
        return K.tab[(size_t)T * K.nn + (size_t)i * K.N + j];
    }
}

// ---------------------------------------------------------------- math glue
MPF_DEV double m_sqrt(double x) { return sqrt(x); }
MPF_DEV float m_sqrt(float x) { return sqrtf(x); }
MPF_DEV double m_abs(double x) { return fabs(x); }
MPF_DEV float m_abs(float x) { return fabsf(x); }
MPF_DEV double m_tan(double x) { return mpf_tan(x); }
MPF_DEV float m_tan(float x) { return tanf(x); }
MPF_DEV double m_acos(double x) { return mpf_acos(x); }
MPF_DEV float m_acos(float x) { return acosf(x); }
MPF_DEV double m_tanh(double x) { return mpf_tanh(x); }
MPF_DEV float m_tanh(float x) { return tanhf(x); }
template <typename Real>
MPF_DEV Real pymax(Real a, Real b) { return (b > a) ? b : a; }

// Quotient a / b: the IEEE division in double; in float on the device, the
// product with the hardware reciprocal (about 1 ulp; the expanded division
// costs five instructions, about 8% of an interface cell), which the FP32
// contract allows (results within 1e-4 of the FP64 oracle).
MPF_DEV double m_div(double a, double b) { return a / b; }
MPF_DEV float m_div(float a, float b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return a * __builtin_amdgcn_rcpf(b);
#else
    return a / b;
#endif
}

// calc_a_from_cos, src/gpu_kernels_3d.py:24-30
template <typename Real>
MPF_DEV Real calc_a(const Consts<Real> &K, Real cost)
{
    Real c2 = cost * cost;
    Real C = m_sqrt(c2 + K.p2);
    Real S = m_sqrt(pymax(Real(1.0) - c2, Real(0.0)) + K.p2);
    return K.mu_a * (Real(1.0) + K.delta_a * (C + K.tan_a0 * S));
}

// calc_b_from_cos, src/gpu_kernels_3d.py:33-55
template <typename Real>
MPF_DEV Real calc_b(const Consts<Real> &K, Real best_cost)
{
    Real c = best_cost;
    if (c > Real(1.0))
        c = Real(1.0);
    else if (c < Real(-1.0))
        c = Real(-1.0);
    Real theta = m_acos(c);
    if (theta >= K.theta_c) return Real(1.0);
    Real x = m_div(theta, K.theta_c);
    const Real eps = Real(1.0e-6);
    const Real half_pi = Real(0.5 * 3.14159265358979311600e+00);
    Real y = half_pi * x;
    if (y > half_pi - eps) y = half_pi - eps;
    if (y < eps) y = eps;
    Real t = m_tan(y);
    return K.ksi + (Real(1.0) - K.ksi) * t * m_tanh(m_div(Real(1.0), t));
}

// ------------------------------------------------------ boundary conditions
// src/gpu_kernels_3d.py:58-91 (x,y periodic; z mirror) and
// src/gpu_kernels.py:84-105 (2D: x periodic, y mirror).
MPF_DEV int bc_pp(int i, int n) { return i < n - 1 ? i + 1 : 0; }
MPF_DEV int bc_pm(int i, int n) { return i > 0 ? i - 1 : n - 1; }
MPF_DEV int bc_mp(int i, int n) { return i < n - 1 ? i + 1 : n - 1; }
MPF_DEV int bc_mm(int i, int n) { return i > 0 ? i - 1 : 0; }

// Result of the {111} facet search at one stencil point for one grain.
template <typename Real>
struct Facet {
    Real cs, ca, nx, ny, nz; // signed cos, |cos|, normal of the best facet
};

// facet_cos_and_nxyz_from_grad (src/gpu_kernels_3d.py:251-301) given the unit
// gradient; best_cos_from_grad_3d (:110-134) returns the same |cos|.
//
// The scan keeps the first facet of largest |cos| (strictly greater
// replaces).  When the rows come in opposite pairs (n111_pairs), rows 4..7
// give the |cos| of rows 3..0 again (negating a dot product only flips its
// sign), so they can never replace and the first four rows decide: n_lo holds
// rows 0..3 (12 values), rows 4..7 are read from n_hi (the table's row 4)
// only when pairs is false.
template <typename Real, int DIM>
MPF_DEV void facet_visit(Facet<Real> &f, Real ux, Real uy, Real uz, Real gx, Real gy, Real gz)
{
    Real dot;
    if constexpr (DIM == 3)
        dot = gx * ux + gy * uy + gz * uz;
    else
        dot = gx * ux + gy * uy;
    Real ad = m_abs(dot);
    if (ad > f.ca) {
        f.ca = ad;
        f.cs = dot;
        f.nx = ux;
        f.ny = uy;
        f.nz = uz;
    }
}

template <typename Real, int DIM>
MPF_DEV Facet<Real> facet_search(const Real *n_lo, const Real *n_hi, bool pairs, Real gx, Real gy, Real gz)
{
    Facet<Real> f;
    f.cs = Real(0.0);
    f.ca = Real(0.0);
    f.nx = Real(1.0);
    f.ny = Real(0.0);
    f.nz = Real(0.0);
#pragma unroll
    for (int t = 0; t < 4; ++t) facet_visit<Real, DIM>(f, n_lo[t * 3], n_lo[t * 3 + 1], n_lo[t * 3 + 2], gx, gy, gz);
    if (!pairs) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
            facet_visit<Real, DIM>(f, n_hi[t * 3], n_hi[t * 3 + 1], n_hi[t * 3 + 2], gx, gy, gz);
    }
    if (f.ca > Real(1.0)) f.ca = Real(1.0);
    if (f.cs > Real(1.0)) f.cs = Real(1.0);
    if (f.cs < Real(-1.0)) f.cs = Real(-1.0);
    return f;
}

// Liquid gradient at one stencil point with its derived quantities.
template <typename Real>
struct LGrad {
    Real gx, gy, gz; // central differences (grad_phi_xyz)
    Real q;          // |grad|^2
    Real inv_g;      // 1/sqrt(q)       (valid when q >= g2_floor, else 0)
    Real inv_g3;     // 1/(q*sqrt(q))
};

template <typename Real, int DIM>
MPF_DEV void lgrad_finish(const Consts<Real> &K, LGrad<Real> &G)
{
    if constexpr (DIM == 3)
        G.q = G.gx * G.gx + G.gy * G.gy + G.gz * G.gz;
    else
        G.q = G.gx * G.gx + G.gy * G.gy;
    if (!(G.q < K.g2_floor)) {
        Real sq = m_sqrt(G.q);
        G.inv_g = m_div(Real(1.0), sq);
        G.inv_g3 = m_div(Real(1.0), G.q * sq);
    } else {
        G.inv_g = G.inv_g3 = Real(0.0);
    }
}

template <typename Real>
MPF_DEV bool lgrad_ok(const Consts<Real> &K, const LGrad<Real> &G)
{
    return !(G.q < K.g2_floor);
}

// A12 quantities at one point: a, coef and dc/dphi (da/dphi = mu_delta*coef*dc).
template <typename Real>
struct A12v {
    Real a, dax, day, daz;
};

// da_dphixyz_A12, src/gpu_kernels_3d.py:304-341 (2D: src/gpu_kernels.py:237-264)
// a_at_point is calc_a(f.ca), already evaluated for eps^2 at the same point.
template <typename Real, int DIM>
MPF_DEV A12v<Real> a12_at(const Consts<Real> &K, const LGrad<Real> &G, const Facet<Real> &f, Real a_at_point)
{
    A12v<Real> r;
    if (!lgrad_ok(K, G)) {
        r.a = r.dax = r.day = r.daz = Real(0.0);
        return r;
    }
    r.a = a_at_point;
    Real udotg;
    if constexpr (DIM == 3)
        udotg = G.gx * f.nx + G.gy * f.ny + G.gz * f.nz;
    else
        udotg = G.gx * f.nx + G.gy * f.ny;
    Real dcx = f.nx * G.inv_g - udotg * G.gx * G.inv_g3;
    Real dcy = f.ny * G.inv_g - udotg * G.gy * G.inv_g3;
    Real c = f.cs;
    Real C = m_sqrt(c * c + K.p2);
    Real S = m_sqrt(pymax(Real(1.0) - c * c, Real(0.0)) + K.p2);
    Real coef = m_div(c, C) - K.tan_a0 * m_div(c, S);
    r.dax = K.mu_delta * coef * dcx;
    r.day = K.mu_delta * coef * dcy;
    if constexpr (DIM == 3) {
        Real dcz = f.nz * G.inv_g - udotg * G.gz * G.inv_g3;
        r.daz = K.mu_delta * coef * dcz;
    } else {
        r.daz = Real(0.0);
    }
    return r;
}

// Reads the liquid field at absolute cell coordinates.
//   Real operator()(int l, int m, int k) const
// Liquid anisotropy context of one cell: gradients at the centre and the face
// neighbours plus the second derivatives at the centre.
template <typename Real, int DIM>
struct LiqCtx {
    static constexpr int NP = (DIM == 3) ? 7 : 5;
    LGrad<Real> G[NP];
    Real v[NP];                         // liquid values at the stencil points
    Real dxx, dyy, dzz, dxy, dxz, dyz;  // d2_phi_xyz at the centre
};

template <typename Real, int DIM, typename LiqAt>
MPF_DEV void lgrad_at(const Consts<Real> &K, const LiqAt &L, int l, int m, int k,
                                         LGrad<Real> &G)
{
    if constexpr (DIM == 3) {
        int lp = bc_pp(l, K.nx), lm = bc_pm(l, K.nx);
        int mp = bc_pp(m, K.ny), mm = bc_pm(m, K.ny);
        int kp = bc_mp(k, K.nz), km = bc_mm(k, K.nz);
        G.gx = m_div(L(lp, m, k) - L(lm, m, k), K.two_dx);
        G.gy = m_div(L(l, mp, k) - L(l, mm, k), K.two_dx);
        G.gz = m_div(L(l, m, kp) - L(l, m, km), K.two_dx);
    } else {
        int lp = bc_pp(l, K.nx), lm = bc_pm(l, K.nx);
        int mp = bc_mp(m, K.ny), mm = bc_mm(m, K.ny);
        G.gx = m_div(L(lp, m, 0) - L(lm, m, 0), K.two_dx);
        G.gy = m_div(L(l, mp, 0) - L(l, mm, 0), K.two_dx);
        G.gz = Real(0.0);
    }
    lgrad_finish<Real, DIM>(K, G);
}

template <typename Real, int DIM, typename LiqAt>
MPF_DEV void build_liq_ctx(const Consts<Real> &K, const LiqAt &L, int l, int m, int k,
                                           LiqCtx<Real, DIM> &X)
{
    int lp = bc_pp(l, K.nx), lm = bc_pm(l, K.nx);
    if constexpr (DIM == 3) {
        int mp = bc_pp(m, K.ny), mm = bc_pm(m, K.ny);
        int kp = bc_mp(k, K.nz), km = bc_mm(k, K.nz);
        lgrad_at<Real, 3>(K, L, l, m, k, X.G[P_C]);
        lgrad_at<Real, 3>(K, L, lp, m, k, X.G[P_XP]);
        lgrad_at<Real, 3>(K, L, lm, m, k, X.G[P_XM]);
        lgrad_at<Real, 3>(K, L, l, mp, k, X.G[P_YP]);
        lgrad_at<Real, 3>(K, L, l, mm, k, X.G[P_YM]);
        lgrad_at<Real, 3>(K, L, l, m, kp, X.G[P_ZP]);
        lgrad_at<Real, 3>(K, L, l, m, km, X.G[P_ZM]);
        Real c = L(l, m, k);
        Real xp = L(lp, m, k), xm = L(lm, m, k), yp = L(l, mp, k), ym = L(l, mm, k);
        Real zp = L(l, m, kp), zm = L(l, m, km);
        X.v[P_C] = c; X.v[P_XP] = xp; X.v[P_XM] = xm; X.v[P_YP] = yp; X.v[P_YM] = ym;
        X.v[P_ZP] = zp; X.v[P_ZM] = zm;
        // d2_phi_xyz, src/gpu_kernels_3d.py:207-248
        X.dxx = (xp - Real(2.0) * c + xm) * K.inv_dx2;
        X.dyy = (yp - Real(2.0) * c + ym) * K.inv_dx2;
        X.dzz = (zp - Real(2.0) * c + zm) * K.inv_dx2;
        X.dxy = (L(lp, mp, k) - L(lp, mm, k) - L(lm, mp, k) + L(lm, mm, k)) * K.q_inv_dx2;
        X.dxz = (L(lp, m, kp) - L(lp, m, km) - L(lm, m, kp) + L(lm, m, km)) * K.q_inv_dx2;
        X.dyz = (L(l, mp, kp) - L(l, mp, km) - L(l, mm, kp) + L(l, mm, km)) * K.q_inv_dx2;
    } else {
        int mp = bc_mp(m, K.ny), mm = bc_mm(m, K.ny);
        lgrad_at<Real, 2>(K, L, l, m, 0, X.G[P_C]);
        lgrad_at<Real, 2>(K, L, lp, m, 0, X.G[P_XP]);
        lgrad_at<Real, 2>(K, L, lm, m, 0, X.G[P_XM]);
        lgrad_at<Real, 2>(K, L, l, mp, 0, X.G[P_YP]);
        lgrad_at<Real, 2>(K, L, l, mm, 0, X.G[P_YM]);
        Real c = L(l, m, 0);
        Real xp = L(lp, m, 0), xm = L(lm, m, 0), yp = L(l, mp, 0), ym = L(l, mm, 0);
        X.v[P_C] = c; X.v[P_XP] = xp; X.v[P_XM] = xm; X.v[P_YP] = yp; X.v[P_YM] = ym;
        // d2_phi_xy, src/gpu_kernels.py:184-202
        X.dxx = (xp - Real(2.0) * c + xm) * K.inv_dx2;
        X.dyy = (yp - Real(2.0) * c + ym) * K.inv_dx2;
        X.dzz = Real(0.0);
        X.dxy = (L(lp, mp, 0) - L(lp, mm, 0) - L(lm, mp, 0) + L(lm, mm, 0)) * K.q_inv_dx2;
        X.dxz = X.dyz = Real(0.0);
    }
}

// lap_sl and b_sl of one active solid s (src/gpu_kernels_3d.py:558-577,
// :709-733; 2D src/gpu_kernels.py:394-408).  Requires K.aniso.
//
// The facet search runs once per stencil point; from it come eps^2 at the
// point (aniso_term1) and the one component of da/dphi that A13 needs there,
// so only two numbers per point stay live.  The full facet and A12 are kept
// for the centre (torque).
template <typename Real, int DIM>
MPF_DEV void solid_liquid_terms(const Consts<Real> &K, const LiqCtx<Real, DIM> &X, int s, Real &lap_sl,
                                Real &b_sl)
{
    const Real *nrow = K.n111() + (size_t)s * 24;
    Real nrm[12];
#pragma unroll
    for (int u = 0; u < 12; ++u) nrm[u] = nrow[u];
    const bool pairs = K.n111_pairs != 0;
    // One stencil point: facet search, eps^2 and A12 (written out per point so
    // that every access to the context is static).
    auto point = [&](const LGrad<Real> &G, Real &e2p, Facet<Real> &f, A12v<Real> &r) {
        f.cs = f.ca = f.ny = f.nz = Real(0.0);
        f.nx = Real(1.0);
        // eps2_at_cell_from_liquid_3d (:137-146): best_cos is 0 below the floor
        if (lgrad_ok(K, G))
            f = facet_search<Real, DIM>(nrm, nrow + 12, pairs, G.gx * G.inv_g, G.gy * G.inv_g, G.gz * G.inv_g);
        Real a = calc_a(K, f.ca);
        Real e = K.eps0_sl * a;
        e2p = e * e;
        if (K.torque)
            r = a12_at<Real, DIM>(K, G, f, a);
        else
            r.a = r.dax = r.day = r.daz = Real(0.0);
    };
    Facet<Real> fc, fo;
    A12v<Real> ac, ro;
    Real e2c, e2xp, e2xm, e2yp, e2ym, e2zp = Real(0.0), e2zm = Real(0.0);
    Real dxp, dxm, dyp, dym, dzp = Real(0.0), dzm = Real(0.0);
    point(X.G[P_C], e2c, fc, ac);
    point(X.G[P_XP], e2xp, fo, ro);
    dxp = ro.dax;
    point(X.G[P_XM], e2xm, fo, ro);
    dxm = ro.dax;
    point(X.G[P_YP], e2yp, fo, ro);
    dyp = ro.day;
    point(X.G[P_YM], e2ym, fo, ro);
    dym = ro.day;
    if constexpr (DIM == 3) {
        point(X.G[P_ZP], e2zp, fo, ro);
        dzp = ro.daz;
        point(X.G[P_ZM], e2zm, fo, ro);
        dzm = ro.daz;
    }
    // aniso_term1_solid_3d (:149-204)
    Real pc = X.v[P_C];
    Real Fx_p = m_div((e2c + e2xp) * (X.v[P_XP] - pc), K.two_dx);
    Real Fx_m = m_div((e2c + e2xm) * (pc - X.v[P_XM]), K.two_dx);
    Real Fy_p = m_div((e2c + e2yp) * (X.v[P_YP] - pc), K.two_dx);
    Real Fy_m = m_div((e2c + e2ym) * (pc - X.v[P_YM]), K.two_dx);
    Real lap_val;
    if constexpr (DIM == 3) {
        Real Fz_p = m_div((e2c + e2zp) * (X.v[P_ZP] - pc), K.two_dx);
        Real Fz_m = m_div((e2c + e2zm) * (pc - X.v[P_ZM]), K.two_dx);
        lap_val = m_div(Fx_p - Fx_m, K.dx) + m_div(Fy_p - Fy_m, K.dx) + m_div(Fz_p - Fz_m, K.dx);
    } else {
        lap_val = m_div(Fx_p - Fx_m, K.dx) + m_div(Fy_p - Fy_m, K.dx);
    }

    // torque_A11_3d (:386-441); 2D torque_A11 (src/gpu_kernels.py:290-332).
    // Below the gradient floor the torque is 0.0 and is still added.
    const LGrad<Real> &G = X.G[P_C];
    Real tq = Real(0.0);
    if (K.torque && lgrad_ok(K, G)) {
        Real phix = G.gx, phiy = G.gy, phiz = G.gz;
        Real q = G.q;
        Real q_x, q_y, q_z;
        if constexpr (DIM == 3) {
            q_x = Real(2.0) * (phix * X.dxx + phiy * X.dxy + phiz * X.dxz);
            q_y = Real(2.0) * (phix * X.dxy + phiy * X.dyy + phiz * X.dyz);
            q_z = Real(2.0) * (phix * X.dxz + phiy * X.dyz + phiz * X.dzz);
        } else {
            q_x = Real(2.0) * (phix * X.dxx + phiy * X.dxy);
            q_y = Real(2.0) * (phix * X.dxy + phiy * X.dyy);
            q_z = Real(0.0);
        }
        Real udotg;
        if constexpr (DIM == 3)
            udotg = phix * fc.nx + phiy * fc.ny + phiz * fc.nz;
        else
            udotg = phix * fc.nx + phiy * fc.ny;
        Real dcx = fc.nx * G.inv_g - udotg * phix * G.inv_g3;
        Real dcy = fc.ny * G.inv_g - udotg * phiy * G.inv_g3;
        Real c = fc.cs;
        Real C = m_sqrt(c * c + K.p2);
        Real S = m_sqrt(pymax(Real(1.0) - c * c, Real(0.0)) + K.p2);
        Real coef = m_div(c, C) - K.tan_a0 * m_div(c, S);
        Real a_val = ac.a;
        // A13_divergence_3d (:344-383)
        Real d_dx = (dxp - dxm) * K.half_over_dx;
        Real d_dy = (dyp - dym) * K.half_over_dx;
        if constexpr (DIM == 3) {
            Real dcz = fc.nz * G.inv_g - udotg * phiz * G.inv_g3;
            Real c_x = dcx * X.dxx + dcy * X.dxy + dcz * X.dxz;
            Real c_y = dcx * X.dxy + dcy * X.dyy + dcz * X.dyz;
            Real c_z = dcx * X.dxz + dcy * X.dyz + dcz * X.dzz;
            Real da_dx = K.mu_delta * coef * c_x;
            Real da_dy = K.mu_delta * coef * c_y;
            Real da_dz = K.mu_delta * coef * c_z;
            Real d_dz = (dzp - dzm) * K.half_over_dx;
            Real Ex = (da_dx * ac.dax * q) + (a_val * d_dx * q) + (a_val * ac.dax * q_x);
            Real Ey = (da_dy * ac.day * q) + (a_val * d_dy * q) + (a_val * ac.day * q_y);
            Real Ez = (da_dz * ac.daz * q) + (a_val * d_dz * q) + (a_val * ac.daz * q_z);
            tq = K.eps0_sq * (Ex + Ey + Ez);
        } else {
            Real c_x = dcx * X.dxx + dcy * X.dxy;
            Real c_y = dcx * X.dxy + dcy * X.dyy;
            Real da_dx = K.mu_delta * coef * c_x;
            Real da_dy = K.mu_delta * coef * c_y;
            Real Ex = (da_dx * ac.dax * q) + (a_val * d_dx * q) + (a_val * ac.dax * q_x);
            Real Ey = (da_dy * ac.day * q) + (a_val * d_dy * q) + (a_val * ac.day * q_y);
            tq = K.eps0_sq * (Ex + Ey);
        }
    }
    if (K.torque) lap_val += tq;
    lap_sl = lap_val;
    // b_sl: kinetic anisotropy from the liquid gradient at the centre
    b_sl = calc_b(K, lgrad_ok(K, G) ? fc.ca : Real(0.0));
}

// Active set of one cell in gathered form.  Index 0 is the liquid (present iff
// liq), indices 1..nsol are the active solids in ascending id order.
template <typename Real, int DIM, int A>
struct Cell {
    static constexpr int NP = (DIM == 3) ? 7 : 5;
    static constexpr int NA = A + 1;
    bool liq;
    int nsol;
    int gid[NA];       // gid[0] = 0
    Real v[NA][NP];    // values at the stencil points
    Real dom1[NP];     // 2D only: values of phase 1 (default dominant grain)
    Real out[NA];      // new values after clip + renormalisation
};

// The first NS solids of a cell held with room for A >= NS (the same values,
// for an update compiled for exactly NS solids).
template <int NS, typename Real, int DIM, int A>
MPF_DEV Cell<Real, DIM, NS> narrow_cell(const Cell<Real, DIM, A> &C)
{
    constexpr int NP = Cell<Real, DIM, A>::NP;
    Cell<Real, DIM, NS> D;
    D.liq = C.liq;
    D.nsol = C.nsol;
#pragma unroll
    for (int t = 0; t <= NS; ++t) {
        D.gid[t] = C.gid[t];
#pragma unroll
        for (int p = 0; p < NP; ++p) D.v[t][p] = C.v[t][p];
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) D.dom1[p] = C.dom1[p];
    return D;
}

// The update of one cell: src/gpu_kernels_3d.py:550-680 / :699-830 (3D) and
// src/gpu_kernels.py:386-515 (2D).  Tcell is the (already cooled) temperature.
// Cells with many active phases (A large) keep the arrays in scratch memory.
// ANISO_TERMS = false compiles the anisotropy terms out; the caller then
// guarantees the cell has no active solid-liquid pair or anisotropy is off.
// EXACT = true: the caller guarantees exactly A active solids and an active
// liquid (XLIQ) or an inactive one (!XLIQ), so every presence test folds at
// compile time.
// G > 1 (register path only): G consecutive lanes of the wavefront (G a
// power of two, G >= A) update the same cell together; lane r of the group
// evaluates the solid-liquid terms of solid r + 1 and the group exchanges
// them, which shortens the serial chain of a cell with several solids.
template <typename Real, int DIM, int A, bool ANISO_TERMS = true, int TC = 0, bool EXACT = false, int G = 1,
          bool XLIQ = true, typename LiqAt>
MPF_DEV void cell_update(const Consts<Real> &K, const LiqAt &L, int l, int m, int k,
                                            double Tcell, Cell<Real, DIM, A> &S)
{
    const bool s_liq = EXACT ? XLIQ : S.liq;
    const int s_nsol = EXACT ? A : S.nsol;
    constexpr int NA = A + 1;
    constexpr bool SMALL = (A <= 8);
    constexpr int UNR_NA = SMALL ? NA : 1;
    constexpr int UNR_A = SMALL ? A : 1;
    // Loops run over the fixed capacity on the register path (unrolled, with
    // a presence test) and only over the occupied entries otherwise.
    const int TMAX = [&] {
        if constexpr (SMALL)
            return NA;
        else
            return s_nsol + 1;
    }();
    const int n_act = (s_liq ? 1 : 0) + s_nsol;
    Real lap_sl[NA], b_sl[NA], lap[NA];
#pragma unroll UNR_NA
    for (int t = 0; t < TMAX; ++t) {
        lap_sl[t] = Real(0.0);
        b_sl[t] = Real(1.0);
    }

    // Laplacian of every active phase (laplacian_3d, :444-459 / inline :620-625)
#pragma unroll UNR_NA
    for (int t = 0; t < TMAX; ++t) {
        const Real *w = S.v[t];
        if constexpr (DIM == 3)
            lap[t] = (w[P_XP] + w[P_XM] + w[P_YP] + w[P_YM] + w[P_ZP] + w[P_ZM] - Real(6.0) * w[P_C]) *
                     K.inv_dx2;
        else
            lap[t] = (w[P_XP] + w[P_XM] + w[P_YP] + w[P_YM] - Real(4.0) * w[P_C]) * K.inv_dx2;
    }

    // Per-solid terms are read only in solid-liquid pairs, so they are needed
    // only when the liquid is active (the reference computes them regardless).
    Real w_sl = K.w0_sl;
    if (s_liq && s_nsol > 0) {
        // The dominant solid's gradient is taken first so that the neighbour
        // values of the other phases are dead before the liquid context.
        // dominant solid and anisotropic w_sl (:579-595; 2D :410-436)
        if constexpr (!ANISO_TERMS) {
        } else if constexpr (DIM == 3) {
            if (K.aniso) {
                int ts = -1;
                Real maxv = Real(-1.0);
#pragma unroll UNR_A
                for (int t = 1; t < NA; ++t) {
                    if (t > s_nsol) break;
                    if (S.v[t][P_C] > maxv) {
                        maxv = S.v[t][P_C];
                        ts = t;
                    }
                }
                if (ts >= 0) {
                    Real sx = Real(0.0), sy = Real(0.0), sz = Real(0.0);
                    int gs = 0;
#pragma unroll UNR_A
                    for (int t = 1; t < NA; ++t) {
                        if (t == ts) {
                            sx = m_div(S.v[t][P_XP] - S.v[t][P_XM], K.two_dx);
                            sy = m_div(S.v[t][P_YP] - S.v[t][P_YM], K.two_dx);
                            sz = m_div(S.v[t][P_ZP] - S.v[t][P_ZM], K.two_dx);
                            gs = S.gid[t];
                        }
                    }
                    Real g2 = sx * sx + sy * sy + sz * sz;
                    Real cmax = Real(0.0);
                    if (!(g2 < K.g2_floor)) {
                        Real inv_g = m_div(Real(1.0), m_sqrt(g2));
                        const Real *nr = K.n111() + (size_t)gs * 24;
                        Facet<Real> f = facet_search<Real, 3>(nr, nr + 12, K.n111_pairs != 0, sx * inv_g,
                                                              sy * inv_g, sz * inv_g);
                        cmax = f.ca;
                    }
                    Real a_loc = calc_a(K, cmax);
                    w_sl = K.w0_sl * (a_loc * a_loc);
                }
            }
        } else {
            int ts = -1;
            int gs = 1;
            Real maxv = Real(0.0);
#pragma unroll UNR_A
            for (int t = 1; t < NA; ++t) {
                if (t > s_nsol) break;
                if (S.v[t][P_C] > maxv) {
                    maxv = S.v[t][P_C];
                    ts = t;
                }
            }
            Real wxp = S.dom1[P_XP], wxm = S.dom1[P_XM], wyp = S.dom1[P_YP], wym = S.dom1[P_YM];
#pragma unroll UNR_A
            for (int t = 1; t < NA; ++t) {
                if (t == ts) {
                    wxp = S.v[t][P_XP]; wxm = S.v[t][P_XM]; wyp = S.v[t][P_YP]; wym = S.v[t][P_YM];
                    gs = S.gid[t];
                }
            }
            Real phix_ = (wxp - wxm) * K.half_over_dx;
            Real phiy_ = (wyp - wym) * K.half_over_dx;
            Real gn2 = phix_ * phix_ + phiy_ * phiy_;
            Real cmax = Real(0.0);
            if (gn2 >= K.g2_floor) {
                Real inv_gn = m_div(Real(1.0), m_sqrt(gn2));
                const Real *nr = K.n111() + (size_t)gs * 24;
                Facet<Real> f = facet_search<Real, 2>(nr, nr + 12, K.n111_pairs != 0, phix_ * inv_gn,
                                                      phiy_ * inv_gn, Real(0.0));
                cmax = f.ca;
            }
            Real a_loc = calc_a(K, cmax);
            w_sl = K.w0_sl * (a_loc * a_loc);
        }

        if (ANISO_TERMS && (K.aniso || DIM == 2)) {
            if constexpr (ANISO_TERMS) {
                LiqCtx<Real, DIM> X;
                build_liq_ctx<Real, DIM>(K, L, l, m, k, X);
                if constexpr (SMALL && G > 1) {
                    static_assert(G >= A && (G & (G - 1)) == 0 && G <= 64, "one lane per solid");
                    const int lane = lane_id();
                    const int r = lane & (G - 1);
                    int gr = 0;
#pragma unroll
                    for (int u = 0; u < A; ++u)
                        if (u == r) gr = S.gid[u + 1];
                    Real ls = Real(0.0), bs = Real(1.0);
                    if (r < s_nsol) solid_liquid_terms<Real, DIM>(K, X, gr, ls, bs);
#pragma unroll
                    for (int u = 0; u < A; ++u) {
                        const Real lu = lane_read(ls, (lane & ~(G - 1)) + u);
                        const Real bu = lane_read(bs, (lane & ~(G - 1)) + u);
                        if (u < s_nsol) {
                            lap_sl[u + 1] = lu;
                            b_sl[u + 1] = bu;
                        }
                    }
                } else if constexpr (SMALL) {
                    // The solids pass through a rotating register queue: A
                    // iterations (uniform trip count, body emitted once), each
                    // taking the head id and appending its result, so every
                    // array access is static.
                    int gq[A];
                    Real lq[A], bq[A];
#pragma unroll
                    for (int u = 0; u < A; ++u) {
                        gq[u] = S.gid[u + 1];
                        lq[u] = Real(0.0);
                        bq[u] = Real(1.0);
                    }
#pragma unroll 1
                    for (int it = 0; it < A; ++it) {
                        Real ls = Real(0.0), bs = Real(1.0);
                        if (it < s_nsol) solid_liquid_terms<Real, DIM>(K, X, gq[0], ls, bs);
#pragma unroll
                        for (int u = 0; u < A - 1; ++u) {
                            gq[u] = gq[u + 1];
                            lq[u] = lq[u + 1];
                            bq[u] = bq[u + 1];
                        }
                        gq[A - 1] = 0;
                        lq[A - 1] = ls;
                        bq[A - 1] = bs;
                    }
#pragma unroll
                    for (int u = 0; u < A; ++u) {
                        if (u < s_nsol) {
                            lap_sl[u + 1] = lq[u];
                            b_sl[u + 1] = bq[u];
                        }
                    }
                } else {
                    for (int t = 1; t <= s_nsol; ++t) {
                        Real ls, bs;
                        solid_liquid_terms<Real, DIM>(K, X, S.gid[t], ls, bs);
                        lap_sl[t] = ls;
                        b_sl[t] = bs;
                    }
                }
            }
        } else {
            // anisotropy off (switchable kernel, :729-731): eps0^2 * lap(phi_L)
#pragma unroll UNR_A
            for (int t = 1; t < NA; ++t) {
                if (t > s_nsol) break;
                lap_sl[t] = K.eps0_sq * lap[0];
                b_sl[t] = Real(1.0);
            }
        }
    }

    // Pairwise evolution (:600-661)
    const double dG = K.Sf * (Tcell - K.T_melt);
    const Real df_sl = Real(-dG), df_ls = Real(dG);
    const Real n_real = Real(n_act);
    const ClassVals<Real> cv = class_vals<TC>(K);
#pragma unroll UNR_NA
    for (int t1 = 0; t1 < TMAX; ++t1) {
        S.out[t1] = Real(0.0);
        const bool p1 = (t1 == 0) ? s_liq : (t1 <= s_nsol);
        if (!p1) continue;
        const int i = S.gid[t1];
        Real dpi = Real(0.0);
#pragma unroll UNR_NA
        for (int t2 = 0; t2 < TMAX; ++t2) {
            const bool p2 = (t2 == 0) ? s_liq : (t2 <= s_nsol);
            if (!p2 || t2 == t1) continue;
            const int j = S.gid[t2];
            Real driving_force = Real(0.0);
            if (i != 0 && j == 0)
                driving_force = df_sl;
            else if (i == 0 && j != 0)
                driving_force = df_ls;
            Real ppp = Real(0.0);
#pragma unroll UNR_NA
            for (int t3 = 0; t3 < TMAX; ++t3) {
                const bool p3 = (t3 == 0) ? s_liq : (t3 <= s_nsol);
                if (!p3) continue;
                const int pk = S.gid[t3];
                const Real phi_k = S.v[t3][P_C];
                Real wik = tab_at<TC, TAB_W>(K, cv, i, pk);
                Real wjk = tab_at<TC, TAB_W>(K, cv, j, pk);
                if (pk == 0) {
                    if (i != 0) wik = w_sl;
                    if (j != 0) wjk = w_sl;
                }
                Real term1 = (wik - wjk) * phi_k;
                Real term2 = Real(0.0);
                if (j == 0 && i != 0) {
                    if (pk == i) term2 = Real(0.5) * lap_sl[t1];
                } else if (i == 0 && j != 0) {
                    if (pk == j) term2 = Real(-0.5) * lap_sl[t2];
                } else {
                    Real aik = tab_at<TC, TAB_A>(K, cv, i, pk), ajk = tab_at<TC, TAB_A>(K, cv, j, pk);
                    Real eps2ik = aik * aik;
                    Real eps2jk = ajk * ajk;
                    term2 = Real(0.5) * (eps2ik - eps2jk) * lap[t3];
                }
                ppp += term1 + term2;
            }
            Real phii_phij = S.v[t1][P_C] * S.v[t2][P_C];
            Real term_force = Real(8.0 / 3.1415926535) * m_sqrt(pymax(phii_phij, Real(0.0))) * driving_force;
            const Real mij_tab = tab_at<TC, TAB_M>(K, cv, i, j);
            Real mij_eff = mij_tab;
            if (i != 0 && j == 0)
                mij_eff = mij_tab * b_sl[t1];
            else if (i == 0 && j != 0)
                mij_eff = mij_tab * b_sl[t2];
            dpi -= m_div(Real(2.0) * mij_eff, n_real) * (ppp - term_force);
        }
        S.out[t1] = S.v[t1][P_C] + K.dt * dpi;
    }

    // Clip and renormalise (:663-680)
    Real s = Real(0.0);
#pragma unroll UNR_NA
    for (int t = 0; t < TMAX; ++t) {
        const bool pr = (t == 0) ? s_liq : (t <= s_nsol);
        if (!pr) continue;
        Real v = S.out[t];
        if (v < Real(0.0)) v = Real(0.0);
        if (v > Real(1.0)) v = Real(1.0);
        S.out[t] = v;
        s += v;
    }
    if (s > Real(1e-20)) {
        Real invs = m_div(Real(1.0), s);
#pragma unroll UNR_NA
        for (int t = 0; t < TMAX; ++t) {
            const bool pr = (t == 0) ? s_liq : (t <= s_nsol);
            if (pr) S.out[t] *= invs;
        }
    } else {
        S.out[0] = Real(1.0);
    }
}

} // namespace mpf
