// mpf_consts.h -- device constants of the update from the C-ABI scalars.
// Shared by the host API and the CPU emulation tests.
#pragma once

#include <stdint.h>
#include <string.h>

#include <type_traits>

#include "mpf.h"
#include "mpf_cell.h"
#include "mpf_math.h"

namespace mpf {

// The reference computes these inline in double (2.0*dx, 1.0/(dx*dx),
// 0.25*inv_dx2, 0.5/dx, eps0*eps0, mu_a*delta_a, p_round**2, tan(a0)); forming
// them once the same way gives identical values.
template <typename R>
inline Consts<R> make_consts(const mpf_scalars &s, int nx, int ny, int nz, int N, int kmax, int aniso,
                             int torque, const void *tab)
{
    Consts<R> K;
    K.nx = nx;
    K.ny = ny;
    K.nz = nz;
    K.N = N;
    K.kmax = kmax;
    K.aniso = aniso;
    K.torque = torque;
    // Derived constants are formed exactly as the reference writes them
    // inline (2.0*dx, 1.0/(dx*dx), 0.25*inv_dx2, 0.5/dx, ...).
    double inv_dx2 = 1.0 / (s.dx * s.dx);
    K.dx = (R)s.dx;
    K.dt = (R)s.dt;
    K.two_dx = (R)(2.0 * s.dx);
    K.inv_dx2 = (R)inv_dx2;
    K.q_inv_dx2 = (R)(0.25 * inv_dx2);
    K.half_over_dx = (R)(0.5 / s.dx);
    K.eps0_sl = (R)s.eps0_sl;
    K.eps0_sq = (R)(s.eps0_sl * s.eps0_sl);
    K.w0_sl = (R)s.w0_sl;
    K.tan_a0 = (R)mpf_tan(s.a0);
    K.delta_a = (R)s.delta_a;
    K.mu_a = (R)s.mu_a;
    K.mu_delta = (R)(s.mu_a * s.delta_a);
    K.p2 = (R)(s.p_round * s.p_round);
    K.g2_floor = (R)s.g2_floor;
    K.ksi = (R)s.ksi;
    K.theta_c = (R)s.theta_c_rad;
    K.Sf = s.Sf;
    K.T_melt = s.T_melt;
    K.tab = (const R *)tab;
    K.nn = (uint32_t)N * (uint32_t)N;
    for (int t = 0; t < 3; ++t) K.cls[t][0] = K.cls[t][1] = R(0);
    K.n111_pairs = 0;
    return K;
}

// True when every grain's n111 rows (grains 1..N-1, [8][3] each, in the
// solver precision) come in opposite pairs: row 7 - t is row t with every
// sign bit flipped, as compute_rotated_n111 (src/orientation_utils.py:182-212)
// produces from (+-1,+-1,+-1)/sqrt(3) in its order.
template <typename R>
inline bool n111_pairs(const R *n111, int N)
{
    using U = typename std::conditional<sizeof(R) == 8, uint64_t, uint32_t>::type;
    const U sign = (U)1 << (8 * sizeof(R) - 1);
    for (int g = 1; g < N; ++g) {
        const R *row = n111 + (size_t)g * 24;
        for (int t = 0; t < 4; ++t)
            for (int c = 0; c < 3; ++c) {
                U a, b;
                memcpy(&a, &row[t * 3 + c], sizeof(R));
                memcpy(&b, &row[(7 - t) * 3 + c], sizeof(R));
                if ((a ^ sign) != b) return false;
            }
    }
    return true;
}

// True when the tables tab = {wij, aij, mij} ([N][N] each, in the solver
// precision) have the class form of build_interaction_matrices
// (src/seed_modes.py:333-380): +0 on the diagonal, one bit pattern for all
// solid-liquid pairs and one for all grain-grain pairs.  The class values go
// to cls[table][0] (solid-liquid) and cls[table][1] (grain-grain).
template <typename R>
inline bool class_tables(const R *tab, int N, R cls[3][2])
{
    const size_t nn = (size_t)N * N;
    const R zero = R(0);
    for (int t = 0; t < 3; ++t) {
        const R *a = tab + t * nn;
        bool have[2] = {false, false};
        cls[t][0] = cls[t][1] = zero;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                const R v = a[(size_t)i * N + j];
                if (i == j) {
                    if (memcmp(&v, &zero, sizeof(R)) != 0) return false;
                    continue;
                }
                const int c = (i == 0 || j == 0) ? 0 : 1;
                if (!have[c]) {
                    cls[t][c] = v;
                    have[c] = true;
                } else if (memcmp(&v, &cls[t][c], sizeof(R)) != 0) {
                    return false;
                }
            }
    }
    return true;
}


} // namespace mpf
