// mpf_consts.h -- device constants of the update from the C-ABI scalars.
// Shared by the host API and the CPU emulation tests.
#pragma once

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
    const R *t = (const R *)tab;
    size_t nn = (size_t)N * N;
    K.wij = t;
    K.aij = t + nn;
    K.mij = t + 2 * nn;
    K.n111 = t + 3 * nn;
    return K;
}


} // namespace mpf
