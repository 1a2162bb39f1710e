// jfa_dissipative.h -- second-order central viscous and heat fluxes at one face, added to the
// convective face flux inside the fused sweeps (solvers/source_term_solver.py:188-350,414-448,
// 504-580; flux assembly space_solver.py:565-587).
//
// Face f along AXIS lies between the cells L = f-1 and R = f:
//   T_f  = (T_L + T_R)/2,  T = p/(rho R)                          (central_2, ideal_gas.py:81-83)
//   mu   = mu(T_f) (POWER or SUTHERLAND law),  mu2 = mu_b - 2/3 mu   (material.py:82-101, :527)
//   g[k][i] = d u_k / d x_i at the face: i == AXIS: (u_R - u_L)/dx (deriv_face_2);
//            i != AXIS: the central derivative (u(+e_i) - u(-e_i))/(2 dx_i) at L and at R
//            (deriv_center_2), averaged (central_2); inactive i: 0
//   tau_k = mu (g[AXIS][k] + g[k][AXIS]) (active k), tau_AXIS += mu2 sum_k g[k][k]
//   F[1+k] += -tau_k;  F[4] += -sum_k tau_k u_f,k + q,  q = -kappa(T_f) (T_R - T_L)/dx,
//   kappa = cp mu / Pr (PRANDTL) or its own POWER / SUTHERLAND law      (material.py:103-124)
// On periodic domains the accesses wrap onto the interior (the halos and edge halos there are
// periodic copies); otherwise (prm.diss_halos) they read the buffer's face and edge halos, which
// the stage's halo scatter and edge-halo fill wrote.
#pragma once

#include "jfa_common.h"

namespace jfa {

__device__ __forceinline__ double law_value(double T, const Law& l) {
    if (l.kind == JFA_LAW_SUTHERLAND) {  // material.py:94-96, power3_2 (math/power_functions.py:29-30)
        const double x = T * l.inv_T_ref;
        return l.ref * ((l.T_ref + l.param) / (T + l.param)) * (x * sqrt(x));
    }
    return l.param == 0.0 ? l.ref : l.ref * pow(T * l.inv_T_ref, l.param);
}

__device__ __forceinline__ double viscosity(double T, const Params& prm) { return law_value(T, prm.mu); }

// thermal conductivity at temperature T, where mu = viscosity(T)
__device__ __forceinline__ double conductivity(double T, double mu, const Params& prm) {
    return prm.kappa.kind == JFA_LAW_PRANDTL ? prm.kappa_fac * mu : law_value(T, prm.kappa);
}

// buffer index of cell c (interior coordinates, -1 and n are halo cells): wrapped onto [0, n) on
// the active axes of a periodic domain, the halo cell itself otherwise
__device__ __forceinline__ int64_t wrapped_index(const Geo& G, const int c[3], int halos) {
    int64_t b = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        int ci = c[i];
        if (G.active[i] && !halos) ci = ci < 0 ? ci + G.n[i] : (ci >= G.n[i] ? ci - G.n[i] : ci);
        b += (int64_t)(ci + G.off[i]) * G.s[i];
    }
    return b;
}

template <class T>
__device__ __forceinline__ void load_velocity(const T* __restrict__ W, int64_t fs, int64_t b, double u[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) u[k] = ld(W, (k + 1) * fs + b);
}

// Face state of the dissipative flux: the two cells' primitives, their temperatures, the face
// temperature (clamped: T_f <= 0 replaced by the precision eps), the viscosity there and, for
// viscous fluxes, the velocity-gradient tensor g[k][i] = d u_k / d x_i at the face.
struct DissState {
    double wl[5], wr[5];
    double TL, TR, Tf;
    bool clamped;
    double mu;
    double g[3][3];
};

// face whose right cell has interior coordinates cr (cr[AXIS] in 0..n, the left cell is cr - e_AXIS)
template <class T, int AXIS>
__device__ __forceinline__ void dissipative_state(const T* __restrict__ W, const Geo& G, const Params& prm,
                                                  const int cr[3], DissState& s) {
    const int64_t fs = G.fs;
    int cl[3] = {cr[0], cr[1], cr[2]};
    cl[AXIS] -= 1;
    const int64_t bl = wrapped_index(G, cl, prm.diss_halos), br = wrapped_index(G, cr, prm.diss_halos);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        s.wl[i] = ld(W, i * fs + bl);
        s.wr[i] = ld(W, i * fs + br);
    }
    s.TL = s.wl[4] / (s.wl[0] * prm.R);
    s.TR = s.wr[4] / (s.wr[0] * prm.R);
    s.Tf = 0.5 * (s.TL + s.TR);
    s.clamped = s.Tf <= 0.0;
    if (s.clamped) s.Tf = 2.220446049250313e-16;  // source_term_solver.py:301 (precision eps)
    s.mu = viscosity(s.Tf, prm);
    if (!(prm.dissip & 1)) return;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (!G.active[i]) {
#pragma unroll
            for (int k = 0; k < 3; ++k) s.g[k][i] = 0.0;
        } else if (i == AXIS) {
#pragma unroll
            for (int k = 0; k < 3; ++k) s.g[k][i] = (s.wr[1 + k] - s.wl[1 + k]) * G.one_dx[i];
        } else {
            const double h = 0.5 * G.one_dx[i];
            int c[3];
            double up[3], um[3], dL[3], dR[3];
            c[0] = cl[0]; c[1] = cl[1]; c[2] = cl[2];
            c[i] += 1;
            load_velocity(W, fs, wrapped_index(G, c, prm.diss_halos), up);
            c[i] -= 2;
            load_velocity(W, fs, wrapped_index(G, c, prm.diss_halos), um);
#pragma unroll
            for (int k = 0; k < 3; ++k) dL[k] = (up[k] - um[k]) * h;
            c[0] = cr[0]; c[1] = cr[1]; c[2] = cr[2];
            c[i] += 1;
            load_velocity(W, fs, wrapped_index(G, c, prm.diss_halos), up);
            c[i] -= 2;
            load_velocity(W, fs, wrapped_index(G, c, prm.diss_halos), um);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                dR[k] = (up[k] - um[k]) * h;
                s.g[k][i] = 0.5 * (dL[k] + dR[k]);
            }
        }
    }
}

// adds the dissipative flux of a face state to F
template <int AXIS>
__device__ __forceinline__ void dissipative_flux(const DissState& s, const Geo& G, const Params& prm, double F[5]) {
    double d4 = 0.0;
    if (prm.dissip & 1) {
        const double mu2 = prm.bulk - 2.0 / 3.0 * s.mu;
        double tau[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) tau[k] = G.active[k] ? s.mu * (s.g[AXIS][k] + s.g[k][AXIS]) : 0.0;
        double div = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (G.active[k]) div = div + s.g[k][k];
        tau[AXIS] = tau[AXIS] + mu2 * div;
        double vel_tau = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (G.active[k]) vel_tau = vel_tau + tau[k] * (0.5 * (s.wl[1 + k] + s.wr[1 + k]));
#pragma unroll
        for (int k = 0; k < 3; ++k) F[1 + k] = F[1 + k] + (-tau[k]);
        d4 = -vel_tau;
    }
    if (prm.dissip & 2) d4 = d4 + (-conductivity(s.Tf, s.mu, prm) * ((s.TR - s.TL) * G.one_dx[AXIS]));
    F[4] = F[4] + d4;
}

// adds the dissipative flux of the face whose right cell has interior coordinates cr to F
template <class T, int AXIS>
__device__ __forceinline__ void dissipative_face(const T* __restrict__ W, const Geo& G, const Params& prm,
                                                 const int cr[3], double F[5]) {
    DissState s;
    dissipative_state<T, AXIS>(W, G, prm, cr, s);
    dissipative_flux<AXIS>(s, G, prm, F);
}

// cell diffusivities for the time-step limits: nu = mu/rho, alpha = kappa/(rho cp)
__device__ __forceinline__ void diffusivities(const double w[5], const Params& prm, double& nu, double& alpha) {
    const double T = w[4] / (w[0] * prm.R);
    const double mu = viscosity(T, prm);
    nu = (prm.dissip & 1) ? mu / w[0] : 0.0;
    alpha = (prm.dissip & 2) ? conductivity(T, mu, prm) / w[0] * prm.inv_cp : 0.0;
}

}  // namespace jfa
