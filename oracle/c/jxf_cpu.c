/*
 * jxf_cpu.c -- C/OpenMP restatement of JAX-Fluids' convective-flux step (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the CPU baseline timed next to the GPU, and a second
 * CPU restatement the tests check against the NumPy oracle.  Nothing in the
 * product links it.
 *
 * Expressions follow the reference in its own evaluation order (citations are to
 * src/jaxfluids/...): WENO5 weno5_base.py:34-50, weno5_js.py:39-58,
 * weno5_z.py:152-166; characteristic projection eigendecomposition.py:139-231,
 * 425-431, 517-521; prim/cons equation_manager.py:93-101,164-171; flux
 * :237-252; HLLC HLLC.py:41-126; signal speeds signal_speeds.py; Rusanov
 * Rusanov.py:35-45; divergence space_solver.py:597-599; RK3 RK3.py:27-62;
 * halos outer/material.py:94-287, boundary_condition.py:564-595; CFL
 * time_step_size.py:103-109,154-155.
 *
 * Layout: (5, Nx+2nh, Ny+2nh, Nz+2nh), z fastest; inactive axes have extent 1.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int n[3];
    int nh;
    double dx[3];
} cgrid;

typedef struct {
    int var;      /* 0 PRIMITIVE 1 CONSERVATIVE 2 CHAR-PRIMITIVE */
    int stencil;  /* 0 JS 1 Z */
    int riemann;  /* 0 HLLC 1 RUSANOV */
    int ss;       /* 0 ARITHMETIC 1 RUSANOV 2 DAVIS 3 EINFELDT 4 TORO */
    double gamma;
    double eps;
} cscheme;

static const int MINOR[3][2] = {{2, 3}, {3, 1}, {1, 2}};

typedef struct {
    int B[3], off[3], act[3];
    int64_t s[3], fs;
} cgeo;

static void geo(const cgrid* g, cgeo* G) {
    for (int i = 0; i < 3; ++i) {
        G->act[i] = g->n[i] > 1;
        G->off[i] = G->act[i] ? g->nh : 0;
        G->B[i] = G->act[i] ? g->n[i] + 2 * g->nh : 1;
    }
    G->s[2] = 1;
    G->s[1] = G->B[2];
    G->s[0] = (int64_t)G->B[1] * G->B[2];
    G->fs = (int64_t)G->B[0] * G->B[1] * G->B[2];
}

static double weno(int z, double a, double b, double c, double d, double e, double eps) {
    double b0 = 13.0 / 12.0 * ((a - 2 * b + c) * (a - 2 * b + c)) + 1.0 / 4.0 * ((a - 4 * b + 3 * c) * (a - 4 * b + 3 * c));
    double b1 = 13.0 / 12.0 * ((b - 2 * c + d) * (b - 2 * c + d)) + 1.0 / 4.0 * ((b - d) * (b - d));
    double b2 = 13.0 / 12.0 * ((c - 2 * d + e) * (c - 2 * d + e)) + 1.0 / 4.0 * ((3 * c - 4 * d + e) * (3 * c - 4 * d + e));
    double a0, a1, a2;
    if (!z) {
        a0 = 0.1 * (1.0 / (b0 * b0 + eps));
        a1 = 0.6 * (1.0 / (b1 * b1 + eps));
        a2 = 0.3 * (1.0 / (b2 * b2 + eps));
    } else {
        double tau = fabs(b0 - b2);
        a0 = 0.1 * (1.0 + tau / (b0 + eps));
        a1 = 0.6 * (1.0 + tau / (b1 + eps));
        a2 = 0.3 * (1.0 + tau / (b2 + eps));
    }
    double one = 1.0 / (a0 + a1 + a2);
    double w0 = a0 * one, w1 = a1 * one, w2 = a2 * one;
    double p0 = (1.0 / 3.0) * a + (-7.0 / 6.0) * b + (11.0 / 6.0) * c;
    double p1 = (-1.0 / 6.0) * b + (5.0 / 6.0) * c + (1.0 / 3.0) * d;
    double p2 = (1.0 / 3.0) * c + (5.0 / 6.0) * d + (-1.0 / 6.0) * e;
    return w0 * p0 + w1 * p1 + w2 * p2;
}

static void cons(const double* W, double g, double* U) {
    double rho = W[0];
    double e = W[4] / (rho * (g - 1.0));
    U[0] = rho;
    U[1] = rho * W[1];
    U[2] = rho * W[2];
    U[3] = rho * W[3];
    U[4] = rho * (0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]) + e);
}

static void prim(const double* U, double g, double* W) {
    double rho = U[0], one = 1.0 / rho;
    double u = U[1] * one, v = U[2] * one, w = U[3] * one;
    double e = U[4] * one - 0.5 * (u * u + v * v + w * w);
    W[0] = rho;
    W[1] = u;
    W[2] = v;
    W[3] = w;
    W[4] = (g - 1.0) * e * rho;
}

static void pflux(const double* W, const double* U, int ax, double* F) {
    int v = ax + 1;
    F[0] = U[v];
    F[1] = U[v] * W[1];
    F[2] = U[v] * W[2];
    F[3] = U[v] * W[3];
    F[v] += W[4];
    F[4] = W[v] * (U[4] + W[4]);
}

static void speeds(int ss, double uL, double uR, double aL, double aR, double rL, double rR, double pL, double pR,
                   double g, double* SL, double* SR) {
    if (ss == 0) {
        double um = 0.5 * (uL + uR), am = 0.5 * (aL + aR);
        *SL = fmin(um - am, uL - aL);
        *SR = fmax(um + am, uR + aR);
    } else if (ss == 1) {
        double sp = fmax(fabs(uL) + aL, fabs(uR) + aR);
        *SL = -sp;
        *SR = sp;
    } else if (ss == 2) {
        *SL = fmin(uL - aL, uR - aR);
        *SR = fmax(uL + aL, uR + aR);
    } else if (ss == 4) {
        double rb = 0.5 * (rL + rR), ab = 0.5 * (aL + aR);
        double ps = fmax(0.0, 0.5 * (pL + pR) - 0.5 * (uR - uL) * rb * ab);
        double gg = (g + 1) * 0.5 / g;
        double qL = ps <= pL ? 1.0 : sqrt(1 + gg * (ps / pL - 1));
        double qR = ps <= pR ? 1.0 : sqrt(1 + gg * (ps / pR - 1));
        *SL = uL - aL * qL;
        *SR = uR + aR * qR;
    } else {
        double sL = sqrt(rL), sR = sqrt(rR);
        double od = 1.0 / (sL + sR);
        double eta2 = 0.5 * sL * sR * od * od;
        double ub = (sL * uL + sR * uR) * od;
        double db = sqrt((sL * aL * aL + sR * aR * aR) * od + eta2 * ((uR - uL) * (uR - uL)));
        *SL = fmin(ub - db, uL - aL);
        *SR = fmax(ub + db, uR + aR);
    }
}

static void star(const double* W, const double* U, double SK, double Ss, int ax, int left, double* F) {
    int v = ax + 1, m0 = MINOR[ax][0], m1 = MINOR[ax][1];
    double pre = (SK - W[v]) / (SK - Ss) * W[0];
    double us[5];
    us[0] = us[1] = us[2] = us[3] = pre;
    us[4] = pre * (U[4] / U[0] + (Ss - W[v]) * (Ss + W[4] / W[0] / (SK - W[v])));
    us[v] *= Ss;
    us[m0] *= W[m0];
    us[m1] *= W[m1];
    pflux(W, U, ax, F);
    double S = left ? fmin(SK, 0.0) : fmax(SK, 0.0);
    for (int i = 0; i < 5; ++i) F[i] = F[i] + S * (us[i] - U[i]);
}

static void riemann(const cscheme* sc, const double* WL, const double* WR, const double* UL, const double* UR, int ax,
                    double* F) {
    int v = ax + 1;
    double g = sc->gamma;
    double cL = sqrt(g * WL[4] / WL[0]), cR = sqrt(g * WR[4] / WR[0]);
    if (sc->riemann == 1) {
        double al = fmax(fabs(WL[v]) + cL, fabs(WR[v]) + cR);
        double FL[5], FR[5];
        pflux(WL, UL, ax, FL);
        pflux(WR, UR, ax, FR);
        for (int i = 0; i < 5; ++i) F[i] = 0.5 * (FL[i] + FR[i]) - 0.5 * al * (UR[i] - UL[i]);
        return;
    }
    double SL, SR;
    speeds(sc->ss, WL[v], WR[v], cL, cR, WL[0], WR[0], WL[4], WR[4], g, &SL, &SR);
    double rdL = WL[0] * (SL - WL[v]), rdR = WR[0] * (SR - WR[v]);
    double Ss = ((WR[4] - WL[4]) + (WL[v] * rdL - WR[v] * rdR)) / (rdL - rdR);
    double FsL[5], FsR[5];
    star(WL, UL, SL, Ss, ax, 1, FsL);
    star(WR, UR, SR, Ss, ax, 0, FsR);
    double sg = (Ss > 0) - (Ss < 0);
    for (int i = 0; i < 5; ++i) F[i] = 0.5 * (1 + sg) * FsL[i] + 0.5 * (1 - sg) * FsR[i];
}

/* one face: window cells -3..2 along ax at buffer index b (cell right of the face) */
static void face_flux(const cscheme* sc, const double* W, const double* U, const cgeo* G, int64_t b, int ax,
                      double* F) {
    double win[6][5], WL[5], WR[5], UL[5], UR[5];
    int64_t sa = G->s[ax];
    const double* src = sc->var == 1 ? U : W;
    for (int k = 0; k < 6; ++k)
        for (int i = 0; i < 5; ++i) win[k][i] = src[i * G->fs + b + (k - 3) * sa];
    double eps = sc->eps, g = sc->gamma;
    int z = sc->stencil;
    if (sc->var == 0 || sc->var == 1) {
        double* QL = sc->var == 0 ? WL : UL;
        double* QR = sc->var == 0 ? WR : UR;
        for (int i = 0; i < 5; ++i) {
            QL[i] = weno(z, win[0][i], win[1][i], win[2][i], win[3][i], win[4][i], eps);
            QR[i] = weno(z, win[5][i], win[4][i], win[3][i], win[2][i], win[1][i], eps);
        }
        if (sc->var == 0) {
            cons(WL, g, UL);
            cons(WR, g, UR);
        } else {
            prim(UL, g, WL);
            prim(UR, g, WR);
        }
    } else {
        int v = ax + 1, m0 = MINOR[ax][0], m1 = MINOR[ax][1];
        double r = 0.5 * (win[2][0] + win[3][0]);
        double p = 0.5 * (win[2][4] + win[3][4]);
        double c = sqrt(g * p / r), cc = c * c;
        double ch[6][5], cL[5], cR[5];
        for (int k = 0; k < 6; ++k) {
            ch[k][0] = -0.5 / c * win[k][v] + 0.5 / (cc * r) * win[k][4];
            ch[k][1] = win[k][0] - 1.0 / cc * win[k][4];
            ch[k][2] = win[k][m0];
            ch[k][3] = win[k][m1];
            ch[k][4] = 0.5 / c * win[k][v] + 0.5 / (cc * r) * win[k][4];
        }
        for (int i = 0; i < 5; ++i) {
            cL[i] = weno(z, ch[0][i], ch[1][i], ch[2][i], ch[3][i], ch[4][i], eps);
            cR[i] = weno(z, ch[5][i], ch[4][i], ch[3][i], ch[2][i], ch[1][i], eps);
        }
        WL[0] = r * (cL[0] + cL[4]) + cL[1];
        WL[v] = c * (-cL[0] + cL[4]);
        WL[m0] = cL[2];
        WL[m1] = cL[3];
        WL[4] = cc * r * (cL[0] + cL[4]);
        WR[0] = r * (cR[0] + cR[4]) + cR[1];
        WR[v] = c * (-cR[0] + cR[4]);
        WR[m0] = cR[2];
        WR[m1] = cR[3];
        WR[4] = cc * r * (cR[0] + cR[4]);
        cons(WL, g, UL);
        cons(WR, g, UR);
    }
    riemann(sc, WL, WR, UL, UR, ax, F);
}

/* rhs (5, Nx, Ny, Nz) = sum over active axes of (1/dx)(F_lo - F_hi); flux scratch (5, faces) */
void jxf_cpu_compute_rhs(const double* W, const double* U, const cgrid* g, const cscheme* sc, double* rhs,
                         double* flux) {
    cgeo G;
    geo(g, &G);
    const int n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
    const int64_t ncell = (int64_t)n0 * n1 * n2;
    int first = 1;
    for (int ax = 0; ax < 3; ++ax) {
        if (!G.act[ax]) continue;
        int fe[3] = {n0, n1, n2};
        fe[ax] += 1;
        const int64_t nf = (int64_t)fe[0] * fe[1] * fe[2];
#pragma omp parallel for collapse(2) schedule(static)
        for (int f0 = 0; f0 < fe[0]; ++f0)
            for (int f1 = 0; f1 < fe[1]; ++f1)
                for (int f2 = 0; f2 < fe[2]; ++f2) {
                    int64_t b = (int64_t)(f0 + G.off[0]) * G.s[0] + (int64_t)(f1 + G.off[1]) * G.s[1] + f2 + G.off[2];
                    double F[5];
                    face_flux(sc, W, U, &G, b, ax, F);
                    int64_t o = ((int64_t)f0 * fe[1] + f1) * fe[2] + f2;
                    for (int i = 0; i < 5; ++i) flux[i * nf + o] = F[i];
                }
        const int64_t st = ax == 0 ? (int64_t)fe[1] * fe[2] : (ax == 1 ? fe[2] : 1);
        const double odx = 1.0 / g->dx[ax];
#pragma omp parallel for collapse(2) schedule(static)
        for (int c0 = 0; c0 < n0; ++c0)
            for (int c1 = 0; c1 < n1; ++c1)
                for (int c2 = 0; c2 < n2; ++c2) {
                    int64_t fo = ((int64_t)c0 * fe[1] + c1) * fe[2] + c2;
                    int64_t r = ((int64_t)c0 * n1 + c1) * n2 + c2;
                    for (int i = 0; i < 5; ++i) {
                        double d = odx * (flux[i * nf + fo] - flux[i * nf + fo + st]);
                        rhs[i * ncell + r] = first ? 0.0 + d : rhs[i * ncell + r] + d;
                    }
                }
        first = 0;
    }
}

static void halo(double* W, double* U, const cgrid* g, const int bc[6], double gamma) {
    cgeo G;
    geo(g, &G);
    for (int loc = 0; loc < 6; ++loc) {
        int ax = loc / 2, high = loc % 2 == 0;
        if (!G.act[ax]) continue;
        int N = g->n[ax], nh = g->nh;
        int e[3] = {g->n[0], g->n[1], g->n[2]};
        e[ax] = nh;
#pragma omp parallel for collapse(2) schedule(static)
        for (int i0 = 0; i0 < e[0]; ++i0)
            for (int i1 = 0; i1 < e[1]; ++i1)
                for (int i2 = 0; i2 < e[2]; ++i2) {
                    int c[3] = {i0, i1, i2};
                    int m = c[ax], dst, src;
                    if (high) {
                        dst = N + nh + m;
                        src = bc[loc] == 0 ? nh + m : (bc[loc] == 1 ? N + nh - 1 : N + nh - 1 - m);
                    } else {
                        dst = m;
                        src = bc[loc] == 0 ? N + m : (bc[loc] == 1 ? nh : 2 * nh - 1 - m);
                    }
                    int64_t bd = 0, bs = 0;
                    for (int a = 0; a < 3; ++a) {
                        bd += (int64_t)(a == ax ? dst : c[a] + G.off[a]) * G.s[a];
                        bs += (int64_t)(a == ax ? src : c[a] + G.off[a]) * G.s[a];
                    }
                    double w[5], u[5];
                    for (int i = 0; i < 5; ++i) w[i] = W[i * G.fs + bs];
                    if (bc[loc] == 2) w[ax + 1] = -w[ax + 1];
                    cons(w, gamma, u);
                    for (int i = 0; i < 5; ++i) {
                        W[i * G.fs + bd] = w[i];
                        U[i * G.fs + bd] = u[i];
                    }
                }
    }
}

/* One TVD-RK3 step in the reference order.  Un, rhs, flux are caller scratch:
 * Un like U, rhs (5, Nx*Ny*Nz), flux (5, max faces). */
void jxf_cpu_rk3_step(double* W, double* U, double* Un, double* rhs, double* flux, const cgrid* g, const cscheme* sc,
                      const int bc[6], double dt) {
    cgeo G;
    geo(g, &G);
    const int64_t tot = 5 * G.fs;
    memcpy(Un, U, sizeof(double) * tot);
    static const double M[3] = {1.0, 0.25, 2.0 / 3.0};
    static const double C[2][2] = {{0.25, 0.75}, {2.0 / 3.0, 1.0 / 3.0}};
    const int n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
    const int64_t ncell = (int64_t)n0 * n1 * n2;
    for (int stage = 0; stage < 3; ++stage) {
        jxf_cpu_compute_rhs(W, U, g, sc, rhs, flux);
        const double ts = dt * M[stage];
#pragma omp parallel for collapse(2) schedule(static)
        for (int c0 = 0; c0 < n0; ++c0)
            for (int c1 = 0; c1 < n1; ++c1)
                for (int c2 = 0; c2 < n2; ++c2) {
                    int64_t b = (int64_t)(c0 + G.off[0]) * G.s[0] + (int64_t)(c1 + G.off[1]) * G.s[1] + c2 + G.off[2];
                    int64_t r = ((int64_t)c0 * n1 + c1) * n2 + c2;
                    double u[5], w[5];
                    for (int i = 0; i < 5; ++i) {
                        double x = U[i * G.fs + b];
                        if (stage > 0) x = C[stage - 1][0] * x + C[stage - 1][1] * Un[i * G.fs + b];
                        u[i] = x + ts * rhs[i * ncell + r];
                    }
                    prim(u, sc->gamma, w);
                    for (int i = 0; i < 5; ++i) {
                        U[i * G.fs + b] = u[i];
                        W[i * G.fs + b] = w[i];
                    }
                }
        halo(W, U, g, bc, sc->gamma);
    }
}

double jxf_cpu_compute_dt(const double* W, const cgrid* g, double gamma, double cfl, double eps) {
    cgeo G;
    geo(g, &G);
    double mx = 0.0;
    double dxmin = 1e300;
    for (int a = 0; a < 3; ++a)
        if (G.act[a] && g->dx[a] < dxmin) dxmin = g->dx[a];
#pragma omp parallel for collapse(2) reduction(max : mx) schedule(static)
    for (int c0 = 0; c0 < g->n[0]; ++c0)
        for (int c1 = 0; c1 < g->n[1]; ++c1)
            for (int c2 = 0; c2 < g->n[2]; ++c2) {
                int64_t b = (int64_t)(c0 + G.off[0]) * G.s[0] + (int64_t)(c1 + G.off[1]) * G.s[1] + c2 + G.off[2];
                double w[5];
                for (int i = 0; i < 5; ++i) w[i] = W[i * G.fs + b];
                double c = sqrt(gamma * w[4] / w[0]);
                double s = 0.0;
                for (int a = 0; a < 3; ++a)
                    if (G.act[a]) s = s + (fabs(w[a + 1]) + c);
                if (s > mx) mx = s;
            }
    return dxmin / (mx + eps) * cfl;
}

int jxf_cpu_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
