/*
 * CPU oracle, second restatement: scalar triple loops over the reference's
 * FDTD step.  TEST INFRASTRUCTURE ONLY (see oracle/reference_math.py for who
 * may call it).  It is written per cell from the index ranges of
 * phonomena/simulation/base_solver.py, independently of the NumPy
 * restatement, and the tests require the two to agree bit for bit:
 *
 *   stage 1  update_T        base_solver.py:329-372   (interior stresses)
 *   stage 2  apply_T_tfbc    base_solver.py:408-433   (k = 0 plane)
 *   stage 3  update_u        base_solver.py:441-463   (interior displacement)
 *   stage 4  apply_u_tfbc    base_solver.py:495-517   (k = 0 plane)
 *   stage 5  apply_u_abc     base_solver.py:539-554   (Mur, fixed face order)
 *   stage 6  time_step       base_solver.py:561-571   (copy rotation)
 *
 * Arrays use the reference shapes in C order (z fastest):
 *   ux (nx-1,ny,nz)  uy (nx,ny-1,nz)  uz (nx,ny,nz-1)
 *   T1..T3 (nx,ny,nz)  T4 (nx,ny-1,nz-1)  T5 (nx-1,ny,nz-1)  T6 (nx-1,ny-1,nz)
 * Material: ids[nx*ny*nz] indexes a table of 13 doubles per material:
 *   C00 C01 C02 C10 C11 C12 C20 C21 C22 C33 C44 C55 rho   (C = 6x6 Voigt).
 * Built with -ffp-contract=off; every expression keeps the reference's
 * left-to-right evaluation order.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define I3(a, b, c, NB, NC) ((((size_t)(a)) * (NB) + (b)) * (NC) + (c))

/* Built with -fopenmp (liboracle_brute_omp.so) the volume loops split over
   x planes; every cell is computed exactly as in the serial build (no
   reductions), so both give the same bits. */
#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#else
#define OMP_FOR
#endif

typedef struct {
    int nx, ny, nz;
    const double *fdx, *fdy, *fdz, *sdx, *sdy, *sdz;
    const uint8_t *ids;
    const double *mat;
    double D;          /* dt**2 as the reference computes it */
    const double *mur; /* 12: xmax(ux,uy,uz) y0(...) ymax(...) zmax(...) */
    const double *damp_x, *damp_y, *damp_z; /* per-axis sponge terms or NULL */
} prob_t;

static double cm(const prob_t *p, int i, int j, int k, int e) {
    return p->mat[13 * p->ids[I3(i, j, k, p->ny, p->nz)] + e];
}

static double rho(const prob_t *p, int i, int j, int k) { return cm(p, i, j, k, 12); }

static double leap(double u, double uo, double coef, double div, double s, int damped) {
    if (!damped) return 2 * u - uo + coef * div;
    return (2 * u - (1 - s) * uo + coef * div) / (1 + s);
}

static double damp(const prob_t *p, int i, int j, int k) {
    return p->damp_x[i] + p->damp_y[j] + p->damp_z[k];
}

/* One step on (u, u_new, u_old) with copy rotation.  w = source sample. */
static void step(const prob_t *p, double w,
                 double *ux, double *uy, double *uz,
                 double *uxn, double *uyn, double *uzn,
                 double *uxo, double *uyo, double *uzo,
                 double *T1, double *T2, double *T3, double *T4, double *T5, double *T6) {
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const int dmp = p->damp_x != NULL;
#define UX(a, b, c) ux[I3(a, b, c, ny, nz)]
#define UY(a, b, c) uy[I3(a, b, c, ny - 1, nz)]
#define UZ(a, b, c) uz[I3(a, b, c, ny, nz - 1)]
#define UXN(a, b, c) uxn[I3(a, b, c, ny, nz)]
#define UYN(a, b, c) uyn[I3(a, b, c, ny - 1, nz)]
#define UZN(a, b, c) uzn[I3(a, b, c, ny, nz - 1)]
#define UXO(a, b, c) uxo[I3(a, b, c, ny, nz)]
#define UYO(a, b, c) uyo[I3(a, b, c, ny - 1, nz)]
#define UZO(a, b, c) uzo[I3(a, b, c, ny, nz - 1)]
#define TN(T, a, b, c) T[I3(a, b, c, ny, nz)]
#define TT4(a, b, c) T4[I3(a, b, c, ny - 1, nz - 1)]
#define TT5(a, b, c) T5[I3(a, b, c, ny, nz - 1)]
#define TT6(a, b, c) T6[I3(a, b, c, ny - 1, nz)]

    /* source line (base_solver.py:251) */
    for (int j = 0; j < ny; j++) UZ(0, j, 0) = w;

    /* stage 1: interior stresses */
    OMP_FOR
    for (int i = 1; i <= nx - 2; i++)
        for (int j = 1; j <= ny - 2; j++)
            for (int k = 1; k <= nz - 2; k++) {
                double dx_ = UX(i, j, k) - UX(i - 1, j, k);
                double dy_ = UY(i, j, k) - UY(i, j - 1, k);
                double dz_ = UZ(i, j, k) - UZ(i, j, k - 1);
                double *Ts[3] = {T1, T2, T3};
                for (int a = 0; a < 3; a++) {
                    double t = cm(p, i, j, k, 3 * a + 0) * dx_ / p->sdx[i - 1];
                    t = t + cm(p, i, j, k, 3 * a + 1) * dy_ / p->sdy[j - 1];
                    t = t + cm(p, i, j, k, 3 * a + 2) * dz_ / p->sdz[k - 1];
                    TN(Ts[a], i, j, k) = t;
                }
            }
    OMP_FOR
    for (int i = 1; i <= nx - 2; i++)
        for (int j = 0; j <= ny - 2; j++)
            for (int k = 0; k <= nz - 2; k++)
                TT4(i, j, k) = cm(p, i, j + 1, k + 1, 9) *
                               ((UY(i, j, k + 1) - UY(i, j, k)) / p->fdz[k] +
                                (UZ(i, j + 1, k) - UZ(i, j, k)) / p->fdy[j]);
    OMP_FOR
    for (int i = 0; i <= nx - 2; i++)
        for (int j = 1; j <= ny - 2; j++)
            for (int k = 0; k <= nz - 2; k++)
                TT5(i, j, k) = cm(p, i + 1, j, k + 1, 10) *
                               ((UX(i, j, k + 1) - UX(i, j, k)) / p->fdz[k] +
                                (UZ(i + 1, j, k) - UZ(i, j, k)) / p->fdx[i]);
    OMP_FOR
    for (int i = 0; i <= nx - 2; i++)
        for (int j = 0; j <= ny - 2; j++)
            for (int k = 1; k <= nz - 2; k++)
                TT6(i, j, k) = cm(p, i + 1, j + 1, k, 11) *
                               ((UX(i, j + 1, k) - UX(i, j, k)) / p->fdy[j] +
                                (UY(i + 1, j, k) - UY(i, j, k)) / p->fdx[i]);

    /* stage 2: traction-free plane k = 0; first spacings only */
    for (int i = 1; i <= nx - 2; i++)
        for (int j = 1; j <= ny - 2; j++) {
            double dx_ = UX(i, j, 0) - UX(i - 1, j, 0);
            double dy_ = UY(i, j, 0) - UY(i, j - 1, 0);
            double dz_ = UZ(i, j, 0) - 0;
            double *Ts[2] = {T1, T2};
            for (int a = 0; a < 2; a++) {
                double t = cm(p, i, j, 0, 3 * a + 0) * dx_ / p->sdx[0];
                t = t + cm(p, i, j, 0, 3 * a + 1) * dy_ / p->sdy[0];
                t = t + cm(p, i, j, 0, 3 * a + 2) * dz_ / p->sdz[0];
                TN(Ts[a], i, j, 0) = t;
            }
            TN(T3, i, j, 0) = 0;
        }
    for (int i = 1; i <= nx - 2; i++)
        for (int j = 0; j <= ny - 2; j++)
            TT4(i, j, 0) = cm(p, i, j + 1, 0, 9) *
                           ((UY(i, j, 1) - UY(i, j, 0)) / p->fdy[0] +
                            (UZ(i, j + 1, 0) - UZ(i, j, 0)) / p->fdz[0]);
    for (int i = 0; i <= nx - 2; i++)
        for (int j = 1; j <= ny - 2; j++)
            TT5(i, j, 0) = cm(p, i + 1, j, 0, 10) *
                           ((UX(i, j, 1) - UX(i, j, 0)) / p->fdx[0] +
                            (UZ(i + 1, j, 0) - UZ(i, j, 0)) / p->fdz[0]);
    for (int i = 0; i <= nx - 2; i++)
        for (int j = 0; j <= ny - 2; j++)
            TT6(i, j, 0) = cm(p, i + 1, j + 1, 0, 11) *
                           ((UX(i, j + 1, 0) - UX(i, j, 0)) / p->fdx[0] +
                            (UY(i + 1, j, 0) - UY(i, j, 0)) / p->fdz[0]);

    /* stage 3: interior displacements */
    OMP_FOR
    for (int i = 0; i <= nx - 2; i++)
        for (int j = 1; j <= ny - 2; j++)
            for (int k = 1; k <= nz - 2; k++) {
                double div = (TN(T1, i + 1, j, k) - TN(T1, i, j, k)) / p->fdx[i];
                div = div + (TT6(i, j, k) - TT6(i, j - 1, k)) / p->sdy[j - 1];
                div = div + (TT5(i, j, k) - TT5(i, j, k - 1)) / p->sdz[k - 1];
                UXN(i, j, k) = leap(UX(i, j, k), UXO(i, j, k), p->D / rho(p, i + 1, j, k), div,
                                    dmp ? damp(p, i, j, k) : 0, dmp);
            }
    OMP_FOR
    for (int i = 1; i <= nx - 2; i++)
        for (int j = 0; j <= ny - 2; j++)
            for (int k = 1; k <= nz - 2; k++) {
                double div = (TT6(i, j, k) - TT6(i - 1, j, k)) / p->sdx[i - 1];
                div = div + (TN(T2, i, j + 1, k) - TN(T2, i, j, k)) / p->fdy[j];
                div = div + (TT4(i, j, k) - TT4(i, j, k - 1)) / p->sdz[k - 1];
                UYN(i, j, k) = leap(UY(i, j, k), UYO(i, j, k), p->D / rho(p, i, j + 1, k), div,
                                    dmp ? damp(p, i, j, k) : 0, dmp);
            }
    OMP_FOR
    for (int i = 1; i <= nx - 2; i++)
        for (int j = 1; j <= ny - 2; j++)
            for (int k = 0; k <= nz - 2; k++) {
                double div = (TT5(i, j, k) - TT5(i - 1, j, k)) / p->sdx[i - 1];
                div = div + (TT4(i, j, k) - TT4(i, j - 1, k)) / p->sdy[j - 1];
                div = div + (TN(T3, i, j, k + 1) - TN(T3, i, j, k)) / p->fdz[k];
                UZN(i, j, k) = leap(UZ(i, j, k), UZO(i, j, k), p->D / rho(p, i, j, k + 1), div,
                                    dmp ? damp(p, i, j, k) : 0, dmp);
            }

    /* stage 4: free-surface displacements (k = 0) */
    for (int i = 0; i <= nx - 2; i++)
        for (int j = 1; j <= ny - 2; j++) {
            double div = (TN(T1, i + 1, j, 0) - TN(T1, i, j, 0)) / p->fdx[0];
            div = div + (TT6(i, j, 0) - TT6(i, j - 1, 0)) / p->sdy[0];
            div = div + (TT5(i, j, 0) - 0) / p->sdz[0];
            UXN(i, j, 0) = leap(UX(i, j, 0), UXO(i, j, 0), p->D / rho(p, i + 1, j, 0), div,
                                dmp ? damp(p, i, j, 0) : 0, dmp);
        }
    for (int i = 1; i <= nx - 2; i++)
        for (int j = 0; j <= ny - 2; j++) {
            double div = (TT6(i, j, 0) - TT6(i - 1, j, 0)) / p->sdx[0];
            div = div + (TN(T2, i, j + 1, 0) - TN(T2, i, j, 0)) / p->fdy[0];
            div = div + (TT4(i, j, 0) - 0) / p->sdz[0];
            UYN(i, j, 0) = leap(UY(i, j, 0), UYO(i, j, 0), p->D / rho(p, i, j + 1, 0), div,
                                dmp ? damp(p, i, j, 0) : 0, dmp);
        }
    for (int i = 1; i <= nx - 2; i++)
        for (int j = 1; j <= ny - 2; j++) {
            /* reference precedence: (a/sdx0 + b/sdy0) + T3[..,1] - T3[..,0]/fdz0 */
            double div = (TT5(i, j, 0) - TT5(i - 1, j, 0)) / p->sdx[0];
            div = div + (TT4(i, j, 0) - TT4(i, j - 1, 0)) / p->sdy[0];
            div = div + TN(T3, i, j, 1);
            div = div - TN(T3, i, j, 0) / p->fdz[0];
            UZN(i, j, 0) = leap(UZ(i, j, 0), UZO(i, j, 0), p->D / rho(p, i, j, 0), div,
                                dmp ? damp(p, i, j, 0) : 0, dmp);
        }

    /* stage 5: Mur faces, x = max, y = 0, y = max, z = max in that order */
    const double *m = p->mur;
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) UXN(nx - 2, j, k) = UX(nx - 3, j, k) + m[0] * (UXN(nx - 3, j, k) - UX(nx - 2, j, k));
    for (int j = 0; j < ny - 1; j++)
        for (int k = 0; k < nz; k++) UYN(nx - 1, j, k) = UY(nx - 2, j, k) + m[1] * (UYN(nx - 2, j, k) - UY(nx - 1, j, k));
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz - 1; k++) UZN(nx - 1, j, k) = UZ(nx - 2, j, k) + m[2] * (UZN(nx - 2, j, k) - UZ(nx - 1, j, k));

    for (int i = 0; i < nx - 1; i++)
        for (int k = 0; k < nz; k++) UXN(i, 0, k) = UX(i, 1, k) + m[3] * (UXN(i, 1, k) - UX(i, 0, k));
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) UYN(i, 0, k) = UY(i, 1, k) + m[4] * (UYN(i, 1, k) - UY(i, 0, k));
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz - 1; k++) UZN(i, 0, k) = UZ(i, 1, k) + m[5] * (UZN(i, 1, k) - UZ(i, 0, k));

    for (int i = 0; i < nx - 1; i++)
        for (int k = 0; k < nz; k++) UXN(i, ny - 1, k) = UX(i, ny - 2, k) + m[6] * (UXN(i, ny - 2, k) - UX(i, ny - 1, k));
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) UYN(i, ny - 2, k) = UY(i, ny - 3, k) + m[7] * (UYN(i, ny - 3, k) - UY(i, ny - 2, k));
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz - 1; k++) UZN(i, ny - 1, k) = UZ(i, ny - 2, k) + m[8] * (UZN(i, ny - 2, k) - UZ(i, ny - 1, k));

    for (int i = 0; i < nx - 1; i++)
        for (int j = 0; j < ny; j++) UXN(i, j, nz - 1) = UX(i, j, nz - 2) + m[9] * (UXN(i, j, nz - 2) - UX(i, j, nz - 1));
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny - 1; j++) UYN(i, j, nz - 1) = UY(i, j, nz - 2) + m[10] * (UYN(i, j, nz - 2) - UY(i, j, nz - 1));
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) UZN(i, j, nz - 2) = UZ(i, j, nz - 3) + m[11] * (UZN(i, j, nz - 3) - UZ(i, j, nz - 2));

    /* stage 6: copy rotation (u_new keeps its values) */
    size_t sx = (size_t)(nx - 1) * ny * nz, sy = (size_t)nx * (ny - 1) * nz, sz = (size_t)nx * ny * (nz - 1);
    memcpy(uxo, ux, sx * 8);
    memcpy(uyo, uy, sy * 8);
    memcpy(uzo, uz, sz * 8);
    memcpy(ux, uxn, sx * 8);
    memcpy(uy, uyn, sy * 8);
    memcpy(uz, uzn, sz * 8);
}

/*
 * Run `steps` steps starting at step index t0.  State in/out: u, u_new, u_old
 * (reference shapes).  T arrays are internal and start at zero, as after
 * Grid.update (grid.py:105-110); they carry no state across steps.
 * probe (6 ints: i0 si ni j0 sj nj) may be NULL; frames receive uz[...,0]
 * after every step with (tt+1) % every == 0.  Returns frames written or -1.
 */
int oracle_run(int nx, int ny, int nz,
               const double *fdx, const double *fdy, const double *fdz,
               const double *sdx, const double *sdy, const double *sdz,
               const uint8_t *ids, const double *mat, double D, const double *mur,
               const double *damp_x, const double *damp_y, const double *damp_z,
               const double *wave, int t0, int steps,
               double *ux, double *uy, double *uz,
               double *uxn, double *uyn, double *uzn,
               double *uxo, double *uyo, double *uzo,
               const int *probe, int every, double *frames) {
    prob_t p = {nx, ny, nz, fdx, fdy, fdz, sdx, sdy, sdz, ids, mat, D, mur, damp_x, damp_y, damp_z};
    size_t nn = (size_t)nx * ny * nz;
    double *T = calloc(3 * nn + (size_t)nx * (ny - 1) * (nz - 1) + (size_t)(nx - 1) * ny * (nz - 1) +
                           (size_t)(nx - 1) * (ny - 1) * nz, sizeof(double));
    if (!T) return -1;
    double *T1 = T, *T2 = T1 + nn, *T3 = T2 + nn, *T4 = T3 + nn;
    double *T5 = T4 + (size_t)nx * (ny - 1) * (nz - 1);
    double *T6 = T5 + (size_t)(nx - 1) * ny * (nz - 1);
    int nf = 0;
    for (int s = 0; s < steps; s++) {
        int tt = t0 + s;
        step(&p, wave[tt], ux, uy, uz, uxn, uyn, uzn, uxo, uyo, uzo, T1, T2, T3, T4, T5, T6);
        if (probe && every > 0 && (tt + 1) % every == 0) {
            for (int a = 0; a < probe[2]; a++)
                for (int b = 0; b < probe[5]; b++)
                    frames[((size_t)nf * probe[2] + a) * probe[5] + b] =
                        uz[I3(probe[0] + probe[1] * a, probe[3] + probe[4] * b, 0, ny, nz - 1)];
            nf++;
        }
    }
    free(T);
    return nf;
}
