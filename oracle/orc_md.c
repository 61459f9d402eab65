/*
 * orc_md.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see orc_md.h).
 *
 * Plain C99, FP64, built with -ffp-contract=off so every expression rounds as written,
 * following the reference's Fortran operation order line by line where it matters for
 * bit-level agreement (neighbour predicates, table index selection, integrator).
 */
#include "orc_md.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MXG 16

/* ------------------------------------------------------------------------------------------ */
/* potentials                                                                                  */
/* ------------------------------------------------------------------------------------------ */

/* Benchmark W FS form (pair polynomial only).  Same branch structure as
 * FS_Ackland_WW.F90:23-91: V for r <= c, rho for r <= d. */
static const double WFS_C = 3.25, WFS_C0 = 47.1346, WFS_C1 = -33.7666, WFS_C2 = 6.2542;
static const double WFS_A = 1.8964, WFS_D = 4.4002;

void orc_pot_wfs_nn(const double* rp, double* potr, double* fpotr) {
    double r = *rp;
    if (r <= WFS_C) {
        double rc = r - WFS_C;
        double poly = WFS_C0 + WFS_C1 * r + WFS_C2 * r * r;
        double v = rc * rc * poly;
        *potr = 0.5 * v;
        *fpotr = -(2.0 * rc * poly + rc * rc * (WFS_C1 + 2.0 * WFS_C2 * r));
    } else {
        *potr = 0.0;
        *fpotr = 0.0;
    }
}

void orc_pot_wfs_ne(const double* rp, double* potb, double* fpotb) {
    double r = *rp;
    if (r <= WFS_D) {
        double rd = r - WFS_D;
        *potb = WFS_A * WFS_A * rd * rd;
        *fpotb = -(WFS_A * WFS_A * 2.0 * rd);
    } else {
        *potb = 0.0;
        *fpotb = 0.0;
    }
}

/* FS_Ackland_WW.F90:11-20 constants expressed in eV and Angstrom. */
static const double ACK_A = 1.896373, ACK_C = 3.25, ACK_C0 = 47.1346499, ACK_C1 = -33.7665655,
                    ACK_C2 = 6.2541999, ACK_D = 4.400224, ACK_B = 90.3, ACK_ALPHA = 1.2,
                    ACK_B0 = 2.7411;

void orc_pot_ackland_ww_nn(const double* rp, double* potr, double* fpotr) {
    double r = *rp, core = 0.0, dcore = 0.0;
    if (r < ACK_B0) {
        double b = ACK_B0 - r, e = exp(-ACK_ALPHA * r);
        core = ACK_B * b * b * b * e;
        dcore = -ACK_B * (3.0 * b * b * e + ACK_ALPHA * b * b * b * e);
    }
    if (r <= ACK_C) {
        double rc = r - ACK_C;
        double poly = ACK_C0 + ACK_C1 * r + ACK_C2 * r * r;
        *potr = 0.5 * (rc * rc * poly + core);
        *fpotr = -(2.0 * rc * poly + rc * rc * (ACK_C1 + 2.0 * ACK_C2 * r) + dcore);
    } else {
        *potr = 0.0;
        *fpotr = 0.0;
    }
}

void orc_pot_ackland_ww_ne(const double* rp, double* potb, double* fpotb) {
    double r = *rp;
    if (r <= ACK_D) {
        double rd = r - ACK_D;
        *potb = ACK_A * ACK_A * rd * rd;
        *fpotb = -(ACK_A * ACK_A * 2.0 * rd);
    } else {
        *potb = 0.0;
        *fpotb = 0.0;
    }
}

void orc_emb_fs(const double* rhop, double* f, double* df) {
    double rho = *rhop;
    if (rho > 0.0) {
        *f = -sqrt(rho);
        *df = -0.5 / sqrt(rho);
    } else {
        *f = 0.0;
        *df = 0.0;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* tables: MD_TypeDef_ForceTable.F90:591-594 (grid), 949-976 (pair), 1043-1048 (embedding)     */
/* ------------------------------------------------------------------------------------------ */

void orc_table_pairwise(int ntab, double rmax, orc_pairfn nn, orc_pairfn ne, double rhomult,
                        double* rhomx_inout, double* potr, double* fpotr, double* potb,
                        double* fpotb) {
    double csi = (double)ntab / sqrt(rmax);
    double csiv = 1.0 / csi;
    potr[0] = fpotr[0] = potb[0] = fpotb[0] = 0.0;
    potr[ntab + 1] = fpotr[ntab + 1] = potb[ntab + 1] = fpotb[ntab + 1] = 0.0;
    for (int i = 1; i <= ntab; ++i) {
        double t = (double)i * csiv;
        double r = t * t;
        double v, dv;
        nn(&r, &v, &dv);
        potr[i] = v * r;
        fpotr[i] = dv * r;
    }
    for (int i = 1; i <= ntab; ++i) {
        if (ne) {
            double t = (double)i * csiv;
            double r = t * t;
            ne(&r, &potb[i], &fpotb[i]);
            if (*rhomx_inout < potb[i] * rhomult) *rhomx_inout = potb[i] * rhomult;
        } else {
            potb[i] = 0.0;
            fpotb[i] = 0.0;
        }
    }
}

void orc_table_embd(int nembd, double rhod, orc_pairfn embd, double* fembd, double* dfembd) {
    fembd[0] = dfembd[0] = 0.0;
    fembd[nembd + 1] = dfembd[nembd + 1] = 0.0;
    for (int i = 1; i <= nembd; ++i) {
        double rho = (double)(i - 1) * rhod;
        embd(&rho, &fembd[i], &dfembd[i]);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* neighbour lists                                                                             */
/* ------------------------------------------------------------------------------------------ */

/* Half stencil MD_NeighborsList.F90:183-185 and full stencil :427-429 (entry 0 = own cell). */
static const int NIX[27] = {0, -1, -1, -1, 0, 0, -1, 1, -1, 0, 1, -1, 0, 1,
                            1, 1, 1, 0, 0, -1, 1, -1, 0, 1, -1, 0, 1};
static const int NIY[27] = {0, 0, -1, 1, 1, 0, 0, 0, -1, -1, -1, 1, 1, 1,
                            0, 1, -1, -1, 0, 0, 0, -1, -1, -1, 1, 1, 1};
static const int NIZ[27] = {0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                            0, 0, 0, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1};

static const double NL_EPS = 1.0e-4;

/* Shared body of Cal_NeighboreListC / Cal_NeighboreList2C. */
static int nlist_cells(int half, int n, const double* xp, const int* ityp, const int* statu,
                       const double lo[3], const double len[3], const int ifpd[3], int ngroup,
                       const double* nb_rm, int mxkvois, int* kvois, int* indi) {
    int nstencil = half ? 14 : 27;
    double rmax = 0.0, rcut2[MXG * MXG];
    int ncell[3];
    if (ngroup > MXG) return ORC_ERR_ARG;
    for (int a = 0; a < ngroup * ngroup; ++a) {
        rcut2[a] = nb_rm[a] * nb_rm[a];
        if (nb_rm[a] > rmax) rmax = nb_rm[a];
    }
    for (int k = 0; k < 3; ++k) {
        ncell[k] = (int)(len[k] / (1.0 * rmax) - NL_EPS);
        if (ncell[k] < 3) ncell[k] = 3;
    }
    int nc = ncell[0] * ncell[1] * ncell[2];
    int* head = (int*)malloc(sizeof(int) * nc);
    int* link = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    int* ident = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    double* xpt = (double*)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    int* ttyp = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    int rc = ORC_OK;
    for (int c = 0; c < nc; ++c) head[c] = -1;
    for (int i = 0; i < n; ++i) {
        link[i] = -1;
        kvois[i] = 0;
    }
    /* linked cells, :255-286 */
    for (int i = 0; i < n; ++i) {
        if ((statu[i] & ORC_STATU_OUTOFBOX) == ORC_STATU_OUTOFBOX) continue;
        int ixyz[3];
        for (int k = 0; k < 3; ++k) {
            double s = (xp[i + k * n] - lo[k]) / len[k] * (double)ncell[k] - NL_EPS;
            ixyz[k] = (int)s;
        }
        for (int k = 0; k < 3; ++k)
            /* the reference accepts ixyz == ncell (a slot past the row); that position is
             * outside the box, so it is reported here as out of box as well */
            if (ixyz[k] < 0 || ixyz[k] >= ncell[k]) rc = ORC_ERR_OUTOFBOX;
        if (rc != ORC_OK) goto done;
        int ic = ixyz[0] + ncell[0] * (ixyz[1] + ncell[1] * ixyz[2]);
        link[i] = head[ic];
        head[ic] = i;
    }
    /* scan cells, :295-382 */
    for (int iz = 0; iz < ncell[2]; ++iz)
        for (int iy = 0; iy < ncell[1]; ++iy)
            for (int ix = 0; ix < ncell[0]; ++ix) {
                int ic0 = ix + ncell[0] * (iy + ncell[1] * iz);
                int id = head[ic0];
                if (id < 0) continue;
                int nn = 0;
                while (id >= 0) {
                    ident[nn] = id;
                    for (int k = 0; k < 3; ++k) xpt[3 * nn + k] = xp[id + k * n];
                    ttyp[nn] = ityp[id];
                    ++nn;
                    id = link[id];
                }
                int n0 = nn;
                for (int s = 1; s < nstencil; ++s) {
                    int j3[3] = {ix + NIX[s], iy + NIY[s], iz + NIZ[s]};
                    double cx[3] = {0.0, 0.0, 0.0};
                    int skip = 0;
                    for (int k = 0; k < 3; ++k) {
                        if (ifpd[k]) {
                            if (j3[k] >= ncell[k]) {
                                j3[k] = 0;
                                cx[k] = len[k];
                            } else if (j3[k] < 0) {
                                j3[k] = ncell[k] - 1;
                                cx[k] = -len[k];
                            }
                        }
                        if (j3[k] < 0 || j3[k] >= ncell[k]) skip = 1;
                    }
                    if (skip) continue;
                    id = head[j3[0] + ncell[0] * (j3[1] + ncell[1] * j3[2])];
                    while (id >= 0) {
                        ident[nn] = id;
                        for (int k = 0; k < 3; ++k) xpt[3 * nn + k] = xp[id + k * n] + cx[k];
                        ttyp[nn] = ityp[id];
                        ++nn;
                        id = link[id];
                    }
                }
                for (int a = 0; a < n0; ++a) {
                    int ida = ident[a], cnt = 0;
                    int jstart = half ? a + 1 : 0;
                    for (int b = jstart; b < nn; ++b) {
                        if (!half && b == a) continue;
                        double dx = xpt[3 * a] - xpt[3 * b];
                        double dy = xpt[3 * a + 1] - xpt[3 * b + 1];
                        double dz = xpt[3 * a + 2] - xpt[3 * b + 2];
                        double r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 < rcut2[ttyp[a] * ngroup + ttyp[b]]) {
                            if (cnt >= mxkvois) {
                                rc = ORC_ERR_OVERFLOW;
                                goto done;
                            }
                            indi[ida + (size_t)cnt * n] = ident[b];
                            ++cnt;
                        }
                    }
                    kvois[ida] = cnt;
                }
            }
done:
    free(head);
    free(link);
    free(ident);
    free(xpt);
    free(ttyp);
    return rc;
}

int orc_nlist_half(int n, const double* xp, const int* ityp, const int* statu, const double lo[3],
                   const double len[3], const int ifpd[3], int ngroup, const double* nb_rm,
                   int mxkvois, int* kvois, int* indi) {
    return nlist_cells(1, n, xp, ityp, statu, lo, len, ifpd, ngroup, nb_rm, mxkvois, kvois, indi);
}

int orc_nlist_full(int n, const double* xp, const int* ityp, const int* statu, const double lo[3],
                   const double len[3], const int ifpd[3], int ngroup, const double* nb_rm,
                   int mxkvois, int* kvois, int* indi) {
    return nlist_cells(0, n, xp, ityp, statu, lo, len, ifpd, ngroup, nb_rm, mxkvois, kvois, indi);
}

static inline double min_image(double s, double box, double hbox, int pd) {
    if (pd && fabs(s) > hbox) s = s - copysign(box, s);
    return s;
}

int orc_nlist_brute(int n, const double* xp, const int* ityp, const int* statu,
                    const double len[3], const int ifpd[3], int ngroup, const double* nb_rm,
                    int mxkvois, int* kvois, int* indi) {
    double hbox[3] = {0.5 * len[0], 0.5 * len[1], 0.5 * len[2]};
    for (int i = 0; i < n; ++i) {
        int cnt = 0;
        kvois[i] = 0;
        if ((statu[i] & ORC_STATU_OUTOFBOX) == ORC_STATU_OUTOFBOX) continue;
        for (int j = 0; j < n; ++j) {
            if (j == i || (statu[j] & ORC_STATU_OUTOFBOX) == ORC_STATU_OUTOFBOX) continue;
            double d[3];
            for (int k = 0; k < 3; ++k)
                d[k] = min_image(xp[i + k * n] - xp[j + k * n], len[k], hbox[k], ifpd[k]);
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            double rm = nb_rm[ityp[i] * ngroup + ityp[j]];
            if (r2 < rm * rm) {
                if (cnt >= mxkvois) return ORC_ERR_OVERFLOW;
                indi[i + (size_t)cnt * n] = j;
                ++cnt;
            }
        }
        kvois[i] = cnt;
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* forces                                                                                      */
/* ------------------------------------------------------------------------------------------ */

#define TAB(arr, k, i) ((arr)[(size_t)(k) * (size_t)(t->ntab + 2) + (size_t)(i)])
#define ETAB(arr, k, i) ((arr)[(size_t)(k) * (size_t)(t->nembd + 2) + (size_t)(i)])

/* MD_FS_ForceTable.F90:646-669 */
static inline double diff_ffun_fs(double den) { return den > 0.0 ? -0.5 / sqrt(den) : 0.0; }

double orc_sum(const double* a, int n) {
    if (n <= 0) return 0.0;
    if (n <= 8) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += a[i];
        return s;
    }
    int h = n / 2;
    return orc_sum(a, h) + orc_sum(a + h, n - h);
}

int orc_force_fs_half(int n, const double* xp, const int* ityp, const int* kvois, const int* indi,
                      int mxkvois, const double len[3], const int ifpd[3], const double* ru,
                      const orc_ftable* t, double* fp, double* epot, double* den, double* dden,
                      double* vtensor, double* vabs) {
    (void)mxkvois;
    int ng = t->ngroup;
    double hbox[3] = {0.5 * len[0], 0.5 * len[1], 0.5 * len[2]};
    double rcut2[MXG * MXG];
    double* er = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    if (ng > MXG) {
        free(er);
        return ORC_ERR_ARG;
    }
    for (int a = 0; a < ng * ng; ++a) rcut2[a] = ru[a] * ru[a];
    for (int i = 0; i < n; ++i) den[i] = 0.0;
    /* preCALFOR, :188-273 */
    for (int i = 0; i < n; ++i) {
        int ti = ityp[i];
        for (int iw = 0; iw < kvois[i]; ++iw) {
            int j = indi[i + (size_t)iw * n];
            double sep[3];
            for (int k = 0; k < 3; ++k)
                sep[k] = min_image(xp[i + k * n] - xp[j + k * n], len[k], hbox[k], ifpd[k]);
            double r2 = sep[0] * sep[0] + sep[1] * sep[1] + sep[2] * sep[2];
            int tj = ityp[j];
            if (r2 <= rcut2[ti * ng + tj]) {
                int ktab = t->kpair[ti * ng + tj], ktab1 = t->kpair[tj * ng + ti];
                double r = sqrt(r2);
                double sk = sqrt(r) * t->csi;
                int kk = (int)sk, k1 = kk + 1;
                double dk = sk - (double)kk;
                double exp1 =
                    (TAB(t->potr, ktab, kk) + dk * (TAB(t->potr, ktab, k1) - TAB(t->potr, ktab, kk))) / r;
                er[i] += exp1;
                er[j] += exp1;
                den[i] += TAB(t->potb, ktab, kk) + dk * (TAB(t->potb, ktab, k1) - TAB(t->potb, ktab, kk));
                den[j] += TAB(t->potb, ktab1, kk) +
                          dk * (TAB(t->potb, ktab1, k1) - TAB(t->potb, ktab1, kk));
            }
        }
    }
    for (int i = 0; i < n; ++i) dden[i] = diff_ffun_fs(den[i]);
    /* CALFOR, :277-395 */
    for (int a = 0; a < 3 * n; ++a) fp[a] = 0.0;
    for (int a = 0; a < 9; ++a) vtensor[a] = 0.0;
    if (vabs)
        for (int a = 0; a < 9; ++a) vabs[a] = 0.0;
    for (int i = 0; i < n; ++i) {
        int ti = ityp[i];
        double denki = dden[i];
        for (int iw = 0; iw < kvois[i]; ++iw) {
            int j = indi[i + (size_t)iw * n];
            double sep[3];
            for (int k = 0; k < 3; ++k)
                sep[k] = min_image(xp[i + k * n] - xp[j + k * n], len[k], hbox[k], ifpd[k]);
            double r2 = sep[0] * sep[0] + sep[1] * sep[1] + sep[2] * sep[2];
            int tj = ityp[j];
            if (r2 <= rcut2[ti * ng + tj]) {
                int ktab = t->kpair[ti * ng + tj], ktab1 = t->kpair[tj * ng + ti];
                double r = sqrt(r2);
                double sk = sqrt(r) * t->csi;
                int kk = (int)sk, kk1 = kk + 1;
                double dk = sk - (double)kk;
                double fortot =
                    (TAB(t->fpotr, ktab, kk) + dk * (TAB(t->fpotr, ktab, kk1) - TAB(t->fpotr, ktab, kk))) / r2 +
                    ((TAB(t->fpotb, ktab1, kk) + dk * (TAB(t->fpotb, ktab1, kk1) - TAB(t->fpotb, ktab1, kk))) * dden[j] +
                     (TAB(t->fpotb, ktab, kk) + dk * (TAB(t->fpotb, ktab, kk1) - TAB(t->fpotb, ktab, kk))) * denki) /
                        r;
                for (int k = 0; k < 3; ++k) {
                    double f = fortot * sep[k];
                    fp[i + k * n] += f;
                    fp[j + k * n] -= f;
                }
                for (int k = 0; k < 3; ++k)
                    for (int k1 = 0; k1 < 3; ++k1) vtensor[k + 3 * k1] += sep[k1] * sep[k] * fortot;
                if (vabs)
                    for (int k = 0; k < 3; ++k)
                        for (int k1 = 0; k1 < 3; ++k1)
                            vabs[k + 3 * k1] += fabs(sep[k1] * sep[k] * fortot);
            }
        }
    }
    /* :181 EPOT = ER - sqrt(DEN) */
    for (int i = 0; i < n; ++i) epot[i] = er[i] - sqrt(den[i]);
    free(er);
    return ORC_OK;
}

/* CAL_FS/EAM_AtomicStress_KERNEL (MD_FS_ForceTable_GPU.F90:1746-1965, EAM twin): for every
 * active atom, AP(i, 3a+b) = sum over its full list of SEP_a SEP_b FORTOT (no factor 1/2), with
 * FORTOT = FPOTR/r^2 + (FPOTB(KTAB) F'_i + FPOTB(KTAB1) F'_j)/r and the single cutoff RU2; zero for
 * inactive atoms.  dden is F' as orc_force_full returns it.  ap is (N, 9) column-major; apabs (may
 * be NULL) receives the sums of the terms' absolute values. */
int orc_atomic_stress_full(int n, const double* xp, const int* ityp, const int* statu,
                           const int* kvois, const int* indi, const double len[3],
                           const int ifpd[3], double ru2max, const orc_ftable* t,
                           const double* dden, double* ap, double* apabs) {
    int ng = t->ngroup;
    double hbox[3] = {0.5 * len[0], 0.5 * len[1], 0.5 * len[2]};
    for (int i = 0; i < n; ++i) {
        double p[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, q[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if ((statu[i] & ORC_STATU_ACTIVE) == ORC_STATU_ACTIVE) {
            int ti = ityp[i];
            for (int iw = 0; iw < kvois[i]; ++iw) {
                int j = indi[i + (size_t)iw * n];
                double s[3];
                for (int k = 0; k < 3; ++k)
                    s[k] = min_image(xp[i + k * n] - xp[j + k * n], len[k], hbox[k], ifpd[k]);
                double r2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
                if (r2 <= ru2max) {
                    int tj = ityp[j];
                    int ktab = t->kpair[ti * ng + tj], ktab1 = t->kpair[tj * ng + ti];
                    double r = sqrt(r2), sk = sqrt(r) * t->csi;
                    int kk = (int)sk;
                    double dk = sk - kk;
                    double fortot =
                        (TAB(t->fpotr, ktab, kk) + dk * (TAB(t->fpotr, ktab, kk + 1) - TAB(t->fpotr, ktab, kk))) / r2 +
                        ((TAB(t->fpotb, ktab, kk) + dk * (TAB(t->fpotb, ktab, kk + 1) - TAB(t->fpotb, ktab, kk))) * dden[i] +
                         (TAB(t->fpotb, ktab1, kk) + dk * (TAB(t->fpotb, ktab1, kk + 1) - TAB(t->fpotb, ktab1, kk))) * dden[j]) /
                            r;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            double v = s[a] * s[b] * fortot;
                            p[3 * a + b] = p[3 * a + b] + v;
                            q[3 * a + b] = q[3 * a + b] + fabs(v);
                        }
                }
            }
        }
        for (int c = 0; c < 9; ++c) {
            ap[i + (size_t)c * n] = p[c];
            if (apabs) apabs[i + (size_t)c * n] = q[c];
        }
    }
    return ORC_OK;
}

/* EAM embedding lookups, MD_EAM_ForceTable_GPU.F90:535-541 (F') and :1628-1631 (F). */
static inline double emb_lerp(const orc_ftable* t, const double* tab, int k, double rho, int* err) {
    double sk = rho / t->rhod + 1.0;
    int kk = (int)(sk + 0.000001);
    if (kk < 0) kk = 0;
    if (kk > t->nembd) {
        *err = 1;
        kk = t->nembd;
    }
    return ETAB(tab, k, kk) + (sk - (double)kk) * (ETAB(tab, k, kk + 1) - ETAB(tab, k, kk));
}

int orc_force_full(int n, const double* xp, const int* ityp, const int* statu, const int* kvois,
                   const int* indi, int mxkvois, const double len[3], const int ifpd[3],
                   double ru2max, const orc_ftable* t, double* fp, double* dden, double* epot) {
    (void)mxkvois;
    int ng = t->ngroup, err = 0;
    double hbox[3] = {0.5 * len[0], 0.5 * len[1], 0.5 * len[2]};
    /* PRECALFOR_KERNEL (FS :341-516, EAM :369-548) */
    for (int i = 0; i < n; ++i) {
        double den0 = 0.0;
        if ((statu[i] & ORC_STATU_ACTIVE) == ORC_STATU_ACTIVE) {
            int ti = ityp[i];
            for (int iw = 0; iw < kvois[i]; ++iw) {
                int j = indi[i + (size_t)iw * n];
                double s[3];
                for (int k = 0; k < 3; ++k)
                    s[k] = min_image(xp[i + k * n] - xp[j + k * n], len[k], hbox[k], ifpd[k]);
                double r2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
                if (r2 <= ru2max) {
                    int ktab = t->kpair[ti * ng + ityp[j]];
                    double r = sqrt(r2), sk = sqrt(r) * t->csi;
                    int kk = (int)sk;
                    den0 = den0 + (TAB(t->potb, ktab, kk) +
                                   (sk - kk) * (TAB(t->potb, ktab, kk + 1) - TAB(t->potb, ktab, kk)));
                }
            }
            /* a non-positive density is passed through unchanged (FS :497-508, EAM :535-541) */
            if (t->pot_type == 0) {
                if (den0 > 0.0) den0 = -0.5 / sqrt(den0);
            } else if (den0 > 0.0) {
                den0 = emb_lerp(t, t->dfembd, t->kembd[ti], den0, &err);
            }
        }
        dden[i] = den0;
    }
    /* CALFORCE_KERNEL (FS :623-806, EAM :646-829) */
    for (int i = 0; i < n; ++i) {
        double f[3] = {0.0, 0.0, 0.0};
        if ((statu[i] & ORC_STATU_ACTIVE) == ORC_STATU_ACTIVE) {
            int ti = ityp[i];
            double denki = dden[i];
            for (int iw = 0; iw < kvois[i]; ++iw) {
                int j = indi[i + (size_t)iw * n];
                double s[3];
                for (int k = 0; k < 3; ++k)
                    s[k] = min_image(xp[i + k * n] - xp[j + k * n], len[k], hbox[k], ifpd[k]);
                double r2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
                if (r2 <= ru2max) {
                    int tj = ityp[j];
                    int ktab = t->kpair[ti * ng + tj], ktab1 = t->kpair[tj * ng + ti];
                    double r = sqrt(r2), sk = sqrt(r) * t->csi;
                    int kk = (int)sk;
                    double dk = sk - kk;
                    double fortot =
                        (TAB(t->fpotr, ktab, kk) + dk * (TAB(t->fpotr, ktab, kk + 1) - TAB(t->fpotr, ktab, kk))) / r2 +
                        ((TAB(t->fpotb, ktab, kk) + dk * (TAB(t->fpotb, ktab, kk + 1) - TAB(t->fpotb, ktab, kk))) * denki +
                         (TAB(t->fpotb, ktab1, kk) + dk * (TAB(t->fpotb, ktab1, kk + 1) - TAB(t->fpotb, ktab1, kk))) * dden[j]) /
                            r;
                    for (int k = 0; k < 3; ++k) f[k] = f[k] + fortot * s[k];
                }
            }
        }
        for (int k = 0; k < 3; ++k) fp[i + k * n] = f[k];
    }
    /* CALEPOT_KERNEL (FS :1451-1609, EAM :1470-1636) */
    for (int i = 0; i < n; ++i) {
        double er0 = 0.0, den0 = 0.0;
        if ((statu[i] & ORC_STATU_ACTIVE) == ORC_STATU_ACTIVE) {
            int ti = ityp[i];
            for (int iw = 0; iw < kvois[i]; ++iw) {
                int j = indi[i + (size_t)iw * n];
                double s[3];
                for (int k = 0; k < 3; ++k)
                    s[k] = min_image(xp[i + k * n] - xp[j + k * n], len[k], hbox[k], ifpd[k]);
                double r2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
                if (r2 <= ru2max) {
                    int ktab = t->kpair[ti * ng + ityp[j]];
                    double r = sqrt(r2), sk = sqrt(r) * t->csi;
                    int kk = (int)sk;
                    double dk = sk - kk;
                    er0 = er0 + (TAB(t->potr, ktab, kk) + dk * (TAB(t->potr, ktab, kk + 1) - TAB(t->potr, ktab, kk))) / r;
                    den0 = den0 + (TAB(t->potb, ktab, kk) + dk * (TAB(t->potb, ktab, kk + 1) - TAB(t->potb, ktab, kk)));
                }
            }
            if (t->pot_type == 0)
                den0 = -sqrt(den0);
            else
                den0 = emb_lerp(t, t->fembd, t->kembd[ti], den0, &err);
        }
        epot[i] = er0 + den0;
    }
    return err ? ORC_ERR_ARG : ORC_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* adjoint of the full-list force and energy (every table kind)                                 */
/* ------------------------------------------------------------------------------------------ */

/* F'(rho) and dF'/drho of the density pass; E_emb(rho) and dE_emb/drho of the energy pass, with
 * the knot (kk) and weight (w = sk - kk) of each EAM lookup. */
static void emb_terms(const orc_ftable* t, int k1, double rho, double* fp, double* dfp,
                      double* de, int* kf, double* wf, int* ke, double* we) {
    *kf = -1;
    *ke = -1;
    if (t->pot_type == 0) {
        if (rho > 0.0) {
            double sr = sqrt(rho);
            *fp = -0.5 / sr;
            *dfp = 0.25 / (rho * sr);
            *de = -0.5 / sr;
        } else {
            *fp = rho;
            *dfp = 1.0;
            *de = 0.0;
        }
        return;
    }
    double sk = rho / t->rhod + 1.0;
    int kk = (int)(sk + 0.000001);
    if (kk < 0) kk = 0;
    if (kk > t->nembd) kk = t->nembd;
    *de = (ETAB(t->fembd, k1, kk + 1) - ETAB(t->fembd, k1, kk)) / t->rhod;
    *ke = kk;
    *we = sk - (double)kk;
    if (rho > 0.0) {
        *fp = ETAB(t->dfembd, k1, kk) + (sk - kk) * (ETAB(t->dfembd, k1, kk + 1) - ETAB(t->dfembd, k1, kk));
        *dfp = (ETAB(t->dfembd, k1, kk + 1) - ETAB(t->dfembd, k1, kk)) / t->rhod;
        *kf = kk;
        *wf = sk - (double)kk;
    } else {
        *fp = rho;
        *dfp = 1.0;
    }
}

int orc_adjoint_full(int n, const double* xp, const int* ityp, const int* statu,
                     const int* kvois, const int* indi, const double len[3], const int ifpd[3],
                     double ru2max, const orc_ftable* t, const double* dldf, double dlde,
                     double* g_potr, double* g_fpotr, double* g_potb, double* g_fpotb,
                     double* g_fembd, double* g_dfembd) {
    int ng = t->ngroup, nt = t->ntab + 2, ne = t->nembd + 2;
    if (t->pot_type != 0 && (!g_fembd || !g_dfembd)) return ORC_ERR_ARG;
    double hbox[3] = {0.5 * len[0], 0.5 * len[1], 0.5 * len[2]};
    double* rho = (double*)calloc((size_t)n, sizeof(double));
    double* fpr = (double*)calloc((size_t)n, sizeof(double));
    double* lam = (double*)calloc((size_t)n, sizeof(double));
    double* g = (double*)calloc((size_t)3 * n, sizeof(double));
    if (!rho || !fpr || !lam || !g) {
        free(rho); free(fpr); free(lam); free(g);
        return ORC_ERR_NOMEM;
    }
#define ACT(i) ((statu[i] & ORC_STATU_ACTIVE) == ORC_STATU_ACTIVE)
#define PAIR_GEOM(i, j)                                                                   \
    double s[3];                                                                          \
    for (int k = 0; k < 3; ++k)                                                           \
        s[k] = min_image(xp[i + k * n] - xp[j + k * n], len[k], hbox[k], ifpd[k]);        \
    double r2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
#define BIN(arr, kd, kk, w, dk)                                                           \
    do {                                                                                  \
        (arr)[(size_t)(kd) * nt + (kk)] += (w) * (1.0 - (dk));                            \
        (arr)[(size_t)(kd) * nt + (kk) + 1] += (w) * (dk);                                \
    } while (0)
    /* G of inactive atoms does not reach the tables (their force is identically zero) */
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) g[i + k * n] = ACT(i) ? dldf[i + k * n] : 0.0;
    /* forward density (kind KPAIR(ti, tj)) and F' (embedding kind KEMBD(ti)) */
    for (int i = 0; i < n; ++i) {
        if (!ACT(i)) continue;
        double den0 = 0.0;
        for (int iw = 0; iw < kvois[i]; ++iw) {
            int j = indi[i + (size_t)iw * n];
            PAIR_GEOM(i, j)
            if (r2 <= ru2max) {
                int kd = t->kpair[ityp[i] * ng + ityp[j]];
                double r = sqrt(r2), sk = sqrt(r) * t->csi;
                int kk = (int)sk;
                den0 = den0 + (TAB(t->potb, kd, kk) + (sk - kk) * (TAB(t->potb, kd, kk + 1) - TAB(t->potb, kd, kk)));
            }
        }
        rho[i] = den0;
        double dfp, de, wf, we;
        int kf, ke;
        emb_terms(t, t->kembd[ityp[i]], den0, &fpr[i], &dfp, &de, &kf, &wf, &ke, &we);
    }
    /* pair terms and lambda: F_i's pair term reads FPOTR and FPOTB of KPAIR(ti, tj) (times F'_i)
     * and FPOTB of KPAIR(tj, ti) (times F'_j); E_i reads POTR of KPAIR(ti, tj) */
    for (int i = 0; i < n; ++i) {
        if (!ACT(i)) continue;
        for (int iw = 0; iw < kvois[i]; ++iw) {
            int j = indi[i + (size_t)iw * n];
            PAIR_GEOM(i, j)
            if (!(r2 <= ru2max)) continue;
            int kd = t->kpair[ityp[i] * ng + ityp[j]], kd1 = t->kpair[ityp[j] * ng + ityp[i]];
            double r = sqrt(r2), sk = sqrt(r) * t->csi;
            int kk = (int)sk;
            double dk = sk - kk;
            double bp = TAB(t->fpotb, kd, kk) + dk * (TAB(t->fpotb, kd, kk + 1) - TAB(t->fpotb, kd, kk));
            double ai = g[i] * s[0] + g[i + n] * s[1] + g[i + 2 * n] * s[2];
            double aj = g[j] * s[0] + g[j + n] * s[1] + g[j + 2 * n] * s[2];
            BIN(g_fpotr, kd, kk, ai / r2, dk);
            BIN(g_fpotb, kd, kk, ai * fpr[i] / r, dk);
            BIN(g_fpotb, kd1, kk, ai * fpr[j] / r, dk);
            BIN(g_potr, kd, kk, dlde / r, dk);
            /* F'_i enters F_i through FPOTB(KPAIR(ti, tj)) and F_j through the same kind */
            lam[i] += (ai - aj) * bp / r;
        }
    }
    /* density terms: mu_i = lambda_i dF'/drho + g dE/drho */
    for (int i = 0; i < n; ++i) {
        if (!ACT(i)) continue;
        double fp, dfp, de, wf = 0.0, we = 0.0;
        int kf, ke;
        int k1 = t->kembd[ityp[i]];
        emb_terms(t, k1, rho[i], &fp, &dfp, &de, &kf, &wf, &ke, &we);
        if (t->pot_type != 0) {
            g_fembd[(size_t)k1 * ne + ke] += dlde * (1.0 - we);
            g_fembd[(size_t)k1 * ne + ke + 1] += dlde * we;
            if (kf >= 0) {
                g_dfembd[(size_t)k1 * ne + kf] += lam[i] * (1.0 - wf);
                g_dfembd[(size_t)k1 * ne + kf + 1] += lam[i] * wf;
            }
        }
        double mu = lam[i] * dfp + dlde * de;
        for (int iw = 0; iw < kvois[i]; ++iw) {
            int j = indi[i + (size_t)iw * n];
            PAIR_GEOM(i, j)
            if (!(r2 <= ru2max)) continue;
            int kd = t->kpair[ityp[i] * ng + ityp[j]];
            double r = sqrt(r2), sk = sqrt(r) * t->csi;
            int kk = (int)sk;
            double dk = sk - kk;
            BIN(g_potb, kd, kk, mu, dk);
        }
    }
#undef BIN
#undef PAIR_GEOM
#undef ACT
    free(rho);
    free(fpr);
    free(lam);
    free(g);
    return ORC_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* integrator: MD_SwopeScheme.F90:105-205 (predictor), :233-280 (correction)                   */
/* ------------------------------------------------------------------------------------------ */

void orc_predictor(int n, double* xp, double* xp1, const double* fp, double* dis, int* statu,
                   const int* ityp, const double* cm, double h, const double lo[3],
                   const double len[3], const int ifpd[3]) {
    double hs2 = h * 0.5, h2s2 = h * hs2;
    static const int fixpos[3] = {ORC_STATU_FIXPOSX, ORC_STATU_FIXPOSY, ORC_STATU_FIXPOSZ};
    static const int fixvel[3] = {ORC_STATU_FIXVELX, ORC_STATU_FIXVELY, ORC_STATU_FIXVELZ};
    for (int k = 0; k < 3; ++k) {
        double ub = lo[k] + len[k];
        for (int j = 0; j < n; ++j) {
            if ((statu[j] & ORC_STATU_ACTIVE) != ORC_STATU_ACTIVE) continue;
            if (statu[j] & fixpos[k]) continue;
            double cm0 = cm[ityp[j]];
            size_t a = (size_t)j + (size_t)k * n;
            double dist = h * xp1[a] + h2s2 * fp[a] / cm0;
            dis[a] = dis[a] + dist;
            xp[a] = xp[a] + dist;
            if (xp[a] > ub) {
                if (ifpd[k])
                    xp[a] = xp[a] - len[k];
                else
                    statu[j] = ORC_STATU_OUTOFBOX | ORC_STATU_REFLECT;
            } else if (xp[a] < lo[k]) {
                if (ifpd[k])
                    xp[a] = xp[a] + len[k];
                else
                    statu[j] = ORC_STATU_OUTOFBOX | ORC_STATU_TRANSMIT;
            }
            if ((statu[j] & fixvel[k]) == 0) xp1[a] = xp1[a] + hs2 * fp[a] / cm0;
        }
    }
}

void orc_correction(int n, double* xp1, const double* fp, const int* statu, const int* ityp,
                    const double* cm, double h, double* ekin) {
    double hs2 = h * 0.5;
    static const int fix[3] = {ORC_STATU_FIXPOSX | ORC_STATU_FIXVELX, ORC_STATU_FIXPOSY | ORC_STATU_FIXVELY,
                               ORC_STATU_FIXPOSZ | ORC_STATU_FIXVELZ};
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < n; ++j) {
            if ((statu[j] & ORC_STATU_ACTIVE) != ORC_STATU_ACTIVE) continue;
            if (statu[j] & fix[k]) continue;
            size_t a = (size_t)j + (size_t)k * n;
            xp1[a] = xp1[a] + hs2 * fp[a] / cm[ityp[j]];
        }
    if (ekin)
        for (int j = 0; j < n; ++j) {
            if ((statu[j] & ORC_STATU_ACTIVE) != ORC_STATU_ACTIVE) continue;
            double vx = xp1[j], vy = xp1[j + n], vz = xp1[j + 2 * (size_t)n];
            ekin[j] = 0.5 * cm[ityp[j]] * (vx * vx + vy * vy + vz * vz);
        }
}

/* DAMPING_KERNEL (MD_DiffScheme_GPU.F90:125-187): a velocity component whose product with the
 * force component is negative becomes zero, and so does a component with FIXPOS set. */
void orc_damping(int n, double* xp1, const double* fp, const int* statu) {
    static const int fixp[3] = {ORC_STATU_FIXPOSX, ORC_STATU_FIXPOSY, ORC_STATU_FIXPOSZ};
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < n; ++j) {
            size_t a = (size_t)j + (size_t)k * n;
            if (xp1[a] * fp[a] < 0.0) xp1[a] = 0.0;
            if (statu[j] & fixp[k]) xp1[a] = 0.0;
        }
}

/* CALEKIN_KERNEL (:851-904): EK = CM/2 |v|^2 for active atoms with no FIXPOS bit, else -1e32. */
void orc_ekin(int n, const double* xp1, const int* statu, const int* ityp, const double* cm,
              double* ek) {
    const int fixpos = ORC_STATU_FIXPOSX | ORC_STATU_FIXPOSY | ORC_STATU_FIXPOSZ;
    for (int j = 0; j < n; ++j) {
        ek[j] = -1.0e32;
        if ((statu[j] & ORC_STATU_ACTIVE) == ORC_STATU_ACTIVE && (statu[j] & fixpos) == 0) {
            double vx = xp1[j], vy = xp1[j + n], vz = xp1[j + 2 * (size_t)n];
            ek[j] = 0.5 * cm[ityp[j]] * (vx * vx + vy * vy + vz * vz);
        }
    }
}

/* VelScaling_KERNEL (:1262-1326) with the factor of VelScaling_DEV (:1390-1448): velocities of
 * active atoms times sqrt(temp 3 kb / 2 / <EK>) (mean over the atoms with EK >= 0), components
 * with FIXPOS or FIXVEL set to zero.  Returns ORC_ERR_ARG when <EK> is not positive. */
int orc_velscale(int n, double* xp1, const int* statu, const int* ityp, const double* cm,
                 double temp, double kb) {
    double* ek = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double sum = 0.0, cnt = 0.0;
    orc_ekin(n, xp1, statu, ityp, cm, ek);
    for (int j = 0; j < n; ++j)
        if (ek[j] >= 0.0) {
            sum += ek[j];
            cnt += 1.0;
        }
    free(ek);
    if (!(cnt > 0.0) || !(sum > 0.0)) return ORC_ERR_ARG;
    double f = sqrt(temp * 3.0 * kb * 0.5 / (sum / cnt));
    static const int fix[3] = {ORC_STATU_FIXPOSX | ORC_STATU_FIXVELX, ORC_STATU_FIXPOSY | ORC_STATU_FIXVELY,
                               ORC_STATU_FIXPOSZ | ORC_STATU_FIXVELZ};
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < n; ++j) {
            if ((statu[j] & ORC_STATU_ACTIVE) != ORC_STATU_ACTIVE) continue;
            size_t a = (size_t)j + (size_t)k * n;
            xp1[a] = (statu[j] & fix[k]) ? 0.0 : xp1[a] * f;
        }
    return ORC_OK;
}

int orc_md_run(int n, double* xp, double* xp1, double* fp, double* dis, int* statu,
               const int* ityp, const double* cm, double h, int nsteps, int nb_uptab, int step0,
               int init, const double lo[3], const double len[3], const int ifpd[3],
               const double* nb_rm, const double* ru, int mxkvois, const orc_ftable* t,
               int* kvois, int* indi, double* epot) {
    int rc;
    double vt[9];
    double* den = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* dden = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (init) {
        rc = orc_nlist_half(n, xp, ityp, statu, lo, len, ifpd, t->ngroup, nb_rm, mxkvois, kvois, indi);
        if (rc) goto out;
        rc = orc_force_fs_half(n, xp, ityp, kvois, indi, mxkvois, len, ifpd, ru, t, fp, epot, den, dden, vt, NULL);
        if (rc) goto out;
    }
    rc = ORC_OK;
    for (int s = step0 + 1; s <= step0 + nsteps; ++s) {
        orc_predictor(n, xp, xp1, fp, dis, statu, ityp, cm, h, lo, len, ifpd);
        if (nb_uptab > 0 && s % nb_uptab == 0) {
            rc = orc_nlist_half(n, xp, ityp, statu, lo, len, ifpd, t->ngroup, nb_rm, mxkvois, kvois, indi);
            if (rc) goto out;
        }
        rc = orc_force_fs_half(n, xp, ityp, kvois, indi, mxkvois, len, ifpd, ru, t, fp, epot, den, dden, vt, NULL);
        if (rc) goto out;
        orc_correction(n, xp1, fp, statu, ityp, cm, h, NULL);
    }
out:
    free(den);
    free(dden);
    return rc;
}

/* ------------------------------------------------------------------------------------------ */
/* RNG: DRAND32.F90:60-85 (L'Ecuyer's combined generator with Schrage's factorisation)         */
/* ------------------------------------------------------------------------------------------ */

double orc_drand32(int* seed) {
    int i1 = seed[0] / 53668;
    seed[0] = 40014 * (seed[0] - i1 * 53668) - i1 * 12211;
    if (seed[0] < 0) seed[0] += 2147483563;
    int i2 = seed[1] / 52774;
    seed[1] = 40692 * (seed[1] - i2 * 52774) - i2 * 3791;
    if (seed[1] < 0) seed[1] += 2147483399;
    int iz = seed[0] - seed[1];
    if (iz < 1) iz += 2147483562;
    return (double)iz * (1.0 / 2147483648.0);
}

void orc_thermalize_mc(int n, const int* ityp, const double* cm, int ngroup, double temp,
                       double kb, int* seed, double* xp1) {
    /* CP_TWOPI = 2*3.141592654 in the reference (MSM_Const.F90:56-59) */
    const double twopi = 2.0 * 3.141592654;
    double wt = 0.0;
    for (int j = 0; j < n; ++j) wt += cm[ityp[j]];
    (void)ngroup;
    for (int k = 0; k < 3; ++k) {
        double vt = 0.0;
        for (int j = 0; j < n; ++j) {
            double v0 = sqrt(2.0 * temp * kb / cm[ityp[j]]);
            double z1 = orc_drand32(seed);
            double z2 = orc_drand32(seed);
            double v = v0 * sqrt(-log(z1)) * cos(twopi * z2);
            xp1[j + (size_t)k * n] = v;
            vt = vt + (cm[ityp[j]] * v) / wt;
        }
        for (int j = 0; j < n; ++j) xp1[j + (size_t)k * n] = xp1[j + (size_t)k * n] - vt;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* device thermalisation: Philox4x32-10 uniforms, reference transformation                    */
/* ------------------------------------------------------------------------------------------ */

/* One block of Philox4x32 with 10 rounds: c' = (hi(M1 c2) ^ c1 ^ k0, lo(M1 c2),
 * hi(M0 c0) ^ c3 ^ k1, lo(M0 c0)), the key bumped by the Weyl constants between rounds. */
void orc_philox4x32_10(const unsigned* ctr, const unsigned* key, unsigned* out) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double unit_open0(uint32_t lo, uint32_t hi) {
    uint64_t m = ((uint64_t)(hi & 0x1FFFFFu) << 32) | (uint64_t)lo;
    return (double)(m + 1u) * 0x1.0p-53;
}

/* Thermalizing_MC_KERNEL (MD_DiffScheme_GPU.F90:1605-1651) and the per-box momentum removal of
 * Thermalizing_MC_DEV (:1781-1800): active atoms' free components V0 sqrt(-ln z1) cos(2 pi z2),
 * V0 = sqrt(2 T kB / CM), FIXVEL/FIXPOS components kept, inactive atoms zero; then
 * v -= sum(m v) / sum(m) over each box (all its atoms). */
void orc_thermalize_philox(int n, int nbox, const int* natoms, const int* ityp, const int* statu,
                           const double* cm, double tkb, unsigned long long seed,
                           unsigned long long draw, double* xp1) {
    const unsigned key[2] = {(unsigned)seed, (unsigned)(seed >> 32)};
    const int fix[3] = {16 | 2, 32 | 4, 64 | 8};
    for (int i = 0; i < n; ++i) {
        int st = statu[i];
        if ((st & 1) != 1) {
            xp1[i] = xp1[n + i] = xp1[2 * n + i] = 0.0;
            continue;
        }
        double v0 = sqrt(2.0 * tkb / cm[ityp[i] - 1]);
        for (int k = 0; k < 3; ++k) {
            if (st & fix[k]) continue;
            unsigned c[4] = {(unsigned)i, (unsigned)k, (unsigned)draw, (unsigned)(draw >> 32)};
            unsigned r[4];
            orc_philox4x32_10(c, key, r);
            double z1 = unit_open0(r[0], r[1]), z2 = unit_open0(r[2], r[3]);
            xp1[k * n + i] = v0 * sqrt(-log(z1)) * cos(6.283185307179586 * z2);
        }
    }
    int off = 0;
    for (int b = 0; b < nbox; ++b) {
        double wt = 0.0, vt[3] = {0.0, 0.0, 0.0};
        for (int i = off; i < off + natoms[b]; ++i) {
            double m = cm[ityp[i] - 1];
            wt += m;
            for (int k = 0; k < 3; ++k) vt[k] += m * xp1[k * n + i];
        }
        if (wt > 0.0)
            for (int i = off; i < off + natoms[b]; ++i)
                for (int k = 0; k < 3; ++k) xp1[k * n + i] -= vt[k] / wt;
        off += natoms[b];
    }
}
