// msm_host.cpp -- host-side helpers of libmsm_hip.so: force-table generation, built-in potential
// forms, the DRAND32 generator and Maxwell thermalisation.  None of this touches the GPU.
//
// Table generation follows MD_TypeDef_ForceTable.F90:591-594 (grid), :949-976 (pair tables) and
// :1043-1048 (embedding tables); potentials follow FS_Ackland_WW.F90:23-91 and
// TB_CR_AuAu.F90:17-53; DRAND32.F90:60-85; MD_TypeDef_SimBox.F90:4865-4942.
#include <cmath>
#include <cstddef>

#include "msm_hip.h"

extern "C" {

int msm_table_pairwise(int ntab, double rmax, msm_pairfn nn, msm_pairfn ne, double rhomult,
                       double* rhomx_inout, double* potr, double* fpotr, double* potb,
                       double* fpotb, int stride) {
    if (ntab < 2 || !(rmax > 0.0) || !nn || !potr || !fpotr || !potb || !fpotb || stride < 1 ||
        !rhomx_inout)
        return MSM_ERR_ARG;
    const double csi = static_cast<double>(ntab) / std::sqrt(rmax);
    const double csiv = 1.0 / csi;
    for (int i = 1; i <= ntab; ++i) {
        const double t = static_cast<double>(i) * csiv;
        const double r = t * t;
        double v = 0.0, dv = 0.0;
        nn(&r, &v, &dv);
        const std::size_t a = static_cast<std::size_t>(i - 1) * stride;
        potr[a] = v * r;
        fpotr[a] = dv * r;
    }
    for (int i = 1; i <= ntab; ++i) {
        const std::size_t a = static_cast<std::size_t>(i - 1) * stride;
        if (ne) {
            const double t = static_cast<double>(i) * csiv;
            const double r = t * t;
            ne(&r, &potb[a], &fpotb[a]);
            if (*rhomx_inout < potb[a] * rhomult) *rhomx_inout = potb[a] * rhomult;
        } else {
            potb[a] = 0.0;
            fpotb[a] = 0.0;
        }
    }
    return MSM_OK;
}

int msm_table_embd(int nembd, double rhod, msm_pairfn embd, double* fembd, double* dfembd,
                   int stride) {
    if (nembd < 2 || !(rhod > 0.0) || !fembd || !dfembd || stride < 1) return MSM_ERR_ARG;
    for (int i = 1; i <= nembd; ++i) {
        const std::size_t a = static_cast<std::size_t>(i - 1) * stride;
        if (embd) {
            const double rho = static_cast<double>(i - 1) * rhod;
            embd(&rho, &fembd[a], &dfembd[a]);
        } else {
            fembd[a] = 0.0;
            dfembd[a] = 0.0;
        }
    }
    return MSM_OK;
}

// ---- W Finnis-Sinclair, benchmark form: pair polynomial without the Ackland core ----------
namespace {
constexpr double kWfsC = 3.25, kWfsC0 = 47.1346, kWfsC1 = -33.7666, kWfsC2 = 6.2542;
constexpr double kWfsA = 1.8964, kWfsD = 4.4002;
// Ackland-Thetford W-W, the reference's coefficients in eV and Angstrom.
constexpr double kAckA = 1.896373, kAckC = 3.25, kAckC0 = 47.1346499, kAckC1 = -33.7665655,
                 kAckC2 = 6.2541999, kAckD = 4.400224, kAckB = 90.3, kAckAlpha = 1.2,
                 kAckB0 = 2.7411;

struct TbSma {
    double a, xi, p, q, r0;
};
// Cleri & Rosato, PRB 48 (1993) 22, Table I (r0 = a/sqrt 2 with a = 3.615 and 3.524 Angstrom).
constexpr TbSma kCu{0.0855, 1.2240, 10.960, 2.2780, 3.615 * 0.70710678118654752};
constexpr TbSma kNi{0.0376, 1.0700, 16.999, 1.1890, 3.524 * 0.70710678118654752};

TbSma tbsma_params(int pair) {
    if (pair == 0) return kCu;
    if (pair == 1) return kNi;
    return TbSma{0.5 * (kCu.a + kNi.a), std::sqrt(kCu.xi * kNi.xi), 0.5 * (kCu.p + kNi.p),
                 0.5 * (kCu.q + kNi.q), 0.5 * (kCu.r0 + kNi.r0)};
}
}  // namespace

void msm_pot_wfs_nn(const double* rp, double* potr, double* fpotr) {
    const double r = *rp;
    if (r <= kWfsC) {
        const double rc = r - kWfsC;
        const double poly = kWfsC0 + kWfsC1 * r + kWfsC2 * r * r;
        *potr = 0.5 * (rc * rc * poly);
        *fpotr = -(2.0 * rc * poly + rc * rc * (kWfsC1 + 2.0 * kWfsC2 * r));
    } else {
        *potr = 0.0;
        *fpotr = 0.0;
    }
}

void msm_pot_wfs_ne(const double* rp, double* potb, double* fpotb) {
    const double r = *rp;
    if (r <= kWfsD) {
        const double rd = r - kWfsD;
        *potb = kWfsA * kWfsA * rd * rd;
        *fpotb = -(kWfsA * kWfsA * 2.0 * rd);
    } else {
        *potb = 0.0;
        *fpotb = 0.0;
    }
}

void msm_pot_ackland_ww_nn(const double* rp, double* potr, double* fpotr) {
    const double r = *rp;
    double core = 0.0, dcore = 0.0;
    if (r < kAckB0) {
        const double b = kAckB0 - r, e = std::exp(-kAckAlpha * r);
        core = kAckB * b * b * b * e;
        dcore = -kAckB * (3.0 * b * b * e + kAckAlpha * b * b * b * e);
    }
    if (r <= kAckC) {
        const double rc = r - kAckC;
        const double poly = kAckC0 + kAckC1 * r + kAckC2 * r * r;
        *potr = 0.5 * (rc * rc * poly + core);
        *fpotr = -(2.0 * rc * poly + rc * rc * (kAckC1 + 2.0 * kAckC2 * r) + dcore);
    } else {
        *potr = 0.0;
        *fpotr = 0.0;
    }
}

void msm_pot_ackland_ww_ne(const double* rp, double* potb, double* fpotb) {
    const double r = *rp;
    if (r <= kAckD) {
        const double rd = r - kAckD;
        *potb = kAckA * kAckA * rd * rd;
        *fpotb = -(kAckA * kAckA * 2.0 * rd);
    } else {
        *potb = 0.0;
        *fpotb = 0.0;
    }
}

void msm_emb_fs(const double* rhop, double* f, double* df) {
    const double rho = *rhop;
    if (rho > 0.0) {
        *f = -std::sqrt(rho);
        *df = -0.5 / std::sqrt(rho);
    } else {
        *f = 0.0;
        *df = 0.0;
    }
}

void msm_pot_tbsma_nn(int pair, const double* rp, double* potr, double* fpotr) {
    const TbSma c = tbsma_params(pair);
    *potr = c.a * std::exp(-c.p * ((*rp / c.r0) - 1.0));
    *fpotr = 2.0 * (*potr) * (c.p / c.r0);
}

void msm_pot_tbsma_ne(int pair, const double* rp, double* potb, double* fpotb) {
    const TbSma c = tbsma_params(pair);
    *potb = (c.xi * c.xi) * std::exp(-(2.0 * c.q) * ((*rp / c.r0) - 1.0));
    *fpotb = *potb * (c.q / c.r0) * 2.0;
}

void msm_pot_tbsma_cucu_nn(const double* r, double* v, double* dv) { msm_pot_tbsma_nn(0, r, v, dv); }
void msm_pot_tbsma_cucu_ne(const double* r, double* v, double* dv) { msm_pot_tbsma_ne(0, r, v, dv); }
void msm_pot_tbsma_nini_nn(const double* r, double* v, double* dv) { msm_pot_tbsma_nn(1, r, v, dv); }
void msm_pot_tbsma_nini_ne(const double* r, double* v, double* dv) { msm_pot_tbsma_ne(1, r, v, dv); }
void msm_pot_tbsma_cuni_nn(const double* r, double* v, double* dv) { msm_pot_tbsma_nn(2, r, v, dv); }
void msm_pot_tbsma_cuni_ne(const double* r, double* v, double* dv) { msm_pot_tbsma_ne(2, r, v, dv); }

// ---- DRAND32: L'Ecuyer's combined multiplicative generator (CACM 31 (1988) 742), with the
// same Schrage factorisation and constants as the reference. ------------------------------
double msm_drand32(int* seed) {
    int i1 = seed[0] / 53668;
    seed[0] = 40014 * (seed[0] - i1 * 53668) - i1 * 12211;
    if (seed[0] < 0) seed[0] += 2147483563;
    int i2 = seed[1] / 52774;
    seed[1] = 40692 * (seed[1] - i2 * 52774) - i2 * 3791;
    if (seed[1] < 0) seed[1] += 2147483399;
    int iz = seed[0] - seed[1];
    if (iz < 1) iz += 2147483562;
    return static_cast<double>(iz) * (1.0 / 2147483648.0);
}

int msm_thermalize_mc(int n, const int* ityp, const double* cm, double temp, double kb,
                      int* seed, double* xp1) {
    if (n < 0 || !ityp || !cm || !seed || !xp1) return MSM_ERR_ARG;
    const double twopi = 2.0 * 3.141592654;  // the reference's CP_TWOPI
    double wt = 0.0;
    for (int j = 0; j < n; ++j) wt += cm[ityp[j] - 1];
    for (int k = 0; k < 3; ++k) {
        double* v = xp1 + static_cast<std::size_t>(k) * n;
        double vt = 0.0;
        for (int j = 0; j < n; ++j) {
            const double m = cm[ityp[j] - 1];
            const double v0 = std::sqrt(2.0 * temp * kb / m);
            const double z1 = msm_drand32(seed);
            const double z2 = msm_drand32(seed);
            v[j] = v0 * std::sqrt(-std::log(z1)) * std::cos(twopi * z2);
            vt = vt + (m * v[j]) / wt;
        }
        for (int j = 0; j < n; ++j) v[j] = v[j] - vt;
    }
    return MSM_OK;
}

}  // extern "C"
