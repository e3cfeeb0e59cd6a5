// Continuum absorption on the device: MT-CKD 4.3 H2O self + foreign, the
// MT-CKD CO2 line-coupling continuum and HITRAN CIA (CO2-CO2, N2-N2, O2-O2),
// restated from src/CrossSections/continuum.jl:138-363 and added per layer to
// the optical depth as src/Sensor/sounder.jl:310-322 does (mid-layer VMR, p, T;
// dz from src/CrossSections/optical_depth.jl:123-130).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define IRS_CONT_H2O    1
#define IRS_CONT_CO2    2
#define IRS_CONT_CO2CIA 4
#define IRS_CONT_N2     8
#define IRS_CONT_O2    16
#define IRS_CIA_MAXB  128      // temperature blocks per CIA pair

// One CIA pair on the device.  Blocks are sorted by temperature; block b holds
// knots nu[off[b] .. off[b+1]) (ascending) and sigma >= 0 (cm^5 molec^-2).
struct CiaDev {
    int nb;
    const double *T, *numin, *numax;   // [nb]
    const int64_t *off;                // [nb + 1]
    const double *nu, *sig;
};

// Per (profile, layer) continuum state written by the host: mid-layer VMRs
// (N2 and O2 are 0.78084 and 0.20946 of the dry fraction 1 - vmr_H2O),
// pressure (hPa), temperature, geometric thickness dz (cm), and the mid-layer
// interpolation weight wM = ln(p_mid/p_k)/ln(p_{k+1}/p_k) (Jacobian chain rule).
struct ContLayer {
    double vh2o, vco2, vn2, vo2, p_mid, T_mid, dz, wM;
};

struct ContArgs {
    float *tau;                        // [n_prof][n_layers][n]   (+= tau_c)
    int64_t n;
    int n_layers, n_prof;
    double an, sn, den;                // grid node i = (an + i*sn)/den
    const ContLayer *lay;              // [n_prof][n_layers]
    int continua;                      // bit mask IRS_CONT_*
    const double *h2o;                 // [IRS_CKD_H2O_N][4]: nu (0, 10, ..), Cs, Cf, texp
    const double *co2;                 // [IRS_CKD_CO2_N][2]: S*xfac, tdep on nu = -4 + 2 i
    CiaDev cia[3];
    // Jacobian outputs for profile 0 (each may be null), [n_layers][n]:
    //   dtlo += d tau_c/d T_mid (1 - wM),  dthi += d tau_c/d T_mid wM
    //   vlo_h/vhi_h, vlo_c/vhi_c: the same for the H2O / CO2 level VMR
    float *dtlo, *dthi, *vlo_h, *vhi_h, *vlo_c, *vhi_c;
    double *tau64;                     // optional FP64 tau instead of tau, [n_prof][n_layers][n]
    int add64;                         // 1: tau64 += tau_c (FP64 forward), 0: tau64 = tau_c (layer export)
};

hipError_t launch_continuum(const ContArgs &a, hipStream_t s);
