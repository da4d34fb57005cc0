/*
 * gegs_media.h -- device-resident physics tables of the transport path.
 *
 * One plain-old-data block holding every table the reference keeps as
 * __device__/__constant__ symbols (commonBlocks dev.h:76-262,
 * media.cu:98-100) in the reference's own index layouts, so the index
 * arithmetic of the samplers carries over unchanged:
 *   electron tables   [qel*MXEKE*MXMED + medium*MXEKE + lelke-1]
 *   range_ep          [(medium*MXEKE + lelke-1)*2 + qel]
 *   photon tables     [medium*MXGE + lgle]
 *   brems/pair dl*    [medium*8 + l-1]
 *   MS q2+ surface    [(k*(MAXQ_MS+1) + j)*(MAXL_MS+1) + i]
 * The block is built on the host (gegs_amd/hatch.py restates the HATCH
 * builders of media.cu and PhysicsInit.cu) and copied to HBM once per
 * context; kernels read it through one const pointer.
 */
#ifndef GEGS_MEDIA_H
#define GEGS_MEDIA_H
#include <stdint.h>

#define GEGS_MXMED 6
#define GEGS_MXGE 200
#define GEGS_MXEKE 150
#define GEGS_MXEL 10
#define GEGS_MXRAYFF 100
#define GEGS_RAYCDFSIZE 100
#define GEGS_MAXL_MS 63
#define GEGS_MAXQ_MS 7
#define GEGS_MAXU_MS 31
#define GEGS_MS_SIZE ((GEGS_MAXU_MS + 1) * (GEGS_MAXQ_MS + 1) * (GEGS_MAXL_MS + 1))
#define GEGS_MXELEMENT 100        /* elements Z = 1 .. 100 (EGSnrc_Parameters.h:59) */
#define GEGS_MXSHELL 6            /* K, L1, L2, L3, <M>, <N> (EGSnrc_Parameters.h:60) */
#define GEGS_MXTRANS 39           /* relaxation transitions (EGSnrc_Parameters.h:62) */
#define GEGS_MXEDGE 16            /* photo-absorption edges per element (EGSnrc_Parameters.h:63) */
#define GEGS_MXTOTSH 1538         /* Compton sub-shells of all elements (EGSnrc_Parameters.h:55) */
#define GEGS_MXMDSH 50            /* Compton sub-shells per medium (EGSnrc_Parameters.h:56) */
#define GEGS_MAX_EII_SHELLS 40    /* EGSnrc_Parameters.h:66 */
#define GEGS_N_EII_BINS 250       /* EGSnrc_Parameters.h:67 */
#define GEGS_MXVAC 10             /* vacancies of a relaxation cascade (EGSnrc_Parameters.h:128) */

typedef struct gegs_float2 { float x, y; } gegs_float2;

typedef struct gegs_rayleigh {
    float xgrid, fcum, b_array, c_array;
} gegs_rayleigh;

typedef struct gegs_media_tables {
    int32_t nmed;
    int32_t eii_flag;
    int32_t nbr_split;
    int32_t ibrdst;
    int32_t iprdst;
    int32_t exact_bca;
    int32_t spin_effects;
    int32_t transport_algorithm;   /* 0 = PRESTA-II, 1 = PRESTA-I */
    float skindepth_for_bca;
    float dllambi;
    float dqmsi;
    float rm;
    /* per-medium scalars */
    float rho[GEGS_MXMED];
    float te[GEGS_MXMED];
    float thmoll[GEGS_MXMED];
    float ae[GEGS_MXMED];
    float ap[GEGS_MXMED];
    float log_ap[GEGS_MXMED];
    float smaxir[GEGS_MXMED];
    float e_max_rr[GEGS_MXMED];
    float zbrang[GEGS_MXMED];
    float lzbrang[GEGS_MXMED];
    float delcm[GEGS_MXMED];
    float blcc[GEGS_MXMED];
    float xcc[GEGS_MXMED];
    int32_t i_do_rr[GEGS_MXMED];
    int32_t meke[GEGS_MXMED];
    int32_t nne[GEGS_MXMED];
    float zelem[GEGS_MXEL * GEGS_MXMED];
    float pz[GEGS_MXEL * GEGS_MXMED];
    float dl1[8 * GEGS_MXMED];
    float dl2[8 * GEGS_MXMED];
    float dl3[8 * GEGS_MXMED];
    float dl4[8 * GEGS_MXMED];
    float dl5[8 * GEGS_MXMED];
    float dl6[8 * GEGS_MXMED];
    float bpar[2 * GEGS_MXMED];
    float sig_e[2 * GEGS_MXMED];
    int32_t sig_ismonotone[2 * GEGS_MXMED];
    gegs_float2 eke01[GEGS_MXMED];
    /* electron tables */
    gegs_float2 sig[2 * GEGS_MXMED * GEGS_MXEKE];
    gegs_float2 dedx[2 * GEGS_MXMED * GEGS_MXEKE];
    gegs_float2 ebr1[GEGS_MXMED * GEGS_MXEKE];
    gegs_float2 pbr1[GEGS_MXMED * GEGS_MXEKE];
    gegs_float2 pbr2[GEGS_MXMED * GEGS_MXEKE];
    gegs_float2 tmxs[GEGS_MXMED * GEGS_MXEKE];
    float range_ep[2 * GEGS_MXMED * GEGS_MXEKE];
    float e_array[GEGS_MXMED * GEGS_MXEKE];
    /* photon tables */
    gegs_float2 ge[GEGS_MXMED];
    gegs_float2 gmfp[GEGS_MXMED * GEGS_MXGE];
    gegs_float2 gbr1[GEGS_MXMED * GEGS_MXGE];
    gegs_float2 gbr2[GEGS_MXMED * GEGS_MXGE];
    gegs_float2 cohe[GEGS_MXMED * GEGS_MXGE];
    gegs_float2 pmax[GEGS_MXMED * GEGS_MXGE];
    gegs_rayleigh rayleigh_data[GEGS_MXMED * GEGS_MXRAYFF];
    uint32_t i_array[GEGS_MXMED * GEGS_RAYCDFSIZE];
    /* screened-Rutherford multiple-scattering q2+ surface */
    float ums[GEGS_MS_SIZE];
    float fms[GEGS_MS_SIZE];
    float wms[GEGS_MS_SIZE];
    int16_t ims[GEGS_MS_SIZE];
    /* photo-electric edges and atomic relaxation (edgset, PhysicsInit.cu:1184-1334):
     *   binding_energies, interaction_prob [(Z-1)*MXSHELL + shell-1], MeV
     *   relaxation_prob                    [(Z-1)*MXTRANS + k-1]
     *   edge_*                             [(Z-1)*MXEDGE + j-1], energies in MeV */
    float binding_energies[GEGS_MXSHELL * GEGS_MXELEMENT];
    float interaction_prob[GEGS_MXSHELL * GEGS_MXELEMENT];
    float relaxation_prob[GEGS_MXTRANS * GEGS_MXELEMENT];
    float edge_energies[GEGS_MXEDGE * GEGS_MXELEMENT];
    float edge_a[GEGS_MXEDGE * GEGS_MXELEMENT];
    float edge_b[GEGS_MXEDGE * GEGS_MXELEMENT];
    float edge_c[GEGS_MXEDGE * GEGS_MXELEMENT];
    float edge_d[GEGS_MXEDGE * GEGS_MXELEMENT];
    int32_t edge_number[GEGS_MXELEMENT];
    /* bound Compton scattering (init_compton, PhysicsInit.cu:1336-1519):
     *   per medium an alias table over its sub-shells [medium*MXMDSH + i-1];
     *   per sub-shell j of the data: Z, relaxation shell, binding energy (units
     *   of m), J0 (units of 1/(m c)) and the erf term of the p_z rejection */
    int32_t n_shell[GEGS_MXMED];
    int32_t shell_array[GEGS_MXMDSH * GEGS_MXMED];
    float eno_array[GEGS_MXMDSH * GEGS_MXMED];
    int32_t eno_atbin_array[GEGS_MXMDSH * GEGS_MXMED];
    int32_t iz_array[GEGS_MXTOTSH];
    int32_t shn_array[GEGS_MXTOTSH];
    float be_array[GEGS_MXTOTSH];
    float jo_array[GEGS_MXTOTSH];
    float erfjo_array[GEGS_MXTOTSH];
    /* electron-impact ionisation (eii_init, PhysicsInit.cu:1643-2047):
     *   eii_no / eii_first [iele*MXMED + medium]; eii / eii_xsection per
     *   EII shell jj: [jj-1] and [(jj-1)*N_EII_BINS + bin-1] */
    int32_t eii_nsh[GEGS_MXMED];
    int32_t eii_no[GEGS_MXEL * GEGS_MXMED];
    int32_t eii_first[GEGS_MXEL * GEGS_MXMED];
    float eii_cons[GEGS_MXMED];
    gegs_float2 eii[GEGS_MAX_EII_SHELLS];
    gegs_float2 eii_xsection[GEGS_N_EII_BINS * GEGS_MAX_EII_SHELLS];
} gegs_media_tables;

/* Region record, 16 bytes (CUDA_EGS.h:577-584). */
typedef struct gegs_region {
    uint16_t med;      /* 0xFFFF = vacuum */
    uint16_t flags;    /* GEGS_F_* bits below */
    float rhof;        /* density / medium density */
    float pcut;        /* total-energy photon cut (MeV) */
    float ecut;        /* total-energy electron cut (MeV) */
} gegs_region;

#define GEGS_VACUUM 0xFFFFu
#define GEGS_F_RAYLEIGH 0x0001u
#define GEGS_F_BOUND_COMPTON 0x0002u
#define GEGS_F_ATOMIC_RELAXATION 0x0010u
#define GEGS_F_PHOTOELECTRON_ANGLE 0x0020u
#define GEGS_F_RANGE_REJECTION 0x0040u
/* Bound Compton variants (compton.cu:87-95, CUDA_EGS.h:566-570): bits 0x000E
 * of the flags select ibcmpl: 0x0002 binding and Doppler broadening with
 * rejected scatters returned to the photon step (the reference default),
 * 0x0006 neither binding nor Doppler, 0x000A rejected scatters resampled;
 * any other value Klein-Nishina. */
#define GEGS_F_BOUND_COMPTON_2 0x0006u
#define GEGS_F_BOUND_COMPTON_3 0x000Au
#define GEGS_F_BOUND_COMPTON_MASK 0x000Eu
/* The reference's default region flags (CUDA_EGS.h:44-51, init.cu:935-1000):
 * Rayleigh, atomic relaxation, photo-electron angles, bound Compton. */
#define GEGS_F_REFERENCE_DEFAULT \
    (GEGS_F_RAYLEIGH | GEGS_F_ATOMIC_RELAXATION | GEGS_F_PHOTOELECTRON_ANGLE | GEGS_F_BOUND_COMPTON)
/* Flags the kernels act on; gegs_load_phantom rejects a region carrying any
 * other bit. */
#define GEGS_F_SUPPORTED (GEGS_F_RAYLEIGH | GEGS_F_BOUND_COMPTON_MASK | GEGS_F_ATOMIC_RELAXATION | \
                          GEGS_F_PHOTOELECTRON_ANGLE | GEGS_F_RANGE_REJECTION)

#endif
