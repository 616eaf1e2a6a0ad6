/*
 * fullmag_fdm.h -- C ABI of the MI355X-native FDM micromagnetic engine
 * (library libfullmag_fdm.so, SONAME libfullmag_fdm.so.0).
 *
 * Drop-in replacement for the reference backend's interface
 * native/include/fullmag_fdm.h (mzrptu/fullmag).  Every type below has the
 * same size, field order and field types as the reference declaration, so the
 * Rust mirror in crates/fullmag-fdm-sys/src/lib.rs:71-262 binds it unchanged;
 * every function has the same name, arguments and return convention.  The
 * reference line that each declaration replaces is cited beside it.
 *
 * Engine-specific additions (device-resident multi-step loop, mixed
 * tolerances, ...) live in fullmag_fdm_ext.h; nothing here deviates from the
 * reference surface.
 *
 * Conventions (reference api.cpp:151-1070):
 *   - return codes: 0 ok, -1 invalid argument/state, -2 device (HIP) error,
 *     -3 internal error, -4 interrupted by the poll callback;
 *   - create() returns a non-NULL handle with last_error set when the plan is
 *     rejected or allocation fails; NULL only for a NULL plan or host OOM;
 *   - plan arrays are only read during create();
 *   - host vectors are AoS [x0,y0,z0,x1,...]; the device keeps SoA;
 *   - one handle is driven by one host thread; snapshot payloads may be
 *     consumed on another thread.
 */
#ifndef FULLMAG_FDM_H
#define FULLMAG_FDM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes (reference fullmag_fdm.h:37-41). */
#define FULLMAG_FDM_OK               0
#define FULLMAG_FDM_ERR_INVALID     -1
#define FULLMAG_FDM_ERR_CUDA        -2 /* device runtime error (HIP here) */
#define FULLMAG_FDM_ERR_INTERNAL    -3
#define FULLMAG_FDM_ERR_INTERRUPTED -4

/* Size of one side of the inter-region exchange table (reference :44). */
#define FULLMAG_FDM_MAX_EXCHANGE_REGIONS 256

/* Enumerations (reference fullmag_fdm.h:48-78). */
typedef enum {
    FULLMAG_FDM_PRECISION_SINGLE = 1,
    FULLMAG_FDM_PRECISION_DOUBLE = 2,
} fullmag_fdm_precision;

typedef enum {
    FULLMAG_FDM_INTEGRATOR_HEUN = 1,
    FULLMAG_FDM_INTEGRATOR_DP45 = 2,
    FULLMAG_FDM_INTEGRATOR_ABM3 = 3,
    FULLMAG_FDM_INTEGRATOR_RK4  = 4,
    FULLMAG_FDM_INTEGRATOR_RK23 = 5,
} fullmag_fdm_integrator;

typedef enum {
    FULLMAG_FDM_OBSERVABLE_M       = 1,
    FULLMAG_FDM_OBSERVABLE_H_EX    = 2,
    FULLMAG_FDM_OBSERVABLE_H_DEMAG = 3,
    FULLMAG_FDM_OBSERVABLE_H_EXT   = 4,
    FULLMAG_FDM_OBSERVABLE_H_EFF   = 5,
} fullmag_fdm_observable;

typedef enum {
    FULLMAG_FDM_SNAPSHOT_SCALAR_F32 = 1,
    FULLMAG_FDM_SNAPSHOT_SCALAR_F64 = 2,
} fullmag_fdm_snapshot_scalar_type;

typedef enum {
    FULLMAG_FDM_BOUNDARY_NONE   = 0,
    FULLMAG_FDM_BOUNDARY_VOLUME = 1,
    FULLMAG_FDM_BOUNDARY_FULL   = 2,
} fullmag_fdm_boundary_correction;

/* Returns non-zero to request an interrupt (reference :80). */
typedef int (*fullmag_fdm_interrupt_poll_fn)(void *user_data);

/* Grid (reference fullmag_fdm.h:84-91). */
typedef struct {
    uint32_t nx, ny, nz;
    double   dx, dy, dz;        /* cell edge lengths, m */
} fullmag_fdm_grid_desc;

/* Material (reference fullmag_fdm.h:93-98). */
typedef struct {
    double saturation_magnetisation; /* Ms, A/m */
    double exchange_stiffness;       /* A, J/m */
    double damping;                  /* Gilbert alpha */
    double gyromagnetic_ratio;       /* gamma, m/(A s) */
} fullmag_fdm_material_desc;

/*
 * Plan descriptor (reference fullmag_fdm.h:100-269).  The field list is
 * fixed by the reference ABI; terms this engine does not evaluate are still
 * accepted in the layout and create() rejects a plan that switches them on.
 */
typedef struct {
    fullmag_fdm_grid_desc      grid;
    fullmag_fdm_material_desc  material;
    fullmag_fdm_precision      precision;
    fullmag_fdm_integrator     integrator;
    int                        disable_precession;
    int                        enable_exchange;
    int                        enable_demag;
    int                        has_external_field;
    double                     external_field_am[3];

    /* uniaxial anisotropy */
    int                        has_uniaxial_anisotropy;
    double                     uniaxial_anisotropy_constant;  /* Ku1, J/m^3 */
    double                     uniaxial_anisotropy_k2;        /* Ku2, J/m^3 */
    double                     anisotropy_axis[3];
    const double              *ku1_field;
    const double              *ku2_field;

    /* cubic anisotropy */
    int                        has_cubic_anisotropy;
    double                     cubic_Kc1;
    double                     cubic_Kc2;
    double                     cubic_Kc3;
    double                     cubic_axis1[3];
    double                     cubic_axis2[3];
    const double              *kc1_field;
    const double              *kc2_field;
    const double              *kc3_field;

    /* DMI */
    int                        has_interfacial_dmi;
    double                     dmi_D_interfacial;
    int                        has_bulk_dmi;
    double                     dmi_D_bulk;

    /* magnetoelastic coupling */
    int                        has_magnetoelastic;
    double                     mel_b1;
    double                     mel_b2;
    double                     mel_strain[6];

    double                     temperature;

    /* Zhang-Li spin-transfer torque */
    double                     current_density_x;
    double                     current_density_y;
    double                     current_density_z;
    double                     stt_degree;
    double                     stt_beta;

    /* Slonczewski spin-transfer torque */
    double                     stt_p_x;
    double                     stt_p_y;
    double                     stt_p_z;
    double                     stt_lambda;
    double                     stt_epsilon_prime;

    /* spin-orbit torque */
    int                        has_sot;
    double                     sot_je;
    double                     sot_xi_dl;
    double                     sot_xi_fl;
    double                     sot_sigma[3];
    double                     sot_thickness;

    /* Oersted field of a cylindrical conductor */
    int                        has_oersted_cylinder;
    double                     oersted_current;
    double                     oersted_radius;
    double                     oersted_center[3];
    double                     oersted_axis[3];
    uint32_t                   oersted_time_dep_kind;
    double                     oersted_time_dep_freq;
    double                     oersted_time_dep_phase;
    double                     oersted_time_dep_offset;
    double                     oersted_time_dep_t_on;
    double                     oersted_time_dep_t_off;

    /*
     * Optional Newell tensor spectra, complex interleaved (re, im) over the
     * padded grid: 2*(2nx*2ny) values for the thin-film form (nz == 1) or
     * 2*(2nx*2ny*2nz) for the 3-D form.  When absent and demag is enabled the
     * engine computes the FP64 Newell tensor itself.
     */
    const double              *demag_kernel_xx_spectrum;
    const double              *demag_kernel_yy_spectrum;
    const double              *demag_kernel_zz_spectrum;
    const double              *demag_kernel_xy_spectrum;
    const double              *demag_kernel_xz_spectrum;
    const double              *demag_kernel_yz_spectrum;
    uint64_t                   demag_kernel_spectrum_len;

    /* geometry: 1 = magnetic cell, 0 = empty */
    const uint8_t             *active_mask;
    uint64_t                   active_mask_len;

    /* region ids and the inter-region exchange table A_ij (J/m) */
    const uint32_t            *region_mask;
    uint64_t                   region_mask_len;
    const double              *exchange_lut;
    uint64_t                   exchange_lut_len;

    /* boundary-correction tiers (only NONE is evaluated by this engine) */
    fullmag_fdm_boundary_correction boundary_correction;
    double                     boundary_phi_floor;
    double                     boundary_delta_min;
    const double              *volume_fraction;
    uint64_t                   volume_fraction_len;
    const double              *face_link_xp;
    const double              *face_link_xm;
    const double              *face_link_yp;
    const double              *face_link_ym;
    const double              *face_link_zp;
    const double              *face_link_zm;
    const double              *delta_xp;
    const double              *delta_xm;
    const double              *delta_yp;
    const double              *delta_ym;
    const double              *delta_zp;
    const double              *delta_zm;
    int                        has_demag_boundary_corr;
    const int32_t             *demag_corr_target_idx;
    const int32_t             *demag_corr_source_idx;
    const double              *demag_corr_tensor;
    uint32_t                   demag_corr_target_count;
    uint32_t                   demag_corr_stencil_size;

    /* initial magnetisation, AoS, 3 * cell_count values */
    const double              *initial_magnetization_xyz;
    uint64_t                   initial_magnetization_len;

    /* adaptive step control; 0 selects the default in brackets */
    double                     adaptive_max_error;  /* [1e-5]  */
    double                     adaptive_dt_min;     /* [1e-18] */
    double                     adaptive_dt_max;     /* [1e-10] */
    double                     adaptive_headroom;   /* [0.8]   */
} fullmag_fdm_plan_desc;

/* Per-step diagnostics (reference fullmag_fdm.h:273-289). */
typedef struct {
    uint64_t step;
    double   time_seconds;
    double   dt_seconds;
    double   exchange_energy_joules;
    double   demag_energy_joules;
    double   external_energy_joules;
    double   anisotropy_energy_joules;
    double   cubic_energy_joules;
    double   dmi_energy_joules;
    double   total_energy_joules;
    double   max_effective_field_amplitude;
    double   max_demag_field_amplitude;
    double   max_rhs_amplitude;
    double   suggested_next_dt;
    uint64_t wall_time_ns;
} fullmag_fdm_step_stats;

/* Device description (reference fullmag_fdm.h:293-299). */
typedef struct {
    char name[128];
    int  compute_capability_major;   /* gfx9xx major/minor here */
    int  compute_capability_minor;
    int  driver_version;
    int  runtime_version;
} fullmag_fdm_device_info;

/* Snapshot payload description (reference fullmag_fdm.h:301-306). */
typedef struct {
    uint64_t cell_count;
    uint32_t component_count;
    uint32_t scalar_bytes;
    fullmag_fdm_snapshot_scalar_type scalar_type;
} fullmag_fdm_snapshot_desc;

typedef struct fullmag_fdm_backend fullmag_fdm_backend;
typedef struct fullmag_fdm_field_snapshot fullmag_fdm_field_snapshot;
typedef struct fullmag_fdm_preview_snapshot fullmag_fdm_preview_snapshot;

/* 1 when a usable gfx9 GPU is present (reference :320). */
int fullmag_fdm_is_available(void);

/* Validate the plan, allocate device state, upload m0 (reference :327). */
fullmag_fdm_backend *fullmag_fdm_backend_create(const fullmag_fdm_plan_desc *plan);

/* One accepted step.  For DP45 dt_seconds is the first trial step and the
 * step controller may shrink it; out_stats->suggested_next_dt carries the
 * controller's proposal (reference :335). */
int fullmag_fdm_backend_step(fullmag_fdm_backend *handle, double dt_seconds,
                             fullmag_fdm_step_stats *out_stats);

/* Poll callback, invoked on the stepping thread between device stages
 * (reference :340). */
int fullmag_fdm_backend_set_interrupt_poll(fullmag_fdm_backend *handle,
                                           fullmag_fdm_interrupt_poll_fn poll_fn,
                                           void *user_data);

/* Download an observable as AoS f64 / f32, out_len = 3*cell_count
 * (reference :349, :359). */
int fullmag_fdm_backend_copy_field_f64(fullmag_fdm_backend *handle,
                                       fullmag_fdm_observable observable,
                                       double *out_xyz, uint64_t out_len);
int fullmag_fdm_backend_copy_field_f32(fullmag_fdm_backend *handle,
                                       fullmag_fdm_observable observable,
                                       float *out_xyz, uint64_t out_len);

/* Bin-averaged preview (reference :371, :386). */
int fullmag_fdm_backend_copy_field_preview_f64(fullmag_fdm_backend *handle,
                                               fullmag_fdm_observable observable,
                                               uint32_t preview_nx, uint32_t preview_ny,
                                               uint32_t preview_nz, uint32_t z_origin,
                                               uint32_t z_stride,
                                               double *out_xyz, uint64_t out_len);
int fullmag_fdm_backend_copy_field_preview_f32(fullmag_fdm_backend *handle,
                                               fullmag_fdm_observable observable,
                                               uint32_t preview_nx, uint32_t preview_ny,
                                               uint32_t preview_nz, uint32_t z_origin,
                                               uint32_t z_stride,
                                               float *out_xyz, uint64_t out_len);

/* Asynchronous snapshots: SoA field payload / AoS preview payload owned by
 * the snapshot until destroy (reference :412, :427, :446, :455, :465, :472). */
fullmag_fdm_field_snapshot *fullmag_fdm_backend_begin_field_snapshot(
    fullmag_fdm_backend *handle, fullmag_fdm_observable observable);
fullmag_fdm_preview_snapshot *fullmag_fdm_backend_begin_preview_snapshot(
    fullmag_fdm_backend *handle, fullmag_fdm_observable observable,
    uint32_t preview_nx, uint32_t preview_ny, uint32_t preview_nz,
    uint32_t z_origin, uint32_t z_stride);
int fullmag_fdm_field_snapshot_wait(fullmag_fdm_field_snapshot *snapshot,
                                    const void **out_data, uint64_t *out_len_bytes,
                                    fullmag_fdm_snapshot_desc *out_desc);
int fullmag_fdm_preview_snapshot_wait(fullmag_fdm_preview_snapshot *snapshot,
                                      const void **out_data, uint64_t *out_len_bytes,
                                      fullmag_fdm_snapshot_desc *out_desc);
void fullmag_fdm_field_snapshot_destroy(fullmag_fdm_field_snapshot *snapshot);
void fullmag_fdm_preview_snapshot_destroy(fullmag_fdm_preview_snapshot *snapshot);

/* Replace m (normalised, inactive cells zeroed); time is not advanced
 * (reference :480, :490). */
int fullmag_fdm_backend_upload_magnetization_f64(fullmag_fdm_backend *handle,
                                                 const double *m_xyz, uint64_t len);
int fullmag_fdm_backend_upload_magnetization_f32(fullmag_fdm_backend *handle,
                                                 const float *m_xyz, uint64_t len);

/* Recompute fields for the current m (reference :499, :506). */
int fullmag_fdm_backend_refresh_observables(fullmag_fdm_backend *handle);
int fullmag_fdm_backend_refresh_demag_observable(fullmag_fdm_backend *handle);

/* Diagnostics of the current state without stepping (reference :516). */
int fullmag_fdm_backend_snapshot_stats(fullmag_fdm_backend *handle,
                                       fullmag_fdm_step_stats *out_stats);

/* Device description (reference :523). */
int fullmag_fdm_backend_get_device_info(fullmag_fdm_backend *handle,
                                        fullmag_fdm_device_info *out_info);

/* Last error text or NULL; valid until the next call on the handle
 * (reference :531). */
const char *fullmag_fdm_backend_last_error(fullmag_fdm_backend *handle);

/* Free everything; NULL is accepted (reference :538). */
void fullmag_fdm_backend_destroy(fullmag_fdm_backend *handle);

#ifdef __cplusplus
}
#endif

#endif /* FULLMAG_FDM_H */
