/*
 * FP64 CPU oracle for the gargantua per-frame passes.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (gargantua_amd/, the
 * C-ABI library, the CLI) links, loads or calls this code. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only
 * as the checker or as the timed CPU baseline, never as the thing measured.
 *
 * It is a from-scratch restatement of the reference's Swift oracle
 * (the Swift files under Sources/Oracle) plus the kernel-only semantics the Swift oracle
 * lacks (fixed point deposit, resolve, sphere jump, gather, starfield),
 * each function citing the reference file:line it follows. The reference's
 * own oracle cannot be built here (Swift is absent), so this restatement is
 * pinned by the reference's known-answer tests and goldens and by closed
 * forms; tests/test_oracle_*.py hold those checks. Functions with no such
 * check are listed as "parity unpinned" in oracle/README.md.
 */
#ifndef GARGANTUA_ORACLE_H
#define GARGANTUA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Hash (Oracle/Hash.swift) ---------------------------------------- */
void orc_pcg4d(const uint32_t in[4], uint32_t out[4]);
void orc_hash_draw(uint64_t seed, uint32_t index, uint32_t event, uint32_t lane, uint32_t out[4]);
double orc_unit(uint32_t x);
void orc_gaussian_pair(double u1, double u2, double out[2]);
uint64_t orc_fnv1a(const void* bytes, uint64_t count);

/* ---- Potential and dynamics (Oracle/Potential.swift, ParticleDynamics.swift) */
double orc_potential_energy(double r, int potential);
double orc_potential_gradient(double r, int potential);
double orc_circular_speed(double r, int potential);
double orc_angular_frequency(double r, int potential);
double orc_circular_energy(double r, int potential);
double orc_circular_angular_momentum(double r, int potential);
/* state = {x, y, z, vx, vy, vz}; one semi-implicit Euler step with drag. */
void orc_particle_step(double state[6], int potential, double dt, double alpha);
int orc_particle_removed(const double before[3], const double after[6]);
double orc_specific_energy(const double state[6], int potential);

/* ---- Disk model and edge (Oracle/DiskModel.swift, DiskEdge.swift) ----- */
double orc_disk_shape(double r, double floor);
double orc_disk_temperature(double r, double peak);
double orc_steady_surface_density(double r);
double orc_radial_number_density(double r);
void orc_radial_inverse_cdf(int count, int resolution, double* out);
/* generator 0: reference spawn (event 0 from the inverse CDF, later events
 * in the annulus, z = 0.05 r N); generator 1: annulus at every event with
 * z = sigma_z N (the benchmark scene generator). */
void orc_spawn(uint32_t index, uint32_t event, uint64_t seed, const double* cdf, int cdf_count,
               int potential, int generator, double sigma_z, double out[6]);
double orc_inner_edge(const double* radii, int64_t n, double bin_width, double fit_inner, double fit_outer);
/* Full FP64 disk run as the isco golden does it (GoldenTests.swift:39-64):
 * returns cylindrical radii after `substeps` substeps. */
void orc_disk_run(uint32_t count, uint64_t seed, int substeps, int threads, double* radii_out, uint64_t* respawns_out);

/* ---- Blackbody (Oracle/Blackbody.swift) --------------------------------- */
void orc_blackbody_xyz(double temperature, double out[3]);
void orc_blackbody_chromaticity(double temperature, double out[2]);
void orc_blackbody_color(double temperature, double out[3]);
void orc_blackbody_table(int count, double tmin, double tmax, double* out_rgb);
double orc_blackbody_coordinate(double temperature);

/* ---- Schwarzschild (Oracle/Schwarzschild.swift) ------------------------- */
typedef struct {
    double v, r, phi, vdot, rdot, phidot;
} orc_ray_state;

typedef struct {
    orc_ray_state s;
    double e1[3], e2[3];
    double energy, angular_momentum;
} orc_ray;

typedef struct {
    double step_factor, step_min, step_max, capture_radius, escape_radius;
    int64_t step_cap;
    double empty_step_factor, empty_step_max;
} orc_settings;

typedef struct {
    orc_ray ray;
    int outcome; /* 0 captured, 1 escaped, 2 exhausted */
    int64_t steps;
    double drift;
} orc_result;

void orc_settings_interactive(orc_settings* s);
void orc_settings_still(orc_settings* s);
double orc_schw_f(double r);
void orc_schw_launch(const double origin[3], const double direction[3], orc_ray* out);
void orc_schw_launch_b(double radius, double b, orc_ray* out);
void orc_schw_step(orc_ray_state* s, double h);
double orc_schw_null_residual(const orc_ray_state* s);
double orc_schw_energy(const orc_ray_state* s);
double orc_schw_angular_momentum(const orc_ray_state* s);
void orc_schw_position(const orc_ray* ray, double out[3]);
void orc_schw_direction(const orc_ray* ray, double out[3]);
void orc_schw_integrate(const orc_ray* launch, const orc_settings* settings, orc_result* out);
double orc_schw_deflection(const orc_ray* launch, const orc_result* result);
double orc_schw_critical_b(double launch_radius, const orc_settings* settings, int iterations);
double orc_redshift_factor(double f, const double direction[3], const double gas[3]);

/* ---- Deflection (Oracle/Deflection.swift) ------------------------------- */
double orc_max_impact_parameter(double r);
double orc_sweep(double r, double b, int intervals);
double orc_table_b(double low, double high, double t);
double orc_table_coordinate(double low, double high, double b);
double orc_periapsis(double b);
double orc_turning_sweep(double r, double b, int intervals);
double orc_escape_sweep(double r, double b, int intervals);
void orc_sweep_table(double r, int count, int intervals, double* out);
void orc_sky_table(double camera_r, double sphere_r, int count, int intervals, double* out);
/* Log-spaced FP64 far-field table: for b_k = b_lo (b_hi / b_lo)^(k/(n-1)),
 * alpha = total deflection of a ray from infinity (2 S(r_p, b) - pi, with
 * S(r_p, b) the sweep from periapsis) and sweep = S(r_obs, b), the sweep
 * from the observer radius to infinity. */
void orc_far_field_table(double r_obs, int count, double b_lo, double b_hi, int intervals,
                         double* b_out, double* alpha_out, double* sweep_out);
/* Strong deflection limit (Bozza 2002, Phys. Rev. D 66, 103001). */
double orc_strong_deflection(double b);

/* ---- Kerr (Oracle/KerrGeodesic.swift) ----------------------------------- */
typedef struct {
    double r, theta, phi, pr, ptheta, energy, angular_momentum;
} orc_kerr_state;

typedef struct {
    double step_factor, step_min, step_max, capture_margin, escape_radius;
    int64_t step_cap;
} orc_kerr_settings;

typedef struct {
    orc_kerr_state state;
    int outcome;
    int64_t steps;
    double drift;
} orc_kerr_result;

void orc_kerr_settings_default(orc_kerr_settings* s, int64_t step_cap);
double orc_kerr_outer_horizon(double spin);
void orc_kerr_launch(double spin, const double origin[3], const double direction[3], orc_kerr_state* out);
void orc_kerr_step(double spin, orc_kerr_state* s, double h);
double orc_kerr_step_length(double spin, const orc_kerr_state* s, const orc_kerr_settings* settings);
double orc_kerr_carter(double spin, const orc_kerr_state* s);
double orc_kerr_null_residual(double spin, const orc_kerr_state* s);
void orc_kerr_position(double spin, const orc_kerr_state* s, double out[3]);
void orc_kerr_direction(double spin, const orc_kerr_state* s, double out[3]);
void orc_kerr_integrate(double spin, const orc_kerr_state* launch, const orc_kerr_settings* settings, orc_kerr_result* out);
double orc_kerr_radius_of(double spin, const double point[3]);
double orc_kerr_bisect_critical_xi(double spin, double radius, double side, const orc_kerr_settings* settings, int iterations);
double orc_kerr_critical_xi(double spin, double rc);

/* ---- Kerr in Cartesian Kerr-Schild coordinates (kerr_schild.cpp) --------- */
typedef struct {
    double x, y, z, px, py, pz; /* position and covariant spatial momentum */
    double energy;              /* -p_t */
    double phase;               /* int a/Delta dr at the launch radius */
} orc_ks_state;

void orc_ks_metric(double spin, double x, double y, double z, double g[16]);
double orc_ks_radius(double spin, double x, double y, double z);
void orc_ks_from_bl(double spin, const orc_kerr_state* s, orc_ks_state* out);
void orc_ks_to_bl(double spin, const orc_ks_state* s, orc_kerr_state* out);
double orc_ks_null_residual(double spin, const orc_ks_state* s);
void orc_ks_step(double spin, orc_ks_state* s, double h);
double orc_ks_step_length(double spin, const orc_ks_state* s, const orc_kerr_settings* settings);
void orc_ks_position(double spin, const orc_ks_state* s, double out[3]);
double orc_ks_drift(double spin, const orc_ks_state* s, double carter0, double l0, double scale);
void orc_ks_integrate(double spin, const orc_ks_state* launch, const orc_kerr_settings* settings, orc_kerr_result* out);
double orc_kerr_critical_eta(double spin, double rc);
void orc_kerr_photon_orbits(double spin, double out[2]);

/* ---- Volume field and lensed render (Oracle/VolumeField.swift, LensedRender.swift) */
typedef struct {
    int size;
    double half_extent;    /* x and y */
    double half_extent_z;  /* z; equal to half_extent for the reference cube */
    const double* emission; /* size^3 x 4, rgb + density, x fastest */
    const double* velocity; /* size^3 x 3 */
} orc_field;

void orc_field_sample(const orc_field* field, const double position[3], double emission_out[4], double velocity_out[3]);
double orc_luminance(const double rgb[3]);
double orc_side_ratio(const double* luminance, int width, int height);

typedef struct {
    double position[3], right[3], up[3], forward[3];
    double tan_half_fov[2];
} orc_camera;

void orc_camera_direction(const orc_camera* c, int px, int py, int width, int height, double out[3]);
/* LensedRender.trace (LensedRender.swift:46-78): gas only, no stars. */
void orc_trace(const double origin[3], const double direction[3], const orc_field* field,
               const orc_settings* settings, double opacity, double slab, int redshift, double out[3]);
void orc_render(int width, int height, const orc_camera* camera, const orc_field* field,
                const orc_settings* settings, double opacity, int redshift, int threads, double* out_rgb);

/* ---- Full pixel shading as the march kernel does it (march.metal) ------ */
/* Uniform block in double precision mirroring MarchUniforms. */
typedef struct {
    double camera_position[3], camera_right[3], camera_up[3], camera_forward[3];
    double tan_half_fov[2];
    int width, height;
    double volume_half_extent, volume_half_extent_z;
    double opacity_scale, star_brightness, drift_budget;
    uint32_t star_seed;
    int redshift;
    double sphere_radius, camera_radius, sphere_max_b, sphere_table_max_b, camera_max_b, sky_min_b, sky_max_b;
    double spin;
    orc_settings geodesic;
    /* sweep tables, SWEEP_TABLE_SIZE entries each, FP64 */
    const double* sphere_table;
    const double* camera_table;
    const double* sky_table;
    const double* blackbody; /* BLACKBODY_TABLE_SIZE x 3 */
} orc_march_params;

typedef struct {
    double color[3];
    double transmittance;
    double drift;
    double depth;
    int64_t steps;
    int outcome;
    int falls_in;
    double sky[3];          /* sky direction used for the starfield */
} orc_pixel;

/* One pixel of marchImage (march.metal:194-288); field may be NULL (no gas). */
void orc_march_pixel(const orc_march_params* p, const orc_field* field, int px, int py, orc_pixel* out);
/* A batch of pixels, multi-threaded; returns total steps. */
int64_t orc_march_pixels(const orc_march_params* p, const orc_field* field, const int32_t* pixels_xy, int64_t count,
                         int threads, orc_pixel* out);
/* One probe ray (march.metal:290-341): results in the probe layout. */
void orc_probe(const orc_march_params* p, const double origin[3], const double direction[3],
               double position_out[3], double direction_out[3], double* drift, int* outcome, int64_t* steps,
               double* energy, double* angular_momentum);
void orc_starfield(const double direction[3], uint32_t seed, const double* blackbody, double out[3]);

/* ---- Bake and walk (bake.metal:17-92, 94-144) ---------------------------- */
typedef struct {
    int64_t count;          /* samples stored (at most the capacity) */
    int falls_in;
    int overflowed;         /* a sample fell beyond the capacity */
    double direction[3];    /* sky direction of the ray's end */
    double depth;           /* min(depth / (2 R_ESCAPE), 1): path to the first sample, or all of it */
} orc_baked_ray;
/* bakeRays for one pixel with the Schwarzschild march of p (spin 0):
 * samples[4 k .. 4 k + 3] = (position, path since the previous sample) of
 * sample k, for k < header->count; `capacity` and `spacing` are the
 * uniforms' bakeCapacity and bakeSpacing. */
void orc_bake_pixel(const orc_march_params* p, int px, int py, int64_t capacity, double spacing, double* samples,
                    orc_baked_ray* header);
/* walkRays for one pixel: composite `count` stored samples against `field`
 * from the camera position of p, then the starfield along `direction` unless
 * the ray falls in. */
void orc_walk_pixel(const orc_march_params* p, const orc_field* field, const double* samples, int64_t count,
                    int falls_in, const double direction[3], double color_out[3]);
void orc_pbr_neutral(const double in[3], double out[3]);

/* ---- Exact replays of the fp32 integer deposit (splat.metal:16-88) ------ */
/* particles: AoS 32 B records (float pos[3], vel[3], temperature, uint32 respawns).
 * emission: 2 x size^3 u32, momentum: 3 x size^3 i32, mass: size^3 u32, all accumulated. */
void orc_splat_replay(const void* particles, uint32_t count, uint32_t size, float half_extent, float half_extent_z,
                      float peak_temperature, float emission_scale, uint32_t exact, uint32_t frame, uint64_t seed,
                      uint32_t* emission, int32_t* momentum, uint32_t* mass);
/* One fp32 sim dispatch replay (sim.metal:44-79) for particles that do not
 * respawn; sets respawned[i] = 1 where the kernel would respawn. */
void orc_sim_replay_step(void* particles, uint32_t count, int potential, float dt, float alpha, float temperature_scale,
                         uint8_t* respawned);

/* ---- Tone curve (present.metal:19-35) -------------------------------- */
double orc_srgb_encode(double linear);

#ifdef __cplusplus
}
#endif

#endif
