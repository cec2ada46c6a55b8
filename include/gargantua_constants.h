/*
 * Every numeric constant of the per-frame passes, in one table shared by the
 * HIP kernels, the C-ABI host code and the FP64 oracle.
 *
 * Units are geometrized, G = c = M = 1, so every length is in units of the
 * black hole mass. Values and their meaning follow the reference's single
 * source of truth, Sources/Oracle/Constants.swift (line given per entry),
 * which the reference renders into Shaders/Constants.h. Here the values are
 * written once as double literals; device code takes GARG_F(NAME), the
 * float nearest to that double, which for every entry equals the float
 * nearest to the decimal literal (checked by tests/test_constants.py).
 */
#ifndef GARGANTUA_CONSTANTS_H
#define GARGANTUA_CONSTANTS_H

#define GARG_F(name) ((float)(name))

/* Schwarzschild geometry (Constants.swift:56-84). */
#define R_SCHWARZSCHILD 2.0                 /* horizon, r = 2M */
#define R_PHOTON_SPHERE 3.0                 /* unstable circular photon orbit */
#define R_ISCO 6.0                          /* innermost stable circular orbit, a = 0 */
#define R_MARGINALLY_BOUND 4.0              /* marginally bound circular orbit */
#define B_CRITICAL 5.196152422706632        /* 3 sqrt(3), shadow radius */
#define WEAK_DEFLECTION_COEFFICIENT 4.0     /* deflection 4M / b for b >> M */

/* Scene radii (Constants.swift:88-121). */
#define R_CAPTURE 2.02                      /* rays and particles end here */
#define KERR_CAPTURE_MARGIN 0.05            /* Kerr rays end this far above r+ */
/* Largest radius of a spinning ray (escape radius, camera, probe origins):
 * the fast forms take one reciprocal of products growing like r^6, which
 * stays far inside float range there. This build's own bound; the
 * reference's scenes stay within r = 60. */
#define KERR_MAX_RADIUS 1.0e5
#define R_ESCAPE 60.0                       /* escaped beyond this radius */
#define R_DISK_OUTER 24.0                   /* outer edge of seeded disk and annulus */
#define R_FEED_INNER 20.0                   /* inner edge of the feeding annulus */
#define DISK_SLAB_HALF_HEIGHT 6.0           /* rays sample only within |z| < 6 */
#define VOLUME_HALF_EXTENT 30.0             /* half width of the splat cube */

/* Particle dynamics (Constants.swift:125-153). */
#define SIM_DT 0.1                          /* fixed timestep per substep */
#define SIM_SUBSTEPS 2u                     /* substeps per step dispatch */
#define DRAG_ALPHA 0.003                    /* drag rate as a fraction of Omega */
#define PLUNGE_SPEED_LIMIT 12.0             /* faster particles count as captured */
#define FEED_VELOCITY_DISPERSION 0.02       /* Gaussian velocity spread / v_c */
#define DISK_ASPECT 0.05                    /* vertical scale height / r */

/* Temperature and colour (Constants.swift:157-185). */
#define T_PEAK_DEFAULT 8000.0               /* peak of the disk temperature profile, K */
#define ZERO_TORQUE_FLOOR 0.01              /* floor on 1 - sqrt(r_isco / r) */
#define CIC_CORNERS 8u                      /* cloud in cell corners */
#define BLACKBODY_T_MIN 1000.0              /* lowest tabulated temperature, K */
#define BLACKBODY_T_MAX 40000.0             /* highest tabulated temperature, K */
#define BLACKBODY_TABLE_SIZE 256u           /* entries, log spaced */

/* Ray marching (Constants.swift:189-265). */
#define STEP_CAP_INTERACTIVE 256u
#define STEP_CAP_STILL 4096u
#define STEP_RADIUS_FACTOR 0.02             /* h = 0.02 r, clamped */
#define STEP_MIN 0.02
#define STEP_MAX 0.5
#define STEP_FACTOR_EMPTY_INTERACTIVE 0.1
#define STEP_MAX_EMPTY_INTERACTIVE 3.0
#define STEP_FACTOR_EMPTY_STILL 0.06
#define STEP_MAX_EMPTY_STILL 1.5
#define MARCH_THREADGROUP_WIDTH 8u
#define MARCH_THREADGROUP_HEIGHT 8u
#define BAKE_SAMPLES 96u                    /* stored samples per ray */
#define BAKE_SPACING_VOXELS 0.5             /* minimum path between samples, voxels */
#define BAKE_FRAMES 4u                      /* frames a progressive rebake spans */
#define SWEEP_TABLE_SIZE 1024u              /* entries per sweep table */
#define DRIFT_BUDGET_INTERACTIVE 1e-4
#define DRIFT_BUDGET_STILL 1e-5

/* Rendering (Constants.swift:269-307). */
#define EXPOSURE_DEFAULT 1.0
#define EMISSION_SCALE 0.0005               /* emission per particle at T_peak */
#define GAS_SPEED_CEILING 0.85              /* beta = 0.85 tanh(v / 0.85) */
#define BEAMING_DENSITY_FLOOR 40.0          /* Doppler blends in with min(rho / 40, 1) */
#define VOLUME_OPACITY 0.0005               /* extinction per unit path per density */
#define STAR_BRIGHTNESS 1.5
#define STILL_SETTLE_STEPS 600u             /* substeps before a still */
#define STILL_TILE 128u

/* Safety (Constants.swift:311-324). */
#define COMMAND_BUFFER_BUDGET_MS 30.0
#define MEMORY_BUDGET_FRACTION 0.4
#define MEMORY_BUDGET_CAP_BYTES 3000000000u

/*
 * Values that live outside Constants.swift in the reference but fix results:
 * the inverse CDF length (Presets.swift:45), the Kerr oracle's default
 * minimum step (KerrGeodesic.swift:176-197; the kernel floor is 0.002,
 * Kerr.h:186), the fixed point scales of the deposit (splat.metal:16-19),
 * the star cube map (Geodesic.h:190-215) and the sweep table integration
 * intervals (DeflectionTables.swift:12, Deflection.swift:23).
 */
#define INVERSE_CDF_COUNT 1024u
#define INVERSE_CDF_RESOLUTION 8192u
#define KERR_STEP_FLOOR 0.002
#define KERR_POLE_SINE 0.08
#define ENERGY_FIXED 16777216.0             /* 2^24 */
#define ENERGY_T_FIXED 1048576.0            /* 2^20 */
#define MOMENTUM_FIXED 4096.0               /* 2^12 */
#define MASS_FIXED 256.0                    /* 2^8 */
#define STAR_CELLS 96.0
#define STAR_THRESHOLD 0x30000000u
#define SWEEP_INTERVALS_SPHERE 2048u
#define SWEEP_INTERVALS_CAMERA 256u
#define TRANSMITTANCE_FLOOR 0.004           /* stop gathering below this */

#endif
