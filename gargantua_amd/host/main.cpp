// The gargantua command line (main.swift, Program.swift:8-23): parse, then
// dispatch. Usage errors exit 64, operational errors 2, failed validation 1.
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/gargantua_constants.h"
#include "host.hpp"

namespace host {

// `gargantua constants`: the constants table the kernels are built with.
void print_constants() {
#define C(name) printf("#define %-30s %.17g\n", #name, (double)(name));
    C(R_SCHWARZSCHILD) C(R_PHOTON_SPHERE) C(R_ISCO) C(R_MARGINALLY_BOUND) C(B_CRITICAL)
    C(WEAK_DEFLECTION_COEFFICIENT) C(R_CAPTURE) C(KERR_CAPTURE_MARGIN) C(R_ESCAPE) C(R_DISK_OUTER)
    C(R_FEED_INNER) C(DISK_SLAB_HALF_HEIGHT) C(VOLUME_HALF_EXTENT) C(SIM_DT) C(SIM_SUBSTEPS) C(DRAG_ALPHA)
    C(PLUNGE_SPEED_LIMIT) C(FEED_VELOCITY_DISPERSION) C(DISK_ASPECT) C(T_PEAK_DEFAULT) C(ZERO_TORQUE_FLOOR)
    C(CIC_CORNERS) C(BLACKBODY_T_MIN) C(BLACKBODY_T_MAX) C(BLACKBODY_TABLE_SIZE) C(STEP_CAP_INTERACTIVE)
    C(STEP_CAP_STILL) C(STEP_RADIUS_FACTOR) C(STEP_MIN) C(STEP_MAX) C(STEP_FACTOR_EMPTY_INTERACTIVE)
    C(STEP_MAX_EMPTY_INTERACTIVE) C(STEP_FACTOR_EMPTY_STILL) C(STEP_MAX_EMPTY_STILL) C(BAKE_SAMPLES)
    C(BAKE_SPACING_VOXELS) C(BAKE_FRAMES) C(SWEEP_TABLE_SIZE) C(DRIFT_BUDGET_INTERACTIVE) C(DRIFT_BUDGET_STILL)
    C(EXPOSURE_DEFAULT) C(EMISSION_SCALE) C(GAS_SPEED_CEILING) C(BEAMING_DENSITY_FLOOR) C(VOLUME_OPACITY)
    C(STAR_BRIGHTNESS) C(STILL_SETTLE_STEPS) C(STILL_TILE) C(MEMORY_BUDGET_FRACTION)
#undef C
}

}  // namespace host

int main(int argc, char** argv) {
    using namespace host;
    std::vector<std::string> args(argv + 1, argv + argc);
    Options o;
    std::string err = parse_options(args, o);
    if (!err.empty()) usage_error(err);
    switch (o.command) {
        case Command::version: line(garg_version()); return 0;
        case Command::constants: print_constants(); return 0;
        case Command::interactive:
            operational("The interactive window is macOS only; on this GPU use --still, bench, validate or serve.");
        case Command::still: run_still(o);
        case Command::bench: run_bench(o);
        case Command::validate: run_validate(o);
        case Command::serve: run_serve(o);
    }
    return 0;
}
