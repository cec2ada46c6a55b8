// Host-built FP64 tables of a context (blackbody colours, inverse CDF).
#pragma once

#include <vector>

namespace garg {

// BLACKBODY_TABLE_SIZE entries of linear sRGB (peak channel 1) plus 1, log
// spaced in temperature between BLACKBODY_T_MIN and BLACKBODY_T_MAX.
void build_blackbody_table(std::vector<float>& rgba);
// INVERSE_CDF_COUNT radii; entry k maps the uniform deviate k / (count - 1).
void build_inverse_cdf(std::vector<float>& table);
// Peak temperature over the unnormalized disk profile at its peak.
double temperature_scale();

}  // namespace garg
