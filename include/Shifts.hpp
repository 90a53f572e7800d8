// Shifts.hpp -- per-channel dispersion delay tables.
// Same signatures and ownership as the reference (include/Shifts.hpp:26-27
// of TRASAL/Dedispersion): the caller owns the returned vector.
#pragma once

#include <vector>

#include <Observation.hpp>

namespace Dedispersion {

// Delay of every channel relative to the highest frequency, in samples per
// unit DM: ((4148.808 * (1/f^2 - 1/f_max^2)) * (S / downsampling)) / (S * t),
// f = minFreq + channel * bandwidth, last channel forced to 0, length
// pad(C, padding/4), padding entries 0.  (src/Shifts.cpp:20-30; the square
// is x * x.)
std::vector<float> *getShifts(AstroData::Observation &observation, const unsigned int padding);
// The same over subbands (f = subbandMinFreq + subband * subbandBandwidth,
// f_max = subbandMaxFreq), length pad(Nsb, padding/4).  (src/Shifts.cpp:32-42)
std::vector<float> *getShiftsStepTwo(AstroData::Observation &observation, const unsigned int padding);

} // namespace Dedispersion
