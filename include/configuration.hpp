// configuration.hpp -- data types of the path (include/configuration.hpp:9-15
// of TRASAL/Dedispersion): 8-bit unsigned input, float intermediate and
// float output.
#pragma once

#include <cstdint>
#include <string>

typedef uint8_t inputDataType;
const std::string inputDataName("uchar");
const uint8_t inputBits = 8;
typedef float intermediateDataType;
const std::string intermediateDataName("float");
typedef float outputDataType;
const std::string outputDataName("float");
