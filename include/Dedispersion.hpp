// Dedispersion.hpp -- C++ API of the MI355X dedispersion engine.
//
// Keeps the reference's C++ names (include/Dedispersion.hpp of
// TRASAL/Dedispersion): DedispersionConf (:34-54), tunedDedispersionConf
// (:56) and readTunedDedispersionConf (:66).  The OpenCL source generators
// getDedispersionOpenCL / getSubbandDedispersionStepOneOpenCL /
// getSubbandDedispersionStepTwoOpenCL (:63-65), whose output a caller
// compiled and launched, are replaced by get...HIP factories returning a
// kernel object that launches the hand-written HIP kernels on a stream.
// Everything here is a thin layer over the C ABI in dedispersion_amd.h;
// errors become exceptions (Dedispersion::HIPError).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <Observation.hpp>
#include <dedispersion_amd.h>

namespace Dedispersion {

class HIPError : public std::runtime_error {
public:
  HIPError(dd_status status, const std::string &message)
      : std::runtime_error(message), status_(status) {}
  dd_status status() const { return status_; }

private:
  dd_status status_;
};

inline void check(dd_status status) {
  if (status != DD_OK) throw HIPError(status, dd_last_error());
}

// Thread/item geometry of a kernel configuration (the reference inherits it
// from isa::OpenCL::KernelConf).  print() writes "t0 t1 t2 i0 i1 i2".
class KernelConf {
public:
  unsigned int getNrThreadsD0() const { return nrThreadsD0; }
  unsigned int getNrThreadsD1() const { return nrThreadsD1; }
  unsigned int getNrThreadsD2() const { return nrThreadsD2; }
  unsigned int getNrItemsD0() const { return nrItemsD0; }
  unsigned int getNrItemsD1() const { return nrItemsD1; }
  unsigned int getNrItemsD2() const { return nrItemsD2; }
  void setNrThreadsD0(unsigned int v) { nrThreadsD0 = v; }
  void setNrThreadsD1(unsigned int v) { nrThreadsD1 = v; }
  void setNrThreadsD2(unsigned int v) { nrThreadsD2 = v; }
  void setNrItemsD0(unsigned int v) { nrItemsD0 = v; }
  void setNrItemsD1(unsigned int v) { nrItemsD1 = v; }
  void setNrItemsD2(unsigned int v) { nrItemsD2 = v; }
  std::string print() const;

private:
  unsigned int nrThreadsD0 = 64;
  unsigned int nrThreadsD1 = 4;
  unsigned int nrThreadsD2 = 1;
  unsigned int nrItemsD0 = 16;
  unsigned int nrItemsD1 = 4;
  unsigned int nrItemsD2 = 1;
};

// include/Dedispersion.hpp:34-54, src/Dedispersion.cpp:83-89.
class DedispersionConf : public KernelConf {
public:
  DedispersionConf();
  ~DedispersionConf();
  bool getSplitBatches() const { return splitBatches; }
  bool getLocalMem() const { return local != 0; }
  unsigned int getUnroll() const { return unroll; }
  void setSplitBatches(bool split) { splitBatches = split; }
  void setLocalMem(bool local) { this->local = local ? 1 : 0; }
  void setUnroll(unsigned int unroll) { this->unroll = unroll; }
  // The kernel selection behind the local flag: dd_conf.local (0 = caches,
  // 1 = pipelined LDS staging, 2 = LDS staging one group at a time,
  // 3 = channel-block patterns); getLocalMem() is true for 1, 2 and 3.
  unsigned int getLocalMode() const { return local; }
  void setLocalMode(unsigned int mode) { local = mode; }
  // "split local unroll t0 t1 t2 i0 i1 i2" (src/Dedispersion.cpp:88); local
  // is the local mode.
  std::string print() const;

  dd_conf toC() const;
  static DedispersionConf fromC(const dd_conf &conf);

private:
  bool splitBatches;
  unsigned int local;
  unsigned int unroll;
};

// include/Dedispersion.hpp:56
typedef std::map<std::string, std::map<unsigned int, Dedispersion::DedispersionConf *> *> tunedDedispersionConf;

// src/Dedispersion.cpp:20-81: lines starting with a letter are
// "device nrDMs split local unroll t0 t1 t2 i0 i1 i2"; entries are heap
// allocated and owned by the caller; the first line for a (device, nrDMs)
// is kept (later ones are dropped, which the reference leaks).  local is read
// as the local mode (0-3; the reference reads a bool).  Throws
// AstroData::FileError if the file cannot be opened.
void readTunedDedispersionConf(tunedDedispersionConf &tunedDedispersion, const std::string &dedispersionFilename);

// Observation -> C descriptor (dispersed sizes taken from the observation
// when set, else derived by the reference's sizing rules).
dd_observation toC(const AstroData::Observation &observation);

// A plan-backed kernel: the HIP counterpart of a compiled reference kernel.
class HIPKernel {
public:
  HIPKernel(const AstroData::Observation &observation, unsigned int padding) {
    dd_observation obs = toC(observation);
    dd_plan *p = nullptr;
    check(dd_plan_create(&p, &obs, padding));
    plan_.reset(p);
  }
  dd_plan *plan() const { return plan_.get(); }
  dd_sizes sizes() const {
    dd_sizes s;
    check(dd_plan_get_sizes(plan_.get(), &s));
    return s;
  }
  // The reference's inputBits: 1, 2, 4 (packed input rows) or 8.
  void setInputBits(unsigned int inputBits) { check(dd_plan_set_input_bits(plan_.get(), inputBits)); }
  void setConf(dd_mode mode, const DedispersionConf &conf) {
    dd_conf c = conf.toC();
    check(dd_plan_set_conf(plan_.get(), mode, &c));
  }
  void setZappedChannels(const std::vector<unsigned int> &zapped) {
    check(dd_plan_set_zapped_channels(plan_.get(), zapped.data()));
  }
  void setBeamMapping(const std::vector<unsigned int> *singleStep, const std::vector<unsigned int> *stepTwo) {
    check(dd_plan_set_beam_mapping(plan_.get(), singleStep ? singleStep->data() : nullptr,
                                   stepTwo ? stepTwo->data() : nullptr));
  }
  void setShifts(const std::vector<float> *shifts, const std::vector<float> *shiftsStepTwo) {
    check(dd_plan_set_shifts(plan_.get(), shifts ? shifts->data() : nullptr,
                             shiftsStepTwo ? shiftsStepTwo->data() : nullptr));
  }

private:
  struct Deleter {
    void operator()(dd_plan *p) const { dd_plan_destroy(p); }
  };
  std::unique_ptr<dd_plan, Deleter> plan_;
};

// Replaces the `dedispersion` kernel: launch(input, output, firstSynthesizedBeam).
class DedispersionKernel : public HIPKernel {
public:
  using HIPKernel::HIPKernel;
  void launch(const uint8_t *input, float *output, unsigned int firstSynthesizedBeam = 0,
              unsigned int nrSynthesizedBeams = 0, void *stream = nullptr) {
    check(dd_dedisperse(plan(), input, output, firstSynthesizedBeam, nrSynthesizedBeams, stream));
  }
};

// Replaces `dedispersionStepOne`: launch(input, output).
class SubbandStepOneKernel : public HIPKernel {
public:
  using HIPKernel::HIPKernel;
  void launch(const uint8_t *input, float *output, void *stream = nullptr) {
    check(dd_subband_step_one(plan(), input, output, stream));
  }
};

// Replaces `dedispersionStepTwo`: launch(input, output, firstSynthesizedBeam).
class SubbandStepTwoKernel : public HIPKernel {
public:
  using HIPKernel::HIPKernel;
  void launch(const float *input, float *output, unsigned int firstSynthesizedBeam = 0,
              unsigned int nrSynthesizedBeams = 0, void *stream = nullptr) {
    check(dd_subband_step_two(plan(), input, output, firstSynthesizedBeam, nrSynthesizedBeams, stream));
  }
};

// Factories with the generators' argument lists (include/Dedispersion.hpp:63-65).
// I must be uint8_t and O float (include/configuration.hpp:9-15); inputBits
// is 1, 2, 4 (packed rows, see dd_plan_set_input_bits) or 8, else throws.
template <typename I, typename O>
DedispersionKernel *getDedispersionHIP(const DedispersionConf &conf, const unsigned int padding,
                                       const uint8_t inputBits, const AstroData::Observation &observation,
                                       std::vector<float> &shifts) {
  static_assert(sizeof(I) == 1 && sizeof(O) == 4, "uint8 input and float output only");
  DedispersionKernel *k = new DedispersionKernel(observation, padding);
  try {
    k->setInputBits(inputBits);
  } catch (...) {
    delete k;
    throw;
  }
  k->setShifts(&shifts, nullptr);
  k->setConf(DD_SINGLE_STEP, conf);
  return k;
}

template <typename I, typename O>
SubbandStepOneKernel *getSubbandDedispersionStepOneHIP(const DedispersionConf &conf, const unsigned int padding,
                                                       const uint8_t inputBits,
                                                       const AstroData::Observation &observation,
                                                       std::vector<float> &shifts) {
  static_assert(sizeof(I) == 1 && sizeof(O) == 4, "uint8 input and float output only");
  SubbandStepOneKernel *k = new SubbandStepOneKernel(observation, padding);
  try {
    k->setInputBits(inputBits);
  } catch (...) {
    delete k;
    throw;
  }
  k->setShifts(&shifts, nullptr);
  k->setConf(DD_STEP_ONE, conf);
  return k;
}

template <typename I>
SubbandStepTwoKernel *getSubbandDedispersionStepTwoHIP(const DedispersionConf &conf, const unsigned int padding,
                                                       const AstroData::Observation &observation,
                                                       std::vector<float> &shifts) {
  static_assert(sizeof(I) == 4, "float intermediate only");
  SubbandStepTwoKernel *k = new SubbandStepTwoKernel(observation, padding);
  k->setShifts(nullptr, &shifts);
  k->setConf(DD_STEP_TWO, conf);
  return k;
}

} // namespace Dedispersion
