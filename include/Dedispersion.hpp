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
#include <type_traits>
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

// ---- the sequential CPU functions (include/Dedispersion.hpp:59-61) ----
//
// The reference exports its CPU loops as templates over the input (I),
// accumulator (L) and output (O) types; callers use them as a fallback or
// as the checker of a device path.  These have the same signatures and the
// same arithmetic: for every output sample the live rows are added in order
// into an L accumulator that starts at 0, and a row's shift is
// (unsigned)((firstDM + dm * DMStep) * delay) with every operation rounded
// separately (DH:81, 124, 168).  The rounding is forced through volatile
// temporaries, so the result does not depend on whether the caller's
// compiler contracts multiply-adds.  Sizes come from the observation's
// getters, dispersed sizes included (set them as src/DedispersionTest.cpp:
// 129, 145, 146 does).  Layouts are those of the device entry points
// (dedispersion_amd.h), which keep the reference's for 8-bit and float
// input; where the reference's own indexing is defective they follow the
// device layout instead:
//   * packed input (inputBits 1, 2, 4; I = uint8_t): rows of
//     pad(ceil(S_disp / (8 / inputBits)), padding) bytes with the beam stride
//     (DH:97 drops the beam stride and rounds the row down);
//   * the step-one output rows are pad(S1, padding / sizeof(O)) elements
//     whatever inputBits is (DH:148 divides S1 by 8 / inputBits).
// inputBits >= 8 reads whole elements of I (uint8_t or float input).
namespace detail {

inline unsigned int shiftOf(float firstDM, float dmStep, unsigned int dm, float delay) {
  volatile float product = static_cast<float>(dm) * dmStep;
  volatile float value = firstDM + product;
  volatile float scaled = value * delay;
  return static_cast<unsigned int>(scaled);
}

// Row length in elements of I: S_disp samples of inputBits bits each.
template <typename I>
inline uint64_t inputPitch(unsigned int dispersedSamples, unsigned int padding, uint8_t inputBits) {
  const unsigned int per = inputBits >= 8 ? 1u : 8u / inputBits;
  return AstroData::pad((dispersedSamples + per - 1) / per, padding / sizeof(I));
}

// Sample x of a row starting at `row`.
template <typename I, typename L> inline L sampleOf(const I *row, unsigned int x, uint8_t inputBits) {
  if (inputBits >= 8) return static_cast<L>(row[x]);
  const unsigned int per = 8u / inputBits;
  const unsigned int byte = static_cast<unsigned int>(row[x / per]);
  return static_cast<L>((byte >> ((x % per) * inputBits)) & ((1u << inputBits) - 1u));
}

} // namespace detail

// include/Dedispersion.hpp:59, body :70-109.  input [beams][C][pitch],
// output [sBeams][D][pad(S / downsampling, padding / sizeof(O))].
template <typename I, typename L, typename O>
void dedispersion(AstroData::Observation &observation, const std::vector<unsigned int> &zappedChannels,
                  const std::vector<unsigned int> &beamMapping, const std::vector<I> &input, std::vector<O> &output,
                  const std::vector<float> &shifts, const unsigned int padding, const uint8_t inputBits) {
  const unsigned int C = observation.getNrChannels();
  const unsigned int samples = observation.getNrSamplesPerBatch() / observation.getDownsampling();
  const unsigned int mapPitch = observation.getNrChannels(padding / sizeof(unsigned int));
  const uint64_t inPitch = detail::inputPitch<I>(observation.getNrSamplesPerDispersedBatch(), padding, inputBits);
  const uint64_t outPitch = AstroData::pad(samples, padding / sizeof(O));
  const float firstDM = observation.getFirstDM(), dmStep = observation.getDMStep();
  for (unsigned int sBeam = 0; sBeam < observation.getNrSynthesizedBeams(); sBeam++) {
    for (unsigned int dm = 0; dm < observation.getNrDMs(); dm++) {
      O *out = output.data() + (static_cast<uint64_t>(sBeam) * observation.getNrDMs() + dm) * outPitch;
      for (unsigned int sample = 0; sample < samples; sample++) {
        L sum = static_cast<L>(0);
        for (unsigned int channel = 0; channel < C; channel++) {
          if (zappedChannels[channel] != 0) continue;
          const unsigned int shift = detail::shiftOf(firstDM, dmStep, dm, shifts[channel]);
          const uint64_t beam = beamMapping[static_cast<uint64_t>(sBeam) * mapPitch + channel];
          const I *row = input.data() + (beam * C + channel) * inPitch;
          sum += detail::sampleOf<I, L>(row, sample + shift, inputBits);
        }
        out[sample] = static_cast<O>(sum);
      }
    }
  }
}

// include/Dedispersion.hpp:60, body :111-153.  The shift of a channel is
// relative to the top channel of its subband (DH:124).  input
// [beams][C][pitch(S1_disp)], output [beams][Dc][Nsb][pad(S1, padding /
// sizeof(O))] with S1 / downsampling samples per row.
template <typename I, typename L, typename O>
void subbandDedispersionStepOne(AstroData::Observation &observation, const std::vector<unsigned int> &zappedChannels,
                                const std::vector<I> &input, std::vector<O> &output, const std::vector<float> &shifts,
                                const unsigned int padding, const uint8_t inputBits) {
  const unsigned int C = observation.getNrChannels();
  const unsigned int subbands = observation.getNrSubbands();
  const unsigned int perSubband = observation.getNrChannelsPerSubband();
  const unsigned int coarseDMs = observation.getNrDMs(true);
  const unsigned int samples = observation.getNrSamplesPerBatch(true) / observation.getDownsampling();
  const uint64_t inPitch = detail::inputPitch<I>(observation.getNrSamplesPerDispersedBatch(true), padding, inputBits);
  const uint64_t outPitch = AstroData::pad(observation.getNrSamplesPerBatch(true), padding / sizeof(O));
  const float firstDM = observation.getFirstDM(true), dmStep = observation.getDMStep(true);
  for (unsigned int beam = 0; beam < observation.getNrBeams(); beam++) {
    for (unsigned int dm = 0; dm < coarseDMs; dm++) {
      for (unsigned int subband = 0; subband < subbands; subband++) {
        const float top = shifts[(subband + 1) * perSubband - 1];
        O *out = output.data() + ((static_cast<uint64_t>(beam) * coarseDMs + dm) * subbands + subband) * outPitch;
        for (unsigned int sample = 0; sample < samples; sample++) {
          L sum = static_cast<L>(0);
          for (unsigned int channel = subband * perSubband; channel < (subband + 1) * perSubband; channel++) {
            if (zappedChannels[channel] != 0) continue;
            const unsigned int shift = detail::shiftOf(firstDM, dmStep, dm, shifts[channel] - top);
            const I *row = input.data() + (static_cast<uint64_t>(beam) * C + channel) * inPitch;
            sum += detail::sampleOf<I, L>(row, sample + shift, inputBits);
          }
          out[sample] = static_cast<O>(sum);
        }
      }
    }
  }
}

// include/Dedispersion.hpp:61, body :155-176.  input [beams][Dc][Nsb]
// [pad(S1 / downsampling, padding / sizeof(I))], output [sBeams][Dc * Df]
// [pad(S / downsampling, padding / sizeof(O))]; no zapping.
template <typename I, typename L, typename O>
void subbandDedispersionStepTwo(AstroData::Observation &observation, const std::vector<unsigned int> &beamMapping,
                                const std::vector<I> &input, std::vector<O> &output, const std::vector<float> &shifts,
                                const unsigned int padding) {
  const unsigned int subbands = observation.getNrSubbands();
  const unsigned int mapPitch = observation.getNrSubbands(padding / sizeof(unsigned int));
  const unsigned int coarseDMs = observation.getNrDMs(true), fineDMs = observation.getNrDMs();
  const unsigned int ds = observation.getDownsampling();
  const unsigned int samples = observation.getNrSamplesPerBatch() / ds;
  const uint64_t inPitch = AstroData::pad(observation.getNrSamplesPerBatch(true) / ds, padding / sizeof(I));
  const uint64_t outPitch = AstroData::pad(samples, padding / sizeof(O));
  const float firstDM = observation.getFirstDM(), dmStep = observation.getDMStep();
  for (unsigned int sBeam = 0; sBeam < observation.getNrSynthesizedBeams(); sBeam++) {
    for (unsigned int coarse = 0; coarse < coarseDMs; coarse++) {
      for (unsigned int dm = 0; dm < fineDMs; dm++) {
        O *out = output.data() + ((static_cast<uint64_t>(sBeam) * coarseDMs + coarse) * fineDMs + dm) * outPitch;
        for (unsigned int sample = 0; sample < samples; sample++) {
          L sum = static_cast<L>(0);
          for (unsigned int subband = 0; subband < subbands; subband++) {
            const unsigned int shift = detail::shiftOf(firstDM, dmStep, dm, shifts[subband]);
            const uint64_t beam = beamMapping[static_cast<uint64_t>(sBeam) * mapPitch + subband];
            const I *row = input.data() + ((beam * coarseDMs + coarse) * subbands + subband) * inPitch;
            sum += static_cast<L>(row[sample + shift]);
          }
          out[sample] = static_cast<O>(sum);
        }
      }
    }
  }
}

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

// Replaces the `dedispersion` kernel: launch(input, output, firstSynthesizedBeam)
// over uint8 (or packed) rows, or float rows for a kernel made with I = float.
class DedispersionKernel : public HIPKernel {
public:
  using HIPKernel::HIPKernel;
  void launch(const uint8_t *input, float *output, unsigned int firstSynthesizedBeam = 0,
              unsigned int nrSynthesizedBeams = 0, void *stream = nullptr) {
    check(dd_dedisperse(plan(), input, output, firstSynthesizedBeam, nrSynthesizedBeams, stream));
  }
  void launch(const float *input, float *output, unsigned int firstSynthesizedBeam = 0,
              unsigned int nrSynthesizedBeams = 0, void *stream = nullptr) {
    check(dd_dedisperse_f32(plan(), input, output, firstSynthesizedBeam, nrSynthesizedBeams, stream));
  }
};

// Replaces `dedispersionStepOne`: launch(input, output).  The step-one delays
// are the CPU function's (relative to each subband's top channel) unless
// setAbsoluteShifts(true) selects the OpenCL kernel's (each channel's own).
class SubbandStepOneKernel : public HIPKernel {
public:
  using HIPKernel::HIPKernel;
  void launch(const uint8_t *input, float *output, void *stream = nullptr) {
    check(dd_subband_step_one(plan(), input, output, stream));
  }
  void launch(const float *input, float *output, void *stream = nullptr) {
    check(dd_subband_step_one_f32(plan(), input, output, stream));
  }
  void setAbsoluteShifts(bool absolute) {
    check(dd_plan_set_step_one_shifts(plan(), absolute ? DD_STEP_ONE_ABSOLUTE : DD_STEP_ONE_RELATIVE));
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

namespace detail {

// The plan's input_bits for input type I and the reference's inputBits:
// uint8_t rows take 1, 2, 4 or 8; float rows (I = float) take whole
// elements, inputBits >= 8 as in the reference's `inputBits >= 8` branch.
template <typename I> unsigned int planInputBits(uint8_t inputBits) {
  static_assert(std::is_same<I, uint8_t>::value || std::is_same<I, float>::value, "uint8_t or float input");
  if (std::is_same<I, float>::value) return inputBits >= 8 ? 32u : 0u;
  return inputBits;
}

// A new kernel with its configuration and tables set; deleted on any error.
template <typename K>
K *makeKernel(const AstroData::Observation &observation, unsigned int padding, unsigned int inputBits,
              dd_mode mode, const DedispersionConf &conf, std::vector<float> *shifts,
              std::vector<float> *shiftsStepTwo) {
  K *k = new K(observation, padding);
  try {
    if (mode != DD_STEP_TWO) k->setInputBits(inputBits);
    k->setShifts(shifts, shiftsStepTwo);
    k->setConf(mode, conf);
  } catch (...) {
    delete k;
    throw;
  }
  return k;
}

} // namespace detail

// Factories with the generators' argument lists (include/Dedispersion.hpp:63-65).
// I is uint8_t (inputBits 1, 2, 4 = packed rows, see dd_plan_set_input_bits,
// or 8) or float (inputBits >= 8); O is float (include/configuration.hpp:9-15).
// Other values throw HIPError.
template <typename I, typename O>
DedispersionKernel *getDedispersionHIP(const DedispersionConf &conf, const unsigned int padding,
                                       const uint8_t inputBits, const AstroData::Observation &observation,
                                       std::vector<float> &shifts) {
  static_assert(sizeof(O) == 4, "float output only");
  return detail::makeKernel<DedispersionKernel>(observation, padding, detail::planInputBits<I>(inputBits),
                                                DD_SINGLE_STEP, conf, &shifts, nullptr);
}

template <typename I, typename O>
SubbandStepOneKernel *getSubbandDedispersionStepOneHIP(const DedispersionConf &conf, const unsigned int padding,
                                                       const uint8_t inputBits,
                                                       const AstroData::Observation &observation,
                                                       std::vector<float> &shifts) {
  static_assert(sizeof(O) == 4, "float output only");
  return detail::makeKernel<SubbandStepOneKernel>(observation, padding, detail::planInputBits<I>(inputBits),
                                                  DD_STEP_ONE, conf, &shifts, nullptr);
}

template <typename I>
SubbandStepTwoKernel *getSubbandDedispersionStepTwoHIP(const DedispersionConf &conf, const unsigned int padding,
                                                       const AstroData::Observation &observation,
                                                       std::vector<float> &shifts) {
  static_assert(std::is_same<I, float>::value, "float intermediate only");
  return detail::makeKernel<SubbandStepTwoKernel>(observation, padding, 8, DD_STEP_TWO, conf, nullptr, &shifts);
}

} // namespace Dedispersion
