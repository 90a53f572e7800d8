// Observation.hpp -- the observation metadata type the dedispersion path
// reads.  The reference takes it from the external AstroData library
// (AstroData::Observation, used at include/Dedispersion.hpp:72-171 and
// src/Shifts.cpp:21-40 of TRASAL/Dedispersion); that library is not part of
// the reference tree, so this library owns a compatible type with the getters
// and setters the path uses, and states their semantics here:
//
//   getMaxFreq()          = minFreq + (C - 1) * channelBandwidth
//   getSubbandBandwidth() = channelsPerSubband * channelBandwidth
//   getSubbandMinFreq()   = minFreq + (channelsPerSubband - 1) * channelBandwidth
//                           (a subband is referenced to its top channel)
//   getSubbandMaxFreq()   = getSubbandMinFreq() + (Nsb - 1) * getSubbandBandwidth()
//   getX(..., padding)    = X rounded up to a multiple of padding (0 = X)
// All arithmetic is single precision, evaluated left to right.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace AstroData {

class FileError : public std::runtime_error {
public:
  explicit FileError(const std::string &message) : std::runtime_error(message) {}
};

// isa::utils::pad semantics: round value up to a multiple of padding.
inline unsigned int pad(unsigned int value, unsigned int padding) {
  return padding == 0 ? value : ((value + padding - 1) / padding) * padding;
}

class Observation {
public:
  Observation() = default;

  // Beams
  unsigned int getNrBeams() const { return nrBeams; }
  unsigned int getNrSynthesizedBeams() const { return nrSynthesizedBeams; }
  void setNrBeams(unsigned int beams) { nrBeams = beams; }
  void setNrSynthesizedBeams(unsigned int beams) { nrSynthesizedBeams = beams; }

  // Time
  unsigned int getNrSamplesPerBatch(bool subbanding = false, unsigned int padding = 0) const {
    return pad(subbanding ? nrSamplesPerBatchSubbanding : nrSamplesPerBatch, padding);
  }
  unsigned int getNrSamplesPerDispersedBatch(bool subbanding = false, unsigned int padding = 0) const {
    return pad(subbanding ? nrSamplesPerDispersedBatchSubbanding : nrSamplesPerDispersedBatch, padding);
  }
  void setNrSamplesPerBatch(unsigned int samples, bool subbanding = false) {
    (subbanding ? nrSamplesPerBatchSubbanding : nrSamplesPerBatch) = samples;
  }
  void setNrSamplesPerDispersedBatch(unsigned int samples, bool subbanding = false) {
    (subbanding ? nrSamplesPerDispersedBatchSubbanding : nrSamplesPerDispersedBatch) = samples;
  }
  float getSamplingTime() const { return samplingTime; }
  void setSamplingTime(float t) { samplingTime = t; }
  unsigned int getDownsampling() const { return downsampling; }
  void setDownsampling(unsigned int factor) { downsampling = factor; }

  // Frequency
  void setFrequencyRange(unsigned int subbands, unsigned int channels, float minFreq,
                         float channelBandwidth) {
    nrSubbands = subbands;
    nrChannels = channels;
    minChannelFrequency = minFreq;
    this->channelBandwidth = channelBandwidth;
  }
  unsigned int getNrChannels(unsigned int padding = 0) const { return pad(nrChannels, padding); }
  unsigned int getNrSubbands(unsigned int padding = 0) const { return pad(nrSubbands, padding); }
  unsigned int getNrChannelsPerSubband() const { return nrSubbands ? nrChannels / nrSubbands : 0; }
  unsigned int getNrZappedChannels() const { return nrZappedChannels; }
  void setNrZappedChannels(unsigned int zapped) { nrZappedChannels = zapped; }
  float getMinFreq() const { return minChannelFrequency; }
  float getChannelBandwidth() const { return channelBandwidth; }
  float getMaxFreq() const {
    float span = static_cast<float>(nrChannels - 1) * channelBandwidth;
    return minChannelFrequency + span;
  }
  float getSubbandBandwidth() const {
    return static_cast<float>(getNrChannelsPerSubband()) * channelBandwidth;
  }
  float getSubbandMinFreq() const {
    float span = static_cast<float>(getNrChannelsPerSubband() - 1) * channelBandwidth;
    return minChannelFrequency + span;
  }
  float getSubbandMaxFreq() const {
    float span = static_cast<float>(nrSubbands - 1) * getSubbandBandwidth();
    return getSubbandMinFreq() + span;
  }

  // Dispersion measures
  void setDMRange(unsigned int dms, float firstDM, float step, bool subbanding = false) {
    if (subbanding) {
      nrDMsSubbanding = dms;
      firstDMSubbanding = firstDM;
      dmStepSubbanding = step;
    } else {
      nrDMs = dms;
      this->firstDM = firstDM;
      dmStep = step;
    }
  }
  unsigned int getNrDMs(bool subbanding = false, unsigned int padding = 0) const {
    return pad(subbanding ? nrDMsSubbanding : nrDMs, padding);
  }
  float getFirstDM(bool subbanding = false) const { return subbanding ? firstDMSubbanding : firstDM; }
  float getDMStep(bool subbanding = false) const { return subbanding ? dmStepSubbanding : dmStep; }

private:
  unsigned int nrBeams = 1;
  unsigned int nrSynthesizedBeams = 1;
  unsigned int nrSamplesPerBatch = 0;
  unsigned int nrSamplesPerDispersedBatch = 0;
  unsigned int nrSamplesPerBatchSubbanding = 0;
  unsigned int nrSamplesPerDispersedBatchSubbanding = 0;
  float samplingTime = 0.0f;
  unsigned int downsampling = 1;
  unsigned int nrChannels = 0;
  unsigned int nrSubbands = 1;
  unsigned int nrZappedChannels = 0;
  float minChannelFrequency = 0.0f;
  float channelBandwidth = 0.0f;
  unsigned int nrDMs = 0;
  float firstDM = 0.0f;
  float dmStep = 0.0f;
  unsigned int nrDMsSubbanding = 0;
  float firstDMSubbanding = 0.0f;
  float dmStepSubbanding = 0.0f;
};

// Reads a zapped-channel file into flags [pad(C, padding/4)] (non-zero =
// zapped) and sets the observation's zapped count.  Format: channel indices
// separated by white space; '#' starts a comment to the end of the line.
// Throws FileError if the file cannot be opened and std::out_of_range for an
// index >= C.  (Reference call site: src/DedispersionTest.cpp:125.)
void readZappedChannels(Observation &observation, const std::string &filename,
                        std::vector<unsigned int> &zappedChannels);

// Fills the synthesized-beam -> input-beam mapping: entry
// [sBeam * pad(rows, padding/4) + row] = sBeam % nrBeams, with rows = C, or
// Nsb when subbanding.  (Reference call sites: src/DedispersionTest.cpp:249,310.)
void generateBeamMapping(const Observation &observation, std::vector<unsigned int> &beamMapping,
                         unsigned int padding, bool subbanding = false);

} // namespace AstroData
