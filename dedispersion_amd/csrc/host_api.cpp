// host_api.cpp -- the C++ layer kept from the reference's API.
//
//   getShifts / getShiftsStepTwo      src/Shifts.cpp:20-42
//   DedispersionConf, print           src/Dedispersion.cpp:83-89
//   readTunedDedispersionConf         src/Dedispersion.cpp:20-81
//   readZappedChannels, generateBeamMapping (AstroData helpers used at
//   src/DedispersionTest.cpp:125, 249, 310)
#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>

#include <Dedispersion.hpp>
#include <Shifts.hpp>

namespace {

// One table entry with the operation order of src/Shifts.cpp:25-26.
float delay_of(float freq, float inverse_high, float samples, float time_unit) {
  float inverse = 1.0f / (freq * freq);
  float delay = 4148.808f * (inverse - inverse_high);
  delay = delay * samples;
  return delay / time_unit;
}

} // namespace

namespace Dedispersion {

std::vector<float> *getShifts(AstroData::Observation &observation, const unsigned int padding) {
  const float max_freq = observation.getMaxFreq();
  const float inverse_high = 1.0f / (max_freq * max_freq);
  const float time_unit = observation.getNrSamplesPerBatch() * observation.getSamplingTime();
  const float samples = static_cast<float>(observation.getNrSamplesPerBatch() / observation.getDownsampling());
  auto *shifts = new std::vector<float>(observation.getNrChannels(padding / sizeof(float)));
  for (unsigned int channel = 0; channel + 1 < observation.getNrChannels(); channel++) {
    float freq = observation.getMinFreq() + (channel * observation.getChannelBandwidth());
    shifts->at(channel) = delay_of(freq, inverse_high, samples, time_unit);
  }
  shifts->at(observation.getNrChannels() - 1) = 0;
  return shifts;
}

std::vector<float> *getShiftsStepTwo(AstroData::Observation &observation, const unsigned int padding) {
  const float max_freq = observation.getSubbandMaxFreq();
  const float inverse_high = 1.0f / (max_freq * max_freq);
  const float time_unit = observation.getNrSamplesPerBatch() * observation.getSamplingTime();
  const float samples = static_cast<float>(observation.getNrSamplesPerBatch() / observation.getDownsampling());
  auto *shifts = new std::vector<float>(observation.getNrSubbands(padding / sizeof(float)));
  for (unsigned int subband = 0; subband + 1 < observation.getNrSubbands(); subband++) {
    float freq = observation.getSubbandMinFreq() + (subband * observation.getSubbandBandwidth());
    shifts->at(subband) = delay_of(freq, inverse_high, samples, time_unit);
  }
  shifts->at(observation.getNrSubbands() - 1) = 0;
  return shifts;
}

std::string KernelConf::print() const {
  return std::to_string(nrThreadsD0) + " " + std::to_string(nrThreadsD1) + " " + std::to_string(nrThreadsD2) + " " +
         std::to_string(nrItemsD0) + " " + std::to_string(nrItemsD1) + " " + std::to_string(nrItemsD2);
}

DedispersionConf::DedispersionConf() : KernelConf(), splitBatches(false), local(0), unroll(1) {}

DedispersionConf::~DedispersionConf() {}

std::string DedispersionConf::print() const {
  return std::to_string(splitBatches) + " " + std::to_string(local) + " " + std::to_string(unroll) + " " +
         KernelConf::print();
}

dd_conf DedispersionConf::toC() const {
  dd_conf c{};
  c.threads_d0 = getNrThreadsD0();
  c.threads_d1 = getNrThreadsD1();
  c.items_d0 = getNrItemsD0();
  c.items_d1 = getNrItemsD1();
  c.unroll = unroll;
  c.local = local;
  c.split_batches = splitBatches ? 1 : 0;
  return c;
}

DedispersionConf DedispersionConf::fromC(const dd_conf &c) {
  DedispersionConf conf;
  conf.setNrThreadsD0(c.threads_d0);
  conf.setNrThreadsD1(c.threads_d1);
  conf.setNrItemsD0(c.items_d0);
  conf.setNrItemsD1(c.items_d1);
  conf.setUnroll(c.unroll);
  conf.setLocalMode(c.local);
  conf.setSplitBatches(c.split_batches != 0);
  return conf;
}

// src/Dedispersion.cpp:20-81.  Fields are separated by single spaces; the
// numeric fields are read with stream extraction (the reference uses
// isa::utils::castToType, a stringstream conversion).
void readTunedDedispersionConf(tunedDedispersionConf &tunedDedispersion, const std::string &dedispersionFilename) {
  std::ifstream file(dedispersionFilename);
  if (!file) throw AstroData::FileError("Impossible to open " + dedispersionFilename);
  std::string line;
  while (std::getline(file, line)) {
    if (line.empty() || !std::isalpha(static_cast<unsigned char>(line[0]))) continue;
    std::istringstream fields(line);
    std::string device;
    unsigned int nrDMs = 0, local = 0, t0 = 0, t1 = 0, t2 = 0, i0 = 0, i1 = 0, i2 = 0, unroll = 0;
    bool split = false;
    fields >> device >> nrDMs >> split >> local >> unroll >> t0 >> t1 >> t2 >> i0 >> i1 >> i2;
    auto *conf = new DedispersionConf();
    conf->setSplitBatches(split);
    conf->setLocalMode(local);
    conf->setUnroll(unroll);
    conf->setNrThreadsD0(t0);
    conf->setNrThreadsD1(t1);
    conf->setNrThreadsD2(t2);
    conf->setNrItemsD0(i0);
    conf->setNrItemsD1(i1);
    conf->setNrItemsD2(i2);
    if (tunedDedispersion.count(device) == 0) {
      auto *container = new std::map<unsigned int, DedispersionConf *>();
      container->insert(std::make_pair(nrDMs, conf));
      tunedDedispersion.insert(std::make_pair(device, container));
    } else if (!tunedDedispersion.at(device)->insert(std::make_pair(nrDMs, conf)).second) {
      delete conf;  // the first entry for (device, nrDMs) stays
    }
  }
}

dd_observation toC(const AstroData::Observation &o) {
  dd_observation c{};
  c.nr_beams = o.getNrBeams();
  c.nr_synthesized_beams = o.getNrSynthesizedBeams();
  c.nr_channels = o.getNrChannels();
  c.nr_subbands = o.getNrSubbands();
  c.min_freq = o.getMinFreq();
  c.channel_bandwidth = o.getChannelBandwidth();
  c.sampling_time = o.getSamplingTime();
  c.downsampling = o.getDownsampling();
  c.nr_samples_per_batch = o.getNrSamplesPerBatch();
  c.nr_dms = o.getNrDMs();
  c.first_dm = o.getFirstDM();
  c.dm_step = o.getDMStep();
  c.nr_dms_subbanding = o.getNrDMs(true);
  c.first_dm_subbanding = o.getFirstDM(true);
  c.dm_step_subbanding = o.getDMStep(true);
  c.nr_samples_per_dispersed_batch = o.getNrSamplesPerDispersedBatch();
  c.nr_samples_per_batch_subbanding = o.getNrSamplesPerBatch(true);
  c.nr_samples_per_dispersed_batch_subbanding = o.getNrSamplesPerDispersedBatch(true);
  return c;
}

} // namespace Dedispersion

namespace AstroData {

void readZappedChannels(Observation &observation, const std::string &filename,
                        std::vector<unsigned int> &zappedChannels) {
  std::ifstream file(filename);
  if (!file) throw FileError("Impossible to open " + filename);
  std::string line;
  unsigned int count = 0;
  while (std::getline(file, line)) {
    std::string body = line.substr(0, line.find('#'));
    std::istringstream tokens(body);
    unsigned long index;
    while (tokens >> index) {
      if (index >= observation.getNrChannels()) throw std::out_of_range("zapped channel index out of range");
      if (zappedChannels.at(index) == 0) count++;
      zappedChannels.at(index) = 1;
    }
  }
  observation.setNrZappedChannels(count);
}

void generateBeamMapping(const Observation &observation, std::vector<unsigned int> &beamMapping,
                         unsigned int padding, bool subbanding) {
  const unsigned int rows = subbanding ? observation.getNrSubbands() : observation.getNrChannels();
  const unsigned int pitch = pad(rows, padding / sizeof(unsigned int));
  beamMapping.assign(static_cast<size_t>(observation.getNrSynthesizedBeams()) * pitch, 0);
  for (unsigned int sBeam = 0; sBeam < observation.getNrSynthesizedBeams(); sBeam++)
    for (unsigned int row = 0; row < rows; row++)
      beamMapping[static_cast<size_t>(sBeam) * pitch + row] = sBeam % observation.getNrBeams();
}

} // namespace AstroData
