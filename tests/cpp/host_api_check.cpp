// Exercises the host-side C++ API of the engine without touching a GPU:
// getShifts/getShiftsStepTwo, DedispersionConf::print, the tuned-conf reader,
// readZappedChannels and generateBeamMapping.  Prints results as text for the
// Python test to compare against the oracle.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>

#include <Dedispersion.hpp>
#include <Shifts.hpp>

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  std::string mode = argv[1];
  if (mode == "shifts") {
    // args: C Nsb minFreq bandwidth samples samplingTime downsampling padding
    AstroData::Observation obs;
    obs.setFrequencyRange(std::atoi(argv[3]), std::atoi(argv[2]), std::strtof(argv[4], nullptr),
                          std::strtof(argv[5], nullptr));
    obs.setNrSamplesPerBatch(std::atoi(argv[6]));
    obs.setSamplingTime(std::strtof(argv[7], nullptr));
    obs.setDownsampling(std::atoi(argv[8]));
    unsigned padding = std::atoi(argv[9]);
    std::unique_ptr<std::vector<float>> s(Dedispersion::getShifts(obs, padding));
    std::unique_ptr<std::vector<float>> s2(Dedispersion::getShiftsStepTwo(obs, padding));
    std::printf("%zu %zu\n", s->size(), s2->size());
    for (float v : *s) { unsigned u; std::memcpy(&u, &v, 4); std::printf("%08x\n", u); }
    for (float v : *s2) { unsigned u; std::memcpy(&u, &v, 4); std::printf("%08x\n", u); }
    return 0;
  }
  if (mode == "conf") {
    // args: file; writes two confs, reads them back, prints them
    Dedispersion::DedispersionConf a, b;
    a.setLocalMem(true); a.setUnroll(4); a.setNrThreadsD0(64); a.setNrThreadsD1(4); a.setNrItemsD0(16); a.setNrItemsD1(2);
    b.setUnroll(2); b.setNrThreadsD0(64); b.setNrThreadsD1(8); b.setNrItemsD0(8); b.setNrItemsD1(1);
    {
      std::ofstream f(argv[2]);
      f << "# device nrDMs split local unroll t0 t1 t2 i0 i1 i2\n";
      f << "MI355X 2048 " << a.print() << "\n";
      f << "\n1234 ignored line\n";
      f << "MI355X 64 " << b.print() << "\n";
      f << "gfx950 32 " << b.print() << "\n";
    }
    Dedispersion::tunedDedispersionConf tuned;
    Dedispersion::readTunedDedispersionConf(tuned, argv[2]);
    for (auto &dev : tuned)
      for (auto &e : *dev.second) {
        std::printf("%s %u %s\n", dev.first.c_str(), e.first, e.second->print().c_str());
        delete e.second;
      }
    for (auto &dev : tuned) delete dev.second;
    try {
      Dedispersion::tunedDedispersionConf none;
      Dedispersion::readTunedDedispersionConf(none, std::string(argv[2]) + ".missing");
      std::printf("no exception\n");
    } catch (AstroData::FileError &) {
      std::printf("FileError\n");
    }
    return 0;
  }
  if (mode == "readconf") {
    // args: file; prints every entry the reader returns
    Dedispersion::tunedDedispersionConf tuned;
    Dedispersion::readTunedDedispersionConf(tuned, argv[2]);
    for (auto &dev : tuned) {
      for (auto &e : *dev.second) {
        std::printf("%s %u %s\n", dev.first.c_str(), e.first, e.second->print().c_str());
        delete e.second;
      }
      delete dev.second;
    }
    return 0;
  }
  if (mode == "zapped") {
    AstroData::Observation obs;
    obs.setFrequencyRange(1, 16, 1400.0f, 1.0f);
    obs.setNrSynthesizedBeams(3);
    obs.setNrBeams(2);
    std::vector<unsigned int> z(obs.getNrChannels(32 / 4), 0);
    AstroData::readZappedChannels(obs, argv[2], z);
    std::printf("%u:", obs.getNrZappedChannels());
    for (unsigned v : z) std::printf(" %u", v);
    std::printf("\n");
    std::vector<unsigned int> m;
    AstroData::generateBeamMapping(obs, m, 32);
    std::printf("%zu:", m.size());
    for (unsigned v : m) std::printf(" %u", v);
    std::printf("\n");
    return 0;
  }
  return 2;
}
