// Exercises the host-side C++ API of the engine without touching a GPU:
// getShifts/getShiftsStepTwo, DedispersionConf::print, the tuned-conf reader,
// readZappedChannels, generateBeamMapping and the CPU functions
// dedispersion / subbandDedispersionStepOne / subbandDedispersionStepTwo.
// Prints results as text, or writes raw files, for the Python test to
// compare against the oracle.
#include <algorithm>
#include <cmath>
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
  if (mode == "templates") {
    // args: OUT_DIR C NSB S D FIRST STEP DC FIRST1 STEP1 BEAMS SBEAMS DS S1 PADDING
    // Runs the CPU functions (include/Dedispersion.hpp:59-61) on seeded
    // inputs and writes inputs and outputs as raw files for the oracle to
    // check: the single step over 8-bit, 2-bit packed and float rows, step
    // one over 8-bit and float rows, and step two over general floats.
    if (argc != 17) return 2;
    const std::string out = argv[2];
    AstroData::Observation obs;
    obs.setFrequencyRange(std::atoi(argv[4]), std::atoi(argv[3]), 1300.0f, 0.5f);
    obs.setNrSamplesPerBatch(std::atoi(argv[5]));
    obs.setSamplingTime(0.00008192f);
    obs.setDMRange(std::atoi(argv[6]), std::strtof(argv[7], nullptr), std::strtof(argv[8], nullptr));
    obs.setDMRange(std::atoi(argv[9]), std::strtof(argv[10], nullptr), std::strtof(argv[11], nullptr), true);
    obs.setNrBeams(std::atoi(argv[12]));
    obs.setNrSynthesizedBeams(std::atoi(argv[13]));
    obs.setDownsampling(std::atoi(argv[14]));
    const unsigned s1 = std::atoi(argv[15]);
    const unsigned padding = std::atoi(argv[16]);
    std::unique_ptr<std::vector<float>> shifts(Dedispersion::getShifts(obs, padding));
    std::unique_ptr<std::vector<float>> shifts2(Dedispersion::getShiftsStepTwo(obs, padding));
    // src/DedispersionTest.cpp:129, 145 (unless S1 is given), 146
    obs.setNrSamplesPerDispersedBatch(static_cast<unsigned int>(std::ceil(
        obs.getNrSamplesPerBatch() + (shifts->at(0) * (obs.getFirstDM() + ((obs.getNrDMs() - 1) * obs.getDMStep()))))));
    obs.setNrSamplesPerBatch(s1 ? s1
                                : static_cast<unsigned int>(std::ceil(
                                      obs.getNrSamplesPerBatch() +
                                      (shifts2->at(0) * (obs.getFirstDM() + ((obs.getNrDMs() - 1) * obs.getDMStep()))))),
                             true);
    obs.setNrSamplesPerDispersedBatch(
        static_cast<unsigned int>(std::ceil(obs.getNrSamplesPerBatch(true) +
                                            (shifts->at(0) * (obs.getFirstDM(true) +
                                                              ((obs.getNrDMs(true) - 1) * obs.getDMStep(true)))))),
        true);
    std::printf("%u %u %u\n", obs.getNrSamplesPerDispersedBatch(), obs.getNrSamplesPerBatch(true),
                obs.getNrSamplesPerDispersedBatch(true));
    const unsigned C = obs.getNrChannels(), B = obs.getNrBeams(), SB = obs.getNrSynthesizedBeams();
    std::vector<unsigned int> zapped(obs.getNrChannels(padding / 4), 0);
    zapped[1] = zapped[C / 2] = 1;
    std::vector<unsigned int> map1, map2;
    AstroData::generateBeamMapping(obs, map1, padding);
    AstroData::generateBeamMapping(obs, map2, padding, true);
    if (SB > 1 && B > 1) map1[obs.getNrChannels(padding / 4) + 2] = 0;
    uint64_t state = 99;
    auto next = [&state]() {
      uint64_t z = (state += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return z ^ (z >> 31);
    };
    auto write = [&](const std::string &name, const void *data, size_t bytes) {
      std::ofstream f(out + "/" + name, std::ios::binary);
      f.write(reinterpret_cast<const char *>(data), static_cast<std::streamsize>(bytes));
    };
    const size_t outPitch = AstroData::pad(obs.getNrSamplesPerBatch() / obs.getDownsampling(), padding / 4);
    const size_t nOut = static_cast<size_t>(SB) * obs.getNrDMs() * outPitch;
    const size_t pitch8 = obs.getNrSamplesPerDispersedBatch(false, padding);
    std::vector<uint8_t> in8(B * C * pitch8);
    for (auto &v : in8) v = static_cast<uint8_t>(next());
    std::vector<float> o(nOut, 0.0f);
    Dedispersion::dedispersion<uint8_t, float, float>(obs, zapped, map1, in8, o, *shifts, padding, 8);
    write("single_u8_in", in8.data(), in8.size());
    write("single_u8_out", o.data(), o.size() * 4);
    const size_t pitch2 = AstroData::pad((obs.getNrSamplesPerDispersedBatch() + 3) / 4, padding);
    std::vector<uint8_t> in2(B * C * pitch2);
    for (auto &v : in2) v = static_cast<uint8_t>(next());
    std::fill(o.begin(), o.end(), 0.0f);
    Dedispersion::dedispersion<uint8_t, float, float>(obs, zapped, map1, in2, o, *shifts, padding, 2);
    write("single_2bit_in", in2.data(), in2.size());
    write("single_2bit_out", o.data(), o.size() * 4);
    const size_t pitchf = obs.getNrSamplesPerDispersedBatch(false, padding / 4);
    std::vector<float> inf(B * C * pitchf);
    for (auto &v : inf) v = static_cast<float>(static_cast<int64_t>(next() % 200001) - 100000) / 37.0f;
    std::fill(o.begin(), o.end(), 0.0f);
    Dedispersion::dedispersion<float, float, float>(obs, zapped, map1, inf, o, *shifts, padding, 32);
    write("single_f32_in", inf.data(), inf.size() * 4);
    write("single_f32_out", o.data(), o.size() * 4);
    if (obs.getNrDMs(true) > 0) {
      const unsigned Dc = obs.getNrDMs(true), Nsb = obs.getNrSubbands();
      const size_t sub8 = obs.getNrSamplesPerDispersedBatch(true, padding);
      const size_t midPitch = AstroData::pad(obs.getNrSamplesPerBatch(true), padding / 4);
      std::vector<uint8_t> s8(B * C * sub8);
      for (auto &v : s8) v = static_cast<uint8_t>(next());
      std::vector<float> mid(static_cast<size_t>(B) * Dc * Nsb * midPitch, 0.0f);
      Dedispersion::subbandDedispersionStepOne<uint8_t, float, float>(obs, zapped, s8, mid, *shifts, padding, 8);
      write("one_u8_in", s8.data(), s8.size());
      write("one_u8_out", mid.data(), mid.size() * 4);
      const size_t subf = obs.getNrSamplesPerDispersedBatch(true, padding / 4);
      std::vector<float> sf(B * C * subf);
      for (auto &v : sf) v = static_cast<float>(static_cast<int64_t>(next() % 200001) - 100000) / 37.0f;
      std::fill(mid.begin(), mid.end(), 0.0f);
      Dedispersion::subbandDedispersionStepOne<float, float, float>(obs, zapped, sf, mid, *shifts, padding, 32);
      write("one_f32_in", sf.data(), sf.size() * 4);
      write("one_f32_out", mid.data(), mid.size() * 4);
      const size_t twoPitch = AstroData::pad(obs.getNrSamplesPerBatch(true) / obs.getDownsampling(), padding / 4);
      std::vector<float> x(static_cast<size_t>(B) * Dc * Nsb * twoPitch);
      for (auto &v : x) v = static_cast<float>(static_cast<int64_t>(next() % 200001) - 100000) / 11.0f;
      std::vector<float> o2(static_cast<size_t>(SB) * Dc * obs.getNrDMs() * outPitch, 0.0f);
      Dedispersion::subbandDedispersionStepTwo<float, float, float>(obs, map2, x, o2, *shifts2, padding);
      write("two_in", x.data(), x.size() * 4);
      write("two_out", o2.data(), o2.size() * 4);
    }
    return 0;
  }
  return 2;
}
