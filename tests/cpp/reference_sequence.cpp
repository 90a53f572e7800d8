// The reference's call sequence, compiled against this engine's headers and
// run on the GPU: tables (src/DedispersionTest.cpp:116-125), sizes
// (Test:129, 145, 146), a tuned configuration read with
// readTunedDedispersionConf (src/Dedispersion.cpp:20-81), the get...HIP
// factories in place of the OpenCL generators (Test:339-343), launches on
// hipMalloc'd buffers (Test:380-399) and the CPU functions
// (include/Dedispersion.hpp:59-61) as the check (Test:403-416, 419-498).
//
// usage: reference_sequence OUT_DIR TUNED_SINGLE TUNED_STEP_ONE TUNED_STEP_TWO ZAPPED_FILE
//
// Every case prints "<case> <mismatching samples> <compared samples>" and
// writes its input and device output to OUT_DIR as raw files, which the
// Python test also compares with the oracle.  The exit status is 0 when
// every case matches its CPU function bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include <Dedispersion.hpp>
#include <Shifts.hpp>
#include <configuration.hpp>

namespace {

int failures = 0;

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    std::exit(2);
  }
}

template <typename T> T *to_device(const std::vector<T> &host) {
  T *d = nullptr;
  hip_check(hipMalloc(reinterpret_cast<void **>(&d), host.size() * sizeof(T)), "hipMalloc");
  hip_check(hipMemcpy(d, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
  return d;
}

template <typename T> T *device_filled(size_t n, T value) {
  std::vector<T> host(n, value);
  return to_device(host);
}

template <typename T> std::vector<T> to_host(const T *d, size_t n) {
  std::vector<T> host(n);
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  hip_check(hipMemcpy(host.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
  return host;
}

template <typename T> void write_raw(const std::string &path, const std::vector<T> &v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char *>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}

// Compares the first `valid` samples of every row of `pitch`.
void compare(const char *name, const std::vector<float> &gpu, const std::vector<float> &cpu, size_t rows,
             size_t pitch, size_t valid) {
  size_t bad = 0;
  for (size_t r = 0; r < rows; r++)
    for (size_t t = 0; t < valid; t++)
      if (std::memcmp(&gpu[r * pitch + t], &cpu[r * pitch + t], sizeof(float)) != 0) bad++;
  std::printf("%s %zu %zu\n", name, bad, rows * valid);
  if (bad) failures++;
}

// A portable generator (splitmix64) for the inputs.
struct Rng {
  uint64_t state;
  uint64_t next() {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
};

Dedispersion::DedispersionConf tuned_conf(const std::string &file, const std::string &device, unsigned int nrDMs) {
  Dedispersion::tunedDedispersionConf tuned;
  Dedispersion::readTunedDedispersionConf(tuned, file);
  Dedispersion::DedispersionConf conf = *tuned.at(device)->at(nrDMs);
  for (auto &d : tuned) {
    for (auto &c : *d.second) delete c.second;
    delete d.second;
  }
  return conf;
}

} // namespace

int main(int argc, char **argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s OUT_DIR TUNED_SINGLE TUNED_STEP_ONE TUNED_STEP_TWO ZAPPED_FILE\n", argv[0]);
    return 2;
  }
  const std::string out = argv[1];
  const unsigned int padding = 128;
  const std::string device = "MI355X";

  // Test:80-89: two beams, three synthesized beams, 256 channels in 8
  // subbands, 1500 samples, 40 DMs from 3.3 in steps of 0.7 and 12 coarse
  // DMs from 12.7 in steps of 4.3 (first DMs away from 0, where a fused
  // multiply-add in the shift rule would show).
  AstroData::Observation observation;
  observation.setNrBeams(2);
  observation.setNrSynthesizedBeams(3);
  observation.setNrSamplesPerBatch(1500);
  observation.setSamplingTime(0.00008192f);
  observation.setFrequencyRange(8, 256, 1300.0f, 0.5f);
  observation.setDMRange(40, 3.3f, 0.7f);
  observation.setDMRange(12, 12.7f, 4.3f, true);

  // Test:116-125
  std::unique_ptr<std::vector<float>> shiftsSingleStep(Dedispersion::getShifts(observation, padding));
  std::unique_ptr<std::vector<float>> shiftsStepOne(Dedispersion::getShifts(observation, padding));
  std::unique_ptr<std::vector<float>> shiftsStepTwo(Dedispersion::getShiftsStepTwo(observation, padding));
  std::vector<unsigned int> zappedChannels(observation.getNrChannels(padding / sizeof(unsigned int)));
  std::vector<unsigned int> beamMappingSingleStep, beamMappingStepTwo;
  AstroData::readZappedChannels(observation, argv[5], zappedChannels);
  AstroData::generateBeamMapping(observation, beamMappingSingleStep, padding);
  AstroData::generateBeamMapping(observation, beamMappingStepTwo, padding, true);
  // a mixed mapping: synthesized beam 2 takes its first 100 channels from beam 0
  for (unsigned int ch = 0; ch < 100; ch++)
    beamMappingSingleStep[2 * observation.getNrChannels(padding / sizeof(unsigned int)) + ch] = 0;

  // Test:129, 145, 146
  observation.setNrSamplesPerDispersedBatch(static_cast<unsigned int>(
      std::ceil(observation.getNrSamplesPerBatch() +
                (shiftsSingleStep->at(0) *
                 (observation.getFirstDM() + ((observation.getNrDMs() - 1) * observation.getDMStep()))))));
  observation.setNrSamplesPerBatch(
      static_cast<unsigned int>(std::ceil(
          observation.getNrSamplesPerBatch() +
          (shiftsStepTwo->at(0) *
           (observation.getFirstDM() + ((observation.getNrDMs() - 1) * observation.getDMStep()))))),
      true);
  observation.setNrSamplesPerDispersedBatch(
      static_cast<unsigned int>(std::ceil(
          observation.getNrSamplesPerBatch(true) +
          (shiftsStepOne->at(0) *
           (observation.getFirstDM(true) + ((observation.getNrDMs(true) - 1) * observation.getDMStep(true)))))),
      true);

  const unsigned int C = observation.getNrChannels();
  const unsigned int S = observation.getNrSamplesPerBatch();
  Rng rng{12345};

  // ---- single step, 8-bit input (Test:339, 380-385, 403) ----
  {
    const size_t inPitch = observation.getNrSamplesPerDispersedBatch(false, padding / sizeof(inputDataType));
    const size_t outPitch = observation.getNrSamplesPerBatch(false, padding / sizeof(outputDataType));
    std::vector<inputDataType> input(observation.getNrBeams() * C * inPitch);
    for (auto &v : input) v = static_cast<inputDataType>(rng.next() & 0xFF);
    std::vector<outputDataType> output_c(observation.getNrSynthesizedBeams() * observation.getNrDMs() * outPitch);
    Dedispersion::DedispersionConf conf = tuned_conf(argv[2], device, observation.getNrDMs());
    std::unique_ptr<Dedispersion::DedispersionKernel> kernel(
        Dedispersion::getDedispersionHIP<inputDataType, outputDataType>(conf, padding, inputBits, observation,
                                                                        *shiftsSingleStep));
    kernel->setZappedChannels(zappedChannels);
    kernel->setBeamMapping(&beamMappingSingleStep, nullptr);
    inputDataType *d_input = to_device(input);
    outputDataType *d_output = device_filled<outputDataType>(output_c.size(), -1.0f);
    kernel->launch(d_input, d_output, 0);
    std::vector<outputDataType> output = to_host(d_output, output_c.size());
    Dedispersion::dedispersion<inputDataType, intermediateDataType, outputDataType>(
        observation, zappedChannels, beamMappingSingleStep, input, output_c, *shiftsSingleStep, padding, inputBits);
    compare("single_step", output, output_c, observation.getNrSynthesizedBeams() * observation.getNrDMs(), outPitch,
            S);
    write_raw(out + "/single_in.u8", input);
    write_raw(out + "/single_out.f32", output);
    // firstSynthesizedBeam = 1, two synthesized beams (Test:385)
    outputDataType *d_sub = device_filled<outputDataType>(2 * observation.getNrDMs() * outPitch, -1.0f);
    kernel->launch(d_input, d_sub, 1, 2);
    std::vector<outputDataType> sub = to_host(d_sub, 2 * observation.getNrDMs() * outPitch);
    std::vector<outputDataType> want(output_c.begin() + observation.getNrDMs() * outPitch, output_c.end());
    compare("single_step_sbeams_1_2", sub, want, 2 * observation.getNrDMs(), outPitch, S);
    (void)hipFree(d_sub);
    (void)hipFree(d_output);
    (void)hipFree(d_input);
  }

  // ---- single step, 4-bit packed input (inputBits 4) ----
  {
    const unsigned int bits = 4;
    const size_t inPitch = AstroData::pad((observation.getNrSamplesPerDispersedBatch() + 1) / 2, padding);
    const size_t outPitch = observation.getNrSamplesPerBatch(false, padding / sizeof(outputDataType));
    std::vector<uint8_t> input(observation.getNrBeams() * C * inPitch);
    for (auto &v : input) v = static_cast<uint8_t>(rng.next() & 0xFF);
    std::vector<float> output_c(observation.getNrSynthesizedBeams() * observation.getNrDMs() * outPitch);
    Dedispersion::DedispersionConf conf = tuned_conf(argv[2], device, observation.getNrDMs());
    std::unique_ptr<Dedispersion::DedispersionKernel> kernel(
        Dedispersion::getDedispersionHIP<uint8_t, float>(conf, padding, bits, observation, *shiftsSingleStep));
    kernel->setZappedChannels(zappedChannels);
    kernel->setBeamMapping(&beamMappingSingleStep, nullptr);
    uint8_t *d_input = to_device(input);
    float *d_output = device_filled<float>(output_c.size(), -1.0f);
    kernel->launch(d_input, d_output);
    std::vector<float> output = to_host(d_output, output_c.size());
    Dedispersion::dedispersion<uint8_t, float, float>(observation, zappedChannels, beamMappingSingleStep, input,
                                                      output_c, *shiftsSingleStep, padding, bits);
    compare("single_step_4bit", output, output_c, observation.getNrSynthesizedBeams() * observation.getNrDMs(),
            outPitch, S);
    (void)hipFree(d_output);
    (void)hipFree(d_input);
  }

  // ---- single step, float input (I = float) ----
  {
    const size_t inPitch = observation.getNrSamplesPerDispersedBatch(false, padding / sizeof(float));
    const size_t outPitch = observation.getNrSamplesPerBatch(false, padding / sizeof(float));
    std::vector<float> input(observation.getNrBeams() * C * inPitch);
    for (auto &v : input) v = static_cast<float>(static_cast<int64_t>(rng.next() % 2000001) - 1000000) / 777.0f;
    std::vector<float> output_c(observation.getNrSynthesizedBeams() * observation.getNrDMs() * outPitch);
    Dedispersion::DedispersionConf conf = tuned_conf(argv[2], device, observation.getNrDMs());
    std::unique_ptr<Dedispersion::DedispersionKernel> kernel(
        Dedispersion::getDedispersionHIP<float, float>(conf, padding, 32, observation, *shiftsSingleStep));
    kernel->setZappedChannels(zappedChannels);
    kernel->setBeamMapping(&beamMappingSingleStep, nullptr);
    float *d_input = to_device(input);
    float *d_output = device_filled<float>(output_c.size(), -1.0f);
    kernel->launch(d_input, d_output);
    std::vector<float> output = to_host(d_output, output_c.size());
    Dedispersion::dedispersion<float, float, float>(observation, zappedChannels, beamMappingSingleStep, input,
                                                    output_c, *shiftsSingleStep, padding, 32);
    compare("single_step_float", output, output_c, observation.getNrSynthesizedBeams() * observation.getNrDMs(),
            outPitch, S);
    write_raw(out + "/float_in.f32", input);
    write_raw(out + "/float_out.f32", output);
    (void)hipFree(d_output);
    (void)hipFree(d_input);
  }

  // ---- step one, then step two on its output (Test:341-343, 388-397, 407-410) ----
  {
    const size_t inPitch = observation.getNrSamplesPerDispersedBatch(true, padding / sizeof(inputDataType));
    const size_t midPitch = observation.getNrSamplesPerBatch(true, padding / sizeof(outputDataType));
    const size_t outPitch = observation.getNrSamplesPerBatch(false, padding / sizeof(outputDataType));
    const unsigned int Dc = observation.getNrDMs(true), Nsb = observation.getNrSubbands();
    std::vector<inputDataType> input(observation.getNrBeams() * C * inPitch);
    for (auto &v : input) v = static_cast<inputDataType>(rng.next() & 0xFF);
    std::vector<outputDataType> mid_c(observation.getNrBeams() * Dc * Nsb * midPitch);
    Dedispersion::DedispersionConf conf1 = tuned_conf(argv[3], device, Dc);
    Dedispersion::DedispersionConf conf2 = tuned_conf(argv[4], device, observation.getNrDMs());
    std::unique_ptr<Dedispersion::SubbandStepOneKernel> one(
        Dedispersion::getSubbandDedispersionStepOneHIP<inputDataType, outputDataType>(conf1, padding, inputBits,
                                                                                      observation, *shiftsStepOne));
    std::unique_ptr<Dedispersion::SubbandStepTwoKernel> two(
        Dedispersion::getSubbandDedispersionStepTwoHIP<outputDataType>(conf2, padding, observation, *shiftsStepTwo));
    one->setZappedChannels(zappedChannels);
    two->setBeamMapping(nullptr, &beamMappingStepTwo);
    inputDataType *d_input = to_device(input);
    outputDataType *d_mid = device_filled<outputDataType>(mid_c.size(), -1.0f);
    const size_t nOut = observation.getNrSynthesizedBeams() * Dc * observation.getNrDMs() * outPitch;
    outputDataType *d_out = device_filled<outputDataType>(nOut, -1.0f);
    one->launch(d_input, d_mid);
    two->launch(d_mid, d_out, 0);
    std::vector<outputDataType> mid = to_host(d_mid, mid_c.size());
    std::vector<outputDataType> output = to_host(d_out, nOut);
    Dedispersion::subbandDedispersionStepOne<inputDataType, intermediateDataType, outputDataType>(
        observation, zappedChannels, input, mid_c, *shiftsStepOne, padding, inputBits);
    compare("step_one", mid, mid_c, observation.getNrBeams() * Dc * Nsb, midPitch,
            observation.getNrSamplesPerBatch(true));
    std::vector<outputDataType> out_c(nOut);
    Dedispersion::subbandDedispersionStepTwo<outputDataType, intermediateDataType, outputDataType>(
        observation, beamMappingStepTwo, mid, out_c, *shiftsStepTwo, padding);
    compare("step_two", output, out_c, observation.getNrSynthesizedBeams() * Dc * observation.getNrDMs(), outPitch,
            S);
    write_raw(out + "/two_in.u8", input);
    write_raw(out + "/two_mid.f32", mid);
    write_raw(out + "/two_out.f32", output);
    (void)hipFree(d_out);
    (void)hipFree(d_mid);
    (void)hipFree(d_input);
  }

  // ---- errors are exceptions at this layer ----
  {
    Dedispersion::DedispersionConf bad = tuned_conf(argv[2], device, observation.getNrDMs());
    bad.setNrThreadsD0(32);
    bool thrown = false;
    try {
      delete Dedispersion::getDedispersionHIP<inputDataType, outputDataType>(bad, padding, inputBits, observation,
                                                                             *shiftsSingleStep);
    } catch (const Dedispersion::HIPError &e) {
      thrown = e.status() == DD_ERR_CONFIGURATION;
    }
    std::printf("invalid_conf_throws %d 1\n", thrown ? 0 : 1);
    if (!thrown) failures++;
  }
  std::printf("sizes %u %u %u\n", observation.getNrSamplesPerDispersedBatch(), observation.getNrSamplesPerBatch(true),
              observation.getNrSamplesPerDispersedBatch(true));
  return failures == 0 ? 0 : 1;
}
