/*
 * dedispersion_amd.h -- C ABI of the MI355X dedispersion engine.
 *
 * The reference (TRASAL/Dedispersion) exposes its device path as OpenCL
 * source generators whose kernels a caller compiles and launches with
 * positional arguments:
 *   single step  (include/Dedispersion.hpp:63, args src/DedispersionTest.cpp:380-385)
 *   step one     (include/Dedispersion.hpp:64, args src/DedispersionTest.cpp:388-391)
 *   step two     (include/Dedispersion.hpp:65, args src/DedispersionTest.cpp:393-397)
 * This header replaces those generators and that launch protocol with a
 * plan object: the observation and padding fix every size and layout, the
 * tables (shifts, zapped channels, beam mapping) become device-resident at
 * plan set-up, and each launch takes device pointers laid out exactly as
 * the reference's buffers are.
 *
 * Conventions: plain C types only; every function returns a dd_status;
 * device pointers are HIP device pointers; a stream argument is a
 * hipStream_t passed as void* (NULL = the default stream).  A plan is used
 * by one host thread at a time; distinct plans may run concurrently.
 */
#ifndef DEDISPERSION_AMD_H
#define DEDISPERSION_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DD_ABI_VERSION 6

typedef enum dd_status {
  DD_OK = 0,
  DD_ERR_INVALID_ARGUMENT = 1, /* a size, pointer or table value is invalid */
  DD_ERR_UNSUPPORTED = 2,      /* valid for the reference, not supported here */
  DD_ERR_HIP = 3,              /* a HIP runtime call failed */
  DD_ERR_OUT_OF_MEMORY = 4,
  DD_ERR_CONFIGURATION = 5,    /* the kernel configuration cannot run */
  DD_ERR_NOT_READY = 6,        /* a required table has not been set */
  DD_ERR_IO = 7                /* a file could not be read or written */
} dd_status;

/* Observation metadata: the AstroData::Observation fields the path reads
 * (include/Dedispersion.hpp:72-171, src/Shifts.cpp:21-40).  Zero for a
 * dispersed/subbanding sample count means "derive it with the reference's
 * sizing rule" (src/DedispersionTest.cpp:129,145,146). */
typedef struct dd_observation {
  uint32_t nr_beams;
  uint32_t nr_synthesized_beams;
  uint32_t nr_channels;
  uint32_t nr_subbands;           /* 1 for single-step only plans */
  float min_freq;                 /* MHz, frequency of channel 0 */
  float channel_bandwidth;        /* MHz */
  float sampling_time;            /* s */
  uint32_t downsampling;          /* >= 1 */
  uint32_t nr_samples_per_batch;  /* S: output samples of single step / step two */
  uint32_t nr_dms;                /* D (single step) or fine DMs Df (step two) */
  float first_dm;
  float dm_step;
  uint32_t nr_dms_subbanding;     /* Dc (step one); 0 for single-step plans */
  float first_dm_subbanding;
  float dm_step_subbanding;
  uint32_t nr_samples_per_dispersed_batch;             /* S_disp, 0 = derive */
  uint32_t nr_samples_per_batch_subbanding;            /* S1, 0 = derive */
  uint32_t nr_samples_per_dispersed_batch_subbanding;  /* S1_disp, 0 = derive */
} dd_observation;

/* Kernel configuration, the fields of Dedispersion::DedispersionConf
 * (include/Dedispersion.hpp:34-54) as they apply to the HIP kernels:
 *   threads_d0  threads along samples (64, one wavefront)
 *   threads_d1  wavefronts along DMs per workgroup
 *   items_d0    samples per thread (4, 8 or 16)
 *   items_d1    DMs per thread (1, 2, 4 or 8)
 *   unroll      channels staged per barrier (1..16)
 *   local       0 = read rows through the caches; 1 = stage rows in LDS with
 *               a double-buffered pipeline (used when unroll <= 2 *
 *               threads_d1 and the row window fits, else as 2); 2 = stage
 *               rows in LDS, one group at a time; 3 = single step over
 *               channel-block patterns (blocks of four live channels whose
 *               relative shifts repeat across a tile of 32 DMs are summed
 *               once per pattern): needs threads_d1 x items_d1 = 32,
 *               items_d1 4 or 8, items_d0 8 or 16, unroll (blocks per
 *               phase) <= 8; plans whose shift table has more than 16
 *               patterns per block and tile, or whose phase of `unroll`
 *               blocks does not fit in LDS, run the staged pipeline.
 *               Local 1 and 2 whose row windows (the tile plus the spread
 *               of its DMs' shifts) do not fit in LDS run as local 0.
 *   split_batches must be 0 (unimplemented in the reference too). */
typedef struct dd_conf {
  uint32_t threads_d0;
  uint32_t threads_d1;
  uint32_t items_d0;
  uint32_t items_d1;
  uint32_t unroll;
  uint32_t local;
  uint32_t split_batches;
} dd_conf;

typedef enum dd_mode {
  DD_SINGLE_STEP = 0,
  DD_STEP_ONE = 1,
  DD_STEP_TWO = 2
} dd_mode;

/* Sizes of the device buffers of a plan, in elements of their type, using
 * the reference layouts (input elements are bytes, or floats for float
 * input):
 *   input        uint8 [beams][C][in_pitch]              (include/Dedispersion.hpp:90)
 *   output       float [sBeams][D][out_pitch]            (include/Dedispersion.hpp:105)
 *   step-one in  uint8 [beams][C][sub_in_pitch]          (include/Dedispersion.hpp:133)
 *   intermediate float [beams][Dc][Nsb][sub_out_pitch]   (include/Dedispersion.hpp:148)
 *   step-two out float [sBeams][Dc*Df][out_pitch]        (include/Dedispersion.hpp:171) */
typedef struct dd_sizes {
  uint32_t nr_samples;                 /* S / downsampling */
  uint32_t nr_dispersed_samples;       /* S_disp */
  uint32_t nr_samples_subbanding;      /* S1 */
  uint32_t nr_dispersed_samples_subbanding; /* S1_disp */
  uint64_t in_pitch;                   /* input elements per channel row, single step */
  uint64_t out_pitch;                  /* floats per DM row */
  uint64_t sub_in_pitch;               /* input elements per channel row, step one */
  uint64_t sub_out_pitch;              /* floats per subband row of the intermediate */
  uint64_t input_elements;             /* single-step input */
  uint64_t output_elements;            /* single-step output */
  uint64_t sub_input_elements;         /* step-one input */
  uint64_t intermediate_elements;      /* step-one output = step-two input */
  uint64_t sub_output_elements;        /* step-two output */
  uint32_t max_shift;                  /* largest integer shift, single step */
  uint32_t max_shift_step_one;
  uint32_t max_shift_step_two;
  uint32_t input_bits;                 /* bits per input sample (dd_plan_set_input_bits) */
  /* Downsampling in the subband path (include/Dedispersion.hpp:119, 148,
   * 163, 169): step one writes S1 / downsampling samples per row into rows of
   * sub_out_pitch = pad(S1) floats, while step two reads rows of
   * step_two_in_pitch = pad(S1 / downsampling) floats.  The two agree only
   * at downsampling 1; each step follows its own CPU function, so a caller
   * chaining them at downsampling > 1 re-lays the rows in between, as a
   * caller of the reference must. */
  uint32_t nr_samples_step_one;        /* S1 / downsampling */
  uint64_t step_two_in_pitch;          /* floats per subband row of step two's input */
  uint64_t step_two_input_elements;    /* step-two input */
} dd_sizes;

typedef struct dd_plan dd_plan;

/* Human-readable text of the last error on this thread. */
const char *dd_last_error(void);
const char *dd_status_string(dd_status status);
uint32_t dd_abi_version(void);

/* Creates a plan on the current HIP device.  padding is the reference's
 * `padding` argument in bytes (a row of T elements is padded to a multiple
 * of padding/sizeof(T); 0 = unpadded).  The shift tables are built on the
 * device from the observation with the arithmetic of src/Shifts.cpp:20-42;
 * all channels start live and every synthesized beam maps to input beam
 * (sBeam % nr_beams). */
dd_status dd_plan_create(dd_plan **plan, const dd_observation *obs, uint32_t padding);
dd_status dd_plan_destroy(dd_plan *plan);
dd_status dd_plan_get_sizes(const dd_plan *plan, dd_sizes *sizes);

/* Bits per input sample, the reference's `inputBits` argument
 * (include/Dedispersion.hpp:59-60, :63-64): 8 (the default), 1, 2, 4 for
 * packed rows, or 32 for float input (the reference's templates with
 * I = float).  Float rows are pad(S_disp, padding / 4) floats; in_pitch,
 * sub_in_pitch and the input sizes then count floats, and only
 * dd_dedisperse_f32 and dd_subband_step_one_f32 take the plan's input (the
 * 8-bit entry points refuse it, and the float ones refuse other plans).  A
 * float row sum adds the live channels in order into a float accumulator
 * that starts at 0, the reference's sequential float sum (:80-106,
 * :123-146), so results are bit-identical to it for any float input.
 * Batches and streams refuse float plans.  Packed rows (1, 2, 4):  Packed rows hold 8 / input_bits samples per byte, sample x
 * in the field at bit (x % (8 / input_bits)) * input_bits of byte
 * x / (8 / input_bits), least significant bit first, read as an unsigned
 * value (include/Dedispersion.hpp:94-101 and :137-144).  A row of the
 * single-step and step-one inputs is then pad(ceil(S_disp / (8 / input_bits)),
 * padding) bytes (the reference pads S_disp / (8 / input_bits) rounded
 * down, :97, :140, which drops the last samples when 8 / input_bits does not
 * divide S_disp); in_pitch, sub_in_pitch and the input sizes count bytes.
 * Beams are input_bits-packed rows of C channels each, at the beam stride of
 * the 8-bit layout (the reference's single-step sub-byte path, :97, omits
 * the beam stride).  The intermediate keeps the 8-bit row padding (:148
 * pads S1 / (8 / input_bits) there, which overlaps rows).  Packed rows are
 * read only by the pipelined kernels, which unpack them while staging:
 * configurations with local 0, 2 or 3 run as local 1 with unroll capped at
 * 2 x threads_d1.  Streams (dd_stream_create) take 8-bit plans only. */
dd_status dd_plan_set_input_bits(dd_plan *plan, uint32_t input_bits);

/* Step one's delay per channel (a plan with a subband path):
 *   DD_STEP_ONE_RELATIVE (the default): the channel's delay minus that of
 *     its subband's top channel, (unsigned)((first + dm * step) *
 *     (shifts[ch] - shifts[top])), the reference's CPU function
 *     (include/Dedispersion.hpp:124); the coarse-DM delay between subbands
 *     is then applied nowhere, so a two-step output dedisperses correctly
 *     only at coarse DM 0;
 *   DD_STEP_ONE_ABSOLUTE: the channel's own delay, (unsigned)((first + dm *
 *     step) * shifts[ch]), the reference's OpenCL step-one kernel
 *     (include/Dedispersion.hpp:595-597, :731, :733), which makes step one
 *     plus step two a dedispersion at DM coarse + fine.
 * The same kernels run either way; only the delay table differs.  (The
 * reference's OpenCL kernel bakes firstDM and DMStep into its source with
 * six decimals, :208-219; this engine uses the float values.) */
typedef enum dd_step_one_shifts {
  DD_STEP_ONE_RELATIVE = 0,
  DD_STEP_ONE_ABSOLUTE = 1
} dd_step_one_shifts;
dd_status dd_plan_set_step_one_shifts(dd_plan *plan, dd_step_one_shifts delays);

/* Kernel configuration per mode (defaults are chosen at creation). */
dd_status dd_plan_set_conf(dd_plan *plan, dd_mode mode, const dd_conf *conf);
dd_status dd_plan_get_conf(const dd_plan *plan, dd_mode mode, dd_conf *conf);

/* Zapped channels, the reference's uint32 flags [pad(C, padding/4)]
 * (src/DedispersionTest.cpp:120); non-zero = zapped.  Host pointer.
 * Compacted into device-resident lists of the live channels (per subband for
 * step one), which the kernels walk. */
dd_status dd_plan_set_zapped_channels(dd_plan *plan, const uint32_t *zapped);
/* Beam mappings (host pointers, NULL = keep): single step
 * uint32 [sBeams][pad(C, padding/4)], step two uint32 [sBeams][pad(Nsb, padding/4)]
 * (src/DedispersionTest.cpp:121-122). */
dd_status dd_plan_set_beam_mapping(dd_plan *plan, const uint32_t *single_step,
                                   const uint32_t *step_two);
/* Replaces the device shift tables with caller tables (host pointers, NULL =
 * keep): shifts [pad(C, padding/4)] as getShifts returns, step-two shifts
 * [pad(Nsb, padding/4)] as getShiftsStepTwo returns. */
dd_status dd_plan_set_shifts(dd_plan *plan, const float *shifts, const float *shifts_step_two);
/* Copies the device tables back (host pointers, NULL = skip).  Integer
 * tables are [rows][dms] with the reference's integer shift per (row, DM):
 * single [C][D], step one [C][Dc] (relative to the subband's top channel),
 * step two [Nsb][Df]. */
dd_status dd_plan_get_shifts(const dd_plan *plan, float *shifts, float *shifts_step_two);
dd_status dd_plan_get_integer_shifts(const dd_plan *plan, dd_mode mode, uint32_t *table);

/* Single-step dedispersion of sBeams [first, first + n) (n = 0: all).
 * Replaces the `dedispersion` kernel (include/Dedispersion.hpp:203-540). */
dd_status dd_dedisperse(dd_plan *plan, const uint8_t *d_input, float *d_output,
                        uint32_t first_synthesized_beam, uint32_t nr_synthesized_beams,
                        void *stream);
/* Step one: channels into subbands at the coarse DMs, with the shift
 * relative to each subband's top channel (include/Dedispersion.hpp:124), or
 * the channel's own (dd_plan_set_step_one_shifts).
 * Replaces `dedispersionStepOne` (include/Dedispersion.hpp:543-882). */
dd_status dd_subband_step_one(dd_plan *plan, const uint8_t *d_input, float *d_intermediate,
                              void *stream);
/* The single step and step one over float input rows (a plan with
 * dd_plan_set_input_bits(plan, 32)); same layouts and semantics as
 * dd_dedisperse and dd_subband_step_one, with I = float
 * (include/Dedispersion.hpp:59-60, generators :63-64 with inputDataType
 * "float"). */
dd_status dd_dedisperse_f32(dd_plan *plan, const float *d_input, float *d_output,
                            uint32_t first_synthesized_beam, uint32_t nr_synthesized_beams,
                            void *stream);
dd_status dd_subband_step_one_f32(dd_plan *plan, const float *d_input, float *d_intermediate,
                                  void *stream);
/* Step two: subbands at the fine DMs.  Replaces `dedispersionStepTwo`
 * (include/Dedispersion.hpp:884-1122). */
dd_status dd_subband_step_two(dd_plan *plan, const float *d_intermediate, float *d_output,
                              uint32_t first_synthesized_beam, uint32_t nr_synthesized_beams,
                              void *stream);

/* ---- the step after: boxcar peak search over dedispersed rows ----
 *
 * No counterpart in the reference, whose output is the dedispersed array;
 * this is the search a pipeline runs next.  For each of nr_rows rows of
 * nr_samples non-negative integer-valued floats (the outputs above; a
 * fractional part is truncated; values below 2^24, and nr_samples x max^2
 * below 2^63), pitch floats apart, it finds the boxcar of width w in {1, 2, 4, ...,
 * 2^(nr_widths - 1)} (nr_widths <= 7) and start t (t + w <= nr_samples)
 * with the highest
 *     snr = (sum_{i<w} y(t + i) - w * mean) / (sigma * sqrt(w)),
 * mean and sigma (population) being those of the row, computed in double
 * from the exact integer sums.  Ties go to the smaller width, then the
 * smaller start; snr is 0 for a constant row.  d_peaks has nr_rows entries
 * (device memory). */
typedef struct dd_peak {
  float snr;
  float mean;
  float sigma;
  uint32_t sample;  /* start t of the best window */
  uint32_t width;   /* its width */
} dd_peak;

dd_status dd_detect_peaks(const float *d_rows, uint64_t nr_rows, uint32_t nr_samples, uint64_t pitch,
                          uint32_t nr_widths, dd_peak *d_peaks, void *stream);

/* The same search over many sets of rows (the outputs of many beams, sizes
 * free to differ) in one persistent launch per 64 sets: the same results as
 * one dd_detect_peaks call per set, without a launch's ramp and drain per
 * set.  A set with nr_rows 0 is skipped.  Also no counterpart in the
 * reference.  `jobs` is host memory, read before the call returns. */
typedef struct dd_detect_job {
  const float *d_rows;
  uint64_t nr_rows;
  uint64_t pitch;       /* floats between rows */
  uint32_t nr_samples;
  uint32_t reserved;    /* 0 */
  dd_peak *d_peaks;     /* nr_rows entries */
} dd_detect_job;

dd_status dd_detect_peaks_batch(const dd_detect_job *jobs, uint32_t nr_jobs, uint32_t nr_widths, void *stream);

/* ---- tuning harness (replaces src/DedispersionTuning.cpp:301-428) ---- */

/* Times one mode under `conf` without changing the plan's configuration:
 * `warmup` untimed launches, then `iterations` launches each bracketed by
 * HIP events on `stream`.  Buffers are those of the launch functions
 * (step two reads d_input as the intermediate).  Returns the mean and the
 * standard deviation of the per-launch times in milliseconds;
 * DD_ERR_CONFIGURATION if the configuration cannot run. */
dd_status dd_tune_time(dd_plan *plan, dd_mode mode, const dd_conf *conf, const void *d_input,
                       void *d_output, uint32_t warmup, uint32_t iterations, void *stream,
                       float *mean_ms, float *std_ms);

/* ---- batches: many plans in one persistent launch per kernel ----
 *
 * The reference launches one kernel per beam and step
 * (src/DedispersionTest.cpp:399-416, one enqueueNDRangeKernel per call).  A
 * batch gathers jobs -- the single-step or two-step path of a plan over
 * device buffers, plans free to differ in DMs and samples -- and runs every
 * tile of every job of one kernel kind in a single persistent launch over
 * work queues: one launch for all single-step jobs, one for all step-one
 * jobs, then one for all step-two jobs.  Results are bit-identical to
 * calling dd_dedisperse / dd_subband_step_one / dd_subband_step_two on each
 * job with the same configuration.  One configuration per kernel kind
 * applies to every job (DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO); it must
 * select the staged pipeline (local = 1, unroll <= 2 * threads_d1), or
 * dd_batch_run fails with DD_ERR_CONFIGURATION; a job whose row windows do
 * not fit it runs beside the batch kernels (dd_batch_set_conf).  Plans and
 * buffers must stay alive and unchanged
 * while they are in a batch.  A batch is used by one host thread at a time. */
typedef struct dd_batch dd_batch;

typedef struct dd_batch_info {
  uint32_t jobs;
  uint32_t tiles[3];        /* tiles per kernel kind (dd_mode order) */
  uint32_t workgroups[3];   /* resident workgroups of each persistent launch */
  uint32_t lds_bytes[3];
  uint64_t units;           /* output elements of the final arrays (DM trials x samples x beams) */
  uint32_t compact_intermediate; /* 1: step one writes uint16 sums (see dd_batch_set_compact_intermediate) */
  uint32_t pipeline_chunks;      /* chunks of the pipelined mode in effect, 0 = off */
  uint32_t solo_jobs;            /* jobs run beside the batch kernels (see dd_batch_set_conf) */
  uint32_t paired_step_two;      /* 1: the compact step two adds subbands in pairs (see dd_batch_set_paired_step_two) */
} dd_batch_info;

dd_status dd_batch_create(dd_batch **batch);
dd_status dd_batch_destroy(dd_batch *batch);
/* Adds the single-step path of `plan`: synthesized beams
 * [first_sbeam, first_sbeam + nr_sbeams) (nr_sbeams 0 = to the last). */
dd_status dd_batch_add_single_step(dd_batch *batch, dd_plan *plan, const uint8_t *d_input, float *d_output,
                                   uint32_t first_sbeam, uint32_t nr_sbeams);
/* Adds the two-step path of `plan`: step one into d_intermediate, then step
 * two of the synthesized beams [first_sbeam, first_sbeam + nr_sbeams). */
dd_status dd_batch_add_two_step(dd_batch *batch, dd_plan *plan, const uint8_t *d_input, float *d_intermediate,
                                float *d_output, uint32_t first_sbeam, uint32_t nr_sbeams);
/* One configuration per kernel kind for all jobs.  Batch kernels are the
 * staged (pipelined) kernels: a configuration with local != 1 or unroll > 2
 * x threads_d1 makes dd_batch_run fail with DD_ERR_CONFIGURATION.  A job
 * whose row windows (the sample tile plus the largest shift spread over a
 * DM tile) do not fit the staged kernels under this configuration runs
 * beside them, in the same run and graph, as the ordinary launches of its
 * plan's own configuration with the float intermediate (dd_batch_info's
 * solo_jobs counts them); fewer DMs per tile (threads_d1 x items_d1)
 * narrow the spread. */
dd_status dd_batch_set_conf(dd_batch *batch, dd_mode mode, const dd_conf *conf);
/* Compact intermediate (on by default): when 255 x the live channels of
 * every subband of every two-step job is below 2^16, step one writes its
 * sums as uint16 into the first half of each job's intermediate buffer
 * (same element offsets) and step two adds them as integers.  The final
 * outputs are bit-identical to the float pipeline; the intermediate
 * buffers then hold uint16 sums.  With enable = 0, or when a job does not
 * qualify, the batch keeps the float intermediate of the reference layout. */
dd_status dd_batch_set_compact_intermediate(dd_batch *batch, int enable);
/* Paired step two (off by default), with the compact intermediate: step two
 * first adds the uint16 sums of subbands 2k and 2k + 1 for every shift
 * difference the pair takes over a tile's DMs, and each DM then adds one
 * row per pair.  It applies when, for every job in the batch kernels, twice
 * 255 x the live channels of a subband is below 2^16, a pair's shift
 * difference takes at most 4 values per DM tile and the intermediate rows are
 * 8-element aligned; otherwise every subband is added on its own.  Results
 * are bit-identical either way.  On the 40-beam MI355X workload it is
 * slower than adding every subband (2.68 against 2.22 ms for the batch's step
 * two), so it is an option.  No counterpart in the reference, whose step two
 * (include/Dedispersion.hpp:155-176) adds every subband. */
dd_status dd_batch_set_paired_step_two(dd_batch *batch, int enable);
/* Pipelined mode (chunks >= 2, the default is 3; 0 or 1 = off): the two-step jobs are split,
 * in order, into `chunks` groups of near-equal work.  Step one of group c
 * and step two of group c run on two batch-owned streams, ordered by an
 * event, so step two of group c overlaps step one of group c + 1; each
 * launch is a plain grid over the group's tiles.  dd_batch_run forks from and
 * joins back into the caller's stream (and its graph holds the whole DAG).
 * dd_batch_run_mode keeps the persistent per-kind launches. */
dd_status dd_batch_set_pipeline(dd_batch *batch, uint32_t chunks);
/* In the pipelined mode, each workgroup of kernel kind `mode` requests at
 * least `bytes` of LDS (0, the default: what its kernel needs).  This caps
 * how many workgroups of the kind share a compute unit, so that workgroups
 * of the other kind fit beside them.  The persistent launches are not
 * affected.  A reserve above the device limit fails with
 * DD_ERR_CONFIGURATION. */
dd_status dd_batch_set_lds_reserve(dd_batch *batch, dd_mode mode, uint32_t bytes);
dd_status dd_batch_get_conf(const dd_batch *batch, dd_mode mode, dd_conf *conf);
/* Builds the work queues (first run, or after a change) and enqueues the
 * batch on `stream`.  use_graph != 0: the launches are captured once into a
 * HIP graph that later runs replay; the graph is rebuilt after any change. */
dd_status dd_batch_run(dd_batch *batch, int use_graph, void *stream);
/* Enqueues the persistent launch of one kernel kind only (no graph): for
 * timing a kind, or running step one and step two of a batch apart.  Step
 * two reads what step one of the same batch wrote: run DD_STEP_ONE first. */
dd_status dd_batch_run_mode(dd_batch *batch, dd_mode mode, void *stream);
/* The same work as one ordinary launch per job and kernel kind, in job
 * order (the reference's launch pattern), each with its plan's own
 * configuration and the float intermediate of the reference layout. */
dd_status dd_batch_run_per_job(dd_batch *batch, void *stream);
dd_status dd_batch_get_info(dd_batch *batch, dd_batch_info *info);
/* Where job `job` (in the order of the dd_batch_add_* calls) runs: beside = 1
 * when it runs beside the batch kernels as its plan's own launches (its
 * intermediate then holds the reference's floats); compact = 1 when its
 * intermediate holds the uint16 sums of the compact intermediate (a
 * two-step job in the batch kernels with compact intermediate in effect). */
dd_status dd_batch_get_job_placement(dd_batch *batch, uint32_t job, uint32_t *beside, uint32_t *compact);

/* Allocation-free run of a batch.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It enqueues on
 * `stream` the persistent launches of dd_batch_run_mode for every kernel
 * kind in order (single step, step one, step two) -- bit-identical results
 * -- with the work queues (job parameter blocks, tile lists, queue bounds,
 * claim counters) in a fixed device array of DD_BATCH_STATIC_BYTES per GPU
 * instead of buffers the batch allocates; it calls no device allocator.  The
 * queues are copied there from host memory on `stream` and the claim
 * counters are cleared by a kernel, so consecutive calls in one stream see
 * only their own queues; every call on a GPU shares that GPU's array, so
 * calls that overlap on one GPU (from several streams or threads) race.
 * Fails with DD_ERR_UNSUPPORTED when the queues need more than
 * DD_BATCH_STATIC_BYTES (the message gives the size) or when a job would run
 * beside the batch kernels (see dd_batch_set_conf).  The plans, their tables
 * and the job buffers are the caller's, set up beforehand. */
#define DD_BATCH_STATIC_BYTES (4u << 20)
dd_status dd_batch_run_static(dd_batch *batch, void *stream);

/* ---- streaming: continuous two-step dedispersion of chunked beams ----
 *
 * No counterpart in the reference (its driver processes one batch,
 * src/DedispersionTest.cpp:129-416).  `plan` is a two-step plan of ONE beam
 * whose nr_samples_per_batch S is the chunk length; every beam of the stream
 * shares it.  Chunk k of a beam is [C][S] uint8 host memory (row pitch
 * host_pitch bytes; pinned memory makes the copy asynchronous) and must stay
 * valid until its completion is reported.  The engine keeps the last
 * H = nr_dispersed_samples_subbanding - S samples of each beam on the device
 * and dedisperses [history | chunk]: the output rows of chunk k hold the
 * beam's samples [k S - H, k S - H + S) (zero history before the first
 * chunk).  Requests are admitted into slots (input, intermediate and output
 * buffers of the plan) while the HBM budget allows, in submission order,
 * chunks of one beam in order; each admitted request runs on one of
 * `hip_streams` streams, so requests admitted while others run overlap.  An
 * engine is used by one host thread at a time. */
typedef struct dd_stream dd_stream;

typedef struct dd_stream_info {
  uint32_t slots;            /* requests that can be in flight at once */
  uint32_t hip_streams;
  uint64_t chunk_samples;    /* S */
  uint64_t history_samples;  /* H */
  uint64_t slot_bytes;       /* device bytes per slot */
  uint64_t history_bytes;    /* device bytes of all beam histories */
  uint32_t queued;           /* submitted, waiting for a slot */
  uint32_t in_flight;        /* admitted, not released */
  uint64_t submitted, completed;
} dd_stream_info;

typedef struct dd_stream_event {
  uint64_t tag;              /* the caller's tag of the request */
  uint32_t beam;
  uint32_t slot;             /* pass to dd_stream_release once the output is consumed */
  uint64_t chunk;            /* index of the chunk within its beam */
  int64_t first_sample;      /* beam sample of output column 0: chunk * S - H */
  float *output;             /* device, step-two output [Dc * Df][out_pitch] of the plan */
  float *intermediate;       /* device, step-one output of the plan */
  double device_seconds;     /* admission to observed completion */
} dd_stream_event;

/* Slots = min(max_slots, (hbm_budget_bytes - history bytes) / slot bytes). */
dd_status dd_stream_create(dd_stream **stream, dd_plan *plan, uint32_t nr_beams, uint64_t hbm_budget_bytes,
                           uint32_t max_slots, uint32_t hip_streams);
dd_status dd_stream_destroy(dd_stream *stream);
dd_status dd_stream_get_info(const dd_stream *stream, dd_stream_info *info);
/* Queues the next chunk of `beam` and admits what can run. */
dd_status dd_stream_submit(dd_stream *stream, uint32_t beam, const uint8_t *host_chunk, uint64_t host_pitch,
                           uint64_t tag);
/* Admits what can run and reports up to max_events newly completed requests. */
dd_status dd_stream_progress(dd_stream *stream, dd_stream_event *events, uint32_t max_events, uint32_t *n_events);
/* Frees the slot of a reported request (its output is overwritten later). */
dd_status dd_stream_release(dd_stream *stream, uint32_t slot);
/* Waits until all admitted work has finished on the device. */
dd_status dd_stream_drain(dd_stream *stream);

/* ---- cost model of the learned tuner ----
 *
 * No counterpart in the reference, whose tuner (src/DedispersionTuning.cpp:
 * 201-427) times every configuration.  A small FP32 network -- `inputs`
 * features, `hidden_layers` ReLU layers of `hidden` units, one linear output
 * -- is trained on the device on timings of that exhaustive sweep and ranks
 * untimed configurations.  Parameters are one flat array: per layer, the
 * weights (fan_in x fan_out, row-major, W[i * fan_out + j] joins input i to
 * unit j) then the biases.  Device pointers are float arrays of n rows. */
typedef struct dd_mlp dd_mlp;
dd_status dd_mlp_create(dd_mlp **net, uint32_t inputs, uint32_t hidden, uint32_t hidden_layers);
dd_status dd_mlp_destroy(dd_mlp *net);
dd_status dd_mlp_param_count(const dd_mlp *net, uint32_t *count);
dd_status dd_mlp_set_params(dd_mlp *net, const float *params); /* host array; resets the optimizer */
dd_status dd_mlp_get_params(const dd_mlp *net, float *params); /* host array */
/* He-uniform weights from `seed`, zero biases. */
dd_status dd_mlp_init(dd_mlp *net, uint64_t seed);
/* d_y[r] = network(d_x[r][0..inputs)) for r < n. */
dd_status dd_mlp_forward(dd_mlp *net, const float *d_x, uint32_t n, float *d_y, void *stream);
/* Mean squared error L = (1/n) sum (y_r - t_r)^2 into d_loss[0] and dL/dparams
 * into d_grad (flat parameter layout). */
dd_status dd_mlp_loss_grad(dd_mlp *net, const float *d_x, const float *d_target, uint32_t n, float *d_loss,
                           float *d_grad, void *stream);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with gradient d_grad. */
dd_status dd_mlp_adam_step(dd_mlp *net, const float *d_grad, float lr, void *stream);
/* `steps` full-batch Adam steps; the loss before each step goes to the host
 * array `losses` [steps] when it is not NULL.  Synchronizes the stream. */
dd_status dd_mlp_train(dd_mlp *net, const float *d_x, const float *d_target, uint32_t n, uint32_t steps, float lr,
                       float *losses, void *stream);

/* ---- synthetic inputs (seeded, reproducible) ---- */

/* Fills rows x cols uint8 values (row pitch `pitch` bytes) with
 * round(mean + std * N(0,1)) clipped to [0, 255]; the normal deviates come
 * from a counter-based generator keyed by (seed, row, col). */
dd_status dd_synth_gaussian_u8(uint8_t *d_data, uint64_t rows, uint64_t cols, uint64_t pitch,
                               float mean, float std_dev, uint64_t seed, void *stream);
/* Fills uint8 values uniform in [0, 255] keyed by (seed, row, col). */
dd_status dd_synth_uniform_u8(uint8_t *d_data, uint64_t rows, uint64_t cols, uint64_t pitch,
                              uint64_t seed, void *stream);
/* Adds `amplitude` (clipped at 255) to a dispersed boxcar: for every channel
 * c of one beam ([C][pitch] bytes), samples [t0 + shift(c), + width) below
 * row_len, where shift(c) = (unsigned)(dm_value * shifts[c]) with the plan's
 * single-step table (include/Dedispersion.hpp:81). */
dd_status dd_synth_inject_pulse(dd_plan *plan, uint8_t *d_beam_input, uint64_t pitch,
                                uint32_t row_len, float dm_value, uint32_t t0, uint32_t width,
                                uint32_t amplitude, void *stream);

/* ---- measurement ----
 *
 * No counterpart in the reference.  Reads `bytes` (a multiple of 16) of a
 * 16-byte aligned device buffer with coalesced 16-byte loads and writes one
 * partial sum per thread to d_sink (DD_READ_PROBE_SINK_FLOATS floats): the
 * read bandwidth a streaming kernel reaches on the device, against which the
 * read-bound kernels are measured beside the nominal HBM peak. */
#define DD_READ_PROBE_SINK_FLOATS 2097152u
dd_status dd_read_probe(const void *d_data, uint64_t bytes, float *d_sink, void *stream);

#ifdef __cplusplus
}
#endif
#endif
