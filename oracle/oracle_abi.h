/* Test infrastructure only: the C ABI of the CPU oracle.
 *
 * Two shared libraries implement exactly this ABI:
 *   oracle/liboracle.so        - the independent CPU restatement (gaffa_oracle.cpp)
 *   oracle/_ref/libgaffa_ref.so - a shim over gaffa's own CPU sources, compiled
 *                                 from the reference tree by oracle/Makefile
 * Tests load both through oracle/oracle.py and check one against the other and
 * against the golden vectors under tests/golden/.  Nothing in the product
 * (gaffa_amd/) may load either library.
 *
 * Every function returns 0 on success or a positive error code naming the C++
 * exception class the reference throws for the same condition:
 *   1 invalid_argument, 2 overflow_error, 3 out_of_range, 4 runtime_error,
 *   5 logic_error, 6 other.
 * orc_last_error() returns the message of the last failure on this thread.
 * Output arrays are caller-owned; a *_count out-parameter always receives the
 * required element count, and a call with a too-small capacity returns 7
 * (ORC_TOO_SMALL) without writing, so callers can size and retry.
 */
#ifndef GAFFA_ORACLE_ABI_H
#define GAFFA_ORACLE_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_OK = 0,
  ORC_INVALID_ARGUMENT = 1,
  ORC_OVERFLOW = 2,
  ORC_OUT_OF_RANGE = 3,
  ORC_RUNTIME = 4,
  ORC_LOGIC = 5,
  ORC_OTHER = 6,
  ORC_TOO_SMALL = 7,
};

/* Same field order and types as gaffa::FfaSearchTask (ffa_plan.h:8-27). */
typedef struct orc_task {
  double downsample_factor;
  double effective_tsamp;
  uint64_t input_nsamples;
  uint64_t prepared_nsamples;
  uint64_t bins;
  uint64_t rows;
  uint64_t rows_eval;
  double period_begin;
  double period_end;
} orc_task;

/* Same fields as gaffa::FfaPeak (ffa_peak.h:8-19). */
typedef struct orc_peak {
  double period;
  double frequency;
  uint64_t width;
  double duty_cycle;
  uint64_t width_index;
  uint64_t period_index;
  uint64_t phase;
  uint64_t shift;
  uint64_t bins;
  float snr;
} orc_peak;

/* gaffa::DmPeak (dm_search.h:16-20). */
typedef struct orc_dm_peak {
  double dm;
  uint64_t dm_index;
  orc_peak peak;
} orc_dm_peak;

/* gaffa::Candidate (candidate.h:12-28). */
typedef struct orc_candidate {
  double dm;
  uint64_t dm_index;
  double period;
  double frequency;
  uint64_t width;
  double duty_cycle;
  float snr;
  uint64_t peak_count;
  uint64_t dm_index_min;
  uint64_t dm_index_max;
  double frequency_min;
  double frequency_max;
  orc_dm_peak best_peak;
} orc_candidate;

/* gaffa::HarmonicOptions and HarmonicRelation (harmonic.h:19-48). */
typedef struct orc_harmonic_options {
  uint64_t max_harmonic;
  uint64_t denominator_max;
  double frequency_tolerance_bins;
  double phase_distance_max;
  double dm_distance_max;
  int32_t use_snr_consistency;
  double snr_distance_max;
} orc_harmonic_options;

typedef struct orc_harmonic {
  int32_t is_harmonic;
  uint64_t parent_index;
  int32_t numerator;
  int32_t denominator;
  double frequency_error_bins;
  double phase_distance;
  double dm_distance;
  double expected_snr;
  double snr_distance;
} orc_harmonic;

/* gaffa::RiptideFfaPlanOptions (ffa_plan.h:34-47). */
typedef struct orc_plan_options {
  double period_min;
  double period_max;
  uint64_t bins_min;
  uint64_t bins_max;
  uint64_t min_periods;
  double duty_cycle_max;
  double width_trial_spacing;
  uint64_t max_tasks;
} orc_plan_options;

/* A two-step riptide preprocess plan (preprocessing.h:11-42) restricted to the
 * shapes make_riptide_preprocess_plan() emits: an optional running-median
 * detrend followed by an optional normalise. */
typedef struct orc_preprocess_plan {
  int32_t detrend;          /* 1: DetrendRunningMedian step present */
  uint64_t window_samples;  /* odd */
  uint64_t min_points;      /* odd */
  int32_t normalise;        /* 1: Normalise step present */
  int32_t reject_constant;  /* NormaliseOptions::reject_constant */
} orc_preprocess_plan;

const char* orc_last_error(void);
const char* orc_implementation(void); /* "port" or "reference" */

/* ---- dedispersion (dedispersion_delay.h, dedispersion.cpp) ---- */
int orc_delay_bins(double dm, double frequency_mhz, double ref_frequency_mhz,
                   double tsamp, int32_t* out);
int orc_multi_dm_delay_table(const double* frequency_mhz, uint64_t nchans,
                             double dm_low, double dm_step, uint64_t ndm,
                             double ref_frequency_mhz, double tsamp,
                             uint64_t chan_begin, uint64_t chan_end,
                             int32_t* out, uint64_t capacity, uint64_t* count);
/* kind: 0 brute force multi-DM, 1 subband. dtype: 0 uint8, 1 uint16, 2 float32.
 * out is uint32 for integer inputs and float for float32 input. */
int orc_dedisperse(int32_t kind, int32_t dtype, const void* samples,
                   uint64_t nsamples, uint64_t nchans,
                   const double* frequency_mhz, double dm_low, double dm_step,
                   uint64_t ndm, double ref_frequency_mhz, double tsamp,
                   uint64_t chan_begin, uint64_t chan_end,
                   uint64_t subband_channels, uint64_t ndm_per_nominal,
                   void* out, uint64_t capacity, uint64_t* out_nsamples);
int orc_dedisperse_single(int32_t dtype, const void* samples,
                          uint64_t nsamples, uint64_t nchans,
                          const double* frequency_mhz, double dm,
                          double ref_frequency_mhz, double tsamp,
                          uint64_t chan_begin, uint64_t chan_end, void* out,
                          uint64_t capacity, uint64_t* out_nsamples);

/* ---- time series (time_series.cpp) ---- */
int orc_time_series_stats(const float* input, uint64_t n, double* mean,
                          double* variance, double* stddev);
int orc_normalise(const float* input, uint64_t n, int32_t reject_constant,
                  float* out);
int orc_downsampled_size(uint64_t n, double factor, uint64_t* out);
int orc_downsampled_variance(uint64_t n, double factor, double* out);
int orc_downsample(const float* input, uint64_t n, double factor, float* out,
                   uint64_t capacity, uint64_t* count);

/* ---- preprocessing (preprocessing.cpp) ---- */
int orc_running_median_window(double window_seconds, double tsamp,
                              uint64_t* window_samples);
int orc_detrend(const float* input, uint64_t n, uint64_t window_samples,
                uint64_t min_points, float* out);
int orc_preprocess(const float* input, uint64_t n, double tsamp,
                   const orc_preprocess_plan* plan, float* out);

/* ---- FFA (ffa_plan.cpp, ffa.cpp, ffa_executor.cpp, ffa_detection.cpp) ---- */
int orc_make_riptide_plan(uint64_t nsamples, double tsamp,
                          const orc_plan_options* options, orc_task* tasks,
                          uint64_t task_capacity, uint64_t* task_count,
                          uint64_t* widths, uint64_t width_capacity,
                          uint64_t* width_count);
int orc_ffa_transform(const float* input, uint64_t rows, uint64_t bins,
                      float* out);
/* The first rows_eval rows of each task's transform, concatenated in plan
 * order, exactly as for_each_ffa_row_cpu streams them. */
int orc_ffa_rows(const float* series, uint64_t n, const orc_task* tasks,
                 uint64_t ntasks, float* out, uint64_t capacity,
                 uint64_t* count);
int orc_task_stdnoise(const orc_task* task, float* out);
int orc_find_ffa_peaks(const float* transform, uint64_t rows, uint64_t bins,
                       const orc_task* task, const uint64_t* widths,
                       uint64_t nwidths, float stdnoise, float snr_threshold,
                       uint64_t max_peaks, orc_peak* peaks, uint64_t capacity,
                       uint64_t* count);
int orc_search_ffa(const float* series, uint64_t n, const orc_task* tasks,
                   uint64_t ntasks, const uint64_t* widths, uint64_t nwidths,
                   float snr_threshold, uint64_t max_peaks, orc_peak* peaks,
                   uint64_t capacity, uint64_t* count);
int orc_sort_peaks(orc_peak* peaks, uint64_t count);

/* ---- candidates (candidate.cpp:175-212, harmonic.cpp:245-311) ----
 * Only the reference build implements these (the port returns ORC_OTHER):
 * the product's own host code is the restatement, checked against it. */
int orc_select_candidates(const orc_dm_peak* peaks, uint64_t npeaks,
                          double observation_seconds, double frequency_radius,
                          uint64_t dm_radius, int32_t across_widths,
                          uint64_t max_candidates, orc_candidate* out,
                          uint64_t capacity, uint64_t* count);
int orc_flag_harmonics(const orc_candidate* candidates, uint64_t n,
                       double observation_seconds, double frequency_low_mhz,
                       double frequency_high_mhz,
                       const orc_harmonic_options* options,
                       orc_candidate* out, orc_harmonic* relations,
                       uint64_t capacity, uint64_t* count);

/* Aligned single-DM spectrum (dedispersion.cpp:249-289): out [n_out][ce - cb]
 * of the input element type (dtype 0 u8, 1 u16, 2 f32). */
int orc_dedisperse_spectrum(int32_t dtype, const void* samples, uint64_t nsamples,
                            uint64_t nchans, const double* freq, double dm, double ref,
                            double tsamp, uint64_t cb, uint64_t ce, void* out,
                            uint64_t capacity, uint64_t* out_nsamples);

/* ---- folding (folding.cpp:300-470; reference build only) ----
 * dtype: 0 u8, 1 u16, 2 u32 (time series only), 3 f32.  Time series:
 * profile/exposure [nbin].  Spectrum: data [nsamples][nchans], cube
 * [nsubint][out_chans][nbin], exposure [nsubint][nbin]. */
int orc_fold_time_series(const void* data, int32_t dtype, uint64_t n, double period,
                         double tsamp, uint64_t nbin, uint64_t start_sample,
                         float* profile, double* exposure);
int orc_fold_spectrum(const void* data, int32_t dtype, uint64_t nsamples, uint64_t nchans,
                      double period, double tsamp, uint64_t nbin, uint64_t subint_samples,
                      uint64_t start_sample, uint64_t output_channels, float* cube,
                      double* exposure, uint64_t cube_capacity, uint64_t* nsubint);

/* ---- DM search (dm_search.cpp) ---- */
int orc_find_dm_peaks(const uint32_t* plane, uint64_t ndm, uint64_t nsamples,
                      const double* dms, double tsamp,
                      const orc_plan_options* plan_options,
                      const orc_preprocess_plan* preprocess,
                      float snr_threshold, uint64_t max_peaks,
                      orc_dm_peak* peaks, uint64_t capacity, uint64_t* count);

/* ---- reference-only: gaffa's SIGPROC reader (io/filterbank.cpp), exported
 * by the ref shim alone to pin the product's native reader. */
typedef struct orc_fb_header {
  int64_t header_size, nsamples;
  int32_t telescope_id, machine_id, data_type, barycentric, pulsarcentric, ibeam, nbeams,
      npuls, nbins, nbits, nifs, nchans;
  double az_start, za_start, src_raj, src_dej, tstart, tsamp, fch1, foff, refdm, period;
  int32_t uses_frequency_table;
  char rawdatafile[256];
  char source_name[256];
} orc_fb_header;

/* Reads `path` with channel_order preserve_file_order (preserve != 0) or
 * frequency_ascending; frequency table to freq (capacity freq_cap), sample
 * bytes to data (capacity data_cap, NULL = header only). */
int orc_read_filterbank(const char* path, int32_t preserve, orc_fb_header* header,
                        double* freq, uint64_t freq_cap, void* data, uint64_t data_cap,
                        uint64_t* data_bytes);

/* ---- oracle-only extras (no reference counterpart; the ref shim returns 6) */
/* FFA transform with FP64 accumulation and the same index schedule. */
int orc_ffa_transform_f64(const double* input, uint64_t rows, uint64_t bins,
                          double* out);

/* Seeded synthetic filterbank, integer-exact (see gaffa_oracle.cpp). */
typedef struct orc_pulsar {
  double period_s;
  double dm;
  double duty;      /* Gaussian FWHM as a fraction of the period */
  double amp_sigma; /* per-channel peak amplitude in units of the noise sigma */
  double phase0;    /* pulse centre phase at t = 0 */
} orc_pulsar;

int orc_synth_filterbank(uint64_t nsamples, uint64_t nchans, double fch1,
                         double foff, double tsamp, uint64_t seed,
                         const orc_pulsar* pulsars, uint64_t npulsars,
                         const uint64_t* rfi_channels, uint64_t nrfi,
                         int32_t rfi_offset, uint64_t t_begin, uint64_t t_count,
                         uint8_t* out);

#ifdef __cplusplus
}
#endif

#endif
