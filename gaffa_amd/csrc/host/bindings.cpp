// pybind11 module gaffa_amd._core: the Python surface of gaffa's _core for
// the search hot path (src/gaffa/python/dedispersion_bindings.cpp:594-672,
// ffa_bindings.cpp:169-217), running on the GPU through the C++ host layer.
// Numpy inputs are zero-copy views that must be C-contiguous; outputs own
// their memory through a capsule; the GIL is released around device work.
#include "gaffa/candidate.h"
#include "gaffa/dedispersion.h"
#include "gaffa/ffa_batch.h"
#include "gaffa/ffa_gpu.h"
#include "gaffa/folding.h"
#include "gaffa/pipeline.h"
#include "gaffa/preprocessing.h"
#include "gaffa/scorer.h"
#include "gaffa/synth.h"

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <optional>
#include <string>
#include <vector>

namespace py = pybind11;
using namespace gaffa;

namespace {

// gaffa's backend names (dedispersion_bindings.cpp:56-64): "cuda" selects the
// device path, here HIP on the GPU, with "gpu" as a synonym; the result
// reports the name the caller used.  "cpu" is gaffa's own CPU code, which
// gaffa_amd does not ship (its CPU restatement is test-only), so it is refused
// with a ValueError rather than computed.
std::string parse_backend(const std::string& value, const char* what) {
  if (value == "gpu" || value == "cuda") return value;
  if (value == "cpu") {
    throw py::value_error(std::string(what) +
                          " backend 'cpu' is gaffa's CPU path, which gaffa_amd does not "
                          "provide; use backend='cuda' (or 'gpu')");
  }
  throw py::value_error(std::string("unknown ") + what + " backend: " + value);
}

// gaffa's reader option names and messages (filterbank_bindings.cpp:23-43).
ChannelOrderPolicy parse_channel_order(const std::string& value) {
  if (value == "frequency_ascending") return ChannelOrderPolicy::FrequencyAscending;
  if (value == "preserve_file_order") return ChannelOrderPolicy::PreserveFileOrder;
  throw py::value_error("unknown channel_order: " + value);
}

ReverseBackend parse_reverse_backend(const std::string& value) {
  if (value == "auto") return ReverseBackend::Auto;
  if (value == "cpu_scalar") return ReverseBackend::CpuScalar;
  if (value == "cpu_openmp") return ReverseBackend::CpuOpenmp;
  throw py::value_error("unknown reverse_backend: " + value);
}

template <typename T>
py::array_t<T> owning_array(std::vector<T>&& v, std::vector<py::ssize_t> shape) {
  auto* owner = new std::vector<T>(std::move(v));
  py::capsule cap(owner, [](void* p) { delete static_cast<std::vector<T>*>(p); });
  std::vector<py::ssize_t> strides(shape.size());
  py::ssize_t s = sizeof(T);
  for (std::size_t i = shape.size(); i-- > 0;) {
    strides[i] = s;
    s *= shape[i];
  }
  return py::array_t<T>(shape, strides, owner->data(), cap);
}

void require_c_contiguous(const py::buffer_info& info, const char* name) {
  py::ssize_t expect = info.itemsize;
  for (py::ssize_t i = info.ndim; i-- > 0;) {
    if (info.shape[i] > 1 && info.strides[i] != expect) {
      throw py::value_error(std::string(name) +
                            " must be C-contiguous; use numpy.ascontiguousarray explicitly if a "
                            "copy is intended");
    }
    expect *= info.shape[i];
  }
}

// ------------------------------------------------------------ dedispersion --
struct PyDedispersedResult {
  py::array data;
  std::string backend;
  double dm_low = 0.0;
  double dm_step = 0.0;
  std::size_t ndm = 0;
  std::size_t nsamples = 0;
  double tsamp = 0.0;
};

struct FilterbankInput {
  HostSampleView<std::uint8_t> samples;
  std::vector<double> freq;
  double tsamp = 0.0;
  py::array keep;
};

// Accepts gaffa_amd.io.Filterbank (or any object with .header and .data):
// data is (nsamples, nifs, nchans) C-contiguous, header carries tsamp, nbits
// and the ascending frequency_table (dedispersion_bindings.cpp:224-259).
FilterbankInput filterbank_input(const py::object& fb) {
  if (!py::hasattr(fb, "header") || !py::hasattr(fb, "data")) {
    throw py::type_error("filterbank must provide .header and .data");
  }
  const py::object header = fb.attr("header");
  const int nbits = header.attr("nbits").cast<int>();
  if (nbits != 8) {
    throw py::type_error("GPU dedispersion supports 8-bit filterbanks (uint8 samples)");
  }
  FilterbankInput in;
  in.tsamp = header.attr("tsamp").cast<double>();
  if (!(in.tsamp > 0.0) || !std::isfinite(in.tsamp)) {
    throw std::invalid_argument("filterbank tsamp must be positive");
  }
  in.freq = header.attr("frequency_table").cast<std::vector<double>>();
  py::object data = fb.attr("data");
  if (!py::isinstance<py::array>(data)) throw py::type_error("filterbank data must be a numpy.ndarray");
  in.keep = py::reinterpret_borrow<py::array>(data);
  const py::buffer_info info = in.keep.request();
  if (info.ndim != 3) throw std::invalid_argument("filterbank data must be a 3D numpy array");
  if (info.itemsize != 1 || info.format != py::format_descriptor<std::uint8_t>::format()) {
    throw std::invalid_argument("filterbank data dtype does not match header");
  }
  if (info.strides[2] != 1 || info.strides[1] != info.shape[2] ||
      info.strides[0] != info.shape[1] * info.shape[2]) {
    throw std::invalid_argument("filterbank data must be C-contiguous in (time, if, channel) order");
  }
  const SampleShape shape{static_cast<std::size_t>(info.shape[0]),
                          static_cast<std::size_t>(info.shape[1]),
                          static_cast<std::size_t>(info.shape[2])};
  if (in.freq.size() != shape.nchans) {
    throw std::invalid_argument("filterbank frequency table length must match nchans");
  }
  in.samples = make_host_sample_view<std::uint8_t>(
      std::span<const std::uint8_t>(static_cast<const std::uint8_t*>(info.ptr),
                                    sample_element_count(shape)),
      shape);
  return in;
}

double reference_frequency(const std::vector<double>& f) {
  if (f.empty()) throw std::invalid_argument("dedispersion requires at least one channel");
  return *std::max_element(f.begin(), f.end());
}

PyDedispersedResult make_result(DedispersedResult<std::uint32_t>&& r, double dm_low,
                                double dm_step, double tsamp, std::string backend) {
  PyDedispersedResult p;
  p.ndm = r.shape.ndm;
  p.nsamples = r.shape.nsamples;
  p.data = owning_array(std::move(r.data), {static_cast<py::ssize_t>(p.ndm),
                                            static_cast<py::ssize_t>(p.nsamples)});
  p.backend = std::move(backend);
  p.dm_low = dm_low;
  p.dm_step = dm_step;
  p.tsamp = tsamp;
  return p;
}

PyDedispersedResult py_single(const py::object& fb, double dm, const std::string& backend,
                              int device_id, std::size_t tpb, std::size_t tile) {
  std::string name = parse_backend(backend, "dedispersion");
  FilterbankInput in = filterbank_input(fb);
  const SingleDmDedispersionPlan plan{dm, reference_frequency(in.freq), in.tsamp, 0,
                                      in.samples.shape.nchans};
  DedispersedResult<std::uint32_t> r;
  {
    py::gil_scoped_release release;
    r = dedisperse_single_dm_gpu(in.samples, in.freq, plan,
                                 GpuDedispersionOptions{device_id, tpb, tile, nullptr});
  }
  return make_result(std::move(r), dm, 0.0, in.tsamp, std::move(name));
}

PyDedispersedResult py_multi(const py::object& fb, double dm_low, double dm_step, std::size_t ndm,
                             const std::string& backend, int device_id, std::size_t tpb,
                             std::size_t tile) {
  std::string name = parse_backend(backend, "dedispersion");
  FilterbankInput in = filterbank_input(fb);
  const MultiDmDedispersionPlan plan{dm_low, dm_step, ndm, reference_frequency(in.freq),
                                     in.tsamp, 0, in.samples.shape.nchans};
  DedispersedResult<std::uint32_t> r;
  {
    py::gil_scoped_release release;
    r = dedisperse_multi_dm_gpu(in.samples, in.freq, plan,
                                GpuDedispersionOptions{device_id, tpb, tile, nullptr});
  }
  return make_result(std::move(r), dm_low, dm_step, in.tsamp, std::move(name));
}

PyDedispersedResult py_subband(const py::object& fb, double dm_low, double dm_step,
                               std::size_t ndm, const std::string& backend,
                               std::size_t subband_channels, std::size_t ndm_per_nominal,
                               int device_id, std::size_t tpb, std::size_t tile) {
  std::string name = parse_backend(backend, "dedispersion");
  FilterbankInput in = filterbank_input(fb);
  const MultiDmDedispersionPlan plan{dm_low, dm_step, ndm, reference_frequency(in.freq),
                                     in.tsamp, 0, in.samples.shape.nchans};
  DedispersedResult<std::uint32_t> r;
  {
    py::gil_scoped_release release;
    r = dedisperse_subband_gpu(in.samples, in.freq, plan,
                               SubbandDedispersionOptions{subband_channels, ndm_per_nominal},
                               GpuDedispersionOptions{device_id, tpb, tile, nullptr});
  }
  return make_result(std::move(r), dm_low, dm_step, in.tsamp, std::move(name));
}

// ----------------------------------------------------------------------- FFA --
const py::array& require_series(const py::object& obj, py::array& holder, int ndim) {
  if (!py::isinstance<py::array>(obj)) throw py::type_error("FFA time_series must be a numpy.ndarray");
  holder = py::reinterpret_borrow<py::array>(obj);
  const py::buffer_info info = holder.request();
  if (info.ndim != ndim) {
    throw py::value_error(ndim == 1 ? "FFA time_series must be a 1D array"
                                    : "FFA batch must be a 2D array");
  }
  if (info.size <= 0) throw py::value_error("FFA time_series must not be empty");
  if (info.itemsize != 4 || info.format != py::format_descriptor<float>::format()) {
    throw py::type_error("FFA time_series must have dtype float32");
  }
  require_c_contiguous(info, "FFA time_series");
  return holder;
}

std::size_t peak_limit(const py::object& max_peaks) {
  if (max_peaks.is_none()) return 0;
  const auto v = max_peaks.cast<long long>();
  if (v <= 0) throw py::value_error("max_peaks must be positive or None");
  return static_cast<std::size_t>(v);
}

// The argument checks of a single-series search, in gaffa's order and words
// (ffa_bindings.cpp:55-71, ffa_executor.cpp:31-35), without searching: the
// Python wrapper runs them before it decides on the backend, as gaffa's
// bindings validate whichever backend was chosen.
void py_check_ffa_search(const py::object& ts, const FfaSearchPlan& plan,
                         const py::object& max_peaks) {
  py::array holder;
  require_series(ts, holder, 1);
  const std::size_t n = static_cast<std::size_t>(holder.request().shape[0]);
  for (const FfaSearchTask& t : plan.tasks)
    if (t.input_nsamples != n)
      throw std::invalid_argument("FFA executor task input_nsamples must match time series size");
  (void)peak_limit(max_peaks);
}

std::vector<FfaPeak> py_ffa_search(const py::object& ts, const FfaSearchPlan& plan, int device_id,
                                   float snr_threshold, const py::object& max_peaks) {
  py::array holder;
  require_series(ts, holder, 1);
  const py::buffer_info info = holder.request();
  const std::size_t limit = peak_limit(max_peaks);
  const std::span<const float> s(static_cast<const float*>(info.ptr),
                                 static_cast<std::size_t>(info.shape[0]));
  py::gil_scoped_release release;
  return search_ffa_gpu_host(s, plan, FfaSearchOptions{snr_threshold, limit}, device_id).peaks;
}

// Batch of equal-length series (rows of a 2D float32 array) -> list of
// (series_index, FfaPeak), sorted by (series, peak order).
std::vector<std::pair<std::size_t, FfaPeak>> py_ffa_search_batch(
    const py::object& batch, const FfaSearchPlan& plan, int device_id, float snr_threshold,
    const py::object& max_peaks, std::size_t series_tile) {
  py::array holder;
  require_series(batch, holder, 2);
  const py::buffer_info info = holder.request();
  const std::size_t ns = static_cast<std::size_t>(info.shape[0]);
  const std::size_t n = static_cast<std::size_t>(info.shape[1]);
  const std::size_t limit = peak_limit(max_peaks);
  std::vector<std::pair<std::size_t, FfaPeak>> out;
  {
    py::gil_scoped_release release;
    GpuDeviceScope scope(device_id);
    GpuDeviceBuffer<float> d = upload(static_cast<const float*>(info.ptr), ns * n);
    GpuFfaExecutionOptions eo;
    eo.series_tile_size = series_tile;
    GpuFfaProgram program(plan, GpuFfaProgramOptions{device_id}, eo);
    const FfaBatchSearchResult r = run_ffa_batch_gpu(
        program, GpuTimeSeriesBatchView{d.data(), ns, n, n, device_id},
        FfaSearchOptions{snr_threshold, limit});
    out.reserve(r.peaks.size());
    for (const auto& p : r.peaks) out.emplace_back(p.series_index, p.peak);
  }
  return out;
}

// rows_eval x bins final transform rows of task `task` for each series.
py::array_t<float> py_ffa_transform_rows(const py::object& batch, const FfaSearchPlan& plan,
                                         std::size_t task, int device_id) {
  py::array holder;
  require_series(batch, holder, 2);
  const py::buffer_info info = holder.request();
  const std::size_t ns = static_cast<std::size_t>(info.shape[0]);
  const std::size_t n = static_cast<std::size_t>(info.shape[1]);
  if (task >= plan.tasks.size()) throw py::index_error("FFA task index out of range");
  const FfaSearchTask& t = plan.tasks[task];
  std::vector<float> host(ns * t.rows_eval * t.bins);
  {
    py::gil_scoped_release release;
    GpuDeviceScope scope(device_id);
    GpuDeviceBuffer<float> d = upload(static_cast<const float*>(info.ptr), ns * n);
    GpuDeviceBuffer<float> out(host.size());
    GpuFfaExecutionOptions eo;
    eo.series_tile_size = ns;
    GpuFfaProgram program(plan, GpuFfaProgramOptions{device_id}, eo);
    if (n != program.execution_plan().input_nsamples()) {
      throw std::invalid_argument("FFA executor task input_nsamples must match time series size");
    }
    program.transform_rows(d.data(), ns, n, task, out.data());
    copy_to_host(host.data(), out.data(), host.size() * 4, nullptr);
  }
  return owning_array(std::move(host), {static_cast<py::ssize_t>(ns),
                                        static_cast<py::ssize_t>(t.rows_eval),
                                        static_cast<py::ssize_t>(t.bins)});
}

// ---------------------------------------------------------------- preprocess --
PreprocessPlan preprocess_plan_from(std::optional<std::size_t> window_samples,
                                    std::size_t min_points, bool normalise, bool reject_constant) {
  PreprocessPlan plan;
  if (window_samples.has_value()) {
    PreprocessStep s;
    s.kind = PreprocessStepKind::DetrendRunningMedian;
    s.detrend_running_median.window_samples = *window_samples;
    s.detrend_running_median.min_points = min_points;
    plan.steps.push_back(s);
  }
  if (normalise) {
    PreprocessStep s;
    s.kind = PreprocessStepKind::Normalise;
    s.normalise.reject_constant = reject_constant;
    plan.steps.push_back(s);
  }
  return plan;
}

// [nseries][nsamples] uint32 or float32 -> preprocessed float32 copy.
py::array_t<float> py_preprocess_batch(const py::array& batch, std::optional<std::size_t> window,
                                       std::size_t min_points, bool normalise,
                                       bool reject_constant, int device_id) {
  const py::buffer_info info = batch.request();
  if (info.ndim != 2) throw py::value_error("preprocess batch must be a 2D array");
  require_c_contiguous(info, "preprocess batch");
  const bool is_u32 = info.format == py::format_descriptor<std::uint32_t>::format() ||
                      (info.itemsize == 4 && (info.format == "I" || info.format == "L"));
  const bool is_f32 = info.format == py::format_descriptor<float>::format();
  if (info.itemsize != 4 || !(is_u32 || is_f32)) {
    throw py::type_error("preprocess batch must have dtype uint32 or float32");
  }
  const std::size_t ns = static_cast<std::size_t>(info.shape[0]);
  const std::size_t n = static_cast<std::size_t>(info.shape[1]);
  std::vector<float> host(ns * n);
  const PreprocessPlan plan = preprocess_plan_from(window, min_points, normalise, reject_constant);
  {
    py::gil_scoped_release release;
    GpuDeviceScope scope(device_id);
    GpuDeviceBuffer<std::uint32_t> d =
        upload(static_cast<const std::uint32_t*>(info.ptr), ns * n);
    if (plan.steps.empty()) {
      if (!is_u32) throw std::invalid_argument("preprocess plan must contain at least one step");
      gaffa_hip_preprocess_args a{};
      a.data = d.data();
      a.nseries = ns;
      a.nsamples = n;
      a.stride = n;
      a.input_is_u32 = 1;
      a.scrunch_factor = 1;
      a.median_window = 1;
      GpuDeviceBuffer<std::uint8_t> ws(gaffa_hip_preprocess_workspace_bytes(ns, n, 1));
      GpuDeviceBuffer<std::int32_t> st(ns);
      hip_ok(gaffa_hip_memset(st.data(), 0, st.bytes(), nullptr), "hipMemset");
      a.workspace = ws.data();
      a.workspace_bytes = ws.bytes();
      a.status = st.data();
      hip_ok(gaffa_hip_detrend_normalise_batch(&a, nullptr), "convert_uint32_to_float");
      hip_ok(gaffa_hip_stream_synchronize(nullptr), "hipStreamSynchronize");
    } else {
      GpuPreprocessExecutionOptions pe;
      pe.series_tile_size = std::min<std::size_t>(ns, 1024);
      GpuPreprocessProgram prog(plan, GpuPreprocessProgramOptions{device_id}, pe);
      prog.run(d.data(), ns, n, n, is_u32);
      prog.synchronize();
    }
    copy_to_host(host.data(), d.data(), host.size() * 4, nullptr);
  }
  return owning_array(std::move(host), {static_cast<py::ssize_t>(ns), static_cast<py::ssize_t>(n)});
}

// ------------------------------------------------------------------ plans --
FfaSearchPlan plan_from_arrays(const py::array& tasks, const std::vector<std::size_t>& widths) {
  // structured array with the fields of FfaSearchTask, in order
  const py::buffer_info info = tasks.request();
  if (info.ndim != 1 || info.itemsize != static_cast<py::ssize_t>(sizeof(FfaSearchTask))) {
    throw py::value_error("tasks must be a 1D structured array of FfaSearchTask records");
  }
  require_c_contiguous(info, "tasks");
  FfaSearchPlan plan;
  plan.width_trials = widths;
  plan.tasks.resize(static_cast<std::size_t>(info.shape[0]));
  std::memcpy(plan.tasks.data(), info.ptr, plan.tasks.size() * sizeof(FfaSearchTask));
  return plan;
}

py::array tasks_to_array(const FfaSearchPlan& plan) {
  std::vector<FfaSearchTask> v = plan.tasks;
  py::dtype dt = py::dtype::of<FfaSearchTask>();
  auto* owner = new std::vector<FfaSearchTask>(std::move(v));
  py::capsule cap(owner, [](void* p) { delete static_cast<std::vector<FfaSearchTask>*>(p); });
  return py::array(dt, {static_cast<py::ssize_t>(owner->size())}, {}, owner->data(), cap);
}

py::dict stage_dict(const StageTimes& t) {
  py::dict d;
  d["dedisperse_ms"] = t.dedisperse_ms;
  d["preprocess_ms"] = t.preprocess_ms;
  d["ffa_ms"] = t.ffa_ms;
  d["total_ms"] = t.total_ms;
  if (t.ffa_tiles != 0) {
    static constexpr const char* names[FfaStageMarks::kStages] = {
        "ffa_prepare_ms", "ffa_subtree_ms", "ffa_levels_ms", "ffa_detect_ms"};
    for (int k = 0; k < FfaStageMarks::kStages; ++k) d[names[k]] = t.ffa_stage_ms[k];
    d["ffa_tiles"] = t.ffa_tiles;
  }
  return d;
}

}  // namespace

struct PyDedispersedSpectrum {
  py::array data;
  std::string backend = "external";
  double dm = 0.0, tsamp = 0.0;
  std::size_t nsamples = 0, nchans = 0, chan_begin = 0, chan_end = 0;
};

PyDedispersedSpectrum py_spectrum_array(const py::array& data, const std::vector<double>& freq,
                                        double tsamp, double dm, std::size_t chan_begin,
                                        std::optional<std::size_t> chan_end, int device_id) {
  const py::array c = py::array::ensure(data, py::array::c_style);
  if (!c || c.ndim() != 2)
    throw std::invalid_argument("spectrum input must be a C-contiguous (nsamples, nchans) array");
  const std::size_t esize = static_cast<std::size_t>(c.itemsize());
  const bool ok = py::isinstance<py::array_t<std::uint8_t>>(c) ||
                  py::isinstance<py::array_t<std::uint16_t>>(c) ||
                  py::isinstance<py::array_t<float>>(c);
  if (!ok) throw py::type_error("spectrum input dtype must be uint8, uint16 or float32");
  const SampleShape shape{static_cast<std::size_t>(c.shape(0)), 1,
                          static_cast<std::size_t>(c.shape(1))};
  const std::size_t ce = chan_end.value_or(shape.nchans);
  const SingleDmDedispersionPlan plan{dm, reference_frequency(freq), tsamp, chan_begin, ce};
  const std::size_t n_out = spectrum_output_nsamples(shape.nsamples, freq, plan);
  const std::size_t nc = ce - chan_begin;
  const void* src = c.data();
  const std::size_t nbytes = static_cast<std::size_t>(c.nbytes());
  py::array out(c.dtype(), {static_cast<py::ssize_t>(n_out), static_cast<py::ssize_t>(nc)});
  void* dst = out.mutable_data();
  {
    py::gil_scoped_release release;  // no Python object is touched below
    GpuDeviceScope scope(device_id);
    GpuDeviceBuffer<std::uint8_t> d_in(nbytes);
    copy_to_device(d_in.data(), src, nbytes, nullptr);
    GpuDeviceBuffer<std::uint8_t> d_out(std::max<std::size_t>(1, n_out * nc * esize));
    dedisperse_spectrum_gpu_raw(d_in.data(), esize, shape, freq, plan, d_out.data(), n_out,
                                GpuDedispersionOptions{device_id, 256, 81920, nullptr});
    copy_to_host(dst, d_out.data(), n_out * nc * esize, nullptr);
  }
  PyDedispersedSpectrum r;
  r.data = out;
  r.backend = "gpu";
  r.dm = dm;
  r.tsamp = tsamp;
  r.nsamples = n_out;
  r.nchans = nc;
  r.chan_begin = chan_begin;
  r.chan_end = ce;
  return r;
}

struct PyFoldedProfile {
  py::array profile, exposure, phase;
  std::size_t nbin = 0;
  double period = 0.0, tsamp = 0.0;
};

struct PyFoldResult {
  py::array cube, exposure, profile, freq_phase, time_phase, phase, time;
  std::size_t nsubint = 0, nchans = 0, nbin = 0;
  double period = 0.0, tsamp = 0.0, tsubint = 0.0;
};

std::vector<float> phase_axis(std::size_t nbin) {
  std::vector<float> p(nbin);
  for (std::size_t i = 0; i < nbin; ++i)
    p[i] = static_cast<float>(static_cast<double>(i) / static_cast<double>(nbin));
  return p;
}

FoldDtype fold_dtype(const py::array& a) {
  if (py::isinstance<py::array_t<std::uint8_t>>(a)) return FoldDtype::U8;
  if (py::isinstance<py::array_t<std::uint16_t>>(a)) return FoldDtype::U16;
  if (py::isinstance<py::array_t<std::uint32_t>>(a)) return FoldDtype::U32;
  if (py::isinstance<py::array_t<float>>(a)) return FoldDtype::F32;
  throw py::type_error("fold input dtype must be uint8, uint16, uint32 or float32");
}

PyFoldedProfile py_fold_profile(const py::array& data, double tsamp, double period,
                                std::size_t nbin, std::size_t dm_index, int device_id) {
  const py::array c = py::array::ensure(data, py::array::c_style);
  if (!c || (c.ndim() != 1 && c.ndim() != 2))
    throw std::invalid_argument("fold input must be a C-contiguous 1D or 2D array");
  const FoldDtype dt = fold_dtype(c);
  const std::size_t rows = c.ndim() == 2 ? static_cast<std::size_t>(c.shape(0)) : 1;
  const std::size_t n = static_cast<std::size_t>(c.shape(c.ndim() - 1));
  if (dm_index >= rows) throw std::out_of_range("dm_index out of range");
  if (n == 0) throw std::invalid_argument("fold input must not be empty");
  const std::size_t esize = static_cast<std::size_t>(c.itemsize());
  const std::uint8_t* src = static_cast<const std::uint8_t*>(c.data()) + dm_index * n * esize;
  FoldedProfile r;
  {
    py::gil_scoped_release release;  // no Python object is touched below
    GpuDeviceScope scope(device_id);
    GpuDeviceBuffer<std::uint8_t> d(n * esize);
    copy_to_device(d.data(), src, n * esize, nullptr);
    r = fold_time_series_gpu(GpuFoldInput{d.data(), dt, n, 1, 1, device_id},
                             FoldOptions{period, tsamp, nbin, 0});
  }
  PyFoldedProfile p;
  p.profile = owning_array(std::move(r.profile), {static_cast<py::ssize_t>(nbin)});
  p.exposure = owning_array(std::move(r.exposure), {static_cast<py::ssize_t>(nbin)});
  p.phase = owning_array(phase_axis(nbin), {static_cast<py::ssize_t>(nbin)});
  p.nbin = nbin;
  p.period = period;
  p.tsamp = tsamp;
  return p;
}

PyFoldResult py_fold_spectrum(const py::array& data, double tsamp, double period,
                              std::size_t nbin, double tsubint, std::size_t output_channels,
                              int device_id) {
  const py::array c = py::array::ensure(data, py::array::c_style);
  if (!c || c.ndim() != 2)
    throw std::invalid_argument("fold spectrum must be a C-contiguous (nsamples, nchans) array");
  const FoldDtype dt = fold_dtype(c);
  if (dt == FoldDtype::U32) throw py::type_error("fold spectrum dtype must be uint8, uint16 or float32");
  const std::size_t ns = static_cast<std::size_t>(c.shape(0));
  const std::size_t nc = static_cast<std::size_t>(c.shape(1));
  const std::size_t nbytes = static_cast<std::size_t>(c.nbytes());
  const void* src = c.data();
  FoldDedispersedSpectrumResult r;
  {
    py::gil_scoped_release release;  // no Python object is touched below
    GpuDeviceScope scope(device_id);
    GpuDeviceBuffer<std::uint8_t> d(nbytes);
    copy_to_device(d.data(), src, nbytes, nullptr);
    r = fold_dedispersed_spectrum_gpu(GpuFoldInput{d.data(), dt, ns, nc, nc, device_id}, tsamp,
                                      FoldDedispersedSpectrumOptions{period, nbin, tsubint,
                                                                     output_channels});
  }
  PyFoldResult p;
  const auto S = static_cast<py::ssize_t>(r.nsubint), C = static_cast<py::ssize_t>(r.nchans),
             B = static_cast<py::ssize_t>(r.nbin);
  p.cube = owning_array(std::move(r.cube.data), {S, C, B});
  p.exposure = owning_array(std::move(r.cube.exposure), {S, B});
  p.profile = owning_array(std::move(r.profile), {B});
  p.freq_phase = owning_array(std::move(r.freq_phase), {C, B});
  p.time_phase = owning_array(std::move(r.time_phase), {S, B});
  p.phase = owning_array(phase_axis(r.nbin), {B});
  std::vector<double> t(r.nsubint);
  for (std::size_t i = 0; i < r.nsubint; ++i) t[i] = (static_cast<double>(i) + 0.5) * r.tsubint;
  p.time = owning_array(std::move(t), {S});
  p.nsubint = r.nsubint;
  p.nchans = r.nchans;
  p.nbin = r.nbin;
  p.period = r.period;
  p.tsamp = r.tsamp;
  p.tsubint = r.tsubint;
  return p;
}

namespace {

// Scorer inputs: float32 C-contiguous cubes (n, 32, 64), patches (n, 128)
// and, when given, labels (n,).
using FloatArray = py::array_t<float, py::array::c_style | py::array::forcecast>;

std::size_t scorer_rows(const FloatArray& cubes, const FloatArray& patches,
                        const FloatArray* labels) {
  if (cubes.ndim() != 3 || cubes.shape(1) != GAFFA_HIP_SCORER_SUBINTS ||
      cubes.shape(2) != GAFFA_HIP_SCORER_PHASE_BINS)
    throw py::value_error("cubes must have shape (n, 32, 64)");
  if (patches.ndim() != 2 || patches.shape(1) != GAFFA_HIP_SCORER_PATCH)
    throw py::value_error("patches must have shape (n, 128)");
  const auto n = static_cast<std::size_t>(cubes.shape(0));
  if (n == 0 || static_cast<std::size_t>(patches.shape(0)) != n)
    throw py::value_error("cubes and patches must hold the same non-zero number of candidates");
  if (labels != nullptr && (labels->ndim() != 1 || static_cast<std::size_t>(labels->shape(0)) != n))
    throw py::value_error("labels must have shape (n,)");
  return n;
}

}  // namespace

PYBIND11_MODULE(_core, m) {
  m.doc() = "gaffa search hot path on AMD Instinct MI355X (HIP, gfx950)";
  PYBIND11_NUMPY_DTYPE(FfaSearchTask, downsample_factor, effective_tsamp, input_nsamples,
                       prepared_nsamples, bins, rows, rows_eval, period_begin, period_end);

  m.attr("DM_PEAK_RECORD_BYTES") = sizeof(DmPeak);
  m.def("gpu_device_count", &gpu_device_count);
  m.def("gpu_runtime_version", &gpu_runtime_version);
  m.def("cuda_device_count", &gpu_device_count, "Alias of gpu_device_count for gaffa scripts.");

  py::register_exception_translator([](std::exception_ptr p) {
    try {
      if (p) std::rethrow_exception(p);
    } catch (const std::out_of_range& e) {
      PyErr_SetString(PyExc_IndexError, e.what());
    } catch (const std::invalid_argument& e) {
      PyErr_SetString(PyExc_ValueError, e.what());
    } catch (const std::overflow_error& e) {
      PyErr_SetString(PyExc_OverflowError, e.what());
    }
  });

  // dedispersion
  py::class_<PyDedispersedResult>(m, "DedispersedResult")
      .def_readonly("data", &PyDedispersedResult::data)
      .def_readonly("backend", &PyDedispersedResult::backend)
      .def_readonly("dm_low", &PyDedispersedResult::dm_low)
      .def_readonly("dm_step", &PyDedispersedResult::dm_step)
      .def_readonly("ndm", &PyDedispersedResult::ndm)
      .def_readonly("nsamples", &PyDedispersedResult::nsamples)
      .def_readonly("tsamp", &PyDedispersedResult::tsamp)
      .def_property_readonly("shape", [](const PyDedispersedResult& r) {
        return py::make_tuple(r.ndm, r.nsamples);
      })
      .def_property_readonly("dtype", [](const PyDedispersedResult& r) { return r.data.attr("dtype"); })
      .def_property_readonly("nbytes", [](const PyDedispersedResult& r) { return r.data.attr("nbytes"); })
      .def("__repr__", [](const PyDedispersedResult& r) {
        return "<DedispersedResult shape=(" + std::to_string(r.ndm) + ", " +
               std::to_string(r.nsamples) + ") backend=" + r.backend + ">";
      });
  // gaffa's defaults (dedispersion_bindings.cpp:648-672): "cpu" for the single,
  // multi and spectrum forms, which gaffa_amd refuses, "cuda" for subband
  m.def("dedisperse_single_dm", &py_single, py::arg("filterbank"), py::kw_only(), py::arg("dm"),
        py::arg("backend") = "cpu", py::arg("device_id") = 0, py::arg("threads_per_block") = 256,
        py::arg("time_tile_samples") = 81920);
  m.def("dedisperse_multi_dm", &py_multi, py::arg("filterbank"), py::kw_only(), py::arg("dm_low"),
        py::arg("dm_step"), py::arg("ndm"), py::arg("backend") = "cpu", py::arg("device_id") = 0,
        py::arg("threads_per_block") = 256, py::arg("time_tile_samples") = 81920);
  m.def("dedisperse_subband", &py_subband, py::arg("filterbank"), py::kw_only(), py::arg("dm_low"),
        py::arg("dm_step"), py::arg("ndm"), py::arg("backend") = "cuda",
        py::arg("subband_channels") = 32, py::arg("ndm_per_nominal") = 32,
        py::arg("device_id") = 0, py::arg("threads_per_block") = 256,
        py::arg("time_tile_samples") = 81920);
  m.def("dedispersion_delay_bins", &dedispersion_delay_bins, py::arg("dm"),
        py::arg("frequency_mhz"), py::arg("ref_frequency_mhz"), py::arg("tsamp"));

  // FFA
  py::class_<FfaSearchPlan>(m, "FfaPlan")
      .def(py::init(&plan_from_arrays), py::arg("tasks"), py::arg("width_trials"))
      .def_property_readonly("task_count", [](const FfaSearchPlan& p) { return p.tasks.size(); })
      .def_property_readonly("width_trials",
                             [](const FfaSearchPlan& p) { return py::tuple(py::cast(p.width_trials)); })
      .def_property_readonly("tasks", &tasks_to_array)
      .def("__repr__", [](const FfaSearchPlan& p) {
        return "<FfaPlan task_count=" + std::to_string(p.tasks.size()) +
               " width_trials=" + std::to_string(p.width_trials.size()) + ">";
      });
  py::class_<FfaPeak>(m, "FfaPeak")
      .def_readonly("period", &FfaPeak::period)
      .def_readonly("frequency", &FfaPeak::frequency)
      .def_readonly("snr", &FfaPeak::snr)
      .def_readonly("width", &FfaPeak::width)
      .def_readonly("duty_cycle", &FfaPeak::duty_cycle)
      .def_readonly("phase", &FfaPeak::phase)
      .def_readonly("shift", &FfaPeak::shift)
      .def_readonly("bins", &FfaPeak::bins)
      .def_readonly("width_index", &FfaPeak::width_index)
      .def_readonly("period_index", &FfaPeak::period_index)
      .def("__repr__", [](const FfaPeak& p) {
        return "<FfaPeak period=" + std::to_string(p.period) + " snr=" + std::to_string(p.snr) + ">";
      });
  m.def(
      "_make_riptide_ffa_plan",
      [](std::size_t nsamples, double tsamp, double period_min, double period_max,
         std::size_t bins_min, std::size_t bins_max, std::size_t min_periods, double duty,
         double spacing, std::size_t max_tasks) {
        return make_riptide_ffa_plan(nsamples, tsamp,
                                     RiptideFfaPlanOptions{period_min, period_max, bins_min,
                                                           bins_max, min_periods, duty, spacing,
                                                           max_tasks});
      },
      py::kw_only(), py::arg("nsamples"), py::arg("tsamp"), py::arg("period_min"),
      py::arg("period_max"), py::arg("bins_min") = 180, py::arg("bins_max") = 256,
      py::arg("min_periods") = 1, py::arg("duty_cycle_max") = 0.20,
      py::arg("width_trial_spacing") = 1.5, py::arg("max_tasks") = 1'000'000);
  m.def("_check_ffa_search", &py_check_ffa_search, py::arg("time_series"), py::arg("plan"),
        py::arg("max_peaks") = py::none());
  m.def("_ffa_search_gpu_host", &py_ffa_search, py::arg("time_series"), py::arg("plan"),
        py::kw_only(), py::arg("device_id") = 0, py::arg("snr_threshold") = 6.0F,
        py::arg("max_peaks") = py::none());
  m.def("_ffa_search_batch_gpu", &py_ffa_search_batch, py::arg("batch"), py::arg("plan"),
        py::kw_only(), py::arg("device_id") = 0, py::arg("snr_threshold") = 6.0F,
        py::arg("max_peaks") = py::none(), py::arg("series_tile") = 0);
  m.def("_ffa_transform_rows_gpu", &py_ffa_transform_rows, py::arg("batch"), py::arg("plan"),
        py::arg("task"), py::kw_only(), py::arg("device_id") = 0);
  m.def("ffa_task_stdnoise", &ffa_task_stdnoise);
  m.def(
      "_ffa_level_layout",
      [](const FfaSearchPlan& plan, std::uint32_t level_rows) {
        // the host-built level tables of a plan (no device needed): items
        // [n, 4], spans [n, 20] and leaf rows [m, 16] as uint32, in launch order
        const GpuFfaExecutionPlan e(plan, FfaItemShape{level_rows, 0});
        auto words = [](const auto& v, std::size_t per) {
          std::vector<std::uint32_t> w(v.size() * per);
          if (!v.empty()) std::memcpy(w.data(), v.data(), w.size() * sizeof(std::uint32_t));
          return owning_array(std::move(w), {static_cast<py::ssize_t>(v.size()),
                                             static_cast<py::ssize_t>(per)});
        };
        static_assert(sizeof(gaffa_hip_ffa_item) == 16 && sizeof(gaffa_hip_ffa_level_span) == 80 &&
                      sizeof(gaffa_hip_ffa_level_row) == 64);
        py::dict d;
        d["items"] = words(e.level_items(), 4);
        d["spans"] = words(e.level_spans(), 20);
        d["leaf_rows"] = words(e.level_leaf_rows(), 16);
        d["level_lds_bytes"] = e.level_lds_bytes();
        return d;
      },
      py::arg("plan"), py::arg("level_rows") = 0,
      "Host-built level work items, their leaf-run spans and leaf rows (tests)");

  // SIGPROC files (native reader / writer)
  m.def(
      "_read_filterbank",
      [](const std::string& path, const std::string& channel_order,
         const std::string& reverse_backend, std::size_t io_buffer_bytes,
         std::size_t openmp_min_rows, bool header_only) {
        FilterbankReadOptions o;
        o.io_buffer_bytes = io_buffer_bytes;
        o.openmp_min_rows = openmp_min_rows;
        o.channel_order = parse_channel_order(channel_order);
        o.reverse_backend = parse_reverse_backend(reverse_backend);
        FilterbankData d;
        {
          py::gil_scoped_release release;
          if (header_only)
            d.header = read_filterbank_header(path, o.channel_order);
          else
            d = read_filterbank(path, o);
        }
        const FilterbankHeader& h = d.header;
        py::dict hd;
        hd["header_size"] = h.header_size;
        hd["nsamples"] = h.nsamples;
        hd["telescope_id"] = h.telescope_id;
        hd["machine_id"] = h.machine_id;
        hd["data_type"] = h.data_type;
        hd["barycentric"] = h.barycentric;
        hd["pulsarcentric"] = h.pulsarcentric;
        hd["ibeam"] = h.ibeam;
        hd["nbeams"] = h.nbeams;
        hd["npuls"] = h.npuls;
        hd["nbins"] = h.nbins;
        hd["nbits"] = h.nbits;
        hd["nifs"] = h.nifs;
        hd["nchans"] = h.nchans;
        hd["az_start"] = h.az_start;
        hd["za_start"] = h.za_start;
        hd["src_raj"] = h.src_raj;
        hd["src_dej"] = h.src_dej;
        hd["tstart"] = h.tstart;
        hd["tsamp"] = h.tsamp;
        hd["fch1"] = h.fch1;
        hd["foff"] = h.foff;
        hd["refdm"] = h.refdm;
        hd["period"] = h.period;
        hd["rawdatafile"] = h.rawdatafile;
        hd["source_name"] = h.source_name;
        hd["frequency_table"] = h.frequency_table;
        hd["uses_frequency_table"] = h.uses_frequency_table;
        hd["reversed"] = h.reversed;
        if (header_only) return py::tuple(py::make_tuple(hd, py::none()));
        const py::dtype dt = h.nbits == 8    ? py::dtype::of<std::uint8_t>()
                             : h.nbits == 16 ? py::dtype::of<std::uint16_t>()
                                             : py::dtype::of<float>();
        const std::size_t esize = static_cast<std::size_t>(h.nbits / 8);
        auto* owner = new std::vector<std::uint8_t>(std::move(d.bytes));
        py::capsule cap(owner, [](void* p) { delete static_cast<std::vector<std::uint8_t>*>(p); });
        const auto ns = static_cast<py::ssize_t>(h.nsamples);
        const auto ni = static_cast<py::ssize_t>(h.nifs);
        const auto nc = static_cast<py::ssize_t>(h.nchans);
        const auto es = static_cast<py::ssize_t>(esize);
        py::array a(dt, {ns, ni, nc}, {ni * nc * es, nc * es, es}, owner->data(), cap);
        return py::tuple(py::make_tuple(hd, a));
      },
      py::arg("path"), py::kw_only(), py::arg("channel_order") = "frequency_ascending",
      py::arg("reverse_backend") = "auto",
      py::arg("io_buffer_bytes") = default_filterbank_io_buffer_bytes,
      py::arg("openmp_min_rows") = 4096, py::arg("header_only") = false);
  m.def(
      "_write_filterbank_u8",
      [](const std::string& path, const py::array_t<std::uint8_t, py::array::c_style>& samples,
         double tsamp, double fch1, double foff, const std::string& source_name, double tstart) {
        if (samples.ndim() != 2) throw std::invalid_argument("samples must be (nsamples, nchans)");
        FilterbankHeader h;
        h.nchans = static_cast<int>(samples.shape(1));
        h.nbits = 8;
        h.tsamp = tsamp;
        h.fch1 = fch1;
        h.foff = foff;
        h.source_name = source_name;
        h.tstart = tstart;
        write_filterbank_u8(path, h, samples.data(), static_cast<std::size_t>(samples.shape(0)));
      },
      py::arg("path"), py::arg("samples"), py::kw_only(), py::arg("tsamp"), py::arg("fch1"),
      py::arg("foff"), py::arg("source_name") = "", py::arg("tstart") = 0.0);

  // aligned dynamic spectrum
  py::class_<PyDedispersedSpectrum>(m, "DedispersedSpectrum")
      .def(py::init([](const py::array& data, double tsamp, double dm, std::size_t chan_begin,
                       std::optional<std::size_t> chan_end, std::string backend) {
             const py::array c = py::array::ensure(data, py::array::c_style);
             if (!c || c.ndim() != 2)
               throw std::invalid_argument("spectrum data must be a C-contiguous 2D array");
             if (!(tsamp > 0.0) || !std::isfinite(tsamp))
               throw std::invalid_argument("spectrum tsamp must be finite and positive");
             if (!(dm >= 0.0)) throw std::invalid_argument("spectrum dm must be non-negative");
             PyDedispersedSpectrum r;
             r.data = c;
             r.backend = std::move(backend);
             r.dm = dm;
             r.tsamp = tsamp;
             r.nsamples = static_cast<std::size_t>(c.shape(0));
             r.nchans = static_cast<std::size_t>(c.shape(1));
             r.chan_begin = chan_begin;
             r.chan_end = chan_end.value_or(chan_begin + r.nchans);
             if (r.chan_end - r.chan_begin != r.nchans)
               throw std::invalid_argument("spectrum channel range must match nchans");
             return r;
           }),
           py::arg("data"), py::kw_only(), py::arg("tsamp"), py::arg("dm") = 0.0,
           py::arg("chan_begin") = 0, py::arg("chan_end") = py::none(),
           py::arg("backend") = "external")
      .def_readonly("data", &PyDedispersedSpectrum::data)
      .def_readonly("backend", &PyDedispersedSpectrum::backend)
      .def_readonly("dm", &PyDedispersedSpectrum::dm)
      .def_readonly("tsamp", &PyDedispersedSpectrum::tsamp)
      .def_readonly("nsamples", &PyDedispersedSpectrum::nsamples)
      .def_readonly("nchans", &PyDedispersedSpectrum::nchans)
      .def_readonly("chan_begin", &PyDedispersedSpectrum::chan_begin)
      .def_readonly("chan_end", &PyDedispersedSpectrum::chan_end)
      .def_property_readonly("shape", [](const PyDedispersedSpectrum& r) {
        return py::make_tuple(r.nsamples, r.nchans);
      });
  m.def("_dedisperse_spectrum_array", &py_spectrum_array, py::arg("data"), py::kw_only(),
        py::arg("frequency_mhz"), py::arg("tsamp"), py::arg("dm"), py::arg("chan_begin") = 0,
        py::arg("chan_end") = py::none(), py::arg("device_id") = 0);
  m.def(
      "dedisperse_spectrum",
      [](const py::object& fb, double dm, const std::string& backend, int device_id, std::size_t,
         std::size_t) {
        std::string name = parse_backend(backend, "dedispersion");
        FilterbankInput in = filterbank_input(fb);
        const py::array flat = in.keep.reshape({static_cast<py::ssize_t>(in.samples.shape.nsamples),
                                                static_cast<py::ssize_t>(in.samples.shape.nchans)});
        PyDedispersedSpectrum r =
            py_spectrum_array(flat, in.freq, in.tsamp, dm, 0, std::nullopt, device_id);
        r.backend = std::move(name);
        return r;
      },
      py::arg("filterbank"), py::kw_only(), py::arg("dm"), py::arg("backend") = "cpu",
      py::arg("device_id") = 0, py::arg("threads_per_block") = 256,
      py::arg("time_tile_samples") = 81920,
      "Aligned single-DM dynamic spectrum (valid samples only), on the GPU");

  // folding
  py::class_<PyFoldedProfile>(m, "FoldedProfile")
      .def_readonly("profile", &PyFoldedProfile::profile)
      .def_readonly("exposure", &PyFoldedProfile::exposure)
      .def_readonly("phase", &PyFoldedProfile::phase)
      .def_readonly("nbin", &PyFoldedProfile::nbin)
      .def_readonly("period", &PyFoldedProfile::period)
      .def_readonly("tsamp", &PyFoldedProfile::tsamp);
  py::class_<PyFoldResult>(m, "FoldResult")
      .def_readonly("cube", &PyFoldResult::cube)
      .def_readonly("exposure", &PyFoldResult::exposure)
      .def_readonly("profile", &PyFoldResult::profile)
      .def_readonly("freq_phase", &PyFoldResult::freq_phase)
      .def_readonly("time_phase", &PyFoldResult::time_phase)
      .def_readonly("phase", &PyFoldResult::phase)
      .def_readonly("time", &PyFoldResult::time)
      .def_readonly("nsubint", &PyFoldResult::nsubint)
      .def_readonly("nchans", &PyFoldResult::nchans)
      .def_readonly("nbin", &PyFoldResult::nbin)
      .def_readonly("period", &PyFoldResult::period)
      .def_readonly("tsamp", &PyFoldResult::tsamp)
      .def_readonly("tsubint", &PyFoldResult::tsubint);
  m.def("_fold_dedispersed_profile", &py_fold_profile, py::arg("data"), py::kw_only(),
        py::arg("tsamp"), py::arg("period"), py::arg("nbin"), py::arg("dm_index") = 0,
        py::arg("device_id") = 0);
  m.def("_fold_dedispersed_spectrum", &py_fold_spectrum, py::arg("data"), py::kw_only(),
        py::arg("tsamp"), py::arg("period"), py::arg("nbin"), py::arg("tsubint"),
        py::arg("output_channels"), py::arg("device_id") = 0);

  // candidate selection and harmonic flagging (host code over the peak list)
  py::class_<Candidate>(m, "Candidate")
      .def_readonly("dm", &Candidate::dm)
      .def_readonly("dm_index", &Candidate::dm_index)
      .def_readonly("period", &Candidate::period)
      .def_readonly("frequency", &Candidate::frequency)
      .def_readonly("width", &Candidate::width)
      .def_readonly("duty_cycle", &Candidate::duty_cycle)
      .def_readonly("snr", &Candidate::snr)
      .def_readonly("peak_count", &Candidate::peak_count)
      .def_readonly("dm_index_min", &Candidate::dm_index_min)
      .def_readonly("dm_index_max", &Candidate::dm_index_max)
      .def_readonly("frequency_min", &Candidate::frequency_min)
      .def_readonly("frequency_max", &Candidate::frequency_max)
      .def_readonly("best_peak", &Candidate::best_peak)
      .def("__repr__", [](const Candidate& c) {
        return "<Candidate dm=" + std::to_string(c.dm) + " period=" + std::to_string(c.period) +
               " snr=" + std::to_string(c.snr) + " peaks=" + std::to_string(c.peak_count) + ">";
      });
  py::class_<HarmonicRelation>(m, "HarmonicRelation")
      .def_readonly("is_harmonic", &HarmonicRelation::is_harmonic)
      .def_readonly("parent_index", &HarmonicRelation::parent_index)
      .def_readonly("numerator", &HarmonicRelation::numerator)
      .def_readonly("denominator", &HarmonicRelation::denominator)
      .def_readonly("frequency_error_bins", &HarmonicRelation::frequency_error_bins)
      .def_readonly("phase_distance", &HarmonicRelation::phase_distance)
      .def_readonly("dm_distance", &HarmonicRelation::dm_distance)
      .def_readonly("expected_snr", &HarmonicRelation::expected_snr)
      .def_readonly("snr_distance", &HarmonicRelation::snr_distance);
  py::class_<HarmonicCandidate>(m, "HarmonicCandidate")
      .def_readonly("candidate", &HarmonicCandidate::candidate)
      .def_readonly("harmonic", &HarmonicCandidate::harmonic);
  m.def(
      "_select_candidates_ranked_records",
      [](const py::array& peaks, const py::array_t<std::uint32_t, py::array::c_style>& ranks,
         double obs, double radius, std::size_t dm_radius, bool across,
         std::size_t max_candidates) {
        if (peaks.itemsize() != static_cast<py::ssize_t>(sizeof(DmPeak)))
          throw std::invalid_argument("peak records must be gaffa DmPeak records");
        const py::array c = py::array::ensure(peaks, py::array::c_style);
        std::vector<DmPeak> v(static_cast<std::size_t>(c.size()));
        if (!v.empty()) std::memcpy(v.data(), c.data(), v.size() * sizeof(DmPeak));
        std::vector<std::uint32_t> r(ranks.data(), ranks.data() + ranks.size());
        std::vector<Candidate> out;
        {
          py::gil_scoped_release release;
          out = select_candidates_ranked_cpu(
              v, r, obs, CandidateSelectionOptions{radius, dm_radius, across, max_candidates});
        }
        return py::bytes(reinterpret_cast<const char*>(out.data()),
                         out.size() * sizeof(Candidate));
      },
      py::arg("peaks"), py::arg("period_rank"), py::arg("observation_seconds"),
      py::arg("frequency_cluster_radius"), py::arg("dm_cluster_radius"),
      py::arg("cluster_across_widths"), py::arg("max_candidates"));
  m.def(
      "_select_candidates_records",
      [](const py::array& peaks, double obs, double radius, std::size_t dm_radius, bool across,
         std::size_t max_candidates) {
        if (peaks.itemsize() != static_cast<py::ssize_t>(sizeof(DmPeak)))
          throw std::invalid_argument("peak records must be gaffa DmPeak records");
        const py::array c = py::array::ensure(peaks, py::array::c_style);
        std::vector<DmPeak> v(static_cast<std::size_t>(c.size()));
        if (!v.empty()) std::memcpy(v.data(), c.data(), v.size() * sizeof(DmPeak));
        std::vector<Candidate> r;
        {
          py::gil_scoped_release release;
          r = select_candidates_cpu(v, obs,
                                    CandidateSelectionOptions{radius, dm_radius, across,
                                                              max_candidates});
        }
        return py::bytes(reinterpret_cast<const char*>(r.data()), r.size() * sizeof(Candidate));
      },
      py::arg("peaks"), py::arg("observation_seconds"), py::arg("frequency_cluster_radius"),
      py::arg("dm_cluster_radius"), py::arg("cluster_across_widths"), py::arg("max_candidates"));
  m.def(
      "_flag_harmonics_records",
      [](const py::array& cands, double obs, double flo, double fhi, std::size_t max_harmonic,
         std::size_t denominator_max, double ftol, double phase_max, double dm_max,
         bool use_snr, double snr_max) {
        if (cands.itemsize() != static_cast<py::ssize_t>(sizeof(Candidate)))
          throw std::invalid_argument("candidate records must be gaffa Candidate records");
        const py::array c = py::array::ensure(cands, py::array::c_style);
        std::vector<Candidate> v(static_cast<std::size_t>(c.size()));
        if (!v.empty()) std::memcpy(v.data(), c.data(), v.size() * sizeof(Candidate));
        const auto r = flag_harmonics_cpu(
            v, HarmonicContext{obs, flo, fhi},
            HarmonicOptions{max_harmonic, denominator_max, ftol, phase_max, dm_max, use_snr,
                            snr_max});
        std::vector<Candidate> out;
        py::list rel;
        for (const auto& h : r) {
          out.push_back(h.candidate);
          rel.append(h.harmonic);
        }
        return py::make_tuple(
            py::bytes(reinterpret_cast<const char*>(out.data()), out.size() * sizeof(Candidate)),
            rel);
      });
  m.def(
      "select_candidates",
      [](const std::vector<DmPeak>& peaks, double obs, double radius, std::size_t dm_radius,
         bool across, std::size_t max_candidates) {
        py::gil_scoped_release release;
        return select_candidates_cpu(
            peaks, obs, CandidateSelectionOptions{radius, dm_radius, across, max_candidates});
      },
      py::arg("peaks"), py::arg("observation_seconds"), py::kw_only(),
      py::arg("frequency_cluster_radius") = 0.1, py::arg("dm_cluster_radius") = 1,
      py::arg("cluster_across_widths") = true, py::arg("max_candidates") = 0);
  m.def(
      "flag_harmonics",
      [](const std::vector<Candidate>& c, double obs, double flo, double fhi,
         std::size_t max_harmonic, std::size_t denominator_max, double ftol, double phase_max,
         double dm_max, bool use_snr, double snr_max) {
        return flag_harmonics_cpu(c, HarmonicContext{obs, flo, fhi},
                                  HarmonicOptions{max_harmonic, denominator_max, ftol, phase_max,
                                                  dm_max, use_snr, snr_max});
      },
      py::arg("candidates"), py::arg("observation_seconds"), py::arg("frequency_low_mhz"),
      py::arg("frequency_high_mhz"), py::kw_only(), py::arg("max_harmonic") = 16,
      py::arg("denominator_max") = 5, py::arg("frequency_tolerance_bins") = 1.5,
      py::arg("phase_distance_max") = 1.0, py::arg("dm_distance_max") = 3.0,
      py::arg("use_snr_consistency") = false, py::arg("snr_distance_max") = 3.0);
  m.def(
      "remove_harmonics",
      [](const std::vector<HarmonicCandidate>& c, std::size_t max_candidates) {
        return remove_harmonics_cpu(c, max_candidates);
      },
      py::arg("candidates"), py::arg("max_candidates") = 0);

  // batched FFA over series with their own plans
  py::class_<GpuFfaBatchedSearch>(m, "FfaBatchedSearch")
      .def(py::init([](std::vector<FfaSearchPlan> plans, std::size_t series_per_tile,
                       std::size_t workspace_bytes, bool use_graph, int device_id) {
             GpuFfaBatchedOptions o;
             o.series_per_tile = series_per_tile;
             o.workspace_bytes_limit = workspace_bytes;
             o.use_graph = use_graph;
             o.device_id = device_id;
             py::gil_scoped_release release;
             return std::make_unique<GpuFfaBatchedSearch>(std::move(plans), o);
           }),
           py::arg("plans"), py::kw_only(), py::arg("series_per_tile") = 16,
           py::arg("workspace_bytes") = std::size_t{32} << 30, py::arg("use_graph") = true,
           py::arg("device_id") = 0)
      .def_property_readonly("nseries", &GpuFfaBatchedSearch::nseries)
      .def_property_readonly("row_length", &GpuFfaBatchedSearch::row_length)
      .def_property_readonly("series_per_tile", &GpuFfaBatchedSearch::series_per_tile)
      .def_property_readonly("offsets", &GpuFfaBatchedSearch::offsets)
      .def_property_readonly("lengths", &GpuFfaBatchedSearch::lengths)
      .def("upload",
           [](GpuFfaBatchedSearch& b, const py::array_t<float, py::array::c_style>& row) {
             if (static_cast<std::size_t>(row.size()) != b.row_length())
               throw std::invalid_argument("batch row length does not match");
             const float* src = row.data();
             py::gil_scoped_release release;
             b.upload(src, b.row_length());
           })
      .def("download_row",
           [](GpuFfaBatchedSearch& b) {
             std::vector<float> host(b.row_length());
             copy_to_host(host.data(), b.row(), host.size() * sizeof(float), nullptr);
             return owning_array(std::move(host), {static_cast<py::ssize_t>(b.row_length())});
           })
      .def(
          "generate",
          [](GpuFfaBatchedSearch& b, std::uint64_t seed, double tsamp,
             const std::vector<double>& period, const std::vector<double>& phase0,
             const std::vector<float>& sigma, const std::vector<float>& amplitude) {
            const std::size_t n = b.nseries();
            if (period.size() != n || phase0.size() != n || sigma.size() != n ||
                amplitude.size() != n)
              throw std::invalid_argument("one injection entry per series is required");
            py::gil_scoped_release release;
            GpuDeviceBuffer<std::uint64_t> off = upload(b.offsets().data(), n);
            GpuDeviceBuffer<std::uint64_t> len = upload(b.lengths().data(), n);
            GpuDeviceBuffer<double> per = upload(period.data(), n);
            GpuDeviceBuffer<double> ph = upload(phase0.data(), n);
            GpuDeviceBuffer<float> sg = upload(sigma.data(), n);
            GpuDeviceBuffer<float> amp = upload(amplitude.data(), n);
            gaffa_hip_synth_series_args a{};
            a.out = b.row();
            a.row_length = b.row_length();
            a.nseries = static_cast<std::uint32_t>(n);
            a.key = seed;
            a.tsamp = tsamp;
            a.offsets = off.data();
            a.lengths = len.data();
            a.period = per.data();
            a.phase0 = ph.data();
            a.sigma = sg.data();
            a.amplitude = amp.data();
            hip_ok(gaffa_hip_synth_series(&a, b.stream()), "synth_series");
            hip_ok(gaffa_hip_stream_synchronize(b.stream()), "hipStreamSynchronize");
          },
          py::arg("seed"), py::arg("tsamp"), py::arg("period"), py::arg("phase0"),
          py::arg("sigma"), py::arg("amplitude"),
          "Seeded N(0,1) series (+ Gaussian pulse trains where amplitude != 0) in the row")
      .def("set_series_per_tile", &GpuFfaBatchedSearch::set_series_per_tile)
      .def("autotune",
           [](GpuFfaBatchedSearch& b, const std::vector<std::size_t>& candidates,
              const std::vector<std::uint32_t>& level_rows,
              const std::vector<std::uint32_t>& terminal_rows) {
             py::gil_scoped_release release;
             return b.autotune(candidates, level_rows, terminal_rows);
           },
           py::arg("candidates"), py::arg("level_rows") = std::vector<std::uint32_t>{8, 16, 32},
           py::arg("terminal_rows") = std::vector<std::uint32_t>{16, 32},
           "Tune series per tile, then rows per level item, then rows per terminal item; "
           "cached per batch shape")
      .def_property_readonly("tuning",
                             [](const GpuFfaBatchedSearch& b) {
                               const FfaBatchTuning& t = b.tuning();
                               py::dict d;
                               d["series_per_tile"] = t.series_per_tile;
                               d["level_rows"] = t.item_shape.level_rows;
                               d["terminal_rows"] = t.item_shape.terminal_rows;
                               d["gpu_ms"] = t.gpu_ms;
                               d["cached"] = t.cached;
                               return d;
                             })
      .def(
          "set_item_shape",
          [](GpuFfaBatchedSearch& b, std::uint32_t level_rows, std::uint32_t terminal_rows) {
            py::gil_scoped_release release;
            b.set_item_shape(FfaItemShape{level_rows, terminal_rows});
          },
          py::arg("level_rows") = 0, py::arg("terminal_rows") = 0,
          "Rows per level / terminal work item (0 = the kernels' defaults)")
      .def(
          "run",
          [](GpuFfaBatchedSearch& b, float snr_threshold, bool per_series) {
            FfaBatchTiming t;
            FfaBatchSearchResult r;
            {
              py::gil_scoped_release release;
              r = per_series ? b.run_per_series(snr_threshold, &t) : b.run(snr_threshold, &t);
            }
            py::list peaks;
            for (const auto& p : r.peaks) peaks.append(py::make_tuple(p.series_index, p.peak));
            py::dict timing;
            timing["gpu_ms"] = t.gpu_ms;
            timing["tiles"] = t.tiles;
            timing["launches"] = t.launches;
            return py::make_tuple(peaks, timing);
          },
          py::arg("snr_threshold") = 6.0F, py::kw_only(), py::arg("per_series") = false,
          "Search every series; returns ([(series, FfaPeak)] in (series, peak) order, timing)");

  // preprocessing
  m.def("running_median_window_samples", [](double seconds, double tsamp) {
    return running_median_window_seconds(seconds, tsamp).window_samples;
  });
  m.def("_preprocess_batch_gpu", &py_preprocess_batch, py::arg("batch"), py::kw_only(),
        py::arg("window_samples") = py::none(), py::arg("min_points") = 101,
        py::arg("normalise") = true, py::arg("reject_constant") = true, py::arg("device_id") = 0);

  // DM search pipeline
  py::class_<DmPeak>(m, "DmPeak")
      .def_readonly("dm", &DmPeak::dm)
      .def_readonly("dm_index", &DmPeak::dm_index)
      .def_readonly("peak", &DmPeak::peak)
      .def("__repr__", [](const DmPeak& p) {
        return "<DmPeak dm=" + std::to_string(p.dm) + " period=" + std::to_string(p.peak.period) +
               " snr=" + std::to_string(p.peak.snr) + ">";
      });
  m.def(
      "_expand_sorted_dm_peaks",
      [](const FfaSearchPlan& plan, const py::array_t<std::uint32_t, py::array::c_style>& series,
         const py::array_t<std::uint32_t, py::array::c_style>& task,
         const py::array_t<std::uint32_t, py::array::c_style>& shift,
         const py::array_t<std::uint32_t, py::array::c_style>& phase,
         const py::array_t<std::uint32_t, py::array::c_style>& width_index,
         const py::array_t<float, py::array::c_style>& snr, const std::vector<double>& dms,
         bool ranked, std::size_t parts) {
        const std::size_t n = static_cast<std::size_t>(series.size());
        if (task.size() != series.size() || shift.size() != series.size() ||
            phase.size() != series.size() || width_index.size() != series.size() ||
            snr.size() != series.size())
          throw std::invalid_argument("compact peak fields must have equal lengths");
        std::vector<gaffa_hip_ffa_peak> c(n);
        for (std::size_t i = 0; i < n; ++i) {
          if (phase.data()[i] > 0xFFFFu || width_index.data()[i] > 0xFFFFu)
            throw std::invalid_argument("phase and width index must fit 16 bits");
          c[i] = gaffa_hip_ffa_peak{series.data()[i], task.data()[i], shift.data()[i],
                                    phase.data()[i] | (width_index.data()[i] << 16),
                                    snr.data()[i]};
        }
        const GpuFfaExecutionPlan e(plan);
        std::vector<DmPeak> out;
        const std::vector<std::uint32_t> ranks =
            ranked ? ffa_period_ranks(plan) : std::vector<std::uint32_t>{};
        if (parts <= 1) {
          expand_sorted_dm_peaks_into(e, c, dms, out, ranks);
          return out;
        }
        // as the search pipeline does it: keys sorted and peaks expanded
        // per part, then one merge
        std::vector<DmPeakKey> keys(n), scratch;
        std::vector<DmPeak> expanded(n);
        std::vector<std::size_t> ends;
        for (std::size_t k = 0; k < parts; ++k) {
          const std::size_t b = n * k / parts, f = n * (k + 1) / parts;
          const std::span<const gaffa_hip_ffa_peak> part(c.data() + b, f - b);
          dm_peak_keys(e, part, b, dms, ranks, keys.data() + b);
          sort_dm_peak_keys(std::span<DmPeakKey>(keys.data() + b, f - b));
          expand_dm_peaks(e, part, dms, expanded.data() + b);
          ends.push_back(f);
        }
        merge_sorted_dm_peaks(expanded, ends, keys, scratch, out);
        return out;
      },
      py::arg("plan"), py::arg("series"), py::arg("task"), py::arg("shift"), py::arg("phase"),
      py::arg("width_index"), py::arg("snr"), py::arg("dms"), py::arg("ranked") = true,
      py::arg("parts") = 1,
      "Host-side expansion + is_better_dm_peak ordering of compact GPU peaks (no device use); "
      "parts > 1 sorts and expands contiguous parts separately and merges them");
  m.def(
      "_find_dm_peaks_gpu",
      [](const py::array& plane, const std::vector<double>& dms, double tsamp, double period_min,
         double period_max, std::size_t bins_min, std::size_t bins_max, std::size_t min_periods,
         double duty, double spacing, std::optional<std::size_t> window, std::size_t min_points,
         bool normalise, float snr_threshold, const py::object& max_peaks, int device_id) {
        const py::buffer_info info = plane.request();
        if (info.ndim != 2 || info.itemsize != 4) throw py::value_error("plane must be 2D uint32");
        require_c_contiguous(info, "plane");
        const std::size_t ndm = static_cast<std::size_t>(info.shape[0]);
        const std::size_t n = static_cast<std::size_t>(info.shape[1]);
        DmSearchOptions o;
        o.plan = RiptideFfaPlanOptions{period_min, period_max, bins_min, bins_max,
                                       min_periods, duty, spacing, 1'000'000};
        o.preprocess = preprocess_plan_from(window, min_points, normalise, true);
        o.snr_threshold = snr_threshold;
        o.max_peaks = peak_limit(max_peaks);
        DmSearchResult r;
        {
          py::gil_scoped_release release;
          GpuDeviceScope scope(device_id);
          GpuDeviceBuffer<std::uint32_t> d =
              upload(static_cast<const std::uint32_t*>(info.ptr), ndm * n);
          r = find_dm_peaks_gpu(GpuDedispersedView<std::uint32_t>{d.data(), {ndm, n}, n, device_id},
                                dms, tsamp, o, device_id);
        }
        return r.peaks;
      },
      py::arg("plane"), py::arg("dms"), py::arg("tsamp"), py::kw_only(), py::arg("period_min"),
      py::arg("period_max"), py::arg("bins_min") = 180, py::arg("bins_max") = 256,
      py::arg("min_periods") = 1, py::arg("duty_cycle_max") = 0.20,
      py::arg("width_trial_spacing") = 1.5, py::arg("window_samples") = py::none(),
      py::arg("min_points") = 101, py::arg("normalise") = true, py::arg("snr_threshold") = 6.0F,
      py::arg("max_peaks") = py::none(), py::arg("device_id") = 0);

  py::class_<SynthPulsar>(m, "SynthPulsar")
      .def(py::init([](double period_s, double dm, double duty, double amp_sigma, double phase0) {
             return SynthPulsar{period_s, dm, duty, amp_sigma, phase0};
           }),
           py::arg("period_s"), py::arg("dm"), py::arg("duty"), py::arg("amp_sigma"),
           py::arg("phase0") = 0.0)
      .def_readonly("period_s", &SynthPulsar::period_s)
      .def_readonly("dm", &SynthPulsar::dm)
      .def_readonly("duty", &SynthPulsar::duty)
      .def_readonly("amp_sigma", &SynthPulsar::amp_sigma)
      .def_readonly("phase0", &SynthPulsar::phase0);
  py::class_<SynthFilterbankSpec>(m, "SynthFilterbankSpec")
      .def(py::init([](std::size_t nsamples, std::size_t nchans, double fch1, double foff,
                       double tsamp, std::uint64_t seed, std::vector<SynthPulsar> pulsars,
                       std::vector<std::size_t> rfi, std::int32_t rfi_offset) {
             return SynthFilterbankSpec{nsamples, nchans, fch1, foff, tsamp, seed,
                                        std::move(pulsars), std::move(rfi), rfi_offset};
           }),
           py::kw_only(), py::arg("nsamples"), py::arg("nchans"), py::arg("fch1"),
           py::arg("foff"), py::arg("tsamp"), py::arg("seed"),
           py::arg("pulsars") = std::vector<SynthPulsar>{},
           py::arg("rfi_channels") = std::vector<std::size_t>{}, py::arg("rfi_offset") = 0)
      .def_readonly("nsamples", &SynthFilterbankSpec::nsamples)
      .def_readonly("nchans", &SynthFilterbankSpec::nchans)
      .def_readonly("tsamp", &SynthFilterbankSpec::tsamp)
      .def_property_readonly("frequencies", &synth_frequencies);
  m.def(
      "_synth_filterbank_gpu",
      [](const SynthFilterbankSpec& spec, std::size_t t_begin, std::optional<std::size_t> t_count,
         int device_id) {
        const std::size_t cnt = t_count.value_or(spec.nsamples - t_begin);
        std::vector<std::uint8_t> host(cnt * spec.nchans);
        {
          py::gil_scoped_release release;
          GpuDeviceScope scope(device_id);
          GpuDeviceBuffer<std::uint8_t> d(host.size());
          synth_filterbank_gpu(spec, d.data(), t_begin, cnt, nullptr);
          copy_to_host(host.data(), d.data(), host.size(), nullptr);
        }
        return owning_array(std::move(host), {static_cast<py::ssize_t>(cnt),
                                              static_cast<py::ssize_t>(spec.nchans)});
      },
      py::arg("spec"), py::kw_only(), py::arg("t_begin") = 0, py::arg("t_count") = py::none(),
      py::arg("device_id") = 0);

  py::class_<DmSegment>(m, "DmSegment")
      .def(py::init([](double lo, double step, std::size_t n) { return DmSegment{lo, step, n}; }),
           py::arg("dm_low"), py::arg("dm_step"), py::arg("ndm"))
      .def_readonly("dm_low", &DmSegment::dm_low)
      .def_readonly("dm_step", &DmSegment::dm_step)
      .def_readonly("ndm", &DmSegment::ndm);
  py::class_<GpuSearchPipeline>(m, "GpuSearchPipeline")
      .def(py::init([](std::size_t nsamples, std::size_t nchans, double tsamp,
                       std::vector<double> freq, std::vector<DmSegment> segments, bool subband,
                       std::size_t subband_channels, std::size_t ndm_per_nominal,
                       std::optional<std::size_t> window, std::size_t min_points, bool normalise,
                       double period_min, double period_max, std::size_t bins_min,
                       std::size_t bins_max, double duty, double spacing, float snr_threshold,
                       std::size_t ffa_workspace_bytes, std::size_t ffa_series_tile,
                       int device_id) {
             SearchPipelineConfig c;
             c.nsamples = nsamples;
             c.nchans = nchans;
             c.tsamp = tsamp;
             c.frequency_mhz = std::move(freq);
             c.dm_segments = std::move(segments);
             c.subband = subband;
             c.subband_options = SubbandDedispersionOptions{subband_channels, ndm_per_nominal};
             c.preprocess = preprocess_plan_from(window, min_points, normalise, true);
             c.ffa = RiptideFfaPlanOptions{period_min, period_max, bins_min, bins_max, 1, duty,
                                           spacing, 1'000'000};
             c.snr_threshold = snr_threshold;
             c.ffa_workspace_bytes_limit = ffa_workspace_bytes;
             c.ffa_series_tile = ffa_series_tile;
             c.device_id = device_id;
             py::gil_scoped_release release;
             return std::make_unique<GpuSearchPipeline>(std::move(c));
           }),
           py::kw_only(), py::arg("nsamples"), py::arg("nchans"), py::arg("tsamp"),
           py::arg("frequency_mhz"), py::arg("dm_segments"), py::arg("subband") = true,
           py::arg("subband_channels") = 32, py::arg("ndm_per_nominal") = 32,
           py::arg("window_samples") = py::none(), py::arg("min_points") = 101,
           py::arg("normalise") = true, py::arg("period_min"), py::arg("period_max"),
           py::arg("bins_min") = 180, py::arg("bins_max") = 256, py::arg("duty_cycle_max") = 0.20,
           py::arg("width_trial_spacing") = 1.5, py::arg("snr_threshold") = 6.0F,
           py::arg("ffa_workspace_bytes") = 32ULL << 30, py::arg("ffa_series_tile") = 0,
           py::arg("device_id") = 0)
      .def("generate", [](GpuSearchPipeline& p, const SynthFilterbankSpec& s) {
        py::gil_scoped_release release;
        p.generate(s);
      })
      .def("upload", [](GpuSearchPipeline& p, const py::array_t<std::uint8_t>& a) {
        const py::buffer_info info = a.request();
        require_c_contiguous(info, "filterbank data");
        py::gil_scoped_release release;
        p.upload(static_cast<const std::uint8_t*>(info.ptr), static_cast<std::size_t>(info.size));
      })
      .def("dedisperse", [](GpuSearchPipeline& p) {
        py::gil_scoped_release release;
        p.dedisperse();
      })
      .def("preprocess", [](GpuSearchPipeline& p) {
        py::gil_scoped_release release;
        p.preprocess();
      })
      .def("search", [](GpuSearchPipeline& p) {
        py::gil_scoped_release release;
        p.search();
      })
      .def(
          "select_candidates",
          [](GpuSearchPipeline& p, double radius, std::size_t dm_radius, bool across,
             std::size_t max_candidates, bool device) {
            py::gil_scoped_release release;
            const CandidateSelectionOptions o{radius, dm_radius, across, max_candidates};
            return device ? p.select_candidates(o) : p.select_candidates_host(o);
          },
          py::kw_only(), py::arg("frequency_cluster_radius") = 0.1,
          py::arg("dm_cluster_radius") = 1, py::arg("cluster_across_widths") = true,
          py::arg("max_candidates") = 0, py::arg("device") = true,
          "Candidates from the last search's peaks (observation = out_nsamples * tsamp); "
          "device=False runs the host selection")
      .def(
          "fold_candidates",
          [](GpuSearchPipeline& p, const std::vector<std::size_t>& dm_index,
             const std::vector<double>& period, std::size_t nbin, std::size_t nsubint) {
            if (dm_index.size() != period.size() || dm_index.empty())
              throw std::invalid_argument("one DM index per period is required");
            if (nsubint == 0) throw std::invalid_argument("fold nsubint must be positive");
            std::vector<FoldJob> jobs;
            for (std::size_t i = 0; i < dm_index.size(); ++i) {
              if (dm_index[i] >= p.ndm()) throw std::out_of_range("dm_index out of range");
              jobs.push_back(FoldJob{dm_index[i] * p.plane_stride(), period[i], 0});
            }
            const std::size_t n = p.out_nsamples();
            const std::size_t sub = (n + nsubint - 1) / nsubint;
            FoldedCube c;
            {
              py::gil_scoped_release release;
              c = fold_batch_gpu(GpuFoldInput{p.plane_u32(), FoldDtype::F32, n, 1, 1, 0}, jobs,
                                 p.tsamp(), nbin, sub, 1, nullptr, p.stream());
            }
            const auto J = static_cast<py::ssize_t>(jobs.size()),
                       S = static_cast<py::ssize_t>(c.nsubint), B = static_cast<py::ssize_t>(nbin);
            return py::make_tuple(owning_array(std::move(c.data), {J, S, B}),
                                  owning_array(std::move(c.exposure), {J, S, B}));
          },
          py::arg("dm_index"), py::arg("period"), py::kw_only(), py::arg("nbin") = 64,
          py::arg("nsubint") = 32,
          "Subint x phase cubes of the preprocessed plane rows (after run())")
      .def("enable_periodogram", &GpuSearchPipeline::enable_periodogram, py::arg("on") = true,
           "Keep the dense S/N periodogram of the following searches (scorer patches)")
      .def_property_readonly("trial_periods",
                             [](const GpuSearchPipeline& p) {
                               const std::vector<double>& v = p.trial_periods();
                               py::array_t<double> out(static_cast<py::ssize_t>(v.size()));
                               std::memcpy(out.mutable_data(), v.data(), v.size() * sizeof(double));
                               return out;
                             })
      .def(
          "periodogram",
          [](const GpuSearchPipeline& p, std::size_t dm_index) {
            const auto rows = static_cast<py::ssize_t>(p.trial_periods().size());
            const auto w = static_cast<py::ssize_t>(p.periodogram_widths());
            py::array_t<float> out({rows, w});
            p.download_periodogram(dm_index, out.mutable_data());
            return out;
          },
          py::arg("dm_index"), "S/N [trial period][width trial] of one DM trial")
      .def(
          "candidate_features",
          [](GpuSearchPipeline& p, const std::vector<Candidate>& c) {
            const std::size_t n = c.size();
            py::array_t<float> cubes({static_cast<py::ssize_t>(n),
                                      static_cast<py::ssize_t>(GAFFA_HIP_SCORER_SUBINTS),
                                      static_cast<py::ssize_t>(GAFFA_HIP_SCORER_PHASE_BINS)});
            py::array_t<float> patches({static_cast<py::ssize_t>(n),
                                        static_cast<py::ssize_t>(GAFFA_HIP_SCORER_PATCH)});
            float* hc = cubes.mutable_data();
            float* hp = patches.mutable_data();
            {
              py::gil_scoped_release release;
              GpuDeviceScope scope(p.device_id());
              GpuDeviceBuffer<float> dc(n * kScorerCubeFloats), dp(n * kScorerPatchFloats);
              p.candidate_features(c, dc.data(), dp.data());
              copy_to_host(hc, dc.data(), dc.bytes(), p.stream());
              copy_to_host(hp, dp.data(), dp.bytes(), p.stream());
              hip_ok(gaffa_hip_stream_synchronize(p.stream()), "hipStreamSynchronize");
            }
            return py::make_tuple(cubes, patches);
          },
          py::arg("candidates"),
          "Scorer inputs (cubes (n, 32, 64), patches (n, 128)) of candidates of the last search")
      .def(
          "fold_candidates_full",
          [](GpuSearchPipeline& p, const std::vector<Candidate>& c, std::size_t nbin,
             std::size_t nsubint, std::size_t output_channels, bool keep) -> py::tuple {
            std::vector<float> cubes;
            double ms = 0.0;
            {
              py::gil_scoped_release release;
              ms = p.fold_candidates_full(c, nbin, nsubint, output_channels,
                                          keep ? &cubes : nullptr);
            }
            if (!keep) return py::make_tuple(ms, py::none());
            const std::size_t nout = output_channels == 0 ? p.nchans() : output_channels;
            return py::make_tuple(
                ms, owning_array(std::move(cubes),
                                 {static_cast<py::ssize_t>(c.size()),
                                  static_cast<py::ssize_t>(nsubint),
                                  static_cast<py::ssize_t>(nout), static_cast<py::ssize_t>(nbin)}));
          },
          py::arg("candidates"), py::kw_only(), py::arg("nbin") = 64, py::arg("nsubint") = 32,
          py::arg("output_channels") = 64, py::arg("keep_cubes") = false,
          "gaffa's full fold of each candidate from the filterbank: (device ms, cubes or None)")
      .def(
          "score_candidates",
          [](GpuSearchPipeline& p, GpuCandidateScorer& scorer, const std::vector<Candidate>& c) {
            const std::size_t n = c.size();
            py::array_t<float> out(static_cast<py::ssize_t>(n));
            float* ho = out.mutable_data();
            {
              py::gil_scoped_release release;
              GpuDeviceScope scope(p.device_id());
              GpuDeviceBuffer<float> dc(n * kScorerCubeFloats), dp(n * kScorerPatchFloats),
                  dz(n);
              p.candidate_features(c, dc.data(), dp.data());
              scorer.forward_device(dc.data(), dp.data(), n, dz.data(), p.stream());
              copy_to_host(ho, dz.data(), dz.bytes(), p.stream());
              hip_ok(gaffa_hip_stream_synchronize(p.stream()), "hipStreamSynchronize");
            }
            return out;
          },
          py::arg("scorer"), py::arg("candidates"),
          "Scorer logits of candidates of the last search, features and network on the GPU")
      .def(
          "run_file",
          [](GpuSearchPipeline& p, const std::string& path, std::size_t chunk_samples) {
            StageTimes t;
            {
              py::gil_scoped_release release;
              t = p.run_file(path, chunk_samples);
            }
            py::dict d = stage_dict(t);
            const FilterbankStreamStats& s = p.stream_stats();
            d["stream_bytes"] = s.bytes;
            d["stream_chunks"] = s.chunks;
            d["stream_read_ms"] = s.read_ms;
            d["stream_wall_ms"] = s.wall_ms;
            return d;
          },
          py::arg("path"), py::arg("chunk_samples") = std::size_t{1} << 18,
          "Stream an 8-bit SIGPROC file in and search it, dedispersing while it arrives")
      .def("set_kernel_timing", &GpuSearchPipeline::set_kernel_timing, py::arg("on"),
           "Record HIP events between the FFA launch stages of every tile")
      .def("run", [](GpuSearchPipeline& p) {
        StageTimes t;
        {
          py::gil_scoped_release release;
          t = p.run();
        }
        return stage_dict(t);
      })
      .def("download_plane_u32", [](const GpuSearchPipeline& p) {
        std::vector<std::uint32_t> host(p.ndm() * p.out_nsamples());
        p.download_plane(host.data());
        return owning_array(std::move(host), {static_cast<py::ssize_t>(p.ndm()),
                                              static_cast<py::ssize_t>(p.out_nsamples())});
      })
      .def("download_plane_f32", [](const GpuSearchPipeline& p) {
        std::vector<std::uint32_t> bits(p.ndm() * p.out_nsamples());
        p.download_plane(bits.data());
        std::vector<float> host(bits.size());
        std::memcpy(host.data(), bits.data(), bits.size() * 4);
        return owning_array(std::move(host), {static_cast<py::ssize_t>(p.ndm()),
                                              static_cast<py::ssize_t>(p.out_nsamples())});
      })
      .def("download_filterbank", [](const GpuSearchPipeline& p, std::size_t nsamples,
                                     std::size_t nchans) {
        std::vector<std::uint8_t> host(nsamples * nchans);
        p.download_filterbank(host.data());
        return owning_array(std::move(host), {static_cast<py::ssize_t>(nsamples),
                                              static_cast<py::ssize_t>(nchans)});
      })
      .def_property_readonly("ndm", &GpuSearchPipeline::ndm)
      .def_property_readonly("segment_ndm_per_nominal",
                             &GpuSearchPipeline::segment_ndm_per_nominal)
      .def_property_readonly("out_nsamples", &GpuSearchPipeline::out_nsamples)
      .def_property_readonly("dms", &GpuSearchPipeline::dms)
      .def_property_readonly("ffa_plan", &GpuSearchPipeline::ffa_plan)
      .def_property_readonly("peaks", &GpuSearchPipeline::peaks)
      .def_property_readonly("peak_count", &GpuSearchPipeline::peak_count)
      .def(
          "wait_peaks",
          [](const GpuSearchPipeline& p) {
            py::gil_scoped_release release;
            p.wait_peaks();
          },
          "Wait for the last search's peaks to be ordered (the search ends by ordering "
          "them on a host thread)")
      .def(
          "peak_records",
          [](const GpuSearchPipeline& p, std::size_t dm_begin, std::size_t dm_end) {
            // raw DmPeak records, in list order, of the DM trials in
            // [dm_begin, dm_end): a filtered sorted list stays sorted
            if (dm_end > p.ndm()) dm_end = p.ndm();
            std::vector<DmPeak> out;
            {
              py::gil_scoped_release release;
              for (const DmPeak& q : p.peaks())
                if (q.dm_index >= dm_begin && q.dm_index < dm_end) out.push_back(q);
            }
            return py::bytes(reinterpret_cast<const char*>(out.data()),
                             out.size() * sizeof(DmPeak));
          },
          py::arg("dm_begin") = 0, py::arg("dm_end") = static_cast<std::size_t>(-1),
          "Peaks of DM trials [dm_begin, dm_end) as packed DmPeak records (bytes)")
      .def_property_readonly("ffa_tile", [](const GpuSearchPipeline& p) {
        return p.ffa_program().tile_capacity();
      })
      .def_property_readonly(
          "ffa_layout",
          [](const GpuSearchPipeline& p) {
            const GpuFfaExecutionPlan& e = p.ffa_program().execution_plan();
            std::vector<std::uint32_t> depth, terminal_levels;
            for (const auto& t : e.tasks()) {
              depth.push_back(t.depth);
              terminal_levels.push_back(t.terminal_levels);
            }
            std::size_t level_launches = 0;
            const auto& off = e.level_offsets();
            for (std::size_t d = 1; d + 1 < off.size(); ++d)
              if (off[d + 1] > off[d]) ++level_launches;
            py::dict d;
            d["task_depth"] = owning_array(std::move(depth),
                                           {static_cast<py::ssize_t>(e.tasks().size())});
            d["terminal_levels"] = owning_array(
                std::move(terminal_levels), {static_cast<py::ssize_t>(e.tasks().size())});
            d["max_depth"] = e.max_depth();
            d["level_launches"] = level_launches;
            d["subtree_items"] = e.subtree_items().size();
            d["level_items"] = e.level_items().size();
            d["terminal_items"] = e.terminal_items().size();
            d["subtree_lds_bytes"] = e.subtree_lds_bytes();
            d["terminal_lds_bytes"] = e.terminal_lds_bytes();
            d["level_lds_bytes"] = e.level_lds_bytes();
            return d;
          },
          "Device layout of the FFA plan: LDS depth per task, launches and items");

  // learned candidate scorer
  py::class_<GpuCandidateScorer>(m, "CandidateScorer")
      .def(py::init<std::size_t, int>(), py::arg("capacity") = 1024, py::arg("device_id") = 0)
      .def_property_readonly("parameter_count", &GpuCandidateScorer::parameter_count)
      .def_property_readonly("capacity", &GpuCandidateScorer::capacity)
      .def_static("layout",
                  [] {
                    const gaffa_hip_scorer_layout l = GpuCandidateScorer::layout();
                    py::dict d;
                    d["weight"] = std::vector<std::uint64_t>(l.weight, l.weight + 6);
                    d["bias"] = std::vector<std::uint64_t>(l.bias, l.bias + 6);
                    d["fan_in"] = std::vector<std::uint64_t>(l.fan_in, l.fan_in + 6);
                    d["nparams"] = l.nparams;
                    return d;
                  })
      .def("initialize", &GpuCandidateScorer::initialize, py::arg("seed"))
      .def_property(
          "parameters",
          [](const GpuCandidateScorer& s) {
            const std::vector<float> p = s.parameters();
            py::array_t<float> out(static_cast<py::ssize_t>(p.size()));
            std::memcpy(out.mutable_data(), p.data(), p.size() * sizeof(float));
            return out;
          },
          [](GpuCandidateScorer& s, const FloatArray& p) {
            if (p.ndim() != 1) throw py::value_error("parameters must be 1D");
            s.set_parameters(std::span<const float>(p.data(), static_cast<std::size_t>(p.size())));
          })
      .def(
          "forward",
          [](GpuCandidateScorer& s, const FloatArray& cubes, const FloatArray& patches) {
            const std::size_t n = scorer_rows(cubes, patches, nullptr);
            const float* c = cubes.data();
            const float* q = patches.data();
            std::vector<float> z;
            {
              py::gil_scoped_release release;
              z = s.forward(c, q, n);
            }
            py::array_t<float> out(static_cast<py::ssize_t>(n));
            std::memcpy(out.mutable_data(), z.data(), n * sizeof(float));
            return out;
          },
          py::arg("cubes"), py::arg("patches"), "Logits (n,) of the candidates")
      .def(
          "gradients",
          [](GpuCandidateScorer& s, const FloatArray& cubes, const FloatArray& patches,
             const FloatArray& labels) {
            const std::size_t n = scorer_rows(cubes, patches, &labels);
            const float* c = cubes.data();
            const float* q = patches.data();
            const float* y = labels.data();
            std::vector<float> g;
            float loss = 0.0f;
            {
              py::gil_scoped_release release;
              loss = s.gradients(c, q, y, n, g);
            }
            py::array_t<float> out(static_cast<py::ssize_t>(g.size()));
            std::memcpy(out.mutable_data(), g.data(), g.size() * sizeof(float));
            return py::make_tuple(loss, out);
          },
          py::arg("cubes"), py::arg("patches"), py::arg("labels"),
          "(mean BCE loss, gradient of every parameter) for one batch")
      .def(
          "fit",
          [](GpuCandidateScorer& s, const FloatArray& cubes, const FloatArray& patches,
             const FloatArray& labels, std::size_t epochs, std::size_t batch, float lr,
             std::uint64_t seed) {
            const std::size_t n = scorer_rows(cubes, patches, &labels);
            ScorerTrainOptions o;
            o.epochs = epochs;
            o.batch = batch;
            o.learning_rate = lr;
            o.seed = seed;
            const float* c = cubes.data();
            const float* q = patches.data();
            const float* y = labels.data();
            py::gil_scoped_release release;
            return s.fit(c, q, y, n, o);
          },
          py::arg("cubes"), py::arg("patches"), py::arg("labels"), py::kw_only(),
          py::arg("epochs") = 4, py::arg("batch") = 256, py::arg("learning_rate") = 1e-3f,
          py::arg("seed") = 1, "Adam training on the GPU; returns the mean loss per epoch");
}
