// Runtime entry points of the C ABI: device queries, memory, streams,
// events and graphs.  Thin, checked wrappers over the HIP runtime.
#include <cstring>

#include "common.hpp"

namespace gaffa_hip {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& message) { g_last_error = message; }
}  // namespace gaffa_hip

using namespace gaffa_hip;

namespace gaffa_hip {
// A kernel fill, for memory that a captured graph resets on every replay
// (memset nodes are avoided in graphs).
// Grid-stride, so one launch covers any count.
__global__ void __launch_bounds__(256) fill_u32_kernel(uint32_t* dst, uint32_t value,
                                                       uint64_t count) {
  const uint64_t step = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < count; i += step)
    dst[i] = value;
}
}  // namespace gaffa_hip

extern "C" {

const char* gaffa_hip_last_error(void) { return g_last_error.c_str(); }

int gaffa_hip_abi_version(void) { return GAFFA_HIP_ABI_VERSION; }

int gaffa_hip_device_count(int32_t* count) {
  return guarded([&] {
    int n = 0;
    const hipError_t err = hipGetDeviceCount(&n);
    *count = err == hipSuccess ? n : 0;
  });
}

int gaffa_hip_device_info_get(int32_t device, gaffa_hip_device_info* out) {
  return guarded([&] {
    check(out != nullptr, "device info output must not be null");
    hipDeviceProp_t p{};
    hip_check(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties");
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->name, p.name, sizeof(out->name) - 1);
    std::strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
    out->compute_units = p.multiProcessorCount;
    out->clock_khz = p.clockRate;
    out->memory_clock_khz = p.memoryClockRate;
    out->memory_bus_width = p.memoryBusWidth;
    out->total_memory = static_cast<int64_t>(p.totalGlobalMem);
    out->l2_bytes = p.l2CacheSize;
    out->lds_bytes_per_block = static_cast<int32_t>(p.sharedMemPerBlock);
    int rt = 0;
    hip_check(hipRuntimeGetVersion(&rt), "hipRuntimeGetVersion");
    out->runtime_version = rt;
  });
}

int gaffa_hip_set_device(int32_t device) {
  return guarded([&] { hip_check(hipSetDevice(device), "hipSetDevice"); });
}

int gaffa_hip_get_device(int32_t* device) {
  return guarded([&] {
    int d = 0;
    hip_check(hipGetDevice(&d), "hipGetDevice");
    *device = d;
  });
}

int gaffa_hip_malloc(void** ptr, uint64_t bytes) {
  return guarded([&] {
    check(ptr != nullptr, "allocation output must not be null");
    *ptr = nullptr;
    if (bytes == 0) return;
    hip_check(hipMalloc(ptr, bytes), "hipMalloc");
  });
}

int gaffa_hip_free(void* ptr) {
  return guarded([&] {
    if (ptr != nullptr) hip_check(hipFree(ptr), "hipFree");
  });
}

int gaffa_hip_memcpy(void* dst, const void* src, uint64_t bytes,
                     gaffa_hip_stream stream) {
  return guarded([&] {
    if (bytes == 0) return;
    check(dst != nullptr && src != nullptr, "memcpy pointers must not be null");
    if (stream == nullptr) {
      hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDefault), "hipMemcpy");
    } else {
      hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, as_stream(stream)),
                "hipMemcpyAsync");
    }
  });
}

int gaffa_hip_memset(void* dst, int32_t value, uint64_t bytes,
                     gaffa_hip_stream stream) {
  return guarded([&] {
    if (bytes == 0) return;
    check(dst != nullptr, "memset pointer must not be null");
    hip_check(hipMemsetAsync(dst, value, bytes, as_stream(stream)), "hipMemsetAsync");
    if (stream == nullptr) hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
  });
}

int gaffa_hip_stream_create(gaffa_hip_stream* stream) {
  return guarded([&] {
    hipStream_t s = nullptr;
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    *stream = reinterpret_cast<gaffa_hip_stream>(s);
  });
}

int gaffa_hip_stream_destroy(gaffa_hip_stream stream) {
  return guarded([&] {
    if (stream != nullptr) hip_check(hipStreamDestroy(as_stream(stream)), "hipStreamDestroy");
  });
}

int gaffa_hip_stream_synchronize(gaffa_hip_stream stream) {
  return guarded(
      [&] { hip_check(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize"); });
}

int gaffa_hip_device_synchronize(void) {
  return guarded([&] { hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); });
}

int gaffa_hip_event_create(gaffa_hip_event* event) {
  return guarded([&] {
    hipEvent_t e = nullptr;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    *event = reinterpret_cast<gaffa_hip_event>(e);
  });
}

int gaffa_hip_event_destroy(gaffa_hip_event event) {
  return guarded([&] {
    if (event != nullptr)
      hip_check(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)), "hipEventDestroy");
  });
}

int gaffa_hip_event_record(gaffa_hip_event event, gaffa_hip_stream stream) {
  return guarded([&] {
    hip_check(hipEventRecord(reinterpret_cast<hipEvent_t>(event), as_stream(stream)),
              "hipEventRecord");
  });
}

int gaffa_hip_event_synchronize(gaffa_hip_event event) {
  return guarded([&] {
    check(event != nullptr, "event must not be null");
    hip_check(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)), "hipEventSynchronize");
  });
}

int gaffa_hip_stream_wait_event(gaffa_hip_stream stream, gaffa_hip_event event) {
  return guarded([&] {
    check(event != nullptr, "event must not be null");
    hip_check(hipStreamWaitEvent(as_stream(stream), reinterpret_cast<hipEvent_t>(event), 0),
              "hipStreamWaitEvent");
  });
}

int gaffa_hip_host_malloc(void** ptr, uint64_t bytes) {
  return guarded([&] {
    check(ptr != nullptr, "host allocation output must not be null");
    *ptr = nullptr;
    if (bytes == 0) return;
    hip_check(hipHostMalloc(ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
  });
}

int gaffa_hip_host_free(void* ptr) {
  return guarded([&] {
    if (ptr != nullptr) hip_check(hipHostFree(ptr), "hipHostFree");
  });
}

int gaffa_hip_event_elapsed_ms(gaffa_hip_event start, gaffa_hip_event stop, float* ms) {
  return guarded([&] {
    hip_check(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)), "hipEventSynchronize");
    hip_check(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                                  reinterpret_cast<hipEvent_t>(stop)),
              "hipEventElapsedTime");
  });
}

int gaffa_hip_fill_u32(uint32_t* dst, uint32_t value, uint64_t count, gaffa_hip_stream stream) {
  return guarded([&] {
    check(dst != nullptr || count == 0, "fill destination must not be null");
    if (count == 0) return;
    const uint64_t blocks = std::min<uint64_t>(ceil_div(count, 256), 1ull << 20);
    hipLaunchKernelGGL(gaffa_hip::fill_u32_kernel, dim3(static_cast<uint32_t>(blocks)),
                       dim3(256), 0, as_stream(stream), dst, value, count);
    launch_check("fill_u32_kernel");
  });
}

int gaffa_hip_graph_begin(gaffa_hip_stream stream) {
  return guarded([&] {
    check(stream != nullptr, "graph capture requires a non-default stream");
    hip_check(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal),
              "hipStreamBeginCapture");
  });
}

int gaffa_hip_graph_end(gaffa_hip_stream stream, gaffa_hip_graph* graph) {
  return guarded([&] {
    hipGraph_t g = nullptr;
    hip_check(hipStreamEndCapture(as_stream(stream), &g), "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    const hipError_t err = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    hip_check(err, "hipGraphInstantiate");
    *graph = reinterpret_cast<gaffa_hip_graph>(exec);
  });
}

int gaffa_hip_graph_launch(gaffa_hip_graph graph, gaffa_hip_stream stream) {
  return guarded([&] {
    hip_check(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph), as_stream(stream)),
              "hipGraphLaunch");
  });
}

int gaffa_hip_graph_destroy(gaffa_hip_graph graph) {
  return guarded([&] {
    if (graph != nullptr)
      hip_check(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph)),
                "hipGraphExecDestroy");
  });
}

}  // extern "C"
