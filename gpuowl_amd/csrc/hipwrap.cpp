// C-ABI device layer over HIP (declared in include/owl_hip.h).
// Replaces src/clwrap.cpp and src/KernelCompiler.cpp: no runtime source
// compilation; kernels come from the offline gfx950 code object linked into
// this library and are found by name in owl_kernel_table.
#include "owl_hip.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>

struct OwlKernelEntry {
  const char* name;
  const void* fn;
  unsigned maxThreads;
};
extern "C" const OwlKernelEntry owl_kernel_table[];
extern "C" const OwlKernelEntry owl_kernel_table_row_a[];
extern "C" const OwlKernelEntry owl_kernel_table_row_b[];
extern "C" const OwlKernelEntry owl_kernel_table_row_c[];
extern "C" const OwlKernelEntry owl_kernel_table_colx_a[];
extern "C" const OwlKernelEntry owl_kernel_table_colx_b[];
extern "C" const OwlKernelEntry owl_kernel_table_colx_c[];
extern "C" const OwlKernelEntry owl_kernel_table_coly_a[];
extern "C" const OwlKernelEntry owl_kernel_table_coly_b[];

struct owl_queue_s {
  hipStream_t stream;
};
struct owl_kernel_s {
  OwlKernelEntry entry;
};
struct owl_event_s {
  hipEvent_t event;
};
struct owl_graph_s {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

static thread_local std::string lastError;

static int fail(int code, const std::string& what) {
  lastError = what;
  return code ? code : -1;
}

#define HIPCHECK(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      return fail((int)e_, std::string(#call) + ": " + hipGetErrorString(e_));         \
    }                                                                                   \
  } while (0)

extern "C" {

const char* owl_last_error(void) { return lastError.c_str(); }

// Used by the engine's C-ABI (capi.cpp) to report its own failures through the same slot.
int owl_set_error(const char* msg) {
  lastError = msg ? msg : "";
  return -1;
}

int owl_hip_device_count(int* count) {
  HIPCHECK(hipGetDeviceCount(count));
  return 0;
}

int owl_hip_set_device(int device) {
  HIPCHECK(hipSetDevice(device));
  return 0;
}

int owl_hip_device_info(int device, char* name, size_t nameLen, size_t* totalMem, size_t* freeMem,
                        int* computeUnits, int* ldsPerBlock) {
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, device));
  if (name && nameLen) {
    snprintf(name, nameLen, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (totalMem) { *totalMem = prop.totalGlobalMem; }
  if (computeUnits) { *computeUnits = prop.multiProcessorCount; }
  if (ldsPerBlock) {
    int v = 0;
    HIPCHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    *ldsPerBlock = v;
  }
  if (freeMem) {
    size_t f = 0, t = 0;
    HIPCHECK(hipMemGetInfo(&f, &t));
    *freeMem = f;
  }
  return 0;
}

int owl_hip_alloc(void** ptr, size_t bytes) {
  *ptr = nullptr;
  HIPCHECK(hipMalloc(ptr, bytes ? bytes : 1));
  return 0;
}

int owl_hip_free(void* ptr) {
  if (ptr) { HIPCHECK(hipFree(ptr)); }
  return 0;
}

int owl_hip_queue_create(owl_queue* queue) {
  auto* q = new owl_queue_s{};
  hipError_t e = hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete q;
    return fail((int)e, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  *queue = q;
  return 0;
}

int owl_hip_queue_destroy(owl_queue queue) {
  if (!queue) { return 0; }
  hipError_t e = hipStreamDestroy(queue->stream);
  delete queue;
  if (e != hipSuccess) { return fail((int)e, std::string("hipStreamDestroy: ") + hipGetErrorString(e)); }
  return 0;
}

int owl_hip_finish(owl_queue queue) {
  HIPCHECK(hipStreamSynchronize(queue->stream));
  return 0;
}

int owl_hip_write(owl_queue queue, void* dst, const void* hostSrc, size_t bytes, int blocking) {
  HIPCHECK(hipMemcpyAsync(dst, hostSrc, bytes, hipMemcpyHostToDevice, queue->stream));
  if (blocking) { HIPCHECK(hipStreamSynchronize(queue->stream)); }
  return 0;
}

int owl_hip_read(owl_queue queue, void* hostDst, const void* src, size_t bytes, int blocking) {
  HIPCHECK(hipMemcpyAsync(hostDst, src, bytes, hipMemcpyDeviceToHost, queue->stream));
  if (blocking) { HIPCHECK(hipStreamSynchronize(queue->stream)); }
  return 0;
}

int owl_hip_copy(owl_queue queue, void* dst, const void* src, size_t bytes) {
  HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, queue->stream));
  return 0;
}

namespace {
const std::initializer_list<const OwlKernelEntry*> kTables = {
    owl_kernel_table,       owl_kernel_table_row_a,  owl_kernel_table_row_b,  owl_kernel_table_row_c,
    owl_kernel_table_colx_a, owl_kernel_table_colx_b, owl_kernel_table_colx_c, owl_kernel_table_coly_a, owl_kernel_table_coly_b};
}  // namespace

int owl_hip_kernel_names(char* out, size_t cap, size_t* needed) {
  std::string all;
  for (const OwlKernelEntry* table : kTables) {
    for (const OwlKernelEntry* e = table; e->name; ++e) {
      if (!all.empty()) { all += ','; }
      all += e->name;
    }
  }
  if (needed) { *needed = all.size() + 1; }
  if (!out || cap < all.size() + 1) { return fail(-1, "kernel name buffer too small"); }
  memcpy(out, all.c_str(), all.size() + 1);
  return 0;
}

int owl_hip_kernel_get(const char* name, owl_kernel* kernel) {
  static std::mutex mtx;
  static std::unordered_map<std::string, owl_kernel_s*> cache;
  std::lock_guard<std::mutex> lock(mtx);
  auto it = cache.find(name);
  if (it != cache.end()) {
    *kernel = it->second;
    return 0;
  }
  for (const OwlKernelEntry* table : kTables) {
    for (const OwlKernelEntry* e = table; e->name; ++e) {
      if (!strcmp(e->name, name)) {
        auto* k = new owl_kernel_s{*e};
        cache[name] = k;
        *kernel = k;
        return 0;
      }
    }
  }
  *kernel = nullptr;
  return fail(-2, std::string("no kernel named '") + name + "'");
}

int owl_hip_kernel_max_threads(owl_kernel kernel, uint32_t* maxThreads) {
  *maxThreads = kernel->entry.maxThreads;
  return 0;
}

int owl_hip_fill32(owl_queue queue, void* dst, uint32_t value, size_t count) {
  owl_kernel k = nullptr;
  if (int rc = owl_hip_kernel_get("fill32", &k)) { return rc; }
  uint32_t n = (uint32_t)count;
  void* p = dst;
  void* args[] = {&p, &n, &value};
  return owl_hip_launch(queue, k, (n + 255) / 256, 1, 256, 0, args);
}

int owl_hip_launch(owl_queue queue, owl_kernel kernel, uint32_t gridX, uint32_t gridY, uint32_t blockX,
                   uint32_t ldsBytes, void** args) {
  if (blockX == 0 || blockX > kernel->entry.maxThreads) {
    return fail(-3, std::string("launch of ") + kernel->entry.name + ": block size " + std::to_string(blockX) +
                        " exceeds the kernel's bound " + std::to_string(kernel->entry.maxThreads));
  }
  if (gridX == 0 || gridY == 0) { return 0; }
  if (ldsBytes > 65536) {
    // Dynamic LDS above 64 KiB must be allowed per function (gfx950 has 160 KiB per CU).
    static std::mutex mtx;
    static std::unordered_map<const void*, uint32_t> allowed;
    std::lock_guard<std::mutex> lock(mtx);
    uint32_t& cur = allowed[kernel->entry.fn];
    if (cur < ldsBytes) {
      HIPCHECK(hipFuncSetAttribute(kernel->entry.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsBytes));
      cur = ldsBytes;
    }
  }
  HIPCHECK(hipLaunchKernel(kernel->entry.fn, dim3(gridX, gridY, 1), dim3(blockX, 1, 1), args, ldsBytes,
                           queue->stream));
  return 0;
}

int owl_hip_event_create(owl_event* event) {
  auto* e = new owl_event_s{};
  hipError_t r = hipEventCreate(&e->event);
  if (r != hipSuccess) {
    delete e;
    return fail((int)r, std::string("hipEventCreate: ") + hipGetErrorString(r));
  }
  *event = e;
  return 0;
}

int owl_hip_event_record(owl_event event, owl_queue queue) {
  HIPCHECK(hipEventRecord(event->event, queue->stream));
  return 0;
}

int owl_hip_event_elapsed_ms(owl_event start, owl_event stop, float* ms) {
  HIPCHECK(hipEventSynchronize(stop->event));
  HIPCHECK(hipEventElapsedTime(ms, start->event, stop->event));
  return 0;
}

int owl_hip_event_destroy(owl_event event) {
  if (!event) { return 0; }
  hipError_t r = hipEventDestroy(event->event);
  delete event;
  if (r != hipSuccess) { return fail((int)r, std::string("hipEventDestroy: ") + hipGetErrorString(r)); }
  return 0;
}

int owl_hip_graph_begin(owl_queue queue) {
  HIPCHECK(hipStreamBeginCapture(queue->stream, hipStreamCaptureModeThreadLocal));
  return 0;
}

int owl_hip_graph_end(owl_queue queue, owl_graph* graph) {
  auto* g = new owl_graph_s{};
  hipError_t r = hipStreamEndCapture(queue->stream, &g->graph);
  if (r != hipSuccess) {
    delete g;
    return fail((int)r, std::string("hipStreamEndCapture: ") + hipGetErrorString(r));
  }
  r = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (r != hipSuccess) {
    hipGraphDestroy(g->graph);
    delete g;
    return fail((int)r, std::string("hipGraphInstantiate: ") + hipGetErrorString(r));
  }
  *graph = g;
  return 0;
}

int owl_hip_graph_launch(owl_graph graph, owl_queue queue) {
  HIPCHECK(hipGraphLaunch(graph->exec, queue->stream));
  return 0;
}

int owl_hip_graph_destroy(owl_graph graph) {
  if (!graph) { return 0; }
  hipGraphExecDestroy(graph->exec);
  hipGraphDestroy(graph->graph);
  delete graph;
  return 0;
}

}  // extern "C"
