/* owl_hip.h -- thin C-ABI device layer over HIP for gfx950.
 *
 * Replaces the reference's OpenCL wrapper and runtime kernel compiler:
 * src/clwrap.h:46-122 (device, buffer, queue, launch, read/write, events) and
 * src/KernelCompiler.h:14-37 (named kernel loading). Kernels are compiled
 * offline for gfx950 into the same shared library and looked up by name.
 *
 * Every function returns 0 on success and a non-zero code on failure; the
 * message of the last failure on the calling thread is owl_last_error().
 * Pointers are device pointers unless named host*. No torch types cross this
 * boundary.
 */
#ifndef OWL_HIP_H
#define OWL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct owl_queue_s* owl_queue;        /* an in-order HIP stream (src/Queue.h:23-63) */
typedef const struct owl_kernel_s* owl_kernel; /* a named kernel (src/Kernel.h:17-79) */
typedef struct owl_event_s* owl_event;        /* a timing event (src/Event.h:9-26) */
typedef struct owl_graph_s* owl_graph;        /* a captured, instantiated launch sequence */

/* Message of the last failure on this thread ("" if none). */
const char* owl_last_error(void);

/* Devices: getAllDeviceIDs / getDevice / getDeviceName / getFreeMem (src/clwrap.h:46-70). */
int owl_hip_device_count(int* count);
int owl_hip_set_device(int device);
int owl_hip_device_info(int device, char* name, size_t nameLen, size_t* totalMem, size_t* freeMem,
                        int* computeUnits, int* ldsPerBlock);

/* Buffers: makeBuf_ / release (src/clwrap.h:72-80, src/Buffer.h:16-102). */
int owl_hip_alloc(void** ptr, size_t bytes);
int owl_hip_free(void* ptr);

/* Queues: makeQueue / finish / flush (src/clwrap.h:82-86). */
int owl_hip_queue_create(owl_queue* queue);
int owl_hip_queue_destroy(owl_queue queue);
int owl_hip_finish(owl_queue queue);

/* Transfers: write / read / copyBuf / fillBuf (src/clwrap.h:88-104). blocking != 0 waits for completion. */
int owl_hip_write(owl_queue queue, void* dst, const void* hostSrc, size_t bytes, int blocking);
int owl_hip_read(owl_queue queue, void* hostDst, const void* src, size_t bytes, int blocking);
int owl_hip_copy(owl_queue queue, void* dst, const void* src, size_t bytes);
int owl_hip_fill32(owl_queue queue, void* dst, uint32_t value, size_t count);

/* Kernels: loadKernel by name / run (src/clwrap.h:106-112, src/clwrap.cpp:314-323).
 * args is an array of pointers to the argument values, in kernel parameter order. */
int owl_hip_kernel_get(const char* name, owl_kernel* kernel);
/* Comma-separated names of every compiled kernel (no clwrap counterpart: the reference
   compiles kernels at run time); *needed = bytes including the terminator. */
int owl_hip_kernel_names(char* out, size_t cap, size_t* needed);
int owl_hip_kernel_max_threads(owl_kernel kernel, uint32_t* maxThreads);
int owl_hip_launch(owl_queue queue, owl_kernel kernel, uint32_t gridX, uint32_t gridY, uint32_t blockX,
                   uint32_t ldsBytes, void** args);

/* Events: getEventNanos (src/clwrap.h:114-118). */
int owl_hip_event_create(owl_event* event);
int owl_hip_event_record(owl_event event, owl_queue queue);
int owl_hip_event_elapsed_ms(owl_event start, owl_event stop, float* ms);
int owl_hip_event_destroy(owl_event event);

/* Graphs (no reference counterpart; the reference enqueues every launch):
 * capture the launches made on queue between begin and end, then replay them. */
int owl_hip_graph_begin(owl_queue queue);
int owl_hip_graph_end(owl_queue queue, owl_graph* graph);
int owl_hip_graph_launch(owl_graph graph, owl_queue queue);
int owl_hip_graph_destroy(owl_graph graph);

#ifdef __cplusplus
}
#endif

#endif
