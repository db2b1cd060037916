/* crdfft runtime plumbing: device memory, copies, streams and events through the library's own HIP runtime.
 *
 * These entry points have no counterpart in the reference (its FFT never leaves the CPU). They exist so that a host
 * program or binding that does not link HIP itself (a ctypes test, a benchmark driver) can hand crdfft_execute device
 * buffers and a stream from the same HIP runtime instance the kernels run on.
 */
#ifndef CRDFFT_RUNTIME_H
#define CRDFFT_RUNTIME_H

#include "crdfft.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct crdfft_device_props
{
    char name[256];
    char arch[64];            /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int32_t compute_units;
    int32_t clock_khz;
    int32_t memory_clock_khz;
    int32_t memory_bus_width;
    uint64_t global_mem_bytes;
    uint64_t l2_bytes;
    int32_t max_lds_per_block;
} crdfft_device_props;

CRDFFT_API crdfft_status crdfft_device_count(int32_t* count);
CRDFFT_API crdfft_status crdfft_device_props_get(int32_t device, crdfft_device_props* props);

CRDFFT_API crdfft_status crdfft_malloc(void** ptr, uint64_t bytes);
CRDFFT_API crdfft_status crdfft_free(void* ptr);
/* *is_device = 1 when ptr points into device (or managed) memory that kernels can read, 0 for host memory (pageable
 * or pinned) and for NULL. The C++ interface uses it to run reference call sites that pass host spans. */
CRDFFT_API crdfft_status crdfft_pointer_is_device(const void* ptr, int32_t* is_device);
/* Synchronous copies (host <-> device) and an asynchronous device copy on `stream`. */
CRDFFT_API crdfft_status crdfft_memcpy_h2d(void* dst, const void* src, uint64_t bytes);
CRDFFT_API crdfft_status crdfft_memcpy_d2h(void* dst, const void* src, uint64_t bytes);
CRDFFT_API crdfft_status crdfft_memcpy_d2d(void* dst, const void* src, uint64_t bytes, void* stream);
/* Streaming device copy with 16-byte accesses (bytes a multiple of 16): the achievable-bandwidth probe. */
CRDFFT_API crdfft_status crdfft_stream_copy(void* dst, const void* src, uint64_t bytes, void* stream);
/* Fills `bytes` with the byte `value` by a kernel on `stream` (no hipMemset node, safe under graph capture). */
CRDFFT_API crdfft_status crdfft_fill_bytes(void* dst, int32_t value, uint64_t bytes, void* stream);

CRDFFT_API crdfft_status crdfft_stream_create(void** stream);
CRDFFT_API crdfft_status crdfft_stream_destroy(void* stream);
CRDFFT_API crdfft_status crdfft_stream_synchronize(void* stream);
CRDFFT_API crdfft_status crdfft_device_synchronize(void);

CRDFFT_API crdfft_status crdfft_event_create(void** event);
CRDFFT_API crdfft_status crdfft_event_destroy(void* event);
CRDFFT_API crdfft_status crdfft_event_record(void* event, void* stream);
CRDFFT_API crdfft_status crdfft_event_synchronize(void* event);
CRDFFT_API crdfft_status crdfft_event_elapsed_ms(float* ms, void* start, void* stop);

/* hipGraph capture of any sequence of executes on one (non-null) stream: crdfft_graph_begin starts capturing,
 * crdfft_graph_end stops and instantiates it (*graph), crdfft_graph_launch replays it on any stream. Every plan's
 * execute is capturable once the plan has executed on that stream before (its per-stream scratch then exists; a
 * first execute on a new stream allocates and must not be captured). The plans record no memset nodes (the
 * cooperative plan clears its control words with a kernel); the Welch PSD of a signal shorter than one frame is the
 * one call that records a memset, and is not meant for capture. */
CRDFFT_API crdfft_status crdfft_graph_begin(void* stream);
CRDFFT_API crdfft_status crdfft_graph_end(void* stream, void** graph);
CRDFFT_API crdfft_status crdfft_graph_launch(void* graph, void* stream);
CRDFFT_API crdfft_status crdfft_graph_destroy(void* graph);

#ifdef __cplusplus
}
#endif

#endif /* CRDFFT_RUNTIME_H */
