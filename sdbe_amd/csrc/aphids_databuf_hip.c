/*
 * Device side of the APHIDS input and output threads on HIP, under the link
 * names of software/aphids/vdif_in_databuf_cuda.cu and vdif_out_databuf_cuda.cu
 * so that the reference's vdif_in_databuf.c / vdif_out_databuf.c wrappers (and
 * the threads that call them) link against it unchanged.
 *
 * Differences from the reference: errors are reported and returned as -1
 * instead of aborting the process; the host-to-device copy moves only the
 * packets the slot holds; the device count comes from the runtime instead of
 * a compile-time NUM_GPU.
 */
#include <stdio.h>
#include <stdlib.h>

#include <hip/hip_runtime_api.h>

#include "aphids_databuf.h"

static int hip_ok(hipError_t e, const char *what) {
    if (e == hipSuccess) return 1;
    hashpipe_error("aphids_databuf_hip", "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
    return -1;
}

static int device_of(int index) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return 0;
    return index % n;
}

static void *new_stream(void) {
    hipStream_t *s = (hipStream_t *)malloc(sizeof(hipStream_t));
    if (!s) return NULL;
    if (hipStreamCreate(s) != hipSuccess) {
        free(s);
        return NULL;
    }
    return s;
}

static void free_stream(void **slot) {
    if (*slot) {
        (void)hipStreamDestroy(*(hipStream_t *)*slot);
        free(*slot);
        *slot = NULL;
    }
}

int get_bgv_cpu_memory_cuda(beng_group_vdif_buffer_t **bgv_buf_cpu, int index) {
    if (hip_ok(hipSetDevice(device_of(index)), "hipSetDevice") != 1) return -1;
    return hip_ok(hipHostMalloc((void **)bgv_buf_cpu, sizeof(beng_group_vdif_buffer_t), hipHostMallocDefault),
                  "hipHostMalloc (group buffer)");
}

int get_bgv_gpu_memory_cuda(beng_group_vdif_buffer_t **bgv_buf_gpu, int index) {
    if (hip_ok(hipSetDevice(device_of(index)), "hipSetDevice") != 1) return -1;
    return hip_ok(hipMalloc((void **)bgv_buf_gpu, sizeof(beng_group_vdif_buffer_t)), "hipMalloc (group buffer)");
}

int transfer_beng_group_to_gpu_cuda(vdif_in_databuf_t *bgc_buf, int index) {
    beng_group_completion_t *g = &bgc_buf->bgc[index];
    const int dev = device_of(index);
    if (hip_ok(hipSetDevice(dev), "hipSetDevice") != 1) return -1;
    g->memcpy_stream = new_stream();
    if (!g->memcpy_stream) return hip_ok(hipErrorOutOfMemory, "hipStreamCreate");
    g->gpu_id = dev;
    int n = g->beng_group_vdif_packet_count;
    if (n < 0) n = 0;
    if (n > BENG_FRAMES_PER_GROUP * VDIF_PER_BENG_FRAME) n = BENG_FRAMES_PER_GROUP * VDIF_PER_BENG_FRAME;
    if (n == 0) return 1;
    return hip_ok(hipMemcpyAsync(g->bgv_buf_gpu, g->bgv_buf_cpu, (size_t)n * sizeof(vdif_in_packet_t),
                                 hipMemcpyHostToDevice, *(hipStream_t *)g->memcpy_stream),
                  "hipMemcpyAsync (group to device)");
}

int check_transfer_beng_group_to_gpu_complete_cuda(vdif_in_databuf_t *bgc_buf, int index) {
    beng_group_completion_t *g = &bgc_buf->bgc[index];
    if (!g->memcpy_stream) return 1;
    if (hip_ok(hipSetDevice(g->gpu_id), "hipSetDevice") != 1) return -1;
    const hipError_t e = hipStreamQuery(*(hipStream_t *)g->memcpy_stream);
    if (e == hipErrorNotReady) return 0;
    if (hip_ok(e, "hipStreamQuery (group to device)") != 1) return -1;
    free_stream(&g->memcpy_stream);
    return 1;
}

int get_vdg_cpu_memory_cuda(vdif_out_data_group_t **vdg_buf_cpu, int index) {
    if (hip_ok(hipSetDevice(device_of(index)), "hipSetDevice") != 1) return -1;
    return hip_ok(hipHostMalloc((void **)vdg_buf_cpu, sizeof(vdif_out_data_group_t), hipHostMallocDefault),
                  "hipHostMalloc (output group)");
}

int transfer_vdif_group_to_cpu_cuda(vdif_out_databuf_t *qs_buf, int index) {
    quantized_storage_t *q = &qs_buf->blocks[index];
    if (hip_ok(hipSetDevice(q->gpu_id), "hipSetDevice") != 1) return -1;
    q->memcpy_stream = new_stream();
    if (!q->memcpy_stream) return hip_ok(hipErrorOutOfMemory, "hipStreamCreate");
    return hip_ok(hipMemcpyAsync(q->vdg_buf_cpu, q->vdg_buf_gpu, sizeof(vdif_out_data_group_t),
                                 hipMemcpyDeviceToHost, *(hipStream_t *)q->memcpy_stream),
                  "hipMemcpyAsync (output group to host)");
}

int check_transfer_vdif_group_to_cpu_complete_cuda(vdif_out_databuf_t *qs_buf, int index) {
    quantized_storage_t *q = &qs_buf->blocks[index];
    if (!q->memcpy_stream) return 1;
    if (hip_ok(hipSetDevice(q->gpu_id), "hipSetDevice") != 1) return -1;
    const int rv = hip_ok(hipStreamSynchronize(*(hipStream_t *)q->memcpy_stream), "hipStreamSynchronize");
    free_stream(&q->memcpy_stream);
    return rv;
}

/* the reference's wrappers (vdif_in_databuf.c:169-190, vdif_out_databuf.c:16-33) */
int get_bgv_cpu_memory(beng_group_vdif_buffer_t **b, int index) { return get_bgv_cpu_memory_cuda(b, index); }
int get_bgv_gpu_memory(beng_group_vdif_buffer_t **b, int index) { return get_bgv_gpu_memory_cuda(b, index); }
int transfer_beng_group_to_gpu(vdif_in_databuf_t *d, int index) { return transfer_beng_group_to_gpu_cuda(d, index); }
int check_transfer_beng_group_to_gpu_complete(vdif_in_databuf_t *d, int index) {
    return check_transfer_beng_group_to_gpu_complete_cuda(d, index);
}
int get_vdg_cpu_memory(vdif_out_data_group_t **b, int index) { return get_vdg_cpu_memory_cuda(b, index); }
int transfer_vdif_group_to_cpu(vdif_out_databuf_t *d, int index) { return transfer_vdif_group_to_cpu_cuda(d, index); }
int check_transfer_vdif_group_to_cpu_complete(vdif_out_databuf_t *d, int index) {
    return check_transfer_vdif_group_to_cpu_complete_cuda(d, index);
}
