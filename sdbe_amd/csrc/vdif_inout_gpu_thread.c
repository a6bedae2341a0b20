/*
 * vdif_inout_gpu_thread -- hashpipe thread plugin, drop-in for the APHIDS GPU
 * stage of the same name (software/aphids/vdif_inout_gpu_thread.cu:745-1059):
 * same name and status key, same databufs (vdif_in_databuf_t in,
 * vdif_out_databuf_t out, include/aphids_databuf.h), same per-slot protocol:
 *
 *   wait_filled(in, i) -> wait_free(out, j) -> select bgc.gpu_id ->
 *   process the group -> set_free(in, i) -> fill out->blocks[j]
 *   (bit_depth 2, N_32bit_words_per_chan, gpu_id, vdg_buf_gpu, header
 *   template copied from the input) -> set_filled(out, j)
 *
 * The group is processed by sdbe_process_group (include/sdbe_stage.h) with the
 * reference threshold policy and payload-only output (the layout of
 * vdif_out_data_group_t), so the reference output thread consumes it as is.
 *
 * Differences from the reference, all deliberate:
 *   - one device output buffer per output slot (the reference reuses a single
 *     buffer for every slot, which the output thread may still be copying);
 *   - HIP errors are reported through hashpipe_error and put the thread into
 *     its error state (the reference ignores most CUDA errors);
 *   - no redis context: status and the trim_from_dc key are not read (this
 *     stage implements the 64/39 resampling, which has no trim).
 */
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include <hip/hip_runtime_api.h>

#include "aphids_databuf.h"
#include "sdbe_stage.h"

#define STATE_ERROR -1
#define STATE_INIT 0
#define STATE_PROC 1
#define MAX_GPUS 8

typedef struct {
    sdbe_handle_t *h;
    vdif_out_data_group_t *out[VDIF_OUT_BUFFER_SIZE];
} gpu_slot_t;

static int ensure_gpu(gpu_slot_t *g, int dev) {
    if (g->h) return 0;
    sdbe_config_t cfg;
    sdbe_default_config(&cfg);
    cfg.thresh_policy = SDBE_THRESH_REFERENCE;
    cfg.out_mode = SDBE_OUT_PAYLOAD;
    int rc = sdbe_create(&cfg, dev, &g->h);
    if (rc != SDBE_OK) {
        hashpipe_error("vdif_inout_gpu_thread", "sdbe_create(device %d): %s", dev, sdbe_error_string(rc));
        g->h = NULL;
        return -1;
    }
    for (int j = 0; j < VDIF_OUT_BUFFER_SIZE; ++j) {
        if (hipMalloc((void **)&g->out[j], sizeof(vdif_out_data_group_t)) != hipSuccess) {
            hashpipe_error("vdif_inout_gpu_thread", "hipMalloc of output slot %d failed", j);
            return -1;
        }
    }
    return 0;
}

static void release_gpus(gpu_slot_t *gs) {
    for (int d = 0; d < MAX_GPUS; ++d) {
        if (!gs[d].h) continue;
        (void)hipSetDevice(d);
        for (int j = 0; j < VDIF_OUT_BUFFER_SIZE; ++j)
            if (gs[d].out[j]) (void)hipFree(gs[d].out[j]);
        sdbe_destroy(gs[d].h);
        gs[d].h = NULL;
    }
}

static void *run_method(hashpipe_thread_args_t *args) {
    vdif_in_databuf_t *db_in = (vdif_in_databuf_t *)args->ibuf;
    vdif_out_databuf_t *db_out = (vdif_out_databuf_t *)args->obuf;
    gpu_slot_t gpus[MAX_GPUS];
    memset(gpus, 0, sizeof(gpus));
    int index_in = 0, index_out = 0, rv;
    int state = STATE_INIT;

    while (run_threads()) {
        if (state == STATE_ERROR) {
            sleep(1);
            continue;
        }
        state = STATE_PROC;
        while ((rv = hashpipe_databuf_wait_filled(&db_in->header, index_in)) != HASHPIPE_OK) {
            if (rv == HASHPIPE_TIMEOUT) {
                if (!run_threads()) goto done;
                continue;
            }
            hashpipe_error(__func__, "error waiting for filled input databuf %d", index_in);
            state = STATE_ERROR;
            break;
        }
        if (state == STATE_ERROR) continue;
        while ((rv = hashpipe_databuf_wait_free(&db_out->header, index_out)) != HASHPIPE_OK) {
            if (rv == HASHPIPE_TIMEOUT) {
                if (!run_threads()) goto done;
                continue;
            }
            hashpipe_error(__func__, "error waiting for free output databuf %d", index_out);
            state = STATE_ERROR;
            break;
        }
        if (state == STATE_ERROR) continue;

        beng_group_completion_t bgc = db_in->bgc[index_in];
        const int dev = bgc.gpu_id;
        if (dev < 0 || dev >= MAX_GPUS || hipSetDevice(dev) != hipSuccess || ensure_gpu(&gpus[dev], dev) != 0) {
            hashpipe_error(__func__, "cannot use GPU %d", dev);
            state = STATE_ERROR;
            continue;
        }
        gpu_slot_t *g = &gpus[dev];
        vdif_out_data_group_t *out = g->out[index_out];
        int rc = sdbe_process_group(g->h, bgc.bgv_buf_gpu, bgc.beng_group_vdif_packet_count, bgc.bfc[0].b, out,
                                    (int64_t)sizeof(vdif_out_data_block_t), NULL, NULL);
        sdbe_stats_t st;
        if (rc == SDBE_OK) rc = sdbe_get_stats(g->h, &st);   /* synchronises */
        if (rc != SDBE_OK) {
            hashpipe_error(__func__, "group processing failed: %s (hip %d)", sdbe_error_string(rc),
                           sdbe_last_hip_error(g->h));
            state = STATE_ERROR;
            continue;
        }
        hashpipe_databuf_set_free(&db_in->header, index_in);
        index_in = (index_in + 1) % db_in->header.n_block;

        quantized_storage_t *qs = &db_out->blocks[index_out];
        qs->bit_depth = 2;
        qs->N_32bit_words_per_chan = SDBE_OUT_FRAMES_PER_GROUP * (VDIF_OUT_PKT_DATA_SIZE / 4);
        qs->gpu_id = dev;
        qs->vdg_buf_gpu = out;
        memcpy(&qs->vdif_header_template, &bgc.vdif_header_template, sizeof(vdif_in_header_t));
        hashpipe_databuf_set_filled(&db_out->header, index_out);
        index_out = (index_out + 1) % db_out->header.n_block;
    }
done:
    release_gpus(gpus);
    return NULL;
}

static hashpipe_thread_desc_t vdif_inout_gpu_thread = {
    .name = "vdif_inout_gpu_thread",
    .skey = "VDIFIO",
    .init = NULL,
    .run = run_method,
    .ibuf_desc = {vdif_in_databuf_create},
    .obuf_desc = {vdif_out_databuf_create},
};

static __attribute__((constructor)) void ctor(void) { register_hashpipe_thread(&vdif_inout_gpu_thread); }
