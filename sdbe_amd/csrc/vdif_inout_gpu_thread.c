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
 * Resampler, chosen at compile time as the reference chooses its rate
 * (sample_rates.h:12-63):
 *   SDBE_PLUGIN_TRIMMED = 0 (default): the 39 -> 64 resampler of the
 *     prototypes (resample.py fft_batched), 6/11 rate only, 4096 Msps out;
 *   SDBE_PLUGIN_TRIMMED = 1: the reference thread's band-trimmed resampler at
 *     SWARM_RATE_FRAC (6, 8, 10 or 11; aphids_resampler_init,
 *     vdif_inout_gpu_thread.cu:184-347), trimming the MHz read once at start
 *     from the status key TRIMDC (the reference reads the redis key
 *     trim_from_dc, :770-781; 0 when the key is absent).
 * N_32bit_words_per_chan is (2 * 16384 * 128 * EXPANSION) / 16 words, as the
 * reference sets it (:1012), with EXPANSION = 64 for the 39 -> 64 resampler.
 *
 * Differences from the reference, all deliberate:
 *   - one device output buffer per output slot (the reference reuses a single
 *     buffer for every slot, which the output thread may still be copying);
 *   - HIP errors are reported through hashpipe_error and put the thread into
 *     its error state (the reference ignores most CUDA errors);
 *   - no redis context: the trim comes from the hashpipe status buffer;
 *   - band-trimmed frames follow each other per block (E per block and
 *     polarisation), where the reference writes E of every
 *     VDIF_OUT_PKTS_PER_BLOCK / 128 payload slots (the same at 10/11 and 11/11).
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

#ifndef SDBE_PLUGIN_TRIMMED
#define SDBE_PLUGIN_TRIMMED 0
#endif
#if SDBE_PLUGIN_TRIMMED
#define PLUGIN_RATE SWARM_RATE_FRAC
#define PLUGIN_EXPANSION EXPANSION_FACTOR
#else
#if SWARM_RATE_FRAC != 6
#error "the 39 -> 64 resampler takes 39-snapshot blocks of the 6/11 rate: build with -DSDBE_PLUGIN_TRIMMED=1"
#endif
#define PLUGIN_RATE SDBE_RATE_64_39
#define PLUGIN_EXPANSION 64
#endif
/* 2-bit words per polarisation and group: 128 blocks of 32768 E samples */
#define PLUGIN_WORDS_PER_CHAN ((2 * 16384 * 128 * PLUGIN_EXPANSION) / (32 / 2))
_Static_assert((PLUGIN_WORDS_PER_CHAN / (VDIF_OUT_PKT_DATA_SIZE / 4)) <= VDIF_OUT_PKTS_PER_BLOCK,
               "a group's frames fit the output slot");

typedef struct {
    sdbe_handle_t *h;
    vdif_out_data_group_t *out[VDIF_OUT_BUFFER_SIZE];
} gpu_slot_t;

static int ensure_gpu(gpu_slot_t *g, int dev, int trim_mhz) {
    if (g->h) return 0;
    sdbe_config_t cfg;
    sdbe_default_config(&cfg);
    cfg.thresh_policy = SDBE_THRESH_REFERENCE;
    cfg.out_mode = SDBE_OUT_PAYLOAD;
    cfg.rate_frac = PLUGIN_RATE;
    cfg.skip_mhz = SDBE_PLUGIN_TRIMMED ? trim_mhz : 0;
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

/* MHz trimmed from DC: status key TRIMDC, read once (0 when absent). */
static int read_trim_mhz(hashpipe_thread_args_t *args) {
    int mhz = 0;
    hashpipe_status_t st = args->st;
    if (!st.buf) return 0;
    hashpipe_status_lock_safe(&st);
    if (!hgeti4(st.buf, "TRIMDC", &mhz)) mhz = 0;
    hashpipe_status_unlock_safe(&st);
    return mhz;
}

static void *run_method(hashpipe_thread_args_t *args) {
    vdif_in_databuf_t *db_in = (vdif_in_databuf_t *)args->ibuf;
    vdif_out_databuf_t *db_out = (vdif_out_databuf_t *)args->obuf;
    gpu_slot_t gpus[MAX_GPUS];
    memset(gpus, 0, sizeof(gpus));
    int index_in = 0, index_out = 0, rv;
    int state = STATE_INIT;
    const int trim_mhz = read_trim_mhz(args);

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
        if (dev < 0 || dev >= MAX_GPUS || hipSetDevice(dev) != hipSuccess || ensure_gpu(&gpus[dev], dev, trim_mhz) != 0) {
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
        qs->N_32bit_words_per_chan = PLUGIN_WORDS_PER_CHAN;
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
