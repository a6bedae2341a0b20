/*
 * vdif_out_file_thread -- hashpipe output thread plugin: the per-group work of
 * the APHIDS output thread (software/aphids/vdif_out_net_thread.c:180-336)
 * writing each polarisation's VDIF frames to a file instead of a TCP socket
 * (the socket transport, sgcomm_net.c tx_frames, is not reproduced; the
 * reference's own file writer, vdif_out_file_thread.c:98-136, writes the
 * slot's bookkeeping struct rather than frames).  Per filled input slot:
 *
 *   a valid header template starts the stream (:203-237): secs, df and
 *     ref_epoch from it, frames skipped to the next second when fewer than a
 *     group's frames remain in this one, the picoseconds hidden in edh_psn
 *     plus SWARM_ZERO_DELAY_OFFSET_PICOSECONDS turned into int32 words to
 *     skip at R2DBE_RATE, frames this group can fill (num_usable) and the
 *     words it carries into the next (leftover), psn = 1;
 *   the group's payloads come to the host (transfer_vdif_group_to_cpu + wait,
 *     :239-245) and the slot is freed;
 *   per polarisation the stream is the previous group's leftover words, then
 *     this group's words from skip_int32 + 2048 skip_frames on; the group's
 *     last leftover words are kept (:247-288);
 *   num_usable frames are stamped (station, pol, bdc / rx sideband, bps =
 *     bit_depth - 1, df wrapping at 125000 frames per second with secs + 1,
 *     psn + 1; the fields init_vdif_out sets, vdif_out_databuf.c:35-67) and
 *     written to the polarisation's file (:290-336).
 *
 * Status keys read at start: VOUTPOL0, VOUTPOL1 (output file of each
 * polarisation, required), STATION (2 characters, default "XX"), BDCSB, RXSB
 * (sideband bits, default 0) -- the reference's redis keys vdif:station,
 * vdif:bdc, vdif:rx.  Written: VOUTGRPS (groups written), VOUTFRMS (frames
 * written per polarisation).
 *
 * Difference from the reference, deliberate: on the first group of a stream
 * that skips whole frames the reference copies skip_frames * 2048 words past
 * the end of the group's buffer (the copy length does not subtract them);
 * here the copy stops at the end of the group.  The frames sent are the same.
 */
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "aphids_databuf.h"

#define STATE_ERROR -1
#define STATE_PROCESS 1
#define WORDS_PER_FRAME (VDIF_OUT_PKT_DATA_SIZE / 4)
#define R2DBE_RATE 4096e6

static int status_i4(hashpipe_thread_args_t *args, const char *key, int dflt) {
    int v = dflt;
    hashpipe_status_t st = args->st;
    if (!st.buf) return dflt;
    hashpipe_status_lock_safe(&st);
    if (!hgeti4(st.buf, key, &v)) v = dflt;
    hashpipe_status_unlock_safe(&st);
    return v;
}

static void status_s(hashpipe_thread_args_t *args, const char *key, char *out, int n) {
    hashpipe_status_t st = args->st;
    out[0] = 0;
    if (!st.buf) return;
    hashpipe_status_lock_safe(&st);
    if (!hgets(st.buf, key, n, out)) out[0] = 0;
    hashpipe_status_unlock_safe(&st);
}

typedef struct {
    uint32_t secs, df, ref_epoch;
    uint64_t psn;
    int leftover;                      /* words carried between groups */
    int carry_n;                       /* words held in carry (0: none yet) */
    uint32_t *carry[VDIF_CHAN];        /* [WORDS_PER_FRAME] each */
} out_stream_t;

/* pinned host buffers of the slots, allocated on first use and kept for the
 * life of the process (the reference's vdg_buf_cpu) */
static vdif_out_data_group_t *g_host[VDIF_OUT_BUFFER_SIZE];

static void *run_method(hashpipe_thread_args_t *args) {
    vdif_out_databuf_t *db_in = (vdif_out_databuf_t *)args->ibuf;
    vdif_out_databuf_t local;          /* host-side view of the slots (qs_buf of the reference) */
    vdif_out_data_group_t **host = g_host;
    uint32_t *stream[VDIF_CHAN] = {0};
    vdif_out_packet_t *frames = NULL;
    FILE *fo[VDIF_CHAN] = {0};
    out_stream_t s;
    memset(&s, 0, sizeof(s));
    memset(&local, 0, sizeof(local));
    int state = STATE_PROCESS, index = 0, rv;
    long groups = 0, frames_out = 0;

    char path[VDIF_CHAN][256], station[16];
    status_s(args, "VOUTPOL0", path[0], sizeof(path[0]));
    status_s(args, "VOUTPOL1", path[1], sizeof(path[1]));
    status_s(args, "STATION", station, sizeof(station));
    if (strlen(station) < 2) strcpy(station, "XX");
    const uint32_t station_id = ((uint32_t)(uint8_t)station[0] << 8) | (uint32_t)(uint8_t)station[1];
    const uint32_t bdc = (uint32_t)status_i4(args, "BDCSB", 0) & 1u, rx = (uint32_t)status_i4(args, "RXSB", 0) & 1u;
    for (int p = 0; p < VDIF_CHAN; ++p) {
        fo[p] = path[p][0] ? fopen(path[p], "wb") : NULL;
        /* a group's words plus the carried ones */
        stream[p] = (uint32_t *)malloc(sizeof(vdif_out_data_block_t) + VDIF_OUT_PKT_DATA_SIZE);
        s.carry[p] = (uint32_t *)malloc(VDIF_OUT_PKT_DATA_SIZE);
        if (!fo[p] || !stream[p] || !s.carry[p]) {
            hashpipe_error("vdif_out_file_thread", "cannot open output '%s' (%s)", path[p], strerror(errno));
            state = STATE_ERROR;
        }
    }
    frames = (vdif_out_packet_t *)malloc(256 * sizeof(vdif_out_packet_t));
    if (!frames) state = STATE_ERROR;

    while (run_threads()) {
        if (state == STATE_ERROR) {
            sleep(1);
            continue;
        }
        while ((rv = hashpipe_databuf_wait_filled(&db_in->header, index)) != HASHPIPE_OK) {
            if (rv == HASHPIPE_TIMEOUT) {
                if (!run_threads()) goto done;
                continue;
            }
            hashpipe_error("vdif_out_file_thread", "error waiting for filled databuf %d", index);
            state = STATE_ERROR;
            break;
        }
        if (state == STATE_ERROR) continue;

        quantized_storage_t *q = &local.blocks[index];
        *q = db_in->blocks[index];
        if (!host[index] && get_vdg_cpu_memory(&host[index], index) != 1) {
            state = STATE_ERROR;
            continue;
        }
        q->vdg_buf_cpu = host[index];
        const int n_words = q->N_32bit_words_per_chan;
        int skip_int32 = 0, skip_frames = 0, num_usable = n_words / WORDS_PER_FRAME;
        if (n_words < 0 || n_words > (int)(sizeof(vdif_out_data_block_t) / 4)) {
            hashpipe_error("vdif_out_file_thread", "N_32bit_words_per_chan %d out of range", n_words);
            state = STATE_ERROR;
            continue;
        }
        if (!q->vdif_header_template.w0.invalid) {
            s.secs = q->vdif_header_template.w0.secs_inre;
            s.df = q->vdif_header_template.w1.df_num_insec;
            if (s.df > (uint32_t)(VDIF_OUT_FRAMES_PER_SECOND - num_usable)) {
                skip_frames = VDIF_OUT_FRAMES_PER_SECOND - (int)s.df;
                s.df = 0;
                s.secs++;
            }
            s.ref_epoch = q->vdif_header_template.w1.ref_epoch;
            const uint64_t skip_ps = q->vdif_header_template.edh_psn + SWARM_ZERO_DELAY_OFFSET_PICOSECONDS;
            const int skip_samples = (int)floor((double)skip_ps * R2DBE_RATE / 1e12);
            skip_int32 = (int)floor((double)skip_samples / (8.0 / (double)q->bit_depth) / (double)sizeof(int32_t));
            num_usable = (n_words - skip_int32) / WORDS_PER_FRAME - skip_frames;
            s.leftover = n_words - skip_int32 - num_usable * WORDS_PER_FRAME - skip_frames * WORDS_PER_FRAME;
            s.psn = 1;
        }
        if (transfer_vdif_group_to_cpu(&local, index) != 1 || check_transfer_vdif_group_to_cpu_complete(&local, index) != 1) {
            state = STATE_ERROR;
            continue;
        }
        hashpipe_databuf_set_free(&db_in->header, index);

        const int L = s.leftover;
        const int start = skip_int32 + skip_frames * WORDS_PER_FRAME;
        int count = n_words - skip_int32 - L;
        if (start + count > n_words) count = n_words - start;
        for (int p = 0; p < VDIF_CHAN; ++p) {
            const uint32_t *src = (const uint32_t *)host[index]->chan[p].datas;
            /* the carried words, at most L of them */
            const int pos = s.carry_n < L ? s.carry_n : L;
            if (pos > 0) memcpy(stream[p], s.carry[p], (size_t)pos * 4);
            if (count > 0) memcpy(stream[p] + pos, src + start, (size_t)count * 4);
            if (L > 0) memcpy(s.carry[p], src + n_words - L, (size_t)L * 4);
        }
        if (L > 0) s.carry_n = L;
        /* stamp and write num_usable frames per polarisation, 256 at a time */
        for (int p = 0; p < VDIF_CHAN && state != STATE_ERROR; ++p) {
            uint32_t secs = s.secs, df = s.df;
            uint64_t psn = s.psn;
            for (int f0 = 0; f0 < num_usable; f0 += 256) {
                const int nf = num_usable - f0 < 256 ? num_usable - f0 : 256;
                for (int k = 0; k < nf; ++k) {
                    uint32_t *h = (uint32_t *)&frames[k].header;
                    h[0] = secs & 0x3FFFFFFFu;
                    h[1] = (df & 0xFFFFFFu) | ((s.ref_epoch & 0x3Fu) << 24);
                    h[2] = (uint32_t)(sizeof(vdif_out_packet_t) / 8);
                    h[3] = station_id | ((uint32_t)(q->bit_depth - 1) << 26);
                    h[4] = (uint32_t)p | (bdc << 1) | (rx << 2) | (2u << 24);
                    h[5] = 0;
                    h[6] = (uint32_t)(psn & 0xFFFFFFFFu);
                    h[7] = (uint32_t)(psn >> 32);
                    memcpy(frames[k].data.data, stream[p] + (size_t)(f0 + k) * WORDS_PER_FRAME, VDIF_OUT_PKT_DATA_SIZE);
                    if (++df == VDIF_OUT_FRAMES_PER_SECOND) {
                        df = 0;
                        secs++;
                    }
                    psn++;
                }
                if (fwrite(frames, sizeof(vdif_out_packet_t), (size_t)nf, fo[p]) != (size_t)nf) {
                    hashpipe_error("vdif_out_file_thread", "error writing '%s'", path[p]);
                    state = STATE_ERROR;
                    break;
                }
            }
            if (p == VDIF_CHAN - 1) {
                s.secs = secs;
                s.df = df;
                s.psn = psn;
            }
        }
        for (int p = 0; p < VDIF_CHAN; ++p)
            if (fo[p]) fflush(fo[p]);
        groups++;
        frames_out += num_usable;
        {
            hashpipe_status_t st = args->st;
            if (st.buf) {
                hashpipe_status_lock_safe(&st);
                hputi4(st.buf, "VOUTGRPS", (int)groups);
                hputi4(st.buf, "VOUTFRMS", (int)frames_out);
                hashpipe_status_unlock_safe(&st);
            }
        }
        index = (index + 1) % db_in->header.n_block;
    }
done:
    for (int p = 0; p < VDIF_CHAN; ++p) {
        if (fo[p]) fclose(fo[p]);
        free(stream[p]);
        free(s.carry[p]);
    }
    free(frames);
    return NULL;
}

static hashpipe_thread_desc_t vdif_out_file_thread = {
    .name = "vdif_out_file_thread",
    .skey = "VDIFOUT",
    .init = NULL,
    .run = run_method,
    .ibuf_desc = {vdif_out_databuf_create},
    .obuf_desc = {NULL},
};

static __attribute__((constructor)) void ctor(void) { register_hashpipe_thread(&vdif_out_file_thread); }
