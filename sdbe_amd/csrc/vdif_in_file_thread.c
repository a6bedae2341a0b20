/*
 * vdif_in_file_thread -- hashpipe input thread plugin: the grouping loop of
 * the APHIDS input thread (software/aphids/vdif_in_net_thread.c:269-500) fed
 * from a file of received data instead of a TCP socket (the socket transport,
 * sgcomm_net.c rx_frames, is not reproduced).  Each "receive" reads a batch
 * of records from the file; then, as the reference does for every batch:
 *
 *   8pac records: unpack_8pac + beng_reform_headers (:292-296)
 *   qsort by B-count, invalid packets first (:48-68, :302)
 *   on the first batch: b_first = first valid B-count, the four group
 *     windows initialised at b_first + 1 + BENG_FRAMES_PER_GROUP * i (:322-339)
 *   per packet: get_beng_group_index_offset; negative -> counted and dropped;
 *     offset > MAX_INDEX_LOOK_AHEAD (2) -> hand the current group off
 *     unfilled; else insert, and hand it off once complete (:374-424)
 *   the first hand-off fills the header template from the current packet and
 *     the count of packets before the start (:399, :418)
 *   hand-off: wait_free(out, i) -> transfer_beng_group_to_gpu -> poll until
 *     the copy is done -> copy the slot's bookkeeping to the output databuf ->
 *     set_filled(out, i) -> re-initialise the slot after the furthest window
 *     (:437-488)
 *
 * Status keys read at start (hashpipe status buffer, in place of the
 * reference's redis context):
 *   VINFILE   path of the input file (required)
 *   VIN8PAC   1: 8504-byte 8pac records with new-format B-engine timestamps
 *             (the wire format, default); 0: 1056-byte packets already in the
 *             old format (no reform; the header template then has no T0)
 *   VINBATCH  records per receive (default 1024)
 *   VINFLUSH  1: at the end of the file hand off the groups still being
 *             filled, in order; 0 (default): stop as the reference does at
 *             end of transmission (STATE_DONE), leaving them
 * Status keys written: VINPKTS (packets received), VINUSED, VINSKIP (before
 * the start), VININVAL, VINAHEAD (too far ahead), VINGRPS (groups handed
 * off), VINDONE (1 at the end of the file).
 *
 * Differences from the reference, deliberate: the B-count comparison of the
 * sort does not truncate the 64-bit difference to int; the template of a
 * group completed by its batch's last packet is taken from that packet (the
 * reference reads one past it); errors are reported and end in the error
 * state instead of continuing with stale buffers.
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "aphids_databuf.h"

#define STATE_ERROR -1
#define STATE_PROCESS 1
#define STATE_DONE 99
#define MAX_INDEX_LOOK_AHEAD 2
#define PAC8_BYTES ((int)sizeof(vdif_8pac_packet_t))

typedef struct {
    beng_group_vdif_buffer_t *cpu[BENG_GROUPS_IN_BUFFER];
    beng_group_vdif_buffer_t *gpu[BENG_GROUPS_IN_BUFFER];
} in_buffers_t;

/* one per process, as the reference's file-scope buffers */
static in_buffers_t g_in;
static vdif_in_databuf_t g_local;     /* the slots being filled */

static int compar_beng(const void *a, const void *b) {
    const vdif_in_packet_t *pa = (const vdif_in_packet_t *)a, *pb = (const vdif_in_packet_t *)b;
    if (pa->header.w0.invalid) return pb->header.w0.invalid ? 0 : -1;
    if (pb->header.w0.invalid) return 1;
    const int64_t ba = get_packet_b_count((vdif_in_header_t *)&pa->header);
    const int64_t bb = get_packet_b_count((vdif_in_header_t *)&pb->header);
    return ba < bb ? -1 : (ba > bb ? 1 : 0);
}

/* The slots' pinned host and device group buffers, allocated once and kept
 * for the life of the process, as the reference's (a later pipeline in the
 * same process reuses them). */
static int init_method(hashpipe_thread_args_t *args) {
    (void)args;
    for (int i = 0; i < BENG_GROUPS_IN_BUFFER; ++i) {
        if ((!g_in.cpu[i] && get_bgv_cpu_memory(&g_in.cpu[i], i) != 1) ||
            (!g_in.gpu[i] && get_bgv_gpu_memory(&g_in.gpu[i], i) != 1)) {
            hashpipe_error("vdif_in_file_thread", "cannot allocate group buffers of slot %d", i);
            return -1;
        }
    }
    return 0;
}

static int status_i4(hashpipe_thread_args_t *args, const char *key, int dflt) {
    int v = dflt;
    hashpipe_status_t st = args->st;
    if (!st.buf) return dflt;
    hashpipe_status_lock_safe(&st);
    if (!hgeti4(st.buf, key, &v)) v = dflt;
    hashpipe_status_unlock_safe(&st);
    return v;
}

typedef struct {
    long all, used, skipped, invalid, ahead, groups;
    int done;
} in_counts_t;

static void publish(hashpipe_thread_args_t *args, const in_counts_t *c) {
    hashpipe_status_t st = args->st;
    if (!st.buf) return;
    hashpipe_status_lock_safe(&st);
    hputi4(st.buf, "VINPKTS", (int)c->all);
    hputi4(st.buf, "VINUSED", (int)c->used);
    hputi4(st.buf, "VINSKIP", (int)c->skipped);
    hputi4(st.buf, "VININVAL", (int)c->invalid);
    hputi4(st.buf, "VINAHEAD", (int)c->ahead);
    hputi4(st.buf, "VINGRPS", (int)c->groups);
    hputi4(st.buf, "VINDONE", c->done);
    hashpipe_status_unlock_safe(&st);
}

/* Hand slot i off to the GPU thread (vdif_in_net_thread.c:437-488).  Returns
 * 0, or -1 on an error, or 1 when the pipeline is stopping. */
static int hand_off(hashpipe_thread_args_t *args, vdif_in_databuf_t *db_out, int i) {
    int rv;
    while ((rv = hashpipe_databuf_wait_free(&db_out->header, i)) != HASHPIPE_OK) {
        if (rv == HASHPIPE_TIMEOUT) {
            if (!run_threads()) return 1;
            continue;
        }
        hashpipe_error("vdif_in_file_thread", "error waiting for free databuf %d", i);
        return -1;
    }
    if (transfer_beng_group_to_gpu(&g_local, i) != 1) return -1;
    while ((rv = check_transfer_beng_group_to_gpu_complete(&g_local, i)) == 0) usleep(100);
    if (rv != 1) return -1;
    memcpy(&db_out->bgc[i], &g_local.bgc[i], sizeof(beng_group_completion_t));
    hashpipe_databuf_set_filled(&db_out->header, i);
    const int prev = (i + BENG_GROUPS_IN_BUFFER - 1) % BENG_GROUPS_IN_BUFFER;
    init_beng_group(&g_local.bgc[i], g_in.cpu[i], g_in.gpu[i],
                    (int64_t)g_local.bgc[prev].bfc[BENG_FRAMES_PER_GROUP - 1].b + 1);
    return 0;
}

/* Read up to `batch` records; returns packets (old format, sorted by B-count)
 * in *pk and their count, 0 at the end of the file, -1 on an error. */
static long receive(FILE *f, int pac8, int batch, vdif_in_packet_t **pk, void *raw) {
    const size_t rec = pac8 ? (size_t)PAC8_BYTES : sizeof(vdif_in_packet_t);
    const size_t n_rec = fread(raw, rec, (size_t)batch, f);
    if (n_rec == 0) return ferror(f) ? -1 : 0;
    const long n = pac8 ? 8 * (long)n_rec : (long)n_rec;
    *pk = (vdif_in_packet_t *)malloc((size_t)n * sizeof(vdif_in_packet_t));
    if (!*pk) return -1;
    if (pac8) {
        unpack_8pac(*pk, (vdif_8pac_packet_t *)raw, (int)n_rec);
        beng_reform_headers(*pk, (int)n);
    } else {
        memcpy(*pk, raw, (size_t)n * sizeof(vdif_in_packet_t));
    }
    qsort(*pk, (size_t)n, sizeof(vdif_in_packet_t), compar_beng);
    return n;
}

static void *run_method(hashpipe_thread_args_t *args) {
    vdif_in_databuf_t *db_out = (vdif_in_databuf_t *)args->obuf;
    in_counts_t cnt;
    memset(&cnt, 0, sizeof(cnt));
    int state = STATE_PROCESS;
    char path[256] = {0};
    const int pac8 = status_i4(args, "VIN8PAC", 1);
    const int batch = status_i4(args, "VINBATCH", 1024);
    const int flush = status_i4(args, "VINFLUSH", 0);
    {
        hashpipe_status_t st = args->st;
        if (st.buf) {
            hashpipe_status_lock_safe(&st);
            if (!hgets(st.buf, "VINFILE", (int)sizeof(path), path)) path[0] = 0;
            hashpipe_status_unlock_safe(&st);
        }
    }
    FILE *f = path[0] ? fopen(path, "rb") : NULL;
    void *raw = batch > 0 ? malloc((size_t)batch * (pac8 ? (size_t)PAC8_BYTES : sizeof(vdif_in_packet_t))) : NULL;
    if (!f || !raw) {
        hashpipe_error("vdif_in_file_thread", "cannot read input file '%s' (%s)", path, f ? "no memory" : strerror(errno));
        state = STATE_ERROR;
    }
    vdif_in_packet_t *pk = NULL;
    long n_pk = 0, idx = 0;
    int64_t b_first = -1;
    int index_db_out = 0, start_copy = 0, first_copy = 1;

    while (run_threads()) {
        if (state == STATE_ERROR || state == STATE_DONE) {
            usleep(state == STATE_DONE ? 10000 : 1000000);
            continue;
        }
        if (n_pk == 0) {
            n_pk = receive(f, pac8, batch, &pk, raw);
            if (n_pk < 0) {
                hashpipe_error("vdif_in_file_thread", "error reading '%s'", path);
                n_pk = 0;
                state = STATE_ERROR;
                continue;
            }
            if (n_pk == 0) {   /* end of transmission */
                if (flush && b_first >= 0) {
                    for (int k = 0; k < BENG_GROUPS_IN_BUFFER; ++k) {
                        const int i = (index_db_out + k) % BENG_GROUPS_IN_BUFFER;
                        if (g_local.bgc[i].beng_group_vdif_packet_count == 0) break;
                        if (first_copy) {
                            fill_vdif_header_template(&g_local.bgc[i].vdif_header_template,
                                                      &g_in.cpu[i]->frames[0].vdif_packets[0], (int)cnt.skipped);
                            first_copy = 0;
                        }
                        const int rc = hand_off(args, db_out, i);
                        if (rc) {
                            state = rc < 0 ? STATE_ERROR : STATE_DONE;
                            break;
                        }
                        cnt.groups++;
                    }
                }
                cnt.done = 1;
                publish(args, &cnt);
                if (state != STATE_ERROR) state = STATE_DONE;
                continue;
            }
            idx = 0;
            cnt.all += n_pk;
            if (b_first == -1) {
                while (idx < n_pk && (b_first = get_packet_b_count(&pk[idx].header)) < 0) ++idx;
                if (b_first < 0) {   /* nothing valid yet (all invalid): receive again */
                    cnt.invalid += n_pk;
                    free(pk);
                    pk = NULL;
                    n_pk = 0;
                    continue;
                }
                for (int i = 0; i < BENG_GROUPS_IN_BUFFER; ++i)
                    init_beng_group(&g_local.bgc[i], g_in.cpu[i], g_in.gpu[i],
                                    b_first + 1 + (int64_t)BENG_FRAMES_PER_GROUP * i);
            }
        }
        while (idx < n_pk) {
            const int off = get_beng_group_index_offset(&g_local, index_db_out, &pk[idx]);
            if (off < 0) {
                if (off == vidErrorPacketInvalid) cnt.invalid++;
                if (off == vidErrorPacketBeforeStartTime) cnt.skipped++;
                if (off == vidErrorPacketTooFarAheadToCare) cnt.ahead++;
                ++idx;
                continue;
            }
            if (off > MAX_INDEX_LOOK_AHEAD) {
                start_copy = 1;
                if (first_copy) {
                    fill_vdif_header_template(&g_local.bgc[index_db_out].vdif_header_template, &pk[idx],
                                              (int)cnt.skipped);
                    first_copy = 0;
                }
            } else {
                cnt.used++;
                insert_vdif_in_beng_group_buffer(&g_local, index_db_out, off, &pk[idx]);
                ++idx;
                if (check_beng_group_complete(&g_local, index_db_out)) {
                    start_copy = 1;
                    if (first_copy) {
                        fill_vdif_header_template(&g_local.bgc[index_db_out].vdif_header_template,
                                                  &pk[idx < n_pk ? idx : idx - 1], (int)cnt.skipped);
                        first_copy = 0;
                    }
                }
            }
            if (start_copy) break;
        }
        if (idx == n_pk) {
            free(pk);
            pk = NULL;
            n_pk = 0;
        }
        if (start_copy) {
            const int rc = hand_off(args, db_out, index_db_out);
            if (rc) {
                state = rc < 0 ? STATE_ERROR : STATE_DONE;
                continue;
            }
            cnt.groups++;
            index_db_out = (index_db_out + 1) % BENG_GROUPS_IN_BUFFER;
            start_copy = 0;
            publish(args, &cnt);
        }
    }
    publish(args, &cnt);
    free(pk);
    free(raw);
    if (f) fclose(f);
    return NULL;
}

static hashpipe_thread_desc_t vdif_in_file_thread = {
    .name = "vdif_in_file_thread",
    .skey = "VDIFIN",
    .init = init_method,
    .run = run_method,
    .ibuf_desc = {NULL},
    .obuf_desc = {vdif_in_databuf_create},
};

static __attribute__((constructor)) void ctor(void) { register_hashpipe_thread(&vdif_in_file_thread); }
