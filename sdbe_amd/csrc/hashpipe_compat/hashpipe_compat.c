/*
 * In-process implementation of the minimal hashpipe API declared in
 * hashpipe.h / hashpipe_databuf.h, plus a harness that runs one registered
 * thread plugin between an input and an output databuf.  Test support only:
 * production loads the plugin into the real hashpipe.
 *
 * Block states: 0 free, 1 filled.  wait_* time out after 250 ms with
 * HASHPIPE_TIMEOUT so that plugins re-check run_threads(), as with hashpipe.
 */
#include <errno.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "hashpipe.h"

#define MAX_DATABUFS 16
#define MAX_THREADS 16

typedef struct {
    hashpipe_databuf_t *d;
    pthread_mutex_t m;
    pthread_cond_t c;
    int *state;
} databuf_slot_t;

static databuf_slot_t g_bufs[MAX_DATABUFS];
static int g_nbufs = 0;
static pthread_mutex_t g_reg = PTHREAD_MUTEX_INITIALIZER;
static hashpipe_thread_desc_t *g_threads[MAX_THREADS];
static int g_nthreads = 0;
static volatile int g_run = 1;

static databuf_slot_t *slot_of(hashpipe_databuf_t *d) {
    for (int i = 0; i < g_nbufs; ++i)
        if (g_bufs[i].d == d) return &g_bufs[i];
    return NULL;
}

hashpipe_databuf_t *hashpipe_databuf_create(int instance_id, int databuf_id, size_t header_size,
                                            size_t block_size, int n_block) {
    (void)instance_id;
    pthread_mutex_lock(&g_reg);
    if (g_nbufs >= MAX_DATABUFS || n_block <= 0) {
        pthread_mutex_unlock(&g_reg);
        return NULL;
    }
    hashpipe_databuf_t *d = (hashpipe_databuf_t *)calloc(1, header_size + block_size * (size_t)n_block);
    if (!d) {
        pthread_mutex_unlock(&g_reg);
        return NULL;
    }
    snprintf(d->data_type, sizeof(d->data_type), "compat-%d", databuf_id);
    d->header_size = header_size;
    d->block_size = block_size;
    d->n_block = n_block;
    databuf_slot_t *s = &g_bufs[g_nbufs++];
    s->d = d;
    pthread_mutex_init(&s->m, NULL);
    pthread_cond_init(&s->c, NULL);
    s->state = (int *)calloc((size_t)n_block, sizeof(int));
    pthread_mutex_unlock(&g_reg);
    return d;
}

void hashpipe_databuf_destroy(hashpipe_databuf_t *d) {
    pthread_mutex_lock(&g_reg);
    for (int i = 0; i < g_nbufs; ++i) {
        if (g_bufs[i].d != d) continue;
        pthread_mutex_destroy(&g_bufs[i].m);
        pthread_cond_destroy(&g_bufs[i].c);
        free(g_bufs[i].state);
        free(d);
        g_bufs[i] = g_bufs[--g_nbufs];
        break;
    }
    pthread_mutex_unlock(&g_reg);
}

static int wait_state(hashpipe_databuf_t *d, int block_id, int want) {
    databuf_slot_t *s = slot_of(d);
    if (!s || block_id < 0 || block_id >= d->n_block) return HASHPIPE_ERR_PARAM;
    struct timespec ts;
    clock_gettime(CLOCK_REALTIME, &ts);
    ts.tv_nsec += 250 * 1000 * 1000;
    if (ts.tv_nsec >= 1000000000L) {
        ts.tv_sec += 1;
        ts.tv_nsec -= 1000000000L;
    }
    int rv = HASHPIPE_OK;
    pthread_mutex_lock(&s->m);
    while (s->state[block_id] != want) {
        if (pthread_cond_timedwait(&s->c, &s->m, &ts) == ETIMEDOUT) {
            rv = s->state[block_id] == want ? HASHPIPE_OK : HASHPIPE_TIMEOUT;
            break;
        }
    }
    pthread_mutex_unlock(&s->m);
    return rv;
}

static int set_state(hashpipe_databuf_t *d, int block_id, int v) {
    databuf_slot_t *s = slot_of(d);
    if (!s || block_id < 0 || block_id >= d->n_block) return HASHPIPE_ERR_PARAM;
    pthread_mutex_lock(&s->m);
    s->state[block_id] = v;
    pthread_cond_broadcast(&s->c);
    pthread_mutex_unlock(&s->m);
    return HASHPIPE_OK;
}

int hashpipe_databuf_wait_free(hashpipe_databuf_t *d, int b) { return wait_state(d, b, 0); }
int hashpipe_databuf_wait_filled(hashpipe_databuf_t *d, int b) { return wait_state(d, b, 1); }
int hashpipe_databuf_set_free(hashpipe_databuf_t *d, int b) { return set_state(d, b, 0); }
int hashpipe_databuf_set_filled(hashpipe_databuf_t *d, int b) { return set_state(d, b, 1); }

int hashpipe_databuf_block_status(hashpipe_databuf_t *d, int b) {
    databuf_slot_t *s = slot_of(d);
    if (!s || b < 0 || b >= d->n_block) return HASHPIPE_ERR_PARAM;
    pthread_mutex_lock(&s->m);
    const int v = s->state[b];
    pthread_mutex_unlock(&s->m);
    return v;
}

int register_hashpipe_thread(hashpipe_thread_desc_t *ptm) {
    pthread_mutex_lock(&g_reg);
    int rv = -1;
    if (g_nthreads < MAX_THREADS) {
        g_threads[g_nthreads++] = ptm;
        rv = 0;
    }
    pthread_mutex_unlock(&g_reg);
    return rv;
}

hashpipe_thread_desc_t *find_hashpipe_thread(const char *name) {
    for (int i = 0; i < g_nthreads; ++i)
        if (strcmp(g_threads[i]->name, name) == 0) return g_threads[i];
    return NULL;
}

int run_threads(void) { return g_run; }

static void vlog(const char *lvl, const char *name, const char *msg, va_list ap) {
    fprintf(stderr, "[%s] %s: ", lvl, name);
    vfprintf(stderr, msg, ap);
    fputc('\n', stderr);
}
void hashpipe_error(const char *name, const char *msg, ...) {
    va_list ap;
    va_start(ap, msg);
    vlog("error", name, msg, ap);
    va_end(ap);
}
void hashpipe_warn(const char *name, const char *msg, ...) {
    va_list ap;
    va_start(ap, msg);
    vlog("warn", name, msg, ap);
    va_end(ap);
}
void hashpipe_info(const char *name, const char *msg, ...) {
    va_list ap;
    va_start(ap, msg);
    vlog("info", name, msg, ap);
    va_end(ap);
}

/* ------------------------------------------------------------------ harness */
/* Run thread plugin `name` between fresh input and output databufs.  The
 * caller fills input blocks through the returned pointers:
 *   hpc_start(name)                -> starts the plugin thread
 *   hpc_in_block(i) / hpc_out_block(j) -> block memory
 *   hpc_in_databuf() / hpc_out_databuf() -> the two databufs
 *   hpc_set_in_filled(i), hpc_wait_out_filled(j, timeout_ms), hpc_set_out_free(j)
 *   hpc_stop()                     -> run_threads() turns false, thread joined */
static hashpipe_thread_args_t g_args;
static pthread_t g_thr;
static int g_started = 0;

static void *thread_main(void *a) {
    hashpipe_thread_args_t *args = (hashpipe_thread_args_t *)a;
    args->thread_desc->run(args);
    return NULL;
}

int hpc_start(const char *name) {
    if (g_started) return -1;
    hashpipe_thread_desc_t *td = find_hashpipe_thread(name);
    if (!td) return -2;
    memset(&g_args, 0, sizeof(g_args));
    g_args.thread_desc = td;
    g_args.ibuf = td->ibuf_desc.create ? td->ibuf_desc.create(0, 1) : NULL;
    g_args.obuf = td->obuf_desc.create ? td->obuf_desc.create(0, 2) : NULL;
    if (!g_args.ibuf || !g_args.obuf) return -3;
    if (td->init && td->init(&g_args) != 0) return -4;
    g_run = 1;
    if (pthread_create(&g_thr, NULL, thread_main, &g_args) != 0) return -5;
    g_started = 1;
    return 0;
}

void *hpc_in_block(int i) {
    hashpipe_databuf_t *d = g_args.ibuf;
    return d ? (char *)d + d->header_size + d->block_size * (size_t)i : NULL;
}

void *hpc_out_block(int j) {
    hashpipe_databuf_t *d = g_args.obuf;
    return d ? (char *)d + d->header_size + d->block_size * (size_t)j : NULL;
}

void *hpc_in_databuf(void) { return g_args.ibuf; }
void *hpc_out_databuf(void) { return g_args.obuf; }
int hpc_set_in_filled(int i) { return hashpipe_databuf_set_filled(g_args.ibuf, i); }
int hpc_in_status(int i) { return hashpipe_databuf_block_status(g_args.ibuf, i); }
int hpc_set_out_free(int j) { return hashpipe_databuf_set_free(g_args.obuf, j); }

int hpc_wait_out_filled(int j, int timeout_ms) {
    for (int waited = 0; waited < timeout_ms; waited += 250) {
        const int rv = hashpipe_databuf_wait_filled(g_args.obuf, j);
        if (rv != HASHPIPE_TIMEOUT) return rv;
    }
    return HASHPIPE_TIMEOUT;
}

int hpc_stop(void) {
    if (!g_started) return -1;
    g_run = 0;
    pthread_join(g_thr, NULL);
    hashpipe_databuf_destroy(g_args.ibuf);
    hashpipe_databuf_destroy(g_args.obuf);
    g_started = 0;
    g_run = 1;
    return 0;
}
