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

/* ------------------------------------------------------------ status buffer */
#define STATUS_RECORDS 256
static char g_status_buf[STATUS_RECORDS * 80 + 1];
static pthread_mutex_t g_status_mu = PTHREAD_MUTEX_INITIALIZER;
static int g_status_init = 0;

static void status_init(void) {
    if (g_status_init) return;
    memset(g_status_buf, ' ', sizeof(g_status_buf) - 1);
    memcpy(g_status_buf, "END", 3);
    g_status_buf[sizeof(g_status_buf) - 1] = 0;
    g_status_init = 1;
}

int hashpipe_status_lock(hashpipe_status_t *s) {
    (void)s;
    return pthread_mutex_lock(&g_status_mu);
}
int hashpipe_status_unlock(hashpipe_status_t *s) {
    (void)s;
    return pthread_mutex_unlock(&g_status_mu);
}

static int is_end(const char *r) { return r[0] == 'E' && r[1] == 'N' && r[2] == 'D' && (r[3] == ' ' || !r[3]); }

/* record of `keyword` (8-character, blank-padded key field) or NULL */
static char *status_find(const char *h, const char *keyword) {
    char key[9];
    snprintf(key, sizeof(key), "%-8.8s", keyword);
    for (const char *r = h; r[0] && !is_end(r); r += 80)
        if (strncmp(r, key, 8) == 0 && r[8] == '=') return (char *)r;
    return NULL;
}

static char *status_end(char *h) {
    char *r = h;
    while (r[0] && !is_end(r)) r += 80;
    return r;
}

static int status_put(char *h, const char *keyword, const char *value) {
    char *r = status_find(h, keyword);
    if (!r) {
        r = status_end(h);
        if (r + 160 > g_status_buf + sizeof(g_status_buf) - 1) return -1;
        memset(r + 80, ' ', 80);
        memcpy(r + 80, "END", 3);
    }
    char rec[81];
    snprintf(rec, sizeof(rec), "%-8.8s= %-70.70s", keyword, value);
    memcpy(r, rec, 80);
    return 0;
}

int hgeti4(const char *hstring, const char *keyword, int *ival) {
    const char *r = status_find(hstring, keyword);
    if (!r) return 0;
    char v[72];
    memcpy(v, r + 10, 70);
    v[70] = 0;
    char *end = NULL;
    long x = strtol(v, &end, 10);
    if (end == v) return 0;
    *ival = (int)x;
    return 1;
}

int hgets(const char *hstring, const char *keyword, int lstr, char *string) {
    const char *r = status_find(hstring, keyword);
    if (!r || lstr < 1) return 0;
    const char *v = r + 10;
    int n = 70;
    if (*v == '\'') {   /* quoted string value */
        ++v;
        const char *q = memchr(v, '\'', 69);
        n = q ? (int)(q - v) : 69;
    }
    while (n > 0 && v[n - 1] == ' ') --n;
    if (n > lstr - 1) n = lstr - 1;
    memcpy(string, v, (size_t)n);
    string[n] = 0;
    return 1;
}

int hputi4(char *hstring, const char *keyword, int ival) {
    char v[32];
    snprintf(v, sizeof(v), "%20d", ival);
    return status_put(hstring, keyword, v);
}

int hputs(char *hstring, const char *keyword, const char *cval) {
    char v[72];
    snprintf(v, sizeof(v), "'%-.67s'", cval);
    return status_put(hstring, keyword, v);
}

/* Harness access to the status buffer (set keys before starting threads,
 * read the ones threads publish). */
int hpc_status_puti4(const char *keyword, int ival) {
    status_init();
    pthread_mutex_lock(&g_status_mu);
    const int rc = hputi4(g_status_buf, keyword, ival);
    pthread_mutex_unlock(&g_status_mu);
    return rc;
}
int hpc_status_puts(const char *keyword, const char *cval) {
    status_init();
    pthread_mutex_lock(&g_status_mu);
    const int rc = hputs(g_status_buf, keyword, cval);
    pthread_mutex_unlock(&g_status_mu);
    return rc;
}
int hpc_status_geti4(const char *keyword, int *ival) {
    status_init();
    pthread_mutex_lock(&g_status_mu);
    const int rc = hgeti4(g_status_buf, keyword, ival);
    pthread_mutex_unlock(&g_status_mu);
    return rc;
}
int hpc_status_clear(void) {
    pthread_mutex_lock(&g_status_mu);
    g_status_init = 0;
    status_init();
    pthread_mutex_unlock(&g_status_mu);
    return 0;
}

/* ------------------------------------------------------------------ harness */
/* Run registered thread plugins the way hashpipe runs a pipeline:
 *   hpc_start(name)                -> one plugin between a fresh input and a
 *                                     fresh output databuf
 *   hpc_pipeline_start(names, n)   -> n plugins in a chain, databuf k between
 *                                     plugin k (output) and plugin k + 1
 *                                     (input), created by plugin k's output
 *                                     constructor (else k + 1's input one)
 *   hpc_in_block(i) / hpc_out_block(j) -> block memory of the first plugin's
 *                                     input / the last plugin's output databuf
 *   hpc_in_databuf() / hpc_out_databuf() -> those two databufs
 *   hpc_set_in_filled(i), hpc_wait_out_filled(j, timeout_ms), hpc_set_out_free(j)
 *   hpc_thread_done(k)             -> 1 once plugin k's run function returned
 *   hpc_stop()                     -> run_threads() turns false, threads joined
 * Every plugin sees the one status buffer (args->st.buf).                  */
#define MAX_PIPE 8
static hashpipe_thread_args_t g_args[MAX_PIPE];
static pthread_t g_thr[MAX_PIPE];
static volatile int g_done[MAX_PIPE];
static hashpipe_databuf_t *g_links[MAX_PIPE + 1];
static int g_nstage = 0;
static int g_started = 0;

static void *thread_main(void *a) {
    hashpipe_thread_args_t *args = (hashpipe_thread_args_t *)a;
    args->thread_desc->run(args);
    g_done[args - g_args] = 1;
    return NULL;
}

static void destroy_links(void) {
    for (int k = 0; k <= MAX_PIPE; ++k) {
        if (g_links[k]) hashpipe_databuf_destroy(g_links[k]);
        g_links[k] = NULL;
    }
}

int hpc_pipeline_start(const char *const *names, int n) {
    if (g_started) return -1;
    if (n < 1 || n > MAX_PIPE) return -6;
    hashpipe_thread_desc_t *td[MAX_PIPE];
    for (int k = 0; k < n; ++k) {
        td[k] = find_hashpipe_thread(names[k]);
        if (!td[k]) return -2;
    }
    status_init();
    memset(g_args, 0, sizeof(g_args));
    memset(g_links, 0, sizeof(g_links));
    /* link 0: the first plugin's input; link n: the last plugin's output */
    for (int k = 0; k <= n; ++k) {
        databuf_createfunc_t f = NULL;
        if (k > 0 && td[k - 1]->obuf_desc.create) f = td[k - 1]->obuf_desc.create;
        else if (k < n && td[k]->ibuf_desc.create) f = td[k]->ibuf_desc.create;
        if (f) {
            g_links[k] = f(0, k + 1);
            if (!g_links[k]) {
                destroy_links();
                return -3;
            }
        }
    }
    for (int k = 0; k < n; ++k) {
        g_args[k].thread_desc = td[k];
        g_args[k].instance_id = 0;
        g_args[k].input_buffer = k;
        g_args[k].output_buffer = k + 1;
        g_args[k].ibuf = td[k]->ibuf_desc.create ? g_links[k] : NULL;
        g_args[k].obuf = td[k]->obuf_desc.create ? g_links[k + 1] : NULL;
        g_args[k].st.buf = g_status_buf;
        g_done[k] = 0;
    }
    for (int k = 0; k < n; ++k)
        if (td[k]->init && td[k]->init(&g_args[k]) != 0) {
            destroy_links();
            return -4;
        }
    g_run = 1;
    for (int k = 0; k < n; ++k) {
        if (pthread_create(&g_thr[k], NULL, thread_main, &g_args[k]) != 0) {
            g_run = 0;
            for (int j = 0; j < k; ++j) pthread_join(g_thr[j], NULL);
            g_run = 1;
            destroy_links();
            return -5;
        }
    }
    g_nstage = n;
    g_started = 1;
    return 0;
}

int hpc_start(const char *name) {
    const char *names[1] = {name};
    return hpc_pipeline_start(names, 1);
}

static void *block_of(hashpipe_databuf_t *d, int i) {
    return d ? (char *)d + d->header_size + d->block_size * (size_t)i : NULL;
}

void *hpc_in_block(int i) { return block_of(g_links[0], i); }
void *hpc_out_block(int j) { return block_of(g_links[g_nstage], j); }
void *hpc_in_databuf(void) { return g_links[0]; }
void *hpc_out_databuf(void) { return g_links[g_nstage]; }
int hpc_set_in_filled(int i) { return hashpipe_databuf_set_filled(g_links[0], i); }
int hpc_in_status(int i) { return hashpipe_databuf_block_status(g_links[0], i); }
int hpc_set_out_free(int j) { return hashpipe_databuf_set_free(g_links[g_nstage], j); }
int hpc_thread_done(int k) { return (k >= 0 && k < g_nstage) ? g_done[k] : -1; }

int hpc_wait_out_filled(int j, int timeout_ms) {
    for (int waited = 0; waited < timeout_ms; waited += 250) {
        const int rv = hashpipe_databuf_wait_filled(g_links[g_nstage], j);
        if (rv != HASHPIPE_TIMEOUT) return rv;
    }
    return HASHPIPE_TIMEOUT;
}

int hpc_stop(void) {
    if (!g_started) return -1;
    g_run = 0;
    for (int k = 0; k < g_nstage; ++k) pthread_join(g_thr[k], NULL);
    destroy_links();
    g_nstage = 0;
    g_started = 0;
    g_run = 1;
    return 0;
}
