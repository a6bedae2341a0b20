/*
 * Minimal stand-in for hashpipe's thread-plugin API (see hashpipe_databuf.h).
 */
#ifndef SDBE_HASHPIPE_COMPAT_H
#define SDBE_HASHPIPE_COMPAT_H

#include <pthread.h>

#include "hashpipe_databuf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hashpipe_status {
    int instance_id;
    int shmid;
    void *lock;
    char *buf;
} hashpipe_status_t;

typedef struct hashpipe_thread_args hashpipe_thread_args_t;
typedef int (*initfunc_t)(hashpipe_thread_args_t *);
typedef void *(*runfunc_t)(hashpipe_thread_args_t *);
typedef hashpipe_databuf_t *(*databuf_createfunc_t)(int, int);

typedef struct {
    databuf_createfunc_t create;
} databuf_desc_t;

typedef struct {
    const char *name;
    const char *skey;
    initfunc_t init;
    runfunc_t run;
    databuf_desc_t ibuf_desc;
    databuf_desc_t obuf_desc;
} hashpipe_thread_desc_t;

struct hashpipe_thread_args {
    hashpipe_thread_desc_t *thread_desc;
    int instance_id;
    int input_buffer;
    int output_buffer;
    unsigned int cpu_mask;
    int finished;
    pthread_cond_t finished_c;
    pthread_mutex_t finished_m;
    hashpipe_status_t st;
    hashpipe_databuf_t *ibuf;
    hashpipe_databuf_t *obuf;
    void *user_data;
};

/* Status buffer: FITS-style 80-character records "KEYWORD = value", ended by
 * an END record, as hashpipe's shared status memory; the lock serialises
 * readers and writers of one process.                                      */
int hashpipe_status_lock(hashpipe_status_t *s);
int hashpipe_status_unlock(hashpipe_status_t *s);
#define hashpipe_status_lock_safe(s)                                         \
    pthread_cleanup_push((void (*)(void *))hashpipe_status_unlock, (void *)(s)); \
    hashpipe_status_lock(s);
#define hashpipe_status_unlock_safe(s) \
    hashpipe_status_unlock(s);         \
    pthread_cleanup_pop(0);
/* hashpipe's fitshead accessors: 1 if the keyword was found, else 0.       */
int hgeti4(const char *hstring, const char *keyword, int *ival);
int hgets(const char *hstring, const char *keyword, int lstr, char *string);
int hputi4(char *hstring, const char *keyword, int ival);
int hputs(char *hstring, const char *keyword, const char *cval);

int register_hashpipe_thread(hashpipe_thread_desc_t *ptm);
hashpipe_thread_desc_t *find_hashpipe_thread(const char *name);
int run_threads(void);
void hashpipe_error(const char *name, const char *msg, ...);
void hashpipe_warn(const char *name, const char *msg, ...);
void hashpipe_info(const char *name, const char *msg, ...);

#ifdef __cplusplus
}
#endif
#endif
