/*
 * Minimal stand-in for hashpipe's databuf API, used to build and test the
 * vdif_inout_gpu_thread plugin where hashpipe is not installed.  Production
 * builds compile the plugin against the real hashpipe headers instead.
 * Blocks live in process memory (not SysV shared memory); the block states
 * are guarded by one mutex and condition variable per databuf.
 */
#ifndef SDBE_HASHPIPE_COMPAT_DATABUF_H
#define SDBE_HASHPIPE_COMPAT_DATABUF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HASHPIPE_OK          0
#define HASHPIPE_TIMEOUT     1
#define HASHPIPE_ERR_GEN    -1
#define HASHPIPE_ERR_SYS    -2
#define HASHPIPE_ERR_PARAM  -3

/* Header at the start of every databuf (hashpipe's public layout). */
typedef struct {
    char data_type[64];
    size_t header_size;
    size_t block_size;
    int n_block;
    int shmid;
    int semid;
} hashpipe_databuf_t;

hashpipe_databuf_t *hashpipe_databuf_create(int instance_id, int databuf_id, size_t header_size,
                                            size_t block_size, int n_block);
int hashpipe_databuf_wait_free(hashpipe_databuf_t *d, int block_id);
int hashpipe_databuf_wait_filled(hashpipe_databuf_t *d, int block_id);
int hashpipe_databuf_set_free(hashpipe_databuf_t *d, int block_id);
int hashpipe_databuf_set_filled(hashpipe_databuf_t *d, int block_id);
int hashpipe_databuf_block_status(hashpipe_databuf_t *d, int block_id);
void hashpipe_databuf_destroy(hashpipe_databuf_t *d);

#ifdef __cplusplus
}
#endif
#endif
