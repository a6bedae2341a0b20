/*
 * aphids_databuf.h -- the hashpipe databuf types the GPU stage exchanges with
 * the APHIDS input and output threads, with byte-identical layouts to the
 * reference declarations at the SWARM rate preset SWARM_RATE_FRAC (default 6:
 * the 6/11 rate, 39 B-frames per group):
 *   input side   software/aphids/vdif_in_databuf.h:15-152
 *   output side  software/aphids/vdif_out_databuf.h:11-144
 *   rate presets software/aphids/sample_rates.h:5-74
 * The layouts are an interface (the other threads of aphids.so share them),
 * so names, field order and widths follow the reference; every size and
 * offset is pinned by a static assertion below.
 *
 * Also declared: the host-side group bookkeeping, header template and
 * B-engine timestamp reform the input thread performs before the GPU stage
 * (vdif_in_databuf.c:56-220, beng_reform.c:16-260, vdif_8pac.c:5-56),
 * implemented in sdbe_amd/csrc/aphids_host.c.
 */
#ifndef SDBE_APHIDS_DATABUF_H
#define SDBE_APHIDS_DATABUF_H

#include <stddef.h>
#include <stdint.h>

#include "hashpipe.h"
#include "hashpipe_databuf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ rate preset */
/* SWARM_RATE_FRAC selects the group size (BENG_FRAMES_PER_GROUP =
 * DECIMATION_FACTOR B-frames) and the output slots per polarisation, as
 * sample_rates.h does.  The reference compiles 11/11 by default
 * (sample_rates.h:12-14); here the default is the 6/11 rate of the 64/39
 * resampler.                                                               */
#ifndef SWARM_RATE_FRAC
#define SWARM_RATE_FRAC 6
#endif
#if SWARM_RATE_FRAC == 6
#define DECIMATION_FACTOR 39
#define EXPANSION_FACTOR 32
#define VDIF_OUT_PKTS_PER_BLOCK 8192
#elif SWARM_RATE_FRAC == 8
#define DECIMATION_FACTOR 13
#define EXPANSION_FACTOR 8
#define VDIF_OUT_PKTS_PER_BLOCK 2048
#elif SWARM_RATE_FRAC == 10
#define DECIMATION_FACTOR 65
#define EXPANSION_FACTOR 64
#define VDIF_OUT_PKTS_PER_BLOCK 8192
#elif SWARM_RATE_FRAC == 11
#define DECIMATION_FACTOR 143
#define EXPANSION_FACTOR 128
#define VDIF_OUT_PKTS_PER_BLOCK 16384
#else
#error "SWARM_RATE_FRAC must be 6, 8, 10 or 11"
#endif

/* -------------------------------------------------------------- constants */
#define BENG_FRAMES_PER_GROUP        DECIMATION_FACTOR
#define VDIF_PER_BENG_FRAME          2048
#define VDIF_IN_PKT_HEADER_SIZE      32
#define VDIF_IN_PKT_DATA_SIZE        1024
#define FID_PER_CHAN                 8
#define CHAN_PER_BENG                256
#define BENG_GROUPS_IN_BUFFER        4
#define VDIF_OUT_PKT_HEADER_SIZE     32
#define VDIF_OUT_PKT_DATA_SIZE       8192
#define VDIF_CHAN                    2
#define VDIF_OUT_BUFFER_SIZE         4
#define VDIF_OUT_FRAMES_PER_SECOND   125000
#define MAGIC_OFFSET_IN_BENG_FFT_WINDOWS 52
#define SWARM_ZERO_DELAY_OFFSET_PICOSECONDS 3715791

typedef enum {
    vidErrorPacketInvalid = -1,
    vidErrorPacketBeforeStartTime = -2,
    vidErrorPacketTooFarAheadToCare = -3
} vdif_in_databuf_error_t;

/* ------------------------------------------------------------- input side */
typedef struct vdif_in_header {
    struct word0_in { uint32_t secs_inre:30; uint32_t legacy:1; uint32_t invalid:1; } w0;
    struct word1_in { uint32_t df_num_insec:24; uint32_t ref_epoch:6; uint32_t UA:2; } w1;
    struct word2_in { uint32_t df_len:24; uint32_t num_channels:5; uint32_t ver:3; } w2;
    struct word3_in { uint32_t stationID:16; uint32_t threadID:10; uint32_t bps:5; uint32_t dt:1; } w3;
    struct beng_hdr { uint32_t b_upper; uint8_t c; uint8_t z; uint8_t f; uint8_t b_lower; } beng;
    uint64_t edh_psn;
} vdif_in_header_t;

typedef struct vdif_in_packet {
    vdif_in_header_t header;
    char data[VDIF_IN_PKT_DATA_SIZE];
} vdif_in_packet_t;

typedef struct beng_frame_vdif_buffer {
    vdif_in_packet_t vdif_packets[VDIF_PER_BENG_FRAME];
} beng_frame_vdif_buffer_t;

typedef struct beng_group_vdif_buffer {
    beng_frame_vdif_buffer_t frames[BENG_FRAMES_PER_GROUP];
} beng_group_vdif_buffer_t;

typedef struct channel_completion {
    uint8_t f0:1; uint8_t f1:1; uint8_t f2:1; uint8_t f3:1;
    uint8_t f4:1; uint8_t f5:1; uint8_t f6:1; uint8_t f7:1;
} channel_completion_t;

typedef struct beng_frame_completion {
    uint64_t b;                                 /* B-engine counter value      */
    uint16_t beng_frame_vdif_packet_count;      /* packets found for it        */
    channel_completion_t cc[CHAN_PER_BENG];     /* bit f of cc[c]: (c, f) found */
} beng_frame_completion_t;

typedef struct beng_group_completion {
    int32_t beng_group_vdif_packet_count;
    beng_group_vdif_buffer_t *bgv_buf_cpu;
    beng_group_vdif_buffer_t *bgv_buf_gpu;      /* device copy the GPU stage reads */
    int gpu_id;
    void *memcpy_stream;
    beng_frame_completion_t bfc[BENG_FRAMES_PER_GROUP];
    vdif_in_header_t vdif_header_template;
} beng_group_completion_t;

typedef struct vdif_in_databuf {
    hashpipe_databuf_t header;
    beng_group_completion_t bgc[BENG_GROUPS_IN_BUFFER];
} vdif_in_databuf_t;

/* ------------------------------------------------------------ output side */
typedef struct vdif_out_header {
    struct word0_out { uint32_t secs_inre:30; uint32_t legacy:1; uint32_t invalid:1; } w0;
    struct word1_out { uint32_t df_num_insec:24; uint32_t ref_epoch:6; uint32_t UA:2; } w1;
    struct word2_out { uint32_t df_len:24; uint32_t num_channels:5; uint32_t ver:3; } w2;
    struct word3_out { uint32_t stationID:16; uint32_t threadID:10; uint32_t bps:5; uint32_t dt:1; } w3;
    struct word4_out {
        uint32_t pol:1; uint32_t bdc_sideband:1; uint32_t rx_sideband:1; uint32_t eud5:21; uint32_t edv5:8;
    } w4;
    struct word5_out { uint32_t PICstatus; } w5;
    uint64_t edh_psn;
} vdif_out_header_t;

typedef struct vdif_out_data { char data[VDIF_OUT_PKT_DATA_SIZE]; } vdif_out_data_t;
typedef struct vdif_out_packet { vdif_out_header_t header; vdif_out_data_t data; } vdif_out_packet_t;
typedef struct vdif_out_packet_block { vdif_out_packet_t packets[VDIF_OUT_PKTS_PER_BLOCK]; } vdif_out_packet_block_t;
typedef struct vdif_out_packet_group { vdif_out_packet_block_t chan[VDIF_CHAN]; } vdif_out_packet_group_t;
typedef struct vdif_out_data_block { vdif_out_data_t datas[VDIF_OUT_PKTS_PER_BLOCK]; } vdif_out_data_block_t;
typedef struct vdif_out_data_group { vdif_out_data_block_t chan[VDIF_CHAN]; } vdif_out_data_group_t;

typedef struct quantized_storage {
    vdif_out_data_group_t *vdg_buf_cpu;
    vdif_out_data_group_t *vdg_buf_gpu;         /* device output of the GPU stage */
    int bit_depth;
    int N_32bit_words_per_chan;
    int gpu_id;
    void *memcpy_stream;
    vdif_out_header_t vdif_header_template;
} quantized_storage_t;

typedef struct vdif_out_databuf {
    hashpipe_databuf_t header;
    quantized_storage_t blocks[VDIF_OUT_BUFFER_SIZE];
} vdif_out_databuf_t;

/* Layout pins (LP64; sizes worked out from the reference declarations). */
#define SDBE_LAYOUT_ASSERT(c, m) _Static_assert(c, m)
#ifdef __cplusplus
#undef SDBE_LAYOUT_ASSERT
#define SDBE_LAYOUT_ASSERT(c, m) static_assert(c, m)
#endif
SDBE_LAYOUT_ASSERT(sizeof(vdif_in_header_t) == 32, "vdif_in_header_t");
SDBE_LAYOUT_ASSERT(offsetof(vdif_in_header_t, beng) == 16 && offsetof(vdif_in_header_t, edh_psn) == 24, "in hdr");
SDBE_LAYOUT_ASSERT(sizeof(vdif_in_packet_t) == 1056, "vdif_in_packet_t");
SDBE_LAYOUT_ASSERT(sizeof(beng_group_vdif_buffer_t) == (size_t)BENG_FRAMES_PER_GROUP * 2048 * 1056,
                   "group buffer = frames * 2048 * 1056 (84,344,832 at 6/11)");
SDBE_LAYOUT_ASSERT(sizeof(channel_completion_t) == 1, "channel_completion_t");
SDBE_LAYOUT_ASSERT(sizeof(beng_frame_completion_t) == 272, "beng_frame_completion_t = 8+2+256 -> 272");
SDBE_LAYOUT_ASSERT(offsetof(beng_frame_completion_t, cc) == 10, "cc offset");
SDBE_LAYOUT_ASSERT(offsetof(beng_group_completion_t, bgv_buf_cpu) == 8, "bgc.bgv_buf_cpu");
SDBE_LAYOUT_ASSERT(offsetof(beng_group_completion_t, gpu_id) == 24, "bgc.gpu_id");
SDBE_LAYOUT_ASSERT(offsetof(beng_group_completion_t, bfc) == 40, "bgc.bfc");
SDBE_LAYOUT_ASSERT(offsetof(beng_group_completion_t, vdif_header_template) == 40 + BENG_FRAMES_PER_GROUP * 272,
                   "bgc.template");
SDBE_LAYOUT_ASSERT(sizeof(beng_group_completion_t) == 72 + BENG_FRAMES_PER_GROUP * 272,
                   "beng_group_completion_t (10680 at 6/11)");
SDBE_LAYOUT_ASSERT(sizeof(vdif_out_header_t) == 32, "vdif_out_header_t");
SDBE_LAYOUT_ASSERT(sizeof(vdif_out_packet_t) == 8224, "vdif_out_packet_t");
SDBE_LAYOUT_ASSERT(sizeof(vdif_out_data_group_t) == 2ul * VDIF_OUT_PKTS_PER_BLOCK * 8192, "vdif_out_data_group_t");
SDBE_LAYOUT_ASSERT(sizeof(vdif_out_packet_group_t) == 2ul * VDIF_OUT_PKTS_PER_BLOCK * 8224, "vdif_out_packet_group_t");
SDBE_LAYOUT_ASSERT(offsetof(quantized_storage_t, bit_depth) == 16, "qs.bit_depth");
SDBE_LAYOUT_ASSERT(offsetof(quantized_storage_t, memcpy_stream) == 32, "qs.memcpy_stream");
SDBE_LAYOUT_ASSERT(sizeof(quantized_storage_t) == 72, "quantized_storage_t");
SDBE_LAYOUT_ASSERT(offsetof(vdif_in_databuf_t, bgc) == sizeof(hashpipe_databuf_t), "in databuf blocks");
SDBE_LAYOUT_ASSERT(offsetof(vdif_out_databuf_t, blocks) == sizeof(hashpipe_databuf_t), "out databuf blocks");

/* --------------------------------------------- databuf constructors (hashpipe) */
hashpipe_databuf_t *vdif_in_databuf_create(int instance_id, int databuf_id);
hashpipe_databuf_t *vdif_out_databuf_create(int instance_id, int databuf_id);

/* ------------------------------------ host bookkeeping (vdif_in_databuf.c) */
int64_t get_packet_b_count(vdif_in_header_t *vdif_pkt_hdr);
void init_beng_group(beng_group_completion_t *bgc, beng_group_vdif_buffer_t *bgv_buf_cpu,
                     beng_group_vdif_buffer_t *bgv_buf_gpu, int64_t b_start);
int get_beng_group_index_offset(vdif_in_databuf_t *bgc_buf, int index_ref, vdif_in_packet_t *vdif_pkt);
int insert_vdif_in_beng_group_buffer(vdif_in_databuf_t *bgc_buf, int index_ref, int offset,
                                     vdif_in_packet_t *vdif_pkt);
int check_beng_group_complete(vdif_in_databuf_t *bgc_buf, int index);
void fill_vdif_header_template(vdif_in_header_t *vdif_hdr_copy, vdif_in_packet_t *vdif_pkt_ref, int n_skipped);

/* ---------------------------------- B-engine timestamps (beng_reform.c) */
typedef struct beng_timestamp { int32_t sec; int32_t clk; } beng_timestamp_t;
void get_beng_t0(beng_timestamp_t *t);
void get_beng_tprev(beng_timestamp_t *t);
void get_beng_bprev(int64_t *b);
void beng_reform_headers(vdif_in_packet_t *pkts, int npkts);
double beng_timestamp_clk_to_float(beng_timestamp_t *t);
void beng_timestamp_increment(beng_timestamp_t *t);
void beng_timestamp_decrement(beng_timestamp_t *t);
/* Forget T0 / TPREV / BPREV (start of a new stream).  No reference
 * counterpart: the reference keeps them for the life of the process.      */
void beng_reform_reset(void);

/* ----------------------------------------------- 8pac (vdif_8pac.c) */
#define VTP_BYTE_SIZE 8
typedef struct vtp_plus_pac { char vtp[VTP_BYTE_SIZE]; vdif_in_packet_t pac; } vtp_plus_pac_t;
typedef struct vdif_8pac_packet { vdif_in_packet_t pkt1; vtp_plus_pac_t pac7[7]; } vdif_8pac_packet_t;
SDBE_LAYOUT_ASSERT(sizeof(vdif_8pac_packet_t) == 1056 + 7 * 1064, "vdif_8pac_packet_t");
void unpack_8pac(vdif_in_packet_t *dest, vdif_8pac_packet_t *src, int n_src);

/* ------------------------------ device side of the input / output threads */
/* HIP implementations under the reference's link names
 * (software/aphids/vdif_in_databuf_cuda.cu:47-107,
 *  vdif_out_databuf_cuda.cu:48-93): pinned host group buffers, device group
 * buffers, and the asynchronous copies the input thread starts after a group
 * completes and the output thread starts for every filled output slot.  Each
 * copy runs on its own stream (default priority), kept in memcpy_stream until
 * the completion check destroys it.  All return 1 on success, -1 on failure.
 * The device of slot `index` is index % (number of HIP devices), as the
 * reference's index % NUM_GPU.                                            */
int get_bgv_cpu_memory_cuda(beng_group_vdif_buffer_t **bgv_buf_cpu, int index);
int get_bgv_gpu_memory_cuda(beng_group_vdif_buffer_t **bgv_buf_gpu, int index);
/* Copies only the packet_count packets the slot holds (the reference copies
 * the whole 84 MB slot); sets gpu_id and memcpy_stream.                    */
int transfer_beng_group_to_gpu_cuda(vdif_in_databuf_t *bgc_buf, int index);
/* 0 while the copy runs, 1 when done (the stream is destroyed then).       */
int check_transfer_beng_group_to_gpu_complete_cuda(vdif_in_databuf_t *bgc_buf, int index);
int get_vdg_cpu_memory_cuda(vdif_out_data_group_t **vdg_buf_cpu, int index);
int transfer_vdif_group_to_cpu_cuda(vdif_out_databuf_t *qs_buf, int index);
/* Waits for the copy, destroys its stream, returns 1.                      */
int check_transfer_vdif_group_to_cpu_complete_cuda(vdif_out_databuf_t *qs_buf, int index);

/* The reference's wrappers of the above (vdif_in_databuf.c:169-190,
 * vdif_out_databuf.c:16-33), for builds without the reference's files.     */
int get_bgv_cpu_memory(beng_group_vdif_buffer_t **bgv_buf_cpu, int index);
int get_bgv_gpu_memory(beng_group_vdif_buffer_t **bgv_buf_gpu, int index);
int transfer_beng_group_to_gpu(vdif_in_databuf_t *bgc_buf, int index);
int check_transfer_beng_group_to_gpu_complete(vdif_in_databuf_t *bgc_buf, int index);
int get_vdg_cpu_memory(vdif_out_data_group_t **vdg_buf_cpu, int index);
int transfer_vdif_group_to_cpu(vdif_out_databuf_t *qs_buf, int index);
int check_transfer_vdif_group_to_cpu_complete(vdif_out_databuf_t *qs_buf, int index);

#ifdef __cplusplus
}
#endif

#endif /* SDBE_APHIDS_DATABUF_H */
