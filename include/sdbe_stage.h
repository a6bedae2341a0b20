/*
 * sdbe_stage.h -- C-ABI of the MI355X (gfx950) SWARM data back-end GPU stage.
 *
 * The stage takes one group of SWARM B-engine-over-VDIF packets (39 B-frames,
 * up to 79,872 packets of 1056 bytes, any order, some missing) resident in
 * device memory and produces 2-bit VDIF output for both polarisations:
 * 128 blocks of 39 snapshots each, resampled 39 -> 64 (2496 -> 4096 Msps),
 * 64 frames of 8192 payload bytes per block and polarisation.
 *
 * Plain pointers and sizes only; device pointers are HIP device addresses,
 * `stream` is a hipStream_t passed as void* (NULL = the default stream).
 * Every call returns SDBE_OK (0) or a negative SDBE_ERR_* code and never
 * throws.  Calls on one handle must not overlap.
 *
 * Reference interfaces replaced (paths relative to the Smithsonian/sdbe tree):
 *   sdbe_create / sdbe_destroy  -> aphids_resampler_init / _destroy
 *                                  (software/aphids/vdif_inout_gpu_thread.cu:184-347, 712-743)
 *   sdbe_process_group          -> the per-slot body of run_method: memset +
 *                                  vdif_to_beng + 128 x [inflate, C2R, R2C, C2R,
 *                                  thresholds, quantize2bit]
 *                                  (vdif_inout_gpu_thread.cu:894-1008), plus the
 *                                  output thread's per-packet header stamping
 *                                  (software/aphids/vdif_out_net_thread.c:294-317)
 *   sdbe_unpack_group_int8      -> vdif_to_beng with its zero fill
 *                                  (vdif_inout_gpu_thread.cu:404-507, 897-906)
 *   sdbe_process_spectra        -> fft_batched + quantize on already-reordered
 *                                  spectra (software/prototyping/cuda/resample.py:67-146,
 *                                  vdif_inout_gpu_thread.cu:923-1008)
 *   sdbe_resample_spectra_f32   -> the three cuFFT stages alone
 *                                  (vdif_inout_gpu_thread.cu:589-709;
 *                                  resample.py:109-123), normalised like numpy
 *   sdbe_stream_state_from_template -> fill_vdif_header_template + the output
 *                                  thread's start alignment
 *                                  (software/aphids/vdif_in_databuf.c:193-220,
 *                                  vdif_out_net_thread.c:203-237)
 */
#ifndef SDBE_STAGE_H
#define SDBE_STAGE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------ */
#define SDBE_OK              0
#define SDBE_ERR_ARG        -1   /* bad argument (NULL, size out of range)   */
#define SDBE_ERR_HIP        -2   /* a HIP runtime call failed                */
#define SDBE_ERR_NOMEM      -3   /* device or host allocation failed         */
#define SDBE_ERR_CAPACITY   -4   /* request exceeds the handle's capacity    */
#define SDBE_ERR_NODEV      -5   /* no such HIP device                       */
#define SDBE_ERR_STATE      -6   /* call not valid in the current state      */

/* ---- geometry (6/11 SWARM rate, 64/39 resampling) ----------------------- */
#define SDBE_CHANNELS            16384  /* complex channels per snapshot     */
#define SDBE_SNAPSHOTS_PER_FRAME 128
#define SDBE_FRAMES_PER_GROUP    39     /* B-frames per group = snapshots per block */
#define SDBE_PKTS_PER_FRAME      2048
#define SDBE_PKT_BYTES           1056
#define SDBE_BLOCKS_PER_GROUP    128
#define SDBE_SAMPLES_PER_BLOCK   2097152 /* output samples per block and pol */
#define SDBE_OUT_FRAMES_PER_BLOCK 64
#define SDBE_OUT_PAYLOAD_BYTES   8192
#define SDBE_OUT_HEADER_BYTES    32
#define SDBE_OUT_FRAMES_PER_GROUP (SDBE_BLOCKS_PER_GROUP * SDBE_OUT_FRAMES_PER_BLOCK)
#define SDBE_MAX_CHUNK           1024   /* block-pols per pipeline chunk, at most
                                           (about 20.5 GB of working buffers at 64/39) */

/* ---- threshold policy --------------------------------------------------- */
/* REFERENCE: mean and ddof=1 sigma of the first block of polarisation 0 are
 *   latched for the whole stream and used for both polarisations, recomputed
 *   while the latched sigma is 0 (vdif_inout_gpu_thread.cu:954-1005).
 * PER_BLOCK: every block and polarisation uses its own mean and sigma.
 * FIXED: caller-given threshold and offset.                                */
#define SDBE_THRESH_REFERENCE 0
#define SDBE_THRESH_PER_BLOCK 1
#define SDBE_THRESH_FIXED     2

/* ---- FFT arithmetic (precision variants) -------------------------------- */
/* FP32: radix-16 butterflies on the vector ALUs in FP32 (the default).
 * BF16X3: radix-16 passes as DFT-as-GEMM on the matrix cores, each operand
 *   split into bf16 high + low parts, three products, FP32 accumulation.
 * BF16 / FP16: one bf16 / f16 product per pass (FP16 with a power-of-two
 *   scale per DFT column), FP32 accumulation.
 * Everything else (twiddles, the other stages, statistics) is unchanged.   */
#define SDBE_FFT_FP32   0
#define SDBE_FFT_BF16X3 1
#define SDBE_FFT_BF16   2
#define SDBE_FFT_FP16   3

/* ---- output layout ------------------------------------------------------ */
/* PAYLOAD: 8192-byte payloads back to back per polarisation, the layout of
 *   vdif_out_data_group_t (software/aphids/vdif_out_databuf.h:110-116).
 * FRAMES: complete VDIF frames (32-byte header + payload) back to back per
 *   polarisation, the layout of vdif_out_packet_block_t (vdif_out_databuf.h:66-98). */
#define SDBE_OUT_PAYLOAD 0
#define SDBE_OUT_FRAMES  1

/* Resampler.  SDBE_RATE_64_39: the 39-to-64 rational resampler of the
 * prototypes (resample.py fft_batched): a block of 39 snapshots becomes
 * 2,097,152 samples, 64 frames per block-pol.  SDBE_RATE_6_11 / _8_11 /
 * _10_11 / _11_11: the band-trimmed path of the aphids GPU thread compiled
 * for that SWARM rate (vdif_inout_gpu_thread.cu:589-709, plans :269-333,
 * sample_rates.h:21-59): per block of D = DECIMATION_FACTOR snapshots, a
 * 32768-point C2R per snapshot, 128 real DFTs of 256 D samples, and from each
 * of them bins skip_chan .. skip_chan + 128 E into a 256 E-point C2R,
 * skip_chan = skip_mhz 1e6 / (SWARM_RATE / 256 D), E = EXPANSION_FACTOR:
 *   6/11:  D = 39,  E = 32:  2496 -> 2048 Msps, 1,048,576 samples,  32 frames per block-pol;
 *   8/11:  D = 13,  E = 8:   3328 -> 2048 Msps,   262,144 samples,   8 frames;
 *   10/11: D = 65,  E = 64:  4160 -> 4096 Msps, 2,097,152 samples,  64 frames;
 *   11/11: D = 143, E = 128: 4576 -> 4096 Msps, 4,194,304 samples, 128 frames.
 * In a band-trimmed handle a "block-pol" of the calls below is D snapshots of
 * one polarisation and a group is D B-frames (BENG_FRAMES_PER_GROUP =
 * DECIMATION_FACTOR, sample_rates.h:67; D * 2048 packets, 128 blocks, 128 E
 * frames per polarisation); sdbe_process_blocks and sdbe_set_output_alignment
 * return SDBE_ERR_STATE.                                                    */
#define SDBE_RATE_64_39 0
/* packet geometry (sdbe_config_t.packet_geometry) */
#define SDBE_GEOM_REFERENCE 0   /* 1056-byte packets: 128 snapshots x 8 channels x 2 pols,
                                   2048 per B-frame (cid 0..255, fid 0..7)          */
#define SDBE_GEOM_8X1024    1   /* 8224-byte packets: 8 snapshots x 1024 channels x 2 pols,
                                   256 per B-frame (cid 0..15: channels 1024 cid ..,
                                   fid 0..15: snapshots 8 fid ..), same header and
                                   byte encoding; no counterpart in the reference      */

#define SDBE_RATE_6_11  6
#define SDBE_RATE_8_11  8
#define SDBE_RATE_10_11 10
#define SDBE_RATE_11_11 11

typedef struct sdbe_config {
    int32_t thresh_policy;      /* SDBE_THRESH_*                                */
    float   fixed_threshold;    /* used with SDBE_THRESH_FIXED                  */
    float   fixed_offset;
    int32_t out_mode;           /* SDBE_OUT_*                                   */
    int32_t chunk_blockpols;    /* block-pols per pipeline chunk (even, 2..SDBE_MAX_CHUNK;
                                   a batch of up to this many runs as one launch
                                   per stage)                                  */
    int32_t max_spectra_blockpols; /* capacity of sdbe_process_spectra (>= 1)   */
    uint16_t station_id;        /* VDIF w3.stationID (2 ASCII chars, "XX")      */
    uint8_t  bdc_sideband;      /* VDIF w4.bdc_sideband                        */
    uint8_t  rx_sideband;       /* VDIF w4.rx_sideband                         */
    int32_t use_graph;          /* sdbe_process_blocks: capture the launch
                                   sequence once per batch shape as a HIP graph
                                   and replay it                               */
    int32_t c1_tiles_per_wg;    /* tuning: stage-C1 tiles per workgroup (1..64, 0 = 8) */
    int32_t max_batch_blockpols;/* capacity of sdbe_process_blocks (0 = 1024)   */
    int32_t fft_precision;      /* SDBE_FFT_*: arithmetic of the 16384-point FFT
                                   passes of stages B and C2                   */
    int32_t rate_frac;          /* resampler: SDBE_RATE_64_39 (0), or the aphids
                                   band-trimmed path of a SWARM rate preset,
                                   SDBE_RATE_6_11 / _8_11 / _10_11 / _11_11    */
    int32_t skip_mhz;           /* band-trimmed path: MHz skipped at the bottom
                                   of the band (aphids_resampler_init skip_mhz) */
    int32_t c1_interleave;      /* tuning: stage-C1 workgroups that take adjacent
                                   tiles at each step (1..64, 0 = 16); 1 = each
                                   workgroup runs c1_tiles_per_wg consecutive
                                   tiles                                       */
    int32_t packet_geometry;    /* SDBE_GEOM_*: the packets sdbe_process_group and
                                   sdbe_unpack_group_int8 take                  */
} sdbe_config_t;

/* VDIF output time state for the next frame to be produced; both
 * polarisations share it.  df wraps at 125000 frames per second.           */
typedef struct sdbe_stream_state {
    uint32_t secs;              /* w0.secs_inre                                 */
    uint32_t df;                /* w1.df_num_insec                              */
    uint32_t ref_epoch;         /* w1.ref_epoch                                 */
    uint32_t reserved;
    uint64_t psn;               /* edh_psn of the next frame                    */
} sdbe_stream_state_t;

/* Counters of the last completed call (filled by sdbe_get_stats).          */
typedef struct sdbe_stats {
    int32_t n_packets;          /* packets handed in                            */
    int32_t n_invalid;          /* w0.invalid set: ignored                      */
    int32_t n_out_of_range;     /* B-count outside the group or fid >= 8        */
    int32_t n_duplicate;        /* repeated (frame, cid, fid): first one used   */
    int32_t n_unique;           /* (frame, cid, fid) slots filled               */
    int32_t n_missing;          /* slots zero-filled                            */
    int32_t n_blockpols;        /* block-pols resampled                         */
    int32_t n_frames_per_pol;   /* frames written per polarisation              */
    double  latched_threshold;  /* REFERENCE policy latch (0 if not latched)    */
    double  latched_offset;
} sdbe_stats_t;

typedef struct sdbe_handle sdbe_handle_t;

/* Fill `cfg` with the reference defaults (REFERENCE policy, payload output,
 * station "XX", chunk of 256 block-pols).                                  */
int sdbe_default_config(sdbe_config_t *cfg);

/* Allocate plans, twiddle tables and scratch on HIP device `device`.        */
int sdbe_create(const sdbe_config_t *cfg, int device, sdbe_handle_t **out);
void sdbe_destroy(sdbe_handle_t *h);

const char *sdbe_error_string(int code);
/* The last HIP error code seen by the handle (0 if none).                   */
int sdbe_last_hip_error(const sdbe_handle_t *h);

int sdbe_set_stream_state(sdbe_handle_t *h, const sdbe_stream_state_t *st);
int sdbe_get_stream_state(const sdbe_handle_t *h, sdbe_stream_state_t *st);

/* Derive the stream state from the first group's 32-byte header template the
 * way the output thread does on its first valid template; also returns the
 * sample alignment it implies (int32 words to skip, frames to skip).        */
int sdbe_stream_state_from_template(const void *template_header32,
                                    int32_t n_words_per_chan,
                                    sdbe_stream_state_t *st,
                                    int32_t *skip_int32, int32_t *skip_frames);

/* Process one group.  d_pkts: n_pkts packets of 1056 bytes (device), at
 * most the group's 39 * 2048 (band-trimmed: D * 2048); with SDBE_GEOM_8X1024
 * packets of 8224 bytes, at most 39 * 256 (D * 256).  b_zero: B-count of
 * the group's first frame (bfc[0].b).  d_out: output, polarisation p at
 * d_out + p * out_pol_stride_bytes, 8192 frames (band-trimmed: 128 E) each of
 * 8192 (PAYLOAD) or 8224 (FRAMES) bytes.  d_bitmap (optional, device):
 * uint8 [frames][256], bit f of [frame][cid] set when packet (frame, cid, f)
 * arrived (the cc[] flags of vdif_in_databuf.h:112-130); with
 * SDBE_GEOM_8X1024 uint16 [frames][16], bit f of [frame][cid].  A 16-byte
 * aligned d_pkts (every hipMalloc'd buffer) lets the first kernel read the
 * packets directly (stage A at 64/39, the per-snapshot C2R band-trimmed);
 * other alignments take the int8-row path (same output).                  */
int sdbe_process_group(sdbe_handle_t *h, const void *d_pkts, int32_t n_pkts,
                       uint64_t b_zero, void *d_out, int64_t out_pol_stride_bytes,
                       uint8_t *d_bitmap, void *stream);

/* Unpack one group into the reference's int8 layout: d_beng0/d_beng1 are
 * int8 [frames][128][16384][2 = (im, re)] (163,577,856 bytes each for 39
 * frames), missing packets zero-filled; d_bitmap optional as above.         */
int sdbe_unpack_group_int8(sdbe_handle_t *h, const void *d_pkts, int32_t n_pkts,
                           uint64_t b_zero, int8_t *d_beng0, int8_t *d_beng1,
                           uint8_t *d_bitmap, void *stream);

/* Resample + quantise + pack n_bp independent block-pols given as int8
 * spectra d_beng [n_bp][39][16384][2 = (im, re)].  Output of block-pol i is
 * 64 frames at d_out + i * 64 * frame_bytes.  Threshold policy treats every
 * block-pol as polarisation 0 of consecutive blocks.                        */
int sdbe_process_spectra(sdbe_handle_t *h, const int8_t *d_beng, int32_t n_bp,
                         void *d_out, void *stream);

/* Resample only: d_samples [n_bp][2097152] float, normalised like numpy's
 * irfft(pad(rfft(concat(irfft(X))))) (the reference's cuFFT output divided by
 * 32768 * 2097152); d_stats (optional) [n_bp][2] double = (mean, ddof=1 sigma)
 * of each block-pol's samples.                                              */
int sdbe_resample_spectra_f32(sdbe_handle_t *h, const int8_t *d_beng, int32_t n_bp,
                              float *d_samples, double *d_stats, void *stream);

/* ---- ragged multi-stream batches ---------------------------------------- */
/* One input stream of a batch (a beam and sideband): n_blocks consecutive
 * 39-snapshot blocks of both polarisations, already reordered, and where its
 * VDIF frames go.  Block b of polarisation p reads snapshots 39 b .. 39 b + 38
 * of d_beng[p] (int8 [..][16384][2 = (im, re)], the layout
 * sdbe_unpack_group_int8 writes) and writes 64 frames at
 * d_out + p * out_pol_stride_bytes + 64 b * frame bytes (8192 or 8224).
 * flip = 1 reverses the channel order first (lower sideband, the spectrum
 * reversal of software/prototyping/rslib.py:124).  Each stream has its own
 * threshold policy; REFERENCE streams latch into their own slot
 * (0 .. SDBE_LATCH_SLOTS - 1, shared with whatever else uses that slot; slot 0
 * is the one sdbe_process_group uses).  *state is the VDIF time of the
 * stream's first frame and is advanced by 64 * n_blocks frames.            */
#define SDBE_LATCH_SLOTS 64
typedef struct sdbe_stream_desc {
    const int8_t *d_beng[2];
    int32_t  n_blocks;
    int32_t  flip;
    void    *d_out;
    int64_t  out_pol_stride_bytes;
    sdbe_stream_state_t *state;
    int32_t  thresh_policy;     /* SDBE_THRESH_*                                */
    float    fixed_threshold;
    float    fixed_offset;
    int32_t  latch_slot;
    uint16_t station_id;
    uint8_t  bdc_sideband;
    uint8_t  rx_sideband;
    int32_t  reserved[3];
} sdbe_stream_desc_t;

/* Resample, quantise, pack (and, in FRAMES mode, stamp) all blocks of all
 * streams in one call: the block-pols of every stream are laid end to end and
 * go through the pipeline in chunks, so streams of uneven length share
 * launches.  With cfg.use_graph the launch sequence is captured as a HIP graph
 * the first time a batch of this size is seen and replayed afterwards (the
 * per-block-pol descriptors are copied in by the graph itself).  No reference
 * counterpart: the reference processes one stream and one block per
 * iteration (vdif_inout_gpu_thread.cu:923-1008).                            */
int sdbe_process_blocks(sdbe_handle_t *h, const sdbe_stream_desc_t *streams, int32_t n_streams,
                        void *stream);

/* Re-framed group output: the output thread's time alignment done on the
 * device (software/aphids/vdif_out_net_thread.c:203-290).  With enable = 1 the
 * next sdbe_process_group starts an aligned stream: its frames begin
 * skip_int32 + 2048 skip_frames words into the group's per-polarisation
 * output (the values sdbe_stream_state_from_template returns), so it writes
 * 8192 - ceil((skip_int32 + 2048 skip_frames) / 2048) frames per
 * polarisation; every later group writes 8192 frames, the first of which
 * opens with the words the previous group could not fill a frame with.  The
 * frames written per polarisation are in sdbe_stats_t.n_frames_per_pol, and
 * FRAMES-mode headers are stamped over exactly those frames from the stream
 * state.  enable = 0 returns to group-aligned output.  No reference
 * counterpart as a call: the reference does this on the host after the copy. */
int sdbe_set_output_alignment(sdbe_handle_t *h, int32_t enable, int32_t skip_int32, int32_t skip_frames);

/* Clear REFERENCE latch slot `slot` (asynchronously, on `stream`).          */
int sdbe_reset_latch(sdbe_handle_t *h, int32_t slot, void *stream);

/* Synchronise the handle's last stream and copy the counters out.          */
int sdbe_get_stats(sdbe_handle_t *h, sdbe_stats_t *out);

/* Per-kernel timing for benchmarks (no reference counterpart; the reference
 * times cuFFT calls with cudaEvents under GPU_DEBUG,
 * vdif_inout_gpu_thread.cu:593-622).  When enabled, every kernel launch is
 * bracketed by two HIP events recorded on its own stream.  sdbe_profile_read
 * synchronises on the last event, returns one record per kernel kind that ran
 * (launch count, summed milliseconds, work units processed: block-pols for
 * the resampler kernels, groups for the unpack kernels) and clears them.   */
typedef struct sdbe_kernel_time {
    char    name[32];
    int32_t launches;
    int32_t reserved;
    double  total_ms;
    double  units;
} sdbe_kernel_time_t;

int sdbe_profile_enable(sdbe_handle_t *h, int enable);
int sdbe_profile_read(sdbe_handle_t *h, sdbe_kernel_time_t *out, int32_t max_records,
                      int32_t *n_records);
/* Bracket only the launches of one kernel kind (its name as in
 * sdbe_kernel_time_t; NULL or "" = every kernel), so that a timed run carries
 * two events per launch of that kernel only.  SDBE_ERR_ARG for an unknown name. */
int sdbe_profile_select(sdbe_handle_t *h, const char *kernel);

/* ---- allocation-free variant of sdbe_process_spectra --------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The same work as sdbe_process_spectra on a handle created
 * with *cfg (chunk_blockpols fixed at 2; max_*, use_graph, rate_frac other than
 * SDBE_RATE_64_39 not accepted or ignored as stated) whose stream state is
 * *state, without a handle: the pipeline's working memory is one fixed device
 * array of at most SDBE_NOALLOC_SCRATCH_BYTES per GPU inside the library, held
 * from the first call on that GPU until the process ends.  It runs on the
 * current HIP device and on `stream`; the first call on a device also uploads
 * the twiddle tables (synchronously).  No call allocates or frees memory.
 * n_bp above SDBE_NOALLOC_MAX_BLOCKPOLS returns SDBE_ERR_CAPACITY.
 * *state is advanced by 64 n_bp frames.  The REFERENCE latch starts from zero
 * in every call (the call's first block-pol latches).  All calls on one GPU
 * share its array: calls that overlap on one GPU (two host threads, or two
 * streams without an ordering between them) race and give undefined output.  */
#define SDBE_NOALLOC_MAX_BLOCKPOLS 64
#define SDBE_NOALLOC_SCRATCH_BYTES 41943040   /* 40 MiB */
int sdbe_process_spectra_noalloc(const sdbe_config_t *cfg, sdbe_stream_state_t *state,
                                 const int8_t *d_beng, int32_t n_bp, void *d_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SDBE_STAGE_H */
