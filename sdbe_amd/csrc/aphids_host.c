/*
 * Host-side work the APHIDS input thread does around the GPU stage, with the
 * behaviour of the reference functions of the same names:
 *   group window / insert / completion / header template
 *       software/aphids/vdif_in_databuf.c:56-220
 *   new-format B-engine timestamp -> B-frame count
 *       software/aphids/beng_reform.c:16-260
 *   8-packet aggregate unpack
 *       software/aphids/vdif_8pac.c:5-56
 * and the two databuf constructors (vdif_in_databuf.c:56-63,
 * vdif_out_databuf.c:10-14).
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "aphids_databuf.h"

#define SWARM_RATE_6_11 (4576e6 * 6 / 11)        /* 2496 Msps */
#define SWARM_CLK_PER_SEC 286000000
#define SWARM_BFRAME_CLK 262144
#define BFRAME_CLK_TOL 1024
#define MAX_COUNTABLE_BFRAME_OFFSET 1000
#define SWARM_TIMESTAMP_OFF_BY_BENG (-1)

/* ---------------------------------------------------------- constructors */
hashpipe_databuf_t *vdif_in_databuf_create(int instance_id, int databuf_id) {
    return hashpipe_databuf_create(instance_id, databuf_id, sizeof(hashpipe_databuf_t),
                                   sizeof(beng_group_completion_t), BENG_GROUPS_IN_BUFFER);
}

hashpipe_databuf_t *vdif_out_databuf_create(int instance_id, int databuf_id) {
    return hashpipe_databuf_create(instance_id, databuf_id, sizeof(hashpipe_databuf_t),
                                   sizeof(quantized_storage_t), VDIF_OUT_BUFFER_SIZE);
}

/* ------------------------------------------------------ group bookkeeping */
int64_t get_packet_b_count(vdif_in_header_t *h) {
    if (h->w0.invalid) return -1;
    return ((int64_t)h->beng.b_upper << 8) | (int64_t)h->beng.b_lower;
}

void init_beng_group(beng_group_completion_t *bgc, beng_group_vdif_buffer_t *bgv_buf_cpu,
                     beng_group_vdif_buffer_t *bgv_buf_gpu, int64_t b_start) {
    bgc->bgv_buf_cpu = bgv_buf_cpu;
    bgc->bgv_buf_gpu = bgv_buf_gpu;
    bgc->beng_group_vdif_packet_count = 0;
    for (int i = 0; i < BENG_FRAMES_PER_GROUP; ++i) {
        memset(&bgc->bfc[i], 0, sizeof(bgc->bfc[i]));
        bgc->bfc[i].b = (uint64_t)(b_start + i);
    }
    bgc->vdif_header_template.w0.invalid = 1;
}

/* Offset (0..3) of the group whose B-count window holds the packet, counted
 * from index_ref; the last offset when the packet lies beyond the four
 * windows but within another four of them; negative codes otherwise. */
int get_beng_group_index_offset(vdif_in_databuf_t *bgc_buf, int index_ref, vdif_in_packet_t *pkt) {
    const int64_t b = get_packet_b_count(&pkt->header);
    int64_t lo = (int64_t)bgc_buf->bgc[index_ref].bfc[0].b;
    int64_t hi = (int64_t)bgc_buf->bgc[index_ref].bfc[BENG_FRAMES_PER_GROUP - 1].b;
    if (pkt->header.w0.invalid) return vidErrorPacketInvalid;
    if (b < lo) return vidErrorPacketBeforeStartTime;
    int offset = -1;
    for (int n = 0; n < BENG_GROUPS_IN_BUFFER; ++n) {
        ++offset;
        if (b >= lo && b <= hi) return offset;
        lo += BENG_FRAMES_PER_GROUP;
        hi += BENG_FRAMES_PER_GROUP;
    }
    if (b > hi + BENG_GROUPS_IN_BUFFER * BENG_FRAMES_PER_GROUP) return vidErrorPacketTooFarAheadToCare;
    return offset;
}

int insert_vdif_in_beng_group_buffer(vdif_in_databuf_t *bgc_buf, int index_ref, int offset,
                                     vdif_in_packet_t *pkt) {
    const int64_t b = get_packet_b_count(&pkt->header);
    int inserted = 0;
    for (int i = offset; i < BENG_GROUPS_IN_BUFFER; ++i) {
        beng_group_completion_t *g = &bgc_buf->bgc[(i + index_ref) % BENG_GROUPS_IN_BUFFER];
        const int64_t lo = (int64_t)g->bfc[0].b, hi = (int64_t)g->bfc[BENG_FRAMES_PER_GROUP - 1].b;
        if (b < lo || b > hi) continue;
        beng_frame_completion_t *f = &g->bfc[b - lo];
        if (f->beng_frame_vdif_packet_count >= VDIF_PER_BENG_FRAME ||
            g->beng_group_vdif_packet_count >= BENG_FRAMES_PER_GROUP * VDIF_PER_BENG_FRAME)
            continue;
        f->beng_frame_vdif_packet_count++;
        ((uint8_t *)f->cc)[pkt->header.beng.c] |= (uint8_t)(1u << pkt->header.beng.f);
        memcpy((char *)g->bgv_buf_cpu + (size_t)g->beng_group_vdif_packet_count * sizeof(vdif_in_packet_t), pkt,
               sizeof(vdif_in_packet_t));
        g->beng_group_vdif_packet_count++;
        ++inserted;
    }
    return inserted;
}

int check_beng_group_complete(vdif_in_databuf_t *bgc_buf, int index) {
    if (index < 0 || index >= BENG_GROUPS_IN_BUFFER) return -1;
    return bgc_buf->bgc[index].beng_group_vdif_packet_count == BENG_FRAMES_PER_GROUP * VDIF_PER_BENG_FRAME;
}

void fill_vdif_header_template(vdif_in_header_t *out, vdif_in_packet_t *ref, int n_skipped) {
    (void)n_skipped;   /* the reference only logs the implied offset */
    beng_timestamp_t t0;
    memcpy(out, ref, sizeof(vdif_in_header_t));
    get_beng_t0(&t0);
    beng_timestamp_increment(&t0);   /* first used packet is one B-frame later */
    out->w0.secs_inre = (uint32_t)t0.sec;
    const double frame_full = (VDIF_OUT_FRAMES_PER_SECOND * 1.0) * beng_timestamp_clk_to_float(&t0);
    const double frame_ceil = ceil(frame_full);
    const double frame_frac = frame_ceil - frame_full;
    out->w1.df_num_insec = (uint32_t)(int)frame_ceil;
    /* picoseconds before that frame boundary, hidden in edh_psn */
    out->edh_psn = (uint64_t)floor(frame_frac * (1e12 / VDIF_OUT_FRAMES_PER_SECOND));
}

/* ------------------------------------------------------ timestamp reform */
typedef struct newbeng_hdr {
    uint32_t b_upper;
    uint32_t c:8;
    uint32_t f:3;
    uint32_t b_lower:21;
} newbeng_hdr_t;

static beng_timestamp_t T0 = {-1, -1};
static beng_timestamp_t TPREV = {-1, -1};
static int64_t BPREV = -1;

void beng_reform_reset(void) {
    T0.sec = T0.clk = -1;
    TPREV.sec = TPREV.clk = -1;
    BPREV = -1;
}

static void read_timestamp(const newbeng_hdr_t *h, beng_timestamp_t *t) {
    const int64_t b = ((int64_t)h->b_upper << 21) | (int64_t)(h->b_lower & 0x1FFFFF);
    t->clk = (int32_t)(b & 0x1FFFFFFF);
    t->sec = (int32_t)((b & 0x001FFFFFE0000000LL) >> 29);
}

/* <0: then earlier than now, 0: same B-frame (within the clock tolerance),
 * >0: later.  Across a second boundary only the clock wrap is checked. */
static int compare_then_now(const beng_timestamp_t *then, const beng_timestamp_t *now) {
    if (then->sec < now->sec) {
        if (then->sec == now->sec - 1 && then->clk + BFRAME_CLK_TOL > SWARM_CLK_PER_SEC &&
            (then->clk + BFRAME_CLK_TOL) % SWARM_CLK_PER_SEC >= now->clk)
            return 0;
        return -1;
    }
    if (then->sec > now->sec) {
        if (now->clk + BFRAME_CLK_TOL > SWARM_CLK_PER_SEC &&
            (now->clk + BFRAME_CLK_TOL) % SWARM_CLK_PER_SEC >= then->clk)
            return 0;
        return 1;
    }
    if (then->clk + BFRAME_CLK_TOL < now->clk) return -1;
    if (now->clk + BFRAME_CLK_TOL < then->clk) return 1;
    return 0;
}

static int64_t offset_then_now(const beng_timestamp_t *then, const beng_timestamp_t *now) {
    const int cmp = compare_then_now(then, now);
    if (cmp == 0) return 0;
    beng_timestamp_t a = cmp < 0 ? *then : *now;
    const beng_timestamp_t b = cmp < 0 ? *now : *then;
    const int64_t step = cmp < 0 ? 1 : -1;
    int64_t off = 0;
    while (compare_then_now(&a, &b) < 0) {
        beng_timestamp_increment(&a);
        off += step;
        if (off > MAX_COUNTABLE_BFRAME_OFFSET || -off > MAX_COUNTABLE_BFRAME_OFFSET) break;
    }
    return off;
}

void get_beng_t0(beng_timestamp_t *t) {
    *t = T0;
    for (int i = 0; i > SWARM_TIMESTAMP_OFF_BY_BENG; --i) beng_timestamp_decrement(t);
}

void get_beng_tprev(beng_timestamp_t *t) { *t = TPREV; }
void get_beng_bprev(int64_t *b) { *b = BPREV; }

void beng_reform_headers(vdif_in_packet_t *pkts, int npkts) {
    for (int i = 0; i < npkts; ++i) {
        vdif_in_packet_t *p = pkts + i;
        if (p->header.w0.invalid) continue;
        newbeng_hdr_t *h = (newbeng_hdr_t *)&p->header.beng;
        if (T0.sec == -1 && T0.clk == -1) {
            read_timestamp(h, &T0);
            TPREV = T0;
            BPREV = 0;
        }
        beng_timestamp_t now;
        read_timestamp(h, &now);
        int off = (int)offset_then_now(&TPREV, &now);
        if (off > MAX_COUNTABLE_BFRAME_OFFSET || off < -MAX_COUNTABLE_BFRAME_OFFSET) {
            p->header.w0.invalid = 1;
            off = -1;
        } else {
            off += (int)BPREV;
            TPREV = now;
            BPREV = off;
        }
        const int c = h->c, f = h->f;
        p->header.beng.f = (uint8_t)f;
        p->header.beng.c = (uint8_t)c;
        p->header.beng.b_lower = (uint8_t)(off & 0xFF);
        p->header.beng.b_upper = (uint32_t)((off & 0x7FFFFF00) >> 8);
    }
}

double beng_timestamp_clk_to_float(beng_timestamp_t *t) { return (double)t->clk / (double)SWARM_CLK_PER_SEC; }

void beng_timestamp_increment(beng_timestamp_t *t) {
    t->clk += SWARM_BFRAME_CLK;
    if (t->clk > SWARM_CLK_PER_SEC) {
        t->sec += 1;
        t->clk = t->clk % SWARM_CLK_PER_SEC;
    }
}

void beng_timestamp_decrement(beng_timestamp_t *t) {
    t->clk -= SWARM_BFRAME_CLK;
    if (t->clk < 0) {
        t->sec -= 1;
        t->clk += SWARM_CLK_PER_SEC;
    }
}

/* -------------------------------------------------------------------- 8pac */
void unpack_8pac(vdif_in_packet_t *dest, vdif_8pac_packet_t *src, int n_src) {
    int n = 0;
    for (int a = 0; a < n_src; ++a) {
        vdif_8pac_packet_t *s = src + a;
        vdif_in_header_t *h1 = &s->pkt1.header;
        const uint32_t newsize = (uint32_t)(((int)h1->w2.df_len - 7) / 8) & 0x00FFFFFF;
        const uint32_t df8 = 8u * (uint32_t)h1->w1.df_num_insec;
        const uint64_t psn8 = 8u * h1->edh_psn;
        h1->w2.df_len = newsize;
        h1->w1.df_num_insec = df8;
        h1->edh_psn = psn8;
        for (int i = 0; i < 7; ++i) {
            vdif_in_header_t *h = &s->pac7[i].pac.header;
            h->w2.df_len = newsize;
            h->w1.df_num_insec = df8 + (uint32_t)i + 1;
            h->edh_psn = psn8 + (uint64_t)i + 1;
            h->w0.invalid = h1->w0.invalid;
        }
        memcpy(dest + n++, &s->pkt1, sizeof(vdif_in_packet_t));
        for (int i = 0; i < 7; ++i) memcpy(dest + n++, &s->pac7[i].pac, sizeof(vdif_in_packet_t));
    }
}
