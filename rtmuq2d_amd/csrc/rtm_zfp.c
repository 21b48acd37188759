/*
 * rtm_zfp.c -- this library's encoder and decoder for ZFP 0.5.5 streams of
 * 2-D float fields (host C).
 *
 * rtmUq2D keeps its observed seismograms as headerless ZFP 0.5.5 streams
 * (SRC/compression_tools.c:33-70, :94-132; ZFP_LIBRARY_0.5.5/include/zfp.h:75-77
 * for the version).  This file writes and reads the same bits without the
 * ZFP library.  The format, as ZFP 0.5.5 defines it for floats in 2-D:
 *
 *   * The field is cut into 4x4 blocks in raster order (block rows of y,
 *     blocks of x within a row), x varying fastest in memory.  A block at
 *     the right or bottom edge is completed from its own values: a missing
 *     4th value copies the 1st, a missing 3rd copies the 2nd, a missing
 *     2nd copies the 1st (rows first, then columns).
 *   * Per block: e = the binary exponent of its largest |value| (frexp, at
 *     least -126; -127 for an all-zero block); p = min(maxprec,
 *     max(0, e - minexp + 6)) bit planes are kept.  If p = 0 or the block is
 *     zero, one 0 bit.  Otherwise a 1 bit, 8 bits of e + 127, and the block
 *     as integers: v * 2^(30-e) truncated to int32 (FP32 product), a
 *     decorrelating integer lifting transform along x then y, the 16
 *     coefficients in sequency order, each mapped to negabinary, and their
 *     bit planes from the most significant down with group tests (see
 *     encode_planes).  A mode's minbits pads a block with zero bits, its
 *     maxbits truncates it.
 *   * Bits are packed least significant first into 64-bit little-endian
 *     words; the stream ends at the next word boundary.
 *
 * The steps and integer operations are fixed by the format (any change
 * changes the bits), so they are alike by necessity; the code is written
 * here.  Conversions that overflow int32 give INT32_MIN, what the x86
 * truncating conversion returns (reached only by blocks of subnormals).
 */
#include "rtm_zfp.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum {
    Z_MIN_BITS = 1,        /* fewest bits of a block               zfp.h:96  */
    Z_MAX_BITS = 16657,    /* most bits of a block                 zfp.h:97  */
    Z_MAX_PREC = 64,       /* most bit planes                      zfp.h:98  */
    Z_MIN_EXP = -1074,     /* smallest exponent                    zfp.h:99  */
    Z_HEADER_BITS = 148,   /* counted in the size bound            zfp.h:115 */
    F_EBITS = 8,           /* exponent bits of a float block */
    F_EBIAS = 127,
    F_IBITS = 32,          /* bits of a block integer */
};
#define NEGA_MASK 0xaaaaaaaau

/* Sequency order of the 16 coefficients of a 4x4 block, as indices i + 4j:
 * by i + j, then by i^2 + j^2, then the larger i first.  Any other order
 * changes the streams, which the tests compare with ZFP's byte for byte. */
static const unsigned char kOrder[16] = {0, 1, 4, 5, 2, 8, 6, 9, 3, 12, 10, 7, 13, 11, 14, 15};

/* ---- bits ------------------------------------------------------------- */

typedef struct {
    unsigned char *out;
    size_t cap, pos;   /* bytes: capacity, written (whole words) */
    uint64_t acc;      /* pending bits, least significant first  */
    unsigned nb;       /* pending bit count, < 64                */
    int full;          /* ran out of room                        */
} zwriter;

static void zw_word(zwriter *w, uint64_t v)
{
    if (w->pos + 8 > w->cap) {
        w->full = 1;
        return;
    }
    for (int b = 0; b < 8; b++) w->out[w->pos + b] = (unsigned char)(v >> (8 * b));
    w->pos += 8;
}

/* the low n bits of v, n <= 64 */
static void zw_bits(zwriter *w, uint64_t v, unsigned n)
{
    if (!n) return;
    if (n < 64) v &= ((uint64_t)1 << n) - 1;
    const unsigned room = 64 - w->nb;
    w->acc |= v << w->nb;
    if (n < room) {
        w->nb += n;
        return;
    }
    zw_word(w, w->acc);
    w->acc = room < 64 ? v >> room : 0;
    w->nb = n - room;
}

static void zw_zeros(zwriter *w, unsigned n)
{
    for (; n >= 64; n -= 64) zw_bits(w, 0, 64);
    zw_bits(w, 0, n);
}

static void zw_end(zwriter *w)
{
    if (w->nb) zw_word(w, w->acc);
    w->acc = 0;
    w->nb = 0;
}

typedef struct {
    const unsigned char *in;
    size_t bytes, pos;  /* bytes: available, consumed (whole words) */
    uint64_t acc;
    unsigned nb;        /* valid bits in acc */
    int past;           /* read beyond the stream */
} zreader;

static uint64_t zr_word(zreader *r)
{
    uint64_t v = 0;
    if (r->pos + 8 > r->bytes) r->past = 1;
    else
        for (int b = 0; b < 8; b++) v |= (uint64_t)r->in[r->pos + b] << (8 * b);
    r->pos += 8;
    return v;
}

static uint64_t zr_bits(zreader *r, unsigned n)
{
    if (!n) return 0;
    const uint64_t lo = n < 64 ? ((uint64_t)1 << n) - 1 : ~(uint64_t)0;
    if (n <= r->nb) {
        const uint64_t v = r->acc & lo;
        r->acc = n < 64 ? r->acc >> n : 0;
        r->nb -= n;
        return v;
    }
    const uint64_t w = zr_word(r);
    const uint64_t v = (r->acc | (w << r->nb)) & lo;   /* nb < n <= 64, so nb < 64 */
    const unsigned take = n - r->nb;                   /* 1 .. 64 bits of w */
    r->acc = take < 64 ? w >> take : 0;
    r->nb = 64 - take;
    return v;
}

static void zr_skip(zreader *r, unsigned n)
{
    for (; n >= 64; n -= 64) (void)zr_bits(r, 64);
    (void)zr_bits(r, n);
}

/* ---- integer arithmetic of the transform (wrapping, as int32 hardware) --- */

static inline int32_t w_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t w_sub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
static inline int32_t w_shl1(int32_t a) { return (int32_t)((uint32_t)a << 1); }
static inline int32_t a_shr1(int32_t a) { return a >> 1; }   /* arithmetic shift */

/* forward lifting step on (x, y, z, w) at stride s */
static void lift_fwd(int32_t *p, int s)
{
    int32_t x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
    x = a_shr1(w_add(x, w)); w = w_sub(w, x);
    z = a_shr1(w_add(z, y)); y = w_sub(y, z);
    x = a_shr1(w_add(x, z)); z = w_sub(z, x);
    w = a_shr1(w_add(w, y)); y = w_sub(y, w);
    w = w_add(w, a_shr1(y)); y = w_sub(y, a_shr1(w));
    p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}

/* its inverse */
static void lift_inv(int32_t *p, int s)
{
    int32_t x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
    y = w_add(y, a_shr1(w)); w = w_sub(w, a_shr1(y));
    y = w_add(y, w); w = w_sub(w_shl1(w), y);
    z = w_add(z, x); x = w_sub(w_shl1(x), z);
    y = w_add(y, z); z = w_sub(w_shl1(z), y);
    w = w_add(w, x); x = w_sub(w_shl1(x), w);
    p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}

static int f_exponent(float x)
{
    if (x > 0) {
        int e;
        (void)frexpf(x, &e);
        return e > 1 - F_EBIAS ? e : 1 - F_EBIAS;
    }
    return -F_EBIAS;
}

static unsigned block_planes(int emax, const rtm_zfp_mode *m)
{
    const int p = emax - m->minexp + 2 * (2 + 1);
    const unsigned up = p > 0 ? (unsigned)p : 0u;
    return up < m->maxprec ? up : m->maxprec;
}

static int32_t to_int32(float v)
{
    if (!(fabsf(v) < 2147483648.0f)) return INT32_MIN;   /* overflow or NaN, as cvttss2si */
    return (int32_t)v;
}

/* ---- bit planes ---------------------------------------------------------- */

/* Planes 31 .. 32-prec of u[16], most significant first, within `budget`
 * bits.  n counts the coefficients already known to be significant (the
 * leading ones in sequency order); their bits of a plane are written as
 * they are.  The rest of the plane is coded by group tests: a 1 says that a
 * remaining coefficient has this bit set and is followed by the remaining
 * bits up to and including the first set one (after which that coefficient
 * is significant too; a set bit of the last coefficient is implied), a 0
 * says that none has.  Returns the bits written. */
static unsigned encode_planes(zwriter *w, unsigned budget, unsigned prec, const uint32_t *u)
{
    const int kmin = prec < (unsigned)F_IBITS ? F_IBITS - (int)prec : 0;
    unsigned left = budget, n = 0;
    for (int k = F_IBITS - 1; k >= kmin && left > 0; k--) {
        uint64_t x = 0;
        for (int i = 0; i < 16; i++) x |= (uint64_t)((u[i] >> k) & 1u) << i;
        const unsigned m = n < left ? n : left;
        left -= m;
        zw_bits(w, x, m);
        x >>= m;
        while (n < 16 && left > 0) {
            left--;
            const unsigned any = x != 0;
            zw_bits(w, any, 1);
            if (!any) break;
            while (n < 15 && left > 0) {
                left--;
                const unsigned b = (unsigned)(x & 1u);
                zw_bits(w, b, 1);
                if (b) break;
                x >>= 1;
                n++;
            }
            x >>= 1;
            n++;
        }
    }
    return budget - left;
}

/* The inverse: u[16] from the planes; returns the bits read. */
static unsigned decode_planes(zreader *r, unsigned budget, unsigned prec, uint32_t *u)
{
    const int kmin = prec < (unsigned)F_IBITS ? F_IBITS - (int)prec : 0;
    unsigned left = budget, n = 0;
    for (int i = 0; i < 16; i++) u[i] = 0;
    for (int k = F_IBITS - 1; k >= kmin && left > 0; k--) {
        const unsigned m = n < left ? n : left;
        left -= m;
        uint64_t x = zr_bits(r, m);
        while (n < 16 && left > 0) {
            left--;
            if (!zr_bits(r, 1)) break;
            while (n < 15 && left > 0) {
                left--;
                if (zr_bits(r, 1)) break;
                n++;
            }
            x += (uint64_t)1 << n;
            n++;
        }
        for (int i = 0; x; i++, x >>= 1) u[i] += (uint32_t)(x & 1u) << k;
    }
    return budget - left;
}

/* ---- blocks ------------------------------------------------------------- */

static void encode_block(zwriter *w, const rtm_zfp_mode *m, const float *f)
{
    float amax = 0.0f;
    for (int i = 0; i < 16; i++) {
        const float a = fabsf(f[i]);
        if (amax < a) amax = a;
    }
    const int emax = f_exponent(amax);
    const unsigned prec = block_planes(emax, m);
    const unsigned e = prec ? (unsigned)(emax + F_EBIAS) : 0u;
    unsigned bits = 1;
    if (!e) {
        zw_bits(w, 0, 1);
        if (m->minbits > bits) zw_zeros(w, m->minbits - bits);
        return;
    }
    bits += F_EBITS;
    zw_bits(w, 2 * (uint64_t)e + 1, bits);
    const float s = ldexpf(1.0f, F_IBITS - 2 - emax);
    int32_t q[16];
    for (int i = 0; i < 16; i++) q[i] = to_int32(s * f[i]);
    for (int y = 0; y < 4; y++) lift_fwd(q + 4 * y, 1);
    for (int x = 0; x < 4; x++) lift_fwd(q + x, 4);
    uint32_t u[16];
    for (int i = 0; i < 16; i++) u[i] = ((uint32_t)q[kOrder[i]] + NEGA_MASK) ^ NEGA_MASK;
    bits += encode_planes(w, m->maxbits > bits ? m->maxbits - bits : 0u, prec, u);
    if (bits < m->minbits) zw_zeros(w, m->minbits - bits);
}

static void decode_block(zreader *r, const rtm_zfp_mode *m, float *f)
{
    unsigned bits = 1;
    if (!zr_bits(r, 1)) {
        for (int i = 0; i < 16; i++) f[i] = 0.0f;
        if (m->minbits > bits) zr_skip(r, m->minbits - bits);
        return;
    }
    bits += F_EBITS;
    const int emax = (int)zr_bits(r, F_EBITS) - F_EBIAS;
    const unsigned prec = block_planes(emax, m);
    uint32_t u[16];
    bits += decode_planes(r, m->maxbits > bits ? m->maxbits - bits : 0u, prec, u);
    if (bits < m->minbits) zr_skip(r, m->minbits - bits);
    int32_t q[16];
    for (int i = 0; i < 16; i++) q[kOrder[i]] = (int32_t)((u[i] ^ NEGA_MASK) - NEGA_MASK);
    for (int x = 0; x < 4; x++) lift_inv(q + x, 4);
    for (int y = 0; y < 4; y++) lift_inv(q + 4 * y, 1);
    const float s = ldexpf(1.0f, emax - (F_IBITS - 2));
    for (int i = 0; i < 16; i++) f[i] = s * (float)q[i];
}

/* complete a run of n < 4 values at stride s */
static void pad_run(float *p, unsigned n, int s)
{
    if (n == 0) p[0] = 0.0f;
    if (n <= 1) p[s] = p[0];
    if (n <= 2) p[2 * s] = p[s];
    if (n <= 3) p[3 * s] = p[0];
}

static void gather(float *q, const float *a, unsigned nx, unsigned ny, unsigned x0, unsigned y0)
{
    const unsigned bx = nx - x0 < 4 ? nx - x0 : 4, by = ny - y0 < 4 ? ny - y0 : 4;
    for (unsigned y = 0; y < by; y++) {
        const float *row = a + (size_t)(y0 + y) * nx + x0;
        for (unsigned x = 0; x < bx; x++) q[4 * y + x] = row[x];
        if (bx < 4) pad_run(q + 4 * y, bx, 1);
    }
    if (by < 4)
        for (unsigned x = 0; x < 4; x++) pad_run(q + x, by, 4);
}

static void scatter(const float *q, float *a, unsigned nx, unsigned ny, unsigned x0, unsigned y0)
{
    const unsigned bx = nx - x0 < 4 ? nx - x0 : 4, by = ny - y0 < 4 ? ny - y0 : 4;
    for (unsigned y = 0; y < by; y++) {
        float *row = a + (size_t)(y0 + y) * nx + x0;
        for (unsigned x = 0; x < bx; x++) row[x] = q[4 * y + x];
    }
}

/* ---- modes, sizes, fields -------------------------------------------------- */

double rtm_zfp_mode_accuracy(rtm_zfp_mode *m, double tol)
{
    int emin = Z_MIN_EXP;
    if (tol > 0) {           /* tol = f 2^e with 0.5 <= f < 1, so 2^(e-1) <= tol */
        (void)frexp(tol, &emin);
        emin--;
    }
    m->minbits = Z_MIN_BITS;
    m->maxbits = Z_MAX_BITS;
    m->maxprec = Z_MAX_PREC;
    m->minexp = emin;
    return tol > 0 ? ldexp(1.0, emin) : 0.0;
}

unsigned rtm_zfp_mode_precision(rtm_zfp_mode *m, unsigned prec)
{
    m->minbits = Z_MIN_BITS;
    m->maxbits = Z_MAX_BITS;
    m->maxprec = prec && prec < Z_MAX_PREC ? prec : Z_MAX_PREC;
    m->minexp = Z_MIN_EXP;
    return m->maxprec;
}

double rtm_zfp_mode_rate(rtm_zfp_mode *m, double rate)
{
    unsigned bits = (unsigned)floor(16.0 * rate + 0.5);
    if (bits < 1 + F_EBITS) bits = 1 + F_EBITS;
    m->minbits = m->maxbits = bits;
    m->maxprec = Z_MAX_PREC;
    m->minexp = Z_MIN_EXP;
    return (double)bits / 16.0;
}

/* most bits one block can take in mode m */
static unsigned block_bound(const rtm_zfp_mode *m)
{
    unsigned b = 1 + F_EBITS + 15 + 16 * (m->maxprec < (unsigned)F_IBITS ? m->maxprec : (unsigned)F_IBITS);
    if (b > m->maxbits) b = m->maxbits;
    if (b < m->minbits) b = m->minbits;
    return b;
}

size_t rtm_zfp2f_max_bytes(const rtm_zfp_mode *m, unsigned nx, unsigned ny)
{
    const size_t mx = ((nx ? nx : 1u) + 3) / 4, my = ((ny ? ny : 1u) + 3) / 4;
    const size_t bits = Z_HEADER_BITS + mx * my * block_bound(m);
    return ((bits + 63) & ~(size_t)63) / 8;
}

/* Encode block rows [r0, r1) into w (no end padding). */
static void encode_rows(zwriter *w, const rtm_zfp_mode *m, const float *a, unsigned nx, unsigned ny, unsigned r0,
                        unsigned r1)
{
    float q[16];
    for (unsigned by = r0; by < r1; by++)
        for (unsigned x0 = 0; x0 < nx; x0 += 4) {
            gather(q, a, nx, ny, x0, 4 * by);
            encode_block(w, m, q);
        }
}

size_t rtm_zfp2f_encode(const rtm_zfp_mode *m, const float *a, unsigned nx, unsigned ny, void *out, size_t cap,
                        int threads)
{
    if (!nx || !ny) return 0;
    const unsigned rows = (ny + 3) / 4, mx = (nx + 3) / 4;
    zwriter w = {(unsigned char *)out, cap, 0, 0, 0, 0};
    int nchunk = threads > 1 ? threads : 1;
    if ((unsigned)nchunk > rows) nchunk = (int)rows;
    if (nchunk <= 1) {
        encode_rows(&w, m, a, nx, ny, 0, rows);
        zw_end(&w);
        return w.full ? 0 : w.pos;
    }
    /* chunks of block rows into their own buffers, then their bits
     * concatenated in order: the same stream as one pass */
    zwriter *part = calloc((size_t)nchunk, sizeof(zwriter));
    if (!part) return 0;
    const size_t per_row = ((size_t)mx * block_bound(m) + 63) / 64 * 8 + 8;
    int ok = 1;
    for (int c = 0; c < nchunk; c++) {
        const unsigned r0 = (unsigned)((uint64_t)rows * c / nchunk), r1 = (unsigned)((uint64_t)rows * (c + 1) / nchunk);
        part[c].cap = per_row * (r1 - r0) + 8;
        part[c].out = malloc(part[c].cap);
        if (!part[c].out) ok = 0;
    }
    if (ok) {
#pragma omp parallel for num_threads(nchunk) schedule(static, 1)
        for (int c = 0; c < nchunk; c++) {
            const unsigned r0 = (unsigned)((uint64_t)rows * c / nchunk), r1 = (unsigned)((uint64_t)rows * (c + 1) / nchunk);
            encode_rows(&part[c], m, a, nx, ny, r0, r1);
        }
        for (int c = 0; c < nchunk && ok; c++) {
            if (part[c].full) ok = 0;
            const size_t words = part[c].pos / 8;
            for (size_t t = 0; t < words; t++) {
                uint64_t v = 0;
                for (int b = 0; b < 8; b++) v |= (uint64_t)part[c].out[8 * t + b] << (8 * b);
                zw_bits(&w, v, 64);
            }
            zw_bits(&w, part[c].acc, part[c].nb);
        }
        zw_end(&w);
    }
    for (int c = 0; c < nchunk; c++) free(part[c].out);
    free(part);
    return ok && !w.full ? w.pos : 0;
}

size_t rtm_zfp2f_decode(const rtm_zfp_mode *m, const void *in, size_t bytes, unsigned nx, unsigned ny, float *a)
{
    if (!nx || !ny) return 0;
    zreader r = {(const unsigned char *)in, bytes, 0, 0, 0, 0};
    float q[16];
    for (unsigned y0 = 0; y0 < ny; y0 += 4)
        for (unsigned x0 = 0; x0 < nx; x0 += 4) {
            decode_block(&r, m, q);
            scatter(q, a, nx, ny, x0, y0);
        }
    return r.past ? 0 : r.pos;
}
