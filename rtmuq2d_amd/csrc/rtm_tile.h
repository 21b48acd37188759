/*
 * rtm_tile.h -- the LDS-tiled step kernel body, written as barrier-separated
 * phases so the same code runs on the GPU (one workgroup per tile) and in the
 * CPU test harness (phases run one after another for all 256 threads).
 *
 * A launch advances any number of simulations ("sims") by one time step.
 * Each sim is described by a SimDesc in device memory (its grid may differ
 * from the others'), and the launch walks a flat list of tiles
 * {sim, x0, z0, class}.  A tile covers 64 x TZ points (TZ = NW*RZ) of the
 * stepped region 4 <= i < X-4, 4 <= k < Z-4; 64*NW threads = 64 columns x
 * NW row groups (NW = 4 or 8 wavefronts), each thread owning RZ consecutive
 * rows of one column.
 *
 * Phase 1  load u^n over the tile plus a halo into LDS: radius 4, or radius
 *          8 across a CPML strip (the strip's psi is needed 4 points out,
 *          and psi itself reads u 4 points out);
 * Phase 2  (strip tiles) psi^{n+1} over the tile plus a radius-4 halo into
 *          LDS, once per point -- replacing the reference's separate first
 *          pass and barrier (SRC/absorbingBoundaryCondition.c:174-217);
 * Phase 3  per point: stencil, CPML corrections in the order L, T, R, B,
 *          source add or receiver replacement, snapshot store or imaging.
 *
 * The arithmetic is the value forms of rtm_step.h, so results are
 * bit-identical to rtm::step_point (and to the reference's FP32 arithmetic).
 * psi^{n+1} of a point outside the update range is 0: in a modeling pass
 * those entries are never written and start at zero.
 */
#ifndef RTM_TILE_H
#define RTM_TILE_H

#include <math.h>

#include "rtm_step.h"

namespace rtm {

/* Global-memory access.  Device pointers reach the kernel through SimDesc
 * (not as kernel arguments), so the compiler cannot infer their address
 * space and would emit flat loads, whose completion counter is shared with
 * LDS; casting to the global address space keeps HBM loads and LDS traffic
 * independent.  Offsets are 32-bit (one field is < 2^31 elements). */
#if defined(__HIP_DEVICE_COMPILE__)
#define RTM_GAS __attribute__((address_space(1)))
#else
#define RTM_GAS
#endif

/* The byte offset is a 32-bit unsigned value (every index here is >= 0 and
 * every array is < 4 GiB; snapshot slices are offset per step on the host
 * side of the view), which lets the compiler use the scalar-base + 32-bit
 * vector-offset addressing form instead of 64-bit address arithmetic per
 * access. */
template <typename T>
RTM_HD T ldg(const T *p, int i)
{
    const unsigned off = (unsigned)i * (unsigned)sizeof(T);
    return *(const RTM_GAS T *)((const RTM_GAS char *)p + off);
}

template <typename T>
RTM_HD void stg(T *p, int i, T v)
{
    const unsigned off = (unsigned)i * (unsigned)sizeof(T);
    *(RTM_GAS T *)((RTM_GAS char *)p + off) = v;
}

/* A store the pass does not read back (snapshot stores): on the device a
 * non-temporal store, which spares the caches for the fields every step
 * re-reads (measured 3% faster on the c1 forward step at TS=1; the matching
 * non-temporal load of the backward pass was slower and is not used). */
template <typename T>
RTM_HD void stg_stream(T *p, int i, T v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    /* through the global address space: a flat store would also count
     * against the LDS counter and make the next LDS wait wait for it */
    const unsigned off = (unsigned)i * (unsigned)sizeof(T);
    __builtin_nontemporal_store(v, (RTM_GAS T *)((RTM_GAS char *)p + off));
#else
    p[i] = v;
#endif
}

RTM_HD int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

enum SimPhase : int { PH_IDLE = 0, PH_FORWARD = 1, PH_BACKWARD = 2 };
enum SimFlags : int {
    F_KAHAN = 1,    /* compensated image accumulation          */
    F_RECORD = 2,   /* fused seismogram recording (forward)     */
    F_SNAP = 4,     /* store snapshots (forward)                */
    F_IMAGE = 8,    /* image against snapshots (backward)       */
    F_RING = 16,    /* boundary saving: forward saves the ring, backward
                       images against the reconstructed source field     */
};
enum TileClass : int { T_XSTRIP = 1, T_ZSTRIP = 2 };
/* In a mixed tile list (IN_MIXED): a 64 x 4 tile, its class in the low bits */
constexpr int T_SMALL = 4;

template <typename R>
struct SimDesc {
    int X, Z, P;
    int TL, TR, TU, TD;
    int nx, nz;
    int nt, TS, nw, nrec;
    int phase;              /* SimPhase */
    int flags;              /* SimFlags */
    int src_j;              /* device index of the source, -1 if none */
    int rec_row;
    R h1, h2, h3;
    /* u at step n is ubuf[n & umask]: two buffers (umask 1, the update
     * overwrites u^{n-1} in place) or four (umask 3, two-step launches:
     * a launch reads two time levels and writes the next two) */
    R *ubuf[4];
    int umask;
    const R *C;
    const R *ax, *bx, *az, *bz;
    R *psiX[2], *psiZ[2];   /* psi^n is psi?[n & 1]                    */
    R *csiX, *csiZ;
    const int *recmap;      /* [X] receiver at column i on rec_row, -1 */
    const float *wav;       /* [nw]                                    */
    const float *seis_t;    /* [nt][nrec] input traces (backward)      */
    float *seis_rec;        /* [nt][nrec] recorded traces (forward)    */
    R *snaps;               /* [nt/TS][nz][nx]                          */
    R *I, *Is, *Ic, *Isc;   /* [nz][nx]                                 */
    int qmode;              /* snapshot store: 0 as R, Q_FP16, Q_BF16   */
    unsigned short *q16;    /* [nt/TS][nz][nx] scaled 16-bit store      */
    unsigned *smax;         /* [nt/TS][kSnapSlots * kSlotStride] max |u| slots (float bits) */
    /* boundary saving: the source wavefield's own buffers (u^n in
     * sbuf[n & 1]) and the ring of width 4 around the image region saved
     * after every forward step, [nt][NR] */
    R *sbuf[2];
    R *ring;
    int NR;
    /* int16 block store (Q_I16): codes in q16, and per snapshot, image row
     * and block of 64 padded columns the block's binary exponent,
     * [nt/TS][nz][nqb]; block b covers padded columns kHalf + 64 b ..
     * kHalf + 64 b + 63, and qb0 is the first block holding image columns */
    signed char *qexp;
    int qb0, nqb;
};

/* The ring of boundary saving: every point within 4 of the image region
 * [TL, TL+nx) x [TU, TU+nz) and outside it (its values, saved every forward
 * step, let the region's interior be stepped backward in time).  Bands: top
 * and bottom (4 rows of nx+8 points each), then left and right (4 columns of
 * nz points each). */
template <typename R>
RTM_HD int ring_count(const SimDesc<R> &S)
{
    return 8 * (S.nx + 8) + 8 * S.nz;
}

template <typename R>
RTM_HD void ring_point(const SimDesc<R> &S, int t, int &i, int &k)
{
    const int W = S.nx + 8;
    if (t < 4 * W) {
        k = S.TU - 4 + t / W; i = S.TL - 4 + t % W;
    } else if (t < 8 * W) {
        t -= 4 * W;
        k = S.TU + S.nz + t / W; i = S.TL - 4 + t % W;
    } else if (t < 8 * W + 4 * S.nz) {
        t -= 8 * W;
        i = S.TL - 4 + t / S.nz; k = S.TU + t % S.nz;
    } else {
        t -= 8 * W + 4 * S.nz;
        i = S.TL + S.nx + t / S.nz; k = S.TU + t % S.nz;
    }
}

/* 16-bit snapshot store.  Each snapshot c is stored as u * 2^(14-e_c),
 * rounded to FP16 or BF16 (nearest, ties to even), where e_c is the binary
 * exponent of the snapshot's max |u| (clamped to [-100, 100]): the largest
 * value lands in [2^14, 2^15), well inside FP16's range, and the power-of-two
 * scale and its inverse are exact, so the only error is the rounding.  The
 * conversions are integer code shared by the device and the CPU harness. */
enum QMode : int { Q_NONE = 0, Q_FP16 = 1, Q_BF16 = 2, Q_I16 = 3 };

/* A snapshot's max |u| is accumulated with atomic max over kSnapSlots
 * counters, kSlotStride words (256 bytes) apart so that concurrent
 * workgroups update different lines; its value is the max of the slots. */
constexpr int kSnapSlots = 32, kSlotStride = 64, kSnapWords = kSnapSlots * kSlotStride;

RTM_HD unsigned f32_bits(float f)
{
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}

RTM_HD float bits_f32(unsigned u)
{
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

RTM_HD int snap_exp(unsigned maxabs_bits)
{
    const int e = (int)((maxabs_bits >> 23) & 0xff) - 127;
    return e < -100 ? -100 : (e > 100 ? 100 : e);
}

RTM_HD float pow2f(int k) { return bits_f32((unsigned)(k + 127) << 23); }   /* k in [-126, 127] */

RTM_HD unsigned short f32_to_f16(float f)
{
    const unsigned x = f32_bits(f);
    const unsigned sign = (x >> 16) & 0x8000u;
    unsigned mant = x & 0x7fffffu;
    const int ex = (int)((x >> 23) & 0xff);
    if (ex == 0xff) return (unsigned short)(sign | 0x7c00u | (mant ? 0x200u : 0u));
    const int e = ex - 112;                   /* FP16 biased exponent */
    if (e >= 31) return (unsigned short)(sign | 0x7c00u);
    if (e <= 0) {                             /* FP16 subnormal or zero */
        if (e < -10) return (unsigned short)sign;
        mant |= 0x800000u;
        const int shift = 14 - e;
        unsigned h = mant >> shift;
        const unsigned rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (unsigned short)(sign | h);
    }
    unsigned h = ((unsigned)e << 10) | (mant >> 13);
    const unsigned rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;   /* a carry rounds into the exponent */
    return (unsigned short)(sign | h);
}

RTM_HD float f16_to_f32(unsigned short h)
{
    const unsigned sign = ((unsigned)h & 0x8000u) << 16;
    const unsigned e = ((unsigned)h >> 10) & 0x1fu, m = (unsigned)h & 0x3ffu;
    if (e == 0) {                             /* m * 2^-24, exact */
        const float v = (float)m * 5.9604644775390625e-8f;
        return sign ? -v : v;
    }
    if (e == 31) return bits_f32(sign | 0x7f800000u | (m << 13));
    return bits_f32(sign | ((e + 112u) << 23) | (m << 13));
}

RTM_HD unsigned short f32_to_bf16(float f)
{
    const unsigned x = f32_bits(f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((x >> 16) | 0x40u);   /* NaN */
    return (unsigned short)((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
}

RTM_HD float bf16_to_f32(unsigned short b) { return bits_f32((unsigned)b << 16); }

/* On the device the conversions are the hardware's (v_cvt_f16_f32,
 * v_cvt_pk_bf16_f32, v_cvt_f32_f16): round to nearest even with subnormals
 * kept (the kernels run with FP16/FP32 denormals on), the results of the
 * integer restatements above, which the host uses. */
RTM_HD unsigned short quant16(float v, float scale, int mode)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float x = v * scale;
    return mode == Q_FP16 ? __builtin_bit_cast(unsigned short, (_Float16)x)
                          : __builtin_bit_cast(unsigned short, (__bf16)x);
#else
    return mode == Q_FP16 ? f32_to_f16(v * scale) : f32_to_bf16(v * scale);
#endif
}

RTM_HD float dequant16(unsigned short q, float scale, int mode)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (mode == Q_FP16 ? (float)__builtin_bit_cast(_Float16, q) : bf16_to_f32(q)) * scale;
#else
    return (mode == Q_FP16 ? f16_to_f32(q) : bf16_to_f32(q)) * scale;
#endif
}

/* The int16 block store.  Each image row of a snapshot is cut into blocks
 * of 64 padded columns aligned with the step kernels' tiles (every tile and
 * column item starts at kHalf + 64 t), so a block is one wavefront row of a
 * tile.  A block with max |u| = f 2^E (0.5 <= f < 1) stores each value as
 * the int16 nearest u 2^(15-E) (ties to even, at most 32767 in magnitude)
 * and E (clamped to [-100, 100]); it reads back as code 2^(E-15).  Powers of
 * two scale exactly, so the rounding is the only error: at most 2^-16 of the
 * block's max, against FP16's 2^-11 of each value, in the same 2 bytes per
 * point plus a byte per block. */
constexpr int kQBlock = 64;

RTM_HD int i16_block(int i) { return (i - kHalf) >> 6; }   /* block of padded column i >= kHalf */
RTM_HD int i16_first_block(int TL) { return i16_block(TL); }
RTM_HD int i16_blocks(int TL, int nx) { return i16_block(TL + nx - 1) - i16_block(TL) + 1; }

RTM_HD int i16_exp(float maxabs)
{
    const int e = (int)((f32_bits(maxabs) >> 23) & 0xff) - 126;   /* maxabs < 2^e for normal values */
    return e < -100 ? -100 : (e > 100 ? 100 : e);
}

RTM_HD short quant_i16(float v, float scale)
{
    float r = rintf(v * scale);   /* the current (nearest-even) rounding mode on device and host */
    r = r > 32767.0f ? 32767.0f : (r < -32767.0f ? -32767.0f : r);
    return (short)r;
}

RTM_HD float dequant_i16(short q, int e) { return (float)q * pow2f(e - 15); }

/* The max over a snapshot's kSnapSlots slots.  On the device every lane of
 * the (fully active) wavefront loads one slot and a butterfly of lane shuffles
 * takes the max, one memory round trip; the host loops. */
RTM_HD unsigned snap_max_bits(const unsigned *slots)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const int lane = (int)(threadIdx.x & 63);
    unsigned m = lane < kSnapSlots ? slots[lane * kSlotStride] : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned o = (unsigned)__shfl_xor((int)m, off);
        m = o > m ? o : m;
    }
    return (unsigned)__builtin_amdgcn_readfirstlane((int)m);
#else
    unsigned m = 0;
    for (int t = 0; t < kSnapSlots; t++) {
        const unsigned v = ldg(slots, t * kSlotStride);
        m = v > m ? v : m;
    }
    return m;
#endif
}

/* The psi^{n+1} planes of a tile: psiX for x-strip tiles, psiZ for z-strip
 * tiles.  Only a corner tile (both strips) needs both at once; the work
 * lists cut corner tiles of four row groups into 64 x 4 tiles
 * (corner_tiles_small), so those tiles keep the two planes in one union and
 * the LDS of a workgroup stays within an eighth of the CU's. */
template <typename R, int RZ, int NW, bool UNI>
struct TilePsiLds {
    static constexpr int TZ = NW * RZ;
    R px[TZ][72];           /* psiX^{n+1}: rows z0 .., columns x0-4 .. x0+67 */
    R pz[TZ + 8][64];       /* psiZ^{n+1}: rows z0-4 .., columns x0 .. x0+63 */
};

template <typename R, int RZ, int NW>
struct TilePsiLds<R, RZ, NW, true> {
    static constexpr int TZ = NW * RZ;
    union {
        R px[TZ][72];
        R pz[TZ + 8][64];
    };
};

/* Tiles of four row groups and more than one row per thread never are corner
 * tiles (the lists cut those into 64 x 4 tiles). */
constexpr bool corner_tiles_small(int rz, int nw) { return rz > 1 && nw == 4; }

template <typename R, int RZ, int NW = 4>
struct TileLds : TilePsiLds<R, RZ, NW, corner_tiles_small(RZ, NW)> {
    static constexpr int TZ = NW * RZ;
    R u[TZ + 16][80];       /* rows z0-8 .. z0+TZ+7, columns x0-8 .. x0+71 */
    R az[TZ + 8], bz[TZ + 8];   /* z profiles of rows z0-4 .. z0+TZ+3 (z-strip tiles) */
};

/* Per-launch view of one sim at step n. */
template <typename R>
struct StepView {
    int n, active, forward;
    const R *u;
    R *U;                   /* u^{n+1} written */
    const R *Uo;            /* u^{n-1} read (U itself with two buffers) */
    const R *psiXi, *psiZi;
    R *psiXo, *psiZo;
    int src;                /* add the source this step */
    R src_w;
    const float *seis_row;  /* backward: traces at tau */
    float *rec_row_out;     /* forward recording row   */
    R *snap;                /* snapshot written (forward) or read (backward) */
    /* 16-bit store: forward, the slot of the previous step's snapshot, filled
     * this step from u^n (= that step's result) with scale 2^(14-e); backward,
     * the slot read, with scale 2^(e-14) */
    unsigned short *qsnap;
    float qscale;
    unsigned *smax;         /* forward, 16-bit store: this step's snapshot max |U| */
    const R *sfield;        /* backward, boundary saving: the reconstructed source field (padded) */
    signed char *qexp;      /* int16 block store: the block exponents of the snapshot in qsnap */
};

/* Q16: the sim keeps a 16-bit snapshot store (SimDesc::qmode != 0).  The
 * kernels dispatch on it per sim, so the FP32-store code carries none of the
 * 16-bit fields. */
template <typename R, bool Q16, bool I16 = false, int UM = 1>
RTM_HD StepView<R> step_view_q(const SimDesc<R> &S, int n)
{
    StepView<R> v{};
    v.n = n;
    v.active = S.phase != PH_IDLE && n >= 0 && n < S.nt;
    if (!v.active) return v;
    v.forward = S.phase == PH_FORWARD;
    v.u = S.ubuf[n & UM];
    v.U = S.ubuf[(n + 1) & UM];
    v.Uo = UM == 1 ? v.U : S.ubuf[(n - 1) & UM];
    v.psiXi = S.psiX[n & 1];
    v.psiXo = S.psiX[(n + 1) & 1];
    v.psiZi = S.psiZ[n & 1];
    v.psiZo = S.psiZ[(n + 1) & 1];
    const size_t nint = (size_t)S.nx * S.nz;
    if (v.forward) {
        v.src = n < S.nw && S.src_j >= 0;
        v.src_w = v.src ? R(ldg(S.wav, n)) : R(0);
        v.rec_row_out = (S.flags & F_RECORD) ? S.seis_rec + (size_t)n * S.nrec : nullptr;
        const bool store = (S.flags & F_SNAP) && n % S.TS == 0;
        if (I16) {   /* stored in the step that computes the snapshot */
            const int c = n / S.TS;
            v.qsnap = store ? S.q16 + (size_t)c * nint : nullptr;
            v.qexp = store ? S.qexp + (size_t)c * S.nz * S.nqb : nullptr;
        } else if (!Q16) {
            v.snap = store ? S.snaps + (size_t)(n / S.TS) * nint : nullptr;
        } else {
            v.smax = store ? S.smax + (size_t)(n / S.TS) * kSnapWords : nullptr;
            if ((S.flags & F_SNAP) && n >= 1 && (n - 1) % S.TS == 0) {
                const int c = (n - 1) / S.TS;
                v.qsnap = S.q16 + (size_t)c * nint;
                v.qscale = pow2f(14 - snap_exp(snap_max_bits(S.smax + (size_t)c * kSnapWords)));
            }
        }
    } else {
        const int tau = S.nt - 1 - n;
        v.seis_row = S.seis_t + (size_t)tau * S.nrec;
        const bool img = (S.flags & F_IMAGE) && tau % S.TS == 0;
        if (I16) {
            const int c = tau / S.TS;
            v.qsnap = img ? S.q16 + (size_t)c * nint : nullptr;
            v.qexp = img ? S.qexp + (size_t)c * S.nz * S.nqb : nullptr;
        } else if (!Q16) {
            /* both fields assigned from selected values: a store through a
             * selected field address would put the whole view in scratch */
            const bool ring = (S.flags & F_RING) != 0;
            const R *sf = S.sbuf[(S.nt - n) & 1];               /* u^{tau+1} */
            R *sn = S.snaps + (size_t)(tau / S.TS) * nint;
            v.sfield = (img && ring) ? sf : nullptr;
            v.snap = (img && !ring) ? sn : nullptr;
        } else if (img) {
            const int c = tau / S.TS;
            v.qsnap = S.q16 + (size_t)c * nint;
            v.qscale = pow2f(snap_exp(snap_max_bits(S.smax + (size_t)c * kSnapWords)) - 14);
        }
    }
    return v;
}

/* The view of a step that stores no snapshot and images nothing (the light
 * tile instances): no snapshot index, so no division by TS in the
 * prologue every wavefront runs. */
template <typename R, int UM = 1>
RTM_HD StepView<R> step_view_light(const SimDesc<R> &S, int n)
{
    /* every descriptor field read up front and combined without branches,
     * so the reads issue together instead of one dependent round trip per
     * condition; an inactive step (n outside [0, nt)) gets a view whose
     * reads stay in bounds (trace row clamped, wavelet sample 0) and the
     * caller stores nothing */
    const int phase = S.phase, nt = S.nt, nw = S.nw, src_j = S.src_j, flags = S.flags, nrec = S.nrec;
    StepView<R> v{};
    v.n = n;
    v.active = (phase != PH_IDLE) & (n >= 0) & (n < nt);
    v.forward = phase == PH_FORWARD;
    v.u = S.ubuf[n & UM];
    v.U = S.ubuf[(n + 1) & UM];
    v.Uo = UM == 1 ? v.U : S.ubuf[(n - 1) & UM];
    v.psiXi = S.psiX[n & 1];
    v.psiXo = S.psiX[(n + 1) & 1];
    v.psiZi = S.psiZ[n & 1];
    v.psiZo = S.psiZ[(n + 1) & 1];
    /* every field from selected values, no branch: fields stored on some
     * paths only would put the whole view in scratch memory */
    v.src = v.forward & (n < nw) & (src_j >= 0);
    const R w = R(ldg(S.wav, v.src ? n : 0));
    v.src_w = v.src ? w : R(0);
    const int nc = clampi(n, 0, nt - 1);
    float *rec = S.seis_rec + (size_t)nc * nrec;
    const float *seis = S.seis_t + (size_t)(nt - 1 - nc) * nrec;
    v.rec_row_out = (v.forward & ((flags & F_RECORD) != 0)) ? rec : nullptr;
    v.seis_row = v.forward ? nullptr : seis;
    return v;
}

template <typename R>
RTM_HD StepView<R> step_view(const SimDesc<R> &S, int n)
{
    if (S.umask == 3) return step_view_q<R, false, false, 3>(S, n);
    return S.qmode == Q_I16 ? step_view_q<R, false, true>(S, n)
           : S.qmode        ? step_view_q<R, true>(S, n)
                            : step_view_q<R, false>(S, n);
}

template <typename R>
RTM_HD bool x_strip(const SimDesc<R> &S, int i)
{
    return (i >= kHalf && i < S.TL - kHalf) || (i >= S.X - S.TR + kHalf && i < S.X - kHalf);
}

template <typename R>
RTM_HD bool z_strip(const SimDesc<R> &S, int k)
{
    return (k >= kHalf && k < S.TU - kHalf) || (k >= S.Z - S.TD + kHalf && k < S.Z - kHalf);
}

/* Class of the tile with origin (x0, z0): which strips its points touch. */
template <typename R>
RTM_HD int tile_class(const SimDesc<R> &S, int x0, int z0, int TZ)
{
    const int x1 = x0 + 64 < S.X - kHalf ? x0 + 64 : S.X - kHalf;   /* exclusive */
    const int z1 = z0 + TZ < S.Z - kHalf ? z0 + TZ : S.Z - kHalf;
    int cls = 0;
    if (x0 < S.TL - kHalf || x1 > S.X - S.TR + kHalf) cls |= T_XSTRIP;
    if (z0 < S.TU - kHalf || z1 > S.Z - S.TD + kHalf) cls |= T_ZSTRIP;
    return cls;
}

/* An interior tile inside the stepped region, halo included: columns
 * x0-4 .. x0+67 and rows z0-4 .. z0+TZ+3 lie in the grid and every point of
 * the tile is stepped (i < X-4, k < Z-4), so its loads need no clamps and
 * its points no bounds tests (the SAFE instances of the tile phases). */
template <typename R>
RTM_HD bool tile_safe(const SimDesc<R> &S, int x0, int z0, int TZ)
{
    return x0 >= kHalf && z0 >= kHalf && x0 + 64 + kHalf <= S.X && z0 + TZ + kHalf <= S.Z;
}

/* Per-thread values fetched from HBM before the first barrier, so that a
 * step costs one memory round trip: the thread's share of the u tile and the
 * C, U_old, csi, psi, profile, imaging and trace values its later phases
 * need.  Every load uses a clamped (always valid) address and its value is
 * selected afterwards: no load is predicated, which keeps the scalar unit
 * free of exec-mask bookkeeping. */
template <typename R, int RZ, int NW = 4>
struct TileRegs {
    static_assert(NW == 4 || NW == 8, "a tile has 4 or 8 row groups");
    R u0[RZ + 16 / NW], u1[RZ + 16 / NW];   /* u rows of the tile+halo, 2 column blocks */
    R C[RZ], Uold[RZ];          /* own points: rows ty*RZ+rr, column tx */
    R csx[RZ], csz[RZ];
    R px0[RZ], px1[RZ];         /* psiX^n at the psi-phase points       */
    R pz[RZ + 8 / NW];          /* psiZ^n at the psi-phase points       */
    R ax0, bx0, ax1, bx1;       /* x profiles at psi columns tx-4, 64+(tx&7)-4 */
    R axo, bxo;                 /* x profiles at the own column         */
    R sv[RZ], Iv[RZ], Isv[RZ];  /* imaging: snapshot and accumulators   */
    R Icv[RZ], Iscv[RZ];        /* Kahan compensations                  */
    int recr;                   /* receiver at the own column, or -1    */
    R recv;                     /* its trace sample (backward)          */
    R azt, bzt;                 /* z-strip tiles: the z profiles at row z0-4+thread (staged in LDS) */
    R uq[RZ];                   /* int16 block store: U at own image points, else 0 */
};

/* The imaging operands of the thread's points (backward, on imaging
 * steps): the snapshot value and the image accumulators.  Part of phase 1,
 * or loaded after the barriers (run_tile's LATE instances: their registers
 * then do not add to phase 1's, which keeps the occupancy of the batched
 * kernels up at the price of a second round trip on imaging steps). */
template <typename R, int RZ, bool Q16 = false, bool I16 = false, int NW = 4>
RTM_HD void tile_fetch_img(TileRegs<R, RZ, NW> &G, const SimDesc<R> &S, const StepView<R> &V, int x0, int z0,
                           int tx, int ty)
{
    const int i = x0 + tx;
    const int ic = clampi(i, 0, S.X - 1);
    if (!V.forward && ((Q16 || I16) ? V.qsnap != nullptr : (V.snap || V.sfield))) {
        const int ii = clampi(i - S.TL, 0, S.nx - 1);
        const int qb = clampi(i16_block(x0) - S.qb0, 0, S.nqb - 1);   /* the tile row's block (I16) */
#pragma unroll
        for (int rr = 0; rr < RZ; rr++) {
            const int kk = clampi(z0 + ty * RZ + rr - S.TU, 0, S.nz - 1);
            const int q = kk * S.nx + ii;
            if (I16) G.sv[rr] = R(dequant_i16((short)ldg(V.qsnap, q), ldg(V.qexp, kk * S.nqb + qb)));
            else if (Q16) G.sv[rr] = R(dequant16(ldg(V.qsnap, q), V.qscale, S.qmode));
            else if (V.sfield) G.sv[rr] = ldg(V.sfield, clampi(z0 + ty * RZ + rr, 0, S.Z - 1) * S.P + ic);
            else G.sv[rr] = ldg(V.snap, q);
            G.Iv[rr] = ldg(S.I, q);
            G.Isv[rr] = ldg(S.Is, q);
            if (S.flags & F_KAHAN) {
                G.Icv[rr] = ldg(S.Ic, q);
                G.Iscv[rr] = ldg(S.Isc, q);
            }
        }
    }
}

/* Phase 1a: every global load of the step (the imaging operands unless
 * IMG_EARLY is false).  tx in [0,64), ty in [0,4) (the wavefront index:
 * uniform within a wavefront).  CLS is the tile class.  SAFE: an interior
 * tile whose halo lies inside the grid (tile_safe), read without clamps or
 * masks. */
template <typename R, int RZ, int CLS, bool Q16 = false, bool I16 = false, int NW = 4, bool IMG_EARLY = true,
          bool SAFE = false>
RTM_HD void tile_fetch(TileRegs<R, RZ, NW> &G, const SimDesc<R> &S, const StepView<R> &V, int x0, int z0,
                       int tx, int ty)
{
    static_assert(!SAFE || CLS == 0, "only interior tiles take the unclamped path");
    constexpr int HX = (CLS & T_XSTRIP) ? 8 : 4, HZ = (CLS & T_ZSTRIP) ? 8 : 4;
    constexpr int NR = RZ + 2 * HZ / NW;             /* (TZ + 2 HZ) / NW rows per thread */
    const int i = x0 + tx;
    const int ic = SAFE ? i : clampi(i, 0, S.X - 1);
    /* the receiver column first: the trace sample it selects is read after
     * the field loads, and waiting for this load then leaves them in flight */
    G.recr = -1;
    G.recv = R(0);
    const int rrow = S.rec_row - (z0 + ty * RZ);
    const bool on_rec = rrow >= 0 && rrow < RZ && (!V.forward || V.rec_row_out);   /* uniform */
    if (on_rec) G.recr = (SAFE || i < S.X - kHalf) ? ldg(S.recmap, ic) : -1;
    /* column blocks: 64 columns, then the 2*HX remaining ones (lanes repeat) */
    const int ia = x0 - HX + tx, ib = x0 - HX + 64 + (tx % (2 * HX));
    const bool oka = SAFE || (ia >= 0 && ia < S.X), okb = SAFE || ib < S.X;
    const int ca = SAFE ? ia : clampi(ia, 0, S.X - 1), cb = SAFE ? ib : clampi(ib, 0, S.X - 1);
#pragma unroll
    for (int t = 0; t < NR; t++) {
        const int k = z0 - HZ + ty + NW * t;
        const bool okk = SAFE || (k >= 0 && k < S.Z);
        const int base = (SAFE ? k : clampi(k, 0, S.Z - 1)) * S.P;
        const R a = ldg(V.u, base + ca), b = ldg(V.u, base + cb);
        G.u0[t] = (okk && oka) ? a : R(0);
        G.u1[t] = (okk && okb) ? b : R(0);
    }
    const bool inX = i < S.TL - kHalf || (i >= S.X - S.TR + kHalf && i < S.X - kHalf);
#pragma unroll
    for (int rr = 0; rr < RZ; rr++) {
        const int k = SAFE ? z0 + ty * RZ + rr : clampi(z0 + ty * RZ + rr, 0, S.Z - 1);
        const int j = k * S.P + ic;
        G.C[rr] = ldg(S.C, j);
        G.Uold[rr] = ldg(V.Uo, j);
        if (CLS & T_XSTRIP) G.csx[rr] = ldg(S.csiX, j);
        if (CLS & T_ZSTRIP) G.csz[rr] = ldg(S.csiZ, j);
    }
    if (CLS & T_XSTRIP) {
        G.axo = inX ? ldg(S.ax, ic) : R(0);
        G.bxo = inX ? ldg(S.bx, ic) : R(0);
        const int pa = x0 - 4 + tx, pb = x0 + 60 + (tx & 7);   /* psi columns c = tx, 64 + (tx & 7) */
        const bool sa = x_strip(S, pa), sb = x_strip(S, pb);
        const int qa = clampi(pa, 0, S.X - 1), qb = clampi(pb, 0, S.X - 1);
        const R a0 = ldg(S.ax, qa), b0 = ldg(S.bx, qa), a1 = ldg(S.ax, qb), b1 = ldg(S.bx, qb);
        G.ax0 = sa ? a0 : R(0);
        G.bx0 = sa ? b0 : R(0);
        G.ax1 = sb ? a1 : R(0);
        G.bx1 = sb ? b1 : R(0);
#pragma unroll
        for (int t = 0; t < RZ; t++) {
            const int k = z0 + ty + NW * t;
            const bool kok = k < S.Z - kHalf;
            const int base = clampi(k, 0, S.Z - 1) * S.P;
            const R va = ldg(V.psiXi, base + qa), vb = ldg(V.psiXi, base + qb);
            G.px0[t] = (kok && sa) ? va : R(0);
            G.px1[t] = (kok && sb) ? vb : R(0);
        }
    }
    if (CLS & T_ZSTRIP) {
        const bool iok = i < S.X - kHalf;
#pragma unroll
        for (int t = 0; t < RZ + 8 / NW; t++) {
            const int k = z0 - 4 + ty + NW * t;
            const R v = ldg(V.psiZi, clampi(k, 0, S.Z - 1) * S.P + ic);
            G.pz[t] = (iok && z_strip(S, k)) ? v : R(0);
        }
        /* the z profiles of the tile's rows and psi halo, read once here
         * instead of after the barriers, row by row (thread ty*64+tx takes
         * row z0-4+ty*64+tx; the rows past TZ+8 are not staged) */
        const int kp = clampi(z0 - 4 + ty * 64 + tx, 0, S.Z - 1);
        G.azt = ldg(S.az, kp);
        G.bzt = ldg(S.bz, kp);
    }
    if (IMG_EARLY) tile_fetch_img<R, RZ, Q16, I16, NW>(G, S, V, x0, z0, tx, ty);
    if (on_rec && !V.forward) {   /* uniform */
        const R v = R(ldg(V.seis_row, G.recr > 0 ? G.recr : 0));
        G.recv = G.recr >= 0 ? v : R(0);
    }
}

/* Phase 1b: the fetched u into LDS (row z0-8+r, column x0-8+c at [r][c]). */
template <typename R, int RZ, int CLS, int NW = 4>
RTM_HD void tile_stage(TileLds<R, RZ, NW> &L, const TileRegs<R, RZ, NW> &G, int tx, int ty)
{
    constexpr int HX = (CLS & T_XSTRIP) ? 8 : 4, HZ = (CLS & T_ZSTRIP) ? 8 : 4;
    constexpr int NR = RZ + 2 * HZ / NW;
    const int cb = 64 + (tx % (2 * HX)) + 8 - HX;   /* lanes beyond 2*HX repeat a column: same value */
#pragma unroll
    for (int t = 0; t < NR; t++) {
        const int r = ty + NW * t + 8 - HZ;
        L.u[r][tx + 8 - HX] = G.u0[t];
        L.u[r][cb] = G.u1[t];
    }
    if ((CLS & T_ZSTRIP) && ty == 0 && tx < NW * RZ + 8) {
        L.az[tx] = G.azt;
        L.bz[tx] = G.bzt;
    }
}

/* Phase 2 (strip classes): psi^{n+1} over the tile plus a radius-4 halo. */
template <typename R, int RZ, int CLS, int NW = 4>
RTM_HD void tile_psi(TileLds<R, RZ, NW> &L, const TileRegs<R, RZ, NW> &G, const SimDesc<R> &S, int x0, int z0,
                     int tx, int ty)
{
    if (CLS & T_XSTRIP) {
        const int cb = 64 + (tx & 7);
        const bool sa = x_strip(S, x0 - 4 + tx), sb = x_strip(S, x0 - 4 + cb);
#pragma unroll
        for (int t = 0; t < RZ; t++) {
            const int r = ty + NW * t;
            const bool kok = z0 + r < S.Z - kHalf;
            const R *la = &L.u[r + 8][tx + 4];
            const R *lb = &L.u[r + 8][cb + 4];
            const R da = d1_val(la[1], la[2], la[3], la[4], la[-1], la[-2], la[-3], la[-4]);
            const R db = d1_val(lb[1], lb[2], lb[3], lb[4], lb[-1], lb[-2], lb[-3], lb[-4]);
            const R va = psi_val(G.ax0, G.bx0, G.px0[t], da, S.h1);
            const R vb = psi_val(G.ax1, G.bx1, G.px1[t], db, S.h1);
            L.px[r][tx] = (kok && sa) ? va : R(0);
            L.px[r][cb] = (kok && sb) ? vb : R(0);
        }
    }
    if (CLS & T_ZSTRIP) {
        const bool iok = x0 + tx < S.X - kHalf;
        const int lc = tx + 8;
#pragma unroll
        for (int t = 0; t < RZ + 8 / NW; t++) {
            const int r = ty + NW * t;
            const int k = z0 - 4 + r;
            R v = R(0);
            if (z_strip(S, k)) {   /* uniform: k is the same for the whole wavefront */
                const int lr = r + 4;
                const R d1 = d1_val(L.u[lr + 1][lc], L.u[lr + 2][lc], L.u[lr + 3][lc], L.u[lr + 4][lc],
                                    L.u[lr - 1][lc], L.u[lr - 2][lc], L.u[lr - 3][lc], L.u[lr - 4][lc]);
                const R pv = psi_val(L.az[r], L.bz[r], G.pz[t], d1, S.h1);   /* row k = z0-4+r */
                v = iok ? pv : R(0);
            }
            L.pz[r][tx] = v;
        }
    }
}

/* Phase 3: the fused update of this thread's RZ points. */
/* XS: what the instance carries: 2 the snapshot store and the imaging, 1
 * the snapshot store only (forward store steps), 0 neither (a light
 * instance for steps that have neither, with fewer registers).
 * SAFE: an interior tile inside the stepped region (tile_safe).  Nothing is
 * stored on an inactive step (a light instance runs those through). */
template <typename R, int RZ, int CLS, bool Q16 = false, bool I16 = false, int NW = 4, int XS = 2,
          bool SAFE = false>
RTM_HD float tile_points(const TileLds<R, RZ, NW> &L, TileRegs<R, RZ, NW> &G, const SimDesc<R> &S,
                         const StepView<R> &V, int x0, int z0, int tx, int ty)
{
    constexpr int HX = (CLS & T_XSTRIP) ? 8 : 4;
    float umax = 0.0f;   /* max |U| over the thread's image points (16-bit store) */
    (void)HX;
    const int i = x0 + tx;
    const bool iok = (SAFE || i < S.X - kHalf) && V.active;
    const int lc = tx + 8;
    /* this column's u from tile row ty*RZ-4 to ty*RZ+RZ+3 */
    R col[RZ + 8];
#pragma unroll
    for (int t = 0; t < RZ + 8; t++) col[t] = L.u[ty * RZ + 4 + t][lc];
    const bool inL = i < S.TL - kHalf, inR = i >= S.X - S.TR + kHalf;
    const bool img_col = i - S.TL >= 0 && i - S.TL < S.nx;
    if (I16) {
#pragma unroll
        for (int rr = 0; rr < RZ; rr++) G.uq[rr] = R(0);
    }
#pragma unroll
    for (int rr = 0; rr < RZ; rr++) {
        const int kl = ty * RZ + rr;         /* tile-relative row (uniform) */
        const int k = z0 + kl;
        if (!SAFE && k >= S.Z - kHalf) break;
        const int j = k * S.P + i;
        const R *lu = &L.u[kl + 8][lc];
        const R uc = col[rr + 4];
        const R xm1 = lu[-1], xp1 = lu[1], xm2 = lu[-2], xp2 = lu[2];
        const R xm3 = lu[-3], xp3 = lu[3], xm4 = lu[-4], xp4 = lu[4];
        const R zm1 = col[rr + 3], zp1 = col[rr + 5], zm2 = col[rr + 2], zp2 = col[rr + 6];
        const R zm3 = col[rr + 1], zp3 = col[rr + 7], zm4 = col[rr + 0], zp4 = col[rr + 8];
        const R Cj = G.C[rr];
        R U = stencil_val(Cj, G.Uold[rr], uc, xm1, xp1, zm1, zp1, xm2, xp2, zm2, zp2, xm3, xp3, zm3, zp3,
                          xm4, xp4, zm4, zp4);
        R corrX = R(0), corrZ = R(0), csx = R(0), csz = R(0), pxo = R(0), pzo = R(0);
        const bool inT = k < S.TU - kHalf, inB = k >= S.Z - S.TD + kHalf;   /* uniform */
        if (CLS & T_XSTRIP) {
            const R *p = &L.px[kl][tx + 4];
            pxo = p[0];
            const R dpsi = d1_val(p[1], p[2], p[3], p[4], p[-1], p[-2], p[-3], p[-4]) * S.h1;
            const R d2 = d2_val(uc, xp1, xp2, xp3, xp4, xm1, xm2, xm3, xm4);
            csx = csi_val(G.axo, G.bxo, G.csx[rr], d2, dpsi, S.h3);
            corrX = corr_val(Cj, S.h2, dpsi, csx);
        }
        if ((CLS & T_ZSTRIP) && (inT || inB)) {
            const int pr = kl + 4;
            pzo = L.pz[pr][tx];
            const R dpsi = d1_val(L.pz[pr + 1][tx], L.pz[pr + 2][tx], L.pz[pr + 3][tx], L.pz[pr + 4][tx],
                                  L.pz[pr - 1][tx], L.pz[pr - 2][tx], L.pz[pr - 3][tx], L.pz[pr - 4][tx])
                           * S.h1;
            const R d2 = d2_val(uc, zp1, zp2, zp3, zp4, zm1, zm2, zm3, zm4);
            csz = csi_val(L.az[kl + 4], L.bz[kl + 4], G.csz[rr], d2, dpsi, S.h3);   /* row k = z0+kl */
            corrZ = corr_val(Cj, S.h2, dpsi, csz);
        }
        /* additions in the reference's strip order: L, T, R, B */
        if (CLS & T_XSTRIP) U = inL ? U + corrX : U;
        if ((CLS & T_ZSTRIP) && inT) U = U + corrZ;
        if (CLS & T_XSTRIP) U = inR ? U + corrX : U;
        if ((CLS & T_ZSTRIP) && inB) U = U + corrZ;
        const bool rec_row = k == S.rec_row && G.recr >= 0;
        if (V.forward) {
            if (V.src && j == S.src_j) U = U + V.src_w * Cj;
        } else {
            U = rec_row ? G.recv : U;
        }
        if (iok) {
            stg(V.U, j, U);
            if ((CLS & T_XSTRIP) && (inL || inR)) {
                stg(V.psiXo, j, pxo);
                stg(S.csiX, j, csx);
            }
            if ((CLS & T_ZSTRIP) && (inT || inB)) {
                stg(V.psiZo, j, pzo);
                stg(S.csiZ, j, csz);
            }
            if (V.forward && rec_row) stg(V.rec_row_out, G.recr, float(U));
            const int kk = k - S.TU;
            const bool img_pt = img_col && kk >= 0 && kk < S.nz;
            if (!XS) {
                /* a light instance: no store and no imaging on this step */
            } else if (I16 && V.forward) {
                /* the int16 block store follows the update (tile_store_i16) */
                if (img_pt) G.uq[rr] = U;
            } else if (Q16 && V.forward) {
                /* the snapshot max, kept on every step (the caller commits it on
                 * store steps only); the 16-bit store of u^n follows the loop */
                if (img_pt) umax = fmaxf(umax, fabsf(float(U)));
            } else if (((Q16 || I16) ? V.qsnap != nullptr : (V.snap || V.sfield)) && img_pt) {
                const int q = kk * S.nx + (i - S.TL);
                if (V.forward) {
                    if (V.snap) stg_stream(V.snap, q, U);
                } else if constexpr (XS == 2) {
                    if (S.flags & F_KAHAN) {
                        R I = G.Iv[rr], Ic = G.Icv[rr], Is = G.Isv[rr], Isc = G.Iscv[rr];
                        const R s = G.sv[rr];
                        kahan_add(I, Ic, s * U);
                        kahan_add(Is, Isc, s * s);
                        stg(S.I, q, I); stg(S.Ic, q, Ic); stg(S.Is, q, Is); stg(S.Isc, q, Isc);
                    } else {
                        const R s = G.sv[rr];
                        stg(S.I, q, G.Iv[rr] + s * U);
                        stg(S.Is, q, G.Isv[rr] + s * s);
                    }
                }
            }
        }
    }
    if (XS && Q16 && V.forward && V.qsnap && img_col) {
        /* 16-bit store of u^n (the previous step's snapshot) from the tile's
         * LDS copy, apart from the update loop: kept inside it, u^n and the
         * store's operands stayed live across the loop and slowed every step */
#pragma unroll
        for (int rr = 0; rr < RZ; rr++) {
            const int k = z0 + ty * RZ + rr;
            const int kk = k - S.TU;
            if (k >= S.Z - kHalf) break;
            if (kk >= 0 && kk < S.nz)
                stg_stream(V.qsnap, kk * S.nx + (i - S.TL),
                           quant16(float(L.u[ty * RZ + rr + 8][lc]), V.qscale, S.qmode));
        }
    }
    return umax;
}

/* Forward, int16 block store: the thread's image points of this step's
 * snapshot.  rmax[rr] is the max |uq[rr]| over the wavefront (one block:
 * the tile row's 64 columns); lane 0 writes the block's exponent. */
template <typename R, int RZ, int NW = 4>
RTM_HD void tile_store_i16(const TileRegs<R, RZ, NW> &G, const SimDesc<R> &S, const StepView<R> &V, const float *rmax,
                           int x0, int z0, int tx, int ty)
{
    const int i = x0 + tx;
    const bool img_col = i - S.TL >= 0 && i - S.TL < S.nx && i < S.X - kHalf;
    const bool blk_img = x0 + kQBlock > S.TL && x0 < S.TL + S.nx;   /* the block holds image columns */
    const int qb = i16_block(x0) - S.qb0;
#pragma unroll
    for (int rr = 0; rr < RZ; rr++) {
        const int k = z0 + ty * RZ + rr, kk = k - S.TU;
        if (k >= S.Z - kHalf) break;
        if (kk < 0 || kk >= S.nz) continue;   /* uniform */
        const int e = i16_exp(rmax[rr]);
        if (img_col)
            stg_stream(V.qsnap, kk * S.nx + (i - S.TL), (unsigned short)quant_i16(float(G.uq[rr]), pow2f(15 - e)));
        if (tx == 0 && blk_img) stg(V.qexp, kk * S.nqb + qb, (signed char)e);
    }
}

}  // namespace rtm

#endif /* RTM_TILE_H */
