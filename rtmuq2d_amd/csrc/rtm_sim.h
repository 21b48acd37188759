// rtm_sim.h -- device-side building blocks shared by the plan (one
// simulation) and the batch/serving layers (many simulations per launch):
// kernels, the per-simulation buffer set, and the step executor that walks a
// tile list with one launch per time step (optionally replayed from a
// hipGraph of a block of steps).  Internal to librtm_hip.so.
#ifndef RTM_SIM_H
#define RTM_SIM_H

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "rtm_internal.h"
#include "rtm_step.h"
#include "rtm_tile.h"
#include "rtm_column.h"
#include "rtm_work.h"
#include "rtm_pair.h"

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            rtm_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),     \
                          __FILE__, __LINE__);                                         \
            return -1;                                                                 \
        }                                                                              \
    } while (0)

namespace rtmd {

using rtm::SimDesc;
using rtm::StepArgs;

constexpr int kBX = 64;   // v1 kernel block: one wavefront along x ...
constexpr int kBZ = 4;    // ... by 4 rows

inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// The tiled step: one workgroup per tile of the flat tile list, any number of
// simulations per launch.  Workgroup b takes tile (b % 8) * per + b / 8 so
// that each XCD (workgroups are dealt round-robin to the 8 XCDs) walks a
// contiguous run of neighbouring tiles and their shared halo rows stay in its
// L2.  `s` is the step within the current block; sim i is at n0[i] + s.
// A snapshot's max |U| for the 16-bit store: wavefront max by butterfly
// shuffles, then a (vector) atomic max on the float bits, whose unsigned
// order is the order of non-negative floats, into one of the snapshot's
// slots (rtm::kSnapSlots, picked by `who`).
__device__ __forceinline__ float wave_max(float m)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    return m;
}

__device__ __forceinline__ void commit_snap_max(unsigned *slots, float m, int who)
{
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomicMax(slots + (who % rtm::kSnapSlots) * rtm::kSlotStride, __float_as_uint(m));
}

// the same for a workgroup of NW wavefronts: one atomic per workgroup
template <int NW = 4>
__device__ __forceinline__ void commit_snap_max_wg(unsigned *slots, float m, float *red)
{
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = red[0];
#pragma unroll
        for (int w = 1; w < NW; w++) t = fmaxf(t, red[w]);
        atomicMax(slots + (blockIdx.x % rtm::kSnapSlots) * rtm::kSlotStride, __float_as_uint(t));
    }
}

// KIND: which steps the instance serves.
//   K_LIGHT  steps that neither store nor image (no store or imaging code:
//            fewer registers, a higher occupancy, and a prologue without
//            the snapshot index).  It has no early exit for an inactive
//            step (its loads stay in bounds and tile_points stores
//            nothing), so its descriptor reads issue together.
//   K_STORE  forward steps that store a snapshot (no imaging code or
//            imaging operands); meeting an imaging step sets err bit 1.
//   K_FULL   every step.
// The executor's schedule (Exec::heavy_step) and the store / imaging tests
// of rtm::step_view_q are the same formulas.
// SAFE: an interior tile inside the stepped region (rtm::tile_safe).
enum StepKind : int { K_LIGHT = 0, K_STORE = 1, K_FULL = 2 };

template <typename R, int RZ, int CLS, bool Q16, bool I16 = false, int NW = 4, int UM = 1, int KIND = K_FULL,
          bool SAFE = false>
__device__ __forceinline__ void run_tile(rtm::TileLds<R, RZ, NW> &L, float *red, const SimDesc<R> &S, int n,
                                         int x0, int z0, int tx, int ty, int *err)
{
    constexpr bool LIGHT = KIND == K_LIGHT;
    const rtm::StepView<R> V = LIGHT ? rtm::step_view_light<R, UM>(S, n) : rtm::step_view_q<R, Q16, I16, UM>(S, n);
    if (!LIGHT && !V.active) return;
    if (KIND == K_STORE && !V.forward && (V.snap || V.sfield || V.qsnap)) {   // uniform: an imaging step
        if (threadIdx.x == 0 && err) atomicOr(err, 1);
        return;
    }
    rtm::TileRegs<R, RZ, NW> G;
    rtm::tile_fetch<R, RZ, CLS, Q16, I16, NW, KIND == K_FULL, SAFE>(G, S, V, x0, z0, tx, ty);
    rtm::tile_stage<R, RZ, CLS, NW>(L, G, tx, ty);
    __syncthreads();
    if (CLS) {
        rtm::tile_psi<R, RZ, CLS, NW>(L, G, S, x0, z0, tx, ty);
        __syncthreads();
    }
    const float m = rtm::tile_points<R, RZ, CLS, Q16, I16, NW, KIND, SAFE>(L, G, S, V, x0, z0, tx, ty);
    if (LIGHT) return;
    if (Q16 && V.smax) commit_snap_max_wg<NW>(V.smax, m, red);
    if (I16 && V.forward && V.qsnap) {   /* uniform: a store step */
        float rmax[RZ];
#pragma unroll
        for (int rr = 0; rr < RZ; rr++) rmax[rr] = wave_max(fabsf(float(G.uq[rr])));   /* the row's block */
        rtm::tile_store_i16<R, RZ, NW>(G, S, V, rmax, x0, z0, tx, ty);
    }
}

template <typename R, int RZ, bool Q16, bool I16 = false, int NW = 4, int UM = 1, int KIND = K_FULL>
__device__ __forceinline__ void run_tile_cls(rtm::TileLds<R, RZ, NW> &L, float *red, const SimDesc<R> &S, int n,
                                             int4 tile, int tx, int ty, int *err)
{
    switch (tile.w) {
    case 0:
        if (rtm::tile_safe(S, tile.y, tile.z, NW * RZ))
            run_tile<R, RZ, 0, Q16, I16, NW, UM, KIND, true>(L, red, S, n, tile.y, tile.z, tx, ty, err);
        else
            run_tile<R, RZ, 0, Q16, I16, NW, UM, KIND>(L, red, S, n, tile.y, tile.z, tx, ty, err);
        break;
    case 1: run_tile<R, RZ, 1, Q16, I16, NW, UM, KIND>(L, red, S, n, tile.y, tile.z, tx, ty, err); break;
    case 2: run_tile<R, RZ, 2, Q16, I16, NW, UM, KIND>(L, red, S, n, tile.y, tile.z, tx, ty, err); break;
    default:
        if constexpr (rtm::corner_tiles_small(RZ, NW)) {
            // the lists cut these corner tiles into 64 x 4 tiles (its LDS
            // holds one psi plane): flag the list as broken
            if (threadIdx.x == 0 && err) atomicOr(err, 2);
        } else {
            run_tile<R, RZ, 3, Q16, I16, NW, UM, KIND>(L, red, S, n, tile.y, tile.z, tx, ty, err);
        }
        break;
    }
}

// A step kernel's LDS: the tile's, and, where corner tiles come as 64 x 4
// tiles (rtm::corner_tiles_small), the small tile's in a union with it.
template <typename R, int RZ, int NW, bool SMALL = rtm::corner_tiles_small(RZ, NW)>
struct StepLds {
    rtm::TileLds<R, RZ, NW> full;
};

template <typename R, int RZ, int NW>
struct StepLds<R, RZ, NW, true> {
    union {
        rtm::TileLds<R, 1> small;
        rtm::TileLds<R, RZ, NW> full;
    };
};

// Q16: every sim of the launch keeps a 16-bit snapshot store (the store
// kind is a property of the whole plan or batch, so the FP32-store kernels
// carry none of that code).
// ONE: the launch steps a single simulation (a plan), so its descriptor's
// address is known before the tile entry arrives and the descriptor loads
// need not wait for it.
// I16: every sim of the launch keeps an int16 block store (FP32 stepping).
// NW: wavefronts per workgroup, one row group each (tiles of 64 x NW*RZ).
// UM: u buffer mask, 1 (two buffers, in-place update) or 3 (four buffers:
// the tiles beside two-step pair items).
// KIND: the steps the instance serves (run_tile); err: its flags.
template <typename R, int RZ, bool Q16, bool ONE = false, bool I16 = false, int NW = 4, int UM = 1, int KIND = K_FULL>
__global__ void __launch_bounds__(64 * NW) k_tiles(const SimDesc<R> *__restrict__ sims,
                                                   const int *__restrict__ n0,
                                                   const int4 *__restrict__ tiles,
                                                   const int *__restrict__ ntiles_p, int s, int *err)
{
    __shared__ StepLds<R, RZ, NW> L;
    __shared__ float red[NW];
    const int ntiles = *ntiles_p;
    const int per = (ntiles + 7) >> 3;
    const int b = blockIdx.x;
    const int t = (b & 7) * per + (b >> 3);
    if ((b >> 3) >= per || t >= ntiles) return;
    const int4 tile = tiles[t];
    if (tile.x < 0) return;   // padding of an XCD's share (balance_tiles_for_xcds)
    // read through the restrict pointer: uniform fields become scalar loads
    // that no store of the step can invalidate
    const int sim = ONE ? 0 : tile.x;
    const SimDesc<R> &S = sims[sim];
    const int n = n0[sim] + s;
    const int tx = threadIdx.x & 63;
    const int ty = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // one row group per wavefront
    if constexpr (rtm::corner_tiles_small(RZ, NW)) {
        if (tile.w & rtm::T_SMALL) {   // a 64 x 4 piece of a corner tile: one row per thread
            const int4 t4 = make_int4(tile.x, tile.y, tile.z, tile.w & 3);
            run_tile_cls<R, 1, Q16, I16, 4, UM, KIND>(L.small, red, S, n, t4, tx, ty, err);
            return;
        }
    }
    run_tile_cls<R, RZ, Q16, I16, NW, UM, KIND>(L.full, red, S, n, tile, tx, ty, err);
}

// A mixed tile list (rtm::IN_MIXED): 64 x 4 tiles (T_SMALL) of any class run
// one row per thread; full 64 x 4RZ tiles are interior ones.  The two shapes
// share the LDS.
template <typename R, int RZ, bool Q16>
__global__ void __launch_bounds__(256) k_tiles_mix(const SimDesc<R> *__restrict__ sims,
                                                   const int *__restrict__ n0,
                                                   const int4 *__restrict__ tiles,
                                                   const int *__restrict__ ntiles_p, int s)
{
    __shared__ union {
        rtm::TileLds<R, 1> small;
        rtm::TileLds<R, RZ> full;
    } L;
    __shared__ float red[4];
    const int ntiles = *ntiles_p;
    const int per = (ntiles + 7) >> 3;
    const int b = blockIdx.x;
    const int t = (b & 7) * per + (b >> 3);
    if ((b >> 3) >= per || t >= ntiles) return;
    int4 tile = tiles[t];
    if (tile.x < 0) return;   // padding of an XCD's share (balance_tiles_for_xcds)
    const SimDesc<R> &S = sims[tile.x];
    const int n = n0[tile.x] + s;
    const int tx = threadIdx.x & 63;
    const int ty = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (tile.w & rtm::T_SMALL) {
        tile.w &= 3;
        run_tile_cls<R, 1, Q16>(L.small, red, S, n, tile, tx, ty, nullptr);
    } else {
        run_tile<R, RZ, 0, Q16>(L.full, red, S, n, tile.y, tile.z, tx, ty, nullptr);
    }
}

// Wavefront shifts by one lane (DPP wave_shr:1 / wave_shl:1): lane l gets
// lane l-1's (l+1's) value; lane 0 (63) keeps `old`.
__device__ __forceinline__ int dpp_shr1_i(int v, int old)
{
    return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int dpp_shl1_i(int v, int old)
{
    return __builtin_amdgcn_update_dpp(old, v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ float lane_shr1(float v, float old)
{
    return __builtin_bit_cast(float, dpp_shr1_i(__builtin_bit_cast(int, v), __builtin_bit_cast(int, old)));
}
__device__ __forceinline__ float lane_shl1(float v, float old)
{
    return __builtin_bit_cast(float, dpp_shl1_i(__builtin_bit_cast(int, v), __builtin_bit_cast(int, old)));
}
__device__ __forceinline__ double lane_shr1(double v, double old)
{
    const int2 a = __builtin_bit_cast(int2, v), o = __builtin_bit_cast(int2, old);
    return __builtin_bit_cast(double, make_int2(dpp_shr1_i(a.x, o.x), dpp_shr1_i(a.y, o.y)));
}
__device__ __forceinline__ double lane_shl1(double v, double old)
{
    const int2 a = __builtin_bit_cast(int2, v), o = __builtin_bit_cast(int2, old);
    return __builtin_bit_cast(double, make_int2(dpp_shl1_i(a.x, o.x), dpp_shl1_i(a.y, o.y)));
}

// One lane's per-row operands for four rows: C, U_old and the halo quad.
template <typename R>
struct ColSet {
    rtm::Pair<R> C[4], O[4];
    rtm::Quad<R> H[4];
};

// Per-wavefront constants of a column item.
template <typename R>
struct ColCtx {
    const rtm::Pair<R> *u2, *C2;
    rtm::Pair<R> *U2;
    const rtm::Quad<R> *u4;
    int Ph, Pq;     // row pitch in pairs and in quads
    int xh, hq;     // the lane's pair index and halo quad index within a row
    int xl;         // the lane's first column
    bool ok;        // the lane's columns belong to the item
    int srow;       // source row this step (-1: none)
};

// Loads of chunk rows k0 .. k0+3: u rows k0+4 .. k0+7 into window slots
// (WS+rr) & 15, and C, U_old, halo of the rows themselves into `st`.  Rows
// past the item (the prefetch after its last chunk) are clamped to its last
// row: the loads are always issued, so every chunk waits for exactly the
// loads of the previous prefetch and never for the one just issued.
template <typename R, int WS>
__device__ __forceinline__ void col_fetch(rtm::Pair<R> (&w)[16], ColSet<R> &st, const ColCtx<R> &X, int k0,
                                          int zB)
{
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        const int k = k0 + rr < zB ? k0 + rr : zB - 1;
        w[(WS + rr) & 15] = rtm::ldg(X.u2 + (k + 4) * X.Ph, X.xh);
        st.C[rr] = rtm::ldg(X.C2 + k * X.Ph, X.xh);
        st.O[rr] = rtm::ldg((const rtm::Pair<R> *)X.U2 + k * X.Ph, X.xh);
        st.H[rr] = rtm::ldg(X.u4 + k * X.Pq, X.hq);
    }
}

// Rows k0 .. k0+3; row k0+rr's window rows k-4 .. k+4 are slots
// (B + rr + j) & 15, j = 0..8.  EXTRAS: some row of the chunk has a source,
// receiver, snapshot or imaging term.
template <typename R, int B, bool EXTRAS, bool Q16, bool IMG>
__device__ __forceinline__ void col_chunk(const rtm::Pair<R> (&w)[16], const ColSet<R> &st, const ColCtx<R> &X,
                                          const SimDesc<R> &S, const rtm::StepView<R> &V,
                                          const rtm::ColRec<R> &rec, int k0, float &umax)
{
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        const int k = k0 + rr;
        rtm::Pair<R> win[9];
#pragma unroll
        for (int j = 0; j < 9; j++) win[j] = w[(B + rr + j) & 15];
        const rtm::Pair<R> c = win[4];
        const rtm::Quad<R> H = st.H[rr];
        rtm::Pair<R> L1, L2, R1, R2;
        L1.a = lane_shr1(c.a, H.z);
        L1.b = lane_shr1(c.b, H.w);
        L2.a = lane_shr1(L1.a, H.x);
        L2.b = lane_shr1(L1.b, H.y);
        R1.a = lane_shl1(c.a, H.x);
        R1.b = lane_shl1(c.b, H.y);
        R2.a = lane_shl1(R1.a, H.z);
        R2.b = lane_shl1(R1.b, H.w);
        rtm::Pair<R> U = rtm::col_update(win, L1, L2, R1, R2, st.C[rr], st.O[rr]);
        if (EXTRAS) rtm::col_extras<R, Q16, IMG>(S, V, rec, k, X.xl, X.ok, X.srow, U, st.C[rr], c, umax);
        if (X.ok) rtm::stg(X.U2 + k * X.Ph, X.xh, U);
    }
}

// Chunk c of the four-chunk cycle (rows k0 .. k0+3): prefetch the next
// chunk into the other set, then compute this one.
template <typename R, int CI, bool Q16, bool IMG>
__device__ __forceinline__ void col_step(rtm::Pair<R> (&w)[16], ColSet<R> &cur, ColSet<R> &nxt, const ColCtx<R> &X,
                                         const SimDesc<R> &S, const rtm::StepView<R> &V, const rtm::ColRec<R> &rec,
                                         int k0, int zB, int ex_lo, int ex_hi, float &umax)
{
    col_fetch<R, (4 * CI + 12) & 15>(w, nxt, X, k0 + 4, zB);
    // rows [ex_lo, ex_hi) carry snapshot/imaging terms; single rows the
    // source or receiver terms
    const bool extras = (k0 < ex_hi && k0 + 4 > ex_lo) || (X.srow >= k0 && X.srow < k0 + 4)
                        || (S.rec_row >= k0 && S.rec_row < k0 + 4 && (!V.forward || V.rec_row_out));
    if (extras) col_chunk<R, 4 * CI, true, Q16, IMG>(w, cur, X, S, V, rec, k0, umax);
    else col_chunk<R, 4 * CI, false, Q16, IMG>(w, cur, X, S, V, rec, k0, umax);
}

// The interior column kernel (rtm_column.h): four independent wavefronts
// per workgroup, each marching one item {sim(+half), x0, zA, zB} with the
// next four rows in flight.  The register window holds 16 rows (12 in use,
// 4 arriving) and its slots cycle every four chunks, so the loop is
// unrolled by four chunks and nothing is copied between registers.  Blocks
// are remapped so that each XCD walks a contiguous run of items
// (consecutive segments of a column share rows).
// One column item (the body of k_columns), for a sim without (Q16 false)
// or with a 16-bit snapshot store; IMG: the instance for imaging steps.
template <typename R, bool Q16, bool IMG>
__device__ __forceinline__ void column_march(const SimDesc<R> &S, const rtm::StepView<R> &V, int4 it, int item)
{
    const int lane = threadIdx.x & 63;
    const int x0 = it.y, zA = it.z, zB = it.w;
    const int P = S.P;
    ColCtx<R> X;
    X.u2 = (const rtm::Pair<R> *)V.u;
    X.C2 = (const rtm::Pair<R> *)S.C;
    X.U2 = (rtm::Pair<R> *)V.U;
    X.u4 = (const rtm::Quad<R> *)V.u;
    X.Ph = P >> 1;
    X.Pq = P >> 2;
    X.xl = x0 + 2 * lane;
    X.xh = (X.xl < P - 2 ? X.xl : P - 2) >> 1;
    X.hq = rtm::clampi(lane == 63 ? x0 + 128 : x0 - 4, 0, P - 4) >> 2;
    X.ok = lane < (rtm::col_half(it.x) ? 32 : 64);
    X.srow = rtm::col_src_row(S, V);
    // rows of the image region carry the snapshot store (forward) or the
    // imaging (backward, from the snapshot store or the rebuilt source field)
    const bool img = Q16 ? (V.qsnap || V.smax) : (V.snap || V.sfield);
    const int ex_lo = img ? S.TU : 0, ex_hi = img ? S.TU + S.nz : 0;
    const rtm::ColRec<R> rec = rtm::col_rec(S, V, X.xl, X.ok, S.rec_row >= zA && S.rec_row < zB);
    float umax = 0.0f;
    rtm::Pair<R> w[16];
#pragma unroll
    for (int r = 0; r < 8; r++) w[r] = rtm::ldg(X.u2 + (zA - 4 + r) * X.Ph, X.xh);
    ColSet<R> A, Bs;
    col_fetch<R, 8>(w, A, X, zA, zB);
    for (int k0 = zA;; k0 += 16) {
        col_step<R, 0, Q16, IMG>(w, A, Bs, X, S, V, rec, k0, zB, ex_lo, ex_hi, umax);
        if (k0 + 4 >= zB) break;
        col_step<R, 1, Q16, IMG>(w, Bs, A, X, S, V, rec, k0 + 4, zB, ex_lo, ex_hi, umax);
        if (k0 + 8 >= zB) break;
        col_step<R, 2, Q16, IMG>(w, A, Bs, X, S, V, rec, k0 + 8, zB, ex_lo, ex_hi, umax);
        if (k0 + 12 >= zB) break;
        col_step<R, 3, Q16, IMG>(w, Bs, A, X, S, V, rec, k0 + 12, zB, ex_lo, ex_hi, umax);
        if (k0 + 16 >= zB) break;
    }
    if (Q16 && V.smax) commit_snap_max(V.smax, umax, item);
}

template <typename R, bool Q16>
__device__ __forceinline__ void column_item(const SimDesc<R> &S, int n, int4 it, int item)
{
    const rtm::StepView<R> V = rtm::step_view_q<R, Q16>(S, n);
    if (!V.active) return;
    // imaging steps (backward, a snapshot or the rebuilt source field) run
    // the instance that loads the imaging operands inside the march
    if (!V.forward && (Q16 ? V.qsnap != nullptr : (V.snap || V.sfield))) column_march<R, Q16, true>(S, V, it, item);
    else column_march<R, Q16, false>(S, V, it, item);
}

template <typename R, bool Q16>
__global__ void __launch_bounds__(256) k_columns(const SimDesc<R> *__restrict__ sims,
                                                 const int *__restrict__ n0,
                                                 const int4 *__restrict__ items,
                                                 const int *__restrict__ nitems_p, int s)
{
    const int nitems = *nitems_p;
    const int nblk = (nitems + 3) >> 2;
    const int per = (nblk + 7) >> 3;
    const int b = blockIdx.x;
    const int t = (b & 7) * per + (b >> 3);
    if ((b >> 3) >= per || t >= nblk) return;
    const int item = t * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (item >= nitems) return;
    const int4 it = items[item];
    const int sim = rtm::col_sim(it.x);
    const SimDesc<R> &S = sims[sim];
    column_item<R, Q16>(S, n0[sim] + s, it, item);
}

// ---- two steps per launch over the inner image region (rtm_pair.h) ----

// Per-wavefront constants of a pair item.
template <typename R>
struct PairCtx {
    const rtm::Pair<R> *u2, *O2, *C2;   // u^n, u^{n-1}, C
    rtm::Pair<R> *U1, *U2;              // u^{n+1}, u^{n+2}
    int Ph;                             // row pitch in pairs
    int xh;                             // the lane's pair index (clamped into the row)
    int xl;                             // the lane's first column
    int zA, zB;                         // the item's rows
    int ulast, clast;                   // last rows the u^n and C/u^{n-1} loads may touch
    bool own;                           // the lane's columns are the item's
};

// Register windows of the march: row r of every window is slot
// (r - zA + 8) mod NS, NS = 8 + 8 PD, where PD is the prefetch distance in
// chunks: with the loop unrolled over NS/4 chunks every slot index is a
// compile-time constant and nothing is copied between registers.
template <int PD>
constexpr int pair_slots() { return 8 + 8 * PD; }

// Loads for chunk c (step-1 rows m0 = zA-4+4c .. m0+3): u^n rows m0+4 ..
// m0+7 and C, u^{n-1} rows m0 .. m0+3, into the slots of chunk c = CI mod
// NS/4.  Rows past the item are clamped, so every chunk issues the same
// loads and waits for exactly those issued PD chunks earlier.
template <typename R, int PD, int CI>
__device__ __forceinline__ void pair_fetch(rtm::Pair<R> (&a)[pair_slots<PD>()], rtm::Pair<R> (&c)[pair_slots<PD>()],
                                           rtm::Pair<R> (&o)[pair_slots<PD>()], const PairCtx<R> &X, int m0)
{
    constexpr int NS = pair_slots<PD>();
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        const int ku = m0 + 4 + rr < X.ulast ? m0 + 4 + rr : X.ulast;
        const int kc = m0 + rr < X.clast ? m0 + rr : X.clast;
        a[(8 + 4 * CI + rr) % NS] = rtm::ldg(X.u2 + ku * X.Ph, X.xh);
        c[(4 + 4 * CI + rr) % NS] = rtm::ldg(X.C2 + kc * X.Ph, X.xh);
        o[(4 + 4 * CI + rr) % NS] = rtm::ldg(X.O2 + kc * X.Ph, X.xh);
    }
}

// x-neighbours of a lane's pair: lanes l-1, l-2, l+1, l+2 (a lane with no
// neighbour keeps its own value; those lanes' results are never used)
template <typename R>
__device__ __forceinline__ void pair_shifts(rtm::Pair<R> c, rtm::Pair<R> &L1, rtm::Pair<R> &L2, rtm::Pair<R> &R1,
                                            rtm::Pair<R> &R2)
{
    L1.a = lane_shr1(c.a, c.a);
    L1.b = lane_shr1(c.b, c.b);
    L2.a = lane_shr1(L1.a, L1.a);
    L2.b = lane_shr1(L1.b, L1.b);
    R1.a = lane_shl1(c.a, c.a);
    R1.b = lane_shl1(c.b, c.b);
    R2.a = lane_shl1(R1.a, R1.a);
    R2.b = lane_shl1(R1.b, R1.b);
}

struct PairRun {
    bool two;           // the second step runs
    int srow1, srow2;   // source rows of the two steps (-1: none)
    int clast;          // last chunk index
};

// Chunk c = CI mod NS/4: prefetch chunk c+PD, u^{n+1} on rows m0 .. m0+3
// (kept in the window b, stored on the item's rows), then, from the third
// chunk on, u^{n+2} on rows m0-4 .. m0-1.  The extras (source, receiver,
// snapshot, imaging) are branched around on rows that carry none.
template <typename R, int PD, bool IMG, int CI>
__device__ __forceinline__ void pair_chunk(rtm::Pair<R> (&a)[pair_slots<PD>()], rtm::Pair<R> (&b)[pair_slots<PD>()],
                                           rtm::Pair<R> (&c)[pair_slots<PD>()], rtm::Pair<R> (&o)[pair_slots<PD>()],
                                           const PairCtx<R> &X, const SimDesc<R> &S, const rtm::StepView<R> &V1,
                                           const rtm::StepView<R> &V2, const rtm::PairRec<R> &rec, const PairRun &P,
                                           int cidx)
{
    constexpr int NS = pair_slots<PD>();
    const int m0 = X.zA - 4 + 4 * cidx;
    pair_fetch<R, PD, (CI + PD) % (NS / 4)>(a, c, o, X, m0 + 4 * PD);
    const bool ex1 = (P.srow1 >= m0 && P.srow1 < m0 + 4) || (S.rec_row >= m0 && S.rec_row < m0 + 4)
                     || (V1.snap && m0 + 4 > X.zA && m0 < X.zB);
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        const int m = m0 + rr;
        rtm::Pair<R> win[9];
#pragma unroll
        for (int j = 0; j < 9; j++) win[j] = a[(4 * CI + rr + j) % NS];
        rtm::Pair<R> L1, L2, R1, R2;
        pair_shifts(win[4], L1, L2, R1, R2);
        const int sm = (4 + 4 * CI + rr) % NS;
        rtm::Pair<R> U = rtm::col_update(win, L1, L2, R1, R2, c[sm], o[sm]);
        const bool mine = m >= X.zA && m < X.zB;   // uniform
        if (ex1) rtm::pair_extras<R, IMG>(S, V1, rec, false, m, X.xl, X.own && mine, P.srow1, U, c[sm]);
        b[sm] = U;
        if (mine && X.own) rtm::stg(X.U1 + m * X.Ph, X.xh, U);
    }
    if (!P.two || cidx < 2) return;
    const int q0 = m0 - 4;
    const bool ex2 = (P.srow2 >= q0 && P.srow2 < q0 + 4) || (S.rec_row >= q0 && S.rec_row < q0 + 4) || V2.snap;
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        const int q = q0 + rr;
        rtm::Pair<R> win[9];
#pragma unroll
        for (int j = 0; j < 9; j++) win[j] = b[(4 * CI + rr + NS - 4 + j) % NS];
        rtm::Pair<R> L1, L2, R1, R2;
        pair_shifts(win[4], L1, L2, R1, R2);
        const int sq = (4 * CI + rr) % NS;
        rtm::Pair<R> U = rtm::col_update(win, L1, L2, R1, R2, c[sq], a[sq]);
        if (ex2) rtm::pair_extras<R, IMG>(S, V2, rec, true, q, X.xl, X.own, P.srow2, U, c[sq]);
        if (X.own) rtm::stg(X.U2 + q * X.Ph, X.xh, U);
    }
}

// One turn of the window cycle: chunks c0 .. c0 + NS/4 - 1; true when the
// item's last chunk has run.
template <typename R, int PD, bool IMG, int CI>
__device__ __forceinline__ bool pair_cycle(rtm::Pair<R> (&a)[pair_slots<PD>()], rtm::Pair<R> (&b)[pair_slots<PD>()],
                                           rtm::Pair<R> (&c)[pair_slots<PD>()], rtm::Pair<R> (&o)[pair_slots<PD>()],
                                           const PairCtx<R> &X, const SimDesc<R> &S, const rtm::StepView<R> &V1,
                                           const rtm::StepView<R> &V2, const rtm::PairRec<R> &rec, const PairRun &P,
                                           int c0)
{
    pair_chunk<R, PD, IMG, CI>(a, b, c, o, X, S, V1, V2, rec, P, c0 + CI);
    if (c0 + CI == P.clast) return true;
    if constexpr (CI + 1 < pair_slots<PD>() / 4)
        return pair_cycle<R, PD, IMG, CI + 1>(a, b, c, o, X, S, V1, V2, rec, P, c0);
    else return false;
}

// One pair item (the body of k_pairs) for a sim with four u buffers at
// steps n and, when `two`, n+1.  IMG: an instance for items whose steps
// image (backward, snapshot steps), the only one that loads inside the march.
template <typename R, int PD, bool IMG>
__device__ __forceinline__ void pair_march(const SimDesc<R> &S, const rtm::StepView<R> &V1, const rtm::StepView<R> &V2,
                                           bool two, int4 it)
{
    constexpr int NS = pair_slots<PD>();
    PairRun P;
    P.two = two;
    const int lane = threadIdx.x & 63;
    const int wp = rtm::pair_wp(it.x);
    PairCtx<R> X;
    X.u2 = (const rtm::Pair<R> *)V1.u;
    X.O2 = (const rtm::Pair<R> *)V1.Uo;
    X.C2 = (const rtm::Pair<R> *)S.C;
    X.U1 = (rtm::Pair<R> *)V1.U;
    X.U2 = (rtm::Pair<R> *)V2.U;
    X.Ph = S.P >> 1;
    X.xl = it.y - 8 + 2 * lane;
    X.xh = rtm::clampi(X.xl, 0, S.P - 2) >> 1;
    X.zA = it.z;
    X.zB = it.w;
    X.ulast = X.zB + 7;
    X.clast = X.zB + 3;
    X.own = lane >= 4 && lane < 4 + wp;
    P.srow1 = rtm::col_src_row(S, V1);
    P.srow2 = P.two ? rtm::col_src_row(S, V2) : -1;
    P.clast = (X.zB - X.zA) / 4 + 1;   // chunks 0 .. clast
    const rtm::PairRec<R> rec =
        rtm::pair_rec(S, V1, V2, P.two, X.xl, S.rec_row >= X.zA - 4 && S.rec_row < X.zB + 4);
    rtm::Pair<R> a[NS], b[NS], c[NS], o[NS];
#pragma unroll
    for (int r = 0; r < 8; r++) a[r] = rtm::ldg(X.u2 + (X.zA - 8 + r) * X.Ph, X.xh);
    pair_fetch<R, PD, 0>(a, c, o, X, X.zA - 4);
    if constexpr (PD >= 2) pair_fetch<R, PD, 1>(a, c, o, X, X.zA);
    if constexpr (PD >= 3) pair_fetch<R, PD, 2>(a, c, o, X, X.zA + 4);
    for (int c0 = 0;; c0 += NS / 4)
        if (pair_cycle<R, PD, IMG, 0>(a, b, c, o, X, S, V1, V2, rec, P, c0)) break;
}

template <typename R, int PD>
__device__ __forceinline__ void pair_item(const SimDesc<R> &S, int n, bool two, int4 it)
{
    const rtm::StepView<R> V1 = rtm::step_view_q<R, false, false, 3>(S, n);
    if (!V1.active) return;
    const rtm::StepView<R> V2 = rtm::step_view_q<R, false, false, 3>(S, n + 1);
    two = two && V2.active;
    if (!V1.forward && (V1.snap || (two && V2.snap))) pair_march<R, PD, true>(S, V1, V2, two, it);
    else pair_march<R, PD, false>(S, V1, V2, two, it);
}

// Two steps (s and s+1 of the block) of every pair item, four independent
// wavefronts per workgroup; the second step only when s+1 < lim (the steps
// of the block or of the eager remainder).  XCD-contiguous runs of items as
// in k_columns.
template <typename R, bool ONE, int PD>
__global__ void __launch_bounds__(256) k_pairs(const SimDesc<R> *__restrict__ sims, const int *__restrict__ n0,
                                               const int4 *__restrict__ items, const int *__restrict__ nitems_p,
                                               int s, int lim)
{
    const int nitems = *nitems_p;
    const int nblk = (nitems + 3) >> 2;
    const int per = (nblk + 7) >> 3;
    const int b = blockIdx.x;
    const int t = (b & 7) * per + (b >> 3);
    if ((b >> 3) >= per || t >= nblk) return;
    const int item = t * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (item >= nitems) return;
    const int4 it = items[item];
    const int sim = ONE ? 0 : rtm::pair_sim(it.x);
    pair_item<R, PD>(sims[sim], n0[sim] + s, s + 1 < lim, it);
}

// Boundary saving, forward: after step n, U = u^{n+1} on the ring (rtm::ring_point)
// into ring[n][*].  grid = (ring blocks, sims).
template <typename R>
__global__ void k_ring_save(const SimDesc<R> *__restrict__ sims, const int *__restrict__ n0, int s)
{
    const SimDesc<R> &S = sims[blockIdx.y];
    if (!(S.flags & rtm::F_RING) || S.phase != rtm::PH_FORWARD) return;
    const int n = n0[blockIdx.y] + s;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < 0 || n >= S.nt || t >= S.NR) return;
    int i, k;
    rtm::ring_point(S, t, i, k);
    S.ring[(size_t)n * S.NR + t] = S.ubuf[(n + 1) & 1][(size_t)k * S.P + i];
}

// Boundary saving, backward step n (1 <= n <= nt-2): the source field one
// step further back.  With m = nt - n, sbuf[m & 1] holds u^m and
// sbuf[(m+1) & 1] u^{m+1}; inside the image region u^{m-1} is the forward
// update solved for its old value, u^{m-1} = stencil(C, u^{m+1}, u^m) plus
// the source term of forward step m (its stencil subtracted U_old and then
// added w[m] C), written over u^{m+1}; on the ring u^{m-1} is the value
// saved by forward step m-2.  The step's imaging reads sbuf[m & 1] only.
template <typename R>
__global__ void k_recon(const SimDesc<R> *__restrict__ sims, const int *__restrict__ n0, int s)
{
    const SimDesc<R> &S = sims[blockIdx.y];
    if (!(S.flags & rtm::F_RING) || S.phase != rtm::PH_BACKWARD) return;
    const int n = n0[blockIdx.y] + s;
    if (n < 1 || n > S.nt - 2) return;
    const int m = S.nt - n;
    const R *cur = S.sbuf[m & 1];
    R *out = S.sbuf[(m + 1) & 1];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nD = S.nx * S.nz;
    if (t < nD) {
        const int i = S.TL + t % S.nx, k = S.TU + t / S.nx;
        const size_t j = (size_t)k * S.P + i;
        const size_t P = S.P;
        R v = rtm::stencil_val(S.C[j], out[j], cur[j], cur[j - 1], cur[j + 1], cur[j - P], cur[j + P],
                               cur[j - 2], cur[j + 2], cur[j - 2 * P], cur[j + 2 * P],
                               cur[j - 3], cur[j + 3], cur[j - 3 * P], cur[j + 3 * P],
                               cur[j - 4], cur[j + 4], cur[j - 4 * P], cur[j + 4 * P]);
        if (m < S.nw && (int)j == S.src_j) v = v + R(S.wav[m]) * S.C[j];
        out[j] = v;
    } else if (t < nD + S.NR) {
        int i, k;
        rtm::ring_point(S, t - nD, i, k);
        out[(size_t)k * S.P + i] = S.ring[(size_t)(m - 2) * S.NR + (t - nD)];
    }
}

// n0[i] += block for every simulation (the last node of a captured block)
static __global__ void k_advance(int *__restrict__ n0, int nsims, int block)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsims) n0[i] += block;
}

// zero fill (used instead of hipMemsetAsync so that the same code is safe to
// capture into a graph)
template <typename T>
__global__ void k_fill_zero(T *__restrict__ p, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) p[t] = T(0);
}

template <typename T>
inline void fill_zero(T *p, size_t n, hipStream_t s)
{
    if (!p || !n) return;
    const size_t blocks = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
    hipLaunchKernelGGL(k_fill_zero<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
}

// The per-point kernel (rtm::step_point), used by the per-step legacy entry
// points and for injection points outside the stepped region.
template <typename R>
__global__ void __launch_bounds__(kBX * kBZ) k_step(StepArgs<R> A)
{
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int k = blockIdx.y * kBZ + threadIdx.y;
    if (i < rtm::kHalf || i >= A.X - rtm::kHalf || k < rtm::kHalf || k >= A.Z - rtm::kHalf) return;
    rtm::step_point(A, i, k);
}

// Top-strip correction of the outer columns i < 4 and i >= X-4
// (rtm::cpml_top_edge_point); 8 columns x (TU-8) rows.
template <typename R>
__global__ void k_top_edges(StepArgs<R> A)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int rows = A.TU - 2 * rtm::kHalf;
    if (t >= 8 * rows) return;
    const int c = t % 8, k = rtm::kHalf + t / 8;
    const int i = c < 4 ? c : A.X - 8 + c;
    rtm::cpml_top_edge_point(A, i, k);
}

template <typename R>
inline void launch_top_edges(const StepArgs<R> &A, hipStream_t s)
{
    const int n = 8 * (A.TU - 2 * rtm::kHalf);
    hipLaunchKernelGGL(k_top_edges<R>, dim3((n + 63) / 64), dim3(64), 0, s, A);
}

// coef = (v*aux)*(v*aux) from the padded velocity in the reference's layout
// ([X][Z], z fastest) into the device layout, SRC/forward_modeling.c:79.
template <typename R>
__global__ void k_coef(const float *__restrict__ v, R *__restrict__ C, R aux, int X, int Z, int P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= X || k >= Z) return;
    const R vv = R(v[(size_t)i * Z + k]);
    C[(size_t)k * P + i] = (vv * aux) * (vv * aux);
}

// receiver replacement at points outside the stepped region,
// SRC/backward_modeling.c:106-110
template <typename R>
__global__ void k_inject_list(R *__restrict__ U, const int *__restrict__ jlist,
                              const int *__restrict__ rlist, int n, const float *__restrict__ seis_row)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) U[jlist[t]] = R(seis_row[rlist[t]]);
}

// source add at a point outside the stepped region, SRC/forward_modeling.c:105-109
template <typename R>
__global__ void k_source_point(R *__restrict__ U, const R *__restrict__ C, int j, const float *__restrict__ w, int n)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) U[j] = U[j] + R(w[n]) * C[j];
}

// seismogram recording for every receiver (also when several share a point)
template <typename R>
__global__ void k_record(const R *__restrict__ U, const int *__restrict__ rec_j, int nrec,
                         float *__restrict__ row)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrec) row[r] = float(U[rec_j[r]]);
}

// imaging of padded fields (legacy normalized_cross_correlation),
// SRC/image_conditions.c:20-33; image layout [nz][nx] x fastest.
template <typename R>
__global__ void k_image_padded(R *__restrict__ I, R *__restrict__ Is, const R *__restrict__ S,
                               const R *__restrict__ Rw, int nx, int nz, int TL, int TU, int P)
{
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    const int kk = blockIdx.y;
    if (ii >= nx || kk >= nz) return;
    const size_t j = (size_t)(kk + TU) * P + (ii + TL);
    const size_t q = (size_t)kk * nx + ii;
    I[q] = I[q] + S[j] * Rw[j];
    if (Is) Is[q] = Is[q] + S[j] * S[j];
}

// ---------------------------------------------------------------------------
// host-side layout helpers (reference [X][Z] z-fastest <-> device [Z][P])
// ---------------------------------------------------------------------------

template <typename Src, typename Dst>
void to_device_layout(const Src *ref, Dst *dev, int X, int Z, int P)
{
    for (int k = 0; k < Z; k++) {
        Dst *row = dev + (size_t)k * P;
        for (int i = 0; i < X; i++) row[i] = Dst(ref[(size_t)i * Z + k]);
        for (int i = X; i < P; i++) row[i] = Dst(0);
    }
}

template <typename Src, typename Dst>
void from_device_layout(const Src *dev, Dst *ref, int X, int Z, int P)
{
    for (int i = 0; i < X; i++)
        for (int k = 0; k < Z; k++) ref[(size_t)i * Z + k] = Dst(dev[(size_t)k * P + i]);
}

// ---------------------------------------------------------------------------
// device memory arena
// ---------------------------------------------------------------------------

// One hipMalloc up front, first-fit sub-allocation in 256-byte units with
// coalescing on free, so that simulations come and go (the request server)
// without hipMalloc / hipFree, the latter of which synchronises the device.
struct DeviceArena {
    char *base = nullptr;
    size_t size = 0, in_use = 0;
    std::map<size_t, size_t> free_;     // offset -> bytes
    std::map<size_t, size_t> live_;     // offset -> bytes
    int init(size_t bytes);
    void release();
    void *alloc(size_t bytes);           // nullptr when no free block is large enough
    void free(void *ptr);
    size_t largest_free() const;
    static size_t rounded(size_t bytes) { return (bytes + 255) / 256 * 256; }
};

// ---------------------------------------------------------------------------
// one simulation's device buffers
// ---------------------------------------------------------------------------

struct SimBuffers {
    rtm_params p{};
    int real = RTM_REAL_F32;
    size_t esz = 4;
    int X = 0, Z = 0, P = 0;
    size_t nfield = 0, nint = 0, nsnap = 0;
    // u^n is u[n & umask]: two buffers, or four for two-step pair launches
    // (u[2], u[3] are allocated when a pair shape is first set)
    void *u[4] = {nullptr, nullptr, nullptr, nullptr};
    int umask = 1;
    void *C = nullptr;
    void *psiX[2] = {nullptr, nullptr}, *psiZ[2] = {nullptr, nullptr};
    void *csiX = nullptr, *csiZ = nullptr;
    void *ax = nullptr, *bx = nullptr, *az = nullptr, *bz = nullptr;
    int *recmap = nullptr;              // [X]
    int *rec_j = nullptr;               // [nrec] device index of each receiver
    int *out_j = nullptr, *out_r = nullptr;
    int n_out = 0;                      // receivers outside the stepped region
    bool rec_dup = false;               // several receivers share a point
    float *wav = nullptr;               // [n_w]
    float *seis_t = nullptr;            // [nt][nrec] backward input (t-major)
    float *seis_rec = nullptr;          // [nt][nrec] forward recording
    void *snaps = nullptr;              // [nsnap][nz][nx] (snapshot mode FP32)
    int qmode = rtm::Q_NONE;            // 16-bit snapshot store (modes FP16, BF16)
    unsigned short *q16 = nullptr;      // [nsnap][nz][nx]
    unsigned *smax = nullptr;           // [nsnap] max |u| bits per snapshot
    signed char *qexp = nullptr;        // int16 block store: [nsnap][nz][nqb] block exponents
    int qb0 = 0, nqb = 0;               // its first block and blocks per image row
    bool ringmode = false;              // boundary saving (snapshot mode BOUNDARY)
    void *sbuf[2] = {nullptr, nullptr}; // the source wavefield's buffers
    void *ring = nullptr;               // [nt][NR] ring values of every forward step
    int NR = 0;
    void *I = nullptr, *Is = nullptr, *Ic = nullptr, *Isc = nullptr;
    std::vector<float> wav_host;
    bool have_velocity = false, have_seis = false, have_snaps = false;
    std::vector<void *> owned;
    DeviceArena *arena = nullptr;       // allocate from here instead of hipMalloc

    // `share` (same geometry): take its profiles, receiver tables, wavelet
    // and trace buffers instead of allocating copies (an ensemble over one
    // shot gather)
    int alloc(const rtm_params &pp, int real_type, const SimBuffers *share = nullptr);
    void release();
    // step with four u buffers (pair shapes) or two; the extra buffers stay
    // allocated once made, and are zeroed by reset_state while in use
    int set_quad(bool on);
    // a pass may run its steps that neither store nor image with the light
    // tile instance: the FP32 and the int16 block stores write a snapshot
    // only in the step that computes it (and boundary saving none in the
    // tile launch); the FP16 / BF16 stores also fill the previous
    // snapshot's slot in the step after it
    bool light_steps() const { return qmode == rtm::Q_NONE || qmode == rtm::Q_I16; }
    int upload_profiles();
    int upload_receivers();
    // uploads; with wait == false the host data must stay alive until the
    // stream has run the copy (the caller synchronises later)
    int set_velocity(const float *vpad, hipStream_t s, bool wait = true);
    int set_wavelet(const float *w, int n);
    int set_seismogram(const float *d, hipStream_t s, bool wait = true);
    // zero the wavefield (u, or with source_field the boundary-saving source
    // buffers), psi and csi: a pass starts at rest
    void reset_state(hipStream_t s, bool source_field = false);
    void reset_snap_scales(hipStream_t s); // zero the 16-bit store's max |u| slots
    // the 16-bit store's last snapshot when it is the pass's final field
    int flush_last_snapshot(hipStream_t s);
    void zero_image(hipStream_t s);
    template <typename R>
    SimDesc<R> desc() const;
};

// plan memory in bytes for params p (the admission footprint)
size_t sim_bytes(const rtm_params &p, size_t esz);

// Work lists covering the stepped region of a grid (rtm::build_work): tiles
// {sim, x0, z0, class}, and interior items for the kernel `interior`
// (rtm::Interior) in segments of at most seg_rows rows.
// Reorder a tile list for the eight XCDs.  Workgroup b runs on XCD b & 7 and
// k_tiles gives XCD k the k-th eighth of the list, so each eighth should cost
// the same: every simulation's tiles are cut into eight slabs of whole tile
// columns of equal cost (a strip tile, which moves about twice the bytes and
// runs the longer chain, counts as three interior tiles), slab k of every
// simulation goes to XCD k, and the eighths are padded to equal length with
// entries k_tiles skips.  Inside a slab the classes with the longest chains
// are dispatched first (corner, z-strip, x-strip, interior), each row-major
// so that neighbouring tiles share the XCD's L2.
void balance_tiles_for_xcds(std::vector<int4> &tiles, int tz);

void append_work(std::vector<int4> &tiles, std::vector<int4> &inner, int sim, const rtm_params &p,
                 int rz, int interior, int seg_rows);
inline void append_tiles(std::vector<int4> &out, int sim, const rtm_params &p, int rz)
{
    std::vector<int4> none;
    append_work(out, none, sim, p, rz, rtm::IN_TILES, 0);
}

// ---------------------------------------------------------------------------
// executor: one launch per time step over a tile list, hipGraph blocks
// ---------------------------------------------------------------------------

// 0 if (rows per thread, interior kernel) is a shape the kernels exist for;
// otherwise -1 with the reason in the error string
int check_shape(int rz, int interior);
// 0 if the snapshot store (enum rtm_snapshot_mode) runs with that interior
// kernel: the int16 block store runs in tile shapes only
int check_store_shape(int snapshot_mode, int interior);
// 0 if a plan or batch of this element type runs with that interior kernel:
// two-step pair items step FP32 fields only
int check_pair_real(int real, int interior);

struct Exec {
    int real = RTM_REAL_F32;
    int rz = 4;                         // rows per thread (tile height 4*rz)
    int interior = rtm::IN_TILES;       // kernel of the interior items
    bool q16 = false;                   // the sims keep 16-bit snapshot stores (FP16 / BF16)
    bool i16 = false;                   // the sims keep int16 block stores (tile shapes only)
    // boundary saving: the pass's extra kernel (rtm::PH_FORWARD: ring save,
    // rtm::PH_BACKWARD: reconstruction, 0: none) and its grid width
    int ring_phase = 0, ring_blocks = 0;
    int cap_sims = 0, cap_tiles = 0, cap_inner = 0;
    int nsims = 0, ntiles = 0, ninner = 0;
    void *d_desc = nullptr;             // SimDesc<R>[cap_sims]
    int *d_n0 = nullptr;                // [cap_sims]
    int4 *d_tiles = nullptr;            // [cap_tiles]
    int *d_ntiles = nullptr;
    int4 *d_inner = nullptr;            // [cap_inner] column items
    int *d_ninner = nullptr;
    std::vector<int> h_n0;
    // Captured blocks of steps, one per boundary-saving phase (index
    // ring_phase), so that a plan alternating between its forward and
    // backward passes keeps both graphs instead of re-capturing at every
    // switch.  The captured launches depend on the lists only through their
    // capacities (grid sizes) and on which lists are non-empty (the kernels
    // read the items and their counts from device memory), so new list
    // contents keep the graphs unless that changes.  A graph is destroyed
    // only after the stream it was last launched on has drained.
    struct Captured {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int block = 0, ring_blocks = 0;
        std::vector<char> heavy;        // per captured step: its StepKind (empty: every step K_FULL)
    };
    // by phase (plans, batches, boundary saving: 0 none, 1 forward, 2
    // backward); 3..6 for the request server's schedules (a group's heavy
    // steps change with its members' passes, and a few patterns recur)
    Captured cap[7];
    int gcap_next = 0;
    hipStream_t g_stream = nullptr;     // where the graphs were last launched
    bool g_tiles = false, g_inner = false;
    // the strip tiles run on `side` beside the interior kernel (they write
    // disjoint points); fork/join through two events, also under capture
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    int init(int real_type, int max_sims, int max_tiles, int rows_per_thread, int max_inner = 0,
             int interior_kernel = rtm::IN_TILES);
    void release();
    int set_tiles(const std::vector<int4> &tiles, hipStream_t s);
    int set_inner(const std::vector<int4> &items, hipStream_t s);
    int upload_descs(const void *descs, int n, hipStream_t s);   // SimDesc<R>[n]
    int set_n0(const std::vector<int> &n0, hipStream_t s);
    // `steps` launches from the current n0 (eager, or in graph blocks of
    // `block` steps when block > 1); n0 advances by `steps` for all sims
    int run(int steps, int block, hipStream_t s);
    // the launch of step s_in_block of a block; with pair items (IN_PAIRS)
    // an even s_in_block also steps them at s_in_block + 1 when that is
    // below `lim` (the block's length).  kind (StepKind): K_LIGHT, no
    // simulation stores or images at this step; K_STORE, some store a
    // snapshot and none images; K_FULL, anything.
    int launch_one(int s_in_block, hipStream_t s, int lim = 1 << 30, int kind = K_FULL);
    // The step schedule of the coming pass (every simulation of the launch at
    // the same step, a store that writes snapshots only in the step that
    // computes them): the forward steps n % ts == 0 store a snapshot (flags
    // F_SNAP), the backward steps with (nt-1-n) % ts == 0 image (F_IMAGE);
    // every other step runs the light tile instance.  phase 0 clears it
    // (every step runs the full instance).
    bool sched = false;
    int sched_phase = 0, sched_nt = 0, sched_ts = 1, sched_flags = 0;
    // The request server's form: one schedule {phase, nt, ts, flags} per
    // simulation slot, each at its own step (h_n0); phase 0 for the slots
    // outside the launch.  A step runs the light instance when no member
    // stores or images at it.
    std::vector<int4> gsched;
    int *d_err = nullptr;               // set by a step launch that met a tile its list should not hold
    void set_schedule(int phase, int nt, int ts, int flags);
    void set_group_schedule(const std::vector<int4> &per_sim);
    int heavy_step(int n) const;        // the StepKind step n of the pass needs (K_FULL without a schedule)
    int heavy_at(int t) const;          // the StepKind step t from the current counters needs
    bool pattern(int block) const;      // the heavy steps repeat from block to block
    int check_err(hipStream_t s);       // -1 with the error set if d_err was set (and clears it)
    // every captured graph, once the stream they ran on has drained
    int drop_graph();
    // the boundary-saving kernel of the coming pass (0: none; its grid
    // width `blocks` matters only for a boundary-saving phase)
    int set_ring_phase(int phase, int blocks);
};

// ---------------------------------------------------------------------------
// autotuner: candidate kernel shapes, timing, and a per-grid-class cache
// ---------------------------------------------------------------------------

// Launch groups: the working set of one step of a grid is about 16 B per
// padded point (u^n, u^{n-1}, C read, u^{n+1} written); a launch whose
// grids' working sets sum to at most kGroupBytes stays within the device's
// 256 MB last-level cache from one step to the next (16 grids of 1801 x 451
// padded points).
constexpr size_t kStepBytesPerPoint = 16;
constexpr size_t kGroupBytes = size_t(224) << 20;

struct Shape {
    int rz, interior, seg;
    int group = 0;      // batches: realizations per launch group (0: all)
    int lanes = 1;      // batches: launch groups in flight at once
};

// the shapes tried: tile heights 4..32 rows, and the column kernel beside
// strip tiles with whole or segmented columns
const std::vector<Shape> &tuner_candidates();

// cache key of a launch: element type, store kind and every sim's grid
std::string tuner_key(int real, int qmode, const std::vector<const rtm_params *> &grids);
bool tuner_lookup(const std::string &key, Shape *out);
void tuner_store(const std::string &key, const Shape &s);

// ms per step of `steps` forward steps of `descs` (from rest, no source;
// stores only if `flags` asks for them) with the executor's current work
// lists; descriptors and step counters are left as uploaded here, so the
// caller re-uploads its own
template <typename R>
int time_steps(Exec &ex, std::vector<SimDesc<R>> descs, int steps, int block, hipStream_t s, double *ms,
               int flags = 0);

}  // namespace rtmd

#endif
