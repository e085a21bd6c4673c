// sdm_kernels.h -- HIP kernels of the SDM step for gfx950 (wave64).
//
// One step of every replica r of a context (LangevinIntegratorSDM::step,
// openmmapi/src/LangevinIntegratorSDM.cpp:153-183) is the launch sequence
//
//   k_bin -> k_scan -> k_scatter -> k_cellsort -> k_bbox -> k_tiles     spatial sort + tiles
//   k_pair                 F0: full LJ+RF sweep at x (FP32 math, fixed-point forces)
//   k_energy64             U0 on the steps that report energies (FP64 pair sums)
//   k_bonded               Ub, Fb: harmonic bonds/angles at x (FP64, gathered per atom)
//   k_ligcand              j-blocks near the displaced atoms at x or at x+d
//   k_denv                 screened pair lists; F1-F0 on the surrounding atoms (FP64)
//   k_dlig, k_deltaA_red   u and F1-F0 on displaced atoms (FP64, pairs with d_i != d_j)
//   k_scalar               SoftCoreF, bias W/W', PotEnergy, noneq schedule/work, sp
//   k_mix                  F += sp (F1 - F0)
//   k_langevin             leapfrog Langevin update, Philox noise (FP64)
//   k_constrain            SHAKE on constrained clusters (FP64)
//
// U1 and F1 are never computed as a second full sweep: pairs whose displacements are
// equal (d_i == d_j) have identical minimum-image distances at x and at x+d, so
// U(x+d) - U(x) and F(x+d) - F(x) are sums over the few pairs with d_i != d_j only.
// Every grid has the replica index in its y (or z) dimension.
#pragma once
#include <type_traits>

#include "sdm_device.h"

namespace sdm {

// Box and cutoff constants of every kernel.  A non-periodic system (CutoffNonPeriodic)
// has a box of 1e30 with zero inverse: every image shift L * round(d / L) is then 0 and
// coordinates are never wrapped.
struct Geom {
    float bx, by, bz, ibx, iby, ibz;  // box and inverse (FP32 path)
    double Lx, Ly, Lz, iLx, iLy, iLz; // FP64 path
    float rc2f, krff, crff;
    double rc2, krf, crf;
    // one periodic shift s for a whole tile (two blocks with centre difference dc and
    // half-extents h_i, h_j) is exact for every pair within the list radius r_l when
    // |dc - s| + h_i + h_j <= L - r_l in each dimension: a pair that s puts k != 0 box
    // lengths off its minimum image then lies at least L - (L - r_l) = r_l away in truth,
    // and at least L/2 >= r_c as evaluated, so it is left out either way.  oix..oiz hold
    // L - r_l less a 1e-3 nm margin for FP32 rounding (1e30 if not periodic)
    float oix, oiy, oiz;
    // LJ sigma combining: 0 Lorentz-Berthelot, 1 geometric (OPLS).  The per-atom FP32 and
    // FP64-record sigma entries hold s/2 for the first, s for the second (the FP64 sigma
    // array always s); the FP32 block kernel takes it as a template parameter
    int geo;
    // sort grid: cell c_d = floor((w_d - go_d) * gs_d), clamped to the grid, w the
    // (wrapped, if periodic) coordinate.  Periodic: go = 0, gs = cells / L; otherwise
    // the grid spans a region around the positions (atoms outside take edge cells)
    double go[3], gs[3];
};

// Tile lists kept over several steps (a Verlet skin).  The lists are built with the
// radius r_l = r_c + skin + 1e-3 nm (r2), and with oix..oiy as Geom's but for that radius
// (a shift exact for every pair within r_l).  A pair within r_c now was within
// r_c + d_i + d_j when the lists were built (d: each atom's displacement since, as a
// minimum image), so the lists hold every pair within r_c while 2 max d <= skin:
// reuse_d2 = (skin / 2)^2 (negative: build on every call).
struct TileRadius {
    float r2, oix, oiy, oiz, reuse_d2;
};
// a block displacement record that forces the next build (after a sort, or non-finite)
constexpr float DISP_STALE = 3.0e38f;
// floats between consecutive replicas' displacement records (dmax[r * DISP_STRIDE]): one
// 256-byte line each, so that the blocks' atomics of different replicas do not queue on
// one line (all on one line cost 62 us per launch at 22 x 356 blocks)
constexpr int DISP_STRIDE = 64;

// Slot-ordered FP64 record of the energy kernel: the position in the same periodic image
// as the slot's FP32 record (so one tile shift serves both), and the parameters.
struct Rec64 {
    double x, y, z, qk, sh, es;  // position; q sqrt(k_e), sigma/2, 2 sqrt(eps)
};

__device__ __forceinline__ int grid_cell(double w, double go, double gs, int n) {
    return (int)fmin(fmax(floor((w - go) * gs), 0.0), (double)(n - 1));
}

// ---------------------------------------------------------------------------
// spatial sort (deterministic): cell rank per atom, counting sort, per-cell sort by
// atom index so the slot order does not depend on atomic ordering.
// ---------------------------------------------------------------------------
__global__ void k_bin(int N, int Np, const double* __restrict__ x64, int* __restrict__ key,
                      int* __restrict__ ccount, const int* __restrict__ cellrank, Geom g,
                      int ncx, int ncy, int ncz, int NC) {
    int r = blockIdx.y;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double* X = x64 + (size_t)r * 3 * Np;
    double wx = X[i], wy = X[Np + i], wz = X[2 * Np + i];
    wx -= g.Lx * floor(wx * g.iLx);
    wy -= g.Ly * floor(wy * g.iLy);
    wz -= g.Lz * floor(wz * g.iLz);
    int cx = grid_cell(wx, g.go[0], g.gs[0], ncx);
    int cy = grid_cell(wy, g.go[1], g.gs[1], ncy);
    int cz = grid_cell(wz, g.go[2], g.gs[2], ncz);
    int k = cellrank[(cx * ncy + cy) * ncz + cz];
    key[(size_t)r * Np + i] = k;
    atomicAdd(&ccount[(size_t)r * NC + k], 1);
}

// exclusive scan of the cell counts, one 1024-thread workgroup per replica; clears the
// counters for the next step.
__global__ __launch_bounds__(1024) void k_scan(int NC, int* __restrict__ ccount,
                                               int* __restrict__ cstart, int* __restrict__ cfill) {
    __shared__ int part[1024];
    int r = blockIdx.x;
    int t = threadIdx.x;
    int per = (NC + 1023) / 1024;
    int beg = min(t * per, NC), end = min(beg + per, NC);
    int* cnt = ccount + (size_t)r * NC;
    int s = 0;
    for (int c = beg; c < end; c++) s += cnt[c];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int v = (t >= o) ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    int* st = cstart + (size_t)r * (NC + 1);
    int* fl = cfill + (size_t)r * NC;
    for (int c = beg; c < end; c++) {
        st[c] = run;
        run += cnt[c];
        cnt[c] = 0;
        fl[c] = 0;
    }
    if (t == 1023) st[NC] = part[1023];
}

__global__ void k_scatter(int N, int Np, int NC, const int* __restrict__ key,
                          const int* __restrict__ cstart, int* __restrict__ cfill,
                          int* __restrict__ orig_s) {
    int r = blockIdx.y;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int k = key[(size_t)r * Np + i];
    int s = cstart[(size_t)r * (NC + 1) + k] + atomicAdd(&cfill[(size_t)r * NC + k], 1);
    if (s >= 0 && s < N) orig_s[(size_t)r * Np + s] = i;  // always, with consistent counts
}

// per cell: order its slots by atom index, then gather the sorted FP32 arrays and clear
// the slot-ordered force accumulator.
__global__ void k_cellsort(int N, int Np, int NC, const int* __restrict__ cstart,
                           int* __restrict__ orig_s, int* __restrict__ slot_of,
                           const double* __restrict__ x64, const float* __restrict__ qs,
                           const float* __restrict__ sh, const float* __restrict__ se,
                           float4* __restrict__ posq_s, float2* __restrict__ par_s,
                           float4* __restrict__ arec, long long* __restrict__ F, Geom g,
                           const double4* __restrict__ par64, Rec64* __restrict__ rec) {
    int r = blockIdx.y;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= NC) return;
    const int* st = cstart + (size_t)r * (NC + 1);
    int beg = max(st[c], 0), end = min(st[c + 1], N);
    int* o = orig_s + (size_t)r * Np;
    for (int a = beg + 1; a < end; a++) {
        int v = o[a];
        int b = a - 1;
        while (b >= beg && o[b] > v) {
            o[b + 1] = o[b];
            b--;
        }
        o[b + 1] = v;
    }
    const double* X = x64 + (size_t)r * 3 * Np;
    for (int s = beg; s < end; s++) {
        int i = o[s];
        slot_of[(size_t)r * Np + i] = s;
        double wx = X[i], wy = X[Np + i], wz = X[2 * Np + i];
        wx -= g.Lx * floor(wx * g.iLx);
        wy -= g.Ly * floor(wy * g.iLy);
        wz -= g.Lz * floor(wz * g.iLz);
        float4 pq = make_float4((float)wx, (float)wy, (float)wz, qs[i]);
        posq_s[(size_t)r * Np + s] = pq;
        if (rec) {  // energy evaluations: the FP64 record in the same image
            double4 pr = par64[i];
            rec[(size_t)r * Np + s] = Rec64{wx, wy, wz, pr.x, pr.y, pr.z};
        }
        par_s[(size_t)r * Np + s] = make_float2(sh[i], se[i]);
        // interleaved copy for the grouped pair kernel: one 32-byte record per slot
        arec[2 * ((size_t)r * Np + s)] = pq;
        arec[2 * ((size_t)r * Np + s) + 1] = make_float4(sh[i], se[i], 0.f, 0.f);
        long long* Fr = F + (size_t)r * 3 * Np;  // forces are accumulated in slot order
        Fr[s] = 0;
        Fr[Np + s] = 0;
        Fr[2 * Np + s] = 0;
    }
}

// bounding box of each 64-slot block (one wave per block).  After a sort the slot
// order is new, so the replica's displacement record (dmax, read by k_tiles) is set to a
// value that makes the next tile-list build unconditional.
__global__ __launch_bounds__(64) void k_bbox(int N, int Np, int NB,
                                             const float4* __restrict__ posq_s,
                                             float4* __restrict__ bb_lo, float4* __restrict__ bb_hi,
                                             float* __restrict__ dmax) {
    int r = blockIdx.y, b = blockIdx.x, lane = threadIdx.x;
    int s = b * 64 + lane;
    bool v = s < N;
    float4 p = posq_s[(size_t)r * Np + s];
    float lx = wave_min(v ? p.x : 3.0e38f), ly = wave_min(v ? p.y : 3.0e38f),
          lz = wave_min(v ? p.z : 3.0e38f);
    float hx = wave_max(v ? p.x : -3.0e38f), hy = wave_max(v ? p.y : -3.0e38f),
          hz = wave_max(v ? p.z : -3.0e38f);
    if (lane == 0) {
        bb_lo[(size_t)r * NB + b] = make_float4(lx, ly, lz, 0.f);
        bb_hi[(size_t)r * NB + b] = make_float4(hx, hy, hz, 0.f);
        if (b == 0) dmax[(size_t)r * DISP_STRIDE] = DISP_STALE;
    }
}

// Between full sorts the slot order of the last sort is kept.  One wave per 64-slot
// block re-gathers its atoms' coordinates from the FP64 positions, each wrapped to the
// periodic image nearest the block's first atom (itself wrapped into the box), so a
// block stays compact when one of its atoms crosses a face of the box; then the block's
// box is recomputed and the slot-ordered force accumulator cleared.  Culling and pair
// distances take minimum images of box centres or of pairs, so they hold for any image.
__global__ __launch_bounds__(64) void k_refresh(int N, int Np, int NB,
                                                const double* __restrict__ x64,
                                                const int* __restrict__ orig_s,
                                                const float* __restrict__ qs, Geom g,
                                                float4* __restrict__ posq_s,
                                                float4* __restrict__ arec,
                                                long long* __restrict__ F,
                                                float4* __restrict__ bb_lo,
                                                float4* __restrict__ bb_hi,
                                                const double4* __restrict__ par64,
                                                Rec64* __restrict__ rec,
                                                const float4* __restrict__ xb,
                                                float* __restrict__ dmax) {
    int r = blockIdx.y, b = blockIdx.x, lane = threadIdx.x;
    int sl = b * 64 + lane;
    bool v = sl < N;  // slot b * 64 (lane 0) is always below N
    const int* O = orig_s + (size_t)r * Np;
    const double* X = x64 + (size_t)r * 3 * Np;
    int i = O[v ? sl : b * 64];
    double wx = X[i], wy = X[Np + i], wz = X[2 * Np + i];
    double rx = __shfl(wx, 0, 64), ry = __shfl(wy, 0, 64), rz = __shfl(wz, 0, 64);
    rx -= g.Lx * floor(rx * g.iLx);
    ry -= g.Ly * floor(ry * g.iLy);
    rz -= g.Lz * floor(rz * g.iLz);
    wx -= g.Lx * rint((wx - rx) * g.iLx);
    wy -= g.Ly * rint((wy - ry) * g.iLy);
    wz -= g.Lz * rint((wz - rz) * g.iLz);
    float4 p = make_float4((float)wx, (float)wy, (float)wz, qs[i]);
    size_t o = (size_t)r * Np + sl;
    if (v) {
        posq_s[o] = p;
        arec[2 * o] = p;
        if (rec) {  // energy evaluations: the FP64 record in the same image
            double4 pr = par64[i];
            rec[o] = Rec64{wx, wy, wz, pr.x, pr.y, pr.z};
        }
        long long* Fr = F + (size_t)r * 3 * Np;
        Fr[sl] = 0;
        Fr[Np + sl] = 0;
        Fr[2 * Np + sl] = 0;
    }
    float lx = wave_min(v ? p.x : 3.0e38f), ly = wave_min(v ? p.y : 3.0e38f),
          lz = wave_min(v ? p.z : 3.0e38f);
    float hx = wave_max(v ? p.x : -3.0e38f), hy = wave_max(v ? p.y : -3.0e38f),
          hz = wave_max(v ? p.z : -3.0e38f);
    if (lane == 0) {
        bb_lo[(size_t)r * NB + b] = make_float4(lx, ly, lz, 0.f);
        bb_hi[(size_t)r * NB + b] = make_float4(hx, hy, hz, 0.f);
    }
    // kept tile lists (xb: the slot coordinates when they were built): the block's
    // largest squared displacement since, as a minimum image (the block may have moved
    // to another image of the box; non-finite counts as stale), into the replica's
    // maximum by a device-scope atomic max (non-negative floats order as their bits).
    // k_scalar sets it back to 0 at the end of each evaluation; where no k_scalar runs in
    // between (pair timing, pair counts) it can only be larger, so lists are rebuilt
    // sooner, never later.  (A last-block reduction needs a device-scope fence per block,
    // which writes back the XCD's L2: +0.2 ms per step at 22 x 22,750 atoms.)  One atomic
    // per block: 356 per replica at 22,750 atoms.
    if (xb) {
        float d2 = 0.f;
        if (v) {
            const float4 q = xb[o];
            float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
            dx -= g.bx * rintf(dx * g.ibx);
            dy -= g.by * rintf(dy * g.iby);
            dz -= g.bz * rintf(dz * g.ibz);
            d2 = dx * dx + dy * dy + dz * dz;
            if (!(d2 <= 1.0e30f)) d2 = DISP_STALE;
        }
        d2 = wave_max(d2);
        if (lane == 0)
            atomicMax(reinterpret_cast<unsigned*>(dmax) + (size_t)r * DISP_STRIDE,
                      __float_as_uint(d2));
    }
}

__device__ __forceinline__ float bbox_gap2(float4 lo_i, float4 hi_i, float4 lo_j, float4 hi_j,
                                           const Geom& g) {
    float cx = 0.5f * (lo_i.x + hi_i.x) - 0.5f * (lo_j.x + hi_j.x);
    float cy = 0.5f * (lo_i.y + hi_i.y) - 0.5f * (lo_j.y + hi_j.y);
    float cz = 0.5f * (lo_i.z + hi_i.z) - 0.5f * (lo_j.z + hi_j.z);
    cx -= g.bx * rintf(cx * g.ibx);
    cy -= g.by * rintf(cy * g.iby);
    cz -= g.bz * rintf(cz * g.ibz);
    float gx = fmaxf(fabsf(cx) - 0.5f * (hi_i.x - lo_i.x) - 0.5f * (hi_j.x - lo_j.x), 0.f);
    float gy = fmaxf(fabsf(cy) - 0.5f * (hi_i.y - lo_i.y) - 0.5f * (hi_j.y - lo_j.y), 0.f);
    float gz = fmaxf(fabsf(cz) - 0.5f * (hi_i.z - lo_i.z) - 0.5f * (hi_j.z - lo_j.z), 0.f);
    return gx * gx + gy * gy + gz * gz;
}

__device__ __forceinline__ int lane_rank(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// min / max over the 8 lanes of each group of 8 (lane ^ 7, lane ^ 2, lane ^ 1 via DPP)
__device__ __forceinline__ float min8(float v) {
    v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xf, 0xf, true)));
    v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xf, 0xf, true)));
    return fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xf, 0xf, true)));
}
__device__ __forceinline__ float max8(float v) {
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xf, 0xf, true)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xf, 0xf, true)));
    return fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xf, 0xf, true)));
}

// tile list of one i-block: every j-block (j >= i for the half list) whose box is within
// the list radius r_l (r_c + skin + slack, TileRadius) of the i-block's box, ascending,
// each with a 64-bit mask of the j atoms that lie within r_l of the box of at least one
// of the block's eight 8-slot groups (atoms outside all of them cannot have a partner
// within r_l in the i-block; the union of the small boxes is a much tighter region than
// the block's box).  Half list: in the diagonal tile an atom only counts for groups that
// start below it.
// NW waves per i-block: the candidate j-blocks (box test) are listed in LDS, each
// wave computes the atom masks of every NW-th candidate, and wave 0 compacts the
// non-empty ones in candidate order.  Results do not depend on NW.
// Kept lists (tr.reuse_d2 >= 0): the workgroup returns at once while no atom of the
// replica has moved more than half the skin since the build (dmax); a build stores the
// block's slot coordinates in xb.
constexpr int TILE_CAND = 256;  // candidates per round
template <int NW>  // waves per i-block workgroup
__global__ __launch_bounds__(64 * NW) void k_tiles(int N, int Np, int NB, int maxj,
                                               const float4* __restrict__ posq_s,
                                               const float4* __restrict__ bb_lo,
                                               const float4* __restrict__ bb_hi,
                                               int* __restrict__ tiles,
                                               unsigned long long* __restrict__ tmask,
                                               int* __restrict__ ntiles,
                                               const float* __restrict__ dmax,
                                               float4* __restrict__ xb, TileRadius tr, Geom g,
                                               int half, int* __restrict__ flags) {
    __shared__ float4 sub[8][2];
    __shared__ int cand[TILE_CAND];
    __shared__ unsigned long long cmk[TILE_CAND];
    __shared__ int ncand_s, base_s;
    int r = blockIdx.y, ib = blockIdx.x;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const float list_r2 = tr.r2;
    // kept lists: every workgroup of the replica reads the same record (written by the
    // slot refresh or the sort before this launch), so all of them keep or all rebuild
    if (tr.reuse_d2 >= 0.f && dmax[(size_t)r * DISP_STRIDE] <= tr.reuse_d2) return;
    const float4* lo = bb_lo + (size_t)r * NB;
    const float4* hi = bb_hi + (size_t)r * NB;
    const float4* P = posq_s + (size_t)r * Np;
    float4 li = lo[ib], hi_i = hi[ib];
    // the eight group boxes (empty groups past N get a negative extent and never match)
    if (wave == 0) {
        int is = ib * 64 + lane;
        bool v = is < N;
        float4 p = P[min(is, N - 1)];
        if (tr.reuse_d2 >= 0.f && v) xb[(size_t)r * Np + is] = p;  // the build coordinates
        float lx = min8(v ? p.x : 3.0e38f), ly = min8(v ? p.y : 3.0e38f), lz = min8(v ? p.z : 3.0e38f);
        float hx = max8(v ? p.x : -3.0e38f), hy = max8(v ? p.y : -3.0e38f), hz = max8(v ? p.z : -3.0e38f);
        if ((lane & 7) == 0) {
            bool any = lx <= hx;
            sub[lane >> 3][0] = make_float4(0.5f * (lx + hx), 0.5f * (ly + hy), 0.5f * (lz + hz), 0.f);
            sub[lane >> 3][1] = any ? make_float4(0.5f * (hx - lx), 0.5f * (hy - ly), 0.5f * (hz - lz), 0.f)
                                    : make_float4(-1.0e30f, -1.0e30f, -1.0e30f, 0.f);
        }
    }
    // atom mask of candidate j-block jb2 (lane = atom)
    auto mask_of = [&](int jb2) {
        int s = jb2 * 64 + lane;
        float4 p = P[s];
        // one periodic shift for the whole tile when it is exact for every pair within
        // the list radius (TileRadius::oix); otherwise the minimum image per group
        float4 lj = lo[jb2], hj = hi[jb2];
        float dcx = 0.5f * (li.x + hi_i.x - lj.x - hj.x);
        float dcy = 0.5f * (li.y + hi_i.y - lj.y - hj.y);
        float dcz = 0.5f * (li.z + hi_i.z - lj.z - hj.z);
        float sx = g.bx * rintf(dcx * g.ibx), sy = g.by * rintf(dcy * g.iby),
              sz = g.bz * rintf(dcz * g.ibz);
        bool one_image =
            fabsf(dcx - sx) + 0.5f * (hi_i.x - li.x + hj.x - lj.x) < tr.oix &&
            fabsf(dcy - sy) + 0.5f * (hi_i.y - li.y + hj.y - lj.y) < tr.oiy &&
            fabsf(dcz - sz) + 0.5f * (hi_i.z - li.z + hj.z - lj.z) < tr.oiz;
        if (one_image) {
            p.x += sx;
            p.y += sy;
            p.z += sz;
        }
        // groups whose box is within rc of the j-block's box (lane q tests group q);
        // only those are tested atom by atom
        unsigned long long rel;
        {
            float4 c = sub[lane & 7][0], h = sub[lane & 7][1];
            float dx = 0.5f * (lj.x + hj.x) - c.x, dy = 0.5f * (lj.y + hj.y) - c.y,
                  dz = 0.5f * (lj.z + hj.z) - c.z;
            if (one_image) {
                dx += sx;
                dy += sy;
                dz += sz;
            } else {
                dx -= g.bx * rintf(dx * g.ibx);
                dy -= g.by * rintf(dy * g.iby);
                dz -= g.bz * rintf(dz * g.ibz);
            }
            float gx = fmaxf(fabsf(dx) - h.x - 0.5f * (hj.x - lj.x), 0.f);
            float gy = fmaxf(fabsf(dy) - h.y - 0.5f * (hj.y - lj.y), 0.f);
            float gz = fmaxf(fabsf(dz) - h.z - 0.5f * (hj.z - lj.z), 0.f);
            rel = __ballot(lane < 8 && h.x >= 0.f && gx * gx + gy * gy + gz * gz < list_r2);
        }
        // the atom tests, with the per-pair image only where the tile needs it and the
        // half-list rule only in the diagonal tile (both uniform: separate loops)
        auto atom_in = [&](auto mic, auto diagc) {
            constexpr bool MI = decltype(mic)::value, DIAG = decltype(diagc)::value;
            bool in = false;
            unsigned long long rq = rel;
            while (rq) {
                int q = __ffsll((long long)rq) - 1;
                rq &= rq - 1ull;
                float4 c = sub[q][0], h = sub[q][1];
                float dx = p.x - c.x, dy = p.y - c.y, dz = p.z - c.z;
                if (MI) {
                    dx -= g.bx * rintf(dx * g.ibx);
                    dy -= g.by * rintf(dy * g.iby);
                    dz -= g.bz * rintf(dz * g.ibz);
                }
                float gx = fmaxf(fabsf(dx) - h.x, 0.f), gy = fmaxf(fabsf(dy) - h.y, 0.f),
                      gz = fmaxf(fabsf(dz) - h.z, 0.f);
                bool inq = gx * gx + gy * gy + gz * gz < list_r2;
                if (DIAG) inq = inq && s > ib * 64 + q * 8;
                in = in || inq;
            }
            return in;
        };
        using T = std::true_type;
        using Fa = std::false_type;
        bool in;
        if (half && jb2 == ib) in = one_image ? atom_in(Fa(), T()) : atom_in(T(), T());
        else in = one_image ? atom_in(Fa(), Fa()) : atom_in(T(), Fa());
        return __ballot(in && (s < N));
    };
    int* out = tiles + ((size_t)r * NB + ib) * maxj;
    unsigned long long* mout = tmask + ((size_t)r * NB + ib) * maxj;
    int count = 0;    // wave 0's running output count
    int base = half ? (ib & ~63) : 0;
    while (base < NB) {
        // list up to TILE_CAND candidates (wave 0), continuing from base
        if (wave == 0) {
            int nc = 0;
            while (base < NB && nc + 64 <= TILE_CAND) {
                int jb = base + lane;
                bool ok = false;
                if (jb < NB && (!half || jb >= ib)) ok = bbox_gap2(li, hi_i, lo[jb], hi[jb], g) < list_r2;
                unsigned long long m = __ballot(ok);
                if (ok) cand[nc + lane_rank(m)] = jb;
                nc += __popcll(m);
                base += 64;
            }
            if (lane == 0) {
                ncand_s = nc;
                base_s = base;
            }
        }
        __syncthreads();
        int nc = ncand_s;
        base = base_s;  // every wave continues from wave 0's position (uniform loop)
        for (int c = wave; c < nc; c += NW) {
            // the candidate is the same for the whole wave: scalar, so that its box
            // loads and the uniform tests above compile to scalar code and branches
            unsigned long long am = mask_of(__builtin_amdgcn_readfirstlane(cand[c]));
            if (lane == 0) cmk[c] = am;
        }
        __syncthreads();
        if (wave == 0) {
            for (int c0 = 0; c0 < nc; c0 += 64) {
                int c = c0 + lane;
                unsigned long long am = c < nc ? cmk[c] : 0ull;
                unsigned long long keep = __ballot(am != 0ull);
                int pos = count + lane_rank(keep);
                if (am != 0ull && pos < maxj) {
                    out[pos] = cand[c];
                    mout[pos] = am;
                }
                count += __popcll(keep);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ntiles[(size_t)r * NB + ib] = min(count, maxj);
        if (count > maxj) atomicOr(&flags[r], 4);
    }
}

// ---------------------------------------------------------------------------
// the hot kernel: U0 and F0 over all pairs at x.  One wave per (replica, i-block):
// lane = i atom (registers), j atoms are wave-uniform (scalar loads) taken from each
// tile's j mask.
//   NEWTON = false: full tile list, every ordered pair once, no j forces, energy halved.
//   NEWTON = true:  half tile list (j-block >= i-block, j > i inside the diagonal tile),
//                   every pair once; the j forces of 8 j atoms at a time are summed over
//                   the wave by a reduce-scatter and added with int64 atomics.
// When one shift is exact for every pair of a tile within the list radius (Geom::oix)
// the periodic image is that shift for the whole tile (applied to the i coordinates);
// otherwise each pair takes the minimum image.  Pair energies are accumulated in FP64 per pair and the per-tile FP32 force and
// reaction-field sums in FP64 in tile order; fixed-point atomics commute, so results are
// bit-reproducible.
// ---------------------------------------------------------------------------
struct PairAcc {
    float fx, fy, fz;  // per-tile force on the lane's i atom
    float sqr2, sq;    // per-tile sums of qq r^2 and qq (reaction-field energy terms)
    double e;          // LJ + qq/r energy in FP64 (net energies cancel strongly)
    float ef;          // the same over the current chunk of 8 pairs, added to e per chunk
    int cnt;
};

// one pair (i on the lane, j uniform); returns F/r (0 outside the cutoff or excluded)
template <bool MINIMG, bool EXCL>
__device__ __forceinline__ float pair_one(float4 pi, float2 qi, float4 pj, float2 qj,
                                          bool extra_ok, const Geom& g, PairAcc& a,
                                          float& dx, float& dy, float& dz) {
    dx = pi.x - pj.x;
    dy = pi.y - pj.y;
    dz = pi.z - pj.z;
    if (MINIMG) {
        dx -= g.bx * rintf(dx * g.ibx);
        dy -= g.by * rintf(dy * g.iby);
        dz -= g.bz * rintf(dz * g.ibz);
    }
    float r2 = dx * dx + dy * dy + dz * dz;
    bool ok = r2 < g.rc2f;
    if (EXCL) ok = ok && extra_ok;
    float ir = __frsqrt_rn(r2);
    float ir2 = ir * ir;
    float sig = qi.x + qj.x;
    float s2 = sig * sig * ir2;
    float s6 = s2 * s2 * s2;
    float es6 = (qi.y * qj.y) * s6;
    float qq = pi.w * pj.w;
    float fr = es6 * (12.0f * s6 - 6.0f) * ir2 + qq * (ir * ir2 - 2.0f * g.krff);
    fr = ok ? fr : 0.0f;
    // E = LJ + qq/r here; the reaction-field terms qq(k_rf r^2 - c_rf) are summed as
    // sq*r^2 and sq and scaled by k_rf, c_rf in FP64 (their FP32 rounding would bias a
    // sum over millions of pairs)
    float ee = ok ? (es6 * (s6 - 1.0f) + qq * ir) : 0.0f;
    float qqm = ok ? qq : 0.0f;
    a.fx += fr * dx;
    a.fy += fr * dy;
    a.fz += fr * dz;
    a.e += (double)ee;
    a.sqr2 += qqm * r2;
    a.sq += qqm;
    a.cnt += ok ? 1 : 0;
    return fr;
}

// One pair, branch free: the result is computed for every lane and cleared with a bit
// mask outside the cutoff (a never-taken branch per pair would cost more and would keep
// the scheduler from overlapping consecutive pairs).  Masked lanes may hold inf or NaN
// before the mask (r = 0 for an unused slot); the AND yields +0.
// ENERGY = false: forces only (the energy terms and their sums are left out).  GEO: the
// geometric sigma rule, qi.x and qj.x holding sigma itself (Geom::geo)
template <bool ENERGY = true, bool GEO = false>
__device__ __forceinline__ float pair_masked(float4 pi, float2 qi, float4 pj, float2 qj,
                                             bool incl, bool minimg, const Geom& g, PairAcc& a,
                                             float& dx, float& dy, float& dz) {
    dx = pi.x - pj.x;
    dy = pi.y - pj.y;
    dz = pi.z - pj.z;
    if (minimg) {  // wave-uniform
        dx -= g.bx * rintf(dx * g.ibx);
        dy -= g.by * rintf(dy * g.iby);
        dz -= g.bz * rintf(dz * g.ibz);
    }
    float r2 = dx * dx + dy * dy + dz * dz;
    unsigned m = (incl && r2 < g.rc2f) ? 0xFFFFFFFFu : 0u;
    float ir = __frsqrt_rn(r2);
    float ir2 = ir * ir;
    float s2;
    if constexpr (GEO) {
        s2 = (qi.x * qj.x) * ir2;
    } else {
        float sig = qi.x + qj.x;
        s2 = sig * sig * ir2;
    }
    float s6 = s2 * s2 * s2;
    float es6 = (qi.y * qj.y) * s6;
    float qq = pi.w * pj.w;
    float fr = es6 * (12.0f * s6 - 6.0f) * ir2 + qq * (ir * ir2 - 2.0f * g.krff);
    fr = __uint_as_float(__float_as_uint(fr) & m);
    a.fx += fr * dx;
    a.fy += fr * dy;
    a.fz += fr * dz;
    a.cnt += (int)(m & 1u);
    // pin the sums here: otherwise the compiler sinks every pair's accumulation to the
    // end of the chunk and keeps all eight pairs' operands live (VGPR pressure)
    if (ENERGY) {
        float ee = es6 * (s6 - 1.0f) + qq * ir;
        ee = __uint_as_float(__float_as_uint(ee) & m);
        float qqm = __uint_as_float(__float_as_uint(qq) & m);
        a.ef += ee;
        a.sqr2 += qqm * r2;
        a.sq += qqm;
        asm volatile("" : "+v"(a.fx), "+v"(a.fy), "+v"(a.fz), "+v"(a.ef), "+v"(a.sqr2), "+v"(a.sq));
    } else {
        asm volatile("" : "+v"(a.fx), "+v"(a.fy), "+v"(a.fz));
    }
    return fr;
}

// a + b taken from the DPP partner lane.  bound_ctrl lets the mov fold into the add as
// its DPP operand; contraction is off so the add is not fused with the product that
// made a (gfx9 has no DPP form of v_fma).
template <int CTRL>
__device__ __forceinline__ float dpp_add(float a, float b) {
#pragma clang fp contract(off)
    return a + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(b), CTRL, 0xf, 0xf, true));
}

// Maps workgroup ids so that consecutive (replica, block) work lands on one XCD:
// workgroups are dispatched to the 8 XCDs round-robin, each XCD has its own L2.
__device__ __forceinline__ int xcd_linear_id(int id, int n) {
    int per = n >> 3, rem = n & 7, x = id & 7, q = id >> 3;
    return x * per + min(x, rem) + q;
}

// s_waitcnt vmcnt(0) alone (expcnt and lgkmcnt at their no-wait maxima).  The
// scheduling barrier keeps the arithmetic before it from being moved past the wait.
__device__ __forceinline__ void wait_vmem() {
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __builtin_amdgcn_sched_barrier(0);
}

__device__ __forceinline__ int next_bit(unsigned long long& m) {
    int j = m ? __ffsll((long long)m) - 1 : -1;
    m &= m - 1ull;
    return j;
}

// full evaluation: all j of one tile, four at a time so their scalar loads overlap
template <bool MINIMG, bool EXCL>
__device__ __forceinline__ void tile_full(const float4* __restrict__ P, const float2* __restrict__ Q,
                                          int jbeg, unsigned long long am, float4 pi, float2 qi,
                                          unsigned long long em, const Geom& g, PairAcc& a) {
    float dx, dy, dz;
    while (am) {
        int j0 = next_bit(am), j1 = next_bit(am), j2 = next_bit(am), j3 = next_bit(am);
        float4 p0 = P[jbeg + j0];
        float2 q0 = Q[jbeg + j0];
        float4 p1 = P[jbeg + max(j1, 0)];
        float2 q1 = Q[jbeg + max(j1, 0)];
        float4 p2 = P[jbeg + max(j2, 0)];
        float2 q2 = Q[jbeg + max(j2, 0)];
        float4 p3 = P[jbeg + max(j3, 0)];
        float2 q3 = Q[jbeg + max(j3, 0)];
        pair_one<MINIMG, EXCL>(pi, qi, p0, q0, !((em >> j0) & 1ull), g, a, dx, dy, dz);
        if (j1 >= 0) pair_one<MINIMG, EXCL>(pi, qi, p1, q1, !((em >> j1) & 1ull), g, a, dx, dy, dz);
        if (j2 >= 0) pair_one<MINIMG, EXCL>(pi, qi, p2, q2, !((em >> j2) & 1ull), g, a, dx, dy, dz);
        if (j3 >= 0) pair_one<MINIMG, EXCL>(pi, qi, p3, q3, !((em >> j3) & 1ull), g, a, dx, dy, dz);
    }
}

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Sum v[24] (8 j atoms x 3 components, index 3k+c) over the 64 lanes so that lanes
// 8k..8k+7 end up holding the three components of j atom k.  Three halving stages split
// the values between lane halves (bit 5: v_permlane32_swap, bit 4: v_permlane16_swap,
// bit 3: DPP row_ror:8), then a DPP sum over the 8 lanes of each group.  The swaps
// exchange the upper half of their first operand with the lower half of their second,
// so r0 + r1 is the pairwise sum of the value each half keeps, with no selects.
__device__ __forceinline__ void reduce_scatter24(const float (&v)[24], int lane, float (&out)[3]) {
    float w[12], x[6];
#pragma unroll
    for (int t = 0; t < 12; t++) {
        auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[t]), __float_as_uint(v[t + 12]),
                                                   false, false);
        w[t] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
#pragma unroll
    for (int t = 0; t < 6; t++) {
        auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(w[t]), __float_as_uint(w[t + 6]),
                                                   false, false);
        x[t] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    bool b3 = (lane & 8) != 0;
#pragma unroll
    for (int t = 0; t < 3; t++) {
        float keep = b3 ? x[t + 3] : x[t];
        float send = b3 ? x[t] : x[t + 3];
        float y = dpp_add<0x128>(keep, send);  // row_ror:8 = lane ^ 8 within 16
        y = dpp_add<0xB1>(y, y);               // quad_perm [1,0,3,2] = lane ^ 1
        y = dpp_add<0x4E>(y, y);               // quad_perm [2,3,0,1] = lane ^ 2
        out[t] = dpp_add<0x141>(y, y);         // row_half_mirror: lane i <- 7 - i within 8
    }
}

// Sum v[12] (4 j atoms x 3 components, index 3k+c) over the 64 lanes so that lanes
// 16k..16k+15 end up holding the three components of j atom k: two halving stages
// (v_permlane32_swap on lane bit 5, v_permlane16_swap on bit 4), then DPP sums over the
// 16 lanes of each row (row_ror:8, the two quad swaps, row_half_mirror).
__device__ __forceinline__ void reduce_scatter12(const float (&v)[12], float (&out)[3]) {
    float w[6];
#pragma unroll
    for (int t = 0; t < 6; t++) {
        auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[t]), __float_as_uint(v[t + 6]),
                                                   false, false);
        w[t] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
#pragma unroll
    for (int t = 0; t < 3; t++) {
        auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(w[t]), __float_as_uint(w[t + 3]),
                                                   false, false);
        float y = __uint_as_float(r[0]) + __uint_as_float(r[1]);
        y = dpp_add<0x128>(y, y);      // row_ror:8 = lane ^ 8 within 16
        y = dpp_add<0xB1>(y, y);       // quad_perm [1,0,3,2] = lane ^ 1
        y = dpp_add<0x4E>(y, y);       // quad_perm [2,3,0,1] = lane ^ 2
        out[t] = dpp_add<0x141>(y, y); // row_half_mirror: lane i <- 7 - i within 8
    }
}

// The same for v[6] (2 j atoms): lanes 32k..32k+31 end up holding j atom k.  Each sum
// pairs lanes in the same order as reduce_scatter24 and reduce_scatter12 (lane ^ 32,
// ^ 16, ^ 8, ^ 1, ^ 2, half mirror), so a j atom's sum does not depend on the width of
// the chunk it was taken in.
__device__ __forceinline__ void reduce_scatter6(const float (&v)[6], float (&out)[3]) {
#pragma unroll
    for (int t = 0; t < 3; t++) {
        auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[t]), __float_as_uint(v[t + 3]),
                                                   false, false);
        float w = __uint_as_float(r[0]) + __uint_as_float(r[1]);
        auto q = __builtin_amdgcn_permlane16_swap(__float_as_uint(w), __float_as_uint(w), false,
                                                   false);
        float y = __uint_as_float(q[0]) + __uint_as_float(q[1]);
        y = dpp_add<0x128>(y, y);      // row_ror:8 = lane ^ 8 within 16
        y = dpp_add<0xB1>(y, y);       // quad_perm [1,0,3,2] = lane ^ 1
        y = dpp_add<0x4E>(y, y);       // quad_perm [2,3,0,1] = lane ^ 2
        out[t] = dpp_add<0x141>(y, y); // row_half_mirror: lane i <- 7 - i within 8
    }
}

// Newton-3 evaluation of one tile: j atoms in groups of 8; their forces are scattered
// to the slot-ordered fixed-point accumulator F (3 planes of Np).  Slots past the last j
// atom re-read the first one and are masked off.
template <bool MINIMG, bool EXCL>
__device__ __forceinline__ void tile_newton(const float4* __restrict__ P, const float2* __restrict__ Q,
                                            int jbeg, unsigned long long am, float4 pi, float2 qi,
                                            bool vi, unsigned long long em, const Geom& g,
                                            PairAcc& a, unsigned long long* __restrict__ Fr,
                                            int Np, int lane, int& bad) {
    while (am) {
        int jj[8];
#pragma unroll
        for (int k = 0; k < 8; k++) jj[k] = next_bit(am);
        float v[24];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            int j = jj[k] < 0 ? jj[0] : jj[k];
            float4 pj = P[jbeg + j];
            float2 qj = Q[jbeg + j];
            bool incl = vi && jj[k] >= 0;
            if (EXCL) incl = incl && !((em >> j) & 1ull);
            float dx, dy, dz;
            float fr = pair_masked(pi, qi, pj, qj, incl, MINIMG, g, a, dx, dy, dz);
            v[3 * k] = -fr * dx;
            v[3 * k + 1] = -fr * dy;
            v[3 * k + 2] = -fr * dz;
        }
        float f[3];
        reduce_scatter24(v, lane, f);
        a.e += (double)a.ef;  // FP32 over 8 pairs, FP64 beyond
        a.ef = 0.f;
        int k = lane >> 3;
        int mine = jj[0];
#pragma unroll
        for (int q = 1; q < 8; q++) mine = (k == q) ? jj[q] : mine;
        if ((lane & 7) == 0 && mine >= 0) {
            int s = jbeg + mine;
            atomicAdd(&Fr[s], fixed_from_float(f[0], &bad));
            atomicAdd(&Fr[Np + s], fixed_from_float(f[1], &bad));
            atomicAdd(&Fr[2 * Np + s], fixed_from_float(f[2], &bad));
        }
    }
}

template <int WPW, bool COUNT, bool NEWTON>
__global__ __launch_bounds__(64 * WPW) void k_pair(
    int R, int N, int Np, int NB, int maxj, const float4* __restrict__ posq_s,
    const float2* __restrict__ par_s, const int* __restrict__ orig_s,
    const int* __restrict__ slot_of, const int* __restrict__ tiles,
    const unsigned long long* __restrict__ tmask, const int* __restrict__ ntiles,
    const float4* __restrict__ bb_lo, const float4* __restrict__ bb_hi,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list, int max_ex,
    long long* __restrict__ F, double* __restrict__ e_pair, int* __restrict__ flags,
    unsigned long long* __restrict__ pair_count, Geom g) {
    extern __shared__ int exs_lds_dyn[];  // [WPW][max_ex][64] exclusion partners' slots
    // wave-uniform by construction; readfirstlane lets the compiler keep everything
    // derived from it (replica, block, tile, j index) in scalar registers
    int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int w = xcd_linear_id(blockIdx.x, gridDim.x) * WPW + wv;
    int lane = threadIdx.x & 63;
    if (w >= R * NB) return;
    int r = w / NB;
    int ib = w - r * NB;
    const float4* P = posq_s + (size_t)r * Np;
    const float2* Q = par_s + (size_t)r * Np;
    unsigned long long* Fr = (unsigned long long*)(F + (size_t)r * 3 * Np);
    int islot = ib * 64 + lane;
    bool vi = islot < N;
    float4 pi = P[islot];
    float2 qi = Q[islot];
    int nex = 0;
    if (vi) {
        int ai = orig_s[(size_t)r * Np + islot];
        int b0 = ex_start[ai];
        nex = ex_start[ai + 1] - b0;
        for (int k = 0; k < nex; k++)
            exs_lds_dyn[(wv * max_ex + k) * 64 + lane] = slot_of[(size_t)r * Np + ex_list[b0 + k]];
    }
    float4 bli = bb_lo[(size_t)r * NB + ib], bhi = bb_hi[(size_t)r * NB + ib];
    double ax = 0.0, ay = 0.0, az = 0.0, aqr2 = 0.0, aq = 0.0;
    PairAcc acc;
    acc.e = 0.0;
    acc.ef = 0.f;
    acc.cnt = 0;
    int bad = 0;
    const int* tl = tiles + ((size_t)r * NB + ib) * maxj;
    const unsigned long long* tm = tmask + ((size_t)r * NB + ib) * maxj;
    int nt = ntiles[(size_t)r * NB + ib];
    for (int t = 0; t < nt; t++) {
        int jb = tl[t];
        unsigned long long am = tm[t];
        unsigned long long em = 0;
        for (int k = 0; k < nex; k++) {
            int sl = exs_lds_dyn[(wv * max_ex + k) * 64 + lane];
            if ((sl >> 6) == jb) em |= 1ull << (sl & 63);
        }
        if (jb == ib) em |= NEWTON ? ((2ull << lane) - 1ull) : (1ull << lane);
        // one periodic shift for the whole tile when it is exact for every pair within
        // the list radius (Geom::oix); otherwise the minimum image per pair
        float4 blj = bb_lo[(size_t)r * NB + jb], bhj = bb_hi[(size_t)r * NB + jb];
        float dcx = 0.5f * (bli.x + bhi.x - blj.x - bhj.x);
        float dcy = 0.5f * (bli.y + bhi.y - blj.y - bhj.y);
        float dcz = 0.5f * (bli.z + bhi.z - blj.z - bhj.z);
        float sx = g.bx * rintf(dcx * g.ibx), sy = g.by * rintf(dcy * g.iby),
              sz = g.bz * rintf(dcz * g.ibz);
        bool one_image =
            fabsf(dcx - sx) + 0.5f * (bhi.x - bli.x + bhj.x - blj.x) < g.oix &&
            fabsf(dcy - sy) + 0.5f * (bhi.y - bli.y + bhj.y - blj.y) < g.oiy &&
            fabsf(dcz - sz) + 0.5f * (bhi.z - bli.z + bhj.z - blj.z) < g.oiz;
        float4 ps = pi;
        if (one_image) {
            ps.x -= sx;
            ps.y -= sy;
            ps.z -= sz;
        }
        acc.fx = acc.fy = acc.fz = acc.sqr2 = acc.sq = 0.f;
        bool ex = __ballot(em != 0ull) != 0ull;
        int jbeg = jb * 64;
        if (NEWTON) {
            if (one_image) {
                if (ex) tile_newton<false, true>(P, Q, jbeg, am, ps, qi, vi, em, g, acc, Fr, Np, lane, bad);
                else tile_newton<false, false>(P, Q, jbeg, am, ps, qi, vi, em, g, acc, Fr, Np, lane, bad);
            } else {
                if (ex) tile_newton<true, true>(P, Q, jbeg, am, ps, qi, vi, em, g, acc, Fr, Np, lane, bad);
                else tile_newton<true, false>(P, Q, jbeg, am, ps, qi, vi, em, g, acc, Fr, Np, lane, bad);
            }
        } else {
            if (one_image) {
                if (ex) tile_full<false, true>(P, Q, jbeg, am, ps, qi, em, g, acc);
                else tile_full<false, false>(P, Q, jbeg, am, ps, qi, em, g, acc);
            } else {
                if (ex) tile_full<true, true>(P, Q, jbeg, am, ps, qi, em, g, acc);
                else tile_full<true, false>(P, Q, jbeg, am, ps, qi, em, g, acc);
            }
        }
        ax += (double)acc.fx;
        ay += (double)acc.fy;
        az += (double)acc.fz;
        aqr2 += (double)acc.sqr2;
        aq += (double)acc.sq;
    }
    if (COUNT) {
        double c = wave_sum(vi ? (double)acc.cnt : 0.0);
        if (lane == 0) atomicAdd(&pair_count[r], (unsigned long long)(c + 0.5));
        return;
    }
    if (vi) {  // atomic in both modes: the bonded kernel may run concurrently
        long long fx = to_fixed(ax, &bad), fy = to_fixed(ay, &bad), fz = to_fixed(az, &bad);
        atomicAdd(&Fr[islot], (unsigned long long)fx);
        atomicAdd(&Fr[Np + islot], (unsigned long long)fy);
        atomicAdd(&Fr[2 * Np + islot], (unsigned long long)fz);
    }
    double es = wave_sum(vi ? (acc.e + g.krf * aqr2 - g.crf * aq) : 0.0);
    if (__ballot(bad) != 0ull && lane == 0) atomicOr(&flags[r], 2);
    if (lane == 0) e_pair[(size_t)r * NB + ib] = NEWTON ? es : 0.5 * es;
}

// ---------------------------------------------------------------------------
// Block kernel, Newton-3, with LDS-staged j tiles.  Same work split as k_pair<.., true>
// (one wave per i-block, half tile list, j wave-uniform; or `split` waves per i-block,
// each taking every split-th tile, with their own i-force atomics and energy slot, for
// batches too small to fill the GPU with one wave per block), but the 64 slot records of
// each j-block are copied global -> LDS by DMA one tile ahead (no registers, no scalar
// loads in the pair loop), and the j forces of a tile are kept in LDS (each j atom's
// sum at its rank in the tile's mask: every j atom falls in exactly one chunk) and
// added to the global accumulator once per tile with coalesced atomics.
//   LDS per wave: two 2 KB tile buffers (separate objects, so reading one never waits
//   for the copy into the other), 768 B of j-force sums, max_ex*256 B of exclusions.
//   vmcnt discipline: the only vector memory operations inside the tile loop are the
//   next tile's copy and the three flush atomics that end each tile.  Loads and
//   stores/atomics share vmcnt but need not retire in issue order, so the copy is
//   waited for with vmcnt(0) before the atomics, not with vmcnt(3) after them.
//   ENERGY = false: forces only; e_pair is not written.
// ---------------------------------------------------------------------------

// At least 5 waves per SIMD (at most 96 VGPRs, allocated in granules of 8): the energy
// variant otherwise takes 97, which leaves room for 4.  The forces-only variant takes
// 84; held to 80 for 6 waves it spills 24 bytes per lane and runs 3.3% slower.
template <int WPW, bool COUNT, bool ENERGY, bool GEO = false>
__global__ __launch_bounds__(64 * WPW) __attribute__((amdgpu_waves_per_eu(5))) void k_pair_n3(
    int R, int N, int Np, int NB, int maxj, int split, const float4* __restrict__ arec,
    const int* __restrict__ orig_s, const int* __restrict__ slot_of, const int* __restrict__ tiles,
    const unsigned long long* __restrict__ tmask, const int* __restrict__ ntiles,
    const float4* __restrict__ bb_lo, const float4* __restrict__ bb_hi,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list, int max_ex,
    long long* __restrict__ F, double* __restrict__ e_pair, int* __restrict__ flags,
    unsigned long long* __restrict__ pair_count, Geom g) {
    // two tile buffers of 64 slot records (2 float4 each) plus 8 zero records: a chunk's
    // entries past the tile's mask read there or the tile's own finite records
    __shared__ float4 tile_a[WPW][144], tile_b[WPW][144];
    __shared__ float facc[WPW][3][64];
    extern __shared__ int exs_dyn[];
    int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int w = xcd_linear_id(blockIdx.x, gridDim.x) * WPW + wv;
    int lane = threadIdx.x & 63;
    if (w >= R * NB * split) return;
    // split waves per (replica, i-block), wave `part` taking tiles part, part + split, ..
    const int part = w % split;
    const int rb = w / split;
    int r = rb / NB;
    int ib = rb - r * NB;
    const float4* A = arec + (size_t)r * Np * 2;
    unsigned long long* Fr = (unsigned long long*)(F + (size_t)r * 3 * Np);
    int* exs = exs_dyn + wv * max_ex * 64;
    int islot = ib * 64 + lane;
    bool vi = islot < N;
    float4 pi = A[2 * islot];
    float4 qi4 = A[2 * islot + 1];
    float2 qi = make_float2(qi4.x, qi4.y);
    int nex = 0;
    // the range of j-blocks that hold an exclusion partner of one of the wave's i atoms
    // (bonded partners are close: mostly this block and its neighbours); tiles outside
    // it skip the exclusion test (wave-uniform)
    int exlo = NB, exhi = -1;
    if (vi) {
        int ai = orig_s[(size_t)r * Np + islot];
        int b0 = ex_start[ai];
        nex = ex_start[ai + 1] - b0;
        for (int k = 0; k < nex; k++) {
            const int sl = slot_of[(size_t)r * Np + ex_list[b0 + k]];
            exs[k * 64 + lane] = sl;
            exlo = min(exlo, sl >> 6);
            exhi = max(exhi, sl >> 6);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        exlo = min(exlo, __shfl_xor(exlo, o, 64));
        exhi = max(exhi, __shfl_xor(exhi, o, 64));
    }
    exlo = __builtin_amdgcn_readfirstlane(exlo);
    exhi = __builtin_amdgcn_readfirstlane(exhi);
    float4 bli = bb_lo[(size_t)r * NB + ib], bhi = bb_hi[(size_t)r * NB + ib];
    if (lane < 16) {
        tile_a[wv][128 + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
        tile_b[wv][128 + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    double ax = 0.0, ay = 0.0, az = 0.0, aqr2 = 0.0, aq = 0.0;
    PairAcc acc;
    acc.e = 0.0;
    acc.ef = 0.f;
    acc.cnt = 0;
    int bad = 0;
    const int* tl = tiles + ((size_t)r * NB + ib) * maxj;
    const unsigned long long* tmk = tmask + ((size_t)r * NB + ib) * maxj;
    int nt = ntiles[(size_t)r * NB + ib];

    // copies the 64 slot records (32 B) of j-block jb into tile buffer B: lane L moves
    // half L&1 of record L>>1 (first instruction) and of record 32 + (L>>1) (second)
    auto stage = [&](int jb, auto bufc) {
        constexpr int B = decltype(bufc)::value;
        float4* dst = B ? tile_b[wv] : tile_a[wv];
        const char* src = (const char*)(A + 2 * jb * 64) + lane * 16;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_global_load_lds((const void*)src,
                                         (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
        __builtin_amdgcn_global_load_lds((const void*)(src + 1024),
                                         (__attribute__((address_space(3))) void*)(dst + 64), 16, 0, 0);
        __builtin_amdgcn_wave_barrier();
    };

    auto do_tile = [&](int t, auto bufc) {
        constexpr int B = decltype(bufc)::value;
        const float4* buf = B ? tile_b[wv] : tile_a[wv];
        int jb = tl[t];
        unsigned long long am = tmk[t];
        const unsigned long long am0 = am;
        if (t + split < nt) stage(tl[t + split], std::integral_constant<int, 1 - B>());
        unsigned long long em = 0;
        if (jb >= exlo && jb <= exhi) {
            for (int k = 0; k < nex; k++) {
                int sl = exs[k * 64 + lane];
                if ((sl >> 6) == jb) em |= 1ull << (sl & 63);
            }
        }
        if (jb == ib) em |= (2ull << lane) - 1ull;
        // one periodic shift for the whole tile when it is exact for every pair within
        // the list radius (Geom::oix); otherwise the minimum image per pair
        float4 blj = bb_lo[(size_t)r * NB + jb], bhj = bb_hi[(size_t)r * NB + jb];
        float dcx = 0.5f * (bli.x + bhi.x - blj.x - bhj.x);
        float dcy = 0.5f * (bli.y + bhi.y - blj.y - bhj.y);
        float dcz = 0.5f * (bli.z + bhi.z - blj.z - bhj.z);
        float sx = g.bx * rintf(dcx * g.ibx), sy = g.by * rintf(dcy * g.iby),
              sz = g.bz * rintf(dcz * g.ibz);
        bool one_image =
            fabsf(dcx - sx) + 0.5f * (bhi.x - bli.x + bhj.x - blj.x) < g.oix &&
            fabsf(dcy - sy) + 0.5f * (bhi.y - bli.y + bhj.y - blj.y) < g.oiy &&
            fabsf(dcz - sz) + 0.5f * (bhi.z - bli.z + bhj.z - blj.z) < g.oiz;
        float4 ps = pi;
        if (one_image) {
            ps.x -= sx;
            ps.y -= sy;
            ps.z -= sz;
        }
        const bool ex = __ballot(em != 0ull) != 0ull;
        // the mask's records to the front of the buffer in mask order (entry e = the e-th
        // set bit), so that a chunk reads its W records at fixed offsets from one address.
        // Every lane reads its own record before any lane writes (one wave: its LDS
        // accesses are carried out in order); the records past the mask stay finite.
        {
            float4* bw = B ? tile_b[wv] : tile_a[wv];
            const float4 r0 = bw[2 * lane], r1 = bw[2 * lane + 1];
            if ((am0 >> lane) & 1ull) {
                const int rank = lane_rank(am0);
                bw[2 * rank] = r0;
                bw[2 * rank + 1] = r1;
            }
        }
        acc.fx = acc.fy = acc.fz = acc.sqr2 = acc.sq = 0.f;
        // all j atoms of the tile, 8 at a time; MI: per-pair minimum image, EX: the tile
        // holds excluded pairs (both wave-uniform, compile-time inside the loop).  ent:
        // entries (j atoms in mask order) taken so far, nvalid: entries of this chunk.
        int ent = 0, nvalid = 0;
        auto chunks = [&](auto mic, auto exc) {
            constexpr bool MI = decltype(mic)::value, EX = decltype(exc)::value;
            // the next W j atoms of the tile, W = 8, 4 or 2: a chunk costs about 310, 166
            // and 100 VALU instructions, so the tail of a tile is taken in narrower ones
            auto chunk = [&](auto wc) {
                constexpr int W = decltype(wc)::value;
                constexpr int SH = W == 8 ? 3 : W == 4 ? 4 : 5;  // lanes per j atom: 1 << SH
                int jj[W];
#pragma unroll
                for (int k = 0; k < W; k++) jj[k] = next_bit(am);
                float v[3 * W];
                const float4* cb = buf + 2 * ent;  // the chunk's records, in entry order
#pragma unroll
                for (int k = 0; k < W; k++) {
                    float4 pj = cb[2 * k];
                    float2 qj = *(const float2*)&cb[2 * k + 1];
                    bool incl = vi && k < nvalid;
                    if (EX) incl = incl && !((em >> (jj[k] & 63)) & 1ull);
                    float dx, dy, dz;
                    float fr =
                        pair_masked<ENERGY, GEO>(ps, qi, pj, qj, incl, MI, g, acc, dx, dy, dz);
                    v[3 * k] = -fr * dx;
                    v[3 * k + 1] = -fr * dy;
                    v[3 * k + 2] = -fr * dz;
                }
                float f[3];
                if constexpr (W == 8) reduce_scatter24(v, lane, f);
                else if constexpr (W == 4) reduce_scatter12(v, f);
                else reduce_scatter6(v, f);
                if (ENERGY) {
                    acc.e += (double)acc.ef;  // FP32 over W pairs, FP64 beyond
                    acc.ef = 0.f;
                }
                // Lanes (lane >> SH) << SH hold the sum of entry ent + (lane >> SH): every
                // j atom of the tile is one entry of one chunk, taken in mask order, so
                // its sum is written (not added) at its rank in the tile's mask.  Written
                // out because the compiler would otherwise put an s_waitcnt vmcnt(0) (a
                // wait for the tile copy in flight) in front of any LDS access it cannot
                // separate from that copy.
                const int kq = lane >> SH;
                if ((lane & ((1 << SH) - 1)) == 0 && kq < nvalid) {
                    unsigned addr = (unsigned)(size_t)&facc[wv][0][ent + kq];
                    asm volatile(
                        "ds_write_b32 %0, %1\n\t"
                        "ds_write_b32 %0, %2 offset:256\n\t"
                        "ds_write_b32 %0, %3 offset:512"
                        :
                        : "v"(addr), "v"(f[0]), "v"(f[1]), "v"(f[2])
                        : "memory");
                }
            };
            while (am) {
                const int n = __popcll(am);
                if (n >= 7) {
                    nvalid = min(n, 8);
                    chunk(std::integral_constant<int, 8>());
                } else if (n >= 3) {
                    nvalid = min(n, 4);
                    chunk(std::integral_constant<int, 4>());
                } else {
                    nvalid = n;
                    chunk(std::integral_constant<int, 2>());
                }
                ent += nvalid;
            }
        };
        using T = std::true_type;
        using Fa = std::false_type;
        if (one_image) {
            if (ex) chunks(Fa(), T());
            else chunks(Fa(), Fa());
        } else {
            if (ex) chunks(T(), T());
            else chunks(T(), Fa());
        }
        ax += (double)acc.fx;
        ay += (double)acc.fy;
        az += (double)acc.fz;
        if (ENERGY) {
            aqr2 += (double)acc.sqr2;
            aq += (double)acc.sq;
        }
        // flush the tile's j forces: lane = j atom, which reads its sum at its rank in
        // the mask (bits of am0 below the lane), coalesced fixed-point atomics
        float fjx, fjy, fjz;
        {
            const int rank = __builtin_amdgcn_mbcnt_hi(
                (unsigned)(am0 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)am0, 0u));
            unsigned addr = (unsigned)(size_t)&facc[wv][0][rank];
            // early-clobber outputs: the first load's result lands while the other two
            // still read the address, so no output may share the address register
            asm volatile(
                "ds_read_b32 %0, %3\n\t"
                "ds_read_b32 %1, %3 offset:256\n\t"
                "ds_read_b32 %2, %3 offset:512\n\t"
                "s_waitcnt lgkmcnt(0)"
                : "=&v"(fjx), "=&v"(fjy), "=&v"(fjz)
                : "v"(addr)
                : "memory");
        }
        // the next tile's records have landed (and the previous tile's atomics, issued a
        // tile earlier, are done): waited for before this tile's atomics, because a load
        // and later atomics share vmcnt but may retire out of order
        wait_vmem();
        if (!COUNT && ((am0 >> lane) & 1ull)) {
            int s = jb * 64 + lane;
            atomicAdd(&Fr[s], fixed_from_float(fjx, &bad));
            atomicAdd(&Fr[Np + s], fixed_from_float(fjy, &bad));
            atomicAdd(&Fr[2 * Np + s], fixed_from_float(fjz, &bad));
        }
    };

    if (part < nt) stage(tl[part], std::integral_constant<int, 0>());
    wait_vmem();  // the first tile, this lane's atom and the block boxes have arrived
    for (int t = part; t < nt; t += 2 * split) {
        do_tile(t, std::integral_constant<int, 0>());
        if (t + split < nt) do_tile(t + split, std::integral_constant<int, 1>());
    }
    if (COUNT) {
        double c = wave_sum(vi ? (double)acc.cnt : 0.0);
        if (lane == 0) atomicAdd(&pair_count[r], (unsigned long long)(c + 0.5));
        return;
    }
    if (vi) {  // atomic: the bonded kernel may run concurrently
        long long fx = to_fixed(ax, &bad), fy = to_fixed(ay, &bad), fz = to_fixed(az, &bad);
        atomicAdd(&Fr[islot], (unsigned long long)fx);
        atomicAdd(&Fr[Np + islot], (unsigned long long)fy);
        atomicAdd(&Fr[2 * Np + islot], (unsigned long long)fz);
    }
    if (__ballot(bad) != 0ull && lane == 0) atomicOr(&flags[r], 2);
    if (ENERGY) {
        double es = wave_sum(vi ? (acc.e + g.krf * aqr2 - g.crf * aq) : 0.0);
        if (lane == 0) e_pair[((size_t)r * NB + ib) * split + part] = es;
    }
}

// ---------------------------------------------------------------------------
// Grouped pair kernel (Newton-3, each pair once).  One wave per (replica, 64-slot
// i-block), split into NG = 64/G groups of G consecutive slots.  Every group streams its
// own j atoms -- those within rc of the GROUP's box rather than the block's -- so far
// fewer of the evaluated lane pairs lie outside the cutoff (a 16-atom group's box is
// about half as wide as a 64-atom block's).
//
// Candidate j-blocks (j-block >= i-block) are found by the wave from the block boxes,
// without a tile list; those that need no periodic shift come first, the rest after
// them.  For each j-block the 64 lanes test its 64 atoms against every group's box and
// append the slots that pass to that group's ring buffer in LDS (entry = slot | flags).
// Full chunks of 8 entries per group are consumed as soon as every group has one; at the
// end, and when a ring would overflow, partial chunks are consumed with per-lane guards.
//
// Within a chunk lane L of a group evaluates ring entry k ^ rot(L) in its register slot
// k.  With that order the j forces are summed over the group's lanes by halving DPP
// stages that never select: at each stage a lane keeps slots [0, n/2) and adds its
// partner's slots [n/2, n), which hold the same j atoms.  The lane then owns the force
// on the j atom of its register slot 0 and adds it with int64 atomics.
// ---------------------------------------------------------------------------
constexpr unsigned RING_SLOT = 0x1FFFFFFFu;
constexpr unsigned RING_EXCL = 1u << 29;    // pair needs j > i and the exclusion test
constexpr unsigned RING_MINIMG = 1u << 30;  // pair needs the per-pair minimum image

// Ring order of lane gl (0..G-1) of a group.  G = 16: stage partners lane^8 (bit 3),
// lane^7 (bits 2,1,0), lane^2 (bit 1), then lane^1 sums equal atoms.  G = 8: lane^7,
// lane^2, lane^1.  rot flips exactly the bit the stage halves on for each partner.
template <int G>
__device__ __forceinline__ int group_rot(int gl) {
    if (G == 16) {
        int b3 = (gl >> 3) & 1, b2 = (gl >> 2) & 1, b1 = (gl >> 1) & 1;
        return 4 * b3 + 2 * b2 + (b1 ^ b2);
    }
    int b2 = (gl >> 2) & 1, b1 = (gl >> 1) & 1, b0 = gl & 1;
    return 4 * b2 + 2 * (b1 ^ b2) + (b0 ^ b2);
}

// v[3k + c]: component c of the force on ring entry k ^ rot(L).  Returns in out[] the
// group's sum for ring entry rot(L) (register slot 0).
template <int G>
__device__ __forceinline__ void group_reduce24(const float (&v)[24], float (&out)[3]) {
    float w[12], x[6];
    if (G == 16) {
#pragma unroll
        for (int t = 0; t < 12; t++) w[t] = dpp_add<0x128>(v[t], v[t + 12]);  // row_ror:8
#pragma unroll
        for (int t = 0; t < 6; t++) x[t] = dpp_add<0x141>(w[t], w[t + 6]);    // half mirror
#pragma unroll
        for (int t = 0; t < 3; t++) {
            float y = dpp_add<0x4E>(x[t], x[t + 3]);                          // lane ^ 2
            out[t] = dpp_add<0xB1>(y, y);                                     // lane ^ 1
        }
    } else {
#pragma unroll
        for (int t = 0; t < 12; t++) w[t] = dpp_add<0x141>(v[t], v[t + 12]);
#pragma unroll
        for (int t = 0; t < 6; t++) x[t] = dpp_add<0x4E>(w[t], w[t + 6]);
#pragma unroll
        for (int t = 0; t < 3; t++) out[t] = dpp_add<0xB1>(x[t], x[t + 3]);
    }
}

// min / max over the G lanes of a group (same butterfly as the force sum)
template <int G>
__device__ __forceinline__ float group_min(float v) {
    if (G == 16) v = fminf(v, dpp_f<0x128>(v));
    v = fminf(v, dpp_f<0x141>(v));
    v = fminf(v, dpp_f<0x4E>(v));
    return fminf(v, dpp_f<0xB1>(v));
}
template <int G>
__device__ __forceinline__ float group_max(float v) {
    if (G == 16) v = fmaxf(v, dpp_f<0x128>(v));
    v = fmaxf(v, dpp_f<0x141>(v));
    v = fmaxf(v, dpp_f<0x4E>(v));
    return fmaxf(v, dpp_f<0xB1>(v));
}



// One chunk: 8 ring entries per group.  SRC 0 / 1: every entry is valid and the records
// were copied to LDS staging buffer `stage` (record of group q, entry k at
// stage[2 (8q + k)]); SRC 2: the records are loaded from A and bit k of `valid` says
// whether register slot k holds an entry.  mi (per-pair minimum image) and ex (entries
// that need the j > i / exclusion test) are wave-uniform.  The pairs run in the order
// 0,4,2,6,1,5,3,7 so that the first two halving stages of the j-force sum fold in as
// soon as both of their slots are done (12 live partial sums instead of 24).  The j
// force of register slot 0's entry is returned in f.
template <int G, int SRC>
__device__ __forceinline__ void grp_chunk(const float4* __restrict__ A,
                                          const float4* __restrict__ stage, const unsigned (&e)[8],
                                          unsigned valid, bool mi, bool ex, int gi, int rot,
                                          float4 pi, float2 qi, int islot, bool vi,
                                          const int* __restrict__ exs, int nex, const Geom& g,
                                          PairAcc& a, float (&f)[3]) {
    constexpr bool STAGED = SRC != 2;
    constexpr int ord[8] = {0, 4, 2, 6, 1, 5, 3, 7};
    constexpr int S1 = G == 16 ? 0x128 : 0x141;  // row_ror:8 | row_half_mirror
    constexpr int S2 = G == 16 ? 0x141 : 0x4E;   // row_half_mirror | lane ^ 2
    constexpr int S3 = G == 16 ? 0x4E : 0xB1;    // lane ^ 2 | lane ^ 1
    float4 pj[8];
    float2 qj[8];
    auto ld = [&](int k) {
        if (STAGED) {
            int ix = 2 * (gi * 8 + (k ^ rot));
            pj[k] = stage[ix];
            qj[k] = *(const float2*)&stage[ix + 1];
        } else {
            // e << 5 = the slot's byte offset (32-byte records; the flag bits shift out)
            const char* rec = (const char*)A + (e[k] << 5);
            pj[k] = *(const float4*)rec;
            qj[k] = *(const float2*)(rec + 16);
        }
    };
    auto pair = [&](int k, float (&v)[3]) {
        bool incl = vi;
        if (!STAGED) incl = incl && ((valid >> k) & 1u);
        if (ex && (e[k] & RING_EXCL)) {
            int js = (int)(e[k] & RING_SLOT);
            bool part = js <= islot;
            for (int x = 0; x < nex; x++) part |= exs[x * 64] == js;
            incl = incl && !part;
        }
        float dx, dy, dz;
        float fr = pair_masked(pi, qi, pj[k], qj[k], incl, mi, g, a, dx, dy, dz);
        v[0] = -fr * dx;
        v[1] = -fr * dy;
        v[2] = -fr * dz;
    };
    // staged records are read two pairs ahead; global ones (rare partial chunks) just
    // before their pair, which keeps this path's registers within the staged one's
    constexpr int AHEAD = STAGED ? 2 : 0;
#pragma unroll
    for (int t = 0; t < AHEAD; t++) ld(ord[t]);
    float w[4][3], x[2][3];
#pragma unroll
    for (int t = 0; t < 8; t += 2) {
        float v0[3], v1[3];
        if (t + AHEAD < 8) ld(ord[t + AHEAD]);
        __builtin_amdgcn_sched_barrier(0);
        pair(ord[t], v0);
        if (t + 1 + AHEAD < 8) ld(ord[t + 1 + AHEAD]);
        __builtin_amdgcn_sched_barrier(0);
        pair(ord[t + 1], v1);
        const int sl = ord[t];  // slot sl pairs with slot sl + 4: 0, 2, 1, 3
#pragma unroll
        for (int c = 0; c < 3; c++) w[sl][c] = dpp_add<S1>(v0[c], v1[c]);
        if (sl == 2)
#pragma unroll
            for (int c = 0; c < 3; c++) x[0][c] = dpp_add<S2>(w[0][c], w[2][c]);
        if (sl == 3)
#pragma unroll
            for (int c = 0; c < 3; c++) x[1][c] = dpp_add<S2>(w[1][c], w[3][c]);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float y = dpp_add<S3>(x[0][c], x[1][c]);
        f[c] = G == 16 ? dpp_add<0xB1>(y, y) : y;
    }
    a.e += (double)a.ef;  // FP32 over 8 pairs, FP64 beyond
    a.ef = 0.f;
}


template <int G, int CAP, int WPW, bool COUNT>
__global__ __launch_bounds__(64 * WPW) void k_pair_grp(
    int R, int N, int Np, int NB, const float4* __restrict__ arec,
    const int* __restrict__ orig_s, const int* __restrict__ slot_of,
    const float4* __restrict__ bb_lo, const float4* __restrict__ bb_hi,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list, int max_ex,
    long long* __restrict__ F, double* __restrict__ e_pair, int* __restrict__ flags,
    unsigned long long* __restrict__ pair_count, float list_r2, Geom g) {
    constexpr int NG = 64 / G;
    extern __shared__ unsigned sdm_lds[];
    // two staging buffers as separate LDS objects: the compiler then knows that reading
    // one does not have to wait for the copy into the other
    __shared__ float4 stage_a[WPW][NG * 16], stage_b[WPW][NG * 16];
    int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int w = xcd_linear_id(blockIdx.x, gridDim.x) * WPW + wv;
    int lane = threadIdx.x & 63;
    if (w >= R * NB) return;
    int r = w / NB;
    int ib = w - r * NB;
    // per wave: rings [NG][CAP] | exclusion partners [max_ex][64] | group boxes [NG][2]
    unsigned* ring = sdm_lds + wv * (NG * CAP + max_ex * 64 + NG * 8);
    int* exs = (int*)(ring + NG * CAP);
    const float4* A = arec + (size_t)r * Np * 2;
    unsigned long long* Fr = (unsigned long long*)(F + (size_t)r * 3 * Np);
    const int gi = lane / G, gl = lane % G;
    const int rot = group_rot<G>(gl);
    int islot = ib * 64 + lane;
    bool vi = islot < N;
    float4 pi = A[2 * islot];
    float4 qi4 = A[2 * islot + 1];
    float2 qi = make_float2(qi4.x, qi4.y);
    int nex = 0;
    if (vi) {
        int ai = orig_s[(size_t)r * Np + islot];
        int b0 = ex_start[ai];
        nex = ex_start[ai + 1] - b0;
        for (int k = 0; k < nex; k++) exs[k * 64 + lane] = slot_of[(size_t)r * Np + ex_list[b0 + k]];
    }
    // group boxes: every lane knows its own group's, the wave holds all of them uniform
    float glx = group_min<G>(vi ? pi.x : 3.0e38f), gly = group_min<G>(vi ? pi.y : 3.0e38f),
          glz = group_min<G>(vi ? pi.z : 3.0e38f);
    float ghx = group_max<G>(vi ? pi.x : -3.0e38f), ghy = group_max<G>(vi ? pi.y : -3.0e38f),
          ghz = group_max<G>(vi ? pi.z : -3.0e38f);
    float gcx = 0.5f * (glx + ghx), gcy = 0.5f * (gly + ghy), gcz = 0.5f * (glz + ghz);
    float gex = 0.5f * (ghx - glx), gey = 0.5f * (ghy - gly), gez = 0.5f * (ghz - glz);
    // every group's box in LDS (center, half extent), read back uniformly when appending
    float4* gbox = (float4*)(exs + max_ex * 64);
    if (gl == 0) {
        gbox[2 * gi] = make_float4(gcx, gcy, gcz, 0.f);
        gbox[2 * gi + 1] = make_float4(gex, gey, gez, 0.f);
    }
    __builtin_amdgcn_wave_barrier();
    // groups with at least one atom (slots are filled in order, so the first lane decides)
    int ngv = min(NG, (N - ib * 64 + G - 1) / G);
    float4 bli = bb_lo[(size_t)r * NB + ib], bhi = bb_hi[(size_t)r * NB + ib];
    // ring head, common to every group and a multiple of 8: a chunk takes 8 entries from
    // every ring, and a ring with fewer has its tail moved up to the new head
    int h = 0;
    int tail[NG];
#pragma unroll
    for (int q = 0; q < NG; q++) tail[q] = 0;
    double ax = 0.0, ay = 0.0, az = 0.0, aqr2 = 0.0, aq = 0.0;
    PairAcc acc;
    acc.fx = acc.fy = acc.fz = acc.sqr2 = acc.sq = 0.f;
    acc.e = 0.0;
    acc.ef = 0.f;
    acc.cnt = 0;
    int bad = 0;
    int nchunk = 0;
    const unsigned rot4 = (unsigned)rot * 4u;
    const bool own_j = G == 8 || (gl & 1) == 0;  // lane that adds its j atom's force
    // staging: lane L copies half L&1 of record (L>>1) + 32 i of the chunk (DMA i)
    int dkoff[NG / 4];
#pragma unroll
    for (int i = 0; i < NG / 4; i++) {
        int ri = i * 32 + (lane >> 1);
        dkoff[i] = (ri >> 3) * CAP + (ri & 7);
    }
    int cur = 0;       // staging buffer of the next full chunk
    bool pre = false;  // its records are already there
    // copies the records of ring entries hd..hd+7 of every group into staging buffer b
    // (global -> LDS without registers; completion is waited for with vmcnt)
    // copies the records of ring entries hd..hd+7 of every group into staging buffer B
    // (global -> LDS without registers; completion is waited for with vmcnt)
    auto stage_chunk = [&](int hd, auto bufc) {
        constexpr int B = decltype(bufc)::value;
        float4* dst = B ? stage_b[wv] : stage_a[wv];
        int hb = hd & (CAP - 1);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int i = 0; i < NG / 4; i++) {
            unsigned ee = ring[dkoff[i] + hb];
            const char* src = (const char*)A + ((ee << 5) | ((unsigned)(lane & 1) << 4));
            __builtin_amdgcn_global_load_lds(
                (const void*)src, (__attribute__((address_space(3))) void*)(dst + i * 64), 16, 0, 0);
        }
        __builtin_amdgcn_wave_barrier();
    };
    auto add_j = [&](unsigned e0, const float (&f)[3]) {
        int s = (int)(e0 & RING_SLOT);
        atomicAdd(&Fr[s], fixed_from_float(f[0], &bad));
        atomicAdd(&Fr[Np + s], fixed_from_float(f[1], &bad));
        atomicAdd(&Fr[2 * Np + s], fixed_from_float(f[2], &bad));
    };
    auto after_chunk = [&]() {
        h += 8;
#pragma unroll
        for (int q = 0; q < NG; q++) tail[q] = max(tail[q], h);
        if ((++nchunk & 3) == 0) {  // FP32 partial sums of 32 pairs go to FP64
            ax += (double)acc.fx;
            ay += (double)acc.fy;
            az += (double)acc.fz;
            aqr2 += (double)acc.sqr2;
            aq += (double)acc.sq;
            acc.fx = acc.fy = acc.fz = acc.sqr2 = acc.sq = 0.f;
        }
    };
    // full chunk from staging buffer B; stages the next one into the other buffer first
    // when every ring holds it (next_ok), so its copy overlaps this chunk's arithmetic.
    // B is a compile-time constant so every LDS access names its buffer.
    auto consume_full = [&](bool next_ok, auto bufc) {
        constexpr int B = decltype(bufc)::value;
        if (!pre) {
            stage_chunk(h, std::integral_constant<int, B>());
            wait_vmem();
        }
        pre = false;
        if (next_ok) {
            stage_chunk(h + 8, std::integral_constant<int, 1 - B>());
            pre = true;
        }
        unsigned base4 = (unsigned)(gi * CAP + (h & (CAP - 1))) * 4u;
        unsigned e[8];
        unsigned fl = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            e[k] = *(const unsigned*)((const char*)ring + (((unsigned)k * 4u) ^ rot4) + base4);
            fl |= e[k];
        }
        bool mi = __ballot((fl & RING_MINIMG) != 0u) != 0ull;
        bool ex = __ballot((fl & RING_EXCL) != 0u) != 0ull;
        float f[3];
        grp_chunk<G, B>(A, B ? stage_b[wv] : stage_a[wv], e, 0xFFu, mi, ex, gi, rot, pi, qi,
                        islot, vi, exs + lane, nex, g, acc, f);
        wait_vmem();  // the next chunk's records have landed; older atomics are done
        if (own_j) add_j(e[0], f);
        cur ^= 1;
        after_chunk();
    };
    // partial chunk (fewer than 8 entries in some ring): records loaded directly
    auto consume_partial = [&]() {
        pre = false;
        int t = tail[0];
#pragma unroll
        for (int q = 1; q < NG; q++) t = (gi == q) ? tail[q] : t;
        int n = t - h;
        unsigned base4 = (unsigned)(gi * CAP + (h & (CAP - 1))) * 4u;
        unsigned e[8];
        unsigned valid = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            bool ok = (k ^ rot) < n;
            unsigned v = *(const unsigned*)((const char*)ring + (((unsigned)k * 4u) ^ rot4) + base4);
            e[k] = ok ? v : (unsigned)islot;  // unused slot: own record
            valid |= ok ? (1u << k) : 0u;
        }
        float f[3];
        grp_chunk<G, 2>(A, nullptr, e, valid, true, true, gi, rot, pi, qi, islot, vi, exs + lane, nex, g, acc, f);
        if (own_j && (valid & 1u)) add_j(e[0], f);
        after_chunk();
    };
    auto min_avail = [&]() {
        int m = tail[0];
#pragma unroll
        for (int q = 1; q < NG; q++) m = min(m, tail[q]);
        return m - h;
    };
    auto max_avail = [&]() {
        int m = tail[0];
#pragma unroll
        for (int q = 1; q < NG; q++) m = max(m, tail[q]);
        return m - h;
    };

    // candidate j-blocks of one pass among blocks base..base+63 (pass 0: the plain
    // difference is the minimum image for every pair of the two boxes; pass 1: the rest)
    auto candidates = [&](int pass, int base) {
        int jb = base + lane;
        bool cand = false;
        if (jb < NB && jb >= ib) {
            float4 blj = bb_lo[(size_t)r * NB + jb], bhj = bb_hi[(size_t)r * NB + jb];
            float dcx = 0.5f * (bli.x + bhi.x - blj.x - bhj.x);
            float dcy = 0.5f * (bli.y + bhi.y - blj.y - bhj.y);
            float dcz = 0.5f * (bli.z + bhi.z - blj.z - bhj.z);
            bool plain = fabsf(dcx) + 0.5f * (bhi.x - bli.x + bhj.x - blj.x) < g.oix &&
                         fabsf(dcy) + 0.5f * (bhi.y - bli.y + bhj.y - blj.y) < g.oiy &&
                         fabsf(dcz) + 0.5f * (bhi.z - bli.z + bhj.z - blj.z) < g.oiz;
            cand = (plain ? 0 : 1) == pass && bbox_gap2(bli, bhi, blj, bhj, g) < list_r2;
        }
        return __ballot(cand);
    };
    // appends the atoms of j-block jb that lie within rc of each group's box
    auto append = [&](int jb, int pass) {
        __builtin_amdgcn_wave_barrier();
        int s = jb * 64 + lane;
        bool vj = s < N;
        float4 pj = A[2 * min(s, N - 1)];
        bool mine = jb == ib;
        for (int k = 0; k < nex; k++) mine |= (exs[k * 64 + lane] >> 6) == jb;
        unsigned fl = (__ballot(mine) != 0ull ? RING_EXCL : 0u) | (pass ? RING_MINIMG : 0u);
#pragma unroll
        for (int q = 0; q < NG; q++) {
            float4 c = gbox[2 * q], h = gbox[2 * q + 1];
            float dx = pj.x - c.x, dy = pj.y - c.y, dz = pj.z - c.z;
            if (pass) {
                dx -= g.bx * rintf(dx * g.ibx);
                dy -= g.by * rintf(dy * g.iby);
                dz -= g.bz * rintf(dz * g.ibz);
            }
            float ux = fmaxf(fabsf(dx) - h.x, 0.f), uy = fmaxf(fabsf(dy) - h.y, 0.f),
                  uz = fmaxf(fabsf(dz) - h.z, 0.f);
            bool in = vj && q < ngv && ux * ux + uy * uy + uz * uz < list_r2;
            if (jb == ib) in = in && s > ib * 64 + q * G;  // j must follow an i of the group
            unsigned long long mq = __ballot(in);
            if (in) ring[q * CAP + ((tail[q] + lane_rank(mq)) & (CAP - 1))] = (unsigned)s | fl;
            tail[q] += __popcll(mq);
        }
        __builtin_amdgcn_wave_barrier();
    };

    // Fill first (every ring 16 deep, so a full chunk can stage the next one), then full
    // chunks; partial chunks only when a ring would overflow or no tile is left.
    int pass = 0, base = ib & ~63;
    unsigned long long tm = candidates(0, base);
    bool tiles_left = true;
    // Everything loaded so far (this lane's atom, the block boxes) has arrived: without
    // this wait the compiler would place one in the loop at the first use of each such
    // value, and it would then also wait for the staging copy in flight every chunk.
    wait_vmem();
    while (true) {
        int mn = min_avail(), mx = max_avail();
        if (tiles_left && mn < 16 && mx <= CAP - 64) {
            while (tm == 0ull && tiles_left) {
                base += 64;
                if (base >= NB) {
                    if (pass == 1) {
                        tiles_left = false;
                        break;
                    }
                    pass = 1;
                    base = ib & ~63;
                }
                tm = candidates(pass, base);
            }
            if (!tiles_left) continue;
            int jb = base + __ffsll((long long)tm) - 1;
            tm &= tm - 1ull;
            append(jb, pass);
            continue;
        }
        if (mn >= 8) {
            if (cur == 0) consume_full(mn >= 16, std::integral_constant<int, 0>());
            else consume_full(mn >= 16, std::integral_constant<int, 1>());
            continue;
        }
        if (mx > 0) {
            consume_partial();
            continue;
        }
        break;
    }
    ax += (double)acc.fx;
    ay += (double)acc.fy;
    az += (double)acc.fz;
    aqr2 += (double)acc.sqr2;
    aq += (double)acc.sq;
    if (COUNT) {
        double c = wave_sum(vi ? (double)acc.cnt : 0.0);
        if (lane == 0) atomicAdd(&pair_count[r], (unsigned long long)(c + 0.5));
        return;
    }
    if (vi) {
        long long fx = to_fixed(ax, &bad), fy = to_fixed(ay, &bad), fz = to_fixed(az, &bad);
        atomicAdd(&Fr[islot], (unsigned long long)fx);
        atomicAdd(&Fr[Np + islot], (unsigned long long)fy);
        atomicAdd(&Fr[2 * Np + islot], (unsigned long long)fz);
    }
    double es = wave_sum(vi ? (acc.e + g.krf * aqr2 - g.crf * aq) : 0.0);
    if (__ballot(bad) != 0ull && lane == 0) atomicOr(&flags[r], 2);
    if (lane == 0) e_pair[(size_t)r * NB + ib] = es;
}

// ---------------------------------------------------------------------------
// Pair energies in FP64 (the energy evaluations: U0 and with it U1 = U0 + u and the
// potential energy).  The force kernels compute in FP32 from FP32 coordinates, whose
// rounding (an ulp is ~5e-7 nm at 4-8 nm) and arithmetic leave ~2^-24 of the summed pair
// magnitudes in U0; where U0 is a small net of large terms that exceeds 1e-5 of U0.  This
// kernel takes the FP64 positions and parameters of every pair with r < rc once, in
// Newton-3 order over the half tile lists, so U0 is exact to FP64 rounding.
//
// One wave per (replica, i-block, part): part p of `split` takes tiles p, p + split, ..
// Per tile the wave screens the mask's j atoms against its 64 i atoms in FP32 (the FP32
// slot records, cutoff widened by 1e-4 nm so no pair inside rc is dropped) and queues
// the surviving (i, j) lane pairs in LDS; every 64 queued pairs are evaluated one per
// lane in FP64 (minimum image, exact cutoff test, LJ + reaction field), so the FP64 work
// is spent on the ~25% of screened pairs that lie inside the cutoff.  half = 0: full tile
// lists (the non-Newton kernel's), of which only tiles with j-block >= i-block are taken.
// ---------------------------------------------------------------------------
constexpr int E64_SPLIT = 4;  // waves (parts) per i-block of the FP64 energy kernel

// Orders the LDS accesses of a one-wave workgroup around this point.  A wave's LDS
// instructions are carried out in issue order, so a compiler barrier is enough: unlike
// __syncthreads() it does not wait for the wave's global loads in flight.
__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

__global__ __launch_bounds__(64) void k_energy64(
    int R, int N, int Np, int NB, int maxj, int split, int half, const float4* __restrict__ arec,
    const Rec64* __restrict__ rec, const int* __restrict__ orig_s,
    const int* __restrict__ slot_of, const int* __restrict__ tiles,
    const unsigned long long* __restrict__ tmask, const int* __restrict__ ntiles,
    const float4* __restrict__ bb_lo, const float4* __restrict__ bb_hi,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list, int max_ex,
    double* __restrict__ e_out, Geom g) {
    __shared__ Rec64 irec[64], jrec[64];
    __shared__ float4 jpos[64];
    __shared__ unsigned short queue[128];
    extern __shared__ int exs64[];  // [max_ex][64] exclusion partners' slots
    const int w = xcd_linear_id(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x;
    if (w >= R * NB * split) return;
    const int part = w % split, rb = w / split;
    const int r = rb / NB, ib = rb - r * NB;
    const float4* A = arec + (size_t)r * Np * 2;
    const Rec64* RC = rec + (size_t)r * Np;
    const int islot = ib * 64 + lane;
    const bool vi = islot < N;
    const bool full_block = ib * 64 + 63 < N;  // uniform: every lane holds an atom
    // this lane's i atom: FP32 record for the screen, FP64 record for the pairs
    const float4 pi = A[2 * islot];
    irec[lane] = RC[vi ? islot : ib * 64];
    int nex = 0;
    int exlo = NB, exhi = -1;  // j-blocks holding exclusion partners (as in k_pair_n3)
    if (vi) {
        int ai = orig_s[(size_t)r * Np + islot];
        int b0 = ex_start[ai];
        nex = ex_start[ai + 1] - b0;
        for (int k = 0; k < nex; k++) {
            const int sl = slot_of[(size_t)r * Np + ex_list[b0 + k]];
            exs64[k * 64 + lane] = sl;
            exlo = min(exlo, sl >> 6);
            exhi = max(exhi, sl >> 6);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        exlo = min(exlo, __shfl_xor(exlo, o, 64));
        exhi = max(exhi, __shfl_xor(exhi, o, 64));
    }
    exlo = __builtin_amdgcn_readfirstlane(exlo);
    exhi = __builtin_amdgcn_readfirstlane(exhi);
    const float rcs = sqrtf(g.rc2f) + 1e-4f;
    const float rc2s = rcs * rcs;
    const float4 bli = bb_lo[(size_t)r * NB + ib], bhi = bb_hi[(size_t)r * NB + ib];
    const int* tl = tiles + ((size_t)r * NB + ib) * maxj;
    const unsigned long long* tmk = tmask + ((size_t)r * NB + ib) * maxj;
    const int nt = ntiles[(size_t)r * NB + ib];
    double e = 0.0;
    unsigned head = 0, tail = 0;  // queue positions (wave-uniform)
    // the queued pairs head .. head + n - 1 (n <= 64), one per lane.  one_image: every
    // pair of the tile takes the shift (sx, sy, sz) (records are in the FP32 records'
    // images); otherwise the minimum image per pair
    double sx = 0.0, sy = 0.0, sz = 0.0;
    bool one_image = true;
    auto evaluate = [&](unsigned n) {
        if ((unsigned)lane < n) {
            unsigned q = queue[(head + lane) & 127u];
            const Rec64 I = irec[q & 63u], J = jrec[q >> 6];
            double dx = (I.x - sx) - J.x, dy = (I.y - sy) - J.y, dz = (I.z - sz) - J.z;
            if (!one_image) {
                dx = min_image(dx, g.Lx, g.iLx);
                dy = min_image(dy, g.Ly, g.iLy);
                dz = min_image(dz, g.Lz, g.iLz);
            }
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < g.rc2) {
                double ir = rsqrt(r2);
                double ir2 = ir * ir;
                double s2;
                if (g.geo) {  // Geom::geo: the records hold sigma itself
                    s2 = I.sh * J.sh * ir2;
                } else {
                    double sig = I.sh + J.sh;
                    s2 = sig * sig * ir2;
                }
                double s6 = s2 * s2 * s2;
                double qq = I.qk * J.qk;
                e += (I.es * J.es) * s6 * (s6 - 1.0) + qq * (ir + g.krf * r2 - g.crf);
            }
        }
        head += n;
    };
    // screen of one j atom against the wave's i atoms, queueing the pairs inside the
    // widened cutoff.  MI: per-pair minimum image (otherwise the tile's shift is in ps);
    // CHK: some lane is padding or has an exclusion partner in the tile
    auto screen = [&](int j, float4 pj, float4 ps, unsigned long long em, auto mic, auto chk) {
        constexpr bool MI = decltype(mic)::value, CHK = decltype(chk)::value;
        float dx = ps.x - pj.x, dy = ps.y - pj.y, dz = ps.z - pj.z;
        if (MI) {
            dx -= g.bx * rintf(dx * g.ibx);
            dy -= g.by * rintf(dy * g.iby);
            dz -= g.bz * rintf(dz * g.ibz);
        }
        bool in = dx * dx + dy * dy + dz * dz < rc2s;
        if (CHK) in = in && vi && !((em >> j) & 1ull);
        const unsigned long long m = __ballot(in);
        if (in) queue[(tail + (unsigned)lane_rank(m)) & 127u] = (unsigned short)(lane | (j << 6));
        tail += (unsigned)__popcll(m);
        if (tail - head >= 64u) {
            lds_order();  // the queue entries are written
            evaluate(64u);
            lds_order();  // and read before they are overwritten
        }
    };
    // the j records of the wave's next tile are loaded into registers while the current
    // tile is screened and evaluated, and stored to LDS when its turn comes
    // (plain vector values, not a struct, so that they stay in registers).  This lane's
    // i record is used here first, so that the loop's waits for it are settled before
    // the first prefetch is in flight (a wait inside the loop would wait for that too)
    asm volatile("" ::"v"(pi.x), "v"(pi.y), "v"(pi.z), "v"(pi.w));
    double2 na = make_double2(0.0, 0.0), nb = na, nc = na;  // a record is 3 x 16 bytes
    float4 npos = make_float4(0.f, 0.f, 0.f, 0.f);
    if (part < nt) {
        const int js = tl[part] * 64 + lane;
        if (js < N) {
            const double2* src = (const double2*)(RC + js);
            na = src[0];
            nb = src[1];
            nc = src[2];
            npos = A[2 * js];
        }
    }
    for (int t = part; t < nt; t += split) {
        const int jb = tl[t];
        unsigned long long am = tmk[t];
        lds_order();  // the previous tile's records are no longer read
        if (jb * 64 + lane < N) {
            jrec[lane] = Rec64{na.x, na.y, nb.x, nb.y, nc.x, nc.y};
            jpos[lane] = npos;
        }
        if (t + split < nt) {
            const int js = tl[t + split] * 64 + lane;
            if (js < N) {
                const double2* src = (const double2*)(RC + js);
                na = src[0];
                nb = src[1];
                nc = src[2];
                npos = A[2 * js];
            }
        }
        if (!half && jb < ib) continue;  // full lists: each block pair once
        unsigned long long em = 0;
        if (jb >= exlo && jb <= exhi) {
            for (int k = 0; k < nex; k++) {
                int sl = exs64[k * 64 + lane];
                if ((sl >> 6) == jb) em |= 1ull << (sl & 63);
            }
        }
        if (jb == ib) em |= (2ull << lane) - 1ull;  // the diagonal tile: j > i only
        const bool chk = !full_block || __ballot(em != 0ull) != 0ull;
        // one periodic shift for the whole tile when it is exact for every pair within
        // the list radius (Geom::oix), as in the pair kernel
        const float4 blj = bb_lo[(size_t)r * NB + jb], bhj = bb_hi[(size_t)r * NB + jb];
        const float dcx = 0.5f * (bli.x + bhi.x - blj.x - bhj.x);
        const float dcy = 0.5f * (bli.y + bhi.y - blj.y - bhj.y);
        const float dcz = 0.5f * (bli.z + bhi.z - blj.z - bhj.z);
        const float kx = rintf(dcx * g.ibx), ky = rintf(dcy * g.iby), kz = rintf(dcz * g.ibz);
        const float fsx = g.bx * kx, fsy = g.by * ky, fsz = g.bz * kz;
        one_image = fabsf(dcx - fsx) + 0.5f * (bhi.x - bli.x + bhj.x - blj.x) < g.oix &&
                    fabsf(dcy - fsy) + 0.5f * (bhi.y - bli.y + bhj.y - blj.y) < g.oiy &&
                    fabsf(dcz - fsz) + 0.5f * (bhi.z - bli.z + bhj.z - blj.z) < g.oiz;
        float4 ps = pi;
        if (one_image) {
            ps.x -= fsx;
            ps.y -= fsy;
            ps.z -= fsz;
            sx = g.Lx * (double)kx;  // the same shift in FP64 (0 if not periodic)
            sy = g.Ly * (double)ky;
            sz = g.Lz * (double)kz;
        } else {
            sx = sy = sz = 0.0;
        }
        lds_order();  // this tile's records are stored
        // the mask's j atoms four at a time: their four position reads are in flight
        // together instead of one LDS round trip per atom
        auto screen4 = [&](auto mic, auto chk) {
            while (am) {
                int jq[4];
                float4 pq[4];
#pragma unroll
                for (int k = 0; k < 4; k++) jq[k] = next_bit(am);
#pragma unroll
                for (int k = 0; k < 4; k++) pq[k] = jpos[jq[k] & 63];
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (jq[k] >= 0) screen(jq[k], pq[k], ps, em, mic, chk);
            }
        };
        using T = std::true_type;
        using Fa = std::false_type;
        if (one_image) {
            if (chk) screen4(Fa(), T());
            else screen4(Fa(), Fa());
        } else {
            if (chk) screen4(T(), T());
            else screen4(T(), Fa());
        }
        lds_order();
        evaluate(tail - head);  // the tile's remaining pairs (its j records change next)
    }
    e = wave_sum(e);
    if (lane == 0) e_out[(size_t)rb * split + part] = e;
}

// ---------------------------------------------------------------------------
// group 1: harmonic bonds and angles at x (FP64), gathered per atom.  One thread per
// (atom, replica) evaluates every term its atom belongs to (a bond by both of its atoms,
// an angle by all three) and keeps only its own atom's share, so the forces need no
// atomics and are summed in a fixed order; a term's energy is counted by its first
// atom.  bcode holds (bond << 1) | role and acode (angle << 2) | role, role being the
// atom's position in the term; the lists of atom i are [bstart[i], bstart[i+1]) and
// [astart[i], astart[i+1]).  The forces go to Fb (FP64, atom order, [R][3][Np]), which
// the integrator adds to the fixed-point accumulator.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bonded(
    int N, int Np, const int* __restrict__ bstart, const int* __restrict__ bcode,
    const int* __restrict__ astart, const int* __restrict__ acode,
    const int* __restrict__ bond_atoms, const double* __restrict__ bond_r0,
    const double* __restrict__ bond_k, const int* __restrict__ ang_atoms,
    const double* __restrict__ ang_t0, const double* __restrict__ ang_k,
    const double* __restrict__ x64, double* __restrict__ Fb, double* __restrict__ e_bond,
    int nbw, int* __restrict__ flags) {
    __shared__ double red[256];
    int r = blockIdx.y;
    int i = blockIdx.x * 256 + threadIdx.x;
    const double* X = x64 + (size_t)r * 3 * Np;
    double e = 0.0;
    if (i < N) {
        double f[3] = {0.0, 0.0, 0.0};
        double xi[3] = {X[i], X[Np + i], X[2 * Np + i]};
        for (int q = bstart[i], qe = bstart[i + 1]; q < qe; q++) {
            int code = bcode[q], t = code >> 1, role = code & 1;
            int j = bond_atoms[2 * t + (role ^ 1)];
            double d[3];
            for (int a = 0; a < 3; a++) d[a] = X[a * Np + j] - xi[a];
            const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const double irr = rsq > 0.0 ? rsqrt(rsq) : 0.0;
            double dr = rsq * irr - bond_r0[t];
            double gg = bond_k[t] * dr * irr;
            for (int a = 0; a < 3; a++) f[a] += gg * d[a];  // pulls i toward j when stretched
            if (role == 0) e += 0.5 * bond_k[t] * dr * dr;
        }
        for (int q = astart[i], qe = astart[i + 1]; q < qe; q++) {
            int code = acode[q], k = code >> 2, role = code & 3;
            int ia = ang_atoms[3 * k], im = ang_atoms[3 * k + 1], ic = ang_atoms[3 * k + 2];
            double v1[3], v2[3];
            for (int a = 0; a < 3; a++) {
                double xm = X[a * Np + im];
                v1[a] = X[a * Np + ia] - xm;
                v2[a] = X[a * Np + ic] - xm;
            }
            // theta = atan2(|v1 x v2|, v1 . v2): one call gives the angle and, with the
            // reciprocal lengths, its sine (acos plus sin took twice the instructions and
            // is worse conditioned near 0 and pi)
            const double i1 = rsqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
            const double i2 = rsqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
            const double dot = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
            const double cx = v1[1] * v2[2] - v1[2] * v2[1], cy = v1[2] * v2[0] - v1[0] * v2[2],
                         cz = v1[0] * v2[1] - v1[1] * v2[0];
            const double cr = sqrt(cx * cx + cy * cy + cz * cz);
            double th = atan2(cr, dot);
            const double cs = dot * i1 * i2, sn = cr * i1 * i2;
            double dth = th - ang_t0[k];
            double gg = ang_k[k] * dth / fmax(sn, 1e-12);
            if (role == 0) e += 0.5 * ang_k[k] * dth * dth;
            const double p12 = gg * i1 * i2, p11 = gg * cs * i1 * i1, p22 = gg * cs * i2 * i2;
            for (int a = 0; a < 3; a++) {
                double fa = p12 * v2[a] - p11 * v1[a];
                double fc = p12 * v1[a] - p22 * v2[a];
                f[a] += role == 0 ? fa : (role == 2 ? fc : -(fa + fc));
            }
        }
        bool bad = false;
        double* Fr = Fb + (size_t)r * 3 * Np;
        for (int a = 0; a < 3; a++) {
            bad |= !(fabs(f[a]) < SDM_FIX_LIMIT);
            Fr[a * Np + i] = f[a];
        }
        if (bad) atomicOr(&flags[r], 2);
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) e_bond[(size_t)r * nbw + blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// group 1: the SDMUtils restraints at x (FP64), fixed-point atomics; one wave per
// (restraint, replica).  Restates SDMUtils.py:32-257 (the energy expressions of the
// CustomCentroidBondForce and the three CustomCompoundBondForces); no periodic images.
// ---------------------------------------------------------------------------
struct DevCentroid {
    int g1, n1, g2, n2;  // ranges of the restraint atom / weight arrays (weights sum to 1)
    int do_angles, pad;
    int ref[6];          // g3, g4, g5 (receptor), g6, g7, g8 (ligand)
    double kfcm, tol, off[3], control;
    double kf[3], lo[3], hi[3];
};

struct DevAlign {
    int a[3], b[3];
    double kfd, kth, kps, off[3];
};

__device__ __forceinline__ double wrap_period(double x, double P) {
    return x - P * floor(x / P + 0.5);
}

__device__ __forceinline__ double dot3(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// (kf/2)(step(dm) max(0,db)^2 + step(-dm) max(0,-da)^2) with db, da, dm = t - hi, t - lo,
// t - (lo+hi)/2 wrapped to the period P (step(0) = 1); *dEdt = its derivative in t
__device__ double flat_bottom(double t, double kf, double lo, double hi, double P,
                              double* dEdt) {
    double db = wrap_period(t - hi, P), da = wrap_period(t - lo, P);
    double dm = wrap_period(t - 0.5 * (lo + hi), P);
    double e = 0.0, g = 0.0;
    if (dm >= 0.0) {
        double m = fmax(0.0, db);
        e += m * m;
        g += m;
    }
    if (-dm >= 0.0) {
        double m = fmax(0.0, -da);
        e += m * m;
        g -= m;
    }
    *dEdt = kf * g;
    return 0.5 * kf * e;
}

// angle at p2 in [0, pi] and its gradient (zero when the three points are collinear)
__device__ double angle_grad(const double* p1, const double* p2, const double* p3, double* g1,
                             double* g2, double* g3) {
    double u[3], v[3], n[3], nu[3], nv[3];
    for (int d = 0; d < 3; d++) {
        u[d] = p1[d] - p2[d];
        v[d] = p3[d] - p2[d];
    }
    cross3(u, v, n);
    double nn = sqrt(dot3(n, n));
    double t = atan2(nn, dot3(u, v));
    cross3(n, u, nu);
    cross3(n, v, nv);
    double su = nn > 0.0 ? 1.0 / (nn * dot3(u, u)) : 0.0;
    double sv = nn > 0.0 ? 1.0 / (nn * dot3(v, v)) : 0.0;
    for (int d = 0; d < 3; d++) {
        g1[d] = -nu[d] * su;
        g3[d] = nv[d] * sv;
        g2[d] = -(g1[d] + g3[d]);
    }
    return t;
}

// dihedral (IUPAC sign) in (-pi, pi] and its gradient (zero when degenerate)
__device__ double dihedral_grad(const double* p1, const double* p2, const double* p3,
                                const double* p4, double* g1, double* g2, double* g3,
                                double* g4) {
    double b1[3], b2[3], b3[3], n1[3], n2[3];
    for (int d = 0; d < 3; d++) {
        b1[d] = p2[d] - p1[d];
        b2[d] = p3[d] - p2[d];
        b3[d] = p4[d] - p3[d];
    }
    cross3(b1, b2, n1);
    cross3(b2, b3, n2);
    double l2 = sqrt(dot3(b2, b2));
    double t = atan2(l2 * dot3(b1, n2), dot3(n1, n2));
    double q1 = dot3(n1, n1), q2 = dot3(n2, n2);
    bool ok = q1 > 0.0 && q2 > 0.0 && l2 > 0.0;
    double s1 = ok ? dot3(b1, b2) / (l2 * l2) : 0.0, s3 = ok ? dot3(b3, b2) / (l2 * l2) : 0.0;
    for (int d = 0; d < 3; d++) {
        g1[d] = ok ? -l2 * n1[d] / q1 : 0.0;
        g4[d] = ok ? l2 * n2[d] / q2 : 0.0;
        g2[d] = (-s1 - 1.0) * g1[d] + s3 * g4[d];
        g3[d] = s1 * g1[d] - (s3 + 1.0) * g4[d];
    }
    return t;
}

// k (1 - cos p) for cos p between the components of x3-x1 and x5-x4 perpendicular to
// x2-x1; adds the forces to f[0..4]
__device__ double psi_term(const double* x1, const double* x2, const double* x3,
                           const double* x4, const double* x5, double k, double (*f)[3]) {
    double d1[3], d0[3], d3[3], n1[3], v[3], w[3];
    for (int d = 0; d < 3; d++) {
        d1[d] = x2[d] - x1[d];
        d0[d] = x3[d] - x1[d];
        d3[d] = x5[d] - x4[d];
    }
    double l1 = sqrt(dot3(d1, d1));
    for (int d = 0; d < 3; d++) n1[d] = d1[d] / l1;
    double p0 = dot3(d0, n1), p3 = dot3(d3, n1);
    for (int d = 0; d < 3; d++) {
        v[d] = d0[d] - p0 * n1[d];
        w[d] = d3[d] - p3 * n1[d];
    }
    double lv = sqrt(dot3(v, v)), lw = sqrt(dot3(w, w));
    double c = dot3(v, w) / (lv * lw);
    for (int d = 0; d < 3; d++) {
        double gv = (w[d] / lw - c * v[d] / lv) / lv;
        double gw = (v[d] / lv - c * w[d] / lw) / lw;
        double gd1 = (-p0 * gv - p3 * gw) / l1;
        // F = -dE/dx = k d(cos p)/dx
        f[0][d] -= k * (gd1 + gv);
        f[1][d] += k * gd1;
        f[2][d] += k * gv;
        f[3][d] -= k * gw;
        f[4][d] += k * gw;
    }
    return k * (1.0 - c);
}

__global__ __launch_bounds__(64) void k_restraints(
    int Np, int ncen, int nrest, const DevCentroid* __restrict__ cen,
    const DevAlign* __restrict__ ali, const int* __restrict__ ratom,
    const double* __restrict__ rw, const double* __restrict__ x64,
    const int* __restrict__ slot_of, long long* __restrict__ F, double* __restrict__ e_rest,
    int* __restrict__ flags) {
    const double kPi = 3.14159265358979323846;
    int k = blockIdx.x, r = blockIdx.y, lane = threadIdx.x;
    const double* X = x64 + (size_t)r * 3 * Np;
    unsigned long long* Fr = (unsigned long long*)(F + (size_t)r * 3 * Np);
    const int* S = slot_of ? slot_of + (size_t)r * Np : nullptr;
    int bad = 0;
    auto addf = [&](int a, const double* f) {
        int s = S ? S[a] : a;
        atomicAdd(&Fr[s], (unsigned long long)to_fixed(f[0], &bad));
        atomicAdd(&Fr[Np + s], (unsigned long long)to_fixed(f[1], &bad));
        atomicAdd(&Fr[2 * Np + s], (unsigned long long)to_fixed(f[2], &bad));
    };
    auto pos = [&](int a, double* p) {
        p[0] = X[a];
        p[1] = X[Np + a];
        p[2] = X[2 * Np + a];
    };
    double e = 0.0;
    if (k < ncen) {
        const DevCentroid& R = cen[k];
        // weighted centres of the two groups (wave butterflies: every lane gets the sums)
        double c[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int i = lane; i < R.n1; i += 64) {
            int a = ratom[R.g1 + i];
            double w = rw[R.g1 + i];
            c[0] += w * X[a];
            c[1] += w * X[Np + a];
            c[2] += w * X[2 * Np + a];
        }
        for (int i = lane; i < R.n2; i += 64) {
            int a = ratom[R.g2 + i];
            double w = rw[R.g2 + i];
            c[3] += w * X[a];
            c[4] += w * X[Np + a];
            c[5] += w * X[2 * Np + a];
        }
        for (int d = 0; d < 6; d++) c[d] = wave_sum(c[d]);
        double dv[3] = {c[0] - R.off[0] - c[3], c[1] - R.off[1] - c[4],
                        c[2] - R.off[2] - c[5]};
        double d12 = sqrt(dot3(dv, dv));
        double gd = 0.0;
        if (d12 >= R.tol && d12 > 0.0) {
            double x = d12 - R.tol;
            e += 0.5 * R.control * R.kfcm * x * x;
            gd = R.control * R.kfcm * x / d12;  // |dE/dc1| / d12
        }
        if (gd != 0.0) {  // the same on every lane
            for (int i = lane; i < R.n1; i += 64) {
                double w = rw[R.g1 + i];
                double f[3] = {-w * gd * dv[0], -w * gd * dv[1], -w * gd * dv[2]};
                addf(ratom[R.g1 + i], f);
            }
            for (int i = lane; i < R.n2; i += 64) {
                double w = rw[R.g2 + i];
                double f[3] = {w * gd * dv[0], w * gd * dv[1], w * gd * dv[2]};
                addf(ratom[R.g2 + i], f);
            }
        }
        if (R.do_angles && lane == 0) {
            double p[6][3], f[6][3] = {}, g[4][3], dE;
            for (int j = 0; j < 6; j++) pos(R.ref[j], p[j]);
            // theta = angle(g3, g6, g7)
            double t = angle_grad(p[0], p[3], p[4], g[0], g[1], g[2]);
            e += R.control * flat_bottom(t, R.kf[0], R.lo[0], R.hi[0], kPi, &dE);
            dE *= R.control;
            const int ta[3] = {0, 3, 4};
            for (int j = 0; j < 3; j++)
                for (int d = 0; d < 3; d++) f[ta[j]][d] -= dE * g[j][d];
            // phi1 = dihedral(g4, g3, g6, g7), phi2 = dihedral(g3, g6, g7, g8)
            const int da[2][4] = {{1, 0, 3, 4}, {0, 3, 4, 5}};
            for (int q = 0; q < 2; q++) {
                const int* ix = da[q];
                t = dihedral_grad(p[ix[0]], p[ix[1]], p[ix[2]], p[ix[3]], g[0], g[1], g[2],
                                  g[3]);
                e += R.control *
                     flat_bottom(t, R.kf[1 + q], R.lo[1 + q], R.hi[1 + q], 2.0 * kPi, &dE);
                dE *= R.control;
                for (int j = 0; j < 4; j++)
                    for (int d = 0; d < 3; d++) f[ix[j]][d] -= dE * g[j][d];
            }
            for (int j = 0; j < 6; j++) addf(R.ref[j], f[j]);
        }
    } else if (lane == 0) {
        const DevAlign& A = ali[k - ncen];
        double a[3][3], b[3][3], fa[3][3] = {}, fb[3][3] = {};
        for (int j = 0; j < 3; j++) {
            pos(A.a[j], a[j]);
            pos(A.b[j], b[j]);
        }
        // displacement: (kfdispl/2)|b0 - offset - a0|^2
        double dd[3];
        for (int d = 0; d < 3; d++) dd[d] = b[0][d] - A.off[d] - a[0][d];
        e += 0.5 * A.kfd * dot3(dd, dd);
        for (int d = 0; d < 3; d++) {
            fb[0][d] -= A.kfd * dd[d];
            fa[0][d] += A.kfd * dd[d];
        }
        // theta: (ktheta/2)(1 - cos t), cos t = unit(b1 - b0).unit(a1 - a0)
        double d1[3], d2[3];
        for (int d = 0; d < 3; d++) {
            d1[d] = b[1][d] - b[0][d];
            d2[d] = a[1][d] - a[0][d];
        }
        double l1 = sqrt(dot3(d1, d1)), l2 = sqrt(dot3(d2, d2));
        double cs = dot3(d1, d2) / (l1 * l2);
        e += 0.5 * A.kth * (1.0 - cs);
        for (int d = 0; d < 3; d++) {
            double f2 = 0.5 * A.kth * (d2[d] / l2 - cs * d1[d] / l1) / l1;
            double f4 = 0.5 * A.kth * (d1[d] / l1 - cs * d2[d] / l2) / l2;
            fb[1][d] += f2;
            fb[0][d] -= f2;
            fa[1][d] += f4;
            fa[0][d] -= f4;
        }
        // psi, symmetrized: (kpsi/4)(1 - cos p) for (b0,b1,b2,a0,a2) and (a0,a1,a2,b0,b2)
        double f5[5][3] = {};
        e += psi_term(b[0], b[1], b[2], a[0], a[2], 0.25 * A.kps, f5);
        for (int d = 0; d < 3; d++) {
            fb[0][d] += f5[0][d];
            fb[1][d] += f5[1][d];
            fb[2][d] += f5[2][d];
            fa[0][d] += f5[3][d];
            fa[2][d] += f5[4][d];
        }
        double h5[5][3] = {};
        e += psi_term(a[0], a[1], a[2], b[0], b[2], 0.25 * A.kps, h5);
        for (int d = 0; d < 3; d++) {
            fa[0][d] += h5[0][d];
            fa[1][d] += h5[1][d];
            fa[2][d] += h5[2][d];
            fb[0][d] += h5[3][d];
            fb[2][d] += h5[4][d];
        }
        for (int j = 0; j < 3; j++) {
            addf(A.a[j], fa[j]);
            addf(A.b[j], fb[j]);
        }
    }
    if (bad) atomicOr(&flags[r], 2);
    if (lane == 0) e_rest[(size_t)r * nrest + k] = e;
}

// ---------------------------------------------------------------------------
// further force-field terms (FP64, one thread per term): periodic torsions, harmonic
// impropers and positional restraints (group 1), the 1-4 exceptions of the nonbonded
// force (group 2: plain LJ + Coulomb, no cutoff, no periodic image).
// ---------------------------------------------------------------------------
struct DevTerms {
    int nt, ni, nfb, nph, nex;  // term index ranges, in this order
    const int* t_atoms;
    const int* t_n;
    const double *t_phase, *t_k;
    const int* i_atoms;
    const double *i_phi0, *i_k;
    const int* fb_atom;
    const double *fb_x0, *fb_k, *fb_tol;
    const int* ph_atom;
    const double *ph_x0, *ph_k;
    const int* ex_atoms;
    const double *ex_qq, *ex_sig, *ex_eps;  // qq includes k_e
};

// group-1 forces go to Fb (slot map Sb, or atom order if null), group-2 forces to F
// (slot map S); energies per workgroup into e1 (group 1) and e2 (group 2)
__global__ __launch_bounds__(256) void k_terms(int Np, DevTerms T, const double* __restrict__ x64,
                                               const int* __restrict__ Sb_all,
                                               long long* __restrict__ Fb,
                                               const int* __restrict__ S_all,
                                               long long* __restrict__ F, double* __restrict__ e1,
                                               double* __restrict__ e2, int nblk, Geom g,
                                               int* __restrict__ flags) {
    __shared__ double red[2][256];
    int r = blockIdx.y, t = blockIdx.x * 256 + threadIdx.x;
    const double* X = x64 + (size_t)r * 3 * Np;
    const int* Sb = Sb_all ? Sb_all + (size_t)r * Np : nullptr;
    const int* S = S_all + (size_t)r * Np;
    unsigned long long* Fbr = (unsigned long long*)(Fb + (size_t)r * 3 * Np);
    unsigned long long* Fr = (unsigned long long*)(F + (size_t)r * 3 * Np);
    int bad = 0;
    double eg1 = 0.0, eg2 = 0.0;
    auto pos = [&](int a, double* p) {
        p[0] = X[a];
        p[1] = X[Np + a];
        p[2] = X[2 * Np + a];
    };
    auto addf = [&](unsigned long long* Fx, const int* map, int a, const double* f) {
        int s = map ? map[a] : a;
        atomicAdd(&Fx[s], (unsigned long long)to_fixed(f[0], &bad));
        atomicAdd(&Fx[Np + s], (unsigned long long)to_fixed(f[1], &bad));
        atomicAdd(&Fx[2 * Np + s], (unsigned long long)to_fixed(f[2], &bad));
    };
    const int e_imp = T.nt + T.ni, e_fb = e_imp + T.nfb, e_ph = e_fb + T.nph, e_ex = e_ph + T.nex;
    if (t < e_imp) {  // torsion or improper: a function of the dihedral
        bool tors = t < T.nt;
        int k = tors ? t : t - T.nt;
        const int* a = tors ? T.t_atoms + 4 * k : T.i_atoms + 4 * k;
        double p[4][3], gr[4][3];
        for (int j = 0; j < 4; j++) pos(a[j], p[j]);
        double phi = dihedral_grad(p[0], p[1], p[2], p[3], gr[0], gr[1], gr[2], gr[3]);
        double dE;
        if (tors) {
            double n = (double)T.t_n[k], arg = n * phi - T.t_phase[k];
            eg1 = T.t_k[k] * (1.0 + cos(arg));
            dE = -T.t_k[k] * n * sin(arg);
        } else {
            double d = phi - T.i_phi0[k];
            eg1 = T.i_k[k] * d * d;
            dE = 2.0 * T.i_k[k] * d;
        }
        for (int j = 0; j < 4; j++) {
            double f[3] = {-dE * gr[j][0], -dE * gr[j][1], -dE * gr[j][2]};
            addf(Fbr, Sb, a[j], f);
        }
    } else if (t < e_fb) {  // flat-bottom positional restraint, minimum image
        int k = t - e_imp, a = T.fb_atom[k];
        double p[3];
        pos(a, p);
        double d[3] = {min_image(p[0] - T.fb_x0[3 * k], g.Lx, g.iLx),
                       min_image(p[1] - T.fb_x0[3 * k + 1], g.Ly, g.iLy),
                       min_image(p[2] - T.fb_x0[3 * k + 2], g.Lz, g.iLz)};
        double dist = sqrt(dot3(d, d)), x = dist - T.fb_tol[k];
        if (x >= 0.0 && dist > 0.0) {
            eg1 = T.fb_k[k] * x * x;
            double c = -2.0 * T.fb_k[k] * x / dist;
            double f[3] = {c * d[0], c * d[1], c * d[2]};
            addf(Fbr, Sb, a, f);
        }
    } else if (t < e_ph) {  // harmonic positional restraint per axis
        int k = t - e_fb, a = T.ph_atom[k];
        double p[3], f[3];
        pos(a, p);
        for (int d = 0; d < 3; d++) {
            double dx = p[d] - T.ph_x0[3 * k + d];
            eg1 += T.ph_k[3 * k + d] * dx * dx;
            f[d] = -2.0 * T.ph_k[3 * k + d] * dx;
        }
        addf(Fbr, Sb, a, f);
    } else if (t < e_ex) {  // 1-4 exception (group 2)
        int k = t - e_ph, i = T.ex_atoms[2 * k], j = T.ex_atoms[2 * k + 1];
        double pi[3], pj[3], d[3];
        pos(i, pi);
        pos(j, pj);
        for (int q = 0; q < 3; q++) d[q] = pi[q] - pj[q];
        double r2 = dot3(d, d), ir2 = 1.0 / r2, ir = sqrt(ir2);
        double s2 = T.ex_sig[k] * T.ex_sig[k] * ir2, s6 = s2 * s2 * s2;
        double e4 = 4.0 * T.ex_eps[k];
        eg2 = e4 * s6 * (s6 - 1.0) + T.ex_qq[k] * ir;
        double fr = (e4 * s6 * (12.0 * s6 - 6.0) + T.ex_qq[k] * ir) * ir2;  // -dE/dr / r
        double fi[3] = {fr * d[0], fr * d[1], fr * d[2]};
        double fj[3] = {-fi[0], -fi[1], -fi[2]};
        addf(Fr, S, i, fi);
        addf(Fr, S, j, fj);
    }
    if (bad) atomicOr(&flags[r], 2);
    red[0][threadIdx.x] = eg1;
    red[1][threadIdx.x] = eg2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            red[0][threadIdx.x] += red[0][threadIdx.x + o];
            red[1][threadIdx.x] += red[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        e1[(size_t)r * nblk + blockIdx.x] = red[0][0];
        e2[(size_t)r * nblk + blockIdx.x] = red[1][0];
    }
}

// ---------------------------------------------------------------------------
// displaced-state passes (FP64)
// ---------------------------------------------------------------------------

// candidate j-blocks: within rc of a displaced class's box at x or at x+d_k.  One
// 256-thread workgroup per replica (4 waves share the class boxes' reductions and the
// scan over the blocks, so the replica's serial chain is short); candidates are written
// in ascending block order.
constexpr int LIGCAND_THREADS = 256;
__global__ __launch_bounds__(LIGCAND_THREADS) void k_ligcand(
    int Np, int NB, int L, int ncls, const int* __restrict__ lig, const int* __restrict__ cls,
    const double* __restrict__ cdispl, const double* __restrict__ x64,
    const float4* __restrict__ bb_lo, const float4* __restrict__ bb_hi, int maxc,
    int* __restrict__ cand, int* __restrict__ ncand, float list_r2, Geom g,
    int* __restrict__ flags, int* __restrict__ cmap) {
    constexpr int NW = LIGCAND_THREADS / 64;
    __shared__ float4 blo[2 * SDM_MAX_CLASSES], bhi[2 * SDM_MAX_CLASSES];
    __shared__ double red[6][NW];
    __shared__ int wcount[NW];
    const int r = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const double* X = x64 + (size_t)r * 3 * Np;
    for (int k = 1; k <= ncls; k++) {
        double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
        for (int li = t; li < L; li += LIGCAND_THREADS) {
            int a = lig[li];
            if (cls[a] != k) continue;
            for (int d = 0; d < 3; d++) {
                double x = X[d * Np + a];
                v[d] = fmin(v[d], x);
                v[3 + d] = fmax(v[3 + d], x);
            }
        }
        for (int d = 0; d < 3; d++) {
            v[d] = wave_min_d(v[d]);
            v[3 + d] = wave_max_d(v[3 + d]);
        }
        if (lane == 0)
            for (int q = 0; q < 6; q++) red[q][wv] = v[q];
        __syncthreads();
        if (t == 0) {
            for (int w = 1; w < NW; w++)
                for (int d = 0; d < 3; d++) {
                    v[d] = fmin(v[d], red[d][w]);
                    v[3 + d] = fmax(v[3 + d], red[3 + d][w]);
                }
            const double* dk = cdispl + 3 * k;
            blo[2 * (k - 1)] = make_float4((float)v[0], (float)v[1], (float)v[2], 0.f);
            bhi[2 * (k - 1)] = make_float4((float)v[3], (float)v[4], (float)v[5], 0.f);
            blo[2 * (k - 1) + 1] = make_float4((float)(v[0] + dk[0]), (float)(v[1] + dk[1]),
                                               (float)(v[2] + dk[2]), 0.f);
            bhi[2 * (k - 1) + 1] = make_float4((float)(v[3] + dk[0]), (float)(v[4] + dk[1]),
                                               (float)(v[5] + dk[2]), 0.f);
        }
        __syncthreads();  // red is reused by the next class
    }
    const float4* lo = bb_lo + (size_t)r * NB;
    const float4* hi = bb_hi + (size_t)r * NB;
    int* out = cand + (size_t)r * maxc;
    int count = 0;  // candidates before this round (uniform)
    for (int base = 0; base < NB; base += LIGCAND_THREADS) {
        int jb = base + t;
        bool ok = false;
        if (jb < NB) {
            float4 lj = lo[jb], hj = hi[jb];
            for (int k = 0; k < 2 * ncls; k++) ok = ok || (bbox_gap2(blo[k], bhi[k], lj, hj, g) < list_r2);
        }
        unsigned long long m = __ballot(ok);
        if (lane == 0) wcount[wv] = __popcll(m);
        __syncthreads();
        int before = count;
        for (int w = 0; w < wv; w++) before += wcount[w];
        int pos = before + lane_rank(m);
        if (ok && pos < maxc) out[pos] = jb;
        // block -> candidate index (-1: none), for the Langevin kernel's gather of the
        // displaced-state forces
        if (cmap && jb < NB) cmap[(size_t)r * NB + jb] = (ok && pos < maxc) ? pos : -1;
        for (int w = 0; w < NW; w++) count += wcount[w];
        __syncthreads();  // wcount is rewritten next round
    }
    if (t == 0) {
        ncand[r] = min(count, maxc);
        if (count > maxc) atomicOr(&flags[r], 4);
    }
}

__device__ __forceinline__ bool excluded(const int* __restrict__ ex_start,
                                         const int* __restrict__ ex_list, int a, int b) {
    for (int k = ex_start[a]; k < ex_start[a + 1]; k++)
        if (ex_list[k] == b) return true;
    return false;
}

struct Atom64 {
    double x, y, z;      // position at x
    double dx, dy, dz;   // displacement
    double qk, sig, se;  // q*sqrt(k_e), sigma, sqrt(epsilon)
    int idx, cls;
};

__device__ __forceinline__ Atom64 load_atom64(const double* __restrict__ X, int Np, int a,
                                              const int* __restrict__ cls,
                                              const double* __restrict__ cdispl,
                                              const double* __restrict__ q,
                                              const double* __restrict__ sig,
                                              const double* __restrict__ eps) {
    Atom64 A;
    A.x = X[a];
    A.y = X[Np + a];
    A.z = X[2 * Np + a];
    A.cls = cls[a];
    A.dx = cdispl[3 * A.cls];
    A.dy = cdispl[3 * A.cls + 1];
    A.dz = cdispl[3 * A.cls + 2];
    A.qk = q[a];
    A.sig = sig[a];
    A.se = sqrt(eps[a]);
    A.idx = a;
    return A;
}

// the two geometries of one pair (i displaced by d_i, j by d_j): returns
// e(x+d) - e(x) and adds F1-F0 on i to (fx,fy,fz).
__device__ __forceinline__ double delta_pair(const Atom64& I, const Atom64& J, const Geom& g,
                                             double& fx, double& fy, double& fz, int* inside_y) {
    double qq = I.qk * J.qk;
    double sig = g.geo ? sqrt(I.sig * J.sig) : 0.5 * (I.sig + J.sig);
    double eps4 = 4.0 * I.se * J.se;
    double ex, fr_x, ey, fr_y;
    double dx = min_image(I.x - J.x, g.Lx, g.iLx);
    double dy = min_image(I.y - J.y, g.Ly, g.iLy);
    double dz = min_image(I.z - J.z, g.Lz, g.iLz);
    pair64(dx, dy, dz, qq, sig, eps4, g.rc2, g.krf, g.crf, &ex, &fr_x);
    double ex_ = min_image((I.x + I.dx) - (J.x + J.dx), g.Lx, g.iLx);
    double ey_ = min_image((I.y + I.dy) - (J.y + J.dy), g.Ly, g.iLy);
    double ez_ = min_image((I.z + I.dz) - (J.z + J.dz), g.Lz, g.iLz);
    *inside_y = pair64(ex_, ey_, ez_, qq, sig, eps4, g.rc2, g.krf, g.crf, &ey, &fr_y);
    fx += fr_y * ex_ - fr_x * dx;
    fy += fr_y * ey_ - fr_x * dy;
    fz += fr_y * ez_ - fr_x * dz;
    return ey - ex;
}

// Up to 64 atoms staged in LDS by one wave (lane t copies atom t), so that the loops over
// partner atoms read LDS broadcasts instead of chains of dependent global loads.
struct AtomTile64 {
    double x[64], y[64], z[64], dx[64], dy[64], dz[64], q[64], sig[64], se[64];
    int cls[64], idx[64];
};

__device__ __forceinline__ void stage_atom(AtomTile64& T, int t, const double* __restrict__ X,
                                           int Np, int a, const int* __restrict__ cls,
                                           const double* __restrict__ cdispl,
                                           const double* __restrict__ q,
                                           const double* __restrict__ sig,
                                           const double* __restrict__ eps) {
    int k = cls[a];
    T.x[t] = X[a];
    T.y[t] = X[Np + a];
    T.z[t] = X[2 * Np + a];
    T.dx[t] = cdispl[3 * k];
    T.dy[t] = cdispl[3 * k + 1];
    T.dz[t] = cdispl[3 * k + 2];
    T.q[t] = q[a];
    T.sig[t] = sig[a];
    T.se[t] = sqrt(eps[a]);
    T.cls[t] = k;
    T.idx[t] = a;
}

__device__ __forceinline__ Atom64 tile_atom(const AtomTile64& T, int t) {
    Atom64 A;
    A.x = T.x[t];
    A.y = T.y[t];
    A.z = T.z[t];
    A.cls = T.cls[t];
    A.dx = T.dx[t];
    A.dy = T.dy[t];
    A.dz = T.dz[t];
    A.qk = T.q[t];
    A.sig = T.sig[t];
    A.se = T.se[t];
    A.idx = T.idx[t];
    return A;
}

// whether atom a has an exclusion partner that is displaced (want_displaced) or not.
// Exclusions across the boundary of the displaced set are rare (a bond from the ligand
// to the receptor), so only atoms that have one search their list per pair.
__device__ __forceinline__ bool excl_with(const int* __restrict__ ex_start,
                                          const int* __restrict__ ex_list,
                                          const int* __restrict__ cls, int a,
                                          bool want_displaced) {
    bool any = false;
    for (int k = ex_start[a]; k < ex_start[a + 1]; k++)
        any = any || ((cls[ex_list[k]] != 0) == want_displaced);
    return any;
}

// The displaced-state pairs: every (displaced atom i, undisplaced atom j) within the
// cutoff at x or at x+d, evaluated in FP64 at both geometries (delta_pair).  Two passes
// evaluate each such pair once from each side, so that every sum has one owner lane and
// a fixed order (no atomics):
//   k_denv  lanes = the 64 slots of candidate block c.  An FP32 screen of the block's
//           undisplaced atoms against every displaced atom (at x and at x+d, with a
//           margin that covers FP32 rounding) gives each lane its 64-bit list of partners
//           per ligand block, whose pairs it then evaluates: F1-F0 on the undisplaced
//           atom (dB).  The screen's ballot per displaced atom is the same list seen
//           from the other side; it goes to mlig[r][c][li].
//   k_dlig  lanes = the displaced atoms of ligand block lb, one candidate block per
//           workgroup: each lane evaluates the pairs of its mlig bits, giving u and F1-F0
//           on the displaced atom, one partial per candidate block (dA), added in block
//           order by k_deltaA_red.
// A few blocks near the displaced atoms hold most pairs, so both spread a block's pairs
// over DW waves (wave w takes the partners with index = w mod DW, ascending) and add
// the waves' partials in wave order.  Excluded pairs are left out of the lists (only
// atoms with an exclusion across the displaced/undisplaced boundary search their
// exclusion list).
constexpr float DSCREEN_MARGIN = 1e-3f;  // nm; FP32 position rounding is below 1e-6 nm
constexpr int DPAIR_WAVES = 4;

template <int DW>
__global__ __launch_bounds__(64 * DW) void k_denv(
    int N, int Np, int L, int LB, const int* __restrict__ lig,
    const int* __restrict__ lb_start, const int* __restrict__ lb_count,
    const int* __restrict__ cls, const double* __restrict__ cdispl,
    const double* __restrict__ q64, const double* __restrict__ sig64,
    const double* __restrict__ eps64, const int* __restrict__ ex_start,
    const int* __restrict__ ex_list, const double* __restrict__ x64,
    const int* __restrict__ orig_s, const int* __restrict__ cand, const int* __restrict__ ncand,
    int maxc, unsigned long long* __restrict__ mlig, double4* __restrict__ dB, Geom g) {
    int c = blockIdx.x, r = blockIdx.y, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (c >= ncand[r]) return;  // uniform over the workgroup
    const double* X = x64 + (size_t)r * 3 * Np;
    int s = cand[(size_t)r * maxc + c] * 64 + lane;
    bool v = s < N;
    int b = v ? orig_s[(size_t)r * Np + s] : 0;
    Atom64 J = load_atom64(X, Np, b, cls, cdispl, q64, sig64, eps64);
    bool env = v && J.cls == 0;
    bool exJ = env && excl_with(ex_start, ex_list, cls, b, true);
    const float jx = (float)J.x, jy = (float)J.y, jz = (float)J.z;
    const float rs = (float)sqrt(g.rc2) + DSCREEN_MARGIN, rs2 = rs * rs;
    __shared__ AtomTile64 T;
    __shared__ float4 P0[64], P1[64];  // the ligand block at x and at x+d, FP32
    __shared__ double4 part[DW][64];
    unsigned long long* ML = mlig + ((size_t)r * maxc + c) * L;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int lb = 0; lb < LB; lb++) {
        const int base = lb_start[lb], n = lb_count[lb];
        __syncthreads();  // the previous ligand block has been read
        if ((int)threadIdx.x < n) {
            int a = lig[base + threadIdx.x];
            stage_atom(T, threadIdx.x, X, Np, a, cls, cdispl, q64, sig64, eps64);
            const double* d = cdispl + 3 * cls[a];
            double x = X[a], y = X[Np + a], z = X[2 * Np + a];
            P0[threadIdx.x] = make_float4((float)x, (float)y, (float)z, 0.f);
            P1[threadIdx.x] =
                make_float4((float)(x + d[0]), (float)(y + d[1]), (float)(z + d[2]), 0.f);
        }
        __syncthreads();
        unsigned long long mine = 0ull, col = 0ull;
        for (int k = wv; k < n; k += DW) {
            float4 p0 = P0[k], p1 = P1[k];
            float dx = jx - p0.x, dy = jy - p0.y, dz = jz - p0.z;
            dx -= g.bx * rintf(dx * g.ibx);
            dy -= g.by * rintf(dy * g.iby);
            dz -= g.bz * rintf(dz * g.ibz);
            float ex = jx - p1.x, ey = jy - p1.y, ez = jz - p1.z;
            ex -= g.bx * rintf(ex * g.ibx);
            ey -= g.by * rintf(ey * g.iby);
            ez -= g.bz * rintf(ez * g.ibz);
            bool in = env && (dx * dx + dy * dy + dz * dz < rs2 || ex * ex + ey * ey + ez * ez < rs2);
            if (exJ && in) in = !excluded(ex_start, ex_list, b, T.idx[k]);
            unsigned long long bal = __ballot(in);
            if (lane == k) col = bal;
            mine |= (unsigned long long)in << k;
        }
        if (lane < n && lane % DW == wv) ML[base + lane] = col;
        while (mine) {  // this lane's partners of this wave, ascending
            int k = __ffsll((long long)mine) - 1;
            mine &= mine - 1ull;
            Atom64 I = tile_atom(T, k);
            int iny;
            delta_pair(J, I, g, fx, fy, fz, &iny);
        }
    }
    part[wv][lane] = make_double4(fx, fy, fz, 0.0);
    __syncthreads();
    if (wv == 0) {
        double4 t = part[0][lane];
        for (int w = 1; w < DW; w++) {
            t.x += part[w][lane].x;
            t.y += part[w][lane].y;
            t.z += part[w][lane].z;
        }
        dB[((size_t)r * maxc + c) * 64 + lane] = t;
    }
}

template <int DW>
__global__ __launch_bounds__(64 * DW) void k_dlig(
    int N, int Np, int L, int LB, const int* __restrict__ lig,
    const int* __restrict__ lb_start, const int* __restrict__ lb_count,
    const int* __restrict__ cls, const double* __restrict__ cdispl,
    const double* __restrict__ q64, const double* __restrict__ sig64,
    const double* __restrict__ eps64, const double* __restrict__ x64,
    const int* __restrict__ orig_s, const int* __restrict__ cand, const int* __restrict__ ncand,
    int maxc, const unsigned long long* __restrict__ mlig, double4* __restrict__ dA, Geom g,
    unsigned long long* __restrict__ pair_count) {
    int c = blockIdx.x, lb = blockIdx.y, r = blockIdx.z;
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (c >= ncand[r]) return;  // uniform; k_deltaA_red reads only the blocks below ncand
    const double* X = x64 + (size_t)r * 3 * Np;
    int li = lb_start[lb] + lane;
    bool v = lane < lb_count[lb];
    // this wave's partners: slots = wv mod DW
    unsigned long long sel = 0ull;
    for (int k = wv; k < 64; k += DW) sel |= 1ull << k;
    unsigned long long m = v ? (mlig[((size_t)r * maxc + c) * L + li] & sel) : 0ull;
    __shared__ AtomTile64 T;
    __shared__ double4 part[DW][64];
    __shared__ int any_s;
    if (threadIdx.x == 0) any_s = 0;
    __syncthreads();
    const bool wany = __ballot(m != 0ull) != 0ull;
    if (lane == 0 && wany) any_s = 1;
    __syncthreads();
    const bool any = any_s != 0;  // uniform over the workgroup
    double du = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    int cnt = 0;
    if (any) {
        int s = cand[(size_t)r * maxc + c] * 64 + lane;
        if (wv == 0 && s < N)
            stage_atom(T, lane, X, Np, orig_s[(size_t)r * Np + s], cls, cdispl, q64, sig64, eps64);
        Atom64 I = load_atom64(X, Np, v ? lig[li] : lig[lb_start[lb]], cls, cdispl, q64, sig64,
                               eps64);
        __syncthreads();
        while (m) {  // this lane's partners of this wave, ascending slot
            int k = __ffsll((long long)m) - 1;
            m &= m - 1ull;
            Atom64 J = tile_atom(T, k);
            int iny;
            du += delta_pair(I, J, g, fx, fy, fz, &iny);
            cnt += iny;
        }
    }
    if (pair_count) {
        double sc = wave_sum((double)cnt);
        if (lane == 0 && any) atomicAdd(&pair_count[r], (unsigned long long)(sc + 0.5));
        return;
    }
    part[wv][lane] = make_double4(fx, fy, fz, du);
    __syncthreads();
    if (wv == 0) {
        double4 t = part[0][lane];
        for (int w = 1; w < DW; w++) {
            t.x += part[w][lane].x;
            t.y += part[w][lane].y;
            t.z += part[w][lane].z;
            t.w += part[w][lane].w;
        }
        dA[(((size_t)r * maxc + c) * LB + lb) * 64 + lane] = t;
    }
}

// per displaced atom: the block partials of k_dlig in order, plus pairs with displaced
// atoms of a different class (D x D); writes dlig and the ligand block's u.
template <int RW>
__global__ __launch_bounds__(64 * RW) void k_deltaA_red(
    int Np, int L, int LB, const int* __restrict__ lig, const int* __restrict__ lb_start,
    const int* __restrict__ lb_count, const int* __restrict__ cls,
    const double* __restrict__ cdispl, const double* __restrict__ q64,
    const double* __restrict__ sig64, const double* __restrict__ eps64,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list,
    const double* __restrict__ x64, const int* __restrict__ ncand, int maxc,
    const double4* __restrict__ dA, double4* __restrict__ dlig, double* __restrict__ u_part,
    Geom g, unsigned long long* __restrict__ pair_count) {
    // RW waves: wave w adds the partials of blocks w, w + RW, .. and the D x D pairs of
    // displaced atoms w, w + RW, ..; the waves' sums are added in wave order
    int lb = blockIdx.x, r = blockIdx.y, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const double* X = x64 + (size_t)r * 3 * Np;
    int li = lb_start[lb] + lane;
    bool v = lane < lb_count[lb];
    Atom64 I = load_atom64(X, Np, v ? lig[li] : lig[lb_start[lb]], cls, cdispl, q64, sig64, eps64);
    double du = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    const int nc = ncand[r];
#pragma unroll 4
    for (int c = wv; c < nc; c += RW) {
        double4 p = dA[(((size_t)r * maxc + c) * LB + lb) * 64 + lane];
        fx += p.x;
        fy += p.y;
        fz += p.z;
        du += p.w;
    }
    __shared__ AtomTile64 T;
    __shared__ double4 part[RW][64];
    int cnt = 0;
    for (int base = 0; base < L; base += 64) {
        int n = min(64, L - base);
        __syncthreads();  // the previous chunk has been read
        if ((int)threadIdx.x < n)
            stage_atom(T, threadIdx.x, X, Np, lig[base + threadIdx.x], cls, cdispl, q64, sig64,
                       eps64);
        __syncthreads();
        for (int k = wv; k < n; k += RW) {
            Atom64 J = tile_atom(T, k);
            if (!v || J.cls == I.cls || excluded(ex_start, ex_list, I.idx, J.idx)) continue;
            int iny;
            du += 0.5 * delta_pair(I, J, g, fx, fy, fz, &iny);
            cnt += (iny && I.idx < J.idx) ? 1 : 0;
        }
    }
    if (pair_count) {
        double s = wave_sum((double)cnt);
        if (lane == 0) atomicAdd(&pair_count[r], (unsigned long long)(s + 0.5));
        return;
    }
    part[wv][lane] = make_double4(fx, fy, fz, du);
    __syncthreads();
    if (wv == 0) {
        double4 t = part[0][lane];
        for (int w = 1; w < RW; w++) {
            t.x += part[w][lane].x;
            t.y += part[w][lane].y;
            t.z += part[w][lane].z;
            t.w += part[w][lane].w;
        }
        if (v) dlig[(size_t)r * L + li] = t;
        double us = wave_sum(v ? t.w : 0.0);
        if (lane == 0) u_part[(size_t)r * LB + lb] = us;
    }
}
constexpr int DRED_WAVES = 4;

// ---------------------------------------------------------------------------
// scalar stage (one workgroup per replica): energies, SoftCoreF, bias, sp, schedule.
// Restates ReferenceSDMKernels.cpp:202-309 and LangevinIntegratorSDM.cpp:125-149.
// ---------------------------------------------------------------------------
__device__ double softcore_dev(int method, double u, double umax, double a, double ub,
                               double* fp) {
    if (u <= ub || method == 0) {
        *fp = 1.0;
        return u;
    }
    if (method == 1) {
        double t = tanh((u - ub) / umax);
        *fp = 1.0 - t * t;
        return umax * t + ub;
    }
    double gu = (u - ub) / (a * (umax - ub));
    double zeta = 1.0 + 2.0 * gu * (gu + 1.0);
    double zetap = pow(zeta, a);
    double s = 4.0 * (2.0 * gu + 1.0) / zeta;
    *fp = s * zetap / ((1.0 + zetap) * (1.0 + zetap));
    return (umax - ub) * (zetap - 1.0) / (zetap + 1.0) + ub;
}

// ne: pair-energy slots per replica (i-blocks times the pair kernel's waves per block)
__global__ __launch_bounds__(256) void k_scalar(
    int ne, int nbw, int LB, const double* __restrict__ e_pair,
    const double* __restrict__ e_bond, const double* __restrict__ u_part,
    DevState* __restrict__ state, DevRecord* __restrict__ rec, int* __restrict__ flags,
    int advance, int energy, int hist_every, int hist_cap, double* __restrict__ uhist,
    const double* __restrict__ e_rest, int nrest, const double* __restrict__ e_t1,
    const double* __restrict__ e_t2, int ntb, float* __restrict__ dmax) {
    __shared__ double red[256];
    int r = blockIdx.x, t = threadIdx.x;
    // the displacement record of kept tile lists restarts with the next slot refresh:
    // this evaluation's k_tiles has read it and kept or rebuilt the lists (after a sort,
    // rebuilt), and the next refresh measures the displacement since that build again
    if (t == 0) dmax[(size_t)r * DISP_STRIDE] = 0.f;
    double s0 = 0.0, s1 = 0.0;
    // energy == 0: the pair kernel left the pair energies out; U0, U1 and pot_energy
    // keep the values of the last evaluation that had them (nothing else depends on U0)
    double U0 = rec[r].U0;
    if (energy) {
        for (int k = t; k < ne; k += 256) s0 += e_pair[(size_t)r * ne + k];
        red[t] = s0;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (t < o) red[t] += red[t + o];
            __syncthreads();
        }
        U0 = red[0];
        __syncthreads();
    }
    for (int k = t; k < nbw; k += 256) s1 += e_bond[(size_t)r * nbw + k];
    red[t] = s1;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    double Ub = red[0];
    if (t != 0) return;
    for (int k = 0; k < nrest; k++) Ub += e_rest[(size_t)r * nrest + k];  // restraints
    for (int k = 0; k < ntb; k++) {  // further force-field terms: groups 1 and 2
        Ub += e_t1[(size_t)r * ntb + k];
        if (energy) U0 += e_t2[(size_t)r * ntb + k];
    }
    double u = 0.0;
    for (int k = 0; k < LB; k++) u += u_part[(size_t)r * LB + k];
    DevState& S = state[r];
    DevRecord& Rr = rec[r];
    // nonequilibrium schedule (ReferenceSDMKernels.cpp:221-245)
    double stepSize = S.step_size;
    double lambdac = S.lambda, dlambdac = 0.0, gamma = 0.0, wbcoeff = lambdac, w0coeff = 0.0;
    double lambda1 = lambdac, lambda2 = lambdac, alpha = 1.0, u0 = 0.0;
    double alpha_now = S.alpha;  // alpha of this step (scheduled while switching)
    if (S.nonequilibrium == 1) {
        lambdac = S.time / S.noneq_tmax;
        S.lambda = lambdac;
        S.lambda1 = S.m_l1 * lambdac + S.b_l1;
        S.lambda2 = S.m_l2 * lambdac + S.b_l2;
        S.u0 = S.m_u0 * lambdac + S.b_u0;
        S.w0coeff = S.m_w0 * lambdac + S.b_w0;
        dlambdac = stepSize / S.noneq_tmax;
        alpha_now = S.alpha + S.m_alpha * lambdac;
    }
    int method = S.bias_method;
    if (method == 1) {
        gamma = S.gamma;
        wbcoeff = S.wbcoeff;
        w0coeff = S.w0coeff;
    } else if (method == 2) {
        lambda1 = S.lambda1;
        lambda2 = S.lambda2;
        alpha = alpha_now;
        u0 = S.u0;
        w0coeff = S.w0coeff;
    }
    double fp;
    double be = softcore_dev(S.softcore_method, u, S.umax, S.acore, S.ubcore, &fp);
    double ebias, bfp;
    if (method == 1) {
        ebias = 0.5 * gamma * be * be + wbcoeff * be + w0coeff;
        bfp = gamma * be + wbcoeff;
    } else if (method == 2) {
        double z = -alpha * (be - u0);  // ee = 1 + e^z
        ebias = (alpha > 0) ? ((lambda2 - lambda1) / alpha) * softplus(z) : 0.0;
        ebias += lambda2 * be + w0coeff;
        bfp = (lambda2 - lambda1) * inv_one_plus_exp(z) + lambda1;
    } else {
        ebias = lambdac * be;
        bfp = lambdac;
    }
    if (S.nonequilibrium == 1) {
        double z = -alpha * (be - u0);
        double lee = softplus(z);
        double dwdl1 = -lee / alpha;
        double dwdl2 = be + lee / alpha;
        double dwdu0 = (lambda2 - lambda1) * (1.0 - inv_one_plus_exp(z));
        // dW/dalpha of the ILogistic bias: -(l2-l1) ln(1+e^z)/alpha^2 - (u'-u0) dW/du0/alpha
        double dwdalpha = (alpha > 0) ? -(lambda2 - lambda1) * lee / (alpha * alpha) -
                                            (be - u0) * dwdu0 / alpha
                                      : 0.0;
        double dwdlambda = dwdl1 * S.m_l1 + dwdl2 * S.m_l2 + dwdu0 * S.m_u0 + 1.0 * S.m_w0 +
                           dwdalpha * S.m_alpha;
        S.work += dlambdac * dwdlambda;
    }
    double sp = bfp * fp;
    double pot = U0 + ebias + Ub;
    int fl = 0;
    if (!isfinite(u) || !isfinite(be) || !isfinite(sp)) fl |= 1;
    if (energy && (!isfinite(U0) || !isfinite(pot))) fl |= 1;
    if (fl) atomicOr(&flags[r], fl);
    if (energy) {
        Rr.U0 = U0;
        Rr.U1 = U0 + u;
        Rr.pot_energy = pot;
    }
    Rr.u = u;
    Rr.Ub = Ub;
    Rr.bind_e = be;
    Rr.fp = fp;
    Rr.W = ebias;
    Rr.Wp = bfp;
    Rr.sp = sp;
    Rr.lambda = lambdac;
    Rr.lambda1 = (method == 2) ? lambda1 : S.lambda1;
    Rr.lambda2 = (method == 2) ? lambda2 : S.lambda2;
    Rr.u0 = (method == 2) ? u0 : S.u0;
    Rr.w0 = S.w0coeff;
    Rr.work = S.work;
    // Langevin parameters (OpenCLSDMKernels.cpp:331-355; equal to
    // ReferenceStochasticDynamicsSDM.cpp:144-148 for friction > 0)
    double kT = SDM_BOLTZ * S.temperature;
    double vscale = exp(-stepSize * S.friction);
    Rr.vscale = vscale;
    Rr.fscale = (S.friction == 0.0) ? stepSize : (1.0 - vscale) / S.friction;
    Rr.noisescale = sqrt(kT * (1.0 - vscale * vscale));
    Rr.dt = stepSize;
    Rr.noise_step = S.step;
    if (hist_every > 0 && advance && (S.step % hist_every) == 0) {
        uhist[(size_t)r * hist_cap + (size_t)(S.hist_n % hist_cap)] = be;
        S.hist_n += 1;
    }
    if (advance) {
        S.time += stepSize;
        S.step += 1;
    }
    Rr.time = S.time;
    Rr.step = S.step;
    Rr.flags = flags[r];
}

// F += sp (F1 - F0): candidate-block atoms from dB, displaced atoms from dlig (the
// accumulator is in slot order).
__global__ __launch_bounds__(64) void k_mix(int N, int Np, int L, int maxc,
                                            const int* __restrict__ lig, const int* __restrict__ cls,
                                            const int* __restrict__ orig_s,
                                            const int* __restrict__ cand, const int* __restrict__ ncand,
                                            const double4* __restrict__ dB,
                                            const double4* __restrict__ dlig,
                                            const DevRecord* __restrict__ rec,
                                            const int* __restrict__ slot_of,
                                            long long* __restrict__ F, int* __restrict__ flags) {
    int r = blockIdx.y, lane = threadIdx.x;
    double sp = rec[r].sp;
    long long* Fr = F + (size_t)r * 3 * Np;
    const int* S = slot_of + (size_t)r * Np;
    int bad = 0;
    if ((int)blockIdx.x < maxc) {
        int c = blockIdx.x;
        if (c >= ncand[r]) return;
        int s = cand[(size_t)r * maxc + c] * 64 + lane;
        if (s >= N) return;
        int b = orig_s[(size_t)r * Np + s];
        if (cls[b] != 0) return;
        double4 d = dB[((size_t)r * maxc + c) * 64 + lane];
        Fr[s] += to_fixed(sp * d.x, &bad);
        Fr[Np + s] += to_fixed(sp * d.y, &bad);
        Fr[2 * Np + s] += to_fixed(sp * d.z, &bad);
    } else {
        int li = (blockIdx.x - maxc) * 64 + lane;
        if (li >= L) return;
        int a = S[lig[li]];
        double4 d = dlig[(size_t)r * L + li];
        Fr[a] += to_fixed(sp * d.x, &bad);
        Fr[Np + a] += to_fixed(sp * d.y, &bad);
        Fr[2 * Np + a] += to_fixed(sp * d.z, &bad);
    }
    if (bad) atomicOr(&flags[r], 2);
}

// leapfrog Langevin (langevin.cl:7-69, double-precision branch).  The force is the
// fixed-point accumulator (slot order) plus, when there are bonded terms, their FP64
// forces Fb (atom order).  Atoms with a constraint (cflag) leave their displacement
// dt*v in dxs for k_constrain.
// MixGather: k_mix's work done here instead, per atom (one launch fewer per step): the
// displaced-state force sp * dF of atom i, from the ligand sums (lig_of: the atom's
// ligand index, -1 for environment atoms) or from its block's environment sums (cmap:
// block -> candidate index of this step, -1 for none), added to the accumulator as k_mix
// adds it, so the accumulator holds the same bits afterwards.
struct MixGather {
    const int* lig_of;
    const int* cmap;
    const double4* dB;
    const double4* dlig;
    int NB, maxc, L;
};

__global__ __launch_bounds__(256) void k_langevin(int N, int Np, const double* __restrict__ invm,
                                                  long long* __restrict__ F,
                                                  const double* __restrict__ Fb,
                                                  double* __restrict__ x64, double* __restrict__ v64,
                                                  const DevRecord* __restrict__ rec,
                                                  const DevState* __restrict__ state,
                                                  const int* __restrict__ slot_of,
                                                  const int* __restrict__ cflag,
                                                  double* __restrict__ dxs, MixGather mg,
                                                  int* __restrict__ flags) {
    int r = blockIdx.y;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const DevRecord& Rr = rec[r];
    if (mg.lig_of) {
        const int li = mg.lig_of[i];
        const int sl = slot_of[(size_t)r * Np + i];
        const double4* dp = nullptr;
        if (li >= 0) {
            dp = mg.dlig + (size_t)r * mg.L + li;
        } else {
            const int cb = mg.cmap[(size_t)r * mg.NB + (sl >> 6)];
            if (cb >= 0) dp = mg.dB + ((size_t)r * mg.maxc + cb) * 64 + (sl & 63);
        }
        if (dp) {
            const double4 d = *dp;
            const double sp = Rr.sp;
            long long* Fm = F + (size_t)r * 3 * Np;
            int bad = 0;
            Fm[sl] += to_fixed(sp * d.x, &bad);
            Fm[Np + sl] += to_fixed(sp * d.y, &bad);
            Fm[2 * Np + sl] += to_fixed(sp * d.z, &bad);
            if (bad) atomicOr(&flags[r], 2);
        }
    }
    double im = invm[i];
    if (im == 0.0) return;
    double vscale = Rr.vscale, fscale = Rr.fscale, noise = Rr.noisescale, dt = Rr.dt;
    double sim = sqrt(im);
    double gz[3];
    gauss3(state[r].seed, (uint32_t)r, (uint64_t)Rr.noise_step, (uint32_t)i, gz);
    const long long* Fr = F + (size_t)r * 3 * Np;
    double* X = x64 + (size_t)r * 3 * Np;
    double* V = v64 + (size_t)r * 3 * Np;
    double inv = 1.0 / dt;
    int slot = slot_of[(size_t)r * Np + i];
    bool con = cflag && cflag[i];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double f = from_fixed(Fr[d * Np + slot]);
        if (Fb) f += Fb[((size_t)r * 3 + d) * Np + i];
        double vel = vscale * V[d * Np + i] + fscale * im * f + noise * sim * gz[d];
        double delta = dt * vel;
        if (con) {
            dxs[(size_t)r * 3 * Np + d * Np + i] = delta;
        } else {
            X[d * Np + i] += delta;
            V[d * Np + i] = inv * delta;
        }
    }
}

// Virtual sites (massless atoms placed from others; the reference's integrator places
// them after every step, ReferenceStochasticDynamicsSDM.cpp:264, from the definitions the
// DMS reader makes, desmonddmsfile75.py _addVirtualSitesToSystem).  Site k: at[k] = (site,
// p1, p2, p3), weights w[3k..3k+2]:
//   VSITE_AVERAGE2:      x = w0 x1 + w1 x2
//   VSITE_AVERAGE3:      x = w0 x1 + w1 x2 + w2 x3
//   VSITE_OUT_OF_PLANE:  x = x1 + w0 r12 + w1 r13 + w2 (r12 x r13),  r1k = xk - x1
// (coordinates as stored, without periodic images, as OpenMM 7 places them).
constexpr int VSITE_AVERAGE2 = 0, VSITE_AVERAGE3 = 1, VSITE_OUT_OF_PLANE = 2;

__device__ __forceinline__ void vsite_cross(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// before the Langevin update: each site's force (fixed-point accumulator plus the FP64
// bonded forces) moves to its parents by the transpose of the site's Jacobian, through
// fixed-point atomics (parents are never sites, so no site's own entry is written by
// another thread); the site's entries are left at zero.  Out of plane: a parent k's share
// is w f plus w2 times the cross-product term's derivative (r13 x f for p2, f x r12 for p3),
// and p1 takes the rest, so the three shares sum to f.
__global__ __launch_bounds__(256) void k_vsite_forces(int nvs, int Np,
                                                      const int* __restrict__ kind,
                                                      const int4* __restrict__ at,
                                                      const double* __restrict__ w,
                                                      const double* __restrict__ x64,
                                                      const int* __restrict__ slot_of,
                                                      long long* __restrict__ F,
                                                      double* __restrict__ Fb,
                                                      int* __restrict__ flags) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (k >= nvs) return;
    const int4 a = at[k];
    const int K = kind[k];
    const double w0 = w[3 * k], w1 = w[3 * k + 1], w2 = w[3 * k + 2];
    long long* Fr = F + (size_t)r * 3 * Np;
    double* Fbr = Fb ? Fb + (size_t)r * 3 * Np : nullptr;
    const int* S = slot_of + (size_t)r * Np;
    const double* X = x64 + (size_t)r * 3 * Np;
    const int ss = S[a.x];
    double f[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        f[d] = from_fixed(Fr[d * Np + ss]);
        Fr[d * Np + ss] = 0;
        if (Fbr) {
            f[d] += Fbr[d * Np + a.x];
            Fbr[d * Np + a.x] = 0.0;
        }
    }
    double g1[3], g2[3], g3[3];
    if (K == VSITE_OUT_OF_PLANE) {
        double r12[3], r13[3], c2[3], c3[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            r12[d] = X[d * Np + a.z] - X[d * Np + a.y];
            r13[d] = X[d * Np + a.w] - X[d * Np + a.y];
        }
        vsite_cross(r13, f, c2);
        vsite_cross(f, r12, c3);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            g2[d] = w0 * f[d] + w2 * c2[d];
            g3[d] = w1 * f[d] + w2 * c3[d];
            g1[d] = f[d] - g2[d] - g3[d];
        }
    } else {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            g1[d] = w0 * f[d];
            g2[d] = w1 * f[d];
            g3[d] = K == VSITE_AVERAGE3 ? w2 * f[d] : 0.0;
        }
    }
    int bad = 0;
    const int s1 = S[a.y], s2 = S[a.z];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&Fr[d * Np + s1]),
                  (unsigned long long)to_fixed(g1[d], &bad));
        atomicAdd(reinterpret_cast<unsigned long long*>(&Fr[d * Np + s2]),
                  (unsigned long long)to_fixed(g2[d], &bad));
    }
    if (K != VSITE_AVERAGE2) {
        const int s3 = S[a.w];
#pragma unroll
        for (int d = 0; d < 3; d++)
            atomicAdd(reinterpret_cast<unsigned long long*>(&Fr[d * Np + s3]),
                      (unsigned long long)to_fixed(g3[d], &bad));
    }
    if (bad) atomicOr(&flags[r], 2);
}

// after the update and the constraints: the sites' positions from their parents'
__global__ __launch_bounds__(256) void k_vsite_pos(int nvs, int Np, const int* __restrict__ kind,
                                                   const int4* __restrict__ at,
                                                   const double* __restrict__ w,
                                                   double* __restrict__ x64) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (k >= nvs) return;
    const int4 a = at[k];
    const int K = kind[k];
    const double w0 = w[3 * k], w1 = w[3 * k + 1], w2 = w[3 * k + 2];
    double* X = x64 + (size_t)r * 3 * Np;
    double p1[3], p2[3], p3[3] = {0.0, 0.0, 0.0}, o[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        p1[d] = X[d * Np + a.y];
        p2[d] = X[d * Np + a.z];
        if (K != VSITE_AVERAGE2) p3[d] = X[d * Np + a.w];
    }
    if (K == VSITE_OUT_OF_PLANE) {
        double r12[3], r13[3], c[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            r12[d] = p2[d] - p1[d];
            r13[d] = p3[d] - p1[d];
        }
        vsite_cross(r12, r13, c);
#pragma unroll
        for (int d = 0; d < 3; d++) o[d] = p1[d] + w0 * r12[d] + w1 * r13[d] + w2 * c[d];
    } else {
#pragma unroll
        for (int d = 0; d < 3; d++)
            o[d] = w0 * p1[d] + w1 * p2[d] + (K == VSITE_AVERAGE3 ? w2 * p3[d] : 0.0);
    }
#pragma unroll
    for (int d = 0; d < 3; d++) X[d * Np + a.x] = o[d];
}

// centre-of-mass motion removal (one workgroup per replica) on steps whose index (the
// step count before the step, as OpenMM's CMMotionRemover tests it in the
// updateContextState call that starts the step, LangevinIntegratorSDM.cpp:158) is a
// multiple of `every`: v -= sum(m v) / sum(m) for every atom with mass.  Launched after
// the scalar stage has advanced the count, so the step's index is state.step - 1.
__global__ __launch_bounds__(1024) void k_cm_remove(int N, int Np, int every,
                                                    const double* __restrict__ mass,
                                                    double* __restrict__ v64,
                                                    const DevState* __restrict__ state) {
    __shared__ double red[4][1024];
    int r = blockIdx.x, t = threadIdx.x;
    if ((state[r].step - 1) % every != 0) return;  // uniform over the workgroup
    double* V = v64 + (size_t)r * 3 * Np;
    double p[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = t; i < N; i += 1024) {
        double m = mass[i];
        p[0] += m * V[i];
        p[1] += m * V[Np + i];
        p[2] += m * V[2 * Np + i];
        p[3] += m;
    }
    for (int d = 0; d < 4; d++) red[d][t] = p[d];
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (t < o)
            for (int d = 0; d < 4; d++) red[d][t] += red[d][t + o];
        __syncthreads();
    }
    double im = red[3][0] > 0.0 ? 1.0 / red[3][0] : 0.0;
    double cx = red[0][0] * im, cy = red[1][0] * im, cz = red[2][0] * im;
    for (int i = t; i < N; i += 1024) {
        if (mass[i] == 0.0) continue;
        V[i] -= cx;
        V[Np + i] -= cy;
        V[2 * Np + i] -= cz;
    }
}

// Position constraints of the Langevin step: SHAKE on one cluster of constrained atoms
// per thread (Ryckaert, Ciccotti, Berendsen, J. Comput. Phys. 23, 327 (1977)): each
// constraint in turn moves its two atoms along their bond vector before the step, in
// inverse proportion to their masses, until every |r_ij|^2 is within tol of d_ij^2.
// Then x = x', v = (x' - x)/dt.  Sets flag 8 if a cluster does not converge.
constexpr int SHAKE_MAX_ATOMS = 16, SHAKE_MAX_CONS = 32, SHAKE_MAX_ITER = 500;
__global__ __launch_bounds__(256) void k_constrain(
    int Np, int ncl, const int* __restrict__ cl_start, const int* __restrict__ cl_atoms,
    const int* __restrict__ con_start, const int* __restrict__ con_ij,
    const double* __restrict__ con_d2, double tol, const double* __restrict__ invm,
    const double* __restrict__ dxs, double* __restrict__ x64, double* __restrict__ v64,
    const DevRecord* __restrict__ rec, int* __restrict__ flags) {
    int r = blockIdx.y, q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncl) return;
    double* X = x64 + (size_t)r * 3 * Np;
    double* V = v64 + (size_t)r * 3 * Np;
    const double* D = dxs + (size_t)r * 3 * Np;
    int a0 = cl_start[q], na = cl_start[q + 1] - a0;
    int c0 = con_start[q], nc = con_start[q + 1] - c0;
    double xo[SHAKE_MAX_ATOMS][3], xn[SHAKE_MAX_ATOMS][3], im[SHAKE_MAX_ATOMS];
    for (int k = 0; k < na; k++) {
        int a = cl_atoms[a0 + k];
        im[k] = invm[a];
        for (int d = 0; d < 3; d++) {
            xo[k][d] = X[d * Np + a];
            xn[k][d] = xo[k][d] + D[d * Np + a];
        }
    }
    bool done = false;
    for (int it = 0; it < SHAKE_MAX_ITER && !done; it++) {
        done = true;
        for (int k = 0; k < nc; k++) {
            int i = con_ij[2 * (c0 + k)], j = con_ij[2 * (c0 + k) + 1];
            double d2 = con_d2[c0 + k];
            double ro[3], rn[3];
            for (int d = 0; d < 3; d++) {
                ro[d] = xo[i][d] - xo[j][d];
                rn[d] = xn[i][d] - xn[j][d];
            }
            double diff = d2 - dot3(rn, rn);
            if (fabs(diff) > tol * d2) done = false;
            double g = diff / (2.0 * (im[i] + im[j]) * dot3(ro, rn));
            for (int d = 0; d < 3; d++) {
                xn[i][d] += g * im[i] * ro[d];
                xn[j][d] -= g * im[j] * ro[d];
            }
        }
    }
    if (!done) atomicOr(&flags[r], 8);
    double inv = 1.0 / rec[r].dt;
    for (int k = 0; k < na; k++) {
        int a = cl_atoms[a0 + k];
        for (int d = 0; d < 3; d++) {
            V[d * Np + a] = inv * (xn[k][d] - xo[k][d]);
            X[d * Np + a] = xn[k][d];
        }
    }
}

// Kinetic energy 0.5 sum m v^2 of every replica (one workgroup per replica), with the
// velocities of constrained atoms first projected onto the constraints, as
// ReferenceSDMKernels.cpp:108-135 does (applyToVelocities on a copy; the stored
// velocities are not changed): per cluster, each constraint in turn removes the relative
// velocity of its two atoms along their current bond vector in inverse proportion to
// their masses (momentum is conserved), repeated until |r . dv| <= tol |r| (|v_i|+|v_j|)
// for every constraint.  The converged result is the mass-weighted projection onto
// r_c . (v_i - v_j) = 0, whatever solver reaches it.  Unconstrained atoms (cflag 0, or no
// clusters) are summed as they are.  Sets flag 8 if a cluster does not converge.
constexpr int KE_MAX_ITER = 500;
__global__ __launch_bounds__(256) void k_kinetic(
    int N, int Np, const double* __restrict__ mass, const double* __restrict__ invm,
    const double* __restrict__ x64, const double* __restrict__ v64, int ncl,
    const int* __restrict__ cl_start, const int* __restrict__ cl_atoms,
    const int* __restrict__ con_start, const int* __restrict__ con_ij,
    const int* __restrict__ cflag, double tol, int* __restrict__ flags,
    double* __restrict__ out) {
    __shared__ double red[256];
    int r = blockIdx.x, t = threadIdx.x;
    const double* V = v64 + (size_t)r * 3 * Np;
    const double* X = x64 + (size_t)r * 3 * Np;
    double e = 0.0;
    for (int i = t; i < N; i += 256) {
        double m = mass[i];
        if (m > 0 && !(ncl > 0 && cflag[i]))
            e += m * (V[i] * V[i] + V[Np + i] * V[Np + i] + V[2 * Np + i] * V[2 * Np + i]);
    }
    bool bad = false;
    for (int q = t; q < ncl; q += 256) {
        int a0 = cl_start[q], na = cl_start[q + 1] - a0;
        int c0 = con_start[q], nc = con_start[q + 1] - c0;
        double xv[SHAKE_MAX_ATOMS][3], vv[SHAKE_MAX_ATOMS][3], im[SHAKE_MAX_ATOMS];
        for (int k = 0; k < na; k++) {
            int a = cl_atoms[a0 + k];
            im[k] = invm[a];
            for (int d = 0; d < 3; d++) {
                xv[k][d] = X[d * Np + a];
                vv[k][d] = V[d * Np + a];
            }
        }
        bool done = false;
        for (int it = 0; it < KE_MAX_ITER && !done; it++) {
            done = true;
            for (int k = 0; k < nc; k++) {
                int i = con_ij[2 * (c0 + k)], j = con_ij[2 * (c0 + k) + 1];
                double rr[3], dv[3];
                for (int d = 0; d < 3; d++) {
                    rr[d] = xv[i][d] - xv[j][d];
                    dv[d] = vv[i][d] - vv[j][d];
                }
                double r2 = dot3(rr, rr), rv = dot3(rr, dv);
                double vs = sqrt(dot3(vv[i], vv[i])) + sqrt(dot3(vv[j], vv[j]));
                if (fabs(rv) > tol * sqrt(r2) * vs) done = false;
                double g = rv / (r2 * (im[i] + im[j]));
                for (int d = 0; d < 3; d++) {
                    vv[i][d] -= g * im[i] * rr[d];
                    vv[j][d] += g * im[j] * rr[d];
                }
            }
        }
        bad = bad || !done;
        for (int k = 0; k < na; k++) {
            double m = mass[cl_atoms[a0 + k]];
            e += m * dot3(vv[k], vv[k]);
        }
    }
    if (bad) atomicOr(&flags[r], 8);
    red[t] = e;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) out[r] = 0.5 * red[0];
}

// ---------------------------------------------------------------------------
// adjoint (parameter gradients of u over many frames).  Frames are [F][N][3] doubles in
// atom order; pair terms are recomputed from the positions, never stored.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_frame_atom(const double* __restrict__ X, int a, double& x,
                                                double& y, double& z) {
    x = X[3 * a];
    y = X[3 * a + 1];
    z = X[3 * a + 2];
}

// u of every frame: one workgroup per frame, pairs (displaced atom i, atom j) with
// d_i != d_j, displaced-displaced pairs once
__global__ __launch_bounds__(256) void k_u_frames(
    int N, int L, const int* __restrict__ lig, const int* __restrict__ cls,
    const double* __restrict__ cdispl, const double* __restrict__ q64,
    const double* __restrict__ sig64, const double* __restrict__ eps64,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list,
    const double* __restrict__ frames, double* __restrict__ u_out, Geom g) {
    __shared__ double red[256];
    int f = blockIdx.x, t = threadIdx.x;
    const double* X = frames + (size_t)f * 3 * N;
    double acc = 0.0;
    for (long long idx = t; idx < (long long)L * N; idx += 256) {
        int i = lig[idx / N];
        int j = (int)(idx % N);
        int ci = cls[i], cj = cls[j];
        if (ci == cj || (cj != 0 && j < i)) continue;
        if (excluded(ex_start, ex_list, i, j)) continue;
        Atom64 I, J;
        load_frame_atom(X, i, I.x, I.y, I.z);
        load_frame_atom(X, j, J.x, J.y, J.z);
        I.dx = cdispl[3 * ci]; I.dy = cdispl[3 * ci + 1]; I.dz = cdispl[3 * ci + 2];
        J.dx = cdispl[3 * cj]; J.dy = cdispl[3 * cj + 1]; J.dz = cdispl[3 * cj + 2];
        I.qk = q64[i]; I.sig = sig64[i]; I.se = sqrt(eps64[i]);
        J.qk = q64[j]; J.sig = sig64[j]; J.se = sqrt(eps64[j]);
        double fx = 0, fy = 0, fz = 0;
        int iny;
        acc += delta_pair(I, J, g, fx, fy, fz, &iny);
    }
    red[t] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) u_out[f] = red[0];
}

// du/d(q_s, sigma_s, eps_s) for selected atom s in frame f: one wave per (s, f)
__global__ __launch_bounds__(64) void k_du_dparams(
    int N, const int* __restrict__ sel, int nsel, const int* __restrict__ cls,
    const double* __restrict__ cdispl, const double* __restrict__ q64,
    const double* __restrict__ sig64, const double* __restrict__ eps64,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list,
    const double* __restrict__ frames, double* __restrict__ grad, Geom g) {
    int k = blockIdx.x, f = blockIdx.y, lane = threadIdx.x;
    int s = sel[k];
    const double* X = frames + (size_t)f * 3 * N;
    int cs = cls[s];
    double xs, ys, zs;
    load_frame_atom(X, s, xs, ys, zs);
    const double ske = sqrt(SDM_ONE_4PI_EPS0);
    double dq = 0.0, dsig = 0.0, deps = 0.0;
    for (int j = lane; j < N; j += 64) {
        int cj = cls[j];
        if (cj == cs || excluded(ex_start, ex_list, s, j)) continue;
        double xj, yj, zj;
        load_frame_atom(X, j, xj, yj, zj);
        double sig = g.geo ? sqrt(sig64[s] * sig64[j]) : 0.5 * (sig64[s] + sig64[j]);
        double eps = sqrt(eps64[s] * eps64[j]);
        // d sigma_sj / d sigma_s: 1/2, or sigma_j / (2 sigma_sj) for the geometric rule
        // (zero with sigma_sj = 0, where the LJ terms vanish)
        const double dsds = g.geo ? (sig > 0.0 ? sig64[j] / (2.0 * sig) : 0.0) : 0.5;
        const bool lj = !g.geo || sig > 0.0;
        for (int geo = 0; geo < 2; geo++) {  // geo 0: at x (sign -), geo 1: at x+d (sign +)
            double dx = min_image((xs + (geo ? cdispl[3 * cs] : 0.0)) - (xj + (geo ? cdispl[3 * cj] : 0.0)), g.Lx, g.iLx);
            double dy = min_image((ys + (geo ? cdispl[3 * cs + 1] : 0.0)) - (yj + (geo ? cdispl[3 * cj + 1] : 0.0)), g.Ly, g.iLy);
            double dz = min_image((zs + (geo ? cdispl[3 * cs + 2] : 0.0)) - (zj + (geo ? cdispl[3 * cj + 2] : 0.0)), g.Lz, g.iLz);
            double r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 < g.rc2)) continue;
            double sgn = geo ? 1.0 : -1.0;
            double ir = 1.0 / sqrt(r2);
            double s2 = sig * sig / r2;
            double s6 = s2 * s2 * s2;
            double s12 = s6 * s6;
            // q64 holds q*sqrt(k_e): d(k_e q_s q_j f)/dq_s = sqrt(k_e) * q64[j] * f
            dq += sgn * ske * q64[j] * (ir + g.krf * r2 - g.crf);
            if (lj) dsig += sgn * dsds * 4.0 * eps * (12.0 * s12 - 6.0 * s6) / sig;
            if (eps > 0.0) deps += sgn * 4.0 * (s12 - s6) * eps64[j] / (2.0 * eps);
        }
    }
    dq = wave_sum(dq);
    dsig = wave_sum(dsig);
    deps = wave_sum(deps);
    if (lane == 0) {
        double* o = grad + ((size_t)f * nsel + k) * 3;
        o[0] = dq;
        o[1] = dsig;
        o[2] = deps;
    }
}

// ---------------------------------------------------------------------------
// diagnostics: dense component forces (FP64) for sdm_compute_components.
// ---------------------------------------------------------------------------
__global__ void k_fill_ll(long long* __restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
__global__ void k_fill_d(double* __restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}
__global__ void k_fill_u64(unsigned long long* __restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// dF (dense, [R][3][Np]) from dB and dlig
__global__ __launch_bounds__(64) void k_scatter_dF(int N, int Np, int L, int maxc,
                                                   const int* __restrict__ lig,
                                                   const int* __restrict__ cls,
                                                   const int* __restrict__ orig_s,
                                                   const int* __restrict__ cand,
                                                   const int* __restrict__ ncand,
                                                   const double4* __restrict__ dB,
                                                   const double4* __restrict__ dlig,
                                                   double* __restrict__ dF) {
    int r = blockIdx.y, lane = threadIdx.x;
    double* D = dF + (size_t)r * 3 * Np;
    if ((int)blockIdx.x < maxc) {
        int c = blockIdx.x;
        if (c >= ncand[r]) return;
        int s = cand[(size_t)r * maxc + c] * 64 + lane;
        if (s >= N) return;
        int b = orig_s[(size_t)r * Np + s];
        if (cls[b] != 0) return;
        double4 d = dB[((size_t)r * maxc + c) * 64 + lane];
        D[b] = d.x;
        D[Np + b] = d.y;
        D[2 * Np + b] = d.z;
    } else {
        int li = (blockIdx.x - maxc) * 64 + lane;
        if (li >= L) return;
        int a = lig[li];
        double4 d = dlig[(size_t)r * L + li];
        D[a] = d.x;
        D[Np + a] = d.y;
        D[2 * Np + a] = d.z;
    }
}

}  // namespace sdm
