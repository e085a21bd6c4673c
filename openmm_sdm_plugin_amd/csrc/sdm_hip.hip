// sdm_hip.hip -- host runtime and C ABI (include/sdm_hip.h) of the SDM HIP path.
//
// A context owns every device buffer of R replicas and one HIP stream (default
// priority).  sdm_step replays a captured hipGraph of one step (all replicas in every
// launch); no memset nodes are captured (counters are cleared by kernels).
#include "sdm_hip.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "sdm_kernels.h"

using namespace sdm;

static_assert(offsetof(DevState, time) == sizeof(sdm_params), "DevState prefix must be sdm_params");
static_assert(offsetof(DevState, seed) == offsetof(sdm_params, seed), "layout");
static_assert(offsetof(DevState, work) == offsetof(sdm_params, work), "layout");
static_assert(offsetof(DevRecord, flags) == offsetof(sdm_energies, flags), "layout");

// Selection staging of sdm_du_dparams_frames_noalloc: a file-scope device array, so the
// entry point allocates nothing.  Each GPU holds its own copy at its own address, which
// host code must obtain with hipGetSymbolAddress (the array's name is a host-side
// stand-in, not a device address); it is resolved once per device, with that device
// current, by the first sdm_create there, and kept in g_scratch_ptr.
static __device__ __align__(256) int32_t scratch[SDM_NOALLOC_MAX_SEL];

namespace {
thread_local std::string g_err;
int g_registered = 0;
std::mutex g_scratch_mu;
std::vector<int32_t*> g_scratch_ptr;  // by device ordinal; nullptr until resolved

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// Device address of `scratch` on `device`, which must be the current device.
int resolve_scratch(int device, int32_t** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0) return fail(SDM_ERR_HIP, "bad device ordinal");
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(SDM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                         hipGetErrorString(e));
        g_scratch_ptr[device] = (int32_t*)ptr;
    }
    *out = g_scratch_ptr[device];
    return SDM_OK;
}
}  // namespace

#define HIPCHK(expr)                                                                    \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return fail(SDM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

#define CHK(expr)              \
    do {                       \
        int _rc = (expr);      \
        if (_rc != SDM_OK) return _rc; \
    } while (0)

struct sdm_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // side streams of the step's fork/join: bonded terms and the displaced-state chain
    // run concurrently with the pair kernel
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t ev_fork0 = nullptr, ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
    hipEvent_t ev_tiles = nullptr;  // the tile lists are built (energy evaluations)
    int N = 0, Np = 0, NB = 0, R = 0, L = 0, LB = 0, ncls = 0;
    // 0: CutoffNonPeriodic (no images; the sort grid spans reg_lo .. reg_hi, the union of
    // the positions' bounding boxes over the replicas, widened by GRID_MARGIN)
    int periodic = 1;
    std::vector<double> reg_box;  // [R][6] per replica lo, hi (empty entries: hi < lo)
    int NC = 0, ncx = 0, ncy = 0, ncz = 0, NC_alloc = 0;
    int maxj = 0, maxc = 0, nbond = 0, nang = 0, nbw = 1, max_ex = 0;
    Geom g{};
    float list_r2 = 0.f;
    // topology (shared by all replicas)
    float *qs = nullptr, *sh = nullptr, *se = nullptr;
    double *q64 = nullptr, *sig64 = nullptr, *eps64 = nullptr, *invm = nullptr, *mass = nullptr;
    double4* par64 = nullptr;  // [Np] (q sqrt(k_e), sigma/2, 2 sqrt(eps), 0): FP64 energies
    Rec64* rec64 = nullptr;    // [R][Np] slot-ordered FP64 records of the energy kernel
    int *ex_start = nullptr, *ex_list = nullptr, *cls = nullptr, *lig = nullptr, *cellrank = nullptr;
    // blocks of the displaced-atom list: entries lb_start[b] .. + lb_count[b] (<= 64),
    // all of one displacement class, so a block's bounding box covers one group of atoms
    int *lb_start = nullptr, *lb_count = nullptr;
    double* cdispl = nullptr;
    int* bond_atoms = nullptr;
    double *bond_r0 = nullptr, *bond_k = nullptr;
    int* ang_atoms = nullptr;
    double *ang_t0 = nullptr, *ang_k = nullptr;
    // per-atom lists of the bonds and angles each atom belongs to (k_bonded's gather):
    // entries (term << 1 | role) and (term << 2 | role) in term order
    int *bstart = nullptr, *bcode = nullptr, *astart = nullptr, *acode = nullptr;
    double* Fb = nullptr;  // [R][3][Np] FP64 bonded forces, atom order
    // per replica
    double *x64 = nullptr, *v64 = nullptr;
    float4* posq_s = nullptr;
    float2* par_s = nullptr;
    float4* arec = nullptr;  // [R][Np][2] slot records of the grouped pair kernel
    int *orig_s = nullptr, *slot_of = nullptr, *key = nullptr;
    int *ccount = nullptr, *cstart = nullptr, *cfill = nullptr;
    float4 *bb_lo = nullptr, *bb_hi = nullptr;
    int *tiles = nullptr, *ntiles = nullptr;
    unsigned long long* tmask = nullptr;
    int* cmap = nullptr;    // [R][NB] block -> candidate index of k_ligcand (-1: none)
    int* lig_of = nullptr;  // [Np] atom -> index in lig (-1: environment)
    float4* xb = nullptr;   // [R][Np] slot coordinates when the tile lists were built
    float* dmax = nullptr;  // [R][DISP_STRIDE] (first float used) largest squared
                            // displacement of the replica's atoms since then (k_refresh),
                            // or stale after a sort (k_bbox)
    long long* F = nullptr;
    double *e_pair = nullptr, *e_bond = nullptr, *u_part = nullptr;
    double* e64 = nullptr;  // [R][NB * E64_SPLIT] FP64 pair-energy slots
    // SDMUtils restraints (group 1): centroid and alignment terms, their group atoms and
    // normalized weights, and one energy slot per (replica, restraint)
    DevCentroid* rcen = nullptr;
    DevAlign* rali = nullptr;
    int* ratom = nullptr;
    double* rw = nullptr;
    double* e_rest = nullptr;
    int ncen = 0, nali = 0;
    // further force-field terms (torsions, impropers, positional restraints, 1-4
    // exceptions): device arrays behind `terms`, per-workgroup energies of groups 1 and 2
    DevTerms terms{};
    int nterm = 0, ntb = 0;
    std::vector<void*> term_bufs;
    double *e_t1 = nullptr, *e_t2 = nullptr;
    // constraints: clusters (atoms, and constraints as local index pairs with squared
    // lengths), the per-atom flag and the displacement scratch of the Langevin update
    int ncl = 0;
    int *cl_start = nullptr, *cl_atoms = nullptr, *con_start = nullptr, *con_ij = nullptr;
    int* cflag = nullptr;
    double *con_d2 = nullptr, *dxs = nullptr;
    double con_tol = 1e-10;
    int cm_every = 0;  // centre-of-mass motion removal interval (0: never)
    int *cand = nullptr, *ncand = nullptr;
    double4 *dA = nullptr, *dlig = nullptr, *dB = nullptr;
    // [R][maxc][L] 64-bit lists: the undisplaced slots of candidate block c that displaced
    // atom li pairs with (k_denv's screen, read by k_dlig)
    unsigned long long* mlig = nullptr;
    DevState* state = nullptr;
    DevRecord* rec = nullptr;
    int* flags = nullptr;
    double* uhist = nullptr;
    int hist_every = 0, hist_cap = 1;
    long long* Fb_sep = nullptr;
    // virtual sites (sdm_set_virtual_sites): kind, (site, p1, p2, p3), 3 weights each
    int nvs = 0;
    int* vs_kind = nullptr;
    int4* vs_at = nullptr;
    double* vs_w = nullptr;
    std::vector<int> h_vs_kind, h_vs_at;
    std::vector<double> h_vs_w;
    double* dFd = nullptr;
    unsigned long long* counters = nullptr;
    double* ke_out = nullptr;
    int* sel_dev = nullptr;
    size_t sel_cap = 0;
    // this device's `scratch` (selection staging of the allocation-free adjoint)
    int32_t* scratch_sel = nullptr;
    sdm_tuning tune{1, 0, 1, 1, 0, 20, 0, 4, 1, 0.f};
    // steps evaluated in the current slot order (0: stale, the next evaluation sorts)
    int since_sort = 0;
    // captured steps: [with a full sort][with the pair energies]
    hipGraph_t graph[2][2] = {};
    hipGraphExec_t gexec[2][2] = {};
    bool components_valid = false;
    std::vector<void*> allocs;
};

namespace {

template <class T>
int dalloc(sdm_context* c, T** p, size_t n) {
    if (n == 0) n = 1;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess)
        return fail(SDM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    // zero-filled on the context's own stream and waited for: the legacy null stream is
    // not ordered with the context's non-blocking streams, and a reused allocation holds
    // stale values (counters that must start at zero would give out-of-range offsets)
    e = hipMemsetAsync(q, 0, n * sizeof(T), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess)
        return fail(SDM_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    c->allocs.push_back(q);
    *p = (T*)q;
    return SDM_OK;
}

template <class T>
int upload(sdm_context* c, T* dst, const T* src, size_t n) {
    if (n == 0) return SDM_OK;
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SDM_OK;
}

template <class T>
int download(sdm_context* c, T* dst, const T* src, size_t n) {
    if (n == 0) return SDM_OK;
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SDM_OK;
}

// frees a buffer obtained from dalloc before the context ends
void dfree(sdm_context* c, void* p) {
    if (!p) return;
    auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
    if (it != c->allocs.end()) c->allocs.erase(it);
    (void)hipFree(p);
}

void drop_graph(sdm_context* c) {
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            if (c->gexec[a][b]) (void)hipGraphExecDestroy(c->gexec[a][b]);
            if (c->graph[a][b]) (void)hipGraphDestroy(c->graph[a][b]);
            c->gexec[a][b] = nullptr;
            c->graph[a][b] = nullptr;
        }
}

// Whether an evaluation sorts (true) or keeps the slot order of the last sort.  Steps
// sort every tune.sort_interval steps; other evaluations (energies, pair counts, kernel
// timing) sort only a stale order and leave the count alone, so the orders that a run
// of steps sees do not depend on them.
bool take_order(sdm_context* c, bool step) {
    bool full = c->since_sort == 0 || (step && c->since_sort >= c->tune.sort_interval);
    if (step) c->since_sort = full ? 1 : c->since_sort + 1;
    return full;
}

// Hilbert index of a cell (Skilling, "Programming the Hilbert curve", AIP Conf. Proc.
// 707, 381 (2004): axes -> transposed index, then bit interleave).  Hilbert-ordered
// 64-slot blocks are near-cubic, so tile boxes are tight.
uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits) {
    uint32_t X[3] = {x, y, z};
    uint32_t M = 1u << (bits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        uint32_t P = Q - 1;
        for (int i = 0; i < 3; i++) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    for (int i = 0; i < 3; i++) X[i] ^= t;
    uint64_t idx = 0;
    for (int b = bits - 1; b >= 0; b--)
        for (int i = 0; i < 3; i++) idx = (idx << 1) | ((X[i] >> b) & 1u);
    return idx;
}

// cell grid + Hilbert rank of every cell (the slot order of the spatial sort).
constexpr double GRID_MARGIN = 0.5;  // nm around the positions of a non-periodic system

// extent of the sort grid: the box, or (non-periodic) the region around the positions
void grid_region(const sdm_context* c, double lo[3], double hi[3]) {
    if (c->periodic) {
        lo[0] = lo[1] = lo[2] = 0.0;
        hi[0] = c->g.Lx;
        hi[1] = c->g.Ly;
        hi[2] = c->g.Lz;
        return;
    }
    for (int d = 0; d < 3; d++) {
        lo[d] = 1e300;
        hi[d] = -1e300;
    }
    for (int r = 0; r < c->R; r++) {
        const double* b = &c->reg_box[6 * (size_t)r];
        if (!(b[0] <= b[3])) continue;  // no positions yet
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], b[d]);
            hi[d] = std::max(hi[d], b[3 + d]);
        }
    }
    for (int d = 0; d < 3; d++) {
        if (!(lo[d] <= hi[d])) lo[d] = hi[d] = 0.0;
        lo[d] -= GRID_MARGIN;
        hi[d] += GRID_MARGIN;
    }
}

int setup_cells(sdm_context* c, int per_dim) {
    double lo[3], hi[3];
    grid_region(c, lo, hi);
    double L[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    int n[3];
    for (int d = 0; d < 3; d++) {
        int v = per_dim > 0 ? per_dim : (int)std::floor(L[d] / 0.3);
        n[d] = std::max(1, std::min(v, 1024));
        c->g.go[d] = lo[d];
        c->g.gs[d] = n[d] / L[d];
    }
    int NC = n[0] * n[1] * n[2];
    int bits = 1;
    while ((1 << bits) < std::max(n[0], std::max(n[1], n[2]))) bits++;
    std::vector<std::pair<uint64_t, int>> code(NC);
    for (int x = 0; x < n[0]; x++)
        for (int y = 0; y < n[1]; y++)
            for (int z = 0; z < n[2]; z++) {
                int cell = (x * n[1] + y) * n[2] + z;
                code[cell] = {hilbert3((uint32_t)x, (uint32_t)y, (uint32_t)z, bits), cell};
            }
    std::sort(code.begin(), code.end());
    std::vector<int> rank(NC);
    for (int k = 0; k < NC; k++) rank[code[k].second] = k;
    if (NC > c->NC_alloc) {
        HIPCHK(hipStreamSynchronize(c->stream));
        for (void* p : {(void*)c->cellrank, (void*)c->ccount, (void*)c->cstart, (void*)c->cfill})
            dfree(c, p);
        CHK(dalloc(c, &c->cellrank, NC));
        CHK(dalloc(c, &c->ccount, (size_t)c->R * NC));
        CHK(dalloc(c, &c->cstart, (size_t)c->R * (NC + 1)));
        CHK(dalloc(c, &c->cfill, (size_t)c->R * NC));
        c->NC_alloc = NC;
    } else {
        // counters must start at zero for the new grid
        HIPCHK(hipMemsetAsync(c->ccount, 0, sizeof(int) * (size_t)c->R * c->NC_alloc, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    CHK(upload(c, c->cellrank, rank.data(), NC));
    c->ncx = n[0];
    c->ncy = n[1];
    c->ncz = n[2];
    c->NC = NC;
    drop_graph(c);
    return SDM_OK;
}

dim3 g2(int n, int per, int R) { return dim3((unsigned)((n + per - 1) / per), (unsigned)R); }

// Tile lists are kept over steps (tuning.list_skin > 0) with the block pair kernels only:
// there every slot refresh is followed by a k_tiles launch, which rebuilds after a sort,
// so the build coordinates (xb) always belong to the lists in use.  The grouped kernel
// builds lists on energy evaluations only; they are built every time.
bool lists_kept(const sdm_context* c) { return c->tune.list_skin > 0.f && c->tune.pair_group == 0; }

// slot order (a full spatial sort, or the last sort's order refreshed) and block boxes
// for every replica
// energy: also the slot-ordered FP64 records of the energy kernel
void enqueue_order(sdm_context* c, bool full, bool energy = false) {
    hipStream_t s = c->stream;
    int N = c->N, Np = c->Np, R = c->R;
    Rec64* rec = energy ? c->rec64 : nullptr;
    if (!full) {
        k_refresh<<<dim3(c->NB, R), 64, 0, s>>>(N, Np, c->NB, c->x64, c->orig_s, c->qs, c->g,
                                                c->posq_s, c->arec, c->F, c->bb_lo, c->bb_hi,
                                                c->par64, rec,
                                                lists_kept(c) ? c->xb : nullptr,
                                                c->dmax);
    } else {
    k_bin<<<g2(N, 256, R), 256, 0, s>>>(N, Np, c->x64, c->key, c->ccount, c->cellrank, c->g,
                                        c->ncx, c->ncy, c->ncz, c->NC);
    k_scan<<<R, 1024, 0, s>>>(c->NC, c->ccount, c->cstart, c->cfill);
    k_scatter<<<g2(N, 256, R), 256, 0, s>>>(N, Np, c->NC, c->key, c->cstart, c->cfill, c->orig_s);
    k_cellsort<<<g2(c->NC, 128, R), 128, 0, s>>>(N, Np, c->NC, c->cstart, c->orig_s, c->slot_of,
                                                c->x64, c->qs, c->sh, c->se, c->posq_s,
                                                c->par_s, c->arec, c->F, c->g, c->par64, rec);
    k_bbox<<<dim3(c->NB, R), 64, 0, s>>>(N, Np, c->NB, c->posq_s, c->bb_lo, c->bb_hi, c->dmax);
    }
}

// whether a launch of one workgroup per (replica, 64-atom block) leaves the GPU short of
// work (fewer than 8 workgroups per CU)
bool few_blocks(const sdm_context* c) { return (long)c->NB * c->R < 2048; }

// build radius and reuse threshold of the tile lists for a skin of tuning.list_skin nm
TileRadius tile_radius(const sdm_context* c) {
    const float skin = lists_kept(c) ? c->tune.list_skin : 0.f;
    const float rl = std::sqrt(c->list_r2) + skin;
    TileRadius tr;
    tr.r2 = rl * rl;
    tr.oix = c->g.oix - skin;
    tr.oiy = c->g.oiy - skin;
    tr.oiz = c->g.oiz - skin;
    tr.reuse_d2 = skin > 0.f ? 0.25f * skin * skin : -1.f;
    return tr;
}

// tile lists of the block pair kernels and of the FP64 energy kernel (the grouped
// kernel finds its j-blocks itself: with it, lists are built for energy evaluations only,
// as half lists).  With a skin, k_tiles keeps a replica's lists while its atoms have
// moved less than half the skin since they were built.
void enqueue_tiles(sdm_context* c, bool energy) {
    if (c->tune.pair_group != 0 && !energy) return;
    const int half = c->tune.pair_group != 0 ? 1 : c->tune.pair_newton3;
    const TileRadius tr = tile_radius(c);
    // One workgroup per (replica, i-block).  Its time is the serial path through it, so
    // when the workgroups do not fill the GPU (fewer than 8 per CU) each gets 8 waves for
    // the candidate masks, otherwise 4 (measured: one replica of 22,750 atoms +6%, 22
    // replicas -1% with 8).
#define SDM_TILES(NW)                                                                         \
    k_tiles<NW><<<dim3(c->NB, c->R), 64 * NW, 0, c->stream>>>(                                \
        c->N, c->Np, c->NB, c->maxj, c->posq_s, c->bb_lo, c->bb_hi, c->tiles, c->tmask,       \
        c->ntiles, c->dmax, c->xb, tr, c->g, half, c->flags)
    if (few_blocks(c)) SDM_TILES(8);
    else SDM_TILES(4);
#undef SDM_TILES
}

// slot order, block boxes and tile lists
void enqueue_sort(sdm_context* c, bool full) {
    enqueue_order(c, full);
    enqueue_tiles(c, false);
}

// FP64 pair energies into e64 (needs the tile lists and FP64 records of this evaluation)
void enqueue_energy64(sdm_context* c, hipStream_t st) {
    const int half = c->tune.pair_group != 0 ? 1 : c->tune.pair_newton3;
    unsigned waves = (unsigned)(c->R * c->NB * E64_SPLIT);
    size_t lds = (size_t)std::max(1, c->max_ex) * 64 * sizeof(int);
    k_energy64<<<waves, 64, lds, st>>>(c->R, c->N, c->Np, c->NB, c->maxj, E64_SPLIT,
                                              half, c->arec, c->rec64, c->orig_s, c->slot_of,
                                              c->tiles, c->tmask, c->ntiles, c->bb_lo,
                                              c->bb_hi, c->ex_start, c->ex_list, c->max_ex,
                                              c->e64, c->g);
}

// ring capacity (entries per group) of the grouped pair kernel
template <int G>
constexpr int grp_cap() {
    return G == 16 ? 256 : 128;
}

template <int G, int W>
void launch_pair_grp(sdm_context* c, bool count, unsigned long long* pc) {
    constexpr int CAP = grp_cap<G>();
    int waves = c->R * c->NB;
    dim3 gr((waves + W - 1) / W);
    // per wave: rings, exclusion partners, group boxes (the staging buffers are static)
    size_t lds = (size_t)W * ((64 / G) * (CAP + 8) + c->max_ex * 64) * sizeof(unsigned);
#define SDM_GRP_ARGS                                                                            \
    c->R, c->N, c->Np, c->NB, c->arec, c->orig_s, c->slot_of, c->bb_lo, c->bb_hi, c->ex_start, \
        c->ex_list, c->max_ex, c->F, c->e_pair, c->flags, pc, c->list_r2, c->g
    if (count) k_pair_grp<G, CAP, W, true><<<gr, 64 * W, lds, c->stream>>>(SDM_GRP_ARGS);
    else k_pair_grp<G, CAP, W, false><<<gr, 64 * W, lds, c->stream>>>(SDM_GRP_ARGS);
#undef SDM_GRP_ARGS
}

template <int G>
void launch_pair_grp_w(sdm_context* c, bool count, unsigned long long* pc) {
    switch (c->tune.pair_waves_per_group) {
        case 2: launch_pair_grp<G, 2>(c, count, pc); break;
        case 4: launch_pair_grp<G, 4>(c, count, pc); break;
        default: launch_pair_grp<G, 1>(c, count, pc); break;
    }
}

// energy = false (block Newton-3 kernel only): forces without the pair energies
void enqueue_pair(sdm_context* c, bool count, bool energy = true) {
    hipStream_t s = c->stream;
    int waves = c->R * c->NB;
    unsigned long long* pc = count ? c->counters : nullptr;
    if (c->tune.pair_group == 16) return launch_pair_grp_w<16>(c, count, pc);
    if (c->tune.pair_group == 8) return launch_pair_grp_w<8>(c, count, pc);
    if (c->tune.pair_newton3) {
        size_t lds = (size_t)c->tune.pair_waves_per_group * c->max_ex * 64 * sizeof(int);
        const int split = c->tune.pair_split;
        waves *= split;
#define SDM_N3_ARGS                                                                             \
    c->R, c->N, c->Np, c->NB, c->maxj, split, c->arec, c->orig_s, c->slot_of, c->tiles,         \
        c->tmask,                                                                           \
        c->ntiles, c->bb_lo, c->bb_hi, c->ex_start, c->ex_list, c->max_ex, c->F, c->e_pair, \
        c->flags, pc, c->g
#define SDM_LAUNCH_N3_G(W, GEO)                                                               \
    do {                                                                                      \
        dim3 gr((waves + W - 1) / W);                                                         \
        if (count) k_pair_n3<W, true, true, GEO><<<gr, 64 * W, lds, s>>>(SDM_N3_ARGS);       \
        else if (energy) k_pair_n3<W, false, true, GEO><<<gr, 64 * W, lds, s>>>(SDM_N3_ARGS); \
        else k_pair_n3<W, false, false, GEO><<<gr, 64 * W, lds, s>>>(SDM_N3_ARGS);           \
    } while (0)
#define SDM_LAUNCH_N3(W)                                                                      \
    do {                                                                                      \
        if (c->g.geo) SDM_LAUNCH_N3_G(W, true);                                               \
        else SDM_LAUNCH_N3_G(W, false);                                                       \
    } while (0)
        switch (c->tune.pair_waves_per_group) {
            case 2: SDM_LAUNCH_N3(2); break;
            case 4: SDM_LAUNCH_N3(4); break;
            default: SDM_LAUNCH_N3(1); break;
        }
#undef SDM_LAUNCH_N3
#undef SDM_LAUNCH_N3_G
#undef SDM_N3_ARGS
        return;
    }
#define SDM_PAIR_ARGS                                                                       \
    c->R, c->N, c->Np, c->NB, c->maxj, c->posq_s, c->par_s, c->orig_s, c->slot_of, c->tiles, \
        c->tmask, c->ntiles, c->bb_lo, c->bb_hi, c->ex_start, c->ex_list, c->max_ex, c->F, c->e_pair,   \
        c->flags, pc, c->g
#define SDM_LAUNCH_PAIR(W)                                                                   \
    do {                                                                                     \
        dim3 gr((waves + W - 1) / W);                                                        \
        if (newton) {                                                                        \
            if (count) k_pair<W, true, true><<<gr, 64 * W, lds, s>>>(SDM_PAIR_ARGS);           \
            else k_pair<W, false, true><<<gr, 64 * W, lds, s>>>(SDM_PAIR_ARGS);                \
        } else {                                                                             \
            if (count) k_pair<W, true, false><<<gr, 64 * W, lds, s>>>(SDM_PAIR_ARGS);          \
            else k_pair<W, false, false><<<gr, 64 * W, lds, s>>>(SDM_PAIR_ARGS);               \
        }                                                                                    \
    } while (0)
    bool newton = c->tune.pair_newton3 != 0;
    size_t lds = (size_t)c->tune.pair_waves_per_group * std::max(1, c->max_ex) * 64 * sizeof(int);
    switch (c->tune.pair_waves_per_group) {
        case 2: SDM_LAUNCH_PAIR(2); break;
        case 4: SDM_LAUNCH_PAIR(4); break;
        default: SDM_LAUNCH_PAIR(1); break;
    }
#undef SDM_PAIR_ARGS
#undef SDM_LAUNCH_PAIR
}

// A step's Langevin kernel adds the displaced-state forces itself (MixGather) instead of a
// k_mix launch, except with virtual sites, whose forces must hold them before
// k_vsite_forces moves them to the parents (measured on 22,750 atoms: one replica 0.104
// against 0.108 ms per step launched directly, 0.109 against 0.112 replayed; 22 replicas
// unchanged within 0.2%).
bool mix_in_langevin(const sdm_context* c, bool advance) {
    return advance && c->L > 0 && c->nvs == 0;
}

// evaluation of one step (everything before the Langevin update).  After the slot order
// and the tile lists the work forks: the pair kernel on the main stream, the bonded terms
// and the displaced-state chain on two side streams; they join before the scalar stage.
// (Forking before the tile lists, which the side work does not need, was measured
// slower on 22 replicas, where the side kernels then slow the tile builder, which is on
// the critical path, and no faster on one.)  pe0/pe1 (optional) are recorded around the
// pair kernel.
void enqueue_eval(sdm_context* c, bool advance, bool components, bool full, bool energy,
                  hipEvent_t pe0 = nullptr, hipEvent_t pe1 = nullptr) {
    // tuning.side_fork: the side work (restraints, force-field terms, the displaced-state
    // chain) needs the slot order, block boxes and cleared accumulator, not the tile
    // lists.  0: it forks onto the two side streams after the tile lists, 1: before them,
    // 2: it runs on the main stream between the tile lists and the pair kernel (no fork;
    // while the pair kernel fills the GPU, side kernels beside it wait for free wave
    // slots and can outlast it), 3: as 1, but the displaced-state chain joins the main
    // stream before the pair kernel (it runs beside the tile lists only), 4: as 1, but the
    // bonded terms fork with the rest (no event at the start of the step).  The bonded
    // terms need only the positions and write their own FP64 array: with 0, 1 and 3 they
    // start with the step, beside the sort (measured: 22 replicas 0.7% faster than 4; one
    // replica 1.6% slower replayed from a graph, 4% launched directly).
    const int mode = c->tune.side_fork;
    hipStream_t s = c->stream;
    hipStream_t s1 = mode == 2 ? s : c->side[0], s2 = mode == 2 ? s : c->side[1];
    int N = c->N, Np = c->Np, R = c->R;
    auto bonded = [&]() {
        if (c->nbond + c->nang > 0)
            k_bonded<<<dim3(c->nbw, R), 256, 0, s1>>>(
                N, Np, c->bstart, c->bcode, c->astart, c->acode, c->bond_atoms, c->bond_r0,
                c->bond_k, c->ang_atoms, c->ang_t0, c->ang_k, c->x64, c->Fb, c->e_bond,
                c->nbw, c->flags);
    };
    if (mode != 2 && mode != 4) {
        (void)hipEventRecord(c->ev_fork0, s);
        (void)hipStreamWaitEvent(s1, c->ev_fork0, 0);
        bonded();
    }
    enqueue_order(c, full, energy);
    const bool tiles_after_fork = mode == 1 || mode == 3 || mode == 4;
    if (!tiles_after_fork) enqueue_tiles(c, energy);
    if (mode != 2) {
        (void)hipEventRecord(c->ev_fork, s);
        (void)hipStreamWaitEvent(s1, c->ev_fork, 0);
        (void)hipStreamWaitEvent(s2, c->ev_fork, 0);
        if (mode == 4) bonded();
    }
    if (tiles_after_fork) enqueue_tiles(c, energy);
    if (mode == 2) bonded();
    if (c->ncen + c->nali > 0)
        k_restraints<<<dim3(c->ncen + c->nali, R), 64, 0, s1>>>(
            Np, c->ncen, c->ncen + c->nali, c->rcen, c->rali, c->ratom, c->rw, c->x64,
            components ? nullptr : c->slot_of, components ? c->Fb_sep : c->F, c->e_rest,
            c->flags);
    if (c->nterm > 0)
        k_terms<<<dim3(c->ntb, R), 256, 0, s1>>>(
            Np, c->terms, c->x64, components ? nullptr : c->slot_of,
            components ? c->Fb_sep : c->F, c->slot_of, c->F, c->e_t1, c->e_t2, c->ntb, c->g,
            c->flags);
    if (c->L > 0) {
        k_ligcand<<<R, LIGCAND_THREADS, 0, s2>>>(Np, c->NB, c->L, c->ncls, c->lig, c->cls, c->cdispl, c->x64,
                                    c->bb_lo, c->bb_hi, c->maxc, c->cand, c->ncand, c->list_r2,
                                    c->g, c->flags, c->cmap);
        k_denv<DPAIR_WAVES><<<dim3(c->maxc, R), 64 * DPAIR_WAVES, 0, s2>>>(
            N, Np, c->L, c->LB, c->lig, c->lb_start, c->lb_count, c->cls, c->cdispl, c->q64,
            c->sig64, c->eps64, c->ex_start, c->ex_list, c->x64, c->orig_s, c->cand, c->ncand,
            c->maxc, c->mlig, c->dB, c->g);
        k_dlig<DPAIR_WAVES><<<dim3(c->maxc, c->LB, R), 64 * DPAIR_WAVES, 0, s2>>>(
            N, Np, c->L, c->LB, c->lig, c->lb_start, c->lb_count, c->cls, c->cdispl, c->q64,
            c->sig64, c->eps64, c->x64, c->orig_s, c->cand, c->ncand, c->maxc, c->mlig, c->dA,
            c->g, nullptr);
        k_deltaA_red<DRED_WAVES><<<dim3(c->LB, R), 64 * DRED_WAVES, 0, s2>>>(
            Np, c->L, c->LB, c->lig, c->lb_start, c->lb_count, c->cls, c->cdispl, c->q64,
            c->sig64, c->eps64, c->ex_start, c->ex_list, c->x64, c->ncand, c->maxc, c->dA,
            c->dlig, c->u_part, c->g, nullptr);
    }
    // the FP64 pair energies need the tile lists and this evaluation's FP64 records, not
    // the pair forces: they run beside the pair kernel on the first side stream, which
    // joins before the scalar stage
    if (energy && mode != 2) {
        (void)hipEventRecord(c->ev_tiles, s);
        (void)hipStreamWaitEvent(s1, c->ev_tiles, 0);
        enqueue_energy64(c, s1);
    }
    if (mode == 3) {
        (void)hipEventRecord(c->ev_join[1], s2);
        (void)hipStreamWaitEvent(s, c->ev_join[1], 0);
    }
    if (pe0) (void)hipEventRecord(pe0, s);
    // forces in FP32; the pair energies, when asked for, in FP64 by their own kernel
    enqueue_pair(c, false, false);
    if (pe1) (void)hipEventRecord(pe1, s);
    if (energy && mode == 2) enqueue_energy64(c, s);
    if (mode != 2) {
        (void)hipEventRecord(c->ev_join[0], s1);
        (void)hipStreamWaitEvent(s, c->ev_join[0], 0);
    }
    if (mode == 0 || mode == 1 || mode == 4) {
        (void)hipEventRecord(c->ev_join[1], s2);
        (void)hipStreamWaitEvent(s, c->ev_join[1], 0);
    }
    k_scalar<<<R, 256, 0, s>>>(c->NB * E64_SPLIT, c->nbw, c->LB, c->e64, c->e_bond, c->u_part, c->state,
                               c->rec, c->flags, advance ? 1 : 0, energy ? 1 : 0,
                               c->hist_every, c->hist_cap, c->uhist, c->e_rest,
                               c->ncen + c->nali, c->e_t1, c->e_t2, c->ntb, c->dmax);
    if (c->L > 0 && !components && !mix_in_langevin(c, advance))
        k_mix<<<dim3(c->maxc + c->LB, R), 64, 0, s>>>(N, Np, c->L, c->maxc, c->lig, c->cls,
                                                       c->orig_s, c->cand, c->ncand, c->dB,
                                                       c->dlig, c->rec, c->slot_of, c->F, c->flags);
}

void enqueue_step(sdm_context* c, bool full, bool energy, hipEvent_t pe0 = nullptr,
                  hipEvent_t pe1 = nullptr) {
    enqueue_eval(c, true, false, full, energy, pe0, pe1);
    if (c->cm_every > 0)
        k_cm_remove<<<c->R, 1024, 0, c->stream>>>(c->N, c->Np, c->cm_every, c->mass, c->v64,
                                                   c->state);
    if (c->nvs > 0)  // forces on virtual sites to their parents
        k_vsite_forces<<<g2(c->nvs, 256, c->R), 256, 0, c->stream>>>(
            c->nvs, c->Np, c->vs_kind, c->vs_at, c->vs_w, c->x64, c->slot_of, c->F, c->Fb,
            c->flags);
    MixGather mg{};
    if (mix_in_langevin(c, true))
        mg = MixGather{c->lig_of, c->cmap, c->dB, c->dlig, c->NB, c->maxc, c->L};
    k_langevin<<<g2(c->N, 256, c->R), 256, 0, c->stream>>>(c->N, c->Np, c->invm, c->F, c->Fb,
                                                           c->x64, c->v64, c->rec, c->state,
                                                           c->slot_of, c->cflag, c->dxs, mg,
                                                           c->flags);
    if (c->ncl > 0)
        k_constrain<<<g2(c->ncl, 256, c->R), 256, 0, c->stream>>>(
            c->Np, c->ncl, c->cl_start, c->cl_atoms, c->con_start, c->con_ij, c->con_d2,
            c->con_tol, c->invm, c->dxs, c->x64, c->v64, c->rec, c->flags);
    if (c->nvs > 0)  // the sites placed from their parents' new positions
        k_vsite_pos<<<g2(c->nvs, 256, c->R), 256, 0, c->stream>>>(c->nvs, c->Np, c->vs_kind,
                                                                  c->vs_at, c->vs_w, c->x64);
}

int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SDM_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return SDM_OK;
}

int build_graph(sdm_context* c, bool full, bool energy) {
    hipGraph_t& graph = c->graph[full][energy];
    hipGraphExec_t& gexec = c->gexec[full][energy];
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(c, full, energy);
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (e != hipSuccess) return fail(SDM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    return SDM_OK;
}

int check_replica(const sdm_context* c, int r, bool allow_all) {
    if (!c) return fail(SDM_ERR_INVALID, "null context");
    if (r == -1 && allow_all) return SDM_OK;
    if (r < 0 || r >= c->R) return fail(SDM_ERR_INVALID, "replica index out of range");
    return SDM_OK;
}

int validate_params(const sdm_params* p) {
    if (!p) return fail(SDM_ERR_INVALID, "null parameters");
    if (p->softcore_method < 0 || p->softcore_method > 2)
        return fail(SDM_ERR_INVALID, "Unknown soft core method");
    if (!(p->step_size > 0)) return fail(SDM_ERR_INVALID, "step size must be positive");
    if (p->friction < 0) return fail(SDM_ERR_INVALID, "friction must be >= 0");
    if (p->temperature < 0) return fail(SDM_ERR_INVALID, "temperature must be >= 0");
    return SDM_OK;
}

// constructor defaults of LangevinIntegratorSDM (LangevinIntegratorSDM.cpp:48-85)
sdm_params default_params() {
    sdm_params p;
    std::memset(&p, 0, sizeof(p));
    p.bias_method = SDM_LINEAR_METHOD;
    p.softcore_method = SDM_NO_SOFTCORE_METHOD;
    p.lambda = 1.0;
    p.gamma = 0.0;
    p.wbcoeff = 1.0;
    p.w0coeff = 0.0;
    p.lambda1 = 1.0;
    p.lambda2 = 1.0;
    p.alpha = 1.0;
    p.u0 = 0.0;
    p.umax = 200.0;
    p.acore = 0.25;
    p.ubcore = 0.0;
    p.nonequilibrium = 0;
    p.noneq_tmax = 1.0;
    p.work = 0.0;
    p.temperature = 300.0;
    p.friction = 1.0;
    p.step_size = 0.001;
    p.seed = 0;
    return p;
}

// Capacity of one i-block's tile list (the j-blocks whose box lies within the list
// radius of its box), so that tile storage grows as N rather than N^2.  At the mean
// density a block of 64 atoms fills v_b = 64 V / N; a generous bound on its reach from its
// centre is 1.5 (3 v_b / 4 pi)^(1/3) (Hilbert-ordered blocks are near-cubic), so candidate
// centres lie within r_l + 2 of those of the i-block's, a sphere that holds NB V_s / V
// blocks on average.  Three times that, for density fluctuations, at least 64 and at most
// NB (a list that would exceed it sets SDM_FLAG_CAPACITY).  Non-periodic systems: NB.
int tile_capacity(const sdm_system_desc* sys, int NB) {
    if (!sys->periodic) return NB;
    const double pi = 3.14159265358979323846;
    double V = sys->box[0] * sys->box[1] * sys->box[2];
    double vb = 64.0 * V / sys->n_atoms;
    double reach = sys->cutoff + 1e-3 + SDM_MAX_LIST_SKIN + 2.0 * 1.5 * std::cbrt(3.0 * vb / (4.0 * pi));
    double nb = NB * (4.0 / 3.0) * pi * reach * reach * reach / V;
    return std::max(1, std::min(NB, std::max(64, (int)std::ceil(3.0 * nb))));
}

int upload_nb_params(sdm_context* c, const double* q, const double* sig, const double* eps) {
    int N = c->N, Np = c->Np;
    std::vector<float> qs(Np, 0.f), sh(Np, 0.f), se(Np, 0.f);
    std::vector<double> q64(Np, 0.0), s64(Np, 0.0), e64(Np, 0.0);
    std::vector<double4> p64(Np, double4{0.0, 0.0, 0.0, 0.0});
    double sk = std::sqrt(SDM_ONE_4PI_EPS0);
    for (int i = 0; i < N; i++) {
        if (!(sig[i] >= 0) || !(eps[i] >= 0) || !std::isfinite(q[i]))
            return fail(SDM_ERR_INVALID, "nonbonded parameters must be finite, sigma/epsilon >= 0");
        qs[i] = (float)(q[i] * sk);
        const double sh64 = c->g.geo ? sig[i] : 0.5 * sig[i];  // the entry each rule combines
        sh[i] = (float)sh64;
        se[i] = (float)(2.0 * std::sqrt(eps[i]));
        q64[i] = q[i] * sk;
        s64[i] = sig[i];
        e64[i] = eps[i];
        p64[i] = double4{q[i] * sk, sh64, 2.0 * std::sqrt(eps[i]), 0.0};
    }
    CHK(upload(c, c->qs, qs.data(), Np));
    CHK(upload(c, c->sh, sh.data(), Np));
    CHK(upload(c, c->se, se.data(), Np));
    CHK(upload(c, c->q64, q64.data(), Np));
    CHK(upload(c, c->sig64, s64.data(), Np));
    CHK(upload(c, c->eps64, e64.data(), Np));
    CHK(upload(c, c->par64, p64.data(), Np));
    return SDM_OK;
}

}  // namespace

extern "C" {

const char* sdm_last_error(void) { return g_err.c_str(); }

int sdm_create(const sdm_system_desc* sys, int32_t n_replicas, sdm_context** out) {
    g_err.clear();
    if (!sys || !out) return fail(SDM_ERR_INVALID, "null argument");
    *out = nullptr;
    int N = sys->n_atoms;
    if (N <= 0) return fail(SDM_ERR_INVALID, "n_atoms must be positive");
    // the grouped pair kernel addresses 32-byte slot records with 32-bit byte offsets
    if (N >= (1 << 26)) return fail(SDM_ERR_UNSUPPORTED, "n_atoms must be below 2^26");
    if (n_replicas <= 0) return fail(SDM_ERR_INVALID, "n_replicas must be positive");
    if (!sys->charge || !sys->sigma || !sys->epsilon || !sys->mass || !sys->displacement)
        return fail(SDM_ERR_INVALID, "per-atom arrays must not be null");
    if (!(sys->cutoff > 0)) return fail(SDM_ERR_INVALID, "cutoff must be positive");
    if (sys->periodic)
        for (int d = 0; d < 3; d++)
            if (!(sys->box[d] >= 2.0 * sys->cutoff))
                return fail(SDM_ERR_INVALID, "the cutoff distance cannot be greater than half the periodic box size");
    if (!(sys->eps_rf > 0)) return fail(SDM_ERR_INVALID, "eps_rf must be positive");
    if (sys->lj_rule != SDM_LJ_ARITHMETIC && sys->lj_rule != SDM_LJ_GEOMETRIC)
        return fail(SDM_ERR_INVALID, "lj_rule must be SDM_LJ_ARITHMETIC or SDM_LJ_GEOMETRIC");
    for (int i = 0; i < N; i++)
        if (!(sys->mass[i] >= 0)) return fail(SDM_ERR_INVALID, "masses must be >= 0");
    // exclusions -> CSR
    std::vector<int> ecount(N + 1, 0);
    for (int k = 0; k < sys->n_exclusions; k++) {
        int a = sys->exclusions[2 * k], b = sys->exclusions[2 * k + 1];
        if (a < 0 || a >= N || b < 0 || b >= N || a == b)
            return fail(SDM_ERR_INVALID, "exclusion index out of range");
        ecount[a + 1]++;
        ecount[b + 1]++;
    }
    int max_ex = 0;
    for (int i = 0; i < N; i++) {
        if (ecount[i + 1] > SDM_MAX_EXCL)
            return fail(SDM_ERR_UNSUPPORTED, "more than 32 exclusions on one atom");
        max_ex = std::max(max_ex, ecount[i + 1]);
        ecount[i + 1] += ecount[i];
    }
    std::vector<int> elist(std::max(1, ecount[N])), efill(N, 0);
    for (int k = 0; k < sys->n_exclusions; k++) {
        int a = sys->exclusions[2 * k], b = sys->exclusions[2 * k + 1];
        elist[ecount[a] + efill[a]++] = b;
        elist[ecount[b] + efill[b]++] = a;
    }
    if (sys->n_bonds >= (1 << 28) || sys->n_angles >= (1 << 28))
        return fail(SDM_ERR_INVALID, "at most 2^28 - 1 bonds and 2^28 - 1 angles");
    for (int k = 0; k < sys->n_bonds; k++)
        for (int j = 0; j < 2; j++)
            if (sys->bond_atoms[2 * k + j] < 0 || sys->bond_atoms[2 * k + j] >= N)
                return fail(SDM_ERR_INVALID, "bond index out of range");
    for (int k = 0; k < sys->n_angles; k++)
        for (int j = 0; j < 3; j++)
            if (sys->angle_atoms[3 * k + j] < 0 || sys->angle_atoms[3 * k + j] >= N)
                return fail(SDM_ERR_INVALID, "angle index out of range");
    // displacement classes
    std::vector<int> cls(N, 0);
    std::vector<double> cd(3, 0.0);
    std::vector<int> lig;
    for (int i = 0; i < N; i++) {
        const double* d = sys->displacement + 3 * i;
        if (!std::isfinite(d[0]) || !std::isfinite(d[1]) || !std::isfinite(d[2]))
            return fail(SDM_ERR_INVALID, "displacement must be finite");
        if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) continue;
        int k = 0;
        for (int m = 1; m < (int)cd.size() / 3; m++)
            if (cd[3 * m] == d[0] && cd[3 * m + 1] == d[1] && cd[3 * m + 2] == d[2]) k = m;
        if (k == 0) {
            if ((int)cd.size() / 3 > SDM_MAX_CLASSES)
                return fail(SDM_ERR_UNSUPPORTED, "more than 8 distinct displacement vectors");
            cd.insert(cd.end(), d, d + 3);
            k = (int)cd.size() / 3 - 1;
        }
        cls[i] = k;
        lig.push_back(i);
    }
    // grouped by class, in atom order within a class; blocks of at most 64 entries that
    // never span two classes (with two guests, a block spanning both would have a
    // bounding box over both, and the displaced-state kernels cull candidate atoms
    // against that box)
    std::stable_sort(lig.begin(), lig.end(), [&](int a, int b) { return cls[a] < cls[b]; });
    std::vector<int> lbs, lbn;
    for (size_t k0 = 0; k0 < lig.size();) {
        size_t k1 = k0;
        while (k1 < lig.size() && cls[lig[k1]] == cls[lig[k0]]) k1++;
        for (size_t o = k0; o < k1; o += 64) {
            lbs.push_back((int)o);
            lbn.push_back((int)std::min<size_t>(64, k1 - o));
        }
        k0 = k1;
    }
    if (lbs.empty()) {
        lbs.push_back(0);
        lbn.push_back(0);
    }

    sdm_context* c = new sdm_context();
    int rc = SDM_OK;
    auto bail = [&](int code) {
        sdm_destroy(c);
        return code;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(SDM_ERR_HIP, "no HIP device"));
    if ((rc = resolve_scratch(c->device, &c->scratch_sel)) != SDM_OK) return bail(rc);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side[0], hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side[1], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork0, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_tiles, hipEventDisableTiming) != hipSuccess)
        return bail(fail(SDM_ERR_HIP, "hipStreamCreate/hipEventCreate failed"));
    c->N = N;
    c->max_ex = max_ex;
    c->Np = (N + 63) / 64 * 64;
    c->NB = c->Np / 64;
    c->R = n_replicas;
    c->L = (int)lig.size();
    c->LB = (int)lbs.size();
    c->ncls = (int)cd.size() / 3 - 1;
    c->nbond = sys->n_bonds;
    c->nang = sys->n_angles;
    c->nbw = std::max(1, (N + 255) / 256);
    c->maxj = tile_capacity(sys, c->NB);
    c->maxc = std::min(c->NB, 256 * std::max(1, c->ncls) + 64);
    Geom& g = c->g;
    g.geo = sys->lj_rule == SDM_LJ_GEOMETRIC ? 1 : 0;
    c->periodic = sys->periodic ? 1 : 0;
    c->reg_box.assign(6 * (size_t)n_replicas, 0.0);
    for (int r = 0; r < n_replicas; r++) {  // no positions yet: empty boxes
        c->reg_box[6 * (size_t)r] = 1.0;
        c->reg_box[6 * (size_t)r + 3] = -1.0;
    }
    if (c->periodic) {
        g.Lx = sys->box[0];
        g.Ly = sys->box[1];
        g.Lz = sys->box[2];
        g.iLx = 1.0 / g.Lx;
        g.iLy = 1.0 / g.Ly;
        g.iLz = 1.0 / g.Lz;
    } else {
        // no images: L * round(d / L) with a zero inverse is 0 for every d
        g.Lx = g.Ly = g.Lz = 1e30;
        g.iLx = g.iLy = g.iLz = 0.0;
    }
    g.bx = (float)g.Lx;
    g.by = (float)g.Ly;
    g.bz = (float)g.Lz;
    g.ibx = (float)g.iLx;
    g.iby = (float)g.iLy;
    g.ibz = (float)g.iLz;
    double rcut = sys->cutoff, erf_ = sys->eps_rf;
    g.rc2 = rcut * rcut;
    g.krf = (erf_ - 1.0) / ((2.0 * erf_ + 1.0) * rcut * rcut * rcut);
    g.crf = 3.0 * erf_ / ((2.0 * erf_ + 1.0) * rcut);
    g.rc2f = (float)g.rc2;
    g.krff = (float)g.krf;
    g.crff = (float)g.crf;
    c->list_r2 = (float)((rcut + 1e-3) * (rcut + 1e-3));
    if (c->periodic) {
        g.oix = (float)(g.Lx - (rcut + 1e-3) - 1e-3);
        g.oiy = (float)(g.Ly - (rcut + 1e-3) - 1e-3);
        g.oiz = (float)(g.Lz - (rcut + 1e-3) - 1e-3);
    } else {
        g.oix = g.oiy = g.oiz = 1e30f;
    }
    int N_ = N, Np = c->Np, R = c->R, NB = c->NB;
    size_t RNp = (size_t)R * Np;
#define A(ptr, n) \
    if ((rc = dalloc(c, &c->ptr, (n))) != SDM_OK) return bail(rc)
    A(qs, Np); A(sh, Np); A(se, Np); A(q64, Np); A(sig64, Np); A(eps64, Np); A(par64, Np);
    A(invm, Np); A(mass, Np); A(ex_start, N_ + 1); A(ex_list, elist.size());
    A(cls, Np); A(lig, std::max(1, c->L)); A(cdispl, cd.size());
    A(lb_start, lbs.size()); A(lb_count, lbn.size());
    A(bond_atoms, 2 * std::max(1, c->nbond)); A(bond_r0, std::max(1, c->nbond));
    A(bond_k, std::max(1, c->nbond)); A(ang_atoms, 3 * std::max(1, c->nang));
    A(ang_t0, std::max(1, c->nang)); A(ang_k, std::max(1, c->nang));
    A(bstart, N + 1); A(bcode, 2 * std::max(1, c->nbond));
    A(astart, N + 1); A(acode, 3 * std::max(1, c->nang));
    if (c->nbond + c->nang > 0) A(Fb, 3 * RNp);
    A(x64, 3 * RNp); A(v64, 3 * RNp); A(posq_s, RNp); A(par_s, RNp); A(arec, 2 * RNp);
    A(orig_s, RNp);
    A(slot_of, RNp); A(key, RNp); A(bb_lo, (size_t)R * NB); A(bb_hi, (size_t)R * NB);
    A(tiles, (size_t)R * NB * c->maxj); A(tmask, (size_t)R * NB * c->maxj);
    A(xb, RNp); A(dmax, (size_t)R * DISP_STRIDE);
    A(cmap, (size_t)R * NB); A(lig_of, Np);
    A(ntiles, (size_t)R * NB); A(F, 3 * RNp);
    A(e_pair, (size_t)R * NB * 32); A(e_bond, (size_t)R * c->nbw); A(u_part, (size_t)R * c->LB);
    A(e64, (size_t)R * NB * E64_SPLIT); A(rec64, RNp);
    A(cand, (size_t)R * c->maxc); A(ncand, R);
    A(dA, (size_t)R * c->maxc * c->LB * 64); A(dlig, (size_t)R * std::max(1, c->L));
    A(mlig, (size_t)R * c->maxc * std::max(1, c->L));
    A(dB, (size_t)R * c->maxc * 64); A(state, R); A(rec, R); A(flags, R);
    A(uhist, (size_t)R * c->hist_cap); A(counters, 2 * (size_t)R); A(ke_out, R);
#undef A
    // topology upload
    if ((rc = upload_nb_params(c, sys->charge, sys->sigma, sys->epsilon)) != SDM_OK) return bail(rc);
    {
        std::vector<double> im(Np, 0.0), m(Np, 0.0);
        for (int i = 0; i < N; i++) {
            m[i] = sys->mass[i];
            im[i] = sys->mass[i] == 0.0 ? 0.0 : 1.0 / sys->mass[i];
        }
        std::vector<int> clsp(Np, 0);
        std::copy(cls.begin(), cls.end(), clsp.begin());
        std::vector<int> orig_pad(RNp, -1);
        if ((rc = upload(c, c->invm, im.data(), Np)) ||
            (rc = upload(c, c->mass, m.data(), Np)) ||
            (rc = upload(c, c->ex_start, ecount.data(), N + 1)) ||
            (rc = upload(c, c->ex_list, elist.data(), elist.size())) ||
            (rc = upload(c, c->cls, clsp.data(), Np)) ||
            (rc = upload(c, c->cdispl, cd.data(), cd.size())) ||
            (rc = upload(c, c->orig_s, orig_pad.data(), RNp)))
            return bail(rc);
        if (c->L > 0 && (rc = upload(c, c->lig, lig.data(), lig.size()))) return bail(rc);
        {
            std::vector<int> lof(Np, -1);
            for (size_t li = 0; li < lig.size(); li++) lof[lig[li]] = (int)li;
            if ((rc = upload(c, c->lig_of, lof.data(), Np))) return bail(rc);
        }
        if ((rc = upload(c, c->lb_start, lbs.data(), lbs.size()))) return bail(rc);
        if ((rc = upload(c, c->lb_count, lbn.data(), lbn.size()))) return bail(rc);
        if (c->nbond > 0 &&
            ((rc = upload(c, c->bond_atoms, sys->bond_atoms, 2 * (size_t)c->nbond)) ||
             (rc = upload(c, c->bond_r0, sys->bond_r0, c->nbond)) ||
             (rc = upload(c, c->bond_k, sys->bond_k, c->nbond))))
            return bail(rc);
        if (c->nang > 0 &&
            ((rc = upload(c, c->ang_atoms, sys->angle_atoms, 3 * (size_t)c->nang)) ||
             (rc = upload(c, c->ang_t0, sys->angle_theta0, c->nang)) ||
             (rc = upload(c, c->ang_k, sys->angle_k, c->nang))))
            return bail(rc);
        // per-atom term lists: counts, exclusive scan, then the entries in term order
        std::vector<int> bst(N + 1, 0), ast(N + 1, 0);
        for (int k = 0; k < c->nbond; k++)
            for (int j = 0; j < 2; j++) bst[sys->bond_atoms[2 * k + j] + 1]++;
        for (int k = 0; k < c->nang; k++)
            for (int j = 0; j < 3; j++) ast[sys->angle_atoms[3 * k + j] + 1]++;
        for (int i = 0; i < N; i++) {
            bst[i + 1] += bst[i];
            ast[i + 1] += ast[i];
        }
        std::vector<int> bco(std::max(1, 2 * c->nbond)), aco(std::max(1, 3 * c->nang));
        std::vector<int> bfill(bst.begin(), bst.end() - 1), afill(ast.begin(), ast.end() - 1);
        for (int k = 0; k < c->nbond; k++)
            for (int j = 0; j < 2; j++) bco[bfill[sys->bond_atoms[2 * k + j]]++] = (k << 1) | j;
        for (int k = 0; k < c->nang; k++)
            for (int j = 0; j < 3; j++) aco[afill[sys->angle_atoms[3 * k + j]]++] = (k << 2) | j;
        if ((rc = upload(c, c->bstart, bst.data(), N + 1)) ||
            (rc = upload(c, c->bcode, bco.data(), bco.size())) ||
            (rc = upload(c, c->astart, ast.data(), N + 1)) ||
            (rc = upload(c, c->acode, aco.data(), aco.size())))
            return bail(rc);
        std::vector<DevState> st(R);
        sdm_params p = default_params();
        for (int r = 0; r < R; r++) {
            std::memset(&st[r], 0, sizeof(DevState));
            std::memcpy(&st[r], &p, sizeof(sdm_params));
        }
        if ((rc = upload(c, c->state, st.data(), R))) return bail(rc);
    }
    if ((rc = setup_cells(c, 0)) != SDM_OK) return bail(rc);
    *out = c;
    return SDM_OK;
}

int sdm_destroy(sdm_context* c) {
    if (!c) return SDM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (void* p : c->allocs) (void)hipFree(p);
    for (hipEvent_t e : {c->ev_fork0, c->ev_fork, c->ev_join[0], c->ev_join[1], c->ev_tiles})
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t st : {c->side[0], c->side[1]})
        if (st) (void)hipStreamDestroy(st);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SDM_OK;
}

int sdm_num_replicas(const sdm_context* c, int32_t* n_replicas, int32_t* n_atoms) {
    if (!c) return fail(SDM_ERR_INVALID, "null context");
    if (n_replicas) *n_replicas = c->R;
    if (n_atoms) *n_atoms = c->N;
    return SDM_OK;
}

int sdm_set_params(sdm_context* c, int32_t replica, const sdm_params* p) {
    CHK(check_replica(c, replica, true));
    CHK(validate_params(p));
    int r0 = replica < 0 ? 0 : replica, r1 = replica < 0 ? c->R : replica + 1;
    for (int r = r0; r < r1; r++)
        HIPCHK(hipMemcpyAsync(&c->state[r], p, sizeof(sdm_params), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SDM_OK;
}

int sdm_get_params(sdm_context* c, int32_t replica, sdm_params* p) {
    CHK(check_replica(c, replica, false));
    if (!p) return fail(SDM_ERR_INVALID, "null output");
    HIPCHK(hipMemcpyAsync(p, &c->state[replica], sizeof(sdm_params), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SDM_OK;
}

int sdm_set_time(sdm_context* c, int32_t replica, double time, int64_t step) {
    CHK(check_replica(c, replica, true));
    int r0 = replica < 0 ? 0 : replica, r1 = replica < 0 ? c->R : replica + 1;
    for (int r = r0; r < r1; r++) {
        HIPCHK(hipMemcpyAsync(&c->state[r].time, &time, sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(&c->state[r].step, &step, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return SDM_OK;
}

int sdm_set_nonbonded_params(sdm_context* c, const double* q, const double* sig, const double* eps) {
    if (!c || !q || !sig || !eps) return fail(SDM_ERR_INVALID, "null argument");
    c->since_sort = 0;  // the slot-ordered parameter copies are rebuilt by the next sort
    return upload_nb_params(c, q, sig, eps);
}

// constraint clusters (connected components of the constraint graph) for k_constrain;
// t == nullptr removes them
static int set_constraints(sdm_context* c, const sdm_forcefield_terms* t) {
    int nc = t ? t->n_constraints : 0;
    std::vector<int> parent(c->N);
    for (int i = 0; i < c->N; i++) parent[i] = i;
    auto find = [&](int a) {
        while (parent[a] != a) a = parent[a] = parent[parent[a]];
        return a;
    };
    std::vector<double> invm(c->N);
    if (nc > 0) CHK(download(c, invm.data(), c->invm, c->N));
    for (int k = 0; k < nc; k++) {
        int a = t->constraint_atoms[2 * k], b = t->constraint_atoms[2 * k + 1];
        if (a == b || !(t->constraint_length[k] > 0.0))
            return fail(SDM_ERR_INVALID, "a constraint needs two atoms and a positive length");
        if (invm[a] == 0.0 || invm[b] == 0.0)
            return fail(SDM_ERR_INVALID, "constrained atoms must have mass");
        parent[find(a)] = find(b);
    }
    // clusters in order of their first atom; atoms and constraints grouped per cluster
    std::vector<int> root_cl(c->N, -1), local(c->N, -1), cl_start{0}, cl_atoms, flag(c->N, 0);
    std::vector<std::vector<int>> members;
    for (int k = 0; k < nc; k++)
        for (int j = 0; j < 2; j++) flag[t->constraint_atoms[2 * k + j]] = 1;
    for (int i = 0; i < c->N; i++) {
        if (!flag[i]) continue;
        int rt = find(i);
        if (root_cl[rt] < 0) {
            root_cl[rt] = (int)members.size();
            members.emplace_back();
        }
        local[i] = (int)members[root_cl[rt]].size();
        members[root_cl[rt]].push_back(i);
    }
    std::vector<std::vector<int>> cons(members.size());
    for (int k = 0; k < nc; k++) cons[root_cl[find(t->constraint_atoms[2 * k])]].push_back(k);
    std::vector<int> con_start{0}, con_ij;
    std::vector<double> con_d2;
    for (size_t q = 0; q < members.size(); q++) {
        if ((int)members[q].size() > SHAKE_MAX_ATOMS || (int)cons[q].size() > SHAKE_MAX_CONS)
            return fail(SDM_ERR_UNSUPPORTED, "a constraint cluster has more than 16 atoms or "
                                             "32 constraints");
        for (int a : members[q]) cl_atoms.push_back(a);
        cl_start.push_back((int)cl_atoms.size());
        for (int k : cons[q]) {
            con_ij.push_back(local[t->constraint_atoms[2 * k]]);
            con_ij.push_back(local[t->constraint_atoms[2 * k + 1]]);
            con_d2.push_back(t->constraint_length[k] * t->constraint_length[k]);
        }
        con_start.push_back((int)con_d2.size());
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    for (void* p : {(void*)c->cl_start, (void*)c->cl_atoms, (void*)c->con_start,
                    (void*)c->con_ij, (void*)c->con_d2, (void*)c->cflag, (void*)c->dxs})
        dfree(c, p);
    c->cl_start = c->cl_atoms = c->con_start = c->con_ij = c->cflag = nullptr;
    c->con_d2 = c->dxs = nullptr;
    c->ncl = 0;
    drop_graph(c);
    if (nc == 0) return SDM_OK;
    if (t->constraint_tol > 0.0) c->con_tol = t->constraint_tol;
    CHK(dalloc(c, &c->cl_start, cl_start.size()));
    CHK(dalloc(c, &c->cl_atoms, cl_atoms.size()));
    CHK(dalloc(c, &c->con_start, con_start.size()));
    CHK(dalloc(c, &c->con_ij, con_ij.size()));
    CHK(dalloc(c, &c->con_d2, con_d2.size()));
    CHK(dalloc(c, &c->cflag, c->N));
    CHK(dalloc(c, &c->dxs, 3 * (size_t)c->R * c->Np));
    CHK(upload(c, c->cl_start, cl_start.data(), cl_start.size()));
    CHK(upload(c, c->cl_atoms, cl_atoms.data(), cl_atoms.size()));
    CHK(upload(c, c->con_start, con_start.data(), con_start.size()));
    CHK(upload(c, c->con_ij, con_ij.data(), con_ij.size()));
    CHK(upload(c, c->con_d2, con_d2.data(), con_d2.size()));
    CHK(upload(c, c->cflag, flag.data(), c->N));
    c->ncl = (int)members.size();
    return SDM_OK;
}

int sdm_set_cm_motion_removal(sdm_context* c, int32_t frequency) {
    if (!c) return fail(SDM_ERR_INVALID, "null context");
    if (frequency < 0) return fail(SDM_ERR_INVALID, "frequency must be >= 0");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->cm_every = frequency;
    drop_graph(c);
    return SDM_OK;
}

int sdm_set_forcefield_terms(sdm_context* c, const sdm_forcefield_terms* t) {
    if (!c) return fail(SDM_ERR_INVALID, "null context");
    sdm_forcefield_terms none{};
    if (!t) t = &none;
    if (t->n_torsions < 0 || t->n_impropers < 0 || t->n_posre_fb < 0 || t->n_posre_harm < 0 ||
        t->n_exceptions < 0 || t->n_constraints < 0)
        return fail(SDM_ERR_INVALID, "negative term count");
    auto ok_atoms = [&](const int32_t* a, int n, int per) {
        if (n > 0 && !a) return false;
        for (int k = 0; k < n * per; k++)
            if (a[k] < 0 || a[k] >= c->N) return false;
        return true;
    };
    auto ok_arr = [&](const void* p, int n) { return n == 0 || p != nullptr; };
    if (!ok_atoms(t->torsion_atoms, t->n_torsions, 4) ||
        !ok_atoms(t->improper_atoms, t->n_impropers, 4) ||
        !ok_atoms(t->posre_fb_atoms, t->n_posre_fb, 1) ||
        !ok_atoms(t->posre_harm_atoms, t->n_posre_harm, 1) ||
        !ok_atoms(t->exception_atoms, t->n_exceptions, 2) ||
        !ok_atoms(t->constraint_atoms, t->n_constraints, 2))
        return fail(SDM_ERR_INVALID, "term atom index out of range");
    if (!ok_arr(t->torsion_n, t->n_torsions) || !ok_arr(t->torsion_phase, t->n_torsions) ||
        !ok_arr(t->torsion_k, t->n_torsions) || !ok_arr(t->improper_phi0, t->n_impropers) ||
        !ok_arr(t->improper_k, t->n_impropers) || !ok_arr(t->posre_fb_x0, t->n_posre_fb) ||
        !ok_arr(t->posre_fb_k, t->n_posre_fb) || !ok_arr(t->posre_fb_tol, t->n_posre_fb) ||
        !ok_arr(t->posre_harm_x0, t->n_posre_harm) || !ok_arr(t->posre_harm_k, t->n_posre_harm) ||
        !ok_arr(t->exception_qq, t->n_exceptions) || !ok_arr(t->exception_sigma, t->n_exceptions) ||
        !ok_arr(t->exception_epsilon, t->n_exceptions) ||
        !ok_arr(t->constraint_length, t->n_constraints))
        return fail(SDM_ERR_INVALID, "null term array");
    if (t->n_exceptions > 0) {  // both atoms of an exception move together under d
        std::vector<int> cls(c->N);
        CHK(download(c, cls.data(), c->cls, c->N));
        for (int k = 0; k < t->n_exceptions; k++)
            if (cls[t->exception_atoms[2 * k]] != cls[t->exception_atoms[2 * k + 1]])
                return fail(SDM_ERR_UNSUPPORTED,
                            "an exception pair's atoms must have the same displacement");
    }
    CHK(set_constraints(c, t->n_constraints > 0 ? t : nullptr));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (void* p : c->term_bufs) dfree(c, p);
    c->term_bufs.clear();
    dfree(c, c->e_t1);
    dfree(c, c->e_t2);
    c->e_t1 = c->e_t2 = nullptr;
    c->terms = DevTerms{};
    c->nterm = c->ntb = 0;
    drop_graph(c);
    c->components_valid = false;
    int nterm = t->n_torsions + t->n_impropers + t->n_posre_fb + t->n_posre_harm + t->n_exceptions;
    if (nterm == 0) return SDM_OK;
    // device copies; the exception charge products include k_e
    auto put = [&](auto** dst, const auto* src, size_t n) {
        using T = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
        T* p = nullptr;
        int rc = dalloc(c, &p, n);
        if (rc == SDM_OK) rc = upload(c, p, src, n);
        if (rc == SDM_OK) c->term_bufs.push_back(p);
        *dst = p;
        return rc;
    };
    DevTerms& T = c->terms;
    T.nt = t->n_torsions;
    T.ni = t->n_impropers;
    T.nfb = t->n_posre_fb;
    T.nph = t->n_posre_harm;
    T.nex = t->n_exceptions;
    if (T.nt) {
        CHK(put(&T.t_atoms, t->torsion_atoms, 4 * (size_t)T.nt));
        CHK(put(&T.t_n, t->torsion_n, T.nt));
        CHK(put(&T.t_phase, t->torsion_phase, T.nt));
        CHK(put(&T.t_k, t->torsion_k, T.nt));
    }
    if (T.ni) {
        CHK(put(&T.i_atoms, t->improper_atoms, 4 * (size_t)T.ni));
        CHK(put(&T.i_phi0, t->improper_phi0, T.ni));
        CHK(put(&T.i_k, t->improper_k, T.ni));
    }
    if (T.nfb) {
        CHK(put(&T.fb_atom, t->posre_fb_atoms, T.nfb));
        CHK(put(&T.fb_x0, t->posre_fb_x0, 3 * (size_t)T.nfb));
        CHK(put(&T.fb_k, t->posre_fb_k, T.nfb));
        CHK(put(&T.fb_tol, t->posre_fb_tol, T.nfb));
    }
    if (T.nph) {
        CHK(put(&T.ph_atom, t->posre_harm_atoms, T.nph));
        CHK(put(&T.ph_x0, t->posre_harm_x0, 3 * (size_t)T.nph));
        CHK(put(&T.ph_k, t->posre_harm_k, 3 * (size_t)T.nph));
    }
    if (T.nex) {
        std::vector<double> qq(T.nex);
        for (int k = 0; k < T.nex; k++) qq[k] = SDM_ONE_4PI_EPS0 * t->exception_qq[k];
        CHK(put(&T.ex_atoms, t->exception_atoms, 2 * (size_t)T.nex));
        CHK(put(&T.ex_qq, qq.data(), T.nex));
        CHK(put(&T.ex_sig, t->exception_sigma, T.nex));
        CHK(put(&T.ex_eps, t->exception_epsilon, T.nex));
    }
    c->nterm = nterm;
    c->ntb = (nterm + 255) / 256;
    CHK(dalloc(c, &c->e_t1, (size_t)c->R * c->ntb));
    CHK(dalloc(c, &c->e_t2, (size_t)c->R * c->ntb));
    return SDM_OK;
}

int sdm_set_restraints(sdm_context* c, const sdm_centroid_restraint* cr, int32_t ncr,
                       const sdm_alignment_restraint* ar, int32_t nar) {
    if (!c) return fail(SDM_ERR_INVALID, "null context");
    if (ncr < 0 || nar < 0 || (ncr > 0 && !cr) || (nar > 0 && !ar))
        return fail(SDM_ERR_INVALID, "bad restraint arrays");
    auto atom_ok = [&](int a) { return a >= 0 && a < c->N; };
    std::vector<double> mass(c->N);
    CHK(download(c, mass.data(), c->mass, c->N));
    std::vector<DevCentroid> hc(ncr);
    std::vector<DevAlign> ha(nar);
    std::vector<int> atoms;
    std::vector<double> wts;
    // one group: its atoms and weights normalized to sum 1 (masses by default)
    auto add_group = [&](const int32_t* at, const double* w, int n, int* beg) {
        if (n <= 0 || !at) return fail(SDM_ERR_INVALID, "empty restraint group");
        double sum = 0.0;
        for (int i = 0; i < n; i++) {
            if (!atom_ok(at[i])) return fail(SDM_ERR_INVALID, "restraint atom out of range");
            sum += w ? w[i] : mass[at[i]];
        }
        if (!(sum > 0.0) || !std::isfinite(sum))
            return fail(SDM_ERR_INVALID, "restraint group weights must sum to > 0");
        *beg = (int)atoms.size();
        for (int i = 0; i < n; i++) {
            atoms.push_back(at[i]);
            wts.push_back((w ? w[i] : mass[at[i]]) / sum);
        }
        return SDM_OK;
    };
    for (int k = 0; k < ncr; k++) {
        const sdm_centroid_restraint& q = cr[k];
        DevCentroid& d = hc[k];
        d = DevCentroid{};
        CHK(add_group(q.lig_atoms, q.lig_weights, q.n_lig, &d.g1));
        CHK(add_group(q.rcpt_atoms, q.rcpt_weights, q.n_rcpt, &d.g2));
        d.n1 = q.n_lig;
        d.n2 = q.n_rcpt;
        d.do_angles = q.do_angles ? 1 : 0;
        for (int j = 0; j < 3; j++) {
            d.ref[j] = q.rcpt_ref[j];
            d.ref[3 + j] = q.lig_ref[j];
            d.off[j] = q.offset[j];
            d.kf[j] = q.kf[j];
            d.lo[j] = q.lo[j];
            d.hi[j] = q.hi[j];
        }
        if (d.do_angles)
            for (int j = 0; j < 6; j++)
                if (!atom_ok(d.ref[j]))
                    return fail(SDM_ERR_INVALID, "restraint reference atom out of range");
        d.kfcm = q.kfcm;
        d.tol = q.tolcm;
        d.control = q.control;
    }
    for (int k = 0; k < nar; k++) {
        const sdm_alignment_restraint& q = ar[k];
        DevAlign& d = ha[k];
        for (int j = 0; j < 3; j++) {
            d.a[j] = q.liga_ref[j];
            d.b[j] = q.ligb_ref[j];
            d.off[j] = q.offset[j];
            if (!atom_ok(d.a[j]) || !atom_ok(d.b[j]))
                return fail(SDM_ERR_INVALID, "alignment reference atom out of range");
        }
        d.kfd = q.kfdispl;
        d.kth = q.ktheta;
        d.kps = q.kpsi;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    dfree(c, c->rcen);
    dfree(c, c->rali);
    dfree(c, c->ratom);
    dfree(c, c->rw);
    dfree(c, c->e_rest);
    c->rcen = nullptr;
    c->rali = nullptr;
    c->ratom = nullptr;
    c->rw = nullptr;
    c->e_rest = nullptr;
    c->ncen = c->nali = 0;
    drop_graph(c);
    if (ncr + nar == 0) return SDM_OK;
    CHK(dalloc(c, &c->rcen, hc.size()));
    CHK(dalloc(c, &c->rali, ha.size()));
    CHK(dalloc(c, &c->ratom, atoms.size()));
    CHK(dalloc(c, &c->rw, wts.size()));
    CHK(dalloc(c, &c->e_rest, (size_t)c->R * (ncr + nar)));
    CHK(upload(c, c->rcen, hc.data(), hc.size()));
    CHK(upload(c, c->rali, ha.data(), ha.size()));
    CHK(upload(c, c->ratom, atoms.data(), atoms.size()));
    CHK(upload(c, c->rw, wts.data(), wts.size()));
    c->ncen = ncr;
    c->nali = nar;
    c->components_valid = false;
    return SDM_OK;
}

static int set_vec(sdm_context* c, double* base, int32_t r, const double* xyz) {
    CHK(check_replica(c, r, false));
    if (!xyz) return fail(SDM_ERR_INVALID, "null array");
    std::vector<double> soa(3 * (size_t)c->Np, 0.0);
    for (int i = 0; i < c->N; i++)
        for (int d = 0; d < 3; d++) {
            double v = xyz[3 * i + d];
            if (!std::isfinite(v)) return fail(SDM_ERR_INVALID, "non-finite coordinate");
            soa[(size_t)d * c->Np + i] = v;
        }
    return upload(c, base + (size_t)r * 3 * c->Np, soa.data(), soa.size());
}

static int get_vec(sdm_context* c, const double* base, int32_t r, double* xyz) {
    CHK(check_replica(c, r, false));
    if (!xyz) return fail(SDM_ERR_INVALID, "null array");
    std::vector<double> soa(3 * (size_t)c->Np);
    CHK(download(c, soa.data(), base + (size_t)r * 3 * c->Np, soa.size()));
    for (int i = 0; i < c->N; i++)
        for (int d = 0; d < 3; d++) xyz[3 * i + d] = soa[(size_t)d * c->Np + i];
    return SDM_OK;
}

int sdm_set_positions(sdm_context* c, int32_t r, const double* xyz) {
    if (c) c->since_sort = 0;  // the slot order is re-sorted from the new positions
    CHK(set_vec(c, c->x64, r, xyz));
    if (c->periodic) return SDM_OK;
    // non-periodic: the sort grid follows the positions
    double* b = &c->reg_box[6 * (size_t)r];
    for (int d = 0; d < 3; d++) {
        b[d] = 1e300;
        b[3 + d] = -1e300;
    }
    for (int i = 0; i < c->N; i++)
        for (int d = 0; d < 3; d++) {
            b[d] = std::min(b[d], xyz[3 * i + d]);
            b[3 + d] = std::max(b[3 + d], xyz[3 * i + d]);
        }
    return setup_cells(c, c->tune.cell_cells_per_dim);
}
int sdm_get_positions(sdm_context* c, int32_t r, double* xyz) { return get_vec(c, c->x64, r, xyz); }
int sdm_set_velocities(sdm_context* c, int32_t r, const double* xyz) { return set_vec(c, c->v64, r, xyz); }
int sdm_get_velocities(sdm_context* c, int32_t r, double* xyz) { return get_vec(c, c->v64, r, xyz); }

}  // extern "C"

namespace sdm {
// [R][N][3] <-> [R][3][Np]
__global__ void k_aos_to_soa(int N, int Np, const double* __restrict__ in, double* __restrict__ out) {
    int r = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int d = 0; d < 3; d++) out[((size_t)r * 3 + d) * Np + i] = in[((size_t)r * N + i) * 3 + d];
}
// per replica: lo (3) and hi (3) of the FP64 positions (one workgroup per replica)
__global__ __launch_bounds__(256) void k_extent(int N, int Np, const double* __restrict__ x64,
                                                double* __restrict__ out) {
    __shared__ double red[6][256];
    int r = blockIdx.x, t = threadIdx.x;
    const double* X = x64 + (size_t)r * 3 * Np;
    double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    for (int i = t; i < N; i += 256)
        for (int d = 0; d < 3; d++) {
            v[d] = fmin(v[d], X[d * Np + i]);
            v[3 + d] = fmax(v[3 + d], X[d * Np + i]);
        }
    for (int k = 0; k < 6; k++) red[k][t] = v[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o)
            for (int k = 0; k < 6; k++)
                red[k][t] = k < 3 ? fmin(red[k][t], red[k][t + o]) : fmax(red[k][t], red[k][t + o]);
        __syncthreads();
    }
    if (t < 6) out[(size_t)r * 6 + t] = red[t][0];
}
__global__ void k_soa_to_aos(int N, int Np, const double* __restrict__ in, double* __restrict__ out) {
    int r = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int d = 0; d < 3; d++) out[((size_t)r * N + i) * 3 + d] = in[((size_t)r * 3 + d) * Np + i];
}
}  // namespace sdm

extern "C" {

int sdm_set_positions_device(sdm_context* c, const double* d_xyz) {
    if (!c || !d_xyz) return fail(SDM_ERR_INVALID, "null argument");
    c->since_sort = 0;
    k_aos_to_soa<<<g2(c->N, 256, c->R), 256, 0, c->stream>>>(c->N, c->Np, d_xyz, c->x64);
    CHK(check_launch());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->periodic) return SDM_OK;
    // non-periodic: the sort grid follows the positions (per replica bounding boxes)
    double* ext = nullptr;
    CHK(dalloc(c, &ext, 6 * (size_t)c->R));
    k_extent<<<c->R, 256, 0, c->stream>>>(c->N, c->Np, c->x64, ext);
    int rc = check_launch();
    if (rc == SDM_OK) rc = download(c, c->reg_box.data(), ext, 6 * (size_t)c->R);
    dfree(c, ext);
    if (rc != SDM_OK) return rc;
    return setup_cells(c, c->tune.cell_cells_per_dim);
}

int sdm_get_positions_device(sdm_context* c, double* d_xyz) {
    if (!c || !d_xyz) return fail(SDM_ERR_INVALID, "null argument");
    k_soa_to_aos<<<g2(c->N, 256, c->R), 256, 0, c->stream>>>(c->N, c->Np, c->x64, d_xyz);
    CHK(check_launch());
    HIPCHK(hipStreamSynchronize(c->stream));
    return SDM_OK;
}

int sdm_compute(sdm_context* c) {
    if (!c) return fail(SDM_ERR_INVALID, "null context");
    enqueue_eval(c, false, false, take_order(c, false), true);
    c->components_valid = false;
    return check_launch();
}

int sdm_step(sdm_context* c, int32_t nsteps) {
    if (!c) return fail(SDM_ERR_INVALID, "null context");
    if (nsteps < 0) return fail(SDM_ERR_INVALID, "nsteps must be >= 0");
    c->components_valid = false;
    if (nsteps == 0) return SDM_OK;
    for (int s = 0; s < nsteps; s++) {
        bool full = take_order(c, true);
        // the pair energies (U0, U1, pot_energy) on the last step of the call, the one
        // whose record sdm_get_energies reads, or on every step if asked
        bool energy = c->tune.energy_every_step || s == nsteps - 1;
        if (c->tune.use_graph) {
            if (!c->gexec[full][energy]) CHK(build_graph(c, full, energy));
            HIPCHK(hipGraphLaunch(c->gexec[full][energy], c->stream));
        } else {
            enqueue_step(c, full, energy);
        }
    }
    return check_launch();
}

int sdm_get_energies(sdm_context* c, sdm_energies* out) {
    if (!c || !out) return fail(SDM_ERR_INVALID, "null argument");
    std::vector<DevRecord> rec(c->R);
    std::vector<int> fl(c->R);
    CHK(download(c, rec.data(), c->rec, c->R));
    CHK(download(c, fl.data(), c->flags, c->R));
    for (int r = 0; r < c->R; r++) {
        sdm_energies& e = out[r];
        const DevRecord& d = rec[r];
        e.U0 = d.U0; e.U1 = d.U1; e.u = d.u; e.Ub = d.Ub; e.bind_e = d.bind_e; e.fp = d.fp;
        e.W = d.W; e.Wp = d.Wp; e.sp = d.sp; e.pot_energy = d.pot_energy; e.lambda = d.lambda;
        e.lambda1 = d.lambda1; e.lambda2 = d.lambda2; e.u0 = d.u0; e.w0 = d.w0; e.work = d.work;
        e.time = d.time; e.step = d.step; e.flags = fl[r]; e.pad = 0;
    }
    return SDM_OK;
}

int sdm_clear_flags(sdm_context* c) {
    if (!c) return fail(SDM_ERR_INVALID, "null context");
    HIPCHK(hipMemsetAsync(c->flags, 0, sizeof(int) * c->R, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SDM_OK;
}

int sdm_compute_components(sdm_context* c) {
    if (!c) return fail(SDM_ERR_INVALID, "null context");
    size_t n3 = 3 * (size_t)c->R * c->Np;
    if (!c->Fb_sep) {
        CHK(dalloc(c, &c->Fb_sep, n3));
        CHK(dalloc(c, &c->dFd, n3));
    }
    k_fill_ll<<<(unsigned)((n3 + 255) / 256), 256, 0, c->stream>>>(c->Fb_sep, n3);
    k_fill_d<<<(unsigned)((n3 + 255) / 256), 256, 0, c->stream>>>(c->dFd, n3);
    enqueue_eval(c, false, true, take_order(c, false), true);
    if (c->L > 0)
        k_scatter_dF<<<dim3(c->maxc + c->LB, c->R), 64, 0, c->stream>>>(
            c->N, c->Np, c->L, c->maxc, c->lig, c->cls, c->orig_s, c->cand, c->ncand, c->dB,
            c->dlig, c->dFd);
    CHK(check_launch());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->components_valid = true;
    return SDM_OK;
}

}  // extern "C"

// The distribution of k_vsite_forces in FP64 on the host, for forces read back (x: one
// replica's positions, [3][Np]; out: [N][3]).  A site's force is left at zero, so forces a
// step has already distributed on the device pass through unchanged.
static void vsite_distribute_host(const sdm_context* c, const std::vector<double>& x,
                                  double* out) {
    const size_t Np = (size_t)c->Np;
    for (int k = 0; k < c->nvs; k++) {
        const int* a = &c->h_vs_at[4 * (size_t)k];
        const double* w = &c->h_vs_w[3 * (size_t)k];
        double f[3], g1[3], g2[3], g3[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < 3; d++) {
            f[d] = out[3 * (size_t)a[0] + d];
            out[3 * (size_t)a[0] + d] = 0.0;
        }
        if (c->h_vs_kind[k] == VSITE_OUT_OF_PLANE) {
            double r12[3], r13[3];
            for (int d = 0; d < 3; d++) {
                r12[d] = x[d * Np + a[2]] - x[d * Np + a[1]];
                r13[d] = x[d * Np + a[3]] - x[d * Np + a[1]];
            }
            const double c2[3] = {r13[1] * f[2] - r13[2] * f[1], r13[2] * f[0] - r13[0] * f[2],
                                  r13[0] * f[1] - r13[1] * f[0]};
            const double c3[3] = {f[1] * r12[2] - f[2] * r12[1], f[2] * r12[0] - f[0] * r12[2],
                                  f[0] * r12[1] - f[1] * r12[0]};
            for (int d = 0; d < 3; d++) {
                g2[d] = w[0] * f[d] + w[2] * c2[d];
                g3[d] = w[1] * f[d] + w[2] * c3[d];
                g1[d] = f[d] - g2[d] - g3[d];
            }
        } else {
            for (int d = 0; d < 3; d++) {
                g1[d] = w[0] * f[d];
                g2[d] = w[1] * f[d];
                if (c->h_vs_kind[k] == VSITE_AVERAGE3) g3[d] = w[2] * f[d];
            }
        }
        for (int d = 0; d < 3; d++) {
            out[3 * (size_t)a[1] + d] += g1[d];
            out[3 * (size_t)a[2] + d] += g2[d];
            if (c->h_vs_kind[k] != VSITE_AVERAGE2) out[3 * (size_t)a[3] + d] += g3[d];
        }
    }
}

extern "C" {

int sdm_set_virtual_sites(sdm_context* c, int32_t n, const int32_t* kind, const int32_t* atoms,
                          const double* weights) {
    if (!c) return fail(SDM_ERR_INVALID, "null context");
    if (n < 0) return fail(SDM_ERR_INVALID, "negative virtual-site count");
    if (n > 0 && (!kind || !atoms || !weights))
        return fail(SDM_ERR_INVALID, "null virtual-site array");
    const int N = c->N;
    std::vector<double> mass(N);
    std::vector<int> cls(N);
    CHK(download(c, mass.data(), c->mass, N));
    CHK(download(c, cls.data(), c->cls, N));
    std::vector<char> site(N, 0);
    std::vector<int> at(4 * (size_t)n);
    for (int k = 0; k < n; k++) {
        if (kind[k] != SDM_VSITE_AVERAGE2 && kind[k] != SDM_VSITE_AVERAGE3 &&
            kind[k] != SDM_VSITE_OUT_OF_PLANE)
            return fail(SDM_ERR_INVALID, "unknown virtual-site kind");
        const int np = kind[k] == SDM_VSITE_AVERAGE2 ? 2 : 3;
        for (int j = 0; j <= np; j++)
            if (atoms[4 * k + j] < 0 || atoms[4 * k + j] >= N)
                return fail(SDM_ERR_INVALID, "virtual-site atom index out of range");
        for (int j = 0; j < 3; j++)
            if (!std::isfinite(weights[3 * k + j]))
                return fail(SDM_ERR_INVALID, "virtual-site weights must be finite");
        const int a = atoms[4 * k];
        if (site[a]) return fail(SDM_ERR_INVALID, "an atom is defined as a virtual site twice");
        if (mass[a] != 0.0) return fail(SDM_ERR_INVALID, "a virtual site must have mass 0");
        site[a] = 1;
        for (int j = 0; j < 4; j++) at[4 * k + j] = atoms[4 * k + (j <= np ? j : 1)];
    }
    for (int k = 0; k < n; k++) {
        const int np = kind[k] == SDM_VSITE_AVERAGE2 ? 2 : 3;
        for (int j = 1; j <= np; j++) {
            const int p = at[4 * k + j];
            if (site[p]) return fail(SDM_ERR_INVALID, "a virtual site's parent is a virtual site");
            if (!(mass[p] > 0.0)) return fail(SDM_ERR_INVALID, "a virtual site's parent has no mass");
            if (cls[p] != cls[at[4 * k]])
                return fail(SDM_ERR_INVALID,
                            "a virtual site and its parents must share one displacement");
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    dfree(c, c->vs_kind);
    dfree(c, c->vs_at);
    dfree(c, c->vs_w);
    c->vs_kind = nullptr;
    c->vs_at = nullptr;
    c->vs_w = nullptr;
    c->nvs = 0;
    c->h_vs_kind.assign(kind, kind + n);
    c->h_vs_at = at;
    c->h_vs_w.assign(weights, weights + 3 * (size_t)n);
    if (n > 0) {
        CHK(dalloc(c, &c->vs_kind, n));
        CHK(dalloc(c, &c->vs_at, n));
        CHK(dalloc(c, &c->vs_w, 3 * (size_t)n));
        CHK(upload(c, c->vs_kind, kind, n));
        CHK(upload(c, c->vs_at, reinterpret_cast<const int4*>(at.data()), n));
        CHK(upload(c, c->vs_w, weights, 3 * (size_t)n));
    }
    c->nvs = n;
    c->components_valid = false;
    drop_graph(c);  // the step's kernels change
    return SDM_OK;
}

int sdm_get_forces(sdm_context* c, int32_t r, int32_t which, double* out) {
    CHK(check_replica(c, r, false));
    if (!out) return fail(SDM_ERR_INVALID, "null output");
    if (which < 0 || which > 3) return fail(SDM_ERR_INVALID, "unknown force component");
    int Np = c->Np, N = c->N;
    size_t off = (size_t)r * 3 * Np;
    std::vector<long long> f(3 * (size_t)Np);
    std::vector<int> slot(Np);
    std::vector<double> fbond(3 * (size_t)Np, 0.0);  // FP64 bonds and angles, atom order
    CHK(download(c, f.data(), c->F + off, f.size()));
    CHK(download(c, slot.data(), c->slot_of + (size_t)r * Np, Np));
    if (c->Fb) CHK(download(c, fbond.data(), c->Fb + off, fbond.size()));
    // forces on virtual sites are reported moved to their parents, as the step uses them
    auto sites = [&]() -> int {
        if (c->nvs == 0) return SDM_OK;
        std::vector<double> x(3 * (size_t)Np);
        CHK(download(c, x.data(), c->x64 + off, x.size()));
        vsite_distribute_host(c, x, out);
        return SDM_OK;
    };
    if (which == SDM_FORCE_HYBRID && !c->components_valid) {
        for (int i = 0; i < N; i++)
            for (int d = 0; d < 3; d++)
                out[3 * i + d] = (double)f[(size_t)d * Np + slot[i]] / SDM_FIX_SCALE +
                                 fbond[(size_t)d * Np + i];
        return sites();
    }
    if (!c->components_valid)
        return fail(SDM_ERR_STATE, "force components require sdm_compute_components");
    std::vector<long long> fb(3 * (size_t)Np);
    std::vector<double> dF(3 * (size_t)Np);
    CHK(download(c, fb.data(), c->Fb_sep + off, fb.size()));
    CHK(download(c, dF.data(), c->dFd + off, dF.size()));
    DevRecord rec;
    CHK(download(c, &rec, c->rec + r, 1));
    for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++) {
            size_t k = (size_t)d * Np + i;
            double f0 = (double)f[(size_t)d * Np + slot[i]] / SDM_FIX_SCALE;
            double b = (double)fb[k] / SDM_FIX_SCALE + fbond[k];
            double v = 0.0;
            switch (which) {
                case SDM_FORCE_HYBRID: v = f0 + rec.sp * dF[k] + b; break;
                case SDM_FORCE_STATE1: v = f0; break;
                case SDM_FORCE_STATE2: v = f0 + dF[k]; break;
                default: v = b; break;
            }
            out[3 * i + d] = v;
        }
    return sites();
}

int sdm_kinetic_energy(sdm_context* c, double* out) {
    if (!c || !out) return fail(SDM_ERR_INVALID, "null argument");
    // the kinetic-energy tolerance of the reference (applyToVelocities(..., 1e-4),
    // ReferenceSDMKernels.cpp:126) is a relative one; the projection here is iterated to
    // 1e-12 of the atoms' speeds, well inside it
    k_kinetic<<<c->R, 256, 0, c->stream>>>(c->N, c->Np, c->mass, c->invm, c->x64, c->v64,
                                           c->ncl, c->cl_start, c->cl_atoms, c->con_start,
                                           c->con_ij, c->cflag, 1e-12, c->flags, c->ke_out);
    CHK(check_launch());
    return download(c, out, c->ke_out, c->R);
}

int sdm_set_u_history(sdm_context* c, int32_t every, int32_t capacity) {
    if (!c) return fail(SDM_ERR_INVALID, "null context");
    if (every < 0 || capacity <= 0) return fail(SDM_ERR_INVALID, "bad history settings");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (capacity > c->hist_cap) {
        dfree(c, c->uhist);
        c->uhist = nullptr;
        CHK(dalloc(c, &c->uhist, (size_t)c->R * capacity));
    }
    c->hist_every = every;
    c->hist_cap = capacity;
    // the ring starts empty: no record of an earlier setting is returned
    std::vector<DevState> st(c->R);
    CHK(download(c, st.data(), c->state, c->R));
    for (auto& x : st) x.hist_n = 0;
    CHK(upload(c, c->state, st.data(), c->R));
    drop_graph(c);
    return SDM_OK;
}

int sdm_get_u_history(sdm_context* c, int32_t r, double* out, int32_t capacity, int32_t* count) {
    CHK(check_replica(c, r, false));
    if (!out || !count || capacity < 0) return fail(SDM_ERR_INVALID, "null argument");
    std::vector<double> ring(c->hist_cap);
    CHK(download(c, ring.data(), c->uhist + (size_t)r * c->hist_cap, ring.size()));
    DevState st;
    CHK(download(c, &st, c->state + r, 1));
    // record k (k = 0, 1, .. since the history was configured) sits in ring slot
    // k % hist_cap; the ring holds the last min(rec_n, hist_cap) of them
    int64_t rec_n = c->hist_every > 0 ? st.hist_n : 0;
    int64_t held = std::min<int64_t>(rec_n, c->hist_cap);
    int64_t n = std::min<int64_t>(held, capacity);
    for (int64_t k = 0; k < n; k++) out[k] = ring[(size_t)((rec_n - n + k) % c->hist_cap)];
    *count = (int32_t)n;
    return SDM_OK;
}

int sdm_set_constraint_tolerance(sdm_context* c, double tol) {
    if (!c) return fail(SDM_ERR_INVALID, "null context");
    if (!(tol > 0.0) || !std::isfinite(tol)) return fail(SDM_ERR_INVALID, "tolerance must be positive");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->con_tol = tol;
    drop_graph(c);  // the tolerance is a kernel argument of the captured steps
    return SDM_OK;
}

int sdm_count_pairs(sdm_context* c, int64_t* p0, int64_t* pdelta) {
    if (!c || !p0 || !pdelta) return fail(SDM_ERR_INVALID, "null argument");
    size_t n = 2 * (size_t)c->R;
    hipStream_t s = c->stream;
    k_fill_u64<<<1, 256, 0, s>>>(c->counters, n);
    enqueue_sort(c, take_order(c, false));
    enqueue_pair(c, true);
    if (c->L > 0) {
        k_ligcand<<<c->R, LIGCAND_THREADS, 0, s>>>(c->Np, c->NB, c->L, c->ncls, c->lig, c->cls, c->cdispl,
                                      c->x64, c->bb_lo, c->bb_hi, c->maxc, c->cand, c->ncand,
                                      c->list_r2, c->g, c->flags, c->cmap);
        k_denv<DPAIR_WAVES><<<dim3(c->maxc, c->R), 64 * DPAIR_WAVES, 0, s>>>(
            c->N, c->Np, c->L, c->LB, c->lig, c->lb_start, c->lb_count, c->cls, c->cdispl,
            c->q64, c->sig64, c->eps64, c->ex_start, c->ex_list, c->x64, c->orig_s, c->cand,
            c->ncand, c->maxc, c->mlig, c->dB, c->g);
        k_dlig<DPAIR_WAVES><<<dim3(c->maxc, c->LB, c->R), 64 * DPAIR_WAVES, 0, s>>>(
            c->N, c->Np, c->L, c->LB, c->lig, c->lb_start, c->lb_count, c->cls, c->cdispl,
            c->q64, c->sig64, c->eps64, c->x64, c->orig_s, c->cand, c->ncand, c->maxc, c->mlig,
            c->dA, c->g, c->counters + c->R);
        k_deltaA_red<DRED_WAVES><<<dim3(c->LB, c->R), 64 * DRED_WAVES, 0, s>>>(
            c->Np, c->L, c->LB, c->lig, c->lb_start, c->lb_count, c->cls, c->cdispl, c->q64,
            c->sig64, c->eps64, c->ex_start, c->ex_list, c->x64, c->ncand, c->maxc, c->dA,
            c->dlig, c->u_part, c->g, c->counters + c->R);
    }
    CHK(check_launch());
    std::vector<unsigned long long> h(n);
    CHK(download(c, h.data(), c->counters, n));
    for (int r = 0; r < c->R; r++) {
        p0[r] = (int64_t)((c->tune.pair_group || c->tune.pair_newton3) ? h[r] : h[r] / 2);
        pdelta[r] = (int64_t)h[c->R + r];
    }
    return SDM_OK;
}

int sdm_time_pair_kernel(sdm_context* c, int32_t nsteps, double* avg_ms) {
    if (!c || !avg_ms || nsteps <= 0) return fail(SDM_ERR_INVALID, "bad argument");
    // The pair kernel alone (spatial sort, then the pair kernel; no bonded or
    // displaced-state work running beside it), nsteps times back to back without
    // synchronisation, each launch between its own pair of events on the context's
    // stream.  Positions do not advance; the force accumulator is left incomplete.
    std::vector<hipEvent_t> ev(2 * (size_t)nsteps, nullptr);
    int rc = SDM_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = fail(SDM_ERR_HIP, "hipEventCreate failed");
    double total = 0.0;
    if (rc == SDM_OK) {
        for (int s = 0; s < nsteps; s++) {
            enqueue_sort(c, take_order(c, false));
            (void)hipEventRecord(ev[2 * s], c->stream);
            enqueue_pair(c, false, false);
            (void)hipEventRecord(ev[2 * s + 1], c->stream);
        }
        rc = check_launch();
        if (rc == SDM_OK && hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail(SDM_ERR_HIP, "hipStreamSynchronize failed");
        for (int s = 0; s < nsteps && rc == SDM_OK; s++) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]) != hipSuccess)
                rc = fail(SDM_ERR_HIP, "hipEventElapsedTime failed");
            total += ms;
        }
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc != SDM_OK) return rc;
    *avg_ms = total / nsteps;
    c->components_valid = false;
    return SDM_OK;
}

int sdm_get_tuning(sdm_context* c, sdm_tuning* t) {
    if (!c || !t) return fail(SDM_ERR_INVALID, "null argument");
    *t = c->tune;
    return SDM_OK;
}

int sdm_set_tuning(sdm_context* c, const sdm_tuning* t) {
    if (!c || !t) return fail(SDM_ERR_INVALID, "null argument");
    if (t->pair_waves_per_group != 1 && t->pair_waves_per_group != 2 && t->pair_waves_per_group != 4)
        return fail(SDM_ERR_INVALID, "pair_waves_per_group must be 1, 2 or 4");
    if (t->cell_cells_per_dim < 0 || t->cell_cells_per_dim > 1024)
        return fail(SDM_ERR_INVALID, "cell_cells_per_dim out of range");
    if (t->pair_group != 0 && t->pair_group != 8 && t->pair_group != 16)
        return fail(SDM_ERR_INVALID, "pair_group must be 0, 8 or 16");
    if (t->sort_interval < 1 || t->sort_interval > 100000)
        return fail(SDM_ERR_INVALID, "sort_interval must be in [1, 100000]");
    if (t->energy_every_step != 0 && t->energy_every_step != 1)
        return fail(SDM_ERR_INVALID, "energy_every_step must be 0 or 1");
    if (t->pair_newton3 != 0 && t->pair_newton3 != 1)
        return fail(SDM_ERR_INVALID, "pair_newton3 must be 0 or 1");
    if (t->pair_split != 1 && t->pair_split != 2 && t->pair_split != 4 && t->pair_split != 8 &&
        t->pair_split != 16 && t->pair_split != 32)
        return fail(SDM_ERR_INVALID, "pair_split must be 1, 2, 4, 8, 16 or 32");
    if (t->use_graph != 0 && t->use_graph != 1)
        return fail(SDM_ERR_INVALID, "use_graph must be 0 or 1");
    if (t->side_fork < 0 || t->side_fork > 4)
        return fail(SDM_ERR_INVALID, "side_fork must be 0, 1, 2, 3 or 4");
    if (!(t->list_skin >= 0.f && t->list_skin <= (float)SDM_MAX_LIST_SKIN))
        return fail(SDM_ERR_INVALID, "list_skin must be in [0, 0.1] nm");
    if (c->g.geo && (t->pair_group != 0 || t->pair_newton3 != 1))
        return fail(SDM_ERR_INVALID, "the geometric LJ rule runs on the block Newton-3 pair "
                                     "kernel (pair_group 0, pair_newton3 1)");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (t->cell_cells_per_dim != c->tune.cell_cells_per_dim) CHK(setup_cells(c, t->cell_cells_per_dim));
    c->tune = *t;
    c->since_sort = 0;
    drop_graph(c);
    return SDM_OK;
}

int sdm_autotune(sdm_context* c, int32_t steps, sdm_tuning* chosen) {
    if (!c || steps <= 0) return fail(SDM_ERR_INVALID, "bad argument");
    size_t n3 = 3 * (size_t)c->R * c->Np;
    std::vector<double> x(n3), v(n3);
    std::vector<DevState> st(c->R);
    std::vector<int> fl(c->R);
    CHK(download(c, x.data(), c->x64, n3));
    CHK(download(c, v.data(), c->v64, n3));
    CHK(download(c, st.data(), c->state, c->R));
    CHK(download(c, fl.data(), c->flags, c->R));
    sdm_tuning best = c->tune;
    double best_ms = 1e300;
    double glo[3], ghi[3];
    grid_region(c, glo, ghi);
    double L = std::min(ghi[0] - glo[0], std::min(ghi[1] - glo[1], ghi[2] - glo[2]));
    const double cell_sides[] = {0.24, 0.3, 0.4};
    const int waves[] = {1, 2, 4};
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // the trials' steps take no u' records (the ring keeps the caller's); the state
    // restored after each trial brings back the record count
    const int hist_every = c->hist_every;
    c->hist_every = 0;
    // one trial: n steps from the saved state after two unmeasured ones (the first
    // sorts and, with use_graph, captures), then the saved state back
    auto trial = [&](const sdm_tuning& t, int n, double& ms_out) -> int {
        CHK(sdm_set_tuning(c, &t));
        CHK(sdm_step(c, 2));
        HIPCHK(hipEventRecord(e0, c->stream));
        CHK(sdm_step(c, n));
        HIPCHK(hipEventRecord(e1, c->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        ms_out = ms;
        CHK(upload(c, c->x64, x.data(), n3));
        CHK(upload(c, c->v64, v.data(), n3));
        c->since_sort = 0;
        CHK(upload(c, c->state, st.data(), c->R));
        CHK(upload(c, c->flags, fl.data(), c->R));
        return SDM_OK;
    };
    // pair kernels: block waves without Newton-3, with Newton-3 and 1, 2, 4, 8, 16 or 32 waves
    // per i-block, grouped waves of 16 and of 8
    const int kernels[][3] = {{0, 0, 1}, {0, 1, 1}, {0, 1, 2}, {0, 1, 4}, {0, 1, 8},
                              {0, 1, 16}, {0, 1, 32}, {16, 1, 1}, {8, 1, 1}};
    std::vector<std::pair<double, sdm_tuning>> timed;
    int rc = SDM_OK;
    for (double side : cell_sides)
        for (int w : waves)
            for (const auto& kn : kernels) {
                if (rc != SDM_OK) break;
                if (c->g.geo && (kn[0] != 0 || kn[1] != 1)) continue;  // see sdm_set_tuning
                sdm_tuning t = c->tune;
                t.pair_waves_per_group = w;
                t.pair_group = kn[0];
                t.pair_newton3 = kn[1];
                t.pair_split = kn[2];
                t.cell_cells_per_dim = std::max(1, (int)std::floor(L / side));
                double ms = 0.0;
                rc = trial(t, steps, ms);
                if (rc == SDM_OK) timed.push_back({ms, t});
            }
    // the three fastest again with trials three times as long (configurations within a
    // few percent of each other are not told apart reliably by short trials), each with
    // the side work forked before or after the tile lists, and with the step replayed
    // from a captured graph and launched kernel by kernel: which is
    // faster depends on the size of the step, and it changes the ranking of the other
    // settings (graph replay of the three-stream step was measured slower than direct
    // launches at 22 x 22,750 atoms, where 8 waves per i-block won under replay and 4
    // under direct launches)
    // Each finalist is timed in two rounds over all of them and ranked by the sum, so
    // that a slow moment of the GPU during one trial does not decide between settings
    // that differ by a few percent.
    std::stable_sort(timed.begin(), timed.end(),
                     [](const auto& p, const auto& q) { return p.first < q.first; });
    std::vector<sdm_tuning> finalists;
    for (size_t k = 0; k < timed.size() && k < 3; k++)
        for (int variant = 0; variant < 10; variant++) {
            sdm_tuning t = timed[k].second;
            t.use_graph = variant & 1;
            t.side_fork = variant >> 1;
            finalists.push_back(t);
        }
    std::vector<double> total(finalists.size(), 0.0);
    for (int round = 0; round < 2 && rc == SDM_OK; round++)
        for (size_t k = 0; k < finalists.size() && rc == SDM_OK; k++) {
            double ms = 0.0;
            rc = trial(finalists[k], 3 * steps, ms);
            total[k] += ms;
        }
    for (size_t k = 0; k < finalists.size() && rc == SDM_OK; k++)
        if (total[k] < best_ms) {
            best_ms = total[k];
            best = finalists[k];
        }
    // then the Verlet skin of the tile lists for the chosen configuration (block kernels
    // only): a wider skin keeps the lists longer but the pair kernel evaluates more pairs
    // outside the cutoff.  Two rounds again, ranked by the sum.
    if (rc == SDM_OK && best.pair_group == 0) {
        const float skins[] = {0.f, 0.01f, 0.02f, 0.03f, 0.04f};
        constexpr int NS = sizeof(skins) / sizeof(skins[0]);
        double tsk[NS] = {};
        for (int round = 0; round < 2 && rc == SDM_OK; round++)
            for (int k = 0; k < NS && rc == SDM_OK; k++) {
                sdm_tuning t = best;
                t.list_skin = skins[k];
                double ms = 0.0;
                rc = trial(t, 3 * steps, ms);
                tsk[k] += ms;
            }
        int kb = 0;
        for (int k = 1; k < NS; k++)
            if (tsk[k] < tsk[kb]) kb = k;
        if (rc == SDM_OK) best.list_skin = skins[kb];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->hist_every = hist_every;
    drop_graph(c);  // graphs captured during the trials record no history
    if (rc != SDM_OK) return rc;
    CHK(sdm_set_tuning(c, &best));
    if (chosen) *chosen = best;
    return SDM_OK;
}

}  // extern "C"

// The adjoint's kernels for a selection already staged on the device at d_sel.
static int du_dparams_launch(sdm_context* c, const double* d_frames, int32_t nframes,
                             const int32_t* d_sel, int32_t nsel, double* d_grad, double* d_u) {
    if (d_u && c->L > 0)
        k_u_frames<<<nframes, 256, 0, c->stream>>>(c->N, c->L, c->lig, c->cls, c->cdispl, c->q64,
                                                   c->sig64, c->eps64, c->ex_start, c->ex_list,
                                                   d_frames, d_u, c->g);
    else if (d_u)
        k_fill_d<<<(nframes + 255) / 256, 256, 0, c->stream>>>(d_u, (size_t)nframes);
    if (d_grad)
        k_du_dparams<<<dim3(nsel, nframes), 64, 0, c->stream>>>(
            c->N, d_sel, nsel, c->cls, c->cdispl, c->q64, c->sig64, c->eps64, c->ex_start,
            c->ex_list, d_frames, d_grad, c->g);
    CHK(check_launch());
    HIPCHK(hipStreamSynchronize(c->stream));
    return SDM_OK;
}

static int du_dparams_check(sdm_context* c, const double* d_frames, int32_t nframes,
                            const int32_t* sel, int32_t nsel, double* d_grad) {
    if (!c || !d_frames || nframes <= 0) return fail(SDM_ERR_INVALID, "bad argument");
    if (d_grad && (!sel || nsel <= 0)) return fail(SDM_ERR_INVALID, "gradient needs selected atoms");
    if (d_grad)
        for (int k = 0; k < nsel; k++)
            if (sel[k] < 0 || sel[k] >= c->N) return fail(SDM_ERR_INVALID, "selected atom out of range");
    return SDM_OK;
}

extern "C" {

int sdm_du_dparams_frames(sdm_context* c, const double* d_frames, int32_t nframes,
                          const int32_t* sel, int32_t nsel, double* d_grad, double* d_u) {
    CHK(du_dparams_check(c, d_frames, nframes, sel, nsel, d_grad));
    if (d_grad) {
        if ((size_t)nsel > c->sel_cap) {
            HIPCHK(hipStreamSynchronize(c->stream));
            dfree(c, c->sel_dev);
            c->sel_dev = nullptr;
            c->sel_cap = 0;
            CHK(dalloc(c, &c->sel_dev, nsel));
            c->sel_cap = nsel;
        }
        HIPCHK(hipMemcpyAsync(c->sel_dev, sel, sizeof(int) * nsel, hipMemcpyHostToDevice, c->stream));
    }
    return du_dparams_launch(c, d_frames, nframes, c->sel_dev, nsel, d_grad, d_u);
}

int sdm_du_dparams_frames_noalloc(sdm_context* c, const double* d_frames, int32_t nframes,
                                  const int32_t* sel, int32_t nsel, double* d_grad,
                                  double* d_u) {
    CHK(du_dparams_check(c, d_frames, nframes, sel, nsel, d_grad));
    if (d_grad && nsel > SDM_NOALLOC_MAX_SEL)
        return fail(SDM_ERR_CAPACITY, "sdm_du_dparams_frames_noalloc: more than "
                                      "SDM_NOALLOC_MAX_SEL (" +
                                          std::to_string(SDM_NOALLOC_MAX_SEL) +
                                          ") selected atoms");
    // every entry the kernels read ([0, nsel)) is written first; nothing needs clearing
    if (d_grad)
        HIPCHK(hipMemcpyAsync(c->scratch_sel, sel, sizeof(int32_t) * nsel,
                              hipMemcpyHostToDevice, c->stream));
    return du_dparams_launch(c, d_frames, nframes, c->scratch_sel, nsel, d_grad, d_u);
}

void registerPlatforms(void) {}
void registerKernelFactories(void) { g_registered++; }
void registerSDMHIPKernelFactories(void) { registerKernelFactories(); }

}  // extern "C"
